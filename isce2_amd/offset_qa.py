"""Offset-field quality control downstream of the dense offsets.

What isce2 does with an offset field before using it, restated on numpy
arrays (one row per offset):

* :meth:`OffsetField.cull` — drop offsets below an SNR threshold
  (isceobj/Location/Offset.py:364-378: ``snr < threshold`` is dropped);
* :func:`offoutliers` — straight-line fits of the across offsets against the
  across position and of the down offsets against the down position, then
  removal of offsets whose perpendicular distance to either line exceeds
  ``distance`` (isceobj/Util/offoutliers/src/offoutliers.F:1-139,
  Offoutliers.py:81-91,258-277);
* :func:`fitoff` — iterative affine fit ``(x2, y2) = M (x1, y1) + t`` with
  n-sigma culling until both RMS residuals fall below ``maxrms``, in the L1
  (default) or L2 norm, and the rotation / scale / skew decomposition of M
  (mroipac/fitoff/src/fitoff.F:1-389, Fitoff.py:57-68);
* :func:`fit_offset_polynomials` — SNR-weighted 2-D polynomial fits of the
  offsets (isceobj/Location/Offset.py:392-511 with
  stdproc/stdproc/offsetpoly/src/offsetpoly.f90).

Positions follow the reference: ``x`` across (range), ``y`` down (azimuth).
Least-squares solves use an SVD whose singular values below 1e-6 of the
largest are dropped, as the reference's svdfit does; the L1 fit is solved as
the dual linear program (HiGHS through scipy) instead of the reference's
modified simplex (ACM algorithm 478): the minimum is the same, and the
minimiser is unique for data in general position.
"""
from __future__ import annotations

import numpy as np

__all__ = ["OffsetField", "offoutliers", "fitoff", "fit_offset_polynomials", "svd_solve"]


class OffsetField:
    """Offsets as columns: across position ``x``, across offset ``dx``, down
    position ``y``, down offset ``dy``, ``snr`` and the covariance terms
    (``cov_across``, ``cov_down``, ``cov_cross``)."""

    def __init__(self, x, dx, y, dy, snr, cov_across=None, cov_down=None, cov_cross=None):
        self.x = np.asarray(x, np.float64).ravel()
        self.dx = np.asarray(dx, np.float64).ravel()
        self.y = np.asarray(y, np.float64).ravel()
        self.dy = np.asarray(dy, np.float64).ravel()
        self.snr = np.asarray(snr, np.float64).ravel()
        n = len(self.x)
        z = np.zeros(n)
        self.cov_across = z if cov_across is None else np.asarray(cov_across, np.float64).ravel()
        self.cov_down = z if cov_down is None else np.asarray(cov_down, np.float64).ravel()
        self.cov_cross = z if cov_cross is None else np.asarray(cov_cross, np.float64).ravel()
        for a in (self.dx, self.y, self.dy, self.snr, self.cov_across, self.cov_down, self.cov_cross):
            if len(a) != n:
                raise ValueError("offset columns differ in length")

    @classmethod
    def from_dense(cls, offset, snr, cov, x_start, y_start, x_step, y_step):
        """From the dense rasters (offset ``[nd][na][2]`` = (down, across), snr
        ``[nd][na]``, cov ``[nd][na][3]`` = (down, across, cross)) and the grid of
        window centres (the ``geom`` of the dense-offsets command line)."""
        offset = np.asarray(offset)
        nd, na = offset.shape[:2]
        yy, xx = np.meshgrid(y_start + y_step * np.arange(nd), x_start + x_step * np.arange(na), indexing="ij")
        cov = np.asarray(cov).reshape(nd, na, 3)
        return cls(xx, offset[..., 1], yy, offset[..., 0], np.asarray(snr).reshape(nd, na),
                   cov_across=cov[..., 1], cov_down=cov[..., 0], cov_cross=cov[..., 2])

    def __len__(self):
        return len(self.x)

    def take(self, idx):
        idx = np.asarray(idx)
        return OffsetField(self.x[idx], self.dx[idx], self.y[idx], self.dy[idx], self.snr[idx],
                           self.cov_across[idx], self.cov_down[idx], self.cov_cross[idx])

    def cull(self, snr=0.0):
        """Keep offsets whose SNR is not below ``snr`` (Offset.py:364-378); NaN SNRs are kept,
        as the reference's ``<`` comparison keeps them."""
        return self.take(np.nonzero(~(self.snr < snr))[0])

    def unpack(self):
        """Rows ``[x, dx, y, dy, snr]`` (OffsetField.unpackOffsets, Offset.py:341-351)."""
        return np.stack([self.x, self.dx, self.y, self.dy, self.snr], axis=1)


def svd_solve(a, b, tol=1e-6):
    """Least-squares solution of ``a @ c = b`` through an SVD, singular values
    below ``tol`` times the largest set to zero (svdfit, offoutliers.F:176-201)."""
    u, w, vt = np.linalg.svd(np.asarray(a, np.float64), full_matrices=False)
    keep = (w > 0) & (w >= tol * (w.max() if w.size else 0.0))
    winv = np.where(keep, 1.0 / np.where(keep, w, 1.0), 0.0)
    return vt.T @ (winv * (u.T @ np.asarray(b, np.float64)))


def _line_distance(t, v, slope, icpt):
    # distance from (t, v) to v = slope t + icpt through the foot of the perpendicular
    # (offoutliers.F:97-120)
    x = (t - slope * icpt + slope * v) / (1.0 + slope * slope)
    y = slope * x + icpt
    return np.sqrt((x - t) ** 2 + (y - v) ** 2)


def offoutliers(field: OffsetField, distance: float, snr_threshold: float = 0.0):
    """Outlier removal of isceobj/Util/offoutliers (offoutliers.F:58-127).

    The field is first culled at ``snr_threshold`` (Offoutliers.addOffsets);
    across offsets are fitted as ``c_ac + s_ac * x`` and down offsets as
    ``c_dn + s_dn * y`` (unit sigmas); an offset is kept when its
    perpendicular distances to both lines are ``<= distance``. Returns a dict
    with ``index`` (into the culled field), ``refined`` (the kept offsets),
    ``average_offset_across`` / ``average_offset_down`` (the intercepts, as
    the reference's getters return them) and the two slopes."""
    f = field.cull(snr_threshold)
    n = len(f)
    if n == 0:
        raise ValueError("no offsets left after the SNR cull")
    ones = np.ones(n)
    c_ac, s_ac = svd_solve(np.stack([ones, f.x], axis=1), f.dx)
    c_dn, s_dn = svd_solve(np.stack([ones, f.y], axis=1), f.dy)
    d_ac = _line_distance(f.x, f.dx, s_ac, c_ac).astype(np.float32)   # the reference keeps real*4 distances
    d_dn = _line_distance(f.y, f.dy, s_dn, c_dn).astype(np.float32)
    idx = np.nonzero((d_ac <= distance) & (d_dn <= distance))[0]
    return dict(index=idx, refined=f.take(idx), average_offset_across=float(c_ac), average_offset_down=float(c_dn),
                slope_across=float(s_ac), slope_down=float(s_dn), culled=f)


def _l1_fit(a, b):
    """min_c sum |a c - b|, through the dual linear program max b.y subject to
    a^T y = 0, -1 <= y <= 1 (as many rows as unknowns, so the simplex is quick);
    the primal solution is minus the multipliers of the equality rows. Strong
    duality is checked: the primal objective must equal the dual one."""
    from scipy.optimize import linprog

    m, n = a.shape
    r = linprog(-b, A_eq=a.T, b_eq=np.zeros(n), bounds=[(-1.0, 1.0)] * m, method="highs")
    if r.status != 0:
        raise RuntimeError(f"L1 fit failed: {r.message}")
    c = -np.asarray(r.eqlin.marginals, np.float64)
    primal, dual = np.abs(a @ c - b).sum(), -r.fun
    if not abs(primal - dual) <= 1e-7 * max(1.0, abs(dual)):
        raise RuntimeError(f"L1 fit: primal {primal} and dual {dual} objectives differ")
    return c


def _sdev(r):
    # sample standard deviation (dmoment, fitoff.F:1009-1040)
    p = len(r)
    if p <= 1:
        raise ValueError("fitoff: at least 2 points are needed for the residual statistics")
    return float(np.sqrt(((r - r.mean()) ** 2).sum() / (p - 1)))


def _affine_analysis(m):
    """Rotation, axis scales and skew of the 2x2 affine matrix by a Householder QR
    (fitoff.F:323-365)."""
    a = np.array([[m[0], m[1]], [m[2], m[3]]], np.float64)
    scale = max(abs(a[0, 0]), abs(a[1, 0]))
    if scale == 0.0:
        return dict(rotation_deg=0.0, scale=(0.0, abs(a[1, 1])), skew=0.0, rotation=np.eye(2))
    col = a[:, 0] / scale
    sigma = np.copysign(np.sqrt((col ** 2).sum()), col[0])
    u = col.copy()
    u[0] += sigma
    c1 = sigma * u[0]
    d1 = -scale * sigma
    tau = (u[0] * a[0, 1] + u[1] * a[1, 1]) / c1
    r12 = a[0, 1] - tau * u[0]
    d2 = a[1, 1] - tau * u[1]
    u2 = 0.5 * (u[0] ** 2 + u[1] ** 2)
    rot = np.array([[1.0 - u[0] ** 2 / u2, -u[0] * u[1] / u2], [-u[0] * u[1] / u2, 1.0 - u[1] ** 2 / u2]])
    if d1 < 0:
        rot[:, 0] = -rot[:, 0]
        d1 = -d1
        r12 = -r12
    elif d2 < 0:
        rot[:, 1] = -rot[:, 1]
        d2 = -d2
    return dict(rotation_deg=float(np.degrees(np.arctan2(rot[1, 0], rot[0, 0]))), scale=(abs(d1), abs(d2)),
                skew=float(r12 / d1), rotation=rot)


def fitoff(field: OffsetField, nsig=1.5, maxrms=0.08, minpoint=50, miniter=3, maxiter=30, l1norm=True):
    """Iterative affine fit with n-sigma culling (fitoff.F:44-389; defaults of Fitoff.py:61-67).

    Each iteration fits ``x + dx = m1 x + m2 y + m5`` and ``y + dy = m3 x +
    m4 y + m6`` (L1 or L2), computes the residual RMS and sample standard
    deviation per axis and, while an RMS exceeds ``maxrms``, keeps the offsets
    whose residuals are below ``nsig`` standard deviations on both axes. It
    stops when both RMS are below ``maxrms`` ("converged"), when ``minpoint``
    or fewer offsets remain ("min_points"), when the relative change of the
    solution stops changing after ``miniter`` iterations ("no_change"), or
    after ``maxiter + 1`` fits ("max_iter")."""
    f = field
    if len(f) < 2:
        raise ValueError("fitoff: need at least 2 points")
    x1, y1 = f.x.copy(), f.y.copy()
    x2, y2 = x1 + f.dx, y1 + f.dy
    keep = np.arange(len(f))
    x_prev = None
    length = length_last = 0.0
    status = "max_iter"
    sol = None
    rmsx = rmsy = np.inf
    it = 0
    for it in range(1, maxiter + 2):
        xa, ya = x1[keep], y1[keep]
        ax = np.stack([xa, ya, np.ones(len(keep))], axis=1)
        if l1norm:
            px, py = _l1_fit(ax, x2[keep]), _l1_fit(ax, y2[keep])
        else:
            px, py = svd_solve(ax, x2[keep], tol=0.0), svd_solve(ax, y2[keep], tol=0.0)
        sol = np.array([px[0], px[1], py[0], py[1], px[2], py[2]])
        resx = ax @ px - x2[keep]
        resy = ax @ py - y2[keep]
        rmsx, rmsy = float(np.sqrt((resx ** 2).mean())), float(np.sqrt((resy ** 2).mean()))
        xsdev, ysdev = _sdev(resx), _sdev(resy)
        tx = nsig * xsdev if rmsx > maxrms else 99999.0
        ty = nsig * ysdev if rmsy > maxrms else 99999.0
        if rmsx > maxrms or rmsy > maxrms:
            keep = keep[(np.abs(resx) < tx) & (np.abs(resy) < ty)]
        if len(keep) <= minpoint:
            status = "min_points"
            break
        if rmsx < maxrms and rmsy < maxrms:
            status = "converged"
            break
        if it > 1:
            length = float(np.sqrt(((sol - x_prev) ** 2).sum() / (x_prev ** 2).sum()) * 100.0)
        if it >= miniter and (length - length_last) == 0.0 and (rmsx > maxrms or rmsy > maxrms):
            status = "no_change"
            break
        length_last = length
        x_prev = sol.copy()
    out = dict(affine=sol, average_offset_across=float(sol[4]), average_offset_down=float(sol[5]),
               index=keep, refined=f.take(keep), rms_x=rmsx, rms_y=rmsy, iterations=it, status=status)
    out.update(_affine_analysis(sol))
    return out


# ---------------------------------------------------------------------------
# SNR-weighted polynomial fits
# ---------------------------------------------------------------------------
# basis of offsetpoly.f90 in coefficient order: 1, r, a, r a, r^2, a^2, r a^2, r^2 a, r^3, a^3
_POLY_TERMS = [(0, 0), (1, 0), (0, 1), (1, 1), (2, 0), (0, 2), (1, 2), (2, 1), (3, 0), (0, 3)]


def _num_coeff(range_order, azimuth_order, max_order):
    # Offset.py:396-412
    if range_order == azimuth_order and range_order <= 3:
        if range_order == 1:
            return 3 if max_order else 4
        if range_order == 2 and max_order:
            return 6
        if range_order == 3 and max_order:
            return 10
    return 0


def fit_offset_polynomials(field: OffsetField, range_order=2, azimuth_order=2, max_order=True):
    """Azimuth- and range-offset polynomials (OffsetField.getFitPolynomials, Offset.py:392-511).

    Returns ``(azimuth_coeffs, range_coeffs, mean_azimuth)`` with ``coeffs[i][j]``
    multiplying ``(y - mean_azimuth)**i * x**j``, ``mean_azimuth`` = the smallest
    down position. Fits weight each offset by ``1 / (1 + 1/snr)`` (offsetpoly.f90:
    ``sig = 1 + 1/snr``); the terms are offsetpoly's for the orders it covers
    (3, 4, 6 or 10 coefficients, Offset.py:396-412) and every ``y^i x^j`` with
    ``i <= azimuth_order``, ``j <= range_order`` otherwise."""
    azmin = float(field.y.min())
    r, a = field.x, field.y - azmin
    sig = 1.0 + 1.0 / field.snr
    nc = _num_coeff(range_order, azimuth_order, max_order)
    if nc == 4:
        terms = [(0, 0), (1, 0), (0, 1), (1, 1)]
    elif nc:
        terms = _POLY_TERMS[:nc]
    else:
        terms = [(j, i) for i in range(azimuth_order + 1) for j in range(range_order + 1)]
    basis = np.stack([r ** pj * a ** pi for (pj, pi) in terms], axis=1) / sig[:, None]
    # columns scaled to unit norm before the SVD (tolerance 1e-12 as the shared
    # svdfit, isceobj/Util/src/svd.F:3): pixel positions to the third power
    # would otherwise put the singular values of a well-posed fit below it
    norm = np.sqrt((basis ** 2).sum(axis=0))
    norm[norm == 0] = 1.0
    out = []
    for v in (field.dy, field.dx):
        c = svd_solve(basis / norm, v / sig, tol=1e-12) / norm
        co = np.zeros((azimuth_order + 1, range_order + 1))
        for (pj, pi), val in zip(terms, c):
            co[pi, pj] = val
        out.append(co)
    return out[0], out[1], azmin


def eval_poly(coeffs, y, x, mean_azimuth=0.0):
    """Evaluate ``sum coeffs[i][j] (y - mean_azimuth)^i x^j``."""
    y = np.asarray(y, np.float64) - mean_azimuth
    x = np.asarray(x, np.float64)
    out = np.zeros(np.broadcast(y, x).shape)
    for i in range(coeffs.shape[0]):
        for j in range(coeffs.shape[1]):
            if coeffs[i, j]:
                out = out + coeffs[i, j] * y ** i * x ** j
    return out
