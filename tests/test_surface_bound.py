"""Column bound used to prune the 16x16 -> 256x256 surface argmax sweep.

The surface kernels evaluate only the columns q whose bound
B_q = sum_k1 |G[k1][q]| (G = the padded spectrum transformed along the column
axis) satisfies B_q (1 + 2^-10) + 2^-11 |a|max >= max(a), a = the 16x16 trimmed
samples. These checks run the oracle's surface oversampling
(oracle.ampcor_oracle.oversample_r2r, cuOverSampler.cpp:141-148) and confirm,
on noise, peaked, flat, two-peak and near-tie surfaces, that (1) B_q bounds
every value of column q, (2) the surface's maximum is at least max(a), and (3)
every column holding the argmax or a value within 2^-12 |a|max of the maximum
is kept, so pruning cannot change the argmax, the tie rule or a small margin.
"""
import numpy as np
import pytest

from oracle import ampcor_oracle as O


def _surfaces(seed):
    rng = np.random.default_rng(seed)
    yy, xx = np.mgrid[0:16, 0:16]
    out = [rng.standard_normal((16, 16)) * 0.2,                                   # noise
           np.exp(-((yy - 7.3) ** 2 + (xx - 8.6) ** 2) / 6.0) + 0.05 * rng.standard_normal((16, 16)),
           np.full((16, 16), 0.25),                                               # flat
           np.zeros((16, 16))]
    two = 0.9 * np.exp(-((yy - 3.0) ** 2 + (xx - 4.0) ** 2) / 3.0) \
        + 0.9 * np.exp(-((yy - 12.0) ** 2 + (xx - 11.0) ** 2) / 3.0)            # two equal peaks
    out.append(two)
    tie = np.exp(-((yy - 8.0) ** 2 + (xx - 7.5) ** 2) / 4.0)                     # symmetric: exact ties
    out.append(tie)
    for _ in range(20):
        c = rng.uniform(2, 13, 2)
        w = rng.uniform(1.5, 8.0)
        out.append(rng.uniform(0.3, 1.0) * np.exp(-((yy - c[0]) ** 2 + (xx - c[1]) ** 2) / w)
                   + rng.uniform(0.0, 0.3) * rng.standard_normal((16, 16)))
    return np.stack(out)


def _columns(adj):
    spec = np.fft.ifft2(adj.astype(np.complex128)) * 256.0
    padded = O._pad_quadrants(spec, 256, 256)
    return np.fft.fft(padded, axis=2)            # G[w][k1][q]


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_column_bound_keeps_argmax_and_near_maxima(seed):
    adj = _surfaces(seed)
    surf = O.oversample_r2r(adj, 256, 256)
    G = _columns(adj)
    B = np.abs(G).sum(axis=1)                                     # [w][q]
    for w in range(len(adj)):
        s = surf[w]
        colmax = np.abs(s).max(axis=0)
        assert np.all(colmax <= B[w] * (1 + 1e-12) + 1e-12), w
        amax, amag = adj[w].max(), np.abs(adj[w]).max()
        assert s.max() >= amax - 1e-9 * max(amag, 1.0)
        keep = ~(B[w] * (1 + 2.0 ** -10) + 2.0 ** -11 * amag < amax)
        near = (s >= s.max() - 2.0 ** -12 * amag).any(axis=0)     # columns with argmax or a close runner-up
        assert np.all(keep[near]), w
        loc = np.unravel_index(np.argmax(s), s.shape)
        assert keep[loc[1]]


def test_column_bound_prunes_correlation_peaks():
    # a sharp, well-sampled peak leaves only a few columns to sweep
    yy, xx = np.mgrid[0:16, 0:16]
    adj = np.exp(-((yy - 7.6) ** 2 + (xx - 8.2) ** 2) / 3.0)[None] * 0.85
    G = _columns(adj)
    B = np.abs(G).sum(axis=1)[0]
    keep = ~(B * (1 + 2.0 ** -10) + 2.0 ** -11 * np.abs(adj).max() < adj.max())
    assert 1 <= keep.sum() <= 64
