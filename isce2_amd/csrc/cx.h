// Complex arithmetic and small helpers shared by the engine's kernels.
#pragma once
#include <hip/hip_runtime.h>
#include <cfloat>
#include <stdint.h>

namespace ampcor {

template <class T>
struct alignas(2 * sizeof(T)) cx {
    T x, y;
};

template <class T> __host__ __device__ __forceinline__ cx<T> mk(T a, T b) { return cx<T>{a, b}; }
template <class T> __host__ __device__ __forceinline__ cx<T> operator+(cx<T> a, cx<T> b) { return {a.x + b.x, a.y + b.y}; }
template <class T> __host__ __device__ __forceinline__ cx<T> operator-(cx<T> a, cx<T> b) { return {a.x - b.x, a.y - b.y}; }
template <class T> __host__ __device__ __forceinline__ cx<T> operator*(cx<T> a, T s) { return {a.x * s, a.y * s}; }
template <class T> __host__ __device__ __forceinline__ cx<T> cmul(cx<T> a, cx<T> b) {
    return {a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x};
}
// conj(a) * b
template <class T> __host__ __device__ __forceinline__ cx<T> cmulc(cx<T> a, cx<T> b) {
    return {a.x * b.x + a.y * b.y, a.x * b.y - a.y * b.x};
}
template <class T> __host__ __device__ __forceinline__ cx<T> conj(cx<T> a) { return {a.x, -a.y}; }
// acc + conj(a) * b as two chained FMAs per component (the accumulator is the addend)
template <class T> __device__ __forceinline__ cx<T> cmac_conj(cx<T> acc, cx<T> a, cx<T> b) {
    return {fma(a.y, b.y, fma(a.x, b.x, acc.x)), fma(-a.y, b.x, fma(a.x, b.y, acc.y))};
}
// multiply by -i (dir=-1) or +i (dir=+1)
template <class T, int DIR> __host__ __device__ __forceinline__ cx<T> mul_i(cx<T> a) {
    return DIR > 0 ? cx<T>{-a.y, a.x} : cx<T>{a.y, -a.x};
}
// Hardware square root (v_sqrt_f32, within 1 ulp). The default sqrtf expands to a
// correctly rounded sequence of ~20 instructions; amplitudes |z| use this one.
__device__ __forceinline__ float fast_sqrt(float x) { return __builtin_amdgcn_sqrtf(x); }

template <class T> __device__ __forceinline__ T amag(cx<T> a);
template <> __device__ __forceinline__ float amag<float>(cx<float> a) { return fast_sqrt(a.x * a.x + a.y * a.y); }
template <> __device__ __forceinline__ double amag<double>(cx<double> a) { return sqrt(a.x * a.x + a.y * a.y); }

__host__ __device__ __forceinline__ int idivup(int a, int b) { return (a + b - 1) / b; }

// 9-bit bit reversal: tie-break key of the 512-thread argmax of cuOffset.cu:23-84
__device__ __forceinline__ uint32_t bitrev9(uint32_t v) { return __brev(v) >> 23; }

// Argmax candidate ordering used by every maxloc in the engine: larger value
// wins; on equal values the smaller (bitrev9(idx % 512), idx) key wins.
__device__ __forceinline__ uint64_t maxloc_key(uint32_t idx) {
    return (uint64_t(bitrev9(idx & 511u)) << 32) | idx;
}

template <class T>
__device__ __forceinline__ bool better(T va, uint64_t ka, T vb, uint64_t kb) {
    return va > vb || (va == vb && ka < kb);
}

// Running top-2 of an argmax with the tie rule above: best (value, key) and
// the largest value among the other candidates. margin() = top-1 minus top-2
// tells how far the winner is from a rounding-level tie.
template <class T>
struct Top2 {
    T v1 = T(-FLT_MAX), v2 = T(-FLT_MAX);
    uint64_t k1 = ~uint64_t(0);
    __device__ __forceinline__ static T mx(T a, T b) { return a > b ? a : b; }
    __device__ __forceinline__ void add(T v, uint64_t k) {
        if (!(v > T(-FLT_MAX))) return;   // -FLT_MAX and NaN are never selected
        if (better(v, k, v1, k1)) {
            v2 = mx(v2, v1);
            v1 = v; k1 = k;
        } else {
            v2 = mx(v2, v);
        }
    }
    // same as add(v, maxloc_key(idx)) for finite v above -FLT_MAX; the key is only formed on a tie
    __device__ __forceinline__ void add_idx(T v, uint32_t idx) {
        const T lo = v < v1 ? v : v1;
        v2 = mx(v2, lo);
        if (v > v1 || (v == v1 && maxloc_key(idx) < k1)) { v1 = v; k1 = maxloc_key(idx); }
    }
    __device__ __forceinline__ void merge(T ov1, uint64_t ok1, T ov2) {
        if (better(ov1, ok1, v1, k1)) {
            v2 = mx(mx(v2, ov2), v1);
            v1 = ov1; k1 = ok1;
        } else {
            v2 = mx(v2, mx(ov1, ov2));
        }
    }
    // workgroup reduction over NTH threads (wave64); shv, shv2, shk hold NTH/64
    // entries each; every thread returns with the workgroup result
    template <int NTH>
    __device__ void reduce(T *shv, T *shv2, uint64_t *shk) {
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) {
            const T a = __shfl_xor(v1, o), b = __shfl_xor(v2, o);
            const uint64_t k = __shfl_xor(k1, o);
            merge(a, k, b);
        }
        __syncthreads();
        if ((threadIdx.x & 63) == 0) {
            shv[threadIdx.x >> 6] = v1; shv2[threadIdx.x >> 6] = v2; shk[threadIdx.x >> 6] = k1;
        }
        __syncthreads();
        v1 = shv[0]; v2 = shv2[0]; k1 = shk[0];
#pragma unroll
        for (int i = 1; i < NTH / 64; ++i) merge(shv[i], shk[i], shv2[i]);
        __syncthreads();
    }
    // index of the winner (0 when no value was eligible)
    __device__ __forceinline__ int index() const { return k1 == ~uint64_t(0) ? 0 : int(k1 & 0xffffffffu); }
    // 0 for exact ties; +inf when there is no runner-up
    __device__ __forceinline__ float margin() const {
        return v2 > T(-FLT_MAX) ? float(v1 - v2) : INFINITY;
    }
};

}  // namespace ampcor
