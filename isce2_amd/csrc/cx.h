// Complex arithmetic and small helpers shared by the engine's kernels.
#pragma once
#include <hip/hip_runtime.h>
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
// multiply by -i (dir=-1) or +i (dir=+1)
template <class T, int DIR> __host__ __device__ __forceinline__ cx<T> mul_i(cx<T> a) {
    return DIR > 0 ? cx<T>{-a.y, a.x} : cx<T>{a.y, -a.x};
}
template <class T> __device__ __forceinline__ T amag(cx<T> a);
template <> __device__ __forceinline__ float amag<float>(cx<float> a) { return sqrtf(a.x * a.x + a.y * a.y); }
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

}  // namespace ampcor
