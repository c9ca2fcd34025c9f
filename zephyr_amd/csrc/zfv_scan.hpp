// Exclusive scan of a device array in two launches, for the adaptation's
// counters: tiles of 1024 elements (256 threads x 4), kernel 1 writes each
// tile's sum, kernel 2 lets every workgroup add the sums of the tiles before
// its own and scan its tile.  Integer sums, so the result does not depend
// on the order of the additions.  The last workgroup also writes the total
// to `total` (device or pinned host memory), which spares a separate read
// of the scan's tail.  No host-side work beyond the two launches; meant for
// up to a few thousand tiles (each workgroup re-reads the tile sums).
// Op: a device functor combining two T and a static `shfl(T, int d)`, the
// value of lane - d of the wavefront.
#pragma once

#include <hip/hip_runtime.h>

namespace zfv {
namespace scan {

constexpr int kThreads = 256;
constexpr int kPer = 4;
constexpr int kTile = kThreads * kPer;

struct AddInt {
    __device__ __forceinline__ int operator()(int a, int b) const { return a + b; }
    __device__ static __forceinline__ int shfl(int v, int d) { return __shfl_up(v, d, 64); }
};

template <class T, class Op>
__device__ __forceinline__ T wave_inclusive(T v, Op op) {
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const T o = Op::shfl(v, d);
        if (lane >= d) v = op(o, v);
    }
    return v;
}

// Exclusive scan of one value per thread over the workgroup; returns the
// thread's exclusive prefix and the workgroup total in `total`.
template <class T, class Op>
__device__ __forceinline__ T block_exclusive(T v, Op op, T zero, T& total) {
    __shared__ T wsum[kThreads / 64];
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    const T inc = wave_inclusive(v, op);
    if (lane == 63) wsum[wid] = inc;
    __syncthreads();
    T before = zero;
    for (int w = 0; w < wid; ++w) before = op(before, wsum[w]);
    total = zero;
    for (int w = 0; w < kThreads / 64; ++w) total = op(total, wsum[w]);
    __syncthreads();
    const T exc = Op::shfl(inc, 1);
    return op(before, lane == 0 ? zero : exc);
}

template <class T, class Op>
__global__ void __launch_bounds__(kThreads) k_tile_sums(const T* __restrict__ in, T* __restrict__ sums, int n,
                                                        Op op, T zero) {
    const int base = blockIdx.x * kTile + threadIdx.x * kPer;
    T v = zero;
#pragma unroll
    for (int k = 0; k < kPer; ++k)
        if (base + k < n) v = op(v, in[base + k]);
    T total;
    (void)block_exclusive(v, op, zero, total);
    if (threadIdx.x == 0) sums[blockIdx.x] = total;
}

template <class T, class Op>
__global__ void __launch_bounds__(kThreads) k_tile_apply(const T* __restrict__ in, T* __restrict__ out,
                                                         const T* __restrict__ sums, int n, int tiles, Op op,
                                                         T zero, T* total_out) {
    // prefix of the tiles before this one
    T p = zero;
    for (int t = threadIdx.x; t < (int)blockIdx.x; t += kThreads) p = op(p, sums[t]);
    T prefix;
    (void)block_exclusive(p, op, zero, prefix);
    const int base = blockIdx.x * kTile + threadIdx.x * kPer;
    T x[kPer];
    T v = zero;
#pragma unroll
    for (int k = 0; k < kPer; ++k) {
        x[k] = (base + k < n) ? in[base + k] : zero;
        v = op(v, x[k]);
    }
    T tile_total;
    T run = op(prefix, block_exclusive(v, op, zero, tile_total));
#pragma unroll
    for (int k = 0; k < kPer; ++k) {
        if (base + k < n) out[base + k] = run;
        run = op(run, x[k]);
    }
    if (total_out && (int)blockIdx.x == tiles - 1 && threadIdx.x == 0) *total_out = op(prefix, tile_total);
}

// Scratch: one T per tile.
inline int tiles_of(int n) { return (n + kTile - 1) / kTile; }

template <class T, class Op>
hipError_t exclusive(const T* in, T* out, T* tile_sums, int n, Op op, T zero, T* total_out, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    const int tiles = tiles_of(n);
    hipLaunchKernelGGL((k_tile_sums<T, Op>), dim3(tiles), dim3(kThreads), 0, s, in, tile_sums, n, op, zero);
    hipLaunchKernelGGL((k_tile_apply<T, Op>), dim3(tiles), dim3(kThreads), 0, s, in, out, tile_sums, n, tiles, op,
                       zero, total_out);
    return hipGetLastError();
}

}  // namespace scan
}  // namespace zfv
