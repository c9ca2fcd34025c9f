// Curve order of an adapted mesh for the step kernels.
//
// The reference's adaptation fills the holes left by coarsened families
// with cells moved from the end of the storage (mesh/amr/swap_lists.h:28-58),
// so after some adaptations the cells of one refined region are spread over
// the whole index range and a wavefront's neighbour gathers touch a cache
// line per lane.  The step's results do not depend on where a cell is
// stored (each cell's update sums its own slots in slot order; the CFL
// minimum is order-free), so the step kernels can run on a copy of the mesh
// whose cells are sorted along a space-filling curve: base cell (b_idx)
// first, then the Morton index of the cell inside its base cell at the
// finest level, z_idx * 4^(lmax - level) (children are numbered 4 z + i in
// Z order, mesh/amr/refining.h).  The reference order stays the mesh's own:
// adaptation, downloads and every other entry point see it.
#include "zfv_internal.hpp"

#include <hipcub/hipcub.hpp>

#include "zfv_scan.hpp"

namespace zfv {

namespace {

inline dim3 grid256(int n) { return dim3((unsigned)((n + 255) / 256)); }

__global__ void __launch_bounds__(256) k_curve_keys(const int* __restrict__ b_idx, const int* __restrict__ z_idx,
                                                    const int* __restrict__ level, int n, int lmax,
                                                    unsigned long long* __restrict__ keys, int* __restrict__ vals) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= n) return;
    const int L = level[c];
    const int shift = 2 * (lmax - (L < 0 ? 0 : (L > lmax ? lmax : L)));
    keys[c] = ((unsigned long long)(unsigned)b_idx[c] << (2 * lmax)) | ((unsigned long long)(unsigned)z_idx[c] << shift);
    vals[c] = c;
}

// Dense curve slots: cell c at base b and level L occupies the finest-level
// Morton range starting at b 4^lmax + z_idx 4^(lmax - L); its index goes to
// the first slot of that range.  Compacting the slots in order gives the
// curve order (the ranges of one mesh's cells are disjoint).
__global__ void __launch_bounds__(256) k_curve_slots(const int* __restrict__ b_idx, const int* __restrict__ z_idx,
                                                     const int* __restrict__ level, int n, int lmax,
                                                     int* __restrict__ slots, int* __restrict__ counts,
                                                     long long n_slots) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= n) return;
    const int L = level[c];
    if (L < 0 || L > lmax) return;   // excluded by the mesh checks; never index out of range
    const long long at = ((long long)b_idx[c] << (2 * lmax)) + ((long long)z_idx[c] << (2 * (lmax - L)));
    if (at >= 0 && at < n_slots) {
        slots[at] = c;
        atomicAdd(&counts[at >> 6], 1);
    }
}

// Compaction of the dense slots in 64-slot chunks (the scatter counts each
// chunk's cells): after a scan of the counts, one wavefront per chunk gives
// every cell its curve position = chunk offset + its rank among the chunk's
// set lanes (ballot).  The write pass also resets the slots to -1 and the
// counts to 0 for the next build.
__global__ void __launch_bounds__(256) k_curve_write(int* __restrict__ slots, int n_chunks,
                                                     int* __restrict__ counts, const int* __restrict__ offsets,
                                                     int* __restrict__ perm, int* __restrict__ iperm) {
    const int chunk = (int)((blockIdx.x * blockDim.x + threadIdx.x) >> 6);
    const int lane = threadIdx.x & 63;
    if (chunk >= n_chunks) return;
    const size_t at = (size_t)chunk * 64 + lane;
    const int c = slots[at];
    const unsigned long long m = __ballot(c >= 0);
    if (c >= 0) {
        const int pos = offsets[chunk] + __popcll(m & ((1ULL << lane) - 1ULL));
        perm[pos] = c;
        iperm[c] = pos;
        slots[at] = -1;
    }
    if (lane == 0) counts[chunk] = 0;
}

__global__ void __launch_bounds__(256) k_curve_inverse(const int* __restrict__ perm, int* __restrict__ iperm, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) iperm[perm[i]] = i;
}

// Curve-ordered copy of the step kernels' mesh arrays: cell i of the copy is
// cell perm[i] of the mesh; neighbour indices are mapped through iperm
// (boundary slots point at the cell itself, undefined slots too).
__global__ void __launch_bounds__(256) k_curve_mesh(DevMesh src, CurveMesh dst, const int* __restrict__ perm,
                                                    const int* __restrict__ iperm, DevState zsrc, DevState zdst,
                                                    const int* __restrict__ wsrc, int* __restrict__ wdst) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= src.n) return;
    const int o = perm[i];
    if (zsrc.r) {   // the state into curve order (gather_state of the same permutation)
        zdst.r[i] = zsrc.r[o];
        zdst.u[i] = zsrc.u[o];
        zdst.v[i] = zsrc.v[o];
        zdst.p[i] = zsrc.p[o];
        zdst.e[i] = zsrc.e[o];
        wdst[i] = wsrc[o];
    }
    const int code = src.code[o];
    // every load is issued before the first store (the source and copy
    // arrays are of the same types, so stores would otherwise order them)
    const double cx = src.cx[o], cy = src.cy[o], vol = src.vol[o], h = src.h[o];
    int adj[8];
    double area[8], fcx[8], fcy[8], w[8], nx[8], ny[8];
    // undefined slots are never read past their code, canonical normals
    // never past it either: only what the step kernels read is copied
#pragma unroll
    for (int s = 0; s < 8; ++s) {
        const int so = s * src.stride + o;
        const int nib = (code >> (4 * s)) & 15;
        const bool def = (nib & 7) != SC_UNDEF;
        adj[s] = def ? src.adj[so] : -1;
        area[s] = def ? src.area[so] : 0.0;
        fcx[s] = def ? src.fcx[so] : 0.0;
        fcy[s] = def ? src.fcy[so] : 0.0;
        w[s] = def ? src.w[so] : 0.0;
        const bool gen = def && (nib & kGeneralNormal);
        nx[s] = gen ? src.nx[so] : 0.0;
        ny[s] = gen ? src.ny[so] : 0.0;
    }
#pragma unroll
    for (int s = 0; s < 8; ++s) {
        const int a = adj[s];
        adj[s] = (a >= 0 && a < src.n) ? iperm[a] : i;
    }
    dst.cx[i] = cx;
    dst.cy[i] = cy;
    dst.vol[i] = vol;
    dst.h[i] = h;
    dst.code[i] = code;
#pragma unroll
    for (int s = 0; s < 8; ++s) {
        const int si = s * dst.stride + i;
        const int nib = (code >> (4 * s)) & 15;
        dst.adj[si] = adj[s];
        if ((nib & 7) == SC_UNDEF) continue;
        dst.area[si] = area[s];
        dst.fcx[si] = fcx[s];
        dst.fcy[si] = fcy[s];
        dst.w[si] = w[s];
        if (nib & kGeneralNormal) {
            dst.nx[si] = nx[s];
            dst.ny[si] = ny[s];
        }
    }
}

// dst[i] = src[idx[i]] for the five state rows and the isolation counters.
__global__ void __launch_bounds__(256) k_gather_state(DevState src, DevState dst, const int* __restrict__ wsrc,
                                                      int* __restrict__ wdst, const int* __restrict__ idx, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int j = idx[i];
    dst.r[i] = src.r[j];
    dst.u[i] = src.u[j];
    dst.v[i] = src.v[j];
    dst.p[i] = src.p[j];
    dst.e[i] = src.e[j];
    wdst[i] = wsrc[j];
}

// dst[k][i] = src[k][idx[i]] for `rows` rows of `stride` doubles.
__global__ void __launch_bounds__(256) k_gather_rows(const double* __restrict__ src, double* __restrict__ dst,
                                                     const int* __restrict__ idx, int n, int rows, int stride) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int j = idx[i];
    for (int k = 0; k < rows; ++k) dst[(size_t)k * stride + i] = src[(size_t)k * stride + j];
}

}  // namespace

hipError_t curve_keys(const int* b_idx, const int* z_idx, const int* level, int n, int lmax,
                      unsigned long long* keys, int* vals, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_curve_keys, grid256(n), dim3(256), 0, s, b_idx, z_idx, level, n, lmax, keys, vals);
    return hipGetLastError();
}

hipError_t curve_slots(const int* b_idx, const int* z_idx, const int* level, int n, int lmax, int* slots,
                       int* counts, long long n_slots, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_curve_slots, grid256(n), dim3(256), 0, s, b_idx, z_idx, level, n, lmax, slots, counts,
                       n_slots);
    return hipGetLastError();
}

int curve_scan_tiles(int n_chunks) { return scan::tiles_of(n_chunks); }

hipError_t curve_compact(int* slots, int n_chunks, int* counts, int* offsets, int* tile_sums, int* perm, int* iperm,
                         hipStream_t s) {
    hipError_t e = scan::exclusive(counts, offsets, tile_sums, n_chunks, scan::AddInt(), 0, (int*)nullptr, s);
    if (e != hipSuccess) return e;
    const dim3 g((unsigned)(((long long)n_chunks * 64 + 255) / 256));
    hipLaunchKernelGGL(k_curve_write, g, dim3(256), 0, s, slots, n_chunks, counts, offsets, perm, iperm);
    return hipGetLastError();
}

hipError_t curve_sort(const unsigned long long* keys_in, unsigned long long* keys_out, const int* vals_in,
                      int* vals_out, int n, int end_bit, void* tmp, size_t& tmp_bytes, hipStream_t s) {
    return hipcub::DeviceRadixSort::SortPairs(tmp, tmp_bytes, keys_in, keys_out, vals_in, vals_out, n, 0, end_bit, s);
}

hipError_t curve_inverse(const int* perm, int* iperm, int n, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_curve_inverse, grid256(n), dim3(256), 0, s, perm, iperm, n);
    return hipGetLastError();
}

hipError_t curve_mesh(const DevMesh& src, const CurveMesh& dst, const int* perm, const int* iperm,
                      const DevState& zsrc, const DevState& zdst, const int* wsrc, int* wdst, hipStream_t s) {
    if (src.n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_curve_mesh, grid256(src.n), dim3(256), 0, s, src, dst, perm, iperm, zsrc, zdst, wsrc, wdst);
    return hipGetLastError();
}

hipError_t gather_state(const DevState& src, const DevState& dst, const int* wsrc, int* wdst, const int* idx, int n,
                        hipStream_t s) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_gather_state, grid256(n), dim3(256), 0, s, src, dst, wsrc, wdst, idx, n);
    return hipGetLastError();
}

hipError_t gather_rows(const double* src, double* dst, const int* idx, int n, int rows, int stride, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_gather_rows, grid256(n), dim3(256), 0, s, src, dst, idx, n, rows, stride);
    return hipGetLastError();
}

}  // namespace zfv
