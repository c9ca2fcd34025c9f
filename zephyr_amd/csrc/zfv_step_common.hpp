// Step-kernel plumbing shared by the single- and multi-material kernels:
// XCD-aware workgroup-to-cell mapping, per-problem tiles of a batch, the
// rotating CFL slots (step_dt / step_prologue / block_min_atomic), state
// loads and stores, and a slot's face frame.
#pragma once

#include "zfv_internal.hpp"

namespace zfv {
namespace {

constexpr unsigned long long kInfBits = 0x7ff0000000000000ULL;   // +inf

__device__ __forceinline__ int xcd_block(int b, int G) {
    const int x = b & 7, q = G >> 3, r = G & 7;
    return x * q + (x < r ? x : r) + (b >> 3);
}

// The cells of one workgroup: tile = XCD-aware block index, cells_per_block
// cells from c0.  In a batch of problems (chunk_pid set) every problem
// starts on a 512-cell boundary, so a workgroup lies in one problem `pid`;
// its cells end at `end`, and `first` marks the problem's first workgroup,
// which does the per-problem step prologue.
struct Tile {
    int c0, end, pid;
    bool first;
};

__device__ __forceinline__ Tile tile_of(const DevMesh& M, int cells_per_block) {
    const int tile = xcd_block((int)blockIdx.x, (int)gridDim.x);
    Tile t;
    t.c0 = tile * cells_per_block;
    if (M.chunk_pid) {
        t.pid = M.chunk_pid[t.c0 >> kChunkShift];
        t.end = M.seg_end[t.pid];
        t.first = t.c0 == M.seg_begin[t.pid];
    } else {
        t.pid = 0;
        t.end = M.n;
        t.first = tile == 0;
    }
    return t;
}

__device__ __forceinline__ Z load_z(const DevState& s, int c) {
    return Z{s.r[c], s.u[c], s.v[c], s.p[c], s.e[c]};
}

__device__ __forceinline__ void store_z(const DevState& s, int c, const Z& z) {
    s.r[c] = z.r;
    s.u[c] = z.u;
    s.v[c] = z.v;
    s.p[c] = z.p;
    s.e[c] = z.e;
}

// h / (|v| + c) (sm_fluid.cpp:105-110, incircle precomputed per cell).
__device__ __forceinline__ double cell_dt(const Z& z, double h, const Eos& eos) {
    double c = eos_c(eos, z.r, z.p);
    return h / (sqrt(vsq(z)) + c);
}

// dt of problem pid: three rotating slots per problem.
__device__ __forceinline__ double step_dt(const StepCtl& ctl, int pid) {
    double dmin = __longlong_as_double((long long)ctl.dtmin[3 * pid + ctl.rd]);
    return smin(ctl.cfl * dmin, *ctl.max_dt);
}

// The axisymmetric pressure source of SmFluid: momentum.y += k coeff p dt
// with coeff = volume / volume_alt (M.axc; nothing for a planar step) and
// k = 0.5 in the predictor (sm_fluid.cpp:166-169, 242-245, 347-350).
template <bool HALF>
__device__ __forceinline__ void axial_source(const DevMesh& M, int c, double p, double dt, Q& q) {
    if (M.axc == nullptr) return;
    const double coeff = M.axc[c];
    q.my = q.my + (HALF ? 0.5 * coeff : coeff) * p * dt;
}

__device__ __forceinline__ void step_prologue(const StepCtl& ctl, const Tile& t, double dt) {
    if (t.first && threadIdx.x == 0) {
        ctl.dt_out[t.pid] = dt;
        ctl.dtmin[3 * t.pid + ctl.clr] = kInfBits;
    }
}

__device__ __forceinline__ void block_min_atomic(unsigned long long* slot, double v) {
    if (!(v < INFINITY)) v = INFINITY;   // NaN and +inf never win
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        double o = __shfl_xor(v, off, 64);
        v = (o < v) ? o : v;
    }
    __shared__ double red[kMaxBlock / 64];
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    if (lane == 0) red[wid] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        double m = red[0];
        const int nw = (int)(blockDim.x >> 6);
        for (int i = 1; i < nw; ++i) m = (red[i] < m) ? red[i] : m;
        if (m < INFINITY) atomicMin(slot, (unsigned long long)__double_as_longlong(m));
    }
}

// Frame of slot s of cell c (code nibble `nib`).
__device__ __forceinline__ Frame slot_frame(const DevMesh& M, int c, int s, int nib) {
    const bool gen = (nib & kGeneralNormal) != 0;
    double nx = 0.0, ny = 0.0;
    if (gen) {
        nx = M.nx[s * M.stride + c];
        ny = M.ny[s * M.stride + c];
    }
    return frame_of(s & 3, gen, nx, ny);
}

}  // namespace
}  // namespace zfv
