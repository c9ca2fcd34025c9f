// Tiled MUSCL step for structured rectangles on gfx950: one kernel per step
// computes, for a tile of kTileI x kTileJ cells, the limited LSM gradient and
// the predictor's half state of the tile and its one-cell halo, keeps them in
// LDS, and then runs the corrector of the tile's cells from LDS, with the
// next step's CFL candidate fused at the end.  The state is read once from
// HBM (the tile plus a two-cell halo, itself mostly served by L2 since the
// neighbouring tiles of one XCD read it too) and written once; the half
// states and gradients never leave the chip.
//
// Every per-cell value is computed by the same operations in the same order
// as the mesh-array kernels (zfv_kernels.hip, through the shared per-face
// terms of zfv_step_faces.hpp): the slot geometry comes from per-column and
// per-row tables that the host checked, cell by cell, against the slot
// arrays, and the neighbours' values from LDS instead of HBM.  The results
// are therefore bit-identical to the two-kernel step.
//
// Layout of a workgroup (kTileThreads = (kTileI + 2) * 64 threads): wavefront
// r takes row i = i0 - 1 + r of the halo box and lane l the cell
// j = j0 - 1 + l, so a wavefront reads 64 consecutive cells; the outputs are
// rows 1..kTileI and lanes 1..kTileJ.  Cell index c = off + i * ny + j.
#include <algorithm>

#include "zfv_internal.hpp"
#include "zfv_step_common.hpp"
#include "zfv_step_faces.hpp"

namespace zfv {

namespace {

constexpr int kBJ = 68;   // state box row: the tile's cells and a two-cell halo on each side (66), padded
constexpr int kHJ = 64;   // half-state / gradient box row: the tile's cells and a one-cell halo

// min over the workgroup of the CFL candidates, then one 64-bit atomicMin on
// the bit pattern (block_min_atomic for workgroups of up to 16 wavefronts).
template <int NW>
__device__ __forceinline__ void tile_min_atomic(unsigned long long* slot, double v, double* red) {
    if (!(v < INFINITY)) v = INFINITY;   // NaN and +inf never win
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        const double o = __shfl_xor(v, off, 64);
        v = (o < v) ? o : v;
    }
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    if (lane == 0) red[wid] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        double m = red[0];
        for (int i = 1; i < NW; ++i) m = (red[i] < m) ? red[i] : m;
        if (m < INFINITY) atomicMin(slot, (unsigned long long)__double_as_longlong(m));
    }
}

__device__ __forceinline__ int wrap(int g, int n) {
    const int m = g % n;
    return m < 0 ? m + n : m;
}

// Problem of tile b: the last problem whose first tile is <= b.
template <int V>
__device__ __forceinline__ int problem_of(const TileMesh& T, int b) {
    int lo = 0, hi = T.n_probs - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (T.prob[mid].tile0[V] <= b) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

// Slot code of side s of cell (i, j) (L, R, B, T = 0..3).
__device__ __forceinline__ int side_type(const TileProb& P, int s, int i, int j) {
    switch (s) {
        case 0: return i > 0 ? SC_INNER : P.t[0];
        case 1: return i < P.nx - 1 ? SC_INNER : P.t[1];
        case 2: return j > 0 ? SC_INNER : P.t[2];
        default: return j < P.ny - 1 ? SC_INNER : P.t[3];
    }
}

// Whether side s reads the neighbour across it (an inner side, or a
// periodic side of a periodic direction); otherwise adj is the cell itself.
__device__ __forceinline__ bool uses_neighbour(const TileProb& P, int s, int t) {
    return t == SC_INNER || (t == SC_PERIODIC && P.per[s >> 1]);
}

__device__ __forceinline__ int side_di(int s) { return s == 0 ? -1 : (s == 1 ? 1 : 0); }
__device__ __forceinline__ int side_dj(int s) { return s == 2 ? -1 : (s == 3 ? 1 : 0); }

// Slot geometry of side s of cell (i, j) from the tables: face centre, area,
// LSM weight, and the neighbour's centre across an inner side.
struct SideGeom {
    double fx, fy, a, w, ncx, ncy;
};

__device__ __forceinline__ SideGeom side_geom(const TileMesh& T, const TileProb& P, int s, int i, int j) {
    const TileCol& C = T.col[P.col + i];
    const TileRow& R = T.row[P.row + j];
    SideGeom g;
    switch (s) {
        case 0:
            g.fx = C.fxL; g.fy = R.fyL; g.a = R.aL; g.w = C.wL;
            g.ncx = i > 0 ? T.col[P.col + i - 1].cx : C.cx; g.ncy = R.cy;
            break;
        case 1:
            g.fx = C.fxR; g.fy = R.fyR; g.a = R.aR; g.w = C.wR;
            g.ncx = i < P.nx - 1 ? T.col[P.col + i + 1].cx : C.cx; g.ncy = R.cy;
            break;
        case 2:
            g.fx = C.fxB; g.fy = R.fyB; g.a = C.aB; g.w = R.wB;
            g.ncx = C.cx; g.ncy = j > 0 ? T.row[P.row + j - 1].cy : R.cy;
            break;
        default:
            g.fx = C.fxT; g.fy = R.fyT; g.a = C.aT; g.w = R.wT;
            g.ncx = C.cx; g.ncy = j < P.ny - 1 ? T.row[P.row + j + 1].cy : R.cy;
            break;
    }
    return g;
}

// LSM data of one slot (lsm_neib + slot_load of the mesh-array kernels) from
// values: zl the state across the slot (the cell's own where adj is itself).
__device__ __forceinline__ SlotL tile_slot(const Frame& fr, int t, const Z& zc, const Z& zl, double ccx, double ccy,
                                           const SideGeom& g) {
    SlotL L;
    const Z zn = (t == SC_INNER || t == SC_PERIODIC) ? zl : frame_boundary(fr, zc, t == SC_WALL);
    double px = g.ncx, py = g.ncy;
    if (t != SC_INNER) frame_symm(fr, ccx, ccy, g.fx, g.fy, px, py);
    L.drx = px - ccx;
    L.dry = py - ccy;
    lsm_drn(fr, L.drx, L.dry, L.dnx, L.dny);
    L.w = g.w;
    L.dF[0] = zn.r - zc.r;
    L.dF[1] = zn.u - zc.u;
    L.dF[2] = zn.v - zc.v;
    L.dF[3] = zn.p - zc.p;
    L.xp = fr.gen || fr.side < 2;
    L.yp = fr.gen || fr.side >= 2;
    return L;
}

}  // namespace

// SmFluid::update with acc 2 (sm_fluid.cpp:81-102) on one tile: compute_grad
// (sm_fluid.cpp:116-125, gradient.h:226-282, 412-474), fluxes_stage1
// (:176-253) and fluxes_stage2 (:255-366) with HLLC, plus the next step's
// CFL candidate (:104-114).
// The workgroup's LDS (declared by each kernel, as separate arrays): the
// state box zs (rho, u, v, p), the half states hs, limited gradients gs
// (d/dx then d/dy of rho, u, v, p) and isolation counters ws of the halo
// box, the CFL reduction red.
#define ZFV_TILE_LDS(TI)                                  \
    __shared__ double zs[4][((TI) + 4) * kBJ];            \
    __shared__ double hs[5][((TI) + 2) * kHJ];            \
    __shared__ double gs[8][((TI) + 2) * kHJ];            \
    __shared__ int ws[((TI) + 2) * kHJ];                  \
    __shared__ double red[(TI) + 2]

// One tile b (of the tile list of shape V) by the whole workgroup.
template <int LIM, int V, int TI>
__device__ __forceinline__ void tile_body(const TileMesh& T, const DevState& in, const DevState& out,
                                          const int* __restrict__ wait_in, int* __restrict__ wait_out,
                                          const Eos& eos, const StepCtl& ctl, int b,
                                          double (&zs)[4][(TI + 4) * kBJ], double (&hs)[5][(TI + 2) * kHJ],
                                          double (&gs)[8][(TI + 2) * kHJ], int (&ws)[(TI + 2) * kHJ],
                                          double (&red)[TI + 2]) {
    constexpr int kTileI = TI, kTileThreads = (TI + 2) * 64;
    constexpr int kBI = TI + 4, kHI = TI + 2;

    const int pid = problem_of<V>(T, b);
    const TileProb P = T.prob[pid];
    const int t = b - P.tile0[V];
    const int ti = t / P.tiles_j, tj = t - ti * P.tiles_j;
    const int i0 = ti * kTileI, j0 = tj * kTileJ;
    const int tid = (int)threadIdx.x;
    const double dt = step_dt(ctl, pid);
    if (t == 0 && tid == 0) {
        ctl.dt_out[pid] = dt;
        ctl.dtmin[3 * pid + ctl.clr] = kInfBits;
    }

    // ---- state box: rows i0-2 .. i0+kTileI+1, cells j0-2 .. j0+kTileJ+1
    for (int k = tid; k < kBI * kBJ; k += kTileThreads) {
        const int li = k / kBJ, lj = k - li * kBJ;
        if (lj >= kTileJ + 4) continue;
        int gi = i0 - 2 + li, gj = j0 - 2 + lj;
        if (P.per[0]) gi = wrap(gi, P.nx);
        else if (gi < 0 || gi >= P.nx) continue;
        if (P.per[1]) gj = wrap(gj, P.ny);
        else if (gj < 0 || gj >= P.ny) continue;
        const int c = P.off + gi * P.ny + gj;
        zs[0][k] = in.r[c];
        zs[1][k] = in.u[c];
        zs[2][k] = in.v[c];
        zs[3][k] = in.p[c];
    }
    __syncthreads();

    // ---- gradient, limiting and predictor on the halo box
    const int r = tid >> 6, l = tid & 63;
    const int hidx = r * kHJ + l;            // this thread's cell in the halo box
    const int bidx = (r + 1) * kBJ + (l + 1);   // ... and in the state box
    {
        int gi = i0 - 1 + r, gj = j0 - 1 + l;
        bool valid = true;
        if (P.per[0]) gi = wrap(gi, P.nx);
        else valid = gi >= 0 && gi < P.nx;
        if (P.per[1]) gj = wrap(gj, P.ny);
        else valid = valid && gj >= 0 && gj < P.ny;
        if (valid) {
            const int c = P.off + gi * P.ny + gj;
            const Z zc{zs[0][bidx], zs[1][bidx], zs[2][bidx], zs[3][bidx], in.e[c]};
            const int wc = wait_in[c];
            const double ccx = T.col[P.col + gi].cx, ccy = T.row[P.row + gj].cy;
            // A slot's LSM data is rebuilt (from LDS and the tables) for
            // each of the two passes rather than held across them: the
            // same values, fewer live registers.
            auto slot = [&](int s) {
                const int ts = side_type(P, s, gi, gj);
                const int nb = uses_neighbour(P, s, ts) ? bidx + side_di(s) * kBJ + side_dj(s) : bidx;
                const Z zl{zs[0][nb], zs[1][nb], zs[2][nb], zs[3][nb], 0.0};
                return tile_slot(frame_of(s, false, 0.0, 0.0), ts, zc, zl, ccx, ccy, side_geom(T, P, s, gi, gj));
            };
            double A00 = 0.0, A01 = 0.0, A10 = 0.0, A11 = 0.0;
            double Fx[4] = {0.0, 0.0, 0.0, 0.0}, Fy[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
            for (int s = 0; s < 4; ++s) lsm_add(slot(s), A00, A01, A10, A11, Fx, Fy);
            if (fabs(A11) < 1e-8) A11 = 1.0;   // A(2,2) = 0 -> 1 always holds in 2D (gradient.h:268-273)
            const double det = A11 * A00 + (-A01) * A10;
            const double invdet = 1.0 / det;
            const double B00 = A11 * invdet, B01 = (-A01) * invdet;
            const double B10 = (-A10) * invdet, B11 = A00 * invdet;
            double gx[4], gy[4];
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                gx[k] = B00 * Fx[k] + B01 * Fy[k];
                gy[k] = B10 * Fx[k] + B11 * Fy[k];
            }
            double Lx[4] = {0.0, 0.0, 0.0, 0.0}, Ly[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
            for (int s = 0; s < 4; ++s) lim_add<LIM>(slot(s), gx, gy, Lx, Ly);
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                gx[k] = B00 * Lx[k] + B01 * Ly[k];
                gy[k] = B10 * Lx[k] + B11 * Ly[k];
                gs[k][hidx] = gx[k];
                gs[4 + k][hidx] = gy[k];
            }
            Z zh = zc;
            if (wc <= 0) {
                double F0 = 0.0, F1 = 0.0, F2 = 0.0, F3 = 0.0;
#pragma unroll
                for (int s = 0; s < 4; ++s) {
                    const SideGeom g = side_geom(T, P, s, gi, gj);
                    const Q f = pred_eval(frame_of(s, false, 0.0, 0.0), g.fx - ccx, g.fy - ccy, g.a, zc, gx, gy, eos);
                    F0 = F0 + f.r;
                    F1 = F1 + f.mx;
                    F2 = F2 + f.my;
                    F3 = F3 + f.E;
                }
                Q q = to_q(zc);
                const double coef = 0.5 * dt / T.vol[c];
                q.r = q.r - coef * F0;
                q.mx = q.mx - coef * F1;
                q.my = q.my - coef * F2;
                q.E = q.E - coef * F3;
                zh = to_z(q, eos);
                if (is_bad(zh, eos)) zh = zc;
            }
            hs[0][hidx] = zh.r;
            hs[1][hidx] = zh.u;
            hs[2][hidx] = zh.v;
            hs[3][hidx] = zh.p;
            hs[4][hidx] = zh.e;
            ws[hidx] = wc;
        }
    }
    __syncthreads();

    // ---- corrector on the tile's cells
    double cand = INFINITY;
    const int gi = i0 - 1 + r, gj = j0 - 1 + l;
    if (r >= 1 && r <= kTileI && l >= 1 && l <= kTileJ && gi < P.nx && gj < P.ny) {
        const int c = P.off + gi * P.ny + gj;
        const Z zch{hs[0][hidx], hs[1][hidx], hs[2][hidx], hs[3][hidx], hs[4][hidx]};
        const int wc = ws[hidx];
        const double ccx = T.col[P.col + gi].cx, ccy = T.row[P.row + gj].cy;
        double gx[4], gy[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            gx[k] = gs[k][hidx];
            gy[k] = gs[4 + k][hidx];
        }
        double F0 = 0.0, F1 = 0.0, F2 = 0.0, F3 = 0.0;
        bool bad = false;
        for (int s = 0; s < 4; ++s) {
            const int ts = side_type(P, s, gi, gj);
            const int nb = uses_neighbour(P, s, ts) ? hidx + side_di(s) * kHJ + side_dj(s) : hidx;
            const SideGeom g = side_geom(T, P, s, gi, gj);
            FaceC F;
            F.hn = Z{hs[0][nb], hs[1][nb], hs[2][nb], hs[3][nb], hs[4][nb]};
            F.wnb = ws[nb];
            F.ncx = g.ncx;
            F.ncy = g.ncy;
            F.fx = g.fx;
            F.fy = g.fy;
            F.a = g.a;
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                F.ngx[k] = gs[k][nb];
                F.ngy[k] = gs[4 + k][nb];
            }
            F.fr = frame_of(s, false, 0.0, 0.0);
            const Q f = corr_eval<3>(F, ts, zch, wc, ccx, ccy, gx, gy, eos, bad);
            F0 = F0 + f.r;
            F1 = F1 + f.mx;
            F2 = F2 + f.my;
            F3 = F3 + f.E;
        }
        const int bc = (r + 1) * kBJ + (l + 1);
        const Z zc{zs[0][bc], zs[1][bc], zs[2][bc], zs[3][bc], in.e[c]};
        Q q = to_q(zc);
        const double coef = dt / T.vol[c];
        q.r = q.r - coef * F0;
        q.mx = q.mx - coef * F1;
        q.my = q.my - coef * F2;
        q.E = q.E - coef * F3;
        Z zn = to_z(q, eos);
        int wn;
        if (is_bad(zn, eos)) {
            zn = zc;
            wn = 3;
            atomicOr(ctl.err, kErrIsolated);
        } else {
            wn = wc > 0 ? wc - 1 : wc;
        }
        store_z(out, c, zn);
        wait_out[c] = wn;
        if (bad) *ctl.err = 1;
        cand = cell_dt(zn, T.h[c], eos);
    }
    tile_min_atomic<TI + 2>(&ctl.dtmin[3 * pid + ctl.wr], cand, red);
}

// Workgroups of (TI + 2) wavefronts; WAVES = the resident waves per SIMD the
// register allocation is sized for.  One workgroup per tile, dealt to the
// XCDs in contiguous ranges of tiles (xcd_block).
template <int LIM, int V, int TI, int WAVES>
__global__ void __launch_bounds__((TI + 2) * 64) __attribute__((amdgpu_waves_per_eu(WAVES)))
k_tile_step(TileMesh T, DevState in, DevState out, const int* __restrict__ wait_in, int* __restrict__ wait_out,
            Eos eos, StepCtl ctl) {
    ZFV_TILE_LDS(TI);
    tile_body<LIM, V, TI>(T, in, out, wait_in, wait_out, eos, ctl, xcd_block((int)blockIdx.x, (int)gridDim.x), zs, hs,
                          gs, ws, red);
}

// The same step by persistent workgroups: the grid is the number of
// workgroups the GPU holds at once (a multiple of 8), and workgroup g, on
// XCD g % 8 (workgroups are dealt to the XCDs round-robin), walks the tiles
// of that XCD's contiguous range (the ranges of xcd_block) with a stride of
// the XCD's workgroup count.  Every workgroup leaves after its last tile.
template <int LIM, int V, int TI, int WAVES>
__global__ void __launch_bounds__((TI + 2) * 64) __attribute__((amdgpu_waves_per_eu(WAVES)))
k_tile_step_p(TileMesh T, DevState in, DevState out, const int* __restrict__ wait_in, int* __restrict__ wait_out,
              Eos eos, StepCtl ctl) {
    ZFV_TILE_LDS(TI);
    const int n = T.n_tiles[V];
    const int g = (int)blockIdx.x, x = g & 7, wgs = (int)gridDim.x >> 3;
    const int q = n >> 3, rem = n & 7;
    const int begin = x * q + (x < rem ? x : rem), end = begin + q + (x < rem ? 1 : 0);
    for (int b = begin + (g >> 3); b < end; b += wgs) {
        tile_body<LIM, V, TI>(T, in, out, wait_in, wait_out, eos, ctl, b, zs, hs, gs, ws, red);
        __syncthreads();   // the next tile's loads overwrite the LDS boxes
    }
}

// Workgroups the GPU holds at once for a persistent tiled-step kernel
// (a multiple of 8, at least 8).
template <int LIM, int V, int TI, int WAVES>
static hipError_t resident_grid(int* grid) {
    static int per_cu = 0;   // one value per kernel instance
    if (per_cu == 0) {
        int nb = 0;
        hipError_t e =
            hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_tile_step_p<LIM, V, TI, WAVES>, (TI + 2) * 64, 0);
        if (e != hipSuccess) return e;
        per_cu = nb > 0 ? nb : 1;
    }
    int dev = 0, cus = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e == hipSuccess) e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    if (e != hipSuccess) return e;
    *grid = std::max(8, (per_cu * cus) / 8 * 8);
    return hipSuccess;
}

template <int V, int TI, int WAVES>
static hipError_t launch_tile_t(int limiter, bool persistent, const TileMesh& T, const DevState& in,
                                const DevState& out, const int* wait_in, int* wait_out, const Eos& eos,
                                const StepCtl& ctl, hipStream_t s) {
    const dim3 b((TI + 2) * 64);
    dim3 g(T.n_tiles[V]);
#define ZFV_TILE(L)                                                                                              \
    do {                                                                                                         \
        if (persistent) {                                                                                        \
            int grid = 0;                                                                                        \
            const hipError_t e = resident_grid<L, V, TI, WAVES>(&grid);                                          \
            if (e != hipSuccess) return e;                                                                       \
            hipLaunchKernelGGL((k_tile_step_p<L, V, TI, WAVES>), dim3(grid), b, 0, s, T, in, out, wait_in,       \
                               wait_out, eos, ctl);                                                              \
        } else {                                                                                                 \
            hipLaunchKernelGGL((k_tile_step<L, V, TI, WAVES>), g, b, 0, s, T, in, out, wait_in, wait_out, eos,   \
                               ctl);                                                                             \
        }                                                                                                        \
    } while (0)
    switch (limiter) {
        case 0: ZFV_TILE(0); break;
        case 1: ZFV_TILE(1); break;
        case 2: ZFV_TILE(2); break;
        case 3: ZFV_TILE(3); break;
        case 4: ZFV_TILE(4); break;
        case 5: ZFV_TILE(5); break;
        case 6: ZFV_TILE(6); break;
        case 7: ZFV_TILE(7); break;
        case 8: ZFV_TILE(8); break;
        case 9: ZFV_TILE(9); break;
        case 10: ZFV_TILE(10); break;
        case 11: ZFV_TILE(11); break;
        default: ZFV_TILE(12); break;
    }
#undef ZFV_TILE
    return hipGetLastError();
}

// Tile shapes: 6 rows at 4 waves per SIMD (two workgroups per CU), 10 rows
// at 3 (one workgroup, no register spills), 14 rows at 4 (one workgroup).
hipError_t launch_tile_step(int limiter, int variant, bool persistent, const TileMesh& T, const DevState& in,
                            const DevState& out, const int* wait_in, int* wait_out, const Eos& eos,
                            const StepCtl& ctl, hipStream_t s) {
    if (variant < 0 || variant >= kTileVariants) return hipErrorInvalidValue;
    if (T.n_tiles[variant] <= 0) return hipSuccess;
    static_assert(kTileRows[0] == 6 && kTileRows[1] == 10 && kTileRows[2] == 14, "tile shapes");
    switch (variant) {
        case 0: return launch_tile_t<0, 6, 4>(limiter, persistent, T, in, out, wait_in, wait_out, eos, ctl, s);
        case 1: return launch_tile_t<1, 10, 3>(limiter, persistent, T, in, out, wait_in, wait_out, eos, ctl, s);
        default: return launch_tile_t<2, 14, 4>(limiter, persistent, T, in, out, wait_in, wait_out, eos, ctl, s);
    }
}

}  // namespace zfv
