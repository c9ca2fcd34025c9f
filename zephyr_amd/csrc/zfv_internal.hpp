// Internal device-data layout shared by the kernels and the C ABI.
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>

#include "zfv_physics.hpp"
#include "zfv_mm.hpp"

namespace zfv {

// Device mesh: cell arrays [cap], slot arrays slot-major [8][cap] (slot s
// of cell c at s*stride + c) so that a wavefront of 64 consecutive cells
// reads slot s of each cell from 64 consecutive addresses.  code[c] packs the
// 8 slot nibbles: bits 0-2 the slot type (SlotCode), bit 3 set when the
// slot's normal is not exactly the canonical side normal (AMR faces whose
// vertices carry round-off); such slots read nx/ny and use the general
// rotation.
struct DevMesh {
    int n;
    int stride;
    const double* __restrict__ cx;
    const double* __restrict__ cy;
    const double* __restrict__ vol;
    const double* __restrict__ h;       // incircle diameter
    const int* __restrict__ code;
    const int* __restrict__ adj;        // [8][stride]
    const double* __restrict__ area;    // [8][stride]
    const double* __restrict__ fcx;     // [8][stride]
    const double* __restrict__ fcy;     // [8][stride]
    const double* __restrict__ w;       // [8][stride] LSM weight area / |drn|^3
    const double* __restrict__ nx;      // [8][stride] outward normal
    const double* __restrict__ ny;
    // Axisymmetric step (SmFluid::set_axial): volume / volume_alt per cell,
    // with vol and area then volume_alt and area_alt; null for a planar step.
    const double* __restrict__ axc;
    // Batches of independent problems (null for one problem): problem of each
    // 64-cell chunk, first cell and end of each problem's cells.  Problems
    // start on 512-cell boundaries.
    const int* __restrict__ chunk_pid;
    const int* __restrict__ seg_begin;
    const int* __restrict__ seg_end;
};

constexpr int kChunkShift = 6;     // 64-cell chunks
constexpr int kSegAlign = 512;     // problem alignment in a batch (cells)
constexpr int kMaxBlock = 512;     // largest workgroup of the step kernels

constexpr int kGeneralNormal = 8;

struct DevState {
    double* r;
    double* u;
    double* v;
    double* p;
    double* e;
};

struct DevGrad {
    double* dx;   // [4][stride]: rho, u, v, p
    double* dy;
};

// Per-launch step control: dt slots hold min_c h/(|v|+c) as the bit pattern
// of a positive double (ordered like uint64), reduced with atomicMin; three
// rotating slots per problem (problem p at 3p), dt_out one per problem.
struct StepCtl {
    unsigned long long* dtmin;   // [3 * problems] rotating slots
    int rd, wr, clr;             // slot read for this step's dt, written for the next, cleared
    double cfl;
    const double* max_dt;        // device value, so captured step graphs follow set_max_dt
    double* dt_out;              // dt of this step
    int* err;                    // sticky numeric-error flag
};

// Mutable mesh view used by the AMR kernels (zfv_amr.hip): the DevMesh
// arrays plus the AMR bookkeeping of the reference's AmrCells
// (mesh/euler/amr_cells.h:74-104): per slot the Boundary id, per cell
// level / flag / b_idx / z_idx / next / idx (idx < 0: undefined cell) and the
// 9 SqQuad vertices ([9][stride]).  flag0 receives the restricted flags of
// the balancing's sort phase.
struct AmrMesh {
    int n;
    int stride;
    double *cx, *cy, *vol, *h;
    int *code, *adj, *bnd;
    double *area, *fcx, *fcy, *w, *nx, *ny;
    double *vx, *vy;
    int *level, *flag, *flag0, *b_idx, *z_idx, *next, *idx;
};

hipError_t amr_criterion(const AmrMesh& A, const DevState& z, const DevGrad& G, int criterion, double thr,
                         hipStream_t s);
// balance_flags: the restricted, balanced flags are left in A.flag0 (A.flag
// keeps the unrestricted ones); the caller swaps the two arrays.
hipError_t amr_balance(const AmrMesh& A, int max_level, hipStream_t s);
// Per-cell marks scanned by the adaptation: refine, coarsening-family head,
// flag != 0.
// Per-cell marks of an adaptation: refined, head of a coarsened family,
// changed; cc: the entries curve position i of the steps' order becomes
// (4 refined, 1 kept or family head, 0 other family member; 0 when the
// order is not carried).
struct AmrMarks {
    int ref, head, nz, cc;
};
// Exclusive scan of the marks (tile_sums: amr_scan3_tiles(count) entries);
// the total of all marks goes to `total` (device or pinned host memory).
int amr_scan3_tiles(int count);
hipError_t amr_scan3(const AmrMarks* in, AmrMarks* out, AmrMarks* tile_sums, int count, AmrMarks* total, hipStream_t s);
// perm (may be null): the steps' curve order of the cells, for cc.
hipError_t amr_mark(const AmrMesh& A, AmrMarks* marks, const int* perm, hipStream_t s);
// fam receives the four children of each coarsened family in z order.
hipError_t amr_positions(const AmrMesh& A, int n_large, int n_short, int n_parents, const AmrMarks* rank,
                         int* hole_at, int* ref_list, int* crs_list, int* fam, unsigned long long* faces,
                         unsigned long long* dt_rd, unsigned long long* dt_wr, hipStream_t s);
hipError_t amr_swapmap(const AmrMesh& A, int n_short, int n_large, int nz_total, const AmrMarks* rank,
                       const int* hole_at, int* swap_src, int* swap_dst, int* err, hipStream_t s);
// z.r null: mesh only (the solver's data are distributed and moved by its
// own kernels); fam: the coarsened families of amr_positions.  perm_new not
// null: the same launch also carries the steps' curve order (perm, of the
// A.n cells before) into perm_new / iperm_new, at the positions the scanned
// marks' cc give (rank).
hipError_t amr_geometry(const AmrMesh& A, const DevState& z, const DevGrad& G, int* wait, const Eos& eos,
                        bool slope, const int* ref_list, int n_refine, const int* fam, int n_parents,
                        const int* perm, const AmrMarks* rank, int* perm_new, int* iperm_new, hipStream_t s);
hipError_t amr_move(const AmrMesh& A, const DevState& z, int* wait, int n_short, int nz_total, const AmrMarks* rank,
                    int bound, const int* swap_src, const int* swap_dst, hipStream_t s);
hipError_t amr_derive(const AmrMesh& A, const DevState& z, const Eos& eos, unsigned long long* counts,
                      unsigned long long* dt_slot, int* err, hipStream_t s);

// Curve-ordered copy of the step kernels' mesh arrays (zfv_order.hip).
struct CurveMesh {
    int stride;
    double *cx, *cy, *vol, *h;
    int *code, *adj;
    double *area, *fcx, *fcy, *w, *nx, *ny;
};

// Keys of the space-filling-curve order (base cell, then Morton index at
// level lmax inside it) and the identity as sort values.
hipError_t curve_keys(const int* b_idx, const int* z_idx, const int* level, int n, int lmax,
                      unsigned long long* keys, int* vals, hipStream_t s);
// Dense-slot form of the same order (when n_base 4^lmax slots fit): every
// cell writes its index into its slot of an array that holds -1 elsewhere
// and counts itself in its 64-slot chunk (counts hold 0 before); the chunks
// are compacted in order into perm and iperm, and slots and counts are
// reset.  tile_sums: curve_scan_tiles(n_chunks) ints.
hipError_t curve_slots(const int* b_idx, const int* z_idx, const int* level, int n, int lmax, int* slots,
                       int* counts, long long n_slots, hipStream_t s);
int curve_scan_tiles(int n_chunks);
hipError_t curve_compact(int* slots, int n_chunks, int* counts, int* offsets, int* tile_sums, int* perm, int* iperm,
                         hipStream_t s);
hipError_t curve_inverse(const int* perm, int* iperm, int n, hipStream_t s);
// Radix sort of (key, value) pairs on bits [0, end_bit); tmp == nullptr
// sizes the scratch.
hipError_t curve_sort(const unsigned long long* keys_in, unsigned long long* keys_out, const int* vals_in,
                      int* vals_out, int n, int end_bit, void* tmp, size_t& tmp_bytes, hipStream_t s);
// The curve-ordered mesh copy (perm: curve -> mesh index, iperm: inverse);
// with zsrc.r not null the same launch gathers the state rows and wait
// counters into curve order (zdst, wdst).
hipError_t curve_mesh(const DevMesh& src, const CurveMesh& dst, const int* perm, const int* iperm,
                      const DevState& zsrc, const DevState& zdst, const int* wsrc, int* wdst, hipStream_t s);
// dst[i] = src[idx[i]]: the five state rows and the wait counters, or
// `rows` rows of a [rows][stride] array.
hipError_t gather_state(const DevState& src, const DevState& dst, const int* wsrc, int* wdst, const int* idx, int n,
                        hipStream_t s);
hipError_t gather_rows(const double* src, double* dst, const int* idx, int n, int rows, int stride, hipStream_t s);

// Launchers (zfv_kernels.hip).
hipError_t launch_fill_u64(unsigned long long* p, int count, unsigned long long v, hipStream_t s);
hipError_t launch_dt(const DevMesh& M, const DevState& z, const Eos& eos, unsigned long long* slot, hipStream_t s);

// Step-kernel layouts: one thread per cell, a lane quad per cell, or one
// thread per cell with the main slots' loads batched (more registers, fewer
// resident waves, fewer dependent memory round trips per cell).
constexpr int kLayoutCell = 0, kLayoutQuad = 1, kLayoutBatch = 2;
constexpr int kLayouts = 3;
// Stage 1 only: the whole MUSCL step in one tiled kernel (structured
// rectangles, HLLC); stage 2 is then not launched.
constexpr int kLayoutTile = 3;

hipError_t launch_flux1(int flux, int layout, int block, const DevMesh& M, const DevState& in, const DevState& out,
                        const Eos& eos, const StepCtl& ctl, hipStream_t s);
hipError_t launch_grad_pred(int limiter, bool predictor, int layout, int block, const DevMesh& M, const DevState& in,
                            const DevGrad& g, const DevState& half, const int* wait, const Eos& eos,
                            const StepCtl& ctl, hipStream_t s, double* raw = nullptr);
hipError_t launch_corr(int flux, int layout, int block, const DevMesh& M, const DevState& init, const DevState& half,
                       const DevGrad& g, const int* wait_in, int* wait_out, const DevState& next,
                       const Eos& eos, const StepCtl& ctl, hipStream_t s);
hipError_t launch_totals(const DevMesh& M, const DevState& z, double* partial, int nblocks, hipStream_t s);

// Structured rectangles (zfv_tile.hip): a Rectangle of nx x ny cells (cell
// off + i ny + j, i along x) whose slot arrays the host has checked against
// per-column and per-row tables, cell by cell, so that the tables give the
// same values the slot arrays hold.  The tiled MUSCL step reads the state,
// the isolation counters, volume and incircle per cell and the geometry from
// the tables.
struct TileCol {   // column i: centre x; face-centre x of slots L, R, B, T; weights of L, R; areas of B, T
    double cx, fxL, fxR, fxB, fxT, wL, wR, aB, aT;
};
struct TileRow {   // row j: centre y; face-centre y of slots L, R, B, T; weights of B, T; areas of L, R
    double cy, fyL, fyR, fyB, fyT, wB, wT, aL, aR;
};
// Tile shapes: kTileJ output cells along j (64-lane wavefronts with a halo
// lane on each side) by kTileRows[v] along i, with (rows + 2) wavefronts
// per workgroup (one per row of the tile and its two halo rows).
constexpr int kTileJ = 62;
constexpr int kTileVariants = 3;
constexpr int kTileRows[kTileVariants] = {6, 10, 14};
struct TileProb {
    int off, nx, ny;     // first cell, columns, rows
    int tiles_j;         // tiles along j
    int tile0[kTileVariants];   // first tile (workgroup) of the problem, per tile shape
    int col, row;        // first column / row table entry
    int t[4];            // slot code of the domain sides L, R, B, T (periodic, wall or copy)
    int per[2];          // x / y periodic (both sides periodic): edge neighbours wrap, else the cell itself
};
struct TileMesh {
    int n_probs;
    int n_tiles[kTileVariants];
    const TileProb* __restrict__ prob;
    const TileCol* __restrict__ col;
    const TileRow* __restrict__ row;
    const double* __restrict__ vol;
    const double* __restrict__ h;
};
// One MUSCL + HLLC step (gradient, limiting, predictor, corrector and the
// next step's CFL candidate) per tile, intermediate values in LDS.
// variant: the tile shape (kTileRows); workgroups of (kTileRows[v] + 2) * 64 threads.
hipError_t launch_tile_step(int limiter, int variant, bool persistent, const TileMesh& T, const DevState& in,
                            const DevState& out, const int* wait_in, int* wait_out, const Eos& eos,
                            const StepCtl& ctl, hipStream_t s);
// Scratch of the no-allocation entry points (zfv_kernels.hip): 4 result
// doubles, then up to kScratchBlocks workgroups' 4 partial sums (4 MiB).
constexpr int kScratchBlocks = 131072;
constexpr int kScratchDoubles = 4 + 4 * kScratchBlocks;
hipError_t scratch_address(double** out);
hipError_t launch_totals_sum(const double* partial, int nblocks, double* out, hipStream_t s);

// Discrete adjoint of the MUSCL + HLLC step (zfv_adjoint.hip).
// One reverse step: z the step's input, znext its output, half / G its half
// state and gradient, slot its dt-minimum slot; lam_in = dJ/dznext ->
// lam_out = dJ/dz.  W [4][stride], ZB [20][stride], NG [16][stride], lz and
// lG are work buffers.
hipError_t adj_step(int limiter, const DevMesh& M, const DevState& z, const DevState& znext, const DevState& half,
                    const DevGrad& G, const int* wait, const Eos& eos, const unsigned long long* slot, double cfl,
                    const double* max_dt, const DevState& lam_in, const DevState& lam_out, const DevState& lz,
                    double* W, const DevGrad& lG, double* ZB, double* NG, double* dtb, int* err, hipStream_t s);
hipError_t adj_seed(const DevMesh& M, const DevState& z, const double* win, const DevState& lam, hipStream_t s);

constexpr int kBlock = 256;

// Multi-material state layer (zfv_mm.hip): per cell rho, u, v, P, e, T, the
// cached sound speed c, and per material [mat][stride] mass fractions b and
// true densities rh (NaN for an absent material).
struct MmLayer {
    double *r, *u, *v, *P, *e, *T, *c, *b, *rh;
};
// Options of one multi-material sweep: the flux of the plain faces (the
// reference's Fluxes numbering: 1 Rusanov, 2 HLL, 3 HLLC), the CrpMode (0
// NONE, 1 V3, 2 V5, 3 VS, 4 PLIC), the faces swept (EuCell::faces(dir):
// 0 all, 1 |n.x| >= 0.7, 2 |n.y| >= 0.7) and the factor on the step's dt
// (0.5 for STRANG's X sweeps).
struct MmOpts {
    int flux = 3, crp = 4, dir = 0;
    double dt_scale = 1.0;
};
hipError_t mm_prepare(const DevMesh& M, const MmLayer& L, const mm::Mix& mx, unsigned long long* slot,
                      hipStream_t s);
// CF: composite Riemann fluxes per [slot][4 + n_mat][stride] from mm_crp,
// or null without composite Riemann problems.  pass 0: every cell; pass 1:
// the cells without a composite face (CF is not read: mm_crp may still be
// writing it on another stream); pass 2: the cells listed by mm_crp_mark
// (cells, count[1]), which pass 1 left.
hipError_t mm_flux1(const DevMesh& M, const MmLayer& in, const MmLayer& out, const mm::Mix& mx,
                    const StepCtl& ctl, const MmOpts& o, const double* CF, int pass, const int* cells,
                    const unsigned int* count, hipStream_t s);
// The faces that take composite Riemann fluxes: list (8 * stride ints of
// work space) and their number (count[0]), with `cells` (stride ints, or
// null) the cells that have such a face and their number (count[1]); then
// the fluxes into CF, over `groups` one-wavefront workgroups.  acc 1 reads
// the cells of `in`; acc 2 the face states extrapolated from `half` with
// the gradients gx / gy ([6 + 2 n_mat][stride], k_mm_grad).
hipError_t mm_crp_mark(const DevMesh& M, const MmLayer& in, const MmLayer& half, const mm::Mix& mx, int acc,
                       const MmOpts& o, const double* gx, const double* gy, int* list, int* cells,
                       unsigned int* count, hipStream_t s);
hipError_t mm_crp(const DevMesh& M, const MmLayer& in, const MmLayer& half, const mm::Mix& mx, int acc,
                  const MmOpts& o, const double* gx, const double* gy, const StepCtl& ctl, const double* nx,
                  const double* ny, const int* list, const unsigned int* count, double* CF, int groups,
                  hipStream_t s);
// Second order: the predictor into `half`, then the corrector into `out`.
hipError_t mm_stage1(const DevMesh& M, const MmLayer& in, const MmLayer& half, const mm::Mix& mx, const StepCtl& ctl,
                     const MmOpts& o, const double* gx, const double* gy, hipStream_t s);
hipError_t mm_flux2(const DevMesh& M, const MmLayer& in, const MmLayer& half, const MmLayer& out, const mm::Mix& mx,
                    const StepCtl& ctl, const MmOpts& o, const double* gx, const double* gy, const double* CF,
                    int pass, const int* cells, const unsigned int* count, hipStream_t s);
hipError_t mm_interface(const DevMesh& M, const MmLayer& L, const mm::Mix& mx, const double* vx, const double* vy,
                        double* nx, double* ny, double* p, hipStream_t s);
// Adaptation of the multi-material state: the limited gradient (rows rho,
// u, v, e, beta[n_mat] of gx / gy, [4 + n_mat][stride]), the refinement
// flags, the split / merge distributor (fam: the four children of each
// coarsened family), the moves of the state rows, and the cached sound
// speed with the next step's dt candidate.
// compute_grad of the first `rows` components (rho, u, v, e, beta[n_mat],
// P, T, rhos[n_mat]) into gx / gy ([6 + 2 n_mat][stride]).
hipError_t mm_grad(const DevMesh& M, const MmLayer& L, const mm::Mix& mx, int crp, int limiter, int rows, double* gx,
                   double* gy, hipStream_t s);
hipError_t mm_flags(const DevMesh& M, const MmLayer& L, const mm::Mix& mx, int* flag, hipStream_t s);
hipError_t mm_distribute(const AmrMesh& A, const MmLayer& L, const mm::Mix& mx, const double* gx, const double* gy,
                         const int* ref_list, int n_refine, const int* fam, int n_parents, int* err, hipStream_t s);
hipError_t mm_move(const MmLayer& L, const mm::Mix& mx, int stride, int n, int n_short, int nz_total,
                   const AmrMarks* rank, int bound, const int* swap_src, const int* swap_dst, hipStream_t s);
hipError_t mm_cache(const DevMesh& M, const MmLayer& L, const mm::Mix& mx, unsigned long long* slot,
                    hipStream_t s);

// Transfer (zfv_transfer.hip): the slopes (MUSCL methods) and the fluxes
// of one update from u into out (mesh order), the interface normals, the
// CFL candidates' minimum into *slot, set_flags into flag.
bool tr_method_built(int method);
hipError_t tr_step(const DevMesh& M, int method, const double* u, double* out, const double* tnx, const double* tny,
                   double* gx, double* gy, const double* fvx, const double* fvy, double dt, int dir, int gnx,
                   int gny, hipStream_t s);
// WENO / WENO_CRP: five-cell stencils on a Rectangle's (i, j) grid.
bool tr_method_structured(int method);
hipError_t tr_normals(const DevMesh& M, const double* u, double* nx, double* ny, hipStream_t s);
hipError_t tr_dt(const DevMesh& M, const double* cvx, const double* cvy, unsigned long long* slot, hipStream_t s);
hipError_t tr_flags(const DevMesh& M, const double* u, int* flag, hipStream_t s);

// Error-word bits: 1 numeric, 2 inconsistent mesh (errors); kErrRobust and
// kErrIsolated are notices for the adjoint (a cell in the robust isolation
// mode, a cell isolated by the corrector's bad-state fallback).
constexpr int kErrRobust = 4;
constexpr int kErrIsolated = 8;

}  // namespace zfv
