// Device-side AMR-2D adaptation for gfx950: refinement flags, flag
// balancing, and the apply phase (positions, child/parent geometry, face
// split/merge, neighbour re-linking, compaction), mirroring zephyr's
// EuMesh::refine (mesh/euler/eu_mesh.cpp:290-359) on one process.
//
// The reference's serial parts become scans: the counters of
// setup_positions (mesh/amr/setup_positions.h:24-64) are exclusive scans of
// the refine / coarsen marks, and the SwapLists (swap_lists.h:28-58) pair
// the k-th hole (ascending) with the k-th surviving cell above n_short
// (descending), both found by scans.  The geometry work of the reference's
// parallel setup_geometry runs as three kernels over compacted lists:
// retained cells (one thread each), refined cells (one thread per child) and
// coarsened families (one thread per family); each thread writes only the
// slots of its own cell and the cells it creates.  The host synchronises once
// per adaptation, to read the new cell count.
// Single process, Cartesian base: every adjacency rotation is 0.
#include "zfv_internal.hpp"

#include "zfv_scan.hpp"

namespace zfv {

namespace {

// AMR-2D index tables (geom/indexing.h).
__constant__ int cSF[4][2] = {{0, 6}, {2, 8}, {0, 2}, {6, 8}};
__constant__ int cCF[8][2] = {{0, 3}, {2, 5}, {0, 1}, {6, 7}, {3, 6}, {5, 8}, {1, 2}, {7, 8}};
__constant__ int cChildBySubface[8] = {0, 1, 0, 2, 2, 3, 1, 3};
__constant__ int cSubfaceBySideChild[4][4] = {{0, -1, 4, -1}, {-1, 1, -1, 5}, {2, 6, -1, -1}, {-1, -1, 3, 7}};
__constant__ int cNeibChild0[8] = {1, 0, 2, 0, 3, 2, 3, 1};
__constant__ int cNextSiblingSide[4] = {1, 3, 2, 0};
__constant__ int cCorners[4][4] = {{0, 1, 3, 4}, {1, 2, 4, 5}, {3, 4, 6, 7}, {4, 5, 7, 8}};
__constant__ int cLinks[8][3] = {{0, 1, 1}, {0, 3, 2}, {1, 0, 0}, {1, 3, 3},
                                 {2, 1, 3}, {2, 2, 0}, {3, 0, 2}, {3, 2, 1}};

constexpr int B_UNDEF = -1, B_INNER = 0, B_PERIODIC = 3;

__device__ __forceinline__ bool is_bnd(int b) { return b != B_INNER && b != B_PERIODIC && b != B_UNDEF; }

__device__ __forceinline__ int adjacent_child(int side, int child) {
    return cNeibChild0[cSubfaceBySideChild[side][child]];
}

struct V2 {
    double x, y;
};

__device__ __forceinline__ int SI(const AmrMesh& A, int c, int s) { return s * A.stride + c; }
__device__ __forceinline__ V2 vtx(const AmrMesh& A, int c, int k) {
    return V2{A.vx[k * A.stride + c], A.vy[k * A.stride + c]};
}

// Line::length / centroid / normal(view) (geom/primitives/line.cpp:18-71) for
// z = 0: tau x (v0 - view) = (0, 0, cz), (0, 0, cz) x tau = (-cz ty, cz tx, 0).
__device__ void line_geom(V2 v0, V2 v1, V2 view, double& area, double& fx, double& fy, double& nx, double& ny) {
    const double tx = v1.x - v0.x, ty = v1.y - v0.y;
    area = sqrt(tx * tx + ty * ty);
    fx = 0.5 * (v0.x + v1.x);
    fy = 0.5 * (v0.y + v1.y);
    const double ax = v0.x - view.x, ay = v0.y - view.y;
    const double cz = tx * ay - ty * ax;
    const double px = 0.0 - cz * ty, py = cz * tx - 0.0;
    const double z = px * px + py * py;
    if (z > 0.0) {
        const double r = sqrt(z);
        nx = px / r;
        ny = py / r;
    } else {
        nx = px;
        ny = py;
    }
}

// AmrCells::set_cell(ic, Quad, axial = false) (amr_cells.cpp:725-785),
// computed in registers: area, centroid, the SqQuad's nine vertices and the
// four sides' face geometry (normals seen from the centroid).
struct QuadG {
    double vol;
    V2 C;
    V2 V[9];
    double area[4], fx[4], fy[4], nx[4], ny[4];
};

__device__ void quad_geom(V2 v00, V2 v01, V2 v10, V2 v11, QuadG& G) {
    // Quad::area: 0.5 |(v11 - v00) x (v10 - v01)| (quad.cpp:322-324)
    const double ax = v11.x - v00.x, ay = v11.y - v00.y, bx = v10.x - v01.x, by = v10.y - v01.y;
    const double cz = ax * by - ay * bx;
    const double area = 0.5 * sqrt(cz * cz);
    G.vol = area;
    // Quad::centroid_oxy over vertices 0, 1, 3, 2 (quad.cpp:358-377)
    const V2 q[4] = {v00, v01, v11, v10};
    double Cx = 0.0, Cy = 0.0;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const V2 a = q[i], b = q[(i + 1) & 3];
        Cx = Cx + (b.y - a.y) * (b.x * b.x + b.x * a.x + a.x * a.x);
        Cy = Cy - (b.x - a.x) * (b.y * b.y + b.y * a.y + a.y * a.y);
    }
    const double d6 = 6.0 * area;
    G.C = V2{Cx / d6, Cy / d6};
    // SqQuad(const Quad&): bilinear vertices (quad.cpp:452-457)
    G.V[0] = v00;
    G.V[1] = V2{0.5 * (v00.x + v01.x), 0.5 * (v00.y + v01.y)};
    G.V[2] = v01;
    G.V[3] = V2{0.5 * (v00.x + v10.x), 0.5 * (v00.y + v10.y)};
    G.V[4] = V2{0.25 * (((v00.x + v01.x) + v10.x) + v11.x), 0.25 * (((v00.y + v01.y) + v10.y) + v11.y)};
    G.V[5] = V2{0.5 * (v01.x + v11.x), 0.5 * (v01.y + v11.y)};
    G.V[6] = v10;
    G.V[7] = V2{0.5 * (v10.x + v11.x), 0.5 * (v10.y + v11.y)};
    G.V[8] = v11;
#pragma unroll
    for (int side = 0; side < 4; ++side)
        line_geom(G.V[cSF[side][0]], G.V[cSF[side][1]], G.C, G.area[side], G.fx[side], G.fy[side], G.nx[side],
                  G.ny[side]);
}

// The cell arrays and vertices of a new cell ic (index ic, flag 0, the given
// tree indices and level).
__device__ void write_cell(const AmrMesh& A, int ic, const QuadG& G, int b_idx, int level, int z_idx) {
    A.idx[ic] = ic;
    A.flag[ic] = 0;
    A.b_idx[ic] = b_idx;
    A.level[ic] = level;
    A.z_idx[ic] = z_idx;
    A.vol[ic] = G.vol;
    A.cx[ic] = G.C.x;
    A.cy[ic] = G.C.y;
#pragma unroll
    for (int k = 0; k < 9; ++k) {
        A.vx[k * A.stride + ic] = G.V[k].x;
        A.vy[k * A.stride + ic] = G.V[k].y;
    }
}

// The eight slots of a new cell: types b, neighbours a; the main sides get
// set_cell's face geometry, except the sides split into two sub-faces
// (split bit per side), which get setup_face_geom's halves with the SqQuad
// centre as view (faces.h:57-97, 121-130).
__device__ void write_slots(const AmrMesh& A, int ic, const QuadG& G, const int* b, const int* a, unsigned split) {
#pragma unroll
    for (int s = 0; s < 8; ++s) {
        A.bnd[SI(A, ic, s)] = b[s];
        A.adj[SI(A, ic, s)] = a[s];
    }
#pragma unroll
    for (int side = 0; side < 4; ++side) {
        const int f0 = SI(A, ic, side);
        if (split & (1u << side)) {
            const int f1 = SI(A, ic, side + 4);
            line_geom(G.V[cCF[side][0]], G.V[cCF[side][1]], G.V[4], A.area[f0], A.fcx[f0], A.fcy[f0], A.nx[f0],
                      A.ny[f0]);
            line_geom(G.V[cCF[side + 4][0]], G.V[cCF[side + 4][1]], G.V[4], A.area[f1], A.fcx[f1], A.fcy[f1],
                      A.nx[f1], A.ny[f1]);
        } else {
            A.area[f0] = G.area[side];
            A.fcx[f0] = G.fx[side];
            A.fcy[f0] = G.fy[side];
            A.nx[f0] = G.nx[side];
            A.ny[f0] = G.ny[side];
        }
    }
}

// can_coarse / get_siblings (mesh/amr/siblings.h:36-131), flags read from `flag`.
__device__ bool can_coarse(const AmrMesh& A, const int* flag, int ic) {
    if (flag[ic] >= 0) return false;
    for (int i = 0; i < 3; ++i) {
        const int jc = A.adj[SI(A, ic, cNextSiblingSide[A.z_idx[ic] & 3])];
        if (A.level[ic] != A.level[jc]) return false;
        if (flag[jc] >= 0) return false;
        ic = jc;
    }
    return true;
}

__device__ void get_siblings(const AmrMesh& A, int ic, int* sibs) {
    int jc = ic;
    for (int i = 0; i < 3; ++i) {
        jc = A.adj[SI(A, jc, cNextSiblingSide[A.z_idx[jc] & 3])];
        sibs[i] = jc;
    }
}

__device__ __forceinline__ int cell_id() { return (int)(blockIdx.x * blockDim.x + threadIdx.x); }

}  // namespace

// ---------------------------------------------------------------- flags
// Refinement criterion.  0: SmFluid::set_flags relative jumps (sm_fluid.cpp:517-643);
// 1: |grad rho| h / rho against `threshold`, G holding the unlimited LSM
// gradient of rho (G.dx[c], G.dy[c]).
__global__ void __launch_bounds__(256) k_amr_criterion(AmrMesh A, DevState z, DevGrad G, int criterion,
                                                       double threshold) {
    const int ic = cell_id();
    if (ic >= A.n) return;
    const double dens = z.r[ic], pres = z.p[ic];
    if (criterion == 1) {
        const double gx = G.dx[ic], gy = G.dy[ic];
        const double xi = sqrt(gx * gx + gy * gy) * A.h[ic] / dens;
        A.flag[ic] = xi > threshold ? 1 : (xi > 0.4 * threshold ? 0 : -1);
        return;
    }
    int flag = -1;
    const double ds = 0.05 * fabs(dens), ps = 0.05 * fabs(pres);
    for (int s = 0; s < 8; ++s) {
        const int b = A.bnd[SI(A, ic, s)];
        if (b == B_UNDEF || is_bnd(b)) continue;
        const int jc = A.adj[SI(A, ic, s)];
        const double dd = fabs(z.r[jc] - dens), dp = fabs(z.p[jc] - pres);
        if (dd > ds || dp > ps) {
            flag = 1;
            break;
        }
        if (dd > 0.4 * ds || dp > 0.4 * ps) flag = 0;
    }
    A.flag[ic] = flag;
}

// base_restriction (balancing_restrictions.h:21-47): reads the unrestricted
// flags (can_coarse looks at the siblings' ones), writes the restricted ones
// to flag0; the balancing rounds then work on flag0 in place, and the caller
// swaps flag and flag0.
__global__ void __launch_bounds__(256) k_amr_restrict(AmrMesh A, int max_level) {
    const int ic = cell_id();
    if (ic >= A.n) return;
    int flag = A.flag[ic];
    if (flag != 0) flag = flag > 0 ? 1 : -1;
    const int lvl = A.level[ic];
    if (lvl + flag < 0) flag = 0;
    else if (lvl + flag > max_level) flag = lvl > max_level ? -1 : 0;
    if (flag < 0 && !can_coarse(A, A.flag, ic)) flag = 0;
    A.flag0[ic] = flag;
}

// Largest wanted level (level + flag) over the face neighbours of cell ic
// (inner and periodic slots); every adjacency load is issued before any
// neighbour's level and flag, so a cell costs two dependent round trips.
__device__ __forceinline__ int max_wanted(const AmrMesh& A, const int* flag, int ic, int code) {
    int jc[8];
#pragma unroll
    for (int s = 0; s < 8; ++s) {
        const int t = (code >> (4 * s)) & 7;
        jc[s] = (t == SC_INNER || t == SC_PERIODIC) ? A.adj[SI(A, ic, s)] : -1;
    }
    int w = -1;
#pragma unroll
    for (int s = 0; s < 8; ++s)
        if (jc[s] >= 0) w = max(w, A.level[jc[s]] + flag[jc[s]]);
    return w;
}

// The balancing rounds of one level (balancing_fast.h:273-390) on the
// restricted flags, in place.  The reference runs round 1 on the cells that
// kept 0 (one becomes 1 if a neighbour wants more than one level finer),
// rounds 2 and 3 on the cells restricted to -1 (1 as in round 1, else 0 if
// a neighbour wants a finer level than the cell's), round 4 on the cells
// still at -1 (0 if a sibling does not coarsen).  PHASE 0 runs rounds 1 and
// 2 together, PHASE 1 round 3, PHASE 2 round 4.  The flags at the end equal
// the reference's: a 1 depends only on finer neighbours, whose flags are
// final before this level's rounds; a 0 of rounds 2-3 depends also on
// same-level neighbours turning 1 during rounds 1-2, and round 3 reads them
// all after the launch that set them, in both orders; a cell's own flag
// changes only in its own thread (0 -> 1, -1 -> 0 or 1), so its value at
// the start of a launch tells which round applies, as the snapshot did.
template <int PHASE>
__global__ void __launch_bounds__(256) k_amr_round(AmrMesh A, int lvl) {
    const int ic = cell_id();
    if (ic >= A.n) return;
    int* flag = A.flag0;
    const int L = A.level[ic], f = flag[ic], code = A.code[ic];
    if (L != lvl) return;
    if (PHASE == 2) {
        if (f != -1) return;
        int sibs[3];
        get_siblings(A, ic, sibs);
        if (flag[sibs[0]] >= 0 || flag[sibs[1]] >= 0 || flag[sibs[2]] >= 0) flag[ic] = 0;
        return;
    }
    if (PHASE == 0 ? (f != 0 && f != -1) : f != -1) return;
    const int w = max_wanted(A, flag, ic, code);
    if (w > lvl + 1) flag[ic] = 1;
    else if (f == -1 && w > lvl) flag[ic] = 0;
}

// Statistics (statistics.h:59-81), setup_positions (setup_positions.h:24-64)
// and the SwapLists (swap_lists.h:28-58) all come from one exclusive scan of
// per-cell marks {refine, coarsening-family head, flag != 0}: its tail gives
// n_refine, n_parents and the number of changed cells; the ranks give the
// positions of children and parents, the rank of each hole (a changed cell
// below n_short) and, in closed form, the number of surviving positions
// above each position >= n_short.
// cc (perm not null): the entries the cell at curve position ic becomes in
// the carried curve order (see k_amr_geometry).
__global__ void __launch_bounds__(256) k_amr_mark(AmrMesh A, AmrMarks* marks, const int* __restrict__ perm) {
    const int ic = cell_id();
    if (ic >= A.n) return;
    const int f = A.flag[ic];
    int cc = 0;
    if (perm) {
        const int o = perm[ic];
        const int fo = A.flag[o];
        cc = fo > 0 ? 4 : (fo < 0 ? ((A.z_idx[o] & 3) == 0 ? 1 : 0) : 1);
    }
    marks[ic] = AmrMarks{f > 0 ? 1 : 0, (f < 0 && (A.z_idx[ic] & 3) == 0) ? 1 : 0, f != 0 ? 1 : 0, cc};
}

// Positions: retained cells keep their index, parents go to n + rank,
// children to n + n_parents + 4 rank.  Also the compacted refine / coarsen
// lists, the four children of each coarsened family in z order (fam) and the
// hole table (k-th hole ascending).  Thread 0 clears the
// face counters and the dt slots the derive pass fills.
__global__ void __launch_bounds__(256) k_amr_positions(AmrMesh A, int n_large, int n_short, int n_parents,
                                                       const AmrMarks* rank, int* hole_at, int* ref_list,
                                                       int* crs_list, int* fam, unsigned long long* faces,
                                                       unsigned long long* dt_rd, unsigned long long* dt_wr) {
    const int ic = cell_id();
    const int n = A.n;
    if (ic == 0) {
        faces[0] = 0ULL;
        faces[1] = 0ULL;
        *dt_rd = 0x7ff0000000000000ULL;
        *dt_wr = 0x7ff0000000000000ULL;
    }
    if (ic >= n_large) return;
    if (ic >= n) {
        A.next[ic] = ic;
        return;
    }
    const int f = A.flag[ic];
    const AmrMarks rk = rank[ic];
    if (f == 0) {
        A.next[ic] = ic;
    } else if (f > 0) {
        A.next[ic] = n + n_parents + 4 * rk.ref;
        ref_list[rk.ref] = ic;
    } else if ((A.z_idx[ic] & 3) == 0) {
        const int pos = n + rk.head;
        int sibs[3];
        get_siblings(A, ic, sibs);
        A.next[ic] = pos;
        for (int k = 0; k < 3; ++k) A.next[sibs[k]] = pos;
        crs_list[rk.head] = ic;
        fam[4 * rk.head] = ic;
        for (int k = 0; k < 3; ++k) fam[4 * rk.head + (A.z_idx[sibs[k]] & 3)] = sibs[k];
    }
    if (ic < n_short && f != 0) hole_at[rk.nz] = ic;
}

// SwapLists: the survivor at position p >= n_short (a new cell, or an old
// cell kept as it is) with r survivors above it fills hole r.  For p >= n
// every position survives; for p < n the survivors above p are the new
// cells plus the retained cells in (p, n).  n_swap = number of holes.
__global__ void __launch_bounds__(256) k_amr_swapmap(AmrMesh A, int n_short, int n_large, int nz_total,
                                                     const AmrMarks* rank, const int* hole_at, int* swap_src,
                                                     int* swap_dst, int* err) {
    const int p = n_short + cell_id();
    if (p >= n_large) return;
    const int n = A.n;
    const int n_swap = n_short >= n ? nz_total : rank[n_short].nz;
    int r;
    if (p >= n) {
        r = n_large - 1 - p;
    } else {
        if (A.flag[p] != 0) return;   // a changed cell leaves its position
        r = (n_large - n) + (n - 1 - p) - (nz_total - rank[p].nz);
    }
    if (r < 0 || r >= n_swap) {
        atomicOr(err, 2);
        return;
    }
    const int h = hole_at[r];
    A.next[p] = h;
    swap_src[r] = p;
    swap_dst[r] = h;
}

// ---------------------------------------------------------------- geometry
namespace {

// setup_face_geom with view = SqQuad centre (faces.h:57-97), from the
// cell's vertices held in registers.
__device__ __forceinline__ void face_geom_v(const AmrMesh& A, int ic, int slot, const V2* V, int a, int b) {
    double area, fx, fy, nx, ny;
    line_geom(V[a], V[b], V[4], area, fx, fy, nx, ny);
    const int f = SI(A, ic, slot);
    A.area[f] = area;
    A.fcx[f] = fx;
    A.fcy[f] = fy;
    A.nx[f] = nx;
    A.ny[f] = ny;
}

// retain_cell (retain.h:17-216).  The loads come in three dependent batches
// (the cell's slots; the neighbours' level, flag and next; the second
// `next` lookups), all issued before the first store: the slot arrays are
// written with the types of the arrays the lookups read, so stores between
// them would serialise every side's chain.
__device__ void retain_cell(const AmrMesh& A, int ic) {
    enum { K_NONE, K_SET, K_SPLIT, K_MERGE, K_SUBS };
    const int lvl_c = A.level[ic], z3 = A.z_idx[ic] & 3, next_c = A.next[ic];
    int b0[4], b1[4], a0[4], a1[4];
#pragma unroll
    for (int side = 0; side < 4; ++side) {
        b0[side] = A.bnd[SI(A, ic, side)];
        b1[side] = A.bnd[SI(A, ic, side + 4)];
        a0[side] = A.adj[SI(A, ic, side)];
        a1[side] = A.adj[SI(A, ic, side + 4)];
    }
    int ln[4], fl[4], nn[4];
#pragma unroll
    for (int side = 0; side < 4; ++side) {
        const bool inner = b0[side] != B_UNDEF && !is_bnd(b0[side]);
        const int jc = a0[side];
        ln[side] = inner ? A.level[jc] : 0;
        fl[side] = inner ? A.flag[jc] : 0;
        nn[side] = inner ? A.next[jc] : 0;
    }
    int kind[4], r0[4], r1[4];
    bool geo = false;
#pragma unroll
    for (int side = 0; side < 4; ++side) {
        kind[side] = K_NONE;
        r0[side] = r1[side] = -1;
        if (b0[side] == B_UNDEF || is_bnd(b0[side])) continue;
        if (lvl_c == ln[side]) {
            if (fl[side] > 0) {
                kind[side] = K_SPLIT;
                r0[side] = A.next[nn[side] + cNeibChild0[side]];
                r1[side] = A.next[nn[side] + cNeibChild0[side + 4]];
            } else {
                kind[side] = K_SET;
                r0[side] = fl[side] == 0 ? nn[side] : A.next[nn[side]];
            }
        } else if (lvl_c < ln[side]) {
            if (fl[side] < 0) {
                kind[side] = K_MERGE;
                r0[side] = A.next[nn[side]];
            } else {
                kind[side] = K_SUBS;
                r0[side] = A.next[a0[side]];
                if (b1[side] != B_UNDEF) r1[side] = A.next[a1[side]];
            }
        } else {
            kind[side] = K_SET;
            r0[side] = fl[side] == 0 ? nn[side] : A.next[nn[side] + adjacent_child(side, z3)];
        }
        geo = geo || kind[side] == K_SPLIT || kind[side] == K_MERGE;
    }
    V2 V[9];
    if (geo) {
#pragma unroll
        for (int k = 0; k < 9; ++k) V[k] = vtx(A, ic, k);
    }
#pragma unroll
    for (int side = 0; side < 4; ++side) {
        const int f0 = SI(A, ic, side), f1 = SI(A, ic, side + 4);
        if (b0[side] == B_UNDEF) continue;
        if (is_bnd(b0[side])) {
            A.adj[f0] = next_c;
            continue;
        }
        switch (kind[side]) {
            case K_SPLIT:   // split_face (faces.h:121-130), then the two sub-neighbours
                A.bnd[f1] = b0[side];
                face_geom_v(A, ic, side, V, cCF[side][0], cCF[side][1]);
                face_geom_v(A, ic, side + 4, V, cCF[side + 4][0], cCF[side + 4][1]);
                A.adj[f0] = r0[side];
                A.adj[f1] = r1[side];
                break;
            case K_MERGE:   // merge_faces (faces.h:137-154)
                A.bnd[f1] = B_UNDEF;
                A.adj[f1] = -1;
                face_geom_v(A, ic, side, V, cSF[side][0], cSF[side][1]);
                A.adj[f0] = r0[side];
                break;
            case K_SUBS:
                A.adj[f0] = r0[side];
                if (b1[side] != B_UNDEF) A.adj[f1] = r1[side];
                break;
            default:
                A.adj[f0] = r0[side];
                break;
        }
    }
}

// SmFluid distributor split "slope" (sm_fluid.cpp:395-458) for a child
// whose centroid is (drx, dry) from the parent's, from the parent's state zp
// and limited gradient (gx, gy of rho, u, v, p): false if the reconstructed
// state is bad.
__device__ bool split_slope(const Z& zp, const double* gx, const double* gy, double drx, double dry,
                            const Eos& eos, Z& q) {
    const double dR = (eos.g - 1.0) * (zp.e - eos.e0);
    const double dE = (eos.g - 1.0) * zp.r;
    const double gex = (gx[3] - dR * gx[0]) / dE, gey = (gy[3] - dR * gy[0]) / dE;
    q = zp;
    q.r = zp.r + gx[0] * drx + gy[0] * dry;
    const double ratio = zp.r / q.r;
    q.u = zp.u + ratio * (gx[1] * drx + gy[1] * dry);
    q.v = zp.v + ratio * (gx[2] * drx + gy[2] * dry);
    const double du = zp.u - q.u, dv = zp.v - q.v;
    q.e = zp.e - 0.5 * (du * du + dv * dv) + ratio * (gex * drx + gey * dry);
    q.p = eos_p(eos, q.r, q.e);
    return !is_bad(q, eos);
}

// One child of refine_cell / make_children / link_siblings (refine.h:36-444):
// child i of parent ip, written only into the child's own cell and slots,
// plus (z.r not null) the child's state by the const or slope distributor
// (sm_fluid.cpp:372-458), decided for all four children together (lanes
// 4k..4k+3 of a wavefront hold one family).  The loads come in three
// dependent batches before the first store: the parent's data and slots;
// the new positions and the outer neighbours' level, flag and position;
// then the positions of those neighbours' children or parents.
__device__ void refine_one(const AmrMesh& A, const DevState& z, const DevGrad& Gd, int* wait, const Eos& eos,
                           bool slope, int ip, int i) {
    // batch 1: the parent
    const int mc = A.next[ip];
    const int lvl = A.level[ip], zi = A.z_idx[ip], bi = A.b_idx[ip];
    const V2 q0 = vtx(A, ip, cCorners[i][0]), q1 = vtx(A, ip, cCorners[i][1]);
    const V2 q2 = vtx(A, ip, cCorners[i][2]), q3 = vtx(A, ip, cCorners[i][3]);
    int pb0[4], pb1[4], pa0[4], pas[4];
#pragma unroll
    for (int side = 0; side < 4; ++side) {
        const int sub = cSubfaceBySideChild[side][i];
        pb0[side] = A.bnd[SI(A, ip, side)];
        pb1[side] = A.bnd[SI(A, ip, side + 4)];
        pa0[side] = A.adj[SI(A, ip, side)];
        pas[side] = A.adj[SI(A, ip, sub >= 0 ? sub : side)];
    }
    Z zp{};
    double gx[4] = {0.0, 0.0, 0.0, 0.0}, gy[4] = {0.0, 0.0, 0.0, 0.0};
    double pcx = 0.0, pcy = 0.0;
    if (z.r) {
        zp = Z{z.r[ip], z.u[ip], z.v[ip], z.p[ip], z.e[ip]};
        if (slope) {
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                gx[k] = Gd.dx[k * A.stride + ip];
                gy[k] = Gd.dy[k * A.stride + ip];
            }
            pcx = A.cx[ip];
            pcy = A.cy[ip];
        }
    }
    // batch 2: positions of the siblings and of the child itself; the outer
    // neighbours (through the parent's side, or its sub-face on a split side)
    const int ich = mc + i;
    const int lk0 = A.next[mc + cLinks[2 * i][2]], lk1 = A.next[mc + cLinks[2 * i + 1][2]];
    const int nself = A.next[ich];
    int jc[4], jl[4], jf[4], jn[4];
#pragma unroll
    for (int side = 0; side < 4; ++side) {
        const int j = pb1[side] == B_UNDEF ? pa0[side] : pas[side];
        jc[side] = (unsigned)j < (unsigned)A.n ? j : ip;
        jl[side] = A.level[jc[side]];
        jf[side] = A.flag[jc[side]];
        jn[side] = A.next[jc[side]];
    }
    // batch 3: the positions behind them (x0 / x1 < 0: not needed)
    int y0[4], y1[4];
    bool splits[4];
#pragma unroll
    for (int side = 0; side < 4; ++side) {
        const bool simple = pb1[side] == B_UNDEF;
        int x0 = -1, x1 = -1;
        splits[side] = false;
        if (simple) {
            if (lvl > jl[side]) x0 = jn[side] + adjacent_child(side, zi & 3);
            else if (jf[side] != 0) x0 = jn[side] + adjacent_child(side, i);
        } else if (jl[side] + jf[side] > lvl + 1) {
            splits[side] = true;
            x0 = jn[side] + cNeibChild0[side];
            x1 = jn[side] + cNeibChild0[side + 4];
        } else if (jf[side] < 0) {
            x0 = jn[side];
        }
        y0[side] = A.next[(unsigned)x0 < (unsigned)A.stride ? x0 : ip];
        y1[side] = A.next[(unsigned)x1 < (unsigned)A.stride ? x1 : ip];
    }

    QuadG G;
    quad_geom(q0, q1, q2, q3, G);
    int b[8], a[8];
#pragma unroll
    for (int s = 0; s < 8; ++s) {
        b[s] = s < 4 ? B_INNER : B_UNDEF;
        a[s] = -1;
    }
    a[cLinks[2 * i][1]] = lk0;
    a[cLinks[2 * i + 1][1]] = lk1;
    unsigned split = 0u;
#pragma unroll
    for (int side = 0; side < 4; ++side) {
        if (cSubfaceBySideChild[side][i] < 0) continue;   // child i does not touch this side
        const int bflag = pb0[side];
        if (bflag == B_UNDEF) continue;
        b[side] = bflag;
        if (is_bnd(bflag)) {
            a[side] = nself;
            continue;
        }
        if (pb1[side] == B_UNDEF) {
            a[side] = (lvl > jl[side] || jf[side] != 0) ? y0[side] : jn[side];
        } else if (splits[side]) {
            b[side + 4] = bflag;
            a[side] = y0[side];
            a[side + 4] = y1[side];
            split |= 1u << side;
        } else {
            a[side] = jf[side] == 0 ? jn[side] : (jf[side] < 0 ? y0[side] : jc[side]);
        }
    }
    write_cell(A, ich, G, bi, lvl + 1, 4 * zi + i);
    write_slots(A, ich, G, b, a, split);
    if (z.r) {
        Z q;
        bool ok = false;
        if (slope) ok = split_slope(zp, gx, gy, G.C.x - pcx, G.C.y - pcy, eos, q);
        const unsigned long long m = __ballot(!ok);
        const bool any_bad = ((m >> (threadIdx.x & 60)) & 0xFULL) != 0ULL;
        if (any_bad) q = zp;
        z.r[ich] = q.r; z.u[ich] = q.u; z.v[ich] = q.v; z.p[ich] = q.p; z.e[ich] = q.e;
        wait[ich] = 0;
    }
    if (i == 0) A.idx[ip] = -1;
}

// coarse_cell / select_children / make_parent (coarse.h:95-404) and the
// conservative merge (sm_fluid.cpp:382-393) for the family ch[0..3]
// (children in z order, from the positions pass).  Loads in three dependent
// batches (the children; the parent's position and the outer neighbours;
// the positions behind them), then the parent's cell and slots.
__device__ void coarse_one(const AmrMesh& A, const DevState& z, int* wait, const Eos& eos, const int* ch) {
    // batch 1: the children
    const int ip = A.next[ch[0]];
    const V2 q0 = vtx(A, ch[0], 0), q1 = vtx(A, ch[1], 2), q2 = vtx(A, ch[2], 6), q3 = vtx(A, ch[3], 8);
    const int lc = A.level[ch[0]], zc0 = A.z_idx[ch[0]], bc = A.b_idx[ch[0]];
    int sb[4], sa[4], sa2[4];
#pragma unroll
    for (int side = 0; side < 4; ++side) {
        const int c1 = ch[cChildBySubface[side]], c2 = ch[cChildBySubface[side + 4]];
        sb[side] = A.bnd[SI(A, c1, side)];
        sa[side] = A.adj[SI(A, c1, side)];
        sa2[side] = A.adj[SI(A, c2, side)];
    }
    Z zk[4];
    double Vk[4];
    if (z.r) {
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const int c = ch[k];
            zk[k] = Z{z.r[c], z.u[c], z.v[c], z.p[c], z.e[c]};
            Vk[k] = A.vol[c];
        }
    }
    // batch 2: the parent's final position, the neighbours across each side
    const int ip_next = A.next[ip];
    int nl[4], nf[4], nn[4], ml[4], mf[4], mn[4];
#pragma unroll
    for (int side = 0; side < 4; ++side) {
        const int j = (unsigned)sa[side] < (unsigned)A.n ? sa[side] : ch[0];
        const int k = (unsigned)sa2[side] < (unsigned)A.n ? sa2[side] : ch[0];
        nl[side] = A.level[j];
        nf[side] = A.flag[j];
        nn[side] = A.next[j];
        ml[side] = A.level[k];
        mf[side] = A.flag[k];
        mn[side] = A.next[k];
    }
    // batch 3: the positions behind them
    const int lp = lc - 1;
    int y0[4], y1[4];
#pragma unroll
    for (int side = 0; side < 4; ++side) {
        int x0 = -1, x1 = -1;
        if (nl[side] + nf[side] <= lp) {
            if (nf[side] != 0) x0 = nn[side];
        } else {
            if (lc > nl[side]) x0 = nn[side] + cNeibChild0[side];
            else if (nf[side] != 0) x0 = nn[side];
            if (lc > ml[side]) x1 = mn[side] + cNeibChild0[side + 4];
            else if (mf[side] != 0) x1 = mn[side];
        }
        y0[side] = A.next[(unsigned)x0 < (unsigned)A.stride ? x0 : ip];
        y1[side] = A.next[(unsigned)x1 < (unsigned)A.stride ? x1 : ip];
    }

    QuadG G;
    quad_geom(q0, q1, q2, q3, G);
    int b[8], a[8];
#pragma unroll
    for (int s = 0; s < 8; ++s) {
        b[s] = s < 4 ? B_INNER : B_UNDEF;
        a[s] = -1;
    }
    unsigned split = 0u;
#pragma unroll
    for (int side = 0; side < 4; ++side) {
        b[side] = sb[side];
        if (is_bnd(sb[side])) {
            a[side] = ip_next;
            continue;
        }
        if (nl[side] + nf[side] <= lp) {
            a[side] = nf[side] == 0 ? nn[side] : y0[side];
            continue;
        }
        b[side + 4] = sb[side];
        a[side] = (lc > nl[side] || nf[side] != 0) ? y0[side] : nn[side];
        a[side + 4] = (lc > ml[side] || mf[side] != 0) ? y1[side] : mn[side];
        split |= 1u << side;
    }
    write_cell(A, ip, G, bc, lc - 1, zc0 / 4);
    write_slots(A, ip, G, b, a, split);
    if (z.r) {
        // merge: q_p = sum_k q(child_k) V_k / V_p
        double q[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const Q qc = to_q(zk[k]);
            q[0] = q[0] + qc.r * Vk[k];
            q[1] = q[1] + qc.mx * Vk[k];
            q[2] = q[2] + qc.my * Vk[k];
            q[3] = q[3] + qc.E * Vk[k];
        }
        const double Vp = G.vol;
        const Z zp = to_z(Q{q[0] / Vp, q[1] / Vp, q[2] / Vp, q[3] / Vp}, eos);
        z.r[ip] = zp.r; z.u[ip] = zp.u; z.v[ip] = zp.v; z.p[ip] = zp.p; z.e[ip] = zp.e;
        wait[ip] = 0;
    }
#pragma unroll
    for (int k = 0; k < 4; ++k) A.idx[ch[k]] = -1;
}

}  // namespace

// The steps' curve order carried across the adaptation.  Refining a cell
// puts its four children, in z order, exactly where the cell was on the
// curve; coarsening a family puts the parent where its four children were
// (the first of them, z_idx % 4 == 0, leads); every other cell keeps its
// place.  So the new order is the old one with each entry expanded to the
// cc entries of its marks, at the scanned position rank[j].cc, and each
// entry's mesh index is the final position the adaptation gives it:
// next[o] for a kept cell, next[next[o] + k] for child k, next[next[o]] for
// a family's parent.
__device__ void carry_curve(const AmrMesh& A, const int* perm, const AmrMarks* rank, int* perm_new, int* iperm_new,
                            int j) {
    const int o = perm[j];
    const int f = A.flag[o];
    const int at = rank[j].cc;
    if (f > 0) {
        const int mc = A.next[o];
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const int p = A.next[mc + k];
            perm_new[at + k] = p;
            iperm_new[p] = at + k;
        }
    } else if (f == 0 || (A.z_idx[o] & 3) == 0) {
        const int p = f == 0 ? A.next[o] : A.next[A.next[o]];
        perm_new[at] = p;
        iperm_new[p] = at;
    }
}

// One launch for the three kinds of cells, by workgroup: the first
// blocks_retain workgroups take the retained cells (one thread each, flag 0),
// the next blocks_refine the refined cells (a lane quad per parent, one lane
// per child; the slope distributor's all-or-nothing fallback is decided
// across the quad), then blocks_coarse the coarsened families (one thread
// per family, its children from `fam`), the rest the carried curve order
// (one thread per old curve position).  With z.r null the kernel does the
// mesh only.
__global__ void __launch_bounds__(256) k_amr_geometry(AmrMesh A, DevState z, DevGrad G, int* wait, Eos eos,
                                                      int slope, const int* ref_list, int n_refine,
                                                      const int* fam, int n_parents, int blocks_retain,
                                                      int blocks_refine, int blocks_coarse, const int* perm,
                                                      const AmrMarks* rank, int* perm_new, int* iperm_new) {
    const int b = (int)blockIdx.x;
    if (b >= blocks_retain + blocks_refine + blocks_coarse) {
        const int j = (b - blocks_retain - blocks_refine - blocks_coarse) * 256 + (int)threadIdx.x;
        if (j < A.n) carry_curve(A, perm, rank, perm_new, iperm_new, j);
        return;
    }
    if (b < blocks_retain) {
        const int ic = b * 256 + (int)threadIdx.x;
        if (ic >= A.n || A.flag[ic] != 0) return;
        retain_cell(A, ic);
        return;
    }
    if (b < blocks_retain + blocks_refine) {
        const int t = (b - blocks_retain) * 256 + (int)threadIdx.x;
        const int pi = t >> 2, i = t & 3;
        if (pi >= n_refine) return;   // whole quads leave together
        refine_one(A, z, G, wait, eos, slope != 0, ref_list[pi], i);
        return;
    }
    const int t = (b - blocks_retain - blocks_refine) * 256 + (int)threadIdx.x;
    if (t >= n_parents) return;
    int ch[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) ch[k] = fam[4 * t + k];
    coarse_one(A, z, wait, eos, ch);
}

// SwapLists::move_elements (swap_lists.h:129-138) + AmrCells::move_item
// (amr_cells.cpp:428-468) + the data rows (init layer, wait).  16 threads
// per move: lanes 0-7 copy one face slot each, lane 8 the vertices, lane 9
// the cell arrays and the state.
__global__ void __launch_bounds__(256) k_amr_move(AmrMesh A, DevState z, int* wait, int n_short, int nz_total,
                                                  const AmrMarks* rank, int bound, const int* swap_src,
                                                  const int* swap_dst) {
    const int t = cell_id();
    const int i = t >> 4, k = t & 15;
    if (i >= bound || k >= 10) return;
    const int n_swap = n_short >= A.n ? nz_total : rank[n_short].nz;   // the holes
    if (i >= n_swap) return;
    const int from = swap_src[i], to = swap_dst[i];
    // each lane loads everything it copies before its first store (sources
    // and destinations are distinct cells, but of the same array types)
    if (k < 8) {
        const int a = SI(A, from, k), b = SI(A, to, k);
        const int bnd = A.bnd[a], adj = A.adj[a];
        const double nx = A.nx[a], ny = A.ny[a], fcx = A.fcx[a], fcy = A.fcy[a], area = A.area[a];
        A.bnd[b] = bnd;
        A.adj[b] = adj;
        A.nx[b] = nx;
        A.ny[b] = ny;
        A.fcx[b] = fcx;
        A.fcy[b] = fcy;
        A.area[b] = area;
    } else if (k == 8) {
        double vx[9], vy[9];
#pragma unroll
        for (int v = 0; v < 9; ++v) {
            vx[v] = A.vx[v * A.stride + from];
            vy[v] = A.vy[v * A.stride + from];
        }
#pragma unroll
        for (int v = 0; v < 9; ++v) {
            A.vx[v * A.stride + to] = vx[v];
            A.vy[v * A.stride + to] = vy[v];
        }
    } else {
        const int flag = A.flag[from], level = A.level[from], b_idx = A.b_idx[from], z_idx = A.z_idx[from];
        const double cx = A.cx[from], cy = A.cy[from], vol = A.vol[from];
        Z q{};
        int w = 0;
        if (z.r) {   // mesh only: the solver moves its own rows
            q = Z{z.r[from], z.u[from], z.v[from], z.p[from], z.e[from]};
            w = wait[from];
        }
        A.next[to] = to;
        A.idx[to] = to;
        A.flag[to] = flag;
        A.level[to] = level;
        A.b_idx[to] = b_idx;
        A.z_idx[to] = z_idx;
        A.cx[to] = cx;
        A.cy[to] = cy;
        A.vol[to] = vol;
        A.idx[from] = -1;
        if (z.r) {
            z.r[to] = q.r; z.u[to] = q.u; z.v[to] = q.v; z.p[to] = q.p; z.e[to] = q.e;
            wait[to] = w;
        }
    }
}

// Correctly rounded x^3 for x > 0: x^2 = p + e1 and p x = q + e2 exactly
// (fused multiply-adds), x^3 = q + (e2 + e1 x) rounded once.
__device__ __forceinline__ double cube_cr(double x) {
    const double p = x * x;
    const double e1 = __builtin_fma(x, x, -p);
    const double q = p * x;
    const double e2 = __builtin_fma(p, x, -q);
    const double t = __builtin_fma(e1, x, e2);
    return q + t;
}

// Derived per-cell data after an adaptation: slot code words (type and
// general-normal bit), LSM weights w = area / |drn|^3 (gradient.h:251-256)
// and incircle diameters (amr_cells.cpp:107-114).  counts[0], counts[1]:
// defined interior (inner + periodic) and boundary slots, for the unique-face
// count interior / 2 + boundary (one atomic per workgroup).  An adjacency
// out of range sets bit 2 of the error word.
__global__ void __launch_bounds__(256) k_amr_derive(AmrMesh A, DevState z, Eos eos, unsigned long long* counts,
                                                    unsigned long long* dt_slot, int* err) {
    const int c = cell_id();
    int n_int = 0, n_bnd = 0;
    double cand = INFINITY;
    if (c < A.n) {
        const double kNX[4] = {-1.0, 1.0, 0.0, 0.0}, kNY[4] = {0.0, 0.0, -1.0, 1.0};
        // loads in two dependent batches, every load unconditional (the
        // slot arrays are read for undefined slots too, and a neighbour
        // index out of range reads the cell itself), the slot types
        // selecting afterwards: the cell's own and slot data, then the
        // neighbours' centres; then the stores (the weights are stored into
        // an array of the loads' type)
        const double cx = A.cx[c], cy = A.cy[c];
        const V2 va = vtx(A, c, 5), vb = vtx(A, c, 3), vp = vtx(A, c, 7), vq = vtx(A, c, 1);
        Z zc{};
        if (z.r) zc = Z{z.r[c], z.u[c], z.v[c], z.p[c], z.e[c]};
        int t[8], nb[8], bl[8], al[8];
        double nx[8], ny[8], fcx[8], fcy[8], area[8], px[8], py[8];
#pragma unroll
        for (int s = 0; s < 8; ++s) {
            const int f = SI(A, c, s);
            bl[s] = A.bnd[f];
            al[s] = A.adj[f];
            nx[s] = A.nx[f];
            ny[s] = A.ny[f];
            fcx[s] = A.fcx[f];
            fcy[s] = A.fcy[f];
            area[s] = A.area[f];
        }
#pragma unroll
        for (int s = 0; s < 8; ++s) {
            switch (bl[s]) {
                case B_UNDEF: t[s] = SC_UNDEF; break;
                case B_INNER: t[s] = SC_INNER; break;
                case B_PERIODIC: t[s] = SC_PERIODIC; break;
                case 1: t[s] = SC_WALL; break;
                default: t[s] = SC_COPY; break;
            }
            const bool def = t[s] != SC_UNDEF;
            nb[s] = def ? al[s] : c;
            nx[s] = def ? nx[s] : 0.0;
            ny[s] = def ? ny[s] : 0.0;
            fcx[s] = def ? fcx[s] : 0.0;
            fcy[s] = def ? fcy[s] : 0.0;
            area[s] = def ? area[s] : 0.0;
        }
        bool ok[8];
#pragma unroll
        for (int s = 0; s < 8; ++s) {
            ok[s] = t[s] != SC_UNDEF && nb[s] >= 0 && nb[s] < A.n && !(t[s] >= SC_WALL && nb[s] != c);
            const bool inner = ok[s] && t[s] == SC_INNER;
            const int j = inner ? nb[s] : c;
            const double qx = A.cx[j], qy = A.cy[j];
            px[s] = inner ? qx : 0.0;
            py[s] = inner ? qy : 0.0;
        }
        int cw = 0;
#pragma unroll
        for (int s = 0; s < 8; ++s) {
            const int f = SI(A, c, s);
            if (t[s] == SC_UNDEF) {
                A.adj[f] = c;
                continue;
            }
            const int side = s & 3;
            const bool gen = !(nx[s] == kNX[side] && ny[s] == kNY[side]);
            cw |= (t[s] | (gen ? kGeneralNormal : 0)) << (4 * s);
            if (t[s] >= SC_WALL) ++n_bnd;
            else ++n_int;
            if (!ok[s]) {   // inconsistent mesh: report, keep in bounds
                A.adj[f] = c;
                atomicOr(err, 2);
                continue;
            }
            double qx = px[s], qy = py[s];
            if (t[s] != SC_INNER) {
                const double d2 = 2.0 * ((fcx[s] - cx) * nx[s] + (fcy[s] - cy) * ny[s]);
                qx = cx + d2 * nx[s];
                qy = cy + d2 * ny[s];
            }
            const double drx = qx - cx, dry = qy - cy;
            const double d = drx * nx[s] + dry * ny[s];
            const double dnx = d * nx[s], dny = d * ny[s];
            const double len = sqrt(dnx * dnx + dny * dny);
            A.w[f] = (1.0 * area[s]) / cube_cr(len);
        }
        A.code[c] = cw;
        const double ax = va.x - vb.x, ay = va.y - vb.y, bx = vp.x - vq.x, by = vp.y - vq.y;
        const double s1 = ax * ax + ay * ay, s2 = bx * bx + by * by;
        const double h = sqrt((s2 < s1) ? s2 : s1);
        A.h[c] = h;
        // CFL candidate of the adapted state (sm_fluid.cpp:104-114)
        if (z.r) cand = h / (sqrt(vsq(zc)) + eos_c(eos, zc.r, zc.p));
    }
    if (!(cand < INFINITY)) cand = INFINITY;   // NaN and +inf never win
    for (int off = 32; off > 0; off >>= 1) {
        n_int += __shfl_xor(n_int, off, 64);
        n_bnd += __shfl_xor(n_bnd, off, 64);
        const double o = __shfl_xor(cand, off, 64);
        cand = (o < cand) ? o : cand;
    }
    __shared__ int red[2][4];
    __shared__ double dred[4];
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    if (lane == 0) {
        red[0][wid] = n_int;
        red[1][wid] = n_bnd;
        dred[wid] = cand;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        const int nw = (blockDim.x + 63) >> 6;
        unsigned long long ti = 0, tb = 0;
        double m = dred[0];
        for (int w = 0; w < nw; ++w) {
            ti += (unsigned long long)red[0][w];
            tb += (unsigned long long)red[1][w];
            m = (dred[w] < m) ? dred[w] : m;
        }
        if (ti) atomicAdd(&counts[0], ti);
        if (tb) atomicAdd(&counts[1], tb);
        if (m < INFINITY) atomicMin(dt_slot, (unsigned long long)__double_as_longlong(m));
    }
}

// ---------------------------------------------------------------- launchers
static inline dim3 grid256(int n) { return dim3((n + 255) / 256); }

hipError_t amr_criterion(const AmrMesh& A, const DevState& z, const DevGrad& G, int criterion, double thr,
                         hipStream_t s) {
    hipLaunchKernelGGL(k_amr_criterion, grid256(A.n), dim3(256), 0, s, A, z, G, criterion, thr);
    return hipGetLastError();
}

hipError_t amr_balance(const AmrMesh& A, int max_level, hipStream_t s) {
    hipLaunchKernelGGL(k_amr_restrict, grid256(A.n), dim3(256), 0, s, A, max_level);
    for (int lvl = max_level - 1; lvl >= 0; --lvl) {
        hipLaunchKernelGGL(k_amr_round<0>, grid256(A.n), dim3(256), 0, s, A, lvl);
        hipLaunchKernelGGL(k_amr_round<1>, grid256(A.n), dim3(256), 0, s, A, lvl);
        hipLaunchKernelGGL(k_amr_round<2>, grid256(A.n), dim3(256), 0, s, A, lvl);
    }
    return hipGetLastError();
}

struct AddMarks {
    __device__ __forceinline__ AmrMarks operator()(const AmrMarks& a, const AmrMarks& b) const {
        return AmrMarks{a.ref + b.ref, a.head + b.head, a.nz + b.nz, a.cc + b.cc};
    }
    __device__ static __forceinline__ AmrMarks shfl(const AmrMarks& v, int d) {
        return AmrMarks{__shfl_up(v.ref, d, 64), __shfl_up(v.head, d, 64), __shfl_up(v.nz, d, 64),
                        __shfl_up(v.cc, d, 64)};
    }
};

int amr_scan3_tiles(int count) { return scan::tiles_of(count); }

hipError_t amr_scan3(const AmrMarks* in, AmrMarks* out, AmrMarks* tile_sums, int count, AmrMarks* total, hipStream_t s) {
    return scan::exclusive(in, out, tile_sums, count, AddMarks(), AmrMarks{0, 0, 0, 0}, total, s);
}

hipError_t amr_mark(const AmrMesh& A, AmrMarks* marks, const int* perm, hipStream_t s) {
    hipLaunchKernelGGL(k_amr_mark, grid256(A.n), dim3(256), 0, s, A, marks, perm);
    return hipGetLastError();
}

hipError_t amr_positions(const AmrMesh& A, int n_large, int n_short, int n_parents, const AmrMarks* rank,
                         int* hole_at, int* ref_list, int* crs_list, int* fam, unsigned long long* faces,
                         unsigned long long* dt_rd, unsigned long long* dt_wr, hipStream_t s) {
    hipLaunchKernelGGL(k_amr_positions, grid256(n_large), dim3(256), 0, s, A, n_large, n_short, n_parents, rank,
                       hole_at, ref_list, crs_list, fam, faces, dt_rd, dt_wr);
    return hipGetLastError();
}

hipError_t amr_swapmap(const AmrMesh& A, int n_short, int n_large, int nz_total, const AmrMarks* rank,
                       const int* hole_at, int* swap_src, int* swap_dst, int* err, hipStream_t s) {
    if (n_large <= n_short) return hipSuccess;
    hipLaunchKernelGGL(k_amr_swapmap, grid256(n_large - n_short), dim3(256), 0, s, A, n_short, n_large, nz_total,
                       rank, hole_at, swap_src, swap_dst, err);
    return hipGetLastError();
}

hipError_t amr_geometry(const AmrMesh& A, const DevState& z, const DevGrad& G, int* wait, const Eos& eos,
                        bool slope, const int* ref_list, int n_refine, const int* fam, int n_parents,
                        const int* perm, const AmrMarks* rank, int* perm_new, int* iperm_new, hipStream_t s) {
    const int br = (A.n + 255) / 256, bf = (4 * n_refine + 255) / 256, bc = (n_parents + 255) / 256;
    const int be = perm_new ? (A.n + 255) / 256 : 0;
    hipLaunchKernelGGL(k_amr_geometry, dim3(br + bf + bc + be), dim3(256), 0, s, A, z, G, wait, eos, slope ? 1 : 0,
                       ref_list, n_refine, fam, n_parents, br, bf, bc, perm, rank, perm_new, iperm_new);
    return hipGetLastError();
}

hipError_t amr_move(const AmrMesh& A, const DevState& z, int* wait, int n_short, int nz_total, const AmrMarks* rank,
                    int bound, const int* swap_src, const int* swap_dst, hipStream_t s) {
    if (bound <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_amr_move, grid256(16 * bound), dim3(256), 0, s, A, z, wait, n_short, nz_total, rank, bound,
                       swap_src, swap_dst);
    return hipGetLastError();
}

hipError_t amr_derive(const AmrMesh& A, const DevState& z, const Eos& eos, unsigned long long* counts,
                      unsigned long long* dt_slot, int* err, hipStream_t s) {
    hipLaunchKernelGGL(k_amr_derive, grid256(A.n), dim3(256), 0, s, A, z, eos, counts, dt_slot, err);
    return hipGetLastError();
}

}  // namespace zfv
