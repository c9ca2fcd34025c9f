// Device-side AMR-2D adaptation for gfx950: refinement flags, flag
// balancing, and the apply phase (positions, child/parent geometry, face
// split/merge, neighbour re-linking, compaction), mirroring zephyr's
// EuMesh::refine (mesh/euler/eu_mesh.cpp:290-359) on one process.
//
// The reference's serial parts become scans: the counters of
// setup_positions (mesh/amr/setup_positions.h:24-64) are exclusive scans of
// the refine / coarsen marks, and the SwapLists (swap_lists.h:28-58) pair
// the k-th hole (ascending) with the k-th surviving cell above n_short
// (descending), both found by scans.  Per-cell work (retain_cell, refine_cell,
// coarse_cell) runs one thread per old cell; each writes only its own slots
// and the cells it creates, as in the reference's parallel setup_geometry.
// Single process, Cartesian base: every adjacency rotation is 0.
#include "zfv_internal.hpp"

#include <hipcub/hipcub.hpp>

namespace zfv {

namespace {

// AMR-2D index tables (geom/indexing.h).
__constant__ int cSF[4][2] = {{0, 6}, {2, 8}, {0, 2}, {6, 8}};
__constant__ int cCF[8][2] = {{0, 3}, {2, 5}, {0, 1}, {6, 7}, {3, 6}, {5, 8}, {1, 2}, {7, 8}};
__constant__ int cChildBySubface[8] = {0, 1, 0, 2, 2, 3, 1, 3};
__constant__ int cChildrenBySide[4][2] = {{0, 2}, {1, 3}, {0, 1}, {2, 3}};
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

__device__ __forceinline__ void slot_undef(const AmrMesh& A, int c, int s) {
    A.bnd[SI(A, c, s)] = B_UNDEF;
    A.adj[SI(A, c, s)] = -1;
}

// AmrCells::set_cell(ic, Quad, axial = false) (amr_cells.cpp:725-785).
__device__ void set_cell(const AmrMesh& A, int ic, V2 v00, V2 v01, V2 v10, V2 v11) {
    A.idx[ic] = -1;
    A.flag[ic] = 0;
    A.b_idx[ic] = -1;
    A.z_idx[ic] = -1;
    A.level[ic] = -1;
    // Quad::area: 0.5 |(v11 - v00) x (v10 - v01)| (quad.cpp:322-324)
    const double ax = v11.x - v00.x, ay = v11.y - v00.y, bx = v10.x - v01.x, by = v10.y - v01.y;
    const double cz = ax * by - ay * bx;
    const double area = 0.5 * sqrt(cz * cz);
    A.vol[ic] = area;
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
    const V2 C{Cx / d6, Cy / d6};
    A.cx[ic] = C.x;
    A.cy[ic] = C.y;
    // SqQuad(const Quad&): bilinear vertices (quad.cpp:452-457)
    const V2 V[9] = {v00,
                     {0.5 * (v00.x + v01.x), 0.5 * (v00.y + v01.y)},
                     v01,
                     {0.5 * (v00.x + v10.x), 0.5 * (v00.y + v10.y)},
                     {0.25 * (((v00.x + v01.x) + v10.x) + v11.x), 0.25 * (((v00.y + v01.y) + v10.y) + v11.y)},
                     {0.5 * (v01.x + v11.x), 0.5 * (v01.y + v11.y)},
                     v10,
                     {0.5 * (v10.x + v11.x), 0.5 * (v10.y + v11.y)},
                     v11};
#pragma unroll
    for (int k = 0; k < 9; ++k) {
        A.vx[k * A.stride + ic] = V[k].x;
        A.vy[k * A.stride + ic] = V[k].y;
    }
    for (int s = 0; s < 8; ++s) slot_undef(A, ic, s);
    for (int side = 0; side < 4; ++side) {
        const int f = SI(A, ic, side);
        line_geom(V[cSF[side][0]], V[cSF[side][1]], C, A.area[f], A.fcx[f], A.fcy[f], A.nx[f], A.ny[f]);
        A.bnd[f] = B_INNER;
    }
}

__device__ __forceinline__ bool simple_face(const AmrMesh& A, int ic, int side) {
    return A.bnd[SI(A, ic, side + 4)] == B_UNDEF;
}

// setup_face_geom with view = SqQuad centre (faces.h:57-97).
__device__ void face_geom(const AmrMesh& A, int ic, int slot, int a, int b) {
    const int f = SI(A, ic, slot);
    line_geom(vtx(A, ic, a), vtx(A, ic, b), vtx(A, ic, 4), A.area[f], A.fcx[f], A.fcy[f], A.nx[f], A.ny[f]);
}

// split_face_impl (faces.h:121-130).
__device__ void split_face(const AmrMesh& A, int ic, int side) {
    const int f0 = SI(A, ic, side), f1 = SI(A, ic, side + 4);
    A.bnd[f1] = A.bnd[f0];
    A.adj[f1] = A.adj[f0];
    face_geom(A, ic, side, cCF[side][0], cCF[side][1]);
    face_geom(A, ic, side + 4, cCF[side + 4][0], cCF[side + 4][1]);
}

// merge_faces_impl (faces.h:137-154).
__device__ void merge_faces(const AmrMesh& A, int ic, int side) {
    slot_undef(A, ic, side + 4);
    face_geom(A, ic, side, cSF[side][0], cSF[side][1]);
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
// flags, writes the restricted ones to flag0 (snapshot of the sort phase).
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

__global__ void __launch_bounds__(256) k_amr_copy_flags(AmrMesh A) {
    const int ic = cell_id();
    if (ic < A.n) A.flag[ic] = A.flag0[ic];
}

// One balancing round at level lvl (balancing_fast.h:273-390): round 1 on
// the cells that kept 0, rounds 2-3 and 4 on the cells restricted to -1.
__global__ void __launch_bounds__(256) k_amr_round(AmrMesh A, int lvl, int round) {
    const int ic = cell_id();
    if (ic >= A.n || A.level[ic] != lvl) return;
    const int f0 = A.flag0[ic];
    if (round == 1) {
        if (f0 != 0) return;
        for (int s = 0; s < 8; ++s) {
            const int b = A.bnd[SI(A, ic, s)];
            if (b == B_UNDEF || is_bnd(b)) continue;
            const int jc = A.adj[SI(A, ic, s)];
            if (A.level[jc] + A.flag[jc] > lvl + 1) {
                A.flag[ic] = 1;
                return;
            }
        }
        return;
    }
    if (f0 != -1 || A.flag[ic] >= 0) return;
    if (round == 4) {
        int sibs[3];
        get_siblings(A, ic, sibs);
        for (int k = 0; k < 3; ++k) {
            if (A.flag[sibs[k]] >= 0) {
                A.flag[ic] = 0;
                break;
            }
        }
        return;
    }
    for (int s = 0; s < 8; ++s) {
        const int b = A.bnd[SI(A, ic, s)];
        if (b == B_UNDEF || is_bnd(b)) continue;
        const int jc = A.adj[SI(A, ic, s)];
        const int wanted = A.level[jc] + A.flag[jc];
        if (wanted > lvl) {
            if (wanted > lvl + 1) {
                A.flag[ic] = 1;
                return;
            }
            A.flag[ic] = 0;
        }
    }
}

// Statistics (statistics.h:59-81) and the marks scanned for positions.
__global__ void __launch_bounds__(256) k_amr_count(AmrMesh A, int* counts) {
    const int ic = cell_id();
    int ret = 0, ref = 0, crs = 0;
    if (ic < A.n) {
        const int f = A.flag[ic];
        ret = f == 0;
        ref = f > 0;
        crs = f < 0;
    }
    // wavefront reduction, then one atomic per wavefront
    for (int off = 32; off > 0; off >>= 1) {
        ret += __shfl_xor(ret, off, 64);
        ref += __shfl_xor(ref, off, 64);
        crs += __shfl_xor(crs, off, 64);
    }
    if ((threadIdx.x & 63) == 0) {
        if (ret) atomicAdd(&counts[0], ret);
        if (ref) atomicAdd(&counts[1], ref);
        if (crs) atomicAdd(&counts[2], crs);
    }
}

__global__ void __launch_bounds__(256) k_amr_mark(AmrMesh A, int n_short, int* ref_mark, int* crs_mark,
                                                  int* hole_mark) {
    const int ic = cell_id();
    if (ic >= A.n) return;
    const int f = A.flag[ic];
    ref_mark[ic] = f > 0;
    crs_mark[ic] = f < 0 && (A.z_idx[ic] & 3) == 0;
    hole_mark[ic] = ic < n_short && f != 0;
}

// setup_positions (setup_positions.h:24-64) from the scans, plus the hole table.
__global__ void __launch_bounds__(256) k_amr_positions(AmrMesh A, int n_large, int n_parents, const int* ref_rank,
                                                       const int* crs_rank, const int* hole_mark,
                                                       const int* hole_rank, int* hole_at) {
    const int ic = cell_id();
    const int n = A.n;
    if (ic >= n_large) return;
    if (ic >= n) {
        A.next[ic] = ic;
        return;
    }
    const int f = A.flag[ic];
    if (f == 0) A.next[ic] = ic;
    else if (f > 0) A.next[ic] = n + n_parents + 4 * ref_rank[ic];
    else if ((A.z_idx[ic] & 3) == 0) {
        const int pos = n + crs_rank[ic];
        int sibs[3];
        get_siblings(A, ic, sibs);
        A.next[ic] = pos;
        for (int k = 0; k < 3; ++k) A.next[sibs[k]] = pos;
    }
    if (hole_mark[ic]) hole_at[hole_rank[ic]] = ic;
}

// Positions p >= n_short that survive (new cells, or old cells kept as
// they are) are taken from the top: the k-th of them from above fills the
// k-th hole (swap_lists.h:33-55).
__global__ void __launch_bounds__(256) k_amr_actual_mark(AmrMesh A, int n_short, int n_large, int* act_mark) {
    const int i = cell_id();
    const int p = n_short + i;
    if (p >= n_large) return;
    act_mark[i] = p >= A.n || A.flag[p] == 0;
}

__global__ void __launch_bounds__(256) k_amr_swapmap(AmrMesh A, int n_short, int n_large, int n_swap,
                                                     const int* act_mark, const int* act_incl, const int* hole_at,
                                                     int* swap_src, int* swap_dst) {
    const int i = cell_id();
    const int p = n_short + i;
    if (p >= n_large || !act_mark[i]) return;
    const int r = n_swap - act_incl[i];
    if (r < 0 || r >= n_swap) return;   // unreachable for consistent counts (checked on the host)
    const int h = hole_at[r];
    A.next[p] = h;
    swap_src[r] = p;
    swap_dst[r] = h;
}

// ---------------------------------------------------------------- geometry
namespace {

// retain_cell (retain.h:17-216).
__device__ void retain_cell(const AmrMesh& A, int ic) {
    const int lvl_c = A.level[ic];
    for (int side = 0; side < 4; ++side) {
        const int iface = SI(A, ic, side);
        const int b = A.bnd[iface];
        if (b == B_UNDEF) continue;
        if (is_bnd(b)) {
            A.adj[iface] = A.next[ic];
            continue;
        }
        const int jc = A.adj[iface];
        const int lvl_n = A.level[jc];
        const int fl = A.flag[jc];
        if (lvl_c == lvl_n) {
            const int neib_next = A.next[jc];
            if (fl > 0) {
                split_face(A, ic, side);
                A.adj[SI(A, ic, side)] = A.next[neib_next + cNeibChild0[side]];
                A.adj[SI(A, ic, side + 4)] = A.next[neib_next + cNeibChild0[side + 4]];
            } else if (fl == 0) {
                A.adj[iface] = neib_next;
            } else {
                A.adj[iface] = A.next[neib_next];
            }
        } else if (lvl_c < lvl_n) {
            if (fl < 0) {
                merge_faces(A, ic, side);
                A.adj[iface] = A.next[A.next[jc]];
            } else {
                for (int sub = side; sub < 8; sub += 4) {
                    const int jf = SI(A, ic, sub);
                    if (A.bnd[jf] != B_UNDEF) A.adj[jf] = A.next[A.adj[jf]];
                }
            }
        } else {
            if (fl == 0) A.adj[iface] = A.next[jc];
            else A.adj[iface] = A.next[A.next[jc] + adjacent_child(side, A.z_idx[ic] & 3)];
        }
    }
}

// SmFluid distributor split "slope" / "const" (sm_fluid.cpp:395-458).
__device__ void split_data(const AmrMesh& A, const DevState& z, const DevGrad& G, int* wait, const Eos& eos,
                           int ip, int mc, bool slope) {
    const Z zp{z.r[ip], z.u[ip], z.v[ip], z.p[ip], z.e[ip]};
    bool bad = !slope;
    if (slope) {
        const int st = A.stride;
        const double gxr = G.dx[ip], gxu = G.dx[st + ip], gxv = G.dx[2 * st + ip], gxp = G.dx[3 * st + ip];
        const double gyr = G.dy[ip], gyu = G.dy[st + ip], gyv = G.dy[2 * st + ip], gyp = G.dy[3 * st + ip];
        const double dR = (eos.g - 1.0) * (zp.e - eos.e0);
        const double dE = (eos.g - 1.0) * zp.r;
        const double gex = (gxp - dR * gxr) / dE, gey = (gyp - dR * gyr) / dE;
        Z ch[4];
        for (int k = 0; k < 4 && !bad; ++k) {
            const int c = mc + k;
            const double drx = A.cx[c] - A.cx[ip], dry = A.cy[c] - A.cy[ip];
            Z q = zp;
            q.r = zp.r + gxr * drx + gyr * dry;
            const double ratio = zp.r / q.r;
            q.u = zp.u + ratio * (gxu * drx + gyu * dry);
            q.v = zp.v + ratio * (gxv * drx + gyv * dry);
            const double du = zp.u - q.u, dv = zp.v - q.v;
            q.e = zp.e - 0.5 * (du * du + dv * dv) + ratio * (gex * drx + gey * dry);
            q.p = eos_p(eos, q.r, q.e);
            if (is_bad(q, eos)) bad = true;
            ch[k] = q;
        }
        if (!bad) {
            for (int k = 0; k < 4; ++k) {
                const int c = mc + k;
                z.r[c] = ch[k].r; z.u[c] = ch[k].u; z.v[c] = ch[k].v; z.p[c] = ch[k].p; z.e[c] = ch[k].e;
                wait[c] = 0;
            }
        }
    }
    if (bad) {
        for (int k = 0; k < 4; ++k) {
            const int c = mc + k;
            z.r[c] = zp.r; z.u[c] = zp.u; z.v[c] = zp.v; z.p[c] = zp.p; z.e[c] = zp.e;
            wait[c] = 0;
        }
    }
}

// refine_cell / make_children / link_siblings (refine.h:36-444).
__device__ void refine_cell(const AmrMesh& A, const DevState& z, const DevGrad& G, int* wait, const Eos& eos,
                            int ip, bool slope) {
    const int mc = A.next[ip];
    V2 V[9];
    for (int k = 0; k < 9; ++k) V[k] = vtx(A, ip, k);
    for (int i = 0; i < 4; ++i)
        set_cell(A, mc + i, V[cCorners[i][0]], V[cCorners[i][1]], V[cCorners[i][2]], V[cCorners[i][3]]);
    const int lvl = A.level[ip], zi = A.z_idx[ip], bi = A.b_idx[ip];
    for (int i = 0; i < 4; ++i) {
        const int ich = mc + i;
        A.idx[ich] = ich;
        A.flag[ich] = 0;
        A.b_idx[ich] = bi;
        A.level[ich] = lvl + 1;
        A.z_idx[ich] = 4 * zi + i;
    }
    for (int l = 0; l < 8; ++l) {
        const int f = SI(A, mc + cLinks[l][0], cLinks[l][1]);
        A.bnd[f] = B_INNER;
        A.adj[f] = A.next[mc + cLinks[l][2]];
    }
    for (int side = 0; side < 4; ++side) {
        const int pf = SI(A, ip, side);
        const int bflag = A.bnd[pf];
        if (bflag == B_UNDEF) continue;
        for (int k = 0; k < 2; ++k) A.bnd[SI(A, mc + cChildrenBySide[side][k], side)] = bflag;
        if (is_bnd(bflag)) {
            for (int k = 0; k < 2; ++k) {
                const int i = cChildrenBySide[side][k];
                A.adj[SI(A, mc + i, side)] = A.next[mc + i];
            }
            continue;
        }
        if (simple_face(A, ip, side)) {
            const int jc = A.adj[pf];
            for (int k = 0; k < 2; ++k) {
                const int i = cChildrenBySide[side][k];
                const int chf = SI(A, mc + i, side);
                if (lvl > A.level[jc]) A.adj[chf] = A.next[A.next[jc] + adjacent_child(side, zi & 3)];
                else if (A.flag[jc] == 0) A.adj[chf] = A.next[jc];
                else A.adj[chf] = A.next[A.next[jc] + adjacent_child(side, i)];
            }
        } else {
            for (int k = 0; k < 2; ++k) {
                const int i = cChildrenBySide[side][k];
                const int chf = SI(A, mc + i, side);
                const int jc = A.adj[SI(A, ip, cSubfaceBySideChild[side][i])];
                const int fl = A.flag[jc];
                if (fl == 0) A.adj[chf] = A.next[jc];
                else if (fl < 0) A.adj[chf] = A.next[A.next[jc]];
                else A.adj[chf] = jc;
            }
        }
    }
    for (int side = 0; side < 4; ++side) {
        const int pf = SI(A, ip, side);
        if (simple_face(A, ip, side) || A.bnd[pf] == B_UNDEF || is_bnd(A.bnd[pf])) continue;
        for (int k = 0; k < 2; ++k) {
            const int i = cChildrenBySide[side][k];
            const int ich = mc + i;
            const int jc = A.adj[SI(A, ip, cSubfaceBySideChild[side][i])];
            if (A.level[jc] + A.flag[jc] > lvl + 1) {
                split_face(A, ich, side);
                const int neib_next = A.next[jc];
                A.adj[SI(A, ich, side)] = A.next[neib_next + cNeibChild0[side]];
                A.adj[SI(A, ich, side + 4)] = A.next[neib_next + cNeibChild0[side + 4]];
            }
        }
    }
    split_data(A, z, G, wait, eos, ip, mc, slope);
    A.idx[ip] = -1;
}

// coarse_cell / select_children / make_parent (coarse.h:95-404) and the
// conservative merge (sm_fluid.cpp:382-393).
__device__ void coarse_cell(const AmrMesh& A, const DevState& z, int* wait, const Eos& eos, int ich) {
    if ((A.z_idx[ich] & 3) != 0) {
        A.idx[ich] = -1;
        return;
    }
    int ch[4];
    ch[0] = ich;
    int sibs[3];
    get_siblings(A, ich, sibs);
    for (int k = 0; k < 3; ++k) ch[A.z_idx[sibs[k]] & 3] = sibs[k];
    const int ip = A.next[ich];
    set_cell(A, ip, vtx(A, ch[0], 0), vtx(A, ch[1], 2), vtx(A, ch[2], 6), vtx(A, ch[3], 8));
    const int ip_next = A.next[ip];
    A.idx[ip] = ip;
    A.flag[ip] = 0;
    A.b_idx[ip] = A.b_idx[ch[0]];
    A.level[ip] = A.level[ch[0]] - 1;
    A.z_idx[ip] = A.z_idx[ch[0]] / 4;
    for (int side = 0; side < 4; ++side) {
        const int some_ch = ch[cChildBySubface[side]];
        const int scf = SI(A, some_ch, side);
        const int pf = SI(A, ip, side);
        A.bnd[pf] = A.bnd[scf];
        if (is_bnd(A.bnd[scf])) {
            A.adj[pf] = ip_next;
            continue;
        }
        const int sn = A.adj[scf];
        const int wanted = A.level[sn] + A.flag[sn];
        if (wanted <= A.level[ip]) {
            A.adj[pf] = A.flag[sn] == 0 ? A.next[sn] : A.next[A.next[sn]];
            slot_undef(A, ip, side + 4);
            continue;
        }
        split_face(A, ip, side);
        for (int sub = side; sub < 8; sub += 4) {
            const int c2 = ch[cChildBySubface[sub]];
            const int jc = A.adj[SI(A, c2, side)];
            const int sf = SI(A, ip, sub);
            if (A.level[c2] > A.level[jc]) A.adj[sf] = A.next[A.next[jc] + cNeibChild0[sub]];
            else if (A.flag[jc] == 0) A.adj[sf] = A.next[jc];
            else A.adj[sf] = A.next[A.next[jc]];
        }
    }
    // merge: q_p = sum_k q(child_k) V_k / V_p
    double q[4] = {0.0, 0.0, 0.0, 0.0};
    for (int k = 0; k < 4; ++k) {
        const int c = ch[k];
        const Q qc = to_q(Z{z.r[c], z.u[c], z.v[c], z.p[c], z.e[c]});
        const double V = A.vol[c];
        q[0] = q[0] + qc.r * V;
        q[1] = q[1] + qc.mx * V;
        q[2] = q[2] + qc.my * V;
        q[3] = q[3] + qc.E * V;
    }
    const double Vp = A.vol[ip];
    const Z zp = to_z(Q{q[0] / Vp, q[1] / Vp, q[2] / Vp, q[3] / Vp}, eos);
    z.r[ip] = zp.r; z.u[ip] = zp.u; z.v[ip] = zp.v; z.p[ip] = zp.p; z.e[ip] = zp.e;
    wait[ip] = 0;
    A.idx[ich] = -1;
}

}  // namespace

__global__ void __launch_bounds__(128) k_amr_geometry(AmrMesh A, DevState z, DevGrad G, int* wait, Eos eos,
                                                      int slope) {
    const int ic = cell_id();
    if (ic >= A.n) return;
    const int f = A.flag[ic];
    if (f == 0) retain_cell(A, ic);
    else if (f > 0) refine_cell(A, z, G, wait, eos, ic, slope != 0);
    else coarse_cell(A, z, wait, eos, ic);
}

// SwapLists::move_elements (swap_lists.h:129-138) + AmrCells::move_item
// (amr_cells.cpp:428-468) + the data rows (init layer, wait).
__global__ void __launch_bounds__(256) k_amr_move(AmrMesh A, DevState z, int* wait, int n_swap,
                                                  const int* swap_src, const int* swap_dst) {
    const int i = cell_id();
    if (i >= n_swap) return;
    const int from = swap_src[i], to = swap_dst[i];
    A.next[to] = to;
    A.idx[to] = to;
    A.flag[to] = A.flag[from];
    A.level[to] = A.level[from];
    A.b_idx[to] = A.b_idx[from];
    A.z_idx[to] = A.z_idx[from];
    A.cx[to] = A.cx[from];
    A.cy[to] = A.cy[from];
    A.vol[to] = A.vol[from];
    for (int s = 0; s < 8; ++s) {
        const int a = SI(A, from, s), b = SI(A, to, s);
        A.bnd[b] = A.bnd[a];
        A.adj[b] = A.adj[a];
        A.nx[b] = A.nx[a];
        A.ny[b] = A.ny[a];
        A.fcx[b] = A.fcx[a];
        A.fcy[b] = A.fcy[a];
        A.area[b] = A.area[a];
    }
    for (int k = 0; k < 9; ++k) {
        A.vx[k * A.stride + to] = A.vx[k * A.stride + from];
        A.vy[k * A.stride + to] = A.vy[k * A.stride + from];
    }
    A.idx[from] = -1;
    z.r[to] = z.r[from]; z.u[to] = z.u[from]; z.v[to] = z.v[from]; z.p[to] = z.p[from]; z.e[to] = z.e[from];
    wait[to] = wait[from];
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
// and incircle diameters (amr_cells.cpp:107-114).
// counts[0]: inconsistency flag; counts[1], counts[2]: defined interior
// (inner + periodic) and boundary slots, for the unique-face count
// interior / 2 + boundary.
__global__ void __launch_bounds__(256) k_amr_derive(AmrMesh A, int* counts) {
    const int c = cell_id();
    int n_int = 0, n_bnd = 0;
    if (c < A.n) {
    const double kNX[4] = {-1.0, 1.0, 0.0, 0.0}, kNY[4] = {0.0, 0.0, -1.0, 1.0};
    int cw = 0;
    const double cx = A.cx[c], cy = A.cy[c];
    for (int s = 0; s < 8; ++s) {
        const int f = SI(A, c, s);
        const int b = A.bnd[f];
        int t;
        switch (b) {
            case B_UNDEF: t = SC_UNDEF; break;
            case B_INNER: t = SC_INNER; break;
            case B_PERIODIC: t = SC_PERIODIC; break;
            case 1: t = SC_WALL; break;
            default: t = SC_COPY; break;
        }
        if (t == SC_UNDEF) {
            A.adj[f] = c;
            continue;
        }
        const int side = s & 3;
        const double nx = A.nx[f], ny = A.ny[f];
        const bool gen = !(nx == kNX[side] && ny == kNY[side]);
        cw |= (t | (gen ? kGeneralNormal : 0)) << (4 * s);
        if (t >= SC_WALL) ++n_bnd;
        else ++n_int;
        const int nb = A.adj[f];
        if (nb < 0 || nb >= A.n || (t >= SC_WALL && nb != c)) {   // inconsistent mesh: report, keep in bounds
            A.adj[f] = c;
            counts[0] = 1;
            continue;
        }
        double px, py;
        if (t == SC_INNER) {
            px = A.cx[nb];
            py = A.cy[nb];
        } else {
            const double d2 = 2.0 * ((A.fcx[f] - cx) * nx + (A.fcy[f] - cy) * ny);
            px = cx + d2 * nx;
            py = cy + d2 * ny;
        }
        const double drx = px - cx, dry = py - cy;
        const double d = drx * nx + dry * ny;
        const double dnx = d * nx, dny = d * ny;
        const double len = sqrt(dnx * dnx + dny * dny);
        A.w[f] = (1.0 * A.area[f]) / cube_cr(len);
    }
    A.code[c] = cw;
    const V2 a = vtx(A, c, 5), b = vtx(A, c, 3), p = vtx(A, c, 7), q = vtx(A, c, 1);
    const double ax = a.x - b.x, ay = a.y - b.y, bx = p.x - q.x, by = p.y - q.y;
    const double s1 = ax * ax + ay * ay, s2 = bx * bx + by * by;
    A.h[c] = sqrt((s2 < s1) ? s2 : s1);
    }
    for (int off = 32; off > 0; off >>= 1) {
        n_int += __shfl_xor(n_int, off, 64);
        n_bnd += __shfl_xor(n_bnd, off, 64);
    }
    if ((threadIdx.x & 63) == 0) {
        if (n_int) atomicAdd(&counts[1], n_int);
        if (n_bnd) atomicAdd(&counts[2], n_bnd);
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
    hipLaunchKernelGGL(k_amr_copy_flags, grid256(A.n), dim3(256), 0, s, A);
    for (int lvl = max_level - 1; lvl >= 0; --lvl) {
        hipLaunchKernelGGL(k_amr_round, grid256(A.n), dim3(256), 0, s, A, lvl, 1);
        hipLaunchKernelGGL(k_amr_round, grid256(A.n), dim3(256), 0, s, A, lvl, 2);
        hipLaunchKernelGGL(k_amr_round, grid256(A.n), dim3(256), 0, s, A, lvl, 3);
        hipLaunchKernelGGL(k_amr_round, grid256(A.n), dim3(256), 0, s, A, lvl, 4);
    }
    return hipGetLastError();
}

hipError_t amr_count(const AmrMesh& A, int* counts, hipStream_t s) {
    hipLaunchKernelGGL(k_amr_count, grid256(A.n), dim3(256), 0, s, A, counts);
    return hipGetLastError();
}

hipError_t amr_scan(const int* in, int* out, int count, bool inclusive, void* tmp, size_t& tmp_bytes,
                    hipStream_t s) {
    if (inclusive) return hipcub::DeviceScan::InclusiveSum(tmp, tmp_bytes, in, out, count, s);
    return hipcub::DeviceScan::ExclusiveSum(tmp, tmp_bytes, in, out, count, s);
}

hipError_t amr_mark(const AmrMesh& A, int n_short, int* ref_mark, int* crs_mark, int* hole_mark, hipStream_t s) {
    hipLaunchKernelGGL(k_amr_mark, grid256(A.n), dim3(256), 0, s, A, n_short, ref_mark, crs_mark, hole_mark);
    return hipGetLastError();
}

hipError_t amr_positions(const AmrMesh& A, int n_large, int n_parents, const int* ref_rank, const int* crs_rank,
                         const int* hole_mark, const int* hole_rank, int* hole_at, hipStream_t s) {
    hipLaunchKernelGGL(k_amr_positions, grid256(n_large), dim3(256), 0, s, A, n_large, n_parents, ref_rank,
                       crs_rank, hole_mark, hole_rank, hole_at);
    return hipGetLastError();
}

hipError_t amr_actual_mark(const AmrMesh& A, int n_short, int n_large, int* act_mark, hipStream_t s) {
    hipLaunchKernelGGL(k_amr_actual_mark, grid256(n_large - n_short), dim3(256), 0, s, A, n_short, n_large,
                       act_mark);
    return hipGetLastError();
}

hipError_t amr_swapmap(const AmrMesh& A, int n_short, int n_large, int n_swap, const int* act_mark,
                       const int* act_incl, const int* hole_at, int* swap_src, int* swap_dst, hipStream_t s) {
    hipLaunchKernelGGL(k_amr_swapmap, grid256(n_large - n_short), dim3(256), 0, s, A, n_short, n_large, n_swap,
                       act_mark, act_incl, hole_at, swap_src, swap_dst);
    return hipGetLastError();
}

hipError_t amr_geometry(const AmrMesh& A, const DevState& z, const DevGrad& G, int* wait, const Eos& eos,
                        bool slope, hipStream_t s) {
    hipLaunchKernelGGL(k_amr_geometry, dim3((A.n + 127) / 128), dim3(128), 0, s, A, z, G, wait, eos,
                       slope ? 1 : 0);
    return hipGetLastError();
}

hipError_t amr_move(const AmrMesh& A, const DevState& z, int* wait, int n_swap, const int* swap_src,
                    const int* swap_dst, hipStream_t s) {
    if (n_swap <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_amr_move, grid256(n_swap), dim3(256), 0, s, A, z, wait, n_swap, swap_src, swap_dst);
    return hipGetLastError();
}

hipError_t amr_derive(const AmrMesh& A, int* counts, hipStream_t s) {
    hipLaunchKernelGGL(k_amr_derive, grid256(A.n), dim3(256), 0, s, A, counts);
    return hipGetLastError();
}

}  // namespace zfv
