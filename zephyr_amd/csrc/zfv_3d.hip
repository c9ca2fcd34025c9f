// The SmFluid step on three-dimensional Cuboid meshes for gfx950: the
// reference's AMR-3D cells (24 face slots, mesh/euler/amr_cells.cpp:796-834)
// as a Cuboid generates them (geom/generator/cuboid.cpp:300-442), with the
// six main slots L, R, B, T, Z (back), F (front) defined and unit normals
// -x, +x, -y, +y, -z, +z.  One thread per cell and stage, as the 2D mesh
// kernels: first-order fluxes, or the limited least-squares gradient fused
// with the predictor and then the corrector (sm_fluid.cpp:81-366).
//
// 3D state Z3 = {rho, u, v, w, p, e} (smf::PState, math/cfd/models.h:24-28);
// arithmetic in the reference's order, left to right, with Eigen's
// squaredNorm as ((u u + v v) + w w) and Matrix3d * Vector3d row by row
// ((R0 v0 + R1 v1) + R2 v2).  For the six unit normals the rotations of
// Rotate::to_local / to_global (rotate.cpp:7-53, the 3D branch for +-z)
// are signed permutations of the velocity: the matrix entries are 0 and +-1
// and the zero products change no non-zero value, so the maps below give
// the reference's values.  The same holds for boundary_value's reflection
// (sm_fluid.cpp:69-79), the mirror point (amr_faces.h:197-199) and
// drn = (dr.n) n (gradient.h:254): only the normal's axis carries a value.
// The least-squares system is then diagonal, and Eigen's cofactor inverse
// (det = (c00 a00 + c10 a10) + c20 a20, B = adj / det) reduces to
// B_aa = cof_aa / ((a11 a22) a00) with cof_00 = a11 a22, cof_11 = a22 a00,
// cof_22 = a00 a11, which is what is computed.
#include "zfv_3d.hpp"
#include "zfv_step_common.hpp"

namespace zfv {

namespace {

struct Z3 {
    double r, u, v, w, p, e;
};
struct Q3 {
    double r, mx, my, mz, E;
};

__device__ __forceinline__ double vsq3(const Z3& z) { return (z.u * z.u + z.v * z.v) + z.w * z.w; }

__device__ __forceinline__ Q3 to_q3(const Z3& z) {
    return Q3{z.r, z.r * z.u, z.r * z.v, z.r * z.w, z.r * (z.e + 0.5 * vsq3(z))};
}

__device__ __forceinline__ Z3 to_z3(const Q3& q, const Eos& eos) {
    Z3 z;
    z.r = q.r;
    z.u = q.mx / z.r;
    z.v = q.my / z.r;
    z.w = q.mz / z.r;
    z.e = q.E / z.r - 0.5 * vsq3(z);
    z.p = eos_p(eos, z.r, z.e);
    return z;
}

__device__ __forceinline__ bool is_bad3(const Z3& z, const Eos& q) {
    if (!isfinite(z.r) || !isfinite(z.p) || !isfinite(z.e)) return true;
    return z.r <= 0.0 || z.p < -q.P0;
}

// Flux(z) along the local x axis (models.cpp:131-137).
__device__ __forceinline__ Q3 phys_flux3(const Z3& z) {
    const double ru = z.r * z.u;
    return Q3{ru, ru * z.u + z.p, ru * z.v, ru * z.w, (z.r * (z.e + 0.5 * vsq3(z)) + z.p) * z.u};
}

__device__ __forceinline__ double qk(const Q3& q, int k) {
    return k == 0 ? q.r : (k == 1 ? q.mx : (k == 2 ? q.my : (k == 3 ? q.mz : q.E)));
}
__device__ __forceinline__ void qset(Q3& q, int k, double v) {
    if (k == 0) q.r = v;
    else if (k == 1) q.mx = v;
    else if (k == 2) q.my = v;
    else if (k == 3) q.mz = v;
    else q.E = v;
}

// HLLC::calc_flux, smf (math/cfd/flux/hllc.cpp:23-83); MODE 1 is HLLC_M
// (hllc.cpp:491-572: the tangential velocities averaged across the contact,
// V_C and W_C, with the matching energy correction), MODE 2 HLLC_LM
// (hllc.cpp:427-489: the outer waves scaled by phi = sin(pi/2 Ma / 0.1)
// below Ma = 0.1; sin is the device's, within an ulp of the host's).
template <int MODE>
__device__ Q3 hllc3(const Z3& zL, const Z3& zR, const Eos& eos, bool& bad) {
    const double c_L = eos_c(eos, zL.r, zL.p), c_R = eos_c(eos, zR.r, zR.p);
    const double S_L = smin(zL.u - c_L, zR.u - c_R);
    const double S_R = smax(zL.u + c_L, zR.u + c_R);
    if (S_L >= 0.0) return phys_flux3(zL);
    if (S_R <= 0.0) return phys_flux3(zR);
    const double a_L = zL.r * (S_L - zL.u);
    const double a_R = zR.r * (S_R - zR.u);
    const double S_C = (zL.p - zR.p + a_R * zR.u - a_L * zL.u) / (a_R - a_L);
    const double rho_sL = a_L / (S_L - S_C);
    const double rho_sR = a_R / (S_R - S_C);
    double E_sL = (zL.e + 0.5 * vsq3(zL)) + (S_C - zL.u) * (S_C + zL.p / a_L);
    double E_sR = (zR.e + 0.5 * vsq3(zR)) + (S_C - zR.u) * (S_C + zR.p / a_R);
    double vL = zL.v, wL = zL.w, vR = zR.v, wR = zR.w;
    if (MODE == 1) {
        const double V_C = (a_R * zR.v - a_L * zL.v) / (a_R - a_L);
        const double W_C = (a_R * zR.w - a_L * zL.w) / (a_R - a_L);
        const double v_L2 = zL.v * zL.v + zL.w * zL.w;
        const double v_R2 = zR.v * zR.v + zR.w * zR.w;
        E_sL += 0.5 * ((a_R * v_R2 - a_L * v_L2) / (a_R - a_L) - v_L2);
        E_sR += 0.5 * ((a_R * v_R2 - a_L * v_L2) / (a_R - a_L) - v_R2);
        vL = vR = V_C;
        wL = wR = W_C;
    }
    const Q3 Q_sL{rho_sL, rho_sL * S_C, rho_sL * vL, rho_sL * wL, rho_sL * E_sL};
    const Q3 Q_sR{rho_sR, rho_sR * S_C, rho_sR * vR, rho_sR * wR, rho_sR * E_sR};
    const Q3 Q_L = to_q3(zL), Q_R = to_q3(zR);
    const Q3 F_L = phys_flux3(zL), F_R = phys_flux3(zR);
    double phi = 1.0;
    if (MODE == 2) {
        const double Ma_local = smax(fabs(zL.u / c_L), fabs(zR.u / c_R));
        const double xi = Ma_local / 0.1;
        phi = xi > 1.0 ? 1.0 : sin(1.5707963267948966 * xi);
    }
    const double aSC = fabs(S_C);
    Q3 F;
#pragma unroll
    for (int k = 0; k < 5; ++k) {
        double t = qk(F_L, k) + qk(F_R, k);
        if (MODE == 2) t = t + phi * S_L * (qk(Q_sL, k) - qk(Q_L, k));
        else t = t + S_L * (qk(Q_sL, k) - qk(Q_L, k));
        t = t + aSC * (qk(Q_sL, k) - qk(Q_sR, k));
        if (MODE == 2) t = t + phi * S_R * (qk(Q_sR, k) - qk(Q_R, k));
        else t = t + S_R * (qk(Q_sR, k) - qk(Q_R, k));
        const double f = 0.5 * t;
        if (isnan(f)) bad = true;   // hllc.cpp:72-80, 474-484, 558-568 throw
        qset(F, k, f);
    }
    return F;
}

// HLL::calc_flux, smf (math/cfd/flux/hll.cpp:23-57).
__device__ Q3 hll3(const Z3& zL, const Z3& zR, const Eos& eos, bool& bad) {
    const double c_L = eos_c(eos, zL.r, zL.p), c_R = eos_c(eos, zR.r, zR.p);
    const double S_L = smin(zL.u - c_L, smin(zR.u - c_R, 0.0));
    const double S_R = smax(zL.u + c_L, smax(zR.u + c_R, 0.0));
    const Q3 Q_L = to_q3(zL), Q_R = to_q3(zR);
    const Q3 F_L = phys_flux3(zL), F_R = phys_flux3(zR);
    const double SLSR = S_L * S_R, den = S_R - S_L;
    Q3 F;
#pragma unroll
    for (int k = 0; k < 5; ++k) {
        const double t = S_R * qk(F_L, k) - S_L * qk(F_R, k) + SLSR * (qk(Q_R, k) - qk(Q_L, k));
        const double f = t / den;
        if (isnan(f)) bad = true;   // hll.cpp:46-54 throws
        qset(F, k, f);
    }
    return F;
}

// Rusanov::calc_flux, smf (math/cfd/flux/rusanov.cpp:13-40).
__device__ Q3 rusanov3(const Z3& zL, const Z3& zR, const Eos& eos) {
    const double c_L = eos_c(eos, zL.r, zL.p), c_R = eos_c(eos, zR.r, zR.p);
    double S_L = zL.u - c_L;
    if (zR.u - c_R < S_L) S_L = zR.u - c_R;
    if (0.0 < S_L) S_L = 0.0;
    double S_R = zL.u + c_L;
    if (S_R < zR.u + c_R) S_R = zR.u + c_R;
    if (S_R < 0.0) S_R = 0.0;
    const double S = smax(fabs(S_L), S_R);
    const Q3 Q_L = to_q3(zL), Q_R = to_q3(zR);
    const Q3 F_L = phys_flux3(zL), F_R = phys_flux3(zR);
    const double hS = 0.5 * S;
    Q3 F;
#pragma unroll
    for (int k = 0; k < 5; ++k) qset(F, k, 0.5 * (qk(F_L, k) + qk(F_R, k)) + hS * (qk(Q_L, k) - qk(Q_R, k)));
    return F;
}

// CIR1 / CIR2 (math/cfd/flux/cir.cpp:16-216) with the sides' w.
template <bool EXT>
__device__ __forceinline__ Q3 cir3(const Z3& zL, const Z3& zR, const Eos& eos) {
    const double zl[6] = {zL.r, zL.u, zL.v, zL.w, zL.p, zL.e}, zr[6] = {zR.r, zR.u, zR.v, zR.w, zR.p, zR.e};
    double F[5];
    cir_flux5<EXT>(zl, zr, eos, F);
    return Q3{F[0], F[1], F[2], F[3], F[4]};
}

// Godunov (math/cfd/flux/godunov.cpp): the exact Riemann solution at the
// face (riemann.cpp:118-459), tangential velocities from the upwind side of
// the contact.
__device__ Q3 godunov3(const Z3& zL, const Z3& zR, const Eos& eos, bool& bad) {
    const RiemannFace sol = riemann_face(zL.r, zL.u, zL.p, zR.r, zR.u, zR.p, eos.g, eos.P0);
    bad |= sol.fail;
    Z3 z;
    z.r = sol.rho;
    z.u = sol.Uf;
    z.v = sol.U > 0.0 ? zL.v : zR.v;
    z.w = sol.U > 0.0 ? zL.w : zR.w;
    z.p = sol.Pf;
    z.e = eos_e(eos, sol.rho, sol.Pf);
    return phys_flux3(z);
}

template <int FLUX>
__device__ __forceinline__ Q3 num_flux3(const Z3& zL, const Z3& zR, const Eos& eos, bool& bad) {
    if (FLUX == 0) return godunov3(zL, zR, eos, bad);
    if (FLUX == 1) return rusanov3(zL, zR, eos);
    if (FLUX == 2) return hll3(zL, zR, eos, bad);
    if (FLUX == 6) return cir3<false>(zL, zR, eos);
    if (FLUX == 7) return cir3<true>(zL, zR, eos);
    if (FLUX == 4) return hllc3<1>(zL, zR, eos, bad);
    if (FLUX == 5) return hllc3<2>(zL, zR, eos, bad);
    return hllc3<0>(zL, zR, eos, bad);
}

// Velocity to the frame of side s (L, R, B, T, Z, F) and back.
__device__ __forceinline__ void to_local3(int s, Z3& z) {
    const double u = z.u, v = z.v, w = z.w;
    switch (s) {
        case 0: z.u = -u; z.v = -v; z.w = w; break;   // n = -x
        case 1: break;                                   // n = +x
        case 2: z.u = -v; z.v = u; z.w = w; break;      // n = -y
        case 3: z.u = v; z.v = -u; z.w = w; break;      // n = +y
        case 4: z.u = -w; z.v = v; z.w = u; break;      // n = -z
        default: z.u = w; z.v = v; z.w = -u; break;     // n = +z
    }
}
__device__ __forceinline__ void to_global3(int s, Q3& f) {
    const double x = f.mx, y = f.my, zz = f.mz;
    switch (s) {
        case 0: f.mx = -x; f.my = -y; f.mz = zz; break;
        case 1: break;
        case 2: f.mx = y; f.my = -x; f.mz = zz; break;
        case 3: f.mx = -y; f.my = x; f.mz = zz; break;
        case 4: f.mx = zz; f.my = y; f.mz = -x; break;
        default: f.mx = -zz; f.my = y; f.mz = x; break;
    }
}

// boundary_value (sm_fluid.cpp:69-79): a wall reflects the velocity
// component along the side's axis, v - 2 v.
__device__ __forceinline__ Z3 wall3(int s, const Z3& z) {
    Z3 r = z;
    if (s < 2) r.u = z.u - 2.0 * z.u;
    else if (s < 4) r.v = z.v - 2.0 * z.v;
    else r.w = z.w - 2.0 * z.w;
    return r;
}

__device__ __forceinline__ double axis_of(const double* v, int s) { return v[s >> 1]; }
__device__ __forceinline__ double sgn_of(int s) { return (s & 1) ? 1.0 : -1.0; }

// The five rows rho, u, v, w, p of a state.
__device__ __forceinline__ double zk(const Z3& z, int k) {
    return k == 0 ? z.r : (k == 1 ? z.u : (k == 2 ? z.v : (k == 3 ? z.w : z.p)));
}
__device__ __forceinline__ void zset(Z3& z, int k, double v) {
    if (k == 0) z.r = v;
    else if (k == 1) z.u = v;
    else if (k == 2) z.v = v;
    else if (k == 3) z.w = v;
    else z.p = v;
}

__device__ __forceinline__ Z3 load3(const Dev3State& s, int c) {
    return Z3{s.r[c], s.u[c], s.v[c], s.w[c], s.p[c], s.e[c]};
}
__device__ __forceinline__ void store3(const Dev3State& s, int c, const Z3& z) {
    s.r[c] = z.r; s.u[c] = z.u; s.v[c] = z.v; s.w[c] = z.w; s.p[c] = z.p; s.e[c] = z.e;
}

// FaceExtraDirect (face_extra.h:135-149) on rho, u, v, w, p: du =
// (gx dx + gy dy) + gz dz; the energy is recomputed from (rho, p).
__device__ __forceinline__ Z3 extrap3(const Z3& z, const double (*g)[5], double dx, double dy, double dz,
                                      const Eos& eos) {
    Z3 r = z;
#pragma unroll
    for (int k = 0; k < 5; ++k) zset(r, k, zk(z, k) + ((g[0][k] * dx + g[1][k] * dy) + g[2][k] * dz));
    r.e = eos_e(eos, r.r, r.p);
    return r;
}

// The CFL minimum of slot ctl.dtmin (the least of its partial minima), then
// dt = min(CFL dmin, max_dt) (sm_fluid.cpp:104-114).  Every thread of the
// workgroup calls it.
__device__ __forceinline__ double block_dt(const Dev3Ctl& ctl) {
    __shared__ double s_min;
    if (threadIdx.x < kSubSlots) {
        double v = __longlong_as_double((long long)ctl.dtmin[kSubStride * threadIdx.x]);
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) {
            const double o = __shfl_xor(v, off, 64);
            v = (o < v) ? o : v;
        }
        if (threadIdx.x == 0) s_min = v;
    }
    __syncthreads();
    return smin(ctl.cfl * s_min, ctl.max_dt);
}

// Workgroup 0 publishes dt and clears the slot the previous step read.
__device__ __forceinline__ void step_epilogue0(const Dev3Ctl& ctl, double dt) {
    if (blockIdx.x != 0) return;
    if (threadIdx.x == 0) *ctl.dt_out = dt;
    if (threadIdx.x < kSubSlots) ctl.reset[kSubStride * threadIdx.x] = kDblMaxBits;
}

// The least candidate of a workgroup of 256 into its partial minimum of the
// slot, one atomic per workgroup; infinities and NaNs are left out, as the reference's
// `if (t < res)` over res = DBL_MAX leaves them.  Every thread of the
// workgroup calls it.
__device__ __forceinline__ void block_min(unsigned long long* slot, double v) {
    if (!(v < INFINITY)) v = INFINITY;
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        const double o = __shfl_xor(v, off, 64);
        v = (o < v) ? o : v;
    }
    __shared__ double red[4];
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        double m = red[0];
        for (int i = 1; i < 4; ++i) m = (red[i] < m) ? red[i] : m;
        if (m < INFINITY)
            atomicMin(slot + kSubStride * (blockIdx.x & (kSubSlots - 1)), (unsigned long long)__double_as_longlong(m));
    }
}

// The cell of this thread: workgroups dealt out so that each XCD (blockIdx
// mod 8) takes one contiguous eighth of the cells, whose y and z neighbours
// (and most x neighbours) then sit in the same L2.
__device__ __forceinline__ int cell_of_thread() {
    return xcd_block((int)blockIdx.x, (int)gridDim.x) * (int)blockDim.x + (int)threadIdx.x;
}

// h / (|v| + c) (sm_fluid.cpp:105-110).
__device__ __forceinline__ double cell_dt3(const Z3& q, double h, const Eos& eos) {
    return h / (sqrt(vsq3(q)) + eos_c(eos, q.r, q.p));
}

// The geometry of cell c: the slot arrays, or (TAB) the Cuboid's tables.
template <bool TAB>
struct Geo3 {
    const Dev3Mesh& M;
    int c, i, j, k, L;
    __device__ __forceinline__ Geo3(const Dev3Mesh& m, int cell) : M(m), c(cell), i(0), j(0), k(0), L(0) {
        if constexpr (TAB) {
            const int plane = M.tny * M.tnz;
            i = c / plane;
            const int r = c - i * plane;
            j = r / M.tnz;
            k = r - j * M.tnz;
            L = M.tnx + M.tny + M.tnz;
        }
    }
    __device__ __forceinline__ int idx(int a) const { return a == 0 ? i : (a == 1 ? j : k); }
    __device__ __forceinline__ int len(int a) const { return a == 0 ? M.tnx : (a == 1 ? M.tny : M.tnz); }
    __device__ __forceinline__ int base(int a) const { return a == 0 ? 0 : (a == 1 ? M.tnx : M.tnx + M.tny); }
    __device__ __forceinline__ int stride(int a) const { return a == 0 ? M.tny * M.tnz : (a == 1 ? M.tnz : 1); }
    __device__ __forceinline__ double centre(int a) const {
        if constexpr (TAB) return M.tab[18 * L + base(a) + idx(a)];
        return a == 0 ? M.cx[c] : (a == 1 ? M.cy[c] : M.cz[c]);
    }
    __device__ __forceinline__ double vol() const {
        if constexpr (TAB) return M.tab[25 * L + 6];
        return M.vol[c];
    }
    __device__ __forceinline__ bool edge(int s) const {
        const int a = s >> 1;
        return (s & 1) ? idx(a) == len(a) - 1 : idx(a) == 0;
    }
    __device__ __forceinline__ int type(int s) const {
        if constexpr (TAB) return edge(s) ? M.bcode[s] : SC_INNER;
        return M.type[s * M.stride + c];
    }
    __device__ __forceinline__ int adj(int s) const {
        if constexpr (TAB) {
            const int a = s >> 1, st = stride(a), up = s & 1;
            if (!edge(s)) return up ? c + st : c - st;
            if (!M.per[a]) return c;
            return up ? c - (len(a) - 1) * st : c + (len(a) - 1) * st;
        }
        return M.adj[s * M.stride + c];
    }
    __device__ __forceinline__ double fc(int s, int a) const {
        if constexpr (TAB) return M.tab[(3 * s + a) * L + base(a) + idx(a)];
        const int so = s * M.stride + c;
        return a == 0 ? M.fcx[so] : (a == 1 ? M.fcy[so] : M.fcz[so]);
    }
    __device__ __forceinline__ double area(int s) const {
        if constexpr (TAB) return M.tab[25 * L + s];
        return M.area[s * M.stride + c];
    }
    __device__ __forceinline__ double w(int s) const {
        if constexpr (TAB) return M.tab[19 * L + s * L + base(s >> 1) + idx(s >> 1)];
        return M.w[s * M.stride + c];
    }
    // component a of the centre of the neighbour nb across inner side s
    __device__ __forceinline__ double ncentre(int s, int nb, int a) const {
        if constexpr (TAB) {
            if (a != (s >> 1)) return centre(a);
            return M.tab[18 * L + base(a) + idx(a) + ((s & 1) ? 1 : -1)];
        }
        return a == 0 ? M.cx[nb] : (a == 1 ? M.cy[nb] : M.cz[nb]);
    }
};

}  // namespace

// SmFluid::compute_dt (sm_fluid.cpp:104-114): h / (|v| + c), min over cells.
__global__ void __launch_bounds__(256) k3_dt(Dev3Mesh M, Dev3State z, Eos eos, unsigned long long* slot) {
    const int c = cell_of_thread();
    double cand = INFINITY;
    if (c < M.n) cand = cell_dt3(load3(z, c), M.h[c], eos);
    block_min(slot, cand);
}

// SmFluid::fluxes, first order (sm_fluid.cpp:127-174); returns the next
// step's CFL candidate of the new state.
template <int FLUX, bool TAB>
__device__ __forceinline__ double flux1_cell(const Dev3Mesh& M, const Dev3State& in, const Dev3State& out,
                                             const Eos& eos, const Dev3Ctl& ctl, double dt, int c) {
    const Geo3<TAB> geo(M, c);
    const Z3 zc = load3(in, c);
    Q3 F{0.0, 0.0, 0.0, 0.0, 0.0};
    bool bad = false;
    for (int s = 0; s < 6; ++s) {
        const int so = s * M.stride + c;
        const int t = geo.type(s);
        Z3 zn = zc;
        if (t == SC_INNER || t == SC_PERIODIC) zn = load3(in, geo.adj(s));
        else if (t == SC_WALL) zn = wall3(s, zc);
        Z3 lm = zc, lp = zn;
        to_local3(s, lm);
        to_local3(s, lp);
        Q3 f = num_flux3<FLUX>(lm, lp, eos, bad);
        to_global3(s, f);
        const double a = geo.area(s);
        F.r = F.r + f.r * a;
        F.mx = F.mx + f.mx * a;
        F.my = F.my + f.my * a;
        F.mz = F.mz + f.mz * a;
        F.E = F.E + f.E * a;
    }
    Q3 q = to_q3(zc);
    const double coef = dt / geo.vol();
    q.r = q.r - coef * F.r;
    q.mx = q.mx - coef * F.mx;
    q.my = q.my - coef * F.my;
    q.mz = q.mz - coef * F.mz;
    q.E = q.E - coef * F.E;
    const Z3 zn = to_z3(q, eos);
    store3(out, c, zn);
    if (bad) *ctl.err = 1;
    return cell_dt3(zn, M.h[c], eos);
}

template <int FLUX, bool TAB>
__global__ void __launch_bounds__(256) k3_flux1(Dev3Mesh M, Dev3State in, Dev3State out, Eos eos, Dev3Ctl ctl) {
    const int c = cell_of_thread();
    const double dt = block_dt(ctl);
    step_epilogue0(ctl, dt);
    const double cand = c < M.n ? flux1_cell<FLUX, TAB>(M, in, out, eos, ctl, dt, c) : INFINITY;
    block_min(ctl.next, cand);
}

// SmFluid::compute_grad (gradient::LSM, gradient.h:226-282, then
// gradient::limiting, gradient.h:412-474) fused with fluxes_stage1
// (sm_fluid.cpp:176-253).  G: [3][5][stride], d/dx, d/dy, d/dz of rho, u,
// v, w, p.
template <int LIM, bool TAB>
__global__ void __launch_bounds__(256) k3_grad_pred(Dev3Mesh M, Dev3State in, double* __restrict__ G, Dev3State half,
                                                    const int* __restrict__ wait, Eos eos, Dev3Ctl ctl) {
    const int c = cell_of_thread();
    const double dt = block_dt(ctl);
    if (c >= M.n) return;
    const Geo3<TAB> geo(M, c);
    const Z3 zc = load3(in, c);
    const double cc[3] = {geo.centre(0), geo.centre(1), geo.centre(2)};
    // per slot: dr along its axis (zero across), the state across it
    double dra[6], w[6];
    Z3 zn[6];
#pragma unroll
    for (int s = 0; s < 6; ++s) {
        const int so = s * M.stride + c;
        const int t = geo.type(s);
        const int nb = geo.adj(s);
        w[s] = geo.w(s);
        zn[s] = (t == SC_INNER || t == SC_PERIODIC) ? load3(in, nb) : (t == SC_WALL ? wall3(s, zc) : zc);
        const double fc[3] = {geo.fc(s, 0), geo.fc(s, 1), geo.fc(s, 2)};
        const double na = sgn_of(s);
        if (t == SC_INNER) {
            const double nc[3] = {geo.ncentre(s, nb, 0), geo.ncentre(s, nb, 1), geo.ncentre(s, nb, 2)};
            dra[s] = axis_of(nc, s) - axis_of(cc, s);
        } else {   // symm_point: p + (2 (f - p).n) n
            const double d = 2.0 * ((axis_of(fc, s) - axis_of(cc, s)) * na);
            dra[s] = (axis_of(cc, s) + d * na) - axis_of(cc, s);
        }
    }
    // A (diagonal) and F: sums over the slots in slot order
    double A[3] = {0.0, 0.0, 0.0};
    double F[3][5] = {};
#pragma unroll
    for (int s = 0; s < 6; ++s) {
        const int a = s >> 1;
        const double na = sgn_of(s);
        const double drn = (dra[s] * na) * na;
        A[a] = A[a] + (w[s] * drn) * dra[s];
#pragma unroll
        for (int k = 0; k < 5; ++k) F[a][k] = F[a][k] + (w[s] * (zk(zn[s], k) - zk(zc, k))) * drn;
    }
    if (fabs(A[2]) < 1e-8) {
        A[2] = 1.0;
        if (fabs(A[1]) < 1e-8) A[1] = 1.0;
    }
    const double c00 = A[1] * A[2], c11 = A[2] * A[0], c22 = A[0] * A[1];
    const double invdet = 1.0 / (c00 * A[0]);
    const double B[3] = {c00 * invdet, c11 * invdet, c22 * invdet};
    double g[3][5];
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
        for (int k = 0; k < 5; ++k) g[a][k] = B[a] * F[a][k];
    // limiting: limiter(2 du_n - dF, dF) dF, re-solved with the same A
    double L[3][5] = {};
#pragma unroll
    for (int s = 0; s < 6; ++s) {
        const int a = s >> 1;
        const double na = sgn_of(s);
        const double drn = (dra[s] * na) * na;
#pragma unroll
        for (int k = 0; k < 5; ++k) {
            const double dF = zk(zn[s], k) - zk(zc, k);
            const double du_n = g[a][k] * dra[s];
            const double lim = limiter2<LIM>(2.0 * du_n - dF, dF) * dF;
            L[a][k] = L[a][k] + (w[s] * lim) * drn;
        }
    }
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
        for (int k = 0; k < 5; ++k) {
            g[a][k] = B[a] * L[a][k];
            G[(size_t)(a * 5 + k) * M.stride + c] = g[a][k];
        }
    // predictor
    Z3 zh = zc;
    if (wait[c] <= 0) {
        Q3 Fl{0.0, 0.0, 0.0, 0.0, 0.0};
        for (int s = 0; s < 6; ++s) {
            const int so = s * M.stride + c;
            Z3 zm = extrap3(zc, g, geo.fc(s, 0) - cc[0], geo.fc(s, 1) - cc[1], geo.fc(s, 2) - cc[2], eos);
            if (is_bad3(zm, eos)) zm = zc;
            to_local3(s, zm);
            Q3 f = phys_flux3(zm);
            to_global3(s, f);
            const double a = geo.area(s);
            Fl.r = Fl.r + f.r * a;
            Fl.mx = Fl.mx + f.mx * a;
            Fl.my = Fl.my + f.my * a;
            Fl.mz = Fl.mz + f.mz * a;
            Fl.E = Fl.E + f.E * a;
        }
        Q3 q = to_q3(zc);
        const double coef = 0.5 * dt / geo.vol();
        q.r = q.r - coef * Fl.r;
        q.mx = q.mx - coef * Fl.mx;
        q.my = q.my - coef * Fl.my;
        q.mz = q.mz - coef * Fl.mz;
        q.E = q.E - coef * Fl.E;
        zh = to_z3(q, eos);
        if (is_bad3(zh, eos)) zh = zc;
    }
    store3(half, c, zh);
}

// SmFluid::fluxes_stage2, corrector (sm_fluid.cpp:255-366): wait read from
// the pre-pass array, written to the other; returns the next step's CFL
// candidate of the new state.
template <int FLUX, bool TAB>
__device__ __forceinline__ double corr_cell(const Dev3Mesh& M, const Dev3State& init, const Dev3State& half,
                                            const double* __restrict__ G, const int* __restrict__ wait_in,
                                            int* __restrict__ wait_out, const Dev3State& next, const Eos& eos,
                                            const Dev3Ctl& ctl, double dt, int c) {
    const Geo3<TAB> geo(M, c);
    const Z3 zc = load3(init, c);
    const Z3 zch = load3(half, c);
    const double cc[3] = {geo.centre(0), geo.centre(1), geo.centre(2)};
    const int wc = wait_in[c];
    double g[3][5];
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
        for (int k = 0; k < 5; ++k) g[a][k] = G[(size_t)(a * 5 + k) * M.stride + c];
    // the six slots' codes, neighbours and neighbours' counters first: one
    // round trip for all of them instead of two per face
    int ts[6], nbs[6];
#pragma unroll
    for (int s = 0; s < 6; ++s) {
        ts[s] = geo.type(s);
        nbs[s] = geo.adj(s);
    }
    bool rob[6];
#pragma unroll
    for (int s = 0; s < 6; ++s) rob[s] = wc > 0 || wait_in[nbs[s]] > 0;
    Q3 F{0.0, 0.0, 0.0, 0.0, 0.0};
    bool bad = false;
#pragma unroll
    for (int s = 0; s < 6; ++s) {
        const int so = s * M.stride + c;
        const int t = ts[s];
        const int nb = nbs[s];
        const bool bound = t >= SC_WALL;
        const double fc[3] = {geo.fc(s, 0), geo.fc(s, 1), geo.fc(s, 2)};
        const Z3 znh = bound ? (t == SC_WALL ? wall3(s, zch) : zch) : load3(half, nb);
        const bool robust = rob[s];
        Z3 zm, zp;
        if (robust) {
            zm = zch;
            zp = znh;
        } else {
            double nc[3];
            if (t == SC_INNER) {
                nc[0] = geo.ncentre(s, nb, 0); nc[1] = geo.ncentre(s, nb, 1); nc[2] = geo.ncentre(s, nb, 2);
            } else {   // the mirror point of the cell centre
                const double na = sgn_of(s);
                const double d = 2.0 * ((axis_of(fc, s) - axis_of(cc, s)) * na);
                nc[0] = cc[0]; nc[1] = cc[1]; nc[2] = cc[2];
                nc[s >> 1] = axis_of(cc, s) + d * na;
            }
            zm = extrap3(zch, g, fc[0] - cc[0], fc[1] - cc[1], fc[2] - cc[2], eos);
            if (is_bad3(zm, eos)) zm = zch;
            if (!bound) {
                double gn[3][5];
#pragma unroll
                for (int a = 0; a < 3; ++a)
#pragma unroll
                    for (int k = 0; k < 5; ++k) gn[a][k] = G[(size_t)(a * 5 + k) * M.stride + nb];
                zp = extrap3(znh, gn, fc[0] - nc[0], fc[1] - nc[1], fc[2] - nc[2], eos);
                if (is_bad3(zp, eos)) zp = znh;
            } else {
                zp = t == SC_WALL ? wall3(s, zm) : zm;
            }
        }
        to_local3(s, zm);
        to_local3(s, zp);
        Q3 f = robust ? hll3(zm, zp, eos, bad) : num_flux3<FLUX>(zm, zp, eos, bad);
        to_global3(s, f);
        const double a = geo.area(s);
        F.r = F.r + f.r * a;
        F.mx = F.mx + f.mx * a;
        F.my = F.my + f.my * a;
        F.mz = F.mz + f.mz * a;
        F.E = F.E + f.E * a;
    }
    Q3 q = to_q3(zc);
    const double coef = dt / geo.vol();
    q.r = q.r - coef * F.r;
    q.mx = q.mx - coef * F.mx;
    q.my = q.my - coef * F.my;
    q.mz = q.mz - coef * F.mz;
    q.E = q.E - coef * F.E;
    Z3 zn = to_z3(q, eos);
    int wn;
    if (is_bad3(zn, eos)) {
        zn = zc;
        wn = 3;
    } else {
        wn = wc > 0 ? wc - 1 : wc;
    }
    store3(next, c, zn);
    wait_out[c] = wn;
    if (bad) *ctl.err = 1;
    return cell_dt3(zn, M.h[c], eos);
}

template <int FLUX, bool TAB>
__global__ void __launch_bounds__(256) k3_corr(Dev3Mesh M, Dev3State init, Dev3State half, const double* __restrict__ G,
                                               const int* __restrict__ wait_in, int* __restrict__ wait_out,
                                               Dev3State next, Eos eos, Dev3Ctl ctl) {
    const int c = cell_of_thread();
    const double dt = block_dt(ctl);
    step_epilogue0(ctl, dt);
    const double cand =
        c < M.n ? corr_cell<FLUX, TAB>(M, init, half, G, wait_in, wait_out, next, eos, ctl, dt, c) : INFINITY;
    block_min(ctl.next, cand);
}

// Per-workgroup partial sums of V q: mass, momentum x, y, z, energy.
__global__ void __launch_bounds__(256) k3_totals(Dev3Mesh M, Dev3State z, double* partial) {
    const int c = cell_of_thread();
    double v[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
    if (c < M.n) {
        const Q3 q = to_q3(load3(z, c));
        const double V = M.vol[c];
        v[0] = V * q.r; v[1] = V * q.mx; v[2] = V * q.my; v[3] = V * q.mz; v[4] = V * q.E;
    }
    __shared__ double red[5][256];
#pragma unroll
    for (int k = 0; k < 5; ++k) red[k][threadIdx.x] = v[k];
    __syncthreads();
    for (int off = 128; off > 0; off >>= 1) {
        if ((int)threadIdx.x < off) {
#pragma unroll
            for (int k = 0; k < 5; ++k) red[k][threadIdx.x] += red[k][threadIdx.x + off];
        }
        __syncthreads();
    }
    if (threadIdx.x < 5) partial[5 * blockIdx.x + threadIdx.x] = red[threadIdx.x][0];
}

// ---------------------------------------------------------------- launchers
static inline dim3 g256(int n) { return dim3((unsigned)((n + 255) / 256)); }

hipError_t launch3_dt(const Dev3Mesh& M, const Dev3State& z, const Eos& eos, unsigned long long* slot,
                      hipStream_t s) {
    hipLaunchKernelGGL(k3_dt, g256(M.n), dim3(256), 0, s, M, z, eos, slot);
    return hipGetLastError();
}

hipError_t launch3_flux1(int flux, const Dev3Mesh& M, const Dev3State& in, const Dev3State& out, const Eos& eos,
                         const Dev3Ctl& ctl, hipStream_t s) {
    const bool tab = M.tnx > 0;
#define ZFV3_F1(F)                                                                                   \
    if (tab) hipLaunchKernelGGL((k3_flux1<F, true>), g256(M.n), dim3(256), 0, s, M, in, out, eos, ctl); \
    else hipLaunchKernelGGL((k3_flux1<F, false>), g256(M.n), dim3(256), 0, s, M, in, out, eos, ctl)
    switch (flux) {
        case 0: ZFV3_F1(0); break;
        case 1: ZFV3_F1(1); break;
        case 6: ZFV3_F1(6); break;
        case 7: ZFV3_F1(7); break;
        case 2: ZFV3_F1(2); break;
        case 4: ZFV3_F1(4); break;
        case 5: ZFV3_F1(5); break;
        default: ZFV3_F1(3); break;
    }
#undef ZFV3_F1
    return hipGetLastError();
}

hipError_t launch3_grad_pred(int limiter, const Dev3Mesh& M, const Dev3State& in, double* G, const Dev3State& half,
                             const int* wait, const Eos& eos, const Dev3Ctl& ctl, hipStream_t s) {
    const bool tab = M.tnx > 0;
#define ZFV3_GP(L)                                                                                          \
    if (tab) hipLaunchKernelGGL((k3_grad_pred<L, true>), g256(M.n), dim3(256), 0, s, M, in, G, half, wait, eos, ctl); \
    else hipLaunchKernelGGL((k3_grad_pred<L, false>), g256(M.n), dim3(256), 0, s, M, in, G, half, wait, eos, ctl)
    switch (limiter) {
        case 0: ZFV3_GP(0); break;
        case 1: ZFV3_GP(1); break;
        case 2: ZFV3_GP(2); break;
        case 3: ZFV3_GP(3); break;
        case 4: ZFV3_GP(4); break;
        case 5: ZFV3_GP(5); break;
        case 6: ZFV3_GP(6); break;
        case 7: ZFV3_GP(7); break;
        case 8: ZFV3_GP(8); break;
        case 9: ZFV3_GP(9); break;
        case 10: ZFV3_GP(10); break;
        case 11: ZFV3_GP(11); break;
        default: ZFV3_GP(12); break;
    }
#undef ZFV3_GP
    return hipGetLastError();
}

hipError_t launch3_corr(int flux, const Dev3Mesh& M, const Dev3State& init, const Dev3State& half, const double* G,
                        const int* wait_in, int* wait_out, const Dev3State& next, const Eos& eos, const Dev3Ctl& ctl,
                        hipStream_t s) {
    const bool tab = M.tnx > 0;
#define ZFV3_C(F)                                                                                               \
    if (tab) hipLaunchKernelGGL((k3_corr<F, true>), g256(M.n), dim3(256), 0, s, M, init, half, G, wait_in, wait_out, \
                                next, eos, ctl);                                                                  \
    else hipLaunchKernelGGL((k3_corr<F, false>), g256(M.n), dim3(256), 0, s, M, init, half, G, wait_in, wait_out,  \
                            next, eos, ctl)
    switch (flux) {
        case 0: ZFV3_C(0); break;
        case 1: ZFV3_C(1); break;
        case 6: ZFV3_C(6); break;
        case 7: ZFV3_C(7); break;
        case 2: ZFV3_C(2); break;
        case 4: ZFV3_C(4); break;
        case 5: ZFV3_C(5); break;
        default: ZFV3_C(3); break;
    }
#undef ZFV3_C
    return hipGetLastError();
}

hipError_t launch3_totals(const Dev3Mesh& M, const Dev3State& z, double* partial, int nblocks, hipStream_t s) {
    hipLaunchKernelGGL(k3_totals, dim3(nblocks), dim3(256), 0, s, M, z, partial);
    return hipGetLastError();
}

}  // namespace zfv
