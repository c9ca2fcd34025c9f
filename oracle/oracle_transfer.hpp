// CPU restatement of zephyr's Transfer solver (math/solver/transfer.h:21-153,
// transfer.cpp:1-719): the transport of a volume fraction u1 by a given
// velocity field with the CRP heuristic and normal-aware face fractions,
// the MUSCL variants with their CRP correction, the interface normals of
// InterfaceRecovery::compute_normals (geom/interface_recovery.cpp:17-45),
// compute_dt, set_flags and the split / merge distributor.  The velocity
// is given as samples at the face centres and cell centres (the values the
// reference's virtual velocity() returns there).  Test infrastructure:
// included at the end of oracle.cpp.
//
// Built methods (Transfer::Method numbering): CRP_V3 0, CRP_V5 1, CRP_SE 2,
// CRP_N1 3, MUSCLd 7, MUSCLd_CRP 9, MUSCL_MC 11, MUSCL_MC_CRP 12, WENO 13,
// WENO_CRP 14 (on a Rectangle's structured grid).  CRP_N2, VOF, VOF_CRP,
// MUSCLn, MUSCLn_CRP (sub-cell polygon reconstruction) are not.

struct orc_tr {
    orc_mesh* m = nullptr;
    int method = 11;
    double cfl = 0.5, dt = 1.0e300;   // Transfer::Transfer (transfer.cpp:43-48)
    int limiter = ORC_LIM_MC;        // m_limiter = "MC"
    std::vector<double> u1, u2, nx, ny, dudx, dudy;
    std::vector<double> fvx, fvy;    // velocity at the face centres [8n]
    std::vector<double> cvx, cvy;    // velocity at the cell centres [n]
};

namespace {

enum {
    kTrV3 = 0, kTrV5 = 1, kTrSE = 2, kTrN1 = 3, kTrN2 = 4, kTrVOF = 5, kTrVOFCRP = 6, kTrMUSCLd = 7, kTrMUSCLn = 8,
    kTrMUSCLdCRP = 9, kTrMUSCLnCRP = 10, kTrMC = 11, kTrMCCRP = 12, kTrWENO = 13, kTrWENOCRP = 14
};

inline bool tr_built(int k) {
    return k == kTrV3 || k == kTrV5 || k == kTrSE || k == kTrN1 || k == kTrMUSCLd || k == kTrMUSCLdCRP ||
           k == kTrMC || k == kTrMCCRP || k == kTrWENO || k == kTrWENOCRP;
}

// WENO<b, e> (math/calc/weno.h:39-120) on three cell averages a, b, c of a
// uniform grid: the face values u(+1/2) (p) and u(-1/2) (m) of the
// quadratic interpolant, w1 a + w2 b + w3 c summed left to right, and the
// smoothness indicator 13/12 d2^2 + 1/4 d1^2.
inline double weno_p_m2(double a, double b, double c) { return ((1.0 / 3.0) * a + (-7.0 / 6.0) * b) + (11.0 / 6.0) * c; }
inline double weno_p_m1(double a, double b, double c) { return ((-1.0 / 6.0) * a + (5.0 / 6.0) * b) + (1.0 / 3.0) * c; }
inline double weno_p_0(double a, double b, double c) { return ((1.0 / 3.0) * a + (5.0 / 6.0) * b) + (-1.0 / 6.0) * c; }
inline double weno_m_m2(double a, double b, double c) { return ((-1.0 / 6.0) * a + (5.0 / 6.0) * b) + (1.0 / 3.0) * c; }
inline double weno_m_m1(double a, double b, double c) { return ((1.0 / 3.0) * a + (5.0 / 6.0) * b) + (-1.0 / 6.0) * c; }
inline double weno_m_0(double a, double b, double c) { return ((11.0 / 6.0) * a + (-7.0 / 6.0) * b) + (1.0 / 3.0) * c; }
inline double weno_beta(double d1, double d2) { return (13.0 / 12.0) * d2 * d2 + (1.0 / 4.0) * d1 * d1; }
inline double weno_b_m2(double a, double b, double c) { return weno_beta((a - 4.0 * b) + 3.0 * c, (a - 2.0 * b) + c); }
inline double weno_b_m1(double a, double b, double c) { return weno_beta(a - c, (a - 2.0 * b) + c); }
inline double weno_b_0(double a, double b, double c) { return weno_beta((3.0 * a - 4.0 * b) + c, (a - 2.0 * b) + c); }

// WENO5 (weno.h:126-170) on u1..u5: the sub-stencils we2 = {u1, u2, u3},
// we1 = {u2, u3, u4}, we0 = {u3, u4, u5}, linear weights 3/10, 6/10, 1/10
// (p) or 1/10, 6/10, 3/10 (m) divided by (1e-8 + beta)^2 and normalised.
inline double weno5(const double* u, bool right) {
    const double b0 = weno_b_0(u[2], u[3], u[4]);
    const double b1 = weno_b_m1(u[1], u[2], u[3]);
    const double b2 = weno_b_m2(u[0], u[1], u[2]);
    const double d0c = right ? 3.0 / 10.0 : 1.0 / 10.0, d1c = 6.0 / 10.0, d2c = right ? 1.0 / 10.0 : 3.0 / 10.0;
    const double e0 = 1.0e-8 + b0, e1 = 1.0e-8 + b1, e2 = 1.0e-8 + b2;
    const double d0 = d0c / (e0 * e0), d1 = d1c / (e1 * e1), d2 = d2c / (e2 * e2);
    const double w0 = d0 / (d0 + d1 + d2), w1 = d1 / (d0 + d1 + d2), w2 = d2 / (d0 + d1 + d2);
    if (right) return (w0 * weno_p_0(u[2], u[3], u[4]) + w1 * weno_p_m1(u[1], u[2], u[3])) + w2 * weno_p_m2(u[0], u[1], u[2]);
    return (w0 * weno_m_0(u[2], u[3], u[4]) + w1 * weno_m_m1(u[1], u[2], u[3])) + w2 * weno_m_m2(u[0], u[1], u[2]);
}

// flux_2D (transfer.cpp:110-126): the volume through a face with the face
// fraction as, bounded by what the upwind cell holds and can take.
inline double tr_flux_2D(double a1, double a2, double S, double V1, double V2, double as, double vn, double dt) {
    const double a = vn > 0.0 ? a1 : a2;
    const double V = vn > 0.0 ? V1 : V2;
    const double gamma = dt * std::abs(vn) * S;
    const double F_min = smax(0.0, gamma - (1.0 - a) * V);
    const double F_max = a * V;
    return f_sign(vn) * f_between(gamma * as, F_min, F_max);
}

// face_fraction_n1 (transfer.cpp:131-143).
inline double tr_face_fraction_n1(double a1, double a2, double n1x, double n1y, double n2x, double n2y, double fnx,
                                  double fny, double vn, double V1, double V2, double S, double dt) {
    const double a = vn > 0.0 ? a1 : a2;
    const double V = vn > 0.0 ? V1 : V2;
    const double cs = vn > 0.0 ? (n1x * fnx + n1y * fny) + 0.0 : -((n2x * fnx + n2y * fny) + 0.0);
    const double C = dt * std::abs(vn) * S / V;
    const double a_sig = average_flux(a, cs, C);
    return f_between(a_sig, a1, a2);
}

// best_face_fraction (transfer.cpp:183-196) and flux_CRP (:198-206): the
// face fraction that reproduces Flux, then flux_2D with it.
inline double tr_flux_CRP(double a1, double a2, double S, double V1, double V2, double vn, double dt, double Flux) {
    double a_sig;
    if (vn == 0.0) {
        a_sig = 0.5 * (a1 + a2);
    } else {
        double lo, hi;
        f_sorted(a1, a2, lo, hi);
        a_sig = f_between(Flux / (dt * vn * S), lo, hi);
    }
    return tr_flux_2D(a1, a2, S, V1, V2, a_sig, vn, dt);
}

// limiters::vanAlbada2 (math/cfd/limiter.cxx), the two-argument form used by
// FaceExtraATvL.
inline double tr_van_albada2(double a, double b) { return lim_binary(ORC_LIM_VANALBADA2, a, b); }

// gradient::LSM<double> and gradient::limiting<double> (gradient.h:226-282,
// 412-474) of u1 with boundary_value = u (transfer.cpp:453-458).
void tr_grad(const orc_tr& T, int c, bool limit, double& gx_out, double& gy_out) {
    const orc_mesh& m = *T.m;
    const double w0 = 1.0;
    const double uc = T.u1[c];
    const Vec3 cell_c{m.cx[c], m.cy[c], 0.0};
    double A[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
    double Fx = 0, Fy = 0, Fz = 0;
    struct G { Vec3 dr, drn; double w, un; } fg[8];
    int nf = 0;
    for (int s8 = 0; s8 < 8; ++s8) {
        const size_t s = 8 * (size_t)c + s8;
        const int b = m.bnd[s];
        if (b == ORC_UNDEFINED) continue;
        const Vec3 normal{m.fnx[s], m.fny[s], 0.0};
        const double un = !slot_is_boundary(b) ? T.u1[m.adj[s]] : uc;
        const Vec3 nc = (b == ORC_INNER) ? Vec3{m.cx[m.adj[s]], m.cy[m.adj[s]], 0.0} : symm_point(m, (int)s, cell_c);
        const Vec3 dr = nc - cell_c;
        const Vec3 drn = scale(dot(dr, normal), normal);
        const double w = w0 * m.area[s] / std::pow(norm(drn), 3);
        const double wd[3] = {w * drn.x, w * drn.y, w * drn.z};
        const double drv[3] = {dr.x, dr.y, dr.z};
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) A[i][j] = A[i][j] + wd[i] * drv[j];
        const double dF = un - uc;
        Fx = Fx + (w * dF) * drn.x;
        Fy = Fy + (w * dF) * drn.y;
        Fz = Fz + (w * dF) * drn.z;
        fg[nf++] = {dr, drn, w, un};
    }
    if (std::abs(A[2][2]) < 1e-8) {
        A[2][2] = 1.0;
        if (std::abs(A[1][1]) < 1e-8) A[1][1] = 1.0;
    }
    auto cof = [&](int i, int j) {
        int i1 = (i + 1) % 3, i2 = (i + 2) % 3, j1 = (j + 1) % 3, j2 = (j + 2) % 3;
        return A[i1][j1] * A[i2][j2] - A[i1][j2] * A[i2][j1];
    };
    const double c00 = cof(0, 0), c10 = cof(1, 0), c20 = cof(2, 0);
    const double det = (c00 * A[0][0] + c10 * A[1][0]) + c20 * A[2][0];
    const double invdet = 1.0 / det;
    double B[3][3];
    B[0][0] = c00 * invdet; B[0][1] = c10 * invdet; B[0][2] = c20 * invdet;
    B[1][0] = cof(0, 1) * invdet; B[1][1] = cof(1, 1) * invdet; B[1][2] = cof(2, 1) * invdet;
    B[2][0] = cof(0, 2) * invdet; B[2][1] = cof(1, 2) * invdet; B[2][2] = cof(2, 2) * invdet;
    const double gx = (B[0][0] * Fx + B[0][1] * Fy) + B[0][2] * Fz;
    const double gy = (B[1][0] * Fx + B[1][1] * Fy) + B[1][2] * Fz;
    const double gz = (B[2][0] * Fx + B[2][1] * Fy) + B[2][2] * Fz;
    if (!limit) {
        gx_out = gx;
        gy_out = gy;
        return;
    }
    double Lx = 0, Ly = 0, Lz = 0;
    for (int f = 0; f < nf; ++f) {
        const G& q = fg[f];
        const double dF_p = q.un - uc;
        const double du_n = (gx * q.dr.x + gy * q.dr.y) + gz * q.dr.z;
        const double dF_m = 2.0 * du_n - dF_p;
        const double dF_lim = lim_binary(T.limiter, dF_m, dF_p) * dF_p;
        Lx = Lx + (q.w * dF_lim) * q.drn.x;
        Ly = Ly + (q.w * dF_lim) * q.drn.y;
        Lz = Lz + (q.w * dF_lim) * q.drn.z;
    }
    gx_out = (B[0][0] * Lx + B[0][1] * Ly) + B[0][2] * Lz;
    gy_out = (B[1][0] * Lx + B[1][1] * Ly) + B[1][2] * Lz;
}

// update_WENO's flux sum of one cell (transfer.cpp:541-658) on the
// Rectangle's grid, cell = gny i + j, indices wrapped periodically
// (EuMesh::operator()(i, j)): for each swept side L, R, B, T (no boundary
// test: a boundary face takes the wrapped stencil too) the WENO5 face value
// of the upwind cell -- the cell itself when vn > 0, else the neighbour
// across the side -- on its five cells along the side's axis, m() on its
// lower face and p() on its upper one, bounded to [0, 1]; the flux
// a_sig vn dt S, with WENO_CRP corrected by flux_CRP.  Returns u2.
double tr_cell_weno(const orc_tr& T, int c, int dir) {
    const orc_mesh& m = *T.m;
    const int nx = m.gnx, ny = m.gny;
    const int i = c / ny, j = c % ny;
    const double a1 = T.u1[c];
    const double V1 = m.vol[c];
    double fluxes = 0.0;
    for (int s = 0; s < 4; ++s) {
        if ((s < 2 && dir == 2) || (s >= 2 && dir == 1)) continue;
        const size_t f = 8 * (size_t)c + s;
        const int nb = m.adj[f];
        const double vn = (T.fvx[f] * m.fnx[f] + T.fvy[f] * m.fny[f]) + 0.0;
        const bool upper = (s & 1) != 0;
        const int k = s < 2 ? i : j;
        const int K = vn > 0.0 ? k : (upper ? k + 1 : k - 1);
        double st[5];
        for (int q = 0; q < 5; ++q) {
            const int kk = K - 2 + q;
            const int ii = s < 2 ? (kk + nx) % nx : i, jj = s < 2 ? j : (kk + ny) % ny;
            st[q] = T.u1[(size_t)ny * ii + jj];
        }
        const bool right = upper ? vn > 0.0 : !(vn > 0.0);
        const double a_sig = f_between(weno5(st, right), 0.0, 1.0);
        const double S = m.area[f];
        double Flux = a_sig * vn * T.dt * S;
        if (T.method == kTrWENOCRP) Flux = tr_flux_CRP(a1, T.u1[nb], S, V1, m.vol[nb], vn, T.dt, Flux);
        fluxes += Flux;
    }
    return T.u1[c] - fluxes / V1;
}

// The flux sum of one cell for the built methods (fluxes_CRP,
// transfer.cpp:208-251; fluxes_MUSCL, :331-395), u2 = u1 - sum / V.
double tr_cell(const orc_tr& T, int c, int dir) {
    const orc_mesh& m = *T.m;
    const double a1 = T.u1[c];
    const double V1 = m.vol[c];
    const double ccx = m.cx[c], ccy = m.cy[c];
    double fluxes = 0.0;
    for (int s8 = 0; s8 < 8; ++s8) {
        const size_t f = 8 * (size_t)c + s8;
        const int b = m.bnd[f];
        if (b == ORC_UNDEFINED || slot_is_boundary(b) || mm_skip(m, f, dir)) continue;
        const int nb = m.adj[f];
        const double a2 = T.u1[nb];
        const double fnx = m.fnx[f], fny = m.fny[f];
        const double vn = (T.fvx[f] * fnx + T.fvy[f] * fny) + 0.0;
        const double S = m.area[f];
        const double V2 = m.vol[nb];
        if (T.method <= kTrN1) {
            double a_sig;
            switch (T.method) {
                case kTrV3: a_sig = face_fraction_v3(a1, a2); break;
                case kTrV5: a_sig = face_fraction_v5(a1, a2); break;
                case kTrN1:
                    a_sig = tr_face_fraction_n1(a1, a2, T.nx[c], T.ny[c], T.nx[nb], T.ny[nb], fnx, fny, vn, V1, V2, S,
                                                T.dt);
                    break;
                default: a_sig = face_fraction_s(a1, a2); break;
            }
            fluxes += tr_flux_2D(a1, a2, S, V1, V2, a_sig, vn, T.dt);
            continue;
        }
        const double fcx = m.fcx[f], fcy = m.fcy[f];
        const double ncx = m.cx[nb], ncy = m.cy[nb];
        double a_sig;
        if (T.method == kTrMC || T.method == kTrMCCRP) {   // FaceExtra::Direct (face_extra.h:135-177)
            const double drcx = fcx - ccx, drcy = fcy - ccy;
            const double drnx = fcx - ncx, drny = fcy - ncy;
            const double du_c = (T.dudx[c] * drcx + T.dudy[c] * drcy) + 0.0 * 0.0;
            const double du_n = (T.dudx[nb] * drnx + T.dudy[nb] * drny) + 0.0 * 0.0;
            a_sig = vn > 0.0 ? a1 + du_c : a2 + du_n;
            a_sig = f_between(a_sig, a1, a2);
        } else {   // FaceExtra::ATvL (face_extra.h:264-339)
            const double rcx = fcx - ccx, rcy = fcy - ccy;
            const double rnx = ncx - fcx, rny = ncy - fcy;
            const double dx_c = std::sqrt((rcx * rcx + rcy * rcy) + 0.0 * 0.0);
            const double dx_n = std::sqrt((rnx * rnx + rny * rny) + 0.0 * 0.0);
            const double du_c = ((T.dudx[c] * rcx + T.dudy[c] * rcy) + 0.0 * 0.0) / dx_c;
            const double du_n = ((T.dudx[nb] * rnx + T.dudy[nb] * rny) + 0.0 * 0.0) / dx_n;
            const double dx = dx_c + dx_n;
            if (vn > 0.0) {
                const double delta = dx_c / dx;
                const double eta = (12.0 * delta * delta - 1.0) / 24.0;
                const double dzp = a2 - a1;
                const double dzm = 2.0 * du_c * dx - dzp;
                const double s = tr_van_albada2(dzm, dzp);
                a_sig = a1 + s * ((0.5 * delta - s * eta) * dzm + (0.5 * delta + s * eta) * dzp);
            } else {
                const double delta = dx_n / dx;
                const double eta = (12.0 * delta * delta - 1.0) / 24.0;
                const double dzm = a2 - a1;
                const double dzp = 2.0 * du_n * dx - dzm;
                const double s = tr_van_albada2(dzm, dzp);
                a_sig = a2 - s * ((0.5 * delta + s * eta) * dzm + (0.5 * delta - s * eta) * dzp);
            }
        }
        double Flux = a_sig * vn * T.dt * S;
        if (T.method == kTrMCCRP || T.method == kTrMUSCLdCRP) Flux = tr_flux_CRP(a1, a2, S, V1, V2, vn, T.dt, Flux);
        fluxes += Flux;
    }
    return T.u1[c] - fluxes / V1;
}

// InterfaceRecovery::compute_normal (interface_recovery.cpp:17-39): zero in
// a pure cell, else -sum face_fraction(uc, un) S n_f over the inner faces,
// normalised.
void tr_normal(orc_tr& T, int c) {
    const orc_mesh& m = *T.m;
    const double uc = T.u1[c];
    T.nx[c] = 0.0;
    T.ny[c] = 0.0;
    if (uc < 1.0e-12 || uc > 1.0 - 1.0e-12) return;
    double nx = 0.0, ny = 0.0;
    for (int s8 = 0; s8 < 8; ++s8) {
        const size_t f = 8 * (size_t)c + s8;
        const int b = m.bnd[f];
        if (b == ORC_UNDEFINED || slot_is_boundary(b)) continue;
        const double un = T.u1[m.adj[f]];
        const double Sx = m.fnx[f] * m.area[f], Sy = m.fny[f] * m.area[f];
        const double ff = face_fraction(uc, un);
        nx = nx - ff * Sx;
        ny = ny - ff * Sy;
    }
    const double zz = (nx * nx + ny * ny) + 0.0 * 0.0;
    if (zz > 0.0) {
        const double r = std::sqrt(zz);
        nx = nx / r;
        ny = ny / r;
    }
    T.nx[c] = nx;
    T.ny[c] = ny;
}

}  // namespace

extern "C" {

orc_tr* orc_tr_create(orc_mesh* m, int method, double cfl) {
    if (!tr_built(method)) return nullptr;
    if ((method == kTrWENO || method == kTrWENOCRP) && (m->gnx < 3 || m->gny < 3 || m->gnx * m->gny != m->n))
        return nullptr;   // the five-cell stencils need a Rectangle of at least 3 x 3
    orc_tr* T = new orc_tr();
    T->m = m;
    T->method = method;
    T->cfl = smax(0.0, smin(cfl, 1.0));   // set_CFL
    const size_t n = m->n;
    T->u1.assign(n, 0.0); T->u2.assign(n, 0.0); T->nx.assign(n, 0.0); T->ny.assign(n, 0.0);
    T->dudx.assign(n, 0.0); T->dudy.assign(n, 0.0);
    T->fvx.assign(8 * n, 0.0); T->fvy.assign(8 * n, 0.0); T->cvx.assign(n, 0.0); T->cvy.assign(n, 0.0);
    return T;
}

void orc_tr_free(orc_tr* T) { delete T; }

void orc_tr_set_velocity(orc_tr* T, const double* fvx, const double* fvy, const double* cvx, const double* cvy) {
    const size_t n = T->m->n;
    T->fvx.assign(fvx, fvx + 8 * n); T->fvy.assign(fvy, fvy + 8 * n);
    T->cvx.assign(cvx, cvx + n); T->cvy.assign(cvy, cvy + n);
}

// The state u1, then update_interface (the normals) as Transfer's problem
// files do after setting it.
void orc_tr_set_state(orc_tr* T, const double* u) {
    T->u1.assign(u, u + T->m->n);
    for (int c = 0; c < T->m->n; ++c) tr_normal(*T, c);
}

void orc_tr_state(const orc_tr* T, double* u, double* nx, double* ny) {
    const size_t n = T->m->n;
    if (u) std::memcpy(u, T->u1.data(), n * 8);
    if (nx) std::memcpy(nx, T->nx.data(), n * 8);
    if (ny) std::memcpy(ny, T->ny.data(), n * 8);
}

void orc_tr_set_dt(orc_tr* T, double dt) { T->dt = dt; }

// compute_dt (transfer.cpp:86-101): CFL * min over cells of
// (V / max face area) / |v(centre)|.
double orc_tr_compute_dt(const orc_tr* T) {
    const orc_mesh& m = *T->m;
    double tau = std::numeric_limits<double>::max();
    for (int c = 0; c < m.n; ++c) {
        double max_area = 0.0;
        for (int s8 = 0; s8 < 8; ++s8) {
            const size_t f = 8 * (size_t)c + s8;
            if (m.bnd[f] == ORC_UNDEFINED) continue;
            max_area = smax(max_area, m.area[f]);
        }
        const double dx = m.vol[c] / max_area;
        const double vx = T->cvx[c], vy = T->cvy[c];
        tau = smin(tau, dx / std::sqrt((vx * vx + vy * vy) + 0.0 * 0.0));
    }
    return T->cfl * tau;
}

// update (transfer.cpp:474-535) for the built methods: slopes for the
// MUSCL methods, the fluxes, u1 = between(u2, 0, 1), the normals.
void orc_tr_update(orc_tr* T, int dir) {
    const int n = T->m->n;
    const bool weno = T->method == kTrWENO || T->method == kTrWENOCRP;
    if (T->method >= kTrMUSCLd && !weno)
        for (int c = 0; c < n; ++c) tr_grad(*T, c, T->method == kTrMC || T->method == kTrMCCRP, T->dudx[c], T->dudy[c]);
    for (int c = 0; c < n; ++c) T->u2[c] = weno ? tr_cell_weno(*T, c, dir) : tr_cell(*T, c, dir);
    for (int c = 0; c < n; ++c) {
        T->u1[c] = f_between(T->u2[c], 0.0, 1.0);
        T->u2[c] = 0.0;
    }
    for (int c = 0; c < n; ++c) tr_normal(*T, c);
}

void orc_tr_slopes(const orc_tr* T, double* dx, double* dy) {
    std::memcpy(dx, T->dudx.data(), (size_t)T->m->n * 8);
    std::memcpy(dy, T->dudy.data(), (size_t)T->m->n * 8);
}

// set_flags (transfer.cpp:668-690): 1 where u1 varies over the cell and its
// inner neighbours, -1 elsewhere.
void orc_tr_set_flags(orc_tr* T) {
    orc_mesh& m = *T->m;
    for (int c = 0; c < m.n; ++c) {
        double lo = T->u1[c], hi = T->u1[c];
        for (int s8 = 0; s8 < 8; ++s8) {
            const size_t f = 8 * (size_t)c + s8;
            const int b = m.bnd[f];
            if (b == ORC_UNDEFINED || slot_is_boundary(b)) continue;
            lo = smin(lo, T->u1[m.adj[f]]);
            hi = smax(hi, T->u1[m.adj[f]]);
        }
        m.flag[c] = hi - lo > 0.0 ? 1 : -1;
    }
}

}  // extern "C"

extern "C" {
// WENO5::p() (right != 0) or m() of five cell averages (the WENO test hook).
double orc_weno5(const double* u, int right) { return weno5(u, right != 0); }
}
