// ORACLE — test infrastructure only.  CPU restatement of zephyr's SmFluid
// step on three-dimensional Cuboid meshes (AMR-3D cells of 24 face slots,
// of which a Cuboid defines the six main ones): the mesh of
// Cuboid::initialize (geom/generator/cuboid.cpp:300-442) with
// AmrCells::incircle_diameter (mesh/euler/amr_cells.cpp:107-121) and the
// 3D weight of gradient::LSM (w0 = 1 / linear_size = 1 / cbrt(volume),
// gradient.h:231-232, amr_cells.h:283-285), the 3D branch of
// Rotate::to_local / to_global (math/cfd/rotate.cpp:7-53), and the step of
// SmFluid::update (sm_fluid.cpp:81-366) written with the same per-face
// terms as the two-dimensional restatement in oracle.cpp.  Included at the
// end of oracle.cpp.

struct orc_mesh3 {
    int n = 0, nx = 0, ny = 0, nz = 0;
    std::vector<double> cx, cy, cz, vol, h;
    // face slots, 24 per cell (AMR-3D, geom/indexing.h:340-372): slot
    // s = side + 6 * k (sides L, R, B, T, Z, F; k the sub-face)
    std::vector<int> bnd, adj;
    std::vector<double> fnx, fny, fnz, fcx, fcy, fcz, area, w;
};

struct orc_smf3 {
    orc_mesh3* m;
    Eos eos;
    int flux, limiter, acc;
    double cfl, dt, max_dt;
    std::vector<PState> init, half, next, d_dx, d_dy, d_dz;
    std::vector<int> wait, wait_next;
};

namespace {

constexpr int kSlots3 = 24;

// Rotate::to_local (rotate.cpp:7-36), both branches: the rotation matrix of
// the normalised normal, then R * v (row i: (R(i,0) v0 + R(i,1) v1) + R(i,2) v2).
inline Vec3 rot3(const Vec3& v, const Vec3& no, bool to_local) {
    const Vec3 normal = normalized(no);
    double R[3][3];
    if (no.z == 0.0) {
        const double c = normal.x, s = normal.y;
        if (to_local) {
            R[0][0] = c;   R[0][1] = s;   R[0][2] = 0.0;
            R[1][0] = -s;  R[1][1] = c;   R[1][2] = 0.0;
        } else {
            R[0][0] = c;   R[0][1] = -s;  R[0][2] = 0.0;
            R[1][0] = s;   R[1][1] = c;   R[1][2] = 0.0;
        }
        R[2][0] = 0.0; R[2][1] = 0.0; R[2][2] = 1.0;
    } else {
        // rotation axis: (0, n_z, -n_y) to the local frame, (0, -n_z, n_y) back
        const Vec3 n = to_local ? normalized(Vec3{0.0, normal.z, -normal.y})
                                : normalized(Vec3{0.0, -normal.z, normal.y});
        const double c = normal.x;
        const double s = std::sqrt(1.0 - c * c);
        R[0][0] = c;            R[0][1] = -s * n.z;                       R[0][2] = s * n.y;
        R[1][0] = s * n.z;      R[1][1] = c + (1 - c) * n.y * n.y;        R[1][2] = (1 - c) * n.y * n.z;
        R[2][0] = -s * n.y;     R[2][1] = (1 - c) * n.y * n.z;            R[2][2] = c + (1 - c) * n.z * n.z;
    }
    return {(R[0][0] * v.x + R[0][1] * v.y) + R[0][2] * v.z,
            (R[1][0] * v.x + R[1][1] * v.y) + R[1][2] * v.z,
            (R[2][0] * v.x + R[2][1] * v.y) + R[2][2] * v.z};
}

inline Vec3 slot_normal(const orc_mesh3& m, size_t s) { return Vec3{m.fnx[s], m.fny[s], m.fnz[s]}; }
inline Vec3 slot_centre(const orc_mesh3& m, size_t s) { return Vec3{m.fcx[s], m.fcy[s], m.fcz[s]}; }
inline Vec3 cell_centre(const orc_mesh3& m, int c) { return Vec3{m.cx[c], m.cy[c], m.cz[c]}; }

// EuFace::symm_point (amr_faces.h:197-199).
inline Vec3 symm_point3(const orc_mesh3& m, size_t s, const Vec3& p) {
    const Vec3 n = slot_normal(m, s);
    double d = 2.0 * dot(slot_centre(m, s) - p, n);
    return p + scale(d, n);
}

inline Vec3 add4(const Vec3& a, const Vec3& b, const Vec3& c, const Vec3& d) { return ((a + b) + c) + d; }

// gradient::LSM and gradient::limiting (gradient.h:226-282, 412-474) on a
// 3D cell: the full 3 x 3 system (no 2D fix-up applies), Eigen's cofactor
// inverse, B = adj * (1/det).
void grad_cell3(orc_smf3& S, int c, Grad& out) {
    const orc_mesh3& m = *S.m;
    const PState zc = S.init[c];
    const Vec3 cell_c = cell_centre(m, c);
    double A[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
    double Fx[6] = {0}, Fy[6] = {0}, Fz[6] = {0};
    struct FaceGeo { Vec3 dr, drn; double w; PState zn; };
    FaceGeo fg[kSlots3];
    int nf = 0;
    for (int k24 = 0; k24 < kSlots3; ++k24) {
        const size_t s = (size_t)kSlots3 * c + k24;
        const int b = m.bnd[s];
        if (b == ORC_UNDEFINED) continue;
        const Vec3 normal = slot_normal(m, s);
        const PState zn = !slot_is_boundary(b) ? S.init[m.adj[s]] : boundary_value(zc, normal, b);
        const Vec3 nc = (b == ORC_INNER) ? cell_centre(m, m.adj[s]) : symm_point3(m, s, cell_c);
        const Vec3 dr = nc - cell_c;
        const Vec3 drn = scale(dot(dr, normal), normal);
        const double w = m.w[s];
        const double wd[3] = {w * drn.x, w * drn.y, w * drn.z};
        const double drv[3] = {dr.x, dr.y, dr.z};
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) A[i][j] = A[i][j] + wd[i] * drv[j];
        for (int k = 0; k < 6; ++k) {
            const double dF = zn.arr()[k] - zc.arr()[k];
            Fx[k] = Fx[k] + (w * dF) * drn.x;
            Fy[k] = Fy[k] + (w * dF) * drn.y;
            Fz[k] = Fz[k] + (w * dF) * drn.z;
        }
        fg[nf++] = {dr, drn, w, zn};
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
    Grad g;
    for (int k = 0; k < 6; ++k) {
        g.x.arr()[k] = (B[0][0] * Fx[k] + B[0][1] * Fy[k]) + B[0][2] * Fz[k];
        g.y.arr()[k] = (B[1][0] * Fx[k] + B[1][1] * Fy[k]) + B[1][2] * Fz[k];
        g.z.arr()[k] = (B[2][0] * Fx[k] + B[2][1] * Fy[k]) + B[2][2] * Fz[k];
    }
    double Lx[6] = {0}, Ly[6] = {0}, Lz[6] = {0};
    for (int f = 0; f < nf; ++f) {
        const FaceGeo& q = fg[f];
        for (int k = 0; k < 6; ++k) {
            const double dF_p = q.zn.arr()[k] - zc.arr()[k];
            const double du_n = (g.x.arr()[k] * q.dr.x + g.y.arr()[k] * q.dr.y) + g.z.arr()[k] * q.dr.z;
            const double dF_m = 2.0 * du_n - dF_p;
            const double dF_lim = lim_binary(S.limiter, dF_m, dF_p) * dF_p;
            Lx[k] = Lx[k] + (q.w * dF_lim) * q.drn.x;
            Ly[k] = Ly[k] + (q.w * dF_lim) * q.drn.y;
            Lz[k] = Lz[k] + (q.w * dF_lim) * q.drn.z;
        }
    }
    for (int k = 0; k < 6; ++k) {
        out.x.arr()[k] = (B[0][0] * Lx[k] + B[0][1] * Ly[k]) + B[0][2] * Lz[k];
        out.y.arr()[k] = (B[1][0] * Lx[k] + B[1][1] * Ly[k]) + B[1][2] * Lz[k];
        out.z.arr()[k] = (B[2][0] * Lx[k] + B[2][1] * Ly[k]) + B[2][2] * Lz[k];
    }
}

inline PState in_local3(const PState& z, const Vec3& n) {
    PState r = z;
    r.v = rot3(z.v, n, true);
    return r;
}

// SmFluid::fluxes, first order (sm_fluid.cpp:127-174).
void fluxes_cell3(orc_smf3& S, int c) {
    const orc_mesh3& m = *S.m;
    const PState z_c = S.init[c];
    QState q_c = to_qstate(z_c);
    Flux flux{0, {0, 0, 0}, 0};
    for (int k24 = 0; k24 < kSlots3; ++k24) {
        const size_t s = (size_t)kSlots3 * c + k24;
        const int b = m.bnd[s];
        if (b == ORC_UNDEFINED) continue;
        const Vec3 normal = slot_normal(m, s);
        const PState z_n = !slot_is_boundary(b) ? S.init[m.adj[s]] : boundary_value(z_c, normal, b);
        Flux lf = num_flux(S.flux, in_local3(z_c, normal), in_local3(z_n, normal), S.eos);
        lf.m = rot3(lf.m, normal, false);
        for (int k = 0; k < 5; ++k) flux.arr()[k] = flux.arr()[k] + lf.arr()[k] * m.area[s];
    }
    const double coef = S.dt / m.vol[c];
    for (int k = 0; k < 5; ++k) q_c.arr()[k] = q_c.arr()[k] - coef * flux.arr()[k];
    S.next[c] = to_pstate(q_c, S.eos);
}

// SmFluid::fluxes_stage1, predictor (sm_fluid.cpp:176-253).
void stage1_cell3(orc_smf3& S, int c) {
    const orc_mesh3& m = *S.m;
    if (S.wait[c] > 0) {
        S.half[c] = S.init[c];
        return;
    }
    const Vec3 cell_c = cell_centre(m, c);
    const PState z_c = S.init[c];
    QState q_c = to_qstate(z_c);
    Flux flux{0, {0, 0, 0}, 0};
    for (int k24 = 0; k24 < kSlots3; ++k24) {
        const size_t s = (size_t)kSlots3 * c + k24;
        if (m.bnd[s] == ORC_UNDEFINED) continue;
        const Vec3 normal = slot_normal(m, s);
        PState zm = extrap(z_c, S.d_dx[c], S.d_dy[c], S.d_dz[c], slot_centre(m, s) - cell_c);
        zm.e = S.eos.energy_rP(zm.rho, zm.p);
        if (is_bad(zm, S.eos)) zm = z_c;
        zm.v = rot3(zm.v, normal, true);
        Flux lf = phys_flux(zm);
        lf.m = rot3(lf.m, normal, false);
        for (int k = 0; k < 5; ++k) flux.arr()[k] = flux.arr()[k] + lf.arr()[k] * m.area[s];
    }
    const double coef = 0.5 * S.dt / m.vol[c];
    for (int k = 0; k < 5; ++k) q_c.arr()[k] = q_c.arr()[k] - coef * flux.arr()[k];
    S.half[c] = to_pstate(q_c, S.eos);
    if (is_bad(S.half[c], S.eos)) S.half[c] = z_c;
}

// SmFluid::fluxes_stage2, corrector (sm_fluid.cpp:255-366), wait read from
// the pre-pass array.
void stage2_cell3(orc_smf3& S, int c) {
    const orc_mesh3& m = *S.m;
    const Vec3 cell_c = cell_centre(m, c);
    const PState z_c = S.init[c];
    const PState z_ch = S.half[c];
    QState q_c = to_qstate(z_c);
    Flux flux{0, {0, 0, 0}, 0};
    for (int k24 = 0; k24 < kSlots3; ++k24) {
        const size_t s = (size_t)kSlots3 * c + k24;
        const int b = m.bnd[s];
        if (b == ORC_UNDEFINED) continue;
        const Vec3 normal = slot_normal(m, s);
        const Vec3 face_c = slot_centre(m, s);
        const int nb = m.adj[s];
        const bool bound = slot_is_boundary(b);
        const PState z_nh = bound ? boundary_value(z_ch, normal, b) : S.half[nb];
        PState zm, zp;
        const bool robust = S.wait[c] > 0 || S.wait[nb] > 0;
        if (robust) {
            zm = z_ch;
            zp = z_nh;
        } else {
            const Vec3 neib_c = (b == ORC_INNER) ? cell_centre(m, nb) : symm_point3(m, s, cell_c);
            zm = extrap(z_ch, S.d_dx[c], S.d_dy[c], S.d_dz[c], face_c - cell_c);
            zm.e = S.eos.energy_rP(zm.rho, zm.p);
            if (is_bad(zm, S.eos)) zm = z_ch;
            if (!bound) {
                zp = extrap(z_nh, S.d_dx[nb], S.d_dy[nb], S.d_dz[nb], face_c - neib_c);
                zp.e = S.eos.energy_rP(zp.rho, zp.p);
                if (is_bad(zp, S.eos)) zp = z_nh;
            } else {
                zp = boundary_value(zm, normal, b);
            }
        }
        zm.v = rot3(zm.v, normal, true);
        zp.v = rot3(zp.v, normal, true);
        Flux lf = robust ? hll(zm, zp, S.eos) : num_flux(S.flux, zm, zp, S.eos);
        lf.m = rot3(lf.m, normal, false);
        for (int k = 0; k < 5; ++k) flux.arr()[k] = flux.arr()[k] + lf.arr()[k] * m.area[s];
    }
    const double coef = S.dt / m.vol[c];
    for (int k = 0; k < 5; ++k) q_c.arr()[k] = q_c.arr()[k] - coef * flux.arr()[k];
    const PState zn = to_pstate(q_c, S.eos);
    if (is_bad(zn, S.eos)) {
        S.next[c] = z_c;
        S.wait_next[c] = 3;
    } else {
        S.next[c] = zn;
        S.wait_next[c] = S.wait[c] > 0 ? S.wait[c] - 1 : S.wait[c];
    }
}

double compute_dt3(orc_smf3& S, int nthreads) {
    const orc_mesh3& m = *S.m;
    double res = std::numeric_limits<double>::max();
    #pragma omp parallel for num_threads(nthreads) schedule(static) reduction(min:res)
    for (int c = 0; c < m.n; ++c) {
        const PState& z = S.init[c];
        const double cs = S.eos.sound_speed_rP(z.rho, z.p);
        const double t = m.h[c] / (norm(z.v) + cs);
        if (t < res) res = t;
    }
    return smin(S.cfl * res, S.max_dt);
}

}  // namespace

extern "C" {

// Cuboid::initialize (cuboid.cpp:300-442): cell c = nz (ny i + j) + k,
// vertices x = xmin + ((xmax - xmin) i) / nx (and y, z alike), the SqCube
// of the eight corners (cube.cpp:231-252) for the centre and the face
// centres, volume hx hy hz, areas hy hz / hx hz / hx hy, unit normals; a
// side of a periodic direction is periodic on both ends.  bounds: L, R, B,
// T, Z (back), F (front).
orc_mesh3* orc_mesh3_cuboid(int nx, int ny, int nz, double xmin, double xmax, double ymin, double ymax,
                            double zmin, double zmax, const int* bounds_in) {
    int bd[6];
    for (int k = 0; k < 6; ++k) bd[k] = bounds_in[k];
    const bool per[3] = {bd[0] == ORC_PERIODIC || bd[1] == ORC_PERIODIC, bd[2] == ORC_PERIODIC || bd[3] == ORC_PERIODIC,
                         bd[4] == ORC_PERIODIC || bd[5] == ORC_PERIODIC};
    for (int d = 0; d < 3; ++d)
        if (per[d]) bd[2 * d] = bd[2 * d + 1] = ORC_PERIODIC;
    orc_mesh3* m = new orc_mesh3();
    const int n = nx * ny * nz;
    m->n = n; m->nx = nx; m->ny = ny; m->nz = nz;
    m->cx.resize(n); m->cy.resize(n); m->cz.resize(n); m->vol.resize(n); m->h.resize(n);
    const size_t ns = (size_t)kSlots3 * n;
    m->bnd.assign(ns, ORC_UNDEFINED); m->adj.assign(ns, -1);
    for (auto* v : {&m->fnx, &m->fny, &m->fnz, &m->fcx, &m->fcy, &m->fcz, &m->area, &m->w}) v->assign(ns, 0.0);
    const double hx = (xmax - xmin) / nx, hy = (ymax - ymin) / ny, hz = (zmax - zmin) / nz;
    auto index = [&](int i, int j, int k) { return nz * (ny * i + j) + k; };
    auto vert = [&](int i, int j, int k) {
        return Vec3{xmin + ((xmax - xmin) * i) / nx, ymin + ((ymax - ymin) * j) / ny, zmin + ((zmax - zmin) * k) / nz};
    };
    for (int c = 0; c < n; ++c) {
        const int i = (c / nz) / ny, j = (c / nz) % ny, k = c % nz;
        const Vec3 v000 = vert(i, j, k), v002 = vert(i + 1, j, k), v020 = vert(i, j + 1, k), v022 = vert(i + 1, j + 1, k);
        const Vec3 v200 = vert(i, j, k + 1), v202 = vert(i + 1, j, k + 1), v220 = vert(i, j + 1, k + 1),
                   v222 = vert(i + 1, j + 1, k + 1);
        const Vec3 cc = scale(0.125, ((((((v000 + v002) + v020) + v022) + v200) + v202) + v220) + v222);
        const Vec3 fc[6] = {scale(0.25, add4(v000, v200, v020, v220)), scale(0.25, add4(v002, v202, v022, v222)),
                            scale(0.25, add4(v000, v200, v002, v202)), scale(0.25, add4(v020, v220, v022, v222)),
                            scale(0.25, add4(v000, v002, v020, v022)), scale(0.25, add4(v200, v202, v220, v222))};
        const Vec3 nrm[6] = {{-1.0, 0.0, 0.0}, {1.0, 0.0, 0.0}, {0.0, -1.0, 0.0},
                             {0.0, 1.0, 0.0},  {0.0, 0.0, -1.0}, {0.0, 0.0, 1.0}};
        const double ar[6] = {hy * hz, hy * hz, hx * hz, hx * hz, hx * hy, hx * hy};
        const int nb[6] = {i == 0 && !per[0] ? c : index((i - 1 + nx) % nx, j, k),
                           i == nx - 1 && !per[0] ? c : index((i + 1) % nx, j, k),
                           j == 0 && !per[1] ? c : index(i, (j - 1 + ny) % ny, k),
                           j == ny - 1 && !per[1] ? c : index(i, (j + 1) % ny, k),
                           k == 0 && !per[2] ? c : index(i, j, (k - 1 + nz) % nz),
                           k == nz - 1 && !per[2] ? c : index(i, j, (k + 1) % nz)};
        const int bt[6] = {i > 0 ? ORC_INNER : bd[0], i < nx - 1 ? ORC_INNER : bd[1],
                           j > 0 ? ORC_INNER : bd[2], j < ny - 1 ? ORC_INNER : bd[3],
                           k > 0 ? ORC_INNER : bd[4], k < nz - 1 ? ORC_INNER : bd[5]};
        m->cx[c] = cc.x; m->cy[c] = cc.y; m->cz[c] = cc.z;
        m->vol[c] = hx * hy * hz;
        // incircle_diameter: sqrt of the least squared distance between
        // opposite face centres (amr_cells.cpp:115-120, math::min = std::min
        // (x, std::min(y, z)))
        m->h[c] = std::sqrt(smin(sqnorm(fc[1] - fc[0]), smin(sqnorm(fc[3] - fc[2]), sqnorm(fc[5] - fc[4]))));
        for (int s6 = 0; s6 < 6; ++s6) {
            const size_t s = (size_t)kSlots3 * c + s6;
            m->bnd[s] = bt[s6];
            m->adj[s] = nb[s6];
            m->fnx[s] = nrm[s6].x; m->fny[s] = nrm[s6].y; m->fnz[s] = nrm[s6].z;
            m->fcx[s] = fc[s6].x; m->fcy[s] = fc[s6].y; m->fcz[s] = fc[s6].z;
            m->area[s] = ar[s6];
        }
    }
    // LSM weights (gradient.h:231-256): w = w0 area / |drn|^3, w0 =
    // 1 / cbrt(volume); dr to the neighbour's centre through an inner
    // face, to the mirror point otherwise
    for (int c = 0; c < n; ++c) {
        const double w0 = 1.0 / std::cbrt(m->vol[c]);
        const Vec3 cell_c = cell_centre(*m, c);
        for (int s6 = 0; s6 < 6; ++s6) {
            const size_t s = (size_t)kSlots3 * c + s6;
            const Vec3 normal = slot_normal(*m, s);
            const Vec3 nc = m->bnd[s] == ORC_INNER ? cell_centre(*m, m->adj[s]) : symm_point3(*m, s, cell_c);
            const Vec3 dr = nc - cell_c;
            const Vec3 drn = scale(dot(dr, normal), normal);
            m->w[s] = w0 * m->area[s] / std::pow(norm(drn), 3);
        }
    }
    return m;
}

void orc_mesh3_free(orc_mesh3* m) { delete m; }
int orc_mesh3_size(const orc_mesh3* m) { return m->n; }

void orc_mesh3_get_cells(const orc_mesh3* m, double* cx, double* cy, double* cz, double* vol, double* h) {
    const size_t b = sizeof(double) * m->n;
    std::memcpy(cx, m->cx.data(), b); std::memcpy(cy, m->cy.data(), b); std::memcpy(cz, m->cz.data(), b);
    std::memcpy(vol, m->vol.data(), b); std::memcpy(h, m->h.data(), b);
}

// per slot, 24 per cell, cell-major
void orc_mesh3_get_slots(const orc_mesh3* m, int* boundary, int* adj, double* nx, double* ny, double* nz,
                         double* fcx, double* fcy, double* fcz, double* area, double* w) {
    const size_t ns = (size_t)kSlots3 * m->n;
    std::memcpy(boundary, m->bnd.data(), sizeof(int) * ns);
    std::memcpy(adj, m->adj.data(), sizeof(int) * ns);
    const std::vector<double>* src[8] = {&m->fnx, &m->fny, &m->fnz, &m->fcx, &m->fcy, &m->fcz, &m->area, &m->w};
    double* dst[8] = {nx, ny, nz, fcx, fcy, fcz, area, w};
    for (int k = 0; k < 8; ++k) std::memcpy(dst[k], src[k]->data(), sizeof(double) * ns);
}

orc_smf3* orc_smf3_create(orc_mesh3* m, double gamma, double P0, double e0, int flux, int limiter, int acc,
                          double cfl) {
    orc_smf3* S = new orc_smf3();
    S->m = m;
    S->eos = {gamma, P0, e0};
    S->flux = flux;
    S->limiter = limiter;
    S->acc = std::min(std::max(1, acc), 2);
    S->cfl = std::max(0.0, std::min(cfl, 1.0));
    S->dt = NAN;
    S->max_dt = std::numeric_limits<double>::max();
    PState zz;
    zero(zz);
    const int n = m->n;
    S->init.assign(n, zz); S->half.assign(n, zz); S->next.assign(n, zz);
    S->d_dx.assign(n, zz); S->d_dy.assign(n, zz); S->d_dz.assign(n, zz);
    S->wait.assign(n, 0); S->wait_next.assign(n, 0);
    return S;
}

void orc_smf3_free(orc_smf3* s) { delete s; }

void orc_smf3_set_state(orc_smf3* s, const double* rho, const double* u, const double* v, const double* w,
                        const double* p, const double* e) {
    for (int c = 0; c < s->m->n; ++c) s->init[c] = PState{rho[c], {u[c], v[c], w[c]}, p[c], e[c]};
}

void orc_smf3_get_state(const orc_smf3* s, double* rho, double* u, double* v, double* w, double* p, double* e) {
    for (int c = 0; c < s->m->n; ++c) {
        const PState& z = s->init[c];
        rho[c] = z.rho; u[c] = z.v.x; v[c] = z.v.y; w[c] = z.v.z; p[c] = z.p; e[c] = z.e;
    }
}

void orc_smf3_get_wait(const orc_smf3* s, int* wait) { std::memcpy(wait, s->wait.data(), sizeof(int) * s->m->n); }
void orc_smf3_set_wait(orc_smf3* s, const int* wait) { std::memcpy(s->wait.data(), wait, sizeof(int) * s->m->n); }

// d/dx, d/dy, d/dz of rho, u, v, w, p: [3][5][n]
void orc_smf3_get_grad(const orc_smf3* s, double* g) {
    const int n = s->m->n;
    const std::vector<PState>* G[3] = {&s->d_dx, &s->d_dy, &s->d_dz};
    for (int d = 0; d < 3; ++d)
        for (int c = 0; c < n; ++c) {
            const PState& q = (*G[d])[c];
            const double v[5] = {q.rho, q.v.x, q.v.y, q.v.z, q.p};
            for (int k = 0; k < 5; ++k) g[((size_t)d * 5 + k) * n + c] = v[k];
        }
}

double orc_smf3_compute_dt(orc_smf3* s, int nthreads) { return compute_dt3(*s, nthreads); }

// SmFluid::update (sm_fluid.cpp:81-102).
int orc_smf3_update(orc_smf3* s, int nthreads) {
    const int n = s->m->n;
    int err = 0;
    s->dt = compute_dt3(*s, nthreads);
    if (s->acc == 1) {
        #pragma omp parallel for num_threads(nthreads) schedule(static) reduction(|:err)
        for (int c = 0; c < n; ++c) {
            g_numeric_error = 0;
            fluxes_cell3(*s, c);
            err |= g_numeric_error;
        }
    } else {
        #pragma omp parallel for num_threads(nthreads) schedule(static)
        for (int c = 0; c < n; ++c) {
            Grad g;
            grad_cell3(*s, c, g);
            s->d_dx[c] = g.x; s->d_dy[c] = g.y; s->d_dz[c] = g.z;
        }
        #pragma omp parallel for num_threads(nthreads) schedule(static)
        for (int c = 0; c < n; ++c) stage1_cell3(*s, c);
        #pragma omp parallel for num_threads(nthreads) schedule(static) reduction(|:err)
        for (int c = 0; c < n; ++c) {
            g_numeric_error = 0;
            stage2_cell3(*s, c);
            err |= g_numeric_error;
        }
        std::swap(s->wait, s->wait_next);
    }
    std::swap(s->init, s->next);
    return err;
}

double orc_smf3_dt(const orc_smf3* s) { return s->dt; }

// sum of V q over the cells: mass, momentum x, y, z, energy
void orc_smf3_totals(const orc_smf3* s, double* out) {
    double t[5] = {0, 0, 0, 0, 0};
    for (int c = 0; c < s->m->n; ++c) {
        const QState q = to_qstate(s->init[c]);
        const double V = s->m->vol[c];
        t[0] += V * q.rho; t[1] += V * q.m.x; t[2] += V * q.m.y; t[3] += V * q.m.z; t[4] += V * q.E;
    }
    for (int k = 0; k < 5; ++k) out[k] = t[k];
}

// Rotate::to_local / to_global of one vector (the rotation test hook).
void orc_rotate3(const double* v, const double* n, int to_local, double* out) {
    const Vec3 r = rot3(Vec3{v[0], v[1], v[2]}, Vec3{n[0], n[1], n[2]}, to_local != 0);
    out[0] = r.x; out[1] = r.y; out[2] = r.z;
}

}  // extern "C"
