// Host side of the three-dimensional SmFluid step: the Cuboid mesh built on
// the host (Cuboid::initialize, geom/generator/cuboid.cpp:300-442, with
// AmrCells::incircle_diameter, mesh/euler/amr_cells.cpp:107-121, and the
// least-squares weights of gradient::LSM, gradient.h:231-256), uploaded
// slot-major, and SmFluid::update (sm_fluid.cpp:81-102) as a sequence of
// kernel launches on the context's stream, behind the zfv3_* C ABI.
#include "../../include/zfv.h"

#include <hip/hip_runtime.h>

#include <climits>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "zfv_3d.hpp"

namespace zfv {
int set_error(int code, const char* msg);   // zfv_last_error's text (zfv_capi.cpp)
}

using namespace zfv;

namespace {

int fail(int code, const std::string& msg) { return zfv::set_error(code, msg.c_str()); }

#define Z3HIP(call)                                                                       \
    do {                                                                                  \
        hipError_t e_ = (call);                                                           \
        if (e_ != hipSuccess)                                                             \
            return fail(ZFV_ERR_HIP, std::string(#call) + ": " + hipGetErrorString(e_)); \
    } while (0)

struct V3 {
    double x, y, z;
};
inline V3 operator+(const V3& a, const V3& b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline V3 operator-(const V3& a, const V3& b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline V3 mul(double k, const V3& a) { return {k * a.x, k * a.y, k * a.z}; }
inline double dot3(const V3& a, const V3& b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }
inline double smin_h(double a, double b) { return (b < a) ? b : a; }

constexpr int kSides = 6;      // L, R, B, T, Z (back), F (front)
constexpr int kSlots24 = 24;   // AMR-3D face slots per cell (geom/indexing.h:340-372)

}  // namespace

struct zfv3_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    // mesh, host copy (cell-major, six main slots per cell)
    int n = 0, nx = 0, ny = 0, nz = 0;
    std::vector<double> cx, cy, cz, vol, h;
    std::vector<int> bnd, adj;                                   // [6n]
    std::vector<double> fcx, fcy, fcz, area, w;                  // [6n]
    // device buffers
    std::vector<void*> allocs;
    double* d_cell = nullptr;     // cx, cy, cz, vol, h: [5][n]
    int* d_type = nullptr;        // [6][n]
    int* d_adj = nullptr;         // [6][n]
    double* d_slot = nullptr;     // fcx, fcy, fcz, area, w: [5][6][n]
    double* d_z[3] = {nullptr, nullptr, nullptr};   // three states of [6][n]
    double* d_G = nullptr;        // [15][n]
    int* d_wait[2] = {nullptr, nullptr};
    unsigned long long* d_slots = nullptr;   // three rotating CFL slots
    double* d_dt = nullptr;
    int* d_err = nullptr;
    double* d_part = nullptr;     // totals partial sums [5][blocks]
    double* d_tab = nullptr;      // the Cuboid's geometry tables, when every cell matches them
    int tab_bcode[6] = {0, 0, 0, 0, 0, 0};
    int tab_per[3] = {0, 0, 0};
    bool use_tables = true;
    int blocks = 0;
    // solver
    bool configured = false, have_state = false, dt_valid = false;
    bool dt_ready = false;        // slot `cur` holds the CFL minimum of the current state
    zfv_smf_params prm{};
    int zi = 0, hi = 1, ni = 2;   // init / half / next buffers
    int wi = 0, cur = 0;
    // kernel timing: an event pair around every launch while enabled
    struct Timed {
        const char* name;
        hipEvent_t a, b;
    };
    bool prof_on = false;
    std::vector<Timed> timed;

    ~zfv3_ctx() {
        drop_events();
        release();
        if (stream) (void)hipStreamDestroy(stream);
    }
    void drop_events() {
        for (auto& t : timed) {
            (void)hipEventDestroy(t.a);
            (void)hipEventDestroy(t.b);
        }
        timed.clear();
    }
    void release() {
        for (void* p : allocs) (void)hipFree(p);
        allocs.clear();
        d_cell = nullptr; d_type = nullptr; d_adj = nullptr; d_slot = nullptr;
        d_z[0] = d_z[1] = d_z[2] = nullptr; d_G = nullptr; d_wait[0] = d_wait[1] = nullptr;
        d_slots = nullptr; d_dt = nullptr; d_err = nullptr; d_part = nullptr; d_tab = nullptr;
    }
    template <class T>
    int alloc(T** p, size_t count) {
        void* q = nullptr;
        Z3HIP(hipMalloc(&q, count * sizeof(T) > 0 ? count * sizeof(T) : 8));
        allocs.push_back(q);
        *p = static_cast<T*>(q);
        return ZFV_OK;
    }
    Dev3Mesh mesh() const {
        const size_t N = (size_t)n, S = (size_t)kSides * n;
        Dev3Mesh M{};
        M.n = n;
        M.stride = n;
        if (d_tab) {
            M.tnx = nx; M.tny = ny; M.tnz = nz;
            M.tab = d_tab;
            for (int k = 0; k < 6; ++k) M.bcode[k] = tab_bcode[k];
            for (int a = 0; a < 3; ++a) M.per[a] = tab_per[a];
        }
        M.cx = d_cell; M.cy = d_cell + N; M.cz = d_cell + 2 * N; M.vol = d_cell + 3 * N; M.h = d_cell + 4 * N;
        M.type = d_type; M.adj = d_adj;
        M.fcx = d_slot; M.fcy = d_slot + S; M.fcz = d_slot + 2 * S; M.area = d_slot + 3 * S; M.w = d_slot + 4 * S;
        return M;
    }
    Dev3State state(int k) const {
        double* b = d_z[k];
        const size_t N = (size_t)n;
        return Dev3State{b, b + N, b + 2 * N, b + 3 * N, b + 4 * N, b + 5 * N};
    }
    Eos eos() const { return Eos{prm.gamma, prm.P0, prm.e0}; }
    unsigned long long* slot(int k) const { return d_slots + (size_t)k * kSlotWords; }
};

namespace {

int need_mesh(zfv3_ctx* ctx) {
    if (!ctx) return fail(ZFV_ERR_ARG, "null context");
    if (ctx->n == 0) return fail(ZFV_ERR_STATE, "no mesh");
    (void)hipSetDevice(ctx->device);
    return ZFV_OK;
}

int need_ready(zfv3_ctx* ctx, bool state) {
    int r = need_mesh(ctx);
    if (r) return r;
    if (!ctx->configured) return fail(ZFV_ERR_STATE, "solver not configured");
    if (state && !ctx->have_state) return fail(ZFV_ERR_STATE, "no state uploaded");
    return ZFV_OK;
}

// Sticky error word of the kernels: bit 0 a NaN flux (hllc.cpp:72-80 and
// hll.cpp:46-54 throw there).
int read_err(zfv3_ctx* ctx) {
    int e = 0;
    Z3HIP(hipMemcpyAsync(&e, ctx->d_err, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
    Z3HIP(hipStreamSynchronize(ctx->stream));
    if (e & 1) return fail(ZFV_ERR_NUMERIC, "numeric error: a flux produced NaN");
    return ZFV_OK;
}

int slot_code(int b) {
    switch (b) {
        case ZFV_INNER: return SC_INNER;
        case ZFV_PERIODIC: return SC_PERIODIC;
        case ZFV_WALL: return SC_WALL;
        default: return SC_COPY;   // ZOE, FUNCTION: the cell's own state
    }
}

// Cuboid::initialize on the host, the operations in the reference's order:
// vertices xmin + ((xmax - xmin) i) / nx, SqCube centre (1/8 of the eight
// corners summed left to right) and face centres (1/4 of four corners),
// volume hx hy hz, areas, incircle diameter, then the LSM weights
// w = (1 / cbrt(V)) area / |drn|^3 with dr to the neighbour's centre through
// an inner face and to the mirror point of the cell centre otherwise.
void build_cuboid(zfv3_ctx* m, int nx, int ny, int nz, double xmin, double xmax, double ymin, double ymax,
                  double zmin, double zmax, const int* bd) {
    const int n = nx * ny * nz;
    m->n = n; m->nx = nx; m->ny = ny; m->nz = nz;
    for (auto* v : {&m->cx, &m->cy, &m->cz, &m->vol, &m->h}) v->assign(n, 0.0);
    const size_t ns = (size_t)kSides * n;
    m->bnd.assign(ns, ZFV_INNER);
    m->adj.assign(ns, 0);
    for (auto* v : {&m->fcx, &m->fcy, &m->fcz, &m->area, &m->w}) v->assign(ns, 0.0);
    const bool per[3] = {bd[0] == ZFV_PERIODIC, bd[2] == ZFV_PERIODIC, bd[4] == ZFV_PERIODIC};
    const double hx = (xmax - xmin) / nx, hy = (ymax - ymin) / ny, hz = (zmax - zmin) / nz;
    auto index = [&](int i, int j, int k) { return nz * (ny * i + j) + k; };
    auto vert = [&](int i, int j, int k) {
        return V3{xmin + ((xmax - xmin) * i) / nx, ymin + ((ymax - ymin) * j) / ny, zmin + ((zmax - zmin) * k) / nz};
    };
    const double ar[6] = {hy * hz, hy * hz, hx * hz, hx * hz, hx * hy, hx * hy};
    for (int c = 0; c < n; ++c) {
        const int i = (c / nz) / ny, j = (c / nz) % ny, k = c % nz;
        // corners: a = x offset, b = y offset, d = z offset
        V3 p[2][2][2];
        for (int a = 0; a < 2; ++a)
            for (int b = 0; b < 2; ++b)
                for (int d = 0; d < 2; ++d) p[a][b][d] = vert(i + a, j + b, k + d);
        const V3 sum8 = ((((((p[0][0][0] + p[1][0][0]) + p[0][1][0]) + p[1][1][0]) + p[0][0][1]) + p[1][0][1]) +
                         p[0][1][1]) + p[1][1][1];
        const V3 cc = mul(0.125, sum8);
        auto quarter = [](const V3& a, const V3& b, const V3& c2, const V3& d) { return mul(0.25, ((a + b) + c2) + d); };
        const V3 fc[6] = {quarter(p[0][0][0], p[0][0][1], p[0][1][0], p[0][1][1]),
                          quarter(p[1][0][0], p[1][0][1], p[1][1][0], p[1][1][1]),
                          quarter(p[0][0][0], p[0][0][1], p[1][0][0], p[1][0][1]),
                          quarter(p[0][1][0], p[0][1][1], p[1][1][0], p[1][1][1]),
                          quarter(p[0][0][0], p[1][0][0], p[0][1][0], p[1][1][0]),
                          quarter(p[0][0][1], p[1][0][1], p[0][1][1], p[1][1][1])};
        const bool lo[3] = {i == 0, j == 0, k == 0}, hi[3] = {i == nx - 1, j == ny - 1, k == nz - 1};
        const int nb[6] = {index((i - 1 + nx) % nx, j, k), index((i + 1) % nx, j, k),
                           index(i, (j - 1 + ny) % ny, k), index(i, (j + 1) % ny, k),
                           index(i, j, (k - 1 + nz) % nz), index(i, j, (k + 1) % nz)};
        m->cx[c] = cc.x; m->cy[c] = cc.y; m->cz[c] = cc.z;
        m->vol[c] = hx * hy * hz;
        auto sq = [](const V3& a) { return dot3(a, a); };
        m->h[c] = std::sqrt(smin_h(sq(fc[1] - fc[0]), smin_h(sq(fc[3] - fc[2]), sq(fc[5] - fc[4]))));
        for (int s = 0; s < kSides; ++s) {
            const int a = s >> 1;
            const bool edge = (s & 1) ? hi[a] : lo[a];
            const size_t o = (size_t)kSides * c + s;
            m->bnd[o] = edge ? bd[s] : ZFV_INNER;
            m->adj[o] = (edge && !per[a]) ? c : nb[s];
            m->fcx[o] = fc[s].x; m->fcy[o] = fc[s].y; m->fcz[o] = fc[s].z;
            m->area[o] = ar[s];
        }
    }
    for (int c = 0; c < n; ++c) {
        const double w0 = 1.0 / std::cbrt(m->vol[c]);
        const V3 pc{m->cx[c], m->cy[c], m->cz[c]};
        for (int s = 0; s < kSides; ++s) {
            const size_t o = (size_t)kSides * c + s;
            V3 nrm{0.0, 0.0, 0.0};
            (s >> 1 == 0 ? nrm.x : (s >> 1 == 1 ? nrm.y : nrm.z)) = (s & 1) ? 1.0 : -1.0;
            V3 q;
            if (m->bnd[o] == ZFV_INNER) {
                const int b = m->adj[o];
                q = V3{m->cx[b], m->cy[b], m->cz[b]};
            } else {   // EuFace::symm_point (amr_faces.h:197-199)
                const double d = 2.0 * dot3(V3{m->fcx[o], m->fcy[o], m->fcz[o]} - pc, nrm);
                q = pc + mul(d, nrm);
            }
            const V3 dr = q - pc;
            const V3 drn = mul(dot3(dr, nrm), nrm);
            m->w[o] = w0 * m->area[o] / std::pow(std::sqrt(dot3(drn, drn)), 3);
        }
    }
}

// The Cuboid's geometry as 1D tables (the layout of Dev3Mesh::tab): each
// value taken from the first cell along its axis, then every cell and slot
// compared with them bit for bit; false (no tables) on any difference.
bool build_tables(const zfv3_ctx* m, std::vector<double>& tab, int bcode[6], int per[3]) {
    const int nx = m->nx, ny = m->ny, nz = m->nz, L = nx + ny + nz;
    const int len[3] = {nx, ny, nz}, base[3] = {0, nx, nx + ny}, st[3] = {ny * nz, nz, 1};
    tab.assign(tab3_size(nx, ny, nz), 0.0);
    auto slot = [](int c, int s) { return (size_t)kSides * c + s; };
    const std::vector<double>* fcs[3] = {&m->fcx, &m->fcy, &m->fcz};
    const std::vector<double>* ccs[3] = {&m->cx, &m->cy, &m->cz};
    for (int a = 0; a < 3; ++a)
        for (int t = 0; t < len[a]; ++t) {
            const int c = t * st[a];
            tab[18 * (size_t)L + base[a] + t] = (*ccs[a])[c];
            for (int s = 0; s < kSides; ++s) tab[(size_t)(3 * s + a) * L + base[a] + t] = (*fcs[a])[slot(c, s)];
        }
    for (int s = 0; s < kSides; ++s) {
        const int b = s >> 1;
        for (int t = 0; t < len[b]; ++t) tab[19 * (size_t)L + (size_t)s * L + base[b] + t] = m->w[slot(t * st[b], s)];
        tab[25 * (size_t)L + s] = m->area[slot(0, s)];
        const int edge_cell = (s & 1) ? (len[b] - 1) * st[b] : 0;
        bcode[s] = slot_code(m->bnd[slot(edge_cell, s)]);
    }
    tab[25 * (size_t)L + 6] = m->vol[0];
    for (int a = 0; a < 3; ++a) per[a] = m->bnd[slot((len[a] - 1) * st[a], 2 * a + 1)] == ZFV_PERIODIC;
    for (int c = 0; c < m->n; ++c) {
        const int idx[3] = {c / st[0], (c % st[0]) / nz, c % nz};
        if (m->vol[c] != tab[25 * (size_t)L + 6]) return false;
        for (int a = 0; a < 3; ++a)
            if ((*ccs[a])[c] != tab[18 * (size_t)L + base[a] + idx[a]]) return false;
        for (int s = 0; s < kSides; ++s) {
            const int b = s >> 1, up = s & 1;
            const bool edge = up ? idx[b] == len[b] - 1 : idx[b] == 0;
            const size_t o = slot(c, s);
            const int code = edge ? bcode[s] : SC_INNER;
            int nb = up ? c + st[b] : c - st[b];
            if (edge) nb = per[b] ? (up ? c - (len[b] - 1) * st[b] : c + (len[b] - 1) * st[b]) : c;
            if (slot_code(m->bnd[o]) != code || m->adj[o] != nb) return false;
            if (m->area[o] != tab[25 * (size_t)L + s]) return false;
            if (m->w[o] != tab[19 * (size_t)L + (size_t)s * L + base[b] + idx[b]]) return false;
            for (int a = 0; a < 3; ++a)
                if ((*fcs[a])[o] != tab[(size_t)(3 * s + a) * L + base[a] + idx[a]]) return false;
        }
    }
    return true;
}

int upload_mesh(zfv3_ctx* ctx) {
    const int n = ctx->n;
    const size_t N = (size_t)n, S = (size_t)kSides * n;
    ctx->release();
    int r;
    if ((r = ctx->alloc(&ctx->d_cell, 5 * N)) || (r = ctx->alloc(&ctx->d_type, S)) ||
        (r = ctx->alloc(&ctx->d_adj, S)) || (r = ctx->alloc(&ctx->d_slot, 5 * S)) ||
        (r = ctx->alloc(&ctx->d_z[0], 6 * N)) || (r = ctx->alloc(&ctx->d_z[1], 6 * N)) ||
        (r = ctx->alloc(&ctx->d_z[2], 6 * N)) || (r = ctx->alloc(&ctx->d_G, 15 * N)) ||
        (r = ctx->alloc(&ctx->d_wait[0], N)) || (r = ctx->alloc(&ctx->d_wait[1], N)) ||
        (r = ctx->alloc(&ctx->d_slots, 3 * (size_t)kSlotWords)) || (r = ctx->alloc(&ctx->d_dt, 1)) || (r = ctx->alloc(&ctx->d_err, 1)))
        return r;
    ctx->blocks = (int)((N + 255) / 256);
    if ((r = ctx->alloc(&ctx->d_part, 5 * (size_t)ctx->blocks))) return r;
    std::vector<double> tab;
    if (ctx->use_tables && build_tables(ctx, tab, ctx->tab_bcode, ctx->tab_per)) {
        if ((r = ctx->alloc(&ctx->d_tab, tab.size()))) return r;
        Z3HIP(hipMemcpy(ctx->d_tab, tab.data(), tab.size() * sizeof(double), hipMemcpyHostToDevice));
    }
    // host staging in the device layout
    std::vector<double> cell(5 * N), slot(5 * S);
    std::vector<int> type(S), adj(S);
    const std::vector<double>* cs[5] = {&ctx->cx, &ctx->cy, &ctx->cz, &ctx->vol, &ctx->h};
    for (int k = 0; k < 5; ++k) std::memcpy(&cell[k * N], cs[k]->data(), N * sizeof(double));
    const std::vector<double>* ss[5] = {&ctx->fcx, &ctx->fcy, &ctx->fcz, &ctx->area, &ctx->w};
    for (int c = 0; c < n; ++c)
        for (int s = 0; s < kSides; ++s) {
            const size_t o = (size_t)kSides * c + s, d = (size_t)s * N + c;
            type[d] = slot_code(ctx->bnd[o]);
            adj[d] = ctx->adj[o];
            for (int k = 0; k < 5; ++k) slot[k * S + d] = (*ss[k])[o];
        }
    Z3HIP(hipMemcpy(ctx->d_cell, cell.data(), cell.size() * sizeof(double), hipMemcpyHostToDevice));
    Z3HIP(hipMemcpy(ctx->d_slot, slot.data(), slot.size() * sizeof(double), hipMemcpyHostToDevice));
    Z3HIP(hipMemcpy(ctx->d_type, type.data(), S * sizeof(int), hipMemcpyHostToDevice));
    Z3HIP(hipMemcpy(ctx->d_adj, adj.data(), S * sizeof(int), hipMemcpyHostToDevice));
    const std::vector<unsigned long long> init(3 * (size_t)kSlotWords, kDblMaxBits);
    Z3HIP(hipMemcpy(ctx->d_slots, init.data(), init.size() * sizeof(unsigned long long), hipMemcpyHostToDevice));
    Z3HIP(hipMemset(ctx->d_err, 0, sizeof(int)));
    Z3HIP(hipMemset(ctx->d_wait[0], 0, N * sizeof(int)));
    Z3HIP(hipMemset(ctx->d_wait[1], 0, N * sizeof(int)));
    Z3HIP(hipMemset(ctx->d_G, 0, 15 * N * sizeof(double)));
    ctx->zi = 0; ctx->hi = 1; ctx->ni = 2; ctx->wi = 0; ctx->cur = 0;
    ctx->have_state = false;
    ctx->dt_valid = false;
    ctx->dt_ready = false;
    return ZFV_OK;
}

// A new state: the current slot is cleared for a fresh CFL reduction.
int state_changed(zfv3_ctx* ctx) {
    const std::vector<unsigned long long> v(kSlotWords, kDblMaxBits);
    Z3HIP(hipMemcpy(ctx->slot(ctx->cur), v.data(), v.size() * sizeof(unsigned long long), hipMemcpyHostToDevice));
    ctx->dt_ready = false;
    ctx->dt_valid = false;
    return ZFV_OK;
}

}  // namespace

extern "C" {

int zfv3_create(int device, zfv3_ctx** out) {
    if (!out) return fail(ZFV_ERR_ARG, "null output pointer");
    *out = nullptr;
    int count = 0;
    Z3HIP(hipGetDeviceCount(&count));
    if (device < 0 || device >= count) return fail(ZFV_ERR_ARG, "device index out of range");
    Z3HIP(hipSetDevice(device));
    zfv3_ctx* ctx = new zfv3_ctx();
    ctx->device = device;
    hipError_t e = hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking);
    if (e != hipSuccess) {
        delete ctx;
        return fail(ZFV_ERR_HIP, std::string("hipStreamCreateWithFlags: ") + hipGetErrorString(e));
    }
    *out = ctx;
    return ZFV_OK;
}

int zfv3_destroy(zfv3_ctx* ctx) {
    if (!ctx) return ZFV_OK;
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    delete ctx;
    return ZFV_OK;
}

int zfv3_mesh_cuboid(zfv3_ctx* ctx, int nx, int ny, int nz, double xmin, double xmax, double ymin, double ymax,
                     double zmin, double zmax, const int32_t* bounds) {
    if (!ctx) return fail(ZFV_ERR_ARG, "null context");
    if (nx < 1 || ny < 1 || nz < 1) return fail(ZFV_ERR_ARG, "nx, ny, nz must be >= 1");
    if ((long long)nx * ny * nz > INT_MAX / kSlots24) return fail(ZFV_ERR_ARG, "mesh too large");
    if (!(xmax > xmin) || !(ymax > ymin) || !(zmax > zmin)) return fail(ZFV_ERR_ARG, "empty box");
    int bd[6] = {ZFV_ZOE, ZFV_ZOE, ZFV_ZOE, ZFV_ZOE, ZFV_ZOE, ZFV_ZOE};
    if (bounds)
        for (int k = 0; k < 6; ++k) {
            if (bounds[k] != ZFV_WALL && bounds[k] != ZFV_ZOE && bounds[k] != ZFV_PERIODIC &&
                bounds[k] != ZFV_FUNCTION)
                return fail(ZFV_ERR_ARG, "bounds must be WALL, ZOE, PERIODIC or FUNCTION");
            bd[k] = bounds[k];
        }
    // a periodic side makes its direction periodic on both ends
    for (int d = 0; d < 3; ++d)
        if (bd[2 * d] == ZFV_PERIODIC || bd[2 * d + 1] == ZFV_PERIODIC) bd[2 * d] = bd[2 * d + 1] = ZFV_PERIODIC;
    Z3HIP(hipSetDevice(ctx->device));
    Z3HIP(hipStreamSynchronize(ctx->stream));
    build_cuboid(ctx, nx, ny, nz, xmin, xmax, ymin, ymax, zmin, zmax, bd);
    return upload_mesh(ctx);
}

int zfv3_set_geometry_tables(zfv3_ctx* ctx, int on) {
    if (!ctx) return fail(ZFV_ERR_ARG, "null context");
    ctx->use_tables = on != 0;
    return ZFV_OK;
}

int zfv3_get_geometry_tables(zfv3_ctx* ctx, int* active) {
    if (!ctx || !active) return fail(ZFV_ERR_ARG, "null argument");
    *active = ctx->d_tab != nullptr;
    return ZFV_OK;
}

int zfv3_mesh_n_cells(zfv3_ctx* ctx, int32_t* n) {
    if (!ctx || !n) return fail(ZFV_ERR_ARG, "null argument");
    *n = ctx->n;
    return ZFV_OK;
}

int zfv3_mesh_download(zfv3_ctx* ctx, double* cx, double* cy, double* cz, double* vol, double* h,
                       int32_t* boundary, int32_t* adj, double* nx, double* ny, double* nz, double* fcx,
                       double* fcy, double* fcz, double* area, double* w) {
    int r = need_mesh(ctx);
    if (r) return r;
    const size_t N = (size_t)ctx->n;
    double* cdst[5] = {cx, cy, cz, vol, h};
    const std::vector<double>* csrc[5] = {&ctx->cx, &ctx->cy, &ctx->cz, &ctx->vol, &ctx->h};
    for (int k = 0; k < 5; ++k)
        if (cdst[k]) std::memcpy(cdst[k], csrc[k]->data(), N * sizeof(double));
    double* sdst[9] = {nx, ny, nz, fcx, fcy, fcz, area, w, nullptr};
    const std::vector<double>* ssrc[8] = {nullptr, nullptr, nullptr, &ctx->fcx, &ctx->fcy, &ctx->fcz, &ctx->area,
                                          &ctx->w};
    for (size_t c = 0; c < N; ++c)
        for (int s24 = 0; s24 < kSlots24; ++s24) {
            const size_t o24 = (size_t)kSlots24 * c + s24;
            const bool main = s24 < kSides;
            const size_t o = (size_t)kSides * c + s24;
            if (boundary) boundary[o24] = main ? ctx->bnd[o] : ZFV_UNDEFINED;
            if (adj) adj[o24] = main ? ctx->adj[o] : -1;
            for (int k = 0; k < 8; ++k) {
                if (!sdst[k]) continue;
                double v = 0.0;
                if (main) {
                    if (k < 3) v = (k == (s24 >> 1)) ? ((s24 & 1) ? 1.0 : -1.0) : 0.0;
                    else v = (*ssrc[k])[o];
                }
                sdst[k][o24] = v;
            }
        }
    return ZFV_OK;
}

int zfv3_smf_configure(zfv3_ctx* ctx, const zfv_smf_params* p) {
    if (!ctx || !p) return fail(ZFV_ERR_ARG, "null argument");
    if (p->flux < ZFV_GODUNOV || p->flux > ZFV_CIR2)
        return fail(ZFV_ERR_UNSUPPORTED, "the 3D step takes the single-material fluxes GODUNOV..CIR2");
    if (p->limiter < ZFV_LIM_NONE || p->limiter > ZFV_LIM_SUPERBEE) return fail(ZFV_ERR_ARG, "unknown limiter");
    if (p->axial) return fail(ZFV_ERR_UNSUPPORTED, "set_axial applies to two-dimensional meshes");
    if (!(p->gamma > 1.0)) return fail(ZFV_ERR_ARG, "gamma must be > 1");
    ctx->prm = *p;
    ctx->prm.acc = p->acc < 1 ? 1 : (p->acc > 2 ? 2 : p->acc);
    ctx->prm.cfl = std::fmax(0.0, std::fmin(p->cfl, 1.0));
    ctx->configured = true;
    return ZFV_OK;
}

int zfv3_smf_set_max_dt(zfv3_ctx* ctx, double max_dt) {
    if (!ctx) return fail(ZFV_ERR_ARG, "null context");
    if (!ctx->configured) return fail(ZFV_ERR_STATE, "solver not configured");
    ctx->prm.max_dt = max_dt;
    return ZFV_OK;
}

int zfv3_state_upload(zfv3_ctx* ctx, const double* rho, const double* u, const double* v, const double* w,
                      const double* p, const double* e) {
    int r = need_ready(ctx, false);
    if (r) return r;
    if (!rho || !u || !v || !w || !p) return fail(ZFV_ERR_ARG, "null state array");
    const size_t N = (size_t)ctx->n;
    std::vector<double> ebuf;
    if (!e) {   // energy_rP (stiffened_gas.cpp:121-122)
        const Eos q = ctx->eos();
        ebuf.resize(N);
        for (size_t c = 0; c < N; ++c) ebuf[c] = q.e0 + (p[c] + q.g * q.P0) / ((q.g - 1.0) * rho[c]);
        e = ebuf.data();
    }
    const double* src[6] = {rho, u, v, w, p, e};
    Z3HIP(hipStreamSynchronize(ctx->stream));
    for (int k = 0; k < 6; ++k)
        Z3HIP(hipMemcpy(ctx->d_z[ctx->zi] + k * N, src[k], N * sizeof(double), hipMemcpyHostToDevice));
    ctx->have_state = true;
    return state_changed(ctx);
}

int zfv3_state_download(zfv3_ctx* ctx, double* rho, double* u, double* v, double* w, double* p, double* e) {
    int r = need_ready(ctx, true);
    if (r) return r;
    const size_t N = (size_t)ctx->n;
    double* dst[6] = {rho, u, v, w, p, e};
    for (int k = 0; k < 6; ++k)
        if (dst[k])
            Z3HIP(hipMemcpyAsync(dst[k], ctx->d_z[ctx->zi] + k * N, N * sizeof(double), hipMemcpyDeviceToHost,
                                 ctx->stream));
    return read_err(ctx);
}

int zfv3_wait_upload(zfv3_ctx* ctx, const int32_t* wait) {
    int r = need_mesh(ctx);
    if (r) return r;
    if (!wait) return fail(ZFV_ERR_ARG, "null wait array");
    Z3HIP(hipStreamSynchronize(ctx->stream));
    Z3HIP(hipMemcpy(ctx->d_wait[ctx->wi], wait, (size_t)ctx->n * sizeof(int), hipMemcpyHostToDevice));
    return ZFV_OK;
}

int zfv3_wait_download(zfv3_ctx* ctx, int32_t* wait) {
    int r = need_mesh(ctx);
    if (r) return r;
    if (!wait) return fail(ZFV_ERR_ARG, "null wait array");
    Z3HIP(hipMemcpyAsync(wait, ctx->d_wait[ctx->wi], (size_t)ctx->n * sizeof(int), hipMemcpyDeviceToHost,
                         ctx->stream));
    return read_err(ctx);
}

int zfv3_grad_download(zfv3_ctx* ctx, double* g) {
    int r = need_mesh(ctx);
    if (r) return r;
    if (!g) return fail(ZFV_ERR_ARG, "null gradient array");
    Z3HIP(hipMemcpyAsync(g, ctx->d_G, 15 * (size_t)ctx->n * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    return read_err(ctx);
}

}  // extern "C"

namespace {

// A launch, between two recorded events while kernel timing is on.
template <class F>
int timed_launch(zfv3_ctx* ctx, const char* name, F&& launch) {
    if (!ctx->prof_on) {
        Z3HIP(launch());
        return ZFV_OK;
    }
    zfv3_ctx::Timed t{name, nullptr, nullptr};
    Z3HIP(hipEventCreate(&t.a));
    hipError_t e = hipEventCreate(&t.b);
    if (e != hipSuccess) {
        (void)hipEventDestroy(t.a);
        return fail(ZFV_ERR_HIP, std::string("hipEventCreate: ") + hipGetErrorString(e));
    }
    ctx->timed.push_back(t);
    Z3HIP(hipEventRecord(t.a, ctx->stream));
    Z3HIP(launch());
    Z3HIP(hipEventRecord(t.b, ctx->stream));
    return ZFV_OK;
}

}  // namespace

extern "C" {

int zfv3_smf_step(zfv3_ctx* ctx, int n_steps) {
    int r = need_ready(ctx, true);
    if (r) return r;
    if (n_steps < 0) return fail(ZFV_ERR_ARG, "n_steps must be >= 0");
    const Dev3Mesh M = ctx->mesh();
    const Eos eos = ctx->eos();
    const int flux = ctx->prm.flux, lim = ctx->prm.limiter;
    hipStream_t s = ctx->stream;
    for (int it = 0; it < n_steps; ++it) {
        const Dev3Ctl ctl{ctx->slot(ctx->cur), ctx->slot((ctx->cur + 1) % 3), ctx->slot((ctx->cur + 2) % 3),
                          ctx->prm.cfl, ctx->prm.max_dt, ctx->d_dt, ctx->d_err};
        const Dev3State z = ctx->state(ctx->zi), h = ctx->state(ctx->hi), nx = ctx->state(ctx->ni);
        int* w0 = ctx->d_wait[ctx->wi];
        int* w1 = ctx->d_wait[ctx->wi ^ 1];
        if (!ctx->dt_ready && (r = timed_launch(ctx, "dt", [&] { return launch3_dt(M, z, eos, ctl.dtmin, s); })))
            return r;
        if (ctx->prm.acc == 1) {
            if ((r = timed_launch(ctx, "flux1", [&] { return launch3_flux1(flux, M, z, nx, eos, ctl, s); }))) return r;
        } else {
            if ((r = timed_launch(ctx, "grad_pred",
                                  [&] { return launch3_grad_pred(lim, M, z, ctx->d_G, h, w0, eos, ctl, s); })))
                return r;
            if ((r = timed_launch(ctx, "corr",
                                  [&] { return launch3_corr(flux, M, z, h, ctx->d_G, w0, w1, nx, eos, ctl, s); })))
                return r;
            ctx->wi ^= 1;
        }
        const int t = ctx->zi;
        ctx->zi = ctx->ni;
        ctx->ni = t;
        ctx->cur = (ctx->cur + 1) % 3;
        ctx->dt_valid = true;
        ctx->dt_ready = true;
    }
    return ZFV_OK;
}

int zfv3_smf_dt(zfv3_ctx* ctx, double* dt) {
    int r = need_ready(ctx, true);
    if (r) return r;
    if (!dt) return fail(ZFV_ERR_ARG, "null output");
    if (!ctx->dt_valid) return fail(ZFV_ERR_STATE, "no step taken since the state was set");
    Z3HIP(hipMemcpyAsync(dt, ctx->d_dt, sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    return read_err(ctx);
}

int zfv3_smf_compute_dt(zfv3_ctx* ctx, double* dt) {
    int r = need_ready(ctx, true);
    if (r) return r;
    if (!dt) return fail(ZFV_ERR_ARG, "null output");
    unsigned long long* slot = ctx->slot(ctx->cur);
    if (!ctx->dt_ready) {
        Z3HIP(launch3_dt(ctx->mesh(), ctx->state(ctx->zi), ctx->eos(), slot, ctx->stream));
        ctx->dt_ready = true;
    }
    std::vector<unsigned long long> bits(kSlotWords);
    Z3HIP(hipMemcpyAsync(bits.data(), slot, bits.size() * sizeof(unsigned long long), hipMemcpyDeviceToHost,
                         ctx->stream));
    Z3HIP(hipStreamSynchronize(ctx->stream));
    double dmin = 0.0;
    for (int k = 0; k < kSubSlots; ++k) {
        double v;
        std::memcpy(&v, &bits[(size_t)kSubStride * k], sizeof(v));
        if (k == 0 || v < dmin) dmin = v;
    }
    const double t = ctx->prm.cfl * dmin;
    *dt = (ctx->prm.max_dt < t) ? ctx->prm.max_dt : t;
    return read_err(ctx);
}

int zfv3_conserved_totals(zfv3_ctx* ctx, double* out5) {
    int r = need_ready(ctx, true);
    if (r) return r;
    if (!out5) return fail(ZFV_ERR_ARG, "null output");
    Z3HIP(launch3_totals(ctx->mesh(), ctx->state(ctx->zi), ctx->d_part, ctx->blocks, ctx->stream));
    std::vector<double> part(5 * (size_t)ctx->blocks);
    Z3HIP(hipMemcpyAsync(part.data(), ctx->d_part, part.size() * sizeof(double), hipMemcpyDeviceToHost,
                         ctx->stream));
    r = read_err(ctx);
    if (r) return r;
    for (int k = 0; k < 5; ++k) out5[k] = 0.0;
    for (int b = 0; b < ctx->blocks; ++b)
        for (int k = 0; k < 5; ++k) out5[k] += part[5 * (size_t)b + k];
    return ZFV_OK;
}

int zfv3_profile_enable(zfv3_ctx* ctx, int on) {
    if (!ctx) return fail(ZFV_ERR_ARG, "null context");
    (void)hipSetDevice(ctx->device);
    Z3HIP(hipStreamSynchronize(ctx->stream));
    ctx->drop_events();
    ctx->prof_on = on != 0;
    return ZFV_OK;
}

int zfv3_profile_read(zfv3_ctx* ctx, const char** names, double* ms, int32_t* launches, int32_t* count) {
    if (!ctx || !count) return fail(ZFV_ERR_ARG, "null argument");
    (void)hipSetDevice(ctx->device);
    Z3HIP(hipStreamSynchronize(ctx->stream));
    std::vector<const char*> nm;
    std::vector<double> tot;
    std::vector<int> cnt;
    for (auto& t : ctx->timed) {
        float el = 0.0f;
        Z3HIP(hipEventElapsedTime(&el, t.a, t.b));
        size_t k = 0;
        while (k < nm.size() && std::strcmp(nm[k], t.name) != 0) ++k;
        if (k == nm.size()) {
            nm.push_back(t.name);
            tot.push_back(0.0);
            cnt.push_back(0);
        }
        tot[k] += el;
        cnt[k] += 1;
    }
    ctx->drop_events();
    const int cap = *count;
    const int used = (int)nm.size() < cap ? (int)nm.size() : cap;
    for (int k = 0; k < used; ++k) {
        if (names) names[k] = nm[k];
        if (ms) ms[k] = tot[k];
        if (launches) launches[k] = cnt[k];
    }
    *count = used;
    return ZFV_OK;
}

int zfv3_synchronize(zfv3_ctx* ctx) {
    if (!ctx) return fail(ZFV_ERR_ARG, "null context");
    (void)hipSetDevice(ctx->device);
    if (!ctx->d_err) return ZFV_OK;
    return read_err(ctx);
}

}  // extern "C"
