// C ABI of the MI355X finite-volume step (include/zfv.h): context, mesh
// upload and validation, the Rectangle AMR-2D generator, and the step
// orchestration over the HIP kernels of zfv_kernels.hip.
#include "../../include/zfv.h"
#include "zfv_internal.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <limits>
#include <map>
#include <mutex>
#include <string>
#include <vector>

using namespace zfv;

namespace {

thread_local std::string g_last_error;

int fail(int code, const std::string& msg) {
    g_last_error = msg;
    return code;
}

#define ZFV_HIP(call)                                                                     \
    do {                                                                                  \
        hipError_t e_ = (call);                                                           \
        if (e_ != hipSuccess)                                                             \
            return fail(ZFV_ERR_HIP, std::string(#call) + ": " + hipGetErrorString(e_)); \
    } while (0)

template <class T>
struct DBuf {
    T* p = nullptr;
    size_t n = 0;
    DBuf() = default;
    // Owns its allocation: movable (containers of buffers may reallocate),
    // never copied.
    DBuf(const DBuf&) = delete;
    DBuf& operator=(const DBuf&) = delete;
    DBuf(DBuf&& o) noexcept : p(o.p), n(o.n) {
        o.p = nullptr;
        o.n = 0;
    }
    DBuf& operator=(DBuf&& o) noexcept {
        if (this != &o) {
            release();
            p = o.p;
            n = o.n;
            o.p = nullptr;
            o.n = 0;
        }
        return *this;
    }
    ~DBuf() { release(); }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
    }
    hipError_t alloc(size_t count) {
        if (count == n && p) return hipSuccess;
        release();
        if (count == 0) return hipSuccess;
        hipError_t e = hipMalloc(&p, count * sizeof(T));
        if (e == hipSuccess) n = count;
        return e;
    }
};

// Host copy of the uploaded mesh in the reference's AmrCells layout.
struct HostMesh {
    int n = 0;
    std::vector<double> cx, cy, vol, vx, vy;
    std::vector<int32_t> level, flag, b_idx, z_idx;
    std::vector<int32_t> bnd, adj;
    std::vector<double> nx, ny, fcx, fcy, area;
    // axisymmetric geometry (AmrCells::volume_alt, AmrFaces::area_alt),
    // empty for a planar mesh
    std::vector<double> vol_alt, area_alt;
};

struct KernelTimer {
    std::string name;
    hipEvent_t t0, t1;
};

}  // namespace

struct zfv_ctx {
    int device = 0;
    hipStream_t own_stream = nullptr;
    hipStream_t stream = nullptr;

    HostMesh hm;
    bool have_mesh = false, have_state = false, configured = false;

    // device mesh (capacity `cap` cells; slot-major arrays have stride cap)
    int cap = 0;
    bool host_stale = false;   // hm out of date after a device-side adaptation
    DBuf<double> cx, cy, vol, h, area, fcx, fcy, w, nx, ny, vx, vy;
    DBuf<int> code, adj, bnd, level, flag, flag0, b_idx, z_idx, next, idx;
    // axisymmetric geometry: volume_alt, area_alt [8][cap], volume / volume_alt
    bool axial_mesh = false;
    DBuf<double> vol_alt, area_alt, axc;
    // AMR scratch
    DBuf<AmrMarks> s_marks, s_rank;
    DBuf<unsigned long long> s_faces;
    DBuf<int> s_hole_at, s_ref_list, s_crs_list, s_swap_src, s_swap_dst;
    DBuf<AmrMarks> s_tiles;           // tile sums of the marks scan
    DBuf<int> s_fam;                  // children of each coarsened family in z order (amr_positions)
    AmrMarks* marks_total = nullptr;  // pinned host: totals of the marks scan
    bool faces_pending = false;   // n_faces still on the device (s_faces)
    // device state
    DBuf<double> st[3][5];   // init / next / half layers
    DBuf<double> gdx, gdy;
    DBuf<double> graw;       // unlimited LSM gradient of rho [2][cap] (refinement criterion)
    DBuf<int> wait[2];
    int cur = 0, wcur = 0;   // index of the init layer / current wait buffer
    DBuf<unsigned long long> dtmin;
    DBuf<double> dt_out, partial;
    DBuf<double> dmax_dt;   // set_max_dt value on the device
    DBuf<int> err;

    int64_t n_faces = 0;   // unique faces: interior slots / 2 + boundary slots

    zfv_smf_params prm{};
    bool dt_valid = false;
    long long step_no = 0;

    // step-kernel layout per stage (kLayoutCell / kLayoutQuad / kLayoutBatch): first-order
    // update, gradient + predictor, corrector
    int layout[3] = {kLayoutCell, kLayoutCell, kLayoutCell};
    int block[3] = {kBlock, kBlock, kBlock};   // workgroup size per stage
    bool layout_set[3] = {false, false, false};   // chosen by zfv_set_launch / zfv_autotune (else the default)

    // Structured rectangles (zfv_tile.hip): the Rectangle problems the host
    // built (set before the upload, empty for an uploaded mesh) and, when
    // the uploaded slot arrays match per-column / per-row tables cell by
    // cell, the tables of the tiled MUSCL step.
    struct RectInfo {
        int off, nx, ny;
    };
    std::vector<RectInfo> rects;
    struct Tiles {
        bool ok = false;
        int n_probs = 0;
        int n_tiles[kTileVariants] = {0, 0, 0};
        DBuf<TileProb> prob;
        DBuf<TileCol> col;
        DBuf<TileRow> row;
    } tiles;
    int tile_variant = 0;
    bool tile_persistent = false;   // tiled step by persistent workgroups (else one workgroup per tile)
    // The tiled step does not store its gradients: after one, the gradients
    // of the last step are recomputed on demand from its input (layer 1 - cur).
    bool grad_pending = false;

    // batch of independent problems (nseg > 1): padded cell offsets, valid
    // cell counts, device tables of the kernels' Tile lookup
    int nseg = 1;
    std::vector<int> seg_begin_h, seg_n_h;
    DBuf<int> chunk_pid, seg_begin, seg_end;

    // captured step graphs: kGraphSteps steps, keyed by the phase they start
    // in (dt slot rotation, layer and wait buffer parity, scheme)
    bool use_graphs = false;
    struct GraphEntry {
        int key;
        hipGraphExec_t exec;
    };
    std::vector<GraphEntry> graphs;
    void graphs_clear() {
        for (auto& g : graphs) (void)hipGraphExecDestroy(g.exec);
        graphs.clear();
    }

    // discrete adjoint: checkpoints of the forward run and reverse-sweep buffers
    struct Adjoint {
        int n_steps = 0, every = 0;
        double win[4] = {0, 0, 0, 0};
        bool ready = false;
        std::vector<DBuf<double>> ckpt;   // [5][n] state every `every` steps
        DBuf<double> seg, final_z;         // segment states [every][5][n] (entry 0 unused), final state [5][n]
        DBuf<double> seg_g, seg_h;         // the steps' gradients [every][8][cap] and half states [every][5][n]
        DBuf<unsigned long long> seg_dt;   // the steps' dt slots [every]
        DBuf<double> lam[2], lz, W, lG, ZB, NG, dtb;
    } adjoint;

    // multi-material solver (MmFluid, first or second order, CrpMode PLIC or
    // NONE): two state layers of (7 + 2 n_mat) rows of cap doubles (a third,
    // the half step, for acc 2), its own dt slots and step count
    struct MmSolver {
        bool configured = false, have_state = false;
        long long epoch = -1;   // mesh epoch of the state
        mm::Mix mix{};
        int crp = ZFV_CRP_PLIC;   // CrpMode (mm_fluid.cpp:18)
        int flux = ZFV_HLLC;      // m_nf of the plain faces
        int split = ZFV_SPLIT_NONE;
        double cfl = 0.7, max_dt = 1.0e300;   // MmFluid defaults (mm_fluid.cpp:16-26)
        int cur = 0;
        long long step_no = 0;
        DBuf<double> st[2], half, dmax_dt, dt_out, inter, cf;
        DBuf<double> gx, gy;          // limited gradients (compute_grad), [6 + 2 n_mat][cap]
        long long grad_epoch = -1;    // mesh epoch of gx / gy
        int grad_rows = 0;            // rows of gx / gy the last compute_grad filled
        int limiter = 3;              // MC (mm_fluid.cpp:20)
        int acc = 1;                  // m_acc (mm_fluid.h:157)
        DBuf<int> crp_list, crp_cells;
        DBuf<unsigned int> crp_count;   // [0] listed faces, [1] listed cells
        DBuf<unsigned long long> dtmin;
        // The composite fluxes on a second stream (same priority), beside
        // the update of the cells without a composite face: `side` and the
        // fork / join events are created on the first overlapped step.
        int overlap = 0;        // zfv_mm_set_overlap mode
        int crp_groups = 0;     // k_mm_crp workgroups, 0: 2048
        hipStream_t side = nullptr;
        hipEvent_t ev_in = nullptr, ev_mark = nullptr, ev_crp = nullptr;
    } mm;
    // Transfer solver: u in two layers, the interface normals, the slopes,
    // the velocity samples (faces slot-major, cells), its dt and CFL
    struct TransferSolver {
        bool configured = false, have_state = false, have_velocity = false;
        long long epoch = -1;   // mesh epoch of the state and velocity
        int method = 11;        // Transfer::Method (MUSCL_MC)
        double cfl = 0.5, dt = 1.0e300;   // Transfer::Transfer (transfer.cpp:43-48)
        int cur = 0;
        DBuf<double> u[2], nx, ny, gx, gy, fvx, fvy, cvx, cvy;
        DBuf<unsigned long long> slot;
    } tr;
    long long mesh_epoch = 0;   // incremented whenever the mesh changes

    // Curve order of the step kernels (zfv_order.hip): after a device
    // adaptation (or the upload of a mesh with refined cells) the steps run
    // on a copy of the mesh sorted along the base-cell / Morton curve; the
    // state layer `cur`, the current wait buffer and the last step's
    // gradients are then held in that order until an entry point that
    // observes or changes per-cell data (downloads, uploads, flags,
    // adaptation, totals, the adjoint) puts them back in the mesh's order.
    struct CurveOrder {
        int mode = 1;             // 0: always the mesh's order; 1: curve order for steps of refined meshes,
                                  // carried across adaptations; 2: as 1, rebuilt in full after each adaptation
        bool refined = false;     // the mesh has refined cells (curve order differs from row order)
        bool keys_ok = false;     // levels, z and base indices of the uploaded mesh form valid curve keys
        int lmax = 0;             // finest level the keys resolve
        int base_bits = 0;        // bits of the largest base-cell index
        long long n_base = 0;     // base cells (largest b_idx + 1)
        bool slots_clean = false; // the dense slot array holds -1 everywhere
        long long epoch = -1;     // mesh epoch of the sorted copy
        long long perm_epoch = -1;  // mesh epoch whose order perm / iperm already hold (carried by zfv_refine)
        int cap = 0;              // stride of the copy (= the mesh's capacity when built)
        bool state = false;       // state layer cur and wait[wcur] in curve order
        bool grad = false;        // gdx / gdy last written in curve order
        DBuf<double> cx, cy, vol, h, area, fcx, fcy, w, nx, ny, gtmp;
        DBuf<int> code, adj, perm, iperm, vals, slots, counts, offsets, tiles;
        DBuf<unsigned long long> keys[2];
        DBuf<unsigned char> tmp;
    } co;

    bool profile = false;
    std::vector<KernelTimer> timers;   // pending event pairs of the last zfv_smf_step
    std::vector<hipEvent_t> event_pool;
    std::map<std::string, std::pair<double, int>> prof;   // name -> (ms, launches)

    DevMesh dmesh() const {
        DevMesh M;
        M.n = hm.n;
        M.stride = cap;
        M.cx = cx.p; M.cy = cy.p; M.vol = vol.p; M.h = h.p;
        M.code = code.p; M.adj = adj.p; M.area = area.p; M.fcx = fcx.p; M.fcy = fcy.p; M.w = w.p;
        M.nx = nx.p; M.ny = ny.p;
        M.chunk_pid = nseg > 1 ? chunk_pid.p : nullptr;
        M.seg_begin = nseg > 1 ? seg_begin.p : nullptr;
        M.seg_end = nseg > 1 ? seg_end.p : nullptr;
        M.axc = nullptr;
        return M;
    }
    // The mesh of the SmFluid update: with set_axial, cell.volume(true) and
    // face.area(true) in place of the planar ones and the source
    // coefficient (an axisymmetric mesh is never adapted, so never curve
    // ordered).
    DevMesh ax(DevMesh M) const {
        if (prm.axial && axial_mesh) {
            M.vol = vol_alt.p;
            M.area = area_alt.p;
            M.axc = axc.p;
        }
        return M;
    }
    // The mesh in the order the state is currently held in, as the step
    // kernels read it.
    DevMesh smesh() const {
        if (!co.state) return ax(dmesh());
        DevMesh M;
        M.n = hm.n;
        M.stride = co.cap;
        M.cx = co.cx.p; M.cy = co.cy.p; M.vol = co.vol.p; M.h = co.h.p;
        M.code = co.code.p; M.adj = co.adj.p; M.area = co.area.p; M.fcx = co.fcx.p; M.fcy = co.fcy.p;
        M.w = co.w.p; M.nx = co.nx.p; M.ny = co.ny.p;
        M.chunk_pid = nullptr;
        M.seg_begin = nullptr;
        M.seg_end = nullptr;
        M.axc = nullptr;
        return M;
    }
    TileMesh tmesh() const {
        TileMesh T;
        T.n_probs = tiles.n_probs;
        for (int v = 0; v < kTileVariants; ++v) T.n_tiles[v] = tiles.n_tiles[v];
        T.prob = tiles.prob.p;
        T.col = tiles.col.p;
        T.row = tiles.row.p;
        T.vol = vol.p;
        T.h = h.p;
        return T;
    }
    // Whether the steps run the tiled kernel: structured rectangles, MUSCL
    // with HLLC, the state in the mesh's order.
    bool tile_step() const {
        return tiles.ok && prm.acc == 2 && prm.flux == ZFV_HLLC && layout[1] == kLayoutTile && !co.state &&
               !prm.axial;
    }
    // Launch configuration of the gradient kernel when it runs on its own.
    int gp_layout() const { return layout[1] == kLayoutTile ? kLayoutCell : layout[1]; }
    int gp_block() const { return layout[1] == kLayoutTile ? kBlock : block[1]; }
    AmrMesh amesh() const {
        AmrMesh A;
        A.n = hm.n;
        A.stride = cap;
        A.cx = cx.p; A.cy = cy.p; A.vol = vol.p; A.h = h.p;
        A.code = code.p; A.adj = adj.p; A.bnd = bnd.p;
        A.area = area.p; A.fcx = fcx.p; A.fcy = fcy.p; A.w = w.p; A.nx = nx.p; A.ny = ny.p;
        A.vx = vx.p; A.vy = vy.p;
        A.level = level.p; A.flag = flag.p; A.flag0 = flag0.p; A.b_idx = b_idx.p; A.z_idx = z_idx.p;
        A.next = next.p; A.idx = idx.p;
        return A;
    }
    DevState layer(int i) const {
        return DevState{st[i][0].p, st[i][1].p, st[i][2].p, st[i][3].p, st[i][4].p};
    }
    Eos eos() const { return Eos{prm.gamma, prm.P0, prm.e0}; }
};

namespace {

int check_ctx(zfv_ctx* ctx) {
    if (!ctx) return fail(ZFV_ERR_ARG, "null context");
    hipError_t e = hipSetDevice(ctx->device);
    if (e != hipSuccess) return fail(ZFV_ERR_HIP, std::string("hipSetDevice: ") + hipGetErrorString(e));
    return ZFV_OK;
}

int code_of_boundary(int b) {
    switch (b) {
        case ZFV_UNDEFINED: return SC_UNDEF;
        case ZFV_INNER: return SC_INNER;
        case ZFV_PERIODIC: return SC_PERIODIC;
        case ZFV_WALL: return SC_WALL;
        case ZFV_ZOE:
        case ZFV_FUNCTION: return SC_COPY;
        default: return -1;
    }
}

// Canonical outward normals of the AMR-2D sides L, R, B, T.
const double kNX[4] = {-1.0, 1.0, 0.0, 0.0};
const double kNY[4] = {0.0, 0.0, -1.0, 1.0};

// Incircle diameter of an adaptive 2D cell from its SqQuad vertices
// (AmrCells::incircle_diameter, mesh/euler/amr_cells.cpp:107-114).
double incircle_2d(const double* X, const double* Y) {
    double ax = X[5] - X[3], ay = Y[5] - Y[3];
    double bx = X[7] - X[1], by = Y[7] - Y[1];
    double a = ax * ax + ay * ay, b = bx * bx + by * by;
    return std::sqrt((b < a) ? b : a);
}

// Allocate device buffers and copy the validated host mesh.
// Tables of the tiled step for the Rectangle problems of ctx->rects: per
// column (i) and per row (j) the centres, face centres, LSM weights and
// areas, taken from the first cell of each column / row; then every cell is
// checked against them (slot codes, adjacency, centres, face geometry,
// weights), so the tables give exactly the values of the slot arrays.  Any
// difference leaves the tiled step off for this mesh.
bool build_tiles(zfv_ctx* ctx, const std::vector<int>& code, const std::vector<double>& w) {
    auto& T = ctx->tiles;
    T.ok = false;
    T.n_probs = 0;
    if (ctx->rects.empty()) return false;
    const HostMesh& m = ctx->hm;
    const size_t n = (size_t)m.n;
    std::vector<TileProb> probs;
    std::vector<TileCol> cols;
    std::vector<TileRow> rows;
    int tile0[kTileVariants] = {0, 0, 0};
    for (const auto& R : ctx->rects) {
        const int nx = R.nx, ny = R.ny, off = R.off;
        if (nx < 1 || ny < 1 || (size_t)off + (size_t)nx * ny > n) return false;
        TileProb P{};
        P.off = off;
        P.nx = nx;
        P.ny = ny;
        P.tiles_j = (ny + kTileJ - 1) / kTileJ;
        for (int v = 0; v < kTileVariants; ++v) {
            P.tile0[v] = tile0[v];
            tile0[v] += ((nx + kTileRows[v] - 1) / kTileRows[v]) * P.tiles_j;
        }
        P.col = (int)cols.size();
        P.row = (int)rows.size();
        const int c00 = off, cLast = off + (nx - 1) * ny + (ny - 1);
        P.t[0] = code[c00] & 7;
        P.t[1] = (code[cLast] >> 4) & 7;
        P.t[2] = (code[c00] >> 8) & 7;
        P.t[3] = (code[cLast] >> 12) & 7;
        for (int s = 0; s < 4; ++s)
            if (P.t[s] != SC_PERIODIC && P.t[s] != SC_WALL && P.t[s] != SC_COPY) return false;
        P.per[0] = P.t[0] == SC_PERIODIC && P.t[1] == SC_PERIODIC;
        P.per[1] = P.t[2] == SC_PERIODIC && P.t[3] == SC_PERIODIC;
        auto W = [&](int s, size_t c) { return w[(size_t)s * n + c]; };
        for (int i = 0; i < nx; ++i) {
            const size_t c = (size_t)off + (size_t)i * ny, f = 8 * c;
            cols.push_back(TileCol{m.cx[c], m.fcx[f + 0], m.fcx[f + 1], m.fcx[f + 2], m.fcx[f + 3], W(0, c), W(1, c),
                                   m.area[f + 2], m.area[f + 3]});
        }
        for (int j = 0; j < ny; ++j) {
            const size_t c = (size_t)off + j, f = 8 * c;
            rows.push_back(TileRow{m.cy[c], m.fcy[f + 0], m.fcy[f + 1], m.fcy[f + 2], m.fcy[f + 3], W(2, c), W(3, c),
                                   m.area[f + 0], m.area[f + 1]});
        }
        const TileCol* C = &cols[P.col];
        const TileRow* Rw = &rows[P.row];
        for (int i = 0; i < nx; ++i)
            for (int j = 0; j < ny; ++j) {
                const size_t c = (size_t)off + (size_t)i * ny + j, f = 8 * c;
                const int ts[4] = {i > 0 ? (int)SC_INNER : P.t[0], i < nx - 1 ? (int)SC_INNER : P.t[1],
                                   j > 0 ? (int)SC_INNER : P.t[2], j < ny - 1 ? (int)SC_INNER : P.t[3]};
                const int want_code = ts[0] | (ts[1] << 4) | (ts[2] << 8) | (ts[3] << 12);
                if (code[c] != want_code) return false;
                const int self = (int)c;
                const bool nb_l = ts[0] == SC_INNER || (ts[0] == SC_PERIODIC && P.per[0]);
                const bool nb_r = ts[1] == SC_INNER || (ts[1] == SC_PERIODIC && P.per[0]);
                const bool nb_b = ts[2] == SC_INNER || (ts[2] == SC_PERIODIC && P.per[1]);
                const bool nb_t = ts[3] == SC_INNER || (ts[3] == SC_PERIODIC && P.per[1]);
                const int adj[4] = {nb_l ? off + ((i - 1 + nx) % nx) * ny + j : self,
                                    nb_r ? off + ((i + 1) % nx) * ny + j : self,
                                    nb_b ? off + i * ny + (j - 1 + ny) % ny : self,
                                    nb_t ? off + i * ny + (j + 1) % ny : self};
                for (int sl = 0; sl < 4; ++sl)
                    if (m.adj[f + sl] != adj[sl]) return false;
                if (m.cx[c] != C[i].cx || m.cy[c] != Rw[j].cy) return false;
                const double fx[4] = {C[i].fxL, C[i].fxR, C[i].fxB, C[i].fxT};
                const double fy[4] = {Rw[j].fyL, Rw[j].fyR, Rw[j].fyB, Rw[j].fyT};
                const double ar[4] = {Rw[j].aL, Rw[j].aR, C[i].aB, C[i].aT};
                const double wt[4] = {C[i].wL, C[i].wR, Rw[j].wB, Rw[j].wT};
                for (int sl = 0; sl < 4; ++sl) {
                    if (m.fcx[f + sl] != fx[sl] || m.fcy[f + sl] != fy[sl] || m.area[f + sl] != ar[sl]) return false;
                    if (W(sl, c) != wt[sl]) return false;
                }
            }
        probs.push_back(P);
    }
    hipStream_t s = ctx->stream;
    if (T.prob.alloc(probs.size()) != hipSuccess || T.col.alloc(cols.size()) != hipSuccess ||
        T.row.alloc(rows.size()) != hipSuccess)
        return false;
    if (hipMemcpyAsync(T.prob.p, probs.data(), probs.size() * sizeof(TileProb), hipMemcpyHostToDevice, s) != hipSuccess ||
        hipMemcpyAsync(T.col.p, cols.data(), cols.size() * sizeof(TileCol), hipMemcpyHostToDevice, s) != hipSuccess ||
        hipMemcpyAsync(T.row.p, rows.data(), rows.size() * sizeof(TileRow), hipMemcpyHostToDevice, s) != hipSuccess)
        return false;
    T.n_probs = (int)probs.size();
    for (int v = 0; v < kTileVariants; ++v) T.n_tiles[v] = tile0[v];
    T.ok = true;
    return true;
}

int upload_mesh(zfv_ctx* ctx) {
    ctx->mesh_epoch += 1;
    const HostMesh& m = ctx->hm;
    const int n = m.n;
    const size_t ns = 8 * (size_t)n;
    std::vector<int> code(n, 0), adj(ns, 0);
    std::vector<double> area(ns, 0.0), fcx(ns, 0.0), fcy(ns, 0.0), w(ns, 0.0), h(n, 0.0);
    std::vector<double> nxs(ns, 0.0), nys(ns, 0.0);

    int64_t n_int = 0, n_bnd = 0;
    for (int c = 0; c < n; ++c) {
        int cw = 0;
        for (int s = 0; s < 8; ++s) {
            size_t i = 8 * (size_t)c + s;
            int t = code_of_boundary(m.bnd[i]);
            if (t < 0) return fail(ZFV_ERR_ARG, "unknown boundary id " + std::to_string(m.bnd[i]));
            size_t j = (size_t)s * n + c;   // slot-major
            if (t == SC_UNDEF) {
                adj[j] = c;
                continue;
            }
            if (t >= SC_WALL) ++n_bnd;
            else ++n_int;
            const int side = s & 3;
            const double nx = m.nx[i], ny = m.ny[i];
            if (!std::isfinite(nx) || !std::isfinite(ny) || !(nx * nx + ny * ny > 0.0))
                return fail(ZFV_ERR_ARG, "cell " + std::to_string(c) + " slot " + std::to_string(s) +
                                             ": degenerate normal");
            const bool gen = !(nx == kNX[side] && ny == kNY[side]);
            cw |= (t | (gen ? kGeneralNormal : 0)) << (4 * s);
            int nb = m.adj[i];
            if (nb < 0 || nb >= n) return fail(ZFV_ERR_ARG, "adjacent index out of range");
            if ((t == SC_WALL || t == SC_COPY) && nb != c)
                return fail(ZFV_ERR_ARG, "boundary slot must reference its own cell");
            adj[j] = nb;
            area[j] = m.area[i];
            fcx[j] = m.fcx[i];
            fcy[j] = m.fcy[i];
            nxs[j] = nx;
            nys[j] = ny;
            // LSM weight w = area / |drn|^3, geometry-only (gradient.h:251-256):
            // dr = (INNER ? neib_c : symm_point(cell_c)) - cell_c, drn = (dr.n) n.
            const double cx = m.cx[c], cy = m.cy[c];
            double px, py;
            if (t == SC_INNER) {
                px = m.cx[nb];
                py = m.cy[nb];
            } else {
                double d2 = 2.0 * ((m.fcx[i] - cx) * nx + (m.fcy[i] - cy) * ny);
                px = cx + d2 * nx;
                py = cy + d2 * ny;
            }
            const double drx = px - cx, dry = py - cy;
            const double d = drx * nx + dry * ny;
            const double dnx = d * nx, dny = d * ny;
            const double len = std::sqrt(dnx * dnx + dny * dny);
            w[j] = (1.0 * m.area[i]) / std::pow(len, 3);
        }
        code[c] = cw;
        h[c] = incircle_2d(&m.vx[9 * (size_t)c], &m.vy[9 * (size_t)c]);
    }
    ctx->cap = n;
    ctx->n_faces = n_int / 2 + n_bnd;
    std::vector<int> bnd_sm(ns), iota(n);
    std::vector<double> vx_vm(9 * (size_t)n), vy_vm(9 * (size_t)n);
    for (int c = 0; c < n; ++c) {
        iota[c] = c;
        for (int s8 = 0; s8 < 8; ++s8) bnd_sm[(size_t)s8 * n + c] = m.bnd[8 * (size_t)c + s8];
        for (int k = 0; k < 9; ++k) {
            vx_vm[(size_t)k * n + c] = m.vx[9 * (size_t)c + k];
            vy_vm[(size_t)k * n + c] = m.vy[9 * (size_t)c + k];
        }
    }

    hipStream_t s = ctx->stream;
    auto put = [&](auto& buf, const auto& host) -> hipError_t {
        hipError_t e = buf.alloc(host.size());
        if (e != hipSuccess) return e;
        return hipMemcpyAsync(buf.p, host.data(), host.size() * sizeof(host[0]), hipMemcpyHostToDevice, s);
    };
    ZFV_HIP(put(ctx->cx, m.cx));
    ZFV_HIP(put(ctx->cy, m.cy));
    ZFV_HIP(put(ctx->vol, m.vol));
    ZFV_HIP(put(ctx->h, h));
    ZFV_HIP(put(ctx->code, code));
    ZFV_HIP(put(ctx->adj, adj));
    ZFV_HIP(put(ctx->area, area));
    ZFV_HIP(put(ctx->fcx, fcx));
    ZFV_HIP(put(ctx->fcy, fcy));
    ZFV_HIP(put(ctx->w, w));
    ZFV_HIP(put(ctx->nx, nxs));
    ZFV_HIP(put(ctx->ny, nys));
    ZFV_HIP(put(ctx->bnd, bnd_sm));
    ZFV_HIP(put(ctx->vx, vx_vm));
    ZFV_HIP(put(ctx->vy, vy_vm));
    ZFV_HIP(put(ctx->level, m.level));
    ZFV_HIP(put(ctx->flag, m.flag));
    ZFV_HIP(put(ctx->flag0, m.flag));
    ZFV_HIP(put(ctx->b_idx, m.b_idx));
    ZFV_HIP(put(ctx->z_idx, m.z_idx));
    ZFV_HIP(put(ctx->next, iota));
    ZFV_HIP(put(ctx->idx, iota));
    for (int l = 0; l < 3; ++l)
        for (int k = 0; k < 5; ++k) ZFV_HIP(ctx->st[l][k].alloc(n));
    ZFV_HIP(ctx->gdx.alloc(4 * (size_t)n));
    ZFV_HIP(ctx->gdy.alloc(4 * (size_t)n));
    // a batch's padding cells are never written by the gradient kernels:
    // zero, so a gradient download is deterministic there
    ZFV_HIP(hipMemsetAsync(ctx->gdx.p, 0, 4 * (size_t)n * sizeof(double), s));
    ZFV_HIP(hipMemsetAsync(ctx->gdy.p, 0, 4 * (size_t)n * sizeof(double), s));
    for (int k = 0; k < 2; ++k) {
        ZFV_HIP(ctx->wait[k].alloc(n));
        ZFV_HIP(hipMemsetAsync(ctx->wait[k].p, 0, n * sizeof(int), s));
    }
    ZFV_HIP(ctx->dtmin.alloc(3 * (size_t)ctx->nseg));
    ZFV_HIP(ctx->dt_out.alloc(ctx->nseg));
    ZFV_HIP(ctx->err.alloc(1));
    ZFV_HIP(hipMemsetAsync(ctx->err.p, 0, sizeof(int), s));
    ZFV_HIP(ctx->partial.alloc(4 * (size_t)((n + kBlock - 1) / kBlock)));
    ctx->axial_mesh = !m.vol_alt.empty();
    if (ctx->axial_mesh) {   // the alternative geometry, slot-major, and volume / volume_alt per cell
        std::vector<double> aa(ns, 0.0), axc(n, 0.0);
        for (int c = 0; c < n; ++c) {
            axc[c] = m.vol[c] / m.vol_alt[c];
            for (int s8 = 0; s8 < 8; ++s8) aa[(size_t)s8 * n + c] = m.area_alt[8 * (size_t)c + s8];
        }
        ZFV_HIP(put(ctx->vol_alt, m.vol_alt));
        ZFV_HIP(put(ctx->area_alt, aa));
        ZFV_HIP(put(ctx->axc, axc));
    }
    build_tiles(ctx, code, w);
    ZFV_HIP(hipStreamSynchronize(s));
    ctx->graphs_clear();
    ctx->grad_pending = false;
    if (!ctx->layout_set[1]) {   // default: the tiled step where it applies
        ctx->layout[1] = ctx->tiles.ok ? kLayoutTile : kLayoutCell;
        ctx->block[1] = ctx->tiles.ok ? (kTileRows[0] + 2) * 64 : kBlock;
    }
    ctx->have_mesh = true;
    ctx->host_stale = false;
    ctx->have_state = false;
    ctx->dt_valid = false;
    ctx->cur = 0;
    ctx->wcur = 0;
    // curve keys need 0 <= level <= 15, 0 <= z_idx < 4^level, b_idx >= 0
    int lmax = 0, bmax = 0;
    bool keys_ok = true;
    for (int c = 0; c < n; ++c) {
        const int L = m.level[c], z = m.z_idx[c], b = m.b_idx[c];
        keys_ok = keys_ok && L >= 0 && L <= 15 && z >= 0 && (long long)z < (1LL << (2 * L)) && b >= 0;
        lmax = std::max(lmax, L);
        bmax = std::max(bmax, b);
    }
    ctx->co.keys_ok = keys_ok && ctx->nseg == 1;
    ctx->co.refined = ctx->co.keys_ok && lmax > 0;
    ctx->co.n_base = (long long)bmax + 1;
    ctx->co.lmax = std::min(lmax, 15);
    int bits = 0;
    while (bits < 31 && (1LL << bits) <= bmax) ++bits;
    ctx->co.base_bits = bits;
    ctx->co.state = false;
    ctx->co.grad = false;
    ctx->co.epoch = -1;
    ctx->co.perm_epoch = -1;
    return ZFV_OK;
}

// Error word written by the kernels: bit 0 numeric (NaN flux / Riemann
// solve), bit 1 inconsistent mesh after an adaptation; bits 2 and 3 are
// notices for the adjoint (a cell in the robust isolation mode, a cell
// isolated by the corrector), not errors of the step.
int read_err(zfv_ctx* ctx) {
    int e = 0;
    ZFV_HIP(hipMemcpyAsync(&e, ctx->err.p, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
    ZFV_HIP(hipStreamSynchronize(ctx->stream));
    if (e & 2) {
        ctx->have_state = false;
        return fail(ZFV_ERR_STATE, "AMR apply produced an inconsistent mesh; state invalidated");
    }
    if (e & 1) return fail(ZFV_ERR_NUMERIC, "numeric error: a flux or Riemann solve produced NaN / did not converge");
    return ZFV_OK;
}

// Growth of the [rows][cap] device arrays to a larger capacity, keeping the
// first n_valid entries of every row.  Transactional: every new array is
// allocated before any is copied or swapped in, so a failed allocation
// leaves the context as it was (old capacity, old arrays); the copies are
// enqueued together and waited for once.
struct GrowSet {
    struct Item {
        void** p;
        size_t* n;
        size_t elem;
        int rows;
        void* np;
    };
    std::vector<Item> items;
    template <class T>
    void add(DBuf<T>& b, int rows) {
        items.push_back(Item{reinterpret_cast<void**>(&b.p), &b.n, sizeof(T), rows, nullptr});
    }
    void drop_new() {
        for (auto& it : items)
            if (it.np) (void)hipFree(it.np);
        for (auto& it : items) it.np = nullptr;
    }
    hipError_t apply(int old_cap, int new_cap, int n_valid, hipStream_t s) {
        for (auto& it : items) {
            hipError_t e = hipMalloc(&it.np, (size_t)it.rows * new_cap * it.elem);
            if (e != hipSuccess) {
                it.np = nullptr;
                drop_new();
                return e;
            }
        }
        if (n_valid > 0)
            for (auto& it : items) {
                if (!*it.p) continue;
                hipError_t e = hipMemcpy2DAsync(it.np, (size_t)new_cap * it.elem, *it.p, (size_t)old_cap * it.elem,
                                                (size_t)n_valid * it.elem, it.rows, hipMemcpyDeviceToDevice, s);
                if (e != hipSuccess) {
                    (void)hipStreamSynchronize(s);
                    drop_new();
                    return e;
                }
            }
        hipError_t e = hipStreamSynchronize(s);
        if (e != hipSuccess) {
            drop_new();
            return e;
        }
        for (auto& it : items) {
            if (*it.p) (void)hipFree(*it.p);
            *it.p = it.np;
            *it.n = (size_t)it.rows * new_cap;
            it.np = nullptr;
        }
        return hipSuccess;
    }
};

// Grow every capacity-sized device array to hold new_cap cells.  On failure
// the context keeps its previous capacity and arrays.
int grow_capacity(zfv_ctx* ctx, int new_cap) {
    if (new_cap <= ctx->cap) return ZFV_OK;
    const int oc = ctx->cap, n = ctx->hm.n;
    GrowSet g;
    for (auto* b : {&ctx->cx, &ctx->cy, &ctx->vol, &ctx->h}) g.add(*b, 1);
    for (auto* b : {&ctx->code, &ctx->level, &ctx->flag, &ctx->flag0, &ctx->b_idx, &ctx->z_idx, &ctx->next,
                    &ctx->idx, &ctx->wait[0], &ctx->wait[1]})
        g.add(*b, 1);
    for (auto* b : {&ctx->area, &ctx->fcx, &ctx->fcy, &ctx->w, &ctx->nx, &ctx->ny}) g.add(*b, 8);
    for (auto* b : {&ctx->adj, &ctx->bnd}) g.add(*b, 8);
    for (auto* b : {&ctx->vx, &ctx->vy}) g.add(*b, 9);
    for (auto* b : {&ctx->gdx, &ctx->gdy}) g.add(*b, 4);
    for (int l = 0; l < 3; ++l)
        for (int k = 0; k < 5; ++k) g.add(ctx->st[l][k], 1);
    auto& M = ctx->mm;   // the multi-material rows keep their [rows][cap] layout
    if (M.configured) {
        const int nm = M.mix.n;
        for (auto* b : {&M.st[0], &M.st[1], &M.half})
            if (b->p) g.add(*b, 7 + 2 * nm);
        for (auto* b : {&M.gx, &M.gy})
            if (b->p) g.add(*b, 6 + 2 * nm);
    }
    DBuf<double> partial;
    ZFV_HIP(partial.alloc(4 * (size_t)((new_cap + kBlock - 1) / kBlock)));
    ZFV_HIP(g.apply(oc, new_cap, n, ctx->stream));
    ctx->graphs_clear();   // buffers moved
    ctx->partial = std::move(partial);
    ctx->cap = new_cap;
    return ZFV_OK;
}

// Copy the device mesh back into the host AmrCells-layout copy.
int sync_host_mesh(zfv_ctx* ctx) {
    if (!ctx->host_stale) return ZFV_OK;
    HostMesh& m = ctx->hm;
    const int n = m.n, cap = ctx->cap;
    hipStream_t s = ctx->stream;
    auto get1 = [&](auto& host, const auto& buf) -> hipError_t {
        host.resize(n);
        return hipMemcpyAsync(host.data(), buf.p, (size_t)n * sizeof(host[0]), hipMemcpyDeviceToHost, s);
    };
    auto getr = [&](auto& host, const auto& buf, int rows) -> hipError_t {
        host.resize((size_t)rows * n);
        return hipMemcpy2DAsync(host.data(), (size_t)n * sizeof(host[0]), buf.p, (size_t)cap * sizeof(host[0]),
                                (size_t)n * sizeof(host[0]), rows, hipMemcpyDeviceToHost, s);
    };
    ZFV_HIP(get1(m.cx, ctx->cx)); ZFV_HIP(get1(m.cy, ctx->cy)); ZFV_HIP(get1(m.vol, ctx->vol));
    ZFV_HIP(get1(m.level, ctx->level)); ZFV_HIP(get1(m.flag, ctx->flag));
    ZFV_HIP(get1(m.b_idx, ctx->b_idx)); ZFV_HIP(get1(m.z_idx, ctx->z_idx));
    std::vector<int32_t> bnd, adj;
    std::vector<double> nx, ny, fcx, fcy, area, vx, vy;
    ZFV_HIP(getr(bnd, ctx->bnd, 8)); ZFV_HIP(getr(adj, ctx->adj, 8));
    ZFV_HIP(getr(nx, ctx->nx, 8)); ZFV_HIP(getr(ny, ctx->ny, 8));
    ZFV_HIP(getr(fcx, ctx->fcx, 8)); ZFV_HIP(getr(fcy, ctx->fcy, 8)); ZFV_HIP(getr(area, ctx->area, 8));
    ZFV_HIP(getr(vx, ctx->vx, 9)); ZFV_HIP(getr(vy, ctx->vy, 9));
    ZFV_HIP(hipStreamSynchronize(s));
    const size_t ns = 8 * (size_t)n;
    m.bnd.resize(ns); m.adj.resize(ns); m.nx.resize(ns); m.ny.resize(ns);
    m.fcx.resize(ns); m.fcy.resize(ns); m.area.resize(ns);
    m.vx.resize(9 * (size_t)n); m.vy.resize(9 * (size_t)n);
    for (int c = 0; c < n; ++c) {
        for (int k = 0; k < 8; ++k) {
            size_t a = 8 * (size_t)c + k, b = (size_t)k * n + c;
            m.bnd[a] = bnd[b]; m.adj[a] = adj[b]; m.nx[a] = nx[b]; m.ny[a] = ny[b];
            m.fcx[a] = fcx[b]; m.fcy[a] = fcy[b]; m.area[a] = area[b];
            if (m.bnd[a] == ZFV_UNDEFINED) {   // undefined slots carry no geometry
                m.adj[a] = -1; m.nx[a] = 0.0; m.ny[a] = 0.0; m.fcx[a] = 0.0; m.fcy[a] = 0.0; m.area[a] = 0.0;
            }
        }
        for (int k = 0; k < 9; ++k) {
            m.vx[9 * (size_t)c + k] = vx[(size_t)k * n + c];
            m.vy[9 * (size_t)c + k] = vy[(size_t)k * n + c];
        }
    }
    ctx->host_stale = false;
    return ZFV_OK;
}

void collect_timers(zfv_ctx* ctx) {
    for (auto& t : ctx->timers) {
        float ms = 0.0f;
        (void)hipEventSynchronize(t.t1);
        (void)hipEventElapsedTime(&ms, t.t0, t.t1);
        auto& slot = ctx->prof[t.name];
        slot.first += ms;
        slot.second += 1;
        ctx->event_pool.push_back(t.t0);
        ctx->event_pool.push_back(t.t1);
    }
    ctx->timers.clear();
}

hipEvent_t pooled_event(zfv_ctx* ctx) {
    if (!ctx->event_pool.empty()) {
        hipEvent_t e = ctx->event_pool.back();
        ctx->event_pool.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    (void)hipEventCreate(&e);
    return e;
}

template <class F>
hipError_t timed_on(zfv_ctx* ctx, hipStream_t s, const char* name, F&& launch) {
    if (!ctx->profile) return launch();
    KernelTimer t;
    t.name = name;
    t.t0 = pooled_event(ctx);
    t.t1 = pooled_event(ctx);
    (void)hipEventRecord(t.t0, s);
    hipError_t e = launch();
    (void)hipEventRecord(t.t1, s);
    ctx->timers.push_back(t);
    return e;
}

template <class F>
hipError_t timed(zfv_ctx* ctx, const char* name, F&& launch) {
    return timed_on(ctx, ctx->stream, name, launch);
}

int ensure_dt(zfv_ctx* ctx) {
    if (ctx->dt_valid) return ZFV_OK;
    int rd = (int)(ctx->step_no % 3);
    ZFV_HIP(launch_fill_u64(ctx->dtmin.p, 3 * ctx->nseg, 0x7ff0000000000000ULL, ctx->stream));
    DevMesh M = ctx->smesh();
    DevState z = ctx->layer(ctx->cur);
    Eos eos = ctx->eos();
    ZFV_HIP(timed(ctx, "dt", [&] { return launch_dt(M, z, eos, ctx->dtmin.p + rd, ctx->stream); }));
    ctx->dt_valid = true;
    return ZFV_OK;
}

StepCtl make_ctl(zfv_ctx* ctx) {
    StepCtl c;
    c.dtmin = ctx->dtmin.p;
    c.rd = (int)(ctx->step_no % 3);
    c.wr = (int)((ctx->step_no + 1) % 3);
    c.clr = (int)((ctx->step_no + 2) % 3);
    c.cfl = ctx->prm.cfl;
    c.max_dt = ctx->dmax_dt.p;
    c.dt_out = ctx->dt_out.p;
    c.err = ctx->err.p;
    return c;
}

int need_ready(zfv_ctx* ctx, bool state = true, bool config = false) {
    int r = check_ctx(ctx);
    if (r) return r;
    if (!ctx->have_mesh) return fail(ZFV_ERR_STATE, "no mesh uploaded");
    if (state && !ctx->have_state) return fail(ZFV_ERR_STATE, "no state uploaded");
    if (config && !ctx->configured) return fail(ZFV_ERR_STATE, "solver not configured");
    if (config && ctx->prm.axial && !ctx->axial_mesh)
        return fail(ZFV_ERR_STATE, "axisymmetric solver on a mesh without the axisymmetric geometry");
    return ZFV_OK;
}

int put_max_dt(zfv_ctx* ctx);

// Sorted copy of the mesh for the step kernels (state in the mesh's order);
// with gather, the same launch also gathers state layer cur and wait[wcur]
// into curve order in the other layer and wait buffer, which become the
// current ones (gather_current).
int build_curve(zfv_ctx* ctx, bool gather = false) {
    auto& co = ctx->co;
    const int n = ctx->hm.n, cap = ctx->cap;
    hipStream_t s = ctx->stream;
    if (co.cap != cap) {
        for (auto* b : {&co.cx, &co.cy, &co.vol, &co.h}) ZFV_HIP(b->alloc(cap));
        for (auto* b : {&co.area, &co.fcx, &co.fcy, &co.w, &co.nx, &co.ny}) ZFV_HIP(b->alloc(8 * (size_t)cap));
        ZFV_HIP(co.code.alloc(cap));
        ZFV_HIP(co.adj.alloc(8 * (size_t)cap));
        for (auto* b : {&co.perm, &co.iperm, &co.vals}) ZFV_HIP(b->alloc(cap));
        for (auto& k : co.keys) ZFV_HIP(k.alloc(cap));
        co.cap = cap;
        co.perm_epoch = -1;
    }
    // The order: carried across the last adaptation (zfv_refine), or dense
    // slots (one per finest-level Morton cell of every base cell) compacted
    // in order when they are at most 64 per cell, else a radix sort of the
    // same keys.
    const long long n_slots = co.n_base << (2 * co.lmax);
    const long long n_chunks = (n_slots + 63) / 64;
    size_t bytes = 0;
    if (co.perm_epoch == ctx->mesh_epoch) {
        // perm / iperm hold this mesh's order already
    } else if (n_slots <= 64LL * std::max(n, 1) && n_chunks * 64 < (1LL << 31)) {
        if (co.slots.n < (size_t)(n_chunks * 64)) {
            ZFV_HIP(co.slots.alloc((size_t)(n_chunks * 64)));
            co.slots_clean = false;
        }
        if (co.counts.n < (size_t)n_chunks) {
            ZFV_HIP(co.counts.alloc((size_t)n_chunks));
            co.slots_clean = false;
        }
        if (co.offsets.n < (size_t)n_chunks) ZFV_HIP(co.offsets.alloc((size_t)n_chunks));
        const size_t tiles = (size_t)std::max(curve_scan_tiles((int)n_chunks), 1);
        if (co.tiles.n < tiles) ZFV_HIP(co.tiles.alloc(tiles));
        if (!co.slots_clean) {   // a completed build leaves every slot at -1 and every count at 0
            ZFV_HIP(hipMemsetAsync(co.slots.p, 0xff, co.slots.n * sizeof(int), s));
            ZFV_HIP(hipMemsetAsync(co.counts.p, 0, co.counts.n * sizeof(int), s));
        }
        co.slots_clean = false;
        ZFV_HIP(curve_slots(ctx->b_idx.p, ctx->z_idx.p, ctx->level.p, n, co.lmax, co.slots.p, co.counts.p, n_slots,
                            s));
        ZFV_HIP(curve_compact(co.slots.p, (int)n_chunks, co.counts.p, co.offsets.p, co.tiles.p, co.perm.p,
                              co.iperm.p, s));
        co.slots_clean = true;
    } else {
        ZFV_HIP(curve_keys(ctx->b_idx.p, ctx->z_idx.p, ctx->level.p, n, co.lmax, co.keys[0].p, co.vals.p, s));
        const int end_bit = std::min(64, 2 * co.lmax + co.base_bits);
        ZFV_HIP(curve_sort(co.keys[0].p, co.keys[1].p, co.vals.p, co.perm.p, n, end_bit, nullptr, bytes, s));
        if (co.tmp.n < bytes) ZFV_HIP(co.tmp.alloc(bytes));
        ZFV_HIP(curve_sort(co.keys[0].p, co.keys[1].p, co.vals.p, co.perm.p, n, end_bit, co.tmp.p, bytes, s));
        ZFV_HIP(curve_inverse(co.perm.p, co.iperm.p, n, s));
    }
    const CurveMesh C{co.cap, co.cx.p, co.cy.p, co.vol.p, co.h.p, co.code.p, co.adj.p,
                      co.area.p, co.fcx.p, co.fcy.p, co.w.p, co.nx.p, co.ny.p};
    const int nxt = 1 - ctx->cur;
    ZFV_HIP(curve_mesh(ctx->dmesh(), C, co.perm.p, co.iperm.p, gather ? ctx->layer(ctx->cur) : DevState{},
                       ctx->layer(nxt), ctx->wait[ctx->wcur].p, ctx->wait[1 - ctx->wcur].p, s));
    if (gather) {
        ctx->cur = nxt;
        ctx->wcur = 1 - ctx->wcur;
    }
    co.epoch = ctx->mesh_epoch;
    ctx->graphs_clear();   // captured steps read the previous copy
    return ZFV_OK;
}

// State layer cur and wait[wcur] gathered through idx into the other layer
// and wait buffer, which become the current ones.
int gather_current(zfv_ctx* ctx, const int* idx) {
    const int nxt = 1 - ctx->cur;
    ZFV_HIP(gather_state(ctx->layer(ctx->cur), ctx->layer(nxt), ctx->wait[ctx->wcur].p, ctx->wait[1 - ctx->wcur].p,
                         idx, ctx->hm.n, ctx->stream));
    ctx->cur = nxt;
    ctx->wcur = 1 - ctx->wcur;
    return ZFV_OK;
}

// Before steps: the state in curve order when the mesh has refined cells.
int to_curve_order(zfv_ctx* ctx) {
    auto& co = ctx->co;
    if (co.state || co.mode == 0 || !co.refined || ctx->nseg > 1) return ZFV_OK;
    if (co.epoch != ctx->mesh_epoch || co.cap != ctx->cap) {
        int r = build_curve(ctx, true);   // gathers the state in the same launch
        if (r) return r;
    } else {
        int r = gather_current(ctx, co.perm.p);
        if (r) return r;
    }
    co.state = true;
    return ZFV_OK;
}

// The state (and wait counters) back in the mesh's order.
int to_mesh_order(zfv_ctx* ctx) {
    auto& co = ctx->co;
    if (!co.state) return ZFV_OK;
    int r = gather_current(ctx, co.iperm.p);
    if (r) return r;
    co.state = false;
    return ZFV_OK;
}

// The last gradients back in the mesh's order.
int grad_to_mesh_order(zfv_ctx* ctx) {
    auto& co = ctx->co;
    if (!co.grad) return ZFV_OK;
    const int n = ctx->hm.n, cap = ctx->cap;
    if (co.epoch != ctx->mesh_epoch || co.cap != cap) return fail(ZFV_ERR_STATE, "gradients of a previous mesh");
    if (co.gtmp.n < 4 * (size_t)cap) ZFV_HIP(co.gtmp.alloc(4 * (size_t)cap));
    for (double* g : {ctx->gdx.p, ctx->gdy.p}) {
        ZFV_HIP(gather_rows(g, co.gtmp.p, co.iperm.p, n, 4, cap, ctx->stream));
        ZFV_HIP(hipMemcpy2DAsync(g, (size_t)cap * 8, co.gtmp.p, (size_t)cap * 8, (size_t)n * 8, 4,
                                 hipMemcpyDeviceToDevice, ctx->stream));
    }
    co.grad = false;
    return ZFV_OK;
}

// Device address of the static scratch array of each device, taken once per
// device at zfv_create (the no-allocation entry points use only this).
std::mutex g_scratch_mu;
std::vector<double*> g_scratch;

int scratch_setup(int device) {
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if ((int)g_scratch.size() <= device) g_scratch.resize(device + 1, nullptr);
    if (g_scratch[device]) return ZFV_OK;
    double* p = nullptr;
    ZFV_HIP(scratch_address(&p));
    if (!p) return fail(ZFV_ERR_HIP, "hipGetSymbolAddress(scratch) returned a null address");
    g_scratch[device] = p;
    return ZFV_OK;
}

double* scratch_of(int device) {
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    return device < (int)g_scratch.size() ? g_scratch[device] : nullptr;
}

}  // namespace

namespace zfv {
int set_error(int code, const char* msg) { return fail(code, msg ? msg : ""); }
}  // namespace zfv

extern "C" {

const char* zfv_last_error(void) { return g_last_error.c_str(); }
const char* zfv_version(void) { return "zfv 0.1 gfx950"; }

int zfv_create(int device, zfv_ctx** out) {
    if (!out) return fail(ZFV_ERR_ARG, "null output pointer");
    int count = 0;
    hipError_t e = hipGetDeviceCount(&count);
    if (e != hipSuccess || count == 0) return fail(ZFV_ERR_HIP, "no HIP device available");
    if (device < 0 || device >= count) return fail(ZFV_ERR_ARG, "device index out of range");
    ZFV_HIP(hipSetDevice(device));
    int rs = scratch_setup(device);   // once per device, with the device current
    if (rs) return rs;
    zfv_ctx* ctx = new zfv_ctx();
    ctx->device = device;
    e = hipStreamCreateWithFlags(&ctx->own_stream, hipStreamNonBlocking);
    if (e != hipSuccess) {
        delete ctx;
        return fail(ZFV_ERR_HIP, std::string("hipStreamCreate: ") + hipGetErrorString(e));
    }
    ctx->stream = ctx->own_stream;
    ctx->prm.max_dt = std::numeric_limits<double>::max();
    if (ctx->dmax_dt.alloc(1) != hipSuccess || put_max_dt(ctx) != ZFV_OK) {
        (void)hipStreamDestroy(ctx->own_stream);
        delete ctx;
        return fail(ZFV_ERR_HIP, "context allocation failed");
    }
    *out = ctx;
    return ZFV_OK;
}

int zfv_destroy(zfv_ctx* ctx) {
    if (!ctx) return ZFV_OK;
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    collect_timers(ctx);
    ctx->graphs_clear();
    for (hipEvent_t e : ctx->event_pool) (void)hipEventDestroy(e);
    for (hipEvent_t e : {ctx->mm.ev_in, ctx->mm.ev_mark, ctx->mm.ev_crp})
        if (e) (void)hipEventDestroy(e);
    if (ctx->mm.side) {
        (void)hipStreamSynchronize(ctx->mm.side);
        (void)hipStreamDestroy(ctx->mm.side);
    }
    if (ctx->marks_total) (void)hipHostFree(ctx->marks_total);
    hipStream_t s = ctx->own_stream;
    delete ctx;
    if (s) (void)hipStreamDestroy(s);
    return ZFV_OK;
}

int zfv_set_stream(zfv_ctx* ctx, void* stream) {
    int r = check_ctx(ctx);
    if (r) return r;
    ZFV_HIP(hipStreamSynchronize(ctx->stream));
    ctx->stream = stream ? (hipStream_t)stream : ctx->own_stream;
    return ZFV_OK;
}

int zfv_synchronize(zfv_ctx* ctx) {
    int r = check_ctx(ctx);
    if (r) return r;
    ZFV_HIP(hipStreamSynchronize(ctx->stream));
    collect_timers(ctx);
    if (ctx->have_mesh) return read_err(ctx);
    return ZFV_OK;
}

static int mesh_upload_impl(zfv_ctx* ctx, const zfv_mesh_soa* M, const double* vol_alt, const double* area_alt);

int zfv_mesh_upload(zfv_ctx* ctx, const zfv_mesh_soa* M) { return mesh_upload_impl(ctx, M, nullptr, nullptr); }

int zfv_mesh_upload_axial(zfv_ctx* ctx, const zfv_mesh_soa* M, const double* vol_alt, const double* area_alt) {
    if (!vol_alt || !area_alt) return fail(ZFV_ERR_ARG, "null volume_alt or area_alt");
    return mesh_upload_impl(ctx, M, vol_alt, area_alt);
}

static int mesh_upload_impl(zfv_ctx* ctx, const zfv_mesh_soa* M, const double* vol_alt, const double* area_alt) {
    int r = check_ctx(ctx);
    if (r) return r;
    if (!M) return fail(ZFV_ERR_ARG, "null mesh");
    const int n = M->n_cells;
    if (n <= 0) return fail(ZFV_ERR_ARG, "mesh has no cells");
    if ((long long)n * 8 > (long long)std::numeric_limits<int>::max())
        return fail(ZFV_ERR_ARG, "mesh too large for 32-bit slot indexing");
    if (!M->cx || !M->cy || !M->volume || !M->vx || !M->vy || !M->boundary || !M->adj || !M->nx ||
        !M->ny || !M->fcx || !M->fcy || !M->area)
        return fail(ZFV_ERR_ARG, "mesh array missing");
    if (vol_alt)
        for (int c = 0; c < n; ++c)
            if (!(vol_alt[c] > 0.0) || !std::isfinite(vol_alt[c]))
                return fail(ZFV_ERR_ARG, "volume_alt must be positive and finite (cells off the axis)");
    HostMesh& m = ctx->hm;
    const size_t ns = 8 * (size_t)n, nv = 9 * (size_t)n;
    ctx->nseg = 1;
    m.n = n;
    m.vol_alt.clear();
    m.area_alt.clear();
    if (vol_alt) {
        m.vol_alt.assign(vol_alt, vol_alt + n);
        m.area_alt.assign(area_alt, area_alt + ns);
    }
    m.cx.assign(M->cx, M->cx + n);
    m.cy.assign(M->cy, M->cy + n);
    m.vol.assign(M->volume, M->volume + n);
    m.vx.assign(M->vx, M->vx + nv);
    m.vy.assign(M->vy, M->vy + nv);
    auto opt = [n](std::vector<int32_t>& dst, const int32_t* src, int fill, bool iota) {
        if (src) dst.assign(src, src + n);
        else {
            dst.assign(n, fill);
            if (iota)
                for (int i = 0; i < n; ++i) dst[i] = i;
        }
    };
    opt(m.level, M->level, 0, false);
    opt(m.flag, M->flag, 0, false);
    opt(m.b_idx, M->b_idx, 0, true);
    opt(m.z_idx, M->z_idx, 0, false);
    m.bnd.assign(M->boundary, M->boundary + ns);
    m.adj.assign(M->adj, M->adj + ns);
    m.nx.assign(M->nx, M->nx + ns);
    m.ny.assign(M->ny, M->ny + ns);
    m.fcx.assign(M->fcx, M->fcx + ns);
    m.fcy.assign(M->fcy, M->fcy + ns);
    m.area.assign(M->area, M->area + ns);
    // An uploaded mesh may be a Rectangle (a restored checkpoint of one, or
    // the reference's own): unrefined cells whose first column holds the
    // same centre x.  build_tiles then checks every cell against the
    // Rectangle's tables and takes the tiled step only if all match.
    ctx->rects.clear();
    bool flat = true;
    for (int c = 0; c < n && flat; ++c) flat = m.level[c] == 0;
    int ny = 1;
    while (flat && ny < n && m.cx[ny] == m.cx[0]) ++ny;
    if (flat && n % ny == 0) ctx->rects.assign(1, zfv_ctx::RectInfo{0, n / ny, ny});
    return upload_mesh(ctx);
}

// Rectangle AMR-2D base mesh (generator::Rectangle, Rectangle::initialize,
// geom/generator/rectangle.cpp:465-...; AmrCells(const Rectangle&),
// mesh/euler/amr_cells_gen.cpp:128-281): cell ic = i*ny + j, slots L,R,B,T
// active, boundary slots reference the cell itself, 9 SqQuad vertices.
// Written into m at cells [off, off + nx*ny) (arrays already sized).
static int check_rectangle(int nx, int ny, double xmin, double xmax, double ymin, double ymax,
                           const int32_t* bounds) {
    if (nx < 1 || ny < 1) return fail(ZFV_ERR_ARG, "Rectangle: nx, ny must be >= 1");
    if (!(xmax > xmin) || !(ymax > ymin)) return fail(ZFV_ERR_ARG, "Rectangle: empty domain");
    if (!bounds) return fail(ZFV_ERR_ARG, "null bounds");
    for (int k = 0; k < 4; ++k)
        if (bounds[k] < ZFV_INNER || bounds[k] > ZFV_FUNCTION || bounds[k] == ZFV_INNER)
            return fail(ZFV_ERR_ARG, "Rectangle: bad boundary id");
    if ((long long)nx * ny * 9 > (long long)std::numeric_limits<int>::max())
        return fail(ZFV_ERR_ARG, "Rectangle too large");
    return ZFV_OK;
}

static void resize_host_mesh(HostMesh& m, int n) {
    m.n = n;
    const size_t ns = 8 * (size_t)n;
    m.cx.assign(n, 0.0); m.cy.assign(n, 0.0); m.vol.assign(n, 0.0);
    m.vx.assign(9 * (size_t)n, 0.0); m.vy.assign(9 * (size_t)n, 0.0);
    m.level.assign(n, 0); m.flag.assign(n, 0); m.b_idx.resize(n); m.z_idx.assign(n, 0);
    for (int c = 0; c < n; ++c) m.b_idx[c] = c;
    m.bnd.assign(ns, ZFV_UNDEFINED); m.adj.assign(ns, -1);
    m.nx.assign(ns, 0.0); m.ny.assign(ns, 0.0); m.fcx.assign(ns, 0.0); m.fcy.assign(ns, 0.0);
    m.area.assign(ns, 0.0);
    m.vol_alt.clear();
    m.area_alt.clear();
}

// AmrCells of a Rectangle (rectangle.cpp:465-603); axial: the y-weighted
// volume and face areas and the y-shifted cell and side-face centres of
// Rectangle::set_axial(true) (rectangle.cpp:590-602).
static void build_rectangle(HostMesh& m, int off, int nx, int ny, double xmin, double xmax, double ymin,
                            double ymax, const int32_t* bounds, bool axial = false) {
    if (axial && m.vol_alt.empty()) {
        m.vol_alt.assign(m.n, 0.0);
        m.area_alt.assign(8 * (size_t)m.n, 0.0);
    }
    const bool xper = bounds[0] == ZFV_PERIODIC && bounds[1] == ZFV_PERIODIC;
    const bool yper = bounds[2] == ZFV_PERIODIC && bounds[3] == ZFV_PERIODIC;
    const double hx = (xmax - xmin) / nx, hy = (ymax - ymin) / ny;
    const double Lx = xmax - xmin, Ly = ymax - ymin;
    for (int i = 0; i < nx; ++i) {
        const double x0 = xmin + (Lx * i) / nx, x1 = xmin + (Lx * (i + 1)) / nx;
        for (int j = 0; j < ny; ++j) {
            const int lc = i * ny + j, c = off + lc;
            const double y0 = ymin + (Ly * j) / ny, y1 = ymin + (Ly * (j + 1)) / ny;
            // SqQuad(v00, v01, v10, v11): rows by y, columns by x (quad.cpp:429-436)
            double* X = &m.vx[9 * (size_t)c];
            double* Y = &m.vy[9 * (size_t)c];
            X[0] = x0; Y[0] = y0;
            X[2] = x1; Y[2] = y0;
            X[6] = x0; Y[6] = y1;
            X[8] = x1; Y[8] = y1;
            X[1] = 0.5 * (x0 + x1); Y[1] = 0.5 * (y0 + y0);
            X[3] = 0.5 * (x0 + x0); Y[3] = 0.5 * (y0 + y1);
            X[5] = 0.5 * (x1 + x1); Y[5] = 0.5 * (y0 + y1);
            X[7] = 0.5 * (x0 + x1); Y[7] = 0.5 * (y1 + y1);
            X[4] = 0.25 * (((x0 + x1) + x0) + x1);
            Y[4] = 0.25 * (((y0 + y0) + y1) + y1);
            m.cx[c] = X[4];
            m.cy[c] = Y[4];
            m.vol[c] = hx * hy;
            m.b_idx[c] = lc;
            const size_t f = 8 * (size_t)c;
            m.bnd[f + 0] = i > 0 ? ZFV_INNER : bounds[0];
            m.bnd[f + 1] = i < nx - 1 ? ZFV_INNER : bounds[1];
            m.bnd[f + 2] = j > 0 ? ZFV_INNER : bounds[2];
            m.bnd[f + 3] = j < ny - 1 ? ZFV_INNER : bounds[3];
            m.adj[f + 0] = off + ((i == 0 && !xper) ? lc : ((i - 1 + nx) % nx) * ny + j);
            m.adj[f + 1] = off + ((i == nx - 1 && !xper) ? lc : ((i + 1) % nx) * ny + j);
            m.adj[f + 2] = off + ((j == 0 && !yper) ? lc : i * ny + (j - 1 + ny) % ny);
            m.adj[f + 3] = off + ((j == ny - 1 && !yper) ? lc : i * ny + (j + 1) % ny);
            const int cen[4] = {3, 5, 1, 7};
            for (int s = 0; s < 4; ++s) {
                m.nx[f + s] = kNX[s];
                m.ny[f + s] = kNY[s];
                m.fcx[f + s] = X[cen[s]];
                m.fcy[f + s] = Y[cen[s]];
                m.area[f + s] = s < 2 ? hy : hx;
            }
            if (axial) {
                m.vol_alt[c] = hx * hy * Y[4];
                m.area_alt[f + 0] = hy * Y[3];
                m.area_alt[f + 1] = hy * Y[5];
                m.area_alt[f + 2] = hx * Y[1];
                m.area_alt[f + 3] = hx * Y[7];
                m.cy[c] += hy * hy / (12.0 * Y[4]);
                m.fcy[f + 0] += hy * hy / (12.0 * Y[3]);
                m.fcy[f + 1] += hy * hy / (12.0 * Y[5]);
            }
        }
    }
}

int zfv_mesh_rectangle(zfv_ctx* ctx, int nx, int ny, double xmin, double xmax,
                       double ymin, double ymax, const int32_t* bounds) {
    int r = check_ctx(ctx);
    if (r) return r;
    r = check_rectangle(nx, ny, xmin, xmax, ymin, ymax, bounds);
    if (r) return r;
    ctx->nseg = 1;
    resize_host_mesh(ctx->hm, nx * ny);
    build_rectangle(ctx->hm, 0, nx, ny, xmin, xmax, ymin, ymax, bounds);
    ctx->rects.assign(1, zfv_ctx::RectInfo{0, nx, ny});
    return upload_mesh(ctx);
}

int zfv_mesh_rectangle_axial(zfv_ctx* ctx, int nx, int ny, double xmin, double xmax, double ymin, double ymax,
                             const int32_t* bounds) {
    int r = check_ctx(ctx);
    if (r) return r;
    r = check_rectangle(nx, ny, xmin, xmax, ymin, ymax, bounds);
    if (r) return r;
    if (!(ymin >= 0.0)) return fail(ZFV_ERR_ARG, "axisymmetric rectangle: ymin must be >= 0 (y is the radius)");
    int32_t b[4] = {bounds[0], bounds[1], bounds[2], bounds[3]};
    if (ymin == 0.0) b[2] = ZFV_WALL;   // Rectangle::set_axial (rectangle.cpp:97-103)
    ctx->nseg = 1;
    resize_host_mesh(ctx->hm, nx * ny);
    build_rectangle(ctx->hm, 0, nx, ny, xmin, xmax, ymin, ymax, b, true);
    ctx->rects.assign(1, zfv_ctx::RectInfo{0, nx, ny});
    return upload_mesh(ctx);
}

int zfv_mesh_axial_download(zfv_ctx* ctx, double* vol_alt, double* area_alt) {
    int r = check_ctx(ctx);
    if (r) return r;
    if (!ctx->have_mesh) return fail(ZFV_ERR_STATE, "no mesh uploaded");
    if (!ctx->axial_mesh) return fail(ZFV_ERR_STATE, "the mesh has no axisymmetric geometry");
    const HostMesh& m = ctx->hm;
    if (vol_alt) std::memcpy(vol_alt, m.vol_alt.data(), (size_t)m.n * sizeof(double));
    if (area_alt) std::memcpy(area_alt, m.area_alt.data(), 8 * (size_t)m.n * sizeof(double));
    return ZFV_OK;
}

int zfv_batch_rectangles(zfv_ctx* ctx, int n_problems, const int32_t* nx, const int32_t* ny, const double* box,
                         const int32_t* bounds) {
    int r = check_ctx(ctx);
    if (r) return r;
    if (n_problems < 1 || !nx || !ny) return fail(ZFV_ERR_ARG, "batch: need n_problems >= 1 and nx, ny");
    const double unit[4] = {0.0, 1.0, 0.0, 1.0};
    const int32_t zoe[4] = {ZFV_ZOE, ZFV_ZOE, ZFV_ZOE, ZFV_ZOE};
    std::vector<int> begin(n_problems + 1, 0), count(n_problems);
    long long total = 0;
    for (int p = 0; p < n_problems; ++p) {
        const double* b = box ? box + 4 * p : unit;
        r = check_rectangle(nx[p], ny[p], b[0], b[1], b[2], b[3], bounds ? bounds + 4 * p : zoe);
        if (r) return r;
        begin[p] = (int)total;
        count[p] = nx[p] * ny[p];
        total += ((long long)count[p] + kSegAlign - 1) / kSegAlign * kSegAlign;
        if (total * 9 > (long long)std::numeric_limits<int>::max()) return fail(ZFV_ERR_ARG, "batch too large");
    }
    begin[n_problems] = (int)total;
    HostMesh& m = ctx->hm;
    resize_host_mesh(m, (int)total);   // padding cells: no faces, zero volume
    for (int p = 0; p < n_problems; ++p) {
        const double* b = box ? box + 4 * p : unit;
        build_rectangle(m, begin[p], nx[p], ny[p], b[0], b[1], b[2], b[3], bounds ? bounds + 4 * p : zoe);
    }
    ctx->nseg = n_problems;
    ctx->seg_begin_h = begin;
    ctx->seg_n_h = count;
    ctx->rects.clear();
    for (int p = 0; p < n_problems; ++p) ctx->rects.push_back(zfv_ctx::RectInfo{begin[p], nx[p], ny[p]});
    r = upload_mesh(ctx);
    if (r) return r;
    const int nchunk = (int)(total >> kChunkShift);
    std::vector<int> cp(nchunk), se(n_problems);
    for (int p = 0; p < n_problems; ++p) {
        se[p] = begin[p] + count[p];
        for (int k = begin[p] >> kChunkShift; k < (begin[p + 1] >> kChunkShift); ++k) cp[k] = p;
    }
    ZFV_HIP(ctx->chunk_pid.alloc(nchunk));
    ZFV_HIP(ctx->seg_begin.alloc(n_problems));
    ZFV_HIP(ctx->seg_end.alloc(n_problems));
    ZFV_HIP(hipMemcpyAsync(ctx->chunk_pid.p, cp.data(), nchunk * sizeof(int), hipMemcpyHostToDevice, ctx->stream));
    ZFV_HIP(hipMemcpyAsync(ctx->seg_begin.p, begin.data(), n_problems * sizeof(int), hipMemcpyHostToDevice,
                           ctx->stream));
    ZFV_HIP(hipMemcpyAsync(ctx->seg_end.p, se.data(), n_problems * sizeof(int), hipMemcpyHostToDevice, ctx->stream));
    ZFV_HIP(hipStreamSynchronize(ctx->stream));
    return ZFV_OK;
}

int zfv_batch_layout(zfv_ctx* ctx, int32_t* n_problems, int32_t* begin, int32_t* count) {
    int r = check_ctx(ctx);
    if (r) return r;
    if (!ctx->have_mesh) return fail(ZFV_ERR_STATE, "no mesh uploaded");
    if (n_problems) *n_problems = ctx->nseg;
    if (ctx->nseg == 1) {
        if (begin) { begin[0] = 0; begin[1] = ctx->hm.n; }
        if (count) count[0] = ctx->hm.n;
        return ZFV_OK;
    }
    if (begin) std::memcpy(begin, ctx->seg_begin_h.data(), (ctx->nseg + 1) * sizeof(int32_t));
    if (count) std::memcpy(count, ctx->seg_n_h.data(), ctx->nseg * sizeof(int32_t));
    return ZFV_OK;
}

int zfv_batch_dt(zfv_ctx* ctx, double* dt) {
    int r = need_ready(ctx, false);
    if (r) return r;
    if (!dt) return fail(ZFV_ERR_ARG, "null output");
    ZFV_HIP(hipMemcpyAsync(dt, ctx->dt_out.p, ctx->nseg * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    ZFV_HIP(hipStreamSynchronize(ctx->stream));
    return read_err(ctx);
}

int zfv_mesh_n_cells(zfv_ctx* ctx, int32_t* n) {
    if (!ctx || !n) return fail(ZFV_ERR_ARG, "null argument");
    *n = ctx->have_mesh ? ctx->hm.n : 0;
    return ZFV_OK;
}

int zfv_mesh_n_faces(zfv_ctx* ctx, int64_t* n) {
    if (!ctx || !n) return fail(ZFV_ERR_ARG, "null argument");
    if (ctx->have_mesh && ctx->faces_pending) {
        unsigned long long fc[2] = {0, 0};
        ZFV_HIP(hipMemcpyAsync(fc, ctx->s_faces.p, sizeof(fc), hipMemcpyDeviceToHost, ctx->stream));
        ZFV_HIP(hipStreamSynchronize(ctx->stream));
        ctx->n_faces = (int64_t)(fc[0] / 2 + fc[1]);
        ctx->faces_pending = false;
    }
    *n = ctx->have_mesh ? ctx->n_faces : 0;
    return ZFV_OK;
}

int zfv_mesh_reserve(zfv_ctx* ctx, int n_cells) {
    int r = need_ready(ctx, false);
    if (r) return r;
    if (n_cells < 0 || n_cells > std::numeric_limits<int>::max() / 9) return fail(ZFV_ERR_ARG, "n_cells out of range");
    if (ctx->nseg > 1) return fail(ZFV_ERR_UNSUPPORTED, "reserve on a batch of problems (no adaptation)");
    if (n_cells <= ctx->cap && ctx->s_marks.n >= (size_t)n_cells) return ZFV_OK;
    // the curve-ordered copy is rebuilt at the new stride: state and
    // gradients back in the mesh's order first
    r = to_mesh_order(ctx);
    if (r) return r;
    r = grad_to_mesh_order(ctx);
    if (r) return r;
    r = grow_capacity(ctx, n_cells);
    if (r) return r;
    const size_t cap = (size_t)ctx->cap, nr = (size_t)std::max(n_cells, 1);
    // adaptation scratch (refine_impl) and the criterion's unlimited gradient
    for (auto* b : {&ctx->s_marks, &ctx->s_rank})
        if (b->n < nr) ZFV_HIP(b->alloc(nr));
    for (auto* b : {&ctx->s_hole_at, &ctx->s_ref_list, &ctx->s_crs_list, &ctx->s_swap_src, &ctx->s_swap_dst, &ctx->s_fam})
        if (b->n < nr) ZFV_HIP(b->alloc(nr));
    if (ctx->s_faces.n < 2) ZFV_HIP(ctx->s_faces.alloc(2));
    const size_t tiles = (size_t)std::max(amr_scan3_tiles(n_cells), 1);
    if (ctx->s_tiles.n < tiles) ZFV_HIP(ctx->s_tiles.alloc(tiles));
    if (!ctx->marks_total) ZFV_HIP(hipHostMalloc((void**)&ctx->marks_total, sizeof(AmrMarks), hipHostMallocDefault));
    if (ctx->graw.n < 2 * cap) ZFV_HIP(ctx->graw.alloc(2 * cap));
    // the curve-ordered copy at the new stride (built at the next step)
    auto& co = ctx->co;
    if (co.mode != 0 && co.keys_ok && co.cap != (int)cap) {
        for (auto* b : {&co.cx, &co.cy, &co.vol, &co.h}) ZFV_HIP(b->alloc(cap));
        for (auto* b : {&co.area, &co.fcx, &co.fcy, &co.w, &co.nx, &co.ny}) ZFV_HIP(b->alloc(8 * cap));
        ZFV_HIP(co.code.alloc(cap));
        ZFV_HIP(co.adj.alloc(8 * cap));
        for (auto* b : {&co.perm, &co.iperm, &co.vals}) ZFV_HIP(b->alloc(cap));
        for (auto& k : co.keys) ZFV_HIP(k.alloc(cap));
        co.cap = (int)cap;
        co.epoch = -1;
        co.perm_epoch = -1;
    }
    if (co.mode != 0 && co.keys_ok && co.gtmp.n < 4 * cap) ZFV_HIP(co.gtmp.alloc(4 * cap));
    return ZFV_OK;
}

int zfv_mesh_download(zfv_ctx* ctx, const zfv_mesh_soa* o) {
    if (!ctx || !o) return fail(ZFV_ERR_ARG, "null argument");
    if (!ctx->have_mesh) return fail(ZFV_ERR_STATE, "no mesh uploaded");
    int rs = sync_host_mesh(ctx);
    if (rs) return rs;
    const HostMesh& m = ctx->hm;
    auto cp = [](const void* dst, const void* src, size_t bytes) {
        if (dst) std::memcpy(const_cast<void*>(dst), src, bytes);
    };
    const size_t n = m.n, ns = 8 * n, nv = 9 * n;
    cp(o->cx, m.cx.data(), n * 8); cp(o->cy, m.cy.data(), n * 8); cp(o->volume, m.vol.data(), n * 8);
    cp(o->vx, m.vx.data(), nv * 8); cp(o->vy, m.vy.data(), nv * 8);
    cp(o->level, m.level.data(), n * 4); cp(o->flag, m.flag.data(), n * 4);
    cp(o->b_idx, m.b_idx.data(), n * 4); cp(o->z_idx, m.z_idx.data(), n * 4);
    cp(o->boundary, m.bnd.data(), ns * 4); cp(o->adj, m.adj.data(), ns * 4);
    cp(o->nx, m.nx.data(), ns * 8); cp(o->ny, m.ny.data(), ns * 8);
    cp(o->fcx, m.fcx.data(), ns * 8); cp(o->fcy, m.fcy.data(), ns * 8);
    cp(o->area, m.area.data(), ns * 8);
    return ZFV_OK;
}

int zfv_smf_configure(zfv_ctx* ctx, const zfv_smf_params* p) {
    int r = check_ctx(ctx);
    if (r) return r;
    if (!p) return fail(ZFV_ERR_ARG, "null params");
    if (p->flux < ZFV_GODUNOV || p->flux > ZFV_CIR2)
        return fail(ZFV_ERR_ARG, "unsupported numerical flux id " + std::to_string(p->flux));
    if (p->limiter < ZFV_LIM_NONE || p->limiter > ZFV_LIM_SUPERBEE)
        return fail(ZFV_ERR_ARG, "unknown limiter id " + std::to_string(p->limiter));
    if (!(p->gamma > 1.0)) return fail(ZFV_ERR_ARG, "EOS gamma must be > 1");
    if (p->axial != 0 && p->axial != 1) return fail(ZFV_ERR_ARG, "axial must be 0 or 1");
    if (p->axial && ctx->have_mesh && !ctx->axial_mesh)
        return fail(ZFV_ERR_STATE, "axisymmetric solver on a mesh without the axisymmetric geometry");
    ctx->prm = *p;
    ctx->prm.acc = std::min(std::max(1, p->acc), 2);               // set_accuracy
    ctx->prm.cfl = std::max(0.0, std::min(p->cfl, 1.0));            // set_CFL
    if (std::isnan(ctx->prm.max_dt)) ctx->prm.max_dt = std::numeric_limits<double>::max();
    r = put_max_dt(ctx);
    if (r) return r;
    ctx->configured = true;
    ctx->dt_valid = false;
    ctx->graphs_clear();
    return ZFV_OK;
}

}  // extern "C"

namespace {
// The value goes to the device in stream order (a one-word fill), so steps
// enqueued before keep the old bound and captured graphs read the new one.
int put_max_dt(zfv_ctx* ctx) {
    unsigned long long bits;
    std::memcpy(&bits, &ctx->prm.max_dt, 8);
    ZFV_HIP(launch_fill_u64(reinterpret_cast<unsigned long long*>(ctx->dmax_dt.p), 1, bits, ctx->stream));
    return ZFV_OK;
}
}  // namespace

extern "C" {

int zfv_smf_get_params(zfv_ctx* ctx, zfv_smf_params* out) {
    int r = check_ctx(ctx);
    if (r) return r;
    if (!out) return fail(ZFV_ERR_ARG, "null output");
    if (!ctx->configured) return fail(ZFV_ERR_STATE, "solver not configured");
    *out = ctx->prm;
    return ZFV_OK;
}

int zfv_smf_set_max_dt(zfv_ctx* ctx, double max_dt) {
    int r = check_ctx(ctx);
    if (r) return r;
    ctx->prm.max_dt = max_dt;
    return put_max_dt(ctx);
}

static int materialize_grads(zfv_ctx* ctx);

static int state_copy(zfv_ctx* ctx, const double* const* src, bool to_device, hipMemcpyKind kind) {
    const int n = ctx->hm.n;
    for (int k = 0; k < 5; ++k)
        if (!src[k]) return fail(ZFV_ERR_ARG, "null state array");
    if (to_device) {   // the last step's input may be overwritten: its gradients first
        int r = materialize_grads(ctx);
        if (r) return r;
    }
    for (int k = 0; k < 5; ++k) {
        double* d = ctx->st[ctx->cur][k].p;
        if (to_device) ZFV_HIP(hipMemcpyAsync(d, src[k], n * sizeof(double), kind, ctx->stream));
        else ZFV_HIP(hipMemcpyAsync(const_cast<double*>(src[k]), d, n * sizeof(double), kind, ctx->stream));
    }
    // A batch's padding cells are never written by the step kernels: give
    // both alternating layers the uploaded values, so a download returns
    // them whatever the parity of the step count.
    if (to_device && ctx->nseg > 1)
        for (int k = 0; k < 5; ++k)
            ZFV_HIP(hipMemcpyAsync(ctx->st[1 - ctx->cur][k].p, ctx->st[ctx->cur][k].p, n * sizeof(double),
                                   hipMemcpyDeviceToDevice, ctx->stream));
    ZFV_HIP(hipStreamSynchronize(ctx->stream));
    return ZFV_OK;
}

int zfv_state_upload(zfv_ctx* ctx, const double* rho, const double* u, const double* v, const double* p,
                     const double* e) {
    int r = need_ready(ctx, false);
    if (r) return r;
    r = to_mesh_order(ctx);
    if (r) return r;
    const double* a[5] = {rho, u, v, p, e};
    r = state_copy(ctx, a, true, hipMemcpyHostToDevice);
    if (r) return r;
    ctx->have_state = true;
    ctx->dt_valid = false;
    return ZFV_OK;
}

int zfv_state_upload_device(zfv_ctx* ctx, const double* rho, const double* u, const double* v,
                            const double* p, const double* e) {
    int r = need_ready(ctx, false);
    if (r) return r;
    r = to_mesh_order(ctx);
    if (r) return r;
    const double* a[5] = {rho, u, v, p, e};
    r = state_copy(ctx, a, true, hipMemcpyDeviceToDevice);
    if (r) return r;
    ctx->have_state = true;
    ctx->dt_valid = false;
    return ZFV_OK;
}

int zfv_state_download(zfv_ctx* ctx, double* rho, double* u, double* v, double* p, double* e) {
    int r = need_ready(ctx);
    if (r) return r;
    r = to_mesh_order(ctx);
    if (r) return r;
    const double* a[5] = {rho, u, v, p, e};
    r = state_copy(ctx, a, false, hipMemcpyDeviceToHost);
    if (r) return r;
    return read_err(ctx);
}

int zfv_state_download_device(zfv_ctx* ctx, double* rho, double* u, double* v, double* p, double* e) {
    int r = need_ready(ctx);
    if (r) return r;
    r = to_mesh_order(ctx);
    if (r) return r;
    const double* a[5] = {rho, u, v, p, e};
    r = state_copy(ctx, a, false, hipMemcpyDeviceToDevice);
    if (r) return r;
    return read_err(ctx);
}

int zfv_wait_upload(zfv_ctx* ctx, const int32_t* wait) {
    int r = need_ready(ctx, false);
    if (r) return r;
    r = to_mesh_order(ctx);
    if (r) return r;
    if (!wait) return fail(ZFV_ERR_ARG, "null wait");
    ZFV_HIP(hipMemcpyAsync(ctx->wait[ctx->wcur].p, wait, ctx->hm.n * sizeof(int), hipMemcpyHostToDevice, ctx->stream));
    ZFV_HIP(hipStreamSynchronize(ctx->stream));
    return ZFV_OK;
}

int zfv_wait_download(zfv_ctx* ctx, int32_t* wait) {
    int r = need_ready(ctx, false);
    if (r) return r;
    r = to_mesh_order(ctx);
    if (r) return r;
    if (!wait) return fail(ZFV_ERR_ARG, "null wait");
    ZFV_HIP(hipMemcpyAsync(wait, ctx->wait[ctx->wcur].p, ctx->hm.n * sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
    ZFV_HIP(hipStreamSynchronize(ctx->stream));
    return ZFV_OK;
}

int zfv_grad_download(zfv_ctx* ctx, double* dx, double* dy) {
    int r = need_ready(ctx);
    if (r) return r;
    r = materialize_grads(ctx);
    if (r) return r;
    r = grad_to_mesh_order(ctx);
    if (r) return r;
    if (!dx || !dy) return fail(ZFV_ERR_ARG, "null output");
    const size_t row = (size_t)ctx->hm.n * sizeof(double), pitch = (size_t)ctx->cap * sizeof(double);
    ZFV_HIP(hipMemcpy2DAsync(dx, row, ctx->gdx.p, pitch, row, 4, hipMemcpyDeviceToHost, ctx->stream));
    ZFV_HIP(hipMemcpy2DAsync(dy, row, ctx->gdy.p, pitch, row, 4, hipMemcpyDeviceToHost, ctx->stream));
    ZFV_HIP(hipStreamSynchronize(ctx->stream));
    return ZFV_OK;
}

// SmFluid::update x n_steps (sm_fluid.cpp:81-102).
// One SmFluid::update (sm_fluid.cpp:81-102) enqueued on the context stream.
static int step_once(zfv_ctx* ctx) {
    const DevMesh M = ctx->smesh();
    const Eos eos = ctx->eos();
    const zfv_smf_params& P = ctx->prm;
    StepCtl ctl = make_ctl(ctx);
    const int nxt = (ctx->cur + 1) % 2;   // layers 0/1 alternate as init/next, 2 is half
    DevState in = ctx->layer(ctx->cur), out = ctx->layer(nxt);
    if (P.acc == 1) {
        ZFV_HIP(timed(ctx, "flux1", [&] {
            return launch_flux1(P.flux, ctx->layout[0], ctx->block[0], M, in, out, eos, ctl, ctx->stream);
        }));
    } else if (ctx->tile_step()) {
        int* w_in = ctx->wait[ctx->wcur].p;
        int* w_out = ctx->wait[1 - ctx->wcur].p;
        const TileMesh T = ctx->tmesh();
        ZFV_HIP(timed(ctx, "tile", [&] {
            return launch_tile_step(P.limiter, ctx->tile_variant, ctx->tile_persistent, T, in, out, w_in, w_out, eos,
                                    ctl, ctx->stream);
        }));
        ctx->wcur = 1 - ctx->wcur;
        ctx->grad_pending = true;   // the step's gradients were not stored
        ctx->co.grad = false;
    } else {
        DevGrad G{ctx->gdx.p, ctx->gdy.p};
        DevState half = ctx->layer(2);
        int* w_in = ctx->wait[ctx->wcur].p;
        int* w_out = ctx->wait[1 - ctx->wcur].p;
        ZFV_HIP(timed(ctx, "grad_pred", [&] {
            return launch_grad_pred(P.limiter, true, ctx->gp_layout(), ctx->gp_block(), M, in, G, half, w_in, eos, ctl,
                                    ctx->stream);
        }));
        ZFV_HIP(timed(ctx, "corr", [&] {
            return launch_corr(P.flux, ctx->layout[2], ctx->block[2], M, in, half, G, w_in, w_out, out, eos, ctl,
                               ctx->stream);
        }));
        ctx->wcur = 1 - ctx->wcur;
        ctx->co.grad = ctx->co.state;
        ctx->grad_pending = false;
    }
    ctx->cur = nxt;
    ctx->step_no += 1;
    return ZFV_OK;
}

// The gradients of the last step, when a tiled step did not store them:
// the gradient kernel on that step's input, layer 1 - cur (intact until the
// next step or upload, before which this runs).
static int materialize_grads(zfv_ctx* ctx) {
    if (!ctx->grad_pending) return ZFV_OK;
    const DevGrad G{ctx->gdx.p, ctx->gdy.p};
    const int prev = 1 - ctx->cur;
    ZFV_HIP(launch_grad_pred(ctx->prm.limiter, false, ctx->gp_layout(), ctx->gp_block(), ctx->dmesh(),
                             ctx->layer(prev), G, ctx->layer(2), ctx->wait[ctx->wcur].p, ctx->eos(), make_ctl(ctx),
                             ctx->stream));
    ctx->grad_pending = false;
    ctx->co.grad = false;
    return ZFV_OK;
}

// One MUSCL step (acc 2) of the adjoint's recompute: the step's kernels and
// dt-slot rotation on explicit buffers (input, half state, gradient,
// output), so the recomputed segment lands where the reverse sweep reads it.
static int recompute_step(zfv_ctx* ctx, const DevState& in, const DevState& half, const DevGrad& G,
                          const DevState& out) {
    const DevMesh M = ctx->smesh();
    const Eos eos = ctx->eos();
    const zfv_smf_params& P = ctx->prm;
    StepCtl ctl = make_ctl(ctx);
    int* w_in = ctx->wait[ctx->wcur].p;
    int* w_out = ctx->wait[1 - ctx->wcur].p;
    ZFV_HIP(launch_grad_pred(P.limiter, true, ctx->gp_layout(), ctx->gp_block(), M, in, G, half, w_in, eos, ctl,
                             ctx->stream));
    ZFV_HIP(launch_corr(P.flux, ctx->layout[2], ctx->block[2], M, in, half, G, w_in, w_out, out, eos, ctl,
                        ctx->stream));
    ctx->wcur = 1 - ctx->wcur;
    ctx->step_no += 1;
    return ZFV_OK;
}

// Steps per captured graph: the dt slots rotate with period 3 and the
// layers and wait buffers swap with period 2.
constexpr int kGraphSteps = 6;

static int graph_for_phase(zfv_ctx* ctx, hipGraphExec_t* out) {
    const int key = (int)(ctx->step_no % 3) | (ctx->cur << 2) | (ctx->wcur << 3) | (ctx->prm.acc << 4) |
                    ((ctx->co.state ? 1 : 0) << 7);
    for (auto& g : ctx->graphs)
        if (g.key == key) {
            *out = g.exec;
            return ZFV_OK;
        }
    const int cur = ctx->cur, wcur = ctx->wcur;
    const long long step_no = ctx->step_no;
    ZFV_HIP(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
    int r = ZFV_OK;
    for (int k = 0; k < kGraphSteps && r == ZFV_OK; ++k) r = step_once(ctx);
    hipGraph_t graph = nullptr;
    const hipError_t ec = hipStreamEndCapture(ctx->stream, &graph);
    ctx->cur = cur;
    ctx->wcur = wcur;
    ctx->step_no = step_no;
    if (r) {
        if (graph) (void)hipGraphDestroy(graph);
        return r;
    }
    if (ec != hipSuccess) return fail(ZFV_ERR_HIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(ec));
    hipGraphExec_t exec = nullptr;
    const hipError_t ei = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    if (ei != hipSuccess) return fail(ZFV_ERR_HIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(ei));
    ctx->graphs.push_back({key, exec});
    *out = exec;
    return ZFV_OK;
}

int zfv_smf_step(zfv_ctx* ctx, int n_steps) {
    int r = need_ready(ctx, true, true);
    if (r) return r;
    if (n_steps < 0) return fail(ZFV_ERR_ARG, "negative step count");
    r = to_curve_order(ctx);
    if (r) return r;
    r = ensure_dt(ctx);
    if (r) return r;
    int it = 0;
    if (ctx->use_graphs && ctx->stream && !ctx->profile) {
        while (n_steps - it >= kGraphSteps) {
            hipGraphExec_t exec = nullptr;
            r = graph_for_phase(ctx, &exec);
            if (r) return r;
            ZFV_HIP(hipGraphLaunch(exec, ctx->stream));
            ctx->step_no += kGraphSteps;   // cur and wcur return to their values after 6 swaps
            it += kGraphSteps;
        }
    }
    for (; it < n_steps; ++it) {
        r = step_once(ctx);
        if (r) return r;
    }
    return ZFV_OK;
}

int zfv_graph_enable(zfv_ctx* ctx, int on) {
    int r = check_ctx(ctx);
    if (r) return r;
    ctx->use_graphs = on != 0;
    if (!ctx->use_graphs) ctx->graphs_clear();
    return ZFV_OK;
}

int zfv_set_cell_order(zfv_ctx* ctx, int mode) {
    int r = check_ctx(ctx);
    if (r) return r;
    if (mode < 0 || mode > 2)
        return fail(ZFV_ERR_ARG, "cell order must be 0 (mesh), 1 (curve for steps) or 2 (curve, rebuilt in full)");
    if (mode == 0 && ctx->have_mesh) {
        r = to_mesh_order(ctx);
        if (r) return r;
        r = grad_to_mesh_order(ctx);
        if (r) return r;
    }
    ctx->co.mode = mode;
    ctx->graphs_clear();
    return ZFV_OK;
}

int zfv_get_curve_order(zfv_ctx* ctx, int32_t* perm) {
    int r = check_ctx(ctx);
    if (r) return r;
    if (!perm) return fail(ZFV_ERR_ARG, "null output");
    if (!ctx->have_mesh) return fail(ZFV_ERR_STATE, "no mesh uploaded");
    auto& co = ctx->co;
    if (co.mode == 0 || !co.refined || ctx->nseg > 1) return fail(ZFV_ERR_STATE, "the steps run in the mesh's order");
    if (co.epoch != ctx->mesh_epoch || co.cap != ctx->cap) {
        r = build_curve(ctx);
        if (r) return r;
    }
    ZFV_HIP(hipMemcpyAsync(perm, co.perm.p, (size_t)ctx->hm.n * sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
    ZFV_HIP(hipStreamSynchronize(ctx->stream));
    return ZFV_OK;
}

int zfv_get_cell_order(zfv_ctx* ctx, int* mode, int* active) {
    int r = check_ctx(ctx);
    if (r) return r;
    if (mode) *mode = ctx->co.mode;
    if (active) *active = ctx->co.state ? 1 : 0;
    return ZFV_OK;
}

int zfv_smf_dt(zfv_ctx* ctx, double* dt) {
    int r = need_ready(ctx, false);
    if (r) return r;
    if (!dt) return fail(ZFV_ERR_ARG, "null output");
    ZFV_HIP(hipMemcpyAsync(dt, ctx->dt_out.p, sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    ZFV_HIP(hipStreamSynchronize(ctx->stream));
    return read_err(ctx);
}

int zfv_smf_compute_grad(zfv_ctx* ctx) {
    int r = need_ready(ctx, true, true);
    if (r) return r;
    r = to_mesh_order(ctx);
    if (r) return r;
    ctx->co.grad = false;
    r = ensure_dt(ctx);
    if (r) return r;
    StepCtl ctl = make_ctl(ctx);
    DevGrad G{ctx->gdx.p, ctx->gdy.p};
    ZFV_HIP(launch_grad_pred(ctx->prm.limiter, false, ctx->gp_layout(), ctx->gp_block(), ctx->dmesh(), ctx->layer(ctx->cur), G,
                             ctx->layer(2),
                             ctx->wait[ctx->wcur].p, ctx->eos(), ctl, ctx->stream));
    ctx->grad_pending = false;
    ZFV_HIP(hipStreamSynchronize(ctx->stream));
    return ZFV_OK;
}

int zfv_smf_compute_dt(zfv_ctx* ctx, double* dt) {
    int r = need_ready(ctx, true, true);
    if (r) return r;
    ctx->dt_valid = false;
    r = ensure_dt(ctx);
    if (r) return r;
    unsigned long long bits = 0;
    ZFV_HIP(hipMemcpyAsync(&bits, ctx->dtmin.p + (ctx->step_no % 3), 8, hipMemcpyDeviceToHost, ctx->stream));
    ZFV_HIP(hipStreamSynchronize(ctx->stream));
    double dmin;
    std::memcpy(&dmin, &bits, 8);
    double v = ctx->prm.cfl * dmin;
    if (dt) *dt = (ctx->prm.max_dt < v) ? ctx->prm.max_dt : v;
    return ZFV_OK;
}

int zfv_conserved_totals(zfv_ctx* ctx, double* out4) {
    int r = need_ready(ctx, true, true);
    if (r) return r;
    r = to_mesh_order(ctx);   // the partial sums run in the mesh's order
    if (r) return r;
    if (!out4) return fail(ZFV_ERR_ARG, "null output");
    const int nb = (ctx->hm.n + kBlock - 1) / kBlock;
    ZFV_HIP(launch_totals(ctx->ax(ctx->dmesh()), ctx->layer(ctx->cur), ctx->partial.p, nb, ctx->stream));
    std::vector<double> part(4 * (size_t)nb);
    ZFV_HIP(hipMemcpyAsync(part.data(), ctx->partial.p, part.size() * 8, hipMemcpyDeviceToHost, ctx->stream));
    ZFV_HIP(hipStreamSynchronize(ctx->stream));
    for (int k = 0; k < 4; ++k) out4[k] = 0.0;
    for (int b = 0; b < nb; ++b)
        for (int k = 0; k < 4; ++k) out4[k] += part[4 * (size_t)b + k];
    return read_err(ctx);
}

int zfv_conserved_totals_noalloc(zfv_ctx* ctx, double* out4) {
    int r = need_ready(ctx, true, true);
    if (r) return r;
    if (!out4) return fail(ZFV_ERR_ARG, "null output");
    const int nb = (ctx->hm.n + kBlock - 1) / kBlock;
    if (nb > kScratchBlocks)
        return fail(ZFV_ERR_ARG, "mesh of " + std::to_string(ctx->hm.n) + " cells exceeds the " +
                                     std::to_string((long long)kScratchBlocks * kBlock) +
                                     " cells of the no-allocation scratch");
    double* sp = scratch_of(ctx->device);
    if (!sp) return fail(ZFV_ERR_STATE, "no scratch address for this device");
    r = to_mesh_order(ctx);   // the partial sums run in the mesh's order
    if (r) return r;
    ZFV_HIP(launch_totals(ctx->ax(ctx->dmesh()), ctx->layer(ctx->cur), sp + 4, nb, ctx->stream));
    ZFV_HIP(launch_totals_sum(sp + 4, nb, sp, ctx->stream));
    double res[4];
    ZFV_HIP(hipMemcpyAsync(res, sp, sizeof(res), hipMemcpyDeviceToHost, ctx->stream));
    ZFV_HIP(hipStreamSynchronize(ctx->stream));
    for (int k = 0; k < 4; ++k) out4[k] = res[k];
    return read_err(ctx);
}

// ---- AMR -------------------------------------------------------------------
int zfv_flags_upload(zfv_ctx* ctx, const int32_t* flags) {
    int r = need_ready(ctx, false);
    if (r) return r;
    if (!flags) return fail(ZFV_ERR_ARG, "null flags");
    ZFV_HIP(hipMemcpyAsync(ctx->flag.p, flags, ctx->hm.n * sizeof(int), hipMemcpyHostToDevice, ctx->stream));
    ZFV_HIP(hipStreamSynchronize(ctx->stream));
    ctx->host_stale = true;
    return ZFV_OK;
}

int zfv_flags_download(zfv_ctx* ctx, int32_t* flags) {
    int r = need_ready(ctx, false);
    if (r) return r;
    if (!flags) return fail(ZFV_ERR_ARG, "null flags");
    ZFV_HIP(hipMemcpyAsync(flags, ctx->flag.p, ctx->hm.n * sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
    ZFV_HIP(hipStreamSynchronize(ctx->stream));
    return ZFV_OK;
}

// SmFluid::set_flags (sm_fluid.cpp:517-643): gradient, then the criterion.
int zfv_set_flags(zfv_ctx* ctx, int criterion, double threshold) {
    int r = need_ready(ctx, true, true);
    if (r) return r;
    if (criterion != 0 && criterion != 1) return fail(ZFV_ERR_ARG, "unknown refinement criterion");
    r = to_mesh_order(ctx);
    if (r) return r;
    ctx->co.grad = false;
    r = ensure_dt(ctx);
    if (r) return r;
    StepCtl ctl = make_ctl(ctx);
    DevGrad G{ctx->gdx.p, ctx->gdy.p};
    if (ctx->graw.n < 2 * (size_t)ctx->cap) ZFV_HIP(ctx->graw.alloc(2 * (size_t)ctx->cap));
    ZFV_HIP(launch_grad_pred(ctx->prm.limiter, false, ctx->gp_layout(), ctx->gp_block(), ctx->dmesh(), ctx->layer(ctx->cur), G,
                             ctx->layer(2),
                             ctx->wait[ctx->wcur].p, ctx->eos(), ctl, ctx->stream, ctx->graw.p));
    ctx->grad_pending = false;
    const DevGrad R{ctx->graw.p, ctx->graw.p + ctx->cap};   // unlimited LSM gradient of rho
    ZFV_HIP(amr_criterion(ctx->amesh(), ctx->layer(ctx->cur), R, criterion, threshold, ctx->stream));
    ctx->host_stale = true;
    return ZFV_OK;
}

// amr::balance_flags (mesh/amr/balancing.h:15-35) alone.
int zfv_balance_flags(zfv_ctx* ctx, int max_level) {
    int r = need_ready(ctx, false);
    if (r) return r;
    if (max_level < 0 || max_level > 15) return fail(ZFV_ERR_ARG, "max_level out of [0, 15]");
    ZFV_HIP(amr_balance(ctx->amesh(), max_level, ctx->stream));
    std::swap(ctx->flag, ctx->flag0);   // the balanced flags
    ctx->host_stale = true;
    return ZFV_OK;
}

// Exclusive scan of the marks; the totals land in pinned host memory
// (ctx->marks_total), valid after the next synchronisation.
static int scan3(zfv_ctx* ctx, const AmrMarks* in, AmrMarks* out, int count) {
    const int tiles = amr_scan3_tiles(count);
    if (ctx->s_tiles.n < (size_t)std::max(tiles, 1)) ZFV_HIP(ctx->s_tiles.alloc(std::max(tiles, 1)));
    if (!ctx->marks_total) ZFV_HIP(hipHostMalloc((void**)&ctx->marks_total, sizeof(AmrMarks), hipHostMallocDefault));
    ZFV_HIP(amr_scan3(in, out, ctx->s_tiles.p, count, ctx->marks_total, ctx->stream));
    return ZFV_OK;
}

// EuMesh::refine (eu_mesh.cpp:290-359) = balance_flags + amr::apply
// (apply.h:43-94).  One host synchronisation: the scan tail that sizes the
// new mesh.  The derive pass also seeds the next step's dt, so no further
// launch is needed before stepping.  Consistency failures are reported by
// the next synchronising call through the error word.
// EuMesh::refine (eu_mesh.cpp:290-359) = balance_flags + amr::apply
// (apply.h:43-94) with the solver's distributor: kRefineConst / kRefineSlope
// (SmFluid, sm_fluid.cpp:372-470) or kRefineMm (MmFluid, mm_fluid.cpp:765-830,
// which reads the gradients of zfv_mm_set_flags).  One host
// synchronisation: the scan totals that size the new mesh.  The derive pass
// (SmFluid) or the cache pass (MmFluid) also seeds the next step's dt.  The
// other solver's state does not follow the mesh and is invalidated.
// Consistency failures are reported by the next synchronising call through
// the error word.
enum { kRefineConst = 0, kRefineSlope = 1, kRefineMm = 2 };

static MmLayer mm_layer(zfv_ctx* ctx, int i);
static int mm_ready(zfv_ctx* ctx, bool state);

static int refine_impl(zfv_ctx* ctx, int max_level, int mode) {
    const bool mm = mode == kRefineMm;
    auto& S = ctx->mm;
    int r = mm ? mm_ready(ctx, true) : need_ready(ctx, true, true);
    if (r) return r;
    if (max_level < 0 || max_level > 15) return fail(ZFV_ERR_ARG, "max_level out of [0, 15]");
    if (ctx->nseg > 1) return fail(ZFV_ERR_UNSUPPORTED, "adaptation of a batch of problems");
    if (ctx->axial_mesh) return fail(ZFV_ERR_UNSUPPORTED, "adaptation of an axisymmetric mesh");
    if (mm && S.grad_epoch != ctx->mesh_epoch)
        return fail(ZFV_ERR_STATE, "multi-material refine needs the gradients of zfv_mm_set_flags on this mesh");
    hipStream_t s = ctx->stream;
    if (max_level == 0) return ZFV_OK;   // EuMesh::adaptive() is false
    if (!mm) {
        r = to_mesh_order(ctx);
        if (r) return r;
        r = materialize_grads(ctx);   // the slope distributor reads them
        if (r) return r;
        r = grad_to_mesh_order(ctx);
        if (r) return r;
    }
    const int n = ctx->hm.n;
    ZFV_HIP(amr_balance(ctx->amesh(), max_level, s));
    std::swap(ctx->flag, ctx->flag0);   // the balanced flags
    ctx->host_stale = true;
    for (auto* b : {&ctx->s_marks, &ctx->s_rank})
        if (b->n < (size_t)n) ZFV_HIP(b->alloc(n));
    // The steps' curve order is carried across the adaptation (k_amr_mark
    // counts, k_amr_geometry places) when the copy of this mesh is current
    // and the capacity does not grow.
    auto& co = ctx->co;
    const bool carry_ok = !mm && co.mode == 1 && co.keys_ok && co.refined && co.epoch == ctx->mesh_epoch &&
                          co.cap == ctx->cap;
    co.perm_epoch = -1;
    ZFV_HIP(amr_mark(ctx->amesh(), ctx->s_marks.p, carry_ok ? co.perm.p : nullptr, s));
    r = scan3(ctx, ctx->s_marks.p, ctx->s_rank.p, n);
    if (r) return r;
    ZFV_HIP(hipStreamSynchronize(s));
    const volatile AmrMarks* tot = ctx->marks_total;
    const int n_refine = tot->ref;
    const int n_parents = tot->head;
    const int nz_total = tot->nz;   // refined + coarsened cells
    const int n_children = 4 * n_refine;
    const long long n_large_ll = (long long)n + n_parents + n_children;
    if (n_large_ll > std::numeric_limits<int>::max() / 9) return fail(ZFV_ERR_ARG, "adapted mesh too large");
    const int n_large = (int)n_large_ll;
    // Statistics (statistics.h:77-80): n_short = retained + children + parents
    const int n_short = (n - nz_total) + n_children + n_parents;
    if (n_large <= n) return ZFV_OK;
    if (nz_total != n_refine + 4 * n_parents) return fail(ZFV_ERR_STATE, "unbalanced coarsening flags");
    if (n_large > ctx->cap) {
        r = grow_capacity(ctx, std::max(n_large, (int)std::min<long long>(2LL * ctx->cap, std::numeric_limits<int>::max() / 9)));
        if (r) return r;
    }
    for (auto* b : {&ctx->s_hole_at, &ctx->s_ref_list, &ctx->s_crs_list, &ctx->s_swap_src, &ctx->s_swap_dst})
        if (b->n < (size_t)std::max(n, 1)) ZFV_HIP(b->alloc(std::max(n, 1)));
    if (ctx->s_faces.n < 2) ZFV_HIP(ctx->s_faces.alloc(2));
    if (ctx->s_fam.n < 4 * (size_t)std::max(n_parents, 1)) ZFV_HIP(ctx->s_fam.alloc(4 * (size_t)std::max(n_parents, 1)));

    AmrMesh A = ctx->amesh();   // A.n = n (cells before adaptation)
    const int rd = (int)(ctx->step_no % 3), wr = (int)((ctx->step_no + 1) % 3);
    ZFV_HIP(amr_positions(A, n_large, n_short, n_parents, ctx->s_rank.p, ctx->s_hole_at.p, ctx->s_ref_list.p,
                          ctx->s_crs_list.p, ctx->s_fam.p, ctx->s_faces.p, ctx->dtmin.p + rd, ctx->dtmin.p + wr, s));
    ZFV_HIP(amr_swapmap(A, n_short, n_large, nz_total, ctx->s_rank.p, ctx->s_hole_at.p, ctx->s_swap_src.p,
                        ctx->s_swap_dst.p, ctx->err.p, s));
    const DevState z = mm ? DevState{} : ctx->layer(ctx->cur);
    int* wait = mm ? nullptr : ctx->wait[ctx->wcur].p;
    DevGrad G{ctx->gdx.p, ctx->gdy.p};
    const bool carried = carry_ok && co.cap == ctx->cap && co.vals.n >= (size_t)n_short && tot->cc == n_short;
    ZFV_HIP(amr_geometry(A, z, G, wait, ctx->eos(), mode == kRefineSlope, ctx->s_ref_list.p, n_refine,
                         ctx->s_fam.p, n_parents, co.perm.p, ctx->s_rank.p, carried ? co.vals.p : nullptr,
                         co.iperm.p, s));
    if (carried) std::swap(co.perm, co.vals);
    const size_t cap = (size_t)ctx->cap;
    if (mm)
        ZFV_HIP(mm_distribute(A, mm_layer(ctx, S.cur), S.mix, S.gx.p, S.gy.p, ctx->s_ref_list.p, n_refine,
                              ctx->s_fam.p, n_parents, ctx->err.p, s));
    // moves: one per hole (the kernel reads the hole count from the ranks);
    // there are at most min(changed cells, positions above n_short) of them
    const int bound = std::min(nz_total, n_large - n_short);
    ZFV_HIP(amr_move(A, z, wait, n_short, nz_total, ctx->s_rank.p, bound, ctx->s_swap_src.p, ctx->s_swap_dst.p, s));
    if (mm)
        ZFV_HIP(mm_move(mm_layer(ctx, S.cur), S.mix, (int)cap, n, n_short, nz_total, ctx->s_rank.p, bound,
                        ctx->s_swap_src.p, ctx->s_swap_dst.p, s));
    ctx->hm.n = n_short;
    ZFV_HIP(amr_derive(ctx->amesh(), z, ctx->eos(), ctx->s_faces.p, ctx->dtmin.p + rd, ctx->err.p, s));
    ctx->faces_pending = true;
    ctx->graphs_clear();   // the cell count changed
    ctx->mesh_epoch += 1;
    ctx->co.perm_epoch = carried ? ctx->mesh_epoch : -1;
    ctx->tiles.ok = false;   // no longer a plain Rectangle
    ctx->rects.clear();
    ctx->co.refined = ctx->co.keys_ok;
    ctx->co.lmax = std::max(ctx->co.lmax, max_level);
    if (mm) {
        const int slot = (int)(S.step_no % 3);
        ZFV_HIP(launch_fill_u64(S.dtmin.p + slot, 1, 0x7ff0000000000000ULL, s));
        ZFV_HIP(mm_cache(ctx->dmesh(), mm_layer(ctx, S.cur), S.mix, S.dtmin.p + slot, s));
        S.epoch = ctx->mesh_epoch;
        ctx->have_state = false;   // the single-material state did not follow
        ctx->co.state = false;
        ctx->co.grad = false;
        ctx->dt_valid = false;
    } else {
        ctx->dt_valid = true;
    }
    return ZFV_OK;
}

int zfv_refine(zfv_ctx* ctx, int max_level, int distributor) {
    if (distributor != 0 && distributor != 1) return fail(ZFV_ERR_ARG, "distributor must be 0 (const) or 1 (slope)");
    return refine_impl(ctx, max_level, distributor == 1 ? kRefineSlope : kRefineConst);
}

// ---- discrete adjoint -------------------------------------------------------
static DevState state_at(double* base, int n) { return DevState{base, base + n, base + 2 * (size_t)n, base + 3 * (size_t)n, base + 4 * (size_t)n}; }

static int copy_state(zfv_ctx* ctx, const DevState& dst, const DevState& src) {
    const size_t b = (size_t)ctx->hm.n * sizeof(double);
    double* const d[5] = {dst.r, dst.u, dst.v, dst.p, dst.e};
    const double* const sr[5] = {src.r, src.u, src.v, src.p, src.e};
    for (int k = 0; k < 5; ++k) ZFV_HIP(hipMemcpyAsync(d[k], sr[k], b, hipMemcpyDeviceToDevice, ctx->stream));
    return ZFV_OK;
}

// The adjoint covers the MUSCL + HLLC step on one Cartesian mesh: no
// sub-face slots, canonical normals, limiters without a derivative table
// other than none / unit / minmod (superbee) / MC are refused.
static int adjoint_supported(zfv_ctx* ctx) {
    const zfv_smf_params& P = ctx->prm;
    if (P.acc != 2 || P.flux != ZFV_HLLC) return fail(ZFV_ERR_UNSUPPORTED, "adjoint: MUSCL (acc 2) with HLLC only");
    const int L = P.limiter;
    if (!(L == ZFV_LIM_NONE || L == ZFV_LIM_UNIT || L == ZFV_LIM_MINMOD || L == ZFV_LIM_MC || L == ZFV_LIM_SUPERBEE))
        return fail(ZFV_ERR_UNSUPPORTED, "adjoint: limiter must be none, unit, minmod, superbee or MC");
    if (ctx->nseg != 1) return fail(ZFV_ERR_UNSUPPORTED, "adjoint of a batch");
    if (P.axial) return fail(ZFV_ERR_UNSUPPORTED, "adjoint of the axisymmetric step");
    int r = sync_host_mesh(ctx);
    if (r) return r;
    const HostMesh& m = ctx->hm;
    for (int c = 0; c < m.n; ++c)
        for (int s8 = 0; s8 < 8; ++s8) {
            const size_t i = 8 * (size_t)c + s8;
            if (s8 >= 4 && m.bnd[i] != ZFV_UNDEFINED) return fail(ZFV_ERR_UNSUPPORTED, "adjoint: adapted mesh");
            if (s8 < 4 && !(m.nx[i] == kNX[s8] && m.ny[i] == kNY[s8]))
                return fail(ZFV_ERR_UNSUPPORTED, "adjoint: non-Cartesian normals");
            // The reverse sweep gathers a face's contributions to the neighbour
            // through the neighbour's opposite slot: every inner or periodic
            // face must be seen from both sides (a periodic side needs its
            // periodic partner side).
            if (s8 < 4 && (m.bnd[i] == ZFV_INNER || m.bnd[i] == ZFV_PERIODIC)) {
                const size_t o = 8 * (size_t)m.adj[i] + (s8 ^ 1);
                if (m.adj[o] != c || m.bnd[o] != m.bnd[i])
                    return fail(ZFV_ERR_UNSUPPORTED, "adjoint: unpaired periodic boundary");
            }
        }
    return ZFV_OK;
}

int zfv_adjoint_forward(zfv_ctx* ctx, int n_steps, int every, const double* window, double* J) {
    int r = need_ready(ctx, true, true);
    if (r) return r;
    if (n_steps < 1 || every < 1 || !window || !J) return fail(ZFV_ERR_ARG, "adjoint_forward: bad arguments");
    r = to_mesh_order(ctx);   // checkpoints and the reverse sweep use the mesh's order
    if (r) return r;
    r = grad_to_mesh_order(ctx);
    if (r) return r;
    r = adjoint_supported(ctx);
    if (r) return r;
    auto& A = ctx->adjoint;
    const int n = ctx->hm.n;
    const int nck = (n_steps + every - 1) / every;
    A.ready = false;
    // the derivative is that of the plain (non-isolating) step: no cell may
    // start in the robust isolation mode, and none may be isolated on the way
    r = read_err(ctx);
    if (r) return r;
    {
        std::vector<int> w(n);
        ZFV_HIP(hipMemcpy(w.data(), ctx->wait[ctx->wcur].p, n * sizeof(int), hipMemcpyDeviceToHost));
        for (int c = 0; c < n; ++c)
            if (w[c] > 0) return fail(ZFV_ERR_UNSUPPORTED, "adjoint: cells in the isolation mode");
    }
    ZFV_HIP(hipMemsetAsync(ctx->err.p, 0, sizeof(int), ctx->stream));
    A.ckpt.resize(nck);
    for (auto& b : A.ckpt)
        if (b.n != 5 * (size_t)n) ZFV_HIP(b.alloc(5 * (size_t)n));
    if (A.final_z.n != 5 * (size_t)n) ZFV_HIP(A.final_z.alloc(5 * (size_t)n));
    r = ensure_dt(ctx);
    if (r) return r;
    for (int i = 0; i < n_steps; ++i) {
        if (i % every == 0) {
            r = copy_state(ctx, state_at(A.ckpt[i / every].p, n), ctx->layer(ctx->cur));
            if (r) return r;
        }
        r = step_once(ctx);
        if (r) return r;
    }
    r = copy_state(ctx, state_at(A.final_z.p, n), ctx->layer(ctx->cur));
    if (r) return r;
    // J = sum over the window's cells of p^2 V, in cell order on the host
    std::vector<double> p(n);
    ZFV_HIP(hipMemcpyAsync(p.data(), ctx->layer(ctx->cur).p, n * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    ZFV_HIP(hipStreamSynchronize(ctx->stream));
    r = read_err(ctx);
    if (r) return r;
    {
        int e = 0;
        ZFV_HIP(hipMemcpy(&e, ctx->err.p, sizeof(int), hipMemcpyDeviceToHost));
        if (e & kErrIsolated) return fail(ZFV_ERR_UNSUPPORTED, "adjoint: a cell was isolated (bad state) during the run");
    }
    const HostMesh& m = ctx->hm;
    double j = 0.0;
    for (int c = 0; c < n; ++c) {
        const double x = m.cx[c], y = m.cy[c];
        if (x >= window[0] && x <= window[1] && y >= window[2] && y <= window[3]) j = j + p[c] * p[c] * m.vol[c];
    }
    *J = j;
    A.n_steps = n_steps;
    A.every = every;
    for (int k = 0; k < 4; ++k) A.win[k] = window[k];
    A.ready = true;
    return ZFV_OK;
}

int zfv_adjoint_backward(zfv_ctx* ctx, double* lr, double* lu, double* lv, double* lp, double* le) {
    int r = need_ready(ctx, true, true);
    if (r) return r;
    auto& A = ctx->adjoint;
    if (!A.ready) return fail(ZFV_ERR_STATE, "adjoint_backward: no forward run to reverse");
    r = to_mesh_order(ctx);
    if (r) return r;
    r = materialize_grads(ctx);   // the reverse sweep reuses the layers
    if (r) return r;
    if (!lr || !lu || !lv || !lp || !le) return fail(ZFV_ERR_ARG, "adjoint_backward: null output");
    const int n = ctx->hm.n, cap = ctx->cap;
    // reverse-sweep buffers: segment states, adjoint states, per-slot face-state adjoints
    if (A.seg.n < (size_t)A.every * 5 * n) ZFV_HIP(A.seg.alloc((size_t)A.every * 5 * n));
    if (A.seg_h.n < (size_t)A.every * 5 * n) ZFV_HIP(A.seg_h.alloc((size_t)A.every * 5 * n));
    if (A.seg_g.n < (size_t)A.every * 8 * cap) ZFV_HIP(A.seg_g.alloc((size_t)A.every * 8 * cap));
    if (A.seg_dt.n < (size_t)A.every) ZFV_HIP(A.seg_dt.alloc(A.every));
    for (auto* b : {&A.lam[0], &A.lam[1], &A.lz})
        if (b->n != 5 * (size_t)n) ZFV_HIP(b->alloc(5 * (size_t)n));
    if (A.W.n != 4 * (size_t)cap) ZFV_HIP(A.W.alloc(4 * (size_t)cap));
    if (A.lG.n != 8 * (size_t)cap) ZFV_HIP(A.lG.alloc(8 * (size_t)cap));
    if (A.ZB.n != 20 * (size_t)cap) ZFV_HIP(A.ZB.alloc(20 * (size_t)cap));
    if (A.NG.n != 16 * (size_t)cap) ZFV_HIP(A.NG.alloc(16 * (size_t)cap));
    if (A.dtb.n != 1) ZFV_HIP(A.dtb.alloc(1));
    hipStream_t s = ctx->stream;
    const DevMesh M = ctx->dmesh();
    const Eos eos = ctx->eos();
    const DevGrad lG{A.lG.p, A.lG.p + 4 * (size_t)cap};
    ZFV_HIP(adj_seed(M, state_at(A.final_z.p, n), A.win, state_at(A.lam[0].p, n), s));
    int cur = 0;
    const int nck = (int)A.ckpt.size();
    const DevState start = ctx->layer(ctx->cur), scratch = ctx->layer(1 - ctx->cur);
    // state entering step j of the segment (j = 0: the checkpoint, copied to the current layer)
    auto seg_state = [&](int j) { return j == 0 ? start : state_at(A.seg.p + (size_t)j * 5 * n, n); };
    for (int seg = nck - 1; seg >= 0; --seg) {
        const int s0 = seg * A.every, len = std::min(A.every, A.n_steps - s0);
        r = copy_state(ctx, start, state_at(A.ckpt[seg].p, n));
        if (r) return r;
        ctx->dt_valid = false;
        r = ensure_dt(ctx);
        if (r) return r;
        // recompute the segment: each step writes its half state, gradient and
        // output straight into the segment buffers (the last output, which is
        // the next checkpoint, into scratch), and its dt slot is kept
        for (int j = 0; j < len; ++j) {
            const int rd = (int)(ctx->step_no % 3);
            ZFV_HIP(hipMemcpyAsync(A.seg_dt.p + j, ctx->dtmin.p + rd, 8, hipMemcpyDeviceToDevice, s));
            double* gj = A.seg_g.p + (size_t)j * 8 * cap;
            r = recompute_step(ctx, seg_state(j), state_at(A.seg_h.p + (size_t)j * 5 * n, n),
                               DevGrad{gj, gj + 4 * (size_t)cap}, j + 1 < len ? seg_state(j + 1) : scratch);
            if (r) return r;
        }
        // reverse through the segment
        const int* wait = ctx->wait[ctx->wcur].p;   // zero: isolation is reported, not differentiated
        const DevState seg_out = seg + 1 < nck ? state_at(A.ckpt[seg + 1].p, n) : state_at(A.final_z.p, n);
        for (int j = len - 1; j >= 0; --j) {
            const DevState h = state_at(A.seg_h.p + (size_t)j * 5 * n, n);
            double* gj = A.seg_g.p + (size_t)j * 8 * cap;
            const DevGrad Gj{gj, gj + 4 * (size_t)cap};
            ZFV_HIP(adj_step(ctx->prm.limiter, M, seg_state(j), j + 1 < len ? seg_state(j + 1) : seg_out, h, Gj, wait,
                             eos, A.seg_dt.p + j, ctx->prm.cfl, ctx->dmax_dt.p, state_at(A.lam[cur].p, n),
                             state_at(A.lam[1 - cur].p, n), state_at(A.lz.p, n), A.W.p, lG, A.ZB.p, A.NG.p,
                             A.dtb.p, ctx->err.p, s));
            cur = 1 - cur;
        }
    }
    const DevState L = state_at(A.lam[cur].p, n);
    double* const out[5] = {lr, lu, lv, lp, le};
    const double* const src[5] = {L.r, L.u, L.v, L.p, L.e};
    for (int k = 0; k < 5; ++k) ZFV_HIP(hipMemcpyAsync(out[k], src[k], n * sizeof(double), hipMemcpyDeviceToHost, s));
    // leave the solver at the forward run's final state
    r = copy_state(ctx, ctx->layer(ctx->cur), state_at(A.final_z.p, n));
    if (r) return r;
    ctx->dt_valid = false;
    ZFV_HIP(hipStreamSynchronize(s));
    int e = 0;
    ZFV_HIP(hipMemcpy(&e, ctx->err.p, sizeof(int), hipMemcpyDeviceToHost));
    if (e & (kErrRobust | kErrIsolated))
        return fail(ZFV_ERR_UNSUPPORTED, "adjoint: a cell was isolated (bad state) during the run");
    return read_err(ctx);
}

// ---- multi-material solver ----------------------------------------------------
static MmLayer mm_rows(zfv_ctx* ctx, double* p) {
    const size_t cap = (size_t)ctx->cap;
    const int nm = ctx->mm.mix.n;
    return MmLayer{p, p + cap, p + 2 * cap, p + 3 * cap, p + 4 * cap, p + 5 * cap, p + 6 * cap, p + 7 * cap,
                   p + (7 + (size_t)nm) * cap};
}

static MmLayer mm_layer(zfv_ctx* ctx, int i) { return mm_rows(ctx, ctx->mm.st[i].p); }

static int mm_ready(zfv_ctx* ctx, bool state) {
    int r = need_ready(ctx, false, false);
    if (r) return r;
    if (!ctx->mm.configured) return fail(ZFV_ERR_STATE, "multi-material solver not configured");
    if (ctx->nseg != 1) return fail(ZFV_ERR_UNSUPPORTED, "multi-material solver on a batch");
    if (state && (!ctx->mm.have_state || ctx->mm.epoch != ctx->mesh_epoch))
        return fail(ZFV_ERR_STATE, "no multi-material state for the current mesh");
    return ZFV_OK;
}

static int mm_alloc(zfv_ctx* ctx) {
    auto& S = ctx->mm;
    const size_t rows = 7 + 2 * (size_t)S.mix.n;
    for (auto& b : S.st) ZFV_HIP(b.alloc(rows * (size_t)ctx->cap));
    ZFV_HIP(S.dtmin.alloc(4));   // three rotating slots and the sink of the non-final sweeps' candidates
    ZFV_HIP(S.dt_out.alloc(1));
    ZFV_HIP(S.dmax_dt.alloc(1));
    unsigned long long bits;
    std::memcpy(&bits, &S.max_dt, 8);
    ZFV_HIP(launch_fill_u64(reinterpret_cast<unsigned long long*>(S.dmax_dt.p), 1, bits, ctx->stream));
    return ZFV_OK;
}

int zfv_mm_configure(zfv_ctx* ctx, int n_mat, const double* mats, double cfl) {
    int r = need_ready(ctx, false, false);
    if (r) return r;
    if (n_mat < 1 || n_mat > mm::kMaxMat || !mats) return fail(ZFV_ERR_ARG, "mm_configure: 1..4 materials");
    auto& S = ctx->mm;
    S.mix = mm::Mix{};
    S.mix.n = n_mat;
    for (int i = 0; i < n_mat; ++i) {
        const double* m = mats + 6 * i;
        if (!(m[0] > 1.0) || !(m[3] > 0.0) || !(m[5] > 0.0))
            return fail(ZFV_ERR_ARG, "mm_configure: material needs gamma > 1, Cv > 0, reference density > 0");
        S.mix.m[i] = mm::Mat{m[0], m[1], m[2], m[3], m[4], m[5]};
    }
    S.cfl = std::max(0.0, std::min(cfl, 1.0));   // set_CFL (mm_fluid.cpp:41-43)
    S.have_state = false;
    S.configured = true;
    return mm_alloc(ctx);
}

int zfv_mm_set_max_dt(zfv_ctx* ctx, double max_dt) {
    int r = mm_ready(ctx, false);
    if (r) return r;
    ctx->mm.max_dt = max_dt;
    unsigned long long bits;
    std::memcpy(&bits, &max_dt, 8);
    ZFV_HIP(launch_fill_u64(reinterpret_cast<unsigned long long*>(ctx->mm.dmax_dt.p), 1, bits, ctx->stream));
    return ZFV_OK;
}

int zfv_mm_set_crp(zfv_ctx* ctx, int mode) {
    int r = mm_ready(ctx, false);
    if (r) return r;
    if (mode != 0 && mode != 1) return fail(ZFV_ERR_ARG, "mm_set_crp: 0 (NONE) or 1 (PLIC)");
    ctx->mm.crp = mode ? ZFV_CRP_PLIC : ZFV_CRP_NONE;
    ctx->mm.flux = ZFV_HLLC;
    return ZFV_OK;
}

int zfv_mm_set_method(zfv_ctx* ctx, int flux) {
    int r = mm_ready(ctx, false);
    if (r) return r;
    if (flux == ZFV_CRP) {
        ctx->mm.crp = ZFV_CRP_PLIC;
        ctx->mm.flux = ZFV_HLLC;
    } else if (flux == ZFV_HLLC || flux == ZFV_HLL || flux == ZFV_RUSANOV) {
        ctx->mm.crp = ZFV_CRP_NONE;
        ctx->mm.flux = flux;
    } else if (flux == ZFV_GODUNOV || (flux > ZFV_HLLC && flux < ZFV_CRP)) {
        return fail(ZFV_ERR_UNSUPPORTED, "mm_set_method: only CRP, HLLC, HLL and Rusanov are built");
    } else {
        return fail(ZFV_ERR_ARG, "mm_set_method: unknown flux");
    }
    return ZFV_OK;
}

int zfv_mm_set_crp_mode(zfv_ctx* ctx, int mode) {
    int r = mm_ready(ctx, false);
    if (r) return r;
    if (mode == ZFV_CRP_MUSCL)
        return fail(ZFV_ERR_UNSUPPORTED, "mm_set_crp_mode: MUSCL has no face fraction in the reference");
    if (mode < ZFV_CRP_NONE || mode > ZFV_CRP_PLIC) return fail(ZFV_ERR_ARG, "mm_set_crp_mode: unknown CrpMode");
    if (mode != ZFV_CRP_NONE) {   // set_crp_mode ignores NONE (mm_fluid.cpp:63-70)
        ctx->mm.crp = mode;
        ctx->mm.flux = ZFV_HLLC;
    }
    return ZFV_OK;
}

int zfv_mm_set_splitting(zfv_ctx* ctx, int split) {
    int r = mm_ready(ctx, false);
    if (r) return r;
    if (split < ZFV_SPLIT_NONE || split > ZFV_SPLIT_STRANG) return fail(ZFV_ERR_ARG, "mm_set_splitting: 0, 1 or 2");
    ctx->mm.split = split;
    return ZFV_OK;
}

int zfv_mm_set_overlap(zfv_ctx* ctx, int mode, int crp_groups) {
    int r = mm_ready(ctx, false);
    if (r) return r;
    if (mode < 0 || mode > 2) return fail(ZFV_ERR_ARG, "mm_set_overlap: mode 0, 1 or 2");
    if (crp_groups < 0 || crp_groups > 65536) return fail(ZFV_ERR_ARG, "mm_set_overlap: crp_groups in 0..65536");
    ZFV_HIP(hipStreamSynchronize(ctx->stream));
    ctx->mm.overlap = mode;
    ctx->mm.crp_groups = crp_groups;
    return ZFV_OK;
}

int zfv_mm_set_state(zfv_ctx* ctx, const double* rho, const double* u, const double* v, const double* P,
                     const double* beta) {
    int r = mm_ready(ctx, false);
    if (r) return r;
    if (!rho || !u || !v || !P || !beta) return fail(ZFV_ERR_ARG, "mm_set_state: null array");
    auto& S = ctx->mm;
    if (S.st[0].n != (7 + 2 * (size_t)S.mix.n) * (size_t)ctx->cap) {
        r = mm_alloc(ctx);   // the mesh grew since configure
        if (r) return r;
    }
    const int n = ctx->hm.n;
    const size_t b = (size_t)n * sizeof(double);
    const MmLayer L = mm_layer(ctx, S.cur);
    hipStream_t s = ctx->stream;
    ZFV_HIP(hipMemcpyAsync(L.r, rho, b, hipMemcpyHostToDevice, s));
    ZFV_HIP(hipMemcpyAsync(L.u, u, b, hipMemcpyHostToDevice, s));
    ZFV_HIP(hipMemcpyAsync(L.v, v, b, hipMemcpyHostToDevice, s));
    ZFV_HIP(hipMemcpyAsync(L.P, P, b, hipMemcpyHostToDevice, s));
    for (int i = 0; i < S.mix.n; ++i)
        ZFV_HIP(hipMemcpyAsync(L.b + (size_t)i * ctx->cap, beta + (size_t)i * n, b, hipMemcpyHostToDevice, s));
    ZFV_HIP(launch_fill_u64(S.dtmin.p, 3, 0x7ff0000000000000ULL, s));
    ZFV_HIP(mm_prepare(ctx->dmesh(), L, S.mix, S.dtmin.p + (S.step_no % 3), s));
    ZFV_HIP(hipStreamSynchronize(s));
    S.have_state = true;
    S.epoch = ctx->mesh_epoch;
    return read_err(ctx);
}

int zfv_mm_state(zfv_ctx* ctx, double* rho, double* u, double* v, double* P, double* e, double* T, double* beta,
                 double* rhos) {
    int r = mm_ready(ctx, true);
    if (r) return r;
    if (!rho || !u || !v || !P || !e || !T || !beta || !rhos) return fail(ZFV_ERR_ARG, "mm_state: null array");
    auto& S = ctx->mm;
    const int n = ctx->hm.n;
    const size_t b = (size_t)n * sizeof(double);
    const MmLayer L = mm_layer(ctx, S.cur);
    hipStream_t s = ctx->stream;
    double* const dst[6] = {rho, u, v, P, e, T};
    const double* const src[6] = {L.r, L.u, L.v, L.P, L.e, L.T};
    for (int k = 0; k < 6; ++k) ZFV_HIP(hipMemcpyAsync(dst[k], src[k], b, hipMemcpyDeviceToHost, s));
    for (int i = 0; i < S.mix.n; ++i) {
        ZFV_HIP(hipMemcpyAsync(beta + (size_t)i * n, L.b + (size_t)i * ctx->cap, b, hipMemcpyDeviceToHost, s));
        ZFV_HIP(hipMemcpyAsync(rhos + (size_t)i * n, L.rh + (size_t)i * ctx->cap, b, hipMemcpyDeviceToHost, s));
    }
    ZFV_HIP(hipStreamSynchronize(s));
    return read_err(ctx);
}

static int mm_grad_alloc(zfv_ctx* ctx) {
    auto& S = ctx->mm;
    const size_t rows = 6 + 2 * (size_t)S.mix.n, cap = (size_t)ctx->cap;
    for (auto* b : {&S.gx, &S.gy})
        if (b->n != rows * cap) ZFV_HIP(b->alloc(rows * cap));
    return ZFV_OK;
}

int zfv_mm_step(zfv_ctx* ctx, int n_steps) {
    int r = mm_ready(ctx, true);
    if (r) return r;
    if (n_steps < 0) return fail(ZFV_ERR_ARG, "negative step count");
    auto& S = ctx->mm;
    const DevMesh M = ctx->dmesh();
    const size_t cap = (size_t)ctx->cap, nm = (size_t)S.mix.n;
    const int acc = S.acc;
    if (acc == 2) {
        r = mm_grad_alloc(ctx);
        if (r) return r;
        const size_t rows = 7 + 2 * nm;
        if (S.half.n != rows * cap) ZFV_HIP(S.half.alloc(rows * cap));
    }
    if (S.crp) {
        if (S.inter.n != 3 * nm * cap) ZFV_HIP(S.inter.alloc(3 * nm * cap));
        if (S.cf.n != 8 * (4 + nm) * cap) ZFV_HIP(S.cf.alloc(8 * (4 + nm) * cap));
        if (S.crp_list.n != 8 * cap) ZFV_HIP(S.crp_list.alloc(8 * cap));
        if (S.crp_count.n != 2) ZFV_HIP(S.crp_count.alloc(2));
        if (S.overlap) {
            if (S.overlap == 2 && S.crp_cells.n != cap) ZFV_HIP(S.crp_cells.alloc(cap));
            if (!S.side) ZFV_HIP(hipStreamCreateWithFlags(&S.side, hipStreamNonBlocking));
            for (hipEvent_t* e : {&S.ev_in, &S.ev_mark, &S.ev_crp})
                if (!*e) ZFV_HIP(hipEventCreateWithFlags(e, hipEventDisableTiming));
        }
    }
    // 0: one stream; 1: the interface on the side stream (it reads only
    // `in`) beside the gradient, predictor and marking; 2: also the
    // composite fluxes there, beside the update of the cells without a
    // composite face, and the other cells after them
    const int mode = S.crp ? S.overlap : 0;
    const int groups = S.crp_groups > 0 ? S.crp_groups : 2048;
    // MmFluid::update (mm_fluid.cpp:112-137): the sweeps of one step
    struct Sweep { int dir; double scale; };
    const Sweep none[] = {{0, 1.0}}, simple[] = {{1, 1.0}, {2, 1.0}}, strang[] = {{1, 0.5}, {2, 1.0}, {1, 0.5}};
    const Sweep* sweeps = S.split == ZFV_SPLIT_SIMPLE ? simple : S.split == ZFV_SPLIT_STRANG ? strang : none;
    const int n_sweeps = S.split == ZFV_SPLIT_SIMPLE ? 2 : S.split == ZFV_SPLIT_STRANG ? 3 : 1;
    const bool plic = S.crp == ZFV_CRP_PLIC;
    for (int k = 0; k < n_steps; ++k) {
        for (int j = 0; j < n_sweeps; ++j) {
            StepCtl c;
            c.dtmin = S.dtmin.p;
            c.rd = (int)(S.step_no % 3);
            // only the last sweep's state is the next step's: earlier sweeps
            // send their CFL candidates to the unread fourth slot
            c.wr = j == n_sweeps - 1 ? (int)((S.step_no + 1) % 3) : 3;
            c.clr = (int)((S.step_no + 2) % 3);
            c.cfl = S.cfl;
            c.max_dt = S.dmax_dt.p;
            c.dt_out = S.dt_out.p;
            c.err = ctx->err.p;
            MmOpts o;
            o.flux = S.flux;
            o.crp = S.crp;
            o.dir = sweeps[j].dir;
            o.dt_scale = sweeps[j].scale;
            const MmLayer in = mm_layer(ctx, S.cur), out = mm_layer(ctx, 1 - S.cur);
            const MmLayer half = acc == 2 ? mm_rows(ctx, S.half.p) : in;
            const double* CF = nullptr;
            double* ir = S.inter.p;
            hipStream_t s1 = ctx->stream, s2 = mode ? S.side : ctx->stream;
            // Every cell is updated by the same code whatever the mode.
            if (mode) {
                ZFV_HIP(hipEventRecord(S.ev_in, s1));   // `in` is complete
                ZFV_HIP(hipStreamWaitEvent(s2, S.ev_in, 0));
            }
            if (plic) {   // interface_recovery (mm_fluid.cpp:143-146, 160-163)
                ZFV_HIP(timed_on(ctx, s2, "mm_interface", [&] {
                    return mm_interface(M, in, S.mix, ctx->vx.p, ctx->vy.p, ir, ir + nm * cap, ir + 2 * nm * cap, s2);
                }));
            }
            if (acc == 2) {   // compute_grad of init (mm_fluid.cpp:154)
                ZFV_HIP(timed(ctx, "mm_grad", [&] {
                    return mm_grad(M, in, S.mix, S.crp, S.limiter, 6 + 2 * (int)nm, S.gx.p, S.gy.p, s1);
                }));
                // fluxes_stage1 (:165)
                ZFV_HIP(timed(ctx, "mm_stage1", [&] {
                    return mm_stage1(M, in, half, S.mix, c, o, S.gx.p, S.gy.p, s1);
                }));
            }
            if (S.crp) {   // the composite fluxes of the sweep's faces
                ZFV_HIP(timed(ctx, "mm_crp_mark", [&] {
                    return mm_crp_mark(M, in, half, S.mix, acc, o, S.gx.p, S.gy.p, S.crp_list.p,
                                       mode == 2 ? S.crp_cells.p : nullptr, S.crp_count.p, s1);
                }));
                // the composite fluxes run where the interface ran (s2),
                // after the marking; mode 1 joins before them
                hipStream_t sc = mode == 2 ? s2 : s1;
                if (mode == 2) {   // s2 waits for the list
                    ZFV_HIP(hipEventRecord(S.ev_mark, s1));
                    ZFV_HIP(hipStreamWaitEvent(s2, S.ev_mark, 0));
                } else if (mode == 1) {   // s1 waits for the interface
                    ZFV_HIP(hipEventRecord(S.ev_crp, s2));
                    ZFV_HIP(hipStreamWaitEvent(s1, S.ev_crp, 0));
                }
                ZFV_HIP(timed_on(ctx, sc, "mm_crp", [&] {
                    return mm_crp(M, in, half, S.mix, acc, o, S.gx.p, S.gy.p, c, ir, ir + nm * cap, S.crp_list.p,
                                  S.crp_count.p, S.cf.p, groups, sc);
                }));
                if (mode == 2) ZFV_HIP(hipEventRecord(S.ev_crp, s2));
                CF = S.cf.p;
            }
            // fluxes_stage2 (:167) or fluxes (:148): every cell, or (mode
            // 2) the cells without a composite face, then, after the join,
            // the others
            const char* name = acc == 2 ? "mm_flux2" : "mm_flux1";
            auto update = [&](int pass) {
                return timed(ctx, pass == 2 ? (acc == 2 ? "mm_flux2_list" : "mm_flux1_list") : name, [&] {
                    return acc == 2 ? mm_flux2(M, in, half, out, S.mix, c, o, S.gx.p, S.gy.p, CF, pass,
                                               S.crp_cells.p, S.crp_count.p, s1)
                                    : mm_flux1(M, in, out, S.mix, c, o, CF, pass, S.crp_cells.p, S.crp_count.p, s1);
                });
            };
            ZFV_HIP(update(mode == 2 ? 1 : 0));
            if (mode == 2) {
                ZFV_HIP(hipStreamWaitEvent(s1, S.ev_crp, 0));
                ZFV_HIP(update(2));
            }
            S.cur = 1 - S.cur;   // MmFluid::swap, after every integrate
        }
        S.step_no += 1;
    }
    // as the reference's d_dx / d_dy, the gradients stay those of the last
    // compute_grad: of the last step's old layer with acc 2
    if (acc == 2 && n_steps > 0) {
        S.grad_epoch = ctx->mesh_epoch;
        S.grad_rows = 6 + 2 * (int)nm;
    }
    return ZFV_OK;
}

int zfv_mm_crp_faces(zfv_ctx* ctx, int64_t* n) {
    int r = mm_ready(ctx, true);
    if (r) return r;
    if (!n) return fail(ZFV_ERR_ARG, "null output");
    unsigned int k = 0;
    if (ctx->mm.crp && ctx->mm.crp_count.p) {
        ZFV_HIP(hipMemcpyAsync(&k, ctx->mm.crp_count.p, sizeof(k), hipMemcpyDeviceToHost, ctx->stream));
        ZFV_HIP(hipStreamSynchronize(ctx->stream));
    }
    *n = k;
    return ZFV_OK;
}

int zfv_mm_dt(zfv_ctx* ctx, double* dt) {
    int r = mm_ready(ctx, true);
    if (r) return r;
    if (!dt) return fail(ZFV_ERR_ARG, "null output");
    ZFV_HIP(hipMemcpyAsync(dt, ctx->mm.dt_out.p, sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    ZFV_HIP(hipStreamSynchronize(ctx->stream));
    return read_err(ctx);
}

// ---- Transfer (math/solver/transfer.h:21-153) -------------------------------
static int tr_ready(zfv_ctx* ctx, bool state) {
    int r = check_ctx(ctx);
    if (r) return r;
    if (!ctx->have_mesh) return fail(ZFV_ERR_STATE, "no mesh uploaded");
    auto& T = ctx->tr;
    if (!T.configured) return fail(ZFV_ERR_STATE, "transfer: not configured");
    if (ctx->nseg != 1) return fail(ZFV_ERR_UNSUPPORTED, "transfer on a batch of problems");
    if (state && (!T.have_state || T.epoch != ctx->mesh_epoch))
        return fail(ZFV_ERR_STATE, "transfer: no state on the current mesh");
    return ZFV_OK;
}

int zfv_transfer_configure(zfv_ctx* ctx, int method, double cfl) {
    int r = check_ctx(ctx);
    if (r) return r;
    if (!ctx->have_mesh) return fail(ZFV_ERR_STATE, "no mesh uploaded");
    if (method < 0 || method > 14) return fail(ZFV_ERR_ARG, "transfer: unknown method");
    if (!tr_method_built(method))
        return fail(ZFV_ERR_UNSUPPORTED, "transfer: CRP_N2, VOF, VOF_CRP, MUSCLn, MUSCLn_CRP not built");
    auto& T = ctx->tr;
    T.method = method;
    T.cfl = std::max(0.0, std::min(cfl, 1.0));   // set_CFL
    T.configured = true;
    return ZFV_OK;
}

int zfv_transfer_set_velocity(zfv_ctx* ctx, const double* fvx, const double* fvy, const double* cvx,
                              const double* cvy) {
    int r = tr_ready(ctx, false);
    if (r) return r;
    if (!fvx || !fvy || !cvx || !cvy) return fail(ZFV_ERR_ARG, "transfer: null velocity array");
    r = to_mesh_order(ctx);
    if (r) return r;
    auto& T = ctx->tr;
    const int n = ctx->hm.n;
    const size_t cap = (size_t)ctx->cap;
    std::vector<double> sx(8 * cap, 0.0), sy(8 * cap, 0.0);
    for (int c = 0; c < n; ++c)
        for (int k = 0; k < 8; ++k) {   // host layout 8c + k to slot-major
            sx[(size_t)k * cap + c] = fvx[8 * (size_t)c + k];
            sy[(size_t)k * cap + c] = fvy[8 * (size_t)c + k];
        }
    for (auto* b : {&T.fvx, &T.fvy}) ZFV_HIP(b->alloc(8 * cap));
    for (auto* b : {&T.cvx, &T.cvy}) ZFV_HIP(b->alloc(cap));
    hipStream_t s = ctx->stream;
    ZFV_HIP(hipMemcpyAsync(T.fvx.p, sx.data(), 8 * cap * sizeof(double), hipMemcpyHostToDevice, s));
    ZFV_HIP(hipMemcpyAsync(T.fvy.p, sy.data(), 8 * cap * sizeof(double), hipMemcpyHostToDevice, s));
    ZFV_HIP(hipMemcpyAsync(T.cvx.p, cvx, n * sizeof(double), hipMemcpyHostToDevice, s));
    ZFV_HIP(hipMemcpyAsync(T.cvy.p, cvy, n * sizeof(double), hipMemcpyHostToDevice, s));
    ZFV_HIP(hipStreamSynchronize(s));
    T.have_velocity = true;
    if (T.epoch != ctx->mesh_epoch) T.have_state = false;
    T.epoch = ctx->mesh_epoch;
    return ZFV_OK;
}

// The fraction u1 (data.u1), then the interface normals (update_interface
// without smoothing, as update_CRP ends, transfer.cpp:493-504).
int zfv_transfer_set_state(zfv_ctx* ctx, const double* u) {
    int r = tr_ready(ctx, false);
    if (r) return r;
    if (!u) return fail(ZFV_ERR_ARG, "transfer: null state");
    auto& T = ctx->tr;
    if (!T.have_velocity || T.epoch != ctx->mesh_epoch) return fail(ZFV_ERR_STATE, "transfer: set the velocity first");
    const size_t cap = (size_t)ctx->cap;
    for (auto* b : {&T.u[0], &T.u[1], &T.nx, &T.ny, &T.gx, &T.gy}) ZFV_HIP(b->alloc(cap));
    if (T.slot.n < 1) ZFV_HIP(T.slot.alloc(1));
    T.cur = 0;
    hipStream_t s = ctx->stream;
    ZFV_HIP(hipMemcpyAsync(T.u[0].p, u, ctx->hm.n * sizeof(double), hipMemcpyHostToDevice, s));
    ZFV_HIP(tr_normals(ctx->dmesh(), T.u[0].p, T.nx.p, T.ny.p, s));
    ZFV_HIP(hipStreamSynchronize(s));
    T.have_state = true;
    return ZFV_OK;
}

int zfv_transfer_state(zfv_ctx* ctx, double* u, double* nx, double* ny) {
    int r = tr_ready(ctx, true);
    if (r) return r;
    auto& T = ctx->tr;
    const size_t b = (size_t)ctx->hm.n * sizeof(double);
    hipStream_t s = ctx->stream;
    if (u) ZFV_HIP(hipMemcpyAsync(u, T.u[T.cur].p, b, hipMemcpyDeviceToHost, s));
    if (nx) ZFV_HIP(hipMemcpyAsync(nx, T.nx.p, b, hipMemcpyDeviceToHost, s));
    if (ny) ZFV_HIP(hipMemcpyAsync(ny, T.ny.p, b, hipMemcpyDeviceToHost, s));
    ZFV_HIP(hipStreamSynchronize(s));
    return ZFV_OK;
}

int zfv_transfer_slopes(zfv_ctx* ctx, double* dx, double* dy) {
    int r = tr_ready(ctx, true);
    if (r) return r;
    if (!dx || !dy) return fail(ZFV_ERR_ARG, "transfer: null output");
    auto& T = ctx->tr;
    const size_t b = (size_t)ctx->hm.n * sizeof(double);
    ZFV_HIP(hipMemcpyAsync(dx, T.gx.p, b, hipMemcpyDeviceToHost, ctx->stream));
    ZFV_HIP(hipMemcpyAsync(dy, T.gy.p, b, hipMemcpyDeviceToHost, ctx->stream));
    ZFV_HIP(hipStreamSynchronize(ctx->stream));
    return ZFV_OK;
}

int zfv_transfer_set_dt(zfv_ctx* ctx, double dt) {
    int r = tr_ready(ctx, false);
    if (r) return r;
    ctx->tr.dt = dt;
    return ZFV_OK;
}

int zfv_transfer_get_dt(zfv_ctx* ctx, double* dt) {
    int r = tr_ready(ctx, false);
    if (r) return r;
    if (!dt) return fail(ZFV_ERR_ARG, "transfer: null output");
    *dt = ctx->tr.dt;
    return ZFV_OK;
}

// compute_dt (transfer.cpp:95-102): CFL times the minimum over the cells.
int zfv_transfer_compute_dt(zfv_ctx* ctx, double* dt) {
    int r = tr_ready(ctx, false);
    if (r) return r;
    if (!dt) return fail(ZFV_ERR_ARG, "transfer: null output");
    auto& T = ctx->tr;
    if (!T.have_velocity || T.epoch != ctx->mesh_epoch) return fail(ZFV_ERR_STATE, "transfer: set the velocity first");
    if (T.slot.n < 1) ZFV_HIP(T.slot.alloc(1));
    hipStream_t s = ctx->stream;
    ZFV_HIP(launch_fill_u64(T.slot.p, 1, 0x7ff0000000000000ULL, s));
    ZFV_HIP(tr_dt(ctx->dmesh(), T.cvx.p, T.cvy.p, T.slot.p, s));
    unsigned long long bits = 0;
    ZFV_HIP(hipMemcpyAsync(&bits, T.slot.p, sizeof(bits), hipMemcpyDeviceToHost, s));
    ZFV_HIP(hipStreamSynchronize(s));
    double tau;
    std::memcpy(&tau, &bits, sizeof(tau));
    if (!(tau < INFINITY)) tau = std::numeric_limits<double>::max();   // no finite candidate
    *dt = T.cfl * tau;
    return ZFV_OK;
}

// n_steps calls of Transfer::update(mesh, dir) with the dt as set
// (transfer.cpp:474-535; dir 0 ANY, 1 X, 2 Y).
int zfv_transfer_update(zfv_ctx* ctx, int dir, int n_steps) {
    int r = tr_ready(ctx, true);
    if (r) return r;
    if (dir < 0 || dir > 2) return fail(ZFV_ERR_ARG, "transfer: dir 0 (ANY), 1 (X) or 2 (Y)");
    if (n_steps < 0) return fail(ZFV_ERR_ARG, "negative step count");
    auto& T = ctx->tr;
    // WENO's stencils index the Rectangle's (i, j) grid (EuMesh::operator()(i, j))
    int gnx = 0, gny = 0;
    if (tr_method_structured(T.method)) {
        if (ctx->rects.size() != 1 || ctx->rects[0].off != 0 ||
            (long long)ctx->rects[0].nx * ctx->rects[0].ny != ctx->hm.n || ctx->rects[0].nx < 3 ||
            ctx->rects[0].ny < 3)
            return fail(ZFV_ERR_UNSUPPORTED, "transfer: WENO needs an unadapted Rectangle of at least 3 x 3 cells");
        gnx = ctx->rects[0].nx;
        gny = ctx->rects[0].ny;
    }
    r = to_mesh_order(ctx);
    if (r) return r;
    const DevMesh M = ctx->dmesh();
    hipStream_t s = ctx->stream;
    for (int k = 0; k < n_steps; ++k) {
        ZFV_HIP(timed(ctx, "tr_step", [&] {
            return tr_step(M, T.method, T.u[T.cur].p, T.u[1 - T.cur].p, T.nx.p, T.ny.p, T.gx.p, T.gy.p, T.fvx.p,
                           T.fvy.p, T.dt, dir, gnx, gny, s);
        }));
        T.cur = 1 - T.cur;
        ZFV_HIP(timed(ctx, "tr_normals", [&] { return tr_normals(M, T.u[T.cur].p, T.nx.p, T.ny.p, s); }));
    }
    return ZFV_OK;
}

// Transfer::set_flags (transfer.cpp:668-690) into the mesh's flags.
int zfv_transfer_set_flags(zfv_ctx* ctx) {
    int r = tr_ready(ctx, true);
    if (r) return r;
    r = to_mesh_order(ctx);
    if (r) return r;
    ZFV_HIP(tr_flags(ctx->dmesh(), ctx->tr.u[ctx->tr.cur].p, ctx->flag.p, ctx->stream));
    ZFV_HIP(hipStreamSynchronize(ctx->stream));
    ctx->host_stale = true;
    return ZFV_OK;
}

// MmFluid::domain (mm_fluid.cpp:686-712) after interface_recovery of the
// current state: per cell the part material idx occupies, as a polygon --
// none (alpha <= 0, or alpha < 1/2 without an interface normal), the whole
// cell (alpha >= 1, or alpha > 1/2 without a normal), else the cell clipped
// by the PLIC line through centre + p n with normal n, keeping the side
// (v - point) . n <= 1e-14 (Polygon::clip, polygon.cpp:266-292).  A
// Rectangle-born mesh is linear: a cell's polygon is its four corners
// (AmrCells::polygon, amr_cells.cpp:147-183).  Host code over the
// downloaded state and interface.
int zfv_mm_domain(zfv_ctx* ctx, int idx, int64_t* n_polys, int64_t* n_verts, int32_t* offsets, double* xs,
                  double* ys) {
    int r = mm_ready(ctx, true);
    if (r) return r;
    const int nm = ctx->mm.mix.n;
    if (idx < 0 || idx >= nm) return fail(ZFV_ERR_ARG, "mm_domain: no such material");
    r = sync_host_mesh(ctx);
    if (r) return r;
    const HostMesh& m = ctx->hm;
    const int n = m.n;
    const size_t sn = (size_t)n, sm = (size_t)nm * n;
    std::vector<double> rho(sn), u(sn), v(sn), P(sn), e(sn), T(sn), beta(sm), rhos(sm), inx(sm), iny(sm), ip(sm);
    r = zfv_mm_state(ctx, rho.data(), u.data(), v.data(), P.data(), e.data(), T.data(), beta.data(), rhos.data());
    if (r) return r;
    r = zfv_mm_interface(ctx, inx.data(), iny.data(), ip.data());
    if (r) return r;
    std::vector<int32_t> off{0};
    std::vector<double> px, py;
    double qx[8], qy[8];
    for (int c = 0; c < n; ++c) {
        const size_t k = (size_t)idx * n + c;
        double alpha = 0.0;   // PState::alpha (models.cpp:224-227)
        if (!std::isnan(rhos[k])) {
            const double a = beta[k] * rho[c] / rhos[k];
            const double lo = (1.0 < a) ? 1.0 : a;
            alpha = (0.0 < lo) ? lo : 0.0;
        }
        const double nx = inx[k], ny = iny[k];
        const bool zero_n = std::abs(nx) <= 1.0e-12 && std::abs(ny) <= 1.0e-12;   // isZero()
        if (alpha <= 0.0 || (alpha < 0.5 && zero_n)) continue;
        const double* X = &m.vx[9 * (size_t)c];
        const double* Y = &m.vy[9 * (size_t)c];
        const int ring[4] = {0, 2, 8, 6};   // vs<-1,-1>, vs<+1,-1>, vs<+1,+1>, vs<-1,+1>
        int cnt = 0;
        if (alpha >= 1.0 || (alpha > 0.5 && zero_n)) {
            for (int j : ring) {
                qx[cnt] = X[j];
                qy[cnt] = Y[j];
                ++cnt;
            }
        } else {
            const double ptx = m.cx[c] + ip[k] * nx, pty = m.cy[c] + ip[k] * ny;
            for (int i = 0; i < 4; ++i) {
                const double x1 = X[ring[i]], y1 = Y[ring[i]];
                const double x2 = X[ring[(i + 1) % 4]], y2 = Y[ring[(i + 1) % 4]];
                const double dot1 = (x1 - ptx) * nx + (y1 - pty) * ny;
                const double dot2 = (x2 - ptx) * nx + (y2 - pty) * ny;
                const double det = (x2 - x1) * nx + (y2 - y1) * ny;
                if (dot1 <= 1.0e-14) {
                    qx[cnt] = x1;
                    qy[cnt] = y1;
                    ++cnt;
                }
                if (dot1 * dot2 < 0.0) {
                    qx[cnt] = (dot2 * x1 - dot1 * x2) / det;
                    qy[cnt] = (dot2 * y1 - dot1 * y2) / det;
                    ++cnt;
                }
            }
        }
        for (int i = 0; i < cnt; ++i) {
            px.push_back(qx[i]);
            py.push_back(qy[i]);
        }
        off.push_back((int32_t)px.size());
    }
    if (n_polys) *n_polys = (int64_t)off.size() - 1;
    if (n_verts) *n_verts = (int64_t)px.size();
    if (offsets) std::memcpy(offsets, off.data(), off.size() * sizeof(int32_t));
    if (xs) std::memcpy(xs, px.data(), px.size() * sizeof(double));
    if (ys) std::memcpy(ys, py.data(), py.size() * sizeof(double));
    return ZFV_OK;
}

int zfv_mm_interface(zfv_ctx* ctx, double* nx, double* ny, double* p) {
    int r = mm_ready(ctx, true);
    if (r) return r;
    if (!nx || !ny || !p) return fail(ZFV_ERR_ARG, "mm_interface: null array");
    auto& S = ctx->mm;
    const size_t cap = (size_t)ctx->cap, nm = (size_t)S.mix.n;
    if (S.inter.n != 3 * nm * cap) ZFV_HIP(S.inter.alloc(3 * nm * cap));
    double* o = S.inter.p;
    hipStream_t s = ctx->stream;
    ZFV_HIP(mm_interface(ctx->dmesh(), mm_layer(ctx, S.cur), S.mix, ctx->vx.p, ctx->vy.p, o, o + nm * cap,
                         o + 2 * nm * cap, s));
    const int n = ctx->hm.n;
    const size_t b = (size_t)n * sizeof(double);
    for (size_t i = 0; i < nm; ++i) {
        ZFV_HIP(hipMemcpyAsync(nx + i * n, o + i * cap, b, hipMemcpyDeviceToHost, s));
        ZFV_HIP(hipMemcpyAsync(ny + i * n, o + (nm + i) * cap, b, hipMemcpyDeviceToHost, s));
        ZFV_HIP(hipMemcpyAsync(p + i * n, o + (2 * nm + i) * cap, b, hipMemcpyDeviceToHost, s));
    }
    ZFV_HIP(hipStreamSynchronize(s));
    return read_err(ctx);
}

int zfv_mm_set_accuracy(zfv_ctx* ctx, int acc) {
    int r = check_ctx(ctx);
    if (r) return r;
    ctx->mm.acc = std::min(std::max(1, acc), 2);   // set_accuracy (mm_fluid.cpp:46-48)
    return ZFV_OK;
}

int zfv_mm_set_limiter(zfv_ctx* ctx, int limiter) {
    int r = check_ctx(ctx);
    if (r) return r;
    if (limiter < 0 || limiter > 12) return fail(ZFV_ERR_ARG, "unknown limiter");
    ctx->mm.limiter = limiter;
    ctx->mm.grad_epoch = -1;
    return ZFV_OK;
}

// MmFluid::set_flags (mm_fluid.cpp:836-885): compute_grad (kept for the
// split distributor of zfv_mm_refine), then the flags.
int zfv_mm_set_flags(zfv_ctx* ctx) {
    int r = mm_ready(ctx, true);
    if (r) return r;
    auto& S = ctx->mm;
    r = mm_grad_alloc(ctx);
    if (r) return r;
    const DevMesh M = ctx->dmesh();
    const MmLayer L = mm_layer(ctx, S.cur);
    // the rows the split distributor reads: the flags read none, and
    // P, T and the true densities' rows would only feed a second-order
    // step, which computes its own
    S.grad_rows = 4 + S.mix.n;
    ZFV_HIP(mm_grad(M, L, S.mix, S.crp, S.limiter, S.grad_rows, S.gx.p, S.gy.p, ctx->stream));
    ZFV_HIP(mm_flags(M, L, S.mix, ctx->flag.p, ctx->stream));
    ctx->host_stale = true;
    S.grad_epoch = ctx->mesh_epoch;
    return ZFV_OK;
}

int zfv_mm_refine(zfv_ctx* ctx, int max_level) { return refine_impl(ctx, max_level, kRefineMm); }

int zfv_mm_grad(zfv_ctx* ctx, double* gx, double* gy) {
    int r = mm_ready(ctx, true);
    if (r) return r;
    if (!gx || !gy) return fail(ZFV_ERR_ARG, "mm_grad: null array");
    auto& S = ctx->mm;
    if (S.grad_epoch != ctx->mesh_epoch) return fail(ZFV_ERR_STATE, "no multi-material gradients for this mesh");
    const int n = ctx->hm.n, rows = S.grad_rows, all = 6 + 2 * S.mix.n;
    const size_t row = (size_t)n * sizeof(double), pitch = (size_t)ctx->cap * sizeof(double);
    ZFV_HIP(hipMemcpy2DAsync(gx, row, S.gx.p, pitch, row, rows, hipMemcpyDeviceToHost, ctx->stream));
    ZFV_HIP(hipMemcpy2DAsync(gy, row, S.gy.p, pitch, row, rows, hipMemcpyDeviceToHost, ctx->stream));
    ZFV_HIP(hipStreamSynchronize(ctx->stream));
    for (size_t k = (size_t)n * rows; k < (size_t)n * all; ++k) gx[k] = gy[k] = NAN;   // rows not computed
    return ZFV_OK;
}

int zfv_limiter_from_name(const char* orig) {
    if (!orig) return -1;
    std::string name(orig);
    for (auto& ch : name) ch = (char)std::tolower((unsigned char)ch);
    auto erase_all = [&](const std::string& pat) {
        size_t pos;
        while ((pos = name.find(pat)) != std::string::npos) name.erase(pos, pat.size());
    };
    erase_all(" ");
    erase_all("-");
    erase_all("_");
    erase_all("lim");
    erase_all("limiter");
    auto has = [&](const char* s) { return name.find(s) != std::string::npos; };
    if (name.empty() || has("off") || has("none") || has("zero") || has("empty")) return ZFV_LIM_NONE;
    if (has("one") || has("unit")) return ZFV_LIM_UNIT;
    if (has("minmod")) return ZFV_LIM_MINMOD;
    if (has("mc")) return ZFV_LIM_MC;
    if (has("leer")) return ZFV_LIM_VANLEER;
    if (has("albada")) return has("2") ? ZFV_LIM_VANALBADA2 : ZFV_LIM_VANALBADA1;
    if (has("charm")) return ZFV_LIM_CHARM;
    if (has("hcus")) return ZFV_LIM_HCUS;
    if (has("hquick")) return ZFV_LIM_HQUICK;
    if (has("koren")) return ZFV_LIM_KOREN;
    if (has("smart")) return ZFV_LIM_SMART;
    if (has("superbee")) return ZFV_LIM_SUPERBEE;
    g_last_error = std::string("Unknown limiter '") + orig + "'";
    return -1;
}

static bool valid_block(int b) { return b == 128 || b == 256 || b == 512; }

static int tile_variant_of(int threads) {
    for (int v = 0; v < kTileVariants; ++v)
        if (threads == (kTileRows[v] + 2) * 64) return v;
    return -1;
}

int zfv_set_launch(zfv_ctx* ctx, int stage, int layout, int block) {
    int r = check_ctx(ctx);
    if (r) return r;
    if (stage < 0 || stage > 2) return fail(ZFV_ERR_ARG, "stage must be 0 (first order), 1 (gradient), 2 (corrector)");
    if (layout == kLayoutTile) {
        if (stage != 1) return fail(ZFV_ERR_ARG, "the tiled layout is a stage-1 layout (it runs the whole step)");
        if (!ctx->have_mesh || !ctx->tiles.ok)
            return fail(ZFV_ERR_UNSUPPORTED, "the tiled layout needs a Rectangle mesh (or a batch of them)");
        const int v = block == 0 ? ctx->tile_variant : tile_variant_of(block);
        if (v < 0) return fail(ZFV_ERR_ARG, "tiled layout: block must be 0 (keep), 512, 768 or 1024");
        ctx->layout[1] = kLayoutTile;
        ctx->block[1] = (kTileRows[v] + 2) * 64;
        ctx->tile_variant = v;
        ctx->layout_set[1] = true;
        ctx->graphs_clear();
        return ZFV_OK;
    }
    if (layout < 0 || layout >= kLayouts)
        return fail(ZFV_ERR_ARG, "layout must be 0 (cell), 1 (quad), 2 (batched) or 3 (tiled, stage 1)");
    if (block != 0 && !valid_block(block)) return fail(ZFV_ERR_ARG, "block must be 0 (keep), 128, 256 or 512");
    if (stage == 1 && ctx->layout[1] == kLayoutTile && block == 0) block = kBlock;
    ctx->layout[stage] = layout;
    if (block) ctx->block[stage] = block;
    ctx->layout_set[stage] = true;
    ctx->graphs_clear();
    return ZFV_OK;
}

int zfv_set_tile_schedule(zfv_ctx* ctx, int persistent) {
    int r = check_ctx(ctx);
    if (r) return r;
    if (persistent != 0 && persistent != 1) return fail(ZFV_ERR_ARG, "tile schedule must be 0 or 1");
    ctx->tile_persistent = persistent != 0;
    ctx->graphs_clear();
    return ZFV_OK;
}

int zfv_get_tile_schedule(zfv_ctx* ctx, int* persistent) {
    int r = check_ctx(ctx);
    if (r) return r;
    if (!persistent) return fail(ZFV_ERR_ARG, "null output");
    *persistent = ctx->tile_persistent ? 1 : 0;
    return ZFV_OK;
}

int zfv_get_launch(zfv_ctx* ctx, int stage, int* layout, int* block) {
    int r = check_ctx(ctx);
    if (r) return r;
    if (stage < 0 || stage > 2) return fail(ZFV_ERR_ARG, "bad stage");
    if (layout) *layout = ctx->layout[stage];
    if (block) *block = ctx->block[stage];
    return ZFV_OK;
}

// Launch autotuner: times every (layout, workgroup size) of the configured
// scheme's step kernels on the current state (reps launches each) and keeps
// the fastest per stage.  The timed launches recompute the same outputs as
// the next step's first launches (same inputs, same dt slot values), so the
// state is not advanced.
int zfv_autotune(zfv_ctx* ctx, int reps, double* ms_out) {
    int r = need_ready(ctx, true, true);
    if (r) return r;
    if (reps <= 0) return fail(ZFV_ERR_ARG, "reps must be positive");
    r = to_curve_order(ctx);   // tune on the order the steps will run in
    if (r) return r;
    r = ensure_dt(ctx);
    if (r) return r;
    hipStream_t s = ctx->stream;
    const DevMesh M = ctx->smesh();
    const Eos eos = ctx->eos();
    const zfv_smf_params& P = ctx->prm;
    const StepCtl ctl = make_ctl(ctx);
    const int nxt = (ctx->cur + 1) % 2;
    const DevState in = ctx->layer(ctx->cur), out = ctx->layer(nxt), half = ctx->layer(2);
    const DevGrad G{ctx->gdx.p, ctx->gdy.p};
    int* w_in = ctx->wait[ctx->wcur].p;
    int* w_out = ctx->wait[1 - ctx->wcur].p;
    hipEvent_t e0 = pooled_event(ctx), e1 = pooled_event(ctx);
    auto time_it = [&](auto&& launch, double& ms) -> int {
        ZFV_HIP(launch());   // warm
        ZFV_HIP(hipEventRecord(e0, s));
        for (int i = 0; i < reps; ++i) ZFV_HIP(launch());
        ZFV_HIP(hipEventRecord(e1, s));
        ZFV_HIP(hipEventSynchronize(e1));
        float f = 0.0f;
        ZFV_HIP(hipEventElapsedTime(&f, e0, e1));
        ms = f / reps;
        return ZFV_OK;
    };
    const int blocks[3] = {128, 256, 512};
    double t[3][kLayouts + 1][3] = {};   // [stage][layout][block]; [1][kLayoutTile][v]: the tiled step, shape v
    const int stages[2] = {P.acc == 1 ? 0 : 1, P.acc == 1 ? -1 : 2};
    for (int st : stages) {
        if (st < 0) continue;
        for (int L = 0; L < kLayouts; ++L)
            for (int bi = 0; bi < 3; ++bi) {
                const int B = blocks[bi];
                if (st == 0 && L == kLayoutBatch) {
                    // the first-order kernel is the same in both cell layouts
                    t[st][L][bi] = t[st][kLayoutCell][bi];
                    continue;
                }
                if (st == 0)
                    r = time_it([&] { return launch_flux1(P.flux, L, B, M, in, out, eos, ctl, s); }, t[st][L][bi]);
                else if (st == 1)
                    r = time_it([&] { return launch_grad_pred(P.limiter, true, L, B, M, in, G, half, w_in, eos, ctl, s); },
                                t[st][L][bi]);
                else
                    r = time_it([&] { return launch_corr(P.flux, L, B, M, in, half, G, w_in, w_out, out, eos, ctl, s); },
                                t[st][L][bi]);
                if (r) return r;
            }
        int bl = 0, bb = 0;
        for (int L = 0; L < kLayouts; ++L)
            for (int bi = 0; bi < 3; ++bi)
                if (t[st][L][bi] < t[st][bl][bb]) {
                    bl = L;
                    bb = bi;
                }
        ctx->layout[st] = bl;
        ctx->block[st] = blocks[bb];
        ctx->layout_set[st] = true;
    }
    // the tiled step (one kernel for the whole step) against the best pair
    // of gradient and corrector kernels
    if (P.acc == 2 && P.flux == ZFV_HLLC && ctx->tiles.ok && !ctx->co.state) {
        // [1][tile][v]: one workgroup per tile; [2][tile][v]: persistent workgroups
        const TileMesh T = ctx->tmesh();
        int bv = 0, bp = 0;
        for (int pers = 0; pers < 2; ++pers)
            for (int v = 0; v < kTileVariants; ++v) {
                r = time_it([&] { return launch_tile_step(P.limiter, v, pers != 0, T, in, out, w_in, w_out, eos, ctl, s); },
                            t[1 + pers][kLayoutTile][v]);
                if (r) return r;
                if (t[1 + pers][kLayoutTile][v] < t[1 + bp][kLayoutTile][bv]) {
                    bv = v;
                    bp = pers;
                }
            }
        const double split = t[1][ctx->layout[1]][ctx->block[1] == 128 ? 0 : (ctx->block[1] == 256 ? 1 : 2)] +
                             t[2][ctx->layout[2]][ctx->block[2] == 128 ? 0 : (ctx->block[2] == 256 ? 1 : 2)];
        if (t[1 + bp][kLayoutTile][bv] < split) {
            ctx->layout[1] = kLayoutTile;
            ctx->block[1] = (kTileRows[bv] + 2) * 64;
            ctx->tile_variant = bv;
            ctx->tile_persistent = bp != 0;
        }
    }
    ctx->event_pool.push_back(e0);
    ctx->event_pool.push_back(e1);
    if (ms_out)
        for (int st = 0; st < 3; ++st)
            for (int L = 0; L <= kLayouts; ++L)
                for (int bi = 0; bi < 3; ++bi) ms_out[(st * (kLayouts + 1) + L) * 3 + bi] = t[st][L][bi];
    ctx->co.grad = ctx->co.state;
    ctx->graphs_clear();
    return read_err(ctx);
}

int zfv_profile_enable(zfv_ctx* ctx, int on) {
    int r = check_ctx(ctx);
    if (r) return r;
    ZFV_HIP(hipStreamSynchronize(ctx->stream));
    collect_timers(ctx);
    ctx->prof.clear();
    ctx->profile = on != 0;
    return ZFV_OK;
}

int zfv_profile_read(zfv_ctx* ctx, const char** names, double* ms, int32_t* launches, int32_t* count) {
    int r = check_ctx(ctx);
    if (r) return r;
    if (!count) return fail(ZFV_ERR_ARG, "null count");
    ZFV_HIP(hipStreamSynchronize(ctx->stream));
    collect_timers(ctx);
    int cap = *count, i = 0;
    for (auto& kv : ctx->prof) {
        if (i < cap) {
            if (names) names[i] = kv.first.c_str();
            if (ms) ms[i] = kv.second.first;
            if (launches) launches[i] = kv.second.second;
        }
        ++i;
    }
    *count = std::min(i, cap);
    return ZFV_OK;
}

}  // extern "C"
