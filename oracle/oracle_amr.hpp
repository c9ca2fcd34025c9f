// ORACLE — test infrastructure only.  Included at the end of oracle.cpp (one
// translation unit).  CPU restatement of zephyr's AMR-2D adaptation:
//   * SmFluid::set_flags (math/solver/sm_fluid.cpp:517-643) and the
//     normalised density-gradient criterion of the AMR workload;
//   * amr::balance_flags_fast<2> (mesh/amr/balancing_fast.h:412-451) with
//     base_restriction (balancing_restrictions.h:21-47), can_coarse /
//     get_siblings (siblings.h:37-131) and rounds 1-4 (balancing_fast.h:273-390);
//   * amr::apply_impl<2> (mesh/amr/apply.h:43-94): Statistics
//     (statistics.h:59-81), SwapLists (swap_lists.h:28-138), setup_positions
//     (setup_positions.h:24-64), retain_cell (retain.h:17-216), refine_cell /
//     make_children / link_siblings (refine.h:36-444), coarse_cell /
//     make_parent (coarse.h:95-404), split_face / merge_faces (faces.h:13-228),
//     AmrCells::set_cell (mesh/euler/amr_cells.cpp:725-785), move_item
//     (amr_cells.cpp:428-468);
//   * SmFluid::distributor "const" / "slope" and merge (sm_fluid.cpp:372-470).
// Single process, Cartesian base: every adjacency rotation is 0.
// The rounds of the balancing run in place, serially; the reference runs
// them in parallel in place, and their outcome does not depend on the order
// (a same-level update missed in round 2 is caught by round 3; round 4 only
// reads flags that no round-4 update changes).

namespace {

// AMR-2D index tables (geom/indexing.h).
constexpr int kSF[4][2] = {{0, 6}, {2, 8}, {0, 2}, {6, 8}};                 // amr::face_nodes2d
constexpr int kCF[8][2] = {{0, 3}, {2, 5}, {0, 1}, {6, 7},
                           {3, 6}, {5, 8}, {1, 2}, {7, 8}};                  // amr::subface_nodes2d
constexpr int kChildBySubface[8] = {0, 1, 0, 2, 2, 3, 1, 3};                 // child_by_subface_2D
constexpr int kChildrenBySide[4][2] = {{0, 2}, {1, 3}, {0, 1}, {2, 3}};     // children_by_face_2D
constexpr int kSubfaceBySideChild[4][4] = {{0, -1, 4, -1}, {-1, 1, -1, 5},
                                           {2, 6, -1, -1}, {-1, -1, 3, 7}};  // subface_by_side_child_2D
constexpr int kNeibChild0[8] = {1, 0, 2, 0, 3, 2, 3, 1};                    // neib_child_by_symm_subface_2D[0]
constexpr int kNextSiblingSide[4] = {1, 3, 2, 0};                           // side_to_next_sibling<2>

inline int adjacent_child(int side, int child) { return kNeibChild0[kSubfaceBySideChild[side][child]]; }

inline Vec3 cross(const Vec3& a, const Vec3& b) {
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}

// Line::length / centroid / normal(view) (geom/primitives/line.cpp:18-71):
// normal = ((v1-v0) x (v0-view)) x (v1-v0), normalised.
inline void line_geom(const Vec3& v0, const Vec3& v1, const Vec3& view,
                      double& area, double& fx, double& fy, double& nx, double& ny) {
    Vec3 tau = v1 - v0;
    area = norm(tau);
    fx = 0.5 * (v0.x + v1.x);
    fy = 0.5 * (v0.y + v1.y);
    Vec3 n = normalized(cross(cross(tau, v0 - view), tau));
    nx = n.x;
    ny = n.y;
}

// The solver's cell data through an adaptation: the Distributor's split and
// merge callbacks (mesh/euler/distributor.h:13-39) plus the storage moves
// and resizes the reference applies to every storable (amr_cells.cpp:428-468).
struct AmrData {
    virtual ~AmrData() = default;
    virtual void resize(int n) = 0;
    virtual void split(const orc_mesh& m, int ip, const int* children) = 0;
    virtual void merge(const orc_mesh& m, const int* children, int ip) = 0;
    virtual void move(int from, int to) = 0;
};

struct AmrWork {
    orc_mesh& m;
    AmrData* D;
    int n;   // cells before adaptation
};

inline Vec3 vert(const orc_mesh& m, int c, int k) {
    return {m.vx[9 * (size_t)c + k], m.vy[9 * (size_t)c + k], 0.0};
}

void resize_cells(orc_mesh& m, AmrData* D, int n) {
    m.cx.resize(n); m.cy.resize(n); m.vol.resize(n);
    m.level.resize(n, 0); m.flag.resize(n, 0); m.b_idx.resize(n, 0); m.z_idx.resize(n, 0);
    m.next.resize(n, 0); m.idx.resize(n, -1);
    m.vx.resize(9 * (size_t)n); m.vy.resize(9 * (size_t)n);
    size_t ns = 8 * (size_t)n;
    m.bnd.resize(ns, ORC_UNDEFINED); m.adj.resize(ns, -1);
    m.fnx.resize(ns, 0.0); m.fny.resize(ns, 0.0); m.fcx.resize(ns, 0.0); m.fcy.resize(ns, 0.0);
    m.area.resize(ns, 0.0);
    m.n = n;
    if (D) D->resize(n);
}

inline void set_slot_undefined(orc_mesh& m, size_t s) {   // AmrFaces::set_undefined
    m.bnd[s] = ORC_UNDEFINED;
    m.adj[s] = -1;
}

// AmrCells::set_cell(ic, Quad, axial = false) (amr_cells.cpp:725-785) for the
// quad with corners v00, v01, v10, v11.
void set_cell(orc_mesh& m, int ic, const Vec3& v00, const Vec3& v01, const Vec3& v10, const Vec3& v11) {
    m.idx[ic] = -1;
    m.flag[ic] = 0; m.b_idx[ic] = -1; m.z_idx[ic] = -1; m.level[ic] = -1;
    // Quad::area (quad.cpp:322-324)
    double area = 0.5 * norm(cross(v11 - v00, v10 - v01));
    m.vol[ic] = area;
    // Quad::centroid_oxy (quad.cpp:358-377), vertices in the order 0, 1, 3, 2
    const Vec3 q[4] = {v00, v01, v10, v11};
    const int ord[4] = {0, 1, 3, 2};
    double Cx = 0.0, Cy = 0.0;
    for (int i = 0; i < 4; ++i) {
        const Vec3& a = q[ord[i]];
        const Vec3& b = q[ord[(i + 1) % 4]];
        Cx = Cx + (b.y - a.y) * (b.x * b.x + b.x * a.x + a.x * a.x);
        Cy = Cy - (b.x - a.x) * (b.y * b.y + b.y * a.y + a.y * a.y);
    }
    double d6 = 6.0 * area;
    m.cx[ic] = Cx / d6;
    m.cy[ic] = Cy / d6;
    // SqQuad(const Quad&) (quad.cpp:452-457): bilinear vertices
    Vec3 V[9] = {v00, scale(0.5, v00 + v01), v01,
                 scale(0.5, v00 + v10), scale(0.25, ((v00 + v01) + v10) + v11), scale(0.5, v01 + v11),
                 v10, scale(0.5, v10 + v11), v11};
    for (int k = 0; k < 9; ++k) { m.vx[9 * (size_t)ic + k] = V[k].x; m.vy[9 * (size_t)ic + k] = V[k].y; }
    // AmrFaces::insert(AMR2D): all 8 slots undefined, then the 4 simple faces
    Vec3 C{m.cx[ic], m.cy[ic], 0.0};
    for (int s8 = 0; s8 < 8; ++s8) set_slot_undefined(m, 8 * (size_t)ic + s8);
    for (int side = 0; side < 4; ++side) {
        size_t f = 8 * (size_t)ic + side;
        line_geom(V[kSF[side][0]], V[kSF[side][1]], C, m.area[f], m.fcx[f], m.fcy[f], m.fnx[f], m.fny[f]);
        m.bnd[f] = ORC_INNER;
    }
}

inline bool simple_face(const orc_mesh& m, int ic, int side) {
    return m.bnd[8 * (size_t)ic + side + 4] == ORC_UNDEFINED;
}

// setup_face_geom for a slot with vertex pair (a, b), view = SqQuad centre.
inline void face_geom(orc_mesh& m, int ic, int slot, int a, int b) {
    size_t f = 8 * (size_t)ic + slot;
    Vec3 C = vert(m, ic, 4);
    line_geom(vert(m, ic, a), vert(m, ic, b), C, m.area[f], m.fcx[f], m.fcy[f], m.fnx[f], m.fny[f]);
}

// split_face_impl (faces.h:121-130): topology of side copied to side[1],
// geometry of both halves.
void split_face(orc_mesh& m, int ic, int side) {
    size_t f0 = 8 * (size_t)ic + side, f1 = f0 + 4;
    m.bnd[f1] = m.bnd[f0];
    m.adj[f1] = m.adj[f0];
    face_geom(m, ic, side, kCF[side][0], kCF[side][1]);
    face_geom(m, ic, side + 4, kCF[side + 4][0], kCF[side + 4][1]);
}

// merge_faces_impl (faces.h:137-154).
void merge_faces(orc_mesh& m, int ic, int side) {
    set_slot_undefined(m, 8 * (size_t)ic + side + 4);
    face_geom(m, ic, side, kSF[side][0], kSF[side][1]);
}

// can_coarse (siblings.h:36-85).
bool can_coarse(const orc_mesh& m, int ic) {
    if (m.flag[ic] >= 0) return false;
    for (int i = 0; i < 3; ++i) {
        int z = m.z_idx[ic] % 4;
        int jc = m.adj[8 * (size_t)ic + kNextSiblingSide[z]];
        if (m.level[ic] != m.level[jc]) return false;
        if (m.flag[jc] >= 0) return false;
        ic = jc;
    }
    return true;
}

// get_siblings (siblings.h:93-131).
void get_siblings(const orc_mesh& m, int ic, int* sibs) {
    int jc = ic;
    for (int i = 0; i < 3; ++i) {
        int z = m.z_idx[jc] % 4;
        jc = m.adj[8 * (size_t)jc + kNextSiblingSide[z]];
        sibs[i] = jc;
    }
}

// base_restriction (balancing_restrictions.h:21-47).
void base_restriction(orc_mesh& m, int ic, int max_level) {
    int flag = m.flag[ic];
    if (flag != 0) flag = flag > 0 ? 1 : -1;
    int lvl = m.level[ic];
    if (lvl + flag < 0) flag = 0;
    else if (lvl + flag > max_level) flag = lvl > max_level ? -1 : 0;
    // can_coarse reads the stored flags (this cell's is still unrestricted)
    if (flag < 0 && !can_coarse(m, ic)) flag = 0;
    m.flag[ic] = flag;
}

// retain_update_flag (balancing_fast.h:273-291).
void retain_update_flag(orc_mesh& m, int ic) {
    for (int s8 = 0; s8 < 8; ++s8) {
        size_t f = 8 * (size_t)ic + s8;
        if (m.bnd[f] == ORC_UNDEFINED || slot_is_boundary(m.bnd[f])) continue;
        int jc = m.adj[f];
        if (m.level[jc] + m.flag[jc] > m.level[ic] + 1) {
            m.flag[ic] = 1;
            return;
        }
    }
}

// coarse_update_flag (balancing_fast.h:297-323).
void coarse_update_flag(orc_mesh& m, int ic) {
    if (m.flag[ic] >= 0) return;
    for (int s8 = 0; s8 < 8; ++s8) {
        size_t f = 8 * (size_t)ic + s8;
        if (m.bnd[f] == ORC_UNDEFINED || slot_is_boundary(m.bnd[f])) continue;
        int jc = m.adj[f];
        int wanted = m.level[jc] + m.flag[jc];
        if (wanted > m.level[ic]) {
            if (wanted > m.level[ic] + 1) {
                m.flag[ic] = 1;
                return;
            }
            m.flag[ic] = 0;
        }
    }
}

// coarse_update_flag_by_sibs (balancing_fast.h:328-344).
void coarse_update_flag_by_sibs(orc_mesh& m, int ic) {
    if (m.flag[ic] >= 0) return;
    int sibs[3];
    get_siblings(m, ic, sibs);
    for (int k = 0; k < 3; ++k) {
        if (m.flag[sibs[k]] >= 0) {
            m.flag[ic] = 0;
            break;
        }
    }
}

// balance_flags_fast<2> (balancing_fast.h:412-451).
void balance_flags(orc_mesh& m, int max_level) {
    const int n = m.n;
    for (int ic = 0; ic < n; ++ic) base_restriction(m, ic, max_level);
    std::vector<std::vector<int>> coarse(std::max(max_level, 0)), retain(std::max(max_level, 0));
    for (int ic = 0; ic < n; ++ic) {
        int lvl = m.level[ic], flag = m.flag[ic];
        if (lvl < max_level && flag < 1) (flag < 0 ? coarse : retain)[lvl].push_back(ic);
    }
    for (int lvl = max_level - 1; lvl >= 0; --lvl) {
        for (int ic : retain[lvl]) retain_update_flag(m, ic);
        for (int ic : coarse[lvl]) coarse_update_flag(m, ic);
        for (int ic : coarse[lvl]) coarse_update_flag(m, ic);
        for (int ic : coarse[lvl]) coarse_update_flag_by_sibs(m, ic);
    }
}

// ---- distributor (sm_fluid.cpp:372-470) -----------------------------------
void split_data(orc_smf* S, int ip, const int* children, bool slope) {
    const orc_mesh& m = *S->m;
    const PState zp = S->init[ip];
    if (!slope) {
        for (int k = 0; k < 4; ++k) { S->init[children[k]] = zp; S->wait[children[k]] = 0; }
        return;
    }
    const PState& gx = S->d_dx[ip];
    const PState& gy = S->d_dy[ip];
    const PState& gz = S->d_dz[ip];
    // IdealGas / StiffenedGas pressure_re derivatives (stiffened_gas.cpp:103-110)
    const double dR = (S->eos.gamma - 1.0) * (zp.e - S->eos.e0);
    const double dE = (S->eos.gamma - 1.0) * zp.rho;
    const Vec3 grad_e{(gx.p - dR * gx.rho) / dE, (gy.p - dR * gy.rho) / dE, (gz.p - dR * gz.rho) / dE};
    bool bad = false;
    for (int k = 0; k < 4; ++k) {
        int ch = children[k];
        Vec3 dr = Vec3{m.cx[ch], m.cy[ch], 0.0} - Vec3{m.cx[ip], m.cy[ip], 0.0};
        PState z = zp;
        z.rho = zp.rho + gx.rho * dr.x + gy.rho * dr.y + gz.rho * dr.z;
        double ratio = zp.rho / z.rho;
        Vec3 dv{(gx.v.x * dr.x + gy.v.x * dr.y) + gz.v.x * dr.z,
                (gx.v.y * dr.x + gy.v.y * dr.y) + gz.v.y * dr.z,
                (gx.v.z * dr.x + gy.v.z * dr.y) + gz.v.z * dr.z};
        z.v = zp.v + scale(ratio, dv);
        z.e = zp.e - 0.5 * sqnorm(zp.v - z.v) + ratio * dot(grad_e, dr);
        z.p = S->eos.pressure_re(z.rho, z.e);
        if (is_bad(z, S->eos)) { bad = true; break; }
        S->init[ch] = z;
        S->wait[ch] = 0;
    }
    if (bad)
        for (int k = 0; k < 4; ++k) { S->init[children[k]] = zp; S->wait[children[k]] = 0; }
}

void merge_data(orc_smf* S, const int* children, int ip) {
    const orc_mesh& m = *S->m;
    double q[5] = {0, 0, 0, 0, 0};
    for (int k = 0; k < 4; ++k) {
        QState qc = to_qstate(S->init[children[k]]);
        double V = m.vol[children[k]];
        for (int j = 0; j < 5; ++j) q[j] = q[j] + qc.arr()[j] * V;
    }
    QState qp;
    for (int j = 0; j < 5; ++j) qp.arr()[j] = q[j] / m.vol[ip];
    S->init[ip] = to_pstate(qp, S->eos);
    S->wait[ip] = 0;
}

// SmFluid's data rows through an adaptation ("const" or "slope" split).
struct SmfData : AmrData {
    orc_smf* S;
    bool slope;
    SmfData(orc_smf* s, bool sl) : S(s), slope(sl) {}
    void resize(int n) override {
        PState zz; std::memset(&zz, 0, sizeof(zz));
        S->init.resize(n, zz); S->half.resize(n, zz); S->next.resize(n, zz);
        S->d_dx.resize(n, zz); S->d_dy.resize(n, zz); S->d_dz.resize(n, zz);
        S->wait.resize(n, 0); S->wait_next.resize(n, 0);
    }
    void split(const orc_mesh&, int ip, const int* ch) override { split_data(S, ip, ch, slope); }
    void merge(const orc_mesh&, const int* ch, int ip) override { merge_data(S, ch, ip); }
    void move(int from, int to) override {
        S->init[to] = S->init[from]; S->wait[to] = S->wait[from];
        S->d_dx[to] = S->d_dx[from]; S->d_dy[to] = S->d_dy[from]; S->d_dz[to] = S->d_dz[from];
    }
};

// ---- apply ------------------------------------------------------------------
// retain_cell (retain.h:17-216).
void retain_cell(orc_mesh& m, int ic) {
    const int lvl_c = m.level[ic];
    for (int side = 0; side < 4; ++side) {
        size_t iface = 8 * (size_t)ic + side;
        if (m.bnd[iface] == ORC_UNDEFINED) continue;
        if (slot_is_boundary(m.bnd[iface])) { m.adj[iface] = m.next[ic]; continue; }
        int jc = m.adj[iface];
        int lvl_n = m.level[jc];
        if (lvl_c == lvl_n) {
            int neib_next = m.next[jc];
            if (m.flag[jc] > 0) {
                split_face(m, ic, side);
                for (int sub : {side, side + 4})
                    m.adj[8 * (size_t)ic + sub] = m.next[neib_next + kNeibChild0[sub]];
            } else if (m.flag[jc] == 0) {
                m.adj[iface] = neib_next;
            } else {
                m.adj[iface] = m.next[neib_next];
            }
        } else if (lvl_c < lvl_n) {
            if (m.flag[jc] < 0) {
                merge_faces(m, ic, side);
                m.adj[iface] = m.next[m.next[jc]];
            } else {
                for (int sub : {side, side + 4}) {
                    size_t jf = 8 * (size_t)ic + sub;
                    m.adj[jf] = m.next[m.adj[jf]];
                }
            }
        } else {
            if (m.flag[jc] == 0) {
                m.adj[iface] = m.next[jc];
            } else {
                int zch = adjacent_child(side, m.z_idx[ic] % 4);
                m.adj[iface] = m.next[m.next[jc] + zch];
            }
        }
    }
}

// refine_cell / make_children / link_siblings (refine.h:36-444).
void refine_cell(AmrWork& W, int ip) {
    orc_mesh& m = W.m;
    const int mc = m.next[ip];
    Vec3 V[9];
    for (int k = 0; k < 9; ++k) V[k] = vert(m, ip, k);
    // SqQuad::children (quad.cpp:641-661) reduced to corners
    const int corners[4][4] = {{0, 1, 3, 4}, {1, 2, 4, 5}, {3, 4, 6, 7}, {4, 5, 7, 8}};
    for (int i = 0; i < 4; ++i)
        set_cell(m, mc + i, V[corners[i][0]], V[corners[i][1]], V[corners[i][2]], V[corners[i][3]]);
    for (int i = 0; i < 4; ++i) {
        int ich = mc + i;
        m.idx[ich] = ich;
        m.flag[ich] = 0;
        m.b_idx[ich] = m.b_idx[ip];
        m.level[ich] = m.level[ip] + 1;
        m.z_idx[ich] = 4 * m.z_idx[ip] + i;
    }
    // link_siblings: (child, side) -> sibling
    const int links[8][3] = {{0, 1, 1}, {0, 3, 2}, {1, 0, 0}, {1, 3, 3},
                             {2, 1, 3}, {2, 2, 0}, {3, 0, 2}, {3, 2, 1}};
    for (auto& L : links) {
        size_t f = 8 * (size_t)(mc + L[0]) + L[1];
        m.bnd[f] = ORC_INNER;
        m.adj[f] = m.next[mc + L[2]];
    }
    for (int side = 0; side < 4; ++side) {
        const int pf = 8 * ip + side;
        const int flag = m.bnd[pf];
        if (flag == ORC_UNDEFINED) continue;
        for (int i : kChildrenBySide[side]) m.bnd[8 * (size_t)(mc + i) + side] = flag;
        if (slot_is_boundary(flag)) {
            for (int i : kChildrenBySide[side]) m.adj[8 * (size_t)(mc + i) + side] = m.next[mc + i];
            continue;
        }
        if (simple_face(m, ip, side)) {
            int jc = m.adj[pf];
            for (int i : kChildrenBySide[side]) {
                size_t chf = 8 * (size_t)(mc + i) + side;
                if (m.level[ip] > m.level[jc]) {
                    int zch = adjacent_child(side, m.z_idx[ip] % 4);
                    m.adj[chf] = m.next[m.next[jc] + zch];
                } else if (m.flag[jc] == 0) {
                    m.adj[chf] = m.next[jc];
                } else {
                    int zch = adjacent_child(side, i);
                    m.adj[chf] = m.next[m.next[jc] + zch];
                }
            }
        } else {
            for (int i : kChildrenBySide[side]) {
                size_t chf = 8 * (size_t)(mc + i) + side;
                int jc = m.adj[8 * (size_t)ip + kSubfaceBySideChild[side][i]];
                if (m.flag[jc] == 0) m.adj[chf] = m.next[jc];
                else if (m.flag[jc] < 0) m.adj[chf] = m.next[m.next[jc]];
                else m.adj[chf] = jc;
            }
        }
    }
    for (int side = 0; side < 4; ++side) {
        const int pf = 8 * ip + side;
        if (simple_face(m, ip, side) || m.bnd[pf] == ORC_UNDEFINED || slot_is_boundary(m.bnd[pf])) continue;
        for (int i : kChildrenBySide[side]) {
            int ich = mc + i;
            int jc = m.adj[8 * (size_t)ip + kSubfaceBySideChild[side][i]];
            if (m.level[jc] + m.flag[jc] > m.level[ich]) {
                split_face(m, ich, side);
                int neib_next = m.next[jc];
                for (int sub : {side, side + 4})
                    m.adj[8 * (size_t)ich + sub] = m.next[neib_next + kNeibChild0[sub]];
            }
        }
    }
    if (W.D) {
        int ch[4] = {mc, mc + 1, mc + 2, mc + 3};
        W.D->split(m, ip, ch);
    }
    m.idx[ip] = -1;
}

// coarse_cell / select_children / make_parent (coarse.h:95-404).
void coarse_cell(AmrWork& W, int ich) {
    orc_mesh& m = W.m;
    if (m.z_idx[ich] % 4 != 0) { m.idx[ich] = -1; return; }
    int ch[4];
    ch[0] = ich;
    int sibs[3];
    get_siblings(m, ich, sibs);
    for (int k = 0; k < 3; ++k) ch[m.z_idx[sibs[k]] % 4] = sibs[k];
    const int ip = m.next[ich];
    // parent_vs -> reduce: corners c0.V0, c1.V2, c2.V6, c3.V8
    set_cell(m, ip, vert(m, ch[0], 0), vert(m, ch[1], 2), vert(m, ch[2], 6), vert(m, ch[3], 8));
    const int ip_next = m.next[ip];
    m.idx[ip] = ip;
    m.flag[ip] = 0;
    m.b_idx[ip] = m.b_idx[ch[0]];
    m.level[ip] = m.level[ch[0]] - 1;
    m.z_idx[ip] = m.z_idx[ch[0]] / 4;
    for (int side = 0; side < 4; ++side) {
        const int some_ch = ch[kChildBySubface[side]];
        const size_t scf = 8 * (size_t)some_ch + side;
        const size_t pf = 8 * (size_t)ip + side;
        m.bnd[pf] = m.bnd[scf];
        if (slot_is_boundary(m.bnd[scf])) { m.adj[pf] = ip_next; continue; }
        const int sn = m.adj[scf];
        const int wanted = m.level[sn] + m.flag[sn];
        if (wanted <= m.level[ip]) {
            m.adj[pf] = m.flag[sn] == 0 ? m.next[sn] : m.next[m.next[sn]];
            set_slot_undefined(m, pf + 4);
            continue;
        }
        split_face(m, ip, side);
        for (int sub : {side, side + 4}) {
            const int c2 = ch[kChildBySubface[sub]];
            const int jc = m.adj[8 * (size_t)c2 + side];
            const size_t sf = 8 * (size_t)ip + sub;
            if (m.level[c2] > m.level[jc]) m.adj[sf] = m.next[m.next[jc] + kNeibChild0[sub]];
            else if (m.flag[jc] == 0) m.adj[sf] = m.next[jc];
            else m.adj[sf] = m.next[m.next[jc]];
        }
    }
    if (W.D) W.D->merge(m, ch, ip);
    m.idx[ich] = -1;
}

// AmrCells::move_item (amr_cells.cpp:428-468) plus the data rows.
void move_cell(orc_mesh& m, AmrData* D, int from, int to) {
    m.next[to] = to; m.idx[to] = to;
    m.flag[to] = m.flag[from]; m.level[to] = m.level[from];
    m.b_idx[to] = m.b_idx[from]; m.z_idx[to] = m.z_idx[from];
    m.cx[to] = m.cx[from]; m.cy[to] = m.cy[from]; m.vol[to] = m.vol[from];
    for (int k = 0; k < 8; ++k) {
        size_t a = 8 * (size_t)from + k, b = 8 * (size_t)to + k;
        m.bnd[b] = m.bnd[a]; m.adj[b] = m.adj[a]; m.fnx[b] = m.fnx[a]; m.fny[b] = m.fny[a];
        m.fcx[b] = m.fcx[a]; m.fcy[b] = m.fcy[a]; m.area[b] = m.area[a];
    }
    for (int k = 0; k < 9; ++k) {
        m.vx[9 * (size_t)to + k] = m.vx[9 * (size_t)from + k];
        m.vy[9 * (size_t)to + k] = m.vy[9 * (size_t)from + k];
    }
    m.idx[from] = -1;
    if (D) D->move(from, to);
}

// apply_impl<2> (apply.h:43-94).  Returns false when nothing changes.
bool apply_flags(orc_mesh& m, AmrData* D) {
    const int n = m.n;
    int n_coarse = 0, n_retain = 0, n_refine = 0;
    for (int ic = 0; ic < n; ++ic) {
        int f = m.flag[ic];
        if (!f) ++n_retain; else if (f > 0) ++n_refine; else ++n_coarse;
    }
    const int n_parents = n_coarse / 4, n_children = 4 * n_refine;
    const int n_large = n + n_parents + n_children;
    const int n_short = n_retain + n_children + n_parents;
    if (n_large <= n) return false;

    // SwapLists (swap_lists.h:28-58)
    std::vector<int> undef, actual;
    const int max_swap = n_large - n_short;
    if (max_swap >= 1) {
        for (int ic = 0; ic < n_short && ic < n; ++ic) {
            if (m.flag[ic] != 0) {
                undef.push_back(ic);
                if ((int)undef.size() >= max_swap) break;
            }
        }
        int jc = n_large - 1;
        while (actual.size() < undef.size() && jc >= n) actual.push_back(jc--);
        while (actual.size() < undef.size() && jc >= 0) {
            if (m.flag[jc] == 0) actual.push_back(jc);
            --jc;
        }
    }
    resize_cells(m, D, n_large);
    // setup_positions (setup_positions.h:24-64)
    int cc = n, rc = n + n_parents;
    for (int ic = 0; ic < n; ++ic) {
        if (m.flag[ic] == 0) { m.next[ic] = ic; continue; }
        if (m.flag[ic] > 0) { m.next[ic] = rc; rc += 4; continue; }
        if (m.z_idx[ic] % 4 == 0) {
            int sibs[3];
            get_siblings(m, ic, sibs);
            m.next[ic] = cc;
            for (int k = 0; k < 3; ++k) m.next[sibs[k]] = cc;
            ++cc;
        }
    }
    for (int ic = n; ic < n_large; ++ic) m.next[ic] = ic;
    for (size_t i = 0; i < actual.size(); ++i) m.next[actual[i]] = undef[i];
    // setup_geometry (setup_geometry.h:19-41)
    AmrWork W{m, D, n};
    for (int ic = 0; ic < n; ++ic) {
        if (m.flag[ic] == 0) retain_cell(m, ic);
        else if (m.flag[ic] > 0) refine_cell(W, ic);
        else coarse_cell(W, ic);
    }
    // move_elements (swap_lists.h:129-138)
    for (size_t i = 0; i < actual.size(); ++i) move_cell(m, D, actual[i], undef[i]);
    resize_cells(m, D, n_short);
    return true;
}

}  // namespace

extern "C" {

// SmFluid::set_flags (sm_fluid.cpp:517-643), criterion 0: relative jumps of
// density and pressure (xi = 0.05, hysteresis 0.4).  Criterion 1: the
// normalised density gradient |grad rho| h / rho against `threshold`
// (+1 above, 0 above 0.4 threshold, -1 otherwise), h the incircle diameter,
// grad rho the unlimited LSM gradient (gradient.h:226-282; a limited one is
// zero next to a jump under minmod).  Both compute the limited gradient too,
// for the slope distributor.
void orc_smf_set_flags(orc_smf* S, int criterion, double threshold, int nthreads) {
    orc_mesh& m = *S->m;
    std::vector<double> raw(2 * (size_t)m.n);
    #pragma omp parallel for num_threads(nthreads) schedule(static)
    for (int c = 0; c < m.n; ++c) {
        Grad g;
        grad_cell(*S, c, g, &raw[2 * (size_t)c]);
        S->d_dx[c] = g.x; S->d_dy[c] = g.y; S->d_dz[c] = g.z;
    }
    for (int ic = 0; ic < m.n; ++ic) {
        const PState& z = S->init[ic];
        if (criterion == 1) {
            double gx = raw[2 * (size_t)ic], gy = raw[2 * (size_t)ic + 1];
            double xi = std::sqrt(gx * gx + gy * gy) * incircle(m, ic) / z.rho;
            m.flag[ic] = xi > threshold ? 1 : (xi > 0.4 * threshold ? 0 : -1);
            continue;
        }
        m.flag[ic] = -1;
        double dens = z.rho, pres = z.p;
        double ds = 0.05 * std::abs(dens), ps = 0.05 * std::abs(pres);
        for (int s8 = 0; s8 < 8; ++s8) {
            size_t f = 8 * (size_t)ic + s8;
            int b = m.bnd[f];
            if (b == ORC_UNDEFINED || slot_is_boundary(b)) continue;
            const PState& zn = S->init[m.adj[f]];
            if (std::abs(zn.rho - dens) > ds || std::abs(zn.p - pres) > ps) { m.flag[ic] = 1; break; }
            if (std::abs(zn.rho - dens) > 0.4 * ds || std::abs(zn.p - pres) > 0.4 * ps) m.flag[ic] = 0;
        }
    }
}

void orc_mesh_set_flags(orc_mesh* m, const int* flags) { std::memcpy(m->flag.data(), flags, sizeof(int) * m->n); }
void orc_mesh_balance_flags(orc_mesh* m, int max_level) { balance_flags(*m, max_level); }

// EuMesh::refine (eu_mesh.cpp:290-359): balance, then apply with the solver's
// distributor (slope = 1: "slope", 0: "const").  S may be NULL (mesh only).
int orc_mesh_refine(orc_mesh* m, orc_smf* S, int max_level, int slope) {
    if (m->axial) return -2;   // the restatement adapts planar meshes only
    balance_flags(*m, max_level);
    if (!S) return apply_flags(*m, nullptr) ? 1 : 0;
    SmfData D(S, slope != 0);
    return apply_flags(*m, &D) ? 1 : 0;
}

void orc_mesh_get_next(const orc_mesh* m, int* next) { std::memcpy(next, m->next.data(), sizeof(int) * m->n); }

}  // extern "C"
