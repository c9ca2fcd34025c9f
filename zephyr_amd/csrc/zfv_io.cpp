// Output and checkpoints of a context, on the host, through the C ABI:
//
// * zfv_vtu_write: VtuFile::save (io/vtu_file.cpp:439-584) for an AMR-2D
//   mesh: the XML UnstructuredGrid header with every array "appended", then
//   "_" and the raw arrays, each behind its 4-byte length.  Cells are quads
//   of their four SqQuad corners (hanging nodes not inserted,
//   fill_adaptive_hex_2D :186-200) or, with polygons set, polygons through
//   the midpoints of their split faces (fill_adaptive_poly_2D :218-252,
//   PvdFile::polyhedral).
// * zfv_pvd_*: PvdFile (io/pvd_file.h:12-75): numbered VTU files in a
//   directory and a PVD collection listing them with their times, rewritten
//   after every save so it is always complete.
// * zfv_backup / zfv_restore: EuMesh::backup (mesh/euler/eu_mesh.cpp:
//   1017-1061, AmrCells::backup amr_cells.cpp:1074-1172): a directory with
//   backup.json describing every array ({"sizeof", "data": {"size",
//   "file"}}) and one raw .bin file per array, in the reference's layout
//   (cells/, cells/faces/, cells/faces/adjacent/, cells/data/); the solver's
//   parameters and state are added ("solver" and cells/data/init, wait).
//   zfv_restore reads such a directory back into a context (the reference
//   has no reader: tests/mesh/backup_read.cpp:23-36 is a stub).
#include "../../include/zfv.h"

#include <cerrno>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <map>
#include <memory>
#include <sstream>
#include <string>
#include <sys/stat.h>
#include <vector>

namespace zfv {
int set_error(int code, const char* msg);   // zfv_last_error's text (zfv_capi.cpp)
}

namespace {

int fail(int code, const std::string& msg) { return zfv::set_error(code, msg.c_str()); }

#define ZIO(call)                                                 \
    do {                                                          \
        int r_ = (call);                                          \
        if (r_ != ZFV_OK) return r_;                              \
    } while (0)

bool make_dirs(const std::string& path) {
    std::string cur;
    std::stringstream ss(path);
    std::string part;
    if (!path.empty() && path[0] == '/') cur = "/";
    while (std::getline(ss, part, '/')) {
        if (part.empty()) continue;
        cur += part + "/";
        if (mkdir(cur.c_str(), 0755) != 0 && errno != EEXIST) return false;
    }
    return true;
}

// Everything a context exposes through the ABI, on the host in mesh order.
struct Snapshot {
    int n = 0;
    std::vector<double> cx, cy, vol, vx, vy, nx, ny, fcx, fcy, area;
    std::vector<int32_t> level, flag, b_idx, z_idx, bnd, adj;
    bool axial = false;                  // the mesh carries volume_alt / area_alt
    std::vector<double> vol_alt, area_alt;
    bool state = false;
    std::vector<double> rho, u, v, p, e;
    std::vector<int32_t> wait;
};

int take(zfv_ctx* ctx, Snapshot& S, bool need_state) {
    int32_t n = 0;
    ZIO(zfv_mesh_n_cells(ctx, &n));
    if (n <= 0) return fail(ZFV_ERR_STATE, "no mesh uploaded");
    S.n = n;
    const size_t ns = 8 * (size_t)n, nv = 9 * (size_t)n;
    S.cx.resize(n); S.cy.resize(n); S.vol.resize(n);
    S.level.resize(n); S.flag.resize(n); S.b_idx.resize(n); S.z_idx.resize(n);
    S.vx.resize(nv); S.vy.resize(nv);
    S.bnd.resize(ns); S.adj.resize(ns);
    S.nx.resize(ns); S.ny.resize(ns); S.fcx.resize(ns); S.fcy.resize(ns); S.area.resize(ns);
    zfv_mesh_soa m{};
    m.n_cells = n;
    m.cx = S.cx.data(); m.cy = S.cy.data(); m.volume = S.vol.data();
    m.vx = S.vx.data(); m.vy = S.vy.data();
    m.level = S.level.data(); m.flag = S.flag.data(); m.b_idx = S.b_idx.data(); m.z_idx = S.z_idx.data();
    m.boundary = S.bnd.data(); m.adj = S.adj.data();
    m.nx = S.nx.data(); m.ny = S.ny.data(); m.fcx = S.fcx.data(); m.fcy = S.fcy.data(); m.area = S.area.data();
    ZIO(zfv_mesh_download(ctx, &m));
    S.vol_alt.resize(n);
    S.area_alt.resize(ns);
    const int ra = zfv_mesh_axial_download(ctx, S.vol_alt.data(), S.area_alt.data());
    if (ra == ZFV_OK) {
        S.axial = true;
    } else if (ra == ZFV_ERR_STATE) {   // a planar mesh
        S.vol_alt.clear();
        S.area_alt.clear();
    } else {
        return ra;
    }
    S.rho.resize(n); S.u.resize(n); S.v.resize(n); S.p.resize(n); S.e.resize(n); S.wait.resize(n);
    const int rs = zfv_state_download(ctx, S.rho.data(), S.u.data(), S.v.data(), S.p.data(), S.e.data());
    if (rs == ZFV_OK) {
        S.state = true;
        ZIO(zfv_wait_download(ctx, S.wait.data()));
    } else if (need_state || rs != ZFV_ERR_STATE) {
        return rs;
    }
    return ZFV_OK;
}

template <class T>
void put_raw(std::string& out, const T* p, size_t count) {
    const uint32_t bytes = (uint32_t)(count * sizeof(T));
    out.append(reinterpret_cast<const char*>(&bytes), sizeof(bytes));
    out.append(reinterpret_cast<const char*>(p), count * sizeof(T));
}

bool little_endian() {
    const uint16_t x = 1;
    return *reinterpret_cast<const uint8_t*>(&x) == 1;
}

}  // namespace

extern "C" {

int zfv_vtu_write(zfv_ctx* ctx, const char* filename, unsigned fields, int polygons) {
    if (!ctx || !filename) return fail(ZFV_ERR_ARG, "null argument");
    if (fields & ~(unsigned)ZFV_VTU_ALL) return fail(ZFV_ERR_ARG, "unknown VTU field bits");
    Snapshot S;
    ZIO(take(ctx, S, (fields & ZFV_VTU_STATE_FIELDS) != 0));
    const int n = S.n;
    // SqQuad vertex k = (iy + 1) * 3 + (ix + 1) for vs<ix, iy> (geom/primitives/quad.h:186-191)
    std::vector<double> pts;
    std::vector<int32_t> conn, offs;
    std::vector<uint8_t> types(n);
    pts.reserve(3 * 4 * (size_t)n);
    offs.resize(n);
    auto push = [&](int c, int k) {
        pts.push_back(S.vx[9 * (size_t)c + k]);
        pts.push_back(S.vy[9 * (size_t)c + k]);
        pts.push_back(0.0);
    };
    auto split = [&](int c, int side) { return S.bnd[8 * (size_t)c + side + 4] != ZFV_UNDEFINED; };
    int np = 0;
    for (int c = 0; c < n; ++c) {
        const int before = np;
        if (!polygons) {
            push(c, 0); push(c, 2); push(c, 8); push(c, 6);
            np += 4;
        } else {
            // counter-clockwise from (-,-), the midpoint of a split side between its corners
            push(c, 0); ++np;
            if (split(c, 2)) { push(c, 1); ++np; }
            push(c, 2); ++np;
            if (split(c, 1)) { push(c, 5); ++np; }
            push(c, 8); ++np;
            if (split(c, 3)) { push(c, 7); ++np; }
            push(c, 6); ++np;
            if (split(c, 0)) { push(c, 3); ++np; }
        }
        offs[c] = np;
        const int nv = np - before;
        types[c] = nv == 4 ? 9 /* VTK_QUAD */ : 7 /* VTK_POLYGON */;
    }
    conn.resize(np);
    for (int i = 0; i < np; ++i) conn[i] = i;

    struct Field {
        const char* name;
        const char* type;
        int comps;
        std::string data;
    };
    std::vector<Field> cf;
    auto add_d = [&](const char* name, int comps, auto&& value) {
        Field f{name, "Float64", comps, {}};
        f.data.resize(sizeof(uint32_t));
        const uint32_t bytes = (uint32_t)((size_t)n * comps * 8);
        std::memcpy(&f.data[0], &bytes, 4);
        for (int c = 0; c < n; ++c)
            for (int k = 0; k < comps; ++k) {
                const double x = value(c, k);
                f.data.append(reinterpret_cast<const char*>(&x), 8);
            }
        cf.push_back(std::move(f));
    };
    auto add_i = [&](const char* name, const std::vector<int32_t>& a) {
        Field f{name, "Int32", 1, {}};
        put_raw(f.data, a.data(), a.size());
        cf.push_back(std::move(f));
    };
    if (fields & ZFV_VTU_DENSITY) add_d("density", 1, [&](int c, int) { return S.rho[c]; });
    if (fields & ZFV_VTU_VELOCITY)
        add_d("velocity", 3, [&](int c, int k) { return k == 0 ? S.u[c] : (k == 1 ? S.v[c] : 0.0); });
    if (fields & ZFV_VTU_PRESSURE) add_d("pressure", 1, [&](int c, int) { return S.p[c]; });
    if (fields & ZFV_VTU_ENERGY) add_d("energy", 1, [&](int c, int) { return S.e[c]; });
    if (fields & ZFV_VTU_WAIT) add_i("wait", S.wait);
    if (fields & ZFV_VTU_LEVEL) add_i("level", S.level);
    if (fields & ZFV_VTU_FLAG) add_i("flag", S.flag);
    if (fields & ZFV_VTU_INDEX) {
        std::vector<int32_t> idx(n);
        for (int c = 0; c < n; ++c) idx[c] = c;
        add_i("index", idx);
    }

    std::string body;
    put_raw(body, pts.data(), pts.size());
    const size_t o_conn = body.size();
    put_raw(body, conn.data(), conn.size());
    const size_t o_offs = body.size();
    put_raw(body, offs.data(), offs.size());
    const size_t o_types = body.size();
    put_raw(body, types.data(), types.size());
    std::vector<size_t> o_fields;
    for (auto& f : cf) {
        o_fields.push_back(body.size());
        body += f.data;
    }
    std::ostringstream h;
    h << "<VTKFile type=\"UnstructuredGrid\" version=\"0.1\" byte_order=\""
      << (little_endian() ? "LittleEndian" : "BigEndian") << "\">\n";
    h << "  <UnstructuredGrid>\n";
    h << "    <Piece NumberOfPoints=\"" << np << "\" NumberOfCells=\"" << n << "\">\n";
    h << "      <Points>\n";
    h << "        <DataArray type=\"Float64\" Name=\"Points\" NumberOfComponents=\"3\" format=\"appended\" offset=\"0\"/>\n";
    h << "      </Points>\n";
    h << "      <Cells>\n";
    h << "        <DataArray type=\"Int32\" Name=\"connectivity\" format=\"appended\" offset=\"" << o_conn << "\"/>\n";
    h << "        <DataArray type=\"Int32\" Name=\"offsets\" format=\"appended\" offset=\"" << o_offs << "\"/>\n";
    h << "        <DataArray type=\"UInt8\" Name=\"types\" format=\"appended\" offset=\"" << o_types << "\"/>\n";
    h << "      </Cells>\n";
    h << "      <CellData>\n";
    for (size_t i = 0; i < cf.size(); ++i) {
        h << "        <DataArray type=\"" << cf[i].type << "\" Name=\"" << cf[i].name;
        if (cf[i].comps > 1) h << "\" NumberOfComponents=\"" << cf[i].comps;
        h << "\" format=\"appended\" offset=\"" << o_fields[i] << "\"/>\n";
    }
    h << "      </CellData>\n";
    h << "    </Piece>\n";
    h << "  </UnstructuredGrid>\n";
    h << "  <AppendedData encoding=\"raw\">\n_";
    std::ofstream f(filename, std::ios::out | std::ios::binary | std::ios::trunc);
    if (!f) return fail(ZFV_ERR_ARG, std::string("cannot open ") + filename);
    const std::string head = h.str();
    f.write(head.data(), (std::streamsize)head.size());
    f.write(body.data(), (std::streamsize)body.size());
    f << "\n  </AppendedData>\n</VTKFile>\n";
    if (!f) return fail(ZFV_ERR_ARG, std::string("write failed: ") + filename);
    return ZFV_OK;
}

}  // extern "C"

struct zfv_pvd {
    std::string dir, name;
    std::vector<std::pair<double, std::string>> steps;
};

extern "C" {

int zfv_pvd_open(const char* filename, const char* directory, zfv_pvd** out) {
    if (!filename || !out) return fail(ZFV_ERR_ARG, "null argument");
    std::string dir = directory && *directory ? directory : "output";
    if (!make_dirs(dir + "/" + filename)) return fail(ZFV_ERR_ARG, "cannot create " + dir);
    auto* p = new zfv_pvd();
    p->dir = dir;
    p->name = filename;
    *out = p;
    return ZFV_OK;
}

int zfv_pvd_save(zfv_pvd* pvd, zfv_ctx* ctx, double time, unsigned fields, int polygons) {
    if (!pvd || !ctx) return fail(ZFV_ERR_ARG, "null argument");
    char num[32];
    std::snprintf(num, sizeof(num), "%zu", pvd->steps.size());
    const std::string rel = pvd->name + "/" + pvd->name + "_" + num + ".vtu";
    ZIO(zfv_vtu_write(ctx, (pvd->dir + "/" + rel).c_str(), fields, polygons));
    pvd->steps.emplace_back(time, rel);
    std::ofstream f(pvd->dir + "/" + pvd->name + ".pvd", std::ios::out | std::ios::trunc);
    if (!f) return fail(ZFV_ERR_ARG, "cannot write the PVD file");
    f << "<?xml version=\"1.0\"?>\n";
    f << "<VTKFile type=\"Collection\" version=\"0.1\" byte_order=\""
      << (little_endian() ? "LittleEndian" : "BigEndian") << "\">\n";
    f << "  <Collection>\n";
    f.precision(17);
    for (const auto& s : pvd->steps)
        f << "    <DataSet timestep=\"" << s.first << "\" group=\"\" part=\"0\" file=\"" << s.second << "\"/>\n";
    f << "  </Collection>\n";
    f << "</VTKFile>\n";
    if (!f) return fail(ZFV_ERR_ARG, "cannot write the PVD file");
    return ZFV_OK;
}

int zfv_pvd_close(zfv_pvd* pvd) {
    delete pvd;
    return ZFV_OK;
}

}  // extern "C"

// ---------------------------------------------------------------- backup
namespace {

struct Writer {
    std::string root;
    std::ostringstream js;
    int fail_code = ZFV_OK;
    std::string fail_msg;

    template <class T>
    void arr(const std::string& dir, const std::string& tab, const std::string& name, const T* p, size_t count,
             size_t elem, const char* end) {
        const std::string rel = dir + "/" + name + ".bin";
        std::ofstream b(root + "/" + rel, std::ios::binary | std::ios::trunc | std::ios::out);
        b.write(reinterpret_cast<const char*>(p), (std::streamsize)(count * sizeof(T)));
        if (!b && fail_code == ZFV_OK) {
            fail_code = ZFV_ERR_ARG;
            fail_msg = "cannot write " + root + "/" + rel;
        }
        js << tab << "\"" << name << "\": {\n";
        js << tab << "  \"sizeof\": " << elem << ",\n";
        js << tab << "  \"data\": { \"size\": " << (count * sizeof(T)) / elem << ", \"file\": \"" << rel << "\" }\n";
        js << tab << "}" << end;
    }
};

// Minimal JSON reader for the backup description this file writes:
// objects, arrays, strings without escapes, numbers, true / false.
struct JVal {
    enum Kind { NUL, NUM, STR, BOOL, OBJ, ARR } kind = NUL;
    double num = 0.0;
    bool b = false;
    std::string str;
    std::map<std::string, std::shared_ptr<JVal>> obj;
    std::vector<std::shared_ptr<JVal>> arr;
    const JVal* get(const std::string& k) const {
        auto it = obj.find(k);
        return it == obj.end() ? nullptr : it->second.get();
    }
};

struct JParser {
    const std::string& s;
    size_t i = 0;
    bool ok = true;
    explicit JParser(const std::string& text) : s(text) {}
    void ws() {
        while (i < s.size() && (s[i] == ' ' || s[i] == '\n' || s[i] == '\t' || s[i] == '\r')) ++i;
    }
    std::shared_ptr<JVal> value() {
        ws();
        auto v = std::make_shared<JVal>();
        if (i >= s.size()) {
            ok = false;
            return v;
        }
        const char c = s[i];
        if (c == '{') {
            v->kind = JVal::OBJ;
            ++i;
            ws();
            if (i < s.size() && s[i] == '}') {
                ++i;
                return v;
            }
            while (ok) {
                ws();
                auto k = value();
                if (k->kind != JVal::STR) {
                    ok = false;
                    break;
                }
                ws();
                if (i >= s.size() || s[i] != ':') {
                    ok = false;
                    break;
                }
                ++i;
                v->obj[k->str] = value();
                ws();
                if (i < s.size() && s[i] == ',') {
                    ++i;
                    continue;
                }
                if (i < s.size() && s[i] == '}') {
                    ++i;
                    break;
                }
                ok = false;
            }
        } else if (c == '[') {
            v->kind = JVal::ARR;
            ++i;
            ws();
            if (i < s.size() && s[i] == ']') {
                ++i;
                return v;
            }
            while (ok) {
                v->arr.push_back(value());
                ws();
                if (i < s.size() && s[i] == ',') {
                    ++i;
                    continue;
                }
                if (i < s.size() && s[i] == ']') {
                    ++i;
                    break;
                }
                ok = false;
            }
        } else if (c == '"') {
            v->kind = JVal::STR;
            const size_t j = s.find('"', i + 1);
            if (j == std::string::npos) {
                ok = false;
                return v;
            }
            v->str = s.substr(i + 1, j - i - 1);
            i = j + 1;
        } else if (s.compare(i, 4, "true") == 0) {
            v->kind = JVal::BOOL;
            v->b = true;
            i += 4;
        } else if (s.compare(i, 5, "false") == 0) {
            v->kind = JVal::BOOL;
            i += 5;
        } else {
            v->kind = JVal::NUM;
            char* end = nullptr;
            v->num = std::strtod(s.c_str() + i, &end);
            if (end == s.c_str() + i) ok = false;
            i = (size_t)(end - s.c_str());
        }
        return v;
    }
};

}  // namespace

extern "C" {

int zfv_backup(zfv_ctx* ctx, const char* root) {
    if (!ctx || !root || !*root) return fail(ZFV_ERR_ARG, "null argument");
    Snapshot S;
    ZIO(take(ctx, S, true));
    zfv_smf_params P{};
    ZIO(zfv_smf_get_params(ctx, &P));
    const std::string R = root;
    for (const char* d : {"cells", "cells/faces", "cells/faces/adjacent", "cells/data"})
        if (!make_dirs(R + "/" + d)) return fail(ZFV_ERR_ARG, "cannot create " + R + "/" + d);
    const int n = S.n;
    const size_t ns = 8 * (size_t)n;
    int lmax = 0;
    for (int c = 0; c < n; ++c) lmax = S.level[c] > lmax ? S.level[c] : lmax;
    // Vector3d arrays (centres, normals, vertices) and the PState layer
    std::vector<double> center(3 * (size_t)n), fnorm(3 * ns), fcent(3 * ns), verts(27 * (size_t)n), init(6 * (size_t)n);
    std::vector<int32_t> face_begin(n + 1), node_begin(n + 1), index(n), rank(n, 0), arank(ns, 0), alien(ns, -1),
        basic(ns);
    for (int c = 0; c < n; ++c) {
        center[3 * c] = S.cx[c];
        center[3 * c + 1] = S.cy[c];
        center[3 * c + 2] = 0.0;
        index[c] = c;
        for (int k = 0; k < 9; ++k) {
            verts[27 * (size_t)c + 3 * k] = S.vx[9 * (size_t)c + k];
            verts[27 * (size_t)c + 3 * k + 1] = S.vy[9 * (size_t)c + k];
            verts[27 * (size_t)c + 3 * k + 2] = 0.0;
        }
        const double z[6] = {S.rho[c], S.u[c], S.v[c], 0.0, S.p[c], S.e[c]};
        for (int k = 0; k < 6; ++k) init[6 * (size_t)c + k] = z[k];
    }
    for (int c = 0; c <= n; ++c) {
        face_begin[c] = 8 * c;
        node_begin[c] = 9 * c;
    }
    for (size_t f = 0; f < ns; ++f) {
        basic[f] = (int32_t)(f / 8);   // the face's own cell
        fnorm[3 * f] = S.nx[f];
        fnorm[3 * f + 1] = S.ny[f];
        fnorm[3 * f + 2] = 0.0;
        fcent[3 * f] = S.fcx[f];
        fcent[3 * f + 1] = S.fcy[f];
        fcent[3 * f + 2] = 0.0;
    }
    Writer W;
    W.root = R;
    auto& js = W.js;
    js.precision(17);
    js << std::boolalpha;
    js << "{\n  \"mesh\": {\n";
    js << "    \"max_level\":  " << lmax << ",\n";
    js << "    \"structured\": false,\n";
    js << "    \"cells\": {\n";
    const std::string t = "      ", t2 = "        ", t3 = "          ";
    js << t << "\"dim\":      2,\n" << t << "\"adaptive\": true,\n" << t << "\"axial\":    " << S.axial << ",\n"
       << t << "\"linear\":   true,\n" << t << "\"size\":     " << n << ",\n";
    W.arr("cells", t, "rank", rank.data(), rank.size(), 4, ",\n");
    W.arr("cells", t, "index", index.data(), index.size(), 4, ",\n");
    W.arr("cells", t, "level", S.level.data(), S.level.size(), 4, ",\n");
    W.arr("cells", t, "flag", S.flag.data(), S.flag.size(), 4, ",\n");
    W.arr("cells", t, "b_idx", S.b_idx.data(), S.b_idx.size(), 4, ",\n");
    W.arr("cells", t, "z_idx", S.z_idx.data(), S.z_idx.size(), 4, ",\n");
    W.arr("cells", t, "center", center.data(), center.size(), 24, ",\n");
    W.arr("cells", t, "volume", S.vol.data(), S.vol.size(), 8, ",\n");
    if (S.axial) W.arr("cells", t, "volume_alt", S.vol_alt.data(), S.vol_alt.size(), 8, ",\n");
    W.arr("cells", t, "face_begin", face_begin.data(), face_begin.size(), 4, ",\n");
    W.arr("cells", t, "node_begin", node_begin.data(), node_begin.size(), 4, ",\n");
    js << t << "\"faces\": {\n";
    js << t2 << "\"adjacent\": {\n";
    W.arr("cells/faces/adjacent", t3, "rank", arank.data(), arank.size(), 4, ",\n");
    W.arr("cells/faces/adjacent", t3, "index", S.adj.data(), S.adj.size(), 4, ",\n");
    W.arr("cells/faces/adjacent", t3, "alien", alien.data(), alien.size(), 4, ",\n");
    W.arr("cells/faces/adjacent", t3, "basic", basic.data(), basic.size(), 4, "\n");
    js << t2 << "},\n";
    W.arr("cells/faces", t2, "boundary", S.bnd.data(), S.bnd.size(), 4, ",\n");
    W.arr("cells/faces", t2, "normal", fnorm.data(), fnorm.size(), 24, ",\n");
    W.arr("cells/faces", t2, "center", fcent.data(), fcent.size(), 24, ",\n");
    W.arr("cells/faces", t2, "area", S.area.data(), S.area.size(), 8, S.axial ? ",\n" : "\n");
    if (S.axial) W.arr("cells/faces", t2, "area_alt", S.area_alt.data(), S.area_alt.size(), 8, "\n");
    js << t << "},\n";
    W.arr("cells", t, "verts", verts.data(), verts.size(), 24, ",\n");
    js << t << "\"data\": {\n";
    W.arr("cells/data", t2, "init", init.data(), init.size(), 48, ",\n");
    W.arr("cells/data", t2, "wait", S.wait.data(), S.wait.size(), 4, "\n");
    js << t << "}\n";
    js << "    }\n  },\n";
    js << "  \"solver\": {\n";
    js << "    \"kind\": \"SmFluid\",\n";
    js << "    \"accuracy\": " << P.acc << ",\n    \"flux\": " << P.flux << ",\n    \"limiter\": " << P.limiter
       << ",\n    \"CFL\": " << P.cfl << ",\n    \"max_dt\": " << P.max_dt << ",\n    \"axial\": " << (P.axial != 0)
       << ",\n";
    js << "    \"eos\": { \"gamma\": " << P.gamma << ", \"P0\": " << P.P0 << ", \"e0\": " << P.e0 << " }\n";
    js << "  }\n}\n";
    if (W.fail_code) return fail(W.fail_code, W.fail_msg);
    std::ofstream f(R + "/backup.json", std::ios::out | std::ios::trunc);
    f << js.str();
    if (!f) return fail(ZFV_ERR_ARG, "cannot write " + R + "/backup.json");
    return ZFV_OK;
}

int zfv_restore(zfv_ctx* ctx, const char* root) {
    if (!ctx || !root) return fail(ZFV_ERR_ARG, "null argument");
    const std::string R = root;
    std::ifstream f(R + "/backup.json");
    if (!f) return fail(ZFV_ERR_ARG, "cannot read " + R + "/backup.json");
    std::stringstream ss;
    ss << f.rdbuf();
    const std::string text = ss.str();
    JParser jp(text);
    auto top = jp.value();
    if (!jp.ok || top->kind != JVal::OBJ) return fail(ZFV_ERR_ARG, "backup.json: not a JSON object");
    const JVal* mesh = top->get("mesh");
    const JVal* cells = mesh ? mesh->get("cells") : nullptr;
    const JVal* solver = top->get("solver");
    if (!cells || !solver) return fail(ZFV_ERR_ARG, "backup.json: no mesh.cells or solver");
    const JVal* sz = cells->get("size");
    const JVal* dim = cells->get("dim");
    if (!sz || sz->kind != JVal::NUM || !dim || dim->num != 2.0)
        return fail(ZFV_ERR_UNSUPPORTED, "backup.json: only single-process AMR-2D meshes are read");
    const long long n = (long long)sz->num;
    if (n <= 0 || n > (1LL << 28)) return fail(ZFV_ERR_ARG, "backup.json: bad cell count");
    // one array: its description, element size and count checked, then the raw file
    auto load = [&](const JVal* node, size_t elem, size_t count, void* dst) -> int {
        if (!node) return fail(ZFV_ERR_ARG, "backup.json: missing array");
        const JVal* so = node->get("sizeof");
        const JVal* d = node->get("data");
        const JVal* dsz = d ? d->get("size") : nullptr;
        const JVal* dfile = d ? d->get("file") : nullptr;
        if (!so || !dsz || !dfile || (size_t)so->num != elem || (size_t)dsz->num != count)
            return fail(ZFV_ERR_ARG, "backup.json: array of unexpected size or element type");
        std::ifstream b(R + "/" + dfile->str, std::ios::binary);
        b.read(static_cast<char*>(dst), (std::streamsize)(elem * count));
        if (!b) return fail(ZFV_ERR_ARG, "cannot read " + R + "/" + dfile->str);
        return ZFV_OK;
    };
    const size_t N = (size_t)n, NS = 8 * N;
    std::vector<int32_t> level(N), flag(N), b_idx(N), z_idx(N), bnd(NS), adj(NS), wait(N);
    std::vector<double> center(3 * N), vol(N), fnorm(3 * NS), fcent(3 * NS), area(NS), verts(27 * N), init(6 * N);
    const JVal* faces = cells->get("faces");
    const JVal* adjacent = faces ? faces->get("adjacent") : nullptr;
    const JVal* data = cells->get("data");
    if (!faces || !adjacent || !data) return fail(ZFV_ERR_ARG, "backup.json: missing faces or data");
    ZIO(load(cells->get("level"), 4, N, level.data()));
    ZIO(load(cells->get("flag"), 4, N, flag.data()));
    ZIO(load(cells->get("b_idx"), 4, N, b_idx.data()));
    ZIO(load(cells->get("z_idx"), 4, N, z_idx.data()));
    ZIO(load(cells->get("center"), 24, N, center.data()));
    ZIO(load(cells->get("volume"), 8, N, vol.data()));
    ZIO(load(adjacent->get("index"), 4, NS, adj.data()));
    ZIO(load(faces->get("boundary"), 4, NS, bnd.data()));
    ZIO(load(faces->get("normal"), 24, NS, fnorm.data()));
    ZIO(load(faces->get("center"), 24, NS, fcent.data()));
    ZIO(load(faces->get("area"), 8, NS, area.data()));
    ZIO(load(cells->get("verts"), 24, 9 * N, verts.data()));
    ZIO(load(data->get("init"), 48, N, init.data()));
    ZIO(load(data->get("wait"), 4, N, wait.data()));
    std::vector<double> cx(N), cy(N), vx(9 * N), vy(9 * N), nx(NS), ny(NS), fcx(NS), fcy(NS);
    for (size_t c = 0; c < N; ++c) {
        cx[c] = center[3 * c];
        cy[c] = center[3 * c + 1];
        for (int k = 0; k < 9; ++k) {
            vx[9 * c + k] = verts[27 * c + 3 * k];
            vy[9 * c + k] = verts[27 * c + 3 * k + 1];
        }
    }
    for (size_t s = 0; s < NS; ++s) {
        nx[s] = fnorm[3 * s];
        ny[s] = fnorm[3 * s + 1];
        fcx[s] = fcent[3 * s];
        fcy[s] = fcent[3 * s + 1];
        if (bnd[s] == ZFV_UNDEFINED) {   // undefined slots carry no geometry: any upload-valid values
            adj[s] = (int32_t)(s / 8);
            nx[s] = 1.0;
            ny[s] = 0.0;
        }
    }
    zfv_mesh_soa m{};
    m.n_cells = (int32_t)n;
    m.cx = cx.data(); m.cy = cy.data(); m.volume = vol.data(); m.vx = vx.data(); m.vy = vy.data();
    m.level = level.data(); m.flag = flag.data(); m.b_idx = b_idx.data(); m.z_idx = z_idx.data();
    m.boundary = bnd.data(); m.adj = adj.data(); m.nx = nx.data(); m.ny = ny.data();
    m.fcx = fcx.data(); m.fcy = fcy.data(); m.area = area.data();
    const JVal* va = cells->get("volume_alt");
    if (va) {   // an axisymmetric mesh: its y-weighted volume and face areas
        std::vector<double> vol_alt(N), area_alt(NS);
        ZIO(load(va, 8, N, vol_alt.data()));
        ZIO(load(faces->get("area_alt"), 8, NS, area_alt.data()));
        ZIO(zfv_mesh_upload_axial(ctx, &m, vol_alt.data(), area_alt.data()));
    } else {
        ZIO(zfv_mesh_upload(ctx, &m));
    }
    auto num = [&](const JVal* o, const char* k, double& out) -> bool {
        const JVal* v = o ? o->get(k) : nullptr;
        if (!v || v->kind != JVal::NUM) return false;
        out = v->num;
        return true;
    };
    double acc, flux, lim, cfl, max_dt, g, P0, e0;
    const JVal* eos = solver->get("eos");
    if (!num(solver, "accuracy", acc) || !num(solver, "flux", flux) || !num(solver, "limiter", lim) ||
        !num(solver, "CFL", cfl) || !num(solver, "max_dt", max_dt) || !num(eos, "gamma", g) || !num(eos, "P0", P0) ||
        !num(eos, "e0", e0))
        return fail(ZFV_ERR_ARG, "backup.json: incomplete solver parameters");
    zfv_smf_params P{};
    P.acc = (int32_t)acc;
    P.flux = (int32_t)flux;
    P.limiter = (int32_t)lim;
    P.cfl = cfl;
    P.max_dt = max_dt;
    P.gamma = g;
    P.P0 = P0;
    P.e0 = e0;
    const JVal* ax = solver->get("axial");
    P.axial = ax && ax->kind == JVal::BOOL && ax->b ? 1 : 0;
    ZIO(zfv_smf_configure(ctx, &P));
    std::vector<double> rho(N), u(N), v(N), p(N), e(N);
    for (size_t c = 0; c < N; ++c) {
        rho[c] = init[6 * c];
        u[c] = init[6 * c + 1];
        v[c] = init[6 * c + 2];
        p[c] = init[6 * c + 4];
        e[c] = init[6 * c + 5];
    }
    ZIO(zfv_state_upload(ctx, rho.data(), u.data(), v.data(), p.data(), e.data()));
    ZIO(zfv_wait_upload(ctx, wait.data()));
    ZIO(zfv_flags_upload(ctx, flag.data()));
    return ZFV_OK;
}

}  // extern "C"
