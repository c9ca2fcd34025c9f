// zfv.hpp — host C++ facade over the zfv C ABI with the shapes of zephyr's
// problem-setup API (generator::Rectangle, EuMesh, SmFluid, IdealGas /
// StiffenedGas), so that a problem file written against the reference
// (problems/single-fluid/smf_test_2D.cpp) switches to the GPU path by
// swapping the includes and the namespace.
//
// Header-only; link against libzfv.so.  Errors of the C ABI become
// std::runtime_error, as the reference throws std::runtime_error.
#pragma once

#include <zfv.h>

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdint>
#include <functional>
#include <limits>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

namespace zfv {
namespace host {

inline void check(int code) {
    if (code != ZFV_OK) throw std::runtime_error(std::string("zfv: ") + zfv_last_error());
}

// geom::Boundary (geom/boundary.h:9-16).
enum class Boundary : int { UNDEFINED = -1, INNER = 0, WALL = 1, ZOE = 2, PERIODIC = 3, FUNCTION = 4 };

// math::Fluxes, the members the step implements (math/cfd/flux/num_flux.h:11-21).
enum class Fluxes : int {
    GODUNOV = ZFV_GODUNOV, RUSANOV = ZFV_RUSANOV, HLL = ZFV_HLL, HLLC = ZFV_HLLC,
    HLLC_M = ZFV_HLLC_M, HLLC_LM = ZFV_HLLC_LM, CIR1 = ZFV_CIR1, CIR2 = ZFV_CIR2
};

// Equations of state: StiffenedGas (phys/matter/eos/stiffened_gas.cpp:103-175);
// IdealGas is P0 = e0 = 0 (ideal_gas.cpp:32-104).
struct Eos {
    double gamma = 1.4, P0 = 0.0, e0 = 0.0;
    using Ptr = std::shared_ptr<Eos>;
    double pressure_re(double rho, double e) const { return (gamma - 1.0) * rho * (e - e0) - gamma * P0; }
    double energy_rP(double rho, double p) const { return e0 + (p + gamma * P0) / ((gamma - 1.0) * rho); }
};
struct IdealGas : Eos {
    explicit IdealGas(double g = 1.4) { gamma = g; }
    static Eos::Ptr create(double g = 1.4) { return std::make_shared<IdealGas>(g); }
};
struct StiffenedGas : Eos {
    StiffenedGas(double g, double p0, double E0 = 0.0) {
        gamma = g;
        P0 = p0;
        e0 = E0;
    }
    static Eos::Ptr create(double g, double p0, double E0 = 0.0) { return std::make_shared<StiffenedGas>(g, p0, E0); }
};

// smf::PState in 2D: {density, velocity (x, y), pressure, energy}.
struct PState {
    double density = 0.0, vx = 0.0, vy = 0.0, pressure = 0.0, energy = 0.0;
};

// generator::Rectangle (geom/generator/rectangle.h): bounds, boundary
// conditions {left, right, bottom, top}, cell counts.
class Rectangle {
public:
    Rectangle(double xmin, double xmax, double ymin, double ymax) : x0_(xmin), x1_(xmax), y0_(ymin), y1_(ymax) {}
    void set_boundaries(std::array<Boundary, 4> b) { bounds_ = b; }
    // set_nx: ny = max(round(nx * height / width), 1) (rectangle.cpp:105-112).
    void set_nx(int nx) {
        nx_ = nx;
        ny_ = std::max(int(std::round(nx * (y1_ - y0_) / (x1_ - x0_))), 1);
    }
    void set_sizes(int nx, int ny) {
        nx_ = nx;
        ny_ = ny;
    }
    // set_axial (rectangle.cpp:97-103): y is the radius; the mesh then
    // carries the axisymmetric geometry (a bottom on the axis is a wall).
    void set_axial(bool axial = true) { axial_ = axial; }
    bool axial() const { return axial_; }
    int nx() const { return nx_; }
    int ny() const { return ny_; }

private:
    friend class EuMesh;
    double x0_, x1_, y0_, y1_;
    int nx_ = 0, ny_ = 0;
    bool axial_ = false;
    std::array<Boundary, 4> bounds_{Boundary::ZOE, Boundary::ZOE, Boundary::ZOE, Boundary::ZOE};
};

// Distributor choice of SmFluid::distributor(type) (sm_fluid.cpp:372-470).
enum class Distributor : int { CONST = 0, SLOPE = 1 };

// EuMesh (mesh/euler/eu_mesh.h): owns one device context, i.e. the device
// mesh and the solver state living on it.
class EuMesh {
public:
    explicit EuMesh(const Rectangle& gen, int device = 0) {
        check(zfv_create(device, &ctx_));
        int32_t b[4] = {(int32_t)gen.bounds_[0], (int32_t)gen.bounds_[1], (int32_t)gen.bounds_[2],
                        (int32_t)gen.bounds_[3]};
        check(gen.axial_ ? zfv_mesh_rectangle_axial(ctx_, gen.nx_, gen.ny_, gen.x0_, gen.x1_, gen.y0_, gen.y1_, b)
                         : zfv_mesh_rectangle(ctx_, gen.nx_, gen.ny_, gen.x0_, gen.x1_, gen.y0_, gen.y1_, b));
    }
    ~EuMesh() {
        if (ctx_) zfv_destroy(ctx_);
    }
    EuMesh(const EuMesh&) = delete;
    EuMesh& operator=(const EuMesh&) = delete;

    void set_max_level(int lvl) { max_level_ = lvl; }
    int max_level() const { return max_level_; }
    void set_distributor(Distributor d) { distributor_ = d; }

    // EuMesh::refine (eu_mesh.cpp:290-359): balance the flags set by
    // SmFluid::set_flags and adapt, moving the state with the distributor.
    void refine() { check(zfv_refine(ctx_, max_level_, (int)distributor_)); }

    int n_cells() const {
        int32_t n = 0;
        check(zfv_mesh_n_cells(ctx_, &n));
        return n;
    }

    // Cell centres of the current mesh (host copy).
    void centers(std::vector<double>& x, std::vector<double>& y) const {
        const int n = n_cells();
        x.resize(n);
        y.resize(n);
        zfv_mesh_soa o{};
        o.n_cells = n;
        o.cx = x.data();
        o.cy = y.data();
        check(zfv_mesh_download(ctx_, &o));
    }

    zfv_ctx* context() const { return ctx_; }

private:
    zfv_ctx* ctx_ = nullptr;
    int max_level_ = 0;
    Distributor distributor_ = Distributor::SLOPE;
};

// SmFluid (math/solver/sm_fluid.h:22-106) on the device.
class SmFluid {
public:
    explicit SmFluid(Eos::Ptr eos) : eos_(std::move(eos)) {}

    // Setters (sm_fluid.cpp:37-67); after add_types they reconfigure the device solver.
    void set_CFL(double c) {
        cfl_ = std::max(0.0, std::min(c, 1.0));
        reconfigure();
    }
    void set_accuracy(int acc) {
        acc_ = std::max(1, std::min(acc, 2));
        reconfigure();
    }
    void set_method(Fluxes f) {
        flux_ = f;
        reconfigure();
    }
    void set_limiter(const std::string& name) {
        int id = zfv_limiter_from_name(name.c_str());
        if (id < 0) throw std::runtime_error("Unknown limiter '" + name + "'");
        limiter_ = id;
        reconfigure();
    }
    void set_max_dt(double dt) {
        max_dt_ = dt;
        if (mesh_) check(zfv_smf_set_max_dt(mesh_->context(), dt));
    }
    // set_axial (sm_fluid.cpp:45-47): the axisymmetric areas, volumes and
    // pressure source; the mesh must come from an axial Rectangle.
    void set_axial(bool axial = true) {
        axial_ = axial;
        reconfigure();
    }
    double CFL() const { return cfl_; }
    int accuracy() const { return acc_; }

    // add_types (sm_fluid.cpp:25-35): binds the solver to the mesh; the
    // layers live on the device.
    void add_types(EuMesh& mesh) {
        mesh_ = &mesh;
        configure();
    }
    Distributor distributor(const std::string& type = "slope") const {
        return type == "const" ? Distributor::CONST : Distributor::SLOPE;
    }

    // Initial data at the cell centres (the reference's for_each over cells
    // writing cell[init]).
    void set_initial(EuMesh& mesh, const std::function<PState(double, double)>& f) {
        if (mesh_ != &mesh) add_types(mesh);
        std::vector<double> x, y;
        mesh.centers(x, y);
        const size_t n = x.size();
        std::vector<double> r(n), u(n), v(n), p(n), e(n);
        for (size_t i = 0; i < n; ++i) {
            PState z = f(x[i], y[i]);
            r[i] = z.density;
            u[i] = z.vx;
            v[i] = z.vy;
            p[i] = z.pressure;
            e[i] = z.energy != 0.0 ? z.energy : eos_->energy_rP(z.density, z.pressure);
        }
        check(zfv_state_upload(mesh.context(), r.data(), u.data(), v.data(), p.data(), e.data()));
    }

    // update (sm_fluid.cpp:81-102); `steps` > 1 enqueues several steps.
    void update(EuMesh& mesh, int steps = 1) { check(zfv_smf_step(bound(mesh), steps)); }
    double dt() const {
        double d = 0.0;
        check(zfv_smf_dt(mesh_->context(), &d));
        return d;
    }
    // set_flags (sm_fluid.cpp:517-643); criterion 1 = |grad rho| h / rho.
    void set_flags(EuMesh& mesh, int criterion = 0, double threshold = 0.1) {
        check(zfv_set_flags(bound(mesh), criterion, threshold));
    }

    // State download (SoA host arrays).
    void state(EuMesh& mesh, std::vector<PState>& out) const {
        const int n = mesh.n_cells();
        std::vector<double> r(n), u(n), v(n), p(n), e(n);
        check(zfv_state_download(mesh.context(), r.data(), u.data(), v.data(), p.data(), e.data()));
        out.resize(n);
        for (int i = 0; i < n; ++i) out[i] = PState{r[i], u[i], v[i], p[i], e[i]};
    }
    std::array<double, 4> totals(EuMesh& mesh) const {
        std::array<double, 4> t{};
        check(zfv_conserved_totals(mesh.context(), t.data()));
        return t;
    }

private:
    zfv_ctx* bound(EuMesh& mesh) {
        if (mesh_ != &mesh) add_types(mesh);
        return mesh.context();
    }
    void reconfigure() {
        if (mesh_) configure();
    }
    void configure() {
        zfv_smf_params p{};
        p.acc = acc_;
        p.flux = (int)flux_;
        p.limiter = limiter_;
        p.axial = axial_ ? 1 : 0;
        p.cfl = cfl_;
        p.max_dt = max_dt_;
        p.gamma = eos_->gamma;
        p.P0 = eos_->P0;
        p.e0 = eos_->e0;
        check(zfv_smf_configure(mesh_->context(), &p));
    }

    Eos::Ptr eos_;
    EuMesh* mesh_ = nullptr;
    double cfl_ = 0.5, max_dt_ = std::numeric_limits<double>::max();
    int acc_ = 1, limiter_ = ZFV_LIM_MC;
    bool axial_ = false;
    Fluxes flux_ = Fluxes::HLLC;
};

}  // namespace host
}  // namespace zfv
