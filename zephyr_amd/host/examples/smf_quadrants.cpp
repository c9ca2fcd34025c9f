// Two-dimensional four-quadrant Riemann problem on an adaptive Rectangle
// mesh, written as the reference's problem files are
// (problems/single-fluid/smf_test_2D.cpp): generator, mesh, solver setters,
// initial adaptation, then update / set_flags / refine per step.
//
// usage: smf_quadrants [nx] [steps] [max_level] [seed]
// Prints the number of cells, the last dt and the conserved totals.
#include <zfv.hpp>

#include <cstdint>
#include <cstdio>
#include <cstdlib>

using namespace zfv::host;

namespace {

// SplitMix64, u = (x >> 11) * 2^-53.
struct SplitMix64 {
    uint64_t x;
    double uniform() {
        x += 0x9E3779B97F4A7C15ULL;
        uint64_t z = x;
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
        z ^= z >> 31;
        return (double)(z >> 11) * (1.0 / 9007199254740992.0);
    }
};

}  // namespace

int main(int argc, char** argv) {
    const int nx = argc > 1 ? std::atoi(argv[1]) : 200;
    const int steps = argc > 2 ? std::atoi(argv[2]) : 100;
    const int max_level = argc > 3 ? std::atoi(argv[3]) : 2;
    const uint64_t seed = argc > 4 ? std::strtoull(argv[4], nullptr, 10) : 2024;

    try {
        auto eos = IdealGas::create(1.4);

        // Quadrant states q = 1..4 (x>=1/2,y>=1/2), (x<=1/2,y>=1/2), (x<=1/2,y<=1/2), (x>=1/2,y<=1/2).
        PState q[4];
        SplitMix64 g{seed};
        for (auto& s : q) {
            s.density = 0.1 + 1.9 * g.uniform();
            s.vx = -1.0 + 2.0 * g.uniform();
            s.vy = -1.0 + 2.0 * g.uniform();
            s.pressure = 0.1 + 1.9 * g.uniform();
            s.energy = eos->energy_rP(s.density, s.pressure);
        }
        auto initial = [&q](double x, double y) -> PState {
            if (x >= 0.5 && y >= 0.5) return q[0];
            if (x <= 0.5 && y >= 0.5) return q[1];
            if (x <= 0.5 && y <= 0.5) return q[2];
            return q[3];
        };

        Rectangle gen(0.0, 1.0, 0.0, 1.0);
        gen.set_boundaries({Boundary::ZOE, Boundary::ZOE, Boundary::ZOE, Boundary::ZOE});
        gen.set_nx(nx);

        EuMesh mesh(gen);

        SmFluid solver(eos);
        solver.set_accuracy(2);
        solver.set_CFL(0.5);
        solver.set_limiter("MC");
        solver.set_method(Fluxes::HLLC);
        solver.add_types(mesh);

        mesh.set_max_level(max_level);
        mesh.set_distributor(solver.distributor("slope"));

        for (int k = 0; k < mesh.max_level() + 1; ++k) {
            solver.set_initial(mesh, initial);
            solver.set_flags(mesh);
            mesh.refine();
        }
        solver.set_initial(mesh, initial);

        double time = 0.0;
        for (int step = 0; step < steps; ++step) {
            solver.update(mesh);
            time += solver.dt();
            if (max_level > 0) {
                solver.set_flags(mesh);
                mesh.refine();
            }
        }
        auto t = solver.totals(mesh);
        std::printf("cells %d time %.17g dt %.17g totals %.17g %.17g %.17g %.17g\n", mesh.n_cells(), time,
                    solver.dt(), t[0], t[1], t[2], t[3]);
    } catch (const std::exception& e) {
        std::fprintf(stderr, "smf_quadrants: %s\n", e.what());
        return 1;
    }
    return 0;
}
