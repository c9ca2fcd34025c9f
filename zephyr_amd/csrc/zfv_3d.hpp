// Device views of a three-dimensional Cuboid mesh and its SmFluid state,
// shared by the 3D kernels (zfv_3d.hip) and their host side (zfv_3d.cpp).
#pragma once

#include <hip/hip_runtime.h>

#include "zfv_physics.hpp"

namespace zfv {

// n cells; six face slots per cell (sides L, R, B, T, Z, F), slot s of cell
// c at s * stride + c.  type: SlotCode; adj: the neighbour (the cell itself
// on a non-periodic boundary); w: the least-squares weight of the slot.
struct Dev3Mesh {
    int n, stride;
    const double *cx, *cy, *cz, *vol, *h;
    const int *type, *adj;
    const double *fcx, *fcy, *fcz, *area, *w;
};

// Primitive state {rho, u, v, w, p, e}, one array of n doubles per field.
struct Dev3State {
    double *r, *u, *v, *w, *p, *e;
};

// Step control: dtmin is the CFL slot this step reads (the bits of a
// positive double, reduced with atomicMin); the stage that ends a step sets
// `reset`, the next step's slot, back to the largest finite double.
struct Dev3Ctl {
    unsigned long long* dtmin;
    unsigned long long* reset;
    double cfl, max_dt;
    double* dt_out;
    int* err;
};

constexpr unsigned long long kDblMaxBits = 0x7fefffffffffffffULL;

hipError_t launch3_dt(const Dev3Mesh& M, const Dev3State& z, const Eos& eos, unsigned long long* slot,
                      hipStream_t s);
hipError_t launch3_flux1(int flux, const Dev3Mesh& M, const Dev3State& in, const Dev3State& out, const Eos& eos,
                         const Dev3Ctl& ctl, hipStream_t s);
hipError_t launch3_grad_pred(int limiter, const Dev3Mesh& M, const Dev3State& in, double* G, const Dev3State& half,
                             const int* wait, const Eos& eos, const Dev3Ctl& ctl, hipStream_t s);
hipError_t launch3_corr(int flux, const Dev3Mesh& M, const Dev3State& init, const Dev3State& half, const double* G,
                        const int* wait_in, int* wait_out, const Dev3State& next, const Eos& eos,
                        const Dev3Ctl& ctl, hipStream_t s);
hipError_t launch3_totals(const Dev3Mesh& M, const Dev3State& z, double* partial, int nblocks, hipStream_t s);

}  // namespace zfv
