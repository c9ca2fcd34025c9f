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
    // A Cuboid's geometry as 1D tables (tnx > 0): every centre, face-centre
    // component and least-squares weight depends on one index only, so the
    // kernels read tables of nx + ny + nz entries instead of the slot arrays
    // (whose values the host checked equal to the tables').  Layout, with
    // L = tnx + tny + tnz and base(a) = 0, tnx, tnx + tny: face centre
    // component a of side s at (3 s + a) L + base(a) + index_a; cell centre
    // component a at 18 L + base(a) + index_a; weight of side s at
    // 19 L + s L + base(s / 2) + index_{s/2}; area of side s at 25 L + s;
    // volume at 25 L + 6.  bcode: the slot code of each side at the edge;
    // per: periodic directions.
    int tnx, tny, tnz;
    const double* tab;
    int bcode[6];
    int per[3];
    const double *cx, *cy, *cz, *vol, *h;
    const int *type, *adj;
    const double *fcx, *fcy, *fcz, *area, *w;
};

// Primitive state {rho, u, v, w, p, e}, one array of n doubles per field.
struct Dev3State {
    double *r, *u, *v, *w, *p, *e;
};

// Step control, three rotating CFL slots of kSlotWords words (the bits of
// positive doubles, reduced with atomicMin): dtmin is the slot this step reads; the stage
// that ends a step reduces the new state's h / (|v| + c) into `next`, the
// slot the following step reads, and sets `reset`, the slot the previous
// step read, back to the largest finite double.
struct Dev3Ctl {
    unsigned long long* dtmin;
    unsigned long long* next;
    unsigned long long* reset;
    double cfl, max_dt;
    double* dt_out;
    int* err;
};

constexpr unsigned long long kDblMaxBits = 0x7fefffffffffffffULL;
// A CFL slot is kSubSlots partial minima, kSubStride words (128 bytes) apart,
// so that the workgroups' atomics spread over as many cache lines; the
// slot's value is their least.
constexpr int kSubSlots = 64;
constexpr int kSubStride = 16;
constexpr int kSlotWords = kSubSlots * kSubStride;

// The number of doubles of a Cuboid's geometry tables.
inline size_t tab3_size(int nx, int ny, int nz) { return 25 * (size_t)(nx + ny + nz) + 7; }

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
