// Internal device-data layout shared by the kernels and the C ABI.
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>

#include "zfv_physics.hpp"

namespace zfv {

// Device mesh: cell arrays [cap], slot arrays slot-major [8][cap] (slot s
// of cell c at s*stride + c) so that a wavefront of 64 consecutive cells
// reads slot s of each cell from 64 consecutive addresses.  code[c] packs the
// 8 slot nibbles: bits 0-2 the slot type (SlotCode), bit 3 set when the
// slot's normal is not exactly the canonical side normal (AMR faces whose
// vertices carry round-off); such slots read nx/ny and use the general
// rotation.
struct DevMesh {
    int n;
    int stride;
    const double* __restrict__ cx;
    const double* __restrict__ cy;
    const double* __restrict__ vol;
    const double* __restrict__ h;       // incircle diameter
    const int* __restrict__ code;
    const int* __restrict__ adj;        // [8][stride]
    const double* __restrict__ area;    // [8][stride]
    const double* __restrict__ fcx;     // [8][stride]
    const double* __restrict__ fcy;     // [8][stride]
    const double* __restrict__ w;       // [8][stride] LSM weight area / |drn|^3
    const double* __restrict__ nx;      // [8][stride] outward normal
    const double* __restrict__ ny;
};

constexpr int kGeneralNormal = 8;

struct DevState {
    double* r;
    double* u;
    double* v;
    double* p;
    double* e;
};

struct DevGrad {
    double* dx;   // [4][stride]: rho, u, v, p
    double* dy;
};

// Per-launch step control: dt slots hold min_c h/(|v|+c) as the bit pattern
// of a positive double (ordered like uint64), reduced with atomicMin.
struct StepCtl {
    unsigned long long* dtmin;   // 3 rotating slots
    int rd, wr, clr;             // slot read for this step's dt, written for the next, cleared
    double cfl, max_dt;
    double* dt_out;              // dt of this step
    int* err;                    // sticky numeric-error flag
};

// Launchers (zfv_kernels.hip).
hipError_t launch_fill_u64(unsigned long long* p, int count, unsigned long long v, hipStream_t s);
hipError_t launch_dt(const DevMesh& M, const DevState& z, const Eos& eos, unsigned long long* slot, hipStream_t s);
hipError_t launch_flux1(int flux, const DevMesh& M, const DevState& in, const DevState& out,
                        const Eos& eos, const StepCtl& ctl, hipStream_t s);
hipError_t launch_grad_pred(int limiter, bool predictor, const DevMesh& M, const DevState& in, const DevGrad& g,
                            const DevState& half, const int* wait, const Eos& eos, const StepCtl& ctl,
                            hipStream_t s);
hipError_t launch_corr(int flux, const DevMesh& M, const DevState& init, const DevState& half,
                       const DevGrad& g, const int* wait_in, int* wait_out, const DevState& next,
                       const Eos& eos, const StepCtl& ctl, hipStream_t s);
hipError_t launch_totals(const DevMesh& M, const DevState& z, double* partial, int nblocks, hipStream_t s);

constexpr int kBlock = 256;

}  // namespace zfv
