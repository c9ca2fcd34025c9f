// ORACLE — test infrastructure only.
//
// A plain-C++ CPU restatement of zephyr's single-material finite-volume step
// (SmFluid) on AMR-2D meshes.  It is the checker for the HIP product path:
// only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
// load it.  The product library never links or calls it.
//
// Every function cites the reference file:line whose behaviour it restates.
// Operation order: where the reference's order is fixed by Eigen expression
// templates, the oracle fixes one documented order (see ORDER notes in
// oracle.cpp).  On Cartesian AMR meshes (axis-aligned normals, z = 0) every
// such ambiguous sum has at most one non-zero term, so the chosen order does
// not change any value.
//
// Parity pins (tests/test_oracle_*.py): EOS closed forms and round trips,
// flux consistency F(z,z) = F_phys(z), supersonic early-outs, mirror
// antisymmetry, HLL/Rusanov closed forms, exact Riemann star states against
// Toro's published tables, the Limiter::check table, LSM exactness on linear
// fields, uniform-state preservation, row symmetry, conservation, Sod
// convergence against the exact solution, and brute-force hand-computed
// 2x2 steps.
#pragma once
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

// Numerical flux ids: the values of zephyr::math::Fluxes
// (reference zephyr/math/cfd/flux/num_flux.h:11-21).
/* Fluxes values (math/cfd/flux/num_flux.h:11-21) */
enum { ORC_GODUNOV = 0, ORC_RUSANOV = 1, ORC_HLL = 2, ORC_HLLC = 3, ORC_HLLC_M = 4, ORC_HLLC_LM = 5, ORC_CIR1 = 6,
       ORC_CIR2 = 7 };

// Limiter ids (reference zephyr/math/cfd/limiter.cxx:163-252 name table).
enum {
    ORC_LIM_NONE = 0, ORC_LIM_UNIT = 1, ORC_LIM_MINMOD = 2, ORC_LIM_MC = 3,
    ORC_LIM_VANLEER = 4, ORC_LIM_VANALBADA1 = 5, ORC_LIM_VANALBADA2 = 6,
    ORC_LIM_CHARM = 7, ORC_LIM_HCUS = 8, ORC_LIM_HQUICK = 9, ORC_LIM_KOREN = 10,
    ORC_LIM_SMART = 11, ORC_LIM_SUPERBEE = 12
};

// Boundary ids: the values of zephyr::geom::Boundary (reference geom/boundary.h:9-16).
enum { ORC_UNDEFINED = -1, ORC_INNER = 0, ORC_WALL = 1, ORC_ZOE = 2, ORC_PERIODIC = 3 };

// ---- point functions (known-answer tests) --------------------------------
// z = {rho, vx, vy, vz, p, e} (smf::PState order), eos = {gamma, P0, e0}.
// Returns 0, or 1 if the reference would have thrown (NaN flux).
int orc_flux(int kind, const double* zL, const double* zR, const double* eos, double* F);
double orc_limiter2(int kind, double a, double b);
double orc_limiter1(int kind, double r);
// Exact Riemann solver (riemann.cpp:366-460).  out = {rho, Uf, Pf, P, U, conv}.
void orc_riemann(double rL, double uL, double pL, double gL, double p0L,
                 double rR, double uR, double pR, double gR, double p0R, double* out);
// Exact self-similar solution at xi = x/t (riemann.cpp density/velocity/pressure
// samplers): out = {rho, u, p}.
void orc_riemann_sample(double rL, double uL, double pL, double rR, double uR, double pR,
                        double gamma, double p0, double xi, double* out);
// EOS: out = {p(rho,e), e(rho,p), c(rho,p), min_pressure}.
void orc_eos(const double* eos, double rho, double e, double p, double* out);

// ---- mesh ----------------------------------------------------------------
typedef struct orc_mesh orc_mesh;
orc_mesh* orc_mesh_rectangle(int nx, int ny, double xmin, double xmax,
                             double ymin, double ymax, const int* bounds /*L,R,B,T*/);
// The same with the axisymmetric geometry when axial != 0: y-weighted
// volume and face areas, y-shifted cell and side-face centres, a bottom on
// the axis (ymin = 0) made a wall (Rectangle::set_axial).
orc_mesh* orc_mesh_rectangle_ax(int nx, int ny, double xmin, double xmax,
                                double ymin, double ymax, const int* bounds, int axial);
int orc_mesh_axial(const orc_mesh* m);
// volume_alt (n) and area_alt (8n) of an axial mesh.
void orc_mesh_get_axial(const orc_mesh* m, double* vol_alt, double* area_alt);
void orc_mesh_free(orc_mesh* m);
int  orc_mesh_size(const orc_mesh* m);
// Cell arrays (n): cx, cy, volume, level, flag, b_idx, z_idx.
void orc_mesh_get_cells(const orc_mesh* m, double* cx, double* cy, double* vol,
                        int* level, int* flag, int* b_idx, int* z_idx);
// Slot arrays (8n, slot s of cell c at 8c+s): boundary, adj, normal, centre, area.
void orc_mesh_get_slots(const orc_mesh* m, int* boundary, int* adj, double* nx, double* ny,
                        double* fcx, double* fcy, double* area);
void orc_mesh_set_slots(orc_mesh* m, const int* boundary, const int* adj, const double* nx, const double* ny,
                        const double* fcx, const double* fcy, const double* area);
// Vertex arrays (9n, SqQuad order).
void orc_mesh_get_verts(const orc_mesh* m, double* vx, double* vy);
// Incircle diameter per cell (amr_cells.cpp:107-114).
void orc_mesh_incircle(const orc_mesh* m, double* h);

// ---- solver (SmFluid restatement) ----------------------------------------
typedef struct orc_smf orc_smf;
orc_smf* orc_smf_create(orc_mesh* m, double gamma, double P0, double e0,
                        int flux, int limiter, int acc, double cfl);
void orc_smf_free(orc_smf* s);
void orc_smf_set_max_dt(orc_smf* s, double max_dt);
// SmFluid::set_axial; -1 if the mesh has no axial geometry.
int orc_smf_set_axial(orc_smf* s, int axial);
void orc_smf_plant_axial(int k);   /* test-only planted mistakes of the axial step, 0 = none */
void orc_smf_set_state(orc_smf* s, const double* rho, const double* u, const double* v,
                       const double* p, const double* e);
void orc_smf_get_state(const orc_smf* s, double* rho, double* u, double* v,
                       double* p, double* e);
void orc_smf_get_wait(const orc_smf* s, int* wait);
void orc_smf_set_wait(orc_smf* s, const int* wait);
// Gradient layers d_dx, d_dy for components {rho, vx, vy, p} (4n each).
void orc_smf_get_grad(const orc_smf* s, double* dx, double* dy);
// One SmFluid::update.  Returns 0, or 1 on a reference-thrown numeric error.
int orc_smf_update(orc_smf* s, int nthreads);
double orc_smf_dt(const orc_smf* s);
// Stage entry points for stage-level tests.
double orc_smf_compute_dt(orc_smf* s, int nthreads);
void orc_smf_compute_grad(orc_smf* s, int nthreads);
// Conserved totals {mass, mom_x, mom_y, energy} = sum V*q.
void orc_smf_totals(const orc_smf* s, double* out);

// ---- AMR (oracle_amr.hpp) ---------------------------------------------------
// SmFluid::set_flags: criterion 0 = relative jumps (sm_fluid.cpp:517-643),
// 1 = normalised limited density gradient |grad rho| h / rho vs threshold.
void orc_smf_set_flags(orc_smf* s, int criterion, double threshold, int nthreads);
void orc_mesh_set_flags(orc_mesh* m, const int* flags);
void orc_mesh_balance_flags(orc_mesh* m, int max_level);
// EuMesh::refine: balance + apply; s may be NULL (mesh only); slope: distributor.
// Returns 1 changed, 0 unchanged, -2 for an axial mesh (not adapted here).
int orc_mesh_refine(orc_mesh* m, orc_smf* s, int max_level, int slope);
void orc_mesh_get_next(const orc_mesh* m, int* next);

// ---- synthetic inputs -----------------------------------------------------
// SplitMix64 uniform stream, u = (x >> 11) * 2^-53.
double orc_splitmix_uniform(uint64_t* state);
// Four-quadrant Riemann data drawn from `seed` (quadrants numbered as
// Riemann2D, phys/tests/test_2D.cpp:72-106); evaluated at cell centres.
void orc_init_quadrants(const orc_mesh* m, uint64_t seed, double gamma,
                        double* rho, double* u, double* v, double* p, double* e);

/* Multi-material solver (MmFluid, first order, CrpMode::NONE) with the
 * PT-equilibrium closure of stiffened gases; mats: 6 doubles per material
 * {gamma, P0, e0, Cv, T0, rho_ref}; per-material arrays are [mat][n]. */
typedef struct orc_mm orc_mm;
orc_mm* orc_mm_create(orc_mesh* m, int n_mat, const double* mats, double cfl);
void orc_mm_free(orc_mm* s);
void orc_mm_set_max_dt(orc_mm* s, double v);
void orc_mm_set_crp(orc_mm* s, int mode);   /* CrpMode: 0 NONE, 1 V3, 2 V5, 3 VS, 4 PLIC (default) */
void orc_mm_set_flux(orc_mm* s, int flux);   /* Fluxes: 1 Rusanov, 2 HLL, 3 HLLC (default) */
void orc_mm_set_splitting(orc_mm* s, int split);   /* DirSplit: 0 NONE, 1 SIMPLE, 2 STRANG */
void orc_mm_set_state(orc_mm* s, const double* rho, const double* u, const double* v, const double* P,
                      const double* beta);
void orc_mm_state(const orc_mm* s, double* rho, double* u, double* v, double* P, double* e, double* T,
                  double* beta, double* rhos);
double orc_mm_compute_dt(orc_mm* s);
int orc_mm_update(orc_mm* s);
double orc_mm_dt(const orc_mm* s);
void orc_mm_interface(orc_mm* s, double* nx, double* ny, double* p);
void orc_mm_closure(int n_mat, const double* mats, double rho, double e, const double* beta, double P0,
                    double T0, double* out);
void orc_mm_closure_rP(int n_mat, const double* mats, double rho, double P, const double* beta, double* out);
void orc_mm_face_flux(const orc_mm* s, int cL, int cR, double nx, double ny, double* F);
int orc_mm_face_flux_kind(const orc_mm* s, int kind, int cL, int cR, double nx, double ny, double* F);
void orc_mm_stiffened_gas(const orc_mm* s, int c, double* out);   /* {gamma, P0} of cell c's state */
double orc_mm_sound_speed(const orc_mm* s, int c);
int orc_mm_crp_flux(const orc_mm* s, int cL, int cR, double hL, double hR, int iA, double a_sig, double dt,
                    double* F);
void orc_mm_wave_speeds(const orc_mm* s, int cL, int cR, double* S3);
int orc_mm_crp_sub(const orc_mm* s, int inverse, int c0, int c1, int c2, double delta, double dt, double* F);
void orc_mm_split(const orc_mm* s, int c, int iA, double* out);
void orc_mm_plant(int k);   /* test-only planted mistakes in the composite-flux chain, 0 = none */
void orc_mm_set_limiter(orc_mm* s, int limiter);
void orc_mm_set_accuracy(orc_mm* s, int acc);   /* 1 or 2 (clamped) */
void orc_mm_set_threads(orc_mm* s, int nthreads);   /* OpenMP threads of the per-cell loops (default 1) */
void orc_mm_compute_grad(orc_mm* s);
void orc_mm_grad(const orc_mm* s, double* gx, double* gy);   /* [6 + 2 n_mat][n]: rho, u, v, e, beta, P, T, rhos */
void orc_mm_set_flags(orc_mm* s);
int orc_mm_refine(orc_mm* s, int max_level);   /* 1 changed, 0 unchanged, -1 a rejected state */
double orc_quad_volume_fraction(double p, double nx, double ny, double A, double B);
double orc_quad_find_section(double alpha, double nx, double ny, double A, double B);
double orc_face_fraction(int kind, double a1, double a2);
double orc_average_flux(double alpha, double cs, double CFL);

/* Transfer (math/solver/transfer.h): a volume fraction carried by a velocity
 * sampled at the face centres (fvx, fvy: 8n, slot s of cell c at 8c + s)
 * and cell centres (cvx, cvy: n).  Methods (Transfer::Method numbers):
 * 0 CRP_V3, 1 CRP_V5, 2 CRP_SE, 3 CRP_N1, 7 MUSCLd, 9 MUSCLd_CRP,
 * 11 MUSCL_MC, 12 MUSCL_MC_CRP; create returns NULL for the others.
 * dir: 0 ANY, 1 X, 2 Y. */
typedef struct orc_tr orc_tr;
orc_tr* orc_tr_create(orc_mesh* m, int method, double cfl);
void orc_tr_free(orc_tr* T);
void orc_tr_set_velocity(orc_tr* T, const double* fvx, const double* fvy, const double* cvx, const double* cvy);
void orc_tr_set_state(orc_tr* T, const double* u);
void orc_tr_state(const orc_tr* T, double* u, double* nx, double* ny);
void orc_tr_set_dt(orc_tr* T, double dt);
double orc_tr_compute_dt(const orc_tr* T);
void orc_tr_update(orc_tr* T, int dir);
double orc_weno5(const double* u, int right);
void orc_tr_slopes(const orc_tr* T, double* dx, double* dy);
void orc_tr_set_flags(orc_tr* T);   /* into the mesh's flags */

/* Three-dimensional Cuboid meshes and the SmFluid step on them
 * (oracle_3d.hpp); 24 face slots per cell, cell-major. */
typedef struct orc_mesh3 orc_mesh3;
typedef struct orc_smf3 orc_smf3;
orc_mesh3* orc_mesh3_cuboid(int nx, int ny, int nz, double xmin, double xmax, double ymin, double ymax,
                            double zmin, double zmax, const int* bounds);
void orc_mesh3_free(orc_mesh3* m);
int orc_mesh3_size(const orc_mesh3* m);
void orc_mesh3_get_cells(const orc_mesh3* m, double* cx, double* cy, double* cz, double* vol, double* h);
void orc_mesh3_get_slots(const orc_mesh3* m, int* boundary, int* adj, double* nx, double* ny, double* nz,
                         double* fcx, double* fcy, double* fcz, double* area, double* w);
orc_smf3* orc_smf3_create(orc_mesh3* m, double gamma, double P0, double e0, int flux, int limiter, int acc,
                          double cfl);
void orc_smf3_free(orc_smf3* s);
void orc_smf3_set_state(orc_smf3* s, const double* rho, const double* u, const double* v, const double* w,
                        const double* p, const double* e);
void orc_smf3_get_state(const orc_smf3* s, double* rho, double* u, double* v, double* w, double* p, double* e);
void orc_smf3_get_wait(const orc_smf3* s, int* wait);
void orc_smf3_set_wait(orc_smf3* s, const int* wait);
void orc_smf3_get_grad(const orc_smf3* s, double* g);
double orc_smf3_compute_dt(orc_smf3* s, int nthreads);
int orc_smf3_update(orc_smf3* s, int nthreads);
double orc_smf3_dt(const orc_smf3* s);
void orc_smf3_totals(const orc_smf3* s, double* out);
void orc_rotate3(const double* v, const double* n, int to_local, double* out);

#ifdef __cplusplus
}
#endif
