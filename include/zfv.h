/* zfv.h — C ABI of the MI355X-native zephyr finite-volume step.
 *
 * Plain pointers and sizes, int status returns (0 = ok), no exceptions and
 * no torch types across the boundary.  One context owns one device mesh, one
 * solver state and one HIP stream; a context is not thread-safe.
 *
 * The reference has no C ABI: its hot path sits behind C++ classes.  Each
 * entry point below names the reference interface it replaces.  A host C++
 * facade with the reference's class shapes (zephyr_amd/host/zfv.hpp:
 * EuMesh, SmFluid, Rectangle, IdealGas/StiffenedGas) is built on these calls.
 */
#ifndef ZFV_H
#define ZFV_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ------------------------------------------------------ */
enum {
    ZFV_OK = 0,
    ZFV_ERR_ARG = 1,         /* bad argument (size, null pointer, enum value)         */
    ZFV_ERR_HIP = 2,         /* HIP runtime failure (allocation, launch, copy)        */
    ZFV_ERR_NUMERIC = 3,     /* a NaN flux: where the reference throws
                                (hllc.cpp:72-80, hll.cpp:46-54, riemann.cpp:222-229)  */
    ZFV_ERR_UNSUPPORTED = 4, /* input outside the supported mesh class                */
    ZFV_ERR_STATE = 5        /* call order (no mesh, no state, not configured)        */
};

/* Numerical flux ids = zephyr::math::Fluxes values (math/cfd/flux/num_flux.h:11-21).
 * The single-material step takes GODUNOV..CIR2 (HLLC_M / HLLC_LM:
 * hllc.cpp:427-572, CIR1 / CIR2: cir.cpp:16-216); CRP is MmFluid's. */
enum { ZFV_GODUNOV = 0, ZFV_RUSANOV = 1, ZFV_HLL = 2, ZFV_HLLC = 3, ZFV_HLLC_M = 4, ZFV_HLLC_LM = 5,
       ZFV_CIR1 = 6, ZFV_CIR2 = 7, ZFV_CRP = 8 };

/* MmFluid's CrpMode (math/solver/mm_fluid.h:36-42) and DirSplit (:29-33). */
enum { ZFV_CRP_NONE = 0, ZFV_CRP_V3 = 1, ZFV_CRP_V5 = 2, ZFV_CRP_VS = 3, ZFV_CRP_PLIC = 4, ZFV_CRP_MUSCL = 5 };
enum { ZFV_SPLIT_NONE = 0, ZFV_SPLIT_SIMPLE = 1, ZFV_SPLIT_STRANG = 2 };

/* Limiter ids; names parsed as Limiter(const std::string&) (math/cfd/limiter.cxx:163-252). */
enum {
    ZFV_LIM_NONE = 0, ZFV_LIM_UNIT = 1, ZFV_LIM_MINMOD = 2, ZFV_LIM_MC = 3,
    ZFV_LIM_VANLEER = 4, ZFV_LIM_VANALBADA1 = 5, ZFV_LIM_VANALBADA2 = 6,
    ZFV_LIM_CHARM = 7, ZFV_LIM_HCUS = 8, ZFV_LIM_HQUICK = 9, ZFV_LIM_KOREN = 10,
    ZFV_LIM_SMART = 11, ZFV_LIM_SUPERBEE = 12
};

/* Boundary ids = zephyr::geom::Boundary values (geom/boundary.h:9-16). */
enum { ZFV_UNDEFINED = -1, ZFV_INNER = 0, ZFV_WALL = 1, ZFV_ZOE = 2, ZFV_PERIODIC = 3, ZFV_FUNCTION = 4 };

/* Host-side AMR-2D mesh in the reference's AmrCells layout
 * (mesh/euler/amr_cells.h:74-104, amr_faces.h:83-101): n cells, 8 face slots
 * per cell (slot s of cell c at index 8*c + s, sides L,R,B,T then their
 * second sub-faces), 9 vertices per cell in SqQuad order (geom/primitives/quad.h:186-191).
 * Normals are arbitrary unit vectors in the plane; slots whose normal is
 * exactly the canonical side normal (-e_x, +e_x, -e_y, +e_y for L, R, B, T)
 * take an exact specialised path.  The adaptation entry points assume a
 * Cartesian base (adjacency rotation 0, as Rectangle builds). */
typedef struct {
    int32_t n_cells;
    const double* cx;        /* [n] cell centre                                    */
    const double* cy;
    const double* volume;    /* [n]                                                */
    const double* vx;        /* [9n] vertices (incircle diameter, AMR geometry)    */
    const double* vy;
    const int32_t* level;    /* [n] may be NULL (all 0)                            */
    const int32_t* flag;     /* [n] may be NULL (all 0)                            */
    const int32_t* b_idx;    /* [n] may be NULL (identity)                         */
    const int32_t* z_idx;    /* [n] may be NULL (all 0)                            */
    const int32_t* boundary; /* [8n] Boundary id, -1 = undefined slot               */
    const int32_t* adj;      /* [8n] neighbour cell (the cell itself on boundaries) */
    const double* nx;        /* [8n] outward normal                                */
    const double* ny;
    const double* fcx;       /* [8n] face centre                                   */
    const double* fcy;
    const double* area;      /* [8n]                                               */
} zfv_mesh_soa;

/* Solver parameters: SmFluid setters (math/solver/sm_fluid.h:40-71, sm_fluid.cpp:16-67)
 * and a stiffened-gas EOS {gamma, P0, e0} (IdealGas = P0 = e0 = 0). */
typedef struct {
    int32_t acc;        /* 1 or 2 (clamped as set_accuracy)                 */
    int32_t flux;       /* ZFV_GODUNOV..ZFV_CIR2                             */
    int32_t limiter;    /* ZFV_LIM_*                                         */
    int32_t axial;      /* set_axial (sm_fluid.cpp:45-47): 1 takes face.area(true),
                           cell.volume(true) and the pressure source of
                           sm_fluid.cpp:166-169; needs zfv_mesh_rectangle_axial */
    double cfl;         /* clamped to [0,1] as set_CFL                      */
    double max_dt;      /* set_max_dt                                        */
    double gamma, P0, e0;
} zfv_smf_params;

typedef struct zfv_ctx zfv_ctx;

/* Text of the last error on this thread. */
const char* zfv_last_error(void);
/* Library version string. */
const char* zfv_version(void);

/* Create / destroy a context on HIP device `device`.  No reference counterpart
 * (EuMesh + SmFluid construction, eu_mesh.cpp:113-115, sm_fluid.cpp:16-35). */
int zfv_create(int device, zfv_ctx** out);
int zfv_destroy(zfv_ctx* ctx);
/* Run the context's work on an external HIP stream (hipStream_t as void*);
 * NULL restores the context's own stream. */
int zfv_set_stream(zfv_ctx* ctx, void* stream);
int zfv_synchronize(zfv_ctx* ctx);

/* Upload a host AmrCells-layout mesh (replaces EuMesh(Generator&), eu_mesh.cpp:70-104). */
int zfv_mesh_upload(zfv_ctx* ctx, const zfv_mesh_soa* mesh);
/* Build the Rectangle AMR-2D base mesh on the host and upload it
 * (replaces generator::Rectangle + Rectangle::initialize, rectangle.cpp:465-...).
 * bounds = {left, right, bottom, top} Boundary ids. */
int zfv_mesh_rectangle(zfv_ctx* ctx, int nx, int ny, double xmin, double xmax,
                       double ymin, double ymax, const int32_t* bounds);
/* The same Rectangle with set_axial(true) (rectangle.cpp:97-103, 590-602):
 * y is the radius (ymin >= 0; a bottom on the axis, ymin = 0, becomes a
 * wall), the cells carry volume_alt = hx hy y_c and the faces area_alt =
 * length * y_face (the y-weighted volume and areas of the body of
 * revolution per radian), and the cell and left / right face centres move
 * to their y-weighted centroids (y + hy^2 / (12 y)).  Such a mesh is not
 * adapted (zfv_refine returns ZFV_ERR_UNSUPPORTED). */
int zfv_mesh_rectangle_axial(zfv_ctx* ctx, int nx, int ny, double xmin, double xmax,
                             double ymin, double ymax, const int32_t* bounds);
/* zfv_mesh_upload of a mesh with the axisymmetric geometry (an axial
 * EuMesh, amr_cells.h:288-290, amr_faces.h:183-185): volume_alt (n, > 0)
 * and area_alt (8n); the centres in `mesh` are the y-weighted ones. */
int zfv_mesh_upload_axial(zfv_ctx* ctx, const zfv_mesh_soa* mesh, const double* vol_alt, const double* area_alt);
/* volume_alt (n) and area_alt (8n, slot s of cell c at 8c + s) of an
 * axisymmetric mesh (AmrCells::volume_alt, AmrFaces::area_alt). */
int zfv_mesh_axial_download(zfv_ctx* ctx, double* vol_alt, double* area_alt);
/* Number of cells of the current mesh (EuMesh::n_cells). */
int zfv_mesh_n_cells(zfv_ctx* ctx, int32_t* n);
/* Number of unique faces of the current mesh (interior face sides / 2 +
 * boundary face sides; a hanging face counts once per fine side), the phi =
 * faces / cells of the bandwidth accounting.  No reference counterpart. */
int zfv_mesh_n_faces(zfv_ctx* ctx, int64_t* n);
/* Reserve device capacity for meshes of up to n_cells cells
 * (AmrCells::reserve_amr, mesh/euler/amr_cells.cpp:566-575): every
 * per-cell array, the adaptation's scratch and the curve-ordered copy of
 * the mesh are sized for n_cells, so adaptations that stay below it
 * allocate nothing.  Without it an adaptation that outgrows the capacity
 * doubles it (one synchronisation).  Not available on a batch. */
int zfv_mesh_reserve(zfv_ctx* ctx, int n_cells);
/* Download the current mesh into caller arrays sized as zfv_mesh_soa
 * (pointers of a zfv_mesh_soa used as outputs; NULL members are skipped). */
int zfv_mesh_download(zfv_ctx* ctx, const zfv_mesh_soa* out);

/* Configure the solver (SmFluid(Eos) + setters + add_types, sm_fluid.cpp:16-67). */
int zfv_smf_configure(zfv_ctx* ctx, const zfv_smf_params* params);
/* The solver's parameters as set (SmFluid's getters: CFL(), ...). */
int zfv_smf_get_params(zfv_ctx* ctx, zfv_smf_params* out);
/* set_max_dt (sm_fluid.cpp:65-67). */
int zfv_smf_set_max_dt(zfv_ctx* ctx, double max_dt);

/* Primitive state of the "init" layer, SoA host arrays of n doubles
 * (cell[part.init] = PState{rho, {u, v, 0}, p, e}, sm_fluid.cpp:25-35). */
int zfv_state_upload(zfv_ctx* ctx, const double* rho, const double* u, const double* v,
                     const double* p, const double* e);
int zfv_state_download(zfv_ctx* ctx, double* rho, double* u, double* v, double* p, double* e);
/* Same, with device pointers (device-to-device copies on the context stream). */
int zfv_state_upload_device(zfv_ctx* ctx, const double* rho, const double* u, const double* v,
                            const double* p, const double* e);
int zfv_state_download_device(zfv_ctx* ctx, double* rho, double* u, double* v, double* p, double* e);
/* Isolation counters (Parts::wait, sm_fluid.h:34-35). */
int zfv_wait_upload(zfv_ctx* ctx, const int32_t* wait);
int zfv_wait_download(zfv_ctx* ctx, int32_t* wait);
/* Limited gradient layers d_dx, d_dy of {rho, u, v, p}: 4n doubles each,
 * component-major (compute_grad, sm_fluid.cpp:116-125). */
int zfv_grad_download(zfv_ctx* ctx, double* d_dx, double* d_dy);

/* `n_steps` calls of SmFluid::update (sm_fluid.cpp:81-102), enqueued without
 * a host synchronisation.  Numeric errors are sticky on the device and are
 * reported by the next synchronising call (dt, download, synchronize). */
int zfv_smf_step(zfv_ctx* ctx, int n_steps);
/* SmFluid::dt(): dt of the last step (synchronises). */
int zfv_smf_dt(zfv_ctx* ctx, double* dt);
/* Only the gradient stage (compute_grad) on the current state. */
int zfv_smf_compute_grad(zfv_ctx* ctx);
/* Only the CFL reduction (compute_dt) on the current state. */
int zfv_smf_compute_dt(zfv_ctx* ctx, double* dt);

/* ---- adaptation (AMR-2D, one process, Cartesian base) ------------------- */
/* Refinement flags per cell (AmrCells::flag, amr_cells.h:81): upload / download. */
int zfv_flags_upload(zfv_ctx* ctx, const int32_t* flags);
int zfv_flags_download(zfv_ctx* ctx, int32_t* flags);
/* SmFluid::set_flags (sm_fluid.cpp:517-643): limited gradient, then flags.
 * criterion 0: the reference's relative jumps of density and pressure
 * (xi = 0.05, hysteresis 0.4); criterion 1: |grad rho| h / rho of the
 * unlimited LSM gradient against `threshold` (+1 above, 0 above 0.4 threshold,
 * else -1).  Both leave the limited gradient for the slope distributor. */
int zfv_set_flags(zfv_ctx* ctx, int criterion, double threshold);
/* amr::balance_flags (mesh/amr/balancing.h:15-35, balancing_fast.h:412-451). */
int zfv_balance_flags(zfv_ctx* ctx, int max_level);
/* EuMesh::refine (eu_mesh.cpp:290-359): balance, then amr::apply
 * (mesh/amr/apply.h:43-94) on the device, moving the state with the SmFluid
 * distributor (sm_fluid.cpp:372-470): distributor 0 = "const", 1 = "slope"
 * (uses the gradient of the last zfv_set_flags / zfv_smf_compute_grad). */
int zfv_refine(zfv_ctx* ctx, int max_level, int distributor);

/* Conserved totals {mass, mom_x, mom_y, energy} = sum V*q (synchronises);
 * V = volume_alt with set_axial, the volume the update divides by. */
int zfv_conserved_totals(zfv_ctx* ctx, double* out4);
/* zfv_conserved_totals without any device allocation, for callers that
 * cannot allocate (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  No counterpart in the reference.  The
 * same sums in the same order (bit-identical results), with the partial
 * sums in the library's static per-device scratch (4 MiB) instead of a
 * context buffer: meshes of at most 33,554,432 cells (ZFV_ERR_ARG beyond).
 * Every call on one device shares that scratch: calls that overlap on one
 * device race. */
int zfv_conserved_totals_noalloc(zfv_ctx* ctx, double* out4);

/* Limiter name -> id with the reference's parsing rules; -1 if unknown
 * (the reference throws, limiter.cxx:249-251). */
int zfv_limiter_from_name(const char* name);

/* ---- batches of independent problems (no reference counterpart; the
 * reference would loop over separate EuMesh objects) ------------------- */
/* One context holding n_problems Rectangle meshes nx[p] x ny[p] (box: 4
 * doubles {xmin, xmax, ymin, ymax} per problem, NULL = unit square; bounds: 4
 * Boundary ids per problem, NULL = ZOE), concatenated, each starting on a
 * 512-cell boundary (padding cells have no faces and zero volume).  Each
 * step kernel runs once per stage over all problems, every problem with its
 * own CFL time step.  State arrays span the padded length
 * (zfv_mesh_n_cells); adaptation is not available on a batch. */
int zfv_batch_rectangles(zfv_ctx* ctx, int n_problems, const int32_t* nx, const int32_t* ny,
                         const double* box, const int32_t* bounds);
/* Layout: *n_problems, begin[n_problems + 1] first cells (padded offsets),
 * count[n_problems] cells of each problem (any pointer may be NULL). */
int zfv_batch_layout(zfv_ctx* ctx, int32_t* n_problems, int32_t* begin, int32_t* count);
/* dt of the last step of every problem (n_problems doubles; synchronises). */
int zfv_batch_dt(zfv_ctx* ctx, double* dt);

/* ---- discrete adjoint (no reference counterpart) ---------------------------- */
/* Gradient of J = integral over the window {xmin, xmax, ymin, ymax} of p^2 at
 * the final time with respect to the initial primitive state, for the MUSCL
 * (acc 2) + HLLC step on a Cartesian mesh (limiter none, unit, minmod,
 * superbee or MC).  zfv_adjoint_forward runs n_steps steps from the current
 * state, keeping a checkpoint every `every` steps, and returns J.
 * zfv_adjoint_backward recomputes each segment from its checkpoint and runs
 * the reverse sweep; it returns dJ/d{rho, u, v, p, e} of the initial state
 * (n doubles each; e as an independent input, so a caller holding e(rho, p)
 * adds the chain rule) and leaves the solver at the final state.  Steps that
 * isolate a cell (bad state) are not differentiated: ZFV_ERR_UNSUPPORTED. */
int zfv_adjoint_forward(zfv_ctx* ctx, int n_steps, int every, const double* window, double* J);
int zfv_adjoint_backward(zfv_ctx* ctx, double* d_rho, double* d_u, double* d_v, double* d_p, double* d_e);

/* ---- multi-material solver (MmFluid, mm_fluid.h:45-163) ------------------- */
/* First-order MmFluid step on the context's mesh (one problem, uniform or
 * adapted): the mmf HLLC flux between pure states and, with composite
 * Riemann problems on (CrpMode::PLIC, the MmFluid default), the CRP flux of
 * faces next to mixed cells (MmFluid::calc_flux, mm_fluid.cpp:339-371, with
 * interface_recovery each step).  Materials are stiffened gases with a
 * temperature; mats = n_mat x {gamma, P0, e0, Cv, T0, rho_ref} (StiffenedGas
 * parameters, stiffened_gas.cpp:11-12, and the reference density that seeds
 * the closure, eos.h:152); the mixture is in pressure and temperature
 * equilibrium (MixturePT, mixture_pt.cpp:199-533).  cfl is clamped to [0, 1]
 * (MmFluid::set_CFL, mm_fluid.cpp:41-43; the reference default is 0.7).
 * Per-material arrays are [n_mat][n]. */
int zfv_mm_configure(zfv_ctx* ctx, int n_mat, const double* mats, double cfl);
/* Composite Riemann problems at mixed faces: 1 = CrpMode::PLIC with HLLC
 * (the default, MmFluid::set_method(Fluxes::CRP)), 0 = CrpMode::NONE with
 * HLLC everywhere (set_method(Fluxes::HLLC)) (mm_fluid.cpp:50-61). */
int zfv_mm_set_crp(zfv_ctx* ctx, int mode);
/* MmFluid::set_method (mm_fluid.cpp:50-61): ZFV_CRP = CrpMode::PLIC with
 * HLLC; ZFV_HLLC, ZFV_HLL or ZFV_RUSANOV = CrpMode::NONE with that mmf flux
 * (hllc.cpp:339-421, hll.cpp:125-160, rusanov.cpp:46-83).  The mmf Godunov
 * flux (godunov.cpp:39-77) and the other Fluxes are not built
 * (ZFV_ERR_UNSUPPORTED). */
int zfv_mm_set_method(zfv_ctx* ctx, int flux);
/* MmFluid::set_crp_mode (mm_fluid.cpp:63-70): ZFV_CRP_V3 / V5 / VS (the
 * heuristic face fractions, sections.cpp:52-84) or ZFV_CRP_PLIC, each with
 * HLLC; ZFV_CRP_NONE leaves the mode unchanged, as in the reference;
 * ZFV_CRP_MUSCL is ZFV_ERR_UNSUPPORTED (the reference's alpha_sigma returns
 * NaN for it, mm_fluid.cpp:244-249). */
int zfv_mm_set_crp_mode(zfv_ctx* ctx, int mode);
/* MmFluid::set_splitting (mm_fluid.cpp:76-78): ZFV_SPLIT_NONE (default),
 * ZFV_SPLIT_SIMPLE (an X sweep then a Y sweep, each with dt) or
 * ZFV_SPLIT_STRANG (X with dt / 2, Y with dt, X with dt / 2); a sweep takes
 * the faces of its direction (EuCell::faces(dir), |n| component >= 0.7,
 * amr_faces.cpp:186-197).  MmFluid::set_gravity has no counterpart: the
 * reference stores g and never reads it. */
int zfv_mm_set_splitting(zfv_ctx* ctx, int split);
/* MmFluid::set_max_dt (mm_fluid.cpp:94-96). */
int zfv_mm_set_max_dt(zfv_ctx* ctx, double max_dt);
/* Launch control of the multi-material step, no counterpart in the
 * reference (MmFluid::fluxes evaluates every face in one loop,
 * mm_fluid.cpp:373-445).  mode 0 (the default): every kernel on the
 * context's stream; 1: the interface reconstruction on a second stream of
 * the context, beside the gradient, predictor and composite-face marking;
 * 2: also the composite fluxes there, beside the update of the cells that
 * have no composite face, the other cells after them.  Results are the
 * same bit for bit in every mode; on MI355X modes 1 and 2 measured slower
 * than 0 (the cross-stream joins cost more than the overlap gains).
 * crp_groups: one-wavefront workgroups of the composite-flux kernel (0:
 * 2048). */
int zfv_mm_set_overlap(zfv_ctx* ctx, int mode, int crp_groups);

/* Cells from {rho, u, v, P, beta}: mmf::PState(rho, v, P, Fractions, mixture)
 * (models.cpp:176-185), beta normalised as Fractions does. */
int zfv_mm_set_state(zfv_ctx* ctx, const double* rho, const double* u, const double* v, const double* P,
                     const double* beta);
/* rho, u, v, P, e, T, beta [n_mat][n] and true densities rhos [n_mat][n]
 * (NaN for an absent material). */
int zfv_mm_state(zfv_ctx* ctx, double* rho, double* u, double* v, double* P, double* e, double* T, double* beta,
                 double* rhos);
/* n_steps MmFluid::update steps (compute_dt + fluxes + swap,
 * mm_fluid.cpp:110-196, 373-445, 682-684). */
int zfv_mm_step(zfv_ctx* ctx, int n_steps);
/* dt of the last step (MmFluid::dt). */
int zfv_mm_dt(zfv_ctx* ctx, double* dt);
/* Faces (cell sides) that took a composite Riemann flux in the last step
 * (0 without composite Riemann problems). */
int zfv_mm_crp_faces(zfv_ctx* ctx, int64_t* n);
/* MmFluid::interface_recovery (mm_fluid.cpp:484-519) of the current state:
 * per material the PLIC normal (nx, ny) and section position p,
 * [n_mat][n] each; zero for pure cells and absent materials. */
int zfv_mm_interface(zfv_ctx* ctx, double* nx, double* ny, double* p);
/* ---- Transfer (math/solver/transfer.h:21-153, transfer.cpp) ---------------
 * A volume fraction u carried by a velocity field over the context's mesh.
 * The reference takes the velocity from a virtual velocity(point); here the
 * caller gives its samples at the face centres (fvx, fvy: 8n, slot s of cell
 * c at 8c + s) and at the cell centres (cvx, cvy: n), the points the
 * reference evaluates it at.  Methods (Transfer::Method numbering): 0
 * CRP_V3, 1 CRP_V5, 2 CRP_SE, 3 CRP_N1, 7 MUSCLd, 9 MUSCLd_CRP, 11 MUSCL_MC,
 * 12 MUSCL_MC_CRP, 13 WENO, 14 WENO_CRP (update_WENO, transfer.cpp:541-666:
 * five-cell stencils on the (i, j) grid of an unadapted Rectangle of at
 * least 3 x 3 cells, indices wrapped periodically; ZFV_ERR_UNSUPPORTED at
 * update on other meshes); 4 CRP_N2, 5 VOF, 6 VOF_CRP, 8 MUSCLn, 10
 * MUSCLn_CRP (sub-cell polygon reconstruction) are not built
 * (ZFV_ERR_UNSUPPORTED).  The interface normals are
 * those of InterfaceRecovery::compute_normals after each update (no
 * smoothing); an adaptation invalidates the state (set it again). */
/* Transfer(), set_method, set_CFL (transfer.cpp:43-71). */
int zfv_transfer_configure(zfv_ctx* ctx, int method, double cfl);
/* The velocity samples (replace Transfer::velocity). */
int zfv_transfer_set_velocity(zfv_ctx* ctx, const double* fvx, const double* fvy, const double* cvx,
                              const double* cvy);
/* data.u1 of every cell, then the interface normals. */
int zfv_transfer_set_state(zfv_ctx* ctx, const double* u);
/* u1 and the interface normals n (each optional). */
int zfv_transfer_state(zfv_ctx* ctx, double* u, double* nx, double* ny);
/* du/dx, du/dy of the last MUSCL update (compute_slopes). */
int zfv_transfer_slopes(zfv_ctx* ctx, double* dx, double* dy);
/* set_dt / get_dt (transfer.cpp:78-84) and compute_dt (:86-102). */
int zfv_transfer_set_dt(zfv_ctx* ctx, double dt);
int zfv_transfer_get_dt(zfv_ctx* ctx, double* dt);
int zfv_transfer_compute_dt(zfv_ctx* ctx, double* dt);
/* n_steps calls of update(mesh, dir), dir 0 ANY, 1 X, 2 Y (transfer.cpp:474-535). */
int zfv_transfer_update(zfv_ctx* ctx, int dir, int n_steps);
/* set_flags (transfer.cpp:668-690) into the mesh's adaptation flags. */
int zfv_transfer_set_flags(zfv_ctx* ctx);

/* MmFluid::domain (mm_fluid.cpp:686-712) after interface_recovery of the
 * current state: the region material idx occupies, one polygon per cell
 * holding it -- the cell, or the cell clipped by its PLIC line (Polygon::clip,
 * polygon.cpp:266-292).  Outputs, each optional: the polygon count, the
 * vertex count, offsets (n_polys + 1: polygon k has vertices offsets[k] ..
 * offsets[k + 1] - 1, counter-clockwise) and the vertices xs, ys.  Call with
 * null arrays for the counts, then with arrays of those sizes; no
 * counterpart of the returned EuMesh's other fields.  Host code. */
int zfv_mm_domain(zfv_ctx* ctx, int idx, int64_t* n_polys, int64_t* n_verts, int32_t* offsets, double* xs,
                  double* ys);
/* MmFluid::set_limiter (mm_fluid.cpp:45-47): limiter number as
 * zfv_limiter_from_name gives it (default MC). */
int zfv_mm_set_limiter(zfv_ctx* ctx, int limiter);
/* MmFluid::set_accuracy (mm_fluid.cpp:46-48): 1 (default) the first-order
 * update; 2 compute_grad, fluxes_stage1 and fluxes_stage2 (mm_fluid.cpp:
 * 153-168, 521-680) with PState::interpolation_update (models.cpp:381-391).
 * Values outside [1, 2] are clamped. */
int zfv_mm_set_accuracy(zfv_ctx* ctx, int acc);
/* MmFluid::set_flags (mm_fluid.cpp:836-885): compute_grad (zero gradients
 * in and next to mixed cells with composite Riemann problems on, else the
 * limited LSM gradient; kept for zfv_mm_refine), then per cell +1 for a
 * mixed cell or a pure cell next to another material, else -1, into the
 * mesh's refinement flags (zfv_flags_download reads them). */
int zfv_mm_set_flags(zfv_ctx* ctx);
/* EuMesh::refine (eu_mesh.cpp:290-359) with MmFluid::distributor
 * (mm_fluid.cpp:765-830): balance_flags, then amr::apply with the
 * multi-material split (from the gradients of the last compute_grad, which
 * must have run on the current mesh) and the conservative merge.  The
 * single-material state, if any, does not follow the mesh. */
int zfv_mm_refine(zfv_ctx* ctx, int max_level);
/* The gradients of the last compute_grad (zfv_mm_set_flags, or a
 * second-order step: those of its old layer, as the reference's d_dx /
 * d_dy): rows rho, u, v, e, beta[n_mat], P, T, rhos[n_mat], [6 + 2 n_mat][n]
 * each for x and y (NaN where a true density's stencil meets a cell without
 * that material).  zfv_mm_set_flags computes only the rows its distributor
 * reads (rho, u, v, e, beta); the others come back as NaN after it. */
int zfv_mm_grad(zfv_ctx* ctx, double* gx, double* gy);

/* ---- output and checkpoints (io/, EuMesh::backup) ------------------------- */
/* Cell fields of a VTU file (VtuFile variables): density, velocity (3
 * components, v_z = 0), pressure, specific internal energy, isolation
 * counter, refinement level, refinement flag, cell index. */
enum {
    ZFV_VTU_DENSITY = 1, ZFV_VTU_VELOCITY = 2, ZFV_VTU_PRESSURE = 4, ZFV_VTU_ENERGY = 8, ZFV_VTU_WAIT = 16,
    ZFV_VTU_LEVEL = 32, ZFV_VTU_FLAG = 64, ZFV_VTU_INDEX = 128,
    ZFV_VTU_STATE_FIELDS = 31, ZFV_VTU_ALL = 255
};
/* VtuFile::save (io/vtu_file.cpp:439-584): the mesh and the chosen cell
 * fields as a VTK UnstructuredGrid with raw appended data.  polygons = 0:
 * every cell a quad of its four corners (AmrCells written with hanging
 * nodes); 1: polygons through the midpoints of split faces
 * (PvdFile::polyhedral). */
int zfv_vtu_write(zfv_ctx* ctx, const char* filename, unsigned fields, int polygons);
/* PvdFile (io/pvd_file.h:12-75): zfv_pvd_open makes directory/filename/
 * (directory NULL or empty: "output"); every zfv_pvd_save writes the next
 * VTU file there and rewrites directory/filename.pvd listing all saved
 * files with their times. */
typedef struct zfv_pvd zfv_pvd;
int zfv_pvd_open(const char* filename, const char* directory, zfv_pvd** out);
int zfv_pvd_save(zfv_pvd* pvd, zfv_ctx* ctx, double time, unsigned fields, int polygons);
int zfv_pvd_close(zfv_pvd* pvd);
/* EuMesh::backup (mesh/euler/eu_mesh.cpp:1017-1061, AmrCells::backup
 * amr_cells.cpp:1074-1172) into directory `root`: backup.json and one raw
 * file per array in the reference's layout, plus the solver's parameters
 * and state (cells/data/init as PState {rho, v, p, e}, cells/data/wait).
 * zfv_restore reads such a directory into a context: mesh, solver
 * parameters, state, isolation counters and flags; the continued run is
 * the same as the original's (the reference has no reader,
 * tests/mesh/backup_read.cpp:23-36). */
int zfv_backup(zfv_ctx* ctx, const char* root);
int zfv_restore(zfv_ctx* ctx, const char* root);

/* ---- launch configuration (no reference counterpart) --------------------- */
/* Step kernels of a stage (0 = first-order update, 1 = gradient + predictor,
 * 2 = corrector): layout 0 = one thread per cell, 1 = four lanes per cell
 * (one face each), 2 = one thread per cell with the main slots' loads
 * issued together (more registers per thread; for meshes that do not fill
 * the GPU), block = workgroup size 128, 256 or 512 (0 keeps the current
 * one).  Layout 3 (stage 1 only) runs the whole MUSCL + HLLC step in one
 * tiled kernel with the half states and gradients in LDS, on meshes built
 * by zfv_mesh_rectangle / zfv_batch_rectangles (before any adaptation;
 * ZFV_ERR_UNSUPPORTED otherwise); its block is the tile shape, 512, 768 or
 * 1024 threads (6, 10 or 14 rows of 62 cells); it is the default there, and
 * schemes it does not cover run the stage kernels.  All choices give
 * bit-identical results. */
int zfv_set_launch(zfv_ctx* ctx, int stage, int layout, int block);
int zfv_get_launch(zfv_ctx* ctx, int stage, int* layout, int* block);
/* Schedule of the tiled step (layout 3): 0 (default) one workgroup per
 * tile; 1 persistent workgroups, as many as the GPU holds at once, each
 * walking the tiles of its XCD's range.  Results are bit-identical.  No
 * counterpart in the reference (its SmFluid::update loops over cells on the
 * CPU, sm_fluid.cpp:81-102). */
int zfv_set_tile_schedule(zfv_ctx* ctx, int persistent);
int zfv_get_tile_schedule(zfv_ctx* ctx, int* persistent);
/* Autotuner: times every (layout, workgroup size) of the configured scheme's
 * step kernels on the current state (`reps` launches each; the state is not
 * advanced), and the tiled step's shapes and schedules where it applies, and
 * keeps the fastest per stage (the tiled step when it beats the best
 * gradient + corrector pair).  ms_out (may be NULL): 36 doubles, ms per
 * launch indexed [stage][layout 0-3][block 128/256/512, or for layout 3 the
 * tile shape 512/768/1024], where [1][3][shape] is the tiled step with one
 * workgroup per tile and [2][3][shape] with persistent workgroups (0 where
 * not timed). */
int zfv_autotune(zfv_ctx* ctx, int reps, double* ms_out);
/* Capture steps into HIP graphs (6 steps per graph, one graph per starting
 * phase, replayed by zfv_smf_step; re-captured after any change of mesh,
 * scheme or launch configuration).  Off by default; profiling timers are
 * not recorded for graph-launched steps. */
int zfv_graph_enable(zfv_ctx* ctx, int on);
/* Cell order of the step kernels.  mode 1 (default): on a mesh with refined
 * cells the steps run on a copy of the mesh sorted along a space-filling
 * curve (base cell, then Morton index inside it), the order carried across
 * each zfv_refine (children where their parent was, a parent where its
 * children were) and the copy rebuilt at the next step; the state is
 * gathered into that order before steps and back into the mesh's order by
 * every entry point that reads or writes per-cell data, so the order is
 * never observable and the results are bit-identical to mode 0 (the mesh's
 * own order throughout).  mode 2: as mode 1 with the order computed afresh
 * (sorted) after each adaptation.  active: 1 while the state is held in
 * curve order.  No counterpart in the reference (its cells stay in storage
 * order, amr_cells.h:74-104). */
int zfv_set_cell_order(zfv_ctx* ctx, int mode);
int zfv_get_cell_order(zfv_ctx* ctx, int* mode, int* active);
/* The steps' cell order of the current mesh: perm[i] (n entries) is the
 * mesh index of the i-th cell along the curve.  ZFV_ERR_STATE when the steps
 * run in the mesh's own order (mode 0, no refined cells, or a batch). */
int zfv_get_curve_order(zfv_ctx* ctx, int32_t* perm);

/* Kernel timing: average device time (ms) per launch of each step kernel
 * over the last zfv_smf_step call when enabled (HIP events on the context
 * stream).  names/ms arrays of length *count (in: capacity, out: used). */
int zfv_profile_enable(zfv_ctx* ctx, int on);
int zfv_profile_read(zfv_ctx* ctx, const char** names, double* ms, int32_t* launches, int32_t* count);

/* ---- three-dimensional meshes (EuMesh over AMR-3D cells, one level) ---- */
/* A context holding a three-dimensional Cuboid mesh and its SmFluid state.
 * No reference counterpart as a handle (EuMesh(Cuboid) + SmFluid,
 * eu_mesh.cpp:70-104, sm_fluid.cpp:16-35). */
typedef struct zfv3_ctx zfv3_ctx;
int zfv3_create(int device, zfv3_ctx** out);
int zfv3_destroy(zfv3_ctx* ctx);
/* Build the Cuboid mesh on the host and upload it (replaces
 * generator::Cuboid + Cuboid::initialize, geom/generator/cuboid.cpp:300-442):
 * cell c = nz (ny i + j) + k; bounds = {left, right, bottom, top, back,
 * front} Boundary ids (WALL, ZOE, PERIODIC, FUNCTION; NULL = ZOE; a periodic
 * side makes its direction periodic on both ends). */
int zfv3_mesh_cuboid(zfv3_ctx* ctx, int nx, int ny, int nz, double xmin, double xmax, double ymin, double ymax,
                     double zmin, double zmax, const int32_t* bounds);
int zfv3_mesh_n_cells(zfv3_ctx* ctx, int32_t* n);
/* Geometry tables: a Cuboid's centres, face centres, weights, areas and
 * volume depend on one index each, so the step kernels can read 1D tables
 * (nx + ny + nz entries per quantity) instead of the per-slot arrays; the
 * tables are used only when every cell's values equal them bit for bit.
 * on = 0 keeps the per-slot arrays for meshes built afterwards (default 1);
 * active: whether the current mesh uses the tables.  No reference
 * counterpart. */
int zfv3_set_geometry_tables(zfv3_ctx* ctx, int on);
int zfv3_get_geometry_tables(zfv3_ctx* ctx, int* active);
/* The mesh in the AMR-3D layout (mesh/euler/amr_cells.h:74-104,
 * amr_faces.h:83-101): per cell cx, cy, cz, volume, incircle diameter h
 * (amr_cells.cpp:107-121); 24 face slots per cell (slot s of cell c at
 * 24 c + s, sides L, R, B, T, Z, F then their sub-faces, which a Cuboid
 * leaves undefined: boundary -1, adj -1, zeros), with the Boundary id, the
 * neighbour (the cell itself on a non-periodic boundary), the outward
 * normal, face centre, area and least-squares weight w = area / (cbrt(V)
 * |drn|^3) (gradient.h:231-256).  NULL arrays are skipped. */
int zfv3_mesh_download(zfv3_ctx* ctx, double* cx, double* cy, double* cz, double* vol, double* h,
                       int32_t* boundary, int32_t* adj, double* nx, double* ny, double* nz, double* fcx,
                       double* fcy, double* fcz, double* area, double* w);
/* SmFluid setters as zfv_smf_configure; flux GODUNOV..CIR2
 * (ZFV_ERR_UNSUPPORTED for CRP), axial 0. */
int zfv3_smf_configure(zfv3_ctx* ctx, const zfv_smf_params* params);
int zfv3_smf_set_max_dt(zfv3_ctx* ctx, double max_dt);
/* Primitive state PState{rho, {u, v, w}, p, e} of the "init" layer, n
 * doubles per field; e NULL = energy_rP(rho, p) (stiffened_gas.cpp:121-122). */
int zfv3_state_upload(zfv3_ctx* ctx, const double* rho, const double* u, const double* v, const double* w,
                      const double* p, const double* e);
int zfv3_state_download(zfv3_ctx* ctx, double* rho, double* u, double* v, double* w, double* p, double* e);
/* Isolation counters (Parts::wait, sm_fluid.h:34-35). */
int zfv3_wait_upload(zfv3_ctx* ctx, const int32_t* wait);
int zfv3_wait_download(zfv3_ctx* ctx, int32_t* wait);
/* Limited gradients of {rho, u, v, w, p} of the last second-order step:
 * [3][5][n] (d/dx, d/dy, d/dz; compute_grad, sm_fluid.cpp:116-125). */
int zfv3_grad_download(zfv3_ctx* ctx, double* g);
/* `n_steps` calls of SmFluid::update (sm_fluid.cpp:81-102), enqueued; NaN
 * fluxes are sticky and reported by the next synchronising call. */
int zfv3_smf_step(zfv3_ctx* ctx, int n_steps);
/* SmFluid::dt() of the last step; compute_dt (sm_fluid.cpp:104-114) alone. */
int zfv3_smf_dt(zfv3_ctx* ctx, double* dt);
int zfv3_smf_compute_dt(zfv3_ctx* ctx, double* dt);
/* {mass, mom_x, mom_y, mom_z, energy} = sum V q (synchronises). */
int zfv3_conserved_totals(zfv3_ctx* ctx, double* out5);
int zfv3_synchronize(zfv3_ctx* ctx);
/* Kernel timing as zfv_profile_enable / zfv_profile_read: total device time
 * (ms) and launches of each step kernel (dt, flux1, grad_pred, corr) since
 * timing was enabled or last read (HIP events on the context stream). */
int zfv3_profile_enable(zfv3_ctx* ctx, int on);
int zfv3_profile_read(zfv3_ctx* ctx, const char** names, double* ms, int32_t* launches, int32_t* count);

#ifdef __cplusplus
}
#endif
#endif /* ZFV_H */
