/*
 * kraken_mi355x.h — C ABI of the MI355X lattice-Boltzmann stream–collide library.
 *
 * This is the drop-in boundary for Kraken.jl's fused stream–collide hot path.  A host
 * (Julia `ccall`, Python `ctypes`, C) hands over host arrays in the reference's own
 * layout — column-major `f[i, j, (k,) q]`, x fastest, one contiguous plane per
 * velocity `q`, q 1-based in Julia and 0-based here — and the library keeps the
 * populations resident in HBM between calls.
 *
 * Reference interfaces replaced (paths relative to the Kraken.jl source tree):
 *   - backend value / dispatch:   src/simulation_runner.jl:110-160, 235-443
 *                                  (`_backend_from_krk_symbol`, `run_simulation`)
 *   - state allocation / init:     src/drivers/basic.jl:51-101 (`initialize_2d/3d`)
 *   - fused single-pass step:      src/kernels/fused_bgk_2d.jl:67, fused_trt_2d.jl:145,
 *                                  fused_thermal_2d.jl:251, dsl/bricks_3d.jl:64-229
 *   - unfused step chains:         src/drivers/basic.jl:121-136 (`run_cavity_2d`),
 *                                  src/drivers/thermal.jl:100-133 (`run_rayleigh_benard_2d`),
 *                                  src/kernels/collide_mrt_2d.jl:118 (`collide_mrt_2d!`)
 *   - moments:                     src/kernels/macroscopic.jl:5-80,
 *                                  src/kernels/thermal_2d.jl:113-138
 *   - residual / Nusselt:          src/drivers/gnf_periodic_steady.jl:186-196,
 *                                  src/drivers/thermal.jl:225-234
 *
 * Conventions
 *   - Every call returns a kmi_status (0 = KMI_OK).  On failure, kmi_last_error(ctx)
 *     (or kmi_last_error(NULL) for calls that have no context yet) names the cause.
 *     Nothing ever falls back to a CPU path: a missing device is an error.
 *   - The library owns device memory; the caller owns every host buffer it passes.
 *   - One context = one device + one HIP stream.  Calls on one context are serialised
 *     by the caller; distinct contexts are independent.
 */
#ifndef KRAKEN_MI355X_H
#define KRAKEN_MI355X_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define KMI_ABI_VERSION 2

typedef enum {
  KMI_OK = 0,
  KMI_ERR_ARG = 1,         /* invalid argument (shape, enum, NULL pointer) */
  KMI_ERR_DEVICE = 2,      /* HIP runtime error or no usable device */
  KMI_ERR_STATE = 3,       /* call not valid in the object's current state */
  KMI_ERR_NOMEM = 4,       /* device allocation failed */
  KMI_ERR_UNSUPPORTED = 5  /* combination the reference does not define */
} kmi_status;

typedef enum { KMI_D2Q9 = 0, KMI_D3Q19 = 1 } kmi_stencil;

/* Population storage in HBM.  Arithmetic is FP64 for F64 and FP32 otherwise; F64 and F32
 * follow the reference's unfused evaluation order exactly, F32S and F16S collide with
 * fused multiply-adds.
 * F32S / F16S store the deviation f_q - w_q (thermal: g_q - w_q T_ref); F16S scales
 * it by 2^15 before rounding to IEEE binary16.  The thermal populations of an F16S
 * lattice are stored as F32S: binary16 cannot resolve per-step temperature increments
 * against deviations of order w_q (T - T_ref). */
typedef enum {
  KMI_STORE_F64 = 0,
  KMI_STORE_F32 = 1,
  KMI_STORE_F32S = 2,
  KMI_STORE_F16S = 3
} kmi_storage;

typedef enum { KMI_BGK = 0, KMI_TRT = 1, KMI_MRT = 2 } kmi_collision;

/* Face treatment of the flow populations.
 *   PERIODIC        — wrap; both faces of the axis must be periodic.
 *   WALL            — halfway bounce-back folded into the pull (collide_stream_2d.jl:75-101).
 *   MOVING_WALL     — halfway bounce-back with the Ladd momentum term
 *                     f_q = f*_opp(q) + 6 w_q rho_w (c_q . u_w)   (stream_periodic_3d.jl:106-166).
 *   ZOUHE_VELOCITY  — bounce-back pull, then the Zou–He velocity closure of the
 *                     reference on that face (D2Q9: boundary_2d.jl:13-44, 135-185, north,
 *                     south, west, x velocity; D3Q19: boundary_3d.jl:152-232, any face,
 *                     any velocity).
 *   ZOUHE_PRESSURE  — bounce-back pull, then the Zou–He pressure closure
 *                     (D2Q9: boundary_2d.jl:206-299, west and east; D3Q19:
 *                     boundary_3d.jl:234-290, any face).
 *   ZOUHE_VELOCITY_PROFILE — D2Q9 north, south or west: per-node wall velocity (u_x, u_y)
 *                     set by kmi_set_face_profile; the two corner nodes of the face keep
 *                     the bounce-back pull (boundary_spatial_2d.jl:6-90).
 *   ZOUHE_PRESSURE_PROFILE — D2Q9 east: per-node density, corner nodes skipped
 *                     (boundary_spatial_2d.jl:104-131).
 *   ZOUHE_VELOCITY_REBUILD — west, D2Q9 or D3Q19: the face nodes off the edges
 *                     (j = 2..Ny-1, k = 2..Nz-1) are rebuilt from the step's pulled
 *                     populations — Zou–He closure for the normal velocity of the per-node
 *                     profile set by kmi_set_face_profile (length ny, or ny*nz with index
 *                     j + ny k), then a local TRT collision at the lattice's rates (D2Q9:
 *                     the regularised form of _trt_collide_local) — replacing the node's
 *                     own update; the edge nodes keep the bounce-back pull
 *                     (apply_bc_rebuild_2d!/_3d!, src/bc/rebuild_2d.jl:1-23,
 *                     rebuild_3d.jl:1-45, specs.jl:7-108).  TRT lattices only.
 *   ZOUHE_PRESSURE_REBUILD — east, D2Q9 or D3Q19: the same with the Zou–He pressure
 *                     closure for `rho` (rebuild_2d.jl:25-50, rebuild_3d.jl:47-89). */
typedef enum {
  KMI_FACE_PERIODIC = 0,
  KMI_FACE_WALL = 1,
  KMI_FACE_MOVING_WALL = 2,
  KMI_FACE_ZOUHE_VELOCITY = 3,
  KMI_FACE_ZOUHE_PRESSURE = 4,
  KMI_FACE_ZOUHE_VELOCITY_PROFILE = 5,
  KMI_FACE_ZOUHE_PRESSURE_PROFILE = 6,
  KMI_FACE_ZOUHE_VELOCITY_REBUILD = 7,
  KMI_FACE_ZOUHE_PRESSURE_REBUILD = 8
} kmi_face_kind;

/* Body force on the flow populations (BGK, D2Q9 or D3Q19).
 *   NONE — none.
 *   GUO  — constant force desc.force with the Guo source term, collide_guo_2d!
 *          (collide_guo_2d.jl:5-72) / collide_guo_3d! (collide_guo_3d.jl:5-190); moments
 *          and reductions report u = (sum c f + F/2) / rho, compute_macroscopic_forced_2d!
 *          / _3d! (macroscopic.jl:84-160), as the generic runner does
 *          (simulation_runner.jl:402-406).
 *   GUO_PLAIN_MOMENTS — the same collision, moments without the F/2 shift, as
 *          run_poiseuille_2d reports them (basic.jl:197-199).
 *   GUO_FIELD — per-node force set by kmi_set_force_field (zero until set):
 *          collide_guo_field_2d! / collide_guo_field_3d! (collide_guo_2d.jl:93-140,
 *          collide_guo_3d.jl:199-325); moments with the per-node F/2 shift,
 *          compute_macroscopic_forced_field_3d! (macroscopic.jl:164-200). */
typedef enum {
  KMI_FORCE_NONE = 0,
  KMI_FORCE_GUO = 1,
  KMI_FORCE_GUO_PLAIN_MOMENTS = 2,
  KMI_FORCE_GUO_FIELD = 3
} kmi_forcing;

/* Cells flagged solid by kmi_set_mask.
 *   BOUNCE_BACK — the cell swaps its pulled populations (bounce_back_2d!, collide_2d.jl).
 *   INERT       — the cell holds the rest equilibrium f_q = w_q (SolidInert brick,
 *                 dsl/bricks.jl:133-154): the wall is represented by the cut links. */
typedef enum { KMI_SOLID_BOUNCE_BACK = 0, KMI_SOLID_INERT = 1 } kmi_solid_mode;

/* Faces: x-, x+, y-, y+, z-, z+ (the reference's west/east/south/north/bottom/top). */
typedef enum {
  KMI_WEST = 0,
  KMI_EAST = 1,
  KMI_SOUTH = 2,
  KMI_NORTH = 3,
  KMI_BOTTOM = 4,
  KMI_TOP = 5
} kmi_face_id;

/* Thermal (g) populations on a non-periodic face: adiabatic = bounce-back,
 * Dirichlet = halfway anti-bounce-back g_q = -g*_opp(q) + 2 w_q T_w with the
 * reference's corner rule in 2D (thermal_2d.jl:140-172, 371-417) and edge rule in 3D
 * (thermal_3d.jl:164-230, 494-535): a diagonal slot whose link crosses two Dirichlet
 * faces is flipped once, with the mean of the two wall temperatures. */
typedef enum { KMI_THERMAL_ADIABATIC = 0, KMI_THERMAL_DIRICHLET = 1 } kmi_thermal_face;

/* Velocity seen by the thermal collision.
 *   SPLIT — the reference's unfused chains: u from the post-stream populations without
 *           the Guo half-force shift (D2Q9: thermal.jl:123-128; D3Q19, the only 3D
 *           chain: run_natural_convection_3d, thermal.jl:500-515).
 *   FUSED — the reference's fused natural-convection kernel: u with the F/2 shift
 *           (fused_thermal_2d.jl:61-72, 197-200). */
typedef enum { KMI_COUPLING_SPLIT = 0, KMI_COUPLING_FUSED = 1 } kmi_thermal_coupling;

typedef struct {
  int32_t kind;     /* kmi_face_kind */
  int32_t thermal;  /* kmi_thermal_face (used when desc.thermal != 0) */
  double u[3];      /* wall velocity (MOVING_WALL) or imposed velocity (ZOUHE_VELOCITY) */
  double rho;       /* wall density (MOVING_WALL) or imposed density (ZOUHE_PRESSURE) */
  double T;         /* wall temperature (thermal DIRICHLET) */
} kmi_face;

typedef struct {
  int32_t stencil;    /* kmi_stencil */
  int32_t storage;    /* kmi_storage */
  int32_t collision;  /* kmi_collision; MRT is D2Q9 only */
  int32_t thermal;    /* 0 / 1: double-distribution temperature field g (Q populations) */
  int64_t nx, ny, nz; /* nz = 1 for D2Q9 */
  double omega;             /* BGK rate  1/(3 nu + 1/2)                     (basic.jl:30) */
  double s_plus, s_minus;   /* TRT rates, see kmi_trt_rates              (fused_trt_2d.jl:129-133) */
  double s_e, s_eps, s_q, s_nu; /* MRT rates                            (collide_mrt_2d.jl:118) */
  double omega_T;           /* thermal BGK rate 1/(3 alpha + 1/2)           (thermal.jl:65) */
  double beta_g;            /* Boussinesq  F_y = beta_g (T - T_ref), y up (thermal_2d.jl:435,
                               thermal_3d.jl:583) */
  double T_ref;             /* buoyancy reference temperature (also the g storage shift) */
  int32_t coupling;         /* kmi_thermal_coupling */
  int32_t n_bc_order;       /* 0: default order W,E,S,N,B,T; else the first n entries of */
  int32_t bc_order[6];      /*   bc_order give the order Zou–He closures are applied in   */
  kmi_face faces[6];        /* indexed by kmi_face_id; z faces ignored for D2Q9           */
  int32_t forcing;          /* kmi_forcing                                                */
  double force[3];          /* body force (F_x, F_y, F_z) for KMI_FORCE_GUO(_PLAIN_MOMENTS) */
  int32_t solid_mode;       /* kmi_solid_mode                                             */
} kmi_desc;

typedef struct kmi_ctx kmi_ctx;
typedef struct kmi_lattice kmi_lattice;

/* Population fields of a lattice. */
typedef enum { KMI_FIELD_F = 0, KMI_FIELD_G = 1 } kmi_field;
/* Element type of a host buffer.  STORAGE: the storage elements themselves, no conversion
 * (F64: double; F32 and F32S: float, F32S holding f - w_q; F16S: the IEEE binary16 bits of
 * 2^15 (f - w_q) as uint16; thermal g of an F16S lattice: F32S floats) — the exact state,
 * for checkpoints that restart bit for bit. */
typedef enum { KMI_HOST_F64 = 0, KMI_HOST_F32 = 1, KMI_HOST_STORAGE = 2 } kmi_host_dtype;

/* Reductions, all accumulated in FP64 with compensated (Neumaier) sums.
 *   MASS            out[0] = sum_cells rho
 *   MOMENTUM        out[0..2] = sum_cells rho u
 *   KINETIC_ENERGY  out[0] = sum_cells 1/2 rho |u|^2
 *   THERMAL         out[0] = sum_cells T
 *   NUSSELT         out[0] = hot-wall Nusselt number; requires two opposite Dirichlet
 *                   faces, the hot one named by `arg` (a kmi_face_id) — estimator of
 *                   thermal.jl:225-234 and 525-533 (second-order one-sided gradient
 *                   -H (-8 T_w + 9 T_1 - T_2) / (3 dT), H the cell count along the
 *                   normal, averaged over the face without its edge rows; a periodic
 *                   wall-parallel axis keeps all its rows)
 *   MAX_SPEED       out[0] = max_cells |u|
 *   DRAG_MEA        out[0..2] = force on the cut-link walls, Mei–Luo–Shyy momentum
 *                   exchange with the Bouzidi arriving population, per link
 *                   c_q (f_q + f~_qbar)   (cylinder_libb.jl:61-104, drag_gpu.jl:346-403)
 *   DRAG_HALFWAY    out[0..2] = sum over cut links of 2 c_q f_q   (halfway-BB momentum
 *                   exchange, cylinder_libb.jl:27-42, li_bb_3d_v2.jl:167-206) */
typedef enum {
  KMI_RED_MASS = 0,
  KMI_RED_MOMENTUM = 1,
  KMI_RED_KINETIC_ENERGY = 2,
  KMI_RED_THERMAL = 3,
  KMI_RED_NUSSELT = 4,
  KMI_RED_MAX_SPEED = 5,
  KMI_RED_DRAG_MEA = 6,
  KMI_RED_DRAG_HALFWAY = 7
} kmi_reduction;

/* ---- library / context ---------------------------------------------------------- */
int kmi_abi_version(void);
/* Number of visible HIP devices (0 is a valid answer; no error). */
int kmi_device_count(int* n);
/* Human-readable description of the last error on `ctx` (NULL: the calling thread's
 * last context-free error).  Never NULL; valid until the next call on the same ctx. */
const char* kmi_last_error(const kmi_ctx* ctx);
int kmi_ctx_create(int device, kmi_ctx** out);
int kmi_ctx_destroy(kmi_ctx* ctx);
int kmi_ctx_synchronize(kmi_ctx* ctx);

/* TRT rates from viscosity and magic parameter: s_plus = 1/(3 nu + 1/2),
 * s_minus = 1/(Lambda/(3 nu) + 1/2)   (src/kernels/fused_trt_2d.jl:129-133). */
void kmi_trt_rates(double nu, double Lambda, double* s_plus, double* s_minus);
/* Fill `desc` with the defaults of a D2Q9/D3Q19 lattice: all faces WALL, BGK from nu. */
void kmi_desc_init(kmi_desc* desc, int32_t stencil, int64_t nx, int64_t ny, int64_t nz,
                   double nu);
/* Check a descriptor without a device: KMI_OK, or the status kmi_lattice_create would
 * return, with the reason in kmi_last_error(NULL). */
int kmi_desc_validate(const kmi_desc* desc);
/* sizeof(kmi_desc) as compiled into the library (host bindings check their layout). */
size_t kmi_desc_sizeof(void);

/* ---- single lattice (configs: cavity, Taylor–Green, Rayleigh–Bénard) ------------- */
/* Allocates f (and g) on the device and sets them to the rest equilibrium f_q = w_q
 * (g_q = w_q T_ref), as initialize_2d/3d (basic.jl:51-101). */
int kmi_lattice_create(kmi_ctx* ctx, const kmi_desc* desc, kmi_lattice** out);
int kmi_lattice_destroy(kmi_lattice* lat);
/* Solid mask, one byte per cell in [i, j, k] order (nonzero = solid); NULL clears it. */
int kmi_set_mask(kmi_lattice* lat, const uint8_t* is_solid);
/* Per-node values of a *_PROFILE face: velocity profile (a = u_x, b = u_y) or pressure
 * profile (a = rho, b ignored), each of face length (nx for south/north, ny for
 * west/east), index 0 at the low end of the face — the arrays _fill_bc_arrays! builds
 * (src/bc/handlers.jl:152-187).  Until set, a velocity profile is zero and a pressure
 * profile equals the face's rho.  May be called between steps (time-dependent profiles,
 * src/bc/handlers.jl:203-208). */
int kmi_set_face_profile(kmi_lattice* lat, int32_t face, const double* a, const double* b);
/* Interpolated bounce-back on cut links (Bouzidi, applied to the pulled populations
 * before the collision: fused_trt_libb_v2_step!, li_bb_2d_v2.jl:35-110, dsl/bricks.jl:516-581;
 * D3Q19 li_bb_3d_v2.jl:13-43, dsl/bricks_3d.jl:366-473).  Link n leaves fluid cell
 * cell[n] (index i + nx (j + ny k)) along direction q[n] (0-based) and meets the wall at
 * the fraction qw[n] in (0, 1] of its length; uw[3n..3n+2] is the wall velocity there.
 * The population opp(q) pulled from across the wall is replaced by
 *   qw <= 1/2:  2qw f_q(x) + (1 - 2qw) f~_q + d
 *   qw >  1/2:  s f_q(x) + (1 - s) f_opp(x) + d s,   s = 1/(2qw)
 * with f the cell's previous post-collision populations, f~_q the pulled population q and
 * d = -6 w_q (c_q . uw), links taken in the reference's pair order.  Each call replaces the previous list; n = 0 clears it.  The
 * cells of the list must not be masked solid.  D2Q9 or D3Q19, BGK or TRT. */
int kmi_set_cut_links(kmi_lattice* lat, int64_t n, const int32_t* cell, const int32_t* q,
                      const double* qw, const double* uw);
/* Per-node body force for KMI_FORCE_GUO_FIELD: the Fx_field, Fy_field (, Fz_field)
 * arrays of collide_guo_field_2d! / collide_guo_field_3d! (collide_guo_2d.jl:93-160,
 * collide_guo_3d.jl:199-332), nx*ny*nz values each in the reference layout.  The values
 * are rounded once to the collision arithmetic (FP64 for F64 storage, FP32 otherwise),
 * as the reference's element-typed field arrays are.  fz is ignored for D2Q9 (may be
 * NULL).  The field persists until the next call. */
int kmi_set_force_field(kmi_lattice* lat, const double* fx, const double* fy, const double* fz);
/* Replace the values (u, rho, T) of one face between steps; the face kind and thermal
 * kind must stay the same (time-dependent scalar conditions, handlers.jl:210-218). */
int kmi_set_face(kmi_lattice* lat, int32_t face, const kmi_face* values);
/* Whole population field in the reference layout [i, j, (k,) q], nx*ny*nz*Q values. */
int kmi_upload(kmi_lattice* lat, int32_t field, const void* host, int32_t host_dtype);
int kmi_download(kmi_lattice* lat, int32_t field, void* host, int32_t host_dtype);
/* f_q = feq_q(rho, u) per cell (equilibrium of basic.jl / d2q9.jl:60-67); any of
 * ux/uy/uz may be NULL (zero), rho may be NULL (one).  Host arrays of nx*ny*nz. */
int kmi_init_equilibrium(kmi_lattice* lat, const double* rho, const double* ux,
                         const double* uy, const double* uz);
/* g_q = w_q T per cell (thermal.jl:74-97). */
int kmi_init_temperature(kmi_lattice* lat, const double* T);
/* Advance `nsteps` single-pass steps (stream + boundaries + collide in one launch each). */
int kmi_step(kmi_lattice* lat, int64_t nsteps);
/* As kmi_step, and also the relative one-step velocity change of the last step,
 * r_u = ||u^k - u^(k-1)||_2 / ||u^k||_2 over fluid cells (gnf_periodic_steady.jl:186-196). */
int kmi_step_residual(kmi_lattice* lat, int64_t nsteps, double* r_u);
/* As kmi_step_residual over the last `window` = m steps (1 <= m <= nsteps):
 * r = ||u^k - u^(k-m)||_2 / (m ||u^k||_2), the stop rule of kmi_batch_run; m = 1 is r_u.
 * No reference counterpart: for single-lattice runs to a residual in FP32 storage, whose
 * one-step r_u has a rounding floor near 4e-7. */
int kmi_step_residual_window(kmi_lattice* lat, int64_t nsteps, int64_t window, double* r_u);
/* Device-timed kmi_step: elapsed milliseconds between HIP events on the lattice stream. */
int kmi_step_timed(kmi_lattice* lat, int64_t nsteps, double* ms);
/* Moments of the stored (post-collision) populations, as compute_macroscopic_2d/3d and
 * compute_temperature_2d (macroscopic.jl:5-80, thermal_2d.jl:113-138).  Any pointer may
 * be NULL.  Host arrays of nx*ny*nz doubles.  The reference's fused kernels
 * (fused_bgk_step!, fused_trt_step!, fused_bgk_2d.jl:26-56, fused_trt_2d.jl:53-66) write the
 * pre-collision moments instead, with rho = 1, u = 0 at solid cells: collisions conserve
 * rho and rho u, so at fluid cells the two differ by rounding only; the solid-cell values
 * are the caller's to set (the host runner does, for its TRT path). */
int kmi_moments(kmi_lattice* lat, double* rho, double* ux, double* uy, double* uz,
                double* T);
int kmi_reduce(kmi_lattice* lat, int32_t kind, int32_t arg, double* out);
/* Steps taken since creation (or since the last upload/init). */
int kmi_lattice_steps(const kmi_lattice* lat, int64_t* steps);
/* Set the step counter (a restored checkpoint resumes its cycle count, state.jl:58-70). */
int kmi_lattice_set_steps(kmi_lattice* lat, int64_t steps);
/* Bytes of device memory held by the lattice: the population buffers, mask, residual
 * snapshot, and the staging block of host transfers (uploads, downloads, moments and
 * initialisations reuse it; it grows to the largest transfer and is kept). */
int kmi_lattice_device_bytes(const kmi_lattice* lat, int64_t* bytes);

/* ---- launch tuning (no reference counterpart: KA picks its own workgroup size;
 *      src/kernels/persistent_bgk_2d.jl:17-56 exposes a `workgroupsize` knob) --------- */
/* Work-group of bx × by cells (bx*by a multiple of 64, at most 256) and non-temporal
 * population stores on/off.  Results do not depend on these settings. */
int kmi_lattice_set_launch(kmi_lattice* lat, int32_t bx, int32_t by, int32_t nontemporal);
int kmi_lattice_get_launch(const kmi_lattice* lat, int32_t* bx, int32_t* by, int32_t* nontemporal);
/* Several steps per launch (temporal blocking); the intermediate populations stay in LDS,
 * so HBM sees one read and one write of the field per launch.  tile 0: off.
 * D3Q19 BGK/TRT lattices without thermal populations, nz >= 4, nx and ny multiples of the
 * tile (zchunks: z-ranges per tile column, 1..nz).  Fully periodic lattices without mask
 * take every tile below; lattices with walls, moving walls (Ladd), constant Zou–He velocity
 * or pressure faces or a solid mask (no forcing, no cut links; not F16S) take tiles 12 and
 * 14, and 9 except in FP64 — a link bounced at a wall reads the cell's own opposite
 * population of the previous step from LDS, as the single-step pull reads it from HBM:
 *   two steps, one work-item per interior cell: 1 = 32x8, 2 = 64x4, 3 = 64x8, 4 = 32x16,
 *     5 = 64x16 (not F64);
 *   two steps, one work-item per cell of the tile and its one-cell ring: 12 = 64x8,
 *     13 = 32x16, 14 = 32x8;
 *   three steps, one work-item per interior cell: 6 = 64x8 (not F64), 7 = 32x8,
 *     8 = 32x16 (not F64);
 *   three steps, one work-item per cell of the tile and its two-cell ring: 9 = 32x8,
 *     10 = 32x16 (not F64), 11 = 64x8 (not F64).
 * D2Q9 lattices without cut links or per-node faces (thermal: not F64), at least the tile
 * plus its ring on each axis, zchunks 1; the tile and a ring of as many cells as steps are
 * advanced in LDS, one work-item per cell: 1 = 28x28 two steps, 2 = 24x24 four, 3 = 56x8
 * four, 4 = 16x16 eight, 5 = 60x12 two, 6 = 8x8 four, 7 = 16x16 four, 8 = 8x8 eight.
 * Results are bit-identical to single-step launches; step counts that are not a multiple
 * of the span end with single steps. */
int kmi_lattice_set_temporal(kmi_lattice* lat, int32_t tile, int32_t zchunks);
int kmi_lattice_get_temporal(const kmi_lattice* lat, int32_t* tile, int32_t* zchunks);
/* Time every candidate shape × store policy (and two-step variant, where it applies) on
 * this lattice for ~steps_per_trial steps each, keep the fastest, and restore the
 * populations, parity and step count. */
int kmi_lattice_autotune(kmi_lattice* lat, int64_t steps_per_trial, double* best_ms_per_step);
/* Device copy bandwidth (float4 loads, non-temporal stores) over `bytes`, GB/s counting
 * read + write, averaged over `reps` copies: the practical HBM ceiling of a 1:1 kernel. */
int kmi_measure_copy_bandwidth(kmi_ctx* ctx, int64_t bytes, int32_t reps, double* gbps);

/* ---- batched lid-driven cavities (no reference counterpart: the reference runs one
 *      cavity per driver call, e.g. test/test_mrt.jl:19-25 and basic.jl:112-144) -------
 * An ensemble of square D2Q9 cavities of uneven side N_c and viscosity nu_c advanced
 * together: every step is ONE launch over a packed list of tiles (tw × th cells of one
 * cavity) walked by a persistent grid; tiles are ordered so that neighbouring tiles
 * share an XCD.  Each cavity: halfway bounce-back walls, lid on the north face
 * (Zou–He as run_cavity_2d, or Ladd), BGK / TRT / MRT; per-cavity residual r_u every
 * `check_every` steps; converged or diverged cavities are retired from the tile list. */
typedef struct kmi_batch kmi_batch;

typedef struct {
  int32_t storage;    /* kmi_storage */
  int32_t collision;  /* kmi_collision (rates from each cavity's nu; MRT s_e/s_eps/s_q below) */
  int32_t lid;        /* KMI_FACE_ZOUHE_VELOCITY or KMI_FACE_MOVING_WALL */
  int32_t pad;
  double u_lid;
  double s_e, s_eps, s_q; /* MRT (collide_mrt_2d.jl:118 defaults 1.4, 1.4, 1.2) */
  double Lambda;          /* TRT magic parameter (3/16) */
} kmi_batch_desc;

typedef enum {
  KMI_CAV_RUNNING = 0,
  KMI_CAV_CONVERGED = 1,
  KMI_CAV_DIVERGED = 2,   /* non-finite residual: retired without touching the others */
  KMI_CAV_MAX_STEPS = 3
} kmi_cavity_status;

int kmi_batch_create(kmi_ctx* ctx, const kmi_batch_desc* desc, int32_t ncav, const int32_t* n,
                     const double* nu, kmi_batch** out);
int kmi_batch_destroy(kmi_batch* b);
/* Per-cavity state in the reference layout f[i, j, q] (n_c × n_c × 9). */
int kmi_batch_upload(kmi_batch* b, int32_t cavity, const void* host, int32_t host_dtype);
int kmi_batch_download(kmi_batch* b, int32_t cavity, void* host, int32_t host_dtype);
int kmi_batch_init_equilibrium(kmi_batch* b, int32_t cavity, const double* rho, const double* ux,
                               const double* uy);
int kmi_batch_moments(kmi_batch* b, int32_t cavity, double* rho, double* ux, double* uy);
/* Mark every cavity running again with zero steps (populations untouched). */
int kmi_batch_reset_status(kmi_batch* b);
/* Put cavities retired at max_steps back to running (step counts kept), so a long run
 * can proceed in chunks of kmi_batch_run calls with growing max_steps. */
int kmi_batch_resume(kmi_batch* b);
/* Advance all running cavities `nsteps` steps (no residual checks). */
int kmi_batch_step(kmi_batch* b, int64_t nsteps);
/* Consistent initialisation of a prescribed velocity field (Mei et al., Comput. Fluids 35
 * (2006) 855): the current velocity of every held cavity (hold[c] != 0; hold = NULL holds
 * all) is kept while the batch runs `iterations` steps, after each of which a held cell's
 * populations get f_q += feq_q(rho, u_held) - feq_q(rho, u).  Density and the
 * non-equilibrium part relax to values consistent with the held field (an feq-only start
 * lacks them at the walls and excites slowly decaying modes).  The steps count for every
 * running cavity, held or not.  No reference counterpart: the reference initialises from
 * feq only (initialize_2d, basic.jl:51-70; gnf_periodic_steady.jl:114-150). */
int kmi_batch_consistent_init(kmi_batch* b, const int32_t* hold, int32_t iterations);
/* Step until every cavity is converged (r < tol), diverged or at max_steps, checking every
 * `check_every` (>= 2) steps; elapsed device milliseconds in *ms (may be NULL).
 * r = ||u^k - u^(k-m)||_2 / (m ||u^k||_2) over the last `window` = m steps (1 <= m <=
 * check_every): m = 1 is the reference's one-step r_u (gnf_periodic_steady.jl:192); a
 * longer window averages FP32 rounding noise, whose one-step floor is ~4e-7. */
int kmi_batch_run(kmi_batch* b, double tol, int64_t check_every, int64_t window, int64_t max_steps,
                  double* ms);
/* Arrays of ncav entries (any may be NULL). */
int kmi_batch_status(const kmi_batch* b, int32_t* status, int64_t* steps, double* residual);
/* Tiles of tw × th cells (tw*th = 256, tw in {16,32,64,128,256}), `groups` persistent
 * work-groups (0: one per tile slot), non-temporal stores on/off. */
int kmi_batch_set_launch(kmi_batch* b, int32_t tw, int32_t th, int32_t groups, int32_t nontemporal);
int kmi_batch_get_launch(const kmi_batch* b, int32_t* tw, int32_t* th, int32_t* groups,
                         int32_t* nontemporal);
/* Steps per launch: 1; 2 (the tile's step t+1, with a one-cell ring, kept in LDS; HBM
 * sees one read and one write per population per two steps); 3 or 4 (the tile and a
 * 3- or 4-cell ring advanced in LDS, one work-item per cell of that region; tiles of width
 * 16, 32 or 64).  Bit-identical to single steps; step counts that are not a multiple end
 * with two-step and single-step launches.  Chosen by kmi_batch_autotune. */
int kmi_batch_set_temporal(kmi_batch* b, int32_t steps_per_launch);
int kmi_batch_get_temporal(const kmi_batch* b, int32_t* steps_per_launch);
/* Time every candidate tile shape × group count × store policy and keep the fastest;
 * populations, parity, status and step counts are restored. */
int kmi_batch_autotune(kmi_batch* b, int64_t steps_per_trial, double* best_ms_per_step);
/* Cells of the running cavities (the work of one step). */
int kmi_batch_active_cells(const kmi_batch* b, int64_t* cells);

/* ---- learned warm start (no reference counterpart; it produces initial fields, the
 *      role of f_init in src/drivers/gnf_periodic_steady.jl:114-150) ------------------
 * A conv decoder: x (in_dim) → dense+SiLU (hidden) → dense+SiLU (latent hw×hw×c0, NHWC)
 * → 3 × [nearest ×2 up-sampling, 3×3 conv] with SiLU after the first two → an output
 * image of (8 hw)² × out_c (NHWC).  Trained and evaluated on the GPU by implicit-GEMM
 * kernels on BF16 MFMA tiles with FP32 accumulation; FP32 master weights, Adam.
 * Flat parameter order: W1[in_dim][hidden], b1, W2[hidden][hw*hw*c0], b2,
 * C1[9*c0][c1] (rows (ky, kx, c)), bc1, C2[9*c1][c2], bc2, C3[9*c2][out_c], bc3. */
typedef struct kmi_nn kmi_nn;

typedef struct {
  int32_t in_dim, hidden, latent_hw, latent_c, c1, c2, out_c;
  int32_t max_batch;
  uint64_t seed; /* He-normal weight initialisation */
} kmi_nn_desc;

void kmi_nn_desc_init(kmi_nn_desc* d); /* 2, 256, 16, 64, 64, 32, 3; batch 32 */
int kmi_nn_create(kmi_ctx* ctx, const kmi_nn_desc* d, kmi_nn** out);
int kmi_nn_destroy(kmi_nn* nn);
int kmi_nn_param_count(const kmi_nn* nn, int64_t* n);
int kmi_nn_get_params(kmi_nn* nn, float* host);
int kmi_nn_set_params(kmi_nn* nn, const float* host); /* also resets the Adam state */
/* y: batch × (8hw)² × out_c floats (NHWC); x: batch × in_dim. */
int kmi_nn_forward(kmi_nn* nn, int32_t batch, const float* x, float* y);
/* MSE loss mean((y − target)²) and, when grad != NULL, d loss / d params (flat order). */
int kmi_nn_loss_grad(kmi_nn* nn, int32_t batch, const float* x, const float* target, double* loss,
                     float* grad);
int kmi_nn_train_step(kmi_nn* nn, int32_t batch, const float* x, const float* target, float lr,
                      double* loss);
/* Device-resident training set and epochs of shuffled mini-batches (Fisher–Yates per epoch
 * from `seed`); mean loss of the last epoch in *mean_loss. */
int kmi_nn_set_dataset(kmi_nn* nn, int32_t n, const float* x, const float* y);
int kmi_nn_train_epochs(kmi_nn* nn, int32_t epochs, int32_t batch, float lr, uint64_t seed,
                        double* mean_loss);
/* C = A·B with BF16 operands on MFMA and FP32 accumulation (host FP32 in/out); A(m, k) =
 * A[m*sam + k*sak], B(k, n) = B[k*sbk + n*sbn]; the K range split over `splits` groups. */
int kmi_gemm_bf16(kmi_ctx* ctx, int32_t M, int32_t N, int32_t K, const float* A, int64_t sam,
                  int64_t sak, const float* B, int64_t sbk, int64_t sbn, int32_t splits, float* C);

/* Bytes of the static device workspace of the allocation-free entry point below (per GPU). */
#define KMI_NOALLOC_WORKSPACE_BYTES (32u * 1024u * 1024u)
/* kmi_gemm_bf16 without any device allocation: no counterpart in the reference.  For callers
 * that cannot allocate device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  The same kernels and result bits as kmi_gemm_bf16; the
 * device copies of A and B (FP32 and BF16) and the split-K partials are carved from a static
 * workspace of KMI_NOALLOC_WORKSPACE_BYTES per GPU, held until the process ends, so the call
 * allocates and frees nothing.  Problems that need more than the workspace are rejected with
 * KMI_ERR_ARG (each buffer rounded up to 256 bytes: 6 (na + nb) + 4 splits M N bytes, na and nb
 * the elements the strides of A and B span).  Every call on one GPU uses the same workspace and
 * leaves its contents behind; calls that overlap on one GPU (from contexts on different
 * streams or from different threads) race. */
int kmi_gemm_bf16_noalloc(kmi_ctx* ctx, int32_t M, int32_t N, int32_t K, const float* A, int64_t sam,
                          int64_t sak, const float* B, int64_t sbk, int64_t sbn, int32_t splits, float* C);

#ifdef __cplusplus
}
#endif
#endif /* KRAKEN_MI355X_H */
