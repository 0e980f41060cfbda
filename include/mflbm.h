/*
 * mflbm.h -- C ABI of the MI355X (gfx950) colour-gradient lattice Boltzmann time step.
 *
 * Drop-in replacement for the GPU boundary of MF-LBM-CUDA.  The reference exposes that
 * boundary as process globals plus void functions; here one opaque handle owns all device
 * state and every entry point returns a status (0 = ok, < 0 = error, message via
 * mflbm_last_error).  Nothing here exits the process.
 *
 * Array layouts on upload/download are the reference's (includes/Idx_cpu.h:49-66):
 * x fastest, then y, z; 1-based interior coordinates with ghost layers:
 *   s1 = (nx+2)(ny+2)(nz+2)   one ghost layer   (curv, rho/u/v/w)
 *   s2 = (nx+4)(ny+4)(nz+4)   two ghost layers  (cn_x/y/z, c_norm, walls)
 *   s4 = (nx+8)(ny+8)(nz+8)   four ghost layers (phi, walls_type, s_n*)
 *   pdf = 38 * s1             [x, y, z, direction e (19), colour g (2)]
 * Element type is double for MFLBM_FP64 handles and float for MFLBM_FP32 handles.
 */
#ifndef MFLBM_H
#define MFLBM_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct mflbm_handle mflbm_handle;

enum mflbm_precision { MFLBM_FP64 = 0, MFLBM_FP32 = 1 };

/* Step constants.  Field meaning follows the reference globals copied to the device by
 * copyConstantData (src/main_iteration_GPU.cu:14-47) and the host flags read by
 * main_iteration_kernel_GPU (src/main_iteration_GPU.cu:1890-1956).  Derived values
 * (la_nui*, cos_theta, phi_inlet, rho_in/out) are passed already derived, as
 * src/IO_multiphase.cpp:200-206 and src/Init_multiphase.cpp:173-214 derive them. */
typedef struct mflbm_params {
    int64_t nx, ny, nz;
    int32_t iper, jper, kper;
    int32_t wall_x_min, wall_x_max, wall_y_min, wall_y_max, wall_z_min, wall_z_max;
    int32_t inlet_bc, outlet_bc; /* inlet: 1 velocity, 2 Zou-He pressure; outlet: 1 convective, 2 Zou-He */
    double lbm_gamma, force_z, la_nui1, la_nui2, lbm_beta, phi_inlet, sa_inject, cos_theta;
    double rho_in, rho_out, relaxation;
    int32_t precision;           /* enum mflbm_precision */
    int32_t flags;               /* MFLBM_FLAG_* */
    int32_t porous_plate;        /* porous_plate_cmd: 0 none, 1 blocks fluid 1, 2 blocks fluid 2 */
    int32_t z_porous_plate;      /* Z_porous_plate, 1 .. nz */
    double uin_avg;              /* mean inlet velocity (velocity inlet, convective outlet):
                                  * capillary_number * surface_tension / fluid1_viscosity */
    int32_t mrt;                 /* collision model, the reference's compile-time `mrt`
                                  * (includes/preprocessor.h:3-4; src/main_iteration_GPU.cu:157-187):
                                  * 1 MRT bounce-back optimised, 2 MRT original (its default),
                                  * 3 SRT, 4 MRT advection optimised; 0 means 2 */
} mflbm_params;

/* Run the colour-gradient chain as four kernels per step, one per reference kernel
 * (phi extrapolation, normals + wetting, normal extrapolation, curvature), writing the
 * dense fields every step.  Default (0): one fused kernel that writes only what the next
 * collision reads; the dense fields are brought up to date when downloaded.  Both give
 * identical results. */
#define MFLBM_FLAG_SPLIT_GRADIENT 1

/* Fields for upload/download. */
enum mflbm_field {
    MFLBM_PDF = 0,     /* 38*s1 */
    MFLBM_PHI = 1,     /* s4 */
    MFLBM_CN_X = 2, MFLBM_CN_Y = 3, MFLBM_CN_Z = 4, MFLBM_C_NORM = 5, /* s2 */
    MFLBM_CURV = 6,    /* s1 */
    MFLBM_S_NX = 7, MFLBM_S_NY = 8, MFLBM_S_NZ = 9, /* s4 */
    MFLBM_WALLS = 10,      /* s2, int8 0/1 */
    MFLBM_WALLS_TYPE = 11, /* s4, int8 -1/0/1/2 */
    MFLBM_W_IN = 12,       /* (nx+2)(ny+2): inlet velocity profile of the velocity inlet (the
                            * reference's W_in plane k = 0); the duct profile at create */
    MFLBM_CONVEC = 13      /* 39 (nx+2)(ny+2): convective-outlet state, f_convec then g_convec
                            * ([e][y][x] each) then phi_convec, the checkpoint's tail */
};

/* Monitor sums (src/Monitor.cpp:17-80), accumulated in FP64 on the device for both
 * precisions.  Caller provides arrays of nz doubles; null pointers are skipped. */
typedef struct mflbm_monitor_out {
    double *fl1, *fl2, *vol1, *vol2, *mass1, *mass2, *pre; /* per z-slice, length nz */
    double umax2;                 /* max |u|^2 over fluid nodes */
    double usq1, usq2;            /* sum |u|^2 where phi > 0.999 / phi < -0.999 */
    int64_t n_fluid;              /* fluid nodes in the interior */
} mflbm_monitor_out;

/* Device time of the kernel groups of a profiled run (milliseconds, summed over steps;
 * divide collide_odd_ms by n_odd, collide_even_ms by n_even, the others by n_odd + n_even).
 * boundary = periodic copies + Zou-He; gradient = the colour-gradient chain. */
typedef struct mflbm_kernel_times {
    double collide_odd_ms, collide_even_ms, boundary_ms, gradient_ms;
    int32_t n_odd, n_even;
} mflbm_kernel_times;

/* Replaces initialization_GPU + MemAllocate_geometry_GPU(1) + MemAllocate_multi_GPU(1)
 * (src/Init_multiphase_GPU.cu:13-116) and the host geometry preparation that feeds them:
 * set_walls domain walls (src/Misc.cpp:67-91) and geometry_preprocessing_new
 * (src/Geometry_preprocessing.cpp:29-403) run on the device.  `geometry` is the
 * nx*ny*nz voxel array of a .dat file (1 solid, 0 pore), x fastest. */
int mflbm_create(const mflbm_params* params, const int8_t* geometry, mflbm_handle** out);

/* Replaces MemAllocate_*_GPU(2). */
void mflbm_destroy(mflbm_handle* h);

const char* mflbm_last_error(const mflbm_handle* h);

/* Initial order parameter of a new run (src/Init_multiphase.cpp:299-358), options 1-5. */
int mflbm_init_phi(mflbm_handle* h, int32_t option, double interface_z0);

/* Option 6, the random phase distribution (src/Init_multiphase.cpp:344-350): phi = 1, or -1
 * where the node's uniform deviate exceeds sa_target (target_fluid1_saturation), over
 * [0, n+1]^3, then the inlet ghost layers as for the other options.  The reference draws
 * rand() / RAND_MAX seeded from time(NULL), so no two of its runs agree; here each node's
 * deviate is a counter-based function of `seed` and the node's s1 index (splitmix64), so a
 * run is reproducible and independent of the launch order. */
int mflbm_init_phi_random(mflbm_handle* h, double sa_target, uint64_t seed);

/* Equilibrium populations from rho,u,v,w (s1 layout, host memory, double; any may be
 * null = rho 1 / velocity 0) and the current phi (src/Init_multiphase.cpp:389-443).
 * This is the warm-start hook. */
int mflbm_init_equilibrium(mflbm_handle* h, const double* rho, const double* u, const double* v, const double* w);

/* As mflbm_init_equilibrium, with the constant terms at the collision's own rest state,
 * rho * (4/21, 13/126, 1/63) (MRT e = -11 rho, e2 = 0), in place of the D3Q19 weights
 * (1/3, 1/18, 1/36) whose fourth-order moment e2 = 3 rho the first collision discards.
 * No counterpart in the reference: it is the initialisation of the warm start, which
 * removes the start-up mass pulse the D3Q19 weights leave in the bounce-back links. */
int mflbm_init_relaxed(mflbm_handle* h, const double* rho, const double* u, const double* v, const double* w);

/* Colour-gradient chain on the current phi: the reference's host color_gradient()
 * (src/Phase_gradient.cpp:15-291) that precedes the first step (src/main.cpp:92). */
int mflbm_color_gradient(mflbm_handle* h);

/* One time step: main_iteration_kernel_GPU without its device-to-host copy
 * (src/main_iteration_GPU.cu:1890-2055).  Parity from ntime (odd = pull/push). */
int mflbm_step(mflbm_handle* h, int32_t ntime);

/* nsteps consecutive steps ntime0 .. ntime0+nsteps-1, replayed from captured HIP graphs. */
int mflbm_run(mflbm_handle* h, int32_t ntime0, int32_t nsteps);

/* Same as mflbm_run but with per-kernel device timing (events between launches). */
int mflbm_profile(mflbm_handle* h, int32_t ntime0, int32_t nsteps, mflbm_kernel_times* out);

/* Host <-> device copies of one field in the reference layout (the reference's
 * cudaMemcpy of src/main_iteration_GPU.cu:2059-2076 and src/Init_multiphase_GPU.cu).
 * bytes must equal the field size. */
int mflbm_download(mflbm_handle* h, int32_t field, void* dst, size_t bytes);
int mflbm_upload(mflbm_handle* h, int32_t field, const void* src, size_t bytes);

/* compute_macro_vars (src/Misc.cpp:222-274) on the device: rho,u,v,w (s1, double) of the
 * current (after-even-step) populations; null outputs are skipped. */
int mflbm_macro(mflbm_handle* h, double* rho, double* u, double* v, double* w);

/* Monitor reductions of src/Monitor.cpp:17-80 on the device in FP64. */
int mflbm_monitor(mflbm_handle* h, mflbm_monitor_out* out);

/* Steady criterion of a single-phase solve (no counterpart in the reference, whose steady
 * monitors follow saturation, phase field or capillary pressure): *change = |u - u_prev| /
 * |u| over the fluid nodes, FP64 L2 norms of the compute_macro_vars velocity, u_prev being
 * the velocity at the previous call; -1 on the first call after a (re)initialisation or
 * a PDF upload.  Reduced on the device in a fixed order. */
int mflbm_velocity_change(mflbm_handle* h, double* change);

/* Waits for all work of the handle. */
int mflbm_sync(mflbm_handle* h);

/* Node counts: 0 fluid interior, 1 fluid in [-1,n+2], 2 solid-boundary in [-2,n+3].
 * which = 3 names the gradient chain the step runs: 0-4 the z-march in one of its workgroup
 * shapes, 5 the two-pass 3-D tiles, 6 the four split kernels.  which = 4: the waves of 64
 * fluid nodes whose links span too many cells for the 16-bit neighbour rows and read the
 * 32-bit ones. */
int64_t mflbm_count(const mflbm_handle* h, int32_t which);

/* Size in elements of a field. */
int64_t mflbm_field_size(const mflbm_handle* h, int32_t field);

/* Plain device-to-device copy of `bytes` (two temporary buffers), repeated `repeats` times;
 * returns the mean device time of one copy.  No counterpart in the reference: it calibrates
 * bandwidth figures and the profiler's FETCH_SIZE / WRITE_SIZE counters on the same GPU. */
int mflbm_bandwidth_probe(mflbm_handle* h, int64_t bytes, int32_t repeats, double* ms_per_copy);

/* ---------------------------------------------------------------------------------------
 * Batched single-phase ensemble.  No counterpart in the reference, which runs one domain
 * per process: here n independent domains of nx*ny*nz -- walls on both x faces, periodic in
 * y and z, body force along z, one fluid -- advance together, one launch per step over all
 * of them, each until its velocity field is steady.  It produces the warm start's training
 * fields and screens permeability.  Every domain starts from rest at the collision's rest
 * weights with the mean density of the reference's cold start (as mflbm_init_relaxed);
 * populations are FP32 deviations from that state, the reductions FP64.
 * ------------------------------------------------------------------------------------- */
typedef struct mflbm_ensemble mflbm_ensemble;

/* geometries: n_domains voxel arrays of nx*ny*nz bytes (1 solid, 0 pore, x fastest), back
 * to back; the x = 1 and x = nx planes are made solid.  Reads nx, ny, nz, la_nui1 (inverse
 * viscosity) and force_z of params.  Per-domain rest densities are written to rho_ref
 * (n_domains doubles) when it is not null. */
int mflbm_ensemble_create(const mflbm_params* params, const int8_t* geometries, int32_t n_domains,
                          double* rho_ref, mflbm_ensemble** out);
void mflbm_ensemble_destroy(mflbm_ensemble* e);
const char* mflbm_ensemble_last_error(const mflbm_ensemble* e);
int64_t mflbm_ensemble_fluid_nodes(const mflbm_ensemble* e);

/* Steps every domain not yet steady; after every `check` steps (even) the relative L2
 * change of each domain's velocity field since the previous check is computed, and a domain
 * whose change is below tol stops there.  Returns when all have stopped or after at most
 * max_steps more steps; a later call continues where this one stopped.  Per domain (arrays
 * of n_domains, null = skip): steps taken so far (negated while not yet steady);
 * permeability nu <w> / F over the whole domain at its last check (FP64); the last relative
 * change.  device_ms (nullable) receives the device time of this call's step kernels. */
int mflbm_ensemble_solve(mflbm_ensemble* e, int32_t check, double tol, int32_t max_steps, int32_t* steps,
                         double* permeability, double* change, double* device_ms);

/* Fields of domains first .. first+count-1 at their last check, written to out as
 * count x 4 x nz x ny x nx floats: u, v, w (compute_macro_vars: j - F/2) and the density
 * minus the domain's rest density (rho_ref of mflbm_ensemble_create); zero at solid nodes. */
int mflbm_ensemble_fields(mflbm_ensemble* e, int32_t first, int32_t count, float* out);

/* Initial state of domains first .. first+count-1 from fields in the layout of
 * mflbm_ensemble_fields (count x 4 x npt: u, v, w, rho - rho_ref): the relaxed equilibrium
 * of mflbm_init_relaxed at those values.  Only before the domains' first step; the first
 * steady test then compares with this state.  (No counterpart in the reference: the warm
 * start of the ensemble.) */
int mflbm_ensemble_init_fields(mflbm_ensemble* e, int32_t first, int32_t count, const float* fields);

#ifdef __cplusplus
}
#endif

#endif /* MFLBM_H */
