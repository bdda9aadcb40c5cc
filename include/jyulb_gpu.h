/*
 * jyulb_gpu.h -- C ABI of the MI355X D3Q19 evolve engine (libjyulb_hip.so).
 *
 * The engine is the device half of the reference's jyulb.isothermal.Solver:
 * everything that Solver.evolve / finalize_init do per lattice site runs here
 * in hand-written HIP kernels for gfx950; the Python Solver (jyulb/isothermal.py)
 * keeps the host-side API, validation and array aliasing and calls these
 * functions through ctypes. Plain C types only; no torch types.
 *
 * Array conventions: host arrays are the *interior* of the reference's
 * halo-padded fields in C order, i.e. the contiguous copy of
 * get_phase()/get_den()/get_vel()/get_acc() (isothermal.pyx:1465-1512):
 *   phase u8 [nx][ny][nz], den f64 [nx][ny][nz], vel/acc f64 [nx][ny][nz][3].
 * Face data use the reference's halo-padded face shapes (isothermal.pyx:253-263):
 *   faces 1X/X1: [ny+2][nz+2](,3), 1Y/Y1: [nx+2][nz+2](,3), 1Z/Z1: [nx+2][ny+2](,3).
 * Population order l = 0..18 as jyulb/D3Q19.pxd:59-79.
 *
 * Every function returns LB_OK (0) or an error code; lb_last_error() returns
 * the message of the last failure on the calling thread. Calls on one handle
 * must be serialised by the caller (the reference Solver is not re-entrant
 * either); device work is ordered on the handle's stream.
 */
#ifndef JYULB_GPU_H
#define JYULB_GPU_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define LB_OK 0
#define LB_EINVAL 1   /* invalid argument: ValueError in the reference */
#define LB_ESTATE 2   /* call out of order: RuntimeError in the reference */
#define LB_EHIP 3     /* HIP runtime failure (no device, launch error, ...) */
#define LB_ENOMEM 4   /* device allocation failed */

/* Storage / arithmetic variants of the step kernel */
#define LB_FP64 0          /* f64 populations, f64 arithmetic: the reference's numerics */
#define LB_FP32_SHIFTED 1  /* f32 storage of f - w*rho0, f32 arithmetic */
#define LB_FP16_STORAGE 2  /* f16 storage of s*(f - w*rho0), f32 arithmetic */
#define LB_FP32_STORAGE 3  /* f32 storage of f - w*rho0, f64 arithmetic */

/* Population layouts in device memory */
#define LB_LAYOUT_ROWS 0        /* f[i][j][l][k]: the 19 populations of a lattice row side by side */
#define LB_LAYOUT_SOA 1         /* f[l][i][j][k] */
#define LB_LAYOUT_SOA_PADDED 2  /* f[l][i][j][k], population stride padded by 2304 bytes */
#define LB_LAYOUT_SPARSE 3      /* fluid sites only: f[l][n] over the compacted fluid sites, with a table of
                                   the 18 odd-step locations per site (the fluid-only AA arrays of
                                   include/kernel/kernel.cpp:227-271), built on the device and read
                                   as 4-bit codes against per-wave bases; allocated by lb_finalize */
#define LB_LAYOUT_AUTO 4        /* chosen at the first lb_finalize: SPARSE when at least 5% of the cells are
                                   solid (LB_AUTO_SPARSE_SOLID in the environment sets the fraction), ROWS
                                   otherwise; results are identical in every layout */

typedef struct lb_solver lb_solver;

typedef struct {
    int32_t lat[3];      /* interior lattice size (nx, ny, nz): Solver(lat_size), isothermal.pyx:156 */
    int32_t rlx_scheme;  /* BGK=0, TRT=1, REGULARIZATION=2 (defs.pxd:86-90) */
    double kvisc;        /* kinematic viscosity; tau = 0.5 + 3 kvisc (isothermal.pyx:273-287) */
    double g[3];         /* gravity */
    int32_t bc[6];       /* per face, WALL=0 PERIODIC=1 FIXED_DEN=2 FIXED_VEL=3 SYMMETRIC=4 */
    int32_t precision;   /* LB_FP64 / LB_FP32_SHIFTED / LB_FP16_STORAGE / LB_FP32_STORAGE */
    double rho0;         /* reference density of the shifted variants (1.0 if 0) */
    double fp16_scale;   /* storage scale s of LB_FP16_STORAGE (0: chosen at lb_finalize) */
    int32_t layout;      /* LB_LAYOUT_* */
} lb_config;

/* Message of the last failed call on this thread ("" if none). */
const char *lb_last_error(void);

/* Number of HIP devices visible (0 if none); never fails. */
int lb_device_count(void);

/* Device state of Solver.__cinit__ (isothermal.pyx:156-298): validates the
 * configuration like _check_bc (isothermal.pyx:378-426) and allocates the
 * device buffers on the current HIP device. */
int lb_create(const lb_config *cfg, lb_solver **out);
int lb_destroy(lb_solver *s);

/* Stream for all device work of this handle (hipStream_t; NULL = default). */
int lb_set_stream(lb_solver *s, void *hip_stream);

/* gravity setter, isothermal.pyx:369-373 */
int lb_set_gravity(lb_solver *s, const double g[3]);

/* Interior phase field as written through get_phase() before finalize_init. */
int lb_upload_phase(lb_solver *s, const uint8_t *phase);

/* den/vel through get_den()/get_vel(): the initial field before lb_finalize,
 * later the wall velocities of SOLID sites and the densities that Ladd's
 * moving-wall term reads in the first step of an evolve call
 * (isothermal.pyx:1117-1119). Device copies are replaced entirely. */
int lb_upload_den_vel(lb_solver *s, const double *den, const double *vel);

/* Per-site acceleration field (get_acc(), isothermal.pyx:1293-1295);
 * NULL means the field is identically zero. */
int lb_upload_acc(lb_solver *s, const double *acc);

/* Zou-He data of one FIXED_DEN/FIXED_VEL face: the arrays that
 * _set_bc_data copies from den/vel (isothermal.pyx:476-502). */
int lb_set_face_data(lb_solver *s, int face, const double *den_face, const double *vel_face);

/* finalize_init, device part (isothermal.pyx:1333-1365): link masks from the
 * phase field (halo passes of _enforce_phase_per_sym_bc folded into address
 * resolution) and the initial populations of _init_f (isothermal.pyx:1028-1105).
 * Returns the number of fluid sites in *nfluid. */
int lb_finalize(lb_solver *s, int64_t *nfluid);

/* The step loop of Solver.evolve (isothermal.pyx:1408-1415): `steps` fused
 * AA stream+collide steps. den/vel of the fluid sites are stored on the last
 * step. Asynchronous on the handle's stream, except that the FP16 storage
 * variant checks its stored magnitudes (and rescales them by a power of two)
 * at the start and every 64 steps (LB_FP16_CHECK), failing with LB_ESTATE if
 * a population overflowed. A recorded adjoint forward pass no longer applies
 * afterwards (see lb_adj_backward). */
int lb_evolve(lb_solver *s, uint32_t steps);

/* Pinned host memory (hipHostMalloc) for the Solver's own arrays, so that
 * the copies of lb_evolve_host run at full DMA speed; lb_host_free releases
 * it. */
int lb_host_alloc(size_t bytes, void **ptr);
int lb_host_free(void *ptr);

/* The Solver's halo-padded host fields (isothermal.pyx:243-248): phase u8
 * [NI][NJ][NK], den f64 [NI][NJ][NK], vel and acc f64 [NI][NJ][NK][3] with
 * NI = nx + 2, NJ = ny + 2, NK = nz + 2. lb_evolve_host reads and writes them
 * in place. Any host memory works; lb_host_alloc memory is faster. They must
 * stay valid while bound; four NULLs unbind them. Allocates a device copy of
 * the four fields. */
int lb_bind_host_fields(lb_solver *s, uint8_t *phase, double *den, double *vel, double *acc);

/* Solver.evolve on the bound host fields (isothermal.pyx:1370-1418), after
 * the caller has refreshed the Zou-He data (_set_bc_data, lb_set_face_data)
 * and gravity. Per call: the host fields are copied to the device as they
 * are; a changed fluid/solid state of any site fails with LB_ESTATE (the
 * reference re-reads phase every step, the engine builds its links once); the
 * acceleration field, the wall velocities (moving walls re-derived when they
 * changed) and the densities Ladd's term reads in the first step come from
 * them; then `steps` steps; then den/vel of the fluid sites are written back
 * into the host fields, with the den/vel halo pass of periodic and symmetric
 * faces (_enforce_den_vel_per_sym_bc, isothermal.pyx:582-686) applied to the
 * state of step steps-1, as the reference leaves it. SOLID sites keep their
 * host values. Synchronous. */
int lb_evolve_host(lb_solver *s, uint32_t steps);

/* The first part of lb_evolve_host alone: the bound host fields taken as the
 * inputs of the next steps (phase check, acceleration field, wall
 * velocities, Ladd densities), e.g. before lb_adj_forward. Synchronous. */
int lb_load_host_fields(lb_solver *s);

/* den/vel of the fluid sites of the last stored step into the bound host
 * fields (SOLID sites and halos keep their host values), followed by the
 * den/vel halo pass if halo_pass != 0. Synchronous. */
int lb_store_host_fields(lb_solver *s, int halo_pass);

/* Re-derive the moving-wall flags after lb_upload_den_vel changed wall
 * velocities (called by the Python Solver before evolve when needed). */
int lb_refresh_walls(lb_solver *s);

/* den/vel of the last evolve (fluid sites; SOLID sites keep the uploaded values). */
int lb_download_den_vel(lb_solver *s, double *den, double *vel);

/* Post-collision populations as the reference's f1 after the same number of
 * steps, unpacked to f64 [19][nx][ny][nz] (entries at SOLID sites are 0). */
int lb_download_post_f(lb_solver *s, double *f);

/* The reference's f2 (isothermal.pxd:150-157): after evolve, the
 * post-collision populations of the step before the last one. The engine
 * keeps one population array, so it copies the state before the last step of
 * every evolve call once lb_keep_previous(s, 1) was called (lb_keep_previous(s,
 * 0) frees the copy); lb_download_previous_post_f unpacks it like
 * lb_download_post_f. LB_ESTATE if no evolve call kept one since
 * lb_finalize. */
int lb_keep_previous(lb_solver *s, int on);
int lb_download_previous_post_f(lb_solver *s, double *f);

/* FP64 reductions over the fluid sites of the last stored den/vel:
 * out[0] = sum den, out[1..3] = sum vel, out[4..6] = sum den*vel,
 * out[7] = max |vel|. (The quantities of calc_flow_info,
 * include/solver/solver.cpp:64-129, and of the examples' mass checks.) */
int lb_reduce(lb_solver *s, double out[8]);

/* calc_flow_info (include/solver/solver.cpp:64-129; the fields of
 * IsothermalFlowInfo, include/solver/solver.h:13-15) over the fluid sites of
 * the last stored den/vel, in FP64 on the device for every precision
 * variant: totals of momentum (j = den u), of |j| = den |u| (tot_j2) and of
 * density; averages of u and of |u| (ave_u2); max |u| (max_u2) and the
 * density extrema, which start from the reference density rho0 as in the
 * reference. fluid_sites (an addition) is the number of sites summed. */
typedef struct {
    double tot_jx, tot_jy, tot_jz, tot_j2, tot_den;
    double ave_ux, ave_uy, ave_uz, ave_u2;
    double max_u2, max_den, min_den;
    double fluid_sites;
} lb_flow_info_t;
int lb_flow_info(lb_solver *s, lb_flow_info_t *info);

/* No counterpart in the reference. The calc_flow_info of lb_flow_info for
 * callers that cannot allocate device memory (code under HIP graph capture,
 * tight serving loops): the fields are caller-provided device arrays in the
 * interior C-order layout (phase u8 [nx][ny][nz], 0 = SOLID; den f64
 * [nx][ny][nz]; vel f64 [nx][ny][nz][3]), the result is written to the
 * device array info_dev (13 doubles in lb_flow_info_t order), and the
 * per-block partials live in a fixed 104 KiB array of the library on each
 * device, so the call allocates nothing; it launches two kernels on
 * hip_stream (NULL: the default stream) and returns without synchronising.
 * At most 2^31 - 1 sites (nx * ny * nz); larger lattices are refused with
 * LB_EINVAL. Calls that overlap on one device share that array and race:
 * order them on one stream. The first call on a device resolves the array's
 * address there (make it outside graph capture). */
int lb_flow_info_fields_noalloc(const uint8_t *phase_dev, const double *den_dev, const double *vel_dev,
                                const int32_t lat[3], double rho0, double *info_dev, void *hip_stream);

/* Steady-state test in FP64: ||u - u_prev||_2 / ||u||_2 over the fluid
 * sites, then u_prev := u. The first call returns 1.0. */
int lb_convergence(lb_solver *s, double *rel_change);

/* Device-pointer hand-off (e.g. torch data_ptr()): same layouts as the host
 * variants, pointers in device memory, ordered on the handle's stream. */
int lb_upload_acc_device(lb_solver *s, const double *acc_dev);
int lb_download_den_vel_device(lb_solver *s, double *den_dev, double *vel_dev);

/* Zou-He face data from device memory (same layout as lb_set_face_data). */
int lb_set_face_data_device(lb_solver *s, int face, const double *den_dev, const double *vel_dev);

/* Objective of the adjoint, evaluated on the last stored den/vel: mean over
 * the fluid sites of the face layer `face` of rho * u*_axis (the mass flux
 * through that layer, e.g. face X1, axis 0 = mean outlet x-flux). FP64. */
int lb_objective(lb_solver *s, int face, int axis, double *J);

/* Objective selector of lb_adj_forward: instead of a face flux, the mean over
 * the fluid sites of Pi_neq:Pi_neq, the squared non-equilibrium momentum flux
 * (proportional to the viscous dissipation) of the last step's collision
 * input; `axis` is ignored. */
#define LB_OBJ_DISSIPATION 6

/* Adjoint, forward pass (no counterpart in the reference: reverse-mode
 * differentiation of `steps` evolve steps with fixed parameters). Runs the
 * steps like lb_evolve, keeping a checkpoint of the population array every
 * `interval` steps, and returns the objective J: lb_objective(face, axis)
 * after the last step, or with face = LB_OBJ_DISSIPATION the dissipation
 * objective. The last segment also records (rho, u*) per site and step
 * (min(interval, steps) x 4 rows of the compute type: nq values per row in
 * the dense layouts, the fluid sites rounded up to 64 in the sparse layout).
 * The steps are
 * pull steps where lb_evolve's are (lb_pull_steps). FP64 or FP32 precision;
 * stationary walls only. */
int lb_adj_forward(lb_solver *s, uint32_t steps, uint32_t interval, int face, int axis, double *J);

/* Adjoint, reverse pass: gradients of J from the last lb_adj_forward with
 * respect to the per-site acceleration field (f64 [nx][ny][nz][3]) and to the
 * velocity data of Zou-He face `face` (f64, face layout x 3; face < 0: none).
 * Each checkpoint segment is recomputed forward, recording (rho, u*) per site
 * and step, then stepped back with the adjoint AA step (from 1 GiB of
 * cotangents its pull form, which writes a second cotangent array;
 * LB_ADJ_PULL=0/1 in the environment decides instead; the results are
 * identical). The last segment uses
 * the forward pass's record instead, unless a call since then changed an input
 * or an earlier reverse pass overwrote it (interval < steps). With interval >=
 * steps nothing is recomputed. The population array is left in its state after
 * the forward pass. Null outputs are skipped. Refused (LB_ESTATE) unless it
 * follows its forward pass directly: an evolve step, an input upload, a
 * finalize or a restore in between invalidates the forward pass. */
int lb_adj_backward(lb_solver *s, int face, double *grad_acc, double *grad_face);
int lb_adj_backward_device(lb_solver *s, int face, double *grad_acc_dev, double *grad_face_dev);

/* One snapshot slot of the device state (populations, den/vel, step count);
 * lb_restore returns to it. Used for finite-difference checks and for
 * re-running a segment (no counterpart in the reference). */
int lb_snapshot(lb_solver *s);
int lb_restore(lb_solver *s);

/* Introspection: number of AA steps taken, parity of the next step (1 = odd),
 * bytes of device memory held, storage scale of the FP16 variant. */
int lb_info(lb_solver *s, uint64_t *steps, int *next_odd, uint64_t *device_bytes, double *fp16_scale);

/* Introspection of the sparse layout's odd-step table (no counterpart in the
 * reference, which has no sparse layout): the number of links, over all
 * fluid sites, coded as a near offset, the wave's alternate location,
 * bounce-back, and a full 32-bit location, in counts[0..3]. LB_ESTATE unless
 * the solver is finalized in the sparse layout. */
int lb_sparse_codes(lb_solver *s, uint64_t counts[4]);

/* Whether the solver steps by pull steps (no counterpart in the reference):
 * a second population array into which each odd step stores line-aligned.
 * The sparse layout takes them from 1 GiB of populations in FP64, FP32 and
 * FP32 storage; LB_SP_PULL=0/1 in the environment decides instead, also for
 * the dense layouts. Results are identical either way. */
int lb_pull_steps(lb_solver *s, int *on);

/* The population layout in use (LB_LAYOUT_*; LB_LAYOUT_AUTO until the first
 * lb_finalize resolves it). */
int lb_layout(lb_solver *s, int *layout);

/* Device-to-device copy of `bytes` (no counterpart in the reference): one
 * 8-byte element per thread, faster than hipMemcpy's device copy on MI355X
 * (5.8 against 4.9 TB/s, read + write, on 2 GiB). The engine uses it for the adjoint's
 * checkpoints and the snapshot slot; bench.py times it as the HBM streaming
 * reference. Asynchronous on hip_stream (NULL: the default stream). */
int lb_stream_copy(void *dst, const void *src, size_t bytes, void *hip_stream);

/* Synchronise the handle's stream. */
int lb_synchronize(lb_solver *s);

#ifdef __cplusplus
}
#endif
#endif
