/*
 * swals_hip.h -- C ABI of the MI355X (gfx950) time-stepping backend for SWALS.
 *
 * The backend replaces the per-step loops behind SWALS's Fortran domain_type API.
 * A Fortran host binds these symbols with ISO_C_BINDING interfaces
 * (ptha_amd/fortran/swals_hip_mod.f90); Python binds them with ctypes
 * (ptha_amd/backend.py). Signatures use plain pointers, sizes and C scalars only.
 *
 * Conventions
 *  - Every function returns 0 on success and a negative SWH_E* code on failure;
 *    swh_last_error() returns the message. A Fortran caller maps a non-zero status to
 *    generic_stop (propagation/SWALS/src/util/stop_mod.f90:8-33), which is how the
 *    reference reports fatal errors.
 *  - Host arrays are in the reference's layout: U is Fortran U(nx, ny, 4),
 *    column-major, i fastest, planes [STG, UH, VH, ELV]
 *    (propagation/SWALS/src/shallow_water/domain_mod.f90:79,332). 2-D fields are
 *    (nx, ny). Indices passed in (gauge cells) are 1-based (i, j) pairs.
 *  - Host arrays are FP64 (the reference's default real(dp) = C_DOUBLE,
 *    global_mod.f90:15-20) in both precision modes; the device converts.
 *  - Calls on one handle must come from one host thread; the handle owns its device
 *    memory and its HIP stream.
 *
 * Reference interfaces replaced are cited per entry point (paths relative to
 * propagation/SWALS/src/shallow_water/ unless stated).
 */
#ifndef SWALS_HIP_H
#define SWALS_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Status codes */
#define SWH_OK 0
#define SWH_EINVAL (-1)    /* invalid argument */
#define SWH_EHIP (-2)      /* HIP runtime error */
#define SWH_ESTATE (-3)    /* call not valid in the current state */
#define SWH_ENOMEM (-4)    /* device allocation failed */

/* timestepping_method ids: the index of the method in the reference's
 * timestepping_metadata table (timestepping_metadata_mod.f90:69-140). */
#define SWH_METHOD_RK2 1
#define SWH_METHOD_RK2N 2
#define SWH_METHOD_MIDPOINT 3
#define SWH_METHOD_EULER 4
#define SWH_METHOD_LINEAR 5
#define SWH_METHOD_LLPNF 6 /* 'leapfrog_linear_plus_nonlinear_friction' */
#define SWH_METHOD_LEAPFROG_NONLINEAR 8 /* 'leapfrog_nonlinear' */

/* Arithmetic of the bulk update */
#define SWH_PREC_FP64 0    /* FP64 throughout (the reference's default dp = C_DOUBLE) */
#define SWH_PREC_FP32K 1   /* FP32 state, Kahan-compensated stage (every method), FP64 metrics/reductions;
                              finite-volume kernels form the clipped cells and fluxes in FP64 */

/* max_U tracking mask bits; names and order as in the reference's
 * gridded_output_variables_max_U (util/netcdf_util_mod.f90:39-41) */
#define SWH_MAXQ_MAX_STAGE 1
#define SWH_MAXQ_MAX_SPEED 2
#define SWH_MAXQ_MAX_FLUX 4
#define SWH_MAXQ_ARRIVAL_TIME 8
#define SWH_MAXQ_TIME_OF_MAX_STAGE 16
#define SWH_MAXQ_MIN_STAGE 32

typedef struct swh_domain swh_domain;

/* Static description of one domain (domain_type fields, domain_mod.f90:137-592).
 * Metric arrays are the FP64 arrays that allocate_quantities builds
 * (domain_mod.f90:933-1014); the device copies them. */
typedef struct swh_domain_desc {
    int32_t nx, ny;                       /* domain%nx(1:2) */
    int32_t timestepping_method;          /* SWH_METHOD_* */
    int32_t precision;                    /* SWH_PREC_* */
    int32_t linear_solver_is_truely_linear;
    int32_t include_coriolis;             /* Coriolis terms (spherical builds without -DNOCORIOLIS) */
    int32_t spherical;                    /* spherical metric source terms of the FV scheme */
    int32_t friction_type;                /* 0 = 'manning', 1 = 'chezy' */
    int32_t exterior_cells_width;         /* domain%exterior_cells_width (default 2) */
    int32_t boundary_exterior[4];         /* N, E, S, W */
    int32_t max_U_mask;                   /* SWH_MAXQ_* bits; 0 disables record_max_U */
    int32_t max_U_update_frequency;       /* domain%max_U_update_frequency (>= 1) */
    double msl_linear;
    double cfl;                           /* <= 0: method default */
    double theta;                         /* <= 0: method default */
    double linear_friction_coeff;
    double maximum_timestep;              /* <= 0: 1e20 (global_mod.f90:49) */
    double arrival_stage_threshold_above_msl_linear;
    const double *distance_bottom_edge;   /* ny+1 */
    const double *distance_left_edge;     /* nx+1 */
    const double *area_cell_y;            /* ny */
    const double *coriolis;               /* ny   (may be NULL when include_coriolis == 0) */
    const double *coriolis_bottom_edge;   /* ny+1 (may be NULL when include_coriolis == 0) */
    const double *tanlat_on_radius_earth; /* ny   (may be NULL when spherical == 0) */
} swh_domain_desc;

/* Scalars of domain_type after the latest step (domain_mod.f90:160-240). */
typedef struct swh_scalars {
    double time;
    double max_dt;
    double evolve_step_dt;
    double dt_last_update;
    double boundary_flux_store[4];
    double boundary_flux_evolve_integral;
    double boundary_flux_time_integral;
    double negative_depth_fix_counter;
    int64_t nsteps_advanced;
} swh_scalars;

/* Message of the last error on this host thread (NUL-terminated, truncated to len). */
int swh_last_error(char *buf, int len);

/* Replaces allocate_quantities' device-side allocation (domain_mod.f90:852-1358). */
int swh_domain_create(const swh_domain_desc *desc, swh_domain **out);
int swh_domain_destroy(swh_domain *h);

/* Host <-> device copies of domain%U (domain_mod.f90:332) and fields. In FP32K the
 * stage is stored with its Kahan term so that it represents the FP64 stage; where depth
 * is stage - elevation (every method but the truely-linear leapfrog ones) it is taken on
 * the FP32 elevation instead: a cell dry in FP64 gets the FP32 elevation as its stage
 * (depth 0), a wet one its FP64 stage but no less than the FP32 elevation. */
int swh_upload_U(swh_domain *h, const double *U);
/* The state without the elevation: stage, UH, VH as (nx, ny) FP64 planes (any may be
 * NULL: zeros); the elevation, friction work and metrics stay as they are (a new
 * initial condition on the same grid, as the reference's set-up of U(:,:,STG:VH) for a
 * new run). In FP32K the stage is stored as swh_upload_U stores it (on the FP32
 * elevation uploaded with U where depth is stage - elevation). */
int swh_upload_state(swh_domain *h, const double *stage, const double *uh, const double *vh);
int swh_download_U(swh_domain *h, double *U);
int swh_upload_manning_squared(swh_domain *h, const double *manning_squared); /* domain_mod.f90:347 */
int swh_set_time(swh_domain *h, double time);

/* Stage boundary: boundary_stage_transmissive_normal_momentum (boundary_mod.f90:78-164)
 * evaluated on the device for the flagged sides (N, E, S, W) with the stage function
 * msl + amp * sin(2 pi t / period). Replaces the host callback domain%boundary_subroutine
 * for this case. */
int swh_set_stage_boundary_sine(swh_domain *h, double amp, double period, double msl, const int32_t sides[4]);

/* Flather boundary: flather_boundary (boundary_mod.f90:502-720) evaluated on the device
 * for the flagged sides (N, E, S, W), at the reference's update_boundary calls (after
 * the leapfrog stage half-step; after each finite-volume update). The outside state is
 * the sea at rest at max(msl_linear, elevation) (no boundary_function), or with
 * use_stage_function the stage msl + amp * sin(2 pi t / period) with zero momenta.
 * Replaces domain%boundary_subroutine => flather_boundary. Needs nx, ny >= 3. */
int swh_set_flather_boundary(swh_domain *h, const int32_t sides[4], int32_t use_stage_function, double amp,
                             double period, double msl);

/* Any of boundary_mod.f90's boundary subroutines on the flagged sides (N, E, S, W), at the
 * reference's update_boundary calls:
 *   SWH_BC_STAGE_TRANSMISSIVE_NORMAL_MOMENTUM  boundary_stage_transmissive_normal_momentum (:78-164)
 *   SWH_BC_FLATHER                             flather_boundary (:502-720)
 *   SWH_BC_TRANSMISSIVE                        transmissive_boundary (:23-75; elevation too)
 *   SWH_BC_STAGE_RADIATION_MOMENTUM            boundary_stage_radiation_momentum (:166-318)
 *   SWH_BC_STAGE_TRANSMISSIVE_MOMENTUM         boundary_stage_transmissive_momentum (:321-405)
 * use_stage_function: the boundary_function is the stage msl + amp sin(2 pi t / period)
 * with zero momenta and the cell's elevation (required by the first and last kinds);
 * else there is none (Flather: the sea at rest; radiation: msl_linear). */
#define SWH_BC_STAGE_TRANSMISSIVE_NORMAL_MOMENTUM 0
#define SWH_BC_FLATHER 1
#define SWH_BC_TRANSMISSIVE 2
#define SWH_BC_STAGE_RADIATION_MOMENTUM 3
#define SWH_BC_STAGE_TRANSMISSIVE_MOMENTUM 4
int swh_set_boundary(swh_domain *h, int32_t kind, const int32_t sides[4], int32_t use_stage_function, double amp,
                     double period, double msl);

/* domain%evolve_one_step(timestep) (domain_mod.f90:1679-1813) for 'linear',
 * 'leapfrog_linear_plus_nonlinear_friction', 'rk2', 'euler', 'midpoint' and 'rk2n'
 * (domain_mod_timestepping_alternatives_include.f90:8-336). timestep <= 0 means
 * "not present": only valid for the finite-volume methods, which then take the CFL
 * step computed on the device. nsteps > 1 repeats the call without host
 * synchronisation. */
int swh_evolve_one_step(swh_domain *h, double timestep);
int swh_evolve_steps(swh_domain *h, double timestep, int32_t nsteps);

/* Forcing patch (forcing_patch_type with apply_forcing_patch, forcing_mod.f90:24-56,
 * 158-239): work(i0:i1, j0:j1, k0:k1) in Fortran order (1-based, variables 1 = STG ..
 * 4 = ELV) is added to U over [start_time, end_time] at a constant rate, each update
 * adding the fraction of the interval its step covers (add_forcing_work_to_U_over_time,
 * :80-156): after the stage half-step for the leapfrog methods (stage and elevation
 * forcing; UH/VH work must be zero), after each update_U for the finite-volume methods
 * (stage, UH, VH; elevation work must be zero). One patch per domain; a new call
 * replaces it. */
int swh_set_forcing_patch(swh_domain *h, double start_time, double end_time, int32_t i0, int32_t i1, int32_t j0,
                          int32_t j1, int32_t k0, int32_t k1, const double *work);
int swh_clear_forcing_patch(swh_domain *h);

int swh_get_scalars(swh_domain *h, swh_scalars *out);

/* volume_interior (domain_mod.f90:1815-1849), FP64 compensated, deterministic. */
int swh_volume_interior(swh_domain *h, double *out);

/* compute_domain_statistics_new (domain_mod.f90:683-849) on the device: over the cells
 * more than exterior_cells_width from the edge, the maximum and minimum stage and speed
 * of wet cells and the potential and kinetic energy / rho (FP64, each row summed in
 * column order and the rows in row order, as the reference's work arrays are). */
typedef struct swh_domain_stats {
    double max_stage, max_speed, min_stage, min_speed;
    double energy_potential_on_rho, energy_kinetic_on_rho, energy_total_on_rho;
} swh_domain_stats;
int swh_domain_statistics(swh_domain *h, swh_domain_stats *out);

/* swh_domain_statistics without any device allocation. No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code under HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc). Same inputs,
 * outputs and kernels; the row sums live in a static device array of the library
 * (786 KB per GPU, held until the process ends) instead of a buffer the handle
 * allocates on first use. Domains of more than 16384 rows are rejected (SWH_EINVAL).
 * Every call on one GPU shares that GPU's array: calls that overlap on one GPU (on
 * different streams or host threads) race and must be serialised by the caller. */
int swh_domain_statistics_noalloc(swh_domain *h, swh_domain_stats *out);

/* The scan of check_multidomain_stability (multidomain_mod.f90:348-455) for one domain:
 * *n_bad = number of interior values of stage, uh, vh, elevation that are NaN or beyond
 * +-HUGE(1.0)/2. The reference stops the run when it is non-zero; the caller decides. */
int swh_domain_check_stability(swh_domain *h, int64_t *n_bad);

/* stationary_timestep_max (domain_mod.f90:1964-2022) on the current state. */
int swh_stationary_timestep_max(swh_domain *h, double *out);

/* One max_U plane (update_max_quantities, domain_mod.f90:2243-2404); which = one
 * SWH_MAXQ_* bit; out is (nx, ny) FP64. */
int swh_download_max_U(swh_domain *h, int32_t which, double *out);

/* Point gauges (point_gauge_mod.f90:208-294): n cells given as 1-based (i, j) pairs.
 * swh_gauges_sample records [STG, UH, VH] of every gauge at the current time into a
 * device time-series buffer; swh_gauges_download returns
 * out[(sample * 3 + var) * n + gauge] and the sample times. */
int swh_gauges_set(swh_domain *h, int32_t n, const int32_t *ij, int32_t max_samples);
int swh_gauges_sample(swh_domain *h);
int swh_gauges_download(swh_domain *h, double *values, double *times, int32_t *nsamples);

int swh_synchronize(swh_domain *h);

/* Dispersion (linear_dispersive_solver_mod.f90, the default form without
 * DISPERSIVE_PEREGRINE_IN_FLUX_FORM; staggered form for the leapfrog methods, cell-centred
 * for 'midpoint', 'rk2', 'euler'), at the domain procedures that
 * multidomain_type's dispersive loop calls (multidomain_mod.f90:1245-1410):
 *   swh_dispersion_enable       domain%use_dispersion = .true. + ds%setup (:144-179), with
 *                               dlon, dlat = domain%dx(1:2), ds%td1, ds%td2 and
 *                               ds%tridiagonal_inner_iter;
 *   swh_dispersion_store_last_U domain%copy_U_to_dispersive_last_timestep (domain_mod.f90:4305);
 *   swh_dispersion_append_extrapolation
 *                               ds%qet%append_value_and_time_for_quadratic_extrapolation(
 *                               domain%time, domain%U(:,:,UH:VH)) (quadratic_extrapolation_mod.f90:52);
 *   swh_dispersion_solve        domain%solve_dispersive_terms (domain_mod.f90:4319-4343).
 * Once enabled, 'midpoint' also solves in its predictor (include.f90:286-297) and update_U
 * leaves the edge cells (domain_mod_update_U_DE1_inner_include.f90:19-31). Not 'rk2n'. */
int swh_dispersion_enable(swh_domain *h, double dlon, double dlat, double td1, double td2, int32_t niter);
int swh_dispersion_store_last_U(swh_domain *h);
int swh_dispersion_append_extrapolation(swh_domain *h);
int swh_dispersion_solve(swh_domain *h, int32_t rhs_is_up_to_date, int32_t estimate_solution_forward_in_time,
                         double forward_time);

/* Output snapshot (the D->H staging behind write_to_output_files / write_grids,
 * domain_mod.f90:2188-2218, netcdf_util_mod.f90:437-506). swh_snapshot packs the current
 * state into a device buffer as FP64 U(nx, ny, 4) and its time, in stream order, and
 * returns at once: the caller keeps stepping. swh_snapshot_download waits for that
 * snapshot only (not for the steps after it) and copies it to U (4 * nx * ny doubles)
 * and *time on a separate copy stream; pinned host memory makes the copy a DMA that
 * overlaps the solver. One buffer per domain: a new snapshot replaces the last. */
int swh_snapshot(swh_domain *h);
int swh_snapshot_download(swh_domain *h, double *U, double *time);

/* Forward run with gauges: samples the gauges set by swh_gauges_set now and after
 * every 'every' steps of nsteps (nsteps/every + 1 samples in all). */
int swh_evolve_record_gauges(swh_domain *h, double timestep, int32_t nsteps, int32_t every);

/* Forward run with gauges on a schedule: samples the gauges at the listed steps
 * (strictly increasing; 0 = now), stepping between them with swh_evolve_steps, all in
 * stream order with no host synchronisation (the write-time rule of the reference's
 * driver gives the steps, see swh_batch_gauges_set_schedule). */
int swh_evolve_record_gauges_schedule(swh_domain *h, double timestep, int32_t nsamples, const int64_t *sample_steps);

/* Device checkpoints of the leapfrog state (no counterpart in the reference): a copy of
 * the current stage, UH, VH (and the FP32 Kahan term) and the time in device memory,
 * by slot (allocated on first use, freed with the domain); restore puts it back. */
int swh_checkpoint_save(swh_domain *h, int32_t slot);
int swh_checkpoint_restore(swh_domain *h, int32_t slot);

/* ---- Adjoint of the 'linear' (truely-linear) leapfrog step ----
 * No counterpart in the reference (which has no adjoint); it serves source inversion
 * over unit sources. The adjoint state (lS, lU, lV) lives on the device; a step
 * applies the exact transpose of one forward step of size dt (walls, no stage
 * boundary). swh_adjoint_run performs the backward sweep of a forward
 * swh_evolve_record_gauges(h, dt, nsteps, every): grad[(sample * ngauges) + g] is
 * d(loss)/d(stage at gauge g in that sample); afterwards lS is d(loss)/d(initial
 * stage) and lU, lV the gradients with respect to the initial UH, VH. */
int swh_adjoint_reset(swh_domain *h);
int swh_adjoint_upload(swh_domain *h, const double *lambda3);   /* 3 planes (nx, ny): lS, lU, lV */
int swh_adjoint_download(swh_domain *h, double *lambda3);
int swh_adjoint_step(swh_domain *h, double dt, int32_t nsteps);
int swh_adjoint_inject(swh_domain *h, const double *values);   /* lS[gauge g] += values[g] */
int swh_adjoint_run(swh_domain *h, double dt, int32_t nsteps, int32_t every, const double *grad);
/* The backward sweep of swh_evolve_record_gauges_schedule(h, dt, nsamples, sample_steps)
 * run to nsteps: grad[(q * ngauges) + g] = d(loss)/d(stage at gauge g in sample q). The
 * gradients are uploaded once and injected on the device. */
int swh_adjoint_run_schedule(swh_domain *h, double dt, int32_t nsteps, int32_t nsamples, const int64_t *sample_steps,
                             const double *grad);

/* Launch shape of the leapfrog step kernel (see swh_batch_set_launch). */
int swh_domain_set_launch(swh_domain *h, int32_t block_x, int32_t rows_per_block, int32_t variant);
/* Launch shape of the two-step leapfrog kernel of swh_evolve_steps: threads per
 * workgroup (64..256, a multiple of 64) and rows per strip (4..4096, or 0: chosen from
 * the kernel's occupancy, the default). Results do not depend on it. No counterpart in
 * the reference (a tuning knob of this backend). */
int swh_domain_set_launch2(swh_domain *h, int32_t block_x, int32_t rows_per_block);

/* ---- Batched unit sources on one shared grid ----
 * PTHA unit-source production runs one serial 'linear' model per source over a shared
 * grid and bathymetry (R/examples/austptha_template/SOURCE_ZONES/TEMPLATE/
 * TSUNAMI_UNIT_SOURCE/run_16.sh:6-36, template/model_namelist.in:7-12), each with its
 * own duration, writing gauge series (point_gauge_mod.f90:247-294). A batch advances all
 * of them together: one launch per step covers every active source; a source stops
 * (its state and gauges freeze) once it has taken its own number of steps. The desc
 * describes the shared grid; method must be SWH_METHOD_LINEAR or SWH_METHOD_LLPNF (the
 * latter with linear_solver_is_truely_linear set: the batch precomputes the friction work
 * from the elevation, so SWH_EINVAL otherwise) and max_U tracking is not available in a
 * batch. */
typedef struct swh_batch swh_batch;

int swh_batch_create(const swh_domain_desc *desc, int32_t nsources, swh_batch **out);
int swh_batch_destroy(swh_batch *b);
/* Shared elevation (nx, ny) and, for LLPNF, manning_squared (nx, ny). */
int swh_batch_upload_elevation(swh_batch *b, const double *elev);
int swh_batch_upload_manning_squared(swh_batch *b, const double *manning_squared);
/* Initial [STG, UH, VH] of source k: 3 planes (nx, ny) in Fortran order. */
int swh_batch_upload_source(swh_batch *b, int32_t k, const double *U3);
int swh_batch_download_source(swh_batch *b, int32_t k, double *U3);
/* Number of leapfrog steps each source takes (its duration / dt). */
int swh_batch_set_nsteps(swh_batch *b, const int32_t *nsteps_per_source);
/* Gauges sampled for every active source at the steps where step % every == 0
 * (step 0 = initial state), up to max_samples per source. */
int swh_batch_gauges_set(swh_batch *b, int32_t n, const int32_t *ij, int32_t every, int32_t max_samples);
/* Gauges sampled for every active source at the listed steps (strictly increasing,
 * step 0 = initial state); a source records the listed steps up to its own step count.
 * This carries the reference driver's write-time rule: generic_model.f90:380-397
 * (examples/generic_example, the unit-source production driver) writes gauges when
 * domain%time - last_write_time >= approximate_writeout_frequency, then advances
 * last_write_time by that frequency, so the samples fall on the first step on or after
 * each write time; the caller forms the steps from the FP64 time sequence of the
 * leapfrog (two half-steps per step). Must be called before the first step. */
int swh_batch_gauges_set_schedule(swh_batch *b, int32_t n, const int32_t *ij, int32_t nsamples,
                                  const int64_t *sample_steps);
/* Advance the batch nsteps leapfrog steps of size dt (the shared fixed step). */
int swh_batch_evolve(swh_batch *b, double dt, int32_t nsteps);
/* Gauge values of source k: values[(sample * 3 + var) * n + gauge]; nsamples out. */
int swh_batch_gauges_download(swh_batch *b, int32_t k, double *values, int32_t *nsamples);
/* Steps taken so far by the batch, and number of sources still active. */
int swh_batch_status(swh_batch *b, int64_t *step, int32_t *n_active);
/* Step-kernel launches so far (the autotuner's included): their number, the sources
 * they advanced summed over launches, and sources x fused steps summed over launches.
 * Each launch reads and writes every advanced source's state once, so differences of
 * these counters give a region's compulsory state traffic and its mean fused depth.
 * No counterpart in the reference (measurement). */
int swh_batch_launch_counts(swh_batch *b, int64_t *launches, int64_t *source_launches, int64_t *source_steps);
/* Launch shape of the leapfrog step kernel: threads per block (64..256, multiple of
 * 64), rows per thread strip, and variant bits: 1 = 16-byte vector lanes (4 floats /
 * 2 doubles per lane), 2 = keep the next row's loads in flight, 4 = non-temporal
 * state traffic, 8 = (batches) consecutive workgroups walk the sources of one tile,
 * 16 = (batches) two fused steps per launch wherever no gauge sample falls between
 * them (friction modes without per-step friction work only), 32 = (single grids)
 * one step per launch in swh_evolve_steps (by default pairs of steps run as one
 * two-step launch where no stage boundary, nesting or max_U update falls between),
 * 64 = (batches, two-step launches) XCD-aware tile order: each of the 8 XCDs walks a
 * contiguous range of tiles, columns fastest, then the sources of a group of
 * 2^((variant >> 7) & 7) sources, then strips, so neighbouring tiles, the sources of a
 * group (same elevation) and consecutive strips share its L2 (bit 2 also keeps the
 * next row's loads in flight in the two-step kernel), bits 1024..4096 = (batches, with
 * 16) up to 2 + ((variant >> 10) & 7) fused steps per launch (at most 6) wherever no
 * gauge sample falls inside them.
 * Results do not depend on the launch shape. */
int swh_batch_set_launch(swh_batch *b, int32_t block_x, int32_t rows_per_block, int32_t variant);
/* Launch-shape autotuner (no counterpart in the reference): times every one-step shape
 * (variant 0..15 x block_x {64,128,256} x rows {16,32,64}) and every fused shape (2..5
 * steps per launch x block_x {64,128,256} x rows {32,64,128,256} x tile order:
 * hardware, or XCD order with sources interleaved in groups of 16) over reps repeats
 * of the next launch on this batch's own grid and running sources, keeps the fastest
 * per cell-update, and returns it in best[3] = {block_x, rows_per_block, variant} with
 * its time per step.
 * The state and step count are left as they were. Needs the step counts and the
 * elevation set. */
int swh_batch_autotune(swh_batch *b, double dt, int32_t reps, int32_t *best, double *best_ms);
int swh_batch_synchronize(swh_batch *b);

/* Earthquake events from the batch's unit-source gauge series, on the device
 * (make_tsunami_event_from_unit_sources, R/examples/austptha_template/SOURCE_ZONES/
 * TEMPLATE/TSUNAMI_EVENTS/sum_tsunami_unit_sources.R:555-711). Event e sums the
 * sources event_src[event_ptr[e] .. event_ptr[e+1]-1] (0-based source indices):
 * with event_w == NULL each with weight 1, the sum then scaled by event_slip[e]
 * (uniform slip); otherwise with weights event_w (stochastic slip; event_slip NULL
 * means 1). The stage has msl subtracted before the sum and added after it.
 * summary SWH_SUPERPOSE_SERIES: out[e][sample][3][gauge] (max_samples samples per
 * event; samples past nvalid[e] are not meaningful); SWH_SUPERPOSE_MAX_STAGE:
 * out[e][gauge] = maximum over the event's valid samples of its stage.
 * nvalid[e] = samples recorded by every source of the event. kernel_ms (may be NULL)
 * returns the device time of the products. */
#define SWH_SUPERPOSE_SERIES 0
#define SWH_SUPERPOSE_MAX_STAGE 1
int swh_batch_superpose(swh_batch *b, int32_t n_events, const int32_t *event_ptr, const int32_t *event_src,
                        const double *event_w, const double *event_slip, double msl, int32_t summary, double *out,
                        int32_t *nvalid, double *kernel_ms);
/* swh_batch_evolve bracketed by HIP events recorded on the batch's stream; returns the
 * device time of the region in milliseconds (for throughput measurement). */
int swh_batch_evolve_timed(swh_batch *b, double dt, int32_t nsteps, double *elapsed_ms);

/* Measurement utility (no counterpart in the reference): a plain device copy
 * kernel of 'bytes' (multiple of 16) between two fresh buffers, repeated reps times
 * after one warm-up; returns the average time per copy. It is the known-size
 * reference for rocprofv3 FETCH_SIZE / WRITE_SIZE and the streaming-bandwidth
 * ceiling quoted next to the roofline. */
int swh_probe_copy(int64_t bytes, int32_t reps, double *ms_per_copy);
/* Measurement reference for VALU issue (no counterpart in the reference): blocks x 256
 * threads each run eight independent FMA chains for iters rounds in FP32 (unpacked) or
 * FP64; returns the mean device time of one launch over reps launches and the FMA
 * wave-instructions per launch. */
int swh_probe_valu(int32_t fp64, int32_t blocks, int32_t iters, int32_t reps, double *ms_per_launch,
                   double *wave_insts);

/* Check utility (no counterpart in the reference): evaluates the FP64 friction pow
 * pow(x[k], y[k]) and the boundary sine sin(t[k]) that the kernels use in place of
 * libm's (where the reference calls glibc: Manning depth^(-7/3),
 * domain_mod_update_U_DE1_inner_include.f90:74-75; the stage-function sine of the
 * boundary conditions) on the device and on the host, for n arguments. */
int swh_probe_math(int32_t n, const double *x, const double *y, const double *t, double *pow_dev, double *sin_dev,
                   double *pow_host, double *sin_host);

/* ---------------------------------------------------------------------------
 * Multidomain nesting on one GPU (multidomain_mod.f90, nested_grid_comms_mod.f90).
 *
 * The host runs the reference's own md%setup (it decides halos, priority domains
 * and the send/receive boxes); these calls hand the result to the device, which
 * then does the whole multidomain step: every domain's sub-steps with flux-integral
 * tracking, packing each send box directly into the receiving box's buffer (one
 * GPU: the point-to-point copy is a device write), unpacking, the staggered-grid
 * wet/dry fix-ups and flux correction. Domains step concurrently on their own
 * streams; the exchange runs on the multidomain's stream.
 * ------------------------------------------------------------------------- */
typedef struct swh_md swh_md;

/* One two_way_nesting_comms_type box (nested_grid_comms_mod.f90:69-219), as seen
 * from the domain that owns it. Indices are 1-based as in the reference. */
typedef struct swh_nest_box {
    int32_t inds[6];                   /* send_inds or recv_inds: iL, iU, jL, jU, kL, kU */
    int32_t neighbour_domain_index;    /* 1-based index of the other domain */
    int32_t neighbour_comms_index;     /* 1-based index of the matching box in the other domain's
                                          receive list (for a send box) or send list (for a receive box) */
    int32_t my_domain_is_finer;
    int32_t equal_cell_ratios;
    int32_t my_domain_staggered_grid;
    int32_t neighbour_domain_staggered_grid;
    int32_t use_wetdry_limiting;
    int32_t buffer_size;               /* size(send_buffer) / size(recv_buffer) */
    int32_t flux_start;                /* send_flux_start / recv_flux_start (1-based) */
    int32_t time_index;                /* send_timeind / recv_timeind (1-based) */
    int32_t flux_rows[4];              /* size(box_flux_integral(d)%x, 1), d = N, S, E, W */
    int32_t flux_nvar;                 /* size(box_flux_integral(d)%x, 2) */
    double cell_ratios[2];             /* my dx / neighbour dx */
} swh_nest_box;

/* Starts a multidomain over existing domain handles (all of one precision).
 * all_dx = md%all_dx_md(1:2, k, 1) for k = 1..n (2*n values), all_methods =
 * md%all_timestepping_methods_md(k, 1), trf = timestepping_refinement_factor
 * (multidomain_mod.f90:171-276, 1093-1212). */
int swh_md_create(int32_t n_domains, swh_domain *const *domains, const double *all_dx, const int32_t *all_methods,
                  const int32_t *timestepping_refinement_factor, swh_md **out);
int swh_md_destroy(swh_md *md);
/* domain%nesting%priority_domain_index and is_priority_domain_not_periodic (nx, ny),
 * domain k (1-based) (nested_grid_comms_mod.f90 domain_nesting_type). */
int swh_md_set_priority(swh_md *md, int32_t k, const int32_t *priority_domain_index,
                        const int32_t *is_priority_domain_not_periodic);
/* domain%flux_correct_dry_cells_outside_priority_domain (domain_mod.f90:283), default 1. */
int swh_md_set_flux_correct_dry_cells(swh_md *md, int32_t k, int32_t on);
/* Appends domain k's next send box (is_recv = 0) or receive box (is_recv = 1), in
 * the order of domain%nesting%send_comms / recv_comms. */
int swh_md_add_box(swh_md *md, int32_t k, int32_t is_recv, const swh_nest_box *box);
/* Pairs the boxes, builds the device tables and the flux-correction work
 * (domain_mod.f90:2766-3612), and allocates the buffers. */
int swh_md_finalize(swh_md *md);
/* md%make_initial_conditions_consistent (multidomain_mod.f90:1917-1931): one halo
 * exchange without flux correction; friction work of truely-linear
 * leapfrog+friction domains is then built from the exchanged elevation. */
int swh_md_make_initial_conditions_consistent(swh_md *md);
/* md%evolve_one_step(dt) (multidomain_mod.f90:1093-1212). */
int swh_md_evolve_one_step(swh_md *md, double dt);
int swh_md_evolve_steps(swh_md *md, double dt, int32_t nsteps);
/* Graph capture of the global step (no counterpart in the reference; the device form of
 * repeating md%evolve_one_step(dt), multidomain_mod.f90:1093-1212): records one global
 * step -- every domain's sub-steps and the exchange, kernels only -- as a HIP graph (two
 * when a domain takes an odd number of sub-steps, so that its buffers alternate), without
 * taking it. swh_md_graph_launch then replays nsteps global steps, each equal bit for bit
 * to swh_md_evolve_one_step(dt), and keeps the host-side times and step counts as eager
 * steps would; eager steps may be mixed in between. The kernels' time arguments come
 * from the device. SWH_ESTATE where a domain's step depends on host-side time or step
 * counts: 'leapfrog_nonlinear', dispersion, max_U every n > 1 steps or before the first
 * step (take one eager step first), and on leapfrog domains arrival-time / time-of-max
 * grids, forcing patches and boundaries with a stage function. Recapture after changing
 * any domain setting or dt. */
int swh_md_graph_capture(swh_md *md, double dt);
int swh_md_graph_launch(swh_md *md, int32_t nsteps);
int swh_md_graph_destroy(swh_md *md);
/* md%get_flow_volume per domain: volume_in_priority_domain (domain_mod.f90:1851-1888). */
int swh_md_volumes(swh_md *md, double *out);
/* print_multidomain's statistics (multidomain_mod.f90:2301-2447): per_domain[k] (n entries,
 * may be NULL) = domain k's compute_domain_statistics over its priority cells; global =
 * their extrema and the energy sums in domain order; *energy_is_finite = 0 when the
 * global total energy is NaN or infinite (the check the reference uses to stop unstable
 * runs). Any output pointer may be NULL. */
int swh_md_statistics(swh_md *md, swh_domain_stats *per_domain, swh_domain_stats *global,
                      int32_t *energy_is_finite);
/* check_multidomain_stability (multidomain_mod.f90:348-455) over every domain: *n_bad = the
 * number of interior values that are NaN or beyond +-HUGE(1.0)/2. */
int swh_md_check_stability(swh_md *md, int64_t *n_bad);
/* The receive buffer of domain k's receive box c (both 1-based), as filled by the
 * last exchange (diagnostics). */
int swh_md_download_recv_buffer(swh_md *md, int32_t k, int32_t c, double *out);
int swh_md_synchronize(swh_md *md);

#ifdef __cplusplus
}
#endif
#endif
