// Internal declarations of the gfx950 SWALS backend (not part of the C ABI).
#pragma once
#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <string>
#include <vector>

#include "swals_hip.h"
#include "swh_mathx.h"


// Friction depth power (Manning depth^(-7/3), Chezy depth^(-2)): FP64 through the
// accurate pow of swh_mathx.h, FP32 through the library's powf.
__device__ inline double pow_friction(double x, double p) { return swhm::pow_acc(x, p); }
__device__ inline float pow_friction(float x, float p) { return powf(x, p); }

namespace swh {

// Physical constants of the reference (global_mod.f90:37-44).
constexpr double kGravity = 9.8;
constexpr double kMinDepth = 1.0e-05;
constexpr double kMaxTimestep = 1.0e+20;

// Plane ids of the state
enum { STG = 0, UH = 1, VH = 2, ELV = 3 };
// max_U plane ids (bit k of the mask <-> plane k)
enum { MX_STAGE = 0, MX_SPEED, MX_FLUX, MX_ARRIVAL, MX_TOMS, MX_MIN_STAGE, N_MAXQ };

void set_error(const std::string &msg);
int hip_fail(hipError_t e, const char *what, const char *file, int line);

#define SWH_HIP(...)                                                                        \
    do {                                                                                    \
        hipError_t e_ = (__VA_ARGS__);                                                      \
        if (e_ != hipSuccess) return ::swh::hip_fail(e_, #__VA_ARGS__, __FILE__, __LINE__); \
    } while (0)

// Device-resident scalars of one domain (FP64 throughout).
struct DevScalars {
    double time;
    double dt;                 // finite volume: the step's base sub-step (device-decided when adaptive)
    double max_dt;             // CFL step from the latest flux computation
    double bf_store[4];        // boundary_flux_store N, E, S, W
    double bf_evolve;          // boundary_flux_evolve_integral
    double bf_time;            // boundary_flux_time_integral
    unsigned long long neg_depth_count;   // negative_depth_fix_counter
    double first_max_dt;       // CFL step of the step's first flux computation (the reported max_dt)
    double time0;              // time at the start of the step
    double stage_dt;           // step applied by the current update_U (dt_last_update)
    double evolve_dt;          // evolve_step_dt = time - time0
    unsigned int bc_done;      // blocks of the averaging stage boundary done (0 between launches)
    unsigned int bf_done;      // boundary-flux line workgroups done (0 between launches)
};

// Boundary-flux sums of a leapfrog step taken by extra workgroups of the step launch
// (swh_bflux.h): the caller asks for 'want' of them (0: none), the launcher places
// ey = bf_rows() rows of its grid's width ahead of the tiles. part: the workgroups'
// partial sums, room for 'cap' workgroups of 4 doubles; t_after: the end time of
// each step of the launch (NaN: from the device time).
struct BfLines {
    int want = 0, cap = 0, ey = 0, n_ext = 2;
    const double *dle = nullptr, *dbe = nullptr;
    DevScalars *sc = nullptr;
    double *part = nullptr;
    double dt = 0.0;
    double t_after[2] = {0.0, 0.0};
};
// Rows of line workgroups for a launch gx workgroups wide: enough for 'want' of them
// (at least 4: one per side), within 'cap'; -1 if they do not fit.
inline int bf_rows(const BfLines &b, int gx)
{
    if (!b.want) return 0;
    if (gx < 1 || b.want < 4) return -1;
    const int ey = (b.want + gx - 1) / gx;
    return ey * gx <= b.cap ? ey : -1;
}

// Bookkeeping of one finite-volume Euler stage (fv_scalar_kernel), after
// one_euler_step / one_midpoint_step / one_rk2n_step
// (domain_mod_timestepping_alternatives_include.f90:8-336):
//   kStageFirst      the step's first flux computation: dt = fixed step or the CFL step,
//                    and its CFL step is the one reported
//   kStageAccBf      boundary_flux_evolve_integral += stage_dt * sum(boundary_flux_store)
//   kStageFromTime0  time = time0 + stage_dt (midpoint's full step from the backup)
//   kStageScaleBfN   first scale boundary_flux_evolve_integral by (n-1)/n (rk2n)
//   kStageBegin      the step's first stage: evolve_one_step's reset of the step's flux
//                    integral and the start time (domain_mod.f90:1713-1715)
//   kStageFinish     the step's last stage of euler/midpoint: the step's bookkeeping
//                    (reported CFL step, evolve_step_dt, time integral) after it
enum : int {
    kStageFirst = 1, kStageAccBf = 2, kStageFromTime0 = 4, kStageScaleBfN = 8, kStageBegin = 16, kStageFinish = 32
};
// rk2n_n_value (domain_mod.f90:88-90)
constexpr int kRk2nN = 5;
// fv_finish_kernel modes
enum : int { kFinishEuler = 0, kFinishRk2 = 1, kFinishRk2n = 2 };

// Per-row step coefficients of the leapfrog scheme, built in FP64 on the host from
// the reference's metric arrays and converted to the storage type.
template <typename T>
struct LFRowCoef {
    const T *sa;   // dt / area(j)                              stage (include.f90:97)
    const T *cu;   // gravity*dt / area(j)                      UH (include.f90:230)
    const T *cv;   // gravity*dt / (0.5*(area(j)+area(j+1)))    VH (include.f90:235-236)
    const T *fu;   // dt*coriolis(j)*0.5                        include.f90:434
    const T *fv;   // dt*coriolis_bottom_edge(j+1)*0.5          include.f90:436-437
    const T *dle;  // distance_left_edge   (nx+1)
    const T *dbe;  // distance_bottom_edge (ny+1)
};

template <typename T>
struct LFArgs {
    int nx, ny, pitch;
    long long src_stride;      // elements between sources' planes (batched runs)
    const T *S, *C, *UH, *VH;  // inputs (C: Kahan compensation of S, FP32 mode only)
    const T *E;                // elevation (shared by all sources)
    T *So, *Co, *UHo, *VHo;    // outputs (ping-pong partners of the inputs)
    LFRowCoef<T> rc;
    const T *fwU, *fwV;        // friction_work (precomputed, truely-linear llpnf)
    const T *nsq;              // manning_squared (not-truely-linear llpnf)
    T msl, dt, dt_lfc, lin_fric_mult, fric_power;
    int bc_mask;               // bit 0 N, 1 E, 2 S, 3 W
    T bc_stage;
    int rows_per_block;
    int src_fast;              // batched runs: block x index walks the sources
    int src_base = 0;          // batched runs: first source of this launch
    // fused max_U update (mask of SWH_MAXQ bits; 0 = none this step)
    int maxq_mask;
    T *mx;                     // N_MAXQ planes of pitch*ny
    T t_after, arrival_stage;
    // forcing patch of this step (apply_forcing after the stage half-step,
    // forcing_mod.f90:80-156): stage += f_frac * fS, elevation += f_frac * fE
    // (planes zero outside the patch; null: no such forcing this step)
    const T *fS = nullptr, *fE = nullptr;
    T f_frac = T(0);
    // 1: the inputs S/C already hold the new stage (formed by lf_stage_kernel and the
    // boundary); the step passes it through and writes only UH/VH
    int stage_given = 0;
    // two-step kernel: 1 = XCD-aware block order (each XCD walks a contiguous range of
    // tiles, columns fastest, then the sources of a group of lf2_group, then strips,
    // then groups), 0 = the hardware's order
    int lf2_order = 0;
    int lf2_group = 1;
    // fused kernel: a gauge sample taken at the level where it falls (batched runs, at
    // most 64 gauges): sample g_sample is written after the launch's g_level-th step;
    // g_ij[2g], g_ij[2g+1] = the gauge's cell (i, j), 0-based; g_val[slot][sample][3][g_n]
    int g_n = 0, g_max = 0;
    int g_level = 0, g_sample = 0;
    const int *g_ij = nullptr;
    double *g_val = nullptr;
    // single grid: the step's boundary-flux sums in extra workgroups of this launch
    BfLines bf;
};

// Kernel launchers (defined in the .hip files)
template <typename T, bool KAHAN>
hipError_t launch_leapfrog(const LFArgs<T> &a, int coriolis, int friction_mode, int truely_linear,
                           int linear_friction, int nsrc, int block_x, int variant, hipStream_t s);

// Two fused leapfrog steps per launch (temporal blocking) for batched runs without
// boundary stage, max_U or flux tracking; friction modes 0 and 1 only.
template <typename T, bool KAHAN>
hipError_t launch_leapfrog2(const LFArgs<T> &a, int coriolis, int friction_mode, int truely_linear,
                            int linear_friction, int nsrc, int block_x, hipStream_t s, int prefetch = 0);
// ns (3 or 4) fused steps per launch (same operations; the state is read and written
// once per ns steps).
template <typename T, bool KAHAN>
hipError_t launch_leapfrogn(const LFArgs<T> &a, int coriolis, int friction_mode, int truely_linear,
                            int linear_friction, int nsrc, int block_x, hipStream_t s, int prefetch, int ns);

// Plain streaming copy of 'bytes' (a multiple of 16) for measurement.
hipError_t launch_probe_copy(const void *src, void *dst, long long bytes, hipStream_t s);
hipError_t launch_probe_valu(int fp64, int blocks, int iters, void *out, hipStream_t s);
// Host transfers: a dense FP64 plane (nx*ny, i fastest) into a pitched plane of T
// (padding columns zeroed; with comp, the FP32 Kahan term that makes dst - comp the FP64
// value; src null: zeros), and back (dst = src - comp).
template <typename T>
hipError_t launch_plane_from_dense(const double *src, T *dst, T *comp, int nx, int ny, int pitch, hipStream_t s);
template <typename T>
hipError_t launch_plane_to_dense(const T *src, const T *comp, double *dst, int nx, int ny, int pitch, hipStream_t s);
hipError_t launch_probe_math(const double *x, const double *y, const double *t, double *pw, double *sn, int n,
                             hipStream_t s);

// Events from unit-source gauge series (swh_superpose.hip): out = W * (X - msl on the
// stage) scaled and shifted per event.
struct SupArgs {
    const double *X;         // gauge series [slot][sample][3][ng]
    long long slot_stride;   // gmax * 3 * ng
    int nsrc, ng, gmax;
    const double *W;         // [Epad][Kpad] event weights by slot, zero padded
    int Kpad;                // nsrc rounded up to a multiple of 4
    int E;                   // events (W, scale and nvalid hold Epad rows)
    const double *scale;     // per event: uniform slip (or 1)
    const int *nvalid;       // per event: samples common to all its sources
    double msl;
    int s_chunk, s_max;      // maximum-stage summary: samples per block, samples to scan
    double *out;             // mode 0: [E][gmax * 3 * ng]; mode 1: [E][ng]
    // set when some slot's series holds a non-finite value: [Epad][Kpad] 1 where the
    // event has such a slot (non-finite values then enter the product as 0, and the
    // outputs that an event's non-finite members reach are NaN)
    const double *M = nullptr;
};
// Forcing patch on a finite-volume state after update_U (forcing_mod.f90:80-156):
// the fraction of [start, end] that the stage (time - stage_dt, time] covers, read
// from the device scalars, times the work planes (S, UH, VH; null: none) added to
// the state over the patch cells (0-based, inclusive).
template <typename T>
hipError_t launch_fv_forcing(T *S, T *UH, T *VH, const T *fS, const T *fUH, const T *fVH, int pitch, int i0, int i1,
                             int j0, int j1, double start, double end, const DevScalars *sc, hipStream_t s,
                             T *C = nullptr);
// elevation += frac * fE over the patch (the leapfrog step's elevation forcing)
// 'leapfrog_nonlinear' momentum half-step (swh_nonlinear.hip). Planes of pitch * ny.
template <typename T>
struct NLArgs {
    int nx, ny, pitch;
    const T *S0;              // stage before the stage half-step (time-lagged depth)
    const T *S, *E;           // stage and elevation after the boundary and forcing
    const T *UH, *VH;         // momenta before the update
    T *UHo, *VHo;             // updated momenta
    const T *nsq;             // manning_squared
    T *DL, *UUH, *UVH, *VUH, *VVH, *fEWu, *fEWv, *fNSu, *fNSv;  // work planes
    const double *dbe, *dle, *area, *cor, *corbe, *tanlat;     // FP64 metrics
    const T *ica, *ica_vh;    // per row: (T)(dt / area(j)), (T)(dt / (0.5 (area(j) + area(j+1))))
    T dt, lfc, fric_power, beta, froude_limit;
    int coriolis, spherical;
};
// depth before the stage half-step (launch first)
template <typename T>
hipError_t launch_nonlinear_depth0(const NLArgs<T> &a, hipStream_t s);
// time-lagged depth, advection work, fluxes, momentum update (after the stage half-step,
// boundary and forcing). fused: fluxes straight from the stage before and after the
// stage half-step (valid while the elevation is unchanged in the step; then
// launch_nonlinear_depth0 is not needed)
template <typename T>
hipError_t launch_nonlinear_momentum(const NLArgs<T> &a, bool fused, hipStream_t s);

// Dispersive solve on the staggered grid (swh_dispersion.hip). Planes of pitch * ny
// except W (4 * nx * ny work for the tridiagonal systems).
template <typename T>
struct DispArgs {
    int nx, ny, pitch;
    T *UH, *VH;                  // current momenta (updated)
    const T *E;                  // current elevation
    const T *LU, *LV, *LE;       // UH, VH, elevation stored before the shallow-water step
    T *RU, *RV;                  // right-hand side
    const T *Y0, *Y1, *Y2;       // extrapolation levels, newest first (UH plane, then VH plane)
    const double *qt;            // their times, oldest first (device)
    T *W;
    const double *dbe, *dle;
    T msl, dlon, dlat, td1, td2;
    int cell;                    // 1: cell-centred (finite-volume) form, 0: staggered
    const DevScalars *sc;        // forward time of the extrapolation from sc->time when set
    double t_fwd;                // else this forward time
};
// make_rhs: explicit part and right-hand side; extrapolate: U(UH:VH) extrapolated in time
// from Y0..Y2; then niter tridiagonal iterations (rows of UH, then columns of VH)
template <typename T>
hipError_t launch_dispersion(const DispArgs<T> &a, int make_rhs, int extrapolate, int niter, hipStream_t s);
// shift the level times down one and record the new level's time (sc->time or t_host)
hipError_t launch_dispersion_qt_append(double *qt, const DevScalars *sc, double t_host, hipStream_t s);

// Dense FP64 copy of the state [S (- C), UH, VH, E] as out[(var * ny + j) * nx + i]
// (the bytes of Fortran U(nx, ny, 4)), with out[4 * nx * ny] = the time: sc->time when
// sc is given, else t_host.
template <typename T>
hipError_t launch_snapshot(const T *S, const T *C, const T *UH, const T *VH, const T *E, int nx, int ny, int pitch,
                           double *out, const DevScalars *sc, double t_host, hipStream_t s);

// New stage only (the stage half-step of the fused step, same per-cell operations),
// S/C -> So/Co over the whole grid, border cells passed through
template <typename T, bool KAHAN>
hipError_t launch_lf_stage(const LFArgs<T> &a, hipStream_t s);
// Flather boundary (boundary_mod.f90:502-720) on the edges flagged in mask (1 N, 2 E,
// 4 S, 8 W): one block, west/east columns then south/north rows. Outside state: the
// sea at rest at max(msl_linear, elevation), or (use_fun) the stage
// fun_msl + amp*sin(2*pi*t/period) with zero momenta, t from the device scalars when
// sc is given, else fun_stage is the host-computed stage. C (Kahan compensation, may
// be null) is zeroed on the cells written.
template <typename T>
hipError_t launch_flather(T *S, T *C, T *UH, T *VH, const T *E, int nx, int ny, int pitch, int mask, double msl_linear,
                          int use_fun, double amp, double period, double fun_msl, double fun_stage,
                          const DevScalars *sc, hipStream_t s);
// transmissive_boundary (kind 2), boundary_stage_radiation_momentum (3),
// boundary_stage_transmissive_momentum (4) on the edges in mask (1 N, 2 E, 4 S, 8 W);
// stag: staggered grid (kinds 3, 4 then set stage and elevation only)
template <typename T>
hipError_t launch_edge_bc(T *S, T *C, T *UH, T *VH, T *E, int nx, int ny, int pitch, int mask, int kind, int use_fun,
                          double amp, double period, double fun_msl, double msl_linear, double fun_stage, int stag,
                          const DevScalars *sc, hipStream_t s);
// Kahan-compensated stage += frac * fS over the patch (a leapfrog step's stage forcing
// applied outside the fused step)
template <typename T>
hipError_t launch_add_scaled_kahan(T *S, T *C, const T *fS, int pitch, int i0, int i1, int j0, int j1, T frac,
                                   hipStream_t s);
template <typename T>
hipError_t launch_add_scaled(T *E, const T *fE, int pitch, int i0, int i1, int j0, int j1, T frac, hipStream_t s);

// rows of the event tile (W, scale and nvalid must hold a multiple of it)
int superpose_event_tile();
hipError_t launch_superpose(const SupArgs &a, int mode, hipStream_t s);
// flags[slot] = 1 where the slot's series (slot_stride values each) holds a non-finite
// value; flags must start at 0.
hipError_t launch_nonfinite_slots(const double *X, int nsrc, long long slot_stride, int *flags, hipStream_t s);

// Gauge values after one step, from the current state (batched runs, or a single grid
// whose sample falls between the two steps of a two-step launch); times: null, or
// times[sample] = t.
template <typename T, bool KAHAN>
hipError_t launch_gauge_step1(const LFArgs<T> &a, int coriolis, int friction_mode, int truely_linear,
                              int linear_friction, int nsrc, int n, const long long *idx, double *values, int gmax,
                              int sample, double *times, double t, hipStream_t s);

template <typename T>
hipError_t launch_bflux_leapfrog(const T *UH, const T *VH, int nx, int ny, int pitch, int n_ext,
                                 const double *dle, const double *dbe, double dt, double time_after,
                                 DevScalars *sc, hipStream_t s);

template <typename T>
hipError_t launch_maxq(const T *S, const T *C, const T *UH, const T *VH, const T *E, int nx, int ny, int pitch,
                       int staggered, int truely_linear, T msl, int mask, T *mx, const DevScalars *tsrc,
                       T t, T arrival_stage, hipStream_t s);

template <typename T>
hipError_t launch_volume(const T *S, const T *C, const T *E, int nx, int ny, int pitch, int n_ext,
                         const double *area, double *row_work, double *out, hipStream_t s,
                         const int *mask = nullptr);

template <typename T>
hipError_t launch_gauges(const T *S, const T *C, const T *UH, const T *VH, int n, const long long *idx,
                         double *values, double *times, int sample, const DevScalars *tsrc, double t,
                         hipStream_t s);

template <typename T>
hipError_t launch_timestep_max(const T *S, const T *E, int nx, int ny, int pitch, int linear, double msl,
                               const double *dbe, const double *dle, double *out, hipStream_t s);

// ---- adjoint of the truely-linear leapfrog step ----
template <typename T>
struct AdjArgs {
    int nx, ny, pitch;
    const T *lS, *lU, *lV;   // adjoints of the step's outputs
    T *oS, *oU, *oV;         // adjoints of the step's inputs
    const T *E;
    LFRowCoef<T> rc;
    T msl, lin_fric_mult;
    int coriolis;
    int rows = 0;            // rows per strip of the fused kernel (0: from its occupancy)
};

template <typename T>
hipError_t launch_adjoint_step(const AdjArgs<T> &a, hipStream_t s);
template <typename T>
hipError_t launch_adjoint_inject(T *lS, int n, const long long *idx, const double *values, hipStream_t s);

// ---- finite-volume (rk2) ----
struct NestStrip;

template <typename T>
struct FVArgs {
    int nx, ny, pitch;           // state planes: pitch x ny
    int fpitch;                  // flux planes: fpitch x (ny+1), fpitch >= nx+1
    T *S, *UH, *VH;              // state being stepped (in place)
    const T *S0, *UH0, *VH0;     // state the fluxes are computed from (clipped where it is read)
    const T *E;
    const T *Sq, *UHq, *VHq;     // state update_U adds to; null: the clipped S0/UH0/VH0
                                 // (midpoint's second update starts from the backup)
    double *bmax;                // per-block maximum of max_speed / (dx * cfl / 2) of the flux pass
    T *fNS, *fEW;                // 3 planes each: mass, UH, VH fluxes (pre-multiplied by edge length)
    double *esU, *esV;           // cell-local explicit sources
    double *esJ;                 // NS edge contribution to the cell below (explicit_source_VH_j_minus_1)
    double *esE;                 // EW edge contribution to the cell on the left (explicit_source_im1_work)
    const T *nsq;                // manning_squared
    const double *dbe, *dle, *area, *cor, *tanlat;
    double theta, half_cfl, lfc;
    double fric_power;
    int spherical, coriolis, n_ext;
    DevScalars *sc;
    int skip_edges = 0;          // dispersive domains: update_U leaves the edge cells
    // rk2's second update with the average folded in: the stored state is
    // 0.5 * (updated + backup) (null: plain update)
    const T *avgS = nullptr, *avgUH = nullptr, *avgVH = nullptr;
    // FP32 state: Kahan compensation planes of S0, S, Sq and avgS (the represented stage
    // is S - C); null in FP64
    const T *C0 = nullptr;
    T *Cw = nullptr;
    const T *Cq = nullptr, *avgC = nullptr;
    // with avgS and a stage boundary after the update: the pre-average UH of columns 1
    // and nx-2 ([2][ny]) then VH of rows 1 and ny-2 ([2][nx]), which the boundary reads
    T *ring = nullptr;
    // flux pass: rows per wavefront of the column-marching kernel (4..64)
    int flux_rows = 8;
    // nesting flux integrals of this stage (x += nest_dt * flux on each strip cell),
    // done by extra blocks of the update launch: null or nest_dt == 0: none
    const NestStrip *nest_strips = nullptr;
    int n_nest_strips = 0, nest_chunks = 0;
    double nest_dt = 0.0;
    double *nest_int = nullptr;
    // set by launch_fv_euler when the stage's step is known before the stage's flux pass
    // (a fixed dt, or any stage after a step's first): the scalar pass runs as workgroup
    // row 0 of the update launch, beside the update, which takes the stage's step from
    // stage_dt (stage_dt_dev: from the step the first stage stored, DevScalars::dt,
    // times s_dt_mult), with the pass's arguments below
    int scalar_row = 0, stage_dt_dev = 0;
    double stage_dt = 0.0, s_fixed_dt = 0.0, s_dt_mult = 1.0;
    int s_mode = 0, s_finish = 0, s_nbmax = 0;
};

template <typename T>
hipError_t launch_fv_euler(const FVArgs<T> &a, double fixed_dt, double dt_mult, int mode, hipStream_t s);
// rk2n: part 0: B0 = (B0 - U) / n;  part 1: U = U + B0 and the step's bookkeeping
// (kFinishRk2n)
template <typename T>
hipError_t launch_fv_rk2n_combine(T *S, T *UH, T *VH, T *S0, T *UH0, T *VH0, int nx, int ny, int pitch, int part,
                                  DevScalars *sc, hipStream_t s, T *C = nullptr, T *C0 = nullptr);
// rk2 average with backup_U; the step's bookkeeping (kFinishRk2) rides along
template <typename T>
hipError_t launch_fv_average(T *S, T *UH, T *VH, const T *S0, const T *UH0, const T *VH0, int nx, int ny,
                             int pitch, DevScalars *sc, hipStream_t s, T *C = nullptr, const T *C0 = nullptr);
template <typename T>
hipError_t launch_fv_stage_bc(T *S, T *UH, T *VH, int nx, int ny, int pitch, int mask, double amp, double period,
                              double msl, const DevScalars *sc, hipStream_t s, T *C = nullptr);
// The stage boundary of rk2's second update when the average with backup_U (S0, UH0,
// VH0, C0) is folded into that update: applies the boundary rules in the order of
// launch_fv_stage_bc to the pre-average values (interior neighbours from the ring the
// update kept) and stores the boundary cells' averages, as the separate average pass
// would leave them.
template <typename T>
hipError_t launch_fv_stage_bc_avg(T *S, T *UH, T *VH, const T *S0, const T *UH0, const T *VH0, const T *ring,
                                  int nx, int ny, int pitch, int mask, double amp, double period, double msl,
                                  DevScalars *sc, hipStream_t s, T *C, const T *C0);
// number of blocks of the flux pass (size of FVArgs::bmax)
int fv_flux_blocks(int nx, int ny);


// ---------------------------------------------------------------------------
// Multidomain nesting (one GPU). All box indices are 1-based and inclusive, as in
// the reference's send_inds / recv_inds (nested_grid_comms_mod.f90:69-219); the
// kernels convert when they address the 0-based pitched planes.
// ---------------------------------------------------------------------------
constexpr int kMaxNestDomains = 16;
constexpr int kNestGradients = 4;   // number_of_gradients_coarse_to_fine
constexpr double kMissingTime = -1.7976931348623157e308;   // -HUGE(1.0_dp)

// Pointers to one domain's current state (the leapfrog ping-pong buffer in use).
template <typename T>
struct NestDomPtr {
    T *S, *C, *UH, *VH;    // C: Kahan compensation of S (represented stage S - C), or null
    const T *E;
    int nx, ny, pitch;
};

template <typename T>
struct NestDoms {
    NestDomPtr<T> d[kMaxNestDomains];
};

// One edge strip of a send or receive box, over which dt * flux * edge is
// accumulated (nested_grid_comms_mod.f90:2308-2461).
struct NestStrip {
    int kind;      // 0: a row of north-south fluxes, 1: a column of east-west fluxes
    int line;      // 0-based row (kind 0) or column (kind 1) in the flux planes; -1: out of range
    int lo, n;     // 0-based start and length along the strip
    int out;       // offset of x(1, 1) of this strip in the domain's integral array; x(r, k) at out + k*n + r
    double edge;   // distance_bottom_edge / distance_left_edge factor (1 when fluxes are pre-multiplied)
};

// Packing of one send box into the receiver's buffer (process_data_to_send_2022Feb,
// nested_grid_comms_mod.f90:1387-2046).
struct NestPack {
    int dom;                  // sending domain (0-based)
    int mode;                 // 0: equal cells, same grid type; 1: fine-to-coarse; 2: coarse-to-fine
    int iL, iU, jL, jU, kL, kU;
    int iR, jR, kR;           // receive-grid extent of the box
    int inv_r[2], r_ip[2];    // nint(1/cell_ratios), nint(cell_ratios)
    double cr[2], inv_cr[2];  // cell_ratios, 1/cell_ratios
    int my_stag, nbr_stag, wetdry;
    int flux_start;           // 0-based offset of the flux section in the buffer
    int time_index;           // 0-based offset of the time
    int strip_out[4], strip_n[4], nvar;   // the box's N, S, E, W integral strips
    int n_data, n_flux;       // counts of data / flux entries written
    long long buf;            // offset of the buffer in the buffer arena
    int block0;               // first block of this box in the batched launch
};

// Unpacking of one receive box (process_received_data, nested_grid_comms_mod.f90:2048-2249).
struct NestUnpack {
    int dom;                 // receiving domain
    int fine;                // my_domain_is_finer: interpolate coarse values + gradients
    int iL, iU, jL, jU, kL, kU;
    int inv_r[2];
    double cr[2], mid[2];
    int niC, njC, nC, n_data;
    int strip_out[4], strip_n[4], nvar;   // the box's own integrals (recv side)
    int err_out;             // offset of the flux error in the error arena (same layout as the strips)
    long long buf;
    int block0;
    int stag_fix;            // 0: none, 1: elevation rule (truely linear), 2: depth rule, 3: nonlinear rule
    double elev_lim;         // msl_linear - minimum_allowed_depth
};

// One flux-correction update (domain_mod.f90:2803-3226): cells p0..p1 along
// 'line', all variables var1..varN, U += sgn * err(k)/area.
struct NestFix {
    int is_ew;        // 0: column ni (cells (ni, p)), 1: row mi (cells (p, mi)); 0-based line
    int line, p0, p1;
    int var1, varN;   // 0-based inclusive; varN < var1: nothing
    double sgn;       // +1 or -1 with the N/S/E/W sign rule folded in
    double area;
    int err;          // offset of err(row, var 0) in the error arena; var k at err + k*stride
    int stride;
};

// The flux-correction work of one domain: its updates 'fix', one per cell (the
// launches take them group by group).
struct NestFixDom {
    const NestFix *fix;
    int dom;
    const double *err;      // the domain's flux-error arena
    const int *pnp;         // is_priority_domain_not_periodic, nx*ny, 0-based row-major
    int dry_all;            // flux_correct_dry_cells_outside_priority_domain
};

// Re-interpolation of a receive box of a finer domain after flux correction
// (domain_mod.f90:3228-3468).
struct NestReint {
    int dom;
    int n0, n1, m0, m1;     // 1-based box
    int inv_r[2];
    int ind[4];             // 1-based copy row/col for N, S, E, W
    int copy;               // offset of the copied strips in the copy arena: [side][var][pos]
    int len_ns, len_ew;     // n1-n0+1, m1-m0+1
    int ncx, ncy;           // coarse cells in x and y
    int block0;
};

template <typename T>
hipError_t launch_nest_zero(double *p, long long n, hipStream_t s);
template <typename T>
hipError_t launch_nest_accum(const NestStrip *strips, int n_strips, int max_n, const T *fNS, const T *fEW,
                             long long plane_ns, long long plane_ew, int pitch, int nvar, double dt, double *integ,
                             hipStream_t s);
template <typename T>
hipError_t launch_nest_pack(const NestDoms<T> &doms, const NestPack *packs, int n_packs, int n_blocks,
                            double *const *integ, double *arena, hipStream_t s);
template <typename T>
hipError_t launch_nest_unpack(const NestDoms<T> &doms, const NestUnpack *ups, int n_ups, int n_blocks,
                              double *const *integ, double *const *err, const double *arena, hipStream_t s);
template <typename T>
hipError_t launch_nest_stag_fix(const NestDoms<T> &doms, const NestUnpack *ups, int n_ups, int n_blocks,
                                hipStream_t s);
template <typename T>
hipError_t launch_nest_fix(const NestDoms<T> &doms, const NestFixDom *fds, const int2 *cells, int n, hipStream_t s);
template <typename T>
hipError_t launch_nest_reint(const NestDoms<T> &doms, const NestReint *rs, int n_rs, int max_len, int n_blocks,
                             double *copy, hipStream_t s);


}  // namespace swh
