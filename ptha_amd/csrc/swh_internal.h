// Internal declarations of the gfx950 SWALS backend (not part of the C ABI).
#pragma once
#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <string>
#include <vector>

#include "swals_hip.h"

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
    double dt;                 // rk2: step taken (device-decided when adaptive)
    double max_dt;             // CFL step from the latest flux computation
    double bf_store[4];        // boundary_flux_store N, E, S, W
    double bf_evolve;          // boundary_flux_evolve_integral
    double bf_time;            // boundary_flux_time_integral
    unsigned long long max_dt_inv_bits;   // running max of 1/dt (non-negative doubles order as u64)
    unsigned long long neg_depth_count;   // negative_depth_fix_counter
    double pad[4];
};

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
    // fused max_U update (mask of SWH_MAXQ bits; 0 = none this step)
    int maxq_mask;
    T *mx;                     // N_MAXQ planes of pitch*ny
    T t_after, arrival_stage;
};

// Kernel launchers (defined in the .hip files)
template <typename T, bool KAHAN>
hipError_t launch_leapfrog(const LFArgs<T> &a, int coriolis, int friction_mode, int truely_linear,
                           int linear_friction, int nsrc, int block_x, int variant, hipStream_t s);

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
                         const double *area, double *row_work, double *out, hipStream_t s);

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
};

template <typename T>
hipError_t launch_adjoint_step(const AdjArgs<T> &a, hipStream_t s);
template <typename T>
hipError_t launch_adjoint_inject(T *lS, int n, const long long *idx, const double *values, hipStream_t s);

// ---- finite-volume (rk2) ----
template <typename T>
struct FVArgs {
    int nx, ny, pitch;           // state planes: pitch x ny
    int fpitch;                  // flux planes: fpitch x (ny+1), fpitch >= nx+1
    T *S, *UH, *VH;              // state being stepped (in place)
    const T *S0, *UH0, *VH0;     // state at the start of the Euler stage (read by the clip pass)
    const T *E;
    T *depth, *u, *v;            // per-cell work
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
};

template <typename T>
hipError_t launch_fv_euler(const FVArgs<T> &a, double fixed_dt, int first_stage, hipStream_t s);
template <typename T>
hipError_t launch_fv_average(T *S, T *UH, T *VH, const T *S0, const T *UH0, const T *VH0, int nx, int ny,
                             int pitch, hipStream_t s);
template <typename T>
hipError_t launch_fv_stage_bc(T *S, T *UH, T *VH, int nx, int ny, int pitch, int mask, double amp, double period,
                              double msl, const DevScalars *sc, hipStream_t s);
hipError_t launch_fv_begin(DevScalars *sc, hipStream_t s);
hipError_t launch_fv_finish(DevScalars *sc, hipStream_t s);

}  // namespace swh
