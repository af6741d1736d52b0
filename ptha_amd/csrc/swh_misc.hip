// Small kernels around the step: boundary-flux sums, max_U update, volume,
// gauges and the stationary CFL step. All reductions are FP64, deterministic
// (fixed assignment of work to lanes and a fixed combine order) and carry a
// two-sum compensation term, so results do not depend on scheduling.
#include <algorithm>

#include "swh_bflux.h"

namespace swh {

// Reduce one DD per thread over a block of blockDim.x threads (power of two <= 256).
__device__ DD block_reduce_dd(DD v)
{
    __shared__ double sh_hi[256], sh_lo[256];
    const int t = threadIdx.x;
    sh_hi[t] = v.hi;
    sh_lo[t] = v.lo;
    __syncthreads();
    for (int w = blockDim.x / 2; w > 0; w >>= 1) {
        if (t < w) {
            DD a{sh_hi[t], sh_lo[t]}, b{sh_hi[t + w], sh_lo[t + w]};
            DD c = dd_add(a, b);
            sh_hi[t] = c.hi;
            sh_lo[t] = c.lo;
        }
        __syncthreads();
    }
    DD r{sh_hi[0], sh_lo[0]};
    __syncthreads();
    return r;
}

// ---- leapfrog boundary flux (domain_mod_linear_solver_include.f90:126-165) ----
// One block of 1024 threads: 4 groups of 256, group g sums side g (N, E, S, W) of the
// pre-momentum UH/VH, each thread a fixed stride of the line with its loads unrolled
// (so a few memory latencies cover the line), then a fixed-order compensated tree per
// group. Thread 0 forms boundary_flux_evolve_integral and adds it to
// boundary_flux_time_integral (domain_mod.f90:1803-1807).
constexpr int kBfThreads = 1024;

template <typename T>
__global__ __launch_bounds__(kBfThreads) void bflux_leapfrog_kernel(const T *__restrict__ UH,
                                                                   const T *__restrict__ VH, int nx, int ny,
                                                                   int pitch, int n_ext,
                                                                   const double *__restrict__ dle,
                                                                   const double *__restrict__ dbe, double dt,
                                                                   double time_after, DevScalars *sc)
{
    constexpr int Q = kBfThreads / 4;
    constexpr int B = 8;   // loads issued together before any is used
    __shared__ double part_hi[4][Q], part_lo[4][Q];
    // thread 0 reads the metric terms and the time integral now, so that their latency
    // overlaps the sums
    double m[4] = {0.0, 0.0, 0.0, 0.0}, bf_time = 0.0;
    if (threadIdx.x == 0) {
        m[0] = dbe[ny - n_ext]; m[1] = dle[nx - n_ext]; m[2] = dbe[n_ext]; m[3] = dle[n_ext];
        bf_time = sc->bf_time;
        // NaN: the step's end time from the device time (graph replays), with the
        // host's operations: (time + dt * 0.5) + 0.5 * dt
        if (time_after != time_after) time_after = (sc->time + dt * 0.5) + 0.5 * dt;
    }
    const int g = threadIdx.x / Q, lane = threadIdx.x % Q;
    // side g: N and S rows of VH, E and W columns of UH; each thread adds every Q'th
    // entry in index order, B loads at a time
    const bool ns = (g == 0 || g == 2);
    const T *src = ns ? VH : UH;
    const long long base = ns ? (long long)((g == 0) ? ny - n_ext - 1 : n_ext - 1) * pitch
                              : (long long)((g == 1) ? nx - n_ext - 1 : n_ext - 1);
    const long long step = ns ? 1 : pitch;
    const int last = ns ? nx - n_ext - 1 : ny - n_ext - 1;
    DD acc{0.0, 0.0};
    for (int i0 = n_ext + lane; i0 <= last; i0 += B * Q) {
        double v[B];
#pragma unroll
        for (int u = 0; u < B; u++) {
            const int i = i0 + u * Q;
            v[u] = i <= last ? (double)src[base + i * step] : 0.0;
        }
#pragma unroll
        for (int u = 0; u < B; u++)
            if (i0 + u * Q <= last) acc = dd_add(acc, v[u]);
    }
    part_hi[g][lane] = acc.hi;
    part_lo[g][lane] = acc.lo;
    __syncthreads();
    for (int w = Q / 2; w > 0; w >>= 1) {
        if (lane < w) {
            const DD c = dd_add(DD{part_hi[g][lane], part_lo[g][lane]}, DD{part_hi[g][lane + w], part_lo[g][lane + w]});
            part_hi[g][lane] = c.hi;
            part_lo[g][lane] = c.lo;
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        double store[4];
        for (int q = 0; q < 4; q++) store[q] = part_hi[q][0] + part_lo[q][0];
        const double b0 = store[0] * m[0], b1 = store[1] * m[1], b2 = -store[2] * m[2], b3 = -store[3] * m[3];
        sc->bf_store[0] = b0;
        sc->bf_store[1] = b1;
        sc->bf_store[2] = b2;
        sc->bf_store[3] = b3;
        const double ev = (((b0 + b1) + b2) + b3) * dt;
        sc->bf_evolve = ev;
        sc->bf_time = bf_time + ev;
        sc->time = time_after;
    }
}

template <typename T>
hipError_t launch_bflux_leapfrog(const T *UH, const T *VH, int nx, int ny, int pitch, int n_ext, const double *dle,
                                 const double *dbe, double dt, double time_after, DevScalars *sc, hipStream_t s)
{
    hipLaunchKernelGGL(bflux_leapfrog_kernel<T>, dim3(1), dim3(kBfThreads), 0, s, UH, VH, nx, ny, pitch, n_ext, dle, dbe, dt,
                       time_after, sc);
    return hipGetLastError();
}
template hipError_t launch_bflux_leapfrog<double>(const double *, const double *, int, int, int, int,
                                                  const double *, const double *, double, double, DevScalars *,
                                                  hipStream_t);
template hipError_t launch_bflux_leapfrog<float>(const float *, const float *, int, int, int, int, const double *,
                                                 const double *, double, double, DevScalars *, hipStream_t);

// ---- forcing patches (forcing_mod.f90:80-156) ----
template <typename T>
__global__ void fv_forcing_kernel(T *S, T *UH, T *VH, const T *fS, const T *fUH, const T *fVH, int pitch, int i0,
                                  int i1, int j0, int j1, double start, double end, const DevScalars *sc, T *C)
{
    // update_U advanced the time by the stage's step: the step covers (time - dt, time]
    const double time = sc->time, dt = sc->stage_dt;
    if (!((time - dt) <= end && time >= start)) return;
    double frac;
    if (start == end) frac = 1.0;
    else frac = (fmin(time, end) - fmax(time - dt, start)) / (end - start);
    const int i = i0 + blockIdx.x * blockDim.x + threadIdx.x;
    const int j = j0 + blockIdx.y;
    if (i > i1 || j > j1) return;
    const long long o = (long long)j * pitch + i;
    if (fS) {
        if (C) {
            // FP32 state: the represented stage S - C, kept compensated
            const double v = ((double)S[o] - (double)C[o]) + frac * (double)fS[o];
            const T sn = (T)v;
            S[o] = sn;
            C[o] = (T)((double)sn - v);
        } else {
            S[o] = (T)((double)S[o] + frac * (double)fS[o]);
        }
    }
    if (fUH) UH[o] = (T)((double)UH[o] + frac * (double)fUH[o]);
    if (fVH) VH[o] = (T)((double)VH[o] + frac * (double)fVH[o]);
}

template <typename T>
__global__ void add_scaled_kernel(T *E, const T *fE, int pitch, int i0, int i1, int j0, int j1, T frac)
{
    const int i = i0 + blockIdx.x * blockDim.x + threadIdx.x;
    const int j = j0 + blockIdx.y;
    if (i > i1 || j > j1) return;
    const long long o = (long long)j * pitch + i;
    E[o] = E[o] + frac * fE[o];
}

template <typename T>
hipError_t launch_fv_forcing(T *S, T *UH, T *VH, const T *fS, const T *fUH, const T *fVH, int pitch, int i0, int i1,
                             int j0, int j1, double start, double end, const DevScalars *sc, hipStream_t s, T *C)
{
    if (i1 < i0 || j1 < j0) return hipSuccess;
    hipLaunchKernelGGL(fv_forcing_kernel<T>, dim3((i1 - i0 + 256) / 256, j1 - j0 + 1), dim3(256), 0, s, S, UH, VH, fS,
                       fUH, fVH, pitch, i0, i1, j0, j1, start, end, sc, C);
    return hipGetLastError();
}

template <typename T>
hipError_t launch_add_scaled(T *E, const T *fE, int pitch, int i0, int i1, int j0, int j1, T frac, hipStream_t s)
{
    if (i1 < i0 || j1 < j0) return hipSuccess;
    hipLaunchKernelGGL(add_scaled_kernel<T>, dim3((i1 - i0 + 256) / 256, j1 - j0 + 1), dim3(256), 0, s, E, fE, pitch,
                       i0, i1, j0, j1, frac);
    return hipGetLastError();
}

template <typename T>
__global__ void add_scaled_kahan_kernel(T *S, T *C, const T *fS, int pitch, int i0, int i1, int j0, int j1, T frac)
{
    const int i = i0 + blockIdx.x * blockDim.x + threadIdx.x;
    const int j = j0 + blockIdx.y;
    if (i > i1 || j > j1) return;
    const long long o = (long long)j * pitch + i;
    const T inc = frac * fS[o];
    if (C) {
        // the same compensated add as the fused step's stage forcing
        const T y = inc - C[o];
        const T t = S[o] + y;
        C[o] = (t - S[o]) - y;
        S[o] = t;
    } else {
        S[o] = S[o] + inc;
    }
}

template <typename T>
hipError_t launch_add_scaled_kahan(T *S, T *C, const T *fS, int pitch, int i0, int i1, int j0, int j1, T frac,
                                   hipStream_t s)
{
    if (i1 < i0 || j1 < j0) return hipSuccess;
    hipLaunchKernelGGL(add_scaled_kahan_kernel<T>, dim3((i1 - i0 + 256) / 256, j1 - j0 + 1), dim3(256), 0, s, S, C,
                       fS, pitch, i0, i1, j0, j1, frac);
    return hipGetLastError();
}
template hipError_t launch_add_scaled_kahan<double>(double *, double *, const double *, int, int, int, int, int,
                                                    double, hipStream_t);
template hipError_t launch_add_scaled_kahan<float>(float *, float *, const float *, int, int, int, int, int, float,
                                                   hipStream_t);

// ---- other boundary subroutines (boundary_mod.f90:23-75, 166-405) ----
// kind 2: transmissive_boundary (every variable, elevation included, from the cell inside);
// kind 3: boundary_stage_radiation_momentum; kind 4: boundary_stage_transmissive_momentum.
// One block: west/east columns, then south/north rows. The boundary stage is the sine
// function (use_fun; t from sc->time, else fun_stage from the host) or msl_linear, the
// boundary elevation the cell's own. C (Kahan compensation, may be null) follows the
// stage: copied with it (kind 2) or zeroed where the stage is set.
template <typename T>
__device__ void edge_bc_cell(T *S, T *C, T *Qn, T *Qt, T *E, long long b, long long n, bool neg, int kind,
                             double stage_b, int stag)
{
    if (kind == 2) {
        S[b] = S[n];
        if (C) C[b] = C[n];
        Qn[b] = Qn[n];
        Qt[b] = Qt[n];
        E[b] = E[n];
        return;
    }
    const double el = (double)E[b];
    S[b] = (T)stage_b;
    if (C) C[b] = T(0);
    if (stag) return;
    if (kind == 3) {
        const double local_h = fmax(stage_b - el, 0.0);
        const double sin_ = C ? (double)S[n] - (double)C[n] : (double)S[n];
        const double q = (sqrt(kGravity * fmax(sin_ - el, 0.0)) - sqrt(kGravity * local_h)) * local_h;
        Qn[b] = (T)(neg ? -q : q);
        Qt[b] = Qt[n];
    } else {
        Qn[b] = Qn[n];
        Qt[b] = Qt[n];
    }
}

template <typename T>
__global__ __launch_bounds__(1024) void edge_bc_kernel(T *S, T *C, T *UH, T *VH, T *E, int nx, int ny, int pitch,
                                                       int m, int kind, int use_fun, double amp, double period,
                                                       double fun_msl, double msl_linear, double fun_stage, int stag,
                                                       const DevScalars *sc)
{
    double stage_b = msl_linear;
    if (use_fun) {
        stage_b = fun_stage;
        if (sc) stage_b = fun_msl + amp * swhm::sin_acc(2.0 * acos(-1.0) * sc->time / period);
    }
    for (int j = threadIdx.x; j < ny; j += blockDim.x) {
        const long long o = (long long)j * pitch;
        if (m & 8) edge_bc_cell<T>(S, C, UH, VH, E, o, o + 1, true, kind, stage_b, stag);
        if (m & 2) edge_bc_cell<T>(S, C, UH, VH, E, o + nx - 1, o + nx - 2, false, kind, stage_b, stag);
    }
    __syncthreads();
    for (int i = threadIdx.x; i < nx; i += blockDim.x) {
        if (m & 4) edge_bc_cell<T>(S, C, VH, UH, E, i, (long long)pitch + i, true, kind, stage_b, stag);
        if (m & 1) {
            const long long o = (long long)(ny - 1) * pitch + i;
            edge_bc_cell<T>(S, C, VH, UH, E, o, o - pitch, false, kind, stage_b, stag);
        }
    }
}

template <typename T>
hipError_t launch_edge_bc(T *S, T *C, T *UH, T *VH, T *E, int nx, int ny, int pitch, int mask, int kind, int use_fun,
                          double amp, double period, double fun_msl, double msl_linear, double fun_stage, int stag,
                          const DevScalars *sc, hipStream_t s)
{
    if (nx < 3 || ny < 3 || kind < 2 || kind > 4) return hipErrorInvalidValue;
    hipLaunchKernelGGL(edge_bc_kernel<T>, dim3(1), dim3(1024), 0, s, S, C, UH, VH, E, nx, ny, pitch, mask, kind,
                       use_fun, amp, period, fun_msl, msl_linear, fun_stage, stag, sc);
    return hipGetLastError();
}
template hipError_t launch_edge_bc<double>(double *, double *, double *, double *, double *, int, int, int, int, int,
                                           int, double, double, double, double, double, int, const DevScalars *,
                                           hipStream_t);
template hipError_t launch_edge_bc<float>(float *, float *, float *, float *, float *, int, int, int, int, int, int,
                                          double, double, double, double, double, int, const DevScalars *,
                                          hipStream_t);

// ---- output snapshot ----
template <typename T>
__global__ void snapshot_kernel(const T *S, const T *C, const T *UH, const T *VH, const T *E, int nx, int ny,
                                int pitch, double *out, const DevScalars *sc, double t_host)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y;
    const long long n = (long long)nx * ny;
    if (i == 0 && j == 0) out[4 * n] = sc ? sc->time : t_host;
    if (i >= nx || j >= ny) return;
    const long long o = (long long)j * pitch + i, q = (long long)j * nx + i;
    out[q] = C ? (double)S[o] - (double)C[o] : (double)S[o];
    out[n + q] = (double)UH[o];
    out[2 * n + q] = (double)VH[o];
    out[3 * n + q] = (double)E[o];
}

template <typename T>
hipError_t launch_snapshot(const T *S, const T *C, const T *UH, const T *VH, const T *E, int nx, int ny, int pitch,
                           double *out, const DevScalars *sc, double t_host, hipStream_t s)
{
    hipLaunchKernelGGL(snapshot_kernel<T>, dim3((nx + 255) / 256, ny), dim3(256), 0, s, S, C, UH, VH, E, nx, ny, pitch,
                       out, sc, t_host);
    return hipGetLastError();
}
template hipError_t launch_snapshot<double>(const double *, const double *, const double *, const double *,
                                            const double *, int, int, int, double *, const DevScalars *, double,
                                            hipStream_t);
template hipError_t launch_snapshot<float>(const float *, const float *, const float *, const float *, const float *,
                                           int, int, int, double *, const DevScalars *, double, hipStream_t);

// ---- Flather boundary (boundary_mod.f90:502-720) ----
// One edge cell b with the cell n inside it. West and south edges (sgn < 0) take the
// incoming characteristic from outside as w1 = vel + sqrt(g/h) stage, east and north
// (sgn > 0) as vel - sqrt(g/h) stage; normal / tang: the momentum planes across and
// along the edge. Arithmetic in FP64 in both precisions, in the reference's order.
template <typename T>
__device__ void flather_cell(T *S, T *C, T *Qn, T *Qt, const T *E, long long b, long long n, bool neg,
                             double msl_linear, int use_fun, double stage_fun)
{
    const double kMin = kMinDepth, kG = kGravity;
    const double eb = (double)E[b];
    double stage_o, vel_o, depth_o;
    if (!use_fun) {
        stage_o = msl_linear > eb ? msl_linear : eb;  // max(msl_linear, elevation)
        vel_o = 0.0;
        depth_o = stage_o - eb;
    } else {
        stage_o = stage_fun;
        depth_o = stage_fun - eb;
        const double q_o = 0.0;  // the boundary function's momenta
        vel_o = depth_o <= kMin ? 0.0 : q_o / depth_o;
    }
    const double sn = C ? (double)S[n] - (double)C[n] : (double)S[n];
    const double depth_inside = sn - (double)E[n];
    double st, qn, qt;
    if ((depth_inside > kMin) && (depth_o > kMin)) {
        const double dinv = 1.0 / depth_inside;
        const double sg = sqrt(kG * dinv);
        const double qin = (double)Qn[n];
        double w1, w2, w3;
        if (neg) {
            w1 = vel_o + sg * stage_o;
            w2 = (qin < 0.0) ? (double)Qt[n] * dinv : 0.0;
            w3 = qin * dinv - sg * sn;
            st = (-w3 + w1) / (2.0 * sg);
        } else {
            w1 = vel_o - sg * stage_o;
            w2 = (qin > 0.0) ? (double)Qt[n] * dinv : 0.0;
            w3 = qin * dinv + sg * sn;
            st = (w3 - w1) / (2.0 * sg);
        }
        qn = (w3 + w1) * 0.5 * depth_inside;
        qt = w2 * depth_inside;
    } else {
        st = stage_o;
        qn = 0.0;
        qt = 0.0;
    }
    S[b] = (T)st;
    if (C) C[b] = T(0);
    Qn[b] = (T)qn;
    Qt[b] = (T)qt;
}

template <typename T>
__global__ __launch_bounds__(1024) void flather_kernel(T *S, T *C, T *UH, T *VH, const T *E, int nx, int ny,
                                                       int pitch, int m, double msl_linear, int use_fun, double amp,
                                                       double period, double fun_msl, double fun_stage,
                                                       const DevScalars *sc)
{
    double stage_fun = fun_stage;
    if (use_fun && sc) {
        const double two_pi = 2.0 * acos(-1.0);
        stage_fun = fun_msl + amp * swhm::sin_acc(two_pi * sc->time / period);
    }
    for (int j = threadIdx.x; j < ny; j += blockDim.x) {
        const long long o = (long long)j * pitch;
        if (m & 8) flather_cell<T>(S, C, UH, VH, E, o, o + 1, true, msl_linear, use_fun, stage_fun);
        if (m & 2) flather_cell<T>(S, C, UH, VH, E, o + nx - 1, o + nx - 2, false, msl_linear, use_fun, stage_fun);
    }
    __syncthreads();
    for (int i = threadIdx.x; i < nx; i += blockDim.x) {
        if (m & 4) flather_cell<T>(S, C, VH, UH, E, i, (long long)pitch + i, true, msl_linear, use_fun, stage_fun);
        if (m & 1) {
            const long long o = (long long)(ny - 1) * pitch + i;
            flather_cell<T>(S, C, VH, UH, E, o, o - pitch, false, msl_linear, use_fun, stage_fun);
        }
    }
}

template <typename T>
hipError_t launch_flather(T *S, T *C, T *UH, T *VH, const T *E, int nx, int ny, int pitch, int mask, double msl_linear,
                          int use_fun, double amp, double period, double fun_msl, double fun_stage,
                          const DevScalars *sc, hipStream_t s)
{
    if (nx < 3 || ny < 3) return hipErrorInvalidValue;
    hipLaunchKernelGGL(flather_kernel<T>, dim3(1), dim3(1024), 0, s, S, C, UH, VH, E, nx, ny, pitch, mask, msl_linear,
                       use_fun, amp, period, fun_msl, fun_stage, sc);
    return hipGetLastError();
}
template hipError_t launch_flather<double>(double *, double *, double *, double *, const double *, int, int, int, int,
                                           double, int, double, double, double, double, const DevScalars *,
                                           hipStream_t);
template hipError_t launch_flather<float>(float *, float *, float *, float *, const float *, int, int, int, int,
                                          double, int, double, double, double, double, const DevScalars *,
                                          hipStream_t);

template hipError_t launch_fv_forcing<double>(double *, double *, double *, const double *, const double *,
                                              const double *, int, int, int, int, int, double, double,
                                              const DevScalars *, hipStream_t, double *);
template hipError_t launch_fv_forcing<float>(float *, float *, float *, const float *, const float *, const float *,
                                             int, int, int, int, int, double, double, const DevScalars *, hipStream_t, float *);
template hipError_t launch_add_scaled<double>(double *, const double *, int, int, int, int, int, double, hipStream_t);
template hipError_t launch_add_scaled<float>(float *, const float *, int, int, int, int, int, float, hipStream_t);

// ---- standalone max_U update (domain_mod.f90:2243-2404) ----
template <typename T>
__global__ void maxq_kernel(const T *__restrict__ S, const T *__restrict__ C, const T *__restrict__ UH,
                            const T *__restrict__ VH, const T *__restrict__ E, int nx, int ny, int pitch,
                            int staggered, int truely_linear, T msl, int m, T *__restrict__ mx,
                            const DevScalars *tsrc, T t_host, T arrival_stage)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y;
    if (i >= nx || j >= ny) return;
    const long long o = (long long)j * pitch + i;
    const long long P = (long long)pitch * ny;
    const T MIN = T(kMinDepth), HALF = T(0.5);
    const T t = tsrc ? (T)tsrc->time : t_host;
    const T s = C ? S[o] - C[o] : S[o];
    const T uh = UH[o], vh = VH[o], e = E[o];
    if (m & SWH_MAXQ_MAX_STAGE) {
        if (m & SWH_MAXQ_TIME_OF_MAX_STAGE) {
            if (s > mx[MX_STAGE * P + o]) {
                mx[MX_STAGE * P + o] = s;
                mx[MX_TOMS * P + o] = t;
            }
        } else if (s > mx[MX_STAGE * P + o]) {
            mx[MX_STAGE * P + o] = s;
        }
    }
    if (m & SWH_MAXQ_MAX_SPEED) {
        T sp;
        if (staggered) {
            // half the depth summed over the cell and its east / north neighbour; on the
            // east column / north row the sum is the cell alone (domain_mod.f90:2311-2320)
            const long long oE = (long long)j * pitch + i + 1;
            const long long oN = (long long)(j + 1) * pitch + i;
            const T h0 = truely_linear ? (msl - e) : (s - e);
            T dE = h0, dN = h0;
            if (i + 1 < nx) dE = h0 + (truely_linear ? (msl - E[oE]) : ((C ? S[oE] - C[oE] : S[oE]) - E[oE]));
            if (j + 1 < ny) dN = h0 + (truely_linear ? (msl - E[oN]) : ((C ? S[oN] - C[oN] : S[oN]) - E[oN]));
            dE = HALF * dE;
            dN = HALF * dN;
            const T iE = dE > MIN ? T(1) / dE : T(0);
            const T iN = dN > MIN ? T(1) / dN : T(0);
            sp = sqrt(uh * uh * (iE * iE) + vh * vh * (iN * iN));
        } else {
            const T ld = s - e;
            const T li = ld > MIN ? T(1) / ld : T(0);
            sp = sqrt((uh * uh + vh * vh) * (li * li));
        }
        if (sp > mx[MX_SPEED * P + o]) mx[MX_SPEED * P + o] = sp;
    }
    if (m & SWH_MAXQ_MAX_FLUX) {
        const T f = sqrt(uh * uh + vh * vh);
        if (f > mx[MX_FLUX * P + o]) mx[MX_FLUX * P + o] = f;
    }
    if (m & SWH_MAXQ_ARRIVAL_TIME) {
        if ((s - e) > MIN && s > arrival_stage && mx[MX_ARRIVAL * P + o] < T(0)) mx[MX_ARRIVAL * P + o] = t;
    }
    if (m & SWH_MAXQ_MIN_STAGE) {
        if (s < mx[MX_MIN_STAGE * P + o]) mx[MX_MIN_STAGE * P + o] = s;
    }
}

template <typename T>
hipError_t launch_maxq(const T *S, const T *C, const T *UH, const T *VH, const T *E, int nx, int ny, int pitch,
                       int staggered, int truely_linear, T msl, int mask, T *mx, const DevScalars *tsrc, T t,
                       T arrival_stage, hipStream_t s)
{
    dim3 block(256), grid((nx + 255) / 256, ny);
    hipLaunchKernelGGL(maxq_kernel<T>, grid, block, 0, s, S, C, UH, VH, E, nx, ny, pitch, staggered, truely_linear,
                       msl, mask, mx, tsrc, t, arrival_stage);
    return hipGetLastError();
}
template hipError_t launch_maxq<double>(const double *, const double *, const double *, const double *,
                                        const double *, int, int, int, int, int, double, int, double *,
                                        const DevScalars *, double, double, hipStream_t);
template hipError_t launch_maxq<float>(const float *, const float *, const float *, const float *, const float *,
                                       int, int, int, int, int, float, int, float *, const DevScalars *, float, float,
                                       hipStream_t);

// ---- volume_interior (domain_mod.f90:1815-1849) ----
// Pass 1: one block per interior row: compensated row sum of (stage - elev), times
// area_cell_y(j). Pass 2: one block sums the rows in order.
template <typename T>
__global__ __launch_bounds__(256) void volume_rows_kernel(const T *__restrict__ S, const T *__restrict__ C,
                                                          const T *__restrict__ E, int nx, int pitch, int n_ext,
                                                          const double *__restrict__ area, double *row_work,
                                                          const int *__restrict__ mask)
{
    const int j = n_ext + blockIdx.x;
    DD acc{0.0, 0.0};
    for (int i = n_ext + threadIdx.x; i <= nx - n_ext - 1; i += blockDim.x) {
        const long long o = (long long)j * pitch + i;
        // priority-domain cells only (domain_mod.f90:1872-1876), mask is nx*ny row-major
        if (mask && mask[(long long)j * nx + i] != 1) continue;
        double sv = (double)S[o];
        if (C) acc = dd_add(acc, -(double)C[o]);
        acc = dd_add(acc, sv - (double)E[o]);
    }
    DD r = block_reduce_dd(acc);
    if (threadIdx.x == 0) {
        DD prod;
        // area * (hi + lo) with the low-order part kept separately
        prod.hi = area[j] * r.hi;
        prod.lo = area[j] * r.lo;
        row_work[2 * blockIdx.x] = prod.hi;
        row_work[2 * blockIdx.x + 1] = prod.lo;
    }
}

__global__ __launch_bounds__(256) void volume_sum_kernel(const double *row_work, int nrows, double *out)
{
    DD acc{0.0, 0.0};
    for (int k = threadIdx.x; k < nrows; k += blockDim.x) acc = dd_add(acc, DD{row_work[2 * k], row_work[2 * k + 1]});
    DD r = block_reduce_dd(acc);
    if (threadIdx.x == 0) *out = r.hi + r.lo;
}

template <typename T>
hipError_t launch_volume(const T *S, const T *C, const T *E, int nx, int ny, int pitch, int n_ext,
                         const double *area, double *row_work, double *out, hipStream_t s, const int *mask)
{
    const int nrows = ny - 2 * n_ext;
    if (nrows <= 0) {
        return hipMemsetAsync(out, 0, sizeof(double), s);
    }
    hipLaunchKernelGGL(volume_rows_kernel<T>, dim3(nrows), dim3(256), 0, s, S, C, E, nx, pitch, n_ext, area,
                       row_work, mask);
    hipLaunchKernelGGL(volume_sum_kernel, dim3(1), dim3(256), 0, s, row_work, nrows, out);
    return hipGetLastError();
}
template hipError_t launch_volume<double>(const double *, const double *, const double *, int, int, int, int,
                                          const double *, double *, double *, hipStream_t, const int *);
template hipError_t launch_volume<float>(const float *, const float *, const float *, int, int, int, int,
                                         const double *, double *, double *, hipStream_t, const int *);

// ---- point gauges (point_gauge_mod.f90:208-245) ----
template <typename T>
__global__ void gauges_kernel(const T *__restrict__ S, const T *__restrict__ C, const T *__restrict__ UH,
                              const T *__restrict__ VH, int n, const long long *__restrict__ idx, double *values,
                              double *times, int sample, const DevScalars *tsrc, double t)
{
    const int g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g == 0) times[sample] = tsrc ? tsrc->time : t;
    if (g >= n) return;
    const long long o = idx[g];
    double *v = values + (long long)sample * 3 * n;
    v[g] = C ? (double)S[o] - (double)C[o] : (double)S[o];
    v[n + g] = (double)UH[o];
    v[2 * n + g] = (double)VH[o];
}

template <typename T>
hipError_t launch_gauges(const T *S, const T *C, const T *UH, const T *VH, int n, const long long *idx,
                         double *values, double *times, int sample, const DevScalars *tsrc, double t, hipStream_t s)
{
    const int nb = n > 0 ? (n + 255) / 256 : 1;
    hipLaunchKernelGGL(gauges_kernel<T>, dim3(nb), dim3(256), 0, s, S, C, UH, VH, n, idx, values, times, sample, tsrc,
                       t);
    return hipGetLastError();
}
template hipError_t launch_gauges<double>(const double *, const double *, const double *, const double *, int,
                                          const long long *, double *, double *, int, const DevScalars *, double,
                                          hipStream_t);
template hipError_t launch_gauges<float>(const float *, const float *, const float *, const float *, int,
                                         const long long *, double *, double *, int, const DevScalars *, double,
                                         hipStream_t);

// ---- stationary_timestep_max (domain_mod.f90:1906-1962) ----
// The minimum of positive doubles is order-independent; it is taken on their bit
// patterns, which order like the values.
template <typename T>
__global__ void timestep_max_kernel(const T *__restrict__ S, const T *__restrict__ E, int nx, int ny, int pitch,
                                    int linear, double msl, const double *__restrict__ dbe,
                                    const double *__restrict__ dle, unsigned long long *out_bits)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y;
    double v = 1.0e300;
    if (i < nx && j < ny) {
        const long long o = (long long)j * pitch + i;
        const double h = linear ? msl - (double)E[o] : (double)S[o] - (double)E[o];
        const double dist = fmin(dbe[j] + dbe[j + 1], dle[i] + dle[i + 1]);
        v = 0.5 * dist / sqrt(kGravity * fmax(h, kMinDepth));
    }
    for (int off = 32; off > 0; off >>= 1) v = fmin(v, __shfl_xor(v, off));
    if ((threadIdx.x & 63) == 0) atomicMin(out_bits, (unsigned long long)__double_as_longlong(v));
}

__global__ void fill_double_kernel(double *p, double v) { *p = v; }

template <typename T>
hipError_t launch_timestep_max(const T *S, const T *E, int nx, int ny, int pitch, int linear, double msl,
                               const double *dbe, const double *dle, double *out, hipStream_t s)
{
    hipLaunchKernelGGL(fill_double_kernel, dim3(1), dim3(1), 0, s, out, 1.0e300);
    hipLaunchKernelGGL(timestep_max_kernel<T>, dim3((nx + 255) / 256, ny), dim3(256), 0, s, S, E, nx, ny, pitch,
                       linear, msl, dbe, dle, (unsigned long long *)out);
    return hipGetLastError();
}
template hipError_t launch_timestep_max<double>(const double *, const double *, int, int, int, int, double,
                                                const double *, const double *, double *, hipStream_t);
template hipError_t launch_timestep_max<float>(const float *, const float *, int, int, int, int, double,
                                               const double *, const double *, double *, hipStream_t);

}  // namespace swh

namespace swh {

// ---- plain streaming copy (measurement reference) ----
// 16-byte loads and stores, grid-stride; used to scale rocprofv3's FETCH_SIZE /
// WRITE_SIZE counters to bytes and to measure this GPU's streaming bandwidth.
__global__ __launch_bounds__(256) void probe_copy_kernel(const float4 *__restrict__ src, float4 *__restrict__ dst,
                                                         long long n)
{
    for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (long long)gridDim.x * blockDim.x)
        dst[t] = src[t];
}

hipError_t launch_probe_copy(const void *src, void *dst, long long bytes, hipStream_t s)
{
    const long long n = bytes / 16;
    const long long blocks = std::min<long long>((n + 255) / 256, 256LL * 64);
    hipLaunchKernelGGL(probe_copy_kernel, dim3((unsigned)blocks), dim3(256), 0, s, (const float4 *)src,
                       (float4 *)dst, n);
    return hipGetLastError();
}

// ---- host transfers: dense FP64 planes <-> pitched planes of the storage type ----
template <typename T>
__global__ void plane_from_dense_kernel(const double *__restrict__ src, T *__restrict__ dst, T *__restrict__ comp,
                                        int nx, int ny, int pitch)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y;
    if (i >= pitch || j >= ny) return;
    const long long o = (long long)j * pitch + i;
    const double v = (i < nx && src) ? src[(long long)j * nx + i] : 0.0;
    const T t = (T)v;
    dst[o] = t;
    if (comp) comp[o] = (T)((double)t - v);   // dst - comp represents v (to about 48 bits in FP32)
}

template <typename T>
__global__ void plane_to_dense_kernel(const T *__restrict__ src, const T *__restrict__ comp, double *__restrict__ dst,
                                      int nx, int ny, int pitch)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y;
    if (i >= nx || j >= ny) return;
    const long long o = (long long)j * pitch + i;
    dst[(long long)j * nx + i] = comp ? (double)src[o] - (double)comp[o] : (double)src[o];
}

template <typename T>
hipError_t launch_plane_from_dense(const double *src, T *dst, T *comp, int nx, int ny, int pitch, hipStream_t s)
{
    hipLaunchKernelGGL(plane_from_dense_kernel<T>, dim3((pitch + 255) / 256, ny), dim3(256), 0, s, src, dst, comp, nx,
                       ny, pitch);
    return hipGetLastError();
}

template <typename T>
hipError_t launch_plane_to_dense(const T *src, const T *comp, double *dst, int nx, int ny, int pitch, hipStream_t s)
{
    hipLaunchKernelGGL(plane_to_dense_kernel<T>, dim3((nx + 255) / 256, ny), dim3(256), 0, s, src, comp, dst, nx, ny,
                       pitch);
    return hipGetLastError();
}

template hipError_t launch_plane_from_dense<double>(const double *, double *, double *, int, int, int, hipStream_t);
template hipError_t launch_plane_from_dense<float>(const double *, float *, float *, int, int, int, hipStream_t);
template hipError_t launch_plane_to_dense<double>(const double *, const double *, double *, int, int, int,
                                                  hipStream_t);
template hipError_t launch_plane_to_dense<float>(const float *, const float *, double *, int, int, int, hipStream_t);

// ---- VALU throughput (measurement reference) ----
// Eight independent fused multiply-add chains per lane, iters rounds: 8 * iters VALU
// FMA wave-instructions per wavefront (FP32 ones unpacked: this code is built without
// packed FP32), the issue rate that the stencil kernels' VALU counts are compared with.
template <typename T>
__global__ __launch_bounds__(256) void probe_valu_kernel(T x, T y, int iters, T *__restrict__ out)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    T a[8];
    for (int k = 0; k < 8; k++) a[k] = (T)(t + k);
    for (int it = 0; it < iters; it++)
        for (int k = 0; k < 8; k++) {
            if constexpr (sizeof(T) == 8) a[k] = __builtin_fma(a[k], x, y);
            else a[k] = __builtin_fmaf(a[k], x, y);
        }
    T r = a[0];
    for (int k = 1; k < 8; k++) r += a[k];
    out[t] = r;
}

hipError_t launch_probe_valu(int fp64, int blocks, int iters, void *out, hipStream_t s)
{
    if (fp64)
        hipLaunchKernelGGL(probe_valu_kernel<double>, dim3(blocks), dim3(256), 0, s, 0.999999, 1e-7, iters,
                           (double *)out);
    else
        hipLaunchKernelGGL(probe_valu_kernel<float>, dim3(blocks), dim3(256), 0, s, 0.999999f, 1e-7f, iters,
                           (float *)out);
    return hipGetLastError();
}

// ---- the FP64 friction pow and boundary sine (swh_mathx.h) on the device ----
// One argument per thread; the same functions compiled for the host give the same bits.
__global__ __launch_bounds__(256) void probe_math_kernel(const double *__restrict__ x, const double *__restrict__ y,
                                                         const double *__restrict__ t, double *__restrict__ pw,
                                                         double *__restrict__ sn, int n)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    pw[k] = swhm::pow_acc(x[k], y[k]);
    sn[k] = swhm::sin_acc(t[k]);
}

hipError_t launch_probe_math(const double *x, const double *y, const double *t, double *pw, double *sn, int n,
                             hipStream_t s)
{
    hipLaunchKernelGGL(probe_math_kernel, dim3((n + 255) / 256), dim3(256), 0, s, x, y, t, pw, sn, n);
    return hipGetLastError();
}

}  // namespace swh
