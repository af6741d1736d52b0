// swh_stats.hip -- domain statistics and the stability scan on the device.
//
// compute_domain_statistics_new (domain_mod.f90:683-849): over the cells more than
// exterior_cells_width from the edge (and, in a multidomain, on the priority domain),
// the maximum and minimum stage and speed of wet cells, and the potential and kinetic
// energy / rho. The reference accumulates each row's energies in column order into
// FP64 row work arrays and then sums the rows in row order; here one wavefront owns a
// row, each lane forms its cell's terms, and the row sums are accumulated lane by lane
// in column order (so the sums round exactly as the reference's loops do), then one
// wavefront adds the rows in order.
//
// check_multidomain_stability (multidomain_mod.f90:348-455): counts the interior values
// of stage, uh, vh and elevation that are NaN or beyond +-HUGE(1.0)/2 (the reference's
// single-precision bound, used to stop a run that has blown up).
#include <cfloat>
#include <mutex>

#include "swh_domain.h"

namespace swh {

namespace {

constexpr double kBignum = 1.7014117331926443e38;   // HUGE(1.0_sp) / 2, as the reference's real(dp) parameter

template <typename T>
struct StatsArgs {
    const T *S, *C, *UH, *VH, *E;
    const int *pmask;   // is_priority_domain_not_periodic (nx, ny) dense, or null
    int nx, ny, pitch, ecw, staggered, tl;
    double msl;
    const double *area;
    double *rows;   // [6][ny]: potential, kinetic, max stage, min stage, max speed^2, min speed^2
};

template <typename T>
__device__ inline double stage_at(const StatsArgs<T> &a, long long o)
{
    return a.C ? (double)a.S[o] - (double)a.C[o] : (double)a.S[o];
}

// One wavefront per interior row (4 rows per block).
template <typename T>
__global__ __launch_bounds__(256) void stats_rows_kernel(StatsArgs<T> a)
{
    const int lane = threadIdx.x & 63;
    const int j = (int)blockIdx.x * 4 + (int)(threadIdx.x >> 6) + a.ecw;
    if (j >= a.ny - a.ecw) return;   // the whole wavefront leaves together
    const int jn = j + 1 < a.ny ? j + 1 : j;   // the row above (the reference reads row j+1)
    const double area_j = a.area[j], area_n = a.area[jn];
    const double c_pot = (area_j * kGravity) * 0.5;
    const double area_v = 0.5 * (area_j + area_n);
    double pot = 0.0, kin = 0.0;
    double mxs = -DBL_MAX, mns = DBL_MAX, mxv = 0.0, mnv = DBL_MAX;
    const int iend = a.nx - a.ecw;
    for (int i0 = a.ecw; i0 < iend; i0 += 64) {
        const int i = i0 + lane;
        int counted = 0, wet = 0;
        double tpot = 0.0, tkin = 0.0;
        if (i < iend) {
            counted = a.pmask ? (a.pmask[(long long)j * a.nx + i] != 0) : 1;
            if (counted) {
                const long long o = (long long)j * a.pitch + i;
                const double s = stage_at(a, o), e = (double)a.E[o];
                wet = a.tl ? (a.msl > e + kMinDepth) : (s > e + kMinDepth);
                const double w0 = s + e, d1 = s - e;
                const double e_flow = wet ? w0 * d1 : 0.0;
                const double w2 = a.msl + e, d3 = a.msl - e;
                const double e_const = (a.msl > e + kMinDepth) ? w2 * d3 : 0.0;
                tpot = c_pot * (e_flow - e_const);
                if (wet) {
                    mxs = fmax(mxs, s);
                    mns = fmin(mns, s);
                    const double uh = (double)a.UH[o], vh = (double)a.VH[o];
                    double sp2;
                    if (a.staggered) {
                        const int ie = i + 1 < a.nx ? i + 1 : i;
                        const long long oe = (long long)j * a.pitch + ie, on = (long long)jn * a.pitch + i;
                        const double ee = (double)a.E[oe], en = (double)a.E[on];
                        double dE, dN;
                        if (a.tl) {
                            dE = 0.5 * ((a.msl - e) + (a.msl - ee));
                            dN = 0.5 * ((a.msl - e) + (a.msl - en));
                        } else {
                            dE = 0.5 * ((s - e) + (stage_at(a, oe) - ee));
                            dN = 0.5 * ((s - e) + (stage_at(a, on) - en));
                        }
                        const double dEi = dE > kMinDepth ? 1.0 / dE : 0.0;
                        const double dNi = dN > kMinDepth ? 1.0 / dN : 0.0;
                        sp2 = uh * uh * (dEi * dEi) + vh * vh * (dNi * dNi);
                        tkin = 0.5 * (area_j * uh * uh * dEi + area_v * vh * vh * dNi);
                    } else {
                        const double dC = s - e;
                        sp2 = (uh * uh + vh * vh) / (dC * dC);
                        tkin = area_j * 0.5 * dC * sp2;
                    }
                    mxv = fmax(mxv, sp2);
                    mnv = fmin(mnv, sp2);
                }
            }
        }
        // the row sums in column order: every lane adds the same values in the same order
        for (int l = 0; l < 64; l++) {
            const double tp = __shfl(tpot, l), tk = __shfl(tkin, l);
            const int cl = __shfl(counted, l), wl = __shfl(wet, l);
            if (cl) pot = pot + tp;
            if (wl) kin = kin + tk;
        }
    }
    for (int off = 32; off > 0; off >>= 1) {
        mxs = fmax(mxs, __shfl_xor(mxs, off));
        mns = fmin(mns, __shfl_xor(mns, off));
        mxv = fmax(mxv, __shfl_xor(mxv, off));
        mnv = fmin(mnv, __shfl_xor(mnv, off));
    }
    if (lane == 0) {
        double *r = a.rows;
        r[j] = pot;
        r[a.ny + j] = kin;
        r[2 * a.ny + j] = mxs;
        r[3 * a.ny + j] = mns;
        r[4 * a.ny + j] = mxv;
        r[5 * a.ny + j] = mnv;
    }
}

// One wavefront: the row sums added in row order (rows outside the interior hold zero
// in the reference's work arrays and add nothing), extrema over rows.
// out: maxstage, maxspeed, minstage, minspeed, potential, kinetic, total.
__global__ __launch_bounds__(64) void stats_final_kernel(const double *__restrict__ rows, int ny, int ecw,
                                                         double *__restrict__ out)
{
    const int lane = threadIdx.x;
    double mxs = -DBL_MAX, mns = DBL_MAX, mxv = 0.0, mnv = DBL_MAX;
    for (int j = ecw + lane; j < ny - ecw; j += 64) {
        mxs = fmax(mxs, rows[2 * ny + j]);
        mns = fmin(mns, rows[3 * ny + j]);
        mxv = fmax(mxv, rows[4 * ny + j]);
        mnv = fmin(mnv, rows[5 * ny + j]);
    }
    for (int off = 32; off > 0; off >>= 1) {
        mxs = fmax(mxs, __shfl_xor(mxs, off));
        mns = fmin(mns, __shfl_xor(mns, off));
        mxv = fmax(mxv, __shfl_xor(mxv, off));
        mnv = fmin(mnv, __shfl_xor(mnv, off));
    }
    if (lane == 0) {
        double pot = 0.0, kin = 0.0;
        for (int j = ecw; j < ny - ecw; j++) {
            pot = pot + rows[j];
            kin = kin + rows[ny + j];
        }
        out[0] = mxs;
        out[1] = sqrt(mxv);
        out[2] = mns;
        out[3] = sqrt(mnv);
        out[4] = pot;
        out[5] = kin;
        out[6] = kin + pot;
    }
}

// Stability scan: interior cells of the four planes; one count per wavefront added to
// the counter (which the first kernel clears).
template <typename T>
__global__ __launch_bounds__(256) void stability_kernel(const T *__restrict__ S, const T *__restrict__ C,
                                                        const T *__restrict__ UH, const T *__restrict__ VH,
                                                        const T *__restrict__ E, int nx, int ny, int pitch, int ecw,
                                                        unsigned long long *__restrict__ count)
{
    const int w = nx - 2 * ecw;
    const long long ncell = (long long)w * (ny - 2 * ecw);
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    int bad = 0;
    if (t < ncell) {
        const int i = ecw + (int)(t % w), j = ecw + (int)(t / w);
        const long long o = (long long)j * pitch + i;
        const double v[4] = {C ? (double)S[o] - (double)C[o] : (double)S[o], (double)UH[o], (double)VH[o],
                             (double)E[o]};
        for (int k = 0; k < 4; k++) bad += (v[k] > kBignum) || (v[k] < -kBignum) || (v[k] != v[k]);
    }
    for (int off = 32; off > 0; off >>= 1) bad += __shfl_xor(bad, off);
    if ((threadIdx.x & 63) == 0 && bad) atomicAdd(count, (unsigned long long)bad);
}

__global__ void clear_u64_kernel(unsigned long long *p) { *p = 0ULL; }

__global__ void stats_empty_kernel(double *out)
{
    out[0] = -DBL_MAX;
    out[1] = 0.0;
    out[2] = DBL_MAX;
    out[3] = sqrt(DBL_MAX);
    out[4] = 0.0;
    out[5] = 0.0;
    out[6] = 0.0;
}

}  // namespace

int domain_statistics_launch(swh_domain *h, const int *pmask, double *rows, double *out, hipStream_t stream)
{
    const int c = h->cur;
    const int ecw = h->n_ext;
    if (h->ny - 2 * ecw <= 0 || h->nx - 2 * ecw <= 0) {
        // no interior cell: the reference's initial values
        hipLaunchKernelGGL(stats_empty_kernel, dim3(1), dim3(1), 0, stream, out);
        SWH_HIP(hipGetLastError());
        return SWH_OK;
    }
    const int blocks = (h->ny - 2 * ecw + 3) / 4;
    if (h->prec == SWH_PREC_FP64) {
        StatsArgs<double> a{(const double *)h->S[c], nullptr, (const double *)h->UH[c], (const double *)h->VH[c],
                            (const double *)h->E, pmask, h->nx, h->ny, h->pitch, ecw, h->leapfrog() ? 1 : 0,
                            h->tl_solver() ? 1 : 0, h->msl, h->d_area, rows};
        hipLaunchKernelGGL(stats_rows_kernel<double>, dim3(blocks), dim3(256), 0, stream, a);
    } else {
        StatsArgs<float> a{(const float *)h->S[c], (const float *)h->C[c], (const float *)h->UH[c],
                           (const float *)h->VH[c], (const float *)h->E, pmask, h->nx, h->ny, h->pitch, ecw,
                           h->leapfrog() ? 1 : 0, h->tl_solver() ? 1 : 0, h->msl, h->d_area, rows};
        hipLaunchKernelGGL(stats_rows_kernel<float>, dim3(blocks), dim3(256), 0, stream, a);
    }
    SWH_HIP(hipGetLastError());
    hipLaunchKernelGGL(stats_final_kernel, dim3(1), dim3(64), 0, stream, (const double *)rows, h->ny, ecw, out);
    SWH_HIP(hipGetLastError());
    return SWH_OK;
}

int domain_stability_launch(swh_domain *h, unsigned long long *count, hipStream_t stream)
{
    const int c = h->cur;
    const int ecw = h->n_ext;
    hipLaunchKernelGGL(clear_u64_kernel, dim3(1), dim3(1), 0, stream, count);
    SWH_HIP(hipGetLastError());
    const long long ncell = (long long)std::max(0, h->nx - 2 * ecw) * std::max(0, h->ny - 2 * ecw);
    if (ncell == 0) return SWH_OK;
    const unsigned blocks = (unsigned)((ncell + 255) / 256);
    if (h->prec == SWH_PREC_FP64)
        hipLaunchKernelGGL(stability_kernel<double>, dim3(blocks), dim3(256), 0, stream, (const double *)h->S[c],
                           (const double *)nullptr, (const double *)h->UH[c], (const double *)h->VH[c],
                           (const double *)h->E, h->nx, h->ny, h->pitch, ecw, count);
    else
        hipLaunchKernelGGL(stability_kernel<float>, dim3(blocks), dim3(256), 0, stream, (const float *)h->S[c],
                           (const float *)h->C[c], (const float *)h->UH[c], (const float *)h->VH[c],
                           (const float *)h->E, h->nx, h->ny, h->pitch, ecw, count);
    SWH_HIP(hipGetLastError());
    return SWH_OK;
}

}  // namespace swh

// ---------------------------------------------------------------------------
// The statistics entry point that allocates nothing (swh_domain_statistics_noalloc):
// its row sums and results live in this file-scope device array instead of the
// handle's lazily allocated buffer. Each GPU holds its own copy until the process
// ends; host code reaches it only through the address hipGetSymbolAddress returns
// for the current device (the array's name is a host-side stand-in).
// ---------------------------------------------------------------------------
// 6 row arrays of up to 16384 rows, 7 results, a spare slot: 786 KB per GPU
constexpr int kScratchMaxRows = 16384;
constexpr size_t kScratchDoubles = 6 * (size_t)kScratchMaxRows + 8;
static __device__ __align__(256) double scratch[kScratchDoubles];

namespace swh {
namespace {

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double *g_scratch_ptr[kMaxDevices] = {};

// The device address of this GPU's scratch array, looked up once per device with
// that device current.
int scratch_on_current_device(double **out)
{
    int dev = 0;
    SWH_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) {
        set_error("swh_domain_statistics_noalloc: device index out of range");
        return SWH_EINVAL;
    }
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void *p = nullptr;
        SWH_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        g_scratch_ptr[dev] = (double *)p;
    }
    *out = g_scratch_ptr[dev];
    return SWH_OK;
}

}  // namespace

}  // namespace swh

extern "C" int swh_domain_statistics_noalloc(swh_domain *h, swh_domain_stats *out)
{
    using namespace swh;
    if (!h || !out) {
        set_error("swh_domain_statistics_noalloc: null argument");
        return SWH_EINVAL;
    }
    if (h->ny > kScratchMaxRows) {
        set_error("swh_domain_statistics_noalloc: the domain has " + std::to_string(h->ny) +
                  " rows, more than the scratch array's maximum of " + std::to_string(kScratchMaxRows) +
                  " (use swh_domain_statistics)");
        return SWH_EINVAL;
    }
    double *buf = nullptr;
    int rc = scratch_on_current_device(&buf);
    if (rc) return rc;
    // every value the final pass reads is written earlier in this call (the interior
    // rows, then the results), so nothing needs clearing
    double *res = buf + 6 * (size_t)h->ny;
    if ((rc = domain_statistics_launch(h, nullptr, buf, res, h->stream))) return rc;
    double v[7];
    SWH_HIP(hipStreamSynchronize(h->stream));
    SWH_HIP(hipMemcpy(v, res, sizeof v, hipMemcpyDeviceToHost));
    out->max_stage = v[0];
    out->max_speed = v[1];
    out->min_stage = v[2];
    out->min_speed = v[3];
    out->energy_potential_on_rho = v[4];
    out->energy_kinetic_on_rho = v[5];
    out->energy_total_on_rho = v[6];
    return SWH_OK;
}
