// Earthquake events from unit-source gauge series on the FP64 matrix cores.
//
// The reference combines unit-source tsunami time series into events on the host
// (R/examples/austptha_template/SOURCE_ZONES/TEMPLATE/TSUNAMI_EVENTS/
// sum_tsunami_unit_sources.R:555-711, make_tsunami_event_from_unit_sources): for every
// event, the flow series of its unit sources are summed with weight 1 and scaled by
// the event slip (uniform slip), or summed with per-source slips (stochastic slip);
// the stage has msl subtracted before the sum and added back after it. Here the
// series stay where the batched solver wrote them ([slot][sample][3][gauge], FP64)
// and every event of a call is one row of a dense weight matrix W (events x slots),
// so the whole set of events is one product W * X on v_mfma_f64_16x16x4f64, with
// the event's scale/offset and (optionally) the maximum over time of the stage fused
// into the epilogue, which is what a hazard summary keeps per gauge.
#include <algorithm>

#include "swh_internal.h"

namespace swh {

namespace {

typedef double d4 __attribute__((ext_vector_type(4)));

constexpr int kTE = 64;  // events per block tile (2 x 2 waves of 32 x 32)
constexpr int kTN = 64;  // columns per block tile

__device__ __forceinline__ void atomic_max_f64(double *p, double v)
{
    unsigned long long *q = reinterpret_cast<unsigned long long *>(p);
    unsigned long long old = __atomic_load_n(q, __ATOMIC_RELAXED);
    while (__longlong_as_double((long long)old) < v) {
        const unsigned long long assumed = old;
        old = atomicCAS(q, assumed, (unsigned long long)__double_as_longlong(v));
        if (old == assumed) break;
    }
}

// MODE 0: out[e][n] for every column n = (sample * 3 + var) * ng + gauge.
// MODE 1: out[e][g] = max over samples s < nvalid[e] of the event's stage at gauge g
//         (blockIdx.z walks chunks of samples; out starts at -inf).
//
// MFMA operand layout (16 x 16 x 4, FP64): lane l supplies A[l % 16][l / 16] (event
// row, source) and B[l / 16][l % 16] (source, column) and holds D[l / 16 + 4 r][l % 16]
// in accumulator element r (checked on gfx950 with a probe of known operands).
template <int MODE, bool FLAG>
__global__ __launch_bounds__(256) void superpose_kernel(SupArgs a)
{
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int li = lane & 15, lk = lane >> 4;
    const int e_base = blockIdx.y * kTE + (wave >> 1) * 32;
    const int c_base = blockIdx.x * kTN + (wave & 1) * 32;
    const long long N = (long long)a.gmax * 3 * a.ng;
    const int ncols = MODE == 0 ? (int)N : a.ng;
    // this lane's B columns (two 16-column tiles)
    long long caddr[2];
    bool cok[2];
    double csub[2];
    for (int tj = 0; tj < 2; tj++) {
        const int c = c_base + tj * 16 + li;
        cok[tj] = c < ncols;
        caddr[tj] = cok[tj] ? c : 0;
        const int var = MODE == 0 ? (int)((caddr[tj] / a.ng) % 3) : 0;
        csub[tj] = var == 0 ? a.msl : 0.0;   // stage - msl is linear
    }
    // this lane's A rows (W is zero padded to a multiple of kTE rows and 4 columns)
    const double *w0 = a.W + (long long)(e_base + li) * a.Kpad + lk;
    const double *w1 = w0 + 16LL * a.Kpad;
    const double *m0 = FLAG ? a.M + (long long)(e_base + li) * a.Kpad + lk : nullptr;
    const double *m1 = FLAG ? m0 + 16LL * a.Kpad : nullptr;
    // epilogue rows of this lane: e = e_base + ti * 16 + lk + 4 * r
    double scale[2][4], offs[2][2][4];
    int nval[2][4];
    for (int ti = 0; ti < 2; ti++)
        for (int r = 0; r < 4; r++) {
            const int e = e_base + ti * 16 + lk + 4 * r;
            scale[ti][r] = a.scale[e];
            nval[ti][r] = a.nvalid[e];
            for (int tj = 0; tj < 2; tj++) offs[ti][tj][r] = csub[tj];
        }

    const int s_lo = MODE == 1 ? blockIdx.z * a.s_chunk : 0;
    const int s_hi = MODE == 1 ? min(s_lo + a.s_chunk, a.s_max) : 1;
    double mx[2][2][4];
    bool nan_seen[2][2][4];
    for (int ti = 0; ti < 2; ti++)
        for (int tj = 0; tj < 2; tj++)
            for (int r = 0; r < 4; r++) {
                mx[ti][tj][r] = -__builtin_huge_val();
                nan_seen[ti][tj][r] = false;
            }

    for (int s = s_lo; s < s_hi; s++) {
        const long long soff = MODE == 1 ? (long long)s * 3 * a.ng : 0;
        d4 acc[2][2], accf[2][2];
        for (int ti = 0; ti < 2; ti++)
            for (int tj = 0; tj < 2; tj++) {
                acc[ti][tj] = d4{0.0, 0.0, 0.0, 0.0};
                accf[ti][tj] = d4{0.0, 0.0, 0.0, 0.0};
            }
        const double *xk = a.X + (long long)lk * a.slot_stride + soff;
        for (int k0 = 0; k0 < a.Kpad; k0 += 4) {
            const bool kok = k0 + lk < a.nsrc;
            const double a0 = w0[k0], a1 = w1[k0];
            double b[2], bf[2];
            for (int tj = 0; tj < 2; tj++) {
                b[tj] = (kok && cok[tj]) ? xk[caddr[tj]] - csub[tj] : 0.0;
                if (FLAG) {
                    // a non-finite value enters as 0 (so 0 weights of other events stay 0)
                    // and is counted in the membership product instead
                    const bool fin = __builtin_isfinite(b[tj]);
                    bf[tj] = fin ? 0.0 : 1.0;
                    b[tj] = fin ? b[tj] : 0.0;
                }
            }
            acc[0][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, b[0], acc[0][0], 0, 0, 0);
            acc[0][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, b[1], acc[0][1], 0, 0, 0);
            acc[1][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, b[0], acc[1][0], 0, 0, 0);
            acc[1][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, b[1], acc[1][1], 0, 0, 0);
            if (FLAG) {
                const double f0 = m0[k0], f1 = m1[k0];
                accf[0][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(f0, bf[0], accf[0][0], 0, 0, 0);
                accf[0][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(f0, bf[1], accf[0][1], 0, 0, 0);
                accf[1][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(f1, bf[0], accf[1][0], 0, 0, 0);
                accf[1][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(f1, bf[1], accf[1][1], 0, 0, 0);
            }
            xk += 4 * a.slot_stride;
        }
        if (FLAG) {
            // an output that one of the event's non-finite members reaches is NaN, as the
            // reference's sum over the event's own sources gives
            for (int ti = 0; ti < 2; ti++)
                for (int tj = 0; tj < 2; tj++)
                    for (int r = 0; r < 4; r++)
                        if (accf[ti][tj][r] != 0.0) acc[ti][tj][r] = __builtin_nan("");
        }
        if (MODE == 0) {
            for (int ti = 0; ti < 2; ti++)
                for (int r = 0; r < 4; r++) {
                    const int e = e_base + ti * 16 + lk + 4 * r;
                    if (e >= a.E) continue;
                    for (int tj = 0; tj < 2; tj++)
                        if (cok[tj])
                            a.out[(long long)e * N + caddr[tj]] = acc[ti][tj][r] * scale[ti][r] + offs[ti][tj][r];
                }
        } else {
            for (int ti = 0; ti < 2; ti++)
                for (int tj = 0; tj < 2; tj++)
                    for (int r = 0; r < 4; r++)
                        if (s < nval[ti][r]) {
                            const double v = acc[ti][tj][r] * scale[ti][r] + offs[ti][tj][r];
                            mx[ti][tj][r] = v > mx[ti][tj][r] ? v : mx[ti][tj][r];
                            if (FLAG && v != v) nan_seen[ti][tj][r] = true;   // a maximum over NaN is NaN
                        }
        }
    }
    if (MODE == 1) {
        for (int ti = 0; ti < 2; ti++)
            for (int r = 0; r < 4; r++) {
                const int e = e_base + ti * 16 + lk + 4 * r;
                if (e >= a.E) continue;
                for (int tj = 0; tj < 2; tj++) {
                    if (!cok[tj]) continue;
                    double *o = a.out + (long long)e * a.ng + caddr[tj];
                    if (FLAG && nan_seen[ti][tj][r])
                        // NaN stays: atomic_max_f64 never replaces it (NaN < v is false)
                        atomicExch(reinterpret_cast<unsigned long long *>(o),
                                   (unsigned long long)__double_as_longlong(__builtin_nan("")));
                    else if (mx[ti][tj][r] > -__builtin_huge_val())
                        atomic_max_f64(o, mx[ti][tj][r]);
                }
            }
    }
}

__global__ void nonfinite_slots_kernel(const double *__restrict__ X, long long slot_stride, int *flags)
{
    const int slot = blockIdx.y;
    const double *x = X + (long long)slot * slot_stride;
    bool bad = false;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < slot_stride;
         i += (long long)gridDim.x * blockDim.x)
        bad |= !__builtin_isfinite(x[i]);
    if (bad) flags[slot] = 1;   // every writer stores the same value
}

__global__ void fill_kernel(double *p, long long n, double v)
{
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

}  // namespace

int superpose_event_tile() { return kTE; }

hipError_t launch_superpose(const SupArgs &a, int mode, hipStream_t s)
{
    if (a.Kpad % 4 || a.Kpad < a.nsrc || a.E <= 0 || a.ng <= 0 || a.gmax <= 0) return hipErrorInvalidValue;
    const int ety = (a.E + kTE - 1) / kTE;
    if (mode == 0) {
        const long long N = (long long)a.gmax * 3 * a.ng;
        if (N >= (1LL << 31)) return hipErrorInvalidValue;
        if (a.M)
            hipLaunchKernelGGL((superpose_kernel<0, true>), dim3((unsigned)((N + kTN - 1) / kTN), ety), dim3(256), 0,
                               s, a);
        else
            hipLaunchKernelGGL((superpose_kernel<0, false>), dim3((unsigned)((N + kTN - 1) / kTN), ety), dim3(256), 0,
                               s, a);
    } else {
        if (a.s_chunk < 1 || a.s_max < 1 || a.s_max > a.gmax) return hipErrorInvalidValue;
        const long long n = (long long)a.E * a.ng;
        hipLaunchKernelGGL(fill_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, a.out, n,
                           -__builtin_huge_val());
        const int nz = (a.s_max + a.s_chunk - 1) / a.s_chunk;
        if (a.M)
            hipLaunchKernelGGL((superpose_kernel<1, true>), dim3((a.ng + kTN - 1) / kTN, ety, nz), dim3(256), 0, s, a);
        else
            hipLaunchKernelGGL((superpose_kernel<1, false>), dim3((a.ng + kTN - 1) / kTN, ety, nz), dim3(256), 0, s,
                               a);
    }
    return hipGetLastError();
}

hipError_t launch_nonfinite_slots(const double *X, int nsrc, long long slot_stride, int *flags, hipStream_t s)
{
    if (nsrc < 1 || nsrc > 65535 || slot_stride < 1) return hipErrorInvalidValue;
    const long long per = (slot_stride + 255) / 256;
    const unsigned gx = (unsigned)std::min<long long>(per, 64);
    hipLaunchKernelGGL(nonfinite_slots_kernel, dim3(gx, nsrc), dim3(256), 0, s, X, slot_stride, flags);
    return hipGetLastError();
}

}  // namespace swh
