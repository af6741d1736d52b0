// Fused staggered-grid leapfrog step for gfx950.
//
// One launch performs a whole 'linear' or 'leapfrog_linear_plus_nonlinear_friction'
// step of propagation/SWALS/src/shallow_water/domain_mod_linear_solver_include.f90:
// the stage half-step (:92-113), the boundary stage (:122, boundary_mod.f90:78-164),
// and the UH/VH half-step (:211-571) with Coriolis (:408-439), semi-implicit friction
// (:441-505) and linear drag (:550-557), plus the optional max_U update
// (domain_mod.f90:2243-2404).
//
// Layout: each field is a pitch x ny plane (i fastest); the state is double-buffered
// (inputs -> outputs) so Coriolis and friction read pre-step UH/VH exactly as the
// reference's in-place loop does with its thread-boundary caches (:257-333).
//
// Work split: one thread owns one column i and marches up a strip of rows, carrying
// row j in registers and loading row j+1 once. The new stage at (i+1, j) and
// (i, j+1), which the momentum update needs, is recomputed from the old fields rather
// than exchanged, so a launch touches each field once from HBM (neighbour columns hit
// in L1/L2) and needs no barriers or LDS. The stencil has no dense-matrix structure,
// so there is no MFMA here: the kernel is HBM-bound.
#include "swh_internal.h"

namespace swh {

template <typename T>
struct RowVals {
    T s0, s1, c0, c1;   // stage and compensation at i, i+1
    T uhm, uh0, uh1;    // UH at max(i-1,0), i, i+1
    T vh0, vh1;         // VH at i, i+1
    T e0, e1;           // ELV at i, i+1
};

template <typename T, bool KAHAN>
__device__ __forceinline__ void load_row(const LFArgs<T> &a, const T *__restrict__ S, const T *__restrict__ C,
                                         const T *__restrict__ UH, const T *__restrict__ VH, int i, int j,
                                         RowVals<T> &r)
{
    const long long o = (long long)j * a.pitch + i;
    const bool has1 = (i + 1) < a.nx;
    r.s0 = S[o];
    r.uh0 = UH[o];
    r.vh0 = VH[o];
    r.e0 = a.E[o];
    r.uhm = (i > 0) ? UH[o - 1] : r.uh0;
    if (has1) {
        r.s1 = S[o + 1];
        r.uh1 = UH[o + 1];
        r.vh1 = VH[o + 1];
        r.e1 = a.E[o + 1];
    } else {
        r.s1 = r.s0; r.uh1 = r.uh0; r.vh1 = r.vh0; r.e1 = r.e0;
    }
    if (KAHAN) {
        r.c0 = C[o];
        r.c1 = has1 ? C[o + 1] : r.c0;
    } else {
        r.c0 = T(0); r.c1 = T(0);
    }
}

// New stage of cell (ii, jj): interior cells get the flux divergence
// (include.f90:97-113); the outer ring keeps its value unless a stage boundary is
// set on its side (boundary_mod.f90:78-164, staggered grid: stage only).
template <typename T, bool KAHAN>
__device__ __forceinline__ void stage_new(const LFArgs<T> &a, int ii, int jj, T s, T c, T uh, T uhm, T vh, T vhm,
                                         T &sn, T &cn)
{
    const bool interior = ii > 0 && ii < a.nx - 1 && jj > 0 && jj < a.ny - 1;
    if (interior) {
        const T delta = a.rc.sa[jj] * ((uh - uhm) * a.rc.dle[ii] + (vh * a.rc.dbe[jj + 1] - vhm * a.rc.dbe[jj]));
        if (KAHAN) {
            // Kahan-compensated s - delta: the represented value is sn - cn.
            const T y = -delta - c;
            const T t = s + y;
            cn = (t - s) - y;
            sn = t;
        } else {
            sn = s - delta;
            cn = T(0);
        }
    } else {
        const int m = a.bc_mask;
        const bool on_bc = ((m & 8) && ii == 0) || ((m & 2) && ii == a.nx - 1) || ((m & 4) && jj == 0) ||
                           ((m & 1) && jj == a.ny - 1);
        if (on_bc) {
            sn = a.bc_stage;
            cn = T(0);
        } else {
            sn = s;
            cn = c;
        }
    }
}

template <typename T>
__device__ __forceinline__ T dmaxT(T x, T y) { return x > y ? x : y; }
template <typename T>
__device__ __forceinline__ T dminT(T x, T y) { return x < y ? x : y; }

template <typename T, bool KAHAN, bool COR, int FRIC, bool TL, bool LINF>
__global__ __launch_bounds__(256) void leapfrog_kernel(LFArgs<T> a)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.nx) return;
    const int j0 = blockIdx.y * a.rows_per_block;
    if (j0 >= a.ny) return;
    const int j1 = min(j0 + a.rows_per_block, a.ny);
    const long long so = (long long)blockIdx.z * a.src_stride;
    const T *__restrict__ S = a.S + so;
    const T *__restrict__ C = KAHAN ? a.C + so : nullptr;
    const T *__restrict__ UH = a.UH + so;
    const T *__restrict__ VH = a.VH + so;
    T *__restrict__ So = a.So + so;
    T *__restrict__ Co = KAHAN ? a.Co + so : nullptr;
    T *__restrict__ UHo = a.UHo + so;
    T *__restrict__ VHo = a.VHo + so;
    const T MIN = T(kMinDepth);
    const T HALF = T(0.5);
    const T msl = a.msl;

    RowVals<T> r, q;
    load_row<T, KAHAN>(a, S, C, UH, VH, i, j0, r);
    // VH one row below the strip (stage at j0, and the Coriolis 'j-1/2' average,
    // which the reference clamps to row 1 at the southern edge: include.f90:295-308)
    T vhb0 = r.vh0, vhb1 = r.vh1;
    if (j0 > 0) {
        const long long o = (long long)(j0 - 1) * a.pitch + i;
        vhb0 = VH[o];
        vhb1 = (i + 1 < a.nx) ? VH[o + 1] : vhb0;
    }
    T sn0, cn0, sn1, cn1;
    stage_new<T, KAHAN>(a, i, j0, r.s0, r.c0, r.uh0, r.uhm, r.vh0, vhb0, sn0, cn0);
    stage_new<T, KAHAN>(a, i + 1, j0, r.s1, r.c1, r.uh1, r.uh0, r.vh1, vhb1, sn1, cn1);
    T vhm0 = vhb0, vhm1 = vhb1;

    for (int j = j0; j < j1; j++) {
        const bool has_up = (j + 1) < a.ny;
        T su0 = sn0, cu0 = cn0, su1 = sn1, cu1 = cn1;
        if (has_up) {
            load_row<T, KAHAN>(a, S, C, UH, VH, i, j + 1, q);
            stage_new<T, KAHAN>(a, i, j + 1, q.s0, q.c0, q.uh0, q.uhm, q.vh0, r.vh0, su0, cu0);
            stage_new<T, KAHAN>(a, i + 1, j + 1, q.s1, q.c1, q.uh1, q.uh0, q.vh1, r.vh1, su1, cu1);
        } else {
            q = r;
        }
        const long long o = (long long)j * a.pitch + i;
        T uh_new = r.uh0, vh_new = r.vh0;
        if (i < a.nx - 1 && j < a.ny - 1) {
            // Depths at the UH (i+1/2) and VH (j+1/2) points: include.f90:257-345
            T h_iph, h_jph, wet_iph, wet_jph;
            if (TL) {
                h_iph = (r.e0 < -MIN + msl && r.e1 < -MIN + msl) ? msl - HALF * (r.e1 + r.e0) : T(0);
                h_jph = (q.e0 < -MIN + msl && r.e0 < -MIN + msl) ? msl - HALF * (q.e0 + r.e0) : T(0);
                wet_iph = T(1);
                wet_jph = T(1);
            } else {
                h_iph = (sn0 - r.e0 > MIN && sn1 - r.e1 > MIN) ? HALF * ((sn1 + sn0) - (r.e1 + r.e0)) : T(0);
                h_jph = (su0 - q.e0 > MIN && sn0 - r.e0 > MIN) ? HALF * ((su0 + sn0) - (q.e0 + r.e0)) : T(0);
                wet_iph = h_iph > T(0) ? T(1) : T(0);
                wet_jph = h_jph > T(0) ? T(1) : T(0);
            }
            T ds_i = sn1 - sn0, ds_j = su0 - sn0;
            if (KAHAN) {
                ds_i = ds_i - (cn1 - cn0);
                ds_j = ds_j - (cu0 - cn0);
            }
            uh_new = r.uh0 * wet_iph - a.rc.cu[j] * h_iph * ds_i * a.rc.dle[i + 1];
            vh_new = r.vh0 * wet_jph - a.rc.cv[j] * h_jph * ds_j * a.rc.dbe[j + 1];
            if (COR || FRIC) {
                const T vh_jmh = HALF * (vhm0 + vhm1);
                const T vh_jph = HALF * (r.vh0 + r.vh1);
                const T uh_ij = HALF * (r.uh0 + r.uhm);
                const T uh_ijp1 = HALF * (q.uh0 + q.uhm);
                if (COR) {
                    const T dhc = a.rc.fu[j] * (h_iph > T(0) ? T(1) : T(0));
                    const T dhcj = a.rc.fv[j] * (h_jph > T(0) ? T(1) : T(0));
                    uh_new = uh_new + dhc * (vh_jmh + vh_jph);
                    vh_new = vh_new - dhcj * (uh_ij + uh_ijp1);
                }
                if (FRIC) {
                    T fwu, fwv;
                    if (FRIC == 1) {
                        fwu = a.fwU[o];
                        fwv = a.fwV[o];
                    } else {
                        // not-truely-linear: friction_work from the new stage each step
                        // (leapfrog_solver_friction_include.f90:60-82)
                        const int jp1 = min(j + 1, a.ny - 1);
                        const T *__restrict__ ns = a.nsq;
                        const T n0 = ns[o];
                        const T dep_i = dmaxT(HALF * (sn0 - r.e0 + sn1 - r.e1), MIN);
                        const T dep_j = dmaxT(HALF * (sn0 - r.e0 + su0 - q.e0), MIN);
                        const T nsq_i = HALF * (n0 + ns[o + 1]);
                        const T nsq_j = HALF * (n0 + ns[(long long)jp1 * a.pitch + i]);
                        fwu = T(kGravity) * nsq_i * pow(dep_i, a.fric_power);
                        fwv = T(kGravity) * nsq_j * pow(dep_j, a.fric_power);
                    }
                    const T vbar = HALF * (vh_jmh + vh_jph);
                    const T ubar = HALF * (uh_ij + uh_ijp1);
                    const T fmu = T(1) / (T(1) + a.dt_lfc + a.dt * fwu * (sqrt(r.uh0 * r.uh0 + vbar * vbar)));
                    const T fmv = T(1) / (T(1) + a.dt_lfc + a.dt * fwv * (sqrt(r.vh0 * r.vh0 + ubar * ubar)));
                    uh_new = fmu * uh_new;
                    vh_new = fmv * vh_new;
                }
            }
            if (LINF) {
                uh_new = a.lin_fric_mult * uh_new;
                vh_new = a.lin_fric_mult * vh_new;
            }
        }
        So[o] = sn0;
        if (KAHAN) Co[o] = cn0;
        UHo[o] = uh_new;
        VHo[o] = vh_new;

        if (a.maxq_mask) {
            // update_max_quantities on the new state (domain_mod.f90:2243-2404),
            // stage_extrema_include_dry_cells = .true.
            const long long P = (long long)a.pitch * a.ny;
            T *__restrict__ mx = a.mx;  // max_U is tracked for single-source runs only
            const int m = a.maxq_mask;
            if (m & SWH_MAXQ_MAX_STAGE) {
                if (m & SWH_MAXQ_TIME_OF_MAX_STAGE) {
                    if (sn0 > mx[MX_STAGE * P + o]) {
                        mx[MX_STAGE * P + o] = sn0;
                        mx[MX_TOMS * P + o] = a.t_after;
                    }
                } else {
                    mx[MX_STAGE * P + o] = dmaxT(mx[MX_STAGE * P + o], sn0);
                }
            }
            if (m & SWH_MAXQ_MAX_SPEED) {
                const T sE = (i + 1 < a.nx) ? sn1 : sn0, eE = (i + 1 < a.nx) ? r.e1 : r.e0;
                const T sN = has_up ? su0 : sn0, eN = has_up ? q.e0 : r.e0;
                T dE, dN;
                if (TL) {
                    dE = HALF * ((msl - r.e0) + (msl - eE));
                    dN = HALF * ((msl - r.e0) + (msl - eN));
                } else {
                    dE = HALF * ((sn0 - r.e0) + (sE - eE));
                    dN = HALF * ((sn0 - r.e0) + (sN - eN));
                }
                const T iE = dE > MIN ? T(1) / dE : T(0);
                const T iN = dN > MIN ? T(1) / dN : T(0);
                const T sp = sqrt(uh_new * uh_new * (iE * iE) + vh_new * vh_new * (iN * iN));
                mx[MX_SPEED * P + o] = dmaxT(mx[MX_SPEED * P + o], sp);
            }
            if (m & SWH_MAXQ_MAX_FLUX)
                mx[MX_FLUX * P + o] = dmaxT(mx[MX_FLUX * P + o], sqrt(uh_new * uh_new + vh_new * vh_new));
            if (m & SWH_MAXQ_ARRIVAL_TIME) {
                if ((sn0 - r.e0) > MIN && sn0 > a.arrival_stage && mx[MX_ARRIVAL * P + o] < T(0))
                    mx[MX_ARRIVAL * P + o] = a.t_after;
            }
            if (m & SWH_MAXQ_MIN_STAGE) mx[MX_MIN_STAGE * P + o] = dminT(mx[MX_MIN_STAGE * P + o], sn0);
        }

        // slide the window up one row
        vhm0 = r.vh0;
        vhm1 = r.vh1;
        r = q;
        sn0 = su0; cn0 = cu0; sn1 = su1; cn1 = cu1;
    }
}

template <typename T, bool KAHAN, bool COR, int FRIC, bool TL, bool LINF>
static hipError_t launch_lf(const LFArgs<T> &a, int nsrc, int block_x, hipStream_t s)
{
    dim3 block(block_x);
    dim3 grid((a.nx + block_x - 1) / block_x, (a.ny + a.rows_per_block - 1) / a.rows_per_block, nsrc);
    hipLaunchKernelGGL((leapfrog_kernel<T, KAHAN, COR, FRIC, TL, LINF>), grid, block, 0, s, a);
    return hipGetLastError();
}

template <typename T, bool KAHAN, bool COR, int FRIC>
static hipError_t dispatch_tl(const LFArgs<T> &a, int tl, int linf, int nsrc, int bx, hipStream_t s)
{
    if (tl) {
        return linf ? launch_lf<T, KAHAN, COR, FRIC, true, true>(a, nsrc, bx, s)
                    : launch_lf<T, KAHAN, COR, FRIC, true, false>(a, nsrc, bx, s);
    }
    return linf ? launch_lf<T, KAHAN, COR, FRIC, false, true>(a, nsrc, bx, s)
                : launch_lf<T, KAHAN, COR, FRIC, false, false>(a, nsrc, bx, s);
}

template <typename T, bool KAHAN>
hipError_t launch_leapfrog(const LFArgs<T> &a, int coriolis, int friction_mode, int truely_linear,
                           int linear_friction, int nsrc, int block_x, hipStream_t s)
{
    // Linear drag is folded into the friction multiplier when friction is on
    // (include.f90:486-503 vs :550-557).
    const int linf = (friction_mode == 0) && linear_friction;
    if (coriolis) {
        switch (friction_mode) {
        case 0: return dispatch_tl<T, KAHAN, true, 0>(a, truely_linear, linf, nsrc, block_x, s);
        case 1: return dispatch_tl<T, KAHAN, true, 1>(a, truely_linear, linf, nsrc, block_x, s);
        default: return dispatch_tl<T, KAHAN, true, 2>(a, truely_linear, linf, nsrc, block_x, s);
        }
    }
    switch (friction_mode) {
    case 0: return dispatch_tl<T, KAHAN, false, 0>(a, truely_linear, linf, nsrc, block_x, s);
    case 1: return dispatch_tl<T, KAHAN, false, 1>(a, truely_linear, linf, nsrc, block_x, s);
    default: return dispatch_tl<T, KAHAN, false, 2>(a, truely_linear, linf, nsrc, block_x, s);
    }
}

template hipError_t launch_leapfrog<double, false>(const LFArgs<double> &, int, int, int, int, int, int,
                                                    hipStream_t);
template hipError_t launch_leapfrog<float, true>(const LFArgs<float> &, int, int, int, int, int, int,
                                                  hipStream_t);

}  // namespace swh
