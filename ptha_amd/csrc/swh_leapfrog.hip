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
#include "swh_bflux.h"

namespace swh {

// Vector types for VEC consecutive columns per lane (16-byte loads and stores).
template <typename T, int VEC>
struct VecOf {
    typedef T type __attribute__((ext_vector_type(VEC)));
};
template <typename T>
struct VecOf<T, 1> {
    typedef T type;
};

// NT: non-temporal (streaming) access for the per-source state, which is far larger
// than the caches; elevation and coefficients use ordinary loads.
template <typename T, int VEC, bool NT>
__device__ __forceinline__ void vload(const T *__restrict__ p, T *out)
{
    typedef typename VecOf<T, VEC>::type V;
    V v;
    if (NT) v = __builtin_nontemporal_load(reinterpret_cast<const V *>(p));
    else v = *reinterpret_cast<const V *>(p);
    if constexpr (VEC == 1) {
        out[0] = v;
    } else {
#pragma unroll
        for (int c = 0; c < VEC; c++) out[c] = v[c];
    }
}

template <typename T, bool NT>
__device__ __forceinline__ T sload(const T *__restrict__ p)
{
    if (NT) return __builtin_nontemporal_load(p);
    return *p;
}

template <typename T, int VEC, bool NT>
__device__ __forceinline__ void vstore(T *__restrict__ p, const T *in, int nvalid)
{
    typedef typename VecOf<T, VEC>::type V;
    if (nvalid >= VEC) {
        V v;
        if constexpr (VEC == 1) {
            v = in[0];
        } else {
#pragma unroll
            for (int c = 0; c < VEC; c++) v[c] = in[c];
        }
        if (NT) __builtin_nontemporal_store(v, reinterpret_cast<V *>(p));
        else *reinterpret_cast<V *>(p) = v;
    } else {
#pragma unroll
        for (int c = 0; c < VEC; c++)
            if (c < nvalid) p[c] = in[c];
    }
}

// One row of the fields a lane needs for columns i .. i+VEC-1 and their neighbours:
// s, c, vh, e at columns i .. i+VEC; uh at columns i-1 .. i+VEC (index 0 = i-1).
template <typename T, int VEC>
struct VRow {
    T s[VEC + 1], c[VEC + 1], vh[VEC + 1], e[VEC + 1];
    T uh[VEC + 2];
};

template <typename T, int VEC, bool KAHAN, bool NT>
__device__ __forceinline__ void load_vrow(const LFArgs<T> &a, const T *__restrict__ S, const T *__restrict__ C,
                                          const T *__restrict__ UH, const T *__restrict__ VH, int i, int j,
                                          VRow<T, VEC> &r)
{
    // i is a multiple of VEC and the row pitch a multiple of 64, so the vector stays
    // inside the (padded) row; the right-hand neighbour is read only when it is a cell.
    const long long o = (long long)j * a.pitch + i;
    vload<T, VEC, NT>(S + o, r.s);
    vload<T, VEC, NT>(UH + o, r.uh + 1);
    vload<T, VEC, NT>(VH + o, r.vh);
    vload<T, VEC, false>(a.E + o, r.e);
    if (KAHAN) vload<T, VEC, NT>(C + o, r.c);
    const bool right = (i + VEC) < a.nx;
    r.s[VEC] = right ? sload<T, NT>(S + o + VEC) : T(0);
    r.uh[VEC + 1] = right ? sload<T, NT>(UH + o + VEC) : T(0);
    r.vh[VEC] = right ? sload<T, NT>(VH + o + VEC) : T(0);
    r.e[VEC] = right ? a.E[o + VEC] : T(0);
    if (KAHAN) r.c[VEC] = right ? sload<T, NT>(C + o + VEC) : T(0);
    else {
#pragma unroll
        for (int c = 0; c <= VEC; c++) r.c[c] = T(0);
    }
    // UH at column i-1, clamped to column 0 at the western edge (include.f90:319-322)
    r.uh[0] = (i > 0) ? sload<T, NT>(UH + o - 1) : r.uh[1];
    if (a.fE) {
        // the elevation after this step's forcing (applied after the stage half-step,
        // so the momentum half-step sees it); the plane is updated after the step
#pragma unroll
        for (int c = 0; c < VEC; c++) r.e[c] = r.e[c] + a.f_frac * a.fE[o + c];
        if (right) r.e[VEC] = r.e[VEC] + a.f_frac * a.fE[o + VEC];
    }
}

// New stage of cell (ii, jj): interior cells get the flux divergence
// (include.f90:97-113); the outer ring keeps its value unless a stage boundary is
// set on its side (boundary_mod.f90:78-164, staggered grid: stage only).
template <typename T, bool KAHAN>
__device__ __forceinline__ void stage_new(const LFArgs<T> &a, int ii, int jj, T s, T c, T uh, T uhm, T vh, T vhm,
                                         T &sn, T &cn)
{
    const bool interior = ii > 0 && ii < a.nx - 1 && jj > 0 && jj < a.ny - 1 && !a.stage_given;
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

// New stage of the VEC+1 cells i .. i+VEC of row jj (row values in q, VH of row jj-1 in vhm).
template <typename T, int VEC, bool KAHAN>
__device__ __forceinline__ void stage_row(const LFArgs<T> &a, int i, int jj, const VRow<T, VEC> &q,
                                         const T *vhm, T *sn, T *cn)
{
#pragma unroll
    for (int c = 0; c <= VEC; c++)
        stage_new<T, KAHAN>(a, i + c, jj, q.s[c], q.c[c], q.uh[c + 1], q.uh[c], q.vh[c], vhm[c], sn[c], cn[c]);
    if (a.fS) {
        // apply_forcing after update_boundary (include.f90:117-119; forcing_mod.f90:80-156)
        const long long o = (long long)jj * a.pitch + i;
#pragma unroll
        for (int c = 0; c <= VEC; c++) {
            if (i + c >= a.nx) continue;
            const T inc = a.f_frac * a.fS[o + c];
            if (KAHAN) {
                const T y = inc - cn[c];
                const T t = sn[c] + y;
                cn[c] = (t - sn[c]) - y;
                sn[c] = t;
            } else {
                sn[c] = sn[c] + inc;
            }
        }
    }
}

template <typename T>
__device__ __forceinline__ T dmaxT(T x, T y) { return x > y ? x : y; }
template <typename T>
__device__ __forceinline__ T dminT(T x, T y) { return x < y ? x : y; }

// The fused step. Each lane owns VEC adjacent columns and marches up a strip of
// rows_per_block rows; PF keeps one more row of loads in flight (register prefetch).
template <typename T, int VEC, bool PF, bool NT, bool KAHAN, bool COR, int FRIC, bool TL, bool LINF>
__global__ __launch_bounds__(256) void leapfrog_kernel(LFArgs<T> a)
{
    // single grid: the first a.bf.ey rows of workgroups take the step's boundary-flux
    // sums (of the input UH/VH) and write nothing else
    if (a.bf.ey && (int)blockIdx.y < a.bf.ey) {
        bf_lines<T, 1>(a.bf, a.nx, a.ny, a.pitch, a.UH, a.VH, [](int, int, T &, T &) {});
        return;
    }
    // Block order: with src_fast, consecutive workgroups are the same tile of
    // consecutive sources (they share the elevation tile in L2).
    const int bsrc = a.src_fast ? blockIdx.x : blockIdx.z;
    const int bxt = a.src_fast ? blockIdx.z : blockIdx.x;
    const int i = (bxt * blockDim.x + threadIdx.x) * VEC;
    if (i >= a.nx) return;
    const int j0 = ((int)blockIdx.y - a.bf.ey) * a.rows_per_block;
    if (j0 >= a.ny) return;
    const int j1 = min(j0 + a.rows_per_block, a.ny);
    const long long so = (long long)(bsrc + a.src_base) * a.src_stride;
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
    const int nvalid = min(VEC, a.nx - i);

    VRow<T, VEC> r, q, p;
    load_vrow<T, VEC, KAHAN, NT>(a, S, C, UH, VH, i, j0, r);
    if (j0 + 1 < a.ny) load_vrow<T, VEC, KAHAN, NT>(a, S, C, UH, VH, i, j0 + 1, q);
    // VH one row below the strip (stage at j0, and the Coriolis 'j-1/2' average,
    // which the reference clamps to row 1 at the southern edge: include.f90:295-308)
    T vhm[VEC + 1];
    if (j0 > 0) {
        const long long o = (long long)(j0 - 1) * a.pitch + i;
        vload<T, VEC, NT>(VH + o, vhm);
        vhm[VEC] = (i + VEC < a.nx) ? VH[o + VEC] : T(0);
    } else {
#pragma unroll
        for (int c = 0; c <= VEC; c++) vhm[c] = r.vh[c];
    }
    T sn[VEC + 1], cn[VEC + 1];
    stage_row<T, VEC, KAHAN>(a, i, j0, r, vhm, sn, cn);

    for (int j = j0; j < j1; j++) {
        const bool has_up = (j + 1) < a.ny;
        if (PF && j + 2 < a.ny) load_vrow<T, VEC, KAHAN, NT>(a, S, C, UH, VH, i, j + 2, p);
        if (!PF && has_up && j > j0) load_vrow<T, VEC, KAHAN, NT>(a, S, C, UH, VH, i, j + 1, q);
        T su[VEC + 1], cu[VEC + 1];
        if (has_up) {
            stage_row<T, VEC, KAHAN>(a, i, j + 1, q, r.vh, su, cu);
        } else {
            q = r;
#pragma unroll
            for (int c = 0; c <= VEC; c++) { su[c] = sn[c]; cu[c] = cn[c]; }
        }
        T uh_out[VEC], vh_out[VEC];
#pragma unroll
        for (int c = 0; c < VEC; c++) {
            const int ii = i + c;
            T uh_new = r.uh[c + 1], vh_new = r.vh[c];
            if (ii < a.nx - 1 && j < a.ny - 1) {
                // Depths at the UH (i+1/2) and VH (j+1/2) points: include.f90:257-345
                const T e0 = r.e[c], e1 = r.e[c + 1], eu = q.e[c];
                const T s0 = sn[c], s1 = sn[c + 1], sup = su[c];
                T h_iph, h_jph, wet_iph, wet_jph;
                if (TL) {
                    h_iph = (e0 < -MIN + msl && e1 < -MIN + msl) ? msl - HALF * (e1 + e0) : T(0);
                    h_jph = (eu < -MIN + msl && e0 < -MIN + msl) ? msl - HALF * (eu + e0) : T(0);
                    wet_iph = T(1);
                    wet_jph = T(1);
                } else {
                    h_iph = (s0 - e0 > MIN && s1 - e1 > MIN) ? HALF * ((s1 + s0) - (e1 + e0)) : T(0);
                    h_jph = (sup - eu > MIN && s0 - e0 > MIN) ? HALF * ((sup + s0) - (eu + e0)) : T(0);
                    wet_iph = h_iph > T(0) ? T(1) : T(0);
                    wet_jph = h_jph > T(0) ? T(1) : T(0);
                }
                T ds_i = s1 - s0, ds_j = sup - s0;
                if (KAHAN) {
                    ds_i = ds_i - (cn[c + 1] - cn[c]);
                    ds_j = ds_j - (cu[c] - cn[c]);
                }
                uh_new = r.uh[c + 1] * wet_iph - a.rc.cu[j] * h_iph * ds_i * a.rc.dle[ii + 1];
                vh_new = r.vh[c] * wet_jph - a.rc.cv[j] * h_jph * ds_j * a.rc.dbe[j + 1];
                if (COR || FRIC) {
                    const T vh_jmh = HALF * (vhm[c] + vhm[c + 1]);
                    const T vh_jph = HALF * (r.vh[c] + r.vh[c + 1]);
                    const T uh_ij = HALF * (r.uh[c + 1] + r.uh[c]);
                    const T uh_ijp1 = HALF * (q.uh[c + 1] + q.uh[c]);
                    if (COR) {
                        const T dhc = a.rc.fu[j] * (h_iph > T(0) ? T(1) : T(0));
                        const T dhcj = a.rc.fv[j] * (h_jph > T(0) ? T(1) : T(0));
                        uh_new = uh_new + dhc * (vh_jmh + vh_jph);
                        vh_new = vh_new - dhcj * (uh_ij + uh_ijp1);
                    }
                    if (FRIC) {
                        const long long o = (long long)j * a.pitch + ii;
                        T fwu, fwv;
                        if (FRIC == 1) {
                            fwu = a.fwU[o];
                            fwv = a.fwV[o];
                        } else {
                            // not-truely-linear: friction_work from the new stage each step
                            // (leapfrog_solver_friction_include.f90:60-82)
                            const T *__restrict__ ns = a.nsq;
                            const T n0 = ns[o];
                            const T dep_i = dmaxT(HALF * (s0 - e0 + s1 - e1), MIN);
                            const T dep_j = dmaxT(HALF * (s0 - e0 + sup - eu), MIN);
                            const T nsq_i = HALF * (n0 + ns[o + 1]);
                            const T nsq_j = HALF * (n0 + ns[o + a.pitch]);
                            fwu = T(kGravity) * nsq_i * pow_friction(dep_i, a.fric_power);
                            fwv = T(kGravity) * nsq_j * pow_friction(dep_j, a.fric_power);
                        }
                        const T vbar = HALF * (vh_jmh + vh_jph);
                        const T ubar = HALF * (uh_ij + uh_ijp1);
                        const T ru = r.uh[c + 1], rv = r.vh[c];
                        const T fmu = T(1) / (T(1) + a.dt_lfc + a.dt * fwu * (sqrt(ru * ru + vbar * vbar)));
                        const T fmv = T(1) / (T(1) + a.dt_lfc + a.dt * fwv * (sqrt(rv * rv + ubar * ubar)));
                        uh_new = fmu * uh_new;
                        vh_new = fmv * vh_new;
                    }
                }
                if (LINF) {
                    uh_new = a.lin_fric_mult * uh_new;
                    vh_new = a.lin_fric_mult * vh_new;
                }
            }
            uh_out[c] = uh_new;
            vh_out[c] = vh_new;
        }
        const long long o = (long long)j * a.pitch + i;
        if (!a.stage_given) {
            vstore<T, VEC, NT>(So + o, sn, nvalid);
            if (KAHAN) vstore<T, VEC, NT>(Co + o, cn, nvalid);
        }
        vstore<T, VEC, NT>(UHo + o, uh_out, nvalid);
        vstore<T, VEC, NT>(VHo + o, vh_out, nvalid);

        if (a.maxq_mask) {
            // update_max_quantities on the new state (domain_mod.f90:2243-2404),
            // stage_extrema_include_dry_cells = .true.; single-source runs only
            const long long P = (long long)a.pitch * a.ny;
            T *__restrict__ mx = a.mx;
            const int m = a.maxq_mask;
#pragma unroll
            for (int c = 0; c < VEC; c++) {
                if (c >= nvalid) continue;
                const long long oc = o + c;
                // the compensated stage is sn - cn (as the stand-alone max_U pass reads it)
                const T s0 = KAHAN ? sn[c] - cn[c] : sn[c], e0 = r.e[c], uhn = uh_out[c], vhn = vh_out[c];
                if (m & SWH_MAXQ_MAX_STAGE) {
                    if (m & SWH_MAXQ_TIME_OF_MAX_STAGE) {
                        if (s0 > mx[MX_STAGE * P + oc]) {
                            mx[MX_STAGE * P + oc] = s0;
                            mx[MX_TOMS * P + oc] = a.t_after;
                        }
                    } else {
                        mx[MX_STAGE * P + oc] = dmaxT(mx[MX_STAGE * P + oc], s0);
                    }
                }
                if (m & SWH_MAXQ_MAX_SPEED) {
                    // depth at the UH / VH points: half the sum over the cell and its
                    // east / north neighbour, which on the east column / north row is
                    // the cell alone (sum over i:min(i+1,nx), domain_mod.f90:2311-2320)
                    const bool rt = (i + c + 1) < a.nx;
                    const T sE = rt ? (KAHAN ? sn[c + 1] - cn[c + 1] : sn[c + 1]) : s0, eE = rt ? r.e[c + 1] : e0;
                    const T sN = has_up ? (KAHAN ? su[c] - cu[c] : su[c]) : s0, eN = has_up ? q.e[c] : e0;
                    const T h0 = TL ? (msl - e0) : (s0 - e0);
                    T dE = h0, dN = h0;
                    if (rt) dE = h0 + (TL ? (msl - eE) : (sE - eE));
                    if (has_up) dN = h0 + (TL ? (msl - eN) : (sN - eN));
                    dE = HALF * dE;
                    dN = HALF * dN;
                    const T iE = dE > MIN ? T(1) / dE : T(0);
                    const T iN = dN > MIN ? T(1) / dN : T(0);
                    const T sp = sqrt(uhn * uhn * (iE * iE) + vhn * vhn * (iN * iN));
                    mx[MX_SPEED * P + oc] = dmaxT(mx[MX_SPEED * P + oc], sp);
                }
                if (m & SWH_MAXQ_MAX_FLUX)
                    mx[MX_FLUX * P + oc] = dmaxT(mx[MX_FLUX * P + oc], sqrt(uhn * uhn + vhn * vhn));
                if (m & SWH_MAXQ_ARRIVAL_TIME) {
                    if ((s0 - e0) > MIN && s0 > a.arrival_stage && mx[MX_ARRIVAL * P + oc] < T(0))
                        mx[MX_ARRIVAL * P + oc] = a.t_after;
                }
                if (m & SWH_MAXQ_MIN_STAGE) mx[MX_MIN_STAGE * P + oc] = dminT(mx[MX_MIN_STAGE * P + oc], s0);
            }
        }

        // slide the window up one row
#pragma unroll
        for (int c = 0; c <= VEC; c++) {
            vhm[c] = r.vh[c];
            sn[c] = su[c];
            cn[c] = cu[c];
        }
        r = q;
        if (PF) q = p;
    }
}

// The stage half-step alone (for steps whose boundary must see the new stage before
// the momentum half-step, e.g. the Flather boundary): stage_new on every cell. With
// a.bf.ey, the first ey rows of workgroups take the step's boundary-flux sums from the
// UH/VH this launch reads (for steps in which nothing changes them before the momentum
// half-step).
template <typename T, bool KAHAN>
__global__ __launch_bounds__(256) void lf_stage_kernel(LFArgs<T> a)
{
    if (a.bf.ey && (int)blockIdx.y < a.bf.ey) {
        bf_lines<T, 1>(a.bf, a.nx, a.ny, a.pitch, a.UH, a.VH, [](int, int, T &, T &) {});
        return;
    }
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = (int)blockIdx.y - a.bf.ey;
    if (i >= a.nx || j >= a.ny) return;
    const long long o = (long long)j * a.pitch + i;
    const T s = a.S[o], c = KAHAN ? a.C[o] : T(0);
    const T uh = a.UH[o], uhm = i > 0 ? a.UH[o - 1] : uh;
    const T vh = a.VH[o], vhm = j > 0 ? a.VH[o - a.pitch] : vh;
    T sn, cn;
    stage_new<T, KAHAN>(a, i, j, s, c, uh, uhm, vh, vhm, sn, cn);
    a.So[o] = sn;
    if (KAHAN) a.Co[o] = cn;
}

template <typename T, bool KAHAN>
hipError_t launch_lf_stage(const LFArgs<T> &a, hipStream_t s)
{
    if (a.stage_given || a.src_stride) return hipErrorInvalidValue;
    LFArgs<T> b = a;
    const int gx = (a.nx + 255) / 256;
    b.bf.ey = bf_rows(a.bf, gx);
    if (b.bf.ey < 0 || (b.bf.ey && (!b.bf.part || !b.bf.sc))) return hipErrorInvalidValue;
    hipLaunchKernelGGL((lf_stage_kernel<T, KAHAN>), dim3(gx, a.ny + b.bf.ey), dim3(256), 0, s, b);
    return hipGetLastError();
}
template hipError_t launch_lf_stage<double, false>(const LFArgs<double> &, hipStream_t);
template hipError_t launch_lf_stage<float, true>(const LFArgs<float> &, hipStream_t);

template <typename T, int VEC, bool PF, bool NT, bool KAHAN, bool COR, int FRIC, bool TL, bool LINF>
static hipError_t launch_lf(const LFArgs<T> &a0, int nsrc, int block_x, hipStream_t s)
{
    dim3 block(block_x);
    const int lanes = (a0.nx + VEC - 1) / VEC;
    const int gx = (lanes + block_x - 1) / block_x;
    LFArgs<T> a = a0;
    a.bf.ey = bf_rows(a.bf, gx);
    if (a.bf.ey < 0 || (a.bf.ey && (nsrc != 1 || a.src_fast || a.src_stride != 0 || !a.bf.part || !a.bf.sc)))
        return hipErrorInvalidValue;
    const int gy = (a.ny + a.rows_per_block - 1) / a.rows_per_block + a.bf.ey;
    dim3 grid = a.src_fast ? dim3(nsrc, gy, gx) : dim3(gx, gy, nsrc);
    hipLaunchKernelGGL((leapfrog_kernel<T, VEC, PF, NT, KAHAN, COR, FRIC, TL, LINF>), grid, block, 0, s, a);
    return hipGetLastError();
}

template <typename T, int VEC, bool PF, bool NT, bool KAHAN, bool COR, int FRIC>
static hipError_t dispatch_tl(const LFArgs<T> &a, int tl, int linf, int nsrc, int bx, hipStream_t s)
{
    if (tl) {
        return linf ? launch_lf<T, VEC, PF, NT, KAHAN, COR, FRIC, true, true>(a, nsrc, bx, s)
                    : launch_lf<T, VEC, PF, NT, KAHAN, COR, FRIC, true, false>(a, nsrc, bx, s);
    }
    return linf ? launch_lf<T, VEC, PF, NT, KAHAN, COR, FRIC, false, true>(a, nsrc, bx, s)
                : launch_lf<T, VEC, PF, NT, KAHAN, COR, FRIC, false, false>(a, nsrc, bx, s);
}

template <typename T, int VEC, bool PF, bool NT, bool KAHAN>
static hipError_t dispatch_shape(const LFArgs<T> &a, int coriolis, int friction_mode, int truely_linear, int linf,
                                 int nsrc, int block_x, hipStream_t s)
{
    if (coriolis) {
        switch (friction_mode) {
        case 0: return dispatch_tl<T, VEC, PF, NT, KAHAN, true, 0>(a, truely_linear, linf, nsrc, block_x, s);
        case 1: return dispatch_tl<T, VEC, PF, NT, KAHAN, true, 1>(a, truely_linear, linf, nsrc, block_x, s);
        default: return dispatch_tl<T, VEC, PF, NT, KAHAN, true, 2>(a, truely_linear, linf, nsrc, block_x, s);
        }
    }
    switch (friction_mode) {
    case 0: return dispatch_tl<T, VEC, PF, NT, KAHAN, false, 0>(a, truely_linear, linf, nsrc, block_x, s);
    case 1: return dispatch_tl<T, VEC, PF, NT, KAHAN, false, 1>(a, truely_linear, linf, nsrc, block_x, s);
    default: return dispatch_tl<T, VEC, PF, NT, KAHAN, false, 2>(a, truely_linear, linf, nsrc, block_x, s);
    }
}

// variant bits: 1 = 16-byte vector lanes (4 floats / 2 doubles), 2 = next-row
// prefetch, 4 = non-temporal state traffic.
template <typename T, bool KAHAN>
hipError_t launch_leapfrog(const LFArgs<T> &a, int coriolis, int friction_mode, int truely_linear,
                           int linear_friction, int nsrc, int block_x, int variant, hipStream_t s)
{
    // Linear drag is folded into the friction multiplier when friction is on
    // (include.f90:486-503 vs :550-557).
    const int linf = (friction_mode == 0) && linear_friction;
    constexpr int V = sizeof(T) == 4 ? 4 : 2;
#define SWH_LF_CASE(VV, PF, NT)                                                                              \
    return dispatch_shape<T, VV, PF, NT, KAHAN>(a, coriolis, friction_mode, truely_linear, linf, nsrc, block_x, s)
    switch (variant & 7) {
    case 0: SWH_LF_CASE(1, false, false);
    case 1: SWH_LF_CASE(V, false, false);
    case 2: SWH_LF_CASE(1, true, false);
    case 3: SWH_LF_CASE(V, true, false);
    case 4: SWH_LF_CASE(1, false, true);
    case 5: SWH_LF_CASE(V, false, true);
    case 6: SWH_LF_CASE(1, true, true);
    default: SWH_LF_CASE(V, true, true);
    }
#undef SWH_LF_CASE
}

template hipError_t launch_leapfrog<double, false>(const LFArgs<double> &, int, int, int, int, int, int, int,
                                                    hipStream_t);
template hipError_t launch_leapfrog<float, true>(const LFArgs<float> &, int, int, int, int, int, int, int,
                                                  hipStream_t);

}  // namespace swh
