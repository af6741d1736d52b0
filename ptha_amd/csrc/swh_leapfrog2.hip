// Two leapfrog steps per launch (temporal blocking) for batched unit sources.
//
// Each step is exactly the step of domain_mod_linear_solver_include.f90 (stage
// half-step :92-113, UH/VH half-step :211-571 with Coriolis :408-439, friction
// :441-505 and linear drag :550-557), computed with the same operations in the
// same order as the one-step kernel, so two steps here equal two one-step launches
// bit for bit. What changes is the schedule: the state is read from HBM once and
// written once per two steps, halving the traffic per cell-update.
//
// Work split: a wavefront owns 61 consecutive columns and holds 64 (one halo column
// to the west, two to the east; the westmost lane reads its own west neighbour's
// UH from memory); each lane owns one column and marches up a strip of rows. At
// row m it forms step 1 at row m (stage, UH) and m-1 (VH), then step 2 at row m-1
// (stage, UH) and m-2 (VH), and writes row m-2. East/west neighbours come
// from the adjacent lanes (cross-lane permutes, no LDS tile, no barriers); the
// stencil reaches one column each way per step, which the 3 halo lanes cover.
// Rows below/above the strip are recomputed by the strip itself (3 extra rows).
#include <atomic>

#include "swh_bflux.h"
#include "swh_launch.h"

namespace swh {

namespace {

constexpr int kOwn = 61;   // columns a wavefront writes (lanes 1..61)

// Lane shifts across the wavefront (the lane at the open end keeps its own value).
// Call them on every lane, outside lane-dependent conditions: a lane that does not
// execute the shuffle leaves the lane reading from it with an undefined value.
template <typename T>
__device__ __forceinline__ T shl1(T v) { return __shfl_down(v, 1); }   // value of lane + 1
template <typename T>
__device__ __forceinline__ T shr1(T v) { return __shfl_up(v, 1); }     // value of lane - 1

// Per-row coefficients (uniform across the wavefront), loaded at a clamped row so
// the loads are always in bounds; 'interior' / 'valid' say whether the row updates.
template <typename T>
struct RowCo {
    T sa, cu, cv, fu, fv, dbe_lo, dbe_hi;   // dbe_lo = dbe(j), dbe_hi = dbe(j+1)
    bool interior, valid;                   // stage row (1..ny-2), momentum row (0..ny-2)
};

// A value that is the same on every lane (a row coefficient), kept in scalar registers.
__device__ __forceinline__ float uniform(float v)
{
    return __builtin_bit_cast(float, __builtin_amdgcn_readfirstlane(__builtin_bit_cast(int, v)));
}
__device__ __forceinline__ double uniform(double v)
{
    const long long b = __builtin_bit_cast(long long, v);
    const int lo = __builtin_amdgcn_readfirstlane((int)(b & 0xffffffff));
    const int hi = __builtin_amdgcn_readfirstlane((int)(b >> 32));
    return __builtin_bit_cast(double, ((long long)hi << 32) | (unsigned int)lo);
}

template <typename T>
__device__ __forceinline__ RowCo<T> row_co(const LFArgs<T> &a, int r)
{
    const int rc = min(max(r, 0), a.ny - 1);
    RowCo<T> q;
    q.sa = a.rc.sa[rc];
    q.cu = a.rc.cu[rc];
    q.cv = a.rc.cv[rc];
    q.fu = a.rc.fu[rc];
    q.fv = a.rc.fv[rc];
    q.dbe_lo = a.rc.dbe[rc];
    q.dbe_hi = a.rc.dbe[rc + 1];
    q.interior = r > 0 && r < a.ny - 1;
    q.valid = r >= 0 && r < a.ny - 1;
    return q;
}

// row_co for a row that is the same on every lane of the wavefront (the fused n-step
// kernel's rows), FP64: the coefficients in scalar registers, so that the rows carried
// across iterations take no vector registers (150 -> 105 VGPRs at three steps, a fourth
// wavefront per SIMD, FP64 batch +3 %). FP32 keeps them as the compiler places them: its
// loop is issue-bound and the scalar-register spills this adds cost more than the
// wavefront it gains (-7 %). Not for per-lane rows (step1_cell).
template <typename T>
__device__ __forceinline__ RowCo<T> row_co_u(const LFArgs<T> &a, int r)
{
    RowCo<T> q = row_co(a, r);
    if (sizeof(T) == 4) return q;
    q.sa = uniform(q.sa);
    q.cu = uniform(q.cu);
    q.cv = uniform(q.cv);
    q.fu = uniform(q.fu);
    q.fv = uniform(q.fv);
    q.dbe_lo = uniform(q.dbe_lo);
    q.dbe_hi = uniform(q.dbe_hi);
    return q;
}

// New stage (include.f90:97-113), computed on every lane and kept where the cell
// is interior; no stage boundary in this kernel.
template <typename T, bool KAHAN>
__device__ __forceinline__ void stage_cell(bool interior, T sa, T dle_i, T dbe_lo, T dbe_hi, T s, T c, T uh, T uhm,
                                           T vh, T vhm, T &sn, T &cn)
{
    const T delta = sa * ((uh - uhm) * dle_i + (vh * dbe_hi - vhm * dbe_lo));
    T t, cc;
    if (KAHAN) {
        const T y = -delta - c;
        t = s + y;
        cc = (t - s) - y;
    } else {
        t = s - delta;
        cc = T(0);
    }
    sn = interior ? t : s;
    cn = interior ? cc : c;
}

// Depth at a face of the truely-linear scheme (include.f90:265-280) from the elevations
// of its two cells (e_far: the east or north one); it depends on the elevation alone.
template <typename T>
__device__ __forceinline__ T tl_face_depth(T e_near, T e_far, T msl)
{
    const T MIN = T(kMinDepth), HALF = T(0.5);
    return (e_near < -MIN + msl && e_far < -MIN + msl) ? msl - HALF * (e_far + e_near) : T(0);
}

// UH at (ii + 1/2, j) from the new stage of cells ii, ii+1 and the pre-step VH of
// rows j-1 (vhm0, vhm1 at ii, ii+1) and j (vh0, vh1). HPRE (truely linear only): e0
// carries the face depth h_iph itself (tl_face_depth of the two elevations), e1 unused.
template <typename T, bool KAHAN, bool COR, int FRIC, bool TL, bool LINF, bool HPRE = false>
__device__ __forceinline__ T mom_u(const LFArgs<T> &a, bool valid, T cu, T fu, T dle_ip1, T fwu, T uh_old, T s0,
                                   T c0, T s1, T c1, T e0, T e1, T vhm0, T vhm1, T vh0, T vh1)
{
    static_assert(TL || !HPRE, "precomputed face depths are for the truely-linear scheme");
    const T MIN = T(kMinDepth), HALF = T(0.5), msl = a.msl;
    T h_iph, wet_iph;
    if (TL) {
        h_iph = HPRE ? e0 : tl_face_depth(e0, e1, msl);
        wet_iph = T(1);
    } else {
        h_iph = (s0 - e0 > MIN && s1 - e1 > MIN) ? HALF * ((s1 + s0) - (e1 + e0)) : T(0);
        wet_iph = h_iph > T(0) ? T(1) : T(0);
    }
    T ds_i = s1 - s0;
    if (KAHAN) ds_i = ds_i - (c1 - c0);
    T uh_new = uh_old * wet_iph - cu * h_iph * ds_i * dle_ip1;
    if (COR || FRIC) {
        const T vh_jmh = HALF * (vhm0 + vhm1);
        const T vh_jph = HALF * (vh0 + vh1);
        if (COR) {
            const T dhc = fu * (h_iph > T(0) ? T(1) : T(0));
            uh_new = uh_new + dhc * (vh_jmh + vh_jph);
        }
        if (FRIC) {
            const T vbar = HALF * (vh_jmh + vh_jph);
            const T fmu = T(1) / (T(1) + a.dt_lfc + a.dt * fwu * (sqrt(uh_old * uh_old + vbar * vbar)));
            uh_new = fmu * uh_new;
        }
    }
    if (LINF) uh_new = a.lin_fric_mult * uh_new;
    return valid ? uh_new : uh_old;
}

// VH at (ii, j + 1/2) from the new stage of rows j, j+1 and the pre-step UH of
// row j (uh_i at ii, uh_im1 at ii-1) and row j+1 (uhu_i, uhu_im1). HPRE: e0 carries
// h_jph itself, eu unused.
template <typename T, bool KAHAN, bool COR, int FRIC, bool TL, bool LINF, bool HPRE = false>
__device__ __forceinline__ T mom_v(const LFArgs<T> &a, bool valid, T cv, T fv, T dbe_hi, T fwv, T vh_old, T s0, T c0,
                                   T sup, T cup, T e0, T eu, T uh_i, T uh_im1, T uhu_i, T uhu_im1)
{
    static_assert(TL || !HPRE, "precomputed face depths are for the truely-linear scheme");
    const T MIN = T(kMinDepth), HALF = T(0.5), msl = a.msl;
    T h_jph, wet_jph;
    if (TL) {
        h_jph = HPRE ? e0 : tl_face_depth(e0, eu, msl);
        wet_jph = T(1);
    } else {
        h_jph = (sup - eu > MIN && s0 - e0 > MIN) ? HALF * ((sup + s0) - (eu + e0)) : T(0);
        wet_jph = h_jph > T(0) ? T(1) : T(0);
    }
    T ds_j = sup - s0;
    if (KAHAN) ds_j = ds_j - (cup - c0);
    T vh_new = vh_old * wet_jph - cv * h_jph * ds_j * dbe_hi;
    if (COR || FRIC) {
        const T uh_ij = HALF * (uh_i + uh_im1);
        const T uh_ijp1 = HALF * (uhu_i + uhu_im1);
        if (COR) {
            const T dhcj = fv * (h_jph > T(0) ? T(1) : T(0));
            vh_new = vh_new - dhcj * (uh_ij + uh_ijp1);
        }
        if (FRIC) {
            const T ubar = HALF * (uh_ij + uh_ijp1);
            const T fmv = T(1) / (T(1) + a.dt_lfc + a.dt * fwv * (sqrt(vh_old * vh_old + ubar * ubar)));
            vh_new = fmv * vh_new;
        }
    }
    if (LINF) vh_new = a.lin_fric_mult * vh_new;
    return valid ? vh_new : vh_old;
}

}  // namespace

// The tile (column block, strip, source) of this workgroup. Workgroups are dispatched
// to the 8 XCDs (each with its own L2) round-robin by linear index; in XCD order the
// i-th workgroup on an XCD takes the i-th tile of that XCD's contiguous range of tiles,
// ordered columns fastest, then the sources of a group of 'group' sources, then
// strips, then groups. Neighbouring tiles of a strip, the sources of a group (which
// read the same elevation and row coefficients) and the next strip (which reads the
// halo rows of this one) thus run on the same L2 shortly after one another.
// ey: rows of workgroups ahead of the tiles that do other work (single grid only).
__device__ __forceinline__ void lf2_tile(int order, int group, int ey, int &bx, int &by, int &bz)
{
    bx = blockIdx.x; by = (int)blockIdx.y - ey; bz = blockIdx.z;
    if (!order) return;
    const unsigned gx = gridDim.x, gy = gridDim.y - ey, gz = gridDim.z;
    const unsigned B = gx * gy * gz;
    const unsigned b = blockIdx.x + gx * ((blockIdx.y - ey) + gy * blockIdx.z);
    const unsigned xcd = b % 8u, slot = b / 8u, n8 = B / 8u, rem = B % 8u;
    const unsigned L = xcd < rem ? xcd * (n8 + 1u) + slot : rem * (n8 + 1u) + (xcd - rem) * n8 + slot;
    const unsigned G = (unsigned)group;
    const unsigned gi = L / (gx * gy * G);                 // source group
    const unsigned within = L - gi * (gx * gy * G);
    const unsigned Gi = min(G, gz - gi * G);               // sources in this group (the last may be short)
    const unsigned t = within / gx;
    bx = (int)(within - t * gx);
    bz = (int)(gi * G + t % Gi);
    by = (int)(t / Gi);
}

template <typename T, bool KAHAN, bool COR, int FRIC, bool TL, bool LINF>
__device__ __forceinline__ void step1_cell(const LFArgs<T> &a, const T *__restrict__ S, const T *__restrict__ C,
                                           const T *__restrict__ UH, const T *__restrict__ VH, int i, int j, T &s_c,
                                           T &c_c, T &uh1, T &vh1);

// PF: the loads of the next row are issued before the current row is computed, so
// their latency overlaps the arithmetic (one row of loads in flight per lane).
// a.bf.ey > 0 (single grid): the first a.bf.ey rows of workgroups take the boundary-flux
// sums of both steps (the second from UH/VH one step ahead on the lines) and write
// nothing else.
template <typename T, bool KAHAN, bool COR, int FRIC, bool TL, bool LINF, bool PF>
__global__ __launch_bounds__(256) void leapfrog2_kernel(LFArgs<T> a)
{
    if (a.bf.ey && (int)blockIdx.y < a.bf.ey) {
        bf_lines<T, 2>(a.bf, a.nx, a.ny, a.pitch, a.UH, a.VH, [&](int i, int j, T &uh1, T &vh1) {
            T s_c, c_c;
            step1_cell<T, KAHAN, COR, FRIC, TL, LINF>(a, a.S, a.C, a.UH, a.VH, i, j, s_c, c_c, uh1, vh1);
        });
        return;
    }
    int bx, by, bz;
    lf2_tile(a.lf2_order, a.lf2_group, a.bf.ey, bx, by, bz);
    const int lane = threadIdx.x & 63;
    const int wg = bx * (blockDim.x >> 6) + (threadIdx.x >> 6);
    if (wg * kOwn >= a.nx) return;                       // whole wavefront past the east edge
    const int col = wg * kOwn - 1 + lane;                // column of this lane (may be -1 or >= nx)
    const int ci = min(max(col, 0), a.nx - 1);           // column it loads (clamped into the row)
    const bool own = lane >= 1 && lane <= kOwn && col < a.nx;
    const int j0 = by * a.rows_per_block;
    if (j0 >= a.ny) return;
    const int j1 = min(j0 + a.rows_per_block, a.ny);
    const long long so = (long long)(bz + a.src_base) * a.src_stride;
    const T *__restrict__ S = a.S + so;
    const T *__restrict__ C = KAHAN ? a.C + so : nullptr;
    const T *__restrict__ UH = a.UH + so;
    const T *__restrict__ VH = a.VH + so;
    T *__restrict__ So = a.So + so;
    T *__restrict__ Co = KAHAN ? a.Co + so : nullptr;
    T *__restrict__ UHo = a.UHo + so;
    T *__restrict__ VHo = a.VHo + so;
    const bool west = col == 0;                          // UH(i-1) clamps to UH(0) (include.f90:319-322)
    // per-lane constants: edge lengths and where the column updates
    const T dle_i = a.rc.dle[min(max(col, 0), a.nx)];
    const T dle_ip1 = a.rc.dle[min(max(col + 1, 0), a.nx)];
    const bool col_int = col > 0 && col < a.nx - 1;      // stage: interior column
    const bool col_mom = col >= 0 && col < a.nx - 1;     // momentum: UH/VH of this column update
    const int cf = min(max(col, 0), a.nx - 1);           // friction-work column

    // registers carried between rows (named by the row they belong to, relative to m)
    T vh0_p = T(0), vh0_p_r = T(0), uh0_p = T(0), uh0_p_l = T(0), e_p = T(0), e_p_r = T(0), e_pp = T(0);
    T s1_p = T(0), c1_p = T(0);
    T uh1_p = T(0), uh1_p_l = T(0), uh1_pp = T(0), uh1_pp_l = T(0);
    T vh1_pp = T(0), vh1_pp_r = T(0);
    T s2_p = T(0), c2_p = T(0), uh2_p = T(0);
    RowCo<T> q_p = row_co(a, j0 - 2), q_pp = row_co(a, j0 - 3);
    T fw_u_p = T(0), fw_v_p = T(0);
    {
        // pre-step VH/UH/E of row j0-2 (feeds only rows the strip does not write)
        const int r = min(max(j0 - 2, 0), a.ny - 1);
        const int o = r * a.pitch + ci;
        vh0_p = VH[o];
        uh0_p = UH[o];
        e_p = a.E[o];
        vh0_p_r = shl1(vh0_p);
        // every lane takes part in a shuffle (a lane that skips it leaves its neighbour
        // reading an inactive lane), then the western edge keeps its own value
        const T uh0_p_sh = shr1(uh0_p);
        uh0_p_l = west ? uh0_p : uh0_p_sh;
        e_p_r = shl1(e_p);
    }

    // the loads of one row (lane 0 also reads its own west neighbour's UH)
    auto load_row = [&](int m, T &s0, T &c0, T &uh0, T &vh0, T &e, T &uhw) {
        const int mr = min(max(m, 0), a.ny - 1);
        const int o = mr * a.pitch + ci;
        s0 = S[o];
        c0 = KAHAN ? C[o] : T(0);
        uh0 = UH[o];
        vh0 = VH[o];
        e = a.E[o];
        uhw = (lane == 0 && col > 0) ? UH[o - 1] : uh0;
    };
    T n_s0, n_c0, n_uh0, n_vh0, n_e, n_uhw;
    if (PF) load_row(j0 - 1, n_s0, n_c0, n_uh0, n_vh0, n_e, n_uhw);

    // offsets within one source's plane fit in 32 bits (uniform base + per-lane offset)
    for (int m = j0 - 1; m <= j1 + 1; m++) {
        const int mr = min(max(m, 0), a.ny - 1);
        T s0, c0, uh0, vh0, e, uh0_mem;
        if (PF) {
            s0 = n_s0; c0 = n_c0; uh0 = n_uh0; vh0 = n_vh0; e = n_e; uh0_mem = n_uhw;
            if (m + 1 <= j1 + 1) load_row(m + 1, n_s0, n_c0, n_uh0, n_vh0, n_e, n_uhw);
        } else {
            load_row(m, s0, c0, uh0, vh0, e, uh0_mem);
        }
        const RowCo<T> q = row_co(a, m);
        // UH(i-1): from the lane to the west; lane 0 has none and reads it itself
        const T uh0_sh = shr1(uh0);
        const T uh0_l = west ? uh0 : (lane == 0 ? uh0_mem : uh0_sh);
        const T vh0_r = shl1(vh0), e_r = shl1(e);
        // VH of the row below, clamped to row 0 at the southern edge (include.f90:295-308)
        const T vhb = m > 0 ? vh0_p : vh0, vhb_r = m > 0 ? vh0_p_r : vh0_r;
        // friction work: UH at row m (and m-1 from the last row), VH at row m-1 (and m-2)
        T fw_u1 = T(0), fw_v1 = T(0);
        if (FRIC) {
            fw_u1 = a.fwU[mr * a.pitch + cf];
            fw_v1 = a.fwV[min(max(m - 1, 0), a.ny - 1) * a.pitch + cf];
        }
        const T fw_u2 = fw_u_p, fw_v2 = fw_v_p;

        // step 1: stage and UH at row m
        T s1, c1;
        stage_cell<T, KAHAN>(col_int && q.interior, q.sa, dle_i, q.dbe_lo, q.dbe_hi, s0, c0, uh0, uh0_l, vh0, vhb, s1,
                             c1);
        const T s1_r = shl1(s1);
        const T c1_r = KAHAN ? shl1(c1) : T(0);
        const T uh1 = mom_u<T, KAHAN, COR, FRIC, TL, LINF>(a, col_mom && q.valid, q.cu, q.fu, dle_ip1, fw_u1, uh0, s1,
                                                            c1, s1_r, c1_r, e, e_r, vhb, vhb_r, vh0, vh0_r);
        const T uh1_sh = shr1(uh1);
        const T uh1_l = west ? uh1 : uh1_sh;
        // step 1: VH at row m-1
        const T vh1 = mom_v<T, KAHAN, COR, FRIC, TL, LINF>(a, col_mom && q_p.valid, q_p.cv, q_p.fv, q_p.dbe_hi, fw_v1,
                                                            vh0_p, s1_p, c1_p, s1, c1, e_p, e, uh0_p, uh0_p_l, uh0,
                                                            uh0_l);
        const T vh1_r = shl1(vh1);
        // step 2: stage and UH at row m-1 (VH of row m-2 clamped to row 0 at the southern edge)
        const T vhb1 = m - 1 > 0 ? vh1_pp : vh1, vhb1_r = m - 1 > 0 ? vh1_pp_r : vh1_r;
        T s2, c2;
        stage_cell<T, KAHAN>(col_int && q_p.interior, q_p.sa, dle_i, q_p.dbe_lo, q_p.dbe_hi, s1_p, c1_p, uh1_p,
                             uh1_p_l, vh1, vhb1, s2, c2);
        const T s2_r = shl1(s2);
        const T c2_r = KAHAN ? shl1(c2) : T(0);
        const T uh2 = mom_u<T, KAHAN, COR, FRIC, TL, LINF>(a, col_mom && q_p.valid, q_p.cu, q_p.fu, dle_ip1, fw_u2,
                                                            uh1_p, s2, c2, s2_r, c2_r, e_p, e_p_r, vhb1, vhb1_r, vh1,
                                                            vh1_r);
        // step 2: VH at row m-2
        const T vh2 = mom_v<T, KAHAN, COR, FRIC, TL, LINF>(a, col_mom && q_pp.valid, q_pp.cv, q_pp.fv, q_pp.dbe_hi,
                                                            fw_v2, vh1_pp, s2_p, c2_p, s2, c2, e_pp, e_p, uh1_pp,
                                                            uh1_pp_l, uh1_p, uh1_p_l);
        const int k = m - 2;
        if (own && k >= j0 && k < j1) {
            const int ow = k * a.pitch + col;
            So[ow] = s2_p;
            if (KAHAN) Co[ow] = c2_p;
            UHo[ow] = uh2_p;
            VHo[ow] = vh2;
        }
        // slide up one row
        s2_p = s2; c2_p = c2; uh2_p = uh2;
        uh1_pp = uh1_p; uh1_pp_l = uh1_p_l; uh1_p = uh1; uh1_p_l = uh1_l;
        vh1_pp = vh1; vh1_pp_r = vh1_r;
        s1_p = s1; c1_p = c1;
        vh0_p = vh0; vh0_p_r = vh0_r; uh0_p = uh0; uh0_p_l = uh0_l;
        e_pp = e_p; e_p = e; e_p_r = e_r;
        q_pp = q_p; q_p = q;
        fw_u_p = fw_u1; fw_v_p = fw_v1;
    }
}

// NS fused steps per launch (NS = 3: the state is read and written once per three
// steps). The same per-cell operations as the one-step kernel, scheduled as in
// leapfrog2_kernel but for any number of levels: at row m, level n (the state after n
// steps) forms its stage and UH at row m-n+1 and its VH at row m-n from level n-1,
// whose stage and UH of row m-n+1 were formed in the previous iteration (level 0: the
// row loaded now) and whose VH of rows m-n+1 (this iteration) and m-n (the previous
// one) are at hand; row m-NS of level NS is written. Each level consumes one lane
// of the wavefront to the west (UH(i-1)) and one to the east (stage(i+1), VH(i+1)), so a
// wavefront owns 65 - 2 NS of its 64 columns; each strip recomputes NS rows below and
// NS - 1 above it.
template <typename T, bool KAHAN, bool COR, int FRIC, bool TL, bool LINF, bool PF, int NS, bool GS = false>
__global__ __launch_bounds__(256, (GS && sizeof(T) == 4) ? 4 : 1) void leapfrogn_kernel(LFArgs<T> a)
{
    constexpr int OWN = 65 - 2 * NS;
    int bx, by, bz;
    lf2_tile(a.lf2_order, a.lf2_group, 0, bx, by, bz);
    const int lane = threadIdx.x & 63;
    const int wg = bx * (blockDim.x >> 6) + (threadIdx.x >> 6);
    if (wg * OWN >= a.nx) return;                        // whole wavefront past the east edge
    const int col = wg * OWN - (NS - 1) + lane;          // column of this lane (may be < 0 or >= nx)
    const int ci = min(max(col, 0), a.nx - 1);           // column it loads (clamped into the row)
    const bool own = lane >= NS - 1 && lane < NS - 1 + OWN && col < a.nx;
    const int j0 = by * a.rows_per_block;
    if (j0 >= a.ny) return;
    const int j1 = min(j0 + a.rows_per_block, a.ny);
    const long long so = (long long)(bz + a.src_base) * a.src_stride;
    const T *__restrict__ S = a.S + so;
    const T *__restrict__ C = KAHAN ? a.C + so : nullptr;
    const T *__restrict__ UH = a.UH + so;
    const T *__restrict__ VH = a.VH + so;
    T *__restrict__ So = a.So + so;
    T *__restrict__ Co = KAHAN ? a.Co + so : nullptr;
    T *__restrict__ UHo = a.UHo + so;
    T *__restrict__ VHo = a.VHo + so;
    const bool west = col == 0;                          // UH(i-1) clamps to UH(0) (include.f90:319-322)
    const T dle_i = a.rc.dle[min(max(col, 0), a.nx)];
    const T dle_ip1 = a.rc.dle[min(max(col + 1, 0), a.nx)];
    const bool col_int = col > 0 && col < a.nx - 1;      // stage: interior column
    const bool col_mom = col >= 0 && col < a.nx - 1;     // momentum: UH/VH of this column update
    const int cf = min(max(col, 0), a.nx - 1);           // friction-work column
    // the gauges in this lane's own column and rows j0..j1-1 (the host sees that there are
    // at most two): indices + 1 in bytes 0 and 1 of gk (0: none), rows in the halves of gr
    int gk = 0, gr = 0;
    if (GS && own) {
        for (int g = 0; g < a.g_n; g++) {
            const int gi = a.g_ij[2 * g], gj = a.g_ij[2 * g + 1];
            if (gi == col && gj >= j0 && gj < j1) {
                if (gk == 0) { gk = g + 1; gr = gj; }
                else { gk |= (g + 1) << 8; gr |= gj << 16; }
            }
        }
    }

    // carried between rows, per level n (0 = the loaded state)
    T s_p[NS + 1], c_p[NS + 1];          // stage of level n at its stage row of the previous iteration
    T uh_p[NS + 1], uh_p_l[NS + 1];      // UH of level n, previous iteration (level 0: row m-1), and its west value
    T uh_pp[NS + 1], uh_pp_l[NS + 1];    // the iteration before that (levels >= 1)
    T vh_p[NS + 1], vh_p_r[NS + 1];      // VH of level n, previous iteration, and its east value
    T e_r[NS + 1], e_rr[NS + 1];         // elevation of row m-k and its east value
    // truely linear: the face depths instead (they depend on the elevation alone, so each
    // is formed once, when its row is loaded, not once per level): h_iph of row m-k,
    // h_jph of the face above row m-1-k, and the elevation of row m-1
    T hi_r[NS], hj_r[NS], e_prev = T(0);
    T fwu_r[NS], fwv_r[NS];              // friction work: UH at row m-k, VH at row m-1-k
    RowCo<T> q_r[NS + 1];                // row coefficients of row m-k
#pragma unroll
    for (int n = 0; n <= NS; n++) {
        s_p[n] = T(0); c_p[n] = T(0); uh_p[n] = T(0); uh_p_l[n] = T(0); uh_pp[n] = T(0); uh_pp_l[n] = T(0);
        vh_p[n] = T(0); vh_p_r[n] = T(0); e_r[n] = T(0); e_rr[n] = T(0);
    }
#pragma unroll
    for (int n = 0; n < NS; n++) { hi_r[n] = T(0); hj_r[n] = T(0); }
#pragma unroll
    for (int n = 0; n < NS; n++) { fwu_r[n] = T(0); fwv_r[n] = T(0); }
    const int m0 = j0 - NS + 1;
    // the rings as they stand after row m0-1 (each iteration first slides them by one)
#pragma unroll
    for (int k = 0; k <= NS; k++) q_r[k] = row_co_u(a, m0 - 1 - k);
    {
        // the row below the first one: pre-step VH/UH/E of row m0-1
        const int r = min(max(m0 - 1, 0), a.ny - 1);
        const int o = r * a.pitch + ci;
        vh_p[0] = VH[o];
        uh_p[0] = UH[o];
        e_r[0] = a.E[o];
        vh_p_r[0] = shl1(vh_p[0]);
        const T uh_sh = shr1(uh_p[0]);
        uh_p_l[0] = west ? uh_p[0] : uh_sh;
        e_rr[0] = shl1(e_r[0]);
        e_prev = e_r[0];
    }
    auto load_row = [&](int m, T &s0, T &c0, T &uh0, T &vh0, T &e, T &uhw) {
        const int mr = min(max(m, 0), a.ny - 1);
        const int o = mr * a.pitch + ci;
        s0 = S[o];
        c0 = KAHAN ? C[o] : T(0);
        uh0 = UH[o];
        vh0 = VH[o];
        e = a.E[o];
        uhw = (lane == 0 && col > 0) ? UH[o - 1] : uh0;
    };
    T n_s0, n_c0, n_uh0, n_vh0, n_e, n_uhw;
    if (PF) load_row(m0, n_s0, n_c0, n_uh0, n_vh0, n_e, n_uhw);
    T cap_s = T(0), cap_c = T(0), cap_uh = T(0), cap_vh = T(0);   // the sampled level (gauges)

    for (int m = m0; m <= j1 + NS - 1; m++) {
        T s0, c0, uh0, vh0, e0, uh0_mem;
        if (PF) {
            s0 = n_s0; c0 = n_c0; uh0 = n_uh0; vh0 = n_vh0; e0 = n_e; uh0_mem = n_uhw;
            if (m + 1 <= j1 + NS - 1) load_row(m + 1, n_s0, n_c0, n_uh0, n_vh0, n_e, n_uhw);
        } else {
            load_row(m, s0, c0, uh0, vh0, e0, uh0_mem);
        }
        // slide the rings of rows: index k now holds row m-k
#pragma unroll
        for (int k = NS; k >= 1; k--) q_r[k] = q_r[k - 1];
        q_r[0] = row_co_u(a, m);
        if constexpr (TL) {
#pragma unroll
            for (int k = NS - 1; k >= 1; k--) { hi_r[k] = hi_r[k - 1]; hj_r[k] = hj_r[k - 1]; }
            hi_r[0] = tl_face_depth(e0, shl1(e0), a.msl);
            hj_r[0] = tl_face_depth(e_prev, e0, a.msl);
            e_prev = e0;
        } else {
#pragma unroll
            for (int k = NS; k >= 1; k--) { e_r[k] = e_r[k - 1]; e_rr[k] = e_rr[k - 1]; }
            e_r[0] = e0;
            e_rr[0] = shl1(e0);
        }
        if (FRIC) {
#pragma unroll
            for (int k = NS - 1; k >= 1; k--) { fwu_r[k] = fwu_r[k - 1]; fwv_r[k] = fwv_r[k - 1]; }
            const int mr = min(max(m, 0), a.ny - 1);
            fwu_r[0] = a.fwU[mr * a.pitch + cf];
            fwv_r[0] = a.fwV[min(max(m - 1, 0), a.ny - 1) * a.pitch + cf];
        }
        // level 0 at row m (the loaded row); UH(i-1) from the west lane, lane 0 from memory
        const T uh0_sh = shr1(uh0);
        const T uh0_l = west ? uh0 : (lane == 0 ? uh0_mem : uh0_sh);
        // inputs of the level being formed: stage/UH of the level below at this level's
        // stage row, its VH there (cur) and one row below (below)
        T in_s = s0, in_c = c0, in_uh = uh0, in_uh_l = uh0_l;
        T in_uhb = uh_p[0], in_uhb_l = uh_p_l[0];     // UH of the level below one row lower
        T vh_cur = vh0, vh_cur_r = shl1(vh0);
        T out_s = T(0), out_c = T(0), out_uh = T(0), out_vh = T(0);
#pragma unroll
        for (int n = 1; n <= NS; n++) {
            const int r = m - n + 1;                  // this level's stage/UH row
            const RowCo<T> &q = q_r[n - 1];           // row r
            const RowCo<T> &qb = q_r[n];              // row r-1
            // VH of the level below at row r-1, clamped to row r at the southern edge
            const T vhb = r > 0 ? vh_p[n - 1] : vh_cur, vhb_r = r > 0 ? vh_p_r[n - 1] : vh_cur_r;
            T sn, cn;
            stage_cell<T, KAHAN>(col_int && q.interior, q.sa, dle_i, q.dbe_lo, q.dbe_hi, in_s, in_c, in_uh, in_uh_l,
                                 vh_cur, vhb, sn, cn);
            const T sn_r = shl1(sn);
            const T cn_r = KAHAN ? shl1(cn) : T(0);
            // the face depths of this level's rows (truely linear: precomputed), or the
            // elevations they are formed from
            const T fu_a = TL ? hi_r[n - 1] : e_r[n - 1], fu_b = TL ? T(0) : e_rr[n - 1];
            const T fv_a = TL ? hj_r[n - 1] : e_r[n], fv_b = TL ? T(0) : e_r[n - 1];
            const T uhn = mom_u<T, KAHAN, COR, FRIC, TL, LINF, TL>(a, col_mom && q.valid, q.cu, q.fu, dle_ip1,
                                                                    FRIC ? fwu_r[n - 1] : T(0), in_uh, sn, cn, sn_r,
                                                                    cn_r, fu_a, fu_b, vhb, vhb_r, vh_cur, vh_cur_r);
            // VH at row r-1: stage of this level at rows r-1 (previous iteration) and r,
            // UH of the level below at rows r-1 and r
            const T vhn = mom_v<T, KAHAN, COR, FRIC, TL, LINF, TL>(a, col_mom && qb.valid, qb.cv, qb.fv, qb.dbe_hi,
                                                                    FRIC ? fwv_r[n - 1] : T(0), vh_p[n - 1], s_p[n],
                                                                    c_p[n], sn, cn, fv_a, fv_b, in_uhb, in_uhb_l,
                                                                    in_uh, in_uh_l);
            if (GS && n == a.g_level) {
                // the sampled level's values at this iteration's rows
                cap_s = sn; cap_c = cn; cap_uh = uhn; cap_vh = vhn;
            }
            const T uhn_sh = shr1(uhn);
            const T uhn_l = west ? uhn : uhn_sh;
            const T vhn_r = shl1(vhn);
            // the level below's VH of this iteration becomes its 'previous' one
            if (n == 1) {
                // level 0: the loaded row's UH/VH become row m-1's
                uh_p[0] = uh0; uh_p_l[0] = uh0_l;
            }
            vh_p[n - 1] = vh_cur; vh_p_r[n - 1] = vh_cur_r;
            if (n == NS) {
                out_s = s_p[n]; out_c = c_p[n]; out_uh = uh_p[n]; out_vh = vhn;
            }
            // inputs of level n+1: this level's stage/UH of the previous iteration (row
            // m-n), this level's VH now (row m-n)
            in_s = s_p[n]; in_c = c_p[n]; in_uh = uh_p[n]; in_uh_l = uh_p_l[n];
            in_uhb = uh_pp[n]; in_uhb_l = uh_pp_l[n];
            vh_cur = vhn; vh_cur_r = vhn_r;
            // slide this level
            s_p[n] = sn; c_p[n] = cn;
            uh_pp[n] = uh_p[n]; uh_pp_l[n] = uh_p_l[n];
            uh_p[n] = uhn; uh_p_l[n] = uhn_l;
        }
        vh_p[NS] = vh_cur; vh_p_r[NS] = vh_cur_r;
        if (GS && gk) {
            // the gauge samples after the launch's g_level-th step: stage and UH of row r,
            // VH of row r-1
            const int r = m - a.g_level + 1;
            double *v = a.g_val + ((long long)(bz + a.src_base) * a.g_max + a.g_sample) * 3 * a.g_n;
            const double st = KAHAN ? (double)cap_s - (double)cap_c : (double)cap_s;
#pragma unroll
            for (int q = 0; q < 2; q++) {
                const int g = ((gk >> (8 * q)) & 255) - 1, row = (gr >> (16 * q)) & 0xffff;
                if (g >= 0 && r == row) { v[g] = st; v[a.g_n + g] = (double)cap_uh; }
                if (g >= 0 && r - 1 == row) v[2 * a.g_n + g] = (double)cap_vh;
            }
        }
        const int k = m - NS;
        if (own && k >= j0 && k < j1) {
            const int ow = k * a.pitch + col;
            So[ow] = out_s;
            if (KAHAN) Co[ow] = out_c;
            UHo[ow] = out_uh;
            VHo[ow] = out_vh;
        }
    }
}

// One step ahead at one cell: [stage (S, C), UH, VH] after one step at (i, j),
// computed from the state (S, C, UH, VH) with the same per-cell operations as the step
// kernels.
template <typename T, bool KAHAN, bool COR, int FRIC, bool TL, bool LINF>
__device__ __forceinline__ void step1_cell(const LFArgs<T> &a, const T *__restrict__ S, const T *__restrict__ C,
                                           const T *__restrict__ UH, const T *__restrict__ VH, int i, int j, T &s_c,
                                           T &c_c, T &uh1, T &vh1)
{
    const T *__restrict__ E = a.E;
    const int ip = min(i + 1, a.nx - 1), im = max(i - 1, 0), jp = min(j + 1, a.ny - 1), jm = max(j - 1, 0);
    auto at = [&](int ii, int jj) { return (long long)jj * a.pitch + ii; };
    auto cval = [&](long long o) { return KAHAN ? C[o] : T(0); };
    // VH of the row below, clamped to row 0 at the southern edge
    const T vhb_i = j > 0 ? VH[at(i, jm)] : VH[at(i, j)];
    const T vhb_ip = j > 0 ? VH[at(ip, jm)] : VH[at(ip, j)];
    const RowCo<T> q = row_co(a, j), qu = row_co(a, jp);
    const T dle_i = a.rc.dle[i], dle_ip = a.rc.dle[ip], dle_ip1 = a.rc.dle[i + 1];
    const bool ci_int = i > 0 && i < a.nx - 1, cip_int = ip > 0 && ip < a.nx - 1;
    // new stage at (i, j), (i+1, j), (i, j+1); UH(i-1) clamps to UH(0) at the west edge
    T s_e, c_e, s_n, c_n;
    {
        const long long o = at(i, j);
        stage_cell<T, KAHAN>(ci_int && q.interior, q.sa, dle_i, q.dbe_lo, q.dbe_hi, S[o], cval(o), UH[o],
                             UH[at(im, j)], VH[o], vhb_i, s_c, c_c);
    }
    {
        const long long o = at(ip, j);
        stage_cell<T, KAHAN>(cip_int && q.interior, q.sa, dle_ip, q.dbe_lo, q.dbe_hi, S[o], cval(o), UH[o],
                             UH[at(i, j)], VH[o], vhb_ip, s_e, c_e);
    }
    {
        const long long o = at(i, jp);
        stage_cell<T, KAHAN>(ci_int && qu.interior, qu.sa, dle_i, qu.dbe_lo, qu.dbe_hi, S[o], cval(o), UH[o],
                             UH[at(im, jp)], VH[o], VH[at(i, j)], s_n, c_n);
    }
    const bool mom = i < a.nx - 1 && q.valid;
    const T fwu = FRIC ? a.fwU[at(i, j)] : T(0), fwv = FRIC ? a.fwV[at(i, j)] : T(0);
    uh1 = mom_u<T, KAHAN, COR, FRIC, TL, LINF>(a, mom, q.cu, q.fu, dle_ip1, fwu, UH[at(i, j)], s_c, c_c, s_e, c_e,
                                               E[at(i, j)], E[at(ip, j)], vhb_i, vhb_ip, VH[at(i, j)],
                                               VH[at(ip, j)]);
    vh1 = mom_v<T, KAHAN, COR, FRIC, TL, LINF>(a, mom, q.cv, q.fv, q.dbe_hi, fwv, VH[at(i, j)], s_c, c_c, s_n, c_n,
                                               E[at(i, j)], E[at(i, jp)], UH[at(i, j)], UH[at(im, j)],
                                               UH[at(i, jp)], UH[at(im, jp)]);
}

// Gauge values one step ahead (so a two-step launch can pass over a sampling step).
// One lane per (gauge, source); values[((slot * gmax + sample) * 3 + var) * n + g];
// times (single grid, or null): times[sample] = t.
template <typename T, bool KAHAN, bool COR, int FRIC, bool TL, bool LINF>
__global__ __launch_bounds__(64) void gauge_step1_kernel(LFArgs<T> a, int n, const long long *__restrict__ idx,
                                                         double *values, int gmax, int sample, double *times,
                                                         double t)
{
    const int g = blockIdx.x * blockDim.x + threadIdx.x;
    if (times && g == 0 && blockIdx.y == 0) times[sample] = t;
    if (g >= n) return;
    const int slot = blockIdx.y + a.src_base;
    const long long so = (long long)slot * a.src_stride;
    const int i = (int)(idx[g] % a.pitch), j = (int)(idx[g] / a.pitch);
    T s_c, c_c, uh1, vh1;
    step1_cell<T, KAHAN, COR, FRIC, TL, LINF>(a, a.S + so, KAHAN ? a.C + so : nullptr, a.UH + so, a.VH + so, i, j,
                                              s_c, c_c, uh1, vh1);
    double *v = values + ((long long)slot * gmax + sample) * 3 * n;
    v[g] = KAHAN ? (double)s_c - (double)c_c : (double)s_c;
    v[n + g] = (double)uh1;
    v[2 * n + g] = (double)vh1;
}

template <typename T, bool KAHAN, bool COR, int FRIC, bool TL, bool LINF>
static hipError_t launch_g1(const LFArgs<T> &a, int nsrc, int n, const long long *idx, double *values, int gmax,
                            int sample, double *times, double t, hipStream_t s)
{
    hipLaunchKernelGGL((gauge_step1_kernel<T, KAHAN, COR, FRIC, TL, LINF>), dim3((n + 63) / 64, nsrc), dim3(64), 0, s,
                       a, n, idx, values, gmax, sample, times, t);
    return hipGetLastError();
}

template <typename T, bool KAHAN, bool COR, int FRIC>
static hipError_t g1_tl(const LFArgs<T> &a, int tl, int linf, int nsrc, int n, const long long *idx, double *values,
                        int gmax, int sample, double *times, double t, hipStream_t s)
{
    if (tl)
        return linf ? launch_g1<T, KAHAN, COR, FRIC, true, true>(a, nsrc, n, idx, values, gmax, sample, times, t, s)
                    : launch_g1<T, KAHAN, COR, FRIC, true, false>(a, nsrc, n, idx, values, gmax, sample, times, t, s);
    return linf ? launch_g1<T, KAHAN, COR, FRIC, false, true>(a, nsrc, n, idx, values, gmax, sample, times, t, s)
                : launch_g1<T, KAHAN, COR, FRIC, false, false>(a, nsrc, n, idx, values, gmax, sample, times, t, s);
}

template <typename T, bool KAHAN>
hipError_t launch_gauge_step1(const LFArgs<T> &a, int coriolis, int friction_mode, int truely_linear,
                              int linear_friction, int nsrc, int n, const long long *idx, double *values, int gmax,
                              int sample, double *times, double t, hipStream_t s)
{
    if (friction_mode > 1 || a.bc_mask || n <= 0 || nsrc <= 0) return hipErrorInvalidValue;
    const int linf = (friction_mode == 0) && linear_friction;
    if (coriolis) {
        return friction_mode
                   ? g1_tl<T, KAHAN, true, 1>(a, truely_linear, linf, nsrc, n, idx, values, gmax, sample, times, t, s)
                   : g1_tl<T, KAHAN, true, 0>(a, truely_linear, linf, nsrc, n, idx, values, gmax, sample, times, t, s);
    }
    return friction_mode
               ? g1_tl<T, KAHAN, false, 1>(a, truely_linear, linf, nsrc, n, idx, values, gmax, sample, times, t, s)
               : g1_tl<T, KAHAN, false, 0>(a, truely_linear, linf, nsrc, n, idx, values, gmax, sample, times, t, s);
}

template hipError_t launch_gauge_step1<double, false>(const LFArgs<double> &, int, int, int, int, int, int,
                                                      const long long *, double *, int, int, double *, double,
                                                      hipStream_t);
template hipError_t launch_gauge_step1<float, true>(const LFArgs<float> &, int, int, int, int, int, int,
                                                    const long long *, double *, int, int, double *, double,
                                                    hipStream_t);

// Rows per strip of a single grid when the caller leaves them open (rows_per_block <= 0):
// see occupancy_rows. Measured on a 2048 x 2048 grid: 12-16 rows for the FP32 kernel, 24
// for the FP64 one, against 32 for a fixed 2048-wavefront target.
template <typename K>
static int lf2_auto_rows(K kernel, int nx, int ny, int block_x)
{
    static std::atomic<int> occ[5];   // workgroups per CU, by block_x / 64
    return occupancy_rows((const void *)kernel, occ, nx, ny, block_x, kOwn, 3);
}

template <typename T, bool KAHAN, bool COR, int FRIC, bool TL, bool LINF>
static hipError_t launch_lf2(const LFArgs<T> &a0, int nsrc, int block_x, hipStream_t s, int pf)
{
    const int waves = (a0.nx + kOwn - 1) / kOwn, wpb = block_x / 64;
    const int gx = (waves + wpb - 1) / wpb;
    LFArgs<T> a = a0;
    a.bf.ey = bf_rows(a.bf, gx);
    if (a.bf.ey < 0) return hipErrorInvalidValue;
    if (a.rows_per_block <= 0) {
        if (nsrc != 1) return hipErrorInvalidValue;
        a.rows_per_block = pf ? lf2_auto_rows(leapfrog2_kernel<T, KAHAN, COR, FRIC, TL, LINF, true>, a.nx, a.ny,
                                              block_x)
                              : lf2_auto_rows(leapfrog2_kernel<T, KAHAN, COR, FRIC, TL, LINF, false>, a.nx, a.ny,
                                              block_x);
    }
    dim3 grid(gx, (a.ny + a.rows_per_block - 1) / a.rows_per_block + a.bf.ey, nsrc);
    if (pf)
        hipLaunchKernelGGL((leapfrog2_kernel<T, KAHAN, COR, FRIC, TL, LINF, true>), grid, dim3(block_x), 0, s, a);
    else
        hipLaunchKernelGGL((leapfrog2_kernel<T, KAHAN, COR, FRIC, TL, LINF, false>), grid, dim3(block_x), 0, s, a);
    return hipGetLastError();
}

template <typename T, bool KAHAN, bool COR, int FRIC>
static hipError_t lf2_tl(const LFArgs<T> &a, int tl, int linf, int nsrc, int bx, hipStream_t s, int pf)
{
    if (tl) return linf ? launch_lf2<T, KAHAN, COR, FRIC, true, true>(a, nsrc, bx, s, pf)
                        : launch_lf2<T, KAHAN, COR, FRIC, true, false>(a, nsrc, bx, s, pf);
    return linf ? launch_lf2<T, KAHAN, COR, FRIC, false, true>(a, nsrc, bx, s, pf)
                : launch_lf2<T, KAHAN, COR, FRIC, false, false>(a, nsrc, bx, s, pf);
}

template <typename T, bool KAHAN>
hipError_t launch_leapfrog2(const LFArgs<T> &a, int coriolis, int friction_mode, int truely_linear,
                            int linear_friction, int nsrc, int block_x, hipStream_t s, int pf)
{
    if (friction_mode > 1 || a.bc_mask || a.maxq_mask || (block_x % 64) || block_x < 64 || block_x > 256 ||
        (long long)a.pitch * a.ny >= (1LL << 31) ||
        (a.bf.want && (nsrc != 1 || a.src_stride != 0 || !a.bf.part || !a.bf.sc)))
        return hipErrorInvalidValue;
    const int linf = (friction_mode == 0) && linear_friction;
    if (coriolis) {
        return friction_mode ? lf2_tl<T, KAHAN, true, 1>(a, truely_linear, linf, nsrc, block_x, s, pf)
                             : lf2_tl<T, KAHAN, true, 0>(a, truely_linear, linf, nsrc, block_x, s, pf);
    }
    return friction_mode ? lf2_tl<T, KAHAN, false, 1>(a, truely_linear, linf, nsrc, block_x, s, pf)
                         : lf2_tl<T, KAHAN, false, 0>(a, truely_linear, linf, nsrc, block_x, s, pf);
}

template hipError_t launch_leapfrog2<double, false>(const LFArgs<double> &, int, int, int, int, int, int, hipStream_t,
                                                    int);

template <typename T, bool KAHAN, bool COR, int FRIC, bool TL, bool LINF, int NS>
static hipError_t launch_lfn(const LFArgs<T> &a, int nsrc, int block_x, hipStream_t s, int pf)
{
    constexpr int OWN = 65 - 2 * NS;
    const int waves = (a.nx + OWN - 1) / OWN, wpb = block_x / 64;
    dim3 grid((waves + wpb - 1) / wpb, (a.ny + a.rows_per_block - 1) / a.rows_per_block, nsrc);
    if (a.g_n && a.g_level >= 1 && a.g_level <= NS)
        // the variant that writes the gauge samples falling inside the launch (with the
        // next row's loads in flight): used only by launches that hold a sample
        hipLaunchKernelGGL((leapfrogn_kernel<T, KAHAN, COR, FRIC, TL, LINF, true, NS, true>), grid, dim3(block_x), 0, s,
                           a);
    else if (pf)
        hipLaunchKernelGGL((leapfrogn_kernel<T, KAHAN, COR, FRIC, TL, LINF, true, NS>), grid, dim3(block_x), 0, s, a);
    else
        hipLaunchKernelGGL((leapfrogn_kernel<T, KAHAN, COR, FRIC, TL, LINF, false, NS>), grid, dim3(block_x), 0, s, a);
    return hipGetLastError();
}

template <typename T, bool KAHAN, bool COR, int FRIC, int NS>
static hipError_t lfn_tl(const LFArgs<T> &a, int tl, int linf, int nsrc, int bx, hipStream_t s, int pf)
{
    if (tl) return linf ? launch_lfn<T, KAHAN, COR, FRIC, true, true, NS>(a, nsrc, bx, s, pf)
                        : launch_lfn<T, KAHAN, COR, FRIC, true, false, NS>(a, nsrc, bx, s, pf);
    return linf ? launch_lfn<T, KAHAN, COR, FRIC, false, true, NS>(a, nsrc, bx, s, pf)
                : launch_lfn<T, KAHAN, COR, FRIC, false, false, NS>(a, nsrc, bx, s, pf);
}

template <typename T, bool KAHAN, int NS>
static hipError_t lfn_cf(const LFArgs<T> &a, int coriolis, int friction_mode, int tl, int linf, int nsrc, int bx,
                         hipStream_t s, int pf)
{
    if (coriolis)
        return friction_mode ? lfn_tl<T, KAHAN, true, 1, NS>(a, tl, linf, nsrc, bx, s, pf)
                             : lfn_tl<T, KAHAN, true, 0, NS>(a, tl, linf, nsrc, bx, s, pf);
    return friction_mode ? lfn_tl<T, KAHAN, false, 1, NS>(a, tl, linf, nsrc, bx, s, pf)
                         : lfn_tl<T, KAHAN, false, 0, NS>(a, tl, linf, nsrc, bx, s, pf);
}

template <typename T, bool KAHAN>
hipError_t launch_leapfrogn(const LFArgs<T> &a, int coriolis, int friction_mode, int truely_linear,
                            int linear_friction, int nsrc, int block_x, hipStream_t s, int pf, int ns)
{
    if (friction_mode > 1 || a.bc_mask || a.maxq_mask || (block_x % 64) || block_x < 64 || block_x > 256 ||
        (long long)a.pitch * a.ny >= (1LL << 31) || ns < 3 || ns > 6)
        return hipErrorInvalidValue;
    const int linf = (friction_mode == 0) && linear_friction;
    switch (ns) {
    case 3: return lfn_cf<T, KAHAN, 3>(a, coriolis, friction_mode, truely_linear, linf, nsrc, block_x, s, pf);
    case 4: return lfn_cf<T, KAHAN, 4>(a, coriolis, friction_mode, truely_linear, linf, nsrc, block_x, s, pf);
    case 5: return lfn_cf<T, KAHAN, 5>(a, coriolis, friction_mode, truely_linear, linf, nsrc, block_x, s, pf);
    default: return lfn_cf<T, KAHAN, 6>(a, coriolis, friction_mode, truely_linear, linf, nsrc, block_x, s, pf);
    }
}

template hipError_t launch_leapfrogn<double, false>(const LFArgs<double> &, int, int, int, int, int, int, hipStream_t,
                                                    int, int);
template hipError_t launch_leapfrog2<float, true>(const LFArgs<float> &, int, int, int, int, int, int, hipStream_t,
                                                  int);
template hipError_t launch_leapfrogn<float, true>(const LFArgs<float> &, int, int, int, int, int, int, hipStream_t,
                                                  int, int);

}  // namespace swh
