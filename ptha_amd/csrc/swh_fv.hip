// Finite-volume 'rk2' step (DE1_low_fr_diffusion fluxes + update_U_manning) for gfx950.
//
// References (paths under propagation/SWALS/src/shallow_water/):
//   compute_depth_and_velocity        domain_mod.f90:1360-1419
//   DE1 fluxes, sources, CFL          domain_mod_compute_fluxes_DE1_inner_include.f90:78-727
//   MC limiter                        extrapolation_limiting_mod.f90:44-79
//   update_U (Manning / Chezy)        domain_mod_update_U_DE1_inner_include.f90:14-99
//   one_euler_step / one_rk2_step     domain_mod_timestepping_alternatives_include.f90:8-146
//
// An Euler stage is three launches: edge fluxes and explicit sources (each cell owns
// its south and west edges; a wavefront walks down the rows of a 61-column strip, the
// clipped depth/velocity and limited gradients of each cell formed about once), a
// one-block scalar pass (boundary-flux sums, CFL step choice, time), and the cell update.
// The CFL step is a running max of 1/dt kept on the device, so an adaptive rk2 step
// never returns to the host. Arithmetic follows the reference's evaluation order.
#include "swh_internal.h"

namespace swh {

constexpr double kLimC1 = 0.1, kLimC2 = 0.3, kLimC3 = 100.0;  // domain_mod.f90:111-117

// MAX / MIN as single v_max_f64 / v_min_f64 instructions: they differ from
// 'a > b ? a : b' only for NaN operands and in the sign of a zero result (max(-0, +0)),
// which no nonzero value of the scheme depends on
__device__ __forceinline__ double fmax_(double a, double b) { return __builtin_fmax(a, b); }
__device__ __forceinline__ double fmin_(double a, double b) { return __builtin_fmin(a, b); }

// ---- clip + depth/velocity of one cell (domain_mod.f90:1374-1412) ----
// Evaluated where it is needed (by the flux pass for a cell and its neighbours, by
// the update for its own cell) instead of being stored: the values are the same
// bits as a stored pass would give, and the state is read once instead of being
// copied into three more planes. Always in FP64: with FP32 state the stored values
// (and the stage's Kahan term) are widened first, so depth = stage - elevation and the
// velocities carry no FP32 rounding (in deep water FP32 depth alone would be off by
// its ulp, 2.4e-4 m at 3000 m).
struct Cell {
    double s, uh, vh, d, u, v;
    bool fixed;   // negative depth clipped (negative_depth_fix_counter)
};

// The stored values of one cell (the flux pass loads them a row ahead of their use).
template <typename T>
struct RawCell {
    T s, c, uh, vh, e;
};

template <typename T>
__device__ __forceinline__ RawCell<T> raw_at(const FVArgs<T> &a, long long o)
{
    RawCell<T> r;
    r.s = a.S0[o];
    r.c = a.C0 ? a.C0[o] : T(0);
    r.uh = a.UH0[o];
    r.vh = a.VH0[o];
    r.e = a.E[o];
    return r;
}

template <typename T>
__device__ __forceinline__ Cell clip_raw(const FVArgs<T> &a, const RawCell<T> &r)
{
    Cell c;
    double s = a.C0 ? (double)r.s - (double)r.c : (double)r.s;
    double uh = r.uh, vh = r.vh;
    const double e = r.e;
    double d = s - e, u, v;
    c.fixed = false;
    if (d > kMinDepth) {
        const double inv = 1.0 / d;
        u = uh * inv;
        v = vh * inv;
    } else {
        if (d < 0.0) {
            d = 0.0;
            s = e;
            c.fixed = true;
        }
        u = 0.0; v = 0.0; uh = 0.0; vh = 0.0;
    }
    c.s = s; c.uh = uh; c.vh = vh; c.d = d; c.u = u; c.v = v;
    return c;
}

template <typename T>
__device__ __forceinline__ Cell cell_at(const FVArgs<T> &a, long long o)
{
    return clip_raw(a, raw_at(a, o));
}

// Wet-dry limiter coefficient (include.f90:751-757)
__device__ __forceinline__ double theta_wd(double theta, double d0, double d1, double d2)
{
    const double mindep = fmin_(fmin_(d0, d1), d2) - kMinDepth;
    const double maxdep = fmax_(fmax_(d0, d1), d2) + kLimC3 * kMinDepth;
    const double tl = (1.0 / (kLimC2 - kLimC1)) * (mindep / maxdep - kLimC1);
    return fmax_(theta * fmin_(1.0, tl), 0.0);
}

// MC limited gradient (extrapolation_limiting_mod.f90:57-74)
__device__ __forceinline__ double mc_grad(double lo, double mid, double hi, double th)
{
    const double a = hi - mid, b = mid - lo;
    const double sa = copysign(1.0, a), sb = copysign(1.0, b);
    const double half_sasb = 0.5 * (sa + sb);
    const double d = fmin_(fabs(a), fabs(b)) * half_sasb * th;
    const double e = 0.5 * (a + b);
    const double c = (d == 0.0) ? 0.0 : e;
    return (d > 0.0) ? fmin_(c, d) : fmax_(c, d);
}

struct Grad4 {
    double st, dp, u, v;
};

// Gradients of (stage, depth, u, v) at a cell from its two neighbours along one
// axis (the caller passes zero gradients for cells on that axis' border).
__device__ __forceinline__ Grad4 grad3(const Cell &lo, const Cell &mid, const Cell &hi, double theta)
{
    Grad4 g;
    const double dl = lo.d, dc = mid.d, dh = hi.d;
    const double th = theta_wd(theta, dl, dc, dh);
    g.st = mc_grad(lo.s, mid.s, hi.s, th);
    g.dp = mc_grad(dl, dc, dh, th);
    g.u = mc_grad(lo.u, mid.u, hi.u, th);
    g.v = mc_grad(lo.v, mid.v, hi.v, th);
    return g;
}

struct EdgeRes {
    double f0, f1, f2;       // mass, momentum (reference UH, VH order)
    double hghh;             // half_g_hh_edge
    double bsp_neg, bsp_pos; // bed-slope pressure terms of the two sides
    double max_speed;
};

// HLL/Audusse edge flux (include.f90:178-432 / 454-718). 'n' = normal velocity
// component, 't' = transverse; ns_edge selects the reference's UH/VH ordering and
// its fr2 summation order (v^2 terms first in both loops).
__device__ __forceinline__ EdgeRes edge_flux(double stage_neg, double stage_pos, double depth_neg, double depth_pos,
                                             double u_neg, double u_pos, double v_neg, double v_pos,
                                             double depth_neg_c, double depth_pos_c, double elev_neg,
                                             double elev_pos, double edge_len, bool ns_edge)
{
    const double EPS = 1.0e-12;
    const double EPS_gs = sqrt(kGravity * kMinDepth);
    const double half_gravity = 0.5 * kGravity;
    EdgeRes r;
    const double z_neg = stage_neg - depth_neg, z_pos = stage_pos - depth_pos;
    const double z_half = fmax_(z_neg, z_pos);
    const double sns = fmax_(stage_neg, z_half), dns = sns - z_half;
    if (dns == 0.0) { u_neg = 0.0; v_neg = 0.0; }
    const double ud_neg = u_neg * dns, vd_neg = v_neg * dns;
    const double gs_neg = sqrt(kGravity * dns);
    const double sps = fmax_(stage_pos, z_half), dps = sps - z_half;
    if (dps == 0.0) { u_pos = 0.0; v_pos = 0.0; }
    const double ud_pos = u_pos * dps, vd_pos = v_pos * dps;
    const double gs_pos = sqrt(kGravity * dps);
    const double vbn = (ns_edge ? v_neg : u_neg) * 1.0;
    const double vbp = (ns_edge ? v_pos : u_pos) * 1.0;
    const double s_max = fmax_(fmax_(vbn + gs_neg, vbp + gs_pos), 0.0);
    const double s_min = fmin_(fmin_(vbn - gs_neg, vbp - gs_pos), 0.0);
    const double denom = s_max - s_min;
    const double inv_denom = (denom > EPS && (gs_pos > EPS_gs || gs_neg > EPS_gs)) ? edge_len / denom : 0.0;
    const double sminsmax = s_min * s_max * 1.0;
    const double fr2 = 1.0 * sqrt(fmax_(0.001, fmin_(1.0, (v_neg * v_neg + v_pos * v_pos + u_neg * u_neg + u_pos * u_pos) /
                                                            (gs_neg * gs_neg + gs_pos * gs_pos + 1.0e-10))));
    const double nd_neg = ns_edge ? vd_neg : ud_neg, nd_pos = ns_edge ? vd_pos : ud_pos;
    r.f0 = (s_max * nd_neg - s_min * nd_pos + sminsmax * (sps - sns)) * inv_denom;
    r.f1 = (s_max * ud_neg * vbn - s_min * ud_pos * vbp + fr2 * sminsmax * (ud_pos - ud_neg)) * inv_denom;
    r.f2 = (s_max * vd_neg * vbn - s_min * vd_pos * vbp + fr2 * sminsmax * (vd_pos - vd_neg)) * inv_denom;
    r.hghh = inv_denom * half_gravity * (s_max * dns * dns - s_min * dps * dps);
    r.bsp_neg = half_gravity * ((dns - depth_neg) * (dns + depth_neg) - (depth_neg + depth_neg_c) * (z_neg - elev_neg));
    r.bsp_pos = half_gravity * ((dps - depth_pos) * (dps + depth_pos) - (depth_pos + depth_pos_c) * (z_pos - elev_pos));
    r.max_speed = fmax_(s_max, -s_min);
    return r;
}

// One edge of a cell that owns it (its south or west edge), from the clipped states
// and limited gradients of the two cells: the three pre-multiplied fluxes, the
// explicit-source contributions to the cell below/left (es_neg: explicit_source_VH_j_
// minus_1 / explicit_source_im1_work) and to the owner (es_pos), and the CFL term
// (include.f90:178-432 north-south, 454-718 east-west).
struct EdgeOut {
    double f0, f1, f2, es_neg, es_pos, mx;
};

template <typename T>
__device__ __forceinline__ EdgeOut ns_edge(const FVArgs<T> &a, int j, const Cell &Sc, const Cell &C,
                                           const Grad4 &gl, const Grad4 &gc, double e_s, double e_c)
{
    const double stage_neg = (double)Sc.s + 0.5 * gl.st;
    const double stage_pos = (double)C.s - 0.5 * gc.st;
    const double depth_neg = (double)Sc.d + 0.5 * gl.dp;
    const double depth_pos = (double)C.d - 0.5 * gc.dp;
    const double u_neg = (double)Sc.u + 0.5 * gl.u, u_pos = (double)C.u - 0.5 * gc.u;
    const double v_neg = (double)Sc.v + 0.5 * gl.v, v_pos = (double)C.v - 0.5 * gc.v;
    const EdgeRes r = edge_flux(stage_neg, stage_pos, depth_neg, depth_pos, u_neg, u_pos, v_neg, v_pos,
                                (double)Sc.d, (double)C.d, e_s, e_c, a.dbe[j], true);
    EdgeOut o;
    o.f0 = r.f0; o.f1 = r.f1; o.f2 = r.f2;
    o.es_neg = r.bsp_neg * a.dbe[j] - r.hghh;
    o.es_pos = 0.0 - r.bsp_pos * a.dbe[j] + r.hghh;
    o.mx = r.max_speed > 1.0e-12 ? r.max_speed * (1.0 / (a.dle[0] * a.half_cfl)) : 0.0;
    return o;
}

template <typename T>
__device__ __forceinline__ EdgeOut ew_edge(const FVArgs<T> &a, int i, int j, const Cell &Wc, const Cell &C,
                                           const Grad4 &el, const Grad4 &ec, double e_w, double e_c)
{
    const double st_neg = (double)Wc.s + 0.5 * el.st;
    const double st_pos = (double)C.s - 0.5 * ec.st;
    const double dp_neg = (double)Wc.d + 0.5 * el.dp;
    const double dp_pos = (double)C.d - 0.5 * ec.dp;
    const double uu_neg = (double)Wc.u + 0.5 * el.u, uu_pos = (double)C.u - 0.5 * ec.u;
    const double vv_neg = (double)Wc.v + 0.5 * el.v, vv_pos = (double)C.v - 0.5 * ec.v;
    const EdgeRes q = edge_flux(st_neg, st_pos, dp_neg, dp_pos, uu_neg, uu_pos, vv_neg, vv_pos, (double)Wc.d,
                                (double)C.d, e_w, e_c, a.dle[i], false);
    EdgeOut o;
    o.f0 = q.f0; o.f1 = q.f1; o.f2 = q.f2;
    o.es_neg = q.bsp_neg * a.dle[i] - q.hghh;
    o.es_pos = 0.0 - q.bsp_pos * a.dle[i] + q.hghh;
    o.mx = q.max_speed > 1.0e-12 ? q.max_speed * (1.0 / (a.dbe[j] * a.half_cfl)) : 0.0;
    return o;
}

// The owner cell's spherical terms: the 1/2 g h^2 metric source, Coriolis and the
// Williamson tan(lat)/R term (include.f90:674-710), added to its explicit sources.
template <typename T>
__device__ __forceinline__ void cell_sources(const FVArgs<T> &a, int j, const Cell &C, double &esu, double &esv)
{
    if (!a.spherical) return;
    const double half_gravity = 0.5 * kGravity;
    const double dc = (double)C.d;
    esv = esv + half_gravity * dc * dc * (a.dbe[j + 1] - a.dbe[j]);
    const double uhc = (double)C.uh, vhc = (double)C.vh;
    double cm;
    if (a.coriolis) {
        cm = a.area[j] * a.cor[j];
        esu = esu + cm * vhc;
        esv = esv - cm * uhc;
    }
    cm = a.area[j] * a.tanlat[j] * (double)C.u;
    esu = esu + cm * vhc;
    esv = esv - cm * uhc;
}

// ---- pass 2: fluxes + explicit sources + CFL ----
// Column marching: one wavefront per block walks `rows` rows of a strip of 64 columns, of which lanes
// 2..62 own their columns (61 per strip; lane 0 and 63 only supply neighbours). At each
// row a lane evaluates one new cell (the row below the next one it handles), the
// north-south gradient of its cell from the three rows it holds, and the east-west
// gradient from its neighbours' cells (lane shuffles); the south edge takes the
// previous row's gradient and cell from registers, the west edge the west lane's. Each
// cell and gradient is formed about once (1.05 and 2.1 per cell, against 2.8 and 2.5 for
// the LDS tile), with no LDS and no barrier; every edge's arithmetic is the LDS
// tile kernel it replaced, so the results are the same bits. The next row's stored
// values are loaded one row ahead of their use.
constexpr int kMOwn = 61;   // owned columns per strip

__device__ __forceinline__ double shfl_from_west(double v) { return __shfl_up(v, 1); }
__device__ __forceinline__ double shfl_from_east(double v) { return __shfl_down(v, 1); }

template <typename T>
__device__ __forceinline__ RawCell<T> raw_clamped(const FVArgs<T> &a, int gi, int j)
{
    const int gj = min(max(j, 0), a.ny - 1);
    return raw_at(a, (long long)gj * a.pitch + gi);
}

template <typename T>
__device__ __forceinline__ Cell cell_clamped(const FVArgs<T> &a, int gi, int j, double &e)
{
    const RawCell<T> r = raw_clamped(a, gi, j);
    e = (double)r.e;
    return clip_raw(a, r);
}

template <typename T>
__device__ __forceinline__ void fv_flux_march_body(const FVArgs<T> &a, int rows)
{
    const int lane = threadIdx.x;
    const int i = blockIdx.x * kMOwn - 2 + lane;
    const int gi = min(max(i, 0), a.nx - 1);
    const int j0 = blockIdx.y * rows, j1 = min(j0 + rows, a.ny);
    const bool own = lane >= 2 && lane <= 62 && i >= 1 && i < a.nx;
    const bool ns_col = i < a.nx;                 // north-south gradients exist in this column
    const bool ew_col = i > 0 && i < a.nx - 1;    // east-west gradients exist in this column
    const Grad4 zero{0.0, 0.0, 0.0, 0.0};
    const long long FP = (long long)a.fpitch * (a.ny + 1), EPn = (long long)a.fpitch * a.ny;
    // rows j0-2, j0-1, j0 and the north-south gradient of row j0-1
    double e_m2, e_m1, e_0;
    const Cell c_m2 = cell_clamped(a, gi, j0 - 2, e_m2);
    Cell c_m1 = cell_clamped(a, gi, j0 - 1, e_m1);
    Cell c_0 = cell_clamped(a, gi, j0, e_0);
    Grad4 g_prev = zero;
    if (ns_col && j0 - 1 > 0 && j0 - 1 < a.ny - 1) g_prev = grad3(c_m2, c_m1, c_0, a.theta);
    double mx = 0.0;
    RawCell<T> next = raw_clamped(a, gi, j0 + 1);   // row j + 1, loaded one row ahead
    for (int j = j0; j < j1; j++) {
        const RawCell<T> cur = next;
        if (j + 1 < j1) next = raw_clamped(a, gi, j + 2);
        const double e_p1 = (double)cur.e;
        const Cell c_p1 = clip_raw(a, cur);
        Grad4 g_ns = zero;
        if (ns_col && j > 0 && j < a.ny - 1) g_ns = grad3(c_m1, c_0, c_p1, a.theta);
        // the west and east neighbours of this row's cell
        Cell w, east;
        w.s = shfl_from_west(c_0.s); w.d = shfl_from_west(c_0.d);
        w.u = shfl_from_west(c_0.u); w.v = shfl_from_west(c_0.v);
        w.uh = 0.0; w.vh = 0.0; w.fixed = false;
        east.s = shfl_from_east(c_0.s); east.d = shfl_from_east(c_0.d);
        east.u = shfl_from_east(c_0.u); east.v = shfl_from_east(c_0.v);
        const double e_w = shfl_from_west(e_0);
        Grad4 g_ew = zero;
        if (ew_col) g_ew = grad3(w, c_0, east, a.theta);
        Grad4 g_ew_w;
        g_ew_w.st = shfl_from_west(g_ew.st); g_ew_w.dp = shfl_from_west(g_ew.dp);
        g_ew_w.u = shfl_from_west(g_ew.u); g_ew_w.v = shfl_from_west(g_ew.v);
        if (own && j >= 1) {
            const long long o = (long long)j * a.pitch + i;
            const long long fo = (long long)j * a.fpitch + i;
            const EdgeOut r = ns_edge(a, j, c_m1, c_0, g_prev, g_ns, e_m1, e_0);
            a.fNS[fo] = (T)r.f0;
            a.fNS[FP + fo] = (T)r.f1;
            a.fNS[2 * FP + fo] = (T)r.f2;
            a.esJ[fo] = r.es_neg;
            const EdgeOut q = ew_edge(a, i, j, w, c_0, g_ew_w, g_ew, e_w, e_0);
            a.fEW[fo] = (T)q.f0;
            a.fEW[EPn + fo] = (T)q.f1;
            a.fEW[2 * EPn + fo] = (T)q.f2;
            a.esE[fo] = q.es_neg;
            double esu = q.es_pos, esv = r.es_pos;
            cell_sources(a, j, c_0, esu, esv);
            a.esU[o] = esu;
            a.esV[o] = esv;
            mx = fmax_(mx, fmax_(r.mx, q.mx));
        }
        c_m1 = c_0; c_0 = c_p1;
        e_m1 = e_0; e_0 = e_p1;
        g_prev = g_ns;
    }
    // block maximum (exact in any order) for the scalar pass
    for (int off = 32; off > 0; off >>= 1) mx = fmax_(mx, __shfl_xor(mx, off));
    if (lane == 0) a.bmax[(long long)blockIdx.y * gridDim.x + blockIdx.x] = mx;
}

template <typename T>
__global__ __launch_bounds__(64) void fv_flux_march_kernel(FVArgs<T> a, int rows)
{
    fv_flux_march_body<T>(a, rows);
}

int fv_flux_march_rows(int rows) { return rows < 4 ? 4 : (rows > 64 ? 64 : rows); }

int fv_flux_grid_blocks(int nx, int ny, int rows)
{
    rows = fv_flux_march_rows(rows);
    return ((nx + kMOwn - 1) / kMOwn) * ((ny + rows - 1) / rows);
}

// capacity of the per-block maxima: the grid of the shortest strips
int fv_flux_blocks(int nx, int ny) { return fv_flux_grid_blocks(nx, ny, 4); }

// ---- pass 3: one block: boundary-flux sums (domain_mod.f90:1483-1499), step choice,
// time and flux-integral bookkeeping (timestepping_alternatives_include.f90:31-50) ----
// 1024 threads: a quarter per boundary line, each thread summing a fixed stride of it
// (loads unrolled so they are in flight together), then a fixed-order tree in LDS:
// deterministic, and a few memory latencies instead of one per 64 elements.
constexpr int kScalarThreads = 1024;

// the step's closing bookkeeping (defined with the kernels that run it)
__device__ void fv_finish(DevScalars &sc, int mode);

// The pass's body for a workgroup of NT threads (kScalarThreads, or 256 inside the
// update launch): the boundary lines are summed over Q = kScalarThreads / 4 lanes each
// whatever NT is (a thread takes every (NT / Q)'th line), so the sums keep one order.
// finish: rk2's closing bookkeeping (kFinishRk2) after the stage's.
template <typename T, int NT>
__device__ __forceinline__ void fv_scalar_body(const T *fNS, const T *fEW, int nx, int ny, int fpitch, int n_ext,
                                               double fixed_dt, double dt_mult, int mode, int finish,
                                               DevScalars *sc, const double *bmax, int nbmax)
{
    constexpr int Q = kScalarThreads / 4;
    constexpr int G = NT / Q;   // lines taken at once
    constexpr int B = 8;        // loads issued together before any is used
    __shared__ double part[4][Q];
    __shared__ double pm[NT / 64];
    // thread 0 reads the step's scalars now, so that their latency overlaps the sums
    // below (neg_depth_count is not touched)
    DevScalars L;
    if (threadIdx.x == 0) {
        L.time = sc->time; L.dt = sc->dt; L.bf_evolve = sc->bf_evolve; L.bf_time = sc->bf_time;
        L.first_max_dt = sc->first_max_dt; L.time0 = sc->time0;
    }
    const int lane = threadIdx.x % Q;
    for (int g = threadIdx.x / Q; g < 4; g += G) {
        // boundary line g: N and S rows of flux_NS, E and W columns of flux_EW; each
        // lane adds every Q'th entry in index order, B loads at a time
        const bool ns = (g == 0 || g == 2);
        const T *src = ns ? fNS : fEW;
        const long long base =
            ns ? (long long)((g == 0) ? ny - n_ext : n_ext) * fpitch : (long long)((g == 1) ? nx - n_ext : n_ext);
        const long long step = ns ? 1 : fpitch;
        const int last = ns ? nx - n_ext - 1 : ny - n_ext - 1;
        double acc = 0.0;
        for (int i0 = n_ext + lane; i0 <= last; i0 += B * Q) {
            double v[B];
#pragma unroll
            for (int u = 0; u < B; u++) {
                const int i = i0 + u * Q;
                v[u] = i <= last ? (double)src[base + i * step] : 0.0;
            }
#pragma unroll
            for (int u = 0; u < B; u++)
                if (i0 + u * Q <= last) acc += v[u];
        }
        part[g][lane] = acc;
    }
    // CFL: maximum over the flux pass's blocks, from 1/maximum_timestep
    // (compute_fluxes_DE1_inner_include.f90:92-95); the entries are >= 0
    double m = 1.0 / kMaxTimestep;
    for (int k0 = threadIdx.x; k0 < nbmax; k0 += B * NT) {
        double v[B];
#pragma unroll
        for (int u = 0; u < B; u++) {
            const int k = k0 + u * NT;
            v[u] = k < nbmax ? bmax[k] : 0.0;
        }
#pragma unroll
        for (int u = 0; u < B; u++) m = fmax_(m, v[u]);
    }
    // the maximum needs no fixed order: across each wavefront by shuffles, then over
    // the wavefronts by thread 0
    for (int off = 32; off > 0; off >>= 1) m = fmax_(m, __shfl_xor(m, off));
    if ((threadIdx.x & 63) == 0) pm[threadIdx.x >> 6] = m;
    __syncthreads();
    for (int w = Q / 2; w > 0; w >>= 1) {
        if (lane < w)
            for (int g = threadIdx.x / Q; g < 4; g += G) part[g][lane] = part[g][lane] + part[g][lane + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        double mx = pm[0];
        for (int w = 1; w < NT / 64; w++) mx = fmax_(mx, pm[w]);
        L.bf_store[0] = part[0][0];
        L.bf_store[1] = part[1][0];
        L.bf_store[2] = -part[2][0];
        L.bf_store[3] = -part[3][0];
        if (mode & kStageBegin) {
            // evolve_one_step resets the step's flux integral (domain_mod.f90:1713-1715);
            // the schemes keep the start time (timestepping_alternatives_include.f90:78, 167, 271)
            L.bf_evolve = 0.0;
            L.time0 = L.time;
        }
        const double max_dt = 1.0 / mx;
        L.max_dt = max_dt;
        if (mode & kStageFirst) {
            L.dt = fixed_dt > 0.0 ? fixed_dt : max_dt;
            L.first_max_dt = max_dt;
        }
        // the stage's step: the base step, or midpoint's half step / rk2n's reduced step
        const double ts = (dt_mult == 1.0) ? L.dt : dt_mult * L.dt;
        L.stage_dt = ts;
        if (mode & kStageScaleBfN) L.bf_evolve = L.bf_evolve * (kRk2nN - 1.0) * (1.0 / (1.0 * kRk2nN));
        if (mode & kStageAccBf)
            L.bf_evolve = L.bf_evolve + ts * (((L.bf_store[0] + L.bf_store[1]) + L.bf_store[2]) + L.bf_store[3]);
        // update_U advances the time (update_U_DE1_inner_include.f90:98)
        L.time = (mode & kStageFromTime0) ? L.time0 + ts : L.time + ts;
        if (mode & kStageFinish) {
            fv_finish(L, kFinishEuler);
            sc->evolve_dt = L.evolve_dt;
        } else if (finish) {
            fv_finish(L, kFinishRk2);
            sc->evolve_dt = L.evolve_dt;
        }
        // (dt changes only in a step's first stage: later stages leave it, which the
        // update beside this pass may be reading)
        if (mode & kStageFirst) sc->dt = L.dt;
        sc->time = L.time; sc->max_dt = L.max_dt;
        for (int q = 0; q < 4; q++) sc->bf_store[q] = L.bf_store[q];
        sc->bf_evolve = L.bf_evolve; sc->bf_time = L.bf_time; sc->first_max_dt = L.first_max_dt;
        sc->time0 = L.time0; sc->stage_dt = L.stage_dt;
    }
}

template <typename T>
__global__ __launch_bounds__(kScalarThreads) void fv_scalar_kernel(const T *fNS, const T *fEW, int nx, int ny,
                                                                   int fpitch, int n_ext, double fixed_dt,
                                                                   double dt_mult, int mode, DevScalars *sc,
                                                                   const double *bmax, int nbmax)
{
    fv_scalar_body<T, kScalarThreads>(fNS, fEW, nx, ny, fpitch, n_ext, fixed_dt, dt_mult, mode, 0, sc, bmax, nbmax);
}

// ---- pass 4: update_U (update_U_DE1_inner_include.f90:38-97) ----
template <typename T>
__global__ void fv_update_kernel(FVArgs<T> a)
{
    if (a.scalar_row && blockIdx.y == 0) {
        // the stage's scalar pass beside the update (the step is fixed: the update takes
        // it from the arguments, and neither touches what the other reads or writes)
        if (blockIdx.x == 0)
            fv_scalar_body<T, 256>(a.fNS, a.fEW, a.nx, a.ny, a.fpitch, a.n_ext, a.s_fixed_dt, a.s_dt_mult, a.s_mode,
                                   a.s_finish, a.sc, a.bmax, a.s_nbmax);
        return;
    }
    const int yb = (int)blockIdx.y - a.scalar_row;   // grid row, then the nesting blocks
    if (yb >= a.ny) {
        // the blocks past the grid: the stage's nesting flux integrals
        // (nested_grid_comms_mod.f90:2308-2461), x += nest_dt * flux on each strip cell
        // (the fluxes are pre-multiplied by the edge length); the update does not touch
        // the flux planes or the integrals, so they run beside it
        const int e = (yb - a.ny) * gridDim.x + blockIdx.x;
        if (e >= a.n_nest_strips * a.nest_chunks) return;
        const NestStrip st = a.nest_strips[e / a.nest_chunks];
        const int r = (e % a.nest_chunks) * blockDim.x + threadIdx.x;
        if (st.line < 0 || r >= st.n) return;
        const long long FP = (long long)a.fpitch * (a.ny + 1), EPn = (long long)a.fpitch * a.ny;
        const long long o = st.kind == 0 ? (long long)st.line * a.fpitch + st.lo + r
                                         : (long long)(st.lo + r) * a.fpitch + st.line;
        for (int k = 0; k < 3; k++) {
            const double f = st.kind == 0 ? (double)a.fNS[k * FP + o] : (double)a.fEW[k * EPn + o];
            double *x = a.nest_int + st.out + (long long)k * st.n + r;
            *x = *x + a.nest_dt * f * st.edge;
        }
        return;
    }
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = yb;
    unsigned long long fixed = 0;
    // with the rk2 average folded in, the step's closing bookkeeping (it does not touch
    // stage_dt, the one scalar this kernel reads; with the scalar pass in this launch,
    // that pass does it); with a stage boundary still to come, which reads the stage's
    // time, the boundary kernel does it
    if (a.avgS && !a.ring && !a.scalar_row && i == 0 && j == 0) fv_finish(*a.sc, kFinishRk2);
    if (i < a.nx && j < a.ny) {
    double dt = a.sc->stage_dt;
    if (a.scalar_row) {
        // the scalar pass's stage step, (dt_mult == 1) ? dt : dt_mult * dt
        const double base = a.stage_dt_dev ? a.sc->dt : a.stage_dt;
        dt = a.stage_dt_dev ? ((a.s_dt_mult == 1.0) ? base : a.s_dt_mult * base) : a.stage_dt;
    }
    const long long o = (long long)j * a.pitch + i;
    const long long fo = (long long)j * a.fpitch + i;
    const long long FP = (long long)a.fpitch * (a.ny + 1);
    const long long EPn = (long long)a.fpitch * a.ny;
    const double inv_cell_area_dt = dt / a.area[j];
    const Cell C = cell_at(a, o);
    double q[3] = {C.s, C.uh, C.vh};
    if (a.Sq) {
        q[0] = a.Cq ? (double)a.Sq[o] - (double)a.Cq[o] : (double)a.Sq[o];
        q[1] = a.UHq[o];
        q[2] = a.VHq[o];
    }
    double out[3];
    if (a.skip_edges && (i == 0 || j == 0 || i == a.nx - 1 || j == a.ny - 1)) {
        // dispersive domains: the edge cells keep their values (update_U_DE1_inner_include.f90:19-31)
        out[0] = q[0];
        out[1] = q[1];
        out[2] = q[2];
        fixed = C.fixed ? 1ull : 0ull;
    } else {
    double un[3];
    for (int kk = 0; kk < 3; kk++) {
        const T *fn = a.fNS + kk * FP;
        const T *fe = a.fEW + kk * EPn;
        un[kk] = q[kk] - inv_cell_area_dt * (((double)fn[fo + a.fpitch] - (double)fn[fo]) +
                                             ((double)fe[fo + 1] - (double)fe[fo]));
    }
    const double depth = (double)C.d;
    if (un[0] <= ((double)a.E[o] + kMinDepth)) {
        un[1] = 0.0;
        un[2] = 0.0;
    } else {
        const double uu = (double)C.u, vv = (double)C.v;
        const double fs = (double)a.nsq[o] * sqrt(uu * uu + vv * vv) * pow_friction(fmax_(depth, kMinDepth), a.fric_power);
        const double implicit_factor = 1.0 / (1.0 + (dt * kGravity) * fmax_(depth, 0.0) * fs + dt * a.lfc);
        un[1] = un[1] + inv_cell_area_dt * (a.esU[o] + a.esE[fo + 1]);
        un[1] = un[1] * implicit_factor;
        un[2] = un[2] + inv_cell_area_dt * (a.esV[o] + a.esJ[fo + a.fpitch]);
        un[2] = un[2] * implicit_factor;
    }
    out[0] = un[0];
    out[1] = un[1];
    out[2] = un[2];
    fixed = C.fixed ? 1ull : 0ull;
    }
    if (a.ring) {
        // pre-average momenta next to the boundary lines, for the stage boundary
        if (i == 1) a.ring[j] = (T)out[1];
        if (i == a.nx - 2) a.ring[a.ny + j] = (T)out[1];
        if (j == 1) a.ring[2 * a.ny + i] = (T)out[2];
        if (j == a.ny - 2) a.ring[2 * a.ny + a.nx + i] = (T)out[2];
    }
    if (a.avgS) {
        // one_rk2_step's average with backup_U (timestepping_alternatives_include.f90:123-134),
        // on the stored values as the separate pass forms it (FP32: UH/VH as stored, the
        // stage with its compensation)
        if (a.Cw) {
            const double avs = a.avgC ? (double)a.avgS[o] - (double)a.avgC[o] : (double)a.avgS[o];
            out[0] = 0.5 * (out[0] + avs);
            out[1] = (double)(T(0.5) * ((T)out[1] + a.avgUH[o]));
            out[2] = (double)(T(0.5) * ((T)out[2] + a.avgVH[o]));
        } else {
            out[0] = 0.5 * (out[0] + (double)a.avgS[o]);
            out[1] = 0.5 * (out[1] + (double)a.avgUH[o]);
            out[2] = 0.5 * (out[2] + (double)a.avgVH[o]);
        }
    }
    const T s_new = (T)out[0];
    a.S[o] = s_new;
    if (a.Cw) a.Cw[o] = (T)((double)s_new - out[0]);   // represented stage S - C = out[0]
    a.UH[o] = (T)out[1];
    a.VH[o] = (T)out[2];
    }
    // negative_depth_fix_counter: wave-level count, one atomic per wave
    for (int off = 32; off > 0; off >>= 1) fixed += __shfl_xor(fixed, off);
    if ((threadIdx.x & 63) == 0 && fixed) atomicAdd(&a.sc->neg_depth_count, fixed);
}

// ---- rk2 average (timestepping_alternatives_include.f90:123-134) ----
template <typename T>
__global__ void fv_average_kernel(T *S, T *UH, T *VH, const T *S0, const T *UH0, const T *VH0, int nx, int ny,
                                  int pitch, DevScalars *sc, T *C, const T *C0)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y;
    if (i == 0 && j == 0) fv_finish(*sc, kFinishRk2);
    if (i >= nx || j >= ny) return;
    const long long o = (long long)j * pitch + i;
    if (C) {
        // FP32 state: the represented stages S - C, kept compensated
        const double v = 0.5 * (((double)S[o] - (double)C[o]) + ((double)S0[o] - (double)C0[o]));
        const T sn = (T)v;
        S[o] = sn;
        C[o] = (T)((double)sn - v);
    } else {
        S[o] = T(0.5) * (S[o] + S0[o]);
    }
    UH[o] = T(0.5) * (UH[o] + UH0[o]);
    VH[o] = T(0.5) * (VH[o] + VH0[o]);
}

// ---- stage boundary on a co-located grid (boundary_mod.f90:78-164) ----
// The reference applies the west/east columns first, then the south/north rows, which
// overwrite the corners. Every boundary cell at once here: a thread takes row j's west
// then east cell, or column i's south then north cell (the same order within a line as
// the reference, for grids of width 2), the column rule leaves the corners on an active
// row line to the row rule, and the row rule takes VH = 0 for the cell next to it on an
// active column line (what the column rule left there); no other cell that one rule
// reads is written by the other.
template <typename T>
__global__ __launch_bounds__(256) void fv_stage_bc_kernel(T *S, T *UH, T *VH, int nx, int ny, int pitch, int m,
                                                          double amp, double period, double msl, const DevScalars *sc,
                                                          T *C)
{
    const double two_pi = 2.0 * acos(-1.0);
    const double stage_d = msl + amp * swhm::sin_acc(two_pi * sc->time / period);
    const T stage = (T)stage_d;
    // FP32 state: the compensation that makes S - C the FP64 boundary stage
    const T cstage = (T)((double)stage - stage_d);
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < ny) {
        const int j = t;
        if ((j == 0 && (m & 4)) || (j == ny - 1 && (m & 1))) return;
        const long long o = (long long)j * pitch;
        if (m & 8) { S[o] = stage; if (C) C[o] = cstage; UH[o] = UH[o + 1]; VH[o] = T(0); }
        if (m & 2) {
            const long long e = o + nx - 1;
            S[e] = stage; if (C) C[e] = cstage; UH[e] = UH[e - 1]; VH[e] = T(0);
        }
    } else if (t < ny + nx) {
        const int i = t - ny;
        const bool col = ((m & 8) && i == 0) || ((m & 2) && i == nx - 1);
        if (m & 4) { S[i] = stage; if (C) C[i] = cstage; VH[i] = col ? T(0) : VH[pitch + i]; UH[i] = T(0); }
        if (m & 1) {
            const long long o = (long long)(ny - 1) * pitch + i;
            S[o] = stage; if (C) C[o] = cstage; VH[o] = col ? T(0) : VH[o - pitch]; UH[o] = T(0);
        }
    }
}

__device__ void fv_finish(DevScalars &sc, int mode)
{
    if (mode == kFinishRk2) {
        // timestepping_alternatives_include.f90:136-142
        sc.time = sc.time0 + sc.dt;
        sc.bf_evolve = 0.5 * sc.bf_evolve;
    } else if (mode == kFinishRk2n) {
        // timestepping_alternatives_include.f90:244-249
        sc.time = sc.time0 + sc.dt * (kRk2nN - 1);
    }
    // the reported CFL step is the first flux computation's (euler and midpoint
    // have only that one as max_dt; rk2 and rk2n restore it)
    sc.max_dt = sc.first_max_dt;
    // evolve_one_step (domain_mod.f90:1799-1803)
    sc.evolve_dt = sc.time - sc.time0;
    sc.bf_time = sc.bf_time + sc.bf_evolve;
}

// Stage boundary with rk2's average folded into the preceding update (the separate
// order: update, boundary, average). The boundary cells' pre-average values follow the
// rules of fv_stage_bc_kernel, with the updated interior neighbours taken from the ring
// (the cells (i, 1) of an active column line hold that column's value, VH = 0); each is
// then averaged with backup_U exactly as fv_average_kernel does. The block that finishes
// last, when every block has read the stage's time, runs the step's closing bookkeeping.
template <typename T>
__device__ __forceinline__ void bc_avg_store(T *S, T *UH, T *VH, const T *S0, const T *UH0, const T *VH0, T *C,
                                             const T *C0, long long o, T stage, T cstage, T uh, T vh)
{
    if (C) {
        const double v = 0.5 * (((double)stage - (double)cstage) + ((double)S0[o] - (double)C0[o]));
        const T sn = (T)v;
        S[o] = sn;
        C[o] = (T)((double)sn - v);
    } else {
        S[o] = T(0.5) * (stage + S0[o]);
    }
    UH[o] = T(0.5) * (uh + UH0[o]);
    VH[o] = T(0.5) * (vh + VH0[o]);
}

template <typename T>
__global__ __launch_bounds__(256) void fv_stage_bc_avg_kernel(T *S, T *UH, T *VH, const T *S0, const T *UH0,
                                                              const T *VH0, const T *ring, int nx, int ny, int pitch,
                                                              int m, double amp, double period, double msl,
                                                              DevScalars *sc, T *C, const T *C0)
{
    const double two_pi = 2.0 * acos(-1.0);
    const double stage_d = msl + amp * swhm::sin_acc(two_pi * sc->time / period);
    const T stage = (T)stage_d;
    const T cstage = C ? (T)((double)stage - stage_d) : T(0);
    const T *rW = ring, *rE = ring + ny, *rS = ring + 2 * ny, *rN = ring + 2 * ny + nx;
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < ny) {
        // row j's west and east cells; corners on an active south/north line are the
        // row rule's
        const int j = t;
        if (!((j == 0 && (m & 4)) || (j == ny - 1 && (m & 1)))) {
            const long long o = (long long)j * pitch;
            if (m & 8) bc_avg_store<T>(S, UH, VH, S0, UH0, VH0, C, C0, o, stage, cstage, rW[j], T(0));
            if (m & 2) bc_avg_store<T>(S, UH, VH, S0, UH0, VH0, C, C0, o + nx - 1, stage, cstage, rE[j], T(0));
        }
    } else if (t < ny + nx) {
        const int i = t - ny;
        // VH of the row next to the line: 0 where an active column line holds that cell
        const bool col = ((m & 8) && i == 0) || ((m & 2) && i == nx - 1);
        if (m & 4) bc_avg_store<T>(S, UH, VH, S0, UH0, VH0, C, C0, i, stage, cstage, T(0), col ? T(0) : rS[i]);
        if (m & 1)
            bc_avg_store<T>(S, UH, VH, S0, UH0, VH0, C, C0, (long long)(ny - 1) * pitch + i, stage, cstage, T(0),
                            col ? T(0) : rN[i]);
    }
    __shared__ bool last;
    __syncthreads();   // this block's threads have read the time
    if (threadIdx.x == 0) {
        const unsigned int done = atomicAdd(&sc->bc_done, 1u);
        last = done == gridDim.x - 1;
        if (last) {
            sc->bc_done = 0;
            fv_finish(*sc, kFinishRk2);
        }
    }
}

// ---- rk2n combination (timestepping_alternatives_include.f90:212-241) ----
template <typename T>
__global__ void fv_rk2n_combine_kernel(T *S, T *UH, T *VH, T *S0, T *UH0, T *VH0, int nx, int ny, int pitch,
                                       int part, DevScalars *sc, T *C, T *C0)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y;
    if (part == 1 && i == 0 && j == 0) fv_finish(*sc, kFinishRk2n);
    if (i >= nx || j >= ny) return;
    const long long o = (long long)j * pitch + i;
    if (part == 0) {
        const T n_inverse = T(1.0 / (1.0 * kRk2nN));
        if (C) {
            const double v = (((double)S0[o] - (double)C0[o]) - ((double)S[o] - (double)C[o])) * (1.0 / (1.0 * kRk2nN));
            const T sn = (T)v;
            S0[o] = sn;
            C0[o] = (T)((double)sn - v);
        } else {
            S0[o] = (S0[o] - S[o]) * n_inverse;
        }
        UH0[o] = (UH0[o] - UH[o]) * n_inverse;
        VH0[o] = (VH0[o] - VH[o]) * n_inverse;
    } else {
        if (C) {
            const double v = ((double)S[o] - (double)C[o]) + ((double)S0[o] - (double)C0[o]);
            const T sn = (T)v;
            S[o] = sn;
            C[o] = (T)((double)sn - v);
        } else {
            S[o] = S[o] + S0[o];
        }
        UH[o] = UH[o] + UH0[o];
        VH[o] = VH[o] + VH0[o];
    }
}

template <typename T>
hipError_t launch_fv_euler(const FVArgs<T> &a, double fixed_dt, double dt_mult, int mode, hipStream_t s)
{
    dim3 block(256), grid((a.nx + 255) / 256, a.ny);
    const int rows = fv_flux_march_rows(a.flux_rows);
    hipLaunchKernelGGL(fv_flux_march_kernel<T>, dim3((a.nx + kMOwn - 1) / kMOwn, (a.ny + rows - 1) / rows), dim3(64),
                       0, s, a, rows);
    const int nbmax = fv_flux_grid_blocks(a.nx, a.ny, a.flux_rows);
    FVArgs<T> u = a;
    if (fixed_dt > 0.0 || !(mode & kStageFirst)) {
        // the stage's step as the scalar pass forms it, known before this stage's
        // fluxes: the fixed step (every stage of a fixed-step step is given the same
        // fixed_dt), or the step a CFL-step's first stage stored; that pass then runs
        // beside the update
        u.scalar_row = 1;
        u.stage_dt_dev = fixed_dt > 0.0 ? 0 : 1;
        u.stage_dt = (dt_mult == 1.0) ? fixed_dt : dt_mult * fixed_dt;
        u.s_fixed_dt = fixed_dt; u.s_dt_mult = dt_mult; u.s_mode = mode; u.s_nbmax = nbmax;
        u.s_finish = (a.avgS && !a.ring) ? 1 : 0;
        grid.y += 1;
    } else {
        hipLaunchKernelGGL(fv_scalar_kernel<T>, dim3(1), dim3(kScalarThreads), 0, s, a.fNS, a.fEW, a.nx, a.ny,
                           a.fpitch, a.n_ext, fixed_dt, dt_mult, mode, a.sc, (const double *)a.bmax, nbmax);
    }
    if (a.nest_strips && a.n_nest_strips > 0 && a.nest_chunks > 0 && a.nest_dt != 0.0)
        grid.y += (a.n_nest_strips * a.nest_chunks + grid.x - 1) / grid.x;   // the nesting blocks
    hipLaunchKernelGGL(fv_update_kernel<T>, grid, block, 0, s, u);
    return hipGetLastError();
}

template <typename T>
hipError_t launch_fv_average(T *S, T *UH, T *VH, const T *S0, const T *UH0, const T *VH0, int nx, int ny, int pitch,
                             DevScalars *sc, hipStream_t s, T *C, const T *C0)
{
    dim3 block(256), grid((nx + 255) / 256, ny);
    hipLaunchKernelGGL(fv_average_kernel<T>, grid, block, 0, s, S, UH, VH, S0, UH0, VH0, nx, ny, pitch, sc, C, C0);
    return hipGetLastError();
}

template <typename T>
hipError_t launch_fv_stage_bc(T *S, T *UH, T *VH, int nx, int ny, int pitch, int mask, double amp, double period,
                              double msl, const DevScalars *sc, hipStream_t s, T *C)
{
    hipLaunchKernelGGL(fv_stage_bc_kernel<T>, dim3((nx + ny + 255) / 256), dim3(256), 0, s, S, UH, VH, nx, ny, pitch,
                       mask, amp, period, msl, sc, C);
    return hipGetLastError();
}

template <typename T>
hipError_t launch_fv_stage_bc_avg(T *S, T *UH, T *VH, const T *S0, const T *UH0, const T *VH0, const T *ring,
                                  int nx, int ny, int pitch, int mask, double amp, double period, double msl,
                                  DevScalars *sc, hipStream_t s, T *C, const T *C0)
{
    hipLaunchKernelGGL(fv_stage_bc_avg_kernel<T>, dim3((nx + ny + 255) / 256), dim3(256), 0, s, S, UH, VH, S0, UH0,
                       VH0, ring, nx, ny, pitch, mask, amp, period, msl, sc, C, C0);
    return hipGetLastError();
}

template <typename T>
hipError_t launch_fv_rk2n_combine(T *S, T *UH, T *VH, T *S0, T *UH0, T *VH0, int nx, int ny, int pitch, int part,
                                  DevScalars *sc, hipStream_t s, T *C, T *C0)
{
    dim3 block(256), grid((nx + 255) / 256, ny);
    hipLaunchKernelGGL(fv_rk2n_combine_kernel<T>, grid, block, 0, s, S, UH, VH, S0, UH0, VH0, nx, ny, pitch, part,
                       sc, C, C0);
    return hipGetLastError();
}

template hipError_t launch_fv_euler<double>(const FVArgs<double> &, double, double, int, hipStream_t);
template hipError_t launch_fv_euler<float>(const FVArgs<float> &, double, double, int, hipStream_t);
template hipError_t launch_fv_rk2n_combine<double>(double *, double *, double *, double *, double *, double *, int,
                                                   int, int, int, DevScalars *, hipStream_t, double *, double *);
template hipError_t launch_fv_rk2n_combine<float>(float *, float *, float *, float *, float *, float *, int, int, int,
                                                  int, DevScalars *, hipStream_t, float *, float *);
template hipError_t launch_fv_average<double>(double *, double *, double *, const double *, const double *,
                                              const double *, int, int, int, DevScalars *, hipStream_t, double *,
                                              const double *);
template hipError_t launch_fv_average<float>(float *, float *, float *, const float *, const float *, const float *,
                                             int, int, int, DevScalars *, hipStream_t, float *, const float *);
template hipError_t launch_fv_stage_bc<double>(double *, double *, double *, int, int, int, int, double, double,
                                               double, const DevScalars *, hipStream_t, double *);
template hipError_t launch_fv_stage_bc<float>(float *, float *, float *, int, int, int, int, double, double, double,
                                              const DevScalars *, hipStream_t, float *);
template hipError_t launch_fv_stage_bc_avg<double>(double *, double *, double *, const double *, const double *,
                                                   const double *, const double *, int, int, int, int, double, double,
                                                   double, DevScalars *, hipStream_t, double *, const double *);
template hipError_t launch_fv_stage_bc_avg<float>(float *, float *, float *, const float *, const float *,
                                                  const float *, const float *, int, int, int, int, double, double,
                                                  double, DevScalars *, hipStream_t, float *, const float *);

}  // namespace swh
