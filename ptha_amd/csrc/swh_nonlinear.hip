// 'leapfrog_nonlinear': the nonlinear staggered leapfrog
// (domain_mod_leapfrog_nonlinear_include.f90:61-826, with FROUDE_LIMITING as
// one_leapfrog_nonlinear_step compiles it, domain_mod_timestepping_alternatives_include.f90:439-447).
//
// The stage half-step is the linear solver's (lf_stage_kernel), then the boundary,
// forcing and boundary fluxes. The momentum half-step runs as four passes over the
// grid, each a pure function of planes the previous pass wrote, so every cell is
// computed independently in the reference's operation order:
//   1. the time-lagged depth: the mean of the depth before and after the stage update
//      on rows 1..ny-1 (row ny keeps the old depth), :265-268;
//   2. the momentum advection work U*UH, U*VH at UH points and V*VH, V*UH at VH points
//      from the pre-update UH/VH, with the reference's extrapolations to the east
//      columns and north rows written as an index map (:274-328);
//   3. the four upwind momentum fluxes with their equal-and-opposite and dry-cell rules
//      (:338-481);
//   4. the update itself: pressure gradient on a wet/dry-switched depth, the flux
//      divergence, Coriolis and the spherical metric terms, semi-implicit friction
//      with friction_work formed from the new stage (as precompute_friction_work with
//      truely_linear = .false.), the Froude limiter, and no outflow from dry cells
//      (:575-815). Column nx and row ny keep their UH/VH (the reference never
//      updates them) apart from the dry-cell rule.
// The old UH/VH stay in the current planes and the update writes the spare ones, so
// the row-carried Coriolis averages of the reference are read directly.
#include "swh_internal.h"
#include "swh_launch.h"

namespace swh {

namespace {

template <typename T>
__device__ __forceinline__ T nl_max0(T a) { return a > T(0) ? a : T(0); }

template <typename T>
__device__ __forceinline__ T dmaxT_nl(T x, T y) { return x > y ? x : y; }

template <typename T>
__global__ void nl_depth0_kernel(NLArgs<T> a)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y;
    if (i >= a.nx || j >= a.ny) return;
    const long long o = (long long)j * a.pitch + i;
    a.DL[o] = nl_max0<T>(a.S0[o] - a.E[o]);
}

template <typename T>
__global__ void nl_depth1_kernel(NLArgs<T> a)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y;
    if (i >= a.nx || j >= a.ny - 1) return;
    const long long o = (long long)j * a.pitch + i;
    a.DL[o] = T(0.5) * (a.DL[o] + nl_max0<T>(a.S[o] - a.E[o]));
}

template <typename T>
__device__ __forceinline__ T nl_inv_depth(T d)
{
    return d > T(kMinDepth) ? T(1) / d : T(0);
}


// The four advection work values of cell (i, j) with the reference's extrapolations
// to the east columns and north rows as index maps (0-based), :274-328.
template <typename T>
__device__ __forceinline__ T nl_uuh(const NLArgs<T> &a, int i, int j)
{
    // U*UH at (i+1/2, j): columns < nx-1; column nx-1 copies nx-2; row ny-1 copies ny-2
    const int jj = j == a.ny - 1 ? a.ny - 2 : j;
    const int ii = i == a.nx - 1 ? a.nx - 2 : i;
    const long long q = (long long)jj * a.pitch + ii;
    const T inv = nl_inv_depth<T>(T(0.5) * (a.DL[q] + a.DL[q + 1]));
    return (a.UH[q] * a.UH[q]) * inv;
}

template <typename T>
__device__ __forceinline__ T nl_uvh(const NLArgs<T> &a, int i, int j)
{
    // U*VH at (i+1/2, j): columns < nx-2; columns nx-2, nx-1 copy nx-3; row ny-1 copies ny-2
    const int jj = j == a.ny - 1 ? a.ny - 2 : j;
    const int ii = i >= a.nx - 2 ? a.nx - 3 : i;
    const int jm = jj > 0 ? jj - 1 : 0;
    const long long q = (long long)jj * a.pitch + ii, qm = (long long)jm * a.pitch + ii;
    const T inv = nl_inv_depth<T>(T(0.5) * (a.DL[q] + a.DL[q + 1]));
    return a.UH[q] * inv * T(0.25) * (a.VH[qm] + a.VH[qm + 1] + a.VH[q] + a.VH[q + 1]);
}

template <typename T>
__device__ __forceinline__ T nl_vvh(const NLArgs<T> &a, int i, int j)
{
    // V*VH at (i, j+1/2): rows < ny-1; row ny-1 copies ny-2
    const int jj = j == a.ny - 1 ? a.ny - 2 : j;
    const long long q = (long long)jj * a.pitch + i;
    const T inv = nl_inv_depth<T>(T(0.5) * (a.DL[q] + a.DL[q + a.pitch]));
    return a.VH[q] * a.VH[q] * inv;
}

template <typename T>
__device__ __forceinline__ T nl_vuh(const NLArgs<T> &a, int i, int j)
{
    // V*UH at (i, j+1/2): columns > 0, rows < ny-2; column 0 copies 1; rows ny-2, ny-1 copy ny-3
    const int jj = j >= a.ny - 2 ? a.ny - 3 : j;
    const int ii = i == 0 ? 1 : i;
    const long long q = (long long)jj * a.pitch + ii;
    const int P = a.pitch;
    const T inv = nl_inv_depth<T>(T(0.5) * (a.DL[q] + a.DL[q + P]));
    return a.VH[q] * inv * T(0.25) * (a.UH[q - 1] + a.UH[q] + a.UH[q - 1 + P] + a.UH[q + P]);
}

// advection work planes (used when the elevation changes within the step)
template <typename T>
__global__ void nl_advection_kernel(NLArgs<T> a)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y;
    if (i >= a.nx || j >= a.ny) return;
    const long long o = (long long)j * a.pitch + i;
    a.UUH[o] = nl_uuh<T>(a, i, j);
    a.UVH[o] = nl_uvh<T>(a, i, j);
    a.VVH[o] = nl_vvh<T>(a, i, j);
    a.VUH[o] = nl_vuh<T>(a, i, j);
}

template <typename T>
__device__ __forceinline__ bool nl_wet(const NLArgs<T> &a, long long q)
{
    return a.S[q] > a.E[q] + T(kMinDepth);
}

// the four upwind momentum fluxes (fNSu / fNSv row j: the edge at the bottom of row j)
// from the advection work planes
template <typename T>
__global__ void nl_flux_kernel(NLArgs<T> a)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y;
    if (i >= a.nx || j >= a.ny) return;
    const int nx = a.nx, ny = a.ny, P = a.pitch;
    const T *UH = a.UH, *VH = a.VH;
    const long long o = (long long)j * P + i;
    const T H = T(0.5), Z = T(0);
    auto UUH = [&](int ii, int jj) { return a.UUH[(long long)jj * P + ii]; };
    auto UVH = [&](int ii, int jj) { return a.UVH[(long long)jj * P + ii]; };
    auto VVH = [&](int ii, int jj) { return a.VVH[(long long)jj * P + ii]; };
    auto VUH = [&](int ii, int jj) { return a.VUH[(long long)jj * P + ii]; };
    // EW flux of UH (rows < ny-1)
    {
        T f = Z;
        if (j < ny - 1) {
            const T dle = (T)a.dle[i];
            if (i == 0 || i == nx - 1) {
                f = UUH(i, j) * dle;
            } else {
                const T us = UH[o] + UH[o - 1];
                const T ul = UUH(i - 1, j), ur = UUH(i, j);
                f = dle * (us > Z ? ul : ur);
                if (us == Z) f = dle * H * (ul + ur);
            }
            if (!nl_wet(a, o)) f = Z;
        }
        a.fEWu[o] = f;
    }
    // NS flux of UH on the bottom edge of row j (the top edge of row j-1; row 0's
    // bottom edge from row 0 alone)
    {
        T f = Z;
        if (i < nx - 1) {
            const int ja = j > 0 ? j - 1 : 0, jb = j;
            const long long qa = (long long)ja * P + i, qb = (long long)jb * P + i;
            const T dbe = (T)a.dbe[j];
            const T vs = VH[qa] + VH[qa + 1];
            const T wa = UVH(i, ja), wb = UVH(i, jb);
            f = dbe * (vs > Z ? wa : wb);
            if (vs == Z) f = dbe * H * (wa + wb);
            if (!(nl_wet(a, qa) && nl_wet(a, qa + 1) && nl_wet(a, qb) && nl_wet(a, qb + 1))) f = Z;
        }
        a.fNSu[o] = f;
    }
    // NS flux of VH on the bottom edge of row j (all rows)
    {
        T f = Z;
        if (i < nx - 1) {
            const int jm = j > 0 ? j - 1 : 0;
            const long long qm = (long long)jm * P + i;
            const T dbe = (T)(0.5 * (a.dbe[j] + a.dbe[j + 1]));
            const T vs = VH[o] + VH[qm];
            const T vm = VVH(i, jm), v0 = VVH(i, j);
            f = dbe * (vs > Z ? vm : v0);
            if (vs == Z) f = dbe * H * (vm + v0);
            if (!nl_wet(a, o)) f = Z;
        }
        a.fNSv[o] = f;
    }
    // EW flux of VH (rows < ny-1): column i from VUH at i-1 and i; column 0 takes
    // column 1's upwind value before its equal-and-opposite and dry rules
    {
        T f = Z;
        if (j < ny - 1) {
            const int ic = i == 0 ? 1 : i;
            const long long q = (long long)j * P + ic;
            const T dle = (T)a.dle[ic - 1];
            const T us = UH[q - 1] + UH[q - 1 + P];
            const T xl = VUH(ic - 1, j), xr = VUH(ic, j);
            f = dle * (us > Z ? xl : xr);
            if (i > 0) {
                if (us == Z) f = dle * H * (xl + xr);
                if (!(nl_wet(a, q - 1) && nl_wet(a, q) && nl_wet(a, q - 1 + P) && nl_wet(a, q + P))) f = Z;
            }
        }
        a.fEWv[o] = f;
    }
}

// The flux pass when the elevation is unchanged in the step (no forcing of it): the
// time-lagged depth, the advection work and the four fluxes in one column march.
// A wavefront owns kNLOwn columns of a strip of rows: lane l holds column
// strip * kNLOwn - 2 + l (lanes 0, 1 hold the two columns west of the strip, lane 63
// the column east of it: the west neighbour's V*UH needs the UH west of it) and the
// wavefront walks the strip's rows upward. Each lane forms the time-lagged
// depth and the four advection work values of its own cell once per row (the
// operations of nl_depth0/1_kernel and nl_uuh .. nl_vuh), takes its west and east
// neighbours' values from the adjacent lanes and the row below's from what it carried.
// The reference's extrapolations at the east columns (U*UH, U*VH), the west column
// (V*UH) and the north rows (U*VH, V*VH, V*UH) pick the adjacent lane's or the carried
// value, which is the value the index map of nl_uuh .. nl_vuh points to, so every flux
// is the one nl_flux_kernel forms from the planes. The rows below the strip that the
// carried values need (one, two where the strip begins at the last row) are walked
// first without storing.
constexpr int kNLOwn = 61;

template <typename T>
__device__ __forceinline__ T nl_shl1(T v) { return __shfl_down(v, 1); }   // value of lane + 1

template <typename T>
__device__ __forceinline__ T nl_shr1(T v) { return __shfl_up(v, 1); }     // value of lane - 1

template <typename T>
struct NLRaw {
    T s0, s, e, uh, vh;
};

template <typename T>
__device__ __forceinline__ NLRaw<T> nl_raw(const NLArgs<T> &a, int ci, int r)
{
    const long long q = (long long)min(r, a.ny - 1) * a.pitch + ci;
    return NLRaw<T>{a.S0[q], a.S[q], a.E[q], a.UH[q], a.VH[q]};
}

template <typename T>
struct NLCell {
    T uh, vh, dl;
    int wet;
};

// the cell of row r: momenta, time-lagged depth (:265-268, as nl_depth0/1_kernel), wet
template <typename T>
__device__ __forceinline__ NLCell<T> nl_cell(const NLRaw<T> &w, int r, int ny)
{
    const T d0 = nl_max0<T>(w.s0 - w.e);
    NLCell<T> c;
    c.uh = w.uh;
    c.vh = w.vh;
    c.dl = r < ny - 1 ? T(0.5) * (d0 + nl_max0<T>(w.s - w.e)) : d0;
    c.wet = w.s > w.e + T(kMinDepth) ? 1 : 0;
    return c;
}

template <typename T>
__global__ __launch_bounds__(64) void nl_flux_march_kernel(NLArgs<T> a, int rows)
{
    const int nx = a.nx, ny = a.ny, P = a.pitch;
    const int lane = threadIdx.x;
    const int i = (int)blockIdx.x * kNLOwn - 2 + lane;
    const int ci = min(max(i, 0), nx - 1);
    const bool own = lane >= 2 && lane <= kNLOwn + 1 && i < nx;
    const int j0 = (int)blockIdx.y * rows;
    if (j0 >= ny) return;
    const int j1 = min(j0 + rows, ny);
    const int js = j0 == 0 ? 0 : (j0 >= ny - 1 ? j0 - 2 : j0 - 1);
    const T H = T(0.5), Z = T(0), Q = T(0.25);
    const T dle_u = (T)a.dle[ci];                 // EW flux of UH: dle(i)
    const T dle_v = (T)a.dle[max(ci, 1) - 1];     // EW flux of VH: dle(max(i, 1) - 1)
    NLCell<T> cur = nl_cell(nl_raw(a, ci, js), js, ny);
    NLCell<T> nxt = nl_cell(nl_raw(a, ci, js + 1), js + 1, ny);
    NLRaw<T> p1 = nl_raw(a, ci, js + 2), p2 = p1;
    // carried: VH of the row below (row 0 on row 0) and its east neighbour, UH west of
    // this row, the row below's wet flag and selected U*VH, V*VH, V*UH
    T vh_m = a.VH[(long long)max(js - 1, 0) * P + ci];
    T vh_me = nl_shl1(vh_m);
    T uh_w = nl_shr1(cur.uh);
    T uvh_m = Z, vvh_m = Z, vuh_m = Z;
    int wet_m = 0;
    for (int j = js; j < j1; j++) {
        if (j + 3 <= j1) p2 = nl_raw(a, ci, j + 3);
        const bool first = j == 0;
        const T dl_e = nl_shl1(cur.dl);
        const T vh_e = nl_shl1(cur.vh);
        const T uh_nw = nl_shr1(nxt.uh);
        // the four advection work values of this lane's cell (nl_uuh .. nl_vuh without
        // their index maps)
        const T inv_u = nl_inv_depth<T>(H * (cur.dl + dl_e));
        const T inv_v = nl_inv_depth<T>(H * (cur.dl + nxt.dl));
        const T uuh_r = (cur.uh * cur.uh) * inv_u;
        const T uvh_r = cur.uh * inv_u * Q * (vh_m + vh_me + cur.vh + vh_e);
        const T vvh_r = cur.vh * cur.vh * inv_v;
        const T vuh_r = cur.vh * inv_v * Q * (uh_w + cur.uh + uh_nw + nxt.uh);
        const T uuh_w = nl_shr1(uuh_r);
        const T uvh_w = nl_shr1(uvh_r);
        const T vuh_e = nl_shl1(vuh_r);
        // the index maps: U*UH column nx-1 from nx-2, U*VH column nx-2 from nx-3 (nx-1's
        // is not used), V*UH column 0 from 1; U*VH, V*VH row ny-1 from ny-2, V*UH rows
        // ny-2, ny-1 from ny-3 (U*UH of row ny-1 is not used)
        const T uuh = i == nx - 1 ? uuh_w : uuh_r;
        T uvh = i == nx - 2 ? uvh_w : uvh_r;
        T vvh = vvh_r;
        T vuh = i == 0 ? vuh_e : vuh_r;
        if (j == ny - 1) {
            uvh = uvh_m;
            vvh = vvh_m;
        }
        if (j >= ny - 2) vuh = vuh_m;
        const T vuh_w = nl_shr1(vuh);
        const int wm = first ? cur.wet : wet_m;
        const int wbits = wm | (cur.wet << 1) | (nxt.wet << 2);
        const int we = nl_shl1(wbits), ww = nl_shr1(wbits);
        if (j >= j0 && own) {
            const long long o = (long long)j * P + i;
            // EW flux of UH (rows < ny-1)
            T f = Z;
            if (j < ny - 1) {
                if (i == 0 || i == nx - 1) {
                    f = uuh * dle_u;
                } else {
                    const T us = cur.uh + uh_w;
                    const T ul = uuh_w, ur = uuh_r;
                    f = dle_u * (us > Z ? ul : ur);
                    if (us == Z) f = dle_u * H * (ul + ur);
                }
                if (!cur.wet) f = Z;
            }
            a.fEWu[o] = f;
            // NS flux of UH on the bottom edge of row j
            f = Z;
            if (i < nx - 1) {
                const T dbe = (T)a.dbe[j];
                const T vs = vh_m + vh_me;
                const T wa = first ? uvh : uvh_m, wb = uvh;
                f = dbe * (vs > Z ? wa : wb);
                if (vs == Z) f = dbe * H * (wa + wb);
                if (!(wm && (we & 1) && cur.wet && (we & 2))) f = Z;
            }
            a.fNSu[o] = f;
            // NS flux of VH on the bottom edge of row j
            f = Z;
            if (i < nx - 1) {
                const T dbe = (T)(0.5 * (a.dbe[j] + a.dbe[j + 1]));
                const T vs = cur.vh + vh_m;
                const T vm = first ? vvh : vvh_m, v0 = vvh;
                f = dbe * (vs > Z ? vm : v0);
                if (vs == Z) f = dbe * H * (vm + v0);
                if (!cur.wet) f = Z;
            }
            a.fNSv[o] = f;
            // EW flux of VH (rows < ny-1)
            f = Z;
            if (j < ny - 1) {
                const T us = i == 0 ? cur.uh + nxt.uh : uh_w + uh_nw;
                const T xl = i == 0 ? vuh : vuh_w, xr = vuh;
                f = dle_v * (us > Z ? xl : xr);
                if (i > 0) {
                    if (us == Z) f = dle_v * H * (xl + xr);
                    if (!((ww & 2) && cur.wet && (ww & 4) && nxt.wet)) f = Z;
                }
            }
            a.fEWv[o] = f;
        }
        vh_m = cur.vh;
        vh_me = vh_e;
        uh_w = uh_nw;
        wet_m = cur.wet;
        uvh_m = uvh;
        vvh_m = vvh;
        vuh_m = vuh;
        cur = nxt;
        nxt = nl_cell(p1, j + 2, ny);
        p1 = p2;
    }
}

template <typename T>
__device__ __forceinline__ T nl_sign1(T x)
{
    return __builtin_copysign(T(1), x);
}

template <typename T>
__global__ void nl_momentum_kernel(NLArgs<T> a)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y;
    if (i >= a.nx || j >= a.ny) return;
    const int nx = a.nx, ny = a.ny, P = a.pitch;
    const T *S = a.S, *E = a.E, *UH = a.UH, *VH = a.VH;
    const long long o = (long long)j * P + i;
    const T MIN = T(kMinDepth), H = T(0.5), Z = T(0), ONE = T(1), G = T(kGravity);
    T uh = UH[o], vh = VH[o];
    if (i < nx - 1 && j < ny - 1) {
        const T dt = a.dt;
        const T ica_dt = a.ica[j];          // (T)(dt / area(j)), formed on the host
        const T ica_dt_g = G * ica_dt;
        const T ica_dt_vh = a.ica_vh[j];    // (T)(dt / (0.5 (area(j) + area(j+1))))
        const T ica_dt_vh_g = G * ica_dt_vh;
        const long long oe = o + 1, on = o + P;
        const T s0 = S[o], e0 = E[o], se = S[oe], ee = E[oe], sn = S[on], en = E[on];
        const T dw_j = sn - s0;
        T h_jph = H * ((sn + s0) - (en + e0));
        if (h_jph < MIN) h_jph = Z;
        T h_iph = H * ((se + s0) - (ee + e0));
        if (h_iph < MIN) h_iph = Z;
        const bool w0 = s0 > e0 + MIN, we = se > ee + MIN, wn = sn > en + MIN;
        const bool jph_strict = wn && w0;
        const T jph_wet = (jph_strict || (wn && sn > e0 + MIN) || (w0 && s0 > en + MIN)) ? ONE : Z;
        const bool iph_strict = we && w0;
        const T iph_wet = (iph_strict || (we && se > e0 + MIN) || (w0 && s0 > ee + MIN)) ? ONE : Z;
        // pre-update averages (the row below j is row 0 itself on row 0)
        const long long or_ = j > 0 ? o - P : o;
        const T vh_jmh = H * (VH[or_] + VH[or_ + 1]);
        const T vh_jph = H * (VH[o] + VH[o + 1]);
        const long long ow = i > 0 ? o - 1 : o;
        const T uh_j = H * (UH[o] + UH[ow]);
        const T uh_jp1 = H * (UH[on] + UH[ow + P]);
        T dt_half_cor = Z, dt_half_cor_jph = Z, t_uvh = Z, t_uuh = Z;
        if (a.coriolis) {
            dt_half_cor = dt * (T)a.cor[j] * H * (iph_strict ? ONE : Z);
            dt_half_cor_jph = dt * (T)a.corbe[j + 1] * H * (jph_strict ? ONE : Z);
        }
        if (a.spherical) {
            T inv = Z;
            if (iph_strict) inv = ONE / h_iph;
            t_uvh = dt * (T)a.tanlat[j] * uh * H * (vh_jmh + vh_jph) * inv;
            inv = Z;
            if (jph_strict) inv = ONE / h_jph;
            t_uuh = dt * H * ((T)a.tanlat[j] + (T)a.tanlat[j + 1]) * H * (uh_jp1 + uh_j) * H * (uh_jp1 + uh_j) * inv;
        }
        // friction_work from the new stage (domain_mod_leapfrog_solver_friction_include.f90:21-85,
        // not truely linear)
        const T nsq0 = a.nsq[o];
        const T dep_i = dmaxT_nl(H * (s0 - e0 + se - ee), MIN);
        const T dep_j = dmaxT_nl(H * (s0 - e0 + sn - en), MIN);
        const T fwu = G * (H * (nsq0 + a.nsq[oe])) * pow_friction(dep_i, a.fric_power);
        const T fwv = G * (H * (nsq0 + a.nsq[on])) * pow_friction(dep_j, a.fric_power);
        const T vbar = H * (vh_jmh + vh_jph);
        const T ubar = H * (uh_j + uh_jp1);
        const T fm_uh = ONE / (ONE + dt * a.lfc + dt * fwu * sqrt(uh * uh + vbar * vbar));
        const T fm_vh = ONE / (ONE + dt * a.lfc + dt * fwv * sqrt(vh * vh + ubar * ubar));
        T u = uh * iph_wet - ica_dt_g * h_iph * iph_wet * (se - s0) * (T)a.dle[i + 1];
        T v = vh * jph_wet - ica_dt_vh_g * h_jph * jph_wet * dw_j * (T)a.dbe[j + 1];
        u = u - ica_dt * a.beta * iph_wet * ((a.fEWu[oe] - a.fEWu[o]) + (a.fNSu[on] - a.fNSu[o]));
        v = v - ica_dt_vh * a.beta * jph_wet * ((a.fEWv[oe] - a.fEWv[o]) + (a.fNSv[on] - a.fNSv[o]));
        if (a.coriolis) {
            u = u + dt_half_cor * (vh_jmh + vh_jph);
            v = v - dt_half_cor_jph * (uh_j + uh_jp1);
        }
        if (a.spherical) {
            u = u + t_uvh;
            v = v - t_uuh;
        }
        u = fm_uh * u;
        v = fm_vh * v;
        // Froude limiter (:763-783)
        {
            T dl = H * (s0 - e0 + se - ee);
            if (dl < MIN) dl = Z;
            const T lim = a.froude_limit * sqrt(dl * G) * dl * nl_sign1(u);
            if (u * u > lim * lim) u = lim;
        }
        {
            T dl = H * (s0 - e0 + sn - en);
            if (dl < MIN) dl = Z;
            const T lim = a.froude_limit * sqrt(dl * G) * dl * nl_sign1(v);
            if (v * v > lim * lim) v = lim;
        }
        uh = u;
        vh = v;
    }
    // no outflow from dry cells (:792-815), rows < ny-1
    if (j < ny - 1) {
        const bool dry0 = S[o] < E[o] + MIN;
        if (dry0 && uh > Z) uh = Z;
        if (i < nx - 1 && S[o + 1] < E[o + 1] + MIN && uh < Z) uh = Z;
        if (dry0 && vh > Z) vh = Z;
        if (S[o + P] < E[o + P] + MIN && vh < Z) vh = Z;
    }
    a.UHo[o] = uh;
    a.VHo[o] = vh;
}

}  // namespace

template <typename T>
hipError_t launch_nonlinear_depth0(const NLArgs<T> &a, hipStream_t s)
{
    if (a.nx < 3 || a.ny < 3) return hipErrorInvalidValue;
    hipLaunchKernelGGL(nl_depth0_kernel<T>, dim3((a.nx + 255) / 256, a.ny), dim3(256), 0, s, a);
    return hipGetLastError();
}

template <typename T>
hipError_t launch_nonlinear_momentum(const NLArgs<T> &a, bool fused, hipStream_t s)
{
    if (a.nx < 3 || a.ny < 3) return hipErrorInvalidValue;
    const dim3 grid((a.nx + 255) / 256, a.ny), block(256);
    if (fused) {
        static std::atomic<int> occ[2];
        const int rows = occupancy_rows((const void *)nl_flux_march_kernel<T>, occ, a.nx, a.ny, 64, kNLOwn, 1);
        hipLaunchKernelGGL(nl_flux_march_kernel<T>, dim3((a.nx + kNLOwn - 1) / kNLOwn, (a.ny + rows - 1) / rows),
                           dim3(64), 0, s, a, rows);
    } else {
        hipLaunchKernelGGL(nl_depth1_kernel<T>, grid, block, 0, s, a);
        hipLaunchKernelGGL(nl_advection_kernel<T>, grid, block, 0, s, a);
        hipLaunchKernelGGL(nl_flux_kernel<T>, grid, block, 0, s, a);
    }
    hipLaunchKernelGGL(nl_momentum_kernel<T>, grid, block, 0, s, a);
    return hipGetLastError();
}

template hipError_t launch_nonlinear_depth0<double>(const NLArgs<double> &, hipStream_t);
template hipError_t launch_nonlinear_depth0<float>(const NLArgs<float> &, hipStream_t);
template hipError_t launch_nonlinear_momentum<double>(const NLArgs<double> &, bool, hipStream_t);
template hipError_t launch_nonlinear_momentum<float>(const NLArgs<float> &, bool, hipStream_t);

}  // namespace swh
