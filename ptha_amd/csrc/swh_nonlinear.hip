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

// Time-lagged depth at cell q of row j: from the plane, or (no elevation forcing in the
// step, so the elevation is unchanged) formed from the old and new stage exactly as
// nl_depth0_kernel and nl_depth1_kernel form it.
template <typename T, bool FLY>
__device__ __forceinline__ T nl_dl(const NLArgs<T> &a, long long q, int j)
{
    if (!FLY) return a.DL[q];
    const T d0 = nl_max0<T>(a.S0[q] - a.E[q]);
    return j < a.ny - 1 ? T(0.5) * (d0 + nl_max0<T>(a.S[q] - a.E[q])) : d0;
}

// The four advection work values of cell (i, j) with the reference's extrapolations
// to the east columns and north rows as index maps (0-based), :274-328.
template <typename T, bool FLY>
__device__ __forceinline__ T nl_uuh(const NLArgs<T> &a, int i, int j)
{
    // U*UH at (i+1/2, j): columns < nx-1; column nx-1 copies nx-2; row ny-1 copies ny-2
    const int jj = j == a.ny - 1 ? a.ny - 2 : j;
    const int ii = i == a.nx - 1 ? a.nx - 2 : i;
    const long long q = (long long)jj * a.pitch + ii;
    const T inv = nl_inv_depth<T>(T(0.5) * (nl_dl<T, FLY>(a, q, jj) + nl_dl<T, FLY>(a, q + 1, jj)));
    return (a.UH[q] * a.UH[q]) * inv;
}

template <typename T, bool FLY>
__device__ __forceinline__ T nl_uvh(const NLArgs<T> &a, int i, int j)
{
    // U*VH at (i+1/2, j): columns < nx-2; columns nx-2, nx-1 copy nx-3; row ny-1 copies ny-2
    const int jj = j == a.ny - 1 ? a.ny - 2 : j;
    const int ii = i >= a.nx - 2 ? a.nx - 3 : i;
    const int jm = jj > 0 ? jj - 1 : 0;
    const long long q = (long long)jj * a.pitch + ii, qm = (long long)jm * a.pitch + ii;
    const T inv = nl_inv_depth<T>(T(0.5) * (nl_dl<T, FLY>(a, q, jj) + nl_dl<T, FLY>(a, q + 1, jj)));
    return a.UH[q] * inv * T(0.25) * (a.VH[qm] + a.VH[qm + 1] + a.VH[q] + a.VH[q + 1]);
}

template <typename T, bool FLY>
__device__ __forceinline__ T nl_vvh(const NLArgs<T> &a, int i, int j)
{
    // V*VH at (i, j+1/2): rows < ny-1; row ny-1 copies ny-2
    const int jj = j == a.ny - 1 ? a.ny - 2 : j;
    const long long q = (long long)jj * a.pitch + i;
    const T inv = nl_inv_depth<T>(T(0.5) * (nl_dl<T, FLY>(a, q, jj) + nl_dl<T, FLY>(a, q + a.pitch, jj + 1)));
    return a.VH[q] * a.VH[q] * inv;
}

template <typename T, bool FLY>
__device__ __forceinline__ T nl_vuh(const NLArgs<T> &a, int i, int j)
{
    // V*UH at (i, j+1/2): columns > 0, rows < ny-2; column 0 copies 1; rows ny-2, ny-1 copy ny-3
    const int jj = j >= a.ny - 2 ? a.ny - 3 : j;
    const int ii = i == 0 ? 1 : i;
    const long long q = (long long)jj * a.pitch + ii;
    const int P = a.pitch;
    const T inv = nl_inv_depth<T>(T(0.5) * (nl_dl<T, FLY>(a, q, jj) + nl_dl<T, FLY>(a, q + P, jj + 1)));
    return a.VH[q] * inv * T(0.25) * (a.UH[q - 1] + a.UH[q] + a.UH[q - 1 + P] + a.UH[q + P]);
}

// advection work planes (used when the elevation changes within the step)
template <typename T>
__global__ void nl_advection_kernel(NLArgs<T> a)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y;
    if (i >= a.nx || j >= a.ny) return;
    const long long o = (long long)j * a.pitch + i;
    a.UUH[o] = nl_uuh<T, false>(a, i, j);
    a.UVH[o] = nl_uvh<T, false>(a, i, j);
    a.VVH[o] = nl_vvh<T, false>(a, i, j);
    a.VUH[o] = nl_vuh<T, false>(a, i, j);
}

template <typename T>
__device__ __forceinline__ bool nl_wet(const NLArgs<T> &a, long long q)
{
    return a.S[q] > a.E[q] + T(kMinDepth);
}

// the four upwind momentum fluxes (fNSu / fNSv row j: the edge at the bottom of row j).
// FUSED: the advection work values are formed where needed (each from the same
// operations as nl_advection_kernel) instead of read from its planes.
template <typename T, bool FUSED>
__global__ void nl_flux_kernel(NLArgs<T> a)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y;
    if (i >= a.nx || j >= a.ny) return;
    const int nx = a.nx, ny = a.ny, P = a.pitch;
    const T *UH = a.UH, *VH = a.VH;
    const long long o = (long long)j * P + i;
    const T H = T(0.5), Z = T(0);
    auto UUH = [&](int ii, int jj) { return FUSED ? nl_uuh<T, true>(a, ii, jj) : a.UUH[(long long)jj * P + ii]; };
    auto UVH = [&](int ii, int jj) { return FUSED ? nl_uvh<T, true>(a, ii, jj) : a.UVH[(long long)jj * P + ii]; };
    auto VVH = [&](int ii, int jj) { return FUSED ? nl_vvh<T, true>(a, ii, jj) : a.VVH[(long long)jj * P + ii]; };
    auto VUH = [&](int ii, int jj) { return FUSED ? nl_vuh<T, true>(a, ii, jj) : a.VUH[(long long)jj * P + ii]; };
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
        const T ica_dt = (T)(a.dt / a.area[j]);
        const T ica_dt_g = G * ica_dt;
        const T ica_dt_vh = (T)(a.dt / (0.5 * (a.area[j] + a.area[j + 1])));
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
        hipLaunchKernelGGL((nl_flux_kernel<T, true>), grid, block, 0, s, a);
    } else {
        hipLaunchKernelGGL(nl_depth1_kernel<T>, grid, block, 0, s, a);
        hipLaunchKernelGGL(nl_advection_kernel<T>, grid, block, 0, s, a);
        hipLaunchKernelGGL((nl_flux_kernel<T, false>), grid, block, 0, s, a);
    }
    hipLaunchKernelGGL(nl_momentum_kernel<T>, grid, block, 0, s, a);
    return hipGetLastError();
}

template hipError_t launch_nonlinear_depth0<double>(const NLArgs<double> &, hipStream_t);
template hipError_t launch_nonlinear_depth0<float>(const NLArgs<float> &, hipStream_t);
template hipError_t launch_nonlinear_momentum<double>(const NLArgs<double> &, bool, hipStream_t);
template hipError_t launch_nonlinear_momentum<float>(const NLArgs<float> &, bool, hipStream_t);

}  // namespace swh
