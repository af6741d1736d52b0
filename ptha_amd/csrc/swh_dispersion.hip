// Dispersion: linear_dispersive_solver_mod.f90 without DISPERSIVE_PEREGRINE_IN_FLUX_FORM
// (the reference's default), staggered (leapfrog) and cell-centred (finite-volume) forms.
//
// solve_dispersive_terms (domain_mod.f90:4319-4343 -> linear_dispersive_solve_TRIDIAG,
// :1193-1354): the explicit part of the dispersive terms from the state stored before the
// shallow-water step and the right-hand side RHS = U - Ax (U itself on the border), the
// solution optionally extrapolated in time from three earlier levels
// (quadratic_extrapolation_mod.f90:91-193), then tridiagonal iterations: every row of UH
// (:663-790) and then every column of VH (:792-921) is one tridiagonal system, solved by
// LAPACK's dgtsv elimination with partial pivoting (the reference vendors it,
// util/lapack_tridiagonal_solves.f:462-673). One lane owns one row (UH) or one column
// (VH) and walks it; its coefficients and right-hand side live in work planes laid out
// so that neighbouring lanes touch neighbouring addresses at each step of the walk.
#include "swh_internal.h"

namespace swh {

namespace {

template <typename T>
__device__ __forceinline__ T dmax_d(T a, T b) { return a > b ? a : b; }
template <typename T>
__device__ __forceinline__ T dmin_d(T a, T b) { return a < b ? a : b; }

// depth at a UH / VH point: msl-relative, zero unless both cells are below msl (:209-210)
template <typename T>
__device__ __forceinline__ T disp_depth(T msl, T e1, T e2)
{
    return (e1 < msl && e2 < msl) ? T(0.5) * (msl - e1 + msl - e2) : T(0);
}

// cell-centred depth: msl-relative where the cell and its two neighbours along the line
// are all below msl (:451, :493)
template <typename T>
__device__ __forceinline__ T disp_depth_cc(T msl, T em, T e0, T ep)
{
    return (em < msl && e0 < msl && ep < msl) ? (msl - e0) : T(0);
}

template <typename T>
__device__ __forceinline__ T disp_premult(const DispArgs<T> &a, T dep, T r)
{
    return dep * dep / (T(3) * r) * dmin_d(T(1), dmax_d(T(0), (dep - a.td2) / (a.td1 - a.td2)));
}

// dgtsv for one right-hand side on strided arrays (element k at p[k * s])
template <typename T>
__device__ void dgtsv_strided(int n, T *dl, T *dg, T *du, T *b, long long s)
{
    T fact, temp;
    for (int i = 0; i < n - 2; i++) {
        const long long k = i * s, k1 = k + s;
        if (fabs(dg[k]) >= fabs(dl[k])) {
            if (dg[k] != T(0)) {
                fact = dl[k] / dg[k];
                dg[k1] = dg[k1] - fact * du[k];
                b[k1] = b[k1] - fact * b[k];
            } else {
                return;
            }
            dl[k] = T(0);
        } else {
            fact = dg[k] / dl[k];
            dg[k] = dl[k];
            temp = dg[k1];
            dg[k1] = du[k] - fact * temp;
            dl[k] = du[k1];
            du[k1] = -fact * dl[k];
            du[k] = temp;
            temp = b[k];
            b[k] = b[k1];
            b[k1] = temp - fact * b[k1];
        }
    }
    if (n > 1) {
        const long long k = (long long)(n - 2) * s, k1 = k + s;
        if (fabs(dg[k]) >= fabs(dl[k])) {
            if (dg[k] != T(0)) {
                fact = dl[k] / dg[k];
                dg[k1] = dg[k1] - fact * du[k];
                b[k1] = b[k1] - fact * b[k];
            } else {
                return;
            }
        } else {
            fact = dg[k] / dl[k];
            dg[k] = dl[k];
            temp = dg[k1];
            dg[k1] = du[k] - fact * temp;
            du[k] = temp;
            temp = b[k];
            b[k] = b[k1];
            b[k1] = temp - fact * b[k1];
        }
    }
    const long long kn = (long long)(n - 1) * s;
    if (dg[kn] == T(0)) return;
    b[kn] = b[kn] / dg[kn];
    if (n > 1) b[kn - s] = (b[kn - s] - du[kn - s] * b[kn]) / dg[kn - s];
    for (int i = n - 3; i >= 0; i--) {
        const long long k = i * s;
        b[k] = (b[k] - du[k] * b[k + s] - dl[k] * b[k + 2 * s]) / dg[k];
    }
}

// explicit part from the stored state (linear_dispersive_staggered_matmult, :182-417)
// and the right-hand side (:1226-1265), per cell
template <typename T>
__global__ void disp_rhs_kernel(DispArgs<T> a)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y;
    if (i >= a.nx || j >= a.ny) return;
    const int nx = a.nx, ny = a.ny, P = a.pitch;
    const long long o = (long long)j * P + i;
    const bool edge = i == 0 || j == 0 || i == nx - 1 || j == ny - 1;
    if (edge) {
        a.RU[o] = a.UH[o];
        a.RV[o] = a.VH[o];
        return;
    }
    const T msl = a.msl, dlon = a.dlon, dlat = a.dlat, dle1 = (T)a.dle[0];
    const T *el = a.LE, *uh = a.LU, *vh = a.LV;
    if (a.cell) {
        // linear_dispersive_cellcentred_matmult (:419-660)
        const T dbe_jm = (T)a.dbe[j - 1], dbe_j = (T)a.dbe[j], dbe_jp = (T)a.dbe[j + 1], dbe_jpp = (T)a.dbe[j + 2];
        {
            const T Rdlon = T(0.5) * (dbe_jp + dbe_j);
            const T Rdlat = Rdlon * dlat / dlon;
            const T cjp1 = T(0.5) * (dbe_jpp + dbe_jp) * dlat / (dlon * dle1);
            const T cjm1 = T(0.5) * (dbe_j + dbe_jm) * dlat / (dlon * dle1);
            const T dep = disp_depth_cc(msl, el[o - 1], el[o], el[o + 1]);
            const T pm = disp_premult(a, dep, Rdlon);
            const T ax = pm * (T(1) / Rdlon * (uh[o + 1] - T(2) * uh[o] + uh[o - 1]) +
                               T(1) / Rdlat * ((cjp1 * vh[o + 1 + P] - cjm1 * vh[o + 1 - P]) * T(0.25) -
                                               (cjp1 * vh[o - 1 + P] - cjm1 * vh[o - 1 - P]) * T(0.25)));
            a.RU[o] = a.UH[o] - ax;
        }
        {
            const T r_dlat = dle1;
            const T Rdlon_jph = dbe_jp, Rdlon_jmh = dbe_j;
            const T Rdlat_jph = Rdlon_jph * dlat / dlon, Rdlat_jmh = Rdlon_jmh * dlat / dlon;
            const T cjp1 = T(0.5) * (dbe_jpp + dbe_jp) * dlat / (dle1 * dlon);
            const T cj = T(0.5) * (dbe_jp + dbe_j) * dlat / (dle1 * dlon);
            const T cjm1 = T(0.5) * (dbe_j + dbe_jm) * dlat / (dle1 * dlon);
            const T dep = disp_depth_cc(msl, el[o - P], el[o], el[o + P]);
            const T pm = disp_premult(a, dep, r_dlat);
            const T ax = pm * ((T(1) / Rdlon_jph * (uh[o + 1 + P] - uh[o - 1 + P] + uh[o + 1] - uh[o - 1]) * T(0.25) -
                                T(1) / Rdlon_jmh * (uh[o + 1 - P] - uh[o - 1 - P] + uh[o + 1] - uh[o - 1]) * T(0.25)) +
                               (T(1) / Rdlat_jph * (cjp1 * vh[o + P] - cj * vh[o]) -
                                T(1) / Rdlat_jmh * (cj * vh[o] - cjm1 * vh[o - P])));
            a.RV[o] = a.VH[o] - ax;
        }
        return;
    }
    // UH at (i+1/2, j)
    {
        const T dbe_j = (T)a.dbe[j], dbe_jp = (T)a.dbe[j + 1];
        const T Rdlon = T(0.5) * (dbe_jp + dbe_j);
        const T Rdlat = Rdlon * dlat / dlon;
        const T cjph = dbe_jp * dlat / (dlon * dle1);
        const T cjmh = dbe_j * dlat / (dlon * dle1);
        const T dep = disp_depth(msl, el[o], el[o + 1]);
        const T pm = disp_premult(a, dep, Rdlon);
        const T ax = pm * (T(1) / Rdlon * (uh[o + 1] - T(2) * uh[o] + uh[o - 1]) +
                           T(1) / Rdlat * ((cjph * vh[o + 1] - cjmh * vh[o + 1 - P]) - (cjph * vh[o] - cjmh * vh[o - P])));
        a.RU[o] = a.UH[o] - ax;
    }
    // VH at (i, j+1/2)
    {
        const T r_dlat = dle1;
        const T dbe_j = (T)a.dbe[j], dbe_jp = (T)a.dbe[j + 1], dbe_jpp = (T)a.dbe[j + 2];
        const T Rdlon_jp1 = T(0.5) * (dbe_jpp + dbe_jp);
        const T Rdlon_j = T(0.5) * (dbe_jp + dbe_j);
        const T Rdlat_jp1 = Rdlon_jp1 * dlat / dlon;
        const T Rdlat_j = Rdlon_j * dlat / dlon;
        const T cjp1h = dbe_jpp * dlat / (dle1 * dlon);
        const T cjph = dbe_jp * dlat / (dle1 * dlon);
        const T cjmh = dbe_j * dlat / (dle1 * dlon);
        const T dep = disp_depth(msl, el[o], el[o + P]);
        const T pm = disp_premult(a, dep, r_dlat);
        const T ax = pm * ((T(1) / Rdlon_jp1 * (uh[o + P] - uh[o + P - 1]) - T(1) / Rdlon_j * (uh[o] - uh[o - 1])) +
                           (T(1) / Rdlat_jp1 * (cjp1h * vh[o + P] - cjph * vh[o]) -
                            T(1) / Rdlat_j * (cjph * vh[o] - cjmh * vh[o - P])));
        a.RV[o] = a.VH[o] - ax;
    }
}

// quadratic extrapolation in time of UH and VH (all cells); the coefficients of
// extrapolate_in_time (quadratic_extrapolation_mod.f90:171-181) from the level times and
// the forward time, both on the device (finite-volume times never leave it)
template <typename T>
__global__ void disp_extrapolate_kernel(DispArgs<T> a)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y;
    if (i >= a.nx || j >= a.ny) return;
    const double t_fwd = a.sc ? a.sc->time : a.t_fwd;
    const double tn2 = a.qt[0] - a.qt[2], tn1 = a.qt[1] - a.qt[2], t1 = t_fwd - a.qt[2];
    const double pm = 1.0 / (tn1 * (tn2 * tn2) - tn2 * (tn1 * tn1));
    const T ca = (T)(1.0 + t1 * pm * (-(tn2 * tn2) + tn1 * tn1) + (t1 * t1) * pm * (tn2 - tn1));
    const T cb = (T)(t1 * pm * (tn2 * tn2) + (t1 * t1) * pm * (-tn2));
    const T cc = (T)(t1 * pm * (-(tn1 * tn1)) + (t1 * t1) * pm * tn1);
    const long long o = (long long)j * a.pitch + i, P2 = (long long)a.pitch * a.ny;
    a.UH[o] = a.Y0[o] * ca + a.Y1[o] * cb + a.Y2[o] * cc;
    a.VH[o] = a.Y0[o + P2] * ca + a.Y1[o + P2] * cb + a.Y2[o + P2] * cc;
}

// one lane per row j of UH (rows 1 .. ny-2): work element (i, j) at i * ny + j
template <typename T>
__global__ void disp_sweep_uh_kernel(DispArgs<T> a)
{
    const int j = 1 + blockIdx.x * blockDim.x + threadIdx.x;
    if (j > a.ny - 2) return;
    const int nx = a.nx, ny = a.ny, P = a.pitch;
    const long long s = ny;
    T *dg = a.W, *lo = a.W + (long long)nx * ny, *up = a.W + 2LL * nx * ny, *b = a.W + 3LL * nx * ny;
    const T msl = a.msl, dlon = a.dlon, dlat = a.dlat, dle1 = (T)a.dle[0];
    const T dbe_j = (T)a.dbe[j], dbe_jp = (T)a.dbe[j + 1];
    const T Rdlon = T(0.5) * (dbe_jp + dbe_j);
    const T Rdlat = Rdlon * dlat / dlon;
    const T cjph = dbe_jp * dlat / (dlon * dle1);
    const T cjmh = dbe_j * dlat / (dlon * dle1);
    const long long row = (long long)j * P;
    const int N = nx;
    dg[j] = T(1); dg[(N - 1) * s + j] = T(1);
    lo[j] = T(0); lo[(N - 1) * s + j] = T(0);
    up[j] = T(0); up[(N - 1) * s + j] = T(0);
    b[j] = a.RU[row]; b[(N - 1) * s + j] = a.RU[row + N - 1];
    // cell-centred form (:961-997)
    const T cjp1 = a.cell ? T(0.5) * ((T)a.dbe[j + 2] + dbe_jp) * dlat / (dlon * dle1) : T(0);
    const T cjm1 = a.cell ? T(0.5) * (dbe_j + (T)a.dbe[j - 1]) * dlat / (dlon * dle1) : T(0);
    for (int i = 1; i <= N - 2; i++) {
        const long long o = row + i, k = i * s + j;
        const T dep = a.cell ? disp_depth_cc(msl, a.E[o - 1], a.E[o], a.E[o + 1]) : disp_depth(msl, a.E[o], a.E[o + 1]);
        const T pm = disp_premult(a, dep, Rdlon);
        dg[k] = T(1) + T(2) * pm * T(1) / Rdlon;
        lo[k] = -pm * T(1) / Rdlon;
        up[k] = -pm * T(1) / Rdlon;
        if (a.cell)
            b[k] = a.RU[o] + pm * (T(1) / Rdlat * ((cjp1 * a.VH[o + 1 + P] - cjm1 * a.VH[o + 1 - P]) * T(0.25) -
                                                  (cjp1 * a.VH[o - 1 + P] - cjm1 * a.VH[o - 1 - P]) * T(0.25)));
        else
            b[k] = a.RU[o] + pm * (T(1) / Rdlat * ((cjph * a.VH[o + 1] - cjmh * a.VH[o + 1 - P]) -
                                                  (cjph * a.VH[o] - cjmh * a.VH[o - P])));
    }
    dgtsv_strided<T>(N, lo + s + j, dg + j, up + j, b + j, s);
    for (int i = 0; i < N; i++) a.UH[row + i] = b[i * s + j];
}

// one lane per column i of VH (columns 1 .. nx-2): work element (i, j) at j * nx + i
template <typename T>
__global__ void disp_sweep_vh_kernel(DispArgs<T> a)
{
    const int i = 1 + blockIdx.x * blockDim.x + threadIdx.x;
    if (i > a.nx - 2) return;
    const int nx = a.nx, ny = a.ny, P = a.pitch;
    const long long s = nx;
    T *dg = a.W, *lo = a.W + (long long)nx * ny, *up = a.W + 2LL * nx * ny, *b = a.W + 3LL * nx * ny;
    const T msl = a.msl, dlon = a.dlon, dlat = a.dlat, dle1 = (T)a.dle[0];
    const int N = ny;
    dg[i] = T(1); dg[(N - 1) * s + i] = T(1);
    lo[i] = T(0); lo[(N - 1) * s + i] = T(0);
    up[i] = T(0); up[(N - 1) * s + i] = T(0);
    b[i] = a.RV[i]; b[(N - 1) * s + i] = a.RV[(long long)(N - 1) * P + i];
    for (int j = 1; j <= N - 2; j++) {
        const long long o = (long long)j * P + i, k = j * s + i;
        const T r_dlat = dle1;
        const T dbe_j = (T)a.dbe[j], dbe_jp = (T)a.dbe[j + 1], dbe_jpp = (T)a.dbe[j + 2];
        if (a.cell) {
            // cell-centred form (:1101-1180)
            const T Rdlon_jph = dbe_jp, Rdlon_jmh = dbe_j;
            const T Rdlat_jph = Rdlon_jph * dlat / dlon, Rdlat_jmh = Rdlon_jmh * dlat / dlon;
            const T cjp1 = T(0.5) * (dbe_jpp + dbe_jp) * dlat / (dle1 * dlon);
            const T cj = T(0.5) * (dbe_jp + dbe_j) * dlat / (dle1 * dlon);
            const T cjm1 = T(0.5) * (dbe_j + (T)a.dbe[j - 1]) * dlat / (dle1 * dlon);
            const T dep = disp_depth_cc(msl, a.E[o - P], a.E[o], a.E[o + P]);
            const T pm = disp_premult(a, dep, r_dlat);
            dg[k] = T(1) + pm * (T(1) / Rdlat_jph * cj + T(1) / Rdlat_jmh * cj);
            lo[k] = -pm * T(1) / Rdlat_jmh * cjm1;
            up[k] = -pm * T(1) / Rdlat_jph * cjp1;
            b[k] = a.RV[o] + pm * ((T(1) / Rdlon_jph *
                                        (a.UH[o + 1 + P] - a.UH[o - 1 + P] + a.UH[o + 1] - a.UH[o - 1]) * T(0.25) -
                                    T(1) / Rdlon_jmh *
                                        (a.UH[o + 1 - P] - a.UH[o - 1 - P] + a.UH[o + 1] - a.UH[o - 1]) * T(0.25)));
            continue;
        }
        const T Rdlon_jp1 = T(0.5) * (dbe_jpp + dbe_jp);
        const T Rdlon_j = T(0.5) * (dbe_jp + dbe_j);
        const T Rdlat_jp1 = Rdlon_jp1 * dlat / dlon;
        const T Rdlat_j = Rdlon_j * dlat / dlon;
        const T cjp1h = dbe_jpp * dlat / (dle1 * dlon);
        const T cjph = dbe_jp * dlat / (dle1 * dlon);
        const T cjmh = dbe_j * dlat / (dle1 * dlon);
        const T dep = disp_depth(msl, a.E[o], a.E[o + P]);
        const T pm = disp_premult(a, dep, r_dlat);
        dg[k] = T(1) + pm * (T(1) / Rdlat_jp1 * cjph + T(1) / Rdlat_j * cjph);
        lo[k] = -pm * T(1) / Rdlat_j * cjmh;
        up[k] = -pm * T(1) / Rdlat_jp1 * cjp1h;
        b[k] = a.RV[o] + pm * ((T(1) / Rdlon_jp1 * (a.UH[o + P] - a.UH[o + P - 1]) -
                                T(1) / Rdlon_j * (a.UH[o] - a.UH[o - 1])));
    }
    dgtsv_strided<T>(N, lo + s + i, dg + i, up + i, b + i, s);
    for (int j = 0; j < N; j++) a.VH[(long long)j * P + i] = b[j * s + i];
}

}  // namespace

__global__ void disp_qt_append_kernel(double *qt, const DevScalars *sc, double t_host)
{
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    qt[0] = qt[1];
    qt[1] = qt[2];
    qt[2] = sc ? sc->time : t_host;
}

template <typename T>
hipError_t launch_dispersion(const DispArgs<T> &a, int make_rhs, int extrapolate, int niter, hipStream_t s)
{
    if (a.nx < 3 || a.ny < 3) return hipErrorInvalidValue;
    const dim3 grid((a.nx + 255) / 256, a.ny), block(256);
    if (make_rhs) hipLaunchKernelGGL(disp_rhs_kernel<T>, grid, block, 0, s, a);
    if (extrapolate) hipLaunchKernelGGL(disp_extrapolate_kernel<T>, grid, block, 0, s, a);
    for (int it = 0; it < niter; it++) {
        hipLaunchKernelGGL(disp_sweep_uh_kernel<T>, dim3((a.ny - 2 + 63) / 64), dim3(64), 0, s, a);
        hipLaunchKernelGGL(disp_sweep_vh_kernel<T>, dim3((a.nx - 2 + 63) / 64), dim3(64), 0, s, a);
    }
    return hipGetLastError();
}

template hipError_t launch_dispersion<double>(const DispArgs<double> &, int, int, int, hipStream_t);
template hipError_t launch_dispersion<float>(const DispArgs<float> &, int, int, int, hipStream_t);

hipError_t launch_dispersion_qt_append(double *qt, const DevScalars *sc, double t_host, hipStream_t s)
{
    hipLaunchKernelGGL(disp_qt_append_kernel, dim3(1), dim3(64), 0, s, qt, sc, t_host);
    return hipGetLastError();
}

}  // namespace swh
