// Discrete adjoint (exact transpose) of the truely-linear leapfrog step.
//
// The forward step (domain_mod_linear_solver_include.f90:92-113, 211-571 with
// truely_linear, Coriolis and linear drag, walls / no stage boundary) is linear in
// x = (S, UH, VH):  S' = S + D(UH, VH) on interior cells;
//   UH' = LF * (UH - aU (S'(i+1) - S'(i)) + bU * [VH(i,jm) + VH(i+1,jm) + VH(i,j) + VH(i+1,j)])
//   VH' = LF * (VH - aV (S'(j+1) - S'(j)) - bV * [UH(i,j) + UH(im,j) + UH(i,j+1) + UH(im,j+1)])
// on 0 <= i < nx-1, 0 <= j < ny-1 (jm = max(j-1,0), im = max(i-1,0)), identity
// elsewhere, with aU = cu(j) h_iph dle(i+1), aV = cv(j) h_jph dbe(j+1),
// bU = fu(j) [h_iph > 0] / 2, bV = fv(j) [h_jph > 0] / 2, LF = 1/(1 + dt*lfc).
// Given the adjoints (lS', lU', lV') of the outputs, the adjoints of the inputs are
//   pass 1:  lS  = lS' + sum over momentum points of the pressure-gradient transpose
//   pass 2:  lU, lV = identity/Coriolis transpose of (mu = LF*l') + divergence
//            transpose of lS,
// each a gather over +-1 neighbours, so no atomics are needed.
#include "swh_internal.h"

namespace swh {

template <typename T>
__device__ __forceinline__ bool mom_domain(const AdjArgs<T> &a, int i, int j)
{
    return i >= 0 && j >= 0 && i < a.nx - 1 && j < a.ny - 1;
}

template <typename T>
__device__ __forceinline__ bool interior(const AdjArgs<T> &a, int i, int j)
{
    return i > 0 && j > 0 && i < a.nx - 1 && j < a.ny - 1;
}

// h_iph, h_jph of the truely-linear scheme (include.f90:265-280)
template <typename T>
__device__ __forceinline__ T h_iph(const AdjArgs<T> &a, int i, int j)
{
    const T MIN = T(kMinDepth);
    const T e0 = a.E[(long long)j * a.pitch + i], e1 = a.E[(long long)j * a.pitch + i + 1];
    return (e0 < -MIN + a.msl && e1 < -MIN + a.msl) ? a.msl - T(0.5) * (e1 + e0) : T(0);
}
template <typename T>
__device__ __forceinline__ T h_jph(const AdjArgs<T> &a, int i, int j)
{
    const T MIN = T(kMinDepth);
    const T e0 = a.E[(long long)j * a.pitch + i], eu = a.E[(long long)(j + 1) * a.pitch + i];
    return (eu < -MIN + a.msl && e0 < -MIN + a.msl) ? a.msl - T(0.5) * (eu + e0) : T(0);
}

// aU * mu_U at momentum point (i, j), zero outside the momentum domain
template <typename T>
__device__ __forceinline__ T aU_mu(const AdjArgs<T> &a, int i, int j)
{
    if (!mom_domain(a, i, j)) return T(0);
    return a.rc.cu[j] * h_iph(a, i, j) * a.rc.dle[i + 1] * (a.lin_fric_mult * a.lU[(long long)j * a.pitch + i]);
}
template <typename T>
__device__ __forceinline__ T aV_mu(const AdjArgs<T> &a, int i, int j)
{
    if (!mom_domain(a, i, j)) return T(0);
    return a.rc.cv[j] * h_jph(a, i, j) * a.rc.dbe[j + 1] * (a.lin_fric_mult * a.lV[(long long)j * a.pitch + i]);
}

// Pass 1: lS(i,j) = lS'(i,j) + aU mu_U (i,j) - aU mu_U (i-1,j) + aV mu_V (i,j) - aV mu_V (i,j-1)
template <typename T>
__global__ void adj_stage_kernel(AdjArgs<T> a)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y;
    if (i >= a.nx || j >= a.ny) return;
    const long long o = (long long)j * a.pitch + i;
    a.oS[o] = a.lS[o] + aU_mu(a, i, j) - aU_mu(a, i - 1, j) + aV_mu(a, i, j) - aV_mu(a, i, j - 1);
}

// Pass 2: adjoints of UH and VH (oS holds the pass-1 result)
template <typename T>
__global__ void adj_momentum_kernel(AdjArgs<T> a)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y;
    if (i >= a.nx || j >= a.ny) return;
    const long long P = a.pitch;
    const long long o = (long long)j * P + i;
    const T LF = a.lin_fric_mult;
    // identity part
    T lu = mom_domain(a, i, j) ? LF * a.lU[o] : a.lU[o];
    T lv = mom_domain(a, i, j) ? LF * a.lV[o] : a.lV[o];
    // divergence transpose: S'(k,l) -= sa(l) [(U(k,l) - U(k-1,l)) dle(k) + V(k,l) dbe(l+1) - V(k,l-1) dbe(l)]
    if (interior(a, i, j)) {
        lu -= a.rc.sa[j] * a.rc.dle[i] * a.oS[o];
        lv -= a.rc.sa[j] * a.rc.dbe[j + 1] * a.oS[o];
    }
    if (interior(a, i + 1, j)) lu += a.rc.sa[j] * a.rc.dle[i + 1] * a.oS[o + 1];
    if (interior(a, i, j + 1)) lv += a.rc.sa[j + 1] * a.rc.dbe[j + 1] * a.oS[o + P];
    if (a.coriolis) {
        const T HALF = T(0.5);
        // UH(i,j) enters VH'(k,l) = ... - bV(k,l) [U(k,l) + U(km,l) + U(k,l+1) + U(km,l+1)]
        // for k in {i, i+1} (km = max(k-1,0)), l in {j, j-1}; at i = 0 the k = 0 point
        // uses U(0,.) twice.
        T acc = T(0);
        for (int dl = 0; dl <= 1; dl++) {
            const int l = j - dl;
            if (l < 0) continue;
            for (int dk = 0; dk <= 1; dk++) {
                const int k = i + dk;
                if (!mom_domain(a, k, l)) continue;
                const int mult = (i == 0 && dk == 0) ? 2 : 1;  // U(0,.) counted as U(k,.) and U(km,.)
                if (dk == 1 && k - 1 != i) continue;
                const T mV = h_jph(a, k, l) > T(0) ? T(1) : T(0);
                const T bV = HALF * a.rc.fv[l] * mV;
                acc += T(mult) * bV * (LF * a.lV[(long long)l * P + k]);
            }
        }
        lu -= acc;
        // VH(i,j) enters UH'(k,l) = ... + bU(k,l) [V(k,lm) + V(k+1,lm) + V(k,l) + V(k+1,l)]
        // for k in {i, i-1}, l in {j, j+1} (lm = max(l-1,0)); at j = 0 the l = 0 point
        // uses V(.,0) twice.
        acc = T(0);
        for (int dl = 0; dl <= 1; dl++) {
            const int l = j + dl;
            for (int dk = 0; dk <= 1; dk++) {
                const int k = i - dk;
                if (!mom_domain(a, k, l)) continue;
                const int mult = (j == 0 && dl == 0) ? 2 : 1;
                const T mU = h_iph(a, k, l) > T(0) ? T(1) : T(0);
                const T bU = HALF * a.rc.fu[l] * mU;
                acc += T(mult) * bU * (LF * a.lU[(long long)l * P + k]);
            }
        }
        lv += acc;
    }
    a.oU[o] = lu;
    a.oV[o] = lv;
}

// Adds values[g] to lS at the gauge cells (duplicates resolved by one thread).
template <typename T>
__global__ void adj_inject_kernel(T *lS, int n, const long long *idx, const double *values)
{
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    for (int g = 0; g < n; g++) lS[idx[g]] += (T)values[g];
}

template <typename T>
hipError_t launch_adjoint_step(const AdjArgs<T> &a, hipStream_t s)
{
    dim3 block(256), grid((a.nx + 255) / 256, a.ny);
    hipLaunchKernelGGL(adj_stage_kernel<T>, grid, block, 0, s, a);
    hipLaunchKernelGGL(adj_momentum_kernel<T>, grid, block, 0, s, a);
    return hipGetLastError();
}

template <typename T>
hipError_t launch_adjoint_inject(T *lS, int n, const long long *idx, const double *values, hipStream_t s)
{
    hipLaunchKernelGGL(adj_inject_kernel<T>, dim3(1), dim3(64), 0, s, lS, n, idx, values);
    return hipGetLastError();
}

template hipError_t launch_adjoint_step<double>(const AdjArgs<double> &, hipStream_t);
template hipError_t launch_adjoint_step<float>(const AdjArgs<float> &, hipStream_t);
template hipError_t launch_adjoint_inject<double>(double *, int, const long long *, const double *, hipStream_t);
template hipError_t launch_adjoint_inject<float>(float *, int, const long long *, const double *, hipStream_t);

}  // namespace swh
