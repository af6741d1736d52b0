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
#include <atomic>

#include "swh_internal.h"
#include "swh_launch.h"

namespace swh {

namespace {

constexpr int kAdjOwn = 61;   // columns a wavefront writes (lanes 1..61)

// Lane shifts across the wavefront; every lane must execute them.
template <typename T>
__device__ __forceinline__ T shl1(T v) { return __shfl_down(v, 1); }   // value of lane + 1
template <typename T>
__device__ __forceinline__ T shr1(T v) { return __shfl_up(v, 1); }     // value of lane - 1

// h_iph / h_jph of the truely-linear scheme (include.f90:265-280) from the elevations of
// the two cells of the face (e_far the east or north one)
template <typename T>
__device__ __forceinline__ T face_depth(T e0, T e_far, T msl)
{
    const T MIN = T(kMinDepth);
    return (e0 < -MIN + msl && e_far < -MIN + msl) ? msl - T(0.5) * (e_far + e0) : T(0);
}

}  // namespace

// One transposed step, fused and column-marching. Each lane owns one column (a
// wavefront owns 61: lane 0 is the western halo, lanes 62-63 the eastern one) and
// marches up a strip of rows, keeping the last three rows in registers; east and west
// neighbours come from the adjacent lanes. At row m (loaded now) it forms
//   h_iph(m), aU mu_U(m) = cu h_iph dle(i+1) LF lU,      (and its west neighbour's)
//   h_jph(m-1), aV mu_V(m-1) = cv h_jph dbe(j+1) LF lV,
//   pass 1 at row m-1: lS_in = lS + aU mu_U(i) - aU mu_U(i-1) + aV mu_V(j) - aV mu_V(j-1),
//   pass 2 at row m-2: lU_in, lV_in = identity / Coriolis transpose of LF l' plus the
//   divergence transpose of lS_in (its own, its east neighbour's, the row above's),
// with the operations and their order of the two-pass form, on 0 <= i < nx-1,
// 0 <= j < ny-1 for the momentum points and the interior for the stage. Each strip
// recomputes one row below and two above it; the next two rows' loads are in flight
// while a row is computed.
// Row coefficients of row r (loaded once per row at a clamped index, carried while the
// row is in use): sa, cu, cv, fu, fv, dbe(r+1); mom: 0 <= r < ny-1 (momentum row),
// in: 0 < r < ny-1 (stage row).
template <typename T>
struct AdjRow {
    T sa, cu, cv, fu, fv, dbe1;
    bool mom, in;
};

// A value that is the same on every lane (a row coefficient), kept in scalar registers
// so that the three rows carried across iterations take no vector registers.
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
__device__ __forceinline__ AdjRow<T> adj_row(const AdjArgs<T> &a, int r)
{
    const int rc = min(max(r, 0), a.ny - 1);
    AdjRow<T> q;
    q.sa = uniform(a.rc.sa[rc]);
    q.cu = uniform(a.rc.cu[rc]);
    q.cv = uniform(a.rc.cv[rc]);
    q.fu = uniform(a.rc.fu[rc]);
    q.fv = uniform(a.rc.fv[rc]);
    q.dbe1 = uniform(a.rc.dbe[rc + 1]);
    q.mom = r >= 0 && r < a.ny - 1;
    q.in = r > 0 && r < a.ny - 1;
    return q;
}

template <typename T, bool COR>
__global__ __launch_bounds__(256) void adj_march_kernel(AdjArgs<T> a)
{
    const int lane = threadIdx.x & 63;
    const int wg = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
    if (wg * kAdjOwn >= a.nx) return;                    // whole wavefront past the east edge
    const int col = wg * kAdjOwn - 1 + lane;             // may be -1 or >= nx
    const int ci = min(max(col, 0), a.nx - 1);           // column it loads
    const bool own = lane >= 1 && lane <= kAdjOwn && col < a.nx;
    const int j0 = blockIdx.y * a.rows;
    if (j0 >= a.ny) return;
    const int j1 = min(j0 + a.rows, a.ny);
    const int nx = a.nx, ny = a.ny;
    const long long P = a.pitch;
    const T LF = a.lin_fric_mult, HALF = T(0.5), msl = a.msl;
    // column parts of the momentum-domain and interior predicates (i, i-1, i+1)
    const bool cmom = col >= 0 && col < nx - 1, cmom_w = col - 1 >= 0 && col - 1 < nx - 1,
               cmom_e = col + 1 >= 0 && col + 1 < nx - 1;
    const bool cint = col > 0 && col < nx - 1, cint_e = col + 1 > 0 && col + 1 < nx - 1;
    const T dle_i = a.rc.dle[min(max(col, 0), nx)], dle_ip1 = a.rc.dle[min(max(col + 1, 0), nx)];
    const T mult_i = T(col == 0 ? 2 : 1);                // U(0,.) counted twice at i = 0

    // carried rows (m-1, m-2, m-3 relative to the row loaded now)
    T ls1 = T(0), lu1 = T(0), lv1 = T(0), e1 = T(0), hI1 = T(0), hIw1 = T(0), luw1 = T(0), aU1 = T(0), aUw1 = T(0);
    T lu2 = T(0), lv2 = T(0), hI2 = T(0), hIw2 = T(0), luw2 = T(0), oS2 = T(0), lve2 = T(0), hJ2 = T(0),
      hJe2 = T(0), aV2 = T(0);
    T lv3 = T(0), lve3 = T(0), hJ3 = T(0), hJe3 = T(0);
    AdjRow<T> q1 = adj_row(a, j0 - 2), q2 = adj_row(a, j0 - 3), q3 = adj_row(a, j0 - 4);

    // the next two rows' loads are issued before this row is computed (two rows in flight)
    auto load_row = [&](int m, T &ls, T &lu, T &lv, T &e) {
        const long long o = (long long)min(max(m, 0), ny - 1) * P + ci;
        ls = a.lS[o];
        lu = a.lU[o];
        lv = a.lV[o];
        e = a.E[o];
    };
    T n_ls, n_lu, n_lv, n_e, nn_ls, nn_lu, nn_lv, nn_e;
    load_row(j0 - 1, n_ls, n_lu, n_lv, n_e);
    load_row(j0, nn_ls, nn_lu, nn_lv, nn_e);
    for (int m = j0 - 1; m <= j1 + 1; m++) {
        const T ls0 = n_ls, lu0 = n_lu, lv0 = n_lv, e0 = n_e;
        n_ls = nn_ls; n_lu = nn_lu; n_lv = nn_lv; n_e = nn_e;
        if (m + 2 <= j1 + 1) load_row(m + 2, nn_ls, nn_lu, nn_lv, nn_e);
        const AdjRow<T> q0 = adj_row(a, m);
        const T e0e = shl1(e0);
        // row m: h_iph and aU mu_U (own and the west neighbour's)
        const T hI0 = face_depth(e0, e0e, msl);
        const T hIw0 = shr1(hI0);
        const T luw0 = shr1(lu0);
        const T aU0 = (cmom && q0.mom) ? q0.cu * hI0 * dle_ip1 * (LF * lu0) : T(0);
        const T aUw0 = shr1(aU0);
        // row m-1: h_jph and aV mu_V
        const T hJ1 = face_depth(e1, e0, msl);
        const T hJe1 = shl1(hJ1);
        const T lve1 = shl1(lv1);
        const T aV1 = (cmom && q1.mom) ? q1.cv * hJ1 * q1.dbe1 * (LF * lv1) : T(0);
        // pass 1 at row m-1 (aU mu_U(i-1) is zero outside the momentum domain: the west
        // lane's value carries its own column's predicate)
        const T oS1 = (((ls1 + aU1) - aUw1) + aV1) - aV2;
        const T oSe2 = shl1(oS2);
        // pass 2 at row r = m-2
        const int r = m - 2;
        const bool mom = cmom && q2.mom;
        T lu_o = mom ? LF * lu2 : lu2;
        T lv_o = mom ? LF * lv2 : lv2;
        if (cint && q2.in) {
            lu_o -= q2.sa * dle_i * oS2;
            lv_o -= q2.sa * q2.dbe1 * oS2;
        }
        if (cint_e && q2.in) lu_o += q2.sa * dle_ip1 * oSe2;
        if (cint && q1.in) lv_o += q1.sa * q2.dbe1 * oS1;
        if (COR) {
            // UH(i,j) in VH'(k,l), l in {j, j-1}, k in {i, i+1} (U(0,.) twice at i = 0);
            // row j-1 is no momentum row when j = 0
            T acc = T(0);
            if (cmom && q2.mom) acc += mult_i * (HALF * q2.fv * (hJ2 > T(0) ? T(1) : T(0))) * (LF * lv2);
            if (cmom_e && q2.mom) acc += T(1) * (HALF * q2.fv * (hJe2 > T(0) ? T(1) : T(0))) * (LF * lve2);
            if (cmom && q3.mom) acc += mult_i * (HALF * q3.fv * (hJ3 > T(0) ? T(1) : T(0))) * (LF * lv3);
            if (cmom_e && q3.mom) acc += T(1) * (HALF * q3.fv * (hJe3 > T(0) ? T(1) : T(0))) * (LF * lve3);
            lu_o -= acc;
            // VH(i,j) in UH'(k,l), l in {j, j+1}, k in {i, i-1} (V(.,0) twice at j = 0)
            acc = T(0);
            const T mult_j = T(r == 0 ? 2 : 1);
            if (cmom && q2.mom) acc += mult_j * (HALF * q2.fu * (hI2 > T(0) ? T(1) : T(0))) * (LF * lu2);
            if (cmom_w && q2.mom) acc += mult_j * (HALF * q2.fu * (hIw2 > T(0) ? T(1) : T(0))) * (LF * luw2);
            if (cmom && q1.mom) acc += T(1) * (HALF * q1.fu * (hI1 > T(0) ? T(1) : T(0))) * (LF * lu1);
            if (cmom_w && q1.mom) acc += T(1) * (HALF * q1.fu * (hIw1 > T(0) ? T(1) : T(0))) * (LF * luw1);
            lv_o += acc;
        }
        if (own && r >= j0 && r < j1) {
            const long long ow = (long long)r * P + col;
            a.oS[ow] = oS2;
            a.oU[ow] = lu_o;
            a.oV[ow] = lv_o;
        }
        // slide up one row
        lv3 = lv2; lve3 = lve2; hJ3 = hJ2; hJe3 = hJe2;
        lu2 = lu1; lv2 = lv1; hI2 = hI1; hIw2 = hIw1; luw2 = luw1; oS2 = oS1; lve2 = lve1; hJ2 = hJ1; hJe2 = hJe1;
        aV2 = aV1;
        ls1 = ls0; lu1 = lu0; lv1 = lv0; e1 = e0; hI1 = hI0; hIw1 = hIw0; luw1 = luw0; aU1 = aU0; aUw1 = aUw0;
        q3 = q2; q2 = q1; q1 = q0;
    }
}

// Adds values[g] to lS at the gauge cells: one lane per gauge; the lane of a cell's
// first gauge adds the values of all gauges at that cell in gauge order (the order of
// one-by-one additions), the others skip.
template <typename T>
__global__ void adj_inject_kernel(T *lS, int n, const long long *idx, const double *values)
{
    const int g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= n) return;
    const long long c = idx[g];
    for (int h = 0; h < g; h++)
        if (idx[h] == c) return;
    T t = lS[c];
    for (int h = g; h < n; h++)
        if (idx[h] == c) t += (T)values[h];
    lS[c] = t;
}

template <typename T, bool COR>
static hipError_t launch_adj_march(const AdjArgs<T> &a0, hipStream_t s)
{
    static std::atomic<int> occ[5];
    constexpr int kBlock = 256;
    AdjArgs<T> a = a0;
    if (a.rows <= 0)
        a.rows = occupancy_rows((const void *)adj_march_kernel<T, COR>, occ, a.nx, a.ny, kBlock, kAdjOwn, 3);
    const int waves = (a.nx + kAdjOwn - 1) / kAdjOwn, wpb = kBlock / 64;
    dim3 grid((waves + wpb - 1) / wpb, (a.ny + a.rows - 1) / a.rows);
    hipLaunchKernelGGL((adj_march_kernel<T, COR>), grid, dim3(kBlock), 0, s, a);
    return hipGetLastError();
}

template <typename T>
hipError_t launch_adjoint_step(const AdjArgs<T> &a, hipStream_t s)
{
    if ((long long)a.pitch * a.ny >= (1LL << 31)) return hipErrorInvalidValue;
    return a.coriolis ? launch_adj_march<T, true>(a, s) : launch_adj_march<T, false>(a, s);
}

template <typename T>
hipError_t launch_adjoint_inject(T *lS, int n, const long long *idx, const double *values, hipStream_t s)
{
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(adj_inject_kernel<T>, dim3((n + 63) / 64), dim3(64), 0, s, lS, n, idx, values);
    return hipGetLastError();
}

template hipError_t launch_adjoint_step<double>(const AdjArgs<double> &, hipStream_t);
template hipError_t launch_adjoint_step<float>(const AdjArgs<float> &, hipStream_t);
template hipError_t launch_adjoint_inject<double>(double *, int, const long long *, const double *, hipStream_t);
template hipError_t launch_adjoint_inject<float>(float *, int, const long long *, const double *, hipStream_t);

}  // namespace swh
