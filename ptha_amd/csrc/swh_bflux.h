// Compensated (double-double) accumulation, and the leapfrog boundary-flux sums
// (domain_mod_linear_solver_include.f90:126-165; domain_mod.f90:1803-1807) taken by
// extra workgroups of the step launch itself, so that they run beside the step's
// tiles instead of as launches of their own between steps.
#pragma once
#include "swh_internal.h"

namespace swh {

struct DD {
    double hi, lo;
};

__device__ __forceinline__ DD dd_add(DD a, double b)
{
    const double s = a.hi + b;
    const double bp = s - a.hi;
    const double err = (a.hi - (s - bp)) + (b - bp);
    return DD{s, a.lo + err};
}

__device__ __forceinline__ DD dd_add(DD a, DD b)
{
    DD r = dd_add(a, b.hi);
    r.lo += b.lo;
    return r;
}

// Boundary-flux lines of a grid with n_ext halo cells, in the order N, E, S, W
// (the order of boundary_flux_store): N and S are VH at rows ny-n_ext-1 and n_ext-1,
// columns n_ext..nx-n_ext-1; E and W are UH at columns nx-n_ext-1 and n_ext-1, rows
// n_ext..ny-n_ext-1.

// The line workgroups of one step launch: blockIdx.y < ey (of nlb = grid.x * ey
// workgroups, which the launch places ahead of its tiles). Workgroup k sums side k % 4
// (its k / 4'th share: each thread every (workgroups of the side * blockDim.x)'th cell)
// for L levels (level 0: the launch's input UH/VH; level 1, two-step launches: UH/VH
// after the first step), reduces over the workgroup (fixed shuffle tree, waves in
// order) and publishes its partial sums part[k][l] (hi, lo). The workgroup that
// arrives last adds, per side and level, the partials of that side's workgroups in
// order (one lane group per side and level, a fixed shuffle tree), then does, level by
// level, what boundary_flux_integral_tstep and the time integral do after a step:
// boundary_flux_store = the four sums times the edge lengths (S and W negated), evolve
// integral = dt * their sum, time integral += it, time = that step's end (NaN: from the
// device time, time + dt/2 + dt/2 as the host forms it). The arrival counter
// (DevScalars::bf_done) is 0 between launches. One cell in flight per thread keeps the
// registers within the step's own.
template <typename T, int L, typename Level1>
__device__ void bf_lines(const BfLines &b, int nx, int ny, int pitch, const T *__restrict__ UH,
                         const T *__restrict__ VH, Level1 level1)
{
    const int ne = b.n_ext;
    const int nlb = (int)gridDim.x * b.ey;
    const int k = blockIdx.x + gridDim.x * blockIdx.y;
    const int q = k & 3, r = k >> 2, nbq = (nlb - q + 3) >> 2;   // side, rank in it, its workgroups
    const bool ns = (q & 1) == 0;
    const int len = max(ns ? nx - 2 * ne : ny - 2 * ne, 0);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    DD acc[L];
#pragma unroll
    for (int l = 0; l < L; l++) acc[l] = DD{0.0, 0.0};
    for (int c = r * blockDim.x + threadIdx.x; c < len; c += nbq * blockDim.x) {
        const int i = ns ? ne + c : (q == 1 ? nx - ne - 1 : ne - 1);
        const int j = ns ? (q == 0 ? ny - ne - 1 : ne - 1) : ne + c;
        const long long o = (long long)j * pitch + i;
        acc[0] = dd_add(acc[0], (double)(ns ? VH[o] : UH[o]));
        if (L == 2) {
            T uh1, vh1;
            level1(i, j, uh1, vh1);
            acc[L - 1] = dd_add(acc[L - 1], (double)(ns ? vh1 : uh1));
        }
    }
    __shared__ double wsum[4][L][2];
    __shared__ int last;
#pragma unroll
    for (int l = 0; l < L; l++) {
        DD x = acc[l];
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) {
            const DD y{__shfl_down(x.hi, off), __shfl_down(x.lo, off)};
            x = dd_add(x, y);
        }
        if (lane == 0) {
            wsum[w][l][0] = x.hi;
            wsum[w][l][1] = x.lo;
        }
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double *part = b.part + (size_t)k * (2 * L);
#pragma unroll
        for (int l = 0; l < L; l++) {
            DD x{wsum[0][l][0], wsum[0][l][1]};
            for (int ww = 1; ww < (int)(blockDim.x >> 6); ww++) x = dd_add(x, DD{wsum[ww][l][0], wsum[ww][l][1]});
            __hip_atomic_store(part + 2 * l, x.hi, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            __hip_atomic_store(part + 2 * l + 1, x.lo, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
        // The partials and the counter are agent-scope atomics, which are performed at
        // the agent's coherence point (not in an XCD's L2): waiting for the partial
        // stores to complete orders them before the increment, and the last
        // workgroup's loads are issued after its increment has returned. This is the
        // ordering an agent-scope release/acquire gives atomics, without that fence's
        // L2 write-back and invalidation (which the step's tiles would pay for).
        asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
        const unsigned int done =
            __hip_atomic_fetch_add(&b.sc->bf_done, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        last = done == (unsigned int)(nlb - 1);
        if (last) __hip_atomic_store(&b.sc->bf_done, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    __syncthreads();
    if (!last) return;
    // lane groups of G (a power of two <= 64): group g = side + 4 * level; its lanes
    // take the side's workgroups rr = lane, lane + G, ... (workgroup k = side + 4 rr)
    __shared__ double st[L][4];
    int G = 64;
    while (G * 4 * L > (int)blockDim.x) G >>= 1;
    const int grp = threadIdx.x / G, gl = threadIdx.x % G;
    if (grp < 4 * L) {
        const int qq = grp & 3, ll = grp >> 2;
        const int nb = (nlb - qq + 3) >> 2;
        DD x{0.0, 0.0};
        for (int rr = gl; rr < nb; rr += G) {
            const double *p = b.part + (size_t)(qq + 4 * rr) * (2 * L) + 2 * ll;
            const double hi = __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            const double lo = __hip_atomic_load(p + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            x = dd_add(x, DD{hi, lo});
        }
        for (int off = G / 2; off > 0; off >>= 1) {
            const DD y{__shfl_down(x.hi, off, G), __shfl_down(x.lo, off, G)};
            x = dd_add(x, y);
        }
        if (gl == 0) st[ll][qq] = x.hi + x.lo;
    }
    __syncthreads();
    if (threadIdx.x != 0) return;
    DevScalars *sc = b.sc;
    const double m0 = b.dbe[ny - ne], m1 = b.dle[nx - ne], m2 = b.dbe[ne], m3 = b.dle[ne];
    double t = sc->time, bf_time = sc->bf_time, ev = 0.0, b0 = 0.0, b1 = 0.0, b2 = 0.0, b3 = 0.0;
#pragma unroll
    for (int l = 0; l < L; l++) {
        b0 = st[l][0] * m0;
        b1 = st[l][1] * m1;
        b2 = -st[l][2] * m2;
        b3 = -st[l][3] * m3;
        ev = (((b0 + b1) + b2) + b3) * b.dt;
        bf_time = bf_time + ev;
        const double ta = b.t_after[l];
        t = ta == ta ? ta : (t + b.dt * 0.5) + 0.5 * b.dt;
    }
    sc->bf_store[0] = b0;
    sc->bf_store[1] = b1;
    sc->bf_store[2] = b2;
    sc->bf_store[3] = b3;
    sc->bf_evolve = ev;
    sc->bf_time = bf_time;
    sc->time = t;
}

}  // namespace swh
