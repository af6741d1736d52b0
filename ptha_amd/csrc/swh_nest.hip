// Multidomain nesting on one GPU: flux-integral tracking through the box edges,
// packing of send boxes straight into the receiver's buffer, unpacking, the
// staggered-grid wet/dry fix-ups, flux correction and the re-interpolation of
// fine receive boxes. Every kernel keeps the reference's FP64 operation order, so
// an FP64 run reproduces the reference's nesting arithmetic.
//
// Box geometry is 1-based and inclusive (the reference's send_inds / recv_inds);
// (i, j) 1-based addresses element (j-1)*pitch + (i-1) of a plane.
#include "swh_internal.h"

namespace swh {

namespace {

template <typename T>
__device__ __forceinline__ double nval(const NestDomPtr<T> &d, int k, int i, int j)
{
    const long long o = (long long)(j - 1) * d.pitch + (i - 1);
    switch (k) {
    case 1: return d.C ? (double)d.S[o] - (double)d.C[o] : (double)d.S[o];
    case 2: return (double)d.UH[o];
    case 3: return (double)d.VH[o];
    default: return (double)d.E[o];
    }
}

template <typename T>
__device__ __forceinline__ void nset(const NestDomPtr<T> &d, int k, int i, int j, double v)
{
    const long long o = (long long)(j - 1) * d.pitch + (i - 1);
    switch (k) {
    case 1:
        if (d.C) {
            // split v into the represented pair: stage = S - C
            const T s = (T)v;
            d.S[o] = s;
            d.C[o] = (T)((double)s - v);
        } else {
            d.S[o] = (T)v;
        }
        break;
    case 2: d.UH[o] = (T)v; break;
    case 3: d.VH[o] = (T)v; break;
    default: const_cast<T *>(d.E)[o] = (T)v; break;
    }
}

__device__ __forceinline__ int imin(int a, int b) { return a < b ? a : b; }
__device__ __forceinline__ int imax(int a, int b) { return a > b ? a : b; }
__device__ __forceinline__ double dmax(double a, double b) { return a > b ? a : b; }
__device__ __forceinline__ double dmin(double a, double b) { return a < b ? a : b; }

// Index of the descriptor whose block range holds block b (descriptors sorted by block0).
template <typename D>
__device__ __forceinline__ int find_desc(const D *ds, int n, int b)
{
    int lo = 0, hi = n - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (ds[mid].block0 <= b) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

}  // namespace

// ---- zero ----
__global__ void nest_zero_kernel(double *p, long long n)
{
    for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (long long)gridDim.x * blockDim.x)
        p[t] = 0.0;
}

template <typename T>
hipError_t launch_nest_zero(double *p, long long n, hipStream_t s)
{
    if (n <= 0) return hipSuccess;
    const long long nb = (n + 255) / 256;
    hipLaunchKernelGGL(nest_zero_kernel, dim3((unsigned)(nb < 1024 ? nb : 1024)), dim3(256), 0, s, p, n);
    return hipGetLastError();
}
template hipError_t launch_nest_zero<double>(double *, long long, hipStream_t);

// ---- flux integral: x(:, k) = x(:, k) + dt * flux(k) * edge (nested_grid_comms_mod.f90:2308-2461) ----
template <typename T>
__global__ __launch_bounds__(256) void nest_accum_kernel(const NestStrip *__restrict__ strips, const T *fNS,
                                                         const T *fEW, long long plane_ns, long long plane_ew,
                                                         int pitch, int nvar, double dt, double *integ)
{
    // one cell of one strip per thread (blockIdx.y picks the strip's chunk), so the
    // read-modify-writes of a strip are in flight together
    const NestStrip st = strips[blockIdx.x];
    const int r = blockIdx.y * blockDim.x + threadIdx.x;
    if (st.line < 0 || r >= st.n) return;
    const long long o = st.kind == 0 ? (long long)st.line * pitch + st.lo + r : (long long)(st.lo + r) * pitch + st.line;
    for (int k = 0; k < nvar; k++) {
        const double f = st.kind == 0 ? (double)fNS[k * plane_ns + o] : (double)fEW[k * plane_ew + o];
        double *x = integ + st.out + (long long)k * st.n + r;
        *x = *x + dt * f * st.edge;
    }
}

template <typename T>
hipError_t launch_nest_accum(const NestStrip *strips, int n_strips, int max_n, const T *fNS, const T *fEW,
                             long long plane_ns, long long plane_ew, int pitch, int nvar, double dt, double *integ,
                             hipStream_t s)
{
    if (n_strips <= 0 || max_n <= 0) return hipSuccess;
    hipLaunchKernelGGL(nest_accum_kernel<T>, dim3(n_strips, (max_n + 255) / 256), dim3(256), 0, s, strips, fNS, fEW,
                       plane_ns, plane_ew, pitch, nvar, dt, integ);
    return hipGetLastError();
}
template hipError_t launch_nest_accum<double>(const NestStrip *, int, int, const double *, const double *, long long,
                                              long long, int, int, double, double *, hipStream_t);
template hipError_t launch_nest_accum<float>(const NestStrip *, int, int, const float *, const float *, long long,
                                             long long, int, int, double, double *, hipStream_t);

// ---- pack (process_data_to_send_2022Feb) ----

// Fine-to-coarse value of one receive-grid cell (ic, jc) (0-based within the box) of
// variable k (1-based), nested_grid_comms_mod.f90:1529-1709.
template <typename T>
__device__ double pack_f2c(const NestDomPtr<T> &d, const NestPack &p, int ic, int jc, int k)
{
    const int r1 = p.inv_r[0], r2 = p.inv_r[1];
    const int i0 = p.iL + ic * r1, j0 = p.jL + jc * r2;   // first fine cell
    const int ci = r1 / 2, cj = r2 / 2;                     // central offsets
    const bool sv = (k == 1 || k == 4);
    if (p.my_stag == 1 && p.nbr_stag == 1) {
        if (sv) return nval(d, k, i0 + ci, j0 + cj);
        double s = 0.0;
        if (k == 2) {
            for (int jn = 0; jn < r2; jn++) s = s + nval(d, 2, i0 + r1 - 1, j0 + jn) * p.cr[1];
        } else {
            for (int in = 0; in < r1; in++) s = s + nval(d, 3, i0 + in, j0 + r2 - 1) * p.cr[0];
        }
        return s;
    }
    if (p.my_stag == 0 && p.nbr_stag == 0) return nval(d, k, i0 + ci, j0 + cj);
    if (p.my_stag == 0 && p.nbr_stag == 1) {
        if (sv) return nval(d, k, i0 + ci, j0 + cj);
        if (k == 2) {
            const int i = i0 + r1 - 1, j = j0 + cj, ip1 = imin(i + 1, d.nx);
            return 0.5 * (nval(d, 2, i, j) + nval(d, 2, ip1, j));
        }
        const int i = i0 + ci, j = j0 + r2 - 1, jp1 = imin(j + 1, d.ny);
        return 0.5 * (nval(d, 3, i, j) + nval(d, 3, i, jp1));
    }
    // staggered to non-staggered
    if (sv) return nval(d, k, i0 + ci, j0 + cj);
    double s = 0.0;
    if (k == 2) {
        const int i = i0 + ci, im1 = imax(i - 1, 1);
        for (int jn = 0; jn < r2; jn++)
            s = s + (0.5 * nval(d, 2, i, j0 + jn) + 0.5 * nval(d, 2, im1, j0 + jn)) * p.cr[1];
    } else {
        const int j = j0 + cj, jm1 = imax(j - 1, 1);
        for (int in = 0; in < r1; in++)
            s = s + (0.5 * nval(d, 3, i0 + in, j) + 0.5 * nval(d, 3, i0 + in, jm1)) * p.cr[0];
    }
    return s;
}

// Wet-dry gradient scale along one line of three depths (nested_grid_comms_mod.f90:1762-1791).
__device__ __forceinline__ double grad_scale3(double a, double b, double c, int n)
{
    double mx = a, mn = a;
    if (n > 1) { mx = dmax(mx, b); mn = dmin(mn, b); }
    if (n > 2) { mx = dmax(mx, c); mn = dmin(mn, c); }
    const double upper = 0.25, lower = 0.05;
    double g = 1.0;
    if (mn < upper * mx) g = dmax(mn / mx - lower, 0.0) / (upper - lower);
    if (mx < 0.0 || mn < 0.0) g = 0.0;
    return g;
}

// Coarse-to-fine: value and four one-sided gradients of cell (i, j) (1-based),
// nested_grid_comms_mod.f90:1720-1929.
template <typename T>
__device__ void pack_c2f(const NestDomPtr<T> &d, const NestPack &p, int i, int j, double *buf)
{
    const int iim1 = imax(1, i - 1), iip1 = imin(d.nx, i + 1);
    const int jjm1 = imax(1, j - 1), jjp1 = imin(d.ny, j + 1);
    double gx = 1.0, gy = 1.0;
    if (p.wetdry) {
        // maxval/minval over the row (iim1:iip1, j) and the column (i, jjm1:jjp1)
        double dx[3], dy[3];
        int nxl = 0, nyl = 0;
        for (int ii = iim1; ii <= iip1; ii++) dx[nxl++] = nval(d, 1, ii, j) - nval(d, 4, ii, j);
        for (int jj = jjm1; jj <= jjp1; jj++) dy[nyl++] = nval(d, 1, i, jj) - nval(d, 4, i, jj);
        gx = grad_scale3(dx[0], nxl > 1 ? dx[1] : 0.0, nxl > 2 ? dx[2] : 0.0, nxl);
        gy = grad_scale3(dy[0], nyl > 1 ? dy[1] : 0.0, nyl > 2 ? dy[2] : 0.0, nyl);
    }
    const int nI = p.iU - p.iL + 1, nJ = p.jU - p.jL + 1;
    for (int k = p.kL; k <= p.kU; k++) {
        double dip, dim, djp, djm;
        if (i > iim1 && i < iip1) {
            dip = nval(d, k, i + 1, j) - nval(d, k, i, j);
            dim = nval(d, k, i, j) - nval(d, k, i - 1, j);
        } else {
            dip = 0.0;
            if (i == iim1) dip = nval(d, k, i + 1, j) - nval(d, k, i, j);
            else if (i == iip1) dip = nval(d, k, i, j) - nval(d, k, i - 1, j);
            dim = dip;
        }
        if (j > jjm1 && j < jjp1) {
            djp = nval(d, k, i, j + 1) - nval(d, k, i, j);
            djm = nval(d, k, i, j) - nval(d, k, i, j - 1);
        } else {
            djp = 0.0;
            if (j == jjm1) djp = nval(d, k, i, j + 1) - nval(d, k, i, j);
            else if (j == jjp1) djp = nval(d, k, i, j) - nval(d, k, i, j - 1);
            djm = djp;
        }
        double uc = nval(d, k, i, j);
        if (p.my_stag == 1) {
            if (k == 2) {
                const double del = (p.nbr_stag == 1) ? 0.5 - p.inv_cr[0] * 0.5 : 0.5;
                uc = uc - dim * del;
                dip = dim;
                djp = 0.5 * (nval(d, k, i, jjp1) - nval(d, k, i, j)) + 0.5 * (nval(d, k, iim1, jjp1) - nval(d, k, iim1, j));
                djm = 0.5 * (nval(d, k, i, j) - nval(d, k, i, jjm1)) + 0.5 * (nval(d, k, iim1, j) - nval(d, k, iim1, jjm1));
            } else if (k == 3) {
                const double del = (p.nbr_stag == 1) ? 0.5 - p.inv_cr[1] * 0.5 : 0.5;
                uc = uc - djm * del;
                djp = djm;
                dip = 0.5 * (nval(d, k, iip1, j) - nval(d, k, i, j)) + 0.5 * (nval(d, k, iip1, jjm1) - nval(d, k, i, jjm1));
                dim = 0.5 * (nval(d, k, i, j) - nval(d, k, iim1, j)) + 0.5 * (nval(d, k, i, jjm1) - nval(d, k, iim1, jjm1));
            }
        }
        if (p.my_stag == 0 && p.nbr_stag == 1) {
            if (k == 2) uc = uc + 0.5 * p.inv_cr[0] * dip;
            else if (k == 3) uc = uc + 0.5 * p.inv_cr[1] * djp;
        }
        dip = dip * gx;
        djp = djp * gy;
        dim = dim * gx;
        djm = djm * gy;
        const long long off = ((long long)nI * nJ * (k - p.kL) + (long long)nI * (j - p.jL) + (i - p.iL)) *
                              (1 + kNestGradients);
        buf[off + 0] = uc;
        buf[off + 1] = dip;
        buf[off + 2] = dim;
        buf[off + 3] = djp;
        buf[off + 4] = djm;
    }
}

// Flux entry q of the packed box-edge integrals (nested_grid_comms_mod.f90:1940-2036).
__device__ double pack_flux(const NestPack &p, const double *integ, int q)
{
    for (int side = 0; side < 4; side++) {
        const int n = p.strip_n[side];
        int per, stride;
        if (p.mode == 0) { per = n; stride = 1; }
        else if (p.mode == 1) { stride = side < 2 ? p.inv_r[0] : p.inv_r[1]; per = n / stride; }
        else { stride = side < 2 ? p.r_ip[0] : p.r_ip[1]; per = n * stride; }
        const int cnt = per * p.nvar;
        if (q >= cnt) { q -= cnt; continue; }
        const int k = q / per, r = q - k * per;
        const double *x = integ + p.strip_out[side] + (long long)k * n;
        if (p.mode == 0) return x[r];
        if (p.mode == 1) {
            double s = 0.0;
            for (int m = 0; m < stride; m++) s = s + x[r * stride + m];
            return s;
        }
        const double inv_stride = 1.0 / stride;
        return 0.0 + x[r / stride] * inv_stride;
    }
    return 0.0;
}

template <typename T>
__global__ __launch_bounds__(256) void nest_pack_kernel(NestDoms<T> doms, const NestPack *__restrict__ packs,
                                                        int n_packs, double *const *integ, double *arena)
{
    const int pi = find_desc(packs, n_packs, blockIdx.x);
    const NestPack p = packs[pi];
    const NestDomPtr<T> &d = doms.d[p.dom];
    const int t = (blockIdx.x - p.block0) * blockDim.x + threadIdx.x;
    double *buf = arena + p.buf;
    const int n_items = p.mode == 2 ? (p.iU - p.iL + 1) * (p.jU - p.jL + 1) : p.n_data;
    if (t < n_items) {
        if (p.mode == 2) {
            const int nI = p.iU - p.iL + 1;
            pack_c2f(d, p, p.iL + t % nI, p.jL + t / nI, buf);
        } else {
            const int ii = t % p.iR, rest = t / p.iR, jj = rest % p.jR, kk = rest / p.jR;
            const int k = p.kL + kk;
            double v;
            if (p.mode == 0) v = nval(d, k, p.iL + ii, p.jL + jj);
            else v = pack_f2c(d, p, ii, jj, k);
            buf[((long long)jj + (long long)kk * p.jR) * p.iR + ii] = v;
        }
    } else if (t < n_items + p.n_flux) {
        buf[p.flux_start + (t - n_items)] = pack_flux(p, integ[p.dom], t - n_items);
    } else if (t == n_items + p.n_flux) {
        buf[p.time_index] = kMissingTime;
    }
}

template <typename T>
hipError_t launch_nest_pack(const NestDoms<T> &doms, const NestPack *packs, int n_packs, int n_blocks,
                            double *const *integ, double *arena, hipStream_t s)
{
    if (n_packs <= 0 || n_blocks <= 0) return hipSuccess;
    hipLaunchKernelGGL(nest_pack_kernel<T>, dim3(n_blocks), dim3(256), 0, s, doms, packs, n_packs, integ, arena);
    return hipGetLastError();
}
template hipError_t launch_nest_pack<double>(const NestDoms<double> &, const NestPack *, int, int, double *const *,
                                             double *, hipStream_t);
template hipError_t launch_nest_pack<float>(const NestDoms<float> &, const NestPack *, int, int, double *const *,
                                            double *, hipStream_t);

// ---- unpack (process_received_data) and flux error ----
template <typename T>
__global__ __launch_bounds__(256) void nest_unpack_kernel(NestDoms<T> doms, const NestUnpack *__restrict__ ups,
                                                          int n_ups, double *const *integ, double *const *err,
                                                          const double *arena)
{
    const int ui = find_desc(ups, n_ups, blockIdx.x);
    const NestUnpack u = ups[ui];
    const NestDomPtr<T> &d = doms.d[u.dom];
    const int t = (blockIdx.x - u.block0) * blockDim.x + threadIdx.x;
    const double *buf = arena + u.buf;
    const int nI = u.iU - u.iL + 1, nJ = u.jU - u.jL + 1, nK = u.kU - u.kL + 1;
    const int n_cells = nI * nJ * nK;
    if (t < n_cells) {
        const int ii = t % nI, rest = t / nI, jj = rest % nJ, kk = rest / nJ;
        double v;
        if (u.fine) {
            const int ci = ii / u.inv_r[0], cj = jj / u.inv_r[1];
            const long long b = ((long long)kk * u.nC + (long long)cj * u.niC + ci) * (1 + kNestGradients);
            const double uc = buf[b], dip = buf[b + 1], dim = buf[b + 2], djp = buf[b + 3], djm = buf[b + 4];
            const int li = ii - ci * u.inv_r[0] + 1, lj = jj - cj * u.inv_r[1] + 1;   // 1-based within the coarse cell
            const double dU_dj = (lj < u.mid[1]) ? djm : djp;
            const double dUj = ((double)lj - u.mid[1]) * u.cr[1] * dU_dj;
            const double dU_di = (li < u.mid[0]) ? dim : dip;
            const double dUi = ((double)li - u.mid[0]) * u.cr[0] * dU_di;
            v = uc + dUj + dUi;
        } else {
            v = buf[t];
        }
        nset(d, u.kL + kk, u.iL + ii, u.jL + jj, v);
        return;
    }
    // recv_box_flux_error = recv integral - received
    int q = t - n_cells;
    long long src = u.n_data;
    for (int side = 0; side < 4; side++) {
        const int cnt = u.strip_n[side] * u.nvar;
        if (q < cnt) {
            const long long o = u.strip_out[side] + q;
            err[u.dom][o] = integ[u.dom][o] - buf[src + q];
            return;
        }
        q -= cnt;
        src += cnt;
    }
}

template <typename T>
hipError_t launch_nest_unpack(const NestDoms<T> &doms, const NestUnpack *ups, int n_ups, int n_blocks,
                              double *const *integ, double *const *err, const double *arena, hipStream_t s)
{
    if (n_ups <= 0 || n_blocks <= 0) return hipSuccess;
    hipLaunchKernelGGL(nest_unpack_kernel<T>, dim3(n_blocks), dim3(256), 0, s, doms, ups, n_ups, integ, err, arena);
    return hipGetLastError();
}
template hipError_t launch_nest_unpack<double>(const NestDoms<double> &, const NestUnpack *, int, int,
                                               double *const *, double *const *, const double *, hipStream_t);
template hipError_t launch_nest_unpack<float>(const NestDoms<float> &, const NestUnpack *, int, int, double *const *,
                                              double *const *, const double *, hipStream_t);

// ---- staggered receivers: zero UH/VH across dry faces (domain_mod.f90:3741-3832) ----
template <typename T>
__global__ __launch_bounds__(256) void nest_stag_fix_kernel(NestDoms<T> doms, const NestUnpack *__restrict__ ups,
                                                            int n_ups)
{
    const int ui = find_desc(ups, n_ups, blockIdx.x);
    const NestUnpack u = ups[ui];
    if (!u.stag_fix) return;
    const NestDomPtr<T> &d = doms.d[u.dom];
    const int t = (blockIdx.x - u.block0) * blockDim.x + threadIdx.x;
    const int nI = u.iU - u.iL + 1, nJ = u.jU - u.jL + 1;
    if (t >= nI * nJ) return;
    const int ii = u.iL + t % nI, jj = u.jL + t / nI;
    const int ip1 = imin(ii + 1, d.nx), jp1 = imin(jj + 1, d.ny);
    if (u.stag_fix == 3) {
        // 'leapfrog_nonlinear' (domain_mod.f90:3834-3882): no outflow from dry cells, and
        // outflow from a wet cell into a dry neighbour only above the neighbour's elevation
        const double s0 = nval(d, 1, ii, jj);
        const bool dry0 = s0 - nval(d, 4, ii, jj) <= kMinDepth;
        const bool has_e = ip1 == ii + 1, has_n = jp1 == jj + 1;
        const double ee = nval(d, 4, ip1, jj), en = nval(d, 4, ii, jp1);
        const bool dry_e = has_e && nval(d, 1, ip1, jj) - ee <= kMinDepth;
        const bool dry_n = has_n && nval(d, 1, ii, jp1) - en <= kMinDepth;
        const double uh = nval(d, 2, ii, jj), vh = nval(d, 3, ii, jj);
        // each rule zeroes a positive or a negative value, so after one has fired no
        // later rule changes it
        const bool zu = (dry0 && uh > 0.0) || (dry_e && uh < 0.0) || (uh > 0.0 && dry_e && s0 <= ee + kMinDepth);
        const bool zv = (dry0 && vh > 0.0) || (dry_n && vh < 0.0) || (vh > 0.0 && dry_n && s0 <= en + kMinDepth);
        if (zu) nset(d, 2, ii, jj, 0.0);
        if (zv) nset(d, 3, ii, jj, 0.0);
        return;
    }
    bool zu, zv;
    if (u.stag_fix == 1) {
        const double lim = u.elev_lim;
        const bool here = nval(d, 4, ii, jj) >= lim;
        zu = here || (nval(d, 4, ip1, jj) >= lim);
        zv = here || (nval(d, 4, ii, jp1) >= lim);
    } else {
        const bool here = (nval(d, 1, ii, jj) - nval(d, 4, ii, jj)) <= kMinDepth;
        zu = here || ((nval(d, 1, ip1, jj) - nval(d, 4, ip1, jj)) <= kMinDepth);
        zv = here || ((nval(d, 1, ii, jp1) - nval(d, 4, ii, jp1)) <= kMinDepth);
    }
    if (zu) nset(d, 2, ii, jj, 0.0);
    if (zv) nset(d, 3, ii, jj, 0.0);
}

template <typename T>
hipError_t launch_nest_stag_fix(const NestDoms<T> &doms, const NestUnpack *ups, int n_ups, int n_blocks,
                                hipStream_t s)
{
    if (n_ups <= 0 || n_blocks <= 0) return hipSuccess;
    hipLaunchKernelGGL(nest_stag_fix_kernel<T>, dim3(n_blocks), dim3(256), 0, s, doms, ups, n_ups);
    return hipGetLastError();
}
template hipError_t launch_nest_stag_fix<double>(const NestDoms<double> &, const NestUnpack *, int, int, hipStream_t);
template hipError_t launch_nest_stag_fix<float>(const NestDoms<float> &, const NestUnpack *, int, int, hipStream_t);

// ---- flux correction (domain_mod.f90:2803-3226) ----
// One launch per group level: the updates of each domain are grouped (direction,
// then receive box, in the reference's loop order; consecutive passes with no cell in
// common share a group), the cells of one group are distinct, and level L holds every
// domain's L-th group, one cell per thread. Launches on one stream run in order, so
// each domain's groups apply one after the other.
template <typename T>
__global__ __launch_bounds__(256) void nest_fix_kernel(NestDoms<T> doms, const NestFixDom *__restrict__ fds,
                                                       const int2 *__restrict__ cells, int n)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const int2 c = cells[t];
    const NestFixDom fd = fds[c.x];
    const NestDomPtr<T> &d = doms.d[fd.dom];
    const NestFix f = fd.fix[c.y];
    for (int p = f.p0; p <= f.p1; p++) {
        const int i = f.is_ew ? p : f.line, j = f.is_ew ? f.line : p;   // 0-based
        bool go = fd.dry_all != 0;
        if (!go) go = nval(d, 1, i + 1, j + 1) >= nval(d, 4, i + 1, j + 1) + kMinDepth;
        if (!go) go = fd.pnp[(long long)j * d.nx + i] == 1;
        if (!go) continue;
        // all variables read before any is written, so their loads are in flight together
        double u[3];
        for (int k = f.var1; k <= f.varN; k++) {
            const double q_k = fd.err[f.err + (long long)k * f.stride] / f.area;
            u[k - f.var1] = nval(d, k + 1, i + 1, j + 1) + f.sgn * q_k;
        }
        for (int k = f.var1; k <= f.varN; k++) nset(d, k + 1, i + 1, j + 1, u[k - f.var1]);
    }
}

template <typename T>
hipError_t launch_nest_fix(const NestDoms<T> &doms, const NestFixDom *fds, const int2 *cells, int n, hipStream_t s)
{
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(nest_fix_kernel<T>, dim3((n + 255) / 256), dim3(256), 0, s, doms, fds, cells, n);
    return hipGetLastError();
}
template hipError_t launch_nest_fix<double>(const NestDoms<double> &, const NestFixDom *, const int2 *, int,
                                            hipStream_t);
template hipError_t launch_nest_fix<float>(const NestDoms<float> &, const NestFixDom *, const int2 *, int, hipStream_t);

// ---- re-interpolation of fine receive boxes (domain_mod.f90:3228-3468) ----
// blockIdx.x: the box; blockIdx.y: a chunk of its 4 sides x 4 variables x L copies,
// one per thread
template <typename T>
__global__ __launch_bounds__(256) void nest_reint_copy_kernel(NestDoms<T> doms, const NestReint *__restrict__ rs,
                                                              double *copy)
{
    const NestReint r = rs[blockIdx.x];
    const NestDomPtr<T> &d = doms.d[r.dom];
    const int L = r.len_ns > r.len_ew ? r.len_ns : r.len_ew;
    const int t = blockIdx.y * blockDim.x + threadIdx.x;
    if (t >= 16 * L) return;
    const int side = t / (4 * L), rem = t - side * 4 * L, k = rem / L, pos = rem - k * L;
    const int len = side < 2 ? r.len_ns : r.len_ew;
    if (pos >= len) return;
    const int i = side < 2 ? r.n0 + pos : r.ind[side];
    const int j = side < 2 ? r.ind[side] : r.m0 + pos;
    copy[r.copy + ((long long)side * 4 + k) * L + pos] = nval(d, k + 1, i, j);
}

template <typename T>
__global__ __launch_bounds__(256) void nest_reint_kernel(NestDoms<T> doms, const NestReint *__restrict__ rs, int n_rs,
                                                         const double *copy)
{
    const int ri = find_desc(rs, n_rs, blockIdx.x);
    const NestReint r = rs[ri];
    const NestDomPtr<T> &d = doms.d[r.dom];
    const int r1 = r.inv_r[0], r2 = r.inv_r[1];
    const int t = (blockIdx.x - r.block0) * blockDim.x + threadIdx.x;
    const int per = r1 * r2;
    if (t >= r.ncx * r.ncy * per) return;
    const int c = t / per, f = t - c * per;
    const int cx = c % r.ncx, cy = c / r.ncx;
    const int iL = r.n0 + cx * r1, jL = r.m0 + cy * r2;
    const int jc = jL + r2 / 2, ic = iL + r1 / 2;
    const int ii = iL + f % r1, jj = jL + f / r1;
    if (ii == ic && jj == jc) return;   // the centre keeps its value: U + 0 + 0
    const int nip = imin(ic + r1, d.nx), nim = imax(ic - r1, 1);
    const int njp = imin(jc + r2, d.ny), njm = imax(jc - r2, 1);
    const int L = r.len_ns > r.len_ew ? r.len_ns : r.len_ew;
    // copied strips: N = 0, S = 1, E = 2, W = 3
    auto cp = [&](int side, int k, int pos) { return copy[r.copy + ((long long)side * 4 + k - 1) * L + pos]; };
    const double d_i_j = nval(d, 1, ic, jc) - nval(d, 4, ic, jc);
    const double d_ip_j = nip < r.n1 ? nval(d, 1, nip, jc) - nval(d, 4, nip, jc) : cp(2, 1, jc - r.m0) - cp(2, 4, jc - r.m0);
    const double d_im_j = nim > r.n0 ? nval(d, 1, nim, jc) - nval(d, 4, nim, jc) : cp(3, 1, jc - r.m0) - cp(3, 4, jc - r.m0);
    const double lim = kMinDepth * 100.0;
    double gx, gy;
    {
        const double mx = dmax(d_i_j, dmax(d_ip_j, d_im_j)), mn = dmin(d_i_j, dmin(d_ip_j, d_im_j));
        if (mx <= lim) gx = 0.0;
        else {
            gx = dmax(0.0, (mn / mx - 0.1)) / (0.25 - 0.1);
            gx = dmin(1.0, gx);
        }
    }
    const double d_i_jp = njp < r.m1 ? nval(d, 1, ic, njp) - nval(d, 4, ic, njp) : cp(0, 1, ic - r.n0) - cp(0, 4, ic - r.n0);
    const double d_i_jm = njm > r.m0 ? nval(d, 1, ic, njm) - nval(d, 4, ic, njm) : cp(1, 1, ic - r.n0) - cp(1, 4, ic - r.n0);
    {
        const double mx = dmax(d_i_j, dmax(d_i_jp, d_i_jm)), mn = dmin(d_i_j, dmin(d_i_jp, d_i_jm));
        if (mx <= lim) gy = 0.0;
        else {
            gy = dmax(0.0, (mn / mx - 0.1)) / (0.25 - 0.1);
            gy = dmin(1.0, gy);
        }
    }
    for (int k = 1; k <= 4; k++) {
        const double uc = nval(d, k, ic, jc);
        double du_di, du_dj;
        if (ii < ic) {
            du_di = nim > r.n0 ? (uc - nval(d, k, nim, jc)) / (double)imax(ic - nim, 1)
                               : (uc - cp(3, k, jc - r.m0)) / (double)imax(ic - nim, 1);
        } else {
            du_di = nip < r.n1 ? (nval(d, k, nip, jc) - uc) / (double)imax(nip - ic, 1)
                               : (cp(2, k, jc - r.m0) - uc) / (double)imax(nip - ic, 1);
        }
        if (jj < jc) {
            du_dj = njm > r.m0 ? (uc - nval(d, k, ic, njm)) / (double)imax(jc - njm, 1)
                               : (uc - cp(1, k, ic - r.n0)) / (double)imax(jc - njm, 1);
        } else {
            du_dj = njp < r.m1 ? (nval(d, k, ic, njp) - uc) / (double)imax(njp - jc, 1)
                               : (cp(0, k, ic - r.n0) - uc) / (double)imax(njp - jc, 1);
        }
        nset(d, k, ii, jj, uc + (double)(ii - ic) * du_di * gx + (double)(jj - jc) * du_dj * gy);
    }
}

template <typename T>
hipError_t launch_nest_reint(const NestDoms<T> &doms, const NestReint *rs, int n_rs, int max_len, int n_blocks,
                             double *copy, hipStream_t s)
{
    if (n_rs <= 0 || max_len <= 0) return hipSuccess;
    hipLaunchKernelGGL(nest_reint_copy_kernel<T>, dim3(n_rs, (16 * max_len + 255) / 256), dim3(256), 0, s, doms, rs,
                       copy);
    if (n_blocks > 0)
        hipLaunchKernelGGL(nest_reint_kernel<T>, dim3(n_blocks), dim3(256), 0, s, doms, rs, n_rs, (const double *)copy);
    return hipGetLastError();
}
template hipError_t launch_nest_reint<double>(const NestDoms<double> &, const NestReint *, int, int, int, double *,
                                              hipStream_t);
template hipError_t launch_nest_reint<float>(const NestDoms<float> &, const NestReint *, int, int, int, double *,
                                             hipStream_t);

}  // namespace swh
