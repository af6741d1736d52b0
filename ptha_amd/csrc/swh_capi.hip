// C ABI of the gfx950 SWALS backend: handle lifetime, host <-> device copies and
// the per-step orchestration that domain%evolve_one_step performs in the reference
// (propagation/SWALS/src/shallow_water/domain_mod.f90:1679-1813).
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "swh_domain.h"

namespace swh {

static thread_local std::string g_err;

void set_error(const std::string &msg) { g_err = msg; }

int hip_fail(hipError_t e, const char *what, const char *file, int line)
{
    char buf[512];
    snprintf(buf, sizeof(buf), "HIP error %d (%s) in %s at %s:%d", (int)e, hipGetErrorString(e), what, file, line);
    g_err = buf;
    return e == hipErrorOutOfMemory ? SWH_ENOMEM : SWH_EHIP;
}

static int fail(int code, const std::string &msg)
{
    g_err = msg;
    return code;
}

}  // namespace swh

using namespace swh;


namespace {

int dmalloc(void **p, size_t bytes)
{
    SWH_HIP(hipMalloc(p, bytes));
    SWH_HIP(hipMemset(*p, 0, bytes));
    return SWH_OK;
}

void dfree(void *p)
{
    if (p) (void)hipFree(p);
}

// Host transfers go through one dense FP64 staging plane on the device; the conversion
// to the storage type and the pitch are applied there.
int stage_buffer(swh_domain *h)
{
    if (h->d_stage) return SWH_OK;
    return dmalloc((void **)&h->d_stage, sizeof(double) * (size_t)h->nx * h->ny);
}

// A dense FP64 plane (src; null: zeros) into the pitched plane dst; with comp (FP32
// state), the Kahan term that makes dst - comp the FP64 values. Returns when done.
template <typename T>
int upload_plane(swh_domain *h, void *dst, const double *src, void *comp = nullptr)
{
    int rc = stage_buffer(h);
    if (rc) return rc;
    if (src) SWH_HIP(hipMemcpy(h->d_stage, src, sizeof(double) * (size_t)h->nx * h->ny, hipMemcpyHostToDevice));
    SWH_HIP(launch_plane_from_dense<T>(src ? h->d_stage : nullptr, (T *)dst, (T *)comp, h->nx, h->ny, h->pitch,
                                       h->stream));
    SWH_HIP(hipStreamSynchronize(h->stream));
    return SWH_OK;
}

template <typename T>
int download_plane(swh_domain *h, const void *src, const void *comp, double *dst)
{
    int rc = stage_buffer(h);
    if (rc) return rc;
    SWH_HIP(launch_plane_to_dense<T>((const T *)src, (const T *)comp, h->d_stage, h->nx, h->ny, h->pitch, h->stream));
    SWH_HIP(hipStreamSynchronize(h->stream));
    SWH_HIP(hipMemcpy(dst, h->d_stage, sizeof(double) * (size_t)h->nx * h->ny, hipMemcpyDeviceToHost));
    return SWH_OK;
}

int upload_plane_any(swh_domain *h, void *dst, const double *src, void *comp = nullptr)
{
    return h->prec == SWH_PREC_FP64 ? upload_plane<double>(h, dst, src)
                                    : upload_plane<float>(h, dst, src, comp);
}

// The FP32K stage to upload where depth is stage - elevation (every method but the
// truely-linear leapfrog ones): the elevation is stored rounded to FP32, so the FP64
// stage itself would leave a dry cell (stage = elevation) a depth of plus or minus that
// rounding. A cell dry in FP64 (stage <= elevation) takes the FP32 elevation as its stage
// (depth exactly 0, as the reference's); a wet one keeps its FP64 stage, at least the
// FP32 elevation (rounding is monotonic, so no depth turns negative). elev: FP64
// elevation (nx * ny); stage null: zeros.
std::vector<double> stage_on_fp32_elevation(const swh_domain *h, const double *stage, const double *elev)
{
    const size_t n = (size_t)h->nx * h->ny;
    std::vector<double> t(n);
    for (size_t k = 0; k < n; k++) {
        const double s = stage ? stage[k] : 0.0, e = elev[k], e32 = (double)(float)e;
        t[k] = s <= e ? e32 : std::max(s, e32);
    }
    return t;
}

// friction_work for the truely-linear leapfrog + friction solver, built once on the
// host in FP64 (domain_mod_leapfrog_solver_friction_include.f90:21-56).
int build_friction_work(swh_domain *h)
{
    if (h->method != SWH_METHOD_LLPNF || !h->truely_linear) return SWH_OK;
    const int nx = h->nx, ny = h->ny;
    const double power = (h->friction_type == 1) ? -2.0 : -7.0 / 3.0;
    std::vector<double> fu((size_t)nx * ny), fv((size_t)nx * ny);
    const double msl = h->msl;
    const double *e = h->h_elev.data(), *n2 = h->h_nsq.data();
    for (int j = 0; j < ny; j++)
        for (int i = 0; i < nx; i++) {
            const int ip1 = std::min(i + 1, nx - 1), jp1 = std::min(j + 1, ny - 1);
            const size_t k = (size_t)j * nx + i;
            double depth_iph = 0.5 * (msl - e[k] + msl - e[(size_t)j * nx + ip1]);
            double depth_jph = 0.5 * (msl - e[k] + msl - e[(size_t)jp1 * nx + i]);
            depth_iph = std::max(depth_iph, kMinDepth);
            depth_jph = std::max(depth_jph, kMinDepth);
            const double nsq_iph = 0.5 * (n2[k] + n2[(size_t)j * nx + ip1]);
            const double nsq_jph = 0.5 * (n2[k] + n2[(size_t)jp1 * nx + i]);
            fu[k] = kGravity * nsq_iph * std::pow(depth_iph, power);
            fv[k] = kGravity * nsq_jph * std::pow(depth_jph, power);
        }
    int rc = upload_plane_any(h, h->fwU, fu.data());
    if (rc) return rc;
    rc = upload_plane_any(h, h->fwV, fv.data());
    if (rc) return rc;
    h->fw_ready = 1;
    return SWH_OK;
}

// Leapfrog per-row coefficients for time step dt, in FP64 with the reference's
// operation order, then converted to the storage type.
template <typename T>
int build_rc(swh_domain *h, double dt)
{
    const int nx = h->nx, ny = h->ny;
    const bool nl = h->method == SWH_METHOD_LEAPFROG_NONLINEAR;
    std::vector<T> buf((size_t)(nl ? 6 : 5) * ny + (nx + 1) + (ny + 1));
    T *sa = buf.data(), *cu = sa + ny, *cv = cu + ny, *fu = cv + ny, *fv = fu + ny, *dle = fv + ny,
      *dbe = dle + nx + 1, *sv = dbe + ny + 1;
    for (int j = 0; j < ny; j++) {
        sa[j] = (T)(dt / h->area[j]);
        // dt over the VH-point area (the nonlinear momentum update,
        // domain_mod_leapfrog_nonlinear_include.f90:575-783)
        if (nl) sv[j] = (j + 1 < ny) ? (T)(dt / (0.5 * (h->area[j] + h->area[j + 1]))) : T(0);
        cu[j] = (T)(kGravity * dt / h->area[j]);
        cv[j] = (j + 1 < ny) ? (T)(kGravity * dt / (0.5 * (h->area[j] + h->area[j + 1]))) : T(0);
        fu[j] = h->coriolis ? (T)(dt * h->cor[j] * 0.5) : T(0);
        fv[j] = h->coriolis ? (T)(dt * h->corbe[j + 1] * 0.5) : T(0);
    }
    for (int i = 0; i <= nx; i++) dle[i] = (T)h->dle[i];
    for (int j = 0; j <= ny; j++) dbe[j] = (T)h->dbe[j];
    SWH_HIP(hipStreamSynchronize(h->stream));
    SWH_HIP(hipMemcpy(h->rc, buf.data(), buf.size() * sizeof(T), hipMemcpyHostToDevice));
    h->rc_dt = dt;
    return SWH_OK;
}

template <typename T>
LFRowCoef<T> rc_view(swh_domain *h)
{
    const int nx = h->nx, ny = h->ny;
    const T *b = (const T *)h->rc;
    LFRowCoef<T> r;
    r.sa = b;
    r.cu = b + ny;
    r.cv = b + 2 * ny;
    r.fu = b + 3 * ny;
    r.fv = b + 4 * ny;
    r.dle = b + 5 * ny;
    r.dbe = b + 5 * ny + nx + 1;
    return r;
}

template <typename T>
int maxq_now(swh_domain *h, const DevScalars *tsrc, double t)
{
    if (!h->maxq_mask) return SWH_OK;
    const int c = h->cur;
    SWH_HIP(launch_maxq<T>((const T *)h->S[c], (const T *)h->C[c], (const T *)h->UH[c], (const T *)h->VH[c],
                           (const T *)h->E, h->nx, h->ny, h->pitch, h->leapfrog() ? 1 : 0, h->tl_solver() ? 1 : 0,
                           (T)h->msl, h->maxq_mask, (T *)h->mx, tsrc, (T)t, (T)(h->msl + h->arrival_thresh),
                           h->stream));
    return SWH_OK;
}

// Fraction of a forcing patch's work applied in a step that ends at `time` and lasts
// `dt` (add_forcing_work_to_U_over_time, forcing_mod.f90:101-120; constant rate).
double forcing_fraction(const swh_domain *h, double time, double dt)
{
    if (!h->f_on) return 0.0;
    if (!((time - dt) <= h->f_end && time >= h->f_start)) return 0.0;
    if (h->f_start == h->f_end) return 1.0;
    return (std::min(time, h->f_end) - std::max(time - dt, h->f_start)) / (h->f_end - h->f_start);
}

// The step's boundary-flux sums in line workgroups of the step launch (of block_x
// threads, a workgroup per block_x cells of each side: one cell per thread);
// t_after1/2: the end times of the launch's steps (NaN: from the device time).
BfLines bf_args(const swh_domain *h, int block_x, double dt, double t_after1, double t_after2)
{
    BfLines b;
    const int lmax = std::max(std::max(h->nx, h->ny) - 2 * h->n_ext, 1);
    b.want = 4 * ((lmax + block_x - 1) / block_x);
    b.cap = h->bf_cap;
    b.n_ext = h->n_ext;
    b.dle = h->d_dle;
    b.dbe = h->d_dbe;
    b.sc = h->sc;
    b.part = h->d_bfpart;
    b.dt = dt;
    b.t_after[0] = t_after1;
    b.t_after[1] = t_after2;
    return b;
}

template <typename T, bool KAHAN>
int leapfrog_step(swh_domain *h, double dt)
{
    if (h->rc_dt != dt) {
        int rc = build_rc<T>(h, dt);
        if (rc) return rc;
    }
    const double time0 = h->time;
    if (h->nsteps == 0) {
        int rc = maxq_now<T>(h, nullptr, h->time);
        if (rc) return rc;
    }
    const int c = h->cur, o = 1 - c;
    const double t_half = h->time + dt * 0.5;
    const double t_after = t_half + 0.5 * dt;
    // nesting flux integrals: dt * (UH, VH) before the update, mass only
    // (domain_mod_linear_solver_include.f90:74-77)
    if (h->n_nest_strips)
        SWH_HIP(launch_nest_accum<T>(h->d_nest_strips, h->n_nest_strips, h->nest_strip_max, (const T *)h->VH[c],
                                     (const T *)h->UH[c], 0, 0, h->pitch, 1, dt, h->d_nest_int, h->stream));
    // Flather and transmissive boundaries: they need the new stage inside the edges before
    // the momentum half-step, and change UH/VH on the edges that the momentum half-step reads
    const bool flather = h->bc_mask && (h->bc_kind == 1 || h->bc_kind == 2);
    // boundary fluxes read the pre-momentum UH/VH (include.f90:126-165)
    // (under graph capture the kernel forms the end time from the device time: NaN)
    const double t_after_arg = h->graph_time ? __builtin_nan("") : t_after;
    LFArgs<T> a;
    a.nx = h->nx; a.ny = h->ny; a.pitch = h->pitch; a.src_stride = 0;
    a.S = (const T *)h->S[c]; a.C = (const T *)h->C[c]; a.UH = (const T *)h->UH[c]; a.VH = (const T *)h->VH[c];
    a.E = (const T *)h->E;
    a.So = (T *)h->S[o]; a.Co = (T *)h->C[o]; a.UHo = (T *)h->UH[o]; a.VHo = (T *)h->VH[o];
    a.rc = rc_view<T>(h);
    a.fwU = (const T *)h->fwU; a.fwV = (const T *)h->fwV; a.nsq = (const T *)h->nsq;
    a.msl = (T)h->msl; a.dt = (T)dt; a.dt_lfc = (T)(dt * h->lfc);
    a.lin_fric_mult = (T)(1.0 / (1.0 + dt * h->lfc));
    a.fric_power = (T)((h->friction_type == 1) ? -2.0 : -7.0 / 3.0);
    a.bc_mask = h->bc_mask;
    // the boundary stage: the sine function, or msl_linear (radiation without a function)
    const double bc_stage = h->bc_use_fun
                                ? h->bc_msl + h->bc_amp * std::sin(2.0 * std::acos(-1.0) * t_half / h->bc_period)
                                : h->msl;
    a.bc_stage = (T)bc_stage;
    a.rows_per_block = h->rows_per_block;
    a.src_fast = 0;
    const bool upd = h->maxq_mask && ((h->nsteps % h->maxq_freq) == 0);
    a.maxq_mask = upd ? h->maxq_mask : 0;
    a.mx = (T *)h->mx;
    a.t_after = (T)t_after;
    a.arrival_stage = (T)(h->msl + h->arrival_thresh);
    int fric_mode = 0;
    if (h->method == SWH_METHOD_LLPNF) fric_mode = h->truely_linear ? 1 : 2;
    if (fric_mode == 1 && !h->fw_ready) return fail(SWH_ESTATE, "friction_work not built: upload U first");
    // forcing applied after the stage half-step, with time at its end (include.f90:117-119)
    const double frac = forcing_fraction(h, t_half, dt);
    if (flather) {
        // stage half-step alone into the output planes, then the boundary
        // (include.f90:92-119: stage, update_boundary, apply_forcing), then the
        // boundary fluxes and the momentum half-step from the edited planes
        a.bc_mask = 0;
        SWH_HIP(launch_lf_stage<T, KAHAN>(a, h->stream));
        T *Cb = KAHAN ? (T *)h->C[o] : nullptr;
        if (h->bc_kind == 1)
            SWH_HIP(launch_flather<T>((T *)h->S[o], Cb, (T *)h->UH[c], (T *)h->VH[c], (const T *)h->E, h->nx, h->ny,
                                      h->pitch, h->bc_mask, h->msl, h->bc_use_fun, h->bc_amp, h->bc_period, h->bc_msl,
                                      bc_stage, nullptr, h->stream));
        else
            SWH_HIP(launch_edge_bc<T>((T *)h->S[o], Cb, (T *)h->UH[c], (T *)h->VH[c], (T *)h->E, h->nx, h->ny,
                                      h->pitch, h->bc_mask, h->bc_kind, h->bc_use_fun, h->bc_amp, h->bc_period,
                                      h->bc_msl, h->msl, bc_stage, 1, nullptr, h->stream));
        if (frac != 0.0 && h->fS)
            SWH_HIP(launch_add_scaled_kahan<T>((T *)h->S[o], Cb, (const T *)h->fS, h->pitch, h->f_i0, h->f_i1,
                                               h->f_j0, h->f_j1, (T)frac, h->stream));
        if (frac != 0.0 && h->fE)
            SWH_HIP(launch_add_scaled<T>((T *)h->E, (const T *)h->fE, h->pitch, h->f_i0, h->f_i1, h->f_j0, h->f_j1,
                                         (T)frac, h->stream));
        a.S = (const T *)h->S[o];
        a.C = (const T *)h->C[o];
        a.stage_given = 1;
        a.bf = bf_args(h, h->block_x, dt, t_after_arg, 0.0);
        SWH_HIP(launch_leapfrog<T, KAHAN>(a, h->coriolis, fric_mode, h->truely_linear, h->lfc != 0.0, 1, h->block_x,
                                          h->variant, h->stream));
    } else {
        if (frac != 0.0) {
            a.fS = (const T *)h->fS; a.fE = (const T *)h->fE; a.f_frac = (T)frac;
        }
        a.bf = bf_args(h, h->block_x, dt, t_after_arg, 0.0);
        SWH_HIP(launch_leapfrog<T, KAHAN>(a, h->coriolis, fric_mode, h->truely_linear, h->lfc != 0.0, 1, h->block_x,
                                          h->variant, h->stream));
        if (frac != 0.0 && h->fE)
            SWH_HIP(launch_add_scaled<T>((T *)h->E, (const T *)h->fE, h->pitch, h->f_i0, h->f_i1, h->f_j0, h->f_j1,
                                         (T)frac, h->stream));
    }
    h->cur = o;
    h->time = t_after;
    h->dt_last_update = dt;
    h->evolve_step_dt = h->time - time0;
    h->nsteps++;
    return SWH_OK;
}

// 'leapfrog_nonlinear' step (domain_mod_leapfrog_nonlinear_include.f90:61-826): the
// depth before the step, the stage half-step alone, the boundary, forcing and boundary
// fluxes as in the linear solver, then the nonlinear momentum passes into the spare
// planes. max_U after the step on the update cadence.
template <typename T, bool KAHAN>
int nonlinear_step(swh_domain *h, double dt)
{
    if (h->rc_dt != dt) {
        int rc = build_rc<T>(h, dt);
        if (rc) return rc;
    }
    const double time0 = h->time;
    if (h->nsteps == 0) {
        int rc = maxq_now<T>(h, nullptr, h->time);
        if (rc) return rc;
    }
    const int c = h->cur, o = 1 - c;
    const double t_half = h->time + dt * 0.5;
    const double t_after = t_half + 0.5 * dt;
    // nesting flux integrals: dt * (UH, VH) before the update, mass only
    // (domain_mod_leapfrog_nonlinear_include.f90:93-102)
    if (h->n_nest_strips)
        SWH_HIP(launch_nest_accum<T>(h->d_nest_strips, h->n_nest_strips, h->nest_strip_max, (const T *)h->VH[c],
                                     (const T *)h->UH[c], 0, 0, h->pitch, 1, dt, h->d_nest_int, h->stream));
    const size_t P = h->plane_elems();
    T *w = (T *)h->nlw;
    NLArgs<T> n;
    n.nx = h->nx; n.ny = h->ny; n.pitch = h->pitch;
    n.S0 = (const T *)h->S[c]; n.S = (const T *)h->S[o]; n.E = (const T *)h->E;
    n.UH = (const T *)h->UH[c]; n.VH = (const T *)h->VH[c]; n.UHo = (T *)h->UH[o]; n.VHo = (T *)h->VH[o];
    n.nsq = (const T *)h->nsq;
    n.DL = w; n.UUH = w + P; n.UVH = w + 2 * P; n.VUH = w + 3 * P; n.VVH = w + 4 * P;
    n.fEWu = w + 5 * P; n.fEWv = w + 6 * P; n.fNSu = w + 7 * P; n.fNSv = w + 8 * P;
    n.dbe = h->d_dbe; n.dle = h->d_dle; n.area = h->d_area; n.cor = h->d_cor; n.corbe = h->d_corbe;
    n.tanlat = h->d_tanlat;
    {
        const LFRowCoef<T> q = rc_view<T>(h);
        n.ica = q.sa;                       // (T)(dt / area(j))
        n.ica_vh = q.dbe + h->ny + 1;       // (T)(dt / (0.5 (area(j) + area(j+1))))
    }
    n.dt = (T)dt; n.lfc = (T)h->lfc;
    n.fric_power = (T)((h->friction_type == 1) ? -2.0 : -7.0 / 3.0);
    n.beta = T(1);            // advection_beta (global_mod.f90:38)
    n.froude_limit = T(10);   // leapfrog_froude_limit (domain_mod.f90:210)
    n.coriolis = h->coriolis; n.spherical = h->spherical;
    // the fused flux pass forms the lagged depth from the old and new stage, which needs
    // the elevation unchanged through the step; with elevation forcing, the depth plane
    const double frac = forcing_fraction(h, t_half, dt);
    // elevation forcing or a transmissive boundary (which copies the elevation) change the
    // elevation within the step
    const bool elev_forced = (frac != 0.0 && h->fE) || (h->bc_mask && h->bc_kind == 2);
    if (elev_forced) SWH_HIP(launch_nonlinear_depth0<T>(n, h->stream));

    LFArgs<T> a;
    a.nx = h->nx; a.ny = h->ny; a.pitch = h->pitch; a.src_stride = 0;
    a.S = (const T *)h->S[c]; a.C = (const T *)h->C[c]; a.UH = (const T *)h->UH[c]; a.VH = (const T *)h->VH[c];
    a.E = (const T *)h->E;
    a.So = (T *)h->S[o]; a.Co = (T *)h->C[o]; a.UHo = (T *)h->UH[o]; a.VHo = (T *)h->VH[o];
    a.rc = rc_view<T>(h);
    const bool flather = h->bc_mask && (h->bc_kind == 1 || h->bc_kind == 2);
    const double bc_stage = h->bc_use_fun
                                ? h->bc_msl + h->bc_amp * std::sin(2.0 * std::acos(-1.0) * t_half / h->bc_period)
                                : h->msl;
    a.bc_mask = flather ? 0 : h->bc_mask;
    a.bc_stage = (T)bc_stage;
    a.maxq_mask = 0;
    // the boundary fluxes read the pre-momentum UH/VH (include.f90:149-165): without a
    // Flather or edge boundary (which change them on the edges first) nothing changes
    // them before the momentum passes, so line workgroups of the stage launch take them
    // (under graph capture the end time from the device time: NaN)
    const double t_after_arg = h->graph_time ? __builtin_nan("") : t_after;
    if (!flather) a.bf = bf_args(h, 256, dt, t_after_arg, 0.0);
    SWH_HIP(launch_lf_stage<T, KAHAN>(a, h->stream));
    T *Cb = KAHAN ? (T *)h->C[o] : nullptr;
    if (flather && h->bc_kind == 1)
        SWH_HIP(launch_flather<T>((T *)h->S[o], Cb, (T *)h->UH[c], (T *)h->VH[c], (const T *)h->E, h->nx, h->ny,
                                  h->pitch, h->bc_mask, h->msl, h->bc_use_fun, h->bc_amp, h->bc_period, h->bc_msl,
                                  bc_stage, nullptr, h->stream));
    else if (flather)
        SWH_HIP(launch_edge_bc<T>((T *)h->S[o], Cb, (T *)h->UH[c], (T *)h->VH[c], (T *)h->E, h->nx, h->ny, h->pitch,
                                  h->bc_mask, h->bc_kind, h->bc_use_fun, h->bc_amp, h->bc_period, h->bc_msl, h->msl,
                                  bc_stage, 1, nullptr, h->stream));
    if (frac != 0.0 && h->fS)
        SWH_HIP(launch_add_scaled_kahan<T>((T *)h->S[o], Cb, (const T *)h->fS, h->pitch, h->f_i0, h->f_i1, h->f_j0,
                                           h->f_j1, (T)frac, h->stream));
    if (frac != 0.0 && h->fE)
        SWH_HIP(launch_add_scaled<T>((T *)h->E, (const T *)h->fE, h->pitch, h->f_i0, h->f_i1, h->f_j0, h->f_j1,
                                     (T)frac, h->stream));
    if (flather)
        SWH_HIP(launch_bflux_leapfrog<T>((const T *)h->UH[c], (const T *)h->VH[c], h->nx, h->ny, h->pitch,
                                         h->n_ext, h->d_dle, h->d_dbe, dt, t_after_arg, h->sc, h->stream));
    SWH_HIP(launch_nonlinear_momentum<T>(n, !elev_forced, h->stream));
    h->cur = o;
    h->time = t_after;
    h->dt_last_update = dt;
    h->evolve_step_dt = h->time - time0;
    h->nsteps++;
    if (h->maxq_mask && ((h->nsteps - 1) % h->maxq_freq) == 0) {
        int rc = maxq_now<T>(h, nullptr, h->time);
        if (rc) return rc;
    }
    return SWH_OK;
}

// Two leapfrog steps in one launch (swh_leapfrog2.hip) for a single grid: possible
// where no stage boundary, nesting strips or per-step friction work are involved and
// max_U is not due after the first of the two steps. The boundary fluxes of the second
// step need UH/VH after the first on the four boundary-flux lines: a small kernel
// forms them there (same per-cell operations) into the output planes, the flux pass
// reads them, and the two-step launch then overwrites the planes.
bool two_step_ok(const swh_domain *h, double dt)
{
    if (!h->two_step || !h->leapfrog() || h->bc_mask || h->n_nest_strips) return false;
    if (h->method == SWH_METHOD_LEAPFROG_NONLINEAR) return false;
    if (h->f_on) {
        // no forcing in either step of the pair
        const double t_half1 = h->time + dt * 0.5, t_after1 = t_half1 + 0.5 * dt;
        if (forcing_fraction(h, t_half1, dt) != 0.0 || forcing_fraction(h, t_after1 + dt * 0.5, dt) != 0.0)
            return false;
    }
    if (h->method == SWH_METHOD_LLPNF && !h->truely_linear) return false;
    if (h->maxq_mask && (h->nsteps % h->maxq_freq) == 0) return false;
    if ((long long)h->pitch * h->ny >= (1LL << 31)) return false;
    return true;
}

// Two leapfrog steps in one launch. mid_sample: also record a gauge sample of the state
// after the first step (taken one step ahead from the current state, with the same
// operations, before the launch).
template <typename T, bool KAHAN>
int leapfrog_step2(swh_domain *h, double dt, bool mid_sample = false)
{
    if (h->rc_dt != dt) {
        int rc = build_rc<T>(h, dt);
        if (rc) return rc;
    }
    int fric_mode = 0;
    if (h->method == SWH_METHOD_LLPNF) fric_mode = 1;
    if (fric_mode == 1 && !h->fw_ready) return fail(SWH_ESTATE, "friction_work not built: upload U first");
    const int c = h->cur, o = 1 - c;
    const double t_half1 = h->time + dt * 0.5, t_after1 = t_half1 + 0.5 * dt;
    const double t_half2 = t_after1 + dt * 0.5, t_after2 = t_half2 + 0.5 * dt;
    LFArgs<T> a;
    a.nx = h->nx; a.ny = h->ny; a.pitch = h->pitch; a.src_stride = 0;
    a.S = (const T *)h->S[c]; a.C = (const T *)h->C[c]; a.UH = (const T *)h->UH[c]; a.VH = (const T *)h->VH[c];
    a.E = (const T *)h->E;
    a.So = (T *)h->S[o]; a.Co = (T *)h->C[o]; a.UHo = (T *)h->UH[o]; a.VHo = (T *)h->VH[o];
    a.rc = rc_view<T>(h);
    a.fwU = (const T *)h->fwU; a.fwV = (const T *)h->fwV; a.nsq = (const T *)h->nsq;
    a.msl = (T)h->msl; a.dt = (T)dt; a.dt_lfc = (T)(dt * h->lfc);
    a.lin_fric_mult = (T)(1.0 / (1.0 + dt * h->lfc));
    a.fric_power = (T)((h->friction_type == 1) ? -2.0 : -7.0 / 3.0);
    a.bc_mask = 0;
    a.bc_stage = T(0);
    a.rows_per_block = h->rows2;
    a.src_fast = 0;
    a.maxq_mask = 0;
    a.mx = (T *)h->mx;
    a.t_after = (T)t_after2;
    a.arrival_stage = (T)(h->msl + h->arrival_thresh);
    const int lin = h->lfc != 0.0;
    // both steps' boundary fluxes in line workgroups of the launch: the first step's from
    // the current UH/VH, the second's from UH/VH one step ahead on the lines
    a.bf = bf_args(h, h->block2, dt, t_after1, t_after2);
    if (mid_sample) {
        if (!h->d_gidx || h->ng < 1) return fail(SWH_ESTATE, "gauges not set");
        if (h->gsamples >= h->gmax) return fail(SWH_ESTATE, "gauge sample buffer full");
        SWH_HIP(launch_gauge_step1<T, KAHAN>(a, h->coriolis, fric_mode, h->truely_linear, lin, 1, h->ng, h->d_gidx,
                                             h->d_gval, h->gmax, h->gsamples, h->d_gtime, t_after1, h->stream));
        h->gsamples++;
    }
    SWH_HIP(launch_leapfrog2<T, KAHAN>(a, h->coriolis, fric_mode, h->truely_linear, lin, 1, h->block2, h->stream));
    h->cur = o;
    h->time = t_after2;
    h->dt_last_update = dt;
    h->evolve_step_dt = t_after2 - t_after1;
    h->nsteps += 2;
    // max_U due after the second step (not after the first: two_step_ok)
    if (h->maxq_mask && ((h->nsteps - 1) % h->maxq_freq) == 0) {
        const int rc = maxq_now<T>(h, nullptr, h->time);
        if (rc) return rc;
    }
    return SWH_OK;
}

template <typename T>
FVArgs<T> fv_args(swh_domain *h, int dst, int src)
{
    FVArgs<T> a;
    a.nx = h->nx; a.ny = h->ny; a.pitch = h->pitch; a.fpitch = h->fpitch;
    a.S = (T *)h->S[dst]; a.UH = (T *)h->UH[dst]; a.VH = (T *)h->VH[dst];
    a.S0 = (const T *)h->S[src]; a.UH0 = (const T *)h->UH[src]; a.VH0 = (const T *)h->VH[src];
    a.C0 = (const T *)h->C[src]; a.Cw = (T *)h->C[dst];   // FP32: compensated stage (null in FP64)
    a.E = (const T *)h->E;
    a.Sq = nullptr; a.UHq = nullptr; a.VHq = nullptr;
    a.bmax = h->d_bmax;
    a.fNS = (T *)h->fNS; a.fEW = (T *)h->fEW;
    a.esU = h->esU; a.esV = h->esV; a.esJ = h->esJ; a.esE = h->esE;
    a.nsq = (const T *)h->nsq;
    a.dbe = h->d_dbe; a.dle = h->d_dle; a.area = h->d_area; a.cor = h->d_cor; a.tanlat = h->d_tanlat;
    a.theta = h->theta; a.half_cfl = 0.5 * h->cfl; a.lfc = h->lfc;
    a.fric_power = (h->friction_type == 1) ? -2.0 : -7.0 / 3.0;
    a.spherical = h->spherical; a.coriolis = h->coriolis; a.n_ext = h->n_ext;
    a.sc = h->sc;
    a.skip_edges = h->disp_on;
    a.flux_rows = h->rows_per_block;
    return a;
}

// One Euler stage of the finite-volume schemes: fluxes from state `src`, update into
// `dst` (from `qsrc`'s values when given, else from the clipped `src`), then the
// nesting flux integrals (`nest_dt` times the pre-multiplied fluxes; 0: not tracked
// in this stage) and the boundary condition (update_U calls update_boundary,
// domain_mod.f90:1564-1582).
template <typename T>
int fv_stage(swh_domain *h, int dst, int src, int qsrc, double fixed_dt, double dt_mult, int mode, double nest_dt,
             int avg_src = -1)
{
    FVArgs<T> a = fv_args<T>(h, dst, src);
    if (qsrc >= 0) {
        a.Sq = (const T *)h->S[qsrc]; a.UHq = (const T *)h->UH[qsrc]; a.VHq = (const T *)h->VH[qsrc];
        a.Cq = (const T *)h->C[qsrc];
    }
    if (avg_src >= 0) {
        // rk2's average with state avg_src in the update (only where nothing but the
        // stage boundary follows the second update: no forcing, no Flather or edge
        // boundary); the stage boundary then averages its own cells from the ring
        if (h->f_on || (h->bc_mask && h->bc_kind != 0))
            return fail(SWH_ESTATE, "internal: folded rk2 average with a forcing or a Flather/edge boundary");
        a.avgS = (const T *)h->S[avg_src]; a.avgUH = (const T *)h->UH[avg_src]; a.avgVH = (const T *)h->VH[avg_src];
        a.avgC = (const T *)h->C[avg_src];
        if (h->bc_mask) a.ring = (T *)h->ring;
    }
    if (h->n_nest_strips && nest_dt != 0.0) {
        // the nesting flux integrals ride on the update launch (after the fluxes)
        a.nest_strips = h->d_nest_strips;
        a.n_nest_strips = h->n_nest_strips;
        a.nest_chunks = (h->nest_strip_max + 255) / 256;
        a.nest_dt = nest_dt;
        a.nest_int = h->d_nest_int;
    }
    SWH_HIP(launch_fv_euler<T>(a, fixed_dt, dt_mult, mode, h->stream));
    T *Cd = (T *)h->C[dst];
    if (h->bc_mask && h->bc_kind == 1)
        SWH_HIP(launch_flather<T>((T *)h->S[dst], Cd, (T *)h->UH[dst], (T *)h->VH[dst], (const T *)h->E, h->nx,
                                  h->ny, h->pitch, h->bc_mask, h->msl, h->bc_use_fun, h->bc_amp, h->bc_period,
                                  h->bc_msl, 0.0, h->sc, h->stream));
    else if (h->bc_mask && h->bc_kind >= 2)
        SWH_HIP(launch_edge_bc<T>((T *)h->S[dst], Cd, (T *)h->UH[dst], (T *)h->VH[dst], (T *)h->E, h->nx, h->ny,
                                  h->pitch, h->bc_mask, h->bc_kind, h->bc_use_fun, h->bc_amp, h->bc_period,
                                  h->bc_msl, h->msl, 0.0, 0, h->sc, h->stream));
    else if (h->bc_mask && avg_src >= 0)
        SWH_HIP(launch_fv_stage_bc_avg<T>((T *)h->S[dst], (T *)h->UH[dst], (T *)h->VH[dst], (const T *)h->S[avg_src],
                                          (const T *)h->UH[avg_src], (const T *)h->VH[avg_src], (const T *)h->ring,
                                          h->nx, h->ny, h->pitch, h->bc_mask, h->bc_amp, h->bc_period, h->bc_msl,
                                          h->sc, h->stream, Cd, (const T *)h->C[avg_src]));
    else if (h->bc_mask)
        SWH_HIP(launch_fv_stage_bc<T>((T *)h->S[dst], (T *)h->UH[dst], (T *)h->VH[dst], h->nx, h->ny, h->pitch,
                                      h->bc_mask, h->bc_amp, h->bc_period, h->bc_msl, h->sc, h->stream, Cd));
    if (h->f_on)
        // update_U then apply_forcing (domain_mod.f90:1579-1580), with the stage's step
        SWH_HIP(launch_fv_forcing<T>((T *)h->S[dst], (T *)h->UH[dst], (T *)h->VH[dst], (const T *)h->fS,
                                     (const T *)h->fUH, (const T *)h->fVH, h->pitch, h->f_i0, h->f_i1, h->f_j0,
                                     h->f_j1, h->f_start, h->f_end, h->sc, h->stream, Cd));
    return SWH_OK;
}

// The finite-volume schemes (domain_mod_timestepping_alternatives_include.f90): the
// state at the start of the step stays in buffer A (backup_U), the result ends in B.
template <typename T>
int dispersion_solve(swh_domain *h, int buf, int rhs_up, int estimate, double ftime, bool ftime_on_device)
{
    const size_t P = h->plane_elems();
    DispArgs<T> a;
    a.nx = h->nx; a.ny = h->ny; a.pitch = h->pitch;
    a.UH = (T *)h->UH[buf]; a.VH = (T *)h->VH[buf]; a.E = (const T *)h->E;
    a.LU = (const T *)h->dlast; a.LV = a.LU + P; a.LE = a.LU + 2 * P;
    a.RU = (T *)h->drhs; a.RV = a.RU + P;
    a.Y0 = (const T *)h->dq[2]; a.Y1 = (const T *)h->dq[1]; a.Y2 = (const T *)h->dq[0];
    a.qt = h->d_qt;
    a.W = (T *)h->dwork;
    a.dbe = h->d_dbe; a.dle = h->d_dle;
    a.msl = (T)h->msl; a.dlon = (T)h->disp_dlon; a.dlat = (T)h->disp_dlat;
    a.td1 = (T)h->disp_td1; a.td2 = (T)h->disp_td2;
    a.cell = h->leapfrog() ? 0 : 1;
    a.sc = ftime_on_device ? h->sc : nullptr;
    a.t_fwd = ftime;
    // no extrapolation before three levels: ten times the iterations instead (:1267-1291)
    int niter = h->disp_iter, extrap = 0;
    if (estimate) {
        if (h->dq_n < 3) niter = 10 * h->disp_iter;
        else extrap = 1;
    }
    SWH_HIP(launch_dispersion<T>(a, rhs_up ? 0 : 1, extrap, niter, h->stream));
    return SWH_OK;
}

// A timestep dt <= 0 means "no timestep argument": the CFL step of the first flux
// computation, decided on the device.
template <typename T>
int fv_step(swh_domain *h, double dt)
{
    if (h->nsteps == 0) {
        int rc = maxq_now<T>(h, h->sc, 0.0);
        if (rc) return rc;
    }
    if (h->n_nest_strips && !(dt > 0.0))
        return fail(SWH_EINVAL, "nested finite-volume domains need the timestep (the multidomain step provides it)");
    const int A = h->cur, B = 1 - A;
    const int acc = kStageAccBf;
    int rc = SWH_OK;
#define STAGE(...) do { rc = fv_stage<T>(h, __VA_ARGS__); if (rc) return rc; } while (0)
    const int first = kStageFirst | kStageBegin;
    switch (h->method) {
    case SWH_METHOD_EULER:
        // one_euler_step (include.f90:8-62)
        STAGE(B, A, -1, dt, 1.0, first | acc | kStageFinish, dt);
        break;
    case SWH_METHOD_RK2:
        // one_rk2_step (include.f90:64-146): two Euler stages with the same step,
        // nesting fluxes tracked half a step each, then the average with backup_U
        STAGE(B, A, -1, dt, 1.0, first | acc, dt * 0.5);
        if (!h->f_on && (!h->bc_mask || h->bc_kind == 0)) {
            // the average folded into the second update (nothing but a stage boundary
            // acts between them, and that averages its own cells)
            STAGE(B, B, -1, dt, 1.0, acc, dt * 0.5, A);
        } else {
            STAGE(B, B, -1, dt, 1.0, acc, dt * 0.5);
            SWH_HIP(launch_fv_average<T>((T *)h->S[B], (T *)h->UH[B], (T *)h->VH[B], (const T *)h->S[A],
                                         (const T *)h->UH[A], (const T *)h->VH[A], h->nx, h->ny, h->pitch, h->sc,
                                         h->stream, (T *)h->C[B], (const T *)h->C[A]));
        }
        break;
    case SWH_METHOD_MIDPOINT:
        // one_midpoint_step (include.f90:257-336): half a step for the fluxes, then the
        // full step from backup_U with them; a dispersive domain solves the dispersive
        // terms on the half-step state first (:286-297), forward time backup_time + dt/2
        // (the device time after the half step)
        STAGE(B, A, -1, dt, 0.5, first, 0.0);
        if (h->disp_on) {
            rc = dispersion_solve<T>(h, B, 0, 1, 0.0, true);
            if (rc) return rc;
        }
        STAGE(B, B, A, dt, 1.0, acc | kStageFromTime0 | kStageFinish, dt);
        break;
    case SWH_METHOD_RK2N: {
        // one_rk2n_step (include.f90:149-254), n = kRk2nN
        const double n_inverse = 1.0 / (1.0 * kRk2nN);
        const double dt1 = dt > 0.0 ? dt / (kRk2nN - 1.0) : -1.0;
        const double nest1 = dt > 0.0 ? dt1 * (kRk2nN - 1.0) * n_inverse : 0.0;
        const double reduce = (1.0 * kRk2nN - 1.0) * n_inverse;
        STAGE(B, A, -1, dt1, 1.0, first | acc, nest1);
        for (int j = 2; j <= kRk2nN - 1; j++) STAGE(B, B, -1, dt1, 1.0, acc, nest1);
        SWH_HIP(launch_fv_rk2n_combine<T>((T *)h->S[B], (T *)h->UH[B], (T *)h->VH[B], (T *)h->S[A], (T *)h->UH[A],
                                          (T *)h->VH[A], h->nx, h->ny, h->pitch, 0, h->sc, h->stream, (T *)h->C[B],
                                          (T *)h->C[A]));
        STAGE(B, B, -1, dt1, reduce, acc | kStageScaleBfN, dt > 0.0 ? reduce * dt1 : 0.0);
        SWH_HIP(launch_fv_rk2n_combine<T>((T *)h->S[B], (T *)h->UH[B], (T *)h->VH[B], (T *)h->S[A], (T *)h->UH[A],
                                          (T *)h->VH[A], h->nx, h->ny, h->pitch, 1, h->sc, h->stream, (T *)h->C[B],
                                          (T *)h->C[A]));
        break;
    }
    default:
        return fail(SWH_EINVAL, "not a finite-volume timestepping_method");
    }
#undef STAGE
    h->cur = B;
    const bool upd = h->maxq_mask && ((h->nsteps % h->maxq_freq) == 0);
    if (upd) {
        rc = maxq_now<T>(h, h->sc, 0.0);
        if (rc) return rc;
    }
    h->nsteps++;
    return SWH_OK;
}

int evolve(swh_domain *h, double dt)
{
    if (h->method == SWH_METHOD_LEAPFROG_NONLINEAR) {
        if (!(dt > 0.0)) return fail(SWH_EINVAL, "timestep must be provided for leapfrog-type evolve_one_step");
        return h->prec == SWH_PREC_FP64 ? nonlinear_step<double, false>(h, dt) : nonlinear_step<float, true>(h, dt);
    }
    if (h->leapfrog()) {
        if (!(dt > 0.0)) return fail(SWH_EINVAL, "timestep must be provided for leapfrog-type evolve_one_step");
        return h->prec == SWH_PREC_FP64 ? leapfrog_step<double, false>(h, dt) : leapfrog_step<float, true>(h, dt);
    }
    return h->prec == SWH_PREC_FP64 ? fv_step<double>(h, dt) : fv_step<float>(h, dt);
}

}  // namespace

namespace swh {

int domain_evolve(swh_domain *h, double dt) { return evolve(h, dt); }

std::string domain_graph_unsupported(const swh_domain *h)
{
    if (h->disp_on) return "dispersion (host-side extrapolation bookkeeping)";
    if (h->maxq_mask && h->maxq_freq != 1) return "max_U updated every n > 1 steps";
    if (h->maxq_mask && h->nsteps == 0) return "max_U before the first step (take one eager step first)";
    if (h->leapfrog()) {
        if (h->maxq_mask & (SWH_MAXQ_ARRIVAL_TIME | SWH_MAXQ_TIME_OF_MAX_STAGE))
            return "leapfrog arrival time / time of max stage (host time)";
        if (h->f_on) return "leapfrog forcing patch (host time)";
        if (h->bc_mask && h->bc_use_fun) return "leapfrog boundary with a stage function (host time)";
    }
    return "";
}

int domain_prepare_dt(swh_domain *h, double dt)
{
    if (!h->leapfrog() || h->rc_dt == dt) return SWH_OK;
    return h->prec == SWH_PREC_FP64 ? build_rc<double>(h, dt) : build_rc<float>(h, dt);
}

// The host-side effects of evolve(h, dt) (leapfrog_step / fv_step), without launches.
void domain_host_advance(swh_domain *h, double dt)
{
    if (h->leapfrog()) {
        const double time0 = h->time;
        const double t_half = h->time + dt * 0.5;
        const double t_after = t_half + 0.5 * dt;
        h->time = t_after;
        h->dt_last_update = dt;
        h->evolve_step_dt = h->time - time0;
    }
    h->cur = 1 - h->cur;
    h->nsteps++;
}

int domain_refresh_elevation(swh_domain *h)
{
    std::vector<double> e((size_t)h->nx * h->ny);
    SWH_HIP(hipStreamSynchronize(h->stream));
    int rc = h->prec == SWH_PREC_FP64 ? download_plane<double>(h, h->E, nullptr, e.data())
                                      : download_plane<float>(h, h->E, nullptr, e.data());
    if (rc) return rc;
    // keep the FP64 host value wherever the device still holds its rounding
    for (size_t k = 0; k < e.size(); k++) {
        const bool same = h->prec == SWH_PREC_FP64 ? e[k] == h->h_elev[k] : (float)h->h_elev[k] == (float)e[k];
        if (!same) h->h_elev[k] = e[k];
    }
    if (h->method == SWH_METHOD_LLPNF && h->truely_linear && !h->h_nsq.empty()) return build_friction_work(h);
    return SWH_OK;
}

int domain_stats_buffer(swh_domain *h)
{
    if (h->d_stats) return SWH_OK;
    return dmalloc((void **)&h->d_stats, sizeof(double) * (6 * (size_t)h->ny + 8));
}

}  // namespace swh

namespace {

void stats_from(const double *v, swh_domain_stats *o)
{
    o->max_stage = v[0];
    o->max_speed = v[1];
    o->min_stage = v[2];
    o->min_speed = v[3];
    o->energy_potential_on_rho = v[4];
    o->energy_kinetic_on_rho = v[5];
    o->energy_total_on_rho = v[6];
}

int sync_scalars(swh_domain *h, DevScalars *out)
{
    SWH_HIP(hipStreamSynchronize(h->stream));
    SWH_HIP(hipMemcpy(out, h->sc, sizeof(DevScalars), hipMemcpyDeviceToHost));
    return SWH_OK;
}

}  // namespace

extern "C" {

int swh_last_error(char *buf, int len)
{
    if (!buf || len <= 0) return SWH_EINVAL;
    snprintf(buf, (size_t)len, "%s", g_err.c_str());
    return SWH_OK;
}

int swh_domain_create(const swh_domain_desc *d, swh_domain **out)
{
    if (!d || !out) return fail(SWH_EINVAL, "null argument");
    *out = nullptr;
    if (d->nx < 3 || d->ny < 3) return fail(SWH_EINVAL, "nx and ny must be >= 3");
    if (!((d->timestepping_method >= SWH_METHOD_RK2 && d->timestepping_method <= SWH_METHOD_LLPNF) ||
          d->timestepping_method == SWH_METHOD_LEAPFROG_NONLINEAR))
        return fail(SWH_EINVAL, "unsupported timestepping_method id");
    if (d->precision != SWH_PREC_FP64 && d->precision != SWH_PREC_FP32K) return fail(SWH_EINVAL, "bad precision");
    if (!d->distance_bottom_edge || !d->distance_left_edge || !d->area_cell_y)
        return fail(SWH_EINVAL, "metric arrays are required");
    if (d->include_coriolis && (!d->coriolis || !d->coriolis_bottom_edge)) return fail(SWH_EINVAL, "coriolis arrays missing");
    if (d->spherical && !d->tanlat_on_radius_earth) return fail(SWH_EINVAL, "tanlat_on_radius_earth missing");
    const int n_ext = d->exterior_cells_width > 0 ? d->exterior_cells_width : 2;
    if (2 * n_ext >= d->nx || 2 * n_ext >= d->ny) return fail(SWH_EINVAL, "domain smaller than its exterior cells");

    swh_domain *h = new swh_domain();
    h->nx = d->nx; h->ny = d->ny;
    h->pitch = (d->nx + 63) / 64 * 64;
    h->fpitch = (d->nx + 1 + 63) / 64 * 64;
    h->method = d->timestepping_method; h->prec = d->precision;
    h->truely_linear = d->linear_solver_is_truely_linear != 0;
    h->coriolis = d->include_coriolis != 0; h->spherical = d->spherical != 0;
    h->friction_type = d->friction_type; h->n_ext = n_ext;
    for (int k = 0; k < 4; k++) h->bex[k] = d->boundary_exterior[k] != 0;
    h->maxq_mask = d->max_U_mask & 63;
    h->maxq_freq = d->max_U_update_frequency > 0 ? d->max_U_update_frequency : 1;
    h->msl = d->msl_linear;
    const bool fv = !h->leapfrog();
    const bool euler = h->method == SWH_METHOD_EULER;
    // timestepping_metadata_mod.f90:73-121
    h->cfl = d->cfl > 0 ? d->cfl : (fv ? (euler ? 0.9 : 0.99) : 0.7);
    h->theta = d->theta > 0 ? d->theta : (fv ? (euler ? 0.9 : 1.6) : 0.0);
    h->lfc = d->linear_friction_coeff;
    h->max_timestep = d->maximum_timestep > 0 ? d->maximum_timestep : kMaxTimestep;
    h->arrival_thresh = d->arrival_stage_threshold_above_msl_linear;
    h->esz = h->prec == SWH_PREC_FP64 ? 8 : 4;
    // default launch shape of the one-step kernel, from measured sweeps over shapes on
    // 400^2 to 3000x1800 grids: 128-thread workgroups; FP64 with 16-byte (two-column)
    // lanes, FP32 with one column per lane and the next row's loads in flight; row strips
    // shortened from 32 rows until the launch has at least 8192 wavefronts (32 per CU),
    // down to one row (the kernel is latency-bound on small grids: 612^2 llpnf FP32K
    // 11 us per step at one row against 50 us at four rows with 16-byte lanes)
    if (h->prec == SWH_PREC_FP64) { h->block_x = 128; h->rows_per_block = 32; h->variant = 1; }
    else { h->block_x = 128; h->rows_per_block = 32; h->variant = 2; }
    if (fv) {
        // finite volume: rows per wavefront of the flux pass, 8, or 4 where a grid would
        // have fewer than 4096 wavefronts of 16-byte lanes at 8 (measured on the beach)
        h->rows_per_block = 8;
        const int vec = h->prec == SWH_PREC_FP64 ? 2 : 4;
        const long long waves_x = ((d->nx + vec - 1) / vec + 63) / 64;
        while (h->rows_per_block > 4 && waves_x * ((d->ny + h->rows_per_block - 1) / h->rows_per_block) < 4096)
            h->rows_per_block /= 2;
    } else {
        const int vec = (h->variant & 1) ? (h->prec == SWH_PREC_FP64 ? 2 : 4) : 1;
        const long long waves_x = ((d->nx + vec - 1) / vec + 63) / 64;
        while (h->rows_per_block > 1 && waves_x * ((d->ny + h->rows_per_block - 1) / h->rows_per_block) < 8192)
            h->rows_per_block /= 2;
    }
    {
        // the two-step kernel: rows per strip chosen by its launcher from the kernel's
        // occupancy (rows2 = 0); workgroups of 128 threads (FP32) or 256 (FP64)
        h->rows2 = 0;
        h->block2 = h->prec == SWH_PREC_FP64 ? 256 : 128;
    }
    const int nx = h->nx, ny = h->ny;
    h->dbe.assign(d->distance_bottom_edge, d->distance_bottom_edge + ny + 1);
    h->dle.assign(d->distance_left_edge, d->distance_left_edge + nx + 1);
    h->area.assign(d->area_cell_y, d->area_cell_y + ny);
    h->cor.assign(ny, 0.0);
    h->corbe.assign(ny + 1, 0.0);
    h->tanlat.assign(ny, 0.0);
    if (h->coriolis) {
        h->cor.assign(d->coriolis, d->coriolis + ny);
        h->corbe.assign(d->coriolis_bottom_edge, d->coriolis_bottom_edge + ny + 1);
    }
    if (h->spherical) h->tanlat.assign(d->tanlat_on_radius_earth, d->tanlat_on_radius_earth + ny);
    h->h_elev.assign((size_t)nx * ny, 0.0);
    h->h_nsq.assign((size_t)nx * ny, 0.0);

    int rc = SWH_OK;
    auto A = [&](void **p, size_t bytes) {
        if (rc == SWH_OK) rc = dmalloc(p, bytes);
    };
    if (hipStreamCreate(&h->stream) != hipSuccess) {
        delete h;
        return fail(SWH_EHIP, "hipStreamCreate failed");
    }
    const size_t P = h->plane_elems() * h->esz;
    A((void **)&h->d_dbe, sizeof(double) * (ny + 1));
    A((void **)&h->d_dle, sizeof(double) * (nx + 1));
    A((void **)&h->d_area, sizeof(double) * ny);
    A((void **)&h->d_cor, sizeof(double) * ny);
    A((void **)&h->d_tanlat, sizeof(double) * ny);
    for (int b = 0; b < 2; b++) {
        A(&h->S[b], P); A(&h->UH[b], P); A(&h->VH[b], P);
        if (h->prec == SWH_PREC_FP32K) A(&h->C[b], P);
    }
    A(&h->E, P);
    A(&h->nsq, P);
    if (h->method == SWH_METHOD_LLPNF) { A(&h->fwU, P); A(&h->fwV, P); }
    if (h->method == SWH_METHOD_LEAPFROG_NONLINEAR) {
        A(&h->nlw, P * 9);
        A((void **)&h->d_corbe, sizeof(double) * (ny + 1));
    }
    if (h->maxq_mask) A(&h->mx, P * N_MAXQ);
    if (fv) {
        const size_t FPn = (size_t)h->fpitch * (ny + 1);
        A((void **)&h->d_bmax, fv_flux_blocks(nx, ny) * sizeof(double));
        A(&h->fNS, FPn * 3 * h->esz);
        A(&h->fEW, FPn * 3 * h->esz);
        A((void **)&h->esU, h->plane_elems() * sizeof(double));
        A((void **)&h->esV, h->plane_elems() * sizeof(double));
        A((void **)&h->esJ, FPn * sizeof(double));
        A((void **)&h->esE, FPn * sizeof(double));
        A(&h->ring, (size_t)2 * (nx + ny) * h->esz);
    }
    // (the nonlinear leapfrog's update adds one row array; the other methods' buffer
    // keeps its size, and with it the placement of what is allocated after it)
    A(&h->rc, h->esz * ((size_t)(h->method == SWH_METHOD_LEAPFROG_NONLINEAR ? 6 : 5) * ny + (nx + 1) + (ny + 1)));
    A((void **)&h->sc, sizeof(DevScalars));
    // line workgroups: at most 4 ceil(max(nx, ny) / 64) wanted, in rows at most
    // ceil(nx / 64) workgroups wide
    h->bf_cap = 4 * ((std::max(nx, ny) + 63) / 64) + (nx + 63) / 64;
    A((void **)&h->d_bfpart, sizeof(double) * 4 * (size_t)h->bf_cap);
    A((void **)&h->d_work, sizeof(double) * (2 * (size_t)ny + 16));
    if (rc) {
        swh_domain_destroy(h);
        return rc;
    }
    if (hipMemcpy(h->d_dbe, h->dbe.data(), sizeof(double) * (ny + 1), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(h->d_dle, h->dle.data(), sizeof(double) * (nx + 1), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(h->d_area, h->area.data(), sizeof(double) * ny, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(h->d_cor, h->cor.data(), sizeof(double) * ny, hipMemcpyHostToDevice) != hipSuccess ||
        (h->d_corbe &&
         hipMemcpy(h->d_corbe, h->corbe.data(), sizeof(double) * (ny + 1), hipMemcpyHostToDevice) != hipSuccess) ||
        hipMemcpy(h->d_tanlat, h->tanlat.data(), sizeof(double) * ny, hipMemcpyHostToDevice) != hipSuccess) {
        swh_domain_destroy(h);
        return fail(SWH_EHIP, "metric upload failed");
    }
    if (h->maxq_mask) {
        // max_U initial values (domain_mod.f90:1119-1136): -0.99*huge(float32), and
        // +0.99*huge(float32) for min_stage
        const double big = (double)(0.99f * 3.40282347e+38f);
        std::vector<double> init((size_t)nx * ny, -big);
        for (int k = 0; k < N_MAXQ; k++) {
            if (k == MX_MIN_STAGE) std::fill(init.begin(), init.end(), big);
            void *plane = (char *)h->mx + (size_t)k * P;
            rc = upload_plane_any(h, plane, init.data());
            if (rc) {
                swh_domain_destroy(h);
                return rc;
            }
        }
    }
    *out = h;
    return SWH_OK;
}

int swh_domain_destroy(swh_domain *h)
{
    if (!h) return SWH_OK;
    if (h->stream) (void)hipStreamSynchronize(h->stream);
    void *ptrs[] = {h->d_dbe, h->d_dle, h->d_area, h->d_cor, h->d_tanlat, h->S[0], h->S[1], h->UH[0], h->UH[1],
                    h->VH[0], h->VH[1], h->C[0], h->C[1], h->E, h->fwU, h->fwV, h->nsq, h->mx, h->depth, h->u,
                    h->v, h->fNS, h->fEW, h->esU, h->esV, h->esJ, h->esE, h->rc, h->sc, h->d_work, h->d_gidx,
                    h->d_gval, h->d_gtime, h->adj[0][0], h->adj[0][1], h->adj[0][2], h->adj[1][0],
                    h->adj[1][1], h->adj[1][2], h->d_gvec, h->d_bmax, h->d_bfpart, h->fS, h->fUH, h->fVH, h->fE,
                    h->nlw, h->d_corbe, h->dlast, h->drhs, h->dq[0], h->dq[1], h->dq[2], h->dwork, h->d_qt,
                    h->d_stats, h->d_ggrad, h->ring, h->d_stage};
    for (void *p : ptrs) dfree(p);
    for (void *p : h->ckpt) dfree(p);
    dfree(h->snap);
    if (h->snap_ev) (void)hipEventDestroy(h->snap_ev);
    if (h->cstream) (void)hipStreamDestroy(h->cstream);
    if (h->stream) (void)hipStreamDestroy(h->stream);
    delete h;
    return SWH_OK;
}

int swh_upload_U(swh_domain *h, const double *U)
{
    if (!h || !U) return fail(SWH_EINVAL, "null argument");
    SWH_HIP(hipStreamSynchronize(h->stream));
    const size_t n = (size_t)h->nx * h->ny;
    const int c = h->cur;
    int rc;
    // FP32 state: the stage with its Kahan term, so that S - C represents the FP64 stage (on
    // the FP32 elevation where depth is stage - elevation)
    if (h->prec != SWH_PREC_FP64 && !h->tl_solver()) {
        const std::vector<double> t = stage_on_fp32_elevation(h, U, U + 3 * n);
        rc = upload_plane_any(h, h->S[c], t.data(), h->C[c]);
    } else {
        rc = upload_plane_any(h, h->S[c], U, h->C[c]);
    }
    if (rc) return rc;
    if ((rc = upload_plane_any(h, h->UH[c], U + n))) return rc;
    if ((rc = upload_plane_any(h, h->VH[c], U + 2 * n))) return rc;
    if ((rc = upload_plane_any(h, h->E, U + 3 * n))) return rc;
    h->h_elev.assign(U + 3 * n, U + 4 * n);
    return build_friction_work(h);
}

int swh_upload_state(swh_domain *h, const double *stage, const double *uh, const double *vh)
{
    if (!h) return fail(SWH_EINVAL, "null argument");
    SWH_HIP(hipStreamSynchronize(h->stream));
    const int c = h->cur;
    int rc;
    if (h->prec != SWH_PREC_FP64 && !h->tl_solver()) {
        if (h->h_elev.size() != (size_t)h->nx * h->ny) return fail(SWH_ESTATE, "upload U (with the elevation) first");
        const std::vector<double> t = stage_on_fp32_elevation(h, stage, h->h_elev.data());
        rc = upload_plane_any(h, h->S[c], t.data(), h->C[c]);
    } else {
        rc = upload_plane_any(h, h->S[c], stage, h->C[c]);
    }
    if (rc) return rc;
    if ((rc = upload_plane_any(h, h->UH[c], uh))) return rc;
    return upload_plane_any(h, h->VH[c], vh);
}

int swh_download_U(swh_domain *h, double *U)
{
    if (!h || !U) return fail(SWH_EINVAL, "null argument");
    SWH_HIP(hipStreamSynchronize(h->stream));
    const size_t n = (size_t)h->nx * h->ny;
    const int c = h->cur;
    int rc;
    if (h->prec == SWH_PREC_FP64) {
        if ((rc = download_plane<double>(h, h->S[c], nullptr, U))) return rc;
        if ((rc = download_plane<double>(h, h->UH[c], nullptr, U + n))) return rc;
        if ((rc = download_plane<double>(h, h->VH[c], nullptr, U + 2 * n))) return rc;
        if ((rc = download_plane<double>(h, h->E, nullptr, U + 3 * n))) return rc;
    } else {
        if ((rc = download_plane<float>(h, h->S[c], h->C[c], U))) return rc;
        if ((rc = download_plane<float>(h, h->UH[c], nullptr, U + n))) return rc;
        if ((rc = download_plane<float>(h, h->VH[c], nullptr, U + 2 * n))) return rc;
        if ((rc = download_plane<float>(h, h->E, nullptr, U + 3 * n))) return rc;
    }
    return SWH_OK;
}

int swh_upload_manning_squared(swh_domain *h, const double *m)
{
    if (!h || !m) return fail(SWH_EINVAL, "null argument");
    SWH_HIP(hipStreamSynchronize(h->stream));
    int rc = upload_plane_any(h, h->nsq, m);
    if (rc) return rc;
    h->h_nsq.assign(m, m + (size_t)h->nx * h->ny);
    return build_friction_work(h);
}

int swh_set_time(swh_domain *h, double t)
{
    if (!h) return fail(SWH_EINVAL, "null argument");
    DevScalars s;
    int rc = sync_scalars(h, &s);
    if (rc) return rc;
    s.time = t;
    SWH_HIP(hipMemcpy(h->sc, &s, sizeof(DevScalars), hipMemcpyHostToDevice));
    h->time = t;
    return SWH_OK;
}

int swh_set_stage_boundary_sine(swh_domain *h, double amp, double period, double msl, const int32_t sides[4])
{
    if (!h || !sides) return fail(SWH_EINVAL, "null argument");
    if (!(period != 0.0)) return fail(SWH_EINVAL, "period must be non-zero");
    h->bc_mask = (sides[0] ? 1 : 0) | (sides[1] ? 2 : 0) | (sides[2] ? 4 : 0) | (sides[3] ? 8 : 0);
    for (int k = 0; k < 4; k++) h->bex[k] = sides[k] != 0;
    h->bc_amp = amp; h->bc_period = period; h->bc_msl = msl;
    h->bc_kind = 0; h->bc_use_fun = 1;
    return SWH_OK;
}

int swh_set_boundary(swh_domain *h, int32_t kind, const int32_t sides[4], int32_t use_stage_function, double amp,
                     double period, double msl)
{
    if (!h || !sides) return fail(SWH_EINVAL, "null argument");
    if (kind < SWH_BC_STAGE_TRANSMISSIVE_NORMAL_MOMENTUM || kind > SWH_BC_STAGE_TRANSMISSIVE_MOMENTUM)
        return fail(SWH_EINVAL, "unknown boundary kind");
    if (!use_stage_function &&
        (kind == SWH_BC_STAGE_TRANSMISSIVE_NORMAL_MOMENTUM || kind == SWH_BC_STAGE_TRANSMISSIVE_MOMENTUM))
        return fail(SWH_EINVAL, "this boundary needs a boundary_function (the stage function)");
    if (use_stage_function && !(period != 0.0)) return fail(SWH_EINVAL, "period must be non-zero");
    if (h->nx < 3 || h->ny < 3) return fail(SWH_EINVAL, "boundaries need at least 3 x 3 cells");
    h->bc_mask = (sides[0] ? 1 : 0) | (sides[1] ? 2 : 0) | (sides[2] ? 4 : 0) | (sides[3] ? 8 : 0);
    for (int k = 0; k < 4; k++) h->bex[k] = sides[k] != 0;
    h->bc_amp = amp; h->bc_period = period; h->bc_msl = msl;
    h->bc_kind = kind; h->bc_use_fun = use_stage_function != 0;
    return SWH_OK;
}

int swh_set_flather_boundary(swh_domain *h, const int32_t sides[4], int32_t use_stage_function, double amp,
                             double period, double msl)
{
    if (!h || !sides) return fail(SWH_EINVAL, "null argument");
    if (use_stage_function && !(period != 0.0)) return fail(SWH_EINVAL, "period must be non-zero");
    if (h->nx < 3 || h->ny < 3) return fail(SWH_EINVAL, "the Flather boundary needs at least 3 x 3 cells");
    h->bc_mask = (sides[0] ? 1 : 0) | (sides[1] ? 2 : 0) | (sides[2] ? 4 : 0) | (sides[3] ? 8 : 0);
    for (int k = 0; k < 4; k++) h->bex[k] = sides[k] != 0;
    h->bc_amp = amp; h->bc_period = period; h->bc_msl = msl;
    h->bc_kind = 1; h->bc_use_fun = use_stage_function != 0;
    return SWH_OK;
}

int swh_evolve_one_step(swh_domain *h, double timestep)
{
    if (!h) return fail(SWH_EINVAL, "null argument");
    return evolve(h, timestep);
}

int swh_evolve_steps(swh_domain *h, double timestep, int32_t nsteps)
{
    if (!h) return fail(SWH_EINVAL, "null argument");
    if (nsteps < 0) return fail(SWH_EINVAL, "nsteps must be >= 0");
    for (int k = 0; k < nsteps;) {
        int rc;
        if (k + 1 < nsteps && timestep > 0.0 && two_step_ok(h, timestep)) {
            rc = h->prec == SWH_PREC_FP64 ? leapfrog_step2<double, false>(h, timestep)
                                          : leapfrog_step2<float, true>(h, timestep);
            k += 2;
        } else {
            rc = evolve(h, timestep);
            k += 1;
        }
        if (rc) return rc;
    }
    return SWH_OK;
}

int swh_set_forcing_patch(swh_domain *h, double start_time, double end_time, int32_t i0, int32_t i1, int32_t j0,
                          int32_t j1, int32_t k0, int32_t k1, const double *work)
{
    if (!h || !work) return fail(SWH_EINVAL, "null argument");
    // apply_forcing_patch_base's checks (forcing_mod.f90:177-205)
    if (i0 > i1 || j0 > j1 || i0 < 1 || j0 < 1 || i1 > h->nx || j1 > h->ny)
        return fail(SWH_EINVAL, "forcing patch indices outside the domain or not ordered");
    if (start_time > end_time) return fail(SWH_EINVAL, "forcing patch start_time > end_time");
    if (!(k0 >= 1 && k1 <= 4 && k0 <= k1)) return fail(SWH_EINVAL, "forcing patch variables must be k0..k1 in 1..4");
    const int ni = i1 - i0 + 1, nj = j1 - j0 + 1, nk = k1 - k0 + 1;
    const size_t np = (size_t)ni * nj;
    auto nonzero = [&](int k) {
        if (k < k0 || k > k1) return false;
        const double *w = work + (size_t)(k - k0) * np;
        for (size_t q = 0; q < np; q++)
            if (w[q] != 0.0) return true;
        return false;
    };
    if (h->leapfrog() && (nonzero(2) || nonzero(3)))
        return fail(SWH_EINVAL, "forcing of UH/VH is not provided for the leapfrog methods");
    if (!h->leapfrog() && nonzero(4))
        return fail(SWH_EINVAL,
                    "elevation forcing needs support_elevation_forcing, which the finite-volume methods here do not have");
    SWH_HIP(hipStreamSynchronize(h->stream));
    void **planes[4] = {&h->fS, &h->fUH, &h->fVH, &h->fE};
    std::vector<double> full((size_t)h->nx * h->ny);
    for (int k = 1; k <= 4; k++) {
        dfree(*planes[k - 1]);
        *planes[k - 1] = nullptr;
        if (!nonzero(k)) continue;
        std::fill(full.begin(), full.end(), 0.0);
        const double *w = work + (size_t)(k - k0) * np;
        for (int j = 0; j < nj; j++)
            for (int i = 0; i < ni; i++) full[(size_t)(j0 - 1 + j) * h->nx + (i0 - 1 + i)] = w[(size_t)j * ni + i];
        int rc = dmalloc(planes[k - 1], h->plane_elems() * h->esz);
        if (!rc) rc = upload_plane_any(h, *planes[k - 1], full.data());
        if (rc) return rc;
    }
    (void)nk;
    h->f_on = 1;
    h->f_start = start_time; h->f_end = end_time;
    h->f_i0 = i0 - 1; h->f_i1 = i1 - 1; h->f_j0 = j0 - 1; h->f_j1 = j1 - 1;
    return SWH_OK;
}

int swh_clear_forcing_patch(swh_domain *h)
{
    if (!h) return fail(SWH_EINVAL, "null argument");
    SWH_HIP(hipStreamSynchronize(h->stream));
    dfree(h->fS); dfree(h->fUH); dfree(h->fVH); dfree(h->fE);
    h->fS = h->fUH = h->fVH = h->fE = nullptr;
    h->f_on = 0;
    return SWH_OK;
}

int swh_get_scalars(swh_domain *h, swh_scalars *out)
{
    if (!h || !out) return fail(SWH_EINVAL, "null argument");
    DevScalars s;
    int rc = sync_scalars(h, &s);
    if (rc) return rc;
    if (h->leapfrog()) {
        out->time = h->time;
        out->max_dt = 0.0;
        out->evolve_step_dt = h->evolve_step_dt;
        out->dt_last_update = h->dt_last_update;
    } else {
        out->time = s.time;
        out->max_dt = s.max_dt;
        out->evolve_step_dt = h->nsteps ? s.evolve_dt : 0.0;
        out->dt_last_update = h->nsteps ? s.stage_dt : 0.0;
        h->time = s.time;
    }
    for (int k = 0; k < 4; k++) out->boundary_flux_store[k] = s.bf_store[k];
    out->boundary_flux_evolve_integral = s.bf_evolve;
    out->boundary_flux_time_integral = s.bf_time;
    out->negative_depth_fix_counter = (double)s.neg_depth_count;
    out->nsteps_advanced = h->nsteps;
    return SWH_OK;
}

int swh_volume_interior(swh_domain *h, double *out)
{
    if (!h || !out) return fail(SWH_EINVAL, "null argument");
    const int c = h->cur;
    double *res = h->d_work + 2 * (size_t)h->ny + 8;
    if (h->prec == SWH_PREC_FP64)
        SWH_HIP(launch_volume<double>((const double *)h->S[c], nullptr, (const double *)h->E, h->nx, h->ny, h->pitch,
                                      h->n_ext, h->d_area, h->d_work, res, h->stream));
    else
        SWH_HIP(launch_volume<float>((const float *)h->S[c], (const float *)h->C[c], (const float *)h->E, h->nx,
                                     h->ny, h->pitch, h->n_ext, h->d_area, h->d_work, res, h->stream));
    SWH_HIP(hipStreamSynchronize(h->stream));
    SWH_HIP(hipMemcpy(out, res, sizeof(double), hipMemcpyDeviceToHost));
    return SWH_OK;
}

int swh_domain_statistics(swh_domain *h, swh_domain_stats *out)
{
    if (!h || !out) return fail(SWH_EINVAL, "null argument");
    int rc = domain_stats_buffer(h);
    if (rc) return rc;
    double *res = h->d_stats + 6 * (size_t)h->ny;
    if ((rc = domain_statistics_launch(h, nullptr, h->d_stats, res, h->stream))) return rc;
    double v[7];
    SWH_HIP(hipStreamSynchronize(h->stream));
    SWH_HIP(hipMemcpy(v, res, sizeof v, hipMemcpyDeviceToHost));
    stats_from(v, out);
    return SWH_OK;
}

int swh_domain_check_stability(swh_domain *h, int64_t *n_bad)
{
    if (!h || !n_bad) return fail(SWH_EINVAL, "null argument");
    int rc = domain_stats_buffer(h);
    if (rc) return rc;
    unsigned long long *cnt = (unsigned long long *)(h->d_stats + 6 * (size_t)h->ny + 7);
    if ((rc = domain_stability_launch(h, cnt, h->stream))) return rc;
    unsigned long long v = 0;
    SWH_HIP(hipStreamSynchronize(h->stream));
    SWH_HIP(hipMemcpy(&v, cnt, sizeof v, hipMemcpyDeviceToHost));
    *n_bad = (int64_t)v;
    return SWH_OK;
}

int swh_stationary_timestep_max(swh_domain *h, double *out)
{
    if (!h || !out) return fail(SWH_EINVAL, "null argument");
    const int c = h->cur;
    double *res = h->d_work + 2 * (size_t)h->ny + 9;
    const int linear = h->tl_solver();
    if (h->prec == SWH_PREC_FP64)
        SWH_HIP(launch_timestep_max<double>((const double *)h->S[c], (const double *)h->E, h->nx, h->ny, h->pitch,
                                            linear, h->msl, h->d_dbe, h->d_dle, res, h->stream));
    else
        SWH_HIP(launch_timestep_max<float>((const float *)h->S[c], (const float *)h->E, h->nx, h->ny, h->pitch,
                                           linear, h->msl, h->d_dbe, h->d_dle, res, h->stream));
    SWH_HIP(hipStreamSynchronize(h->stream));
    double v;
    SWH_HIP(hipMemcpy(&v, res, sizeof(double), hipMemcpyDeviceToHost));
    v = v * h->cfl;
    // stationary_timestep_max, domain_mod.f90:2000-2013
    if (!h->leapfrog()) v = v * 0.5;
    if (h->method == SWH_METHOD_RK2N) v = v * (kRk2nN - 1.0);
    *out = v;
    return SWH_OK;
}

int swh_download_max_U(swh_domain *h, int32_t which, double *out)
{
    if (!h || !out) return fail(SWH_EINVAL, "null argument");
    if (!h->maxq_mask) return fail(SWH_ESTATE, "max_U is not tracked (max_U_mask == 0)");
    int k = -1;
    for (int b = 0; b < N_MAXQ; b++)
        if (which == (1 << b)) k = b;
    if (k < 0 || !(h->maxq_mask & which)) return fail(SWH_EINVAL, "max_U plane not tracked");
    SWH_HIP(hipStreamSynchronize(h->stream));
    const void *plane = (const char *)h->mx + (size_t)k * h->plane_elems() * h->esz;
    return h->prec == SWH_PREC_FP64 ? download_plane<double>(h, plane, nullptr, out)
                                    : download_plane<float>(h, plane, nullptr, out);
}

int swh_gauges_set(swh_domain *h, int32_t n, const int32_t *ij, int32_t max_samples)
{
    if (!h || (n > 0 && !ij) || n < 0 || max_samples < 1) return fail(SWH_EINVAL, "bad gauge arguments");
    std::vector<long long> idx(n);
    for (int g = 0; g < n; g++) {
        const int i = ij[2 * g] - 1, j = ij[2 * g + 1] - 1;
        if (i < 0 || i >= h->nx || j < 0 || j >= h->ny) return fail(SWH_EINVAL, "gauge cell outside the domain");
        idx[g] = (long long)j * h->pitch + i;
    }
    SWH_HIP(hipStreamSynchronize(h->stream));
    dfree(h->d_gidx); dfree(h->d_gval); dfree(h->d_gtime);
    h->d_gidx = nullptr; h->d_gval = nullptr; h->d_gtime = nullptr;
    int rc = dmalloc((void **)&h->d_gidx, sizeof(long long) * (n > 0 ? n : 1));
    if (!rc) rc = dmalloc((void **)&h->d_gval, sizeof(double) * 3 * (size_t)(n > 0 ? n : 1) * max_samples);
    if (!rc) rc = dmalloc((void **)&h->d_gtime, sizeof(double) * max_samples);
    if (rc) return rc;
    if (n > 0) SWH_HIP(hipMemcpy(h->d_gidx, idx.data(), sizeof(long long) * n, hipMemcpyHostToDevice));
    h->ng = n; h->gmax = max_samples; h->gsamples = 0;
    return SWH_OK;
}

int swh_gauges_sample(swh_domain *h)
{
    if (!h) return fail(SWH_EINVAL, "null argument");
    if (!h->d_gidx) return fail(SWH_ESTATE, "gauges not set");
    if (h->gsamples >= h->gmax) return fail(SWH_ESTATE, "gauge sample buffer full");
    const int c = h->cur;
    const DevScalars *tsrc = h->leapfrog() ? nullptr : h->sc;
    if (h->prec == SWH_PREC_FP64)
        SWH_HIP(launch_gauges<double>((const double *)h->S[c], nullptr, (const double *)h->UH[c],
                                      (const double *)h->VH[c], h->ng, h->d_gidx, h->d_gval, h->d_gtime, h->gsamples,
                                      tsrc, h->time, h->stream));
    else
        SWH_HIP(launch_gauges<float>((const float *)h->S[c], (const float *)h->C[c], (const float *)h->UH[c],
                                     (const float *)h->VH[c], h->ng, h->d_gidx, h->d_gval, h->d_gtime, h->gsamples,
                                     tsrc, h->time, h->stream));
    h->gsamples++;
    return SWH_OK;
}

int swh_gauges_download(swh_domain *h, double *values, double *times, int32_t *nsamples)
{
    if (!h || !values || !times || !nsamples) return fail(SWH_EINVAL, "null argument");
    SWH_HIP(hipStreamSynchronize(h->stream));
    if (h->gsamples > 0) {
        SWH_HIP(hipMemcpy(values, h->d_gval, sizeof(double) * 3 * (size_t)h->ng * h->gsamples,
                          hipMemcpyDeviceToHost));
        SWH_HIP(hipMemcpy(times, h->d_gtime, sizeof(double) * h->gsamples, hipMemcpyDeviceToHost));
    }
    *nsamples = h->gsamples;
    return SWH_OK;
}

int swh_dispersion_enable(swh_domain *h, double dlon, double dlat, double td1, double td2, int32_t niter)
{
    if (!h) return fail(SWH_EINVAL, "null argument");
    // the methods the reference allows (domain_mod.f90:1313-1327); 'cliffs' is not built
    if (h->method == SWH_METHOD_RK2N) return fail(SWH_EINVAL, "dispersion is not supported for 'rk2n'");
    if (niter < 1 || !(dlon > 0.0) || !(dlat > 0.0)) return fail(SWH_EINVAL, "bad dispersion parameters");
    if (!(td1 != td2)) return fail(SWH_EINVAL, "td1 must differ from td2");
    if (!h->disp_on) {
        SWH_HIP(hipStreamSynchronize(h->stream));
        const size_t P = h->plane_elems() * h->esz;
        int rc = dmalloc(&h->dlast, 3 * P);
        if (!rc) rc = dmalloc(&h->drhs, 2 * P);
        for (int k = 0; k < 3 && !rc; k++) rc = dmalloc(&h->dq[k], 2 * P);
        if (!rc) rc = dmalloc(&h->dwork, 4 * (size_t)h->nx * h->ny * h->esz);
        if (!rc) rc = dmalloc((void **)&h->d_qt, 3 * sizeof(double));
        if (rc) return rc;
        h->disp_on = 1;
    }
    h->disp_td1 = td1; h->disp_td2 = td2; h->disp_iter = niter;
    h->disp_dlon = dlon; h->disp_dlat = dlat;
    h->dq_n = 0;
    return SWH_OK;
}

int swh_dispersion_store_last_U(swh_domain *h)
{
    if (!h) return fail(SWH_EINVAL, "null argument");
    if (!h->disp_on) return fail(SWH_ESTATE, "dispersion not enabled (swh_dispersion_enable)");
    const size_t P = h->plane_elems() * h->esz;
    char *d = (char *)h->dlast;
    const int c = h->cur;
    SWH_HIP(hipMemcpyAsync(d, h->UH[c], P, hipMemcpyDeviceToDevice, h->stream));
    SWH_HIP(hipMemcpyAsync(d + P, h->VH[c], P, hipMemcpyDeviceToDevice, h->stream));
    SWH_HIP(hipMemcpyAsync(d + 2 * P, h->E, P, hipMemcpyDeviceToDevice, h->stream));
    return SWH_OK;
}

int swh_dispersion_append_extrapolation(swh_domain *h)
{
    if (!h) return fail(SWH_EINVAL, "null argument");
    if (!h->disp_on) return fail(SWH_ESTATE, "dispersion not enabled (swh_dispersion_enable)");
    void *old = h->dq[0];
    h->dq[0] = h->dq[1];
    h->dq[1] = h->dq[2];
    h->dq[2] = old;
    if (h->dq_n < 3) h->dq_n++;
    SWH_HIP(launch_dispersion_qt_append(h->d_qt, h->leapfrog() ? nullptr : h->sc, h->time, h->stream));
    const size_t P = h->plane_elems() * h->esz;
    const int c = h->cur;
    SWH_HIP(hipMemcpyAsync(h->dq[2], h->UH[c], P, hipMemcpyDeviceToDevice, h->stream));
    SWH_HIP(hipMemcpyAsync((char *)h->dq[2] + P, h->VH[c], P, hipMemcpyDeviceToDevice, h->stream));
    return SWH_OK;
}

int swh_dispersion_solve(swh_domain *h, int32_t rhs_is_up_to_date, int32_t estimate_solution_forward_in_time,
                         double forward_time)
{
    if (!h) return fail(SWH_EINVAL, "null argument");
    if (!h->disp_on) return fail(SWH_ESTATE, "dispersion not enabled (swh_dispersion_enable)");
    return h->prec == SWH_PREC_FP64 ? dispersion_solve<double>(h, h->cur, rhs_is_up_to_date,
                                                               estimate_solution_forward_in_time, forward_time, false)
                                    : dispersion_solve<float>(h, h->cur, rhs_is_up_to_date,
                                                              estimate_solution_forward_in_time, forward_time, false);
}

int swh_snapshot(swh_domain *h)
{
    if (!h) return fail(SWH_EINVAL, "null argument");
    const size_t n = (size_t)h->nx * h->ny;
    if (!h->snap) {
        SWH_HIP(hipStreamSynchronize(h->stream));
        int rc = dmalloc((void **)&h->snap, (4 * n + 1) * sizeof(double));
        if (rc) return rc;
        SWH_HIP(hipEventCreateWithFlags(&h->snap_ev, hipEventDisableTiming));
        SWH_HIP(hipStreamCreate(&h->cstream));
    }
    const int c = h->cur;
    const DevScalars *tsrc = h->leapfrog() ? nullptr : h->sc;
    if (h->prec == SWH_PREC_FP64)
        SWH_HIP(launch_snapshot<double>((const double *)h->S[c], nullptr, (const double *)h->UH[c],
                                        (const double *)h->VH[c], (const double *)h->E, h->nx, h->ny, h->pitch,
                                        h->snap, tsrc, h->time, h->stream));
    else
        SWH_HIP(launch_snapshot<float>((const float *)h->S[c], (const float *)h->C[c], (const float *)h->UH[c],
                                       (const float *)h->VH[c], (const float *)h->E, h->nx, h->ny, h->pitch, h->snap,
                                       tsrc, h->time, h->stream));
    SWH_HIP(hipEventRecord(h->snap_ev, h->stream));
    h->snap_ok = 1;
    return SWH_OK;
}

int swh_snapshot_download(swh_domain *h, double *U, double *time)
{
    if (!h || !U) return fail(SWH_EINVAL, "null argument");
    if (!h->snap_ok) return fail(SWH_ESTATE, "no snapshot taken (swh_snapshot)");
    const size_t n = (size_t)h->nx * h->ny;
    double t = 0.0;
    SWH_HIP(hipStreamWaitEvent(h->cstream, h->snap_ev, 0));
    SWH_HIP(hipMemcpyAsync(U, h->snap, 4 * n * sizeof(double), hipMemcpyDeviceToHost, h->cstream));
    SWH_HIP(hipMemcpyAsync(&t, h->snap + 4 * n, sizeof(double), hipMemcpyDeviceToHost, h->cstream));
    SWH_HIP(hipStreamSynchronize(h->cstream));
    if (time) *time = t;
    return SWH_OK;
}

int swh_synchronize(swh_domain *h)
{
    if (!h) return fail(SWH_EINVAL, "null argument");
    SWH_HIP(hipStreamSynchronize(h->stream));
    return SWH_OK;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Adjoint of the truely-linear leapfrog step (see swh_adjoint.hip)
// ---------------------------------------------------------------------------
namespace {

int adj_check(swh_domain *h)
{
    if (h->method != SWH_METHOD_LINEAR || !h->truely_linear)
        return fail(SWH_ESTATE, "the adjoint is provided for the truely-linear 'linear' method");
    if (h->bc_mask) return fail(SWH_ESTATE, "the adjoint does not support a stage boundary");
    if (!h->adj[0][0]) {
        const size_t P = h->plane_elems() * h->esz;
        for (int b = 0; b < 2; b++)
            for (int k = 0; k < 3; k++) {
                int rc = dmalloc(&h->adj[b][k], P);
                if (rc) return rc;
            }
    }
    return SWH_OK;
}

template <typename T>
int adj_steps(swh_domain *h, double dt, int nsteps)
{
    if (h->rc_dt != dt) {
        int rc = build_rc<T>(h, dt);
        if (rc) return rc;
    }
    AdjArgs<T> a;
    a.nx = h->nx; a.ny = h->ny; a.pitch = h->pitch;
    a.E = (const T *)h->E;
    a.rc = rc_view<T>(h);
    a.msl = (T)h->msl;
    a.lin_fric_mult = (T)(1.0 / (1.0 + dt * h->lfc));
    a.coriolis = h->coriolis;
    for (int k = 0; k < nsteps; k++) {
        const int c = h->adj_cur, o = 1 - c;
        a.lS = (const T *)h->adj[c][0]; a.lU = (const T *)h->adj[c][1]; a.lV = (const T *)h->adj[c][2];
        a.oS = (T *)h->adj[o][0]; a.oU = (T *)h->adj[o][1]; a.oV = (T *)h->adj[o][2];
        SWH_HIP(launch_adjoint_step<T>(a, h->stream));
        h->adj_cur = o;
    }
    return SWH_OK;
}

int adj_inject_host(swh_domain *h, const double *values)
{
    if (!h->d_gvec) {
        int rc = dmalloc((void **)&h->d_gvec, sizeof(double) * (h->ng > 0 ? h->ng : 1));
        if (rc) return rc;
    }
    SWH_HIP(hipMemcpyAsync(h->d_gvec, values, sizeof(double) * h->ng, hipMemcpyHostToDevice, h->stream));
    // the host buffer must stay valid until the copy has been consumed
    SWH_HIP(hipStreamSynchronize(h->stream));
    void *lS = h->adj[h->adj_cur][0];
    if (h->prec == SWH_PREC_FP64) SWH_HIP(launch_adjoint_inject<double>((double *)lS, h->ng, h->d_gidx, h->d_gvec, h->stream));
    else SWH_HIP(launch_adjoint_inject<float>((float *)lS, h->ng, h->d_gidx, h->d_gvec, h->stream));
    return SWH_OK;
}

}  // namespace

extern "C" {

int swh_adjoint_reset(swh_domain *h)
{
    if (!h) return fail(SWH_EINVAL, "null argument");
    int rc = adj_check(h);
    if (rc) return rc;
    SWH_HIP(hipStreamSynchronize(h->stream));
    for (int b = 0; b < 2; b++)
        for (int k = 0; k < 3; k++) SWH_HIP(hipMemset(h->adj[b][k], 0, h->plane_elems() * h->esz));
    h->adj_cur = 0;
    return SWH_OK;
}

int swh_adjoint_upload(swh_domain *h, const double *l3)
{
    if (!h || !l3) return fail(SWH_EINVAL, "null argument");
    int rc = adj_check(h);
    if (rc) return rc;
    SWH_HIP(hipStreamSynchronize(h->stream));
    const size_t n = (size_t)h->nx * h->ny;
    for (int k = 0; k < 3; k++)
        if ((rc = upload_plane_any(h, h->adj[h->adj_cur][k], l3 + k * n))) return rc;
    return SWH_OK;
}

int swh_adjoint_download(swh_domain *h, double *l3)
{
    if (!h || !l3) return fail(SWH_EINVAL, "null argument");
    int rc = adj_check(h);
    if (rc) return rc;
    SWH_HIP(hipStreamSynchronize(h->stream));
    const size_t n = (size_t)h->nx * h->ny;
    for (int k = 0; k < 3; k++) {
        rc = h->prec == SWH_PREC_FP64 ? download_plane<double>(h, h->adj[h->adj_cur][k], nullptr, l3 + k * n)
                                      : download_plane<float>(h, h->adj[h->adj_cur][k], nullptr, l3 + k * n);
        if (rc) return rc;
    }
    return SWH_OK;
}

int swh_adjoint_step(swh_domain *h, double dt, int32_t nsteps)
{
    if (!h || !(dt > 0.0) || nsteps < 0) return fail(SWH_EINVAL, "bad adjoint step arguments");
    int rc = adj_check(h);
    if (rc) return rc;
    return h->prec == SWH_PREC_FP64 ? adj_steps<double>(h, dt, nsteps) : adj_steps<float>(h, dt, nsteps);
}

int swh_adjoint_inject(swh_domain *h, const double *values)
{
    if (!h || !values) return fail(SWH_EINVAL, "null argument");
    if (!h->d_gidx) return fail(SWH_ESTATE, "gauges not set");
    int rc = adj_check(h);
    if (rc) return rc;
    return adj_inject_host(h, values);
}

int swh_adjoint_run(swh_domain *h, double dt, int32_t nsteps, int32_t every, const double *grad)
{
    if (!h || !grad || !(dt > 0.0) || nsteps < 0 || every < 1) return fail(SWH_EINVAL, "bad adjoint run arguments");
    if (!h->d_gidx) return fail(SWH_ESTATE, "gauges not set");
    int rc = adj_check(h);
    if (rc) return rc;
    const int ng = h->ng;
    // backward sweep: the sample taken after forward step n (n % every == 0) is
    // injected before the transposed step n is applied
    for (int n = nsteps; n >= 0; n--) {
        if (n % every == 0) {
            if ((rc = adj_inject_host(h, grad + (size_t)(n / every) * ng))) return rc;
        }
        if (n > 0) {
            rc = h->prec == SWH_PREC_FP64 ? adj_steps<double>(h, dt, 1) : adj_steps<float>(h, dt, 1);
            if (rc) return rc;
        }
    }
    return SWH_OK;
}

int swh_evolve_record_gauges_schedule(swh_domain *h, double timestep, int32_t nsamples, const int64_t *sample_steps)
{
    if (!h || nsamples < 1 || !sample_steps || !(timestep > 0.0)) return fail(SWH_EINVAL, "bad arguments");
    if (!h->d_gidx) return fail(SWH_ESTATE, "gauges not set");
    for (int q = 0; q < nsamples; q++)
        if (sample_steps[q] < 0 || (q > 0 && sample_steps[q] <= sample_steps[q - 1]) ||
            sample_steps[q] > (int64_t)INT32_MAX)
            return fail(SWH_EINVAL, "sample steps must be >= 0 and strictly increasing");
    if (h->gsamples + nsamples > h->gmax) return fail(SWH_ESTATE, "gauge sample buffer too small");
    // steps in pairs wherever two_step_ok allows, whatever the sample spacing: a sample
    // due between the two steps of a pair is taken one step ahead by the pair itself
    const long long last = sample_steps[nsamples - 1];
    long long n = 0;
    int q = 0, rc = SWH_OK;
    if (sample_steps[0] == 0) {
        rc = swh_gauges_sample(h);
        q = 1;
    }
    while (!rc && n < last) {
        const bool mid = q < nsamples && sample_steps[q] == n + 1;
        if (n + 2 <= last && two_step_ok(h, timestep) && (!mid || h->ng > 0)) {
            rc = h->prec == SWH_PREC_FP64 ? leapfrog_step2<double, false>(h, timestep, mid)
                                          : leapfrog_step2<float, true>(h, timestep, mid);
            if (mid) q++;
            n += 2;
        } else {
            rc = evolve(h, timestep);
            n += 1;
        }
        if (!rc && q < nsamples && sample_steps[q] == n) {
            rc = swh_gauges_sample(h);
            q++;
        }
    }
    return rc;
}

int swh_checkpoint_save(swh_domain *h, int32_t slot)
{
    if (!h || slot < 0 || slot > 100000) return fail(SWH_EINVAL, "bad checkpoint slot");
    if (!h->leapfrog()) return fail(SWH_ESTATE, "checkpoints are provided for the leapfrog methods");
    const size_t P = h->plane_elems() * h->esz;
    if ((int)h->ckpt.size() <= slot) {
        h->ckpt.resize(slot + 1, nullptr);
        h->ckpt_time.resize(slot + 1, 0.0);
    }
    if (!h->ckpt[slot]) {
        int rc = dmalloc(&h->ckpt[slot], 4 * P);
        if (rc) return rc;
    }
    const int c = h->cur;
    char *dst = (char *)h->ckpt[slot];
    SWH_HIP(hipMemcpyAsync(dst, h->S[c], P, hipMemcpyDeviceToDevice, h->stream));
    SWH_HIP(hipMemcpyAsync(dst + P, h->UH[c], P, hipMemcpyDeviceToDevice, h->stream));
    SWH_HIP(hipMemcpyAsync(dst + 2 * P, h->VH[c], P, hipMemcpyDeviceToDevice, h->stream));
    if (h->C[c]) SWH_HIP(hipMemcpyAsync(dst + 3 * P, h->C[c], P, hipMemcpyDeviceToDevice, h->stream));
    h->ckpt_time[slot] = h->time;
    return SWH_OK;
}

int swh_checkpoint_restore(swh_domain *h, int32_t slot)
{
    if (!h || slot < 0 || slot >= (int)h->ckpt.size() || !h->ckpt[slot])
        return fail(SWH_EINVAL, "no checkpoint in that slot");
    const size_t P = h->plane_elems() * h->esz;
    const int c = h->cur;
    const char *src = (const char *)h->ckpt[slot];
    SWH_HIP(hipMemcpyAsync(h->S[c], src, P, hipMemcpyDeviceToDevice, h->stream));
    SWH_HIP(hipMemcpyAsync(h->UH[c], src + P, P, hipMemcpyDeviceToDevice, h->stream));
    SWH_HIP(hipMemcpyAsync(h->VH[c], src + 2 * P, P, hipMemcpyDeviceToDevice, h->stream));
    if (h->C[c]) SWH_HIP(hipMemcpyAsync(h->C[c], src + 3 * P, P, hipMemcpyDeviceToDevice, h->stream));
    h->time = h->ckpt_time[slot];
    return SWH_OK;
}

int swh_adjoint_run_schedule(swh_domain *h, double dt, int32_t nsteps, int32_t nsamples, const int64_t *sample_steps,
                             const double *grad)
{
    if (!h || !grad || !sample_steps || !(dt > 0.0) || nsteps < 0 || nsamples < 0)
        return fail(SWH_EINVAL, "bad adjoint run arguments");
    if (!h->d_gidx) return fail(SWH_ESTATE, "gauges not set");
    for (int q = 0; q < nsamples; q++)
        if (sample_steps[q] < 0 || sample_steps[q] > nsteps || (q > 0 && sample_steps[q] <= sample_steps[q - 1]))
            return fail(SWH_EINVAL, "sample steps must be in 0..nsteps and strictly increasing");
    int rc = adj_check(h);
    if (rc) return rc;
    const int ng = h->ng;
    // every sample's gradient on the device at once (no host round trip in the sweep)
    const size_t need = (size_t)(nsamples > 0 ? nsamples : 1) * (ng > 0 ? ng : 1);
    if (h->n_ggrad < need) {
        SWH_HIP(hipStreamSynchronize(h->stream));
        dfree(h->d_ggrad);
        h->d_ggrad = nullptr;
        h->n_ggrad = 0;
        if ((rc = dmalloc((void **)&h->d_ggrad, sizeof(double) * need))) return rc;
        h->n_ggrad = need;
    }
    if (nsamples > 0 && ng > 0) {
        SWH_HIP(hipMemcpyAsync(h->d_ggrad, grad, sizeof(double) * nsamples * ng, hipMemcpyHostToDevice, h->stream));
        // grad must stay valid until the copy has been consumed
        SWH_HIP(hipStreamSynchronize(h->stream));
    }
    // backward sweep: the sample taken after forward step n is injected before the
    // transposed step n is applied
    int q = nsamples - 1;
    for (int n = nsteps; n >= 0; n--) {
        if (q >= 0 && sample_steps[q] == n) {
            void *lS = h->adj[h->adj_cur][0];
            const double *v = h->d_ggrad + (size_t)q * ng;
            if (h->prec == SWH_PREC_FP64)
                SWH_HIP(launch_adjoint_inject<double>((double *)lS, ng, h->d_gidx, v, h->stream));
            else
                SWH_HIP(launch_adjoint_inject<float>((float *)lS, ng, h->d_gidx, v, h->stream));
            q--;
        }
        if (n > 0) {
            rc = h->prec == SWH_PREC_FP64 ? adj_steps<double>(h, dt, 1) : adj_steps<float>(h, dt, 1);
            if (rc) return rc;
        }
    }
    return SWH_OK;
}

int swh_evolve_record_gauges(swh_domain *h, double timestep, int32_t nsteps, int32_t every)
{
    if (!h || nsteps < 0 || every < 1) return fail(SWH_EINVAL, "bad arguments");
    if (!h->d_gidx) return fail(SWH_ESTATE, "gauges not set");
    if (h->leapfrog() && timestep > 0.0) {
        // the schedule 0, every, 2 every, ... (steps in pairs), then the steps after it
        std::vector<int64_t> steps;
        for (long long n = 0; n <= nsteps; n += every) steps.push_back(n);
        int rc = swh_evolve_record_gauges_schedule(h, timestep, (int32_t)steps.size(), steps.data());
        if (!rc && nsteps > steps.back()) rc = swh_evolve_steps(h, timestep, (int32_t)(nsteps - steps.back()));
        return rc;
    }
    int rc = swh_gauges_sample(h);
    if (rc) return rc;
    for (int n = 1; n <= nsteps; n++) {
        if ((rc = evolve(h, timestep))) return rc;
        if (n % every == 0 && (rc = swh_gauges_sample(h))) return rc;
    }
    return SWH_OK;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Batched unit sources: one shared grid (metrics, elevation, friction work, row
// coefficients held by an embedded single-grid handle) and nsrc state sets laid out
// source-major, each source a pitch x ny plane per field. Sources are kept in slots
// sorted by decreasing step count, so the sources still running at any step are a
// prefix of the slots and one launch with gridDim.z = n_active covers them.
// ---------------------------------------------------------------------------
struct swh_batch {
    swh_domain *g = nullptr;  // shared grid
    int nsrc = 0;
    void *S[2] = {}, *UH[2] = {}, *VH[2] = {}, *C[2] = {};
    std::vector<int> nsteps_slot;  // steps of the source in each slot (descending)
    std::vector<int> slot_of;      // user index -> slot
    int nsteps_set = 0;
    long long step = 0;
    int cur = 0;                   // buffer holding the running sources' state
    std::vector<int> slot_buf;     // buffer holding each finished source's final state
    // gauges
    int ng = 0, gmax = 0;
    std::vector<long long> sched;  // steps (ascending) at which the gauges are sampled
    long long *d_gidx = nullptr;
    int *d_gij = nullptr;          // the gauges' (i, j), 0-based, for the fused kernel's sampling
    std::vector<int> h_gij;
    double *d_gval = nullptr;      // [slot][sample][3][ng]
    int block_x = 256, rows_per_block = 32, variant = 2, src_fast = 0;
    int two_step = 0;              // two fused steps per launch where possible
    int fuse = 2;                  // most fused steps per launch (2..4) where two_step is on
    int xcd_order = 0;             // two-step launches: XCD-aware tile order
    int src_group = 1;             // ... interleaving the tiles of this many sources
    // step-kernel launches so far: count, sources summed over launches, and source
    // steps (sources x fused steps) summed over launches
    long long n_launch = 0, src_launch = 0, src_steps = 0;
};

namespace {

template <typename T>
__global__ void batch_gauges_kernel(const T *__restrict__ S, const T *__restrict__ C, const T *__restrict__ UH,
                                    const T *__restrict__ VH, long long stride, int n,
                                    const long long *__restrict__ idx, double *values, int gmax, int sample)
{
    const int g = blockIdx.x * blockDim.x + threadIdx.x;
    const int slot = blockIdx.y;
    if (g >= n) return;
    const long long o = (long long)slot * stride + idx[g];
    double *v = values + ((long long)slot * gmax + sample) * 3 * n;
    v[g] = C ? (double)S[o] - (double)C[o] : (double)S[o];
    v[n + g] = (double)UH[o];
    v[2 * n + g] = (double)VH[o];
}

// Number of scheduled samples at steps <= s (so the index of the sample due at step s,
// if there is one, is this count - 1).
int samples_upto(const swh_batch *b, long long s)
{
    return (int)(std::upper_bound(b->sched.begin(), b->sched.end(), s) - b->sched.begin());
}

bool sample_due(const swh_batch *b, long long s)
{
    return b->ng && std::binary_search(b->sched.begin(), b->sched.end(), s);
}

int batch_active(const swh_batch *b, long long step)
{
    // number of slots whose step count exceeds 'step' (slots sorted descending)
    int lo = 0, hi = b->nsrc;
    while (lo < hi) {
        int mid = (lo + hi) / 2;
        if (b->nsteps_slot[mid] > step) lo = mid + 1;
        else hi = mid;
    }
    return lo;
}

template <typename T, bool KAHAN>
int batch_sample(swh_batch *b, int n_active, int buf)
{
    if (!sample_due(b, b->step) || n_active <= 0) return SWH_OK;
    const long long sample = samples_upto(b, b->step) - 1;
    swh_domain *g = b->g;
    dim3 grid((b->ng + 63) / 64, n_active);
    hipLaunchKernelGGL(batch_gauges_kernel<T>, grid, dim3(64), 0, g->stream, (const T *)b->S[buf],
                       KAHAN ? (const T *)b->C[buf] : (const T *)nullptr, (const T *)b->UH[buf], (const T *)b->VH[buf],
                       (long long)g->plane_elems(), b->ng, b->d_gidx, b->d_gval, b->gmax, (int)sample);
    SWH_HIP(hipGetLastError());
    return SWH_OK;
}

template <typename T, bool KAHAN>
int batch_args(swh_batch *b, double dt, LFArgs<T> &a, int &fric_mode)
{
    swh_domain *g = b->g;
    if (g->rc_dt != dt) {
        int rc = build_rc<T>(g, dt);
        if (rc) return rc;
    }
    fric_mode = 0;
    if (g->method == SWH_METHOD_LLPNF) fric_mode = 1;
    if (fric_mode == 1 && !g->fw_ready) return fail(SWH_ESTATE, "friction_work not built: upload elevation first");
    a.nx = g->nx; a.ny = g->ny; a.pitch = g->pitch; a.src_stride = (long long)g->plane_elems();
    a.E = (const T *)g->E;
    a.rc = rc_view<T>(g);
    a.fwU = (const T *)g->fwU; a.fwV = (const T *)g->fwV; a.nsq = (const T *)g->nsq;
    a.msl = (T)g->msl; a.dt = (T)dt; a.dt_lfc = (T)(dt * g->lfc);
    a.lin_fric_mult = (T)(1.0 / (1.0 + dt * g->lfc));
    a.fric_power = (T)((g->friction_type == 1) ? -2.0 : -7.0 / 3.0);
    a.bc_mask = 0; a.bc_stage = T(0);
    a.maxq_mask = 0; a.mx = nullptr; a.t_after = T(0); a.arrival_stage = T(0);
    return SWH_OK;
}

// One batched step reading buffer c and writing the other.
template <typename T, bool KAHAN>
int batch_launch(swh_batch *b, LFArgs<T> &a, int fric_mode, int n_active, int c)
{
    swh_domain *g = b->g;
    const int o = 1 - c;
    a.rows_per_block = b->rows_per_block;
    a.src_fast = b->src_fast;
    a.S = (const T *)b->S[c]; a.C = (const T *)b->C[c]; a.UH = (const T *)b->UH[c]; a.VH = (const T *)b->VH[c];
    a.So = (T *)b->S[o]; a.Co = (T *)b->C[o]; a.UHo = (T *)b->UH[o]; a.VHo = (T *)b->VH[o];
    SWH_HIP(launch_leapfrog<T, KAHAN>(a, g->coriolis, fric_mode, g->truely_linear, g->lfc != 0.0, n_active,
                                      b->block_x, b->variant, g->stream));
    b->n_launch++; b->src_launch += n_active; b->src_steps += n_active;
    return SWH_OK;
}

// The step arguments with buffer c as the input state.
template <typename T>
const LFArgs<T> &batch_cur_args(swh_batch *b, LFArgs<T> &a, int c)
{
    a.S = (const T *)b->S[c]; a.C = (const T *)b->C[c]; a.UH = (const T *)b->UH[c]; a.VH = (const T *)b->VH[c];
    return a;
}

// One launch advancing slots [first, first + n) by two steps (reads buffer c).
template <typename T, bool KAHAN>
int batch_launch2(swh_batch *b, LFArgs<T> &a, int fric_mode, int first, int n, int c)
{
    swh_domain *g = b->g;
    const int o = 1 - c;
    a.rows_per_block = b->rows_per_block;
    a.src_fast = 0;
    a.src_base = first;
    a.S = (const T *)b->S[c]; a.C = (const T *)b->C[c]; a.UH = (const T *)b->UH[c]; a.VH = (const T *)b->VH[c];
    a.So = (T *)b->S[o]; a.Co = (T *)b->C[o]; a.UHo = (T *)b->UH[o]; a.VHo = (T *)b->VH[o];
    a.lf2_order = b->xcd_order;
    a.lf2_group = b->src_group;
    SWH_HIP(launch_leapfrog2<T, KAHAN>(a, g->coriolis, fric_mode, g->truely_linear, g->lfc != 0.0, n, b->block_x,
                                       g->stream, (b->variant & 2) ? 1 : 0));
    b->n_launch++; b->src_launch += n; b->src_steps += 2LL * n;
    a.src_base = 0;
    a.lf2_order = 0;
    return SWH_OK;
}

// The fused kernel's lanes hold two gauges each (8-bit indices, 16-bit rows): at most
// two gauges may share a column within one row strip of the launch shape.
bool gauges_fit_lanes(const swh_batch *b)
{
    if (b->ng > 255 || b->g->ny > 65535) return false;
    const int rows = std::max(1, b->rows_per_block);
    for (int p = 0; p < b->ng; p++) {
        int same = 0;
        for (int q = 0; q < b->ng; q++)
            same += b->h_gij[2 * q] == b->h_gij[2 * p] && b->h_gij[2 * q + 1] / rows == b->h_gij[2 * p + 1] / rows;
        if (same > 2) return false;
    }
    return true;
}

// One launch advancing slots [first, first + n) by ns (3 to 6) steps (reads buffer c);
// with gauges_inside, the gauge samples due after any of those steps are taken by the
// launch itself.
template <typename T, bool KAHAN>
int batch_launchn(swh_batch *b, LFArgs<T> &a, int fric_mode, int first, int n, int c, int ns,
                  bool gauges_inside = false)
{
    swh_domain *g = b->g;
    const int o = 1 - c;
    a.rows_per_block = b->rows_per_block;
    a.src_fast = 0;
    a.src_base = first;
    a.S = (const T *)b->S[c]; a.C = (const T *)b->C[c]; a.UH = (const T *)b->UH[c]; a.VH = (const T *)b->VH[c];
    a.So = (T *)b->S[o]; a.Co = (T *)b->C[o]; a.UHo = (T *)b->UH[o]; a.VHo = (T *)b->VH[o];
    a.lf2_order = b->xcd_order;
    a.lf2_group = b->src_group;
    if (gauges_inside) {
        // the one sample that falls after one of the launch's steps (the caller checks)
        a.g_n = b->ng; a.g_max = b->gmax; a.g_ij = b->d_gij; a.g_val = b->d_gval;
        a.g_level = 0;
        for (int q = 1; q <= ns && !a.g_level; q++)
            if (sample_due(b, b->step + q)) {
                a.g_level = q;
                a.g_sample = samples_upto(b, b->step + q) - 1;
            }
    }
    const hipError_t e = launch_leapfrogn<T, KAHAN>(a, g->coriolis, fric_mode, g->truely_linear, g->lfc != 0.0, n,
                                                    b->block_x, g->stream, (b->variant & 2) ? 1 : 0, ns);
    a.g_n = 0; a.g_level = 0;
    SWH_HIP(e);
    b->n_launch++; b->src_launch += n; b->src_steps += (long long)ns * n;
    a.src_base = 0;
    a.lf2_order = 0;
    return SWH_OK;
}

template <typename T, bool KAHAN>
int batch_evolve(swh_batch *b, double dt, int nsteps)
{
    LFArgs<T> a;
    int fric_mode = 0;
    int rc = batch_args<T, KAHAN>(b, dt, a, fric_mode);
    if (rc) return rc;
    if (b->step == 0) {
        rc = batch_sample<T, KAHAN>(b, batch_active(b, -1), b->cur);
        if (rc) return rc;
    }
    const bool can2 = b->two_step && fric_mode <= 1 && (long long)b->g->plane_elems() < (1LL << 31);
    const int depth = can2 ? std::max(2, std::min(6, b->fuse)) : 1;   // most steps per launch
    for (int k = 0; k < nsteps;) {
        const int n_active = batch_active(b, b->step);
        if (n_active == 0) break;
        const int c = b->cur, o = 1 - c;
        const bool sample_next = sample_due(b, b->step + 1);
        int n_last;   // sources that took the last step of this round
        const int n2 = can2 ? batch_active(b, b->step + 1) : 0;   // sources with two or more steps left
        // a round of the full depth that every running source takes to its end, the
        // gauge samples inside it written by the launch itself
        int Dg = 0;
        if (depth >= 3 && gauges_fit_lanes(b)) {
            Dg = std::min<long long>(depth, nsteps - k);
            while (Dg >= 3 && batch_active(b, b->step + Dg - 1) < n_active) Dg--;
            // at most one sample inside the round (the launch writes one)
            int due = 0;
            for (int q = 1; q <= Dg; q++) due += sample_due(b, b->step + q) ? 1 : 0;
            if (due > 1) Dg = 0;
        }
        if (Dg >= 3) {
            rc = batch_launchn<T, KAHAN>(b, a, fric_mode, 0, n_active, c, Dg, true);
            if (rc) return rc;
            b->step += Dg;
            k += Dg;
            b->cur = o;
            for (int sl = batch_active(b, b->step); sl < n_active; sl++) b->slot_buf[sl] = o;
            continue;
        }
        // otherwise the deepest round (3 to 6 steps) that no gauge sample falls inside
        int D = 0;
        if (depth >= 3 && k + 2 < nsteps && !sample_next && !sample_due(b, b->step + 2) &&
            batch_active(b, b->step + 2) > 0) {
            D = 3;
            while (D < depth && k + D < nsteps && !sample_due(b, b->step + D) && batch_active(b, b->step + D) > 0) D++;
        }
        if (D >= 3) {
            // D steps in one launch for the sources with D or more steps left; those with
            // fewer take them in shallower launches, all reading this round's buffer
            int hi = n_active;   // sources [0, hi) have at least d steps left in the loop below
            for (int d = 1; d <= D; d++) {
                const int nd = d < D ? batch_active(b, b->step + d) : 0;   // sources with more than d steps left
                const int lo = nd;
                if (hi > lo) {
                    if (d == 1) {
                        a.src_base = lo;
                        rc = batch_launch<T, KAHAN>(b, a, fric_mode, hi - lo, c);
                        a.src_base = 0;
                    } else if (d == 2) {
                        rc = batch_launch2<T, KAHAN>(b, a, fric_mode, lo, hi - lo, c);
                    } else {
                        rc = batch_launchn<T, KAHAN>(b, a, fric_mode, lo, hi - lo, c, d);
                    }
                    if (rc) return rc;
                }
                hi = lo;
            }
            n_last = batch_active(b, b->step + D - 1);
            b->step += D;
            k += D;
        } else if (n2 > 0 && k + 1 < nsteps) {
            // those take both steps in one launch; the ones with one step left take
            // it in a one-step launch. A gauge sample due after the first step is
            // taken one step ahead from the current state (same operations).
            if (sample_next) {
                a.src_base = 0;
                SWH_HIP(launch_gauge_step1<T, KAHAN>(batch_cur_args(b, a, c), b->g->coriolis, fric_mode,
                                                     b->g->truely_linear, b->g->lfc != 0.0, n_active, b->ng,
                                                     b->d_gidx, b->d_gval, b->gmax,
                                                     samples_upto(b, b->step + 1) - 1, nullptr, 0.0,
                                                     b->g->stream));
            }
            rc = batch_launch2<T, KAHAN>(b, a, fric_mode, 0, n2, c);
            if (rc) return rc;
            if (n_active > n2) {
                a.src_base = n2;
                rc = batch_launch<T, KAHAN>(b, a, fric_mode, n_active - n2, c);
                a.src_base = 0;
                if (rc) return rc;
            }
            b->step += 2;
            k += 2;
            n_last = n2;
        } else {
            rc = batch_launch<T, KAHAN>(b, a, fric_mode, n_active, c);
            if (rc) return rc;
            b->step++;
            k++;
            n_last = n_active;
        }
        b->cur = o;
        for (int sl = batch_active(b, b->step); sl < n_active; sl++) b->slot_buf[sl] = o;
        if (sample_due(b, b->step) && n_last > 0) {
            rc = batch_sample<T, KAHAN>(b, n_last, o);
            if (rc) return rc;
        }
    }
    return SWH_OK;
}

// Launch-shape autotuner: times every candidate shape on the batch's own grid and
// sources by repeating the next step(s) (each launch reads the current buffer and
// writes the other, which the next real step overwrites), so the state and step count
// are unchanged. One-step shapes: variant 0..15 x block_x {64,128,256} x rows
// {16,32,64}; then the fused kernels: 2..5 steps per launch x block_x x rows
// {32,...,256} x tile order (hardware, or XCD order with sources interleaved in groups
// of 16), compared per cell-update of the running sources.
template <typename T, bool KAHAN>
int batch_autotune(swh_batch *b, double dt, int reps, int32_t *best, double *best_ms)
{
    swh_domain *g = b->g;
    LFArgs<T> a;
    int fric_mode = 0;
    int rc = batch_args<T, KAHAN>(b, dt, a, fric_mode);
    if (rc) return rc;
    const int n_active = batch_active(b, b->step);
    if (n_active == 0) return fail(SWH_ESTATE, "autotune: no source is still running");
    const int c = b->cur;
    hipEvent_t e0, e1;
    SWH_HIP(hipEventCreate(&e0));
    SWH_HIP(hipEventCreate(&e1));
    const int bxs[3] = {64, 128, 256};
    double bms = 1e300;   // ms per step of n_active sources
    int bb = b->block_x, br = b->rows_per_block, bv = b->variant | (b->src_fast << 3) | (b->two_step << 4);
    int bx_order = b->xcd_order, b_group = b->src_group, b_fuse = b->fuse;
    // one launch of depth D (1..5) over the sources with D or more steps left
    auto launch = [&](int D, int nD) -> int {
        if (D == 1) return batch_launch<T, KAHAN>(b, a, fric_mode, nD, c);
        if (D == 2) return batch_launch2<T, KAHAN>(b, a, fric_mode, 0, nD, c);
        return batch_launchn<T, KAHAN>(b, a, fric_mode, 0, nD, c, D);
    };
    auto timed = [&](int D, double &ms_out) -> int {
        const int nD = batch_active(b, b->step + D - 1);
        if (nD <= 0) { ms_out = 1e300; return SWH_OK; }
        int r = launch(D, nD);   // warm-up
        if (r) return r;
        if (hipEventRecord(e0, g->stream) != hipSuccess) return fail(SWH_EHIP, "event record");
        for (int k = 0; k < reps; k++)
            if ((r = launch(D, nD))) return r;
        float ms = 0.f;
        if (hipEventRecord(e1, g->stream) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
            hipEventElapsedTime(&ms, e0, e1) != hipSuccess)
            return fail(SWH_EHIP, "event timing");
        // per step of n_active sources
        ms_out = (double)ms / reps / D * ((double)n_active / nD);
        return SWH_OK;
    };
    const int rws1[3] = {16, 32, 64};
    for (int v = 0; v < 16 && rc == SWH_OK; v++)
        for (int ib = 0; ib < 3 && rc == SWH_OK; ib++)
            for (int ir = 0; ir < 3 && rc == SWH_OK; ir++) {
                b->block_x = bxs[ib]; b->rows_per_block = rws1[ir]; b->variant = v & 7; b->src_fast = (v >> 3) & 1;
                double ms;
                rc = timed(1, ms);
                if (!rc && ms < bms) { bms = ms; bb = bxs[ib]; br = rws1[ir]; bv = v; }
            }
    if (!rc && fric_mode <= 1 && (long long)g->plane_elems() < (1LL << 31)) {
        // fused kernels with the next row's loads in flight
        const int v1 = bv | 2;
        b->variant = v1 & 7; b->src_fast = (v1 >> 3) & 1;
        const int orders[2][2] = {{0, 1}, {1, 16}};
        const int rws[4] = {32, 64, 128, 256};
        for (int D = 2; D <= 5 && rc == SWH_OK; D++)
            for (int io = 0; io < 2 && rc == SWH_OK; io++)
                for (int ib = 0; ib < 3 && rc == SWH_OK; ib++)
                    for (int ir = 0; ir < 4 && rc == SWH_OK; ir++) {
                        b->block_x = bxs[ib]; b->rows_per_block = rws[ir];
                        b->xcd_order = orders[io][0]; b->src_group = orders[io][1];
                        double ms;
                        rc = timed(D, ms);
                        if (!rc && ms < bms) {
                            bms = ms; bb = bxs[ib]; br = rws[ir]; bv = v1 | 16;
                            bx_order = orders[io][0]; b_group = orders[io][1]; b_fuse = D;
                        }
                    }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    b->block_x = bb; b->rows_per_block = br; b->variant = bv & 7; b->src_fast = (bv >> 3) & 1;
    b->two_step = (bv >> 4) & 1;
    b->xcd_order = bx_order; b->src_group = b_group;
    b->fuse = b->two_step ? b_fuse : 2;
    if (b->two_step) {
        if (bx_order) {
            int lg = 0;
            while ((1 << lg) < b_group) lg++;
            bv |= 64 | (lg << 7);
        }
        bv |= (b->fuse - 2) << 10;
    }
    if (rc) return rc;
    if (best) { best[0] = bb; best[1] = br; best[2] = bv; }
    if (best_ms) *best_ms = bms;
    return SWH_OK;
}

}  // namespace

extern "C" {

int swh_batch_create(const swh_domain_desc *desc, int32_t nsources, swh_batch **out)
{
    if (!desc || !out || nsources < 1) return fail(SWH_EINVAL, "bad batch arguments");
    *out = nullptr;
    if (desc->timestepping_method != SWH_METHOD_LINEAR && desc->timestepping_method != SWH_METHOD_LLPNF)
        return fail(SWH_EINVAL, "batches support the leapfrog methods only");
    if (desc->timestepping_method == SWH_METHOD_LLPNF && !desc->linear_solver_is_truely_linear)
        return fail(SWH_EINVAL, "batches of 'leapfrog_linear_plus_nonlinear_friction' need "
                                "linear_solver_is_truely_linear (friction work precomputed from the elevation)");
    swh_domain_desc d = *desc;
    d.max_U_mask = 0;
    swh_batch *b = new swh_batch();
    int rc = swh_domain_create(&d, &b->g);
    if (rc) {
        delete b;
        return rc;
    }
    b->nsrc = nsources;
    // default launch: five (FP32) / three (FP64) fused steps per launch, 2-wave blocks,
    // XCD tile order with sources in groups of 16 (the fastest shapes measured on the
    // 512-source 1200^2 workload; swh_batch_autotune refines it per batch)
    b->block_x = 256; b->two_step = 1; b->src_fast = 0;
    b->xcd_order = 1; b->src_group = 16;
    if (b->g->prec == SWH_PREC_FP64) { b->block_x = 128; b->rows_per_block = 128; b->variant = 3; b->fuse = 3; }
    else { b->block_x = 128; b->rows_per_block = 256; b->variant = 3; b->fuse = 5; }
    const size_t bytes = b->g->plane_elems() * b->g->esz * (size_t)nsources;
    const bool kahan = b->g->prec == SWH_PREC_FP32K;
    for (int k = 0; k < 2 && rc == SWH_OK; k++) {
        rc = dmalloc(&b->S[k], bytes);
        if (!rc) rc = dmalloc(&b->UH[k], bytes);
        if (!rc) rc = dmalloc(&b->VH[k], bytes);
        if (!rc && kahan) rc = dmalloc(&b->C[k], bytes);
    }
    if (rc) {
        swh_batch_destroy(b);
        return rc;
    }
    b->nsteps_slot.assign(nsources, 0);
    b->slot_buf.assign(nsources, 0);
    b->slot_of.resize(nsources);
    for (int k = 0; k < nsources; k++) b->slot_of[k] = k;
    *out = b;
    return SWH_OK;
}

int swh_batch_destroy(swh_batch *b)
{
    if (!b) return SWH_OK;
    if (b->g && b->g->stream) (void)hipStreamSynchronize(b->g->stream);
    for (int k = 0; k < 2; k++) { dfree(b->S[k]); dfree(b->UH[k]); dfree(b->VH[k]); dfree(b->C[k]); }
    dfree(b->d_gidx); dfree(b->d_gval); dfree(b->d_gij);
    swh_domain_destroy(b->g);
    delete b;
    return SWH_OK;
}

int swh_batch_upload_elevation(swh_batch *b, const double *elev)
{
    if (!b || !elev) return fail(SWH_EINVAL, "null argument");
    swh_domain *g = b->g;
    SWH_HIP(hipStreamSynchronize(g->stream));
    int rc = upload_plane_any(g, g->E, elev);
    if (rc) return rc;
    g->h_elev.assign(elev, elev + (size_t)g->nx * g->ny);
    return build_friction_work(g);
}

int swh_batch_upload_manning_squared(swh_batch *b, const double *m)
{
    if (!b) return fail(SWH_EINVAL, "null argument");
    return swh_upload_manning_squared(b->g, m);
}

int swh_batch_set_nsteps(swh_batch *b, const int32_t *ns)
{
    if (!b || !ns) return fail(SWH_EINVAL, "null argument");
    if (b->step != 0) return fail(SWH_ESTATE, "step counts must be set before the batch starts");
    std::vector<int> order(b->nsrc);
    for (int k = 0; k < b->nsrc; k++) {
        if (ns[k] < 0) return fail(SWH_EINVAL, "negative step count");
        order[k] = k;
    }
    std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return ns[x] > ns[y]; });
    for (int s = 0; s < b->nsrc; s++) {
        b->slot_of[order[s]] = s;
        b->nsteps_slot[s] = ns[order[s]];
    }
    b->nsteps_set = 1;
    return SWH_OK;
}

int swh_batch_upload_source(swh_batch *b, int32_t k, const double *U3)
{
    if (!b || !U3 || k < 0 || k >= b->nsrc) return fail(SWH_EINVAL, "bad source upload");
    if (!b->nsteps_set) return fail(SWH_ESTATE, "call swh_batch_set_nsteps before uploading sources");
    if (b->step != 0) return fail(SWH_ESTATE, "sources must be uploaded before the batch starts");
    swh_domain *g = b->g;
    SWH_HIP(hipStreamSynchronize(g->stream));
    const size_t n = (size_t)g->nx * g->ny, off = (size_t)b->slot_of[k] * g->plane_elems() * g->esz;
    const int c = b->cur;
    int rc;
    if ((rc = upload_plane_any(g, (char *)b->S[c] + off, U3, b->C[c] ? (char *)b->C[c] + off : nullptr))) return rc;
    if ((rc = upload_plane_any(g, (char *)b->UH[c] + off, U3 + n))) return rc;
    if ((rc = upload_plane_any(g, (char *)b->VH[c] + off, U3 + 2 * n))) return rc;
    return SWH_OK;
}

int swh_batch_download_source(swh_batch *b, int32_t k, double *U3)
{
    if (!b || !U3 || k < 0 || k >= b->nsrc) return fail(SWH_EINVAL, "bad source download");
    swh_domain *g = b->g;
    SWH_HIP(hipStreamSynchronize(g->stream));
    const int slot = b->slot_of[k];
    const int buf = b->nsteps_slot[slot] > b->step ? b->cur : b->slot_buf[slot];
    const size_t n = (size_t)g->nx * g->ny, off = (size_t)slot * g->plane_elems() * g->esz;
    int rc;
    if (g->prec == SWH_PREC_FP64) {
        if ((rc = download_plane<double>(g, (char *)b->S[buf] + off, nullptr, U3))) return rc;
        if ((rc = download_plane<double>(g, (char *)b->UH[buf] + off, nullptr, U3 + n))) return rc;
        if ((rc = download_plane<double>(g, (char *)b->VH[buf] + off, nullptr, U3 + 2 * n))) return rc;
    } else {
        if ((rc = download_plane<float>(g, (char *)b->S[buf] + off, (char *)b->C[buf] + off, U3))) return rc;
        if ((rc = download_plane<float>(g, (char *)b->UH[buf] + off, nullptr, U3 + n))) return rc;
        if ((rc = download_plane<float>(g, (char *)b->VH[buf] + off, nullptr, U3 + 2 * n))) return rc;
    }
    return SWH_OK;
}

int swh_batch_gauges_set(swh_batch *b, int32_t n, const int32_t *ij, int32_t every, int32_t max_samples)
{
    if (!b || every < 1 || max_samples < 1) return fail(SWH_EINVAL, "bad gauge arguments");
    std::vector<int64_t> steps(max_samples);
    for (int q = 0; q < max_samples; q++) steps[q] = (int64_t)q * every;
    return swh_batch_gauges_set_schedule(b, n, ij, max_samples, steps.data());
}

int swh_batch_gauges_set_schedule(swh_batch *b, int32_t n, const int32_t *ij, int32_t nsamples,
                                  const int64_t *sample_steps)
{
    if (!b || n < 0 || (n > 0 && !ij) || nsamples < 1 || !sample_steps) return fail(SWH_EINVAL, "bad gauge arguments");
    for (int q = 0; q < nsamples; q++)
        if (sample_steps[q] < 0 || (q > 0 && sample_steps[q] <= sample_steps[q - 1]))
            return fail(SWH_EINVAL, "sample steps must be >= 0 and strictly increasing");
    if (b->step > 0) return fail(SWH_ESTATE, "set the gauges before the first step");
    const int max_samples = nsamples;
    swh_domain *g = b->g;
    std::vector<long long> idx(n);
    for (int q = 0; q < n; q++) {
        const int i = ij[2 * q] - 1, j = ij[2 * q + 1] - 1;
        if (i < 0 || i >= g->nx || j < 0 || j >= g->ny) return fail(SWH_EINVAL, "gauge cell outside the domain");
        idx[q] = (long long)j * g->pitch + i;
    }
    std::vector<int> gij(2 * (size_t)n);
    for (int q = 0; q < n; q++) { gij[2 * q] = ij[2 * q] - 1; gij[2 * q + 1] = ij[2 * q + 1] - 1; }
    SWH_HIP(hipStreamSynchronize(g->stream));
    dfree(b->d_gidx); dfree(b->d_gval); dfree(b->d_gij);
    b->d_gidx = nullptr; b->d_gval = nullptr; b->d_gij = nullptr;
    int rc = dmalloc((void **)&b->d_gidx, sizeof(long long) * (n > 0 ? n : 1));
    if (!rc) rc = dmalloc((void **)&b->d_gij, sizeof(int) * 2 * (n > 0 ? n : 1));
    if (!rc) rc = dmalloc((void **)&b->d_gval, sizeof(double) * 3 * (size_t)(n > 0 ? n : 1) * max_samples * b->nsrc);
    if (rc) return rc;
    if (n > 0) SWH_HIP(hipMemcpy(b->d_gidx, idx.data(), sizeof(long long) * n, hipMemcpyHostToDevice));
    if (n > 0) SWH_HIP(hipMemcpy(b->d_gij, gij.data(), sizeof(int) * 2 * n, hipMemcpyHostToDevice));
    b->h_gij = gij;
    b->ng = n; b->gmax = max_samples;
    b->sched.assign(sample_steps, sample_steps + nsamples);
    return SWH_OK;
}

int swh_batch_evolve(swh_batch *b, double dt, int32_t nsteps)
{
    if (!b) return fail(SWH_EINVAL, "null argument");
    if (!(dt > 0.0) || nsteps < 0) return fail(SWH_EINVAL, "dt must be > 0 and nsteps >= 0");
    if (!b->nsteps_set) return fail(SWH_ESTATE, "call swh_batch_set_nsteps first");
    return b->g->prec == SWH_PREC_FP64 ? batch_evolve<double, false>(b, dt, nsteps)
                                       : batch_evolve<float, true>(b, dt, nsteps);
}

int swh_batch_gauges_download(swh_batch *b, int32_t k, double *values, int32_t *nsamples)
{
    if (!b || !values || !nsamples || k < 0 || k >= b->nsrc) return fail(SWH_EINVAL, "bad gauge download");
    SWH_HIP(hipStreamSynchronize(b->g->stream));
    const int slot = b->slot_of[k];
    const long long taken = std::min<long long>(b->step, b->nsteps_slot[slot]);
    const long long ns = samples_upto(b, taken);
    if (b->ng > 0 && ns > 0)
        SWH_HIP(hipMemcpy(values, b->d_gval + (size_t)slot * b->gmax * 3 * b->ng, sizeof(double) * 3 * b->ng * ns,
                          hipMemcpyDeviceToHost));
    *nsamples = (int32_t)ns;
    return SWH_OK;
}

int swh_batch_launch_counts(swh_batch *b, int64_t *launches, int64_t *source_launches, int64_t *source_steps)
{
    if (!b || !launches || !source_launches || !source_steps) return fail(SWH_EINVAL, "null argument");
    *launches = b->n_launch;
    *source_launches = b->src_launch;
    *source_steps = b->src_steps;
    return SWH_OK;
}

int swh_batch_status(swh_batch *b, int64_t *step, int32_t *n_active)
{
    if (!b || !step || !n_active) return fail(SWH_EINVAL, "null argument");
    *step = b->step;
    *n_active = batch_active(b, b->step);
    return SWH_OK;
}

static int check_launch(int32_t block_x, int32_t rows_per_block, int32_t variant)
{
    if (block_x < 64 || block_x > 256 || (block_x % 64) || rows_per_block < 1 || variant < 0 || variant > 8191)
        return fail(SWH_EINVAL, "block_x must be 64..256 (multiple of 64), rows_per_block >= 1, variant 0..8191");
    return SWH_OK;
}

int swh_probe_copy(int64_t bytes, int32_t reps, double *ms_per_copy)
{
    if (bytes < 16 || (bytes % 16) || reps < 1 || !ms_per_copy) return fail(SWH_EINVAL, "swh_probe_copy: bad argument");
    void *a = nullptr, *b = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    int rc = dmalloc(&a, (size_t)bytes);
    if (!rc) rc = dmalloc(&b, (size_t)bytes);
    if (!rc && (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess))
        rc = fail(SWH_EHIP, "swh_probe_copy: event creation");
    float ms = 0.f;
    if (!rc && launch_probe_copy(a, b, bytes, nullptr) != hipSuccess) rc = fail(SWH_EHIP, "probe copy launch");
    if (!rc && hipEventRecord(e0, nullptr) != hipSuccess) rc = fail(SWH_EHIP, "event record");
    for (int k = 0; k < reps && !rc; k++)
        if (launch_probe_copy(a, b, bytes, nullptr) != hipSuccess) rc = fail(SWH_EHIP, "probe copy launch");
    if (!rc && (hipEventRecord(e1, nullptr) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
                hipEventElapsedTime(&ms, e0, e1) != hipSuccess))
        rc = fail(SWH_EHIP, "event timing");
    if (!rc) *ms_per_copy = (double)ms / reps;
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    dfree(a);
    dfree(b);
    return rc;
}

int swh_probe_valu(int32_t fp64, int32_t blocks, int32_t iters, int32_t reps, double *ms_per_launch,
                   double *wave_insts)
{
    if (blocks < 1 || blocks > (1 << 20) || iters < 1 || reps < 1 || !ms_per_launch || !wave_insts)
        return fail(SWH_EINVAL, "swh_probe_valu: bad argument");
    void *out = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    int rc = dmalloc(&out, (size_t)blocks * 256 * (fp64 ? 8 : 4));
    if (!rc && (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess))
        rc = fail(SWH_EHIP, "swh_probe_valu: event creation");
    float ms = 0.f;
    if (!rc && launch_probe_valu(fp64, blocks, iters, out, nullptr) != hipSuccess) rc = fail(SWH_EHIP, "probe launch");
    if (!rc && hipEventRecord(e0, nullptr) != hipSuccess) rc = fail(SWH_EHIP, "event record");
    for (int k = 0; k < reps && !rc; k++)
        if (launch_probe_valu(fp64, blocks, iters, out, nullptr) != hipSuccess) rc = fail(SWH_EHIP, "probe launch");
    if (!rc && (hipEventRecord(e1, nullptr) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
                hipEventElapsedTime(&ms, e0, e1) != hipSuccess))
        rc = fail(SWH_EHIP, "event timing");
    if (!rc) {
        *ms_per_launch = (double)ms / reps;
        *wave_insts = (double)blocks * 4 * 8 * iters;   // 4 wavefronts per block, 8 FMAs per round
    }
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    dfree(out);
    return rc;
}

int swh_probe_math(int32_t n, const double *x, const double *y, const double *t, double *pow_dev, double *sin_dev,
                   double *pow_host, double *sin_host)
{
    if (n < 1 || !x || !y || !t || !pow_dev || !sin_dev || !pow_host || !sin_host)
        return fail(SWH_EINVAL, "swh_probe_math: bad argument");
    for (int k = 0; k < n; k++) {
        pow_host[k] = swhm::pow_acc(x[k], y[k]);
        sin_host[k] = swhm::sin_acc(t[k]);
    }
    double *d = nullptr;
    const size_t B = sizeof(double) * (size_t)n;
    int rc = dmalloc((void **)&d, 5 * B);
    if (rc) return rc;
    hipError_t e = hipMemcpy(d, x, B, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(d + n, y, B, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(d + 2 * (size_t)n, t, B, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = launch_probe_math(d, d + n, d + 2 * (size_t)n, d + 3 * (size_t)n, d + 4 * (size_t)n, n, nullptr);
    if (e == hipSuccess) e = hipMemcpy(pow_dev, d + 3 * (size_t)n, B, hipMemcpyDeviceToHost);
    if (e == hipSuccess) e = hipMemcpy(sin_dev, d + 4 * (size_t)n, B, hipMemcpyDeviceToHost);
    dfree(d);
    if (e != hipSuccess) return hip_fail(e, "swh_probe_math", __FILE__, __LINE__);
    return SWH_OK;
}

int swh_batch_autotune(swh_batch *b, double dt, int32_t reps, int32_t *best, double *best_ms)
{
    if (!b || !(dt > 0.0) || reps < 1) return fail(SWH_EINVAL, "swh_batch_autotune: bad argument");
    if (!b->nsteps_set) return fail(SWH_ESTATE, "swh_batch_autotune: set the step counts first");
    return b->g->prec == SWH_PREC_FP64 ? batch_autotune<double, false>(b, dt, reps, best, best_ms)
                                       : batch_autotune<float, true>(b, dt, reps, best, best_ms);
}

int swh_batch_set_launch(swh_batch *b, int32_t block_x, int32_t rows_per_block, int32_t variant)
{
    if (!b) return fail(SWH_EINVAL, "null argument");
    int rc = check_launch(block_x, rows_per_block, variant);
    if (rc) return rc;
    b->block_x = block_x; b->rows_per_block = rows_per_block;
    b->variant = variant & 7; b->src_fast = (variant >> 3) & 1; b->two_step = (variant >> 4) & 1;
    b->xcd_order = (variant >> 6) & 1;
    b->src_group = 1 << ((variant >> 7) & 7);
    b->fuse = 2 + ((variant >> 10) & 7);
    return SWH_OK;
}

int swh_domain_set_launch(swh_domain *h, int32_t block_x, int32_t rows_per_block, int32_t variant)
{
    if (!h) return fail(SWH_EINVAL, "null argument");
    int rc = check_launch(block_x, rows_per_block, variant);
    if (rc) return rc;
    h->block_x = block_x; h->rows_per_block = rows_per_block; h->variant = variant & 7;
    h->two_step = (variant & 32) ? 0 : 1;
    return SWH_OK;
}

int swh_domain_set_launch2(swh_domain *h, int32_t block_x, int32_t rows_per_block)
{
    if (!h) return fail(SWH_EINVAL, "null argument");
    if (block_x < 64 || block_x > 256 || (block_x % 64) || (rows_per_block != 0 && rows_per_block < 4) ||
        rows_per_block > 4096)
        return fail(SWH_EINVAL,
                    "two-step launch: block_x in 64..256 (a multiple of 64), rows_per_block 0 (chosen from the "
                    "kernel's occupancy) or in 4..4096");
    h->block2 = block_x;
    h->rows2 = rows_per_block;
    return SWH_OK;
}

int swh_batch_evolve_timed(swh_batch *b, double dt, int32_t nsteps, double *elapsed_ms)
{
    if (!b || !elapsed_ms) return fail(SWH_EINVAL, "null argument");
    hipEvent_t e0, e1;
    SWH_HIP(hipEventCreate(&e0));
    SWH_HIP(hipEventCreate(&e1));
    SWH_HIP(hipEventRecord(e0, b->g->stream));
    int rc = swh_batch_evolve(b, dt, nsteps);
    SWH_HIP(hipEventRecord(e1, b->g->stream));
    SWH_HIP(hipEventSynchronize(e1));
    float ms = 0.0f;
    SWH_HIP(hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    *elapsed_ms = ms;
    return rc;
}

int swh_batch_superpose(swh_batch *b, int32_t n_events, const int32_t *event_ptr, const int32_t *event_src,
                        const double *event_w, const double *event_slip, double msl, int32_t summary, double *out,
                        int32_t *nvalid, double *kernel_ms)
{
    if (!b || n_events < 1 || !event_ptr || !event_src || !out || !nvalid)
        return fail(SWH_EINVAL, "bad superposition arguments");
    if (summary != SWH_SUPERPOSE_SERIES && summary != SWH_SUPERPOSE_MAX_STAGE)
        return fail(SWH_EINVAL, "summary must be SWH_SUPERPOSE_SERIES or SWH_SUPERPOSE_MAX_STAGE");
    if (b->ng < 1 || !b->d_gval) return fail(SWH_ESTATE, "no gauges recorded (swh_batch_gauges_set)");
    if (event_ptr[0] != 0) return fail(SWH_EINVAL, "event_ptr[0] must be 0");
    for (int e = 0; e < n_events; e++)
        if (event_ptr[e + 1] < event_ptr[e]) return fail(SWH_EINVAL, "event_ptr must be non-decreasing");
    for (int q = 0; q < event_ptr[n_events]; q++)
        if (event_src[q] < 0 || event_src[q] >= b->nsrc) return fail(SWH_EINVAL, "event names a source outside the batch");
    swh_domain *g = b->g;
    SWH_HIP(hipStreamSynchronize(g->stream));
    // samples each slot has recorded (as swh_batch_gauges_download reports them)
    std::vector<int> ns_slot(b->nsrc);
    int ns_all = 0;
    for (int k = 0; k < b->nsrc; k++) {
        const long long taken = std::min<long long>(b->step, b->nsteps_slot[k]);
        ns_slot[k] = samples_upto(b, taken);
        ns_all = std::max(ns_all, ns_slot[k]);
    }
    const int TE = superpose_event_tile();
    const int Kpad = (b->nsrc + 3) / 4 * 4;
    const long long ncol = summary == SWH_SUPERPOSE_SERIES ? (long long)b->gmax * 3 * b->ng : b->ng;
    // events per launch: bounded device output (<= 2^28 doubles) and weight matrix
    long long chunk = std::max<long long>(TE, ((1LL << 28) / std::max<long long>(ncol, Kpad)) / TE * TE);
    chunk = std::min<long long>(chunk, (n_events + TE - 1) / TE * TE);
    double *d_w = nullptr, *d_scale = nullptr, *d_out = nullptr, *d_m = nullptr;
    int *d_nv = nullptr, *d_flags = nullptr;
    // slots whose series hold a non-finite value (a source that blew up): only the
    // events that contain them may see it
    std::vector<int> flagged(b->nsrc, 0);
    bool any_flagged = false;
    int rc = dmalloc((void **)&d_flags, sizeof(int) * b->nsrc);
    if (!rc) {
        hipError_t he = hipMemsetAsync(d_flags, 0, sizeof(int) * b->nsrc, g->stream);
        if (he == hipSuccess)
            he = launch_nonfinite_slots(b->d_gval, b->nsrc, (long long)b->gmax * 3 * b->ng, d_flags, g->stream);
        if (he == hipSuccess) he = hipStreamSynchronize(g->stream);
        if (he == hipSuccess) he = hipMemcpy(flagged.data(), d_flags, sizeof(int) * b->nsrc, hipMemcpyDeviceToHost);
        if (he != hipSuccess) rc = hip_fail(he, "superposition non-finite scan", __FILE__, __LINE__);
        for (int k = 0; k < b->nsrc; k++) any_flagged |= flagged[k] != 0;
    }
    if (!rc) rc = dmalloc((void **)&d_w, sizeof(double) * chunk * Kpad);
    if (!rc && any_flagged) rc = dmalloc((void **)&d_m, sizeof(double) * chunk * Kpad);
    if (!rc) rc = dmalloc((void **)&d_scale, sizeof(double) * chunk);
    if (!rc) rc = dmalloc((void **)&d_nv, sizeof(int) * chunk);
    if (!rc) rc = dmalloc((void **)&d_out, sizeof(double) * chunk * ncol);
    hipEvent_t t0 = nullptr, t1 = nullptr;
    double total_ms = 0.0;
    if (!rc && (hipEventCreate(&t0) != hipSuccess || hipEventCreate(&t1) != hipSuccess))
        rc = fail(SWH_EHIP, "hipEventCreate failed");
    std::vector<double> W, M, scale;
    std::vector<int> nv;
    for (long long e0 = 0; !rc && e0 < n_events; e0 += chunk) {
        const int E = (int)std::min<long long>(chunk, n_events - e0);
        W.assign((size_t)chunk * Kpad, 0.0);
        if (any_flagged) M.assign((size_t)chunk * Kpad, 0.0);
        scale.assign(chunk, 0.0);
        nv.assign(chunk, 0);
        int smax = 1;
        for (int e = 0; e < E; e++) {
            const int ev = (int)e0 + e;
            int n_e = event_ptr[ev + 1] > event_ptr[ev] ? b->gmax : ns_all;
            for (int q = event_ptr[ev]; q < event_ptr[ev + 1]; q++) {
                const int slot = b->slot_of[event_src[q]];
                // uniform slip: unit weights, the slip scales the sum
                // (sum_tsunami_unit_sources.R:658-669); stochastic: per-source slips (:672-687)
                W[(size_t)e * Kpad + slot] += event_w ? event_w[q] : 1.0;
                if (any_flagged && flagged[slot]) M[(size_t)e * Kpad + slot] = 1.0;
                n_e = std::min(n_e, ns_slot[slot]);
            }
            scale[e] = event_slip ? event_slip[ev] : 1.0;
            nv[e] = n_e;
            nvalid[ev] = n_e;
            smax = std::max(smax, n_e);
        }
        if (hipMemcpy(d_w, W.data(), sizeof(double) * W.size(), hipMemcpyHostToDevice) != hipSuccess ||
            hipMemcpy(d_scale, scale.data(), sizeof(double) * chunk, hipMemcpyHostToDevice) != hipSuccess ||
            hipMemcpy(d_nv, nv.data(), sizeof(int) * chunk, hipMemcpyHostToDevice) != hipSuccess ||
            (any_flagged &&
             hipMemcpy(d_m, M.data(), sizeof(double) * M.size(), hipMemcpyHostToDevice) != hipSuccess)) {
            rc = fail(SWH_EHIP, "superposition upload failed");
            break;
        }
        SupArgs a;
        a.X = b->d_gval; a.slot_stride = (long long)b->gmax * 3 * b->ng;
        a.nsrc = b->nsrc; a.ng = b->ng; a.gmax = b->gmax;
        a.W = d_w; a.Kpad = Kpad; a.E = E; a.scale = d_scale; a.nvalid = d_nv; a.msl = msl;
        a.s_chunk = 16; a.s_max = std::min(smax, b->gmax);
        a.out = d_out;
        a.M = any_flagged ? d_m : nullptr;
        hipError_t he = hipEventRecord(t0, g->stream);
        if (he == hipSuccess) he = launch_superpose(a, summary == SWH_SUPERPOSE_SERIES ? 0 : 1, g->stream);
        if (he == hipSuccess) he = hipEventRecord(t1, g->stream);
        if (he == hipSuccess) he = hipEventSynchronize(t1);
        float ms = 0.0f;
        if (he == hipSuccess) he = hipEventElapsedTime(&ms, t0, t1);
        if (he == hipSuccess)
            he = hipMemcpy(out + e0 * ncol, d_out, sizeof(double) * E * ncol, hipMemcpyDeviceToHost);
        if (he != hipSuccess) {
            rc = hip_fail(he, "superposition", __FILE__, __LINE__);
            break;
        }
        total_ms += ms;
    }
    if (t0) (void)hipEventDestroy(t0);
    if (t1) (void)hipEventDestroy(t1);
    dfree(d_w); dfree(d_scale); dfree(d_nv); dfree(d_out); dfree(d_m); dfree(d_flags);
    if (!rc && kernel_ms) *kernel_ms = total_ms;
    return rc;
}

int swh_batch_synchronize(swh_batch *b)
{
    if (!b) return fail(SWH_EINVAL, "null argument");
    SWH_HIP(hipStreamSynchronize(b->g->stream));
    return SWH_OK;
}

}  // extern "C"
