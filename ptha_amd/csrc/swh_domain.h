// The device-side state of one SWALS domain (shared by the C ABI translation units).
#pragma once
#include <vector>

#include "swh_internal.h"

using namespace swh;

struct swh_domain {
    int nx = 0, ny = 0, pitch = 0, fpitch = 0;
    int method = 0, prec = 0, truely_linear = 1, coriolis = 0, spherical = 0, friction_type = 0, n_ext = 2;
    int bex[4] = {1, 1, 1, 1};
    int maxq_mask = 0, maxq_freq = 1;
    double msl = 0.0, cfl = 0.7, theta = 1.6, lfc = 0.0, max_timestep = kMaxTimestep, arrival_thresh = 0.01;
    size_t esz = 8;  // bytes per state element
    hipStream_t stream = nullptr;
    // host copies of the FP64 metrics
    std::vector<double> dbe, dle, area, cor, corbe, tanlat;
    std::vector<double> h_elev, h_nsq;
    // device metrics (FP64)
    double *d_dbe = nullptr, *d_dle = nullptr, *d_area = nullptr, *d_cor = nullptr, *d_tanlat = nullptr;
    // state: ping-pong S, UH, VH (+ Kahan C), shared E
    void *S[2] = {}, *UH[2] = {}, *VH[2] = {}, *C[2] = {}, *E = nullptr;
    int cur = 0;
    void *fwU = nullptr, *fwV = nullptr, *nsq = nullptr, *mx = nullptr;
    // FV work
    void *depth = nullptr, *u = nullptr, *v = nullptr, *fNS = nullptr, *fEW = nullptr;
    double *esU = nullptr, *esV = nullptr, *esJ = nullptr, *esE = nullptr;
    double *d_bmax = nullptr;   // per-block CFL maxima of the flux pass
    // rk2 with a stage boundary and the average folded into the second update: the
    // pre-average UH of columns 1 and nx-2 and VH of rows 1 and ny-2 (element type T)
    void *ring = nullptr;
    // dispersion (leapfrog methods, midpoint/rk2/euler): stored UH/VH/elevation, right-hand side, three
    // extrapolation levels (oldest first) with their times, tridiagonal work
    int disp_on = 0, disp_iter = 2;
    double disp_td1 = 0.0, disp_td2 = -1.0, disp_dlon = 1.0, disp_dlat = 1.0;
    void *dlast = nullptr, *drhs = nullptr, *dq[3] = {}, *dwork = nullptr;
    double *d_qt = nullptr;   // level times (device: finite-volume times live there)
    int dq_n = 0;   // extrapolation levels stored (0..3)
    // output snapshot: dense FP64 [4][ny][nx] copy of the state plus its time, packed on
    // the domain stream; downloaded on a copy stream while the solver runs on
    double *snap = nullptr;
    hipEvent_t snap_ev = nullptr;
    hipStream_t cstream = nullptr;
    int snap_ok = 0;
    // 'leapfrog_nonlinear': nine work planes (swh_nonlinear.hip) and coriolis_bottom_edge
    void *nlw = nullptr;
    double *d_corbe = nullptr;
    // leapfrog row coefficients for rc_dt
    void *rc = nullptr;
    double rc_dt = -1.0;
    DevScalars *sc = nullptr;
    double *d_work = nullptr;  // reductions scratch (2*ny + 8 doubles)
    double *d_stage = nullptr; // host-transfer staging: one dense FP64 plane (nx*ny), lazily
    double *d_stats = nullptr; // statistics scratch: 6 row arrays, 7 results, a counter (6*ny + 8 doubles)
    // host-side scalars (leapfrog keeps time on the host; rk2 on the device)
    double time = 0.0, evolve_step_dt = 0.0, dt_last_update = 0.0;
    long long nsteps = 0;
    // set while a multidomain step is captured into a graph: the leapfrog boundary-flux
    // kernel forms the step's end time from the device time instead of the host's
    int graph_time = 0;
    int fw_ready = 0;
    // forcing patch (forcing_patch_type, forcing_mod.f90:24-56): work planes over the
    // whole grid, zero outside the patch; bounds 0-based inclusive
    void *fS = nullptr, *fUH = nullptr, *fVH = nullptr, *fE = nullptr;
    int f_on = 0, f_i0 = 0, f_i1 = -1, f_j0 = 0, f_j1 = -1;
    double f_start = 0.0, f_end = 0.0;
    // boundary
    int bc_mask = 0;
    double bc_amp = 0.0, bc_period = 1.0, bc_msl = 0.0;
    // 0: transmissive stage boundary (sine stage); 1: Flather, whose outside state is the
    // sine stage when bc_use_fun, else the sea at rest at msl_linear
    int bc_kind = 0, bc_use_fun = 1;
    // gauges
    int ng = 0, gmax = 0, gsamples = 0;
    long long *d_gidx = nullptr;
    double *d_gval = nullptr, *d_gtime = nullptr;
    // adjoint state (ping-pong lS, lU, lV), allocated on first use
    void *adj[2][3] = {};
    int adj_cur = 0;
    double *d_gvec = nullptr;  // one sample of gauge values for injection
    double *d_ggrad = nullptr;     // a whole backward sweep's gauge gradients (swh_adjoint_run_schedule)
    size_t n_ggrad = 0;
    // device checkpoints of the state (S, C, UH, VH planes of the current buffer), by slot
    std::vector<void *> ckpt;
    std::vector<double> ckpt_time;
    // multidomain nesting: box-edge strips whose flux integrals each step accumulates
    const swh::NestStrip *d_nest_strips = nullptr;
    int n_nest_strips = 0;
    int nest_strip_max = 0;   // longest strip (cells), for the accumulation's grid
    double *d_nest_int = nullptr;
    // launch shape of the step kernel
    int block_x = 256, rows_per_block = 32, variant = 2;
    // two leapfrog steps per launch in swh_evolve_steps (where nothing is needed between
    // them): rows per strip of that kernel (0: chosen from its occupancy) and threads
    // per workgroup
    int two_step = 1, rows2 = 0, block2 = 256;
    // partial sums of the boundary-flux line workgroups of the step launches (16
    // doubles per workgroup, room for bf_cap of them)
    double *d_bfpart = nullptr;
    int bf_cap = 0;

    // staggered leapfrog family (time kept on the host, fixed timestep)
    bool leapfrog() const
    {
        return method == SWH_METHOD_LINEAR || method == SWH_METHOD_LLPNF || method == SWH_METHOD_LEAPFROG_NONLINEAR;
    }
    // depths relative to msl_linear ('linear' / llpnf with linear_solver_is_truely_linear)
    bool tl_solver() const { return truely_linear && (method == SWH_METHOD_LINEAR || method == SWH_METHOD_LLPNF); }
    size_t plane_elems() const { return (size_t)pitch * ny; }
};

namespace swh {
// One domain%evolve_one_step on the device (leapfrog or rk2), as swh_evolve_one_step.
int domain_evolve(swh_domain *h, double dt);
// Re-reads the elevation from the device and rebuilds the leapfrog friction work
// from it (the reference builds friction_work on the first step, after any halo
// exchange has set the elevation in the halos).
int domain_refresh_elevation(swh_domain *h);
// compute_domain_statistics_new on the device (swh_stats.hip): rows = 6*ny doubles of
// scratch, out = 7 doubles (max stage, max speed, min stage, min speed, potential,
// kinetic, total energy / rho); pmask = priority mask (nx, ny) or null.
int domain_statistics_launch(swh_domain *h, const int *pmask, double *rows, double *out, hipStream_t stream);
// Stability scan: count (cleared here) of interior values that are NaN or beyond +-HUGE(1.0)/2.
int domain_stability_launch(swh_domain *h, unsigned long long *count, hipStream_t stream);
// Allocates h->d_stats on first use.
int domain_stats_buffer(swh_domain *h);
// Graph capture of multidomain steps: whether a step of h can be captured (its kernels'
// arguments do not depend on the host time or step count; empty string: yes), the
// per-dt set-up a capture must not contain (leapfrog row coefficients), and the host
// bookkeeping of one domain_evolve without its launches (for graph replays).
std::string domain_graph_unsupported(const swh_domain *h);
int domain_prepare_dt(swh_domain *h, double dt);
void domain_host_advance(swh_domain *h, double dt);
}  // namespace swh
