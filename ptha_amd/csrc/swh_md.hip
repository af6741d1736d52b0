// Multidomain on one GPU: the device side of multidomain_type's step
// (propagation/SWALS/src/shallow_water/multidomain_mod.f90:1093-1212). The host
// hands over what the reference's md%setup decided (boxes, priority domains, cell
// sizes); this file turns it into device tables once and then runs each
// multidomain step as: per-domain sub-steps on the domains' own streams, then one
// exchange on the multidomain stream (pack -> unpack -> staggered fix-ups -> flux
// correction -> re-interpolation of fine receive boxes).
#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstdio>
#include <string>
#include <unordered_set>
#include <vector>

#include "swh_domain.h"

namespace {

int md_fail(int code, const std::string &msg)
{
    swh::set_error(msg);
    return code;
}

template <typename V>
int upload_vec(V **dst, const std::vector<V> &v)
{
    *dst = nullptr;
    if (v.empty()) return SWH_OK;
    SWH_HIP(hipMalloc((void **)dst, v.size() * sizeof(V)));
    SWH_HIP(hipMemcpy(*dst, v.data(), v.size() * sizeof(V), hipMemcpyHostToDevice));
    return SWH_OK;
}

inline int nint(double x) { return (int)std::lround(x); }

// timestepping_metadata_mod.f90:118,131,144,155
bool mass_only(int m) { return m == 5 || m == 6 || m == 8; }
bool fc_unsupported(int m) { return m == 7; }

}  // namespace

struct swh_md {
    int nd = 0, prec = 0;
    std::vector<swh_domain *> dom;
    std::vector<double> all_dx;           // 2 per domain
    std::vector<int> methods, trf, dry_all;
    std::vector<std::vector<int>> pdi, pnp;   // (nx, ny) Fortran order == row-major [j][i]
    std::vector<std::vector<swh_nest_box>> send, recv;
    bool finalized = false;
    double time = 0.0;
    // device tables
    hipStream_t stream = nullptr;
    hipEvent_t ev_x = nullptr;
    std::vector<hipEvent_t> ev_dom;
    double *arena = nullptr;
    std::vector<long long> recv_buf_off;      // flattened (domain, recv box) -> arena offset
    std::vector<int> recv_base;               // first flattened index of each domain's recv boxes
    std::vector<double *> d_int, d_err;
    std::vector<long long> n_int;
    std::vector<NestStrip *> d_strips;
    std::vector<int> n_strips;
    std::vector<std::vector<int>> send_strip0, recv_strip0;   // strip-table index of each box's N strip
    std::vector<std::vector<NestStrip>> h_strips;
    double **d_int_tab = nullptr, **d_err_tab = nullptr;
    NestPack *d_packs = nullptr;
    int n_packs = 0, pack_blocks = 0;
    NestUnpack *d_ups = nullptr, *d_fixups = nullptr;
    int n_ups = 0, up_blocks = 0, n_fixups = 0, fixup_blocks = 0;
    NestFixDom *d_fixdoms = nullptr;
    int n_fixdoms = 0;
    std::vector<int2 *> fix_level;            // per group level: (fixdom, update) of every domain's group
    std::vector<int> fix_level_n;
    std::vector<void *> owned;                // fix and group-level arrays
    std::vector<int *> d_pnp;
    NestReint *d_reint = nullptr;
    int n_reint = 0, reint_blocks = 0, reint_max_len = 0;
    double *d_copy = nullptr;
    double *d_vol = nullptr;
    // captured global steps (swh_md_graph_capture): one graph, or two when the domains'
    // ping-pong buffers alternate between global steps; the device copy of the
    // multidomain time the graphs read and advance; graph_dt the captured step
    hipGraphExec_t gexec[2] = {nullptr, nullptr};
    int n_graphs = 0;
    std::vector<int> cap_cur;                 // each domain's current buffer when captured
    double graph_dt = 0.0;
    bool capturing = false;
    double *d_time = nullptr;
};

namespace {

template <typename T>
NestDoms<T> dom_ptrs(swh_md *md)
{
    NestDoms<T> t{};
    for (int k = 0; k < md->nd; k++) {
        swh_domain *h = md->dom[k];
        const int c = h->cur;
        t.d[k].S = (T *)h->S[c];
        t.d[k].C = (T *)h->C[c];
        t.d[k].UH = (T *)h->UH[c];
        t.d[k].VH = (T *)h->VH[c];
        t.d[k].E = (const T *)h->E;
        t.d[k].nx = h->nx;
        t.d[k].ny = h->ny;
        t.d[k].pitch = h->pitch;
    }
    return t;
}

// The four edge strips (N, S, E, W) of one box of domain h, with the flux-plane
// rows/columns the reference's boundary_flux_integral_tstep reads
// (nested_grid_comms_mod.f90:2336-2386): leapfrog domains read UH/VH with
// lower index 2 and multiply by the edge length, finite-volume domains read their
// pre-multiplied flux_NS/flux_EW with lower index 1.
void box_strips(const swh_domain *h, const swh_nest_box &b, long long &next, std::vector<NestStrip> &out)
{
    const int iL = b.inds[0], iU = b.inds[1], jL = b.inds[2], jU = b.inds[3];
    const bool lf = h->leapfrog();
    const int dns = lf ? 1 : 0;
    const int ns_size = lf ? h->ny : h->ny + 1, ew_size = lf ? h->nx : h->nx + 1;
    const int nsv[4] = {jU + 1, jL, iU + 1, iL};   // 1-based edge index of N, S, E, W
    for (int side = 0; side < 4; side++) {
        NestStrip s{};
        s.kind = side < 2 ? 0 : 1;
        const int ns = nsv[side];
        const int idx = ns - dns;                     // 1-based row/column in the flux array
        const int size = side < 2 ? ns_size : ew_size;
        s.line = (idx >= 1 && idx <= size) ? idx - 1 : -1;
        s.lo = side < 2 ? iL - 1 : jL - 1;
        s.n = b.flux_rows[side];
        s.edge = lf ? (side < 2 ? h->dbe[ns - 1] : h->dle[ns - 1]) : 1.0;
        s.out = (int)next;
        next += (long long)s.n * b.flux_nvar;
        out.push_back(s);
    }
}

// compute_offset_inside_or_out (domain_mod.f90:3486-3608) for a single image.
void offset_inside_or_out(const swh_md *md, int my, int dm_outside, int out_index, int nbr_index, bool is_ew, int &dm,
                          int &dx_ratio, int &var1, int &varN)
{
    const int i1 = is_ew ? 0 : 1;
    auto dx = [&](int c, int d) { return md->all_dx[2 * (d - 1) + c]; };
    const double area_out = dx(0, out_index) * dx(1, out_index);
    const double area_nbr = dx(0, nbr_index) * dx(1, nbr_index);
    if (area_out > 1.25 * area_nbr) dm = dm_outside;
    else if (area_out < 0.8 * area_nbr) dm = 0;
    else dm = dm_outside > 0 ? 0 : dm_outside;
    int cor, notcor;
    if (dm == dm_outside) { cor = out_index; notcor = nbr_index; }
    else { cor = nbr_index; notcor = out_index; }
    const int cor_tsi = md->methods[cor - 1], notcor_tsi = md->methods[notcor - 1];
    if (dx(i1, cor) > 0.8 * dx(i1, my)) dx_ratio = nint(dx(i1, cor) / dx(i1, my));
    else dx_ratio = 0;
    if (fc_unsupported(cor_tsi) || fc_unsupported(notcor_tsi)) { var1 = 0; varN = -1; }
    else if (mass_only(cor_tsi)) { var1 = 0; varN = 0; }
    else { var1 = 0; varN = 2; }
}

// The flux-correction updates of domain k in the reference's loop order
// (domain_mod.f90:2803-3226): N, S, E, W passes, each over the receive boxes; one
// entry per corrected cell.
int build_fixes(swh_md *md, int k, std::vector<NestFix> &fix, std::vector<int> &seg)
{
    const swh_domain *h = md->dom[k];
    const int nx = h->nx, ny = h->ny, my = k + 1;
    const std::vector<int> &pdi = md->pdi[k];
    auto P = [&](int i, int j) { return pdi[(size_t)(j - 1) * nx + (i - 1)]; };   // 1-based
    seg.clear();
    fix.clear();
    for (int dir = 0; dir < 4; dir++) {
        const bool ew = dir >= 2;
        const int dmo = (dir == 0 || dir == 2) ? 1 : -1;
        for (size_t c = 0; c < md->recv[k].size(); c++) {
            const swh_nest_box &b = md->recv[k][c];
            const int n0 = b.inds[0], n1 = b.inds[1], m0 = b.inds[2], m1 = b.inds[3];
            const int nbr = b.neighbour_domain_index;
            const int start = (int)fix.size();
            const bool run = dir == 0 ? m1 < ny : dir == 1 ? m0 > 1 : dir == 2 ? n1 < nx : n0 > 1;
            if (run) {
                const int lo = ew ? m0 : n0, hi = ew ? m1 : n1;
                for (int q = lo; q <= hi; q++) {
                    const int edge = dir == 0 ? m1 : dir == 1 ? m0 : dir == 2 ? n1 : n0;
                    const int out_index = ew ? P(edge + dmo, q) : P(q, edge + dmo);
                    if (out_index == nbr) continue;
                    int dm, dx_ratio, var1, varN;
                    offset_inside_or_out(md, my, dmo, out_index, nbr, ew, dm, dx_ratio, var1, varN);
                    if (dx_ratio <= 0) continue;
                    int p0, p1, sg;
                    if (dmo > 0) {
                        if (dm > 0) { p0 = edge + dm; p1 = edge + dm + dx_ratio - 1; sg = 1; }
                        else { p0 = edge + dm - dx_ratio + 1; p1 = edge + dm; sg = -1; }
                    } else {
                        if (dm < 0) { p0 = edge + dm - dx_ratio + 1; p1 = edge + dm; sg = 1; }
                        else { p0 = edge + dm; p1 = edge + dm + dx_ratio - 1; sg = -1; }
                    }
                    if (out_index != my && dm != 0) sg = -sg;
                    const int lim = ew ? nx : ny;
                    if (p0 < 1 || p1 > lim)
                        return md_fail(SWH_EINVAL, "flux correction would leave the domain (box " +
                                                       std::to_string(c + 1) + " of domain " + std::to_string(my) + ")");
                    if ((ew ? P(p0, q) : P(q, p0)) != my) sg = -sg;
                    double area;
                    if (!ew) {
                        area = 0.0;
                        for (int p = p0; p <= p1; p++) area = area + h->area[p - 1];
                    } else {
                        area = h->area[q - 1] * (double)dx_ratio;
                    }
                    NestFix f{};
                    f.is_ew = ew ? 1 : 0;
                    f.line = q - 1;
                    f.p0 = p0 - 1;
                    f.p1 = p1 - 1;
                    f.var1 = var1;
                    f.varN = varN;
                    // N and E subtract sg * err / area, S and W add it
                    f.sgn = (dir == 0 || dir == 2) ? -(double)sg : (double)sg;
                    f.area = area;
                    const NestStrip &st = md->h_strips[k][md->recv_strip0[k][c] + dir];
                    f.err = st.out + (q - lo);
                    f.stride = st.n;
                    // one entry per cell, so that a segment's cells go to separate threads
                    for (int p = p0 - 1; p <= p1 - 1; p++) {
                        f.p0 = p;
                        f.p1 = p;
                        fix.push_back(f);
                    }
                }
            }
            if ((int)fix.size() > start) seg.push_back(start);
        }
    }
    seg.push_back((int)fix.size());
    // consecutive segments with no cell in common are merged: a cell's update reads only
    // that cell (and the read-only error arena), so the merged group gives the same
    // values as the segments one after the other, with fewer barriers
    std::vector<int> merged;
    std::unordered_set<long long> cells;
    for (size_t g = 0; g + 1 < seg.size(); g++) {
        bool clash = merged.empty();
        for (int q = seg[g]; q < seg[g + 1] && !clash; q++) {
            const NestFix &f = fix[q];
            const long long key = f.is_ew ? (long long)f.line * nx + f.p0 : (long long)f.p0 * nx + f.line;
            clash = cells.count(key) != 0;
        }
        if (clash) {
            merged.push_back(seg[g]);
            cells.clear();
        }
        for (int q = seg[g]; q < seg[g + 1]; q++) {
            const NestFix &f = fix[q];
            cells.insert(f.is_ew ? (long long)f.line * nx + f.p0 : (long long)f.p0 * nx + f.line);
        }
    }
    merged.push_back((int)fix.size());
    seg.swap(merged);
    return SWH_OK;
}

template <typename T>
int exchange(swh_md *md, bool flux_correct)
{
    const NestDoms<T> doms = dom_ptrs<T>(md);
    hipStream_t s = md->stream;
    SWH_HIP(launch_nest_pack<T>(doms, md->d_packs, md->n_packs, md->pack_blocks, md->d_int_tab, md->arena, s));
    SWH_HIP(launch_nest_unpack<T>(doms, md->d_ups, md->n_ups, md->up_blocks, md->d_int_tab, md->d_err_tab, md->arena,
                                  s));
    SWH_HIP(launch_nest_stag_fix<T>(doms, md->d_fixups, md->n_fixups, md->fixup_blocks, s));
    if (flux_correct) {
        for (size_t L = 0; L < md->fix_level.size(); L++)
            SWH_HIP(launch_nest_fix<T>(doms, md->d_fixdoms, md->fix_level[L], md->fix_level_n[L], s));
        SWH_HIP(launch_nest_reint<T>(doms, md->d_reint, md->n_reint, md->reint_max_len, md->reint_blocks, md->d_copy,
                                     s));
    }
    return SWH_OK;
}

int exchange_any(swh_md *md, bool flux_correct)
{
    return md->prec == SWH_PREC_FP64 ? exchange<double>(md, flux_correct) : exchange<float>(md, flux_correct);
}

// Domain streams -> multidomain stream.
int join_domains(swh_md *md)
{
    for (int k = 0; k < md->nd; k++) {
        SWH_HIP(hipEventRecord(md->ev_dom[k], md->dom[k]->stream));
        SWH_HIP(hipStreamWaitEvent(md->stream, md->ev_dom[k], 0));
    }
    return SWH_OK;
}

// Multidomain stream -> domain streams.
int fork_domains(swh_md *md)
{
    SWH_HIP(hipEventRecord(md->ev_x, md->stream));
    for (int k = 0; k < md->nd; k++) SWH_HIP(hipStreamWaitEvent(md->dom[k]->stream, md->ev_x, 0));
    return SWH_OK;
}

__global__ void md_set_time_kernel(DevScalars *sc, double t) { sc->time = t; }
// under graph replay: the end time from the device's multidomain time (same operation
// as the host's md->time + dt)
__global__ void md_set_time_dev_kernel(DevScalars *sc, const double *md_time, double dt) { sc->time = *md_time + dt; }
__global__ void md_advance_time_kernel(double *md_time, double dt) { *md_time = *md_time + dt; }

// One global step (multidomain_mod.f90:1093-1212: each domain's sub-steps, then the
// halo exchange with flux correction). While capturing, the domain streams are forked
// from the multidomain stream at the start (so that their work joins the capture) and
// the times come from the device.
int md_step(swh_md *md, double dt)
{
    const double tend = md->time + dt;
    int rc;
    if (md->capturing && (rc = fork_domains(md))) return rc;
    for (int k = 0; k < md->nd; k++) {
        swh_domain *h = md->dom[k];
        SWH_HIP(launch_nest_zero<double>(md->d_int[k], md->n_int[k], h->stream));
        const int nt = md->trf[k];
        const double dt_local = dt / (1.0 * nt);
        for (int i = 0; i < nt; i++) {
            rc = swh::domain_evolve(h, dt_local);
            if (rc) return rc;
        }
        // numerically identical time for all domains (multidomain_mod.f90:1174-1175)
        h->time = tend;
        if (md->capturing)
            hipLaunchKernelGGL(md_set_time_dev_kernel, dim3(1), dim3(1), 0, h->stream, h->sc, md->d_time, dt);
        else
            hipLaunchKernelGGL(md_set_time_kernel, dim3(1), dim3(1), 0, h->stream, h->sc, tend);
        SWH_HIP(hipGetLastError());
    }
    rc = join_domains(md);
    if (rc) return rc;
    rc = exchange_any(md, true);
    if (rc) return rc;
    if (md->capturing) {
        hipLaunchKernelGGL(md_advance_time_kernel, dim3(1), dim3(1), 0, md->stream, md->d_time, dt);
        SWH_HIP(hipGetLastError());
    } else if ((rc = fork_domains(md))) {
        return rc;
    }
    md->time = tend;
    return SWH_OK;
}

void graph_release(swh_md *md)
{
    for (auto &g : md->gexec)
        if (g) {
            (void)hipGraphExecDestroy(g);
            g = nullptr;
        }
    md->n_graphs = 0;
    md->cap_cur.clear();
}

// The captured graph that continues from the domains' current buffers: 0 (as at the
// capture), 1 (after an odd number of global steps, when some domain takes an odd
// number of sub-steps), -1 (neither: the buffers were changed otherwise).
int graph_for_state(const swh_md *md)
{
    for (int g = 0; g < md->n_graphs; g++) {
        bool ok = true;
        for (int k = 0; k < md->nd && ok; k++) {
            const int want = (g == 1 && md->trf[k] % 2 == 1) ? 1 - md->cap_cur[k] : md->cap_cur[k];
            ok = md->dom[k]->cur == want;
        }
        if (ok) return g;
    }
    return -1;
}

// Host bookkeeping of one replayed global step (md_step's, without launches).
void md_host_advance(swh_md *md, double dt)
{
    const double tend = md->time + dt;
    for (int k = 0; k < md->nd; k++) {
        swh_domain *h = md->dom[k];
        const double dt_local = dt / (1.0 * md->trf[k]);
        for (int i = 0; i < md->trf[k]; i++) swh::domain_host_advance(h, dt_local);
        h->time = tend;
    }
    md->time = tend;
}

int graph_capture(swh_md *md, double dt)
{
    graph_release(md);
    for (int k = 0; k < md->nd; k++) {
        const std::string why = swh::domain_graph_unsupported(md->dom[k]);
        if (!why.empty())
            return md_fail(SWH_ESTATE, "swh_md_graph_capture: domain " + std::to_string(k + 1) + ": " + why);
    }
    // per-dt set-up outside the capture
    for (int k = 0; k < md->nd; k++) {
        const int rc = swh::domain_prepare_dt(md->dom[k], dt / (1.0 * md->trf[k]));
        if (rc) return rc;
    }
    for (auto h : md->dom) SWH_HIP(hipStreamSynchronize(h->stream));
    SWH_HIP(hipStreamSynchronize(md->stream));
    if (!md->d_time) SWH_HIP(hipMalloc((void **)&md->d_time, sizeof(double)));
    SWH_HIP(hipMemcpy(md->d_time, &md->time, sizeof(double), hipMemcpyHostToDevice));
    // a second graph when the buffers alternate (an odd number of sub-steps somewhere)
    bool odd = false;
    for (int k = 0; k < md->nd; k++) odd |= (md->trf[k] % 2) == 1;
    const int ng = odd ? 2 : 1;
    // host state before the capture: the captured steps are not taken
    struct Saved { double time, esd, dtl; long long nsteps; int cur; };
    std::vector<Saved> saved;
    for (auto h : md->dom) saved.push_back({h->time, h->evolve_step_dt, h->dt_last_update, h->nsteps, h->cur});
    const double md_time = md->time;
    int rc = SWH_OK;
    for (int g = 0; g < ng && rc == SWH_OK; g++) {
        hipGraph_t graph = nullptr;
        if (hipStreamBeginCapture(md->stream, hipStreamCaptureModeThreadLocal) != hipSuccess) {
            rc = md_fail(SWH_EHIP, "swh_md_graph_capture: hipStreamBeginCapture failed");
            break;
        }
        md->capturing = true;
        for (auto h : md->dom) h->graph_time = 1;
        rc = md_step(md, dt);
        for (auto h : md->dom) h->graph_time = 0;
        md->capturing = false;
        const hipError_t e = hipStreamEndCapture(md->stream, &graph);
        if (rc == SWH_OK && e != hipSuccess) rc = md_fail(SWH_EHIP, "swh_md_graph_capture: hipStreamEndCapture failed");
        if (rc == SWH_OK && hipGraphInstantiate(&md->gexec[g], graph, nullptr, nullptr, 0) != hipSuccess)
            rc = md_fail(SWH_EHIP, "swh_md_graph_capture: hipGraphInstantiate failed");
        if (graph) (void)hipGraphDestroy(graph);
    }
    for (int k = 0; k < md->nd; k++) {
        swh_domain *h = md->dom[k];
        h->time = saved[k].time; h->evolve_step_dt = saved[k].esd; h->dt_last_update = saved[k].dtl;
        h->nsteps = saved[k].nsteps; h->cur = saved[k].cur;
    }
    md->time = md_time;
    if (rc) {
        graph_release(md);
        return rc;
    }
    md->n_graphs = ng;
    md->cap_cur.clear();
    for (auto h : md->dom) md->cap_cur.push_back(h->cur);
    md->graph_dt = dt;
    return SWH_OK;
}

void md_free(swh_md *md)
{
    auto F = [](void *p) {
        if (p) (void)hipFree(p);
    };
    if (md->stream) (void)hipStreamSynchronize(md->stream);
    F(md->arena);
    for (auto p : md->d_int) F(p);
    for (auto p : md->d_err) F(p);
    for (auto p : md->d_strips) F(p);
    for (auto p : md->d_pnp) F(p);
    for (auto p : md->owned) F(p);
    F(md->d_int_tab);
    F(md->d_err_tab);
    F(md->d_packs);
    F(md->d_ups);
    F(md->d_fixups);
    F(md->d_fixdoms);
    F(md->d_reint);
    F(md->d_copy);
    F(md->d_vol);
    F(md->d_time);
    graph_release(md);
    for (auto e : md->ev_dom)
        if (e) (void)hipEventDestroy(e);
    if (md->ev_x) (void)hipEventDestroy(md->ev_x);
    if (md->stream) (void)hipStreamDestroy(md->stream);
    for (auto h : md->dom) {
        h->d_nest_strips = nullptr;
        h->n_nest_strips = 0;
        h->nest_strip_max = 0;
        h->d_nest_int = nullptr;
    }
}

int finalize(swh_md *md)
{
    const int nd = md->nd;
    // ---- pairing and buffers: one buffer per receive box ----
    md->recv_base.assign(nd + 1, 0);
    for (int k = 0; k < nd; k++) md->recv_base[k + 1] = md->recv_base[k] + (int)md->recv[k].size();
    long long total = 0;
    md->recv_buf_off.assign(md->recv_base[nd], 0);
    for (int k = 0; k < nd; k++)
        for (size_t c = 0; c < md->recv[k].size(); c++) {
            md->recv_buf_off[md->recv_base[k] + c] = total;
            total += md->recv[k][c].buffer_size;
        }
    SWH_HIP(hipMalloc((void **)&md->arena, std::max<long long>(total, 1) * sizeof(double)));
    SWH_HIP(hipMemset(md->arena, 0, std::max<long long>(total, 1) * sizeof(double)));

    // ---- strips and integral / error arenas ----
    md->d_int.assign(nd, nullptr);
    md->d_err.assign(nd, nullptr);
    md->n_int.assign(nd, 0);
    md->d_strips.assign(nd, nullptr);
    md->n_strips.assign(nd, 0);
    md->h_strips.assign(nd, {});
    md->send_strip0.assign(nd, {});
    md->recv_strip0.assign(nd, {});
    for (int k = 0; k < nd; k++) {
        swh_domain *h = md->dom[k];
        long long next = 0;
        for (auto &b : md->send[k]) {
            md->send_strip0[k].push_back((int)md->h_strips[k].size());
            box_strips(h, b, next, md->h_strips[k]);
        }
        for (auto &b : md->recv[k]) {
            md->recv_strip0[k].push_back((int)md->h_strips[k].size());
            box_strips(h, b, next, md->h_strips[k]);
        }
        for (auto &st : md->h_strips[k]) {
            const int len = st.kind == 0 ? h->nx : h->ny;
            if (st.lo < 0 || st.lo + st.n > len) return md_fail(SWH_EINVAL, "nesting box edge outside its domain");
        }
        md->n_int[k] = next;
        const size_t bytes = std::max<long long>(next, 1) * sizeof(double);
        SWH_HIP(hipMalloc((void **)&md->d_int[k], bytes));
        SWH_HIP(hipMemset(md->d_int[k], 0, bytes));
        SWH_HIP(hipMalloc((void **)&md->d_err[k], bytes));
        SWH_HIP(hipMemset(md->d_err[k], 0, bytes));
        int rc = upload_vec(&md->d_strips[k], md->h_strips[k]);
        if (rc) return rc;
        md->n_strips[k] = (int)md->h_strips[k].size();
    }
    {
        int rc = upload_vec(&md->d_int_tab, md->d_int);
        if (rc) return rc;
        rc = upload_vec(&md->d_err_tab, md->d_err);
        if (rc) return rc;
    }

    // ---- packs ----
    std::vector<NestPack> packs;
    int blocks = 0;
    for (int k = 0; k < nd; k++)
        for (size_t c = 0; c < md->send[k].size(); c++) {
            const swh_nest_box &b = md->send[k][c];
            const int rk = b.neighbour_domain_index - 1, rc_ = b.neighbour_comms_index - 1;
            if (rk < 0 || rk >= nd || rc_ < 0 || rc_ >= (int)md->recv[rk].size())
                return md_fail(SWH_EINVAL, "send box names a receive box that was not added");
            const swh_nest_box &rb = md->recv[rk][rc_];
            if (rb.buffer_size != b.buffer_size || rb.neighbour_domain_index != k + 1 ||
                rb.neighbour_comms_index != (int)c + 1)
                return md_fail(SWH_EINVAL, "send and receive boxes do not match");
            NestPack p{};
            p.dom = k;
            p.iL = b.inds[0]; p.iU = b.inds[1]; p.jL = b.inds[2]; p.jU = b.inds[3]; p.kL = b.inds[4]; p.kU = b.inds[5];
            p.cr[0] = b.cell_ratios[0]; p.cr[1] = b.cell_ratios[1];
            const bool eq = b.equal_cell_ratios != 0;
            p.my_stag = b.my_domain_staggered_grid;
            p.nbr_stag = b.neighbour_domain_staggered_grid;
            p.wetdry = b.use_wetdry_limiting;
            if (eq) {
                p.inv_cr[0] = p.inv_cr[1] = 1.0;
                p.inv_r[0] = p.inv_r[1] = p.r_ip[0] = p.r_ip[1] = 1;
                p.iR = p.iU - p.iL + 1;
                p.jR = p.jU - p.jL + 1;
            } else {
                p.inv_cr[0] = 1.0 / p.cr[0];
                p.inv_cr[1] = 1.0 / p.cr[1];
                p.inv_r[0] = nint(p.inv_cr[0]); p.inv_r[1] = nint(p.inv_cr[1]);
                p.r_ip[0] = nint(p.cr[0]); p.r_ip[1] = nint(p.cr[1]);
                p.iR = nint((p.iU - p.iL + 1) * p.cr[0]);
                p.jR = nint((p.jU - p.jL + 1) * p.cr[1]);
            }
            p.kR = p.kU - p.kL + 1;
            p.mode = (eq && p.my_stag == p.nbr_stag) ? 0 : ((b.my_domain_is_finer || eq) ? 1 : 2);
            const int nI = p.iU - p.iL + 1, nJ = p.jU - p.jL + 1;
            if (p.mode == 1 && (nI % p.inv_r[0] || nJ % p.inv_r[1]))
                return md_fail(SWH_EINVAL, "fine-to-coarse box is not a whole number of coarse cells");
            p.n_data = p.mode == 2 ? nI * nJ * p.kR * (1 + kNestGradients) : p.iR * p.jR * p.kR;
            p.nvar = b.flux_nvar;
            const int s0 = md->send_strip0[k][c];
            p.n_flux = 0;
            for (int side = 0; side < 4; side++) {
                const NestStrip &st = md->h_strips[k][s0 + side];
                p.strip_out[side] = st.out;
                p.strip_n[side] = st.n;
                int per;
                if (p.mode == 0) per = st.n;
                else if (p.mode == 1) {
                    const int stride = side < 2 ? p.inv_r[0] : p.inv_r[1];
                    if (st.n % stride) return md_fail(SWH_EINVAL, "flux strip is not a whole number of coarse cells");
                    per = st.n / stride;
                } else per = st.n * (side < 2 ? p.r_ip[0] : p.r_ip[1]);
                p.n_flux += per * p.nvar;
            }
            p.flux_start = b.flux_start - 1;
            p.time_index = b.time_index - 1;
            if (p.flux_start != p.n_data || p.flux_start + p.n_flux != p.time_index ||
                p.time_index + 1 != b.buffer_size)
                return md_fail(SWH_EINVAL, "send box buffer layout does not match its geometry");
            p.buf = md->recv_buf_off[md->recv_base[rk] + rc_];
            const int items = (p.mode == 2 ? nI * nJ : p.n_data) + p.n_flux + 1;
            p.block0 = blocks;
            blocks += (items + 255) / 256;
            packs.push_back(p);
        }
    md->n_packs = (int)packs.size();
    md->pack_blocks = blocks;
    int rc = upload_vec(&md->d_packs, packs);
    if (rc) return rc;

    // ---- unpacks and staggered fix-ups ----
    std::vector<NestUnpack> ups, fixups;
    blocks = 0;
    int fblocks = 0;
    for (int k = 0; k < nd; k++) {
        swh_domain *h = md->dom[k];
        for (size_t c = 0; c < md->recv[k].size(); c++) {
            const swh_nest_box &b = md->recv[k][c];
            NestUnpack u{};
            u.dom = k;
            u.fine = b.my_domain_is_finer;
            u.iL = b.inds[0]; u.iU = b.inds[1]; u.jL = b.inds[2]; u.jU = b.inds[3]; u.kL = b.inds[4]; u.kU = b.inds[5];
            u.cr[0] = b.cell_ratios[0];
            u.cr[1] = b.cell_ratios[1];
            const int nI = u.iU - u.iL + 1, nJ = u.jU - u.jL + 1, nK = u.kU - u.kL + 1;
            if (u.iL < 1 || u.iU > h->nx || u.jL < 1 || u.jU > h->ny || u.kL < 1 || u.kU > 4)
                return md_fail(SWH_EINVAL, "receive box outside its domain");
            if (u.fine) {
                u.inv_r[0] = nint(1.0 / u.cr[0]);
                u.inv_r[1] = nint(1.0 / u.cr[1]);
                u.mid[0] = 0.5 * (u.inv_r[0] + 1);
                u.mid[1] = 0.5 * (u.inv_r[1] + 1);
                if (nI % u.inv_r[0] || nJ % u.inv_r[1])
                    return md_fail(SWH_EINVAL, "fine receive box is not a whole number of coarse cells");
                u.niC = nI / u.inv_r[0];
                u.njC = nJ / u.inv_r[1];
                u.nC = u.niC * u.njC;
                u.n_data = nint((double)(nI * nJ * nK) * (u.cr[0] * u.cr[1]) * (1 + kNestGradients));
            } else {
                u.inv_r[0] = u.inv_r[1] = 1;
                u.n_data = nI * nJ * nK;
            }
            u.nvar = b.flux_nvar;
            const int s0 = md->recv_strip0[k][c];
            int nflux = 0;
            for (int side = 0; side < 4; side++) {
                const NestStrip &st = md->h_strips[k][s0 + side];
                u.strip_out[side] = st.out;
                u.strip_n[side] = st.n;
                nflux += st.n * u.nvar;
            }
            if (u.n_data + 1 != b.flux_start || u.n_data + nflux + 1 != b.time_index)
                return md_fail(SWH_EINVAL, "receive box buffer layout does not match its geometry");
            u.err_out = u.strip_out[0];
            u.buf = md->recv_buf_off[md->recv_base[k] + c];
            u.block0 = blocks;
            blocks += (nI * nJ * nK + nflux + 255) / 256;
            // domain_mod.f90:3741-3832
            if (h->method == SWH_METHOD_LEAPFROG_NONLINEAR) {
                u.stag_fix = 3;
            } else if (h->leapfrog()) {
                u.stag_fix = h->truely_linear ? 1 : 2;
                u.elev_lim = h->msl - kMinDepth;
            }
            ups.push_back(u);
            if (u.stag_fix) {
                NestUnpack f = u;
                f.block0 = fblocks;
                fblocks += (nI * nJ + 255) / 256;
                fixups.push_back(f);
            }
        }
    }
    md->n_ups = (int)ups.size();
    md->up_blocks = blocks;
    rc = upload_vec(&md->d_ups, ups);
    if (rc) return rc;
    // the fix-up launch takes its own descriptor list (only staggered receivers)
    rc = upload_vec(&md->d_fixups, fixups);
    if (rc) return rc;
    md->n_fixups = (int)fixups.size();
    md->fixup_blocks = fblocks;

    // ---- priority masks, flux correction ----
    md->d_pnp.assign(nd, nullptr);
    std::vector<NestFixDom> fds;
    std::vector<std::vector<int>> fd_seg;
    for (int k = 0; k < nd; k++) {
        swh_domain *h = md->dom[k];
        if (md->pdi[k].empty()) return md_fail(SWH_ESTATE, "swh_md_set_priority was not called for every domain");
        rc = upload_vec(&md->d_pnp[k], md->pnp[k]);
        if (rc) return rc;
        std::vector<NestFix> fix;
        std::vector<int> seg;
        rc = build_fixes(md, k, fix, seg);
        if (rc) return rc;
        if (fix.empty()) continue;
        NestFix *dfix = nullptr;
        rc = upload_vec(&dfix, fix);
        if (rc) return rc;
        md->owned.push_back(dfix);
        fd_seg.push_back(seg);
        NestFixDom fd{};
        fd.fix = dfix;
        fd.dom = k;
        fd.err = md->d_err[k];
        fd.pnp = md->d_pnp[k];
        fd.dry_all = md->dry_all[k];
        fds.push_back(fd);
        (void)h;
    }
    md->n_fixdoms = (int)fds.size();
    rc = upload_vec(&md->d_fixdoms, fds);
    if (rc) return rc;
    // group level L: every domain's L-th group
    md->fix_level.clear();
    md->fix_level_n.clear();
    for (size_t L = 0;; L++) {
        std::vector<int2> cells;
        bool any = false;
        for (size_t f = 0; f < fd_seg.size(); f++) {
            const std::vector<int> &sg = fd_seg[f];
            if (L + 1 >= sg.size()) continue;
            any = true;
            for (int q = sg[L]; q < sg[L + 1]; q++) cells.push_back(make_int2((int)f, q));
        }
        if (!any) break;
        int2 *dc = nullptr;
        rc = upload_vec(&dc, cells);
        if (rc) return rc;
        md->owned.push_back(dc);
        md->fix_level.push_back(dc);
        md->fix_level_n.push_back((int)cells.size());
    }

    // ---- re-interpolation of fine receive boxes ----
    std::vector<NestReint> rs;
    blocks = 0;
    long long copy_n = 0;
    for (int k = 0; k < nd; k++) {
        swh_domain *h = md->dom[k];
        for (size_t c = 0; c < md->recv[k].size(); c++) {
            const swh_nest_box &b = md->recv[k][c];
            if (!b.my_domain_is_finer) continue;
            NestReint r{};
            r.dom = k;
            r.n0 = b.inds[0]; r.n1 = b.inds[1]; r.m0 = b.inds[2]; r.m1 = b.inds[3];
            r.inv_r[0] = nint(1.0 / b.cell_ratios[0]);
            r.inv_r[1] = nint(1.0 / b.cell_ratios[1]);
            r.ind[0] = std::min(r.m1 + 1 + r.inv_r[1] / 2, h->ny);
            r.ind[1] = std::max(r.m0 - 1 - r.inv_r[1] / 2, 1);
            r.ind[2] = std::min(r.n1 + 1 + r.inv_r[0] / 2, h->nx);
            r.ind[3] = std::max(r.n0 - 1 - r.inv_r[0] / 2, 1);
            r.len_ns = r.n1 - r.n0 + 1;
            r.len_ew = r.m1 - r.m0 + 1;
            const int L = std::max(r.len_ns, r.len_ew);
            r.copy = (int)copy_n;
            copy_n += 16LL * L;
            const int lastx = r.n1 - r.inv_r[0] + 1, lasty = r.m1 - r.inv_r[1] + 1;
            r.ncx = lastx >= r.n0 ? (lastx - r.n0) / r.inv_r[0] + 1 : 0;
            r.ncy = lasty >= r.m0 ? (lasty - r.m0) / r.inv_r[1] + 1 : 0;
            r.block0 = blocks;
            blocks += (r.ncx * r.ncy * r.inv_r[0] * r.inv_r[1] + 255) / 256;
            rs.push_back(r);
        }
    }
    md->n_reint = (int)rs.size();
    md->reint_blocks = blocks;
    md->reint_max_len = 0;
    for (const auto &r : rs) md->reint_max_len = std::max(md->reint_max_len, std::max(r.len_ns, r.len_ew));
    rc = upload_vec(&md->d_reint, rs);
    if (rc) return rc;
    SWH_HIP(hipMalloc((void **)&md->d_copy, std::max<long long>(copy_n, 1) * sizeof(double)));
    SWH_HIP(hipMalloc((void **)&md->d_vol, std::max(nd, 1) * sizeof(double)));

    // ---- hooks into the domains' steps ----
    for (int k = 0; k < nd; k++) {
        swh_domain *h = md->dom[k];
        h->d_nest_strips = md->d_strips[k];
        h->n_nest_strips = md->n_strips[k];
        h->nest_strip_max = 0;
        for (const auto &st : md->h_strips[k]) h->nest_strip_max = std::max(h->nest_strip_max, st.n);
        h->d_nest_int = md->d_int[k];
    }
    md->finalized = true;
    return SWH_OK;
}

}  // namespace

extern "C" {

int swh_md_create(int32_t n_domains, swh_domain *const *domains, const double *all_dx, const int32_t *all_methods,
                  const int32_t *timestepping_refinement_factor, swh_md **out)
{
    if (!out || !domains || !all_dx || !all_methods || !timestepping_refinement_factor)
        return md_fail(SWH_EINVAL, "swh_md_create: null argument");
    *out = nullptr;
    if (n_domains < 1 || n_domains > kMaxNestDomains)
        return md_fail(SWH_EINVAL, "swh_md_create: 1 to " + std::to_string(kMaxNestDomains) + " domains");
    swh_md *md = new swh_md();
    md->nd = n_domains;
    for (int k = 0; k < n_domains; k++) {
        if (!domains[k]) {
            delete md;
            return md_fail(SWH_EINVAL, "swh_md_create: null domain");
        }
        if (timestepping_refinement_factor[k] < 1) {
            delete md;
            return md_fail(SWH_EINVAL, "swh_md_create: timestepping_refinement_factor must be >= 1");
        }
        md->dom.push_back(domains[k]);
        md->trf.push_back(timestepping_refinement_factor[k]);
        md->methods.push_back(all_methods[k]);
        md->all_dx.push_back(all_dx[2 * k]);
        md->all_dx.push_back(all_dx[2 * k + 1]);
    }
    md->prec = domains[0]->prec;
    for (int k = 1; k < n_domains; k++)
        if (domains[k]->prec != md->prec) {
            delete md;
            return md_fail(SWH_EINVAL, "swh_md_create: all domains must use one precision");
        }
    md->dry_all.assign(n_domains, 1);
    md->pdi.assign(n_domains, {});
    md->pnp.assign(n_domains, {});
    md->send.assign(n_domains, {});
    md->recv.assign(n_domains, {});
    md->time = domains[0]->time;
    if (hipStreamCreate(&md->stream) != hipSuccess || hipEventCreateWithFlags(&md->ev_x, hipEventDisableTiming)) {
        md_free(md);
        delete md;
        return md_fail(SWH_EHIP, "swh_md_create: stream/event creation failed");
    }
    md->ev_dom.assign(n_domains, nullptr);
    for (int k = 0; k < n_domains; k++)
        if (hipEventCreateWithFlags(&md->ev_dom[k], hipEventDisableTiming) != hipSuccess) {
            md_free(md);
            delete md;
            return md_fail(SWH_EHIP, "swh_md_create: event creation failed");
        }
    *out = md;
    return SWH_OK;
}

int swh_md_destroy(swh_md *md)
{
    if (!md) return SWH_OK;
    md_free(md);
    delete md;
    return SWH_OK;
}

int swh_md_set_priority(swh_md *md, int32_t k, const int32_t *priority_domain_index,
                        const int32_t *is_priority_domain_not_periodic)
{
    if (!md || !priority_domain_index || !is_priority_domain_not_periodic || k < 1 || k > md->nd)
        return md_fail(SWH_EINVAL, "swh_md_set_priority: bad argument");
    if (md->finalized) return md_fail(SWH_ESTATE, "swh_md_set_priority after swh_md_finalize");
    const size_t n = (size_t)md->dom[k - 1]->nx * md->dom[k - 1]->ny;
    md->pdi[k - 1].assign(priority_domain_index, priority_domain_index + n);
    md->pnp[k - 1].assign(is_priority_domain_not_periodic, is_priority_domain_not_periodic + n);
    for (size_t q = 0; q < n; q++)
        if (md->pdi[k - 1][q] < 1 || md->pdi[k - 1][q] > md->nd)
            return md_fail(SWH_EINVAL, "swh_md_set_priority: priority_domain_index out of range");
    return SWH_OK;
}

int swh_md_set_flux_correct_dry_cells(swh_md *md, int32_t k, int32_t on)
{
    if (!md || k < 1 || k > md->nd) return md_fail(SWH_EINVAL, "swh_md_set_flux_correct_dry_cells: bad argument");
    if (md->finalized) return md_fail(SWH_ESTATE, "swh_md_set_flux_correct_dry_cells after swh_md_finalize");
    md->dry_all[k - 1] = on ? 1 : 0;
    return SWH_OK;
}

int swh_md_add_box(swh_md *md, int32_t k, int32_t is_recv, const swh_nest_box *box)
{
    if (!md || !box || k < 1 || k > md->nd) return md_fail(SWH_EINVAL, "swh_md_add_box: bad argument");
    if (md->finalized) return md_fail(SWH_ESTATE, "swh_md_add_box after swh_md_finalize");
    const swh_domain *h = md->dom[k - 1];
    const int32_t *d = box->inds;
    if (d[0] < 1 || d[1] > h->nx || d[0] > d[1] || d[2] < 1 || d[3] > h->ny || d[2] > d[3] || d[4] < 1 || d[5] > 4 ||
        d[4] > d[5])
        return md_fail(SWH_EINVAL, "swh_md_add_box: box outside the domain");
    if (box->cell_ratios[0] <= 0.0 || box->cell_ratios[1] <= 0.0 || box->flux_nvar < 1 || box->flux_nvar > 4 ||
        box->buffer_size < 1)
        return md_fail(SWH_EINVAL, "swh_md_add_box: bad box description");
    if (box->flux_rows[0] != d[1] - d[0] + 1 || box->flux_rows[1] != d[1] - d[0] + 1 ||
        box->flux_rows[2] != d[3] - d[2] + 1 || box->flux_rows[3] != d[3] - d[2] + 1)
        return md_fail(SWH_EINVAL, "swh_md_add_box: flux strip lengths differ from the box");
    (is_recv ? md->recv : md->send)[k - 1].push_back(*box);
    return SWH_OK;
}

int swh_md_finalize(swh_md *md)
{
    if (!md) return md_fail(SWH_EINVAL, "swh_md_finalize: null handle");
    if (md->finalized) return md_fail(SWH_ESTATE, "swh_md_finalize called twice");
    for (auto h : md->dom) SWH_HIP(hipStreamSynchronize(h->stream));
    return finalize(md);
}

int swh_md_make_initial_conditions_consistent(swh_md *md)
{
    if (!md || !md->finalized) return md_fail(SWH_ESTATE, "swh_md_make_initial_conditions_consistent: not finalized");
    int rc = join_domains(md);
    if (rc) return rc;
    rc = exchange_any(md, false);
    if (rc) return rc;
    SWH_HIP(hipStreamSynchronize(md->stream));
    for (auto h : md->dom) {
        rc = swh::domain_refresh_elevation(h);
        if (rc) return rc;
    }
    return fork_domains(md);
}

int swh_md_evolve_one_step(swh_md *md, double dt)
{
    if (!md || !md->finalized) return md_fail(SWH_ESTATE, "swh_md_evolve_one_step: not finalized");
    if (!(dt > 0.0)) return md_fail(SWH_EINVAL, "swh_md_evolve_one_step: dt must be positive");
    return md_step(md, dt);
}

int swh_md_graph_capture(swh_md *md, double dt)
{
    if (!md || !md->finalized) return md_fail(SWH_ESTATE, "swh_md_graph_capture: not finalized");
    if (!(dt > 0.0)) return md_fail(SWH_EINVAL, "swh_md_graph_capture: dt must be positive");
    return graph_capture(md, dt);
}

int swh_md_graph_launch(swh_md *md, int32_t nsteps)
{
    if (!md || !md->finalized) return md_fail(SWH_ESTATE, "swh_md_graph_launch: not finalized");
    if (md->n_graphs == 0) return md_fail(SWH_ESTATE, "swh_md_graph_launch: no captured step (swh_md_graph_capture)");
    if (nsteps < 0) return md_fail(SWH_EINVAL, "swh_md_graph_launch: nsteps must be >= 0");
    if (nsteps == 0) return SWH_OK;
    // the graphs read the device time: the host's, as eager steps left it
    int rc = join_domains(md);
    if (rc) return rc;
    SWH_HIP(hipMemcpyAsync(md->d_time, &md->time, sizeof(double), hipMemcpyHostToDevice, md->stream));
    SWH_HIP(hipStreamSynchronize(md->stream));
    for (int32_t n = 0; n < nsteps; n++) {
        const int g = graph_for_state(md);
        if (g < 0) {
            (void)fork_domains(md);
            return md_fail(SWH_ESTATE, "swh_md_graph_launch: the domains' buffers no longer match the capture");
        }
        SWH_HIP(hipGraphLaunch(md->gexec[g], md->stream));
        md_host_advance(md, md->graph_dt);
    }
    return fork_domains(md);
}

int swh_md_graph_destroy(swh_md *md)
{
    if (!md) return md_fail(SWH_EINVAL, "swh_md_graph_destroy: null argument");
    graph_release(md);
    return SWH_OK;
}

int swh_md_evolve_steps(swh_md *md, double dt, int32_t nsteps)
{
    for (int32_t n = 0; n < nsteps; n++) {
        const int rc = swh_md_evolve_one_step(md, dt);
        if (rc) return rc;
    }
    return SWH_OK;
}

int swh_md_volumes(swh_md *md, double *out)
{
    if (!md || !out || !md->finalized) return md_fail(SWH_EINVAL, "swh_md_volumes: bad argument");
    int rc = join_domains(md);
    if (rc) return rc;
    for (int k = 0; k < md->nd; k++) {
        swh_domain *h = md->dom[k];
        const int c = h->cur;
        if (md->prec == SWH_PREC_FP64)
            SWH_HIP(launch_volume<double>((const double *)h->S[c], nullptr, (const double *)h->E, h->nx, h->ny,
                                          h->pitch, h->n_ext, h->d_area, h->d_work, md->d_vol + k, md->stream,
                                          md->d_pnp[k]));
        else
            SWH_HIP(launch_volume<float>((const float *)h->S[c], (const float *)h->C[c], (const float *)h->E, h->nx,
                                         h->ny, h->pitch, h->n_ext, h->d_area, h->d_work, md->d_vol + k, md->stream,
                                         md->d_pnp[k]));
    }
    SWH_HIP(hipStreamSynchronize(md->stream));
    SWH_HIP(hipMemcpy(out, md->d_vol, md->nd * sizeof(double), hipMemcpyDeviceToHost));
    return fork_domains(md);
}

int swh_md_statistics(swh_md *md, swh_domain_stats *per_domain, swh_domain_stats *global, int32_t *energy_is_finite)
{
    if (!md || !md->finalized) return md_fail(SWH_EINVAL, "swh_md_statistics: bad argument");
    int rc = join_domains(md);
    if (rc) return rc;
    for (int k = 0; k < md->nd; k++) {
        swh_domain *h = md->dom[k];
        if ((rc = domain_stats_buffer(h))) return rc;
        if ((rc = domain_statistics_launch(h, md->d_pnp[k], h->d_stats, h->d_stats + 6 * (size_t)h->ny, md->stream)))
            return rc;
    }
    SWH_HIP(hipStreamSynchronize(md->stream));
    // print_multidomain's accumulation over domains, in domain order
    double gmaxs = -DBL_MAX, gmins = DBL_MAX, gmaxv = 0.0, gminv = DBL_MAX, gpot = 0.0, gkin = 0.0, gtot = 0.0;
    for (int k = 0; k < md->nd; k++) {
        swh_domain *h = md->dom[k];
        double v[7];
        SWH_HIP(hipMemcpy(v, h->d_stats + 6 * (size_t)h->ny, sizeof v, hipMemcpyDeviceToHost));
        if (per_domain) {
            swh_domain_stats &o = per_domain[k];
            o.max_stage = v[0]; o.max_speed = v[1]; o.min_stage = v[2]; o.min_speed = v[3];
            o.energy_potential_on_rho = v[4]; o.energy_kinetic_on_rho = v[5]; o.energy_total_on_rho = v[6];
        }
        gmaxs = std::max(gmaxs, v[0]);
        gmins = std::min(gmins, v[2]);
        gmaxv = std::max(gmaxv, v[1]);
        gminv = std::min(gminv, v[3]);
        gpot = gpot + v[4];
        gkin = gkin + v[5];
        gtot = gtot + v[6];
    }
    if (global) {
        global->max_stage = gmaxs; global->max_speed = gmaxv; global->min_stage = gmins; global->min_speed = gminv;
        global->energy_potential_on_rho = gpot; global->energy_kinetic_on_rho = gkin;
        global->energy_total_on_rho = gtot;
    }
    if (energy_is_finite) *energy_is_finite = (gtot > DBL_MAX || gtot != gtot) ? 0 : 1;
    return fork_domains(md);
}

int swh_md_check_stability(swh_md *md, int64_t *n_bad)
{
    if (!md || !n_bad || !md->finalized) return md_fail(SWH_EINVAL, "swh_md_check_stability: bad argument");
    int rc = join_domains(md);
    if (rc) return rc;
    for (int k = 0; k < md->nd; k++) {
        swh_domain *h = md->dom[k];
        if ((rc = domain_stats_buffer(h))) return rc;
        if ((rc = domain_stability_launch(h, (unsigned long long *)(h->d_stats + 6 * (size_t)h->ny + 7), md->stream)))
            return rc;
    }
    SWH_HIP(hipStreamSynchronize(md->stream));
    int64_t total = 0;
    for (int k = 0; k < md->nd; k++) {
        swh_domain *h = md->dom[k];
        unsigned long long v = 0;
        SWH_HIP(hipMemcpy(&v, h->d_stats + 6 * (size_t)h->ny + 7, sizeof v, hipMemcpyDeviceToHost));
        total += (int64_t)v;
    }
    *n_bad = total;
    return fork_domains(md);
}

int swh_md_download_recv_buffer(swh_md *md, int32_t k, int32_t c, double *out)
{
    if (!md || !out || !md->finalized || k < 1 || k > md->nd || c < 1 || c > (int)md->recv[k - 1].size())
        return md_fail(SWH_EINVAL, "swh_md_download_recv_buffer: bad argument");
    SWH_HIP(hipStreamSynchronize(md->stream));
    const long long off = md->recv_buf_off[md->recv_base[k - 1] + c - 1];
    SWH_HIP(hipMemcpy(out, md->arena + off, md->recv[k - 1][c - 1].buffer_size * sizeof(double),
                      hipMemcpyDeviceToHost));
    return SWH_OK;
}

int swh_md_synchronize(swh_md *md)
{
    if (!md) return md_fail(SWH_EINVAL, "swh_md_synchronize: null handle");
    SWH_HIP(hipStreamSynchronize(md->stream));
    for (auto h : md->dom) SWH_HIP(hipStreamSynchronize(h->stream));
    return SWH_OK;
}

}  // extern "C"
