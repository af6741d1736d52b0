"""Benchmark of the SWALS leapfrog hot path on one MI355X.

Workload (the batched unit-source configuration): 512 Gaussian unit sources (1 m,
sigma 25 km, 32 x 16 lattice at 50 km) on one shared 1200 x 1200 spherical grid at
1 arcmin, 'linear' leapfrog with Coriolis, FP32 state with Kahan-compensated stage,
power-law bathymetry with a shelf ramp to land, reflective walls, 64 gauges every
60 s, durations U(2 h, 12 h).

A step is one leapfrog step of every source (all 512 active over the timed steps).
`value` is cell-updates per second over the timed steps (Gcell-updates/s, inputs
resident in HBM). The same steps are also timed in FP64 ('fp64' object); the whole
configuration (every source to its own duration) gives unit-sources/hour; the
compiled reference solver on the host cores gives `cpu_baseline`.
"""
import argparse
import json
import math
import os
import subprocess
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "Gcell-updates/s/GPU FP32-comp vs FP64 leapfrog; % HBM peak; unit-sources/hr"
HBM_PEAK_GBS = 8000.0
MFMA_FP64_TFLOPS = 78.6   # MI355X dense FP64 matrix peak
# Algorithmic HBM bytes per cell-update of the fused batched leapfrog step (one read
# and one write of stage, Kahan term, uh, vh per source; elevation shared):
BYTES_PER_CELL = {"fp32k": 32.0, "fp64": 48.0}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def linear_timestep(elev, metrics, cfl=0.7):
    """linear_timestep_max (domain_mod.f90:1906-1931), vectorised (sqrt and min are
    exact, so this equals the reference's value)."""
    dbe, dle = metrics["distance_bottom_edge"], metrics["distance_left_edge"]
    dist = np.minimum((dbe[:-1] + dbe[1:])[:, None], (dle[:-1] + dle[1:])[None, :])
    ts = (0.5 * dist / np.sqrt(9.8 * np.maximum(0.0 - elev, 1.0e-05))).min()
    return float(ts * cfl)


def build_batch(cfg, precision, nsources, nsteps_per_source, sample_steps, block=None):
    from ptha_amd import configs as C
    from ptha_amd.batch import UnitSourceBatch
    b = UnitSourceBatch(cfg["nx"], cfg["ny"], cfg["lw"], cfg["ll"], nsources, precision=precision)
    b.set_elevation(cfg["elev"])
    b.set_nsteps(nsteps_per_source[:nsources])
    for k in range(nsources):
        b.set_source(k, C.unit_source_stage(cfg, k))
    b.setup_point_gauges(cfg["gauges_xy"], sample_steps=sample_steps)
    if block:
        b.set_launch(*block)
    return b


def time_steps(b, dt, steps, warmup):
    if warmup:
        b.evolve(dt, warmup)
    b.synchronize()
    t0 = time.perf_counter()
    ms = b.evolve_timed(dt, steps)
    b.synchronize()
    wall = time.perf_counter() - t0
    return wall, ms


REF_BUILDS = {
    "ref_fast": "amdflang -O3 -ffast-math -fopenmp (the reference's own default, -Ofast)",
    "ref": "amdflang -O2 -ffp-contract=off -fopenmp (the FP64 parity build)",
}


def cpu_baseline_leg(cfg, dt, budget_s):
    """Times the compiled reference solver on one unit source of the same grid, built two
    ways (oracle/_ref/libswals_ref_sph_fast.so with the reference's own default
    optimisation, and the parity build libswals_ref_sph.so), each for half the budget in
    a child process so OMP_NUM_THREADS applies; the faster build is the baseline and
    both figures are in the sample text."""
    runs = {}
    for kind in REF_BUILDS:
        try:
            runs[kind] = _cpu_ref_run(kind, dt, budget_s / len(REF_BUILDS))
        except Exception as e:  # a build may be absent from a tree
            log("cpu baseline (%s) unavailable: %s" % (kind, e))
    if not runs:
        raise RuntimeError("no reference build could be timed")
    best = max(runs, key=lambda k: runs[k]["gcell"])
    o = runs[best]
    others = "; ".join("%s: %.4f Gcell/s over %d steps (%.1f s)" % (REF_BUILDS[k], r["gcell"], r["steps"], r["seconds"])
                       for k, r in runs.items())
    return dict(value=o["gcell"], unit="Gcell-updates/s", cores=o["cores"], kind="reference",
                sample="compiled reference SWALS 'linear' spherical+Coriolis leapfrog (OpenMP), one unit source "
                       "of the 1200x1200 1-arcmin grid, no max_U/gauges; value from the faster build. "
                       "Builds timed: %s. CPU model: %s" % (others, cpu_model()))


def _cpu_ref_run(kind, dt, budget_s):
    code = r"""
import json, os, sys, time, math
import numpy as np
sys.path.insert(0, %(root)r); sys.path.insert(0, os.path.join(%(root)r, 'oracle'))
import pyoracle
from ptha_amd import configs as C
cfg = C.unit_sources()
nx, ny = cfg['nx'], cfg['ny']
d = pyoracle.CpuDomain(%(kind)r, nx, ny, cfg['lw'], cfg['ll'], method='linear', spherical=True)
d.set_record_max_U(False)
U = np.zeros((4, ny, nx)); U[3] = cfg['elev']
U[0] = C.gaussian_on_sphere(cfg['x'], cfg['y'], *cfg['centres'][0], cfg['sigma_km']*1e3)
U[0][U[3] >= 0] = 0.0
d.set_U(U)
dt = %(dt)r
t0 = time.perf_counter(); d.evolve(dt, 20); t1 = time.perf_counter()
n = max(20, int(%(budget)r / max(t1 - t0, 1e-9) * 20))
t0 = time.perf_counter(); d.evolve(dt, n); el = time.perf_counter() - t0
print(json.dumps(dict(steps=n, seconds=el, cells=nx*ny)))
""" % dict(root=ROOT, dt=dt, budget=budget_s, kind=kind)
    env = dict(os.environ)
    ncores = int(env.get("OMP_NUM_THREADS", "0")) or len(os.sched_getaffinity(0))
    env["OMP_NUM_THREADS"] = str(ncores)
    env.setdefault("OMP_PROC_BIND", "true")
    r = subprocess.run([sys.executable, "-c", code], env=env, capture_output=True, text=True, timeout=600)
    if r.returncode != 0:
        raise RuntimeError("cpu baseline failed: " + r.stderr[-2000:])
    o = json.loads(r.stdout.strip().splitlines()[-1])
    return dict(gcell=o["cells"] * o["steps"] / o["seconds"] / 1e9, steps=o["steps"], seconds=o["seconds"],
                cores=ncores)


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def read_traffic(name):
    """The dominant kernel's committed rocprofv3 counter summary (profiles/), if present:
    measured HBM bytes per launch, the compulsory bytes of that launch and its VALU issue
    figures."""
    p = os.path.join(ROOT, "profiles", "traffic_%s.json" % name)
    if os.path.exists(p):
        with open(p) as f:
            return json.load(f)
    return {}


def roofline(r):
    """The HBM roofline of the batched step kernel. A fused launch advances each of its
    sources by several steps while reading and writing their state (stage, Kahan term,
    UH, VH) once, so its compulsory traffic is 32 B per cell per launch and source (the
    shared elevation, 4 B per cell per launch, is left out). 'achieved' is that traffic,
    summed over the timed region's launches, over the region's device time (HIP events);
    'per_step_equivalent_gbs' counts 32 B per cell-update instead (what a kernel taking
    one step per pass would have to move), which exceeds the peak when launches fuse
    steps. 'traffic' is the measured HBM bytes of one full launch (rocprofv3 FETCH_SIZE and
    WRITE_SIZE scaled by a plain copy kernel) and 'achieved_measured' scales 'achieved' by
    its ratio to that launch's compulsory bytes. 'valu_issue_frac' is the same launch's
    VALU wave-instructions over 256 CUs x its busy cycles (one wave-instruction per CU
    per cycle)."""
    t = read_traffic("leapfrog_batch_fp32k")
    bpc = BYTES_PER_CELL["fp32k"]
    out = {"bound": "hbm", "achieved": round(r["comp_gbs"], 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
           "frac": round(r["comp_gbs"] / HBM_PEAK_GBS, 4), "traffic": t.get("bytes_per_launch"),
           "bytes_basis": "one read and one write of S, C, UH, VH per source and fused launch (32 B per cell)",
           "steps_per_launch": round(r["steps_per_launch"], 3), "launches": r["launches"],
           "per_step_equivalent_gbs": round(r["achieved_gbs"], 1),
           "per_step_equivalent_frac": round(r["achieved_gbs"] / HBM_PEAK_GBS, 4),
           "algorithmic_bytes_per_cell_update": bpc}
    if t.get("bytes_per_launch") and t.get("compulsory_bytes_per_launch"):
        ratio = t["bytes_per_launch"] / t["compulsory_bytes_per_launch"]
        gbs = r["comp_gbs"] * ratio
        out.update(traffic_over_compulsory=round(ratio, 4), achieved_measured=round(gbs, 1),
                   frac_measured=round(gbs / HBM_PEAK_GBS, 4))
    if t.get("valu_issue_frac"):
        out["valu_issue_frac"] = t["valu_issue_frac"]
    return out


def copy_probe(mib):
    """A plain device copy kernel of known size (swh_probe_copy): the reference
    point that scales rocprofv3's FETCH_SIZE / WRITE_SIZE counters to bytes, and
    the streaming bandwidth this GPU reaches with a copy."""
    import ctypes
    from ptha_amd import backend as B
    lib = B.load()
    ms = ctypes.c_double()
    n = mib * (1 << 20)
    B.check(lib.swh_probe_copy(n, 10, ctypes.byref(ms)), "swh_probe_copy")
    log("copy probe: %d MiB read + %d MiB written per copy, %.3f ms, %.0f GB/s" % (mib, mib, ms.value,
                                                                                2 * n / ms.value / 1e6))
    return 2 * n / ms.value / 1e6


def valu_probe():
    """VALU issue reference (swh_probe_valu): FMA wave-instructions per second for FP32
    (unpacked) and FP64 chains, 32 wavefronts per CU; the rate that the step kernels'
    SQ_INSTS_VALU counts are compared with."""
    import ctypes
    from ptha_amd import backend as B
    lib = B.load()
    out = {}
    for name, fp64 in (("fp32", 0), ("fp64", 1)):
        ms, wi = ctypes.c_double(), ctypes.c_double()
        B.check(lib.swh_probe_valu(fp64, 2048, 20000, 5, ctypes.byref(ms), ctypes.byref(wi)), "swh_probe_valu")
        out[name] = dict(ms=round(ms.value, 4), wave_insts=wi.value, g_wave_insts_per_s=round(wi.value / ms.value / 1e6, 1))
        log("VALU probe %s: %.0f FMA wave-instructions in %.3f ms = %.1f G/s" % (name, wi.value, ms.value,
                                                                               wi.value / ms.value / 1e6))
    return out


def secondary_legs(steps):
    """Throughput of the other single-grid paths (not the headline workload):
    rk2 on the 2000 x 1000 wet/dry beach (FP64 and FP32 state) and one 2048^2
    spherical 'linear' grid (the adjoint-inversion grid) in FP32 + Kahan and FP64, and
    'leapfrog_nonlinear' on that grid.
    Timed with host wall clock around synchronised step loops (inputs resident)."""
    from ptha_amd import configs as C
    from ptha_amd.domain import Domain
    out = {}
    cb = C.beach_rk2()
    for prec in ("fp64", "fp32k"):
        d = Domain(cb["nx"], cb["ny"], cb["lw"], cb["ll"], "rk2", precision=prec, max_U_variables=())
        d.U = cb["U"]
        d.set_manning_squared(cb["manning_squared"])
        bc = cb["stage_bc"]
        d.set_stage_boundary_sine(bc["amp"], bc["period"], sides=bc["sides"])
        d.evolve_steps(None, 5)
        d.synchronize()
        t0 = time.perf_counter()
        d.evolve_steps(None, steps)
        d.synchronize()
        wall = time.perf_counter() - t0
        out["rk2_beach_" + prec] = dict(gcell_per_s=round(cb["nx"] * cb["ny"] * steps / wall / 1e9, 3),
                                         ms_per_step=round(1e3 * wall / steps, 4))
        d.close()
    ca = C.adjoint_inversion()
    from ptha_amd.domain import build_metrics
    m = build_metrics(ca["nx"], ca["ny"], ca["lw"], ca["ll"], True)
    dt = linear_timestep(ca["elev"], m)
    # compulsory bytes per cell and launch: read S, (C,) UH, VH, E, write S, (C,) UH, VH;
    # swh_evolve_steps takes two steps per launch, so per cell-update half of that
    for prec, bpc in (("fp32k", 36.0 / 2), ("fp64", 56.0 / 2)):
        d = Domain(ca["nx"], ca["ny"], ca["lw"], ca["ll"], "linear", precision=prec, spherical=True,
                   max_U_variables=())
        U = np.zeros((4, ca["ny"], ca["nx"]))
        U[3] = ca["elev"]
        d.U = U
        d.evolve_steps(dt, 5)
        d.synchronize()
        t0 = time.perf_counter()
        d.evolve_steps(dt, steps)
        d.synchronize()
        wall = time.perf_counter() - t0
        g = ca["nx"] * ca["ny"] * steps / wall / 1e9
        out["linear_2048sq_" + prec] = dict(gcell_per_s=round(g, 3), ms_per_step=round(1e3 * wall / steps, 4),
                                            hbm_frac=round(g * bpc / HBM_PEAK_GBS, 4), steps_per_launch=2,
                                            bytes_per_cell_update=bpc)
        d.close()
    # the nonlinear leapfrog on the same grid (Manning 0.025, a 1 m hump), stationary CFL step
    X, Y = np.meshgrid(ca["x"], ca["y"])
    for prec in ("fp64", "fp32k"):
        d = Domain(ca["nx"], ca["ny"], ca["lw"], ca["ll"], "leapfrog_nonlinear", precision=prec, spherical=True,
                   max_U_variables=())
        U = np.zeros((4, ca["ny"], ca["nx"]))
        U[3] = ca["elev"]
        U[0] = np.maximum(np.exp(-((X - ca["src_lon"].mean()) ** 2 + (Y - ca["src_lat"].mean()) ** 2) / 0.5), U[3])
        d.U = U
        d.set_manning_squared(np.full((ca["ny"], ca["nx"]), 0.025 ** 2))
        dtn = d.stationary_timestep_max()
        d.evolve_steps(dtn, 5)
        d.synchronize()
        t0 = time.perf_counter()
        d.evolve_steps(dtn, steps)
        d.synchronize()
        wall = time.perf_counter() - t0
        out["nonlinear_2048sq_" + prec] = dict(gcell_per_s=round(ca["nx"] * ca["ny"] * steps / wall / 1e9, 3),
                                               ms_per_step=round(1e3 * wall / steps, 4))
        d.close()
    # adjoint source inversion grid: forward (gauges every 17 steps, device checkpoints
    # every 256) and the adjoint sweep, FP32 + Kahan, 2000 steps. An inversion calls both
    # once per iteration on the same map: the first calls (which also allocate the
    # checkpoints and the adjoint state and upload the elevation) are reported apart from
    # repeated ones.
    from ptha_amd.adjoint import GaugeMap
    nsteps = 2000
    gm = GaugeMap(ca["nx"], ca["ny"], ca["lw"], ca["ll"], ca["elev"], dt, nsteps, 17, ca["gauges_xy"],
                  precision="fp32k", spherical=True)
    X, Y = np.meshgrid(ca["x"], ca["y"])
    eta0 = np.exp(-((X - ca["src_lon"].mean()) ** 2 + (Y - ca["src_lat"].mean()) ** 2) / 0.5) * (ca["elev"] < 0)
    t0 = time.perf_counter()
    g = gm.forward(eta0)
    t1 = time.perf_counter()
    lam = gm.adjoint(np.ones_like(g))
    t2 = time.perf_counter()
    g = gm.forward(eta0)
    t3 = time.perf_counter()
    lam = gm.adjoint(np.ones_like(g))
    t4 = time.perf_counter()
    cells = ca["nx"] * ca["ny"] * nsteps
    out["adjoint_2048sq_fp32k"] = dict(forward_gcell_per_s=round(cells / (t3 - t2) / 1e9, 3),
                                        adjoint_gcell_per_s=round(cells / (t4 - t3) / 1e9, 3),
                                        first_forward_gcell_per_s=round(cells / (t1 - t0) / 1e9, 3),
                                        first_adjoint_gcell_per_s=round(cells / (t2 - t1) / 1e9, 3), steps=nsteps,
                                        gauge_every=17)
    assert np.isfinite(lam).all()
    gm.close()
    return out


def superpose_leg(b, nsrc, max_samples, n_events=4096, seed=20261203):
    """Earthquake events from the finished batch's gauge series on the FP64 matrix
    cores: rupture patches of 1-8 x 1-16 unit sources of the 32 x 16 lattice
    (uniform slip, then stochastic slip), summarised as the maximum stage per gauge;
    and 64 events as whole series. FLOP/s counts the dense product the kernel does
    (2 x events x sources (padded to 4) x columns), against the FP64 MFMA peak."""
    rng = np.random.default_rng(seed)
    nxs, nys = 32, max(1, nsrc // 32)
    events, weights = [], []
    for _ in range(n_events):
        w, l = int(rng.integers(1, min(8, nys) + 1)), int(rng.integers(1, 17))
        i0, j0 = int(rng.integers(0, nxs - l + 1)), int(rng.integers(0, nys - w + 1))
        src = [j * nxs + i for j in range(j0, j0 + w) for i in range(i0, i0 + l) if j * nxs + i < nsrc]
        events.append(src)
        weights.append(rng.lognormal(0.0, 0.5, size=len(src)))
    slip = rng.uniform(0.5, 10.0, size=n_events)
    ng = b._ng
    kpad = (nsrc + 3) // 4 * 4
    out = {}
    for name, kw in (("uniform", dict(slip=slip)), ("stochastic", dict(weights=weights))):
        b.superpose(events[:64], summary="max_stage", **{k: v[:64] for k, v in kw.items()})
        _, nv, ms = b.superpose(events, summary="max_stage", timed=True, **kw)
        flops = 2.0 * n_events * kpad * float(nv.max()) * ng
        out["max_stage_" + name] = dict(events=n_events, kernel_ms=round(ms, 3),
                                        events_per_s=round(n_events / (ms * 1e-3), 1),
                                        dense_tflops=round(flops / (ms * 1e-3) / 1e12, 2),
                                        mfma_fp64_frac=round(flops / (ms * 1e-3) / 1e12 / MFMA_FP64_TFLOPS, 4))
    _, nv, ms = b.superpose(events[:64], slip=slip[:64], summary="series", timed=True)
    flops = 2.0 * 64 * kpad * float(max_samples) * 3 * ng
    out["series_64_events"] = dict(kernel_ms=round(ms, 3), dense_tflops=round(flops / (ms * 1e-3) / 1e12, 2))
    out["_events"] = (events[:64], slip[:64])
    log("superpose", out)
    return out


def cpu_baseline_superpose(b, events, slip):
    """CPU baseline of the event summation: the oracle's restatement of the
    reference's R summation (numpy, one core) on 64 events (maximum stage)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import superpose as SP
    need = sorted({k for ev in events for k in ev})
    series = {k: b.gauge_series(k) for k in need}
    ns = {k: len(v) for k, v in series.items()}
    t0 = time.perf_counter()
    for e in range(len(events)):
        SP.event_max_stage(series, ns, events[e], slip=slip[e])
    return dict(value=round(len(events) / (time.perf_counter() - t0), 1), unit="events/s", cores=1, kind="port",
                sample="%d events (maximum stage at every gauge) of the same series" % len(events))


def load_nested_tables(name="nested_c3_tables.npz"):
    """Nesting geometry of the nested workload (what the reference's md%setup decides
    for ptha_amd.configs.nested_c3; written by tests/golden/make_nested_tables.py)."""
    z = np.load(os.path.join(ROOT, "tests", "golden", name))
    tables = json.loads(bytes(z["tables_json"]).decode())
    arrays = {k: z[k] for k in z.files if k != "tables_json"}
    return tables, arrays


def nested_leg(prec, steps, warmup):
    """Three-level nested multidomain on the GPU: Gcell-updates/s over all domains'
    cells and sub-steps, wall time per global step, and the relative change of the
    total priority-domain volume over the timed steps."""
    from ptha_amd import configs as C
    from ptha_amd.domain import Domain
    from ptha_amd.multidomain import MultiDomain
    tables, arrays = load_nested_tables()
    cfg = C.nested_c3(scale=tables["scale"])
    doms = []
    cells = 0
    for k, td in enumerate(tables["domains"]):
        d = td["dims"]
        dom = Domain(d["nx"], d["ny"], td["lw"], td["ll"], timestepping_method=tables["methods"][k], precision=prec,
                     spherical=True, max_U_variables=(), exterior_cells_width=d["exterior_cells_width"])
        dom.U = cfg["state"](dom.x, dom.y)
        if tables["methods"][k] != "linear":
            dom.set_manning_squared(np.full((d["ny"], d["nx"]), cfg["manning_n"] ** 2))
        doms.append(dom)
        cells += d["nx"] * d["ny"] * tables["trf"][k]
    prio = [(arrays["pdi_%d" % k].astype(np.int32), arrays["pnp_%d" % k].astype(np.int32)) for k in range(len(doms))]
    md = MultiDomain(doms, arrays["all_dx"], arrays["all_methods"], tables["trf"],
                     [td["send"] for td in tables["domains"]], [td["recv"] for td in tables["domains"]], prio)
    md.make_initial_conditions_consistent()
    dt = tables["dt"]
    md.evolve_steps(dt, max(warmup, 1))
    md.synchronize()
    v0 = md.volumes()
    t0 = time.perf_counter()
    md.evolve_steps(dt, steps)
    md.synchronize()
    wall = time.perf_counter() - t0
    v1 = md.volumes()
    # the same global steps replayed from the captured graphs (swh_md_graph_launch)
    md.graph_capture(dt)
    md.graph_launch(2)
    md.synchronize()
    t0 = time.perf_counter()
    md.graph_launch(steps)
    md.synchronize()
    gwall = time.perf_counter() - t0
    out = dict(gcell_per_s=round(cells * steps / wall / 1e9, 3), ms_per_global_step=round(1e3 * wall / steps, 3),
               simulated_s_per_wall_s=round(dt * steps / wall, 1), cell_updates_per_global_step=cells,
               relative_volume_change=float(abs(v1.sum() - v0.sum()) / v0.sum()),
               graph_gcell_per_s=round(cells * steps / gwall / 1e9, 3),
               graph_ms_per_global_step=round(1e3 * gwall / steps, 3))
    md.close()
    for d in doms:
        d.close()
    return out


def nested_main(args):
    res = {}
    for prec in (["fp32k"] if args.no_fp64 else ["fp32k", "fp64"]):
        log("nested multidomain", prec)
        res[prec] = nested_leg(prec, args.steps, args.warmup)
        log(prec, res[prec])
    r = res["fp32k"]
    tables, _ = load_nested_tables()
    out = {
        "metric": METRIC,
        "value": r["gcell_per_s"],
        "unit": "Gcell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": r["ms_per_global_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "three-level nested multidomain: 'linear' parent 3000x1800 at 2 arcmin, "
                               "'leapfrog_linear_plus_nonlinear_friction' child at 1/3, 'rk2' grandchild at 1/15 "
                               "(in water on the shelf), FP32 + Kahan-compensated stage",
                   "dt_s": tables["dt"], "timestepping_refinement_factor": tables["trf"],
                   "grids": [[td["dims"]["nx"], td["dims"]["ny"]] for td in tables["domains"]]},
        "nested": res,
    }
    print(json.dumps(out))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--sources", type=int, default=512)
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--no-full", action="store_true", help="skip the whole-configuration unit-sources/hr leg")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=20.0)
    ap.add_argument("--block", type=str, default="",
                    help="block_x,rows_per_block,variant of the step kernel (default: the batch's autotuned shape)")
    ap.add_argument("--workload", choices=["unit_sources", "nested", "paths"], default="unit_sources",
                    help="unit_sources (the default measurement), the three-level nested multidomain, or "
                         "paths: only the single-grid legs (rk2 beach, 2048^2 linear, nonlinear, adjoint)")
    ap.add_argument("--no-secondary", action="store_true",
                    help="skip the rk2 beach and single 2048^2 grid throughput legs")
    ap.add_argument("--no-nested", action="store_true",
                    help="skip the three-level nested multidomain leg of the default run")
    ap.add_argument("--valu-probe", action="store_true", help="after the timed legs, measure the VALU FMA issue rate")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="after the timed legs, run a plain device copy of MIB MiB (for calibrating "
                         "rocprofv3 FETCH_SIZE / WRITE_SIZE) and log its bandwidth")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU only: this build has no multi-GPU path; rerun with --gpus 1",
              file=sys.stderr)
        sys.exit(3)

    if args.workload == "nested":
        return nested_main(args)
    if args.workload == "paths":
        out = secondary_legs(args.steps)
        if args.copy_probe:
            copy_probe(args.copy_probe)
        if args.valu_probe:
            valu_probe()
        print(json.dumps(out))
        return

    from ptha_amd import configs as C
    from ptha_amd.domain import build_metrics
    from ptha_amd.batch import reference_write_schedule
    block = tuple(int(v) for v in args.block.split(",")) if args.block else None
    log("building workload")
    cfg = C.unit_sources()
    nx, ny = cfg["nx"], cfg["ny"]
    m = build_metrics(nx, ny, cfg["lw"], cfg["ll"], True)
    dt = linear_timestep(cfg["elev"], m)
    # steps and gauge write steps of the reference's unit-source driver
    # (generic_model.f90:380-405): every source stops before the first step past its
    # duration; gauges are written at the first step on or after each 60 s
    plans = [reference_write_schedule(dt, float(d), cfg["gauge_interval"]) for d in cfg["durations"]]
    nsteps_src = np.array([p[0] for p in plans], dtype=np.int32)
    sched = plans[int(np.argmax(nsteps_src))][1]
    nsrc = args.sources
    max_samples = len(sched)
    total_steps = max(args.warmup + args.steps, 1)
    if nsteps_src[:nsrc].min() < total_steps:
        raise SystemExit("timed steps exceed the shortest source")

    cells = nx * ny
    res = {}
    tuned = {}
    for prec in (["fp32k"] if args.no_fp64 else ["fp32k", "fp64"]):
        log("timing", prec)
        b = build_batch(cfg, prec, nsrc, nsteps_src, sched, block)
        shape = block
        if not block:
            shape, tune_ms = b.autotune(dt, reps=3)
            log(prec, "autotuned launch shape (block_x, rows, variant) =", shape, "%.3f ms/step" % tune_ms)
        tuned[prec] = shape
        if args.warmup:
            b.evolve(dt, args.warmup)
        c0 = b.launch_counts()
        wall, ms = time_steps(b, dt, args.steps, 0)
        c1 = b.launch_counts()
        step, nact = b.status()
        assert nact == nsrc
        gcell = nsrc * cells * args.steps / wall / 1e9
        launch_ms = ms / args.steps
        alg = BYTES_PER_CELL[prec] * nsrc * cells
        nl, sl, ss = (c1[i] - c0[i] for i in range(3))
        assert ss == nsrc * args.steps
        res[prec] = dict(wall=wall, event_ms=ms, gcell=gcell, launch_ms=launch_ms,
                         achieved_gbs=alg / (launch_ms * 1e-3) / 1e9, launches=nl, steps_per_launch=ss / sl,
                         comp_gbs=BYTES_PER_CELL[prec] * cells * sl / (ms * 1e-3) / 1e9)
        log(prec, "Gcell/s %.1f  ms/step %.3f  HBM %.0f GB/s" % (gcell, 1e3 * wall / args.steps,
                                                                 res[prec]["achieved_gbs"]))
        b.close()

    full = None
    if not args.no_full:
        log("whole configuration (every source to its duration)")
        b = build_batch(cfg, "fp32k", nsrc, nsteps_src, sched, tuned.get("fp32k"))
        b.synchronize()
        t0 = time.perf_counter()
        chunk = 2000
        while True:
            step, nact = b.status()
            if nact == 0:
                break
            b.evolve(dt, chunk)
            b.synchronize()
            log("  step", b.status()[0], "active", b.status()[1])
        wall = time.perf_counter() - t0
        total_cell_updates = float(cells) * float(nsteps_src[:nsrc].sum())
        full = dict(unit_sources_per_hr=nsrc / (wall / 3600.0), seconds=wall,
                    gcell_per_s=total_cell_updates / wall / 1e9, sources=nsrc,
                    gauge_samples_per_source=int(b.gauge_series(0).shape[0]))
        log("earthquake events from the unit-source gauge series")
        superpose = superpose_leg(b, nsrc, max_samples)
        ev64, slip64 = superpose.pop("_events")
        if not args.no_cpu:
            try:
                superpose["cpu_baseline"] = cpu_baseline_superpose(b, ev64, slip64)
            except Exception as e:  # the oracle may be absent from a tree
                log("superposition cpu baseline unavailable:", e)
        b.close()

    if args.copy_probe:
        copy_probe(args.copy_probe)
    if args.valu_probe:
        valu_probe()

    secondary = None
    if not args.no_secondary:
        log("secondary legs: rk2 beach, single 2048^2 grid")
        secondary = secondary_legs(50)
    nested = None
    if not args.no_nested:
        log("nested multidomain leg (eager steps and graph replays)")
        nested = {prec: nested_leg(prec, 20, 2) for prec in ("fp32k", "fp64")}
        log("nested", nested)

    cpu = None
    if not args.no_cpu:
        try:
            log("cpu baseline")
            cpu = cpu_baseline_leg(cfg, dt, args.cpu_seconds)
        except Exception as e:  # the reference library may be absent from a tree
            log("cpu baseline unavailable:", e)

    r32 = res["fp32k"]
    out = {
        "metric": METRIC,
        "value": round(r32["gcell"], 3),
        "unit": "Gcell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(1e3 * r32["wall"] / args.steps, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "batched unit sources: %d Gaussian sources on one 1200x1200 spherical 1-arcmin "
                               "'linear' leapfrog grid (Coriolis), FP32 state + Kahan-compensated stage" % nsrc,
                   "nx": nx, "ny": ny, "sources": nsrc, "dt_s": dt, "gauges": 64, "gauge_write_s": cfg["gauge_interval"],
                   "launch_fp32k": list(tuned["fp32k"])},
        "roofline": roofline(r32),
        "cpu_baseline": cpu,
    }
    if "fp64" in res:
        r64 = res["fp64"]
        out["fp64"] = {"value": round(r64["gcell"], 3), "ms_per_step": round(1e3 * r64["wall"] / args.steps, 4),
                       "achieved_gbs": round(r64["comp_gbs"], 1),
                       "frac": round(r64["comp_gbs"] / HBM_PEAK_GBS, 4),
                       "steps_per_launch": round(r64["steps_per_launch"], 3),
                       "per_step_equivalent_gbs": round(r64["achieved_gbs"], 1),
                       "fp32k_over_fp64": round(r32["gcell"] / r64["gcell"], 3)}
    if secondary:
        out["secondary"] = secondary
    if nested:
        out["nested"] = nested
    if full:
        out["superpose"] = superpose
        out["unit_sources_per_hr"] = round(full["unit_sources_per_hr"], 1)
        out["full_run"] = {k: (round(v, 3) if isinstance(v, float) else v) for k, v in full.items()}
    print(json.dumps(out))


if __name__ == "__main__":
    main()
