"""GPU: domain statistics and the stability scan on the device.

* compute_domain_statistics_new (domain_mod.f90:683-849): after some steps on the GPU,
  the state is handed to the compiled reference's own domain_type and its statistics
  are computed there; the device's max/min stage and speed and its potential, kinetic
  and total energy must equal them bit for bit (the device sums each row in column
  order and the rows in row order, as the reference does) -- staggered truely-linear
  ('linear'), staggered with depth from the stage (llpnf not truely linear,
  'leapfrog_nonlinear') and co-located with dry cells ('rk2' beach), FP64 and FP32K.
* print_multidomain (multidomain_mod.f90:2301-2447) on the nested three-level
  workload: each domain's statistics over its priority cells against the reference
  multidomain holding the same states, the global sums, energy_is_finite.
* check_multidomain_stability's scan (multidomain_mod.f90:348-455): zero on a sound
  state; planted NaN and out-of-range values in the interior are counted, values in
  the exterior cells are not.
"""
import json
import os
import subprocess
import sys
import tempfile

import numpy as np
import pytest

import helpers as H
import pyoracle

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

KEYS = ("max_stage", "max_speed", "min_stage", "min_speed", "energy_potential_on_rho", "energy_kinetic_on_rho",
        "energy_total_on_rho")


def _cases():
    sph = H.small_spherical()
    ntl = dict(sph, method="leapfrog_linear_plus_nonlinear_friction",
               manning_squared=np.full((sph["ny"], sph["nx"]), 0.03 ** 2))
    nl = H.nonlinear_leapfrog_cases()["land_sph"][0]()
    return {
        "linear_sph": (sph, 40, {}),
        "llpnf_ntl_sph": (ntl, 40, {"linear_solver_is_truely_linear": False}),
        "leapfrog_nonlinear_sph": (nl, 40, {}),
        "rk2_beach": (H.small_beach(), 60, {}),
    }


@pytest.mark.parametrize("precision", ["fp64", "fp32k"])
@pytest.mark.parametrize("case", list(_cases()))
def test_domain_statistics_vs_reference(oracle_built, case, precision):
    print("CASE statistics", case, precision, flush=True)
    cfg, nsteps, kw = _cases()[case]
    if not pyoracle.available("ref", cfg["spherical"]):
        pytest.skip("compiled reference not built")
    g = H.gpu_domain(cfg, precision=precision, **kw)
    if cfg["method"] in H.FV_METHODS:
        g.evolve_steps(None, nsteps)
    else:
        dt = g.stationary_timestep_max()
        g.evolve_steps(dt, nsteps)
    got = g.statistics()
    U = g.U
    ref_kw = {"truely_linear": False} if kw.get("linear_solver_is_truely_linear") is False else {}
    r = pyoracle.CpuDomain("ref", cfg["nx"], cfg["ny"], cfg["lw"], cfg["ll"], method=cfg["method"],
                           spherical=cfg["spherical"], **ref_kw)
    r.set_U(U)
    want = r.statistics()
    print(case, precision, json.dumps(got), flush=True)
    for k in KEYS:
        assert got[k] == want[k], (k, got[k], want[k])
    assert got["max_speed"] > 0 and got["energy_kinetic_on_rho"] > 0
    assert g.check_stability() == 0
    g.close()


def test_stability_scan_counts_planted_values():
    print("CASE stability scan", flush=True)
    cfg = H.small_spherical()
    g = H.gpu_domain(cfg)
    dt = g.stationary_timestep_max()
    g.evolve_steps(dt, 10)
    assert g.check_stability() == 0
    U = g.U
    U[0, 20, 30] = np.nan
    U[1, 40, 50] = 2e38
    U[2, 41, 50] = -1.8e38
    U[1, 42, 50] = 1.6e38        # inside the bound: not counted
    U[0, 0, 10] = np.nan         # exterior cell (width 2): not scanned
    U[2, 30, cfg["nx"] - 1] = np.inf
    g.U = U
    assert g.check_stability() == 3
    g.close()


def _ref_md_stats(states_npz):
    """Per-domain statistics of the reference multidomain holding the given states
    (child process: the reference's setup runs shell commands)."""
    out = tempfile.mktemp(suffix=".npz")
    code = r"""
import sys, numpy as np
sys.path.insert(0, %(root)r)
from oracle.pyoracle import RefMultidomain
from ptha_amd import configs
cfg = configs.nested_small()
md = RefMultidomain(cfg["domains"], spherical=cfg["spherical"])
z = np.load(%(inp)r)
res = {}
for k in range(md.nd):
    md.set_U(k, z["U_%%d" %% k])
    res["stats_%%d" %% k] = md.statistics(k)
np.savez(%(out)r, **res)
""" % dict(root=ROOT, inp=states_npz, out=out)
    subprocess.run([sys.executable, "-c", code], check=True, cwd=tempfile.mkdtemp(), timeout=300)
    z = np.load(out)
    return [z["stats_%d" % k] for k in range(len(z.files))]


@pytest.mark.parametrize("precision", ["fp64", "fp32k"])
def test_multidomain_statistics_vs_reference(oracle_built, precision):
    print("CASE multidomain statistics", precision, flush=True)
    if not pyoracle.available("ref", True):
        pytest.skip("compiled reference not built")
    from test_gpu_multidomain import build_gpu
    prefix = os.path.join(tempfile.mkdtemp(), "run")
    subprocess.run([sys.executable, os.path.join(ROOT, "tests", "md_ref_run.py"), prefix, "1"], check=True,
                   env=dict(os.environ, OMP_NUM_THREADS="1"), timeout=300)
    with open(prefix + ".json") as f:
        tables = json.load(f)
    arrays = dict(np.load(prefix + ".npz"))
    doms, md = build_gpu(tables, arrays, precision)
    md.make_initial_conditions_consistent()
    md.evolve_steps(tables["dt"], 8)
    per, glob, finite = md.statistics()
    assert finite
    assert md.check_stability() == 0
    states = tempfile.mktemp(suffix=".npz")
    np.savez(states, **{"U_%d" % k: d.U for k, d in enumerate(doms)})
    want = _ref_md_stats(states)
    for k in range(len(doms)):
        got = np.array([per[k][q] for q in KEYS])
        print("domain", k, precision, json.dumps(per[k]), flush=True)
        assert np.array_equal(got, want[k]), (k, got, want[k])
    # print_multidomain's global values
    assert glob["max_stage"] == max(p["max_stage"] for p in per)
    assert glob["min_speed"] == min(p["min_speed"] for p in per)
    tot = 0.0
    for p in per:
        tot = tot + p["energy_total_on_rho"]
    assert glob["energy_total_on_rho"] == tot
    md.close()
    for d in doms:
        d.close()
