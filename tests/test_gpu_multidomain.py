"""GPU parity of the multidomain (nested grids) step against the reference's own
multidomain_type (compiled from the reference sources, oracle/_ref).

The reference runs in a child process (tests/md_ref_run.py) on the three-level
nest of ``ptha_amd.configs.nested_small``: 'linear' parent, 'leapfrog_linear_plus_
nonlinear_friction' child, 'rk2' grandchild (3 sub-steps per global step), with a
shore inside the grandchild. The GPU gets the same geometry, nesting boxes and
priority arrays that the reference's md%setup produced, and the same initial U.

Tolerances (FP64): the leapfrog domains and every exchange use the reference's
operation order, so the initial exchange must match bit for bit; after steps the
rk2 grandchild (and through nesting everything) may differ at round-off level,
so stage/UH/VH are compared at 1e-9 relative to the field's scale. Mass: the sum
of the priority-domain volumes changes by at most 1e-12 relative.
"""
import json
import os
import subprocess
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF_LIB = os.path.join(ROOT, "oracle", "_ref", "libswals_ref_sph.so")
NSTEPS = 20

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ref(tmp_path_factory):
    if not os.path.exists(REF_LIB):
        pytest.skip("compiled reference not built (make -C oracle ref)")
    prefix = str(tmp_path_factory.mktemp("mdref") / "run")
    env = dict(os.environ, OMP_NUM_THREADS="1")
    print("reference multidomain run:", prefix, flush=True)
    subprocess.run([sys.executable, os.path.join(ROOT, "tests", "md_ref_run.py"), prefix, str(NSTEPS)], check=True,
                   env=env, timeout=300)
    with open(prefix + ".json") as f:
        tables = json.load(f)
    arrays = dict(np.load(prefix + ".npz"))
    return tables, arrays


def build_gpu(tables, arrays, precision):
    from ptha_amd import configs
    from ptha_amd.domain import Domain
    from ptha_amd.multidomain import MultiDomain

    cfg = configs.nested_small()
    doms = []
    for k, td in enumerate(tables["domains"]):
        d = td["dims"]
        dom = Domain(d["nx"], d["ny"], td["lw"], td["ll"], timestepping_method=tables["methods"][k],
                     precision=precision, spherical=True, max_U_variables=(),
                     exterior_cells_width=d["exterior_cells_width"])
        dom.U = arrays["U0_%d" % k]
        if tables["methods"][k] != "linear":
            dom.set_manning_squared(np.full((d["ny"], d["nx"]), cfg["manning_n"] ** 2))
        doms.append(dom)
    trf = [td["dims"]["timestepping_refinement_factor"] for td in tables["domains"]]
    prio = [(arrays["pdi_%d" % k], arrays["pnp_%d" % k]) for k in range(len(doms))]
    md = MultiDomain(doms, arrays["all_dx"], arrays["all_methods"], trf,
                     [td["send"] for td in tables["domains"]], [td["recv"] for td in tables["domains"]], prio)
    return doms, md


def _close(doms, md):
    md.close()
    for d in doms:
        d.close()


def test_metrics_match_reference(ref):
    from ptha_amd.domain import build_metrics
    tables, arrays = ref
    for k, td in enumerate(tables["domains"]):
        m = build_metrics(td["dims"]["nx"], td["dims"]["ny"], td["lw"], td["ll"], True, True)
        for name in ["distance_bottom_edge", "distance_left_edge", "area_cell_y", "coriolis", "coriolis_bottom_edge"]:
            np.testing.assert_array_equal(m[name], arrays["%s_%d" % (name, k)], err_msg="%s of domain %d" % (name, k))


def test_initial_exchange_bit_exact(ref):
    tables, arrays = ref
    print("multidomain fp64: initial exchange", flush=True)
    doms, md = build_gpu(tables, arrays, "fp64")
    try:
        md.make_initial_conditions_consistent()
        for k, d in enumerate(doms):
            for c in range(len(tables["domains"][k]["recv"])):
                got = md.recv_buffer(k, c)
                exp = arrays["rbuf0_%d_%d" % (k, c)]
                n = tables["domains"][k]["recv"][c]["time_index"] - 1
                np.testing.assert_array_equal(got[:n], exp[:n], err_msg="recv buffer %d of domain %d" % (c, k))
            np.testing.assert_array_equal(d.U, arrays["Ucons_%d" % k], err_msg="U of domain %d" % k)
    finally:
        _close(doms, md)


def _rel(a, b):
    scale = max(np.abs(b).max(), 1e-30)
    return np.abs(a - b).max() / scale


def test_steps_match_reference_fp64(ref):
    tables, arrays = ref
    dt = tables["dt"]
    print("multidomain fp64: %d steps" % NSTEPS, flush=True)
    doms, md = build_gpu(tables, arrays, "fp64")
    try:
        md.make_initial_conditions_consistent()
        v0 = md.volumes()
        md.evolve_one_step(dt)
        for k, d in enumerate(doms):
            U = d.U
            for v in range(3):
                assert _rel(U[v], arrays["U1_%d" % k][v]) < 1e-9, (k, v, _rel(U[v], arrays["U1_%d" % k][v]))
        md.evolve_steps(dt, NSTEPS - 1)
        for k, d in enumerate(doms):
            U = d.U
            for v in range(3):
                err = _rel(U[v], arrays["UN_%d" % k][v])
                assert err < 1e-9, (k, v, err)
        v = md.volumes()
        np.testing.assert_allclose(v0, arrays["vol0"], rtol=1e-12)
        np.testing.assert_allclose(v, arrays["volN"], rtol=1e-9)
        assert abs(v.sum() - v0.sum()) <= 1e-12 * v0.sum()
    finally:
        _close(doms, md)


def test_steps_fp32k_error_and_mass(ref):
    """FP32 with compensated stage against the FP64 reference: stage within 1e-3 m
    after 20 steps of a 1 m wave, mass change below 1e-6 relative."""
    tables, arrays = ref
    dt = tables["dt"]
    print("multidomain fp32k: %d steps" % NSTEPS, flush=True)
    doms, md = build_gpu(tables, arrays, "fp32k")
    try:
        md.make_initial_conditions_consistent()
        v0 = md.volumes()
        md.evolve_steps(dt, NSTEPS)
        for k, d in enumerate(doms):
            U = d.U
            wet = arrays["UN_%d" % k][3] < 0
            err = np.abs(U[0] - arrays["UN_%d" % k][0])[wet].max()
            assert err < 1e-3, (k, err)
        v = md.volumes()
        assert abs(v.sum() - v0.sum()) <= 1e-6 * v0.sum()
    finally:
        _close(doms, md)


@pytest.fixture(scope="module", params=["euler", "midpoint", "rk2n"])
def ref_fv(request, tmp_path_factory):
    """The same nest with another finite-volume method in the grandchild."""
    if not os.path.exists(REF_LIB):
        pytest.skip("compiled reference not built (make -C oracle ref)")
    prefix = str(tmp_path_factory.mktemp("mdref_" + request.param) / "run")
    env = dict(os.environ, OMP_NUM_THREADS="1")
    print("reference multidomain run, grandchild", request.param, prefix, flush=True)
    subprocess.run([sys.executable, os.path.join(ROOT, "tests", "md_ref_run.py"), prefix, str(NSTEPS),
                    "flow:" + request.param], check=True, env=env, timeout=300)
    with open(prefix + ".json") as f:
        tables = json.load(f)
    return tables, dict(np.load(prefix + ".npz"))


def test_steps_other_fv_grandchild_fp64(ref_fv):
    """20 global steps with an 'euler', 'midpoint' or 'rk2n' grandchild (their own
    nesting flux-integral bookkeeping, timestepping_alternatives_include.f90:8-336)
    against the reference's multidomain, at the bar of the rk2 nest."""
    tables, arrays = ref_fv
    dt = tables["dt"]
    print("multidomain fp64, grandchild", tables["methods"][2], flush=True)
    doms, md = build_gpu(tables, arrays, "fp64")
    try:
        md.make_initial_conditions_consistent()
        v0 = md.volumes()
        md.evolve_steps(dt, NSTEPS)
        for k, d in enumerate(doms):
            U = d.U
            for v in range(3):
                err = _rel(U[v], arrays["UN_%d" % k][v])
                assert err < 1e-9, (k, v, err)
        v = md.volumes()
        np.testing.assert_allclose(v, arrays["volN"], rtol=1e-9)
        assert abs(v.sum() - v0.sum()) <= 1e-12 * v0.sum()
    finally:
        _close(doms, md)


@pytest.fixture(scope="module")
def ref_nl(tmp_path_factory):
    """The same nest with a 'leapfrog_nonlinear' child (Manning n = 0.03)."""
    if not os.path.exists(REF_LIB):
        pytest.skip("compiled reference not built (make -C oracle ref)")
    prefix = str(tmp_path_factory.mktemp("mdref_nl") / "run")
    env = dict(os.environ, OMP_NUM_THREADS="1")
    print("reference multidomain run, leapfrog_nonlinear child", prefix, flush=True)
    subprocess.run([sys.executable, os.path.join(ROOT, "tests", "md_ref_run.py"), prefix, str(NSTEPS),
                    "child:leapfrog_nonlinear"], check=True, env=env, timeout=300)
    with open(prefix + ".json") as f:
        tables = json.load(f)
    return tables, dict(np.load(prefix + ".npz"))


def test_steps_nonlinear_child_fp64(ref_nl):
    """20 global steps with a 'leapfrog_nonlinear' child: its nesting flux integrals
    (mass only, domain_mod_leapfrog_nonlinear_include.f90:93-102), its receive-side
    wet/dry rules (no outflow from dry cells, outflow into a dry neighbour only above
    its elevation, domain_mod.f90:3834-3882) and the nesting layer the reference's
    setup chose for it, against the reference's multidomain at the bar of the rk2 nest
    (measured: bit-identical on every domain); the initial exchange bit for bit."""
    tables, arrays = ref_nl
    assert tables["methods"][1] == "leapfrog_nonlinear"
    dt = tables["dt"]
    print("multidomain fp64, leapfrog_nonlinear child", flush=True)
    doms, md = build_gpu(tables, arrays, "fp64")
    try:
        md.make_initial_conditions_consistent()
        for k, d in enumerate(doms):
            np.testing.assert_array_equal(d.U, arrays["Ucons_%d" % k], err_msg="U of domain %d" % k)
        v0 = md.volumes()
        md.evolve_one_step(dt)
        for k, d in enumerate(doms):
            U = d.U
            for v in range(3):
                err = _rel(U[v], arrays["U1_%d" % k][v])
                assert err < 1e-9, (k, v, err)
        md.evolve_steps(dt, NSTEPS - 1)
        for k, d in enumerate(doms):
            U = d.U
            for v in range(3):
                err = _rel(U[v], arrays["UN_%d" % k][v])
                print("domain", k, "variable", v, "relative error", err, flush=True)
                assert err < 1e-9, (k, v, err)
        v = md.volumes()
        np.testing.assert_allclose(v, arrays["volN"], rtol=1e-9)
        assert abs(v.sum() - v0.sum()) <= 1e-12 * v0.sum()
    finally:
        _close(doms, md)


def test_steps_nonlinear_child_fp32k(ref_nl):
    """The nest with a 'leapfrog_nonlinear' child in FP32 with compensated stage
    against the FP64 reference: the bars of the FP32K nest (stage within 1e-3 m of a
    1 m wave after 20 steps, mass change below 1e-6 relative; measured 1.3e-7 m)."""
    tables, arrays = ref_nl
    dt = tables["dt"]
    print("multidomain fp32k, leapfrog_nonlinear child", flush=True)
    doms, md = build_gpu(tables, arrays, "fp32k")
    try:
        md.make_initial_conditions_consistent()
        v0 = md.volumes()
        md.evolve_steps(dt, NSTEPS)
        for k, d in enumerate(doms):
            wet = arrays["UN_%d" % k][3] < 0
            err = np.abs(d.U[0] - arrays["UN_%d" % k][0])[wet].max()
            print("domain", k, "fp32k stage error (m)", err, flush=True)
            assert err < 1e-3, (k, err)
        v = md.volumes()
        assert abs(v.sum() - v0.sum()) <= 1e-6 * v0.sum()
    finally:
        _close(doms, md)


@pytest.mark.parametrize("precision", ["fp64", "fp32k"])
def test_nonlinear_child_graph_replay_equals_eager(ref_nl, precision):
    """The nest with a 'leapfrog_nonlinear' child replayed from HIP graphs (captured
    after one eager step) equals eager global steps bit for bit: state and time of
    every domain."""
    tables, arrays = ref_nl
    dt = tables["dt"]
    print("multidomain graph replay, leapfrog_nonlinear child", precision, flush=True)
    doms, md = build_gpu(tables, arrays, precision)
    try:
        md.make_initial_conditions_consistent()
        md.evolve_steps(dt, 6)
        md.synchronize()
        want = [(d.U.copy(), d.time) for d in doms]
    finally:
        _close(doms, md)
    doms, md = build_gpu(tables, arrays, precision)
    try:
        md.make_initial_conditions_consistent()
        md.evolve_steps(dt, 1)
        md.graph_capture(dt)
        md.graph_launch(2)
        md.evolve_steps(dt, 1)
        md.graph_launch(2)
        md.synchronize()
        for k, d in enumerate(doms):
            assert np.array_equal(d.U, want[k][0]), ("state", k)
            assert d.time == want[k][1], ("time", k)
        md.graph_destroy()
    finally:
        _close(doms, md)


@pytest.fixture(scope="module")
def ref_planar(tmp_path_factory):
    if not os.path.exists(REF_LIB):
        pytest.skip("compiled reference not built (make -C oracle ref)")
    prefix = str(tmp_path_factory.mktemp("mdplanar") / "run")
    env = dict(os.environ, OMP_NUM_THREADS="1")
    print("reference planar exchange:", prefix, flush=True)
    subprocess.run([sys.executable, os.path.join(ROOT, "tests", "md_ref_run.py"), prefix, "0", "planar"], check=True,
                   env=env, timeout=300)
    with open(prefix + ".json") as f:
        tables = json.load(f)
    return tables, dict(np.load(prefix + ".npz"))


def test_planar_field_exchange(ref_planar):
    """The reference's nesting invariant (multidomain_mod.f90:4088-4176): after one
    exchange of a planar field offset by the domain index, every priority or
    received cell holds the planar value of its priority domain (round-off only;
    the reference allows err**2 <= spacing(1.0)). The GPU exchange must also equal
    the reference's exchange bit for bit."""
    tables, arrays = ref_planar
    print("multidomain planar exchange", flush=True)
    doms, md = build_gpu(tables, arrays, "fp64")
    x0, y0 = 100.0, -60.0
    off = [9.0, 3.0, 6.0, 0.0]
    try:
        md.make_initial_conditions_consistent()
        for k, d in enumerate(doms):
            U = d.U
            np.testing.assert_array_equal(U, arrays["Ucons_%d" % k], err_msg="domain %d" % k)
            td = tables["domains"][k]
            pdi = arrays["pdi_%d" % k]
            X, Y = np.meshgrid(d.x, d.y)
            act = pdi == k + 1
            for b in td["recv"]:
                iL, iU, jL, jU = np.asarray(b["inds"]).ravel()[:4]
                act[jL - 1:jU, iL - 1:iU] = True
            for v in range(4):
                exp = (X - x0) * 1.0e-4 + (Y - y0) * 3.0e-5 + pdi + off[v]
                err = np.abs(U[v] - exp)[act].max()
                assert err * err <= np.spacing(1.0), (k, v, err)
    finally:
        _close(doms, md)
