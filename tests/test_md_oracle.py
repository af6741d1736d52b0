"""CPU: the reference's multidomain (compiled from the reference sources) as the
oracle of the nesting rows, pinned by what the mathematics and the reference's own
tests fix, and the buffer-layout rules the device tables rely on.

* planar-field invariant (the reference's test, multidomain_mod.f90:4088-4176, with
  a positive depth so the wet-dry limiter stays off): after one exchange every
  priority or received cell holds the planar value of its priority domain;
* mass: the sum of the priority-domain volumes is unchanged by 20 global steps of a
  closed three-level nest (flux correction makes nesting conservative);
* layout: each send box's buffer is [data | flux strips | time] with the sizes the
  device pack/unpack compute from the box geometry, and it pairs with exactly one
  receive box of the same size.
"""
import json
import os
import subprocess
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF_LIB = os.path.join(ROOT, "oracle", "_ref", "libswals_ref_sph.so")


def _run(tmp_path, mode, nsteps):
    if not os.path.exists(REF_LIB):
        pytest.skip("compiled reference not built")
    prefix = str(tmp_path / mode)
    env = dict(os.environ, OMP_NUM_THREADS="1")
    args = [sys.executable, os.path.join(ROOT, "tests", "md_ref_run.py"), prefix, str(nsteps)]
    if mode == "planar":
        args.append("planar")
    subprocess.run(args, check=True, env=env, timeout=300)
    with open(prefix + ".json") as f:
        return json.load(f), dict(np.load(prefix + ".npz"))


def test_planar_invariant(tmp_path):
    tables, arrays = _run(tmp_path, "planar", 0)
    off = [9.0, 3.0, 6.0, 0.0]
    for k, td in enumerate(tables["domains"]):
        d = td["dims"]
        lw, ll = td["lw"], td["ll"]
        x = ll[0] + (np.arange(1, d["nx"] + 1) - 0.5) / d["nx"] * lw[0]
        y = ll[1] + (np.arange(1, d["ny"] + 1) - 0.5) / d["ny"] * lw[1]
        X, Y = np.meshgrid(x, y)
        pdi = arrays["pdi_%d" % k]
        act = pdi == k + 1
        for b in td["recv"]:
            iL, iU, jL, jU = np.asarray(b["inds"]).ravel()[:4]
            act[jL - 1:jU, iL - 1:iU] = True
        U = arrays["Ucons_%d" % k]
        for v in range(4):
            exp = (X - 100.0) * 1.0e-4 + (Y + 60.0) * 3.0e-5 + pdi + off[v]
            err = np.abs(U[v] - exp)[act].max()
            assert err * err <= np.spacing(1.0), (k, v, err)


def test_mass_conservation_and_layout(tmp_path):
    tables, arrays = _run(tmp_path, "flow", 20)
    v0, vN = arrays["vol0"], arrays["volN"]
    assert abs(vN.sum() - v0.sum()) <= 1e-12 * v0.sum()
    nint = lambda x: int(round(x))
    doms = tables["domains"]
    for k, td in enumerate(doms):
        for c, b in enumerate(td["send"]):
            (iL, iU), (jL, jU), (kL, kU) = np.asarray(b["inds"])
            cr, eq = b["cell_ratios"], b["equal_cell_ratios"]
            nI, nJ, nK = iU - iL + 1, jU - jL + 1, kU - kL + 1
            inv = [1, 1] if eq else [nint(1 / cr[0]), nint(1 / cr[1])]
            rip = [1, 1] if eq else [nint(cr[0]), nint(cr[1])]
            mode = 0 if (eq and b["my_staggered"] == b["neighbour_staggered"]) else \
                (1 if (b["my_domain_is_finer"] or eq) else 2)
            n_data = nI * nJ * nK * 5 if mode == 2 else (nI if eq else nint(nI * cr[0])) * \
                (nJ if eq else nint(nJ * cr[1])) * nK
            fs = np.asarray(b["flux_shape"])
            assert list(fs[:, 0]) == [nI, nI, nJ, nJ]
            n_flux = 0
            for side in range(4):
                n = fs[side, 0]
                if mode == 0:
                    per = n
                elif mode == 1:
                    st = inv[0] if side < 2 else inv[1]
                    assert n % st == 0
                    per = n // st
                else:
                    per = n * (rip[0] if side < 2 else rip[1])
                n_flux += per * fs[side, 1]
            assert b["flux_start"] - 1 == n_data
            assert b["time_index"] - 1 == n_data + n_flux
            assert b["buffer_size"] == b["time_index"]
            rb = doms[b["neighbour_index"] - 1]["recv"][b["neighbour_comms_index"] - 1]
            assert rb["buffer_size"] == b["buffer_size"]
            assert (rb["neighbour_index"], rb["neighbour_comms_index"]) == (k + 1, c + 1)


def test_nested_tables_fixture_matches_setup():
    """Structure of the committed nesting tables of the nested workload
    (tests/golden/nested_c3_tables.npz, written from the reference's md%setup):
    full-size parent, priority arrays over the three domains, every domain has
    priority cells, and no receive box holds its own domain's priority cells."""
    z = np.load(os.path.join(ROOT, "tests", "golden", "nested_c3_tables.npz"))
    tables = json.loads(bytes(z["tables_json"]).decode())
    assert tables["scale"] == 1 and len(tables["domains"]) == 3
    assert [td["dims"]["nx"] for td in tables["domains"]][0] == 3000
    for k, td in enumerate(tables["domains"]):
        pdi = z["pdi_%d" % k]
        assert pdi.shape == (td["dims"]["ny"], td["dims"]["nx"])
        assert set(np.unique(pdi)) <= {1, 2, 3}
        assert (pdi == k + 1).any()
        for b in td["recv"]:
            (iL, iU), (jL, jU), _ = np.asarray(b["inds"])
            # a receive box never holds this domain's own priority cells
            assert not (pdi[jL - 1:jU, iL - 1:iU] == k + 1).any()
