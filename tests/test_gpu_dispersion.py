"""GPU: the staggered dispersive solve (swh_dispersion_*, linear_dispersive_solver_mod.f90)
driven as multidomain_mod.f90:1245-1410 drives one dispersive domain (extrapolation level
appended, U stored, the shallow-water step, outer dispersive iterations), against the
oracle, whose restatement is bit-identical to the compiled reference on the same cases
(tests/test_oracle_vs_reference.py::test_dispersion_equals_compiled_reference).

* FP64: bit-identical state ('linear' Cartesian channel with land, with and without the
  td1/td2 depth taper and with two outer iterations; a spherical basin;
  'leapfrog_nonlinear' on the channel);
* the cell-centred form with 'midpoint' (dispersion also in its predictor, at the device
  time; edge cells left by update_U): FP64 bit-identical (frictionless channel);
* FP32 + Kahan: within 1e-5 m (stage; measured 1.2e-6) and 5e-3 m^2/s (momenta;
  measured 9.8e-4, they reach 50 m^2/s) of the FP64 oracle on the channel.
"""
import numpy as np
import pytest

import helpers as H

pytestmark = pytest.mark.gpu
CASES = H.dispersive_cases()


def _runs(case, precision):
    make, nsteps, td1, td2, outer = CASES[case]
    cfg = make()
    o, dt, _ = H.cpu_run("oracle", cfg, 0)
    o.dispersion_setup(td1, td2, 2)
    g = H.gpu_domain(cfg, precision=precision)
    g.dispersion_enable(td1, td2, 2)
    for _ in range(nsteps):
        o.dispersive_step(dt, outer)
        g.dispersive_step(dt, outer)
    return o.get_U(), g


LF_CASES = sorted(k for k in CASES if "midpoint" not in k)
FV_CASES = sorted(k for k in CASES if "midpoint" in k)


@pytest.mark.parametrize("case", LF_CASES)
def test_dispersion_fp64_bitexact(oracle_built, case):
    print("CASE dispersion fp64", case, flush=True)
    Uo, g = _runs(case, "fp64")
    Ug = g.U
    for k in range(4):
        assert np.array_equal(Ug[k], Uo[k]), (k, np.abs(Ug[k] - Uo[k]).max())
    g.close()


@pytest.mark.parametrize("case", FV_CASES)
def test_dispersion_midpoint_fp64(oracle_built, case):
    """Cell-centred form inside 'midpoint' (the predictor solve with the device time,
    edge cells left by update_U) and the outer solves."""
    print("CASE dispersion midpoint fp64", case, flush=True)
    Uo, g = _runs(case, "fp64")
    Ug = g.U
    for k in range(4):
        assert np.array_equal(Ug[k], Uo[k]), (k, np.abs(Ug[k] - Uo[k]).max())
    g.close()


def test_dispersion_fp32k(oracle_built):
    print("CASE dispersion fp32k channel_linear", flush=True)
    Uo, g = _runs("channel_linear", "fp32k")
    Ug = g.U
    errs = [float(np.abs(Ug[k] - Uo[k]).max()) for k in range(3)]
    print("errors", errs, flush=True)
    assert errs[0] <= 1e-5
    assert max(errs[1:]) <= 5e-3
    g.close()


def test_dispersion_rejected_for_rk2n():
    from ptha_amd.backend import BackendError
    g = H.gpu_domain(dict(H.small_beach(), method="rk2n"))
    with pytest.raises(BackendError):
        g.dispersion_enable()
    g.close()
