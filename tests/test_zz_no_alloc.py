"""GPU: swh_domain_statistics_noalloc (the statistics entry point that allocates no
device memory: its row sums live in the library's static device array) against
swh_domain_statistics (row sums in the handle's own buffer), bit for bit, on two
calls in a row where the second domain is smaller than the first, so the second
call finds the first call's values in the shared array; and the row limit."""
import ctypes

import numpy as np
import pytest

import helpers as H
from ptha_amd import backend as B

pytestmark = pytest.mark.gpu

KEYS = ("max_stage", "max_speed", "min_stage", "min_speed", "energy_potential_on_rho", "energy_kinetic_on_rho",
        "energy_total_on_rho")


def _both(g):
    lib = B.load()
    a, b = B.DomainStats(), B.DomainStats()
    B.check(lib.swh_domain_statistics_noalloc(g._h, ctypes.byref(a)), "swh_domain_statistics_noalloc")
    B.check(lib.swh_domain_statistics(g._h, ctypes.byref(b)), "swh_domain_statistics")
    return [getattr(a, k) for k in KEYS], [getattr(b, k) for k in KEYS]


def test_noalloc_statistics_equal_allocating_entry_point():
    print("CASE no-alloc statistics: large then smaller domain", flush=True)
    big = H.small_spherical(nx=180, ny=160, arcmin=5.0)
    small = H.small_beach(nx=60, ny=30)
    g1 = H.gpu_domain(big)
    g1.evolve_steps(g1.stationary_timestep_max(), 30)
    na, a = _both(g1)
    assert na == a, (na, a)
    assert a[1] > 0.0
    g2 = H.gpu_domain(small)
    g2.evolve_steps(None, 20)
    nb, b = _both(g2)
    assert nb == b, (nb, b)
    # a call on the first domain again after the smaller one
    na2, _ = _both(g1)
    assert na2 == a
    g1.close()
    g2.close()


def test_noalloc_statistics_reject_too_many_rows():
    print("CASE no-alloc statistics: row limit", flush=True)
    from ptha_amd.domain import Domain
    lib = B.load()
    g = Domain(8, 16385, (8000.0, 16385000.0), (0.0, 0.0), "linear", max_U_variables=())
    st = B.DomainStats()
    rc = lib.swh_domain_statistics_noalloc(g._h, ctypes.byref(st))
    assert rc == -1
    buf = ctypes.create_string_buffer(512)
    lib.swh_last_error(buf, 512)
    assert b"16384" in buf.value
    g.close()
