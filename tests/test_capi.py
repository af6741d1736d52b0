"""The C-ABI library loads and exports every function that include/swals_hip.h
declares (no device calls: this runs without a GPU)."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    hdr = open(os.path.join(ROOT, "include", "swals_hip.h")).read()
    return sorted(set(re.findall(r"^int (swh_\w+)\(", hdr, flags=re.M)))


def test_library_exports_every_declared_symbol():
    path = os.path.join(ROOT, "ptha_amd", "libswals_hip.so")
    if not os.path.exists(path):
        pytest.fail("libswals_hip.so is not built (run __graft_entry__.build())")
    lib = ctypes.CDLL(path)
    names = _declared()
    assert len(names) >= 30
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_python_binding_declares_every_symbol():
    from ptha_amd import backend
    bound = {n for n, _, _ in backend.SYMBOLS}
    assert set(_declared()) <= bound, set(_declared()) - bound


def test_error_reporting_without_device():
    """Argument checks fail before any device call, with a message."""
    from ptha_amd import backend as B
    lib = B.load()
    h = ctypes.c_void_p()
    d = B.DomainDesc()
    d.nx, d.ny = 2, 2
    rc = lib.swh_domain_create(ctypes.byref(d), ctypes.byref(h))
    assert rc == -1
    buf = ctypes.create_string_buffer(256)
    lib.swh_last_error(buf, 256)
    assert b"nx and ny" in buf.value


def test_argument_checks_of_md_batch_and_probe_without_device():
    """The multidomain, batch-tuning and copy-probe entry points reject bad arguments
    on the host, before any device call."""
    from ptha_amd import backend as B
    lib = B.load()
    ms = ctypes.c_double()
    assert lib.swh_probe_copy(15, 1, ctypes.byref(ms)) == -1          # not a multiple of 16
    assert lib.swh_probe_copy(1024, 0, ctypes.byref(ms)) == -1        # no repetitions
    h = ctypes.c_void_p()
    assert lib.swh_md_create(0, None, None, None, None, ctypes.byref(h)) == -1
    assert lib.swh_md_finalize(None) == -1
    assert lib.swh_md_evolve_one_step(None, 1.0) != 0
    assert lib.swh_batch_autotune(None, 1.0, 1, None, None) == -1
    assert lib.swh_batch_superpose(None, 1, None, None, None, None, 0.0, 0, None, None, None) == -1


def test_batch_rejects_llpnf_without_truely_linear_solver():
    """A 'leapfrog_linear_plus_nonlinear_friction' batch needs the truely-linear solver
    (its friction work is precomputed from the elevation): rejected on the host with a
    message naming the flag, before any device call."""
    from ptha_amd import backend as B
    lib = B.load()
    d = B.DomainDesc()
    d.nx, d.ny = 64, 64
    d.timestepping_method = 6   # SWH_METHOD_LLPNF
    d.linear_solver_is_truely_linear = 0
    h = ctypes.c_void_p()
    assert lib.swh_batch_create(ctypes.byref(d), 4, ctypes.byref(h)) == -1
    buf = ctypes.create_string_buffer(256)
    lib.swh_last_error(buf, 256)
    assert b"linear_solver_is_truely_linear" in buf.value
    assert not h.value
    n = ctypes.c_int64()
    assert lib.swh_batch_launch_counts(None, ctypes.byref(n), ctypes.byref(n), ctypes.byref(n)) == -1


def test_md_graph_entry_points_check_arguments_without_device():
    from ptha_amd import backend as B
    lib = B.load()
    assert lib.swh_md_graph_capture(None, 1.0) != 0
    assert lib.swh_md_graph_launch(None, 1) != 0
    assert lib.swh_md_graph_destroy(None) == -1
