"""GPU: the Fortran boundary. A Fortran program sets up the reference's own
domain_type (compiled from the reference sources), passes c_loc of its metric arrays
and U to the backend through swals_hip_mod, advances the domain on the GPU, and
compares with domain%evolve_one_step on the CPU: FP64 must be bit-identical."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "oracle", "_ref", "dropin_driver_sph")

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("method", ["linear", "llpnf"])
def test_fortran_dropin_bitwise(method):
    print("CASE fortran dropin", method, flush=True)
    if not os.path.exists(EXE):
        pytest.skip("Fortran drop-in driver not built (needs the reference sources at build time)")
    r = subprocess.run([EXE, method, "200"], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr)
    assert r.returncode == 0
    assert "BITWISE_EQUAL" in r.stdout
