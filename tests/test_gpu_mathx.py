"""GPU: the FP64 friction pow and boundary sine of the kernels (ptha_amd/csrc/swh_mathx.h)
give on the device the same bits as the same functions built for the host (which
tests/test_mathx.py pins to correct rounding), over 2^20 seeded arguments in the
ranges the configurations reach."""
import math

import numpy as np
import pytest

from ptha_amd import backend as B

pytestmark = pytest.mark.gpu


def test_device_math_equals_host_build():
    print("CASE device pow/sin vs host build", flush=True)
    lib = B.load()
    n = 1 << 20
    rng = np.random.default_rng(20261205)
    x = np.exp(rng.uniform(math.log(1e-5), math.log(1e4), n))
    y = np.where(rng.random(n) < 0.5, -7.0 / 3.0, -2.0)
    t = rng.uniform(-1000.0, 1000.0, n)
    out = [np.empty(n) for _ in range(4)]

    def p(a):
        return a.ctypes.data_as(B._PD)

    B.check(lib.swh_probe_math(n, p(x), p(y), p(t), *[p(o) for o in out]), "swh_probe_math")
    pow_dev, sin_dev, pow_host, sin_host = out
    assert np.array_equal(pow_dev, pow_host)
    assert np.array_equal(sin_dev, sin_host)
    # and close to the library functions (they differ only in the last bit, rarely)
    assert np.allclose(pow_dev, x ** y, rtol=3e-16, atol=0)
    assert np.allclose(sin_dev, np.sin(t), rtol=0, atol=3e-16)
