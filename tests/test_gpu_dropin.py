"""The C++ drop-in layer (include/ompblas_dropin.hpp: amt::mtm_helper,
amt::mtv_helper, amt::dot_prod_helper with the reference's signatures) run
from its C++ test program, which mirrors the reference's Catch2 suites
(test/test.mtm.cpp, test/test.mtv.cpp, test/test.dot_prod.cpp) with
integer brute force as the oracle, on host and on device pointers."""
import os
import subprocess

import pytest

import helpers

pytestmark = pytest.mark.gpu

BIN = os.path.join(os.path.dirname(helpers.pkg().LIB_PATH), "test_dropin")


def _run(args, timeout):
    assert os.path.exists(BIN), "build the package first (make in openmp-blas_amd)"
    out = subprocess.run([BIN, *args], capture_output=True, text=True, timeout=timeout)
    print(out.stdout[-4000:], out.stderr[-2000:])
    return out


def test_dropin_quick(dev):
    out = _run(["--quick"], 120)
    assert out.returncode == 0
    assert "PASSED" in out.stdout


def test_dropin_full_reference_sizes(dev):
    """Every size of the reference suites, plus the 32 Ki x 32 Ki mtv cases."""
    out = _run([], 300)
    assert out.returncode == 0
    assert "PASSED" in out.stdout
