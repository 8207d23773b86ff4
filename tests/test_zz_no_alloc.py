"""The allocation-free force entry point against fgx_disk_force, through the
C ABI: bit-identical results on two calls in a row (the second on a smaller
disk, so stale scratch past its rings would show), and a clear refusal of a
disk larger than the scratch."""
import ctypes

import numpy as np
import pytest

from cudaprotoplanetarydisksfloat_amd import F32, F64, Batch, configs
from cudaprotoplanetarydisksfloat_amd.fgx import lib

pytestmark = pytest.mark.gpu

FGX_ERANGE = -6


def _call(fn, disk, x, y, sm, m):
    out = (ctypes.c_double * 44)()
    rc = fn(disk.h, x, y, sm, m, out)
    return rc, np.array(out[:])


@pytest.mark.parametrize("precision", [F32, F64], ids=["fp32", "fp64"])
def test_noalloc_force_equals_force(gpu, precision):
    big = Batch(*configs.load(*configs.plain_disk(nrad=64, nsec=192)), precision)
    big.step(2)
    small = Batch(*configs.load(*configs.plain_disk(nrad=24, nsec=96, noise=1e-2)), precision)
    for disk, pt in ((big, (1.0, 0.1, 0.03, 1e-3)), (small, (0.8, -0.3, 0.02, 5e-4)), (big, (0.0, 0.0, 0.0, 0.0))):
        rc, ref = _call(lib.fgx_disk_force, disk, *pt)
        rc2, got = _call(lib.fgx_disk_force_noalloc, disk, *pt)
        assert rc == 0 and rc2 == 0, (rc, rc2)
        assert np.any(ref != 0)
        assert np.array_equal(ref, got), np.max(np.abs(ref - got))


def test_noalloc_force_refuses_a_disk_past_its_scratch(gpu):
    disk = Batch(*configs.load(*configs.plain_disk(nrad=4097, nsec=64)), F32)
    rc, _ = _call(lib.fgx_disk_force_noalloc, disk, 1.0, 0.0, 0.03, 1e-3)
    assert rc == FGX_ERANGE
    assert b"4096 rings" in lib.fgx_last_error()
    # the allocating entry point still serves it
    rc, ref = _call(lib.fgx_disk_force, disk, 1.0, 0.0, 0.03, 1e-3)
    assert rc == 0 and np.any(ref != 0)
