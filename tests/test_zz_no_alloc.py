"""GPU: nyxpm_gather_accel_noalloc, the entry point that allocates nothing,
against nyxpm_gather_accel through the C ABI: bit for bit (same kernel) on two
calls in a row, the second on a smaller box, into host and device memory; and
its size limit."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _ctx_after_step(nyx, n, seed, precision):
    pos, vel, m = nyx.ics.lattice_gaussian(n, seed, sigma_cells=1.0)
    c = nyx.PMContext(nyx.BoxGeom((n, n, n)), precision=precision)
    c.upload(0, pos, vel, m_const=m)
    c.pm_init(1.0, 1.0, 1e-10)
    c.pm_steps(1, 1e-3, rel_tol=1e-10)
    return c


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_noalloc_gather_equals_gather(nyx, precision):
    import torch
    L = nyx.load()
    for n in (64, 40):  # the second call on a smaller input
        print(f"noalloc gather {precision} {n}^3", flush=True)
        with _ctx_after_step(nyx, n, 7 + n, precision) as c:
            ref = np.full((3, n ** 3), np.nan)
            assert L.nyxpm_gather_accel(c.h, 0, ref.ctypes.data) == 0, L.nyxpm_last_error(c.h)
            got = np.full((3, n ** 3), np.nan)
            rc = L.nyxpm_gather_accel_noalloc(c.h, 0, got.ctypes.data)
            assert rc == 0, L.nyxpm_last_error(c.h)
            assert np.array_equal(got, ref)
            dev = torch.full((3 * n ** 3,), float("nan"), dtype=torch.float64, device="cuda")
            torch.cuda.synchronize()
            rc = L.nyxpm_gather_accel_noalloc(c.h, 0, ctypes.c_void_p(dev.data_ptr()))
            assert rc == 0, L.nyxpm_last_error(c.h)
            assert np.array_equal(dev.cpu().numpy().reshape(3, -1), ref)


def test_noalloc_gather_rejects_boxes_above_its_array(nyx):
    L = nyx.load()
    n = 128  # 2^21 particles, above the 2^20 the staging array takes
    pos, vel, m = nyx.ics.lattice_gaussian(n, 3, sigma_cells=0.5)
    with nyx.PMContext(nyx.BoxGeom((n, n, n))) as c:
        c.upload(0, pos, vel, m_const=m)
        c.pm_init(1.0, 1.0, 1e-8)
        out = np.zeros((3, n ** 3))
        assert L.nyxpm_gather_accel_noalloc(c.h, 0, out.ctypes.data) == nyx._lib.NYXPM_ERR_ARG
        assert b"above the 1048576" in L.nyxpm_last_error(c.h)
        assert not out.any()
