"""GPU: Nyx's binary particle file (InitFromBinaryFile) through the C ABI.

tests/golden/32.nyx is the reference's Exec/LyA/32.nyx, unchanged (32^3 particles
in an 8 Mpc box: int64 NP, int32 DM = 3, int32 NX = 4, then x y z m vx vy vz in
FP64 per particle).  numpy reads the same bytes independently as the check."""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
FILE = os.path.join(HERE, "golden", "32.nyx")


def _numpy_read(path, dtype=np.float64):
    with open(path, "rb") as fh:
        np_ = int(np.frombuffer(fh.read(8), np.int64)[0])
        dm, nx = (int(v) for v in np.frombuffer(fh.read(8), np.int32))
        a = np.frombuffer(fh.read(), dtype).reshape(np_, dm + nx).astype(np.float64)
    return a


def _write(path, a, dtype, nx=4):
    with open(path, "wb") as fh:
        fh.write(np.int64(a.shape[0]).tobytes())
        fh.write(np.array([3, nx], np.int32).tobytes())
        fh.write(np.ascontiguousarray(a, dtype=dtype).tobytes())


def test_load_32nyx_matches_file(nyx):
    a = _numpy_read(FILE)
    assert a.shape == (32768, 7)
    with nyx.PMContext(nyx.BoxGeom((32, 32, 32), (0, 0, 0), (8, 8, 8))) as c:
        pc = nyx.DarkMatterParticleContainer(c)
        assert pc.InitFromBinaryFile(FILE, 4) == 32768
        p, v = c.download()
        for d in range(3):
            assert np.array_equal(p[d], a[:, d])
            assert np.array_equal(v[d], a[:, 4 + d])
        assert c.sum_particle_mass() == pytest.approx(a[:, 3].sum(), rel=1e-15)


def test_load_fp32_and_extra_columns(nyx, tmp_path):
    r = np.random.default_rng(3)
    a = np.column_stack([r.random((500, 3)) * 8, np.full(500, 2.5), r.normal(0, 1, (500, 3)),
                         r.normal(0, 1, (500, 2))])
    f32 = tmp_path / "p32.nyx"
    _write(f32, a, np.float32, nx=6)  # two extra columns after the velocity are skipped
    with nyx.PMContext(nyx.BoxGeom((16, 16, 16), (0, 0, 0), (8, 8, 8))) as c:
        assert c.load_binary(f32) == 500
        p, v = c.download()
        ref = a.astype(np.float32).astype(np.float64)
        for d in range(3):
            assert np.array_equal(p[d], ref[:, d])
            assert np.array_equal(v[d], ref[:, 4 + d])


def test_load_errors(nyx, tmp_path):
    bad = tmp_path / "bad.nyx"
    with open(bad, "wb") as fh:
        fh.write(np.int64(10).tobytes() + np.array([3, 4], np.int32).tobytes() + b"\0" * 100)
    with nyx.PMContext(nyx.BoxGeom((16, 16, 16))) as c:
        with pytest.raises(nyx.NyxPMError):
            c.load_binary(bad)
        with pytest.raises(nyx.NyxPMError):
            c.load_binary(tmp_path / "missing.nyx")
        two_d = tmp_path / "dm2.nyx"
        with open(two_d, "wb") as fh:
            fh.write(np.int64(1).tobytes() + np.array([2, 4], np.int32).tobytes() + b"\0" * 48)
        with pytest.raises(nyx.NyxPMError):
            c.load_binary(two_d)


def test_32nyx_deposit_and_step_vs_oracle(nyx, orc):
    # the LyA particle set, per-particle masses, one canonical KDK step with
    # 4 pi G = 1e-10 (so that the step stays small in these units) against the oracle
    a = _numpy_read(FILE)
    n, hi = (32, 32, 32), (8.0, 8.0, 8.0)
    pos = [a[:, d].copy() for d in range(3)]
    vel = [a[:, 4 + d].copy() for d in range(3)]
    m = a[:, 3].copy()
    with nyx.PMContext(nyx.BoxGeom(n, (0, 0, 0), hi)) as c:
        c.load_binary(FILE)
        rho = c.assign_density()
        ref = orc.deposit(n, (0, 0, 0), hi, pos[0], pos[1], pos[2], m=m)
        assert np.abs(rho - ref).max() <= 1e-13 * np.abs(ref).max()
        c.pm_init(1.0, 1e-10, 1e-10)
        c.pm_steps(1, 1e-3, 1.0, 1.0, four_pi_G=1e-10, rel_tol=1e-10)
        p, v = c.download()
    phi, _, _ = orc.solve_for_phi(n, (0, 0, 0), hi, pos, m=m, four_pi_G=1e-10, rel_tol=1e-10)
    orc.pm_step(n, (0, 0, 0), hi, pos, vel, phi, m=m, four_pi_G=1e-10, dt=1e-3, rel_tol=1e-10)
    for d in range(3):
        assert np.abs(p[d] - pos[d]).max() <= 1e-12 * 8
        assert np.abs(v[d] - vel[d]).max() <= 1e-12 * np.abs(vel[d]).max()
