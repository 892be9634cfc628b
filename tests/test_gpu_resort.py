"""GPU: the incremental re-sort (tiles with slack; after a kick-drift only the
particles that changed tile move) against a full sort after every kick-drift,
bit for bit, including the fall-back to a full sort when a tile's slack runs
out or the emptied slots pass their limit."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _hot(nyx, n, seed, vscale, mass=False):
    """Lattice GRF particles plus random velocities of rms vscale (box units per
    unit time): at dt = 1e-3 and tiles of 8-12 cells a large share changes tile per step."""
    pos, vel, m = nyx.ics.grf_lattice(n, seed, -2.0, 0.5, 0.1)
    r = np.random.default_rng(seed)
    vel = [v + vscale * r.standard_normal(len(v)) for v in vel]
    mm = m * r.uniform(0.5, 1.5, len(pos[0])) if mass else None
    return pos, vel, mm, m


def _run(nyx, geoms, inputs, steps, slack, precision="fp64", pair=True):
    with nyx.PMContext([nyx.BoxGeom(g) for g in geoms], precision=precision, resort_slack=slack,
                       pair_kicks=pair) as c:
        for b, (pos, vel, mm, m) in enumerate(inputs):
            if pos is not None:
                c.upload(b, pos, vel, m=mm, m_const=0.0 if mm is not None else m)
        c.pm_init(1.0, 1.0, 1e-10)
        c.timing_reset()
        c.timing_enable(True)
        its = c.pm_steps(steps, 1e-3, rel_tol=1e-10)
        c.sync()
        c.timing_enable(False)
        counts = {k: c.timing(k)[0] for k in ("migrate", "scatter")}
        out = [c.download(b) for b in range(len(geoms))]
        phi = [c.get_phi(b) for b in range(len(geoms))]
    return out, phi, list(its), counts


def _same(a, b):
    for box in range(len(a[0])):
        assert np.array_equal(a[1][box], b[1][box])
        for d in range(3):
            assert np.array_equal(a[0][box][0][d], b[0][box][0][d])
            assert np.array_equal(a[0][box][1][d], b[0][box][1][d])
    assert a[2] == b[2]


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
@pytest.mark.parametrize("pair", [True, False])
def test_incremental_equals_full_sort(nyx, precision, pair):
    n = (32, 32, 32)
    inp = [_hot(nyx, 32, 11, 0.5)]  # ~0.6% change tile per step: no slot limit reached
    full = _run(nyx, [n], inp, 6, -1.0, precision, pair)
    inc = _run(nyx, [n], inp, 6, 0.0, precision, pair)
    _same(full, inc)
    # default slack: every step re-sorted incrementally after the first full sort
    assert full[3]["migrate"] == 0 and full[3]["scatter"] >= 6
    assert inc[3]["migrate"] >= 5 and inc[3]["scatter"] <= 1


def test_slack_overflow_falls_back_to_full_sort(nyx):
    # 1/1024 slack per particle (plus 16 slots per tile) cannot take a step of
    # fast particles: k_migrate reports the overflow and a full sort follows
    n = (32, 32, 32)
    inp = [_hot(nyx, 32, 12, 20.0, mass=True)]
    full = _run(nyx, [n], inp, 4, -1.0)
    tight = _run(nyx, [n], inp, 4, 1.0 / 1024)
    _same(full, tight)
    assert tight[3]["migrate"] >= 3 and tight[3]["scatter"] >= 3


def test_incremental_batch_with_empty_box(nyx):
    geoms = [(16, 16, 16), (32, 32, 32), (24, 24, 24)]
    inputs = [_hot(nyx, 16, 21, 3.0), (None, None, None, None), _hot(nyx, 24, 23, 3.0, mass=True)]
    full = _run(nyx, geoms, inputs, 5, -1.0)
    inc = _run(nyx, geoms, inputs, 5, 0.0)
    _same(full, inc)
    assert inc[3]["migrate"] >= 4


def test_calls_between_kick_drift_and_deposit(nyx):
    # after move_kick_drift the tile changes wait in a list; downloading,
    # gathering and estimating dt see the particles, the deposit re-sorts
    inp = _hot(nyx, 32, 31, 5.0)
    outs = []
    for slack in (-1.0, 0.0):
        with nyx.PMContext(nyx.BoxGeom((32, 32, 32)), resort_slack=slack) as c:
            c.upload(0, inp[0], inp[1], m_const=inp[3])
            c.pm_init(1.0, 1.0, 1e-10)
            c.move_kick_drift(1e-3)
            c.assign_density(0)
            c.solve_for_phi(1.0, 1.0, 1e-10)
            c.move_kick(1e-3)
            c.move_kick_drift(1e-3)
            d1 = c.download(0)
            g1 = c.gather_accel(0)
            dt1 = c.est_time_step(1.0, 0.5)
            rho = c.assign_density(0)
            c.move_kick_drift(1e-3)  # kick-drift twice in a row
            c.move_kick_drift(1e-3)
            rho2 = c.assign_density(0)
            d2 = c.download(0)
            outs.append((d1, g1, dt1, rho, rho2, d2))
    a, b = outs
    for q in (0, 5):
        for d in range(3):
            assert np.array_equal(a[q][0][d], b[q][0][d])
            assert np.array_equal(a[q][1][d], b[q][1][d])
    assert np.array_equal(a[1], b[1])
    assert np.array_equal(np.asarray(a[2]), np.asarray(b[2]))
    assert np.array_equal(a[3], b[3])
    assert np.array_equal(a[4], b[4])


def test_timing_select_times_only_the_named_stages(nyx):
    # stage timing restricted to one stage: its launches are counted, the other
    # stages' are not, and the results are those of an untimed run
    inp = _hot(nyx, 32, 41, 0.5)
    outs = []
    for sel in (None, "deposit", ""):
        with nyx.PMContext(nyx.BoxGeom((32, 32, 32))) as c:
            c.upload(0, inp[0], inp[1], m_const=inp[3])
            c.pm_init(1.0, 1.0, 1e-10)
            c.timing_select(sel if sel != "" else None)
            c.timing_enable(sel is not None)
            c.pm_steps(3, 1e-3, rel_tol=1e-10)
            c.sync()
            c.timing_enable(False)
            counts = {k: c.timing(k)[0] for k in ("deposit", "kick_pair", "residual")}
            outs.append((c.download(0), counts))
    (a, ca), (b, cb), (d, cd) = outs
    assert ca == {"deposit": 0, "kick_pair": 0, "residual": 0}
    assert cb["deposit"] >= 3 and cb["kick_pair"] == 0 and cb["residual"] == 0
    assert cd["deposit"] >= 3 and cd["kick_pair"] >= 2 and cd["residual"] >= 3
    for q in (b, d):
        for d_ in range(3):
            assert np.array_equal(a[0][d_], q[0][d_])
            assert np.array_equal(a[1][d_], q[1][d_])
