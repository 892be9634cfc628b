"""GPU: the initial guess of the solves inside nyxpm_pm_steps.  The default
(previous phi, as Nyx) and the linear extrapolation 2 phi_n - phi_{n-1} meet the
same stopping rule, so trajectories agree to the solver tolerance; the
extrapolation needs no more V-cycles than Nyx's guess."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _inputs(nyx, n, seed):
    return nyx.ics.plummer_clustered(n, seed, n_spheres=4)


def _run(nyx, n, inp, steps, precision, guess):
    pos, vel, m = inp
    with nyx.PMContext(nyx.BoxGeom((n, n, n)), precision=precision, phi_guess=guess) as c:
        c.upload(0, pos, vel, m_const=m)
        c.pm_init(1.0, 1.0, 1e-10)
        its = []
        for _ in range(steps):
            its.append(int(c.pm_steps(1, 2e-3, rel_tol=1e-10)[0]))
        g = c.gather_accel()
        # accuracy of the final phi: an FP64 solve of the same density from zero
        c.set_precision("fp64")
        c.assign_density(want=False)
        c.solve_for_phi(1.0, 1.0, 1e-12, 0.0, use_guess=False)
        g64 = c.gather_accel()
        out = c.download()
    return out, g, g64, its


@pytest.mark.parametrize("precision", ["mixed", "fp64"])
def test_extrapolated_guess_meets_the_same_rule(nyx, precision):
    n = 48
    inp = _inputs(nyx, n, 41)
    prev = _run(nyx, n, inp, 8, precision, "previous")
    extr = _run(nyx, n, inp, 8, precision, "extrapolate")
    # FP64: the solver tolerance; mixed: its steps deposit with FP32 weights, so
    # against the FP64 density's solve the contract's 1e-6
    bound = 1e-8 if precision == "fp64" else 1e-6
    for out, g, g64, _ in (prev, extr):
        assert np.linalg.norm(g - g64) / np.linalg.norm(g64) <= bound
    # trajectories: FP64 to the solver tolerance; mixed interpolates g in FP32,
    # where phi differing at the solver tolerance can round a corner's g to the
    # neighbouring FP32 value (~6e-8 relative per step)
    tol = 1e-9 if precision == "fp64" else 1e-7
    for d in range(3):
        for q in range(2):
            a, b = prev[0][q][d], extr[0][q][d]
            assert np.abs(a - b).max() <= tol * max(np.abs(b).max(), 1e-300)
    # no more V-cycles than the previous-phi guess over the run (the first step
    # has no phi_{n-1} and uses phi_n)
    assert sum(extr[3]) <= sum(prev[3])


def test_default_guess_is_nyx_guess(nyx):
    # the default is Nyx's guess in both precisions (bit for bit)
    n = 32
    inp = _inputs(nyx, n, 43)
    a = _run(nyx, n, inp, 4, "fp64", "default")
    b = _run(nyx, n, inp, 4, "fp64", "previous")
    c = _run(nyx, n, inp, 4, "mixed", "default")
    d = _run(nyx, n, inp, 4, "mixed", "previous")
    for x, y in ((a, b), (c, d)):
        assert x[3] == y[3]
        for dd in range(3):
            assert np.array_equal(x[0][0][dd], y[0][0][dd])
            assert np.array_equal(x[0][1][dd], y[0][1][dd])


def test_guess_reset_by_set_phi(nyx):
    # after set_phi the next step has no phi_{n-1}: it starts from the phi given,
    # exactly as a fresh context given the same state does
    n = 32
    pos, vel, m = _inputs(nyx, n, 47)
    with nyx.PMContext(nyx.BoxGeom((n, n, n)), precision="mixed", phi_guess="extrapolate") as c:
        c.upload(0, pos, vel, m_const=m)
        c.pm_init(1.0, 1.0, 1e-10)
        c.pm_steps(3, 2e-3, rel_tol=1e-10)
        (x, v), phi = c.download(), c.get_phi()
        c.set_phi(phi)
        c.pm_steps(2, 2e-3, rel_tol=1e-10)
        r1 = c.download()
    with nyx.PMContext(nyx.BoxGeom((n, n, n)), precision="mixed", phi_guess="extrapolate") as c:
        c.upload(0, x, v, m_const=m)
        c.set_phi(phi)
        c.pm_steps(2, 2e-3, rel_tol=1e-10)
        r2 = c.download()
    for d in range(3):
        assert np.array_equal(r1[0][d], r2[0][d])
        assert np.array_equal(r1[1][d], r2[1][d])


@pytest.mark.parametrize("steps", [4, 5])
def test_graph_replay_with_extrapolation(nyx, steps):
    # the guess kernels read and set a device flag, so captured steps extrapolate
    # exactly as eager ones (replays with eager work and synchronisations between)
    n = 32
    pos, vel, m = _inputs(nyx, n, 53)
    outs = []
    for graph in (False, True):
        with nyx.PMContext(nyx.BoxGeom((n, n, n)), precision="mixed", phi_guess="extrapolate") as c:
            c.upload(0, pos, vel, m_const=m)
            c.pm_init(1.0, 1.0, 1e-10)
            its = c.pm_steps(steps, 2e-3, rel_tol=1e-10, use_graph=graph)
            outs.append((c.download(), c.get_phi(), list(its)))
    (a, pa, ia), (b, pb, ib) = outs
    assert ia == ib
    assert np.array_equal(pa, pb)
    for d in range(3):
        assert np.array_equal(a[0][d], b[0][d])
        assert np.array_equal(a[1][d], b[1][d])


@pytest.mark.parametrize("ir_first", [1, 3, 5, 7])
def test_mixed_refinement_schedule_meets_the_rule(nyx, orc, ir_first):
    # mixed precision runs ir_first FP32 V-cycles on one FP64 residual before the
    # first FP64 stopping test, then one per test: every schedule stops on the
    # same FP64 rule (the oracle's residual of the returned phi below 1e-10 max|f|
    # from a zero guess), runs at least min(ir_first, the FP64 count) cycles and
    # gives the FP64 solve's accelerations to the mixed bar.  The 16^3 box runs
    # its whole V-cycle in the one-workgroup coarse kernel, which carries the
    # correction from cycle to cycle the same way.
    for n in (48, 16):
        nn = (n, n, n)
        pos, vel, m = nyx.ics.plummer_clustered(n, 50 + n, n_spheres=4)
        with nyx.PMContext(nyx.BoxGeom(nn), precision="mixed", ir_first=ir_first) as c:
            c.upload(0, pos, vel, m_const=m)
            c.assign_density(want=False)
            it, _ = c.solve_for_phi(1.0, 1.0, 1e-10, 0.0, use_guess=False)
            g = c.gather_accel()
            phi = c.get_phi()
            c.set_precision("fp64")
            it64, _ = c.solve_for_phi(1.0, 1.0, 1e-10, 0.0, use_guess=False)
            g64 = c.gather_accel()
        assert int(it[0]) >= min(ir_first, int(it64[0]))
        assert np.linalg.norm(g - g64) / np.linalg.norm(g64) <= 1e-6
        f = orc.deposit(nn, (0, 0, 0), (1, 1, 1), *pos, m_const=m)
        orc.correct_rhs(nn, (0, 0, 0), (1, 1, 1), f, orc.mass_offset((0, 0, 0), (1, 1, 1), m * n ** 3))
        orc.scale_rhs(f, 1.0, 1.0)
        _, rmax = orc.residual(nn, [1.0 / n] * 3, f, phi)
        # the GPU's solvability sum differs from the oracle's in the last bits
        assert rmax <= 1e-10 * np.abs(f).max() * (1 + 1e-6)
