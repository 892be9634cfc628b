"""GPU: batches of independent boxes in one launch per phase, and hipGraph-
captured steps, against solo eager runs (bit for bit) and the oracle."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

SIZES = [(16, 16, 16), (24, 24, 24), (32, 32, 32)]


def _inputs(nyx, n, seed, mass=False):
    pos, vel, m = nyx.ics.grf_lattice(n[0], seed, -2.0, 0.5, 0.1)
    if mass:
        r = np.random.default_rng(seed)
        mm = m * r.uniform(0.5, 1.5, len(pos[0]))
        return pos, vel, mm, None
    return pos, vel, None, m


def _upload(c, box, inp):
    pos, vel, mm, m = inp
    c.upload(box, pos, vel, m=mm, m_const=m if m is not None else 0.0)


def _run(nyx, geoms, inputs, steps, precision="fp64", use_graph=False):
    with nyx.PMContext([nyx.BoxGeom(g) for g in geoms], precision=precision) as c:
        for b, inp in enumerate(inputs):
            _upload(c, b, inp)
        c.pm_init(1.0, 1.0, 1e-10)
        its = c.pm_steps(steps, 1e-3, rel_tol=1e-10, use_graph=use_graph)
        out = [c.download(b) for b in range(len(geoms))]
        phis = [c.get_phi(b) for b in range(len(geoms))]
    return out, phis, its


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_batch_equals_solo(nyx, precision):
    inputs = [_inputs(nyx, n, 100 + i, mass=(i == 1)) for i, n in enumerate(SIZES)]
    batch, bphi, bits = _run(nyx, SIZES, inputs, 2, precision)
    for b, n in enumerate(SIZES):
        solo, sphi, sits = _run(nyx, [n], [inputs[b]], 2, precision)
        assert bits[b] == sits[0]
        assert np.array_equal(bphi[b], sphi[0])
        for d in range(3):
            assert np.array_equal(batch[b][0][d], solo[0][0][d])
            assert np.array_equal(batch[b][1][d], solo[0][1][d])


def test_batch_vs_oracle(nyx, orc):
    inputs = [_inputs(nyx, n, 200 + i) for i, n in enumerate(SIZES)]
    batch, _, _ = _run(nyx, SIZES, inputs, 1)
    for b, n in enumerate(SIZES):
        pos, vel, _, m = inputs[b]
        p = [a.copy() for a in pos]
        v = [a.copy() for a in vel]
        phi, _, _ = orc.solve_for_phi(n, (0, 0, 0), (1, 1, 1), p, m_const=m, rel_tol=1e-10)
        orc.pm_step(n, (0, 0, 0), (1, 1, 1), p, v, phi, m_const=m, dt=1e-3, rel_tol=1e-10)
        for d in range(3):
            assert np.abs(batch[b][0][d] - p[d]).max() <= 1e-12 * np.abs(p[d]).max()
            assert np.abs(batch[b][1][d] - v[d]).max() <= 1e-12 * np.abs(v[d]).max()


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
@pytest.mark.parametrize("steps", [1, 4, 5])
def test_graph_replay_equals_eager(nyx, precision, steps):
    # 4 and 5 steps replay the captured two-step graph twice, with eager work and
    # synchronisations (convergence checks, the odd last step) in between
    inputs = [_inputs(nyx, n, 300 + i) for i, n in enumerate(SIZES)]
    eager, ephi, eits = _run(nyx, SIZES, inputs, steps, precision, use_graph=False)
    graph, gphi, gits = _run(nyx, SIZES, inputs, steps, precision, use_graph=True)
    assert list(eits) == list(gits)
    for b in range(len(SIZES)):
        assert np.array_equal(ephi[b], gphi[b])
        for d in range(3):
            assert np.array_equal(eager[b][0][d], graph[b][0][d])
            assert np.array_equal(eager[b][1][d], graph[b][1][d])


def test_precision_switch_in_one_context(nyx):
    pos, vel, _, m = _inputs(nyx, (32, 32, 32), 7)
    with nyx.PMContext(nyx.BoxGeom((32, 32, 32))) as c:
        c.upload(0, pos, vel, m_const=m)
        c.pm_init(1.0, 1.0, 1e-10)
        g64 = c.gather_accel()
        c.set_precision("mixed")
        it, _ = c.solve_for_phi(1.0, 1.0, 1e-10, 0.0, use_guess=False)
        gmx = c.gather_accel()
        c.set_precision("fp64")
        c.solve_for_phi(1.0, 1.0, 1e-10, 0.0, use_guess=False)
        g64b = c.gather_accel()
    assert np.array_equal(g64, g64b)
    rel = np.sqrt(np.sum((gmx - g64) ** 2) / np.sum(g64 ** 2))
    assert rel < 1e-6


def test_autotune_keeps_results(nyx):
    # the autotuner only changes launch shapes; results stay bit for bit
    sizes = [(32, 32, 32), (64, 64, 64), (96, 96, 96)]  # fused sweeps on the two finest depths
    inputs = [_inputs(nyx, n, 400 + i) for i, n in enumerate(sizes)]
    outs = []
    for tune in (False, True):
        with nyx.PMContext([nyx.BoxGeom(g) for g in sizes], precision="mixed") as c:
            if tune:
                rep = c.autotune()
                assert "fp32.d0" in rep and "fp64.d0" in rep and all(v["us"] > 0 for v in rep.values())
            for b, inp in enumerate(inputs):
                _upload(c, b, inp)
            c.pm_init(1.0, 1.0, 1e-10)
            c.pm_steps(2, 1e-3, rel_tol=1e-10, use_graph=True)
            outs.append([c.download(b) for b in range(len(sizes))])
    for b in range(len(sizes)):
        for d in range(3):
            assert np.array_equal(outs[0][b][0][d], outs[1][b][0][d])
            assert np.array_equal(outs[0][b][1][d], outs[1][b][1][d])


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_particle_tile_autotune_keeps_state(nyx, precision):
    # the particle tile re-chosen in the middle of a run: the particles go out to
    # a copy and come back in upload order, so positions, velocities and phi are
    # bit for bit those before it; the steps after it agree with the run that kept
    # its tile to rounding (the fold adds a boundary cell's parts from the tiles
    # that hold them, in FP64, so the density's last bit depends on the tiling);
    # one box with per-particle masses
    sizes = [(32, 32, 32), (48, 32, 24)]
    inputs = [_inputs(nyx, sizes[0], 500), _inputs(nyx, sizes[1], 501, mass=True)]
    outs = []
    for tune in (False, True):
        with nyx.PMContext([nyx.BoxGeom(g) for g in sizes], precision=precision) as c:
            for b, inp in enumerate(inputs):
                _upload(c, b, inp)
            c.pm_init(1.0, 1.0, 1e-10)
            c.pm_steps(2, 1e-3, rel_tol=1e-10)
            if tune:
                before = ([c.download(b) for b in range(len(sizes))], [c.get_phi(b) for b in range(len(sizes))])
                rep = c.autotune(fp64=False, fp32=False)
                assert set(rep) == {"tile"} and rep["tile"]["us"] > 0 and len(rep["tile"]["shape"]) == 3
                for b in range(len(sizes)):
                    assert np.array_equal(before[1][b], c.get_phi(b))
                    pa = c.download(b)
                    for d in range(3):
                        assert np.array_equal(before[0][b][0][d], pa[0][d])
                        assert np.array_equal(before[0][b][1][d], pa[1][d])
                assert c.sum_particle_mass(1) == pytest.approx(inputs[1][2].sum(), rel=1e-15)
            its = c.pm_steps(2, 1e-3, rel_tol=1e-10)
            outs.append(([c.download(b) for b in range(len(sizes))], [c.get_phi(b) for b in range(len(sizes))],
                         [int(v) for v in its]))
    assert outs[0][2] == outs[1][2]
    for b in range(len(sizes)):
        phi0, phi1 = outs[0][1][b], outs[1][1][b]
        assert np.abs(phi0 - phi1).max() <= 1e-9 * np.abs(phi0).max()
        for d in range(3):
            assert np.abs(outs[0][0][b][0][d] - outs[1][0][b][0][d]).max() <= 1e-12
            v0, v1 = outs[0][0][b][1][d], outs[1][0][b][1][d]
            assert np.abs(v0 - v1).max() <= 1e-9 * np.abs(v0).max()


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
@pytest.mark.parametrize("steps,use_graph", [(1, False), (3, False), (6, True)])
def test_pair_kicks_equal_separate(nyx, precision, steps, use_graph):
    # the closing kick of a step fused with the next step's kick-drift gives the
    # same bits as the two passes; fast particles cross the periodic boundary, so
    # the re-interpolation after the shift is exercised
    # cubic cells in both boxes (the second is 0.75 x 1.25 x 0.5)
    geoms = [nyx.BoxGeom((32, 32, 32)), nyx.BoxGeom((24, 40, 16), (0.0, 0.0, 0.0), (0.75, 1.25, 0.5))]
    inputs = []
    for i, g in enumerate(geoms):
        r = np.random.default_rng(500 + i)
        npart = g.ncell
        pos = [r.uniform(g.prob_lo[d], g.prob_hi[d], npart) for d in range(3)]
        vel = [r.normal(0.0, 8.0, npart) for _ in range(3)]
        inputs.append((pos, vel, None, 1.0 / npart))
    outs = []
    for pair in (False, True):
        with nyx.PMContext(geoms, precision=precision, pair_kicks=pair) as c:
            for b, inp in enumerate(inputs):
                _upload(c, b, inp)
            c.pm_init(1.0, 1.0, 1e-10)
            its = c.pm_steps(steps, 1e-3, rel_tol=1e-10, use_graph=use_graph)
            outs.append(([c.download(b) for b in range(2)], [c.get_phi(b) for b in range(2)], its))
    (sep, sphi, sits), (fus, fphi, fits) = outs
    assert list(sits) == list(fits)
    for b in range(2):
        assert np.array_equal(sphi[b], fphi[b])
        for d in range(3):
            assert np.array_equal(sep[b][0][d], fus[b][0][d])
            assert np.array_equal(sep[b][1][d], fus[b][1][d])


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_coarse_workgroup_equals_per_level(nyx, precision):
    # the one-workgroup coarse V-cycle (levels of <= 4096 cells down to the bottom)
    # gives the same bits as one launch per level; 16^3 runs its whole cycle there,
    # 48x24x40 reaches it at 12x6x10 above an odd 6x3x5 bottom
    geoms = [nyx.BoxGeom((16, 16, 16)), nyx.BoxGeom((32, 32, 32)),
             nyx.BoxGeom((48, 24, 40), (0.0, 0.0, 0.0), (1.0, 0.5, 40 / 48)), nyx.BoxGeom((96, 96, 96))]
    inputs = []
    for i, g in enumerate(geoms):
        r = np.random.default_rng(600 + i)
        npart = g.ncell
        pos = [r.uniform(g.prob_lo[d], g.prob_hi[d], npart) for d in range(3)]
        vel = [r.normal(0.0, 0.5, npart) for _ in range(3)]
        inputs.append((pos, vel, None, 1.0 / npart))
    outs = []
    for coarse in ("per_level", "workgroup"):
        with nyx.PMContext(geoms, precision=precision, coarse=coarse) as c:
            for b, inp in enumerate(inputs):
                _upload(c, b, inp)
            c.pm_init(1.0, 1.0, 1e-10)
            its = c.pm_steps(2, 1e-3, rel_tol=1e-10)
            outs.append(([c.download(b) for b in range(len(geoms))], [c.get_phi(b) for b in range(len(geoms))],
                         list(its)))
    (a, aphi, ait), (b_, bphi, bit) = outs
    assert ait == bit
    for b in range(len(geoms)):
        assert np.array_equal(aphi[b], bphi[b])
        for d in range(3):
            assert np.array_equal(a[b][0][d], b_[b][0][d])
            assert np.array_equal(a[b][1][d], b_[b][1][d])


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
@pytest.mark.parametrize("use_graph", [False, True])
def test_est_time_step_fused_equals_standalone(nyx, precision, use_graph):
    # the estimate the closing kicks keep (particle_cfl) equals a separate pass over
    # the same state and leaves the trajectory unchanged; fast particles cross the
    # periodic boundary (deferred paths); a box without particles gives -1e50
    geoms = [nyx.BoxGeom((32, 32, 32)), nyx.BoxGeom((24, 40, 16), (0.0, 0.0, 0.0), (0.75, 1.25, 0.5)),
             nyx.BoxGeom((16, 16, 16))]
    inputs = []
    for i, g in enumerate(geoms[:2]):
        r = np.random.default_rng(700 + i)
        npart = g.ncell
        pos = [r.uniform(g.prob_lo[d], g.prob_hi[d], npart) for d in range(3)]
        vel = [r.normal(0.0, 8.0, npart) for _ in range(3)]
        inputs.append((pos, vel, None, 1.0 / npart))
    res = []
    for cfl in (0.3, 0.0):
        with nyx.PMContext(geoms, precision=precision, particle_cfl=cfl) as c:
            for b, inp in enumerate(inputs):
                _upload(c, b, inp)
            c.pm_init(1.0, 1.0, 1e-10)
            c.pm_steps(5, 1e-3, rel_tol=1e-10, use_graph=use_graph)
            dt = c.est_time_step(1.0, 0.3)
            c.move_kick_drift(1e-3, 1.0, 1.0)
            c.move_kick(1e-3, 1.0, 1.0)
            dt2 = c.est_time_step(1.0, 0.3)
            res.append((dt, dt2, [c.download(b) for b in range(2)]))
    (fa, fb, fp), (sa, sb, sp) = res
    assert np.array_equal(fa, sa) and np.array_equal(fb, sb)
    assert fa[2] == -1e50 and 0 < fa[0] < 1e50 and 0 < fa[1] < 1e50
    for b in range(2):
        for d in range(3):
            assert np.array_equal(fp[b][0][d], sp[b][0][d])
            assert np.array_equal(fp[b][1][d], sp[b][1][d])


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
@pytest.mark.parametrize("use_graph", [False, True])
def test_residual_in_first_sweep_equals_separate_pass(nyx, precision, use_graph):
    # the level-0 residual formed inside the first sweep of the next V-cycle (with
    # its norms) gives the same bits as k_residual0 followed by that sweep
    geoms = [nyx.BoxGeom((32, 32, 32)), nyx.BoxGeom((64, 64, 64)), nyx.BoxGeom((96, 96, 96))]
    inputs = [_inputs(nyx, g.n_cell, 700 + i, mass=(i == 1)) for i, g in enumerate(geoms)]
    outs = []
    for residual in ("separate", "in_sweep"):
        with nyx.PMContext(geoms, precision=precision, residual=residual) as c:
            for b, inp in enumerate(inputs):
                _upload(c, b, inp)
            c.pm_init(1.0, 1.0, 1e-10)
            its = c.pm_steps(5, 1e-3, rel_tol=1e-10, use_graph=use_graph)
            outs.append(([c.download(b) for b in range(len(geoms))], [c.get_phi(b) for b in range(len(geoms))],
                         list(its)))
    (a, aphi, ait), (b_, bphi, bit) = outs
    assert ait == bit
    for b in range(len(geoms)):
        assert np.array_equal(aphi[b], bphi[b])
        for d in range(3):
            assert np.array_equal(a[b][0][d], b_[b][0][d])
            assert np.array_equal(a[b][1][d], b_[b][1][d])
