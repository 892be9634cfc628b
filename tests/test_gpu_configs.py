"""GPU: the mixed-precision workloads, Nyx's two-solve step and captured steps,
against the oracle and against eager runs.

* The 512^3 clustered workload's recipe (20% uniform, 80% in Plummer spheres
  with scale radii log-uniform on 1-8 cells, sigma_v 0.01, dt 1e-3, rel tol
  1e-10) in mixed precision: at 64^3 against the oracle's FP64 steps, and at
  full size (512^3 cells and particles) through size-independent properties:
  mixed vs FP64 accelerations within 1e-6 relative L2 after each of three
  steps, momentum conservation at the solver tolerance, mass conservation and
  a bitwise repeatable deposit.
* The 1024^3 GRF workload (P(k) ~ k^-1.5, rms 1 cell, v = 0.1 psi, dt 5e-4) at
  full size: mixed vs FP64 accelerations within 1e-6 relative L2 after each of
  its first steps.
* Nyx's step that re-solves the old phi from a zero guess
  (Source/TimeStep/Nyx_advance_particles.cpp:148-170) against the oracle's.
* A mixed-precision batch of boxes against the oracle.
* Captured (hipGraph) steps: calls that leave the other particle buffer live,
  and a solve of a replay that does not converge within the recorded cycles.

Tolerances: FP64 paths 1e-12 relative (positions, velocities, accelerations);
mixed paths 1e-6 relative L2 in the accelerations against FP64, which bounds
the velocity change of a step by dt * 1e-6 * |g| per component and the
position change by dt times that.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def relmax(a, b):
    return np.abs(np.asarray(a) - np.asarray(b)).max() / max(np.abs(np.asarray(b)).max(), 1e-300)


def rel_l2(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return float(np.sqrt(np.sum((a - b) ** 2) / np.sum(b ** 2)))


def oracle_accel(orc, n, pos, g):
    """The oracle's interpolation of the gravity field g at positions pos:
    update_particles' kick with a_prev = 0, a_cur = 1 and dt = 2 sets v to the
    interpolated value exactly (0.5 * 2 = 1, times val, times 1/1)."""
    p = [np.ascontiguousarray(a, dtype=np.float64).copy() for a in pos]
    v = [np.zeros(len(p[0])) for _ in range(3)]
    orc.update_particles(n, (0, 0, 0), (1, 1, 1), p, v, g, 2.0, 0.0, 1.0, 0)
    return np.stack(v)


def oracle_steps(orc, n, pos, vel, m, nsteps, dt, rel_tol, nyx_faithful=False):
    """State after each oracle KDK step: (pos, vel, grid g, phi, V-cycles)."""
    pos = [a.copy() for a in pos]
    vel = [a.copy() for a in vel]
    if nyx_faithful:
        phi = np.zeros((n[2], n[1], n[0]))
    else:
        phi, _, _ = orc.solve_for_phi(n, (0, 0, 0), (1, 1, 1), pos, m_const=m, rel_tol=rel_tol)
    out = []
    for _ in range(nsteps):
        g, st = orc.pm_step(n, (0, 0, 0), (1, 1, 1), pos, vel, phi, m_const=m, dt=dt, rel_tol=rel_tol,
                            nyx_faithful=nyx_faithful, want_g=True)
        out.append(([a.copy() for a in pos], [a.copy() for a in vel], g.copy(), phi.copy(), int(st[0])))
    return out


# ---------------------------------------------------------------------------
# 512^3 clustered workload (mixed precision)
# ---------------------------------------------------------------------------

def _clustered(nyx, n):
    return nyx.ics.plummer_clustered(n, nyx.ics.SEED_BASE + 2, n_spheres=max(1, int(2000 * (n / 512) ** 3)))


def test_clustered_recipe_mixed_vs_oracle(nyx, orc):
    # the recipe at 64^3 (the sphere count scaled with the volume: 3 spheres of
    # ~70,000 particles, so the cores are as dense as at 512^3 per sphere)
    n = 64
    nn = (n, n, n)
    dt, steps = 1e-3, 3
    pos, vel, m = _clustered(nyx, n)
    ref = oracle_steps(orc, nn, pos, vel, m, steps, dt, 1e-10)
    with nyx.PMContext(nyx.BoxGeom(nn), precision="mixed") as c:
        c.upload(0, pos, vel, m_const=m)
        rho = c.assign_density()
        c.pm_init(1.0, 1.0, 1e-10)
        for s in range(steps):
            c.pm_steps(1, dt, rel_tol=1e-10)
            p, v = c.download()
            g = c.grav_vector()
            ga = c.gather_accel()
            rp, rv, rg, _, _ = ref[s]
            gmax = np.abs(rg).max()
            assert rel_l2(g, rg) <= 1e-6, s
            assert rel_l2(ga, oracle_accel(orc, nn, p, rg)) <= 1e-6, s
            dv = (s + 1) * dt * 1e-6 * gmax
            for d in range(3):
                assert np.abs(v[d] - rv[d]).max() <= dv, (s, d)
                assert np.abs(p[d] - rp[d]).max() <= dt * dv + 1e-15, (s, d)
    # the deposit of the mixed context agrees with the oracle's to the FP64 bar
    assert relmax(rho, orc.deposit(nn, (0, 0, 0), (1, 1, 1), *pos, m_const=m)) <= 1e-13


def _momentum_ratio(torch, g, N):
    """|sum_p m g_p| / sum_p m |g_p| for equal masses (g: 3N, component-major)."""
    gv = g.view(3, N)
    return (torch.linalg.vector_norm(gv.sum(dim=1)) / torch.linalg.vector_norm(gv, dim=0).sum()).item()


@pytest.mark.timeout(900)
def test_clustered_512_mixed_full_size(nyx):
    # the workload at its own size: 512^3 cells and particles, inputs made on the
    # device; after each of three mixed steps an FP64 solve of the same density
    # from a zero guess gives the reference accelerations
    import torch
    n = 512
    N = n ** 3
    pos, vel, m = nyx.ics.plummer_clustered_device(n, nyx.ics.SEED_BASE + 2)
    c = nyx.PMContext(nyx.BoxGeom((n, n, n)), precision="mixed")
    try:
        c.upload_device(0, N, [t.data_ptr() for t in pos + vel], m_const=m)
        c.sync()
        del pos, vel
        torch.cuda.empty_cache()
        rho_a = torch.empty(N, dtype=torch.float64, device="cuda")
        rho_b = torch.empty(N, dtype=torch.float64, device="cuda")
        c.assign_density_device(rho_a.data_ptr())
        c.assign_density_device(rho_b.data_ptr())
        assert torch.equal(rho_a, rho_b)  # bitwise repeatable
        mass = rho_a.sum().item() / N  # sum rho dV, dV = 1/N, total mass 1
        assert abs(mass - 1.0) <= 1e-12
        del rho_a, rho_b
        c.pm_init(1.0, 1.0, 1e-10)
        gm = torch.empty(3 * N, dtype=torch.float64, device="cuda")
        g64 = torch.empty(3 * N, dtype=torch.float64, device="cuda")
        phi_m = torch.empty(N, dtype=torch.float64, device="cuda")
        rels, moms = [], []
        for s in range(3):
            its = c.pm_steps(1, 1e-3, rel_tol=1e-10)
            c.gather_accel_device(gm.data_ptr())
            c.get_phi_device(phi_m.data_ptr())
            c.set_precision("fp64")
            c.assign_density(want=False)
            c.solve_for_phi(1.0, 1.0, 1e-10, 0.0, use_guess=False)
            c.gather_accel_device(g64.data_ptr())
            c.set_phi_device(phi_m.data_ptr())
            c.set_precision("mixed")
            rel = (torch.linalg.vector_norm(gm - g64) / torch.linalg.vector_norm(g64)).item()
            rels.append(rel)
            moms.append((_momentum_ratio(torch, gm, N), _momentum_ratio(torch, g64, N)))
            print(f"512^3 clustered step {s + 1}: {int(its[0])} V-cycles, accel rel-L2 {rel:.3e}, "
                  f"momentum ratio mixed {moms[-1][0]:.3e} fp64 {moms[-1][1]:.3e}", flush=True)
        assert max(rels) <= 1e-6
        # exactly 0 for the exact discrete solution of an FP64 deposit; the
        # mixed steps deposit with FP32 weights (not the exact transpose of the
        # FP64 interpolation)
        assert max(mm[1] for mm in moms) <= 1e-8
        assert max(mm[0] for mm in moms) <= 1e-7
        # the step kept the particles: the mass is unchanged
        rho = torch.empty(N, dtype=torch.float64, device="cuda")
        c.assign_density_device(rho.data_ptr())
        assert abs(rho.sum().item() / N - 1.0) <= 1e-12
        del rho, gm, g64, phi_m
    finally:
        c.close()
        torch.cuda.empty_cache()


# ---------------------------------------------------------------------------
# 1024^3 GRF workload (mixed precision)
# ---------------------------------------------------------------------------

@pytest.mark.timeout(900)
def test_grf1024_mixed_per_step_accel(nyx):
    # 1024^3 cells and particles (~60 GB of particle data), inputs made on the
    # device; dense tiles (a full sort after every kick-drift) leave room for
    # the 2 x 26 GB of comparison arrays
    import torch
    n = 1024
    N = n ** 3
    steps = 3
    pos, vel, m = nyx.ics.grf_lattice_device(n, nyx.ics.SEED_BASE + 4, -1.5, 1.0, 0.1)
    c = nyx.PMContext(nyx.BoxGeom((n, n, n)), precision="mixed", resort_slack=-1.0)
    try:
        c.upload_device(0, N, [t.data_ptr() for t in pos + vel], m_const=m)
        c.sync()
        del pos, vel
        torch.cuda.empty_cache()
        c.pm_init(1.0, 1.0, 1e-10)
        gm = torch.empty(3 * N, dtype=torch.float64, device="cuda")
        g64 = torch.empty(3 * N, dtype=torch.float64, device="cuda")
        phi_m = torch.empty(N, dtype=torch.float64, device="cuda")
        rels = []
        for s in range(steps):
            its = c.pm_steps(1, 5e-4, rel_tol=1e-10)
            c.gather_accel_device(gm.data_ptr())
            c.get_phi_device(phi_m.data_ptr())
            c.set_precision("fp64")
            c.assign_density(want=False)
            c.solve_for_phi(1.0, 1.0, 1e-10, 0.0, use_guess=False)
            c.gather_accel_device(g64.data_ptr())
            c.set_phi_device(phi_m.data_ptr())
            c.set_precision("mixed")
            rels.append((torch.linalg.vector_norm(gm - g64) / torch.linalg.vector_norm(g64)).item())
            print(f"1024^3 GRF step {s + 1}: {int(its[0])} V-cycles, accel rel-L2 {rels[-1]:.3e}", flush=True)
        assert max(rels) <= 1e-6
        assert min(rels) > 0.0  # two independent solves, not the same iterate
        del gm, g64, phi_m
    finally:
        c.close()
        torch.cuda.empty_cache()


# ---------------------------------------------------------------------------
# Nyx's two-solve step
# ---------------------------------------------------------------------------

@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_nyx_faithful_vs_oracle(nyx, orc, precision):
    # the old phi re-solved from a zero guess at the start of every step, then
    # the kick-drift, the new solve from it and the kick, as the oracle's
    # nyx_faithful step; no pm_init (each step makes its own old phi)
    n = (32, 32, 32)
    dt, steps = 1e-3, 3
    pos, vel, m = nyx.ics.lattice_gaussian(32, 5, sigma_cells=1.0)
    ref = oracle_steps(orc, n, pos, vel, m, steps, dt, 1e-10, nyx_faithful=True)
    with nyx.PMContext(nyx.BoxGeom(n), precision=precision) as c:
        c.upload(0, pos, vel, m_const=m)
        for s in range(steps):
            it = c.pm_steps(1, dt, rel_tol=1e-10, nyx_faithful=True)
            p, v = c.download()
            g = c.grav_vector()
            rp, rv, rg, _, rit = ref[s]
            if precision == "fp64":
                assert int(it[0]) == rit
                assert relmax(g, rg) <= 1e-12
                for d in range(3):
                    assert relmax(p[d], rp[d]) <= 1e-12
                    assert relmax(v[d], rv[d]) <= 1e-12
            else:
                gmax = np.abs(rg).max()
                assert rel_l2(g, rg) <= 1e-6
                dv = (s + 1) * dt * 1e-6 * gmax
                for d in range(3):
                    assert np.abs(v[d] - rv[d]).max() <= dv
                    assert np.abs(p[d] - rp[d]).max() <= dt * dv + 1e-15


@pytest.mark.parametrize("steps", [4, 5])
def test_nyx_faithful_graph_equals_eager(nyx, steps):
    # captured two-solve steps (4 solves per replay) replay the eager ones bit for bit
    pos, vel, m = nyx.ics.lattice_gaussian(32, 8, sigma_cells=1.0)
    outs = []
    for graph in (False, True):
        with nyx.PMContext(nyx.BoxGeom((32, 32, 32))) as c:
            c.upload(0, pos, vel, m_const=m)
            its = c.pm_steps(steps, 1e-3, rel_tol=1e-10, nyx_faithful=True, use_graph=graph)
            outs.append((c.download(), c.get_phi(), list(its)))
    (a, pa, ia), (b, pb, ib) = outs
    assert ia == ib
    assert np.array_equal(pa, pb)
    for d in range(3):
        assert np.array_equal(a[0][d], b[0][d])
        assert np.array_equal(a[1][d], b[1][d])


# ---------------------------------------------------------------------------
# mixed-precision batch
# ---------------------------------------------------------------------------

def test_mixed_batch_vs_oracle(nyx, orc):
    sizes = [(16, 16, 16), (24, 24, 24), (32, 32, 32)]
    inputs = [nyx.ics.grf_lattice(s[0], 210 + i, -2.0, 0.5, 0.1) for i, s in enumerate(sizes)]
    dt, steps = 1e-3, 2
    with nyx.PMContext([nyx.BoxGeom(s) for s in sizes], precision="mixed") as c:
        for b, (pos, vel, m) in enumerate(inputs):
            c.upload(b, pos, vel, m_const=m)
        c.pm_init(1.0, 1.0, 1e-10)
        c.pm_steps(steps, dt, rel_tol=1e-10)
        out = [(c.download(b), c.grav_vector(b), c.gather_accel(b)) for b in range(len(sizes))]
    for b, n in enumerate(sizes):
        pos, vel, m = inputs[b]
        rp, rv, rg, _, _ = oracle_steps(orc, n, pos, vel, m, steps, dt, 1e-10)[-1]
        (p, v), g, ga = out[b]
        assert rel_l2(g, rg) <= 1e-6
        assert rel_l2(ga, oracle_accel(orc, n, p, rg)) <= 1e-6
        dv = steps * dt * 1e-6 * np.abs(rg).max()
        for d in range(3):
            assert np.abs(v[d] - rv[d]).max() <= dv
            assert np.abs(p[d] - rp[d]).max() <= dt * dv + 1e-15


# ---------------------------------------------------------------------------
# captured steps
# ---------------------------------------------------------------------------

@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_graph_calls_with_odd_eager_steps(nyx, precision):
    # a call of 5 captured steps runs 1 eager step, one replay of 2 and 2 eager
    # steps, so it ends with the other particle buffer live; the second call's
    # replay must not reuse the graph captured for the first buffer
    pos, vel, m = nyx.ics.plummer_clustered(32, 91, n_spheres=6)
    outs = []
    for graph in (False, True):
        with nyx.PMContext(nyx.BoxGeom((32, 32, 32)), precision=precision) as c:
            c.upload(0, pos, vel, m_const=m)
            c.pm_init(1.0, 1.0, 1e-10)
            its = []
            for _ in range(3):
                its.append(list(c.pm_steps(5, 2e-3, rel_tol=1e-10, use_graph=graph)))
            outs.append((c.download(), c.get_phi(), its))
    (a, pa, ia), (b, pb, ib) = outs
    assert ia == ib
    assert np.array_equal(pa, pb)
    for d in range(3):
        assert np.array_equal(a[0][d], b[0][d])
        assert np.array_equal(a[1][d], b[1][d])


def test_graph_reports_an_earlier_solve_of_the_replay(nyx):
    # Nyx's two-solve step: the old phi from a zero guess needs more V-cycles
    # than the new solve started from it.  With fewer cycles recorded than the
    # first solve needs, the first solve of each captured step stops above its
    # target while the replay's last solve converges; the failure must still be
    # reported (forced from the host through opts.graph_cycles)
    pos, vel, m = nyx.ics.lattice_gaussian(32, 9, sigma_cells=1.0)
    with nyx.PMContext(nyx.BoxGeom((32, 32, 32))) as c:
        c.upload(0, pos, vel, m_const=m)
        c.assign_density(want=False)
        z, _ = c.solve_for_phi(1.0, 1.0, 1e-10, 0.0, use_guess=False)
        w = c.pm_steps(1, 1e-3, rel_tol=1e-10, nyx_faithful=True)
    z, w = int(z[0]), int(w[0])
    assert w < z
    with nyx.PMContext(nyx.BoxGeom((32, 32, 32)), graph_cycles=w) as c:
        c.upload(0, pos, vel, m_const=m)
        with pytest.raises(nyx.NyxPMError) as ei:
            c.pm_steps(4, 1e-3, rel_tol=1e-10, nyx_faithful=True, use_graph=True)
        assert ei.value.code == 4
        assert "recorded" in str(ei.value)
