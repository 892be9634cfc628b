"""CPU: pin the oracle's particle update (moveKickDrift / moveKick) and the
whole KDK step against exact arithmetic, closed forms and conservation laws."""
import numpy as np
import pytest


def test_interp_and_kick_drift_golden(orc, golden):
    n, plo, phi = golden["rect_n"], golden["rect_plo"], golden["rect_phi"]
    dt, a_prev, a_cur = golden["rect_kd_params"]
    pos = [a.copy() for a in golden["rect_pos"]]
    vel = [a.copy() for a in golden["rect_vel"]]
    orc.update_particles(n, plo, phi, pos, vel, golden["rect_g"], dt, a_prev, a_cur, 1)
    vref, xref = golden["rect_kd_v"], golden["rect_kd_x"]
    for d in range(3):
        assert np.abs(vel[d] - vref[d]).max() <= 1e-13 * np.abs(vref).max()
        assert np.abs(pos[d] - xref[d]).max() <= 1e-15 * np.abs(xref).max()


def test_kick_only_does_not_move(orc, golden):
    n, plo, phi = golden["rect_n"], golden["rect_plo"], golden["rect_phi"]
    pos = [a.copy() for a in golden["rect_pos"]]
    vel = [a.copy() for a in golden["rect_vel"]]
    orc.update_particles(n, plo, phi, pos, vel, golden["rect_g"], 0.1, 1.0, 1.0, 0)
    for d in range(3):
        assert np.array_equal(pos[d], golden["rect_pos"][d])
        ref = golden["rect_vel"][d] + 0.05 * golden["rect_gp"][d]
        assert np.abs(vel[d] - ref).max() < 1e-13 * np.abs(ref).max()


def test_uniform_and_linear_fields_interpolate_exactly(orc):
    n = (10, 10, 10)
    r = np.random.default_rng(6)
    NP = 200
    pos = [0.25 + 0.5 * r.random(NP) for _ in range(3)]  # away from the periodic seam
    vel = [np.zeros(NP) for _ in range(3)]
    g = np.stack([np.full((10, 10, 10), 3.0), np.zeros((10, 10, 10)), np.zeros((10, 10, 10))])
    z, y, x = np.meshgrid(*[(np.arange(10) + 0.5) / 10] * 3, indexing="ij")
    g[1] = 2.0 * x - 1.0 * z  # linear in cell-centre coordinates
    orc.update_particles(n, [0] * 3, [1] * 3, [p.copy() for p in pos], vel, g, 2.0, 1.0, 1.0, 0)
    assert np.abs(vel[0] - 3.0).max() < 1e-14
    assert np.abs(vel[1] - (2.0 * pos[0] - pos[2])).max() < 1e-13


def test_periodic_wrap(orc):
    n = [8, 8, 8]
    x = np.array([-0.25, 1.0, 1.25, 0.5, -1e-18, 0.999999999])
    y = np.zeros(6)
    z = np.zeros(6)
    orc.enforce_periodic(n, [0] * 3, [1] * 3, [x, y, z])
    rhi = orc.domain_rhi(0.0, 1.0, 8)
    assert rhi < 1.0 and int(np.floor(rhi * 8)) == 7
    assert np.allclose(x[:4], [0.75, 0.0, 0.25, 0.5])
    assert 0.0 <= x[4] <= rhi  # tiny negative: shifted up, clamped below 1
    assert x[5] == 0.999999999


def _two_body(orc):
    # two equal masses, mirror symmetric about the box centre
    n = (16, 16, 16)
    pos = [np.array([0.3, 0.7]), np.array([0.5, 0.5]), np.array([0.5, 0.5])]
    vel = [np.zeros(2), np.zeros(2), np.zeros(2)]
    return n, pos, vel


def test_two_body_symmetry_and_momentum(orc):
    n, pos, vel = _two_body(orc)
    phi, _, st = orc.solve_for_phi(n, [0] * 3, [1] * 3, pos, m_const=0.5, rel_tol=1e-13)
    for _ in range(3):
        g, st = orc.pm_step(n, [0] * 3, [1] * 3, pos, vel, phi, m_const=0.5, dt=1e-2, rel_tol=1e-13,
                            want_g=True)
    # attraction: the particles move towards each other, symmetrically
    assert vel[0][0] > 0 > vel[0][1]
    assert abs(vel[0][0] + vel[0][1]) < 1e-12 * abs(vel[0][0])
    assert abs((pos[0][0] - 0.5) + (pos[0][1] - 0.5)) < 1e-14
    assert np.abs(vel[1]).max() < 1e-14 and np.abs(vel[2]).max() < 1e-14


def test_momentum_conservation_random(orc):
    from nyx_amd import ics
    pos, vel, m = ics.lattice_gaussian(16, 11, sigma_cells=2.0)
    n = (16, 16, 16)
    phi, _, st = orc.solve_for_phi(n, [0] * 3, [1] * 3, pos, m_const=m, rel_tol=1e-13)
    g = orc.grav_vector(n, [0] * 3, [1] * 3, phi)
    # force on the particles: gather with dt=2, v=0 -> v = g(x)
    v = [np.zeros_like(pos[0]) for _ in range(3)]
    orc.update_particles(n, [0] * 3, [1] * 3, [p.copy() for p in pos], v, g, 2.0, 1.0, 1.0, 0)
    for d in range(3):
        tot = (m * v[d]).sum()
        scale = (m * np.abs(v[d])).sum()
        assert abs(tot) < 1e-9 * scale


def test_step_conserves_mass_and_keeps_positions_in_box(orc):
    from nyx_amd import ics
    pos, vel, m = ics.lattice_gaussian(16, 12)
    vel = [v * 50 for v in vel]  # cross the periodic boundaries
    n = (16, 16, 16)
    phi, _, _ = orc.solve_for_phi(n, [0] * 3, [1] * 3, pos, m_const=m)
    for _ in range(2):
        orc.pm_step(n, [0] * 3, [1] * 3, pos, vel, phi, m_const=m, dt=1e-2)
    for d in range(3):
        assert pos[d].min() >= 0.0 and pos[d].max() < 1.0
    rho = orc.deposit(n, [0] * 3, [1] * 3, *pos, m_const=m)
    assert abs(rho.mean() - 1.0) < 1e-13


def test_faithful_and_canonical_agree(orc):
    # re-solving the old phi from zero (Nyx) gives the same step to solver tolerance
    from nyx_amd import ics
    pos, vel, m = ics.lattice_gaussian(16, 13, sigma_cells=1.0)
    n = (16, 16, 16)
    phi, _, _ = orc.solve_for_phi(n, [0] * 3, [1] * 3, pos, m_const=m, rel_tol=1e-12)
    p1 = [a.copy() for a in pos]; v1 = [a.copy() for a in vel]; ph1 = phi.copy()
    p2 = [a.copy() for a in pos]; v2 = [a.copy() for a in vel]; ph2 = phi.copy()
    orc.pm_step(n, [0] * 3, [1] * 3, p1, v1, ph1, m_const=m, dt=1e-3, rel_tol=1e-12)
    orc.pm_step(n, [0] * 3, [1] * 3, p2, v2, ph2, m_const=m, dt=1e-3, rel_tol=1e-12, nyx_faithful=True)
    for d in range(3):
        dv = np.abs(v1[d] - v2[d]).max()
        assert dv < 1e-9 * np.abs(v1[d]).max()


# ---------------------------------------------------------------------------
# estTimestep (NyxParticleContainer.H:327-410)
# ---------------------------------------------------------------------------

def test_est_time_step_velocity_known_answer(orc):
    # g = 0: dt = cfl / max_d(|v_d| dxi_d / a); n = 8 on the unit box, a = 2:
    # |v| (0.2, 0.5, 0.1) -> 0.5 * 8 / 2 = 2 -> 0.5 / 2 = 0.25; a slower particle
    # does not change the minimum
    n = (8, 8, 8)
    g = np.zeros((3, 8, 8, 8))
    pos = [np.array([0.3, 0.9]), np.array([0.5, 0.1]), np.array([0.7, 0.4])]
    vel = [np.array([0.2, 0.01]), np.array([-0.5, 0.0]), np.array([0.1, -0.02])]
    assert orc.est_time_step(n, (0, 0, 0), (1, 1, 1), pos, vel, g, a=2.0, cfl=0.5) == 0.25
    # anisotropic cells: the y spacing (dxi = 16) dominates
    assert orc.est_time_step((8, 16, 8), (0, 0, 0), (1, 1, 1), pos, vel, np.zeros((3, 8, 16, 8)),
                             a=1.0, cfl=0.5) == 0.5 / (0.5 * 16)


def test_est_time_step_no_particles(orc):
    e = np.array([])
    assert orc.est_time_step((8, 8, 8), (0, 0, 0), (1, 1, 1), [e, e, e], [e, e, e],
                             np.zeros((3, 8, 8, 8))) == -1e50


def test_est_time_step_gravity_single_mode(orc):
    # phi = A cos(2 pi x + s): the face-averaged gradient at cell i is
    # g_x = -n A sin(2 pi (i + 1/2) / n + s) sin(2 pi / n) (closed form); at rest the
    # estimate is 1 / sqrt(|g_x| n) of the particle's cell floor(x n) -- cell 3
    # here, near its upper edge, where the CIC index floor(x n + 1/2) is 4
    n, A, sh = 8, 0.7, 0.3
    x = (np.arange(n) + 0.5) / n
    phi = np.broadcast_to(A * np.cos(2 * np.pi * x + sh)[None, None, :], (n, n, n)).copy()
    g = orc.grav_vector((n, n, n), (0, 0, 0), (1, 1, 1), phi)
    gx = lambda i: -n * A * np.sin(2 * np.pi * (i + 0.5) / n + sh) * np.sin(2 * np.pi / n)
    assert np.allclose(g[0, 0, 0, :], [gx(i) for i in range(n)], rtol=0, atol=1e-13)
    z = np.zeros(1)
    pos = [np.array([4.0 / n - 1e-9]), np.array([0.5]), np.array([0.5])]
    dt = orc.est_time_step((n, n, n), (0, 0, 0), (1, 1, 1), pos, [z, z, z], g, a=1.0, cfl=0.5)
    assert abs(dt - 1 / np.sqrt(abs(gx(3)) * n)) <= 1e-14
    assert abs(dt - 1 / np.sqrt(abs(gx(4)) * n)) > 1e-3  # not the CIC cell
    # a fast particle: the velocity term wins
    vel = [np.array([100.0]), z, z]
    dt2 = orc.est_time_step((n, n, n), (0, 0, 0), (1, 1, 1), pos, vel, g, a=1.0, cfl=0.5)
    assert dt2 == 0.5 / (100.0 * n)
