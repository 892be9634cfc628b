"""CPU: pin the oracle's CIC deposit (AssignDensity) against exact arithmetic,
closed forms and invariants."""
import numpy as np
import pytest


def _dep(orc, g, key, pos, m):
    return orc.deposit(g[f"{key}_n"], g[f"{key}_plo"], g[f"{key}_phi"], *pos, m=m)


@pytest.mark.parametrize("name", ["twoline", "mass8"])
def test_reference_particle_files_exact(orc, golden, name):
    # Nyx's own ParticleOnlyTest particle files, exact rational expected density
    rho = orc.deposit(golden["box32_n"], golden["box32_plo"], golden["box32_phi"],
                      *golden[f"{name}_pos"], m=golden[f"{name}_mass"])
    ref = golden[f"{name}_rho"]
    assert np.abs(rho - ref).max() <= 1e-14 * np.abs(ref).max()


def test_rect_box_exact(orc, golden):
    # non-cubic cells, offset prob_lo, particles up to 0.3 cell outside the box
    rho = _dep(orc, golden, "rect", golden["rect_pos"], golden["rect_mass"])
    ref = golden["rect_rho"]
    assert np.abs(rho - ref).max() <= 1e-14 * np.abs(ref).max()


def test_cell_centre_puts_all_mass_in_cell(orc):
    n = [8, 8, 8]
    for (i, j, k) in [(0, 0, 0), (3, 5, 7), (7, 7, 7)]:
        x, y, z = [(np.array([c + 0.5]) / 8.0) for c in (i, j, k)]
        rho = orc.deposit(n, [0, 0, 0], [1, 1, 1], x, y, z, m_const=2.0)
        expect = np.zeros((8, 8, 8))
        expect[k, j, i] = 2.0 * 512
        assert np.array_equal(rho, expect)


def test_cell_corner_splits_in_eight(orc):
    n = [8, 8, 8]
    x = y = z = np.array([0.0])  # corner shared by cells 7 and 0 in each dimension
    rho = orc.deposit(n, [0, 0, 0], [1, 1, 1], x, y, z, m_const=1.0)
    nz = np.argwhere(rho)
    assert len(nz) == 8
    assert set(map(tuple, nz)) == {(a, b, c) for a in (0, 7) for b in (0, 7) for c in (0, 7)}
    assert np.allclose(rho[rho > 0], 512 / 8.0, rtol=0, atol=1e-12)


def test_lattice_gives_uniform_density(orc):
    from nyx_amd import ics
    x, y, z = ics.lattice_positions(16)
    rho = orc.deposit([16] * 3, [0] * 3, [1] * 3, x, y, z, m_const=1.0 / 16 ** 3)
    assert np.abs(rho - 1.0).max() < 1e-13


def test_translation_by_whole_cells(orc):
    r = np.random.default_rng(1)
    n = [12, 10, 8]
    pos = [r.random(500) for _ in range(3)]
    rho = orc.deposit(n, [0] * 3, [1] * 3, *pos, m_const=1.0)
    shift = (3, 2, 5)
    pos2 = [np.mod(pos[d] + shift[d] / n[d], 1.0) for d in range(3)]
    rho2 = orc.deposit(n, [0] * 3, [1] * 3, *pos2, m_const=1.0)
    ref = np.roll(rho, shift=(shift[2], shift[1], shift[0]), axis=(0, 1, 2))
    assert np.abs(rho2 - ref).max() < 1e-9 * np.abs(ref).max()


def test_mass_conservation(orc):
    r = np.random.default_rng(2)
    n = [16, 16, 16]
    pos = [r.random(5000) for _ in range(3)]
    m = r.uniform(0.1, 3.0, 5000)
    rho = orc.deposit(n, [0] * 3, [2, 2, 2], *pos, m=m)
    dv = (2.0 / 16) ** 3
    assert abs(rho.sum() * dv - m.sum()) < 1e-12 * m.sum()


def test_adjoint_of_interpolation(orc):
    # sum_c rho_c q_c dV == sum_p m_p CIC(q)(x_p): deposit and the kick's gather
    # use the same weights (checked through the oracle's particle update with a
    # field in the first component and dt = 2 so that v = val).
    r = np.random.default_rng(3)
    n = [8, 12, 6]
    plo, phi = [0.0, -1.0, 0.5], [1.0, 2.0, 1.25]
    NP = 400
    pos = [plo[d] + r.random(NP) * (phi[d] - plo[d]) for d in range(3)]
    m = r.uniform(0.5, 1.5, NP)
    q = r.standard_normal((6, 12, 8))
    rho = orc.deposit(n, plo, phi, *pos, m=m)
    dv = np.prod([(phi[d] - plo[d]) / n[d] for d in range(3)])
    g = np.stack([q, np.zeros_like(q), np.zeros_like(q)])
    p2 = [a.copy() for a in pos]
    v = [np.zeros(NP) for _ in range(3)]
    orc.update_particles(n, plo, phi, p2, v, g, 2.0, 1.0, 1.0, 0)
    lhs = (rho * q).sum() * dv
    rhs = (m * v[0]).sum()
    assert abs(lhs - rhs) < 1e-12 * np.abs(m * v[0]).sum()


def test_empty_and_single_thread_equivalence(orc):
    rho = orc.deposit([4, 4, 4], [0] * 3, [1] * 3, np.array([]), np.array([]), np.array([]))
    assert not rho.any()
    r = np.random.default_rng(4)
    pos = [r.random(3000) for _ in range(3)]
    nt = orc.num_threads()
    a = orc.deposit([10, 10, 10], [0] * 3, [1] * 3, *pos)
    orc.set_num_threads(1)
    try:
        b = orc.deposit([10, 10, 10], [0] * 3, [1] * 3, *pos)
    finally:
        orc.set_num_threads(nt)
    assert np.array_equal(a, b)  # summation order is independent of thread count


# ---------------------------------------------------------------------------
# AssignDensityAndVels (DarkMatterParticleContainer.cpp:456-460)
# ---------------------------------------------------------------------------

def test_density_and_vels_uniform_velocity(orc):
    # every particle moves with the same v: every cell with mass has exactly that
    # velocity (up to the rounding of a weighted mean), empty cells 0; comp 0 is
    # the plain deposit
    r = np.random.default_rng(5)
    n = (8, 6, 10)
    pos = [r.random(300) * 0.5 for _ in range(3)]  # part of the box stays empty
    vel = [np.full(300, 0.3), np.full(300, -1.25), np.full(300, 2.0)]
    m = r.uniform(0.5, 2.0, 300)
    out = orc.deposit_vels(n, (0, 0, 0), (1, 1, 1), pos, vel, m=m)
    rho = orc.deposit(n, (0, 0, 0), (1, 1, 1), pos[0], pos[1], pos[2], m=m)
    assert out.shape == (4, 10, 6, 8)
    has = out[0] > 0
    assert has.any() and (~has).any()
    for d, v in enumerate((0.3, -1.25, 2.0)):
        assert np.allclose(out[1 + d][has], v, rtol=1e-14, atol=0)
        assert np.all(out[1 + d][~has] == 0)
    assert np.array_equal(out[0], rho)


def test_density_and_vels_known_answers(orc):
    n = (4, 4, 4)
    z = np.zeros(1)
    # a particle at a cell centre: all its mass and velocity in that cell
    pos = [np.array([0.375]), np.array([0.625]), np.array([0.125])]
    out = orc.deposit_vels(n, (0, 0, 0), (1, 1, 1), pos, [np.array([1.5]), np.array([-2.0]), np.array([0.25])],
                           m_const=2.0)
    assert out[0, 0, 2, 1] == 2.0 * 64 and np.count_nonzero(out[0]) == 1
    assert (out[1, 0, 2, 1], out[2, 0, 2, 1], out[3, 0, 2, 1]) == (1.5, -2.0, 0.25)
    # two particles sharing a cell: the mass-weighted mean of their velocities
    pos = [np.array([0.375, 0.375]), np.array([0.625, 0.625]), np.array([0.125, 0.125])]
    out = orc.deposit_vels(n, (0, 0, 0), (1, 1, 1), pos, [np.array([1.0, 4.0]), z.repeat(2), z.repeat(2)],
                           m=np.array([1.0, 3.0]))
    assert out[1, 0, 2, 1] == (1.0 * 1.0 + 3.0 * 4.0) / 4.0


def test_density_and_vels_momentum_conservation(orc):
    # sum over cells of (mass-weighted velocity) * (cell mass) = total momentum
    r = np.random.default_rng(9)
    n = (12, 12, 12)
    pos = [r.random(2000) for _ in range(3)]
    vel = [r.normal(0, 1, 2000) for _ in range(3)]
    m = r.uniform(0.1, 1.0, 2000)
    out = orc.deposit_vels(n, (0, 0, 0), (1, 1, 1), pos, vel, m=m)
    cell_mass = out[0] / 12 ** 3
    for d in range(3):
        assert abs(np.sum(out[1 + d] * cell_mass) - np.sum(m * vel[d])) <= 1e-12 * np.sum(m * np.abs(vel[d]))
