"""CPU oracle for the Nyx particle-mesh gravity step.

TEST INFRASTRUCTURE ONLY: this package is the checker.  Only ``tests/``,
``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of ``bench.py`` may
import it; the product path (``nyx_amd``) never does.

* ``lib()`` -- ctypes handle on ``libnyxpm_oracle.so``, the C/OpenMP FP64
  restatement of Nyx + AMReX on one periodic level (see ``nyxpm_oracle.c``).
* ``fft_solve`` -- exact solve of the same discrete periodic 7-point operator by
  FFT (numpy FP64), an independent check of the multigrid fixed point.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

_d = ctypes.c_double
_i64 = ctypes.c_int64
_int = ctypes.c_int
_pd = ctypes.POINTER(ctypes.c_double)
_pi = ctypes.POINTER(ctypes.c_int)
_vp = ctypes.c_void_p


def build(force: bool = False) -> str:
    """Compile the oracle with its Makefile; returns the library path."""
    so = os.path.join(_HERE, "libnyxpm_oracle.so")
    if force or not os.path.exists(so):
        if force and os.path.exists(so):
            os.remove(so)
        subprocess.run(["make", "-s", "-C", _HERE, "libnyxpm_oracle.so"], check=True)
    return so


def lib():
    global _LIB
    if _LIB is not None:
        return _LIB
    so = os.path.join(_HERE, "libnyxpm_oracle.so")
    if not os.path.exists(so):
        build()
    L = ctypes.CDLL(so)
    L.orc_deposit_cic.argtypes = [_vp, _vp, _vp, _i64, _vp, _vp, _vp, _vp, _d, _vp]
    L.orc_deposit_cic.restype = _int
    L.orc_sum_particle_mass.argtypes = [_i64, _vp, _d]
    L.orc_sum_particle_mass.restype = _d
    L.orc_mass_offset.argtypes = [_vp, _vp, _d]
    L.orc_mass_offset.restype = _d
    L.orc_correct_rhs_using_offset.argtypes = [_vp, _vp, _vp, _vp, _d]
    L.orc_correct_rhs_using_offset.restype = _d
    L.orc_scale_rhs.argtypes = [_i64, _vp, _d, _d]
    L.orc_scale_rhs.restype = None
    L.orc_mg_levels.argtypes = [_vp, _vp]
    L.orc_mg_levels.restype = _int
    L.orc_bottom_green.argtypes = [_vp, _vp, _vp]
    L.orc_bottom_green.restype = None
    L.orc_residual.argtypes = [_vp, _vp, _vp, _vp, _vp]
    L.orc_residual.restype = _d
    L.orc_mg_solve.argtypes = [_vp, _vp, _vp, _vp, _d, _d, _int, _int, _int, _int, _vp, _vp]
    L.orc_mg_solve.restype = _int
    L.orc_grav_vector.argtypes = [_vp, _vp, _vp, _vp, _vp]
    L.orc_grav_vector.restype = None
    L.orc_deposit_cic_vels.argtypes = [_vp, _vp, _vp, _i64, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _d, _vp]
    L.orc_deposit_cic_vels.restype = _int
    L.orc_est_time_step.argtypes = [_vp, _vp, _vp, _i64, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _d, _d]
    L.orc_est_time_step.restype = _d
    L.orc_update_particles.argtypes = [_vp, _vp, _vp, _i64, _vp, _vp, _vp, _vp, _vp, _vp,
                                       _vp, _d, _d, _d, _int]
    L.orc_update_particles.restype = None
    L.orc_domain_rhi.argtypes = [_d, _d, _int]
    L.orc_domain_rhi.restype = _d
    L.orc_enforce_periodic.argtypes = [_vp, _vp, _vp, _i64, _vp, _vp, _vp]
    L.orc_enforce_periodic.restype = None
    L.orc_solve_for_phi.argtypes = [_vp, _vp, _vp, _i64, _vp, _vp, _vp, _vp, _d, _d, _d,
                                    _d, _d, _vp, _vp, _vp]
    L.orc_solve_for_phi.restype = _int
    L.orc_pm_step.argtypes = [_vp, _vp, _vp, _i64, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _d,
                              _vp, _d, _d, _d, _d, _d, _d, _int, _vp, _vp]
    L.orc_pm_step.restype = _int
    L.orc_num_threads.argtypes = []
    L.orc_num_threads.restype = _int
    L.orc_set_num_threads.argtypes = [_int]
    L.orc_set_num_threads.restype = None
    _LIB = L
    return L


def _p(a):
    return None if a is None else a.ctypes.data


def _geom(n, plo, phi):
    n = np.ascontiguousarray(n, dtype=np.int32)
    plo = np.ascontiguousarray(plo, dtype=np.float64)
    phi = np.ascontiguousarray(phi, dtype=np.float64)
    return n, plo, phi


def _c64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def grid_shape(n):
    """numpy shape (nz, ny, nx) of a cell field with x fastest."""
    return (int(n[2]), int(n[1]), int(n[0]))


def deposit(n, plo, phi, x, y, z, m=None, m_const=1.0):
    n, plo, phi = _geom(n, plo, phi)
    x, y, z = _c64(x), _c64(y), _c64(z)
    m = None if m is None else _c64(m)
    rho = np.empty(grid_shape(n), dtype=np.float64)
    rc = lib().orc_deposit_cic(_p(n), _p(plo), _p(phi), len(x), _p(x), _p(y), _p(z), _p(m),
                               float(m_const), _p(rho))
    if rc:
        raise RuntimeError(f"orc_deposit_cic failed ({rc})")
    return rho


def mass_offset(plo, phi, total_mass):
    _, plo, phi = _geom([1, 1, 1], plo, phi)
    return lib().orc_mass_offset(_p(plo), _p(phi), float(total_mass))


def sum_particle_mass(np_, m=None, m_const=1.0):
    m = None if m is None else _c64(m)
    return lib().orc_sum_particle_mass(int(np_), _p(m), float(m_const))


def correct_rhs(n, plo, phi, rhs, mass_off):
    """In place on ``rhs`` (float64, C-contiguous). Returns the solvability sum."""
    n, plo, phi = _geom(n, plo, phi)
    assert rhs.dtype == np.float64 and rhs.flags.c_contiguous
    return lib().orc_correct_rhs_using_offset(_p(n), _p(plo), _p(phi), _p(rhs), float(mass_off))


def scale_rhs(rhs, four_pi_G, a):
    assert rhs.dtype == np.float64 and rhs.flags.c_contiguous
    lib().orc_scale_rhs(rhs.size, _p(rhs), float(four_pi_G), float(a))


def mg_levels(n):
    n = np.ascontiguousarray(n, dtype=np.int32)
    dims = np.zeros((32, 3), dtype=np.int32)
    L = lib().orc_mg_levels(_p(n), _p(dims))
    return [tuple(int(v) for v in dims[l]) for l in range(L)]


def bottom_green(n, dh):
    n = np.ascontiguousarray(n, dtype=np.int32)
    dh = _c64(dh)
    G = np.empty(grid_shape(n), dtype=np.float64)
    lib().orc_bottom_green(_p(n), _p(dh), _p(G))
    return G


def residual(n, dx, f, phi):
    n = np.ascontiguousarray(n, dtype=np.int32)
    dx = _c64(dx)
    dh = _c64((1.0 / dx) ** 2)
    f, phi = _c64(f), _c64(phi)
    r = np.empty_like(f)
    mx = lib().orc_residual(_p(n), _p(dh), _p(f), _p(phi), _p(r))
    return r, mx


def mg_solve(n, dx, f, phi0=None, rel_tol=1e-10, abs_tol=0.0, max_iter=200, fixed_iters=0,
             nu1=2, nu2=2):
    """Returns (phi, iters, stats[resnorm0, rhsnorm, final, target], rc)."""
    n = np.ascontiguousarray(n, dtype=np.int32)
    dx = _c64(dx)
    f = _c64(f)
    phi = np.zeros_like(f) if phi0 is None else _c64(phi0).copy()
    it = np.zeros(1, dtype=np.int32)
    st = np.zeros(4, dtype=np.float64)
    rc = lib().orc_mg_solve(_p(n), _p(dx), _p(f), _p(phi), float(rel_tol), float(abs_tol),
                            int(max_iter), int(fixed_iters), int(nu1), int(nu2), _p(it), _p(st))
    return phi, int(it[0]), st, rc


def grav_vector(n, plo, phi_hi, phi):
    n, plo, phi_hi = _geom(n, plo, phi_hi)
    phi = _c64(phi)
    g = np.empty((3,) + grid_shape(n), dtype=np.float64)
    lib().orc_grav_vector(_p(n), _p(plo), _p(phi_hi), _p(phi), _p(g))
    return g


def deposit_vels(n, plo, phi, pos, vel, m=None, m_const=1.0):
    """AssignDensityAndVels: (4, nz, ny, nx) = density, mass-weighted vx, vy, vz."""
    n, plo, phi = _geom(n, plo, phi)
    pos = [_c64(p) for p in pos]
    vel = [_c64(v) for v in vel]
    m = None if m is None else _c64(m)
    out = np.empty((4,) + grid_shape(n))
    rc = lib().orc_deposit_cic_vels(_p(n), _p(plo), _p(phi), len(pos[0]), _p(pos[0]), _p(pos[1]), _p(pos[2]),
                                    _p(vel[0]), _p(vel[1]), _p(vel[2]), _p(m), float(m_const), _p(out))
    if rc:
        raise RuntimeError(f"orc_deposit_cic_vels failed ({rc})")
    return out


def est_time_step(n, plo, phi_hi, pos, vel, g, a=1.0, cfl=0.5):
    """NyxParticleContainer::estTimestep with the gravity field g (3, nz, ny, nx)."""
    n, plo, phi_hi = _geom(n, plo, phi_hi)
    pos = [_c64(p) for p in pos]
    vel = [_c64(v) for v in vel]
    g = _c64(g)
    return lib().orc_est_time_step(_p(n), _p(plo), _p(phi_hi), len(pos[0]), _p(pos[0]), _p(pos[1]),
                                   _p(pos[2]), _p(vel[0]), _p(vel[1]), _p(vel[2]), _p(g), float(a),
                                   float(cfl))


def update_particles(n, plo, phi_hi, pos, vel, g, dt, a_prev, a_cur, do_move):
    """In place on pos/vel, each a list of three float64 arrays."""
    n, plo, phi_hi = _geom(n, plo, phi_hi)
    g = _c64(g)
    for a in list(pos) + list(vel):
        assert a.dtype == np.float64 and a.flags.c_contiguous
    lib().orc_update_particles(_p(n), _p(plo), _p(phi_hi), len(pos[0]), _p(pos[0]), _p(pos[1]),
                               _p(pos[2]), _p(vel[0]), _p(vel[1]), _p(vel[2]), _p(g), float(dt),
                               float(a_prev), float(a_cur), int(do_move))


def enforce_periodic(n, plo, phi_hi, pos):
    n, plo, phi_hi = _geom(n, plo, phi_hi)
    lib().orc_enforce_periodic(_p(n), _p(plo), _p(phi_hi), len(pos[0]), _p(pos[0]), _p(pos[1]),
                               _p(pos[2]))


def domain_rhi(lo, hi, n):
    return lib().orc_domain_rhi(float(lo), float(hi), int(n))


def solve_for_phi(n, plo, phi_hi, pos, m=None, m_const=1.0, four_pi_G=1.0, a=1.0,
                  rel_tol=1e-10, abs_tol=0.0, phi0=None, want_rho=False):
    n, plo, phi_hi = _geom(n, plo, phi_hi)
    x, y, z = (_c64(p) for p in pos)
    m = None if m is None else _c64(m)
    phi = np.zeros(grid_shape(n)) if phi0 is None else _c64(phi0).copy()
    rho = np.empty(grid_shape(n)) if want_rho else None
    st = np.zeros(8)
    rc = lib().orc_solve_for_phi(_p(n), _p(plo), _p(phi_hi), len(x), _p(x), _p(y), _p(z), _p(m),
                                 float(m_const), float(four_pi_G), float(a), float(rel_tol),
                                 float(abs_tol), _p(phi), _p(rho), _p(st))
    if rc:
        raise RuntimeError(f"orc_solve_for_phi failed ({rc})")
    return phi, rho, st


def pm_step(n, plo, phi_hi, pos, vel, phi, m=None, m_const=1.0, dt=1e-3, a_old=1.0, a_new=1.0,
            four_pi_G=1.0, rel_tol=1e-10, abs_tol=0.0, nyx_faithful=False, want_g=False):
    """One KDK step in place on pos/vel/phi (float64 C-contiguous arrays)."""
    n, plo, phi_hi = _geom(n, plo, phi_hi)
    m = None if m is None else _c64(m)
    g = np.empty((3,) + grid_shape(n)) if want_g else None
    st = np.zeros(12)
    rc = lib().orc_pm_step(_p(n), _p(plo), _p(phi_hi), len(pos[0]), _p(pos[0]), _p(pos[1]),
                           _p(pos[2]), _p(vel[0]), _p(vel[1]), _p(vel[2]), _p(m), float(m_const),
                           _p(phi), float(dt), float(a_old), float(a_new), float(four_pi_G),
                           float(rel_tol), float(abs_tol), int(bool(nyx_faithful)), _p(g), _p(st))
    if rc:
        raise RuntimeError(f"orc_pm_step failed ({rc})")
    return g, st


def set_num_threads(t):
    lib().orc_set_num_threads(int(t))


def num_threads():
    return lib().orc_num_threads()


# --------------------------------------------------------------------------
# Independent exact solve of the discrete operator (numpy FFT, FP64)
# --------------------------------------------------------------------------

def laplacian_eigenvalues(n, dx):
    """lambda_k = sum_d (2 cos(2 pi k_d / n_d) - 2) / dx_d^2 on the (nz, ny, nx) grid."""
    nx, ny, nz = (int(v) for v in n)
    lx = (2.0 * np.cos(2.0 * np.pi * np.arange(nx) / nx) - 2.0) / dx[0] ** 2
    ly = (2.0 * np.cos(2.0 * np.pi * np.arange(ny) / ny) - 2.0) / dx[1] ** 2
    lz = (2.0 * np.cos(2.0 * np.pi * np.arange(nz) / nz) - 2.0) / dx[2] ** 2
    return lz[:, None, None] + ly[None, :, None] + lx[None, None, :]


def fft_solve(n, dx, f):
    """Mean-free phi with L phi = f - mean(f), L the periodic 7-point Laplacian."""
    lam = laplacian_eigenvalues(n, dx)
    fh = np.fft.fftn(np.asarray(f, dtype=np.float64))
    lam = lam.copy()
    lam[0, 0, 0] = 1.0
    ph = fh / lam
    ph[0, 0, 0] = 0.0
    return np.real(np.fft.ifftn(ph))


def apply_laplacian(n, dx, phi):
    """numpy 7-point periodic Laplacian (independent of the C code)."""
    out = np.zeros_like(phi)
    for ax, d in ((2, 0), (1, 1), (0, 2)):
        out += (np.roll(phi, 1, axis=ax) - 2.0 * phi + np.roll(phi, -1, axis=ax)) / dx[d] ** 2
    return out
