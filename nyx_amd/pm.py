"""Host-side interface of the MI355X particle-mesh library.

``PMContext`` wraps the C ABI (include/nyxpm.h) one call per entry point.
``DarkMatterParticleContainer`` and ``Gravity`` mirror the Nyx interfaces the
library replaces, with the same method names and argument meaning:

* ``DarkMatterParticleContainer`` -- NyxParticleContainerBase
  (Source/Particle/NyxParticleContainer.H:9-47) as implemented by
  DarkMatterParticleContainer (Source/Particle/DarkMatterParticleContainer.H:7-39).
* ``Gravity`` -- the single-level periodic part of class Gravity
  (Source/Gravity/Gravity.H:12-155).

Error behaviour: where Nyx calls amrex::Abort (e.g. a lost particle in
moveKickDrift, DarkMatterParticleContainer.cpp:173, or an MLMG that fails to
converge), these raise ``NyxPMError``; the process is never terminated.
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass

import numpy as np

from ._lib import (Geom, NyxPMError, Opts, NYXPM_FP64, NYXPM_MIXED, f64, load, ptr)


def _shape(n):
    return (int(n[2]), int(n[1]), int(n[0]))


@dataclass
class BoxGeom:
    """Geometry of one periodic box (amr.n_cell, geometry.prob_lo/hi)."""
    n_cell: tuple
    prob_lo: tuple = (0.0, 0.0, 0.0)
    prob_hi: tuple = (1.0, 1.0, 1.0)

    @property
    def shape(self):
        return _shape(self.n_cell)

    @property
    def ncell(self):
        return int(np.prod(self.n_cell))


def integrate_comoving_a(old_a, dt, omega_m=1.0, omega_r=0.0, h=0.675, comoving_type=1):
    """Nyx::integrate_comoving_a (Source/Driver/comoving.cpp:630-689)."""
    out = ctypes.c_double()
    rc = load().nyxpm_integrate_comoving_a(float(old_a), float(dt), float(omega_m), float(omega_r), float(h),
                                           int(comoving_type), ctypes.byref(out))
    if rc:
        raise NyxPMError(rc, "integrate_comoving_a: invalid arguments")
    return out.value


class PMContext:
    """One library context: a batch of independent periodic boxes on one GPU."""

    def __init__(self, geoms, precision="fp64", tile=None, nu1=2, nu2=2, max_iter=200, smoother="fused",
                 kick="tiled", sweep_shape=None, sweep_zchunk=0, kick_chunk=0, pair_kicks=True,
                 coarse="workgroup", particle_cfl=0.0, resort_slack=0.0, phi_guess="default",
                 residual="separate", graph_cycles=0, ir_first=0, sweep2=0, sweep2_shape=None, verbose=0):
        """resort_slack: free slots per particle each tile gets at a full sort, for
        the incremental re-sort after kick-drifts (0: library default 0.25;
        negative: a full sort after every kick-drift).  phi_guess: initial guess of
        each pm_steps solve, "default" = "previous" (Nyx's) or "extrapolate"
        (2 phi_n - phi_{n-1}).  ir_first: mixed precision, FP32 V-cycles before a
        solve's first FP64 residual test (0: library default; 1: after every
        cycle).  sweep2: FP32 smoothing pairs as one pass (0: on, > 1: on with
        that many z-planes per block, < 0: off); sweep2_shape: its column,
        (32, 16), (64, 16) or (32, 32).  verbose: the library reports to stderr
        (the autotuner's trials)."""
        if isinstance(geoms, BoxGeom):
            geoms = [geoms]
        self.geoms = list(geoms)
        self.L = load()
        arr = (Geom * len(self.geoms))()
        for g, src in zip(arr, self.geoms):
            for d in range(3):
                g.n_cell[d] = int(src.n_cell[d])
                g.prob_lo[d] = float(src.prob_lo[d])
                g.prob_hi[d] = float(src.prob_hi[d])
        o = Opts()
        o.precision = NYXPM_MIXED if precision == "mixed" else NYXPM_FP64
        o.nu1, o.nu2, o.max_iter = int(nu1), int(nu2), int(max_iter)
        o.smoother = 1 if smoother == "half" else 0
        o.kick = 1 if kick == "global" else 0
        if sweep_shape is not None:
            o.sweep_tx, o.sweep_ty = int(sweep_shape[0]), int(sweep_shape[1])
        o.sweep_zchunk = int(sweep_zchunk)
        o.kick_chunk = int(kick_chunk)
        o.separate_kicks = 0 if pair_kicks else 1
        o.coarse = 1 if coarse == "per_level" else 0
        o.particle_cfl = float(particle_cfl)
        o.resort_slack = float(resort_slack)
        o.phi_guess = {"default": 0, "previous": 1, "extrapolate": 2}[phi_guess]
        o.residual = 1 if residual == "in_sweep" else 0
        o.graph_cycles = int(graph_cycles)
        o.ir_first = int(ir_first)
        o.sweep2 = int(sweep2)
        if sweep2_shape is not None:
            o.sweep2_tx, o.sweep2_ty = int(sweep2_shape[0]), int(sweep2_shape[1])
        o.verbose = int(verbose)
        if tile is not None:
            for d in range(3):
                o.tile[d] = int(tile[d])
        h = ctypes.c_void_p()
        rc = self.L.nyxpm_create_batch(ctypes.byref(h), len(self.geoms), ctypes.byref(arr), ctypes.byref(o))
        self.h = h
        if rc:
            msg = self.L.nyxpm_last_error(h).decode() if h.value else "create failed"
            if h.value:
                self.L.nyxpm_destroy(h)
            self.h = None
            raise NyxPMError(rc, msg)
        self.precision = precision

    # -- plumbing -----------------------------------------------------------
    def _ck(self, rc):
        if rc:
            raise NyxPMError(rc, self.L.nyxpm_last_error(self.h).decode())

    def close(self):
        if getattr(self, "h", None):
            self.L.nyxpm_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    @property
    def nbox(self):
        return len(self.geoms)

    def set_precision(self, precision):
        self._ck(self.L.nyxpm_set_precision(self.h, NYXPM_MIXED if precision == "mixed" else NYXPM_FP64))
        self.precision = precision

    def sync(self):
        self._ck(self.L.nyxpm_sync(self.h))

    # -- particles ----------------------------------------------------------
    def upload(self, box, pos, vel, m=None, m_const=1.0):
        x, y, z = (f64(a) for a in pos)
        vx, vy, vz = (f64(a) for a in vel)
        m = None if m is None else f64(m)
        self._ck(self.L.nyxpm_particles_upload(self.h, box, len(x), ptr(x), ptr(y), ptr(z), ptr(vx), ptr(vy),
                                               ptr(vz), ptr(m), float(m_const)))

    def upload_device(self, box, np_, ptrs, m_ptr=None, m_const=1.0):
        """ptrs: six device addresses (x, y, z, vx, vy, vz)."""
        self._ck(self.L.nyxpm_particles_upload_device(self.h, box, int(np_), *[ctypes.c_void_p(p) for p in ptrs],
                                                      ctypes.c_void_p(m_ptr) if m_ptr else None, float(m_const)))

    def load_binary(self, path, box=0):
        """Nyx's binary particle file (InitFromBinaryFile) into box `box`; returns NP."""
        n = ctypes.c_int64()
        self._ck(self.L.nyxpm_particles_load_binary(self.h, box, str(path).encode(), ctypes.byref(n)))
        return n.value

    def load_ascii(self, path, box=0):
        """Nyx's ASCII particle file (InitFromAsciiFile) into box `box`; returns NP."""
        n = ctypes.c_int64()
        self._ck(self.L.nyxpm_particles_load_ascii(self.h, box, str(path).encode(), ctypes.byref(n)))
        return n.value

    def write_amrex(self, dir, name="DM", checkpoint=True, box=0):
        """AMReX particle checkpoint (NyxCheckpoint) or plotfile (WriteNyxPlotFile)."""
        self._ck(self.L.nyxpm_particles_write_amrex(self.h, box, str(dir).encode(), str(name).encode(),
                                                    int(bool(checkpoint))))

    def read_amrex(self, dir, name="DM", box=0):
        """Restart from an AMReX particle checkpoint or plotfile; returns NP."""
        n = ctypes.c_int64()
        self._ck(self.L.nyxpm_particles_read_amrex(self.h, box, str(dir).encode(), str(name).encode(),
                                                   ctypes.byref(n)))
        return n.value

    def num_particles(self, box=0):
        return int(self.L.nyxpm_num_particles(self.h, box))

    def download_device(self, addrs, box=0):
        """Particles of box `box` in upload order into six device arrays
        (x, y, z, vx, vy, vz addresses; None skips a component)."""
        self._ck(self.L.nyxpm_particles_download(self.h, box, *[ctypes.c_void_p(int(a)) if a else None
                                                                for a in addrs]))

    def download(self, box=0):
        n = self.num_particles(box)
        out = [np.empty(n) for _ in range(6)]
        self._ck(self.L.nyxpm_particles_download(self.h, box, *[ptr(a) for a in out]))
        return out[:3], out[3:]

    # -- AssignDensity / mass ------------------------------------------------
    def assign_density(self, box=0, want=True):
        if not want:
            self._ck(self.L.nyxpm_assign_density(self.h, -1, None))
            return None
        rho = np.empty(self.geoms[box].shape)
        self._ck(self.L.nyxpm_assign_density(self.h, box, ptr(rho)))
        return rho

    def assign_density_and_vels(self, box=0):
        """(4, nz, ny, nx): CIC density and mass-weighted cell velocity (vx, vy, vz)."""
        out = np.empty((4,) + self.geoms[box].shape)
        self._ck(self.L.nyxpm_assign_density_and_vels(self.h, box, ptr(out)))
        return out

    def sum_particle_mass(self, box=0):
        v = ctypes.c_double()
        self._ck(self.L.nyxpm_sum_particle_mass(self.h, box, ctypes.byref(v)))
        return v.value

    # -- gravity ------------------------------------------------------------
    def solve_for_phi(self, a=1.0, four_pi_G=1.0, rel_tol=1e-10, abs_tol=0.0, use_guess=True):
        it = np.zeros(self.nbox, dtype=np.int32)
        res = np.zeros(self.nbox)
        self._ck(self.L.nyxpm_solve_for_phi(self.h, float(a), float(four_pi_G), float(rel_tol), float(abs_tol),
                                            int(bool(use_guess)), ptr(it), ptr(res)))
        return it, res

    def solve_poisson(self, rhs, phi0=None, rel_tol=1e-10, abs_tol=0.0, fixed_iters=0, box=0):
        rhs = f64(rhs)
        phi = np.zeros_like(rhs) if phi0 is None else f64(phi0).copy()
        it = ctypes.c_int()
        norms = np.zeros(4)
        rc = self.L.nyxpm_solve_poisson(self.h, box, ptr(rhs), ptr(phi), float(rel_tol), float(abs_tol),
                                        int(fixed_iters), ctypes.byref(it), ptr(norms))
        self._ck(rc)
        return phi, it.value, norms

    def grav_vector(self, box=0):
        g = np.empty((3,) + self.geoms[box].shape)
        self._ck(self.L.nyxpm_get_grav_vector(self.h, box, ptr(g)))
        return g

    def gather_accel(self, box=0):
        """Accelerations at the particles (3, np), upload order."""
        g = np.empty((3, self.num_particles(box)))
        self._ck(self.L.nyxpm_gather_accel(self.h, box, ptr(g)))
        return g

    def gather_accel_noalloc(self, box=0):
        """gather_accel through nyxpm_gather_accel_noalloc (no device allocation;
        at most 2^20 particles in the box)."""
        g = np.empty((3, self.num_particles(box)))
        self._ck(self.L.nyxpm_gather_accel_noalloc(self.h, box, ptr(g)))
        return g

    def est_time_step(self, a=1.0, cfl=0.5):
        """NyxParticleContainer::estTimestep per box (array of nbox; -1e50 for a box
        without particles)."""
        out = np.zeros(self.nbox)
        self._ck(self.L.nyxpm_est_time_step(self.h, float(a), float(cfl), ptr(out)))
        return out

    # -- device-resident variants: `addr` is a device address (e.g. a torch
    #    tensor's data_ptr()) of an FP64 array of the right length; the library
    #    reads or writes it in place (unified addressing).
    def gather_accel_device(self, addr, box=0):
        self._ck(self.L.nyxpm_gather_accel(self.h, box, ctypes.c_void_p(int(addr))))

    def get_phi_device(self, addr, box=0):
        self._ck(self.L.nyxpm_get_phi(self.h, box, ctypes.c_void_p(int(addr))))

    def set_phi_device(self, addr, box=0):
        self._ck(self.L.nyxpm_set_phi(self.h, box, ctypes.c_void_p(int(addr))))

    def assign_density_device(self, addr, box=0):
        self._ck(self.L.nyxpm_assign_density(self.h, box, ctypes.c_void_p(int(addr))))

    def get_phi(self, box=0):
        a = np.empty(self.geoms[box].shape)
        self._ck(self.L.nyxpm_get_phi(self.h, box, ptr(a)))
        return a

    def set_phi(self, phi, box=0):
        self._ck(self.L.nyxpm_set_phi(self.h, box, ptr(f64(phi))))

    def set_density(self, rho, box=0):
        self._ck(self.L.nyxpm_set_density(self.h, box, ptr(f64(rho))))

    def get_rhs(self, box=0):
        a = np.empty(self.geoms[box].shape)
        self._ck(self.L.nyxpm_get_rhs(self.h, box, ptr(a)))
        return a

    # -- particle update ----------------------------------------------------
    def move_kick_drift(self, dt, a_old=1.0, a_half=1.0):
        self._ck(self.L.nyxpm_move_kick_drift(self.h, float(dt), float(a_old), float(a_half)))

    def move_kick(self, dt, a_new=1.0, a_half=1.0):
        self._ck(self.L.nyxpm_move_kick(self.h, float(dt), float(a_new), float(a_half)))

    def redistribute(self):
        self._ck(self.L.nyxpm_redistribute(self.h))

    # -- whole step ---------------------------------------------------------
    def pm_init(self, a=1.0, four_pi_G=1.0, rel_tol=1e-10, abs_tol=0.0):
        self._ck(self.L.nyxpm_pm_init(self.h, float(a), float(four_pi_G), float(rel_tol), float(abs_tol)))

    def pm_steps(self, nsteps, dt, a_old=1.0, a_new=1.0, four_pi_G=1.0, rel_tol=1e-10, abs_tol=0.0,
                 nyx_faithful=False, use_graph=False):
        it = np.zeros(self.nbox, dtype=np.int32)
        self._ck(self.L.nyxpm_pm_steps(self.h, int(nsteps), float(dt), float(a_old), float(a_new),
                                       float(four_pi_G), float(rel_tol), float(abs_tol), int(bool(nyx_faithful)),
                                       int(bool(use_graph)), ptr(it)))
        return it

    # -- tuning ---------------------------------------------------------------
    def pm_steps_comoving(self, nsteps, dt, a, omega_m=1.0, omega_r=0.0, h=0.675, comoving_type=1,
                          four_pi_G=1.0, rel_tol=1e-10, abs_tol=0.0, nyx_faithful=False):
        """nsteps KDK steps with a(t) integrated as Nyx does; returns (a_end, iters)."""
        it = np.zeros(self.nbox, dtype=np.int32)
        av = ctypes.c_double(a)
        self._ck(self.L.nyxpm_pm_steps_comoving(self.h, int(nsteps), float(dt), ctypes.byref(av), float(omega_m),
                                                float(omega_r), float(h), int(comoving_type), float(four_pi_G),
                                                float(rel_tol), float(abs_tol), int(bool(nyx_faithful)), ptr(it)))
        return av.value, it

    def autotune(self, fp64=True, fp32=True, particles=True):
        """Time every fused-sweep column shape and z-chunk per multigrid depth on
        this context's levels, and (particles, with particles loaded) the particle
        tile for the deposit and kicks, and keep the fastest; returns the choices.
        The smoother shapes give the same bits; the particle tile changes the
        density's last bit at tile boundaries (as the tile option does)."""
        import json
        buf = ctypes.create_string_buffer(8192)
        mask = (1 if fp64 else 0) | (2 if fp32 else 0) | (4 if particles else 0)
        self._ck(self.L.nyxpm_autotune(self.h, mask, buf, len(buf)))
        return json.loads(buf.value.decode())

    # -- measurement --------------------------------------------------------
    def timing_enable(self, on=True):
        self._ck(self.L.nyxpm_timing_enable(self.h, int(bool(on))))

    def timing_select(self, names=None):
        """Time only the named stages (an iterable or comma-separated string of
        stage names); None times every stage."""
        if names is not None and not isinstance(names, str):
            names = ",".join(names)
        self._ck(self.L.nyxpm_timing_select(self.h, names.encode() if names else None))

    def timing_reset(self):
        self._ck(self.L.nyxpm_timing_reset(self.h))

    def timing(self, name):
        n = ctypes.c_int64()
        mean = ctypes.c_double()
        tot = ctypes.c_double()
        self._ck(self.L.nyxpm_timing_get(self.h, name.encode(), ctypes.byref(n), ctypes.byref(mean),
                                         ctypes.byref(tot)))
        return n.value, mean.value, tot.value

    def device_bytes(self):
        return int(self.L.nyxpm_device_bytes(self.h))

    def sort_stats(self):
        """{'full', 'incremental', 'moved', 'holes'}: re-sort counters since creation."""
        out = np.zeros(4, dtype=np.int64)
        self._ck(self.L.nyxpm_sort_stats(self.h, ptr(out)))
        return dict(zip(("full", "incremental", "moved", "holes"), (int(v) for v in out)))


# ---------------------------------------------------------------------------
# Nyx-shaped mirror
# ---------------------------------------------------------------------------

class DarkMatterParticleContainer:
    """NyxParticleContainerBase / DarkMatterParticleContainer on the library.

    Particles live in HBM inside ``ctx``; level is always 0 (one periodic level).
    """

    def __init__(self, ctx: PMContext, box: int = 0):
        self.ctx = ctx
        self.box = box

    def InitFromArrays(self, pos, vel, mass=None, mass_const=1.0):
        """Stands in for the Ascii/Binary particle readers (NyxParticles.cpp:379-744)."""
        self.ctx.upload(self.box, pos, vel, mass, mass_const)

    def InitFromAsciiFile(self, file, extradata=4):
        """NyxParticleContainer's InitFromAsciiFile(file, AMREX_SPACEDIM + 1)."""
        if extradata != 4:
            raise ValueError("extradata must be AMREX_SPACEDIM + 1 (mass and velocity)")
        return self.ctx.load_ascii(file, self.box)

    def NyxCheckpoint(self, dir, name="DM"):
        """NyxParticleContainer::NyxCheckpoint (NyxParticleContainer.H:508-517)."""
        self.ctx.write_amrex(dir, name, True, self.box)

    def WriteNyxPlotFile(self, dir, name="DM"):
        """NyxParticleContainer::WriteNyxPlotFile (NyxParticleContainer.H:497-505)."""
        self.ctx.write_amrex(dir, name, False, self.box)

    def Restart(self, dir, name="DM"):
        """ParticleContainer::Restart as Nyx calls it (NyxParticles.cpp:586)."""
        return self.ctx.read_amrex(dir, name, self.box)

    def InitFromBinaryFile(self, file, extradata=4):
        """NyxParticleContainer's InitFromBinaryFile(file, AMREX_SPACEDIM + 1)."""
        if extradata != 4:
            raise ValueError("extradata must be AMREX_SPACEDIM + 1 (mass and velocity)")
        return self.ctx.load_binary(file, self.box)

    def AssignDensitySingleLevel(self, level=0, ncomp=1, particle_lvl_offset=0):
        """NyxParticleContainer.H:90-93: CIC mass density (x 1/vol), periodic."""
        if level != 0 or ncomp != 1:
            raise NotImplementedError("one level, ncomp=1")
        return self.ctx.assign_density(self.box)

    def AssignDensityAndVels(self, lev_min=0):
        """DarkMatterParticleContainer::AssignDensityAndVels: [density, vx, vy, vz]
        of the single level, as a list with one (4, nz, ny, nx) array."""
        return [self.ctx.assign_density_and_vels(self.box)]

    def AssignDensity(self, lev_min=0, ncomp=1, finest_level=-1, ngrow=1):
        """NyxParticleContainer.H:94-97; returns [rho] for the single level."""
        return [self.AssignDensitySingleLevel(lev_min, ncomp)]

    def moveKickDrift(self, acceleration=None, level=0, timestep=0.0, a_old=1.0, a_half=1.0, where_width=0):
        """DarkMatterParticleContainer.cpp:75-181.  The acceleration is the cell-
        centred gradient of the context's current phi, computed inside the fused
        kernel; an explicit field is not accepted."""
        if acceleration is not None:
            raise NotImplementedError("acceleration comes from the context's phi (fused gradient)")
        self.ctx.move_kick_drift(timestep, a_old, a_half)

    def moveKick(self, acceleration=None, level=0, timestep=0.0, a_new=1.0, a_half=1.0):
        """DarkMatterParticleContainer.cpp:183-247 (plus the periodic shift)."""
        if acceleration is not None:
            raise NotImplementedError("acceleration comes from the context's phi (fused gradient)")
        self.ctx.move_kick(timestep, a_new, a_half)

    def Redistribute(self, lev_min=0, lev_max=-1, nGrow=0):
        """Periodic shift of positions (NyxParticleContainer.H:114-121)."""
        self.ctx.redistribute()

    RedistributeLocal = Redistribute

    def estTimestep(self, acceleration=None, a=1.0, lev=0, cfl=0.5):
        """NyxParticleContainer::estTimestep (NyxParticleContainer.H:327-410).  The
        acceleration argument is not read: the gravity of the context's phi is
        used, as Nyx passes get_new_data(Gravity_Type).  -1e50 without particles."""
        return float(self.ctx.est_time_step(a, cfl)[self.box])

    def sumParticleMass(self, level=0):
        """NyxParticleContainer.H:70-79."""
        return self.ctx.sum_particle_mass(self.box)

    def GetParticles(self):
        """(pos, vel) host copies in upload order."""
        return self.ctx.download(self.box)


class Gravity:
    """Single-level periodic Gravity (Source/Gravity/Gravity.H) on the library.

    ``four_pi_G`` replaces the compile-time Gconst (Ggravity = -four_pi_G,
    Gravity.cpp:104); sl_tol/ml_tol are the gravity.sl_tol/ml_tol inputs.
    """

    def __init__(self, ctx: PMContext, four_pi_G=1.0, sl_tol=1e-12, ml_tol=1e-12, box=0):
        self.ctx = ctx
        self.four_pi_G = four_pi_G
        self.sl_tol = sl_tol
        self.ml_tol = ml_tol
        self.box = box
        self.level_solver_resnorm = [0.0]
        self.last_iters = None

    def set_mass_offset(self, time=0.0):
        """Gravity.cpp:975-1030: sum(m)/ProbSize (kept per box by the library)."""
        g = self.ctx.geoms[self.box]
        size = np.prod(np.asarray(g.prob_hi) - np.asarray(g.prob_lo))
        return self.ctx.sum_particle_mass(self.box) / size

    def solve_for_phi(self, level=0, a=1.0, use_previous_phi_as_guess=True, rel_tol=None):
        """Gravity.cpp:267-315 on the density of the last AssignDensity: mean
        removal, x Ggravity, x 1/a, MLMG solve.  Returns the final residual norm."""
        tol = self.sl_tol if rel_tol is None else rel_tol
        it, res = self.ctx.solve_for_phi(a, self.four_pi_G, tol, 0.0, use_previous_phi_as_guess)
        self.last_iters = it
        self.level_solver_resnorm[0] = float(res[self.box])
        return float(res[self.box])

    def solve_for_new_phi(self, particles: DarkMatterParticleContainer, level=0, a=1.0):
        """Gravity.cpp:234-265: AddParticlesToRhs + solve_for_phi (guess = phi)."""
        particles.ctx.assign_density(want=False)
        return self.solve_for_phi(level, a, True, self.sl_tol)

    def solve_for_old_phi(self, particles: DarkMatterParticleContainer, level=0, a=1.0):
        """Gravity.cpp:202-232 as called from Nyx_advance_particles.cpp:168 (zero guess)."""
        particles.ctx.assign_density(want=False)
        return self.solve_for_phi(level, a, False, self.ml_tol)

    def get_new_grav_vector(self, level=0, time=0.0):
        """Gravity.cpp:817-850: cell-centred average of face gradients of phi."""
        return self.ctx.grav_vector(self.box)

    get_old_grav_vector = get_new_grav_vector

    def get_phi(self):
        return self.ctx.get_phi(self.box)


def advance_particles_only(particles: DarkMatterParticleContainer, gravity: Gravity, dt, a_old=1.0, a_new=1.0,
                           nyx_faithful=False):
    """Nyx_advance_particles.cpp:137-268 + post_timestep redistribute, composed
    from the mirror calls (the fused library step is PMContext.pm_steps)."""
    a_half = 0.5 * (a_old + a_new)
    if nyx_faithful:
        gravity.solve_for_old_phi(particles, 0, a_old)
    particles.moveKickDrift(None, 0, dt, a_old, a_half)
    gravity.solve_for_new_phi(particles, 0, a_new)
    particles.moveKick(None, 0, dt, a_new, a_half)
