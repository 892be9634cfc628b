"""Seeded synthetic initial conditions for the PM workloads (numpy, FP64).

Every generator draws from numpy's counter-based Philox bit generator keyed by
the seed, so the CPU oracle and the GPU library are fed bit-identical inputs.
Box: periodic unit cube [0,1)^3 unless stated; particle masses are equal and
normalised to mean density 1 (m = 1/N_p).

* ``lattice_gaussian``  -- cell-centred lattice + i.i.d. Gaussian displacements,
  i.i.d. Gaussian velocities (64^3 CPU workload).
* ``grf_lattice``       -- lattice + curl-free Gaussian-random-field displacement
  psi_k = i k Phi_k with displacement power P(k) ~ k^index for 0 < |k| < k_Nyq
  (modes at or beyond the Nyquist frequency, and k = 0, are zero), scaled so that
  the rms of |psi| is ``rms_cells`` cells; velocity = vel_factor * psi.
* ``plummer_clustered`` -- a uniform fraction plus Plummer spheres with uniform
  centres and log-uniform scale radii, radii drawn from the inverse Plummer CDF
  truncated at 16 scale radii, isotropic directions, wrapped periodically.
"""
from __future__ import annotations

import numpy as np

SEED_BASE = 0x4E5958


def rng(seed):
    return np.random.Generator(np.random.Philox(key=int(seed)))


def _wrap(a):
    a = np.mod(a, 1.0)
    a[a >= 1.0] = 0.0
    return a


def lattice_positions(n):
    g = (np.arange(n) + 0.5) / n
    z, y, x = np.meshgrid(g, g, g, indexing="ij")
    return x.ravel(), y.ravel(), z.ravel()


def lattice_gaussian(n, seed, sigma_cells=0.25, vel_sigma=0.01):
    r = rng(seed)
    x, y, z = lattice_positions(n)
    d = r.standard_normal((3, n ** 3)) * (sigma_cells / n)
    pos = [_wrap(x + d[0]), _wrap(y + d[1]), _wrap(z + d[2])]
    vel = list(r.standard_normal((3, n ** 3)) * vel_sigma)
    return pos, vel, 1.0 / n ** 3


def grf_displacement(n, seed, index=-2.0, rms_cells=0.5):
    """Curl-free displacement field (3, n, n, n) in box units."""
    r = rng(seed)
    w = r.standard_normal((n, n, n))
    wk = np.fft.rfftn(w)
    kz = np.fft.fftfreq(n) * n
    ky = np.fft.fftfreq(n) * n
    kx = np.fft.rfftfreq(n) * n
    KZ, KY, KX = np.meshgrid(kz, ky, kx, indexing="ij")
    k2 = KX ** 2 + KY ** 2 + KZ ** 2
    kmag = np.sqrt(k2)
    amp = np.zeros_like(kmag)
    keep = (kmag > 0) & (kmag < n / 2) & (np.abs(KX) < n / 2) & (np.abs(KY) < n / 2) & (np.abs(KZ) < n / 2)
    # |psi_k|^2 = k^2 |Phi_k|^2 ~ k^index  ->  |Phi_k| ~ k^((index-2)/2)
    amp[keep] = kmag[keep] ** ((index - 2.0) / 2.0)
    phik = wk * amp
    psi = np.empty((3, n, n, n))
    for d, K in enumerate((KX, KY, KZ)):
        psi[d] = np.fft.irfftn(1j * K * phik, s=(n, n, n), axes=(0, 1, 2))
    rms = np.sqrt(np.mean(np.sum(psi ** 2, axis=0)))
    psi *= (rms_cells / n) / rms
    return psi


def grf_lattice(n, seed, index=-2.0, rms_cells=0.5, vel_factor=0.1):
    psi = grf_displacement(n, seed, index, rms_cells)
    x, y, z = lattice_positions(n)
    pos = [_wrap(x + psi[0].ravel()), _wrap(y + psi[1].ravel()), _wrap(z + psi[2].ravel())]
    vel = [vel_factor * psi[d].ravel() for d in range(3)]
    return pos, vel, 1.0 / n ** 3


def plummer_clustered(n, seed, n_spheres=2000, frac_uniform=0.2, rs_cells=(1.0, 8.0), trunc=16.0,
                      vel_sigma=0.01, n_part=None, chunk=1 << 24):
    """n^3 particles (or n_part): frac_uniform uniform, the rest in Plummer spheres."""
    N = n ** 3 if n_part is None else int(n_part)
    r = rng(seed)
    n_uni = int(round(frac_uniform * N))
    n_clu = N - n_uni
    centres = r.random((n_spheres, 3))
    rs = np.exp(r.uniform(np.log(rs_cells[0]), np.log(rs_cells[1]), n_spheres)) / n
    # equal shares, the remainder to the first spheres
    per = np.full(n_spheres, n_clu // n_spheres, dtype=np.int64)
    per[: n_clu - per.sum()] += 1
    owner = np.repeat(np.arange(n_spheres, dtype=np.int32), per)
    pos = [np.empty(N), np.empty(N), np.empty(N)]
    u = r.random((3, n_uni))
    for d in range(3):
        pos[d][:n_uni] = u[d]
    umax = trunc ** 3 / (trunc ** 2 + 1.0) ** 1.5  # enclosed mass fraction at trunc*r_s
    for s in range(0, n_clu, chunk):
        e = min(n_clu, s + chunk)
        m = e - s
        o = owner[s:e]
        uu = r.random(m) * umax
        uu = np.maximum(uu, 1e-300)
        rad = rs[o] / np.sqrt(uu ** (-2.0 / 3.0) - 1.0)
        cth = r.uniform(-1.0, 1.0, m)
        ph = r.uniform(0.0, 2.0 * np.pi, m)
        sth = np.sqrt(1.0 - cth ** 2)
        off = (rad * sth * np.cos(ph), rad * sth * np.sin(ph), rad * cth)
        for d in range(3):
            pos[d][n_uni + s:n_uni + e] = _wrap(centres[o, d] + off[d])
    vel = list(r.standard_normal((3, N)) * vel_sigma)
    return pos, vel, 1.0 / N


def config(idx, n=None, seed=None):
    """Inputs of workload idx (0: 64^3 lattice+Gaussian, 1: 256^3 GRF k^-2,
    2: 512^3 Plummer-clustered, 4: 1024^3 GRF k^-1.5) at size n (default the
    workload's own)."""
    seed = SEED_BASE + idx if seed is None else seed
    if idx == 0:
        return lattice_gaussian(n or 64, seed)
    if idx == 1:
        return grf_lattice(n or 256, seed, -2.0, 0.5, 0.1)
    if idx == 2:
        return plummer_clustered(n or 512, seed)
    if idx == 4:
        return grf_lattice(n or 1024, seed, -1.5, 1.0, 0.1)
    raise ValueError(idx)


def grf_lattice_device(n, seed, index=-1.5, rms_cells=1.0, vel_factor=0.1, device="cuda"):
    """grf_lattice generated on the GPU (torch FFTs), for sizes where the host
    path would be slow or memory-bound (1024^3: 6 x 8 GiB of particle data).

    Same construction as grf_displacement + grf_lattice (curl-free psi_k = i k
    Phi_k, |psi_k|^2 ~ k^index for 0 < |k| < k_Nyq, rms |psi| = rms_cells cells,
    v = vel_factor * psi, lattice at cell centres, particles in z-major lattice
    order); the white noise comes from torch's device generator seeded with
    `seed`, so the realisation differs from the numpy one but repeats exactly for
    a given seed on a given build.  Returns ([x, y, z], [vx, vy, vz], m) as
    contiguous FP64 device tensors of n^3 elements."""
    import torch

    g = torch.Generator(device=device)
    g.manual_seed(int(seed))
    w = torch.randn((n, n, n), generator=g, dtype=torch.float64, device=device)
    phik = torch.fft.rfftn(w)
    del w
    f = torch.fft.fftfreq(n, d=1.0 / n, dtype=torch.float64, device=device)
    fr = torch.fft.rfftfreq(n, d=1.0 / n, dtype=torch.float64, device=device)
    KZ, KY, KX = f.view(n, 1, 1), f.view(1, n, 1), fr.view(1, 1, n // 2 + 1)
    kmag = torch.sqrt(KX * KX + KY * KY + KZ * KZ)
    keep = (kmag > 0) & (kmag < n / 2) & (KX.abs() < n / 2) & (KY.abs() < n / 2) & (KZ.abs() < n / 2)
    amp = torch.where(keep, kmag.clamp_min(1e-300) ** ((index - 2.0) / 2.0), torch.zeros_like(kmag))
    del kmag, keep
    phik *= amp
    del amp
    psi = []
    for K in (KX, KY, KZ):
        psi.append(torch.fft.irfftn(phik * (1j * K), s=(n, n, n)))
    del phik
    rms = torch.sqrt(sum((p * p).sum() for p in psi) / n ** 3).item()
    scale = (rms_cells / n) / rms
    lat = (torch.arange(n, dtype=torch.float64, device=device) + 0.5) / n
    views = (lat.view(1, 1, n), lat.view(1, n, 1), lat.view(n, 1, 1))
    pos, vel = [], []
    for d in range(3):
        p = psi[d]
        p *= scale
        x = torch.remainder(p + views[d], 1.0)
        x[x >= 1.0] = 0.0
        pos.append(x.reshape(-1))
        p *= vel_factor
        vel.append(p.reshape(-1))
    psi.clear()
    if str(device).startswith("cuda"):
        torch.cuda.empty_cache()  # hand the FFT temporaries back before the library allocates
    return pos, vel, 1.0 / n ** 3


def plummer_clustered_device(n, seed, n_spheres=2000, frac_uniform=0.2, rs_cells=(1.0, 8.0), trunc=16.0,
                             vel_sigma=0.01, device="cuda"):
    """plummer_clustered generated on the GPU (torch), for the 512^3 workload's
    full-size tests: n^3 particles, frac_uniform of them uniform in the unit box,
    the rest in equal shares over n_spheres Plummer spheres with uniform centres
    and scale radii log-uniform on rs_cells (in cells), radii from the inverse
    Plummer CDF truncated at trunc scale radii, isotropic directions, positions
    wrapped periodically; velocities i.i.d. N(0, vel_sigma^2).  The draws come
    from torch's device generator seeded with `seed` (the realisation differs
    from the numpy one).  Returns ([x, y, z], [vx, vy, vz], m) as FP64 device
    tensors of n^3 elements."""
    import torch

    N = n ** 3
    g = torch.Generator(device=device)
    g.manual_seed(int(seed))
    f64 = torch.float64
    n_uni = int(round(frac_uniform * N))
    n_clu = N - n_uni
    centres = torch.rand((n_spheres, 3), generator=g, dtype=f64, device=device)
    lo, hi = np.log(rs_cells[0]), np.log(rs_cells[1])
    rs = torch.exp(lo + (hi - lo) * torch.rand(n_spheres, generator=g, dtype=f64, device=device)) / n
    per = torch.full((n_spheres,), n_clu // n_spheres, dtype=torch.int64, device=device)
    per[: n_clu - int(per.sum().item())] += 1
    owner = torch.repeat_interleave(torch.arange(n_spheres, device=device), per)
    umax = trunc ** 3 / (trunc ** 2 + 1.0) ** 1.5  # enclosed mass fraction at trunc * r_s
    uu = (torch.rand(n_clu, generator=g, dtype=f64, device=device) * umax).clamp_min(1e-300)
    rad = rs[owner] / torch.sqrt(uu.pow_(-2.0 / 3.0) - 1.0)
    del uu
    cth = torch.rand(n_clu, generator=g, dtype=f64, device=device) * 2.0 - 1.0
    ph = torch.rand(n_clu, generator=g, dtype=f64, device=device) * (2.0 * np.pi)
    sth = torch.sqrt(1.0 - cth * cth)
    off = (rad * sth * torch.cos(ph), rad * sth * torch.sin(ph), rad * cth)
    del rad, cth, ph, sth
    pos = []
    for d in range(3):
        x = torch.empty(N, dtype=f64, device=device)
        x[:n_uni] = torch.rand(n_uni, generator=g, dtype=f64, device=device)
        c = torch.remainder(centres[owner, d] + off[d], 1.0)
        c[c >= 1.0] = 0.0
        x[n_uni:] = c
        del c
        pos.append(x)
    del off, owner
    vel = [torch.randn(N, generator=g, dtype=f64, device=device) * vel_sigma for _ in range(3)]
    if str(device).startswith("cuda"):
        torch.cuda.empty_cache()
    return pos, vel, 1.0 / N
