"""Write the golden vectors in tests/golden/golden.npz.

Expected outputs are computed in exact rational arithmetic (fractions.Fraction on
the exact binary values of the FP64 inputs), i.e. independently of the FP64 C
oracle and of the GPU library.  Because CIC weights are continuous in the
particle position, the exact result differs from any correct FP64 evaluation by
a few ulps only, so the tests compare at ~1e-14 relative.

Inputs:
* ``twoline``  -- the 16 particles of the data file particle_file.2line (a copy
  of Exec/ParticleOnlyTest/particle_file.2line: count, then x y z mass vx vy vz
  per line) in Nyx's box of that test (inputs.2line: prob_lo 0, prob_hi 32,
  n_cell 16).
* ``mass8``    -- the 8 particles of particle_file.mass (a copy of
  Exec/ParticleOnlyTest/particle_file.mass) in the box of inputs.mass.nosub
  (same geometry).
* ``rect``     -- 300 seeded particles in a non-cubic box (8 x 6 x 4 cells,
  prob_lo (-1, 0.5, 2), prob_hi (3, 2, 3)), some slightly outside the box as
  after a drift, random masses; plus a random potential phi on that grid and the
  exact gravity field, interpolated accelerations and kick/drift results.

Run:  python tests/golden/make_golden.py   (writes golden.npz next to this file)
"""
from __future__ import annotations

import os
from fractions import Fraction as F

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))

def parse_ascii(path):
    with open(path) as fh:
        lines = [l.split() for l in fh.read().strip().splitlines()]
    n = int(lines[0][0])
    a = np.array([[float(v) for v in l] for l in lines[1:1 + n]])
    return a[:, 0:3].T.copy(), a[:, 3].copy(), a[:, 4:7].T.copy()


def floor_frac(v):
    i = v.numerator // v.denominator
    return i, v - i


def exact_deposit(n, plo, phi, pos, mass):
    """rho (nz, ny, nx) as floats: CIC with l = (x-plo)/dx + 1/2, weights on
    cells i-1, i; periodic; divided by the exact cell volume."""
    n = [int(v) for v in n]
    plo = [F(v) for v in plo]
    L = [F(phi[d]) - plo[d] for d in range(3)]
    dxi = [F(n[d]) / L[d] for d in range(3)]
    acc = {}
    for p in range(len(mass)):
        m = F(float(mass[p]))
        idx, w = [], []
        for d in range(3):
            l = (F(float(pos[d][p])) - plo[d]) * dxi[d] + F(1, 2)
            i, fr = floor_frac(l)
            idx.append(i)
            w.append((1 - fr, fr))
        for kk in (0, 1):
            for jj in (0, 1):
                for ii in (0, 1):
                    c = ((idx[0] + ii - 1) % n[0], (idx[1] + jj - 1) % n[1], (idx[2] + kk - 1) % n[2])
                    acc[c] = acc.get(c, F(0)) + w[0][ii] * w[1][jj] * w[2][kk] * m
    vol = L[0] * L[1] * L[2] / (n[0] * n[1] * n[2])
    rho = np.zeros((n[2], n[1], n[0]))
    for (i, j, k), v in acc.items():
        rho[k, j, i] = float(v / vol)
    return rho


def exact_grav(n, plo, phi_hi, phi):
    """g_d(c) = 0.5*((phi(c)-phi(c-e_d)) + (phi(c+e_d)-phi(c))) / dx_d, as Fractions."""
    n = [int(v) for v in n]
    dxi = [F(n[d]) / (F(phi_hi[d]) - F(plo[d])) for d in range(3)]
    P = {}
    for k in range(n[2]):
        for j in range(n[1]):
            for i in range(n[0]):
                P[(i, j, k)] = F(float(phi[k, j, i]))
    g = [{}, {}, {}]
    for k in range(n[2]):
        for j in range(n[1]):
            for i in range(n[0]):
                c = (i, j, k)
                for d in range(3):
                    up = list(c); up[d] = (up[d] + 1) % n[d]
                    dn = list(c); dn[d] = (dn[d] - 1) % n[d]
                    g[d][c] = F(1, 2) * ((P[c] - P[tuple(dn)]) * dxi[d] + (P[tuple(up)] - P[c]) * dxi[d])
    return g


def exact_interp(n, plo, phi_hi, g, pos):
    n = [int(v) for v in n]
    plo_f = [F(v) for v in plo]
    dxi = [F(n[d]) / (F(phi_hi[d]) - plo_f[d]) for d in range(3)]
    out = np.zeros((3, len(pos[0])))
    exact = []
    for p in range(len(pos[0])):
        idx, s = [], []
        for d in range(3):
            l = (F(float(pos[d][p])) - plo_f[d]) * dxi[d] + F(1, 2)
            i, fr = floor_frac(l)
            idx.append(i)
            s.append((1 - fr, fr))
        vals = []
        for d in range(3):
            v = F(0)
            for kk in (0, 1):
                for jj in (0, 1):
                    for ii in (0, 1):
                        c = ((idx[0] - ii) % n[0], (idx[1] - jj) % n[1], (idx[2] - kk) % n[2])
                        v += s[0][1 - ii] * s[1][1 - jj] * s[2][1 - kk] * g[d][c]
            vals.append(v)
            out[d, p] = float(v)
        exact.append(vals)
    return out, exact


def main():
    out = {}
    geom32 = dict(n=np.array([16, 16, 16], np.int32), plo=np.zeros(3), phi=np.full(3, 32.0))
    for name, fname in (("twoline", "particle_file.2line"), ("mass8", "particle_file.mass")):
        pos, m, vel = parse_ascii(os.path.join(HERE, fname))
        rho = exact_deposit(geom32["n"], geom32["plo"], geom32["phi"], pos, m)
        out[f"{name}_pos"] = pos
        out[f"{name}_vel"] = vel
        out[f"{name}_mass"] = m
        out[f"{name}_rho"] = rho
    out["box32_n"] = geom32["n"]
    out["box32_plo"] = geom32["plo"]
    out["box32_phi"] = geom32["phi"]

    r = np.random.Generator(np.random.Philox(key=20240611))
    n = np.array([8, 6, 4], np.int32)
    plo = np.array([-1.0, 0.5, 2.0])
    phi_hi = np.array([3.0, 2.0, 3.0])
    L = phi_hi - plo
    NP = 300
    pos = np.empty((3, NP))
    for d in range(3):
        # inside the box, plus some up to 0.3 cell outside on either side
        pos[d] = plo[d] + r.uniform(-0.3 / n[d], 1.0 + 0.3 / n[d], NP) * L[d]
    # the box corner, a cell centre and a cell corner, exactly
    pos[:, 0] = plo + L * np.array([0.0, 0.0, 0.0])
    pos[:, 1] = plo + L * (np.array([3.5, 2.5, 1.5]) / n)
    pos[:, 2] = plo + L * (np.array([4.0, 3.0, 2.0]) / n)
    mass = r.uniform(0.5, 2.0, NP)
    vel = r.standard_normal((3, NP))
    rho = exact_deposit(n, plo, phi_hi, pos, mass)
    phi = r.standard_normal((n[2], n[1], n[0]))
    g = exact_grav(n, plo, phi_hi, phi)
    gcell = np.zeros((3, n[2], n[1], n[0]))
    for d in range(3):
        for (i, j, k), v in g[d].items():
            gcell[d, k, j, i] = float(v)
    gp, gexact = exact_interp(n, plo, phi_hi, g, pos)
    # kick-drift with a_prev, a_cur != 1 (exact): v' = (a_prev v + dt/2 g)/a_cur ; x' = x + dt/a_cur v'
    dt, a_prev, a_cur = 0.0123, 0.8, 0.9
    vnew = np.zeros((3, NP))
    xnew = np.zeros((3, NP))
    for p in range(NP):
        for d in range(3):
            v = (F(a_prev) * F(float(vel[d, p])) + F(dt) / 2 * gexact[p][d]) / F(a_cur)
            vnew[d, p] = float(v)
            xnew[d, p] = float(F(float(pos[d, p])) + F(dt) / F(a_cur) * v)
    out.update(rect_n=n, rect_plo=plo, rect_phi=phi_hi, rect_pos=pos, rect_mass=mass, rect_vel=vel,
               rect_rho=rho, rect_pot=phi, rect_g=gcell, rect_gp=gp, rect_kd_v=vnew, rect_kd_x=xnew,
               rect_kd_params=np.array([dt, a_prev, a_cur]))
    np.savez(os.path.join(HERE, "golden.npz"), **out)
    print("wrote", os.path.join(HERE, "golden.npz"), sorted(out))


if __name__ == "__main__":
    main()
