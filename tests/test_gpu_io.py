"""GPU: Nyx's binary particle file (InitFromBinaryFile) through the C ABI.

tests/golden/32.nyx is the reference's Exec/LyA/32.nyx, unchanged (32^3 particles
in an 8 Mpc box: int64 NP, int32 DM = 3, int32 NX = 4, then x y z m vx vy vz in
FP64 per particle).  numpy reads the same bytes independently as the check."""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
FILE = os.path.join(HERE, "golden", "32.nyx")


def _numpy_read(path, dtype=np.float64):
    with open(path, "rb") as fh:
        np_ = int(np.frombuffer(fh.read(8), np.int64)[0])
        dm, nx = (int(v) for v in np.frombuffer(fh.read(8), np.int32))
        a = np.frombuffer(fh.read(), dtype).reshape(np_, dm + nx).astype(np.float64)
    return a


def _write(path, a, dtype, nx=4):
    with open(path, "wb") as fh:
        fh.write(np.int64(a.shape[0]).tobytes())
        fh.write(np.array([3, nx], np.int32).tobytes())
        fh.write(np.ascontiguousarray(a, dtype=dtype).tobytes())


def test_load_32nyx_matches_file(nyx):
    a = _numpy_read(FILE)
    assert a.shape == (32768, 7)
    with nyx.PMContext(nyx.BoxGeom((32, 32, 32), (0, 0, 0), (8, 8, 8))) as c:
        pc = nyx.DarkMatterParticleContainer(c)
        assert pc.InitFromBinaryFile(FILE, 4) == 32768
        p, v = c.download()
        for d in range(3):
            assert np.array_equal(p[d], a[:, d])
            assert np.array_equal(v[d], a[:, 4 + d])
        assert c.sum_particle_mass() == pytest.approx(a[:, 3].sum(), rel=1e-15)


def test_load_fp32_and_extra_columns(nyx, tmp_path):
    r = np.random.default_rng(3)
    a = np.column_stack([r.random((500, 3)) * 8, np.full(500, 2.5), r.normal(0, 1, (500, 3)),
                         r.normal(0, 1, (500, 2))])
    f32 = tmp_path / "p32.nyx"
    _write(f32, a, np.float32, nx=6)  # two extra columns after the velocity are skipped
    with nyx.PMContext(nyx.BoxGeom((16, 16, 16), (0, 0, 0), (8, 8, 8))) as c:
        assert c.load_binary(f32) == 500
        p, v = c.download()
        ref = a.astype(np.float32).astype(np.float64)
        for d in range(3):
            assert np.array_equal(p[d], ref[:, d])
            assert np.array_equal(v[d], ref[:, 4 + d])


def test_load_errors(nyx, tmp_path):
    bad = tmp_path / "bad.nyx"
    with open(bad, "wb") as fh:
        fh.write(np.int64(10).tobytes() + np.array([3, 4], np.int32).tobytes() + b"\0" * 100)
    with nyx.PMContext(nyx.BoxGeom((16, 16, 16))) as c:
        with pytest.raises(nyx.NyxPMError):
            c.load_binary(bad)
        with pytest.raises(nyx.NyxPMError):
            c.load_binary(tmp_path / "missing.nyx")
        two_d = tmp_path / "dm2.nyx"
        with open(two_d, "wb") as fh:
            fh.write(np.int64(1).tobytes() + np.array([2, 4], np.int32).tobytes() + b"\0" * 48)
        with pytest.raises(nyx.NyxPMError):
            c.load_binary(two_d)


def test_32nyx_deposit_and_step_vs_oracle(nyx, orc):
    # the LyA particle set, per-particle masses, one canonical KDK step with
    # 4 pi G = 1e-10 (so that the step stays small in these units) against the oracle
    a = _numpy_read(FILE)
    n, hi = (32, 32, 32), (8.0, 8.0, 8.0)
    pos = [a[:, d].copy() for d in range(3)]
    vel = [a[:, 4 + d].copy() for d in range(3)]
    m = a[:, 3].copy()
    with nyx.PMContext(nyx.BoxGeom(n, (0, 0, 0), hi)) as c:
        c.load_binary(FILE)
        rho = c.assign_density()
        ref = orc.deposit(n, (0, 0, 0), hi, pos[0], pos[1], pos[2], m=m)
        assert np.abs(rho - ref).max() <= 1e-13 * np.abs(ref).max()
        c.pm_init(1.0, 1e-10, 1e-10)
        c.pm_steps(1, 1e-3, 1.0, 1.0, four_pi_G=1e-10, rel_tol=1e-10)
        p, v = c.download()
    phi, _, _ = orc.solve_for_phi(n, (0, 0, 0), hi, pos, m=m, four_pi_G=1e-10, rel_tol=1e-10)
    orc.pm_step(n, (0, 0, 0), hi, pos, vel, phi, m=m, four_pi_G=1e-10, dt=1e-3, rel_tol=1e-10)
    for d in range(3):
        assert np.abs(p[d] - pos[d]).max() <= 1e-12 * 8
        assert np.abs(v[d] - vel[d]).max() <= 1e-12 * np.abs(vel[d]).max()


# ---------------------------------------------------------------------------
# ASCII particle files (InitFromAsciiFile)
# ---------------------------------------------------------------------------
# tests/golden/ic_sb_32.ascii is the reference's Exec/MiniSB/ic_sb_32.ascii,
# unchanged (32,686 particles of the Santa Barbara initial conditions in a
# 64 Mpc box of 32^3 cells, inputs.32: prob_hi 64, n_cell 32); numpy parses the
# same text independently as the check.

SB = os.path.join(HERE, "golden", "ic_sb_32.ascii")
TWOLINE = os.path.join(HERE, "golden", "particle_file.2line")


def _numpy_ascii(path):
    with open(path) as fh:
        tok = fh.read().split()
    n = int(tok[0])
    a = np.array([float(t) for t in tok[1:1 + 7 * n]]).reshape(n, 7)
    return a


def test_load_ascii_sb32_matches_file_and_oracle(nyx, orc):
    a = _numpy_ascii(SB)
    assert a.shape == (32686, 7)
    n, hi = (32, 32, 32), (64.0, 64.0, 64.0)
    with nyx.PMContext(nyx.BoxGeom(n, (0, 0, 0), hi)) as c:
        pc = nyx.DarkMatterParticleContainer(c)
        assert pc.InitFromAsciiFile(SB, 4) == 32686
        p, v = c.download()
        rho = c.assign_density()
    inside = np.all((a[:, :3] >= 0) & (a[:, :3] < 64.0), axis=1)
    assert inside.all()  # the file's particles lie in the box: read unchanged
    for d in range(3):
        assert np.array_equal(p[d], a[:, d])
        assert np.array_equal(v[d], a[:, 4 + d])
    ref = orc.deposit(n, (0, 0, 0), hi, a[:, 0], a[:, 1], a[:, 2], m=a[:, 3])
    assert np.abs(rho - ref).max() <= 1e-13 * np.abs(ref).max()


def test_load_ascii_twoline_matches_golden(nyx, golden):
    with nyx.PMContext(nyx.BoxGeom(tuple(golden["box32_n"]), tuple(golden["box32_plo"]),
                                   tuple(golden["box32_phi"]))) as c:
        assert c.load_ascii(TWOLINE) == 16
        p, v = c.download()
        rho = c.assign_density()
        assert c.sum_particle_mass() == pytest.approx(golden["twoline_mass"].sum(), rel=1e-15)
    for d in range(3):
        assert np.array_equal(p[d], golden["twoline_pos"][d])
        assert np.array_equal(v[d], golden["twoline_vel"][d])
    assert np.abs(rho - golden["twoline_rho"]).max() <= 1e-14 * np.abs(golden["twoline_rho"]).max()


def test_load_ascii_shifts_outside_particles_and_errors(nyx, orc, tmp_path):
    # particles within one period outside the box come back shifted into it
    # (the periodic shift the oracle restates); a short or malformed file is an
    # error, not a crash
    r = np.random.default_rng(8)
    a = np.column_stack([r.uniform(-0.5, 1.5, (200, 3)), np.full(200, 0.25), r.normal(0, 1, (200, 3))])
    f = tmp_path / "p.ascii"
    with open(f, "w") as fh:
        fh.write("200\n" + "\n".join(" ".join(repr(float(x)) for x in row) for row in a) + "\n")
    with nyx.PMContext(nyx.BoxGeom((8, 8, 8))) as c:
        assert c.load_ascii(f) == 200
        p, v = c.download()
        ref = [a[:, d].copy() for d in range(3)]
        orc.enforce_periodic((8, 8, 8), (0, 0, 0), (1, 1, 1), ref)
        for d in range(3):
            assert np.array_equal(p[d], ref[d])
            assert np.array_equal(v[d], a[:, 4 + d])
        short = tmp_path / "short.ascii"
        short.write_text("3\n0.1 0.2 0.3 1 0 0 0\n0.5 0.5\n")
        with pytest.raises(nyx.NyxPMError):
            c.load_ascii(short)
        bad = tmp_path / "bad.ascii"
        bad.write_text("1\n0.1 0.2 zz 1 0 0 0\n")
        with pytest.raises(nyx.NyxPMError):
            c.load_ascii(bad)
        with pytest.raises(nyx.NyxPMError):
            c.load_ascii(tmp_path / "missing.ascii")


# ---------------------------------------------------------------------------
# AMReX particle checkpoint / plotfile (NyxCheckpoint, WriteNyxPlotFile, Restart)
# ---------------------------------------------------------------------------

def _read_amrex_particles(base):
    """Independent numpy reader of an AMReX particle directory, following the
    layout the reference's own converter reads (Util/Converters/
    plotfile2HDF5_particles_amr.py: Header fields in order, then per grid
    "file count offset"; a checkpoint's int components (id, cpu) before the
    reals of each grid; real type from the version string's last "_" token)."""
    with open(os.path.join(base, "Header")) as fh:
        lines = [l.strip() for l in fh.read().splitlines()]
    q = 0
    version = lines[q]; q += 1
    dim = int(lines[q]); q += 1
    nre = int(lines[q]); q += 1
    rnames = lines[q:q + nre]; q += nre
    nie = int(lines[q]); q += 1
    inames = lines[q:q + nie]; q += nie
    chk = bool(int(lines[q])); q += 1
    npart = int(lines[q]); q += 1
    next_id = int(lines[q]); q += 1
    finest = int(lines[q]); q += 1
    ngr = [int(lines[q + l]) for l in range(finest + 1)]; q += finest + 1
    grids = [tuple(int(t) for t in lines[q + g].split()) for g in range(sum(ngr))]
    rt = np.float64 if version.split("_")[-1] == "double" else np.float32
    nint, nreal = (2 + nie if chk else 0), dim + nre
    ints, reals = [], []
    for which, count, where in grids:
        if count == 0:
            continue
        with open(os.path.join(base, "Level_0", "DATA_%05d" % which), "rb") as fh:
            fh.seek(where)
            if chk:
                ints.append(np.frombuffer(fh.read(4 * nint * count), np.int32).reshape(count, nint))
            reals.append(np.frombuffer(fh.read(np.dtype(rt).itemsize * nreal * count), rt).reshape(count, nreal))
    return {"version": version, "dim": dim, "rnames": rnames, "inames": inames, "checkpoint": chk,
            "np": npart, "next_id": next_id, "finest": finest, "grids": grids,
            "ints": np.concatenate(ints) if ints else None, "reals": np.concatenate(reals)}


def test_checkpoint_and_plotfile_round_trip(nyx, tmp_path):
    # the state after a few steps (per-particle masses) written as a checkpoint
    # and as a plotfile, read back by the independent reader and by Restart
    pos, vel, m = nyx.ics.plummer_clustered(16, 31, n_spheres=3)
    mm = m * np.random.default_rng(31).uniform(0.5, 1.5, len(pos[0]))
    n = (16, 16, 16)
    with nyx.PMContext(nyx.BoxGeom(n)) as c:
        c.upload(0, pos, vel, m=mm)
        c.pm_init(1.0, 1.0, 1e-10)
        c.pm_steps(3, 2e-3, rel_tol=1e-10)
        p, v = c.download()
        pc = nyx.DarkMatterParticleContainer(c)
        pc.NyxCheckpoint(str(tmp_path / "chk00003"), "DM")
        pc.WriteNyxPlotFile(str(tmp_path / "plt00003"), "DM")
    npart = len(pos[0])
    for kind, chk in (("chk00003", True), ("plt00003", False)):
        r = _read_amrex_particles(str(tmp_path / kind / "DM"))
        assert r["version"].endswith("_double") and r["dim"] == 3
        assert r["rnames"] == ["mass", "xvel", "yvel", "zvel"] and r["inames"] == []
        assert r["checkpoint"] == chk and r["np"] == npart and r["next_id"] == npart + 1 and r["finest"] == 0
        R = r["reals"]
        for d in range(3):
            assert np.array_equal(R[:, d], p[d])
            assert np.array_equal(R[:, 4 + d], v[d])
        assert np.array_equal(R[:, 3], mm)
        if chk:
            assert np.array_equal(r["ints"][:, 0], np.arange(1, npart + 1)) and not r["ints"][:, 1].any()
        else:
            assert (tmp_path / kind / "DM" / "precision").read_text().strip() == "NATIVE"
        ph = (tmp_path / kind / "DM" / "Level_0" / "Particle_H").read_text().split()
        assert ph[0] == "(1" and "(15,15,15)" in ph
        with nyx.PMContext(nyx.BoxGeom(n)) as c2:
            assert nyx.DarkMatterParticleContainer(c2).Restart(str(tmp_path / kind), "DM") == npart
            p2, v2 = c2.download()
            assert c2.sum_particle_mass() == pytest.approx(mm.sum(), rel=1e-15)
        for d in range(3):
            assert np.array_equal(p2[d], p[d])
            assert np.array_equal(v2[d], v[d])


def test_restart_from_several_grids_in_single_precision(nyx, tmp_path):
    # a checkpoint as AMReX lays out a level split into grids over two DATA
    # files, FP32 reals and one extra int component, written here with numpy;
    # Restart reads the grids in Header order
    r = np.random.default_rng(12)
    npart = 1000
    R = np.column_stack([r.random((npart, 3)), r.uniform(1, 2, npart), r.normal(0, 1, (npart, 3)),
                         r.random(npart)]).astype(np.float32)  # one more real after zvel
    base = tmp_path / "chk" / "DM"
    (base / "Level_0").mkdir(parents=True)
    counts = [300, 0, 450, 250]
    files = [0, 0, 1, 1]
    offsets, blobs = [], {0: b"", 1: b""}
    start = 0
    for g, cnt in enumerate(counts):
        f = files[g]
        offsets.append(len(blobs[f]))
        ints = np.zeros((cnt, 3), np.int32)
        ints[:, 0] = np.arange(start + 1, start + cnt + 1)
        blobs[f] += ints.tobytes() + R[start:start + cnt].tobytes()
        start += cnt
    for f, b in blobs.items():
        (base / "Level_0" / ("DATA_%05d" % f)).write_bytes(b)
    hdr = ["Version_Two_Dot_Zero_single", "3", "5", "mass", "xvel", "yvel", "zvel", "extra", "1", "tag", "1",
           str(npart), str(npart + 1), "0", str(len(counts))]
    hdr += [f"{files[g]} {counts[g]} {offsets[g]}" for g in range(len(counts))]
    (base / "Header").write_text("\n".join(hdr) + "\n")
    with nyx.PMContext(nyx.BoxGeom((8, 8, 8))) as c:
        assert c.read_amrex(str(tmp_path / "chk"), "DM") == npart
        p, v = c.download()
    ref = R.astype(np.float64)
    for d in range(3):
        assert np.array_equal(p[d], ref[:, d])
        assert np.array_equal(v[d], ref[:, 4 + d])
