#!/usr/bin/env python3
"""PM KDK throughput on one MI355X: mixed precision vs FP64, 512^3 clustered.

Workload (the configuration the headline metric is quoted on): 512^3 cells and
512^3 dark-matter particles in the periodic unit box, 20% uniform and 80% in 2000
Plummer spheres (log-uniform scale radii 1-8 cells), velocities N(0, 0.01^2),
4 pi G = 1, a = 1, dt = 1e-3, multigrid to rel 1e-10 (FP64 residual).  A step is
one canonical KDK step: kick-drift, tile sort, deposit, RHS, solve (previous phi
as guess), kick, periodic shift.  Inputs are resident in HBM before the timed
region.

Prints one JSON line (driver contract) with:
  value          mixed-precision particle-steps/s
  fp64           the FP64 path's particle-steps/s and V-cycles
  accel_rel_l2   |g_mixed - g_fp64|_2 / |g_fp64|_2 over all particles, same state
  roofline       the dominant kernel (largest total time) against 8 TB/s
  stages         per-kernel-class mean time and achieved GB/s
  cpu_baseline   the CPU restatement (oracle/) timed on this host on a sample
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

import numpy as np  # noqa: E402

METRIC = "PM KDK particle-steps/s at 512^3, mixed vs FP64 GPU at ≤1e-6 accel rel-L2"
HBM_PEAK_GBS = 8000.0


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


TILE = (16, 8, 8)  # the library's default deposit/kick tile


def tile_of(tuned):
    """The deposit/kick tile of a run: the autotuned one if tuned, else the default."""
    t = (tuned or {}).get("tile")
    return tuple(t["shape"]) if t else TILE


def tile_halo_cells(n, tile):
    """Low-halo cells over all tiles of a box: sum of (bx+1)(by+1)(bz+1) - bx*by*bz
    over the tiles (the last tile along a dimension may be partial)."""
    sizes = []
    for d in range(3):
        full, rem = divmod(n[d], tile[d])
        sizes.append([tile[d]] * full + ([rem] if rem else []))
    return sum((bx + 1) * (by + 1) * (bz + 1) - bx * by * bz
               for bx in sizes[0] for by in sizes[1] for bz in sizes[2])


def coarse_top_cells(ncell):
    """Cells of the first level of at most 4096 cells (halving every dimension
    per level): where the one-workgroup coarse V-cycle takes over."""
    while ncell > 4096:
        ncell //= 8
    return ncell


def kernel_bytes(name, ncell, npart, halo_cells, mixed_sz, moved=0):
    """Algorithmic (compulsory) bytes of one launch of a kernel class: one read
    or write per array per pass.  Multigrid stages carry their depth (".dD": a
    level of ncell / 8^D cells); `moved` is the particles an incremental re-sort
    moves per launch."""
    sz = mixed_sz
    N, P = ncell, npart
    base, _, dep = name.partition(".d")
    if dep:
        Nl = N / 8 ** int(dep)
        per_cell = {
            "smooth": sz + sz / 2 + sz / 2,   # read e (all), r (colour), write e (colour)
            "sweep": 3 * sz,                  # fused red+black: read e, r; write e
            "sweep0": 2 * sz,                 # first sweep from e = 0: read r; write e
            "sweep_p": 3 * sz + sz / 8,       # + coarse correction read (prolongation on load)
            "sweep2": 3 * sz,                 # two sweeps in one pass: read e, r; write e
            "sweep2_0": 2 * sz,               # two sweeps from e = 0: read r; write e
            "sweep2_p": 3 * sz + sz / 8,      # two post-smoothing sweeps, prolongation on load
            "sweep2_u": 2 * sz + sz / 8 + 16,  # the same, phi += e (read/write FP64 phi) instead of e out
            "resid_restrict": 2 * sz + sz / 8,  # read e, r; write coarse r
            "prolong": 2 * sz + sz / 8,       # read/write e, read coarse e
            "fill": sz,
        }
        if base == "sweep_u" and dep == "0":
            return (2 * sz + 16) * N          # read e, r; phi += e (read/write FP64 phi)
        return per_cell[base] * Nl if base in per_cell else None
    table = {
        "residual": (16 + sz) * N,                         # read phi, f (FP64); write r
        "res_sweep0": (16 + 2 * sz) * N,                   # residual + first sweep: write r and e
        "update": (16 + sz) * N,                           # read/write phi, read e
        "guess": 24 * N,                                   # read phi, prev; write both
        "deposit": 24 * P + 8 * N + 8 * halo_cells,        # read positions; write own + halos
        "fold": 8 * N + 8 * halo_cells + 8 * N,            # read own + halos; write f (rho kept in own)
        "kick_drift": 96 * P + 8 * N,                      # read/write x,v; read phi
        "kick_pair": 96 * P + 8 * N,                       # kick + next kick-drift: same traffic
        "kick": 72 * P + 8 * N,                            # read x,v; write v; read phi
        "scatter": 4 * P + 2 * 56 * P,                     # key; move x,v,id
        "keys": 24 * P + 4 * P,                            # read positions; write key
        "migrate": 2 * 52 * moved + 8 * moved,             # move x,v,id of the listed particles; the list
        "coarse": 2 * sz * coarse_top_cells(N),            # the coarse workgroup: its top level's r in, e out
    }
    return table.get(name)


def measured_traffic(stage, n):
    """HBM bytes per launch of `stage` at 512^3 from the committed rocprofv3 PMC
    summary (FETCH_SIZE + WRITE_SIZE, calibrated against a copy of known size;
    profiles/pm_step_512_clustered.md gives the commands), or None."""
    path = os.path.join(ROOT, "profiles", "pmc_traffic_512_clustered.json")
    if n != 512 or not os.path.exists(path):
        return None
    with open(path) as fh:
        e = json.load(fh).get("stages", {}).get(stage)
    return e["traffic_bytes"] if e else None


def dominant_stage(st, steps):
    """The stage with the largest share of a steady-state step.  A call of n steps
    runs one opening kick-drift, n-1 paired kicks and one closing kick; in steady
    state every step has exactly one paired kick."""
    def per_step(kv):
        nm, e = kv
        if nm in ("kick_drift", "kick"):
            return 0.0
        if nm == "kick_pair":
            return e["mean_ms"]
        return e["total_ms"] / steps
    return max(st.items(), key=per_step)[0]


def run_timed(ctx, precision, steps, warmup, dt, tol, use_graph,
              timed_stages=("kick_pair", "residual", "sweep.d0", "sweep2.d0", "sweep2_p.d0")):
    """Warm-up steps, then the `steps` timed steps, in which only the launches of
    `timed_stages` (the candidates for the dominant kernel) are bracketed by HIP
    events (two stream markers per launch: events around every stage would add
    their own gaps to the timed region).  The dominant one is the stage with the
    largest share of a steady-state step in these steps."""
    ctx.set_precision(precision)
    if warmup:
        ctx.pm_steps(warmup, dt, rel_tol=tol, use_graph=use_graph)
    ctx.sync()
    ctx.timing_reset()
    ctx.timing_select(list(timed_stages))
    ctx.timing_enable(True)
    s0 = ctx.sort_stats()
    ctx.sync()
    t0 = time.perf_counter()
    it = ctx.pm_steps(steps, dt, rel_tol=tol, use_graph=use_graph)
    ctx.sync()
    t1 = time.perf_counter()
    ctx.timing_enable(False)
    cand = {}
    for nm in timed_stages:
        cnt, mean, tot = ctx.timing(nm)
        if cnt:
            cand[nm] = {"launches": cnt, "mean_ms": round(mean, 4), "total_ms": round(tot, 2)}
    ctx.timing_select(None)
    s1 = ctx.sort_stats()
    # particle re-sorts inside the timed region: full counting sorts, incremental
    # ones, and the particles moved per step by the latter
    resort = {"full": s1["full"] - s0["full"], "incremental": s1["incremental"] - s0["incremental"],
              "moved_per_step": round((s1["moved"] - s0["moved"]) / max(1, steps), 1)}
    dname = dominant_stage(cand, steps)
    dom = {"stage": dname, "launches": cand[dname]["launches"], "mean_ms": cand[dname]["mean_ms"],
           "timed_in": "timed steps", "candidates": cand}
    return t1 - t0, int(it[0]), resort, dom


def run_instrumented(ctx, precision, steps, dt, tol, use_graph, names, ncell, npart, halo_cells, sz, dom):
    """`steps` further steps with HIP events around every stage: the stage table.
    The dominant stage is read off it; if it is not the stage timed inside the
    timed steps, its mean comes from these steps (reported in the result)."""
    ctx.set_precision(precision)
    ctx.timing_reset()
    ctx.timing_select(None)
    ctx.timing_enable(True)
    s0 = ctx.sort_stats()
    ctx.sync()
    t0 = time.perf_counter()
    ctx.pm_steps(steps, dt, rel_tol=tol, use_graph=use_graph)
    ctx.sync()
    t_inst = time.perf_counter() - t0
    ctx.timing_enable(False)
    s1 = ctx.sort_stats()
    inc = max(1, s1["incremental"] - s0["incremental"])
    st = stage_table(ctx, names, ncell, npart, halo_cells, sz, t_inst * 1e3, (s1["moved"] - s0["moved"]) / inc)
    dname = dominant_stage(st, steps)
    if dom is None or dname not in dom.get("candidates", {}):
        # a stage outside the timed candidates leads the stage table: report it
        # with its mean from these steps
        dom = {"stage": dname, "launches": st[dname]["launches"], "mean_ms": st[dname]["mean_ms"],
               "timed_in": "instrumented steps"}
    return st, dom, t_inst


def stage_table(ctx, names, ncell, npart, halo_cells, sz, elapsed_ms, moved=0):
    out = {}
    for nm in names:
        cnt, mean, tot = ctx.timing(nm)
        if cnt == 0:
            continue
        b = kernel_bytes(nm, ncell, npart, halo_cells, sz, moved)
        e = {"launches": cnt, "mean_ms": round(mean, 4), "total_ms": round(tot, 2),
             "share": round(tot / elapsed_ms, 4)}
        if b is not None and mean > 0:
            e["gbs"] = round(b / (mean * 1e-3) / 1e9, 1)
            e["bytes_per_launch"] = int(b)
        out[nm] = e
    return out


def host_cpu():
    """CPU model and core counts of this host (/proc/cpuinfo, as lscpu reads it)."""
    model, cores, logical = None, set(), 0
    try:
        phys = core = None
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                k, _, v = line.partition(":")
                k, v = k.strip(), v.strip()
                if k == "processor":
                    logical += 1
                elif k == "model name" and model is None:
                    model = v
                elif k == "physical id":
                    phys = v
                elif k == "core id":
                    core = v
                    cores.add((phys, core))
    except OSError:
        pass
    try:
        usable = len(os.sched_getaffinity(0))
    except Exception:
        usable = os.cpu_count() or 1
    return {"model": model, "physical_cores": len(cores) or None, "logical_cpus": logical or None,
            "usable_cpus": usable}


def stream_copy(gib=4, reps=10):
    """Device-to-device copy bandwidth of this GPU (torch copy_ of `gib` GiB, read
    plus write bytes per copy), for the roofline beside the 8 TB/s peak."""
    import torch
    n = gib * 2 ** 27
    src = torch.ones(n, dtype=torch.float64, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(2):
        dst.copy_(src)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    e0.record()
    for _ in range(reps):
        dst.copy_(src)
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / reps
    del src, dst
    torch.cuda.empty_cache()
    return {"gbs": round(2 * n * 8 / (ms * 1e-3) / 1e9, 1), "bytes_per_copy": 2 * n * 8, "ms": round(ms, 4),
            "what": f"torch copy_ of {gib} GiB FP64 on the current stream, mean of {reps}"}


def cpu_baseline(n, pos, vel, phi, m, threads, dt, tol):
    """The oracle (the C/OpenMP FP64 restatement of the Nyx CPU path) on this
    host: one canonical KDK step of the workload from the same state as the GPU's
    timed steps (particles and phi), then one step of Nyx's two-solve form (old
    phi re-solved from a zero guess, Nyx_advance_particles.cpp:148-170)."""
    import oracle
    oracle.set_num_threads(threads)
    nn = (n, n, n)
    pos = [np.ascontiguousarray(a) for a in pos]
    vel = [np.ascontiguousarray(a) for a in vel]
    phi = np.ascontiguousarray(phi)
    t0 = time.perf_counter()
    _, st = oracle.pm_step(nn, (0, 0, 0), (1, 1, 1), pos, vel, phi, m_const=m, dt=dt, rel_tol=tol)
    t1 = time.perf_counter()
    _, st2 = oracle.pm_step(nn, (0, 0, 0), (1, 1, 1), pos, vel, phi, m_const=m, dt=dt, rel_tol=tol,
                            nyx_faithful=True)
    t2 = time.perf_counter()
    cpu = host_cpu()
    return {"value": round(n ** 3 / (t1 - t0), 1), "unit": "particle-steps/s", "cores": threads,
            "kind": "port", "cpu_model": cpu["model"], "host_physical_cores": cpu["physical_cores"],
            "host_usable_cpus": cpu["usable_cpus"],
            "nyx_faithful_value": round(n ** 3 / (t2 - t1), 1),
            "sample": f"one canonical KDK step of the same {n}^3 clustered workload from the state the GPU's "
                      f"timed steps start from (FP64 deposit, MG to rel {tol:g} with {int(st[0])} V-cycles, "
                      f"gradient, kick-drift, kick) in {t1 - t0:.2f} s (deposit {st[4]:.2f} s, solve "
                      f"{st[5]:.2f} s, gravity + kicks {st[6]:.2f} s); then one step of Nyx's two-solve form "
                      f"(old phi from a zero guess, {int(st2[7])} + {int(st2[0])} V-cycles) in {t2 - t1:.2f} s; "
                      f"{threads} OpenMP threads"}


class Snapshot:
    """Particles (upload order) and phi of a context, kept in device memory, so
    that each precision is timed from the same state."""

    def __init__(self, ctx, n):
        import torch
        N = n ** 3
        self.N = N
        self.parts = [torch.empty(N, dtype=torch.float64, device="cuda") for _ in range(6)]
        self.phi = torch.empty(N, dtype=torch.float64, device="cuda")
        ctx.download_device([t.data_ptr() for t in self.parts])
        ctx.get_phi_device(self.phi.data_ptr())
        ctx.sync()

    def restore(self, ctx, m):
        ctx.upload_device(0, self.N, [t.data_ptr() for t in self.parts], m_const=m)
        ctx.set_phi_device(self.phi.data_ptr())
        ctx.sync()

    def host(self):
        a = [t.cpu().numpy() for t in self.parts]
        return a[:3], a[3:], self.phi.cpu().numpy()


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n", type=int, default=512)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--graph", action="store_true")
    ap.add_argument("--tol", type=float, default=1e-10)
    ap.add_argument("--json-out", default=None)
    ap.add_argument("--modes", default="mixed,fp64", help="precisions to time (comma list)")
    ap.add_argument("--no-accuracy", action="store_true", help="skip the mixed-vs-FP64 acceleration check")
    ap.add_argument("--no-tune", action="store_true", help="skip the smoother-shape autotuner")
    ap.add_argument("--no-stages", action="store_true", help="skip the per-stage tables")
    ap.add_argument("--kick-chunk", type=int, default=0, help="particles per kick work item (0: default)")
    ap.add_argument("--phi-guess", default="default", choices=["default", "previous", "extrapolate"],
                    help="initial guess of each step's solve (default = previous phi, as Nyx)")
    ap.add_argument("--trace-summary", default=None, metavar="KERNEL_TRACE_CSV",
                    help="instead of running: summarise a rocprofv3 kernel trace of a bench run over that "
                         "run's timed mixed steps (needs --from-json with the run's JSON line)")
    ap.add_argument("--from-json", default=None)
    ap.add_argument("--pmc-summary", nargs=4, default=None,
                    metavar=("FETCH_CSV", "WRITE_CSV", "COPY_FETCH_CSV", "COPY_WRITE_CSV"),
                    help="instead of running: HBM traffic per launch of each kernel (largest grid) from "
                         "rocprofv3 --pmc FETCH_SIZE and WRITE_SIZE counter collections of a 512^3 bench run, "
                         "scaled by the same counters of a 1 GiB copy, against the kernel's algorithmic bytes")
    ap.add_argument("--workload", default="clustered512",
                    choices=["clustered512", "batch48", "grf256", "grf1024"],
                    help="clustered512: the headline line; batch48: 48 boxes (16 each of 64^3, 96^3, "
                         "128^3) in one launch per phase vs one after another; grf256: FP64 drop-in, "
                         "256^3 GRF, 10 steps; grf1024: 1024^3 GRF, 50 mixed steps with per-step "
                         "accel rel-L2 against FP64 and the step-50 power spectrum check (--n to shrink)")
    a = ap.parse_args()
    if a.trace_summary:
        return trace_summary(a.trace_summary, a.from_json)
    if a.pmc_summary:
        jp = (a.from_json or "").split(",") if a.from_json else [None, None]
        stages = pmc_summary(*a.pmc_summary, fetch_json=jp[0], write_json=jp[-1])
        if a.json_out:
            with open(a.json_out, "w") as fh:
                json.dump({"source": "bench.py --pmc-summary of rocprofv3 FETCH_SIZE / WRITE_SIZE runs "
                                     "(profiles/pm_step_512_clustered.md)", "stages": stages}, fh, indent=1)
        return None
    if a.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU only: the library runs each box on one GPU and has no "
              "path that spreads a step over several; run with --gpus 1", file=sys.stderr)
        sys.exit(2)
    if a.workload == "batch48":
        return main_batch(a)
    if a.workload == "grf1024":
        return main_grf1024(a)
    if a.workload == "grf256":
        return main_grf256(a)

    import torch
    import nyx_amd
    from nyx_amd import ics
    n = a.n
    dt = 1e-3
    npart = ncell = n ** 3
    nsph = max(1, int(2000 * (n / 512) ** 3))
    log(f"generating {n}^3 clustered particles on the device")
    t = time.perf_counter()
    pos, vel, m = ics.plummer_clustered_device(n, ics.SEED_BASE + 2, n_spheres=nsph)
    log(f"  {time.perf_counter() - t:.1f} s")
    ctx = nyx_amd.PMContext(nyx_amd.BoxGeom((n, n, n)), precision="fp64", kick_chunk=a.kick_chunk,
                            phi_guess=a.phi_guess)
    ctx.upload_device(0, npart, [t_.data_ptr() for t_ in pos + vel], m_const=m)
    ctx.sync()
    del pos, vel
    torch.cuda.empty_cache()
    log("device bytes", ctx.device_bytes())
    tuned = None
    if not a.no_tune:
        t = time.perf_counter()
        tuned = ctx.autotune()
        log(f"autotuned smoother shapes in {time.perf_counter() - t:.1f} s: {tuned}")

    # accuracy: same particles, same density; FP64 solve vs mixed solve
    log("FP64 initial solve")
    ctx.set_precision("fp64")
    ctx.pm_init(1.0, 1.0, a.tol)
    rel_l2 = None
    g64 = gmx = None
    if not a.no_accuracy:
        g64 = torch.empty(3 * npart, dtype=torch.float64, device="cuda")
        gmx = torch.empty(3 * npart, dtype=torch.float64, device="cuda")
        ctx.gather_accel_device(g64.data_ptr())
        ctx.set_precision("mixed")
        it_m, _ = ctx.solve_for_phi(1.0, 1.0, a.tol, 0.0, use_guess=False)
        ctx.gather_accel_device(gmx.data_ptr())
        rel_l2 = (torch.linalg.vector_norm(gmx - g64) / torch.linalg.vector_norm(g64)).item()
        log(f"accel rel-L2 mixed vs fp64 = {rel_l2:.3e}  (mixed cold-start cycles {int(it_m[0])})")
        ctx.set_precision("fp64")
        ctx.solve_for_phi(1.0, 1.0, a.tol, 0.0, use_guess=False)  # back to the FP64 phi
    modes = a.modes.split(",")
    # warm-up steps, then the state every timed run starts from
    warm_prec = "mixed" if "mixed" in modes else "fp64"
    ctx.set_precision(warm_prec)
    if a.warmup:
        ctx.pm_steps(a.warmup, dt, rel_tol=a.tol, use_graph=a.graph)
    snap = Snapshot(ctx, n)
    log(f"{a.warmup} {warm_prec} warm-up steps; state saved")

    names = ["residual", "norm", "update", "deposit", "fold", "rhs", "kick_drift", "kick_pair", "kick", "scatter",
             "scan", "coarse", "migrate", "res_sweep0", "guess",
             "items", "keys", "bottom"] + \
            [f"{k}.d{d}" for k in ("sweep0", "sweep", "sweep_p", "sweep_u", "smooth", "resid_restrict", "prolong",
                                    "fill", "sweep2", "sweep2_0", "sweep2_p", "sweep2_u")
             for d in range(0, 12)]
    halo_cells = tile_halo_cells((n, n, n), tile_of(tuned))

    def settle(prec):
        # same state, then one untimed step of this precision (its first sort is
        # a full one after the restore)
        snap.restore(ctx, m)
        ctx.set_precision(prec)
        ctx.pm_steps(1, dt, rel_tol=a.tol, use_graph=a.graph)
        ctx.sync()

    res = {}
    windows = {}
    for prec in ("mixed", "fp64"):
        if prec not in modes:
            continue
        settle(prec)
        log(f"{prec}: {a.steps} timed steps from the saved state")
        w0 = time.monotonic_ns()
        t_, it_, rs_, dom_ = run_timed(ctx, prec, a.steps, 0, dt, a.tol, a.graph)
        windows[prec] = [w0, time.monotonic_ns()]
        res[prec] = {"t": t_, "it": it_, "rs": rs_, "dom": dom_}
        log(f"  {t_ / a.steps * 1e3:.2f} ms/step, {it_} V-cycles")
        if prec == "mixed" and not a.no_accuracy:
            # the phi the timed mixed steps left against an FP64 solve of the same
            # density from a zero guess
            ctx.gather_accel_device(gmx.data_ptr())
            ctx.set_precision("fp64")
            ctx.assign_density(want=False)
            ctx.solve_for_phi(1.0, 1.0, a.tol, 0.0, use_guess=False)
            ctx.gather_accel_device(g64.data_ptr())
            res[prec]["rel_end"] = (torch.linalg.vector_norm(gmx - g64) / torch.linalg.vector_norm(g64)).item()
            log(f"  accel rel-L2 mixed vs fp64 after the timed steps = {res[prec]['rel_end']:.3e}")
    del g64, gmx
    torch.cuda.empty_cache()
    # the stage tables: the same steps again from the same state, every stage
    # bracketed by events
    st = {}
    if not a.no_stages:
        for prec in ("mixed", "fp64"):
            if prec not in modes:
                continue
            settle(prec)
            st[prec], res[prec]["dom"], ti = run_instrumented(ctx, prec, a.steps, dt, a.tol, a.graph, names, ncell,
                                                              npart, halo_cells, 4 if prec == "mixed" else 8,
                                                              res[prec]["dom"])
            log(f"  {prec} stage table: {a.steps} instrumented steps, {ti / a.steps * 1e3:.2f} ms/step")
    ctx.close()
    first = "mixed" if "mixed" in res else "fp64"
    other = "fp64" if "fp64" in res else first
    R, R64 = res[first], res[other]
    v_mix = npart * a.steps / R["t"]
    v_64 = npart * a.steps / R64["t"]
    dom_mix = R["dom"]
    dname, dmean = dom_mix["stage"], dom_mix["mean_ms"]
    b = kernel_bytes(dname, ncell, npart, halo_cells, 4 if first == "mixed" else 8)
    achieved = b / (dmean * 1e-3) / 1e9 if b and dmean > 0 else None
    roof = {"bound": "hbm", "kernel": dname, "achieved": round(achieved, 1) if achieved else None,
            "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4) if achieved else None,
            "traffic": measured_traffic(dname, n), "bytes_per_launch": int(b) if b else None,
            "mean_launch_ms": dmean, "launches_timed": dom_mix["launches"], "timed_in": dom_mix["timed_in"]}

    def smoother_roof(stt, sz):
        """The finest level's smoothing sweep: a full red+black sweep per pass
        (sweep.d0), or two per pass (sweep2.d0, FP32 levels): then `achieved` is
        the pass's compulsory bytes (e, r in, e out once for both sweeps) per
        time and `per_sweep_equivalent` the bytes of the two single-sweep passes
        it replaces per time."""
        stt = stt or {}

        def hbm(key, mean_ms):
            # the bytes the pass moves to and from HBM (committed PMC summary)
            tb = measured_traffic(key, n)
            if not tb:
                return {}
            g = tb / (mean_ms * 1e-3) / 1e9
            return {"hbm_traffic_per_launch": tb, "hbm_achieved": round(g, 1), "hbm_frac": round(g / HBM_PEAK_GBS, 4)}

        e = stt.get("sweep.d0")
        if e:
            bb = kernel_bytes("sweep.d0", ncell, npart, halo_cells, sz)
            ach = bb / (e["mean_ms"] * 1e-3) / 1e9
            return {"kernel": "sweep.d0 (red+black sweep, finest level)", "achieved": round(ach, 1),
                    "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(ach / HBM_PEAK_GBS, 4),
                    "bytes_per_launch": int(bb), "mean_launch_ms": e["mean_ms"],
                    **hbm(("fp64:" if sz == 8 else "") + "sweep.d0", e["mean_ms"])}
        e = stt.get("sweep2.d0")
        if not e:
            return None
        bb = kernel_bytes("sweep2.d0", ncell, npart, halo_cells, sz)
        ach = bb / (e["mean_ms"] * 1e-3) / 1e9
        return {"kernel": "sweep2.d0 (two red+black sweeps in one pass, finest level)", "achieved": round(ach, 1),
                "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(ach / HBM_PEAK_GBS, 4),
                "bytes_per_launch": int(bb), "mean_launch_ms": e["mean_ms"],
                "per_sweep_equivalent": {"achieved": round(2 * ach, 1), "frac": round(2 * ach / HBM_PEAK_GBS, 4),
                                         "bytes_of_two_passes": int(2 * bb)},
                **hbm("sweep2.d0", e["mean_ms"])}

    copy = stream_copy()
    log(f"device copy {copy['gbs']} GB/s")
    roof["frac_of_copy"] = round(achieved / copy["gbs"], 4) if achieved else None

    cpu = None
    if not a.no_cpu:
        info = host_cpu()
        # the OpenMP threads this job may use (the box sets OMP_NUM_THREADS to its
        # share of the host), else every usable CPU
        threads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or info["usable_cpus"]
        log(f"CPU baseline at {n}^3 on {threads} threads ({info})")
        hp, hv, hphi = snap.host()
        del snap
        torch.cuda.empty_cache()
        cpu = cpu_baseline(n, hp, hv, hphi.reshape(n, n, n), m, threads, dt, a.tol)
        log("  ", cpu["value"], "particle-steps/s")

    out = {
        "metric": METRIC, "value": round(v_mix, 1), "unit": "particle-steps/s", "n_gpus": 1,
        "steps": a.steps, "warmup": a.warmup, "ms_per_step": round(R["t"] / a.steps * 1e3, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32+f64",
        "data": "synthetic",
        "config": {"workload": f"{n}^3 cells, {n}^3 particles, 20% uniform + 80% in {nsph} Plummer spheres, "
                               f"mixed precision (FP32 V-cycles in FP64 refinement, FP64 particles), dt 1e-3, "
                               f"rel tol {a.tol:g}",
                   "cells": ncell, "particles": npart, "graph": bool(a.graph), "autotuned": tuned},
        "fp64": {"value": round(v_64, 1), "ms_per_step": round(R64["t"] / a.steps * 1e3, 3), "vcycles": R64["it"],
                 "dtype": "f64", "resort": R64["rs"]},
        "same_state": f"each precision restarts from the state after the {a.warmup} warm-up steps (particles "
                      f"and phi restored), runs one untimed step, then the {a.steps} timed steps",
        "resort": R["rs"],
        "mixed_over_fp64": round(v_mix / v_64, 3),
        "vcycles_mixed": R["it"],
        "accel_rel_l2": rel_l2,
        "accel_rel_l2_after_steps": R.get("rel_end"),
        "phi_guess": a.phi_guess,
        "roofline": roof,
        "stream_copy": copy,
        "smoother_roofline": {"mixed": smoother_roof(st.get("mixed"), 4), "fp64": smoother_roof(st.get("fp64"), 8)},
        "timed_windows_monotonic_ns": windows,
        "stages_note": "stage tables: the same steps again from the same state with HIP events around every "
                       "stage; the timed steps bracket only the paired kick and the level-0 residual (the "
                       "dominant-kernel candidates)",
        "stages_mixed": st.get("mixed", {}),
        "stages_fp64": st.get("fp64", {}),
        "cpu_baseline": cpu,
    }
    emit(out, a)


def _stage_of_kernel(name):
    """Stage name (as the library's timing classes) of a kernel in a rocprofv3
    trace, from its demangled name; None for kernels outside the step."""
    import re
    s = name
    if "k_rb_sweep2" in s:
        mt = re.search(r"k_rb_sweep2<(\w+), *(\d+), *(\d+), *(\d+)", s)
        mode = int(mt.group(4)) if mt else -1
        return {1: "sweep2_0", 0: "sweep2", 2: "sweep2_p", 6: "sweep2_u"}.get(mode, "sweep2?"), \
            (mt.group(1) if mt else "?")
    if "k_rb_sweep" in s:
        mt = re.search(r"k_rb_sweep<(\w+), *(\d+), *(\d+), *(\d+)>", s)
        mode = int(mt.group(4)) if mt else -1
        return {1: "sweep0", 2: "sweep_p", 4: "sweep_u", 6: "sweep_p", 0: "sweep"}.get(mode, "sweep?"), \
            (mt.group(1) if mt else "?")
    for k, v in (("k_residual0", "residual"), ("k_kick_tiled", None), ("k_deposit_mom", "deposit_mom"),
                 ("k_deposit", "deposit"), ("k_fold", "fold"), ("k_resid_restrict_z", "resid_restrict"),
                 ("k_resid_restrict", "resid_restrict"), ("k_coarse_vcycle", "coarse"), ("k_migrate", "migrate"),
                 ("k_scatter", "scatter"), ("k_keys", "keys"), ("k_make_items", "items"),
                 ("k_rhs_sum", "rhs"), ("k_check", "check"), ("k_update0", "update"),
                 ("k_kick_deferred", "kick_deferred"), ("k_fill", "fill"), ("k_scan", "scan"),
                 ("k_tile", "scan"), ("k_prolong", "prolong"), ("k_gsrb_half", "smooth"), ("k_bottom", "bottom"),
                 ("k_graph_note", "check"), ("k_phi_guess", "guess")):
        if k in s:
            if k == "k_kick_tiled":
                mt = re.search(r"k_kick_tiled<(\d+)", s)
                return {"0": "kick", "1": "kick_drift", "2": "kick_pair"}.get(mt.group(1) if mt else "", "kick?"), ""
            prec = "float" if "<float" in s else ("double" if "<double" in s else "")
            return v, prec
    return None, ""


def _ceil(a, b):
    return -(-a // b)


def _split_by_level(values, levels, counts=None):
    """Dispatches of one kernel and grid size that several levels launch alike,
    split over those levels, the largest values (durations or counter readings,
    ~8x apart from one level to the next) on the finest: by the launch count of
    each level where known (counts), else at the gaps of more than 3x."""
    import math
    levels = sorted(levels)
    v = sorted(values, reverse=True)
    if len(levels) == 1:
        return {levels[0]: v}
    out, start = {}, 0
    if counts and all(l in counts for l in levels):
        for i, lev in enumerate(levels):
            end = len(v) if i + 1 == len(levels) else min(len(v), start + counts[lev])
            out[lev] = v[start:end]
            start = end
        return out
    gaps = [i for i in range(1, len(v)) if v[i - 1] > 3.0 * max(v[i], 1e-30)]
    gaps = sorted(sorted(gaps, key=lambda i: -math.log(v[i - 1] / max(v[i], 1e-30)))[:len(levels) - 1])
    for lev, end in zip(levels, gaps + [len(v)]):
        out[lev] = v[start:end]
        start = end
    return out


def kernel_levels(name, grid, n, tuned, sw2_zc=128):
    """Every level kernel_level could assign (the finest first)."""
    levs = []
    for d in range(12):
        lev = kernel_level(name, grid, n, tuned, sw2_zc, only=d)
        if lev is not None:
            levs.append(lev)
    return levs or [kernel_level(name, grid, n, tuned, sw2_zc)]


def kernel_level(name, grid, n, tuned, sw2_zc=128, only=None):
    """Multigrid level (0 = finest) of a launch of a level kernel of a box of n^3
    cells, from its grid size (threads) and the launch shapes the library uses
    (the autotuned fused-sweep shape and z-chunk per precision and level, as the
    bench line's config holds them); 0 for the kernels that only run on the
    finest level, None if no level matches."""
    import re
    mt = re.search(r"k_rb_sweep2<(\w+), *(\d+), *(\d+), *(\d+)>", name)
    if mt:
        tx, ty = int(mt.group(2)), int(mt.group(3))
        nt = (tx + 8) // 2 * (ty + 8)
        for d in (range(12) if only is None else [only]):
            nd = n >> d
            if nd < 8:
                break
            cfg = (tuned or {}).get(f"sweep2.d{d}")
            if cfg and tuple(cfg["shape"]) != (tx, ty):
                continue
            zc = min(cfg["zchunk"] if cfg else sw2_zc, nd)
            if nd % tx == 0 and nd % ty == 0 and (nd // tx) * (nd // ty) * _ceil(nd, zc) * nt == grid:
                return d
        return None
    mt = re.search(r"k_rb_sweep<(\w+), *(\d+), *(\d+), *(\d+)>", name)
    if mt:
        pr = "fp32" if mt.group(1) == "float" else "fp64"
        tx, ty = int(mt.group(2)), int(mt.group(3))
        for d in (range(12) if only is None else [only]):
            nd = n >> d
            if nd < 2:
                break
            cfg = (tuned or {}).get(f"{pr}.d{d}")
            if not cfg or tuple(cfg["shape"]) != (tx, ty) or nd % tx or nd % ty:
                continue
            nzc = _ceil(nd, cfg["zchunk"])
            zc = _ceil(nd, nzc)
            if (nd // tx) * (nd // ty) * _ceil(nd, zc) * (tx * ty // 2) == grid:
                return d
        return None
    mt = re.search(r"k_resid_restrict_z<(\w+), *(\d+), *(\d+)>", name)
    if mt:
        pr = "fp32" if mt.group(1) == "float" else "fp64"
        cx, cy = int(mt.group(2)), int(mt.group(3))
        for d in (range(12) if only is None else [only]):
            nd = n >> d
            if nd < 2 * cx:
                break
            cols = (nd // (2 * cx)) * (nd // (2 * cy))
            cz = nd // 2
            cfg = (tuned or {}).get(f"rrz.{pr}.d{d}")
            if cfg:
                zc = min(cfg["zchunk"], cz)
            else:
                zc = min(16, cz)
                while zc > 2 and cols * _ceil(cz, zc) < 1024:
                    zc //= 2
            if cols * _ceil(cz, zc) * cx * cy == grid:
                return d
        return None
    if "k_resid_restrict<" in name:
        for d in (range(12) if only is None else [only]):
            nd = n >> d
            if nd < 2:
                break
            if _ceil(nd ** 3 // 8, 256) * 256 == grid:
                return d
        return None
    return 0 if only in (None, 0) else None


def trace_summary(csv_path, json_path):
    """Per kernel and level: launches, mean duration and achieved GB/s
    (algorithmic bytes of that level per launch / mean duration) over the timed
    steps of each precision of a bench run, from a rocprofv3 kernel trace (CSV)
    of that run and the run's JSON line (timed windows, autotuned shapes)."""
    import csv
    with open(json_path) as fh:
        run = json.loads(fh.read().strip().splitlines()[-1])
    n = round(run["config"]["cells"] ** (1.0 / 3.0))
    tuned = run["config"].get("autotuned")
    steps = run["steps"]
    halo_cells = tile_halo_cells((n, n, n), tile_of(tuned))
    out_rows = []
    for prec in ("mixed", "fp64"):
        w = run.get("timed_windows_monotonic_ns", {}).get(prec)
        if not w:
            continue
        agg = {}
        with open(csv_path) as fh:
            for r in csv.DictReader(fh):
                t0, t1 = int(r["Start_Timestamp"]), int(r["End_Timestamp"])
                if t0 < w[0] or t1 > w[1]:
                    continue
                key = (r["Kernel_Name"], int(r["Grid_Size_X"]) * int(r["Grid_Size_Y"]) * int(r["Grid_Size_Z"]))
                agg.setdefault(key, []).append((t1 - t0) * 1e-6)
        sz = 4 if prec == "mixed" else 8
        stab = run.get(f"stages_{prec}") or {}
        rows = []
        for (nm, grid), durs in agg.items():
            stage, tprec = _stage_of_kernel(nm)
            groups = {None: durs}
            if stage:
                levs = kernel_levels(nm, grid, n, tuned)
                # the launches per level of the instrumented steps (the same number of steps)
                known = [l for l in levs if f"{stage}.d{l}" in stab]
                counts = {l: stab[f"{stage}.d{l}"]["launches"] for l in known}
                groups = _split_by_level(durs, known or levs, counts if known else None)
            for lev, dl in groups.items():
                if not dl:
                    continue
                b = None
                if stage:
                    key = stage
                    if lev is not None and (stage.startswith("sweep") or stage == "resid_restrict"):
                        key = f"{stage}.d{lev}"
                    b = kernel_bytes(key, n ** 3, n ** 3, halo_cells, 4 if tprec == "float" else sz)
                cnt, tot = len(dl), sum(dl)
                rows.append((tot, nm, grid, stage, lev, cnt, tot / cnt, b))
        rows.sort(reverse=True)
        tot_all = sum(r[0] for r in rows)
        out_rows.append(f"### {prec} timed steps ({steps} steps, window {w[1] - w[0]} ns, kernels "
                        f"{tot_all / steps:.3f} ms per step)\n")
        out_rows.append("| kernel | grid | stage | level | launches | mean us | ms per step | GB/s | of 8 TB/s |")
        out_rows.append("|---|---|---|---|---|---|---|---|---|")
        for tot, nm, grid, stage, lev, cnt, mean, b in rows:
            short = nm.split("(")[0].replace("void ", "").replace("nyxpm::", "")[:60]
            gbs = f"{b / (mean * 1e-3) / 1e9:.0f}" if b else ""
            frac = f"{b / (mean * 1e-3) / 1e9 / HBM_PEAK_GBS:.2f}" if b else ""
            out_rows.append(f"| `{short}` | {grid} | {stage or ''} | {'' if lev is None else lev} | {cnt} | "
                            f"{mean * 1e3:.1f} | {tot / steps:.3f} | {gbs} | {frac} |")
        out_rows.append("")
    print("\n".join(out_rows))


def pmc_summary(fetch_csv, write_csv, copy_fetch_csv, copy_write_csv, fetch_json=None, write_json=None, n=512):
    """Per kernel on the finest level the mean FETCH_SIZE and WRITE_SIZE (KiB per
    dispatch) of two rocprofv3 counter collections of bench runs, each scaled by
    what the counter reads for a copy of 1 GiB (__amd_rocclr_copyBuffer), and the
    ratio to the algorithmic bytes.  The level of a dispatch comes from its grid
    size and the autotuned launch shapes of its run's bench line (fetch_json,
    write_json; without them, the largest grid of each kernel), so that the
    autotuner's trial launches are left out."""
    import csv

    def tuned_of(path):
        if not path:
            return None, n, []
        with open(path) as fh:
            run = json.loads(fh.read().strip().splitlines()[-1])
        wins = [tuple(w) for w in run.get("timed_windows_monotonic_ns", {}).values() if w]
        return run["config"].get("autotuned"), round(run["config"]["cells"] ** (1.0 / 3.0)), wins

    def load(path, jpath):
        tuned, nn, wins = tuned_of(jpath)
        rows = {}
        with open(path) as fh:
            for r in csv.DictReader(fh):
                # only the dispatches of the timed steps (not the autotuner's trials)
                if wins and not any(a0 <= int(r["Start_Timestamp"]) <= a1 for a0, a1 in wins):
                    continue
                nm = r["Kernel_Name"].split("(")[0].replace("void ", "").replace("nyxpm::", "")
                rows.setdefault((nm, int(r["Grid_Size"])), []).append(float(r["Counter_Value"]))
        out = {}
        for (nm, grid), v in rows.items():
            if jpath:
                levs = kernel_levels(nm, grid, nn, tuned)
                if 0 in levs:
                    out.setdefault(nm, []).extend(_split_by_level(v, levs)[0])
            elif nm not in out or grid > out[nm][0]:
                out[nm] = (grid, v)
        if not jpath:
            out = {k: v[1] for k, v in out.items()}
        return out, nn

    def copy_kib(path):
        with open(path) as fh:
            for r in csv.DictReader(fh):
                if "copyBuffer" in r["Kernel_Name"]:
                    return float(r["Counter_Value"])
        raise SystemExit(f"{path}: no __amd_rocclr_copyBuffer dispatch")

    (F, nn), (W, _) = load(fetch_csv, fetch_json), load(write_csv, write_json)
    tuned_f = tuned_of(fetch_json)[0]
    gib_kib = 2.0 ** 30 / 1024.0
    sf, sw = gib_kib / copy_kib(copy_fetch_csv), gib_kib / copy_kib(copy_write_csv)
    N = nn ** 3
    halo = tile_halo_cells((nn, nn, nn), tile_of(tuned_f))
    stages = {}
    print(f"calibration: a 1 GiB copy reads FETCH_SIZE x{sf:.3f}, WRITE_SIZE x{sw:.3f} of its bytes\n")
    print("| kernel (finest level) | dispatches | FETCH KiB | WRITE KiB | traffic GB (scaled) | algorithmic GB | ratio |")
    print("|---|---|---|---|---|---|---|")
    for nm in sorted(F):
        fv, wv = F[nm], W.get(nm, [])
        if not wv:
            continue
        fk, wk = sum(fv) / len(fv), sum(wv) / len(wv)
        gb = (fk * sf + wk * sw) * 1024.0 / 1e9
        stage, _ = _stage_of_kernel(nm)
        sz = 4 if "float" in nm else 8
        key = stage + ".d0" if stage and (stage.startswith("sweep") or stage == "resid_restrict") else stage
        alg = kernel_bytes(key, N, N, halo, sz) if key else None
        tail = f"{alg / 1e9:.3f} | {gb / (alg / 1e9):.2f} |" if alg else " | |"
        print(f"| `{nm}` | {len(fv)} | {fk:.4g} | {wk:.4g} | {gb:.3f} | " + tail)
        # the mixed-precision step's kernels (FP32 multigrid, FP32-weight deposit
        # and interpolation) for roofline.traffic
        mixed_kernel = {"fold": nm.startswith("k_fold<false>"),
                        "deposit": nm.startswith("k_deposit<true>")}.get(key, "double" not in nm)
        if key and alg:
            # the mixed step's kernels under the stage name, the FP64 step's
            # under "fp64:" + the stage name
            stages[key if mixed_kernel else "fp64:" + key] = {
                "kernel": nm, "fetch_bytes": int(fk * sf * 1024), "write_bytes": int(wk * sw * 1024),
                "traffic_bytes": int(gb * 1e9), "algorithmic_bytes": int(alg)}
    return stages


def emit(out, a):
    line = json.dumps(out)
    if a.json_out:
        with open(a.json_out, "w") as fh:
            fh.write(line + "\n")
    print(line, flush=True)


def main_batch(a):
    """48 independent periodic boxes (16 each of 64^3, 96^3 and 128^3 cells,
    particles = cells), lattice + GRF displacements (P(k) ~ k^-2, rms 0.5 cell,
    v = 0.1 psi), each with its own seed; 20 hipGraph-captured KDK steps in one
    launch per phase, FP64 then mixed, against the same boxes advanced one after
    another (one context each, same step code)."""
    import nyx_amd
    from nyx_amd import ics
    sizes = [64] * 16 + [96] * 16 + [128] * 16
    steps = a.steps if a.steps != 5 else 20
    log("generating 48 GRF boxes")
    inputs = [ics.grf_lattice(n, ics.SEED_BASE + 3 + 1000 * b, -2.0, 0.5, 0.1) for b, n in enumerate(sizes)]
    npart = sum(n ** 3 for n in sizes)
    res = {}
    for prec in ("fp64", "mixed"):
        ctx = nyx_amd.PMContext([nyx_amd.BoxGeom((n, n, n)) for n in sizes], precision=prec)
        for b, (pos, vel, m) in enumerate(inputs):
            ctx.upload(b, pos, vel, m_const=m)
        tuned = ctx.autotune(fp64=prec == "fp64", fp32=prec == "mixed")  # with the particle tile
        log(prec, "autotuned", tuned)
        ctx.pm_init(1.0, 1.0, a.tol)
        ctx.pm_steps(2, 1e-3, rel_tol=a.tol, use_graph=True)  # capture + warm up
        ctx.sync()
        t0 = time.perf_counter()
        it = ctx.pm_steps(steps, 1e-3, rel_tol=a.tol, use_graph=True)
        ctx.sync()
        tb = time.perf_counter() - t0
        accel = None
        if prec == "mixed":
            # the accelerations after the timed steps against an FP64 solve of the
            # same densities from a zero guess, over all boxes' particles
            gm = [ctx.gather_accel(b) for b in range(len(sizes))]
            ctx.set_precision("fp64")
            ctx.assign_density(want=False)
            ctx.solve_for_phi(1.0, 1.0, a.tol, 0.0, use_guess=False)
            g64 = [ctx.gather_accel(b) for b in range(len(sizes))]
            num = sum(float(np.sum((x - y) ** 2)) for x, y in zip(gm, g64))
            den = sum(float(np.sum(y ** 2)) for y in g64)
            worst = max(float(np.linalg.norm(x - y) / np.linalg.norm(y)) for x, y in zip(gm, g64))
            accel = {"rel_l2_all_boxes": (num / den) ** 0.5, "rel_l2_worst_box": worst}
            log("mixed vs fp64 accel", accel)
            del gm, g64
        ctx.close()
        # one after another
        tseq = 0.0
        for b, (pos, vel, m) in enumerate(inputs):
            c1 = nyx_amd.PMContext(nyx_amd.BoxGeom((sizes[b],) * 3), precision=prec)
            c1.upload(0, pos, vel, m_const=m)
            c1.autotune(fp64=prec == "fp64", fp32=prec == "mixed")  # every box's own context
            c1.pm_init(1.0, 1.0, a.tol)
            c1.pm_steps(2, 1e-3, rel_tol=a.tol, use_graph=True)
            c1.sync()
            t0 = time.perf_counter()
            c1.pm_steps(steps, 1e-3, rel_tol=a.tol, use_graph=True)
            c1.sync()
            tseq += time.perf_counter() - t0
            c1.close()
        res[prec] = {"batched": round(npart * steps / tb, 1), "sequential": round(npart * steps / tseq, 1),
                     "speedup": round(tseq / tb, 3), "ms_per_step_batched": round(tb / steps * 1e3, 3),
                     "vcycles_last_step": [int(v) for v in it[:1]] + [int(it[16]), int(it[32])]}
        if accel:
            res[prec]["accel_vs_fp64"] = accel
        log(prec, res[prec])
    out = {"metric": "batched PM KDK particle-steps/s, 48 boxes (16x64^3 + 16x96^3 + 16x128^3)",
           "value": res["mixed"]["batched"], "unit": "particle-steps/s", "n_gpus": 1, "steps": steps,
           "warmup": 2, "ms_per_step": res["mixed"]["ms_per_step_batched"], "higher_is_better": True,
           "scaling": "weak", "vs_baseline": None, "dtype": "f32+f64", "data": "synthetic",
           "config": {"workload": "48 periodic boxes, lattice + GRF (k^-2, rms 0.5 cell), hipGraph-captured steps",
                      "particles": npart, "graph": True},
           "fp64": res["fp64"], "mixed": res["mixed"]}
    emit(out, a)


def main_grf256(a):
    """The FP64 drop-in configuration: 256^3 cells and particles, lattice + GRF
    (P(k) ~ k^-2, rms 0.5 cell, v = 0.1 psi), 10 KDK steps, dt 1e-3, tol 1e-10."""
    import nyx_amd
    from nyx_amd import ics
    n = 256
    steps = 10
    pos, vel, m = ics.grf_lattice(n, ics.SEED_BASE + 1, -2.0, 0.5, 0.1)
    ctx = nyx_amd.PMContext(nyx_amd.BoxGeom((n, n, n)), precision="fp64")
    ctx.upload(0, pos, vel, m_const=m)
    ctx.pm_init(1.0, 1.0, a.tol)
    ctx.sync()
    t0 = time.perf_counter()
    it = ctx.pm_steps(steps, 1e-3, rel_tol=a.tol, use_graph=a.graph)
    ctx.sync()
    t = time.perf_counter() - t0
    ctx.close()
    out = {"metric": "FP64 PM KDK particle-steps/s, 256^3 GRF drop-in", "value": round(n ** 3 * steps / t, 1),
           "unit": "particle-steps/s", "n_gpus": 1, "steps": steps, "warmup": 0,
           "ms_per_step": round(t / steps * 1e3, 3), "higher_is_better": True, "scaling": "weak",
           "vs_baseline": None, "dtype": "f64", "data": "synthetic",
           "config": {"workload": "256^3 cells/particles, GRF k^-2 rms 0.5 cell, 10 steps", "graph": bool(a.graph)},
           "vcycles_last_step": int(it[0])}
    emit(out, a)


def power_spectrum_device(rho, n):
    """P(k) of delta = rho/mean - 1 on an n^3 periodic grid (torch, on rho's
    device): |delta_k|^2 averaged over |k| shells of width 2 pi / L (integer
    bins of |k| L / 2 pi, rounded), rfft modes with 0 < kx < n/2 counted twice.
    Returns (bin centres, P) as numpy arrays for bins 1 .. n/2."""
    import torch
    d = rho.view(n, n, n)
    d = d / d.mean() - 1.0
    dk = torch.fft.rfftn(d)
    del d
    p = (dk.real * dk.real + dk.imag * dk.imag) / float(n) ** 6
    del dk
    f = torch.fft.fftfreq(n, d=1.0 / n, dtype=torch.float64, device=rho.device)
    fr = torch.fft.rfftfreq(n, d=1.0 / n, dtype=torch.float64, device=rho.device)
    kz, ky, kx = f.view(n, 1, 1), f.view(1, n, 1), fr.view(1, 1, n // 2 + 1)
    kb = torch.round(torch.sqrt(kx * kx + ky * ky + kz * kz)).to(torch.int64)
    wt = torch.where((kx > 0) & (kx < n / 2), 2.0, 1.0).to(torch.float64).expand_as(p)
    nb = n // 2 + 1
    kb = kb.clamp_max(nb).reshape(-1)
    ps = torch.zeros(nb + 1, dtype=torch.float64, device=rho.device).index_add_(0, kb, (p * wt).reshape(-1))
    cnt = torch.zeros(nb + 1, dtype=torch.float64, device=rho.device).index_add_(0, kb, wt.reshape(-1))
    ps, cnt = ps.cpu().numpy(), cnt.cpu().numpy()
    k = np.arange(1, nb)
    return k, ps[1:nb] / np.maximum(cnt[1:nb], 1.0)


def main_grf1024(a):
    """1024^3 mixed precision: lattice + GRF (P(k) ~ k^-1.5, rms 1 cell, v = 0.1
    psi), dt 5e-4, 50 KDK steps.  After every step of the mixed trajectory an
    FP64 solve of the same density from a zero guess (FP64 V-cycles only) gives
    the reference accelerations: rel-L2 = |g_mixed - g_fp64|_2 / |g_fp64|_2 over
    all particles (the mixed phi is put back before the next step).  Then an FP64 trajectory
    from the same initial conditions; the step-50 density power spectra (CIC rho
    -> FFT -> |delta_k|^2 in |k| shells of width 2 pi) of the two must agree to
    0.1% per bin below half the Nyquist wavenumber.  Initial conditions are made
    on the GPU (ics.grf_lattice_device) and handed over as device pointers."""
    import torch
    import nyx_amd
    from nyx_amd import ics
    n = a.n if a.n != 512 else 1024
    steps = a.steps if a.steps != 5 else 50
    dt, tol = 5e-4, a.tol
    N = n ** 3
    seed = ics.SEED_BASE + 4

    def fresh_ics(ctx):
        pos, vel, m = ics.grf_lattice_device(n, seed, -1.5, 1.0, 0.1)
        ctx.upload_device(0, N, [t.data_ptr() for t in pos + vel], m_const=m)
        ctx.sync()
        del pos, vel
        torch.cuda.empty_cache()

    log(f"grf1024: n={n}, {steps} steps, dt={dt}")
    # the checked trajectory with dense tiles (a full sort after every
    # kick-drift): its per-step accuracy check holds 60 GB of comparison arrays,
    # and slack slots would add ~35 GB; the two timed trajectories keep the
    # library's default slack (incremental re-sorts)
    slack = -1.0
    ctx = nyx_amd.PMContext(nyx_amd.BoxGeom((n, n, n)), precision="mixed", resort_slack=slack)
    tuned = ctx.autotune(fp64=True, fp32=True) if not a.no_tune else None
    fresh_ics(ctx)
    log("ICs uploaded; device bytes", ctx.device_bytes())
    ctx.pm_init(1.0, 1.0, tol)
    gm = torch.empty(3 * N, dtype=torch.float64, device="cuda")
    g64 = torch.empty(3 * N, dtype=torch.float64, device="cuda")
    phi_m = torch.empty(N, dtype=torch.float64, device="cuda")
    rel, t_steps, its = [], 0.0, []
    for s in range(steps):
        ctx.sync()
        t0 = time.perf_counter()
        it = ctx.pm_steps(1, dt, rel_tol=tol)
        ctx.sync()
        t_steps += time.perf_counter() - t0
        its.append(int(it[0]))
        ctx.gather_accel_device(gm.data_ptr())
        ctx.get_phi_device(phi_m.data_ptr())
        ctx.set_precision("fp64")
        ctx.assign_density(want=False)
        ctx.solve_for_phi(1.0, 1.0, tol, 0.0, use_guess=False)
        ctx.gather_accel_device(g64.data_ptr())
        ctx.set_phi_device(phi_m.data_ptr())
        ctx.set_precision("mixed")
        num = torch.linalg.vector_norm(gm - g64).item()
        den = torch.linalg.vector_norm(g64).item()
        rel.append(num / den)
        log(f"  step {s + 1}: {its[-1]} V-cycles, accel rel-L2 {rel[-1]:.3e}")
    del gm, g64, phi_m
    torch.cuda.empty_cache()
    rho = torch.empty(N, dtype=torch.float64, device="cuda")
    ctx.assign_density_device(rho.data_ptr())
    k, p_mix = power_spectrum_device(rho, n)
    del rho
    ctx.close()  # a fresh context per trajectory (one set of particle buffers at a time)
    torch.cuda.empty_cache()
    # the mixed trajectory again from the same initial conditions, all steps in
    # one call (paired kicks), as the FP64 trajectory below is timed
    ctx = nyx_amd.PMContext(nyx_amd.BoxGeom((n, n, n)), precision="mixed")
    if not a.no_tune:
        ctx.autotune(fp64=False, fp32=True)
    fresh_ics(ctx)
    ctx.pm_init(1.0, 1.0, tol)
    ctx.sync()
    t0 = time.perf_counter()
    ctx.pm_steps(steps, dt, rel_tol=tol)
    ctx.sync()
    t_mix = time.perf_counter() - t0
    ctx.close()
    torch.cuda.empty_cache()
    log(f"mixed: {t_mix / steps * 1e3:.2f} ms/step in one call; FP64 trajectory")
    ctx = nyx_amd.PMContext(nyx_amd.BoxGeom((n, n, n)), precision="fp64")
    if not a.no_tune:
        ctx.autotune(fp64=True, fp32=False)
    fresh_ics(ctx)
    ctx.pm_init(1.0, 1.0, tol)
    ctx.sync()
    t0 = time.perf_counter()
    ctx.pm_steps(steps, dt, rel_tol=tol)
    ctx.sync()
    t64 = time.perf_counter() - t0
    rho = torch.empty(N, dtype=torch.float64, device="cuda")
    ctx.assign_density_device(rho.data_ptr())
    _, p_64 = power_spectrum_device(rho, n)
    del rho
    ctx.close()
    low = k < n / 4  # below half the Nyquist wavenumber (k_Nyq = n/2 in these units)
    dev = np.abs(p_mix[low] / p_64[low] - 1.0)
    out = {"metric": "PM KDK at 1024^3 GRF: mixed vs FP64 accel rel-L2 per step and step-50 P(k)",
           "value": round(N * steps / t_mix, 1), "unit": "particle-steps/s (mixed, all steps in one call)",
           "n_gpus": 1, "steps": steps, "warmup": 0, "ms_per_step": round(t_mix / steps * 1e3, 3),
           "ms_per_step_checked": round(t_steps / steps * 1e3, 3),
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32+f64",
           "data": "synthetic",
           "config": {"workload": f"{n}^3 cells/particles, GRF k^-1.5 rms 1 cell, v = 0.1 psi, dt {dt}, "
                                  f"{steps} steps, rel tol {tol:g}", "autotuned": tuned},
           "fp64": {"value": round(N * steps / t64, 1), "ms_per_step": round(t64 / steps * 1e3, 3)},
           "mixed_over_fp64": round(t64 / t_mix, 3),
           "accel_rel_l2": {"max": max(rel), "last": rel[-1], "per_step": [float(f"{v:.3e}") for v in rel]},
           "vcycles_mixed": its,
           "power_spectrum": {"bins_compared": int(low.sum()), "max_rel_dev": float(dev.max()),
                              "pass": bool(dev.max() <= 1e-3)}}
    emit(out, a)


if __name__ == "__main__":
    main()
