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


TILE = (12, 8, 8)  # the library's default deposit/kick tile


def tile_halo_cells(n, tile):
    """Low-halo cells over all tiles of a box: sum of (bx+1)(by+1)(bz+1) - bx*by*bz
    over the tiles (the last tile along a dimension may be partial)."""
    sizes = []
    for d in range(3):
        full, rem = divmod(n[d], tile[d])
        sizes.append([tile[d]] * full + ([rem] if rem else []))
    return sum((bx + 1) * (by + 1) * (bz + 1) - bx * by * bz
               for bx in sizes[0] for by in sizes[1] for bz in sizes[2])


def kernel_bytes(name, ncell, npart, halo_cells, mixed_sz):
    """Algorithmic (compulsory) bytes of one launch of a kernel class at level 0."""
    sz = mixed_sz
    N, P = ncell, npart
    table = {
        "smooth.d0": (sz + sz / 2 + sz / 2) * N,          # read e (all), r (colour), write e (colour)
        "sweep.d0": 3 * sz * N,                            # fused red+black: read e, r; write e
        "sweep0.d0": 2 * sz * N,                           # first sweep from e = 0: read r; write e
        "sweep_p.d0": (3 * sz + sz / 8) * N,               # + coarse correction read (prolongation on load)
        "sweep_u.d0": (2 * sz + 16) * N,                   # read e, r; phi += e (read/write FP64 phi)
        "resid_restrict.d0": (2 * sz + sz / 8) * N,        # read e, r; write coarse r
        "prolong.d0": (2 * sz + sz / 8) * N,               # read/write e, read coarse e
        "fill.d0": sz * N,
        "residual": (16 + sz) * N,                         # read phi, f (FP64); write r
        "res_sweep0": (16 + 2 * sz) * N,                   # residual + first sweep: write r and e
        "update": (16 + sz) * N,                           # read/write phi, read e
        "deposit": 24 * P + 8 * N + 8 * halo_cells,        # read positions; write own + halos
        "fold": 8 * N + 8 * halo_cells + 8 * N,            # read own + halos; write f (rho kept in own)
        "kick_drift": 96 * P + 8 * N,                      # read/write x,v; read phi
        "kick_pair": 96 * P + 8 * N,                       # kick + next kick-drift: same traffic
        "kick": 72 * P + 8 * N,                            # read x,v; write v; read phi
        "scatter": 4 * P + 2 * 56 * P,                     # key; move x,v,id
        "keys": 24 * P + 4 * P,
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


def run_timed(ctx, precision, steps, warmup, dt, tol, use_graph, timed_stages=("kick_pair", "residual")):
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
    ctx.sync()
    t0 = time.perf_counter()
    ctx.pm_steps(steps, dt, rel_tol=tol, use_graph=use_graph)
    ctx.sync()
    t_inst = time.perf_counter() - t0
    ctx.timing_enable(False)
    st = stage_table(ctx, names, ncell, npart, halo_cells, sz, t_inst * 1e3)
    dname = dominant_stage(st, steps)
    if dom is None or dname not in dom.get("candidates", {}):
        # a stage outside the timed candidates leads the stage table: report it
        # with its mean from these steps
        dom = {"stage": dname, "launches": st[dname]["launches"], "mean_ms": st[dname]["mean_ms"],
               "timed_in": "instrumented steps"}
    return st, dom, t_inst


def stage_table(ctx, names, ncell, npart, halo_cells, sz, elapsed_ms):
    out = {}
    for nm in names:
        cnt, mean, tot = ctx.timing(nm)
        if cnt == 0:
            continue
        b = kernel_bytes(nm, ncell, npart, halo_cells, sz)
        e = {"launches": cnt, "mean_ms": round(mean, 4), "total_ms": round(tot, 2),
             "share": round(tot / elapsed_ms, 4)}
        if b is not None and mean > 0:
            e["gbs"] = round(b / (mean * 1e-3) / 1e9, 1)
            e["bytes_per_launch"] = int(b)
        out[nm] = e
    return out


def cpu_baseline(n_cpu, threads, steps):
    import oracle
    from nyx_amd import ics
    oracle.set_num_threads(threads)
    pos, vel, m = ics.plummer_clustered(n_cpu, ics.SEED_BASE + 2, n_spheres=max(1, int(2000 * (n_cpu / 512) ** 3)))
    n = (n_cpu,) * 3
    pos = [np.ascontiguousarray(a) for a in pos]
    vel = [np.ascontiguousarray(a) for a in vel]
    phi, _, _ = oracle.solve_for_phi(n, (0, 0, 0), (1, 1, 1), pos, m_const=m, rel_tol=1e-10)
    t0 = time.perf_counter()
    stages = np.zeros(3)
    for _ in range(steps):
        _, st = oracle.pm_step(n, (0, 0, 0), (1, 1, 1), pos, vel, phi, m_const=m, dt=1e-3, rel_tol=1e-10)
        stages += st[4:7]
    t = time.perf_counter() - t0
    return {"value": round(n_cpu ** 3 * steps / t, 1), "unit": "particle-steps/s", "cores": threads,
            "kind": "port",
            "sample": f"{steps} canonical KDK steps (deposit, MG to rel 1e-10 with {int(st[0])} V-cycles in the "
                      f"last, gradient, kick-drift, kick) of the same clustered recipe at {n_cpu}^3 "
                      f"cells/particles ({max(1, int(2000 * (n_cpu / 512) ** 3))} Plummer spheres), FP64, "
                      f"{t:.2f} s; stage s: deposit {stages[0]:.2f} solve {stages[1]:.2f} "
                      f"grav+kick {stages[2]:.2f}"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n", type=int, default=512)
    ap.add_argument("--cpu-n", type=int, default=384)
    ap.add_argument("--cpu-steps", type=int, default=4)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--graph", action="store_true")
    ap.add_argument("--tol", type=float, default=1e-10)
    ap.add_argument("--json-out", default=None)
    ap.add_argument("--modes", default="mixed,fp64", help="precisions to time (comma list)")
    ap.add_argument("--no-accuracy", action="store_true", help="skip the mixed-vs-FP64 acceleration check")
    ap.add_argument("--no-tune", action="store_true", help="skip the smoother-shape autotuner")
    ap.add_argument("--kick-chunk", type=int, default=0, help="particles per kick work item (0: default)")
    ap.add_argument("--phi-guess", default="default", choices=["default", "previous", "extrapolate"],
                    help="initial guess of each step's solve (default = previous phi, as Nyx)")
    ap.add_argument("--workload", default="clustered512",
                    choices=["clustered512", "batch48", "grf256", "grf1024"],
                    help="clustered512: the headline line; batch48: 48 boxes (16 each of 64^3, 96^3, "
                         "128^3) in one launch per phase vs one after another; grf256: FP64 drop-in, "
                         "256^3 GRF, 10 steps; grf1024: 1024^3 GRF, 50 mixed steps with per-step "
                         "accel rel-L2 against FP64 and the step-50 power spectrum check (--n to shrink)")
    a = ap.parse_args()
    if a.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU only: this build has no multi-GPU path (see DESIGN.md, "
              "'How the path would shard'); run with --gpus 1", file=sys.stderr)
        sys.exit(2)
    if a.workload == "batch48":
        return main_batch(a)
    if a.workload == "grf1024":
        return main_grf1024(a)
    if a.workload == "grf256":
        return main_grf256(a)

    import nyx_amd
    from nyx_amd import ics
    n = a.n
    dt = 1e-3
    log(f"generating {n}^3 clustered particles")
    t = time.perf_counter()
    pos, vel, m = ics.plummer_clustered(n, ics.SEED_BASE + 2, n_spheres=max(1, int(2000 * (n / 512) ** 3)))
    log(f"  {time.perf_counter() - t:.1f} s")
    ctx = nyx_amd.PMContext(nyx_amd.BoxGeom((n, n, n)), precision="fp64", kick_chunk=a.kick_chunk,
                            phi_guess=a.phi_guess)
    ctx.upload(0, pos, vel, m_const=m)
    del pos, vel
    npart = n ** 3
    ncell = n ** 3
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
    if not a.no_accuracy:
        g64 = ctx.gather_accel()
        ctx.set_precision("mixed")
        it_m, _ = ctx.solve_for_phi(1.0, 1.0, a.tol, 0.0, use_guess=False)
        gmx = ctx.gather_accel()
        rel_l2 = float(np.sqrt(np.sum((gmx - g64) ** 2)) / np.sqrt(np.sum(g64 ** 2)))
        del g64, gmx
        log(f"accel rel-L2 mixed vs fp64 = {rel_l2:.3e}  (mixed cold-start cycles {int(it_m[0])})")
    modes = a.modes.split(",")

    names = ["residual", "norm", "update", "deposit", "fold", "rhs", "kick_drift", "kick_pair", "kick", "scatter",
             "scan", "coarse", "migrate", "res_sweep0", "guess",
             "items", "keys", "bottom"] + \
            [f"{k}.d{d}" for k in ("sweep0", "sweep", "sweep_p", "sweep_u", "smooth", "resid_restrict", "prolong",
                                    "fill")
             for d in range(0, 12)]
    halo_cells = tile_halo_cells((n, n, n), TILE)

    t_mix = t_64 = None
    it_mix = it_64 = None
    rs_mix = rs_64 = None
    rel_l2_end = None
    st_mix, st_64 = {}, {}
    dom_mix = dom_64 = None
    if "mixed" in modes:
        log(f"mixed: {a.warmup} warmup + {a.steps} timed steps")
        t_mix, it_mix, rs_mix, dom_mix = run_timed(ctx, "mixed", a.steps, a.warmup, dt, a.tol, a.graph)
        log(f"  {t_mix / a.steps * 1e3:.2f} ms/step, {it_mix} V-cycles")
        if not a.no_accuracy:
            # the state the timed mixed steps left: their phi against an FP64 solve of
            # the same density from a zero guess
            gmx = ctx.gather_accel()
            ctx.set_precision("fp64")
            ctx.assign_density(want=False)
            ctx.solve_for_phi(1.0, 1.0, a.tol, 0.0, use_guess=False)
            g64 = ctx.gather_accel()
            rel_l2_end = float(np.sqrt(np.sum((gmx - g64) ** 2)) / np.sqrt(np.sum(g64 ** 2)))
            del gmx, g64
            log(f"  accel rel-L2 mixed vs fp64 after the timed steps = {rel_l2_end:.3e}")
    if "fp64" in modes:
        log(f"fp64: {a.warmup} warmup + {a.steps} timed steps")
        t_64, it_64, rs_64, dom_64 = run_timed(ctx, "fp64", a.steps, a.warmup, dt, a.tol, a.graph)
        log(f"  {t_64 / a.steps * 1e3:.2f} ms/step, {it_64} V-cycles")
    # the stage tables: further steps of each mode, every stage bracketed by events
    if "mixed" in modes:
        st_mix, dom_mix, ti = run_instrumented(ctx, "mixed", a.steps, dt, a.tol, a.graph, names, ncell, npart,
                                               halo_cells, 4, dom_mix)
        log(f"  mixed stage table: {a.steps} instrumented steps, {ti / a.steps * 1e3:.2f} ms/step")
    if "fp64" in modes:
        st_64, dom_64, ti = run_instrumented(ctx, "fp64", a.steps, dt, a.tol, a.graph, names, ncell, npart,
                                             halo_cells, 8, dom_64)
        log(f"  fp64 stage table: {a.steps} instrumented steps, {ti / a.steps * 1e3:.2f} ms/step")
    ctx.close()
    if t_mix is None:  # FP64 only: report it as the value
        t_mix, it_mix, st_mix, rs_mix, dom_mix = t_64, it_64, st_64, rs_64, dom_64
    if t_64 is None:
        t_64, it_64, st_64, rs_64, dom_64 = t_mix, it_mix, st_mix, rs_mix, dom_mix

    v_mix = npart * a.steps / t_mix
    v_64 = npart * a.steps / t_64
    # dominant kernel (largest share of a steady-state step, from the stage table);
    # its mean launch duration is the one measured inside the timed steps
    dname, dmean = dom_mix["stage"], dom_mix["mean_ms"]
    b = kernel_bytes(dname, ncell, npart, halo_cells, 4)
    achieved = b / (dmean * 1e-3) / 1e9 if b and dmean > 0 else None
    roof = {"bound": "hbm", "kernel": dname, "achieved": round(achieved, 1) if achieved else None,
            "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4) if achieved else None,
            "traffic": measured_traffic(dname, n), "bytes_per_launch": int(b) if b else None,
            "mean_launch_ms": dmean, "launches_timed": dom_mix["launches"], "timed_in": dom_mix["timed_in"]}

    def smoother_roof(st, sz):
        e = st.get("sweep.d0")
        if not e:
            return None
        b = kernel_bytes("sweep.d0", ncell, npart, halo_cells, sz)
        ach = b / (e["mean_ms"] * 1e-3) / 1e9
        return {"kernel": "sweep.d0 (fused red+black sweep, finest level)", "achieved": round(ach, 1),
                "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(ach / HBM_PEAK_GBS, 4),
                "bytes_per_launch": int(b), "mean_launch_ms": e["mean_ms"]}

    cpu = None
    if not a.no_cpu:
        try:
            aff = len(os.sched_getaffinity(0))
        except Exception:
            aff = os.cpu_count() or 1
        threads = max(1, min(16, aff))
        log(f"CPU baseline at {a.cpu_n}^3 on {threads} threads")
        cpu = cpu_baseline(a.cpu_n, threads, a.cpu_steps)
        log("  ", cpu["value"], "particle-steps/s")

    out = {
        "metric": METRIC, "value": round(v_mix, 1), "unit": "particle-steps/s", "n_gpus": 1,
        "steps": a.steps, "warmup": a.warmup, "ms_per_step": round(t_mix / a.steps * 1e3, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32+f64",
        "data": "synthetic",
        "config": {"workload": f"{n}^3 cells, {n}^3 particles, 20% uniform + 80% in "
                               f"{max(1, int(2000 * (n / 512) ** 3))} Plummer spheres, mixed precision "
                               f"(FP32 V-cycles in FP64 refinement, FP64 particles), dt 1e-3, rel tol {a.tol:g}",
                   "cells": ncell, "particles": npart, "graph": bool(a.graph), "autotuned": tuned},
        "fp64": {"value": round(v_64, 1), "ms_per_step": round(t_64 / a.steps * 1e3, 3), "vcycles": it_64,
                 "dtype": "f64", "resort": rs_64},
        "resort": rs_mix,
        "mixed_over_fp64": round(v_mix / v_64, 3),
        "vcycles_mixed": it_mix,
        "accel_rel_l2": rel_l2,
        "accel_rel_l2_after_steps": rel_l2_end,
        "phi_guess": a.phi_guess,
        "roofline": roof,
        "smoother_roofline": {"mixed": smoother_roof(st_mix, 4), "fp64": smoother_roof(st_64, 8)},
        "stages_note": "stage tables: further steps of each mode (as many as timed), run after both "
                       "timed runs with HIP events around every stage; the timed steps bracket only "
                       "the paired kick and the level-0 residual (the dominant-kernel candidates)",
        "stages_mixed": st_mix,
        "stages_fp64": st_64,
        "cpu_baseline": cpu,
    }
    line = json.dumps(out)
    if a.json_out:
        with open(a.json_out, "w") as fh:
            fh.write(line + "\n")
    print(line, flush=True)


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
        tuned = ctx.autotune(fp64=prec == "fp64", fp32=prec == "mixed")
        log(prec, "autotuned", tuned)
        for b, (pos, vel, m) in enumerate(inputs):
            ctx.upload(b, pos, vel, m_const=m)
        ctx.pm_init(1.0, 1.0, a.tol)
        ctx.pm_steps(2, 1e-3, rel_tol=a.tol, use_graph=True)  # capture + warm up
        ctx.sync()
        t0 = time.perf_counter()
        it = ctx.pm_steps(steps, 1e-3, rel_tol=a.tol, use_graph=True)
        ctx.sync()
        tb = time.perf_counter() - t0
        ctx.close()
        # one after another
        tseq = 0.0
        for b, (pos, vel, m) in enumerate(inputs):
            c1 = nyx_amd.PMContext(nyx_amd.BoxGeom((sizes[b],) * 3), precision=prec)
            if b in (0, 16, 32):  # tune once per box size; the context is rebuilt per box
                seq_tuned = c1.autotune(fp64=prec == "fp64", fp32=prec == "mixed")
            c1.upload(0, pos, vel, m_const=m)
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
    # dense tiles (a full sort after every kick-drift): the per-step accuracy
    # check holds 60 GB of comparison arrays, and slack slots would add ~35 GB
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
    ctx.close()  # a fresh context for the FP64 trajectory (one set of particle buffers at a time)
    torch.cuda.empty_cache()
    log("mixed trajectory done; FP64 trajectory")
    ctx = nyx_amd.PMContext(nyx_amd.BoxGeom((n, n, n)), precision="fp64", resort_slack=slack)
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
           "value": round(N * steps / t_steps, 1), "unit": "particle-steps/s (mixed, one step per call)",
           "n_gpus": 1, "steps": steps, "warmup": 0, "ms_per_step": round(t_steps / steps * 1e3, 3),
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32+f64",
           "data": "synthetic",
           "config": {"workload": f"{n}^3 cells/particles, GRF k^-1.5 rms 1 cell, v = 0.1 psi, dt {dt}, "
                                  f"{steps} steps, rel tol {tol:g}", "autotuned": tuned},
           "fp64": {"value": round(N * steps / t64, 1), "ms_per_step": round(t64 / steps * 1e3, 3)},
           "accel_rel_l2": {"max": max(rel), "last": rel[-1], "per_step": [float(f"{v:.3e}") for v in rel]},
           "vcycles_mixed": its,
           "power_spectrum": {"bins_compared": int(low.sum()), "max_rel_dev": float(dev.max()),
                              "pass": bool(dev.max() <= 1e-3)}}
    emit(out, a)


if __name__ == "__main__":
    main()
