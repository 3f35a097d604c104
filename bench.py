#!/usr/bin/env python
"""Headline benchmark: zone-updates/s of the 3D MHD blast at 256^3 in FP64 on one MI355X.

Workload: 256^3 periodic cube [-0.5, 0.5]^3, gamma = 5/3, rho = 1 with seeded 1e-3
multiplicative noise, p = 0.1 ambient / 10 for r < 0.1, B = (1/sqrt2, 1/sqrt2, 0);
PPM + HLLD + RK3 + CT + fofc + flattening, CFL 0.3 (synthetic seeded initial state from the
library's own problem generator). A step is one full RK3 step (three stages) of the
device-resident pipeline; the timed region is K steps between two events on the
pipeline's stream, after W warm-up steps, bracketed by stream synchronisation, with no other
event inside it.

Output: one JSON line. `kernels` is a second pass over the same W + K steps (a fresh mesh
follows the same trajectory bitwise, fofc firings included) with HIP events around every
launch on the pipeline's stream: per-class milliseconds per step, the fofc branch included.
`roofline` is for the dominant kernel (the face-flux kernel), its average launch time from
that pass; its `traffic` is read from the committed rocprofv3 counter summary of the same
workload. `step_roofline` applies the reference's traffic model (1040 B per cell per RK stage,
README.md:45-47) to the whole step; `stream_probe` is the device's streaming rate with the
stage kernels' access pattern (chq_stream_probe); `cpu_baseline` times the CPU oracle on this
host's cores on a few steps of the same workload; `mixed` is the same workload on the
mixed-precision path (FP32 reconstruction, Riemann solve and EMFs; FP64 con2prim, RK combine,
CT); `batch` the batched Orszag-Tang ensemble; `serve` the served stream.
"""
import argparse
import json
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)

METRIC = "zone-updates/s at 256³ FP64 & % HBM roofline; served runs/s + p99 latency"
HBM_PEAK = 8.0e12                 # MI355X HBM3E, bytes/s
MODEL_BYTES_PER_ZONE_STAGE = 1040  # README.md:45-47 (130 doubles per cell per RK stage)
# Algorithmic bytes of one flux launch (one direction): the reference's face-flux traffic, 52.0
# doubles per active cell per RK stage over the three directions (reads of the 6 primitives
# and 2 cell-B stencils plus the face-normal B, writes of 6 fluxes and 2 face EMFs), i.e.
# 52.0 / 3 doubles per active cell per launch.
FLUX_DOUBLES_PER_CELL_STAGE = 52.0

# Memory traffic per flux-kernel launch (mean of the x, y and z launches) comes from the
# committed rocprofv3 counter summary of this workload (read + write MiB per launch, scaled by
# the counters' reading of the streaming probe's known bytes).
TRAFFIC_PROFILE = os.path.join(HERE, "profiles", "blast256_fp64_hbm_traffic.md")
FLUX_KERNELS = ("void chq::fluxk::k_flux_x<double", "void chq::fluxk::k_flux_m<double")
STREAM_PROBE = os.path.join(HERE, "chasqui_amd", "_lib", "chq_stream_probe")

DECK = """
[mesh]
nx1 = 256
nx2 = 256
nx3 = 256
x1min = -0.5
x1max = 0.5
x2min = -0.5
x2max = 0.5
x3min = -0.5
x3max = 0.5
[time]
integrator = rk3
cfl = 0.3
tlim = 1.0e9
nlim = -1
ndiag = 0
[mhd]
gamma = 1.6666666666666667
reconstruct = ppm
rsolver = hlld
flatten = true
fofc = true
[problem]
name = blast
pres_in = 10.0
pres_out = 0.1
radius = 0.1
b0 = 1.0
angle = 45.0
perturb = 1.0e-3
seed = 42
"""


def cpu_baseline(n, steps):
    """The CPU oracle (plain C++/OpenMP restatement of the reference) on this host's cores, on
    a bounded sample of the same workload (`steps` RK3 steps of the same 256^3 blast)."""
    os.environ.setdefault("OMP_PROC_BIND", "spread")
    os.environ.setdefault("OMP_PLACES", "threads")
    sys.path.insert(0, os.path.join(HERE, "oracle"))
    import pyoracle as po
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    po.set_threads(threads)
    model = ""
    try:
        with open("/proc/cpuinfo") as f:
            model = next((ln.split(":", 1)[1].strip() for ln in f if ln.startswith("model name")), "")
    except OSError:
        pass
    o = po.OracleSim(DECK, [f"mesh.nx1={n}", f"mesh.nx2={n}", f"mesh.nx3={n}"])
    t0 = time.perf_counter()
    taken = o.step(steps)
    wall = time.perf_counter() - t0
    o.close()
    return {"value": n ** 3 * taken / wall, "unit": "zone-updates/s", "cores": threads, "kind": "port",
            "sample": f"{taken} RK3 steps of the same {n}^3 blast on {threads} OpenMP threads "
                      f"({os.cpu_count()} logical CPUs, {model}), {wall:.1f} s wall "
                      f"(oracle/, -O3 -ffp-contract=off, OMP_PROC_BIND=spread)"}


def _blast_mesh(n, precision):
    from chasqui_amd.deck import Deck
    deck = Deck(DECK, [f"mesh.nx1={n}", f"mesh.nx2={n}", f"mesh.nx3={n}", f"mhd.precision={precision}"])
    mesh = deck.make_mesh()
    deck.close()
    return mesh


def run(n, precision, steps, warmup, breakdown=False, state=None):
    """K timed steps of the device-resident pipeline on the n^3 blast (after W warm-up steps),
    with no per-launch events in the timed region; returns (ms, status, breakdown). With
    `breakdown`, the same W + K steps run again on a fresh mesh (the same trajectory, bitwise:
    the same fofc firings) with HIP events around every launch, giving per-class times on the
    pipeline's stream: flux, corner, update, bc, c2p, and fofc_branch (the repair, edge EMFs,
    re-update, boundaries and con2prim of stages that flagged cells; near zero when none did).
    With `state` (a dict), the final conserved state of the active cells is left in
    state["u0"]."""
    mesh = _blast_mesh(n, precision)
    mesh.step(warmup)
    mesh.sync()
    ms = mesh.step_timed(steps)
    st = mesh.status()
    if state is not None:
        u = mesh.download("u0")
        ng = (u.shape[3] - n) // 2
        act = slice(ng, ng + n)
        state["u0"] = u[:5, act, act, act].copy()
        state["cycle"] = st["cycle"]
        del u
    mesh.close()
    kernels = {}
    if breakdown:
        mesh = _blast_mesh(n, precision)
        mesh.step(warmup)
        mesh.sync()
        mesh.kernel_timing("all")
        ims = mesh.step_timed(steps)
        bd = mesh.kernel_timing_breakdown()
        mesh.kernel_timing(None)
        st2 = mesh.status()
        mesh.close()
        if st2["fofc_total"] != st["fofc_total"] or st2["t"] != st["t"]:
            raise RuntimeError("the instrumented pass did not follow the timed trajectory")
        for cls, (kms, kn) in bd.items():
            kernels["fofc_branch" if cls == "repair" else cls] = {
                "ms_per_step": kms / steps, "launches_per_step": kn / steps, "total_ms": kms, "launches": kn}
        kernels["instrumented_ms_per_step"] = ims / steps
    return ms, st, kernels


def profile_traffic(path, kernel_prefixes):
    """Mean read + write MiB per launch of the kernels whose names start with one of
    `kernel_prefixes`, from a committed rocprofv3 traffic summary (markdown table rows
    | `kernel` | launches | read MiB | write MiB |), in bytes; None if absent."""
    import re
    try:
        text = open(path).read()
    except OSError:
        return None
    vals = []
    for m in re.finditer(r"^\|\s*`([^`]+)`\s*\|\s*(\d+)\s*\|\s*([\d.]+)\s*\|\s*([\d.]+)\s*\|", text, re.M):
        name = m.group(1)
        if any(name.startswith(p) for p in kernel_prefixes):
            vals.append((float(m.group(3)) + float(m.group(4))) * 2 ** 20)
    return sum(vals) / len(vals) if vals else None


OT_DECK = """
[mesh]
nx1 = 64
nx2 = 64
x1min = 0.0
x1max = 1.0
x2min = 0.0
x2max = 1.0
[time]
integrator = rk3
cfl = 0.4
tlim = 0.5
ndiag = 0
[mhd]
gamma = 1.6666666666666667
reconstruct = ppmx
rsolver = hlld
[problem]
name = orszag_tang
"""

BW_DECK = """
[mesh]
nx1 = 400
x1min = -0.5
x1max = 0.5
bc_x1 = outflow
[time]
integrator = rk3
cfl = 0.4
tlim = 0.1
ndiag = 0
[mhd]
gamma = 2.0
reconstruct = ppm
rsolver = hlld
[problem]
name = shock_tube
"""


def ensemble_overrides(nruns, seed):
    """Batched Orszag-Tang ensemble: sizes drawn uniformly from {64 ... 512}^2, per-run seeded
    +-5% perturbation of the pressure and the field amplitude."""
    import numpy as np
    rng = np.random.default_rng(seed)
    sizes = [64, 96, 128, 192, 256, 384, 512]
    p0, b0 = 5.0 / (12.0 * np.pi), 1.0 / np.sqrt(4.0 * np.pi)
    out = []
    for _ in range(nruns):
        n = int(rng.choice(sizes))
        fp, fb = 1.0 + 0.05 * rng.uniform(-1, 1), 1.0 + 0.05 * rng.uniform(-1, 1)
        out.append([f"mesh.nx1={n}", f"mesh.nx2={n}", f"problem.pres={float(p0 * fp)!r}",
                    f"problem.b0={float(b0 * fb)!r}"])
    return out


def active_rho_e(m):
    """Density and total energy of a mesh's active cells (host copies)."""
    u = m.download("u0")
    box = (slice(m.ks, m.ke + 1), slice(m.js, m.je + 1), slice(m.is_, m.ie + 1))
    return u[0][box].copy(), u[4][box].copy()


def rel_l1(states, ref):
    """Relative L1 distance of density and total energy, summed over runs, from `ref`."""
    import numpy as np
    out = {}
    for q, name in ((0, "rho"), (1, "E")):
        num = sum(float(np.abs(a[q] - b[q]).sum()) for a, b in zip(states, ref))
        den = sum(float(np.abs(b[q]).sum()) for b in ref)
        out[name] = num / den
    return out


def ensemble(nruns, steps, precision, seed=7, tune=True, ref=None):
    """Config: `nruns` independent 2D Orszag-Tang runs (PPMX + HLLD + RK3) advanced as one batch
    (one launch per kernel class and stage, hipGraph replay), against the same runs stepped one
    after another on their own. Returns zone-updates/s of both and the speedup, and the runs'
    final density and energy; with `ref` (those of the FP64 batch) also the relative L1
    distance of this batch's final state from it (the same runs, the same number of steps)."""
    from chasqui_amd import Batch
    from chasqui_amd.deck import Deck
    ovs = ensemble_overrides(nruns, seed)
    cells = 0
    meshes = []
    for ov in ovs:
        d = Deck(OT_DECK, ov + [f"mhd.precision={precision}"])
        m = d.make_mesh()
        d.close()
        meshes.append(m)
        cells += int(ov[0].split("=")[1]) ** 2
    # sequential: each run alone, K steps (one warm-up step each, untimed)
    seq_ms = 0.0
    for m in meshes:
        m.step(1)
        m.sync()
        seq_ms += m.step_timed(steps)
    b = Batch(nruns)
    for m in meshes:
        b.add(m)
    b.step(1)
    b.sync()
    tuned = None
    if tune:
        tuned = b.tune(reps=3)
    b.step(1)
    b.sync()
    ms = b.step_timed(steps)
    grids = b.grids()
    b.close()
    fin = [active_rho_e(m) for m in meshes]   # after the timed steps
    for m in meshes:
        m.close()
    zu = float(cells) * steps
    res = {"runs": nruns, "cells": cells, "steps": steps, "precision": precision,
           "batched_zone_updates_per_s": zu / (ms * 1e-3), "batched_ms_per_step": ms / steps,
           "sequential_zone_updates_per_s": zu / (seq_ms * 1e-3), "sequential_ms_per_step": seq_ms / steps,
           "speedup_vs_sequential": seq_ms / ms, "tuned_sum_ms": tuned, "grids": grids}
    if ref is not None:
        res["l1_rel_vs_fp64"] = rel_l1(fin, ref)
    return res, fin


def ensembles(nruns, steps):
    """The FP64 and mixed batches of the same runs; the mixed one carries its L1 distance from
    the FP64 one."""
    r64, fin64 = ensemble(nruns, steps, "fp64")
    r32, _ = ensemble(nruns, steps, "mixed", ref=fin64)
    return {"fp64": r64, "mixed": r32}


def shard(n=256, ranks=8, steps=20, warmup=2):
    """One GPU's share of the n^3 blast split into `ranks` slabs along z: an n x n x (n/ranks)
    mesh with the same cell size, cut through the blast centre, periodic ghosts standing in for
    the halo exchange a multi-GPU run would do after each update. FP64, the headline scheme;
    zone-updates/s of the slab."""
    from chasqui_amd.deck import Deck
    nz = n // ranks
    half = 0.5 * nz / n
    d = Deck(DECK, [f"mesh.nx1={n}", f"mesh.nx2={n}", f"mesh.nx3={nz}", f"mesh.x3min={-half!r}",
                    f"mesh.x3max={half!r}"])
    mesh = d.make_mesh()
    d.close()
    mesh.step(warmup)
    mesh.sync()
    ms = mesh.step_timed(steps)
    mesh.close()
    return {"shape": [n, n, nz], "ranks": ranks, "ms_per_step": ms / steps,
            "zone_updates_per_s": float(n) * n * nz * steps / (ms * 1e-3),
            "note": "one of the slabs of the headline mesh; the halo exchange is not included"}


def batch_floor(nruns=64, steps=50):
    """Per-step cost of batches whose work is negligible, where the launches of a batched step
    set the time: `nruns` 2D Orszag-Tang runs of 32^2 and `nruns` 1D Brio-Wu tubes of 400 cells
    (PPM + HLLD + RK3, FP64), stepped as a batch with the captured graph; ms per batched step."""
    from chasqui_amd import Batch
    from chasqui_amd.deck import Deck
    res = {}
    for name, text, ov in (("orszag_tang_32x32", OT_DECK, ["mesh.nx1=32", "mesh.nx2=32", "mhd.reconstruct=ppm"]),
                           ("brio_wu_400", BW_DECK, ["mesh.nx1=400"])):
        meshes = []
        for _ in range(nruns):
            d = Deck(text, ov + ["time.nlim=-1", "time.tlim=1e9"])
            meshes.append(d.make_mesh())
            d.close()
        b = Batch(nruns)
        for m in meshes:
            b.add(m)
        b.step(2)
        b.sync()
        res[name + "_ms_per_step"] = b.step_timed(steps) / steps
        b.close()
        for m in meshes:
            m.close()
    res["runs"] = nruns
    return res


def serve_jobs(njobs, seed=11):
    """Served stream: seeded mix of 1D Brio-Wu shock tubes (400-4000 cells) run to their
    tlim = 0.1, 2D Orszag-Tang (128^2-1024^2) and 3D blasts (64^3-192^3) capped at 100 and 50
    steps so the stream finishes in minutes; reconstruction drawn from {ppm, wenoz}."""
    import numpy as np
    rng = np.random.default_rng(seed)
    jobs = []
    for _ in range(njobs):
        kind = rng.choice(["1d", "2d", "3d"])
        rec = str(rng.choice(["ppm", "wenoz"]))
        if kind == "1d":
            n = int(rng.integers(400, 4001))
            jobs.append((BW_DECK, [f"mesh.nx1={n}", f"mhd.reconstruct={rec}"]))
        elif kind == "2d":
            n = int(rng.choice([128, 256, 384, 512, 768, 1024]))
            jobs.append((OT_DECK, [f"mesh.nx1={n}", f"mesh.nx2={n}", f"mhd.reconstruct={rec}", "time.nlim=100"]))
        else:
            n = int(rng.choice([64, 96, 128, 160, 192]))
            jobs.append((DECK, [f"mesh.nx1={n}", f"mesh.nx2={n}", f"mesh.nx3={n}", f"mhd.reconstruct={rec}",
                                "time.nlim=50", "time.tlim=1e9"]))
    return jobs


def serve(njobs, rate_factor=8.0, seed=11, budget_gb=200.0, slots=64, chunk=8):
    """Sequential plain-FP64 runs (one after another, alone) against the continuous-batching
    server fed by seeded Poisson arrivals at `rate_factor` x the sequential completion rate,
    mixed precision with FP64 fallback. Returns runs/s of both and the server's p50/p99
    time-to-solution."""
    import numpy as np
    from chasqui_amd.deck import Deck
    from chasqui_amd.serve import Server
    jobs = serve_jobs(njobs, seed)
    # every fifth job's FP64 result is kept (density and energy; the copy is not timed) for the
    # accuracy of the mixed-precision runs
    sample = set(range(0, njobs, 5))
    ref = {}
    copy_s = 0.0
    t0 = time.perf_counter()
    for i, (text, ov) in enumerate(jobs):
        d = Deck(text, ov + ["mhd.precision=fp64"])
        m = d.make_mesh()
        d.close()
        m.run(chunk=chunk)
        if i in sample:
            t1 = time.perf_counter()
            ref[i] = active_rho_e(m)
            copy_s += time.perf_counter() - t1
        m.close()
    seq_s = time.perf_counter() - t0 - copy_s
    seq_rate = njobs / seq_s
    rng = np.random.default_rng(seed + 1)
    gaps = rng.exponential(1.0 / (rate_factor * seq_rate), size=njobs)
    t_arr = np.cumsum(gaps) - gaps[0]

    def play(precision):
        arrivals = [(float(t), text, ov, precision) for t, (text, ov) in zip(t_arr, jobs)]
        srv = Server(int(budget_gb * 2 ** 30), slots=slots, chunk=chunk)
        t0 = time.perf_counter()
        recs = srv.play(arrivals)
        span = time.perf_counter() - t0
        phases = srv.stats()
        srv.close()
        return recs, span, phases

    # the same arrivals served in FP64: the batching and serving gain without the precision gain
    recs64, span64, _ = play("fp64")
    recs, span, phases = play("mixed")
    tts = np.array([r["tts"] for r in recs])
    tts64 = np.array([r["tts"] for r in recs64])
    # accuracy of the served mixed-precision runs: a served run's result is bitwise the same run
    # stepped alone (tests/test_gpu_serve.py), so the sampled jobs are rerun alone in mixed
    # precision and compared with their FP64 results
    errs = []
    for i in sorted(ref):
        text, ov = jobs[i]
        d = Deck(text, ov + ["mhd.precision=mixed"])
        m = d.make_mesh()
        d.close()
        m.run(chunk=chunk)
        errs.append(rel_l1([active_rho_e(m)], [ref[i]]))
        m.close()
    acc = {"jobs": len(errs)}
    for q in ("rho", "E"):
        v = np.array([e[q] for e in errs])
        acc[q + "_median"] = float(np.median(v))
        acc[q + "_max"] = float(v.max())
    return {"jobs": njobs, "sequential_fp64_runs_per_s": seq_rate, "sequential_s": seq_s,
            "served_runs_per_s": njobs / span, "served_makespan_s": span,
            "speedup_vs_sequential_fp64": (njobs / span) / seq_rate,
            "served_fp64_runs_per_s": njobs / span64,
            "served_fp64_speedup_vs_sequential_fp64": (njobs / span64) / seq_rate,
            "served_fp64_tts_p50_s": float(np.percentile(tts64, 50)),
            "served_fp64_tts_p99_s": float(np.percentile(tts64, 99)),
            "arrival_rate_per_s": rate_factor * seq_rate,
            "tts_p50_s": float(np.percentile(tts, 50)), "tts_p99_s": float(np.percentile(tts, 99)),
            "l1_rel_vs_fp64": acc,
            "fell_back": int(sum(r["fell_back"] for r in recs)),
            "failed": int(sum(r["state"] != "done" for r in recs)),
            "budget_gb": budget_gb, "slots_per_batch": slots, "chunk_steps": chunk,
            "host_phases_s": {k: phases[k] for k in ("host_launch_s", "host_admit_s", "host_poll_wait_s",
                                                     "host_retire_s")}}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--n", type=int, default=256, help="cells per side (256 = the headline workload)")
    ap.add_argument("--cpu-steps", type=int, default=6, help="RK3 steps of the CPU baseline sample (~10 s)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-mixed", action="store_true", help="skip the mixed-precision side measurement")
    ap.add_argument("--ensemble-runs", type=int, default=512,
                    help="runs of the batched Orszag-Tang ensemble measured beside the headline (0: skip it and "
                         "the other side measurements: the batch floor and the one-slab shard)")
    ap.add_argument("--serve-jobs", type=int, default=100,
                    help="jobs of the served stream measured beside the headline (0: skip)")
    ap.add_argument("--serve-rate", type=float, default=8.0,
                    help="Poisson arrival rate of the served stream, in multiples of the sequential FP64 rate")
    ap.add_argument("--serve-chunk", type=int, default=8, help="steps per serving-loop iteration")
    ap.add_argument("--no-stream-probe", action="store_true",
                    help="skip the HBM streaming probe (chasqui_amd/_lib/chq_stream_probe) reported beside the roofline")
    ap.add_argument("--only", choices=["batch", "serve"], default=None,
                    help="print only the batched-ensemble or served-stream measurement")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (one mesh per device; "
              "sharding is a design note only). Run with --gpus 1.", file=sys.stderr)
        sys.exit(2)

    stream = None
    if not args.no_stream_probe and args.only is None and os.path.exists(STREAM_PROBE):
        # child process, before this process touches the GPU
        import subprocess
        r = subprocess.run([STREAM_PROBE], capture_output=True, text=True, timeout=120)
        if r.returncode == 0:
            stream = json.loads(r.stdout)
    if args.only == "batch":
        res = ensembles(args.ensemble_runs, 10)
        res["batch_floor"] = batch_floor()
        print(json.dumps(res))
        return
    if args.only == "serve":
        print(json.dumps(serve(args.serve_jobs, rate_factor=args.serve_rate, chunk=args.serve_chunk)))
        return
    n = args.n
    fstate = {} if not args.no_mixed else None
    ms, st, breakdown = run(n, "fp64", args.steps, args.warmup, breakdown=True, state=fstate)
    zone_updates = float(n) ** 3 * args.steps
    value = zone_updates / (ms * 1e-3)
    flux = breakdown["flux"]
    avg_flux_s = flux["total_ms"] * 1e-3 / max(flux["launches"], 1)
    flux_bytes = FLUX_DOUBLES_PER_CELL_STAGE / 3 * 8 * float(n) ** 3
    achieved = flux_bytes / avg_flux_s
    roofline = {"bound": "hbm", "achieved": achieved / 1e9, "peak": HBM_PEAK / 1e9, "unit": "GB/s",
                "frac": achieved / HBM_PEAK,
                "traffic": profile_traffic(TRAFFIC_PROFILE, FLUX_KERNELS) if n == 256 else None,
                "traffic_source": os.path.relpath(TRAFFIC_PROFILE, HERE),
                "kernel": "k_flux_x / k_flux_m (face flux, PPM + flattening + HLLD, one direction per launch)",
                "bytes_per_launch": flux_bytes, "avg_launch_ms": avg_flux_s * 1e3, "launches": flux["launches"]}
    model = value * 3 * MODEL_BYTES_PER_ZONE_STAGE
    out = {
        "metric": METRIC,
        "value": value,
        "unit": "zone-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded blast initial state from the library's problem generator)",
        "config": {"workload": f"mhd_blast_3d_{n}^3_ppm_hlld_rk3_ct_fofc_flatten_cfl0.3",
                   "cells": [n, n, n], "recon": "ppm", "rsolver": "hlld", "integrator": "rk3",
                   "fofc": True, "flatten": True, "cfl": 0.3, "precision": "fp64",
                   "fofc_corrections_in_run": st["fofc_total"], "cycles_run": st["cycle"]},
        "roofline": roofline,
        "step_roofline": {"model": "README.md:45-47, 1040 B/cell/RK stage", "achieved": model / 1e9,
                          "peak": HBM_PEAK / 1e9, "unit": "GB/s", "frac": model / HBM_PEAK},
        "kernels": breakdown,
    }
    if stream is not None:
        out["stream_probe"] = {k: v["gb_per_s"] for k, v in stream.items() if isinstance(v, dict)}
        out["stream_probe"]["unit"] = "GB/s"
    if not args.no_mixed:
        # the same workload with FP32 reconstruction/Riemann/EMF and FP64 con2prim/RK/CT
        mstate = {}
        mms, mst, mk = run(n, "mixed", args.steps, args.warmup, breakdown=True, state=mstate)
        mval = zone_updates / (mms * 1e-3)
        # relative L1 distance of the final mixed state from the FP64 one (same step count)
        import numpy as np
        a, b = mstate["u0"], fstate["u0"]
        l1 = {f: float(np.sum(np.abs(a[q] - b[q])) / np.sum(np.abs(b[q])))
              for f, q in (("rho", 0), ("E", 4))}
        l1["momentum"] = float(np.sum(np.abs(a[1:4] - b[1:4])) / np.sum(np.abs(b[1:4])))
        del mstate, fstate
        out["mixed"] = {"value": mval, "unit": "zone-updates/s", "dtype": "f32 flux/EMF, f64 state",
                        "ms_per_step": mms / args.steps, "speedup_vs_fp64": mval / value,
                        "flux_avg_launch_ms": mk["flux"]["total_ms"] / max(mk["flux"]["launches"], 1),
                        "fofc_corrections_in_run": mst["fofc_total"], "kernels": mk,
                        "l1_rel_vs_fp64": l1, "l1_cycles": mst["cycle"]}
    if args.ensemble_runs > 0:
        if n == 256:
            out["shard"] = shard(n)
        out["batch"] = ensembles(args.ensemble_runs, 10)
        out["batch_floor"] = batch_floor()
    if args.serve_jobs > 0:
        out["serve"] = serve(args.serve_jobs, rate_factor=args.serve_rate, chunk=args.serve_chunk)
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(n, args.cpu_steps)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
