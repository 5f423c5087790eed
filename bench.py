#!/usr/bin/env python3
"""Benchmark of the LES step on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload NAME]

A step is one pass of the hot path velnw -> bondv1 -> velfg -> feedbf ->
les -> adam -> press over the whole domain, with every field resident in
device memory (main.f95:260-296).  Default workload: the 300x300x90 plain
LES domain (dx=dy=4 m, dz 2 m x1.05 capped at 20 m, 30x30 city blocks of
8x8 cells with 2-cell streets, lognormal heights median 16 m sigma 0.5,
log-law inflow u*=0.5 m/s z0=0.1 m with +-1% noise, dt=0.05 s, Cs=0.1,
SOR omega=1, at most 50 iterations, stop at sum(reltmp^2) < 1e-4).

W untimed warmup steps (n = 1..W) are followed by K timed steps
(n = W+1..W+K) bracketed by device synchronisation.  During the timed steps
the dominant kernel (the SOR sweep) is bracketed by HIP events on the
domain's stream; its algorithmic bytes per launch divided by its average
duration gives the roofline figure.  A second short pass with every launch
timed gives the per-kernel table.  The CPU baseline runs the reference's
own Fortran subroutines (flang build under oracle/_ref) -- or, when that
build is absent, the C restatement -- on one host core for a bounded sample
of the same steps, starting from the GPU's state after the warmup.

Prints one JSON line.
"""
from __future__ import annotations

import argparse
import dataclasses
import json
import os
import re
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT / "mpi-les_amd"))

import numpy as np  # noqa: E402

from lesamd import api, setup_gen  # noqa: E402

HBM_PEAK_GBS = 8000.0
METRIC = "steps/s + Mcell-updates/s at 300x300x90 FP32, %HBM peak; served p50/p99"

# Algorithmic bytes per interior cell of each stage: every live reference
# array crossing the stage boundary counted once, FP32 (velnw 40, velfg+vel2
# 60, feedbf 64, les+boundsm 64, adam 48, press rhs 28, SOR 12 per
# iteration, pressure-mean removal 12).
STEP_BYTES_FIXED = 40 + 60 + 64 + 64 + 48 + 28 + 12
SOR_BYTES_PER_ITER = 12
SOR_RUN = 200  # back-to-back SOR launches timed for the dominant kernel's duration
# per launch of each kernel class, bytes per interior cell (sor: per iteration)
CLASS_BYTES = {
    "velnw": 40, "vel2": 48, "velfg": 12, "feedbf": 64, "les_sm": 40, "les_visc": 24, "adam": 48,
    "rhs": 28, "sor": SOR_BYTES_PER_ITER, "pav": 4, "padj": 8,
}
STATE_FIELDS = ("u", "v", "w", "p", "f", "g", "h", "usum", "vsum", "wsum", "fold", "gold", "hold", "rhs",
                "sm", "uspd", "vspd") + tuple(f"diu{q}" for q in range(1, 10))


def _time_steps(be, n0, nsteps, budget_s):
    """Wall time of each of up to nsteps steps of a CPU backend, stopping
    once budget_s is spent; the reference's progress prints are kept off
    stdout (the one-line JSON output)."""
    times = []
    sys.stdout.flush()
    saved = os.dup(1)
    devnull = os.open(os.devnull, os.O_WRONLY)
    os.dup2(devnull, 1)
    try:
        t_all = time.perf_counter()
        for n in range(n0, n0 + nsteps):
            t = time.perf_counter()
            be.step(n, 10 ** 9)
            times.append(time.perf_counter() - t)
            if time.perf_counter() - t_all > budget_s:
                break
        if hasattr(be, "flush"):
            be.flush()
    finally:
        os.dup2(saved, 1)
        os.close(saved)
        os.close(devnull)
    return times


def host_threads() -> int:
    """Threads for the multi-core baseline: OMP_NUM_THREADS when set (the
    process's share of a shared host), else the cores this process may run on."""
    env = os.environ.get("OMP_NUM_THREADS", "")
    if env.isdigit() and int(env) > 0:
        return int(env)
    return len(os.sched_getaffinity(0))


def cpu_baseline(setup, state, n0, nsteps, budget_s):
    """Time the CPU checker on this host for up to nsteps steps (stopping
    once budget_s is spent), from the GPU state `state` before step n0: the
    reference's own subroutines (or the C restatement) on one core, and the
    C restatement's OpenMP build on every thread of the process's share."""
    sys.path.insert(0, str(ROOT))
    from oracle import oracle
    w = setup.workload
    kind = "reference" if oracle.ref_available("large") and w.im <= 300 and w.jm <= 300 and w.km <= 105 else "port"
    if kind == "reference":
        ref_setup = dataclasses.replace(setup, inflow=None)
        be = oracle.RefBackend(ref_setup, "large")
        what = "reference Fortran subroutines (flang -O2, oracle/_ref/large) with its fixed 5 m/s inflow"
    else:
        be = oracle.OracleBackend(setup, 32, redmode=0)
        what = "C restatement (gcc -O2, FP32)"
    be.load(state)
    times = _time_steps(be, n0, nsteps, budget_s)
    del be
    sps = len(times) / sum(times)
    out = {
        "value": round(sps, 4),
        "unit": "steps/s",
        "cores": 1,
        "kind": kind,
        "sample": f"{len(times)} steps (n={n0}..{n0 + len(times) - 1}) of the {w.im}x{w.jm}x{w.km} domain on one "
                  f"host core, {what}, started from the GPU state after the warmup; "
                  f"{sum(times):.1f} s of CPU time",
        "mcell_updates_per_s": round(sps * w.cells / 1e6, 3),
    }
    if oracle.oracle_lib_path(omp=True).exists():
        nthr = host_threads()
        os.environ["OMP_NUM_THREADS"] = str(nthr)  # read by the OpenMP runtime when the library loads
        par = oracle.OracleBackend(setup, 32, redmode=1, omp=True)
        par.load(state)
        pt = _time_steps(par, n0, 4 * nsteps, budget_s)
        psps = len(pt) / sum(pt)
        out["all_cores"] = {
            "value": round(psps, 4), "unit": "steps/s", "cores": nthr, "kind": "port",
            "sample": f"{len(pt)} steps (n={n0}..{n0 + len(pt) - 1}) of the same domain and inflow as the GPU, "
                      f"C restatement (gcc -O2, FP32 fields, FP64 sums) with its k loops on {nthr} OpenMP "
                      f"threads; {sum(pt):.1f} s wall time",
            "mcell_updates_per_s": round(psps * w.cells / 1e6, 3),
        }
    return out


def profiled_traffic(kernel_class: str, workload: str, sor_rows: int = 4):
    """HBM bytes per launch of the dominant kernel, as measured by the
    rocprofv3 FETCH_SIZE and WRITE_SIZE counters (scaled by a copy of known
    size) and recorded in profiles/<kernel>_<workload>.md; None when no such
    measurement is recorded for this kernel and workload.  sor_rows: the SOR
    kernel in use (4 two iterations per launch, 2 or 1 one)."""
    if kernel_class == "sor":
        if os.environ.get("LES_SOR_FUSED", "1") == "0":
            name = "sor_sweep"
        else:
            name = {1: "sor_iter", 2: "sor_iter2"}.get(sor_rows, "sor_pair")
    else:
        name = kernel_class
    f = ROOT / "profiles" / f"{name}_{workload}.md"
    if not f.exists():
        return None
    m = re.search(r"^traffic_bytes_per_launch:\s*(\d+)", f.read_text(), flags=re.M)
    return int(m.group(1)) if m else None


def _tuning_progress(done: int, total: int):
    if done % 10 == 0 or done == total:
        print(f"[bench] tuned {done} of {total} grid shapes", file=sys.stderr, flush=True)


def _served_progress(done: int, total: int):
    if done % 16 == 0 or done == total:
        print(f"[bench] served {done} of {total} requests", file=sys.stderr, flush=True)


SERVE_BATCH_CELLS = 2 << 20  # domains above this many cells fill the GPU alone and are not batched


def serve_measurement(nreq: int, nsteps: int, sizes=None, rate: float | None = None, streams: bool = True) -> dict:
    """Mixed-size requests (nx, ny in 96..512, nz in 64..128, seeded urban
    set-ups with cover, heights and log-law inflow drawn per request; or the
    given sizes), arriving as a Poisson process of the given rate (requests
    per second; all at t=0 when None): served with continuous batching (the
    headline figures: domains of at most SERVE_BATCH_CELLS cells stepped as
    one batch, larger ones on their own streams alongside), optionally
    concurrently on streams (one stream and step graph per domain), and one
    after another, admission by device memory in every mode, domains built
    and released on a worker thread.  Every grid shape's SOR launch shape is
    tuned before the timed runs, so that no mode pays for it.  Reports the
    aggregate cell updates per second of the requests that completed, the
    time-to-result percentiles and the requests that failed (a diverged
    flow) in each mode."""
    from lesamd import serving
    reqs = serving.sample_requests(nreq, nsteps, rate=rate, seed=2024, sizes=sizes)
    srv = serving.Server(budget_bytes=200 << 30, burst=4)
    print(f"[bench] serving {nreq} requests x {nsteps} steps, "
          f"{'Poisson arrivals at %g/s' % rate if rate else 'all at t=0'}", file=sys.stderr, flush=True)
    srv.warm(reqs, progress=_tuning_progress)
    srv.on_result = _served_progress
    seq = serving.summarize(srv.run(reqs, sequential=True))
    con = serving.summarize(srv.run(reqs)) if streams else None
    bat = serving.summarize(srv.run_batched(reqs, max_batch_cells=SERVE_BATCH_CELLS))
    offered = sum(r.cells * r.nsteps for r in reqs) / max(reqs[-1].arrival, 1e-9) / 1e6 if rate else None
    out = {
        "requests": nreq, "steps_each": nsteps,
        "arrivals": f"Poisson, {rate:g} requests/s" if rate else "all at t=0",
        "offered_mcell_updates_per_s": round(offered, 1) if offered else None,
        "mode": f"continuous batching of the domains of at most {SERVE_BATCH_CELLS} cells (each stage one launch "
                "for all of them, SOR one launch per iteration), larger domains on their own streams",
        "mcell_updates_per_s": round(bat["mcell_updates_per_s"], 1),
        "p50_s": round(bat["p50_s"], 4), "p99_s": round(bat["p99_s"], 4), "failed": bat["failed"],
        "speedup_vs_sequential": round(bat["mcell_updates_per_s"] / seq["mcell_updates_per_s"], 3),
        "sequential_mcell_updates_per_s": round(seq["mcell_updates_per_s"], 1),
        "sequential_p50_s": round(seq["p50_s"], 4), "sequential_p99_s": round(seq["p99_s"], 4),
        "sequential_failed": seq["failed"],
        "fp64_oracle_error": "the same as one-at-a-time runs: served fields are bit-identical to the domain "
                             "stepped alone (tests/test_gpu_batch.py, tests/test_gpu_configs.py)",
    }
    if con is not None:
        out.update({"streams_mcell_updates_per_s": round(con["mcell_updates_per_s"], 1),
                    "streams_p50_s": round(con["p50_s"], 4), "streams_p99_s": round(con["p99_s"], 4)})
    return out


def poisson_sweep(rates: list[float], nreq: int, nsteps: int) -> int:
    """The mixed-size request stream of the serving measurement arriving as a
    Poisson process, swept over arrival rates (requests per second): at each
    rate the same seeded requests are served with continuous batching (as in
    the serving measurement) and one after another, after every grid shape
    has been tuned.  Reports per rate and mode the aggregate cell updates per
    second over the run (first arrival to last completion), the latency
    percentiles and the mean wait from arrival to admission; admission
    saturates where the wait and p99 climb while the throughput flattens."""
    from lesamd import serving
    srv = serving.Server(budget_bytes=200 << 30, burst=4)
    srv.on_result = _served_progress
    out = {"requests": nreq, "steps_each": nsteps, "batch_cells_max": SERVE_BATCH_CELLS, "sweep": []}
    for rate in rates:
        reqs = serving.sample_requests(nreq, nsteps, rate=rate, seed=2024)
        srv.warm(reqs, progress=_tuning_progress)
        rec = {"rate_per_s": rate, "offered_mcell_updates_per_s":
               round(sum(r.cells * r.nsteps for r in reqs) / max(reqs[-1].arrival, 1e-9) / 1e6, 1)}
        for tag, fn in (("batched", lambda: srv.run_batched(reqs, max_batch_cells=SERVE_BATCH_CELLS)),
                        ("sequential", lambda: srv.run(reqs, sequential=True))):
            print(f"[bench] poisson rate {rate}/s {tag}", file=sys.stderr, flush=True)
            res = fn()
            s = serving.summarize(res)
            rec[tag] = {"mcell_updates_per_s": round(s["mcell_updates_per_s"], 1),
                        "p50_s": round(s["p50_s"], 4), "p99_s": round(s["p99_s"], 4),
                        "mean_wait_s": round(float(np.mean([r.admitted - r.arrival for r in res])), 4)}
        out["sweep"].append(rec)
    print(json.dumps(out), flush=True)
    return 0


def defect_measure(setup, p: np.ndarray, rhs: np.ndarray) -> float:
    """sum((cn1*d)^2) over the interior in FP64, d = rhs - B(p) with B the
    pressure operator of the SOR (press.f95:95-98 at its fixed point, the
    ghosts of p as press leaves them), cn1 = 1/((ax+by)+cz) in FP32 as the
    kernels evaluate it: the measure press mode 1 stops on, evaluated on any
    result p (a constant shift of p does not change it)."""
    w = setup.workload
    im, jm, km = w.im, w.jm, w.km
    c = setup.coeffs
    ax, by, cz = (np.asarray(t, np.float32) for t in c["_cn1_terms"])
    s32 = (ax[:, None, None] + by[None, :, None]) + cz[None, None, :]
    cn = (np.float32(1.0) / s32).astype(np.float64)
    P = np.asarray(p, np.float64)
    co = {n: np.asarray(c[n], np.float64) for n in ("cn2l", "cn2s", "cn3l", "cn3s", "cn4l", "cn4s")}
    I, J, K = slice(1, im + 1), slice(1, jm + 1), slice(1, km + 1)
    b = (co["cn2l"][:im, None, None] * P[2:im + 2, J, K] + co["cn2s"][:im, None, None] * P[0:im, J, K]
         + co["cn3l"][None, :jm, None] * P[I, 2:jm + 2, K] + co["cn3s"][None, :jm, None] * P[I, 0:jm, K]
         + co["cn4l"][None, None, :km] * P[I, J, 2:km + 2] + co["cn4s"][None, None, :km] * P[I, J, 0:km]
         - P[I, J, K] / cn)
    rd = cn * (np.asarray(rhs, np.float64)[I, J, K] - b)
    return float(np.sum(rd * rd))


def precision_comparison(K: int, W: int, configs) -> int:
    """The precision-512x512x128 workload (dx=dy=2 m, 35 % cover, u*=0.6,
    z0=0.3, dt=0.02; SOR at most 100 iterations, stop at 1e-6) stepped with
    the reference's FP32 SOR and with the FP64 defect correction for each
    (outer, inner, omega) in configs (at most `outer` steps of `inner` FP32
    SOR iterations with relaxation omega, stopping once sum((cn1*d)^2) <
    1e-6): time per step from graph replay, and per step of the timed run
    the measure the stop test compares with 1e-6 and the sweeps taken.  The
    reference's stop measure is sum(reltmp^2), the size of the last
    iteration's update; the defect correction's is sum((cn1*d)^2) of its
    iterate; result_defect_measure evaluates the latter on every press's
    final p, so the presses compare on one measure."""
    w = setup_gen.workloads()["precision-512x512x128"]
    out = {"workload": w.name, "grid": [w.im, w.jm, w.km], "steps": K, "warmup": W, "target": w.pjuge}
    # the reference's SOR as shipped (100 iterations), and with 1024 to show
    # where its FP32 iterate stops improving
    runs = [("fp32_sor", None), ("fp32_sor_1024", None)]
    runs += [(f"fp64_dc_{o}x{i}_omega_{om:g}", (o, i, om)) for o, i, om in configs]
    for tag, cfg in runs:
        print(f"[bench] precision: {tag}", file=sys.stderr, flush=True)
        wt = dataclasses.replace(w, nmaxp=1024) if tag == "fp32_sor_1024" else w
        setup = setup_gen.build(wt, dense=False)
        dom = api.Domain(setup)
        om = None
        if cfg is not None:
            outer, inner, om = cfg
            dom.set_press_mode(1, outer=outer, inner=inner)
            dom.set_press_omega(om)
        nmax = W + K + 1
        for n in range(1, W + 1):
            dom.step(n, nmax)
        dom.set_graph(True)
        dom.step(W + 1, nmax)
        dom.synchronize()
        t0 = time.perf_counter()
        for n in range(W + 2, W + K + 2):
            dom.step(n, nmax)
        dom.synchronize()
        dt_ = (time.perf_counter() - t0) / K
        # one more step with its statistics read back (graph replays as timed)
        sweeps, finals = [], []
        for n in range(W + K + 2, W + K + 4):
            dom.step(n, nmax)
            st = dom.press_stats()
            h = st["sor"]
            if om is None:
                sweeps.append(st["iters"])
            else:
                conv = len(h) and h[-1] < w.pjuge
                sweeps.append(inner * (len(h) - 1 if conv else len(h)))
            finals.append(float(h[-1]))
        rec = {"ms_per_step": round(1e3 * dt_, 3), "mcell_updates_per_s": round(w.cells / dt_ / 1e6, 1),
               "sweeps_per_step": sweeps, "final_residual_measure": finals,
               "reached_target": [f < w.pjuge for f in finals],
               # the same measure for every press: sum((cn1*d)^2) of the last step's result p
               "result_defect_measure": float(f"{defect_measure(setup, dom.download('p'), dom.download('rhs')):.4e}")}
        if om is not None:
            rec["defect_history_last_step"] = [float(f"{x:.3e}") for x in h]
        elif len(h) > 100:
            rec["residual_measure_at_iteration"] = {str(q): float(f"{h[q - 1]:.4e}")
                                                    for q in (100, 200, 400, 800, len(h)) if q <= len(h)}
            rec["smallest_residual_measure"] = float(f"{float(np.min(h)):.4e}")
        out[tag] = rec
        dom.close()
    print(json.dumps(out), flush=True)
    return 0


ACCURACY_SERIAL_CELLS = 10_000_000  # above this the FP64 oracle runs on all host cores


def accuracy_table(wname: str, nsteps: int, dc: tuple | None, dc_omega: float = 1.0) -> int:
    """Per-field error against the FP64 oracle: the workload stepped on the
    GPU (FP32 stencils, FP64 sums), by the C restatement as the reference
    ships it (FP32 everywhere, sequential FP32 sums; equal to the
    reference's own build bit for bit) and by the restatement's FP64
    instantiation (pinned bit for bit to the reference with every real
    promoted to kind 8), from the same seeded set-up.  Reports per step the
    relative L2 error of u, v, w, p, sm against FP64 for the GPU, for the
    FP32 reference and, with dc = (outer, inner), for the GPU's FP64
    defect-correction press.  Above ACCURACY_SERIAL_CELLS cells the FP64
    oracle runs on the OpenMP build of the restatement (its stencils are
    bit for bit the serial build's, its FP64 sums rounded in another order)
    and the FP32 reference row, whose sequential FP32 sums cannot be split
    over threads, is left out."""
    sys.path.insert(0, str(ROOT))
    from oracle import oracle
    w = dataclasses.replace(setup_gen.workloads()[wname], nsteps=nsteps)
    big = w.im * w.jm * w.km > ACCURACY_SERIAL_CELLS
    if big:
        os.environ["OMP_NUM_THREADS"] = str(host_threads())  # read when the library loads
    s = setup_gen.build(w)
    fields = ("u", "v", "w", "p", "sm")

    def rel(a, b):
        a = np.asarray(a, np.float64)
        b = np.asarray(b, np.float64)
        return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))

    runs = {"gpu": api.Domain(s)}
    if dc:
        runs["gpu_defect_correction"] = api.Domain(s)
        runs["gpu_defect_correction"].set_press_mode(1, outer=dc[0], inner=dc[1])
        runs["gpu_defect_correction"].set_press_omega(dc_omega)
    o64 = oracle.OracleBackend(s, 64, redmode=0, omp=big)
    r32 = None if big else oracle.OracleBackend(s, 32, redmode=0)
    table = {k: [] for k in list(runs) + ([] if big else ["reference_fp32"])}
    print(f"[bench] accuracy: {w.name}, {nsteps} steps", file=sys.stderr, flush=True)
    for n in range(1, nsteps + 1):
        for d in runs.values():
            d.step(n, nsteps)
        o64.step(n, nsteps)
        ref = o64.fetch(list(fields))
        for k, d in runs.items():
            table[k].append({f: rel(d.download(f), ref[f]) for f in fields})
        if r32 is not None:
            r32.step(n, nsteps)
            got32 = r32.fetch(list(fields))
            table["reference_fp32"].append({f: rel(got32[f], ref[f]) for f in fields})
        print(f"[bench] accuracy step {n}: gpu p {table['gpu'][-1]['p']:.2e}", file=sys.stderr, flush=True)
    for d in runs.values():
        d.close()
    out = {"accuracy": {"workload": w.name, "grid": [w.im, w.jm, w.km], "steps": nsteps,
                        "oracle": "FP64 restatement (orc64, bit for bit the kind-promoted reference build)"
                                  + (f", OpenMP build on {host_threads()} threads" if big else ""),
                        "metric": "relative L2 error per field after each step",
                        "defect_correction": {"outer": dc[0], "inner": dc[1], "omega": dc_omega} if dc else None,
                        "rel_l2": {k: [{f: float(f"{x:.3e}") for f, x in row.items()} for row in v]
                                   for k, v in table.items()}}}
    print(json.dumps(out), flush=True)
    return 0


def copy_calibration(mib: int) -> int:
    """Ten device-to-device copies of `mib` MiB through the HIP runtime: a
    kernel of known traffic (read mib, write mib) for scaling the HBM byte
    counters."""
    import ctypes
    hip = ctypes.CDLL("libamdhip64.so")
    n = mib << 20
    a, b = ctypes.c_void_p(), ctypes.c_void_p()
    assert hip.hipMalloc(ctypes.byref(a), ctypes.c_size_t(n)) == 0
    assert hip.hipMalloc(ctypes.byref(b), ctypes.c_size_t(n)) == 0
    hip.hipMemset(a, 1, ctypes.c_size_t(n))
    for _ in range(10):
        assert hip.hipMemcpy(b, a, ctypes.c_size_t(n), 3) == 0  # device to device
    hip.hipDeviceSynchronize()
    hip.hipFree(a)
    hip.hipFree(b)
    print(json.dumps({"copy_calibration_bytes": n, "copies": 10}))
    return 0


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--workload", default="plain-300x300x90", choices=sorted(setup_gen.workloads()))
    ap.add_argument("--cpu-steps", type=int, default=8, help="max CPU-baseline steps")
    ap.add_argument("--cpu-budget", type=float, default=20.0, help="CPU-baseline time budget [s]")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--eager", action="store_true", help="launch kernels one by one (no graph replay)")
    ap.add_argument("--no-tune", action="store_true", help="keep the default SOR launch shape (no autotuning)")
    ap.add_argument("--serve-rate", type=float, default=4.0,
                    help="Poisson arrival rate of the served mixed stream, requests per second (0: all at t=0); "
                         "the default is past the rate where admission saturates")
    ap.add_argument("--serve-requests", type=int, default=64,
                    help="mixed-size requests of the serving measurement (0: skip it)")
    ap.add_argument("--serve-steps", type=int, default=200, help="time steps per served request")
    ap.add_argument("--precision", action="store_true",
                    help="only compare the reference SOR with the FP64 defect-correction press on the "
                         "precision-512x512x128 workload (prints its own JSON line)")
    ap.add_argument("--accuracy", type=int, default=0, metavar="STEPS",
                    help="only report the per-field error against the FP64 oracle over this many steps of the "
                         "workload (GPU, FP32 reference, and with --dc the defect-correction press); prints its "
                         "own JSON line")
    ap.add_argument("--dc", default="", metavar="OUTER,INNER",
                    help="with --accuracy: also the FP64 defect-correction press with these outer and inner steps")
    ap.add_argument("--dc-omegas", default="", metavar="OMEGA",
                    help="with --accuracy: the inner relaxation of that press (default 1)")
    ap.add_argument("--dc-configs", default="4x25@1,4x25@1.9,64x32@1.9", metavar="LIST",
                    help="with --precision: defect-correction runs OUTERxINNER@OMEGA, comma-separated (the "
                         "default: the reference's 100 sweeps as 4 x 25 at omega 1 and 1.9, and up to 2048 "
                         "sweeps at 1.9)")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run device-to-device copies of this size (HBM counter calibration)")
    ap.add_argument("--serve-poisson", default="", metavar="RATES",
                    help="only serve the mixed request stream with Poisson arrivals at these comma-separated "
                         "rates (requests per second), batched and sequential (prints its own JSON line)")
    args = ap.parse_args(argv)
    if args.gpus != 1:
        print("bench.py: this build runs on exactly one GPU (one domain per device); "
              f"--gpus {args.gpus} is not supported", file=sys.stderr)
        return 2
    if int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py: multi-rank launch is not supported; run with one GPU", file=sys.stderr)
        return 2
    if args.serve_poisson:
        return poisson_sweep([float(x) for x in args.serve_poisson.split(",")], args.serve_requests,
                             args.serve_steps)
    if args.copy_calibration:
        return copy_calibration(args.copy_calibration)
    if args.accuracy:
        dc = tuple(int(x) for x in args.dc.split(",")) if args.dc else None
        om = float(args.dc_omegas.split(",")[0]) if args.dc_omegas else 1.0
        return accuracy_table(args.workload, args.accuracy, dc, om)
    if args.precision:
        cfgs = []
        for c in args.dc_configs.split(","):
            oi, om = c.split("@")
            o, i = oi.split("x")
            cfgs.append((int(o), int(i), float(om)))
        return precision_comparison(args.steps if args.steps != 100 else 5, args.warmup if args.warmup != 10 else 2,
                                    cfgs)

    w = setup_gen.workloads()[args.workload]
    # the CPU baseline needs the host set-up arrays (masks, cn1); without it
    # the light set-up (masks built on the device, cn1 recomputed) suffices
    setup = setup_gen.build(w, dense=not args.no_cpu_baseline)
    dom = api.Domain(setup)
    W, K = args.warmup, args.steps
    nmax = W + 2 * K + 4
    print(f"[bench] {w.name}: {w.cells} cells, {dom.nbytes() / 2**30:.2f} GiB on device; "
          f"warmup {W}, timed {K}", file=sys.stderr, flush=True)
    for n in range(1, W + 1):
        dom.step(n, nmax)
    dom.synchronize()
    # the SOR launch shape and the stencil kernels' k levels per thread for
    # this grid shape (les_domain_autotune), picked before the timed steps as
    # a served domain gets them at admission
    tuned = None if args.no_tune else dom.autotune(reps=10)
    snapshot = None if args.no_cpu_baseline else {n: dom.download(n) for n in STATE_FIELDS}

    # timed region: steps replayed from the captured graph of the step
    dom.set_graph(not args.eager)
    dom.step(W + 1, nmax)  # capture (not timed)
    dom.synchronize()
    t0 = time.perf_counter()
    for n in range(W + 2, W + K + 2):
        dom.step(n, nmax)
    dom.synchronize()
    t1 = time.perf_counter()
    st = dom.press_stats()
    # the dominant kernel (the SOR iterations, class "sor"; iteration 1 on
    # entry ghosts is class "sor1") bracketed by HIP events on the domain's stream over K
    # more steps launched eagerly (events cannot bracket kernels inside a
    # graph), one event pair per launch: the pair's marker packets add a
    # few us to each launch.  The kernel's own duration is then timed with
    # one pair around a run of back-to-back launches (below).
    dom.set_timing(True, ["sor", "sor1"])
    dom.timing()
    iters_run = []  # SOR iterations each of these steps executed (read from the device)
    for n in range(W + K + 2, W + 2 * K + 2):
        dom.step(n, nmax)
        iters_run.append(dom.press_stats()["iters"])
    kt = dom.timing()
    # short second pass with every launch timed: the per-kernel table
    dom.set_timing(True)
    dom.timing()
    nprof = 2
    for n in range(W + 2 * K + 2, W + 2 * K + 2 + nprof):
        dom.step(n, nmax)
    kt_all = dom.timing()
    dom.set_timing(False)
    # a non-finite state makes every later SOR launch exit at once (its stop
    # flag is set), which would time empty launches: refuse to report them
    if st["nonfinite"] or dom.press_stats()["nonfinite"]:
        print("bench.py: non-finite SOR residual (the flow diverged); no figures reported", file=sys.stderr)
        return 1
    elapsed = t1 - t0
    sps = K / elapsed
    L = st["iters"]
    # dominant kernel = the class with the largest total time
    dom_cls = max((c for c in kt if kt[c][1] > 0), key=lambda c: kt[c][0])
    dom_ms, dom_n = kt[dom_cls]
    avg_s = dom_ms / 1e3 / dom_n
    alg_bytes = CLASS_BYTES.get(dom_cls, 0) * w.cells
    fused = os.environ.get("LES_SOR_FUSED", "1") != "0"
    # SOR iterations per launch: 2 for the two-iteration kernel (the tuned or
    # default shape reports rows 4), 1 for one fused iteration, 1/2 for the
    # colour sweeps (LES_SOR_FUSED=0)
    sor_rows = (tuned or {}).get("rows", 4)  # untuned: the default shape (two iterations per launch)
    ipl = 0.5 if not fused else (2 if sor_rows == 4 else 1)
    if dom_cls == "sor":
        # Class "sor" holds the SOR launches after the one holding iteration
        # 1 (class "sor1"), or all colour sweeps (LES_SOR_FUSED=0).
        # Launches after a step's stop test exit at once: the iterations
        # that ran come from each step's own count, and the launches' total
        # time is divided among the launches that carried them.
        first = {0.5: 0, 1: 1, 2: 2}[ipl]
        iters = sum(max(0, it - first) for it in iters_run)
        alg_bytes = SOR_BYTES_PER_ITER * w.cells * ipl
        avg_s = dom_ms / 1e3 / max(1, round(iters / ipl))
    in_step_us = avg_s * 1e6
    timing_method = "one HIP event pair per launch in K eager steps"
    if dom_cls == "sor" and fused and round(iters / ipl) == dom_n:
        # every launch carried its iterations: the kernel's duration from
        # SOR_RUN back-to-back iterations of the same kernel and launch shape
        # on the domain's stream between one event pair (les_domain_time_sor
        # reports per iteration)
        avg_s = dom.time_sor(SOR_RUN) * ipl / 1e6
        timing_method = (f"one HIP event pair around {SOR_RUN} back-to-back iterations "
                         f"({int(SOR_RUN / ipl)} launches) on the domain's stream")
    achieved = alg_bytes / avg_s / 1e9
    step_bytes = w.cells * (STEP_BYTES_FIXED + SOR_BYTES_PER_ITER * L)
    per_kernel = {c: {"ms_per_step": round(v[0] / nprof, 4), "launches_per_step": v[1] / nprof,
                      "us_per_launch": round(1e3 * v[0] / v[1], 2)} for c, v in kt_all.items() if v[1]}

    result = {
        "metric": METRIC,
        "value": round(sps, 3),
        "unit": "steps/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": W,
        "ms_per_step": round(1e3 * elapsed / K, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded city, log-law inflow)",
        "config": {"workload": w.name, "grid": [w.im, w.jm, w.km], "dx": w.dx, "dt": w.dt,
                   "nmaxp": w.nmaxp, "pjuge": w.pjuge, "seed": w.seed},
        "mcell_updates_per_s": round(sps * w.cells / 1e6, 2),
        "sor_iters_last_step": L,
        "hbm_frac_step": round(step_bytes * sps / 1e9 / HBM_PEAK_GBS, 4),
        "roofline": {
            "kernel": dom_cls,
            "bound": "hbm",
            "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": profiled_traffic(dom_cls, w.name, sor_rows),
            "alg_bytes_per_launch": int(alg_bytes),
            "sor_iterations_per_launch": ipl if dom_cls == "sor" else None,
            "avg_launch_us": round(avg_s * 1e6, 3),
            "timing": timing_method,
            "in_step_us_per_launch": round(in_step_us, 3),
        },
        "kernels": per_kernel,
    }
    if tuned is not None:
        result["sor_launch"] = {"kernel": {4: "k_sor_pair (two iterations per launch)",
                                           2: "k_sor_iter2", 1: "k_sor_iter"}[tuned["rows"]],
                                "rows_per_thread": min(tuned["rows"], 2), "tile_rows": tuned["tile_rows"],
                                "tile_pairs": tuned["tile_width"], "k_levels_per_block": tuned["kchunk"],
                                "tuned_us_per_iter": round(tuned["us_per_iter"], 2),
                                "default_us_per_iter": round(tuned["us_per_iter_default"], 2)}
        result["stencil_launch"] = {"k_levels_per_thread": {"velfg_feedbf": tuned["velfg_kchunk"],
                                                            "les_sm": tuned["sm_kchunk"],
                                                            "les_visc_adam": tuned["visc_kchunk"]},
                                    "note": "0: the automatic chunk (about 36 blocks per compute unit)",
                                    "tuned_us": round(tuned["stencil_us"], 2),
                                    "default_us": round(tuned["stencil_us_default"], 2)}
    dom.close()
    if args.serve_requests > 0:
        result["served"] = serve_measurement(args.serve_requests, args.serve_steps,
                                             rate=args.serve_rate if args.serve_rate > 0 else None, streams=False)
        # small domains only (nx, ny in 96..160, nz 64), where one domain
        # cannot fill the GPU: 24 requests of 40 steps arriving at once
        small = [(a, b, 64) for a in (96, 128, 160) for b in (96, 128, 160)]
        result["served_small"] = serve_measurement(24, 40, sizes=small)
    if snapshot is not None:
        result["cpu_baseline"] = cpu_baseline(setup, snapshot, W + 1, args.cpu_steps, args.cpu_budget)
    print(json.dumps(result), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
