"""Benchmark: explicit bond-based peridynamics steps on one MI355X.

Workload: the mixed-mode three-point-bend beam 500x100x50 mm at dx = 1.25 mm (400x40x80 =
1,280,000 nodes, 74.5M bonds, offset notch, exponential softening), ADR with the penetrator
staircase of 5e-6 m per 1000 iterations over 20,000 iterations (elastic loading, peak and
softening); one step = stretch -> law -> force -> nodal gather -> ADR update -> contact.  Both forms
run the whole history: FP32-mixed (headline `value`) and plain FP64.

Reported (one JSON line on stdout):
  value            bond updates/s of the FP32-mixed form = nBonds * K / seconds (inputs resident)
  fp64             the same for the plain FP64 form
  precision        FP32-mixed against FP64 on the load history: peak of the increment-averaged load
                   curve, the FP64 run's own rounding sensitivity beside it, and per-result errors
                   (load, nodal force, flags, broken bonds, damage field) from identical FP64 states
  roofline         dominant kernel (bond kernel, mixed form): the bytes of a node-centric CSR layout,
                   E*(4 + w_state + w_kappa + n_par*w) + N*180 with E = 2*nBonds entries, divided by
                   the kernel's average launch time from HIP events on the solver's stream, against
                   8 TB/s (a fraction above 1 means the stencil layout moves fewer bytes than that
                   model); `frac_layout_bytes` is the same with the stencil layout's own bytes per
                   launch; `traffic` is the kernel's measured HBM bytes per launch (rocprofv3
                   FETCH_SIZE + WRITE_SIZE, each scaled by the same counter's reading for a device
                   copy of known size) and `traffic_gbs` that per launch time; `copy_gbs` is this
                   GPU's device-to-device copy rate (bytes read plus written), the on-box peak beside
                   the published one
  cpu_baseline     the oracle's OpenMP restatement of the reference CPU passes timed on the host on
                   a bounded number of steps of the same workload
"""
from __future__ import annotations

import argparse
import csv
import glob
import json
import os
import shutil
import subprocess
import sys
import tempfile
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def csr_model_bytes(case, mixed: bool) -> int:
    """Bytes one bond-kernel launch must move in the node-centric layout with kappa stored per
    entry: E*(4 + w_state + w_kappa + n_par*w) + N*180."""
    w = 4 if mixed else 8
    n_par = int(isinstance(case.s0, np.ndarray)) + int(isinstance(case.sc, np.ndarray))
    E = 2 * case.nBonds
    return E * (4 + w + w + n_par * w) + case.nNodes * 180


def run_form(case, precision, steps, warmup, chunk=None, keep=(), retune=False):
    """Build the solver, run `warmup` untimed steps, then exactly `steps` timed steps in chunks of
    `chunk` (each bracketed by synchronisations; the times are summed).  Between chunks, untimed,
    the state at the step counts in `keep` is checkpointed to host memory.  With `retune` (ADR on the
    stencil layout) the autotuner reruns at the start of every chunk, inside the timed window, so
    the launch variant follows the load history (XCD-contiguous block order keeps neighbour rows in
    one L2; launch order spreads a localised process zone over the XCDs).  Returns the solver (open)
    with the timing, the whole per-step history from step 1 and the checkpoints."""
    import torch
    from bb_pd_amd import _lib
    from bb_pd_amd.solver import Solver
    t0 = time.time()
    sol = Solver(case, precision)
    log(f"  setup {time.time() - t0:.1f}s, info {sol.info()}")
    chunk = chunk or steps
    keep = set(keep)
    ckpts = {}
    rows = []
    sol.step(warmup)
    sol.sync()
    if warmup in keep:
        ckpts[warmup] = sol.checkpoint()
    tune = None
    tunable = case.integrator == _lib.ADR and sol.layout().startswith("stencil")
    if tunable:
        tune = sol.autotune(5)
        log(f"  autotune: variant {tune[0]} of {['%.3f' % x for x in tune[1]]} ms")
    picks = []
    sol.set_timing(True)
    dt, done = 0.0, 0
    while done < steps:
        # chunks end on multiples of `chunk` counted from step 0 (load-increment boundaries)
        at = warmup + done
        n = min((at // chunk + 1) * chunk - at, steps - done)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        if retune and tunable and done > 0:
            picks.append(sol.autotune(2)[0])
        sol.step(n)
        sol.sync()
        torch.cuda.synchronize()
        dt += time.perf_counter() - t0
        done += n
        rows.append(sol.history())
        if len(rows) % 10 == 0:
            log(f"    {warmup + done} steps")
        if warmup + done in keep:
            ckpts[warmup + done] = sol.checkpoint()
    avg_ms, launches = sol.kernel_time()
    sol.set_timing(False)
    info = sol.info()
    hist = {k: np.concatenate([r[k] for r in rows]) for k in rows[0]}
    mixed = precision == _lib.MIXED
    model = csr_model_bytes(case, mixed)
    achieved = model / (avg_ms * 1e-3) / 1e9 if avg_ms > 0 else 0.0
    return dict(solver=sol, seconds=dt, ms_per_step=dt * 1e3 / steps, bond_updates_per_s=case.nBonds * steps / dt,
                kernel_ms=avg_ms, launches=launches, bytes_model=model, bytes_layout=info["bond_kernel_bytes"],
                achieved_gbs=achieved, frac=achieved / HBM_PEAK_GBS, info=info, hist=hist, ckpts=ckpts,
                layout=sol.layout(), autotune=tune, retune_picks=picks)


def copy_bandwidth(gib: float = 2.0, reps: int = 10) -> float:
    """This GPU's device-to-device copy rate (GB/s, bytes read plus bytes written) over a buffer of
    `gib` GiB: the on-box reference beside the published 8 TB/s peak."""
    import torch
    n = int(gib * (1 << 30)) // 8
    a = torch.ones(n, dtype=torch.float64, device="cuda")
    b = torch.empty_like(a)
    b.copy_(a)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        b.copy_(a)
    e1.record()
    e1.synchronize()
    ms = e0.elapsed_time(e1) / reps
    del a, b
    torch.cuda.empty_cache()
    return 2 * n * 8 / (ms * 1e-3) / 1e9


def increment_curve(R: np.ndarray, ipi: int) -> np.ndarray:
    """Load per increment: the static reaction averaged over the last half of each increment's
    iterations (row 0 = step 1)."""
    n = len(R) // ipi
    h = ipi // 2
    return np.array([R[i * ipi + ipi - h:(i + 1) * ipi].mean() for i in range(n)])


def one_step_from(sol, st):
    """Restore a checkpoint, run one step, return (nodal force, static reaction, H, flag, fail,
    damage field, damage energy); the damage observables are computed on the device."""
    sol.history()
    sol.restore(st["step"], **{k: v for k, v in st.items() if k != "step"})
    sol.step(1)
    sol.sync()
    R = float(sol.history()["reaction_static"][-1])
    F = sol.nodal()[2]
    H, flag, fail = sol.bonds()
    D, W = sol.damage()
    return F, R, H, flag, fail, D, W


def precision_study(case, fp64, mixed, ipi):
    """FP32-mixed against FP64 on the load history (the load is the penetrator family's static
    reaction, calculatereactionforceFast.m:3-5).

    * free-running: peak of the increment-averaged load curve of each form;
    * rounding floors: the FP64 run restarted after increment 1 with one coordinate moved by one ulp,
      and with every coordinate moved by 2^-24 relative (the size of FP32 rounding), each run to the
      end; their peak differences are what such perturbations alone do to this load history;
    * same state: from FP64 states along the history, incl. the FP64 peak increment, one step of each
      form; errors of the load, nodal forces, flags, broken bonds and damage field."""
    A, B = fp64["solver"], mixed["solver"]
    total = len(fp64["hist"]["reaction_static"])
    c64 = increment_curve(fp64["hist"]["reaction_static"], ipi)
    cm = increment_curve(mixed["hist"]["reaction_static"], ipi)
    p64, pm = int(np.argmax(c64)), int(np.argmax(cm))
    # same-state samples: the checkpoints plus the state at the end of the FP64 peak increment
    ck = dict(fp64["ckpts"])
    peak_step = (p64 + 1) * ipi
    if peak_step not in ck:
        base = max(k for k in ck if k <= peak_step)
        A.history()
        A.restore(base, **{k: v for k, v in ck[base].items() if k != "step"})
        A.step(peak_step - base)
        A.sync()
        ck[peak_step] = A.checkpoint()
    same = []
    for k in sorted(ck):
        st = ck[k]
        Fa, Ra, Ha, fla, faa, Da, Wa = one_step_from(A, st)
        Fb, Rb, Hb, flb, fab, Db, Wb = one_step_from(B, st)
        same.append(dict(step=k, load=Ra, load_err=(Rb - Ra) / Ra if Ra else None,
                         nodal_force_err=float(np.abs(Fb - Fa).max() / np.abs(Fa).max()),
                         flag_flips=int((fla != flb).sum()),
                         broken_fp64=int(((Ha == 1) | (faa == 0)).sum()),
                         broken_mixed=int(((Hb == 1) | (fab == 0)).sum()),
                         damage_err=float(np.abs(Db - Da).max()),
                         damage_energy=Wa, damage_energy_err=(Wb - Wa) / Wa if Wa else None))
        log(f"  same state at step {k}: {same[-1]}")
    # rounding floor: restart FP64 after increment 1 with one ulp on the largest deflection
    st = {k: (v.copy() if isinstance(v, np.ndarray) else v) for k, v in ck[ipi].items()}
    i = int(np.argmax(np.abs(st["u"][:, 2])))
    st["x"][i, 2] = np.nextafter(st["x"][i, 2], np.inf)
    A.history()
    A.restore(st["step"], **{k: v for k, v in st.items() if k != "step"})
    left = total - st["step"]
    while left > 0:
        n = min(10 * ipi, left)
        A.step(n)
        A.sync()
        left -= n
        log(f"    rounding-floor run: {total - left} steps")
    R2 = np.concatenate([fp64["hist"]["reaction_static"][:st["step"]], A.history()["reaction_static"]])
    c2 = increment_curve(R2, ipi)
    p2 = int(np.argmax(c2))
    # perturbation of FP32 rounding size: every coordinate moved by 2^-24 relative, random sign
    st = {k: (v.copy() if isinstance(v, np.ndarray) else v) for k, v in ck[ipi].items()}
    sgn = np.where(np.random.default_rng(7).random(st["x"].shape) < 0.5, -1.0, 1.0)
    st["x"] = np.asfortranarray(st["x"] * (1 + sgn * 2.0 ** -24))
    st["u"] = np.asfortranarray(st["x"] - case.X)
    A.history()
    A.restore(st["step"], **{k: v for k, v in st.items() if k != "step"})
    left = total - st["step"]
    while left > 0:
        n = min(10 * ipi, left)
        A.step(n)
        A.sync()
        left -= n
        log(f"    FP32-sized perturbation run: {total - left} steps")
    R3 = np.concatenate([fp64["hist"]["reaction_static"][:st["step"]], A.history()["reaction_static"]])
    c3 = increment_curve(R3, ipi)
    p3 = int(np.argmax(c3))
    peak = c64[p64]
    at_peak = next(x for x in same if x["step"] == peak_step)
    return dict(
        load="static reaction of the penetrator family, mean over the last half of each load increment",
        increments=len(c64), iterations_per_increment=ipi,
        curve_fp64=[round(float(v), 3) for v in c64], curve_mixed=[round(float(v), 3) for v in cm],
        peak_fp64=float(peak), peak_increment_fp64=p64 + 1, peak_mixed=float(cm[pm]), peak_increment_mixed=pm + 1,
        peak_load_err=float((cm[pm] - peak) / peak),
        fp64_rounding_floor=float((c2[p2] - peak) / peak),
        fp64_fp32_sized_floor=float((c3[p3] - peak) / peak),
        curve_err_max=float(np.abs(cm - c64).max() / peak),
        curve_floor_max=float(np.abs(c2 - c64).max() / peak),
        curve_fp32_sized_floor_max=float(np.abs(c3 - c64).max() / peak),
        curve_fp64_ulp=[round(float(v), 3) for v in c2], curve_fp64_fp32_sized=[round(float(v), 3) for v in c3],
        same_state_peak_load_err=at_peak["load_err"],
        same_state=same)


def cpu_baseline(case, target_s=20.0):
    """The oracle's OpenMP passes (reference layout, nBonds intermediates) on the same beam."""
    from oracle import oracle as O
    threads = min(16, os.cpu_count() or 1)
    run = O.OracleRun(case, nodal_threads=threads)
    run.step_omp(1, threads)                     # first touch
    n, t = 0, 0.0
    t0 = time.perf_counter()
    while t < target_s and n < 50:
        run.step_omp(1, threads)
        n += 1
        t = time.perf_counter() - t0
    return dict(value=case.nBonds * n / t, unit="bond updates/s", cores=threads, kind="port",
                sample=f"{n} steps of the same {case.nBonds}-bond beam (passes A-E in the reference's "
                       f"Euler-Cromer order, exponential law, elastic state), oracle/bbpd_oracle.c "
                       f"or_step_omp with {threads} OpenMP threads")


def _pmc(counter: str, args: list, kernel_match: str, last: int = 0):
    """Average value of one counter per dispatch of the matching kernel, from a rocprofv3 child run
    of this script (the program itself follows `--`); `last` > 0 averages the last `last` such
    dispatches only (the timed steps, not the warm-up or the autotuner's trial launches)."""
    exe = shutil.which("rocprofv3")
    if not exe:
        return None
    with tempfile.TemporaryDirectory() as d:
        cmd = [exe, "--pmc", counter, "-d", d, "-o", "run", "--output-format", "csv", "--",
               sys.executable, os.path.abspath(__file__), *args]
        try:
            subprocess.run(cmd, check=True, stdout=subprocess.DEVNULL, stderr=subprocess.DEVNULL, timeout=240)
        except (subprocess.SubprocessError, OSError):
            return None
        per = {}
        for f in glob.glob(os.path.join(d, "**", "*counter_collection.csv"), recursive=True):
            with open(f) as fh:
                for r in csv.DictReader(fh):
                    if kernel_match in r["Kernel_Name"] and r["Counter_Name"] == counter:
                        k = int(r["Dispatch_Id"])
                        per[k] = per.get(k, 0.0) + float(r["Counter_Value"])
        vals = [per[k] for k in sorted(per)]
        if last > 0:
            vals = vals[-last:]
        return sum(vals) / len(vals) if vals else None


def measure_traffic(form: str):
    """Bytes per launch of the bond kernel between L2 and memory over a short run's timed steps:
    FETCH_SIZE and WRITE_SIZE (KiB per dispatch), each scaled by bytes / counter for a device copy of
    1 GiB (read 1 GiB, write 1 GiB).  The fetches include those that the memory-side cache serves."""
    gib = 1 << 30
    cal_args = ["--calibrate-copy-gib", "1"]
    steps = 5
    kargs = ["--form", form, "--steps", str(steps), "--warmup", "2", "--no-cpu-baseline", "--no-traffic"]
    cf = _pmc("FETCH_SIZE", cal_args, "copyBuffer")
    cw = _pmc("WRITE_SIZE", cal_args, "copyBuffer")
    kf = _pmc("FETCH_SIZE", kargs, "k_bond", last=steps)
    kw = _pmc("WRITE_SIZE", kargs, "k_bond", last=steps)
    if None in (cf, cw, kf, kw) or cf <= 0 or cw <= 0:
        return None
    sf, sw = gib / cf, gib / cw
    return dict(fetch_bytes=kf * sf, write_bytes=kw * sw, bytes=kf * sf + kw * sw, fetch_scale=sf, write_scale=sw)


def size_effect_family(seeds: int = 16):
    """The 64-instance batch: depths 50 / 100 / 200 / 400 mm x `seeds` seeded s0 perturbations.  Each
    depth's lattice is built once; the seeds differ only in their per-bond s0."""
    import copy
    from bb_pd_amd import beam
    cases = []
    for depth in (50, 100, 200, 400):
        base = beam.case_size_effect(depth, 0)
        for seed in range(seeds):
            c = copy.copy(base)
            if seed:
                U = beam.splitmix_uniform(seed, np.arange(base.nBonds, dtype=np.uint64))
                c.s0 = 1.0e-4 * (1 + 0.01 * (2 * U - 1))
                c.name = f"beam_d{depth}_seed{seed}"
            cases.append(c)
    return cases


def run_batch(cases, precision, steps, warmup, chunk=50):
    """Admit every case into one batch, then time `steps` graph-captured steps of all of them."""
    import torch
    from bb_pd_amd.batch import Batch
    slots = sum((c.nNodes + 255) // 256 * 256 for c in cases)
    t0 = time.time()
    b = Batch(cases[0], precision, max_regions=len(cases), max_slots=slots, history_rows=4096)
    for c in cases:
        b.admit(c)
    log(f"  admitted {len(cases)} problems in {time.time() - t0:.1f}s, info {b.info()}")
    b.step(warmup)
    b.sync()
    tune = b.autotune(5)
    log(f"  autotune: variant {tune[0]} of {['%.3f' % x for x in tune[1]]} ms")
    # bond-kernel duration from HIP events over eager steps (graph replays record none)
    b.set_timing(True)
    b.step(10)
    b.sync()
    avg_ms, launches = b.kernel_time()
    b.set_timing(False)
    done, dt = 0, 0.0
    while done < steps:
        n = min(chunk, steps - done)
        b.run_graph(n, 0)  # capture outside the timed region
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        b.run_graph(n, 1)
        b.sync()
        torch.cuda.synchronize()
        dt += time.perf_counter() - t0
        done += n
    nb = sum(c.nBonds for c in cases)
    st = b.status()
    b.close()
    return dict(seconds=dt, ms_per_step=dt * 1e3 / steps, bond_updates_per_s=nb * steps / dt, kernel_ms=avg_ms,
                launches=launches, steps_done=[x["step"] for x in st[:len(cases)]],
                autotune={"variant": tune[0], "ms": tune[1]})


def run_serve(n_requests: int, hbm_gb: float, ipi: int, increment_um: float, load: float = 0.8):
    """Config 4: the request mix served by the continuous-batching service.  First the closed-loop
    capacity (all requests queued at t = 0), then Poisson arrivals at `load` x that capacity; both
    forms, the same requests."""
    from bb_pd_amd import _lib
    from bb_pd_amd.serve import Request, request_case, serve
    rng = np.random.default_rng(2024)
    t0 = time.time()
    cases = [request_case(rng, ipi, increment_um) for _ in range(n_requests)]
    log(f"built {n_requests} requests ({sum(c.nBonds for c in cases)} bonds) in {time.time() - t0:.1f}s")
    out = {}
    for name, prec in (("mixed", _lib.MIXED), ("fp64", _lib.FP64)):
        reqs = [Request(i, c) for i, c in enumerate(cases)]
        cap = serve(reqs, prec, hbm_gb * 1e9, ipi, closed_loop=True)
        lam = load * cap["requests_per_s"]
        arr = np.cumsum(np.random.default_rng(7).exponential(1 / lam, n_requests))
        reqs = [Request(i, c, arrival=float(a)) for i, (c, a) in enumerate(zip(cases, arr))]
        res = serve(reqs, prec, hbm_gb * 1e9, ipi)
        log(f"{name}: capacity {cap['requests_per_s']:.2f} req/s; at {load:.0%}: p50 {res['p50_s']:.2f} s, "
            f"p99 {res['p99_s']:.2f} s, {res['bond_updates_per_s']:.3e} bond updates/s, {res['slots']} slots")
        steps = [r.steps for r in res["done"]]
        out[name] = dict(capacity_requests_per_s=cap["requests_per_s"], arrival_rate=lam, p50_s=res["p50_s"],
                         p99_s=res["p99_s"], requests_per_s=res["requests_per_s"],
                         bond_updates_per_s=res["bond_updates_per_s"], slots=res["slots"],
                         capacity_bond_updates_per_s=cap["bond_updates_per_s"],
                         steps_min=int(min(steps)), steps_max=int(max(steps)))
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=19980,
                    help="timed steps (default: the rest of the 20,000-iteration load history)")
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-traffic", action="store_true", help="skip the rocprofv3 traffic measurement")
    ap.add_argument("--no-precision", action="store_true", help="skip the FP32-mixed vs FP64 load study")
    ap.add_argument("--form", choices=["both", "fp64", "mixed"], default="both",
                    help="which precision forms to run (profiling runs one)")
    ap.add_argument("--small", action="store_true", help="dx = 2.5 mm notched trilinear beam instead")
    ap.add_argument("--iters-per-increment", type=int, default=1000,
                    help="ADR iterations per load increment (the load protocol)")
    ap.add_argument("--increment-um", type=float, default=5.0, help="penetrator travel per increment (um)")
    ap.add_argument("--batch", action="store_true",
                    help="the 64-instance size-effect batch (4 depths x 16 seeds) in one launch per kernel")
    ap.add_argument("--serve", action="store_true",
                    help="the serving workload: Poisson-arrival beam solves at 80%% of measured capacity")
    ap.add_argument("--requests", type=int, default=48, help="serving: number of requests")
    ap.add_argument("--hbm-gb", type=float, default=2.0, help="serving: device bytes the batch may hold")
    ap.add_argument("--calibrate-copy-gib", type=float, default=0.0,
                    help="only copy a device buffer of this many GiB 5 times (scales HBM counters)")
    args = ap.parse_args()
    if args.calibrate_copy_gib > 0:
        import torch
        n = int(args.calibrate_copy_gib * (1 << 30)) // 8
        a = torch.ones(n, dtype=torch.float64, device="cuda")
        b = torch.empty_like(a)
        for _ in range(5):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_per_launch": n * 8, "launches": 5}), flush=True)
        return
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs one GPU per process: --gpus must be 1 "
                                   "(the multi-GPU slab decomposition is a design, not built)"}))
        sys.exit(2)
    from bb_pd_amd import _lib, beam
    if args.serve:
        ipi = args.iters_per_increment if args.iters_per_increment != 1000 else 200
        inc = args.increment_um if args.increment_um != 5.0 else 25.0
        res = run_serve(args.requests, args.hbm_gb, ipi, inc)
        m, f = res["mixed"], res["fp64"]
        print(json.dumps({
            "metric": "p99 time-to-solution of beam fracture solves at 80% of sustained capacity (FP32-mixed; FP64 beside)",
            "value": m["p99_s"], "unit": "s", "n_gpus": 1, "higher_is_better": False, "scaling": "weak",
            "vs_baseline": None, "dtype": "f32 bond arithmetic, f64 state", "data": "synthetic beams generated in-process",
            "config": {"workload": "serving: depth {50,100,200,400} mm, notch ratio {0,0.2,0.4}, E~U[30,45] GPa, "
                                   "s0~U[0.8,1.2]e-4, bilinear, dx 5 mm, ADR",
                       "requests": args.requests, "protocol": f"{inc:g} um per {ipi} ADR iterations",
                       "stop": "load < 20% of peak or 0.5 mm", "hbm_budget_gb": args.hbm_gb},
            "mixed": m, "fp64": f,
        }), flush=True)
        return
    if args.batch:
        t0 = time.time()
        cases = size_effect_family()
        nb = sum(c.nBonds for c in cases)
        log(f"built {len(cases)} beams, {sum(c.nNodes for c in cases)} nodes, {nb} bonds in {time.time() - t0:.1f}s")
        res = {}
        for name, prec in (("mixed", _lib.MIXED), ("fp64", _lib.FP64)):
            res[name] = run_batch(cases, prec, args.steps, args.warmup)
            log(f"{name}: {res[name]['ms_per_step']:.3f} ms/step, bond kernel {res[name]['kernel_ms']:.3f} ms")
        m, f = res["mixed"], res["fp64"]
        print(json.dumps({
            "metric": "aggregate bond updates/s, size-effect batch in one launch per kernel (FP32-mixed; FP64 beside)",
            "value": m["bond_updates_per_s"], "unit": "bond updates/s", "n_gpus": 1, "steps": args.steps,
            "warmup": args.warmup, "ms_per_step": m["ms_per_step"], "higher_is_better": True, "scaling": "weak",
            "vs_baseline": None, "dtype": "f32 bond arithmetic, f64 state", "data": "synthetic beams generated in-process",
            "config": {"workload": "size-effect family: depth 50/100/200/400 mm x 16 seeds, dx 5 mm, bilinear, ADR",
                       "instances": len(cases), "nodes": sum(c.nNodes for c in cases), "bonds": nb,
                       "steps": "hipGraph-captured chunks of 50"},
            "kernel_ms": m["kernel_ms"], "autotune": m["autotune"],
            "fp64": {"value": f["bond_updates_per_s"], "ms_per_step": f["ms_per_step"], "kernel_ms": f["kernel_ms"],
                     "autotune": f["autotune"]},
            "mixed_over_fp64": m["bond_updates_per_s"] / f["bond_updates_per_s"],
        }), flush=True)
        return
    t0 = time.time()
    proto = dict(iters_per_increment=args.iters_per_increment, increment=-args.increment_um * 1e-6)
    case = beam.case_trilinear_dx2p5(**proto) if args.small else beam.case_exponential_dx1p25(**proto)
    ipi = case.itersPerIncrement
    total = args.warmup + args.steps
    log(f"built {case.name}: {case.nNodes} nodes {case.nBonds} bonds in {time.time() - t0:.1f}s")
    # the rocprofv3 child runs go first, while this process has not touched the GPU
    traffic = None if (args.no_traffic or args.form != "both") else measure_traffic("mixed")
    if traffic:
        log(f"traffic: {traffic['bytes'] / 1e9:.3f} GB per launch (fetch x{traffic['fetch_scale']:.2f}, "
            f"write x{traffic['write_scale']:.2f})")
    copy_gbs = copy_bandwidth() if args.form == "both" else None
    if copy_gbs:
        log(f"device copy: {copy_gbs:.0f} GB/s")
    study = args.form == "both" and not args.no_precision and total >= 2 * ipi
    # FP64 states kept for the precision study: after increment 1 and every fifth increment
    keep = sorted({ipi * j for j in range(1, total // ipi + 1) if j == 1 or j % 5 == 0}) if study else ()
    chunk = ipi  # load increments: checkpoints for the precision study, and autotuner reruns
    fp64 = mixed = None
    if args.form in ("both", "fp64"):
        fp64 = run_form(case, _lib.FP64, args.steps, args.warmup, chunk, keep, retune=True)
        log(f"fp64: {fp64['ms_per_step']:.3f} ms/step, kernel {fp64['kernel_ms']:.3f} ms, "
            f"{fp64['achieved_gbs']:.0f} GB/s (model bytes)")
    if args.form in ("both", "mixed"):
        mixed = run_form(case, _lib.MIXED, args.steps, args.warmup, chunk, retune=True)
        log(f"mixed: {mixed['ms_per_step']:.3f} ms/step, kernel {mixed['kernel_ms']:.3f} ms, "
            f"{mixed['achieved_gbs']:.0f} GB/s (model bytes)")
    if args.form != "both":
        res = fp64 or mixed
        res["solver"].close()
        print(json.dumps({"form": args.form, "result": {k: v for k, v in res.items()
                                                         if k not in ("solver", "info", "hist", "ckpts", "autotune")}}),
              flush=True)
        return
    prec = None
    if study:
        t0 = time.time()
        prec = precision_study(case, fp64, mixed, ipi)
        log(f"precision study {time.time() - t0:.0f}s: peak fp64 {prec['peak_fp64']:.1f} (inc "
            f"{prec['peak_increment_fp64']}), mixed {prec['peak_mixed']:.1f}, err {prec['peak_load_err']:.2e}, "
            f"fp64 rounding floor {prec['fp64_rounding_floor']:.2e}, FP32-sized "
            f"{prec['fp64_fp32_sized_floor']:.2e}, same-state peak err "
            f"{prec['same_state_peak_load_err']:.2e}")
    fp64["solver"].close()
    mixed["solver"].close()
    n_hist = int(fp64["hist"]["reaction_static"].size)
    cpu = None
    if not args.no_cpu_baseline:
        cpu = cpu_baseline(case)
        log(f"cpu baseline: {cpu['value']:.3e} bond updates/s on {cpu['cores']} threads")
    head = mixed
    out = {
        "metric": "bond updates/s (FP32-mixed; FP64 beside) and HBM roofline fraction of the bond kernel",
        "value": head["bond_updates_per_s"],
        "unit": "bond updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": head["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 bond arithmetic, f64 state",
        "data": "synthetic beam lattice generated in-process",
        "config": {
            "workload": case.name,
            "nodes": case.nNodes,
            "bonds": case.nBonds,
            "law": "exponential (k=25, alpha=0.25)" if case.law == _lib.LAW_EXPONENTIAL else "trilinear",
            "integrator": "ADR",
            "load": f"penetrator staircase {abs(case.increment) * 1e6:g} um per {ipi} iterations",
            "state": f"steps {args.warmup + 1}..{total} of a {n_hist}-step load history (elastic, peak, "
                     f"softening)",
        },
        "fp64": {"value": fp64["bond_updates_per_s"], "ms_per_step": fp64["ms_per_step"],
                 "kernel_ms": fp64["kernel_ms"], "bytes_model": fp64["bytes_model"],
                 "bytes_layout": fp64["bytes_layout"], "achieved_gbs": fp64["achieved_gbs"], "frac": fp64["frac"],
                 "retune_picks": {str(v): fp64["retune_picks"].count(v) for v in sorted(set(fp64["retune_picks"]))}},
        "mixed_over_fp64": head["bond_updates_per_s"] / fp64["bond_updates_per_s"],
        "mixed_vs_fp64_load_err": prec["same_state_peak_load_err"] if prec else None,
        "precision": prec,
        "roofline": {"bound": "hbm", "achieved": head["achieved_gbs"], "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": head["frac"], "traffic": traffic["bytes"] if traffic else None,
                     "traffic_fetch": traffic["fetch_bytes"] if traffic else None,
                     "traffic_write": traffic["write_bytes"] if traffic else None,
                     "frac_layout_bytes": head["bytes_layout"] / (head["kernel_ms"] * 1e-3) / 1e9 / HBM_PEAK_GBS,
                     "copy_gbs": copy_gbs,
                     "traffic_gbs": traffic["bytes"] / (head["kernel_ms"] * 1e-3) / 1e9 if traffic else None,
                     "kernel": f"bond kernel, FP32-mixed, {head['layout']} layout", "kernel_ms": head["kernel_ms"],
                     "autotune": {"variant": head["autotune"][0], "ms": head["autotune"][1],
                                  "retune_picks": {str(v): head["retune_picks"].count(v)
                                                   for v in sorted(set(head["retune_picks"]))}}
                                 if head["autotune"] else None,
                     "bytes_model": head["bytes_model"], "bytes_layout": head["bytes_layout"]},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
