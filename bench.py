"""Benchmark: explicit bond-based peridynamics steps on one MI355X.

Workload: the mixed-mode three-point-bend beam 500x100x50 mm at dx = 1.25 mm (400x40x80 =
1,280,000 nodes, 74.5M bonds, offset notch, exponential softening) under ADR on the equilibrium
staircase: the penetrator moves 5 um per load increment, and an increment ends once the equilibrium
ratio (out-of-balance force over the free DOFs against the load) falls below 1e-4, after at least
50 and at most 6000 iterations.  One step = stretch -> law -> force -> nodal gather -> ADR update
-> contact -> record.  Timed steps run in HIP-graph-captured chunks (a few eager steps per chunk
carry HIP events around the bond kernel).

Reported (one JSON line on stdout):
  value            bond updates/s of the FP32-mixed form (= nBonds * K / seconds, inputs resident)
                   when the precision gate holds, else of the plain FP64 form
  fp64             the same for the plain FP64 form
  precision        FP32-mixed against FP64 on the free-running load history: the peak of the load at
                   the increment ends (converged states), with the FP64 run's own rounding floors
                   beside it (one ulp on one coordinate, 2^-24 relative noise on every coordinate,
                   each after the first increment), flipped softening flags, and per-result errors
                   from identical FP64 states (load, nodal force, flags, broken bonds, damage field)
  roofline         dominant kernel (the bond kernel of the headline form): its algorithmic bytes per
                   launch at the states of the timed window (the stencil layout's history only where
                   a slice's history word is set) over its average launch time from HIP events,
                   against 8 TB/s; `traffic` = measured bytes beyond L2 per launch (rocprofv3
                   FETCH_SIZE + WRITE_SIZE of the same launch variant at a state of the window, each
                   scaled by the same counter's reading for a device copy of known size); `valu_busy`
                   the VALU issue utilisation of the same launches; `frac_csr_model` the same kernel
                   time against the bytes of a node-centric CSR layout, for comparison
  cpu_baseline     the oracle's OpenMP restatement of the reference CPU passes (oracle/bbpd_oracle.c
                   or_step_omp) timed on the host cores, all physical cores and 8 threads, bound
                   close to cores, median of 5, at an elastic and a post-peak state of the same beam
"""
from __future__ import annotations

import argparse
import csv
import glob
import json
import os
import platform
import shutil
import subprocess
import sys
import tempfile
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
GATE = 0.005  # FP32-mixed peak load within 0.5% of FP64


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def make_case(args):
    from bb_pd_amd import beam
    proto = dict(schedule=2, iters_per_increment=args.iters_max, increment=-args.increment_um * 1e-6,
                 eq_tol=args.eq_tol, iters_per_increment_min=args.iters_min)
    return beam.case_trilinear_dx2p5(**proto) if args.small else beam.case_exponential_dx1p25(**proto)


def protocol_text(args):
    return (f"equilibrium staircase: {args.increment_um:g} um per increment, next increment once the "
            f"equilibrium ratio < {args.eq_tol:g} after >= {args.iters_min} iterations, at most "
            f"{args.iters_max}")


def csr_model_bytes(case, mixed: bool) -> int:
    """Bytes one bond-kernel launch must move in a node-centric CSR layout with the neighbour index,
    kappa and the per-bond parameters stored per entry: E*(4 + w_state + w_kappa + n_par*w) + N*180."""
    w = 4 if mixed else 8
    n_par = int(isinstance(case.s0, np.ndarray)) + int(isinstance(case.sc, np.ndarray))
    E = 2 * case.nBonds
    return E * (4 + w + w + n_par * w) + case.nNodes * 180


def eager_count(n: int) -> int:
    """Eager steps (with HIP events around the bond kernel) at the start of a timed chunk of n."""
    return min(n, max(2, n // 20))


def curve_of(hist):
    """(step, load-point displacement, load) at the increment ends of a history (row 0 = step 1)."""
    e = np.nonzero(hist["inc_end"])[0]
    return e + 1, hist["disp"][e], hist["reaction_static"][e]


def past_peak(loads, after: int = 2, frac: float = 0.9) -> bool:
    """The converged load curve has passed its peak by `after` increments and fallen below frac of it."""
    if len(loads) < 3:
        return False
    m = int(np.argmax(loads))
    return len(loads) - 1 - m >= after and loads[-1] < frac * loads[m]


class Run:
    """One solver stepped through the load history: the timed window first, then (untimed) as far as
    the precision study needs.  Keeps the per-step history from step 1 and, for the reference run,
    checkpoints at chunk boundaries after chosen increment ends."""

    def __init__(self, case, precision, keep_checkpoints=False):
        from bb_pd_amd.solver import Solver
        t0 = time.time()
        self.case = case
        self.sol = Solver(case, precision)
        self.mixed = precision == 1
        self.rows = []
        self.ckpts = []            # (increments done, bond bytes at that state, checkpoint)
        self.keep = keep_checkpoints
        self.marks = [1, 5, 10, 15]  # checkpoint at the first chunk boundary after these increment ends
        self.step = 0
        log(f"  setup {time.time() - t0:.1f}s, {self.sol.layout()}, info {self.sol.info()}")

    def hist(self):
        return {k: np.concatenate([r[k] for r in self.rows]) for k in self.rows[0]}

    def n_incs(self):
        return int(sum(int(r["inc_end"].sum()) for r in self.rows))

    def loads(self):
        return np.concatenate([r["reaction_static"][r["inc_end"]] for r in self.rows]) if self.rows else np.zeros(0)

    def _record(self):
        self.rows.append(self.sol.history())
        if self.keep and self.marks and self.n_incs() >= self.marks[0]:
            n = self.n_incs()
            self.marks = [m for m in self.marks if m > n]
            self.ckpts.append((n, self.sol.bond_bytes(), self.sol.checkpoint()))

    def timed(self, steps, warmup, chunk):
        import torch
        sol = self.sol
        sol.step(warmup)
        sol.sync()
        self.step = warmup
        self._record()
        tunable = sol.layout().startswith("stencil")
        tune = sol.autotune(5) if tunable else None
        variant = tune[0] if tune else 0
        bb = [sol.bond_bytes()]
        sol.set_timing(True)
        n0 = min(chunk, steps)
        if n0 - eager_count(n0) > 0:
            sol.run_graph(n0 - eager_count(n0), 0)   # the first chunk's graph, captured before the clock
        dt, done, picks = 0.0, 0, []
        while done < steps:
            n = min(chunk, steps - done)
            e = eager_count(n)
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            if done and tunable:
                # the launch variant follows the load history; a change recaptures the graph
                variant = sol.autotune(2)[0]
                picks.append(variant)
            sol.step(e)
            if n - e:
                sol.run_graph(n - e, 1)
            sol.sync()
            torch.cuda.synchronize()
            dt += time.perf_counter() - t0
            done += n
            self.step += n
            self._record()
            bb.append(sol.bond_bytes())
            if len(self.rows) % 5 == 0:
                log(f"    {self.step} steps, {self.n_incs()} increments")
        kernel_ms, launches = sol.kernel_time()
        sol.set_timing(False)
        self.softening_at_window_end = int(sol.observables()["n_softening"])
        # bytes per launch along the window: the mean of the samples at the chunk boundaries
        bytes_layout = float(np.mean([b["bytes"] for b in bb]))
        return dict(seconds=dt, ms_per_step=dt * 1e3 / steps, bond_updates_per_s=self.case.nBonds * steps / dt,
                    kernel_ms=kernel_ms, launches=launches, bytes_layout=bytes_layout,
                    history_offsets_set=bb[-1]["offsets_set"] / max(bb[-1]["offsets_total"], 1),
                    variant=variant, autotune=tune,
                    retune_picks={str(v): picks.count(v) for v in sorted(set(picks))})

    def advance(self, until, max_steps, chunk=2000):
        """Untimed graph-captured chunks until `until(self)` or max_steps."""
        while self.step < max_steps and not until(self):
            n = min(chunk, max_steps - self.step)
            self.sol.run_graph(n, 1)
            self.sol.sync()
            self.step += n
            self._record()
            if len(self.rows) % 10 == 0:
                log(f"    {self.step} steps, {self.n_incs()} increments")

    def close(self):
        self.sol.close()


def window_label(run, warmup, steps, peak_inc):
    """What the timed window covered, from its own history: increments, softening, the peak."""
    h = run.hist()
    lo, hi = warmup, warmup + steps
    incs_before = int(h["inc_end"][:lo].sum())
    incs_in = int(h["inc_end"][lo:hi].sum())
    first = incs_before + 1
    last = incs_before + incs_in + (0 if h["inc_end"][hi - 1] else 1)
    parts = [f"steps {lo + 1}..{hi} of the load history: load increments {first}..{last}, "
             f"{run.softening_at_window_end} softening bonds at the window's end"]
    if peak_inc:
        if last < peak_inc:
            parts.append(f"before the peak (increment {peak_inc})")
        elif first > peak_inc:
            parts.append(f"after the peak (increment {peak_inc})")
        else:
            parts.append(f"elastic loading, the peak (increment {peak_inc}) and softening")
    return ", ".join(parts)


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


# ------------------------------------------------------------------------------------------------
# Precision study
# ------------------------------------------------------------------------------------------------
def one_step_from(sol, st):
    """Restore a checkpoint, run one step; (nodal force, static reaction, H, flag, fail, damage field,
    damage energy), the damage observables computed on the device."""
    sol.restore(st["step"], **{k: v for k, v in st.items() if k != "step"})
    sol.history()
    sol.step(1)
    sol.sync()
    R = float(sol.history()["reaction_static"][-1])
    F = sol.nodal()[2]
    H, flag, fail = sol.bonds()
    D, W = sol.damage()
    return F, R, H, flag, fail, D, W


def perturbed(case, st, kind, seed=7):
    st = {k: (v.copy() if isinstance(v, np.ndarray) else v) for k, v in st.items()}
    if kind == "ulp":
        i = int(np.argmax(np.abs(st["u"][:, 2])))
        st["x"][i, 2] = np.nextafter(st["x"][i, 2], np.inf)
        st["u"][i, 2] = st["x"][i, 2] - case.X[i, 2]
    else:
        sgn = np.where(np.random.default_rng(seed).random(st["x"].shape) < 0.5, -1.0, 1.0)
        st["x"] = np.asfortranarray(st["x"] * (1 + sgn * 2.0 ** -24))
        st["u"] = np.asfortranarray(st["x"] - case.X)
    return st


def floor_run(ref, st, n_incs, max_steps):
    """The FP64 reference restarted from a perturbed state and run to the same increment count;
    returns its load curve (the reference's loads before the restart, then its own)."""
    sol = ref.sol
    sol.restore(st["step"], **{k: v for k, v in st.items() if k != "step"})
    sol.history()
    h = ref.hist()
    loads = list(h["reaction_static"][:st["step"]][h["inc_end"][:st["step"]]])
    step = st["step"]
    while len(loads) < n_incs and step < max_steps:
        sol.run_graph(2000, 1)
        sol.sync()
        step += 2000
        r = sol.history()
        loads += list(r["reaction_static"][r["inc_end"]])
    return np.array(loads[:n_incs])


def precision_study(case, ref, mix, max_steps, floor_samples=0):
    """FP32-mixed against FP64 on the load history (the load is the penetrator family's static
    reaction, calculatereactionforceFast.m:3-5, at the increment ends of the equilibrium staircase).
    `floor_samples` more FP64 runs with FP32-sized noise (other seeds) give the spread of FP64's own
    peak under rounding-level perturbations."""
    t0 = time.time()
    ref.advance(lambda r: past_peak(r.loads()), max_steps)
    n = ref.n_incs()
    mix.advance(lambda r: r.n_incs() >= n, max_steps)
    c64, cm = ref.loads()[:n], mix.loads()[:n]
    m = min(len(c64), len(cm))
    p64, pm = int(np.argmax(c64)), int(np.argmax(cm))
    peak = float(c64[p64])
    log(f"  free runs {time.time() - t0:.0f}s: FP64 peak {peak:.2f} N at increment {p64 + 1}, mixed "
        f"{cm[pm]:.2f} N at {pm + 1}")
    # flags at the end of the free runs (the same step count)
    H64, fl64, fa64 = ref.sol.bonds()
    Hm, flm, fam = mix.sol.bonds()
    end_state = dict(step_fp64=ref.step, step_mixed=mix.step, softening_fp64=int(fl64.sum()),
                     softening_mixed=int(flm.sum()), flag_flips=int((fl64 != flm).sum()),
                     broken_fp64=int(((H64 == 1) | (fa64 == 0)).sum()),
                     broken_mixed=int(((Hm == 1) | (fam == 0)).sum()))
    final_ck = ref.sol.checkpoint()
    # same state: one step of each form from FP64 states along the history
    samples = [(c[0], c[2]) for c in ref.ckpts] + [(n, final_ck)]
    same = []
    for incs, st in samples:
        Fa, Ra, Ha, fla, faa, Da, Wa = one_step_from(ref.sol, st)
        Fb, Rb, Hb, flb, fab, Db, Wb = one_step_from(mix.sol, st)
        same.append(dict(step=int(st["step"]), increments=int(incs), load=Ra,
                         load_err=(Rb - Ra) / Ra if Ra else None,
                         nodal_force_err=float(np.abs(Fb - Fa).max() / np.abs(Fa).max()),
                         flag_flips=int((fla != flb).sum()),
                         broken_fp64=int(((Ha == 1) | (faa == 0)).sum()),
                         broken_mixed=int(((Hb == 1) | (fab == 0)).sum()),
                         damage_err=float(np.abs(Db - Da).max()),
                         damage_energy=Wa, damage_energy_err=(Wb - Wa) / Wa if Wa else None))
        log(f"  same state at step {st['step']}: {same[-1]}")
    near = min(same, key=lambda x: abs(x["increments"] - (p64 + 1)))
    # rounding floors of the FP64 run itself, restarted after its first increment
    first = ref.ckpts[0][2]
    c_ulp = floor_run(ref, perturbed(case, first, "ulp"), n, max_steps)
    c_noise = floor_run(ref, perturbed(case, first, "noise"), n, max_steps)
    err = float((cm[pm] - peak) / peak)
    extra = []
    for seed in range(floor_samples):
        c = floor_run(ref, perturbed(case, first, "noise", seed=100 + seed), n, max_steps)
        extra.append(float((c.max() - peak) / peak))
        log(f"  FP32-sized floor, seed {100 + seed}: peak {c.max():.2f} N ({extra[-1]:+.2e})")
    out = dict(
        protocol=None,
        load="static reaction of the penetrator family at the load-increment ends (converged states)",
        increments=n, peak_fp64=peak, peak_increment_fp64=p64 + 1, peak_mixed=float(cm[pm]),
        peak_increment_mixed=pm + 1, peak_load_err=err,
        fp64_rounding_floor=float((c_ulp.max() - peak) / peak),
        fp64_fp32_sized_floor=float((c_noise.max() - peak) / peak),
        curve_err_max=float(np.abs(cm[:m] - c64[:m]).max() / peak),
        curve_err_max_to_peak=float(np.abs(cm[:p64 + 1] - c64[:p64 + 1]).max() / peak),
        curve_floor_max=float(np.abs(c_ulp - c64[:len(c_ulp)]).max() / peak),
        curve_fp32_sized_floor_max=float(np.abs(c_noise - c64[:len(c_noise)]).max() / peak),
        curve_fp64=[round(float(v), 3) for v in c64], curve_mixed=[round(float(v), 3) for v in cm],
        curve_fp64_ulp=[round(float(v), 3) for v in c_ulp],
        curve_fp64_fp32_sized=[round(float(v), 3) for v in c_noise],
        increment_end_steps_fp64=[int(x) for x in curve_of(ref.hist())[0][:n]],
        end_state=end_state, same_state_peak_load_err=near["load_err"], same_state=same,
        **(dict(fp64_fp32_sized_floor_samples=extra) if floor_samples else {}),
        gate={"peak_load_err_max": GATE, "passed": abs(err) <= GATE})
    if floor_samples:
        # every FP64 run's peak against the unperturbed one (0), the mixed form's beside them
        runs = [0.0, out["fp64_rounding_floor"], out["fp64_fp32_sized_floor"]] + extra
        out["fp64_peak_err_range"] = [min(runs), max(runs)]
        out["mixed_within_fp64_range"] = min(runs) <= err <= max(runs)
    log(f"  precision study {time.time() - t0:.0f}s: peak err {err:+.2e}, floors ulp "
        f"{out['fp64_rounding_floor']:+.2e}, FP32-sized {out['fp64_fp32_sized_floor']:+.2e}")
    return out, samples


# ------------------------------------------------------------------------------------------------
# Counters (rocprofv3 child runs of this script) and the CPU baseline (child run)
# ------------------------------------------------------------------------------------------------
def save_state(path, st):
    np.savez(path, step=st["step"], u=st["u"], v=st["v"], x=st["x"], F=st["F"], H=st["H"],
             flag=st["flag"].astype(np.uint8), fail=st["fail"].astype(np.uint8),
             load_state=np.array(st.get("load_state", (1, 0)), dtype=np.int64))


def load_state_file(path):
    z = np.load(path)
    st = dict(step=int(z["step"]), u=np.asfortranarray(z["u"]), v=np.asfortranarray(z["v"]),
              x=np.asfortranarray(z["x"]), F=np.asfortranarray(z["F"]), Fp=np.asfortranarray(z["F"]),
              H=z["H"], flag=z["flag"].astype(np.float64), fail=z["fail"].astype(np.float64))
    st["load_state"] = tuple(int(v) for v in z["load_state"])
    return st


def _pmc(counters: list, args: list, kernel_match: str, last: int = 0):
    """Average value of each counter per dispatch of the matching kernel, from a rocprofv3 child run
    of this script (the program itself follows `--`); `last` > 0 averages the last `last` such
    dispatches only (the measured steps, not the autotuner's or a restore's launches)."""
    exe = shutil.which("rocprofv3")
    if not exe:
        return None
    with tempfile.TemporaryDirectory() as d:
        cmd = [exe, "--pmc", *counters, "-d", d, "-o", "run", "--output-format", "csv", "--",
               sys.executable, os.path.abspath(__file__), *args]
        try:
            subprocess.run(cmd, check=True, stdout=subprocess.DEVNULL, stderr=subprocess.DEVNULL, timeout=300)
        except (subprocess.SubprocessError, OSError) as e:
            log(f"  rocprofv3 run failed: {e}")
            return None
        per = {}
        for f in glob.glob(os.path.join(d, "**", "*counter_collection.csv"), recursive=True):
            with open(f) as fh:
                for r in csv.DictReader(fh):
                    if kernel_match in r["Kernel_Name"] and r["Counter_Name"] in counters:
                        k = int(r["Dispatch_Id"])
                        per.setdefault(k, {}).setdefault(r["Counter_Name"], 0.0)
                        per[k][r["Counter_Name"]] += float(r["Counter_Value"])
        keys = sorted(per)
        if last > 0:
            keys = keys[-last:]
        if not keys:
            return None
        return {c: sum(per[k].get(c, 0.0) for k in keys) / len(keys) for c in counters}


def measure_counters(form: str, variant: int, state_path, warmup: int, small: bool):
    """The bond kernel's bytes beyond L2 per launch (FETCH_SIZE and WRITE_SIZE, KiB per dispatch,
    each scaled by bytes / counter for a 1 GiB device copy) and its VALU issue utilisation
    (SQ_ACTIVE_INST_VALU x 4 cycles over 1,024 SIMDs, per kernel cycle of GRBM_GUI_ACTIVE summed over
    the 8 XCDs), over 5 steps of the given launch variant from the given state."""
    gib = 1 << 30
    steps = 5
    kargs = ["--form", form, "--steps", str(steps), "--launch-variant", str(variant), "--no-cpu-baseline",
             "--no-traffic", "--no-precision"]
    kargs += ["--state-from", state_path, "--warmup", "0"] if state_path else ["--warmup", str(warmup)]
    if small:
        kargs.append("--small")
    cal = ["--calibrate-copy-gib", "1"]
    cf = _pmc(["FETCH_SIZE"], cal, "copyBuffer")
    cw = _pmc(["WRITE_SIZE"], cal, "copyBuffer")
    kf = _pmc(["FETCH_SIZE"], kargs, "k_bond", last=steps)
    kw = _pmc(["WRITE_SIZE"], kargs, "k_bond", last=steps)
    kv = _pmc(["SQ_ACTIVE_INST_VALU", "SQ_INSTS_VALU", "SQ_WAVES", "GRBM_GUI_ACTIVE"], kargs, "k_bond", last=steps)
    if None in (cf, cw, kf, kw) or cf["FETCH_SIZE"] <= 0 or cw["WRITE_SIZE"] <= 0:
        return None
    sf, sw = gib / cf["FETCH_SIZE"], gib / cw["WRITE_SIZE"]
    out = dict(fetch_bytes=kf["FETCH_SIZE"] * sf, write_bytes=kw["WRITE_SIZE"] * sw,
               bytes=kf["FETCH_SIZE"] * sf + kw["WRITE_SIZE"] * sw, fetch_scale=sf, write_scale=sw)
    if kv and kv["GRBM_GUI_ACTIVE"] > 0:
        cycles = kv["GRBM_GUI_ACTIVE"] / 8
        out.update(valu_busy=kv["SQ_ACTIVE_INST_VALU"] * 4 / 1024 / cycles, valu_insts=kv["SQ_INSTS_VALU"],
                   waves=kv["SQ_WAVES"], kernel_cycles=cycles)
    return out


def physical_cores():
    """Physical cores of the host (lscpu), the CPUs this process may use, and the CPU model."""
    cores, model = None, platform.processor() or ""
    try:
        out = subprocess.run(["lscpu", "-p=CORE,SOCKET"], capture_output=True, text=True, timeout=10).stdout
        cores = len({ln for ln in out.splitlines() if ln and not ln.startswith("#")})
        info = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        for ln in info.splitlines():
            if ln.startswith("Model name:"):
                model = ln.split(":", 1)[1].strip()
    except (OSError, subprocess.SubprocessError):
        pass
    avail = len(os.sched_getaffinity(0))
    return cores or avail, avail, model


def cpu_baseline(args, states: dict):
    """Child runs of this script (--cpu-child), one per thread count, each timing both states."""
    cores, avail, model = physical_cores()
    tmax = min(cores, avail)
    results = {}
    with tempfile.TemporaryDirectory() as d:
        paths = {}
        for name, st in states.items():
            paths[name] = os.path.join(d, f"{name}.npz")
            save_state(paths[name], st)
        for T in sorted({tmax, 8}):
            env = dict(os.environ, OMP_NUM_THREADS=str(T), OMP_PROC_BIND="close", OMP_PLACES="cores")
            cmd = [sys.executable, os.path.abspath(__file__), "--cpu-child", str(T)]
            for name in states:
                cmd += ["--cpu-state", f"{name}={paths[name]}"]
            if args.small:
                cmd.append("--small")
            try:
                r = subprocess.run(cmd, env=env, capture_output=True, text=True, timeout=900)
            except subprocess.SubprocessError as e:
                log(f"  cpu child failed: {e}")
                continue
            if r.returncode != 0:
                log(f"  cpu child failed: {r.stderr[-2000:]}")
                continue
            results[T] = json.loads(r.stdout.strip().splitlines()[-1])
            log(f"  cpu baseline, {T} threads: {results[T]}")
    if tmax not in results:
        return None
    per = results[tmax]
    value = float(np.mean([v["median"] for v in per.values()]))
    return dict(value=value, unit="bond updates/s", cores=tmax, threads=tmax, kind="port",
                physical_cores=cores, cpus_available=avail, cpu_model=model,
                binding="OMP_PROC_BIND=close OMP_PLACES=cores",
                by_threads={str(T): {k: v["median"] for k, v in res.items()} for T, res in results.items()},
                sample=(f"median of 5 samples of {next(iter(per.values()))['steps']} steps each, at the "
                        f"{' and '.join(per)} states of the same {args.bonds}-bond beam taken from the FP64 "
                        f"run; value = mean of the two medians at {tmax} threads: oracle/bbpd_oracle.c "
                        f"or_step_omp (passes A-E of the reference CPU MEX in its step order, exponential "
                        f"law; the MATLAB-side contact, zeroing and damage-energy loop are not timed)"))


def cpu_child(args):
    """Time or_step_omp at each given state: 5 samples, the median in bond updates/s."""
    from oracle import oracle as O
    from bb_pd_amd import beam
    T = args.cpu_child
    case = make_case(args)
    # the reference's CPU path steps with Euler-Cromer at the stable time step of
    # Gregoire_beam_tests.m:244, which from a loaded state moves it by little over a sample (the ADR
    # pseudo time step of 1 would not)
    case.DT = beam.euler_cromer_dt(case)
    out = {}
    for spec in args.cpu_state:
        name, path = spec.split("=", 1)
        st = load_state_file(path)
        run = O.OracleRun(case, nodal_threads=T)
        for k in ("u", "v", "x", "F", "H", "flag", "fail"):
            np.copyto(getattr(run, k), st[k])
        run.step_omp(1, T)   # first touch
        for k in ("u", "v", "x", "H", "flag", "fail"):
            np.copyto(getattr(run, k), st[k])
        # steps per sample: about 1 s of work
        t0 = time.perf_counter()
        run.step_omp(1, T)
        per = time.perf_counter() - t0
        n = max(1, int(1.0 / per))
        rates = []
        for _ in range(5):
            for k in ("u", "v", "x", "H", "flag", "fail"):
                np.copyto(getattr(run, k), st[k])
            t0 = time.perf_counter()
            run.step_omp(n, T)
            rates.append(case.nBonds * n / (time.perf_counter() - t0))
        out[name] = dict(median=float(np.median(rates)), samples=[float(x) for x in rates], steps=n)
        del run
    print(json.dumps(out), flush=True)


# ------------------------------------------------------------------------------------------------
# Batch and serving workloads
# ------------------------------------------------------------------------------------------------
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
    """The request mix served by the continuous-batching service.  First the closed-loop capacity
    (all requests queued at t = 0), then Poisson arrivals at `load` x that capacity; both forms, the
    same requests.  Each request's inputs are generated after it arrives."""
    from bb_pd_amd import _lib
    from bb_pd_amd.serve import Request, request_params, serve
    rng = np.random.default_rng(2024)
    params = [request_params(rng) for _ in range(n_requests)]
    out = {}
    for name, prec in (("mixed", _lib.MIXED), ("fp64", _lib.FP64)):
        reqs = [Request(i, params=p) for i, p in enumerate(params)]
        cap = serve(reqs, prec, hbm_gb * 1e9, chunk=ipi, ipi=ipi, increment_um=increment_um, closed_loop=True,
                    log=log)
        lam = load * cap["requests_per_s"]
        arr = np.cumsum(np.random.default_rng(7).exponential(1 / lam, n_requests))
        reqs = [Request(i, params=p, arrival=float(a)) for i, (p, a) in enumerate(zip(params, arr))]
        res = serve(reqs, prec, hbm_gb * 1e9, chunk=ipi, ipi=ipi, increment_um=increment_um, log=log)
        log(f"{name}: capacity {cap['requests_per_s']:.2f} req/s; at {load:.0%}: p50 {res['p50_s']:.2f} s, "
            f"p99 {res['p99_s']:.2f} s, {res['bond_updates_per_s']:.3e} bond updates/s, {res['slots']} slots")
        steps = [r.steps for r in res["done"]]
        out[name] = dict(capacity_requests_per_s=cap["requests_per_s"], arrival_rate=lam, p50_s=res["p50_s"],
                         p90_s=res["p90_s"], p99_s=res["p99_s"], max_s=res["max_s"],
                         requests_per_s=res["requests_per_s"], bond_updates_per_s=res["bond_updates_per_s"],
                         capacity_bond_updates_per_s=cap["bond_updates_per_s"], slots=res["slots"],
                         served=res["requests"], rejected=res["rejected"],
                         build_s_mean=res["build_s_mean"], prepare_s_mean=res["prepare_s_mean"],
                         admit_s_mean=res["admit_s_mean"],
                         admission_wait_s_mean=res["admission_wait_s_mean"], wall_s=res["wall_s"],
                         wall_split=res["wall_split"],
                         steps_min=int(min(steps)), steps_max=int(max(steps)))
    return out


# ------------------------------------------------------------------------------------------------
def profile_child(args):
    """One form for a counter run: optionally restore a state, fix the launch variant, run `steps`."""
    from bb_pd_amd import _lib
    from bb_pd_amd.solver import Solver
    case = make_case(args)
    with Solver(case, _lib.FP64 if args.form == "fp64" else _lib.MIXED) as sol:
        if args.state_from:
            st = load_state_file(args.state_from)
            sol.restore(st["step"], **{k: v for k, v in st.items() if k != "step"})
        sol.step(args.warmup)
        sol.step(args.steps)
        sol.sync()
        print(json.dumps({"form": args.form, "layout": sol.layout(), "steps": args.steps}), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=34000,
                    help="timed steps (default: the load history through the peak)")
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-traffic", action="store_true", help="skip the rocprofv3 counter runs")
    ap.add_argument("--no-precision", action="store_true", help="skip the FP32-mixed vs FP64 load study")
    ap.add_argument("--floor-samples", type=int, default=0,
                    help="precision study: this many more FP64 runs under FP32-sized noise (the spread of FP64's peak)")
    ap.add_argument("--form", choices=["both", "fp64", "mixed"], default="both",
                    help="which precision forms to run (a counter run runs one)")
    ap.add_argument("--small", action="store_true", help="dx = 2.5 mm notched trilinear beam instead")
    ap.add_argument("--increment-um", type=float, default=5.0, help="penetrator travel per increment (um)")
    ap.add_argument("--eq-tol", type=float, default=1e-4, help="equilibrium ratio that ends an increment")
    ap.add_argument("--iters-min", type=int, default=50, help="least ADR iterations per increment")
    ap.add_argument("--iters-max", type=int, default=6000, help="most ADR iterations per increment")
    ap.add_argument("--chunk", type=int, default=2000, help="steps per timed chunk")
    ap.add_argument("--max-study-steps", type=int, default=120000)
    ap.add_argument("--launch-variant", type=int, default=None, help="counter runs: fixed bond-kernel variant")
    ap.add_argument("--state-from", default=None, help="counter runs: restore this state file first")
    ap.add_argument("--save-state", default=None,
                    help="single-form runs: write the state after the timed steps to this .npz (for counter runs)")
    ap.add_argument("--cpu-child", type=int, default=0, help=argparse.SUPPRESS)
    ap.add_argument("--cpu-state", action="append", default=[], help=argparse.SUPPRESS)
    ap.add_argument("--batch", action="store_true",
                    help="the 64-instance size-effect batch (4 depths x 16 seeds) in one launch per kernel")
    ap.add_argument("--serve", action="store_true",
                    help="the serving workload: Poisson-arrival beam solves at 80%% of measured capacity")
    ap.add_argument("--requests", type=int, default=240, help="serving: number of requests")
    ap.add_argument("--hbm-gb", type=float, default=2.0, help="serving: device bytes the batch may hold")
    ap.add_argument("--serve-ipi", type=int, default=200, help="serving: ADR iterations per increment")
    ap.add_argument("--serve-increment-um", type=float, default=25.0, help="serving: travel per increment")
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
    if args.cpu_child:
        cpu_child(args)
        return
    if args.launch_variant is not None:
        os.environ["BBPD_LAUNCH_VARIANT"] = str(args.launch_variant)
        profile_child(args)
        return
    from bb_pd_amd import _lib
    if args.serve:
        res = run_serve(args.requests, args.hbm_gb, args.serve_ipi, args.serve_increment_um)
        m, f = res["mixed"], res["fp64"]
        print(json.dumps({
            "metric": "p99 time-to-solution of beam fracture solves at 80% of sustained capacity (FP32-mixed; FP64 beside)",
            "value": m["p99_s"], "unit": "s", "n_gpus": 1, "higher_is_better": False, "scaling": "weak",
            "vs_baseline": None, "dtype": "f32 bond arithmetic, f64 state", "data": "synthetic beams generated in-process",
            "config": {"workload": "serving: depth {50,100,200,400} mm, notch ratio {0,0.2,0.4}, E~U[30,45] GPa, "
                                   "s0~U[0.8,1.2]e-4, bilinear, dx 5 mm, ADR",
                       "requests": args.requests,
                       "protocol": f"{args.serve_increment_um:g} um per {args.serve_ipi} ADR iterations",
                       "stop": "load < 20% of peak or 0.5 mm", "hbm_budget_gb": args.hbm_gb,
                       "time_to_solution": "arrival to retirement, input generation after arrival included",
                       "input_generation": "inputs generated and laid out for the batch (bbpd_batch_prepare) by 8 host threads, at most 16 queued requests ahead"},
            "mixed": m, "fp64": f,
        }), flush=True)
        return
    if args.batch:
        t0 = time.time()
        cases = size_effect_family()
        nb = sum(c.nBonds for c in cases)
        log(f"built {len(cases)} beams, {sum(c.nNodes for c in cases)} nodes, {nb} bonds in {time.time() - t0:.1f}s")
        steps = args.steps if args.steps != 34000 else 19980
        res = {}
        for name, prec in (("mixed", _lib.MIXED), ("fp64", _lib.FP64)):
            res[name] = run_batch(cases, prec, steps, args.warmup)
            log(f"{name}: {res[name]['ms_per_step']:.3f} ms/step, bond kernel {res[name]['kernel_ms']:.3f} ms")
        m, f = res["mixed"], res["fp64"]
        print(json.dumps({
            "metric": "aggregate bond updates/s, size-effect batch in one launch per kernel (FP32-mixed; FP64 beside)",
            "value": m["bond_updates_per_s"], "unit": "bond updates/s", "n_gpus": 1, "steps": steps,
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
    case = make_case(args)
    args.bonds = case.nBonds
    log(f"built {case.name}: {case.nNodes} nodes {case.nBonds} bonds in {time.time() - t0:.1f}s")
    if args.form != "both":
        run = Run(case, _lib.FP64 if args.form == "fp64" else _lib.MIXED)
        res = run.timed(args.steps, args.warmup, args.chunk)
        if args.save_state:
            save_state(args.save_state, run.sol.checkpoint())
        run.close()
        print(json.dumps({"form": args.form, "result": res}), flush=True)
        return
    copy_gbs = copy_bandwidth()
    log(f"device copy: {copy_gbs:.0f} GB/s")
    fp64 = Run(case, _lib.FP64, keep_checkpoints=True)
    r64 = fp64.timed(args.steps, args.warmup, args.chunk)
    log(f"fp64: {r64['ms_per_step']:.3f} ms/step, bond kernel {r64['kernel_ms']:.3f} ms")
    mixed = Run(case, _lib.MIXED)
    rm = mixed.timed(args.steps, args.warmup, args.chunk)
    log(f"mixed: {rm['ms_per_step']:.3f} ms/step, bond kernel {rm['kernel_ms']:.3f} ms")
    prec, samples = None, []
    if not args.no_precision:
        prec, samples = precision_study(case, fp64, mixed, args.max_study_steps, args.floor_samples)
        prec["protocol"] = protocol_text(args)
    # a state inside the timed window for the counter runs: the FP64 checkpoint nearest to the
    # window's middle, or none (the counter run then steps from 0 as the window did)
    mid = args.warmup + args.steps // 2
    cand = [c for c in fp64.ckpts if c[2]["step"] <= args.warmup + args.steps]
    state_ck = min(cand, key=lambda c: abs(c[2]["step"] - mid)) if cand else None
    gate_ok = bool(prec and prec["gate"]["passed"])
    head, head_name, head_run = (rm, "mixed", mixed) if gate_ok else (r64, "fp64", fp64)
    peak_inc = prec["peak_increment_fp64"] if prec else None
    state = window_label(head_run, args.warmup, args.steps, peak_inc)
    cpu_states = {}
    if samples:
        cpu_states["elastic"] = samples[1][1] if len(samples) > 2 else samples[0][1]
        cpu_states["post_peak"] = samples[-1][1]
    else:
        cpu_states["window_end"] = fp64.sol.checkpoint()
    traffic = None
    if not args.no_traffic:
        with tempfile.TemporaryDirectory() as d:
            path = None
            if state_ck is not None:
                path = os.path.join(d, "state.npz")
                save_state(path, state_ck[2])
            traffic = measure_counters(head_name, head["variant"], path, args.warmup, args.small)
        if traffic:
            traffic["state"] = (f"FP64 checkpoint at step {state_ck[2]['step']} restored" if state_ck
                                else f"steps {args.warmup + 1}..{args.warmup + 5} from the start")
            traffic["launch_variant"] = head["variant"]
            log(f"counters: {traffic}")
    fp64.close()
    mixed.close()
    cpu = None if args.no_cpu_baseline else cpu_baseline(args, cpu_states)
    achieved = head["bytes_layout"] / (head["kernel_ms"] * 1e-3) / 1e9
    model = csr_model_bytes(case, head_name == "mixed")
    out = {
        "metric": "bond updates/s (FP32-mixed when its peak load is within 0.5% of FP64, else FP64; FP64 beside) "
                  "and HBM roofline fraction of the bond kernel",
        "value": head["bond_updates_per_s"],
        "unit": "bond updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": head["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 bond arithmetic, f64 state" if head_name == "mixed" else "f64",
        "data": "synthetic beam lattice generated in-process",
        "config": {
            "workload": case.name,
            "nodes": case.nNodes,
            "bonds": case.nBonds,
            "law": "exponential (k=25, alpha=0.25)" if case.law == _lib.LAW_EXPONENTIAL else "trilinear",
            "integrator": "ADR, hipGraph-captured steps",
            "load": protocol_text(args),
            "state": state,
            "headline_form": head_name,
        },
        "fp64": {"value": r64["bond_updates_per_s"], "ms_per_step": r64["ms_per_step"], "kernel_ms": r64["kernel_ms"],
                 "bytes_layout": r64["bytes_layout"],
                 "frac": r64["bytes_layout"] / (r64["kernel_ms"] * 1e-3) / 1e9 / HBM_PEAK_GBS,
                 "retune_picks": r64["retune_picks"]},
        "mixed": {"value": rm["bond_updates_per_s"], "ms_per_step": rm["ms_per_step"], "kernel_ms": rm["kernel_ms"],
                  "bytes_layout": rm["bytes_layout"],
                  "frac": rm["bytes_layout"] / (rm["kernel_ms"] * 1e-3) / 1e9 / HBM_PEAK_GBS,
                  "retune_picks": rm["retune_picks"]},
        "mixed_over_fp64": rm["bond_updates_per_s"] / r64["bond_updates_per_s"],
        "mixed_vs_fp64_load_err": prec["peak_load_err"] if prec else None,
        "mixed_vs_fp64_same_state_load_err": prec["same_state_peak_load_err"] if prec else None,
        "precision": prec,
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic["bytes"] if traffic else None,
                     "kernel": f"bond kernel, {head_name}, stencil layout", "kernel_ms": head["kernel_ms"],
                     "bytes_per_launch": head["bytes_layout"],
                     "bytes_per_launch_note": "the stencil layout's algorithmic bytes at the window's states: node "
                                              "records, lane masks and nodal state per node, history only for "
                                              "(slice, offset) pairs whose history word is set",
                     "history_offsets_set_at_window_end": head["history_offsets_set"],
                     "traffic_fetch": traffic["fetch_bytes"] if traffic else None,
                     "traffic_write": traffic["write_bytes"] if traffic else None,
                     "traffic_state": traffic["state"] if traffic else None,
                     "traffic_gbs": traffic["bytes"] / (head["kernel_ms"] * 1e-3) / 1e9 if traffic else None,
                     "valu_busy": traffic.get("valu_busy") if traffic else None,
                     "limiter": "VALU instruction issue (valu_busy), not HBM bytes",
                     "launch_variant": head["variant"],
                     "copy_gbs": copy_gbs,
                     "frac_csr_model": model / (head["kernel_ms"] * 1e-3) / 1e9 / HBM_PEAK_GBS,
                     "bytes_csr_model": model},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
