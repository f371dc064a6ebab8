#!/usr/bin/env python3
"""Benchmark of the AeroFlux step data plane on one MI355X.

Default workload: 256^3 MAC grid, H=1/256, DT=0.001, nu=1e-5, seeded
Taylor-Green + U(+-1e-3) noise, FP64, the reference's default 24 red-black
sweep pairs per projection.  W warm-up Step()s, then K timed Step()s
bracketed by device synchronisation.  Prints ONE JSON line:

* value: cell-updates/s = NX*NY*NZ*K / wall seconds (bench_test.go:32);
* step_bandwidth: the schedule's own algorithmic bytes per cell per step
  (schedule_bytes) times the rate, and its fraction of the 8 TB/s peak;
  reference_layout_model: the same rate under the reference layout's byte
  model, an FP64-equivalent figure only;
* roofline: the dominant kernel (the red-black colour pass on the
  colour-split layout, 12 B/cell: other colour, divergence, own colour)
  timed with HIP events on the engine's stream;
* cpu_baseline: the C restatement of the Go reference timed on this host
  on a bounded sample of the same workload;
* mixed / mixed_to_target: the same workload with FP32 red-black smoothing
  inside FP64 defect correction, one cycle, and iterated to the FP64 run's
  final max divergence, with speed-up and per-field error against FP64;
* fp64_64, mixed_512, batched, served, direct_solver: the other workloads.
"""
from __future__ import annotations

import argparse
import ctypes
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBPS = 8000.0


def reference_layout_bytes(iters: int) -> int:
    """Algorithmic FP64 bytes per interior cell per Step on the reference's
    interleaved arena: 6 sweeps x (3 reads + 3 writes) x 8 B, plus per
    projection divergence 32 B, 2*iters colour passes of 24 B (interleaved
    colours move whole lines), gradient 56 B.  A model of the reference's
    layout, not of what this build moves (its colour-split passes move half
    of the 24 B), so its rate is an FP64-equivalent figure, not traffic."""
    return 288 + 2 * (88 + 48 * iters)


def schedule_bytes(kind: str, iters: int = 24) -> int:
    """Algorithmic bytes per interior cell per Step of the schedule this build
    runs (one read of each input and one write of each output per kernel;
    O(N^2) ghost fills not counted), for a single engine and for a batch alike:

    * advection: 6 sweeps x (read U, V, W + write Tmp0..2) = 288;
    * fp64_split, per projection: divergence 44 (U, V, W 24, Div 8, split
      right-hand side 8, and the solve's first red pass, which depends only
      on the right-hand side, 4), second pass 12, 2*iters-2 passes of 12,
      merge into P 16, gradient 56 = 104 + 24*iters; in the first
      projection Div is not stored (nothing reads it before the second
      projection rewrites it: 8 fewer); a single engine's gradients read the
      split solution, so the first projection skips the merge (16 fewer)
      and the second writes P from the gradient (8 fewer: the merge's read
      of the split solution); a batch keeps the merges (fp64_split_batch);
    * fp64_arena (small grids): divergence 32, first pass 16, 2*iters-1
      passes of 24, gradient 56 = 80 + 48*iters;
    * mixed (one FP32 cycle), per projection: divergence 38 (FP32 rhs 4,
      first FP32 red pass 2), second pass 6, 2*iters-2 passes of 6, merge
      12, gradient 56 = 100 + 12*iters; the first projection stores no Div
      (8 fewer); a single engine's gradients read the FP32 correction
      instead of P (4 fewer each), the first projection skips the merge (12
      fewer) and the second writes P from the gradient instead (4 fewer: the
      merge's read of the correction); a batch keeps the merges
      (mixed_batch);
    * mixed_checked: the one-cycle schedule with the second projection's
      first cycle checked by the FP64 residual of p = e: a single engine's
      check reads the correction and Div (12 more); a batch's merge also
      reads Div (8 more, mixed_checked_batch); the first projection runs
      unchecked.  Later cycles, which run only when the check asks for them,
      are not counted (a lower bound);
    * direct (FFT passes): divergence folded into the first pass, five
      passes 80, gradient 56, per projection 32 + 80 + 56."""
    batch = kind.endswith("_batch")
    base = kind[:-len("_batch")] if batch else kind
    proj = {"fp64_split": 104 + 24 * iters, "fp64_arena": 80 + 48 * iters,
            "mixed": 100 + 12 * iters, "mixed_checked": 100 + 12 * iters,
            "direct": 32 + 80 + 56}[base]
    saved = {"fp64_split": 8 + (0 if batch else 16 + 8),
             "mixed": 8 + (0 if batch else (12 + 4) + (4 + 4)),
             "mixed_checked": 8 + (-8 if batch else (12 + 4) + (4 + 4) - 12)}.get(base, 0)
    return 288 + 2 * proj - saved


def bandwidth(value: float, bytes_per_cell: int, kind: str) -> dict:
    """Step-level bandwidth of a cell-updates/s rate under the schedule's
    own algorithmic bytes, and its fraction of the 8 TB/s HBM peak."""
    g = value * bytes_per_cell / 1e9
    return {"step_bandwidth": {"schedule": kind, "bytes_per_cell_step": bytes_per_cell,
                               "gbps": g, "frac_of_peak": g / HBM_PEAK_GBPS}}


def fp64_kind(e) -> str:
    return "fp64_split" if e.solver_tile()[0] == 0 else "fp64_arena"


def hip_sync() -> None:
    rc = ctypes.CDLL("libamdhip64.so").hipDeviceSynchronize()
    if rc != 0:
        raise RuntimeError(f"hipDeviceSynchronize failed: {rc}")


def load_ic(e, f) -> None:
    for fid, name in ((1, "U"), (2, "V"), (3, "W")):
        e.Mesh.SetData(e.Mesh.field(fid), f[name])
    e.apply_velocity_bc()


def traffic_from_profile(name: str):
    """HBM bytes per launch measured by rocprofv3 --pmc (FETCH_SIZE and
    WRITE_SIZE, scaled by a copy kernel of known size) for this workload,
    read from the committed profile summary when present."""
    path = os.path.join(ROOT, "profiles", name)
    if not os.path.exists(path):
        return None
    with open(path) as fh:
        d = json.load(fh)
    return d.get("traffic_bytes_per_launch")


def dominant_kernel(e, cells: int):
    """Name and algorithmic bytes per launch of the solve's red-black kernel.

    Colour-split layout (default): one colour pass reads the other colour's
    N/2 values and this colour's N/2 divergences and writes N/2 values:
    3 * 8 * N/2 = 12 B per cell.  Arena layout: interleaved colours force
    whole-line traffic, read P + read Div + write P = 24 B per cell per pass.
    """
    k, _, _ = e.solver_tile()
    if k == 0:
        return "k_split_pass<double,0> (one colour pass, colour-split layout)", 12 * cells
    if k < 0:
        return "k_rbgs<true> (one colour pass, arena layout)", 24 * cells
    return f"k_rbgs_tb<double,{k}> ({k} passes per launch, arena layout)", 24 * k * cells


def profile_kernels(af, ic, n: int, reps: int) -> None:
    """Launch sequence for rocprofv3 counter runs: a plain device-to-device
    copy of a known size (the calibration kernel), then `reps` launches of
    the dominant red-black kernel on a warmed 256^3 engine."""
    hip = ctypes.CDLL("libamdhip64.so")
    nbytes = 512 << 20
    a, b = ctypes.c_void_p(), ctypes.c_void_p()
    for p in (a, b):
        if hip.hipMalloc(ctypes.byref(p), ctypes.c_size_t(nbytes)) != 0:
            raise RuntimeError("hipMalloc failed")
    if hip.hipMemset(a, 1, ctypes.c_size_t(nbytes)) != 0:
        raise RuntimeError("hipMemset failed")
    for _ in range(3):
        if hip.hipMemcpy(b, a, ctypes.c_size_t(nbytes), 3) != 0:  # hipMemcpyDeviceToDevice
            raise RuntimeError("hipMemcpy failed")
    hip_sync()
    hip.hipFree(a)
    hip.hipFree(b)
    # The colour-split layout pinned (TileK = -2): New picks the layout by
    # timing, and under a counter-collecting profiler that timing can come
    # out the other way, which would profile different kernels in the fetch
    # and write runs.  Unprofiled, New picks the split layout at this size.
    cfg = dict(NX=n, NY=n, NZ=n, H=1.0 / n, DT=0.001, Viscosity=1e-5, TileK=-2)
    e = af.New(af.Config(**cfg))
    load_ic(e, ic.taylor_green(n, n, n, 1, 1.0 / n, seed=1))
    e.StepN(2)
    t = e.time_kernel(0, reps)
    name, nb = dominant_kernel(e, n ** 3)
    t_sweep = e.time_kernel(1, reps)
    # The direct solve's five FFT passes (16 B/cell each), reps + 1 solves.
    e.EnableDirectSolver()
    t_direct = e.time_kernel(6, reps)
    method = e.direct_solver_method()
    for _ in range(reps + 1):  # projections: the first pass also computes the divergence
        e.project()
    e.Close()
    print(json.dumps({"profile_kernels": name, "copy_bytes": nbytes, "copies": 3,
                      "kernel_launches": reps + 1, "avg_ms": t,
                      "algorithmic_bytes_per_launch": nb,
                      "sweep_x_launches": reps + 1, "sweep_x_avg_ms": t_sweep,
                      "sweep_algorithmic_bytes_per_launch": 48 * n ** 3,
                      "direct_solves": reps + 1, "direct_solve_avg_ms": t_direct,
                      "direct_projections": reps + 1,
                      "direct_method": method,
                      "dct_pass_algorithmic_bytes_per_launch": 16 * n ** 3}), flush=True)


def profile_batch_pass(af, n_eng: int, reps: int) -> None:
    """Launch sequence for rocprofv3 counter runs of the batched colour
    pass: a 512 MiB copy (af_measure_stream kind 1, the calibration kernel,
    4 launches), then on the 96 uneven engines, in FP64 and in mixed
    precision, one batched step and one colour pass over every slot timed as
    the step runs it: group by group, 1 + reps launches per solve group."""
    from aeroflux_amd import workload

    out = {"profile_batch_pass": True, "copy_bytes": 512 << 20,
           "copy_gbps": af.MeasureStream(1, 64 << 20, 3), "passes": []}
    for prec in (af.PRECISION_FP64, af.PRECISION_MIXED):
        cfgs = [dict(c, Precision=prec) for c in workload.batch_engines(n_eng, seed=3)]
        need = sum(af.SlotBytes(af.Config(**c), prec) for c in cfgs)
        b = af.Batch(af.BatchConfig(Precision=prec, MemoryBudget=need + (64 << 20),
                                    MaxSlots=n_eng))
        for c in cfgs:
            b.Add(af.Config(**c))
        b.StepN(1)
        info = b.info()
        t = b.time_stage(0, reps)
        cells = sum(c["NX"] * c["NY"] * c["NZ"] for c in cfgs)
        per = 12 if prec == af.PRECISION_FP64 else 6
        out["passes"].append({"precision": "fp64" if prec == af.PRECISION_FP64 else "mixed",
                              "cells": cells, "solve_groups": info["SolveGroups"],
                              "launches_per_group": reps + 1, "pass_ms": t,
                              "algorithmic_bytes": per * cells})
        b.Close()
    print(json.dumps(out), flush=True)


def pct(a, q) -> float:
    return float(np.percentile(np.asarray(a, dtype=np.float64), q)) if len(a) else 0.0


def batched_leg(af, ic, n_eng: int, steps: int, warmup: int, with_plain: bool = True) -> dict:
    """Batched uneven engines: n_eng engines with each axis drawn from
    {32,48,64,96,128}, H = 1/max(dims), nu log-uniform in [1e-6,1e-4],
    DT = min(0.004, 0.5 H^2/(6 nu)), seeded Taylor-Green + noise per engine.
    The same engines are stepped by
      * the plain FP64 GPU path: each engine alone (New, StepN), one after
        the other;
      * an FP64 batch (bit-identical to the plain path);
      * a one-cycle mixed batch (the same 24 sweep pairs in FP32);
      * a checked mixed batch whose step leaves at most that engine's final
        max|div| on the FP64 path: the second projection's FP64 residual
        check after 24 FP32 sweep pairs, then while it is missed the FP32
        iteration continued by 2 pairs, then FP64 defect-correction cycles
        of 2 pairs (4 checks at most).
    Each reports cell-updates/s and each engine's final max|div|.  Without
    the plain path (with_plain False, for profiling the batched kernels
    alone) the FP64 batch's final max|div| -- the same bits -- sets the
    targets, and no speed-up against the plain path is reported."""
    from aeroflux_amd import workload

    engines = workload.batch_engines(n_eng, seed=3)
    fields = [ic.taylor_green(e["NX"], e["NY"], e["NZ"], 1, e["H"], seed=100 + i)
              for i, e in enumerate(engines)]
    cells = sum(e["NX"] * e["NY"] * e["NZ"] for e in engines)
    out = {"engines": n_eng, "cells": cells, "steps": steps, "warmup": warmup}

    # Plain FP64 path: device time of StepN only (creation and uploads are
    # outside the timed region; one synchronisation per engine).
    el_plain, d_plain = 0.0, []
    for e, f in zip(engines, fields) if with_plain else ():
        eng = af.New(af.Config(**e))
        load_ic(eng, f)
        if warmup:
            eng.StepN(warmup)
        hip_sync()
        t0 = time.perf_counter()
        eng.StepN(steps)
        hip_sync()
        el_plain += time.perf_counter() - t0
        d_plain.append(eng.MaxDivergence())
        eng.Close()
    d_plain = np.array(d_plain)
    plain = cells * steps / el_plain if with_plain else None
    if with_plain:
        out["plain_fp64"] = {"value": plain, "unit": "cell-updates/s",
                             "ms_per_step": 1e3 * el_plain / steps,
                             "how": "each engine alone through New/StepN, one after the other"}

    def run(prec, bc_extra, cfg_extra):
        need = sum(af.SlotBytes(af.Config(**e, Precision=prec, **cfg_extra(i)), prec)
                   for i, e in enumerate(engines))
        b = af.Batch(af.BatchConfig(Precision=prec, MemoryBudget=need + (64 << 20),
                                    MaxSlots=max(n_eng, 8), **bc_extra))
        slots = []
        for i, (e, f) in enumerate(zip(engines, fields)):
            s = b.Add(af.Config(**e, Precision=prec, **cfg_extra(i)))
            for fid, key in ((1, "U"), (2, "V"), (3, "W")):
                b.SetData(s, fid, f[key])
            b.apply_velocity_bc(s)
            slots.append(s)
        b.Tune()
        if warmup:
            b.StepN(warmup)
        hip_sync()
        t0 = time.perf_counter()
        b.StepN(steps)
        hip_sync()
        el = time.perf_counter() - t0
        d = np.array([b.MaxDivergence(s) for s in slots])
        cyc = ([b.mixed_stats(s)[1] for s in slots] if prec == af.PRECISION_MIXED else None)
        t_pass = b.time_stage(0, 20)
        # Every stage of the step, timed alone over all slots (the divergence
        # and gradient last: they change Div and the velocity).
        sw = [b.time_stage(k, 5) for k in (1, 5, 6)]
        t_vbc, t_pbc, t_merge = b.time_stage(4, 10), b.time_stage(7, 10), b.time_stage(8, 10)
        t_res = b.time_stage(9, 10) if prec == af.PRECISION_MIXED else 0.0
        t_div, t_grad = b.time_stage(2, 10), b.time_stage(3, 10)
        # The divergence writes the first red pass; the solve launches the rest.
        npass = 2 * (bc_extra.get("MixedSweeps", 24) if prec == af.PRECISION_MIXED else 24) - 1
        bkind = ("fp64_split_batch" if prec == af.PRECISION_FP64 else
                 ("mixed_checked_batch" if bc_extra.get("MixedChecked") else "mixed_batch"))
        model = t_vbc + 2 * sum(sw) + 2 * (t_div + npass * t_pass + t_merge + t_pbc + t_grad)
        info = b.info()
        pass_bytes = (6 if prec == af.PRECISION_MIXED else 12) * cells
        res = {
            "value": cells * steps / el,
            "unit": "cell-updates/s",
            "ms_per_step": 1e3 * el / steps,
            "speedup_vs_plain_fp64": cells * steps / el / plain if plain else None,
            **bandwidth(cells * steps / el,
                        schedule_bytes(bkind), bkind),
            "pass_roofline": {"bound": "hbm", "achieved": pass_bytes / (t_pass * 1e-3) / 1e9,
                              "peak": HBM_PEAK_GBPS, "unit": "GB/s",
                              "frac": pass_bytes / (t_pass * 1e-3) / 1e9 / HBM_PEAK_GBPS,
                              "algorithmic_bytes_per_launch": pass_bytes,
                              "launches_per_pass": info["SolveGroups"],
                              "note": "one colour pass over every slot; with several solve "
                                      "groups it runs as one launch per group, each group's "
                                      "passes back to back, and the time is their sum"},
            "stages": {"rbgs_pass_ms": t_pass, "advect_sweep_ms": sw[0],
                       "advect_sweep_xyz_ms": sw, "velocity_bc_ms": t_vbc,
                       "divergence_ms": t_div, "gradient_ms": t_grad, "merge_ms": t_merge,
                       "pressure_bc_ms": t_pbc, "residual_ms": t_res,
                       "one_cycle_step_model_ms": model,
                       "model": "1 velocity BC + 2 x (X+Y+Z sweeps, which apply the BC to "
                                "their output) + 2 x (divergence, which writes the first red pass, + "
                                f"{npass} passes + merge + pressure BC + gradient, which applies the BC)"},
            "tuned": {"sweep_shape": info["SweepShape"], "sweep_staged": info["SweepStaged"],
                      "staged_slots": info["StagedSlots"],
                      "workgroups": [info["SweepWorkgroups"], info["PassWorkgroups"]],
                      "solve_groups": info["SolveGroups"]},
            "pool_bytes": info["BytesInUse"],
        }
        if cyc is not None:
            res["cycles_last_solve"] = {"min": int(min(cyc)), "max": int(max(cyc))}
        b.Close()
        if ref is not None:
            rel = (d - ref) / ref
            res["final_max_divergence_vs_plain_fp64"] = {
                "rel_diff_max_abs": float(np.abs(rel).max()), "rel_diff_mean": float(rel.mean()),
                "engines_at_or_below": int((d <= ref).sum())}
        return res, d

    none = lambda i: {}
    ref = d_plain if with_plain else None
    out["fp64"], d64 = run(af.PRECISION_FP64, {}, none)
    if with_plain:
        out["fp64"]["final_max_divergence_equal_to_plain"] = bool(np.array_equal(d64, d_plain))
    else:
        d_plain = ref = d64  # the FP64 batch gives the plain path's bits
    out["mixed"], _ = run(af.PRECISION_MIXED, {}, none)
    out["mixed"]["schedule"] = "24 FP32 sweep pairs in one FP64 defect-correction cycle"
    out["mixed_to_target"], _ = run(
        af.PRECISION_MIXED,
        {"MixedChecked": 1, "MixedSweeps": 24, "MixedFixSweeps": 2, "MixedMaxCycles": 4},
        lambda i: {"MixedSweeps": 24, "MixedFixSweeps": 2, "MixedMaxCycles": 4,
                   "MixedTarget": float(d_plain[i])})
    out["mixed_to_target"]["schedule"] = (
        "24 FP32 sweep pairs per projection; the second projection's FP64 residual check, then "
        "while max|div| > the engine's final max|div| on the plain FP64 path: the FP32 "
        "iteration continued by 2 pairs, then FP64 defect-correction cycles of 2 pairs (at "
        "most 4 checks)")
    out["speedup_mixed_vs_fp64_batch"] = out["mixed"]["value"] / out["fp64"]["value"]
    return out


def big_leg(af, n: int, steps: int) -> dict:
    """Mixed-precision projection at n^3 (H=1/n, DT=0.0005, nu=1e-5), seeded
    band-limited initial velocity (|k| <= 8, |k|^-5/3, RMS 1, built on the
    device), against the FP64 path -- which is bit-exact to the oracle -- on
    the same initial condition: cell-updates/s of both, final max|div| of
    both, and each output field's max and RMS error relative to FP64."""
    cfg = dict(NX=n, NY=n, NZ=n, H=1.0 / n, DT=0.0005, Viscosity=1e-5)
    out = {"grid": [n, n, n], "steps": steps, "initial_condition": "band-limited seed 5"}
    fields = {}
    runs = [("fp64", {}), ("mixed", {"Precision": af.PRECISION_MIXED}), ("mixed_to_target", None)]
    for name, extra in runs:
        if extra is None:
            # Iterated to the FP64 run's final max|div|: 24 FP32 sweep pairs,
            # the second projection's FP64 residual check, then at most one
            # pair of the FP32 iteration continued and one more check.
            extra = {"Precision": af.PRECISION_MIXED, "MixedSweeps": 24, "MixedFixSweeps": 1,
                     "MixedMaxCycles": 2,
                     "MixedTarget": out["fp64"]["final_max_divergence"]}
        e = af.New(af.Config(**cfg, **extra))
        e.init_band_limited(5)
        e.StepN(1)
        hip_sync()
        t0 = time.perf_counter()
        e.StepN(steps - 1)
        hip_sync()
        el = time.perf_counter() - t0
        out[name] = {"value": n ** 3 * (steps - 1) / el, "unit": "cell-updates/s",
                     "ms_per_step": 1e3 * el / (steps - 1),
                     "final_max_divergence": e.MaxDivergence()}
        kind = {"fp64": "fp64_split", "mixed": "mixed"}.get(name, "mixed_checked")
        out[name].update(bandwidth(out[name]["value"], schedule_bytes(kind), kind))
        if name == "mixed_to_target":
            out[name]["cycles_last_solve"] = e.mixed_stats()[1]
        fields[name] = {k: e.Mesh.Data(e.Mesh.field(i)) for i, k in enumerate("PUVW")}
        e.Close()
        if name == "fp64":
            continue
        err = {}
        for k in "PUVW":
            a, r = fields[name][k], fields["fp64"][k]
            d = a - r
            err[k] = {"max_rel": float(np.abs(d).max() / np.abs(r).max()),
                      "rms_rel": float(np.sqrt(np.mean(d * d)) / np.sqrt(np.mean(r * r)))}
        del fields[name]
        out[name]["error_vs_fp64"] = err
        out[name]["speedup_vs_fp64"] = out[name]["value"] / out["fp64"]["value"]
        d64, dmx = out["fp64"]["final_max_divergence"], out[name]["final_max_divergence"]
        out[name]["final_max_divergence_rel_diff"] = (dmx - d64) / d64
    out["mixed_to_target"]["error_note"] = (
        "the correction cycles carry the iterate past the FP64 path's 24 unconverged sweep "
        "pairs, so these errors are that path's own remainder, not FP32 rounding (the one-cycle "
        "run's errors are the rounding)")
    # The same run with the direct pressure solve (FFT passes at 512^3): an
    # exact projection, so its fields differ from the 24-sweep solve's by the
    # iteration's error; only rates and the final max|div| are compared.
    e = af.New(af.Config(**cfg))
    e.init_band_limited(5)
    e.EnableDirectSolver()
    e.StepN(1)
    hip_sync()
    t0 = time.perf_counter()
    e.StepN(steps - 1)
    hip_sync()
    el = time.perf_counter() - t0
    dv = n ** 3 * (steps - 1) / el
    out["direct"] = {"value": dv, "unit": "cell-updates/s", "ms_per_step": 1e3 * el / (steps - 1),
                     "speedup_vs_fp64": dv / out["fp64"]["value"],
                     "final_max_divergence": e.MaxDivergence(),
                     "transforms": {af.DIRECT_FFT: "FFT passes",
                                    af.DIRECT_DENSE: "dense GEMMs"}[e.direct_solver_method()],
                     "pressure_solve_ms": e.time_kernel(6, 3)}
    e.Close()
    out["timing"] = "steps 2..N timed (the first step also builds the step graph)"
    return out


def small_leg(af, ic, n: int, steps: int, workers: int, cpu: bool) -> dict:
    """The 64^3 FP64 workload (H=1/64, DT=0.004, nu=1e-5, seeded Taylor-Green
    + noise), 100 steps on the GPU and, unless skipped, the same 100 steps
    on the C oracle (Z-slab pool of `workers` threads): both rates, and
    whether all eight fields end bit-identical.  The 19 MB arena stays in the
    last-level cache, so its bandwidth is not an HBM figure."""
    h = 1.0 / n
    cfg = dict(NX=n, NY=n, NZ=n, H=h, DT=0.004, Viscosity=1e-5)
    f = ic.taylor_green(n, n, n, 1, h, seed=1)
    e, el = run_gpu(af, cfg, f, 0, steps)
    gpu = [e.Mesh.Data(e.Mesh.field(k)) for k in range(8)]
    kind = fp64_kind(e)
    e.Close()
    value = n ** 3 * steps / el
    out = {"grid": [n, n, n], "steps": steps, "value": value, "unit": "cell-updates/s",
           "ms_per_step": 1e3 * el / steps, "cache_resident": True,
           **bandwidth(value, schedule_bytes(kind), kind)}
    if cpu:
        from oracle import oracle
        o = oracle.OracleEngine(dict(cfg, Workers=workers))
        for fid, name in ((1, "U"), (2, "V"), (3, "W")):
            o.field(fid)[...] = f[name]
        o.apply_velocity_bc()
        t0 = time.perf_counter()
        for _ in range(steps):
            if o.step() != 0:
                raise RuntimeError("oracle step faulted")
        cel = time.perf_counter() - t0
        same = all(np.array_equal(gpu[k].view(np.uint64).ravel(),
                                  o.field(k).view(np.uint64).ravel()) for k in range(8))
        o.close()
        out["cpu_baseline"] = {"value": n ** 3 * steps / cel, "unit": "cell-updates/s",
                               "cores": workers, "kind": "port",
                               "sample": f"all {steps} steps, same initial condition"}
        out["bit_identical_to_oracle"] = same
    return out


def serve_summary(jobs, res, st, kind: str = "mixed_batch") -> dict:
    ok = [r for r in res if r["Status"] == 0]
    lat = [r["Done"] - j["Arrival"] for j, r in zip(jobs, res) if r["Status"] == 0]
    que = [r["Admitted"] - j["Arrival"] for j, r in zip(jobs, res) if r["Status"] == 0]
    svc = [r["Done"] - r["Admitted"] for r in ok]
    return {
        "completed": len(ok),
        "faulted_or_refused": len(jobs) - len(ok),
        "wall_s": st["Wall"],
        "drain_s": st["Wall"] - (jobs[-1]["Arrival"] if jobs else 0.0),
        "value": st["CellUpdates"] / st["Wall"] if st["Wall"] > 0 else 0.0,
        "unit": "cell-updates/s",
        "latency_s": {"p50": pct(lat, 50), "p99": pct(lat, 99), "max": max(lat, default=0.0)},
        "queueing_s": {"p50": pct(que, 50), "p99": pct(que, 99)},
        "service_s": {"p50": pct(svc, 50), "p99": pct(svc, 99)},
        "batched_steps": st["Steps"],
        "max_resident_jobs": st["MaxResident"],
        "peak_pool_bytes": st["PeakBytes"],
        "preemptions": st["Preemptions"],
        **bandwidth(st["CellUpdates"] / st["Wall"] if st["Wall"] > 0 else 0.0,
                    schedule_bytes(kind), kind),
    }


# Serving policies: name -> (ServePolicy, description, output key).
SERVE_POLICIES = {
    "fcfs": (0, "first come first served; every resident job advances each batched step",
             "first_come_first_served"),
    "steps": (2, "admission first come first served; each batched step advances the resident "
                 "jobs with the least remaining work (cells x steps left), up to 32 Mi "
                 "interior cells", "least_work_stepping"),
    "least": (3, "least remaining work (cells x steps left) admitted first, without "
                 "preemption, and stepped first, up to 32 Mi interior cells per batched step",
              "least_work_first"),
    "preempt": (1, "least remaining work (cells x steps left) admitted first; a resident job "
                   "with more than 4x the work of a queued job that cannot be placed is "
                   "snapshotted to host memory and requeued (at most once per job)",
                "least_work_first_preempting"),
}


def serve_once(af, jobs, budget: float, policy: int, label: str, checked: bool = False):
    extra = {"MixedChecked": 1, "MixedSweeps": 24, "MixedFixSweeps": 2,
             "MixedMaxCycles": 4} if checked else {}
    b = af.Batch(af.BatchConfig(Precision=af.PRECISION_MIXED, MemoryBudget=int(budget),
                                MaxSlots=4096, ServePolicy=policy, **extra))

    def report(so_far, done):
        print(f"served ({label}): {so_far['Wall']:.0f} s, {done}/{len(jobs)} jobs done, "
              f"{so_far['Steps']} batched steps", file=sys.stderr, flush=True)

    res, st = af.Serve(b, jobs, realtime=True, progress=report, progress_interval=30.0)
    info = b.info()
    b.Close()
    return res, st, info


def served_leg(af, n_jobs: int, rate: float, budget: float, n_plain: int,
               policies: str) -> dict:
    """Served step-job stream: n_jobs jobs arriving as a Poisson process at
    `rate`/s, axis sizes from {32,64,96,128,192,256}, N uniform in [10,200],
    nu log-uniform in [1e-6,1e-4], DT = min(0.004, 0.5 H^2/(6 nu)), seeded
    band-limited initial velocity (|k| <= 8, |k|^-5/3, RMS 1) built on the
    device at admission, mixed-precision projection (one cycle of the same
    24 sweep pairs in FP32), arenas admitted first come first served under
    the HBM budget.  The first n_plain jobs are also run on the plain FP64
    path, each alone (New, init_band_limited, StepN; StepN timed), for its
    cell-updates/s and each job's final max|div|, against which every served
    job's is compared; then those n_plain jobs are served again in a checked
    batch, each iterated to its own FP64 final max|div| (24 FP32 pairs, the
    second projection's FP64 residual check, then the FP32 iteration
    continued by 2 pairs and FP64 defect-correction cycles of 2 pairs, at
    most 4 checks).

    `policies` is a comma list of least (least remaining work admitted
    first and stepped first, up to the step cell budget), fcfs (first come
    first served, every resident job advances each batched step), steps
    (admission as fcfs; each batched step advances the resident jobs with the
    least remaining work, up to the step cell budget) and preempt (least
    remaining work admitted first, preempting through host snapshots); the
    first is the main run, the others run the same stream again and compare
    their results with it."""
    from aeroflux_amd import workload

    names = [p for p in policies.split(",") if p]
    jobs = workload.served_jobs(n_jobs, seed=7, rate=rate)
    first = SERVE_POLICIES[names[0]][0]
    label = SERVE_POLICIES[names[0]][1]
    res, st, info = serve_once(af, jobs, budget, first, label)
    plain = None
    if n_plain > 0:
        el, cu, rel, d_alone = 0.0, 0, [], []
        for j, r in list(zip(jobs, res))[:n_plain]:
            e = af.New(af.Config(NX=j["NX"], NY=j["NY"], NZ=j["NZ"], H=j["H"], DT=j["DT"],
                                 Viscosity=j["Viscosity"]))
            e.init_band_limited(j["Seed"])
            hip_sync()
            t0 = time.perf_counter()
            e.StepN(j["Steps"])
            hip_sync()
            el += time.perf_counter() - t0
            cu += j["NX"] * j["NY"] * j["NZ"] * j["Steps"]
            d = e.MaxDivergence()
            e.Close()
            d_alone.append(d)
            if r["Status"] == 0:
                rel.append((r["MaxDivergence"] - d) / d)
        rel = np.array(rel)
        plain = {"jobs": n_plain, "value": cu / el, "unit": "cell-updates/s",
                 "how": "each job alone through New/init_band_limited/StepN, one after the other",
                 "served_mixed_final_max_divergence_rel_diff": {
                     "max_abs": float(np.abs(rel).max()) if rel.size else 0.0,
                     "mean": float(rel.mean()) if rel.size else 0.0,
                     "jobs_at_or_below_fp64": int((rel <= 0).sum()),
                     "jobs_compared": int(rel.size)}}
        # The same jobs served again, each iterated to its own FP64 final
        # max|div| (a checked batch with per-job targets).
        cjobs = [dict(j, MixedTarget=d) for j, d in zip(jobs[:n_plain], d_alone)]
        cres, cst, _ = serve_once(af, cjobs, budget, af.SERVE_FCFS,
                                  "checked, first come first served", checked=True)
        crel = np.array([(r["MaxDivergence"] - j["MixedTarget"]) / j["MixedTarget"]
                         for j, r in zip(cjobs, cres) if r["Status"] == 0])
        plain["served_checked"] = {
            "jobs": n_plain,
            "schedule": "24 FP32 sweep pairs per projection; the second projection's FP64 "
                        "residual check, then while max|div| > the job's final max|div| on the "
                        "FP64 path: the FP32 iteration continued by 2 pairs, then FP64 "
                        "defect-correction cycles of 2 pairs (at most 4 checks)",
            **serve_summary(cjobs, cres, cst, "mixed_checked_batch"),
            "speedup_vs_plain_fp64": (cst["CellUpdates"] / cst["Wall"] / plain["value"]
                                      if cst["Wall"] > 0 else None),
            "final_max_divergence_vs_fp64": {
                "jobs_at_or_below_fp64": int((crel <= 1e-9).sum()),
                "jobs_compared": int(crel.size),
                "rel_diff_max": float(crel.max()) if crel.size else 0.0,
                "rel_diff_mean": float(crel.mean()) if crel.size else 0.0}}
    demand = sum(j["NX"] * j["NY"] * j["NZ"] * j["Steps"] for j in jobs)
    out = {
        "jobs": n_jobs,
        "arrival_rate_per_s": rate,
        "last_arrival_s": jobs[-1]["Arrival"] if jobs else 0.0,
        "offered_cell_updates_per_s": demand / jobs[-1]["Arrival"] if jobs else 0.0,
        "policy": label,
        "schedule": "24 FP32 sweep pairs per projection in one FP64 defect-correction cycle",
        **serve_summary(jobs, res, st),
        "memory_budget_bytes": int(budget),
        "tuned": {"sweep_shape": info["SweepShape"], "sweep_staged": info["SweepStaged"],
                  "workgroups": [info["SweepWorkgroups"], info["PassWorkgroups"]]},
        "plain_fp64_sample": plain,
        "speedup_vs_plain_fp64": (st["CellUpdates"] / st["Wall"] / plain["value"]
                                  if plain and st["Wall"] > 0 else None),
    }
    for name in names[1:]:
        pol, plabel, key = SERVE_POLICIES[name]
        pres, pst, _ = serve_once(af, jobs, budget, pol, plabel)
        same = all((a["MaxDivergence"], a["KineticEnergy"], a["Status"]) ==
                   (c["MaxDivergence"], c["KineticEnergy"], c["Status"])
                   for a, c in zip(res, pres))
        out[key] = {
            "policy": plabel,
            **serve_summary(jobs, pres, pst),
            "speedup_vs_plain_fp64": (pst["CellUpdates"] / pst["Wall"] / plain["value"]
                                      if plain and pst["Wall"] > 0 else None),
            "results_identical_to_main_run": same,
        }
    return out


def cpu_baseline(cfg: dict, f: dict, seconds: float, workers: int) -> dict:
    from oracle import oracle

    ocfg = dict(cfg, Workers=workers)
    o = oracle.OracleEngine(ocfg)
    for fid, name in ((1, "U"), (2, "V"), (3, "W")):
        o.field(fid)[...] = f[name]
    o.apply_velocity_bc()
    steps = 0
    t0 = time.perf_counter()
    while True:
        if o.step() != 0:
            raise RuntimeError("oracle step faulted")
        steps += 1
        el = time.perf_counter() - t0
        if el >= seconds or steps >= 50:
            break
    o.close()
    cells = cfg["NX"] * cfg["NY"] * cfg["NZ"]
    return {
        "value": cells * steps / el,
        "unit": "cell-updates/s",
        "cores": workers,
        "kind": "port",
        "sample": f"{steps} Step() calls of the same {cfg['NX']}x{cfg['NY']}x{cfg['NZ']} workload "
                  f"(first steps from the same initial condition), {el:.1f} s, "
                  f"C restatement of stencil.go/project.go, Z-slab pool of {workers} threads",
    }


def run_gpu(af, cfg: dict, f: dict, warmup: int, steps: int, direct: bool = False):
    e = af.New(af.Config(**cfg))
    load_ic(e, f)
    if direct:
        e.EnableDirectSolver()
    if warmup:
        e.StepN(warmup)
    hip_sync()
    t0 = time.perf_counter()
    e.StepN(steps)
    hip_sync()
    el = time.perf_counter() - t0
    return e, el


def main() -> int:
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--n", type=int, default=256, help="grid edge (default 256)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-mixed", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=10.0)
    ap.add_argument("--cpu-workers", type=int, default=int(os.environ.get("AF_CPU_WORKERS", "0")),
                    help="oracle threads for the CPU baseline (0: the host cores this process "
                         "may use, capped at 16, the CPU share of one GPU)")
    ap.add_argument("--profile-kernels", action="store_true",
                    help="only run the copy-calibration + dominant-kernel launch sequence")
    ap.add_argument("--profile-batch-pass", action="store_true",
                    help="only run the copy-calibration + batched colour pass launch sequence")
    ap.add_argument("--no-batched", action="store_true")
    ap.add_argument("--batched-only", action="store_true",
                    help="only the batched leg, printed as its own JSON object (for profiling "
                         "the batched kernels; with --batched-no-plain no engine steps alone)")
    ap.add_argument("--batched-no-plain", action="store_true")
    ap.add_argument("--batched-engines", type=int, default=96)
    ap.add_argument("--batched-steps", type=int, default=50)
    ap.add_argument("--target-sweeps", type=int, default=24,
                    help="FP32 sweep pairs of the first cycle of the iterated-to-target run")
    ap.add_argument("--no-direct", action="store_true")
    ap.add_argument("--no-small", action="store_true", help="skip the 64^3 leg")
    ap.add_argument("--no-big", action="store_true")
    ap.add_argument("--big-n", type=int, default=512)
    ap.add_argument("--big-steps", type=int, default=100)
    ap.add_argument("--no-served", action="store_true")
    ap.add_argument("--served-jobs", type=int, default=400,
                    help="jobs in the served stream (the full stream is 10000)")
    ap.add_argument("--served-rate", type=float, default=200.0)
    ap.add_argument("--served-budget", type=float, default=200e9)
    ap.add_argument("--served-policy", default="least,fcfs",
                    help="comma list of serving policies, the first the main run: least "
                         "(least remaining work admitted and stepped first), fcfs (lock step), "
                         "steps (first come first served admission, least work stepped "
                         "first), preempt (least work admitted first, preempting)")
    ap.add_argument("--served-plain-jobs", type=int, default=60,
                    help="jobs of the stream also run alone on the plain FP64 path")
    args = ap.parse_args()
    if args.cpu_workers <= 0:
        args.cpu_workers = min(16, len(os.sched_getaffinity(0)))

    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print("bench.py: this build measures one GPU only; the step data plane is a single-GPU "
              "path here (multi-GPU sharding is a design note, not a measured mode).",
              file=sys.stderr)
        return 2

    import aeroflux_amd as af
    from aeroflux_amd import ic

    af.load_library()
    n = args.n
    if args.profile_kernels:
        profile_kernels(af, ic, n, 10)
        return 0
    if args.profile_batch_pass:
        profile_batch_pass(af, args.batched_engines, 5)
        return 0
    if args.batched_only:
        print(json.dumps({"batched": batched_leg(af, ic, args.batched_engines,
                                                 args.batched_steps, 2,
                                                 with_plain=not args.batched_no_plain)}),
              flush=True)
        return 0
    h = 1.0 / n
    cfg = dict(NX=n, NY=n, NZ=n, H=h, DT=0.001, Viscosity=1e-5)
    iters = 24
    cells = n ** 3
    f = ic.taylor_green(n, n, n, 1, h, seed=1)

    e, el = run_gpu(af, cfg, f, args.warmup, args.steps)
    d64 = e.MaxDivergence()
    ms_per_step = 1e3 * el / args.steps
    value = cells * args.steps / el
    # The final state, before the kernel timings below overwrite it.
    fp64_fields = None
    if not args.no_mixed:
        fp64_fields = {k: e.Mesh.Data(e.Mesh.field(i)) for i, k in enumerate("PUVW")}

    # Dominant kernel: one launch of the red-black kernel the solve runs.
    reps = 40
    t_rb = e.time_kernel(0, reps)
    rb_name, rb_bytes = dominant_kernel(e, cells)
    achieved = rb_bytes / (t_rb * 1e-3) / 1e9
    kernels = {
        "rbgs_pass_ms": t_rb,
        "advect_sweep_ms": e.time_kernel(1, 10),
        "divergence_ms": e.time_kernel(2, 10),
        "gradient_ms": e.time_kernel(3, 10),
        "pressure_solve_ms": e.time_kernel(6, 10),
    }
    # What a plain streaming kernel with the pass's pattern (two reads, one
    # write) and the gradient's (four reads, three writes) attain from HBM:
    # 512 MiB arrays, so the 256 MB last-level cache holds none of the
    # working set (the pass's own four split arrays, 268 MB at 256^3, partly
    # fit it).
    stream_n = 64 << 20
    stream_2r1w = af.MeasureStream(af.STREAM_2R1W, stream_n, 20)
    stream_4r3w = af.MeasureStream(af.STREAM_4R3W, stream_n, 10)
    tuned = {"sweep_shape": e.sweep_shape()[0] + 1, "sweep_slab_rows": e.sweep_slab(),
             "pass_slab_workgroup_rows": e.pass_slab()}
    kind = fp64_kind(e)
    e.Close()

    out = {
        "metric": "cell-updates/s & HBM GB/s at 256³ FP64 vs mixed; p50/p99 step-job latency",
        "value": value,
        "unit": "cell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic: seeded Taylor-Green + U(-1e-3,1e-3) noise (SplitMix64 seed 1)",
        "config": {
            "workload": f"plain FP64 step, {n}^3 MAC grid, H=1/{n}, DT=0.001, nu=1e-5, "
                        f"{iters} red-black sweep pairs, Strang X,Y,Z,[P],Z,Y,X,[P]",
            "grid": [n, n, n],
            "gauss_seidel_iters": iters,
            "precision": "fp64",
        },
        **bandwidth(value, schedule_bytes(kind, iters), kind),
        "reference_layout_model": {
            "bytes_per_cell_step": reference_layout_bytes(iters),
            "fp64_equivalent_gbps": value * reference_layout_bytes(iters) / 1e9,
            "note": "the reference's interleaved-arena byte model; this build's colour-split "
                    "passes move half its pass bytes, so this is not traffic and not a fraction "
                    "of peak"},
        "final_max_divergence": d64,
        "roofline": {
            "kernel": rb_name,
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBPS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBPS,
            "traffic": traffic_from_profile(f"split_pass_{n}_pmc.json"),
            "algorithmic_bytes_per_launch": rb_bytes,
            "avg_launch_ms": t_rb,
            "attainable_same_pattern": stream_2r1w,
            "frac_of_attainable": achieved / stream_2r1w,
        },
        "stream_gbps": {"2r1w_512MiB_arrays": stream_2r1w, "4r3w_512MiB_arrays": stream_4r3w},
        "kernels": kernels,
        "tuned": tuned,
    }

    if not args.no_small:
        out["fp64_64"] = small_leg(af, ic, 64, 100, args.cpu_workers, not args.no_cpu_baseline)

    if not args.no_mixed:
        # Same 24 sweep pairs, one FP32 cycle: the iterate is the FP64 one up
        # to FP32 rounding, so the final max|div| matches the FP64 run's to
        # that rounding (both reported).
        mcfg = dict(cfg, Precision=af.PRECISION_MIXED, MixedSweeps=iters, MixedTarget=0.0)
        m, mel = run_gpu(af, mcfg, f, args.warmup, args.steps)
        dmx = m.MaxDivergence()
        errs = {}
        for i, k in enumerate("PUVW"):
            a = m.Mesh.Data(m.Mesh.field(i))
            ref = fp64_fields[k]
            errs[k] = float(np.abs(a - ref).max() / max(np.abs(ref).max(), 1e-300))
        _, cyc = m.mixed_stats()
        m.Close()
        mval = cells * args.steps / mel
        out["mixed"] = {
            "value": mval,
            "unit": "cell-updates/s",
            "ms_per_step": 1e3 * mel / args.steps,
            "speedup_vs_fp64": mval / value,
            "final_max_divergence": dmx,
            "fp64_final_max_divergence": d64,
            "schedule": f"{iters} FP32 red-black sweep pairs (colour-split) in one FP64 "
                        "defect-correction cycle per projection",
            "final_max_divergence_rel_diff": (dmx - d64) / d64,
            "cycles_last_solve": cyc,
            "max_rel_error_vs_fp64": errs,
            **bandwidth(mval, schedule_bytes("mixed", iters), "mixed"),
        }
        # Iterated to the FP64 path's final max|div|: a short first cycle, then
        # correction cycles until the FP64 residual says the projection's
        # max|div| is at most that value.
        tcfg = dict(cfg, Precision=af.PRECISION_MIXED, MixedSweeps=args.target_sweeps,
                    MixedFixSweeps=1, MixedMaxCycles=2, MixedTarget=d64)
        t, tel = run_gpu(af, tcfg, f, args.warmup, args.steps)
        dt_ = t.MaxDivergence()
        terrs = {}
        for i, k in enumerate("PUVW"):
            a = t.Mesh.Data(t.Mesh.field(i))
            ref = fp64_fields[k]
            terrs[k] = float(np.abs(a - ref).max() / max(np.abs(ref).max(), 1e-300))
        _, tcyc = t.mixed_stats()
        t.Close()
        tval = cells * args.steps / tel
        out["mixed_to_target"] = {
            "value": tval,
            "unit": "cell-updates/s",
            "ms_per_step": 1e3 * tel / args.steps,
            "speedup_vs_fp64": tval / value,
            "target_max_divergence": d64,
            "final_max_divergence": dt_,
            "schedule": f"{args.target_sweeps} FP32 sweep pairs per projection; the second "
                        "projection's FP64 residual check (read-only), then at most one more "
                        "pair of the same FP32 iteration and one more check if max|div| is "
                        "still above the FP64 run's final value",
            "cycles_last_solve": tcyc,
            "max_rel_error_vs_fp64": terrs,
            **bandwidth(tval, schedule_bytes("mixed_checked", iters), "mixed_checked"),
        }
    del fp64_fields

    if not args.no_direct:
        # The same step with the exact (direct) pressure solve in place of
        # the 24 red-black sweep pairs: a different, exactly projected field.
        dsteps = args.steps
        de, delt = run_gpu(af, cfg, f, args.warmup, dsteps, direct=True)
        ddiv = de.MaxDivergence()
        method = de.direct_solver_method()
        t_solve = de.time_kernel(6, 10)
        t_dense = None
        if method == af.DIRECT_FFT:
            de.EnableDirectSolverMethod(af.DIRECT_DENSE)
            t_dense = de.time_kernel(6, 10)
        de.Close()
        dval = cells * dsteps / delt
        out["direct_solver"] = {
            "value": dval,
            "unit": "cell-updates/s",
            "steps": dsteps,
            "ms_per_step": 1e3 * delt / dsteps,
            "speedup_vs_fp64": dval / value,
            "pressure_solve_ms": t_solve,
            "dense_transform_pressure_solve_ms": t_dense,
            "fp64_pressure_solve_ms": kernels["pressure_solve_ms"],
            "transforms": {af.DIRECT_FFT: "FFT passes", af.DIRECT_DENSE: "dense GEMMs"}[method],
            # 6 sweeps 288 B + per projection divergence 32, solve 80 (FFT
            # passes), gradient 56.
            **bandwidth(dval, schedule_bytes("direct"), "direct"),
            "final_max_divergence": ddiv,
            "fp64_final_max_divergence": d64,
            "solver": "pinned Neumann system solved exactly: orthonormal DCT-II along each "
                      "axis (FFT passes for power-of-two extents, else FP64 GEMMs), spectral "
                      "division, inverse DCT",
        }

    if not args.no_big:
        out["mixed_512"] = big_leg(af, args.big_n, args.big_steps)
    if not args.no_batched:
        out["batched"] = batched_leg(af, ic, args.batched_engines, args.batched_steps, 2)
    if not args.no_served:
        out["served"] = served_leg(af, args.served_jobs, args.served_rate, args.served_budget,
                                   min(args.served_plain_jobs, args.served_jobs),
                                   args.served_policy)

    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(cfg, f, args.cpu_seconds, args.cpu_workers)

    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
