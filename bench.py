"""Headline benchmark: FP64 cell-updates/s of the structured-block explicit update.

Workload (one GPU): the flux-path bandwidth stress case, a single 3-D block of
256x128x128 cells of ideal air, Mach-6 free stream (p = 2e3 Pa, T = 220 K) with seeded
+/-5 % Fourier perturbations of density and temperature, supersonic inflow on the west
face and simple outflow elsewhere, MUSCL (l2r2) + van Albada, AUSMDV, predictor-corrector
(2 stages) with the reference's CFL controller.  A step is one full explicit update of
all cells (dt control, ghost cells, two flux/update stages); a cell-update is one cell
advanced one step.  Inputs are resident in HBM before timing starts.

Prints one JSON line (see the repository's benchmark contract).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFLOPS = 78.6


def stage_bytes_per_cell(dims=3):
    """Algorithmic HBM bytes per cell for one fused predictor-corrector stage (FP64).

    Per cell and stage the kernel must read the 6 reconstruction inputs of the flow
    state (rho, vx, vy, vz, u, a), U0 (5), the geometry it owns (3 faces x (n, t1, t2,
    area) = 30, volume, 3 lengths), and write the decoded flow state (8); stage 1 also
    writes dUdt0 (5), stage 2 reads dUdt0 (5) and writes U (5).  Average of the two.
    """
    ncq = 5 if dims == 3 else 4
    nf = 3 if dims == 3 else 2
    geom = nf * 10 + 1 + 3
    s1 = 6 + ncq + geom + 8 + ncq
    s2 = 6 + ncq + ncq + geom + 8 + ncq
    return 8.0 * (s1 + s2) / 2.0


def face_bytes_per_face(dims=3):
    """Algorithmic HBM bytes per face of the face-flux kernel (FP64).

    Each face reads its frame and area (10 doubles) and writes F (ncq doubles).  Its four
    stencil cells are shared with the neighbouring faces of the same direction, so per
    face one cell's worth of reconstruction inputs is new: rho, vx, vy, vz, u, a and the
    length along the direction (7 doubles).
    """
    ncq = 5 if dims == 3 else 4
    return 8.0 * (10 + ncq + 7)


def cpu_baseline(seconds_target=15.0):
    """The oracle (C++ restatement of the D path, one thread as in the reference's one-block
    mode) timed on a bounded sample: a 256x128x8 slab of the same Mach-6 box."""
    from gdtk_amd import cases
    from oracle import oracle as O
    case = cases.mach6_box(256, 128, 8, seed=1234)
    g = O.IdealGas()
    s = O.Sim(g, 3, "ausmdv")
    s.add_block(case)
    s.run(1)  # first step (dt set-up) untimed
    steps = 0
    t0 = time.perf_counter()
    while True:
        s.run(1)
        steps += 1
        el = time.perf_counter() - t0
        if el > seconds_target or steps >= 50:
            break
    return dict(value=case.ncells * steps / el, unit="cell-updates/s", cores=1, kind="port",
                sample=f"oracle, 1 thread, 256x128x8 slab of the Mach-6 box, {steps} steps in {el:.1f} s")


def measured_traffic(kernel, items):
    """HBM bytes per launch of the kernel from the committed counter summary, if one exists
    for this kernel and workload (profiles/*_traffic.json, written from rocprofv3 --pmc
    FETCH_SIZE and WRITE_SIZE runs scaled by a copy kernel of known size)."""
    path = os.path.join(ROOT, "profiles", "flux_mach6_256x128x128_traffic.json")
    try:
        with open(path) as f:
            d = json.load(f)
        if d.get("kernel") == kernel and abs(d.get("items_per_launch", 0) - items) < 0.5:
            return d["hbm_bytes_per_launch"]
    except (OSError, ValueError, KeyError):
        pass
    return None


def copy_probe(nbytes=1 << 30):
    """A plain device copy of known size for HBM counter calibration (FETCH_SIZE/WRITE_SIZE)."""
    import torch
    a = torch.ones(nbytes // 8, dtype=torch.float64, device="cuda")
    b = torch.empty_like(a)
    for _ in range(3):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"copy_probe_bytes": nbytes}), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--flux", default="ausmdv", choices=["ausmdv", "efm"])
    ap.add_argument("--size", default="256,128,128")
    ap.add_argument("--copy-probe", action="store_true",
                    help="run only a device copy of known size (HBM counter calibration)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU only (asked for {args.gpus}); "
              "the multi-GPU sharding is described in DESIGN.md", file=sys.stderr)
        sys.exit(2)
    if args.copy_probe:
        copy_probe()
        return
    nic, njc, nkc = (int(x) for x in args.size.split(","))
    from gdtk_amd import Simulation, cases
    case = cases.mach6_box(nic, njc, nkc, seed=1234)
    sim = Simulation(3, flux=args.flux)
    sim.add_block(case)
    print(f"bench: {nic}x{njc}x{nkc} = {case.ncells} cells, flux={args.flux}, warmup {args.warmup}", file=sys.stderr,
          flush=True)
    # The autotuner (part of set-up, not timed) picks the stage kernels for this block.
    tune = sim.autotune()
    print("bench: autotune", tune, file=sys.stderr, flush=True)
    if args.warmup:
        sim.run(args.warmup)
    # Timed region: K whole steps; run() synchronises on the device before returning.
    t0 = time.perf_counter()
    n, st = sim.run(args.steps)
    el = time.perf_counter() - t0
    ms_per_step = 1e3 * el / args.steps
    value = case.ncells * args.steps / el
    # Dominant kernel: timed with HIP events around each of its launches on the library's
    # stream (the stream every kernel of the library runs on).
    sim.profile(True)
    sim.run(min(10, args.steps))
    ks = sim.kernel_stats()
    sim.profile(False)
    dom = max(("face_flux", "fused_stage"), key=lambda k: ks[k]["ms"])
    kd = ks[dom]
    avg_ms = kd["ms"] / max(kd["launches"], 1)
    if dom == "face_flux":
        bpu, unit_name, kname = face_bytes_per_face(3), "face", "k_face_flux<3,%d>" % (1 if args.flux == "efm" else 0)
    else:
        bpu, unit_name, kname = stage_bytes_per_cell(3), "cell", "k_fused_stage<3,%d,*>" % (1 if args.flux == "efm" else 0)
    achieved = bpu * kd["items_per_launch"] / (avg_ms * 1e-3) / 1e9
    traffic = measured_traffic(kname, kd["items_per_launch"])
    out = {
        "metric": "cell-updates/s (flux+chem) FP64",
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
        "data": "synthetic",
        "config": {"workload": f"Mach-6 3-D box {nic}x{njc}x{nkc}, ideal air, l2r2+van Albada, {args.flux}, pc RK2",
                   "cells": case.ncells, "flux": args.flux, "stage_kernels": tune["best"],
                   "kchunk": tune["kchunk"], "final_dt": st["dt"]},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic, "kernel": kname,
                     "avg_launch_ms": avg_ms, "bytes_per_" + unit_name: bpu,
                     "items_per_launch": kd["items_per_launch"]},
        "kernels": {k: {"launches": v["launches"], "avg_ms": v["ms"] / max(v["launches"], 1)} for k, v in ks.items()
                    if v["launches"]},
    }
    if not args.no_cpu_baseline:
        try:
            out["cpu_baseline"] = cpu_baseline()
        except Exception as e:  # the baseline is reported, never required
            out["cpu_baseline"] = {"error": str(e)}
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
