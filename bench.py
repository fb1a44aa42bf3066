"""Headline benchmark: FP64 cell-updates/s of the structured-block explicit update with
operator-split chemistry (flux + chem).

Headline workload (one GPU): the coupled 2-D Mach-10 cylinder, R = 1 cm, 5-species air
(N2, O2, NO, N, O; Park 1-T mechanism) at p = 500 Pa, T = 300 K, on 8 blocks of 256x128
cells joined by exchange boundaries; MUSCL (l2r2) + van Albada, AUSMDV, predictor-corrector
(2 stages) with the reference's CFL controller, then the alpha-QSS chemistry step of every
cell (operator split).  A step is one flux step plus one chemistry step of all cells; a
cell-update is one cell advanced one such step.  The flow is started from the free stream
and spun up (set-up, untimed) until the bow shock and its reacting shock layer stand off
the body; inputs are resident in HBM before timing starts.

Sub-legs (compact records, summarised last in the JSON line): the Sod tube (200x100, 500
steps), the Mach-6 box (256x128x128, AUSMDV and EFM), 64 uneven blocks, 4 Mi chemistry
cells with the alpha-QSS integrator and the learned surrogate, the cylinder's surrogate
run against the integrator run, adaptive flux, Eilmer 5 numerics, flow files and the
reference's n90 case.

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


def cpu_baseline(seconds_target=15.0, flux="ausmdv"):
    """The oracle (C++ restatement of the D path, one thread as in the reference's one-block
    mode) timed on a bounded sample: a 256x128x8 slab of the same Mach-6 box."""
    from gdtk_amd import cases
    from oracle import oracle as O
    case = cases.mach6_box(256, 128, 8, seed=1234)
    g = O.IdealGas()
    s = O.Sim(g, 3, flux)
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
    s.close()
    return dict(value=case.ncells * steps / el, unit="cell-updates/s", cores=1, kind="port",
                sample=f"oracle, 1 thread, {flux}, 256x128x8 slab of the Mach-6 box, {steps} steps in {el:.1f} s")


def source_digest(names):
    """A digest of the library sources that define a kernel, so that a committed counter
    summary is used only for the kernel build it was measured on."""
    import hashlib
    h = hashlib.sha256()
    for n in names:
        with open(os.path.join(ROOT, "gdtk_amd", "csrc", n), "rb") as f:
            h.update(f.read())
    return h.hexdigest()[:16]


# Sources of each kernel whose HBM traffic is taken from profiles/traffic.json.
KERNEL_SOURCES = {
    "k_face_flux_batch": ("flux_kernels.hip", "gdtk_device.hpp", "kernel_common.hpp", "block.hpp"),
    "k_face_flux_species_pair": ("flux_species.hip", "flux_species.hpp", "chem_device.hpp", "gdtk_device.hpp",
                            "kernel_common.hpp", "block.hpp"),
}


def measured_traffic(kernel, items):
    """HBM bytes per launch of the kernel from the committed counter summary
    (profiles/traffic.json: rocprofv3 --pmc FETCH_SIZE and WRITE_SIZE runs scaled by a copy
    kernel of known size), if a record exists for this kernel, this launch size and the
    current sources of the kernel; None otherwise."""
    path = os.path.join(ROOT, "profiles", "traffic.json")
    base = kernel.split("<")[0]
    try:
        with open(path) as f:
            recs = json.load(f)
        digest = source_digest(KERNEL_SOURCES[base])
        for d in recs:
            if d.get("kernel") == kernel and abs(d.get("items_per_launch", 0) - items) < 0.5 \
                    and d.get("source_digest") == digest:
                return d["hbm_bytes_per_launch"]
    except (OSError, ValueError, KeyError):
        pass
    return None


def split_rate_flops(mech):
    """Algorithmic FP64 flops of one eval_split_rates (reaction_mechanism.d:83-96) for the
    mechanism record: per reaction, the collider sum (a multiply and an add per efficiency,
    then one multiply each for the forward and backward rates), nu multiplies per listed
    reactant or product (the concentration power and the product), and per species with a
    nonzero net coefficient two multiply-adds into q and L."""
    f = 0
    for ir in range(mech.nr):
        r = mech.r[ir]
        if r.collider:
            f += 2 * r.neff + 2
        f += sum(r.reac_nu[k] for k in range(r.nreac)) + sum(r.prod_nu[k] for k in range(r.nprod))
        nu = {}
        for k in range(r.nreac):
            nu[r.reac_sp[k]] = nu.get(r.reac_sp[k], 0) - r.reac_nu[k]
        for k in range(r.nprod):
            nu[r.prod_sp[k]] = nu.get(r.prod_sp[k], 0) + r.prod_nu[k]
        f += 4 * sum(1 for v in nu.values() if v != 0)
    return f


def alpha_qss_flops(mech, nsp):
    """Algorithmic FP64 flops (add, multiply, divide each 1) of the alpha-QSS integration,
    as (per step attempt, per corrector iteration), counted from chemistry_update.d:533-655
    and the controller :162-277.  Per attempt: the rates at y0, per species p0 (2), alpha
    (3 + 19 for the two cubic polynomials and their quotient) and the predictor (8); the
    step suggestion (4 per species, 15 for three Newton square-root iterations, 3); the
    mass-fraction test (2 per species, nsp - 1 + 2 for the pairwise sum and its check) and 5
    for the controller.  Per corrector: the rates at yp, per species pp (2), alpha-bar (22),
    q-tilde (4), p-bar (2), the corrector (8) and the convergence test (3)."""
    fs = split_rate_flops(mech)
    per_attempt = fs + nsp * (2 + 22 + 8) + (4 * nsp + 15 + 3) + (2 * nsp + nsp - 1 + 2) + 5
    per_iter = fs + nsp * (2 + 22 + 4 + 2 + 8 + 3)
    return per_attempt, per_iter


def chem_cpu_baseline(gas, mech, cells, dt, seconds_target=15.0):
    """The oracle's batched thermochemical increment (OpenMP over cells, as the reference
    runs blocks in parallel) on a bounded sample of the same cells, chunk by chunk."""
    import numpy as np

    from oracle import oracle as O
    n = cells["rho"].size
    threads = max(1, min(16, os.cpu_count() or 1))
    chunk = 32768
    Y = cells["massf"].reshape(-1, n)
    done = 0
    t0 = time.perf_counter()
    while done < n:
        sl = slice(done, min(n, done + chunk))
        m = sl.stop - sl.start
        O.chem_step_batch(gas, mech, cells["rho"][sl], cells["T"][sl], cells["u"][sl], cells["p"][sl],
                          np.ascontiguousarray(Y[:, sl]).reshape(-1), np.full(m, -1.0), dt, nthreads=threads)
        done += m
        if time.perf_counter() - t0 > seconds_target:
            break
    el = time.perf_counter() - t0
    return dict(value=done / el, unit="cells/s", cores=threads, kind="port",
                sample=f"oracle, {threads} OpenMP threads, first {done} of the {n} cells, dt_chem = -1, "
                       f"{el:.1f} s")


def bench_chem(ncells, steps, warmup, with_cpu=True, seed=2024, dt=1.0e-7):
    """Plain chemistry, uneven stiffness: alpha-QSS over dt on independent 5-species air
    cells (gdtk_amd.chemcases.air_cells), each timed step from the same resident inputs with
    dt_chem = -1, out of place."""
    import numpy as np
    import torch

    from gdtk_amd import data as gdata
    from gdtk_amd.chem import Chemistry, gas_record, mech_record
    from gdtk_amd.chemcases import air_cells
    gas = gas_record(gdata.load("tpg-air-5sp"))
    mech = mech_record(gdata.load("park-air-5sp-1T"))
    ch = Chemistry(gas, mech)
    cells = air_cells(ncells, seed=seed)
    dev = torch.device("cuda:0")
    t = {k: torch.tensor(v, dtype=torch.float64, device=dev) for k, v in cells.items()}
    dtc = torch.full((ncells,), -1.0, dtype=torch.float64, device=dev)
    out = {k: torch.empty_like(t[k]) for k in ("T", "u", "p", "massf")}
    dto = torch.empty_like(dtc)
    st = torch.empty(ncells, dtype=torch.int32, device=dev)
    stream = torch.cuda.current_stream()
    ptrs = (t["rho"].data_ptr(), t["T"].data_ptr(), t["u"].data_ptr(), t["p"].data_ptr(), t["massf"].data_ptr(),
            dtc.data_ptr())

    def step():
        ch.step_device(ncells, *ptrs, dt, out["T"].data_ptr(), out["u"].data_ptr(), out["p"].data_ptr(),
                       out["massf"].data_ptr(), dto.data_ptr(), st.data_ptr(), stream.cuda_stream)

    tune = ch.autotune(ncells, *ptrs, dt)
    torch.cuda.synchronize()
    best = min(tune, key=lambda r: r["ms"])
    print("bench: chem autotune", best, file=sys.stderr, flush=True)
    for _ in range(warmup):
        step()
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    kern = {"ms_setup": 0.0, "ms_integrate": 0.0, "ms_finish": 0.0}
    e0.record(stream)
    t0 = time.perf_counter()
    for _ in range(steps):
        step()
    e1.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ms_total = e0.elapsed_time(e1)
    # per-kernel durations (HIP events the library records around each kernel on `stream`)
    for _ in range(3):
        step()
        s = ch.stats()
        for k in kern:
            kern[k] += s[k] / 3.0
    failed = int((st != 0).sum().item())
    per_att, per_it = alpha_qss_flops(mech, gas.nsp)
    flops = per_att * s["attempts_total"] + per_it * s["iters_total"]
    achieved = flops / (kern["ms_integrate"] * 1e-3) / 1e12
    res = {
        "metric": "chem cells/s (alpha-QSS, FP64)",
        "value": ncells * steps / (ms_total * 1e-3),
        "unit": "cells/s",
        "steps": steps,
        "warmup": warmup,
        "ms_per_step": ms_total / steps,
        "wall_ms_per_step": 1e3 * wall / steps,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"plain chemistry, uneven stiffness: {ncells} cells of 5-species air (Park 1-T), "
                               "T log-uniform 1500-12000 K, p log-uniform 1e3-1e6 Pa, dt 1e-7 s, dt_chem -1",
                   "cells": ncells, "dt": dt, "failed_cells": failed,
                   "substeps_total": s["substeps_total"], "substeps_max": s["substeps_max"],
                   "attempts_total": s["attempts_total"], "corrector_iters_total": s["iters_total"],
                   "tuning": {k: best[k] for k in ("persistent", "chunk", "refill", "blocks_per_cu")}},
        "roofline": {"bound": "fp64", "achieved": achieved, "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved / FP64_PEAK_TFLOPS, "traffic": None,
                     "kernel": "k_chem_integrate<5,5>", "avg_launch_ms": kern["ms_integrate"],
                     "flops_per_attempt": per_att, "flops_per_corrector": per_it, "flops_per_launch": flops},
        "kernels": kern,
    }
    if with_cpu:
        try:
            res["cpu_baseline"] = chem_cpu_baseline(gas, mech, cells, dt)
        except Exception as e:  # the baseline is reported, never required
            res["cpu_baseline"] = {"error": str(e)}
    return res


MLP_FLOPS_PER_CELL = 2 * (8 * 256 + 256 * 256 + 256 * 256 + 256 * 5)  # 268 800 per network
BF16_PEAK_TFLOPS = 2500.0
DY_BOUND, T_BOUND, SPEEDUP_TARGET = 1.0e-3, 5.0e-3, 5.0


def bench_surrogate(ncells, steps, seed=2024, dt=1.0e-7, n_train=1 << 22, tols=(2.0e-4, 5.0e-4, 2.0e-3), log=None):
    """The learned step on the same cells as the alpha-QSS measurement: two networks trained
    (set-up, untimed) on n_train integrator samples of independently seeded cells, then, for
    each disagreement tolerance, K timed steps from the same resident inputs compared cell
    by cell with the integrator.  The first tolerance whose max |dY| error is <= 1e-3 is the
    reported operating point (the largest tolerance if none is)."""
    import numpy as np
    import torch

    from gdtk_amd import data as gdata
    from gdtk_amd.chem import Chemistry, gas_record, mech_record
    from gdtk_amd.chemcases import air_cells
    from gdtk_amd.surrogate import integrator_samples, train_air_surrogate, with_tolerance
    gd = gdata.load("tpg-air-5sp")
    ch = Chemistry(gas_record(gd), mech_record(gdata.load("park-air-5sp-1T")))
    sur0, info = train_air_surrogate(ch, gd, n_train=n_train, dt=dt, seed=77, log=log)
    dev = torch.device("cuda:0")
    cells = air_cells(ncells, seed=seed)
    t_in, ref, ref_st = integrator_samples(ch, cells, dt, dev)
    out = {k: torch.empty_like(t_in[k]) for k in ("T", "u", "p", "massf")}
    dtc = torch.full((ncells,), -1.0, dtype=torch.float64, device=dev)
    dto = torch.empty_like(dtc)
    st = torch.empty(ncells, dtype=torch.int32, device=dev)
    path = torch.empty(ncells, dtype=torch.int32, device=dev)
    stream = torch.cuda.current_stream()
    args = (t_in["rho"].data_ptr(), t_in["T"].data_ptr(), t_in["u"].data_ptr(), t_in["p"].data_ptr(),
            t_in["massf"].data_ptr(), dtc.data_ptr(), dt, out["T"].data_ptr(), out["u"].data_ptr(),
            out["p"].data_ptr(), out["massf"].data_ptr(), dto.data_ptr(), st.data_ptr(), path.data_ptr(),
            stream.cuda_stream)
    Tref = ref["T"].cpu().numpy()
    rs = ref_st.cpu().numpy()
    sweep = []
    for tol in tols:
        sur = with_tolerance(sur0, tol)
        sur.step_device(ncells, *args)  # warm-up (packs the BF16 weights)
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        for _ in range(steps):
            sur.step_device(ncells, *args)
        e1.record(stream)
        torch.cuda.synchronize()
        ms = e0.elapsed_time(e1) / steps
        s = sur.stats()
        P = path.cpu().numpy()
        ss = st.cpu().numpy()
        dY = (out["massf"] - ref["massf"]).abs().reshape(5, ncells).max(dim=0).values.cpu().numpy()
        Terr = np.abs(out["T"].cpu().numpy() - Tref) / Tref
        ok = (rs == 0) & (ss == 0)  # accuracy over cells both complete
        okT = ok & (Tref > 300.0)   # T error where the integrator's T is physical
        achieved = 2 * MLP_FLOPS_PER_CELL * ncells / (s["ms_network"] * 1e-3) / 1e12
        sweep.append({
            "disagree_tol": tol, "value": ncells / (ms * 1e-3), "ms_per_step": ms,
            "max_abs_dY": float(dY[ok].max()), "p999_abs_dY": float(np.percentile(dY[ok], 99.9)),
            "p99_abs_dY": float(np.percentile(dY[ok], 99)), "max_rel_T_err": float(Terr[okT].max()),
            "p999_rel_T_err": float(np.percentile(Terr[okT], 99.9)),
            "fallback_fraction": float(np.count_nonzero(P == 1)) / ncells,
            "network_cells": int(np.count_nonzero(P == 0)), "failed_cells": int(np.count_nonzero(ss != 0)),
            "integrator_failed_cells": int(np.count_nonzero(rs != 0)),
            "false_failures": int(np.count_nonzero((rs == 0) & (ss != 0))),
            "false_successes": int(np.count_nonzero((rs != 0) & (ss == 0))),
            "kernels": {k: s[k] for k in ("ms_network", "ms_project", "ms_integrator")},
            "mfma_tflops": achieved})
        if log:
            log(f"surrogate tol {tol:g}: {sweep[-1]['value']:.3e} cells/s, fallback "
                f"{sweep[-1]['fallback_fraction']:.3f}, max dY {sweep[-1]['max_abs_dY']:.2e}")
        sur.close()
    # The target operating point: max |dY| <= 1e-3 and T error <= 0.5 % over the cells that
    # both complete.  Where no tolerance reaches it, value is None and the nearest point (the
    # smallest max |dY|) is reported beside it.
    meets = [r for r in sweep if r["max_abs_dY"] <= DY_BOUND and r["max_rel_T_err"] <= T_BOUND]
    nearest = min(sweep, key=lambda r: r["max_abs_dY"])
    pick = meets[0] if meets else nearest
    return {
        "metric": "chem cells/s (surrogate + alpha-QSS fallback)",
        "value": pick["value"] if meets else None,
        "meets_target": bool(meets),
        "target": {"max_abs_dY": DY_BOUND, "max_rel_T_err": T_BOUND},
        "nearest": nearest,
        "unit": "cells/s",
        "ms_per_step": pick["ms_per_step"],
        "steps": steps,
        "dtype": "bf16 network (fp32 accumulate), f64 projection/T(u)/fallback",
        "operating_point": pick,
        "sweep": sweep,
        "training": {"samples": info["samples"], "sample_s": info["sample_s"], "train_s": info["train_s"],
                     "final_loss": [float(h[-1]) for h in info["loss"]]},
        "roofline": {"bound": "mfma", "achieved": pick["mfma_tflops"], "peak": BF16_PEAK_TFLOPS, "unit": "TFLOP/s",
                     "frac": pick["mfma_tflops"] / BF16_PEAK_TFLOPS, "traffic": None,
                     "kernel": "k_mlp_infer<256,3> x2 + features", "avg_launch_ms": pick["kernels"]["ms_network"],
                     "flops_per_cell": 2 * MLP_FLOPS_PER_CELL},
    }


def bench_batched(nblocks, steps, warmup, flux="ausmdv", log=None):
    """Batched uneven blocks: nblocks blocks with per-block edges drawn from 16..192 cells
    (seeded; cases.batched_ragged), in pairs joined by exchange boundaries, each with its
    own Mach 2-10 free stream; every stage is one ghost-cell launch and one face and one cell
    launch (or one persistent stage launch) over all blocks, each step one captured graph.
    Returns a sub-record with its own throughput and the dominant kernel's roofline."""
    from gdtk_amd import Simulation, cases
    t0 = time.perf_counter()
    blocks = cases.batched_ragged(nblocks)
    ncells = sum(b.ncells for b in blocks)
    sim = Simulation(3, flux=flux)
    for q, b in enumerate(blocks):
        sim.add_block(b)
        if log and q % 16 == 15:
            log(f"batched: {q + 1} blocks set up, {time.perf_counter() - t0:.1f} s")
    if log:
        log(f"batched: {len(blocks)} blocks, {ncells} cells, set-up {time.perf_counter() - t0:.1f} s")
    tune = sim.autotune(dt=0.25 * min(sim.block_dt(q, 1.0, 0.5, False)[0] for q in range(len(blocks))), reps=3)
    if log:
        log(f"batched autotune {tune}")
    if warmup:
        sim.run(warmup)
    t0 = time.perf_counter()
    _, st = sim.run(steps)
    el = time.perf_counter() - t0
    sim.profile(True)
    sim.run(min(5, steps))
    ks = sim.kernel_stats()
    sim.profile(False)
    dom = max(("face_flux", "fused_stage"), key=lambda k: ks[k]["ms"])
    kd = ks[dom]
    avg_ms = kd["ms"] / max(kd["launches"], 1)
    bpu = stage_bytes_per_cell(3) if dom == "fused_stage" else face_bytes_per_face(3)
    achieved = bpu * kd["items_per_launch"] / (avg_ms * 1e-3) / 1e9
    sim.close()
    return {
        "value": ncells * steps / el, "unit": "cell-updates/s", "steps": steps, "ms_per_step": 1e3 * el / steps,
        "config": {"workload": f"batched uneven blocks: {nblocks} blocks, per-block edges 16-192, pairs joined "
                               f"by exchange, Mach 2-10 per block, {flux}, pc RK2",
                   "blocks": len(blocks), "cells": ncells, "stage_kernels": tune["best"], "kchunk": tune["kchunk"],
                   "tile_x": tune.get("tile_x"),
                   "graph": True, "final_dt": st["dt"]},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "kernel": f"k_fused_batch<3,{1 if flux == 'efm' else 0},*>"
                     if dom == "fused_stage" else dom, "avg_launch_ms": avg_ms,
                     "items_per_launch": kd["items_per_launch"]},
        "kernels": {k: {"launches": v["launches"], "avg_ms": v["ms"] / max(v["launches"], 1)} for k, v in ks.items()
                    if v["launches"]},
    }


def bench_adaptive(nic, njc, nkc, steps, warmup=2, log=None):
    """The headline Mach-6 box with the adaptive EFM/AUSMDV calculator: shock detection
    (in stage 1, a launch over all faces for the detector's decisions, then one over all
    cells for the indicators) plus the face kernels that choose EFM on marked faces.
    Returns a sub-record with the throughput, the share of marked cells and the detection's
    roofline over its two launches (algorithmic bytes per cell: the cell's velocity and
    sound speed, the n/t1/t2 frames of its three low-side faces and the 1-byte indicator)."""
    import numpy as np
    from gdtk_amd import Simulation, cases
    case = cases.mach6_box(nic, njc, nkc, seed=1234)
    sim = Simulation(3, flux="adaptive")
    sim.add_block(case)
    if warmup:
        sim.run(warmup)
    t0 = time.perf_counter()
    _, st = sim.run(steps)
    el = time.perf_counter() - t0
    sim.profile(True)
    sim.run(min(5, steps))
    ks = sim.kernel_stats()
    sim.profile(False)
    marked = float(np.mean(sim.get(0, "S")))
    kd = ks["shock_detect"]
    avg_ms = kd["ms"] / max(kd["launches"], 1)
    bpc = 4 * 8 + 3 * 9 * 8 + 1
    achieved = bpc * kd["items_per_launch"] / (avg_ms * 1e-3) / 1e9
    sim.close()
    if log:
        log(f"adaptive: {case.ncells * steps / el:.3e} cell-updates/s, marked {marked:.3f}")
    return {
        "value": case.ncells * steps / el, "unit": "cell-updates/s", "steps": steps, "ms_per_step": 1e3 * el / steps,
        "marked_cell_fraction": marked,
        "config": {"workload": f"Mach-6 3-D box {nic}x{njc}x{nkc}, ideal air, l2r2+van Albada, "
                               "adaptive_efm_ausmdv (PJ detector, strict), pc RK2", "cells": case.ncells,
                   "final_dt": st["dt"]},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "kernel": "k_shock_faces_batch<3> + k_shock_cells_batch<3>", "avg_launch_ms": avg_ms,
                     "bytes_per_cell": bpc, "items_per_launch": kd["items_per_launch"]},
        "kernels": {k: {"launches": v["launches"], "avg_ms": v["ms"] / max(v["launches"], 1)} for k, v in ks.items()
                    if v["launches"]},
    }


def bench_lmr(nic, njc, nkc, steps, warmup=2, log=None):
    """The headline Mach-6 box with Eilmer 5 (lmr) numerics: the face kernel computes the
    dimensionalised van Albada epsilon per reconstructed scalar and AUSMDV's smooth-maximum
    sound speed.  Returns the throughput and the face kernel's average launch."""
    from gdtk_amd import Simulation, cases
    case = cases.mach6_box(nic, njc, nkc, seed=1234)
    sim = Simulation(3, flux="ausmdv", numerics="lmr")
    sim.add_block(case)
    if warmup:
        sim.run(warmup)
    t0 = time.perf_counter()
    _, st = sim.run(steps)
    el = time.perf_counter() - t0
    sim.profile(True)
    sim.run(min(5, steps))
    ks = sim.kernel_stats()
    sim.profile(False)
    sim.close()
    if log:
        log(f"lmr numerics: {case.ncells * steps / el:.3e} cell-updates/s")
    return {
        "value": case.ncells * steps / el, "unit": "cell-updates/s", "steps": steps, "ms_per_step": 1e3 * el / steps,
        "config": {"workload": f"Mach-6 3-D box {nic}x{njc}x{nkc}, ideal air, l2r2+van Albada with Eilmer 5 "
                               "epsilon scaling, ausmdv with smooth-maximum sound speed, pc RK2",
                   "cells": case.ncells, "final_dt": st["dt"]},
        "kernels": {k: {"launches": v["launches"], "avg_ms": v["ms"] / max(v["launches"], 1)} for k, v in ks.items()
                    if v["launches"]},
    }


def bench_flow_io(nic, njc, nkc, text_size=(64, 64, 64), log=None):
    """Flow-solution files of the Mach-6 box: write_solution and the restart read + cold start
    of one block, raw binary at the headline size and gzip text on a smaller box (the text
    path formats every value with %.18e on the host threads and compresses with zlib).
    Wall-clock times around the library calls, file on the local disk."""
    import os
    import tempfile
    from gdtk_amd import Simulation, cases
    out = {}
    for fmt, size in (("rawbinary", (nic, njc, nkc)), ("gziptext", text_size)):
        case = cases.mach6_box(*size, seed=1234)
        sim = Simulation(3, flux="ausmdv")
        sim.add_block(case)
        sim.run(2)
        with tempfile.TemporaryDirectory() as d:
            fn = os.path.join(d, "b0000.flow")
            t0 = time.perf_counter()
            sim.write_flow(0, fn, fmt)
            tw = time.perf_counter() - t0
            nbytes = os.path.getsize(fn)
            t0 = time.perf_counter()
            sim.read_flow(0, fn, fmt)
            tr = time.perf_counter() - t0
        sim.close()
        out[fmt] = {"cells": case.ncells, "file_bytes": nbytes, "write_s": tw, "read_s": tr,
                    "write_cells_per_s": case.ncells / tw, "read_cells_per_s": case.ncells / tr}
        if log:
            log(f"flow io {fmt}: {case.ncells} cells, write {tw:.2f} s, read {tr:.2f} s, {nbytes / 1e6:.0f} MB")
    return out


def bench_n90(cpu_steps=300, log=None):
    """The reference's own reacting integration case, examples/eilmer/2D/cylinder-dlr-n90
    (N2/N, 2 reactions, adaptive_hanel_ausmdv, 4 blocks of 60x10, 100 us): time to solution
    on the GPU, the step count and probe state the reference's n90-test.rb checks, and the
    oracle's steps/s on one core over its first cpu_steps steps (the CPU leg of the same run)."""
    import numpy as np
    from gdtk_amd import Simulation, cases, data
    from gdtk_amd.chem import Chemistry, mech_record
    blocks, rec, info = cases.n90_blocks()
    mech = mech_record(data.load("nitrogen-2sp-2r"))
    sim = Simulation(2, flux="adaptive_hanel_ausmdv", tpg=rec, dt_init=info["dt_init"])
    for b in blocks:
        sim.add_block(b)
    sim.set_chemistry(Chemistry(rec, mech))
    t0 = time.perf_counter()
    _, st = sim.run(40000, info["max_time"])
    el = time.perf_counter() - t0
    best = (np.inf, 0, 0)
    for q, b in enumerate(blocks):
        xc = cases.cell_centres(b)
        d = np.hypot(xc[:, 0] - info["probe"][0], xc[:, 1] - info["probe"][1])
        i = int(np.argmin(d))
        if d[i] < best[0]:
            best = (d[i], q, i)
    _, q, i = best
    probe = {"rho": sim.get(q, "rho")[i], "vel.x": sim.get(q, "vx")[i], "p": sim.get(q, "p")[i],
             "massf_N": sim.get(q, 301)[i], "T": sim.get(q, "T")[i]}
    ncell = sum(b.ncells for b in blocks)
    sim.close()
    out = {"steps": st["step"], "reference_steps": "4428 +/- 30", "wall_s": el, "steps_per_s": st["step"] / el,
           "cell_updates_per_s": ncell * st["step"] / el, "probe": probe,
           "reference_probe": {"rho": 1.9090e-02, "vel.x": 5.797e+02, "p": 5.3735e+04, "massf_N": 1.749e-02,
                               "T": 9.3205e+03}}
    if cpu_steps:
        from oracle import oracle as O
        o = O.Sim(O.TPG(rec), 2, "adaptive_hanel_ausmdv", dt_init=info["dt_init"])
        for b in blocks:
            o.add_block(b)
        o.set_chemistry(O.Chem(rec, mech))
        t0 = time.perf_counter()
        o.run(cpu_steps)
        cel = time.perf_counter() - t0
        out["cpu_baseline"] = {"steps_per_s": cpu_steps / cel, "cores": 1, "kind": "port",
                               "sample": f"first {cpu_steps} steps of the same run, oracle on one core"}
    if log:
        log(f"n90: {st['step']} steps in {el:.2f} s on the GPU")
    return out


def _cell_fields(sim, nblocks, names):
    import numpy as np
    return {n: np.concatenate([sim.get(q, n) for q in range(nblocks)]) for n in names}


def per_result_accuracy(ch, sur, fields, dt):
    """The chemistry step of one state two ways, on the same inputs: the alpha-QSS integrator
    and the surrogate (routing included).  Over the cells the networks answered: max |dY|
    and relative T error against the integrator, and the two paths' times."""
    import numpy as np
    import torch
    dev = torch.device("cuda:0")
    n = fields["rho"].size
    t = {k: torch.tensor(np.ascontiguousarray(v), dtype=torch.float64, device=dev) for k, v in fields.items()}
    outs = []
    stream = torch.cuda.current_stream()
    for which in ("integrator", "surrogate"):
        o = {k: torch.empty_like(t[k]) for k in ("T", "u", "p", "massf")}
        dto = torch.empty(n, dtype=torch.float64, device=dev)
        st = torch.empty(n, dtype=torch.int32, device=dev)
        path = torch.empty(n, dtype=torch.int32, device=dev)
        args = (n, t["rho"].data_ptr(), t["T"].data_ptr(), t["u"].data_ptr(), t["p"].data_ptr(), t["massf"].data_ptr(),
                t["dt_chem"].data_ptr(), dt, o["T"].data_ptr(), o["u"].data_ptr(), o["p"].data_ptr(),
                o["massf"].data_ptr(), dto.data_ptr(), st.data_ptr())
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        torch.cuda.synchronize()
        e0.record(stream)
        if which == "integrator":
            ch.step_device(*args, stream.cuda_stream)
        else:
            sur.step_device(*args, path.data_ptr(), stream.cuda_stream)
        e1.record(stream)
        torch.cuda.synchronize()
        outs.append((o, st, path, e0.elapsed_time(e1)))
    (oi, sti, _, ms_i), (os_, sts, path, ms_s) = outs
    phases = sur.stats()
    P = path.cpu().numpy()
    net = (P == 0) & (sti.cpu().numpy() == 0)
    nsp = ch.nsp
    dY = np.abs(os_["massf"].cpu().numpy().reshape(nsp, n) - oi["massf"].cpu().numpy().reshape(nsp, n)).max(axis=0)
    Ti = oi["T"].cpu().numpy()
    Te = np.abs(os_["T"].cpu().numpy() - Ti) / Ti
    return {"cells": int(n), "network_cells": int(net.sum()), "integrator_ms": ms_i, "surrogate_ms": ms_s,
            "surrogate_phases_ms": {k: phases[k] for k in ("ms_network", "ms_project", "ms_integrator")},
            "speedup": ms_i / ms_s, "max_abs_dY": float(dY[net].max()) if net.any() else 0.0,
            "p999_abs_dY": float(np.quantile(dY[net], 0.999)) if net.any() else 0.0,
            "max_rel_T_err": float(Te[net].max()) if net.any() else 0.0}


def bench_coupled(steps, snapshot_every=100, tols=(5.0e-4, 2.0e-3), route_ratio=8.0, route_T=2500.0, log=None):
    """The coupled 2-D Mach-10 cylinder (5-species air, 8 blocks of 256 x 128 cells): flux
    steps with operator-split chemistry, first through the FP64 alpha-QSS integrator, then
    through the learned surrogate.  The surrogate answers the cells whose integration would
    be long (the flow step spans more than route_ratio of the cell's last suggested
    chemistry substep, or, for a cell with no suggestion yet, such as one just behind the
    shock, T above route_T) and the integrator the rest; its networks are trained on (state, dt,
    result) samples of the integrator run's own trajectory (every `snapshot_every` steps,
    every reacting cell's state (T above the 300 K free stream) integrated at 0.5, 1 and 2
    times the flow step: about 2e7 samples over 5000 steps).  Reports both
    runs' cell-update and chemistry rates, the chemistry speed-up, per result (the same
    states through both paths at the end of the run) max |dY| and T error of the network
    cells against the integrator, and the runs' shock stand-off and stagnation temperature."""
    import numpy as np

    from gdtk_amd import Simulation, cases
    from gdtk_amd import data as gdata
    from gdtk_amd.chem import Chemistry, mech_record
    from gdtk_amd.surrogate import train_surrogate, with_tolerance
    blocks, rec, info = cases.cylinder_blocks()
    nb = len(blocks)
    ncell = sum(b.ncells for b in blocks)
    gd = gdata.load("tpg-air-5sp")
    ch = Chemistry(rec, mech_record(gdata.load("park-air-5sp-1T")))
    state_names = ["rho", "T", "u", "p", "dt_chem", 300, 301, 302, 303, 304]

    def make(sur=None):
        g = Simulation(2, tpg=rec, dt_init=1.0e-10, dt_max=1.0e-6)
        for b in blocks:
            g.add_block(b)
        g.set_chemistry(ch, sur)
        return g

    def run(g, sample=None):
        done, el, st = 0, 0.0, None
        while done < steps:
            t0 = time.perf_counter()
            n, st = g.run(min(snapshot_every, steps - done))
            el += time.perf_counter() - t0
            done += n
            if sample is not None:
                sample.append((_cell_fields(g, nb, state_names), st["dt"]))
            if log and done % 1000 == 0:
                sl = cases.stagnation_line(g, blocks)
                log(f"cylinder: {done} steps, t = {st['time']:.3e} s, stand-off {sl['standoff'] * 1e3:.3f} mm, "
                    f"T_stag {sl['T_stag']:.0f} K")
        return el, st

    def as_cells(snap):
        return {"rho": snap["rho"], "T": snap["T"], "u": snap["u"], "p": snap["p"], "dt_chem": snap["dt_chem"],
                "massf": np.ascontiguousarray(np.stack([snap[300 + s] for s in range(5)])).reshape(-1)}

    pool = []
    gi = make()
    el_i, st_i = run(gi, pool)
    sl_i = cases.stagnation_line(gi, blocks)
    cs_i = gi.chem_stats()
    gi.close()
    # training samples: every reacting cell of every snapshot (the cells the router sends to
    # the networks are among them); with ~2e7 samples a few passes suffice
    cells = {k: [] for k in ("rho", "T", "u", "p", "massf")}
    n_routed = 0
    for snap, dt_snap in pool:
        d = snap["dt_chem"]
        hot = snap["T"] > 300.0
        n_routed += int(np.count_nonzero(hot & np.where(d > 0.0, dt_snap > route_ratio * d, snap["T"] > route_T)))
        for k in ("rho", "T", "u", "p"):
            cells[k].append(snap[k][hot])
        cells["massf"].append(np.stack([snap[300 + s][hot] for s in range(5)]))
    pool.clear()
    cells = {k: (np.concatenate(v, axis=1) if k == "massf" else np.concatenate(v)) for k, v in cells.items()}
    nsamp = cells["T"].size
    cells["massf"] = np.ascontiguousarray(cells["massf"]).reshape(-1)
    dt_typ = st_i["dt"]
    sur, tinfo = train_surrogate(ch, gd, cells, (0.5 * dt_typ, dt_typ, 2.0 * dt_typ), epochs=(2, 1, 1), log=log)
    del cells
    chem_rate_i = ncell * cs_i["steps"] / (cs_i["ms"] * 1e-3)
    results = []
    for tol in tols:
        s_tol = with_tolerance(sur, tol, route_ratio=route_ratio, route_T=route_T)
        g = make(s_tol)
        el_s, st_s = run(g)
        sl_s = cases.stagnation_line(g, blocks)
        cs_s = g.chem_stats()
        final = as_cells(_cell_fields(g, nb, state_names))
        g.close()
        acc = per_result_accuracy(ch, s_tol, final, st_s["dt"])
        chem_rate_s = ncell * cs_s["steps"] / (cs_s["ms"] * 1e-3)
        r = {"disagree_tol": tol, "route_ratio": route_ratio, "route_T": route_T, "cell_updates_per_s": ncell * steps / el_s,
             "chem_cells_per_s": chem_rate_s, "chem_speedup_vs_alpha_qss": chem_rate_s / chem_rate_i,
             "chem_ms_per_step": cs_s["ms"] / cs_s["steps"], "network_cells_sent_back_last_step":
                 cs_s["surrogate_queued"], "per_result": acc,
             "standoff_m": sl_s["standoff"], "standoff_rel_diff": sl_s["standoff"] / sl_i["standoff"] - 1.0,
             "T_stag_K": sl_s["T_stag"], "T_stag_rel_diff": sl_s["T_stag"] / sl_i["T_stag"] - 1.0,
             "final_time": st_s["time"]}
        r["meets_target"] = bool(r["chem_speedup_vs_alpha_qss"] >= SPEEDUP_TARGET and acc["max_abs_dY"] <= DY_BOUND
                                 and acc["max_rel_T_err"] <= T_BOUND)
        results.append(r)
        s_tol.close()
        if log:
            log(f"cylinder surrogate tol {tol:g}: chemistry {r['chem_speedup_vs_alpha_qss']:.2f}x, per result "
                f"{acc['network_cells']} network cells, max dY {acc['max_abs_dY']:.2e}, T err {acc['max_rel_T_err']:.2e}")
    return {
        "meets_target": any(r["meets_target"] for r in results),
        "target": {"chem_speedup_vs_alpha_qss": SPEEDUP_TARGET, "max_abs_dY": DY_BOUND, "max_rel_T_err": T_BOUND},
        "config": {"workload": "2-D Mach-10 cylinder, R = 1 cm, p = 500 Pa, T = 300 K, 5-species air (Park 1-T), "
                               "8 blocks of 256x128, AUSMDV, pc RK2 + operator-split chemistry",
                   "cells": ncell, "steps": steps, "final_time": st_i["time"]},
        "alpha_qss": {"cell_updates_per_s": ncell * steps / el_i, "chem_cells_per_s": chem_rate_i,
                      "chem_ms_per_step": cs_i["ms"] / cs_i["steps"], "standoff_m": sl_i["standoff"],
                      "T_stag_K": sl_i["T_stag"]},
        "surrogate_training": {"samples": int(tinfo["samples"]), "trajectory_states": int(nsamp),
                               "routed_states": n_routed, "snapshots": steps // snapshot_every,
                               "sample_s": tinfo["sample_s"], "train_s": tinfo["train_s"]},
        "surrogate": results,
    }


def bench_two_temperature(ncells, dt=1.0e-7, reps=3, with_cpu=True, seed=77, log=None):
    """The two-temperature air reactor (TwoTemperatureAirKinetics, Gupta 5-species r1-r6,
    RKF, tight coupling) over ncells post-shock states (T 4000-14000 K, Tv 500-9000 K,
    p 1-30 kPa, partly dissociated air; no substep history), timed with HIP events around
    the kernel; the oracle on a bounded sample of the same states as CPU baseline."""
    import os

    import numpy as np
    import torch

    from gdtk_amd import data as gdata
    from gdtk_amd.chem import gas_record, mech_record
    from gdtk_amd.two_temp import TwoTemperatureReactor, tt_params
    from oracle import oracle as O
    gd, md = gdata.load("tpg-air-5sp"), gdata.load("gupta-air-5sp-2T")
    gas, mech = gas_record(gd), mech_record(md)
    prm = tt_params(gd["species"])
    mol, A = [prm.molecular[i] for i in range(5)], [prm.A_MW[i] for i in range(5)]
    ot = O.TT(gas, mech, mol, A)
    rng = np.random.default_rng(seed)
    t0 = time.perf_counter()
    Y = rng.dirichlet([7.0, 2.0, 0.3, 0.3, 0.3], size=ncells).T.copy()
    T = rng.uniform(4000.0, 14000.0, ncells)
    Tv = rng.uniform(500.0, 9000.0, ncells)
    p = rng.uniform(1.0e3, 3.0e4, ncells)
    cols = {k: np.zeros(ncells) for k in ("rho", "u", "uv")}
    for i in range(ncells):
        st = ot.update("pT", Y[:, i], p=p[i], T=T[i], Tv=Tv[i])
        cols["rho"][i], cols["u"][i], cols["uv"][i] = st["rho"], st["u"], st["uv"]
    cells = dict(rho=cols["rho"], T=T, Tv=Tv, u=cols["u"], uv=cols["uv"], p=p,
                 massf=np.ascontiguousarray(Y).reshape(-1), dt_chem=np.full(ncells, -1.0))
    if log:
        log(f"two-temperature: {ncells} states in {time.perf_counter() - t0:.1f} s")
    dev = torch.device("cuda:0")
    tin = {k: torch.tensor(v, dtype=torch.float64, device=dev) for k, v in cells.items()}
    tout = {k: torch.empty_like(tin[k]) for k in ("T", "Tv", "u", "uv", "p", "massf", "dt_chem")}
    status = torch.empty(ncells, dtype=torch.int32, device=dev)
    r = TwoTemperatureReactor(gas, mech, prm)
    stream = torch.cuda.current_stream()
    ins = [tin[k].data_ptr() for k in ("rho", "T", "Tv", "u", "uv", "p", "massf", "dt_chem")]
    outs = [tout[k].data_ptr() for k in ("T", "Tv", "u", "uv", "p", "massf", "dt_chem")]
    ms = []
    for rep in range(reps + 1):
        r.step_device(ncells, *ins, dt, *outs, status.data_ptr(), stream.cuda_stream)
        torch.cuda.synchronize()
        st = r.stats()
        if rep:
            ms.append(st["ms"])
    failed = int((status != 0).sum().item())
    ms_avg = sum(ms) / len(ms)
    res = {"workload": "two-temperature air reactor, 5-species Gupta r1-r6 (Park rates, s = 0.5), RKF + tight coupling",
           "cells": ncells, "dt": dt, "ms": ms_avg, "cells_per_s": ncells / (ms_avg * 1e-3),
           "substeps_total": st["substeps_total"], "substeps_max": st["substeps_max"],
           "attempts_total": st["attempts_total"], "failed": failed, "dtype": "f64"}
    res["value"], res["unit"] = res["cells_per_s"], "cells/s"
    if with_cpu:
        threads = min(16, os.cpu_count() or 1)
        m = min(ncells, 2048)
        sub = {k: (v[:m] if k != "massf" else np.ascontiguousarray(v.reshape(5, ncells)[:, :m]).reshape(-1))
               for k, v in cells.items()}
        done, el = 0, 0.0
        while el < 8.0 and done < 64 * m:
            t0 = time.perf_counter()
            O.tt_step_batch(gas, mech, mol, A, sub["rho"], sub["T"], sub["Tv"], sub["u"], sub["uv"], sub["p"],
                            sub["massf"], sub["dt_chem"], dt, nthreads=threads)
            el += time.perf_counter() - t0
            done += m
        res["cpu_baseline"] = {"value": done / el, "unit": "cells/s", "cores": threads, "kind": "port",
                               "sample": f"oracle on the first {m} of the same states, {done} cell-steps in "
                                         f"{el:.1f} s on {threads} OpenMP threads"}
        res["speedup_vs_cpu"] = res["cells_per_s"] / res["cpu_baseline"]["value"]
    r.close()
    if log:
        log(f"two-temperature: {res['cells_per_s']:.4g} cells/s ({ms_avg:.2f} ms for {ncells} cells), "
            f"{st['substeps_total'] / ncells:.1f} substeps per cell, max {st['substeps_max']}")
    return res


def copy_probe(nbytes=1 << 30):
    """A plain device copy of known size for HBM counter calibration (FETCH_SIZE/WRITE_SIZE)."""
    import torch
    a = torch.ones(nbytes // 8, dtype=torch.float64, device="cuda")
    b = torch.empty_like(a)
    for _ in range(3):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"copy_probe_bytes": nbytes}), flush=True)


def species_face_bytes(dims, nsp):
    """Algorithmic HBM bytes per face of the multi-species face-flux kernel (FP64): the face's
    frame and area (10 doubles), F (ncq doubles), and one new stencil cell per face (its
    stencil cells are shared with the neighbouring faces of the same direction): internal
    energy, the velocity components, the nsp species densities, the cell length along the
    direction, and the cell's T (the Newton start) and sound speed."""
    ncq = (5 if dims == 3 else 4) + nsp
    return 8.0 * (10 + ncq + 3 + dims + nsp + 1)


def coupled_cpu_baseline(g, blocks, rec, mech, dt, seconds_target=12.0):
    """The oracle's coupled step (flux + chemistry, one thread, as the reference runs one
    block per thread) on the GPU run's spun-up state: every block's flow state handed over,
    then whole steps at the GPU's dt until about seconds_target has passed."""
    from oracle import oracle as O
    gas = O.TPG(rec)
    t0 = time.perf_counter()
    o = O.Sim(gas, 2, "ausmdv", dt_init=1.0e-10, dt_max=1.0e-6, mutate_cell_velocities=False)
    for b in blocks:
        o.add_block(b)
    import numpy as np
    for q, b in enumerate(blocks):
        mf = np.stack([g.get(q, 300 + s) for s in range(5)], axis=1)
        o.set_fs(q, g.get(q, "rho"), g.get(q, "vx"), g.get(q, "vy"), g.get(q, "vz"), g.get(q, "p"), g.get(q, "a"),
                 g.get(q, "u"), g.get(q, "T"), mf)
    o.set_chemistry(O.Chem(rec, mech))
    setup = time.perf_counter() - t0
    ncell = sum(b.ncells for b in blocks)
    steps, el = 0, 0.0
    while el < seconds_target and steps < 20:
        t1 = time.perf_counter()
        o.step_fixed_dt(dt)
        el += time.perf_counter() - t1
        steps += 1
    o.close()
    return dict(value=ncell * steps / el, unit="cell-updates/s", cores=1, kind="port",
                sample=f"oracle, 1 thread, {steps} whole steps (flux + alpha-QSS chemistry) of all {ncell} cells "
                       f"from the GPU run's spun-up state at its dt, {el:.1f} s (hand-over {setup:.1f} s untimed)")


def bench_coupled_headline(steps, warmup, spinup, with_cpu=True, log=None):
    """The headline: coupled cylinder flux + chemistry steps with the alpha-QSS integrator,
    timed after the spin-up; the dominant kernel's roofline from HIP events around its
    launches on the library's stream."""
    from gdtk_amd import Simulation, cases
    from gdtk_amd import data as gdata
    from gdtk_amd.chem import Chemistry, mech_record
    blocks, rec, info = cases.cylinder_blocks()
    ncell = sum(b.ncells for b in blocks)
    mech = mech_record(gdata.load("park-air-5sp-1T"))
    ch = Chemistry(rec, mech)
    g = Simulation(2, tpg=rec, dt_init=1.0e-10, dt_max=1.0e-6)
    for b in blocks:
        g.add_block(b)
    g.set_chemistry(ch)
    t0 = time.perf_counter()
    g.run(spinup)
    sl = cases.stagnation_line(g, blocks)
    if log:
        log(f"headline: spin-up {spinup} steps in {time.perf_counter() - t0:.1f} s, t = {g.state()['time']:.3e} s, "
            f"stand-off {sl['standoff'] * 1e3:.3f} mm, T_stag {sl['T_stag']:.0f} K")
    if warmup:
        g.run(warmup)
    cs0 = g.chem_stats()
    t0 = time.perf_counter()
    _, st = g.run(steps)
    el = time.perf_counter() - t0
    cs1 = g.chem_stats()
    nprof = max(1, min(10, steps))
    g.profile(True)
    cs2 = g.chem_stats()
    # the chemistry statistics describe the last reactor call: sampled after every profiled step
    integ_ms, attempts, iters = 0.0, 0, 0
    for _ in range(nprof):
        g.run(1)
        c1 = ch.stats()
        integ_ms += c1["ms_integrate"]
        attempts += c1["attempts_total"]
        iters += c1["iters_total"]
    ks = g.kernel_stats()
    cs3 = g.chem_stats()
    g.profile(False)
    chs = dict(ms_integrate=integ_ms / nprof, attempts_total=attempts / nprof, iters_total=iters / nprof)
    per_step = {k: v["ms"] / nprof for k, v in ks.items() if v["launches"]}
    per_step["chemistry"] = (cs3["ms"] - cs2["ms"]) / max(cs3["steps"] - cs2["steps"], 1)
    hot = sum(int((g.get(q, "T") > 2000.0).sum()) for q in range(len(blocks)))
    face = ks["face_flux"]
    face_avg = face["ms"] / max(face["launches"], 1)
    bpf = species_face_bytes(2, 5)
    face_ach = bpf * face["items_per_launch"] / (face_avg * 1e-3) / 1e9
    kname = "k_face_flux_species_pair<2,0,5,false>"
    roof_face = {"bound": "hbm", "achieved": face_ach, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                 "frac": face_ach / HBM_PEAK_GBS, "traffic": measured_traffic(kname, face["items_per_launch"]),
                 "kernel": kname, "avg_launch_ms": face_avg, "bytes_per_face": bpf,
                 "items_per_launch": face["items_per_launch"],
                 "note": "bound by instruction issue, not HBM: per face two Newton T(u) solves over the CEA "
                         "curves of 5 species; VALU 58 % busy, 55 % of it FP64 "
                         "(profiles/coupled_cylinder_headline_counters.txt)"}
    per_att, per_it = alpha_qss_flops(mech, 5)
    chem_flops = per_att * chs["attempts_total"] + per_it * chs["iters_total"]
    chem_ach = chem_flops / max(chs["ms_integrate"] * 1e-3, 1e-12) / 1e12
    roof_chem = {"bound": "fp64", "achieved": chem_ach, "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                 "frac": chem_ach / FP64_PEAK_TFLOPS, "traffic": None, "kernel": "k_chem_integrate<5,5,0,false,true>",
                 "avg_launch_ms": chs["ms_integrate"], "flops_per_launch": chem_flops,
                 "note": "latency bound: at a 2 ns flow step most cells take one substep and the launch lasts as "
                         "long as the slowest cell's serial attempts (VALU 14 % busy)"}
    # the kernel with the most time per step: the face kernel (two launches) or the integrator
    dominant = "face_flux" if 2 * face_avg >= chs["ms_integrate"] else "chemistry"
    out = {
        "metric": "cell-updates/s (flux+chem) FP64",
        "value": ncell * steps / el,
        "unit": "cell-updates/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": warmup,
        "ms_per_step": 1e3 * el / steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "coupled 2-D Mach-10 cylinder, R = 1 cm, p = 500 Pa, T = 300 K, 5-species air "
                               "(Park 1-T), 8 blocks of 256x128, l2r2+van Albada, AUSMDV, pc RK2 + operator-split "
                               "alpha-QSS chemistry", "cells": ncell, "spinup_steps": spinup,
                   "time_at_start": g.state()["time"] - 0.0, "final_dt": st["dt"],
                   "cells_above_2000K": hot, "chem_ms_per_step": (cs1["ms"] - cs0["ms"]) / max(
                       cs1["steps"] - cs0["steps"], 1)},
        "roofline": roof_face if dominant == "face_flux" else roof_chem,
        "roofline_other": roof_chem if dominant == "face_flux" else roof_face,
        "kernels_ms_per_step": per_step,
    }
    if with_cpu:
        try:
            out["cpu_baseline"] = coupled_cpu_baseline(g, blocks, rec, mech, st["dt"])
        except Exception as e:  # the baseline is reported, never required
            out["cpu_baseline"] = {"error": str(e)}
    g.close()
    return out


def bench_box(nic, njc, nkc, flux, steps, warmup, with_cpu=True, cpu_seconds=15.0, log=None):
    """The Mach-6 box (flux-path bandwidth stress, ideal air) with one flux calculator: the
    autotuner's stage kernels, K timed steps, the face kernel's roofline by HIP events."""
    from gdtk_amd import Simulation, cases
    case = cases.mach6_box(nic, njc, nkc, seed=1234)
    sim = Simulation(3, flux=flux)
    sim.add_block(case)
    tune = sim.autotune()
    if warmup:
        sim.run(warmup)
    t0 = time.perf_counter()
    _, st = sim.run(steps)
    el = time.perf_counter() - t0
    sim.profile(True)
    sim.run(min(10, steps))
    ks = sim.kernel_stats()
    sim.profile(False)
    dom = max(("face_flux", "fused_stage"), key=lambda k: ks[k]["ms"])
    kd = ks[dom]
    avg_ms = kd["ms"] / max(kd["launches"], 1)
    fc = 1 if flux == "efm" else 0
    if dom == "face_flux":
        bpu, unit_name, kname = face_bytes_per_face(3), "face", "k_face_flux_batch<3,%d,false>" % fc
    else:
        bpu, unit_name, kname = stage_bytes_per_cell(3), "cell", "k_fused_batch<3,%d,*>" % fc
    achieved = bpu * kd["items_per_launch"] / (avg_ms * 1e-3) / 1e9
    out = {
        "value": case.ncells * steps / el, "unit": "cell-updates/s", "steps": steps, "warmup": warmup,
        "ms_per_step": 1e3 * el / steps,
        "config": {"workload": f"Mach-6 3-D box {nic}x{njc}x{nkc}, ideal air, l2r2+van Albada, {flux}, pc RK2",
                   "cells": case.ncells, "flux": flux, "stage_kernels": tune["best"], "final_dt": st["dt"]},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": measured_traffic(kname, kd["items_per_launch"]),
                     "kernel": kname, "avg_launch_ms": avg_ms, "bytes_per_" + unit_name: bpu,
                     "items_per_launch": kd["items_per_launch"]},
        "whole_step_hbm_frac": stage_bytes_per_cell(3) * 2 * case.ncells / (1e-3 * el / steps) / 1e9 / HBM_PEAK_GBS,
        "kernels": {k: {"launches": v["launches"], "avg_ms": v["ms"] / max(v["launches"], 1)} for k, v in ks.items()
                    if v["launches"]},
    }
    sim.close()
    if with_cpu:
        try:
            out["cpu_baseline"] = cpu_baseline(cpu_seconds, flux)
        except Exception as e:  # the baseline is reported, never required
            out["cpu_baseline"] = {"error": str(e)}
    if log:
        log(f"box {flux}: {out['value']:.3e} cell-updates/s, face kernel {avg_ms * 1e3:.0f} us "
            f"({out['roofline']['frac']:.3f} of HBM peak)")
    return out


def bench_sod(steps=500, cpu_seconds=10.0, log=None):
    """The plain method, small: the 2-D Sod tube of 200x100 cells, AUSMDV, CFL 0.5, 500 steps
    from the initial split on the GPU; the oracle on one thread (the reference's mode for one
    block) over a bounded number of steps of the same run."""
    from gdtk_amd import Simulation, cases
    from oracle import oracle as O
    case = cases.sod_2d(200, 100)
    sim = Simulation(2, flux="ausmdv")
    sim.add_block(case)
    t0 = time.perf_counter()
    _, st = sim.run(steps)
    el = time.perf_counter() - t0
    sim.close()
    o = O.Sim(O.IdealGas(), 2, "ausmdv")
    o.add_block(case)
    n, t1 = 0, time.perf_counter()
    while n < steps and time.perf_counter() - t1 < cpu_seconds:
        o.run(10)
        n += 10
    cel = time.perf_counter() - t1
    o.close()
    out = {"value": case.ncells * steps / el, "unit": "cell-updates/s", "steps": steps, "ms_per_step": 1e3 * el / steps,
           "config": {"workload": "2-D Sod tube 200x100, ideal air, l2r2+van Albada, ausmdv, pc RK2, CFL 0.5, "
                                  "from the initial split", "cells": case.ncells, "final_time": st["time"],
                      "final_dt": st["dt"]},
           "cpu_baseline": {"value": case.ncells * n / cel, "unit": "cell-updates/s", "cores": 1, "kind": "port",
                            "sample": f"oracle, 1 thread, the first {n} of the {steps} steps, {cel:.1f} s"}}
    if log:
        log(f"sod: {out['value']:.3e} cell-updates/s on the GPU (latency-bound grid), oracle "
            f"{out['cpu_baseline']['value']:.3e}")
    return out


def _batched_slab_rate(args):
    """Worker of the batched CPU baseline: the oracle on one thread over an 8-layer slab of
    one block of the batched case, for about `seconds`."""
    nblocks, q, seconds = args
    sys.path.insert(0, ROOT)
    from gdtk_amd import cases
    from oracle import oracle as O
    b = cases.batched_ragged(nblocks, only=q)[0]
    slab = cases.k_slab(b, min(8, b.nkc))
    o = O.Sim(O.IdealGas(), 3, "ausmdv")
    o.add_block(slab)
    o.run(1)
    n, t0 = 0, time.perf_counter()
    while time.perf_counter() - t0 < seconds and n < 200:
        o.run(1)
        n += 1
    el = time.perf_counter() - t0
    o.close()
    return slab.ncells * n / el


def batched_cpu_baseline(nblocks, seconds=10.0):
    """The oracle over the batched blocks in the reference's mode (one thread per block,
    threads = min(host cores, blocks)): as many worker processes, each timing an 8-layer slab
    of its own block; the rates add up."""
    import multiprocessing as mp
    nproc = max(1, min(16, os.cpu_count() or 1, nblocks))
    qs = [int(q) for q in range(0, nblocks, max(1, nblocks // nproc))][:nproc]
    with mp.get_context("spawn").Pool(nproc) as pool:
        rates = pool.map(_batched_slab_rate, [(nblocks, q, seconds) for q in qs])
    return dict(value=float(sum(rates)), unit="cell-updates/s", cores=nproc, kind="port",
                sample=f"oracle, {nproc} processes of one thread, each an 8-layer k-slab of one block "
                       f"(blocks {qs[0]}..{qs[-1]} step {max(1, nblocks // nproc)}), about {seconds:.0f} s each")


def _leg_summary(r):
    """Compact figures of a sub-leg record."""
    out = {}
    for k in ("value", "unit", "ms_per_step"):
        if k in r:
            out[k] = r[k]
    if "roofline" in r:
        out["roofline"] = {k: r["roofline"][k] for k in ("kernel", "bound", "frac") if k in r["roofline"]}
    cb = r.get("cpu_baseline")
    if isinstance(cb, dict) and "value" in cb:
        out["cpu_baseline"] = {"value": cb["value"], "cores": cb["cores"]}
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--spinup", type=int, default=1500,
                    help="untimed steps of the headline cylinder from the free stream (set-up)")
    ap.add_argument("--size", default="256,128,128", help="cells of the Mach-6 box sub-legs")
    ap.add_argument("--box-steps", type=int, default=20)
    ap.add_argument("--copy-probe", action="store_true",
                    help="run only a device copy of known size (HBM counter calibration)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--headline-only", action="store_true", help="measure only the headline coupled cylinder")
    ap.add_argument("--box-only", action="store_true", help="measure only the Mach-6 box (AUSMDV and EFM)")
    ap.add_argument("--flux", default="ausmdv", choices=["ausmdv", "efm"], help="flux calculator of --box-only")
    ap.add_argument("--no-sod", action="store_true", help="skip the Sod tube")
    ap.add_argument("--chem-cells", type=int, default=1 << 22,
                    help="cells of the chemistry workload (0 skips the chemistry measurement)")
    ap.add_argument("--chem-steps", type=int, default=5)
    ap.add_argument("--chem-only", action="store_true", help="measure only the chemistry workload")
    ap.add_argument("--no-surrogate", action="store_true", help="skip the learned-surrogate measurement")
    ap.add_argument("--surrogate-train", type=int, default=1 << 22, help="integrator samples to train on")
    ap.add_argument("--batched", type=int, default=64,
                    help="blocks of the batched uneven-blocks workload (0 skips it)")
    ap.add_argument("--batched-steps", type=int, default=20)
    ap.add_argument("--batched-only", action="store_true", help="measure only the batched uneven-blocks workload")
    ap.add_argument("--coupled-steps", type=int, default=5000,
                    help="steps of the cylinder's integrator and surrogate runs (0 skips them)")
    ap.add_argument("--coupled-only", action="store_true", help="measure only the cylinder's surrogate comparison")
    ap.add_argument("--adaptive-steps", type=int, default=20,
                    help="steps of the adaptive EFM/AUSMDV run on the Mach-6 box (0 skips it)")
    ap.add_argument("--adaptive-only", action="store_true", help="measure only the adaptive flux calculator")
    ap.add_argument("--lmr-steps", type=int, default=20,
                    help="steps of the Eilmer 5 (lmr) numerics run on the Mach-6 box (0 skips it)")
    ap.add_argument("--lmr-only", action="store_true", help="measure only the Eilmer 5 (lmr) numerics")
    ap.add_argument("--no-flow-io", action="store_true", help="skip the flow-file write/read measurement")
    ap.add_argument("--no-n90", action="store_true", help="skip the reference's n90 reacting cylinder case")
    ap.add_argument("--n90-only", action="store_true", help="run only the n90 case")
    ap.add_argument("--flow-io-only", action="store_true", help="measure only the flow-file write/read")
    ap.add_argument("--tt-cells", type=int, default=1 << 18,
                    help="cells of the two-temperature reactor leg (0 skips it)")
    ap.add_argument("--tt-only", action="store_true", help="measure only the two-temperature reactor")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU only (asked for {args.gpus}); "
              "the multi-GPU sharding is described in DESIGN.md", file=sys.stderr)
        sys.exit(2)
    if args.copy_probe:
        copy_probe()
        return

    def log(msg):
        print("bench:", msg, file=sys.stderr, flush=True)

    def emit(rec):
        print(json.dumps(rec, default=float), flush=True)

    with_cpu = not args.no_cpu_baseline
    nic, njc, nkc = (int(x) for x in args.size.split(","))
    nblk = args.batched

    def chem_legs():
        res = bench_chem(args.chem_cells, args.chem_steps, 1, with_cpu)
        if not args.no_surrogate:
            sres = bench_surrogate(args.chem_cells, args.chem_steps, n_train=args.surrogate_train, log=log)
            sres["speedup_vs_alpha_qss"] = (sres["value"] / res["value"]) if sres["value"] else None
            res["surrogate"] = sres
        return res

    if args.chem_only:
        return emit(chem_legs())
    if args.coupled_only:
        return emit(bench_coupled(args.coupled_steps, log=log))
    if args.batched_only:
        r = bench_batched(nblk, args.batched_steps, 2, "ausmdv", log)
        if with_cpu:
            r["cpu_baseline"] = batched_cpu_baseline(nblk)
        return emit(r)
    if args.n90_only:
        return emit(bench_n90(300 if with_cpu else 0, log=log))
    if args.flow_io_only:
        return emit(bench_flow_io(nic, njc, nkc, log=log))
    if args.tt_only:
        return emit(bench_two_temperature(args.tt_cells, with_cpu=with_cpu, log=log))
    if args.lmr_only:
        return emit(bench_lmr(nic, njc, nkc, args.lmr_steps, log=log))
    if args.adaptive_only:
        return emit(bench_adaptive(nic, njc, nkc, args.adaptive_steps, log=log))
    if args.box_only:
        return emit(bench_box(nic, njc, nkc, args.flux, args.steps, args.warmup, with_cpu, log=log))

    out = bench_coupled_headline(args.steps, args.warmup, args.spinup, with_cpu, log=log)
    log(f"headline: {out['value']:.4e} cell-updates/s, {out['ms_per_step']:.3f} ms/step, "
        f"dominant {out['roofline']['kernel']} at {out['roofline']['frac']:.3f} of peak")
    if args.headline_only:
        return emit(out)
    legs = {}
    if not args.no_sod:
        legs["sod_200x100"] = bench_sod(log=log)
    legs["box_ausmdv"] = bench_box(nic, njc, nkc, "ausmdv", args.box_steps, 2, with_cpu, log=log)
    legs["box_efm"] = bench_box(nic, njc, nkc, "efm", args.box_steps, 2, with_cpu, cpu_seconds=8.0, log=log)
    if nblk > 0:
        r = bench_batched(nblk, args.batched_steps, 2, "ausmdv", log)
        if with_cpu:
            try:
                r["cpu_baseline"] = batched_cpu_baseline(nblk)
            except Exception as e:  # the baseline is reported, never required
                r["cpu_baseline"] = {"error": str(e)}
        legs["batched"] = r
    if args.chem_cells > 0:
        legs["chem"] = chem_legs()
    if args.coupled_steps > 0:
        legs["cylinder_surrogate"] = bench_coupled(args.coupled_steps, log=log)
    if args.adaptive_steps > 0:
        legs["adaptive"] = bench_adaptive(nic, njc, nkc, args.adaptive_steps, log=log)
    if args.lmr_steps > 0:
        legs["lmr"] = bench_lmr(nic, njc, nkc, args.lmr_steps, log=log)
    if args.tt_cells > 0:
        legs["two_temperature"] = bench_two_temperature(args.tt_cells, with_cpu=with_cpu, log=log)
    if not args.no_flow_io:
        legs["flow_io"] = bench_flow_io(nic, njc, nkc, log=log)
    if not args.no_n90:
        legs["n90"] = bench_n90(300 if with_cpu else 0, log=log)
    out["legs"] = legs
    # compact per-leg figures last, so that they survive a truncated head of the line
    summ = {k: _leg_summary(v) for k, v in legs.items()}
    if "chem" in legs and "surrogate" in legs["chem"]:
        su = legs["chem"]["surrogate"]
        summ["chem_surrogate"] = {"value": su["value"], "meets_target": su["meets_target"],
                                  "speedup_vs_alpha_qss": legs["chem"]["surrogate"]["speedup_vs_alpha_qss"],
                                  "nearest": {k: su["nearest"][k] for k in ("disagree_tol", "max_abs_dY",
                                                                              "max_rel_T_err", "value")}}
    if "cylinder_surrogate" in legs:
        cyl = legs["cylinder_surrogate"]
        summ["cylinder_surrogate"] = {"meets_target": cyl["meets_target"],
                                      "points": [{"disagree_tol": r["disagree_tol"],
                                                  "chem_speedup_vs_alpha_qss": r["chem_speedup_vs_alpha_qss"],
                                                  "max_abs_dY": r["per_result"]["max_abs_dY"],
                                                  "max_rel_T_err": r["per_result"]["max_rel_T_err"],
                                                  "T_stag_rel_diff": r["T_stag_rel_diff"]}
                                                 for r in cyl["surrogate"]]}
    out["summary"] = {"headline": {"value": out["value"], "ms_per_step": out["ms_per_step"],
                                   "roofline": {k: out["roofline"][k] for k in ("kernel", "bound", "frac")},
                                   "cpu_baseline": out.get("cpu_baseline", {}).get("value")}, **summ}
    emit(out)


if __name__ == "__main__":
    main()
