#!/usr/bin/env python3
"""Benchmark of the gfx950 VL+CT step: zone-updates/s on one MI355X.

Workload (default): plain FP64 VL+CT on the 3D Orszag-Tang vortex in one
periodic 256^3 block with 3 ghost zones (gamma = 5/3, PLM with theta = 2,
HLLD, CT, CFL 0.4).  A step is one Cello cycle of the method on the
block: the CFL reduction (timestep), the periodic ghost refresh and the
two-stage VL+CT compute.  The timed loop uses the fused call
(enzoe_vlct_compute_timestep): each step refreshes, computes and leaves the
next cycle's CFL dt, taken inside the final cell update (bit for bit the
separate timestep's).  The state lives in HBM for the whole run.

Prints one JSON line (metric, value, roofline, cpu_baseline, ...).
"""
import argparse
import ctypes
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "enzo-e_amd"))

import numpy as np  # noqa: E402

import enzoe_vlct as ev  # noqa: E402
import problems  # noqa: E402

HBM_PEAK_GBS = 8000.0
# Algorithmic HBM bytes per zone-update of a fused FP64 MHD VL+CT step:
# stage 1 reads and writes the 11-word state (8 cell + 3 face words), stage 2
# reads the half-step state and the start-of-step state and writes the new
# state: 55 words = 440 B.
STEP_BYTES_PER_ZONE_FP64 = 440.0
# The mixed step stores its half-step state as 8 FP32 primitives and the 3
# FP64 face fields (56 B instead of 88): 88 + 56 + 56 + 88 + 88 = 376 B.
STEP_BYTES_PER_ZONE_MIXED = 376.0
STEP_BYTES = {("mhd", "fp64"): STEP_BYTES_PER_ZONE_FP64, ("mhd", "mixed"): STEP_BYTES_PER_ZONE_MIXED,
              ("mhd", "fp64_fma"): STEP_BYTES_PER_ZONE_FP64,
              # hydro: 5 words per state, FP64 25 words; mixed half-step 5 x 4 B
              ("hd", "fp64"): 200.0, ("hd", "mixed"): 40.0 + 20.0 + 20.0 + 40.0 + 40.0,
              ("hd", "fp64_fma"): 200.0}
# Algorithmic bytes per cell of one flux launch: it must read the 11-word
# state of the cell and write 7 flux components on each of its 3 low faces.
FLUX_BYTES_PER_CELL = {0: (11 + 21) * 8.0, 1: 11 * 8.0 + 21 * 4.0, 2: (11 + 21) * 8.0}
# hydro: read 5 words, write 5 fluxes on 3 faces; step: 25 words per zone
FLUX_BYTES_PER_CELL_HD = {0: (5 + 15) * 8.0, 1: 5 * 8.0 + 15 * 4.0, 2: (5 + 15) * 8.0}
STEP_BYTES_PER_ZONE_HD = 200.0
WORKLOADS = ["orszag_tang", "fast_wave", "batched_amr", "turbulence", "smr_flux_correct",
             "kelvin_helmholtz", "output", "drop_in"]


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--n", type=int, default=256, help="active cells per axis")
    # fp64_fma: FP64 with multiply-adds contracted (not bit-identical to the
    # reference; held to 1e-12 relative per cell); Orszag-Tang only
    ap.add_argument("--precision", choices=["fp64", "mixed", "fp64_fma"], default="fp64")
    # launch structure of the stages (enzoe_vlct_set_fused_ct: 0 separate
    # launches, 1-4 fused CT, 5 one persistent launch per stage) for the
    # Orszag-Tang block (its primary run), the fast-wave and batched-AMR
    # workloads; -1: the autotuner's choice
    ap.add_argument("--structure", type=int, default=-1)
    # shape of the flux launch (enzoe_vlct_set_flux_shape: 0 one thread per
    # cell, 1 one face per thread, 2 z pencils for the nearest-neighbour
    # stage) for the Orszag-Tang run; -1: the autotuner's choice
    ap.add_argument("--flux-shape", type=int, default=-1)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-compare", action="store_true",
                    help="skip the run at the other precision")
    ap.add_argument("--cpu-n", type=int, default=128,
                    help="edge of the Orszag-Tang block timed on the CPU")
    ap.add_argument("--cpu-steps", type=int, default=30)
    ap.add_argument("--workload", choices=WORKLOADS, default="orszag_tang",
                    help="orszag_tang (default): one 256^3 block; fast_wave: "
                         "32^3 fast eigenmode for one period; batched_amr: "
                         "512/1536/3000 blocks of 32^3 on 3 levels with 2:1 "
                         "sub-cycling; turbulence: 512 blocks of 64^3, Mach 5; "
                         "kelvin_helmholtz: 512 blocks of 128^3, hydro; "
                         "smr_flux_correct: 64 x (1 coarse + 4 fine) 32^3 hydro "
                         "blocks with coarse/fine ghost transfers and flux "
                         "corrections; output: HDF5 block output and restart "
                         "input of the 256^3 Orszag-Tang state; drop_in: the C++ "
                         "layer's Method::timestep + Method::compute on the "
                         "Orszag-Tang block, fused and stage by stage")
    ap.add_argument("--batched-steps", type=int, default=50,
                    help="steps of the batched mixed-precision leg (turbulence, 512 "
                         "blocks of 64^3) of the default run")
    ap.add_argument("--dual-energy", action="store_true",
                    help="orszag_tang with the dual-energy formalism on (eta 0.001) "
                         "as the measured step (for profiling that path)")
    ap.add_argument("--launch-mask", type=int, default=-1,
                    help="orszag_tang: fixed launch order of the stage kernels (bit mask "
                         "k_flux 1, k_edge 2, k_faces 4, k_cells 8 in the tiled "
                         "XCD-contiguous order) instead of the autotuner's choice")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run the plain copy kernel (HBM counter scale)")
    return ap.parse_args()


def make_state(n, method):
    blk, cw = problems.orszag_tang(n)
    if method.dual_energy:
        # the specific internal energy of the state (setup_eint_,
        # EnzoInitialInclinedWave.cpp:495-538); ghost zones come from the refresh
        rho = blk["density"]
        v2 = sum(np.square(blk[k]) for k in ("velocity_x", "velocity_y", "velocity_z"))
        b2 = sum(np.square(blk[k]) for k in ("bfield_x", "bfield_y", "bfield_z"))
        blk["internal_energy"] = blk["total_energy"] - 0.5 * v2 - np.divide(
            0.5 * b2, rho, out=np.zeros_like(rho), where=rho > 0)
    grid = ev.DeviceGrid(method, (1, 1, 1), (n, n, n), [cw])
    grid.upload_from([blk])
    grid.refresh()
    ev.synchronize()
    return grid


def step(grid):
    m = grid.method
    dt = m.timestep(grid.blocks)
    grid.refresh()
    m.compute(grid.blocks, dt)
    return dt


def fused_step(grid, dt):
    """One cycle with the CFL reduction of the next one fused into the
    final cell update: refresh, compute with `dt`, return the next dt."""
    grid.refresh()
    return grid.method.compute_timestep(grid.blocks, dt)


COPY_WORDS = 1 << 26  # 512 MiB read + 512 MiB written per copy launch


def copy_bandwidth(reps=5):
    """Achievable HBM copy rate: a plain one-word-per-thread copy kernel of
    known size (also the scale for the FETCH_SIZE / WRITE_SIZE counters)."""
    nbytes = COPY_WORDS * 8
    src, dst = ev.DeviceBuffer(nbytes), ev.DeviceBuffer(nbytes)
    ev.hip_check(ev.hip().hipMemset(src.ptr, 0, ctypes.c_size_t(nbytes)),
                 "hipMemset")
    L = ev.lib()
    L.enzoe_calibration_copy.argtypes = [ctypes.c_void_p, ctypes.c_void_p,
                                         ctypes.c_size_t, ctypes.c_void_p]
    ev.check(L.enzoe_calibration_copy(dst.ptr, src.ptr, COPY_WORDS, None),
             "copy")
    ev.synchronize()
    ev.profiling(True)
    for _ in range(reps):
        ev.check(L.enzoe_calibration_copy(dst.ptr, src.ptr, COPY_WORDS, None),
                 "copy")
    ev.synchronize()
    ms = ev.profile_report()["k_calibration_copy"][0]
    ev.profiling(False)
    src.free()
    dst.free()
    return {"kernel": "k_calibration_copy", "bytes_read": nbytes,
            "bytes_written": nbytes, "avg_ms": ms,
            "GBs": 2 * nbytes / (ms * 1e-3) / 1e9}


# kernel launches of one VL+CT step (refresh, two stages, CFL fused into the
# final cell update)
STEP_LAUNCHES = {"k_refresh": 1, "k_flux_nn": 1, "k_flux_plm": 1, "k_edge": 2,
                 "k_faces": 2, "k_cells": 2}


def measured_step_traffic(workload, mhd=True):
    """HBM bytes of one step from the committed counter summary: its
    step_traffic line, else the sum of the step's launches, or None when a
    kernel of the step has no entry."""
    path = os.path.join(ROOT, "profiles", "hbm_traffic.txt")
    if os.path.exists(path):
        with open(path) as f:
            for line in f:
                p = line.split()
                if len(p) == 3 and p[0] == "step_traffic" and p[1] == workload:
                    return float(p[2])
    tot = 0.0
    for k, n in STEP_LAUNCHES.items():
        if not mhd and k in ("k_edge", "k_faces"):
            continue
        t = measured_traffic(k, workload)
        if t is None:
            return None
        tot += n * t
    return tot


def step_roofline(prec_name, zones, ms_per_step, gpu_ms_per_step, workload, mhd=True):
    """The step against the HBM roofline as the metric defines it: the
    algorithmic bytes of a fully fused step (440 B per MHD zone-update in
    FP64) over the measured wall time per step."""
    bpz = STEP_BYTES[("mhd" if mhd else "hd", prec_name)]
    achieved = bpz * zones / (ms_per_step * 1e-3) / 1e9
    return {"bound": "hbm", "scope": "step", "achieved": achieved, "peak": HBM_PEAK_GBS,
            "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
            "traffic": measured_step_traffic(workload, mhd),
            "bytes_per_zone": bpz, "algorithmic_bytes_per_step": bpz * zones,
            "kernel_ms_per_step": gpu_ms_per_step,
            "achieved_over_kernel_time": bpz * zones / (gpu_ms_per_step * 1e-3) / 1e9}


def measured_traffic(kernel, workload):
    """HBM bytes per launch of `kernel` on `workload` from the committed
    rocprofv3 counter summary (FETCH_SIZE + WRITE_SIZE, scaled by the copy
    kernel's reading), or None when no summary covers it."""
    path = os.path.join(ROOT, "profiles", "hbm_traffic.txt")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        for line in f:
            p = line.split()
            if len(p) == 4 and p[0] == "traffic" and p[1] == kernel \
                    and p[2] == workload:
                return float(p[3])
    return None


# wave64 FP64 (or non-packed FP32) VALU instructions MI355X issues per
# second: 256 CUs x 4 SIMDs x 2.4 GHz / 4 cycles per instruction
VALU_PEAK_WAVE_INSTR = 256 * 4 * 2.4e9 / 4


def measured_valu(kernel, workload, launch_ms):
    """The flux kernel's issue rate against the VALU peak: wave-instructions
    per launch from the committed rocprofv3 summary (SQ_INSTS_VALU) over the
    launch time measured here, or None when no summary covers it."""
    path = os.path.join(ROOT, "profiles", "valu_k_flux.txt")
    if not os.path.exists(path) or not launch_ms or launch_ms != launch_ms:
        return None
    with open(path) as f:
        for line in f:
            p = line.split()
            if len(p) == 4 and p[0] == "valu" and p[1] == kernel and p[2] == workload:
                rate = float(p[3]) / (launch_ms * 1e-3)
                return {"wave_instructions_per_launch": float(p[3]), "achieved": rate,
                        "peak": VALU_PEAK_WAVE_INSTR, "unit": "wave-instructions/s",
                        "frac": rate / VALU_PEAK_WAVE_INSTR}
    return None


def _oracle_rate(n, nsteps, threads):
    """Zone-updates/s of the oracle on an n^3 Orszag-Tang block."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle as po
    cfg = po.make_config(riemann="hlld", recon=("nn", "plm"), theta=2.0,
                         mhd=True, courant=0.4)
    blk, cw = problems.orszag_tang(n)
    grid = po.Grid(cfg, (1, 1, 1), (n, n, n), (0, 0, 0), (1.0, 1.0, 1.0))
    view = grid.block_arrays(0)
    for k, v in blk.items():
        view[k][...] = v
    grid.refresh()
    grid.cycle(0.0, 1e9, nthreads=threads)  # warm the scratch
    t0 = time.perf_counter()
    t = 0.0
    for _ in range(nsteps):
        t += grid.cycle(t, 1e9, nthreads=threads)
    el = time.perf_counter() - t0
    return n ** 3 * nsteps / el, el


def cpu_baseline(n, nsteps):
    """The oracle (plain-C restatement of the reference loops, OpenMP over
    the z loops) on the host cores, and on one core on a smaller sample."""
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    rate, el = _oracle_rate(n, nsteps, threads)
    n1, s1 = 64, 10
    rate1, el1 = _oracle_rate(n1, s1, 1)
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"value": rate, "unit": "zone-updates/s",
            "cores": threads, "kind": "port", "cpu_model": model,
            "host_cpus": os.cpu_count(),
            "sample": "oracle (C restatement of the reference loops, OpenMP "
                      "over z) on the same Orszag-Tang problem at %d^3, %d "
                      "steps, %.1f s" % (n, nsteps, el),
            "one_core": {"value": rate1, "cores": 1,
                         "sample": "same at %d^3, %d steps, %.1f s" % (n1, s1, el1)}}


def cpu_baseline_sample(builder, n, mhd, gamma, riemann, nsteps):
    """The oracle on a sample of a batched workload's own blocks: the first
    2x2x2 blocks of the domain as a periodic grid of their own (ghost zones
    refreshed within the sample), `nsteps` cycles with the CFL dt, OpenMP
    over the z loops of each block on the host cores."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle as po
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    cfg = po.make_config(gamma=gamma, riemann=riemann, recon=("nn", "plm"), theta=2.0,
                         mhd=mhd, courant=0.4)
    order = [(bx, by, bz) for bz in range(2) for by in range(2) for bx in range(2)]
    blk0, cw = builder(order[0])
    grid = po.Grid(cfg, (2, 2, 2), (n, n, n), (0.0, 0.0, 0.0),
                   tuple(2 * n * c for c in cw))
    for i, ib in enumerate(order):
        arrs = blk0 if i == 0 else builder(ib)[0]
        view = grid.block_arrays(i)
        for k, v in arrs.items():
            if k in view:
                view[k][...] = v
    grid.refresh()
    grid.cycle(0.0, 1e9, nthreads=threads)  # warm the scratch
    t0 = time.perf_counter()
    t = 0.0
    for _ in range(nsteps):
        t += grid.cycle(t, 1e9, nthreads=threads)
    el = time.perf_counter() - t0
    return {"value": 8 * n ** 3 * nsteps / el, "unit": "zone-updates/s", "cores": threads,
            "kind": "port",
            "sample": "oracle (C restatement of the reference loops, OpenMP over z) on "
                      "2x2x2 of the workload's %d^3 blocks, %d steps, %.1f s"
                      % (n, nsteps, el)}


METRIC = "zone-updates/s/GPU (batched FP32-mixed vs plain FP64 VL+CT) + HBM GB/s"


PRECISIONS = {"fp64": 0, "mixed": 1, "fp64_fma": 2}  # ENZOE_PREC_*


def timed_run(precision, n, steps, warmup, dual_energy=False, launch_mask=-1, keep=False,
              structure=-1, flux_shape=-1):
    """Warm-up, then `steps` timed steps with per-kernel HIP-event timing.
    dual_energy: the "modern" dual-energy formalism (eta 0.001) on.
    launch_mask >= 0: that launch order instead of the autotuner's.
    keep: the result holds the grid ("grid", "method") after the run.
    structure >= 0: that launch structure (enzoe_vlct_set_fused_ct) instead
    of the autotuner's."""
    prec = PRECISIONS[precision]
    if launch_mask >= 0:
        os.environ["ENZOE_XCD_REMAP"] = str(launch_mask)
    method = ev.VlctMethod(mhd=True, riemann="hlld", reconstruct="plm",
                           theta=2.0, courant=0.4, gamma=5.0 / 3.0,
                           precision=prec,
                           dual_energy_eta=0.001 if dual_energy else None)
    if structure >= 0:
        method.set_fused_ct(structure)
    if flux_shape >= 0:
        method.set_flux_shape(flux_shape)
    grid = make_state(n, method)
    dt = method.timestep(grid.blocks)  # the first cycle's CFL reduction
    for _ in range(warmup):
        dt = fused_step(grid, dt)
    ev.synchronize()
    # launch order of the stage kernels for this block shape (timed on a
    # copy of the block, outside the timed region)
    if launch_mask >= 0:
        mask = launch_mask
    else:
        mask = method.autotune(grid.blocks, dt)
    choice = method.launch_choice((n + 6,) * 3)
    ev.profiling(True)
    ev.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        dt = fused_step(grid, dt)
    ev.synchronize()
    el = time.perf_counter() - t0
    prof = ev.profile_report()
    ev.profiling(False)
    kept = {"grid": grid, "method": method} if keep else {}
    del grid, method
    zones = n ** 3
    m = n + 6
    # dominant kernel: the flux kernel of the full (PLM) stage
    dom = "k_flux_plm"
    dom_ms = prof.get(dom, (float("nan"), 0))[0]
    cells = (m - 4) ** 3  # flux box of the PLM stage: [2, m-2)^3
    dom_bytes = FLUX_BYTES_PER_CELL[prec] * cells
    gpu_step_ms = sum(v[0] * v[1] for v in prof.values()) / steps
    return {"precision": precision, "value": zones * steps / el,
            "ms_per_step": el * 1e3 / steps, "prof": prof, "dom": dom,
            "dom_ms": dom_ms, "dom_bytes": dom_bytes,
            "achieved": dom_bytes / (dom_ms * 1e-3) / 1e9,
            "gpu_step_ms": gpu_step_ms,
            "step_gbs": STEP_BYTES_PER_ZONE_FP64 * zones / (gpu_step_ms * 1e-3) / 1e9,
            "workload": "orszag_tang_%d3_%s%s_vlct_plm_hlld" % (
                n, precision, "_dual_energy" if dual_energy else ""),
            "launch_mask": mask, "launch_choice": choice, **kept}


def _roofline(prof, steps, cells_per_step, prec, mhd, workload):
    """Dominant-kernel roofline: algorithmic flux-kernel bytes of a step
    over the flux kernel's HIP-event time in that step."""
    dom = "k_flux_plm"
    avg_ms, count = prof.get(dom, (float("nan"), 0))
    per_step_ms = avg_ms * count / steps
    table = FLUX_BYTES_PER_CELL if mhd else FLUX_BYTES_PER_CELL_HD
    nbytes = table[prec] * cells_per_step
    achieved = nbytes / (per_step_ms * 1e-3) / 1e9
    return {"bound": "hbm", "kernel": dom, "achieved": achieved,
            "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
            "traffic": measured_traffic(dom, workload),
            "bytes_per_step": nbytes, "kernel_ms_per_step": per_step_ms}


def _both_rooflines(prof, steps, elapsed, zones_per_step, cells_per_step, precision, mhd,
                    workload):
    """The step against the HBM roofline (algorithmic bytes per zone-update
    over the wall time per step), with the flux kernel's own figure nested."""
    prec = ev.ENZOE_PREC_DOUBLE if precision == "fp64" else ev.ENZOE_PREC_MIXED
    gpu_ms = sum(v[0] * v[1] for v in prof.values()) / steps
    return dict(step_roofline(precision, zones_per_step, elapsed * 1e3 / steps, gpu_ms,
                              workload, mhd),
                dominant_kernel=_roofline(prof, steps, cells_per_step, prec, mhd, workload))


def _profiled(fn, steps, warmup):
    for _ in range(warmup):
        fn()
    ev.synchronize()
    ev.profiling(True)
    ev.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        fn()
    ev.synchronize()
    el = time.perf_counter() - t0
    prof = ev.profile_report()
    ev.profiling(False)
    return el, prof


def run_fast_wave(a):
    """Config 1: 32^3 fast eigenmode along (1,1,1), evolved one period."""
    n = 32
    blk, cw, cf, period = problems.fast_wave_111(n)
    out = {}
    for precision in ("fp64", "mixed"):
        prec = ev.ENZOE_PREC_DOUBLE if precision == "fp64" else ev.ENZOE_PREC_MIXED
        method = ev.VlctMethod(mhd=True, riemann="hlld", reconstruct="plm",
                               theta=2.0, courant=0.4, precision=prec)
        grid = ev.DeviceGrid(method, (1, 1, 1), (n, n, n), [cw])
        grid.upload_from([blk])
        grid.refresh()
        ev.synchronize()
        if a.structure >= 0:
            method.set_fused_ct(a.structure)
        else:
            method.autotune(grid.blocks, method.timestep(grid.blocks))
        # eager cycles (host dt) on a copy, then the captured-graph cycles
        eager = ev.DeviceGrid(method, (1, 1, 1), (n, n, n), [cw])
        eager.upload_from([blk])
        eager.refresh()
        ev.synchronize()
        t0 = time.perf_counter()
        ncycles, _ = eager.evolve(0.0, period)
        ev.synchronize()
        el_eager = time.perf_counter() - t0
        t0 = time.perf_counter()
        nrep, t_end = grid.evolve_graph(0.0, period, ncycles=8)
        ev.synchronize()
        el_graph = time.perf_counter() - t0
        # the same device-driven cycles enqueued directly (no graph)
        direct = ev.DeviceGrid(method, (1, 1, 1), (n, n, n), [cw])
        direct.upload_from([blk])
        direct.refresh()
        ev.synchronize()
        t0 = time.perf_counter()
        ncalls, t_end_direct = direct.evolve_device(0.0, period, ncycles=8)
        ev.synchronize()
        el_direct = time.perf_counter() - t0
        assert t_end_direct == t_end
        el = min(el_graph, el_direct)
        fin = grid.blocks[0].download(["density", "velocity_x", "total_energy"])
        s = (slice(3, -3),) * 3
        l1 = {k: float(np.abs(fin[k][s] - blk[k][s]).mean()) for k in fin}
        method.check_persistent()
        out[precision] = {"value": n ** 3 * ncycles / el, "cycles": ncycles,
                          "seconds": el, "eager_seconds": el_eager,
                          "graph_seconds": el_graph, "direct_device_seconds": el_direct,
                          "driver": "graph" if el_graph <= el_direct else "direct",
                          "launch_choice": method.launch_choice((n + 6,) * 3),
                          "graph_replays_of_8_cycles": nrep, "t_end": t_end,
                          "l1_vs_initial": l1}
    return {
        "metric": METRIC, "value": out["fp64"]["value"], "unit": "zone-updates/s",
        "n_gpus": 1, "steps": out["fp64"]["cycles"], "warmup": 0,
        "ms_per_step": out["fp64"]["seconds"] * 1e3 / out["fp64"]["cycles"],
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64", "data": "synthetic (fast magnetosonic eigenmode, amplitude 1e-6)",
        "config": {"workload": "fast_wave_32_one_period", "block": [n, n, n],
                   "riemann": "hlld", "reconstruct": "plm", "theta": 2.0,
                   "courant": 0.4, "period": period, "fast_speed": cf},
        "precision_comparison": {"fp64": out["fp64"], "mixed": out["mixed"]},
    }


def run_batched_amr(a):
    """Config 3: 512 / 1536 / 3000 independent 32^3 blocks on three levels,
    2:1 sub-cycling (launches of 5048, 3000, 4536, 3000 blocks per coarse
    step), one batched launch sequence per substep with per-block dt, the
    whole coarse step (dt included) on the device and replayed from a HIP
    graph (enzoe_vlct_levels)."""
    import workloads
    res = {}
    for precision in (a.precision, "mixed" if a.precision == "fp64" else "fp64"):
        prec = ev.ENZOE_PREC_DOUBLE if precision == "fp64" else ev.ENZOE_PREC_MIXED
        method = ev.VlctMethod(mhd=True, riemann="hlld", reconstruct="plm",
                               theta=2.0, courant=0.4, precision=prec)
        lv = workloads.SubcycledLevels(method)
        if a.structure >= 0:
            method.set_fused_ct(a.structure)
        else:
            method.autotune(lv.levels[-1], method.timestep(lv.levels[-1]) / 4)
        # the levels prepared on the device: a coarse step is kernels only
        # (per-level CFL, the dt kernel, the substeps); per-kernel times from
        # eager steps, the rate from a captured graph of `steps` coarse steps
        plan = lv.device_levels()
        _, prof = _profiled(plan.step, a.steps, a.warmup)
        plan.capture(a.steps)
        plan.replay()  # first replay: the graph's upload
        ev.synchronize()
        t0 = time.perf_counter()
        plan.replay()
        ev.synchronize()
        el = time.perf_counter() - t0
        zones = lv.zones_per_coarse_step()
        cells = sum(34 ** 3 * c * 2 ** l for l, c in enumerate(lv.counts))
        res[precision] = {"value": zones * a.steps / el,
                          "ms_per_step": el * 1e3 / a.steps,
                          "kernels_ms": {k: round(v[0], 4) for k, v in sorted(prof.items())},
                          "roofline": _both_rooflines(prof, a.steps, el, zones, cells,
                                                      precision, True,
                                                      "batched_amr_%s" % precision),
                          "launch_blocks": lv.launch_sizes(),
                          "launch_choice": method.launch_choice((38, 38, 38)),
                          "driver": "HIP graph of %d device-driven coarse steps" % a.steps}
        plan.destroy()
        method.check_persistent()
        del lv, method, plan
        if a.no_compare:
            break
    r = res[a.precision]
    out = {
        "metric": METRIC, "value": r["value"], "unit": "zone-updates/s", "n_gpus": 1,
        "steps": a.steps, "warmup": a.warmup, "ms_per_step": r["ms_per_step"],
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64" if a.precision == "fp64" else "f32-riemann/f64-update",
        "data": "synthetic (seeded random smooth MHD blocks: lognormal rho, Mach 2, beta 1)",
        "config": {"workload": "batched_amr_32_512_1536_3000_%s" % a.precision,
                   "blocks_per_level": [512, 1536, 3000], "block": [32, 32, 32],
                   "subcycling": "2:1", "launch_blocks": r["launch_blocks"],
                   "launch_choice": r["launch_choice"]},
        "roofline": r["roofline"], "kernels_ms": r["kernels_ms"],
    }
    if len(res) == 2:
        f, m = res["fp64"], res["mixed"]
        out["precision_comparison"] = {"fp64_value": f["value"], "mixed_value": m["value"],
                                       "mixed_over_fp64": m["value"] / f["value"],
                                       "mixed_kernels_ms": m["kernels_ms"],
                                       "mixed_launch_choice": m["launch_choice"]}
    return out


def run_uniform(a):
    """Configs 4 and 5: periodic grids of 512 blocks (turbulence 64^3 MHD,
    Kelvin-Helmholtz 128^3 hydro); the primary precision is timed, the other
    one run beside it, and the relative L1 between them reported."""
    import workloads
    if a.workload == "turbulence":
        nb, n, mhd, gamma, rs = (8, 8, 8), 64, True, 5.0 / 3.0, "hlld"
        spec = problems.turbulence_spectrum()
        norm = problems.turbulence_vrms_norm(spec, mach=5.0)
        builder = lambda ib: problems.turbulence_block(spec, norm, ib, nb=nb, n=n)
        data = "synthetic (k^-2 solenoidal velocity k=1..8, Mach 5, beta 0.1)"
    else:
        nb, n, mhd, gamma, rs = (8, 8, 8), 128, False, 1.4, "hllc"
        builder = lambda ib: problems.kelvin_helmholtz_block(ib, nb=nb, n=n)
        data = "synthetic (Kelvin-Helmholtz slab, density ratio 2, 1e-2 v_y noise)"
    precs = [a.precision] + ([] if a.no_compare else
                             ["mixed" if a.precision == "fp64" else "fp64"])
    grids, res = {}, {}
    first = None
    for precision in precs:
        prec = ev.ENZOE_PREC_DOUBLE if precision == "fp64" else ev.ENZOE_PREC_MIXED
        method = ev.VlctMethod(mhd=mhd, riemann=rs, reconstruct="plm", theta=2.0,
                               courant=0.4, gamma=gamma, precision=prec)
        if first is None:
            g = workloads.uniform_grid(method, nb, n, builder)
            first = g
        else:
            g = ev.DeviceGrid(method, nb, (n, n, n), first.blocks[0].desc.cell_width[:],
                              bc=("periodic",) * 3)
            workloads.copy_grid_state(first, g)
        grids[precision] = g
    for precision in precs:
        g = grids[precision]
        prec = ev.ENZOE_PREC_DOUBLE if precision == "fp64" else ev.ENZOE_PREC_MIXED
        g.method.autotune(g.blocks, g.method.timestep(g.blocks))
        state = {"dt": g.method.timestep(g.blocks)}

        def one(g=g, state=state):
            state["dt"] = fused_step(g, state["dt"])
        el, prof = _profiled(one, a.steps, a.warmup)
        nblk = nb[0] * nb[1] * nb[2]
        cells = (n + 2) ** 3 * nblk
        res[precision] = {"value": n ** 3 * nblk * a.steps / el,
                          "ms_per_step": el * 1e3 / a.steps,
                          "kernels_ms": {k: round(v[0], 4) for k, v in sorted(prof.items())},
                          "roofline": _both_rooflines(prof, a.steps, el,
                                                      n ** 3 * nblk, cells, precision, mhd,
                                                      "%s_%s" % (a.workload, precision))}
    r = res[a.precision]
    cpu = None
    if not a.no_cpu_baseline:
        cpu = cpu_baseline_sample(builder, n, mhd, gamma, rs, 3 if n > 64 else 5)
    out = {
        "metric": METRIC, "value": r["value"], "unit": "zone-updates/s", "n_gpus": 1,
        "steps": a.steps, "warmup": a.warmup, "ms_per_step": r["ms_per_step"],
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64" if a.precision == "fp64" else "f32-riemann/f64-update",
        "data": data,
        "config": {"workload": "%s_%dx%d3_%s" % (a.workload, nb[0] * nb[1] * nb[2], n,
                                                 a.precision),
                   "blocks": list(nb), "block": [n, n, n], "riemann": rs,
                   "reconstruct": "plm", "theta": 2.0, "courant": 0.4, "gamma": gamma},
        "roofline": r["roofline"], "kernels_ms": r["kernels_ms"],
    }
    if len(res) == 2:
        f, m = res["fp64"], res["mixed"]
        names = (["density", "velocity_x", "velocity_y", "velocity_z", "total_energy"]
                 + (["bfield_x", "bfield_y", "bfield_z"] if mhd else []))
        l1, l1_mag = workloads.relative_l1(grids["mixed"], grids["fp64"], names)
        out["precision_comparison"] = {
            "fp64_value": f["value"], "mixed_value": m["value"],
            "mixed_over_fp64": m["value"] / f["value"],
            "mixed_kernels_ms": m["kernels_ms"], "fp64_kernels_ms": f["kernels_ms"],
            "relative_l1_mixed_vs_fp64": l1, "max_relative_l1": max(l1.values()),
            "l1_mixed_vs_fp64_by_vector_magnitude": l1_mag,
            "max_l1_by_vector_magnitude": max(l1_mag.values())}
    if cpu is not None:
        out["cpu_baseline"] = cpu
    return out


def run_drop_in(a):
    """The C++ drop-in layer (EnzoMethodMHDVlct of libenzoe_vlct_host.so)
    driven as Cello drives it, per cycle: the ghost refresh,
    Method::timestep on the block (its CFL pass, dt read back), then
    Method::compute.  device_fused = true runs the fused launch sequence;
    false runs the reference's own composition stage by stage through the
    plugin objects (reconstructor, Riemann solver, CT, integration update),
    every call one single-operation entry of the C ABI.  Beside them the
    library's own cycle, whose final cell update also takes the next dt.
    FP64, PLM + HLLD + CT, the Orszag-Tang block of --n cells."""
    n = a.n
    params = ("mhd_choice = constrained_transport; riemann_solver = hlld; "
              "reconstruct_method = plm; theta_limiter = 2.0; courant = 0.4; "
              "device_fused = %s")
    res = {}
    for mode in ("library", "host_fused", "host_staged"):
        method = ev.VlctMethod(mhd=True, riemann="hlld", reconstruct="plm", theta=2.0,
                               courant=0.4, gamma=5.0 / 3.0)
        grid = make_state(n, method)
        state = {"dt": method.timestep(grid.blocks)}
        if mode == "library":
            def one(grid=grid, state=state):
                state["dt"] = fused_step(grid, state["dt"])
        else:
            hm = ev.HostMethod(params % ("true" if mode == "host_fused" else "false"),
                               gamma=5.0 / 3.0)

            def one(grid=grid, hm=hm):
                grid.refresh()
                hm.compute(grid.blocks, hm.timestep(grid.blocks[0]))
        for _ in range(a.warmup):
            one()
        ev.synchronize()
        t0 = time.perf_counter()
        for _ in range(a.steps):
            one()
        ev.synchronize()
        el = time.perf_counter() - t0
        res[mode] = {"value": n ** 3 * a.steps / el, "ms_per_step": el * 1e3 / a.steps}
        print("%s: %.3e zone-updates/s" % (mode, res[mode]["value"]), file=sys.stderr,
              flush=True)
        del grid, method, one
    lib = res["library"]["value"]
    return {
        "metric": METRIC, "value": res["host_fused"]["value"], "unit": "zone-updates/s",
        "n_gpus": 1, "steps": a.steps, "warmup": a.warmup,
        "ms_per_step": res["host_fused"]["ms_per_step"], "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic (3D Orszag-Tang vortex, seeded 1e-3 v_z noise)",
        "config": {"workload": "drop_in_orszag_tang_%d3_fp64_vlct_plm_hlld" % n,
                   "block": [n, n, n], "driver": "Method::timestep + Method::compute "
                   "per cycle (C++ layer)"},
        "drop_in": {mode: dict(r, over_library=r["value"] / lib) for mode, r in res.items()},
    }


def run_smr_flux_correct(a):
    """Static mesh refinement with flux corrections (hydro, the reference's
    VL+CT + flux_correct pairing): 64 units of one coarse 32^3 block and the
    four fine 32^3 blocks across its high-x face (320 blocks, ghost depth 4).
    A step: the ghost zones on both sides of each refinement face (1280
    restrictions, then 1280 prolongations: one launch each), dt, the VL+CT
    step storing the final-stage face fluxes (one batched launch sequence),
    the 256 fine-to-coarse flux sends (one launch) and the 64 coarse-block
    corrections (one launch).  The plain step (dt and VL+CT only) is timed
    beside it."""
    import workloads
    n, g, units = 32, 4, 64
    names = ["density", "velocity_x", "velocity_y", "velocity_z", "total_energy"]
    comp = problems.two_level_composite(n, g)
    res = {}
    for store in (True, False):
        prec = ev.ENZOE_PREC_DOUBLE if a.precision == "fp64" else ev.ENZOE_PREC_MIXED
        method = ev.VlctMethod(mhd=False, riemann="hllc", reconstruct="plm", theta=2.0,
                               courant=0.4, gamma=5.0 / 3.0, density_floor=1e-30,
                               pressure_floor=1e-30, precision=prec)
        blocks, fds = [], []
        for _ in range(units):
            for (_, _, _, dx, arr) in comp:
                m = n + 2 * g
                db = ev.DeviceBlock(m, m, m, g, (dx, dx, dx), mhd=False)
                db.upload({k: v for k, v in arr.items() if k in db.shapes})
                blocks.append(db)
                if store:
                    fds.append(ev.FluxData((n, n, n)))
        links, entries, restricts, prolongs = [], [], [], []
        if store:
            for u in range(units):
                fines = {}
                for j, (_, child, _, _, _) in enumerate(comp[1:], start=1):
                    links.append((fds[5 * u + j], fds[5 * u], 0, 0, child))
                    fines[(child[1], child[2])] = blocks[5 * u + j]
                entries.append((blocks[5 * u], fds[5 * u], [0, 1, 0, 0, 0, 0]))
                r, p_ = workloads.composite_interface_transfers(blocks[5 * u], fines, n, g,
                                                                names)
                restricts += r
                prolongs += p_
            # tables built once, outside the timed steps
            restricts, prolongs = ev.transfer_table(restricts), ev.transfer_table(prolongs)
            links, entries = ev.FluxLinks(links), ev.FluxCorrections(entries)

        def one_step():
            if store:
                ev.transfer_batch(method, restricts)
                ev.transfer_batch(method, prolongs)
            dt = method.timestep(blocks)
            if store:
                ev.compute_store_fluxes(method, blocks, dt, fds)
                ev.flux_send_to_coarse_batch(method, links)
                ev.flux_correct_blocks(method, entries)
            else:
                method.compute(blocks, dt)
        el, prof = _profiled(one_step, a.steps, a.warmup)
        zones = n ** 3 * len(blocks)
        res[store] = {"value": zones * a.steps / el, "ms_per_step": el * 1e3 / a.steps,
                      "kernels_ms": {k: round(v[0], 4) for k, v in sorted(prof.items())},
                      "roofline": _roofline(prof, a.steps, (n + 2 * g - 4) ** 3 * len(blocks),
                                            prec, False,
                                            "smr_flux_correct_%s" % a.precision)}
        del blocks, fds, method
    r, plain = res[True], res[False]
    fc_ms = sum(r["kernels_ms"].get(k, 0.0) for k in
                ("k_store_fluxes", "k_send_to_coarse", "k_flux_correct",
                 "k_restrict_batch", "k_prolong_batch"))
    return {
        "metric": METRIC, "value": r["value"], "unit": "zone-updates/s", "n_gpus": 1,
        "steps": a.steps, "warmup": a.warmup, "ms_per_step": r["ms_per_step"],
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64" if a.precision == "fp64" else "f32-riemann/f64-update",
        "data": "synthetic (smooth hydro state sampled on a coarse and four fine blocks)",
        "config": {"workload": "smr_flux_correct_64x(1+4)x32_%s" % a.precision,
                   "blocks": 5 * units, "block": [n, n, n], "ghost": g, "riemann": "hllc",
                   "reconstruct": "plm", "theta": 2.0, "courant": 0.4,
                   "flux_links": 4 * units, "corrected_blocks": units,
                   "ghost_transfers": 2 * len(names) * 4 * units},
        "roofline": r["roofline"], "kernels_ms": r["kernels_ms"],
        "refinement_kernels_ms_per_step": fc_ms,
        "plain_step": {"value": plain["value"], "ms_per_step": plain["ms_per_step"],
                       "kernels_ms": plain["kernels_ms"]},
        "refinement_overhead": plain["value"] / r["value"] - 1.0,
    }


def run_output(a):
    """HDF5 block output of the device-resident 256^3 Orszag-Tang MHD block
    (12 fields with ghost zones, 1.73 GB) and its restart input, against
    the same file written from host copies of the fields."""
    import shutil
    import tempfile
    import enzoe_io as eio
    method = ev.VlctMethod(mhd=True, riemann="hlld", reconstruct="plm", theta=2.0,
                           courant=0.4)
    grid = make_state(a.n, method)
    step(grid)
    ev.synchronize()
    blk = grid.blocks[0]
    nbytes = sum(int(np.prod(s)) * 8 for s in blk.shapes.values())
    host = eio.HostBlock(blk.download(), blk.ghost)
    reps = max(1, min(a.steps, 3))
    tmp = tempfile.mkdtemp(prefix="enzoe_output_")
    cwd = os.getcwd()
    try:
        os.chdir(tmp)
        times = {"device": [], "host": [], "read": []}
        back = ev.DeviceBlock(blk.mx, blk.my, blk.mz, blk.ghost, (1.0 / a.n,) * 3, mhd=True)
        for r in range(reps):
            for kind, b in (("device", blk), ("host", host)):
                t0 = time.perf_counter()
                path = eio.OutputData(None, name=("%s-%%d.h5" % kind, "count")).write(
                    [b], (0, 0, 0), (1, 1, 1), (1, 1, 1), 1, 0.0, 0.0, count=r)
                times[kind].append(time.perf_counter() - t0)
            t0 = time.perf_counter()
            eio.read_block("./device-%d.h5" % r, "B__", back)
            times["read"].append(time.perf_counter() - t0)
            for k in ("device-%d" % r, "host-%d" % r):
                for ext in (".h5", ".block_list", ".file_list"):
                    os.remove(k + ext)
        got = back.download()
        same = all(np.array_equal(got[k], host.arrays[k]) for k in got)
    finally:
        os.chdir(cwd)
        shutil.rmtree(tmp, ignore_errors=True)
    best = {k: min(v) for k, v in times.items()}
    return {
        "metric": "HDF5 block output rate", "value": nbytes / best["device"] / 1e9,
        "unit": "GB/s", "n_gpus": 1, "steps": reps, "warmup": 0,
        "ms_per_step": best["device"] * 1e3, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic (Orszag-Tang vortex after one cycle)",
        "config": {"workload": "output_orszag_tang_%d3_12_fields" % a.n, "bytes": nbytes,
                   "file_system": "temporary directory of the host"},
        "host_write_GBs": nbytes / best["host"] / 1e9,
        "restart_read_GBs": nbytes / best["read"] / 1e9,
        "device_write_over_host_write": best["host"] / best["device"],
        "read_back_identical": bool(same),
    }


def batched_mixed_leg(steps, cpu=True):
    """Config 4 (512 blocks of 64^3, decaying Mach-5 MHD turbulence, beta
    0.1): `steps` fused steps in FP64 (its own CFL dt) and in mixed
    precision (the FP64 run's dt sequence, so both states stay comparable;
    the same fused step with its CFL reduction, whose dt is not used), both
    timed over the whole run; then the relative L1 of the mixed state
    against the FP64 one per field, on the device.  cpu: the oracle timed
    on a sample of the same blocks beside it."""
    import workloads
    nb, n = (8, 8, 8), 64
    spec = problems.turbulence_spectrum()
    norm = problems.turbulence_vrms_norm(spec, mach=5.0)
    vel = problems.turbulence_domain_velocity(spec, norm, nb[0] * n)
    builder = lambda ib: problems.turbulence_block_from_domain(vel, ib, n)
    grids = {}
    for precision in ("fp64", "mixed"):
        prec = ev.ENZOE_PREC_DOUBLE if precision == "fp64" else ev.ENZOE_PREC_MIXED
        method = ev.VlctMethod(mhd=True, riemann="hlld", reconstruct="plm", theta=2.0,
                               courant=0.4, precision=prec)
        if precision == "fp64":
            g = workloads.uniform_grid(method, nb, n, builder)
        else:
            g = ev.DeviceGrid(method, nb, (n, n, n), grids["fp64"].blocks[0].desc.cell_width[:],
                              bc=("periodic",) * 3)
            workloads.copy_grid_state(grids["fp64"], g)
        grids[precision] = g
    cpu_rate = cpu_baseline_sample(builder, n, True, 5.0 / 3.0, "hlld", 5) if cpu else None
    del vel
    res, dts = {}, []
    zones = n ** 3 * nb[0] * nb[1] * nb[2]
    for precision in ("fp64", "mixed"):
        g = grids[precision]
        m = g.method
        dt0 = m.timestep(g.blocks)
        mask = m.autotune(g.blocks, dt0)  # on a copy; also takes the scratch
        ev.synchronize()
        ev.profiling(True)
        t0 = time.perf_counter()
        if precision == "fp64":
            dt = dt0
            for _ in range(steps):
                dts.append(dt)
                g.refresh()
                dt = m.compute_timestep(g.blocks, dt)
        else:
            for dt in dts:
                g.refresh()
                m.compute_timestep(g.blocks, dt)
        ev.synchronize()
        el = time.perf_counter() - t0
        prof = ev.profile_report()
        ev.profiling(False)
        gpu_ms = sum(v[0] * v[1] for v in prof.values()) / steps
        wl = "turbulence_%s" % precision
        res[precision] = {"value": zones * steps / el, "ms_per_step": el * 1e3 / steps,
                          "launch_mask": mask,
                          "kernels_ms": {k: round(v[0], 4) for k, v in sorted(prof.items())},
                          "roofline": step_roofline(precision, zones, el * 1e3 / steps, gpu_ms,
                                                    wl)}
    per, mag = grids["mixed"].method.state_l1(grids["mixed"].blocks, grids["fp64"].blocks)
    del grids
    out = {"workload": "turbulence_512x64^3_mach5_beta0.1_vlct_plm_hlld", "steps": steps,
           "fp64": res["fp64"], "mixed": res["mixed"],
           "mixed_over_fp64": res["mixed"]["value"] / res["fp64"]["value"],
           "relative_l1_mixed_vs_fp64": per,
           "l1_mixed_vs_fp64_by_vector_magnitude": mag,
           "max_l1_by_vector_magnitude": max(mag.values()),
           "l1_definition": "sum|a-b| / sum|b| over the active cells; for a velocity or "
                            "B component the denominator is the sum of |b_vec|"}
    if cpu_rate is not None:
        out["cpu_baseline"] = cpu_rate
    return out


def main():
    a = parse()
    if a.gpus != 1:
        print("bench.py runs on one GPU only (this build has no multi-GPU "
              "path); got --gpus %d" % a.gpus, file=sys.stderr)
        sys.exit(2)
    if ev.device_count() < 1:
        print("no HIP device visible", file=sys.stderr)
        sys.exit(1)
    if a.copy_calibration:
        print(json.dumps(copy_bandwidth()))
        return
    if a.workload != "orszag_tang":
        run = {"fast_wave": run_fast_wave, "batched_amr": run_batched_amr,
               "turbulence": run_uniform, "kelvin_helmholtz": run_uniform,
               "smr_flux_correct": run_smr_flux_correct, "output": run_output,
               "drop_in": run_drop_in}[a.workload]
        out = run(a)
        out["copy_kernel_GBs"] = copy_bandwidth()["GBs"]
        print(json.dumps(out))
        return
    if a.dual_energy and a.precision == "mixed":
        print("the dual-energy formalism runs in FP64 only", file=sys.stderr)
        sys.exit(2)
    compare = not a.no_compare and a.precision != "fp64_fma"
    r = timed_run(a.precision, a.n, a.steps, a.warmup, dual_energy=a.dual_energy,
                  launch_mask=a.launch_mask, keep=compare and a.precision == "fp64",
                  structure=a.structure, flux_shape=a.flux_shape)
    other = dual = fma = None
    if compare:
        other = timed_run("mixed" if a.precision == "fp64" else "fp64", a.n,
                          a.steps, a.warmup, keep=a.precision == "mixed")
        fp64 = r if a.precision == "fp64" else other
        # the contracted FP64 build on the same steps, its state against the
        # reference-parity FP64 one cell by cell
        fma = timed_run("fp64_fma", a.n, a.steps, a.warmup, keep=True)
        per, mag, mx = fma["method"].state_l1(fma["grid"].blocks, fp64["grid"].blocks,
                                              max_rel=True)
        fma["vs_fp64"] = {"max_rel_per_cell": mx, "l1_by_vector_magnitude": mag}
        for x in (fma, fp64):
            del x["grid"], x["method"]
        dual = timed_run("fp64", a.n, a.steps, a.warmup, dual_energy=True)
    copy = copy_bandwidth()
    out = {
        "metric": METRIC,
        "value": r["value"],
        "unit": "zone-updates/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": r["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": {"fp64": "f64", "fp64_fma": "f64-fma"}.get(a.precision,
                                                            "f32-riemann/f64-update"),
        "data": "synthetic (3D Orszag-Tang vortex, seeded 1e-3 v_z noise)",
        "config": {"workload": r["workload"],
                   "block": [a.n, a.n, a.n], "ghost": 3, "riemann": "hlld",
                   "reconstruct": "plm", "theta": 2.0, "courant": 0.4,
                   "precision": a.precision, "launch_mask": r["launch_mask"],
                   "launch_choice": r["launch_choice"],
                   "dual_energy": a.dual_energy},
        # the step against the roofline as the metric defines it (440 B per
        # zone-update, fully fused); the dominant kernel's own figure beside it
        "roofline": dict(step_roofline(a.precision, a.n ** 3, r["ms_per_step"],
                                       r["gpu_step_ms"], r["workload"]),
                         dominant_kernel={
                             "kernel": r["dom"],
                             "achieved": r["achieved"],
                             "peak": HBM_PEAK_GBS,
                             "unit": "GB/s",
                             "frac": r["achieved"] / HBM_PEAK_GBS,
                             "traffic": measured_traffic(r["dom"], r["workload"]),
                             "bytes_per_launch": r["dom_bytes"],
                             "avg_launch_ms": r["dom_ms"],
                             # the kernel's real bound: FP64 instruction issue
                             "valu_issue": measured_valu(r["dom"], r["workload"], r["dom_ms"]),
                         }),
        "kernels_ms": {k: round(v[0], 4) for k, v in sorted(r["prof"].items())},
        "copy_kernel_GBs": copy["GBs"],
    }
    if other is not None:
        fp64, mixed = (r, other) if a.precision == "fp64" else (other, r)
        out["precision_comparison"] = {
            "fp64_value": fp64["value"], "mixed_value": mixed["value"],
            "mixed_over_fp64": mixed["value"] / fp64["value"],
            "mixed_ms_per_step": mixed["ms_per_step"],
            "mixed_kernels_ms": {k: round(v[0], 4)
                                 for k, v in sorted(mixed["prof"].items())},
            "mixed_roofline": step_roofline("mixed", a.n ** 3, mixed["ms_per_step"],
                                            mixed["gpu_step_ms"], mixed["workload"]),
            "mixed_dominant_GBs": mixed["achieved"],
            "mixed_traffic": measured_traffic(mixed["dom"], mixed["workload"]),
        }
        # the metric's batched leg: the Mach-5 turbulence workload (512 blocks
        # of 64^3) for its 50 steps in FP64 and in mixed precision, with the
        # mixed state's L1 against the FP64 one per field
        out["batched_mixed"] = batched_mixed_leg(a.batched_steps, cpu=not a.no_cpu_baseline)
    if dual is not None:
        fp64 = r if a.precision == "fp64" else other
        out["dual_energy_comparison"] = {
            "workload": dual["workload"], "value": dual["value"],
            "over_fp64": dual["value"] / fp64["value"],
            "ms_per_step": dual["ms_per_step"],
            "kernels_ms": {k: round(v[0], 4) for k, v in sorted(dual["prof"].items())},
        }
    if fma is not None:
        fp64 = r if a.precision == "fp64" else other
        out["contracted_fp64_comparison"] = {
            "workload": fma["workload"], "value": fma["value"],
            "over_fp64": fma["value"] / fp64["value"],
            "ms_per_step": fma["ms_per_step"],
            "kernels_ms": {k: round(v[0], 4) for k, v in sorted(fma["prof"].items())},
            "bar": "1e-12 relative per cell (velocity and B components over the "
                   "cell's vector magnitude)",
            "vs_fp64_after_steps": a.warmup + a.steps,
            **fma["vs_fp64"],
        }
    if not a.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(a.cpu_n, a.cpu_steps)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
