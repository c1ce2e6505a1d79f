"""Benchmark: photon packets per second on one MI355X for the octree configuration.

Workload (spiral_octree): stellar exponential disc + Sersic bulge (B/T 0.2), both discs with a two-arm
spiral (pitch 20 deg, weight 0.3); dust disc face-on tau (default 2); octree over +-20 kpc refined to
level 8 by dust mass fraction 1e-6; four 512x512 frames at 0/45/75/89 deg; HG g=0.6, albedo 0.6.
A step is one pass of the transport loop over a batch of packets (default 2e7: the configuration's
1e8 packets are the five timed steps) with all inputs resident in HBM. The timed region is the mixed-precision
path. The FP64 baseline is timed on one step of the same packets with both transport pipelines (the
faster counts). Image errors: mixed vs FP64 on identical packet streams, and both against an FP64
reference run with ten times the packets (the equal-image-error condition of the metric).

Prints one JSON line (see the project notes for the field meanings).
"""
import argparse
import json
import math
import os
import platform
import subprocess
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK = 8.0e12
# algorithmic bytes per cell crossing: tree leaf walk 6 bounds + kappa*rho + one link; Voronoi about 15.5
# neighbours x (4-byte id + 24-byte FP64 site) + kappa*rho; Cartesian kappa*rho + one wall coordinate
BYTES_PER_CROSSING = {"tree": {"mixed": 6 * 4 + 4 + 4, "fp64": 6 * 8 + 8 + 4},
                      "voronoi": {"mixed": 438, "fp64": 442},
                      "cartesian": {"mixed": 8, "fp64": 16}}
BYTES_PER_ABSORPTION = 16   # FP64 read-modify-write of one (cell, wavelength) entry per own-path crossing
BYTES_PER_PEELOFF = 16   # FP64 read-modify-write of one pixel
# HBM bytes per cell crossing of the walk kernel, measured with rocprofv3 FETCH_SIZE / WRITE_SIZE (scaled by
# the copy-kernel calibration) on this workload
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "k_walks_spiral_octree_traffic.json")


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--workload", default="spiral_octree",
                    choices=("disc_cartesian", "spiral_octree", "spiral_voronoi", "fit_population", "lognormal_kdtree",
                             "fitskirt"),
                    help="spiral_octree: octree over a spiral galaxy (the metric's workload); disc_cartesian: plain "
                         "Cartesian 50^3 (the CPU-runnable case); spiral_voronoi: 1e6-site Voronoi grid with the "
                         "absorption table; fit_population: batched population + fit; lognormal_kdtree: kd-tree over a "
                         "log-normal field; fitskirt: a FitSKIRT run (a step is one generation of the genetic "
                         "algorithm, evaluated in batched launches)")
    ap.add_argument("--popsize", type=int, default=100, help="fitskirt: population size (FitSKIRT's default)")
    ap.add_argument("--members", type=int, default=256, help="fit_population: population size")
    ap.add_argument("--one-wavelength", action="store_true",
                    help="spiral_voronoi: run only the 0.55 micron wavelength instead of all 50 per step")
    ap.add_argument("--member-packets", type=float, nargs=2, default=(1e5, 1e7),
                    help="fit_population: log-uniform range of packets per member")
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--packets", type=float, default=None,
                    help="packets per step (default 2e7 for spiral_octree, 1e6 for the others)")
    ap.add_argument("--tau", type=float, default=2.0)
    ap.add_argument("--tau-variants", type=float, nargs="*", default=[0.5, 5.0],
                    help="spiral_octree: the other face-on optical depths to time (two steps each, rate only)")
    ap.add_argument("--maxlevel", type=int, default=8)
    ap.add_argument("--pixels", type=int, default=512)
    ap.add_argument("--cpu-seconds", type=float, default=10.0,
                    help="target duration of each multi-threaded CPU baseline sample")
    ap.add_argument("--ref-mult", type=int, default=10,
                    help="the equal-image-error reference runs this many times the timed steps' packets in FP64")
    ap.add_argument("--slot-groups", type=int, default=2, choices=(1, 2),
                    help="slot groups of the timed steps (2: alternating on two streams; 1: one stream, as the "
                         "roofline step)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--tau-cut", type=float, default=0.0)
    ap.add_argument("--pipeline", type=int, default=1, choices=(0, 1),
                    help="1: event and walk kernels (default); 0: single lane-pipeline kernel (the autotuner, "
                         "unless switched off or given --tuning, picks the faster for the mixed steps)")
    ap.add_argument("--walk-order", type=int, default=0, choices=(0, 1, 2),
                    help="1: the walk kernel takes each pass's walks sorted by a coherence key; 2: sorted and "
                         "claimed by XCD segment; 0 (default): as the event kernel appended them")
    ap.add_argument("--no-equal-error", action="store_true",
                    help="skip the FP64 reference run with ten times the packets")
    ap.add_argument("--no-autotune", action="store_true",
                    help="keep the default launch parameters instead of tuning them on a sample first")
    ap.add_argument("--tuning", default=None, metavar="BLOCK,SLOTS,COMPACT",
                    help="fixed launch parameters of the mixed steps instead of the autotuner (e.g. the ones a "
                         "default run picked, for a profiled run whose kernel list holds no tuning trials)")
    ap.add_argument("--autotune-sample", type=float, default=None,
                    help="packets per autotuner trial (default: one step; a quarter of that for FP64)")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="first run a plain device copy of MIB MiB (5 times), the reference kernel for scaling "
                         "the profiler's FETCH_SIZE / WRITE_SIZE counters to HBM bytes")
    return ap.parse_args()


def copy_calibration(mib):
    import torch
    n = mib * (1 << 20) // 4
    src = torch.ones(n, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(5):
        dst.copy_(src)
    torch.cuda.synchronize()
    log("copy calibration: 5 copies of %d MiB (%d bytes read and %d written per copy)" % (mib, 4 * n, 4 * n))


def frames(w, ell=0):
    ins = w["instr"]
    return [ins.frame(i)[ell].copy() for i in range(len(ins.instruments))]


def run_steps(sim, first, nsteps, per, npp, precision, ell=0):
    """Runs nsteps batches of packets [first + s per, first + (s + 1) per) at wavelength ell (or at each of a list
    of wavelengths in turn); returns the summed counters and the transport kernel's device seconds."""
    ells = list(ell) if isinstance(ell, (list, tuple, range)) else [ell]
    tot = None
    for s in range(nsteps):
        st = None
        for e in ells:
            c = sim.run_packets(e, first + s * per, per, npp, precision)
            st = c if st is None else {k: st[k] + c[k] for k in c}
        tot = st if tot is None else {k: tot[k] + st[k] for k in st}
        log("  step %d/%d (%s): %d packets, %.3f s on the device" % (
            s + 1, nsteps, "mixed" if precision == 1 else "fp64", st["packets"], st["kernel_us"] * 1e-6))
    return tot, (tot["transport_us"] * 1e-6 if tot else 0.0)


def _cpu_model():
    cpu = platform.processor() or platform.machine()
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return cpu


def _time_reference(sc, mc0, threads, seconds, start=2000.0):
    """Runs the oracle's reference mode (MT19937 streams, the reference's chunk rule) on a sample of the
    workload sized to take about `seconds`; returns (packets, crossings, seconds)."""
    mc = np.array(mc0)
    n = start
    while True:
        mc[0] = n
        t0 = time.perf_counter()
        _, c = sc.run_reference(mc, threads)
        dt = time.perf_counter() - t0
        if dt * 3 > seconds or n >= 1e7:
            return float(c[0]), float(c[1] + c[2]), dt
        n = min(1e7, n * max(2.0, 0.9 * seconds / max(dt, 1e-3)))


def cpu_baseline(w, seconds):
    """The CPU restatement of the reference loop on this host: the -O2 build (the reference's qmake release
    level) on the host threads this process may use, capped at 16 (the CPU share of one GPU on these
    machines), the -O3 -march=x86-64-v3 build on as many, and the -O2 build on one thread."""
    from tests import oracle_lib as O
    nproc = os.cpu_count() or 1
    try:
        avail = len(os.sched_getaffinity(0))
    except AttributeError:
        avail = nproc
    threads = max(1, min(avail, 16))
    cpu = _cpu_model()
    mc = np.array(w["sim"].mc)
    sc = O.scene_from_workload(w)
    pk, cr, dt = _time_reference(sc, mc, threads, seconds)
    sc3 = O.scene_from_workload(w, "v3")
    pk3, cr3, dt3 = _time_reference(sc3, mc, threads, seconds)
    pk1, cr1, dt1 = _time_reference(sc, mc, 1, seconds / 4, start=500.0)
    return {"value": pk / dt, "unit": "packets/s", "cores": threads, "kind": "port",
            "sample": "%d packets of the same workload, MT19937 streams and the reference's chunk rule, %d threads "
                      "on %s, %.1f s" % (int(pk), threads, cpu, dt),
            "crossings_per_s": cr / dt, "cpu_model": cpu, "nproc": nproc, "affinity_cpus": avail,
            "compiler": "g++ -O2 -ffp-contract=off (the reference's qmake release level)",
            "o3_x86_64_v3": {"value": pk3 / dt3, "unit": "packets/s", "cores": threads,
                             "crossings_per_s": cr3 / dt3,
                             "sample": "%d packets, %d threads, %.1f s" % (int(pk3), threads, dt3),
                             "compiler": "g++ -O3 -march=x86-64-v3 (AVX2, FMA; -march=native of the build host is "
                                         "not portable to this machine)"},
            "single_thread": {"value": pk1 / dt1, "unit": "packets/s", "crossings_per_s": cr1 / dt1,
                              "sample": "%d packets, 1 thread, %.1f s" % (int(pk1), dt1)}}


def population(a):
    """fit_population: one batched launch over the population against member-by-member runs, then the fit."""
    from skirt7_amd import skirt as S
    from skirt7_amd import workloads
    from skirt7_amd.capi import SG_MIXED
    t0 = time.perf_counter()
    pop = workloads.fit_population(members=a.members, packets=tuple(a.member_packets))
    ms = pop["members"]
    counts = [m["packages"] for m in ms]
    total = int(sum(counts))
    log("setup: %d members (%d packets in all) built in %.1f s on %s" % (len(ms), total, time.perf_counter() - t0,
                                                                         S.device_name()))
    sims = [m["sim"] for m in ms]
    S.run_batch(sims[:4], [min(c, 100000) for c in counts[:4]], precision=SG_MIXED)   # warmup
    for m in ms:
        m["instr"].clear()
    t0 = time.perf_counter()
    cb = S.run_batch(sims, counts, precision=SG_MIXED)
    t_batch = time.perf_counter() - t0
    log("batched: %d packets in %.3f s: %.3e packets/s" % (total, t_batch, total / t_batch))
    t0 = time.perf_counter()
    for m in ms:
        m["sim"].run_packets(0, 0, m["packages"], m["packages"], SG_MIXED)
    t_seq = time.perf_counter() - t0
    log("member by member: %.3f s: %.3e packets/s" % (t_seq, total / t_seq))
    tw = pop["target"]
    tw["sim"].run_packets(0, 0, tw["packages"], tw["packages"], SG_MIXED)
    ref, gain = workloads.target_image(tw["instr"].frame(0)[0], pop["target_counts"])
    S.fit_frames([m["instr"] for m in ms[:2]], ref, pop["kernel"], gain * 1e-2, gain * 1e2)   # warmup
    t0 = time.perf_counter()
    lum, chi2 = S.fit_frames([m["instr"] for m in ms], ref, pop["kernel"], gain * 1e-2, gain * 1e2)
    t_fit = time.perf_counter() - t0
    log("fit of %d frames (%dx%d, PSF %dx%d): %.3f s" % (len(ms), ref.shape[1], ref.shape[0], pop["kernel"].dimension,
                                                         pop["kernel"].dimension, t_fit))
    cross = cb["own_crossings"] + cb["peel_crossings"] + cb["rewalk_crossings"]
    out = {"metric": "packets/s on 1 MI355X, mixed vs FP64 GPU at equal image error; % HBM peak",
           "value": total / t_batch, "unit": "packets/s", "n_gpus": 1, "steps": 1, "warmup": 1,
           "ms_per_step": 1e3 * t_batch, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
           "dtype": "f32 (walk, weights) + f64 (wall fallback, tallies)", "data": "synthetic",
           "config": {"workload": "fitting population of %d Cartesian galaxies, %d packets, one batched launch "
                                  "per round" % (len(ms), total), "members": len(ms), "packets": total},
           "member_by_member_packets_per_s": total / t_seq, "fit_seconds": t_fit,
           "crossings_per_s": cross / t_batch, "counters": cb,
           "best_fit": {"member": int(np.argmin(chi2)), "chi2": float(np.min(chi2))}}
    print(json.dumps(out), flush=True)


def fitskirt(a):
    """fitskirt: OligoFitScheme::runSelf (OligoFitScheme.cpp:27-34) on the GPU. The population's first
    evaluation and `warmup` generations are untimed; each timed step is one generation: the genetic algorithm's
    step, the new individuals' adjusted ski files parsed and taken by pooled members in place, one batched
    transport launch and the PSF-convolved golden-section fit of every member."""
    import tempfile

    from skirt7_amd import fitskirt as F
    from skirt7_amd import skirt as S
    from skirt7_amd import workloads
    from skirt7_amd.capi import SG_MIXED
    res = tempfile.mkdtemp()
    os.makedirs(os.path.join(res, "SED", "Sun"))
    with open(os.path.join(res, "SED", "Sun", "SunSED.dat"), "w") as f:
        f.write(workloads.SUN_STANDIN)
    packets = a.packets or 1e6
    prob = workloads.fitskirt_edgeon(packages=packets)
    t0 = time.perf_counter()
    data, kernel, gain = workloads.fitskirt_observation(prob, res)
    log("observation built in %.1f s on %s" % (time.perf_counter() - t0, S.device_name()))
    sim = F.AdjustableSkirtSimulation(prob["ski"], resources=res, precision=SG_MIXED)
    optim = F.Optimization(popsize=a.popsize, generations=a.warmup + a.steps, pmut=0.03, pcross=0.65)
    image = F.ReferenceImage(data, kernel, [0.2 * gain], [5.0 * gain])
    scheme = F.OligoFitScheme(sim, prob["ranges"], [image], optim, fixedSeed=True)
    optim.setup(scheme.ranges, 4357)
    t0 = time.perf_counter()
    optim.initialize()
    scheme._evaluate_pending()
    log("initial population of %d evaluated in %.2f s (pool built)" % (a.popsize, time.perf_counter() - t0))
    for _ in range(a.warmup):
        optim.step()
        scheme._evaluate_pending()
    sim.timing.update(prepare=0.0, transport=0.0, fit=0.0, packets=0)
    n0 = len(scheme.records)
    t0 = time.perf_counter()
    for _ in range(a.steps):
        tg = time.perf_counter()
        optim.step()
        scheme._evaluate_pending()
        log("generation %d: %d individuals, %.3f s, best chi2 %.6g" % (optim.generation(), len(scheme.records) -
            n0, time.perf_counter() - tg, optim.best()[1]))
    wall = time.perf_counter() - t0
    tm = sim.timing
    individuals = len(scheme.records) - n0
    best, score = optim.best()
    out = {"metric": "packets/s on 1 MI355X, mixed vs FP64 GPU at equal image error; % HBM peak",
           "value": tm["packets"] / wall, "unit": "packets/s", "n_gpus": 1, "steps": a.steps, "warmup": a.warmup,
           "ms_per_step": 1e3 * wall / a.steps, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
           "dtype": "f32 (walk, weights) + f64 (wall fallback, tallies)", "data": "synthetic",
           "config": {"workload": "FitSKIRT run: population %d, %d generations timed, %g packets per individual, "
                                  "Cartesian %dx%dx%d, %dx%d frame, 6 parameters" % (a.popsize, a.steps, packets, 100,
                                                                                       100, 30, 256, 128),
                      "popsize": a.popsize, "packets_per_individual": packets},
           "individuals_per_s": individuals / wall,
           "seconds_per_generation": {"total": wall / a.steps, "host_prepare": tm["prepare"] / a.steps,
                                      "transport": tm["transport"] / a.steps, "fit": tm["fit"] / a.steps,
                                      "ga_and_other": (wall - tm["prepare"] - tm["transport"] - tm["fit"]) / a.steps},
           "pool_rebuilds": sim.rebuilt,
           "best": {"chi2": score, "values": {r.label: (v if not r.quantityString() else
                                                        sim.units.out(r.quantityString(), v))
                                              for r, v in zip(scheme.ranges, best)}}}
    print(json.dumps(out), flush=True)


def main():
    a = parse()
    if a.gpus != 1:
        print("bench.py measures one GPU only; run it with --gpus 1 (multi-GPU runs are not part of this build)",
              file=sys.stderr)
        sys.exit(2)
    from skirt7_amd import skirt as S
    from skirt7_amd import workloads
    from skirt7_amd.capi import SG_FP64, SG_MIXED

    if a.workload == "fit_population":
        population(a)
        return
    if a.workload == "fitskirt":
        fitskirt(a)
        return
    if a.copy_calibration > 0:
        copy_calibration(a.copy_calibration)
    S.set_pipeline(a.pipeline)
    S.set_walk_order(a.walk_order)
    S.set_slot_groups(a.slot_groups)
    if a.packets is None:
        a.packets = 2e7 if a.workload == "spiral_octree" else 1e6
    per = int(a.packets)
    total = per * (a.steps + a.warmup)
    t0 = time.perf_counter()
    ell = 0
    ells = None
    if a.workload == "disc_cartesian":
        w = workloads.disc_cartesian(pixels=min(a.pixels, 256), packages=total)
        desc = "Cartesian 50^3 grid, exponential disc + Sersic bulge, face-on tau 1"
    elif a.workload == "spiral_octree":
        w = workloads.spiral_octree(tau=a.tau, maxlevel=a.maxlevel, pixels=a.pixels, packages=total, tau_cut=a.tau_cut)
        desc = "octree over a spiral galaxy, face-on tau %.3g" % a.tau
    elif a.workload == "spiral_voronoi":
        w = workloads.spiral_voronoi(pixels=min(a.pixels, 256), packages=total)
        ell = int(np.argmin(np.abs(w["wavelengths"] - 0.55)))
        if a.one_wavelength:
            desc = "Voronoi grid of 1e6 sites, 50 wavelengths (run at %.3g micron), absorption table stored" % (
                w["wavelengths"][ell])
        else:
            # a step runs its packets at every wavelength (each wavelength its own Npp packets, as the reference)
            ells = list(range(len(w["wavelengths"])))
            desc = "Voronoi grid of 1e6 sites, all 50 wavelengths per step (1e6 x 50 absorption table stored)"
    else:
        w = workloads.lognormal_kdtree(pixels=min(a.pixels * 2, 1024), packages=total)
        desc = ("kd-tree (alternating binary tree, TopDown) over 1e7 SPH particles sampled from a log-normal field, "
                "exact mass-in-box densities")
    ells = ells or [ell]
    ncell = w["grid"].numCells()
    gkind = "voronoi" if w["kind"] == "voronoi" else "cartesian" if w["kind"] == "cartesian" else "tree"
    setup_s = time.perf_counter() - t0
    log("setup: %s with %d cells built in %.1f s on %s" % (a.workload, ncell, setup_s, S.device_name()))
    sim = w["sim"]
    npp = total
    sample = int(a.autotune_sample or per)
    tuning = {}
    if a.tuning:
        b, sl, c = (int(float(x)) for x in a.tuning.split(","))
        S.set_tuning(b, 0, c)
        S.set_slots(sl)
        tuning["mixed"] = {"pipeline": a.pipeline, "block": b, "slots": sl, "compact_lanes": c, "fixed": True}
        a.no_autotune = True
    elif not a.no_autotune:
        # launch parameters picked on a sample of the workload's own packets, outside the timed region
        t0 = time.perf_counter()
        tuning["mixed"] = S.autotune(sim, ell, 0, sample, SG_MIXED, choose_pipeline=True)
        a.pipeline = tuning["mixed"]["pipeline"]
        log("autotune (mixed, %d-packet trials, %.1f s): %s" % (sample, time.perf_counter() - t0, tuning["mixed"]))

    # warmup (mixed), then the timed mixed steps
    run_steps(sim, 0, a.warmup, per, npp, SG_MIXED, ells)
    w["instr"].clear()
    if w.get("store_abs"):
        w["dust"].clear()
    t0 = time.perf_counter()
    cm, kmix = run_steps(sim, a.warmup * per, a.steps, per, npp, SG_MIXED, ells)
    wall = time.perf_counter() - t0
    img_mixed = frames(w, ell)
    pk = cm["packets"]
    cross = cm["own_crossings"] + cm["peel_crossings"] + cm["rewalk_crossings"]
    log("mixed: %d packets in %.3f s (transport kernel %.3f s in %d launches, emission + transport %.3f s): "
        "%.3e packets/s, %.3e crossings/s" % (pk, wall, kmix, cm["transport_launches"], cm["kernel_us"] * 1e-6,
                                               pk / wall, cross / kmix))

    out = {"metric": "packets/s on 1 MI355X, mixed vs FP64 GPU at equal image error; % HBM peak",
           "value": pk / wall, "unit": "packets/s", "n_gpus": 1, "steps": a.steps, "warmup": a.warmup,
           "ms_per_step": 1e3 * wall / a.steps, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
           "dtype": "f32 (walk, weights) + f64 (wall fallback, tallies)", "data": "synthetic",
           "config": {"workload": "%s, %d cells, %d packets per step, %d frames of %d^2" % (
                          desc, ncell, per * len(ells), len(w["instr"].instruments), w["instr"].instruments[0].Nx),
                      "packets_per_step": per * len(ells), "wavelengths_per_step": len(ells), "cells": ncell,
                      "instruments": len(w["instr"].instruments),
                      "pixels": w["instr"].instruments[0].Nx, "tuning": tuning},
           # whole setup before the timed region: synthetic input generation, the grid built (tree subdivision
           # on the device for the octree and kd-tree workloads), cell densities, scene upload
           "setup_s": setup_s}
    def algorithmic_bytes(c):
        """Algorithmic bytes of a run's walks: crossings x bytes per crossing, peel-off tallies, absorption."""
        b = (c["own_crossings"] + c["peel_crossings"] + c["rewalk_crossings"]) * BYTES_PER_CROSSING[gkind]["mixed"]
        b += c["peeloffs"] * BYTES_PER_PEELOFF
        if w.get("store_abs"):
            b += c["own_crossings"] * BYTES_PER_ABSORPTION
        return b

    # Three figures for the walk kernel. (1) frac: its algorithmic bytes per launch over its own launch
    # duration, measured on a step run with one slot group on one stream, so that each launch has the device to
    # itself (HIP events on that stream). (2) overlapped: the same ratio over the timed steps, whose two slot
    # groups' passes share the device, so a launch's duration includes time shared with the other stream's
    # kernels. (3) device: the timed steps' algorithmic bytes over their wall time.
    bytes_steps = algorithmic_bytes(cm)
    launches = max(1, cm["transport_launches"])
    kwalk = cm["walk_kernel_us"] * 1e-6
    S.set_slot_groups(1)
    w["instr"].clear()
    c1, _ = run_steps(sim, a.warmup * per, 1, per, npp, SG_MIXED, ells)
    S.set_slot_groups(a.slot_groups)
    w["instr"].clear()
    b1 = algorithmic_bytes(c1)
    l1 = max(1, c1["transport_launches"])
    k1 = c1["walk_kernel_us"] * 1e-6
    achieved = b1 / k1
    cross1 = c1["own_crossings"] + c1["peel_crossings"] + c1["rewalk_crossings"]
    log("walk kernel alone (one slot group, one stream): %d launches, %.3f ms each, %.3e B each: %.0f GB/s" % (
        l1, 1e3 * k1 / l1, b1 / l1, achieved / 1e9))
    traffic, tsrc = None, None
    try:
        with open(TRAFFIC_FILE) as f:
            tr = json.load(f)
        if a.pipeline == tr["pipeline"] and a.workload == tr["workload_name"]:
            # HBM bytes per crossing measured by rocprofv3 on this workload, times this run's crossings per launch
            traffic = (tr["fetch_bytes_per_crossing"] + tr["write_bytes_per_crossing"]) * cross1 / l1
            tsrc = "rocprofv3 FETCH_SIZE + WRITE_SIZE per crossing of k_walks on this workload (%s) x this run's " \
                   "crossings per launch" % os.path.relpath(TRAFFIC_FILE, ROOT)
    except (OSError, KeyError, ValueError):
        traffic = None
    out["roofline"] = {"bound": "hbm", "achieved": achieved / 1e9, "peak": HBM_PEAK / 1e9, "unit": "GB/s",
                       "frac": achieved / HBM_PEAK, "traffic": traffic, "traffic_source": tsrc,
                       "kernel": "k_walks" if a.pipeline == 1 else "k_mc",
                       "measured_on": "one step of %d packets with one slot group on one stream (each launch alone "
                                      "on the device)" % per,
                       "algorithmic_bytes_per_launch": b1 / l1, "kernel_ms_per_launch": 1e3 * k1 / l1,
                       "launches": l1,
                       "overlapped": {"frac": bytes_steps / kwalk / HBM_PEAK, "kernel_ms_per_launch": 1e3 * kwalk / launches,
                                      "note": "timed steps, two slot groups on two streams: a launch's duration "
                                              "includes time shared with the other stream's kernels"},
                       "device": {"frac": bytes_steps / wall / HBM_PEAK, "achieved": bytes_steps / wall / 1e9,
                                  "note": "algorithmic bytes of the timed steps over their wall time"}}
    out["crossings_per_s"] = cross / kmix
    out["transport_span_ms_per_launch"] = 1e3 * kmix / launches
    out["counters"] = cm

    if a.workload == "spiral_octree" and a.tau_variants:
        # the configuration's other face-on optical depths: same grid refinement, the density scaled; the
        # launch parameters of the headline run, one warmup and two timed steps each
        out["tau_variants"] = {}
        for tv in a.tau_variants:
            wt = workloads.spiral_octree(tau=tv, maxlevel=a.maxlevel, pixels=a.pixels, packages=3 * per,
                                         tau_cut=a.tau_cut)
            run_steps(wt["sim"], 0, 1, per, 3 * per, SG_MIXED, ell)
            t0 = time.perf_counter()
            ct, _ = run_steps(wt["sim"], per, 2, per, 3 * per, SG_MIXED, ell)
            dt = time.perf_counter() - t0
            crt = ct["own_crossings"] + ct["peel_crossings"] + ct["rewalk_crossings"]
            out["tau_variants"]["%g" % tv] = {"packets_per_s": ct["packets"] / dt, "crossings_per_s": crt / dt,
                                              "device_frac": algorithmic_bytes(ct) / dt / HBM_PEAK,
                                              "packets": ct["packets"], "scatterings": ct["scatterings"]}
            log("tau %g: %.3e packets/s, %.3e crossings/s" % (tv, ct["packets"] / dt, crt / dt))
            del wt

    if not a.no_fp64:
        # FP64 baseline rate: one timed step's packets with each pipeline (autotuned); the faster one counts
        first = a.warmup * per
        fp64 = {}
        for p in (1, 0):
            S.set_pipeline(p)
            if not a.no_autotune:
                S.set_tuning(-1, -1, -1)
                S.set_slots(0)
                tuning["fp64_pipeline%d" % p] = S.autotune(sim, ell, first, max(1, sample // 4), SG_FP64)
                log("autotune (fp64, pipeline %d): %s" % (p, tuning["fp64_pipeline%d" % p]))
            w["instr"].clear()
            t0 = time.perf_counter()
            c64, _ = run_steps(sim, first, 1, per, npp, SG_FP64, ells)
            dt = time.perf_counter() - t0
            fp64[p] = (c64["packets"] / dt, c64)
            log("  fp64 step (pipeline %d): %d packets, %.3f s" % (p, c64["packets"], dt))
        best = max(fp64, key=lambda p: fp64[p][0])
        fp64_rate, c64 = fp64[best]
        out["fp64"] = {"packets_per_s": fp64_rate, "pipeline": best,
                       "packets_per_s_by_pipeline": {str(p): fp64[p][0] for p in fp64},
                       "crossings_per_s": (c64["own_crossings"] + c64["peel_crossings"] + c64["rewalk_crossings"])
                       / (c64["transport_us"] * 1e-6),
                       "speedup_mixed_over_fp64": out["value"] / fp64_rate}
        log("fp64: %.3e packets/s (pipeline %d; %s); mixed/fp64 = %.2fx" % (
            fp64_rate, best, {p: "%.3e" % fp64[p][0] for p in fp64}, out["value"] / fp64_rate))
        if not a.no_equal_error and len(ells) == 1:
            # Equal image error. The mixed image of the timed steps (their steps x per packets) and the FP64
            # image of the same packets (identical random streams), both against an FP64 reference run of
            # ref_mult times as many fresh packets; every image normalised to a full run.
            S.set_pipeline(best)
            w["instr"].clear()
            t0 = time.perf_counter()
            for s_ in range(a.steps):
                sim.run_packets(ell, first + s_ * per, per, npp, SG_FP64)
            log("  fp64 image of the timed steps' packets: %.1f s" % (time.perf_counter() - t0))
            n_img = a.steps * per
            img64 = [x * (npp / n_img) for x in frames(w, ell)]
            img32 = [x * (npp / n_img) for x in img_mixed]
            nref = a.ref_mult * n_img
            ref0 = (a.warmup + a.steps) * per
            w["instr"].clear()
            t0 = time.perf_counter()
            for k in range(a.ref_mult * a.steps):
                sim.run_packets(ell, ref0 + k * per, per, nref, SG_FP64)
                if (k + 1) % 10 == 0:
                    log("  fp64 reference: %d of %d packets, %.1f s" % ((k + 1) * per, nref, time.perf_counter() - t0))
            imgref = frames(w, ell)
            S.set_pipeline(a.pipeline)

            def rms(x, y):
                return float(np.sqrt(np.mean((x - y) ** 2)) / np.sqrt(np.mean(y ** 2)))
            e32 = [rms(x, r) for x, r in zip(img32, imgref)]
            e64 = [rms(x, r) for x, r in zip(img64, imgref)]
            d32 = [rms(x, y) for x, y in zip(img32, img64)]
            # the error of n packets alone (the reference's own noise, 1/ref_mult of the variance, removed), and
            # the packet count at which the worst frame's error would be 1%
            worst = max(e32)
            own = worst / math.sqrt(1.0 + 1.0 / a.ref_mult)
            out["fp64"]["equal_image_error"] = {
                "packets": n_img, "reference_packets": nref, "rms_error_mixed": e32, "rms_error_fp64": e64,
                "rms_difference_mixed_vs_fp64_same_packets": d32, "bound": 0.01,
                "within_bound": bool(worst <= 0.01),
                "packets_for_1pct_rms": float(n_img * (own / 0.01) ** 2)}
            log("image rms error vs FP64 with %dx packets: mixed %s, fp64 %s; mixed vs fp64 on the same packets %s" % (
                a.ref_mult, ["%.4f" % e for e in e32], ["%.4f" % e for e in e64], ["%.1e" % e for e in d32]))

    if not a.no_cpu:
        out["cpu_baseline"] = cpu_baseline(w, a.cpu_seconds)
        log("cpu: %.3e packets/s on %d threads (-O2), %.3e (-O3 x86-64-v3), %.3e on 1 thread" % (
            out["cpu_baseline"]["value"], out["cpu_baseline"]["cores"], out["cpu_baseline"]["o3_x86_64_v3"]["value"],
            out["cpu_baseline"]["single_thread"]["value"]))
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
