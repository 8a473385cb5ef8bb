"""Benchmark: particle-steps/s of the plain DEM step at 500K spheres on one MI355X.

Workload: 500,000 lognormal spheres (mean d = 1 mm, sigma 0.1, clipped to [0.7, 1.3] mm) inserted
at random without overlap into a cylinder of radius 40 mean-d with a flat floor; E = 1e7 Pa,
nu = 0.29, rho = 2500 kg/m^3, sliding friction 0.3, rolling friction 0.002 d, restitution 0.6,
dt = 1e-6 s, FP64, reference rebuild policy (displacement trigger).

The timed steps (`--steps`, the headline `value`) run on that workload poured into a bed
(200,000 steps on the GPU before timing): the packed regime, P/N = 7.6 candidate pairs and about
2.5 contacts per sphere -- the regime the reference's PCCP timings describe.  Inputs are
resident in HBM before timing.  One step = [list rebuild when triggered] + pair forces +
ordered sums, walls, gravity, integration.

Other legs, each a JSON object beside the headline: the workload's own 50,000-step run from
its random insertion (`full_run`; its first 300 steps are `loose_start`), the 2M bimodal box's
20,000-step run (`c2`), the list rebuild stage by stage (`rebuild`), the reference-named drop-in
entry points in the reference driver's order (`dropin`), the batched form (`batched`), the
adjoint form and calibration (`adjoint`), and the CPU baseline (`cpu_baseline`: the oracle, the
same PCCP algorithm in C with OpenMP, on the host cores).

Prints ONE JSON line.
"""
import argparse
import importlib.util
import json
import math
import os
import platform
import sys
import threading
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
PKG_DIR = os.path.join(ROOT, "discrete-element-method-for-packing-cuda-c-code_amd")
METRIC = "particle-steps/s at 500K spheres (FP64, plain vs batched); adjoint/forward cost"
PUBLISHED_500K = 9.8e6      # GPU-PCCP at 500K spheres, reference Figure 2(a)
HBM_PEAK_GBS = 8000.0


def load_pkg():
    spec = importlib.util.spec_from_file_location(
        "dem_amd", os.path.join(PKG_DIR, "__init__.py"), submodule_search_locations=[PKG_DIR])
    mod = importlib.util.module_from_spec(spec)
    sys.modules["dem_amd"] = mod
    spec.loader.exec_module(mod)
    return mod


def log(msg):
    print(msg, file=sys.stderr, flush=True)


LOOSE = 300   # the first steps of the full run, reported on their own (loose start)


def kernel_bytes(P, N):
    """Algorithmic bytes per launch: one touch per array element with the reference's element
    sizes (double3 24, double 8, uint 4, uint2 8); gathers of particle data by pairs counted
    once per particle; for a fused kernel, its unique external inputs and outputs.
    fused step (k_step): per particle, reads pos, pdis, angv, qi (4 x 24), rad, rmass, inert
                   (3 x 8), two CSR offsets (2 x 4), the wall-candidate flag 4; writes pos,
                   pdis, angv, qi (4 x 24) and the contact count 4 (= 232); per pair, reads
                   the pair 8, the transposed index 4, the history (dispt 24 + fricp 8) and
                   writes the new history 32 (= 76).
    split path, pair kernel: pair 8 + history r/w 2*32 + output (F, Ta, Tb, flag) 80 per
                   pair; gathered particle data (pos, pdis, angv, rad, rmass = 88 B) once
                   per particle.
    split path, particle kernel: pair output 80 + transposed index 4 per pair; per particle
                   record r/w 2*128, wall history r/w 2*32, two CSR offsets 8, contact
                   count 4."""
    return {"step": 76 * P + 232 * N, "pair": 152 * P + 88 * N, "particle": 84 * P + 332 * N}


def reference_step_bytes(P, N):
    """The reference's own non-rebuild step (pair force, reduction, walls + gravity,
    integration as separate kernels with their per-pair arrays) by the same counting rule:
    268 P + 636 N bytes."""
    return 268 * P + 636 * N


def rebuild_bytes(N, P, P_old, C):
    """Algorithmic bytes of each rebuild stage (same counting rule; kernels and arrays of this
    build): hash reads pos 24 and the index 4, writes (cell, index) 8 and the arrival rank 4,
    adds to the cell count 4, and the counts are zeroed first (4 per cell); the sort scans the
    counts (read and write 4 per cell) and scatters each pair (reads it 8, its rank 4 and its
    cell's start 4, writes it 8); the cell ranges are the scan's output (no stage of their
    own); reorder moves the 128-byte record and 32-byte wall history, reads the scattered pair
    8, its cell's bounds 8 and the inverse map, writes the sorted pair 8, the inverse map,
    wall flag, compact position (32 + 16), index and old slot; the count reads the compact
    positions, indices and keys and writes two counts; two scans read and write n words
    each; the fill reads what the count reads plus both offsets and writes the pair (8),
    partner (4) and transposed entry (4); the remap reads the pair, both members' indices and
    old slots, the old offsets, the old pair it finds (8) and its history (32), and writes the
    history (32); the transpose reads the offsets, the transposed entry and the pair it finds,
    and writes two words per entry."""
    return {"hash": 44 * N + 4 * C, "sort": 24 * N + 8 * C, "cell_range": 0,
            "reorder": 416 * N, "nbr_count": 48 * N + 8 * C, "scans": 16 * N,
            "nbr_fill": 56 * N + 8 * C + 16 * P, "remap": 12 * N + 80 * P + 8 * P_old,
            "transpose": 8 * N + 24 * P}


TRAFFIC_PROFILES = [os.path.join(ROOT, "profiles", f) for f in (
    "hbm_traffic_step_500k_lognormal_cylinder.md", "bed_step_500k_lognormal_cylinder.md")]


def measured_traffic(kernel: str, n: int, p_avg: float):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 --pmc summaries
    (FETCH_SIZE and WRITE_SIZE scaled by a known-size copy) of the 500K workload: the profile
    whose pair count is nearest the run's average, scaled to that average by its ratio of
    measured to algorithmic bytes.  None for other sizes."""
    if n != 500_000:
        return None, None
    key = f"traffic_bytes_per_launch dem::{kernel}<false> = "
    best = None
    for path in TRAFFIC_PROFILES:
        if not os.path.exists(path):
            continue
        for line in open(path):
            if line.startswith(key):
                b, _, p = line[len(key):].partition(" P = ")
                p = float(p)
                if best is None or abs(p - p_avg) < abs(best[1] - p_avg):
                    best = (float(b), p, os.path.basename(path))
    if best is None:
        return None, None
    name = {"k_step": "step", "k_pair_force": "pair", "k_particle": "particle"}[kernel]
    ratio = best[0] / kernel_bytes(best[1], n)[name]
    return ratio * kernel_bytes(p_avg, n)[name], (
        f"{best[2]} at P = {best[1]:.0f}: {ratio:.2f} x algorithmic bytes, scaled to the run's "
        f"average P")


def dominant_kernel(tm, N, label):
    """Dominant kernel of a run timed with OPT_TIME_STEPS (events around each step's kernels
    on the context stream): its average duration, algorithmic bytes at the average pair count
    and the roofline object."""
    cand = []
    sp, fu = tm["split"], tm["fused"]
    if sp["steps"]:
        pa = sp["pairs_sum"] / sp["steps"]
        cand.append(("pair", "k_pair_force", sp["pair_walls_ms"], sp["steps"], pa))
        cand.append(("particle", "k_particle", sp["particle_ms"], sp["steps"], pa))
    if fu["steps"]:
        cand.append(("step", "k_step", fu["step_ms"], fu["steps"], fu["pairs_sum"] / fu["steps"]))
    dom, kname, tot_ms, launches, p_avg = max(cand, key=lambda c: c[2])
    ms = tot_ms / launches
    kb = kernel_bytes(p_avg, N)[dom]
    achieved = kb / (ms * 1e-3) / 1e9
    traffic, src = measured_traffic(kname, N, p_avg)
    log(f"{label}: dominant kernel {kname}, {launches} launches, {ms * 1e3:.1f} us average at "
        f"P {p_avg:.0f}: {achieved:.0f} GB/s algorithmic ({achieved / HBM_PEAK_GBS:.3f} of peak)")
    return {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic, "traffic_source": src,
            "kernel": kname, "launches": launches, "avg_us": ms * 1e3, "avg_pairs": p_avg,
            "bytes_per_launch": kb}


# ---------------------------------------------------------------------------------------
# Headline: the poured 500K bed
# ---------------------------------------------------------------------------------------

def dense_bed(dem, w, settle_steps, load=None, save=None, graph=True):
    """The 500K workload poured into a bed on the GPU (graph steps), or loaded from a file
    a previous run saved (so a profiler can run the timed window alone)."""
    if load:
        d = np.load(load)
        ws = dem.Workload(w.name + "-settled", w.instances, d["pos"], d["pdis"], d["angv"], w.rad,
                          w.rmass, w.inert, meta=dict(w.meta, bed_top=float(d["pos"][:, 2].max()),
                                                      settle_steps=int(d["settle_steps"])))
        log(f"bed loaded from {load}")
        return ws
    t0 = time.perf_counter()
    ws = dem.settle_bed(w, steps=settle_steps, interval=50, tap=0.0, graph=graph)
    log(f"bed poured in {time.perf_counter() - t0:.1f} s ({settle_steps} steps, top "
        f"{ws.meta['bed_top']:.1f})")
    if save:
        np.savez(save, pos=ws.pos, pdis=ws.pdis, angv=ws.angv, settle_steps=settle_steps)
        log(f"bed saved to {save}")
    return ws


def new_context(dem, ws, split_opt):
    ctx = dem.Context(ws.instances)
    if split_opt is not None:
        ctx.set_option(dem.OPT_FUSED_STEP, 0 if split_opt else 1)
    ctx.load(ws)
    return ctx


# Steps the headline context runs on the loaded bed before its warm-up: the bed is handed over
# as positions, displacements and spins, so the contact histories (and the zero-history bits)
# start empty and the first steps after loading still rebuild them -- part of preparing the
# poured state, like the pour itself, not of the timed workload.
BED_RESUME_STEPS = 200


def bench_headline(dem, ws, steps, warmup, split_opt):
    """`steps` timed steps of the poured bed after BED_RESUME_STEPS steps that rebuild its
    contact histories and `warmup` more (displacement trigger: the first step builds the
    lists), wall clock around device work only; then the same steps again from the same
    state with HIP events around every step's kernels (the dominant kernel's duration; the
    events cost a few percent, so `value` is from the clean pass; the trajectory is the same
    bit for bit); then one rebuild timed stage by stage."""
    N = ws.n
    ctx = new_context(dem, ws, split_opt)
    ctx.step(BED_RESUME_STEPS + warmup)
    ctx.sync()
    s0 = ctx.stats()
    t0 = time.perf_counter()
    ctx.step(steps)
    ctx.sync()
    el = time.perf_counter() - t0
    s1 = ctx.stats()
    value = N * steps / el
    log(f"timed {steps} steps of the bed in {el * 1e3:.2f} ms: {value:.4e} particle-steps/s, "
        f"P={s1.npairs} (P/N {s1.npairs / N:.2f}), {s1.rebuilds - s0.rebuilds} rebuilds, "
        f"path {'fused' if s1.step_path else 'split'}")
    ctx.close()
    ctx = new_context(dem, ws, split_opt)
    ctx.step(BED_RESUME_STEPS + warmup)
    ctx.set_option(dem.OPT_TIME_STEPS, 1)
    ctx.step_timing()
    ctx.step(steps)
    ctx.sync()
    tm = ctx.step_timing()
    roof = dominant_kernel(tm, N, "bed")
    contacts = int(ctx.contacts().sum())
    # one rebuild of the bed, stage by stage
    rb = ctx.time_rebuild()
    ctx.close()
    P = s1.npairs
    ref_b = reference_step_bytes(P, N)
    out = {"value": value, "ms_per_step": el / steps * 1e3, "npairs": P, "contacts": contacts,
           "rebuilds_timed": int(s1.rebuilds - s0.rebuilds),
           "step_path": "fused" if s1.step_path else "split",
           "reference_step_bytes": {"bytes": ref_b,
                                    "GB_s_over_step": ref_b / (el / steps) / 1e9,
                                    "frac": ref_b / (el / steps) / 1e9 / HBM_PEAK_GBS}}
    return out, roof, rebuild_object(rb)


def rebuild_object(rb):
    """Stage table of one rebuild (HIP events between the stages; each event adds a gap of a
    few microseconds) and the whole rebuild timed again with events at its two phases' ends
    only (`total_ms`, the figure the roofline fraction uses)."""
    b = rebuild_bytes(rb["n"], rb["npairs"], rb["npairs_old"], rb["ncells"])
    stages = {}
    tot_b = st_ms = 0.0
    for k in rb:
        if k not in b or b[k] == 0:
            continue
        ms = rb[k]
        gbs = b[k] / (ms * 1e-3) / 1e9 if ms > 0 else None
        stages[k] = {"ms": round(ms, 4), "bytes": b[k], "GB_s": round(gbs, 1) if gbs else None,
                     "frac": round(gbs / HBM_PEAK_GBS, 4) if gbs else None}
        tot_b += b[k]
        st_ms += ms
    tot_ms = rb["phases_ms"]
    log("rebuild: " + ", ".join(f"{k} {v['ms'] * 1e3:.0f} us" for k, v in stages.items()) +
        f"; stages sum {st_ms:.3f} ms; whole rebuild {tot_ms:.3f} ms, "
        f"{tot_b / (tot_ms * 1e-3) / 1e9:.0f} GB/s")
    return {"n": rb["n"], "npairs": rb["npairs"], "npairs_old": rb["npairs_old"],
            "ncells": rb["ncells"], "sort": "counting sort by cell", "total_ms": round(tot_ms, 4),
            "stages_sum_ms": round(st_ms, 4), "bytes": tot_b,
            "GB_s": round(tot_b / (tot_ms * 1e-3) / 1e9, 1),
            "frac": round(tot_b / (tot_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4), "stages": stages}


# ---------------------------------------------------------------------------------------
# The workload's own runs: 500K from its random insertion, 2M bimodal box
# ---------------------------------------------------------------------------------------

def bench_run(dem, w, steps, split_opt, label, loose=0):
    """`steps` steps of workload `w` from its initial state (displacement trigger), then the
    same steps again with events around every step's kernels."""
    ctx = new_context(dem, w, split_opt)
    ctx.step(1)                      # the first list (not timed)
    ctx.sync()
    s0 = ctx.stats()
    t0 = time.perf_counter()
    done, el_loose, t_last = 0, None, t0
    while done < steps:
        k = min(steps - done, loose if (done == 0 and loose) else 1000)
        ctx.step(k)
        done += k
        now = time.perf_counter()
        if el_loose is None:
            ctx.sync()
            el_loose = time.perf_counter() - t0
        if now - t_last > 10.0:
            log(f"  {label}: {done} / {steps} steps ({now - t0:.1f} s)")
            t_last = now
    ctx.sync()
    el = time.perf_counter() - t0
    s1 = ctx.stats()
    ctx.close()
    ctx = new_context(dem, w, split_opt)
    ctx.step(1)
    ctx.set_option(dem.OPT_TIME_STEPS, 1)
    ctx.step_timing()
    for k0 in range(0, steps, 5000):
        ctx.step(min(5000, steps - k0))
    ctx.sync()
    tm = ctx.step_timing()
    ctx.close()
    out = {"workload": w.name, "n": w.n, "steps": steps, "value": w.n * steps / el,
           "ms_per_step": el / steps * 1e3, "npairs_start": s0.npairs, "npairs_end": s1.npairs,
           "rebuilds": int(s1.rebuilds - s0.rebuilds),
           "kernel_totals_ms": {"fused_step": tm["fused"]["step_ms"],
                                "fused_steps": tm["fused"]["steps"],
                                "split_pair_walls": tm["split"]["pair_walls_ms"],
                                "split_particle": tm["split"]["particle_ms"],
                                "split_steps": tm["split"]["steps"], "rebuild": tm["rebuild_ms"],
                                "rebuilds": tm["rebuilds"]},
           "roofline": dominant_kernel(tm, w.n, label)}
    if loose:
        out["loose_start"] = {"steps": loose, "value": w.n * loose / el_loose,
                              "ms_per_step": el_loose / loose * 1e3}
    log(f"{label}: {steps} steps in {el:.2f} s: {out['value']:.4e} particle-steps/s, P "
        f"{s0.npairs} -> {s1.npairs}, {out['rebuilds']} rebuilds")
    return out


# ---------------------------------------------------------------------------------------
# Drop-in: the reference-named entry points in the reference driver's order
# ---------------------------------------------------------------------------------------

def bench_dropin(dem, ws, steps, width=64):
    """The 11 reference-named entry points called in dempacking.cu:535-739 order on
    caller-owned device buffers (the reference's AoS double3 layout, rows of `width`
    neighbours), on the poured bed: zero the totals; on a rebuild step calcHash, RadixSort,
    findCellStart, neighborarray, prefixsum, pairlist; then calculateforcepair, sumallforcei,
    calculateforceiw, updateposition (which returns the rebuild flag to the host).  The
    neighbour test is the reference's fixed cutoff for unit spheres (1.55 d)."""
    import ctypes
    L = dem.lib()
    D = dem.DeviceArray
    inst = ws.instances[0]
    n = ws.n
    S = n * width
    f64, u32 = np.float64, np.uint32
    pos, pdis, angv = (D.from_host(getattr(ws, k)) for k in ("pos", "pdis", "angv"))
    rad, rmass, inert = (D.from_host(getattr(ws, k)) for k in ("rad", "rmass", "inert"))
    C = inst.gsize[0] * inst.gsize[1] * inst.gsize[2]
    hash0, hash1 = D(2 * n, u32), D(2 * n, u32)
    cstart = D(C, u32)
    anei, aref = D(S, u32), D(S, u32)
    njgi, njli, sjgi, oldsjgi = D(n, u32), D(n, u32), D(n, u32), D(n, u32)
    lp, ln, oldln, cpair = D(S, u32), D(S, u32), D(S, u32), D(S, u32)
    dispt, olddispt, fp, tpi, tpj = (D(3 * S, f64) for _ in range(5))
    fricp, oldfricp = D(S, f64), D(S, f64)
    disptw, fricpw = D(3 * n, f64), D(n, f64)
    forcei, torquei, qi = D(3 * n, f64), D(3 * n, f64), D(3 * n, f64)
    contacti = D(n, u32)
    G = (ctypes.c_uint * 3)(*inst.gsize)
    cell = (ctypes.c_double * 3)(*inst.cell)
    orig = (ctypes.c_double * 3)(*inst.origin)
    wsz = (ctypes.c_double * 3)(*inst.walls.size)
    L.setMaterialProperties(ctypes.byref(inst.mat.p), ctypes.byref(inst.mat.w))
    L.setWallGeometry(ctypes.byref(inst.walls))
    iall = ctypes.c_uint(1)
    maxhz = ctypes.c_double(0.0)
    P = ctypes.c_uint(0)
    L.dem_last_error()
    dt = inst.dt

    def one_step():
        forcei.zero()
        torquei.zero()
        contacti.zero()
        rebuilt = iall.value != 0
        if rebuilt:
            qi.zero()
            L.calcHash(pos.p, hash0.p, G, cell, orig, n)
            L.RadixSort(hash0.p, hash1.p, n, 32)
            L.findCellStart(hash0.p, cstart.p, n, C)
            L.neighborarray(pos.p, hash0.p, cstart.p, G, cell, orig, width, n, anei.p, njgi.p,
                            njli.p)
            L.prefixsum(n, njgi.p, sjgi.p, oldsjgi.p, ctypes.byref(P))
            L.pairlist(sjgi.p, oldsjgi.p, oldln.p, anei.p, njgi.p, njli.p, olddispt.p, oldfricp.p,
                       width, width, n, aref.p, lp.p, ln.p, dispt.p, fricp.p)
        L.calculateforcepair(pos.p, pdis.p, angv.p, rad.p, rmass.p, lp.p, ln.p, oldln.p, dispt.p,
                             fricp.p, olddispt.p, oldfricp.p, dt, fp.p, tpi.p, tpj.p, cpair.p,
                             P.value)
        L.sumallforcei(width, n, njgi.p, njli.p, fp.p, tpi.p, tpj.p, cpair.p, aref.p, forcei.p,
                       torquei.p, contacti.p)
        L.calculateforceiw(wsz, pos.p, pdis.p, angv.p, rad.p, rmass.p, dt, P.value, n, disptw.p,
                           fricpw.p, forcei.p, torquei.p)
        L.updateposition(wsz, n, pos.p, pdis.p, angv.p, rmass.p, inert.p, forcei.p, torquei.p,
                         qi.p, dt, 0.0, ctypes.byref(iall), ctypes.byref(maxhz))
        return rebuilt

    one_step()                         # builds the lists
    for _ in range(4):
        one_step()
    err = L.dem_last_error()
    if err:
        raise dem.DemError(err, "drop-in entry points")
    L.dem_device_sync()
    t0 = time.perf_counter()
    reb = 0
    for _ in range(steps):
        reb += one_step()
    L.dem_device_sync()
    el = time.perf_counter() - t0
    err = L.dem_last_error()
    if err:
        raise dem.DemError(err, "drop-in entry points")
    out = {"workload": ws.name, "n": n, "steps": steps, "value": n * steps / el,
           "ms_per_step": el / steps * 1e3, "npairs": int(P.value), "rebuilds": reb,
           "row_width": width}
    log(f"drop-in: {out['value']:.4e} particle-steps/s ({out['ms_per_step']:.3f} ms per step, "
        f"P {P.value}, {reb} rebuilds)")
    return out


# ---------------------------------------------------------------------------------------
# Batched and adjoint forms
# ---------------------------------------------------------------------------------------

def bench_batched(dem, steps, plain_steps, interval, tune_steps=48, graph=True):
    """Batched form (256 instances, hipGraph-captured, autotuned skin / launch shapes / rebuild
    interval) over `steps` steps, against the same instances stepped as per-instance plain
    contexts (same skin and fixed rebuild interval) over `plain_steps` steps: host loop, and
    hipGraph replays.  graph=False steps the batched context with the host loop too and skips
    the per-instance graph replays (the same kernels, for profilers that cannot trace graph
    launches)."""
    t0 = time.perf_counter()
    w = dem.config3(rebuild_interval=interval)
    log(f"batched workload: {len(w.instances)} instances, n={w.n} ({time.perf_counter() - t0:.1f} s)")
    big = dem.Context(w.instances)
    big.load(w)
    tune = None
    skin = dem.SKIN
    if tune_steps:
        big.set_option(dem.OPT_REBUILD_INTERVAL, 0)
        tune = big.autotune(tune_steps)
        interval = max(1, min(int(tune["interval"]), 200))
        skin = tune["skin"]
        big.set_option(dem.OPT_REBUILD_INTERVAL, interval)
        log(f"autotune: {tune}")
    steps = max(interval, (steps // interval) * interval)
    plain_steps = max(interval, (plain_steps // interval) * interval)
    adv = big.step_graph if graph else big.step
    adv(2 * interval)
    big.sync()
    t0 = time.perf_counter()
    adv(steps)
    big.sync()
    tb = time.perf_counter() - t0
    nb = w.n
    flag = big.stats().last_rebuild_flag
    # the dominant kernel's roofline: the same schedule through the host loop with HIP events
    # around every step's kernels (2 rebuild intervals)
    big.set_option(dem.OPT_TIME_STEPS, 1)
    big.step_timing()
    big.step(2 * interval)
    big.sync()
    roof = dominant_kernel(big.step_timing(), nb, "batched")
    del big
    # per-instance plain launches: one context per instance, run back to back
    offs = [0]
    for inst in w.instances:
        offs.append(offs[-1] + inst.n)
    ctxs = []
    for k, inst in enumerate(w.instances):
        c = dem.Context([inst])
        sl = slice(offs[k], offs[k + 1])
        c.set_state(w.pos[sl], w.pdis[sl], w.angv[sl], w.rad[sl], w.rmass[sl], w.inert[sl])
        c.set_option(dem.OPT_SKIN, skin)
        c.set_option(dem.OPT_REBUILD_INTERVAL, interval)
        c.step(2 * interval)
        ctxs.append(c)
    for c in ctxs:
        c.sync()
    log(f"batched {nb * steps / tb:.4e} p-steps/s over {steps} steps; {len(ctxs)} per-instance "
        f"contexts ready")
    t0 = time.perf_counter()
    for c in ctxs:
        c.step(plain_steps)
    tp = time.perf_counter() - t0
    tg = None
    if graph:
        t0 = time.perf_counter()
        for c in ctxs:
            c.step_graph(plain_steps)
        tg = time.perf_counter() - t0
    vb, vp = nb * steps / tb, nb * plain_steps / tp
    vg = nb * plain_steps / tg if tg else None
    log(f"batched {vb:.4e} p-steps/s; per-instance plain {vp:.4e}; ratio {vb / vp:.2f}x"
        + (f" (per-instance graphs {vg:.4e}: {vb / vg:.2f}x)" if vg else ""))
    return {"workload": w.name, "instances": len(w.instances), "n": nb, "steps": steps,
            "per_instance_steps": plain_steps, "rebuild_interval": interval, "value": vb,
            "per_instance_plain": vp, "per_instance_graph": vg, "ratio_vs_plain": vb / vp,
            "ratio_vs_plain_graph": vb / vg if vg else None, "ms_per_step": tb / steps * 1e3,
            "trigger_in_last_group": int(flag), "autotune": tune, "roofline": roof}


def bench_adjoint(dem, nsteps=5000, every=100, fd_rel=1e-4, graph=True, roll_floor=3.0):
    """Differentiable form on the 20K-sphere bed: checkpointed forward + adjoint backward
    against the plain forward (cost ratio), and the adjoint gradient of phi and CN in
    (mu, log E, e, mu_r) against central differences of the plain forward -- with the
    form's rolling floor (roll_floor = k; 0: the reference's rolling moment, regularised in
    the reverse pass only)."""
    import copy
    A = dem.adjoint
    t0 = time.perf_counter()
    w = dem.settle_bed(dem.config4(), tap=0.0, graph=graph)
    log(f"adjoint workload: n={w.n}, bed top {w.meta['bed_top']:.2f} ({time.perf_counter() - t0:.1f} s)")
    obs = dem.make_obs(zhi=w.meta["bed_top"] - 3.0, margin=2.0, zlo=2.0, eps=0.5, eps_c=1e-3)
    prob = A.PackingProblem(w, nsteps, every, obs, roll_floor=roll_floor)
    th0 = [0.3, math.log(1e7), 0.6, 0.002]
    prob.evaluate([th0])          # warm-up
    phi_p, cn_p = prob.evaluate([th0])
    t_plain = prob.timing["plain_s"]
    ms_plain = prob.ctx.last_ms()
    phi, cn = prob.forward([th0])
    g_phi = prob.backward([1.0], [0.0])[0]
    t_fwd, t_bwd = prob.timing["forward_s"], prob.timing["backward_s"]
    info = prob.ctx.adj_info()
    g_cn = prob.backward([0.0], [1.0])[0]
    assert phi[0] == phi_p[0] and cn[0] == cn_p[0]
    log(f"adjoint: forward {t_fwd:.3f} s, backward {t_bwd:.3f} s; central differences next")
    r = A.fd_check(prob, [th0], rel=fd_rel, which=(0, 1, 3))
    base = prob.mmap.material(th0)
    gam = base.p.dampnp
    vals = []
    for sgn in (1, -1):
        m = copy.deepcopy(base)
        m.p.dampnp = m.p.damptp = m.w.dampnw = m.w.damptw = gam * (1 + sgn * fd_rel)
        prob.ctx.set_material(0, m)
        prob.ctx.load(w)
        prob.ctx.step(nsteps)
        vals.append(prob.ctx.observe(obs))
    dg = prob.mmap.dgamma_de(0.6)
    r["fd_phi"][2] = (vals[0][0][0] - vals[1][0][0]) / (2 * fd_rel * gam) * dg
    r["fd_cn"][2] = (vals[0][1][0] - vals[1][1][0]) / (2 * fd_rel * gam) * dg
    rel = lambda a, b: abs(a - b) / max(abs(a), abs(b), 1e-300)
    out = {"workload": w.name, "n": w.n, "steps": nsteps, "checkpoint_every": every,
           "roll_floor": roll_floor, "phi": float(phi[0]), "cn": float(cn[0]),
           "plain_s": t_plain, "forward_s": t_fwd, "backward_s": t_bwd,
           "cost_ratio": (t_fwd + t_bwd) / t_plain, "backward_over_forward": t_bwd / t_fwd,
           "device_ms": {"plain": ms_plain, "forward": info["forward_ms"],
                         "backward": info["backward_ms"]},
           "checkpoint_bytes": info["checkpoint_bytes"], "tape_bytes": info["tape_bytes"],
           "params": list(A.PARAMS), "fd_rel_step": fd_rel,
           "grad_phi": g_phi.tolist(), "fd_phi": r["fd_phi"].tolist(),
           "grad_cn": g_cn.tolist(), "fd_cn": r["fd_cn"].tolist(),
           "rel_err_phi": [rel(a, b) for a, b in zip(g_phi, r["fd_phi"])],
           "rel_err_cn": [rel(a, b) for a, b in zip(g_cn, r["fd_cn"])],
           # error against the dominant sensitivity: |adj - fd| x natural scale of the
           # parameter / max over parameters of |fd| x its natural scale
           "norm_err_phi": norm_err(g_phi, r["fd_phi"], th0),
           "norm_err_cn": norm_err(g_cn, r["fd_cn"], th0)}
    log(f"adjoint: cost ratio {out['cost_ratio']:.2f}x (plain {t_plain:.3f} s, fwd {t_fwd:.3f} s, "
        f"bwd {t_bwd:.3f} s); rel err phi {np.round(out['rel_err_phi'], 4)} cn "
        f"{np.round(out['rel_err_cn'], 4)}; normalised phi {np.round(out['norm_err_phi'], 4)} "
        f"cn {np.round(out['norm_err_cn'], 4)}")
    return out


def norm_err(adj, fd, theta):
    nat = np.array([theta[0], 1.0, theta[2], theta[3]])
    fd = np.asarray(fd)
    top = np.max(np.abs(fd * nat))
    return (np.abs(np.asarray(adj) - fd) * nat / top).tolist()


def bench_calibration(dem, ninst=64, n=2000, side=12.0, nsteps=5000, every=100, iters=24,
                      seed=7, graph=True, roll_floor=3.0):
    """Calibration of (mu, log E, e, mu_r) to the observables of a bed run by the plain method
    at seeded hidden parameters: `ninst` seeded starts in one batched context (replicas of one
    settled bed), one forward + adjoint backward per iteration."""
    A = dem.adjoint
    t0 = time.perf_counter()
    w1 = dem.settle_bed(dem.config4(n=n, side=side, seed=seed), tap=0.0, graph=graph)
    obs = dem.make_obs(zhi=w1.meta["bed_top"] - 2.0, margin=1.5, zlo=1.5, eps=0.5, eps_c=1e-3)
    rng = np.random.default_rng(seed)

    def draw(k):
        return np.stack([rng.uniform(0.2, 0.6, k), np.log(rng.uniform(3e6, 3e7, k)),
                         rng.uniform(0.4, 0.8, k), rng.uniform(0.001, 0.005, k)], axis=1)
    hidden = draw(1)[0]
    tgt = A.PackingProblem(w1, nsteps, every, obs, roll_floor=roll_floor)
    tphi, tcn = tgt.evaluate([hidden])
    del tgt
    insts = [copy_instance(w1.instances[0]) for _ in range(ninst)]
    rep = lambda a: np.concatenate([a] * ninst)
    big = dem.Workload(f"calibration-{ninst}x{n}", insts, rep(w1.pos), rep(w1.pdis), rep(w1.angv),
                       rep(w1.rad), rep(w1.rmass), rep(w1.inert), rebuild_interval=0)
    prob = A.PackingProblem(big, nsteps, every, obs, roll_floor=roll_floor)
    starts = draw(ninst)
    log(f"calibration: {ninst} x {n} spheres, target phi {tphi[0]:.5f} cn {tcn[0]:.4f} "
        f"({time.perf_counter() - t0:.1f} s setup)")
    t0 = time.perf_counter()
    res = A.calibrate(prob, float(tphi[0]), float(tcn[0]), starts, iters=iters,
                      lr=(0.05, 0.25, 0.05, 0.001), scale=(1e-3, 0.05), decay=0.9,
                      progress=lambda h: log(f"  calibration iter {h['iter']}: loss {h['loss']:.4f} "
                                             f"(fwd {h['forward_s']:.2f} s, bwd {h['backward_s']:.2f} s)"))
    el = time.perf_counter() - t0
    h0, h1 = res["history"][0], res["history"][-1]
    best = [min(h["loss_per_instance"][k] for h in res["history"]) for k in range(ninst)]
    out = {"instances": ninst, "n_per_instance": n, "steps": nsteps, "iters": iters,
           "hidden": hidden.tolist(), "target": [float(tphi[0]), float(tcn[0])],
           "loss_first": h0["loss"], "loss_last": h1["loss"],
           "median_instance_loss_first": float(np.median(h0["loss_per_instance"])),
           "median_instance_loss_best": float(np.median(best)),
           "median_instance_loss_last": float(np.median(h1["loss_per_instance"])),
           "instances_loss_below_1": int(np.sum(np.asarray(h1["loss_per_instance"]) < 1.0)),
           "seconds": el, "seconds_per_iter": el / iters,
           "particle_steps_per_s_fwd_bwd": big.n * nsteps * iters / el}
    log(f"calibration: loss {h0['loss']:.3f} -> {h1['loss']:.3f} in {iters} iterations "
        f"({el:.1f} s)")
    return out


def copy_instance(inst):
    import ctypes
    c = type(inst)()
    ctypes.pointer(c)[0] = inst
    return c


# ---------------------------------------------------------------------------------------
# CPU baseline
# ---------------------------------------------------------------------------------------

def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor() or "unknown"


def cpu_baseline(dem, ws, w0, budget_s, w2=None):
    """The oracle (this build's C restatement of the same PCCP algorithm, OpenMP, built with
    -O3 -march=x86-64-v3 -fopenmp -ffp-contract=off) on the GPU box's host cores: the poured
    500K bed -- the same state the GPU's timed steps start from -- on all threads and on one
    thread (a bounded sample of steps each, after the first, list-building step), the
    2,000-sphere configuration's full 20,000-step run on all threads and on one thread, and
    (w2) a bounded sample of the 2M bimodal box from its initial state, as its GPU leg starts,
    on all threads."""
    import ctypes
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as orc
    gomp = ctypes.CDLL("libgomp.so.1")
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    out = {"cpu_model": cpu_model(), "nproc": os.cpu_count(), "threads": threads}

    def bed_rate(th, budget, wk=ws, width=64):
        gomp.omp_set_num_threads(th)
        sim = orc.OracleSim(wk.instances[0], wk.pos, wk.pdis, wk.angv, wk.rad, wk.rmass,
                            wk.inert, width=width)
        assert sim.step(1) == 0
        t0 = time.perf_counter()
        assert sim.step(1) == 0
        per = time.perf_counter() - t0
        k = max(1, min(200, int(budget / max(per, 1e-6))))
        t0 = time.perf_counter()
        assert sim.step(k) == 0
        el = time.perf_counter() - t0
        P = sim.scalars()["npairs"]
        assert sim.scalars()["overflow"] == 0
        del sim
        return wk.n * k / el, k, P

    def c0_rate(th):
        gomp.omp_set_num_threads(th)
        sim = orc.OracleSim(w0.instances[0], w0.pos, w0.pdis, w0.angv, w0.rad, w0.rmass, w0.inert)
        t0 = time.perf_counter()
        for it in range(w0.steps):
            sim.set_iall(1 if it % w0.rebuild_interval == 0 else 0)
            assert sim.step(1) == 0
        el = time.perf_counter() - t0
        return w0.n * w0.steps / el

    v_all, k_all, _ = bed_rate(threads, budget_s)
    log(f"cpu baseline: poured bed {v_all:.4e} p-steps/s on {threads} threads ({k_all} steps)")
    v_one, k_one, _ = bed_rate(1, 0.5 * budget_s)
    log(f"cpu baseline: poured bed {v_one:.4e} p-steps/s on 1 thread ({k_one} steps)")
    c2 = None
    if w2 is not None:
        v2, k2, P2 = bed_rate(threads, 0.5 * budget_s, w2, 128)
        c2 = {"n": w2.n, "steps": k2, "value_all_threads": v2, "npairs": P2}
        log(f"cpu baseline: 2M bimodal box {v2:.4e} p-steps/s on {threads} threads ({k2} steps)")
    c0_all = c0_rate(threads)
    c0_one = c0_rate(1)
    log(f"cpu baseline: 2,000-sphere full run {c0_all:.4e} ({threads} threads), {c0_one:.4e} "
        f"(1 thread) p-steps/s")
    gomp.omp_set_num_threads(threads)
    out.update({"value": v_all, "unit": "particle-steps/s", "cores": threads, "kind": "port",
                "sample": f"{k_all} steps of the poured 500000-sphere bed (the state the GPU's "
                          f"timed steps start from), after its list-building first step; oracle "
                          f"C/OpenMP",
                "one_thread": {"value": v_one, "steps": k_one},
                "config0_full_run": {"n": w0.n, "steps": w0.steps, "value_all_threads": c0_all,
                                     "value_one_thread": c0_one},
                "config2_sample": c2,
                "build": "-O3 -march=x86-64-v3 -fopenmp -ffp-contract=off"})
    return out


# ---------------------------------------------------------------------------------------

def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000,
                    help="timed steps of the poured 500K bed (the headline)")
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--n", type=int, default=500_000)
    ap.add_argument("--settle", type=int, default=200_000, help="pour steps before timing")
    ap.add_argument("--bed-save", default=None, help="save the poured bed (.npz)")
    ap.add_argument("--bed-load", default=None, help="load the poured bed (.npz)")
    ap.add_argument("--bed-only", action="store_true",
                    help="run only the headline steps on the bed (no other leg, no JSON line)")
    ap.add_argument("--no-full-run", action="store_true")
    ap.add_argument("--full-steps", type=int, default=50_000)
    ap.add_argument("--no-c2", action="store_true")
    ap.add_argument("--c2-steps", type=int, default=20_000)
    ap.add_argument("--no-dropin", action="store_true")
    ap.add_argument("--dropin-steps", type=int, default=200)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=8.0)
    ap.add_argument("--no-batched", action="store_true")
    ap.add_argument("--batched-steps", type=int, default=10_000)
    ap.add_argument("--batched-plain-steps", type=int, default=480)
    ap.add_argument("--batched-interval", type=int, default=20)
    ap.add_argument("--no-adjoint", action="store_true")
    ap.add_argument("--no-calibration", action="store_true")
    ap.add_argument("--calib-instances", type=int, default=64)
    ap.add_argument("--calib-n", type=int, default=2000)
    ap.add_argument("--fd-rel", type=float, default=1e-4,
                    help="relative step of the adjoint leg's central differences")
    ap.add_argument("--roll-floor", type=float, default=3.0,
                    help="differentiable form: rolling floor k (0: the reference's moment)")
    ap.add_argument("--adjoint-only", action="store_true",
                    help="run only the adjoint leg (and calibration unless --no-calibration)")
    ap.add_argument("--split", action="store_true",
                    help="force the split path (pair kernel + gather); default: auto")
    ap.add_argument("--fused", action="store_true", help="force the fused step; default: auto")
    ap.add_argument("--no-graph", action="store_true",
                    help="host-loop steps wherever the legs use hipGraph replays (same kernels, "
                         "same results; for profilers that cannot trace graph launches)")
    ap.add_argument("--copy-probe", type=int, default=0,
                    help="also run 3 device-to-device copies of this many bytes (the known-size "
                         "kernel that scales the HBM traffic counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}); this build has no "
              f"multi-GPU path", file=sys.stderr)
        sys.exit(2)
    dem = load_pkg()
    if args.copy_probe:
        L = dem.lib()
        nb = int(args.copy_probe)
        a, b = L.dem_dev_alloc(nb), L.dem_dev_alloc(nb)
        assert a and b, "copy probe allocation failed"
        dem._check(L.dem_memzero(a, nb), "dem_memzero")
        for _ in range(3):
            dem._check(L.dem_memcpy(b, a, nb, 3), "dem_memcpy")
        dem._check(L.dem_device_sync(), "dem_device_sync")
        L.dem_dev_free(a)
        L.dem_dev_free(b)
        log(f"copy probe: 3 device-to-device copies of {nb} bytes")
    split_opt = True if args.split else (False if args.fused else None)
    graph = not args.no_graph
    if args.adjoint_only:
        a = bench_adjoint(dem, graph=graph, roll_floor=args.roll_floor, fd_rel=args.fd_rel)
        if not args.no_calibration:
            a["calibration"] = bench_calibration(dem, ninst=args.calib_instances, n=args.calib_n,
                                                 graph=graph, roll_floor=args.roll_floor)
        log(json.dumps(a))
        return
    t0 = time.perf_counter()
    w = dem.config1(n=args.n)
    log(f"workload {w.name}: n={w.n} height={w.meta['height']:.1f} ({time.perf_counter() - t0:.1f} s)")
    # the 2M box's random insertion runs on the host while the GPU pours the 500K bed
    c2_box = {}
    gen = None
    if not (args.no_c2 or args.bed_only):
        gen = threading.Thread(target=lambda: c2_box.update(w=dem.config2()), daemon=True)
        gen.start()
    ws = dense_bed(dem, w, args.settle, args.bed_load, args.bed_save, graph)
    head, roof, rebuild = bench_headline(dem, ws, args.steps, args.warmup, split_opt)
    if args.bed_only:
        log(json.dumps({"bed": head, "roofline": roof, "rebuild": rebuild}))
        return
    full = None
    if not args.no_full_run:
        full = bench_run(dem, w, args.full_steps, split_opt, "full run", loose=LOOSE)
    c2 = w2 = None
    if gen is not None:
        gen.join()
        w2 = c2_box.pop("w")
        c2 = bench_run(dem, w2, args.c2_steps, split_opt, "2M bimodal")
    dropin = None
    if not args.no_dropin:
        dropin = bench_dropin(dem, ws, args.dropin_steps)
    batched = None
    if not args.no_batched:
        batched = bench_batched(dem, args.batched_steps, args.batched_plain_steps,
                                args.batched_interval, graph=graph)
    adjoint = None
    if not args.no_adjoint:
        log("importing torch for the autograd layer (the first import on a fresh machine can "
            "take a minute or two)")
        t0 = time.perf_counter()
        import torch  # noqa: F401
        log(f"torch imported ({time.perf_counter() - t0:.1f} s)")
        adjoint = bench_adjoint(dem, graph=graph, roll_floor=args.roll_floor, fd_rel=args.fd_rel)
        if not args.no_calibration:
            adjoint["calibration"] = bench_calibration(dem, ninst=args.calib_instances,
                                                       n=args.calib_n, graph=graph,
                                                       roll_floor=args.roll_floor)
    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(dem, ws, dem.config0(), args.cpu_seconds, w2)
    N = ws.n
    out = {
        "metric": METRIC,
        "value": head["value"],
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": head["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": head["value"] / PUBLISHED_500K,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "plain-500k-lognormal-cylinder-poured-bed", "n": N,
                   "settle_steps": ws.meta["settle_steps"], "resume_steps": BED_RESUME_STEPS,
                   "npairs": head["npairs"],
                   "pairs_per_particle": round(head["npairs"] / N, 3),
                   "contacts_per_particle": round(head["contacts"] / N, 3),
                   "rebuild_policy": "displacement trigger |q| > 0.275",
                   "rebuilds_timed": head["rebuilds_timed"], "step_path": head["step_path"]},
        "roofline": roof,
        "cpu_baseline": cpu,
        "reference_step_bytes": head["reference_step_bytes"],
        "rebuild": rebuild,
        "full_run": full,
        "c2": c2,
        "dropin": dropin,
        "batched": batched,
        "adjoint": adjoint,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
