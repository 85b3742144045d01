"""Benchmark of the alignment core on one MI355X.

One "step" = one batched banded Gotoh + traceback pass (band offset = half =
b, run_dispatch(Some((b,1,1,b))) semantics of crates/map/src/extend.rs:293)
over the uneven-length workload: 2,000,000 read-window pairs, reads
U{50..250}, band U{16,32,48,64}, window = read + 2*band, 1% substitutions
with A->G weighted 3x, 0.2% indels of 1-5 (seed 0xE5A00003).  Inputs are
resident in HBM before the timed region.  value = GCUPS over the in-band
cell census sum_pairs L*(2b+1).

Reported beside it (same JSON line): the plain one-launch-per-call form on the
10,000-pair 100/164-nt workload, the batched form on that workload, the
batched EA-Myers filter on 1,000,000 150/214-nt pairs and its plain
one-launch-per-pair form, the integer-VALU
roofline of the Gotoh kernel, and the CPU restatement of the reference path
timed on this host's cores on a bounded sample of the same workload.

    python bench.py [--gpus 1] [--steps 5] [--warmup 2]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

# MI355X integer VALU peak: 256 CUs x 64 int32 lane-ops/clk x 2.4 GHz.
VALU_PEAK_TOPS = 256 * 64 * 2.4e9 / 1e12
# One wave64 issues at most one VALU instruction per 4 cycles on its SIMD
# (16 lanes a cycle): 1/1024 of the chip, the ceiling of a one-wave call.
ONE_WAVE_PEAK_TOPS = 16 * 2.4e9 / 1e12
GOTOH_OPS_PER_CELL = 18      # E 3, F 3, subst 1, diag 1, H 3, dirs 2, hdir 3, pack 1, best 1
MYERS_OPS_PER_WORD_COL = 31  # per 64-bit word and text column, +3 per column


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def band_cells(hb) -> int:
    m = hb.read_len.astype(np.int64)
    n = hb.win_len.astype(np.int64)
    off = hb.offset.astype(np.int64)
    half = np.maximum(hb.half.astype(np.int64), 0)
    interior = (off - half >= 0) & (off + m + half <= n)
    total = int(np.sum(m[interior] * (2 * half[interior] + 1)))
    from esperanto_amd.batch import band_cells as bc
    for k in np.nonzero(~interior)[0]:
        total += bc(int(m[k]), int(n[k]), int(off[k]), int(half[k]))
    return total


def timed_steps(fn_launch, db, steps: int, warmup: int):
    """Runs warmup + timed steps; returns (wall s/step, device ms/step, per-launch
    timings of every timed step)."""
    for _ in range(warmup):
        fn_launch()
        db.sync()
    db.sync()
    kms, tims = [], []
    t0 = time.perf_counter()
    for _ in range(steps):
        fn_launch()
        db.sync()
        kms.append(db.last_kernel_ms())
        tims.append(db.last_timings())
    db.sync()
    wall = (time.perf_counter() - t0) / steps
    return wall, float(np.mean(kms)), tims


def _graph_steps(g, db, steps: int, warmup: int) -> float:
    """Wall seconds per replay of a captured batch, each replay waited for."""
    for _ in range(warmup):
        g.launch()
        db.sync()
    t0 = time.perf_counter()
    for _ in range(steps):
        g.launch()
        db.sync()
    return (time.perf_counter() - t0) / steps


def read_traffic(path: str):
    """HBM bytes per band-kernel launch from the committed PMC summary, if any."""
    try:
        for line in open(path):
            if line.startswith("traffic_bytes_per_launch:"):
                return int(float(line.split(":", 1)[1]))
    except OSError:
        pass
    return None


def bench_window(args, buf, params):
    """Spliced window DP and mate-window rescue over the 2x150-nt pair workload
    (R1 across a GT-AG intron of 60-50,000 nt: two 300-nt anchor windows; R2
    rescued in the mate window, reference geometry 2*(750+m)+m nt)."""
    from esperanto_amd.reference import DeviceReference
    from esperanto_amd.rescue import DeviceMateBatch
    from esperanto_amd.splice import DeviceSpliceBatch
    from esperanto_amd.synth_window import splice_rescue_workload
    n = args.window_pairs
    log(f"splice/rescue workload: {n} pairs on a {args.window_ref_mb} Mb reference")
    ref, sp, mt, _ = splice_rescue_workload(n, ref_len=args.window_ref_mb * 1_000_000)
    dref = DeviceReference(ref, buf)
    out = {"pairs": n, "reference_bases": int(ref.lens.sum()),
           "reference_device_bytes": dref.device_bytes()}
    from esperanto_amd import _native
    from esperanto_amd.batch import tune_prune_shape
    for name, db in (("spliced", DeviceSpliceBatch(sp, dref, buf)),
                     ("mate_rescue", DeviceMateBatch(mt, dref, buf))):
        # the pruned passes' stripe and head shape, tuned on the pipeline itself
        # (untimed; results do not depend on it)
        tune = None
        if not args.no_tune:
            def run_once():
                db.launch(params)
                db.sync()
            tune = tune_prune_shape(buf, run_once, db.last_timings, reps=1)
            log(f"{name}: pruned-pass shape tuned: {tune}")
        w, k, tims = timed_steps(lambda: db.launch(params), db, max(args.steps, 3), 1)
        stages = {}
        for st in tims:
            for kind, half, ms in st:
                key = {0: "plan", 1: f"band{half}", 2: "strip", 3: "myers", 5: "prep",
                       6: "select_finish", 7: "escalated_full", 11: "escalated_unpruned",
                       12: "escalation_sort", 13: "escalated_wave"}.get(kind, str(kind))
                stages.setdefault(key, []).append(ms)
        nsteps = len(tims)
        stages = {k2: round(float(np.sum(v)) / nsteps, 3) for k2, v in stages.items()}
        out[name] = {"pairs_per_s": round(n / w), "ms_per_step": round(w * 1e3, 3),
                     "device_ms": round(k, 3), "stages_ms": stages}
        log(f"{name}: {n / w / 1e6:.2f} M pairs/s, {w * 1e3:.2f} ms/step, stages {stages}")
        ps = db.last_plan_stats()
        out[name]["prune_shape_tune"] = tune
        out[name]["escalated_to_full"] = ps["escalated"]
        out[name]["escalated_pruned"] = ps["full_pruned"]
        out[name]["band_attempts"] = ps["band_pairs"][30]
        # Full-matrix reruns: one more (untimed) launch with the cell census
        # on gives the cells the passes really computed; their rate over the
        # timed full-matrix stages (sort, pruned passes, unpruned pass) is the
        # stage's GCUPS and roofline fraction.
        esc_ms = sum(stages.get(k, 0.0) for k in ("escalation_sort", "escalated_full",
                                                    "escalated_wave", "escalated_unpruned"))
        if esc_ms and ps["escalated"]:
            buf.set_option(_native.ESP_OPT_FULL_CENSUS, 1)
            db.launch(params)
            db.sync()
            cs = db.last_plan_stats()
            buf.set_option(_native.ESP_OPT_FULL_CENSUS, 0)
            comp, iss, mat = (cs["full_cells_computed"], cs["full_lane_cells_issued"],
                              cs["full_cells_matrix"])
            tops = comp * GOTOH_OPS_PER_CELL / (esc_ms / 1e3) / 1e12
            out[name]["full_matrix"] = {
                "ms": round(esc_ms, 3), "cells_computed": comp, "lane_cells_issued": iss,
                "cells_of_full_matrices": mat,
                "computed_fraction": round(comp / mat, 4) if mat else None,
                "issue_efficiency": round(comp / iss, 4) if iss else None,
                "gcups": round(comp / (esc_ms / 1e3) / 1e9, 1),
                "roofline_frac": round(tops / VALU_PEAK_TOPS, 4)}
            log(f"{name}: full-matrix census {out[name]['full_matrix']}")
        log(f"{name}: routing {ps}")
        # The whole pipeline recorded as one HIP graph and replayed.
        g = db.capture(params)
        gw = _graph_steps(g, db, max(args.steps, 3), 1)
        g.check()
        g.destroy()
        out[name]["graph_ms_per_step"] = round(gw * 1e3, 3)
        out[name]["graph_pairs_per_s"] = round(n / gw)
        log(f"{name}: graph replay {n / gw / 1e6:.2f} M pairs/s, {gw * 1e3:.2f} ms/step")
        if name == "mate_rescue":
            rec = db.raw_results()[0]
            out[name]["placed_fraction"] = round(float((rec["placed"] == 1).mean()), 4)
        db.free()
        buf.set_option(_native.ESP_OPT_FULL_PRUNE_SHAPE, 0)  # the next leg starts from the default
    if not args.no_cpu:
        from oracle import oracle
        threads = min(16, os.cpu_count() or 1)
        o = oracle.OracleRef(ref)
        nsample = min(n, args.window_cpu_pairs)
        sub_sp = _subset_splice(sp, nsample)
        sub_mt = _subset_mate(mt, nsample)
        t0 = time.perf_counter()
        oracle.spliced_batch(o, sub_sp, params, nthreads=threads)
        t1 = time.perf_counter()
        oracle.mate_batch(o, sub_mt, params, nthreads=threads)
        t2 = time.perf_counter()
        out["cpu_port"] = {"spliced_pairs_per_s": round(nsample / (t1 - t0)),
                           "mate_rescue_pairs_per_s": round(nsample / (t2 - t1)),
                           "cores": threads, "sample_pairs": nsample}
        log(f"cpu port: {out['cpu_port']}")
    dref.close()
    return out


def bench_strip(args, buf, params):
    """Band halves the band kernel has no instance for (20, 40, 56): the C3
    shape (reads U{50..250}, window = read + 2 x band, 1% substitutions with
    A->G weighted 3x, 0.2% indels of 1-5), every pair through the int32 strip
    kernel.  GCUPS over the strip launch and its fraction of the int32 peak
    at 18 ops a cell."""
    from esperanto_amd.batch import DeviceBatch, synth_pairs
    hb = synth_pairs(0xE5A00006, args.strip_pairs, 50, 250, [20, 40, 56], sub_rate=0.01, ag_weight=3.0,
                     indel_rate=0.002, indel_max=5)
    cells = band_cells(hb)
    db = DeviceBatch(hb, buf)
    wall, kms, tims = timed_steps(lambda: db.launch_gotoh(params, 0), db, max(args.steps, 3), 1)
    st = db.last_plan_stats()
    db.free()
    strip_ms = float(np.mean([sum(t[2] for t in s if t[0] == 2) for s in tims]))
    tops = cells * GOTOH_OPS_PER_CELL / (strip_ms / 1e3) / 1e12
    out = {"pairs": hb.npairs, "halves": [20, 40, 56], "cells": cells, "strip_pairs": st["strip_pairs"],
           "ms_per_step": round(wall * 1e3, 3), "strip_kernel_ms": round(strip_ms, 3),
           "gcups": round(cells / (strip_ms / 1e3) / 1e9, 1), "roofline_frac": round(tops / VALU_PEAK_TOPS, 4)}
    log(f"strip kernel (halves 20/40/56): {out}")
    return out


def bench_plain(args, buf, params, c1=None):
    """The plain one-launch-per-call form (esp_band, one wavefront per pair)
    on the first C1 pairs: calls/s with each call's copies and sync, then each
    call's HIP-event kernel time (a second pass, so that the read-back stays
    out of the wall-clock rate), GCUPS and the fraction of one wave's issue
    ceiling."""
    from esperanto_amd.batch import band_cells as one_band_cells, workload_c1
    from esperanto_amd.extend import band as plain_band
    if c1 is None:
        c1 = workload_c1(10_000)
    ncalls = min(args.plain_calls, c1.npairs)
    pairs = [c1.pair(k) for k in range(ncalls)]
    plain_band(pairs[0][0], pairs[0][1], params, (32, 32), buf)
    t0 = time.perf_counter()
    for r, w in pairs:
        plain_band(r, w, params, (32, 32), buf)
    pps = ncalls / (time.perf_counter() - t0)
    kms = []
    for r, w in pairs:
        plain_band(r, w, params, (32, 32), buf)
        kms.append(buf.last_kernel_ms())
    cells = sum(one_band_cells(len(r), len(w), 32, 32) for r, w in pairs)
    ks = float(np.sum(kms)) / 1e3
    tops = cells * GOTOH_OPS_PER_CELL / ks / 1e12
    kern = {"kernel": "k_gotoh_wave", "calls": ncalls, "us_per_call": round(1e6 * ks / ncalls, 2),
            "gcups": round(cells / ks / 1e9, 4), "roofline_frac_chip": round(tops / VALU_PEAK_TOPS, 7),
            "roofline_frac_one_wave": round(tops / ONE_WAVE_PEAK_TOPS, 4)}
    log(f"plain form: {pps:.0f} pairs/s, kernel {kern['us_per_call']} us/call, {kern['gcups']} GCUPS")
    return {"pairs_per_s": round(pps), "kernel": kern}


def bench_junction(args, buf, params):
    """align_spliced over chains (segment_chain, refine_junction_read per seam,
    the guards, then the spliced pipeline), try_intron_chain_placement (MEMs,
    fragments, EA-Myers refinements and verification, junction, DP) and
    reextend_with (window from the clips, extend_hint, acceptance) over
    junction_workload reads on a 64 Mb reference: reads/s per call with
    inputs resident, device ms and stage times, and the first reads checked
    against the oracle."""
    from esperanto_amd.junction import (DeviceChainBatch, DeviceIntronChainBatch, DeviceReextendBatch,
                                        JunctionLib)
    from esperanto_amd.reference import DeviceReference
    from esperanto_amd.synth_window import junction_workload
    n = args.junction_reads
    log(f"junction workload: {n} reads on a {args.window_ref_mb} Mb reference")
    w = junction_workload(n, ref_len=args.window_ref_mb * 1_000_000)
    dref = DeviceReference(w["reference"], buf)
    jlib = JunctionLib(dref, [tuple(int(x) for x in q) for q in w["quads"]])
    k = w["k"]
    legs = (
        ("align_spliced", DeviceChainBatch(dref, w["reads"], w["read_off"], w["read_len"], w["contig"],
                                           w["minus"], w["chain_anchors"], w["chain_off"]),
         lambda db: db.launch(jlib, k, extend_params=params)),
        ("intron_chain", DeviceIntronChainBatch(dref, w["reads"], w["read_off"], w["read_len"],
                                                w["pool_anchors"], w["pool_off"]),
         lambda db: db.launch(jlib, k, extend_params=params)),
        ("reextend", DeviceReextendBatch(dref, w["rx_reads"], w["read_off"], w["read_len"], w["contig"],
                                         w["rx_pos"], w["rx_cigar"], w["rx_cig_off"], w["rx_cig_n"]),
         lambda db: db.launch(params, 0)),
    )
    out = {"reads": n, "reference_bases": int(args.window_ref_mb * 1_000_000)}
    kinds = {0: "plan", 1: "band30", 2: "strip", 3: "myers", 5: "prep", 6: "select_finish",
             7: "escalated_full", 11: "escalated_unpruned", 12: "escalation_sort", 13: "escalated_wave"}
    for name, db, launch in legs:
        wall, kms, tims = timed_steps(lambda: launch(db), db, max(args.steps, 3), 1)
        stages = {}
        for st in tims:
            for kind, half, ms in st:
                stages.setdefault(kinds.get(kind, str(kind)), []).append(ms)
        stages = {k2: round(float(np.sum(v)) / len(tims), 3) for k2, v in stages.items()}
        rec = db.records()
        ok = int((rec["status"] == 1).sum())
        out[name] = {"reads_per_s": round(n / wall), "ms_per_step": round(wall * 1e3, 3),
                     "device_ms": round(kms, 3), "stages_ms": stages, "status_1_fraction": round(ok / n, 4)}
        log(f"{name}: {n / wall / 1e6:.2f} M reads/s, {wall * 1e3:.2f} ms/step, device {kms:.2f} ms, "
            f"status 1 on {ok / n:.3f}, stages {stages}")
    if not args.no_cpu:
        # the first reads against the oracle, timed as the CPU port (one thread)
        from oracle import oracle
        o = oracle.OracleRef(w["reference"])
        jl = oracle.JLib([tuple(int(x) for x in q) for q in w["quads"]])
        m = int(w["read_len"][0])
        ns = min(n, args.junction_check)
        cpu = {}
        chains, icres, rxres = legs[0][1].records(), legs[1][1].records(), legs[2][1].records()
        t0 = time.perf_counter()
        for c in range(ns):
            a0, a1 = int(w["chain_off"][c]), int(w["chain_off"][c + 1])
            anc = [(int(w["chain_anchors"][2 * i]), int(w["chain_anchors"][2 * i + 1])) for i in range(a0, a1)]
            r = oracle.align_spliced(o, jl, bytes(w["reads"][m * c:m * c + m]), 0, False, anc, k, params=params)
            assert r["status"] == int(chains[c]["status"]) and r["genomic_start"] == int(chains[c]["genomic_start"]), c
        t1 = time.perf_counter()
        for c in range(ns):
            a0, a1 = int(w["pool_off"][c]), int(w["pool_off"][c + 1])
            pool = [tuple(int(x) for x in w["pool_anchors"][4 * i:4 * i + 4]) for i in range(a0, a1)]
            r = oracle.intron_chain(o, jl, bytes(w["reads"][m * c:m * c + m]), pool, k, params=params)
            assert r["status"] == int(icres[c]["status"]) and r["score"] == int(icres[c]["score"]), c
        t2 = time.perf_counter()
        for c in range(ns):
            r = oracle.reextend_with(o, bytes(w["rx_reads"][m * c:m * c + m]), 0, int(w["rx_pos"][c]),
                                     [int(x) for x in w["rx_cigar"][3 * c:3 * c + 3]], params, 0)
            assert r["status"] == int(rxres[c]["status"]) and r["pos"] == int(rxres[c]["pos"]), c
        t3 = time.perf_counter()
        cpu = {"align_spliced_reads_per_s": round(ns / (t1 - t0)), "intron_chain_reads_per_s": round(ns / (t2 - t1)),
               "reextend_reads_per_s": round(ns / (t3 - t2)), "cores": 1, "sample_reads": ns,
               "matches_device": True}
        out["cpu_port"] = cpu
        log(f"junction cpu port (and device check on {ns} reads): {cpu}")
    for _, db, _ in legs:
        db.free()
    jlib.close()
    dref.close()
    return out


def _copy_calibration(buf, src, nbytes):
    """One plain device copy of nbytes (at most src's size): the yardstick that
    the FETCH_SIZE / WRITE_SIZE counters are scaled by."""
    import ctypes
    from esperanto_amd import _native
    from esperanto_amd.batch import DeviceArray
    nbytes = min(nbytes, src.nbytes)
    dst = DeviceArray(buf, nbytes)
    _native.check(_native.lib().esp_device_copy(buf.handle, ctypes.c_void_p(dst.ptr),
                                                ctypes.c_void_p(src.ptr), nbytes), buf.handle)
    log(f"copy calibration: {nbytes} bytes")
    dst.free()


def bench_myers(args, buf):
    """Batched EA-Myers on the 150-nt / 214-nt workload (distance and end):
    (pairs/s, kernel ms, TOP/s at the algorithmic 31 ops per 64-bit word-column
    + 3 per column, pairs/s of the plain one-launch-per-pair form)."""
    from esperanto_amd.batch import DeviceBatch, workload_c2
    c2 = workload_c2(args.myers_pairs)
    d2 = DeviceBatch(c2, buf)
    w2, k2, _ = timed_steps(lambda: d2.launch_myers(False), d2, max(args.steps, 3), 1)
    myers_pps = c2.npairs / w2
    words = 3
    myers_ops = c2.npairs * 214 * (MYERS_OPS_PER_WORD_COL * words + 3)
    myers_tops = myers_ops / (k2 / 1e3) / 1e12
    log(f"myers: {myers_pps/1e6:.1f} M pairs/s (kernel {k2:.2f} ms)")
    # The same batch with its read-length bound stated: no host round trip.
    maxm = int(c2.read_len.max())
    wb, kb, _ = timed_steps(lambda: d2.launch_myers(False, max_read_len=maxm), d2, max(args.steps, 3), 1)
    args._myers_bounded = (c2.npairs / wb, kb)
    log(f"myers, bound {maxm} stated: {c2.npairs / wb / 1e6:.1f} M pairs/s (kernels {kb:.2f} ms)")
    d2.free()
    # The plain form: one esp_myers_best_end call (one launch) per pair.
    from esperanto_amd.myers_ea import infix_best_end
    ncalls = min(args.plain_calls, c2.npairs)
    pairs = [c2.pair(k) for k in range(ncalls)]
    infix_best_end(pairs[0][0], pairs[0][1], buf)
    t0 = time.perf_counter()
    for r, w in pairs:
        infix_best_end(r, w, buf)
    plain_pps = ncalls / (time.perf_counter() - t0)
    log(f"myers plain: {plain_pps:.0f} pairs/s ({ncalls} calls)")
    return myers_pps, k2, myers_tops, plain_pps


def bench_walk(args, buf, params):
    """The three placement walks (rescore_with, recount_mm_ea_in,
    exact_kmer_support) over the placements of the pair workload: per pair,
    R1 as uM (intron)N (m-u)M and R2 as 150M, half of the planted junctions
    in the library."""
    from esperanto_amd.reference import DeviceReference
    from esperanto_amd.synth_window import placement_walk_workload
    from esperanto_amd.walk import DeviceJunctions, DeviceWalkBatch
    n = args.walk_pairs
    log(f"placement walks: {2 * n} placements on a {args.window_ref_mb} Mb reference")
    ref, pl, jn = placement_walk_workload(n, ref_len=args.window_ref_mb * 1_000_000)
    dref = DeviceReference(ref, buf)
    dj = DeviceJunctions(jn, buf)
    db = DeviceWalkBatch(pl, dref, buf)
    w, k, tims = timed_steps(lambda: db.launch(params, None, dj), db, max(args.steps, 3), 1)
    walk_ms = float(np.mean([sum(t[2] for t in st if t[0] == 8) for st in tims]))
    # algorithmic bytes: query bytes, 37 B of per-placement arrays, CIGAR
    # words, the 2-bit reference under the M columns (every query base sits
    # in an M column here), 64 B of records out
    nb = int(pl.read_len.sum()) + 37 * pl.n + 4 * int(pl.cig_n.sum()) + 64 * pl.n
    nb += (int(pl.read_len.sum()) + 3) // 4
    out = {"placements": pl.n, "placements_per_s": round(pl.n / w), "ms_per_step": round(w * 1e3, 3),
           "kernel_ms": round(walk_ms, 3), "algorithmic_bytes": nb,
           "kernel_gb_per_s": round(nb / (walk_ms / 1e3) / 1e9, 1), "junctions": len(jn)}
    log(f"walks: {pl.n / w / 1e6:.2f} M placements/s, kernel {walk_ms:.3f} ms")
    if args.copy_bytes:
        _copy_calibration(buf, db.d["reads"], args.copy_bytes)
    db.free()
    # The Kadane fast lane over the R2 placements as chains (anchors every 15
    # bases on the true diagonal, the last one at m - 15).
    from esperanto_amd.fast_lane import DeviceFastBatch
    r2 = np.arange(1, pl.n, 2)
    m = int(pl.read_len[1])
    qp = np.unique(np.append(np.arange(0, m - 15 + 1, 15), m - 15)).astype(np.int64)
    na = qp.shape[0]
    anchors = np.stack([np.broadcast_to(qp, (r2.shape[0], na)),
                        pl.pos[r2].astype(np.int64)[:, None] + qp[None, :]], axis=2)
    fb = DeviceFastBatch(dref, pl.reads, pl.read_off[r2], pl.read_len[r2], pl.contig[r2], anchors,
                         np.arange(r2.shape[0], dtype=np.uint64) * na, np.full(r2.shape[0], na, np.uint32))
    wf, _, tf = timed_steps(lambda: fb.launch(params), fb, max(args.steps, 3), 1)
    fast_ms = float(np.mean([sum(t[2] for t in st if t[0] == 9) for st in tf]))
    rec, _ = fb.results()
    out["fast_lane"] = {"chains": int(r2.shape[0]), "chains_per_s": round(r2.shape[0] / wf),
                        "kernel_ms": round(fast_ms, 3), "accepted": round(float((rec["status"] == 1).mean()), 4)}
    log(f"fast lane: {r2.shape[0] / wf / 1e6:.1f} M chains/s, kernel {fast_ms:.3f} ms, "
        f"accepted {out['fast_lane']['accepted']}")
    fb.free()
    dj.close()
    if not args.no_cpu:
        from oracle import oracle
        threads = min(16, os.cpu_count() or 1)
        o = oracle.OracleRef(ref)
        t0 = time.perf_counter()
        oracle.walk_batch(o, pl, params, oracle.SPLICE_DEFAULTS, jn, 15, threads)
        dt = time.perf_counter() - t0
        out["cpu_port"] = {"placements_per_s": round(pl.n / dt), "cores": threads, "sample": pl.n}
        log(f"walks cpu port: {out['cpu_port']}")
    dref.close()
    return out


def _serve_open_loop(srv, submit, pool, sizes, peak, seconds, rng):
    """Open-loop Poisson arrivals at 25/50/90% of peak (pairs/s) for the given
    time each; latency = completion minus scheduled arrival."""
    import queue
    import threading
    from esperanto_amd.server import now_ns
    loads = []
    for frac in (0.25, 0.5, 0.9):
        rate = frac * peak / float(sizes.mean())  # requests per second
        nreq = max(1, int(rate * seconds))
        gaps = rng.exponential(1.0 / rate, nreq)
        order = rng.integers(0, len(pool), nreq)
        q = queue.Queue()
        lat = []

        def completer():
            while True:
                item = q.get()
                if item is None:
                    return
                t_sched, t = item
                lat.append((t.wait().done_ns - t_sched) / 1e6)

        th = threading.Thread(target=completer)
        th.start()
        start = now_ns() + 20_000_000
        sched = start + np.cumsum(gaps * 1e9).astype(np.int64)
        pairs = 0
        for k in range(nreq):
            while True:
                d = int(sched[k]) - now_ns()
                if d <= 0:
                    break
                if d > 300_000:
                    time.sleep((d - 200_000) / 1e9)
            r = pool[order[k]]
            q.put((int(sched[k]), submit(r)))
            pairs += r["read_len"].shape[0]
        q.put(None)
        th.join()
        span = (max(start, now_ns()) - start) / 1e9
        lat_a = np.array(lat)
        loads.append({"offered_fraction": frac, "offered_pairs_per_s": round(frac * peak),
                      "achieved_pairs_per_s": round(pairs / span), "requests": nreq,
                      "p50_ms": round(float(np.percentile(lat_a, 50)), 3),
                      "p99_ms": round(float(np.percentile(lat_a, 99)), 3),
                      "max_ms": round(float(lat_a.max()), 3)})
        log(f"serving load {frac:.2f}: {loads[-1]}")
    return loads


def bench_serving(args, buf, params, c3_pairs_per_s=None):
    """Serving loop: a pool of mixed requests (log-uniform 1-4,096 pairs of 50-250-nt
    reads; 10% mate-rescue requests with ~2 kb windows) over a 256 Mb reference
    resident in HBM.  For each batch-size cap (pairs per batch: the default and a
    small one that trades throughput for latency): closed loop first (the whole
    pool at once: that configuration's peak), then open-loop Poisson arrivals at
    25/50/90% of that peak for a fixed time."""
    from esperanto_amd.reference import DeviceReference
    from esperanto_amd.server import Server, now_ns
    from esperanto_amd.synth_window import serving_workload
    log(f"serving workload: {args.serve_requests} requests on a {args.serve_ref_mb} Mb reference")
    ref, pool = serving_workload(args.serve_requests, ref_len=args.serve_ref_mb * 1_000_000)
    dref = DeviceReference(ref, buf)
    sizes = np.array([r["read_len"].shape[0] for r in pool])
    out = {"requests_in_pool": len(pool), "pairs_in_pool": int(sizes.sum()),
           "workers": args.serve_workers, "mem_budget_bytes": 32 << 30, "caps": []}
    rng = np.random.default_rng(5)
    for cap in (args.serve_batch_pairs, args.serve_small_batch_pairs):
        srv = Server(dref, params, mem_budget=32 << 30, workers=args.serve_workers, max_batch_pairs=cap)

        def submit(r):
            return srv.submit_arrays(r["kind"], r["reads"], r["read_len"], r["contig"],
                                     r.get("win_start"), r.get("win_end"), r.get("hint"),
                                     r.get("mate_pos"))

        # warm-up: one untimed pass of the whole pool grows every worker's
        # scratch and pinned staging to the batch sizes the timed passes use
        for t in [submit(r) for r in pool]:
            t.wait()
        t0 = now_ns()
        tickets = [submit(r) for r in pool]
        done = max(t.wait().done_ns for t in tickets)
        peak = float(sizes.sum()) / ((done - t0) / 1e9)
        log(f"serving cap {cap}: closed loop {len(pool)} requests, {sizes.sum()} pairs, {peak / 1e6:.2f} M pairs/s")
        verified = 0
        if not args.no_cpu and cap == args.serve_batch_pairs:
            verified = _serve_verify(ref, pool, tickets, params)
        del tickets
        loads = _serve_open_loop(srv, submit, pool, sizes, peak, args.serve_seconds, rng)
        st = srv.stats()
        srv.close()
        out["caps"].append({"max_batch_pairs": cap, "closed_loop_pairs_per_s": round(peak), "loads": loads,
                            "verified_requests_bit_exact": verified,
                            "max_batch_requests": st["max_batch_requests"], "batches": st["batches"],
                            "peak_inflight_bytes": st["peak_inflight_bytes"]})
    dref.close()
    # the default cap's figures at the top level
    out.update({k: out["caps"][0][k] for k in ("closed_loop_pairs_per_s", "loads", "verified_requests_bit_exact",
                                               "max_batch_requests", "batches", "peak_inflight_bytes")})
    if c3_pairs_per_s:
        # Loads at fractions of the batched C3 rate overload this mix: its pairs
        # are extend_hint calls (a band attempt, with full-matrix reruns for
        # the escalated ones) and 10% carry ~2 kb rescue windows, far more work
        # per pair than a C3 band pair, so even 25% of that rate exceeds the
        # mix's closed-loop peak and the queue would grow without bound.
        out["c3_batched_pairs_per_s"] = round(c3_pairs_per_s)
        out["c3_25pct_over_mix_peak"] = round(0.25 * c3_pairs_per_s / out["closed_loop_pairs_per_s"], 2)
    return out


def _serve_verify(ref, pool, tickets, params):
    """Bit-exactness of 24 sampled requests against the oracle."""
    from oracle import oracle
    from esperanto_amd.batch import PairBatch
    from esperanto_amd.rescue import MateRequests
    o = oracle.OracleRef(ref)
    rng = np.random.default_rng(1)
    verified = 0
    for j in rng.choice(len(pool), min(24, len(pool)), replace=False):
        r, t = pool[j], tickets[j]
        n = r["read_len"].shape[0]
        off = np.zeros(n + 1, dtype=np.int64)
        off[1:] = np.cumsum(r["read_len"])
        rec, cig, coff, cn = t.raw()
        if r["kind"] == 0:
            reads = [r["reads"][off[k]:off[k + 1]].tobytes() for k in range(n)]
            wins = [oracle.decode(o, 0, int(r["win_start"][k]), int(r["win_end"][k])) for k in range(n)]
            pb = PairBatch.from_pairs([(reads[k], wins[k], 32, 0) for k in range(n)])
            e, _, ec, eo, en, _ = oracle.band_batch(pb, params, 2, nthreads=16)
            got = rec.astype(np.int64)
            got[:, 3] = rec[:, 3].view(np.int32)
            ok = np.array_equal(got, e) and np.array_equal(cn, en)
        else:
            mq = MateRequests(r["reads"], off[:-1].astype(np.uint64), r["read_len"], r["contig"],
                              r["mate_pos"])
            e, ec, eo, en = oracle.mate_batch(o, mq, params, nthreads=16)
            ok = [x.tolist() for x in rec] == [x.tolist() for x in e] and np.array_equal(cn, en)
        ok = ok and all(np.array_equal(cig[int(coff[k]):int(coff[k]) + int(cn[k])],
                                       ec[int(coff[k]):int(coff[k]) + int(cn[k])]) for k in range(n))
        if not ok:
            raise RuntimeError(f"serving result of request {j} differs from the CPU path")
        verified += 1
    return verified


def _subset_splice(sp, k):
    from esperanto_amd.splice import SpliceRequests
    e1 = int(sp.exon_off[k])
    r1 = int(sp.read_off[k - 1] + sp.read_len[k - 1])
    return SpliceRequests(sp.reads[:r1], sp.read_off[:k], sp.read_len[:k], sp.contig[:k],
                          sp.hint[:k], sp.exons[:e1], sp.exon_off[:k + 1])


def _subset_mate(mt, k):
    from esperanto_amd.rescue import MateRequests
    r1 = int(mt.read_off[k - 1] + mt.read_len[k - 1])
    return MateRequests(mt.reads[:r1], mt.read_off[:k], mt.read_len[:k], mt.contig[:k],
                        mt.mate_pos[:k])


def bench_cpu(args, params):
    """The reference's CPU route restated in C (oracle/esp_oracle.c): run_dispatch
    sends these pure pairs to the AVX2 i16 anti-diagonal kernel avx2_band
    (crates/map/src/extend.rs:293-344, :1526-2026), one worker per thread with
    static chunks as rayon's map_init would give.  Timed on all host threads (up
    to 16) and on one, plus the scalar run_band restatement on all threads;
    each leg is a bounded prefix of the same uneven workload."""
    from oracle import oracle
    from esperanto_amd.batch import workload_c3
    threads = min(16, os.cpu_count() or 1)
    avx = oracle.avx2_available()

    def leg(route, nthreads, seconds):
        oracle.set_route(route)
        try:
            probe = workload_c3(2000)
            t0 = time.perf_counter()
            oracle.band_batch(probe, params, 0, nthreads=nthreads)
            per_pair = (time.perf_counter() - t0) / probe.npairs
            n = int(min(args.pairs, max(2000, seconds / max(per_pair, 1e-9))))
            sample = workload_c3(n)
            cells = band_cells(sample)
            t0 = time.perf_counter()
            oracle.band_batch(sample, params, 0, nthreads=nthreads)
            dt = time.perf_counter() - t0
        finally:
            oracle.set_route(0)
        return cells / dt / 1e9, n, n / dt

    route = 1 if avx else 0
    g_all, n_all, pps_all = leg(route, threads, args.cpu_seconds)
    g_one, n_one, pps_one = leg(route, 1, args.cpu_seconds / 3)
    g_scalar, n_scalar, _ = leg(0, threads, args.cpu_seconds / 3)
    name = ("C restatement of the reference's AVX2 route (run_dispatch -> avx2_band i16 anti-diagonal kernel "
            "+ finish_extension)" if avx else "scalar C restatement of run_dispatch/run_band (no AVX2 on this host)")
    log(f"cpu baseline: {name}: {g_all:.3f} GCUPS on {threads} threads ({n_all} pairs), "
        f"{g_one:.3f} on 1 thread ({n_one} pairs); scalar port {g_scalar:.3f} on {threads} threads")
    return {"value": round(g_all, 3), "unit": "GCUPS", "cores": threads, "kind": "port",
            "sample": f"first {n_all} pairs of the same uneven workload (seed 0xE5A00003), {name}, "
                      f"{pps_all:.0f} pairs/s",
            "route": "avx2_band" if avx else "run_band",
            "single_thread_gcups": round(g_one, 3), "single_thread_sample_pairs": n_one,
            "scalar_port_gcups": round(g_scalar, 3), "scalar_port_sample_pairs": n_scalar}


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--pairs", type=int, default=2_000_000)
    ap.add_argument("--myers-pairs", type=int, default=1_000_000)
    ap.add_argument("--plain-calls", type=int, default=300)
    ap.add_argument("--cpu-seconds", type=float, default=10.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--window-pairs", type=int, default=500_000,
                    help="read pairs of the splice + mate-rescue leg (0 skips it)")
    ap.add_argument("--window-ref-mb", type=int, default=64)
    ap.add_argument("--window-cpu-pairs", type=int, default=20_000)
    ap.add_argument("--only-window", action="store_true",
                    help="run only the splice + mate-rescue leg")
    ap.add_argument("--serve-requests", type=int, default=1500,
                    help="requests in the serving pool (0 skips the serving leg)")
    ap.add_argument("--serve-ref-mb", type=int, default=256)
    ap.add_argument("--serve-seconds", type=float, default=4.0)
    ap.add_argument("--serve-workers", type=int, default=3)
    ap.add_argument("--serve-batch-pairs", type=int, default=131072)
    ap.add_argument("--serve-small-batch-pairs", type=int, default=8192,
                    help="the smaller batch-size cap the serving leg also runs (latency over throughput)")
    ap.add_argument("--only-serve", action="store_true", help="run only the serving leg")
    ap.add_argument("--walk-pairs", type=int, default=1_000_000,
                    help="read pairs of the placement-walk leg (two placements each; 0 skips it)")
    ap.add_argument("--only-walk", action="store_true", help="run only the placement-walk leg")
    ap.add_argument("--junction-reads", type=int, default=200_000,
                    help="reads of the align_spliced / intron-chain / reextend leg (0 skips it)")
    ap.add_argument("--junction-check", type=int, default=2000,
                    help="reads of that leg checked against the oracle and timed on the CPU")
    ap.add_argument("--only-junction", action="store_true", help="run only the junction leg")
    ap.add_argument("--strip-pairs", type=int, default=200_000,
                    help="pairs of the strip-kernel leg (band halves without a band-kernel instance; 0 skips it)")
    ap.add_argument("--only-strip", action="store_true", help="run only the strip-kernel leg")
    ap.add_argument("--only-plain", action="store_true",
                    help="run only the plain one-launch-per-call form on the 10k C1 pairs")
    ap.add_argument("--no-graph", action="store_true", help="skip the graph replay of the Gotoh batch")
    ap.add_argument("--only-myers", action="store_true", help="run only the batched Myers leg")
    ap.add_argument("--only-gotoh", action="store_true",
                    help="time only the batched Gotoh workload (profiling runs)")
    ap.add_argument("--no-tune", action="store_true",
                    help="keep the band grids at the occupancy limit and the pruned passes at their default shape")
    ap.add_argument("--band-grid", type=str, default="",
                    help="blocks per CU for halves 16,30,32,48,64 (an earlier autotune's choice; skips tuning)")
    ap.add_argument("--copy-bytes", type=int, default=0,
                    help="also run one plain device copy of this size (counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only: this build has no multi-GPU path "
              f"(--gpus {args.gpus} requested); pairs shard independently across GPUs, "
              f"see the sharding design notes.", file=sys.stderr)
        return 2

    from esperanto_amd.batch import DeviceBatch, workload_c1, workload_c2, workload_c3
    from esperanto_amd.extend import ExtendBuffer, band as plain_band
    from esperanto_amd.records import ExtendParams

    params = ExtendParams()
    buf = ExtendBuffer(0)
    if args.only_window:
        print(json.dumps(bench_window(args, buf, params)), flush=True)
        buf.close()
        return 0
    if args.only_serve:
        print(json.dumps(bench_serving(args, buf, params)), flush=True)
        buf.close()
        return 0
    if args.only_myers:
        pps, kms, tops, plain = bench_myers(args, buf)
        print(json.dumps({"myers_pairs_per_s": round(pps), "kernel_ms": round(kms, 3),
                          "valu_tops": round(tops, 2), "plain_pairs_per_s": round(plain)}), flush=True)
        buf.close()
        return 0
    if args.only_walk:
        print(json.dumps(bench_walk(args, buf, params)), flush=True)
        buf.close()
        return 0
    if args.only_strip:
        print(json.dumps(bench_strip(args, buf, params)), flush=True)
        buf.close()
        return 0
    if args.only_plain:
        print(json.dumps(bench_plain(args, buf, params)), flush=True)
        buf.close()
        return 0
    if args.only_junction:
        print(json.dumps(bench_junction(args, buf, params)), flush=True)
        buf.close()
        return 0

    log(f"generating {args.pairs} uneven pairs")
    hb = workload_c3(args.pairs)
    cells = band_cells(hb)
    log(f"cells={cells} uploading")
    db = DeviceBatch(hb, buf)
    tune = None
    if args.band_grid:
        # a grid chosen by an earlier autotune run, applied without tuning
        from esperanto_amd import _native
        import ctypes
        tune = {}
        for half, k in zip((16, 30, 32, 48, 64), (int(x) for x in args.band_grid.split(","))):
            _native.check(_native.lib().esp_ctx_set_option(buf.handle, 3, ctypes.c_int64(half << 16 | k)),
                          buf.handle)
            tune[half] = (k,)
        log(f"band grids set: {tune}")
    elif not args.no_tune:
        # the band kernels' grids, tuned on this batch (a pipeline tunes on
        # its first chunk: the best persistent grid depends on the batch size)
        tune = db.autotune_band(params, reps=1)
        log(f"autotune (half: chosen blocks/CU, occupancy limit, ms at limit, ms chosen): {tune}")
    wall, kms, tims = timed_steps(lambda: db.launch_gotoh(params, 0), db, args.steps, args.warmup)
    gcups = cells / wall / 1e9
    kernel_gcups = cells / (kms / 1e3) / 1e9
    # Dominant kernel: the band kernel (one launch per band half per step).
    band_ms = float(np.mean([sum(t[2] for t in st if t[0] == 1) for st in tims]))
    plan_ms = float(np.mean([sum(t[2] for t in st if t[0] == 0) for st in tims]))
    strip_ms = float(np.mean([sum(t[2] for t in st if t[0] == 2) for st in tims]))
    per_half = {}
    for st in tims:
        for kind, half, ms in st:
            if kind == 1:
                per_half.setdefault(half, []).append(ms)
    per_half = {int(h): round(float(np.mean(v)), 3) for h, v in sorted(per_half.items())}
    achieved_tops = cells * GOTOH_OPS_PER_CELL / (band_ms / 1e3) / 1e12
    log(f"gotoh batched: {wall*1e3:.2f} ms/step, device {kms:.2f} ms (plan {plan_ms:.3f}, "
        f"band {band_ms:.2f} {per_half}, strip {strip_ms:.3f}), {gcups:.1f} GCUPS")
    pairs_per_s = hb.npairs / wall
    # The same batch captured once as a HIP graph and replayed.
    graph_wall = float("nan")
    if not args.no_graph:
        g3 = db.capture_gotoh(params, 0)
        graph_wall = _graph_steps(g3, db, args.steps, args.warmup)
        g3.check()
        g3.destroy()
        log(f"gotoh batched, graph replay: {graph_wall*1e3:.2f} ms/step, {cells / graph_wall / 1e9:.1f} GCUPS")
    if args.copy_bytes:
        _copy_calibration(buf, db.d["wins"], args.copy_bytes)
    db.free()
    if args.only_gotoh:
        print(json.dumps({"gcups": round(gcups, 2), "band_ms": round(band_ms, 3),
                          "per_half_ms": per_half, "plan_ms": round(plan_ms, 3)}), flush=True)
        buf.close()
        return 0

    # C1 batched and plain forms.
    c1 = workload_c1(10_000)
    d1 = DeviceBatch(c1, buf)
    w1, k1, _ = timed_steps(lambda: d1.launch_gotoh(params, 0), d1, max(args.steps, 3), 1)
    c1_batched_pps = c1.npairs / w1
    g1 = d1.capture_gotoh(params, 0)
    w1g = _graph_steps(g1, d1, max(args.steps, 10), 2)
    g1.check()
    g1.destroy()
    c1_graph_pps = c1.npairs / w1g
    log(f"c1 graph replay: {c1_graph_pps:.0f} pairs/s ({w1g * 1e3:.3f} ms/step)")
    d1.free()
    pl = bench_plain(args, buf, params, c1)
    plain_pps = pl["pairs_per_s"]
    myers_pps, k2, myers_tops, myers_plain_pps = bench_myers(args, buf)

    # Splice + mate-rescue window pipelines on a 64 Mb reference.
    win = None
    if args.window_pairs:
        win = bench_window(args, buf, params)
    walks = None
    if args.walk_pairs:
        walks = bench_walk(args, buf, params)
    junction = None
    if args.junction_reads:
        junction = bench_junction(args, buf, params)
    strip = None
    if args.strip_pairs:
        strip = bench_strip(args, buf, params)
    serve = None
    if args.serve_requests:
        serve = bench_serving(args, buf, params, pairs_per_s)

    cpu = None
    if not args.no_cpu:
        cpu = bench_cpu(args, params)

    out = {
        "metric": "Gotoh GCUPS + Myers pairs/s, batched vs plain on 1 MI355X; p99 latency",
        "value": round(gcups, 2),
        "unit": "GCUPS",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(wall * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int32",
        "data": "synthetic",
        "config": {"workload": "batched uneven banded Gotoh+traceback: 2,000,000 pairs, reads "
                               "U{50..250}, band U{16,32,48,64}, window=read+2*band",
                   "pairs": hb.npairs, "cells": cells, "kernel": "k_gotoh_band",
                   "pairs_per_s": round(pairs_per_s),
                   "device_ms": round(kms, 3), "planner_ms": round(plan_ms, 3),
                   "band_kernel_ms": round(band_ms, 3), "band_kernel_ms_per_half": per_half,
                   "kernel_gcups": round(kernel_gcups, 2),
                   "autotune": {str(h): v[0] for h, v in tune.items()} if tune else None,
                   "graph_ms_per_step": None if args.no_graph else round(graph_wall * 1e3, 3),
                   "graph_gcups": None if args.no_graph else round(cells / graph_wall / 1e9, 2),
                   "c1_batched_pairs_per_s": round(c1_batched_pps),
                   "c1_graph_pairs_per_s": round(c1_graph_pps),
                   "c1_plain_pairs_per_s": round(plain_pps),
                   "batched_over_plain": round(c1_batched_pps / plain_pps, 1),
                   "c1_plain_kernel": pl["kernel"],
                   "myers_pairs_per_s": round(myers_pps),
                   "myers_kernel_ms": round(k2, 3),
                   "myers_valu_tops": round(myers_tops, 2),
                   "myers_bounded_pairs_per_s": round(args._myers_bounded[0]),
                   "myers_plain_pairs_per_s": round(myers_plain_pps),
                   "myers_batched_over_plain": round(myers_pps / myers_plain_pps, 1),
                   "splice_rescue": win,
                   "placement_walks": walks,
                   "junction": junction,
                   "strip_kernel": strip,
                   "serving": serve},
        "roofline": {"bound": "valu", "achieved": round(achieved_tops, 3),
                     "peak": round(VALU_PEAK_TOPS, 2), "unit": "TOP/s",
                     "frac": round(achieved_tops / VALU_PEAK_TOPS, 4),
                     "traffic": read_traffic(os.path.join(ROOT, "profiles",
                                                          "gotoh_band_uneven_2M_traffic.txt"))},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)
    buf.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
