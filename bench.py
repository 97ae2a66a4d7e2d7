"""Benchmark: the hmmsearch filter pipeline (MSV -> bias -> Viterbi -> Forward
-> Backward) of one M=400 query over a resident 2M-sequence synthetic database
on one MI355X.

A step is one p7_gpu_Pipeline call over the whole database, inputs already in
HBM. value = nres * M / seconds per step, in GCUPS (the reference's Mc/sec
convention, HMMER/src/p7_pipeline.c:1791-1792, divided by 1000). Prints one
JSON line.

    python bench.py [--gpus 1] [--steps K] [--warmup W]
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

# The batched and serving legs run up to sixteen stage chains (and several
# query slots) on their own HIP streams. HIP maps streams onto
# GPU_MAX_HW_QUEUES hardware queues per process (4 by default); streams that
# share a queue run their kernels in queue order, so one chain's long-target
# tail holds up the chains behind it. The runtime reads the variable when it
# initialises, so main() sets it from --hw-queues before the library (or
# torch) touches the GPU. 32 is the most this build asks for.
MAX_HW_QUEUES = 32

PROFILE = os.path.join(ROOT, "tests", "golden", "synth", "synth_M400.hmm")
PEAK_PK16_TOPS = 78.6      # MI355X packed 16-bit VALU ops/s (256 CU x 64 lanes x 2 x 2.4 GHz)
PEAK_HBM_GBS = 8000.0
SSV_OPS_PER_CELL = 2       # the ungapped pass: one saturating add + one max per cell (ssvfilter.c:433-435)


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def host_cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


# The reference's own SSE build (HMMER 3.1b2, -O3 -msse2), timed on one core of
# an 8-core Intel Xeon with LuxC.hmm (M = 400): per-stage filter rates and
# hmmsearch --cpu 1 / --cpu 8 over 200k sequences of L = 300. Quoted beside the
# port, not measured here: the reference's binary is not on this host.
REFERENCE_SSE_PER_CORE = {"host": "8-core Intel Xeon (build container), 1 thread per core",
                          "workload": "LuxC.hmm (M = 400) vs i.i.d. sequences of L = 400 (filters), "
                                      "200k x L = 300 (hmmsearch)",
                          "msv_gcups": "37-40", "ssv_gcups": 40.0, "viterbi_gcups": 2.1,
                          "forward_gcups": 0.19, "backward_gcups": 0.16,
                          "hmmsearch_cpu1_gcups": 17.9, "hmmsearch_cpu8_gcups": 57.1}


def cpu_baseline(om, db, M, gpu_stage, gpu_usc=None, seconds_target=15.0, max_threads=16):
    """The oracle's pipeline on a bounded sample, on the host cores; GCUPS.
    The sample's stages are also compared with the device run's stages
    (gpu_stage, in database order): oracle_mismatches counts the targets whose
    stage differs, oracle_usc_mismatches those whose MSV score differs."""
    import oracle
    t = om.tables()
    prof, keep = oracle.make_profile(t, oracle.bias_hmm(t), t["evparam"], Lmax=int(db.lengths.max()) + 1)
    nthreads = max(1, min(max_threads, os.cpu_count() or 1))
    n = 2000
    while True:
        sub = db.subset(np.arange(n))
        t0 = time.perf_counter()
        ref = oracle.pipeline_db(prof, sub.residues, sub.offsets, sub.lengths, nthreads)
        dt = time.perf_counter() - t0
        if dt >= seconds_target / 4 or n >= db.nseq:
            break
        n = min(db.nseq, int(n * max(2.0, seconds_target / 4 / max(dt, 1e-3))))
    cells = float(sub.nres) * M
    mism = int(np.sum(np.asarray(gpu_stage[:n]) != ref["stage"].astype(np.uint8)))
    usc_mism = None if gpu_usc is None else int(np.sum(np.asarray(gpu_usc[:n]) != ref["usc"]))
    return {"value": cells / dt / 1e9, "unit": "GCUPS", "cores": nthreads, "kind": "port",
            "cpu_model": host_cpu_model(),
            "sample": f"first {n} of {db.nseq} database sequences ({sub.nres} residues), full filter pipeline "
                      f"(oracle/p7_oracle.c restatement, scalar C), {nthreads} threads, {dt:.1f} s",
            "per_thread_gcups": round(cells / dt / 1e9 / nthreads, 3),
            "oracle_compared": n, "oracle_mismatches": mism, "oracle_usc_mismatches": usc_mism,
            "oracle_pass_counts": {"msv": int((ref["stage"] >= 1).sum()), "bias": int((ref["stage"] >= 2).sum()),
                                   "vit": int((ref["stage"] >= 3).sum()), "fwd": int((ref["stage"] >= 4).sum())},
            "reference_sse_per_core_quoted": REFERENCE_SSE_PER_CORE}


def serving_leg(p7gpu, synth, nseq, nq, rate, nslots, seed=11, tune=True):
    """A resident database of nseq i.i.d. sequences, 0.5% of them planted
    homologs of a hidden set of 1000 profiles (M ~ lognormal, median 250,
    clipped to [50, 3000]). Queries, drawn uniformly from 16 calibrated
    members of the set and 4 decoy profiles outside it, arrive as a Poisson
    stream and are served by p7_gpu_server_* (several in flight); the same
    queries run one by one through a single pipeline give the plain
    per-query rate. Both use the tuner's shapes for the queries' length
    classes, measured first. Reports p50/p99 latency, aggregate GCUPS of
    both, and whether every query's pass set is identical in both."""
    import tempfile
    import threading
    tmp = tempfile.mkdtemp(prefix="p7q_")
    t0 = time.perf_counter()
    fam, qpaths, dpaths = synth.serving_family(p7gpu, tmp, nfam=1000, nquery=16, ndecoy=4, seed=seed)
    db = synth.make_db(nseq, fam, frac_homolog=0.005, seed_db=seed + 1, seed_plant=seed + 2)
    del fam
    sdb = p7gpu.SeqDB(db.residues, db.offsets, db.lengths)
    log(f"serving leg: resident database of {db.nseq} sequences, {db.nres} residues "
        f"({time.perf_counter() - t0:.1f} s to generate and upload)")
    del db
    oms = [p7gpu.OProfile.from_file(pth) for pth in qpaths + dpaths]
    Ms = [om.tables()["M"] for om in oms]
    if tune:
        p7gpu.autotune(oms, sdb)                 # as a server would load its device's tuning file
    rng = np.random.default_rng(seed)
    which = rng.integers(0, len(oms), nq)
    arrivals = np.cumsum(rng.exponential(1.0 / rate, nq))
    cells = float(sum(Ms[q] for q in which)) * sdb.nres

    # warm the device copies of every profile, then the plain per-query path
    pli = p7gpu.Pipeline()
    for om in oms:
        pli.run(om, sdb, scores=False)
    p7gpu.lib.p7_gpu_Synchronize()
    t0 = time.perf_counter()
    plain = [pli.run(oms[q], sdb, scores=False)["stage"] for q in which]
    t_plain = time.perf_counter() - t0

    srv = p7gpu.Server(sdb, nslots=nslots)
    tickets = [None] * nq
    def feeder():
        start = time.perf_counter()
        for i in range(nq):
            dt = arrivals[i] - (time.perf_counter() - start)
            if dt > 0:
                time.sleep(dt)
            tickets[i] = srv.submit(oms[which[i]], scores=False)
    t0 = time.perf_counter()
    th = threading.Thread(target=feeder)
    th.start()
    lat, same = [], True
    for i in range(nq):
        while tickets[i] is None:
            time.sleep(0.0005)
        out, st, ms = srv.wait(tickets[i])
        lat.append(ms)
        same &= bool(np.array_equal(out["stage"], plain[i]))
    th.join()
    t_served = time.perf_counter() - t0
    # saturation: every query queued at once, one query per slot launch (the
    # server's default), then up to four
    t_sat = {}
    for c in (1, 4):
        srv.set_coalesce(c)
        t0 = time.perf_counter()
        tk = [srv.submit(oms[q], scores=False) for q in which]
        for i, t in enumerate(tk):
            out, st, ms = srv.wait(t)
            same &= bool(np.array_equal(out["stage"], plain[i]))
        t_sat[c] = time.perf_counter() - t0
    slots = srv.slots
    srv.close()
    import shutil
    shutil.rmtree(tmp, ignore_errors=True)
    lat = np.array(lat)
    return {"nseq": sdb.nseq, "nres": sdb.nres, "homologs": "0.5% planted from a hidden set of 1000 profiles",
            "query_profiles": "16 members of the set + 4 decoys, drawn uniformly",
            "queries": nq, "arrival_rate_qps": rate, "slots": slots, "profiles_M": Ms,
            "p50_ms": round(float(np.percentile(lat, 50)), 3), "p99_ms": round(float(np.percentile(lat, 99)), 3),
            "gcups_served_stream": round(cells / t_served / 1e9, 1),
            "gcups_served_saturated": round(cells / t_sat[1] / 1e9, 1),
            "gcups_served_saturated_four_per_launch": round(cells / t_sat[4] / 1e9, 1),
            "gcups_plain_per_query": round(cells / t_plain / 1e9, 1),
            "identical_pass_sets": same}


def batched_leg(p7gpu, synth, nprof, nseq, reps=3, seed=42, tune=True):
    """Many profiles of uneven M against one database with a long-sequence tail:
    the tuner picks each length class's ungapped-pass shape, then the same
    profiles run one by one through a pipeline (plain per-query drop-in) and
    as one p7_gpu_PipelineBatch call (one persistent launch per kernel shape,
    concurrent stage chains), launched directly and replayed from a captured
    HIP graph. Reports the aggregate GCUPS of each and whether every profile's
    pass set is the same in all."""
    import tempfile
    import shutil
    tmp = tempfile.mkdtemp(prefix="p7b_")
    t0 = time.perf_counter()
    paths = synth.batch_family(p7gpu, tmp, nprof=nprof, seed=seed)
    oms = [p7gpu.OProfile.from_file(pth) for pth in paths]
    Ms = [om.tables()["M"] for om in oms]
    hmms = [synth.read_hmm_probs(pth) for pth in paths[:8]]
    t_prof = time.perf_counter() - t0
    t0 = time.perf_counter()
    db = synth.make_db(nseq, hmms, frac_homolog=0.01, seed_db=seed + 1, seed_plant=seed + 2,
                       long_frac=0.02, long_range=(5000, 35000))
    sdb = p7gpu.SeqDB(db.residues, db.offsets, db.lengths)
    t_db = time.perf_counter() - t0
    log(f"batched leg: {nprof} profiles (M {min(Ms)}..{max(Ms)}) in {t_prof:.1f} s, "
        f"db {db.nseq} sequences / {db.nres} residues in {t_db:.1f} s")
    cells = float(sum(Ms)) * db.nres

    t0 = time.perf_counter()
    if tune:
        p7gpu.autotune(oms, sdb)
    t_tune = time.perf_counter() - t0
    shapes = {}
    for M in Ms:
        lanes, grab, _, _ = p7gpu.tune_get(M)
        shapes[f"{lanes}x{grab}"] = shapes.get(f"{lanes}x{grab}", 0) + 1

    pli = p7gpu.Pipeline()
    for om in oms:                                   # device copies of every profile
        pli.run(om, sdb, scores=False)
    p7gpu.lib.p7_gpu_Synchronize()
    t0 = time.perf_counter()
    plain = [pli.run(om, sdb, scores=False)["stage"] for om in oms]
    t_plain = time.perf_counter() - t0
    log(f"batched leg: plain per-query pass {t_plain:.3f} s")

    same = True
    res = {}
    for graph in (False, True):                      # direct launches on eight streams; captured graph
        pb = p7gpu.Pipeline()
        t0 = time.perf_counter()
        pb.run_batch(oms, sdb, scores=False, graph=graph)   # first call (captures the graph)
        t_first = time.perf_counter() - t0
        t0 = time.perf_counter()
        for _ in range(reps):
            outs, sts = pb.run_batch(oms, sdb, scores=False, graph=graph)
        res[graph] = ((time.perf_counter() - t0) / reps, t_first, pb.batch_graphed)
        for j in range(nprof):
            same &= bool(np.array_equal(outs[j]["stage"], plain[j]))
        log(f"batched leg: graph={graph} {res[graph][0]:.3f} s per batch")
    shutil.rmtree(tmp, ignore_errors=True)
    t_batch = res[False][0]
    return {"profiles": nprof, "M_min": min(Ms), "M_max": max(Ms), "M_mean": round(float(np.mean(Ms)), 1),
            "nseq": db.nseq, "nres": db.nres, "long_tail": "2% of sequences with L in [5000, 35000]",
            "cells": cells, "tune_s": round(t_tune, 3), "tuned_shapes_lanes_x_grab": shapes,
            "ms_batch": round(t_batch * 1000.0, 2), "ms_batch_graph": round(res[True][0] * 1000.0, 2),
            "graph_capture_and_first_run_s": round(res[True][1], 3), "graphed": res[True][2],
            "ms_plain_per_query_total": round(t_plain * 1000.0, 2),
            "gcups_batched": round(cells / t_batch / 1e9, 1),
            "gcups_batched_graph": round(cells / res[True][0] / 1e9, 1),
            "gcups_plain_per_query": round(cells / t_plain / 1e9, 1),
            "identical_pass_sets": same}


def domain_leg(p7gpu, om, sdb, steps, ms_plain):
    """The same pipeline with domain definition of the F3 survivors on
    (posterior decoding of the special states, the region heuristics, the
    regions' envelope scores, multidomain regions' sampled traces and
    clusters, and every envelope's null2 correction and alignment): their
    device time, the whole step, what they found, and the host time of the
    hit-list scores built from them."""
    pli = p7gpu.Pipeline(domains=True)
    pli.run(om, sdb, scores=False)
    st = pli.stats()                                 # counters of one call
    p7gpu.lib.p7_gpu_Synchronize()
    ms_dom, ms_n2, ms_md, ms_fwd, ms_bck, t0 = [], [], [], [], [], time.perf_counter()
    for _ in range(steps):
        pli.run(om, sdb, scores=False)
        sx = pli.stats()
        ms_dom.append(sx["ms_domains"]); ms_n2.append(sx["ms_null2"]); ms_md.append(sx["ms_multidomain"])
        ms_fwd.append(sx["ms_fwd"]); ms_bck.append(sx["ms_bck"])
    p7gpu.lib.p7_gpu_Synchronize()
    ms_step = (time.perf_counter() - t0) * 1000.0 / steps
    t1 = time.perf_counter()
    ts = pli.target_scores()
    ms_hits = (time.perf_counter() - t1) * 1000.0
    _, n2st = pli.region_null2()
    return {"steps": steps, "ms_domains": round(float(np.mean(ms_dom)), 4), "ms_null2": round(float(np.mean(ms_n2)), 4),
            "ms_multidomain": round(float(np.mean(ms_md)), 4),
            "ms_per_step": round(ms_step, 4), "ms_fwd": round(float(np.mean(ms_fwd)), 4),
            "ms_bck": round(float(np.mean(ms_bck)), 4), "ms_per_step_without": round(ms_plain, 4),
            "targets_decoded": st["n_past_fwd"], "regions": st["n_regions"], "multidomain_regions": st["n_clustered"],
            "envelopes_rescored": int(np.sum(n2st == 0)), "targets_scored": int(np.sum(ts["status"] == 0)),
            "ms_hit_scores_host": round(ms_hits, 3)}


def precision_leg(p7gpu, synth, Ms=(2000, 3000, 4000, 5000), nseq=2000, nsub=100, seed=91):
    """Precision stress: profiles of M = 2000..5000 (calibrated by this library)
    against sequences of L = 2k..35k with 20% indel-rich homologs. The GPU
    pipeline runs over every sequence; on a length-stratified subset the
    oracle's pipeline (its fp32 restatement of the reference) runs too, and
    for the twelve shortest sequences the FP64 log-space Forward. Reports per
    model: identical pass sets on the subset, the max and mean
    |Forward - oracle| and |Backward - oracle| in nats, and the max and mean
    |Forward - FP64| in nats."""
    import tempfile
    import shutil
    import oracle
    tmp = tempfile.mkdtemp(prefix="p7p_")
    rng = np.random.default_rng(seed)
    out = []
    for M in Ms:
        h = synth.random_profile(M, rng, f"prec_M{M}")
        path = os.path.join(tmp, h.name + ".hmm")
        synth.write_hmm(path, [h])
        h.stats = synth.calibrate_with(p7gpu, path, seed=seed)
        synth.write_hmm(path, [h])
        hh = synth.read_hmm_probs(path)
        om = p7gpu.OProfile.from_file(path)
        t = om.tables()
        db = synth.make_db(nseq, [hh], frac_homolog=0.2, seed_db=seed + M, seed_plant=seed + M + 1,
                           median=8000.0, sigma=0.8, lo=2000, hi=35000, indel_rich=True)
        sdb = p7gpu.SeqDB(db.residues, db.offsets, db.lengths)
        pli = p7gpu.Pipeline()
        t0 = time.perf_counter()
        got = pli.run(om, sdb)
        t_gpu = time.perf_counter() - t0
        order = np.argsort(db.lengths)
        sub_idx = np.sort(order[np.linspace(0, nseq - 1, nsub).astype(int)])
        sub = db.subset(sub_idx)
        prof, keep = oracle.make_profile(t, oracle.bias_hmm(t), t["evparam"], Lmax=int(db.lengths.max()) + 1)
        t0 = time.perf_counter()
        ref = oracle.pipeline_db(prof, sub.residues, sub.offsets, sub.lengths, nthreads=min(16, os.cpu_count() or 1))
        t_orc = time.perf_counter() - t0
        g = {k: v[sub_idx] for k, v in got.items()}
        same = bool(np.array_equal(g["stage"], ref["stage"]))
        errs = {}
        for k in ("fwdsc", "bcksc"):
            m = ~np.isnan(ref[k]) & ~np.isnan(g[k])
            d = np.abs(g[k][m].astype(np.float64) - ref[k][m])
            errs[k] = (float(d.max()) if d.size else None, float(d.mean()) if d.size else None, int(m.sum()))
        # the FP64 log-space Forward of the twelve shortest sequences against the
        # device Forward of the same sequences (each length-configured as in the pipeline)
        short = order[:12]
        fsc, _ = p7gpu.forward_batch(om, sdb, short.astype(np.int64))
        e64 = []
        for j, i in enumerate(short):
            om.gm.reconfig_length(int(db.lengths[i]))
            tsc, rsc, xsc = om.gm.scores()
            e64.append(abs(float(fsc[j]) - oracle.forward_log64(db.seq(int(i)), tsc, rsc, xsc)))
        st = pli.stats()
        out.append({"M": M, "nseq": nseq, "nres": int(db.nres), "subset": nsub, "gpu_s": round(t_gpu, 3),
                    "oracle_subset_s": round(t_orc, 1), "pass_sets_identical_on_subset": same,
                    "n_past_fwd": st["n_past_fwd"],
                    "fwd_vs_oracle_fp32_max": errs["fwdsc"][0], "fwd_vs_oracle_fp32_mean": errs["fwdsc"][1],
                    "fwd_compared": errs["fwdsc"][2],
                    "bck_vs_oracle_fp32_max": errs["bcksc"][0], "bck_vs_oracle_fp32_mean": errs["bcksc"][1],
                    "bck_compared": errs["bcksc"][2],
                    "fwd_vs_fp64_max": max(e64) if e64 else None,
                    "fwd_vs_fp64_mean": float(np.mean(e64)) if e64 else None, "fp64_compared": len(e64)})
        log(f"precision: {out[-1]}")
    shutil.rmtree(tmp, ignore_errors=True)
    return out


def calibrate_copy(mib=256, reps=3):
    """Plain device-to-device copies of a known size: what FETCH_SIZE and WRITE_SIZE
    report for them gives the factor from counter KiB to HBM bytes."""
    import torch
    a = torch.ones(mib * 1024 * 1024 // 4, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    for _ in range(reps):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes_read": mib * 1024 * 1024, "copy_bytes_written": mib * 1024 * 1024, "reps": reps}))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--nseq", type=int, default=2_000_000)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--traffic", default=os.path.join(ROOT, "profiles", "k_ssv_traffic.json"),
                    help="JSON with the HBM bytes of one k_ssv launch, from rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE")
    ap.add_argument("--serve-queries", type=int, default=100, help="queries in the serving leg (0: skip it)")
    ap.add_argument("--serve-rate", type=float, default=20.0, help="Poisson arrival rate of the serving leg, queries/s")
    ap.add_argument("--serve-nseq", type=int, default=10_000_000, help="sequences of the serving leg's resident database")
    ap.add_argument("--serve-slots", type=int, default=4)
    ap.add_argument("--domain-steps", type=int, default=5, help="steps of the pipeline with domain decoding (0: skip)")
    ap.add_argument("--batch-profiles", type=int, default=64, help="profiles in the batched leg (0: skip it)")
    ap.add_argument("--batch-nseq", type=int, default=500_000, help="database sequences of the batched leg")
    ap.add_argument("--no-tune", action="store_true", help="skip the tuner (keep the built-in ungapped-pass shape)")
    ap.add_argument("--only-batched", action="store_true",
                    help="only run the batched leg and print its JSON line (for a kernel trace of that leg alone)")
    ap.add_argument("--only-serving", action="store_true", help="only run the serving leg and print its JSON line")
    ap.add_argument("--precision", action="store_true",
                    help="only run the precision stress (M = 2000..5000, L = 2k..35k) and print its JSON line")
    ap.add_argument("--hw-queues", type=int, default=MAX_HW_QUEUES,
                    help="GPU_MAX_HW_QUEUES for this process: hardware queues the HIP streams map onto")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="only run three plain 256 MiB device copies (to scale the FETCH_SIZE/WRITE_SIZE counters)")
    args = ap.parse_args()
    if not 1 <= args.hw_queues <= MAX_HW_QUEUES:
        ap.error(f"--hw-queues must be in 1..{MAX_HW_QUEUES}")
    os.environ["GPU_MAX_HW_QUEUES"] = str(args.hw_queues)
    if args.calibrate_copy:
        calibrate_copy()
        return
    if args.precision:
        from biobench2_amd import p7gpu, synth
        p7gpu.init(0)
        print(json.dumps({"precision_stress": precision_leg(p7gpu, synth)}), flush=True)
        return
    if args.only_batched:
        from biobench2_amd import p7gpu, synth
        p7gpu.init(0)
        print(json.dumps({"batched": batched_leg(p7gpu, synth, args.batch_profiles, args.batch_nseq,
                                                 tune=not args.no_tune),
                          "hw_queues": os.environ.get("GPU_MAX_HW_QUEUES")}), flush=True)
        return
    if args.only_serving:
        from biobench2_amd import p7gpu, synth
        p7gpu.init(0)
        print(json.dumps({"serving": serving_leg(p7gpu, synth, args.serve_nseq, args.serve_queries, args.serve_rate,
                                                 args.serve_slots, tune=not args.no_tune),
                          "hw_queues": os.environ.get("GPU_MAX_HW_QUEUES")}), flush=True)
        return
    if args.gpus != 1:
        print(f"bench.py runs on one GPU: --gpus {args.gpus} is not supported by this build "
              "(the database would shard by sequence ranges across GPUs; only the one-GPU path is built)",
              file=sys.stderr)
        sys.exit(2)

    from biobench2_amd import p7gpu, synth
    p7gpu.init(0)
    om = p7gpu.OProfile.from_file(PROFILE)
    M = om.tables()["M"]
    hmm = synth.read_hmm_probs(PROFILE)
    t0 = time.time()
    db = synth.make_db(args.nseq, [hmm], frac_homolog=0.01, seed_db=43, seed_plant=44)
    log(f"db: {db.nseq} sequences, {db.nres} residues, generated in {time.time() - t0:.1f} s")
    sdb = p7gpu.SeqDB(db.residues, db.offsets, db.lengths)
    pli = p7gpu.Pipeline()
    if not args.no_tune:
        t0 = time.perf_counter()
        p7gpu.autotune([om], sdb)                 # the ungapped pass's shape for M = 400 on this device
        lanes, grab, us, _ = p7gpu.tune_get(M)
        log(f"tuner: {lanes} lanes per target, grab {grab} ({time.perf_counter() - t0:.2f} s; us per shape {us.tolist()})")
    lanes, grab, _, _ = p7gpu.tune_get(M)

    for w in range(args.warmup):
        pli.run(om, sdb, scores=False)
        log(f"warmup {w + 1}/{args.warmup} done")
    p7gpu.lib.p7_gpu_Synchronize()
    ms_ssv, ms_stage = [], {k: [] for k in ("ms_msv", "ms_bias", "ms_vit", "ms_fwd", "ms_bck", "ms_total")}
    t0 = time.perf_counter()
    for k in range(args.steps):
        out = pli.run(om, sdb, scores=False)
        st = pli.stats()
        ms_ssv.append(st["ms_ssv"])
        for key in ms_stage:
            ms_stage[key].append(st[key])
    p7gpu.lib.p7_gpu_Synchronize()
    t1 = time.perf_counter()
    log(f"timed {args.steps} steps")

    ms_per_step = (t1 - t0) * 1000.0 / args.steps
    cells = float(db.nres) * M
    value = cells / (ms_per_step / 1000.0) / 1e9
    stage = out["stage"]
    ssv_ms = float(np.mean(ms_ssv))
    ssv_tops = cells * SSV_OPS_PER_CELL / (ssv_ms / 1000.0) / 1e12
    per_stage = {}
    for key, cname in (("ms_msv", "cells_msv"), ("ms_bias", "cells_bias"), ("ms_vit", "cells_vit"),
                       ("ms_fwd", "cells_fwd"), ("ms_bck", "cells_bck")):
        ms = float(np.mean(ms_stage[key]))
        per_stage[key[3:]] = {"ms": round(ms, 4), "cells": st[cname],
                              "gcups": round(st[cname] / (ms / 1000.0) / 1e9, 2) if ms > 0 else None}
    traffic = None
    if args.traffic and os.path.exists(args.traffic):
        traffic = json.load(open(args.traffic)).get("bytes_per_launch")
    line = {
        "metric": "filter GCUPS (MSV/Vit/Fwd) at M=400 vs CPU; served-query p50/p99 latency",
        "value": round(value, 2),
        "unit": "GCUPS",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u8/i16/f32",
        "data": "synthetic (seeded profile generator M=400; lognormal-length i.i.d. database, 1% planted homologs)",
        "config": {"workload": "hmmsearch filter pipeline, one query M=400 vs 2M-sequence database",
                   "nseq": db.nseq, "nres": db.nres, "M": M, "F1": 0.02, "F2": 1e-3, "F3": 1e-5},
        "roofline": {"bound": "valu", "kernel": f"k_ssv (ungapped MSV pass, packed int16, {lanes} lanes per target)",
                     "achieved": round(ssv_tops, 3), "peak": PEAK_PK16_TOPS, "unit": "TOP/s",
                     "frac": round(ssv_tops / PEAK_PK16_TOPS, 4), "traffic": traffic,
                     "ms_per_launch": round(ssv_ms, 4), "ops_per_cell": SSV_OPS_PER_CELL,
                     "residue_stream_GBs": round(db.nres / (ssv_ms / 1000.0) / 1e9, 2)},
        "stages": per_stage,
        "hw_queues": os.environ.get("GPU_MAX_HW_QUEUES"),
        "pass_counts": {"msv": int((stage >= 1).sum()), "bias": int((stage >= 2).sum()),
                        "vit": int((stage >= 3).sum()), "fwd": int((stage >= 4).sum()),
                        "msv_full_rescored": st["n_msv_full"]},
    }
    if args.domain_steps > 0:
        line["domain_decoding"] = domain_leg(p7gpu, om, sdb, args.domain_steps, ms_per_step)
    if args.batch_profiles > 0:
        log("batched leg ...")
        line["batched"] = batched_leg(p7gpu, synth, args.batch_profiles, args.batch_nseq, tune=not args.no_tune)
    if args.serve_queries > 0:
        log("serving leg ...")
        line["serving"] = serving_leg(p7gpu, synth, args.serve_nseq, args.serve_queries, args.serve_rate,
                                      args.serve_slots, tune=not args.no_tune)
    if not args.no_cpu:
        log("cpu baseline ...")
        chk = pli.run(om, sdb, scores=True)          # outside the timed region: the scores to compare
        line["cpu_baseline"] = cpu_baseline(om, db, M, chk["stage"], chk["usc"], args.cpu_seconds)
        line["oracle_mismatches"] = line["cpu_baseline"]["oracle_mismatches"]
    else:
        line["cpu_baseline"] = None
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
