"""Benchmark: single-end mapping throughput of the HIP path on one MI355X, plain and
with the learned candidate prefilter.

Workloads (probebwa_amd/synth.py):
- uneven_se_100mbp (default): a 100 Mbp synthetic genome over 10 contigs
  (uniform background, 5% segmental duplications at 0.5-3% divergence, 5-15 bp
  homopolymer runs every ~2 kb) and 2,000,000 single-end reads with lengths
  U{80..150}, Phred qualities decaying 38 -> 22 along the read (+ N(0,3) noise),
  substitutions at the Phred-implied rate.  The prefilter trains on 131,072
  other reads of the same generator.
- prefilter_se_2mbp: the 2.1 Mbp probe genome (2 Mbp uniform + 50 copies of a
  2 kb unit at 1% divergence); the prefilter trains on 500,000 reads and is
  evaluated on 2,000,000 held-out reads (L ~ U{80..150}, 1% substitutions,
  0.05% indels, Q ~ U{20..40}).

A step maps one batch of `--batch` reads that are already resident in HBM
(`pbw_reads_upload` once, then `pbw_map_se_device` per step; calls return
after the records are in host memory): seed -> cluster -> banded DP -> MAPQ.
`value` = reads mapped per second over the K timed steps of the plain path.
The same K steps are then timed with the trained prefilter on (`learned`),
and every record of them is compared with the plain path's.

Also reported: the dominant kernel's roofline position (HIP-event timed inside
the library on its own stream), and the CPU oracle's reads/s on a bounded
sample of the same reads on the host cores, with the sample's SAM-field
concordance between the HIP path and the oracle.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload NAME]
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_HBM_GBPS = 8000.0
# integer VALU: 256 CUs x 4 SIMD x 16 lanes x 2.4 GHz (no packed int32)
PEAK_INT32_TOPS = 256 * 64 * 2.4e9 / 1e12
DP_OPS_PER_CELL = 20  # VALU ops of one cell of the reference recurrence (smith_waterman.rs:204-233)


def log(msg):
    print("[bench] " + msg, file=sys.stderr, flush=True)


TRAFFIC_PROFILE = os.path.join(ROOT, "profiles", "pmc_traffic_uneven_se_100mbp.txt")
TRAFFIC_BATCH = 262144  # reads per launch the profile was taken at


VALU_PROFILE = os.path.join(ROOT, "profiles", "pmc_valu_uneven_se_100mbp.txt")


PROFILED_WORKLOAD = "uneven_se_100mbp"  # the workload the committed counter summaries were taken on


def measured_valu(kernel, batch, workload=PROFILED_WORKLOAD):
    """VALU issue counters of `kernel` from the committed rocprofv3 summary (taken at the
    default batch): fraction of the VALU issue capacity used, int32 share of it, active lanes."""
    if workload != PROFILED_WORKLOAD or batch != TRAFFIC_BATCH or not os.path.exists(VALU_PROFILE):
        return None
    for line in open(VALU_PROFILE):
        f = line.strip().split(",")
        if len(f) == 10 and f[0] == kernel:
            return {"valu_busy": float(f[7]), "int32_util": float(f[8]), "lanes_active": float(f[9]),
                    "valu_insts": int(f[1]), "int32_insts": int(f[2])}
    return None


def measured_traffic(kernel, batch, workload=PROFILED_WORKLOAD):
    """HBM bytes per launch of `kernel` (FETCH_SIZE + WRITE_SIZE, copy-calibrated) from the
    committed rocprofv3 counter summary, when it was taken on this workload and batch size; else None."""
    if workload != PROFILED_WORKLOAD or batch != TRAFFIC_BATCH or not os.path.exists(TRAFFIC_PROFILE):
        return None
    for line in open(TRAFFIC_PROFILE):
        f = line.strip().split(",")
        if len(f) == 7 and f[0] == kernel:
            return int(float(f[5]))
    return None


def bench_pe(args):
    """Paired-end joint posterior on the probe genome: 1M FR pairs of 2x150 bp (10% with one
    mate inside a repeat copy), one pbw_map_pe call over all of them (500 serial warm-up
    pairs, then 2000-pair batches on a frozen insert-size snapshot, as the reference)."""
    import probebwa_amd as P
    from probebwa_amd import synth
    n = args.reads or 1_000_000
    t0 = time.time()
    contigs = synth.genome_probe_se_2mbp()
    spans = synth.probe_repeat_spans()
    g = P.Genome()
    for name, s in contigs:
        g.add_contig(name, s.tobytes())
    m1, m2 = synth.reads_pairs_probe(contigs, n=n, seed=3, repeat_spans=spans)
    r1, r2 = P.ReadBatch.from_arrays(*m1), P.ReadBatch.from_arrays(*m2)
    log("pairs %.1f s (%d pairs)" % (time.time() - t0, n))
    ctx = P.Context(g, device=0)
    w = min(n, 10000)
    ctx.map_pe(r1.subset(0, w), r2.subset(0, w))
    t1 = time.perf_counter()
    recs = ctx.map_pe(r1, r2)
    dt = time.perf_counter() - t1
    st = ctx.stats()
    log("mapped %d pairs in %.2f s" % (n, dt))
    cpu = None
    concord = None
    if not args.no_cpu:
        from oracle import pbw_oracle as O
        S = min(n, 20000)
        go = O.Genome()
        for name, s in contigs:
            go.add_contig(name, s.tobytes())
        m = O.Mapper(go, O.HashTable(go), threads=args.cpu_threads)
        a = O.ReadBatch.from_arrays(*[x for x in m1[:4]]).subset(0, S)
        b = O.ReadBatch.from_arrays(*[x for x in m2[:4]]).subset(0, S)
        t2 = time.perf_counter()
        ro = m.map_pe(a, b)
        cdt = time.perf_counter() - t2
        cpu = {"value": round(S / cdt, 1), "unit": "pairs/s", "cores": args.cpu_threads, "kind": "port",
               "sample": "first %d pairs of the workload, oracle (C++ restatement) on %d threads, %.1f s"
                         % (S, args.cpu_threads, cdt)}
        concord = float(P.primary_identical(ro, P.Records(recs.rec[:2 * S], recs.cigar)).mean())
    out = {
        "metric": "pairs/s on 1 GPU, paired-end joint posterior",
        "value": round(n / dt, 1), "unit": "pairs/s", "n_gpus": 1, "steps": 1, "warmup": 1,
        "ms_per_step": round(dt * 1000.0, 3), "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "int32+f64", "data": "synthetic (probe genome seed 1, pairs seed 3)",
        "config": {"workload": "probe_pe_2mbp: 2.1 Mbp probe genome, %d FR pairs of 2x150 bp, 10%% anchored in "
                               "repeat copies" % n, "pair_batch": 2000, "serial_warmup_pairs": 500},
        "reads_per_s": round(2 * n / dt, 1),
        "dp_cells": int(st["dp_cells"]), "wide_jobs": int(st["wide_jobs"]), "wide_cells": int(st["wide_cells"]),
        "stage_ms": {k: round(st[k], 1) for k in ("ms_seed", "ms_cluster", "ms_dp", "ms_pair", "ms_total")},
        "cpu_baseline": cpu, "sample_sam_concordance": concord,
    }
    print(json.dumps(out), flush=True)


WORKLOADS = {
    "uneven_se_100mbp": {"reads": 2_000_000, "train_reads": 131072,
                         "desc": "single-end, 100 Mbp / 10 contigs, 2M reads of 80-150 bp"},
    "prefilter_se_2mbp": {"reads": 2_000_000, "train_reads": 500_000,
                          "desc": "single-end, 2.1 Mbp probe genome, 2M held-out reads of 80-150 bp"},
    "genome_se_3gbp": {"reads": 10_000_000, "train_reads": 131072,
                       "desc": "single-end, 3.1 Gbp / 24 contigs (45% interspersed repeats, 5% segmental "
                               "duplications), 10M reads of 60-250 bp"},
}


def make_workload(name, n, n_train):
    from probebwa_amd import synth
    if name == "uneven_se_100mbp":
        contigs = synth.genome_uneven_100mbp()
        return (contigs, lambda: synth.reads_uneven_se(contigs, n=n, seed=12),
                lambda: synth.reads_uneven_se(contigs, n=n_train, seed=99), "seeds 2/12/99")
    if name == "genome_se_3gbp":
        contigs = synth.genome_scale_3gbp()
        return (contigs, lambda: synth.reads_uneven_se(contigs, n=n, seed=55, lmin=60, lmax=250),
                lambda: synth.reads_uneven_se(contigs, n=n_train, seed=59, lmin=60, lmax=250), "seeds 5/55/59")
    contigs = synth.genome_probe_se_2mbp()
    return (contigs, lambda: synth.reads_probe_uneven_se(contigs, n=n, seed=42),
            lambda: synth.reads_probe_uneven_se(contigs, n=n_train, seed=41), "seeds 1/42/41")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=8)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--batch", type=int, default=262144)
    ap.add_argument("--workload", choices=sorted(WORKLOADS) + ["probe_pe_2mbp"], default="uneven_se_100mbp")
    ap.add_argument("--reads", type=int, default=None)
    ap.add_argument("--cpu-sample", type=int, default=131072)
    ap.add_argument("--cpu-threads", type=int, default=min(16, os.cpu_count() or 1))
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-learned", action="store_true")
    ap.add_argument("--autotune", action="store_true",
                    help="tune launch shapes on the first 524288 reads before the timed steps")
    ap.add_argument("--train-reads", type=int, default=None)
    ap.add_argument("--label", choices=["loglk", "score"], default="loglk")
    ap.add_argument("--margin", type=float, default=40.0)
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only copy 1 GiB device to device three times (the known-size run that scales "
                         "rocprofv3's FETCH_SIZE / WRITE_SIZE counters to bytes)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU only: this build has no multi-GPU path (--gpus %d requested)" % args.gpus,
              file=sys.stderr)
        sys.exit(2)
    if args.copy_calibration:
        import torch
        a = torch.ones(1 << 28, dtype=torch.float32, device="cuda")  # 1 GiB
        b = torch.empty_like(a)
        torch.cuda.synchronize()
        for _ in range(3):
            b.copy_(a)
        torch.cuda.synchronize()
        print("copied %d bytes x3" % (4 << 28), flush=True)
        return

    if args.workload == "probe_pe_2mbp":
        bench_pe(args)
        return
    import probebwa_amd as P

    wl = WORKLOADS[args.workload]
    n_reads = args.reads or wl["reads"]
    n_train = args.train_reads or wl["train_reads"]
    t0 = time.time()
    contigs, gen_reads, gen_train, seeds = make_workload(args.workload, n_reads, n_train)
    gp = P.Genome()
    for name, s in contigs:
        gp.add_contig(name, s.tobytes())
    log("genome %.1f s" % (time.time() - t0))
    t0 = time.time()
    seq, so, qual, qo = gen_reads()
    log("reads %.1f s (%d reads, %d bases)" % (time.time() - t0, n_reads, int(so[-1])))
    t0 = time.time()
    ctx = P.Context(gp, batch_reads=args.batch, device=0)
    log("context (index build + upload) %.1f s, %d index entries" % (time.time() - t0, ctx.total_entries))

    dreads = ctx.upload_reads(P.ReadBatch.from_arrays(seq, so, qual, qo))
    n = n_reads
    B = min(args.batch, n)

    tuned = None
    if args.autotune:
        t1 = time.time()
        rep = ctx.autotune(dreads, 0, min(n, 524288))
        tuned = {"best": rep["best"], "best_ms": round(rep["best_ms"], 2), "trials": len(rep["trials"]),
                 "default_ms": round(rep["trials"][0]["ms"], 2), "seconds": round(time.time() - t1, 1)}
        B = min(ctx.tune.batch_reads, n) if args.batch == 262144 else B
        log("autotune %.1f s: %s" % (time.time() - t1, rep["best"]))

    def step(s):
        lo = (s * B) % n
        hi = min(n, lo + B)
        return ctx.map_se_device(dreads, lo, hi), ctx.stats()

    for w in range(args.warmup):
        step(w)
        log("warmup step %d done" % w)
    agg = {}
    plain_recs = []
    t0 = time.perf_counter()
    for k in range(args.steps):
        recs, st = step(k)
        plain_recs.append(recs)
        for key, v in st.items():
            agg[key] = agg.get(key, 0) + v
        log("step %d: %d reads in %.1f ms" % (k, st["reads"], st["ms_total"]))
    dt = time.perf_counter() - t0
    reads_done = int(agg["reads"])
    value = reads_done / dt

    # dominant kernel + roofline (HIP events on the library's stream)
    ms_dp = agg["ms_k_dp"] / max(agg["n_k_dp"], 1)
    ms_seed = agg["ms_k_seed"] / max(agg["n_k_seed"], 1)
    cells_per_launch = agg["dp_cells"] / max(agg["n_k_dp"], 1)
    gather_bytes = 8 * (agg["lookups"] + agg["positions"] + agg["filtered_hits"])
    gbytes_per_launch = gather_bytes / max(agg["n_k_seed"], 1)
    dp_gcups = cells_per_launch / (ms_dp * 1e-3) / 1e9 if ms_dp > 0 else 0.0
    gather_gbps = gbytes_per_launch / (ms_seed * 1e-3) / 1e9 if ms_seed > 0 else 0.0
    if agg["ms_k_dp"] >= agg["ms_k_seed"]:
        ach = dp_gcups * DP_OPS_PER_CELL / 1000.0
        roof = {"kernel": "k_dp_quad", "bound": "valu-int32", "achieved": round(ach, 3), "peak": round(PEAK_INT32_TOPS, 2),
                "unit": "Tops/s", "frac": round(ach / PEAK_INT32_TOPS, 4), "traffic": measured_traffic("pbw::k_dp_quad", B, args.workload),
                "per_launch": {"cells": cells_per_launch, "ms": ms_dp, "ops_per_cell": DP_OPS_PER_CELL},
                "counters": measured_valu("pbw::k_dp_quad", B, args.workload)}
    else:
        roof = {"kernel": "k_seed", "bound": "hbm", "achieved": round(gather_gbps, 2), "peak": PEAK_HBM_GBPS,
                "unit": "GB/s", "frac": round(gather_gbps / PEAK_HBM_GBPS, 4), "traffic": measured_traffic("pbw::k_seed", B, args.workload),
                "per_launch": {"bytes": gbytes_per_launch, "ms": ms_seed}}
    gather_roof = {"kernel": "k_seed", "bound": "hbm", "achieved": round(gather_gbps, 2), "peak": PEAK_HBM_GBPS,
                   "unit": "GB/s", "frac": round(gather_gbps / PEAK_HBM_GBPS, 4),
                   "traffic": measured_traffic("pbw::k_seed", B, args.workload)}

    learned = None
    if not args.no_learned:
        # learned candidate prefilter: trained on a separate read set of the same genome (labels from
        # the plain DP over every candidate), then the same steps timed with it on
        t1 = time.time()
        tseq, tso, tqual, tqo = gen_train()
        rep = ctx.prefilter_train(P.ReadBatch.from_arrays(tseq, tso, tqual, tqo), label=args.label, margin=args.margin,
                                  recall=0.9999, epochs=20, batch=4096, lr=1e-3)
        del tseq, tqual
        log("prefilter trained in %.1f s on %d rows (%d positive), held-out skip %.3f recall %.5f"
            % (time.time() - t1, rep["rows"], rep["positives"], rep["val_skip"], rep["val_recall"]))
        ctx.set_prefilter(enabled=1, audit_period=64)
        for w in range(args.warmup):
            step(w)
        agg2 = {}
        learned_recs = []
        t1 = time.perf_counter()
        for k in range(args.steps):
            recs, st = step(k)
            learned_recs.append(recs)
            for key, v in st.items():
                agg2[key] = agg2.get(key, 0) + v
            log("learned step %d: %d reads in %.1f ms" % (k, st["reads"], st["ms_total"]))
        dt2 = time.perf_counter() - t1
        ctx.set_prefilter(enabled=0)
        same = np.concatenate([P.primary_identical(a, b) for a, b in zip(plain_recs, learned_recs)])
        del learned_recs
        label = ("log-likelihood within %g nats of the read's best" if args.label == "loglk"
                 else "DP score within %g units of the read's best") % args.margin
        learned = {
            "reads_per_s": round(int(agg2["reads"]) / dt2, 1),
            "ms_per_step": round(dt2 * 1000.0 / args.steps, 3),
            "speedup_vs_plain": round((int(agg2["reads"]) / dt2) / value, 3),
            "sam_identical": round(float(same.mean()), 6),
            "sam_differing_reads": int((~same).sum()),
            "reads_compared": int(len(same)),
            "audit_reads": int(agg2["pf_audit_reads"]),
            "audit_mismatches": int(agg2["pf_audit_mismatches"]),
            "fallback_reads": int(agg2["pf_fallback_reads"]),
            "fallback_rate": round(agg2["pf_fallback_reads"] / max(1, agg2["reads"]), 5),
            "dp_jobs_aligned_frac": round(agg2["pf_jobs_aligned"] / max(1, agg2["pf_jobs"]), 5),
            "dp_cells": int(agg2["dp_cells"]),
            "dp_cells_skipped_frac": round(1.0 - agg2["dp_cells"] / max(1, agg["dp_cells"]), 5),
            "stage_ms": {k: round(agg2[k], 1) for k in ("ms_seed", "ms_cluster", "ms_pf", "ms_dp", "ms_mapq",
                                                         "ms_total")},
            "training": {"reads": n_train, "rows": int(rep["rows"]), "positives": int(rep["positives"]),
                         "label": label, "epochs": 20, "batch": 4096, "lr": 1e-3,
                         "threshold": round(float(rep["threshold"]), 5), "heldout_recall": round(rep["val_recall"], 6),
                         "heldout_skip": round(rep["val_skip"], 5), "loss_first_epoch": round(rep["loss_first"], 6),
                         "loss_last_epoch": round(rep["loss_last"], 6)},
        }

    cpu = None
    concord = None
    if not args.no_cpu:
        from oracle import pbw_oracle as O

        S = min(args.cpu_sample, B)
        t1 = time.time()
        go = O.Genome()
        for name, s in contigs:
            go.add_contig(name, s.tobytes())
        ht = O.HashTable(go, progress=True)
        m = O.Mapper(go, ht, threads=args.cpu_threads)
        log("oracle index %.1f s" % (time.time() - t1))
        rb = O.ReadBatch.from_arrays(seq[:int(so[S])], so[:S + 1], qual[:int(qo[S])], qo[:S + 1])
        t1 = time.perf_counter()
        ro = m.map_se(rb)
        cdt = time.perf_counter() - t1
        cpu = {"value": round(S / cdt, 1), "unit": "reads/s", "cores": args.cpu_threads, "kind": "port",
               "sample": "first %d reads of the workload, oracle (C++ restatement) on %d threads, %.1f s"
                         % (S, args.cpu_threads, cdt)}
        first = plain_recs[0]
        concord = float(P.primary_identical(ro, P.Records(first.rec[:S], first.cigar)).mean())

    out = {
        "metric": "reads/s on 1 GPU, learned vs plain @≥99.9% SAM concordance; DP GCUPS",
        "value": round(value, 1),
        "unit": "reads/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(dt * 1000.0 / args.steps, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int32+f64",
        "data": "synthetic (%s generator, %s)" % (args.workload, seeds),
        "config": {"workload": "%s: plain GPU path, %s" % (args.workload, wl["desc"]), "batch_reads": B,
                   "reads_per_step": B, "path": "plain (no learned prefilter)"},
        "dp_gcups": round(dp_gcups, 2),
        "gather_gbps": round(gather_gbps, 2),
        "roofline": roof,
        "roofline_gather": gather_roof,
        "autotune": tuned,
        "learned": learned,
        "learned_vs_plain": learned["speedup_vs_plain"] if learned else None,
        "cpu_baseline": cpu,
        "sample_sam_concordance": concord,
        "stage_ms": {k: round(agg[k], 1) for k in ("ms_seed", "ms_cluster", "ms_dp", "ms_mapq", "ms_total")},
        "counters": {k: int(agg[k]) for k in ("reads", "lookups", "positions", "filtered_hits", "dp_cells",
                                               "seed_overflow_reads", "tier2_cluster_reads", "band_dp_jobs",
                                               "generic_dp_jobs")},
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
