"""Benchmark: single-end mapping throughput of the HIP path on one MI355X, plain and
with the learned candidate prefilter.

Workloads (probebwa_amd/synth.py):
- genome_se_3gbp (default): a 3.1 Gbp synthetic genome over 24 contigs (45%
  interspersed repeats from a 300 bp family at 10-20% divergence and a 6 kb family at
  5-15%, 5% segmental duplications), single-end reads of 60-250 bp with Phred
  qualities decaying 38 -> 22 along the read (+ N(0,3) noise) and substitutions at
  the Phred-implied rate; plus a paired-end leg of 65,536 FR pairs of 2 x 150 on the
  same genome.  The prefilter trains on 131,072 other reads of the same generator.
- uneven_se_100mbp: a 100 Mbp genome over 10 contigs (uniform background, 5%
  segmental duplications at 0.5-3% divergence, 5-15 bp homopolymer runs every
  ~2 kb), 2,000,000 reads of 80-150 bp with the same read model.
- prefilter_se_2mbp: the 2.1 Mbp probe genome (2 Mbp uniform + 50 copies of a 2 kb
  unit at 1% divergence); the prefilter trains on 500,000 reads and is evaluated on
  2,000,000 held-out reads (L ~ U{80..150}, 1% substitutions, 0.05% indels,
  Q ~ U{20..40}).
- probe_pe_2mbp: 1,000,000 FR pairs of 2 x 150 on the probe genome (pairs/s), with
  the k_dp_wide roofline.  The default genome run also maps a bounded 200,000 pairs of
  this workload (the oracle on the first 8,000) into its `probe_pe_2mbp` field.

A step maps one batch of `--batch` reads that are already resident in HBM
(`pbw_reads_upload` once, then `pbw_map_se_device` per step; calls return after the
records are in host memory): seed -> cluster -> banded DP -> MAPQ.  `value` = reads
mapped per second over the K timed steps of the plain path.  The reads' upload over
PCIe is timed on its own and reported beside it (`pcie_inclusive_reads_per_s`).
The same K steps are then timed with the trained prefilter on (`learned`), and every
record of them is compared with the plain path's.

Also reported: the roofline position of the kernel that takes the most time per step
(seed gather against HBM, clustering against HBM, or the banded DP against the
integer VALU), HIP-event timed inside the library on its own stream; and the CPU
oracle's reads/s on bounded samples of the same reads on the host cores (16 threads,
and 1 thread), with the sample's SAM-field concordance between the HIP path and the
oracle.

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
CAND_BYTES = 20 * 16 + 1  # a read's candidate list written by the clustering kernels
PROFILE_BATCH = 262144  # reads per launch the committed counter summaries were taken at
HOST_THREADS = min(16, os.cpu_count() or 1)  # the GPU host's CPU share is 16 threads


def log(msg):
    print("[bench] " + msg, file=sys.stderr, flush=True)


def _profile_rows(kind, workload, batch, ncols):
    path = os.path.join(ROOT, "profiles", "pmc_%s_%s.txt" % (kind, workload))
    if batch != PROFILE_BATCH or not os.path.exists(path):
        return {}
    rows = {}
    for line in open(path):
        f = line.strip().split(",")
        if len(f) == ncols and not f[0].startswith("#") and f[0] != "kernel":
            rows[f[0]] = f
    return rows


def measured_traffic(kernel, batch, workload):
    """HBM bytes per step of `kernel` (FETCH_SIZE + WRITE_SIZE, copy-calibrated) from the
    committed rocprofv3 counter summary of this workload at the default batch (summed over
    the kernel's dispatches in the timed step); else None.  `kernel` may be a tuple: the
    sum over those kernels."""
    rows = _profile_rows("traffic", workload, batch, 7)
    names = kernel if isinstance(kernel, tuple) else (kernel,)
    if not all(k in rows for k in names):
        return None
    return sum(int(float(rows[k][5])) for k in names)


def measured_valu(kernel, batch, workload):
    """VALU issue counters of `kernel` from the committed rocprofv3 summary (its dispatches in
    the timed step): fraction of the VALU issue capacity used, int32 share of it, active lanes."""
    f = _profile_rows("valu", workload, batch, 10).get(kernel)
    if not f:
        return None
    return {"valu_busy": float(f[7]), "int32_util": float(f[8]), "lanes_active": float(f[9]),
            "valu_insts": int(f[1]), "int32_insts": int(f[2])}


PE_PROFILE_PAIRS = 200_000  # the pair count profiles/pmc_traffic_probe_pe_2mbp.txt was taken at


def wide_traffic(pairs):
    """HBM bytes (FETCH_SIZE + WRITE_SIZE, copy-calibrated) of every k_dp_wide dispatch of the
    timed map_pe call, from the committed counter summary of probe_pe_2mbp at its profiled
    pair count; else None."""
    path = os.path.join(ROOT, "profiles", "pmc_traffic_probe_pe_2mbp.txt")
    if pairs != PE_PROFILE_PAIRS or not os.path.exists(path):
        return None
    tot = 0
    for line in open(path):
        f = line.strip().split(",")
        if len(f) == 8 and f[0].startswith("pbw::k_dp_wide"):
            tot += int(f[6])
    return tot or None


def wide_roofline(st, pairs=None):
    """The paired-end mate-window DP (k_dp_wide: rescue windows around the mate's placement,
    score rows, wave traceback and likelihood) against the int32 VALU peak: its cells x the
    reference recurrence's ops per cell over the summed duration of its launches, timed with
    HIP events on the stream it runs on (one span per pair batch, which also holds the tiled
    kernel's launch for the few windows k_dp_wide hands on: 1.5 of 49.6 ms in the counter
    run of the probe genome).  `traffic`: the measured HBM bytes of the call's
    k_dp_wide dispatches on the probe genome at the profiled pair count (mostly its traceback
    bits, 4 per cell)."""
    ms, cells = float(st["ms_k_wide"]), int(st["wide_cells"])
    ach = cells / (ms * 1e-3) * DP_OPS_PER_CELL / 1e12 if ms > 0 else 0.0
    return {"kernel": "k_dp_wide", "bound": "valu-int32", "unit": "Tops/s", "peak": round(PEAK_INT32_TOPS, 2),
            "achieved": round(ach, 3), "frac": round(ach / PEAK_INT32_TOPS, 4),
            "traffic": wide_traffic(pairs) if pairs is not None else None,
            "per_launch": {"cells": cells, "pair_batches": int(st["n_k_wide"]), "ms": round(ms, 3),
                           "ops_per_cell": DP_OPS_PER_CELL}}


def probe_pe_run(args, n, cpu_pairs):
    """Paired-end joint posterior on the probe genome: n FR pairs of 2x150 bp (10% with one
    mate inside a repeat copy), one pbw_map_pe call over all of them (500 serial warm-up
    pairs, then 2000-pair batches on a frozen insert-size snapshot, as the reference); the
    oracle on the first `cpu_pairs` pairs (0: none) for the CPU baseline and the SAM
    concordance.  Returns the result line's fields."""
    import probebwa_amd as P
    from probebwa_amd import synth
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
    if cpu_pairs > 0:
        from oracle import pbw_oracle as O
        S = min(n, cpu_pairs)
        go = O.Genome()
        for name, s in contigs:
            go.add_contig(name, s.tobytes())
        m = O.Mapper(go, O.HashTable(go, threads=args.cpu_threads), threads=args.cpu_threads)
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
        "roofline": wide_roofline(st, n), "cpu_baseline": cpu, "sample_sam_concordance": concord,
    }
    return out


def bench_pe(args):
    """The probe_pe_2mbp workload on its own: 1M pairs (or --reads), the oracle on 20,000."""
    n = args.reads or 1_000_000
    print(json.dumps(probe_pe_run(args, n, 0 if args.no_cpu else 20000)), flush=True)


WORKLOADS = {
    "genome_se_3gbp": {"reads": 10_000_000, "train_reads": 131072, "pe_pairs": 65536,
                       "desc": "single-end, 3.1 Gbp / 24 contigs (45% interspersed repeats, 5% segmental "
                               "duplications), reads of 60-250 bp"},
    "uneven_se_100mbp": {"reads": 2_000_000, "train_reads": 131072, "pe_pairs": 0,
                         "desc": "single-end, 100 Mbp / 10 contigs, 2M reads of 80-150 bp"},
    "prefilter_se_2mbp": {"reads": 2_000_000, "train_reads": 500_000, "pe_pairs": 0,
                          "desc": "single-end, 2.1 Mbp probe genome, 2M held-out reads of 80-150 bp"},
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


def stage_rooflines(agg, B, workload):
    """Roofline position of each stage's kernels per launch, from the library's HIP-event times
    and the oracle-defined algorithmic work units, and the stage that takes the most time."""
    n_seed, n_cl, n_dp = (max(agg[k], 1) for k in ("n_k_seed", "n_k_cluster", "n_k_dp"))
    ms_seed, ms_cl, ms_dp = agg["ms_k_seed"] / n_seed, agg["ms_k_cluster"] / n_cl, agg["ms_k_dp"] / n_dp
    gather_bytes = 8 * (agg["lookups"] + agg["positions"] + agg["filtered_hits"]) / n_seed
    cl_bytes = (8 * agg["filtered_hits"] + CAND_BYTES * agg["reads"]) / n_cl
    cells = agg["dp_cells"] / n_dp
    seed = {"kernel": "k_seed", "bound": "hbm", "unit": "GB/s", "peak": PEAK_HBM_GBPS,
            "achieved": round(gather_bytes / (ms_seed * 1e-3) / 1e9, 2) if ms_seed > 0 else 0.0,
            "traffic": measured_traffic("pbw::k_seed<false>", B, workload),
            "per_launch": {"bytes": gather_bytes, "ms": ms_seed,
                           "unit": "8 B per lookup + 8 B per position + 8 B per filtered hit"}}
    cluster = {"kernel": "k_cluster (all tiers)", "bound": "hbm", "unit": "GB/s", "peak": PEAK_HBM_GBPS,
               "achieved": round(cl_bytes / (ms_cl * 1e-3) / 1e9, 2) if ms_cl > 0 else 0.0,
               "traffic": measured_traffic(("pbw::k_cluster<false 128u>", "pbw::k_cluster<false 512u>",
                                            "pbw::k_cg_group", "pbw::k_cg_sort", "pbw::k_cg_select"),
                                           B, workload),
               "per_launch": {"bytes": cl_bytes, "ms": ms_cl,
                              "unit": "8 B per hit key read + 321 B of candidates per read written"}}
    ach = cells / (ms_dp * 1e-3) * DP_OPS_PER_CELL / 1e12 if ms_dp > 0 else 0.0
    dp = {"kernel": "DP (k_dp_quad rows, k_dp_walk + k_dp_lk beside them)", "bound": "valu-int32", "unit": "Tops/s",
          "peak": round(PEAK_INT32_TOPS, 2), "achieved": round(ach, 3),
          "traffic": measured_traffic(("pbw::k_dp_quad", "pbw::k_dp_walk", "pbw::k_dp_lk"), B, workload),
          "per_launch": {"cells": cells, "ms": ms_dp, "ops_per_cell": DP_OPS_PER_CELL},
          "counters": measured_valu("pbw::k_dp_quad", B, workload)}
    for r in (seed, cluster, dp):
        r["frac"] = round(r["achieved"] / r["peak"], 4)
    times = {"seed": ms_seed, "cluster": ms_cl, "dp": ms_dp}
    stages = {"seed": seed, "cluster": cluster, "dp": dp}
    top = max(times, key=times.get)
    return stages[top], stages, cells / (ms_dp * 1e-3) / 1e9 if ms_dp > 0 else 0.0, seed["achieved"]


def run_steps(ctx, dreads, B, n, k, label):
    agg, recs_all = {}, []
    t0 = time.perf_counter()
    for s in range(k):
        lo = (s * B) % n
        hi = min(n, lo + B)
        recs = ctx.map_se_device(dreads, lo, hi)
        st = ctx.stats()
        recs_all.append(recs)
        for key, v in st.items():
            agg[key] = agg.get(key, 0) + v
        log("%s step %d: %d reads in %.1f ms" % (label, s, st["reads"], st["ms_total"]))
    return agg, recs_all, time.perf_counter() - t0


def pe_leg(ctx, contigs, pairs):
    """Paired-end throughput on the same genome and context: one map_pe call over `pairs` FR
    pairs (500 serial warm-up pairs, then 2000-pair frozen-snapshot batches)."""
    import probebwa_amd as P
    from probebwa_amd import synth
    a, b = synth.reads_pairs_genome(contigs, n=pairs, seed=58)
    r1, r2 = P.ReadBatch.from_arrays(*a), P.ReadBatch.from_arrays(*b)
    ctx.map_pe(r1.subset(0, 2500), r2.subset(0, 2500))  # warm-up call
    t0 = time.perf_counter()
    recs = ctx.map_pe(r1, r2)
    dt = time.perf_counter() - t0
    st = ctx.stats()
    proper = float(((recs.rec["flag"] & 2) != 0).mean())
    log("paired-end: %d pairs in %.2f s" % (pairs, dt))
    return {"pairs": pairs, "pairs_per_s": round(pairs / dt, 1), "reads_per_s": round(2 * pairs / dt, 1),
            "seconds": round(dt, 3), "proper_pair_frac": round(proper, 4), "wide_jobs": int(st["wide_jobs"]),
            "wide_cells": int(st["wide_cells"]), "wide_roofline": wide_roofline(st),
            "insert_model": [round(x, 3) for x in ctx.insert_model[:2]],
            "stage_ms": {k: round(st[k], 1) for k in ("ms_seed", "ms_cluster", "ms_dp", "ms_pair", "ms_total")},
            "data": "FR pairs of 2x150 (fragment N(350,50) clipped to [200,600]), seed 58"}


def e2e_leg(ctx, reads, n, batch):
    """End to end, FASTQ in -> SAM out (the map subcommand, src/main.rs:130-195): the first n
    reads written as a FASTQ file, then parsed, mapped and written as SAM by pbw_map_files in
    batches of `batch` reads (a quarter of the timed steps' batch: parsing the next batch and
    formatting the last overlap the mapping of the current one, so more batches keep the three
    stages busier); one untimed call, then one timed."""
    import shutil
    import tempfile
    import probebwa_amd as P
    seq, so, qual, qo = reads
    d = tempfile.mkdtemp(prefix="pbw_e2e_")
    try:
        fq, sam = os.path.join(d, "reads.fq"), os.path.join(d, "out.sam")
        q33 = (np.asarray(qual[:qo[n]], np.uint16) + 33).astype(np.uint8).tobytes()
        sb = np.asarray(seq[:so[n]], np.uint8).tobytes()
        with open(fq, "wb") as f:
            for i in range(n):
                f.write(b"@r%d\n%s\n+\n%s\n" % (i, sb[so[i]:so[i + 1]], q33[qo[i]:qo[i + 1]]))
        P.map_files(ctx, fq, out_path=sam, command_line="bench", batch_reads=batch)
        rep = P.map_files(ctx, fq, out_path=sam, command_line="bench", batch_reads=batch)
        log("end to end: %d reads in %.2f s" % (n, rep["ms_total"] / 1000.0))
        return {"reads": n, "reads_per_s": round(n / (rep["ms_total"] / 1000.0), 1),
                "ms_total": round(rep["ms_total"], 1), "ms_parse": round(rep["ms_parse"], 1),
                "ms_map": round(rep["ms_map"], 1), "ms_format_write": round(rep["ms_write"], 1),
                "fastq_bytes": int(rep["bytes_in"]), "sam_bytes": int(rep["bytes_out"]), "batch_reads": batch,
                "data": "the workload's first %d reads as FASTQ (Phred+33), SAM to a local file" % n}
    finally:
        shutil.rmtree(d, ignore_errors=True)


def cpu_leg(args, contigs, reads, first_recs, sample, sample_1t):
    """The oracle (C++ restatement of the reference CPU path) on bounded samples of the timed
    reads: `sample` reads on HOST_THREADS threads (the baseline) and `sample_1t` on one; the
    sample's SAM concordance with the HIP path; and how many of the sample's records change
    when the oracle's cluster group order switches from the emulated FxHashMap order to
    (contig, forward before reverse) -- the reads whose records depend on the one ordering
    no reference test pins."""
    import probebwa_amd as P
    from oracle import pbw_oracle as O
    seq, so, qual, qo = reads
    t1 = time.time()
    go = O.Genome()
    for name, s in contigs:
        go.add_contig(name, s.tobytes())
    ht = O.HashTable(go, progress=True, threads=args.cpu_threads)
    log("oracle index %.1f s (%d threads)" % (time.time() - t1, args.cpu_threads))

    def sub(S):
        return O.ReadBatch.from_arrays(seq[:int(so[S])], so[:S + 1], qual[:int(qo[S])], qo[:S + 1])

    m = O.Mapper(go, ht, threads=args.cpu_threads)
    t1 = time.perf_counter()
    ro = m.map_se(sub(sample))
    cdt = time.perf_counter() - t1
    log("oracle %d reads on %d threads in %.1f s" % (sample, args.cpu_threads, cdt))
    cpu = {"value": round(sample / cdt, 1), "unit": "reads/s", "cores": args.cpu_threads, "kind": "port",
           "sample": "first %d reads of the timed reads, oracle (C++ restatement of the reference CPU path) "
                     "on %d threads, %.1f s" % (sample, args.cpu_threads, cdt)}
    concord = float(P.primary_identical(ro, P.Records(first_recs.rec[:sample], first_recs.cigar)).mean())
    O.set_group_order(1)
    try:
        rl = O.Mapper(go, ht, threads=args.cpu_threads).map_se(sub(sample))
    finally:
        O.set_group_order(0)
    changed = int((~P.primary_identical(ro, rl)).sum() + ((ro.rec["log_lk"] != rl.rec["log_lk"])
                                                           & P.primary_identical(ro, rl)).sum())
    m1 = O.Mapper(go, ht, threads=1)
    t1 = time.perf_counter()
    m1.map_se(sub(sample_1t))
    cdt1 = time.perf_counter() - t1
    log("oracle %d reads on 1 thread in %.1f s" % (sample_1t, cdt1))
    cpu1 = {"value": round(sample_1t / cdt1, 1), "unit": "reads/s", "cores": 1, "kind": "port",
            "sample": "first %d reads of the timed reads, oracle on 1 thread, %.1f s" % (sample_1t, cdt1)}
    return cpu, cpu1, concord, {"reads": sample, "records_changed": changed}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=8)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--batch", type=int, default=262144)
    ap.add_argument("--workload", choices=sorted(WORKLOADS) + ["probe_pe_2mbp"], default="genome_se_3gbp")
    ap.add_argument("--reads", type=int, default=None,
                    help="reads generated (default: enough for the timed steps, at most the workload's count)")
    ap.add_argument("--cpu-sample", type=int, default=None)
    ap.add_argument("--cpu-threads", type=int, default=HOST_THREADS)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-learned", action="store_true")
    ap.add_argument("--no-pe", action="store_true")
    ap.add_argument("--no-e2e", action="store_true", help="skip the FASTQ -> SAM end-to-end leg")
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

    wl = WORKLOADS[args.workload]
    n_reads = args.reads or min(wl["reads"], args.batch * max(args.steps, 1))
    n_train = args.train_reads or wl["train_reads"]
    t0 = time.time()
    contigs, gen_reads, gen_train, seeds = make_workload(args.workload, n_reads, n_train)
    log("genome %.1f s" % (time.time() - t0))
    t0 = time.time()
    seq, so, qual, qo = gen_reads()
    log("reads %.1f s (%d reads, %d bases)" % (time.time() - t0, n_reads, int(so[-1])))

    import probebwa_amd as P
    t0 = time.time()
    gp = P.Genome()
    for name, s in contigs:
        gp.add_contig(name, s.tobytes())
    ctx = P.Context(gp, batch_reads=args.batch, device=0)
    log("context (genome packing, index build on the device, upload) %.1f s, %d index entries"
        % (time.time() - t0, ctx.total_entries))

    batch_all = P.ReadBatch.from_arrays(seq, so, qual, qo)
    dreads = ctx.upload_reads(batch_all)  # first upload: allocation
    dreads = None
    t0 = time.perf_counter()
    dreads = ctx.upload_reads(batch_all)
    t_up = time.perf_counter() - t0
    n = n_reads
    B = min(args.batch, n)
    log("reads upload %.3f s (%.2f GB)" % (t_up, (int(so[-1]) * 2) / 1e9))

    tuned = None
    if args.autotune:
        t1 = time.time()
        rep = ctx.autotune(dreads, 0, min(n, 524288))
        tuned = {"best": rep["best"], "best_ms": round(rep["best_ms"], 2), "trials": len(rep["trials"]),
                 "default_ms": round(rep["trials"][0]["ms"], 2), "seconds": round(time.time() - t1, 1)}
        B = min(ctx.tune.batch_reads, n) if args.batch == 262144 else B
        log("autotune %.1f s: %s" % (time.time() - t1, rep["best"]))

    run_steps(ctx, dreads, B, n, args.warmup, "warmup")
    agg, plain_recs, dt = run_steps(ctx, dreads, B, n, args.steps, "plain")
    reads_done = int(agg["reads"])
    value = reads_done / dt
    up_per_step = t_up * min(1.0, reads_done / n)  # PCIe time of the reads the steps mapped
    roof, stages, dp_gcups, gather_gbps = stage_rooflines(agg, B, args.workload)

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
        run_steps(ctx, dreads, B, n, args.warmup, "learned warmup")
        agg2, learned_recs, dt2 = run_steps(ctx, dreads, B, n, args.steps, "learned")
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

    pe = None
    if wl["pe_pairs"] and not args.no_pe:
        pe = pe_leg(ctx, contigs, wl["pe_pairs"])

    e2e = None
    if not args.no_e2e:
        e2e = e2e_leg(ctx, (seq, so, qual, qo), min(3 * B, n), max(B // 4, 16384))

    cpu = cpu1 = concord = order = None
    if not args.no_cpu:
        big = args.workload == "genome_se_3gbp"
        S = min(args.cpu_sample or (32768 if big else 131072), B)
        S1 = min(2048 if big else 16384, S)
        del ctx, dreads
        cpu, cpu1, concord, order = cpu_leg(args, contigs, (seq, so, qual, qo), plain_recs[0], S, S1)

    # the paired-end workload on the probe genome, bounded: 200,000 pairs, the oracle on 8,000
    probe_pe = None
    if args.workload == "genome_se_3gbp" and not args.no_pe:
        probe_pe = probe_pe_run(args, 200_000, 0 if args.no_cpu else 8000)

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
                   "reads_per_step": B, "reads_generated": n, "path": "plain (no learned prefilter)"},
        "pcie_inclusive_reads_per_s": round(reads_done / (dt + up_per_step), 1),
        "upload_ms_per_step": round(up_per_step * 1000.0 / args.steps, 3),
        "dp_gcups": round(dp_gcups, 2),
        "gather_gbps": round(gather_gbps, 2),
        "roofline": roof,
        "roofline_stages": stages,
        "autotune": tuned,
        "learned": learned,
        "learned_vs_plain": learned["speedup_vs_plain"] if learned else None,
        "paired_end": pe,
        "probe_pe_2mbp": probe_pe,
        "end_to_end_fastq_to_sam": e2e,
        "cpu_baseline": cpu,
        "cpu_baseline_1_thread": cpu1,
        "sample_sam_concordance": concord,
        "cluster_group_order_sensitivity": order,
        "stage_ms": {k: round(agg[k], 1) for k in ("ms_seed", "ms_cluster", "ms_dp", "ms_mapq", "ms_total")},
        "kernel_ms_per_step": {k: round(agg["ms_k_" + k] / args.steps, 3) for k in ("seed", "cluster", "dp")},
        "counters": {k: int(agg[k]) for k in ("reads", "lookups", "positions", "filtered_hits", "dp_cells",
                                               "seed_overflow_reads", "tier2_cluster_reads", "band_dp_jobs",
                                               "generic_dp_jobs", "mapq_near_boundary")},
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
