"""Writes tests/golden/genome_3gbp.npz: the oracle's records for a sample of the
genome-scale workload (3.1 Gbp over 24 contigs, probebwa_amd.synth.genome_scale_3gbp
seed 5), so that a GPU test can check the HIP path at that scale without building the
oracle's index on the GPU host.

Contents:
  genome_digest      digest of the generated genome (drift guard)
  se_digest          digest of the single-end sample (4096 reads, L ~ U{60..250}, seed 505)
  se_rec, se_cigar   the oracle's records for them (SingleEndMapper::map per read)
  se_sam_sha256      SHA-256 of their SAM text (format_sam, default header)
  pe_digest          digest of the paired sample (1024 FR pairs of 2 x 150, seed 508)
  pe_rec, pe_cigar   the oracle's map_paired records (interleaved r1, r2)
  pe_sam_sha256      SHA-256 of their SAM text (read names p<i> for both mates)
  pe_model           insert model after the call (mean, sd, n_observed)

Run from the repository root (about 64 GB of host memory, several minutes):
    python tests/golden/make_genome_3gbp.py [--threads N]
"""
import argparse
import hashlib
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

OUT = os.path.join(ROOT, "tests", "golden", "genome_3gbp.npz")
SE_READS, SE_SEED = 4096, 505
PE_PAIRS, PE_SEED = 1024, 508


def genome_digest(contigs):
    h = hashlib.sha256()
    for name, s in contigs:
        h.update(name.encode() + len(s).to_bytes(8, "little"))
        h.update(s[:1 << 20].tobytes())
        h.update(s[-(1 << 20):].tobytes())
        h.update(s[::997].tobytes())
    return h.hexdigest()


def reads_digest(*arrays):
    h = hashlib.sha256()
    for a in arrays:
        h.update(np.ascontiguousarray(a).tobytes())
    return h.hexdigest()


def samples(contigs):
    from probebwa_amd import synth
    se = synth.reads_uneven_se(contigs, n=SE_READS, seed=SE_SEED, lmin=60, lmax=250)
    pe = synth.reads_pairs_genome(contigs, n=PE_PAIRS, seed=PE_SEED)
    return se, pe


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--threads", type=int, default=min(16, os.cpu_count() or 1))
    args = ap.parse_args()
    from oracle import pbw_oracle as O
    from probebwa_amd import synth

    t0 = time.time()
    contigs = synth.genome_scale_3gbp()
    gd = genome_digest(contigs)
    se, pe = samples(contigs)
    print("genome + samples %.1f s" % (time.time() - t0), flush=True)
    go = O.Genome()
    for name, s in contigs:
        go.add_contig(name, s.tobytes())
    del contigs
    t0 = time.time()
    ht = O.HashTable(go, progress=True, threads=args.threads)
    print("oracle index %.1f s, %d entries" % (time.time() - t0, ht.total_entries), flush=True)
    m = O.Mapper(go, ht, threads=args.threads)
    t0 = time.time()
    rse = O.ReadBatch.from_arrays(*se)
    ro = m.map_se(rse)
    se_sam = O.format_sam(go, rse, ro)
    print("oracle SE %.1f s" % (time.time() - t0), flush=True)
    t0 = time.time()
    ids = ["p%d" % i for i in range(PE_PAIRS)]
    r1, r2 = O.ReadBatch.from_arrays(*pe[0], ids=ids), O.ReadBatch.from_arrays(*pe[1], ids=ids)
    rp = m.map_pe(r1, r2)
    pe_sam = O.format_sam(go, O.interleave(r1, r2), rp)
    print("oracle PE %.1f s" % (time.time() - t0), flush=True)
    np.savez_compressed(
        OUT, genome_digest=np.array(gd), se_digest=np.array(reads_digest(*se)), se_rec=ro.rec, se_cigar=ro.cigar,
        se_sam_sha256=np.array(hashlib.sha256(se_sam).hexdigest()),
        pe_digest=np.array(reads_digest(*pe[0], *pe[1])), pe_rec=rp.rec, pe_cigar=rp.cigar,
        pe_sam_sha256=np.array(hashlib.sha256(pe_sam).hexdigest()), pe_model=np.array(m.insert_model, np.float64))
    print("wrote %s" % OUT, flush=True)


if __name__ == "__main__":
    main()
