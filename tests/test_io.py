"""Host I/O: the read-file reader (src/io/fastx.rs:51-305), SAM text (src/io/sam.rs:37-143),
BAM with BGZF and BAI (src/io/bam.rs:19-57) and the map subcommand's file path
(src/main.rs:130-195, src/lib.rs:65-185).

The reader is checked against the oracle's restatement of fastx.rs on hand-made files whose
expected records are also written out by hand.  BAM is decoded here with Python's gzip (BGZF
is multi-member gzip) and parsed field by field against the SAM text; the index is checked by
querying it for regions and comparing with a brute-force overlap scan.  The reference writes
BAM through noodles (noodles-sam 0.86 / noodles-bam 0.91, absent here), so byte parity with
its BAM and BAI output is unpinned: the checks pin the SAM/BAM/BAI specifications instead.
"""
import gzip
import os
import struct

import numpy as np
import pytest

import probebwa_amd as P
from oracle import pbw_oracle as O

FQ_CASES = {
    # multi-line sequence and quality, '+name' separator, lower case, CRLF, '@@' ids,
    # /1 /2 suffixes, an empty read with a blank quality line, quality below the offset,
    # a Unicode no-break space ending the id, a final record without its newline
    "mixed": (b"@r1/1 extra words\nACGTacgt\nNNAC\n+r1\nIIIIIIII\nIIII\n"
              b"@@r2/2\r\nGATTACA\r\n+\r\n!!!!!!!\r\n"
              b"@empty\n\n+\n\n"
              b"@low\nACGT\n+\n\x1f\x21\x22\x23\n"
              b"@nbsp\xc2\xa0tail\nTTTT\n+\nABCD\n"
              b"@last\nCCCC\n+\n55"),
    "quality_wraps": b"@q\nACGTACGTAC\n+\nIIIII\nIIIII\n@n\nA\n+\nI\n",
}


def _expect_mixed():
    return [("r1", b"ACGTACGTNNAC", bytes([40] * 12)),
            ("r2", b"GATTACA", bytes([0] * 7)),
            ("empty", b"", b""),
            ("low", b"ACGT", bytes([0, 0, 1, 2])),
            ("nbsp", b"TTTT", bytes([32, 33, 34, 35])),
            ("last", b"CCCC", bytes([20, 20]))]


def _records(rb):
    return [(rb.ids[i], rb.seq_of(i), rb.qual_of(i)) for i in range(rb.n)]


def _write(tmp_path, name, data, gz=False, members=1):
    path = tmp_path / (name + (".gz" if gz else ""))
    if gz:
        cut = [len(data) * k // members for k in range(members + 1)]
        path.write_bytes(b"".join(gzip.compress(data[cut[k]:cut[k + 1]]) for k in range(members)))
    else:
        path.write_bytes(data)
    return str(path)


@pytest.mark.parametrize("gz,members", [(False, 1), (True, 1), (True, 3)])
def test_fastq_reader_known_records(tmp_path, gz, members):
    path = _write(tmp_path, "mixed.fq", FQ_CASES["mixed"], gz, members)
    got = _records(P.read_fastx(path))
    assert got == _expect_mixed()
    if not gz:  # the oracle reads plain files
        assert _records(O.read_reads(path)) == got


def test_fastq_reader_matches_oracle_and_batches(tmp_path):
    rng = np.random.default_rng(5)
    lines = []
    for i in range(500):
        L = int(rng.integers(0, 90))
        seq = bytes(rng.choice(list(b"ACGTNacgtn"), L).astype(np.uint8))
        qual = bytes(rng.integers(33, 75, L).astype(np.uint8))
        cut = int(rng.integers(0, L + 1))
        sep = b"\r\n" if i % 7 == 0 else b"\n"
        lines.append(b"@read%d/%d desc%s%s%s%s+%s" % (i, 1 + i % 2, sep, seq[:cut], sep + seq[cut:] if cut < L else b"",
                                                     sep, sep))
        lines.append(qual + sep)
    path = _write(tmp_path, "rand.fq", b"".join(lines))
    want = _records(O.read_reads(path, quality_offset=33, adapter="AGATCGGAAG"))
    got = _records(P.read_fastx(path, adapter="AGATCGGAAG"))
    assert got == want and len(got) == 500
    assert _records(P.read_fastx(path, adapter="AGATCGGAAG", batch=7)) == got
    assert _records(P.read_fastx(path, quality_offset=64)) == _records(O.read_reads(path, quality_offset=64))


def test_adapter_trimming_cases(tmp_path):
    """strip_adapter (fastx.rs:97-117): longest overlap first, one mismatch per 10 bases, >= 6."""
    ad = "AGATCGGAAGAGC"
    data = (b"@exact\nTTTTTAGATCGGAAGAGC\n+\nIIIIIIIIIIIIIIIIII\n"       # whole adapter at the end
            b"@mm\nTTTTTAGATCGGTAGAGC\n+\nIIIIIIIIIIIIIIIIII\n"          # one mismatch in 13
            b"@short\nTTTTTTTTTTAGATC\n+\nIIIIIIIIIIIIIII\n"             # 5-base overlap: kept
            b"@six\nTTTTTTTTTAGATCG\n+\nIIIIIIIIIIIIIII\n")              # 6-base overlap: trimmed
    path = _write(tmp_path, "ad.fq", data)
    got = {i: s for i, s, _q in _records(P.read_fastx(path, adapter=ad))}
    assert got == {"exact": b"TTTTT", "mm": b"TTTTT", "short": b"TTTTTTTTTTAGATC", "six": b"TTTTTTTTT"}
    assert _records(P.read_fastx(path, adapter=ad)) == _records(O.read_reads(path, adapter=ad))


def test_fasta_reads_and_genome_fasta(tmp_path):
    data = b">a first\nACGT\n\nacg\n>b\n>c  x\nTTT\nGG\n"
    path = _write(tmp_path, "r.fa", data)
    got = _records(P.read_fastx(path, fasta=True))
    assert got == [("a", b"ACGTACG", bytes([40] * 7)), ("b", b"", b""), ("c", b"TTTGG", bytes([40] * 5))]
    assert _records(O.read_reads(path, fasta=True)) == got
    gzp = _write(tmp_path, "g.fa", data, gz=True)
    g1, g2 = P.Genome().add_fasta(path), P.Genome().add_fasta(gzp)
    assert g1.names == g2.names == ["a", "b", "c"]
    assert [g2.get_sequence(i, 0, 99) for i in range(3)] == [b"ACGTACG", b"", b"TTTGG"]


def test_reader_rejects_invalid_utf8(tmp_path):
    path = _write(tmp_path, "bad.fq", b"@ok\nACGT\n+\nIIII\n@bad\xff\nACGT\n+\nIIII\n")
    with pytest.raises(P.PbwError):
        P.read_fastx(path)
    with pytest.raises(IOError):
        O.read_reads(path)


# ---------------------------------------------------------------------------
# SAM text and BAM from records built by hand (no GPU)
# ---------------------------------------------------------------------------
def _genome():
    g = P.Genome()
    g.add_contig("chrA", b"ACGT" * 30000)
    g.add_contig("chrB", b"GATTACA" * 10000)
    g.add_contig("chrA", b"TTTT" * 100)   # a repeated name: the sort ranks it with its last index
    go = O.Genome()
    go.add_contig("chrA", b"ACGT" * 30000)
    go.add_contig("chrB", b"GATTACA" * 10000)
    go.add_contig("chrA", b"TTTT" * 100)
    return g, go


def _synthetic(n, seed, qual_lo=33, qual_hi=60):
    """Reads and records as a mapping call could return them, built at random."""
    rng = np.random.default_rng(seed)
    seqs, quals, ids, cig = [], [], [], []
    rec = np.zeros(n, P.RECORD_DTYPE)
    lens = [120000, 70000, 400]
    for i in range(n):
        L = int(rng.integers(20, 160))
        seqs.append(bytes(rng.choice(list(b"ACGTN"), L, p=[.24, .24, .24, .24, .04]).astype(np.uint8)))
        quals.append(bytes(rng.integers(qual_lo, qual_hi, L).astype(np.uint8)))
        ids.append("q%d:%d" % (seed, i))
        r = rec[i]
        if rng.random() < 0.15:
            r["contig"], r["pos"], r["flag"], r["mate_contig"] = -1, 0, 4, -1
            continue
        c = int(rng.integers(0, 3))
        ops = []
        left = L
        while left > 0:
            k = int(min(left, rng.integers(1, 40)))
            op = int(rng.choice([0, 0, 0, 1, 2, 4]))
            if op in (1, 4) and k >= left:
                op = 0
            ops.append((k << 4) | op)
            if op != 2:
                left -= k
        r["contig"] = c
        r["pos"] = int(rng.integers(1, lens[c] - 300)) if rng.random() < 0.9 else int(rng.integers(16380, 16390))
        r["flag"] = int(rng.choice([0, 16, 1 | 2 | 64 | 32, 1 | 128 | 16]))
        r["mapq"] = int(rng.integers(0, 61))
        r["strand"] = 1 if r["flag"] & 16 else 0
        r["seq_rc"] = r["strand"]
        r["cigar_off"], r["cigar_n"] = len(cig), len(ops)
        cig += ops
        r["mate_contig"] = c if r["flag"] & 1 else -1
        r["mate_pos"] = int(r["pos"]) + 200 if r["flag"] & 1 else 0
        r["tlen"] = 350 if r["flag"] & 64 else (-350 if r["flag"] & 128 else 0)
    reads = P.ReadBatch(seqs, quals, ids)
    return reads, P.Records(rec, np.asarray(cig if cig else [0], np.uint32)[:len(cig)])


def test_sam_text_matches_oracle_and_lossy_utf8():
    g, go = _genome()
    reads, recs = _synthetic(300, 1, qual_lo=11, qual_hi=250)  # QUAL bytes >= 0x80: from_utf8_lossy
    want = O.format_sam(go, O.ReadBatch.from_arrays(reads.seq, reads.seq_off, reads.qual, reads.qual_off,
                                                     reads.ids), O.Records(recs.rec, recs.cigar),
                        command_line="probebwa map", read_group="ID:g1,SM:s")
    got = P.format_sam(g, reads, recs, command_line="probebwa map", read_group="ID:g1,SM:s")
    assert got == want
    # each line's QUAL is the raw bytes decoded as UTF-8 with U+FFFD for invalid parts
    lines = [ln for ln in got.decode("utf-8").split("\n")[:-1] if not ln.startswith("@")]
    for i, line in enumerate(lines):
        q = reads.qual_of(i)
        if recs.rec[i]["seq_rc"]:
            q = q[::-1]
        assert line.split("\t")[10] == q.decode("utf-8", "replace")
    with pytest.raises(P.PbwError, match="ID tag"):
        P.format_sam(g, reads, recs, read_group="SM:s")  # ReadGroup::parse needs an ID


def reg2bin(beg, end):
    end -= 1
    for shift, base in ((14, 4681), (17, 585), (20, 73), (23, 9), (26, 1)):
        if beg >> shift == end >> shift:
            return base + (beg >> shift)
    return 0


def reg2bins(beg, end):
    end -= 1
    out = [0]
    for shift, base in ((26, 1), (23, 9), (20, 73), (17, 585), (14, 4681)):
        out += list(range(base + (beg >> shift), base + (end >> shift) + 1))
    return out


def read_bgzf(path):
    data = open(path, "rb").read()
    blocks, at = [], 0
    while at < len(data):
        assert data[at:at + 4] == b"\x1f\x8b\x08\x04" and data[at + 12:at + 14] == b"BC"
        bsize = struct.unpack_from("<H", data, at + 16)[0] + 1
        blocks.append((at, gzip.decompress(data[at:at + bsize])))
        at += bsize
    assert blocks[-1][1] == b"" and data[-28:] == bytes.fromhex("1f8b08040000000000ff0600424302001b0003000000000000000000")
    return blocks


def parse_bam(blocks):
    """Records with their virtual offsets: list of (voff_begin, voff_end, fields dict)."""
    stream = b"".join(b for _, b in blocks)
    starts = []
    u = 0
    for coff, b in blocks:
        starts.append((u, coff))
        u += len(b)

    def voff(x):
        k = max(i for i, (u0, _c) in enumerate(starts) if u0 <= x) if x < u else len(starts) - 1
        u0, c = starts[k]
        if x == u:
            return c << 16
        return (c << 16) | (x - u0)

    assert stream[:4] == b"BAM\1"
    lt = struct.unpack_from("<i", stream, 4)[0]
    text = stream[8:8 + lt]
    at = 8 + lt
    nref = struct.unpack_from("<i", stream, at)[0]
    at += 4
    refs = []
    for _ in range(nref):
        ln = struct.unpack_from("<i", stream, at)[0]
        name = stream[at + 4:at + 4 + ln - 1].decode()
        lref = struct.unpack_from("<i", stream, at + 4 + ln)[0]
        refs.append((name, lref))
        at += 8 + ln
    recs = []
    while at < len(stream):
        bs = struct.unpack_from("<i", stream, at)[0]
        body = stream[at + 4:at + 4 + bs]
        ref, pos, lrn, mapq, bin_, ncig, flag, lseq, nref_, npos, tlen = struct.unpack_from("<iiBBHHHiiii", body, 0)
        o = 32
        name = body[o:o + lrn - 1].decode()
        o += lrn
        cig = list(struct.unpack_from("<%dI" % ncig, body, o))
        o += 4 * ncig
        packed = body[o:o + (lseq + 1) // 2]
        o += (lseq + 1) // 2
        seq = "".join("=ACMGRSVTWYHKDBN"[(packed[k // 2] >> (4 * (1 - k % 2))) & 15] for k in range(lseq))
        qual = body[o:o + lseq]
        o += lseq
        recs.append((voff(at), voff(at + 4 + bs), dict(ref=ref, pos=pos, mapq=mapq, bin=bin_, flag=flag, tlen=tlen,
                                                        nref=nref_, npos=npos, name=name, cigar=cig, seq=seq,
                                                        qual=qual, tags=body[o:])))
        at += 4 + bs
    return text, refs, recs


def _sam_fields(sam):
    lines = sam.decode().split("\n")
    return [ln.split("\t") for ln in lines if ln and not ln.startswith("@")]


def record_rows(names, reads, recs, rg_id=None):
    """The SAM fields of each record (format_sam, sam.rs:37-58), built from the records
    rather than parsed from text: QUAL holds raw Phred bytes, which may include tab or newline."""
    ops, comp = "MIDNSHP=X", {65: 84, 84: 65, 67: 71, 71: 67}
    rows = []
    for i in range(len(recs.rec)):
        r = recs.rec[i]
        s, q = reads.seq_of(i), reads.qual_of(i)
        if r["seq_rc"]:
            s, q = bytes(comp.get(c, 78) for c in s[::-1]), q[::-1]
        cig = "".join("%d%s" % (w >> 4, ops[w & 15]) for w in recs.cigar[r["cigar_off"]:r["cigar_off"] + r["cigar_n"]])
        row = [reads.ids[i], str(r["flag"]), names[r["contig"]] if r["contig"] >= 0 else "*", str(r["pos"]),
               str(r["mapq"]), cig or "*", names[r["mate_contig"]] if r["mate_contig"] >= 0 else "*",
               str(r["mate_pos"]), str(r["tlen"]), s.decode("latin-1"), q.decode("latin-1")]
        if rg_id:
            row.append("RG:Z:" + rg_id)
        rows.append(row)
    return rows


def _check_bam_against_sam(bam_recs, sam_rows, names, raw_qual):
    ops = "MIDNSHP=X"
    assert len(bam_recs) == len(sam_rows)
    for (_v0, _v1, b), f in zip(bam_recs, sam_rows):
        assert b["name"] == f[0] and b["flag"] == int(f[1]) and b["mapq"] == int(f[4])
        assert (names[b["ref"]] if b["ref"] >= 0 else "*") == f[2] and b["pos"] + 1 == int(f[3])
        assert ("".join("%d%s" % (w >> 4, ops[w & 15]) for w in b["cigar"]) or "*") == f[5]
        assert (names[b["nref"]] if b["nref"] >= 0 else "*") == f[6] and b["npos"] + 1 == int(f[7])
        assert b["tlen"] == int(f[8]) and b["seq"] == f[9]
        q = bytes(x + (0 if raw_qual else 33) for x in b["qual"])
        assert q == f[10].encode("latin-1")
        span = sum(w >> 4 for w in b["cigar"] if w & 15 in (0, 2, 3, 7, 8))
        assert b["bin"] == (4680 if b["pos"] < 0 else reg2bin(b["pos"], b["pos"] + max(span, 1)))
        tag = [x for x in f[11:] if x.startswith("RG:Z:")]
        assert b["tags"] == (b"RGZ" + tag[0][5:].encode() + b"\0" if tag else b"")


@pytest.mark.parametrize("index", [False, True])
def test_bam_decodes_to_the_sam_text(tmp_path, index):
    g, _go = _genome()
    reads, recs = _synthetic(1500, 2)
    path = str(tmp_path / "o.bam")
    P.write_bam(g, reads, recs, path, command_line="probebwa map", read_group="ID:rg7", index=index, threads=4)
    text, refs, brecs = parse_bam(read_bgzf(path))
    assert refs == list(zip(g.names, g.lengths))
    sam = P.format_sam(g, reads, recs, command_line="probebwa map", read_group="ID:rg7")
    rows = _sam_fields(sam)
    if index:
        # sort_for_coordinate_order (sam.rs:109-124): stable, by the last index of the contig's
        # name, then position; unmapped last
        rank = {n: i for i, n in enumerate(g.names)}
        key = lambda f: (1 << 62, 0) if f[3] == "0" else (rank[f[2]], int(f[3]))
        rows = sorted(rows, key=key)
        assert text.startswith(b"@HD\tVN:1.6\tSO:coordinate\n")
    else:
        assert text == sam[:len(text)] and text.startswith(b"@HD\tVN:1.6\tSO:unsorted\n")
    _check_bam_against_sam(brecs, rows, g.names, raw_qual=False)


def test_bai_queries_match_a_scan(tmp_path):
    g, _go = _genome()
    reads, recs = _synthetic(4000, 3)
    path = str(tmp_path / "s.bam")
    P.write_bam(g, reads, recs, path, index=True, threads=3)
    _text, refs, brecs = parse_bam(read_bgzf(path))
    bai = open(path + ".bai", "rb").read()
    assert bai[:4] == b"BAI\1"
    at, nref = 8, struct.unpack_from("<i", bai, 4)[0]
    assert nref == len(refs)
    index = []
    for _ in range(nref):
        nbin = struct.unpack_from("<i", bai, at)[0]
        at += 4
        bins, meta = {}, None
        for _ in range(nbin):
            b, nch = struct.unpack_from("<Ii", bai, at)
            at += 8
            ch = [struct.unpack_from("<QQ", bai, at + 16 * k) for k in range(nch)]
            at += 16 * nch
            if b == 37450:
                meta = ch
            else:
                bins[b] = ch
        nintv = struct.unpack_from("<i", bai, at)[0]
        lin = list(struct.unpack_from("<%dQ" % nintv, bai, at + 4))
        at += 4 + 8 * nintv
        index.append((bins, lin, meta))
    n_no_coor = struct.unpack_from("<Q", bai, at)[0]
    assert at + 8 == len(bai)
    placed = [r for r in brecs if r[2]["ref"] >= 0]
    assert n_no_coor == len(brecs) - len(placed)

    def span(r):
        s = sum(w >> 4 for w in r["cigar"] if w & 15 in (0, 2, 3, 7, 8))
        return r["pos"], r["pos"] + max(s, 1)

    for ref, (bins, lin, meta) in enumerate(index):
        on = [r for r in placed if r[2]["ref"] == ref]
        if on:
            assert meta[1] == (sum(1 for r in on if not r[2]["flag"] & 4), sum(1 for r in on if r[2]["flag"] & 4))
    rng = np.random.default_rng(9)
    for _ in range(300):
        ref = int(rng.integers(0, 3))
        beg = int(rng.integers(0, refs[ref][1]))
        end = beg + int(rng.integers(1, 40000))
        want = {r[0] for r in placed if r[2]["ref"] == ref and span(r[2])[0] < end and span(r[2])[1] > beg}
        bins, lin, _m = index[ref]
        minoff = lin[beg >> 14] if (beg >> 14) < len(lin) else (lin[-1] if lin else 0)
        chunks = [c for b in reg2bins(beg, end) for c in bins.get(b, []) if c[1] > minoff]
        got = set()
        for r in placed:
            if r[2]["ref"] != ref or not any(c0 <= r[0] < c1 for c0, c1 in chunks):
                continue
            s0, s1 = span(r[2])
            if s0 < end and s1 > beg:
                got.add(r[0])
        assert got == want


def test_bam_rejects_what_noodles_rejects(tmp_path):
    g, _go = _genome()
    reads, recs = _synthetic(50, 4, qual_lo=20, qual_hi=45)  # raw Phred bytes below '!'
    with pytest.raises(P.PbwError, match="QUAL"):
        P.write_bam(g, reads, recs, str(tmp_path / "x.bam"))
    P.write_bam(g, reads, recs, str(tmp_path / "y.bam"), raw_qual=True)
    _t, _r, brecs = parse_bam(read_bgzf(str(tmp_path / "y.bam")))
    for (_a, _b, b), i in zip(brecs, range(reads.n)):
        q = reads.qual_of(i)
        assert b["qual"] == (q[::-1] if recs.rec[i]["seq_rc"] else q)
    bad = P.ReadBatch([reads.seq_of(0)], [reads.qual_of(0)], ["has@at"])
    one = P.Records(recs.rec[:1].copy(), recs.cigar)
    with pytest.raises(P.PbwError, match="QNAME"):
        P.write_bam(g, bad, one, str(tmp_path / "z.bam"), raw_qual=True)
