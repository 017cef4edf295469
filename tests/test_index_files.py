"""Persisted index files: GenomeIndex `.stidx` and HashTable `.sthash`
(src/index/genome.rs:80-94, src/index/hashtable.rs:66-80), both written by bincode 1.3
with its default options.

The byte layouts are restated here from the serde types (src/types.rs:107-121,
src/index/hashtable.rs:53-64) and bincode 1's published encoding: little-endian
fixed-width integers, usize as u64, a u64 length before every String / Vec / map, struct
fields in declaration order, a map as its length and then (key, value) pairs in any order.
No file written by the reference itself exists in this repository, so byte parity with
real reference files is unpinned; the fixtures pin the layout as restated.
"""
import os
import struct

import numpy as np
import pytest

import probebwa_amd as P
from probebwa_amd import synth


def _u32(v):
    return struct.pack("<I", v)


def _u64(v):
    return struct.pack("<Q", v)


def _str(s):
    b = s.encode()
    return _u64(len(b)) + b


def stidx_bytes(species, assembly, contigs):
    """GenomeIndex as bincode: Contig::from_ascii packing (genome.rs:17-45)."""
    out = _str(species) + _str(assembly) + _u64(len(contigs))
    total = 0
    for name, seq in contigs:
        packed = bytearray((len(seq) + 3) // 4)
        runs, start = [], None
        for p, b in enumerate(seq):
            c = chr(b).upper()
            if c in "ACGT":
                packed[p // 4] |= "ACGT".index(c) << (2 * (p % 4))
                if start is not None:
                    runs.append((start, p))
                    start = None
            elif start is None:
                start = p
        if start is not None:
            runs.append((start, len(seq)))
        out += _str(name) + _u64(len(packed)) + bytes(packed) + _u64(len(runs))
        out += b"".join(_u32(s) + _u32(e) for s, e in runs) + _u64(len(seq))
        total += len(seq)
    return out + _u64(total)


def slice_of(seq):
    """Contig::slice (genome.rs:47-68) of a whole contig: upper-case ACGT, 'N' elsewhere."""
    return bytes(c if c in b"ACGT" else ord("N") for c in bytes(seq).upper())


FIXTURE = [
    ("chr1", b"ACGTacgtNNNNRYACGTTTGCA"),   # lower case, an N run, IUPAC codes
    ("empty", b""),
    ("nstart", b"NNACG"),                    # a run at the start, length not a multiple of 4
    ("nend", b"GATTACANN"),                  # a run to the end
    ("allN", b"NNNNNNN"),
    ("chrM", b"ACGT" * 40 + b"A"),
]


def test_stidx_fixture_loads(tmp_path):
    path = tmp_path / "g.stidx"
    path.write_bytes(stidx_bytes("Homo sapiens", "GRCh38", FIXTURE))
    g = P.Genome.load_stidx(str(path))
    assert g.names == [n for n, _ in FIXTURE]
    assert g.lengths == [len(s) for _, s in FIXTURE]
    for i, (_, s) in enumerate(FIXTURE):
        assert g.get_sequence(i, 0, len(s) + 10) == slice_of(s)
    assert g.get_sequence(0, 4, 10) == slice_of(FIXTURE[0][1])[4:10]


def test_stidx_save_is_byte_exact(tmp_path):
    g = P.Genome()
    for name, s in FIXTURE:
        g.add_contig(name, s)
    path = tmp_path / "w.stidx"
    g.save_stidx(str(path), species="Homo sapiens", assembly="GRCh38")
    assert path.read_bytes() == stidx_bytes("Homo sapiens", "GRCh38", FIXTURE)
    g2 = P.Genome.load_stidx(str(path))
    assert [g2.get_sequence(i, 0, 1 << 20) for i in range(len(FIXTURE))] == [slice_of(s) for _, s in FIXTURE]


@pytest.mark.parametrize("damage", ["truncated", "trailing", "total", "huge_len", "short_packed"])
def test_stidx_rejects_malformed(tmp_path, damage):
    good = stidx_bytes("s", "a", FIXTURE[:1])
    if damage == "truncated":
        bad = good[:-3]
    elif damage == "trailing":
        bad = good + b"\0"
    elif damage == "total":
        bad = good[:-8] + _u64(999)
    elif damage == "huge_len":
        bad = _u64(1 << 60) + good[8:]
    else:  # packed bytes shorter than the length needs
        name, seq = FIXTURE[0]
        body = _str(name) + _u64(2) + b"\0\0" + _u64(0) + _u64(len(seq))
        bad = _str("s") + _str("a") + _u64(1) + body + _u64(len(seq))
    path = tmp_path / "bad.stidx"
    path.write_bytes(bad)
    with pytest.raises(P.PbwError):
        P.Genome.load_stidx(str(path))


# ---------------------------------------------------------------------------
# .sthash: a context's bucket table (device)
# ---------------------------------------------------------------------------
M32 = 0xFFFFFFFF


def mix(h):
    """hashtable.rs:117-124."""
    h ^= h >> 16
    h = (h * 0x7FEB352D) & M32
    h ^= h >> 15
    h = (h * 0x846CA68B) & M32
    h ^= h >> 16
    return h


def test_mix_inverse_table():
    """The loader recovers each bucket's code from its key through the inverse of mix:
    the restated mix equals the oracle's hash of the same code."""
    from oracle import pbw_oracle as O
    rng = np.random.default_rng(2)
    for _ in range(200):
        kmer = bytes(synth.random_bases(rng, 15))
        assert mix(P.code_of(kmer)) == O.hash_kmer(kmer)


def sthash_bytes(buckets, k=15, stride=5, order=None, rng=None):
    """HashTable as bincode: table (FxHashMap<u32, (u32, u32)>), positions, k, seed_stride,
    total_entries.  `buckets` is [(key, [packed...])]; the positions of the buckets are laid
    out in a shuffled order and the map's entries listed in another, as a hash map may."""
    rng = rng or np.random.default_rng(0)
    lay = rng.permutation(len(buckets)) if order is None else order
    positions, where = [], {}
    for i in lay:
        where[i] = (len(positions), len(buckets[i][1]))
        positions += list(buckets[i][1])
    out = _u64(len(buckets))
    for i in rng.permutation(len(buckets)):
        off, cnt = where[i]
        out += _u32(buckets[i][0]) + _u32(off) + _u32(cnt)
    out += _u64(len(positions)) + np.asarray(positions, dtype="<u8").tobytes()
    return out + _u64(k) + _u64(stride) + _u64(len(positions))


def parse_sthash(data):
    n = struct.unpack_from("<Q", data, 0)[0]
    ent = np.frombuffer(data, dtype="<u4", count=3 * n, offset=8).reshape(n, 3)
    at = 8 + 12 * n
    npos = struct.unpack_from("<Q", data, at)[0]
    pos = np.frombuffer(data, dtype="<u8", count=npos, offset=at + 8)
    k, stride, total = struct.unpack_from("<QQQ", data, at + 8 + 8 * npos)
    assert at + 8 + 8 * npos + 24 == len(data)
    return {int(key): pos[off:off + cnt] for key, off, cnt in ent}, k, stride, total


def _oracle_table(contigs):
    from oracle import pbw_oracle as O
    from test_gpu_parity import pair_of_genomes
    go, gp = pair_of_genomes(contigs)
    ht = O.HashTable(go)
    return go, gp, ht, [(h, b) for h, b in ht.buckets()]


@pytest.mark.gpu
def test_sthash_save_matches_oracle_buckets(tmp_path):
    contigs = synth.small_genome(8, n_contigs=3, length=30000)
    _go, gp, ht, buckets = _oracle_table(contigs)
    ctx = P.Context(gp)
    path = tmp_path / "x.sthash"
    ctx.save_sthash(str(path))
    table, k, stride, total = parse_sthash(path.read_bytes())
    assert (k, stride, total) == (15, 5, ht.total_entries)
    assert len(table) == len(buckets)
    for h, b in buckets:
        assert np.array_equal(table[h], np.asarray(b, np.uint64)), "bucket %08x" % h


@pytest.mark.gpu
def test_context_from_sthash_maps_identically(tmp_path):
    from oracle import pbw_oracle as O
    from test_gpu_parity import compare_records
    contigs = synth.small_genome(9, n_contigs=4, length=25000)
    go, gp, ht, buckets = _oracle_table(contigs)
    path = tmp_path / "o.sthash"
    path.write_bytes(sthash_bytes(buckets, rng=np.random.default_rng(7)))
    loaded = P.Context(gp, sthash=str(path))
    built = P.Context(gp)
    assert loaded.total_entries == built.total_entries == ht.total_entries
    rng = np.random.default_rng(4)
    for h, b in buckets[:: max(1, len(buckets) // 1500)]:
        code = None
        cid, pos = (int(b[0]) >> 40) & ((1 << 23) - 1), int(b[0]) & ((1 << 40) - 1)
        kmer = np.asarray(contigs[cid][1][pos:pos + 15], np.uint8)
        if int(b[0]) >> 63:
            kmer = synth.revcomp(kmer)
        code = P.code_of(kmer.tobytes())
        assert mix(code) == h
        assert list(loaded.lookup_code(code)) == list(b)
    seq, so, qual, qo = synth.sample_reads(contigs, 600, 91)
    reads = P.ReadBatch.from_arrays(seq, so, qual, qo)
    ro = O.Mapper(go, ht).map_se(O.ReadBatch.from_arrays(seq, so, qual, qo))
    compare_records(ro, loaded.map_se(reads), label="from .sthash")
    compare_records(ro, built.map_se(reads), label="built")
    del rng


@pytest.mark.gpu
@pytest.mark.parametrize("damage", ["k16", "over_cap", "duplicate", "not_a_code", "off_genome", "truncated"])
def test_sthash_rejects_malformed(tmp_path, damage):
    contigs = synth.small_genome(10, n_contigs=1, length=3000)
    from test_gpu_parity import pair_of_genomes
    _go, gp = pair_of_genomes(contigs)
    seq = bytes(contigs[0][1])
    good = [(mix(P.code_of(seq[p:p + 15])), [p]) for p in (0, 5, 10)]
    buckets, k = good, 15
    if damage == "k16":
        k = 16
    elif damage == "over_cap":
        buckets = good + [(mix(P.code_of(b"ACGTACGTACGTACG")), list(range(301)))]
    elif damage == "duplicate":
        buckets = good + [(good[0][0], [20])]
    elif damage == "not_a_code":
        buckets = good + [(mix((1 << 30) + 5), [20])]
    elif damage == "off_genome":
        buckets = good + [(mix(P.code_of(b"TTTTTTTTTTTTTTT")), [(7 << 40) | 20])]
    data = sthash_bytes(buckets, k=k)
    if damage == "truncated":
        data = data[:-5]
    path = tmp_path / "bad.sthash"
    path.write_bytes(data)
    with pytest.raises(P.PbwError):
        P.Context(gp, sthash=str(path))
    assert os.path.exists(path)
