"""CPU-only checks of the product library: it loads and exports every symbol
include/probebwa_amd.h declares (no compute calls: this container has no GPU)."""
import ctypes
import os
import re

import probebwa_amd as P

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_symbols():
    text = open(os.path.join(ROOT, "include", "probebwa_amd.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(pbw_[a-z_0-9]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(P.LIB_PATH)
    syms = header_symbols()
    assert len(syms) >= 20
    for s in syms:
        assert hasattr(lib, s), s
    assert sorted(P.EXPORTED_SYMBOLS) == syms


def test_default_options_match_reference():  # src/types.rs:233-249
    o = P.MapperOptions()
    assert (o.substitution_rate, o.insert_size_mean, o.insert_size_sd, o.sv_prior_phred) == (0.001, 250.0, 60.0, 55.0)
    assert (o.max_indel_len, o.gap_open_phred, o.gap_extend_phred) == (30, 40, 3)


def test_genome_host_side():
    g = P.Genome().add_contig("c1", "ACGTN" * 10).add_contig("c2", "acgt")
    assert g.names == ["c1", "c2"] and g.lengths == [50, 4]


def test_code_of_matches_hash_packing():
    assert P.code_of("AAAAAAAAAAAAAAA") == 0
    assert P.code_of("TTTTTTTTTTTTTTT") == (1 << 30) - 1
    assert P.code_of("CAAAAAAAAAAAAAA") == 1 << 28


def test_rust_crate_declares_the_abi():
    """probebwa_amd/rust (the probebwa-amd-sys crate; cargo is absent here, so it is not
    built): every function of the header is declared in its extern block, and every
    #[repr(C)] struct has the C struct's field count."""
    src = open(os.path.join(ROOT, "probebwa_amd", "rust", "src", "lib.rs")).read()
    ext = src[src.index('extern "C" {'):]
    ext = ext[:ext.index("\n}\n")]
    declared = sorted(set(re.findall(r"pub fn (pbw_[a-z_0-9]+)\(", ext)))
    assert declared == header_symbols()
    text = re.sub(r"/\*.*?\*/", "", open(os.path.join(ROOT, "include", "probebwa_amd.h")).read(), flags=re.S)
    for body, name in re.findall(r"typedef struct \{(.*?)\} (pbw_[a-z_]+);", text, flags=re.S):
        n_c = sum(len(decl.split(",")) for decl in body.split(";") if decl.strip())
        m = re.search(r"pub struct %s \{(.*?)\n\}" % name, src, flags=re.S)
        assert m, name
        n_r = len(re.findall(r"pub [a-z_0-9]+:", m.group(1)))
        assert n_r == n_c, (name, n_r, n_c)
