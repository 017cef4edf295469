"""Inputs of the reference's own tests, regenerated exactly.

The reference tests draw genomes from an in-test Xorshift64 generator
(tests/mapping.rs:48-65, tests/cli.rs:180-192, examples/tune_homopolymer.rs:13-30)
and write FASTQ with quality 'I' (Q40, tests/mapping.rs:21-29).  This module
restates those generators so the same genomes and reads can be fed to the
oracle and to the HIP path.
"""

MASK64 = (1 << 64) - 1
BASES = "ACGT"


class Xorshift64:
    """tests/mapping.rs:48-65."""

    def __init__(self, seed):
        self.s = seed & MASK64

    def next_u64(self):
        x = self.s
        x ^= (x << 13) & MASK64
        x ^= x >> 7
        x ^= (x << 17) & MASK64
        self.s = x
        return x

    def next_base(self):
        return BASES[self.next_u64() % 4]

    def gen_seq(self, n):
        return "".join(BASES[self.next_u64() % 4] for _ in range(n))


def revcomp(s):
    """tests/mapping.rs:31-35."""
    return "".join({"A": "T", "T": "A", "C": "G", "G": "C"}.get(c, "N") for c in reversed(s))


def q40(seq):
    """Quality 'I' (Phred+33) for every base, as write_fastq does."""
    return bytes([40] * len(seq))


def homopolymer_scenario(seed, run_len, del_len, n_errors):
    """examples/tune_homopolymer.rs:44-73 scenario(): (genome, read, true_pos)."""
    rng = Xorshift64(seed)
    prefix = rng.gen_seq(200)
    run_base = BASES[rng.next_u64() % 4]
    run = run_base * run_len if run_len > 1 else rng.gen_seq(1)
    suffix = rng.gen_seq(200)
    genome = prefix + run + suffix
    del_start = len(prefix)
    read_start = max(del_start - 37, 0)
    read_end = min(del_start + 40, len(genome))
    a = min(del_start + del_len, len(genome))
    b = max(read_end, del_start + del_len)
    read = list(genome[read_start:del_start] + genome[a:b])
    for _ in range(n_errors):
        i = rng.next_u64() % len(read)
        base = BASES[rng.next_u64() % 4]
        while base == read[i]:
            base = BASES[rng.next_u64() % 4]
        read[i] = base
    return genome, "".join(read), read_start + 1


def homopolymer_battery():
    """examples/tune_homopolymer.rs:95-140: the 285 (run_len, del_len, n_errors, seed) cases."""
    cases = []
    for run_len in (1, 3, 5, 8, 12):
        for del_len in (1, 2, 3, 5, 8):
            if del_len >= run_len and run_len > 1:
                continue
            for n_errors in (0, 2, 4):
                for seed in range(5):
                    vs = seed * 1000 + run_len * 10 + del_len + n_errors * 100000
                    cases.append((run_len, del_len, n_errors, seed, vs))
    return cases


# tests/integration.rs:132-133: real E. coli window and read (900 bp / 100 bp).
ECOLI_WINDOW = (
    "TCCCTGCTGGTGCCACCAGCGTAGACCGTCTGGTGACGTTGGAAGTGCTGTCAGAACCGGGAGCCAGCGCCATTGACCGGATTCTGAAACTGATCGAAGAAGCCGAAGAGCG"
    "TCGCGCTCCCATTGAGCGGTTTATCGACCGTTTCAGCCGTATCTATACGCCCGCGATTATGGCCGTCGCTCTGCTGGTGACGCTGGTGCCACCGCTGCTGTTTGCCGCCAGC"
    "TGGCAGGAGTGGATTTATAAAGGGCTGACGCTGCTGCTGATTGGCTGCCCGTGTGCGTTAGTTATCTCAACGCCTGCGGCGATTACCTCCGGGCTGGCGGCGGCAGCGCGTC"
    "GTGGGGCGTTGATTAAAGGCGGAGCGGCGCTGGAACAGCTGGGTCGTGTTACTCAGGTGGCGTTTGATAAAACCGGTACGCTGACCGTCGGTAAACCGCGCGTTACCGCGAT"
    "TCATCCGGCAACGGGTATTAGTGAATCTGAACTGCTGACACTGGCGGCGGCGGTCGAGCAAGGCGCGACGCATCCACTGGCGCAAGCCATCGTACGCGAAGCACAGGTTGCT"
    "GAACTCGCCATTCCCACCGCCGAATCACAGCGGGCGCTGGTCGGGTCTGGCATTGAAGCGCAGGTTAACGGTGAGCGCGTATTGATTTGCGCTGCCGGGAAACATCCCGCTG"
    "ATGCATTTACTGGTTTAATTAACGAACTGGAAAGCGCCGGGCAAACGGTAGTGCTGGTAGTACGTAACGATGACGTGCTTGGTGTCATTGCGTTACAGGATACCCTGCGCGC"
    "CGATGCTGCAACTGCCATCAGTGAACTGAACGCGCTGGGCGTCAAAGGGGTGATCCTCACCGGCGATAATCCACGCGCAGCGGCGGCAATTGCCGGGGAGCTGGGGCTGGAG"
    "TTTA"
)
ECOLI_READ = "CGGCGCGGCCCCTGGCGGAGGCGGTCGCCCGCGAAGCACAGGTTGCTCACCTCGCCCTTCCCACCGCCGAATCACAGCGGGCGCTGGTCGGGTCTGGCAT"
