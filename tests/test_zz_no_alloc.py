"""pbw_align_batch_noalloc, the entry point that allocates nothing, against pbw_align_batch
(the batched SmithWatermanAligner::align, src/align/smith_waterman.rs:113,163): identical
scores, CIGAR words and offsets on two calls in a row (the second smaller), in a fresh
process or after any other test; a batch beyond its fixed region is refused and both
entry points keep working."""
import numpy as np
import pytest

import probebwa_amd as P
from test_gpu_parity import random_align_jobs

pytestmark = pytest.mark.gpu


def test_noalloc_align_equals_align_batch():
    g = P.Genome().add_contig("x", "ACGT" * 50)
    ctx = P.Context(g)
    rng = np.random.default_rng(2024)
    first = random_align_jobs(rng, 600) + random_align_jobs(rng, 200, wide=True)
    second = random_align_jobs(rng, 150) + random_align_jobs(rng, 40, wide=True)
    long_windows = [(b"ACGT" * 25, bytes([30] * 100), b"TTGCA" * 400, 2000, 900)] * 8  # tiled kernel, 2000 columns
    for jobs in (first, second, long_windows):
        want = ctx.align_batch(jobs)
        got = ctx.align_batch_noalloc(jobs)
        assert got == want
    big = [(b"A" * 1000, bytes([30] * 1000), b"A" * 8000, 8000, 0)] * 2000  # beyond the 16 MiB host stage
    with pytest.raises(P.PbwError, match="no-allocation"):
        ctx.align_batch_noalloc(big)
    assert ctx.align_batch_noalloc(second) == ctx.align_batch(second)
