"""The algorithmic units bench.py prices the sampler's roofline with: the
storage form the library picks for a model (block form for the normal
projection up to N = 256, N = 104 with the fused measurement) and the HBM
bytes of one walker sweep in each form."""
import os
import sys

import pytest

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
import bench  # noqa: E402


@pytest.mark.parametrize("projection,N,fused,block", [
    ("normal", 100, False, True), ("normal", 100, True, True), ("normal", 256, False, True),
    ("normal", 144, True, False), ("normal_exact", 100, False, False), ("transverse", 36, False, False)])
def test_sampler_form(projection, N, fused, block):
    assert bench.sampler_form(projection, N, fused) is block


def test_sweep_bytes():
    # block form: each of the N proposals reads the (N/2)^2 entries of F^-1 of
    # the C copies (8 C N^2 / 4 bytes), an accepted one writes them once
    C, N, A = 16, 100, 0.3
    per_proposal_read = 8 * C * (N // 2) ** 2
    assert bench.sweep_bytes(True, C, N, A) == pytest.approx(N * per_proposal_read * (1 + A))
    assert bench.sweep_bytes(True, C, N, A) == pytest.approx(4.16e7)
    # folded skew storage: N^2 / 2 entries per copy read per proposal, an
    # accepted one reads and writes them once more
    assert bench.sweep_bytes(False, C, N, A) == pytest.approx(N * 8 * C * N * N / 2 * (1 + 2 * A))
    # the 16x16 workload at its measured acceptance
    assert bench.sweep_bytes(True, 16, 256, 0.2907616848848304) == pytest.approx(6.9297e8, rel=1e-4)
