"""bench.py's workload lists and algorithmic byte counts: the uneven batch list (seed 3 + i,
k in [10, 18], depth in [4, 24]), the round-pair split of a phase, and the per-pair byte
counts against a direct count of the tables each pass reads and writes."""
import os
import sys

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
import bench  # noqa: E402


def test_batch_specs_shape():
    specs = bench.batch_specs(512)
    assert len(specs) == 512
    assert [s for s, _, _ in specs] == list(range(3, 515))
    assert min(k for _, _, k in specs) == 10 and max(k for _, _, k in specs) == 18
    assert min(d for _, d, _ in specs) == 4 and max(d for _, d, _ in specs) == 24
    # k and depth are drawn as two arrays of `count`: a slice of the workload is a prefix of
    # the 512-proof list, not batch_specs(32)
    assert bench.batch_specs(32) != specs[:32]


def test_pair_split():
    # pairs run while the table has more than 2048 rows; K rounded up to even
    assert [bench.pair_count(n) for n in (10, 11, 12, 13, 14, 15, 18, 20)] == [0, 0, 1, 1, 2, 2, 4, 5]


def _direct(n, R, S, phase, j):
    """Bytes read and written by pair j, counted table by table."""
    N = 1 << n
    hrow = S if phase == 1 else 4
    if j == 0:
        src = N
        reads = src * S + R * src * hrow  # f: the shared layer input; h: per repetition
        return reads
    src = N >> (2 * (j - 1))
    dst = src >> 2
    f_read = src * S if j == 1 else R * src * S
    return f_read + R * src * hrow + R * dst * (S + hrow)


def test_pair_bytes_match_a_direct_count():
    for n in (12, 14, 18, 20):
        for phase in (1, 2):
            for j in range(bench.pair_count(n)):
                assert bench.pair_task_bytes(n, 3, 64, phase, j) == _direct(n, 3, 64, phase, j), (n, phase, j)


def test_single_proof_roofline_bytes():
    # 24 x 2^20 M31x16: phase-1 double-fold pairs j = 1..4 of each layer, 96 launches
    per_layer = sum(bench.pair_task_bytes(20, 3, 64, 1, j) for j in range(1, 5))
    assert per_layer == (64 + 192 + 96 + 120 + 30 + 7.5) * 2 ** 20
    total = bench.circuit_pair_bytes([20] * 24, 3, 64, phases=(1,), first=False)
    assert total == 24 * per_layer
    assert abs(total / 96 - 133.6e6) < 0.1e6
