"""The autotuner's table emission (betanoox_amd/autotune.py) on synthetic
sweep data (no GPU): size rows, ties, thresholds and the emitted header."""
import os

from betanoox_amd import autotune


def _rec(b, n_slices, ms_by_id, gen="elastic_cube", n=0):
    # iteration times: the SpMV time plus the same vector-kernel time for every shape
    table = [{"id": i, "slices_per_block": 4, "unroll": 4, "nontemporal": int(i >= 9), "column_bytes": 2,
              "ms": ms, "gbps": 1.0, "iter_ms": ms + 0.001} for i, ms in enumerate(ms_by_id)]
    chosen = min(range(len(ms_by_id)), key=lambda i: ms_by_id[i])
    return {"gen": gen, "b": b, "n": n or n_slices, "n_nodes": 64 * n_slices, "n_slices": n_slices,
            "chosen": chosen, "table": table}


def _ms(base, best_id=None, gain=0.0):
    v = [base] * 18
    if best_id is not None:
        v[best_id] = base * (1 - gain)
    return v


def test_rows_keep_default_within_tie_and_switch_beyond():
    recs = [
        _rec(3, 100, _ms(0.006, 2, 0.01)),    # 1% better: a tie, shape 0 stays
        _rec(3, 1000, _ms(0.02, 3, 0.005)),   # tie
        _rec(3, 8000, _ms(0.10, 9, 0.15)),    # streaming clearly faster
        _rec(3, 20000, _ms(0.25, 9, 0.10)),
    ]
    rows = autotune._rows(recs)
    assert [shape for _, shape in rows] == [0, 9]
    # the boundary lies between the last size of the first row and the first of the second
    assert 1000 < rows[0][0] < 8000
    assert rows[-1][0] >= 1 << 40


def test_emit_header(tmp_path):
    data = {"single": [_rec(3, 500, _ms(0.01)), _rec(3, 9000, _ms(0.1, 9, 0.2)),
                       _rec(1, 30000, _ms(0.08, 11, 0.2), gen="diffusion_cube")],
            "multi": [], "reps": 10}
    path = os.path.join(str(tmp_path), "tuned.hpp")
    autotune.emit_header(data, path)
    txt = open(path).read()
    assert "constexpr TunedSellRow kTunedSell[] = {" in txt
    assert "{3, " in txt and "{1, 1LL << 40, 11}," in txt
    assert txt.count("1LL << 40") == 2   # one open-ended row per block size measured
    md = os.path.join(str(tmp_path), "sweep.md")
    autotune.emit_md(data, md)
    assert "| elastic_cube |" in open(md).read()


def test_committed_table_covers_every_block_size():
    hdr = open(autotune.HEADER).read()
    for b in (1, 2, 3):
        assert "{%d, 1LL << 40," % b in hdr
