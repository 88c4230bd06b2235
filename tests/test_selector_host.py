"""Host logic of the format selector (no GPU): targets, regret, the
row-length marks behind the hybrid widths, tuning labels and rules."""
import math

import numpy as np
import pytest

from matrix_format_performance_amd import selector as sel


def test_soft_targets_and_regret():
    t = np.array([[2.0, 1.0, math.inf, 1.05, 4, 4, 4, 4]])
    p = sel.soft_targets(t, tau=0.05)
    assert p.shape == (1, 8) and p[0, 2] == 0 and abs(p.sum() - 1) < 1e-6
    assert p[0, 1] / p[0, 3] == pytest.approx(math.e, rel=1e-5)  # 5% slower = one tau
    assert np.allclose(sel.regret([1.1, 2.0], [1.0, 2.0]), [0.1, 0.0])


def test_row_length_marks_follow_the_autotuner():
    # lengths 0..99 shuffled: ceil mean 50 (4950 / 100 rounded up), quantiles at floor(q * 99)
    rng = np.random.default_rng(0)
    lens = rng.permutation(100)
    rp = np.concatenate([[0], np.cumsum(lens)])
    m = sel.row_length_marks(rp)
    assert m == {"avg": 50, "q75": 74, "q90": 89, "q99": 98}


def test_tuning_labels():
    base = {"lanes": 8, "nnz_per_wg": 2040, "hyb_width": 0, "scoo_slice_size": 512, "flags": 0}
    marks = {"avg": 5, "q75": 7, "q90": 7, "q99": 30}
    assert sel.tuning_label("csr_vector", base, marks) == "lanes8"
    assert sel.tuning_label("csr_adaptive", dict(base, nnz_per_wg=1020), marks) == "W/2"
    assert sel.tuning_label("hyb", base, marks) == "rule"
    assert sel.tuning_label("hyb", dict(base, hyb_width=7), marks) == "q75"  # first matching mark
    assert sel.tuning_label("scoo", dict(base, flags=4, scoo_slice_size=16384), marks) == "cs16384p0"
    assert sel.tuning_label("scoo", dict(base, flags=4, scoo_slice_size=8192, scoo_pieces=2,
                                         scoo_threads=1024), marks) == "cs8192p2t1024"
    assert sel.tuning_label("scoo", dict(base, flags=2), marks) == "default"
    assert sel.tuning_label("ell", base, marks) == "default"


def _rec(mean_bin_feat, rows_feat, cands):
    f = [0.0] * 64
    f[3], f[0] = mean_bin_feat, rows_feat
    best = min(c[2] for c in cands if c[2] is not None)
    return {"features": f, "cands": cands, "exhaustive_ms": best}


def test_rules_pick_the_label_with_the_lowest_mean_slowdown_per_key():
    # mean bin 2 (log2(1+mean) in [2,3)): lanes4 best; mean bin 5: lanes32 best
    recs = []
    for k in range(4):
        recs.append(_rec(2.5 / 16, 10.2 / 32, [["csr_vector", "lanes4", 1.0, True], ["csr_vector", "lanes32", 1.5, False]]))
        recs.append(_rec(5.5 / 16, 12.2 / 32, [["csr_vector", "lanes4", 2.0, True], ["csr_vector", "lanes32", 1.0, False]]))
    rules = sel.fit_tuning_rules(recs)
    assert rules["csr_vector"]["2,10"][0] == "lanes4"
    assert rules["csr_vector"]["5,12"][0] == "lanes32"
    f = [0.0] * 64
    f[3], f[0] = 5.2 / 16, 20.5 / 32  # unseen row bin: falls back to the mean bin
    assert sel.choose_label(rules, "csr_vector", f, {"lanes4", "lanes32"}) == "lanes32"
    # a label the matrix lacks is skipped
    assert sel.choose_label(rules, "csr_vector", f, {"lanes4"}) == "lanes4"
    probs = np.zeros((2, 8))
    probs[:, sel.FORMAT_ORDER.index("csr_vector")] = 1
    ev = sel.evaluate_with_rules(recs[:2], probs, rules)
    assert ev["mean_regret"] == 0 and ev["within_5pct"] == 1


def test_rules_split_by_column_spread():
    # same mean and row bins: banded matrices (feature 58 small) prefer the
    # small column-sorted slices, spread-out ones the large slices
    recs = []
    for k in range(4):
        a = _rec(3.5 / 16, 20.2 / 32, [["scoo", "cs4096p0", 1.0, True], ["scoo", "cs16384p0", 1.4, False]])
        b = _rec(3.5 / 16, 20.2 / 32, [["scoo", "cs4096p0", 2.0, True], ["scoo", "cs16384p0", 1.0, False]])
        a["features"][58], b["features"][58] = 0.01, 0.8
        recs += [a, b]
    rules = sel.fit_tuning_rules(recs)
    assert rules["scoo"]["3,20,0"][0] == "cs4096p0"
    assert rules["scoo"]["3,20,2"][0] == "cs16384p0"
    f = [0.0] * 64
    f[3], f[0], f[58] = 3.5 / 16, 20.2 / 32, 0.5
    assert sel.choose_label(rules, "scoo", f, {"cs4096p0", "cs16384p0"}) == "cs16384p0"


def test_rules_prefer_plan_sized_pieces_unless_a_fixed_count_wins_clearly():
    # the column-sorted SCOO with its pieces chosen by the plan (p0) ranks
    # ahead of a fixed piece count that is better by less than the margin
    # (a near tie), and behind one that is better by more
    def recs(fixed_ms):
        return [_rec(3.5 / 16, 20.2 / 32, [["scoo", "cs16384p0t1024", 1.0, True],
                                         ["scoo", "cs16384p2t1024", fixed_ms, False]]) for _ in range(4)]
    near = sel.fit_tuning_rules(recs(0.998))
    assert near["scoo"]["3,20"][:2] == ["cs16384p0t1024", "cs16384p2t1024"]
    assert sel.fit_tuning_rules(recs(0.998), self_sized_margin=0.0)["scoo"]["3,20"][0] == "cs16384p2t1024"
    clear = sel.fit_tuning_rules(recs(0.8))  # p0 is 25% slower than p2
    assert clear["scoo"]["3,20"][0] == "cs16384p2t1024"
    assert sel.SELF_SIZED_MARGIN == 0.1
