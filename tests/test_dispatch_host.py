"""Host-side dispatcher functions that need no device: the segment length
bins, the configuration table's feasibility rules and the model rows."""
import numpy as np


def test_seg_bins_brute_force(clops):
    """bin(len) = 4 floor(log2 len) + the two bits after the leading one
    (lengths 0 and 1 in bin 0, capped at 63); bin_length(b) is the least
    length in bin b, 0 for bins no length reaches.  Checked against the
    definition by enumeration over 0..2^17."""
    d = clops.Dispatch
    lens = np.arange(0, (1 << 17) + 1)
    got = np.array([d.seg_bin(int(x)) for x in lens])
    want = np.zeros(lens.size, dtype=np.int64)
    for i, x in enumerate(lens):
        if x >= 2:
            lg = int(x).bit_length() - 1
            frac = int(x) / (1 << lg)  # in [1, 2)
            want[i] = min(4 * lg + int((frac - 1) * 4), 63)
    assert np.array_equal(got, want)
    first = {}
    for x, b in zip(lens, got):
        first.setdefault(int(b), int(x) if x else 1)
    for b in range(clops.SEG_BINS):
        assert d.seg_bin_length(b) == first.get(b, 0), b
    assert d.seg_bin(1 << 40) == 63
    # Bins are monotone in the length.
    assert np.all(np.diff(got) >= 0)


def test_config_table_rules(clops):
    d = clops.Dispatch
    names = [d.config_name(c) for c in range(clops.N_CONFIGS)]
    assert len(set(names)) == clops.N_CONFIGS
    seg = [c for c in range(clops.N_CONFIGS) if d.feasible(c, clops.OP_SEGSORT, 1000, 4)]
    assert seg == list(range(12, 24))
    # Tile 32 holds 32-bit keys only.
    assert [c for c in seg if not d.feasible(c, clops.OP_SEGSORT, 1000, 8)] == [20, 21, 22, 23]
    # Digit widths are sort configurations of any size.
    for c in (9, 10, 11):
        assert d.feasible(c, clops.OP_SORT, 1 << 28, 8) and not d.feasible(c, clops.OP_SCAN, 10, 4)
    fx = d.fixed(clops.OP_SEGSORT, 1000, 4)
    assert fx in seg and names[fx].startswith("segmented tile 256x")
    assert names[12] == "segmented tile 256x8, 2/CU" and names[23] == "segmented tile 256x32, 5/CU"
    for op in (clops.OP_SORT, clops.OP_SCAN, clops.OP_SEGSORT):
        for w in (4, 8):
            assert d.feasible(d.fixed(op, 5000, w), op, 5000, w)


def test_segment_rows(clops):
    d = clops.Dispatch
    r = d.segment_row(1023, 8, 17, 2)  # tile 16, 3 blocks per CU
    assert r.size == clops.SEG_INPUTS
    assert np.allclose(r, [np.log2(1024) / 16, 1.0, 0.5, 0.75, 0, 0, 1])
    f = np.arange(clops.N_FEATURES, dtype=np.float32)
    row = d.model_row(f, 23)
    assert row.size == clops.MODEL_INPUTS
    assert np.array_equal(row[:clops.N_FEATURES], f)
    assert row[clops.N_FEATURES + 23] == 1 and row[clops.N_FEATURES:].sum() == 1
