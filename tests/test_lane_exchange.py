"""The cooperative kernels' butterfly reductions (group_model.h: Group::sum / Group::max, and the
pivot search of batched.h: lu_rows) run their steps below 16 lanes as DPP row rotations
(row_ror:k, lane i reads lane (i -+ k) mod 16 of its row) instead of xor shuffles (lane i reads
lane i ^ k). The claim that makes this exact: with the offsets taken in decreasing order, the
value a rotation reads equals the value the xor partner holds, so every lane combines the same
operands in the same order and the result is bit-identical to the xor butterfly. Checked here on
the CPU by simulating both exchanges lane by lane over random FP64 data (sums, maxima and the
pivot search's lexicographic (|a|, -position) maximum), for groups of 16, 32 and 64 lanes and
both rotation directions. A plausible mistake -- increasing offsets, or a rotation used before the
larger steps -- makes the two disagree, which the last test shows."""
import numpy as np
import pytest


def partner_xor(i, off):
    return i ^ off


def partner_ror(i, off, sign):
    row = i & ~15
    return row + ((i - row + sign * off) % 16)


def butterfly(vals, G, combine, offsets, rotate_below16, sign=1):
    """One wave of len(vals) lanes in groups of G; returns the per-lane results."""
    x = list(vals)
    for off in offsets:
        new = []
        for i in range(len(x)):
            j = partner_ror(i, off, sign) if (rotate_below16 and off < 16) else partner_xor(i, off)
            assert (j // G) == (i // G), "exchange stays inside the group"
            new.append(combine(x[i], x[j]))
        x = new
    return x


def fsum(a, b):
    return float(np.float64(a) + np.float64(b))


def fmax_(a, b):
    return max(a, b)


def pivot_max(a, b):
    # (value, position, source) keyed by (value, -position), as lu_rows
    (va, pa, sa), (vb, pb, sb) = a, b
    return b if (vb > va or (vb == va and pb < pa)) else a


@pytest.mark.parametrize("G", [16, 32, 64])
@pytest.mark.parametrize("sign", [1, -1])
def test_rotation_butterfly_is_bit_identical_to_xor(G, sign):
    rng = np.random.default_rng(G + (sign > 0))
    offsets = [o for o in (32, 16, 8, 4, 2, 1) if o < G]
    for trial in range(20):
        v = (rng.standard_normal(64) * 10.0 ** rng.integers(-8, 8, 64)).tolist()
        for comb in (fsum, fmax_):
            ref = butterfly(v, G, comb, offsets, rotate_below16=False)
            got = butterfly(v, G, comb, offsets, rotate_below16=True, sign=sign)
            assert [np.float64(a).tobytes() for a in ref] == [np.float64(a).tobytes() for a in got]
            # every lane of a group holds the same bits
            for g in range(64 // G):
                assert len({np.float64(a).tobytes() for a in ref[g * G:(g + 1) * G]}) == 1
        # pivot search: |a| with ties, distinct positions
        mag = np.abs(rng.integers(0, 4, 64)).astype(float).tolist()
        pos = [int(p) for p in rng.permutation(64)]
        keys = [(mag[i], pos[i], i) for i in range(64)]
        ref = butterfly(keys, G, pivot_max, offsets, rotate_below16=False)
        got = butterfly(keys, G, pivot_max, offsets, rotate_below16=True, sign=sign)
        assert ref == got


def test_rotation_needs_decreasing_offsets():
    """With increasing offsets a rotation reads a lane that has not yet been made equal to the xor
    partner, and the sums differ from the xor butterfly's in the last bits."""
    rng = np.random.default_rng(5)
    differs = 0
    for trial in range(50):
        v = (rng.standard_normal(64) * 10.0 ** rng.integers(-8, 8, 64)).tolist()
        ref = butterfly(v, 32, fsum, [1, 2, 4, 8, 16], rotate_below16=False)
        got = butterfly(v, 32, fsum, [1, 2, 4, 8, 16], rotate_below16=True)
        differs += ref != got
    assert differs > 0
