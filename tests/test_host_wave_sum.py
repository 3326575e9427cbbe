"""The wave reduction's lane-exchange pattern (gpx_wave_sum in geepee_amd/csrc/gpx_common.h),
simulated on the host from the lane-swap widths and DPP control words written in the header:
every step pairs lanes symmetrically, every lane ends with each of the 64 lanes' values exactly
once, and every lane's sum has the same addition tree up to the order of the two operands of
each addition (so, IEEE addition being commutative, all lanes hold the same bits)."""
import os
import re

import numpy as np

HDR = os.path.join(os.path.dirname(os.path.abspath(__file__)), '..', 'geepee_amd', 'csrc', 'gpx_common.h')


def _wave_sum_steps():
    src = open(HDR).read()
    body = src[src.index('double gpx_wave_sum(double v)'):]
    body = body[:body.index('}')]
    return [(k, int(a, 0)) for k, a in re.findall(r'gpx_(swap_add|dpp_add)<(0x[0-9A-Fa-f]+|\d+)>', body)]


def _partner(kind, arg, lane):
    if kind == 'swap_add':
        # permlane32/16 swaps pair lane l with l ^ 32 / l ^ 16
        assert arg in (16, 32)
        return lane ^ arg
    r, base = lane % 16, lane - lane % 16
    if arg == 0x140:  # row_mirror
        return base + 15 - r
    if arg == 0x141:  # row_half_mirror
        return base + (7 - r if r < 8 else 23 - r)
    if arg < 0x100:  # quad_perm [s0, s1, s2, s3]: lane 4k + j reads lane 4k + s_j
        sel = [(arg >> (2 * j)) & 3 for j in range(4)]
        return lane - lane % 4 + sel[lane % 4]
    raise AssertionError('unexpected DPP control word %#x' % arg)


def test_wave_sum_steps_pair_lanes_and_cover_the_wave_once():
    steps = _wave_sum_steps()
    assert len(steps) == 6
    cnt = np.eye(64, dtype=np.int64)  # cnt[l, s]: how often lane s's value is in lane l's sum
    for kind, arg in steps:
        p = [_partner(kind, arg, l) for l in range(64)]
        assert all(p[p[l]] == l and p[l] != l for l in range(64)), (kind, hex(arg))
        cnt = cnt + cnt[p]
    assert (cnt == 1).all()


def test_wave_sum_gives_every_lane_the_same_addition_tree():
    val = [('lane', l) for l in range(64)]
    for kind, arg in _wave_sum_steps():
        val = [frozenset((val[l], val[_partner(kind, arg, l)])) for l in range(64)]
    assert len(set(val)) == 1


def test_a_wrong_control_word_is_caught():
    # row_half_mirror written as a second row_mirror pairs the same lanes twice
    steps = [(k, 0x140 if (k, a) == ('dpp_add', 0x141) else a) for k, a in _wave_sum_steps()]
    cnt = np.eye(64, dtype=np.int64)
    for kind, arg in steps:
        cnt = cnt + cnt[[_partner(kind, arg, l) for l in range(64)]]
    assert not (cnt == 1).all()
