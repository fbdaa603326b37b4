"""The algorithmic op counts behind the bench's roofline, pinned by hand.

The roofline's numerator is the oracle's tally of the reference
algorithm's arithmetic: one op per IEEE add, sub, mul, div, sqrt, min or
compare (abs not counted), along the sequential code path of the f32 lane
kernel (crates/sesmesh/src/field/kernel.rs:837-952) and of the f64
eval_local (crates/sesmesh/src/field/point.rs:422-489, 686-738). Here the
tally at single points of one- and two-sphere sets is compared with counts
made by hand from those code paths, for each branch:

f32 lane: inside test 9 per inside-group sphere (3 sub, 3 mul, 2 add, cmp);
S candidate 13 (3 sub, 3 mul, 2 add, sqrt, sub, mul, 2 cmp), +7 when
projected (div, 3 mul, 3 add), +9 per blocker tested; C candidate 25, +7
when projected, +9 per blocker; V 9; +2 for the result (sqrt, sub).
f64 eval_local: inside test 10 per sphere until the first inside (3 sub,
3 mul, 2 add, sqrt, cmp); try_sphere 12, +7 past the distance test, +11 per
blocker; try_circle 27, +10 past the distance test (e1 fallback 4), +11 per
blocker; try_vertex 10.
"""
import numpy as np

import oracle

PROBE, SPACING = 1.4, 0.5  # band = sqrt(3) 0.5 1.01 = 0.8747; reach = probe + band + rho


def counts(xyzr, point):
    return oracle.Surface(np.array(xyzr, float), PROBE, SPACING, 10, threads=1, stage=1).count_point(point, 0.01)


def test_single_sphere_op_counts():
    one = [[0.0, 0.0, 0.0, 1.7]]  # expanded radius R = 3.1
    # Beyond R + rho: no sphere in the ball's inside group, nothing evaluated.
    assert counts(one, (4.0, 0, 0)) == (0, 0)
    # In the inside group (d < R + rho) but outside R: the inside test only.
    assert counts(one, (3.105, 0, 0)) == (9, 10)
    # Inside R near the surface: inside test, the sphere candidate and its
    # projection, the result: 9 + 13 + 7 + 2; f64: 10 + 12 + 7.
    assert counts(one, (3.05, 0, 0)) == (31, 29)
    assert counts(one, (2.0, 0, 0)) == (31, 29)
    # Deep inside, |d - R| beyond the reach: the sphere is no candidate.
    assert counts(one, (0.5, 0, 0)) == (11, 10)
    assert counts(one, (0.0, 0, 0)) == (11, 10)


def test_two_sphere_op_counts():
    two = [[0.0, 0.0, 0.0, 1.7], [3.0, 0.0, 0.0, 1.7]]
    # Inside A only; candidates A (lb 0.2), B (lb 1.07), circle AB (1.51).
    # f32: 9 + A (13 + 7 + 9 for blocker B, free) + B (13, lb^2 above the
    # best) + circle (25, above the best) + 2 = 78.
    # f64: 10 + A (12 + 7 + 11) + B (12) + circle (27) = 79.
    assert counts(two, (0.0, 2.9, 0)) == (78, 79)
    # The saddle: inside both (f32 tests both, 18; f64 stops at A, 10); each
    # sphere projection is blocked by the other (13 + 7 + 9 each; f64
    # 12 + 7 + 11 each); the circle wins (25 + 7, no blockers; f64 27 + 10);
    # + 2.
    assert counts(two, (1.5, 2.6, 0)) == (18 + 29 + 29 + 32 + 2, 10 + 30 + 30 + 37)


def test_mean_counts_are_weighted_branch_counts():
    """Over a brick range the tally is the sum of per-sample counts: more
    bricks never lower it, and evaluated samples and evaluations add up."""
    one = np.array([[0.0, 0.0, 0.0, 1.7]])
    s = oracle.Surface(one, PROBE, SPACING, 10, threads=1, stage=2)
    a = s.count_ops(0, 1, 1)
    b = s.count_ops(0, s.n_bricks, 1)
    assert 0 < a["f32_ops"] <= b["f32_ops"] and 0 < a["samples"] <= b["samples"]
    # every evaluated sample costs at least the inside test of its sphere or nothing,
    # at most the single-sphere maximum of 31
    assert b["f32_ops"] <= 31 * b["samples"] and b["f64_ops"] <= 29 * b["evals"]
