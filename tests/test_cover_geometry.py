"""The hidden-candidate rule of the gathers (cap_covered / arc_covered in
sesmesh_amd/csrc/dev_common.hpp) and the blocker-pruning rule (cap_reaches /
arc_reaches), restated in numpy and checked against brute force.

A class S or C candidate is dropped from a sub-block's lists when one of its
blockers contains, with a margin, the projection of every point of the
sub-block's ball onto it. The meshes stay byte-identical only if that claim
is never wrong, so the restated tests are checked on random configurations
against the farthest projection found by sampling the ball (interior and
boundary), and the division-free cap test against its trigonometric form
cos(min(theta + alpha, pi)). A blocker is left out of a candidate's bitset
when no projection of the ball comes within t of it; that claim is checked
against the nearest sampled projection. The reference has neither rule: its
lane kernel tests every candidate and every blocker of its row
(field/kernel.rs:837-952).
"""
import numpy as np

MARGIN = 4e-3


def cap_covered(b, rb, c, R, k, lim):
    v = b - c
    h2 = v @ v - rb * rb
    if not h2 > 0:
        return False
    e = k - c
    d2, dot = e @ e, v @ e
    if dot <= 0 and dot * dot >= d2 * h2:
        m = lim - R
        return m > 0 and d2 < m * m
    cr = np.linalg.norm(np.cross(v, e))
    return (R * R + d2 - lim * lim) * (h2 + rb * rb) < 2 * R * (dot * np.sqrt(h2) - cr * rb)


def cap_covered_trig(b, rb, c, R, k, lim):
    v = b - c
    D = np.linalg.norm(v)
    if not D > rb:
        return False
    e = k - c
    d = np.linalg.norm(e)
    sa = rb / D
    ca = np.sqrt(max(0.0, 1 - sa * sa))
    ct = np.clip(v @ e / (D * d), -1, 1) if d > 0 else 1.0
    cphi = -1.0 if ct < -ca else ct * ca - np.sqrt(max(0.0, 1 - ct * ct)) * sa
    return R * R + d * d - 2 * R * d * cphi < lim * lim


def arc_covered(b, rb, m, u, r, k, lim):
    y = b - m
    w = y - u * (y @ u)
    W2 = w @ w
    h2 = W2 - rb * rb
    f = k - m
    p = f - u * (f @ u)
    L2, wp = p @ p, w @ p
    A0 = r * r + f @ f - lim * lim
    if not h2 > 0 or (wp <= 0 and wp * wp >= L2 * h2):
        return A0 < 0 and A0 * A0 > 4 * r * r * L2
    cr = np.linalg.norm(np.cross(w, p))
    return A0 * W2 < 2 * r * (wp * np.sqrt(h2) - cr * rb)


def cap_reaches(b, rb, c, R, k, t):
    v = b - c
    D2 = v @ v
    h2 = D2 - rb * rb
    e = k - c
    d2, dot = e @ e, v @ e
    if not h2 > 0 or (dot >= 0 and dot * dot >= d2 * h2):
        lo, hi = R - t, R + t
        return d2 < hi * hi and not (lo > 0 and d2 <= lo * lo)
    cr = np.linalg.norm(np.cross(v, e))
    return (R * R + d2 - t * t) * D2 < 2 * R * (dot * np.sqrt(h2) + cr * rb)


def arc_reaches(b, rb, m, u, r, k, t):
    y = b - m
    w = y - u * (y @ u)
    W2 = w @ w
    h2 = W2 - rb * rb
    f = k - m
    p = f - u * (f @ u)
    L2, wp = p @ p, w @ p
    A0 = r * r + f @ f - t * t
    if not h2 > 0 or (wp >= 0 and wp * wp >= L2 * h2):
        return A0 < 0 or A0 * A0 < 4 * r * r * L2
    cr = np.linalg.norm(np.cross(w, p))
    return A0 * W2 < 2 * r * (wp * np.sqrt(h2) + cr * rb)


def unit(v):
    return v / np.linalg.norm(v)


def ball_points(rng, b, rb, n):
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1)[:, None]
    rad = rb * rng.uniform(0, 1, n) ** (1 / 3)
    rad[: n // 2] = rb
    return b + d * rad[:, None]


def test_cap_rule_is_sound_and_matches_its_trigonometric_form():
    rng = np.random.default_rng(11)
    covered = 0
    for _ in range(1500):
        c = rng.normal(size=3) * 3
        R = rng.uniform(1.5, 3.5)
        b = c + unit(rng.normal(size=3)) * rng.uniform(0.5, 6)
        rb = rng.uniform(0.2, 3)
        k = c + unit(rng.normal(size=3)) * rng.uniform(0, 5)
        lim = rng.uniform(1.5, 4.5) - MARGIN
        got = cap_covered(b, rb, c, R, k, lim)
        assert got == cap_covered_trig(b, rb, c, R, k, lim)
        if got:
            covered += 1
            x = ball_points(rng, b, rb, 2000)
            q = c + (x - c) / np.linalg.norm(x - c, axis=1)[:, None] * R
            assert np.linalg.norm(q - k, axis=1).max() < lim
    assert covered > 200  # the rule is not vacuous


def test_arc_rule_is_sound():
    rng = np.random.default_rng(12)
    covered = 0
    for _ in range(1500):
        m = rng.normal(size=3) * 3
        u = unit(rng.normal(size=3))
        r = rng.uniform(0.5, 3)
        e1 = unit(np.cross(u, rng.normal(size=3)))
        ang = rng.uniform(0, 2 * np.pi)
        inplane = np.cos(ang) * e1 + np.sin(ang) * np.cross(u, e1)
        b = m + inplane * rng.uniform(0, 6) + u * rng.normal() * 2
        rb = rng.uniform(0.2, 3)
        k = m + rng.normal(size=3) * 2.5
        lim = rng.uniform(1.5, 5) - MARGIN
        if arc_covered(b, rb, m, u, r, k, lim):
            covered += 1
            x = ball_points(rng, b, rb, 2000)
            y = x - m
            w = y - np.outer(y @ u, u)
            q = m + w / np.linalg.norm(w, axis=1)[:, None] * r
            # A point on the axis projects along e1; only a ball that reaches
            # the axis holds one.
            yb = b - m
            if np.linalg.norm(yb - u * (yb @ u)) <= rb:
                q = np.vstack([q, m + e1 * r])
            assert np.linalg.norm(q - k, axis=1).max() < lim
    assert covered > 200


def test_full_circle_and_far_side_cases():
    # A ball around the axis sees the whole circle: covered only if the
    # blocker holds every circle point.
    m, u, r = np.zeros(3), np.array([0.0, 0.0, 1.0]), 1.0
    b = np.array([0.0, 0.0, 0.5])
    assert arc_covered(b, 0.3, m, u, r, np.array([0.0, 0.0, 0.2]), 1.2)
    assert not arc_covered(b, 0.3, m, u, r, np.array([0.5, 0.0, 0.2]), 1.2)
    # A ball around the sphere's centre projects onto every direction.
    assert not cap_covered(np.zeros(3), 0.5, np.array([0.2, 0.0, 0.0]), 2.0, np.zeros(3), 10.0)
    # The cap reaching past the far side: the farthest point is at R + d.
    c, R = np.zeros(3), 1.0
    b = np.array([2.0, 0.0, 0.0])
    k = np.array([-0.5, 0.0, 0.0])
    assert cap_covered(b, 1.9, c, R, k, 1.6) and not cap_covered(b, 1.9, c, R, k, 1.4)


def test_reach_rules_never_drop_a_blocker_that_can_block():
    rng = np.random.default_rng(13)
    dropped = 0
    for _ in range(1500):
        c = rng.normal(size=3) * 3
        R = rng.uniform(1.5, 3.5)
        b = c + unit(rng.normal(size=3)) * rng.uniform(0.3, 6)
        rb = rng.uniform(0.2, 3)
        k = c + unit(rng.normal(size=3)) * rng.uniform(0, 7)
        t = rng.uniform(1.0, 4.0)
        if not cap_reaches(b, rb, c, R, k, t):
            dropped += 1
            x = ball_points(rng, b, rb, 2000)
            q = c + (x - c) / np.linalg.norm(x - c, axis=1)[:, None] * R
            assert np.linalg.norm(q - k, axis=1).min() >= t
    assert dropped > 200
    dropped = 0
    for _ in range(1500):
        m = rng.normal(size=3) * 3
        u = unit(rng.normal(size=3))
        r = rng.uniform(0.5, 3)
        e1 = unit(np.cross(u, rng.normal(size=3)))
        ang = rng.uniform(0, 2 * np.pi)
        inplane = np.cos(ang) * e1 + np.sin(ang) * np.cross(u, e1)
        b = m + inplane * rng.uniform(0, 6) + u * rng.normal() * 2
        rb = rng.uniform(0.2, 3)
        k = m + rng.normal(size=3) * 3
        t = rng.uniform(1.0, 4.0)
        if not arc_reaches(b, rb, m, u, r, k, t):
            dropped += 1
            x = ball_points(rng, b, rb, 2000)
            y = x - m
            w = y - np.outer(y @ u, u)
            q = m + w / np.linalg.norm(w, axis=1)[:, None] * r
            yb = b - m
            if np.linalg.norm(yb - u * (yb @ u)) <= rb:
                q = np.vstack([q, m + e1 * r])
            assert np.linalg.norm(q - k, axis=1).min() >= t
    assert dropped > 200
