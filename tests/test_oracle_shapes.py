"""Pins the oracle's triangle (shapes/triangle.rs) and sphere (shapes/sphere.rs,
core/efloat.rs) tests against known answers, f64 closed forms and the
watertightness and error-bound properties the reference's algorithm guarantees."""
from fractions import Fraction

import numpy as np
import pytest

from conftest import ulp_diff


def ray(o, d, t_max=np.inf):
    import pyoracle
    return pyoracle.make_ray(o, d, t_max)


def moller_f64(p0, p1, p2, o, d):
    """Textbook ray-triangle in f64: (t, b0, b1, b2) or None."""
    p0, p1, p2, o, d = (np.asarray(x, np.float64) for x in (p0, p1, p2, o, d))
    e1, e2 = p1 - p0, p2 - p0
    h = np.cross(d, e2)
    a = np.dot(e1, h)
    if abs(a) < 1e-300:
        return None
    f = 1.0 / a
    s = o - p0
    u = f * np.dot(s, h)
    q = np.cross(s, e1)
    v = f * np.dot(d, q)
    t = f * np.dot(e2, q)
    return t, 1 - u - v, u, v


def test_triangle_known_answer(oracle):
    p0, p1, p2 = np.array([0, 0, 0]), np.array([1, 0, 0]), np.array([0, 1, 0])
    ok, out = oracle.triangle_test(p0, p1, p2, ray([0.25, 0.25, -1], [0, 0, 1]))
    assert ok
    t, b0, b1, b2 = out
    assert t == 1.0 and b0 == 0.5 and b1 == 0.25 and b2 == 0.25
    # from behind, other winding, scaled direction
    ok, out = oracle.triangle_test(p0, p2, p1, ray([0.25, 0.25, 2], [0, 0, -4]))
    assert ok and out[0] == 0.5


def test_triangle_t_range_and_self_hit(oracle):
    p0, p1, p2 = np.array([0, 0, 0]), np.array([1, 0, 0]), np.array([0, 1, 0])
    assert not oracle.triangle_test(p0, p1, p2, ray([0.25, 0.25, -1], [0, 0, 1], t_max=0.5))[0]
    assert oracle.triangle_test(p0, p1, p2, ray([0.25, 0.25, -1], [0, 0, 1], t_max=1.0))[0]
    assert not oracle.triangle_test(p0, p1, p2, ray([0.25, 0.25, -1], [0, 0, -1]))[0]   # behind the origin
    # origin on the surface: t ~ 0 is inside the delta_t error bound -> rejected
    assert not oracle.triangle_test(p0, p1, p2, ray([0.25, 0.25, 0], [0.3, 0.2, 1]))[0]
    # degenerate (collinear) triangle never reports a hit
    assert not oracle.triangle_test(p0, p1, np.array([2, 0, 0]), ray([0.5, 0, -1], [0, 0, 1]))[0]


def test_triangle_against_f64_closed_form(oracle):
    rng = np.random.default_rng(3)
    n_hit = 0
    for _ in range(3000):
        p = rng.normal(size=(3, 3)).astype(np.float32)
        o = (rng.normal(size=3) * 3).astype(np.float32)
        w = rng.uniform(-0.3, 1.0, size=3)
        tgt = (w / w.sum()) @ p.astype(np.float64)   # sometimes outside the triangle
        d = (tgt - o).astype(np.float32)
        ref = moller_f64(p[0], p[1], p[2], o, d)
        ok, out = oracle.triangle_test(p[0], p[1], p[2], ray(o, d))
        if ref is None:
            continue
        t, b0, b1, b2 = ref
        margin = min(b0, b1, b2)
        if margin > 1e-4 and t > 1e-3:
            assert ok, (p, o, d)
            n_hit += 1
            # f32 error grows as the ray grazes the plane: scale by 1/|cos|
            nrm = np.cross(p[1].astype(np.float64) - p[0], p[2].astype(np.float64) - p[0])
            cos = abs(nrm @ d) / (np.linalg.norm(nrm) * np.linalg.norm(d.astype(np.float64)))
            assert abs(out[0] - t) <= (2e-6 * abs(t) + 1e-6) / max(cos, 1e-3)
            assert np.allclose(out[1:], [b0, b1, b2], atol=1e-4 / max(cos, 1e-2))
        elif margin < -1e-4 or t < -1e-3:
            assert not ok, (p, o, d)
    assert n_hit > 200


def test_triangle_watertight_on_shared_edges_and_vertices(oracle):
    """Rays through a point of an edge shared by two triangles never slip
    through the crack (the point lies inside their union when seen along d)."""
    rng = np.random.default_rng(4)
    n_checked = 0
    for _ in range(1000):
        a, b, c, e = rng.normal(size=(4, 3)).astype(np.float32)
        s = np.float32(rng.random())
        tgt = (a * (1 - s) + b * s).astype(np.float32)
        o = (tgt + rng.normal(size=3) * 3).astype(np.float32)
        d = (tgt - o).astype(np.float32)
        h1 = oracle.triangle_test(a, b, c, ray(o, d))[0]
        h2 = oracle.triangle_test(b, a, e, ray(o, d))[0]
        # c and e on opposite sides of the edge: the edge point is inside the union
        nrm = np.cross(b - a, d)
        if np.dot(c - a, nrm) * np.dot(e - a, nrm) < 0 and abs(np.dot(c - a, nrm)) > 1e-3 \
                and abs(np.dot(e - a, nrm)) > 1e-3:
            assert h1 or h2
            n_checked += 1
    assert n_checked > 200


def test_triangle_f64_edge_retry_is_used(oracle):
    """A ray through a vertex makes an f32 edge function exactly 0; the f64
    retry (triangle.rs:192-204) must still produce a consistent hit."""
    p0, p1, p2 = np.array([0, 0, 0], np.float32), np.array([1, 0, 0], np.float32), np.array([0, 1, 0], np.float32)
    ok, out = oracle.triangle_test(p0, p1, p2, ray([0, 0, -1], [0, 0, 1]))
    assert ok and out[0] == 1.0 and out[1] == 1.0


def test_sphere_known_answers(oracle):
    # the object-space origin is first advanced along d past its error bound
    # (transform.rs:396-402), so t is 4 minus that offset (gamma(3)*5)
    ok, out = oracle.sphere_intersect([0, 0, 0], 1.0, ray([0, 0, -5], [0, 0, 1]))
    assert ok and 4.0 - 2e-6 < out[0] <= 4.0
    assert np.allclose(out[4:7], [0, 0, -1], atol=1e-6)   # outward normal at the entry point
    ok, out = oracle.sphere_intersect([0, 0, 0], 1.0, ray([0, 0, 0], [1, 0, 0]))  # from inside
    assert ok and out[0] == 1.0   # zero origin: no offset
    assert not oracle.sphere_intersect([0, 0, 0], 1.0, ray([0, 2, -5], [0, 0, 1]))[0]  # clear miss
    assert not oracle.sphere_intersect([0, 0, 0], 1.0, ray([0, 0, -5], [0, 0, -1]))[0]  # behind
    assert not oracle.sphere_intersect([0, 0, 0], 1.0, ray([0, 0, -5], [0, 0, 1], t_max=3.9))[0]


def test_sphere_against_f64_roots(oracle):
    rng = np.random.default_rng(5)
    checked = 0
    for _ in range(3000):
        c = rng.normal(size=3).astype(np.float32)
        r = np.float32(np.exp(rng.normal() - 1))
        o = (c + rng.normal(size=3) * 3 * r).astype(np.float32)
        d = (c + rng.normal(size=3) * 0.7 * r - o).astype(np.float32)   # mostly aimed at the sphere
        oc = o.astype(np.float64) - c
        dd = d.astype(np.float64)
        A, B, C = dd @ dd, 2 * dd @ oc, oc @ oc - float(r) ** 2
        disc = B * B - 4 * A * C
        ok, out = oracle.sphere_intersect(c, r, ray(o, d))
        if disc < 0:
            if disc < -1e-6 * B * B:
                assert not ok
            continue
        roots = sorted([(-B - np.sqrt(disc)) / (2 * A), (-B + np.sqrt(disc)) / (2 * A)])
        want = roots[0] if roots[0] > 1e-4 else (roots[1] if roots[1] > 1e-4 else None)
        if want is None or disc < 1e-6 * B * B:
            continue
        assert ok
        # f32 error of the root is amplified by |B|/sqrt(disc) for grazing rays
        # and the object-space origin carries gamma(3)*(|o|+|c|) of error, in units of |d|
        amp = abs(B) / np.sqrt(disc)
        o_err = 4 * oracle.gamma(3) * (np.abs(o).sum() + np.abs(c).sum()) / np.sqrt(A)
        assert abs(out[0] - want) <= (4e-6 + 2e-7 * amp) * abs(want) + o_err + 1e-6
        checked += 1
    assert checked > 500


def test_efloat_quadratic_interval_contains_exact_roots(oracle):
    rng = np.random.default_rng(6)
    for _ in range(2000):
        a, b, c = (np.float32(x) for x in rng.normal(size=3) * np.exp(rng.normal(size=3)))
        ok, out = oracle.quadratic(a, 0, b, 0, c, 0)
        A, B, C = Fraction(float(a)), Fraction(float(b)), Fraction(float(c))
        disc = B * B - 4 * A * C
        if disc < 0:
            assert not ok
            continue
        assert ok
        sq = Fraction(np.sqrt(float(disc)))
        r = sorted([(-B - sq) / (2 * A), (-B + sq) / (2 * A)])
        # interval bounds bracket the (near-exact) roots; f64 sqrt error ~1e-16 relative
        for k, root in enumerate(r):
            v, lo, hi = out[3 * k:3 * k + 3]
            tol = abs(float(root)) * 1e-12 + 1e-30
            assert lo - tol <= float(root) <= hi + tol, (a, b, c, out, float(root))
            assert lo <= v <= hi
