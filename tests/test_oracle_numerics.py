"""Pins the oracle's numerics (core/pbrt.rs, core/rng.rs, core/geometry/ray.rs)
against independent definitions: numpy's nextafter, the published PCG32
reference algorithm and its known-answer vector, and closed forms."""
import numpy as np
import pytest

M64 = (1 << 64) - 1


def pcg32_reference(initstate, initseq):
    """pcg32_srandom_r / pcg32_random_r as published at pcg-random.org."""
    st = [0, ((initseq << 1) | 1) & M64]

    def nxt():
        old = st[0]
        st[0] = (old * 6364136223846793005 + st[1]) & M64
        xs = (((old >> 18) ^ old) >> 27) & 0xFFFFFFFF
        rot = old >> 59
        return ((xs >> rot) | (xs << ((-rot) & 31))) & 0xFFFFFFFF

    nxt()
    st[0] = (st[0] + initstate) & M64
    nxt()
    return nxt


def test_pcg32_reference_known_answer():
    # pcg32-demo output for pcg32_srandom_r(42, 54)
    g = pcg32_reference(42, 54)
    assert [g() for _ in range(6)] == [0xA15C02B7, 0x7B47F409, 0xBA1D3330, 0x83D2F293, 0xBFA4784B, 0xCBED606E]


@pytest.mark.parametrize("seq", [0, 1, 7, 12345, 2 ** 40 + 3])
def test_oracle_pcg32_equals_reference(oracle, seq):
    # pbrt's RNG(seq) == pcg32_srandom_r(PCG32_DEFAULT_STATE, seq) (rng.rs:30-36)
    g = pcg32_reference(0x853C49E6748FEA9B, seq)
    ref = [g() for _ in range(64)]
    assert list(oracle.pcg32(seq, 64)) == ref


@pytest.mark.parametrize("skip", [1, 2, 3, 17, 1000, 65537])
def test_pcg32_advance_equals_stepping(oracle, skip):
    full = oracle.pcg32(5, skip + 8)
    adv = oracle.pcg32(5, 8, skip=skip)
    assert np.array_equal(full[skip:], adv)


def test_uniform_float_range(oracle):
    u = oracle.pcg32_float(3, 100000)
    assert u.min() >= 0 and u.max() < 1
    ints = oracle.pcg32(3, 100000)
    expect = np.minimum(ints.astype(np.float32) * np.float32(2.0 ** -32), np.float32(float.fromhex("0x1.fffffep-1")))
    assert np.array_equal(u, expect.astype(np.float32))
    assert abs(u.mean() - 0.5) < 0.005


def test_gamma_known_values(oracle):
    eps = np.float32(2.0 ** -24)
    for n in (1, 2, 3, 5, 7):
        expect = np.float32(np.float32(n) * eps) / np.float32(np.float32(1) - np.float32(n) * eps)
        assert np.float32(oracle.gamma(n)) == expect
    assert np.isclose(oracle.gamma(3), 3 * 2.0 ** -24, rtol=1e-6)


def test_next_float_matches_nextafter(oracle):
    rng = np.random.default_rng(0)
    vals = np.concatenate([rng.normal(size=2000) * 10 ** rng.uniform(-30, 30, 2000),
                           [0.0, 1.0, -1.0, 1e-45, -1e-45, 3.4e38, -3.4e38]]).astype(np.float32)
    for v in vals:
        up = np.float32(oracle.next_float_up(float(v)))
        dn = np.float32(oracle.next_float_down(float(v)))
        assert up == np.nextafter(v, np.float32(np.inf)), v
        assert dn == np.nextafter(v, np.float32(-np.inf)), v
    assert oracle.next_float_up(np.inf) == np.inf
    assert oracle.next_float_down(-np.inf) == -np.inf
    # -0 steps up to the smallest denormal, +0 steps down to its negative
    assert np.float32(oracle.next_float_up(-0.0)) == np.float32(1e-45)
    assert np.float32(oracle.next_float_down(0.0)) == np.float32(-1e-45)


def test_offset_ray_origin_moves_to_the_side_of_w(oracle):
    rng = np.random.default_rng(1)
    for _ in range(500):
        p = rng.normal(size=3).astype(np.float32)
        n = rng.normal(size=3)
        n = (n / np.linalg.norm(n)).astype(np.float32)
        pe = np.abs(rng.normal(size=3) * 1e-6).astype(np.float32)
        w = rng.normal(size=3).astype(np.float32)
        po = oracle.offset_ray_origin(p, pe, n, w)
        side = np.sign(np.dot(w, n))
        d = float(np.dot(np.abs(n), pe))
        # moved by at least the error-bound distance along +-n, and rounded away from p
        assert side * np.dot((po.astype(np.float64) - p), n) >= d * 0.999
        off = n * d * (1 if side >= 0 else -1)
        for i in range(3):
            if off[i] > 0:
                assert po[i] > p[i]
            elif off[i] < 0:
                assert po[i] < p[i]
