"""Pins the oracle's Sobol' sampler: its generator matrices and
interval-to-index tables (built from the Joe-Kuo direction numbers) against
the reference's own tables (core/sobolmatrices.rs, through the digests and
entries in tests/golden/sobol_tables.npz), then by the properties that define
the sampler: every sample lands in its pixel, a pixel's samples are a (0,2)-
net, the values are the binary expansions the matrices give; then lower error
than the random sampler at equal spp."""
import hashlib
import os

import numpy as np
import pytest

from rs_illume_amd import scenes as S

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "sobol_tables.npz")


def digest(arrs):
    h = hashlib.sha256()
    for x in arrs:
        h.update(np.ascontiguousarray(x).astype(x.dtype.newbyteorder("<")).tobytes())
    return np.frombuffer(h.digest(), np.uint8)


def test_tables_equal_the_reference(oracle):
    g = np.load(GOLDEN)
    c, vdc, inv = oracle.sobol_tables()
    assert np.array_equal(digest([c.reshape(-1)]), g["matrices_sha256"])
    assert np.array_equal(c[g["sample_dims"]], g["sample_matrices"])
    vdc_rows = [vdc[m, :n] for m, n in enumerate(g["vdc_lengths"])]
    inv_rows = [inv[m, :n] for m, n in enumerate(g["vdc_inv_lengths"])]
    assert [len(r) for r in vdc_rows] == [52 - 2 * m for m in range(1, 26)]
    assert [len(r) for r in inv_rows] == [2 * m for m in range(1, 27)]
    assert np.array_equal(digest(vdc_rows), g["vdc_sha256"])
    assert np.array_equal(digest(inv_rows), g["vdc_inv_sha256"])
    assert np.array_equal(np.concatenate(vdc_rows[:4]), g["vdc_m1_m4"])
    assert np.array_equal(np.concatenate(inv_rows[:4]), g["vdc_inv_m1_m4"])


def test_dimension_0_is_van_der_corput_and_values_follow_the_matrices(oracle):
    c, _, _ = oracle.sobol_tables()
    for a in (1, 2, 3, 6, 12345, 2 ** 40 + 7):
        # columns 32..51 of dimension 0 are zero: only the low 32 bits are reversed
        rev = int("{:032b}".format(a & 0xFFFFFFFF)[::-1], 2)
        assert oracle.sobol_value(a, 0) == min(np.float32(rev) * np.float32(2.0 ** -32), np.float32(1 - 2 ** -24))
        for dim in (1, 5, 700):
            v = 0
            for i in range(52):
                if (a >> i) & 1:
                    v ^= int(c[dim, i])
            assert oracle.sobol_value(a, dim) == np.float32(min(v * 2.0 ** -32, 1 - 2 ** -24))


def test_first_points_of_dimension_pairs_are_nets(oracle):
    """The first 2^k points of dimensions (0, 1) of a Sobol' sequence form a
    (0, k, 2)-net: every 2^i x 2^(k-i) elementary box holds one point."""
    k = 8
    pts = np.array([[oracle.sobol_value(a, 0), oracle.sobol_value(a, 1)] for a in range(2 ** k)])
    for i in range(k + 1):
        cells = set(zip((pts[:, 0] * 2 ** i).astype(int), (pts[:, 1] * 2 ** (k - i)).astype(int)))
        assert len(cells) == 2 ** k, i


@pytest.mark.parametrize("bounds", [(0, 0, 16, 16), (0, 0, 20, 12), (-2, -2, 37, 21)])
def test_every_sample_lands_in_its_pixel(oracle, bounds):
    x0, y0, w, h = bounds
    res = 1 << max(0, (max(w, h) - 1).bit_length())
    for (x, y) in ((x0, y0), (x0 + 1, y0 + 2), (x0 + w - 1, y0 + h - 1), (x0 + w // 2, y0 + 3)):
        for s in (0, 1, 2, 7, 33):
            idx, u = oracle.sobol_sample(bounds, (x, y), s, 2)
            fx = oracle.sobol_value(idx, 0) * res + x0
            fy = oracle.sobol_value(idx, 1) * res + y0
            assert int(np.floor(fx)) == x and int(np.floor(fy)) == y, (x, y, s, fx, fy)
            assert 0 <= u[0] < 1 and 0 <= u[1] < 1
            assert abs(u[0] - (fx - x)) < 1e-6 and abs(u[1] - (fy - y)) < 1e-6
            assert idx >> (2 * (res.bit_length() - 1)) == s  # the sample number is the index's high bits


def test_pixel_samples_are_stratified(oracle):
    """The 2^k samples of one pixel are a (0, k, 2)-net inside the pixel."""
    for (x, y) in ((0, 0), (5, 9)):
        u = np.array([oracle.sobol_sample((0, 0, 16, 16), (x, y), s, 2)[1] for s in range(16)])
        for i in range(5):
            cells = set(zip((u[:, 0] * 2 ** i).astype(int), (u[:, 1] * 2 ** (4 - i)).astype(int)))
            assert len(cells) == 16


def test_sobol_lowers_error_at_equal_spp(oracle):
    desc = S.cornell_box(0)
    sc = oracle.OracleRenderScene(desc.arrays())
    cam = S.cornell_camera(48, 48)
    ref, _ = sc.render_sum(cam, 2048, 5, seed=99)
    ref /= 2048
    errs = {}
    for smp in ("random", "sobol"):
        img, _ = sc.render_sum(cam, 64, 5, seed=1, sampler=smp)
        errs[smp] = S.rel_mse(img / 64, ref)
    assert errs["sobol"] < 0.7 * errs["random"], errs
