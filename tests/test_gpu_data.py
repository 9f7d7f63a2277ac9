"""Dataset kernels on the GPU against the oracle: decode and per-SNP moments
bit-exact, band X^T X exact (int8 MFMA), the X^T r scan within its stated
fixed-point bound, synthetic genotypes byte-identical to the oracle's
generator."""
import numpy as np
import pytest

import oracle
from fixtures import plinktest, small_modelspace, pack_genotypes

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def bmg():
    import bmagwa_amd
    bmagwa_amd.lib()
    return bmagwa_amd


def random_body(rng, n, p, with_missing=False):
    G = rng.integers(0, 3, size=(p, n))
    if with_missing:
        G[rng.random(size=(p, n)) < 0.02] = -1
    return G, pack_genotypes(G)


@pytest.mark.parametrize("n,p", [(5, 10), (30, 7), (37, 23), (1000, 64), (4099, 50)])
def test_decode_bit_exact(bmg, n, p):
    rng = np.random.default_rng(n * 7 + p)
    G, body = random_body(rng, n, p, with_missing=True)
    y = rng.normal(size=n)
    ctx = bmg.Ctx(n, p, body, y, types=("A",))
    for coding in range(4):
        got = ctx.decode(np.arange(p), coding)
        for j in range(p):
            assert np.array_equal(got[j], oracle.decode(body, n, j, coding)), (j, coding)
    assert ctx.missing_count() == int((G < 0).sum())
    assert np.array_equal(ctx.read_body(), body)


def test_decode_fixtures(bmg):
    for (body, y), n, p in [(plinktest(), 5, 10), (small_modelspace(), 30, 7)]:
        ctx = bmg.Ctx(n, p, body, y)
        for c in range(4):
            got = ctx.decode(np.arange(p), c)
            for j in range(p):
                assert np.array_equal(got[j], oracle.decode(body, n, j, c))


@pytest.mark.parametrize("types", [("A",), ("AH",), ("A", "H", "D", "R", "AH"), ("D",)])
def test_self_moments_bit_exact(bmg, types):
    rng = np.random.default_rng(11)
    n, p = 333, 41
    G, body = random_body(rng, n, p, with_missing=True)
    y = rng.normal(size=n)
    ctx = bmg.Ctx(n, p, body, y, types=types)
    d = oracle.Data(n, p, body, y)
    dm = oracle.DataModel(d, list(types))
    assert np.array_equal(ctx.self_moments(), dm.self_moments())
    st, so = ctx.stats(), d.stats()
    for k in so:
        assert st[k] == so[k], k


def test_recode_matches_oracle(bmg):
    rng = np.random.default_rng(4)
    n, p = 101, 30
    G = rng.integers(0, 3, size=(p, n))
    G[:10] = np.where(rng.random(size=(10, n)) < 0.8, 2, G[:10])   # allele freq > 0.5
    body = pack_genotypes(G)
    y = rng.normal(size=n)
    ctx = bmg.Ctx(n, p, body, y, recode=True)
    d = oracle.Data(n, p, body, y, recode=True)
    assert np.array_equal(ctx.read_body(), d.body_now())


@pytest.mark.parametrize("W,ca,cb,p", [(8, 0, 0, 300), (40, 0, 0, 300), (17, 0, 1, 300), (100, 3, 2, 300),
                                        (1, 0, 0, 300), (21, 0, 0, 389), (64, 1, 1, 300), (200, 2, 0, 700),
                                        (500, 0, 0, 700), (560, 3, 3, 900), (560, 1, 2, 700),
                                        (100, 0, 0, 60), (33, 2, 1, 129)])
def test_band_xtx_exact(bmg, W, ca, cb, p):
    """Every column-tile bucket (W up to 560), same and different codings,
    row blocks cut by p, missing genotypes."""
    rng = np.random.default_rng(W + p)
    n = 517
    G, body = random_body(rng, n, p, with_missing=True)
    ctx = bmg.Ctx(n, p, body, rng.normal(size=n))
    got = ctx.band_xtx(W, ca, cb)
    ref = np.zeros(p * W, dtype=np.int64)
    oracle.lib().orc_band_xtx(oracle.u8(body), (n + 3) // 4, n, p, W, ca, cb,
                              ref.ctypes.data_as(oracle.C.POINTER(oracle.C.c_int64)))
    assert np.array_equal(got.astype(np.int64), ref.reshape(p, W))


@pytest.mark.parametrize("n,p,nvec,digits", [(1000, 300, 1, 5), (2053, 700, 3, 5), (777, 129, 4, 4)])
def test_scan_xtr_fixed_point_bound(bmg, n, p, nvec, digits):
    rng = np.random.default_rng(n)
    G, body = random_body(rng, n, p)
    ctx = bmg.Ctx(n, p, body, rng.normal(size=n))
    R = rng.normal(size=(nvec, n)) * 3.0
    got = ctx.scan_xtr(R, coding=0, digits=digits)
    X = G.astype(np.float64)
    ref = R @ X.T
    # error bound: |r_i - R_i 2^e| <= 2^(e-1) per element, 2^e = 2^(ceil log2 max|r|) / 2^(8D-2)
    for v in range(nvec):
        m = np.abs(R[v]).max()
        e = np.frexp(m)[1]
        step = 2.0 ** (e - (8 * digits - 2))
        bound = 0.5 * step * X.sum(1) + 1e-12 * np.abs(ref[v])
        assert np.all(np.abs(got[v] - ref[v]) <= bound)


def test_scan_xtr_is_exact_on_representable_residuals(bmg):
    """Residuals that are exact multiples of the fixed-point step give the
    exact dot product (linearity and integer exactness of the MFMA path)."""
    rng = np.random.default_rng(2)
    n, p = 640, 200
    G, body = random_body(rng, n, p)
    ctx = bmg.Ctx(n, p, body, rng.normal(size=n))
    k = rng.integers(-1000, 1000, size=n).astype(np.float64)
    r = k / 1024.0
    got = ctx.scan_xtr(r[None, :], coding=0, digits=4)[0]
    assert np.array_equal(got, G.astype(np.float64) @ r)
    got2 = ctx.scan_xtr(np.stack([r, 2 * r]), coding=0, digits=4)
    assert np.array_equal(got2[1], 2 * got2[0])


def test_scan_shapes_all_exact(bmg):
    """Every shape the scan autotuner may pick (tiles per wave, LDS tile
    width, so one or several individual tiles added atomically, and the
    number of workgroups looping over the SNP blocks) gives the exact dot
    products on representable residuals; the tuned shape is one of them."""
    rng = np.random.default_rng(11)
    n, p = 20011, 1500
    G, body = random_body(rng, n, p)
    ctx = bmg.Ctx(n, p, body, rng.normal(size=n))
    k = rng.integers(-1000, 1000, size=(3, n)).astype(np.float64)
    R = k / 1024.0
    ref = np.stack([(G.astype(np.float64) @ R[v]) for v in range(3)])
    tuned = ctx.scan_xtr(R, coding=0, digits=5)
    assert np.array_equal(tuned, ref)
    shape = ctx.scan_shape(3, 5)
    assert shape["tiles_per_wave"] in (2, 4) and shape["workgroups"] >= 1
    assert shape["waves"] in (4, 8, 16)
    w32 = ((n + 15) // 16 + 31) // 32 * 32
    for waves in (4, 8, 16):
        for nt in (2, 4):
            for tw in (32, 288, 640):
                for gx in (1, 5, 4096):
                    ctx.scan_set_shape(3, 5, nt, tw, gx, waves)
                    got = ctx.scan_xtr(R, coding=0, digits=5)
                    assert np.array_equal(got, ref), (waves, nt, tw, gx)
                    sh = ctx.scan_shape(3, 5)
                    assert sh["n_tiles"] == -(-w32 // tw)
    with pytest.raises(bmg.BmgError):
        ctx.scan_set_shape(3, 5, 3, 272, 8)
    with pytest.raises(bmg.BmgError):
        ctx.scan_set_shape(3, 5, 2, 704, 8)
    with pytest.raises(bmg.BmgError):
        ctx.scan_set_shape(3, 5, 2, 272, 8)  # not a multiple of 32
    with pytest.raises(bmg.BmgError):
        ctx.scan_set_shape(3, 5, 2, 272, 8, waves=6)


@pytest.mark.parametrize("coding", [0, 1, 2, 3])
def test_scan_codings_exact(bmg, coding):
    """The scan decodes each genotype coding (additive, heterozygous,
    dominant, recessive) to the values of the exact decode (itself
    bit-exact against the oracle above), missing genotypes as 0."""
    rng = np.random.default_rng(5 + coding)
    n, p = 3001, 700
    G, body = random_body(rng, n, p, with_missing=True)
    ctx = bmg.Ctx(n, p, body, rng.normal(size=n))
    X = np.concatenate([ctx.decode(np.arange(j, min(p, j + 100)), coding) for j in range(0, p, 100)])
    X[X < 0] = 0.0  # missing genotypes enter the scan as 0
    r = rng.integers(-500, 500, size=(2, n)).astype(np.float64) / 64.0
    got = ctx.scan_xtr(r, coding=coding, digits=4)
    assert np.array_equal(got, r @ X.T)


@pytest.mark.parametrize("n,p,ld", [(37, 50, None), (203, 300, (0.9, 0.9, 10, 40)),
                                    (64, 500, (0.7, 0.95, 20, 100))])
def test_synthetic_genotypes_match_oracle(bmg, n, p, ld):
    from bmagwa_amd import synth
    thresh, rho, maf = synth.snp_layout(p, (0.05, 0.5), ld, 123)
    ctx = bmg.Ctx(n, p, synth=(99, thresh, rho))
    ref = oracle.synth_genotypes(99, n, p, thresh, rho)
    assert np.array_equal(ctx.read_body(), ref)
