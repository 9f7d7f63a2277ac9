"""Decoding, missing-value index and per-SNP moments in the oracle, against
the reference's known answers (src/tests/data_tests.hpp,
data_model_tests.hpp, precomputed_snp_covariances_tests.hpp)."""
import os

import numpy as np
import pytest

import oracle
from fixtures import (GOLDEN, PLINKTEST_G, SMALL_MODELSPACE_G, SMALL_MODELSPACE_Y,
                      pack_genotypes, plinktest, small_modelspace)

CODING_OF = {0: lambda g: g, 1: lambda g: np.where(g < 0, -1, (g == 1) * 1.0),
             2: lambda g: np.where(g < 0, -1, (g > 0) * 1.0),
             3: lambda g: np.where(g < 0, -1, (g == 2) * 1.0)}


def test_bed_header_and_fixture_bytes():
    raw = open(os.path.join(GOLDEN, "plinktest.bed"), "rb").read()
    assert raw.hex() == "6c1b01" + "7e00ea00ba028403fc03ae00cb02e202a200a802"


def test_decode_small_modelspace_kat():
    body, y = small_modelspace()
    assert np.allclose(y, SMALL_MODELSPACE_Y)
    for j in range(7):
        g = SMALL_MODELSPACE_G[j]
        for c in range(4):
            assert np.array_equal(oracle.decode(body, 30, j, c), CODING_OF[c](g))


def test_decode_plinktest_kat_with_missing():
    body, y = plinktest()
    for j in range(10):
        for c in range(4):
            assert np.array_equal(oracle.decode(body, 5, j, c), CODING_OF[c](PLINKTEST_G[j]))
    d = oracle.Data(5, 10, body, y)
    # src/tests/data_tests.hpp:171-191
    for j in range(10):
        locs, pc = d.missing(j)
        if j == 0:
            assert locs == [3] and list(pc) == [1, 2, 4]
        elif j == 3:
            assert locs == [1] and list(pc) == [2, 3, 4]
        else:
            assert locs == []


def test_decode_all_bytes_against_reference_tables():
    """All 256 byte values x 4 codings vs src/genotype_tables.hpp."""
    tables = np.load(os.path.join(GOLDEN, "genotype_tables.npy"))
    body = np.arange(256, dtype=np.uint8)   # one SNP with n = 1024
    for c in range(4):
        out = oracle.decode(body, 1024, 0, c).reshape(256, 4)
        assert np.array_equal(out, tables[c].astype(np.float64))


def test_decode_tail_byte_partial():
    G = np.random.default_rng(3).integers(0, 3, size=(4, 7))
    body = pack_genotypes(G)
    for j in range(4):
        assert np.array_equal(oracle.decode(body, 7, j, 0), G[j])


def test_datamodel_fills_missing_from_miss_val():
    body, y = plinktest()
    d = oracle.Data(5, 10, body, y)
    dm = oracle.DataModel(d, ["A"])
    # missing values start at 0 (src/data_model.hpp:75-88)
    assert list(dm.get(0, 0)) == [1, 2, 2, 0, 0]
    oracle.lib().orc_datamodel_set_missing(dm.h, 0, 1, 2)
    assert list(dm.get(0, 0)) == [1, 2, 2, 2, 0]
    assert list(dm.get(1, 0)) == [1, 0, 0, 0, 0]   # H: 2 -> 0
    assert list(dm.get(2, 0)) == [1, 1, 1, 1, 0]   # D
    assert list(dm.get(3, 0)) == [0, 1, 1, 1, 0]   # R


@pytest.mark.parametrize("types,offset", [(["A"], 2), (["A", "AH"], 5),
                                          (["A", "H", "D", "R", "AH"], 9), (["AH"], 5)])
def test_self_moments_closed_forms(types, offset):
    """src/tests/precomputed_snp_covariances_tests.hpp:30-150."""
    body, y = plinktest()
    d = oracle.Data(5, 10, body, y)
    dm = oracle.DataModel(d, types)
    off, ot = dm.offset()
    assert off == offset
    sm = dm.self_moments()
    for j in range(10):
        row = list(sm[j])
        for t in range(4):
            if ot[t] < 0:
                continue
            x = dm.get(t, j)
            s = row.pop(0)
            css = row.pop(0)
            assert s == pytest.approx(x.mean() * 5, abs=1e-12)
            assert css == pytest.approx(x.var(ddof=1) * 4, abs=1e-10)
        if "AH" in types:
            a, h = dm.get(0, j), dm.get(1, j)
            assert row.pop(0) == pytest.approx(a @ h - 5 * a.mean() * h.mean(), abs=1e-10)
        assert row == []


def test_update_prexx_matches_recompute():
    """src/tests/data_model_tests.hpp:190-320: after setting missing values,
    the updated moments equal a recomputation."""
    body, y = plinktest()
    d = oracle.Data(5, 10, body, y)
    dm = oracle.DataModel(d, ["A", "H", "D", "R", "AH"])
    base = dm.self_moments()
    for v in (0, 1, 2):
        oracle.lib().orc_datamodel_set_missing(dm.h, 0, 1, v)
        oracle.lib().orc_datamodel_set_missing(dm.h, 3, 1, v)
        fresh = dm.self_moments()
        for j in (0, 3):
            pre = base[j].copy()
            oracle.lib().orc_datamodel_update_prexx(dm.h, j, oracle.ptr(pre))
            assert np.allclose(pre, fresh[j], atol=1e-10)


@pytest.mark.parametrize("ca,cb", [(0, 0), (0, 1), (1, 1), (3, 2)])
def test_band_xtx_bruteforce(ca, cb):
    rng = np.random.default_rng(5)
    n, p, W = 37, 23, 6
    G = rng.integers(0, 3, size=(p, n))
    body = pack_genotypes(G)
    out = np.zeros(p * W, dtype=np.int64)
    oracle.lib().orc_band_xtx(oracle.u8(body), (n + 3) // 4, n, p, W, ca, cb,
                              out.ctypes.data_as(oracle.C.POINTER(oracle.C.c_int64)))
    Xa = np.stack([CODING_OF[ca](g) for g in G]).astype(np.int64)
    Xb = np.stack([CODING_OF[cb](g) for g in G]).astype(np.int64)
    ref = np.zeros((p, W), dtype=np.int64)
    for j in range(p):
        for dl in range(W):
            if j + dl < p:
                ref[j, dl] = Xa[j] @ Xb[j + dl]
    assert np.array_equal(out.reshape(p, W), ref)
    if ca == cb:   # diagonal = sum of squares
        assert np.array_equal(out.reshape(p, W)[:, 0], (Xa * Xa).sum(1))
