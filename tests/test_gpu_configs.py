"""The benchmark configurations at their full sizes on the GPU: the AH
sampler with the 500-SNP band at n = 5,000, p = 100,000; 64 chains on each of
8 phenotypes measured on uneven subsets of a 20,000-individual pool
(p = 200,000) in one batch, graph against eager and against the oracle; the
n = 10,000, p = 500,000 headline dataset against the oracle; and the
n = 50,000, p = 1,000,000 dataset (12.5 GB packed) with its exact scan,
band rows and two 256-column chains against the oracle."""
import numpy as np
import pytest

import oracle
from bmagwa_amd import synth

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def bmg():
    import bmagwa_amd
    bmagwa_amd.lib()
    return bmagwa_amd


def log(msg):
    print(msg, flush=True)


def device_dataset(bmg, n, p, maf, ld, n_causal, h2, seed, types="A", dominance=False):
    thresh, rho, _ = synth.snp_layout(p, maf, ld, seed)
    ctx = bmg.Ctx(n, p, synth=(seed, thresh, rho), types=tuple(types.split(",")))
    # additive decoding of the causal SNPs; the dominance effect acts on the
    # heterozygotes (additive value 1)
    y, causal, _ = synth.phenotype(lambda s: ctx.decode(np.asarray(s), 0), n, p, n_causal, h2, seed,
                                   dominance=dominance)
    ctx.set_y(y)
    return ctx, y, causal


def oracle_chain(cfg, body, y, n_iter):
    s = oracle.Sampler(cfg, oracle.Data(cfg["n"], cfg["m_g"], body, y))
    s.init_flat()
    assert s.setup() == 0
    s.iterate(n_iter)
    return s


def compare(tr_g, tr_o, n_iter):
    assert len(tr_g["move"]) == n_iter
    first_bad = None
    for key in ("move", "jd", "ms", "size"):
        neq = np.flatnonzero(tr_g[key] != tr_o[key])
        if len(neq):
            first_bad = neq[0] if first_bad is None else min(first_bad, neq[0])
    assert first_bad is None, "diverged at iteration %d" % first_bad
    assert np.allclose(tr_g["loglik"], tr_o["loglik"], rtol=1e-10, atol=0)


def exact_xtr(body, n, snps, r_int, coding=0):
    """int64 X^T r on the CPU for the given SNPs."""
    return np.array([int(np.dot(oracle.decode(body, n, int(j), coding).astype(np.int64), r_int))
                     for j in snps], dtype=np.int64)


def test_c2_ah_sampler_and_band(bmg):
    """n = 5,000, p = 100,000, rho = 0.9 blocks of 50, 50 causal SNPs with
    dominance effects, h2 = 0.4, A + dominance (AH) coding: 300 iterations of
    the device sampler equal the oracle's; the W = 500 band of the A and the
    H coding is exact on sampled rows and its diagonal is sum x^2 of the
    self moments."""
    n, p, W = 5000, 100000, 500
    seed = synth.seed_of(2)
    ctx, y, causal = device_dataset(bmg, n, p, (0.02, 0.5), (0.9, 0.9, 50, 50), 50, 0.4, seed,
                                    types="AH", dominance=True)
    body = ctx.read_body()
    log("c2 dataset n=%d p=%d AH ready" % (n, p))
    pre = ctx.self_moments()
    rng = np.random.default_rng(2)
    rows = np.sort(rng.choice(p - W, size=6, replace=False))
    stride = (n + 3) // 4
    for ca, cb in ((0, 0), (1, 1), (0, 1)):
        band = ctx.band_xtx(W, ca, cb)
        if ca == cb:
            # diagonal = sum x^2 = centred square sum + (sum x)^2 / n
            sx, css = (pre[:, 0], pre[:, 1]) if ca == 0 else (pre[:, 2], pre[:, 3])
            assert np.allclose(band[:, 0], css + sx * sx / n, rtol=1e-12, atol=1e-6)
        for j in rows:
            sub = body[j * stride:(j + W + 1) * stride]
            ref = np.zeros((W + 1) * W, dtype=np.int64)
            oracle.lib().orc_band_xtx(oracle.u8(sub), stride, n, W + 1, W, ca, cb,
                                      ref.ctypes.data_as(oracle.C.POINTER(oracle.C.c_int64)))
            assert np.array_equal(band[j], ref[:W].astype(np.int32)), (ca, cb, j)
    log("c2 band checked")
    n_iter = 300
    cfg = synth.sampler_cfg(n, p, 50, types="AH", seed=seed, chain=0, n_rao=100, n_rao_burnin=2,
                            thin=10, do_n_iter=n_iter)
    b = bmg.Batch([ctx], [0], [cfg], dict(trace_every_iter=1))
    b.setup()
    b.run(n_iter)
    tr = b.trace(0)
    s = oracle_chain(cfg, body, y, n_iter)
    compare(tr, s.trace(), n_iter)
    assert np.allclose(b.p_rao(0), s.p_rao(), rtol=1e-8, atol=1e-10)
    log("c2 trajectory equal to the oracle over %d iterations, final size %d" % (n_iter, tr["size"][-1]))


def test_c3_uneven_phenotypes_one_batch(bmg):
    """64 chains on each of 8 phenotypes, each phenotype measured on its own
    subset of a 20,000-individual pool (n from 2,000 to 20,000), p = 200,000
    SNPs, rho = 0.9 blocks of 50, 5-100 causal SNPs, h2 ~ U(0.1, 0.6), all
    512 chains in one launch per segment: the hipGraph replay equals the
    eager run for every chain, and three chains (smallest, middle and
    largest n) equal the oracle."""
    N, p = 20000, 200000
    seed = synth.seed_of(3)
    thresh, rho, _ = synth.snp_layout(p, (0.02, 0.5), (0.9, 0.9, 50, 50), seed)
    pool = bmg.Ctx(N, p, synth=(seed, thresh, rho))
    rng = np.random.default_rng(seed)
    sizes = np.sort(rng.integers(2000, 20001, size=8))
    sizes[0], sizes[-1] = 2000, 20000
    subsets, ys = [], []
    for k, nk in enumerate(sizes):
        sub = np.sort(rng.choice(N, size=int(nk), replace=False)).astype(np.uint32)
        nc = int(rng.integers(5, 101))
        h2 = float(rng.uniform(0.1, 0.6))
        dec = lambda s: pool.decode(np.asarray(s), 0)[:, sub]
        y, _, _ = synth.phenotype(dec, int(nk), p, nc, h2, seed + k)
        subsets.append(sub)
        ys.append(y)
    log("c3 pool N=%d p=%d, subsets %s" % (N, p, list(map(int, sizes))))
    n_iter, per = 400, 64
    cfgs, chain_ctx = [], []
    for k, nk in enumerate(sizes):
        for c in range(per):
            i = len(cfgs)
            cfgs.append(synth.sampler_cfg(int(nk), p, 50, seed=5000 + i, chain=i, n_rao=100, n_rao_burnin=2,
                                          thin=10, do_n_iter=n_iter))
            chain_ctx.append(k)
    out = {}
    bodies = None
    for g in (0, 1):
        ctxs = [bmg.Ctx.subset(pool, subsets[k], ys[k]) for k in range(8)]
        if bodies is None:
            bodies = [ctxs[k].read_body() for k in (0, 4, 7)]
        b = bmg.Batch(ctxs, chain_ctx, cfgs, dict(trace_every_iter=1, use_graph=g))
        b.setup()
        b.run(200)
        b.run(n_iter - 200)
        st = b.graph_stats()
        assert (st["replays"] >= 2) if g else (st["replays"] == 0), st
        out[g] = [(b.trace(c), b.p_rao(c)) for c in range(len(cfgs))]
        b.close()
        for x in ctxs:
            x.close()
        log("c3 run with use_graph=%d done (%d chains)" % (g, len(cfgs)))
    for c in range(len(cfgs)):
        (t0, r0), (t1, r1) = out[0][c], out[1][c]
        for key in t0:
            assert np.array_equal(t0[key], t1[key]), (c, key)
        assert np.array_equal(r0, r1), c
    for bi, k in enumerate((0, 4, 7)):
        c = k * per + 1
        s = oracle_chain(cfgs[c], bodies[bi], ys[k], n_iter)
        compare(out[1][c][0], s.trace(), n_iter)
        assert np.allclose(out[1][c][1], s.p_rao(), rtol=1e-8, atol=1e-10)
        log("c3 chain %d (n=%d) equal to the oracle" % (c, sizes[k]))


def test_c4_headline_dataset_against_oracle(bmg):
    """n = 10,000, p = 500,000, rho ~ U(0.7, 0.95) in blocks of 20-100, 100
    causal SNPs, h2 = 0.5 (the benchmark's dataset): two chains of 500
    iterations, RB block included, equal the oracle's."""
    n, p = 10000, 500000
    seed = 20261202 + 4000
    ctx, y, causal = device_dataset(bmg, n, p, (0.02, 0.5), (0.7, 0.95, 20, 100), 100, 0.5, seed)
    body = ctx.read_body()
    log("c4 dataset ready")
    n_iter = 500
    cfgs = [synth.sampler_cfg(n, p, 100, seed=100 + c, chain=c, n_rao=250, n_rao_burnin=1, thin=10,
                              do_n_iter=n_iter) for c in range(2)]
    b = bmg.Batch([ctx], [0, 0], cfgs, dict(trace_every_iter=1))
    b.setup()
    b.run(n_iter)
    for c in range(2):
        s = oracle_chain(cfgs[c], body, y, n_iter)
        compare(b.trace(c), s.trace(), n_iter)
        assert np.allclose(b.p_rao(c), s.p_rao(), rtol=1e-8, atol=1e-10)
        log("c4 chain %d equal to the oracle" % c)


def test_c5_largest_dataset(bmg):
    """n = 50,000, p = 1,000,000 (12.5 GB packed), MAF ~ U(0.01, 0.5), 300
    causal SNPs: the scan is exact on representable residuals (sampled SNPs
    against int64 sums), band rows equal the oracle's, and two chains of the
    256-column kernel with the 200-SNP limit equal the oracle's for 100
    iterations (no RB block inside the run)."""
    n, p = 50000, 1000000
    seed = synth.seed_of(5)
    ctx, y, causal = device_dataset(bmg, n, p, (0.01, 0.5), (0.7, 0.95, 20, 100), 300, 0.5, seed)
    log("c5 dataset on the device")
    rng = np.random.default_rng(5)
    r_int = rng.integers(-3, 4, size=n)
    got = ctx.scan_xtr(r_int[None, :].astype(np.float64), coding=0, digits=5)[0]
    body = ctx.read_body()
    snps = np.sort(rng.choice(p, size=1500, replace=False))
    assert np.array_equal(got[snps], exact_xtr(body, n, snps, r_int).astype(np.float64))
    log("c5 scan exact on %d sampled SNPs" % len(snps))
    W = 64
    band = ctx.band_xtx(W, 0, 0)
    stride = (n + 3) // 4
    for j in np.sort(rng.choice(p - W, size=4, replace=False)):
        sub = body[j * stride:(j + W + 1) * stride]
        ref = np.zeros((W + 1) * W, dtype=np.int64)
        oracle.lib().orc_band_xtx(oracle.u8(sub), stride, n, W + 1, W, 0, 0,
                                  ref.ctypes.data_as(oracle.C.POINTER(oracle.C.c_int64)))
        assert np.array_equal(band[j], ref[:W].astype(np.int32)), j
    del band
    log("c5 band rows checked")
    n_iter = 100
    cfgs = [synth.sampler_cfg(n, p, 300, seed=900 + c, chain=c, n_rao=1000, n_rao_burnin=1, thin=10,
                              do_n_iter=n_iter) for c in range(2)]
    b = bmg.Batch([ctx], [0, 0], cfgs, dict(trace_every_iter=1, capacity=256, max_model_size=200))
    b.setup()
    b.run(n_iter)
    sizes = [int(b.trace(c)["size"].max()) for c in range(2)]
    assert max(sizes) < 180, sizes      # the 200-SNP limit never binds in this window
    d = oracle.Data(n, p, body, y)
    del body
    for c in range(2):
        s = oracle.Sampler(cfgs[c], d)
        s.init_flat()
        assert s.setup() == 0
        s.iterate(n_iter)
        compare(b.trace(c), s.trace(), n_iter)
        log("c5 chain %d equal to the oracle over %d iterations" % (c, n_iter))
