"""The allocation-free scan entry point (bmg_scan_xtr_dev_noalloc) gives the
same X^T R as bmg_scan_xtr_dev, bit for bit, on two calls in a row whose
second call is smaller (fewer residual vectors on a smaller dataset), so the
shared scratch must not carry state from the first call into the second."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _scan_both(bmg, ctx, n, p, nvec, coding, seed):
    n_pad = ctx.n_pad()
    rng = np.random.default_rng(seed)
    r = np.zeros((nvec, n_pad))
    r[:, :n] = rng.normal(size=(nvec, n)) * rng.uniform(0.1, 10.0, size=(nvec, 1))
    r_dev = bmg.DevBuf(r.nbytes)
    r_dev.upload(r)
    outs = []
    for noalloc in (False, True):
        o = bmg.DevBuf(nvec * p * 8)
        ctx.scan_xtr_dev(r_dev.ptr, nvec, o.ptr, coding=coding, digits=5, noalloc=noalloc)
        outs.append(o.download(np.float64, nvec * p).reshape(nvec, p))
        o.free()
    r_dev.free()
    return outs


def test_noalloc_scan_matches_allocating_scan_twice():
    import bmagwa_amd as bmg
    from bmagwa_amd import synth
    cases = [(3000, 20000, 3, 0, 11), (700, 5000, 1, 2, 12)]   # second call smaller
    for n, p, nvec, coding, seed in cases:
        print("noalloc case n=%d p=%d nvec=%d coding=%d" % (n, p, nvec, coding), flush=True)
        thresh, rho, _ = synth.snp_layout(p, (0.05, 0.5), (0.7, 0.95, 20, 100), seed)
        ctx = bmg.Ctx(n, p, synth=(seed, thresh, rho), types=("A",))
        ref, got = _scan_both(bmg, ctx, n, p, nvec, coding, seed)
        assert np.array_equal(ref, got)
        assert np.abs(ref).max() > 0
        ctx.close()


def test_noalloc_scan_rejects_oversized_inputs():
    import bmagwa_amd as bmg
    from bmagwa_amd import synth
    n, p = 16, 1047553          # one SNP past the scratch's limit
    thresh, rho, _ = synth.snp_layout(p, (0.05, 0.5), None, 3)
    ctx = bmg.Ctx(n, p, synth=(3, thresh, rho), types=("A",))
    r_dev = bmg.DevBuf(ctx.n_pad() * 8)
    r_dev.upload(np.ones(ctx.n_pad()))
    o = bmg.DevBuf(p * 8)
    with pytest.raises(bmg.BmgError, match="allocation-free"):
        ctx.scan_xtr_dev(r_dev.ptr, 1, o.ptr, noalloc=True)
    o.free()
    r_dev.free()
    ctx.close()
