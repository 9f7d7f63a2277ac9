"""PIP effective sample size of inclusion-indicator traces.

The reference defines no ESS (its only diagnostic is the MCMC-frequency PIP of
bmagwa_postprocess.py:79-123).  This module fixes the definition used by the
benchmark:

* a chain's thinned trace gives, for every monitored SNP j, the indicator
  series gamma_j(t) (is j in the model at thinned sample t);
* `ess_multichain`: the multi-chain ESS of Vehtari, Gelman, Simpson,
  Carpenter and Buerkner (2021, "Rank-normalization, folding, and
  localization", eq. 10-11, without the rank normalisation and chain
  splitting): autocorrelations from the chains' mean autocovariance against
  the pooled variance var+ = (T-1)/T W + B/T, so chains that sit in
  different states lower the ESS; Geyer's (1992) initial monotone sequence
  truncates the sum.  A SNP none of whose chains ever changes has no
  information about an uncertain PIP: ESS 0.  The benchmark monitors SNPs
  whose PIP lies in [0.05, 0.95] and reports the median (and minimum) over
  them divided by the wall time of the sampling;
* `ess_geyer`: the single-series form (a constant series counts as fully
  informative there), kept for the oracle comparison.
"""
import numpy as np


def indicator_matrix(sizes, loci, snps, burn=0):
    """sizes[t], loci (concatenated) -> bool matrix (len(snps), T - burn)."""
    snps = np.asarray(snps, dtype=np.int64)
    sizes = np.asarray(sizes, dtype=np.int64)
    loci = np.asarray(loci, dtype=np.int64)
    T = len(sizes)
    M = np.zeros((len(snps), max(T - burn, 0)), dtype=bool)
    if T <= burn or len(snps) == 0:
        return M
    off = int(sizes[:burn].sum())
    lo = loci[off:]
    t = np.repeat(np.arange(T - burn), sizes[burn:])
    # position of each locus among the monitored SNPs (-1: not monitored)
    order = np.argsort(snps)
    k = np.searchsorted(snps[order], lo)
    k = np.minimum(k, len(snps) - 1)
    hit = snps[order][k] == lo
    M[order[k[hit]], t[hit]] = True
    return M


def ess_geyer(X):
    """Row-wise Geyer initial-monotone-sequence ESS of a (m, T) array."""
    X = np.asarray(X, dtype=np.float64)
    m, T = X.shape
    out = np.full(m, float(T))
    if T < 4:
        return out
    Xc = X - X.mean(axis=1, keepdims=True)
    var = (Xc * Xc).mean(axis=1)
    live = var > 0
    if not live.any():
        return out
    nfft = 1 << int(np.ceil(np.log2(2 * T)))
    F = np.fft.rfft(Xc[live], n=nfft, axis=1)
    ac = np.fft.irfft(F * np.conj(F), n=nfft, axis=1)[:, :T] / T
    rho = ac / var[live][:, None]
    npairs = (T - 1) // 2
    # Gamma_k = rho_2k + rho_2k+1, summed up to the first non-positive one,
    # each clipped to the running minimum (initial monotone sequence)
    g = rho[:, 0:2 * npairs:2] + rho[:, 1:2 * npairs + 1:2]
    pos = g > 0
    first_bad = np.where(pos.all(axis=1), npairs, np.argmin(pos, axis=1))
    g = np.minimum.accumulate(np.where(pos, g, 0.0), axis=1)
    keep = np.arange(npairs)[None, :] < first_bad[:, None]
    s = np.where(keep, g, 0.0).sum(axis=1)
    res = T / np.maximum(-1.0 + 2.0 * s, 1.0 / T)
    out[live] = res
    return out


def ess_multichain(X):
    """Multi-chain ESS of each row of X, a (chains, m, T) array of series
    (chains x monitored SNPs x samples); returns (m,) ESS over all chains."""
    X = np.asarray(X, dtype=np.float64)
    C, m, T = X.shape
    out = np.zeros(m)
    if T < 4:
        return out
    mean_c = X.mean(axis=2)                       # (C, m)
    Xc = X - mean_c[:, :, None]
    nfft = 1 << int(np.ceil(np.log2(2 * T)))
    acov = np.zeros((m, T))
    for c in range(C):                            # chain by chain: memory stays (m, nfft)
        F = np.fft.rfft(Xc[c], n=nfft, axis=1)
        acov += np.fft.irfft(F * np.conj(F), n=nfft, axis=1)[:, :T] / T
    acov /= C                                     # mean over chains of the biased autocovariance
    W = acov[:, 0] * T / (T - 1)                  # mean within-chain variance
    var_plus = W * (T - 1) / T
    if C > 1:
        var_plus = var_plus + mean_c.var(axis=0, ddof=1)
    live = var_plus > 0
    if not live.any():
        return out
    rho = 1.0 - (W[live, None] - acov[live]) / var_plus[live, None]
    rho[:, 0] = 1.0
    npairs = (T - 1) // 2
    g = rho[:, 0:2 * npairs:2] + rho[:, 1:2 * npairs + 1:2]
    pos = g > 0
    first_bad = np.where(pos.all(axis=1), npairs, np.argmin(pos, axis=1))
    g = np.minimum.accumulate(np.where(pos, g, 0.0), axis=1)
    keep = np.arange(npairs)[None, :] < first_bad[:, None]
    s = np.where(keep, g, 0.0).sum(axis=1)
    tau = np.maximum(-1.0 + 2.0 * s, 1.0 / (C * T))
    out[live] = C * T / tau
    return out


def pips(M):
    return M.mean(axis=1) if M.shape[1] else np.zeros(M.shape[0])
