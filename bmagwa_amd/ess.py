"""PIP effective sample size of inclusion-indicator traces.

The reference defines no ESS (its only diagnostic is the MCMC-frequency PIP of
bmagwa_postprocess.py:79-123).  This module fixes the definition used by the
benchmark:

* a chain's thinned trace gives, for every monitored SNP j, the indicator
  series gamma_j(t) (is j in the model at thinned sample t);
* the ESS of one series is Geyer's (1992) initial monotone sequence
  estimator on its autocorrelations; a constant series (never or always
  included) counts as fully informative, ESS = number of samples;
* ESS is summed over chains; the benchmark reports the median (and minimum)
  over the monitored SNPs divided by the wall time of the sampling.
"""
import numpy as np


def indicator_matrix(sizes, loci, snps, burn=0):
    """sizes[t], loci (concatenated) -> bool matrix (len(snps), T - burn)."""
    snps = np.asarray(snps)
    pos = {int(j): k for k, j in enumerate(snps)}
    T = len(sizes)
    M = np.zeros((len(snps), max(T - burn, 0)), dtype=bool)
    off = np.concatenate([[0], np.cumsum(sizes.astype(np.int64))])
    for t in range(burn, T):
        for j in loci[off[t]:off[t + 1]]:
            k = pos.get(int(j))
            if k is not None:
                M[k, t - burn] = True
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
    res = []
    for r in rho:
        npairs = (T - 1) // 2
        g = r[0:2 * npairs:2] + r[1:2 * npairs + 1:2]
        s, prev = 0.0, np.inf
        for gk in g:
            if gk <= 0:
                break
            gk = min(gk, prev)
            prev = gk
            s += gk
        tau = max(-1.0 + 2.0 * s, 1.0 / T)
        res.append(T / tau)
    out[live] = res
    return out


def pips(M):
    return M.mean(axis=1) if M.shape[1] else np.zeros(M.shape[0])
