"""Synthetic GWAS datasets of the benchmark configurations.

Genotypes follow the latent-AR(1) haplotype model implemented by the
engine's generator kernel (bmg_ctx_create_synthetic): per SNP a MAF, a
threshold thresh = Phi^-1(MAF) on a latent Gaussian, and an AR coefficient
rho_j linking SNP j to j-1 inside LD blocks (rho_j = 0 starts a block).  The
phenotype is y = X beta + e with causal SNPs drawn without replacement,
beta ~ N(0, 1) and var(e) set by the heritability.  Everything is seeded as
seed = 20261202 + 1000 * config + replicate.
"""
import numpy as np
from scipy.special import ndtri

CONFIGS = {
    # n, p, maf_lo, block: (rho_lo, rho_hi, len_lo, len_hi) or None, causal, h2
    1: dict(n=1000, p=10000, maf=(0.05, 0.5), ld=None, causal=10, h2=0.3, types="A"),
    2: dict(n=5000, p=100000, maf=(0.02, 0.5), ld=(0.9, 0.9, 50, 50), causal=50, h2=0.4,
            types="AH"),
    3: dict(n=20000, p=200000, maf=(0.02, 0.5), ld=(0.9, 0.9, 50, 50), causal=(5, 100),
            h2=(0.1, 0.6), types="A"),
    4: dict(n=10000, p=500000, maf=(0.02, 0.5), ld=(0.7, 0.95, 20, 100), causal=100, h2=0.5,
            types="A"),
    5: dict(n=50000, p=1000000, maf=(0.01, 0.5), ld=(0.7, 0.95, 20, 100), causal=300, h2=0.5,
            types="A"),
}


def seed_of(config, replicate=0):
    return 20261202 + 1000 * config + replicate


def snp_layout(p, maf, ld, seed):
    """thresh[p], rho[p], maf[p] for the genotype generator."""
    rng = np.random.default_rng(seed)
    m = rng.uniform(maf[0], maf[1], size=p)
    thresh = ndtri(m)
    rho = np.zeros(p)
    if ld is not None:
        r_lo, r_hi, l_lo, l_hi = ld
        j = 0
        while j < p:
            length = int(rng.integers(l_lo, l_hi + 1))
            r = rng.uniform(r_lo, r_hi) if r_hi > r_lo else r_lo
            rho[j + 1:j + length] = r
            j += length
    return thresh, rho, m


def phenotype(decode_causal, n, p, n_causal, h2, seed, dominance=False):
    """y = sum beta_j x_j (+ beta_H h_j) + e; decode_causal(snps) -> (k, n)
    additive genotypes.  Returns y, causal indices, effects."""
    rng = np.random.default_rng(seed + 7)
    causal = np.sort(rng.choice(p, size=n_causal, replace=False))
    beta = rng.normal(size=n_causal)
    X = decode_causal(causal)
    g = beta @ X
    if dominance:
        bh = rng.normal(size=n_causal)
        g = g + bh @ (X == 1).astype(np.float64)
    vg = g.var()
    s2 = vg * (1 - h2) / h2 if vg > 0 else 1.0
    y = g + rng.normal(scale=np.sqrt(s2), size=n)
    return y, causal, beta


def sampler_cfg(n, p, n_causal, types="A", seed=1, chain=0, **over):
    """Plain PMV settings of the benchmark: e_qg = #causal,
    var_qg = #causal^2, the reference example's move settings, no delayed
    rejection unless `delay_rejection` is given (the reference example uses
    10)."""
    e_qg = max(float(n_causal), 1.0)
    var_qg = max(float(n_causal) ** 2, 2.0 * e_qg)
    cfg = dict(n=n, m_g=p, m_e=0, types=types, seed=seed, chain=chain,
               do_n_iter=100000, n_rao=500, n_rao_burnin=20, thin=10,
               n_sample_tau2_and_missing=10, delay_rejection=0, max_move_size=20,
               p_move_size=0.2, p_move_size_nbs=0.7, p_move_size_nbc=0.25,
               max_SNP_neighborhood_size=20, adapt_p_move_size=1, p_move_size_acpt_goal=0.0,
               adaptation=0, flat_proposal_dist=0, use_individual_tau2=1,
               types_prior=[1, 1, 1, 1, 1], e_qg=e_qg, var_qg=var_qg, R2mode_sigma2=0.2,
               nu_sigma2=1, nu_tau2=[5, 5, 5, 5], s2_tau2=[0.05, 0.05, 0.05, 0.05],
               eh_tau2=[0, 0, 0, 0], mu_alpha=1.0, inv_tau2_e_const_val=0.0,
               inv_tau2_e_val=0.001)
    cfg.update(over)
    return cfg
