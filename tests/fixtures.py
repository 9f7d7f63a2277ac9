"""Fixture datasets and known answers transcribed from the reference's tests.

Data files are the reference's own test fixtures (tests/golden/*.bed, *.fam);
the expected matrices are the literals of src/tests/data_tests.hpp."""
import os

import numpy as np

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")

# src/tests/data_tests.hpp:40-47 (rows = SNP, n = 30)
SMALL_MODELSPACE_G = np.array([
    [2,1,0,2,1,2,2,1,2,2,1,0,2,1,1,0,2,2,1,2,2,2,1,2,2,2,2,2,2,1],
    [1,2,2,2,2,1,2,2,1,2,2,2,1,2,2,2,2,2,0,2,2,2,2,2,2,2,2,1,1,2],
    [1,1,1,1,1,1,1,2,1,2,1,1,1,1,1,0,0,0,0,1,1,0,1,2,2,2,1,2,2,2],
    [2,2,2,2,2,2,2,2,1,2,2,2,0,2,2,1,2,2,2,1,2,2,2,1,1,2,2,2,2,2],
    [2,0,2,0,1,2,2,1,1,1,1,1,1,0,1,1,0,2,1,1,1,2,2,0,2,1,2,1,1,2],
    [2,1,2,2,2,1,0,1,0,2,2,2,0,2,2,2,1,2,2,2,2,1,1,0,2,2,2,2,2,1],
    [0,1,1,1,1,1,1,1,1,0,1,1,1,1,1,1,1,1,2,1,2,1,2,2,1,1,1,0,2,1]])

# src/tests/data_tests.hpp:49-79
SMALL_MODELSPACE_Y = np.array([
    0.290913, 1.72566, -1.14075, -0.775, 0.283488, -0.0173902, 0.359494,
    0.91743, 1.17233, 0.0284589, -0.3621, -0.63587, -0.518061, -0.175326,
    -0.0953311, 1.88926, -0.311815, -0.935637, 0.342461, 0.401009, 0.348721,
    1.65077, 1.53112, 2.58911, -0.162152, -0.196501, -0.719778, 0.384602,
    1.43458, 0.207898])

# src/tests/data_tests.hpp:130-140 (rows = SNP, n = 5, -1 = missing)
PLINKTEST_G = np.array([
    [1,2,2,-1,0], [1,1,1,2,0], [1,1,2,1,1], [0,-1,0,1,2], [0,2,2,2,2],
    [1,2,1,1,0], [2,1,0,2,1], [1,0,1,2,1], [1,0,1,1,0], [0,1,1,1,1]])


def load_fixture(name, n, p):
    from oracle import read_bed, read_fam_pheno
    body = read_bed(os.path.join(GOLDEN, name + ".bed"), n, p)
    y = read_fam_pheno(os.path.join(GOLDEN, name + ".fam"))
    return body, y


def plinktest():
    return load_fixture("plinktest", 5, 10)


def small_modelspace():
    return load_fixture("small_modelspace", 30, 7)


def pack_genotypes(G):
    """G: p x n matrix in {-1,0,1,2} -> packed body (reference code mapping)."""
    code = {0: 0, -1: 1, 1: 2, 2: 3}
    p, n = G.shape
    stride = (n + 3) // 4
    body = np.zeros(p * stride, dtype=np.uint8)
    for j in range(p):
        for i in range(n):
            body[j * stride + i // 4] |= code[int(G[j, i])] << (2 * (i % 4))
    return body


# src/tests/testdata/small_modelspace_tau2.ini
SMALL_MODELSPACE_TAU2_INI = dict(
    n=30, m_g=7, m_e=0, recode_g_to_minor_allele_count=1,
    do_n_iter=20000000, n_rao=250, n_rao_burnin=20000, adaptation=0,
    delay_rejection=20, verbosity=100000, thin=50, save_beta=0,
    n_sample_tau2_and_missing=10, seed=1245, types="A",
    types_prior=[1, 1, 1, 1, 1], e_qg=2, var_qg=2, use_individual_tau2=0,
    R2mode_sigma2=0.20, nu_sigma2=1, nu_tau2=[3, 3, 3, 3],
    eh_tau2=[0.10, 0.10, 0.10, 0.10], s2_tau2=[0.02, 0, 0, 0],
    mu_alpha=1.0, inv_tau2_e_const_val=0, inv_tau2_e_val=1)

# src/tests/testdata/plinktest.ini
PLINKTEST_INI = dict(
    n=5, m_g=10, m_e=0, recode_g_to_minor_allele_count=0,
    do_n_iter=20000000, n_rao=250, n_rao_burnin=20000, adaptation=0,
    delay_rejection=20, thin=50, save_beta=0, n_sample_tau2_and_missing=10,
    seed=1245, types="A,AH", types_prior=[1, 1, 1, 1, 1], e_qg=2, var_qg=2,
    use_individual_tau2=0, R2mode_sigma2=0.20, nu_sigma2=1,
    nu_tau2=[3, 3, 3, 3], eh_tau2=[0.10, 0.10, 0.10, 0.10],
    s2_tau2=[0.02, 0, 0, 0], mu_alpha=1.0, inv_tau2_e_const_val=0,
    inv_tau2_e_val=1)
