"""CPU checks of the measurement model's sampler (hpup.likelihood): the
attitude noise Z = R Q, Q ~ MF(kappa I), has the matrix-Fisher moments
E[Q_ii] = (dc/ds_i)/c at s = (kappa, kappa, kappa) (the normalizer and its
gradient from pdf_MF_normal, pinned by Haar quadrature), and Z R^T is
independent of R."""
import numpy as np

import hpup
from hpup.likelihood import sample_mf_isotropic


def test_isotropic_matrix_fisher_sampler_moments():
    rng = np.random.default_rng(0)
    kappa, n = 20.0, 40000
    R = np.repeat(np.eye(3)[None], n, axis=0)
    Q = sample_mf_isotropic(rng, R, kappa)
    c, dc, _ = hpup.pdf_MF_normal(np.array([kappa] * 3), bool_scaled=False)
    mom = dc / c  # E[Q_ii]
    est = np.einsum("nii->ni", Q).mean(axis=0)
    se = np.einsum("nii->ni", Q).std(axis=0) / np.sqrt(n)
    assert np.all(np.abs(est - mom) < 5 * se + 1e-4), (est, mom, se)
    # orthogonality and det = 1
    assert np.abs(np.einsum("nij,nkj->nik", Q, Q) - np.eye(3)).max() < 1e-12
    assert np.abs(np.linalg.det(Q) - 1).max() < 1e-12


def test_sampler_is_equivariant():
    rng1, rng2 = np.random.default_rng(3), np.random.default_rng(3)
    n = 1000
    R0 = hpup.psvd(np.random.default_rng(9).standard_normal((3, 3)))[0]
    Q = sample_mf_isotropic(rng1, np.repeat(np.eye(3)[None], n, axis=0), 20.0)
    Z = sample_mf_isotropic(rng2, np.repeat(R0[None], n, axis=0), 20.0)
    assert np.abs(Z - np.einsum("ij,njk->nik", R0, Q)).max() < 1e-12
