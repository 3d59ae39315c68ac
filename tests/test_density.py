"""Initial densities (hpup.density, torch on the CPU here): the drivers'
matrix-Fisher x Gaussian component and the weighted mixture of components
(the bimodal initial density of the 3-D hybrid workload)."""
import numpy as np
import pytest

from hpup import density as D


def test_mixture_is_the_weighted_sum_of_unit_mass_components():
    import torch
    B, dim, N, L = 3, 3, 6, 12.0
    rng = np.random.default_rng(4)
    comps = [dict(s=(12, 6, 2), U=D.random_rotation(rng), mu=(0.5, 0.0, -0.3)),
             dict(s=(12, 6, 2), U=D.random_rotation(rng), Sigma=np.diag([0.6, 0.9, 1.2]))]
    f = D.mixture_density(B, dim, N, L, comps, (0.6, 0.4), device="cpu")
    _, w = D.beta_weights(B)
    wj = np.broadcast_to(w[None, :, None], (2 * B,) * 3).reshape(-1, order="F")
    mass = lambda g: float((g.reshape(N ** dim, -1).numpy() * wj[None, :]).sum()) * (L / N) ** dim
    assert abs(mass(f) - 1.0) < 1e-12
    parts = []
    for c in comps:
        fR, fx = D.matrix_fisher_gaussian(B, dim, N, L, c["s"], c["U"], Sigma=c.get("Sigma"), mu=c.get("mu"),
                                          device="cpu")
        parts.append(D.grid_density(fR, fx))
        assert abs(mass(parts[-1]) - 1.0) < 1e-12
    assert torch.allclose(f, 0.6 * parts[0] + 0.4 * parts[1], rtol=1e-14, atol=0)
    with pytest.raises(ValueError):
        D.mixture_density(B, dim, N, L, comps, (1.0,), device="cpu")


def test_weights_on_a_single_component_is_an_error():
    with pytest.raises(ValueError, match="mixture_density"):
        D.matrix_fisher_gaussian(2, 2, 4, 4.0, (1, 1, 1), np.eye(3), weights=(0.5, 0.5), device="cpu")
