"""gyrobias_propagate (gyrobias_propagate/kernel.cu:8-120) on the GPU against the
oracle's Euler step of gyro_bias.m's integrate() (the refcpu 3-D derivative
with no damping, gyroscopic or gravity term plus the gyro random walk)."""
import numpy as np
import pytest

import hpup
import oracle as O
from helpers import density

pytestmark = pytest.mark.gpu


def _print(*a):
    print(*a, flush=True)


@pytest.mark.parametrize("B,N", [(3, 4), (4, 6)])
def test_gyrobias_step_vs_oracle(B, N):
    _print("case gyrobias", B, N)
    L = 2.0
    f = density(B, 3, N, L, (4, 4, 4), seed=3, Sigma=np.eye(3) * 0.04, mu=np.array([0, 0, 0.2]))
    d = O.wigner_table(B)
    _, _, _, w = O.grid(B)
    F = O.fft_forward(B, f, d, w, 3, N)
    X = O.build_X(3, N, L)
    u = O.getu(B - 1)
    G1 = 0.5 * np.diag([0.1, 0.12, 0.08]) ** 2
    G2 = 0.5 * np.diag([0.05, 0.05, 0.06]) ** 2
    G2[0, 1] = G2[1, 0] = 1e-4
    dt = 0.01
    ref = O.gyrobias_step(B, N, F, X, dt, u, G1, G2)
    out = hpup.gyrobias_propagate(F, X, dt, u, G1, G2)
    M = 2 * B - 1
    lmax = B - 1
    mask = np.zeros((M, M, B), dtype=bool)  # valid (m, n, l) entries
    for l in range(B):
        mask[lmax - l:lmax + l + 1, lmax - l:lmax + l + 1, l] = True
    a, b = out[mask], ref[mask]
    assert np.linalg.norm(a - b) <= 1e-10 * np.linalg.norm(b)
    assert np.abs(out[~mask]).max() == 0.0
