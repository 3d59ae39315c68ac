"""pendulum_plot_getc on the GPU (marginal, SO(3) transform and point
evaluation through the C ABI) against the oracle's getc (getc.cpp:7-40) on a
random density and random Euler angles, and against the closed form of a
degree-one marginal at the reduced model's full grid (B = 15, 32^2)."""
import numpy as np
import pytest

import hpup
import oracle as O
from helpers import density
from test_getc import closed_form_c, degree_one_density, random_euler

pytestmark = pytest.mark.gpu


def _print(*a):
    print(*a, flush=True)


@pytest.mark.parametrize("dim,B,N", [(2, 5, 6), (3, 4, 4)])
def test_getc_vs_oracle(dim, B, N):
    _print("case getc vs oracle", dim, B, N)
    L = 8.0
    f = density(B, dim, N, L, (8, 4, 2), seed=7)
    e = random_euler(np.random.default_rng(2), 6, 4)
    c = hpup.pendulum_plot_getc(f, e, L)
    ref = O.getc(f, e, L)
    assert np.abs(c - ref).max() <= 1e-12 * np.abs(ref).max()


def test_getc_degree_one_full_size():
    _print("case getc degree one B=15 N=32")
    B, N, L = 15, 32, 8.0
    f, A, H = degree_one_density(B, 2, N, L, seed=11)
    e = random_euler(np.random.default_rng(4), 50, 40)
    c = hpup.pendulum_plot_getc(f, e, L)
    ref = closed_form_c(e, A, H)
    assert np.abs(c - ref).max() <= 1e-11 * np.abs(ref).max()


def test_getc_empty_angles():
    _print("case getc empty")
    f = density(3, 2, 4, 6.0, (4, 2, 1), seed=1)
    c = hpup.pendulum_plot_getc(f, np.zeros((3, 0, 5, 3)), 6.0)
    assert c.shape == (0, 3)
