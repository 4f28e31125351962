"""Periodic staggered operators (oracle side, CPU): each scheme pinned by its
order of accuracy on a smooth periodic function sampled at edges and cells,
which fixes every tap's offset and weight (a misplaced tap or a wrong weight
drops the order to 2 or less), and the exactly repeated edge plane."""
import numpy as np
import pytest

from oracle import pade_dense as pd

L = 2 * np.pi


def _sample(n, where):
    dz = L / n
    z = (np.arange(n) + (0.5 if where == "c" else 0.0)) * dz
    return z, dz


def _f(z):
    return np.sin(3 * z) + 0.5 * np.cos(2 * z + 0.3)


def _df(z):
    return 3 * np.cos(3 * z) - np.sin(2 * z + 0.3)


def _d2f(z):
    return -9 * np.sin(3 * z) - 2.0 * np.cos(2 * z + 0.3)


CASES = {
    "d1_e2c": ("e", "c", _df), "d1_c2e": ("c", "e", _df),
    "d2_c2c": ("c", "c", _d2f), "d2_e2e": ("e", "e", _d2f),
    "interp_e2c": ("e", "c", _f), "interp_c2e": ("c", "e", _f),
}


@pytest.mark.parametrize("op", sorted(CASES))
def test_sixth_order(op):
    src, dst, exact = CASES[op]
    errs = []
    for n in (64, 128):
        zi, dz = _sample(n, src)
        zo, _ = _sample(n, dst)
        got = pd.stagg_operator(op, n, dz) @ _f(zi)
        errs.append(np.abs(got - exact(zo)).max())
    order = np.log2(errs[0] / errs[1])
    assert order > 5.5, (op, errs, order)
    assert errs[1] < 1e-7


def test_edge_plane_repeats():
    rng = np.random.default_rng(1)
    n = 12
    fc = np.asfortranarray(rng.standard_normal((3, 2, n)))
    out = pd.stagg_apply("d1_c2e", fc, 0.1)
    assert out.shape == (3, 2, n + 1)
    assert np.array_equal(out[:, :, n], out[:, :, 0])
    fe = np.concatenate([fc, fc[:, :, :1]], axis=2)
    assert pd.stagg_apply("interp_e2c", fe, 0.1).shape == (3, 2, n)
