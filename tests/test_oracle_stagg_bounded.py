"""Bounded staggered operators (oracle side, CPU).

The oracle restates the reference's right-hand-side files and tridiagonal
set-up statement by statement (src/derivatives/STAGG_CD06_files/).  These
checks pin that restatement against what the mathematics fixes:

  * even / odd ends: each closure row is the interior scheme applied to the
    mirror image of the data about the wall, so the bounded operator on
    parity-consistent data equals the periodic operator (itself pinned by
    its order of accuracy) on the mirror-extended data of period 4n,
    restricted to the domain -- every closure row of A and B, for every
    combination of even and odd ends, is fixed by this;
  * sided ends: the one-sided closures are exact on polynomials up to the
    degree their construction gives (cubic for the edge-to-cell and
    cell-to-edge derivative and the edge-to-cell interpolation, quartic for
    the cell derivative, quadratic for the cell-to-edge interpolation); a
    wrong coefficient or index breaks that exactness;
  * operators without a sided closure (edge first derivative, both second
    derivatives) ignore the sided flags, as the reference does.
"""
import numpy as np
import pytest

from oracle import pade_dense as pd

PARITIES = [(True, True), (True, False), (False, True), (False, False)]   # (bot even, top even)
PERIODIC_OF = {"d1_e2c": ("stagg", "d1_e2c"), "d1_c2e": ("stagg", "d1_c2e"),
               "d2_c2c": ("stagg", "d2_c2c"), "d2_e2e": ("stagg", "d2_e2e"),
               "interp_e2c": ("stagg", "interp_e2c"), "interp_c2e": ("stagg", "interp_c2e"),
               "d1_c2c": ("cd06", "d1"), "d1_e2e": ("cd06", "d1")}


def _extend(f, edge, sb, st):
    """Mirror extension of a cell (n) or edge (n + 1) array to one period of
    4n points: walls at the cell faces (cells) or on the end edges (edges)."""
    n = len(f) - (1 if edge else 0)
    N = 4 * n
    out = np.empty(N)
    for j in range(-2 * n, 2 * n):
        if edge:
            if 0 <= j <= n:
                v = f[j]
            elif n < j < 2 * n:
                v = st * f[2 * n - j]
            elif -n <= j < 0:
                v = sb * f[-j]
            else:
                v = sb * st * f[2 * n + j]
        else:
            if 0 <= j < n:
                v = f[j]
            elif n <= j < 2 * n:
                v = st * f[2 * n - 1 - j]
            elif -n <= j < 0:
                v = sb * f[-1 - j]
            else:
                v = sb * st * f[2 * n + j]
        out[j % N] = v
    return out


def _periodic(op, N, dx):
    fam, which = PERIODIC_OF[op]
    if fam == "stagg":
        return pd.stagg_operator(which, N, dx)
    return pd.dense_operator(fam, which, N, dx, periodic=True)


@pytest.mark.parametrize("op", pd.STAGG_OPS_BOUNDED)
@pytest.mark.parametrize("par", PARITIES)
@pytest.mark.parametrize("n", [5, 6, 12, 33])
def test_mirror_ends_equal_periodic_extension(op, par, n):
    rng = np.random.default_rng(n)
    dx = 0.37
    edge_in, edge_out = pd.STAGG_EDGE_IN[op], pd.STAGG_EDGE_OUT[op]
    sb, st = (1.0 if par[0] else -1.0), (1.0 if par[1] else -1.0)
    f = rng.standard_normal(n + 1 if edge_in else n)
    if edge_in:             # an odd edge field vanishes on the wall
        if sb < 0:
            f[0] = 0.0
        if st < 0:
            f[n] = 0.0
    D = pd.stagg_bounded_operator(op, n, dx, (par[0], False), (par[1], False))
    got = D @ f
    ext = _periodic(op, 4 * n, dx) @ _extend(f, edge_in, sb, st)
    want = ext[: (n + 1 if edge_out else n)]
    np.testing.assert_allclose(got, want, rtol=0, atol=1e-12 * max(1.0, np.abs(want).max()))


def test_mirror_extension_helper_is_a_mirror():
    # the helper itself: an even cell field extends to an even periodic one
    f = np.arange(1.0, 6.0)
    e = _extend(f, False, 1.0, 1.0)
    assert np.array_equal(e[:5], f) and np.array_equal(e[5:10], f[::-1])
    assert np.array_equal(e[-5:], f[::-1])
    g = np.array([0.0, 1.0, 2.0, 3.0])      # odd-odd edge field, n = 3
    e = _extend(g, True, -1.0, -1.0)
    assert np.array_equal(e[:4], g) and e[4] == -2.0 and e[-1] == -1.0


SIDED_DEGREE = {"d1_e2c": 3, "d1_c2e": 3, "d1_c2c": 4, "interp_e2c": 3, "interp_c2e": 2}


@pytest.mark.parametrize("op", sorted(SIDED_DEGREE))
@pytest.mark.parametrize("n", [6, 24])
def test_sided_closures_exact_on_polynomials(op, n):
    dx = 1.0 / n
    zE = np.arange(n + 1) * dx
    zC = (np.arange(n) + 0.5) * dx
    zin = zE if pd.STAGG_EDGE_IN[op] else zC
    zout = zE if pd.STAGG_EDGE_OUT[op] else zC
    S = (True, True)
    D = pd.stagg_bounded_operator(op, n, dx, S, S)
    for d in range(SIDED_DEGREE[op] + 1):
        f = (zin - 0.3) ** d
        if op.startswith("d1"):
            want = d * (zout - 0.3) ** (d - 1) if d else np.zeros_like(zout)
        else:
            want = (zout - 0.3) ** d
        np.testing.assert_allclose(D @ f, want, rtol=0, atol=1e-10 * max(1.0, np.abs(want).max()))
    # one degree more is not exact: the closure rows are what is tested
    f = (zin - 0.3) ** (SIDED_DEGREE[op] + 1)
    d = SIDED_DEGREE[op] + 1
    want = d * (zout - 0.3) ** (d - 1) if op.startswith("d1") else (zout - 0.3) ** d
    assert np.abs(D @ f - want).max() > 1e-8


@pytest.mark.parametrize("op", ["d1_e2e", "d2_c2c", "d2_e2e"])
@pytest.mark.parametrize("par", PARITIES)
def test_sided_flags_ignored_without_sided_closure(op, par):
    a = pd.stagg_bounded_operator(op, 10, 0.5, (par[0], False), (par[1], False))
    b = pd.stagg_bounded_operator(op, 10, 0.5, (par[0], True), (par[1], True))
    assert np.array_equal(a, b)


@pytest.mark.parametrize("op", sorted(SIDED_DEGREE))
def test_mixed_ends_rows_come_from_each_end(op):
    # sided at the bottom, mirror at the top: the bottom rows are those of the
    # both-sided operator, the top rows those of the both-mirror operator
    n = 16
    A_s, B_s = pd.stagg_bounded_matrices(op, n, 0.2, (True, True), (True, True))
    A_m, B_m = pd.stagg_bounded_matrices(op, n, 0.2, (True, False), (True, False))
    A_x, B_x = pd.stagg_bounded_matrices(op, n, 0.2, (True, True), (True, False))
    m = A_x.shape[0]
    assert np.array_equal(A_x[:4], A_s[:4]) and np.array_equal(B_x[:4], B_s[:4])
    assert np.array_equal(A_x[m - 4:], A_m[m - 4:]) and np.array_equal(B_x[m - 4:], B_m[m - 4:])
