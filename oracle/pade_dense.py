"""Dense-matrix CPU oracle for the compact-difference operator path.

TEST INFRASTRUCTURE ONLY.  Nothing in the product (``padeops_amd``) may import
this module; only ``tests/``, ``__graft_entry__.smoke()`` and the CPU-baseline
leg of ``bench.py`` use it, and only as the checker.

Every operator is written as ``A y = B f`` with ``A`` banded (cyclic when
periodic) and ``B`` an explicit stencil matrix, both assembled row by row from
the coefficient formulas of the reference and solved with LAPACK.  No banded
factorisation is involved, so this oracle is independent of the LU / Thomas /
partitioned solvers it checks.

Row conventions are 0-based: reference row ``i`` (1-based) is row ``i-1`` here.

Pinning (tests/test_oracle_*.py):
  * periodic rows  -> closed-form modified wavenumbers / transfer functions
    (``src/utilities/spectral.F90:938-1010``, ``tests/test_cf90.F90:108-116``),
  * bc = +-1 rows  -> a generic mirror-extension construction written
    independently in the tests (the reference's explicit rows must equal it),
  * bc = 0 rows    -> polynomial exactness of every closure row and the
    discrete conservation identity (column sums of the weighted B vanish),
  * cd06 second derivative -> the reference has no implementation
    (``src/derivatives/cd06.F90:330-343,492-513,745-773,841-853``); the
    operator defined here is pinned only by its closed form ("parity unpinned"
    with respect to reference output),
  * periodic staggered operators (``src/derivatives/cd06stagg.F90``) -> sixth
    order of accuracy on smooth periodic data at edges and cells,
  * bounded staggered operators (``STAGG_CD06_files/``) -> even / odd ends
    equal the periodic operator on the mirror-extended data (period 4n);
    one-sided ends exact on polynomials to the degree of their construction,
  * periodic Gaussian filter (``src/filters/gaussian.F90``) -> its transfer
    function agf + 2 sum_m c_m cos(m k) on every Fourier mode,
  * complex z filter (``src/filters/cf90.F90:1230-1251``) -> composition of
    the pinned A (0, 0) and B (bc1, bcn); its periodic case is unpinned (the
    reference reads an unallocated factorization there).
"""
from __future__ import annotations

import numpy as np

# ----------------------------------------------------------------------------
# Coefficients (restated from the reference's parameter expressions)
# ----------------------------------------------------------------------------


def cd06_coeffs():
    """cd06 coefficients, ``src/derivatives/cd06.F90:14-59``."""
    c = {}
    c["alpha_d1"] = 1.0 / 3.0
    c["a_d1"] = (14.0 / 9.0) / 2.0          # multiplies (f[i+1]-f[i-1])
    c["b_d1"] = (1.0 / 9.0) / 4.0           # multiplies (f[i+2]-f[i-2])
    c["alpha_d2"] = 2.0 / 11.0
    c["a_d2"] = 12.0 / 11.0                 # multiplies (f[i+1]-2f[i]+f[i-1])
    c["b_d2"] = (3.0 / 11.0) / 4.0          # multiplies (f[i+2]-2f[i]+f[i-2])
    # one-sided closure (Lele): cd06.F90:29-59
    alpha, p, q, r, s = 3.0, 17.0 / 6.0, 1.5, 1.5, -1.0 / 6.0
    ah = c["alpha_d1"]
    q_p, alpha_p = 0.75, 0.25
    alpha_pp = (((40 * ah - 1) * q + 7 * (4 * ah - 1) * s)
                / (16 * (ah + 2) * q + 8 * (1 - 4 * ah) * s))
    q_pp = (alpha_pp + 2) / 3.0
    r_pp = (4 * alpha_pp - 1) / 12.0
    w1 = (2 * ah + 1) / (2 * (q + s))
    w2 = ((8 * ah + 7) * q - 6 * (2 * ah + 1) * r + (8 * ah + 7) * s) / (9 * (q + s))
    w3 = (4 * (ah + 2) * q + 2 * (1 - 4 * ah) * s) / (9 * (q + s))
    c.update(dict(alpha_b=alpha, p=p, q=q, r=r, s=s, q_p=q_p, alpha_p=alpha_p,
                  alpha_pp=alpha_pp, q_pp=q_pp, r_pp=r_pp, w1=w1, w2=w2, w3=w3))
    return c


def cd10_coeffs():
    """cd10 coefficients, ``src/derivatives/cd10.F90:16-104``."""
    c = {}
    c["alpha_d1"] = 1.0 / 2.0
    c["beta_d1"] = 1.0 / 20.0
    c["a_d1"] = (17.0 / 12.0) / 2.0
    c["b_d1"] = (101.0 / 150.0) / 4.0
    c["c_d1"] = (1.0 / 100.0) / 6.0
    c["alpha_d2"] = 334.0 / 899.0
    c["beta_d2"] = 43.0 / 1798.0
    c["a_d2"] = 1065.0 / 1798.0
    c["b_d2"] = (1038.0 / 899.0) / 4.0
    c["c_d2"] = (79.0 / 1798.0) / 9.0
    # first-derivative closures, cd10.F90:35-80
    alpha, p, q, r, s = 3.0, -17.0 / 6.0, 1.5, 1.5, -1.0 / 6.0
    qh, rh, sh = c["a_d1"], c["b_d1"], c["c_d1"]
    q_p, alpha_p = 0.75, 0.25
    alpha_ppp = (8 * rh - 175 * sh) / (18 * rh - 550 * sh)
    beta_ppp = (1.0 / 20.0) * (-3 + 8 * alpha_ppp)
    q_ppp = (1.0 / 12.0) * (12 - 7 * alpha_ppp)
    r_ppp = (1.0 / 600.0) * (568 * alpha_ppp - 183)
    s_ppp = (1.0 / 300.0) * (9 * alpha_ppp - 4)
    t = (s * (rh + 2 * sh) - q * (qh + rh + sh))
    u = (q + s) * (qh + rh - sh * (q_ppp / s_ppp - 1))
    alpha_pp = (17 * t / (72 * u) - 8.0 / 9.0) / (19 * t / (24 * u) - 1.0 / 3.0)
    beta_pp = (1.0 / 12.0) * (-1 + 3 * alpha_pp)
    q_pp = (2.0 / 18.0) * (8 - 3 * alpha_pp)
    r_pp = (1.0 / 72.0) * (-17 + 57 * alpha_pp)
    w1 = (qh + 2 * rh + 3 * sh) / (q + s)
    w2 = (1 / q_p) * (rh + sh * (1 + q_ppp / s_ppp) - r * (qh + 2 * rh + 3 * sh) / (q + s))
    w3 = (qh + rh + sh * (1 - q_ppp / s_ppp)) / r_pp
    w4 = sh / s_ppp
    c.update(dict(alpha_b=alpha, p=p, q=q, r=r, s=s, q_p=q_p, alpha_p=alpha_p,
                  alpha_ppp=alpha_ppp, beta_ppp=beta_ppp, q_ppp=q_ppp, r_ppp=r_ppp,
                  s_ppp=s_ppp, alpha_pp=alpha_pp, beta_pp=beta_pp, q_pp=q_pp,
                  r_pp=r_pp, w1=w1, w2=w2, w3=w3, w4=w4))
    # second-derivative closures, cd10.F90:85-104
    b1_alpha = 11.0
    c["b1_alpha"] = b1_alpha
    c["b1"] = [(11 * b1_alpha + 35) / 12, -(5 * b1_alpha + 26) / 3, (b1_alpha + 19) / 2,
               (b1_alpha - 14) / 3, (11 - b1_alpha) / 12]
    c["b2_alpha"] = 1.0 / 10.0
    c["b2_a"] = 4 * (1 - c["b2_alpha"]) / 3
    b3_alpha = 344.0 / 1179.0
    c["b3_alpha"] = b3_alpha
    c["b3_beta"] = (38.0 * b3_alpha - 9.0) / 214.0
    c["b3_a"] = (696 - 1191 * b3_alpha) / 428
    c["b3_b"] = ((2454 * b3_alpha - 294) / 535) / 4.0
    return c


def cf90_coeffs():
    """cf90 coefficients (5-significant-digit decimals), ``src/filters/cf90.F90:16-48``."""
    return dict(alpha=0.66624, beta=0.16688, a=0.99965, b=0.66652, c=0.16674, d=4.0e-5,
                e=-5.0e-6,
                b2_alpha=0.4997, b2_a=0.9997, b2_b=0.49985,
                b3_alpha=0.66624, b3_beta=0.16688, b3_a=0.99952, b3_b=0.66656, b3_c=0.1668,
                b4_alpha=0.66624, b4_beta=0.16688, b4_a=0.99968, b4_b=0.66652, b4_c=0.16672,
                b4_d=4.0e-5)


# ----------------------------------------------------------------------------
# Interior stencils: offset -> coefficient (before 1/dx scaling)
# ----------------------------------------------------------------------------


def interior(kind, which):
    """Interior (A band, B stencil) of each operator.

    A band is a dict offset->coef; B stencil a dict offset->coef.
    """
    if kind == "cd06" and which == "d1":
        c = cd06_coeffs()
        al = c["alpha_d1"]
        A = {-1: al, 0: 1.0, 1: al}
        B = {1: c["a_d1"], -1: -c["a_d1"], 2: c["b_d1"], -2: -c["b_d1"]}
        return A, B
    if kind == "cd06" and which == "d2":
        c = cd06_coeffs()
        al = c["alpha_d2"]
        A = {-1: al, 0: 1.0, 1: al}
        a, b = c["a_d2"], c["b_d2"]
        B = {0: -2 * (a + b), 1: a, -1: a, 2: b, -2: b}
        return A, B
    if kind == "cd10" and which == "d1":
        c = cd10_coeffs()
        al, be = c["alpha_d1"], c["beta_d1"]
        A = {-2: be, -1: al, 0: 1.0, 1: al, 2: be}
        B = {}
        for m, k in ((1, "a_d1"), (2, "b_d1"), (3, "c_d1")):
            B[m] = c[k]
            B[-m] = -c[k]
        return A, B
    if kind == "cd10" and which == "d2":
        c = cd10_coeffs()
        al, be = c["alpha_d2"], c["beta_d2"]
        A = {-2: be, -1: al, 0: 1.0, 1: al, 2: be}
        B = {0: -2 * (c["a_d2"] + c["b_d2"] + c["c_d2"])}
        for m, k in ((1, "a_d2"), (2, "b_d2"), (3, "c_d2")):
            B[m] = c[k]
            B[-m] = c[k]
        return A, B
    if kind == "cf90" and which == "filter":
        c = cf90_coeffs()
        A = {-2: c["beta"], -1: c["alpha"], 0: 1.0, 1: c["alpha"], 2: c["beta"]}
        B = {0: c["a"]}
        for m, k in ((1, "b"), (2, "c"), (3, "d"), (4, "e")):
            B[m] = c[k]
            B[-m] = c[k]
        return A, B
    if kind == "gaussian" and which == "filter":
        # explicit Gaussian filter, src/filters/gaussian.F90:11-17 (periodic
        # stencil :137-185): A = I
        g = gaussian_coeffs()
        B = {0: g[0]}
        for m in range(1, 5):
            B[m] = B[-m] = g[m]
        return {0: 1.0}, B
    raise ValueError(f"unknown operator {kind}/{which}")


def gaussian_coeffs():
    """agf..egf of ``src/filters/gaussian.F90:11-17`` (offsets 0..4)."""
    return (3565.0 / 10368.0, 3091.0 / 12960.0, 1997.0 / 25920.0, 149.0 / 12960.0, 107.0 / 103680.0)


def scale_power(kind, which):
    return {"d1": 1, "d2": 2, "filter": 0}[which]


# ----------------------------------------------------------------------------
# Boundary rows of the bounded operators, restated row by row
# Rows are transcribed with the reference's 1-based indices through ``_acc``;
# each entry is (row0, {col0: coef}).  Right-end rows follow by reflection
# i -> n-1-i, with the sign of each reference routine (odd operators negate
# the mirrored RHS rows).
# ----------------------------------------------------------------------------


def _acc(*terms):
    """Sum (coef, col_1based) terms into a {col_0based: coef} row."""
    row = {}
    for coef, col in terms:
        row[col - 1] = row.get(col - 1, 0.0) + coef
    return row


def _left_rows_cd06_d1(bc_init1):
    """cd06 bounded: LHS rows 1-3 only for init bc1 == 0 (cd06.F90:279-292),
    otherwise the interior rows stay ("Incomplete", cd06.F90:288-292);
    RHS rows 1-3 always use the one-sided closure (cd06.F90:563-571)."""
    c = cd06_coeffs()
    w1, w2, w3 = c["w1"], c["w2"], c["w3"]
    A = []
    if bc_init1 == 0:
        A = [(0, _acc((w1, 1), (w1 * c["alpha_b"], 2))),
             (1, _acc((w2 * c["alpha_p"], 1), (w2, 2), (w2 * c["alpha_p"], 3))),
             (2, _acc((w3 * c["alpha_pp"], 2), (w3, 3), (w3 * c["alpha_pp"], 4)))]
    a1, b1, c1, d1 = w1 * (-c["p"]), w1 * c["q"], w1 * c["r"], w1 * c["s"]
    a2 = w2 * c["q_p"]
    a3, b3 = w3 * c["q_pp"], w3 * c["r_pp"]
    B = [(0, _acc((a1, 1), (b1, 2), (c1, 3), (d1, 4))),
         (1, _acc((a2, 3), (-a2, 1))),
         (2, _acc((a3, 4), (-a3, 2), (b3, 5), (-b3, 1)))]
    return A, B, -1      # right rows: negated mirror, cd06.F90:577-585


def _left_rows_cd10_d1(bc1):
    c = cd10_coeffs()
    ah, bh = c["alpha_d1"], c["beta_d1"]
    a, b, cc = c["a_d1"], c["b_d1"], c["c_d1"]
    if bc1 == 0:      # cd10.F90:450-473 (LHS), :1164-1176 (RHS)
        w1, w2, w3, w4 = c["w1"], c["w2"], c["w3"], c["w4"]
        A = [(0, _acc((w1, 1), (w1 * c["alpha_b"], 2))),
             (1, _acc((w2 * c["alpha_p"], 1), (w2, 2), (w2 * c["alpha_p"], 3))),
             (2, _acc((w3 * c["beta_pp"], 1), (w3 * c["alpha_pp"], 2), (w3, 3),
                      (w3 * c["alpha_pp"], 4), (w3 * c["beta_pp"], 5))),
             (3, _acc((w4 * c["beta_ppp"], 2), (w4 * c["alpha_ppp"], 3), (w4, 4),
                      (w4 * c["alpha_ppp"], 5), (w4 * c["beta_ppp"], 6)))]
        a1, b1, c1, d1 = w1 * c["p"], w1 * c["q"], w1 * c["r"], w1 * c["s"]
        a2 = w2 * c["q_p"]
        a3, b3 = w3 * c["q_pp"], w3 * c["r_pp"]
        a4, b4, c4 = w4 * c["q_ppp"], w4 * c["r_ppp"], w4 * c["s_ppp"]
        B = [(0, _acc((a1, 1), (b1, 2), (c1, 3), (d1, 4))),
             (1, _acc((a2, 3), (-a2, 1))),
             (2, _acc((a3, 4), (-a3, 2), (b3, 5), (-b3, 1))),
             (3, _acc((a4, 5), (-a4, 3), (b4, 6), (-b4, 2), (c4, 7), (-c4, 1)))]
        return A, B, -1
    if bc1 == 1:      # cd10.F90:474-487 (LHS), :1177-1191 (RHS)
        A = [(0, _acc((1.0, 1))),
             (1, _acc((ah, 1), (1.0 - bh, 2), (ah, 3), (bh, 4)))]
        B = [(0, {}),
             (1, _acc((a, 3), (-a, 1), (b, 4), (-b, 2), (cc, 5), (-cc, 3))),
             (2, _acc((a, 4), (-a, 2), (b, 5), (-b, 1), (cc, 6), (-cc, 2))),
             (3, _acc((a, 5), (-a, 3), (b, 6), (-b, 2), (cc, 7), (-cc, 1)))]
        return A, B, -1
    if bc1 == -1:     # cd10.F90:489-500 (LHS), :1192-1206 (RHS)
        A = [(0, _acc((1.0, 1), (2 * ah, 2), (2 * bh, 3))),
             (1, _acc((ah, 1), (1.0 + bh, 2), (ah, 3), (bh, 4)))]
        B = [(0, _acc((a, 2), (a, 2), (b, 3), (b, 3), (cc, 4), (cc, 4))),
             (1, _acc((a, 3), (-a, 1), (b, 4), (b, 2), (cc, 5), (cc, 3))),
             (2, _acc((a, 4), (-a, 2), (b, 5), (-b, 1), (cc, 6), (cc, 2))),
             (3, _acc((a, 5), (-a, 3), (b, 6), (-b, 2), (cc, 7), (-cc, 1)))]
        return A, B, -1
    raise ValueError("bc must be -1, 0 or 1")


def _left_rows_cd10_d2(bc1):
    c = cd10_coeffs()
    al, be = c["alpha_d2"], c["beta_d2"]
    a, b, cc = c["a_d2"], c["b_d2"], c["c_d2"]
    if bc1 == 0:      # cd10.F90:601-618 (LHS), :1658-1667 (RHS)
        A = [(0, _acc((1.0, 1), (c["b1_alpha"], 2))),
             (1, _acc((c["b2_alpha"], 1), (1.0, 2), (c["b2_alpha"], 3))),
             (2, _acc((c["b3_beta"], 1), (c["b3_alpha"], 2), (1.0, 3), (c["b3_alpha"], 4),
                      (c["b3_beta"], 5)))]
        b1 = c["b1"]
        a2 = c["b2_a"]
        a3, b3 = c["b3_a"], c["b3_b"]
        B = [(0, _acc((b1[0], 1), (b1[1], 2), (b1[2], 3), (b1[3], 4), (b1[4], 5))),
             (1, _acc((a2, 3), (-2 * a2, 2), (a2, 1))),
             (2, _acc((a3, 4), (-2 * a3, 3), (a3, 2), (b3, 5), (-2 * b3, 3), (b3, 1)))]
        return A, B, 1
    if bc1 == 1:      # cd10.F90:619-630 (LHS), :1668-1679 (RHS)
        A = [(0, _acc((1.0, 1), (2 * al, 2), (2 * be, 3))),
             (1, _acc((al, 1), (1.0 + be, 2), (al, 3), (be, 4)))]
        B = [(0, _acc((a, 2), (-2 * a, 1), (a, 2), (b, 3), (-2 * b, 1), (b, 3),
                      (cc, 4), (-2 * cc, 1), (cc, 4))),
             (1, _acc((a, 3), (-2 * a, 2), (a, 1), (b, 4), (-2 * b, 2), (b, 2),
                      (cc, 5), (-2 * cc, 2), (cc, 3))),
             (2, _acc((a, 4), (-2 * a, 3), (a, 2), (b, 5), (-2 * b, 3), (b, 1),
                      (cc, 6), (-2 * cc, 3), (cc, 2)))]
        return A, B, 1
    if bc1 == -1:     # cd10.F90:631-642 (LHS), :1680-1691 (RHS)
        A = [(0, _acc((1.0, 1))),
             (1, _acc((al, 1), (1.0 - be, 2), (al, 3), (be, 4)))]
        B = [(0, _acc((a, 2), (-2 * a, 1), (-a, 2), (b, 3), (-2 * b, 1), (-b, 3),
                      (cc, 4), (-2 * cc, 1), (-cc, 4))),
             (1, _acc((a, 3), (-2 * a, 2), (a, 1), (b, 4), (-2 * b, 2), (-b, 2),
                      (cc, 5), (-2 * cc, 2), (-cc, 3))),
             (2, _acc((a, 4), (-2 * a, 3), (a, 2), (b, 5), (-2 * b, 3), (b, 1),
                      (cc, 6), (-2 * cc, 3), (-cc, 2)))]
        return A, B, 1
    raise ValueError("bc must be -1, 0 or 1")


def _left_rows_cf90(bc1):
    c = cf90_coeffs()
    al, be = c["alpha"], c["beta"]
    a, b, cc, d, e = c["a"], c["b"], c["c"], c["d"], c["e"]
    if bc1 == 0:      # cf90.F90:295-320 (LHS), :702-715 (RHS)
        A = [(0, _acc((1.0, 1))),
             (1, _acc((c["b2_alpha"], 1), (1.0, 2), (c["b2_alpha"], 3))),
             (2, _acc((c["b3_beta"], 1), (c["b3_alpha"], 2), (1.0, 3), (c["b3_alpha"], 4),
                      (c["b3_beta"], 5))),
             (3, _acc((c["b4_beta"], 2), (c["b4_alpha"], 3), (1.0, 4), (c["b4_alpha"], 5),
                      (c["b4_beta"], 6)))]
        B = [(0, _acc((1.0, 1))),
             (1, _acc((c["b2_a"], 2), (c["b2_b"], 3), (c["b2_b"], 1))),
             (2, _acc((c["b3_a"], 3), (c["b3_b"], 4), (c["b3_b"], 2), (c["b3_c"], 5),
                      (c["b3_c"], 1))),
             (3, _acc((c["b4_a"], 4), (c["b4_b"], 5), (c["b4_b"], 3), (c["b4_c"], 6),
                      (c["b4_c"], 2), (c["b4_d"], 7), (c["b4_d"], 1)))]
        return A, B, 1
    if bc1 == 1:      # cf90.F90:321-332 (LHS), :716-736 (RHS)
        A = [(0, _acc((1.0, 1), (2 * al, 2), (2 * be, 3))),
             (1, _acc((al, 1), (1.0 + be, 2), (al, 3), (be, 4)))]
        B = [(0, _acc((a, 1), (b, 2), (b, 2), (cc, 3), (cc, 3), (d, 4), (d, 4), (e, 5), (e, 5))),
             (1, _acc((a, 2), (b, 3), (b, 1), (cc, 4), (cc, 2), (d, 5), (d, 3), (e, 6), (e, 4))),
             (2, _acc((a, 3), (b, 4), (b, 2), (cc, 5), (cc, 1), (d, 6), (d, 2), (e, 7), (e, 3))),
             (3, _acc((a, 4), (b, 5), (b, 3), (cc, 6), (cc, 2), (d, 7), (d, 1), (e, 8), (e, 2)))]
        return A, B, 1
    if bc1 == -1:     # cf90.F90:333-344 (LHS), :737-757 (RHS)
        A = [(0, _acc((1.0, 1))),
             (1, _acc((al, 1), (1.0 - be, 2), (al, 3), (be, 4)))]
        B = [(0, _acc((a, 1), (b, 2), (-b, 2), (cc, 3), (-cc, 3), (d, 4), (-d, 4), (e, 5), (-e, 5))),
             (1, _acc((a, 2), (b, 3), (b, 1), (cc, 4), (-cc, 2), (d, 5), (-d, 3), (e, 6), (-e, 4))),
             (2, _acc((a, 3), (b, 4), (b, 2), (cc, 5), (cc, 1), (d, 6), (-d, 2), (e, 7), (-e, 3))),
             (3, _acc((a, 4), (b, 5), (b, 3), (cc, 6), (cc, 2), (d, 7), (d, 1), (e, 8), (-e, 2)))]
        return A, B, 1
    raise ValueError("bc must be -1, 0 or 1")


def _left_rows_gaussian(bc1):
    """Bounded Gaussian filter, src/filters/gaussian.F90 (A = I): bc 0 rows
    :192-208 with the closure weights :19-36; bc 1 rows :209-231; bc -1 rows
    :232-254."""
    a, b, c, d, e = gaussian_coeffs()
    if bc1 == 0:
        B = [(0, _acc((5.0 / 6.0, 1), (1.0 / 6.0, 2))),
             (1, _acc((2.0 / 3.0, 2), (1.0 / 6.0, 3), (1.0 / 6.0, 1))),
             (2, _acc((31.0 / 64.0, 3), (7.0 / 32.0, 4), (7.0 / 32.0, 2), (5.0 / 128.0, 5), (5.0 / 128.0, 1))),
             (3, _acc((17.0 / 48.0, 4), (15.0 / 64.0, 5), (15.0 / 64.0, 3), (7.0 / 96.0, 6), (7.0 / 96.0, 2),
                      (1.0 / 64.0, 7), (1.0 / 64.0, 1)))]
        return [], B, 1
    if bc1 in (1, -1):
        s = float(bc1)
        B = [(0, _acc((a, 1), (b, 2), (s * b, 2), (c, 3), (s * c, 3), (d, 4), (s * d, 4), (e, 5), (s * e, 5))),
             (1, _acc((a, 2), (b, 3), (b, 1), (c, 4), (s * c, 2), (d, 5), (s * d, 3), (e, 6), (s * e, 4))),
             (2, _acc((a, 3), (b, 4), (b, 2), (c, 5), (c, 1), (d, 6), (s * d, 2), (e, 7), (s * e, 3))),
             (3, _acc((a, 4), (b, 5), (b, 3), (c, 6), (c, 2), (d, 7), (d, 1), (e, 8), (s * e, 2)))]
        return [], B, 1
    raise ValueError("bc must be -1, 0 or 1")


def _left_rows(kind, which, bc):
    if kind == "gaussian" and which == "filter":
        return _left_rows_gaussian(bc)
    if kind == "cd06" and which == "d1":
        return _left_rows_cd06_d1(bc)
    if kind == "cd10" and which == "d1":
        return _left_rows_cd10_d1(bc)
    if kind == "cd10" and which == "d2":
        return _left_rows_cd10_d2(bc)
    if kind == "cf90" and which == "filter":
        return _left_rows_cf90(bc)
    raise ValueError(f"no bounded closure for {kind}/{which}")


# ----------------------------------------------------------------------------
# Matrix assembly
# ----------------------------------------------------------------------------


def min_n(kind, periodic):
    """Minimum grid size accepted by ``init`` (cd06.F90:153-159,176-182,
    cd10.F90:217-226,263-272, cf90.F90:121-129,159-167)."""
    if kind == "cd06":
        return 6 if periodic else 8
    if kind == "cd10":
        return 8
    if kind == "gaussian":
        return 9
    return 10


def init_error(kind, n, periodic):
    """Error code ``init`` returns for size n (0 = ok)."""
    if n == 1 or n >= min_n(kind, periodic):
        return 0
    return {"cd06": 3 if periodic else 2, "cd10": 2, "cf90": 7, "gaussian": 21}[kind]


def assemble(kind, which, n, dx=1.0, periodic=True, bc1=0, bcn=0):
    """Return dense (A, B) with 1/dx^p folded into B.

    For ``cd06`` bounded, ``bc1``/``bcn`` are the *init-time* boundary codes
    (they only select the LHS closure; the RHS always uses the one-sided rows),
    for ``cd10`` and ``cf90`` they are the apply-time codes.
    """
    if which == "d2" and kind == "cd06" and not periodic:
        raise ValueError("bounded cd06 d2 is defined as d1 applied twice")
    Aint, Bint = interior(kind, which)
    A = np.zeros((n, n))
    B = np.zeros((n, n))
    if periodic:
        for i in range(n):
            for o, v in Aint.items():
                A[i, (i + o) % n] += v
            for o, v in Bint.items():
                B[i, (i + o) % n] += v
    else:
        for i in range(n):
            for o, v in Aint.items():
                if 0 <= i + o < n:
                    A[i, i + o] += v
            for o, v in Bint.items():
                if 0 <= i + o < n:
                    B[i, i + o] += v
        for side, bc in ((0, bc1), (1, bcn)):
            Al, Bl, sgn = _left_rows(kind, which, bc)
            # left rows as given; right rows by reflection i -> n-1-i with
            # the routine's sign (odd operators negate the mirrored RHS rows).
            for row, cols in Al:
                r = row if side == 0 else n - 1 - row
                A[r, :] = 0.0
                for col, v in cols.items():
                    A[r, col if side == 0 else n - 1 - col] = v
            for row, cols in Bl:
                r = row if side == 0 else n - 1 - row
                B[r, :] = 0.0
                for col, v in cols.items():
                    B[r, col if side == 0 else n - 1 - col] = v * (1.0 if side == 0 else sgn)
    B *= dx ** (-scale_power(kind, which))
    return A, B


def dense_operator(kind, which, n, dx=1.0, periodic=True, bc1=0, bcn=0):
    """D = A^{-1} B (n x n), FP64."""
    if kind == "cd06" and which == "d2" and not periodic:
        D1 = dense_operator("cd06", "d1", n, dx, False, bc1, bcn)
        return D1 @ D1
    A, B = assemble(kind, which, n, dx, periodic, bc1, bcn)
    return np.linalg.solve(A, B)


def filter3_complex_operator(n, periodic=True, bc1=None, bcn=None):
    """The (n x n) real operator that ``filter3`` applies to each part of a
    complex(rkind) array, ``src/filters/cf90.F90:1230-1251``:

    * the right-hand side uses (bc1, bcn) only when ``bcn_`` is present
      (``:1242-1246``), else (0, 0);
    * the solve always uses ``penta_nn`` (``:1249``), the (0, 0) left-hand
      side, whatever codes formed the right-hand side.

    For a periodic cf90 the reference solves with ``penta_nn`` too, which a
    periodic operator never allocates (only ``LU``, ``:118-120``; the
    ``penta_*`` arrays are allocated for bounded ones, ``:135``); the periodic
    filter is used here instead (parity unpinned for that case: the
    reference's result is undefined).  The bounded mixed operator is pinned in
    tests/test_oracle_complex_filter.py: with (0, 0) it is the real filter,
    and its rows away from the ends are the periodic filter's."""
    if periodic:
        return dense_operator("cf90", "filter", n, 1.0, True)
    b1, bn = (0, 0) if bcn is None else (0 if bc1 is None else bc1, bcn)
    A_nn, _ = assemble("cf90", "filter", n, 1.0, False, 0, 0)
    _, B = assemble("cf90", "filter", n, 1.0, False, b1, bn)
    return np.linalg.solve(A_nn, B)


def apply_axis(D, f, axis):
    """Apply an (n x n) operator along a 0-based axis of a Fortran-ordered
    array ``f`` (shape (n1, n2, n3) in Fortran index order)."""
    out = np.tensordot(D, f, axes=([1], [axis]))
    return np.moveaxis(out, 0, axis)


def apply(kind, which, f, axis, dx=1.0, periodic=True, bc1=0, bcn=0):
    """Reference result of dd1/dd2/dd3 (axis 0/1/2), d2d*, filter* on f."""
    n = f.shape[axis]
    if n == 1:
        return f.copy() if which == "filter" else np.zeros_like(f)
    D = dense_operator(kind, which, n, dx, periodic, bc1, bcn)
    return apply_axis(D, f, axis)


# ----------------------------------------------------------------------------
# Closed forms of the periodic operators (eigenvalues of the circulant D)
# ----------------------------------------------------------------------------


def modified_wavenumber(kind, which, k, dx):
    """Eigenvalue of the periodic operator on exp(i k x): returns complex.

    d1 -> i*k'(k); d2 -> -k''(k); filter -> T(k).  Written from the interior
    stencils (spectral.F90:938-1010 ship the same formulas for cd10/cd06/cf90).
    """
    w = np.asarray(k, dtype=np.float64) * dx
    Aint, Bint = interior(kind, which)
    num = sum(v * np.exp(1j * o * w) for o, v in Bint.items())
    den = sum(v * np.exp(1j * o * w) for o, v in Aint.items())
    return num / den / dx ** scale_power(kind, which)


def spectral_apply(kind, which, f, axis, dx):
    """Periodic operator applied through the FFT along ``axis`` (exact up to
    round-off because Fourier modes are eigenvectors of the circulant D)."""
    n = f.shape[axis]
    k = np.fft.fftfreq(n, d=1.0 / n) * (2 * np.pi / (n * dx))
    lam = modified_wavenumber(kind, which, k, dx)
    if which == "d1" and n % 2 == 0:
        lam = lam.copy()
        lam[n // 2] = lam[n // 2].real  # Nyquist: d1 symbol is odd, exactly 0 in exact arithmetic
    shape = [1, 1, 1]
    shape[axis] = n
    F = np.fft.fft(f, axis=axis)
    out = np.fft.ifft(F * lam.reshape(shape), axis=axis)
    return out.real


# ----------------------------------------------------------------------------
# Periodic staggered operators (src/derivatives/cd06stagg.F90)
# ----------------------------------------------------------------------------

# RHS taps as {input offset: weight}, output point i of a cell array lying
# between input edges i and i+1, output edge j between input cells j-1 and j;
# the LHS is the cyclic tridiagonal (alpha, 1, alpha).
def _stagg_scheme(op):
    if op in ("d1_e2c", "d1_c2e"):
        a, b = 63.0 / 62.0, (17.0 / 62.0) / 3.0              # cd06stagg.F90:307,355
        taps = ({0: -a, 1: a, -1: -b, 2: b} if op == "d1_e2c"      # :313-320
                else {-1: -a, 0: a, -2: -b, 1: b})                   # :361-373
        return 9.0 / 62.0, taps, 1                             # alpha :174
    if op in ("d2_c2c", "d2_e2e"):
        a, b = 12.0 / 11.0, (3.0 / 11.0) / 4.0               # :411,471
        return 2.0 / 11.0, {0: -2.0 * (a + b), 1: a, -1: a, 2: b, -2: b}, 2   # :417-429, :175
    if op in ("interp_e2c", "interp_c2e"):
        a, b = 1.5 / 2.0, 0.1 / 2.0                           # :532-536,578-582
        taps = ({0: a, 1: a, -1: b, 2: b} if op == "interp_e2c"    # :538-545
                else {-1: a, 0: a, -2: b, 1: b})                     # :585-597
        return 3.0 / 10.0, taps, 0                             # alpha :176
    raise ValueError(op)


def stagg_operator(op, n, dx):
    """(n x n) D = A^{-1} B of a periodic staggered operation on the n distinct
    points of a period (an edge array's (n+1)-th plane repeats its first)."""
    al, taps, p = _stagg_scheme(op)
    A = np.zeros((n, n))
    B = np.zeros((n, n))
    for i in range(n):
        A[i, i] = 1.0
        A[i, (i - 1) % n] += al
        A[i, (i + 1) % n] += al
        for o, v in taps.items():
            B[i, (i + o) % n] += v
    return np.linalg.solve(A, B * dx ** (-p))


STAGG_EDGE_IN = {"d1_e2c": True, "d1_c2e": False, "d2_c2c": False, "d2_e2e": True,
                 "interp_e2c": True, "interp_c2e": False}
STAGG_EDGE_OUT = {"d1_e2c": False, "d1_c2e": True, "d2_c2c": False, "d2_e2e": True,
                  "interp_e2c": False, "interp_c2e": True}


def stagg_apply(op, f, dx):
    """Apply along z (last axis) of a Fortran-ordered (n1, n2, n or n+1) array."""
    n = f.shape[2] - (1 if STAGG_EDGE_IN[op] else 0)
    out = apply_axis(stagg_operator(op, n, dx), f[:, :, :n], 2)
    if STAGG_EDGE_OUT[op]:
        out = np.concatenate([out, out[:, :, :1]], axis=2)
    return np.asfortranarray(out)


# ----------------------------------------------------------------------------
# Bounded staggered operators (cd06stagg.F90 init_nonperiodic :197-232)
# ----------------------------------------------------------------------------
#
# Restated statement by statement from the reference's right-hand-side files
# (STAGG_CD06_files/*RHS*_common.F90) and tridiagonal set-up
# (STAGG_CD06_files/ComputeTri_allRoutines.F90), in their order of
# assignment, with the reference's 1-based indices kept in the helpers below.
# The solve is a dense LAPACK solve of the (ddn, dg, dup) tridiagonal matrix,
# so the reference's Thomas sweeps (TridiagSolver_allRoutines.F90) are not
# restated.  The ends are named as the reference does: "bot" is the k = 1
# end, "top" the k = n end; each is (even, sided).

STAGG_OPS_BOUNDED = ("d1_e2c", "d1_c2e", "d1_c2c", "d1_e2e", "d2_c2c", "d2_e2e",
                     "interp_e2c", "interp_c2e")
STAGG_EDGE_IN.update({"d1_c2c": False, "d1_e2e": True})
STAGG_EDGE_OUT.update({"d1_c2c": False, "d1_e2e": True})

# module parameters, cd06stagg.F90:18-56
_SG_W0S, _SG_W1S = 223.0 / 186.0, 61.0 / 62.0


def _stagg_c2c_closure():
    """Lele's one-sided closure of the cell first derivative, cd06stagg.F90:18-53."""
    p, q, r, s = 17.0 / 6.0, 1.5, 1.5, -1.0 / 6.0
    ah = 1.0 / 3.0
    q_p, alpha_p = 0.75, 0.25
    alpha_pp = (((40 * ah - 1) * q + 7 * (4 * ah - 1) * s)
                / (16 * (ah + 2) * q + 8 * (1 - 4 * ah) * s))
    q_pp = (alpha_pp + 2) / 3.0
    r_pp = (4 * alpha_pp - 1) / 12.0
    w1 = (2 * ah + 1) / (2 * (q + s))
    w2 = ((8 * ah + 7) * q - 6 * (2 * ah + 1) * r + (8 * ah + 7) * s) / (9 * (q + s))
    w3 = (4 * (ah + 2) * q + 2 * (1 - 4 * ah) * s) / (9 * (q + s))
    return dict(p=p, q=q, r=r, s=s, q_p=q_p, alpha_p=alpha_p, alpha_pp=alpha_pp,
                q_pp=q_pp, r_pp=r_pp, w1=w1, w2=w2, w3=w3)


class _Rows:
    """Dense rows addressed with the reference's 1-based indices."""

    def __init__(self, m, ncol):
        self.M = np.zeros((m, ncol))

    def set(self, i, terms):
        self.M[i - 1, :] = 0.0
        self.add(i, terms)

    def add(self, i, terms):
        for c, v in terms:
            self.M[i - 1, c - 1] += v


def _stagg_rhs(op, n, dx, bot, top):
    nE = n + 1
    bot_even, bot_sided = bot
    top_even, top_sided = top
    W0S, W1S = _SG_W0S, _SG_W1S
    if op == "d1_e2c":                  # D1RHS_E2C_common.F90:1-40
        a, b = 63.0 / 62.0, (17.0 / 62.0) / 3.0
        alpha1, alpha0 = 37.0 / 183.0, -1.0
        a0 = (alpha0 - 23.0) / 24.0
        b0 = (-9.0 * alpha0 + 7.0) / 8.0
        c0 = (9.0 * alpha0 + 1.0) / 8.0
        d0 = -(alpha0 + 1.0) / 24.0
        a1 = (3.0 / 8.0) * (3.0 - 2.0 * alpha1)
        b1 = (1.0 / 8.0) * (-1.0 + 22.0 * alpha1)
        R = _Rows(n, nE)
        for i in range(2, n):
            R.set(i, [(i + 2, b), (i - 1, -b), (i + 1, a), (i, -a)])
        if bot_sided:
            R.set(1, [(1, W0S * a0), (2, W0S * b0), (3, W0S * c0), (4, W0S * d0)])
            R.set(2, [(1, -W1S * b1 / 3), (2, -W1S * a1), (3, W1S * a1), (4, W1S * b1 / 3)])
        elif bot_even:
            R.set(1, [(3, b), (2, -b), (2, a), (1, -a)])
        else:
            R.set(1, [(3, b), (2, b), (2, a), (1, -a)])
        if top_sided:
            R.set(n, [(nE, -W0S * a0), (nE - 1, -W0S * b0), (nE - 2, -W0S * c0), (nE - 3, -W0S * d0)])
            R.set(n - 1, [(nE - 3, -W1S * b1 / 3), (nE - 2, -W1S * a1), (nE - 1, W1S * a1),
                          (nE, W1S * b1 / 3)])
        elif top_even:
            R.set(n, [(nE - 1, b), (nE - 2, -b), (nE, a), (nE - 1, -a)])
        else:
            R.set(n, [(nE - 1, -b), (nE - 2, -b), (nE, a), (nE - 1, -a)])
        return R.M / dx
    if op == "d1_c2e":                  # D1RHS_C2E_common.F90:1-36
        a, b = 63.0 / 62.0, (17.0 / 62.0) / 3.0
        a0, b0, c0, d0 = -71.0 / 24.0, 47.0 / 8.0, -31.0 / 8.0, 23.0 / 24.0
        a1 = 12.0 / 11.0
        R = _Rows(nE, n)
        for i in range(2, nE):
            R.set(i, [(i, a), (i - 1, -a)])
        for i in range(3, nE - 1):
            R.add(i, [(i + 1, b), (i - 2, -b)])
        if bot_sided:
            R.set(1, [(1, a0), (2, b0), (3, c0), (4, d0)])
            R.set(2, [(2, a1), (1, -a1)])
        elif bot_even:
            R.set(1, [])
            R.add(2, [(3, b), (1, -b)])
        else:
            R.set(1, [(2, 2 * b), (1, 2 * a)])
            R.add(2, [(3, b), (1, b)])
        if top_sided:
            R.set(nE - 1, [(n, a1), (n - 1, -a1)])
            R.set(nE, [(n, -a0), (n - 1, -b0), (n - 2, -c0), (n - 3, -d0)])
        elif top_even:
            R.add(nE - 1, [(n, b), (n - 2, -b)])
            R.set(nE, [])
        else:
            R.add(nE - 1, [(n, -b), (n - 2, -b)])
            R.set(nE, [(n - 1, -2 * b), (n, -2 * a)])
        return R.M / dx
    if op == "d1_c2c":                  # D1RHS_C2C_common.F90:1-71
        a, b = (14.0 / 9.0) / 2.0, (1.0 / 9.0) / 4.0
        k = _stagg_c2c_closure()
        a_np_3, b_np_3 = k["w3"] * k["q_pp"], k["w3"] * k["r_pp"]
        a_np_2 = k["w2"] * k["q_p"]
        a_np_1, b_np_1 = k["w1"] * (-k["p"]), k["w1"] * k["q"]
        c_np_1, d_np_1 = k["w1"] * k["r"], k["w1"] * k["s"]
        R = _Rows(n, n)
        if bot_sided:
            R.set(1, [(1, a_np_1), (2, b_np_1), (3, c_np_1), (4, d_np_1)])
            R.set(2, [(3, a_np_2), (1, -a_np_2)])
            R.set(3, [(4, a_np_3), (2, -a_np_3), (5, b_np_3), (1, -b_np_3)])
        elif bot_even:
            R.set(1, [(3, b), (2, -b), (2, a), (1, -a)])
            R.set(2, [(4, b), (1, -b), (3, a), (1, -a)])
            R.set(3, [(5, b), (1, -b), (4, a), (2, -a)])
        else:
            R.set(1, [(3, b), (2, b), (2, a), (1, a)])
            R.set(2, [(4, b), (1, b), (3, a), (1, -a)])
            R.set(3, [(5, b), (1, -b), (4, a), (2, -a)])
        for i in range(4, n - 2):
            R.set(i, [(i + 2, b), (i - 2, -b), (i + 1, a), (i - 1, -a)])
        if top_sided:
            R.set(n - 2, [(n - 1, a_np_3), (n - 3, -a_np_3), (n, b_np_3), (n - 4, -b_np_3)])
            R.set(n - 1, [(n, a_np_2), (n - 2, -a_np_2)])
            R.set(n, [(n, -a_np_1), (n - 1, -b_np_1), (n - 2, -c_np_1), (n - 3, -d_np_1)])
        elif top_even:
            R.set(n - 2, [(n, b), (n - 4, -b), (n - 1, a), (n - 3, -a)])
            R.set(n - 1, [(n, b), (n - 3, -b), (n, a), (n - 2, -a)])
            R.set(n, [(n - 1, b), (n - 2, -b), (n, a), (n - 1, -a)])
        else:
            R.set(n - 2, [(n, b), (n - 4, -b), (n - 1, a), (n - 3, -a)])
            R.set(n - 1, [(n, -b), (n - 3, -b), (n, a), (n - 2, -a)])
            R.set(n, [(n - 1, -b), (n - 2, -b), (n, -a), (n - 1, -a)])
        return R.M / dx
    if op == "d1_e2e":                  # D1RHS_E2E_common.F90:1-22
        a, b = (14.0 / 9.0) / 2.0, (1.0 / 9.0) / 4.0
        R = _Rows(nE, nE)
        for i in range(2, nE):
            R.set(i, [(i + 1, a), (i - 1, -a)])
        for i in range(3, nE - 1):
            R.add(i, [(i + 2, b), (i - 2, -b)])
        if bot_even:
            R.set(1, [])
            R.add(2, [(4, b), (2, -b)])
        else:
            R.set(1, [(3, 2 * b), (2, 2 * a)])
            R.add(2, [(4, b), (2, b)])
        if top_even:
            R.set(nE, [])
            R.add(nE - 1, [(nE - 1, b), (nE - 3, -b)])
        else:
            R.set(nE, [(nE - 2, -2 * b), (nE - 1, -2 * a)])
            R.add(nE - 1, [(nE - 1, -b), (nE - 3, -b)])
        return R.M / dx
    if op in ("d2_c2c", "d2_e2e"):      # D2RHS_C2C_common.F90:1-26, D2RHS_E2E_common.F90:1-25
        a, b = 12.0 / 11.0, (3.0 / 11.0) / 4.0
        m = n if op == "d2_c2c" else nE
        R = _Rows(m, m)
        for i in range(2, m):
            R.set(i, [(i + 1, a), (i - 1, a)])
        for i in range(3, m - 1):
            R.add(i, [(i + 2, b), (i - 2, b)])
        if op == "d2_c2c":
            if bot_even:
                R.set(1, [(3, b), (2, b), (2, a), (1, a)])
                R.add(2, [(4, b), (1, b)])
            else:
                R.set(1, [(3, b), (2, -b), (2, a), (1, -a)])
                R.add(2, [(4, b), (1, -b)])
            if top_even:
                R.set(m, [(m - 1, b), (m - 2, b), (m, a), (m - 1, a)])
                R.add(m - 1, [(m, b), (m - 3, b)])
            else:
                R.set(m, [(m - 1, -b), (m - 2, b), (m, -a), (m - 1, a)])
                R.add(m - 1, [(m, -b), (m - 3, b)])
        else:
            if bot_even:
                R.set(1, [(3, 2 * b), (2, 2 * a)])
                R.add(2, [(4, b), (2, b)])
            else:
                R.set(1, [])
                R.add(2, [(4, b), (2, -b)])
            if top_even:
                R.set(m, [(m - 2, 2 * b), (m - 1, 2 * a)])
                R.add(m - 1, [(m - 1, b), (m - 3, b)])
            else:
                R.set(m, [])
                R.add(m - 1, [(m - 1, -b), (m - 3, b)])
        R.M -= 2.0 * (a + b) * np.eye(m)
        return R.M / dx ** 2
    if op == "interp_c2e":              # InterpRHS_C2E_common.F90:1-32
        b, a = 0.1 / 2.0, 1.5 / 2.0
        a0, b0, c0 = 15.0 / 8.0, -5.0 / 4.0, 3.0 / 8.0
        alpha1 = 1.0 / 6.0
        a1 = (9.0 + 10.0 * alpha1) / 8.0
        R = _Rows(nE, n)
        for i in range(2, nE):
            R.set(i, [(i, a), (i - 1, a)])
        for i in range(3, nE - 1):
            R.add(i, [(i + 1, b), (i - 2, b)])
        if bot_sided:
            R.set(1, [(1, a0), (2, b0), (3, c0)])
            R.set(2, [(2, a1 / 2), (1, a1 / 2)])
        elif bot_even:
            R.set(1, [(2, 2 * b), (1, 2 * a)])
            R.add(2, [(3, b), (1, b)])
        else:
            R.set(1, [])
            R.add(2, [(3, b), (1, -b)])
        if top_sided:
            R.set(nE, [(n, a0), (n - 1, b0), (n - 2, c0)])
            R.set(nE - 1, [(n, a1 / 2), (n - 1, a1 / 2)])
        elif top_even:
            R.set(nE, [(n - 1, 2 * b), (n, 2 * a)])
            R.add(nE - 1, [(n, b), (n - 2, b)])
        else:
            R.set(nE, [])
            R.add(nE - 1, [(n, -b), (n - 2, b)])
        return R.M
    if op == "interp_e2c":              # InterpRHS_E2C_common.F90:1-26
        b, a = 0.1 / 2.0, 1.5 / 2.0
        alpha0 = 1.0
        a0 = (5.0 - alpha0) / 16.0
        b0 = (9.0 * alpha0 + 15.0) / 16.0
        c0 = (9.0 * alpha0 - 5.0) / 16.0
        d0 = (1.0 - alpha0) / 16.0
        R = _Rows(n, nE)
        if bot_sided:
            R.set(1, [(1, a0), (2, b0), (3, c0), (4, d0)])
        elif bot_even:
            R.set(1, [(3, b), (2, b), (2, a), (1, a)])
        else:
            R.set(1, [(3, b), (2, -b), (2, a), (1, a)])
        for i in range(2, n):
            R.set(i, [(i + 2, b), (i - 1, b), (i + 1, a), (i, a)])
        if top_sided:
            R.set(n, [(nE, a0), (nE - 1, b0), (nE - 2, c0), (nE - 3, d0)])
        elif top_even:
            R.set(n, [(nE - 1, b), (nE - 2, b), (nE, a), (nE - 1, a)])
        else:
            R.set(n, [(nE - 1, -b), (nE - 2, b), (nE, a), (nE - 1, a)])
        return R.M
    raise ValueError(op)


def _stagg_lhs(op, n, bot, top):
    """(ddn, dg, dup) of ComputeTri_allRoutines.F90 as a dense matrix."""
    nE = n + 1
    bot_even, bot_sided = bot
    top_even, top_sided = top
    W0S, W1S = _SG_W0S, _SG_W1S
    alpha = {"d1_e2c": 9.0 / 62.0, "d1_c2e": 9.0 / 62.0, "d1_c2c": 1.0 / 3.0,
             "d1_e2e": 1.0 / 3.0, "d2_c2c": 2.0 / 11.0, "d2_e2e": 2.0 / 11.0,
             "interp_e2c": 3.0 / 10.0, "interp_c2e": 3.0 / 10.0}[op]
    m = n if op in ("d1_e2c", "d1_c2c", "d2_c2c", "interp_e2c") else nE
    ddn = np.full(m + 1, alpha)     # 1-based: element 0 unused
    dg = np.ones(m + 1)
    dup = np.full(m + 1, alpha)
    if op == "d1_e2c":                  # ComputeTri_allRoutines.F90:1-48
        alpha1, alpha0 = 37.0 / 183.0, -1.0
        if top_sided:
            ddn[m], dg[m] = W0S * alpha0, W0S
            ddn[m - 1], dup[m - 1], dg[m - 1] = W1S * alpha1, W1S * alpha1, W1S
        elif top_even:
            dup[m], dg[m] = 0.0, 1.0 - alpha
        else:
            dg[m], dup[m] = 1.0 + alpha, 0.0
        if bot_sided:
            dup[1], dg[1] = W0S * alpha0, W0S
            ddn[2], dup[2], dg[2] = W1S * alpha1, W1S * alpha1, W1S
        elif bot_even:
            ddn[1], dg[1] = 0.0, 1.0 - alpha
        else:
            ddn[1], dg[1] = 0.0, 1.0 + alpha
    elif op == "d1_c2e":                # :50-96
        if top_sided:
            ddn[m], dup[m] = 0.0, 0.0
            ddn[m - 1], dup[m - 1] = 1.0 / 22.0, 1.0 / 22.0
        elif top_even:
            dup[m], ddn[m] = 0.0, 0.0
        else:
            ddn[m], dup[m] = 2 * alpha, 0.0
        if bot_sided:
            dup[1], ddn[1] = 0.0, 0.0
            dup[2], ddn[2] = 1.0 / 22.0, 1.0 / 22.0
        elif bot_even:
            ddn[1], dup[1] = 0.0, 0.0
        else:
            ddn[1], dup[1] = 0.0, 2 * alpha
    elif op == "d1_c2c":                # :98-168
        k = _stagg_c2c_closure()
        alpha_low = 3.0
        if top_sided:
            dup[m], dup[m - 1], dup[m - 2] = 0.0, k["w2"] * k["alpha_p"], k["w3"] * k["alpha_pp"]
            dg[m], dg[m - 1], dg[m - 2] = k["w1"], k["w2"], k["w3"]
            ddn[m], ddn[m - 1], ddn[m - 2] = (k["w1"] * alpha_low, k["w2"] * k["alpha_p"],
                                              k["w3"] * k["alpha_pp"])
        elif top_even:
            dg[m] = 1.0 - alpha
        else:
            dg[m] = 1.0 + alpha
        if bot_sided:
            ddn[1], ddn[2], ddn[3] = 0.0, k["w2"] * k["alpha_p"], k["w3"] * k["alpha_pp"]
            dg[1], dg[2], dg[3] = k["w1"], k["w2"], k["w3"]
            dup[1], dup[2], dup[3] = (k["w1"] * alpha_low, k["w2"] * k["alpha_p"],
                                      k["w3"] * k["alpha_pp"])
        elif bot_even:
            dg[1] = 1.0 - alpha
        else:
            dg[1] = 1.0 + alpha
    elif op == "d1_e2e":                # :170-208
        if top_even:
            dg[m], ddn[m] = 1.0, 0.0
        else:
            dg[m], ddn[m] = 1.0, 2 * alpha
        if bot_even:
            dg[1], dup[1] = 1.0, 0.0
        else:
            dg[1], dup[1] = 1.0, 2 * alpha
    elif op == "interp_c2e":            # :210-255
        alpha1 = 1.0 / 6.0
        if top_sided:
            dup[m], ddn[m] = 0.0, 0.0
            dup[m - 1], ddn[m - 1] = alpha1, alpha1
        elif top_even:
            ddn[m], dg[m] = 2 * alpha, 1.0
        else:
            ddn[m], dg[m] = 0.0, 1.0
        if bot_sided:
            dup[1], ddn[1], dg[1] = 0.0, 0.0, 1.0
            dup[2], ddn[2], dg[2] = alpha1, alpha1, 1.0
        elif bot_even:
            dup[1], dg[1] = 2 * alpha, 1.0
        else:
            dup[1], dg[1] = 0.0, 1.0
    elif op == "interp_e2c":            # :258-299
        alpha0 = 1.0
        if top_sided:
            ddn[m], dup[m] = alpha0, 0.0
        elif top_even:
            dg[m] = 1.0 + alpha
        else:
            dg[m] = 1.0 - alpha
        if bot_sided:
            dup[1], ddn[1] = alpha0, 0.0
        elif bot_even:
            dg[1] = 1.0 + alpha
        else:
            dg[1] = 1.0 - alpha
    elif op == "d2_e2e":                # :302-335
        if top_even:
            dg[m], ddn[m] = 1.0, 2 * alpha
        else:
            dg[m], ddn[m] = 1.0, 0.0
        if bot_even:
            dg[1], dup[1] = 1.0, 2 * alpha
        else:
            dg[1], dup[1] = 1.0, 0.0
    elif op == "d2_c2c":                # :337-366
        dg[m] = 1.0 + alpha if top_even else 1.0 - alpha
        dg[1] = 1.0 + alpha if bot_even else 1.0 - alpha
    else:
        raise ValueError(op)
    A = np.diag(dg[1:])
    for i in range(2, m + 1):
        A[i - 1, i - 2] = ddn[i]
    for i in range(1, m):
        A[i - 1, i] = dup[i]
    return A


def stagg_bounded_matrices(op, n, dx, bot, top):
    """(A, B) of a bounded staggered operation: A (m x m), B (m x n_in)."""
    return _stagg_lhs(op, n, bot, top), _stagg_rhs(op, n, dx, bot, top)


def stagg_bounded_operator(op, n, dx, bot, top):
    """D = A^{-1} B (m x n_in) of a bounded staggered operation."""
    A, B = stagg_bounded_matrices(op, n, dx, bot, top)
    return np.linalg.solve(A, B)


def stagg_bounded_apply(op, f, dx, bot, top):
    """Apply along z (last axis) of a Fortran-ordered (n1, n2, n or n+1) array."""
    n = f.shape[2] - (1 if STAGG_EDGE_IN[op] else 0)
    D = stagg_bounded_operator(op, n, dx, bot, top)
    return np.asfortranarray(np.einsum("ij,abj->abi", D, f))
