// Host-side planning: operator coefficients, band/stencil tables, partitioned
// solve factors and the dense operator.  All arithmetic is FP64.
//
// Coefficients follow the reference's parameter expressions:
//   cd06  src/derivatives/cd06.F90:14-59
//   cd10  src/derivatives/cd10.F90:16-104
//   cf90  src/filters/cf90.F90:16-48
// Bounded rows with bc = +-1 are produced by a mirror extension of the
// interior scheme (even/odd reflection about the boundary node), which is
// what the reference's explicit bc = +-1 rows encode
// (cd10.F90:474-500,528-555,619-642,662-685; cf90.F90:321-344,370-389 and the
// matching RHS rows).  bc = 0 rows are the one-sided closures.
#include "pade_plan.h"

#include <cmath>
#include <cstdlib>
#include <cstring>
#include <utility>

namespace pade {
namespace {

struct Entry {
    int col;
    double v;
};
struct Row {
    int row;
    std::vector<Entry> e;
};

// Interior scheme of each operator: A band (offsets -2..2) and B stencil
// (offsets -4..4), before any 1/dx scaling.
struct Interior {
    double a[5] = {0};
    double b[9] = {0};
    int H = 0;
    int parity = 1;  // +1: output has the parity of the input (d2, filter); -1: d1
    int dxpow = 0;
};

Interior interior(int kind, int oper)
{
    Interior s;
    if (kind == CD06 && oper == D1) {
        const double al = 1.0 / 3.0, a = (14.0 / 9.0) / 2.0, b = (1.0 / 9.0) / 4.0;
        s.a[1] = al; s.a[2] = 1.0; s.a[3] = al;
        s.b[5] = a; s.b[3] = -a; s.b[6] = b; s.b[2] = -b;
        s.H = 2; s.parity = -1; s.dxpow = 1;
    } else if (kind == CD06 && oper == D2) {
        // Lele's sixth-order tridiagonal second derivative (the reference lists
        // the coefficients, cd06.F90:18-21, but never solves with them).
        const double al = 2.0 / 11.0, a = 12.0 / 11.0, b = (3.0 / 11.0) / 4.0;
        s.a[1] = al; s.a[2] = 1.0; s.a[3] = al;
        s.b[4] = -2.0 * (a + b); s.b[3] = s.b[5] = a; s.b[2] = s.b[6] = b;
        s.H = 2; s.parity = 1; s.dxpow = 2;
    } else if (kind == CD10 && oper == D1) {
        const double al = 0.5, be = 1.0 / 20.0;
        const double a = (17.0 / 12.0) / 2.0, b = (101.0 / 150.0) / 4.0, c = (1.0 / 100.0) / 6.0;
        s.a[0] = be; s.a[1] = al; s.a[2] = 1.0; s.a[3] = al; s.a[4] = be;
        s.b[5] = a; s.b[3] = -a; s.b[6] = b; s.b[2] = -b; s.b[7] = c; s.b[1] = -c;
        s.H = 3; s.parity = -1; s.dxpow = 1;
    } else if (kind == CD10 && oper == D2) {
        const double al = 334.0 / 899.0, be = 43.0 / 1798.0;
        const double a = 1065.0 / 1798.0, b = (1038.0 / 899.0) / 4.0, c = (79.0 / 1798.0) / 9.0;
        s.a[0] = be; s.a[1] = al; s.a[2] = 1.0; s.a[3] = al; s.a[4] = be;
        s.b[4] = -2.0 * (a + b + c);
        s.b[3] = s.b[5] = a; s.b[2] = s.b[6] = b; s.b[1] = s.b[7] = c;
        s.H = 3; s.parity = 1; s.dxpow = 2;
    } else if (kind == CF90 && oper == FILTER) {
        const double al = 6.6624e-1, be = 1.6688e-1;
        s.a[0] = be; s.a[1] = al; s.a[2] = 1.0; s.a[3] = al; s.a[4] = be;
        s.b[4] = 9.9965e-1;
        s.b[3] = s.b[5] = 6.6652e-1;
        s.b[2] = s.b[6] = 1.6674e-1;
        s.b[1] = s.b[7] = 4.0e-5;
        s.b[0] = s.b[8] = -5.0e-6;
        s.H = 4; s.parity = 1; s.dxpow = 0;
    } else if (kind == GAUSSIAN && oper == FILTER) {
        // explicit Gaussian filter of width 4 dx (src/filters/gaussian.F90:11-17
        // coefficients, :137-185 periodic stencil, bounded rows below): A = I
        s.a[2] = 1.0;
        s.b[4] = 3565.0 / 10368.0;
        s.b[3] = s.b[5] = 3091.0 / 12960.0;
        s.b[2] = s.b[6] = 1997.0 / 25920.0;
        s.b[1] = s.b[7] = 149.0 / 12960.0;
        s.b[0] = s.b[8] = 107.0 / 103680.0;
        s.H = 4; s.parity = 1; s.dxpow = 0;
    } else if (kind == CD06STAGG) {
        // periodic staggered sixth-order scheme, cd06stagg.F90:170-195 (LHS
        // alphas) and the ComputeZ*_periodic RHS routines: output point i of a
        // cell array sits between input edges i and i + 1, output edge j between
        // input cells j - 1 and j.  Index offsets below are input minus output.
        double al = 0.0;
        if (oper == D1_E2C || oper == D1_C2E) {
            al = 9.0 / 62.0;                                    // :174
            const double a = 63.0 / 62.0, b = (17.0 / 62.0) / 3.0;   // :307, :355
            const int o = oper == D1_E2C ? 0 : -1;              // E2C :313-320, C2E :361-373
            s.b[4 + o] = -a; s.b[5 + o] = a; s.b[3 + o] = -b; s.b[6 + o] = b;
            s.parity = -1; s.dxpow = 1;
        } else if (oper == D2_C2C || oper == D2_E2E) {
            al = 2.0 / 11.0;                                    // :175
            const double a = 12.0 / 11.0, b = (3.0 / 11.0) / 4.0;  // :411, :471
            s.b[4] = -2.0 * (a + b); s.b[3] = s.b[5] = a; s.b[2] = s.b[6] = b;
            s.parity = 1; s.dxpow = 2;
        } else if (oper == INTERP_E2C || oper == INTERP_C2E) {
            al = 3.0 / 10.0;                                    // :176
            const double a = 1.5 / 2.0, b = 0.1 / 2.0;          // :532-536, :578-582
            const int o = oper == INTERP_E2C ? 0 : -1;          // E2C :538-545, C2E :585-597
            s.b[4 + o] = s.b[5 + o] = a; s.b[3 + o] = s.b[6 + o] = b;
            s.parity = 1; s.dxpow = 0;
        } else if (oper == D1_C2C || oper == D1_E2E) {
            // the collocated sixth-order first derivative on cells or edges
            // (STAGG_CD06_files/D1RHS_C2C_common.F90:1, D1RHS_E2E_common.F90:1;
            // alpha ComputeTri_allRoutines.F90:103,174)
            al = 1.0 / 3.0;
            const double a = (14.0 / 9.0) / 2.0, b = (1.0 / 9.0) / 4.0;
            s.b[5] = a; s.b[3] = -a; s.b[6] = b; s.b[2] = -b;
            s.parity = -1; s.dxpow = 1;
        } else {
            s.H = -1;
            return s;
        }
        s.a[1] = al; s.a[2] = 1.0; s.a[3] = al;
        s.H = 2;
    } else {
        s.H = -1;
    }
    return s;
}

// One-sided (bc = 0) closure rows at the left end, unscaled by dx.
void closure_rows(int kind, int oper, std::vector<Row> &A, std::vector<Row> &B)
{
    A.clear();
    B.clear();
    if (kind == CD06 && oper == D1) {
        // Lele's boundary scheme with the interior alpha = 1/3
        const double ah = 1.0 / 3.0;
        const double al = 3.0, p = 17.0 / 6.0, q = 1.5, r = 1.5, s = -1.0 / 6.0;
        const double qp = 0.75, alp = 0.25;
        const double alpp = ((40 * ah - 1) * q + 7 * (4 * ah - 1) * s) /
                            (16 * (ah + 2) * q + 8 * (1 - 4 * ah) * s);
        const double qpp = (alpp + 2) / 3.0, rpp = (4 * alpp - 1) / 12.0;
        const double w1 = (2 * ah + 1) / (2 * (q + s));
        const double w2 = ((8 * ah + 7) * q - 6 * (2 * ah + 1) * r + (8 * ah + 7) * s) / (9 * (q + s));
        const double w3 = (4 * (ah + 2) * q + 2 * (1 - 4 * ah) * s) / (9 * (q + s));
        A.push_back({0, {{0, w1}, {1, w1 * al}}});
        A.push_back({1, {{0, w2 * alp}, {1, w2}, {2, w2 * alp}}});
        A.push_back({2, {{1, w3 * alpp}, {2, w3}, {3, w3 * alpp}}});
        B.push_back({0, {{0, -w1 * p}, {1, w1 * q}, {2, w1 * r}, {3, w1 * s}}});
        B.push_back({1, {{2, w2 * qp}, {0, -w2 * qp}}});
        B.push_back({2, {{3, w3 * qpp}, {1, -w3 * qpp}, {4, w3 * rpp}, {0, -w3 * rpp}}});
    } else if (kind == CD10 && oper == D1) {
        const double qh = (17.0 / 12.0) / 2.0, rh = (101.0 / 150.0) / 4.0, sh = (1.0 / 100.0) / 6.0;
        const double al = 3.0, p = -17.0 / 6.0, q = 1.5, r = 1.5, s = -1.0 / 6.0;
        const double qp = 0.75, alp = 0.25;
        const double al3 = (8 * rh - 175 * sh) / (18 * rh - 550 * sh);
        const double be3 = (-3 + 8 * al3) / 20.0;
        const double q3 = (12 - 7 * al3) / 12.0;
        const double r3 = (568 * al3 - 183) / 600.0;
        const double s3 = (9 * al3 - 4) / 300.0;
        const double num = s * (rh + 2 * sh) - q * (qh + rh + sh);
        const double den = (q + s) * (qh + rh - sh * (q3 / s3 - 1));
        const double al2 = (17 * num / (72 * den) - 8.0 / 9.0) / (19 * num / (24 * den) - 1.0 / 3.0);
        const double be2 = (-1 + 3 * al2) / 12.0;
        const double q2 = (8 - 3 * al2) / 9.0;
        const double r2 = (-17 + 57 * al2) / 72.0;
        const double w1 = (qh + 2 * rh + 3 * sh) / (q + s);
        const double w2 = (rh + sh * (1 + q3 / s3) - r * (qh + 2 * rh + 3 * sh) / (q + s)) / qp;
        const double w3 = (qh + rh + sh * (1 - q3 / s3)) / r2;
        const double w4 = sh / s3;
        A.push_back({0, {{0, w1}, {1, w1 * al}}});
        A.push_back({1, {{0, w2 * alp}, {1, w2}, {2, w2 * alp}}});
        A.push_back({2, {{0, w3 * be2}, {1, w3 * al2}, {2, w3}, {3, w3 * al2}, {4, w3 * be2}}});
        A.push_back({3, {{1, w4 * be3}, {2, w4 * al3}, {3, w4}, {4, w4 * al3}, {5, w4 * be3}}});
        B.push_back({0, {{0, w1 * p}, {1, w1 * q}, {2, w1 * r}, {3, w1 * s}}});
        B.push_back({1, {{2, w2 * qp}, {0, -w2 * qp}}});
        B.push_back({2, {{3, w3 * q2}, {1, -w3 * q2}, {4, w3 * r2}, {0, -w3 * r2}}});
        B.push_back({3, {{4, w4 * q3}, {2, -w4 * q3}, {5, w4 * r3}, {1, -w4 * r3},
                         {6, w4 * s3}, {0, -w4 * s3}}});
    } else if (kind == CD10 && oper == D2) {
        const double b1al = 11.0;
        const double b1a = (11 * b1al + 35) / 12, b1b = -(5 * b1al + 26) / 3, b1c = (b1al + 19) / 2,
                     b1d = (b1al - 14) / 3, b1e = (11 - b1al) / 12;
        const double b2al = 0.1, b2a = 4 * (1 - b2al) / 3;
        const double b3al = 344.0 / 1179.0, b3be = (38.0 * b3al - 9.0) / 214.0;
        const double b3a = (696 - 1191 * b3al) / 428, b3b = ((2454 * b3al - 294) / 535) / 4.0;
        A.push_back({0, {{0, 1.0}, {1, b1al}}});
        A.push_back({1, {{0, b2al}, {1, 1.0}, {2, b2al}}});
        A.push_back({2, {{0, b3be}, {1, b3al}, {2, 1.0}, {3, b3al}, {4, b3be}}});
        B.push_back({0, {{0, b1a}, {1, b1b}, {2, b1c}, {3, b1d}, {4, b1e}}});
        B.push_back({1, {{2, b2a}, {1, -2 * b2a}, {0, b2a}}});
        B.push_back({2, {{3, b3a}, {2, -2 * (b3a + b3b)}, {1, b3a}, {4, b3b}, {0, b3b}}});
    } else if (kind == CF90 && oper == FILTER) {
        A.push_back({0, {{0, 1.0}}});
        A.push_back({1, {{0, 4.997e-1}, {1, 1.0}, {2, 4.997e-1}}});
        A.push_back({2, {{0, 1.6688e-1}, {1, 6.6624e-1}, {2, 1.0}, {3, 6.6624e-1}, {4, 1.6688e-1}}});
        A.push_back({3, {{1, 1.6688e-1}, {2, 6.6624e-1}, {3, 1.0}, {4, 6.6624e-1}, {5, 1.6688e-1}}});
        B.push_back({0, {{0, 1.0}}});
        B.push_back({1, {{1, 9.997e-1}, {2, 4.9985e-1}, {0, 4.9985e-1}}});
        B.push_back({2, {{2, 9.9952e-1}, {3, 6.6656e-1}, {1, 6.6656e-1}, {4, 1.668e-1}, {0, 1.668e-1}}});
        B.push_back({3, {{3, 9.9968e-1}, {4, 6.6652e-1}, {2, 6.6652e-1}, {5, 1.6672e-1},
                         {1, 1.6672e-1}, {6, 4.0e-5}, {0, 4.0e-5}}});
    } else if (kind == GAUSSIAN && oper == FILTER) {
        // explicit closure rows of the bounded Gaussian filter (A = I),
        // gaussian.F90:19-36 weights, rows :192-208
        B.push_back({0, {{0, 5.0 / 6.0}, {1, 1.0 / 6.0}}});
        B.push_back({1, {{1, 2.0 / 3.0}, {2, 1.0 / 6.0}, {0, 1.0 / 6.0}}});
        B.push_back({2, {{2, 31.0 / 64.0}, {3, 7.0 / 32.0}, {1, 7.0 / 32.0}, {4, 5.0 / 128.0}, {0, 5.0 / 128.0}}});
        B.push_back({3, {{3, 17.0 / 48.0}, {4, 15.0 / 64.0}, {2, 15.0 / 64.0}, {5, 7.0 / 96.0}, {1, 7.0 / 96.0},
                         {6, 1.0 / 64.0}, {0, 1.0 / 64.0}}});
    }
}

// Mirror rows (bc = +-1): the interior scheme on the reflected grid.  For A,
// the output parity is bc*parity; for B the input parity is bc.
void mirror_rows(const Interior &s, int bc, std::vector<Row> &A, std::vector<Row> &B)
{
    A.clear();
    B.clear();
    const double sa = double(bc * s.parity), sb = double(bc);
    for (int i = 0; i < kBandW; i++) {
        double acc[16] = {0};
        for (int o = -2; o <= 2; o++) {
            int c = i + o;
            double v = s.a[o + 2];
            if (c < 0) { c = -c; v *= sa; }
            acc[c] += v;
        }
        Row r{i, {}};
        for (int c = 0; c < 16; c++) if (acc[c] != 0.0) r.e.push_back({c, acc[c]});
        // an odd reflection pins the boundary value itself: the row reduces to
        // the identity (reference: penta rows with d = 1 and zero neighbours)
        A.push_back(r);
    }
    for (int i = 0; i < s.H; i++) {
        double acc[16] = {0};
        for (int o = -s.H; o <= s.H; o++) {
            int c = i + o;
            double v = s.b[o + 4];
            if (c < 0) { c = -c; v *= sb; }
            acc[c] += v;
        }
        Row r{i, {}};
        for (int c = 0; c < 16; c++) if (acc[c] != 0.0) r.e.push_back({c, acc[c]});
        B.push_back(r);
    }
}

// ------------------------------------------------------------------------
// Bounded staggered operators (cd06stagg.F90:197-232, init_nonperiodic).
//
// An even or odd end is the interior scheme applied to the mirror image of
// the data about the wall: a cell array's wall lies half a cell beyond its
// end point, an edge array's wall on its end point.  Input and output are
// reflected with their own parities (a first derivative flips it).  This
// reproduces every even / odd row of the reference's right-hand sides
// (STAGG_CD06_files/*RHS*_common.F90) and tridiagonal systems
// (ComputeTri_allRoutines.F90).  A one-sided end replaces the first rows by
// the closure tables below (bottom end; the top end uses them reflected,
// negated for a first derivative).  Order of assignment follows the
// reference: right-hand side bottom rows, then top rows (the top wins where
// they meet on very short grids); left-hand side top, then bottom.

struct TriRow {
    double dn, dg, up;
};
struct SidedClosure {
    int nrows = 0;                       // 0: the operation has no one-sided form
    std::vector<Row> B;                  // bottom rows, columns from the wall
    TriRow A[3] = {};
};

SidedClosure stagg_sided(int oper)
{
    SidedClosure c;
    const double w0s = 223.0 / 186.0, w1s = 61.0 / 62.0;            // cd06stagg.F90:56
    if (oper == D1_E2C) {
        // D1RHS_E2C_common.F90:2-5,12-16; ComputeTri_allRoutines.F90:6,26-29
        const double al1 = 37.0 / 183.0, al0 = -1.0;
        const double a0 = (al0 - 23.0) / 24.0, b0 = (7.0 - 9.0 * al0) / 8.0;
        const double c0 = (9.0 * al0 + 1.0) / 8.0, d0 = -(al0 + 1.0) / 24.0;
        const double a1 = 3.0 * (3.0 - 2.0 * al1) / 8.0, b1 = (22.0 * al1 - 1.0) / 8.0;
        c.nrows = 2;
        c.B = {{0, {{0, w0s * a0}, {1, w0s * b0}, {2, w0s * c0}, {3, w0s * d0}}},
               {1, {{0, -w1s * b1 / 3.0}, {1, -w1s * a1}, {2, w1s * a1}, {3, w1s * b1 / 3.0}}}};
        c.A[0] = {0.0, w0s, w0s * al0};
        c.A[1] = {w1s * al1, w1s, w1s * al1};
    } else if (oper == D1_C2E) {
        // D1RHS_C2E_common.F90:2-4,12-14; ComputeTri_allRoutines.F90:71-74
        const double a1 = 12.0 / 11.0;
        c.nrows = 2;
        c.B = {{0, {{0, -71.0 / 24.0}, {1, 47.0 / 8.0}, {2, -31.0 / 8.0}, {3, 23.0 / 24.0}}},
               {1, {{0, -a1}, {1, a1}}}};
        c.A[0] = {0.0, 1.0, 0.0};
        c.A[1] = {1.0 / 22.0, 1.0, 1.0 / 22.0};
    } else if (oper == D1_C2C) {
        // Lele's closure, cd06stagg.F90:18-53; D1RHS_C2C_common.F90:9-27;
        // ComputeTri_allRoutines.F90:135-146
        const double p = 17.0 / 6.0, q = 1.5, r = 1.5, s = -1.0 / 6.0, ah = 1.0 / 3.0;
        const double q_p = 0.75, al_p = 0.25;
        const double al_pp = ((40 * ah - 1) * q + 7 * (4 * ah - 1) * s) / (16 * (ah + 2) * q + 8 * (1 - 4 * ah) * s);
        const double q_pp = (al_pp + 2) / 3.0, r_pp = (4 * al_pp - 1) / 12.0;
        const double w1 = (2 * ah + 1) / (2 * (q + s));
        const double w2 = ((8 * ah + 7) * q - 6 * (2 * ah + 1) * r + (8 * ah + 7) * s) / (9 * (q + s));
        const double w3 = (4 * (ah + 2) * q + 2 * (1 - 4 * ah) * s) / (9 * (q + s));
        c.nrows = 3;
        c.B = {{0, {{0, -w1 * p}, {1, w1 * q}, {2, w1 * r}, {3, w1 * s}}},
               {1, {{0, -w2 * q_p}, {2, w2 * q_p}}},
               {2, {{0, -w3 * r_pp}, {1, -w3 * q_pp}, {3, w3 * q_pp}, {4, w3 * r_pp}}}};
        c.A[0] = {0.0, w1, 3.0 * w1};
        c.A[1] = {w2 * al_p, w2, w2 * al_p};
        c.A[2] = {w3 * al_pp, w3, w3 * al_pp};
    } else if (oper == INTERP_C2E) {
        // InterpRHS_C2E_common.F90:2-3,8-10; ComputeTri_allRoutines.F90:215,236-237
        const double a1 = (9.0 + 10.0 / 6.0) / 8.0;
        c.nrows = 2;
        c.B = {{0, {{0, 15.0 / 8.0}, {1, -5.0 / 4.0}, {2, 3.0 / 8.0}}},
               {1, {{0, a1 / 2.0}, {1, a1 / 2.0}}}};
        c.A[0] = {0.0, 1.0, 0.0};
        c.A[1] = {1.0 / 6.0, 1.0, 1.0 / 6.0};
    } else if (oper == INTERP_E2C) {
        // InterpRHS_E2C_common.F90:2-5,7; ComputeTri_allRoutines.F90:263,279
        const double al0 = 1.0;
        c.nrows = 1;
        c.B = {{0, {{0, (5.0 - al0) / 16.0}, {1, (9.0 * al0 + 15.0) / 16.0},
                    {2, (9.0 * al0 - 5.0) / 16.0}, {3, (1.0 - al0) / 16.0}}}};
        c.A[0] = {0.0, 1.0, al0};
    }
    return c;
}

// Dense B (m x nin, before 1/dx^p) and tridiagonal A (m rows) of a bounded
// staggered operation.  bc: +-1 even / odd end, +-2 one-sided closure
// (parity by the sign).
int stagg_dense(int oper, int n, int bc1, int bcn, const Interior &s, std::vector<double> &B,
                std::vector<TriRow> &A, int &m_out, int &nin_out)
{
    auto ok = [](int b) { return b == 1 || b == -1 || b == 2 || b == -2; };
    if (!ok(bc1) || !ok(bcn)) return 324;
    const bool ein = stagg_edge_in(oper), eout = stagg_edge_out(oper);
    const int nin = n + (ein ? 1 : 0), m = n + (eout ? 1 : 0);
    m_out = m;
    nin_out = nin;
    const int sb = bc1 > 0 ? 1 : -1, st = bcn > 0 ? 1 : -1;
    const int ob = sb * s.parity, ot = st * s.parity;     // output parities
    // mirror image of index j of an array of len points -> (index, sign)
    auto mirror = [](int j, bool edge, int len, int lo, int hi, int &sg) {
        sg = 1;
        if (j < 0) { sg = lo; return edge ? -j : -1 - j; }
        if (j >= len) { sg = hi; return edge ? 2 * (len - 1) - j : 2 * len - 1 - j; }
        return j;
    };
    B.assign((size_t)m * nin, 0.0);
    A.assign(m, TriRow{0.0, 0.0, 0.0});
    auto mirror_row_B = [&](int i) {
        double *r = &B[(size_t)i * nin];
        for (int c = 0; c < nin; c++) r[c] = 0.0;
        for (int k = 0; k < 9; k++) {
            if (s.b[k] == 0.0) continue;
            int sg;
            const int j = mirror(i - 4 + k, ein, nin, sb, st, sg);
            r[j] += sg * s.b[k];
        }
    };
    for (int i = 0; i < m; i++) {
        mirror_row_B(i);
        double row[3] = {s.a[1], s.a[2], s.a[3]};
        double full[3] = {0.0, 0.0, 0.0};     // coefficients of y[i-1], y[i], y[i+1]
        for (int o = -1; o <= 1; o++) {
            int sg;
            const int j = mirror(i + o, eout, m, ob, ot, sg);
            full[j - i + 1] += sg * row[o + 1];
        }
        A[i] = {full[0], full[1], full[2]};
    }
    const SidedClosure sc = stagg_sided(oper);
    const double bsign = double(s.parity);
    auto put_B = [&](const Row &r, bool top) {
        const int i = top ? m - 1 - r.row : r.row;
        double *d = &B[(size_t)i * nin];
        for (int c = 0; c < nin; c++) d[c] = 0.0;
        for (const Entry &e : r.e) d[top ? nin - 1 - e.col : e.col] += (top ? bsign : 1.0) * e.v;
    };
    const bool bot_sided = sc.nrows && std::abs(bc1) == 2, top_sided = sc.nrows && std::abs(bcn) == 2;
    if (bot_sided)
        for (const Row &r : sc.B) put_B(r, false);
    if (top_sided) {
        for (const Row &r : sc.B) put_B(r, true);
    } else if (bot_sided) {
        for (int q = 0; q < sc.nrows; q++) mirror_row_B(m - 1 - q);
    }
    if (top_sided)
        for (int q = 0; q < sc.nrows; q++) A[m - 1 - q] = {sc.A[q].up, sc.A[q].dg, sc.A[q].dn};
    if (bot_sided)
        for (int q = 0; q < sc.nrows; q++) A[q] = sc.A[q];
    return 0;
}

// OpMatrices of the dense (B, A) restricted to rows [0, m): explicit rows for
// every row that is not the plain stencil; right rows relative to input
// point nin - 1 (generic form) or to point m (system form, `rect`).
int stagg_tables(const std::vector<double> &B, const std::vector<TriRow> &A, int m, int nin, double dx,
                 const Interior &s, bool rect, OpMatrices &out)
{
    out = OpMatrices();
    out.n = m;
    out.n_in = nin;
    out.periodic = false;
    out.rect = rect;
    out.H = s.H;
    out.band.assign((size_t)m * 5, 0.0);
    for (int i = 0; i < m; i++) {
        if (i > 0) out.band[(size_t)i * 5 + 1] = A[i].dn;
        out.band[(size_t)i * 5 + 2] = A[i].dg;
        if (i + 1 < m) out.band[(size_t)i * 5 + 3] = A[i].up;
    }
    const double scale = std::pow(dx, -double(s.dxpow));
    for (int k = 0; k < 9; k++) out.stencil[k] = s.b[k] * scale;
    // rows that are not the plain stencil go to the explicit tables
    auto plain = [&](int i) {
        const double *r = &B[(size_t)i * nin];
        for (int c = 0; c < nin; c++) {
            const int k = c - i + 4;
            const double want = (k >= 0 && k < 9) ? s.b[k] : 0.0;
            if (r[c] != want) return false;
        }
        for (int k = 0; k < 9; k++) {
            const int c = i - 4 + k;
            if (s.b[k] != 0.0 && (c < 0 || c >= nin)) return false;
        }
        return true;
    };
    const int h = (m + 1) / 2;
    int nbl = 0, nbr = 0;
    for (int i = 0; i < h; i++)
        if (!plain(i)) nbl = i + 1;
    for (int i = m - 1; i >= h; i--)
        if (!plain(i)) nbr = m - i;
    if (nbl > kBRows || nbr > kBRows) return 21;
    for (int q = 0; q < nbl; q++)
        for (int c = 0; c < nin; c++) {
            const double v = B[(size_t)q * nin + c];
            if (v == 0.0) continue;
            if (c >= kBCols) return 21;
            out.bl[q][c] = v * scale;
        }
    if (rect && nbr > kBRows - 1) return 21;     // the last table row holds the tail row
    for (int q = 0; q < nbr; q++)
        for (int c = 0; c < nin; c++) {
            const double v = B[(size_t)(m - 1 - q) * nin + c];
            if (v == 0.0) continue;
            const int cc = rect ? m - c : nin - 1 - c;
            if (cc < 0 || cc >= kBCols) return 21;
            out.br[q][cc] = v * scale;
        }
    out.nbl = nbl;
    out.nbr = nbr;
    return 0;
}

int build_stagg_bounded(int oper, int n, double dx, int bc1, int bcn, const Interior &s, OpMatrices &out)
{
    std::vector<double> B;
    std::vector<TriRow> A;
    int m = 0, nin = 0;
    if (int rc = stagg_dense(oper, n, bc1, bcn, s, B, A, m, nin)) return rc;
    return stagg_tables(B, A, m, nin, dx, s, false, out);
}

}  // namespace

int build_stagg_system(int oper, int n, double dx, int bc1, int bcn, OpMatrices &out)
{
    const Interior s = interior(CD06STAGG, oper);
    if (s.H < 0) return 21;
    std::vector<double> B;
    std::vector<TriRow> A;
    int m = 0, nin = 0;
    if (int rc = stagg_dense(oper, n, bc1, bcn, s, B, A, m, nin)) return rc;
    if (!stagg_edge_out(oper)) return stagg_tables(B, A, m, nin, dx, s, true, out);
    // n + 1 outputs: eliminate the last row L from row L - 1 (its only neighbour)
    const int L = m - 1;
    const double tdn = A[L].dn, tdg = A[L].dg;
    const double f = A[L - 1].up / tdg;
    A[L - 1].dg -= f * A[L].dn;
    A[L - 1].up = 0.0;
    std::vector<double> Bt(B.begin() + (size_t)L * nin, B.begin() + (size_t)m * nin);
    for (int c = 0; c < nin; c++) B[(size_t)(L - 1) * nin + c] -= f * Bt[c];
    B.resize((size_t)L * nin);
    A.resize(L);
    if (int rc = stagg_tables(B, A, L, nin, dx, s, true, out)) return rc;
    const double scale = std::pow(dx, -double(s.dxpow));
    for (int c = 0; c < nin; c++) {
        if (Bt[c] == 0.0) continue;
        const int cc = L - c;
        if (cc < 0 || cc >= kBCols) return 21;
        out.br[kBRows - 1][cc] = Bt[c] * scale;
    }
    out.tail = true;
    out.tail_dn = tdn;
    out.tail_idg = 1.0 / tdg;
    return 0;
}

int init_error(int kind, int n, bool periodic)
{
    if (n == 1) return 0;
    if (kind == CD06) return periodic ? (n >= 6 ? 0 : 3) : (n >= 8 ? 0 : 2);
    if (kind == CD10) return n >= 8 ? 0 : 2;
    if (kind == CF90) return n >= 10 ? 0 : 7;
    // the Gaussian stencil spans 9 points (gaussian.F90:137-185); the
    // reference's init checks nothing and would index out of its array below
    if (kind == GAUSSIAN) return n >= 9 ? 0 : 21;
    // cd06stagg.F90:182-184: GracefulExit 21 for nx <= 4
    if (kind == CD06STAGG) return n > 4 ? 0 : 21;
    return 1;
}

int build_operator(int kind, int oper, int n, double dx, bool periodic, int bc1, int bcn,
                   OpMatrices &out)
{
    Interior s = interior(kind, oper);
    if (s.H < 0) return 21;
    if (kind == CD06 && oper == D2 && !periodic) return 21;  // composed from d1 by the caller
    if (kind == CD06STAGG && !periodic) return build_stagg_bounded(oper, n, dx, bc1, bcn, s, out);
    // the reference builds the collocated staggered-array derivatives for the
    // bounded form only (cd06stagg.F90:224-227)
    if (kind == CD06STAGG && (oper == D1_C2C || oper == D1_E2E)) return 21;
    if (kind != CD06 && !periodic) {
        auto ok = [](int b) { return b == 0 || b == 1 || b == -1; };
        if (!ok(bc1) || !ok(bcn)) return 324;
    }
    out = OpMatrices();
    out.n = n;
    out.periodic = periodic;
    out.H = s.H;
    out.band.assign((size_t)n * 5, 0.0);
    const double scale = std::pow(dx, -double(s.dxpow));
    for (int k = 0; k < 9; k++) out.stencil[k] = s.b[k] * scale;
    for (int i = 0; i < n; i++)
        for (int o = -2; o <= 2; o++) {
            int j = i + o;
            if (periodic || (j >= 0 && j < n)) out.band[(size_t)i * 5 + o + 2] = s.a[o + 2];
        }
    if (periodic) return 0;

    for (int side = 0; side < 2; side++) {
        const int bc = side == 0 ? bc1 : bcn;
        std::vector<Row> Ar, Br;
        if (kind == CD06) {
            // the reference's cd06 always evaluates the one-sided RHS rows
            // (cd06.F90:563-585) and uses the closure LHS rows only for init
            // bc == 0 (cd06.F90:279-312); other codes keep the interior rows.
            std::vector<Row> Ac;
            closure_rows(kind, oper, Ac, Br);
            if (bc == 0) Ar = Ac;
        } else if (bc == 0) {
            closure_rows(kind, oper, Ar, Br);
        } else {
            mirror_rows(s, bc, Ar, Br);
        }
        const double bsign = side == 0 ? 1.0 : double(s.parity);
        for (const Row &r : Ar) {
            const int i = side == 0 ? r.row : n - 1 - r.row;
            for (int o = -2; o <= 2; o++) out.band[(size_t)i * 5 + o + 2] = 0.0;
            for (const Entry &e : r.e) {
                const int c = side == 0 ? e.col : n - 1 - e.col;
                out.band[(size_t)i * 5 + (c - i) + 2] = e.v;
            }
        }
        int nrows = 0;
        for (const Row &r : Br) {
            nrows = std::max(nrows, r.row + 1);
            double *dst = side == 0 ? out.bl[r.row] : out.br[r.row];
            for (int c = 0; c < kBCols; c++) dst[c] = 0.0;
            for (const Entry &e : r.e) dst[e.col] += e.v * scale * bsign;
        }
        if (side == 0) out.nbl = nrows; else out.nbr = nrows;
    }
    return 0;
}

void transpose_operator(const OpMatrices &f, OpMatrices &t)
{
    const int n = f.n;
    t = OpMatrices();
    t.n = n;
    t.periodic = f.periodic;
    t.H = f.H;
    t.band.assign((size_t)n * 5, 0.0);
    for (int i = 0; i < n; i++)
        for (int o = -2; o <= 2; o++) {
            int j = i + o;
            if (f.periodic) j = ((j % n) + n) % n;
            else if (j < 0 || j >= n) continue;
            // A^T(i, i+o) = A(i+o, i): row j, offset -o
            t.band[(size_t)i * 5 + o + 2] = f.band[(size_t)j * 5 + (-o) + 2];
        }
    for (int k = 0; k < 9; k++) t.stencil[k] = f.stencil[8 - k];
    if (f.periodic) return;
    // B(i, j) of the bounded forward operator (any row kind)
    auto Bentry = [&](int i, int j) -> double {
        if (i < 0 || i >= n || j < 0 || j >= n) return 0.0;
        if (i < f.nbl) return j < kBCols ? f.bl[i][j] : 0.0;
        if (i >= n - f.nbr) {
            const int q = n - 1 - i, c = n - 1 - j;
            return (c >= 0 && c < kBCols) ? f.br[q][c] : 0.0;
        }
        const int o = j - i;
        return (o >= -f.H && o <= f.H) ? f.stencil[o + 4] : 0.0;
    };
    t.nbl = std::min(kBRows, f.nbl + f.H);
    t.nbr = std::min(kBRows, f.nbr + f.H);
    // B^T(j, c) = B(c, j); right rows are stored reflected
    for (int j = 0; j < t.nbl; j++)
        for (int c = 0; c < kBCols; c++) t.bl[j][c] = Bentry(c, j);
    for (int j = 0; j < t.nbr; j++)
        for (int c = 0; c < kBCols; c++) t.br[j][c] = Bentry(n - 1 - c, n - 1 - j);
}

namespace {

// Band LU (no pivoting) of rows [r0, r0+M) restricted to that block.
// lu[t*5 + {0..4}] = l1, l2, 1/pivot, u1/pivot, u2/pivot.
int band_lu_block(const OpMatrices &op, int r0, int M, double *lu)
{
    const int n = op.n;
    std::vector<double> U((size_t)M * 3, 0.0);  // U rows: offsets 0..2
    for (int t = 0; t < M; t++) {
        const int i = r0 + t;
        double row[5];
        for (int o = -2; o <= 2; o++) {
            const int tt = t + o;
            row[o + 2] = (tt >= 0 && tt < M) ? op.band[(size_t)i * 5 + o + 2] : 0.0;
        }
        (void)n;
        double l1 = 0.0, l2 = 0.0;
        if (t >= 2) {
            const double *u = &U[(size_t)(t - 2) * 3];
            l2 = row[0] / u[0];
            // row t-2 spans columns t-2 .. t  -> offsets -2..0 of row t
            row[0] -= l2 * u[0];
            row[1] -= l2 * u[1];
            row[2] -= l2 * u[2];
        }
        if (t >= 1) {
            const double *u = &U[(size_t)(t - 1) * 3];
            l1 = row[1] / u[0];
            row[1] -= l1 * u[0];
            row[2] -= l1 * u[1];
            row[3] -= l1 * u[2];
        }
        if (row[2] == 0.0 || !std::isfinite(row[2])) return 2;
        U[(size_t)t * 3 + 0] = row[2];
        U[(size_t)t * 3 + 1] = row[3];
        U[(size_t)t * 3 + 2] = row[4];
        const double ip = 1.0 / row[2];
        lu[t * 5 + 0] = l1;
        lu[t * 5 + 1] = l2;
        lu[t * 5 + 2] = ip;
        lu[t * 5 + 3] = row[3] * ip;
        lu[t * 5 + 4] = row[4] * ip;
    }
    return 0;
}

void band_lu_solve(const double *lu, int M, double *x)
{
    for (int t = 0; t < M; t++) {
        double z = x[t];
        if (t >= 1) z -= lu[t * 5 + 0] * x[t - 1];
        if (t >= 2) z -= lu[t * 5 + 1] * x[t - 2];
        x[t] = z;
    }
    for (int t = M - 1; t >= 0; t--) {
        double g = x[t] * lu[t * 5 + 2];
        if (t + 1 < M) g -= lu[t * 5 + 3] * x[t + 1];
        if (t + 2 < M) g -= lu[t * 5 + 4] * x[t + 2];
        x[t] = g;
    }
}

// In-place Gauss-Jordan inverse with partial pivoting (small dense matrices).
int invert(std::vector<double> &a, int m)
{
    std::vector<double> inv((size_t)m * m, 0.0);
    for (int i = 0; i < m; i++) inv[(size_t)i * m + i] = 1.0;
    for (int c = 0; c < m; c++) {
        int piv = c;
        for (int r = c + 1; r < m; r++)
            if (std::fabs(a[(size_t)r * m + c]) > std::fabs(a[(size_t)piv * m + c])) piv = r;
        if (a[(size_t)piv * m + c] == 0.0) return 2;
        if (piv != c)
            for (int k = 0; k < m; k++) {
                std::swap(a[(size_t)piv * m + k], a[(size_t)c * m + k]);
                std::swap(inv[(size_t)piv * m + k], inv[(size_t)c * m + k]);
            }
        const double d = 1.0 / a[(size_t)c * m + c];
        for (int k = 0; k < m; k++) {
            a[(size_t)c * m + k] *= d;
            inv[(size_t)c * m + k] *= d;
        }
        for (int r = 0; r < m; r++) {
            if (r == c) continue;
            const double f = a[(size_t)r * m + c];
            if (f == 0.0) continue;
            for (int k = 0; k < m; k++) {
                a[(size_t)r * m + k] -= f * a[(size_t)c * m + k];
                inv[(size_t)r * m + k] -= f * inv[(size_t)c * m + k];
            }
        }
    }
    a.swap(inv);
    return 0;
}

}  // namespace

int spike_factor(const OpMatrices &op, int M, SpikeFactors &out)
{
    const int n = op.n;
    // partitions of at least 8 rows; a bounded pencil solved whole (one
    // partition, no reduced system) may be as short as 4 (the staggered
    // operators on n = 5..7 cells)
    const bool whole_bounded = !op.periodic && M == n && n >= 4;
    if ((M < 8 && !whole_bounded) || n % M != 0) return 1;
    const int P = n / M;
    out = SpikeFactors();
    out.n = n;
    out.M = M;
    out.P = P;
    out.periodic = op.periodic;
    out.lu.assign((size_t)P * M * 5, 0.0);
    out.S.assign((size_t)P * M * 4, 0.0);
    out.Q.assign((size_t)P * 4 * 4 * P, 0.0);
    std::vector<double> col(M);
    for (int p = 0; p < P; p++) {
        double *lu = &out.lu[(size_t)p * M * 5];
        if (int rc = band_lu_block(op, p * M, M, lu)) return rc;
        // coupling columns: rows of this block that reference columns outside it
        for (int c = 0; c < 4; c++) {
            std::fill(col.begin(), col.end(), 0.0);
            bool any = false;
            for (int t = 0; t < M; t++) {
                const int i = p * M + t;
                for (int o = -2; o <= 2; o++) {
                    const int tt = t + o;
                    if (tt >= 0 && tt < M) continue;
                    const double v = op.band[(size_t)i * 5 + o + 2];
                    if (v == 0.0) continue;
                    // tt = -2,-1 -> tail of left neighbour (c = 0,1);
                    // tt = M, M+1 -> head of right neighbour (c = 2,3)
                    const int cc = tt < 0 ? tt + 2 : 2 + (tt - M);
                    if (cc == c) { col[t] += v; any = true; }
                }
            }
            if (any) band_lu_solve(lu, M, col.data());
            for (int t = 0; t < M; t++) out.S[((size_t)p * M + t) * 4 + c] = any ? col[t] : 0.0;
        }
    }
    if (!op.periodic && P == 1) return 0;
    // reduced system over the tips: unknown 4q + k, k = 0,1 head rows (0,1),
    // k = 2,3 tail rows (M-2, M-1)
    const int m = 4 * P;
    std::vector<double> R((size_t)m * m, 0.0);
    auto nbr_index = [&](int p, int c) -> int {
        if (c < 2) {
            int q = p - 1;
            if (q < 0) { if (!op.periodic) return -1; q += P; }
            return 4 * q + 2 + c;
        }
        int q = p + 1;
        if (q >= P) { if (!op.periodic) return -1; q -= P; }
        return 4 * q + (c - 2);
    };
    const int trow[4] = {0, 1, M - 2, M - 1};
    for (int p = 0; p < P; p++)
        for (int k = 0; k < 4; k++) {
            const int r = 4 * p + k;
            R[(size_t)r * m + r] += 1.0;
            for (int c = 0; c < 4; c++) {
                const int j = nbr_index(p, c);
                if (j < 0) continue;
                R[(size_t)r * m + j] += out.S[((size_t)p * M + trow[k]) * 4 + c];
            }
        }
    if (int rc = invert(R, m)) return rc;
    for (int p = 0; p < P; p++)
        for (int c = 0; c < 4; c++) {
            const int j = nbr_index(p, c);
            for (int k = 0; k < m; k++)
                out.Q[((size_t)p * 4 + c) * m + k] = j < 0 ? 0.0 : R[(size_t)j * m + k];
        }
    return 0;
}

int dense_operator(const OpMatrices &op, std::vector<double> &D)
{
    const int n = op.n;
    // dense A (row-major) and B
    std::vector<double> A((size_t)n * n, 0.0), B((size_t)n * n, 0.0);
    for (int i = 0; i < n; i++)
        for (int o = -2; o <= 2; o++) {
            int j = i + o;
            if (op.periodic) j = ((j % n) + n) % n;
            else if (j < 0 || j >= n) continue;
            A[(size_t)i * n + j] += op.band[(size_t)i * 5 + o + 2];
        }
    for (int i = 0; i < n; i++) {
        const bool left = !op.periodic && i < op.nbl;
        const bool right = !op.periodic && i >= n - op.nbr;
        if (left) {
            for (int c = 0; c < kBCols && c < n; c++) B[(size_t)i * n + c] += op.bl[i][c];
        } else if (right) {
            const int q = n - 1 - i;
            for (int c = 0; c < kBCols && c < n; c++) B[(size_t)i * n + (n - 1 - c)] += op.br[q][c];
        } else {
            for (int o = -op.H; o <= op.H; o++) {
                int j = i + o;
                if (op.periodic) j = ((j % n) + n) % n;
                else if (j < 0 || j >= n) continue;
                B[(size_t)i * n + j] += op.stencil[o + 4];
            }
        }
    }
    if (int rc = invert(A, n)) return rc;
    D.assign((size_t)n * n, 0.0);
    for (int i = 0; i < n; i++)
        for (int k = 0; k < n; k++) {
            const double a = A[(size_t)i * n + k];
            if (a == 0.0) continue;
            const double *b = &B[(size_t)k * n];
            double *d = &D[(size_t)i * n];
            for (int j = 0; j < n; j++) d[j] += a * b[j];
        }
    return 0;
}

}  // namespace pade
