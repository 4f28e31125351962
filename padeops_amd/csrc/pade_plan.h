// Host-side planning for the compact-difference operators.
//
// An operator instance is written as  A y = B f  with A a band matrix of
// half-width 2 (cyclic when periodic; cd06's tridiagonal A has zero outer
// bands) and B a 9-point stencil plus explicit boundary rows.  The planner
// turns (kind, operation, n, dx, periodicity, bc) into
//   * the stencil and boundary-row tables of B (or of B^T for the adjoint),
//   * the factors of a partitioned ("SPIKE") solve of A (or A^T): P
//     partitions of M rows, each with its own band LU, coupling columns S_p
//     and the rows Q_p of the inverse reduced system,
// everything in FP64.  The device kernels consume these tables.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

namespace pade {

enum Kind { CD06 = 1, CD10 = 2, CF90 = 3, CD06STAGG = 4, GAUSSIAN = 5 };
// cd06stagg: edge (E) and cell (C) staggered z operators; an E array holds
// n + 1 planes (periodic: the last a copy of the first).  D1_C2C and D1_E2E
// exist for the bounded form only, as in the reference.
enum Oper { D1 = 1, D2 = 2, FILTER = 3,
            D1_E2C = 11, D1_C2E = 12, D2_C2C = 13, D2_E2E = 14, INTERP_E2C = 15, INTERP_C2E = 16,
            D1_C2C = 17, D1_E2E = 18 };

// Input and output lengths of a staggered operation on n cells.
inline bool stagg_edge_in(int oper)
{
    return oper == D1_E2C || oper == D2_E2E || oper == INTERP_E2C || oper == D1_E2E;
}
inline bool stagg_edge_out(int oper)
{
    return oper == D1_C2E || oper == D2_E2E || oper == INTERP_C2E || oper == D1_E2E;
}

constexpr int kBandW = 2;      // band half-width handled by the solver
constexpr int kStencilH = 4;   // max stencil half-width (cf90)
constexpr int kBRows = 8;      // max special boundary rows (B^T: closure rows + H)
constexpr int kBCols = 16;     // max columns of a special boundary row
constexpr int kLuStride = 5;   // l1 l2 ip u1 u2 per row
constexpr int kSStride = 4;    // S columns: tail(p-1)[0..1], head(p+1)[0..1]

// Dense description of one operator: A band rows and B = stencil + boundary rows.
struct OpMatrices {
    int n = 0;                           // output points = rows of A
    int n_in = 0;                        // input points (columns of B); 0: n
    bool periodic = true;
    // band[i][o+2] = A(i, i+o) (indices wrapped when periodic)
    std::vector<double> band;            // n * 5
    double stencil[9] = {0};             // B interior, offsets -4..4 (1/dx^p folded in)
    int H = 0;                           // stencil half-width actually used
    int nbl = 0, nbr = 0;                // special rows at the left / right end
    // left rows q: out[q] = sum_c bl[q][c] * in[c];
    // right rows q: out[n-1-q] = sum_c br[q][c] * in[n_in-1-c]
    double bl[kBRows][kBCols] = {{0}};
    double br[kBRows][kBCols] = {{0}};
    // System form of a bounded staggered operator (strided-row kernels): the
    // right rows read out[n-1-q] = sum_c br[q][c] * in[n-c] (in may hold
    // n + 1 points), and with `tail` one more output row follows the n solved
    // ones: out[n] = (sum_c br[kBRows-1][c] * in[n-c] - tail_dn * y[n-1]) * tail_idg.
    bool rect = false, tail = false;
    double tail_dn = 0.0, tail_idg = 0.0;
};

// Partitioned solve factors (all FP64 on the host).
struct SpikeFactors {
    int n = 0, M = 0, P = 0;
    bool periodic = true;
    std::vector<double> lu;   // P * M * 5
    std::vector<double> S;    // P * M * 4
    std::vector<double> Q;    // P * 4 * (4P)
};

// Build A and B of the forward operator.  For cd06 bounded, bc1/bcn are the
// init-time codes (they only choose the LHS closure); for bounded cd06stagg
// the init-time end codes (+-1: even / odd, +-2: one-sided closure, parity
// by the sign); otherwise the apply-time codes.  Returns 0 or an error code
// (324 for a bad bc, 21 unsupported).
int build_operator(int kind, int oper, int n, double dx, bool periodic, int bc1, int bcn,
                   OpMatrices &out);

// System form of a bounded staggered operation for the partitioned
// strided-row kernels: the cell-to-edge operators (n + 1 outputs) leave their
// last output row out of the solve (eliminated from the row above it) and
// recover it afterwards; rows and columns as described with OpMatrices::rect.
int build_stagg_system(int oper, int n, double dx, int bc1, int bcn, OpMatrices &out);

// B^T and A^T of an operator (adjoint).
void transpose_operator(const OpMatrices &fwd, OpMatrices &adj);

// Factor A for a partitioned solve with partition size M (n % M == 0).
int spike_factor(const OpMatrices &op, int M, SpikeFactors &out);

// Error code init() returns for this size (0 ok, 2/3/7 as the reference).
int init_error(int kind, int n, bool periodic);

// Dense D = A^{-1} B (row-major n x n), FP64, for the MFMA form.
int dense_operator(const OpMatrices &op, std::vector<double> &D);

}  // namespace pade
