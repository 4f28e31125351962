/*
 * CPU restatement of the reference's banded compact-operator algorithm.
 *
 * TEST INFRASTRUCTURE ONLY: used by tests/ as a second oracle and by the
 * CPU-baseline leg of bench.py as the timed CPU path.  The product never links
 * or calls it.
 *
 * Algorithm followed (per pencil, FP64):
 *   1. RHS: interior stencil (with periodic wrap) plus explicit boundary rows,
 *      as in Compute{X,Y,Z}D1RHS / D2RHS / RHS
 *      (src/derivatives/cd10.F90:1098-2027, cd06.F90:515-742,
 *       src/filters/cf90.F90:610-1018).
 *   2. Periodic: bordered LU of the cyclic band matrix -- the last w unknowns
 *      form the border, the leading (n-w) block is factored as a band LU
 *      without pivoting, the border is solved through its w x w Schur
 *      complement with running dot products during the forward sweep, and the
 *      back substitution carries the border columns
 *      (cd10.F90:351-427 factor, :712-821 sweeps; cd06.F90:221-262, :345-429;
 *       cf90.F90:198-274, :421-530).
 *   3. Bounded: band LU without pivoting, forward and backward sweeps
 *      (cd10.F90:429-709, :823-901; cd06.F90:264-327, :432-490;
 *       cf90.F90:276-418, :532-608).
 *
 * The coefficient tables (band rows of A, interior stencil and boundary rows
 * of B) are supplied by the caller; oracle/pade_dense.py builds them from the
 * reference's coefficient formulas.  Loop order per axis follows the
 * reference: x pencils one at a time, y vectorised over i per k-slab, z
 * vectorised over (i,j) planes.  OpenMP spreads independent pencils/slabs.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define MAXW 2
#define MAXK 9

typedef struct {
    int n, w, periodic;
    /* band LU of the leading block (size m = n or n-w):
       L: unit lower, subdiagonals l[i][1..w] (coefficient of z[i-d]),
       U: diag piv[i] (stored inverted), superdiagonals u[i][1..w] (pre-scaled by 1/piv) */
    double *l;      /* m x MAXW */
    double *ipiv;   /* m */
    double *u;      /* m x MAXW */
    /* periodic border */
    double *V;      /* m x MAXW : L^{-1} A12 */
    double *Wt;     /* MAXW x m : A21 U^{-1} */
    double Sinv[MAXW][MAXW];
    /* RHS */
    int H;
    double st[9];   /* offsets -4..4 */
    int nbl, nbr;
    double bl[4][MAXK], br[4][MAXK];
} oracle_op;

/* A is given as band rows a[i][o+w] = A(i, i+o) (wrapped when periodic). */
static int factor(oracle_op *op, const double *band)
{
    const int n = op->n, w = op->w;
    const int m = op->periodic ? n - w : n;
    op->l = calloc((size_t)m * MAXW, sizeof(double));
    op->ipiv = calloc((size_t)m, sizeof(double));
    op->u = calloc((size_t)m * MAXW, sizeof(double));
    if (!op->l || !op->ipiv || !op->u) return 1;
    /* working copy of the leading m x m band: rows store A(i, i+o), o=-w..w */
    double *a = malloc((size_t)m * (2 * w + 1) * sizeof(double));
    if (!a) return 1;
    for (int i = 0; i < m; i++)
        for (int o = -w; o <= w; o++) {
            int j = i + o;
            a[i * (2 * w + 1) + o + w] = (j >= 0 && j < m) ? band[i * (2 * w + 1) + o + w] : 0.0;
        }
    /* Doolittle band LU without pivoting */
    for (int i = 0; i < m; i++) {
        for (int d = w; d >= 1; d--) {          /* eliminate A(i, i-d) using row i-d */
            int r = i - d;
            if (r < 0) continue;
            double piv = a[r * (2 * w + 1) + w];
            double mult = a[i * (2 * w + 1) + w - d] / piv;
            op->l[i * MAXW + d - 1] = mult;
            for (int o = 0; o <= w; o++) {       /* row r columns r..r+w */
                int col = r + o;                  /* offset in row i: col - i */
                int oi = col - i;
                if (oi < -w || oi > w) continue;
                a[i * (2 * w + 1) + oi + w] -= mult * a[r * (2 * w + 1) + o + w];
            }
            a[i * (2 * w + 1) + w - d] = 0.0;
        }
        double piv = a[i * (2 * w + 1) + w];
        if (piv == 0.0) { free(a); return 2; }
        op->ipiv[i] = 1.0 / piv;
        for (int d = 1; d <= w; d++) op->u[i * MAXW + d - 1] = a[i * (2 * w + 1) + w + d] / piv;
    }
    free(a);
    if (!op->periodic) return 0;
    /* border: A12 (m x w) columns m..n-1, A21 (w x m) rows m..n-1, A22 */
    op->V = calloc((size_t)m * MAXW, sizeof(double));
    op->Wt = calloc((size_t)MAXW * m, sizeof(double));
    if (!op->V || !op->Wt) return 1;
    double A22[MAXW][MAXW] = {{0}};
    for (int c = 0; c < w; c++) {
        /* column m+c of A restricted to rows 0..m-1 */
        double *col = calloc((size_t)m, sizeof(double));
        for (int i = 0; i < m; i++)
            for (int o = -w; o <= w; o++)
                if (((i + o) % n + n) % n == m + c) col[i] += band[i * (2 * w + 1) + o + w];
        /* V[:,c] = L^{-1} col */
        for (int i = 0; i < m; i++) {
            double s = col[i];
            for (int d = 1; d <= w; d++) if (i - d >= 0) s -= op->l[i * MAXW + d - 1] * op->V[(i - d) * MAXW + c];
            op->V[i * MAXW + c] = s;
        }
        free(col);
    }
    for (int rr = 0; rr < w; rr++) {
        int i = m + rr;
        double *row = calloc((size_t)m, sizeof(double));
        for (int o = -w; o <= w; o++) {
            int j = ((i + o) % n + n) % n;
            if (j < m) row[j] += band[i * (2 * w + 1) + o + w];
            else A22[rr][j - m] += band[i * (2 * w + 1) + o + w];
        }
        /* Wt[rr,:] = row U^{-1}: solve x U = row (forward over columns) */
        for (int j = 0; j < m; j++) {
            double s = row[j];
            for (int d = 1; d <= w; d++)
                if (j - d >= 0) s -= op->Wt[rr * m + j - d] * (op->u[(j - d) * MAXW + d - 1] / op->ipiv[j - d]);
            op->Wt[rr * m + j] = s * op->ipiv[j];
        }
        free(row);
    }
    /* S = A22 - Wt V */
    double S[MAXW][MAXW];
    for (int r = 0; r < w; r++)
        for (int c = 0; c < w; c++) {
            double s = A22[r][c];
            for (int j = 0; j < m; j++) s -= op->Wt[r * m + j] * op->V[j * MAXW + c];
            S[r][c] = s;
        }
    if (w == 1) {
        op->Sinv[0][0] = 1.0 / S[0][0];
    } else {
        double det = S[0][0] * S[1][1] - S[0][1] * S[1][0];
        op->Sinv[0][0] = S[1][1] / det;
        op->Sinv[0][1] = -S[0][1] / det;
        op->Sinv[1][0] = -S[1][0] / det;
        op->Sinv[1][1] = S[0][0] / det;
    }
    return 0;
}

/* ------------------------------------------------------------------------ */
/* RHS along one pencil, strided access                                      */

static void rhs_pencil(const oracle_op *op, const double *f, long st, double *r, long rs)
{
    const int n = op->n, H = op->H;
    if (op->periodic) {
        for (int i = 0; i < n; i++) {
            double s = 0.0;
            for (int o = -H; o <= H; o++) {
                int j = i + o;
                j = (j < 0) ? j + n : (j >= n ? j - n : j);
                s += op->st[o + 4] * f[j * st];
            }
            r[i * rs] = s;
        }
        return;
    }
    for (int i = op->nbl; i < n - op->nbr; i++) {
        double s = 0.0;
        for (int o = -H; o <= H; o++) s += op->st[o + 4] * f[(i + o) * st];
        r[i * rs] = s;
    }
    for (int q = 0; q < op->nbl; q++) {
        double s = 0.0;
        for (int c = 0; c < MAXK && c < n; c++) s += op->bl[q][c] * f[c * st];
        r[q * rs] = s;
    }
    for (int q = 0; q < op->nbr; q++) {
        double s = 0.0;
        for (int c = 0; c < MAXK && c < n; c++) s += op->br[q][c] * f[(n - 1 - c) * st];
        r[(n - 1 - q) * rs] = s;
    }
}

/* Interior-specialised stencils: the reference writes each operator's
   stencil out explicitly; these keep the CPU baseline at that cost. */
static void rhs_interior_vec(const oracle_op *op, const double *const *fp, double *r, long len)
{
    /* fp[o+4] points at f(:, i+o) of a slab row; r gets the stencil */
    const int H = op->H;
    const double c0 = op->st[4];
    if (H == 3 && c0 == 0.0) {
        const double a = op->st[5], b = op->st[6], c = op->st[7];
        const double *m1 = fp[3], *p1 = fp[5], *m2 = fp[2], *p2 = fp[6], *m3 = fp[1], *p3 = fp[7];
        for (long t = 0; t < len; t++) r[t] = a * (p1[t] - m1[t]) + b * (p2[t] - m2[t]) + c * (p3[t] - m3[t]);
        return;
    }
    for (long t = 0; t < len; t++) {
        double s = 0.0;
        for (int o = -H; o <= H; o++) s += op->st[o + 4] * fp[o + 4][t];
        r[t] = s;
    }
}

/* ------------------------------------------------------------------------ */
/* Solve along one pencil (x direction), in place, strided */

static void solve_pencil(const oracle_op *op, double *y, long st)
{
    const int n = op->n, w = op->w;
    const int m = op->periodic ? n - w : n;
    double acc[MAXW] = {0.0, 0.0};
    for (int i = 0; i < m; i++) {
        double z = y[i * st];
        for (int d = 1; d <= w; d++) if (i - d >= 0) z -= op->l[i * MAXW + d - 1] * y[(i - d) * st];
        y[i * st] = z;
        if (op->periodic) for (int r = 0; r < w; r++) acc[r] += op->Wt[r * m + i] * z;
    }
    double b[MAXW] = {0.0, 0.0};
    if (op->periodic) {
        double s[MAXW];
        for (int r = 0; r < w; r++) s[r] = y[(m + r) * st] - acc[r];
        for (int r = 0; r < w; r++) {
            b[r] = 0.0;
            for (int c = 0; c < w; c++) b[r] += op->Sinv[r][c] * s[c];
        }
        for (int r = 0; r < w; r++) y[(m + r) * st] = b[r];
    }
    for (int i = m - 1; i >= 0; i--) {
        double z = y[i * st];
        if (op->periodic) for (int c = 0; c < w; c++) z -= op->V[i * MAXW + c] * b[c];
        z *= op->ipiv[i];
        for (int d = 1; d <= w; d++) if (i + d < m) z -= op->u[i * MAXW + d - 1] * y[(i + d) * st];
        y[i * st] = z;
    }
}

/* Solve along j for a slab f(n1, n, .): vectorised over i as the reference */
static void solve_slab(const oracle_op *op, double *y, long len, long st, double *acc)
{
    const int n = op->n, w = op->w;
    const int m = op->periodic ? n - w : n;
    if (op->periodic) memset(acc, 0, sizeof(double) * (size_t)len * MAXW);
    for (int i = 0; i < m; i++) {
        double *yi = y + i * st;
        for (int d = 1; d <= w; d++) {
            if (i - d < 0) continue;
            const double l = op->l[i * MAXW + d - 1];
            const double *yd = y + (i - d) * st;
            for (long t = 0; t < len; t++) yi[t] -= l * yd[t];
        }
        if (op->periodic)
            for (int r = 0; r < w; r++) {
                const double wt = op->Wt[r * m + i];
                double *a = acc + r * len;
                for (long t = 0; t < len; t++) a[t] += wt * yi[t];
            }
    }
    if (op->periodic) {
        for (long t = 0; t < len; t++) {
            double s[MAXW], b[MAXW];
            for (int r = 0; r < w; r++) s[r] = y[(m + r) * st + t] - acc[r * len + t];
            for (int r = 0; r < w; r++) {
                b[r] = 0.0;
                for (int c = 0; c < w; c++) b[r] += op->Sinv[r][c] * s[c];
            }
            for (int r = 0; r < w; r++) y[(m + r) * st + t] = b[r];
        }
    }
    for (int i = m - 1; i >= 0; i--) {
        double *yi = y + i * st;
        if (op->periodic)
            for (int c = 0; c < w; c++) {
                const double v = op->V[i * MAXW + c];
                const double *yb = y + (m + c) * st;
                for (long t = 0; t < len; t++) yi[t] -= v * yb[t];
            }
        const double ip = op->ipiv[i];
        for (long t = 0; t < len; t++) yi[t] *= ip;
        for (int d = 1; d <= w; d++) {
            if (i + d >= m) continue;
            const double u = op->u[i * MAXW + d - 1];
            const double *yd = y + (i + d) * st;
            for (long t = 0; t < len; t++) yi[t] -= u * yd[t];
        }
    }
}

static void rhs_slab(const oracle_op *op, const double *f, double *r, long len, long st)
{
    /* f, r: (len, n) with stride st between consecutive points along the pencil */
    const int n = op->n, H = op->H;
    const double *fp[9];
    if (op->periodic) {
        for (int i = 0; i < n; i++) {
            for (int o = -4; o <= 4; o++) {
                int j = i + o;
                j = ((j % n) + n) % n;
                fp[o + 4] = f + (long)j * st;
            }
            rhs_interior_vec(op, fp, r + (long)i * st, len);
        }
        return;
    }
    for (int i = op->nbl; i < n - op->nbr; i++) {
        for (int o = -4; o <= 4; o++) {
            int j = i + o;
            if (j < 0) j = 0;
            if (j >= n) j = n - 1;
            fp[o + 4] = f + (long)j * st;
        }
        (void)H;
        rhs_interior_vec(op, fp, r + (long)i * st, len);
    }
    for (int q = 0; q < op->nbl; q++) {
        double *rq = r + (long)q * st;
        for (long t = 0; t < len; t++) rq[t] = 0.0;
        for (int c = 0; c < MAXK && c < n; c++) {
            const double b = op->bl[q][c];
            if (b == 0.0) continue;
            const double *fc = f + (long)c * st;
            for (long t = 0; t < len; t++) rq[t] += b * fc[t];
        }
    }
    for (int q = 0; q < op->nbr; q++) {
        double *rq = r + (long)(n - 1 - q) * st;
        for (long t = 0; t < len; t++) rq[t] = 0.0;
        for (int c = 0; c < MAXK && c < n; c++) {
            const double b = op->br[q][c];
            if (b == 0.0) continue;
            const double *fc = f + (long)(n - 1 - c) * st;
            for (long t = 0; t < len; t++) rq[t] += b * fc[t];
        }
    }
}

/* ------------------------------------------------------------------------ */
/* Public entry: create / apply / destroy                                    */

void *oracle_create(int n, int w, int periodic, const double *band, int H, const double *st9,
                    int nbl, const double *bl, int nbr, const double *br)
{
    oracle_op *op = calloc(1, sizeof(oracle_op));
    if (!op) return NULL;
    op->n = n; op->w = w; op->periodic = periodic; op->H = H;
    for (int k = 0; k < 9; k++) op->st[k] = st9[k];
    op->nbl = nbl; op->nbr = nbr;
    for (int q = 0; q < nbl; q++) for (int c = 0; c < MAXK; c++) op->bl[q][c] = bl[q * MAXK + c];
    for (int q = 0; q < nbr; q++) for (int c = 0; c < MAXK; c++) op->br[q][c] = br[q * MAXK + c];
    if (factor(op, band) != 0) { free(op); return NULL; }
    return op;
}

void oracle_destroy(void *h)
{
    oracle_op *op = h;
    if (!op) return;
    free(op->l); free(op->ipiv); free(op->u); free(op->V); free(op->Wt);
    free(op);
}

/* axis 1: f(n, na, nb); axis 2: f(na, n, nb); axis 3: f(na, nb, n) */
int oracle_apply(void *h, int axis, const double *f, double *df, long na, long nb, int nthreads)
{
    const oracle_op *op = h;
    const long n = op->n;
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#else
    (void)nthreads;
#endif
    if (axis == 1) {
        const long np = na * nb;
#pragma omp parallel for schedule(static)
        for (long q = 0; q < np; q++) {
            rhs_pencil(op, f + q * n, 1, df + q * n, 1);
            solve_pencil(op, df + q * n, 1);
        }
        return 0;
    }
    if (axis == 2) {
#pragma omp parallel
        {
            double *acc = malloc(sizeof(double) * (size_t)na * MAXW);
#pragma omp for schedule(static)
            for (long k = 0; k < nb; k++) {
                rhs_slab(op, f + k * na * n, df + k * na * n, na, na);
                solve_slab(op, df + k * na * n, na, na, acc);
            }
            free(acc);
        }
        return 0;
    }
    if (axis == 3) {
        const long plane = na * nb;
        const long blk = 4096;
        const long nblk = (plane + blk - 1) / blk;
#pragma omp parallel
        {
            double *acc = malloc(sizeof(double) * (size_t)blk * MAXW);
#pragma omp for schedule(static)
            for (long b = 0; b < nblk; b++) {
                long off = b * blk;
                long len = (plane - off < blk) ? plane - off : blk;
                rhs_slab(op, f + off, df + off, len, plane);
                solve_slab(op, df + off, len, plane, acc);
            }
            free(acc);
        }
        return 0;
    }
    return 1;
}

int oracle_max_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}
