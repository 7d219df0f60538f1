/*
 * hmx_oracle.c -- CPU restatement of the KernelMatrix compression and mul! of
 * HierarchicalMatrices.jl (v0.2.2), used as the parity oracle.
 *
 * TEST INFRASTRUCTURE ONLY: loaded by tests/, the smoke check and bench.py's
 * cpu_baseline leg, and never by the product library.
 *
 * Every function follows the cited reference lines loop for loop, with the same
 * operation order and no FMA contraction (built with -ffp-contract=off), so the
 * generated factors are bit-identical to what the Julia code computes from the
 * same nodes, and the traversal mul! reproduces the reference's summation
 * order. The only places where the last bit can differ from Julia are sinpi
 * (computed here in long double, then rounded) and log (libm).
 *
 * Pinning (what checks this file, see tests/test_oracle.py):
 *   or_chebyshevpoints / weights  -> mpmath at 50 digits (<= 1 ulp)
 *   or_blockrank                  -> the reference's value table (Float16..BigFloat)
 *   or_indsplit / or_tree_build   -> brute-force tiling, admissibility, block counts
 *   far factors U, V, F           -> partition of unity, polynomial reproduction,
 *                                    closed form F = f(t_m, s_n)
 *   leaf kernels                  -> restated known-answer tests of test/runtests.jl:15-52
 *                                    (exact on small integers) and numpy GEMV
 *   traversal mul!                -> dense long-or_real K*b within the interpolation bound
 */
#include "hmx_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <float.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* ---------------------------------------------------------------- kernels */

/* examples/Kernel.jl:34-37. Julia lowers (x-y)^2 to d*d and (x-y)^3 to (d*d)*d
 * (Base.literal_pow), and inv(d) to 1/d. */
or_real or_kernel(int kernel_id, or_real x, or_real y) {
    or_real d = x - y;
    switch (kernel_id) {
    case OR_CAUCHY: return R(1) / d;
    case OR_COULOMB: return R(1) / (d * d);
    case OR_COULOMB3: return R(1) / ((d * d) * d);
    case OR_LOG: return OR_LOGF(OR_FABS(d));
    default: return NAN;
    }
}

/* src/HierarchicalMatrices.jl:5 in or_real arithmetic; for mant_bits <= 52 the
 * result equals the one computed in the narrower type (values 4, 10, 20, 100
 * for Float16, Float32, Float64, BigFloat(256)). */
int or_blockrank(int mant_bits) {
    double inveps = ldexp(1.0, mant_bits);
    double v = 0.5 * (log(inveps) / log(3.0 + sqrt(8.0)));
    return 2 * (int)lround(v);
}

or_real or_sinpi(or_real q) {
    const long double PI_L = 3.14159265358979323846264338327950288419716939937510L;
    return (or_real)sinl(PI_L * (long double)q);
}

/* src/BarycentricMatrix.jl:92-111 */
void or_chebyshevpoints(or_real *x, int n, int kind) {
    for (int k = 0; k < n; ++k) x[k] = 0.0;
    int nd2 = n / 2;
    if (kind == 1) {
        for (int k = 1; k <= nd2; ++k)
            x[k - 1] = or_sinpi(((or_real)(n - 2 * k) + R(1)) / (or_real)(2 * n));
        for (int k = 1; k <= nd2; ++k) x[n - k] = -x[k - 1];
    } else {
        for (int k = 1; k <= nd2; ++k)
            x[k - 1] = or_sinpi(((or_real)(n - 2 * k) + R(1)) / (or_real)(2 * (n - 1)));
        for (int k = 1; k <= nd2; ++k) x[n - k] = -x[k - 1];
    }
}

/* src/BarycentricMatrix.jl:114-136 */
void or_chebyshevbarycentricweights(or_real *l, int n, int kind) {
    for (int k = 0; k < n; ++k) l[k] = 0.0;
    int nd2 = n / 2;
    if (kind == 1) {
        for (int k = 1; k <= nd2 + 1 && k <= n; ++k)
            l[k - 1] = or_sinpi(((or_real)(2 * k) - R(1)) / (or_real)(2 * n));
        for (int k = 1; k <= nd2; ++k) l[n - k] = l[k - 1];
        for (int k = 2; k <= n; k += 2) l[k - 1] *= R(-1);
    } else {
        for (int k = 0; k < n; ++k) l[k] = 1.0;
        for (int k = 2; k <= n; k += 2) l[k - 1] *= R(-1);
        l[0] *= R(0.5);
        l[n - 1] *= R(0.5);
    }
}

/* src/BarycentricMatrix.jl:299-307 */
int or_indsplit(const or_real *x, int64_t nx, int64_t first, int64_t len, or_real a,
                or_real b, int64_t out[4]) {
    int64_t last = first + len - 1;
    int64_t i = first;
    or_real ab2 = R(0.5) * (a + b);
    for (;;) {
        if (i < 1 || i > nx) return -1; /* Julia BoundsError */
        if (!(x[i - 1] >= ab2)) break;
        i += 1;
        if (i > last) break;
    }
    int64_t l1 = i - 1 - first + 1, l2 = last - i + 1;
    out[0] = first;
    out[1] = l1 > 0 ? l1 : 0;
    out[2] = i;
    out[3] = l2 > 0 ? l2 : 0;
    return 0;
}

/* ------------------------------------------------------------------- tree */

typedef struct or_node or_node;

typedef struct {
    int code;          /* OR_CODE_* */
    int depth;         /* depth of this block (root children have depth 1) */
    int64_t i0, m;     /* 1-based first row index into x, and row count */
    int64_t j0, n;     /* 1-based first column index into y, and column count */
    or_real a, b, c, d; /* intervals the block was generated with */
    or_node *node;     /* code KM */
    or_real *A;         /* dense m x n, column-major (store_factors only) */
    or_real *tx, *ty;   /* far: mapped nodes (always) */
    or_real *U, *V, *F; /* far factors (store_factors only) */
} or_child;

struct or_node {
    int type, depth;
    or_child ch[4]; /* (1,1) (1,2) (2,1) (2,2) */
};

typedef struct {
    or_child *c;
    int64_t row0, col0; /* 0-based traversal offsets */
} or_flat;

struct or_tree {
    int kernel_id, r, bs, store, quirk, err;
    const or_real *x, *y; /* borrowed during build only */
    or_real *xs, *ys;     /* owned copies */
    int64_t nx, ny;
    or_real *xhat, *lam; /* chebyshevpoints / weights of rank r */
    or_child root;
    int64_t nd, nk1, nk2, ndense, nfar, dense_entries, far_rows, far_cols;
    int maxdepth;
    or_flat *flat;
    int64_t nflat, capflat;
};

static void *xmalloc(size_t n) {
    void *p = malloc(n ? n : 1);
    return p;
}

/* src/KernelMatrix.jl:57-60 (and :82, :104): T[f(x[i], y[j]) for i in ir, j in jr] */
static void gen_dense(or_tree *t, or_child *ch) {
    ch->A = (or_real *)xmalloc(sizeof(or_real) * (size_t)(ch->m * ch->n));
    for (int64_t j = 0; j < ch->n; ++j)
        for (int64_t i = 0; i < ch->m; ++i)
            ch->A[i + ch->m * j] = or_kernel(t->kernel_id, t->xs[ch->i0 - 1 + i], t->ys[ch->j0 - 1 + j]);
}

/* nodes of BarycentricPoly2D, src/BarycentricMatrix.jl:147-166 */
static void gen_nodes(or_tree *t, or_child *ch) {
    int r = t->r;
    ch->tx = (or_real *)xmalloc(sizeof(or_real) * r);
    ch->ty = (or_real *)xmalloc(sizeof(or_real) * r);
    or_real ab2 = R(0.5) * (ch->a + ch->b), ba2 = R(0.5) * (ch->b - ch->a);
    for (int p = 0; p < r; ++p) ch->tx[p] = ab2 + ba2 * t->xhat[p];
    or_real cd2 = R(0.5) * (ch->c + ch->d), dc2 = R(0.5) * (ch->d - ch->c);
    for (int q = 0; q < r; ++q) ch->ty[q] = cd2 + dc2 * t->xhat[q];
}

/* one row of update!, src/BarycentricMatrix.jl:256-267 (U) / :272-283 (V):
 * row[m] = lam[m]*inv(p - t[m]), then row[m] /= sum(row) */
static void basis_row(const or_real *tn, const or_real *lam, int r, or_real p, or_real *row) {
    for (int m = 0; m < r; ++m) row[m] = lam[m] * (R(1) / (p - tn[m]));
    or_real s = 0.0;
    for (int m = 0; m < r; ++m) s += row[m];
    for (int m = 0; m < r; ++m) row[m] /= s;
}

/* F of BarycentricPoly2D (src/BarycentricMatrix.jl:168-175) and update!
 * (src/BarycentricMatrix.jl:248-287). */
static void gen_far(or_tree *t, or_child *ch) {
    int r = t->r;
    ch->F = (or_real *)xmalloc(sizeof(or_real) * r * r);
    for (int n = 0; n < r; ++n)
        for (int m = 0; m < r; ++m) ch->F[m + r * n] = or_kernel(t->kernel_id, ch->tx[m], ch->ty[n]);
    ch->U = (or_real *)xmalloc(sizeof(or_real) * (size_t)(ch->m * r));
    ch->V = (or_real *)xmalloc(sizeof(or_real) * (size_t)(ch->n * r));
    or_real row[256];
    for (int64_t i = 0; i < ch->m; ++i) {
        basis_row(ch->tx, t->lam, r, t->xs[ch->i0 - 1 + i], row);
        for (int m = 0; m < r; ++m) ch->U[i + ch->m * m] = row[m];
    }
    for (int64_t j = 0; j < ch->n; ++j) {
        basis_row(ch->ty, t->lam, r, t->ys[ch->j0 - 1 + j], row);
        for (int m = 0; m < r; ++m) ch->V[j + ch->n * m] = row[m];
    }
}

static void set_dense(or_tree *t, or_child *ch, int depth, int64_t i0, int64_t m, int64_t j0,
                      int64_t n) {
    memset(ch, 0, sizeof(*ch));
    ch->code = OR_CODE_DENSE;
    ch->depth = depth;
    ch->i0 = i0; ch->m = m; ch->j0 = j0; ch->n = n;
    t->ndense++;
    t->dense_entries += m * n;
    if (t->store) gen_dense(t, ch);
}

/* BarycentricMatrix2D(T, f, a, b, c, d, x, y, ir, jr), src/BarycentricMatrix.jl:236-246 */
static void set_far(or_tree *t, or_child *ch, int depth, or_real a, or_real b, or_real c, or_real d,
                    int64_t i0, int64_t m, int64_t j0, int64_t n) {
    memset(ch, 0, sizeof(*ch));
    ch->code = OR_CODE_BARY;
    ch->depth = depth;
    ch->i0 = i0; ch->m = m; ch->j0 = j0; ch->n = n;
    ch->a = a; ch->b = b; ch->c = c; ch->d = d;
    t->nfar++;
    t->far_rows += m;
    t->far_cols += n;
    gen_nodes(t, ch);
    if (t->store) gen_far(t, ch);
}

static or_node *build_node(or_tree *t, int type, int depth, int64_t i0, int64_t ni, int64_t j0,
                           int64_t nj, or_real a, or_real b, or_real c, or_real d);

static void set_node(or_tree *t, or_child *ch, int type, int depth, int64_t i0, int64_t ni,
                     int64_t j0, int64_t nj, or_real a, or_real b, or_real c, or_real d) {
    memset(ch, 0, sizeof(*ch));
    ch->code = OR_CODE_KM;
    ch->depth = depth;
    ch->i0 = i0; ch->m = ni; ch->j0 = j0; ch->n = nj;
    ch->a = a; ch->b = b; ch->c = c; ch->d = d;
    ch->node = build_node(t, type, depth, i0, ni, j0, nj, a, b, c, d);
}

/* KernelMatrix / KernelMatrix1 / KernelMatrix2, src/KernelMatrix.jl:49-116 */
static or_node *build_node(or_tree *t, int type, int depth, int64_t i0, int64_t ni, int64_t j0,
                           int64_t nj, or_real a, or_real b, or_real c, or_real d) {
    if (t->err) return NULL;
    if (depth > t->maxdepth) t->maxdepth = depth;
    if (depth > 200) { t->err = 2; return NULL; } /* the reference would recurse without bound */
    if (type == OR_NODE_D) t->nd++;
    else if (type == OR_NODE_K1) t->nk1++;
    else t->nk2++;
    if (ni == 0 || nj == 0) t->quirk = 1;
    int64_t is[4], js[4];
    if (or_indsplit(t->xs, t->nx, i0, ni, a, b, is) || or_indsplit(t->ys, t->ny, j0, nj, c, d, js)) {
        t->err = 1;
        return NULL;
    }
    int64_t ir1f = is[0], ir1n = is[1], ir2f = is[2], ir2n = is[3];
    int64_t jr1f = js[0], jr1n = js[1], jr2f = js[2], jr2n = js[3];
    or_real ab2 = R(0.5) * (a + b), cd2 = R(0.5) * (c + d);
    int bs = t->bs;
    int leaf = ir1n < bs && ir2n < bs && jr1n < bs && jr2n < bs;
    or_node *H = (or_node *)calloc(1, sizeof(or_node));
    H->type = type;
    H->depth = depth;
    int cd = depth + 1;
    if (type == OR_NODE_D) {
        if (leaf) {
            set_dense(t, &H->ch[0], cd, ir1f, ir1n, jr1f, jr1n);
            set_dense(t, &H->ch[1], cd, ir1f, ir1n, jr2f, jr2n);
            set_dense(t, &H->ch[2], cd, ir2f, ir2n, jr1f, jr1n);
            set_dense(t, &H->ch[3], cd, ir2f, ir2n, jr2f, jr2n);
        } else {
            set_node(t, &H->ch[0], OR_NODE_D, cd, ir1f, ir1n, jr1f, jr1n, a, ab2, c, cd2);
            set_node(t, &H->ch[1], OR_NODE_K1, cd, ir1f, ir1n, jr2f, jr2n, a, ab2, cd2, d);
            set_node(t, &H->ch[2], OR_NODE_K2, cd, ir2f, ir2n, jr1f, jr1n, ab2, b, c, cd2);
            set_node(t, &H->ch[3], OR_NODE_D, cd, ir2f, ir2n, jr2f, jr2n, ab2, b, cd2, d);
        }
    } else if (type == OR_NODE_K1) {
        set_far(t, &H->ch[0], cd, a, ab2, c, cd2, ir1f, ir1n, jr1f, jr1n);
        set_far(t, &H->ch[1], cd, a, ab2, cd2, d, ir1f, ir1n, jr2f, jr2n);
        if (leaf) set_dense(t, &H->ch[2], cd, ir2f, ir2n, jr1f, jr1n);
        else set_node(t, &H->ch[2], OR_NODE_K1, cd, ir2f, ir2n, jr1f, jr1n, ab2, b, c, cd2);
        set_far(t, &H->ch[3], cd, ab2, b, cd2, d, ir2f, ir2n, jr2f, jr2n);
    } else {
        set_far(t, &H->ch[0], cd, a, ab2, c, cd2, ir1f, ir1n, jr1f, jr1n);
        if (leaf) set_dense(t, &H->ch[1], cd, ir1f, ir1n, jr2f, jr2n);
        else set_node(t, &H->ch[1], OR_NODE_K2, cd, ir1f, ir1n, jr2f, jr2n, a, ab2, cd2, d);
        set_far(t, &H->ch[2], cd, ab2, b, c, cd2, ir2f, ir2n, jr1f, jr1n);
        set_far(t, &H->ch[3], cd, ab2, b, cd2, d, ir2f, ir2n, jr2f, jr2n);
    }
    return H;
}

/* size(H, k) of the @hierarchical type, src/hierarchical.jl:33-47,76-97: rows
 * summed down the last column, columns along the first row. */
static int64_t child_size(const or_child *ch, int k);
static int64_t node_size(const or_node *H, int k) {
    if (k == 1) return child_size(&H->ch[1], 1) + child_size(&H->ch[3], 1);
    return child_size(&H->ch[0], 2) + child_size(&H->ch[1], 2);
}
static int64_t child_size(const or_child *ch, int k) {
    if (ch->code == OR_CODE_KM) return node_size(ch->node, k);
    if (ch->code == OR_CODE_NONE) return 0;
    return k == 1 ? ch->m : ch->n;
}

static void flat_push(or_tree *t, or_child *c, int64_t r0, int64_t c0) {
    if (t->nflat == t->capflat) {
        t->capflat = t->capflat ? 2 * t->capflat : 1024;
        t->flat = (or_flat *)realloc(t->flat, sizeof(or_flat) * (size_t)t->capflat);
    }
    t->flat[t->nflat].c = c;
    t->flat[t->nflat].row0 = r0;
    t->flat[t->nflat].col0 = c0;
    t->nflat++;
}

/* the generated traversal of src/KernelMatrix.jl:17-45, recording the leaf
 * calls mul!(u, block, v, istart + p, jstart + q) in order */
static void flatten(or_tree *t, or_node *H, int64_t istart, int64_t jstart) {
    int64_t p = 0;
    for (int m = 0; m < 2; ++m) {
        int64_t q = 0;
        for (int n = 0; n < 2; ++n) {
            or_child *c = &H->ch[2 * m + n];
            if (c->code == OR_CODE_KM) flatten(t, c->node, istart + p, jstart + q);
            else if (c->code != OR_CODE_NONE) flat_push(t, c, istart + p - 1, jstart + q - 1);
            q += child_size(&H->ch[n], 2);
        }
        p += child_size(&H->ch[2 * m + 1], 1);
    }
}

or_tree *or_tree_build(int kernel_id, const or_real *x, const or_real *y, int64_t nx, int64_t ny,
                       or_real a, or_real b, or_real c, or_real d, int blocksize, int rank,
                       int store_factors, int *quirk) {
    if (rank < 1 || rank > 256 || blocksize < 1) return NULL;
    or_tree *t = (or_tree *)calloc(1, sizeof(or_tree));
    t->kernel_id = kernel_id;
    t->r = rank;
    t->bs = blocksize;
    t->store = store_factors;
    t->nx = nx;
    t->ny = ny;
    t->xs = (or_real *)xmalloc(sizeof(or_real) * (size_t)nx);
    t->ys = (or_real *)xmalloc(sizeof(or_real) * (size_t)ny);
    memcpy(t->xs, x, sizeof(or_real) * (size_t)nx);
    memcpy(t->ys, y, sizeof(or_real) * (size_t)ny);
    t->xhat = (or_real *)xmalloc(sizeof(or_real) * rank);
    t->lam = (or_real *)xmalloc(sizeof(or_real) * rank);
    or_chebyshevpoints(t->xhat, rank, 1);
    or_chebyshevbarycentricweights(t->lam, rank, 1);
    /* KernelMatrix(f, x, y, a, b, c, d), src/KernelMatrix.jl:47 */
    set_node(t, &t->root, OR_NODE_D, 0, 1, nx, 1, ny, a, b, c, d);
    t->root.depth = 0;
    if (quirk) *quirk = t->quirk;
    if (t->err) {
        or_tree_free(t);
        return NULL;
    }
    flatten(t, t->root.node, 1, 1);
    return t;
}

static void free_child(or_child *c) {
    if (c->code == OR_CODE_KM && c->node) {
        for (int k = 0; k < 4; ++k) free_child(&c->node->ch[k]);
        free(c->node);
    }
    free(c->A); free(c->tx); free(c->ty); free(c->U); free(c->V); free(c->F);
}

void or_tree_free(or_tree *t) {
    if (!t) return;
    free_child(&t->root);
    free(t->xs); free(t->ys); free(t->xhat); free(t->lam); free(t->flat);
    free(t);
}

void or_tree_stats(const or_tree *t, int64_t out[9]) {
    out[0] = t->nd; out[1] = t->nk1; out[2] = t->nk2;
    out[3] = t->ndense; out[4] = t->nfar; out[5] = t->dense_entries;
    out[6] = t->far_rows; out[7] = t->far_cols; out[8] = t->maxdepth;
}

int64_t or_tree_blocks(const or_tree *t, int64_t *ints, or_real *dbls, int64_t cap) {
    for (int64_t k = 0; k < t->nflat && k < cap; ++k) {
        const or_child *c = t->flat[k].c;
        int64_t *o = ints + 6 * k;
        o[0] = c->code; o[1] = c->depth; o[2] = t->flat[k].row0; o[3] = t->flat[k].col0;
        o[4] = c->m; o[5] = c->n;
        or_real *g = dbls + 4 * k;
        g[0] = c->a; g[1] = c->b; g[2] = c->c; g[3] = c->d;
    }
    return t->nflat;
}

int or_tree_block_data(const or_tree *t, int64_t bidx, or_real *A, or_real *U, or_real *V, or_real *F) {
    if (!t->store || bidx < 0 || bidx >= t->nflat) return -1;
    const or_child *c = t->flat[bidx].c;
    if (c->code == OR_CODE_DENSE) {
        if (A) memcpy(A, c->A, sizeof(or_real) * (size_t)(c->m * c->n));
    } else {
        int r = t->r;
        if (U) memcpy(U, c->U, sizeof(or_real) * (size_t)(c->m * r));
        if (V) memcpy(V, c->V, sizeof(or_real) * (size_t)(c->n * r));
        if (F) memcpy(F, c->F, sizeof(or_real) * (size_t)(r * r));
    }
    return 0;
}

/* ------------------------------------------------------------ leaf kernels */

/* src/algebra.jl:37-48 */
void or_dense_mul(or_real *y, const or_real *A, int64_t m, int64_t n, const or_real *x, int64_t istart,
                  int64_t jstart, int64_t incx, int64_t incy) {
    int64_t ishift = istart - incy, jshift = jstart - incx;
    for (int64_t j = 1; j <= n; ++j) {
        or_real xj = x[jshift + j * incx - 1];
        for (int64_t i = 1; i <= m; ++i) y[ishift + i * incy - 1] += A[(i - 1) + m * (j - 1)] * xj;
    }
}

/* src/algebra.jl:52-65 (transpose(A) with A m x n) */
void or_dense_tmul(or_real *y, const or_real *A, int64_t m, int64_t n, const or_real *x,
                   int64_t istart, int64_t jstart, int64_t incx, int64_t incy) {
    int64_t ishift = istart - incy, jshift = jstart - incx;
    for (int64_t i = 1; i <= n; ++i) {
        or_real yi = 0.0;
        for (int64_t j = 1; j <= m; ++j) yi += A[(j - 1) + m * (i - 1)] * x[jshift + j * incx - 1];
        y[ishift + i * incy - 1] += yi;
    }
}

/* src/algebra.jl:110-131 (U m x r, V n x r, S = diag) */
void or_lowrank_mul(or_real *y, const or_real *U, const or_real *S, const or_real *V, int64_t m,
                    int64_t n, int64_t r, const or_real *x, int64_t istart, int64_t jstart,
                    int64_t incx, int64_t incy) {
    int64_t ishift = istart - incy, jshift = jstart - incx;
    or_real *temp = (or_real *)xmalloc(sizeof(or_real) * (size_t)r);
    for (int64_t k = 0; k < r; ++k) {
        temp[k] = 0.0;
        for (int64_t j = 1; j <= n; ++j) temp[k] += V[(j - 1) + n * k] * x[jshift + j * incx - 1];
        temp[k] *= S[k];
    }
    for (int64_t k = 0; k < r; ++k) {
        or_real tk = temp[k];
        for (int64_t i = 1; i <= m; ++i) y[ishift + i * incy - 1] += U[(i - 1) + m * k] * tk;
    }
    free(temp);
}

/* src/algebra.jl:137-159: adjoint(L) = V S U', size n x m */
void or_lowrank_adjmul(or_real *y, const or_real *U, const or_real *S, const or_real *V, int64_t m,
                       int64_t n, int64_t r, const or_real *x, int64_t istart, int64_t jstart,
                       int64_t incx, int64_t incy) {
    int64_t ishift = istart - incy, jshift = jstart - incx;
    or_real *temp = (or_real *)xmalloc(sizeof(or_real) * (size_t)r);
    for (int64_t k = 0; k < r; ++k) {
        temp[k] = 0.0;
        for (int64_t j = 1; j <= m; ++j) temp[k] += U[(j - 1) + m * k] * x[jshift + j * incx - 1];
        temp[k] *= S[k];
    }
    for (int64_t k = 0; k < r; ++k) {
        or_real tk = temp[k];
        for (int64_t i = 1; i <= n; ++i) y[ishift + i * incy - 1] += V[(i - 1) + n * k] * tk;
    }
    free(temp);
}

/* src/algebra.jl:243-277 */
void or_bary_mul(or_real *u, const or_real *U, const or_real *F, const or_real *V, int64_t m,
                 int64_t n, int64_t r, const or_real *v, int64_t istart, int64_t jstart) {
    int64_t ishift = istart - 1, jshift = jstart - 1;
    or_real *t1 = (or_real *)xmalloc(sizeof(or_real) * (size_t)r);
    or_real *t2 = (or_real *)xmalloc(sizeof(or_real) * (size_t)r);
    for (int64_t k = 0; k < r; ++k) {
        or_real s = 0.0;
        for (int64_t j = 1; j <= n; ++j) s += V[(j - 1) + n * k] * v[jshift + j - 1];
        t1[k] = s;
        t2[k] = 0.0;
    }
    for (int64_t l = 0; l < r; ++l) {
        or_real t1l = t1[l];
        for (int64_t k = 0; k < r; ++k) t2[k] += F[k + r * l] * t1l;
    }
    for (int64_t k = 0; k < r; ++k) {
        or_real t2k = t2[k];
        for (int64_t i = 1; i <= m; ++i) u[ishift + i - 1] += U[(i - 1) + m * k] * t2k;
    }
    free(t1);
    free(t2);
}

/* evaluate(B::BarycentricPoly2D, x, y), src/BarycentricMatrix.jl:180-205:
 * tx, ty the mapped nodes, lx, ly the weights, F r x r column-major. */
or_real or_bary_evaluate(int r, const or_real *tx, const or_real *ty, const or_real *lx,
                         const or_real *ly, const or_real *F, or_real x, or_real y) {
    or_real temp1 = R(0), temp2 = R(0), temp3, ret = R(0);
    for (int p = 0; p < r; ++p) temp1 += lx[p] * (R(1) / (x - tx[p]));
    for (int q = 0; q < r; ++q) temp2 += ly[q] * (R(1) / (y - ty[q]));
    for (int n = 0; n < r; ++n) {
        temp3 = R(0);
        for (int m = 0; m < r; ++m) temp3 += lx[m] * F[m + r * n] * (R(1) / (x - tx[m]));
        temp3 *= R(1) / temp1;
        ret += ly[n] * temp3 * (R(1) / (y - ty[n]));
    }
    ret *= R(1) / temp2;
    return ret;
}

/* --------------------------------------------------------------- matvecs */

int or_tree_matvec(or_tree *t, or_real *u, const or_real *v) {
    if (!t->store) return -1;
    for (int64_t k = 0; k < t->nflat; ++k) {
        or_child *c = t->flat[k].c;
        int64_t is = t->flat[k].row0 + 1, js = t->flat[k].col0 + 1;
        if (c->code == OR_CODE_DENSE) or_dense_mul(u, c->A, c->m, c->n, v, is, js, 1, 1);
        else or_bary_mul(u, c->U, c->F, c->V, c->m, c->n, t->r, v, is, js);
    }
    return 0;
}

/* temp2 = F * (V' v_block) of src/algebra.jl:249-265, with V rows regenerated
 * (bit-identical to update!) unless stored. */
static void far_temp2(const or_tree *t, const or_flat *fb, const or_real *v, or_real *t2) {
    const or_child *c = fb->c;
    int r = t->r;
    or_real t1[256], row[256];
    for (int k = 0; k < r; ++k) t1[k] = 0.0;
    for (int64_t j = 0; j < c->n; ++j) {
        or_real vj = v[fb->col0 + j];
        if (c->V) {
            for (int k = 0; k < r; ++k) t1[k] += c->V[j + c->n * k] * vj;
        } else {
            basis_row(c->ty, t->lam, r, t->ys[c->j0 - 1 + j], row);
            for (int k = 0; k < r; ++k) t1[k] += row[k] * vj;
        }
    }
    for (int k = 0; k < r; ++k) t2[k] = 0.0;
    for (int l = 0; l < r; ++l) {
        or_real t1l = t1[l];
        for (int k = 0; k < r; ++k) {
            or_real F = c->F ? c->F[k + r * l] : or_kernel(t->kernel_id, c->tx[k], c->ty[l]);
            t2[k] += F * t1l;
        }
    }
}

/* phase-2 worker: u[row] accumulated over blocks in traversal order */
static void row_accumulate(const or_tree *t, const or_real *v, const or_real *t2all,
                           const int64_t *faridx, int64_t row, or_real *acc) {
    int r = t->r;
    or_real rowb[256];
    for (int64_t k = 0; k < t->nflat; ++k) {
        const or_flat *fb = &t->flat[k];
        const or_child *c = fb->c;
        if (row < fb->row0 || row >= fb->row0 + c->m) continue;
        int64_t i = row - fb->row0;
        if (c->code == OR_CODE_DENSE) {
            for (int64_t j = 0; j < c->n; ++j) {
                or_real A = c->A ? c->A[i + c->m * j]
                                : or_kernel(t->kernel_id, t->xs[c->i0 - 1 + i], t->ys[c->j0 - 1 + j]);
                *acc += A * v[fb->col0 + j];
            }
        } else {
            const or_real *t2 = t2all + (size_t)r * (size_t)faridx[k];
            if (c->U) {
                for (int kk = 0; kk < r; ++kk) *acc += c->U[i + c->m * kk] * t2[kk];
            } else {
                basis_row(c->tx, t->lam, r, t->xs[c->i0 - 1 + i], rowb);
                for (int kk = 0; kk < r; ++kk) *acc += rowb[kk] * t2[kk];
            }
        }
    }
}

static or_real *all_temp2(const or_tree *t, const or_real *v, int64_t **faridx_out, int nthreads) {
    int64_t *faridx = (int64_t *)xmalloc(sizeof(int64_t) * (size_t)t->nflat);
    int64_t nf = 0;
    for (int64_t k = 0; k < t->nflat; ++k) faridx[k] = t->flat[k].c->code == OR_CODE_BARY ? nf++ : -1;
    int64_t *list = (int64_t *)xmalloc(sizeof(int64_t) * (size_t)(nf + 1));
    for (int64_t k = 0; k < t->nflat; ++k)
        if (faridx[k] >= 0) list[faridx[k]] = k;
    or_real *t2 = (or_real *)xmalloc(sizeof(or_real) * (size_t)(nf * t->r + 1));
#ifdef _OPENMP
    if (nthreads <= 0) nthreads = omp_get_max_threads();
#pragma omp parallel for schedule(dynamic, 4) num_threads(nthreads)
#endif
    for (int64_t f = 0; f < nf; ++f) far_temp2(t, &t->flat[list[f]], v, t2 + (size_t)t->r * f);
    free(list);
    *faridx_out = faridx;
    return t2;
}

int or_tree_matvec_omp(or_tree *t, or_real *u, const or_real *v, int nthreads) {
    if (!t->store) return -1;
    int64_t *faridx;
    or_real *t2 = all_temp2(t, v, &faridx, nthreads);
    /* bucket blocks by 256-row chunks, keeping traversal order */
    const int64_t CH = 256;
    int64_t nch = (t->nx + CH - 1) / CH;
    int64_t *cnt = (int64_t *)calloc((size_t)nch + 1, sizeof(int64_t));
    for (int64_t k = 0; k < t->nflat; ++k) {
        const or_flat *fb = &t->flat[k];
        if (fb->c->m == 0) continue;
        for (int64_t ch = fb->row0 / CH; ch <= (fb->row0 + fb->c->m - 1) / CH; ++ch) cnt[ch + 1]++;
    }
    for (int64_t ch = 0; ch < nch; ++ch) cnt[ch + 1] += cnt[ch];
    int64_t *lst = (int64_t *)xmalloc(sizeof(int64_t) * (size_t)(cnt[nch] + 1));
    int64_t *fill = (int64_t *)calloc((size_t)nch, sizeof(int64_t));
    for (int64_t k = 0; k < t->nflat; ++k) {
        const or_flat *fb = &t->flat[k];
        if (fb->c->m == 0) continue;
        for (int64_t ch = fb->row0 / CH; ch <= (fb->row0 + fb->c->m - 1) / CH; ++ch)
            lst[cnt[ch] + fill[ch]++] = k;
    }
    int r = t->r;
#ifdef _OPENMP
    if (nthreads <= 0) nthreads = omp_get_max_threads();
#pragma omp parallel for schedule(dynamic, 1) num_threads(nthreads)
#endif
    for (int64_t ch = 0; ch < nch; ++ch) {
        int64_t lo = ch * CH, hi = lo + CH < t->nx ? lo + CH : t->nx;
        for (int64_t e = cnt[ch]; e < cnt[ch + 1]; ++e) {
            const or_flat *fb = &t->flat[lst[e]];
            const or_child *c = fb->c;
            int64_t r0 = fb->row0 > lo ? fb->row0 : lo;
            int64_t r1 = fb->row0 + c->m < hi ? fb->row0 + c->m : hi;
            if (c->code == OR_CODE_DENSE) {
                for (int64_t j = 0; j < c->n; ++j) {
                    or_real xj = v[fb->col0 + j];
                    for (int64_t i = r0; i < r1; ++i) u[i] += c->A[(i - fb->row0) + c->m * j] * xj;
                }
            } else {
                const or_real *tt = t2 + (size_t)r * (size_t)faridx[lst[e]];
                for (int kk = 0; kk < r; ++kk) {
                    or_real tk = tt[kk];
                    for (int64_t i = r0; i < r1; ++i) u[i] += c->U[(i - fb->row0) + c->m * kk] * tk;
                }
            }
        }
    }
    free(cnt); free(lst); free(fill); free(t2); free(faridx);
    return 0;
}

int or_tree_matvec_rows(or_tree *t, const or_real *v, const int64_t *rows, int64_t nrows,
                        or_real *out, int nthreads) {
    int64_t *faridx;
    or_real *t2 = all_temp2(t, v, &faridx, nthreads);
#ifdef _OPENMP
    if (nthreads <= 0) nthreads = omp_get_max_threads();
#pragma omp parallel for schedule(dynamic, 1) num_threads(nthreads)
#endif
    for (int64_t k = 0; k < nrows; ++k) {
        or_real acc = 0.0;
        row_accumulate(t, v, t2, faridx, rows[k], &acc);
        out[k] = acc;
    }
    free(t2);
    free(faridx);
    return 0;
}

/* getindex, src/hierarchical.jl:120-147 with the leaf getindex of
 * src/BarycentricMatrix.jl:222-234 and a dense lookup */
or_real or_tree_getindex(or_tree *t, int64_t i, int64_t j) {
    or_node *H = t->root.node;
    int64_t ii = i + 1, jj = j + 1; /* 1-based, as in the reference */
    for (;;) {
        int m = 1, n = 1;
        while (m <= 2) {
            int64_t rr = child_size(&H->ch[2 * (m - 1) + 1], 1);
            if (ii > rr) { ii -= rr; m += 1; } else break;
        }
        while (n <= 2) {
            int64_t s = child_size(&H->ch[n - 1], 2);
            if (jj > s) { jj -= s; n += 1; } else break;
        }
        if (m > 2 || n > 2) return 0.0;
        or_child *c = &H->ch[2 * (m - 1) + (n - 1)];
        if (c->code == OR_CODE_KM) { H = c->node; continue; }
        if (c->code == OR_CODE_NONE) return 0.0;
        if (c->code == OR_CODE_DENSE) {
            if (c->A) return c->A[(ii - 1) + c->m * (jj - 1)];
            return or_kernel(t->kernel_id, t->xs[c->i0 - 2 + ii], t->ys[c->j0 - 2 + jj]);
        }
        int r = t->r;
        or_real ur[256], vr[256];
        basis_row(c->tx, t->lam, r, t->xs[c->i0 - 2 + ii], ur);
        basis_row(c->ty, t->lam, r, t->ys[c->j0 - 2 + jj], vr);
        or_real ret = 0.0;
        for (int k = 0; k < r; ++k) {
            or_real tmp = 0.0;
            for (int l = 0; l < r; ++l) {
                or_real F = c->F ? c->F[k + r * l] : or_kernel(t->kernel_id, c->tx[k], c->ty[l]);
                tmp += F * vr[l];
            }
            ret += ur[k] * tmp;
        }
        return ret;
    }
}

/* ------------------------------------------------------------ dense oracle */

void or_dense_matvec_rows(int kernel_id, const or_real *x, const or_real *y, int64_t ny,
                          const or_real *v, const int64_t *rows, int64_t nrows, or_real *out,
                          int nthreads) {
#ifdef _OPENMP
    if (nthreads <= 0) nthreads = omp_get_max_threads();
#pragma omp parallel for schedule(dynamic, 4) num_threads(nthreads)
#endif
    for (int64_t k = 0; k < nrows; ++k) {
        int64_t i = rows ? rows[k] : k;
        long double acc = 0.0L;
        or_real xi = x[i];
        for (int64_t j = 0; j < ny; ++j) acc += (long double)or_kernel(kernel_id, xi, y[j]) * (long double)v[j];
        out[k] = (or_real)acc;
    }
}

/* ------------------------------------------------ EvenBarycentricMatrix */

/* src/BarycentricMatrix.jl:41: (a+b)/2 + (b-a)*x[k]/2 -- Int/Int is a Float64
 * division; (b-a)*x[k] and /2 are in T; the sum promotes to Float64. */
void or_even_nodes(int64_t a, int64_t b, int n, double *t) {
    or_real *x = (or_real *)xmalloc(sizeof(or_real) * (size_t)n);
    or_chebyshevpoints(x, n, 1);
    double mid = (double)(a + b) / 2.0;
    for (int k = 0; k < n; ++k) {
        or_real h = R(b - a) * x[k];
        h = h / R(2);
        t[k] = mid + (double)h;
    }
    free(x);
}

/* 2i-a-b-(b-a)*x[k]: Int arithmetic, then Int - T with (b-a)*x[k] in T */
static or_real even_den(int64_t i, int64_t a, int64_t b, or_real xk) {
    return R(2 * i - a - b) - R(b - a) * xk;
}

/* src/BarycentricMatrix.jl:23-37 */
void or_even_weights(int64_t a, int64_t b, int n, or_real *w, or_real *W) {
    int64_t m = b - a + 1;
    or_real *x = (or_real *)xmalloc(sizeof(or_real) * (size_t)n);
    or_real *lam = (or_real *)xmalloc(sizeof(or_real) * (size_t)n);
    or_chebyshevpoints(x, n, 1);
    or_chebyshevbarycentricweights(lam, n, 1);
    for (int64_t r = 0; r < m; ++r) w[r] = R(0);
    for (int64_t i = a; i <= b; ++i)
        for (int k = 0; k < n; ++k) w[i - a] += lam[k] * (R(1) / even_den(i, a, b, x[k]));
    for (int64_t i = a; i <= b; ++i)
        for (int k = 0; k < n; ++k)
            W[k + (int64_t)n * (i - a)] = lam[k] * (R(1) / (even_den(i, a, b, x[k]) * w[i - a]));
    free(x);
    free(lam);
}

/* src/BarycentricMatrix.jl:62-75: the same w, W stored m x n */
void or_barycentricmatrix_weights(int64_t a, int64_t b, int n, or_real *W) {
    int64_t m = b - a + 1;
    or_real *x = (or_real *)xmalloc(sizeof(or_real) * (size_t)n);
    or_real *lam = (or_real *)xmalloc(sizeof(or_real) * (size_t)n);
    or_real *w = (or_real *)xmalloc(sizeof(or_real) * (size_t)m);
    or_chebyshevpoints(x, n, 1);
    or_chebyshevbarycentricweights(lam, n, 1);
    for (int64_t r = 0; r < m; ++r) w[r] = R(0);
    for (int64_t i = a; i <= b; ++i)
        for (int k = 0; k < n; ++k) w[i - a] += lam[k] * (R(1) / even_den(i, a, b, x[k]));
    for (int k = 0; k < n; ++k)
        for (int64_t i = a; i <= b; ++i)
            W[(i - a) + m * k] = lam[k] * (R(1) / (even_den(i, a, b, x[k]) * w[i - a]));
    free(x);
    free(lam);
    free(w);
}

/* src/BarycentricMatrix.jl:49-58 */
or_real or_even_getindex(const or_real *W, const or_real *F, int64_t m, int64_t p, int n,
                         int64_t i, int64_t j) {
    or_real ret = R(0);
    if (((m + p + i + j) & 1) == 0)
        for (int k = 0; k < n; ++k) ret += F[(j - 1) + p * k] * W[k + (int64_t)n * (i - 1)];
    return ret;
}

/* beta_k = sum over j = j1, j1+2, ... <= p of v[jshift+j] F[j,k]
 * (src/algebra.jl:174-179 and its three twins) */
static void even_beta(or_real *beta, const or_real *F, int64_t p, int n, const or_real *v,
                      int64_t jshift, int64_t j1) {
    for (int k = 0; k < n; ++k) {
        or_real bk = R(0);
        for (int64_t j = j1; j <= p; j += 2) bk += v[jshift + j - 1] * F[(j - 1) + p * k];
        beta[k] = bk;
    }
}

/* u[ishift+i] += sum_k beta_k W[k,i] for i = i1, i1+2, ... <= m
 * (src/algebra.jl:181-187 and its three twins) */
static void even_apply(or_real *u, const or_real *W, int64_t m, int n, const or_real *beta,
                       int64_t ishift, int64_t i1) {
    for (int64_t i = i1; i <= m; i += 2) {
        or_real ui = R(0);
        for (int k = 0; k < n; ++k) ui += beta[k] * W[k + (int64_t)n * (i - 1)];
        u[ishift + i - 1] += ui;
    }
}

/* src/algebra.jl:168-239 */
void or_even_mul(or_real *u, const or_real *W, const or_real *F, int64_t m, int64_t p, int n,
                 const or_real *v, int64_t istart, int64_t jstart) {
    int64_t ishift = istart - 1, jshift = jstart - 1;
    or_real *beta = (or_real *)calloc((size_t)(n > 0 ? n : 1), sizeof(or_real));
    if (!beta) abort();
    if (((ishift + jshift) & 1) == 0) {
        even_beta(beta, F, p, n, v, jshift, 1);
        even_apply(u, W, m, n, beta, ishift, 1);
        even_beta(beta, F, p, n, v, jshift, 2);
        even_apply(u, W, m, n, beta, ishift, 2);
    } else {
        even_beta(beta, F, p, n, v, jshift, 2);
        even_apply(u, W, m, n, beta, ishift, 1);
        even_beta(beta, F, p, n, v, jshift, 1);
        even_apply(u, W, m, n, beta, ishift, 2);
    }
    free(beta);
}
