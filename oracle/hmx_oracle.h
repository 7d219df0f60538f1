/*
 * hmx_oracle.h -- CPU restatement of HierarchicalMatrices.jl's KernelMatrix path.
 *
 * TEST INFRASTRUCTURE ONLY. This library is the checker: only tests/, the smoke
 * check and the cpu_baseline leg of bench.py may load it. The product library
 * (libhmx) never links it and never falls back to it.
 *
 * Conventions: all index ranges are Julia-style, 1-based `first` plus a length,
 * so that the restated code reads like the reference. Matrices are column-major.
 */
#ifndef HMX_ORACLE_H
#define HMX_ORACLE_H

#include <stdint.h>

/* The same source builds two libraries: liboracle.so with or_real = double
 * (Float64) and liboracle_f32.so with or_real = float (Float32, every
 * operation rounded to single precision as Julia does for T = Float32).
 * liboracle_f32.so is built with -DOR_F32. */
#ifdef OR_F32
typedef float or_real;
#define OR_LOGF logf
#define OR_FABS fabsf
#else
typedef double or_real;
#define OR_LOGF log
#define OR_FABS fabs
#endif
#define R(v) ((or_real)(v))

#ifdef __cplusplus
extern "C" {
#endif

enum { OR_CAUCHY = 0, OR_COULOMB = 1, OR_COULOMB3 = 2, OR_LOG = 3 };
enum { OR_CODE_NONE = 0, OR_CODE_KM = 1, OR_CODE_BARY = 2, OR_CODE_DENSE = 3 };
enum { OR_NODE_D = 0, OR_NODE_K1 = 1, OR_NODE_K2 = 2 };

/* scalar kernels of examples/Kernel.jl:34-37 */
or_real or_kernel(int kernel_id, or_real x, or_real y);

/* BLOCKRANK for an IEEE type with `mant_bits` explicit mantissa bits
 * (src/HierarchicalMatrices.jl:5): 2*round(0.5*log_{3+sqrt 8}(1/eps)). */
int or_blockrank(int mant_bits);

/* first/second-kind Chebyshev points, src/BarycentricMatrix.jl:92-111 */
void or_chebyshevpoints(or_real *out, int n, int kind);
/* barycentric weights, src/BarycentricMatrix.jl:114-136 */
void or_chebyshevbarycentricweights(or_real *out, int n, int kind);
/* sinpi used by the two functions above */
or_real or_sinpi(or_real q);

/* indsplit, src/BarycentricMatrix.jl:299-307. x is 1-based in the sense that
 * index i reads x[i-1]. Writes (first1,len1,first2,len2). Returns 0, or -1 when
 * the reference would raise BoundsError. */
int or_indsplit(const or_real *x, int64_t nx, int64_t first, int64_t len,
                or_real a, or_real b, int64_t out[4]);

/* KernelMatrix tree, src/KernelMatrix.jl:47-116 */
typedef struct or_tree or_tree;

/* Build the tree. store_factors=1 also generates every dense block, F, U, V
 * (ref layout); 0 keeps only the skeleton (ranges, intervals) for the
 * row-sampled matvec. Returns NULL on failure; *quirk is set to 1 if an empty
 * range went through indsplit (the reference then returns a length-1 range). */
or_tree *or_tree_build(int kernel_id, const or_real *x, const or_real *y, int64_t nx,
                       int64_t ny, or_real a, or_real b, or_real c, or_real d,
                       int blocksize, int rank, int store_factors, int *quirk);
void or_tree_free(or_tree *t);

/* Counts: [0]=D nodes [1]=K1 [2]=K2 [3]=dense blocks [4]=far blocks
 * [5]=dense entries [6]=sum far rows [7]=sum far cols [8]=max depth */
void or_tree_stats(const or_tree *t, int64_t out[9]);

/* Leaf blocks in traversal order (src/KernelMatrix.jl:17-45). Each record:
 * code, depth, row0 (0-based), col0, m, n, then 4 doubles a,b,c,d (far only).
 * Returns the number of blocks; writes at most cap records. */
int64_t or_tree_blocks(const or_tree *t, int64_t *ints /* 6 per block */,
                       or_real *dbls /* 4 per block */, int64_t cap);

/* Faithful traversal mul! (src/KernelMatrix.jl:14-45): u += H*v with the
 * reference's loop order. Needs store_factors=1. */
int or_tree_matvec(or_tree *t, or_real *u, const or_real *v);

/* Same values, OpenMP over row chunks: phase 1 forms F*(V'v) per far block,
 * phase 2 accumulates each row in traversal order, so every u[i] is
 * bit-identical to or_tree_matvec. nthreads<=0 uses the OpenMP default. */
int or_tree_matvec_omp(or_tree *t, or_real *u, const or_real *v, int nthreads);

/* Row-sampled faithful mul!: out[k] = (H*v)[rows[k]] (0-based rows), with the
 * reference's per-row summation order; far-field factors are regenerated on
 * the fly (bit-identical to the stored ones), so it works at any n. */
int or_tree_matvec_rows(or_tree *t, const or_real *v, const int64_t *rows,
                        int64_t nrows, or_real *out, int nthreads);

/* getindex(H,i,j), src/hierarchical.jl:120-147 (0-based i,j here) */
or_real or_tree_getindex(or_tree *t, int64_t i, int64_t j);

/* Copy generated factors of the b-th block (traversal order) out: dense A
 * (m*n), or far U (m*r), V (n*r), F (r*r). Needs store_factors=1. */
int or_tree_block_data(const or_tree *t, int64_t bidx, or_real *A, or_real *U, or_real *V,
                       or_real *F);

/* Leaf kernels, src/algebra.jl. istart/jstart are 1-based as in Julia. */
void or_dense_mul(or_real *y, const or_real *A, int64_t m, int64_t n, const or_real *x,
                  int64_t istart, int64_t jstart, int64_t incx, int64_t incy);
void or_dense_tmul(or_real *y, const or_real *A, int64_t m, int64_t n, const or_real *x,
                   int64_t istart, int64_t jstart, int64_t incx, int64_t incy);
void or_lowrank_mul(or_real *y, const or_real *U, const or_real *S, const or_real *V,
                    int64_t m, int64_t n, int64_t r, const or_real *x, int64_t istart,
                    int64_t jstart, int64_t incx, int64_t incy);
void or_lowrank_adjmul(or_real *y, const or_real *U, const or_real *S, const or_real *V,
                       int64_t m, int64_t n, int64_t r, const or_real *x, int64_t istart,
                       int64_t jstart, int64_t incx, int64_t incy);
void or_bary_mul(or_real *u, const or_real *U, const or_real *F, const or_real *V, int64_t m,
                 int64_t n, int64_t r, const or_real *v, int64_t istart, int64_t jstart);

/* evaluate(B::BarycentricPoly2D, x, y), src/BarycentricMatrix.jl:180-205: the
 * 2-D barycentric interpolant of F (r x r column-major, F[m + r*n] =
 * f(tx_m, ty_n)) on the mapped nodes tx, ty with weights lx, ly. */
or_real or_bary_evaluate(int r, const or_real *tx, const or_real *ty, const or_real *lx,
                         const or_real *ly, const or_real *F, or_real x, or_real y);

/* Dense oracle: out[k] = sum_j f(x[rows[k]], y[j]) * v[j] accumulated in long
 * or_real (rows==NULL means all rows 0..nrows-1). OpenMP over rows. */
void or_dense_matvec_rows(int kernel_id, const or_real *x, const or_real *y, int64_t ny,
                          const or_real *v, const int64_t *rows, int64_t nrows,
                          or_real *out, int nthreads);

/* ------------------------------------------------ EvenBarycentricMatrix
 * src/BarycentricMatrix.jl:5-89 and its mul!, src/algebra.jl:168-239. The
 * rows are the integer grid a..b (m = b-a+1), the columns c..d (p = d-c+1),
 * n Chebyshev nodes. f is the caller's function: F (p x n, column-major,
 * F[j,k] = f(T, t_k, j)) is an input here. */

/* Nodes where the reference evaluates f: t_k = (a+b)/2 + (b-a)*x_k/2
 * (src/BarycentricMatrix.jl:41). (a+b)/2 is an Int/Int division, which Julia
 * carries out in Float64 for every T, so t is Float64 in both builds. */
void or_even_nodes(int64_t a, int64_t b, int n, double *t);

/* w (m) and W (n x m column-major, W[k + n*(i-a)]) of
 * EvenBarycentricMatrix(T, f, a, b, c, d), src/BarycentricMatrix.jl:23-37:
 * w_i = sum_k lambda_k * inv(2i-a-b-(b-a)x_k), W[k,i] = lambda_k *
 * inv((2i-a-b-(b-a)x_k) * w_i). */
void or_even_weights(int64_t a, int64_t b, int n, or_real *w, or_real *W);

/* W of barycentricmatrix(T, f, a, b, c, d), src/BarycentricMatrix.jl:62-75:
 * the same values, m x n column-major (W[(i-a) + m*k]). */
void or_barycentricmatrix_weights(int64_t a, int64_t b, int n, or_real *W);

/* getindex(B, i, j), src/BarycentricMatrix.jl:49-58 (1-based i, j): nonzero
 * only when size(B,1)+size(B,2)+i+j is even. */
or_real or_even_getindex(const or_real *W, const or_real *F, int64_t m, int64_t p, int n,
                         int64_t i, int64_t j);

/* mul!(u, B, v, istart, jstart), src/algebra.jl:168-239 (1-based istart,
 * jstart): the row and column parity classes chosen by iseven(ishift+jshift),
 * loop for loop. */
void or_even_mul(or_real *u, const or_real *W, const or_real *F, int64_t m, int64_t p, int n,
                 const or_real *v, int64_t istart, int64_t jstart);

#ifdef __cplusplus
}
#endif
#endif
