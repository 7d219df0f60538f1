// hmx_even.h -- device layout of a set of EvenBarycentricMatrix leaves.
//
// An EvenBarycentricMatrix (src/BarycentricMatrix.jl:5-46) is an m x p block
// on the integer grids a..b (rows) and c..d (columns) with entries
// sum_k F[j, k] W[k, i], kept only where the global row and column indices
// have the same parity (mul!, src/algebra.jl:168-239). A plan holds any
// number of such leaves placed in one nrows x ncols matrix; a product
// accumulates every leaf:
//
//   up      task (leaf, <= HM_EVEN_COLS columns): beta[par][k] = sum over the
//           task's columns J of parity par of F[j, k] v[J]; one wavefront.
//   reduce  leaves with several tasks: beta = sum of the task partials, in
//           task order.
//   down    segment (rows covered by the same leaves, <= 256 rows): one
//           thread per row I sums W[:, i] . beta[I % 2] over the segment's
//           leaves in leaf order and adds it to u[I]; no atomics.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define HM_EVEN_COLS 4096  // columns per upward task (one wavefront)
#define HM_EVEN_ROWS 256   // rows per downward task (one thread each)
#define HM_EVEN_KB 4       // right-hand sides the workspace holds per launch
#define HM_EVEN_RMAX 64    // largest rank

// W (m x n) and F (p x n) of a leaf are kept in 64-row tiles, k-major inside
// a tile: entry (i, k) at even_tile(i, k, n); a tile holds 64*n values (the
// last one padded). A wavefront's 64 rows of one k are one contiguous run.
#define EVEN_TILE 64
#ifdef __HIPCC__
__host__ __device__
#endif
inline int64_t even_tile(int64_t i, int64_t k, int64_t n) {
    return (i >> 6) * (EVEN_TILE * n) + k * EVEN_TILE + (i & (EVEN_TILE - 1));
}
inline int64_t even_tiled_size(int64_t m, int64_t n) {
    return ((m + EVEN_TILE - 1) / EVEN_TILE) * EVEN_TILE * n;
}

struct ELeaf {
    int64_t row0, col0;  // placement (0-based) in the plan's matrix
    int64_t m, p;        // b - a + 1, d - c + 1
    int64_t woff, foff;  // offsets of W (m x n) and F (p x n), both in 64-row tiles
    int64_t a, b;        // row grid
    int64_t soff;        // offset of F in the column-major staging copy (plan build)
};

struct EUp {
    int32_t leaf;
    int32_t single;  // 1: the leaf has only this task (writes beta directly)
    int64_t j0, len; // columns j0 .. j0 + len - 1 of the leaf
    int64_t slot;    // partial slot (tasks of several-task leaves)
};

struct ERed {
    int32_t leaf, ntask;
    int64_t slot0;  // the leaf's tasks use slots slot0 .. slot0 + ntask - 1
};

struct EDown {
    int64_t r0;         // first row (plan coordinates)
    int32_t len;        // rows (<= HM_EVEN_ROWS)
    int32_t lbeg, lend; // range in the leaf list
    int32_t pad;
};

// Chebyshev nodes and barycentric weights of the first kind, by value
template <class T>
struct ENodes {
    T x[HM_EVEN_RMAX];
    T lam[HM_EVEN_RMAX];
};

template <class T>
struct EvenDev {
    int n;                      // rank
    const ELeaf *leaves;
    const EUp *up;
    int32_t nup;
    const ERed *red;
    int32_t nred;
    const EDown *down;
    int32_t ndown;
    const int32_t *list;        // leaf ids per downward segment
    const T *W, *F;
    T *partial;                 // nslots * HM_EVEN_KB * 2 * n
    T *beta;                    // nleaves * HM_EVEN_KB * 2 * n
    int64_t nrows;
    int covered;                // 1: the segments cover every row 0 .. nrows-1
};

// sets the message of hm_last_error and returns code (defined in plan.cpp)
int hmx_set_error(int code, const char *msg);

// W (and optionally w, at w[row0 + i]) of the leaves [0, nleaves) into W + leaf.woff:
// layout 0 stores W[k + n*i] (EvenBarycentricMatrix), 1 stores W[i + m*k]
// (barycentricmatrix), 2 the 64-row tiles of even_tile (plans).
hipError_t hmx_even_gen_weights(const ELeaf *leaves, int64_t nleaves, int64_t max_m,
                                const ENodes<double> &nodes, int n, int layout, double *W,
                                double *w, hipStream_t s);
hipError_t hmx_even_gen_weights(const ELeaf *leaves, int64_t nleaves, int64_t max_m,
                                const ENodes<float> &nodes, int n, int layout, float *W,
                                float *w, hipStream_t s);
// F of the leaves from the column-major staging copy (F[j + p*k] at soff)
// into the 64-row tiles of the pool (even_tile(j, k, n) at foff)
hipError_t hmx_even_tile_F(const ELeaf *leaves, int64_t nleaves, int64_t max_pn, int n,
                           const double *stage, double *F, hipStream_t s);
hipError_t hmx_even_tile_F(const ELeaf *leaves, int64_t nleaves, int64_t max_pn, int n,
                           const float *stage, float *F, hipStream_t s);
// u[q*ldu + (ioff + I)*incu] (+)= sum over leaves of the leaf products, q < nq
// (launched in groups of up to HM_EVEN_KB right-hand sides)
hipError_t hmx_even_matvec(const EvenDev<double> &p, double *u, const double *v, int64_t ioff,
                           int64_t joff, int64_t incu, int64_t incv, int64_t ldu, int64_t ldv,
                           int nq, int beta, hipStream_t s);
hipError_t hmx_even_matvec(const EvenDev<float> &p, float *u, const float *v, int64_t ioff,
                           int64_t joff, int64_t incu, int64_t incv, int64_t ldu, int64_t ldv,
                           int nq, int beta, hipStream_t s);
