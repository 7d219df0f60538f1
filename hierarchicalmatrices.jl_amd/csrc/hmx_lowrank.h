// hmx_lowrank.h -- batched LowRankMatrix arithmetic with recompression
// (src/LowRankMatrix.jl:64-171) on the device.
//
// Every operation of the reference reduces to: thin QR of tall factors
// (Householder, one workgroup per factor), an SVD of a small core (one-sided
// Jacobi in LDS, one workgroup per core), getrank, and products of tall
// factors with the small singular-vector matrices. Items of one call run in
// the same launches; sizes may differ per item.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define LR_MAX 64  // largest core dimension (r1 + r2 for sums)
// per-item small slot (doubles): RU, RV, SU, SV, and the compact-WY factors
// TU, TV (Q = I - V T V^T) (LR_MAX x LR_MAX each, ld LR_MAX), tau U, tau V,
// S (LR_MAX each), rank
#define LR_RU 0
#define LR_RV (LR_MAX * LR_MAX)
#define LR_SU (2 * LR_MAX * LR_MAX)
#define LR_SV (3 * LR_MAX * LR_MAX)
#define LR_WU (4 * LR_MAX * LR_MAX)
#define LR_WV (5 * LR_MAX * LR_MAX)
#define LR_TU (6 * LR_MAX * LR_MAX)
#define LR_TV (LR_TU + LR_MAX)
#define LR_S (LR_TV + LR_MAX)
#define LR_RANK (LR_S + LR_MAX)
#define LR_SMALL (LR_RANK + 8)

struct LRItem {
    int32_t m, n, k, r1, r2;
    int32_t cu, cv;   // columns of the tall factors XU (m x cu) and XV (n x cv) to factor
    int32_t p, q;     // the core is p x q
    int32_t tr;       // SVDTRUNC with m < n: factor A^T
    const double *U1, *S1, *V1, *U2, *S2, *V2, *A;
    double *Uo, *So, *Vo;
    int64_t xu, xv;   // offsets of XU, XV in the tall workspace
    int64_t small;    // offset of the small slot
};

// LDS (doubles) available to the in-LDS factorization: M*C + 2*C*(C+1) must
// fit (160 KB per workgroup on gfx950, less the static arrays)
#define LR_LDS_DOUBLES 20352

// op codes: the enum hm_lr_op of include/hmx.h. rank (device, count values)
// receives each result's rank; max_rows = the largest m or n of the items;
// lds_qr = the largest in-LDS factorization (doubles, 0: none), big_qr = some
// factor does not fit; lds_svd = the largest Jacobi workspace (doubles);
// max_c = the largest factor width (cu, cv) of the items.
hipError_t hmx_lowrank_run(int op, const LRItem *items, int64_t count, double *tall,
                           double *small, int64_t *rank, int64_t max_rows, int64_t lds_qr,
                           bool big_qr, int64_t lds_svd, int64_t max_c, hipStream_t s);
