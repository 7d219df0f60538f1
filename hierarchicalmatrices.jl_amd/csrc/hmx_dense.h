// hmx_dense.h -- batched dense leaf steps of the hierarchical Cholesky
// factorization (src/cholesky.jl:12-228) on the device.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

// One item of a batched call. potrf: A (n x n, lda). trsm: R (n x n, ldr),
// B in A (n x k, lda). gemm: C in A (n x k, lda) = beta C + alpha op(R)
// diag(D) op(B) with inner dimension kk (D == nullptr: the identity).
struct DenseItem {
    int32_t n, k, kk, pad;
    double *A;
    int64_t lda;
    const double *R;
    int64_t ldr;
    const double *B;
    int64_t ldb;
    const double *D;
};

// the items of a launch: a device table, or (items == nullptr) one item
// passed by value
struct DenseItems {
    const DenseItem *items;
    DenseItem one;
#ifdef __HIPCC__
    __device__ DenseItem get(int t) const { return items ? items[t] : one; }
#endif
};

hipError_t hmx_dense_potrf(DenseItems its, int64_t count, int64_t *info, hipStream_t s);
hipError_t hmx_dense_trsm_t(DenseItems its, int64_t count, int64_t max_k, hipStream_t s);
hipError_t hmx_dense_gemm(int ta, int tb, double alpha, double beta, DenseItems its,
                          int64_t count, int64_t max_tiles, hipStream_t s);
