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
    // the item of workgroup-uniform index t, its fields in scalar registers
    // (the table and the by-value copy live in different address spaces, so
    // the loaded values are not known to be uniform otherwise)
    __device__ DenseItem get(int t) const {
        const DenseItem d = items ? items[t] : one;
        DenseItem u;
        u.n = __builtin_amdgcn_readfirstlane(d.n);
        u.k = __builtin_amdgcn_readfirstlane(d.k);
        u.kk = __builtin_amdgcn_readfirstlane(d.kk);
        u.pad = 0;
        u.A = (double *)uni64((uint64_t)d.A);
        u.lda = (int64_t)uni64((uint64_t)d.lda);
        u.R = (const double *)uni64((uint64_t)d.R);
        u.ldr = (int64_t)uni64((uint64_t)d.ldr);
        u.B = (const double *)uni64((uint64_t)d.B);
        u.ldb = (int64_t)uni64((uint64_t)d.ldb);
        u.D = (const double *)uni64((uint64_t)d.D);
        return u;
    }
    __device__ static uint64_t uni64(uint64_t v) {
        const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)v);
        const uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)(v >> 32));
        return ((uint64_t)hi << 32) | lo;
    }
#endif
};

hipError_t hmx_dense_potrf(DenseItems its, int64_t count, int64_t *info, hipStream_t s);
hipError_t hmx_dense_trsm_t(DenseItems its, int64_t count, int64_t max_n, int64_t max_k,
                            hipStream_t s);
hipError_t hmx_dense_gemm(int ta, int tb, double alpha, double beta, DenseItems its,
                          int64_t count, int64_t max_tiles, hipStream_t s);
