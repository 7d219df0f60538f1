// dense.hip -- gfx950 kernels of the dense leaf steps of the hierarchical
// Cholesky factorization (src/cholesky.jl:12-228): batched Cholesky of
// symmetric positive definite blocks, batched triangular solves with R^T,
// and batched FP64 GEMMs for the Schur-complement updates. One call covers
// many items of free sizes; column-major, leading dimensions per item.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "hmx_dense.h"

namespace {

constexpr int NB = 32;   // panel width of the blocked factorization and solve
constexpr int WGP = 256;

// lane l's value broadcast to the wavefront (v_readlane into a scalar register)
__device__ inline double readlane_d(double v, int l) {
    const int lo = __builtin_amdgcn_readlane(__double2loint(v), l);
    const int hi = __builtin_amdgcn_readlane(__double2hiint(v), l);
    return __hiloint2double(hi, lo);
}

// The NB x NB (or smaller, kb) diagonal block at (k0, k0) of A, factored in
// place by wavefront 0: lane c holds column c of the block's upper triangle
// in registers; step j broadcasts the pivot and row j with readlane. Then
// R_kk goes to Rs (LDS, Rs[i][c]). Returns false when a pivot is not
// positive (*bad = k0 + j + 1).
__device__ inline void diag_factor(double *A, int64_t ld, int k0, int kb, double (*Rs)[NB + 1],
                                   int *bad) {
    const int lane = threadIdx.x;
    if (lane < 64) {
        double d[NB];
#pragma unroll
        for (int i = 0; i < NB; ++i)
            d[i] = (lane < kb && i <= lane) ? A[(k0 + i) + ld * (k0 + lane)] : 0.0;
        int fail = 0;
#pragma unroll
        for (int j = 0; j < NB; ++j) {
            if (j < kb && !fail) {
                const double djj = readlane_d(d[j], j);
                if (!(djj > 0.0)) {
                    fail = j + 1;
                } else {
                    const double r = sqrt(djj);
                    if (lane == j) d[j] = r;
                    else if (lane > j) d[j] /= r;
#pragma unroll
                    for (int i = j + 1; i < NB; ++i) {
                        const double rji = readlane_d(d[j], i);  // R[j][i] from lane i
                        if (lane >= i) d[i] -= rji * d[j];
                    }
                }
            }
        }
        if (lane < kb) {
#pragma unroll
            for (int i = 0; i < NB; ++i) {
                if (i < kb) {
                    const double v = i <= lane ? d[i] : 0.0;
                    Rs[i][lane] = v;
                    A[(k0 + i) + ld * (k0 + lane)] = v;
                }
            }
        }
        if (lane == 0 && fail) *bad = k0 + fail;
    }
}

// Upper Cholesky in place, R^T R = A, from A's upper triangle (Julia's
// cholesky(Symmetric(A)).U); the strictly lower triangle is set to zero.
// Blocked right-looking, panels of NB columns: the diagonal block by one
// wavefront (diag_factor), the panel row by forward substitution with R_kk
// (one thread per column), then the trailing upper triangle takes P^T P in
// 64 x 64 tiles (4 x 4 per thread). The panel row P is staged in LDS once
// when it has at most PL_MAX columns, else tile by tile. One workgroup per
// item; info[t] = 0 or the failing pivot.
constexpr int PL_MAX = 448;
__device__ inline void syrk_tile(double *A, int64_t ld, int n, int i0, int j0,
                                 const double (*Pl)[PL_MAX + 1], int pi, int pj) {
    const int tid = threadIdx.x, tx = tid % 16, ty = tid / 16;
    double acc[4][4];
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b) acc[a][b] = 0.0;
#pragma unroll 8
    for (int t = 0; t < NB; ++t) {
        double a4[4], b4[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            a4[q] = Pl[t][pi + tx + 16 * q];
            b4[q] = Pl[t][pj + ty + 16 * q];
        }
#pragma unroll
        for (int a = 0; a < 4; ++a)
#pragma unroll
            for (int b = 0; b < 4; ++b) acc[a][b] += a4[a] * b4[b];
    }
#pragma unroll
    for (int a = 0; a < 4; ++a) {
        const int gi = i0 + tx + 16 * a;
#pragma unroll
        for (int b = 0; b < 4; ++b) {
            const int gj = j0 + ty + 16 * b;
            if (gi < n && gj < n && gi <= gj) A[gi + ld * gj] -= acc[a][b];
        }
    }
}

__global__ __launch_bounds__(WGP) void k_potrf(DenseItems its, int64_t *info) {
    const DenseItem it = its.get(blockIdx.x);
    double *A = it.A;
    const int n = it.n;
    const int64_t ld = it.lda;
    const int tid = threadIdx.x;
    __shared__ double Rs[NB][NB + 1];
    __shared__ double Pl[NB][PL_MAX + 1];
    __shared__ int bad;
    if (tid == 0) bad = 0;
    __syncthreads();
    for (int k0 = 0; k0 < n; k0 += NB) {
        const int kb = min(NB, n - k0);
        diag_factor(A, ld, k0, kb, Rs, &bad);
        __syncthreads();
        if (bad) break;
        const int s0 = k0 + kb, m = n - s0;
        const bool staged = m <= PL_MAX;
        // panel row: A[k0:k0+kb, c] := R_kk^{-T} A[k0:k0+kb, c] for c >= s0
        for (int c = s0 + tid; c < n; c += WGP) {
            double p[NB];
#pragma unroll
            for (int i = 0; i < NB; ++i) p[i] = i < kb ? A[(k0 + i) + ld * c] : 0.0;
#pragma unroll
            for (int i = 0; i < NB; ++i) {
                if (i < kb) {
                    double v = p[i];
#pragma unroll
                    for (int t = 0; t < i; ++t) v -= Rs[t][i] * p[t];
                    p[i] = v / Rs[i][i];
                }
            }
#pragma unroll
            for (int i = 0; i < NB; ++i) {
                if (i < kb) A[(k0 + i) + ld * c] = p[i];
                if (staged) Pl[i][c - s0] = i < kb ? p[i] : 0.0;
            }
        }
        __syncthreads();
        const int nt = (m + 63) / 64;
        if (staged) {
            // zero the staged columns past m (read by the last tile)
            for (int e = tid; e < NB * 64; e += WGP) {
                const int t = e % NB, r = m + e / NB;
                if (r < nt * 64 && r < PL_MAX + 1) Pl[t][r] = 0.0;
            }
            __syncthreads();
            for (int tj = 0; tj < nt; ++tj)
                for (int ti = 0; ti <= tj; ++ti)
                    syrk_tile(A, ld, n, s0 + 64 * ti, s0 + 64 * tj, Pl, 64 * ti, 64 * tj);
        } else {
            for (int tj = 0; tj < nt; ++tj) {
                for (int ti = 0; ti <= tj; ++ti) {
                    const int i0 = s0 + 64 * ti, j0 = s0 + 64 * tj;
                    for (int e = tid; e < NB * 64; e += WGP) {
                        const int t = e % NB, r = e / NB;
                        Pl[t][r] = (t < kb && i0 + r < n) ? A[(k0 + t) + ld * (i0 + r)] : 0.0;
                        Pl[t][64 + r] = (t < kb && j0 + r < n) ? A[(k0 + t) + ld * (j0 + r)] : 0.0;
                    }
                    __syncthreads();
                    syrk_tile(A, ld, n, i0, j0, Pl, 0, 64);
                    __syncthreads();
                }
            }
        }
        __syncthreads();
    }
    __syncthreads();
    for (int k = 0; k < n; ++k)
        for (int i = k + 1 + tid; i < n; i += WGP) A[i + ld * k] = 0.0;
    if (tid == 0) info[blockIdx.x] = bad;
}

// B := R^{-T} B (R upper n x n, B n x k): blocked forward substitution with
// R^T. Workgroup (item, block of NC columns): the NB x NC block of rows
// [k0, k0+kb) is solved with R_kk (LDS) by NC threads, then every row below
// takes B[r, :] -= R[k0:k0+kb, r]^T P, one thread per row.
constexpr int WGT = 256, NC = 16;
__global__ __launch_bounds__(WGT) void k_trsm_t(DenseItems its) {
    const DenseItem it = its.get(blockIdx.x);
    const int n = it.n, kc = it.k;
    const int c0 = blockIdx.y * NC;
    if (c0 >= kc) return;
    const int nc = min(NC, kc - c0);
    const double *R = it.R;
    const int64_t ldr = it.ldr, ldb = it.lda;
    double *B = it.A + ldb * c0;
    const int tid = threadIdx.x;
    __shared__ double Rs[NB][NB + 1];
    __shared__ double Ps[NB][NC];
    for (int k0 = 0; k0 < n; k0 += NB) {
        const int kb = min(NB, n - k0);
        __syncthreads();
        for (int e = tid; e < NB * NB; e += WGT) {
            const int a = e % NB, b = e / NB;
            Rs[a][b] = (a < kb && b < kb) ? R[(k0 + a) + ldr * (k0 + b)] : 0.0;
        }
        __syncthreads();
        if (tid < nc) {
            double p[NB];
#pragma unroll
            for (int i = 0; i < NB; ++i) p[i] = i < kb ? B[(k0 + i) + ldb * tid] : 0.0;
#pragma unroll
            for (int i = 0; i < NB; ++i) {
                if (i < kb) {
                    double v = p[i];
#pragma unroll
                    for (int t = 0; t < i; ++t) v -= Rs[t][i] * p[t];
                    p[i] = v / Rs[i][i];
                }
            }
#pragma unroll
            for (int i = 0; i < NB; ++i) {
                Ps[i][tid] = p[i];
                if (i < kb) B[(k0 + i) + ldb * tid] = p[i];
            }
        }
        __syncthreads();
        for (int r = k0 + kb + tid; r < n; r += WGT) {
            double rr[NB];
#pragma unroll
            for (int t = 0; t < NB; ++t) rr[t] = t < kb ? R[(k0 + t) + ldr * r] : 0.0;
            for (int c = 0; c < nc; ++c) {
                double v = 0.0;
#pragma unroll
                for (int t = 0; t < NB; ++t) v += rr[t] * Ps[t][c];
                B[r + ldb * c] -= v;
            }
        }
    }
}

// C = beta C + alpha op(A) diag(D) op(B), op = identity or transpose, D
// optional; 64 x 64 tiles of C per workgroup, 4 x 4 per thread, K in steps of
// 16 through LDS.
constexpr int TM = 64, TK = 16;
__global__ __launch_bounds__(256) void k_gemm(int ta, int tb, double alpha, double beta,
                                              DenseItems its) {
    const DenseItem it = its.get(blockIdx.x);
    const int M = it.n, N = it.k, K = it.kk;
    const int tiles_m = (M + TM - 1) / TM, tiles_n = (N + TM - 1) / TM;
    const int tile = blockIdx.y;
    if (tile >= tiles_m * tiles_n) return;
    const int i0 = (tile % tiles_m) * TM, j0 = (tile / tiles_m) * TM;
    const double *A = it.R, *B = it.B;
    double *C = it.A;
    const int64_t lda = it.ldr, ldb = it.ldb, ldc = it.lda;
    __shared__ double As[TK][TM + 1], Bs[TK][TM + 1];
    const int tid = threadIdx.x, tx = tid % 16, ty = tid / 16;
    double acc[4][4];
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b) acc[a][b] = 0.0;
    for (int k0 = 0; k0 < K; k0 += TK) {
        for (int e = tid; e < TK * TM; e += 256) {
            const int kk = e % TK, r = e / TK;
            const int gi = i0 + r, gk = k0 + kk;
            double av = 0.0;
            if (gi < M && gk < K) {
                av = ta ? A[gk + lda * gi] : A[gi + lda * gk];
                if (it.D) av *= it.D[gk];
            }
            As[kk][r] = av;
            const int gj = j0 + r;
            double bv = 0.0;
            if (gj < N && gk < K) bv = tb ? B[gj + ldb * gk] : B[gk + ldb * gj];
            Bs[kk][r] = bv;
        }
        __syncthreads();
#pragma unroll
        for (int kk = 0; kk < TK; ++kk) {
            double a[4], b[4];
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                a[q] = As[kk][tx + 16 * q];
                b[q] = Bs[kk][ty + 16 * q];
            }
#pragma unroll
            for (int p = 0; p < 4; ++p)
#pragma unroll
                for (int q = 0; q < 4; ++q) acc[p][q] += a[p] * b[q];
        }
        __syncthreads();
    }
#pragma unroll
    for (int p = 0; p < 4; ++p) {
        const int gi = i0 + tx + 16 * p;
        if (gi >= M) continue;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int gj = j0 + ty + 16 * q;
            if (gj >= N) continue;
            double *c = C + gi + ldc * gj;
            *c = (beta == 0.0 ? 0.0 : beta * *c) + alpha * acc[p][q];
        }
    }
}

}  // namespace

hipError_t hmx_dense_potrf(DenseItems its, int64_t count, int64_t *info, hipStream_t s) {
    if (count > 0) k_potrf<<<(unsigned)count, WGP, 0, s>>>(its, info);
    return hipGetLastError();
}

hipError_t hmx_dense_trsm_t(DenseItems its, int64_t count, int64_t max_k, hipStream_t s) {
    if (count > 0 && max_k > 0)
        k_trsm_t<<<dim3((unsigned)count, (unsigned)((max_k + NC - 1) / NC)), WGT, 0, s>>>(its);
    return hipGetLastError();
}

hipError_t hmx_dense_gemm(int ta, int tb, double alpha, double beta, DenseItems its,
                          int64_t count, int64_t max_tiles, hipStream_t s) {
    if (count > 0 && max_tiles > 0)
        k_gemm<<<dim3((unsigned)count, (unsigned)max_tiles), 256, 0, s>>>(ta, tb, alpha, beta,
                                                                          its);
    return hipGetLastError();
}
