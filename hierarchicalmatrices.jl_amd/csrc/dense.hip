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

// lane l's value broadcast to the wavefront (v_readlane into a scalar register)
__device__ inline double readlane_d(double v, int l) {
    const int lo = __builtin_amdgcn_readlane(__double2loint(v), l);
    const int hi = __builtin_amdgcn_readlane(__double2hiint(v), l);
    return __hiloint2double(hi, lo);
}

// 1/sqrt(x) for x > 0: the hardware estimate refined by two Newton steps
// y <- y + y (1 - x y^2) / 2 (each step squares the relative error, so the
// result is within an ulp or two of the rounded value). Three dependent
// operations per step instead of the longer correctly rounded square root
// and division, which sit on the factorization's serial chain.
__device__ inline double rsqrt_nr(double x) {
    double y = __builtin_amdgcn_rsq(x);
#pragma unroll
    for (int it = 0; it < 2; ++it) {
        const double e = fma(-x * y, y, 1.0);
        y = fma(0.5 * y, e, y);
    }
    return y;
}

// The NB x NB (or smaller, kb) diagonal block at (k0, k0) of A, factored in
// place by wavefront 0: lane c holds column c of the block's upper triangle
// in registers. Step j takes the pivot from lane j (readlane), scales row j
// (the j-th register of every lane) by 1/r_jj (rsqrt_nr), stores it as row j
// of Rs (LDS) and updates each column with row j read back from Rs (one batch
// of reads, the same addresses for all lanes). Rs ends holding R_kk
// (Rs[i][c], c < NB) and 1/r_jj in its last column (Rs[j][NB]); the
// workgroup writes R_kk back to A. Called by every thread; *bad = k0 + j + 1
// when pivot j is not positive.
__device__ __attribute__((noinline)) void diag_factor(double *A, int64_t ld, int k0, int kb, double (*Rs)[NB + 1],
                                   int *bad, int nthreads) {
    const int tid = threadIdx.x;
    if (tid < 64) {
        // (loads from clamped addresses, then selects: no divergent branches)
        const double *col = A + k0 + ld * (k0 + (tid < kb ? tid : 0));
        double d[NB];
#pragma unroll
        for (int i = 0; i < NB; ++i) {
            const double v = col[i < kb ? i : 0];
            d[i] = (tid < kb && i <= tid) ? v : 0.0;
        }
        int fail = 0;
#pragma unroll
        for (int j = 0; j < NB; ++j) {
            if (j < kb && !fail) {
                const double djj = readlane_d(d[j], j);
                if (!(djj > 0.0)) {
                    fail = j + 1;
                } else {
                    const double y = rsqrt_nr(djj);  // 1 / r_jj
                    const double r = djj * y, q = d[j] * y;
                    d[j] = tid == j ? r : (tid > j ? q : d[j]);
                    if (tid < NB) Rs[j][tid] = d[j];
                    if (tid == 0) Rs[j][NB] = y;
                    __builtin_amdgcn_wave_barrier();
                    // row j back from Rs, all reads issued before the
                    // updates; the entries below the diagonal (lane c < i)
                    // are updated too and never read
                    double rj[NB];
#pragma unroll
                    for (int i = j + 1; i < NB; ++i) rj[i] = Rs[j][i];
#pragma unroll
                    for (int i = j + 1; i < NB; ++i) d[i] -= rj[i] * d[j];
                    __builtin_amdgcn_wave_barrier();
                    __builtin_amdgcn_sched_barrier(0);
                }
            }
        }
        if (tid == 0 && fail) *bad = k0 + fail;
    }
    __syncthreads();
    for (int e = tid; e < kb * kb; e += nthreads) {
        const int i = e % kb, c = e / kb;
        A[(k0 + i) + ld * (k0 + c)] = i <= c ? Rs[i][c] : 0.0;
    }
}

// Upper Cholesky in place, R^T R = A, from A's upper triangle (Julia's
// cholesky(Symmetric(A)).U); the strictly lower triangle is set to zero.
// Blocked right-looking, panels of NB columns: the diagonal block by
// wavefront 0 (diag_factor), the panel row by forward substitution with R_kk
// (panel_solve, one thread per column), then the trailing upper triangle
// takes P^T P in 64 x 64 tiles (4 x 4 per thread), two tiles at a time (one
// per group of 256 threads). The panel row P stays in LDS for the update
// when it has at most PL_LD columns; otherwise each group stages its tile's
// columns in its own part of the same buffer. One workgroup per item;
// info[t] = 0 or the failing pivot.
// Pl: NB rows of up to PL_LD columns, row stride PL_LD + 1
constexpr int PL_LD = 512, PL_ST = PL_LD + 1, WGF = 512, NGRP = WGF / 256;
constexpr int PG_LD = 128;  // per-group staging: NB x 128 (two 64-column sides)
static_assert(NGRP * PG_LD <= PL_ST, "group staging exceeds the panel buffer");
// A[i0.., j0..] -= Pi^T Pj for the group's 64 x 64 tile; Pl columns pi.., pj..
// (the tile's entries are loaded before the products, so their latency
// overlaps the arithmetic)
template <int LDP>
__device__ inline void syrk_tile(double *A, int64_t ld, int n, int i0, int j0,
                                 const double (*Pl)[LDP], int pi, int pj, int gt) {
    const int tx = gt % 16, ty = gt / 16;
    double acc[4][4];
#pragma unroll
    for (int a = 0; a < 4; ++a) {
        const int gi = i0 + tx + 16 * a;
#pragma unroll
        for (int b = 0; b < 4; ++b) {
            const int gj = j0 + ty + 16 * b;
            acc[a][b] = (gi < n && gj < n && gi <= gj) ? A[gi + ld * gj] : 0.0;
        }
    }
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
            for (int b = 0; b < 4; ++b) acc[a][b] -= a4[a] * b4[b];
    }
#pragma unroll
    for (int a = 0; a < 4; ++a) {
        const int gi = i0 + tx + 16 * a;
#pragma unroll
        for (int b = 0; b < 4; ++b) {
            const int gj = j0 + ty + 16 * b;
            if (gi < n && gj < n && gi <= gj) A[gi + ld * gj] = acc[a][b];
        }
    }
}

// panel row A[k0:k0+kb, c] := R_kk^{-T} A[k0:k0+kb, c] for c >= s0: one
// thread per column, the column in registers (one batch of loads), each row
// of R_kk read from Rs in one batch, the pivots applied as the reciprocals
// diag_factor left in Rs[t][NB]; the result also goes to Pl for the
// trailing update when the panel fits there. Kept out of line so that its
// registers are allocated apart from the other steps of the factorization.
__device__ __attribute__((noinline)) void panel_solve(double *A, int64_t ld, int n, int k0, int kb,
                                                      int s0, bool staged,
                                                      const double (*Rs)[NB + 1],
                                                      double (*Pl)[PL_ST]) {
    const int tid = threadIdx.x;
    for (int c = s0 + tid; c < n; c += WGF) {
        double *pc = A + k0 + ld * c;
        double p[NB];
#pragma unroll
        for (int i = 0; i < NB; ++i) {
            const double v = pc[i < kb ? i : 0];
            p[i] = i < kb ? v : 0.0;
        }
#pragma unroll
        for (int t = 0; t < NB; ++t) {
            if (t < kb) {
                double rt[NB + 1];
#pragma unroll
                for (int i = t + 1; i <= NB; ++i) rt[i] = Rs[t][i];
                p[t] = p[t] * rt[NB];  // 1 / r_tt
#pragma unroll
                for (int i = t + 1; i < NB; ++i) p[i] -= rt[i] * p[t];
            }
            __builtin_amdgcn_sched_barrier(0);  // one step's reads at a time
        }
#pragma unroll
        for (int i = 0; i < NB; ++i) {
            if (i < kb) pc[i] = p[i];
            if (staged) Pl[i][c - s0] = i < kb ? p[i] : 0.0;
        }
    }
}

__global__ __launch_bounds__(WGF) void k_potrf(DenseItems its, int64_t *info) {
    const DenseItem it = its.get(blockIdx.x);
    double *A = it.A;
    const int n = it.n;
    const int64_t ld = it.lda;
    const int tid = threadIdx.x;
    const int grp = tid / 256, gt = tid % 256;
    __shared__ double Rs[NB][NB + 1];
    __shared__ double Pbuf[NB * PL_ST];
    __shared__ int bad;
    double (*Pl)[PL_ST] = reinterpret_cast<double (*)[PL_ST]>(Pbuf);
    // group grp's staging, when P is not staged whole: NB x PG_LD from
    // Pbuf + grp * NB * PG_LD
    double (*Pg)[PG_LD] = reinterpret_cast<double (*)[PG_LD]>(Pbuf + grp * NB * PG_LD);
    if (tid == 0) bad = 0;
    __syncthreads();
    for (int k0 = 0; k0 < n; k0 += NB) {
        const int kb = min(NB, n - k0);
        diag_factor(A, ld, k0, kb, Rs, &bad, WGF);
        __syncthreads();
        if (bad) break;
        const int s0 = k0 + kb, m = n - s0;
        const bool staged = m <= PL_LD;
        panel_solve(A, ld, n, k0, kb, s0, staged, Rs, Pl);
        const int nt = (m + 63) / 64, ntiles = nt * (nt + 1) / 2;
        if (staged) {
            // zero the staged columns past m (read by the last tile)
            for (int e = tid; e < NB * 64; e += WGF) {
                const int t = e % NB, r = m + e / NB;
                if (r < nt * 64 && r < PL_LD) Pl[t][r] = 0.0;  // r >= m: not read by the write-back
            }
        }
        __syncthreads();
        // upper tiles (ti <= tj) in order, tile q to group q % NGRP
        for (int q0 = 0; q0 < ntiles; q0 += NGRP) {
            const int q = q0 + grp;
            int tj = 0;
            while ((tj + 1) * (tj + 2) / 2 <= q) ++tj;
            const int ti = q - tj * (tj + 1) / 2;
            const int i0 = s0 + 64 * ti, j0 = s0 + 64 * tj;
            if (staged) {
                if (q < ntiles) syrk_tile<PL_ST>(A, ld, n, i0, j0, Pl, 64 * ti, 64 * tj, gt);
            } else {
                if (q < ntiles) {
                    for (int e = gt; e < NB * 64; e += 256) {
                        const int t = e % NB, r = e / NB;
                        Pg[t][r] = (t < kb && i0 + r < n) ? A[(k0 + t) + ld * (i0 + r)] : 0.0;
                        Pg[t][64 + r] = (t < kb && j0 + r < n) ? A[(k0 + t) + ld * (j0 + r)] : 0.0;
                    }
                }
                __syncthreads();
                if (q < ntiles) syrk_tile<PG_LD>(A, ld, n, i0, j0, Pg, 0, 64, gt);
                __syncthreads();
            }
        }
        __syncthreads();
    }
    __syncthreads();
    for (int64_t e = tid; e < (int64_t)n * n; e += WGF) {
        const int i = (int)(e % n), k = (int)(e / n);
        if (i > k) A[i + ld * k] = 0.0;
    }
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

// B := R^{-T} B for n <= TS_MAX, one workgroup per (item, block of NC
// columns). The column block of B stays in LDS for the whole solve. The
// inverses of the NB x NB diagonal blocks of R are formed first (two blocks
// per wavefront, one column of the inverse per lane, by back substitution),
// so that each panel step is a product with inv(R_kk)^T
// rather than a chain of substitutions. The rows below then take
// B[r, :] -= R[k0:k0+kb, r]^T X_k, one thread per row, with R's column slab
// of the next panel loaded while the current one is applied.
constexpr int TS_MAX = 256, TS_NB = TS_MAX / NB;
constexpr int TS_LDS = TS_NB * NB * (NB + 1) + (TS_MAX + NB) * (NC + 1);  // doubles
static_assert(WGT == TS_MAX && TS_NB == 2 * (WGT / 64), "one row per thread, two blocks per wave");

__device__ inline void trsm_slab(const double *R, int64_t ldr, int n, int k, int r,
                                 double (&dst)[NB]) {
    const int k0 = k * NB, kb = min(NB, n - k0);
    const bool live = r >= k0 + kb && r < n;
#pragma unroll
    for (int t = 0; t < NB; ++t) dst[t] = (live && t < kb) ? R[(k0 + t) + ldr * r] : 0.0;
}

__global__ __launch_bounds__(WGT) void k_trsm_t_lds(DenseItems its) {
    extern __shared__ double lds[];
    double (*Ri)[NB][NB + 1] = reinterpret_cast<double (*)[NB][NB + 1]>(lds);
    double (*Bs)[NC + 1] = reinterpret_cast<double (*)[NC + 1]>(lds + TS_NB * NB * (NB + 1));
    double (*Xs)[NC + 1] = Bs + TS_MAX;
    const DenseItem it = its.get(blockIdx.x);
    const int n = it.n, kc = it.k;
    const int c0 = blockIdx.y * NC;
    if (c0 >= kc || n <= 0) return;
    const int nc = min(NC, kc - c0);
    const double *R = it.R;
    const int64_t ldr = it.ldr, ldb = it.lda;
    double *B = it.A + ldb * c0;
    const int tid = threadIdx.x;
    const int nb = (n + NB - 1) / NB;
    // B's block (rows past n zero; thread r loads row r) and the diagonal
    // blocks of R (the part past n the identity; thread (b, c) loads column
    // c of block b and forms 1/R_cc into the padding column) into LDS
    {
        const int r = tid;
#pragma unroll
        for (int c = 0; c < NC; ++c) {
            const bool live = r < n && c < nc;
            const double v = B[(live ? r : 0) + ldb * (live ? c : 0)];
            Bs[r][c] = live ? v : 0.0;
        }
        const int b = tid / NB, c = tid % NB, gc = b * NB + c;
        if (b < nb) {
            const double *col = R + ldr * (gc < n ? gc : 0) + b * NB;
#pragma unroll
            for (int i = 0; i < NB; ++i) {
                const int gi = b * NB + i;
                const double v = col[(gi < n && gc < n) ? i : 0];
                Ri[b][i][c] = (gi < n && gc < n) ? (i <= c ? v : 0.0) : (i == c ? 1.0 : 0.0);
            }
            Ri[b][c][NB] = 1.0 / Ri[b][c][c];
        }
    }
    __syncthreads();
    {
        // two blocks per wavefront (one per half): lane c forms column c of
        // the inverse by back substitution (R_bb[i][t] from Rs, the same
        // address for the lanes of a half; the pivots applied as the
        // reciprocals in the padding column). A wave reads and writes only
        // its own blocks.
        const int lane = tid % 64, c = lane % NB, b = 2 * (tid / 64) + lane / NB;
        double x[NB];
        if (b < nb) {
            const double (*Rb)[NB + 1] = Ri[b];
#pragma unroll
            for (int i = NB - 1; i >= 0; --i) {
                double s0 = 0.0, s1 = 0.0;
#pragma unroll
                for (int t = i + 1; t < NB; ++t) {
                    if ((t - i) & 1) s0 += Rb[i][t] * x[t];
                    else s1 += Rb[i][t] * x[t];
                }
                x[i] = i > c ? 0.0 : (i == c ? 1.0 : -(s0 + s1)) * Rb[i][NB];
            }
        }
        __syncthreads();
        if (b < nb) {
#pragma unroll
            for (int t = 0; t < NB; ++t) Ri[b][t][c] = x[t];
        }
    }
    const int r = tid;
    double rr[NB], rn[NB];
    trsm_slab(R, ldr, n, 0, r, rr);
    __syncthreads();
    for (int k = 0; k < nb; ++k) {
        const int k0 = k * NB, kb = min(NB, n - k0);
        if (k + 1 < nb) trsm_slab(R, ldr, n, k + 1, r, rn);
        // X = inv(R_kk)^T B[k0:k0+NB, :]
        for (int e = tid; e < NB * nc; e += WGT) {
            const int i = e % NB, c = e / NB;
            double x[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
            for (int t = 0; t < NB; ++t) x[t % 4] += Ri[k][t][i] * Bs[k0 + t][c];
            Xs[i][c] = (x[0] + x[1]) + (x[2] + x[3]);
        }
        __syncthreads();
        if (r >= k0 + kb && r < n) {
            for (int c = 0; c < nc; ++c) {
                double a = 0.0;
#pragma unroll
                for (int t = 0; t < NB; ++t) a += rr[t] * Xs[t][c];
                Bs[r][c] -= a;
            }
        } else if (r >= k0 && r < k0 + kb) {
            for (int c = 0; c < nc; ++c) Bs[r][c] = Xs[r - k0][c];
        }
        __syncthreads();
#pragma unroll
        for (int t = 0; t < NB; ++t) rr[t] = rn[t];
    }
    for (int e = tid; e < n * nc; e += WGT) {
        const int q = e % n, c = e / n;
        B[q + ldb * c] = Bs[q][c];
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
    constexpr int PER = TK * TM / 256;  // entries of each operand per thread and K step
    double acc[4][4];
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b) acc[a][b] = 0.0;
    // the next K step's entries are loaded into registers while the current
    // one is multiplied from LDS
    double av[PER], bv[PER];
    auto load = [&](int k0) {
#pragma unroll
        for (int q = 0; q < PER; ++q) {
            const int e = tid + 256 * q, kk = e % TK, r = e / TK;
            const int gi = i0 + r, gj = j0 + r, gk = k0 + kk;
            double x = 0.0, y = 0.0;
            if (gi < M && gk < K) {
                x = ta ? A[gk + lda * gi] : A[gi + lda * gk];
                if (it.D) x *= it.D[gk];
            }
            if (gj < N && gk < K) y = tb ? B[gj + ldb * gk] : B[gk + ldb * gj];
            av[q] = x;
            bv[q] = y;
        }
    };
    load(0);
    for (int k0 = 0; k0 < K; k0 += TK) {
#pragma unroll
        for (int q = 0; q < PER; ++q) {
            const int e = tid + 256 * q, kk = e % TK, r = e / TK;
            As[kk][r] = av[q];
            Bs[kk][r] = bv[q];
        }
        __syncthreads();
        if (k0 + TK < K) load(k0 + TK);
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
    if (count > 0) k_potrf<<<(unsigned)count, WGF, 0, s>>>(its, info);
    return hipGetLastError();
}

hipError_t hmx_dense_trsm_t(DenseItems its, int64_t count, int64_t max_n, int64_t max_k,
                            hipStream_t s) {
    if (count <= 0 || max_k <= 0) return hipSuccess;
    const dim3 grid((unsigned)count, (unsigned)((max_k + NC - 1) / NC));
    if (max_n <= TS_MAX) {
        const size_t bytes = sizeof(double) * (size_t)TS_LDS;
        hipError_t e = hipFuncSetAttribute((const void *)k_trsm_t_lds,
                                           hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes);
        if (e != hipSuccess) return e;
        k_trsm_t_lds<<<grid, WGT, bytes, s>>>(its);
    } else {
        k_trsm_t<<<grid, WGT, 0, s>>>(its);
    }
    return hipGetLastError();
}

hipError_t hmx_dense_gemm(int ta, int tb, double alpha, double beta, DenseItems its,
                          int64_t count, int64_t max_tiles, hipStream_t s) {
    if (count > 0 && max_tiles > 0)
        k_gemm<<<dim3((unsigned)count, (unsigned)max_tiles), 256, 0, s>>>(ta, tb, alpha, beta,
                                                                          its);
    return hipGetLastError();
}
