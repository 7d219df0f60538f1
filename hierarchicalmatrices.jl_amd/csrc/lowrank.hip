// lowrank.hip -- gfx950 kernels of the batched LowRankMatrix arithmetic
// (src/LowRankMatrix.jl:64-171). One workgroup per item (and side) in every
// launch:
//   k_lr_pack    tall factors to factor: [U1 U2], [V1 V2], A or A^T
//   k_lr_gemm    tall x small products: A U2 and A^T V1 before the QR, and
//                U1 SU, V1 SU, V2 SV after the SVD
//   k_lr_gram    M = X1^T U2 (cores of L1 * L2 and L1' * L2)
//   k_lr_qr      Householder QR in place (LAPACK dlarfg convention: beta =
//                -sign(alpha) ||x||, tau = (beta - alpha)/beta, v(1) = 1)
//   k_lr_wy      T of the compact WY form Q = I - V T V^T (LAPACK dlarft
//                recurrence over the Gram matrix V^T V)
//   k_lr_svd     core, one-sided Jacobi SVD in LDS, descending order, getrank
//   k_lr_wz      W = T (V^T [S_; 0]) for the compact WY product
//   k_lr_applyq  Q [S_; 0] = [S_; 0] - V W, one pass over V
#include "hmx.h"
#include "hmx_lowrank.h"

namespace {

constexpr int WG = 256;
constexpr int LD = LR_MAX + 1;  // LDS leading dimension of the Jacobi matrices

__device__ inline double wave_sum(double v) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}


struct TallView {
    double *X;    // M x C, column-major, ld M
    int M, C;
    double *R;    // C x C (ld LR_MAX) in the small slot
    double *T;    // C x C compact-WY factor (ld LR_MAX)
    double *tau;  // C
};

__device__ inline TallView tall_view(const LRItem &it, int side, double *tall, double *small) {
    double *sl = small + it.small;
    TallView v;
    v.X = tall + (side ? it.xv : it.xu);
    v.M = side ? it.n : it.m;
    v.C = side ? it.cv : it.cu;
    v.R = sl + (side ? LR_RV : LR_RU);
    v.T = sl + (side ? LR_WV : LR_WU);
    v.tau = sl + (side ? LR_TV : LR_TU);
    return v;
}

// [U1 U2] -> XU, [V1 V2] -> XV (sums); A -> XU or A^T -> XV (svdtrunc)
__global__ __launch_bounds__(WG) void k_lr_pack(int op, const LRItem *items, double *tall) {
    const LRItem it = items[blockIdx.x];
    const int side = blockIdx.y;
    const int C = side ? it.cv : it.cu;
    if (C == 0) return;
    const int M = side ? it.n : it.m;
    double *X = tall + (side ? it.xv : it.xu);
    const int64_t cnt = (int64_t)M * C;
    for (int64_t e = (int64_t)blockIdx.z * WG + threadIdx.x; e < cnt; e += (int64_t)gridDim.z * WG) {
        const int i = (int)(e % M), c = (int)(e / M);
        double v;
        if (op == HM_LR_SVDTRUNC) {
            v = side ? it.A[c + (int64_t)it.m * i] : it.A[e];
        } else {
            const double *F1 = side ? it.V1 : it.U1, *F2 = side ? it.V2 : it.U2;
            v = c < it.r1 ? F1[i + (int64_t)M * c] : F2[i + (int64_t)M * (c - it.r1)];
        }
        X[e] = v;
    }
}

// out (rows x R, ld rows) = X (rows x P) B (P x R); X(i, t) = X[i*si + t*st]
// (X == nullptr: the identity), B(t, c) = B[t*b0 + c*b1]. B goes through
// LDS in chunks of 64 rows; one thread per output row.
struct GemmDesc {
    const double *X;
    int64_t si, st;
    const double *B;
    int64_t b0, b1;
    double *out;
    int rows, P, R;
};

__device__ inline bool gemm_desc(int op, const LRItem &it, int stage, int side, double *tall,
                                 double *small, GemmDesc &g) {
    const double *sl = small + it.small;
    g.X = nullptr;
    g.si = 1;
    g.st = 0;
    if (stage == 0) {  // before the QR: the tall factor of A * L or L * A
        if (op == HM_LR_DMUL && side == 0) {
            g.X = it.A; g.si = 1; g.st = it.m;
            g.B = it.U2; g.b0 = 1; g.b1 = it.k;
            g.out = tall + it.xu; g.rows = it.m; g.P = it.k; g.R = it.r2;
            return true;
        }
        if (op == HM_LR_MULD && side == 1) {
            g.X = it.A; g.si = it.k; g.st = 1;
            g.B = it.V1; g.b0 = 1; g.b1 = it.k;
            g.out = tall + it.xv; g.rows = it.n; g.P = it.k; g.R = it.r1;
            return true;
        }
        return false;
    }
    // after the SVD: sides without a QR factor
    if ((side ? it.cv : it.cu) != 0) return false;
    g.R = (int)sl[LR_RANK];
    g.B = sl + (side ? LR_SV : LR_SU);
    g.b0 = 1;
    g.b1 = LR_MAX;
    if (side == 0) {
        g.out = it.Uo; g.rows = it.m;
        if (op == HM_LR_MUL || op == HM_LR_MULD) { g.X = it.U1; g.st = it.m; g.P = it.r1; }
        else if (op == HM_LR_ADJMUL) { g.X = it.V1; g.st = it.m; g.P = it.r1; }
        else { g.X = nullptr; g.P = it.m; }  // svdtrunc of A^T: U = SU
    } else {
        g.out = it.Vo; g.rows = it.n;
        if (op == HM_LR_MUL || op == HM_LR_ADJMUL || op == HM_LR_DMUL) {
            g.X = it.V2; g.st = it.n; g.P = it.r2;
        } else { g.X = nullptr; g.P = it.n; }  // svdtrunc of A: V = SV
    }
    return true;
}

template <int RB>
__global__ __launch_bounds__(WG) void k_lr_gemm(int op, int stage, const LRItem *items,
                                                double *tall, double *small) {
    const LRItem it = items[blockIdx.x];
    GemmDesc g;
    if (!gemm_desc(op, it, stage, blockIdx.y, tall, small, g)) return;
    const int row0 = blockIdx.z * WG;
    if (row0 >= g.rows) return;
    const int i = row0 + threadIdx.x;
    const bool valid = i < g.rows;
    __shared__ double bs[64 * RB];
    double acc[RB];
#pragma unroll
    for (int c = 0; c < RB; ++c) acc[c] = 0.0;
    for (int t0 = 0; t0 < g.P; t0 += 64) {
        const int tl = min(64, g.P - t0);
        __syncthreads();
        for (int e = threadIdx.x; e < tl * g.R; e += WG) {
            const int t = e % tl, c = e / tl;
            bs[t + 64 * c] = g.B[(int64_t)(t0 + t) * g.b0 + (int64_t)c * g.b1];
        }
        __syncthreads();
        if (valid) {
            for (int t = 0; t < tl; ++t) {
                const double x = g.X ? g.X[(int64_t)i * g.si + (int64_t)(t0 + t) * g.st]
                                     : (i == t0 + t ? 1.0 : 0.0);
#pragma unroll
                for (int c = 0; c < RB; ++c)
                    if (c < g.R) acc[c] += x * bs[t + 64 * c];
            }
        }
    }
    if (valid) {
#pragma unroll
        for (int c = 0; c < RB; ++c)
            if (c < g.R) g.out[i + (int64_t)g.rows * c] = acc[c];
    }
}

// M (r1 x r2, ld LR_MAX, in the RU place) = X1^T U2 over k rows; X1 = V1
// (L1 * L2) or U1 (L1' * L2). Rows stream through LDS in chunks of 32.
__global__ __launch_bounds__(WG) void k_lr_gram(int op, const LRItem *items, double *small) {
    const LRItem it = items[blockIdx.x];
    const double *X1 = op == HM_LR_MUL ? it.V1 : it.U1;
    const double *X2 = it.U2;
    const int K = it.k, A = it.r1, B = it.r2;
    __shared__ double t1[32 * LR_MAX], t2[32 * LR_MAX];
    double acc[16];
#pragma unroll
    for (int o = 0; o < 16; ++o) acc[o] = 0.0;
    const int nout = A * B;
    for (int t0 = 0; t0 < K; t0 += 32) {
        const int tl = min(32, K - t0);
        for (int e = threadIdx.x; e < tl * A; e += WG) {
            const int t = e % tl, a = e / tl;
            t1[t + 32 * a] = X1[(int64_t)a * K + t0 + t];
        }
        for (int e = threadIdx.x; e < tl * B; e += WG) {
            const int t = e % tl, b = e / tl;
            t2[t + 32 * b] = X2[(int64_t)b * K + t0 + t];
        }
        __syncthreads();
#pragma unroll
        for (int o = 0; o < 16; ++o) {
            const int e = threadIdx.x + o * WG;
            if (e < nout) {
                const int a = e % A, b = e / A;
                double s = 0.0;
                for (int t = 0; t < tl; ++t) s += t1[t + 32 * a] * t2[t + 32 * b];
                acc[o] += s;
            }
        }
        __syncthreads();
    }
    double *M = small + it.small + LR_RU;
#pragma unroll
    for (int o = 0; o < 16; ++o) {
        const int e = threadIdx.x + o * WG;
        if (e < nout) M[(e % A) + LR_MAX * (e / A)] = acc[o];
    }
}

// Householder QR in place (LAPACK dlarfg convention: beta = -sign(alpha) *
// norm, tau = (beta - alpha) / beta, v = [1; x[j+1:] / (alpha - beta)]), R
// (upper C x C, ld LR_MAX) and tau into the small slot. Group g of NT/32
// threads takes the columns j + g, j + g + 32, ... of step j; a step has two
// barriers. First every group forms the dot product of the tail of x_j
// with its column's tail by shuffles (q = 0 is |x_j[j+1:]|^2). Then every
// thread knows beta and tau, and the groups take x_k -= w_k v with
// w_k = tau (x_k[j] + v[j+1:] . x_k[j+1:]), v read as the unscaled tail
// times 1/(alpha - beta), each writing its R[j, k]. The tail of x_j is scaled
// (and beta stored) during the next step, which does not read column j.
template <int NT>
__device__ inline void qr_steps(double *X, int M, int C, const TallView &T, double *sums) {
    constexpr int GS = NT / 32, NG = 32;  // threads per column, groups
    const int tid = threadIdx.x, grp = tid / GS, l = tid % GS;
    double pscal = 1.0, pbeta = 0.0;  // the previous column's pending scale and beta
    for (int j = 0; j < C; ++j) {
        const double *xj = X + (int64_t)j * M;
        if (j > 0) {
            double *xp = X + (int64_t)(j - 1) * M;
            for (int i = j + tid; i < M; i += NT) xp[i] *= pscal;
            if (tid == 0) xp[j - 1] = pbeta;
        }
        const int nk = C - j;
        for (int q = grp; q < nk; q += NG) {
            const double *xk = X + (int64_t)(j + q) * M;
            double sq = 0.0;
            for (int i = j + 1 + l; i < M; i += GS) sq += xj[i] * xk[i];
            for (int o = 1; o < GS; o <<= 1) sq += __shfl_xor(sq, o, 64);
            if (l == 0) sums[q] = sq;
        }
        __syncthreads();
        const double xn2 = sums[0], alpha = xj[j];
        double beta = alpha, t = 0.0, scal = 1.0;
        if (xn2 > 0.0) {
            beta = -copysign(sqrt(alpha * alpha + xn2), alpha);
            t = (beta - alpha) / beta;
            scal = 1.0 / (alpha - beta);
        }
        for (int q = 1 + grp; q < nk; q += NG) {
            double *xk = X + (int64_t)(j + q) * M;
            if (t != 0.0) {
                const double wk = t * (xk[j] + scal * sums[q]);
                for (int i = j + 1 + l; i < M; i += GS) xk[i] -= wk * (scal * xj[i]);
                if (l == 0) {
                    const double r = xk[j] - wk;
                    xk[j] = r;
                    T.R[j + LR_MAX * (j + q)] = r;
                }
            } else if (l == 0) {
                T.R[j + LR_MAX * (j + q)] = xk[j];
            }
        }
        if (tid == 0) {
            T.tau[j] = t;
            T.R[j + LR_MAX * j] = beta;
        }
        pscal = t != 0.0 ? scal : 1.0;
        pbeta = beta;
        __syncthreads();
    }
    if (C > 0) {
        double *xp = X + (int64_t)(C - 1) * M;
        for (int i = C + tid; i < M; i += NT) xp[i] *= pscal;
        if (tid == 0) xp[C - 1] = pbeta;
    }
    for (int e = tid; e < C * C; e += NT) {
        const int a = e % C, b = e / C;
        if (a > b) T.R[a + LR_MAX * b] = 0.0;
    }
    __syncthreads();
}

// T of Q = H_0 ... H_{C-1} = I - V T V^T (dlarft, forward, columnwise):
// T[j][j] = tau_j, T[0:j, j] = -tau_j T[0:j, 0:j] (V^T v_j)[0:j]. The Gram
// entries (V^T V)[a][b], a < b, are dot products over rows >= b (v_b is zero
// above row b and 1 at row b), 8 threads each. G and Ts: LDS, ld gld.
template <int NT>
__device__ inline void wy_steps(const double *X, int M, int C, const TallView &T, double *G,
                                double *Ts, int gld) {
    constexpr int GS = NT / 32;
    const int tid = threadIdx.x, grp = tid / GS, l8 = tid % GS;
    const int npair = C * C;
    for (int e0 = 0; e0 < npair; e0 += 32) {
        const int e = e0 + grp;
        const int a = e % C, b = e / C;
        double sacc = 0.0;
        if (e < npair && a < b) {
            const double *va = X + (int64_t)a * M, *vb = X + (int64_t)b * M;
            if (l8 == 0) sacc = va[b];
            for (int i = b + 1 + l8; i < M; i += GS) sacc += va[i] * vb[i];
        }
        for (int o = 1; o < GS; o <<= 1) sacc += __shfl_xor(sacc, o, 64);
        if (e < npair && a < b && l8 == 0) G[a + gld * b] = sacc;
    }
    __syncthreads();
    // the recurrence by one wavefront (C <= 64 lanes): its LDS reads and
    // writes stay in program order, so no workgroup barrier per column
    if (tid < 64) {
        for (int j = 0; j < C; ++j) {
            const double t = T.tau[j];
            if (tid < j) {
                double sacc = 0.0;
                for (int l = tid; l < j; ++l) sacc += Ts[tid + gld * l] * G[l + gld * j];
                Ts[tid + gld * j] = -t * sacc;
            }
            if (tid == j) Ts[j + gld * j] = t;
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
        }
    }
    __syncthreads();
    for (int e = tid; e < C * C; e += NT) {
        const int a = e % C, b = e / C;
        T.T[a + LR_MAX * b] = a <= b ? Ts[a + gld * b] : 0.0;
    }
}

// doubles of LDS the in-LDS factorization of an M x C factor needs
__host__ __device__ inline int64_t lr_lds_need(int64_t M, int64_t C) {
    return M * C + 2 * C * (C + 1);
}

// Householder QR of a factor too large for LDS, in global memory (it stays
// L2-resident), with the same reflectors and convention as qr_steps but
// row-parallel: every thread owns rows i = tid, tid + WGR, ... For column j,
// one pass over the rows forms |x_j[j+1:]|^2 and the dot products
// x_j[j+1:] . x_k[j+1:] of all trailing columns (KCH at a time, summed over
// the wavefronts in LDS); then v = [1; x_j[j+1:] / (alpha - beta)] and
// w_k = tau (x_k[j] + v[j+1:] . x_k[j+1:]) follow, and a second pass scales
// the tail of x_j and takes x_k -= w_k v from every trailing column.
constexpr int WGR = 1024, KCH = 8;
__global__ __launch_bounds__(WGR) void k_lr_qr(const LRItem *items, double *tall, double *small) {
    const LRItem it = items[blockIdx.x];
    const TallView T = tall_view(it, blockIdx.y, tall, small);
    if (T.C == 0 || lr_lds_need(T.M, T.C) <= LR_LDS_DOUBLES) return;
    __shared__ double red[(WGR / 64) * KCH], sums[LR_MAX], wk[LR_MAX];
    double *X = T.X;
    const int M = T.M, C = T.C;
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    for (int j = 0; j < C; ++j) {
        double *xj = X + (int64_t)j * M;
        const int nk = C - j;  // q = 0: the norm of x_j's tail; q > 0: column j + q
        for (int k0 = 0; k0 < nk; k0 += KCH) {
            double part[KCH];
#pragma unroll
            for (int q = 0; q < KCH; ++q) part[q] = 0.0;
            for (int i = j + 1 + tid; i < M; i += WGR) {
                const double xi = xj[i];
#pragma unroll
                for (int q = 0; q < KCH; ++q)
                    if (k0 + q < nk) part[q] += xi * X[(int64_t)(j + k0 + q) * M + i];
            }
#pragma unroll
            for (int q = 0; q < KCH; ++q) {
                const double sq = wave_sum(part[q]);
                if (lane == 0) red[wv * KCH + q] = sq;
            }
            __syncthreads();
            if (tid < KCH && k0 + tid < nk) {
                double sq = 0.0;
                for (int w = 0; w < WGR / 64; ++w) sq += red[w * KCH + tid];
                sums[k0 + tid] = sq;
            }
            __syncthreads();
        }
        const double xn2 = sums[0], alpha = xj[j];
        double beta = alpha, t = 0.0, scal = 0.0;
        if (xn2 > 0.0) {
            beta = -copysign(sqrt(alpha * alpha + xn2), alpha);
            t = (beta - alpha) / beta;
            scal = 1.0 / (alpha - beta);
        }
        if (tid > 0 && tid < nk) wk[tid] = t * (X[(int64_t)(j + tid) * M + j] + scal * sums[tid]);
        __syncthreads();
        if (t != 0.0) {
            for (int i = j + tid; i < M; i += WGR) {
                const double vi = i == j ? 1.0 : xj[i] * scal;
                if (i > j) xj[i] = vi;
                int q = 1;
                for (; q + 3 < nk; q += 4) {
                    double *x0 = X + (int64_t)(j + q) * M + i;
                    const double a0 = x0[0], a1 = x0[M], a2 = x0[2 * (int64_t)M],
                                 a3 = x0[3 * (int64_t)M];
                    x0[0] = a0 - wk[q] * vi;
                    x0[M] = a1 - wk[q + 1] * vi;
                    x0[2 * (int64_t)M] = a2 - wk[q + 2] * vi;
                    x0[3 * (int64_t)M] = a3 - wk[q + 3] * vi;
                }
                for (; q < nk; ++q) X[(int64_t)(j + q) * M + i] -= wk[q] * vi;
            }
        }
        __syncthreads();
        if (tid == 0) {
            xj[j] = beta;
            T.tau[j] = t;
            T.R[j + LR_MAX * j] = beta;
        }
        if (tid > 0 && tid < nk) T.R[j + LR_MAX * (j + tid)] = X[(int64_t)(j + tid) * M + j];
        __syncthreads();
    }
    for (int e = tid; e < C * C; e += WGR) {
        const int a = e % C, b = e / C;
        if (a > b) T.R[a + LR_MAX * b] = 0.0;
    }
}

// T of the compact WY form for a factor too large for LDS: the Gram entries
// G[a][b] = v_a . v_b (a < b; v_b is 1 at row b and zero above) row-parallel
// as in k_lr_qr, one pass over the rows below b per KCH columns a, then the
// dlarft recurrence of wy_steps in LDS.
__global__ __launch_bounds__(WGR) void k_lr_wy(const LRItem *items, double *tall, double *small) {
    const LRItem it = items[blockIdx.x];
    const TallView T = tall_view(it, blockIdx.y, tall, small);
    if (T.C == 0 || lr_lds_need(T.M, T.C) <= LR_LDS_DOUBLES) return;
    __shared__ double G[LR_MAX * LD], Ts[LR_MAX * LD], red[(WGR / 64) * KCH];
    const double *X = T.X;
    const int M = T.M, C = T.C;
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    for (int b = 1; b < C; ++b) {
        const double *vb = X + (int64_t)b * M;
        for (int a0 = 0; a0 < b; a0 += KCH) {
            double part[KCH];
#pragma unroll
            for (int q = 0; q < KCH; ++q) part[q] = 0.0;
            for (int i = b + 1 + tid; i < M; i += WGR) {
                const double x = vb[i];
#pragma unroll
                for (int q = 0; q < KCH; ++q)
                    if (a0 + q < b) part[q] += X[(int64_t)(a0 + q) * M + i] * x;
            }
#pragma unroll
            for (int q = 0; q < KCH; ++q) {
                const double sq = wave_sum(part[q]);
                if (lane == 0) red[wv * KCH + q] = sq;
            }
            __syncthreads();
            if (tid < KCH && a0 + tid < b) {
                double sq = X[(int64_t)(a0 + tid) * M + b];  // v_a at row b (v_b[b] = 1)
                for (int w = 0; w < WGR / 64; ++w) sq += red[w * KCH + tid];
                G[(a0 + tid) + LD * b] = sq;
            }
            __syncthreads();
        }
    }
    for (int j = 0; j < C; ++j) {
        const double t = T.tau[j];
        if (tid < j) {
            double sacc = 0.0;
            for (int l = tid; l < j; ++l) sacc += Ts[tid + LD * l] * G[l + LD * j];
            Ts[tid + LD * j] = -t * sacc;
        }
        if (tid == j) Ts[j + LD * j] = t;
        __syncthreads();
    }
    for (int e = tid; e < C * C; e += WGR) {
        const int a = e % C, b = e / C;
        T.T[a + LR_MAX * b] = a <= b ? Ts[a + LD * b] : 0.0;
    }
}

// factors that fit: copy to LDS, QR and T there, copy the reflectors back.
// One workgroup per CU fits beside that much LDS, so it has 1024 threads
// (4 waves per SIMD) and 32 threads per trailing column.
constexpr int WGQ = 1024;
__global__ __launch_bounds__(WGQ) void k_lr_qr_lds(const LRItem *items, double *tall,
                                                   double *small) {
    const LRItem it = items[blockIdx.x];
    const TallView T = tall_view(it, blockIdx.y, tall, small);
    if (T.C == 0 || lr_lds_need(T.M, T.C) > LR_LDS_DOUBLES) return;
    extern __shared__ double dyn[];
    __shared__ double sums[LR_MAX];
    const int M = T.M, C = T.C;
    const int64_t cnt = (int64_t)M * C;
    double *xs = dyn;
    double *G = dyn + cnt, *Ts = G + C * (C + 1);
    for (int64_t e = threadIdx.x; e < cnt; e += WGQ) xs[e] = T.X[e];
    __syncthreads();
    qr_steps<WGQ>(xs, M, C, T, sums);
    __syncthreads();
    wy_steps<WGQ>(xs, M, C, T, G, Ts, C + 1);
    for (int64_t e = threadIdx.x; e < cnt; e += WGQ) T.X[e] = xs[e];
}

// entry (a, b) of the core whose SVD gives the result
__device__ inline double core_entry(int op, const LRItem &it, const double *sl, int a, int b) {
    const double *RU = sl + LR_RU, *RV = sl + LR_RV;
    switch (op) {
    case HM_LR_ADD:
    case HM_LR_SUB: {  // RU diag([S1; +-S2]) RV^T
        double g = 0.0;
        for (int t = a > b ? a : b; t < it.cu; ++t) {
            const double d = t < it.r1 ? it.S1[t]
                                       : (op == HM_LR_SUB ? -it.S2[t - it.r1] : it.S2[t - it.r1]);
            g += (RU[a + LR_MAX * t] * d) * RV[b + LR_MAX * t];
        }
        return g;
    }
    case HM_LR_MUL:
    case HM_LR_ADJMUL:  // diag(S1) M diag(S2)
        return (it.S1[a] * RU[a + LR_MAX * b]) * it.S2[b];
    case HM_LR_DMUL:  // R diag(S2)
        return RU[a + LR_MAX * b] * it.S2[b];
    case HM_LR_MULD:  // diag(S1) R^T
        return it.S1[a] * RV[b + LR_MAX * a];
    default:  // svdtrunc: R, or R^T when A^T was factored
        return it.tr ? RV[b + LR_MAX * a] : RU[a + LR_MAX * b];
    }
}

// One-sided Jacobi (Hestenes) on the core: the columns of W (P x Q, Q <= P;
// W = G, or G^T when G has more columns than rows) are rotated pairwise in
// round-robin order until every pair is orthogonal to 1e-15 relative; V
// accumulates the rotations. Then S = column norms sorted descending,
// U = W V-columns / S, getrank. Writes S, SU (p x min), SV (q x min), the
// rank, and the result's S.
__global__ __launch_bounds__(WG) void k_lr_svd(int op, const LRItem *items, double *small) {
    const LRItem it = items[blockIdx.x];
    double *sl = small + it.small;
    const int p = it.p, q = it.q, tid = threadIdx.x;
    const bool trn = q > p;
    const int P = trn ? q : p, Q = trn ? p : q;
    // W: P x Q (ld P + 1), V: Q x Q (ld Q + 1), nrm: Q
    extern __shared__ double dyn[];
    const int LW = P + 1, LV = Q + 1;
    double *W = dyn, *V = dyn + LW * Q, *nrm = V + LV * Q;
    __shared__ int flag;
    for (int e = tid; e < p * q; e += WG) {
        const int a = e % p, b = e / p;
        const double g = core_entry(op, it, sl, a, b);
        if (trn) W[b + LW * a] = g;
        else W[a + LW * b] = g;
    }
    for (int e = tid; e < Q * Q; e += WG) {
        const int a = e % Q, b = e / Q;
        V[a + LV * b] = a == b ? 1.0 : 0.0;
    }
    __syncthreads();
    const int Qe = Q + (Q & 1), npairs = Qe / 2;
    const int grp = tid >> 3, l8 = tid & 7;
    for (int sweep = 0; sweep < 60 && Q > 1; ++sweep) {
        if (tid == 0) flag = 0;
        __syncthreads();
        for (int rd = 0; rd < Qe - 1; ++rd) {
            if (grp < npairs) {
                int a, b;
                if (grp == 0) {
                    a = 0;
                    b = 1 + rd;
                } else {
                    a = 1 + (rd + grp) % (Qe - 1);
                    b = 1 + (rd - grp + Qe - 1) % (Qe - 1);
                }
                if (a < Q && b < Q) {
                    double al = 0.0, be = 0.0, ga = 0.0;
                    for (int i = l8; i < P; i += 8) {
                        const double x = W[i + LW * a], y = W[i + LW * b];
                        al += x * x;
                        be += y * y;
                        ga += x * y;
                    }
                    for (int o = 1; o < 8; o <<= 1) {
                        al += __shfl_xor(al, o, 64);
                        be += __shfl_xor(be, o, 64);
                        ga += __shfl_xor(ga, o, 64);
                    }
                    if (ga != 0.0 && fabs(ga) > 1e-15 * sqrt(al * be)) {
                        if (l8 == 0) flag = 1;
                        const double zeta = (be - al) / (2.0 * ga);
                        const double t = copysign(1.0, zeta) / (fabs(zeta) + hypot(1.0, zeta));
                        const double c = 1.0 / sqrt(1.0 + t * t), s = c * t;
                        for (int i = l8; i < P; i += 8) {
                            const double x = W[i + LW * a], y = W[i + LW * b];
                            W[i + LW * a] = c * x - s * y;
                            W[i + LW * b] = s * x + c * y;
                        }
                        for (int i = l8; i < Q; i += 8) {
                            const double x = V[i + LV * a], y = V[i + LV * b];
                            V[i + LV * a] = c * x - s * y;
                            V[i + LV * b] = s * x + c * y;
                        }
                    }
                }
            }
            __syncthreads();
        }
        const int f = flag;
        __syncthreads();
        if (!f) break;
    }
    for (int j = tid; j < Q; j += WG) {
        double s = 0.0;
        for (int i = 0; i < P; ++i) s += W[i + LW * j] * W[i + LW * j];
        nrm[j] = sqrt(s);
    }
    __syncthreads();
    double *Uo = sl + (trn ? LR_SV : LR_SU);  // P x Q
    double *Vo = sl + (trn ? LR_SU : LR_SV);  // Q x Q
    double *S = sl + LR_S;
    for (int j = tid; j < Q; j += WG) {
        const double sj = nrm[j];
        int pos = 0;
        for (int k = 0; k < Q; ++k) pos += (nrm[k] > sj) || (nrm[k] == sj && k < j);
        S[pos] = sj;
        for (int i = 0; i < P; ++i) Uo[i + LR_MAX * pos] = sj > 0.0 ? W[i + LW * j] / sj : 0.0;
        for (int i = 0; i < Q; ++i) Vo[i + LR_MAX * pos] = V[i + LV * j];
    }
    __syncthreads();
    if (tid == 0) {
        // getrank, src/LowRankMatrix.jl:70-80: tol = length(S) * eps(S[1])
        const double tol = (double)Q * (nextafter(S[0], INFINITY) - S[0]);
        int r = Q;
        while (r >= 1 && !(S[r - 1] > tol)) --r;
        sl[LR_RANK] = (double)r;
    }
    __syncthreads();
    const int r = (int)sl[LR_RANK];
    for (int c = tid; c < r; c += WG) it.So[c] = S[c];
}

// W = T (V^T [B; 0]) (C x r) for out = Q [B; 0] = [B; 0] - V W, with B the
// first r columns of SU or SV (C x r) and V unit lower triangular in its top
// C rows. W goes to the RU / RV place of the small slot (the SVD has used
// R by now), where k_lr_applyq reads it with wave-uniform addresses.
__global__ __launch_bounds__(WG) void k_lr_wz(const LRItem *items, double *tall, double *small,
                                              int ldw) {
    const LRItem it = items[blockIdx.x];
    const int side = blockIdx.y;
    const TallView T = tall_view(it, side, tall, small);
    if (T.C == 0) return;
    double *sl = small + it.small;
    const double *B = sl + (side ? LR_SV : LR_SU);
    double *Wg = sl + (side ? LR_RV : LR_RU);
    const int M = T.M, C = T.C, tid = threadIdx.x;
    const int R = (int)sl[LR_RANK];
    const double *X = T.X;
    // the top C rows of V, B's first r columns and T staged in LDS (one
    // batch of loads each) before the products; ldw = the call's largest C
    // plus one (dynamic LDS, 4 * C * ldw doubles)
    extern __shared__ double wzs[];
    const int LW = ldw, CW = ldw - 1;
    double *Z = wzs, *Vt = Z + CW * LW, *Bs = Vt + CW * LW, *Ts = Bs + CW * LW;
    for (int e = tid; e < C * C; e += WG) {
        const int i = e % C, a = e / C;
        Vt[i + LW * a] = X[(int64_t)a * M + i];
        Ts[i + LW * a] = T.T[i + LR_MAX * a];
    }
    for (int e = tid; e < C * R; e += WG) {
        const int i = e % C, k = e / C;
        Bs[i + LW * k] = B[i + LR_MAX * k];
    }
    __syncthreads();
    for (int e = tid; e < C * R; e += WG) {
        const int a = e % C, k = e / C;
        double z = Bs[a + LW * k];  // V[a][a] = 1, V[i][a] stored below
        for (int i = a + 1; i < C; ++i) z += Vt[i + LW * a] * Bs[i + LW * k];
        Z[a + LW * k] = z;
    }
    __syncthreads();
    for (int e = tid; e < C * R; e += WG) {
        const int a = e % C, k = e / C;
        double wv = 0.0;
        for (int l = a; l < C; ++l) wv += Ts[a + LW * l] * Z[l + LW * k];
        Wg[k + LR_MAX * a] = wv;  // row a contiguous: W[a][0..r)
    }
}

// out[i, :] = [B; 0][i, :] - V[i, :] W, one thread per row i, with RB
// accumulators (RB >= rank R); W (C x r) is read at wave-uniform addresses,
// and the reflector entries of four columns are loaded before their products
template <int RB>
__device__ inline void applyq_row(const double *X, int M, int C, int R, int i, const double *Wg,
                                  const double *B, double *out) {
    double acc[RB];
#pragma unroll
    for (int k = 0; k < RB; ++k) acc[k] = 0.0;
    for (int t0 = 0; t0 < C; t0 += 4) {
        double v[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            // V[i][t]: 0 above the diagonal, 1 on it, stored below
            const int t = t0 + u;
            const bool below = t < C && i > t && i < M;
            const double x = X[below ? (int64_t)t * M + i : 0];
            v[u] = below ? x : (t < C && i == t ? 1.0 : 0.0);
        }
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            if (t0 + u < C) {
                const double *wr = Wg + LR_MAX * (t0 + u);
#pragma unroll
                for (int k = 0; k < RB; ++k) acc[k] += v[u] * wr[k];
            }
        }
    }
    if (i < M) {
#pragma unroll
        for (int k = 0; k < RB; ++k)
            if (k < R) out[i + (int64_t)M * k] = (i < C ? B[i + LR_MAX * k] : 0.0) - acc[k];
    }
}

__global__ __launch_bounds__(WG) void k_lr_applyq(const LRItem *items, double *tall,
                                                  double *small) {
    const LRItem it = items[blockIdx.x];
    const int side = blockIdx.y;
    const TallView T = tall_view(it, side, tall, small);
    if (T.C == 0) return;
    const double *sl = small + it.small;
    const double *B = sl + (side ? LR_SV : LR_SU);
    const double *Wg = sl + (side ? LR_RV : LR_RU);
    double *out = side ? it.Vo : it.Uo;
    const int M = T.M, C = T.C;
    const int R = (int)sl[LR_RANK];
    const int i = blockIdx.z * WG + threadIdx.x;
    if (blockIdx.z * WG >= M) return;
    // W's columns past R hold finite values (or zeros) and are not stored
    if (R <= 16) applyq_row<16>(T.X, M, C, R, i, Wg, B, out);
    else if (R <= 32) applyq_row<32>(T.X, M, C, R, i, Wg, B, out);
    else applyq_row<LR_MAX>(T.X, M, C, R, i, Wg, B, out);
}

__global__ void k_lr_ranks(const LRItem *items, int64_t count, const double *small,
                           int64_t *rank) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < count) rank[t] = (int64_t)small[items[t].small + LR_RANK];
}

}  // namespace

hipError_t hmx_lowrank_run(int op, const LRItem *items, int64_t count, double *tall,
                           double *small, int64_t *rank, int64_t max_rows, int64_t lds_qr,
                           bool big_qr, int64_t lds_svd, int64_t max_c, hipStream_t s) {
    if (count <= 0) return hipSuccess;
    const unsigned n = (unsigned)count;
    const unsigned rowblk = (unsigned)((max_rows + WG - 1) / WG);
    if (op == HM_LR_ADD || op == HM_LR_SUB || op == HM_LR_SVDTRUNC)
        k_lr_pack<<<dim3(n, 2, rowblk < 64 ? rowblk : 64), WG, 0, s>>>(op, items, tall);
    if (op == HM_LR_DMUL || op == HM_LR_MULD)
        k_lr_gemm<LR_MAX><<<dim3(n, 2, rowblk), WG, 0, s>>>(op, 0, items, tall, small);
    if (op == HM_LR_MUL || op == HM_LR_ADJMUL) {
        k_lr_gram<<<n, WG, 0, s>>>(op, items, small);
    } else {
        if (lds_qr > 0) {
            const size_t bytes = sizeof(double) * (size_t)lds_qr;
            (void)hipFuncSetAttribute((const void *)k_lr_qr_lds,
                                      hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes);
            k_lr_qr_lds<<<dim3(n, 2), WGQ, bytes, s>>>(items, tall, small);
        }
        if (big_qr) {
            k_lr_qr<<<dim3(n, 2), WGR, 0, s>>>(items, tall, small);
            k_lr_wy<<<dim3(n, 2), WGR, 0, s>>>(items, tall, small);
        }
    }
    k_lr_svd<<<n, WG, sizeof(double) * (size_t)lds_svd, s>>>(op, items, small);
    if (op != HM_LR_MUL && op != HM_LR_ADJMUL) {
        const int ldw = (int)max_c + 1;
        const size_t wz_bytes = sizeof(double) * 4 * (size_t)max_c * ldw;
        if (wz_bytes > 65536)
            (void)hipFuncSetAttribute((const void *)k_lr_wz,
                                      hipFuncAttributeMaxDynamicSharedMemorySize, (int)wz_bytes);
        k_lr_wz<<<dim3(n, 2), WG, wz_bytes, s>>>(items, tall, small, ldw);
        k_lr_applyq<<<dim3(n, 2, rowblk), WG, 0, s>>>(items, tall, small);
    }
    k_lr_gemm<LR_MAX><<<dim3(n, 2, rowblk), WG, 0, s>>>(op, 1, items, tall, small);
    k_lr_ranks<<<(n + 255) / 256, 256, 0, s>>>(items, count, small, rank);
    return hipGetLastError();
}
