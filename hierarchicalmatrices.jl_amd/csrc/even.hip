// even.hip -- gfx950 kernels of the EvenBarycentricMatrix leaves
// (src/BarycentricMatrix.jl:5-89, mul! src/algebra.jl:168-239).
//
// Every pass is bound by HBM bandwidth: a product reads F (p x n) and W
// (n x m) once per leaf and does 2 flops per 8-byte factor entry.
#include "hmx_even.h"

namespace {

constexpr int WG = 256;

// W[k, i] = lambda_k * inv((2i-a-b-(b-a)x_k) * w_i), w_i = sum_k lambda_k *
// inv(2i-a-b-(b-a)x_k), src/BarycentricMatrix.jl:23-37 (and :62-75 for the
// m x n layout of barycentricmatrix), in the reference's operation order.
template <class T>
__global__ __launch_bounds__(WG) void k_even_gen_w(const ELeaf *leaves, int64_t nleaves,
                                                   ENodes<T> nd, int n, int layout, T *W, T *w) {
#pragma clang fp contract(off)
    for (int64_t t = blockIdx.y; t < nleaves; t += gridDim.y) {
        const ELeaf L = leaves[t];
        const T ba = (T)(L.b - L.a);
        for (int64_t i = (int64_t)blockIdx.x * WG + threadIdx.x; i < L.m;
             i += (int64_t)gridDim.x * WG) {
            const T c0 = (T)(2 * (L.a + i) - L.a - L.b);
            T s = T(0);
            for (int k = 0; k < n; ++k) s += nd.lam[k] * (T(1) / (c0 - ba * nd.x[k]));
            if (w) w[L.row0 + i] = s;
            T *out = W + L.woff;
            for (int k = 0; k < n; ++k) {
                const T v = nd.lam[k] * (T(1) / ((c0 - ba * nd.x[k]) * s));
                if (layout == 0) out[k + (int64_t)n * i] = v;
                else if (layout == 1) out[i + L.m * k] = v;
                else out[even_tile(i, k, n)] = v;
            }
        }
    }
}

// column-major staging copy of F -> 64-row tiles
template <class T>
__global__ __launch_bounds__(WG) void k_even_tile_F(const ELeaf *leaves, int64_t nleaves, int n,
                                                    const T *stage, T *F) {
    for (int64_t t = blockIdx.y; t < nleaves; t += gridDim.y) {
        const ELeaf L = leaves[t];
        const int64_t cnt = L.p * n;
        for (int64_t e = (int64_t)blockIdx.x * WG + threadIdx.x; e < cnt;
             e += (int64_t)gridDim.x * WG) {
            const int64_t k = e / L.p, j = e - k * L.p;
            F[L.foff + even_tile(j, k, n)] = stage[L.soff + e];
        }
    }
}

template <class T>
__device__ inline T wave_xor(T v, int o) {
    return __shfl_xor(v, o, 64);
}

// Upward pass: one wavefront per task. Lane l takes the task's columns
// l, l + 64, ..., so all its columns share one global parity; lanes of equal
// parity are summed with xor shuffles 2..32, after which every even lane holds
// the sums of one parity and every odd lane those of the other. Lane l then
// writes k = kb + l/2 of its parity. Output e = (q*2 + par)*n + k of the
// leaf's beta (single-task leaves) or of the task's partial slot.
template <class T, int RB, int KB>
__global__ __launch_bounds__(64) void k_even_up(EvenDev<T> p, const T *v, int64_t joff,
                                               int64_t incv, int64_t ldv, int nq) {
    const EUp tk = p.up[blockIdx.x];
    const ELeaf L = p.leaves[tk.leaf];
    const int n = p.n;
    const int lane = threadIdx.x;
    const T *F = p.F + L.foff;
    const int64_t jg = joff + L.col0;  // global column of the leaf's column 0
    T acc[KB][RB];
#pragma unroll
    for (int q = 0; q < KB; ++q)
#pragma unroll
        for (int k = 0; k < RB; ++k) acc[q][k] = T(0);
    const int64_t jend = tk.j0 + tk.len;
    for (int64_t j = tk.j0 + lane; j < jend; j += 64) {
        // all loads of the column first, then the products: each load is
        // guarded by k < n, and a guard around load and use would wait on
        // every load in turn. F is read once per product: non-temporal
        // loads (W, shared by the leaves of one size, stays cached instead)
        T x[KB], f[RB];
#pragma unroll
        for (int q = 0; q < KB; ++q) x[q] = q < nq ? v[q * ldv + (jg + j) * incv] : T(0);
        const T *Fj = F + even_tile(j, 0, n);
#pragma unroll
        for (int k = 0; k < RB; ++k) f[k] = k < n ? __builtin_nontemporal_load(Fj + k * EVEN_TILE) : T(0);
#pragma unroll
        for (int k = 0; k < RB; ++k)
#pragma unroll
            for (int q = 0; q < KB; ++q) acc[q][k] += f[k] * x[q];
    }
#pragma unroll
    for (int q = 0; q < KB; ++q)
#pragma unroll
        for (int k = 0; k < RB; ++k)
            if (k < n)
#pragma unroll
                for (int o = 2; o < 64; o <<= 1) acc[q][k] += wave_xor(acc[q][k], o);
    const int par = (int)((jg + tk.j0 + lane) & 1);
    const int64_t stride = (int64_t)HM_EVEN_KB * 2 * n;
    T *out = tk.single ? p.beta + (int64_t)tk.leaf * stride : p.partial + tk.slot * stride;
#pragma unroll
    for (int kb = 0; kb < RB; kb += 32) {
        const int kk = kb + (lane >> 1);
        if (kk < n) {
#pragma unroll
            for (int q = 0; q < KB; ++q) {
                T val = T(0);
#pragma unroll
                for (int k = kb; k < kb + 32 && k < RB; ++k)
                    if (k == kk) val = acc[q][k];
                out[(q * 2 + par) * n + kk] = val;
            }
        }
    }
}

// beta of a leaf with several tasks: the sum of its partials; thread
// (g, e) sums slots g, g + G, ... of entry e, then the G partial sums are
// added in g order. With more than 256 entries (2 parities x kb right-hand
// sides x rank n) each thread sums whole entries e, e + 256, ... in slot order.
template <class T>
__global__ __launch_bounds__(256) void k_even_reduce(EvenDev<T> p, int kb) {
    const ERed r = p.red[blockIdx.x];
    const int64_t stride = (int64_t)HM_EVEN_KB * 2 * p.n;
    const int E = 2 * kb * p.n;
    if (E > 256) {  // block-uniform: no thread reaches the barrier below
        for (int e = threadIdx.x; e < E; e += 256) {
            T s = T(0);
            for (int t = 0; t < r.ntask; ++t) s += p.partial[(r.slot0 + t) * stride + e];
            p.beta[(int64_t)r.leaf * stride + e] = s;
        }
        return;
    }
    const int G = 256 / E;
    const int e = threadIdx.x % E, g = threadIdx.x / E;
    __shared__ T part[256];
    T s = T(0);
    if (g < G)
        for (int t = g; t < r.ntask; t += G) s += p.partial[(r.slot0 + t) * stride + e];
    part[threadIdx.x] = s;
    __syncthreads();
    if (threadIdx.x < E) {
        T tot = T(0);
        for (int q = 0; q < G; ++q) tot += part[q * E + threadIdx.x];
        p.beta[(int64_t)r.leaf * stride + threadIdx.x] = tot;
    }
}

// Downward pass: one thread per row of a segment. Row I pairs with the
// columns of its own parity (I + J even), so it takes beta[I % 2]. W is in
// 64-row tiles, so a wavefront's loads of one k are one contiguous run; the
// current leaf's beta is staged in LDS (two distinct addresses per wave).
template <class T, int RB, int KB>
__global__ __launch_bounds__(WG) void k_even_down(EvenDev<T> p, T *u, int64_t ioff, int64_t incu,
                                                  int64_t ldu, int nq, int store) {
    __shared__ T sb[2 * KB * RB];
    const EDown d = p.down[blockIdx.x];
    const int tid = threadIdx.x;
    const bool valid = tid < d.len;
    const int n = p.n;
    const int64_t r = d.r0 + (valid ? tid : 0);
    const int par = (int)((ioff + r) & 1);
    const int64_t stride = (int64_t)HM_EVEN_KB * 2 * n;
    T s[KB];
#pragma unroll
    for (int q = 0; q < KB; ++q) s[q] = T(0);
    for (int e = d.lbeg; e < d.lend; ++e) {
        const int32_t id = p.list[e];
        const ELeaf L = p.leaves[id];
        __syncthreads();
        for (int t = tid; t < 2 * KB * n; t += WG) sb[t] = p.beta[(int64_t)id * stride + t];
        __syncthreads();
        if (valid) {
            const T *wr = p.W + L.woff + even_tile(r - L.row0, 0, n);
            T wk[RB];
#pragma unroll
            for (int k = 0; k < RB; ++k) wk[k] = k < n ? wr[k * EVEN_TILE] : T(0);
#pragma unroll
            for (int q = 0; q < KB; ++q) {
                const T *bq = sb + (q * 2 + par) * n;
                T a = T(0);
#pragma unroll
                for (int k = 0; k < RB; ++k)
                    if (k < n) a += wk[k] * bq[k];
                s[q] += a;
            }
        }
    }
    if (valid) {
#pragma unroll
        for (int q = 0; q < KB; ++q) {
            if (q < nq) {
                T *uo = u + q * ldu + (ioff + r) * incu;
                *uo = store ? s[q] : *uo + s[q];
            }
        }
    }
}

template <class T>
__global__ __launch_bounds__(WG) void k_even_zero(T *u, int64_t ioff, int64_t incu, int64_t ldu,
                                                  int64_t nrows, int nq) {
    const int64_t i = (int64_t)blockIdx.x * WG + threadIdx.x;
    if (i >= nrows) return;
    for (int q = 0; q < nq; ++q) u[q * ldu + (ioff + i) * incu] = T(0);
}

template <class T, int RB, int KB>
hipError_t even_matvec_t(const EvenDev<T> &p, T *u, const T *v, int64_t ioff, int64_t joff,
                         int64_t incu, int64_t incv, int64_t ldu, int64_t ldv, int nq, int beta,
                         hipStream_t s) {
    for (int q0 = 0; q0 < nq; q0 += KB) {
        const int nb = nq - q0 < KB ? nq - q0 : KB;
        T *uq = u + q0 * ldu;
        const T *vq = v + q0 * ldv;
        // beta = 0: rows of a segment are written by the downward pass;
        // rows no leaf covers are cleared first
        const int store = beta == 0 && p.covered;
        if (beta == 0 && !p.covered && p.nrows > 0)
            k_even_zero<T><<<(unsigned)((p.nrows + WG - 1) / WG), WG, 0, s>>>(uq, ioff, incu, ldu,
                                                                          p.nrows, nb);
        if (p.nup > 0) k_even_up<T, RB, KB><<<p.nup, 64, 0, s>>>(p, vq, joff, incv, ldv, nb);
        if (p.nred > 0) k_even_reduce<T><<<p.nred, 256, 0, s>>>(p, nb);
        if (p.ndown > 0) k_even_down<T, RB, KB><<<p.ndown, WG, 0, s>>>(p, uq, ioff, incu, ldu, nb,
                                                                  store);
    }
    return hipGetLastError();
}

template <class T>
hipError_t gen_weights_t(const ELeaf *leaves, int64_t nleaves, int64_t max_m, const ENodes<T> &nodes,
                         int n, int layout, T *W, T *w, hipStream_t s) {
    if (nleaves <= 0 || max_m <= 0) return hipSuccess;
    const int64_t bx = (max_m + WG - 1) / WG;
    dim3 grid((unsigned)(bx < 1024 ? bx : 1024), (unsigned)(nleaves < 65535 ? nleaves : 65535));
    k_even_gen_w<T><<<grid, WG, 0, s>>>(leaves, nleaves, nodes, n, layout, W, w);
    return hipGetLastError();
}

template <class T>
hipError_t tile_F_t(const ELeaf *leaves, int64_t nleaves, int64_t max_pn, int n, const T *stage,
                    T *F, hipStream_t s) {
    if (nleaves <= 0 || max_pn <= 0) return hipSuccess;
    const int64_t bx = (max_pn + WG - 1) / WG;
    dim3 grid((unsigned)(bx < 4096 ? bx : 4096), (unsigned)(nleaves < 65535 ? nleaves : 65535));
    k_even_tile_F<T><<<grid, WG, 0, s>>>(leaves, nleaves, n, stage, F);
    return hipGetLastError();
}

template <class T>
hipError_t matvec_dispatch(const EvenDev<T> &p, T *u, const T *v, int64_t ioff, int64_t joff,
                           int64_t incu, int64_t incv, int64_t ldu, int64_t ldv, int nq, int beta,
                           hipStream_t s) {
    const int n = p.n;
    if (nq == 1) {  // one right-hand side: the single-column kernels (fewer registers)
        if (n <= 8) return even_matvec_t<T, 8, 1>(p, u, v, ioff, joff, incu, incv, ldu, ldv, nq, beta, s);
        if (n <= 16) return even_matvec_t<T, 16, 1>(p, u, v, ioff, joff, incu, incv, ldu, ldv, nq, beta, s);
        if (n <= 24) return even_matvec_t<T, 24, 1>(p, u, v, ioff, joff, incu, incv, ldu, ldv, nq, beta, s);
        if (n <= 32) return even_matvec_t<T, 32, 1>(p, u, v, ioff, joff, incu, incv, ldu, ldv, nq, beta, s);
    }
    if (n <= 8) return even_matvec_t<T, 8, 4>(p, u, v, ioff, joff, incu, incv, ldu, ldv, nq, beta, s);
    if (n <= 16) return even_matvec_t<T, 16, 4>(p, u, v, ioff, joff, incu, incv, ldu, ldv, nq, beta, s);
    if (n <= 24) return even_matvec_t<T, 24, 2>(p, u, v, ioff, joff, incu, incv, ldu, ldv, nq, beta, s);
    if (n <= 32) return even_matvec_t<T, 32, 2>(p, u, v, ioff, joff, incu, incv, ldu, ldv, nq, beta, s);
    if (n <= 48) return even_matvec_t<T, 48, 1>(p, u, v, ioff, joff, incu, incv, ldu, ldv, nq, beta, s);
    return even_matvec_t<T, 64, 1>(p, u, v, ioff, joff, incu, incv, ldu, ldv, nq, beta, s);
}

}  // namespace

hipError_t hmx_even_gen_weights(const ELeaf *leaves, int64_t nleaves, int64_t max_m,
                                const ENodes<double> &nodes, int n, int layout, double *W,
                                double *w, hipStream_t s) {
    return gen_weights_t<double>(leaves, nleaves, max_m, nodes, n, layout, W, w, s);
}
hipError_t hmx_even_gen_weights(const ELeaf *leaves, int64_t nleaves, int64_t max_m,
                                const ENodes<float> &nodes, int n, int layout, float *W,
                                float *w, hipStream_t s) {
    return gen_weights_t<float>(leaves, nleaves, max_m, nodes, n, layout, W, w, s);
}

hipError_t hmx_even_tile_F(const ELeaf *leaves, int64_t nleaves, int64_t max_pn, int n,
                           const double *stage, double *F, hipStream_t s) {
    return tile_F_t<double>(leaves, nleaves, max_pn, n, stage, F, s);
}
hipError_t hmx_even_tile_F(const ELeaf *leaves, int64_t nleaves, int64_t max_pn, int n,
                           const float *stage, float *F, hipStream_t s) {
    return tile_F_t<float>(leaves, nleaves, max_pn, n, stage, F, s);
}

hipError_t hmx_even_matvec(const EvenDev<double> &p, double *u, const double *v, int64_t ioff,
                           int64_t joff, int64_t incu, int64_t incv, int64_t ldu, int64_t ldv,
                           int nq, int beta, hipStream_t s) {
    return matvec_dispatch<double>(p, u, v, ioff, joff, incu, incv, ldu, ldv, nq, beta, s);
}
hipError_t hmx_even_matvec(const EvenDev<float> &p, float *u, const float *v, int64_t ioff,
                           int64_t joff, int64_t incu, int64_t incv, int64_t ldu, int64_t ldv,
                           int nq, int beta, hipStream_t s) {
    return matvec_dispatch<float>(p, u, v, ioff, joff, incu, incv, ldu, ldv, nq, beta, s);
}
