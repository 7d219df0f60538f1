// kernels_impl.h -- gfx950 kernels of the KernelMatrix compression and product
// (templates; instantiated per rank bucket by inst_r*.hip).
//
// One product y = beta*y + H x runs four passes over the flattened tree:
//   up         per task (a column cluster and <= 8192 of its columns), one
//              workgroup: w_task = B[cols, :]^T x[cols] reduced in registers,
//              across the wavefront by recursive halving and across the four
//              wavefronts in LDS; a cluster's only task writes w, with
//              several each task writes its own partial slot (no atomics)
//   reduce     per column cluster with several tasks: w = sum of its slots
//              (folded into the translation of single right-hand sides)
//   translate  per row cluster: z = sum over its far blocks of F w
//   down       per row chunk (one wavefront, one lane per row): near field of
//              the chunk's dense blocks plus U z of every row cluster covering
//              the chunk, written once as beta*y + acc
// The transpose runs the same passes with rows and columns exchanged.
//
// Three modes:
//   stored   streams A, U, V, F from HBM (U, V as update! forms them,
//            src/BarycentricMatrix.jl:248-287)
//   hybrid   streams A and F; the far-field basis is applied in Chebyshev form
//   free     streams nothing but points; A and F are regenerated as well
// Chebyshev form: on first-kind nodes x_k = cos((2k-1)pi/(2r)) the barycentric
// Lagrange basis equals l_k(x) = sum_m A[k][m] T_m(x) with
// A[k][m] = (m ? 2 : 1) T_m(x_k) / r (discrete orthogonality). So the upward
// pass accumulates moments mu_m = sum_j T_m(xh_j) x_j by the three-term
// recurrence, the reduce maps them to node values w = A mu, the translate
// maps z to coefficients c = A^T z, and the downward pass evaluates
// sum_m c_m T_m(xh_i) by Clenshaw: two FMAs per term, no division. Values
// equal the barycentric ones up to rounding.
//
// Every per-wavefront quantity (chunk descriptors, block and cluster records,
// the input vector entries and the z coefficients a lane multiplies with) is
// addressed through wave-uniform indices, so the compiler issues scalar loads
// for it and the vector memory path only carries the streamed factors.
//
// Compression kernels follow the reference's operation order exactly and are
// compiled without FMA contraction, so the stored factors are bit-identical to
// the ones the reference computes from the same points and nodes.
#pragma once
#include "hmx_device.h"

#include <math.h>

#include <type_traits>

namespace hmxk {
namespace {  // internal linkage: every translation unit keeps its own copies

constexpr int WAVE = 64;
constexpr int WG = 256;
constexpr int WPB = WG / WAVE;  // wavefronts per workgroup

__device__ __forceinline__ int uniform(int v) { return __builtin_amdgcn_readfirstlane(v); }

__device__ __forceinline__ int wave_id() {
    return uniform((int)(blockIdx.x * WPB + threadIdx.x / WAVE));
}

// wave-uniform copies of lane l's value (l uniform)
__device__ __forceinline__ int rl_i(int v, int l) { return __builtin_amdgcn_readlane(v, l); }
__device__ __forceinline__ int64_t rl_l(int64_t v, int l) {
    const int lo = __builtin_amdgcn_readlane((int)(uint32_t)(uint64_t)v, l);
    const int hi = __builtin_amdgcn_readlane((int)(uint32_t)((uint64_t)v >> 32), l);
    return (int64_t)(((uint64_t)(uint32_t)hi << 32) | (uint64_t)(uint32_t)lo);
}
__device__ __forceinline__ double rl_d(double v, int l) {
    return __longlong_as_double(rl_l(__double_as_longlong(v), l));
}

// Scalar helpers for the two element types (FP64 plans and FP32 plans).
template <class T> struct V2Of;
template <> struct V2Of<double> { using type = double2; };
template <> struct V2Of<float> { using type = float2; };
template <class T> using V2 = typename V2Of<T>::type;

__device__ __forceinline__ double fmaT(double a, double b, double c) { return __builtin_fma(a, b, c); }
__device__ __forceinline__ float fmaT(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
__device__ __forceinline__ double logT(double v) { return log(v); }
__device__ __forceinline__ float logT(float v) { return logf(v); }
__device__ __forceinline__ double fabsT(double v) { return fabs(v); }
__device__ __forceinline__ float fabsT(float v) { return fabsf(v); }

// sum over n slots p[t * stride], t = 0, 1, ... in that order; the loads of
// eight slots are issued before their additions
template <class T>
__device__ __forceinline__ T sum_slots(const T *p, int64_t stride, int n) {
    T s = T(0);
    int t = 0;
    for (; t + 8 <= n; t += 8) {
        T v[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) v[u] = p[(int64_t)(t + u) * stride];
#pragma unroll
        for (int u = 0; u < 8; ++u) s += v[u];
    }
    for (; t < n; ++t) s += p[(int64_t)t * stride];
    return s;
}

// v_m = sum_k A[m][k] mu_k, k ascending, where lane base + k holds mu_k: the
// node values of one cluster from its moments (A's row loaded eight entries
// at a time, so the loads do not wait on one another)
template <class T>
__device__ __forceinline__ T cheb_nodes(const T *chebA, T mu, int r, int m, int base) {
    const int mm = m < r ? m : 0;
    T v = T(0);
    for (int k0 = 0; k0 < r; k0 += 8) {
        T a[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) a[u] = chebA[mm * r + min(k0 + u, r - 1)];
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            const T mk = __shfl(mu, (base + min(k0 + u, r - 1)) & (WAVE - 1));
            if (k0 + u < r) v = fmaT(a[u], mk, v);
        }
    }
    return v;
}

// Reciprocal: hardware estimate refined by Newton-Raphson steps
// r <- r + r(1 - d r) (Markstein's FMA formulation), two in FP64, one in
// FP32; error <= 1 ulp.
__device__ __forceinline__ double rcpT(double d) {
    double r = __builtin_amdgcn_rcp(d);
    double e = __builtin_fma(-d, r, 1.0);
    r = __builtin_fma(r, e, r);
    e = __builtin_fma(-d, r, 1.0);
    return __builtin_fma(r, e, r);
}
__device__ __forceinline__ float rcpT(float d) {
    const float r = __builtin_amdgcn_rcpf(d);
    const float e = __builtin_fmaf(-d, r, 1.0f);
    return __builtin_fmaf(r, e, r);
}

// f(x, y) with the reference's rounding sequence (examples/Kernel.jl:34-37).
template <class T>
__device__ __forceinline__ T kern_exact(int kid, T x, T y) {
#pragma clang fp contract(off)
    T d = x - y;
    switch (kid) {
    case 0: return T(1) / d;
    case 1: return T(1) / (d * d);
    case 2: return T(1) / ((d * d) * d);
    default: return logT(fabsT(d));
    }
}

// f(x, y) for the matrix-free passes.
template <class T, int KID>
__device__ __forceinline__ T kern_fast(T x, T y) {
    T d = x - y;
    if constexpr (KID == 0) return rcpT(d);
    else if constexpr (KID == 1) return rcpT(d * d);
    else if constexpr (KID == 2) return rcpT(d * d * d);
    else return logT(fabsT(d));
}

// Value f(x, y) and derivative df/dx of the kernels, from one reciprocal of
// d = x - y: Cauchy 1/d, -1/d^2; Coulomb 1/d^2, -2/d^3; 1/d^3, -3/d^4;
// log|d|, 1/d. All are functions of x - y, so df/dy = -df/dx.
template <class T, int KID>
__device__ __forceinline__ void kern_val_dx(T d, T &f, T &fx) {
    const T r = rcpT(d);
    if constexpr (KID == 0) { f = r; fx = -r * r; }
    else if constexpr (KID == 1) { const T r2 = r * r; f = r2; fx = -T(2) * r2 * r; }
    else if constexpr (KID == 2) { const T r2 = r * r; f = r2 * r; fx = -T(3) * r2 * r2; }
    else { f = logT(fabsT(d)); fx = r; }
}

// ---------------------------------------------------------------- reductions
// Recursive-halving reduction of P values held by every lane: at each xor
// offset O the lane pair exchanges half of its values, so after log2 steps
// every lane holds P/2^s partial sums of distinct indices. Odd counts finish
// with plain xor-reductions. `base` receives the first index a lane holds.
// Only offsets >= OMIN take part, so lanes that differ in the low bits (the
// right-hand side index of the multi-RHS passes) keep separate sums.
template <int P, int O, int OMIN = 1>
struct Halve {
    template <class T>
    static __device__ __forceinline__ void run(T *acc, int lane, int &base) {
        if constexpr (O >= OMIN && O > 0) {
            if constexpr (P % 2 == 0) {
                constexpr int H = P / 2;
                const bool hi = (lane & O) != 0;
#pragma unroll
                for (int i = 0; i < H; ++i) {
                    T send = hi ? acc[i] : acc[i + H];
                    T keep = hi ? acc[i + H] : acc[i];
                    acc[i] = keep + __shfl_xor(send, O);
                }
                if (hi) base += H;
                Halve<H, O / 2, OMIN>::run(acc, lane, base);
            } else {
#pragma unroll
                for (int i = 0; i < P; ++i) acc[i] += __shfl_xor(acc[i], O);
                Halve<P, O / 2, OMIN>::run(acc, lane, base);
            }
        }
    }
};
template <int P, int O, int OMIN = 1>
struct HalveShape {  // values left per lane, and the xor-mask of the plain steps
    static constexpr bool stop = O < OMIN || O == 0;
    static constexpr int left = stop ? P : (P % 2 == 0 ? HalveShape<P / 2, O / 2, OMIN>::left
                                                       : HalveShape<P, O / 2, OMIN>::left);
    static constexpr int mask = stop ? 0 : (P % 2 == 0 ? HalveShape<P / 2, O / 2, OMIN>::mask
                                                       : (O | HalveShape<P, O / 2, OMIN>::mask));
};
template <int P, int OMIN>
struct HalveShape<P, 0, OMIN> {
    static constexpr int left = P;
    static constexpr int mask = 0;
};

// ------------------------------------------------------------- compression

template <class T>
__global__ __launch_bounds__(WG) void k_gen_dense(DevPlanT<T> p) {
    const DDense D = p.dense[blockIdx.x];
    const int32_t tot = D.m * D.n;
    for (int32_t e = threadIdx.x; e < tot; e += WG) {
        int32_t j = e / D.m, i = e - j * D.m;
        p.A[D.aoff + tile_index(i, j, D.n)] = kern_exact<T>(p.kid, p.x[D.row0 + i], p.y[D.col0 + j]);
    }
}

// update! (src/BarycentricMatrix.jl:256-287): one lane per point of a chunk,
// for every cluster covering the chunk.
template <class T, int RMAX>
__global__ __launch_bounds__(WG) void k_gen_basis(const DChunk *ch, int nch, const DFar *fl,
                                                 const DCluster *cl, const T *nodes,
                                                 const T *lam, const T *pts,
                                                 T *B, int r) {
#pragma clang fp contract(off)
    const int wid = wave_id();
    const int lane = threadIdx.x % WAVE;
    if (wid >= nch) return;
    const DChunk c = ch[wid];
    if (lane >= c.np) return;
    const T xi = pts[c.p0 + lane];
    for (int e = c.far_beg; e < c.far_end; ++e) {
        const DFar f = fl[e];
        const DCluster C = cl[f.cl];
        const T *t = nodes + (int64_t)f.cl * r;
        T row[RMAX];
        T s = T(0);
#pragma unroll
        for (int k = 0; k < RMAX; ++k)
            if (k < r) row[k] = lam[k] * (T(1) / (xi - t[k]));
#pragma unroll
        for (int k = 0; k < RMAX; ++k)
            if (k < r) s += row[k];
        T *out = B + C.boff + tile_index(f.off + lane, 0, r);
#pragma unroll
        for (int k = 0; k < RMAX; ++k)
            if (k < r) out[k * HM_TILE] = row[k] / s;
    }
}

// F[m, n] = f(t_m, s_n), src/BarycentricMatrix.jl:168-175
template <class T>
__global__ __launch_bounds__(WG) void k_gen_F(DevPlanT<T> p) {
    const DFarBlk b = p.far[blockIdx.x];
    const int r = p.r;
    const T *tx = p.rnodes + (int64_t)b.rc * r;
    const T *ty = p.cnodes + (int64_t)b.cc * r;
    for (int e = threadIdx.x; e < r * r; e += WG) {
        int n = e / r, m = e - n * r;
        p.F[b.foff + e] = kern_exact<T>(p.kid, tx[m], ty[n]);
    }
}

template <class T>
__global__ __launch_bounds__(WG) void k_count_nonfinite(const T *v, int64_t n,
                                                        unsigned long long *blockcount) {
    __shared__ unsigned int sh[WG];
    unsigned int cnt = 0;
    for (int64_t i = (int64_t)blockIdx.x * WG + threadIdx.x; i < n; i += (int64_t)gridDim.x * WG)
        cnt += isfinite(v[i]) ? 0u : 1u;
    sh[threadIdx.x] = cnt;
    __syncthreads();
    for (int o = WG / 2; o > 0; o /= 2) {
        if ((int)threadIdx.x < o) sh[threadIdx.x] += sh[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) blockcount[blockIdx.x] = sh[0];
}

// ------------------------------------------------------------------ product
// Multi-RHS layout: KB right-hand sides per launch. Input element (j, q) is
// x[j*incx + q*ldx], output (i, q) is y[i*incy + q*ldy]; the workspace keeps
// the KB vectors of a slot or cluster together: partial[(slot*KB + q)*r + m],
// w and z[(cluster*KB + q)*r + m].

// Upward pass: partial[task.slot] = B[task points, :]^T x[task points] with B =
// V (forward) or U (transpose) read from HBM, or, in Chebyshev form, the
// moments sum_j T_m(xh_j) x_j with xh = (p - center) * inv_halfwidth. A lane
// works on one (point, right-hand side) pair: lane = point*KB + q.
// A cluster's only task writes its cluster's w itself (in Chebyshev form
// w = A mu, the reduce's arithmetic in the reduce's order); the reduce then
// runs over the clusters with several tasks only.
template <class T, int RMAX, bool CHEB, int KB>
__global__ __launch_bounds__(WG) void k_up(const DTask *tasks, const DCluster *cl, const T *B,
                                          const V2<T> *cmapv, const T *pts, const T *x,
                                          int64_t incx, int64_t ldx, T *partial, T *w,
                                          const T *chebA, int r) {
    constexpr int PP = WAVE / KB;  // points per wavefront step
    __shared__ T sh[WPB][KB][RMAX];
    __shared__ T mus[KB][RMAX];
    const DTask Tk = tasks[blockIdx.x];
    const DCluster C = cl[Tk.cl];
    const int tid = threadIdx.x, lane = tid % WAVE, wv = tid / WAVE;
    const int q = lane % KB, pp = lane / KB;
    T acc[RMAX];
#pragma unroll
    for (int k = 0; k < RMAX; ++k) acc[k] = T(0);
    const int64_t first = C.first + Tk.off;
    const T *xq = x + (int64_t)q * ldx;
    if constexpr (!CHEB) {
        for (int i = wv * PP + pp; i < Tk.len; i += WPB * PP) {
            const T xv = xq[(first + i) * incx];
            const T *bp = B + C.boff + tile_index(Tk.off + i, 0, r);
            // columns past r re-read column r-1 (in bounds) and are dropped,
            // so all RMAX loads issue back to back without a branch
            T bv[RMAX];
#pragma unroll
            for (int k = 0; k < RMAX; ++k) bv[k] = __builtin_nontemporal_load(bp + min(k, r - 1) * HM_TILE);
#pragma unroll
            for (int k = 0; k < RMAX; ++k) acc[k] = fmaT(k < r ? bv[k] : T(0), xv, acc[k]);
        }
    } else {
        const V2<T> cm = cmapv[Tk.cl];
        for (int i = wv * PP + pp; i < Tk.len; i += WPB * PP) {
            const T xv = xq[(first + i) * incx];
            const T xh = (pts[first + i] - cm.x) * cm.y;
            const T x2 = xh + xh;
            T t0 = xv, t1 = xh * xv;  // T_m(xh) * x, by the recurrence
            acc[0] += t0;
            if (RMAX > 1) acc[1] += t1;
#pragma unroll
            for (int k = 2; k < RMAX; ++k) {
                const T t2 = fmaT(x2, t1, -t0);
                acc[k] += t2;
                t0 = t1;
                t1 = t2;
            }
        }
    }
    int base = 0;
    Halve<RMAX, WAVE / 2, KB>::run(acc, lane, base);
    constexpr int LEFT = HalveShape<RMAX, WAVE / 2, KB>::left;
    constexpr int MASK = HalveShape<RMAX, WAVE / 2, KB>::mask;
    if ((lane & MASK) == 0) {
#pragma unroll
        for (int i = 0; i < LEFT; ++i) sh[wv][q][base + i] = acc[i];
    }
    __syncthreads();
    // KB * r reaches 512 (8 right-hand sides at rank 64): threads loop over it
    const bool single = Tk.single != 0;  // uniform over the workgroup
    for (int e = tid; e < KB * r; e += WG) {
        const int qq = e / r, m = e % r;
        T s = T(0);
#pragma unroll
        for (int wi = 0; wi < WPB; ++wi) s += sh[wi][qq][m];
        if (!single) partial[(Tk.slot * KB + qq) * r + m] = s;
        else if constexpr (CHEB) mus[qq][m] = s;
        else w[((int64_t)Tk.cl * KB + qq) * r + m] = s;
    }
    if constexpr (CHEB) {
        if (single) {
            __syncthreads();
            for (int e = tid; e < KB * r; e += WG) {
                const int qq = e / r, m = e % r;
                T v = T(0);
                for (int k = 0; k < r; ++k) v = fmaT(chebA[m * r + k], mus[qq][k], v);
                w[((int64_t)Tk.cl * KB + qq) * r + m] = v;
            }
        }
    }
}

// w[(c, q)] = sum over the task slots of cluster c, in slot order; in
// Chebyshev form the slots hold moments and w = A mu. One wavefront per
// cluster of the list ids: lane = g*r + m handles row m of right-hand sides
// q = g, g + G, ... (G = 64 / r groups).
template <class T, bool CHEB, int KB>
__global__ __launch_bounds__(WG) void k_reduce(const DCluster *cl, const int32_t *ids, int n,
                                              const T *partial, const T *chebA, T *w, int r) {
    const int wid = wave_id();
    const int lane = threadIdx.x % WAVE;
    if (wid >= n) return;
    const int cid = ids[wid];
    const DCluster C = cl[cid];
    const int G = WAVE / r;
    const int m = lane % r, g = lane / r;
    const bool act = g < G;
    for (int q0 = 0; q0 < KB; q0 += G) {
        const int q = q0 + g;
        const bool aq = act && q < KB;
        T acc = T(0);
        if (aq) acc = sum_slots(partial + (C.pbase * KB + q) * r + m, (int64_t)KB * r, C.nslots);
        if constexpr (CHEB) acc = cheb_nodes(chebA, acc, r, m, g * r);
        if (aq) w[((int64_t)cid * KB + q) * r + m] = acc;
    }
}

// z[(d, q)] = sum over far blocks b of cluster d of F_b w[(other side, q)]
// (forward: F w[cc]; transpose: F^T w[rc]); in Chebyshev form z is then
// mapped to coefficients c = A^T z. One wavefront per cluster: lane = g*r + m
// owns row m of F for right-hand sides q = g, g + G, ...
template <class T, int RMAX, bool FREE, int KID, bool TRANS, bool CHEB, int KB>
__global__ __launch_bounds__(WG) void k_translate(DevPlanT<T> p, const DCluster *cl, int ncl,
                                                 const int32_t *far_list, const T *w,
                                                 T *z) {
    const int wid = wave_id();
    const int lane = threadIdx.x % WAVE;
    const int r = p.r;
    const int G = WAVE / r;
    const int m = lane % r, g = lane / r;
    const bool act = g < G;
    // one right-hand side: the G lane groups of a wavefront take G clusters;
    // several: one cluster per wavefront, its groups taking the right-hand sides
    const int ci = KB == 1 ? wid * G + g : wid;
    if (KB == 1 ? (!act || ci >= ncl) : wid >= ncl) return;
    const DCluster C = cl[ci];
    for (int q0 = 0; q0 < KB; q0 += (KB == 1 ? 1 : G)) {
        const int q = KB == 1 ? 0 : q0 + g;
        const bool aq = KB == 1 || (act && q < KB);
        const int qa = aq ? q : 0;
        T acc = T(0);
        // stored F: two blocks' loads in flight; regenerated F: no unrolling
        constexpr int UNR = FREE ? 1 : 2;
#pragma unroll UNR
        for (int e = C.far_beg; e < C.far_end; ++e) {
            const DTrRec b = (TRANS ? p.tr_cc : p.tr_rc)[e];
            const int src = b.src;
            const T *ws = w + ((int64_t)src * KB + qa) * r;
            T fv[RMAX];
            if constexpr (!FREE) {
                const T *F = p.F + b.foff;
#pragma unroll
                for (int l = 0; l < RMAX; ++l) {
                    const int lc = min(l, r - 1);
                    fv[l] = __builtin_nontemporal_load(TRANS ? F + (lc + r * m) : F + (m + r * lc));
                }
            } else {
                // forward: F[m, l] = f(tx_m, ty_l); transpose: lane row m is F's column
                const T *own = (TRANS ? p.cnodes : p.rnodes) + (int64_t)ci * r;
                const T *oth = (TRANS ? p.rnodes : p.cnodes) + (int64_t)src * r;
                const T to = own[m];
#pragma unroll
                for (int l = 0; l < RMAX; ++l) {
                    const T t = oth[min(l, r - 1)];
                    fv[l] = TRANS ? kern_fast<T, KID>(t, to) : kern_fast<T, KID>(to, t);
                }
            }
#pragma unroll
            for (int l = 0; l < RMAX; ++l)
                acc = fmaT(l < r ? fv[l] : T(0), ws[min(l, r - 1)], acc);
        }
        if constexpr (CHEB) {  // coefficients c = A^T z
            T c = T(0);
            for (int k = 0; k < r; ++k) {
                const T zk = __shfl(acc, (g * r + k) & (WAVE - 1));
                c = fmaT(p.chebA[k * r + m], zk, c);
            }
            acc = c;
        }
        if (aq) z[((int64_t)ci * KB + q) * r + m] = acc;
    }
}

// Translation of one right-hand side with stored F, one wavefront per
// cluster: lane (m, h) = (lane % MW, lane / MW) sums F[m, l] w_l over the h-th
// part of the block's columns l (NH parts of LH), the parts added with a
// shuffle, then (Chebyshev form) c = A^T z. The cluster's block records are
// read first, one per lane, so the F loads of a block issue without waiting
// on its record; two blocks' loads are in flight, each lane holding LH of a
// block's F values (12 at rank 24) and lane l its source cluster's w_l (shared
// by shuffles). A source cluster with several upward tasks is summed here from
// its partial slots in slot order (and, in Chebyshev form, mapped by A), the
// reduce's arithmetic in the reduce's order, so no reduce runs before this
// pass. Few registers keep many wavefronts resident to cover the memory
// latency. FREE: F is regenerated from the two clusters' mapped nodes
// (F[m, l] = f(tx_m, ty_l) forward) instead of loaded.
template <class T, int RMAX, bool TRANS, bool CHEB, bool FREE = false, int KID = 0>
__global__ __launch_bounds__(WG) void k_translate1(DevPlanT<T> p, const DCluster *cl, int ncl,
                                                  const T *partial, const T *w, T *z) {
    constexpr int NH = RMAX <= 32 ? 2 : 1;  // column parts
    constexpr int MW = WAVE / NH;           // lanes per part
    constexpr int LHMAX = (RMAX + NH - 1) / NH;
    const int wid = wave_id();
    const int lane = threadIdx.x % WAVE;
    if (wid >= ncl) return;
    const int r = p.r;
    const int m = lane % MW, h = lane / MW;
    const int LH = (r + NH - 1) / NH;
    const int l0 = h * LH, l1 = min(r, l0 + LH);
    const bool act = m < r;
    const int mm = act ? m : 0;
    const int lw = min(lane, r - 1);
    const DCluster C = cl[wid];
    const DTrRec *recs = TRANS ? p.tr_cc : p.tr_rc;
    T acc = T(0);
    for (int e0 = C.far_beg; e0 < C.far_end; e0 += WAVE) {
        const int ne = min(WAVE, C.far_end - e0);
        int64_t b_foff = 0, b_pbase = 0;
        int b_src = 0, b_ns = 1;
        if (lane < ne) {
            const DTrRec b = recs[e0 + lane];
            b_foff = b.foff;
            b_pbase = b.pbase;
            b_src = b.src;
            b_ns = b.nslots;
        }
        auto load = [&](int e, T *fv, T &wl) {
            const int ns = rl_i(b_ns, e);
            if (ns > 1) {  // a source with several tasks (wave-uniform branch)
                wl = sum_slots(partial + rl_l(b_pbase, e) * r + lw, (int64_t)r, ns);
                if constexpr (CHEB) wl = cheb_nodes(p.chebA, wl, r, lw, 0);
            } else {
                wl = w[(int64_t)rl_i(b_src, e) * r + lw];
            }
            if constexpr (FREE) {
                // forward: F[m, l] = f(tx_m, ty_l); transpose: lane row m is F's column
                const T to = (TRANS ? p.cnodes : p.rnodes)[(int64_t)wid * r + mm];
                const T *oth = (TRANS ? p.rnodes : p.cnodes) + (int64_t)rl_i(b_src, e) * r;
#pragma unroll
                for (int j = 0; j < LHMAX; ++j) {
                    const T t = oth[min(l0 + j, r - 1)];
                    fv[j] = TRANS ? kern_fast<T, KID>(t, to) : kern_fast<T, KID>(to, t);
                }
            } else {
                const T *F = p.F + rl_l(b_foff, e);
#pragma unroll
                for (int j = 0; j < LHMAX; ++j) {
                    const int l = min(l0 + j, r - 1);  // in bounds; the extra terms are dropped
                    fv[j] = __builtin_nontemporal_load(TRANS ? F + (l + r * mm) : F + (mm + r * l));
                }
            }
        };
        auto apply = [&](const T *fv, T wl) {
#pragma unroll
            for (int j = 0; j < LHMAX; ++j) {
                const T wv = __shfl(wl, min(l0 + j, r - 1));
                if (l0 + j < l1) acc = fmaT(fv[j], wv, acc);
            }
        };
        int e = 0;
        for (; e + 2 <= ne; e += 2) {
            T fa[LHMAX], fb[LHMAX], wa, wb;
            load(e, fa, wa);
            load(e + 1, fb, wb);
            apply(fa, wa);
            apply(fb, wb);
        }
        if (e < ne) {
            T fa[LHMAX], wa;
            load(e, fa, wa);
            apply(fa, wa);
        }
    }
    if constexpr (NH == 2) acc += __shfl_xor(acc, MW);
    if constexpr (CHEB) {  // coefficients c = A^T z, A's column loaded eight rows at a time
        T c = T(0);
#pragma unroll
        for (int k0 = 0; k0 < RMAX; k0 += 8) {
            T a[8];
#pragma unroll
            for (int u = 0; u < 8; ++u) a[u] = p.chebA[min(k0 + u, r - 1) * r + mm];
#pragma unroll
            for (int u = 0; u < 8; ++u) {
                const T zk = __shfl(acc, k0 + u);
                if (k0 + u < r) c = fmaT(a[u], zk, c);
            }
        }
        acc = c;
    }
    if (act && h == 0) z[(int64_t)wid * r + m] = acc;
}

// Matrix-free near field of one lane and one right-hand side: sum_j f(p, q_j)
// v_j (or f(q_j, p) v_j when TR). For the power kernels 1/d^m two terms share
// one reciprocal: v1/d1^m + v2/d2^m = (v1 d2^m + v2 d1^m) / (d1 d2)^m. The
// block's columns are staged through LDS 64 at a time (one coalesced load per
// lane of the points and of v, then broadcast reads; the pairs start at even
// columns, as over the whole block). sq and sx: this wavefront's 64-entry LDS
// tiles.
template <class T, int KID, bool TR>
__device__ __forceinline__ T near_free_staged(T p, const T *qs, const T *v, int64_t inc, int n,
                                              bool act, T *sq, T *sx, int lane) {
    T acc = T(0);
    for (int jt = 0; jt < n; jt += WAVE) {
        const int nj = min(WAVE, n - jt);
        if (lane < nj) {
            sq[lane] = qs[jt + lane];
            sx[lane] = v[(int64_t)(jt + lane) * inc];
        }
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
        if constexpr (KID <= 2) {
            int j = 0;
            for (; j + 2 <= nj; j += 2) {
                T d1 = TR ? sq[j] - p : p - sq[j];
                T d2 = TR ? sq[j + 1] - p : p - sq[j + 1];
                if constexpr (KID == 1) { d1 *= d1; d2 *= d2; }
                if constexpr (KID == 2) { d1 *= d1 * d1; d2 *= d2 * d2; }
                const T num = fmaT(sx[j], d2, sx[j + 1] * d1);
                acc = fmaT(num, rcpT(d1 * d2), acc);
            }
            if (j < nj) {
                const T f = TR ? kern_fast<T, KID>(sq[j], p) : kern_fast<T, KID>(p, sq[j]);
                acc = fmaT(f, sx[j], acc);
            }
        } else {
            for (int j = 0; j < nj; ++j) {
                const T f = TR ? kern_fast<T, KID>(sq[j], p) : kern_fast<T, KID>(p, sq[j]);
                acc = fmaT(f, sx[j], acc);
            }
        }
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
    }
    return act ? acc : T(0);
}


// Downward pass. Forward: output = rows; near blocks read by rows; far basis U
// (or its Chebyshev form). Transpose (KB = 1 only): output = columns; near
// blocks read by columns; far basis V. With one right-hand side the stored
// blocks and bases are read with non-temporal loads (the streaming hint keeps
// them from displacing the vectors and coefficients the caches hold; it took
// this pass from 6.15 to 6.77 TB/s at n = 2^20); with several per launch the
// hint measured slower and the loads are plain.
template <class T, int RMAX, bool FREE, int KID, bool TRANS, bool CHEB, int KB>
__global__ __launch_bounds__(WG) void k_down(DevPlanT<T> p, const DChunk *ch, int nch,
                                            const DNear *nl, const DFar *fl,
                                            const DCluster *cl, const T *B,
                                            const V2<T> *cmapv, const T *pts_out,
                                            const T *pts_in, const T *z, MvArgsT<T> a) {
    static_assert(!TRANS || KB == 1, "the transpose handles one right-hand side per pass");
    constexpr int UN = KB == 1 ? 16 : 8;
    // per-wavefront staging of right-hand-side tiles and far coefficients
    constexpr int XS = KB == 1 ? (FREE ? 2 * WAVE : 1) : (WAVE * KB > KB * RMAX ? WAVE * KB : KB * RMAX);
    __shared__ T xs_all[WPB][XS];
    const int wid = wave_id();
    const int lane = threadIdx.x % WAVE;
    T *xs = xs_all[(threadIdx.x / WAVE) % WPB];
    (void)xs;
    if (wid >= nch) return;
    const int r = p.r;
    const DChunk c = ch[wid];
    const bool act = lane < c.np;
    T po = T(0);
    if constexpr (FREE || CHEB) po = act ? pts_out[c.p0 + lane] : T(0);
    T acc[KB];
#pragma unroll
    for (int q = 0; q < KB; ++q) acc[q] = T(0);

    // ---- near field
    const int near_end = a.part >= 2 ? c.near_beg : c.near_end;
    for (int e = c.near_beg; e < near_end; ++e) {
        const DNear s = nl[e];
        const DDense D = p.dense[s.blk];
        if constexpr (!TRANS) {
            const T *xin = a.x + D.col0 * a.incx;
            if constexpr (!FREE) {
                const T *Aj = p.A + D.aoff + tile_index(s.off + lane, 0, D.n);
                if constexpr (KB == 1) {
                    int j = 0;
                    for (; j + UN <= D.n; j += UN) {
                        T av[UN];
#pragma unroll
                        for (int u = 0; u < UN; ++u)
                            av[u] = act ? __builtin_nontemporal_load(Aj + (j + u) * HM_TILE) : T(0);
#pragma unroll
                        for (int u = 0; u < UN; ++u)
                            acc[0] = fmaT(av[u], xin[(int64_t)(j + u) * a.incx], acc[0]);
                    }
                    for (; j < D.n; ++j)
                        acc[0] = fmaT(act ? __builtin_nontemporal_load(Aj + j * HM_TILE) : T(0),
                                      xin[(int64_t)j * a.incx], acc[0]);
                } else {
                    for (int jt = 0; jt < D.n; jt += WAVE) {
                        const int nj = min(WAVE, D.n - jt);
                        if (lane < nj) {
#pragma unroll
                            for (int q = 0; q < KB; ++q)
                                xs[lane * KB + q] = xin[(int64_t)(jt + lane) * a.incx + q * a.ldx];
                        }
                        __builtin_amdgcn_wave_barrier();
                        __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
                        int j = 0;
                        for (; j + UN <= nj; j += UN) {
                            T av[UN];
#pragma unroll
                            for (int u = 0; u < UN; ++u) av[u] = act ? Aj[(jt + j + u) * HM_TILE] : T(0);
#pragma unroll
                            for (int u = 0; u < UN; ++u)
#pragma unroll
                                for (int q = 0; q < KB; ++q)
                                    acc[q] = fmaT(av[u], xs[(j + u) * KB + q], acc[q]);
                        }
                        for (; j < nj; ++j) {
                            const T av = act ? Aj[(jt + j) * HM_TILE] : T(0);
#pragma unroll
                            for (int q = 0; q < KB; ++q) acc[q] = fmaT(av, xs[j * KB + q], acc[q]);
                        }
                        __builtin_amdgcn_wave_barrier();
                        __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
                    }
                }
            } else if constexpr (KB == 1) {
                acc[0] += near_free_staged<T, KID, false>(po, pts_in + D.col0, xin, a.incx, D.n, act,
                                                          xs, xs + WAVE, lane);
            } else {
                // right-hand sides of a 64-column tile staged in LDS, read as
                // broadcasts; one kernel value per (row, column)
                const T *yin = pts_in + D.col0;
                for (int jt = 0; jt < D.n; jt += WAVE) {
                    const int nj = min(WAVE, D.n - jt);
                    if (lane < nj) {
#pragma unroll
                        for (int q = 0; q < KB; ++q)
                            xs[lane * KB + q] = xin[(int64_t)(jt + lane) * a.incx + q * a.ldx];
                    }
                    __builtin_amdgcn_wave_barrier();
                    __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
                    for (int j = 0; j < nj; ++j) {
                        const T f = act ? kern_fast<T, KID>(po, yin[jt + j]) : T(0);
#pragma unroll
                        for (int q = 0; q < KB; ++q) acc[q] = fmaT(f, xs[j * KB + q], acc[q]);
                    }
                    __builtin_amdgcn_wave_barrier();
                    __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
                }
            }
                } else {
            const T *xin = a.x + D.row0 * a.incx;
            if constexpr (!FREE) {
                // lanes run down the rows of the block, 64 accumulators for the
                // chunk's columns, then a recursive-halving transpose-reduce
                T accT[WAVE];
#pragma unroll
                for (int q = 0; q < WAVE; ++q) accT[q] = T(0);
                for (int ib = 0; ib < D.m; ib += WAVE) {
                    const bool ra = ib + lane < D.m;
                    const T xv = ra ? xin[(int64_t)(ib + lane) * a.incx] : T(0);
                    const T *Ac = p.A + D.aoff + tile_index(ib + lane, s.off, D.n);
#pragma unroll
                    for (int q = 0; q < WAVE; ++q)
                        if (q < c.np) accT[q] += (ra ? __builtin_nontemporal_load(Ac + q * HM_TILE) : T(0)) * xv;
                }
                int base = 0;
                Halve<WAVE, WAVE / 2>::run(accT, lane, base);
                acc[0] += accT[0];
            } else {
                acc[0] += near_free_staged<T, KID, true>(po, pts_in + D.row0, xin, a.incx, D.m, act,
                                                         xs, xs + WAVE, lane);
            }
        }
    }

    // ---- far field: acc += B[chunk rows, :] z[cluster]
    const int far_end = a.part == 1 ? c.far_beg : c.far_end;
    for (int e = c.far_beg; e < far_end; ++e) {
        const DFar f = fl[e];
        const T *zc = z + (int64_t)f.cl * KB * r;
        if constexpr (!CHEB) {
            const DCluster C = cl[f.cl];
            const T *bp = B + C.boff + tile_index(f.off + lane, 0, r);
            const T *zs = zc;
            if constexpr (KB > 1) {
                for (int idx = lane; idx < KB * r; idx += WAVE) xs[idx] = zc[idx];
                __builtin_amdgcn_wave_barrier();
                __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
                zs = xs;
            }
            int k = 0;
            for (; k + 8 <= r; k += 8) {
                T bv[8];
#pragma unroll
                for (int u = 0; u < 8; ++u)
                    bv[u] = act ? __builtin_nontemporal_load(bp + (k + u) * HM_TILE) : T(0);
#pragma unroll
                for (int u = 0; u < 8; ++u)
#pragma unroll
                    for (int q = 0; q < KB; ++q) acc[q] = fmaT(bv[u], zs[q * r + k + u], acc[q]);
            }
            for (; k < r; ++k) {
                const T bv = act ? __builtin_nontemporal_load(bp + k * HM_TILE) : T(0);
#pragma unroll
                for (int q = 0; q < KB; ++q) acc[q] = fmaT(bv, zs[q * r + k], acc[q]);
            }
            if constexpr (KB > 1) {
                __builtin_amdgcn_wave_barrier();
                __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
            }
        } else {
            const V2<T> cm = cmapv[f.cl];
            const T xh = (po - cm.x) * cm.y;
            const T x2 = xh + xh;
            if constexpr (KB == 1) {
                // Clenshaw: b_k = c_k + 2 xh b_{k+1} - b_{k+2}; sum = c_0 + xh b_1 - b_2
                T b1 = T(0), b2 = T(0);
#pragma unroll
                for (int k = RMAX - 1; k >= 1; --k) {
                    const T ck = k < r ? zc[min(k, r - 1)] : T(0);
                    const T b0 = fmaT(x2, b1, ck - b2);
                    b2 = b1;
                    b1 = b0;
                }
                const T sum = fmaT(xh, b1, zc[0] - b2);
                if (act) acc[0] += sum;
            } else {
                // coefficients of the cluster (KB x r) staged in LDS; T_k(xh) once,
                // then one dot product per right-hand side
                for (int idx = lane; idx < KB * r; idx += WAVE) xs[idx] = zc[idx];
                __builtin_amdgcn_wave_barrier();
                __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
                T t0 = T(1), t1 = xh;
#pragma unroll
                for (int k = 0; k < RMAX; ++k) {
                    const T tk = k == 0 ? t0 : (k == 1 ? t1 : T(0));
                    T tv = tk;
                    if (k >= 2) {
                        tv = fmaT(x2, t1, -t0);
                        t0 = t1;
                        t1 = tv;
                    }
                    const T tu = (act && k < r) ? tv : T(0);
#pragma unroll
                    for (int q = 0; q < KB; ++q) acc[q] = fmaT(tu, xs[q * r + min(k, r - 1)], acc[q]);
                }
                __builtin_amdgcn_wave_barrier();
                __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
            }
        }
    }
    if (act) {
        const bool keep = a.beta || a.part == 2;
#pragma unroll
        for (int q = 0; q < KB; ++q) {
            T *yo = a.y + (c.p0 + lane) * a.incy + q * a.ldy;
            *yo = (keep ? *yo : T(0)) + acc[q];
        }
    }
}

// Downward pass of the gradients. With p the lane's output point (x_i forward,
// y_j transposed), the chunk computes
//   val_p = sum_q f(.,.) in_q + sum_clusters S_c(xh(p))        (the product)
//   der_p = sum_q d f/d p in_q + sum_clusters S_c'(xh(p)) / halfwidth_c
// where S_c is the Chebyshev series of the cluster's coefficients, so der_p is
// the exact derivative of val_p with respect to the point p. The near field is
// regenerated from the points. DM = 1 writes y = beta*y + s .* der; DM = 2
// writes y = beta*y + val and y2 = beta*y2 + s .* der (s null: ones).
// One chunk per workgroup: its WPB wavefronts take every WPB-th column of each
// near block and every WPB-th covering cluster (WPB times the wavefronts of a
// wavefront per chunk, whose column loop waits on its loads), and their sums
// are added in LDS in wavefront order.
template <class T, int RMAX, int KID, bool TRANS, int DM>
__global__ __launch_bounds__(WG) void k_down_grad(DevPlanT<T> p, const DChunk *ch, int nch,
                                                 const DNear *nl, const DFar *fl,
                                                 const V2<T> *cmapv, const T *pts_out,
                                                 const T *pts_in, const T *z, MvArgsT<T> a,
                                                 T *y2, int64_t incy2, const T *sv,
                                                 int64_t incs) {
    __shared__ T red[2][WPB][WAVE];
    __shared__ T stq[WPB][2][WAVE];  // a wavefront's columns: points, then x
    const int wv = uniform((int)(threadIdx.x / WAVE));
    const int lane = threadIdx.x % WAVE;
    if ((int)blockIdx.x >= nch) return;  // workgroup-uniform
    const int r = p.r;
    const DChunk c = ch[blockIdx.x];
    const bool act = lane < c.np;
    const T po = act ? pts_out[c.p0 + lane] : T(0);
    T accv = T(0), accd = T(0);
    for (int e = c.near_beg; e < c.near_end; ++e) {
        const DNear s = nl[e];
        const DDense D = p.dense[s.blk];
        const int64_t q0 = TRANS ? D.row0 : D.col0;
        const int nq = TRANS ? D.m : D.n;
        const T *xin = a.x + q0 * a.incx;
        const T *qin = pts_in + q0;
        // this wavefront's columns wv, wv + WPB, ..., staged 64 at a time
        // (one coalesced load per lane, then broadcast reads), in order
        const int nw = nq > wv ? (nq - wv + WPB - 1) / WPB : 0;
        for (int jt = 0; jt < nw; jt += WAVE) {
            const int nj = min(WAVE, nw - jt);
            if (lane < nj) {
                const int j = wv + (jt + lane) * WPB;
                stq[wv][0][lane] = qin[j];
                stq[wv][1][lane] = xin[(int64_t)j * a.incx];
            }
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
            for (int u = 0; u < nj; ++u) {
                const T qj = stq[wv][0][u];
                const T d = TRANS ? qj - po : po - qj;
                T f, fx;
                kern_val_dx<T, KID>(d, f, fx);
                const T v = stq[wv][1][u];
                if constexpr (DM == 2) accv = fmaT(f, v, accv);
                accd = fmaT(TRANS ? -fx : fx, v, accd);
            }
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
        }
    }
    for (int e = c.far_beg + wv; e < c.far_end; e += WPB) {
        const DFar f = fl[e];
        const T *zc = z + (int64_t)f.cl * r;
        const V2<T> cm = cmapv[f.cl];
        const T xh = (po - cm.x) * cm.y;
        const T x2 = xh + xh;
        T b1 = T(0), b2 = T(0), d1 = T(0), d2 = T(0);
#pragma unroll
        for (int k = RMAX - 1; k >= 1; --k) {
            const T ck = k < r ? zc[min(k, r - 1)] : T(0);
            const T b0 = fmaT(x2, b1, ck - b2);
            const T d0 = fmaT(x2, d1, T(2) * b1 - d2);
            b2 = b1; b1 = b0;
            d2 = d1; d1 = d0;
        }
        if constexpr (DM == 2) accv += fmaT(xh, b1, zc[0] - b2);
        accd = fmaT(fmaT(xh, d1, b1 - d2), cm.y, accd);
    }
    red[0][wv][lane] = accv;
    red[1][wv][lane] = accd;
    __syncthreads();
    if (wv != 0) return;
    accv = red[0][0][lane];
    accd = red[1][0][lane];
#pragma unroll
    for (int u = 1; u < WPB; ++u) {
        accv += red[0][u][lane];
        accd += red[1][u][lane];
    }
    if (act) {
        const int64_t i = c.p0 + lane;
        const T sc = sv ? sv[i * incs] : T(1);  // wave-uniform: no s means ones
        T *yo = a.y + i * a.incy;
        if constexpr (DM == 1) {
            *yo = (a.beta ? *yo : T(0)) + sc * accd;
        } else {
            *yo = (a.beta ? *yo : T(0)) + accv;
            T *yd = y2 + i * incy2;
            *yd = (a.beta ? *yd : T(0)) + sc * accd;
        }
    }
}

// getindex: one thread per requested entry
template <class T>
__global__ __launch_bounds__(WG) void k_entries(DevPlanT<T> p, int64_t n, const int64_t *ii,
                                               const int64_t *jj, const int32_t *kind,
                                               const int32_t *id, T *out) {
#pragma clang fp contract(off)
    const int64_t t = (int64_t)blockIdx.x * WG + threadIdx.x;
    if (t >= n) return;
    const int64_t i = ii[t], j = jj[t];
    const int r = p.r;
    T v = T(0);
    if (kind[t] == 1) {
        const DDense D = p.dense[id[t]];
        v = p.A ? p.A[D.aoff + tile_index(i - D.row0, j - D.col0, D.n)] : kern_exact<T>(p.kid, p.x[i], p.y[j]);
    } else if (kind[t] == 2) {
        const DFarBlk b = p.far[id[t]];
        const DCluster R = p.rcl[b.rc], Cc = p.ccl[b.cc];
        T ur[64], vr[64];
        if (p.U) {
            for (int k = 0; k < r; ++k) {
                ur[k] = p.U[R.boff + tile_index(i - R.first, k, r)];
                vr[k] = p.V[Cc.boff + tile_index(j - Cc.first, k, r)];
            }
        } else {
            T su = T(0), sv = T(0);
            for (int k = 0; k < r; ++k) {
                ur[k] = p.lam[k] * (T(1) / (p.x[i] - p.rnodes[(int64_t)b.rc * r + k]));
                vr[k] = p.lam[k] * (T(1) / (p.y[j] - p.cnodes[(int64_t)b.cc * r + k]));
                su += ur[k];
                sv += vr[k];
            }
            for (int k = 0; k < r; ++k) { ur[k] /= su; vr[k] /= sv; }
        }
        for (int k = 0; k < r; ++k) {
            T tmp = T(0);
            for (int l = 0; l < r; ++l) {
                T F = p.F ? p.F[b.foff + k + (int64_t)r * l]
                               : kern_exact<T>(p.kid, p.rnodes[(int64_t)b.rc * r + k],
                                            p.cnodes[(int64_t)b.cc * r + l]);
                tmp += F * vr[l];
            }
            v += ur[k] * tmp;
        }
    }
    out[t] = v;
}

inline int nblk(int64_t waves) { return (int)((waves + WPB - 1) / WPB); }
// wavefronts of k_translate: one cluster per lane group of r lanes when one
// right-hand side is translated, one per wavefront otherwise
inline int64_t tr_waves(int64_t ncl, int r, int kb) {
    const int G = kb == 1 ? WAVE / r : 1;
    return (ncl + G - 1) / G;
}

template <class T, int RMAX>
hipError_t gen_basis_r(const DevPlanT<T> &p, hipStream_t s) {
    if (p.ncch > 0 && p.V)
        k_gen_basis<T, RMAX><<<nblk(p.ncch), WG, 0, s>>>(p.cch, p.ncch, p.cfar, p.ccl, p.cnodes, p.lam,
                                                     p.y, p.V, p.r);
    if (p.nrch > 0 && p.U)
        k_gen_basis<T, RMAX><<<nblk(p.nrch), WG, 0, s>>>(p.rch, p.nrch, p.rfar, p.rcl, p.rnodes, p.lam,
                                                     p.x, p.U, p.r);
    return hipGetLastError();
}

// y[i incy] += t[i] for i < n: a split product's far field, added once the
// near field has written y
template <class T>
__global__ __launch_bounds__(WG) void k_add_far(T *y, int64_t incy, const T *t, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * WG + threadIdx.x;
    if (i < n) y[i * incy] += t[i];
}

// MODE: 0 stored, 1 hybrid, 2 free
template <class T, int RMAX, int MODE, int KID, int KB>
hipError_t matvec_t(const DevPlanT<T> &p, const MvArgsT<T> &a, bool tr, hipStream_t s, const hipEvent_t *ev) {
    constexpr bool CHEB = MODE != 0;
    constexpr bool FREE = MODE == 2;
    auto mark = [&](int q) {
        if (ev) (void)hipEventRecord(ev[q], s);
    };
    // "in" = the side summed over (columns forward), "out" = the side written
    const DTask *in_task = tr ? p.rtask : p.ctask;
    const int in_ntask = tr ? p.nrtask : p.nctask;
    const DCluster *in_cl = tr ? p.rcl : p.ccl;
    const int32_t *in_multi = tr ? p.multi_r : p.multi_c;  // clusters with several tasks
    const int in_nmulti = tr ? p.nmulti_r : p.nmulti_c;
    const T *in_B = tr ? p.U : p.V;
    const V2<T> *in_map = tr ? p.rmap : p.cmap;
    const T *in_pts = tr ? p.x : p.y;
    const DChunk *out_ch = tr ? p.cch : p.rch;
    const int out_n = tr ? p.ncch : p.nrch;
    const DNear *out_near = tr ? p.cnear : p.rnear;
    const DFar *out_far = tr ? p.cfar : p.rfar;
    const DCluster *out_cl = tr ? p.ccl : p.rcl;
    const int out_ncl = tr ? p.ncc : p.nrc;
    const T *out_B = tr ? p.V : p.U;
    const V2<T> *out_map = tr ? p.cmap : p.rmap;
    const T *out_pts = tr ? p.y : p.x;
    const int32_t *out_far_list = tr ? p.far_by_cc : p.far_by_rc;

    auto down = [&](const MvArgsT<T> &da, hipStream_t ds) {
        if (tr) {
            if constexpr (KB == 1)
                k_down<T, RMAX, FREE, KID, true, CHEB, 1><<<nblk(out_n), WG, 0, ds>>>(
                    p, out_ch, out_n, out_near, out_far, out_cl, out_B, out_map, out_pts, in_pts, a.z, da);
        } else {
            k_down<T, RMAX, FREE, KID, false, CHEB, KB><<<nblk(out_n), WG, 0, ds>>>(
                p, out_ch, out_n, out_near, out_far, out_cl, out_B, out_map, out_pts, in_pts, a.z, da);
        }
    };
    // the near field depends on x alone: with a side stream it runs beside the
    // far-field chain (up, reduce, translate), and the far field adds onto y
    // after the join. One right-hand side, matrix-free or Float32: the
    // downward far field also runs beside the near field, into yt, added onto
    // y after the join (measured faster there; slower for an FP64 hybrid
    // product, whose streamed near field it slows, and with several
    // right-hand sides, whose add costs more than the overlap gains)
    constexpr bool FAR_AHEAD = KB == 1 && (FREE || std::is_same<T, float>::value);
    const bool split = a.side && (!FAR_AHEAD || a.yt) && !ev && out_n > 0 && p.nfar > 0 &&
                       in_ntask > 0;
    const int64_t out_rows = tr ? p.ny : p.nx;
    if (split) {
        hipError_t e = hipEventRecord(a.fork, s);
        if (e == hipSuccess) e = hipStreamWaitEvent(a.side, a.fork, 0);
        if (e != hipSuccess) return e;
        MvArgsT<T> an = a;
        an.part = 1;
        down(an, a.side);
        e = hipEventRecord(a.join, a.side);
        if (e != hipSuccess) return e;
    }
    mark(0);
    if (p.nfar > 0 && in_ntask > 0) {
        k_up<T, RMAX, CHEB, KB><<<in_ntask, WG, 0, s>>>(in_task, in_cl, in_B, in_map, in_pts, a.x,
                                                      a.incx, a.ldx, a.partial, a.w, p.chebA,
                                                      p.r);
        mark(1);
        // one right-hand side: the translation sums the slots of clusters with
        // several tasks itself (stored or regenerated F)
        constexpr bool T1 = KB == 1;
        if (!T1 && in_nmulti > 0)
            k_reduce<T, CHEB, KB><<<nblk(in_nmulti), WG, 0, s>>>(in_cl, in_multi, in_nmulti,
                                                                 a.partial, p.chebA, a.w, p.r);
        mark(2);
        if constexpr (T1) {
            if (tr)
                k_translate1<T, RMAX, true, CHEB, FREE, KID><<<nblk(out_ncl), WG, 0, s>>>(
                    p, out_cl, out_ncl, a.partial, a.w, a.z);
            else
                k_translate1<T, RMAX, false, CHEB, FREE, KID><<<nblk(out_ncl), WG, 0, s>>>(
                    p, out_cl, out_ncl, a.partial, a.w, a.z);
        } else {  // several right-hand sides (forward only)
            k_translate<T, RMAX, FREE, KID, false, CHEB, KB>
                <<<nblk(tr_waves(out_ncl, p.r, KB)), WG, 0, s>>>(p, out_cl, out_ncl, out_far_list,
                                                                a.w, a.z);
        }
    } else {
        mark(1);
        mark(2);
    }
    mark(3);
    if (split && FAR_AHEAD) {
        MvArgsT<T> af = a;
        af.part = 3;
        af.y = a.yt;
        af.incy = 1;
        af.beta = 0;
        down(af, s);
        const hipError_t e = hipStreamWaitEvent(s, a.join, 0);
        if (e != hipSuccess) return e;
        k_add_far<T><<<(int)((out_rows + WG - 1) / WG), WG, 0, s>>>(a.y, a.incy, a.yt, out_rows);
    } else if (split) {
        const hipError_t e = hipStreamWaitEvent(s, a.join, 0);
        if (e != hipSuccess) return e;
        MvArgsT<T> af = a;
        af.part = 2;
        down(af, s);
    } else if (out_n > 0) {
        down(a, s);
    }
    mark(4);
    return hipGetLastError();
}

template <class T, int RMAX, int KID>
hipError_t grad_t(const DevPlanT<T> &p, const MvArgsT<T> &a, bool tr, T *y2, int64_t incy2,
                  const T *sv, int64_t incs, hipStream_t s) {
    const DTask *in_task = tr ? p.rtask : p.ctask;
    const int in_ntask = tr ? p.nrtask : p.nctask;
    const DCluster *in_cl = tr ? p.rcl : p.ccl;
    const V2<T> *in_map = tr ? p.rmap : p.cmap;
    const T *in_pts = tr ? p.x : p.y;
    const DChunk *out_ch = tr ? p.cch : p.rch;
    const int out_n = tr ? p.ncch : p.nrch;
    const DNear *out_near = tr ? p.cnear : p.rnear;
    const DFar *out_far = tr ? p.cfar : p.rfar;
    const DCluster *out_cl = tr ? p.ccl : p.rcl;
    const int out_ncl = tr ? p.ncc : p.nrc;
    const V2<T> *out_map = tr ? p.cmap : p.rmap;
    const T *out_pts = tr ? p.y : p.x;
    if (p.nfar > 0 && in_ntask > 0) {
        k_up<T, RMAX, true, 1><<<in_ntask, WG, 0, s>>>(in_task, in_cl, nullptr, in_map, in_pts, a.x,
                                                     a.incx, 0, a.partial, a.w, p.chebA, p.r);
        // the translation sums the slots of clusters with several tasks
        // itself, F regenerated from the nodes
        if (tr)
            k_translate1<T, RMAX, true, true, true, KID><<<nblk(out_ncl), WG, 0, s>>>(
                p, out_cl, out_ncl, a.partial, a.w, a.z);
        else
            k_translate1<T, RMAX, false, true, true, KID><<<nblk(out_ncl), WG, 0, s>>>(
                p, out_cl, out_ncl, a.partial, a.w, a.z);
    }
    if (out_n > 0) {
        if (tr)
            k_down_grad<T, RMAX, KID, true, 2><<<out_n, WG, 0, s>>>(
                p, out_ch, out_n, out_near, out_far, out_map, out_pts, in_pts, a.z, a, y2, incy2, sv, incs);
        else if (y2)  // the product and the x-derivative sums together
            k_down_grad<T, RMAX, KID, false, 2><<<out_n, WG, 0, s>>>(
                p, out_ch, out_n, out_near, out_far, out_map, out_pts, in_pts, a.z, a, y2, incy2, sv, incs);
        else
            k_down_grad<T, RMAX, KID, false, 1><<<out_n, WG, 0, s>>>(
                p, out_ch, out_n, out_near, out_far, out_map, out_pts, in_pts, a.z, a, y2, incy2, sv, incs);
    }
    return hipGetLastError();
}

template <class T, int RMAX>
hipError_t grad_r(const DevPlanT<T> &p, const MvArgsT<T> &a, bool tr, T *y2, int64_t incy2,
                  const T *sv, int64_t incs, hipStream_t s) {
    switch (p.kid) {
    case 0: return grad_t<T, RMAX, 0>(p, a, tr, y2, incy2, sv, incs, s);
    case 1: return grad_t<T, RMAX, 1>(p, a, tr, y2, incy2, sv, incs, s);
    case 2: return grad_t<T, RMAX, 2>(p, a, tr, y2, incy2, sv, incs, s);
    default: return grad_t<T, RMAX, 3>(p, a, tr, y2, incy2, sv, incs, s);
    }
}


// ------------------------------------------ multi-RHS passes on the MFMA units
// Forward products of up to 32 right-hand sides per launch, in G groups of 16
// (the N of the 16x16xK shapes): every operand a pass generates or streams
// (Chebyshev terms, regenerated or stored near-field entries, F) feeds the G
// groups' products, so a launch of 32 costs one launch of 16 in vector-unit
// work and in streamed bytes. nq <= 16 G right-hand sides are valid; the
// others read and write zeros. Far field in Chebyshev form; every contraction
// on the matrix cores:
//   up         mu = T(points)^T X per task, one wavefront per task; a cluster
//              with a single task applies A at once (w = A mu), the others
//              leave partial sums for the reduce
//   reduce     w = A sum(mu) over the clusters with several tasks
//   translate  z = sum_b F_b w_b and c = A^T z, one wavefront per row cluster
//   down       near field A_blk X and far field T(rows) C, one wavefront per
//              64-row chunk, written once
// PREC selects the arithmetic of the far-field products T X and T C:
//   2: FP64 on v_mfma_f64_16x16x4f64
//   0: FP32 operands on v_mfma_f32_16x16x4f32, FP32 accumulation
//   1: split BF16 (x = hi + lo, hi = bf16(x), lo = bf16(x - hi); products
//      hi*hi + hi*lo + lo*hi) on v_mfma_f32_16x16x32_bf16, FP32 accumulation
//   3: split BF16 in three (x = hi + mid + lo, each the bf16 rounding of what
//      the previous pieces leave; products hi*hi + hi*mid + mid*hi + hi*lo +
//      lo*hi + mid*mid, the terms down to 2^-16 of hi*hi), FP32 accumulation:
//      operands are carried to FP32's 24 bits, where two pieces carry 16
// The FP32 accumulators of the reduced-precision products hold one step's
// sum only (64 points upward, one cluster's series of one row tile downward)
// and are added into the FP64 accumulators after every step, so FP32
// rounding does not grow with the number of points or clusters. Near field,
// reduce and translate run in FP64 (v_mfma_f64_16x16x4f64) for every PREC;
// point differences and the Chebyshev recurrence are FP64.
// Operand layouts (lane l, j16 = l % 16, g = l / 16): A[i][k] comes from the
// lane with i = j16 and k = g (K = 4 shapes) or k = 8g + v, v < 8 (K = 32);
// B[k][j] from the lane with j = j16 and the same k. Results: the FP32 shapes
// hold D[4g + v][j16] in register v, the FP64 shape holds D[g + 4v][j16]. The
// reduced-precision products take row rho(i) = i/4 + 4 (i % 4) into A-row slot
// i, so their register v holds row g + 4v, the FP64 shape's row: they add onto
// the FP64 accumulators lane for lane.
typedef float mf4_t __attribute__((ext_vector_type(4)));
typedef double md4_t __attribute__((ext_vector_type(4)));
typedef __bf16 mbf8_t __attribute__((ext_vector_type(8)));
constexpr int MQ = 16;     // right-hand sides per group (the MFMA's N)
constexpr int MGMAX = 2;   // groups per launch at most
__device__ __forceinline__ int rho(int s) { return (s >> 2) + 4 * (s & 3); }

__device__ __forceinline__ __bf16 bf16_hi(float x) { return (__bf16)x; }
__device__ __forceinline__ __bf16 bf16_lo(float x) { return (__bf16)(x - (float)(__bf16)x); }
// three-way split: x - hi and x - hi - mid are exact in FP32 (each drops the
// leading 8 significant bits of a value with at most 24)
__device__ __forceinline__ void bf16_split3(float x, __bf16 &h, __bf16 &m, __bf16 &l) {
    h = (__bf16)x;
    const float r1 = x - (float)h;
    m = (__bf16)r1;
    l = (__bf16)(r1 - (float)m);
}
// D += A B over split operands: PREC 1 three products (hi*hi, hi*lo, lo*hi),
// PREC 3 six (down to mid*mid)
template <int PREC>
__device__ __forceinline__ mf4_t mfma_split(const mbf8_t &ah, const mbf8_t &am, const mbf8_t &al,
                                            const mbf8_t &bh, const mbf8_t &bm, const mbf8_t &bl,
                                            mf4_t d) {
    d = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, bh, d, 0, 0, 0);
    if constexpr (PREC == 1) {
        d = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, bl, d, 0, 0, 0);
        d = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al, bh, d, 0, 0, 0);
    } else {
        d = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, bm, d, 0, 0, 0);
        d = __builtin_amdgcn_mfma_f32_16x16x32_bf16(am, bh, d, 0, 0, 0);
        d = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, bl, d, 0, 0, 0);
        d = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al, bh, d, 0, 0, 0);
        d = __builtin_amdgcn_mfma_f32_16x16x32_bf16(am, bm, d, 0, 0, 0);
    }
    return d;
}
// the 8 values of a split operand: PREC 1 uses (hi, lo) with mid unused,
// PREC 3 (hi, mid, lo)
template <int PREC>
__device__ __forceinline__ void split8(const float *v, mbf8_t &h, mbf8_t &m, mbf8_t &l) {
#pragma unroll
    for (int q = 0; q < 8; ++q) {
        if constexpr (PREC == 1) {
            h[q] = bf16_hi(v[q]);
            l[q] = bf16_lo(v[q]);
            m[q] = (__bf16)0.f;
        } else {
            __bf16 a, b, c;
            bf16_split3(v[q], a, b, c);
            h[q] = a;
            m[q] = b;
            l[q] = c;
        }
    }
}
__device__ __forceinline__ md4_t mfma64(double a, double b, md4_t c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}
__device__ __forceinline__ md4_t add_f32_acc(md4_t d, mf4_t f) {
    return md4_t{d[0] + (double)f[0], d[1] + (double)f[1], d[2] + (double)f[2],
                 d[3] + (double)f[3]};
}
__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
}

// The nested far field's node values w (of column clusters) and series z (of
// row clusters) are stored in FP64 for the FP64 form and rounded once to FP32
// for the reduced-precision forms (whose moments already carry FP32 terms):
// the passes that move them, the translation above all, read and write half
// the bytes, and every sum over them still runs in FP64. The chains of
// moments and coefficients between tree levels stay FP64.
template <int PREC> using NodeT = std::conditional_t<PREC == 2, double, float>;

// Results of the FP64 shape held as Mo[gi][t][v] = X[16t + g + 4v][16 gi + j16]
// (rank index m = 16t + g + 4v, right-hand side q = 16 gi + j16), stored as
// dst[q r + m], the layout of partial slots, w and z, or (KQ, the nested
// moments and coefficients) as dst[m 16G + q], so that a lane's right-hand
// sides are contiguous.
template <int RP, int G, bool KQ = false, class S = double>
__device__ __forceinline__ void store_ranks(const md4_t (&Mo)[G][RP / 16], S *dst, int r,
                                            int j16, int g) {
#pragma unroll
    for (int gi = 0; gi < G; ++gi) {
#pragma unroll
        for (int t = 0; t < RP / 16; ++t)
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int m = t * 16 + g + 4 * v;
                if (m < r) {
                    if constexpr (KQ) dst[(int64_t)m * (MQ * G) + gi * MQ + j16] = (S)Mo[gi][t][v];
                    else dst[(int64_t)(gi * MQ + j16) * r + m] = (S)Mo[gi][t][v];
                }
            }
    }
}

// Node values w = A mu of one cluster from its moments held as in
// store_ranks: Mo[gi][t][v] is B[k0 + g][j16] of the K-step k0 = 16t + 4v, so
// the product takes mu from registers; A's rows are loaded once for the
// groups. Stored as dst[q r + m], or (KQ, the nested far field's) as
// dst[m 16G + q] (FP64, or rounded to FP32 once).
template <int RP, int G, class S, bool KQ = false>
__device__ __forceinline__ void nodes_from_moments(const md4_t (&Mo)[G][RP / 16],
                                                   const double *chebA, S *dst, int r,
                                                   int j16, int g) {
    constexpr int MT = RP / 16;
    double av[MT][MT][4];
#pragma unroll
    for (int tt = 0; tt < MT; ++tt)
#pragma unroll
        for (int t = 0; t < MT; ++t)
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int m = tt * 16 + j16, k = t * 16 + 4 * v + g;
                av[tt][t][v] = (m < r && k < r) ? chebA[m * r + k] : 0.0;
            }
    md4_t W[G][MT];
#pragma unroll
    for (int gi = 0; gi < G; ++gi)
#pragma unroll
        for (int tt = 0; tt < MT; ++tt) {
            W[gi][tt] = md4_t{0.0, 0.0, 0.0, 0.0};
#pragma unroll
            for (int t = 0; t < MT; ++t)
#pragma unroll
                for (int v = 0; v < 4; ++v) W[gi][tt] = mfma64(av[tt][t][v], Mo[gi][t][v], W[gi][tt]);
        }
    store_ranks<RP, G, KQ>(W, dst, r, j16, g);
}

// xt[j][q] = x[j*incx + q*ldx] for q < nq, 0 beyond (q < 16 G): the
// right-hand sides of one launch packed row by row, so every MFMA B operand
// of the upward and downward passes is a coalesced load.
template <int G>
__global__ __launch_bounds__(WG) void k_pack_x(const double *x, int64_t incx, int64_t ldx, int nq,
                                              int64_t n, double *xt) {
    constexpr int MQT = MQ * G;
    const int64_t j = (int64_t)blockIdx.x * WG + threadIdx.x;
    if (j >= n) return;
    double v[MQT];
#pragma unroll
    for (int q = 0; q < MQT; ++q) v[q] = q < nq ? x[j * incx + q * ldx] : 0.0;
    double2 *o = reinterpret_cast<double2 *>(xt + j * MQT);
#pragma unroll
    for (int q = 0; q < MQT / 2; ++q) o[q] = make_double2(v[2 * q], v[2 * q + 1]);
}

// Upward pass, one wavefront (the whole workgroup) per task: the moments
// mu[m][q] = sum_j T_m(xh_j) x[j, q] of the task's points. The wavefront forms
// T for PTS points (one per lane, FP64 recurrence) in LDS, then contracts over
// the points on MFMA, each T operand feeding the G groups.
// NEST: the tasks are the column tree's leaves (task.cl the node, task.slot
// its column cluster or -1); the moments go to mom[node] and, for a leaf that
// is a cluster, w = A mu to w[cluster].
template <int RP, int PREC, int G, bool NEST, class WS>
__global__ __launch_bounds__(WAVE) void k_up_m(const DTask *tasks, const double *pts,
                                              const double *xt, int64_t ldxt, int nq,
                                              const double *chebA, double *partial, WS *w,
                                              int r, double *mom) {
    static_assert(NEST || std::is_same<WS, double>::value, "per-level w is FP64");
    using TL = typename std::conditional<PREC == 2, double, float>::type;
    constexpr bool BF = PREC == 1 || PREC == 3;  // split-BF16 operands (K = 32 shape)
    constexpr int MT = RP / 16;
    constexpr int MQT = MQ * G;
    constexpr int PTS = (PREC == 2 && RP > 32) ? 32 : 64;  // points per step
    __shared__ __align__(16) TL Tl[RP][PTS + 1];
    const DTask Tk = tasks[blockIdx.x];
    const int lane = threadIdx.x;
    const int j16 = lane % 16, g = lane / 16;
    // the A-operand row (rank) of this lane: j16 for FP64, rho(j16) otherwise
    const int arow = PREC == 2 ? j16 : rho(j16);
    const double2 cm = make_double2(Tk.cx, Tk.cy);
    const int64_t first = Tk.first;
    md4_t D64[G][MT];
#pragma unroll
    for (int gi = 0; gi < G; ++gi)
#pragma unroll
        for (int t = 0; t < MT; ++t) D64[gi][t] = md4_t{0.0, 0.0, 0.0, 0.0};
    for (int base = 0; base < Tk.len; base += PTS) {
        // this lane's B operands of the step, loaded before the recurrence:
        // x[point][16 gi + j16] for the points k0 + g (K = 4) or k0 + 8g + v (K = 32)
        double bx[G][PTS / 4];
#pragma unroll
        for (int u = 0; u < PTS / 4; ++u) {
            const int i = base + (BF ? (u / 8) * 32 + 8 * g + (u % 8) : 4 * u + g);
#pragma unroll
            for (int gi = 0; gi < G; ++gi) {
                const int q = gi * MQ + j16;
                // in bounds for every lane, issued unconditionally, zeroed after
                const bool ok = i < Tk.len && q < nq;
                const double v = xt[(first + (ok ? i : 0)) * ldxt + (ok ? q : 0)];
                bx[gi][u] = ok ? v : 0.0;
            }
        }
        if (lane < PTS) {
            const int i = base + lane;
            const bool ok = i < Tk.len;
            const double xh = ok ? (pts[first + i] - cm.x) * cm.y : 0.0;
            double t0 = 1.0, t1 = xh;
#pragma unroll
            for (int m = 0; m < RP; ++m) {
                double tv = m == 0 ? t0 : t1;
                if (m >= 2) {
                    tv = __builtin_fma(xh + xh, t1, -t0);
                    t0 = t1;
                    t1 = tv;
                }
                Tl[m][lane] = (ok && m < r) ? (TL)tv : (TL)0;
            }
        }
        wave_sync();
        if constexpr (PREC == 2) {
#pragma unroll
            for (int k0 = 0; k0 < PTS; k0 += 4)
#pragma unroll
                for (int t = 0; t < MT; ++t) {
                    const double av = Tl[t * 16 + arow][k0 + g];
#pragma unroll
                    for (int gi = 0; gi < G; ++gi) D64[gi][t] = mfma64(av, bx[gi][k0 / 4], D64[gi][t]);
                }
        } else if constexpr (PREC == 0) {
#pragma unroll
            for (int t = 0; t < MT; ++t) {
                mf4_t D32[G];
#pragma unroll
                for (int gi = 0; gi < G; ++gi) D32[gi] = mf4_t{0.f, 0.f, 0.f, 0.f};
#pragma unroll
                for (int k0 = 0; k0 < PTS; k0 += 4) {
                    const float av = Tl[t * 16 + arow][k0 + g];
#pragma unroll
                    for (int gi = 0; gi < G; ++gi)
                        D32[gi] = __builtin_amdgcn_mfma_f32_16x16x4f32(av, (float)bx[gi][k0 / 4], D32[gi], 0, 0, 0);
                }
#pragma unroll
                for (int gi = 0; gi < G; ++gi) D64[gi][t] = add_f32_acc(D64[gi][t], D32[gi]);
            }
        } else {
#pragma unroll
            for (int t = 0; t < MT; ++t) {
                mf4_t D32[G];
#pragma unroll
                for (int gi = 0; gi < G; ++gi) D32[gi] = mf4_t{0.f, 0.f, 0.f, 0.f};
#pragma unroll
                for (int k0 = 0; k0 < PTS; k0 += 32) {
                    float tv[8];
#pragma unroll
                    for (int v = 0; v < 8; ++v) tv[v] = Tl[t * 16 + arow][k0 + 8 * g + v];
                    mbf8_t ah, am, al;
                    split8<PREC>(tv, ah, am, al);
#pragma unroll
                    for (int gi = 0; gi < G; ++gi) {
                        float bv[8];
#pragma unroll
                        for (int v = 0; v < 8; ++v) bv[v] = (float)bx[gi][(k0 / 32) * 8 + v];
                        mbf8_t bh, bm, bl;
                        split8<PREC>(bv, bh, bm, bl);
                        D32[gi] = mfma_split<PREC>(ah, am, al, bh, bm, bl, D32[gi]);
                    }
                }
#pragma unroll
                for (int gi = 0; gi < G; ++gi) D64[gi][t] = add_f32_acc(D64[gi][t], D32[gi]);
            }
        }
        wave_sync();
    }
    // D64 holds mu as store_ranks and nodes_from_moments read it
    if constexpr (NEST) {
        store_ranks<RP, G, true>(D64, mom + (int64_t)Tk.cl * MQT * r, r, j16, g);
        if (Tk.slot >= 0)
            nodes_from_moments<RP, G, WS, true>(D64, chebA, w + Tk.slot * MQT * r, r, j16, g);
    } else if (Tk.single) {
        nodes_from_moments<RP, G>(D64, chebA, w + (int64_t)Tk.cl * MQT * r, r, j16, g);
    } else {  // a multi-task cluster's slot gets mu itself
        store_ranks<RP, G>(D64, partial + Tk.slot * MQT * r, r, j16, g);
    }
}

// Nested far field, level kernels. Moments combine up the column tree
// (mu_P = M_0 mu_0 + M_1 mu_1: the halves' series re-expanded on the parent's
// interval) and coefficients go down the row tree (c_N = z[its row cluster] +
// M_h^T c_parent), one wavefront per node. A node's result, held as the MFMA
// result layout (value v of lane (j16, g) is row 16t + 4v + g), is the B
// operand of the K-step 16t + 4v as it stands, so a wavefront can go on to the
// next depth from registers: the upward launches cover two depths each, and
// the moments of the lower one never reach memory.

// W += M_h mu, mu in the result layout (rows past r zero)
template <int RP, int G>
__device__ __forceinline__ void m2m_acc(const double *Mh, const md4_t (&mu)[G][RP / 16], int r,
                                        int j16, int g, md4_t (&W)[G][RP / 16]) {
    constexpr int MT = RP / 16;
#pragma unroll
    for (int tt = 0; tt < MT; ++tt) {
        double av[MT][4];
#pragma unroll
        for (int t = 0; t < MT; ++t)
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int m = tt * 16 + j16, k = t * 16 + 4 * v + g;
                const double a = Mh[min(m, r - 1) * r + min(k, r - 1)];
                av[t][v] = (m < r && k < r) ? a : 0.0;
            }
#pragma unroll
        for (int gi = 0; gi < G; ++gi)
#pragma unroll
            for (int t = 0; t < MT; ++t)
#pragma unroll
                for (int v = 0; v < 4; ++v) W[gi][tt] = mfma64(av[t][v], mu[gi][t][v], W[gi][tt]);
    }
}

// The stored moments of node C in the result layout (zero past r)
template <int RP, int G>
__device__ __forceinline__ void load_moments(const double *mom, int C, int r, int j16, int g,
                                             md4_t (&mu)[G][RP / 16]) {
    constexpr int MQT = MQ * G;
    const double *src = mom + (int64_t)C * MQT * r;
#pragma unroll
    for (int gi = 0; gi < G; ++gi)
#pragma unroll
        for (int t = 0; t < RP / 16; ++t)
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int k = t * 16 + 4 * v + g;
                const double b = src[(int64_t)min(k, r - 1) * MQT + gi * MQ + j16];
                mu[gi][t][v] = k < r ? b : 0.0;
            }
}

// mu_N from its two halves' stored moments (an empty half's stay zero)
template <int RP, int G>
__device__ __forceinline__ void m2m_node(const DTNode &N, const double *M, const double *mom, int r,
                                         int j16, int g, md4_t (&W)[G][RP / 16]) {
#pragma unroll
    for (int gi = 0; gi < G; ++gi)
#pragma unroll
        for (int t = 0; t < RP / 16; ++t) W[gi][t] = md4_t{0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int h = 0; h < 2; ++h) {
        md4_t mu[G][RP / 16];
        load_moments<RP, G>(mom, N.child[h], r, j16, g, mu);
        m2m_acc<RP, G>(M + (int64_t)h * r * r, mu, r, j16, g, W);
    }
}

// Upward, one depth: the moments of its internal column nodes (stored), and
// w = A mu of those that are column clusters. TWO: also the depth below, each
// internal half formed from its own halves in registers (its w stored, its
// moments not: only this node reads them); every load is issued before the
// first product, so the halves' chains overlap.
template <int RP, int G, class WS, bool TWO>
__global__ __launch_bounds__(WG) void k_m2m(const DTNode *tn, const int32_t *ids, int n,
                                           const double *M, double *mom, const double *chebA,
                                           WS *w, int r) {
    constexpr int MT = RP / 16;
    constexpr int MQT = MQ * G;
    const int wid = wave_id();
    const int lane = threadIdx.x % WAVE;
    if (wid >= n) return;
    const int j16 = lane % 16, g = lane / 16;
    const int P = ids[wid];
    const DTNode N = tn[P];
    md4_t W[G][MT];
    if constexpr (!TWO) {
        m2m_node<RP, G>(N, M, mom, r, j16, g, W);
    } else {
        const DTNode Cn[2] = {tn[N.child[0]], tn[N.child[1]]};
        // internal halves (wave-uniform); a leaf's moments come from the
        // upward pass, an empty half's are zero
        bool in[2];
        md4_t qm[2][2][G][MT];  // per half: its halves' moments, or its own in [0]
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            in[h] = Cn[h].child[0] >= 0 && Cn[h].len > 0;
            load_moments<RP, G>(mom, in[h] ? Cn[h].child[0] : N.child[h], r, j16, g, qm[h][0]);
            if (in[h]) load_moments<RP, G>(mom, Cn[h].child[1], r, j16, g, qm[h][1]);
        }
#pragma unroll
        for (int gi = 0; gi < G; ++gi)
#pragma unroll
            for (int t = 0; t < MT; ++t) W[gi][t] = md4_t{0.0, 0.0, 0.0, 0.0};
        md4_t mu[2][G][MT];
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            if (in[h]) {
#pragma unroll
                for (int gi = 0; gi < G; ++gi)
#pragma unroll
                    for (int t = 0; t < MT; ++t) mu[h][gi][t] = md4_t{0.0, 0.0, 0.0, 0.0};
                m2m_acc<RP, G>(M, qm[h][0], r, j16, g, mu[h]);
                m2m_acc<RP, G>(M + (int64_t)r * r, qm[h][1], r, j16, g, mu[h]);
            } else {
#pragma unroll
                for (int gi = 0; gi < G; ++gi)
#pragma unroll
                    for (int t = 0; t < MT; ++t) mu[h][gi][t] = qm[h][0][gi][t];
            }
            m2m_acc<RP, G>(M + (int64_t)h * r * r, mu[h], r, j16, g, W);
        }
#pragma unroll
        for (int h = 0; h < 2; ++h)
            if (in[h] && Cn[h].cl >= 0)
                nodes_from_moments<RP, G, WS, true>(mu[h], chebA, w + (int64_t)Cn[h].cl * MQT * r, r,
                                                    j16, g);
    }
    store_ranks<RP, G, true>(W, mom + (int64_t)P * MQT * r, r, j16, g);
    if (N.cl >= 0) nodes_from_moments<RP, G, WS, true>(W, chebA, w + (int64_t)N.cl * MQT * r, r, j16, g);
}

// z[N's row cluster] (zero if it is none) in the result layout (the FP64
// result layout, as store_ranks reads it)
template <int RP, int G, class WS>
__device__ __forceinline__ void load_z(const DTNode &N, const WS *z, int r, int j16, int g,
                                       md4_t (&W)[G][RP / 16]) {
    constexpr int MQT = MQ * G;
#pragma unroll
    for (int gi = 0; gi < G; ++gi)
#pragma unroll
        for (int t = 0; t < RP / 16; ++t)
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int m = t * 16 + g + 4 * v;
                double c = 0.0;
                if (N.cl >= 0 && m < r) c = (double)z[((int64_t)N.cl * r + m) * MQT + gi * MQ + j16];
                W[gi][t][v] = c;
            }
}

// W += M_h^T c_P, c_P in the result layout
template <int RP, int G>
__device__ __forceinline__ void l2l_acc(const double *Mh, const md4_t (&cp)[G][RP / 16], int r,
                                        int j16, int g, md4_t (&W)[G][RP / 16]) {
    constexpr int MT = RP / 16;
    // A[m][k] = M_h[k][m], B[k][q] = c_P[k][q]
#pragma unroll
    for (int tt = 0; tt < MT; ++tt) {
        double av[MT][4];
#pragma unroll
        for (int t = 0; t < MT; ++t)
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int m = tt * 16 + j16, k = t * 16 + 4 * v + g;
                const double a = Mh[min(k, r - 1) * r + min(m, r - 1)];
                av[t][v] = (m < r && k < r) ? a : 0.0;
            }
#pragma unroll
        for (int gi = 0; gi < G; ++gi)
#pragma unroll
            for (int t = 0; t < MT; ++t)
#pragma unroll
                for (int v = 0; v < 4; ++v) W[gi][tt] = mfma64(av[t][v], cp[gi][t][v], W[gi][tt]);
    }
}

// Downward, row node N: c_N = z[N's row cluster] (if it is one) + M_h^T c_P
// from its parent's stored coefficients
template <int RP, int G, class WS>
__device__ __forceinline__ void l2l_node(const DTNode &N, const double *M, const WS *z,
                                         const double *cof, int r, int j16, int g,
                                         md4_t (&W)[G][RP / 16]) {
    md4_t cp[G][RP / 16];
    if (N.parent >= 0) load_moments<RP, G>(cof, N.parent, r, j16, g, cp);
    load_z<RP, G, WS>(N, z, r, j16, g, W);
    if (N.parent >= 0) l2l_acc<RP, G>(M + (int64_t)N.half * r * r, cp, r, j16, g, W);
}

// A chunk's row leaf's coefficients (l2l_node from its parent's) into the
// wavefront's LDS, rows of r + 1 (padded against bank conflicts), where the
// downward pass's far field reads them.
template <int RP, int G, class WS>
__device__ __forceinline__ void leaf_series_to_lds(const DevPlan &p, int leaf, const WS *z, int r,
                                                   int j16, int g, double *Dl) {
    const DTNode N = p.xtn[leaf];
    md4_t W[G][RP / 16];
    l2l_node<RP, G, WS>(N, p.nestM, z, p.mx, r, j16, g, W);
#pragma unroll
    for (int gi = 0; gi < G; ++gi)
#pragma unroll
        for (int t = 0; t < RP / 16; ++t)
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int m = t * 16 + g + 4 * v;
                if (m < r) Dl[(gi * MQ + j16) * (r + 1) + m] = W[gi][t][v];
            }
    wave_sync();
}

// Downward, one depth: the coefficients of its internal row nodes (a chunk's
// leaf forms its own in the downward pass). Two depths per launch measured
// slower here: the halves' products wait on this node's, and their registers
// cost the occupancy.
template <int RP, int G, class WS>
__global__ __launch_bounds__(WG) void k_l2l(const DTNode *tn, const int32_t *ids, int n,
                                           const double *M, const WS *z, double *cof, int r) {
    constexpr int MQT = MQ * G;
    const int wid = wave_id();
    const int lane = threadIdx.x % WAVE;
    if (wid >= n) return;
    const int j16 = lane % 16, g = lane / 16;
    const int id = ids[wid];
    const DTNode N = tn[id];
    md4_t W[G][RP / 16];
    l2l_node<RP, G, WS>(N, M, z, cof, r, j16, g, W);
    store_ranks<RP, G, true>(W, cof + (int64_t)id * MQT * r, r, j16, g);
}

// Translation, one wavefront per row cluster: Z = sum over the cluster's far
// blocks of F_b W_b (F stored, or regenerated from the mapped nodes; each F
// operand feeds the G groups), then C = A^T Z. The FP64 result layout of Z is
// the B-operand layout of the second product, so Z feeds it from registers.
template <int RP, bool FREE, int KID, int G, class WS, bool KQ>
__global__ __launch_bounds__(WG) void k_translate_m(DevPlan p, const DCluster *cl, int ncl,
                                                   const WS *w, WS *z) {
    constexpr int MT = RP / 16;
    constexpr int MQT = MQ * G;
    const int wid = wave_id();
    const int lane = threadIdx.x % WAVE;
    if (wid >= ncl) return;
    const int r = p.r;
    const int j16 = lane % 16, g = lane / 16;
    const DCluster C = cl[wid];
    md4_t Z[G][MT];
#pragma unroll
    for (int gi = 0; gi < G; ++gi)
#pragma unroll
        for (int t = 0; t < MT; ++t) Z[gi][t] = md4_t{0.0, 0.0, 0.0, 0.0};
    for (int e0 = C.far_beg; e0 < C.far_end; e0 += WAVE) {
        // the far blocks' translation records (in this cluster's list
        // order), one per lane, read together
        const int ne = min(WAVE, C.far_end - e0);
        int b_cc = 0;
        int64_t b_foff = 0;
        if (lane < ne) {
            const DTrRec b = p.tr_rc[e0 + lane];
            b_cc = b.src;
            b_foff = b.foff;
        }
        for (int e = 0; e < ne; ++e) {
            const int cc = rl_i(b_cc, e);
            const WS *ws = w + (int64_t)cc * MQT * r;
            const double *F = FREE ? nullptr : p.F + rl_l(b_foff, e);
            const double *tx = p.rnodes + (int64_t)wid * r;
            const double *ty = p.cnodes + (int64_t)cc * r;
#pragma unroll
            for (int k0 = 0; k0 < RP; k0 += 16) {
                // four K-steps: operands first (every load in bounds and issued
                // unconditionally, the terms past r zeroed after), then the products
                double bv[G][4], av[4][MT];
#pragma unroll
                for (int h = 0; h < 4; ++h) {
                    const int l = k0 + 4 * h + g;
                    const bool lok = l < r;
                    const int lc = min(l, r - 1);
#pragma unroll
                    for (int gi = 0; gi < G; ++gi) {
                        const double v = KQ ? (double)ws[lc * MQT + gi * MQ + j16]
                                            : (double)ws[(gi * MQ + j16) * r + lc];
                        bv[gi][h] = lok ? v : 0.0;
                    }
                    const double tyl = FREE ? ty[lc] : 0.0;
#pragma unroll
                    for (int t = 0; t < MT; ++t) {
                        const int m = t * 16 + j16;
                        const bool ok = lok && m < r;
                        double v;
                        if constexpr (FREE) v = kern_fast<double, KID>(tx[min(m, r - 1)], tyl);
                        else v = __builtin_nontemporal_load(F + min(m, r - 1) + (int64_t)r * lc);
                        av[h][t] = ok ? v : 0.0;
                    }
                }
#pragma unroll
                for (int h = 0; h < 4; ++h)
#pragma unroll
                    for (int t = 0; t < MT; ++t)
#pragma unroll
                        for (int gi = 0; gi < G; ++gi) Z[gi][t] = mfma64(av[h][t], bv[gi][h], Z[gi][t]);
            }
        }
    }
    // C[m][q] = sum_k A[k][m] Z[k][q]; this lane holds Z[16t + g + 4v][j16] in
    // Z[t][v], which is B[k0 + g][j16] of the K-step k0 = 16t + 4v
#pragma unroll
    for (int gi = 0; gi < G; ++gi) {
        md4_t Cc[MT];
#pragma unroll
        for (int t = 0; t < MT; ++t) Cc[t] = md4_t{0.0, 0.0, 0.0, 0.0};
#pragma unroll
        for (int t = 0; t < MT; ++t)
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int k = t * 16 + 4 * v + g;
#pragma unroll
                for (int tt = 0; tt < MT; ++tt) {
                    const int m = tt * 16 + j16;
                    const double av = (k < r && m < r) ? p.chebA[k * r + m] : 0.0;
                    Cc[tt] = mfma64(av, Z[gi][t][v], Cc[tt]);
                }
            }
#pragma unroll
        for (int t = 0; t < MT; ++t)
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int m = t * 16 + g + 4 * v;
                if (m < r) {
                    if constexpr (KQ) z[((int64_t)wid * r + m) * MQT + gi * MQ + j16] = (WS)Cc[t][v];
                    else z[((int64_t)wid * MQT + gi * MQ + j16) * r + m] = (WS)Cc[t][v];
                }
            }
    }
}

// Wavefronts per SIMD that the downward pass asks the register allocator
// for. Left alone, it holds the 64 FP64 accumulators of a 64-row x 32-RHS
// chunk in AGPRs beside up to 130 VGPRs and runs three wavefronts (two for the
// per-level forms); the nested forms up to rank 16 fit four (G = 2) or five
// (G = 1). Above rank 16 (where a forced count was measured slower or would
// spill) and for the matrix-free log kernel the allocator keeps its choice.
constexpr int down_m_waves(int rp, bool free, int kid, int g, bool nest) {
    return (free && kid == 3) || !nest || rp > 16 ? 1 : g == 1 ? 5 : 4;
}

// Downward pass, one wavefront per 64-row chunk. Rows 16t + j16 (t < 4) are
// this lane's A-operand rows of the FP64 products (rows 16t + rho(j16) of the
// reduced-precision ones). Near field: per 8 columns of a dense block the
// block entries (stored, or regenerated from the points) enter the FP64 MFMA
// directly, each feeding the G groups. Far field: for each covering cluster,
// T_k(xh) of the lane's rows by the FP64 recurrence, times the cluster's
// coefficients C[k][q]. The 64 x 16G result goes through LDS, 16 rows at a
// time, and is written once: whole rows per store when a row's right-hand
// sides are contiguous (ldy = 1), else 16 rows of one right-hand side per
// store.
// NEST: the far field of a chunk is its row leaf's series alone (the
// coefficients of every covering cluster pushed down the row tree into p.mx).
template <int RP, int PREC, bool FREE, int KID, int G, bool NEST, class WS>
__global__ __launch_bounds__(WG) __attribute__((amdgpu_waves_per_eu(down_m_waves(RP, FREE, KID, G, NEST)))) void k_down_m(DevPlan p, const DChunk *ch, int nch,
                                              const DNear *nl, const DFar *fl,
                                              const double2 *cmapv, const double *pts_out,
                                              const double *pts_in, const double *xt,
                                              int64_t ldxt, const WS *z, MvArgs a, int nq) {
    static_assert(NEST || std::is_same<WS, double>::value, "per-level z is FP64");
    constexpr bool BF = PREC == 1 || PREC == 3;  // split-BF16 operands (K = 32 shape)
    constexpr int KSEL = BF ? ((RP + 31) / 32) * 8 : RP / 4;  // T values per lane and row tile
    constexpr int MQT = MQ * G;
    // stored blocks stream through LDS slabs, except where the slab path's
    // registers would cost the occupancy (two groups above rank 16, or one
    // above rank 32): there the block entries load straight into the A operand
    constexpr bool SLAB = !FREE && (RP <= 16 || (RP <= 32 && G == 1));
    // per wavefront: one 16-row output tile; with NEST first the leaf's
    // coefficients, rows of r + 1 (padded against bank conflicts)
    constexpr int LDSW0 = (NEST && (RP + 1) * MQT > 16 * MQT) ? (RP + 1) * MQT : 16 * MQT;
    // stored blocks: an 8-column slab of A and 8 rows of x first
    constexpr int LDSW = (SLAB && LDSW0 < 8 * HM_TILE + 8 * MQT) ? 8 * HM_TILE + 8 * MQT : LDSW0;
    __shared__ double ds_all[WPB][LDSW];
    const int wid = wave_id();
    const int lane = threadIdx.x % WAVE;
    double *Dl = ds_all[uniform((int)(threadIdx.x / WAVE)) % WPB];  // a scalar address
    if (wid >= nch) return;
    const int r = p.r;
    const DChunk c = ch[wid];
    const int j16 = lane % 16, g = lane / 16;
    const int frow = PREC == 2 ? j16 : rho(j16);  // the far-field A-operand row
    // the lane's row points: the regenerated near field needs them (po), the
    // far field its A-operand rows' (pf, loaded after the near field when
    // the near field is streamed, so no registers hold them meanwhile)
    double po[4], pf[4];
    bool rok[4], fok[4];
#pragma unroll
    for (int t = 0; t < 4; ++t) {
        rok[t] = t * 16 + j16 < c.np;
        fok[t] = t * 16 + frow < c.np;
        po[t] = (FREE && rok[t]) ? pts_out[c.p0 + t * 16 + j16] : 0.0;
    }
    auto load_pf = [&]() {
#pragma unroll
        for (int t = 0; t < 4; ++t)
            pf[t] = (FREE && PREC == 2) ? po[t] : (fok[t] ? pts_out[c.p0 + t * 16 + frow] : 0.0);
    };
    const int nt = (c.np + 15) / 16;  // row tiles holding rows of this chunk (wave-uniform)
    // NEST: the chunk's row leaf forms its coefficients (l2l_node from its
    // parent's) into LDS, where the far field reads them: before the near
    // field while no accumulators are live (matrix-free), or after it (stored
    // blocks pass through the same LDS)
    const int leaf = NEST ? p.chunk_leaf[wid] : 0;
    if constexpr (NEST && !SLAB) leaf_series_to_lds<RP, G, WS>(p, leaf, z, r, j16, g, Dl);
    md4_t D[G][4];
#pragma unroll
    for (int gi = 0; gi < G; ++gi)
#pragma unroll
        for (int t = 0; t < 4; ++t) D[gi][t] = md4_t{0.0, 0.0, 0.0, 0.0};
    // ---- near field, FP64: two K-steps (8 columns) per step; the chunk's
    // block records are read together, one per lane
    for (int e0 = c.near_beg; e0 < c.near_end; e0 += WAVE) {
        const int ne = min(WAVE, c.near_end - e0);
        int64_t d_aoff = 0, d_col0 = 0;
        int d_n = 0, d_off = 0, d_m = 1;
        if (lane < ne) {
            const DNear sn = nl[e0 + lane];
            const DDense Dd = p.dense[sn.blk];
            d_aoff = Dd.aoff;
            d_col0 = Dd.col0;
            d_n = Dd.n;
            if constexpr (SLAB) d_m = Dd.m;
            d_off = sn.off;
        }
        if constexpr (!SLAB) {
            // entries regenerated from the points (or loaded) straight into
            // the A operand
            for (int e = 0; e < ne; ++e) {
                const int64_t col0 = rl_l(d_col0, e);
                const int bn = rl_i(d_n, e);
                const double *xtb = xt + col0 * ldxt;
                const double *ab = FREE ? nullptr : p.A + rl_l(d_aoff, e);
                // element (i, cc) of a stored block sits at ab[rowoff(i) + 64 cc]
                uint32_t rowoff[4];
#pragma unroll
                for (int t = 0; t < 4; ++t) {
                    const int i = rl_i(d_off, e) + t * 16 + j16;
                    rowoff[t] = FREE ? 0u : (uint32_t)((i >> 6) * HM_TILE * bn + (i & (HM_TILE - 1)));
                }
                for (int k0 = 0; k0 < bn; k0 += 8) {
                    double bv[G][2], av[2][4];
#pragma unroll
                    for (int h = 0; h < 2; ++h) {
                        const int cc = k0 + 4 * h + g;
                        const bool cok = cc < bn;
#pragma unroll
                        for (int gi = 0; gi < G; ++gi) {
                            const int q = gi * MQ + j16;
                            bv[gi][h] = (cok && q < nq) ? xtb[(int64_t)cc * ldxt + q] : 0.0;
                        }
                        double yc = 0.0;
                        if constexpr (FREE) yc = cok ? pts_in[col0 + cc] : 0.0;
#pragma unroll
                        for (int t = 0; t < 4; ++t) {
                            av[h][t] = 0.0;
                            if (rok[t] && cok) {
                                if constexpr (FREE) av[h][t] = kern_fast<double, KID>(po[t], yc);
                                else av[h][t] = __builtin_nontemporal_load(ab + rowoff[t] + (uint32_t)(cc * HM_TILE));
                            }
                        }
                    }
#pragma unroll
                    for (int h = 0; h < 2; ++h)
#pragma unroll
                        for (int t = 0; t < 4; ++t)
                            if (t < nt)
#pragma unroll
                                for (int gi = 0; gi < G; ++gi) D[gi][t] = mfma64(av[h][t], bv[gi][h], D[gi][t]);
                }
            }
        } else {
            // stored blocks: one slab of 8 columns (the chunk's 64 rows of the
            // block, 4 KB) and the matching 8 rows of x (8 x 16G values) at a
            // time through this wavefront's LDS, copied by global_load_lds
            // (no registers held); the next slab's copies are issued as soon
            // as this slab's operands are in registers, and run while it
            // multiplies. A lane copies 16 bytes per instruction (two rows of
            // a column: the chunk starts on an even row of the block, and a
            // 64-row tile, padded in storage, never splits such a pair; two
            // values of a row of x: nq even, x 16-byte aligned) or else one
            // 4-byte word. Rows past the block's last are clamped to it (or,
            // in pairs, to the chunk's first) and columns past its last to
            // that one: in-bounds copies, zeroed at the operand. Element
            // (i, cc) of a block sits at rowoff(i) + 64 cc.
            using lds_ptr = __attribute__((address_space(3))) void *;
            constexpr int WR = 2 * MQT;     // 4-byte words per row of x
            constexpr int RPI = WAVE / WR;  // rows of x per 4-byte copy instruction
            constexpr int LR = MQT / 2;     // lanes per row of x, 16-byte copies
            constexpr int RPI16 = WAVE / LR;
            double *Sa = Dl;                // [8][64]: column-major slab of A
            double *Sx = Dl + 8 * HM_TILE;  // [8][16G]: rows of x
            const bool x16 = nq % 2 == 0 && ldxt % 2 == 0 &&
                             (reinterpret_cast<uintptr_t>(xt) & 15) == 0;  // wave-uniform
            uint32_t voa[2] = {0, 0};
            // this lane's part of a row of x: row lane / WR (lane / LR) of an
            // instruction; value min(q, nq - 1), half lane % 2 (values
            // min(2 (lane % LR), nq - 2) and the next)
            const uint32_t vox =
                x16 ? 8u * (uint32_t)min(2 * (lane % LR), nq - 2)
                    : 8u * (uint32_t)min((lane % WR) / 2, nq - 1) + 4u * (uint32_t)(lane & 1);
            auto setup = [&](int e) {  // this lane's rows of block e, as byte offsets
                const int bn = rl_i(d_n, e), boff = rl_i(d_off, e), bm = rl_i(d_m, e);
                if ((boff & 1) == 0) {
                    int i = boff + 2 * (lane % 32);
                    if (i > bm - 1) i = boff;
                    voa[0] = 8u * (uint32_t)((i >> 6) * HM_TILE * bn + (i & (HM_TILE - 1)));
                } else {
#pragma unroll
                    for (int hh = 0; hh < 2; ++hh) {
                        const int i = min(boff + 32 * hh + lane / 2, bm - 1);
                        voa[hh] = 8u * (uint32_t)((i >> 6) * HM_TILE * bn + (i & (HM_TILE - 1))) +
                                  4u * (uint32_t)(lane & 1);
                    }
                }
            };
            auto issue = [&](int e, int k0) {
                const int bn = rl_i(d_n, e);
                if (bn == 0) return;  // an empty block (wave-uniform): nothing to copy
                const char *ab = reinterpret_cast<const char *>(p.A + rl_l(d_aoff, e));
                if ((rl_i(d_off, e) & 1) == 0) {
#pragma unroll
                    for (int j = 0; j < 4; ++j) {
                        const int cc = min(k0 + 2 * j + lane / 32, bn - 1);
                        __builtin_amdgcn_global_load_lds(ab + (int64_t)cc * (8 * HM_TILE) + voa[0],
                                                         (lds_ptr)(Sa + 128 * j), 16, 0, 0);
                    }
                } else {
#pragma unroll
                    for (int j = 0; j < 16; ++j) {
                        const int cc = min(k0 + j / 2, bn - 1);
                        __builtin_amdgcn_global_load_lds(ab + (int64_t)cc * (8 * HM_TILE) + voa[j % 2],
                                                         (lds_ptr)(Sa + 32 * j), 4, 0, 0);
                    }
                }
                const char *xb = reinterpret_cast<const char *>(xt + rl_l(d_col0, e) * ldxt);
                if (x16) {
#pragma unroll
                    for (int jx = 0; jx < 8 / RPI16; ++jx) {
                        const int cc = min(k0 + jx * RPI16 + lane / LR, bn - 1);
                        __builtin_amdgcn_global_load_lds(xb + (int64_t)cc * ldxt * 8 + vox,
                                                         (lds_ptr)(Sx + jx * RPI16 * MQT), 16, 0, 0);
                    }
                } else {
#pragma unroll
                    for (int jx = 0; jx < 8 / RPI; ++jx) {
                        const int cc = min(k0 + jx * RPI + lane / WR, bn - 1);
                        __builtin_amdgcn_global_load_lds(xb + (int64_t)cc * ldxt * 8 + vox,
                                                         (lds_ptr)(Sx + jx * RPI * MQT), 4, 0, 0);
                    }
                }
            };
            int e = 0, k0 = 0;
            if (ne > 0) {
                setup(0);
                issue(0, 0);
            }
            while (e < ne) {
                const int bn = rl_i(d_n, e);
                // the slab's copies complete before these reads: the compiler
                // does not count global_load_lds among the writes a LDS read
                // waits for, so the wait is explicit (and a barrier to moving
                // the reads above it)
                asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
                double av[2][4], bv[2][G];
#pragma unroll
                for (int h = 0; h < 2; ++h) {
                    const bool cok = k0 + 4 * h + g < bn;
#pragma unroll
                    for (int t = 0; t < 4; ++t) {
                        const double v = Sa[(4 * h + g) * HM_TILE + 16 * t + j16];
                        av[h][t] = (rok[t] && cok) ? v : 0.0;
                    }
#pragma unroll
                    for (int gi = 0; gi < G; ++gi) {
                        const int q = gi * MQ + j16;
                        const double v = Sx[(4 * h + g) * MQT + q];
                        bv[h][gi] = (cok && q < nq) ? v : 0.0;
                    }
                }
                int e2 = e, k2 = k0 + 8;
                if (k2 >= bn) {
                    ++e2;
                    k2 = 0;
                }
                // the next slab's copies overwrite this one after the reads
                // have returned their data
                asm volatile("s_waitcnt lgkmcnt(0)" ::: "memory");
                if (e2 < ne) {
                    if (e2 != e) setup(e2);
                    issue(e2, k2);
                }
                __builtin_amdgcn_sched_barrier(0);
#pragma unroll
                for (int h = 0; h < 2; ++h)
#pragma unroll
                    for (int t = 0; t < 4; ++t)
                        if (t < nt)
#pragma unroll
                            for (int gi = 0; gi < G; ++gi) D[gi][t] = mfma64(av[h][t], bv[h][gi], D[gi][t]);
                e = e2;
                k0 = k2;
            }
            wave_sync();  // every lane's slab reads done before LDS is reused
        }
    }
    if constexpr (NEST && SLAB) leaf_series_to_lds<RP, G, WS>(p, leaf, z, r, j16, g, Dl);
    // ---- far field
    load_pf();
    if constexpr (NEST) {
        // one FP64 series per row: this lane's T_k (k = g + 4u; T_g and
        // T_{g+4} from the prefix T_0..T_7, then T_{k+8} = 2 T_4 T_{k+4} - T_k)
        // times the leaf's coefficients, read from LDS at each use rather than
        // held across the four row tiles
        const DTNode L = p.xtn[leaf];
#pragma unroll
        for (int t = 0; t < 4; ++t) {
            if (t >= nt) break;
            const double xh = (pf[t] - L.cx) * L.cy;
            const double x2 = xh + xh;
            const double c2 = __builtin_fma(x2, xh, -1.0), c3 = __builtin_fma(x2, c2, -xh);
            const double c4 = __builtin_fma(x2, c3, -c2), c5 = __builtin_fma(x2, c4, -c3);
            const double c6 = __builtin_fma(x2, c5, -c4), c7 = __builtin_fma(x2, c6, -c5);
            double ta = g == 0 ? 1.0 : g == 1 ? xh : g == 2 ? c2 : c3;  // T_g
            double tb = g == 0 ? c4 : g == 1 ? c5 : g == 2 ? c6 : c7;   // T_{g+4}
            const double t4x2 = c4 + c4;
#pragma unroll
            for (int u = 0; u < RP / 4; ++u) {
                const int k = 4 * u + g;
                const double sel = (fok[t] && k < r) ? ta : 0.0;
#pragma unroll
                for (int gi = 0; gi < G; ++gi) {
                    const int q = gi * MQ + j16;
                    const double v = Dl[q * (r + 1) + min(k, r - 1)];
                    D[gi][t] = mfma64(sel, (k < r && q < nq) ? v : 0.0, D[gi][t]);
                }
                const double tn = __builtin_fma(t4x2, tb, -ta);  // T_{k+8}
                ta = tb;
                tb = tn;
            }
        }
    }
    if constexpr (!NEST) {
        // every covering cluster, its records one per lane; the coefficients
        // of the next cluster load while this one is applied
        auto load_bz = [&](double (*bz)[KSEL], int cl) {
#pragma unroll
            for (int gi = 0; gi < G; ++gi) {
                const int q = gi * MQ + j16;
                const double *zc = z + ((int64_t)cl * MQT + q) * r;  // C[k][q]
#pragma unroll
                for (int u = 0; u < KSEL; ++u) {
                    const int k = BF ? (u / 8) * 32 + 8 * g + (u % 8) : 4 * u + g;
                    bz[gi][u] = (k < r && q < nq) ? zc[min(k, r - 1)] : 0.0;
                }
            }
        };
        for (int e0 = c.far_beg; e0 < c.far_end; e0 += WAVE) {
            const int ne = min(WAVE, c.far_end - e0);
            int f_cl = 0;
            double f_cx = 0.0, f_cy = 0.0;
            if (lane < ne) {
                f_cl = fl[e0 + lane].cl;
                const double2 cmv = cmapv[f_cl];
                f_cx = cmv.x;
                f_cy = cmv.y;
            }
            double bz[G][KSEL], bzn[G][KSEL] = {};
            load_bz(bz, rl_i(f_cl, 0));
            for (int e = 0; e < ne; ++e) {
                const double2 cm = make_double2(rl_d(f_cx, e), rl_d(f_cy, e));
                if (e + 1 < ne) load_bz(bzn, rl_i(f_cl, e + 1));
#pragma unroll
                for (int t = 0; t < 4; ++t) {
                    if (t >= nt) break;
                    const double xh = (pf[t] - cm.x) * cm.y;
                    double sel[KSEL];
                    if constexpr (!BF) {
                        // this lane's T_k, k = g + 4u: T_g and T_{g+4} from the prefix
                        // T_0..T_7, then T_{k+4} = 2 T_4 T_k - T_{k-4} (stride-4 recurrence)
                        const double x2 = xh + xh;
                        const double c2 = __builtin_fma(x2, xh, -1.0), c3 = __builtin_fma(x2, c2, -xh);
                        const double c4 = __builtin_fma(x2, c3, -c2), c5 = __builtin_fma(x2, c4, -c3);
                        const double c6 = __builtin_fma(x2, c5, -c4), c7 = __builtin_fma(x2, c6, -c5);
                        double ta = g == 0 ? 1.0 : g == 1 ? xh : g == 2 ? c2 : c3;  // T_g
                        double tb = g == 0 ? c4 : g == 1 ? c5 : g == 2 ? c6 : c7;   // T_{g+4}
                        const double t4x2 = c4 + c4;
#pragma unroll
                        for (int u = 0; u < KSEL; ++u) {
                            sel[u] = (fok[t] && 4 * u + g < r) ? ta : 0.0;
                            const double tn = __builtin_fma(t4x2, tb, -ta);  // T_{g+4u+8}
                            ta = tb;
                            tb = tn;
                        }
                    } else {
#pragma unroll
                        for (int u = 0; u < KSEL; ++u) sel[u] = 0.0;
                        double t0 = 1.0, t1 = xh;
#pragma unroll
                        for (int k = 0; k < RP; ++k) {
                            double tv = k == 0 ? t0 : t1;
                            if (k >= 2) {
                                tv = __builtin_fma(xh + xh, t1, -t0);
                                t0 = t1;
                                t1 = tv;
                            }
                            const bool mine = ((k & 31) >> 3) == g;
                            const int u = (k >> 5) * 8 + (k & 7);
                            if (mine && fok[t] && k < r) sel[u] = tv;
                        }
                    }
                    if constexpr (PREC == 2) {
#pragma unroll
                        for (int u = 0; u < KSEL; ++u)
#pragma unroll
                            for (int gi = 0; gi < G; ++gi) D[gi][t] = mfma64(sel[u], bz[gi][u], D[gi][t]);
                    } else if constexpr (PREC == 0) {
                        mf4_t D32[G];
#pragma unroll
                        for (int gi = 0; gi < G; ++gi) D32[gi] = mf4_t{0.f, 0.f, 0.f, 0.f};
#pragma unroll
                        for (int u = 0; u < KSEL; ++u)
#pragma unroll
                            for (int gi = 0; gi < G; ++gi)
                                D32[gi] = __builtin_amdgcn_mfma_f32_16x16x4f32((float)sel[u], (float)bz[gi][u], D32[gi], 0, 0, 0);
                        // this cluster's FP32 sums into the FP64 accumulators
#pragma unroll
                        for (int gi = 0; gi < G; ++gi) D[gi][t] = add_f32_acc(D[gi][t], D32[gi]);
                    } else {
                        mf4_t D32[G];
#pragma unroll
                        for (int gi = 0; gi < G; ++gi) D32[gi] = mf4_t{0.f, 0.f, 0.f, 0.f};
#pragma unroll
                        for (int s8 = 0; s8 < KSEL / 8; ++s8) {
                            float av[8];
#pragma unroll
                            for (int v = 0; v < 8; ++v) av[v] = (float)sel[s8 * 8 + v];
                            mbf8_t ah, am, al;
                            split8<PREC>(av, ah, am, al);
#pragma unroll
                            for (int gi = 0; gi < G; ++gi) {
                                float bv[8];
#pragma unroll
                                for (int v = 0; v < 8; ++v) bv[v] = (float)bz[gi][s8 * 8 + v];
                                mbf8_t bh, bm, bl;
                                split8<PREC>(bv, bh, bm, bl);
                                D32[gi] = mfma_split<PREC>(ah, am, al, bh, bm, bl, D32[gi]);
                            }
                        }
#pragma unroll
                        for (int gi = 0; gi < G; ++gi) D[gi][t] = add_f32_acc(D[gi][t], D32[gi]);
                    }
                }
#pragma unroll
                for (int gi = 0; gi < G; ++gi)
#pragma unroll
                    for (int u = 0; u < KSEL; ++u) bz[gi][u] = bzn[gi][u];
            }
        }
    }
    if constexpr (NEST) wave_sync();  // the output tiles reuse the coefficients' LDS
    // ---- result, one 16-row tile at a time through LDS
#pragma unroll
    for (int t = 0; t < 4; ++t) {
        if (t >= nt) break;
#pragma unroll
        for (int gi = 0; gi < G; ++gi)
#pragma unroll
            for (int v = 0; v < 4; ++v) Dl[(g + 4 * v) * MQT + gi * MQ + j16] = D[gi][t][v];
        wave_sync();
        if (a.ldy == 1) {
            // a row's right-hand sides are contiguous: each store writes
            // WAVE / MQT whole rows
            constexpr int RPS = WAVE / MQT;
            const int q = lane % MQT;
#pragma unroll
            for (int s = 0; s < 16 / RPS; ++s) {
                const int row = s * RPS + lane / MQT;
                if (t * 16 + row < c.np && q < nq) {
                    double *yo = a.y + (c.p0 + t * 16 + row) * a.incy + q;
                    *yo = (a.beta ? *yo : 0.0) + Dl[row * MQT + q];
                }
            }
        } else {
            const int row = j16;
            if (t * 16 + row < c.np) {
                const int64_t i = c.p0 + t * 16 + row;
#pragma unroll
                for (int s4 = 0; s4 < 4 * G; ++s4) {
                    const int q = g + 4 * s4;
                    if (q < nq) {
                        double *yo = a.y + i * a.incy + q * a.ldy;
                        *yo = (a.beta ? *yo : 0.0) + Dl[row * MQT + q];
                    }
                }
            }
        }
        wave_sync();
    }
}

// MODE 1: stored dense blocks (stored and hybrid plans); MODE 2: matrix-free.
// With the plan's interval trees (p.nest) the far field is nested: column
// leaves' moments, combined up the column tree, translated, and the
// coefficients pushed down the row tree to each chunk's leaf; else every
// covering cluster is applied per level.
template <int RP, int PREC, int MODE, int KID, int G>
hipError_t matvec_m_t(const DevPlan &p, const MvArgs &a, int nq, hipStream_t s) {
    constexpr bool FREE = MODE == 2;
    constexpr int MQT = MQ * G;
    // rows of x are read in place when the right-hand sides of a row are
    // contiguous (ldx = 1, as for a row-major n x k array), else packed
    const double *xr = a.x;
    int64_t ldxr = a.incx;
    if (!(a.ldx == 1 && a.incx >= nq)) {
        if (p.ny > 0)
            k_pack_x<G><<<(int)((p.ny + WG - 1) / WG), WG, 0, s>>>(a.x, a.incx, a.ldx, nq, p.ny, a.xt);
        xr = a.xt;
        ldxr = MQT;
    }
    const bool nest = p.nest && p.nfar > 0;
    using WS = NodeT<PREC>;  // nested w and z (the workspace holds them in its first bytes)
    WS *wn = reinterpret_cast<WS *>(a.w), *zn = reinterpret_cast<WS *>(a.z);
    if (p.nfar > 0 && nest) {
        if (p.nyleaf > 0)
            k_up_m<RP, PREC, G, true, WS><<<p.nyleaf, WAVE, 0, s>>>(
                p.yleaf, p.y, xr, ldxr, nq, p.chebA, a.partial, wn, p.r, p.my);
        // moments: two depths per launch (the even depths, each with the one
        // below it) where the second node's registers fit, else one
        constexpr bool TWO = RP <= 32;
        for (int d = TWO ? (p.nylev - 1) & ~1 : p.nylev - 1; d >= 0; d -= TWO ? 2 : 1) {
            const int off = p.ylev_off_h[d], cnt = p.ylev_off_h[d + 1] - off;
            if (cnt > 0)
                k_m2m<RP, G, WS, TWO><<<nblk(cnt), WG, 0, s>>>(p.ytn, p.ylev + off, cnt, p.nestM,
                                                                p.my, p.chebA, wn, p.r);
        }
        k_translate_m<RP, FREE, KID, G, WS, true><<<nblk(p.nrc), WG, 0, s>>>(p, p.rcl, p.nrc, wn, zn);
        for (int d = 0; d < p.nxlev; ++d) {
            const int off = p.xlev_off_h[d], cnt = p.xlev_off_h[d + 1] - off;
            if (cnt > 0)
                k_l2l<RP, G, WS><<<nblk(cnt), WG, 0, s>>>(p.xtn, p.xlev + off, cnt, p.nestM, zn,
                                                           p.mx, p.r);
        }
    } else if (p.nfar > 0 && p.nctask > 0) {
        k_up_m<RP, PREC, G, false, double><<<p.nctask, WAVE, 0, s>>>(
            p.ctask, p.y, xr, ldxr, nq, p.chebA, a.partial, a.w, p.r, nullptr);
        if (p.nmulti_c > 0)
            k_reduce<double, true, MQT><<<nblk(p.nmulti_c), WG, 0, s>>>(
                p.ccl, p.multi_c, p.nmulti_c, a.partial, p.chebA, a.w, p.r);
        k_translate_m<RP, FREE, KID, G, double, false><<<nblk(p.nrc), WG, 0, s>>>(p, p.rcl, p.nrc, a.w,
                                                                           a.z);
    }
    if (p.nrch > 0) {
        // nested: one leaf series per row, evaluated in FP64 for every PREC
        // (a reduced-precision form saves nothing on one series and costs the
        // near field registers)
        if (nest)
            k_down_m<RP, 2, FREE, KID, G, true, WS><<<nblk(p.nrch), WG, 0, s>>>(
                p, p.rch, p.nrch, p.rnear, p.rfar, p.rmap, p.x, p.y, xr, ldxr, zn, a, nq);
        else
            k_down_m<RP, PREC, FREE, KID, G, false, double><<<nblk(p.nrch), WG, 0, s>>>(
                p, p.rch, p.nrch, p.rnear, p.rfar, p.rmap, p.x, p.y, xr, ldxr, a.z, a, nq);
    }
    return hipGetLastError();
}

template <int RP, int PREC, int G>
hipError_t matvec_m_p(const DevPlan &p, const MvArgs &a, int mode, int nq, hipStream_t s) {
    if (mode != 2) return matvec_m_t<RP, PREC, 1, 0, G>(p, a, nq, s);
    switch (p.kid) {
    case 0: return matvec_m_t<RP, PREC, 2, 0, G>(p, a, nq, s);
    case 1: return matvec_m_t<RP, PREC, 2, 1, G>(p, a, nq, s);
    case 2: return matvec_m_t<RP, PREC, 2, 2, G>(p, a, nq, s);
    default: return matvec_m_t<RP, PREC, 2, 3, G>(p, a, nq, s);
    }
}

template <int RP, int PREC>
hipError_t matvec_m_g(const DevPlan &p, const MvArgs &a, int mode, int nq, hipStream_t s) {
    if (nq > MQ) return matvec_m_p<RP, PREC, 2>(p, a, mode, nq, s);
    return matvec_m_p<RP, PREC, 1>(p, a, mode, nq, s);
}

template <int RP>
hipError_t matvec_m_r(const DevPlan &p, const MvArgs &a, int mode, int prec, int nq,
                      hipStream_t s) {
    if (prec == 0) return matvec_m_g<RP, 0>(p, a, mode, nq, s);
    if (prec == 1) return matvec_m_g<RP, 1>(p, a, mode, nq, s);
    if (prec == 3) return matvec_m_g<RP, 3>(p, a, mode, nq, s);
    return matvec_m_g<RP, 2>(p, a, mode, nq, s);
}

template <class T, int RMAX, int KB>
hipError_t matvec_rk(const DevPlanT<T> &p, const MvArgsT<T> &a, bool tr, int mode, hipStream_t s,
                     const hipEvent_t *ev) {
    if (mode == 0) return matvec_t<T, RMAX, 0, 0, KB>(p, a, tr, s, ev);
    if (mode == 1) return matvec_t<T, RMAX, 1, 0, KB>(p, a, tr, s, ev);
    switch (p.kid) {
    case 0: return matvec_t<T, RMAX, 2, 0, KB>(p, a, tr, s, ev);
    case 1: return matvec_t<T, RMAX, 2, 1, KB>(p, a, tr, s, ev);
    case 2: return matvec_t<T, RMAX, 2, 2, KB>(p, a, tr, s, ev);
    default: return matvec_t<T, RMAX, 2, 3, KB>(p, a, tr, s, ev);
    }
}

template <class T, int RMAX>
hipError_t matvec_r(const DevPlanT<T> &p, const MvArgsT<T> &a, bool tr, int mode, int kb, hipStream_t s,
                    const hipEvent_t *ev) {
    if (tr || kb == 1) return matvec_rk<T, RMAX, 1>(p, a, tr, mode, s, ev);
    if (kb == 2) return matvec_rk<T, RMAX, 2>(p, a, false, mode, s, ev);
    if (kb == 4) return matvec_rk<T, RMAX, 4>(p, a, false, mode, s, ev);
    return matvec_rk<T, RMAX, 8>(p, a, false, mode, s, ev);
}

}  // namespace
}  // namespace hmxk

// per-rank-bucket entry points (inst_r*.hip: FP64 plans, ranks up to 64;
// inst_f32_r*.hip: FP32 plans, ranks up to 32)
#define HMX_DECLARE_BUCKET(R)                                                                   \
    hipError_t hmx_matvec_r##R(const DevPlan &p, const MvArgs &a, bool tr, int mode, int kb,    \
                               hipStream_t s, const hipEvent_t *ev);                            \
    hipError_t hmx_grad_r##R(const DevPlan &p, const MvArgs &a, bool tr, double *y2,            \
                             int64_t incy2, const double *sv, int64_t incs, hipStream_t s);     \
    hipError_t hmx_gen_basis_r##R(const DevPlan &p, hipStream_t s);                              \
    hipError_t hmx_matvec_m_r##R(const DevPlan &p, const MvArgs &a, int mode, int prec, int nq,  \
                                 hipStream_t s);
#define HMX_DECLARE_BUCKET_F32(R)                                                               \
    hipError_t hmx_matvec_f32_r##R(const DevPlanF &p, const MvArgsF &a, bool tr, int mode,      \
                                   int kb, hipStream_t s, const hipEvent_t *ev);                \
    hipError_t hmx_grad_f32_r##R(const DevPlanF &p, const MvArgsF &a, bool tr, float *y2,       \
                                 int64_t incy2, const float *sv, int64_t incs, hipStream_t s);  \
    hipError_t hmx_gen_basis_f32_r##R(const DevPlanF &p, hipStream_t s);
HMX_DECLARE_BUCKET(8)
HMX_DECLARE_BUCKET(16)
HMX_DECLARE_BUCKET(24)
HMX_DECLARE_BUCKET(32)
HMX_DECLARE_BUCKET(48)
HMX_DECLARE_BUCKET(64)
HMX_DECLARE_BUCKET_F32(8)
HMX_DECLARE_BUCKET_F32(16)
HMX_DECLARE_BUCKET_F32(32)
