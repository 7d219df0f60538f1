// kernels.hip -- gfx950 kernels of the KernelMatrix compression and product.
//
// One product y = beta*y + H x runs four passes over the flattened tree:
//   up         per task (a column cluster and <= 8192 of its columns), one
//              workgroup: w_task = B[cols, :]^T x[cols] reduced in registers,
//              across the wavefront by recursive halving and across the four
//              wavefronts in LDS; each task owns one partial slot (no atomics)
//   reduce     per column cluster: w = sum of its task slots
//   translate  per row cluster: z = sum over its far blocks of F w
//   down       per row chunk (one wavefront, one lane per row): near field of
//              the chunk's dense blocks plus U z of every row cluster covering
//              the chunk, written once as beta*y + acc
// The transpose runs the same passes with rows and columns exchanged.
//
// Stored mode streams A, U, V, F from HBM. Matrix-free mode regenerates them
// from the points: kernel values f(x_i, y_j) and the barycentric basis
// lam_k/(x - t_k) / sum_m lam_m/(x - t_m) (src/BarycentricMatrix.jl:248-287).
//
// Every per-wavefront quantity (chunk descriptors, block and cluster records,
// the input vector entries and the z coefficients a lane multiplies with) is
// addressed through wave-uniform indices, so the compiler issues scalar loads
// for it and the vector memory path only carries the streamed factors.
//
// Compression kernels follow the reference's operation order exactly and are
// compiled without FMA contraction, so the stored factors are bit-identical to
// the ones the reference computes from the same points and nodes.
#include "hmx_device.h"

#include <math.h>

namespace {

constexpr int WAVE = 64;
constexpr int WG = 256;
constexpr int WPB = WG / WAVE;  // wavefronts per workgroup

__device__ __forceinline__ int uniform(int v) { return __builtin_amdgcn_readfirstlane(v); }

__device__ __forceinline__ int wave_id() {
    return uniform((int)(blockIdx.x * WPB + threadIdx.x / WAVE));
}

// Reciprocal: hardware estimate refined by two Newton-Raphson steps
// r <- r + r(1 - d r) (Markstein's FMA formulation); error <= 1 ulp.
__device__ __forceinline__ double rcp64(double d) {
    double r = __builtin_amdgcn_rcp(d);
    double e = __builtin_fma(-d, r, 1.0);
    r = __builtin_fma(r, e, r);
    e = __builtin_fma(-d, r, 1.0);
    return __builtin_fma(r, e, r);
}

// f(x, y) with the reference's rounding sequence (examples/Kernel.jl:34-37).
__device__ __forceinline__ double kern_exact(int kid, double x, double y) {
#pragma clang fp contract(off)
    double d = x - y;
    switch (kid) {
    case 0: return 1.0 / d;
    case 1: return 1.0 / (d * d);
    case 2: return 1.0 / ((d * d) * d);
    default: return log(fabs(d));
    }
}

// f(x, y) for the matrix-free passes.
template <int KID>
__device__ __forceinline__ double kern_fast(double x, double y) {
    double d = x - y;
    if constexpr (KID == 0) return rcp64(d);
    else if constexpr (KID == 1) return rcp64(d * d);
    else if constexpr (KID == 2) return rcp64(d * d * d);
    else return log(fabs(d));
}

// ---------------------------------------------------------------- reductions
// Recursive-halving reduction of P values held by every lane: at each xor
// offset O the lane pair exchanges half of its values, so after log2 steps
// every lane holds P/2^s partial sums of distinct indices. Odd counts finish
// with plain xor-reductions. `base` receives the first index a lane holds.
template <int P, int O>
struct Halve {
    static __device__ __forceinline__ void run(double *acc, int lane, int &base) {
        if constexpr (O > 0) {
            if constexpr (P % 2 == 0) {
                constexpr int H = P / 2;
                const bool hi = (lane & O) != 0;
#pragma unroll
                for (int i = 0; i < H; ++i) {
                    double send = hi ? acc[i] : acc[i + H];
                    double keep = hi ? acc[i + H] : acc[i];
                    acc[i] = keep + __shfl_xor(send, O);
                }
                if (hi) base += H;
                Halve<H, O / 2>::run(acc, lane, base);
            } else {
#pragma unroll
                for (int i = 0; i < P; ++i) acc[i] += __shfl_xor(acc[i], O);
                Halve<P, O / 2>::run(acc, lane, base);
            }
        }
    }
};
template <int P, int O>
struct HalveShape {  // values left per lane, and the xor-mask of the plain steps
    static constexpr int left = (O == 0) ? P : (P % 2 == 0 ? HalveShape<P / 2, O / 2>::left
                                                           : HalveShape<P, O / 2>::left);
    static constexpr int mask = (O == 0) ? 0 : (P % 2 == 0 ? HalveShape<P / 2, O / 2>::mask
                                                           : (O | HalveShape<P, O / 2>::mask));
};
template <int P>
struct HalveShape<P, 0> {
    static constexpr int left = P;
    static constexpr int mask = 0;
};

// ------------------------------------------------------------- compression

__global__ __launch_bounds__(WG) void k_gen_dense(DevPlan p) {
    const DDense D = p.dense[blockIdx.x];
    const int32_t tot = D.m * D.n;
    for (int32_t e = threadIdx.x; e < tot; e += WG) {
        int32_t j = e / D.m, i = e - j * D.m;
        p.A[D.aoff + e] = kern_exact(p.kid, p.x[D.row0 + i], p.y[D.col0 + j]);
    }
}

// update! (src/BarycentricMatrix.jl:256-287): one lane per point of a chunk,
// for every cluster covering the chunk.
template <int RMAX>
__global__ __launch_bounds__(WG) void k_gen_basis(const DChunk *ch, int nch, const DFar *fl,
                                                 const DCluster *cl, const double *nodes,
                                                 const double *lam, const double *pts,
                                                 double *B, int r) {
#pragma clang fp contract(off)
    const int wid = wave_id();
    const int lane = threadIdx.x % WAVE;
    if (wid >= nch) return;
    const DChunk c = ch[wid];
    if (lane >= c.np) return;
    const double xi = pts[c.p0 + lane];
    for (int e = c.far_beg; e < c.far_end; ++e) {
        const DFar f = fl[e];
        const DCluster C = cl[f.cl];
        const double *t = nodes + (int64_t)f.cl * r;
        double row[RMAX];
        double s = 0.0;
#pragma unroll
        for (int k = 0; k < RMAX; ++k)
            if (k < r) row[k] = lam[k] * (1.0 / (xi - t[k]));
#pragma unroll
        for (int k = 0; k < RMAX; ++k)
            if (k < r) s += row[k];
        double *out = B + C.boff + f.off + lane;
#pragma unroll
        for (int k = 0; k < RMAX; ++k)
            if (k < r) out[(int64_t)k * C.len] = row[k] / s;
    }
}

// F[m, n] = f(t_m, s_n), src/BarycentricMatrix.jl:168-175
__global__ __launch_bounds__(WG) void k_gen_F(DevPlan p) {
    const DFarBlk b = p.far[blockIdx.x];
    const int r = p.r;
    const double *tx = p.rnodes + (int64_t)b.rc * r;
    const double *ty = p.cnodes + (int64_t)b.cc * r;
    for (int e = threadIdx.x; e < r * r; e += WG) {
        int n = e / r, m = e - n * r;
        p.F[b.foff + e] = kern_exact(p.kid, tx[m], ty[n]);
    }
}

__global__ __launch_bounds__(WG) void k_count_nonfinite(const double *v, int64_t n,
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

// Upward pass: partial[task.slot] = B[task points, :]^T x[task points], B = V
// (forward) or U (transpose), stored or regenerated.
template <int RMAX, bool FREE, int KID>
__global__ __launch_bounds__(WG) void k_up(const DTask *tasks, const DCluster *cl, const double *B,
                                          const double *nodes, const double *lam,
                                          const double *pts, const double *x, int64_t incx,
                                          double *partial, int r) {
    __shared__ double sh[WPB][RMAX];
    const DTask T = tasks[blockIdx.x];
    const DCluster C = cl[T.cl];
    const int tid = threadIdx.x, lane = tid % WAVE, wv = tid / WAVE;
    double acc[RMAX];
#pragma unroll
    for (int k = 0; k < RMAX; ++k) acc[k] = 0.0;
    const int64_t first = C.first + T.off;
    if constexpr (!FREE) {
        const double *bp = B + C.boff + T.off;
        for (int i = tid; i < T.len; i += WG) {
            const double xv = x[(first + i) * incx];
#pragma unroll
            for (int k = 0; k < RMAX; ++k)
                if (k < r) acc[k] = __builtin_fma(bp[i + (int64_t)k * C.len], xv, acc[k]);
        }
    } else {
        const double *t = nodes + (int64_t)T.cl * r;
        for (int i = tid; i < T.len; i += WG) {
            const double xv = x[(first + i) * incx];
            const double pt = pts[first + i];
            double wk[RMAX];
            double S = 0.0;
#pragma unroll
            for (int k = 0; k < RMAX; ++k) {
                wk[k] = k < r ? lam[k] * rcp64(pt - t[k]) : 0.0;
                S += wk[k];
            }
            const double sc = xv * rcp64(S);
#pragma unroll
            for (int k = 0; k < RMAX; ++k) acc[k] = __builtin_fma(wk[k], sc, acc[k]);
        }
    }
    int base = 0;
    Halve<RMAX, WAVE / 2>::run(acc, lane, base);
    constexpr int LEFT = HalveShape<RMAX, WAVE / 2>::left;
    constexpr int MASK = HalveShape<RMAX, WAVE / 2>::mask;
    if ((lane & MASK) == 0) {
#pragma unroll
        for (int i = 0; i < LEFT; ++i) sh[wv][base + i] = acc[i];
    }
    __syncthreads();
    if (tid < r) {
        double s = 0.0;
#pragma unroll
        for (int q = 0; q < WPB; ++q) s += sh[q][tid];
        partial[T.slot * r + tid] = s;
    }
}

// w[c] = sum over the task slots of cluster c
__global__ __launch_bounds__(WG) void k_reduce(const DCluster *cl, int ncl, const double *partial,
                                              double *w, int r) {
    const int wid = wave_id();
    const int lane = threadIdx.x % WAVE;
    if (wid >= ncl) return;
    const DCluster C = cl[wid];
    const int G = WAVE / r;  // lane groups; lane = g*r + k
    double acc = 0.0;
    if (lane < G * r) {
        const int k = lane % r, g = lane / r;
        for (int t = g; t < C.nslots; t += G) acc += partial[(C.pbase + t) * r + k];
    }
    double tot = acc;
    for (int g = 1; g < G; ++g) tot += __shfl(acc, (lane + g * r) & (WAVE - 1));
    if (lane < r) w[(int64_t)wid * r + lane] = tot;
}

// z[d] = sum over far blocks b of cluster d of F_b w[other side of b]
// (forward: F w[cc]; transpose: F^T w[rc]).
template <bool FREE, int KID, bool TRANS>
__global__ __launch_bounds__(WG) void k_translate(DevPlan p, const DCluster *cl, int ncl,
                                                 const int32_t *far_list, const double *w,
                                                 double *z) {
    const int wid = wave_id();
    const int lane = threadIdx.x % WAVE;
    if (wid >= ncl) return;
    const int r = p.r;
    const DCluster C = cl[wid];
    const bool act = lane < r;
    double acc = 0.0;
    for (int e = C.far_beg; e < C.far_end; ++e) {
        const DFarBlk b = p.far[far_list[e]];
        const int src = TRANS ? b.rc : b.cc;
        const double *ws = w + (int64_t)src * r;
        if constexpr (!FREE) {
            const double *F = p.F + b.foff;
            for (int l = 0; l < r; ++l) {
                double f = act ? (TRANS ? F[l + (int64_t)r * lane] : F[lane + (int64_t)r * l]) : 0.0;
                acc = __builtin_fma(f, ws[l], acc);
            }
        } else {
            // forward: lane = m, F[m, l] = f(tx_m, ty_l); transpose: lane = l, sum over m
            const double *own = TRANS ? p.cnodes + (int64_t)b.cc * r : p.rnodes + (int64_t)b.rc * r;
            const double *oth = TRANS ? p.rnodes + (int64_t)b.rc * r : p.cnodes + (int64_t)b.cc * r;
            const double to = act ? own[lane] : 0.0;
            for (int l = 0; l < r; ++l) {
                double f = 0.0;
                if (act) f = TRANS ? kern_fast<KID>(oth[l], to) : kern_fast<KID>(to, oth[l]);
                acc = __builtin_fma(f, ws[l], acc);
            }
        }
    }
    if (act) z[(int64_t)wid * r + lane] = acc;
}

// Downward pass. Forward: output = rows; near blocks read by rows; far basis U.
// Transpose: output = columns; near blocks read by columns; far basis V.
template <bool FREE, int KID, bool TRANS>
__global__ __launch_bounds__(WG) void k_down(DevPlan p, const DChunk *ch, int nch,
                                            const DNear *nl, const DFar *fl,
                                            const DCluster *cl, const double *B,
                                            const double *nodes, const double *pts_out,
                                            const double *pts_in, const double *z, MvArgs a) {
    constexpr int UN = 16;
    const int wid = wave_id();
    const int lane = threadIdx.x % WAVE;
    if (wid >= nch) return;
    const int r = p.r;
    const DChunk c = ch[wid];
    const bool act = lane < c.np;
    double po = 0.0;
    if constexpr (FREE) po = act ? pts_out[c.p0 + lane] : 0.0;
    double acc = 0.0;

    // ---- near field
    for (int e = c.near_beg; e < c.near_end; ++e) {
        const DNear s = nl[e];
        const DDense D = p.dense[s.blk];
        if constexpr (!TRANS) {
            const double *xin = a.x + D.col0 * a.incx;
            if constexpr (!FREE) {
                const double *Aj = p.A + D.aoff + s.off + lane;
                int j = 0;
                for (; j + UN <= D.n; j += UN) {
                    double av[UN];
#pragma unroll
                    for (int u = 0; u < UN; ++u) av[u] = act ? Aj[(int64_t)(j + u) * D.m] : 0.0;
#pragma unroll
                    for (int u = 0; u < UN; ++u) acc = __builtin_fma(av[u], xin[(int64_t)(j + u) * a.incx], acc);
                }
                for (; j < D.n; ++j)
                    acc = __builtin_fma(act ? Aj[(int64_t)j * D.m] : 0.0, xin[(int64_t)j * a.incx], acc);
            } else {
                const double *yin = pts_in + D.col0;
                for (int j = 0; j < D.n; ++j) {
                    const double f = kern_fast<KID>(po, yin[j]);
                    acc = __builtin_fma(act ? f : 0.0, xin[(int64_t)j * a.incx], acc);
                }
            }
        } else {
            const double *xin = a.x + D.row0 * a.incx;
            if constexpr (!FREE) {
                // lanes run down the rows of the block, 64 accumulators for the
                // chunk's columns, then a recursive-halving transpose-reduce
                double accT[WAVE];
#pragma unroll
                for (int q = 0; q < WAVE; ++q) accT[q] = 0.0;
                const double *Ac = p.A + D.aoff + (int64_t)s.off * D.m;
                for (int ib = 0; ib < D.m; ib += WAVE) {
                    const bool ra = ib + lane < D.m;
                    const double xv = ra ? xin[(int64_t)(ib + lane) * a.incx] : 0.0;
#pragma unroll
                    for (int q = 0; q < WAVE; ++q)
                        if (q < c.np) accT[q] += (ra ? Ac[ib + lane + (int64_t)q * D.m] : 0.0) * xv;
                }
                int base = 0;
                Halve<WAVE, WAVE / 2>::run(accT, lane, base);
                acc += accT[0];
            } else {
                const double *pin = pts_in + D.row0;
                for (int i = 0; i < D.m; ++i) {
                    const double f = kern_fast<KID>(pin[i], po);
                    acc = __builtin_fma(act ? f : 0.0, xin[(int64_t)i * a.incx], acc);
                }
            }
        }
    }

    // ---- far field: acc += B[chunk rows, :] z[cluster]
    for (int e = c.far_beg; e < c.far_end; ++e) {
        const DFar f = fl[e];
        const double *zc = z + (int64_t)f.cl * r;
        if constexpr (!FREE) {
            const DCluster C = cl[f.cl];
            const double *bp = B + C.boff + f.off + lane;
            int k = 0;
            for (; k + 8 <= r; k += 8) {
                double bv[8];
#pragma unroll
                for (int u = 0; u < 8; ++u) bv[u] = act ? bp[(int64_t)(k + u) * C.len] : 0.0;
#pragma unroll
                for (int u = 0; u < 8; ++u) acc = __builtin_fma(bv[u], zc[k + u], acc);
            }
            for (; k < r; ++k) acc = __builtin_fma(act ? bp[(int64_t)k * C.len] : 0.0, zc[k], acc);
        } else {
            const double *tc = nodes + (int64_t)f.cl * r;
            double num = 0.0, den = 0.0;
            for (int k = 0; k < r; ++k) {
                const double wk = p.lam[k] * rcp64(po - tc[k]);
                den += wk;
                num = __builtin_fma(wk, zc[k], num);
            }
            if (act) acc += num * rcp64(den);
        }
    }
    if (act) {
        double *yo = a.y + (c.p0 + lane) * a.incy;
        *yo = (a.beta ? *yo : 0.0) + acc;
    }
}

// getindex: one thread per requested entry
__global__ __launch_bounds__(WG) void k_entries(DevPlan p, int64_t n, const int64_t *ii,
                                               const int64_t *jj, const int32_t *kind,
                                               const int32_t *id, double *out) {
#pragma clang fp contract(off)
    const int64_t t = (int64_t)blockIdx.x * WG + threadIdx.x;
    if (t >= n) return;
    const int64_t i = ii[t], j = jj[t];
    const int r = p.r;
    double v = 0.0;
    if (kind[t] == 1) {
        const DDense D = p.dense[id[t]];
        v = p.A ? p.A[D.aoff + (i - D.row0) + (int64_t)D.m * (j - D.col0)] : kern_exact(p.kid, p.x[i], p.y[j]);
    } else if (kind[t] == 2) {
        const DFarBlk b = p.far[id[t]];
        const DCluster R = p.rcl[b.rc], Cc = p.ccl[b.cc];
        double ur[64], vr[64];
        if (p.U) {
            for (int k = 0; k < r; ++k) {
                ur[k] = p.U[R.boff + (i - R.first) + (int64_t)k * R.len];
                vr[k] = p.V[Cc.boff + (j - Cc.first) + (int64_t)k * Cc.len];
            }
        } else {
            double su = 0.0, sv = 0.0;
            for (int k = 0; k < r; ++k) {
                ur[k] = p.lam[k] * (1.0 / (p.x[i] - p.rnodes[(int64_t)b.rc * r + k]));
                vr[k] = p.lam[k] * (1.0 / (p.y[j] - p.cnodes[(int64_t)b.cc * r + k]));
                su += ur[k];
                sv += vr[k];
            }
            for (int k = 0; k < r; ++k) { ur[k] /= su; vr[k] /= sv; }
        }
        for (int k = 0; k < r; ++k) {
            double tmp = 0.0;
            for (int l = 0; l < r; ++l) {
                double F = p.F ? p.F[b.foff + k + (int64_t)r * l]
                               : kern_exact(p.kid, p.rnodes[(int64_t)b.rc * r + k],
                                            p.cnodes[(int64_t)b.cc * r + l]);
                tmp += F * vr[l];
            }
            v += ur[k] * tmp;
        }
    }
    out[t] = v;
}

inline int nblk(int64_t waves) { return (int)((waves + WPB - 1) / WPB); }

template <int RMAX>
hipError_t gen_basis_r(const DevPlan &p, hipStream_t s) {
    if (p.ncch > 0)
        k_gen_basis<RMAX><<<nblk(p.ncch), WG, 0, s>>>(p.cch, p.ncch, p.cfar, p.ccl, p.cnodes, p.lam,
                                                     p.y, p.V, p.r);
    if (p.nrch > 0)
        k_gen_basis<RMAX><<<nblk(p.nrch), WG, 0, s>>>(p.rch, p.nrch, p.rfar, p.rcl, p.rnodes, p.lam,
                                                     p.x, p.U, p.r);
    return hipGetLastError();
}

template <int RMAX, bool FREE, int KID>
hipError_t matvec_t(const DevPlan &p, const MvArgs &a, bool tr, hipStream_t s, const hipEvent_t *ev) {
    auto mark = [&](int q) {
        if (ev) (void)hipEventRecord(ev[q], s);
    };
    // "in" = the side summed over (columns forward), "out" = the side written
    const DTask *in_task = tr ? p.rtask : p.ctask;
    const int in_ntask = tr ? p.nrtask : p.nctask;
    const DCluster *in_cl = tr ? p.rcl : p.ccl;
    const int in_ncl = tr ? p.nrc : p.ncc;
    const double *in_B = tr ? p.U : p.V;
    const double *in_nodes = tr ? p.rnodes : p.cnodes;
    const double *in_pts = tr ? p.x : p.y;
    const DChunk *out_ch = tr ? p.cch : p.rch;
    const int out_n = tr ? p.ncch : p.nrch;
    const DNear *out_near = tr ? p.cnear : p.rnear;
    const DFar *out_far = tr ? p.cfar : p.rfar;
    const DCluster *out_cl = tr ? p.ccl : p.rcl;
    const int out_ncl = tr ? p.ncc : p.nrc;
    const double *out_B = tr ? p.V : p.U;
    const double *out_nodes = tr ? p.cnodes : p.rnodes;
    const double *out_pts = tr ? p.y : p.x;
    const int32_t *out_far_list = tr ? p.far_by_cc : p.far_by_rc;

    mark(0);
    if (p.nfar > 0 && in_ntask > 0) {
        k_up<RMAX, FREE, KID><<<in_ntask, WG, 0, s>>>(in_task, in_cl, in_B, in_nodes, p.lam,
                                                       in_pts, a.x, a.incx, a.partial, p.r);
        mark(1);
        k_reduce<<<nblk(in_ncl), WG, 0, s>>>(in_cl, in_ncl, a.partial, a.w, p.r);
        mark(2);
        if (tr)
            k_translate<FREE, KID, true><<<nblk(out_ncl), WG, 0, s>>>(p, out_cl, out_ncl,
                                                                     out_far_list, a.w, a.z);
        else
            k_translate<FREE, KID, false><<<nblk(out_ncl), WG, 0, s>>>(p, out_cl, out_ncl,
                                                                      out_far_list, a.w, a.z);
    } else {
        mark(1);
        mark(2);
    }
    mark(3);
    if (out_n > 0) {
        if (tr)
            k_down<FREE, KID, true><<<nblk(out_n), WG, 0, s>>>(p, out_ch, out_n, out_near, out_far,
                                                              out_cl, out_B, out_nodes, out_pts,
                                                              in_pts, a.z, a);
        else
            k_down<FREE, KID, false><<<nblk(out_n), WG, 0, s>>>(p, out_ch, out_n, out_near, out_far,
                                                               out_cl, out_B, out_nodes, out_pts,
                                                               in_pts, a.z, a);
    }
    mark(4);
    return hipGetLastError();
}

template <int RMAX, bool FREE>
hipError_t matvec_k(const DevPlan &p, const MvArgs &a, bool tr, hipStream_t s, const hipEvent_t *ev) {
    if constexpr (!FREE) {
        return matvec_t<RMAX, false, 0>(p, a, tr, s, ev);
    } else {
        switch (p.kid) {
        case 0: return matvec_t<RMAX, true, 0>(p, a, tr, s, ev);
        case 1: return matvec_t<RMAX, true, 1>(p, a, tr, s, ev);
        case 2: return matvec_t<RMAX, true, 2>(p, a, tr, s, ev);
        default: return matvec_t<RMAX, true, 3>(p, a, tr, s, ev);
        }
    }
}

template <int RMAX>
hipError_t matvec_r(const DevPlan &p, const MvArgs &a, bool tr, bool free_, hipStream_t s,
                    const hipEvent_t *ev) {
    return free_ ? matvec_k<RMAX, true>(p, a, tr, s, ev) : matvec_k<RMAX, false>(p, a, tr, s, ev);
}

}  // namespace

hipError_t hmx_launch_compress(const DevPlan &p, hipStream_t s) {
    if (p.ndense > 0 && p.A) k_gen_dense<<<p.ndense, WG, 0, s>>>(p);
    if (p.nfar > 0 && p.F) k_gen_F<<<p.nfar, WG, 0, s>>>(p);
    hipError_t err = hipGetLastError();
    if (err != hipSuccess || !p.U) return err;
    if (p.r <= 8) return gen_basis_r<8>(p, s);
    if (p.r <= 16) return gen_basis_r<16>(p, s);
    if (p.r <= 24) return gen_basis_r<24>(p, s);
    if (p.r <= 32) return gen_basis_r<32>(p, s);
    if (p.r <= 48) return gen_basis_r<48>(p, s);
    return gen_basis_r<64>(p, s);
}

hipError_t hmx_launch_matvec(const DevPlan &p, const MvArgs &a, bool tr, bool free_,
                             hipStream_t s, const hipEvent_t *ev) {
    if (p.r <= 8) return matvec_r<8>(p, a, tr, free_, s, ev);
    if (p.r <= 16) return matvec_r<16>(p, a, tr, free_, s, ev);
    if (p.r <= 24) return matvec_r<24>(p, a, tr, free_, s, ev);
    if (p.r <= 32) return matvec_r<32>(p, a, tr, free_, s, ev);
    if (p.r <= 48) return matvec_r<48>(p, a, tr, free_, s, ev);
    return matvec_r<64>(p, a, tr, free_, s, ev);
}

hipError_t hmx_launch_entries(const DevPlan &p, int64_t n, const int64_t *i, const int64_t *j,
                              const int32_t *kind, const int32_t *id, double *out,
                              hipStream_t s) {
    if (n <= 0) return hipSuccess;
    k_entries<<<(int)((n + WG - 1) / WG), WG, 0, s>>>(p, n, i, j, kind, id, out);
    return hipGetLastError();
}

hipError_t hmx_launch_count_nonfinite(const double *v, int64_t n, unsigned long long *blockcount,
                                      hipStream_t s) {
    // caller provides 1024 slots
    if (n <= 0) return hipSuccess;
    int64_t nb = (n + WG - 1) / WG;
    if (nb > 1024) nb = 1024;
    k_count_nonfinite<<<(int)nb, WG, 0, s>>>(v, n, blockcount);
    return hipGetLastError();
}
