// kernels.hip -- public launchers of the gfx950 kernels: dispatch on the rank
// bucket to the instantiations in inst_r*.hip (FP64 plans: 8, 16, 24, 32, 48,
// 64) and inst_f32_r*.hip (FP32 plans: 8, 16, 32).
#include "hmx.h"
#include "kernels_impl.h"

hipError_t hmx_launch_compress(const DevPlan &p, hipStream_t s) {
    // HM_USER plans take A and F from the caller (hm_plan_set_leaf_data)
    const bool gen = p.kid != HM_USER;
    if (gen && p.ndense > 0 && p.A) hmxk::k_gen_dense<double><<<p.ndense, hmxk::WG, 0, s>>>(p);
    if (gen && p.nfar > 0 && p.F) hmxk::k_gen_F<double><<<p.nfar, hmxk::WG, 0, s>>>(p);
    hipError_t err = hipGetLastError();
    if (err != hipSuccess || (!p.U && !p.V)) return err;
    if (p.r <= 8) return hmx_gen_basis_r8(p, s);
    if (p.r <= 16) return hmx_gen_basis_r16(p, s);
    if (p.r <= 24) return hmx_gen_basis_r24(p, s);
    if (p.r <= 32) return hmx_gen_basis_r32(p, s);
    if (p.r <= 48) return hmx_gen_basis_r48(p, s);
    return hmx_gen_basis_r64(p, s);
}

hipError_t hmx_launch_matvec(const DevPlan &p, const MvArgs &a, bool tr, int mode, int kb,
                             hipStream_t s, const hipEvent_t *ev) {
    if (p.r <= 8) return hmx_matvec_r8(p, a, tr, mode, kb, s, ev);
    if (p.r <= 16) return hmx_matvec_r16(p, a, tr, mode, kb, s, ev);
    if (p.r <= 24) return hmx_matvec_r24(p, a, tr, mode, kb, s, ev);
    if (p.r <= 32) return hmx_matvec_r32(p, a, tr, mode, kb, s, ev);
    if (p.r <= 48) return hmx_matvec_r48(p, a, tr, mode, kb, s, ev);
    return hmx_matvec_r64(p, a, tr, mode, kb, s, ev);
}

hipError_t hmx_launch_grad(const DevPlan &p, const MvArgs &a, bool tr, double *y2, int64_t incy2,
                           const double *sv, int64_t incs, hipStream_t s) {
    if (p.r <= 8) return hmx_grad_r8(p, a, tr, y2, incy2, sv, incs, s);
    if (p.r <= 16) return hmx_grad_r16(p, a, tr, y2, incy2, sv, incs, s);
    if (p.r <= 24) return hmx_grad_r24(p, a, tr, y2, incy2, sv, incs, s);
    if (p.r <= 32) return hmx_grad_r32(p, a, tr, y2, incy2, sv, incs, s);
    if (p.r <= 48) return hmx_grad_r48(p, a, tr, y2, incy2, sv, incs, s);
    return hmx_grad_r64(p, a, tr, y2, incy2, sv, incs, s);
}

hipError_t hmx_launch_entries(const DevPlan &p, int64_t n, const int64_t *i, const int64_t *j,
                              const int32_t *kind, const int32_t *id, double *out,
                              hipStream_t s) {
    if (n <= 0) return hipSuccess;
    hmxk::k_entries<double><<<(int)((n + hmxk::WG - 1) / hmxk::WG), hmxk::WG, 0, s>>>(p, n, i, j, kind, id, out);
    return hipGetLastError();
}

hipError_t hmx_launch_count_nonfinite(const double *v, int64_t n, unsigned long long *blockcount,
                                      hipStream_t s) {
    // caller provides 1024 slots
    if (n <= 0) return hipSuccess;
    int64_t nb = (n + hmxk::WG - 1) / hmxk::WG;
    if (nb > 1024) nb = 1024;
    hmxk::k_count_nonfinite<double><<<(int)nb, hmxk::WG, 0, s>>>(v, n, blockcount);
    return hipGetLastError();
}

hipError_t hmx_launch_matvec_m(const DevPlan &p, const MvArgs &a, int mode, int prec, int nq,
                               hipStream_t s) {
    if (p.r <= 8) return hmx_matvec_m_r8(p, a, mode, prec, nq, s);
    if (p.r <= 16) return hmx_matvec_m_r16(p, a, mode, prec, nq, s);
    if (p.r <= 24) return hmx_matvec_m_r24(p, a, mode, prec, nq, s);
    if (p.r <= 32) return hmx_matvec_m_r32(p, a, mode, prec, nq, s);
    if (p.r <= 48) return hmx_matvec_m_r48(p, a, mode, prec, nq, s);
    return hmx_matvec_m_r64(p, a, mode, prec, nq, s);
}

// ------------------------------------------------------------- FP32 plans

hipError_t hmx_launch_compress(const DevPlanF &p, hipStream_t s) {
    const bool gen = p.kid != HM_USER;
    if (gen && p.ndense > 0 && p.A) hmxk::k_gen_dense<float><<<p.ndense, hmxk::WG, 0, s>>>(p);
    if (gen && p.nfar > 0 && p.F) hmxk::k_gen_F<float><<<p.nfar, hmxk::WG, 0, s>>>(p);
    hipError_t err = hipGetLastError();
    if (err != hipSuccess || (!p.U && !p.V)) return err;
    if (p.r <= 8) return hmx_gen_basis_f32_r8(p, s);
    if (p.r <= 16) return hmx_gen_basis_f32_r16(p, s);
    if (p.r <= 32) return hmx_gen_basis_f32_r32(p, s);
    return hipErrorInvalidValue;
}

hipError_t hmx_launch_matvec(const DevPlanF &p, const MvArgsF &a, bool tr, int mode, int kb,
                             hipStream_t s, const hipEvent_t *ev) {
    if (p.r <= 8) return hmx_matvec_f32_r8(p, a, tr, mode, kb, s, ev);
    if (p.r <= 16) return hmx_matvec_f32_r16(p, a, tr, mode, kb, s, ev);
    if (p.r <= 32) return hmx_matvec_f32_r32(p, a, tr, mode, kb, s, ev);
    return hipErrorInvalidValue;
}

hipError_t hmx_launch_grad(const DevPlanF &p, const MvArgsF &a, bool tr, float *y2, int64_t incy2,
                           const float *sv, int64_t incs, hipStream_t s) {
    if (p.r <= 8) return hmx_grad_f32_r8(p, a, tr, y2, incy2, sv, incs, s);
    if (p.r <= 16) return hmx_grad_f32_r16(p, a, tr, y2, incy2, sv, incs, s);
    if (p.r <= 32) return hmx_grad_f32_r32(p, a, tr, y2, incy2, sv, incs, s);
    return hipErrorInvalidValue;
}

hipError_t hmx_launch_entries(const DevPlanF &p, int64_t n, const int64_t *i, const int64_t *j,
                              const int32_t *kind, const int32_t *id, float *out, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    hmxk::k_entries<float><<<(int)((n + hmxk::WG - 1) / hmxk::WG), hmxk::WG, 0, s>>>(p, n, i, j, kind, id, out);
    return hipGetLastError();
}

hipError_t hmx_launch_count_nonfinite(const float *v, int64_t n, unsigned long long *blockcount,
                                      hipStream_t s) {
    if (n <= 0) return hipSuccess;
    int64_t nb = (n + hmxk::WG - 1) / hmxk::WG;
    if (nb > 1024) nb = 1024;
    hmxk::k_count_nonfinite<float><<<(int)nb, hmxk::WG, 0, s>>>(v, n, blockcount);
    return hipGetLastError();
}

// ------------------------------------------------------- diagonal scaling

namespace {

// A[i, j] *= b[col0 + j] (side 0) or b[row0 + i] (side 1), one block per workgroup
__global__ __launch_bounds__(hmxk::WG) void k_scale_dense(DevPlan p, int side, const double *b) {
    const DDense D = p.dense[blockIdx.x];
    const int32_t tot = D.m * D.n;
    for (int32_t e = threadIdx.x; e < tot; e += hmxk::WG) {
        const int32_t j = e / D.m, i = e - j * D.m;
        const double s = side == 0 ? b[D.col0 + j] : b[D.row0 + i];
        double *a = p.A + D.aoff + tile_index(i, j, D.n);
        *a = *a * s;
    }
}

// B[i, k] *= b[first + i] for every basis row of one cluster
__global__ __launch_bounds__(hmxk::WG) void k_scale_basis(const DCluster *cl, double *B, int r,
                                                         const double *b) {
    const DCluster C = cl[blockIdx.x];
    const int32_t tot = C.len * r;
    for (int32_t e = threadIdx.x; e < tot; e += hmxk::WG) {
        const int32_t k = e / C.len, i = e - k * C.len;
        double *v = B + C.boff + tile_index(i, k, r);
        *v = *v * b[C.first + i];
    }
}

__global__ void k_zero_status(int32_t *status) { *status = 0; }

// New point values: each side's points are copied into the plan's array (d)
// and checked as they are read: monotone and finite (t < n - 1 compares t and
// t + 1), then every split of its tree. Workgroups [0, bx) take x's points and
// splits, the rest y's; any failing thread stores 1 (the status is zeroed
// before the launch).
template <class T>
__global__ __launch_bounds__(hmxk::WG) void k_copy_check_points(
    const T *x, T *dx, int64_t nx, const DSplit *xs, int64_t nxs, const T *y, T *dy, int64_t ny,
    const DSplit *ys, int64_t nys, int64_t bx, int32_t *status) {
    const bool sx = (int64_t)blockIdx.x < bx;  // workgroup-uniform
    const T *p = sx ? x : y;
    T *d = sx ? dx : dy;
    const int64_t n = sx ? nx : ny, nsp = sx ? nxs : nys;
    const DSplit *sp = sx ? xs : ys;
    const int64_t t = ((int64_t)blockIdx.x - (sx ? 0 : bx)) * hmxk::WG + threadIdx.x;
    bool bad = false;
    if (t < n) {
        const T v0 = p[t];
        d[t] = v0;
        const double v = (double)v0;
        bad = !isfinite(v) || (t + 1 < n && !(v >= (double)p[t + 1]));
    } else if (t < n + nsp) {
        const DSplit q = sp[t - n];
        if (q.len == 0) {
            bad = !(q.first < n && (double)p[q.first] < q.mid);
        } else {
            if (q.at > q.first) bad |= !((double)p[q.at - 1] >= q.mid);
            if (q.at < q.first + q.len) bad |= !((double)p[q.at] < q.mid);
        }
    }
    if (bad) *status = 1;
}

// one workgroup per job of hm_plan_add_col
__global__ __launch_bounds__(hmxk::WG) void k_add_col(DevPlan p, AddColJobs J, const double *u) {
    const AddColJob q = J.job[blockIdx.x];
    if (q.kind == 3) {
        for (int32_t i = threadIdx.x; i < q.m; i += hmxk::WG) {
            double *a = p.A + q.aoff + tile_index(i, q.jl, q.n);
            *a = *a + u[q.row0 + i];
        }
    } else {
        for (int32_t i = threadIdx.x; i < q.m; i += hmxk::WG)
            p.U[q.uoff + tile_index(i, q.c, p.r)] = u[q.row0 + i];
        if (threadIdx.x == 0) {
            p.V[q.voff + tile_index(q.jl, q.c, p.r)] = 1.0;
            p.F[q.foff + q.c + (int64_t)p.r * q.c] = 1.0;
        }
    }
}

}  // namespace

template <class T>
hipError_t hmx_launch_update_points(const T *x, T *dx, int64_t nx, const DSplit *xs, int64_t nxs,
                                    const T *y, T *dy, int64_t ny, const DSplit *ys, int64_t nys,
                                    int32_t *status, hipStream_t s) {
    k_zero_status<<<1, 1, 0, s>>>(status);
    const int64_t bx = (nx + nxs + hmxk::WG - 1) / hmxk::WG, by = (ny + nys + hmxk::WG - 1) / hmxk::WG;
    if (bx + by > 0)
        k_copy_check_points<T><<<(unsigned)(bx + by), hmxk::WG, 0, s>>>(x, dx, nx, xs, nxs, y, dy,
                                                                        ny, ys, nys, bx, status);
    return hipGetLastError();
}
template hipError_t hmx_launch_update_points<double>(const double *, double *, int64_t,
                                                     const DSplit *, int64_t, const double *,
                                                     double *, int64_t, const DSplit *, int64_t,
                                                     int32_t *, hipStream_t);
template hipError_t hmx_launch_update_points<float>(const float *, float *, int64_t,
                                                    const DSplit *, int64_t, const float *, float *,
                                                    int64_t, const DSplit *, int64_t, int32_t *,
                                                    hipStream_t);

hipError_t hmx_launch_refit(const DevPlan &p, hipStream_t s) {
    if (p.kid != HM_USER && p.ndense > 0 && p.A) hmxk::k_gen_dense<double><<<p.ndense, hmxk::WG, 0, s>>>(p);
    hipError_t err = hipGetLastError();
    if (err != hipSuccess || (!p.U && !p.V)) return err;
    if (p.r <= 8) return hmx_gen_basis_r8(p, s);
    if (p.r <= 16) return hmx_gen_basis_r16(p, s);
    if (p.r <= 24) return hmx_gen_basis_r24(p, s);
    if (p.r <= 32) return hmx_gen_basis_r32(p, s);
    if (p.r <= 48) return hmx_gen_basis_r48(p, s);
    return hmx_gen_basis_r64(p, s);
}

hipError_t hmx_launch_refit(const DevPlanF &p, hipStream_t s) {
    if (p.kid != HM_USER && p.ndense > 0 && p.A) hmxk::k_gen_dense<float><<<p.ndense, hmxk::WG, 0, s>>>(p);
    hipError_t err = hipGetLastError();
    if (err != hipSuccess || (!p.U && !p.V)) return err;
    if (p.r <= 8) return hmx_gen_basis_f32_r8(p, s);
    if (p.r <= 16) return hmx_gen_basis_f32_r16(p, s);
    if (p.r <= 32) return hmx_gen_basis_f32_r32(p, s);
    return hipErrorInvalidValue;
}

hipError_t hmx_launch_add_col(const DevPlan &p, const AddColJobs &J, const double *u, hipStream_t s) {
    if (J.count > 0) k_add_col<<<J.count, hmxk::WG, 0, s>>>(p, J, u);
    return hipGetLastError();
}

hipError_t hmx_launch_scale(const DevPlan &p, int side, const double *b, hipStream_t s) {
    if (p.ndense > 0 && p.A) k_scale_dense<<<p.ndense, hmxk::WG, 0, s>>>(p, side, b);
    if (side == 0 && p.ncc > 0 && p.V) k_scale_basis<<<p.ncc, hmxk::WG, 0, s>>>(p.ccl, p.V, p.r, b);
    if (side == 1 && p.nrc > 0 && p.U) k_scale_basis<<<p.nrc, hmxk::WG, 0, s>>>(p.rcl, p.U, p.r, b);
    return hipGetLastError();
}
