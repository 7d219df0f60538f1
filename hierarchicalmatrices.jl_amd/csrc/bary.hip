// bary.hip -- the standalone barycentric pieces of src/BarycentricMatrix.jl on
// the device, outside a KernelMatrix plan:
//   hm_bary_evaluate   evaluate(B::BarycentricPoly2D, x, y) (:180-205) at many
//                      points, one thread per point
//   hm_bary_basis      the update! basis of one BarycentricMatrix2D side
//                      (:248-297): B[i, m] = (lambda_m / (p_i - t_m)) / sum_k
//                      lambda_k / (p_i - t_k), one thread per point
// Both follow the reference's operation order with IEEE division and no FMA
// contraction, so their values are those the reference computes.
#include "hmx.h"

#include <hip/hip_runtime.h>

#include <cstdio>

int hmx_set_error(int code, const char *msg);

namespace {

constexpr int BWG = 256;
constexpr int BASIS_RMAX = 64;

template <class T>
__global__ __launch_bounds__(BWG) void k_bary_evaluate(int r, const T *xr, const T *yr,
                                                       const T *lx, const T *ly, const T *F,
                                                       int64_t n, const T *px, const T *py,
                                                       T *out) {
#pragma clang fp contract(off)
    const int64_t t = (int64_t)blockIdx.x * BWG + threadIdx.x;
    if (t >= n) return;
    const T x = px[t], y = py[t];
    T temp1 = T(0), temp2 = T(0), ret = T(0);
    for (int p = 0; p < r; ++p) temp1 += lx[p] * (T(1) / (x - xr[p]));
    for (int q = 0; q < r; ++q) temp2 += ly[q] * (T(1) / (y - yr[q]));
    const T it1 = T(1) / temp1;
    for (int c = 0; c < r; ++c) {
        T temp3 = T(0);
        for (int m = 0; m < r; ++m) temp3 += lx[m] * F[m + (int64_t)r * c] * (T(1) / (x - xr[m]));
        temp3 *= it1;
        ret += ly[c] * temp3 * (T(1) / (y - yr[c]));
    }
    ret *= T(1) / temp2;
    out[t] = ret;
}

template <class T>
__global__ __launch_bounds__(BWG) void k_bary_basis(int r, const T *nodes, const T *lam, int64_t n,
                                                    const T *pts, T *B, int64_t ldb) {
#pragma clang fp contract(off)
    const int64_t i = (int64_t)blockIdx.x * BWG + threadIdx.x;
    if (i >= n) return;
    const T p = pts[i];
    T row[BASIS_RMAX];
    T s = T(0);
#pragma unroll
    for (int m = 0; m < BASIS_RMAX; ++m)
        if (m < r) row[m] = lam[m] * (T(1) / (p - nodes[m]));
#pragma unroll
    for (int m = 0; m < BASIS_RMAX; ++m)
        if (m < r) s += row[m];
#pragma unroll
    for (int m = 0; m < BASIS_RMAX; ++m)
        if (m < r) B[i + ldb * m] = row[m] / s;
}

int bfail(int code, const char *what, hipError_t e) {
    char buf[256];
    snprintf(buf, sizeof(buf), "%s: %s", what, e == hipSuccess ? "bad arguments" : hipGetErrorString(e));
    return hmx_set_error(code, buf);
}

template <class T>
int evaluate_t(int r, const T *tx, const T *ty, const T *lamx, const T *lamy, const T *F,
               int64_t n, const T *px, const T *py, T *out, void *stream) {
    if (r < 1 || n < 0) return bfail(HM_ERR_INVALID, "bary_evaluate", hipSuccess);
    if (n == 0) return HM_OK;
    if (!tx || !ty || !lamx || !lamy || !F || !px || !py || !out)
        return bfail(HM_ERR_INVALID, "bary_evaluate", hipSuccess);
    const int64_t nb = (n + BWG - 1) / BWG;
    if (nb > 0x7fffffff) return bfail(HM_ERR_INVALID, "bary_evaluate", hipSuccess);
    k_bary_evaluate<T><<<(unsigned)nb, BWG, 0, (hipStream_t)stream>>>(r, tx, ty, lamx, lamy, F, n,
                                                                        px, py, out);
    const hipError_t e = hipGetLastError();
    return e == hipSuccess ? HM_OK : bfail(HM_ERR_HIP, "bary_evaluate launch", e);
}

template <class T>
int basis_t(int r, const T *nodes, const T *lam, int64_t n, const T *pts, T *B, int64_t ldb,
            void *stream) {
    if (r < 1 || n < 0 || ldb < n) return bfail(HM_ERR_INVALID, "bary_basis", hipSuccess);
    if (r > BASIS_RMAX) {
        char buf[96];
        snprintf(buf, sizeof(buf), "bary_basis: rank %d above %d", r, BASIS_RMAX);
        return hmx_set_error(HM_ERR_UNSUPPORTED, buf);
    }
    if (n == 0) return HM_OK;
    if (!nodes || !lam || !pts || !B) return bfail(HM_ERR_INVALID, "bary_basis", hipSuccess);
    const int64_t nb = (n + BWG - 1) / BWG;
    if (nb > 0x7fffffff) return bfail(HM_ERR_INVALID, "bary_basis", hipSuccess);
    k_bary_basis<T><<<(unsigned)nb, BWG, 0, (hipStream_t)stream>>>(r, nodes, lam, n, pts, B, ldb);
    const hipError_t e = hipGetLastError();
    return e == hipSuccess ? HM_OK : bfail(HM_ERR_HIP, "bary_basis launch", e);
}

}  // namespace

extern "C" {

int hm_bary_evaluate(int r, const double *tx, const double *ty, const double *lamx,
                     const double *lamy, const double *F, int64_t n, const double *px,
                     const double *py, double *out, void *stream) {
    return evaluate_t(r, tx, ty, lamx, lamy, F, n, px, py, out, stream);
}

int hm_bary_evaluate_f32(int r, const float *tx, const float *ty, const float *lamx,
                         const float *lamy, const float *F, int64_t n, const float *px,
                         const float *py, float *out, void *stream) {
    return evaluate_t(r, tx, ty, lamx, lamy, F, n, px, py, out, stream);
}

int hm_bary_basis(int r, const double *nodes, const double *lam, int64_t n, const double *pts,
                  double *B, int64_t ldb, void *stream) {
    return basis_t(r, nodes, lam, n, pts, B, ldb, stream);
}

int hm_bary_basis_f32(int r, const float *nodes, const float *lam, int64_t n, const float *pts,
                      float *B, int64_t ldb, void *stream) {
    return basis_t(r, nodes, lam, n, pts, B, ldb, stream);
}

}  // extern "C"
