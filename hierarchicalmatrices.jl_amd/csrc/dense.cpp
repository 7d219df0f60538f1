// dense.cpp -- host side of the batched dense leaf steps (hm_dense_* of
// include/hmx.h): argument checks, the item table, launches.
#include "hmx.h"
#include "hmx_dense.h"
#include "hmx_even.h"

#include <algorithm>
#include <cstdarg>
#include <cstdio>
#include <mutex>
#include <vector>

// Item table of hm_dense_gemm_noalloc: a fixed device array, so that call
// allocates nothing. Each device holds its own copy until the process ends;
// every call on a device shares it (calls that overlap on one device race,
// and the array keeps the last call's items).
static __device__ __align__(256) DenseItem scratch[HM_DENSE_NOALLOC_MAX];

namespace {

int dfail(int code, const char *fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    return hmx_set_error(code, buf);
}

// runs fn(items, info) on the stream. One item without an info result goes
// by value and the call returns without waiting; otherwise the item table is
// copied to the device and the call synchronizes the stream.
template <class F>
int run_items(const std::vector<DenseItem> &items, hipStream_t s, const char *what, F fn,
              int64_t *info_host = nullptr) {
    const size_t count = items.size();
    if (count == 1 && !info_host) {
        DenseItems its{nullptr, items[0]};
        const hipError_t e = fn(its, nullptr);
        if (e != hipSuccess) return dfail(HM_ERR_HIP, "%s: %s", what, hipGetErrorString(e));
        return HM_OK;
    }
    DenseItem *d = nullptr;
    int64_t *dinfo = nullptr;
    hipError_t e = hipSuccess;
    if (count > 1) e = hipMallocAsync((void **)&d, sizeof(DenseItem) * count, s);
    if (e == hipSuccess && info_host) e = hipMallocAsync((void **)&dinfo, sizeof(int64_t) * count, s);
    if (e == hipSuccess && d)
        e = hipMemcpyAsync(d, items.data(), sizeof(DenseItem) * count, hipMemcpyHostToDevice, s);
    DenseItems its{d, items[0]};
    if (e == hipSuccess) e = fn(its, dinfo);
    if (e == hipSuccess && info_host)
        e = hipMemcpyAsync(info_host, dinfo, sizeof(int64_t) * count, hipMemcpyDeviceToHost, s);
    if (d) (void)hipFreeAsync(d, s);
    if (dinfo) (void)hipFreeAsync(dinfo, s);
    const hipError_t es = hipStreamSynchronize(s);
    if (e == hipSuccess) e = es;
    if (e != hipSuccess) return dfail(HM_ERR_HIP, "%s: %s", what, hipGetErrorString(e));
    return HM_OK;
}

// The device address of `scratch` on the current device, looked up once per
// device (the host-side name is a stand-in, not a device address).
constexpr int NOALLOC_MAX_DEVICES = 64;
std::mutex g_scratch_mu;
DenseItem *g_scratch[NOALLOC_MAX_DEVICES] = {};

int scratch_items(DenseItem **out) {
    int dev = -1;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return dfail(HM_ERR_HIP, "dense_gemm_noalloc: hipGetDevice: %s", hipGetErrorString(e));
    if (dev < 0 || dev >= NOALLOC_MAX_DEVICES)
        return dfail(HM_ERR_UNSUPPORTED, "dense_gemm_noalloc: device %d beyond %d", dev, NOALLOC_MAX_DEVICES);
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch[dev]) {
        void *p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return dfail(HM_ERR_HIP, "dense_gemm_noalloc: hipGetSymbolAddress: %s", hipGetErrorString(e));
        g_scratch[dev] = (DenseItem *)p;
    }
    *out = g_scratch[dev];
    return HM_OK;
}

// the item table of a batched GEMM (hm_dense_gemm's checks) into items[0, count)
int gemm_items(int transa, int transb, int64_t count, const int64_t *mnk, const double *const *A,
               const int64_t *lda, const double *const *D, const double *const *B,
               const int64_t *ldb, double *const *C, const int64_t *ldc, const char *what,
               DenseItem *items, int64_t &max_tiles) {
    if ((transa != 0 && transa != 1) || (transb != 0 && transb != 1) || count < 0 ||
        (count > 0 && (!mnk || !A || !lda || !B || !ldb || !C || !ldc)))
        return dfail(HM_ERR_INVALID, "%s: bad arguments", what);
    max_tiles = 0;
    for (int64_t t = 0; t < count; ++t) {
        const int64_t m = mnk[3 * t], n = mnk[3 * t + 1], k = mnk[3 * t + 2];
        const int64_t arows = transa ? k : m, brows = transb ? n : k;
        if (m < 0 || n < 0 || k < 0 || m > INT32_MAX || n > INT32_MAX || k > INT32_MAX ||
            ldc[t] < std::max<int64_t>(1, m) || lda[t] < std::max<int64_t>(1, arows) ||
            ldb[t] < std::max<int64_t>(1, brows) || (m > 0 && n > 0 && !C[t]) ||
            (m > 0 && n > 0 && k > 0 && (!A[t] || !B[t])))
            return dfail(HM_ERR_INVALID, "%s: item %lld: bad sizes or pointers", what, (long long)t);
        items[(size_t)t] = DenseItem{(int32_t)m, (int32_t)n, (int32_t)k, 0, C[t], ldc[t], A[t],
                                     lda[t], B[t], ldb[t], D ? D[t] : nullptr};
        max_tiles = std::max(max_tiles, ((m + 63) / 64) * ((n + 63) / 64));
    }
    return HM_OK;
}

}  // namespace

extern "C" {

int hm_dense_potrf(int64_t count, const int64_t *n, double *const *A, const int64_t *lda,
                   int64_t *info, void *stream) {
    if (count < 0 || (count > 0 && (!n || !A || !lda || !info)))
        return dfail(HM_ERR_INVALID, "dense_potrf: bad count or NULL arrays");
    if (count == 0) return HM_OK;
    std::vector<DenseItem> items((size_t)count);
    for (int64_t t = 0; t < count; ++t) {
        if (n[t] < 0 || n[t] > INT32_MAX || lda[t] < std::max<int64_t>(1, n[t]) || (n[t] > 0 && !A[t]))
            return dfail(HM_ERR_INVALID, "dense_potrf: item %lld: bad n, lda or A", (long long)t);
        items[(size_t)t] = DenseItem{(int32_t)n[t], 0, 0, 0, A[t], lda[t], nullptr, 0, nullptr, 0,
                                     nullptr};
    }
    int rc = run_items(items, (hipStream_t)stream, "dense_potrf",
                       [&](const DenseItems &d, int64_t *di) {
                           return hmx_dense_potrf(d, count, di, (hipStream_t)stream);
                       },
                       info);
    if (rc) return rc;
    for (int64_t t = 0; t < count; ++t)
        if (info[t]) return dfail(HM_ERR_INVALID, "dense_potrf: item %lld is not positive definite (pivot %lld)", (long long)t, (long long)info[t]);
    return HM_OK;
}

int hm_dense_potrf_dev(int64_t count, const int64_t *n, double *const *A, const int64_t *lda,
                       int64_t *info, void *stream) {
    if (count < 0 || (count > 0 && (!n || !A || !lda || !info)))
        return dfail(HM_ERR_INVALID, "dense_potrf_dev: bad count or NULL arrays");
    if (count == 0) return HM_OK;
    std::vector<DenseItem> items((size_t)count);
    for (int64_t t = 0; t < count; ++t) {
        if (n[t] < 0 || n[t] > INT32_MAX || lda[t] < std::max<int64_t>(1, n[t]) || (n[t] > 0 && !A[t]))
            return dfail(HM_ERR_INVALID, "dense_potrf_dev: item %lld: bad n, lda or A", (long long)t);
        items[(size_t)t] = DenseItem{(int32_t)n[t], 0, 0, 0, A[t], lda[t], nullptr, 0, nullptr, 0,
                                     nullptr};
    }
    hipStream_t s = (hipStream_t)stream;
    hipError_t e = hipSuccess;
    if (count == 1) {
        e = hmx_dense_potrf(DenseItems{nullptr, items[0]}, 1, info, s);
    } else {
        // the table goes up with the stream (the copy from pageable host
        // memory has read it when the call returns) and is freed in order
        DenseItem *d = nullptr;
        e = hipMallocAsync((void **)&d, sizeof(DenseItem) * (size_t)count, s);
        if (e == hipSuccess)
            e = hipMemcpyAsync(d, items.data(), sizeof(DenseItem) * (size_t)count,
                               hipMemcpyHostToDevice, s);
        if (e == hipSuccess) e = hmx_dense_potrf(DenseItems{d, items[0]}, count, info, s);
        if (d) (void)hipFreeAsync(d, s);
    }
    if (e != hipSuccess) return dfail(HM_ERR_HIP, "dense_potrf_dev: %s", hipGetErrorString(e));
    return HM_OK;
}

int hm_dense_trsm_t(int64_t count, const int64_t *n, const int64_t *k, const double *const *R,
                    const int64_t *ldr, double *const *B, const int64_t *ldb, void *stream) {
    if (count < 0 || (count > 0 && (!n || !k || !R || !ldr || !B || !ldb)))
        return dfail(HM_ERR_INVALID, "dense_trsm_t: bad count or NULL arrays");
    if (count == 0) return HM_OK;
    std::vector<DenseItem> items((size_t)count);
    int64_t max_k = 0, max_n = 0;
    for (int64_t t = 0; t < count; ++t) {
        if (n[t] < 0 || k[t] < 0 || n[t] > INT32_MAX || k[t] > INT32_MAX ||
            ldr[t] < std::max<int64_t>(1, n[t]) || ldb[t] < std::max<int64_t>(1, n[t]) ||
            (n[t] > 0 && k[t] > 0 && (!R[t] || !B[t])))
            return dfail(HM_ERR_INVALID, "dense_trsm_t: item %lld: bad sizes or pointers", (long long)t);
        items[(size_t)t] = DenseItem{(int32_t)n[t], (int32_t)k[t], 0, 0, B[t], ldb[t], R[t], ldr[t],
                                     nullptr, 0, nullptr};
        max_k = std::max(max_k, n[t] > 0 ? k[t] : 0);
        max_n = std::max(max_n, n[t]);
    }
    return run_items(items, (hipStream_t)stream, "dense_trsm_t",
                     [&](const DenseItems &d, int64_t *) {
                         return hmx_dense_trsm_t(d, count, max_n, max_k, (hipStream_t)stream);
                     });
}

int hm_dense_gemm(int transa, int transb, int64_t count, const int64_t *mnk, double alpha,
                  const double *const *A, const int64_t *lda, const double *const *D,
                  const double *const *B, const int64_t *ldb, double beta, double *const *C,
                  const int64_t *ldc, void *stream) {
    if ((transa != 0 && transa != 1) || (transb != 0 && transb != 1) || count < 0)
        return dfail(HM_ERR_INVALID, "dense_gemm: bad arguments");
    std::vector<DenseItem> items((size_t)count);
    int64_t max_tiles = 0;
    if (int rc = gemm_items(transa, transb, count, mnk, A, lda, D, B, ldb, C, ldc, "dense_gemm",
                            items.data(), max_tiles))
        return rc;
    if (count == 0) return HM_OK;
    return run_items(items, (hipStream_t)stream, "dense_gemm",
                     [&](const DenseItems &d, int64_t *) {
                         return hmx_dense_gemm(transa, transb, alpha, beta, d, count, max_tiles,
                                               (hipStream_t)stream);
                     });
}

int hm_dense_gemm_noalloc(int transa, int transb, int64_t count, const int64_t *mnk, double alpha,
                          const double *const *A, const int64_t *lda, const double *const *D,
                          const double *const *B, const int64_t *ldb, double beta,
                          double *const *C, const int64_t *ldc, void *stream) {
    if (count > HM_DENSE_NOALLOC_MAX)
        return dfail(HM_ERR_UNSUPPORTED, "dense_gemm_noalloc: %lld items, at most %d per call",
                     (long long)count, HM_DENSE_NOALLOC_MAX);
    // the host table is static too, held while it is filled and copied (the
    // copy from pageable host memory has read it when hipMemcpyAsync returns)
    static DenseItem host_items[HM_DENSE_NOALLOC_MAX];
    static std::mutex host_mu;
    std::lock_guard<std::mutex> lock(host_mu);
    int64_t max_tiles = 0;
    if (int rc = gemm_items(transa, transb, count, mnk, A, lda, D, B, ldb, C, ldc,
                            "dense_gemm_noalloc", host_items, max_tiles))
        return rc;
    if (count == 0) return HM_OK;
    hipStream_t s = (hipStream_t)stream;
    hipError_t e = hipSuccess;
    if (count == 1) {  // one item goes by value, as in hm_dense_gemm
        e = hmx_dense_gemm(transa, transb, alpha, beta, DenseItems{nullptr, host_items[0]}, 1,
                           max_tiles, s);
    } else {
        DenseItem *d = nullptr;
        if (int rc = scratch_items(&d)) return rc;
        e = hipMemcpyAsync(d, host_items, sizeof(DenseItem) * (size_t)count,
                           hipMemcpyHostToDevice, s);
        if (e == hipSuccess)
            e = hmx_dense_gemm(transa, transb, alpha, beta, DenseItems{d, host_items[0]}, count,
                               max_tiles, s);
    }
    if (e != hipSuccess) return dfail(HM_ERR_HIP, "dense_gemm_noalloc: %s", hipGetErrorString(e));
    return HM_OK;
}

}  // extern "C"
