// lowrank.cpp -- host side of the batched LowRankMatrix arithmetic
// (hm_lowrank_op of include/hmx.h): shape checks, the item table, workspace.
#include "hmx.h"
#include "hmx_even.h"
#include "hmx_lowrank.h"

#include <algorithm>
#include <cstdarg>
#include <cstdio>
#include <vector>

namespace {

int lfail(int code, const char *fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    return hmx_set_error(code, buf);
}

const char *op_name(int op) {
    static const char *names[] = {"add", "sub", "mul", "adjmul", "dmul", "muld", "svdtrunc"};
    return op >= 0 && op <= HM_LR_SVDTRUNC ? names[op] : "?";
}

}  // namespace

extern "C" {

int hm_lowrank_op(int op, int64_t count, const int64_t *dims, const double *const *in,
                  double *const *out, int64_t *rank, void *stream) {
    if (op < HM_LR_ADD || op > HM_LR_SVDTRUNC) return lfail(HM_ERR_INVALID, "lowrank_op: unknown op %d", op);
    if (count < 0 || (count > 0 && (!dims || !in || !out || !rank)))
        return lfail(HM_ERR_INVALID, "lowrank_op: bad count or NULL arrays");
    if (count == 0) return HM_OK;
    if (count >= (int64_t)1 << 31) return lfail(HM_ERR_INVALID, "lowrank_op: too many items");
    std::vector<LRItem> items((size_t)count);
    int64_t tall = 0, max_rows = 0, lds_qr = 0, lds_svd = 0, max_c = 1;
    bool big_qr = false;
    for (int64_t t = 0; t < count; ++t) {
        const int64_t *d = dims + 5 * t;
        const int64_t m = d[0], n = d[1], k = d[2], r1 = d[3], r2 = d[4];
        const double *const *pi = in + 7 * t;
        double *const *po = out + 3 * t;
        LRItem &it = items[(size_t)t];
        it = LRItem{};
        const bool prod = op == HM_LR_MUL || op == HM_LR_ADJMUL || op == HM_LR_DMUL || op == HM_LR_MULD;
        if (m < 1 || n < 1 || m > INT32_MAX || n > INT32_MAX || (prod && (k < 1 || k > INT32_MAX)))
            return lfail(HM_ERR_INVALID, "lowrank_op %s: item %lld has bad sizes", op_name(op), (long long)t);
        it.m = (int32_t)m; it.n = (int32_t)n; it.k = (int32_t)k;
        it.U1 = pi[0]; it.S1 = pi[1]; it.V1 = pi[2]; it.U2 = pi[3]; it.S2 = pi[4]; it.V2 = pi[5];
        it.A = pi[6];
        it.Uo = po[0]; it.So = po[1]; it.Vo = po[2];
        const bool need1 = op != HM_LR_DMUL && op != HM_LR_SVDTRUNC;
        const bool need2 = op != HM_LR_MULD && op != HM_LR_SVDTRUNC;
        const bool needA = op == HM_LR_DMUL || op == HM_LR_MULD || op == HM_LR_SVDTRUNC;
        if (need1 && (r1 < 1 || r1 > LR_MAX)) return lfail(HM_ERR_UNSUPPORTED, "lowrank_op %s: item %lld: rank r1 = %lld outside 1..%d", op_name(op), (long long)t, (long long)r1, LR_MAX);
        if (need2 && (r2 < 1 || r2 > LR_MAX)) return lfail(HM_ERR_UNSUPPORTED, "lowrank_op %s: item %lld: rank r2 = %lld outside 1..%d", op_name(op), (long long)t, (long long)r2, LR_MAX);
        if ((need1 && (!it.U1 || !it.S1 || !it.V1)) || (need2 && (!it.U2 || !it.S2 || !it.V2)) ||
            (needA && !it.A) || !it.Uo || !it.So || !it.Vo)
            return lfail(HM_ERR_INVALID, "lowrank_op %s: item %lld: NULL operand or output", op_name(op), (long long)t);
        it.r1 = need1 ? (int32_t)r1 : 0;
        it.r2 = need2 ? (int32_t)r2 : 0;
        switch (op) {
        case HM_LR_ADD:
        case HM_LR_SUB: {
            const int64_t c = r1 + r2;
            if (c > LR_MAX || m < c || n < c)
                return lfail(HM_ERR_UNSUPPORTED, "lowrank_op %s: item %lld: r1 + r2 = %lld must be <= %d and <= m, n", op_name(op), (long long)t, (long long)c, LR_MAX);
            it.cu = it.cv = (int32_t)c;
            it.p = it.q = (int32_t)c;
            break;
        }
        case HM_LR_MUL:
        case HM_LR_ADJMUL:
            it.p = (int32_t)r1;
            it.q = (int32_t)r2;
            break;
        case HM_LR_DMUL:
            if (m < r2) return lfail(HM_ERR_UNSUPPORTED, "lowrank_op dmul: item %lld: m < rank", (long long)t);
            it.cu = (int32_t)r2;
            it.p = it.q = (int32_t)r2;
            break;
        case HM_LR_MULD:
            if (n < r1) return lfail(HM_ERR_UNSUPPORTED, "lowrank_op muld: item %lld: n < rank", (long long)t);
            it.cv = (int32_t)r1;
            it.p = it.q = (int32_t)r1;
            break;
        default:  // svdtrunc
            if (std::min(m, n) > LR_MAX)
                return lfail(HM_ERR_UNSUPPORTED, "lowrank_op svdtrunc: item %lld: min(m, n) > %d", (long long)t, LR_MAX);
            if (m >= n) { it.cu = (int32_t)n; it.p = it.q = (int32_t)n; }
            else { it.cv = (int32_t)m; it.tr = 1; it.p = it.q = (int32_t)m; }
        }
        it.xu = tall;
        tall += (int64_t)it.cu * m;
        max_c = std::max<int64_t>(max_c, std::max(it.cu, it.cv));
        it.xv = tall;
        tall += (int64_t)it.cv * n;
        it.small = t * LR_SMALL;
        max_rows = std::max(max_rows, std::max(m, n));
        const int64_t sides[2][2] = {{m, it.cu}, {n, it.cv}};
        for (const auto &sd : sides) {
            if (sd[1] == 0) continue;
            const int64_t need = sd[0] * sd[1] + 2 * sd[1] * (sd[1] + 1);
            if (need <= LR_LDS_DOUBLES) lds_qr = std::max(lds_qr, need);
            else big_qr = true;
        }
        const int64_t P = std::max(it.p, it.q), Q = std::min(it.p, it.q);
        lds_svd = std::max(lds_svd, (P + 1) * Q + (Q + 1) * Q + Q);
    }
    hipStream_t s = (hipStream_t)stream;
    LRItem *d_items = nullptr;
    double *d_tall = nullptr, *d_small = nullptr;
    int64_t *d_rank = nullptr;
    hipError_t e = hipMallocAsync((void **)&d_items, sizeof(LRItem) * (size_t)count, s);
    if (e == hipSuccess && tall > 0) e = hipMallocAsync((void **)&d_tall, sizeof(double) * (size_t)tall, s);
    if (e == hipSuccess) e = hipMallocAsync((void **)&d_small, sizeof(double) * (size_t)(count * LR_SMALL), s);
    if (e == hipSuccess) e = hipMallocAsync((void **)&d_rank, sizeof(int64_t) * (size_t)count, s);
    if (e == hipSuccess)
        e = hipMemcpyAsync(d_items, items.data(), sizeof(LRItem) * (size_t)count, hipMemcpyHostToDevice, s);
    if (e == hipSuccess)
        e = hmx_lowrank_run(op, d_items, count, d_tall, d_small, d_rank, max_rows, lds_qr, big_qr,
                            lds_svd, max_c, s);
    if (e == hipSuccess)
        e = hipMemcpyAsync(rank, d_rank, sizeof(int64_t) * (size_t)count, hipMemcpyDeviceToHost, s);
    hipError_t e2 = hipSuccess;
    void *bufs[] = {d_items, d_tall, d_small, d_rank};
    for (void *b : bufs)
        if (b) {
            hipError_t ef = hipFreeAsync(b, s);
            if (e2 == hipSuccess) e2 = ef;
        }
    hipError_t es = hipStreamSynchronize(s);
    if (e == hipSuccess) e = e2;
    if (e == hipSuccess) e = es;
    if (e != hipSuccess) return lfail(HM_ERR_HIP, "lowrank_op %s: %s", op_name(op), hipGetErrorString(e));
    return HM_OK;
}

}  // extern "C"
