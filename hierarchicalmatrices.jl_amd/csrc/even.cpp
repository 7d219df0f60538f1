// even.cpp -- host side of the EvenBarycentricMatrix leaves: plan layout and
// the hm_even_* entry points of include/hmx.h.
#include "hmx.h"
#include "hmx_even.h"

#include <algorithm>
#include <cstdarg>
#include <cstdio>
#include <set>
#include <vector>

namespace {

int efail(int code, const char *fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    return hmx_set_error(code, buf);
}

#define HM_EHIP(expr)                                                                      \
    do {                                                                                   \
        hipError_t e_ = (expr);                                                            \
        if (e_ != hipSuccess) return efail(HM_ERR_HIP, "%s: %s", #expr, hipGetErrorString(e_)); \
    } while (0)

// grid values stay exact in Float64 through 2i - a - b
constexpr int64_t GRID_MAX = (int64_t)1 << 50;

int check_grid(int64_t a, int64_t b, const char *what) {
    if (b < a) return efail(HM_ERR_INVALID, "%s: empty grid %lld:%lld", what, (long long)a, (long long)b);
    if (a < -GRID_MAX || b > GRID_MAX)
        return efail(HM_ERR_INVALID, "%s: grid %lld:%lld beyond +-2^50", what, (long long)a,
                     (long long)b);
    return HM_OK;
}

ENodes<double> nodes_of(int n) {
    ENodes<double> nd{};
    hm_chebyshevpoints(nd.x, n, 1);
    hm_chebyshevbarycentricweights(nd.lam, n, 1);
    return nd;
}

template <class T>
int dalloc(T **p, int64_t count) {
    *p = nullptr;
    if (count <= 0) return HM_OK;
    if (hipMalloc((void **)p, sizeof(T) * (size_t)count) != hipSuccess) {
        (void)hipGetLastError();
        *p = nullptr;
        return efail(HM_ERR_OOM, "hipMalloc of %lld bytes failed",
                     (long long)(sizeof(T) * (size_t)count));
    }
    return HM_OK;
}

template <class T>
int upload(T **d, const std::vector<T> &h) {
    if (int rc = dalloc(d, (int64_t)h.size())) return rc;
    if (!h.empty()) HM_EHIP(hipMemcpy(*d, h.data(), sizeof(T) * h.size(), hipMemcpyHostToDevice));
    return HM_OK;
}

}  // namespace

struct hm_even {
    int64_t nrows = 0, ncols = 0;
    int n = 0;
    std::vector<ELeaf> leaves;
    std::vector<EUp> up;
    std::vector<ERed> red;
    std::vector<EDown> down;
    std::vector<int32_t> list;
    // wsize, fsize: the tiled pools; wentries, fentries: m*n and p*n summed
    int64_t wsize = 0, wentries = 0, fsize = 0, fentries = 0, nslots = 0;
    int64_t covered_rows = 0;  // rows inside some segment
    ELeaf *d_leaves = nullptr;
    EUp *d_up = nullptr;
    ERed *d_red = nullptr;
    EDown *d_down = nullptr;
    int32_t *d_list = nullptr;
    double *d_W = nullptr, *d_F = nullptr, *d_partial = nullptr, *d_beta = nullptr;

    ~hm_even() {
        void *bufs[] = {d_leaves, d_up, d_red, d_down, d_list, d_W, d_F, d_partial, d_beta};
        for (void *b : bufs)
            if (b) (void)hipFree(b);
    }

    EvenDev<double> dev() const {
        EvenDev<double> p{};
        p.n = n;
        p.leaves = d_leaves;
        p.up = d_up;
        p.nup = (int32_t)up.size();
        p.red = d_red;
        p.nred = (int32_t)red.size();
        p.down = d_down;
        p.ndown = (int32_t)down.size();
        p.list = d_list;
        p.W = d_W;
        p.F = d_F;
        p.partial = d_partial;
        p.beta = d_beta;
        p.nrows = nrows;
        p.covered = covered_rows == nrows;
        return p;
    }

    // upward tasks (<= HM_EVEN_COLS columns) and the reduce list
    void make_up() {
        for (size_t t = 0; t < leaves.size(); ++t) {
            const ELeaf &L = leaves[t];
            const int64_t nt = (L.p + HM_EVEN_COLS - 1) / HM_EVEN_COLS;
            if (nt > 1) red.push_back(ERed{(int32_t)t, (int32_t)nt, nslots});
            for (int64_t q = 0; q < nt; ++q) {
                EUp u{};
                u.leaf = (int32_t)t;
                u.single = nt == 1;
                u.j0 = q * HM_EVEN_COLS;
                u.len = std::min<int64_t>(HM_EVEN_COLS, L.p - u.j0);
                u.slot = nt == 1 ? -1 : nslots++;
                up.push_back(u);
            }
        }
    }

    // row segments covered by the same leaves (a sweep over the leaf row
    // ranges; each segment lists its leaves in leaf order), cut into
    // downward tasks of <= HM_EVEN_ROWS rows
    void make_down() {
        std::vector<std::pair<int64_t, int32_t>> ev;  // (row, +leaf+1 start / -(leaf+1) end)
        for (size_t t = 0; t < leaves.size(); ++t) {
            ev.push_back({leaves[t].row0, (int32_t)t + 1});
            ev.push_back({leaves[t].row0 + leaves[t].m, -((int32_t)t + 1)});
        }
        std::sort(ev.begin(), ev.end());
        std::set<int32_t> active;
        size_t e = 0;
        while (e < ev.size()) {
            const int64_t r = ev[e].first;
            while (e < ev.size() && ev[e].first == r) {
                const int32_t c = ev[e].second;
                if (c > 0) active.insert(c - 1);
                else active.erase(-c - 1);
                ++e;
            }
            if (e == ev.size() || active.empty()) continue;
            const int64_t rend = ev[e].first;
            const int32_t lbeg = (int32_t)list.size();
            for (int32_t id : active) list.push_back(id);
            const int32_t lend = (int32_t)list.size();
            covered_rows += rend - r;
            for (int64_t r0 = r; r0 < rend; r0 += HM_EVEN_ROWS)
                down.push_back(EDown{r0, (int32_t)std::min<int64_t>(HM_EVEN_ROWS, rend - r0),
                                     lbeg, lend, 0});
        }
    }
};

extern "C" {

int hm_even_nodes(int64_t a, int64_t b, int n, double *t) {
    if (n < 1 || n > HM_EVEN_RMAX || !t) return efail(HM_ERR_INVALID, "even_nodes: bad n or t");
    if (int rc = check_grid(a, b, "even_nodes")) return rc;
    double x[HM_EVEN_RMAX];
    hm_chebyshevpoints(x, n, 1);
    // (a+b)/2 is a Float64 division of integers; (b-a)*x_k/2 in Float64
    const double mid = (double)(a + b) / 2.0;
    for (int k = 0; k < n; ++k) t[k] = mid + ((double)(b - a) * x[k]) / 2.0;
    return HM_OK;
}

int hm_even_weights(int64_t a, int64_t b, int n, int layout, double *W, double *w, void *stream) {
    if (n < 1 || n > HM_EVEN_RMAX || !W || (layout != 0 && layout != 1))
        return efail(HM_ERR_INVALID, "even_weights: bad n, layout or W");
    if (int rc = check_grid(a, b, "even_weights")) return rc;
    hipStream_t s = (hipStream_t)stream;
    const int64_t m = b - a + 1;
    ELeaf L{0, 0, m, 0, 0, 0, a, b};
    ELeaf *d = nullptr;
    if (int rc = dalloc(&d, 1)) return rc;
    hipError_t e = hipMemcpyAsync(d, &L, sizeof(L), hipMemcpyHostToDevice, s);
    if (e == hipSuccess) e = hmx_even_gen_weights(d, 1, m, nodes_of(n), n, layout, W, w, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    (void)hipFree(d);
    if (e != hipSuccess) return efail(HM_ERR_HIP, "even_weights: %s", hipGetErrorString(e));
    return HM_OK;
}

int hm_even_create(int64_t nrows, int64_t ncols, int64_t nleaves, const int64_t *row0,
                   const int64_t *col0, const int64_t *abcd, int n, const double *const *F,
                   void *stream, hm_even_t *out) {
    if (!out) return efail(HM_ERR_INVALID, "even_create: NULL out");
    *out = nullptr;
    if (nrows < 0 || ncols < 0 || nleaves < 0 || n < 1 || n > HM_EVEN_RMAX)
        return efail(HM_ERR_INVALID, "even_create: bad sizes or rank (1..%d)", HM_EVEN_RMAX);
    if (nleaves > 0 && (!row0 || !col0 || !abcd || !F))
        return efail(HM_ERR_INVALID, "even_create: NULL leaf arrays");
    if (nleaves >= (int64_t)1 << 30) return efail(HM_ERR_INVALID, "even_create: too many leaves");
    hm_even *P = new hm_even();
    P->nrows = nrows;
    P->ncols = ncols;
    P->n = n;
    int64_t max_m = 0, max_pn = 0;
    for (int64_t t = 0; t < nleaves; ++t) {
        const int64_t a = abcd[4 * t], b = abcd[4 * t + 1], c = abcd[4 * t + 2], d = abcd[4 * t + 3];
        int rc = check_grid(a, b, "even_create rows");
        if (!rc) rc = check_grid(c, d, "even_create columns");
        if (rc) {
            delete P;
            return rc;
        }
        ELeaf L{row0[t], col0[t], b - a + 1, d - c + 1, P->wsize, P->fsize, a, b, P->fentries};
        if (L.row0 < 0 || L.col0 < 0 || L.row0 + L.m > nrows || L.col0 + L.p > ncols || !F[t]) {
            delete P;
            return efail(HM_ERR_INVALID, "even_create: leaf %lld (%lld x %lld at %lld, %lld) "
                         "outside the %lld x %lld matrix or without F", (long long)t,
                         (long long)L.m, (long long)L.p, (long long)L.row0, (long long)L.col0,
                         (long long)nrows, (long long)ncols);
        }
        P->wsize += even_tiled_size(L.m, n);
        P->wentries += L.m * n;
        P->fsize += even_tiled_size(L.p, n);
        P->fentries += L.p * n;
        max_m = std::max(max_m, L.m);
        max_pn = std::max(max_pn, L.p * n);
        P->leaves.push_back(L);
    }
    P->make_up();
    P->make_down();
    hipStream_t s = (hipStream_t)stream;
    int rc = HM_OK;
#define HM_ETRY(expr) do { rc = (expr); if (rc != HM_OK) { delete P; return rc; } } while (0)
    HM_ETRY(upload(&P->d_leaves, P->leaves));
    HM_ETRY(upload(&P->d_up, P->up));
    HM_ETRY(upload(&P->d_red, P->red));
    HM_ETRY(upload(&P->d_down, P->down));
    HM_ETRY(upload(&P->d_list, P->list));
    HM_ETRY(dalloc(&P->d_W, P->wsize));
    HM_ETRY(dalloc(&P->d_F, P->fsize));
    const int64_t stride = (int64_t)HM_EVEN_KB * 2 * n;
    HM_ETRY(dalloc(&P->d_partial, P->nslots * stride));
    HM_ETRY(dalloc(&P->d_beta, nleaves * stride));
    // F: the caller's column-major arrays into a staging copy, then into tiles
    double *stage = nullptr;
    HM_ETRY(dalloc(&stage, P->fentries));
    hipError_t e = hipSuccess;
    for (int64_t t = 0; t < nleaves && e == hipSuccess; ++t) {
        const ELeaf &L = P->leaves[t];
        e = hipMemcpy(stage + L.soff, F[t], sizeof(double) * (size_t)(L.p * n),
                      hipMemcpyHostToDevice);
    }
    if (e == hipSuccess) e = hmx_even_tile_F(P->d_leaves, nleaves, max_pn, n, stage, P->d_F, s);
    if (e == hipSuccess)
        e = hmx_even_gen_weights(P->d_leaves, nleaves, max_m, nodes_of(n), n, 2, P->d_W, nullptr,
                                 s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (stage) (void)hipFree(stage);
    if (e != hipSuccess) {
        delete P;
        return efail(HM_ERR_HIP, "even_create: %s", hipGetErrorString(e));
    }
#undef HM_ETRY
    *out = P;
    return HM_OK;
}

int hm_even_matvec(hm_even_t P, double *u, const double *v, int64_t ioff, int64_t joff,
                   int64_t incu, int64_t incv, int64_t k, int64_t ldu, int64_t ldv, int beta,
                   void *stream) {
    if (!P) return efail(HM_ERR_INVALID, "even_matvec: NULL plan");
    if (k < 0 || ioff < 0 || joff < 0 || incu < 1 || incv < 1 || (beta != 0 && beta != 1))
        return efail(HM_ERR_INVALID, "even_matvec: bad offsets, strides, k or beta");
    if (k == 0) return HM_OK;
    if (!u || !v) return efail(HM_ERR_INVALID, "even_matvec: NULL vectors");
    if (k > (int64_t)1 << 30) return efail(HM_ERR_INVALID, "even_matvec: k too large");
    hipError_t e = hmx_even_matvec(P->dev(), u, v, ioff, joff, incu, incv, ldu, ldv, (int)k,
                                   beta, (hipStream_t)stream);
    if (e != hipSuccess) return efail(HM_ERR_HIP, "even_matvec: %s", hipGetErrorString(e));
    return HM_OK;
}

int hm_even_leaf_data(hm_even_t P, int64_t leaf, double *W, double *F) {
    if (!P || leaf < 0 || leaf >= (int64_t)P->leaves.size())
        return efail(HM_ERR_INVALID, "even_leaf_data: bad plan or leaf");
    const ELeaf &L = P->leaves[leaf];
    const int n = P->n;
    HM_EHIP(hipDeviceSynchronize());
    if (W) {
        std::vector<double> tiles((size_t)even_tiled_size(L.m, n));
        HM_EHIP(hipMemcpy(tiles.data(), P->d_W + L.woff, sizeof(double) * tiles.size(),
                          hipMemcpyDeviceToHost));
        for (int64_t i = 0; i < L.m; ++i)
            for (int k = 0; k < n; ++k) W[k + (int64_t)n * i] = tiles[(size_t)even_tile(i, k, n)];
    }
    if (F) {
        std::vector<double> tiles((size_t)even_tiled_size(L.p, n));
        HM_EHIP(hipMemcpy(tiles.data(), P->d_F + L.foff, sizeof(double) * tiles.size(),
                          hipMemcpyDeviceToHost));
        for (int k = 0; k < n; ++k)
            for (int64_t j = 0; j < L.p; ++j) F[j + L.p * k] = tiles[(size_t)even_tile(j, k, n)];
    }
    return HM_OK;
}

int hm_even_info(hm_even_t P, int64_t *info, int n) {
    if (!P || !info || n < 0) return efail(HM_ERR_INVALID, "even_info: bad arguments");
    const int64_t v[HM_EVEN_INFO_COUNT] = {
        P->nrows, P->ncols, (int64_t)P->leaves.size(), P->n, P->wsize, P->fsize,
        (int64_t)P->up.size(), (int64_t)P->red.size(), (int64_t)P->down.size(),
        (int64_t)P->list.size()};
    for (int i = 0; i < n && i < HM_EVEN_INFO_COUNT; ++i) info[i] = v[i];
    return HM_OK;
}

int hm_even_destroy(hm_even_t P) {
    delete P;
    return HM_OK;
}

}  // extern "C"
