// even.cpp -- host side of the EvenBarycentricMatrix leaves: plan layout and
// the hm_even_* entry points of include/hmx.h.
#include "hmx.h"
#include "hmx_cheb.h"
#include "hmx_even.h"

#include <algorithm>
#include <cstdarg>
#include <cstdio>
#include <map>
#include <set>
#include <type_traits>
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

// 2i - a - b is formed in 64-bit integers and converted to T once, with
// round-to-nearest as the reference's Int -> T promotion; the grid bound
// keeps 2i in range and exact in Float64
constexpr int64_t GRID_MAX = (int64_t)1 << 50;

int check_grid(int64_t a, int64_t b, const char *what) {
    if (b < a) return efail(HM_ERR_INVALID, "%s: empty grid %lld:%lld", what, (long long)a, (long long)b);
    if (a < -GRID_MAX || b > GRID_MAX)
        return efail(HM_ERR_INVALID, "%s: grid %lld:%lld beyond +-2^50", what, (long long)a,
                     (long long)b);
    return HM_OK;
}

// first-kind nodes and weights in the element type T
template <class T>
ENodes<T> nodes_of(int n) {
    ENodes<T> nd{};
    cheb_points_t<T>(nd.x, n);
    cheb_weights_t<T>(nd.lam, n);
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
    // wsize, fsize: the tiled pools; wentries, fentries: m*n over the
    // distinct W blocks and p*n over the leaves
    int64_t wsize = 0, wentries = 0, fsize = 0, fentries = 0, nslots = 0;
    int64_t covered_rows = 0;  // rows inside some segment
    int64_t nwblocks = 0;      // distinct W blocks (one per leaf row count)
    ELeaf *d_leaves = nullptr;
    EUp *d_up = nullptr;
    ERed *d_red = nullptr;
    EDown *d_down = nullptr;
    int32_t *d_list = nullptr;
    bool f32 = false;  // element type of W, F and the products
    void *d_W = nullptr, *d_F = nullptr, *d_partial = nullptr, *d_beta = nullptr;

    ~hm_even() {
        void *bufs[] = {d_leaves, d_up, d_red, d_down, d_list, d_W, d_F, d_partial, d_beta};
        for (void *b : bufs)
            if (b) (void)hipFree(b);
    }

    template <class T>
    EvenDev<T> dev() const {
        EvenDev<T> p{};
        p.n = n;
        p.leaves = d_leaves;
        p.up = d_up;
        p.nup = (int32_t)up.size();
        p.red = d_red;
        p.nred = (int32_t)red.size();
        p.down = d_down;
        p.ndown = (int32_t)down.size();
        p.list = d_list;
        p.W = (const T *)d_W;
        p.F = (const T *)d_F;
        p.partial = (T *)d_partial;
        p.beta = (T *)d_beta;
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

template <class T>
int even_nodes_t(int64_t a, int64_t b, int n, double *t) {
    if (n < 1 || n > HM_EVEN_RMAX || !t) return efail(HM_ERR_INVALID, "even_nodes: bad n or t");
    if (int rc = check_grid(a, b, "even_nodes")) return rc;
    T x[HM_EVEN_RMAX];
    cheb_points_t<T>(x, n);
    // (a+b)/2 is a Float64 division of integers; (b-a)*x_k and /2 are in T;
    // the sum promotes to Float64 (src/BarycentricMatrix.jl:41)
    const double mid = (double)(a + b) / 2.0;
    for (int k = 0; k < n; ++k) {
        T h = (T)(b - a) * x[k];
        h = h / T(2);
        t[k] = mid + (double)h;
    }
    return HM_OK;
}

template <class T>
int even_weights_t(int64_t a, int64_t b, int n, int layout, T *W, T *w, void *stream) {
    if (n < 1 || n > HM_EVEN_RMAX || !W || (layout != 0 && layout != 1))
        return efail(HM_ERR_INVALID, "even_weights: bad n, layout or W");
    if (int rc = check_grid(a, b, "even_weights")) return rc;
    hipStream_t s = (hipStream_t)stream;
    const int64_t m = b - a + 1;
    ELeaf L{0, 0, m, 0, 0, 0, a, b};
    ELeaf *d = nullptr;
    if (int rc = dalloc(&d, 1)) return rc;
    hipError_t e = hipMemcpyAsync(d, &L, sizeof(L), hipMemcpyHostToDevice, s);
    if (e == hipSuccess) e = hmx_even_gen_weights(d, 1, m, nodes_of<T>(n), n, layout, W, w, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    (void)hipFree(d);
    if (e != hipSuccess) return efail(HM_ERR_HIP, "even_weights: %s", hipGetErrorString(e));
    return HM_OK;
}

template <class T>
int even_create_t(int64_t nrows, int64_t ncols, int64_t nleaves, const int64_t *row0,
                  const int64_t *col0, const int64_t *abcd, int n, const T *const *F,
                  void *stream, hm_even_t *out) {
    constexpr bool F32 = std::is_same_v<T, float>;
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
    P->f32 = F32;
    int64_t max_m = 0, max_pn = 0;
    // W depends on the rows only through i - a and b - a (2i - a - b =
    // 2(i - a) - (b - a)), so leaves with the same row count share one
    // bit-identical W: one block per distinct m, formed once (reps)
    std::map<int64_t, int64_t> w_of_m;
    std::vector<ELeaf> reps;
    for (int64_t t = 0; t < nleaves; ++t) {
        const int64_t a = abcd[4 * t], b = abcd[4 * t + 1], c = abcd[4 * t + 2], d = abcd[4 * t + 3];
        int rc = check_grid(a, b, "even_create rows");
        if (!rc) rc = check_grid(c, d, "even_create columns");
        if (rc) {
            delete P;
            return rc;
        }
        ELeaf L{row0[t], col0[t], b - a + 1, d - c + 1, 0, P->fsize, a, b, P->fentries};
        if (L.row0 < 0 || L.col0 < 0 || L.row0 + L.m > nrows || L.col0 + L.p > ncols || !F[t]) {
            delete P;
            return efail(HM_ERR_INVALID, "even_create: leaf %lld (%lld x %lld at %lld, %lld) "
                         "outside the %lld x %lld matrix or without F", (long long)t,
                         (long long)L.m, (long long)L.p, (long long)L.row0, (long long)L.col0,
                         (long long)nrows, (long long)ncols);
        }
        const auto w = w_of_m.find(L.m);
        if (w != w_of_m.end()) {
            L.woff = w->second;
        } else {
            L.woff = P->wsize;
            w_of_m.emplace(L.m, L.woff);
            reps.push_back(L);
            P->wsize += even_tiled_size(L.m, n);
            P->wentries += L.m * n;
        }
        P->fsize += even_tiled_size(L.p, n);
        P->fentries += L.p * n;
        max_m = std::max(max_m, L.m);
        max_pn = std::max(max_pn, L.p * n);
        P->leaves.push_back(L);
    }
    P->nwblocks = (int64_t)reps.size();
    P->make_up();
    P->make_down();
    hipStream_t s = (hipStream_t)stream;
    int rc = HM_OK;
    T *dW = nullptr, *dF = nullptr, *dpart = nullptr, *dbeta = nullptr;
#define HM_ETRY(expr) do { rc = (expr); if (rc != HM_OK) { delete P; return rc; } } while (0)
    HM_ETRY(upload(&P->d_leaves, P->leaves));
    HM_ETRY(upload(&P->d_up, P->up));
    HM_ETRY(upload(&P->d_red, P->red));
    HM_ETRY(upload(&P->d_down, P->down));
    HM_ETRY(upload(&P->d_list, P->list));
    HM_ETRY(dalloc(&dW, P->wsize));
    P->d_W = dW;
    HM_ETRY(dalloc(&dF, P->fsize));
    P->d_F = dF;
    const int64_t stride = (int64_t)HM_EVEN_KB * 2 * n;
    HM_ETRY(dalloc(&dpart, P->nslots * stride));
    P->d_partial = dpart;
    HM_ETRY(dalloc(&dbeta, nleaves * stride));
    P->d_beta = dbeta;
    // F: the caller's column-major arrays into a staging copy, then into tiles
    T *stage = nullptr;
    HM_ETRY(dalloc(&stage, P->fentries));
    hipError_t e = hipSuccess;
    for (int64_t t = 0; t < nleaves && e == hipSuccess; ++t) {
        const ELeaf &L = P->leaves[t];
        e = hipMemcpy(stage + L.soff, F[t], sizeof(T) * (size_t)(L.p * n), hipMemcpyHostToDevice);
    }
    if (e == hipSuccess) e = hmx_even_tile_F(P->d_leaves, nleaves, max_pn, n, stage, dF, s);
    ELeaf *d_reps = nullptr;
    if (e == hipSuccess && !reps.empty()) {
        if (dalloc(&d_reps, (int64_t)reps.size()) != HM_OK) e = hipErrorOutOfMemory;
        if (e == hipSuccess)
            e = hipMemcpy(d_reps, reps.data(), sizeof(ELeaf) * reps.size(), hipMemcpyHostToDevice);
        if (e == hipSuccess)
            e = hmx_even_gen_weights(d_reps, (int64_t)reps.size(), max_m, nodes_of<T>(n), n, 2, dW,
                                     (T *)nullptr, s);
    }
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (stage) (void)hipFree(stage);
    if (d_reps) (void)hipFree(d_reps);
    if (e != hipSuccess) {
        delete P;
        return efail(HM_ERR_HIP, "even_create: %s", hipGetErrorString(e));
    }
#undef HM_ETRY
    *out = P;
    return HM_OK;
}

template <class T>
int even_matvec_t(hm_even_t P, T *u, const T *v, int64_t ioff, int64_t joff, int64_t incu,
                  int64_t incv, int64_t k, int64_t ldu, int64_t ldv, int beta, void *stream) {
    if (!P) return efail(HM_ERR_INVALID, "even_matvec: NULL plan");
    if (P->f32 != std::is_same_v<T, float>)
        return efail(HM_ERR_INVALID, "even_matvec: the plan holds %s data; use the %s entry point",
                     P->f32 ? "Float32" : "Float64", P->f32 ? "_f32" : "Float64");
    if (k < 0 || ioff < 0 || joff < 0 || incu < 1 || incv < 1 || (beta != 0 && beta != 1))
        return efail(HM_ERR_INVALID, "even_matvec: bad offsets, strides, k or beta");
    if (k == 0) return HM_OK;
    if (!u || !v) return efail(HM_ERR_INVALID, "even_matvec: NULL vectors");
    if (k > (int64_t)1 << 30) return efail(HM_ERR_INVALID, "even_matvec: k too large");
    hipError_t e = hmx_even_matvec(P->dev<T>(), u, v, ioff, joff, incu, incv, ldu, ldv, (int)k,
                                   beta, (hipStream_t)stream);
    if (e != hipSuccess) return efail(HM_ERR_HIP, "even_matvec: %s", hipGetErrorString(e));
    return HM_OK;
}

template <class T>
int even_leaf_data_t(hm_even_t P, int64_t leaf, T *W, T *F) {
    if (!P || leaf < 0 || leaf >= (int64_t)P->leaves.size())
        return efail(HM_ERR_INVALID, "even_leaf_data: bad plan or leaf");
    if (P->f32 != std::is_same_v<T, float>)
        return efail(HM_ERR_INVALID, "even_leaf_data: element type of the plan differs");
    const ELeaf &L = P->leaves[leaf];
    const int n = P->n;
    HM_EHIP(hipDeviceSynchronize());
    if (W) {
        std::vector<T> tiles((size_t)even_tiled_size(L.m, n));
        HM_EHIP(hipMemcpy(tiles.data(), (const T *)P->d_W + L.woff, sizeof(T) * tiles.size(),
                          hipMemcpyDeviceToHost));
        for (int64_t i = 0; i < L.m; ++i)
            for (int k = 0; k < n; ++k) W[k + (int64_t)n * i] = tiles[(size_t)even_tile(i, k, n)];
    }
    if (F) {
        std::vector<T> tiles((size_t)even_tiled_size(L.p, n));
        HM_EHIP(hipMemcpy(tiles.data(), (const T *)P->d_F + L.foff, sizeof(T) * tiles.size(),
                          hipMemcpyDeviceToHost));
        for (int k = 0; k < n; ++k)
            for (int64_t j = 0; j < L.p; ++j) F[j + L.p * k] = tiles[(size_t)even_tile(j, k, n)];
    }
    return HM_OK;
}

extern "C" {

int hm_even_nodes(int64_t a, int64_t b, int n, double *t) { return even_nodes_t<double>(a, b, n, t); }
int hm_even_nodes_f32(int64_t a, int64_t b, int n, double *t) { return even_nodes_t<float>(a, b, n, t); }

int hm_even_weights(int64_t a, int64_t b, int n, int layout, double *W, double *w, void *stream) {
    return even_weights_t<double>(a, b, n, layout, W, w, stream);
}
int hm_even_weights_f32(int64_t a, int64_t b, int n, int layout, float *W, float *w, void *stream) {
    return even_weights_t<float>(a, b, n, layout, W, w, stream);
}

int hm_even_create(int64_t nrows, int64_t ncols, int64_t nleaves, const int64_t *row0,
                   const int64_t *col0, const int64_t *abcd, int n, const double *const *F,
                   void *stream, hm_even_t *out) {
    return even_create_t<double>(nrows, ncols, nleaves, row0, col0, abcd, n, F, stream, out);
}
int hm_even_create_f32(int64_t nrows, int64_t ncols, int64_t nleaves, const int64_t *row0,
                       const int64_t *col0, const int64_t *abcd, int n, const float *const *F,
                       void *stream, hm_even_t *out) {
    return even_create_t<float>(nrows, ncols, nleaves, row0, col0, abcd, n, F, stream, out);
}

int hm_even_matvec(hm_even_t P, double *u, const double *v, int64_t ioff, int64_t joff,
                   int64_t incu, int64_t incv, int64_t k, int64_t ldu, int64_t ldv, int beta,
                   void *stream) {
    return even_matvec_t<double>(P, u, v, ioff, joff, incu, incv, k, ldu, ldv, beta, stream);
}
int hm_even_matvec_f32(hm_even_t P, float *u, const float *v, int64_t ioff, int64_t joff,
                       int64_t incu, int64_t incv, int64_t k, int64_t ldu, int64_t ldv, int beta,
                       void *stream) {
    return even_matvec_t<float>(P, u, v, ioff, joff, incu, incv, k, ldu, ldv, beta, stream);
}

int hm_even_leaf_data(hm_even_t P, int64_t leaf, double *W, double *F) {
    return even_leaf_data_t<double>(P, leaf, W, F);
}
int hm_even_leaf_data_f32(hm_even_t P, int64_t leaf, float *W, float *F) {
    return even_leaf_data_t<float>(P, leaf, W, F);
}

int hm_even_info(hm_even_t P, int64_t *info, int n) {
    if (!P || !info || n < 0) return efail(HM_ERR_INVALID, "even_info: bad arguments");
    const int64_t v[HM_EVEN_INFO_COUNT] = {
        P->nrows, P->ncols, (int64_t)P->leaves.size(), P->n, P->wentries, P->fentries,
        (int64_t)P->up.size(), (int64_t)P->red.size(), (int64_t)P->down.size(),
        (int64_t)P->list.size(), P->f32 ? 1 : 0, P->nwblocks};
    for (int i = 0; i < n && i < HM_EVEN_INFO_COUNT; ++i) info[i] = v[i];
    return HM_OK;
}

int hm_even_destroy(hm_even_t P) {
    delete P;
    return HM_OK;
}

}  // extern "C"
