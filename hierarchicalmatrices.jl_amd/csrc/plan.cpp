// plan.cpp -- host side of libhmx: the KernelMatrix block tree, its flattened
// device layout, and the C ABI declared in include/hmx.h.
#include "hmx.h"
#include "hmx_device.h"
#include "hmx_cheb.h"

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <memory>
#include <string>
#include <type_traits>
#include <unordered_map>
#include <vector>

namespace {

thread_local std::string g_err;

int fail(int code, const char *fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_err = buf;
    return code;
}

#define HM_HIP(expr)                                                                      \
    do {                                                                                  \
        hipError_t e_ = (expr);                                                           \
        if (e_ != hipSuccess) return fail(HM_ERR_HIP, "%s: %s", #expr, hipGetErrorString(e_)); \
    } while (0)

// sin(pi q) evaluated in long double and rounded once, for the Chebyshev
// points and weights (the reference calls Julia's sinpi).
double sinpi_ld(double q) {
    const long double PI_L = 3.14159265358979323846264338327950288419716939937510L;
    return (double)sinl(PI_L * (long double)q);
}

// src/BarycentricMatrix.jl:92-111
void cheb_points(double *x, int n, int kind) {
    std::fill(x, x + n, 0.0);
    const int nd2 = n / 2;
    const double den = kind == 1 ? (double)(2 * n) : (double)(2 * (n - 1));
    for (int k = 1; k <= nd2; ++k) x[k - 1] = sinpi_ld(((double)(n - 2 * k) + 1.0) / den);
    for (int k = 1; k <= nd2; ++k) x[n - k] = -x[k - 1];
}

// src/BarycentricMatrix.jl:114-136
void cheb_weights(double *l, int n, int kind) {
    std::fill(l, l + n, 0.0);
    const int nd2 = n / 2;
    if (kind == 1) {
        for (int k = 1; k <= nd2 + 1 && k <= n; ++k)
            l[k - 1] = sinpi_ld(((double)(2 * k) - 1.0) / (double)(2 * n));
        for (int k = 1; k <= nd2; ++k) l[n - k] = l[k - 1];
    } else {
        for (int k = 0; k < n; ++k) l[k] = 1.0;
    }
    for (int k = 2; k <= n; k += 2) l[k - 1] = -l[k - 1];
    if (kind != 1) {
        l[0] *= 0.5;
        l[n - 1] *= 0.5;
    }
}

bool non_increasing(const double *p, int64_t n, int64_t *bad) {
    for (int64_t i = 0; i < n; ++i) {
        if (!std::isfinite(p[i]) || (i + 1 < n && !(p[i] >= p[i + 1]))) {
            *bad = i;
            return false;
        }
    }
    return true;
}

struct Leaf {
    int code, depth;
    int64_t row0, m, col0, n;
    double a, b, c, d;
};

struct HCluster {
    int depth;
    int64_t first, len;
    double lo, hi;  // interval as passed to the block: (a, b) with a the upper end
};

// The tree rule of KernelMatrix / KernelMatrix1 / KernelMatrix2
// (src/KernelMatrix.jl:49-116), recorded as flat tables.
struct Builder {
    const double *x, *y;
    int64_t nx, ny;
    int bs, r;
    bool f32 = false;  // Float32 plan: points and intervals are floats, midpoints rounded to float
    std::vector<Leaf> leaves;
    std::vector<HCluster> rcl, ccl;
    std::unordered_map<uint64_t, int32_t> rmap, cmap;
    std::vector<DFarBlk> far;
    std::vector<int> far_depth;
    std::vector<DDense> dense;
    // every indsplit of the tree (row and column side): the tree stays the
    // same for other points exactly when each split falls at the same index
    std::vector<DSplit> xsplits, ysplits;
    int64_t nd = 0, nk1 = 0, nk2 = 0, far_rows = 0, far_cols = 0, dense_entries = 0;
    int maxdepth = 0;
    int err = 0;
    std::string msg;
    // the interval trees of both sides (every range node() splits, with its
    // two halves); tok turns false if a lookup meets an inconsistent node, and
    // the plan then keeps the per-level far field
    struct TNode {
        int depth;
        int64_t first, len;
        double lo, hi;
        int32_t parent, child[2], half;
        int32_t cl;  // the row / column cluster with this range, or -1
    };
    std::vector<TNode> xt, yt;
    std::unordered_map<uint64_t, int32_t> xtm, ytm;
    bool tok = true;

    static uint64_t tkey(int depth, double lo, double hi) {
        uint64_t a, b;
        std::memcpy(&a, &lo, 8);
        std::memcpy(&b, &hi, 8);
        uint64_t h = a * 0x9E3779B97F4A7C15ull ^ (b + 0x632BE59BD9B4E019ull + (h_mix(a) << 6));
        return h_mix(h ^ (uint64_t)depth * 0xD1B54A32D192ED03ull);
    }
    static uint64_t h_mix(uint64_t v) {
        v ^= v >> 30; v *= 0xBF58476D1CE4E5B9ull;
        v ^= v >> 27; v *= 0x94D049BB133111EBull;
        return v ^ (v >> 31);
    }
    int32_t tfind(const std::vector<TNode> &t, const std::unordered_map<uint64_t, int32_t> &m,
                  int depth, double lo, double hi) const {
        auto it = m.find(tkey(depth, lo, hi));
        if (it == m.end()) return -1;
        const TNode &q = t[it->second];
        return (q.depth == depth && q.lo == lo && q.hi == hi) ? it->second : -2;
    }
    // the node (depth, interval), made on first sight as half `half` of parent
    void tadd(std::vector<TNode> &t, std::unordered_map<uint64_t, int32_t> &m, int depth,
              int64_t first, int64_t len, double lo, double hi, int32_t parent, int half) {
        const int32_t f = tfind(t, m, depth, lo, hi);
        if (f == -2) { tok = false; return; }
        if (f >= 0) {
            const TNode &q = t[f];
            if (q.first != first || q.len != len || q.parent != parent) tok = false;
            return;
        }
        const int32_t id = (int32_t)t.size();
        t.push_back({depth, first, len, lo, hi, parent, {-1, -1}, half, -1});
        m.emplace(tkey(depth, lo, hi), id);
        if (parent >= 0) t[parent].child[half] = id;
    }

    // indsplit (src/BarycentricMatrix.jl:299-307) on a validated non-increasing
    // range: the linear scan stops at the first point below the midpoint, which
    // a binary search finds directly.
    // half(T)*(a+b) in the element type (src/BarycentricMatrix.jl:300,
    // src/KernelMatrix.jl:51); float values are exact in double
    double mid(double a, double b) const {
        if (f32) return (double)(0.5f * ((float)a + (float)b));
        return 0.5 * (a + b);
    }

    int64_t split(const double *p, int64_t np, int64_t first, int64_t len, double a, double b) {
        const double ab2 = mid(a, b);
        if (len == 0) {
            if (first >= np || p[first] >= ab2) {
                err = HM_ERR_DEGENERATE;
                char buf[160];
                snprintf(buf, sizeof(buf),
                         "empty cluster at index %lld reached indsplit: the reference would "
                         "read out of its range", (long long)first);
                msg = buf;
            }
            return 0;
        }
        const double *lo = p + first, *hi = p + first + len;
        const double *it = std::partition_point(lo, hi, [ab2](double v) { return v >= ab2; });
        return (int64_t)(it - lo);
    }

    int32_t cluster(std::vector<HCluster> &cl, std::unordered_map<uint64_t, int32_t> &mp, int depth,
                    int64_t first, int64_t len, double lo, double hi) {
        const uint64_t key = ((uint64_t)depth << 42) | (uint64_t)first;
        auto it = mp.find(key);
        if (it != mp.end()) {
            const HCluster &h = cl[it->second];
            if (h.len != len || h.lo != lo || h.hi != hi) {
                err = HM_ERR_INVALID;
                msg = "internal: inconsistent cluster";
            }
            return it->second;
        }
        const int32_t id = (int32_t)cl.size();
        cl.push_back({depth, first, len, lo, hi});
        mp.emplace(key, id);
        return id;
    }

    void put_dense(int depth, int64_t i0, int64_t m, int64_t j0, int64_t n) {
        leaves.push_back({3, depth, i0, m, j0, n, 0, 0, 0, 0});
        if (m > 0 && n > 0) dense.push_back({i0, j0, (int32_t)m, (int32_t)n, 0});
        dense_entries += m * n;
    }

    // BarycentricMatrix2D(T, f, a, b, c, d, x, y, ir, jr), src/BarycentricMatrix.jl:236
    void put_far(int depth, double a, double b, double c, double d, int64_t i0, int64_t m,
                 int64_t j0, int64_t n) {
        leaves.push_back({2, depth, i0, m, j0, n, a, b, c, d});
        far_rows += m;
        far_cols += n;
        if (m > 0 && n > 0) {
            const int32_t rc = cluster(rcl, rmap, depth, i0, m, a, b);
            const int32_t cc = cluster(ccl, cmap, depth, j0, n, c, d);
            far.push_back({rc, cc, 0});
            if (tok) {
                const int32_t xn = tfind(xt, xtm, depth, a, b), yn = tfind(yt, ytm, depth, c, d);
                if (xn < 0 || yn < 0) {
                    tok = false;
                } else {
                    xt[xn].cl = rc;
                    yt[yn].cl = cc;
                }
            }
            far_depth.push_back(depth);
        }
    }

    void node(int type, int depth, int64_t i0, int64_t ni, int64_t j0, int64_t nj, double a,
              double b, double c, double d) {
        if (err) return;
        maxdepth = std::max(maxdepth, depth);
        if (depth > 200) {
            err = HM_ERR_DEGENERATE;
            msg = "block tree deeper than 200 levels";
            return;
        }
        if (type == 0) ++nd; else if (type == 1) ++nk1; else ++nk2;
        const int64_t ni1 = split(x, nx, i0, ni, a, b), ni2 = ni - ni1;
        const int64_t nj1 = split(y, ny, j0, nj, c, d), nj2 = nj - nj1;
        if (err) return;
        xsplits.push_back({i0, ni, i0 + ni1, mid(a, b)});
        ysplits.push_back({j0, nj, j0 + nj1, mid(c, d)});
        const int64_t i2 = i0 + ni1, j2 = j0 + nj1;
        const double ab2 = mid(a, b), cd2 = mid(c, d);
        if (tok) {  // this node's ranges and their halves in the interval trees
            const int32_t px = tfind(xt, xtm, depth, a, b), py = tfind(yt, ytm, depth, c, d);
            if (px < 0 || py < 0) {
                tok = false;
            } else {
                tadd(xt, xtm, depth + 1, i0, ni1, a, ab2, px, 0);
                tadd(xt, xtm, depth + 1, i2, ni2, ab2, b, px, 1);
                tadd(yt, ytm, depth + 1, j0, nj1, c, cd2, py, 0);
                tadd(yt, ytm, depth + 1, j2, nj2, cd2, d, py, 1);
            }
        }
        const bool leaf = ni1 < bs && ni2 < bs && nj1 < bs && nj2 < bs;
        const int cd_ = depth + 1;
        if (type == 0) {  // KernelMatrix, src/KernelMatrix.jl:49-70
            if (leaf) {
                put_dense(cd_, i0, ni1, j0, nj1);
                put_dense(cd_, i0, ni1, j2, nj2);
                put_dense(cd_, i2, ni2, j0, nj1);
                put_dense(cd_, i2, ni2, j2, nj2);
            } else {
                node(0, cd_, i0, ni1, j0, nj1, a, ab2, c, cd2);
                node(1, cd_, i0, ni1, j2, nj2, a, ab2, cd2, d);
                node(2, cd_, i2, ni2, j0, nj1, ab2, b, c, cd2);
                node(0, cd_, i2, ni2, j2, nj2, ab2, b, cd2, d);
            }
        } else if (type == 1) {  // KernelMatrix1, src/KernelMatrix.jl:72-93
            put_far(cd_, a, ab2, c, cd2, i0, ni1, j0, nj1);
            put_far(cd_, a, ab2, cd2, d, i0, ni1, j2, nj2);
            if (leaf) put_dense(cd_, i2, ni2, j0, nj1);
            else node(1, cd_, i2, ni2, j0, nj1, ab2, b, c, cd2);
            put_far(cd_, ab2, b, cd2, d, i2, ni2, j2, nj2);
        } else {  // KernelMatrix2, src/KernelMatrix.jl:95-116
            put_far(cd_, a, ab2, c, cd2, i0, ni1, j0, nj1);
            if (leaf) put_dense(cd_, i0, ni1, j2, nj2);
            else node(2, cd_, i0, ni1, j2, nj2, a, ab2, cd2, d);
            put_far(cd_, ab2, b, c, cd2, i2, ni2, j0, nj1);
            put_far(cd_, ab2, b, cd2, d, i2, ni2, j2, nj2);
        }
    }
};

// Stable counting sort of the indices 0 .. n-1 by an integer key in
// [0, nkeys): the order std::stable_sort by that key gives, in O(n + nkeys).
template <class Key>
std::vector<int32_t> counting_order(size_t n, int64_t nkeys, Key key) {
    std::vector<int32_t> cnt((size_t)nkeys + 1, 0), out(n);
    for (size_t k = 0; k < n; ++k) cnt[(size_t)key(k) + 1]++;
    for (int64_t q = 0; q < nkeys; ++q) cnt[q + 1] += cnt[q];
    for (size_t k = 0; k < n; ++k) out[cnt[(size_t)key(k)]++] = (int32_t)k;
    return out;
}

// Chunks over one side: boundaries of every block and cluster, split to
// HM_CHUNK; then CSR lists of the dense blocks and clusters covering each.
struct Side {
    std::vector<DChunk> ch;
    std::vector<DNear> near;
    std::vector<DFar> farl;
    std::vector<DTask> tasks;
    int64_t nslots = 0;
};

void build_side(int64_t n, const std::vector<DDense> &dense, bool rows,
                std::vector<DCluster> &cl, Side &S) {
    std::vector<int64_t> bnd;
    const size_t nb = 2 * (dense.size() + cl.size()) + 2;
    const bool marked = (size_t)n + 1 <= 64 * nb;
    if (marked) {
        // sorted distinct boundaries from a mark per point: linear in n
        std::vector<uint8_t> mark((size_t)n + 1, 0);
        mark[0] = mark[n] = 1;
        for (const DDense &D : dense) {
            const int64_t s = rows ? D.row0 : D.col0, l = rows ? D.m : D.n;
            mark[s] = mark[s + l] = 1;
        }
        for (const DCluster &C : cl) mark[C.first] = mark[C.first + C.len] = 1;
        for (int64_t i = 0; i <= n; ++i)
            if (mark[i]) bnd.push_back(i);
    } else {
        bnd.reserve(nb);
        bnd.push_back(0);
        bnd.push_back(n);
        for (const DDense &D : dense) {
            const int64_t s = rows ? D.row0 : D.col0, l = rows ? D.m : D.n;
            bnd.push_back(s);
            bnd.push_back(s + l);
        }
        for (const DCluster &C : cl) {
            bnd.push_back(C.first);
            bnd.push_back(C.first + C.len);
        }
        std::sort(bnd.begin(), bnd.end());
        bnd.erase(std::unique(bnd.begin(), bnd.end()), bnd.end());
    }
    std::vector<int64_t> p0;
    for (size_t k = 0; k + 1 < bnd.size(); ++k)
        for (int64_t s = bnd[k]; s < bnd[k + 1]; s += HM_CHUNK) p0.push_back(s);
    const int64_t nch = (int64_t)p0.size();
    S.ch.assign(nch, DChunk{});
    for (int64_t c = 0; c < nch; ++c) {
        S.ch[c].p0 = p0[c];
        const int64_t e = c + 1 < nch ? p0[c + 1] : n;
        S.ch[c].np = (int32_t)(e - p0[c]);
    }
    // every block and cluster starts and ends on a boundary, and every
    // boundary but n starts a chunk: index the chunks by their first point
    // (only boundary entries are written or read)
    std::unique_ptr<int32_t[]> pos;
    if (marked) {
        pos.reset(new int32_t[(size_t)n + 1]);
        for (int64_t c = 0; c < nch; ++c) pos[p0[c]] = (int32_t)c;
        pos[n] = (int32_t)nch;
    }
    auto range = [&](int64_t s, int64_t l, int64_t &c0, int64_t &c1) {
        if (pos) {
            c0 = pos[s];
            c1 = pos[s + l];
            return;
        }
        c0 = std::lower_bound(p0.begin(), p0.end(), s) - p0.begin();
        c1 = std::lower_bound(p0.begin(), p0.end(), s + l) - p0.begin();
    };
    std::vector<int32_t> ncount(nch + 1, 0), fcount(nch + 1, 0);
    for (const DDense &D : dense) {
        int64_t c0, c1;
        range(rows ? D.row0 : D.col0, rows ? D.m : D.n, c0, c1);
        for (int64_t c = c0; c < c1; ++c) ncount[c + 1]++;
    }
    // clusters in order of depth so every chunk lists coarse levels first
    int32_t maxd = 0;
    for (const DCluster &C : cl) maxd = std::max(maxd, (int32_t)C.depth);
    const std::vector<int32_t> order =
        counting_order(cl.size(), (int64_t)maxd + 1, [&](size_t k) { return cl[k].depth; });
    for (int32_t id : order) {
        const DCluster &C = cl[id];
        int64_t c0, c1;
        range(C.first, C.len, c0, c1);
        for (int64_t c = c0; c < c1; ++c) fcount[c + 1]++;
    }
    for (int64_t c = 0; c < nch; ++c) {
        ncount[c + 1] += ncount[c];
        fcount[c + 1] += fcount[c];
    }
    S.near.assign(ncount[nch], DNear{});
    S.farl.assign(fcount[nch], DFar{});
    std::vector<int32_t> nfill(nch, 0), ffill(nch, 0);
    for (size_t b = 0; b < dense.size(); ++b) {
        const DDense &D = dense[b];
        const int64_t s = rows ? D.row0 : D.col0;
        int64_t c0, c1;
        range(s, rows ? D.m : D.n, c0, c1);
        for (int64_t c = c0; c < c1; ++c)
            S.near[ncount[c] + nfill[c]++] = DNear{(int32_t)b, (int32_t)(p0[c] - s)};
    }
    for (int32_t id : order) {
        const DCluster &C = cl[id];
        int64_t c0, c1;
        range(C.first, C.len, c0, c1);
        for (int64_t c = c0; c < c1; ++c)
            S.farl[fcount[c] + ffill[c]++] = DFar{id, (int32_t)(p0[c] - C.first)};
    }
    // upward-pass tasks: each cluster split into ranges of <= HM_TASK_POINTS
    // points with consecutive slots; launched largest first
    int64_t slot = 0;
    S.tasks.clear();
    for (size_t id = 0; id < cl.size(); ++id) {
        DCluster &C = cl[id];
        C.pbase = slot;
        C.nslots = 0;
        for (int64_t o = 0; o < C.len; o += HM_TASK_POINTS) {
            const int32_t len = (int32_t)std::min<int64_t>(HM_TASK_POINTS, C.len - o);
            DTask T{};
            T.cl = (int32_t)id;
            T.off = (int32_t)o;
            T.len = len;
            T.slot = slot++;
            T.first = C.first + o;
            S.tasks.push_back(T);
            C.nslots++;
        }
    }
    for (DTask &T : S.tasks) T.single = cl[T.cl].nslots == 1 ? 1 : 0;
    S.nslots = slot;
    {
        // largest first, stable: key HM_TASK_POINTS - len for len in 1 .. HM_TASK_POINTS
        const std::vector<int32_t> by_len = counting_order(
            S.tasks.size(), HM_TASK_POINTS, [&](size_t k) { return HM_TASK_POINTS - S.tasks[k].len; });
        std::vector<DTask> sorted(S.tasks.size());
        for (size_t k = 0; k < by_len.size(); ++k) sorted[k] = S.tasks[by_len[k]];
        S.tasks.swap(sorted);
    }
    for (int64_t c = 0; c < nch; ++c) {
        S.ch[c].near_beg = ncount[c];
        S.ch[c].near_end = ncount[c + 1];
        S.ch[c].far_beg = fcount[c];
        S.ch[c].far_end = fcount[c + 1];
    }
}

template <class T>
int upload(T **dst, const std::vector<T> &src) {
    *dst = nullptr;
    if (src.empty()) return HM_OK;
    HM_HIP(hipMalloc((void **)dst, sizeof(T) * src.size()));
    HM_HIP(hipMemcpy(*dst, src.data(), sizeof(T) * src.size(), hipMemcpyHostToDevice));
    return HM_OK;
}

template <class T>
int dalloc(T **dst, int64_t n) {
    *dst = nullptr;
    if (n <= 0) return HM_OK;
    hipError_t e = hipMalloc((void **)dst, sizeof(T) * (size_t)n);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        *dst = nullptr;
        return fail(HM_ERR_OOM, "hipMalloc of %lld %d-byte values failed: %s", (long long)n,
                    (int)sizeof(T), hipGetErrorString(e));
    }
    return HM_OK;
}

// Packs a plan's tables into one device allocation: copied tables first, filled
// by one synchronous copy from a host staging buffer, then uninitialised
// workspace. A plan then pays one hipMalloc, one copy and one hipFree for all
// of them instead of one of each per table (hipFree waits for the device, so
// per-table frees serialise a rebuild-per-step caller). Offsets are 256-byte
// aligned; the pointers stored through add()/reserve() are interior pointers
// into the arena and must never be freed on their own.
struct Staging {
    struct Item {
        void **dst;
        const void *src;  // nullptr for workspace
        size_t bytes;
    };
    std::vector<Item> items;

    template <class T>
    void add(T **dst, const std::vector<T> &src) {
        *dst = nullptr;
        if (!src.empty()) items.push_back({(void **)dst, src.data(), sizeof(T) * src.size()});
    }
    template <class T>
    void reserve(T **dst, int64_t n) {
        *dst = nullptr;
        if (n > 0) items.push_back({(void **)dst, nullptr, sizeof(T) * (size_t)n});
    }
    // sources must stay alive until commit; *arena owns the memory afterwards
    int commit(void **arena) {
        *arena = nullptr;
        const auto up = [](size_t v) { return (v + 255) & ~size_t(255); };
        size_t copied = 0, total = 0;
        for (const Item &I : items)
            if (I.src) copied += up(I.bytes);
        total = copied;
        for (const Item &I : items)
            if (!I.src) total += up(I.bytes);
        if (!total) return HM_OK;
        std::vector<unsigned char> host(copied);
        std::vector<size_t> off(items.size());
        size_t c = 0, w = copied;
        for (size_t q = 0; q < items.size(); ++q) {
            const Item &I = items[q];
            if (I.src) {
                off[q] = c;
                std::memcpy(host.data() + c, I.src, I.bytes);
                c += up(I.bytes);
            } else {
                off[q] = w;
                w += up(I.bytes);
            }
        }
        void *d = nullptr;
        const hipError_t e = hipMalloc(&d, total);
        if (e != hipSuccess) {
            (void)hipGetLastError();
            return fail(HM_ERR_OOM, "hipMalloc of %zu bytes of plan tables failed: %s", total,
                        hipGetErrorString(e));
        }
        *arena = d;
        if (copied) HM_HIP(hipMemcpy(d, host.data(), copied, hipMemcpyHostToDevice));
        for (size_t q = 0; q < items.size(); ++q) *items[q].dst = (unsigned char *)d + off[q];
        return HM_OK;
    }
};

// Device arrays whose element type is the plan's (double or float), and the
// device view built from them.
template <class T>
struct PlanBufs {
    using V2 = typename Real2<T>::type;
    T *x = nullptr, *y = nullptr, *lam = nullptr, *rnodes = nullptr, *cnodes = nullptr,
      *chebA = nullptr;
    V2 *rmap = nullptr, *cmap = nullptr;
    T *A = nullptr, *U = nullptr, *V = nullptr, *F = nullptr;
    T *partial = nullptr, *w = nullptr, *z = nullptr;
    T *xt = nullptr;  // packed right-hand sides of the MFMA passes (ny x 16)
    T *yt = nullptr;  // a split product's far field (hybrid and matrix-free plans)
    // nested far field of the MFMA passes (Float64 plans): the re-expansion
    // matrices, moments per column node, coefficients per row node
    T *nestM = nullptr, *my = nullptr, *mx = nullptr;
    // x .. chebA and the workspace partial .. xt live in this one allocation
    // (Staging); A, U, V and F are allocated on their own
    void *arena = nullptr;
    DevPlanT<T> dp{};

    void release() {
        void *b[] = {arena, A, U, V, F};
        for (void *q : b)
            if (q) (void)hipFree(q);
    }
};

}  // namespace

// the thread-local message of hm_last_error, for the other host files
int hmx_set_error(int code, const char *msg) {
    g_err = msg;
    return code;
}

struct hm_plan {
    int kid = 0, r = 0, bs = 0;
    unsigned flags = 0;
    int64_t nx = 0, ny = 0;
    // host copies
    std::vector<Leaf> leaves;
    std::vector<DDense> dense;
    std::vector<DCluster> rcl, ccl;
    std::vector<DFarBlk> far;
    std::vector<int32_t> far_by_rc_h;
    std::vector<int64_t> row_off, col_off;  // instance offsets (count + 1 each)
    Side rows, cols;
    int64_t nd = 0, nk1 = 0, nk2 = 0, far_rows = 0, far_cols = 0, dense_entries = 0;
    int maxdepth = 0;
    int64_t nonfinite = 0;
    int64_t nonfinite_basis = 0;  // of the generated factors alone (HM_USER plans: U and V)
    // device: element-typed arrays (b64 for Float64 plans, b32 for Float32)
    bool f32 = false;
    PlanBufs<double> b64;
    PlanBufs<float> b32;
    DCluster *d_rcl = nullptr, *d_ccl = nullptr;
    DFarBlk *d_far = nullptr;
    int32_t *d_far_by_rc = nullptr, *d_far_by_cc = nullptr;
    DTrRec *d_tr_rc = nullptr, *d_tr_cc = nullptr;  // translation records in by_rc / by_cc order
    DDense *d_dense = nullptr;
    DChunk *d_rch = nullptr, *d_cch = nullptr;
    DNear *d_rnear = nullptr, *d_cnear = nullptr;
    DFar *d_rfar = nullptr, *d_cfar = nullptr;
    DTask *d_rtask = nullptr, *d_ctask = nullptr;
    std::vector<int32_t> multi_c;     // column clusters with more than one upward task
    std::vector<int32_t> multi_r;     // row clusters with more than one upward task
    int32_t *d_multi_c = nullptr, *d_multi_r = nullptr;
    // nested far field of the MFMA passes (DTNode): the interval trees, the
    // internal nodes of both trees by depth (CSR over depth), one task per
    // non-empty column leaf, each row chunk's leaf
    bool nest = false;
    std::vector<DTNode> xtn, ytn;
    std::vector<int32_t> xlev, xlev_off, ylev, ylev_off;
    std::vector<DTask> yleaf;
    std::vector<int32_t> chunk_leaf;
    DTNode *d_xtn = nullptr, *d_ytn = nullptr;
    int32_t *d_xlev = nullptr, *d_ylev = nullptr, *d_chunk_leaf = nullptr;
    DTask *d_yleaf = nullptr;
    void *d_tables = nullptr;  // the one allocation holding d_rcl .. d_yleaf (Staging)
    // side stream and fork / join events of products whose near field runs
    // beside the far-field chain; made at the first such product that is not
    // being captured into a graph
    hipStream_t side = nullptr;
    hipEvent_t fork = nullptr, join = nullptr;
    bool profile = false;
    bool from_blocks = false;  // built by hm_plan_from_blocks: stored form only
    // plans built from blocks: per leaf, its dense or far block id (-1: an
    // empty leaf, or a rank-0 one), and per far block its active rank (the
    // columns of its U, V and F in use; the rest, up to r, are zero)
    std::vector<int32_t> leaf_blk, far_rank;
    // plans from points: the splits of the tree, for hm_plan_update_points
    std::vector<DSplit> xsplits, ysplits;
    DSplit *d_xsplits = nullptr, *d_ysplits = nullptr;
    int32_t *d_status = nullptr;  // the plan's own status word (in d_tables)
    int max_kb = HM_MAX_KB;  // right-hand sides per forward launch (1, 2, 4 or 8)
    int tuned_kb = HM_MAX_KB, tuned_mode = -1;  // hm_plan_tune's choice
    std::vector<hipEvent_t> prof_events;  // 5 per product

    void drop_events() {
        for (hipEvent_t e : prof_events) (void)hipEventDestroy(e);
        prof_events.clear();
    }

    ~hm_plan() {
        drop_events();
        if (fork) (void)hipEventDestroy(fork);
        if (join) (void)hipEventDestroy(join);
        if (side) (void)hipStreamDestroy(side);
        if (d_tables) (void)hipFree(d_tables);
        b64.release();
        b32.release();
    }

    template <class T>
    PlanBufs<T> &bufs() {
        if constexpr (std::is_same_v<T, double>) return b64;
        else return b32;
    }

    // the element-type-independent part of the device view
    template <class T>
    void fill_structure(DevPlanT<T> &dp) const {
        dp.kid = kid; dp.r = r; dp.nx = nx; dp.ny = ny;
        dp.rcl = d_rcl; dp.ccl = d_ccl;
        dp.nrc = (int32_t)rcl.size(); dp.ncc = (int32_t)ccl.size();
        dp.far = d_far; dp.nfar = (int32_t)far.size();
        dp.far_by_rc = d_far_by_rc; dp.far_by_cc = d_far_by_cc;
        dp.tr_rc = d_tr_rc; dp.tr_cc = d_tr_cc;
        dp.dense = d_dense; dp.ndense = (int32_t)dense.size();
        dp.rch = d_rch; dp.cch = d_cch;
        dp.nrch = (int32_t)rows.ch.size(); dp.ncch = (int32_t)cols.ch.size();
        dp.rnear = d_rnear; dp.cnear = d_cnear;
        dp.rfar = d_rfar; dp.cfar = d_cfar;
        dp.rtask = d_rtask; dp.ctask = d_ctask;
        dp.nrtask = (int32_t)rows.tasks.size(); dp.nctask = (int32_t)cols.tasks.size();
        dp.nslots_r = rows.nslots; dp.nslots_c = cols.nslots;
        dp.multi_r = d_multi_r; dp.multi_c = d_multi_c;
        dp.nmulti_r = (int32_t)multi_r.size(); dp.nmulti_c = (int32_t)multi_c.size();
        dp.nest = nest ? 1 : 0;
        dp.xtn = d_xtn; dp.ytn = d_ytn;
        dp.xlev = d_xlev; dp.ylev = d_ylev;
        dp.xlev_off_h = xlev_off.data(); dp.ylev_off_h = ylev_off.data();
        dp.nxlev = xlev_off.empty() ? 0 : (int32_t)xlev_off.size() - 1;
        dp.nylev = ylev_off.empty() ? 0 : (int32_t)ylev_off.size() - 1;
        dp.chunk_leaf = d_chunk_leaf;
        dp.yleaf = d_yleaf; dp.nyleaf = (int32_t)yleaf.size();
    }

    // The nested far field's tables from the builder's interval trees; nest
    // stays false (the per-level far field is used) unless every half is
    // exactly half of its parent's interval (to 1e-12), every column leaf
    // fits one upward task and every row chunk lies in one row leaf.
    template <class TN>
    void build_nest(const std::vector<TN> &xt, const std::vector<TN> &yt, bool tok) {
        nest = false;
        if (!tok || f32 || xt.empty() || yt.empty()) return;
        auto conv = [](const std::vector<TN> &t, std::vector<DTNode> &o) {
            o.resize(t.size());
            for (size_t k = 0; k < t.size(); ++k) {
                const TN &q = t[k];
                DTNode d{};
                d.first = q.first;
                d.len = (int32_t)q.len;
                d.parent = q.parent;
                d.child[0] = q.child[0];
                d.child[1] = q.child[1];
                d.cl = q.cl;
                d.half = q.half;
                d.cx = 0.5 * (q.lo + q.hi);
                d.cy = 1.0 / (0.5 * (q.hi - q.lo));
                o[k] = d;
            }
            for (size_t k = 0; k < t.size(); ++k) {
                const DTNode &d = o[k];
                if ((d.child[0] < 0) != (d.child[1] < 0) || !std::isfinite(d.cy)) return false;
                if (d.parent < 0) continue;
                const DTNode &P = o[d.parent];
                const double alpha = P.cy / d.cy, beta = (d.cx - P.cx) * P.cy;
                if (!(std::fabs(alpha - 0.5) <= 1e-12) ||
                    !(std::fabs(beta - (d.half ? 0.5 : -0.5)) <= 1e-12))
                    return false;
            }
            return true;
        };
        if (!conv(xt, xtn) || !conv(yt, ytn)) return;
        auto by_depth = [](const std::vector<TN> &t, bool internal_only, std::vector<int32_t> &ids,
                           std::vector<int32_t> &off) {
            int maxd = 0;
            for (const TN &q : t) maxd = std::max(maxd, q.depth);
            off.assign((size_t)maxd + 2, 0);
            ids.clear();
            for (int dd = 0; dd <= maxd; ++dd) {
                for (size_t k = 0; k < t.size(); ++k) {
                    const TN &q = t[k];
                    if (q.depth != dd || q.len == 0) continue;
                    if (internal_only && q.child[0] < 0) continue;
                    ids.push_back((int32_t)k);
                }
                off[dd + 1] = (int32_t)ids.size();
            }
        };
        by_depth(yt, true, ylev, ylev_off);
        by_depth(xt, true, xlev, xlev_off);  // leaves form theirs in the downward pass
        yleaf.clear();
        for (size_t k = 0; k < ytn.size(); ++k) {
            const DTNode &d = ytn[k];
            if (d.child[0] >= 0 || d.len == 0) continue;
            if (d.len > HM_TASK_POINTS) return;
            DTask T{};
            T.cl = (int32_t)k;
            T.off = 0;
            T.len = d.len;
            T.single = 1;
            T.slot = d.cl;
            T.first = d.first;
            T.cx = d.cx;
            T.cy = d.cy;
            yleaf.push_back(T);
        }
        const std::vector<DChunk> &ch = rows.ch;
        chunk_leaf.assign(ch.size(), -1);
        for (size_t k = 0; k < xtn.size(); ++k) {
            const DTNode &d = xtn[k];
            if (d.child[0] >= 0 || d.len == 0) continue;
            size_t c = std::lower_bound(ch.begin(), ch.end(), d.first,
                                        [](const DChunk &a, int64_t v) { return a.p0 < v; }) -
                       ch.begin();
            for (; c < ch.size() && ch[c].p0 < d.first + d.len; ++c) {
                if (ch[c].p0 + ch[c].np > d.first + d.len) return;
                chunk_leaf[c] = (int32_t)k;
            }
        }
        for (int32_t v : chunk_leaf)
            if (v < 0) return;
        nest = true;
    }

    // the clusters with more than one upward task (their slots are summed
    // before the translation reads them)
    void list_multi() {
        for (size_t k = 0; k < ccl.size(); ++k)
            if (ccl[k].nslots > 1) multi_c.push_back((int32_t)k);
        for (size_t k = 0; k < rcl.size(); ++k)
            if (rcl[k].nslots > 1) multi_r.push_back((int32_t)k);
    }
    // translation records of the far blocks in by_rc order (forward: the
    // source is the column cluster) and in by_cc order (transposed: the row
    // cluster)
    void tr_records(const std::vector<int32_t> &by_rc, const std::vector<int32_t> &by_cc,
                    std::vector<DTrRec> &tr_rc, std::vector<DTrRec> &tr_cc) const {
        tr_rc.resize(by_rc.size());
        tr_cc.resize(by_cc.size());
        for (size_t e = 0; e < by_rc.size(); ++e) {
            const DFarBlk &b = far[by_rc[e]];
            const DCluster &C = ccl[b.cc];
            tr_rc[e] = DTrRec{b.foff, C.pbase, b.cc, C.nslots};
        }
        for (size_t e = 0; e < by_cc.size(); ++e) {
            const DFarBlk &b = far[by_cc[e]];
            const DCluster &C = rcl[b.rc];
            tr_cc[e] = DTrRec{b.foff, C.pbase, b.rc, C.nslots};
        }
    }

    int elem_bytes() const { return f32 ? 4 : 8; }

    int64_t basis_rows(const std::vector<DCluster> &cl) const {
        int64_t s = 0;
        for (const DCluster &C : cl) s += C.len;
        return s;
    }
};

extern "C" {

const char *hm_last_error(void) { return g_err.c_str(); }

int hm_version(void) { return 100; }

int hm_blockrank(int mant_bits) {
    if (mant_bits < 1 || mant_bits > 4096) return -1;
    const double v = 0.5 * (std::log(std::ldexp(1.0, mant_bits)) / std::log(3.0 + std::sqrt(8.0)));
    return 2 * (int)std::lround(v);
}

int hm_chebyshevpoints(double *out, int n, int kind) {
    if (!out || n < 1 || (kind != 1 && kind != 2) || (kind == 2 && n < 2))
        return fail(HM_ERR_INVALID, "chebyshevpoints: bad arguments");
    cheb_points(out, n, kind);
    return HM_OK;
}

int hm_chebyshevbarycentricweights(double *out, int n, int kind) {
    if (!out || n < 1 || (kind != 1 && kind != 2))
        return fail(HM_ERR_INVALID, "chebyshevbarycentricweights: bad arguments");
    cheb_weights(out, n, kind);
    return HM_OK;
}

int hm_chebyshevpoints_f32(float *out, int n, int kind) {
    if (!out || n < 1 || (kind != 1 && kind != 2) || (kind == 2 && n < 2))
        return fail(HM_ERR_INVALID, "chebyshevpoints: bad arguments");
    cheb_points_t<float>(out, n, kind);
    return HM_OK;
}

int hm_chebyshevbarycentricweights_f32(float *out, int n, int kind) {
    if (!out || n < 1 || (kind != 1 && kind != 2))
        return fail(HM_ERR_INVALID, "chebyshevbarycentricweights: bad arguments");
    cheb_weights_t<float>(out, n, kind);
    return HM_OK;
}

int hm_indsplit(const double *x, int64_t nx, int64_t first, int64_t len, double a, double b,
                int64_t out[4]) {
    if (!x || !out || len < 0 || first < 1) return fail(HM_ERR_INVALID, "indsplit: bad arguments");
    if (len > 0 && first + len - 1 > nx) return fail(HM_ERR_INVALID, "indsplit: range out of bounds");
    int64_t bad;
    if (len > 0 && !non_increasing(x + first - 1, len, &bad))
        return fail(HM_ERR_NOT_DESCENDING, "indsplit: x not non-increasing at %lld",
                    (long long)(first + bad));
    Builder B{};
    const int64_t l1 = B.split(x, nx, first - 1, len, a, b);
    if (B.err) return fail(B.err, "%s", B.msg.c_str());
    out[0] = first;
    out[1] = l1;
    out[2] = first + l1;
    out[3] = len - l1;
    return HM_OK;
}

struct Inst {
    const double *x, *y;
    int64_t nx, ny;
    double a, b, c, d;
};



// The re-expansion of a half's Chebyshev series into its parent's: with the
// half's variable s and the parent's s/2 + beta (beta = -1/2 for half 0, +1/2
// for half 1), T_k(s/2 + beta) = sum_m M[h][k][m] T_m(s) exactly for k < r, by
// T_{k+1}(u) = 2 u T_k(u) - T_{k-1}(u) and s T_m(s) = (T_{m+1}(s) + T_{|m-1|}(s)) / 2;
// formed in long double and rounded once. Layout M[(h r + k) r + m].
extern "C++" {
template <class T>
std::vector<T> nest_matrices(int r) {
    std::vector<T> out((size_t)2 * r * r);
    for (int h = 0; h < 2; ++h) {
        const long double beta = h ? 0.5L : -0.5L, alpha = 0.5L;
        std::vector<std::vector<long double>> P((size_t)r, std::vector<long double>((size_t)r, 0.0L));
        P[0][0] = 1.0L;
        if (r > 1) {
            P[1][0] = beta;
            P[1][1] = alpha;
        }
        for (int k = 1; k + 1 < r; ++k) {
            std::vector<long double> sp((size_t)r, 0.0L);  // s * P[k]
            for (int m = 0; m <= k; ++m) {
                const long double c = P[k][m];
                if (m == 0) {
                    sp[1] += c;
                } else {
                    sp[m + 1] += 0.5L * c;
                    sp[m - 1] += 0.5L * c;
                }
            }
            for (int j = 0; j < r; ++j) P[k + 1][j] = 2.0L * alpha * sp[j] + 2.0L * beta * P[k][j] - P[k - 1][j];
        }
        for (int k = 0; k < r; ++k)
            for (int m = 0; m < r; ++m) out[((size_t)h * r + k) * r + m] = (T)P[k][m];
    }
    return out;
}
}  // extern "C++"

// The element-typed part of a plan: nodes, weights, cluster maps and the
// Chebyshev-form matrix in T, the device arrays, the device view, and the
// compression (dense blocks, F, U, V) for stored and hybrid plans.
extern "C++" {
template <class T>
int setup_typed(hm_plan *P, const double *x, const double *y, const std::vector<HCluster> &hr,
                const std::vector<HCluster> &hc, unsigned flags, int64_t nA, int64_t nU,
                int64_t nV, int64_t nF, hipStream_t s) {
#pragma clang fp contract(off)
    using V2 = typename Real2<T>::type;
    const int r = P->r;
    PlanBufs<T> &b = P->bufs<T>();
    std::vector<T> xhat(r), lam(r);
    cheb_points_t(xhat.data(), r);
    cheb_weights_t(lam.data(), r);
    // mapped nodes t = ab2 + ba2 * xhat with ab2 = (a+b)/2, ba2 = (b-a)/2 in T
    // (src/BarycentricMatrix.jl:156-166), and per cluster (center,
    // 1/halfwidth): xh = (p - center)/halfwidth maps the interval onto [-1, 1]
    // with the node order of the reference
    auto nodes = [&](const std::vector<HCluster> &h, std::vector<T> &t, std::vector<V2> &m) {
        t.resize(h.size() * r);
        m.resize(h.size());
        for (size_t k = 0; k < h.size(); ++k) {
            const T lo = (T)h[k].lo, hi = (T)h[k].hi;
            const T ab2 = T(0.5) * (lo + hi), ba2 = T(0.5) * (hi - lo);
            for (int q = 0; q < r; ++q) t[k * r + q] = ab2 + ba2 * xhat[q];
            m[k].x = ab2;
            m[k].y = T(1) / ba2;
        }
    };
    std::vector<T> rn, cn;
    std::vector<V2> rmap, cmap;
    nodes(hr, rn, rmap);
    nodes(hc, cn, cmap);
    // A[k][m] = (m ? 2 : 1) T_m(xhat_k) / r, T_m(cos t) = cos(m t), t_k = (2k-1) pi / (2r)
    std::vector<T> chebA((size_t)r * r);
    {
        const long double PI_L = 3.14159265358979323846264338327950288419716939937510L;
        for (int k = 1; k <= r; ++k)
            for (int m = 0; m < r; ++m)
                chebA[(size_t)(k - 1) * r + m] =
                    (T)((m ? 2.0L : 1.0L) * cosl((long double)m * (2 * k - 1) * PI_L / (2.0L * r)) / r);
    }
    const std::vector<T> xv(x, x + P->nx), yv(y, y + P->ny);
    std::vector<T> nestM;
    if (std::is_same_v<T, double> && P->nest) nestM = nest_matrices<T>(r);
    int rc;
#define HM_TRY_T(expr) do { rc = (expr); if (rc != HM_OK) return rc; } while (0)
    const int64_t nslots = std::max(P->rows.nslots, P->cols.nslots);
    const int64_t ncl = std::max(P->rcl.size(), P->ccl.size());
    {
        Staging st;
        st.add(&b.x, xv);
        st.add(&b.y, yv);
        st.add(&b.lam, lam);
        st.add(&b.rnodes, rn);
        st.add(&b.cnodes, cn);
        st.add(&b.rmap, rmap);
        st.add(&b.cmap, cmap);
        st.add(&b.chebA, chebA);
        if (std::is_same_v<T, double> && P->nest) {
            st.add(&b.nestM, nestM);
            st.reserve(&b.my, (int64_t)P->ytn.size() * r * HM_WS_RHS);
            st.reserve(&b.mx, (int64_t)P->xtn.size() * r * HM_WS_RHS);
        }
        st.reserve(&b.partial, nslots * r * HM_WS_RHS);
        st.reserve(&b.w, ncl * r * HM_WS_RHS);
        st.reserve(&b.z, ncl * r * HM_WS_RHS);
        if (std::is_same_v<T, double>) st.reserve(&b.xt, P->ny * HM_WS_RHS);
        if ((flags & HM_PLAN_MATRIX_FREE) || (std::is_same_v<T, float> && (flags & HM_PLAN_HYBRID)))
            st.reserve(&b.yt, std::max(P->nx, P->ny));
        HM_TRY_T(st.commit(&b.arena));
    }
    if (b.my) {  // the moments of empty column nodes are read as zero
        const hipError_t e0 =
            hipMemsetAsync(b.my, 0, sizeof(T) * P->ytn.size() * r * HM_WS_RHS, s);
        if (e0 != hipSuccess) return fail(HM_ERR_HIP, "memset: %s", hipGetErrorString(e0));
    }
    const bool stored = !(flags & HM_PLAN_MATRIX_FREE);
    const bool hybrid = (flags & HM_PLAN_HYBRID) != 0;
    if (stored) {
        HM_TRY_T(dalloc(&b.A, nA));
        HM_TRY_T(dalloc(&b.F, nF));
        if (!hybrid) {
            HM_TRY_T(dalloc(&b.U, nU));
            HM_TRY_T(dalloc(&b.V, nV));
        }
    }
#undef HM_TRY_T
    DevPlanT<T> &dp = b.dp;
    P->fill_structure(dp);
    dp.x = b.x; dp.y = b.y; dp.lam = b.lam;
    dp.rnodes = b.rnodes; dp.cnodes = b.cnodes;
    dp.rmap = b.rmap; dp.cmap = b.cmap; dp.chebA = b.chebA;
    dp.nestM = b.nestM; dp.my = b.my; dp.mx = b.mx;
    dp.A = b.A; dp.U = b.U; dp.V = b.V; dp.F = b.F;

    if (stored) {
        // tile padding (rows beyond a block's or cluster's length) is zeroed
        bool padded = false;
        for (const DDense &D : P->dense) padded |= (D.m % HM_TILE) != 0;
        for (const DCluster &C : P->rcl) padded |= (C.len % HM_TILE) != 0;
        for (const DCluster &C : P->ccl) padded |= (C.len % HM_TILE) != 0;
        const bool user = P->kid == HM_USER;  // A and F come from hm_plan_set_leaf_data
        if (padded || user) {
            T *arrs[] = {b.A, b.U, b.V, user ? b.F : nullptr};
            const int64_t lens[] = {nA, nU, nV, nF};
            for (int q = 0; q < 4; ++q)
                if (lens[q] > 0 && arrs[q]) {
                    hipError_t e0 = hipMemsetAsync((void *)arrs[q], 0, sizeof(T) * lens[q], s);
                    if (e0 != hipSuccess) return fail(HM_ERR_HIP, "memset: %s", hipGetErrorString(e0));
                }
        }
        hipError_t e = hmx_launch_compress(dp, s);
        if (e != hipSuccess) return fail(HM_ERR_HIP, "compression launch: %s", hipGetErrorString(e));
        // count non-finite factors (coincident points give Inf, as in the reference)
        unsigned long long *d_cnt = nullptr;
        e = hipMalloc((void **)&d_cnt, 1024 * sizeof(unsigned long long));
        if (e == hipSuccess) {
            std::vector<unsigned long long> h(1024);
            int64_t total = 0;
            const T *arrs[] = {user ? nullptr : b.A, b.U, b.V, user ? nullptr : b.F};
            const int64_t lens[] = {nA, nU, nV, nF};
            for (int q = 0; q < 4 && e == hipSuccess; ++q) {
                if (lens[q] <= 0 || !arrs[q]) continue;
                e = hipMemsetAsync(d_cnt, 0, 1024 * sizeof(unsigned long long), s);
                if (e == hipSuccess) e = hmx_launch_count_nonfinite(arrs[q], lens[q], d_cnt, s);
                if (e == hipSuccess) e = hipMemcpyAsync(h.data(), d_cnt, 1024 * sizeof(unsigned long long), hipMemcpyDeviceToHost, s);
                if (e == hipSuccess) e = hipStreamSynchronize(s);
                for (auto v : h) total += (int64_t)v;
            }
            (void)hipFree(d_cnt);
            P->nonfinite = total;
            P->nonfinite_basis = total;
        }
        if (e != hipSuccess) return fail(HM_ERR_HIP, "compression: %s", hipGetErrorString(e));
    }
    hipError_t e = hipStreamSynchronize(s);
    if (e != hipSuccess) return fail(HM_ERR_HIP, "compression sync: %s", hipGetErrorString(e));
    return HM_OK;
}
}  // extern "C++"

// Builds every instance's tree, then merges them into one forest over the
// concatenated index space (instance t owns rows [rows_t, rows_t + nx_t) and
// columns [cols_t, cols_t + ny_t)); blocks never cross instances, so every
// pass of a product covers all instances in one launch.
static int make_plan(int kernel_id, const Inst *inst, int64_t count, int blocksize, int rank,
                     unsigned flags, void *stream, hm_plan_t *out, bool f32 = false) {
    if (!out) return fail(HM_ERR_INVALID, "plan pointer is NULL");
    *out = nullptr;
    if (kernel_id < HM_CAUCHY || kernel_id > HM_USER) return fail(HM_ERR_INVALID, "unknown kernel id %d", kernel_id);
    if (kernel_id == HM_USER && (flags & HM_PLAN_MATRIX_FREE))
        return fail(HM_ERR_UNSUPPORTED, "HM_USER: f is the caller's host function, so the factors cannot be regenerated on the device (matrix-free)");
    if (count < 1) return fail(HM_ERR_INVALID, "no instances");
    if (blocksize < 1) return fail(HM_ERR_INVALID, "blocksize must be >= 1");
    if (rank < 1 || rank > 64) return fail(HM_ERR_UNSUPPORTED, "rank %d outside 1..64", rank);
    if (f32 && rank > 32) return fail(HM_ERR_UNSUPPORTED, "Float32 plans take rank <= 32 (got %d)", rank);
    if (flags & ~15u) return fail(HM_ERR_INVALID, "unknown plan flags 0x%x", flags);
    if ((flags & HM_PLAN_MATRIX_FREE) && (flags & HM_PLAN_HYBRID))
        return fail(HM_ERR_INVALID, "HM_PLAN_MATRIX_FREE and HM_PLAN_HYBRID exclude each other");
    if (flags & HM_PLAN_FAR_F32) return fail(HM_ERR_UNSUPPORTED, "FP32 far-field plans are not built by this entry point");
    int64_t NX = 0, NY = 0;
    for (int64_t t = 0; t < count; ++t) {
        const Inst &I = inst[t];
        if (!I.x || !I.y || I.nx < 1 || I.ny < 1)
            return fail(HM_ERR_DEGENERATE, "instance %lld: KernelMatrix needs non-empty x and y (the reference indexes x[1])", (long long)t);
        int64_t bad;
        if (!non_increasing(I.x, I.nx, &bad))
            return fail(HM_ERR_NOT_DESCENDING, "instance %lld: x must be finite and non-increasing (index %lld)", (long long)t, (long long)bad);
        if (!non_increasing(I.y, I.ny, &bad))
            return fail(HM_ERR_NOT_DESCENDING, "instance %lld: y must be finite and non-increasing (index %lld)", (long long)t, (long long)bad);
        NX += I.nx;
        NY += I.ny;
    }
    if (NX >= (1LL << 40) || NY >= (1LL << 40)) return fail(HM_ERR_INVALID, "too many points");
    const double *x = nullptr, *y = nullptr;
    std::vector<double> xcat, ycat;
    if (count == 1) {
        x = inst[0].x;
        y = inst[0].y;
    } else {
        xcat.reserve(NX);
        ycat.reserve(NY);
        for (int64_t t = 0; t < count; ++t) {
            xcat.insert(xcat.end(), inst[t].x, inst[t].x + inst[t].nx);
            ycat.insert(ycat.end(), inst[t].y, inst[t].y + inst[t].ny);
        }
        x = xcat.data();
        y = ycat.data();
    }
    const int64_t nx = NX, ny = NY;
    Builder B{};
    std::vector<int64_t> row_off(count + 1, 0), col_off(count + 1, 0);
    for (int64_t t = 0; t < count; ++t) {
        const Inst &I = inst[t];
        Builder Bt{};
        Bt.x = I.x; Bt.y = I.y; Bt.nx = I.nx; Bt.ny = I.ny; Bt.bs = blocksize; Bt.r = rank;
        Bt.f32 = f32;
        {
            // a banded tree holds about 18 leaves, 12 far blocks and 4 clusters
            // per side for every blocksize points
            const size_t q = (size_t)(std::max(I.nx, I.ny) / std::max(blocksize, 1)) + 16;
            Bt.leaves.reserve(20 * q);
            Bt.far.reserve(14 * q);
            Bt.far_depth.reserve(14 * q);
            Bt.dense.reserve(8 * q);
            Bt.rcl.reserve(5 * q);
            Bt.ccl.reserve(5 * q);
            Bt.rmap.reserve(5 * q);
            Bt.cmap.reserve(5 * q);
        }
        Bt.tadd(Bt.xt, Bt.xtm, 0, 0, I.nx, I.a, I.b, -1, 0);
        Bt.tadd(Bt.yt, Bt.ytm, 0, 0, I.ny, I.c, I.d, -1, 0);
        Bt.node(0, 0, 0, I.nx, 0, I.ny, I.a, I.b, I.c, I.d);
        if (Bt.err) {
            return count == 1 ? fail(Bt.err, "%s", Bt.msg.c_str())
                              : fail(Bt.err, "instance %lld: %s", (long long)t, Bt.msg.c_str());
        }
        const int64_t ro = row_off[t], co = col_off[t];
        row_off[t + 1] = ro + I.nx;
        col_off[t + 1] = co + I.ny;
        if (count == 1) {  // offsets are zero: take the tree as built
            B = std::move(Bt);
            break;
        }
        const int32_t rb = (int32_t)B.rcl.size(), cb = (int32_t)B.ccl.size();
        for (Leaf L : Bt.leaves) { L.row0 += ro; L.col0 += co; B.leaves.push_back(L); }
        for (DDense D : Bt.dense) { D.row0 += ro; D.col0 += co; B.dense.push_back(D); }
        for (HCluster h : Bt.rcl) { h.first += ro; B.rcl.push_back(h); }
        for (HCluster h : Bt.ccl) { h.first += co; B.ccl.push_back(h); }
        for (DFarBlk f : Bt.far) { f.rc += rb; f.cc += cb; B.far.push_back(f); }
        for (DSplit q : Bt.xsplits) { q.first += ro; q.at += ro; B.xsplits.push_back(q); }
        for (DSplit q : Bt.ysplits) { q.first += co; q.at += co; B.ysplits.push_back(q); }
        {
            auto merge = [](std::vector<Builder::TNode> &dst, const std::vector<Builder::TNode> &src,
                            int64_t off, int32_t cbase) {
                const int32_t base = (int32_t)dst.size();
                for (Builder::TNode q : src) {
                    q.first += off;
                    if (q.cl >= 0) q.cl += cbase;
                    if (q.parent >= 0) q.parent += base;
                    for (int h = 0; h < 2; ++h)
                        if (q.child[h] >= 0) q.child[h] += base;
                    dst.push_back(q);
                }
            };
            merge(B.xt, Bt.xt, ro, rb);
            merge(B.yt, Bt.yt, co, cb);
            B.tok = B.tok && Bt.tok;
        }
        B.nd += Bt.nd; B.nk1 += Bt.nk1; B.nk2 += Bt.nk2;
        B.far_rows += Bt.far_rows; B.far_cols += Bt.far_cols; B.dense_entries += Bt.dense_entries;
        B.maxdepth = std::max(B.maxdepth, Bt.maxdepth);
    }

    hm_plan *P = new hm_plan();
    P->kid = kernel_id; P->r = rank; P->bs = blocksize; P->flags = flags;
    P->nx = nx; P->ny = ny;
    P->f32 = f32;
    P->nd = B.nd; P->nk1 = B.nk1; P->nk2 = B.nk2; P->maxdepth = B.maxdepth;
    P->far_rows = B.far_rows; P->far_cols = B.far_cols; P->dense_entries = B.dense_entries;
    P->leaves = std::move(B.leaves);
    P->xsplits = std::move(B.xsplits);
    P->ysplits = std::move(B.ysplits);
    P->row_off = std::move(row_off);
    P->col_off = std::move(col_off);
    P->dense = std::move(B.dense);
    P->far = std::move(B.far);
    const int r = rank;

    // storage offsets
    int64_t off = 0;
    for (DDense &D : P->dense) { D.aoff = off; off += tiled_size(D.m, D.n); }
    const int64_t nA = off;
    auto mk = [&](const std::vector<HCluster> &hc, std::vector<DCluster> &dc) {
        dc.resize(hc.size());
        int64_t o = 0;
        for (size_t k = 0; k < hc.size(); ++k) {
            dc[k] = DCluster{};
            dc[k].first = hc[k].first;
            dc[k].len = (int32_t)hc[k].len;
            dc[k].boff = o;
            dc[k].depth = hc[k].depth;
            o += tiled_size(hc[k].len, r);
        }
        return o;
    };
    const int64_t nU = mk(B.rcl, P->rcl), nV = mk(B.ccl, P->ccl);
    for (size_t k = 0; k < P->far.size(); ++k) P->far[k].foff = (int64_t)k * r * r;
    const int64_t nF = (int64_t)P->far.size() * r * r;

    // far blocks grouped by row / column cluster
    std::vector<int32_t> by_rc = counting_order(P->far.size(), (int64_t)P->rcl.size(),
                                                [&](size_t k) { return P->far[k].rc; });
    std::vector<int32_t> by_cc = counting_order(P->far.size(), (int64_t)P->ccl.size(),
                                                [&](size_t k) { return P->far[k].cc; });
    for (auto &C : P->rcl) C.far_beg = C.far_end = 0;
    for (auto &C : P->ccl) C.far_beg = C.far_end = 0;
    for (size_t k = 0; k < by_rc.size(); ++k) {
        DCluster &C = P->rcl[P->far[by_rc[k]].rc];
        if (C.far_end == 0) C.far_beg = (int32_t)k;
        C.far_end = (int32_t)k + 1;
    }
    for (size_t k = 0; k < by_cc.size(); ++k) {
        DCluster &C = P->ccl[P->far[by_cc[k]].cc];
        if (C.far_end == 0) C.far_beg = (int32_t)k;
        C.far_end = (int32_t)k + 1;
    }

    build_side(nx, P->dense, true, P->rcl, P->rows);
    build_side(ny, P->dense, false, P->ccl, P->cols);
    P->list_multi();
    P->build_nest(B.xt, B.yt, B.tok);
    // the cluster maps the Float64 MFMA upward pass reads from its task
    // record: (a+b)/2 and 1/((b-a)/2) as the Float64 cluster maps form them
    if (!f32) {
        for (DTask &T : P->cols.tasks) {
            const HCluster &h = B.ccl[T.cl];
            const double ab2 = 0.5 * (h.lo + h.hi), ba2 = 0.5 * (h.hi - h.lo);
            T.cx = ab2;
            T.cy = 1.0 / ba2;
        }
        for (DTask &T : P->rows.tasks) {
            const HCluster &h = B.rcl[T.cl];
            const double ab2 = 0.5 * (h.lo + h.hi), ba2 = 0.5 * (h.hi - h.lo);
            T.cx = ab2;
            T.cy = 1.0 / ba2;
        }
    }

    if (flags & HM_PLAN_TREE_ONLY) {
        P->far_by_rc_h = std::move(by_rc);
        *out = P;
        return HM_OK;
    }
    std::vector<DTrRec> tr_rc, tr_cc;
    P->tr_records(by_rc, by_cc, tr_rc, tr_cc);
    int rc_;
#define HM_TRY(expr) do { rc_ = (expr); if (rc_ != HM_OK) { delete P; return rc_; } } while (0)
    {
        Staging st;
        st.add(&P->d_rcl, P->rcl);
        st.add(&P->d_ccl, P->ccl);
        st.add(&P->d_far, P->far);
        st.add(&P->d_far_by_rc, by_rc);
        st.add(&P->d_far_by_cc, by_cc);
        st.add(&P->d_tr_rc, tr_rc);
        st.add(&P->d_tr_cc, tr_cc);
        st.add(&P->d_dense, P->dense);
        st.add(&P->d_rch, P->rows.ch);
        st.add(&P->d_cch, P->cols.ch);
        st.add(&P->d_rnear, P->rows.near);
        st.add(&P->d_cnear, P->cols.near);
        st.add(&P->d_rfar, P->rows.farl);
        st.add(&P->d_cfar, P->cols.farl);
        st.add(&P->d_rtask, P->rows.tasks);
        st.add(&P->d_ctask, P->cols.tasks);
        st.add(&P->d_multi_c, P->multi_c);
        st.add(&P->d_multi_r, P->multi_r);
        st.add(&P->d_xtn, P->xtn);
        st.add(&P->d_ytn, P->ytn);
        st.add(&P->d_xlev, P->xlev);
        st.add(&P->d_ylev, P->ylev);
        st.add(&P->d_chunk_leaf, P->chunk_leaf);
        st.add(&P->d_yleaf, P->yleaf);
        st.add(&P->d_xsplits, P->xsplits);
        st.add(&P->d_ysplits, P->ysplits);
        st.reserve(&P->d_status, 1);
        HM_TRY(st.commit(&P->d_tables));
    }
    P->far_by_rc_h = std::move(by_rc);
    if (f32) HM_TRY(setup_typed<float>(P, x, y, B.rcl, B.ccl, flags, nA, nU, nV, nF, (hipStream_t)stream));
    else HM_TRY(setup_typed<double>(P, x, y, B.rcl, B.ccl, flags, nA, nU, nV, nF, (hipStream_t)stream));
#undef HM_TRY
    *out = P;
    return HM_OK;
}

int hm_plan_from_points(int kernel_id, const double *x, int64_t nx, const double *y, int64_t ny,
                        double a, double b, double c, double d, int blocksize, int rank,
                        unsigned flags, void *stream, hm_plan_t *out) {
    const Inst I{x, y, nx, ny, a, b, c, d};
    return make_plan(kernel_id, &I, 1, blocksize, rank, flags, stream, out);
}

int hm_plan_from_points_f32(int kernel_id, const float *x, int64_t nx, const float *y, int64_t ny,
                            float a, float b, float c, float d, int blocksize, int rank,
                            unsigned flags, void *stream, hm_plan_t *out) {
    if (out) *out = nullptr;
    if (!x || !y || nx < 1 || ny < 1)
        return fail(HM_ERR_DEGENERATE, "KernelMatrix needs non-empty x and y (the reference indexes x[1])");
    // every float is exact in double; the tree and the factors are formed in
    // float arithmetic from these values
    const std::vector<double> xd(x, x + nx), yd(y, y + ny);
    const Inst I{xd.data(), yd.data(), nx, ny, a, b, c, d};
    return make_plan(kernel_id, &I, 1, blocksize, rank, flags, stream, out, true);
}

// A plan from leaves (hm_plan_from_blocks and its Float32 twin). Leaf kinds:
// 3 dense, 2 low rank U diag(S) V', 4 barycentric U F V' (src/algebra.jl:
// 243-277) with the rank x rank core F passed in S[t]. Low-rank and barycentric
// leaves become far blocks with private row and column clusters; their cores
// sit in the plan's r x r F blocks (r = the largest leaf rank, the rest zero).
extern "C++" {
template <class T>
static int from_blocks_t(int64_t nrows, int64_t ncols, int64_t nleaves, const int32_t *kind,
                         const int64_t *row0, const int64_t *col0, const int64_t *m,
                         const int64_t *n, const int64_t *rank, const T *const *A,
                         const T *const *U, const T *const *S, const T *const *V,
                         void *stream, hm_plan_t *out) {
    constexpr bool F32 = std::is_same_v<T, float>;
    constexpr int RMAX = F32 ? 32 : 64;
    if (!out) return fail(HM_ERR_INVALID, "plan pointer is NULL");
    *out = nullptr;
    if (nrows < 1 || ncols < 1 || nleaves < 0 || (nleaves > 0 && (!kind || !row0 || !col0 || !m || !n)))
        return fail(HM_ERR_INVALID, "plan_from_blocks: bad arguments");
    if (nrows >= (1LL << 40) || ncols >= (1LL << 40)) return fail(HM_ERR_INVALID, "too many rows");
    int R = 1;
    for (int64_t t = 0; t < nleaves; ++t) {
        if (kind[t] != 2 && kind[t] != 3 && kind[t] != 4)
            return fail(HM_ERR_INVALID, "leaf %lld: kind must be 2, 3 or 4", (long long)t);
        if (m[t] < 0 || n[t] < 0 || row0[t] < 0 || col0[t] < 0 || row0[t] + m[t] > nrows || col0[t] + n[t] > ncols)
            return fail(HM_ERR_INVALID, "leaf %lld out of range", (long long)t);
        if (m[t] >= (1LL << 31) || n[t] >= (1LL << 31)) return fail(HM_ERR_INVALID, "leaf %lld too large", (long long)t);
        if (kind[t] == 3 && m[t] * n[t] > 0 && (!A || !A[t])) return fail(HM_ERR_INVALID, "leaf %lld: no data", (long long)t);
        if (kind[t] != 3) {
            if (!rank || rank[t] < 0) return fail(HM_ERR_INVALID, "leaf %lld: bad rank", (long long)t);
            if (rank[t] > RMAX)
                return fail(HM_ERR_UNSUPPORTED, "leaf %lld: rank %lld > %d", (long long)t, (long long)rank[t], RMAX);
            if (rank[t] > 0 && (!U || !S || !V || !U[t] || !S[t] || !V[t]))
                return fail(HM_ERR_INVALID, "leaf %lld: no factors", (long long)t);
            R = std::max<int>(R, (int)rank[t]);
        }
    }
    hm_plan *P = new hm_plan();
    P->from_blocks = true;
    P->f32 = F32;
    P->kid = 0; P->r = R; P->bs = 0; P->flags = 0;
    P->nx = nrows; P->ny = ncols;
    P->row_off = {0, nrows};
    P->col_off = {0, ncols};
    std::vector<HCluster> hr, hc;
    std::vector<int64_t> lr_leaf;  // leaf index of each far block
    for (int64_t t = 0; t < nleaves; ++t) {
        P->leaves.push_back({kind[t] == 3 ? 3 : 2, 1, row0[t], m[t], col0[t], n[t], 0, 0, 0, 0});
        P->leaf_blk.push_back(-1);
        if (m[t] == 0 || n[t] == 0) continue;
        if (kind[t] == 3) {
            P->leaf_blk.back() = (int32_t)P->dense.size();
            P->dense.push_back({row0[t], col0[t], (int32_t)m[t], (int32_t)n[t], 0});
            P->dense_entries += m[t] * n[t];
        } else if (rank[t] > 0) {
            const int32_t rc = (int32_t)hr.size(), cc = (int32_t)hc.size();
            P->leaf_blk.back() = (int32_t)P->far.size();
            P->far_rank.push_back((int32_t)rank[t]);
            hr.push_back({1, row0[t], m[t], 0, 0});
            hc.push_back({1, col0[t], n[t], 0, 0});
            P->far.push_back({rc, cc, 0});
            lr_leaf.push_back(t);
            P->far_rows += m[t];
            P->far_cols += n[t];
        }
    }
    const int r = R;
    int64_t off = 0;
    for (DDense &D : P->dense) { D.aoff = off; off += tiled_size(D.m, D.n); }
    const int64_t nA = off;
    auto mk = [&](const std::vector<HCluster> &h, std::vector<DCluster> &dc) {
        dc.resize(h.size());
        int64_t o = 0;
        for (size_t k = 0; k < h.size(); ++k) {
            dc[k] = DCluster{};
            dc[k].first = h[k].first;
            dc[k].len = (int32_t)h[k].len;
            dc[k].boff = o;
            dc[k].depth = 1;
            dc[k].far_beg = (int32_t)k;
            dc[k].far_end = (int32_t)k + 1;
            o += tiled_size(h[k].len, r);
        }
        return o;
    };
    const int64_t nU = mk(hr, P->rcl), nV = mk(hc, P->ccl);
    for (size_t k = 0; k < P->far.size(); ++k) P->far[k].foff = (int64_t)k * r * r;
    const int64_t nF = (int64_t)P->far.size() * r * r;
    std::vector<int32_t> by_rc(P->far.size()), by_cc(P->far.size());
    for (size_t k = 0; k < P->far.size(); ++k) by_rc[k] = by_cc[k] = (int32_t)k;
    build_side(nrows, P->dense, true, P->rcl, P->rows);
    build_side(ncols, P->dense, false, P->ccl, P->cols);
    P->list_multi();
    // host-side tiling of the data
    std::vector<T> hA(nA, T(0)), hU(nU, T(0)), hV(nV, T(0)), hF(nF, T(0));
    {
        size_t d = 0;
        for (int64_t t = 0; t < nleaves; ++t) {
            if (kind[t] != 3 || m[t] == 0 || n[t] == 0) continue;
            const DDense &D = P->dense[d++];
            for (int64_t j = 0; j < D.n; ++j)
                for (int64_t i = 0; i < D.m; ++i) hA[D.aoff + tile_index(i, j, D.n)] = A[t][i + D.m * j];
        }
        for (size_t f = 0; f < lr_leaf.size(); ++f) {
            const int64_t t = lr_leaf[f], k = rank[t];
            const DCluster &Cr = P->rcl[f], &Cc = P->ccl[f];
            for (int64_t c = 0; c < k; ++c) {
                for (int64_t i = 0; i < m[t]; ++i) hU[Cr.boff + tile_index(i, c, r)] = U[t][i + m[t] * c];
                for (int64_t j = 0; j < n[t]; ++j) hV[Cc.boff + tile_index(j, c, r)] = V[t][j + n[t] * c];
                if (kind[t] == 2) hF[(int64_t)f * r * r + c + (int64_t)r * c] = S[t][c];
                else
                    for (int64_t c2 = 0; c2 < k; ++c2)
                        hF[(int64_t)f * r * r + c2 + (int64_t)r * c] = S[t][c2 + k * c];
            }
        }
    }
    std::vector<DTrRec> tr_rc, tr_cc;
    P->tr_records(by_rc, by_cc, tr_rc, tr_cc);
    int rc_;
#define HM_TRY(expr) do { rc_ = (expr); if (rc_ != HM_OK) { delete P; return rc_; } } while (0)
    {
        Staging st;
        st.add(&P->d_rcl, P->rcl);
        st.add(&P->d_ccl, P->ccl);
        st.add(&P->d_far, P->far);
        st.add(&P->d_far_by_rc, by_rc);
        st.add(&P->d_far_by_cc, by_cc);
        st.add(&P->d_tr_rc, tr_rc);
        st.add(&P->d_tr_cc, tr_cc);
        st.add(&P->d_dense, P->dense);
        st.add(&P->d_rch, P->rows.ch);
        st.add(&P->d_cch, P->cols.ch);
        st.add(&P->d_rnear, P->rows.near);
        st.add(&P->d_cnear, P->cols.near);
        st.add(&P->d_rfar, P->rows.farl);
        st.add(&P->d_cfar, P->cols.farl);
        st.add(&P->d_rtask, P->rows.tasks);
        st.add(&P->d_ctask, P->cols.tasks);
        st.add(&P->d_multi_c, P->multi_c);
        st.add(&P->d_multi_r, P->multi_r);
        st.add(&P->d_xtn, P->xtn);
        st.add(&P->d_ytn, P->ytn);
        st.add(&P->d_xlev, P->xlev);
        st.add(&P->d_ylev, P->ylev);
        st.add(&P->d_chunk_leaf, P->chunk_leaf);
        st.add(&P->d_yleaf, P->yleaf);
        HM_TRY(st.commit(&P->d_tables));
    }
    P->far_by_rc_h = std::move(by_rc);
    PlanBufs<T> &b = P->bufs<T>();
    HM_TRY(upload(&b.A, hA));
    HM_TRY(upload(&b.U, hU));
    HM_TRY(upload(&b.V, hV));
    HM_TRY(upload(&b.F, hF));
    const int64_t nslots = std::max(P->rows.nslots, P->cols.nslots);
    const int64_t ncl = std::max(P->rcl.size(), P->ccl.size());
    {
        Staging st;
        st.reserve(&b.partial, std::max<int64_t>(nslots, 1) * r * HM_WS_RHS);
        st.reserve(&b.w, std::max<int64_t>(ncl, 1) * r * HM_WS_RHS);
        st.reserve(&b.z, std::max<int64_t>(ncl, 1) * r * HM_WS_RHS);
        HM_TRY(st.commit(&b.arena));
    }
#undef HM_TRY
    DevPlanT<T> &dp = b.dp;
    P->fill_structure(dp);
    dp.A = b.A; dp.U = b.U; dp.V = b.V; dp.F = b.F;
    (void)stream;
    *out = P;
    return HM_OK;
}
}  // extern "C++"

int hm_plan_from_blocks(int64_t nrows, int64_t ncols, int64_t nleaves, const int32_t *kind,
                        const int64_t *row0, const int64_t *col0, const int64_t *m,
                        const int64_t *n, const int64_t *rank, const double *const *A,
                        const double *const *U, const double *const *S,
                        const double *const *V, void *stream, hm_plan_t *out) {
    return from_blocks_t<double>(nrows, ncols, nleaves, kind, row0, col0, m, n, rank, A, U, S, V,
                                 stream, out);
}

int hm_plan_from_blocks_f32(int64_t nrows, int64_t ncols, int64_t nleaves, const int32_t *kind,
                            const int64_t *row0, const int64_t *col0, const int64_t *m,
                            const int64_t *n, const int64_t *rank, const float *const *A,
                            const float *const *U, const float *const *S,
                            const float *const *V, void *stream, hm_plan_t *out) {
    return from_blocks_t<float>(nrows, ncols, nleaves, kind, row0, col0, m, n, rank, A, U, S, V,
                                stream, out);
}

// Dense leaves and F of an HM_USER plan from the caller's values of f.
extern "C++" {
template <class T>
static int check_type(hm_plan_t P, const char *what);

template <class T>
static int set_leaf_data_t(hm_plan_t P, const T *A, int64_t nA, const T *F, int64_t nF,
                           void *stream) {
    if (!P) return fail(HM_ERR_INVALID, "set_leaf_data: NULL plan");
    if (P->kid != HM_USER || P->from_blocks)
        return fail(HM_ERR_INVALID, "set_leaf_data: only plans built with kernel HM_USER");
    if (P->flags & (HM_PLAN_TREE_ONLY | HM_PLAN_MATRIX_FREE))
        return fail(HM_ERR_INVALID, "set_leaf_data: the plan stores no dense blocks and F");
    if (int rc = check_type<T>(P, "set_leaf_data")) return rc;
    const int64_t r = P->r;
    int64_t wantA = 0;
    for (const DDense &D : P->dense) wantA += (int64_t)D.m * D.n;
    const int64_t wantF = (int64_t)P->far.size() * r * r;
    if (nA != wantA || nF != wantF)
        return fail(HM_ERR_INVALID, "set_leaf_data: %lld dense and %lld far values given, the plan has %lld and %lld",
                    (long long)nA, (long long)nF, (long long)wantA, (long long)wantF);
    if ((nA > 0 && !A) || (nF > 0 && !F)) return fail(HM_ERR_INVALID, "set_leaf_data: NULL data");
    PlanBufs<T> &b = P->bufs<T>();
    int64_t tot = 0;
    for (const DDense &D : P->dense) tot = std::max<int64_t>(tot, D.aoff + tiled_size(D.m, D.n));
    std::vector<T> hA((size_t)tot, T(0));
    int64_t nonfinite = 0;
    const T *src = A;
    for (const DDense &D : P->dense) {
        for (int64_t j = 0; j < D.n; ++j)
            for (int64_t i = 0; i < D.m; ++i) {
                const T v = src[i + (int64_t)D.m * j];
                nonfinite += std::isfinite(v) ? 0 : 1;
                hA[(size_t)(D.aoff + tile_index(i, j, D.n))] = v;
            }
        src += (int64_t)D.m * D.n;
    }
    for (int64_t q = 0; q < nF; ++q) nonfinite += std::isfinite(F[q]) ? 0 : 1;
    hipStream_t s = (hipStream_t)stream;
    if (tot > 0) HM_HIP(hipMemcpyAsync(b.A, hA.data(), sizeof(T) * (size_t)tot, hipMemcpyHostToDevice, s));
    if (nF > 0) HM_HIP(hipMemcpyAsync(b.F, F, sizeof(T) * (size_t)nF, hipMemcpyHostToDevice, s));
    HM_HIP(hipStreamSynchronize(s));  // the host staging is released on return
    P->nonfinite = P->nonfinite_basis + nonfinite;
    return HM_OK;
}
}  // extern "C++"

int hm_plan_set_leaf_data(hm_plan_t P, const double *A, int64_t nA, const double *F, int64_t nF,
                          void *stream) {
    return set_leaf_data_t(P, A, nA, F, nF, stream);
}

int hm_plan_set_leaf_data_f32(hm_plan_t P, const float *A, int64_t nA, const float *F,
                              int64_t nF, void *stream) {
    return set_leaf_data_t(P, A, nA, F, nF, stream);
}

extern "C++" {
template <class T>
static int update_points_t(hm_plan_t P, const T *x, const T *y, int32_t *status, void *stream) {
    if (!P || !x || !y) return fail(HM_ERR_INVALID, "update_points: NULL argument");
    if (P->from_blocks) return fail(HM_ERR_UNSUPPORTED, "update_points: a plan built from blocks has no points");
    if (P->flags & HM_PLAN_TREE_ONLY) return fail(HM_ERR_INVALID, "update_points: tree-only plan");
    if (int rc = check_type<T>(P, "update_points")) return rc;
    hipStream_t s = (hipStream_t)stream;
    PlanBufs<T> &b = P->bufs<T>();
    int32_t *st = status ? status : P->d_status;
    hipError_t e = hmx_launch_update_points(x, b.x, P->nx, P->d_xsplits, (int64_t)P->xsplits.size(),
                                            y, b.y, P->ny, P->d_ysplits, (int64_t)P->ysplits.size(),
                                            st, s);
    // the factors that depend on the points (dense blocks; U and V of stored
    // plans) are formed again; F depends on the intervals alone
    if (e == hipSuccess && !(P->flags & HM_PLAN_MATRIX_FREE)) e = hmx_launch_refit(b.dp, s);
    if (e != hipSuccess) return fail(HM_ERR_HIP, "update_points launch: %s", hipGetErrorString(e));
    return HM_OK;
}
}  // extern "C++"

int hm_plan_update_points(hm_plan_t P, const double *x, const double *y, int32_t *status,
                          void *stream) {
    return update_points_t(P, x, y, status, stream);
}

int hm_plan_update_points_f32(hm_plan_t P, const float *x, const float *y, int32_t *status,
                              void *stream) {
    return update_points_t(P, x, y, status, stream);
}

int hm_plan_add_col(hm_plan_t P, const double *u, int64_t ioff, int64_t j, void *stream) {
    if (!P || !u || ioff < 0) return fail(HM_ERR_INVALID, "add_col: bad arguments");
    if (!P->from_blocks)
        return fail(HM_ERR_UNSUPPORTED, "add_col: only plans built from blocks (the reference defines add_col! for HierarchicalMatrix)");
    if (P->f32) return fail(HM_ERR_INVALID, "add_col: a Float32 plan (Float64 entry point)");
    if (j < 0 || j >= P->ny) return fail(HM_ERR_INVALID, "add_col: column %lld outside 0..%lld", (long long)j, (long long)P->ny - 1);
    // the leaves holding column j: dense ones add u, low-rank ones take it as
    // one more rank term (u in U, e_j in V, 1 on F's diagonal) in their next
    // free column; all are checked before anything changes
    std::vector<AddColJob> jobs;
    for (size_t t = 0; t < P->leaves.size(); ++t) {
        const Leaf &L = P->leaves[t];
        if (L.m == 0 || L.n == 0 || j < L.col0 || j >= L.col0 + L.n) continue;
        const int32_t id = P->leaf_blk[t];
        AddColJob q{};
        q.m = (int32_t)L.m;
        q.jl = (int32_t)(j - L.col0);
        q.row0 = L.row0;
        if (L.code == 3) {
            const DDense &D = P->dense[id];
            q.kind = 3;
            q.n = D.n;
            q.aoff = D.aoff;
        } else {
            if (id < 0 || P->far_rank[id] >= P->r)
                return fail(HM_ERR_UNSUPPORTED,
                            "add_col: leaf %zu has no free rank column (the plan's rank is %d); rebuild the plan from the leaves",
                            t, P->r);
            const DFarBlk &F = P->far[id];
            q.kind = 2;
            q.c = P->far_rank[id];
            q.uoff = P->rcl[F.rc].boff;
            q.voff = P->ccl[F.cc].boff;
            q.foff = F.foff;
            q.far = id;
        }
        jobs.push_back(q);
    }
    hipStream_t s = (hipStream_t)stream;
    for (size_t b = 0; b < jobs.size(); b += HM_ADD_COL_JOBS) {
        AddColJobs J{};
        J.count = (int32_t)std::min<size_t>(HM_ADD_COL_JOBS, jobs.size() - b);
        for (int q = 0; q < J.count; ++q) J.job[q] = jobs[b + q];
        const hipError_t e = hmx_launch_add_col(P->b64.dp, J, u + ioff, s);
        if (e != hipSuccess) return fail(HM_ERR_HIP, "add_col launch: %s", hipGetErrorString(e));
    }
    for (const AddColJob &q : jobs)
        if (q.kind == 2) ++P->far_rank[q.far];
    return HM_OK;
}

int hm_plan_scale(hm_plan_t P, int side, const double *b, int64_t off, void *stream) {
    if (!P || !b || off < 0 || (side != 0 && side != 1))
        return fail(HM_ERR_INVALID, "scale: bad arguments");
    if (!P->from_blocks)
        return fail(HM_ERR_UNSUPPORTED, "scale: only plans built from blocks (the reference has no scale! for KernelMatrix leaves)");
    if (P->f32) return fail(HM_ERR_INVALID, "scale: a Float32 plan (Float64 entry point)");
    hipError_t e = hmx_launch_scale(P->b64.dp, side, b + off, (hipStream_t)stream);
    if (e != hipSuccess) return fail(HM_ERR_HIP, "scale launch: %s", hipGetErrorString(e));
    return HM_OK;
}

int hm_plan_batched(int kernel_id, int64_t count, const double *const *x, const int64_t *nx,
                    const double *const *y, const int64_t *ny, const double *intervals,
                    int blocksize, int rank, unsigned flags, void *stream, hm_plan_t *out) {
    if (out) *out = nullptr;
    if (count < 1 || !x || !nx || !y || !ny || !intervals)
        return fail(HM_ERR_INVALID, "plan_batched: bad arguments");
    std::vector<Inst> inst(count);
    for (int64_t t = 0; t < count; ++t)
        inst[t] = Inst{x[t], y[t], nx[t], ny[t], intervals[4 * t], intervals[4 * t + 1],
                       intervals[4 * t + 2], intervals[4 * t + 3]};
    return make_plan(kernel_id, inst.data(), count, blocksize, rank, flags, stream, out);
}

int hm_plan_batch_offsets(hm_plan_t P, int64_t *row_offsets, int64_t *col_offsets) {
    if (!P) return fail(HM_ERR_INVALID, "batch_offsets: NULL plan");
    for (size_t t = 0; t < P->row_off.size(); ++t) {
        if (row_offsets) row_offsets[t] = P->row_off[t];
        if (col_offsets) col_offsets[t] = P->col_off[t];
    }
    return HM_OK;
}

int hm_plan_destroy(hm_plan_t plan) {
    if (!plan) return HM_OK;
    if (!(plan->flags & HM_PLAN_TREE_ONLY)) (void)hipDeviceSynchronize();
    delete plan;
    return HM_OK;
}

int hm_plan_info(hm_plan_t P, int64_t *info, int n) {
    if (!P || !info || n < 0) return fail(HM_ERR_INVALID, "plan_info: bad arguments");
    int64_t v[HM_INFO_COUNT] = {0};
    const int64_t r = P->r;
    const int64_t bu = P->basis_rows(P->rcl), bv = P->basis_rows(P->ccl);
    int64_t dense_live = 0, basis_live = 0;
    for (const DDense &D : P->dense) dense_live += tiled_size(D.m, D.n);
    for (const DCluster &C : P->rcl) basis_live += tiled_size(C.len, r);
    for (const DCluster &C : P->ccl) basis_live += tiled_size(C.len, r);
    if (P->flags & HM_PLAN_HYBRID) basis_live = 0;
    if (P->flags & HM_PLAN_MATRIX_FREE) dense_live = basis_live = 0;
    const int64_t eb = P->elem_bytes();
    const int64_t vec = eb * (P->nx + P->ny);
    v[HM_INFO_NROWS] = P->nx;
    v[HM_INFO_NCOLS] = P->ny;
    v[HM_INFO_RANK] = P->r;
    v[HM_INFO_BLOCKSIZE] = P->bs;
    v[HM_INFO_KERNEL] = P->kid;
    v[HM_INFO_N_D] = P->nd;
    v[HM_INFO_N_K1] = P->nk1;
    v[HM_INFO_N_K2] = P->nk2;
    int64_t ndense = 0, nfar = 0;
    for (const Leaf &L : P->leaves) (L.code == 3 ? ndense : nfar)++;
    v[HM_INFO_N_DENSE] = ndense;
    v[HM_INFO_N_FAR] = nfar;
    v[HM_INFO_N_ROWCLUSTERS] = (int64_t)P->rcl.size();
    v[HM_INFO_N_COLCLUSTERS] = (int64_t)P->ccl.size();
    v[HM_INFO_DEPTH] = P->maxdepth;
    v[HM_INFO_DENSE_ENTRIES] = P->dense_entries;
    v[HM_INFO_BASIS_ROWS_U] = bu;
    v[HM_INFO_BASIS_ROWS_V] = bv;
    const int64_t fbytes = (P->flags & HM_PLAN_MATRIX_FREE) ? 0 : (int64_t)P->far.size() * r * r;
    v[HM_INFO_BYTES_STORED] = eb * (dense_live + basis_live + fbytes) + vec;
    v[HM_INFO_BYTES_REF_LAYOUT] = eb * (P->dense_entries + r * (P->far_rows + P->far_cols) + nfar * r * r) + vec;
    v[HM_INFO_FLAGS] = P->flags;
    v[HM_INFO_N_ROWCHUNKS] = (int64_t)P->rows.ch.size();
    v[HM_INFO_N_COLCHUNKS] = (int64_t)P->cols.ch.size();
    v[HM_INFO_NONFINITE] = P->nonfinite;
    v[HM_INFO_N_INSTANCES] = (int64_t)P->row_off.size() - 1;
    v[HM_INFO_FROM_BLOCKS] = P->from_blocks ? 1 : 0;
    v[HM_INFO_FLOAT32] = P->f32 ? 1 : 0;
    v[HM_INFO_NESTED] = P->nest ? 1 : 0;
    for (int k = 0; k < n && k < HM_INFO_COUNT; ++k) info[k] = v[k];
    for (int k = HM_INFO_COUNT; k < n; ++k) info[k] = 0;
    return HM_OK;
}

int hm_plan_blocks(hm_plan_t P, int64_t *ints, double *dbls, int64_t cap, int64_t *count) {
    if (!P) return fail(HM_ERR_INVALID, "plan_blocks: NULL plan");
    if (count) *count = (int64_t)P->leaves.size();
    for (int64_t k = 0; k < cap && k < (int64_t)P->leaves.size(); ++k) {
        const Leaf &L = P->leaves[k];
        if (ints) {
            int64_t *o = ints + 6 * k;
            o[0] = L.code; o[1] = L.depth; o[2] = L.row0; o[3] = L.col0; o[4] = L.m; o[5] = L.n;
        }
        if (dbls) {
            double *g = dbls + 4 * k;
            g[0] = L.a; g[1] = L.b; g[2] = L.c; g[3] = L.d;
        }
    }
    return HM_OK;
}

// right-hand sides per launch of the MFMA passes, and the default number of
// right-hand sides from which forward products use them
constexpr int MFMA_KB = 16;      // the tuner's code for the MFMA passes
constexpr int MFMA_LAUNCH = 32;  // right-hand sides per MFMA launch (two groups of 16)
constexpr int64_t HM_MFMA_MIN_K = 8;

// A product entry point for element type T may only be called on a plan of
// that type.
extern "C++" {
template <class T>
static int check_type(hm_plan_t P, const char *what) {
    const bool want32 = std::is_same_v<T, float>;
    if (P->f32 != want32)
        return fail(HM_ERR_INVALID, "%s: the plan holds %s data; use the %s entry point", what,
                    P->f32 ? "Float32" : "Float64", P->f32 ? "_f32" : "Float64");
    return HM_OK;
}
}  // extern "C++"

extern "C++" {
// Hybrid and matrix-free products whose near field holds at least this many
// entries run it on a side stream beside the far-field chain; below it the
// fork and join cost more than the overlap returns.
constexpr int64_t HM_SPLIT_MIN_ENTRIES = 1LL << 24;

// Whether this product may split its near field onto the plan's side stream:
// a large near field, no pass events, and a launch stream that is not being
// captured (a captured product keeps one stream). Makes the side stream and
// its events on first use; any failure leaves the product on one stream.
static bool split_near_field(hm_plan_t P, hipStream_t s) {
    if (P->profile || P->dense_entries < HM_SPLIT_MIN_ENTRIES) return false;
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    if (hipStreamIsCapturing(s, &cs) != hipSuccess) {
        (void)hipGetLastError();
        return false;
    }
    if (cs != hipStreamCaptureStatusNone) return false;
    if (!P->side) {
        hipStream_t st = nullptr;
        hipEvent_t f = nullptr, j = nullptr;
        if (hipStreamCreateWithFlags(&st, hipStreamNonBlocking) != hipSuccess ||
            hipEventCreateWithFlags(&f, hipEventDisableTiming) != hipSuccess ||
            hipEventCreateWithFlags(&j, hipEventDisableTiming) != hipSuccess) {
            (void)hipGetLastError();
            if (f) (void)hipEventDestroy(f);
            if (st) (void)hipStreamDestroy(st);
            return false;
        }
        P->side = st;
        P->fork = f;
        P->join = j;
    }
    return true;
}

template <class T>
static MvArgsT<T> ws_args(hm_plan_t P) {
    MvArgsT<T> a{};
    PlanBufs<T> &b = P->bufs<T>();
    a.partial = b.partial;
    a.w = b.w;
    a.z = b.z;
    a.xt = b.xt;
    a.yt = b.yt;
    return a;
}
}  // extern "C++"

extern "C++" {
template <class T>
static int matvec_common(hm_plan_t P, T *y, const T *x, int64_t ioff, int64_t joff,
                         int64_t incy, int64_t incx, int64_t k, int64_t ldy, int64_t ldx, int beta,
                         unsigned flags, void *stream, bool tr) {
    if (!P) return fail(HM_ERR_INVALID, "matvec: NULL plan");
    if (P->flags & HM_PLAN_TREE_ONLY) return fail(HM_ERR_INVALID, "matvec: tree-only plan has no device data");
    if (int rc = check_type<T>(P, "matvec")) return rc;
    if (k < 0 || incx < 1 || incy < 1 || ioff < 0 || joff < 0 || (beta != 0 && beta != 1))
        return fail(HM_ERR_INVALID, "matvec: bad offsets, strides, k or beta");
    if (flags & ~255u) return fail(HM_ERR_INVALID, "matvec: unknown flags 0x%x", flags);
    if ((flags & HM_MV_MFMA) && (flags & HM_MV_NO_MFMA))
        return fail(HM_ERR_INVALID, "matvec: HM_MV_MFMA and HM_MV_NO_MFMA exclude each other");
    const int prec = (flags & HM_MV_FAR_BF16X6) ? 3 : (flags & HM_MV_FAR_BF16X3) ? 1
                   : (flags & HM_MV_FAR_F32) ? 0 : -1;
    if (__builtin_popcount(flags & (HM_MV_FAR_BF16X3 | HM_MV_FAR_F32 | HM_MV_FAR_BF16X6)) > 1)
        return fail(HM_ERR_INVALID, "matvec: choose one far-field precision");
    if (k == 0) return HM_OK;
    if (!y || !x) return fail(HM_ERR_INVALID, "matvec: NULL vector");
    if (k > 1 && (ldx < 1 || ldy < 1)) return fail(HM_ERR_INVALID, "matvec: bad ld");
    int mode = 0;
    if ((P->flags & HM_PLAN_MATRIX_FREE) || (flags & HM_MV_MATRIX_FREE)) mode = 2;
    else if ((P->flags & HM_PLAN_HYBRID) || (flags & HM_MV_CHEB_BASIS)) mode = 1;
    if (P->from_blocks && (mode != 0 || (flags & (HM_MV_MATRIX_FREE | HM_MV_CHEB_BASIS))))
        return fail(HM_ERR_UNSUPPORTED, "a plan built from blocks has no points: stored form only");
    if (P->kid == HM_USER && mode == 2)
        return fail(HM_ERR_UNSUPPORTED, "HM_USER plan: f is a host function, no matrix-free product");
    int max_kb = P->max_kb;
    if ((flags & HM_MV_TUNED) && P->tuned_mode >= 0) {
        if (!P->from_blocks) mode = std::max(mode, P->tuned_mode);
        max_kb = std::min(P->tuned_kb, HM_MAX_KB);
    }
    hipStream_t s = (hipStream_t)stream;
    // forward products of a Float64 KernelMatrix plan may run on the MFMA
    // passes (16 right-hand sides per launch): always for a reduced-precision
    // far field, by default from HM_MFMA_MIN_K right-hand sides on
    bool use_m = prec >= 0;
    if constexpr (std::is_same_v<T, double>) {
        const bool can_m = !tr && !P->from_blocks;
        if (prec >= 0 && !can_m)
            return fail(HM_ERR_UNSUPPORTED, "reduced-precision far field: forward KernelMatrix products only");
        if (prec < 0 && can_m && !(flags & HM_MV_NO_MFMA)) {
            if (flags & HM_MV_MFMA) use_m = true;
            else if ((flags & HM_MV_TUNED) && P->tuned_mode >= 0) use_m = P->tuned_kb == MFMA_KB;
            else use_m = k >= HM_MFMA_MIN_K;
        }
        if (use_m) {
            const int pm = prec >= 0 ? prec : 2;
            for (int64_t q = 0; q < k; q += MFMA_LAUNCH) {
                MvArgs a = ws_args<double>(P);
                a.y = y + ioff + q * ldy;
                a.x = x + joff + q * ldx;
                a.incy = incy;
                a.incx = incx;
                a.ldy = ldy;
                a.ldx = ldx;
                a.beta = beta;
                const int nq = (int)std::min<int64_t>(MFMA_LAUNCH, k - q);
                hipError_t e = hmx_launch_matvec_m(P->b64.dp, a, mode, pm, nq, s);
                if (e != hipSuccess) return fail(HM_ERR_HIP, "matvec launch: %s", hipGetErrorString(e));
            }
            return HM_OK;
        }
    } else {
        if (prec >= 0) return fail(HM_ERR_UNSUPPORTED, "reduced-precision far field: Float64 plans only");
        if (flags & HM_MV_MFMA) return fail(HM_ERR_UNSUPPORTED, "MFMA passes: Float64 plans only");
    }
    // a streamed near field (stored form) competes with the upward pass for
    // HBM and gains nothing on a side stream; the computed ones (hybrid's
    // Clenshaw far field, matrix-free kernel values) overlap the chain
    const bool use_side = mode != 0 && split_near_field(P, s);
    for (int64_t q = 0; q < k;) {
        // forward products take up to HM_MAX_KB right-hand sides per launch
        int kb = 1;
        if (!tr && max_kb > 1)
            while (kb * 2 <= max_kb && q + kb * 2 <= k) kb *= 2;
        MvArgsT<T> a = ws_args<T>(P);
        a.y = y + ioff + q * ldy;
        a.x = x + joff + q * ldx;
        a.incy = incy;
        a.incx = incx;
        a.ldy = ldy;
        a.ldx = ldx;
        a.beta = beta;
        if (use_side) {
            a.side = P->side;
            a.fork = P->fork;
            a.join = P->join;
        }
        const hipEvent_t *ev = nullptr;
        if (P->profile) {
            const size_t at = P->prof_events.size();
            for (int e5 = 0; e5 < 5; ++e5) {
                hipEvent_t h;
                HM_HIP(hipEventCreate(&h));
                P->prof_events.push_back(h);
            }
            ev = P->prof_events.data() + at;
        }
        hipError_t e = hmx_launch_matvec(P->bufs<T>().dp, a, tr, mode, kb, s, ev);
        if (e != hipSuccess) return fail(HM_ERR_HIP, "matvec launch: %s", hipGetErrorString(e));
        q += kb;
    }
    return HM_OK;
}
}  // extern "C++"

int hm_matvec(hm_plan_t P, double *y, const double *x, int64_t ioff, int64_t joff, int64_t incy,
              int64_t incx, int64_t k, int64_t ldy, int64_t ldx, int beta, unsigned flags,
              void *stream) {
    return matvec_common(P, y, x, ioff, joff, incy, incx, k, ldy, ldx, beta, flags, stream, false);
}

int hm_matvec_t(hm_plan_t P, double *y, const double *x, int64_t ioff, int64_t joff,
                int64_t incy, int64_t incx, int64_t k, int64_t ldy, int64_t ldx, int beta,
                unsigned flags, void *stream) {
    return matvec_common(P, y, x, ioff, joff, incy, incx, k, ldy, ldx, beta, flags, stream, true);
}

int hm_matvec_f32(hm_plan_t P, float *y, const float *x, int64_t ioff, int64_t joff,
                  int64_t incy, int64_t incx, int64_t k, int64_t ldy, int64_t ldx, int beta,
                  unsigned flags, void *stream) {
    return matvec_common(P, y, x, ioff, joff, incy, incx, k, ldy, ldx, beta, flags, stream, false);
}

int hm_matvec_t_f32(hm_plan_t P, float *y, const float *x, int64_t ioff, int64_t joff,
                    int64_t incy, int64_t incx, int64_t k, int64_t ldy, int64_t ldx, int beta,
                    unsigned flags, void *stream) {
    return matvec_common(P, y, x, ioff, joff, incy, incx, k, ldy, ldx, beta, flags, stream, true);
}

extern "C++" {
template <class T>
static int grad_x_common(hm_plan_t P, T *g, const T *b, const T *s, int64_t ioff, int64_t joff,
                         int64_t incg, int64_t incb, int64_t incs, int beta, void *stream) {
    if (!P || !g || !b || !s) return fail(HM_ERR_INVALID, "grad_x: NULL argument");
    if (P->flags & HM_PLAN_TREE_ONLY) return fail(HM_ERR_INVALID, "grad_x: tree-only plan");
    if (P->from_blocks) return fail(HM_ERR_UNSUPPORTED, "grad_x: a plan built from blocks has no points");
    if (P->kid == HM_USER) return fail(HM_ERR_UNSUPPORTED, "grad_x: f is a host function (no derivative on the device)");
    if (int rc = check_type<T>(P, "grad_x")) return rc;
    if (incg < 1 || incb < 1 || incs < 1 || ioff < 0 || joff < 0 || (beta != 0 && beta != 1))
        return fail(HM_ERR_INVALID, "grad_x: bad offsets, strides or beta");
    MvArgsT<T> a = ws_args<T>(P);
    a.y = g + ioff;
    a.x = b + joff;
    a.incy = incg;
    a.incx = incb;
    a.beta = beta;
    hipError_t e = hmx_launch_grad(P->bufs<T>().dp, a, false, (T *)nullptr, 0, s, incs, (hipStream_t)stream);
    if (e != hipSuccess) return fail(HM_ERR_HIP, "grad_x launch: %s", hipGetErrorString(e));
    return HM_OK;
}
}  // extern "C++"

extern "C++" {
template <class T>
static int t_grad_common(hm_plan_t P, T *gb, T *gy, const T *r, const T *s, int64_t incgb,
                         int64_t incgy, int64_t incr, int64_t incs, int beta, void *stream) {
    if (!P || !gb || !gy || !r || !s) return fail(HM_ERR_INVALID, "t_grad: NULL argument");
    if (P->flags & HM_PLAN_TREE_ONLY) return fail(HM_ERR_INVALID, "t_grad: tree-only plan");
    if (P->from_blocks) return fail(HM_ERR_UNSUPPORTED, "t_grad: a plan built from blocks has no points");
    if (P->kid == HM_USER) return fail(HM_ERR_UNSUPPORTED, "t_grad: f is a host function (no derivative on the device)");
    if (int rc = check_type<T>(P, "t_grad")) return rc;
    if (incgb < 1 || incgy < 1 || incr < 1 || incs < 1 || (beta != 0 && beta != 1))
        return fail(HM_ERR_INVALID, "t_grad: bad strides or beta");
    MvArgsT<T> a = ws_args<T>(P);
    a.y = gb;
    a.x = r;
    a.incy = incgb;
    a.incx = incr;
    a.beta = beta;
    hipError_t e = hmx_launch_grad(P->bufs<T>().dp, a, true, gy, incgy, s, incs, (hipStream_t)stream);
    if (e != hipSuccess) return fail(HM_ERR_HIP, "t_grad launch: %s", hipGetErrorString(e));
    return HM_OK;
}
}  // extern "C++"

extern "C++" {
template <class T>
static int mv_grad_common(hm_plan_t P, T *y, T *g, const T *b, const T *s, int64_t incy,
                          int64_t incg, int64_t incb, int64_t incs, int beta, void *stream) {
    if (!P || !y || !g || !b) return fail(HM_ERR_INVALID, "matvec_grad: NULL argument");
    if (P->flags & HM_PLAN_TREE_ONLY) return fail(HM_ERR_INVALID, "matvec_grad: tree-only plan");
    if (P->from_blocks) return fail(HM_ERR_UNSUPPORTED, "matvec_grad: a plan built from blocks has no points");
    if (P->kid == HM_USER) return fail(HM_ERR_UNSUPPORTED, "matvec_grad: f is a host function (no derivative on the device)");
    if (int rc = check_type<T>(P, "matvec_grad")) return rc;
    if (incy < 1 || incg < 1 || incb < 1 || (s && incs < 1) || (beta != 0 && beta != 1))
        return fail(HM_ERR_INVALID, "matvec_grad: bad strides or beta");
    MvArgsT<T> a = ws_args<T>(P);
    a.y = y;
    a.x = b;
    a.incy = incy;
    a.incx = incb;
    a.beta = beta;
    hipError_t e = hmx_launch_grad(P->bufs<T>().dp, a, false, g, incg, s, s ? incs : 0, (hipStream_t)stream);
    if (e != hipSuccess) return fail(HM_ERR_HIP, "matvec_grad launch: %s", hipGetErrorString(e));
    return HM_OK;
}
}  // extern "C++"

int hm_matvec_grad_x(hm_plan_t P, double *g, const double *b, const double *s, int64_t ioff,
                     int64_t joff, int64_t incg, int64_t incb, int64_t incs, int beta, void *stream) {
    return grad_x_common(P, g, b, s, ioff, joff, incg, incb, incs, beta, stream);
}

int hm_matvec_grad(hm_plan_t P, double *y, double *g, const double *b, const double *s,
                   int64_t incy, int64_t incg, int64_t incb, int64_t incs, int beta, void *stream) {
    return mv_grad_common(P, y, g, b, s, incy, incg, incb, incs, beta, stream);
}

int hm_matvec_grad_f32(hm_plan_t P, float *y, float *g, const float *b, const float *s,
                       int64_t incy, int64_t incg, int64_t incb, int64_t incs, int beta,
                       void *stream) {
    return mv_grad_common(P, y, g, b, s, incy, incg, incb, incs, beta, stream);
}

int hm_matvec_t_grad(hm_plan_t P, double *gb, double *gy, const double *r, const double *s,
                     int64_t incgb, int64_t incgy, int64_t incr, int64_t incs, int beta,
                     void *stream) {
    return t_grad_common(P, gb, gy, r, s, incgb, incgy, incr, incs, beta, stream);
}

int hm_matvec_grad_x_f32(hm_plan_t P, float *g, const float *b, const float *s, int64_t ioff,
                         int64_t joff, int64_t incg, int64_t incb, int64_t incs, int beta,
                         void *stream) {
    return grad_x_common(P, g, b, s, ioff, joff, incg, incb, incs, beta, stream);
}

int hm_matvec_t_grad_f32(hm_plan_t P, float *gb, float *gy, const float *r, const float *s,
                         int64_t incgb, int64_t incgy, int64_t incr, int64_t incs, int beta,
                         void *stream) {
    return t_grad_common(P, gb, gy, r, s, incgb, incgy, incr, incs, beta, stream);
}

struct hm_graph {
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
};

extern "C++" {
template <class T>
static int graph_common(hm_plan_t P, T *y, const T *x, int64_t ioff, int64_t joff, int64_t incy,
                        int64_t incx, int64_t k, int64_t ldy, int64_t ldx, int beta,
                        unsigned flags, hm_graph_t *out) {
    if (!out) return fail(HM_ERR_INVALID, "graph: NULL output");
    *out = nullptr;
    if (!P || (P->flags & HM_PLAN_TREE_ONLY)) return fail(HM_ERR_INVALID, "graph: bad plan");
    if (P->profile) return fail(HM_ERR_INVALID, "graph: profiling is on (it records events)");
    hipStream_t s;
    HM_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    hipError_t e = hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal);
    if (e != hipSuccess) {
        (void)hipStreamDestroy(s);
        return fail(HM_ERR_HIP, "graph capture: %s", hipGetErrorString(e));
    }
    const int rc = matvec_common(P, y, x, ioff, joff, incy, incx, k, ldy, ldx, beta, flags, s, false);
    hipGraph_t g = nullptr;
    e = hipStreamEndCapture(s, &g);
    (void)hipStreamDestroy(s);
    if (rc != HM_OK) {
        if (g) (void)hipGraphDestroy(g);
        return rc;
    }
    if (e != hipSuccess) return fail(HM_ERR_HIP, "graph capture: %s", hipGetErrorString(e));
    hm_graph *G = new hm_graph();
    G->graph = g;
    e = hipGraphInstantiate(&G->exec, g, nullptr, nullptr, 0);
    if (e != hipSuccess) {
        (void)hipGraphDestroy(g);
        delete G;
        return fail(HM_ERR_HIP, "graph instantiate: %s", hipGetErrorString(e));
    }
    *out = G;
    return HM_OK;
}
}  // extern "C++"

int hm_graph_matvec(hm_plan_t P, double *y, const double *x, int64_t ioff, int64_t joff,
                    int64_t incy, int64_t incx, int64_t k, int64_t ldy, int64_t ldx, int beta,
                    unsigned flags, hm_graph_t *out) {
    return graph_common(P, y, x, ioff, joff, incy, incx, k, ldy, ldx, beta, flags, out);
}

int hm_graph_matvec_f32(hm_plan_t P, float *y, const float *x, int64_t ioff, int64_t joff,
                        int64_t incy, int64_t incx, int64_t k, int64_t ldy, int64_t ldx, int beta,
                        unsigned flags, hm_graph_t *out) {
    return graph_common(P, y, x, ioff, joff, incy, incx, k, ldy, ldx, beta, flags, out);
}

int hm_graph_launch(hm_graph_t G, void *stream) {
    if (!G) return fail(HM_ERR_INVALID, "graph_launch: NULL graph");
    HM_HIP(hipGraphLaunch(G->exec, (hipStream_t)stream));
    return HM_OK;
}

int hm_graph_destroy(hm_graph_t G) {
    if (!G) return HM_OK;
    if (G->exec) (void)hipGraphExecDestroy(G->exec);
    if (G->graph) (void)hipGraphDestroy(G->graph);
    delete G;
    return HM_OK;
}

extern "C++" {
template <class T>
static int tune_common(hm_plan_t P, T *y, const T *x, int64_t k, void *stream, int64_t out[3]) {
    hipStream_t s = (hipStream_t)stream;
    struct Events {  // destroyed on every exit path
        hipEvent_t e0 = nullptr, e1 = nullptr;
        ~Events() {
            if (e0) (void)hipEventDestroy(e0);
            if (e1) (void)hipEventDestroy(e1);
        }
    } ev;
    HM_HIP(hipEventCreate(&ev.e0));
    HM_HIP(hipEventCreate(&ev.e1));
    const int base_mode = (P->flags & HM_PLAN_MATRIX_FREE) ? 2 : (P->flags & HM_PLAN_HYBRID) ? 1 : 0;
    // a plan built from blocks has no points: its stored form is the only one
    // nor is f on the device for HM_USER plans: no matrix-free candidate
    const int last_mode = P->from_blocks ? 0 : P->kid == HM_USER ? 1 : 2;
    float best = 1e30f;
    int best_kb = 1, best_mode = base_mode;
    const int saved_kb = P->max_kb;
    struct Restore {  // the plan's launch shape is restored on every exit path
        hm_plan_t P;
        int kb;
        ~Restore() { P->max_kb = kb; }
    } restore{P, saved_kb};
    // candidates: 1, 2, 4, 8 right-hand sides per launch on the vector units,
    // and 16 per launch on the MFMA passes (Float64 KernelMatrix plans)
    const bool mfma_ok = std::is_same_v<T, double> && !P->from_blocks;
    for (int mode = base_mode; mode <= last_mode; ++mode) {
        const unsigned fl0 = mode == 2 ? HM_MV_MATRIX_FREE : mode == 1 ? HM_MV_CHEB_BASIS : 0;
        for (int kb = 1; kb <= MFMA_KB; kb *= 2) {
            if (kb == MFMA_KB) {
                if (!mfma_ok || mode == 0 || k < 2) break;  // MFMA passes: Chebyshev form
            } else if (kb > 1 && kb > k) {
                continue;
            }
            const unsigned fl = fl0 | (kb == MFMA_KB ? HM_MV_MFMA : (mfma_ok ? HM_MV_NO_MFMA : 0));
            P->max_kb = std::min(kb, HM_MAX_KB);
            int rc = matvec_common(P, y, x, 0, 0, 1, 1, k, P->nx, P->ny, 0, fl, s, false);
            if (rc != HM_OK) return rc;
            float ms = 0.f;
            HM_HIP(hipEventRecord(ev.e0, s));
            for (int rep = 0; rep < 3; ++rep) {
                rc = matvec_common(P, y, x, 0, 0, 1, 1, k, P->nx, P->ny, 0, fl, s, false);
                if (rc != HM_OK) return rc;
            }
            HM_HIP(hipEventRecord(ev.e1, s));
            HM_HIP(hipEventSynchronize(ev.e1));
            HM_HIP(hipEventElapsedTime(&ms, ev.e0, ev.e1));
            if (ms < best) { best = ms; best_kb = kb; best_mode = mode; }
        }
    }
    P->tuned_kb = best_kb;
    P->tuned_mode = best_mode;
    if (out) {
        out[0] = best_kb;
        out[1] = best_mode;
        out[2] = (int64_t)(1000.0f * best / 3.0f);
    }
    return HM_OK;
}
}  // extern "C++"

int hm_plan_tune(hm_plan_t P, void *y, const void *x, int64_t k, void *stream, int64_t out[3]) {
    if (!P || !y || !x || k < 1 || (P->flags & HM_PLAN_TREE_ONLY))
        return fail(HM_ERR_INVALID, "tune: bad arguments");
    if (P->f32) return tune_common(P, (float *)y, (const float *)x, k, stream, out);
    return tune_common(P, (double *)y, (const double *)x, k, stream, out);
}

int hm_plan_profile(hm_plan_t P, int enable) {
    if (!P) return fail(HM_ERR_INVALID, "profile: NULL plan");
    if (P->flags & HM_PLAN_TREE_ONLY) return fail(HM_ERR_INVALID, "profile: tree-only plan");
    HM_HIP(hipDeviceSynchronize());
    P->drop_events();
    P->profile = enable != 0;
    return HM_OK;
}

int hm_plan_profile_read(hm_plan_t P, double ms[4], int64_t *count) {
    if (!P || !ms) return fail(HM_ERR_INVALID, "profile_read: bad arguments");
    for (int q = 0; q < 4; ++q) ms[q] = 0.0;
    const size_t nprod = P->prof_events.size() / 5;
    if (nprod) HM_HIP(hipEventSynchronize(P->prof_events.back()));
    for (size_t t = 0; t < nprod; ++t) {
        for (int q = 0; q < 4; ++q) {
            float f = 0.f;
            HM_HIP(hipEventElapsedTime(&f, P->prof_events[5 * t + q], P->prof_events[5 * t + q + 1]));
            ms[q] += f;
        }
    }
    if (count) *count = (int64_t)nprod;
    return HM_OK;
}

extern "C++" {
template <class T>
static int entries_common(hm_plan_t P, int64_t n, const int64_t *i, const int64_t *j, T *out,
                          void *stream) {
    if (!P || n < 0) return fail(HM_ERR_INVALID, "entries: bad arguments");
    if (int rc = check_type<T>(P, "entries")) return rc;
    if (P->flags & HM_PLAN_TREE_ONLY) return fail(HM_ERR_INVALID, "entries: tree-only plan has no device data");
    if (n == 0) return HM_OK;
    if (!i || !j || !out) return fail(HM_ERR_INVALID, "entries: NULL pointer");
    hipStream_t s = (hipStream_t)stream;
    std::vector<int64_t> hi(n), hj(n);
    HM_HIP(hipMemcpyAsync(hi.data(), i, n * sizeof(int64_t), hipMemcpyDeviceToHost, s));
    HM_HIP(hipMemcpyAsync(hj.data(), j, n * sizeof(int64_t), hipMemcpyDeviceToHost, s));
    HM_HIP(hipStreamSynchronize(s));
    std::vector<int32_t> kind(n, 0), id(n, -1);
    const Side &R = P->rows;
    for (int64_t t = 0; t < n; ++t) {
        const int64_t ii = hi[t], jj = hj[t];
        if (ii < 0 || ii >= P->nx || jj < 0 || jj >= P->ny)
            return fail(HM_ERR_INVALID, "entries: index (%lld, %lld) out of range", (long long)ii, (long long)jj);
        auto it = std::upper_bound(R.ch.begin(), R.ch.end(), ii,
                                   [](int64_t v, const DChunk &c) { return v < c.p0; });
        const DChunk &c = *(it - 1);
        for (int e = c.near_beg; e < c.near_end && !kind[t]; ++e) {
            const DDense &D = P->dense[R.near[e].blk];
            if (jj >= D.col0 && jj < D.col0 + D.n) { kind[t] = 1; id[t] = R.near[e].blk; }
        }
        for (int e = c.far_beg; e < c.far_end && !kind[t]; ++e) {
            const DCluster &C = P->rcl[R.farl[e].cl];
            for (int32_t q = C.far_beg; q < C.far_end && !kind[t]; ++q) {
                const int32_t f = P->far_by_rc_h[q];
                const DCluster &Cc = P->ccl[P->far[f].cc];
                if (jj >= Cc.first && jj < Cc.first + Cc.len) { kind[t] = 2; id[t] = f; }
            }
        }
    }
    int32_t *d_kind = nullptr, *d_id = nullptr;
    HM_HIP(hipMalloc((void **)&d_kind, n * sizeof(int32_t)));
    hipError_t e = hipMalloc((void **)&d_id, n * sizeof(int32_t));
    if (e == hipSuccess) e = hipMemcpyAsync(d_kind, kind.data(), n * sizeof(int32_t), hipMemcpyHostToDevice, s);
    if (e == hipSuccess) e = hipMemcpyAsync(d_id, id.data(), n * sizeof(int32_t), hipMemcpyHostToDevice, s);
    if (e == hipSuccess) e = hmx_launch_entries(P->bufs<T>().dp, n, i, j, d_kind, d_id, out, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    (void)hipFree(d_kind);
    if (d_id) (void)hipFree(d_id);
    if (e != hipSuccess) return fail(HM_ERR_HIP, "entries: %s", hipGetErrorString(e));
    return HM_OK;
}
}  // extern "C++"

int hm_entries(hm_plan_t P, int64_t n, const int64_t *i, const int64_t *j, double *out,
               void *stream) {
    return entries_common(P, n, i, j, out, stream);
}

int hm_entries_f32(hm_plan_t P, int64_t n, const int64_t *i, const int64_t *j, float *out,
                   void *stream) {
    return entries_common(P, n, i, j, out, stream);
}

}  // extern "C"
