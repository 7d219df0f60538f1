// hmx_device.h -- host/device interface between the plan builder and the
// gfx950 kernels.
#pragma once
#include <hip/hip_runtime.h>
#include "hmx_internal.h"

template <class T> struct Real2;
template <> struct Real2<double> { using type = double2; };
template <> struct Real2<float> { using type = float2; };

// Device-side view of a plan (all pointers are device pointers). T is the
// element type of the plan: double (Float64) or float (Float32).
template <class T>
struct DevPlanT {
    using V2 = typename Real2<T>::type;
    int kid;       // enum hm_kernel
    int r;         // Chebyshev rank
    int64_t nx, ny;
    const T *x, *y;          // points (non-increasing)
    const T *lam;            // r barycentric weights
    const DCluster *rcl;     // row clusters
    const DCluster *ccl;     // col clusters
    int32_t nrc, ncc;
    const T *rnodes;         // nrc*r mapped nodes (x side)
    const T *cnodes;         // ncc*r mapped nodes (y side)
    const V2 *rmap;          // per row cluster: (center, 1/halfwidth) of its interval
    const V2 *cmap;          // per col cluster
    const T *chebA;          // r x r, A[k][m] = (m ? 2 : 1) T_m(xhat_k) / r (row-major)
    const DFarBlk *far;
    int32_t nfar;
    const int32_t *far_by_rc;  // far ids grouped by row cluster (CSR in DCluster.far_*)
    const int32_t *far_by_cc;  // far ids grouped by col cluster
    // the translation records in far_by_rc / far_by_cc order, so a pass
    // walking a cluster's list reads each record with one load
    const DTrRec *tr_rc, *tr_cc;
    const DDense *dense;
    int32_t ndense;
    const DChunk *rch, *cch;   // row chunks, col chunks
    int32_t nrch, ncch;
    const DNear *rnear, *cnear;
    const DFar *rfar, *cfar;
    const DTask *rtask, *ctask;  // upward-pass tasks over row / col clusters
    int32_t nrtask, nctask;
    int64_t nslots_r, nslots_c;
    T *A, *U, *V, *F;          // stored factors (null when matrix-free)
    // the row / column clusters with more than one upward task (the others'
    // tasks write w themselves)
    const int32_t *multi_r, *multi_c;
    int32_t nmulti_r, nmulti_c;
    // nested far field of the MFMA passes (Float64 plans from points; nest
    // is 0 when the plan's tree does not allow it): the interval trees of
    // both sides, their internal node ids by depth, one upward task per
    // non-empty column leaf (task.cl = the leaf node, task.slot = its column
    // cluster or -1), each row chunk's row leaf, the two re-expansion matrices
    // T_k(s/2 -+ 1/2) = sum_m M[h][k][m] T_m(s), and the moments per column
    // node and coefficients per internal row node (r x 32 each; a row leaf's
    // are formed by the downward pass). The level offsets are host arrays,
    // read by the launching code.
    int nest;
    const DTNode *xtn, *ytn;
    const int32_t *xlev, *ylev;
    const int32_t *xlev_off_h, *ylev_off_h;  // host: CSR offsets over depth
    int32_t nxlev, nylev;                    // depths in each
    const int32_t *chunk_leaf;
    const DTask *yleaf;
    int32_t nyleaf;
    const T *nestM;
    T *my, *mx;
};

template <class T>
struct MvArgsT {
    T *y;
    const T *x;
    int64_t incy, incx;
    int64_t ldy, ldx;  // stride between right-hand sides
    int beta;
    T *partial;  // >= max(nslots) * r * kb
    T *w;        // >= max(nrc, ncc) * r * kb
    T *z;        // >= max(nrc, ncc) * r * kb
    T *xt;       // >= ny * 16: right-hand sides packed row by row (MFMA passes)
    T *yt;       // >= max(nx, ny): a split product's far field (one RHS, matrix-free or Float32)
    // downward pass: 0 near and far field; 1 near field only (y = beta*y +
    // near); 2 far field only (y += far); 3 far field only (y = far)
    int part;
    // when set and no pass events are requested, the near field runs on `side`
    // (forked from the launch stream through `fork`) beside the upward,
    // reduce and translate passes; after `join` the far field is added onto
    // y (one right-hand side of a matrix-free or Float32 plan: written into
    // yt beside the near field first)
    hipStream_t side;
    hipEvent_t fork, join;
};

using DevPlan = DevPlanT<double>;
using DevPlanF = DevPlanT<float>;
using MvArgs = MvArgsT<double>;
using MvArgsF = MvArgsT<float>;

// compression: dense blocks, U, V, F (bit-exact operation order)
hipError_t hmx_launch_compress(const DevPlan &p, hipStream_t s);
// one product y = beta*y + op(H) x for a single right-hand side; mode 0
// stored, 1 hybrid (stored A, F; Chebyshev-form basis), 2 matrix-free. When
// ev is not null, ev[0..4] are recorded before up, reduce, translate, down and
// after down.
// kb in {1, 2, 4, 8}: right-hand sides handled by this launch (forward only).
hipError_t hmx_launch_matvec(const DevPlan &p, const MvArgs &a, bool transpose, int mode, int kb,
                             hipStream_t s, const hipEvent_t *ev = nullptr);
// forward product of up to 32 right-hand sides (nq; two groups of 16 when
// nq > 16) on the MFMA units, far
// field in Chebyshev form: prec 2 FP64, 0 FP32, 1 split BF16 for the far-field
// products (near field, reduce and translate FP64); mode 0/1 use the stored
// dense blocks, 2 regenerates them.
hipError_t hmx_launch_matvec_m(const DevPlan &p, const MvArgs &a, int mode, int prec, int nq,
                               hipStream_t s);
// gradients (matrix-free near field, Chebyshev-form far field):
//   forward (tr = false): y = beta*y + s .* (d/dx_i of (H x)_i)
//   transpose (tr = true): y = beta*y + H^T x, y2 = beta*y2 + s .* (d/dy_j of (H^T x)_j)
hipError_t hmx_launch_grad(const DevPlan &p, const MvArgs &a, bool transpose, double *y2,
                           int64_t incy2, const double *s, int64_t incs, hipStream_t stream);
// entries: kind[t] = 0 none, 1 dense (id), 2 far (id); rows/cols 0-based
hipError_t hmx_launch_entries(const DevPlan &p, int64_t n, const int64_t *i, const int64_t *j,
                              const int32_t *kind, const int32_t *id, double *out, hipStream_t s);
// number of non-finite values among generated factors (after compression)
hipError_t hmx_launch_count_nonfinite(const double *v, int64_t n, unsigned long long *count,
                                      hipStream_t s);

// in-place diagonal scaling of stored dense blocks and bases (plans built
// from blocks): side 0 scales columns (A and V rows), 1 rows (A and U rows);
// b is indexed by the global column (side 0) or row (side 1)
hipError_t hmx_launch_scale(const DevPlan &p, int side, const double *b, hipStream_t s);

// hm_plan_update_points: *status = 0 when the points x (nx) and y (ny) are
// new points x, y copied into the plan's arrays dx, dy and checked on the way:
// status 0 if finite, non-increasing and split as the tree's recorded splits
// do, else 1 (the status word is cleared by a kernel first, so the launch is
// capturable); refit re-forms the dense blocks and the stored bases from the
// plan's points
template <class T>
hipError_t hmx_launch_update_points(const T *x, T *dx, int64_t nx, const DSplit *xs, int64_t nxs,
                                    const T *y, T *dy, int64_t ny, const DSplit *ys, int64_t nys,
                                    int32_t *status, hipStream_t s);
hipError_t hmx_launch_refit(const DevPlanT<double> &p, hipStream_t s);
hipError_t hmx_launch_refit(const DevPlanT<float> &p, hipStream_t s);

// add_col! on a plan built from blocks: per job (one leaf holding the column),
// a dense leaf adds u[row0 + i] to its column jl; a low-rank leaf writes
// u[row0 + i] into column c of its U, 1 into row jl of column c of its V and
// 1 into F[c, c] (u is offset to the matrix's first row).
struct AddColJob {
    int32_t kind, m, n, c, jl, far;
    int64_t row0, aoff, uoff, voff, foff;
};
#define HM_ADD_COL_JOBS 24  // per launch (the jobs go by value, about 1.3 KB)
struct AddColJobs {
    AddColJob job[HM_ADD_COL_JOBS];
    int32_t count;
};
hipError_t hmx_launch_add_col(const DevPlan &p, const AddColJobs &J, const double *u, hipStream_t s);

// The same passes for FP32 plans (ranks up to 32; no reduced-precision form).
hipError_t hmx_launch_compress(const DevPlanF &p, hipStream_t s);
hipError_t hmx_launch_matvec(const DevPlanF &p, const MvArgsF &a, bool transpose, int mode, int kb,
                             hipStream_t s, const hipEvent_t *ev = nullptr);
hipError_t hmx_launch_grad(const DevPlanF &p, const MvArgsF &a, bool transpose, float *y2,
                           int64_t incy2, const float *s, int64_t incs, hipStream_t stream);
hipError_t hmx_launch_entries(const DevPlanF &p, int64_t n, const int64_t *i, const int64_t *j,
                              const int32_t *kind, const int32_t *id, float *out, hipStream_t s);
hipError_t hmx_launch_count_nonfinite(const float *v, int64_t n, unsigned long long *count,
                                      hipStream_t s);
