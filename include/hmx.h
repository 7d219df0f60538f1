/*
 * hmx.h -- C ABI of libhmx, the MI355X (gfx950) implementation of the
 * KernelMatrix compression and hierarchical matrix-vector product of
 * HierarchicalMatrices.jl.
 *
 * Every entry point takes plain pointers and sizes, returns an int status
 * (HM_OK == 0) and never aborts; hm_last_error() gives a thread-local message
 * for the last failing call. Vector arguments of the products are DEVICE
 * pointers; point arrays of the constructors are HOST pointers (the block tree
 * is built on the host from them, as the reference builds it). `stream` is a
 * hipStream_t (NULL = the null stream); all work is stream-ordered.
 *
 * Index conventions: offsets are 0-based (the reference's 1-based istart/jstart
 * minus one); strides inc >= 1 step between consecutive entries of one vector;
 * ld steps between right-hand sides.
 */
#ifndef HMX_H
#define HMX_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct hm_plan *hm_plan_t;

enum hm_status {
    HM_OK = 0,
    HM_ERR_INVALID = 1,        /* bad argument */
    HM_ERR_NOT_DESCENDING = 2, /* points not non-increasing: the reference's
                                  indsplit (src/BarycentricMatrix.jl:299-307)
                                  recurses without bound on such input */
    HM_ERR_DEGENERATE = 3,     /* an empty cluster reached indsplit; the
                                  reference then forms an out-of-range block */
    HM_ERR_OOM = 4,
    HM_ERR_HIP = 5,
    HM_ERR_UNSUPPORTED = 6
};

/* kernels of examples/Kernel.jl:34-37, and HM_USER: an arbitrary host
 * function f (the reference's KernelMatrix(f, ...) takes any Julia f,
 * src/KernelMatrix.jl:47). A plan of HM_USER builds the tree and the bases U, V
 * (which do not depend on f) on the device; the caller evaluates f on the
 * dense leaves and at the far blocks' node pairs and hands the values over
 * with hm_plan_set_leaf_data. Such plans have no matrix-free form and no
 * gradients (HM_ERR_UNSUPPORTED). */
enum hm_kernel { HM_CAUCHY = 0, HM_COULOMB = 1, HM_COULOMB3 = 2, HM_LOG = 3, HM_USER = 4 };

/* plan flags (hm_plan_from_points) */
enum hm_plan_flags {
    HM_PLAN_STORED = 0,          /* generate and keep dense blocks, U, V, F in HBM */
    HM_PLAN_MATRIX_FREE = 1,     /* keep only points and nodes; regenerate in every product */
    HM_PLAN_FAR_F32 = 2,         /* stored far-field factors in FP32 (mixed precision) */
    HM_PLAN_TREE_ONLY = 4,       /* build the block tree and its tables on the host only:
                                    no device memory, no products (for inspection) */
    HM_PLAN_HYBRID = 8           /* keep dense blocks and F in HBM; apply the far-field
                                    basis in Chebyshev form at product time (no U, V) */
};

/* product flags (hm_matvec) */
enum hm_mv_flags {
    HM_MV_DEFAULT = 0,
    HM_MV_MATRIX_FREE = 1,       /* regenerate factors even if the plan stores them */
    HM_MV_CHEB_BASIS = 2,        /* apply the far-field basis in Chebyshev form even if
                                    the plan stores U and V */
    HM_MV_TUNED = 4,             /* use the launch shape and mode chosen by hm_plan_tune */
    HM_MV_FAR_F32 = 8,           /* far field (Chebyshev form) on FP32 MFMA, FP32 accumulate
                                    per step (64 points upward, one cluster downward) and
                                    FP64 across steps; points, differences, near field,
                                    reduce, F and translation stay FP64 */
    HM_MV_FAR_BF16X3 = 16,       /* far field on BF16 MFMA with operands split hi + lo
                                    (hi*hi + hi*lo + lo*hi), FP32 accumulate: operands
                                    carry 16 bits */
    HM_MV_MFMA = 32,             /* forward: the MFMA passes, up to 32 right-hand sides
                                    per launch in two groups of 16 that share every
                                    generated or streamed operand (far field in
                                    Chebyshev form); the default when k >= 8 */
    HM_MV_NO_MFMA = 64,          /* forward: never the MFMA passes (up to 8 right-hand
                                    sides per launch on the vector units) */
    HM_MV_FAR_BF16X6 = 128       /* far field on BF16 MFMA with operands split in three,
                                    hi + mid + lo (six products, down to mid*mid), FP32
                                    accumulate: operands carry FP32's 24 bits */
};

/* Message of the last failing call on this thread ("" if none). */
const char *hm_last_error(void);

/* Library version, major*10000 + minor*100 + patch. */
int hm_version(void);

/* BLOCKRANK for an IEEE type with `mant_bits` stored mantissa bits.
 * Replaces BLOCKRANK(T), src/HierarchicalMatrices.jl:5. BLOCKSIZE(T) = 4*rank
 * (src/HierarchicalMatrices.jl:7). */
int hm_blockrank(int mant_bits);

/* Replaces chebyshevpoints(T, n; kind), src/BarycentricMatrix.jl:92-111. */
int hm_chebyshevpoints(double *out, int n, int kind);

/* Replaces chebyshevbarycentricweights(T, n; kind), src/BarycentricMatrix.jl:114-136. */
int hm_chebyshevbarycentricweights(double *out, int n, int kind);
/* The same for T = Float32 (the quotient formed in Float32, sinpi rounded
 * once to Float32). */
int hm_chebyshevpoints_f32(float *out, int n, int kind);
int hm_chebyshevbarycentricweights_f32(float *out, int n, int kind);

/* Replaces indsplit(x, ir, a, b), src/BarycentricMatrix.jl:299-307.
 * ir = first:first+len-1 (1-based first, as in Julia); out = {first1, len1,
 * first2, len2}. Returns HM_ERR_NOT_DESCENDING if x[ir] is not non-increasing
 * and HM_ERR_DEGENERATE where the reference would read out of range. */
int hm_indsplit(const double *x, int64_t nx, int64_t first, int64_t len, double a, double b,
                int64_t out[4]);

/* Replaces KernelMatrix(f, x, y, a, b, c, d), src/KernelMatrix.jl:47-116, with
 * f one of enum hm_kernel, and BLOCKSIZE/BLOCKRANK given explicitly
 * (blocksize = leaf threshold, rank = Chebyshev rank). Builds the block tree on
 * the host with the reference's rule, lays it out level by level in HBM and
 * (unless HM_PLAN_MATRIX_FREE) runs the compression kernels on `stream`:
 * dense blocks (src/KernelMatrix.jl:57-60,82,104), F (src/BarycentricMatrix.jl:168-175)
 * and U, V (update!, src/BarycentricMatrix.jl:248-297), one U per row cluster
 * and one V per column cluster. x (length nx) and y (length ny) are host arrays
 * in non-increasing order; (a, b) and (c, d) are their intervals given as
 * (upper, lower), as in examples/Kernel.jl:66. */
int hm_plan_from_points(int kernel_id, const double *x, int64_t nx, const double *y, int64_t ny,
                        double a, double b, double c, double d, int blocksize, int rank,
                        unsigned flags, void *stream, hm_plan_t *plan);

/* The dense leaves and far-field cores of a plan built with kernel HM_USER
 * (stored or hybrid form), in the order of hm_plan_blocks (the reference's
 * traversal, src/KernelMatrix.jl:17-45), leaves with an empty side skipped:
 * A holds every dense leaf's f(x[i], y[j]) (m x n column-major, one after the
 * other, nA values in all), F every far leaf's F[m, n] = f(t_m, s_n) at its
 * mapped Chebyshev nodes (src/BarycentricMatrix.jl:147-178; r x r column-major,
 * nF = far leaves x r x r). Host arrays; synchronizes `stream`. */
int hm_plan_set_leaf_data(hm_plan_t plan, const double *A, int64_t nA, const double *F,
                          int64_t nF, void *stream);
int hm_plan_set_leaf_data_f32(hm_plan_t plan, const float *A, int64_t nA, const float *F,
                              int64_t nF, void *stream);

/* A HierarchicalMatrix given by its leaves (src/HierarchicalMatrix.jl:1-52 with
 * LowRankMatrix and Matrix leaves, src/LowRankMatrix.jl:28-38), or the leaves
 * of a KernelMatrix built by the reference from an arbitrary Julia f
 * (src/KernelMatrix.jl:47-116, whose far leaves are BarycentricMatrix2D,
 * src/BarycentricMatrix.jl:208-218): leaf t sits at (row0[t], col0[t]) with size
 * m[t] x n[t] in an nrows x ncols matrix, placed where the reference's traversal
 * offsets put it (src/HierarchicalMatrix.jl:45,47; src/KernelMatrix.jl:27-29).
 *   kind[t] = 3: dense, A[t] is m x n column-major.
 *   kind[t] = 2: low rank U Diagonal(S) V' with U[t] m x rank[t], S[t] rank[t],
 *                V[t] n x rank[t] (column-major).
 *   kind[t] = 4: barycentric U F V' (the leaf mul! of src/algebra.jl:243-277:
 *                t1 = V'v, t2 = F t1, u += U t2) with U[t] m x rank[t],
 *                V[t] n x rank[t] and the rank[t] x rank[t] core F passed in
 *                S[t] (column-major, F[i + rank*j]).
 * Leaves may be missing (unassigned blocks are zero). Host pointers; the data
 * is copied into the plan. Products: hm_matvec / hm_matvec_t in stored form
 * (the leaf loops of src/algebra.jl:37-48,110-131,243-277 and their
 * transposes :52-65,137-159). Ranks up to 64. */
int hm_plan_from_blocks(int64_t nrows, int64_t ncols, int64_t nleaves, const int32_t *kind,
                        const int64_t *row0, const int64_t *col0, const int64_t *m,
                        const int64_t *n, const int64_t *rank, const double *const *A,
                        const double *const *U, const double *const *S,
                        const double *const *V, void *stream, hm_plan_t *plan);
/* The same for T = Float32 (the reference's leaf loops are generic in T,
 * src/algebra.jl:37-48,110-131): every leaf and every product in FP32;
 * products through hm_matvec_f32 / hm_matvec_t_f32. Ranks up to 32. */
int hm_plan_from_blocks_f32(int64_t nrows, int64_t ncols, int64_t nleaves, const int32_t *kind,
                            const int64_t *row0, const int64_t *col0, const int64_t *m,
                            const int64_t *n, const int64_t *rank, const float *const *A,
                            const float *const *U, const float *const *S,
                            const float *const *V, void *stream, hm_plan_t *plan);

/* Diagonal scaling of a plan built from blocks, in place on the device
 * (src/HierarchicalMatrix.jl:15-16,54-102 with the leaf forms of
 * src/algebra.jl:282-321):
 *   side = 0: H := H * Diagonal(b[off], ..., b[off + ncols - 1])
 *             (rmul!, scale!(H, b, jstart) with off = jstart - 1): dense leaves
 *             A[i, j] *= b[off + col0 + j], LowRank leaves V[j, :] *= b[off + col0 + j];
 *   side = 1: H := Diagonal(b[off], ..., b[off + nrows - 1]) * H
 *             (lmul!, scale!(b, H, istart) with off = istart - 1): dense leaves
 *             A[i, j] *= b[off + row0 + i], LowRank leaves U[i, :] *= b[off + row0 + i].
 * b is a device vector. Only Float64 plans of hm_plan_from_blocks: the reference
 * defines no scale! for the KernelMatrix leaves (HM_ERR_UNSUPPORTED). A
 * barycentric leaf (kind 4) scales its U or V rows as a low-rank leaf does. */
int hm_plan_scale(hm_plan_t plan, int side, const double *b, int64_t off, void *stream);

/* New point values for a plan of hm_plan_from_points (Float64) or its _f32
 * twin (no counterpart in the reference, which rebuilds KernelMatrix(f, x, y,
 * ...) for other points): x (nx) and y (ny) are DEVICE arrays of the plan's
 * element type. The block tree depends on the points only through the
 * comparisons of indsplit (src/BarycentricMatrix.jl:299-307); the plan keeps
 * its tree and checks on the device that the new points are finite,
 * non-increasing and fall on the same side of every split, writing 0 to the
 * DEVICE int *status when they do and 1 when they do not (status NULL: a
 * word inside the plan). With 1 the plan's products are not those of the
 * KernelMatrix of the new points: rebuild it. Dense blocks and the stored
 * bases U, V are formed again from the new points; F (which depends on the
 * intervals only) stays. Stream-ordered: no synchronization, no allocation,
 * and no memset (the status word is cleared by a kernel), so the call can be
 * captured into a graph. */
int hm_plan_update_points(hm_plan_t plan, const double *x, const double *y, int32_t *status,
                          void *stream);
int hm_plan_update_points_f32(hm_plan_t plan, const float *x, const float *y, int32_t *status,
                              void *stream);

/* add_col!(H, u, istart, j) (src/HierarchicalMatrix.jl:104-146 with the leaf
 * forms of src/algebra.jl:324-380) in place on the device, for a Float64 plan
 * of hm_plan_from_blocks: column j (0-based) of H gains u[ioff + i], i < nrows
 * (ioff = istart - 1). Dense leaves holding column j add their part of u to it
 * (add_col!); low-rank and barycentric leaves take u e_j' as one more rank
 * term (the reference's add_col: U1 = [U u], S1 = [S; 1], V1 = [V e_j]) in
 * their next free column: a leaf whose rank already equals the plan's rank
 * has none, and the call then changes nothing and returns HM_ERR_UNSUPPORTED
 * (rebuild the plan from the updated leaves). u is a device vector;
 * stream-ordered, allocates nothing. */
int hm_plan_add_col(hm_plan_t plan, const double *u, int64_t ioff, int64_t j, void *stream);

/* Many independent KernelMatrix instances in one plan (no counterpart in the
 * reference, which builds one tree per call of src/KernelMatrix.jl:47).
 * Instance t has points x[t] (nx[t]), y[t] (ny[t]) and intervals
 * intervals[4t..4t+3] = (a, b, c, d); kernel, blocksize and rank are shared.
 * The trees are merged into one forest over the concatenated index space, so
 * every pass of hm_matvec is one grouped launch over all instances: the
 * product's vectors are the instances' vectors concatenated (offsets from
 * hm_plan_batch_offsets). */
int hm_plan_batched(int kernel_id, int64_t count, const double *const *x, const int64_t *nx,
                    const double *const *y, const int64_t *ny, const double *intervals,
                    int blocksize, int rank, unsigned flags, void *stream, hm_plan_t *plan);

/* row_offsets[0..count] and col_offsets[0..count] of a batched plan
 * (instance t owns rows [row_offsets[t], row_offsets[t+1])). */
int hm_plan_batch_offsets(hm_plan_t plan, int64_t *row_offsets, int64_t *col_offsets);

/* Frees every device buffer of the plan (synchronises the device first). */
int hm_plan_destroy(hm_plan_t plan);

/* Plan facts, written to info[0..n): see enum hm_info. Replaces size(H) and
 * blocksize(H) of src/hierarchical.jl:31-47 for the root. */
enum hm_info {
    HM_INFO_NROWS = 0, HM_INFO_NCOLS, HM_INFO_RANK, HM_INFO_BLOCKSIZE, HM_INFO_KERNEL,
    HM_INFO_N_D, HM_INFO_N_K1, HM_INFO_N_K2, HM_INFO_N_DENSE, HM_INFO_N_FAR,
    HM_INFO_N_ROWCLUSTERS, HM_INFO_N_COLCLUSTERS, HM_INFO_DEPTH,
    HM_INFO_DENSE_ENTRIES, HM_INFO_BASIS_ROWS_U, HM_INFO_BASIS_ROWS_V,
    HM_INFO_BYTES_STORED,      /* bytes of every stored array read by one product */
    HM_INFO_BYTES_REF_LAYOUT,  /* same arrays in the reference's per-block layout */
    HM_INFO_FLAGS, HM_INFO_N_ROWCHUNKS, HM_INFO_N_COLCHUNKS, HM_INFO_NONFINITE,
    HM_INFO_N_INSTANCES, HM_INFO_FROM_BLOCKS,
    HM_INFO_FLOAT32,           /* 1 for a plan built by hm_plan_from_points_f32 */
    HM_INFO_NESTED,            /* 1 when the MFMA passes apply the far field through the
                                  interval trees (each point in one leaf's moments, each
                                  row in one leaf's series); 0: per covering cluster */
    HM_INFO_COUNT
};
int hm_plan_info(hm_plan_t plan, int64_t *info, int n);

/* Leaf blocks in the order the reference's traversal mul! visits them
 * (src/KernelMatrix.jl:17-45). ints: 6 per block = code (2 far, 3 dense),
 * depth, row0, col0, m, n; dbls: 4 per block = a, b, c, d (far blocks).
 * *count receives the total; at most cap records are written. */
int hm_plan_blocks(hm_plan_t plan, int64_t *ints, double *dbls, int64_t cap, int64_t *count);

/* y[ioff + i*incy + q*ldy] = beta*y[...] + sum_j H[i,j] x[joff + j*incx + q*ldx]
 * for q < k right-hand sides, beta in {0, 1}. beta = 1 is the reference's
 * accumulating mul!(u, H, v, istart, jstart) (src/KernelMatrix.jl:14-45, the
 * leaf loops src/algebra.jl:34-48 and :243-277); beta = 0 is H*v
 * (src/KernelMatrix.jl:5-8). Hybrid and matrix-free products with at least
 * 2^24 near-field entries run the near field on a stream the plan owns,
 * forked from and joined back to `stream` by events; a product stays on
 * `stream` while it is being captured or pass profiling is on. */
int hm_matvec(hm_plan_t plan, double *y, const double *x, int64_t ioff, int64_t joff,
              int64_t incy, int64_t incx, int64_t k, int64_t ldy, int64_t ldx, int beta,
              unsigned flags, void *stream);

/* Transpose product y = beta*y + H^T x (same argument meaning, y indexed by
 * columns of H). No counterpart for the tree in the reference; its leaf forms
 * are src/algebra.jl:50-82 (dense) and :133-159 (low rank). */
int hm_matvec_t(hm_plan_t plan, double *y, const double *x, int64_t ioff, int64_t joff,
                int64_t incy, int64_t incx, int64_t k, int64_t ldy, int64_t ldx, int beta,
                unsigned flags, void *stream);

/* Gradient of the product with respect to the row points x (no counterpart in
 * the reference; it serves the rrule of L(x, y, b)):
 *   g[ioff + i*incg] = beta*g[...] + s[i*incs] * d/dx_i sum_j H(x, y)[i, j] b[joff + j*incb]
 * i.e. the exact derivative of the computed product (near field regenerated
 * with df/dx, far field through the derivative of its Chebyshev-form basis).
 * With s = the output cotangent this is the x-gradient. One right-hand side. */
int hm_matvec_grad_x(hm_plan_t plan, double *g, const double *b, const double *s, int64_t ioff,
                     int64_t joff, int64_t incg, int64_t incb, int64_t incs, int beta, void *stream);

/* The product and its x-derivative sums in one pass (no counterpart in the
 * reference; the forward counterpart of hm_matvec_t_grad):
 *   y[i*incy] = beta*y[...] + (H b)_i,
 *   g[i*incg] = beta*g[...] + s[i*incs] * d/dx_i (H b)_i   (s NULL: ones)
 * with H b computed as hm_matvec_grad_x differentiates it (near field
 * regenerated, far field in Chebyshev form). A rrule that keeps g from the
 * forward (s NULL) forms the x-gradient as r .* g, with no pass of its own.
 * One right-hand side. */
int hm_matvec_grad(hm_plan_t plan, double *y, double *g, const double *b, const double *s,
                   int64_t incy, int64_t incg, int64_t incb, int64_t incs, int beta, void *stream);

/* Adjoint pass for the b- and y-gradients (no counterpart in the reference):
 *   gb = beta*gb + H^T r,   gy[j] = beta*gy[j] + s[j] * d/dy_j (H^T r)_j
 * with r the output cotangent; s = b gives the y-gradient. gb and gy are
 * indexed by columns (gb[j*incgb], gy[j*incgy], s[j*incs], r[i*incr]). */
int hm_matvec_t_grad(hm_plan_t plan, double *gb, double *gy, const double *r, const double *s,
                     int64_t incgb, int64_t incgy, int64_t incr, int64_t incs, int beta,
                     void *stream);

/* Executable HIP graph of one hm_matvec call (no counterpart in the
 * reference): the launches of hm_matvec(plan, y, x, ...) with these exact
 * arguments are captured once and replayed by hm_graph_launch on any stream.
 * The graph holds only kernel nodes (no memset or copy nodes); the plan must
 * outlive it, and replays on one plan must not overlap (they share the
 * plan's workspace). */
typedef struct hm_graph *hm_graph_t;
int hm_graph_matvec(hm_plan_t plan, double *y, const double *x, int64_t ioff, int64_t joff,
                    int64_t incy, int64_t incx, int64_t k, int64_t ldy, int64_t ldx, int beta,
                    unsigned flags, hm_graph_t *graph);
int hm_graph_launch(hm_graph_t graph, void *stream);
int hm_graph_destroy(hm_graph_t graph);

/* Autotuner (no counterpart in the reference): times hm_matvec on this plan
 * for k right-hand sides (device buffers y, x as in hm_matvec with unit
 * stride and ld = n) over the candidate launch shapes (right-hand sides per
 * launch 1, 2, 4, 8 on the vector units; 16 on the MFMA passes, Float64
 * KernelMatrix plans) and far-field forms (the plan's own, Chebyshev basis,
 * matrix-free), keeps the fastest for later products that pass HM_MV_TUNED,
 * and writes its choice to out[0] = right-hand sides per launch, out[1] =
 * mode (0 stored, 1 Chebyshev basis, 2 matrix-free), out[2] = its time in
 * microseconds. Every candidate computes the same product up to rounding. */
int hm_plan_tune(hm_plan_t plan, void *y, const void *x, int64_t k, void *stream,
                 int64_t out[3]);  /* y, x hold the plan's element type */

/* Per-pass timing of the products of this plan (bench support, no counterpart
 * in the reference). enable=1 starts recording HIP events around the four
 * passes of every later product (up, reduce, translate, down) on its stream;
 * enable=0 stops and discards. hm_plan_profile_read synchronises, returns the
 * summed milliseconds per pass in ms[0..3] and the number of products. */
int hm_plan_profile(hm_plan_t plan, int enable);
int hm_plan_profile_read(hm_plan_t plan, double ms[4], int64_t *count);

/* out[t] = H[i[t], j[t]] (0-based) for t < n; i, j, out are device pointers.
 * Replaces getindex(H, i, j), src/hierarchical.jl:120-147 with
 * src/BarycentricMatrix.jl:222-234 at far blocks. */
int hm_entries(hm_plan_t plan, int64_t n, const int64_t *i, const int64_t *j, double *out,
               void *stream);

/* ---------------------------------------------------------------- Float32
 * KernelMatrix{Float32}: the reference's constructors and mul! are generic in
 * the element type T (src/KernelMatrix.jl:47-116, src/BarycentricMatrix.jl:
 * 92-136,147-178,248-307, src/algebra.jl:34-48,243-277), so a Float32 tree
 * splits at Float32 midpoints, its nodes, weights, F, U, V and dense blocks are
 * Float32 values formed in Float32 arithmetic, and mul! accumulates in
 * Float32. These entry points build and apply such a plan: points, intervals
 * and vectors are float; the tree rule, layout and passes are those of the
 * Float64 plan, with every stored array and every accumulation in FP32.
 * Ranks up to 32 (BLOCKRANK(Float32) = 10). Reduced-precision far-field flags
 * do not apply. The Float64 product entry points reject a Float32 plan and
 * the _f32 ones a Float64 plan (HM_ERR_INVALID). */
int hm_plan_from_points_f32(int kernel_id, const float *x, int64_t nx, const float *y, int64_t ny,
                            float a, float b, float c, float d, int blocksize, int rank,
                            unsigned flags, void *stream, hm_plan_t *plan);
int hm_matvec_f32(hm_plan_t plan, float *y, const float *x, int64_t ioff, int64_t joff,
                  int64_t incy, int64_t incx, int64_t k, int64_t ldy, int64_t ldx, int beta,
                  unsigned flags, void *stream);
int hm_matvec_t_f32(hm_plan_t plan, float *y, const float *x, int64_t ioff, int64_t joff,
                    int64_t incy, int64_t incx, int64_t k, int64_t ldy, int64_t ldx, int beta,
                    unsigned flags, void *stream);
int hm_matvec_grad_x_f32(hm_plan_t plan, float *g, const float *b, const float *s, int64_t ioff,
                         int64_t joff, int64_t incg, int64_t incb, int64_t incs, int beta,
                         void *stream);
int hm_matvec_grad_f32(hm_plan_t plan, float *y, float *g, const float *b, const float *s,
                       int64_t incy, int64_t incg, int64_t incb, int64_t incs, int beta,
                       void *stream);
int hm_matvec_t_grad_f32(hm_plan_t plan, float *gb, float *gy, const float *r, const float *s,
                         int64_t incgb, int64_t incgy, int64_t incr, int64_t incs, int beta,
                         void *stream);
int hm_graph_matvec_f32(hm_plan_t plan, float *y, const float *x, int64_t ioff, int64_t joff,
                        int64_t incy, int64_t incx, int64_t k, int64_t ldy, int64_t ldx, int beta,
                        unsigned flags, hm_graph_t *graph);
int hm_entries_f32(hm_plan_t plan, int64_t n, const int64_t *i, const int64_t *j, float *out,
                   void *stream);

/* ------------------------------------------------ barycentric pieces
 * BarycentricPoly2D and BarycentricMatrix2D (src/BarycentricMatrix.jl:139-297)
 * outside a KernelMatrix plan. All arrays are DEVICE arrays; stream-ordered.
 * The reference's operation order with IEEE division. */

/* out[t] = evaluate(B, px[t], py[t]) for t < n (src/BarycentricMatrix.jl:180-205):
 * the 2-D barycentric interpolant of F (r x r column-major, F[m + r*n] =
 * f(tx_m, ty_n)) on the mapped nodes tx, ty with weights lamx, lamy. */
int hm_bary_evaluate(int r, const double *tx, const double *ty, const double *lamx,
                     const double *lamy, const double *F, int64_t n, const double *px,
                     const double *py, double *out, void *stream);
int hm_bary_evaluate_f32(int r, const float *tx, const float *ty, const float *lamx,
                         const float *lamy, const float *F, int64_t n, const float *px,
                         const float *py, float *out, void *stream);

/* The update! basis of one side of a BarycentricMatrix2D
 * (src/BarycentricMatrix.jl:248-297): B[i + ldb*m] = (lam_m / (pts_i - nodes_m))
 * / sum_k lam_k / (pts_i - nodes_k) for i < n, m < r (r <= 64). */
int hm_bary_basis(int r, const double *nodes, const double *lam, int64_t n, const double *pts,
                  double *B, int64_t ldb, void *stream);
int hm_bary_basis_f32(int r, const float *nodes, const float *lam, int64_t n, const float *pts,
                      float *B, int64_t ldb, void *stream);

/* ------------------------------------------------ EvenBarycentricMatrix
 * EvenBarycentricMatrix(T, f, a, b, c, d) (src/BarycentricMatrix.jl:5-46) is an
 * m x p block, m = b-a+1 rows on the integer grid a..b and p = d-c+1 columns
 * on c..d, of rank n: W (n x m) interpolates in the row index at the n
 * first-kind Chebyshev nodes t_k = (a+b)/2 + (b-a)*x_k/2, and F (p x n) holds
 * f(T, t_k, j). f is the caller's function, so F comes from the caller; W is
 * formed on the device. Ranks 1..64; Float64 entry points and Float32 twins
 * (suffix _f32: W, F, u, v in Float32, computed in Float32 as the reference's
 * EvenBarycentricMatrix{Float32}). A plan is of one element type. */
typedef struct hm_even *hm_even_t;

/* t_k, k = 1..n, the arguments at which F[j, k] = f(T, t_k, j) is evaluated
 * (src/BarycentricMatrix.jl:41). Host output. */
int hm_even_nodes(int64_t a, int64_t b, int n, double *t);
/* The same for T = Float32: x_k and (b-a)*x_k/2 in Float32, (a+b)/2 and the
 * sum in Float64, as the reference's promotion gives. */
int hm_even_nodes_f32(int64_t a, int64_t b, int n, double *t);

/* W of the grid a..b into the DEVICE array W, and optionally the sums
 * w (m values, device, NULL to skip): layout 0 stores W[k + n*(i-a)], the n x m
 * W of EvenBarycentricMatrix (src/BarycentricMatrix.jl:23-37); layout 1 stores
 * W[(i-a) + m*k], the m x n W of barycentricmatrix (src/BarycentricMatrix.jl:
 * 61-89, whose LowRankMatrix(W, I, F) then goes to hm_plan_from_blocks).
 * Synchronizes `stream`. */
int hm_even_weights(int64_t a, int64_t b, int n, int layout, double *W, double *w, void *stream);
int hm_even_weights_f32(int64_t a, int64_t b, int n, int layout, float *W, float *w, void *stream);

/* A plan of nleaves EvenBarycentricMatrix leaves of rank n placed in one
 * nrows x ncols matrix: leaf t spans rows row0[t] .. row0[t] + m - 1 and
 * columns col0[t] .. col0[t] + p - 1, with grids abcd[4t .. 4t+3] = (a, b, c, d)
 * and F[t] (host, p x n column-major). Leaves may overlap; products add
 * them all. W is formed on the device, once per distinct leaf row count
 * (the values depend on the rows only through i - a and b - a). */
int hm_even_create(int64_t nrows, int64_t ncols, int64_t nleaves, const int64_t *row0,
                   const int64_t *col0, const int64_t *abcd, int n, const double *const *F,
                   void *stream, hm_even_t *plan);
int hm_even_create_f32(int64_t nrows, int64_t ncols, int64_t nleaves, const int64_t *row0,
                       const int64_t *col0, const int64_t *abcd, int n, const float *const *F,
                       void *stream, hm_even_t *plan);

/* Replaces mul!(u, B, v, istart, jstart) (src/algebra.jl:168-239) for every
 * leaf at once: u[q*ldu + (ioff + I)*incu] (+)= sum over leaves and columns J
 * with (ioff + I) + (joff + J) even of B_t(I - row0, J - col0) *
 * v[q*ldv + (joff + J)*incv], for right-hand sides q < k. The parity rule is
 * the reference's: its iseven(ishift + jshift) branches pair row i with
 * column j exactly when ishift + i + jshift + j is even. beta = 1 accumulates
 * (the reference's mul!), beta = 0 overwrites rows 0 .. nrows-1 of each
 * column. u, v: device. The plan's element type must match the entry point's
 * (HM_ERR_INVALID otherwise). */
int hm_even_matvec(hm_even_t plan, double *u, const double *v, int64_t ioff, int64_t joff,
                   int64_t incu, int64_t incv, int64_t k, int64_t ldu, int64_t ldv, int beta,
                   void *stream);
int hm_even_matvec_f32(hm_even_t plan, float *u, const float *v, int64_t ioff, int64_t joff,
                       int64_t incu, int64_t incv, int64_t k, int64_t ldu, int64_t ldv, int beta,
                       void *stream);

/* Host copies of leaf t's W (n x m) and F (p x n); either may be NULL. */
int hm_even_leaf_data(hm_even_t plan, int64_t leaf, double *W, double *F);
int hm_even_leaf_data_f32(hm_even_t plan, int64_t leaf, float *W, float *F);

/* info[0..]: nrows, ncols, nleaves, rank, W entries (stored: leaves with the
 * same row count share one bit-identical W), F entries, upward tasks,
 * reduced leaves, downward tasks, segment-list entries, element type (0:
 * Float64, 1: Float32), distinct W blocks. */
enum { HM_EVEN_INFO_COUNT = 12 };
int hm_even_info(hm_even_t plan, int64_t *info, int n);
int hm_even_destroy(hm_even_t plan);

/* ------------------------------------------------ LowRankMatrix algebra
 * Batched recompression arithmetic of src/LowRankMatrix.jl:64-171 on the
 * device: `count` independent items of one operation. A LowRankMatrix is
 * U diag(S) V^T with U m x r, V n x r (column-major, leading dimensions m and
 * n). As in the reference, each result is (Q_U SU)[:, 1:r], S[1:r],
 * (Q_V SV)[:, 1:r] from a thin QR of the tall factors and an SVD of the small
 * core, truncated by getrank (src/LowRankMatrix.jl:70-80: r = the largest
 * index with S[r] > length(S) * eps(S[1])). Singular vectors are determined
 * up to sign, as LAPACK's are. */
enum hm_lr_op {
    HM_LR_ADD = 0,      /* L1 + L2 (src/LowRankMatrix.jl:95-102), both m x n */
    HM_LR_SUB = 1,      /* L1 - L2 (src/LowRankMatrix.jl:104-111) */
    HM_LR_MUL = 2,      /* L1 * L2, L1 m x k, L2 k x n (src/LowRankMatrix.jl:113-118) */
    HM_LR_ADJMUL = 3,   /* L1' * L2, L1 k x m, L2 k x n (src/LowRankMatrix.jl:120-126) */
    HM_LR_DMUL = 4,     /* A * L2, A m x k dense (src/LowRankMatrix.jl:128-134; a Diagonal
                           A, :144-150, is passed as its dense form) */
    HM_LR_MULD = 5,     /* L1 * A, A k x n dense (src/LowRankMatrix.jl:136-142,152-158) */
    HM_LR_SVDTRUNC = 6  /* convert(LowRankMatrix, A), A m x n dense (src/LowRankMatrix.jl:82-86) */
};

/* dims: 5 per item, (m, n, k, r1, r2) (k, r1 or r2 unused where the operation
 * has no such operand). in: 7 DEVICE pointers per item, (U1, S1, V1, U2, S2,
 * V2, A), NULL where unused. out: 3 DEVICE pointers per item, (U, S, V) with
 * room for rmax columns: rmax = r1 + r2 (sums), min(r1, r2) (products of two
 * LowRankMatrix), r2 (A * L2), r1 (L1 * A), min(m, n) (svdtrunc); only the
 * first rank[t] columns are written. rank: HOST array of count values.
 * Limits: ranks 1..64; sums need r1 + r2 <= 64 and <= min(m, n); A * L2 needs
 * m >= r2, L1 * A needs n >= r1; svdtrunc needs min(m, n) <= 64
 * (HM_ERR_UNSUPPORTED otherwise). Synchronizes `stream`. */
int hm_lowrank_op(int op, int64_t count, const int64_t *dims, const double *const *in,
                  double *const *out, int64_t *rank, void *stream);

/* ------------------------------------------------ dense leaf steps of the
 * hierarchical Cholesky factorization (src/cholesky.jl:12-228), batched:
 * `count` independent items of free sizes, column-major DEVICE arrays with
 * per-item leading dimensions (host arrays of sizes and pointers).
 * hm_dense_potrf synchronizes `stream` (it reports the pivot); the other two
 * are stream-ordered and return at once for count == 1, and synchronize for
 * count > 1. */

/* Replaces hierarchicalcholesky(A::Matrix) = Matrix(cholesky(Symmetric(A)).U)
 * (src/cholesky.jl:18-20): A[t] (n x n) is overwritten by the upper factor R,
 * R^T R = A, read from A's upper triangle; the strict lower triangle is set
 * to zero. info (host, count): 0, or j + 1 when pivot j is not positive, in
 * which case the call returns HM_ERR_INVALID (Julia's PosDefException). */
int hm_dense_potrf(int64_t count, const int64_t *n, double *const *A, const int64_t *lda,
                   int64_t *info, void *stream);
/* The same factorization with info in DEVICE memory (count values, written
 * by the kernel) and no synchronization: the call is stream-ordered, so a
 * recursion of many leaves runs without a wait per leaf and the caller reads
 * the info values once. Factors after a failed pivot hold NaNs; every later
 * step on them ends normally. */
int hm_dense_potrf_dev(int64_t count, const int64_t *n, double *const *A, const int64_t *lda,
                       int64_t *info, void *stream);

/* Replaces solvetransposed(R::Matrix, b) = UpperTriangular(R)' \ b and its
 * Matrix / LowRankMatrix right-hand sides (src/cholesky.jl:116-142):
 * B[t] (n x k) := R[t]^{-T} B[t], R upper n x n. */
int hm_dense_trsm_t(int64_t count, const int64_t *n, const int64_t *k, const double *const *R,
                    const int64_t *ldr, double *const *B, const int64_t *ldb, void *stream);

/* C[t] = beta C[t] + alpha op(A[t]) diag(D[t]) op(B[t]) (op = transpose when
 * trans = 1; D NULL, or D[t] NULL, for no diagonal), mnk = (m, n, k) per
 * item: the products of the Schur-complement updates A22 - R12' R12 and of
 * b2 - R12' x1 (src/cholesky.jl:22-95,150-228), with a LowRankMatrix's
 * U diag(S) V' applied through D. */
int hm_dense_gemm(int transa, int transb, int64_t count, const int64_t *mnk, double alpha,
                  const double *const *A, const int64_t *lda, const double *const *D,
                  const double *const *B, const int64_t *ldb, double beta, double *const *C,
                  const int64_t *ldc, void *stream);

/* hm_dense_gemm without allocating: no counterpart in the reference; for
 * callers that cannot allocate (code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc). Same arguments, checks and result
 * as hm_dense_gemm. The item table goes to a fixed device array of
 * HM_DENSE_NOALLOC_MAX items (each GPU holds its own copy until the process
 * ends); a larger count returns HM_ERR_UNSUPPORTED. Every call on one GPU
 * shares that array, which keeps the last call's items: calls that overlap
 * on one GPU (other host threads, other streams) race. */
#define HM_DENSE_NOALLOC_MAX 8192
int hm_dense_gemm_noalloc(int transa, int transb, int64_t count, const int64_t *mnk, double alpha,
                          const double *const *A, const int64_t *lda, const double *const *D,
                          const double *const *B, const int64_t *ldb, double beta,
                          double *const *C, const int64_t *ldc, void *stream);

#ifdef __cplusplus
}
#endif
#endif
