// hmx_internal.h -- device layout of a flattened KernelMatrix tree.
//
// The reference keeps a tree of 2x2 nodes with an `assigned` code matrix per
// node (src/hierarchical.jl:49-52) and walks it depth first for every product
// (src/KernelMatrix.jl:17-45). Here the tree is flattened once, at plan time,
// into level-ordered tables in HBM:
//
//   row clusters   one per distinct (depth, row range) that is the row side of
//                  a far block; owns U (len x r, column-major) and its mapped
//                  Chebyshev nodes. Far blocks that share a row cluster share
//                  one U (the reference stores identical copies per block).
//   col clusters   the same for the column side and V.
//   far blocks     (row cluster, col cluster, F offset); F is r x r column-major.
//   dense blocks   (row0, col0, m, n, offset); A is m x n column-major.
//   row chunks     <= 64 consecutive rows that every dense block and row
//                  cluster either covers completely or misses; each lists its
//                  dense blocks (near) and row clusters (far). One wavefront
//                  owns one chunk, so the output needs no atomics.
//   col chunks     the same over columns.
//   tasks          (cluster, range of <= HM_TASK_POINTS points): one workgroup
//                  of the upward pass reduces B[range, :]^T x[range] into one
//                  partial slot; a cluster's slots are consecutive.
#pragma once
#include <stdint.h>

#define HM_CHUNK 64          // rows per chunk = wavefront width
#define HM_TASK_POINTS 8192  // points per upward-pass task
#define HM_MAX_KB 8          // right-hand sides per forward launch (FP64 passes)
#define HM_WS_RHS 32         // right-hand sides the workspace holds (an MFMA launch: 32)

// Stored matrices (dense blocks m x n, bases len x r) are kept in 64-row
// tiles, column-major inside a tile: element (i, k) of a matrix with `cols`
// columns sits at tile_index(i, k, cols). A wavefront that owns 64 rows then
// reads one contiguous 64*cols run.
#define HM_TILE 64
#ifdef __HIPCC__
__host__ __device__
#endif
inline int64_t tile_index(int64_t i, int64_t k, int64_t cols) {
    return (i >> 6) * (HM_TILE * cols) + k * HM_TILE + (i & (HM_TILE - 1));
}
inline int64_t tiled_size(int64_t rows, int64_t cols) {
    return ((rows + HM_TILE - 1) / HM_TILE) * HM_TILE * cols;
}

struct DChunk {
    int64_t p0;        // first row (or column)
    int32_t np;        // number of rows in the chunk (<= HM_CHUNK)
    int32_t near_beg;  // CSR range into DNear
    int32_t near_end;
    int32_t far_beg;   // CSR range into DFar
    int32_t far_end;
    int32_t pad;
};

struct DNear {
    int32_t blk;  // dense block id
    int32_t off;  // offset of the chunk inside the block (rows, or columns when transposed)
};

struct DFar {
    int32_t cl;   // cluster id
    int32_t off;  // offset of the chunk inside the cluster
};

struct DTask {
    int32_t cl;      // cluster id
    int32_t off;     // first point of the task inside the cluster
    int32_t len;     // number of points
    int32_t single;  // 1 if the task is its cluster's only one
    int64_t slot;    // partial-sum slot
    int64_t first;   // first point of the task (cluster first + off)
    double cx, cy;   // the cluster's (center, 1/halfwidth), Float64 plans
};

struct DDense {
    int64_t row0, col0;
    int32_t m, n;
    int64_t aoff;  // offset of A in the dense store
};

struct DCluster {
    int64_t first;  // first point index
    int64_t boff;   // offset of the basis (U or V) in the basis store
    int64_t pbase;  // first partial-sum slot (its tasks' slots are consecutive)
    int32_t len;    // number of points
    int32_t nslots; // number of tasks of this cluster
    int32_t far_beg, far_end;  // CSR range into the far-block list grouped by this cluster
    int32_t depth, pad;
};

struct DFarBlk {
    int32_t rc, cc;  // row cluster, col cluster
    int64_t foff;    // offset of F
};

// A far block as the translation of one side reads it (the records of a
// cluster's list in order): F's offset and the source cluster (column cluster
// forward, row cluster transposed) with its task count and first partial-sum
// slot, so a source with several upward tasks is summed from its slots.
struct DTrRec {
    int64_t foff;
    int64_t pbase;
    int32_t src;
    int32_t nslots;
};

// A node of one side's interval tree: a range that the tree rule splits (or
// a half it leaves whole), with its two halves. The nested far field of the
// MFMA passes sums Chebyshev moments from the halves into their parent and
// pushes expansion coefficients from a parent into its halves, one fixed
// re-expansion per half, so each point enters one leaf's moments and each row
// evaluates one leaf's series.
struct DTNode {
    int64_t first;
    int32_t len;
    int32_t parent;    // -1 at a root
    int32_t child[2];  // -1 at a leaf
    int32_t cl;        // the row / column cluster with this range, or -1
    int32_t half;      // which half of its parent (0: the half holding the upper interval end)
    double cx, cy;     // (center, 1/halfwidth) of its interval, as the cluster maps form them
};

// One indsplit of the tree: the range [first, first + len) of one side split
// at index `at` by the midpoint `mid` (points at or above it first). For other
// points the split falls at the same index exactly when p[at - 1] >= mid (if
// at > first) and p[at] < mid (if at < first + len); an empty range needs
// p[first] < mid, as the reference reads the point after it.
struct DSplit {
    int64_t first, len, at;
    double mid;
};
