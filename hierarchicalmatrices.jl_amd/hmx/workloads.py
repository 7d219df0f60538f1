"""Synthetic inputs of the benchmark and test workloads, from a documented RNG.

Random numbers come from splitmix64 (Steele, Lea, Flood 2014): state_k =
seed + k*0x9E3779B97F4A7C15 (k = 1, 2, ...), mixed by the published finaliser.
Uniforms are the top 53 bits times 2^-53; normals are Box-Muller pairs
(u1 in (0, 1], u2 in [0, 1)). Every workload is a pure function of its seed.

Points are returned in non-increasing order, as the reference's indsplit
requires (src/BarycentricMatrix.jl:299-307); the intervals are (upper, lower),
as examples/Kernel.jl:66 passes them.
"""
import numpy as np

_GOLDEN = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)


def splitmix64(seed, n, offset=0):
    with np.errstate(over="ignore"):
        k = np.arange(offset + 1, offset + n + 1, dtype=np.uint64)
        z = np.uint64(seed) + k * _GOLDEN
        z = (z ^ (z >> np.uint64(30))) * _M1
        z = (z ^ (z >> np.uint64(27))) * _M2
        return z ^ (z >> np.uint64(31))


def uniform(seed, n, offset=0):
    """n uniforms in [0, 1)."""
    return (splitmix64(seed, n, offset) >> np.uint64(11)).astype(np.float64) * 2.0 ** -53


def normal(seed, n, offset=0):
    """n standard normals (Box-Muller on a pair of uniform streams)."""
    m = (n + 1) // 2
    u = uniform(seed, 2 * m, offset)
    u1 = 1.0 - u[0::2]  # (0, 1]
    u2 = u[1::2]
    rad = np.sqrt(-2.0 * np.log(u1))
    z = np.empty(2 * m)
    z[0::2] = rad * np.cos(2.0 * np.pi * u2)
    z[1::2] = rad * np.sin(2.0 * np.pi * u2)
    return z[:n]


def shifted_grid(n):
    """x_i = (i - 1/4)/n, y_j = (j - 3/4)/n for i, j = n..1 (non-increasing;
    x_i - y_j = (i - j + 1/2)/n never vanishes); intervals (1, 0)."""
    i = np.arange(n, 0, -1, dtype=np.float64)
    return (i - 0.25) / n, (i - 0.75) / n, (1.0, 0.0, 1.0, 0.0)


def config_plain(n=4096, blocksize=64, rank=16, seed=0):
    """Cauchy kernel on the shifted grid, b ~ N(0, 1) (seed 0)."""
    x, y, iv = shifted_grid(n)
    return dict(kernel="cauchykernel", x=x, y=y, intervals=iv, blocksize=blocksize, rank=rank,
                b=normal(seed, n))


def config_headline(seed=0):
    """n = 2^20, leaf threshold 256, rank 24, single RHS."""
    return config_plain(n=1 << 20, blocksize=256, rank=24, seed=seed)


def sorted_uniform_points(n, seed):
    """x = sort(U[0,1]) non-increasing, y = x + 0.5/n; intervals = extrema."""
    x = np.sort(uniform(seed, n))[::-1].copy()
    y = x + 0.5 / n
    return x, y, (float(x[0]), float(x[-1]), float(y[0]), float(y[-1]))


def config_batched(count=256, nmin_log2=10, nmax_log2=16, k=32, rank=16, blocksize=64, seed=0):
    """Independent Cauchy instances with n log-uniform in [2^10, 2^16]."""
    u = uniform(seed, count)
    ns = np.round(2.0 ** (nmin_log2 + (nmax_log2 - nmin_log2) * u)).astype(np.int64)
    inst = []
    for t, n in enumerate(ns):
        x, y, iv = sorted_uniform_points(int(n), seed * 1000003 + 1 + t)
        b = normal(seed * 1000003 + 7 + t, int(n) * k).reshape(int(n), k)
        inst.append(dict(x=x, y=y, intervals=iv, b=b))
    return dict(kernel="cauchykernel", rank=rank, blocksize=blocksize, k=k, instances=inst)


def config_gradient(n=1 << 16, blocksize=64, rank=16, seed=0):
    """Shifted grid perturbed by U(-0.1/n, 0.1/n); b, t ~ N(0, 1)."""
    x, y, iv = shifted_grid(n)
    x = x + (uniform(seed + 11, n) - 0.5) * 0.2 / n
    y = y + (uniform(seed + 12, n) - 0.5) * 0.2 / n
    return dict(kernel="cauchykernel", x=x, y=y, intervals=iv, blocksize=blocksize, rank=rank,
                b=normal(seed + 13, n), t=normal(seed + 14, n))


def config_scale(n=1 << 22, k=8, blocksize=512, rank=32, seed=1):
    """log|x - y| on sorted uniform points, y = x + 0.5/n."""
    x, y, iv = sorted_uniform_points(n, seed)
    return dict(kernel="logkernel", x=x, y=y, intervals=iv, blocksize=blocksize, rank=rank,
                b=normal(seed + 5, n * k).reshape(n, k), k=k)


def even_hodlr_leaves(N, levels):
    """Off-diagonal blocks of an N x N matrix split dyadically: at level l the
    blocks (2s, 2s+1) and (2s+1, 2s) of size N / 2^l, l = 1..levels. Each is
    (row0, col0, a, b, c, d), 0-based placement and the block's 1-based global
    row and column indices as its integer grids."""
    out = []
    for l in range(1, levels + 1):
        s = N >> l
        for q in range(0, 1 << l, 2):
            for bi, bj in ((q, q + 1), (q + 1, q)):
                r0, c0 = bi * s, bj * s
                out.append((r0, c0, r0 + 1, r0 + s, c0 + 1, c0 + s))
    return out


def gaussian_even_kernel(N):
    """f(T, t, j) = exp(-((t - j)/N)^2): entire, so rank-20 interpolation on
    any sub-interval of 1..N is accurate to round-off."""
    def f(T, t, j):
        return np.exp(-((t - j) / N) ** 2)
    return f


def config_even(N=1 << 20, levels=10, rank=20, seed=0):
    """EvenBarycentricMatrix leaves of an N x N off-diagonal dyadic partition
    (2046 leaves at N = 2^20, levels 10: 1024-point blocks at the finest
    level), rank 20 = BLOCKRANK(Float64), Gaussian f, v ~ N(0, 1)."""
    return dict(N=N, levels=levels, rank=rank, leaves=even_hodlr_leaves(N, levels),
                f=gaussian_even_kernel(N), v=normal(seed + 21, N))


def spd_points(n, seed=0):
    """Sorted uniform points on [0, 1] (ascending) for the SPD test matrices."""
    return np.sort(uniform(seed + 31, n))


def spd_kernel(x, ell=0.25, shift=1.0):
    """K[i, j] = exp(-((x_i - x_j)/ell)^2) + shift * delta_ij: a Gaussian
    kernel matrix (positive semidefinite) made positive definite by the shift."""
    d = (x[:, None] - x[None, :]) / ell
    return np.exp(-d * d) + shift * np.eye(len(x))


def spd_hierarchical(K, leaf, tol=1e-13, dense_offdiag=(), dense_diag_at=()):
    """A 2 x 2 HierarchicalMatrix of the SPD matrix K stored by its upper
    blocks, as src/cholesky.jl expects: diagonal blocks recurse until they
    are at most `leaf` wide (then Matrix), the (1, 2) block is a
    LowRankMatrix from the SVD of K's block truncated at tol * sigma_1 (a
    Matrix at the depths in dense_offdiag), the (2, 1) block is unassigned.
    Diagonal blocks at the depths in dense_diag_at stay dense whatever their
    size."""
    from .hierarchical import HierarchicalMatrix, LowRankMatrix
    from .kernelmatrix import Block

    def build(r0, r1, depth):
        n = r1 - r0
        if n <= leaf or depth in dense_diag_at:
            return np.asfortranarray(K[r0:r1, r0:r1])
        s = r0 + n // 2
        H = HierarchicalMatrix(2, 2)
        H[Block(1), Block(1)] = build(r0, s, depth + 1)
        B = K[r0:s, s:r1]
        if depth in dense_offdiag:
            H[Block(1), Block(2)] = np.asfortranarray(B)
        else:
            U, S, Vt = np.linalg.svd(B, full_matrices=False)
            r = max(1, int(np.sum(S > tol * S[0])))
            H[Block(1), Block(2)] = LowRankMatrix(U[:, :r], S[:r], Vt[:r].T)
        H[Block(2), Block(2)] = build(s, r1, depth + 1)
        return H

    return build(0, K.shape[0], 0)
