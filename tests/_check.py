"""Host restatement of the surrogate's sampled exact check (bf_exact.hip,
check_sample / check_kernel / ring_check_kernel): per 64x64 tile, one sample
per 8x8 cell at a hashed position; a tile falls back when any in-image sample
differs from the exact filter by more than the tolerance."""
import numpy as np

SALT = 0x5bd1e995
M32 = 0xFFFFFFFF


def hash32(x):
    """The kernels' 32-bit integer hash, on numpy uint64 arrays or ints."""
    x = np.asarray(x, np.uint64) & M32
    x ^= x >> np.uint64(16)
    x = (x * np.uint64(0x7feb352d)) & np.uint64(M32)
    x ^= x >> np.uint64(15)
    x = (x * np.uint64(0x846ca68b)) & np.uint64(M32)
    x ^= x >> np.uint64(16)
    return x


def sample_positions(tx, ty, salt=SALT):
    """(X, Y) int64 arrays of shape (tx*ty, 64): the sample of each cell."""
    tiles = np.arange(tx * ty, dtype=np.uint64)[:, None]
    cells = np.arange(64, dtype=np.uint64)[None, :]
    h = hash32((tiles * np.uint64(64) + cells + np.uint64(salt)) & np.uint64(M32))
    sx = (cells & np.uint64(7)) * np.uint64(8) + (h & np.uint64(7))
    sy = (cells >> np.uint64(3)) * np.uint64(8) + ((h >> np.uint64(3)) & np.uint64(7))
    X = (tiles % np.uint64(tx)) * np.uint64(64) + sx
    Y = (tiles // np.uint64(tx)) * np.uint64(64) + sy
    return X.astype(np.int64), Y.astype(np.int64)


def expected_fallback(sur, ex, tol, salt=SALT):
    """Boolean per tile: some sampled |sur - ex| > tol.  Returns (bad, tx)."""
    h, w = ex.shape
    tx, ty = (w + 63) // 64, (h + 63) // 64
    X, Y = sample_positions(tx, ty, salt)
    inside = (X < w) & (Y < h)
    Xc, Yc = np.minimum(X, w - 1), np.minimum(Y, h - 1)
    d = np.abs(sur[Yc, Xc].astype(np.int32) - ex[Yc, Xc].astype(np.int32))
    bad = np.any(inside & (d > tol), axis=1)
    return bad, tx


def apply_fallback(sur, ex, bad, tx):
    """The surrogate output with every flagged tile replaced by the exact one."""
    out = sur.copy()
    for tile in np.flatnonzero(bad):
        X0, Y0 = tile % tx * 64, tile // tx * 64
        out[Y0:Y0 + 64, X0:X0 + 64] = ex[Y0:Y0 + 64, X0:X0 + 64]
    return out
