"""Seeded synthetic inputs for the FK20 path (uniform Fr coefficients, setup secrets).

Coefficients are uniform modulo r by rejection sampling of 255-bit values drawn from
numpy's PCG64 (r < 2^255, so about 10% of draws are rejected).  The layout is the
reference ABI: fr_t = 4 little-endian uint64 limbs, standard form (FK20-CUDA fr.cuh:24).
"""
import numpy as np

R_MOD = 0x73EDA753299D7D483339D80809A1D80553BDA402FFFE5BFEFFFFFFFF00000001
_R_LIMBS = np.array([(R_MOD >> (64 * i)) & 0xFFFFFFFFFFFFFFFF for i in range(4)], dtype=np.uint64)


def _less_than_r(x):
    """Row-wise x < r for x of shape (n, 4)."""
    lt = np.zeros(x.shape[0], dtype=bool)
    eq = np.ones(x.shape[0], dtype=bool)
    for i in range(3, -1, -1):
        lt |= eq & (x[:, i] < _R_LIMBS[i])
        eq &= x[:, i] == _R_LIMBS[i]
    return lt


def random_fr(count, seed):
    """`count` uniform Fr elements as a (count, 4) uint64 array."""
    rng = np.random.Generator(np.random.PCG64(seed))
    out = np.empty((count, 4), dtype=np.uint64)
    filled = 0
    while filled < count:
        need = count - filled
        draw = rng.integers(0, 2**64, size=(need + need // 8 + 16, 4), dtype=np.uint64)
        draw[:, 3] &= np.uint64(0x7FFFFFFFFFFFFFFF)
        good = draw[_less_than_r(draw)][:need]
        out[filled : filled + good.shape[0]] = good
        filled += good.shape[0]
    return out


P_MOD = 0x1A0111EA397FE69A4B1BA7B6434BACD764774B84F38512BF6730D2A0F6B0F6241EABFFFEB153FFFFB9FEFFFFFFFFAAAB
_P_LIMBS = np.array([(P_MOD >> (64 * i)) & 0xFFFFFFFFFFFFFFFF for i in range(6)], dtype=np.uint64)


def random_fp_nonzero(count, seed):
    """`count` uniform non-zero Fp elements as a (count, 6) uint64 array (rejection sampling
    of 381-bit draws)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    out = np.empty((count, 6), dtype=np.uint64)
    filled = 0
    while filled < count:
        need = count - filled
        draw = rng.integers(0, 2**64, size=(need + need // 4 + 16, 6), dtype=np.uint64)
        draw[:, 5] &= np.uint64((1 << 61) - 1)
        lt = np.zeros(draw.shape[0], dtype=bool)
        eq = np.ones(draw.shape[0], dtype=bool)
        for i in range(5, -1, -1):
            lt |= eq & (draw[:, i] < _P_LIMBS[i])
            eq &= draw[:, i] == _P_LIMBS[i]
        good = draw[lt & draw.any(axis=1)][:need]
        out[filled : filled + good.shape[0]] = good
        filled += good.shape[0]
    return out


def random_polys(rows, N, seed):
    """`rows` polynomials of N uniform Fr coefficients: (rows, N, 4)."""
    return random_fr(rows * N, seed).reshape(rows, N, 4)


def random_secret(seed):
    """A setup secret tau, uniform in [1, r)."""
    while True:
        t = random_fr(1, seed)[0]
        v = sum(int(w) << (64 * i) for i, w in enumerate(t.tolist()))
        if v:
            return v
        seed += 1


def fr_to_int(a):
    return sum(int(w) << (64 * i) for i, w in enumerate(np.asarray(a, dtype=np.uint64).tolist()))


def int_to_fr(v):
    return np.array([(v >> (64 * i)) & 0xFFFFFFFFFFFFFFFF for i in range(4)], dtype=np.uint64)
