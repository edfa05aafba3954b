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


# BLS12-381 G1 generator as a g1p_t (x, y, z = 1), standard form (FK20-CUDA g1p.cu:95-115)
G1_GEN = np.zeros(18, dtype=np.uint64)
G1_GEN[:6] = [0xFB3AF00ADB22C6BB, 0x6C55E83FF97A1AEF, 0xA14E3A3F171BAC58,
              0xC3688C4F9774B905, 0x2695638C4FA9AC0F, 0x17F1D3A73197D794]
G1_GEN[6:12] = [0x0CAA232946C5E7E1, 0xD03CC744A2888AE4, 0x00DB18CB2C04B3ED,
                0xFCF5E095D5D00AF6, 0xA09E30ED741D8AE4, 0x08B3F481E3AAA0F1]
G1_GEN[12] = 1


def device_srs(count, tau):
    """SRS [tau^i]G, i < count, computed on the GPU (fk20_g1p_op mul): a DeviceBuffer of
    count g1p_t (FK20Py kzg_proofs.py:22-28 generate_setup, G1 half)."""
    from . import lib as L
    pw, acc = [], 1
    for _ in range(count):
        pw.append(int_to_fr(acc))
        acc = acc * tau % R_MOD
    gens = L.DeviceBuffer.from_array(np.tile(G1_GEN, (count, 1)))
    ks = L.DeviceBuffer.from_array(np.array(pw))
    srs = L.DeviceBuffer(count * 144)
    L.call("fk20_g1p_op", 3, srs.ptr, gens.ptr, None, ks.ptr, count, None)
    L.synchronize()
    gens.free()
    ks.free()
    return srs


def device_setup(N, tau):
    """xext_fft (16, N/8) of the secret tau on the GPU: SRS [tau^i]G (i <= N), then
    fk20_setup2xext_fft (FK20Py fk20_multi.py:82-85 toeplitz_part1).  A DeviceBuffer."""
    from . import lib as L
    srs = device_srs(N + 1, tau)
    xext = L.DeviceBuffer(16 * (N // 8) * 144)
    L.call("fk20_setup2xext_fft", xext.ptr, srs.ptr, N, None)
    L.synchronize()
    srs.free()
    return xext


def curve_points(count, seed):
    """`count` points of E(Fp): y^2 = x^3 + 4 from random x with x^3 + 4 a square (p = 3 mod 4,
    so y = (x^3 + 4)^((p+1)/4)), as standard-form g1p_t (count, 18) with z = 1.  Such points
    lie outside the order-r subgroup G1 unless their cofactor part vanishes (probability
    1/h, h ~ 2^126)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    out = np.zeros((count, 18), dtype=np.uint64)
    n = 0
    while n < count:
        x = int.from_bytes(rng.bytes(48), "little") % P_MOD
        a = (x * x * x + 4) % P_MOD
        y = pow(a, (P_MOD + 1) // 4, P_MOD)
        if y * y % P_MOD != a:
            continue
        for c, v in enumerate((x, y, 1)):
            for i in range(6):
                out[n, 6 * c + i] = (v >> (64 * i)) & 0xFFFFFFFFFFFFFFFF
        n += 1
    return out
