"""GPU parity of every libfk20 entry point against the oracle and the reference KATs.

All calls go through the C ABI (libfk20.so via ctypes).  G1 results are compared as
canonical affine points (bit-exact), Fr results as canonical integers (bit-exact).
"""
import numpy as np
import pytest

import oracle as O
from conftest import hex_fp, hex_fr, hex_g1p, load_golden
from fk20_amd import ops
from fk20_amd import synthetic as S

pytestmark = pytest.mark.gpu

R = O.R_MOD
P = O.P_MOD


def aff(pts):
    """Canonical affine points computed by the oracle (independent of the GPU)."""
    return O.to_affine(np.ascontiguousarray(pts, dtype=np.uint64).reshape(-1, 18))


def assert_same_points(got, exp):
    a, b = aff(got), aff(exp)
    bad = np.nonzero(np.any(a != b, axis=1))[0]
    assert bad.size == 0, "mismatch at %s" % bad[:10].tolist()


def ints(a):
    return [O.limbs_to_int(x) for x in np.asarray(a).reshape(-1, np.asarray(a).shape[-1])]


# ---------------------------------------------------------------- arithmetic KATs

@pytest.mark.parametrize("field", ["fp", "fr"])
def test_field_kats_gpu(field):
    kat = load_golden(field + "_kat.json")["ops"]
    mod, conv, op = (P, hex_fp, ops.fp_op) if field == "fp" else (R, hex_fr, ops.fr_op)
    names = {"reduce6": "reduce", "reduce4": "reduce", "neg": "neg", "sqr": "sqr", "inv": "inv"}
    checked = 0
    for name, arrs in kat.items():
        base = name.split("_", 1)[-1]
        if base in names:
            q = np.array([conv(h) for h in arrs["q"]])
            got = ints(op(names[base], q))
            assert got == [int(h, 16) % mod for h in arrs["a"]], name
            checked += len(got)
        elif base in ("add", "sub", "mul"):
            x = np.array([conv(h) for h in arrs["x"]])
            y = np.array([conv(h) for h in arrs["y"]])
            got = ints(op(base, x, y))
            assert got == [int(h, 16) % mod for h in arrs["a"]], name
            checked += len(got)
        elif base in ("x2", "x3"):
            q = np.array([conv(h) for h in arrs["q"]])
            m = np.array([conv(hex(int(base[1])))] * len(q))
            got = ints(op("mul", q, m))
            assert got == [int(h, 16) % mod for h in arrs["a"]], name
            checked += len(got)
    assert checked > 60


def test_fp_mul_random_gpu():
    rng = np.random.default_rng(3)
    xs = [int.from_bytes(rng.bytes(48), "little") for _ in range(256)]
    ys = [int.from_bytes(rng.bytes(48), "little") for _ in range(256)]
    # edge values: 0, 1, p-1, p, 2^384-1
    edges = [0, 1, P - 1, P, 2**384 - 1]
    xs[: len(edges)] = edges
    ys[: len(edges)] = edges[::-1]
    a = np.array([O.fp(x) for x in xs])
    b = np.array([O.fp(y) for y in ys])
    for name, fn in [("mul", lambda x, y: x * y), ("add", lambda x, y: x + y), ("sub", lambda x, y: x - y)]:
        got = ints(ops.fp_op(name, a, b))
        assert got == [fn(x, y) % P for x, y in zip(xs, ys)], name
    got = ints(ops.fp_op("sqr", a))
    assert got == [x * x % P for x in xs]


def test_g1_kats_gpu():
    k = load_golden("g1_kat.json")
    pt = lambda name: hex_g1p(k[name])
    G = O.g1p_gen()
    x0, x2, x3, x24, x25 = (pt(n) for n in ["g1p_x0", "g1p_x2", "g1p_x3", "g1p_x24", "g1p_x25"])
    assert_same_points(ops.g1p_op("dbl", G[None]), x2[None])
    assert_same_points(ops.g1p_op("add", np.array([G, G, x0, G]), np.array([G, x2, x0, x0])),
                       np.array([x2, x3, O.g1p_inf(), G]))
    p = x3[None]
    for _ in range(3):
        p = ops.g1p_op("dbl", p)
    assert_same_points(p, x24[None])
    assert_same_points(ops.g1p_op("sub", x24[None], x2[None]), O.g1p_op("mul", G, O.fr(22))[None])
    ks = np.array([O.fr(25), O.fr(25 + R), O.fr(R), O.fr(R - 1)])
    got = ops.g1p_op("mul", np.array([G] * 4), k=ks)
    assert_same_points(got, np.array([x25, x25, O.g1p_inf(), O.g1p_op("neg", G)]))
    w = O.fr_root_of_unity(512)
    assert_same_points(ops.g1p_op("mul", G[None], k=w[None]), pt("g1p_wG")[None])
    # mixed addition with affine operands, including affine infinity (0, 0)
    a = O.to_affine(np.array([G, x2]))
    a_inf = np.zeros(12, dtype=np.uint64)
    qa = np.stack([a[0], a[1], a_inf])  # (3, 12): g1a_t stride
    got = ops.g1p_op("add_affine", np.array([G, x3, x25]), qa)
    assert_same_points(got, np.array([x2, O.g1p_op("mul", G, O.fr(5)), x25]))
    # on-curve and equality predicates
    bad = x2.copy()
    bad[0] ^= np.uint64(1)
    isp = ops.g1p_op("is_point", np.array([G, x0, bad, x25]))
    assert isp[:, 0].tolist() == [1, 1, 0, 1]
    eq = ops.g1p_op("eq", np.array([x25, x25, x0]), np.array([O.g1p_op("mul", G, O.fr(25)), x24, O.g1p_inf()]))
    assert eq[:, 0].tolist() == [1, 0, 1]


def test_to_affine_gpu():
    G = O.g1p_gen()
    pts = np.array([O.g1p_op("mul", G, O.fr(i + 3)) for i in range(8)] + [O.g1p_inf()])
    # non-normalised projective representatives: scale by z = 7
    scaled = pts.copy()
    for i in range(8):
        for c in range(3):
            v = O.limbs_to_int(pts[i, 6 * c : 6 * c + 6]) * 7 % P
            scaled[i, 6 * c : 6 * c + 6] = O.fp(v)
    got = ops.to_affine(scaled)
    assert np.array_equal(got, aff(pts))


# ---------------------------------------------------------------- transforms

def test_fr_fft_kat_gpu():
    k = load_golden("fr_fft_kat.json")
    q = np.array([hex_fr(h) for h in k["q"]])[None]
    a = np.array([hex_fr(h) for h in k["a"]])[None]
    assert ints(ops.fr_fft(q)) == [int(h, 16) % R for h in k["a"]]
    assert ints(ops.fr_fft(a, inverse=True)) == [int(h, 16) % R for h in k["q"]]
    assert ints(ops.fr_fft(q, inplace=True)) == [int(h, 16) % R for h in k["a"]]


@pytest.mark.parametrize("n", [128, 256, 512])
def test_fr_fft_random_gpu(n):
    x = S.random_fr(3 * n, 40 + n).reshape(3, n, 4)
    got = ops.fr_fft(x)
    for c in range(3):
        assert np.array_equal(got[c], O.fr_fft(x[c]))
    back = ops.fr_fft(got, inverse=True)
    assert np.array_equal(back, x)


def test_g1_fft_kat_gpu():
    k = load_golden("g1p_fft_kat.json")
    gin = np.array([hex_g1p(p) for p in k["input"]])[None]
    gout = np.array([hex_g1p(p) for p in k["output"]])[None]
    assert_same_points(ops.g1p_fft(gin), gout)
    assert_same_points(ops.g1p_fft(gout, inverse=True), gin)


@pytest.mark.parametrize("n", [128, 256])
def test_g1_fft_random_gpu(n):
    G = O.g1p_gen()
    xs = S.random_fr(n, 70 + n)
    pts = np.array([O.g1p_op("mul", G, x) for x in xs])[None]
    got = ops.g1p_fft(pts)
    # FFT([x]G) = [FFT(x)]G, the scalar side computed by the oracle
    exp = np.array([O.g1p_op("mul", G, s) for s in O.fr_fft(xs)])
    assert_same_points(got[0], exp)
    assert_same_points(ops.g1p_fft(got, inverse=True), pts)


# ---------------------------------------------------------------- FK20 stages

@pytest.fixture(scope="module")
def case1024():
    N = 1024
    tau = S.random_secret(21)
    poly = S.random_polys(2, N, 22)
    xext = O.xext_fft_from_tau(tau, N)
    return N, tau, poly, xext


@pytest.mark.parametrize("N", [1024, 2048, 4096])
def test_toeplitz_gpu(N):
    poly = S.random_polys(2, N, N)
    got = ops.poly2toeplitz_coefficients(poly)
    for r in range(2):
        assert np.array_equal(got[r], O.poly2toeplitz_coefficients(poly[r]))
    got_fft = ops.poly2toeplitz_coefficients_fft(poly)
    for r in range(2):
        tc = O.poly2toeplitz_coefficients(poly[r])
        assert np.array_equal(got_fft[r], np.array([O.fr_fft(tc[i]) for i in range(16)]))


def test_msm_gpu(case1024):
    N, _, poly, xext = case1024
    n2 = N // 8
    tc_fft = np.array([[O.fr_fft(t) for t in O.poly2toeplitz_coefficients(p)] for p in poly])
    got = ops.msm(tc_fft, xext)
    for r in range(2):
        assert_same_points(got[r], O.msm(tc_fft[r], xext))


def test_hext_fft2h_fft_gpu(case1024):
    N, _, _, xext = case1024
    hext = xext[:2]  # any G1 vectors
    got = ops.hext_fft2h_fft(hext)
    for r in range(2):
        assert_same_points(got[r], O.hext_fft2h_fft(hext[r]))
    # the split path: IFFT, hext2h, FFT
    h = ops.g1p_fft(hext, inverse=True)
    h = ops.hext2h(h)
    assert_same_points(ops.g1p_fft(h), got)


def test_poly2h_fft_gpu_1024(case1024):
    N, tau, poly, xext = case1024
    got = ops.poly2h_fft(poly, xext)
    exp = O.poly2h_fft(poly, xext)
    assert_same_points(got, exp)


def test_poly2h_fft_gpu_4096_closed_form():
    """N = 4096 row against the scalar-shadow closed form h_fft[k] = [s_k] G."""
    N = 4096
    tau = S.random_secret(31)
    poly = S.random_polys(1, N, 32)
    xext = O.xext_fft_from_tau(tau, N)
    got = ops.poly2h_fft(poly, xext)[0]
    s = O.fk20_shadow(poly[0], tau)
    G = O.g1p_gen()
    exp = np.array([O.g1p_op("mul", G, x) for x in s])
    assert_same_points(got, exp)


def test_setup2xext_fft_gpu():
    N = 1024
    tau = S.random_secret(41)
    srs = O.setup_srs(tau, N + 1)
    got = ops.setup2xext_fft(srs, N)
    assert_same_points(got, O.xext_fft_from_tau(tau, N))


def test_zero_rows_gpu(case1024):
    N, _, _, xext = case1024
    got = ops.poly2h_fft(np.zeros((1, N, 4), dtype=np.uint64), xext)
    assert all(O.g1p_is_inf(p) for p in got[0])


@pytest.mark.parametrize("limbs,rbits", [(29, 406), (32, 384)])
def test_fp_mul_probe_layouts_exact(limbs, rbits):
    """Both limb layouts of fk20_fp_mul_probe (the measurement behind the library's choice of
    14 x 29-bit limbs) compute exact Montgomery products: every lane of one block against
    Python integers."""
    from fk20_amd.lib import DeviceBuffer, call, synchronize
    P = O.P_MOD
    rng = np.random.default_rng(limbs)
    xs = [int.from_bytes(rng.bytes(48), "little") % P for _ in range(1024)]
    xs[1] = P - 1
    y = int.from_bytes(rng.bytes(48), "little") % P

    def words(v):
        return np.array([(v >> (32 * i)) & 0xFFFFFFFF for i in range(12)], dtype=np.uint32)

    xb = DeviceBuffer.from_array(np.stack([words(x) for x in xs]))
    yb = DeviceBuffer.from_array(words(y))
    out = DeviceBuffer(256 * 48)
    iters = 5
    call("fk20_fp_mul_probe", limbs, out.ptr, xb.ptr, yb.ptr, 1, iters, None)
    synchronize()
    got = out.download((256, 12), np.uint32)
    for b in (xb, yb, out):
        b.free()
    rinv = pow(2, -rbits, P)
    for lane in range(256):
        a = xs[lane]
        c1 = a * pow(y * rinv, iters, P) % P
        c2 = a * y * rinv % P * pow(y * rinv, iters, P) % P
        v = sum(int(w) << (32 * i) for i, w in enumerate(got[lane]))
        assert v % P == c1 * c2 * rinv % P and v < 2 * P, lane
