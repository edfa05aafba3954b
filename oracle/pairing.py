"""BLS12-381 optimal ate pairing and the KZG multi-proof check, in pure Python.

TEST INFRASTRUCTURE ONLY (the oracle): used by tests/ to verify proofs end to end, never by
the product.  Restates FK20Py's verifier, kzg_proofs.py:207-232 (check_proof_multi: interpolate
the claimed evaluations on the coset, then
    e(commitment - [I(s)]_1, [1]_2) == e(proof, [s^n - x^n]_2)
checked as a product of two pairings equal to one), with py_ecc's pairing replaced by this
module's (py_ecc is not available).  Arithmetic on Python integers; a check costs ~1-2 s.

Tower: Fp2 = Fp[u]/(u^2 + 1), Fp6 = Fp2[v]/(v^3 - (u + 1)), Fp12 = Fp6[w]/(w^2 - v).
G2 lives on the twist E': y^2 = x^3 + 4 (u + 1); psi(x, y) = (x w^-2, y w^-3) maps it to
E(Fp12): y^2 = x^3 + 4.  Miller loop over |z| = 0xd201000000010000 (z < 0, so the result is
conjugated), lines scaled by w^3 (an Fp4 element, removed by the final exponentiation),
final exponentiation by (p^12 - 1) / r directly.
"""

P = 0x1A0111EA397FE69A4B1BA7B6434BACD764774B84F38512BF6730D2A0F6B0F6241EABFFFEB153FFFFB9FEFFFFFFFFAAAB
R = 0x73EDA753299D7D483339D80809A1D80553BDA402FFFE5BFEFFFFFFFF00000001
Z_ABS = 0xD201000000010000  # |z|, z = -0xd201000000010000
FINAL_EXP = (P ** 12 - 1) // R

# ---------------------------------------------------------------- Fp2 = (a, b) = a + b u

def f2(a, b=0):
    return (a % P, b % P)


F2_ZERO, F2_ONE = (0, 0), (1, 0)


def f2_add(x, y):
    return ((x[0] + y[0]) % P, (x[1] + y[1]) % P)


def f2_sub(x, y):
    return ((x[0] - y[0]) % P, (x[1] - y[1]) % P)


def f2_neg(x):
    return ((-x[0]) % P, (-x[1]) % P)


def f2_mul(x, y):
    return ((x[0] * y[0] - x[1] * y[1]) % P, (x[0] * y[1] + x[1] * y[0]) % P)


def f2_scale(x, k):
    return (x[0] * k % P, x[1] * k % P)


def f2_inv(x):
    n = pow(x[0] * x[0] + x[1] * x[1], P - 2, P)
    return (x[0] * n % P, (-x[1]) * n % P)


def f2_mul_xi(x):
    """x * (u + 1)"""
    return ((x[0] - x[1]) % P, (x[0] + x[1]) % P)


# ---------------------------------------------------------------- Fp6 = (c0, c1, c2) over v

F6_ZERO = (F2_ZERO, F2_ZERO, F2_ZERO)
F6_ONE = (F2_ONE, F2_ZERO, F2_ZERO)


def f6_add(x, y):
    return tuple(f2_add(a, b) for a, b in zip(x, y))


def f6_sub(x, y):
    return tuple(f2_sub(a, b) for a, b in zip(x, y))


def f6_neg(x):
    return tuple(f2_neg(a) for a in x)


def f6_mul(x, y):
    t = [F2_ZERO] * 5
    for i in range(3):
        for j in range(3):
            t[i + j] = f2_add(t[i + j], f2_mul(x[i], y[j]))
    # v^3 = xi
    return (f2_add(t[0], f2_mul_xi(t[3])), f2_add(t[1], f2_mul_xi(t[4])), t[2])


def f6_mul_v(x):
    """x * v"""
    return (f2_mul_xi(x[2]), x[0], x[1])


def f6_inv(x):
    a, b, c = x
    # standard cubic-extension inverse
    A = f2_sub(f2_mul(a, a), f2_mul_xi(f2_mul(b, c)))
    B = f2_sub(f2_mul_xi(f2_mul(c, c)), f2_mul(a, b))
    C = f2_sub(f2_mul(b, b), f2_mul(a, c))
    F = f2_add(f2_mul(a, A), f2_mul_xi(f2_add(f2_mul(c, B), f2_mul(b, C))))
    Fi = f2_inv(F)
    return (f2_mul(A, Fi), f2_mul(B, Fi), f2_mul(C, Fi))


# ---------------------------------------------------------------- Fp12 = (a, b) = a + b w

F12_ONE = (F6_ONE, F6_ZERO)


def f12_mul(x, y):
    a, b = x
    c, d = y
    # (a + b w)(c + d w) = (ac + bd v) + (ad + bc) w
    return (f6_add(f6_mul(a, c), f6_mul_v(f6_mul(b, d))), f6_add(f6_mul(a, d), f6_mul(b, c)))


def f12_sqr(x):
    return f12_mul(x, x)


def f12_conj(x):
    return (x[0], f6_neg(x[1]))


def f12_inv(x):
    a, b = x
    # (a + b w)^-1 = (a - b w) / (a^2 - b^2 v)
    n = f6_inv(f6_sub(f6_mul(a, a), f6_mul_v(f6_mul(b, b))))
    return (f6_mul(a, n), f6_neg(f6_mul(b, n)))


def f12_pow(x, e):
    r = F12_ONE
    for bit in bin(e)[2:]:
        r = f12_sqr(r)
        if bit == "1":
            r = f12_mul(r, x)
    return r


# ---------------------------------------------------------------- G2 on the twist (affine, None = infinity)

B2 = f2(4, 4)  # 4 (u + 1)
G2_GEN = (f2(0x024AA2B2F08F0A91260805272DC51051C6E47AD4FA403B02B4510B647AE3D1770BAC0326A805BBEFD48056C8C121BDB8,
             0x13E02B6052719F607DACD3A088274F65596BD0D09920B61AB5DA61BBDC7F5049334CF11213945D57E5AC7D055D042B7E),
          f2(0x0CE5D527727D6E118CC9CDC6DA2E351AADFD9BAA8CBDD3A76D429A695160D12C923AC9CC3BACA289E193548608B82801,
             0x0606C4A02EA734CC32ACD2B02BC28B99CB3E287E85A763AF267492AB572E99AB3F370D275CEC1DA1AAA9075FF05F79BE))


def g2_on_curve(Q):
    if Q is None:
        return True
    x, y = Q
    return f2_sub(f2_mul(y, y), f2_add(f2_mul(f2_mul(x, x), x), B2)) == F2_ZERO


def g2_neg(Q):
    return None if Q is None else (Q[0], f2_neg(Q[1]))


def g2_add(A, B):
    if A is None:
        return B
    if B is None:
        return A
    if A[0] == B[0]:
        if A[1] == B[1] and A[1] != F2_ZERO:
            lam = f2_mul(f2_scale(f2_mul(A[0], A[0]), 3), f2_inv(f2_scale(A[1], 2)))
        else:
            return None
    else:
        lam = f2_mul(f2_sub(B[1], A[1]), f2_inv(f2_sub(B[0], A[0])))
    x = f2_sub(f2_sub(f2_mul(lam, lam), A[0]), B[0])
    return (x, f2_sub(f2_mul(lam, f2_sub(A[0], x)), A[1]))


def g2_mul(Q, k):
    k %= R
    acc = None
    for bit in bin(k)[2:] if k else "":
        acc = g2_add(acc, acc)
        if bit == "1":
            acc = g2_add(acc, Q)
    return acc


# ---------------------------------------------------------------- pairing

def _line(lam, T, xP, yP):
    """w^3 * l(P) for the line of slope lam (on the twist) through T, P = (xP, yP) in E(Fp):
    (lam x_T - y_T) + (-lam xP) v + yP v w."""
    a0 = f2_sub(f2_mul(lam, T[0]), T[1])
    a1 = f2_neg(f2_scale(lam, xP))
    return ((a0, a1, F2_ZERO), (F2_ZERO, f2(yP), F2_ZERO))


def miller_loop(Pa, Q):
    """f_{|z|, Q}(P), conjugated for z < 0.  Pa = (x, y) affine integers or None, Q on E'."""
    if Pa is None or Q is None:
        return F12_ONE
    xP, yP = Pa
    f, T = F12_ONE, Q
    for bit in bin(Z_ABS)[3:]:
        lam = f2_mul(f2_scale(f2_mul(T[0], T[0]), 3), f2_inv(f2_scale(T[1], 2)))
        f = f12_mul(f12_sqr(f), _line(lam, T, xP, yP))
        T = g2_add(T, T)
        if bit == "1":
            lam = f2_mul(f2_sub(Q[1], T[1]), f2_inv(f2_sub(Q[0], T[0])))
            f = f12_mul(f, _line(lam, T, xP, yP))
            T = g2_add(T, Q)
    return f12_conj(f)


def final_exponentiation(f):
    return f12_pow(f, FINAL_EXP)


def pairing(Pa, Q):
    return final_exponentiation(miller_loop(Pa, Q))


def pairing_product_is_one(pairs):
    """prod e(P_i, Q_i) == 1 with one final exponentiation."""
    f = F12_ONE
    for Pa, Q in pairs:
        f = f12_mul(f, miller_loop(Pa, Q))
    return final_exponentiation(f) == F12_ONE


# ---------------------------------------------------------------- KZG multi-proof check

def _inv(a):
    return pow(a, R - 2, R)


def _fft_inverse(vals, root):
    """Natural-order inverse DFT over Fr by the O(n^2) definition (n = 16 here)."""
    n = len(vals)
    rinv = _inv(root)
    ninv = _inv(n)
    return [sum(v * pow(rinv, i * j, R) for j, v in enumerate(vals)) * ninv % R for i in range(n)]


def check_proof_multi(commitment, proof, x, ys, srs_g1_affine, g2_s_n, g1_affine_add, g1_affine_mul):
    """kzg_proofs.py:207-232 with affine G1 points as (x, y) integer pairs (None = infinity):
    commitment, proof in G1; x the coset shift; ys the claimed f(x w^i), w of order n = len(ys);
    srs_g1_affine[i] = [s^i]_1 for i < n; g2_s_n = [s^n]_2.  The G1 arithmetic is supplied
    by the caller (the C oracle's), the pairing is this module's."""
    n = len(ys)
    root = pow(5, (R - 1) // n, R)
    coeffs = _fft_inverse([y % R for y in ys], root)
    coeffs = [c * _inv(pow(x, i, R)) % R for i, c in enumerate(coeffs)]
    interp = None
    for i, c in enumerate(coeffs):
        interp = g1_affine_add(interp, g1_affine_mul(srs_g1_affine[i], c))
    c_minus_i = g1_affine_add(commitment, None if interp is None else (interp[0], (-interp[1]) % P))
    xn = pow(x, n, R)
    s_n_minus_xn = g2_add(g2_s_n, g2_neg(g2_mul(G2_GEN, xn)))
    neg_c = None if c_minus_i is None else (c_minus_i[0], (-c_minus_i[1]) % P)
    return pairing_product_is_one([(neg_c, G2_GEN), (proof, s_n_minus_xn)])
