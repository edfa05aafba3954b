"""The radix-2^2 transform schedule of the per-stage launches, executed on Fr scalars in
place of G1 points (the same butterfly positions, w^+-128 butterfly twiddles and point-wise
twiddle layers, from the work model's mirror of the kernels' index maps), gives exactly the
DFTs the G1 transforms must compute: the full FFT and unscaled IFFT of the drop-ins and the
IFFT -> zero upper half -> FFT of the FK20 path, for every size 2^3 .. 2^9.  The references
are the oracle's Fr transforms (pinned by the reference's fr_fft test vector)."""
import random

import numpy as np
import pytest

import oracle as O
from fk20_amd import workload as W

R = O.R_MOD
W512 = pow(5, (R - 1) // 512, R)


def tw(v, e, inverse):
    return v * pow(W512, (512 - e) % 512 if inverse else e, R) % R


def positions(u, L, b):
    """Butterfly b of local stage u (DIT 0..L-2, middle L-1, DIF L..2L-2): positions p0, p1."""
    if u + 1 < L:
        s = u
        j = b & ((1 << s) - 1)
        p0 = ((b >> s) << (s + 1)) + j
        return p0, p0 + (1 << s)
    if u + 1 == L:
        return b, b + (1 << (L - 1))
    lh = 2 * L - 2 - u
    j = b & ((1 << lh) - 1)
    p0 = ((b >> lh) << (lh + 1)) + j
    return p0, p0 + (1 << lh)


def dit(D, L, u):
    sd = L - 1 - u
    if u + 1 < L and sd % 2 == 1:  # the transposed DIF stage's twiddle layer first
        for q, e in W.r22_layer(sd, L):
            D[q] = tw(D[q], e, True)
    for b in range(1 << (L - 1)):
        p0, p1 = positions(u, L, b)
        idx = W.r22_butterfly_idx(u, L, b)
        x = tw(D[p1], idx, True) if idx else D[p1]
        D[p0], D[p1] = (D[p0] + x) % R, (D[p0] - x) % R


def dif(D, L, u):
    for b in range(1 << (L - 1)):
        p0, p1 = positions(u, L, b)
        idx = W.r22_butterfly_idx(u, L, b)
        a, c = D[p0], D[p1]
        D[p0] = (a + c) % R
        D[p1] = tw((a - c) % R, idx, False) if idx else (a - c) % R
    s = u - (L - 1)
    if s % 2 == 1:  # then its own layer
        for q, e in W.r22_layer(s, L):
            D[q] = tw(D[q], e, False)


def bitrev(x, L):
    return int(format(x, "0%db" % L)[::-1], 2)


def fr_transform(x, inverse):
    """Oracle DFT (inverse unscaled) as Python ints."""
    arr = np.array([O.fr(v) for v in x], dtype=np.uint64)
    out = [O.limbs_to_int(v) for v in O.fr_fft(arr, inverse=inverse)]
    n = len(x)
    return [v * n % R for v in out] if inverse else out


@pytest.mark.parametrize("L", range(3, 10))
def test_radix22_schedule_is_the_dft(L):
    n = 1 << L
    rng = random.Random(L)
    x = [rng.randrange(R) for _ in range(n)]
    fwd, inv = fr_transform(x, False), fr_transform(x, True)
    # drop-in FFT: DIF stages 0 .. L-1 (stage 0 through the middle's positions), bit-reversed out
    D = list(x)
    for u in range(L - 1, 2 * L - 1):
        dif(D, L, u)
    assert [D[bitrev(k, L)] for k in range(n)] == fwd
    # drop-in IFFT: bit-reversed in, DIT stages 0 .. L-1 (stage L-1 through the middle's positions)
    D = [0] * n
    for j in range(n):
        D[bitrev(j, L)] = x[j]
    for u in range(L):
        dit(D, L, u)
    assert D == inv
    # FK20: IFFT, upper half of h dropped, FFT; the middle fuses the last IFFT and first FFT stage
    D = [0] * n
    for j in range(n):
        D[bitrev(j, L)] = x[j]
    for u in range(L - 1):
        dit(D, L, u)
    for b in range(n // 2):
        p0, p1 = positions(L - 1, L, b)
        idx = W.r22_butterfly_idx(L - 1, L, b)
        o = tw(D[p1], idx, True) if idx else D[p1]
        e = (D[p0] + o) % R
        D[p0], D[p1] = e, (tw(e, idx, False) if idx else e)
    for u in range(L, 2 * L - 1):
        dif(D, L, u)
    h = inv[: n // 2] + [0] * (n // 2)
    assert [D[bitrev(k, L)] for k in range(n)] == fr_transform(h, False)


def test_radix22_twiddle_counts():
    """Generic (non-w^+-128) twiddles per transform of 512: 1 196 in the layers against 1 538
    in the radix-2 schedule's butterflies."""
    L = 9
    layer = [e for sd in range(1, L, 2) for _, e in W.r22_layer(sd, L)]
    generic = sum(1 for e in layer if e % 256 != 128)
    assert generic == 1196
    radix2 = sum(1 for s in range(1, L) for b in range(256)
                 if (b & ((1 << s) - 1)) << (8 - s) not in (0, 128))
    assert radix2 == 1538


def quarter_order(b, L, lh):
    """Mirror of fast_kernels.hip quarter_order: launch index b of a w^+-128 stage (span
    2^lh) -> (block, offset j), the untwiddled half of the launch first."""
    c, r = b >> (L - 2), b & ((1 << (L - 2)) - 1)
    return r >> (lh - 1), (r & ((1 << (lh - 1)) - 1)) + (c << (lh - 1))


@pytest.mark.parametrize("L", range(3, 10))
def test_quarter_order_is_a_bijection_in_bounds(L):
    """The kernels renumber the butterflies of every w^+-128 stage: each launch index maps to a
    distinct butterfly of the stage, both positions stay below n2, and the second half of the
    launch is exactly the twiddled half (offset j >= span / 2)."""
    n2 = 1 << L
    for lh in range(1, L):
        seen = set()
        for b in range(n2 // 2):
            blk, j = quarter_order(b, L, lh)
            p0 = (blk << (lh + 1)) + j
            assert 0 <= j < (1 << lh) and 0 <= p0 and p0 + (1 << lh) < n2
            assert (2 * j >= (1 << lh)) == (b >= n2 // 4)
            seen.add((blk, j))
        assert len(seen) == n2 // 2
