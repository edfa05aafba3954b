"""The radix-2^2 transform schedule of the per-stage launches, executed on Fr scalars in
place of G1 points, gives exactly the DFTs the G1 transforms must compute: the full FFT and
unscaled IFFT of the drop-ins and the IFFT -> zero upper half -> FFT of the FK20 path, for
every size 2^3 .. 2^9.  Butterfly positions, butterfly twiddles and the point-wise twiddle
layers come from the kernels' own index maps (schedule.hip.h, compiled by g++ through
tests/native/arith_host.cpp), so drift in the device functions fails here on the CPU.  The
references are the oracle's Fr transforms (pinned by the reference's fr_fft test vector).
The work model's Python copies of the same maps (fk20_amd.workload) are checked against the
compiled ones."""
import ctypes
import random

import numpy as np
import pytest

import oracle as O
from fk20_amd import workload as W

R = O.R_MOD
W512 = pow(5, (R - 1) // 512, R)


class Maps:
    def __init__(self, lib):
        self.lib = lib
        self.out = (ctypes.c_uint * 3)()

    def butterfly(self, u, L, b, r22=True):
        """(p0, p1, twiddle exponent in units of w512) of butterfly b of local stage u."""
        self.lib.host_stage_positions(int(r22), u, L, b, self.out)
        return self.out[0], self.out[1], self.out[2]

    def layer(self, sd, L):
        """[(position, exponent)] of the twiddle layer of odd DIF stage sd."""
        n = self.lib.host_layer_count(sd, L)
        res = []
        for k in range(n):
            self.lib.host_layer_twiddle(sd, L, k, self.out)
            res.append((self.out[0], self.out[1]))
        return res


@pytest.fixture(scope="module")
def M(arith_host):
    arith_host.host_layer_count.restype = ctypes.c_uint
    return Maps(arith_host)


def tw(v, e, inverse):
    return v * pow(W512, (512 - e) % 512 if inverse else e, R) % R


def dit(M, D, L, u):
    sd = L - 1 - u
    if u + 1 < L and sd % 2 == 1:  # the transposed DIF stage's twiddle layer first
        for q, e in M.layer(sd, L):
            D[q] = tw(D[q], e, True)
    for b in range(1 << (L - 1)):
        p0, p1, idx = M.butterfly(u, L, b)
        x = tw(D[p1], idx, True) if idx else D[p1]
        D[p0], D[p1] = (D[p0] + x) % R, (D[p0] - x) % R


def dif(M, D, L, u):
    for b in range(1 << (L - 1)):
        p0, p1, idx = M.butterfly(u, L, b)
        a, c = D[p0], D[p1]
        D[p0] = (a + c) % R
        D[p1] = tw((a - c) % R, idx, False) if idx else (a - c) % R
    s = u - (L - 1)
    if s % 2 == 1:  # then its own layer
        for q, e in M.layer(s, L):
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
def test_radix22_schedule_is_the_dft(M, L):
    n = 1 << L
    rng = random.Random(L)
    x = [rng.randrange(R) for _ in range(n)]
    fwd, inv = fr_transform(x, False), fr_transform(x, True)
    # drop-in FFT: DIF stages 0 .. L-1 (stage 0 through the middle's positions), bit-reversed out
    D = list(x)
    for u in range(L - 1, 2 * L - 1):
        dif(M, D, L, u)
    assert [D[bitrev(k, L)] for k in range(n)] == fwd
    # drop-in IFFT: bit-reversed in, DIT stages 0 .. L-1 (stage L-1 through the middle's positions)
    D = [0] * n
    for j in range(n):
        D[bitrev(j, L)] = x[j]
    for u in range(L):
        dit(M, D, L, u)
    assert D == inv
    # FK20: IFFT, upper half of h dropped, FFT; the middle fuses the last IFFT and first FFT stage
    D = [0] * n
    for j in range(n):
        D[bitrev(j, L)] = x[j]
    for u in range(L - 1):
        dit(M, D, L, u)
    for b in range(n // 2):
        p0, p1, idx = M.butterfly(L - 1, L, b)
        o = tw(D[p1], idx, True) if idx else D[p1]
        e = (D[p0] + o) % R
        D[p0], D[p1] = e, (tw(e, idx, False) if idx else e)
    for u in range(L, 2 * L - 1):
        dif(M, D, L, u)
    h = inv[: n // 2] + [0] * (n // 2)
    assert [D[bitrev(k, L)] for k in range(n)] == fr_transform(h, False)


@pytest.mark.parametrize("L", [3, 6, 9])
def test_radix2_schedule_is_the_dft(M, L):
    """The persistent chain's radix-2 map (every butterfly carries its full twiddle)."""
    n = 1 << L
    rng = random.Random(100 + L)
    x = [rng.randrange(R) for _ in range(n)]
    D = list(x)
    for u in range(L - 1, 2 * L - 1):
        for b in range(n // 2):
            p0, p1, idx = M.butterfly(u, L, b, r22=False)
            a, c = D[p0], D[p1]
            D[p0], D[p1] = (a + c) % R, tw((a - c) % R, idx, False)
    assert [D[bitrev(k, L)] for k in range(n)] == fr_transform(x, False)


def test_radix22_twiddle_counts(M):
    """Generic (non-w^+-128) twiddles per transform of 512: 1 196 in the layers against 1 538
    in the radix-2 schedule's butterflies (the FFT's nine DIF stages: the middle's map is the
    first)."""
    L = 9
    layer = [e for sd in range(1, L, 2) for _, e in M.layer(sd, L)]
    generic = sum(1 for e in layer if e % 256 != 128)
    assert generic == 1196
    radix2 = sum(1 for u in range(L - 1, 2 * L - 1) for b in range(256)
                 if M.butterfly(u, L, b, r22=False)[2] not in (0, 128))
    assert radix2 == 1538


@pytest.mark.parametrize("L", range(3, 10))
def test_stage_maps_are_bijections_in_bounds(M, L):
    """Every stage's butterflies (including the renumbered w^+-128 stages, whose twiddled half
    comes second in launch order) cover each position exactly once, in bounds; the layers
    touch distinct positions below n2."""
    n2 = 1 << L
    for u in range(2 * L - 1):
        seen = []
        for b in range(n2 // 2):
            p0, p1, idx = M.butterfly(u, L, b)
            assert 0 <= p0 < p1 < n2 and idx in (0, 128)
            seen += [p0, p1]
            if idx and u + 1 != L:
                assert b >= n2 // 4  # the twiddled half of a renumbered stage comes second
        assert sorted(seen) == list(range(n2))
    for sd in range(1, L, 2):
        qs = [q for q, _ in M.layer(sd, L)]
        assert len(set(qs)) == len(qs) and all(0 < q < n2 for q in qs)


@pytest.mark.parametrize("L", range(3, 10))
def test_work_model_copies_match_the_kernels(M, L):
    """fk20_amd.workload's Python maps (used for the executed-work figures) equal the
    compiled ones."""
    n2 = 1 << L
    for u in range(2 * L - 1):
        for b in range(n2 // 2):
            assert W.r22_butterfly_idx(u, L, b) == M.butterfly(u, L, b)[2]
    for sd in range(1, L, 2):
        assert W.r22_layer(sd, L) == M.layer(sd, L)
