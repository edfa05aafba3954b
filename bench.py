#!/usr/bin/env python3
"""FK20 multi-proof benchmark on one MI355X: rows proved per second, N = 4096, 512 rows.

A step is one pass of the FK20 path over one batch of synthetic rows (uniform Fr
coefficients, one seeded setup): polynomial -> Toeplitz coefficients -> 16 Fr NTTs per
row -> MSM -> G1 IFFT -> zero upper half -> G1 FFT -> h_fft.  Inputs are resident in HBM
before timing starts.  Forms:
  plain  (P2) the reference's six-launch sequence with grid = rows (fk20_poly2h_fft.cu:47-80)
  P1          the same sequence issued once per row (rows = 1), timed on a bounded sample
The dominant kernel is timed with HIP events on its own stream inside the timed steps.
cpu_baseline: the oracle (CPU restatement of FK20Py fk20_multi) on the host cores over a
bounded sample of the same rows; its outputs also check the GPU rows bit-exactly.
"""
import argparse
import ctypes
import json
import os
import statistics
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "fk20-cuda_amd"))

import numpy as np  # noqa: E402

from fk20_amd import lib as L  # noqa: E402
from fk20_amd import synthetic as S  # noqa: E402
from fk20_amd import workload as WL  # noqa: E402

R_MOD = S.R_MOD
METRIC = "FK20 rows proved/sec (N=4096, 512 rows) batched vs plain; p50/p99 latency, bit-exact"


def log(msg):
    print("[bench] " + msg, file=sys.stderr, flush=True)


class Batch:
    """Device-resident inputs, setup and outputs of one batch."""

    def __init__(self, rows, N, seed):
        self.rows, self.N, self.n2 = rows, N, N // 8
        self.poly_host = S.random_polys(rows, N, seed)
        self.poly = L.DeviceBuffer.from_array(self.poly_host)
        self.h_fft = L.DeviceBuffer(rows * self.n2 * 144)
        self.tau = S.random_secret(seed + 1)
        self.xext = self._setup()

    def _setup(self):
        """SRS [tau^i]G on the GPU, then xext_fft by fk20_setup2xext_fft (FK20Py toeplitz_part1)."""
        N = self.N
        pw, acc = [], 1
        for _ in range(N + 1):
            pw.append(S.int_to_fr(acc))
            acc = acc * self.tau % R_MOD
        G = np.zeros(18, dtype=np.uint64)
        G[:6] = [0xFB3AF00ADB22C6BB, 0x6C55E83FF97A1AEF, 0xA14E3A3F171BAC58,
                 0xC3688C4F9774B905, 0x2695638C4FA9AC0F, 0x17F1D3A73197D794]
        G[6:12] = [0x0CAA232946C5E7E1, 0xD03CC744A2888AE4, 0x00DB18CB2C04B3ED,
                   0xFCF5E095D5D00AF6, 0xA09E30ED741D8AE4, 0x08B3F481E3AAA0F1]
        G[12] = 1
        gens = L.DeviceBuffer.from_array(np.tile(G, (N + 1, 1)))
        ks = L.DeviceBuffer.from_array(np.array(pw))
        srs = L.DeviceBuffer((N + 1) * 144)
        L.call("fk20_g1p_op", 3, srs.ptr, gens.ptr, None, ks.ptr, N + 1, None)
        xext = L.DeviceBuffer(16 * self.n2 * 144)
        L.call("fk20_setup2xext_fft", xext.ptr, srs.ptr, N, None)
        L.synchronize()
        self.xext_host = xext.download((16, self.n2, 18))
        for b in (gens, ks, srs):
            b.free()
        return xext


class Scratch:
    def __init__(self, rows, n2):
        self.tc = L.DeviceBuffer(rows * 16 * n2 * 32)
        self.h = L.DeviceBuffer(rows * n2 * 144)


def plain_step(b, sc, stream, ev):
    """The six-launch sequence (fk20_poly2h_fft.cu:57-77) with an event between stages."""
    s = stream.handle
    rows, N, n2 = b.rows, b.N, b.n2
    ev[0].record(stream)
    L.call("fk20_poly2toeplitz_coefficients", sc.tc.ptr, b.poly.ptr, rows, N, s)
    L.call("fr_fft", sc.tc.ptr, sc.tc.ptr, rows * 16, n2, s)
    ev[1].record(stream)
    L.call("fk20_msm", sc.h.ptr, sc.tc.ptr, b.xext.ptr, rows, n2, s)
    ev[2].record(stream)
    L.call("g1p_ift", sc.h.ptr, sc.h.ptr, rows, n2, s)
    ev[3].record(stream)
    L.call("fk20_hext2h", sc.h.ptr, rows, n2, s)
    L.call("g1p_fft", b.h_fft.ptr, sc.h.ptr, rows, n2, s)
    ev[4].record(stream)


def time_plain(b, steps, warmup):
    stream = L.Stream()
    sc = Scratch(b.rows, b.n2)
    ev = [L.Event() for _ in range(5)]
    for _ in range(warmup):
        plain_step(b, sc, stream, ev)
    stream.synchronize()
    step_ms, stage_ms = [], {"ntt": [], "msm": [], "ifft": [], "fft": []}
    L.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        plain_step(b, sc, stream, ev)
        # events of this step (waits for the step; the next step is issued after)
        stage_ms["ntt"].append(ev[0].elapsed_ms(ev[1]))
        stage_ms["msm"].append(ev[1].elapsed_ms(ev[2]))
        stage_ms["ifft"].append(ev[2].elapsed_ms(ev[3]))
        stage_ms["fft"].append(ev[3].elapsed_ms(ev[4]))
        step_ms.append(ev[0].elapsed_ms(ev[4]))
    stream.synchronize()
    L.synchronize()
    wall = time.perf_counter() - t0
    for e in ev:
        e.destroy()
    stream.destroy()
    return step_ms, stage_ms, wall


def time_p1(b, sample_rows):
    """Per-row plain launch sequence: fk20_poly2h_fft with rows = 1, row after row."""
    stream = L.Stream()
    e0, e1 = L.Event(), L.Event()
    row_bytes_in, row_bytes_out = b.N * 32, b.n2 * 144
    # warm-up row
    L.call("fk20_poly2h_fft", b.h_fft.ptr, b.poly.ptr, b.xext.ptr, 1, b.N, stream.handle)
    stream.synchronize()
    lat = []
    for r in range(sample_rows):
        e0.record(stream)
        L.call("fk20_poly2h_fft", b.h_fft.ptr + r * row_bytes_out, b.poly.ptr + r * row_bytes_in,
               b.xext.ptr, 1, b.N, stream.handle)
        e1.record(stream)
        lat.append(e0.elapsed_ms(e1))
    stream.synchronize()
    e0.destroy()
    e1.destroy()
    stream.destroy()
    return lat


def cpu_baseline(b, sample_rows, threads):
    """Oracle on host cores over the first sample_rows rows; checks those GPU rows bit-exactly."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O

    O.lib()
    t0 = time.perf_counter()
    exp = O.poly2h_fft(b.poly_host[:sample_rows], b.xext_host, nthreads=threads)
    el = time.perf_counter() - t0
    got = b.h_fft.download((b.rows, b.n2, 18))[:sample_rows]
    same = np.array_equal(O.to_affine(got.reshape(-1, 18)), O.to_affine(exp.reshape(-1, 18)))
    return {
        "value": sample_rows / el,
        "unit": "rows/s",
        "cores": threads,
        "kind": "port",
        "sample": "%d rows of N=%d from the timed batch, one oracle row per thread (%.1f s)"
                  % (sample_rows, b.N, el),
    }, same


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--rows", type=int, default=512)
    ap.add_argument("--N", type=int, default=4096)
    ap.add_argument("--seed", type=int, default=2024)
    ap.add_argument("--p1-rows", type=int, default=4, help="rows timed for the per-row form P1")
    ap.add_argument("--cpu-rows", type=int, default=16, help="rows in the CPU baseline sample")
    ap.add_argument("--cpu-threads", type=int, default=16)
    ap.add_argument("--no-cpu", action="store_true")
    args = ap.parse_args()

    if args.gpus != 1:
        print("bench.py runs on one GPU only: FK20 rows are independent and this build measures "
              "the single-GPU path (--gpus %d requested)" % args.gpus, file=sys.stderr)
        sys.exit(2)
    if L.device_count() < 1:
        sys.exit("no HIP device visible")

    log("generating %d rows, N=%d, setup on device" % (args.rows, args.N))
    b = Batch(args.rows, args.N, args.seed)
    log("timing plain batched form: %d warmup + %d steps" % (args.warmup, args.steps))
    step_ms, stage_ms, wall = time_plain(b, args.steps, args.warmup)
    ms = statistics.median(step_ms)
    value = args.rows / (ms / 1e3)
    log("plain: %.1f ms/step, %.1f rows/s; stages %s" % (
        ms, value, {k: round(statistics.median(v), 2) for k, v in stage_ms.items()}))

    log("timing P1 (per-row launches) on %d rows" % args.p1_rows)
    p1 = time_p1(b, args.p1_rows)
    p1_rate = 1e3 / statistics.median(p1)

    # roofline of the dominant kernel: the stage with the largest median time
    stage_med = {k: statistics.median(v) for k, v in stage_ms.items()}
    dom = max(stage_med, key=stage_med.get)
    work = WL.stage_mul32(args.N)[dom] * args.rows
    achieved = work / (stage_med[dom] / 1e3) / 1e12
    peak = WL.PEAK_MUL32_PER_S / 1e12
    roofline = {
        "bound": "valu",
        "kernel": {"msm": "k_msm_plain", "ifft": "k_g1_fft<9,true>", "fft": "k_g1_fft<9,false>",
                   "ntt": "k_fr_ntt<9,false>"}[dom],
        "achieved": round(achieved, 4),
        "peak": round(peak, 2),
        "unit": "Tmul32/s",
        "frac": round(achieved / peak, 4),
        "traffic": None,
    }

    cpu, exact = None, None
    if not args.no_cpu:
        log("cpu baseline: oracle over %d rows with %d threads" % (args.cpu_rows, args.cpu_threads))
        cpu, exact = cpu_baseline(b, args.cpu_rows, args.cpu_threads)

    out = {
        "metric": METRIC,
        "value": round(value, 3),
        "unit": "rows/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u32",
        "data": "synthetic: uniform Fr coefficients (PCG64 seed %d), setup from a seeded secret" % args.seed,
        "config": {"workload": "fk20_batch_plain", "rows": args.rows, "N": args.N, "l": 16,
                   "fft_size": args.N // 8, "form": "plain batched grid (P2)"},
        "forms": {
            "plain_P2_rows_per_s": round(value, 3),
            "per_row_P1_rows_per_s": round(p1_rate, 3),
            "per_row_P1_latency_ms_p50": round(statistics.median(p1), 3),
            "p1_sample_rows": args.p1_rows,
        },
        "stage_ms_p50": {k: round(v, 3) for k, v in stage_med.items()},
        "wall_s": round(wall, 3),
        "roofline": roofline,
        "cpu_baseline": cpu,
        "bit_exact_rows_checked": (args.cpu_rows if cpu else 0),
        "bit_exact": exact,
    }
    print(json.dumps(out), flush=True)
    if exact is False:
        sys.exit("GPU rows differ from the oracle")


if __name__ == "__main__":
    main()
