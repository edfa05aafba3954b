#!/usr/bin/env python3
"""FK20 multi-proof benchmark on one MI355X: rows proved per second, N = 4096, 512 rows.

A step is one pass of the FK20 path over one batch of synthetic rows (uniform Fr
coefficients, one seeded setup): polynomial -> Toeplitz coefficients -> 16 Fr NTTs per
row -> MSM -> G1 IFFT -> zero upper half -> G1 FFT -> h_fft.  Inputs are resident in HBM
before timing starts.  Forms measured:
  F   batched path (fk20_prove_rows): resident window tables, comb MSM, column-batched G1
      transforms with endomorphism-split twiddles.  `value` is this form.
  F+graph  the same launch sequence replayed from a captured hipGraph.
  P2  the reference's six-launch sequence with grid = rows (fk20_poly2h_fft.cu:47-80).
  P1  the same sequence issued once per row (rows = 1), timed on a bounded sample.
Every launch of F is timed with HIP events on the stream it runs on; the dominant kernel's
roofline uses those durations.  cpu_baseline: the oracle (CPU restatement of FK20Py
fk20_multi) on the host cores over a bounded sample of the same rows; its outputs also
check the GPU rows bit-exactly.
"""
import argparse
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
from fk20_amd.prover import LAUNCH_KINDS, Prover  # noqa: E402

R_MOD = S.R_MOD
METRIC = "FK20 rows proved/sec (N=4096, 512 rows) batched vs plain; p50/p99 latency, bit-exact"
G1_GEN = S.G1_GEN


def log(msg):
    print("[bench] " + msg, file=sys.stderr, flush=True)


TRAFFIC_JSON = os.path.join(ROOT, "profiles", "fk20_batch_hbm_traffic.json")
FPMUL_JSON = os.path.join(ROOT, "profiles", "fp_mul_rate.json")
MAD_JSON = os.path.join(ROOT, "profiles", "mad_rate.json")
# launch kind (as timed by the library) -> kernel symbol prefix in rocprofv3 output
KIND_SYMBOL = {"k_dit_stage": "fk20::k_dit_multi<", "k_dif_stage": "fk20::k_dif_multi<",
               "k_middle": "fk20::k_middle_multi<", "k_msm_comb": "fk20::k_msm_multi",
               "k_toeplitz_fft_T": "fk20::k_toeplitz_fft_multi", "k_fft_output": "fk20::k_output_multi",
               "k_transform_chain": "fk20::k_chain<", "k_twiddle_layer": "fk20::k_twiddle_multi<"}


def _kernel_entry(kernels, kind):
    """The profile entry of a launch kind: exact kernel name, or any instance of a template
    (symbols ending in '<')."""
    sym = KIND_SYMBOL.get(kind)
    hits = [v for k, v in kernels.items() if sym and (k == sym or (sym.endswith("<") and k.startswith(sym)))]
    return hits[0] if len(hits) == 1 else None


def measured_traffic(kernel, N, rows, window):
    """HBM bytes per launch of `kernel` from the calibrated FETCH_SIZE/WRITE_SIZE passes
    (fk20_amd.profsum traffic), if they were taken on this workload; else None."""
    try:
        with open(TRAFFIC_JSON) as f:
            t = json.load(f)
    except OSError:
        return None
    if t.get("config") != {"N": N, "rows": rows, "window": window}:
        return None
    k = _kernel_entry(t["kernels"], kernel)
    if not k:
        return None
    return {"bytes": k["fetch_bytes"] + k["write_bytes"], "fetch_bytes": k["fetch_bytes"],
            "write_bytes": k["write_bytes"], "fetch_scale": t["fetch_scale"], "write_scale": t["write_scale"],
            "source": os.path.relpath(TRAFFIC_JSON, ROOT)}


VALU_JSON = os.path.join(ROOT, "profiles", "fk20_batch_valu_counters.json")


def measured_valu(kernel, N, rows, window, avg_launch_s):
    """Executed VALU instructions per launch of `kernel` from the SQ counter pass
    (fk20_amd.profsum valu), if taken on this workload; the rate uses this run's own launch
    time.  Else None."""
    try:
        with open(VALU_JSON) as f:
            t = json.load(f)
    except OSError:
        return None
    if t.get("config") != {"N": N, "rows": rows, "window": window}:
        return None
    k = _kernel_entry(t["kernels"], kernel)
    if not k:
        return None
    lane_rate = k["valu_wave_instr"] * 64 / avg_launch_s
    peak = WL.PEAK_MUL32_PER_S
    return {"valu_wave_instr_per_launch": k["valu_wave_instr"], "int64_share": round(k["int64_share"], 4),
            "lane_instr_per_s": round(lane_rate, 1), "issue_peak_per_s": peak,
            "issue_frac": round(lane_rate / peak, 4),
            "valu_busy": round(k["valu_busy"], 4) if k.get("valu_busy") is not None else None,
            "source": os.path.relpath(VALU_JSON, ROOT)}


def pct(v, q):
    v = sorted(v)
    return v[min(len(v) - 1, int(round(q / 100.0 * (len(v) - 1))))]


device_setup = S.device_setup


class Batch:
    """Device-resident inputs, setup and outputs of one batch."""

    def __init__(self, rows, N, seed):
        self.rows, self.N, self.n2 = rows, N, N // 8
        self.poly_host = S.random_polys(rows, N, seed)
        self.poly = L.DeviceBuffer.from_array(self.poly_host)
        self.h_fft = L.DeviceBuffer(rows * self.n2 * 144)
        self.tau = S.random_secret(seed + 1)
        self.xext = device_setup(N, self.tau)
        self.xext_host = self.xext.download((16, self.n2, 18))


# ---------------------------------------------------------------- F: batched path

def time_fused(b, prover, sid, steps, warmup):
    """Step times of the batched path (HIP events around each whole step, the library's default
    launch layout), then one pass per step with per-launch events for the kernel breakdown
    (single-stream sequence: launches are timed one after another), then graph replay."""
    stream = L.Stream()
    e0, e1 = L.Event(), L.Event()
    for _ in range(warmup):
        prover.prove_device(sid, b.h_fft, b.poly, b.rows, stream=stream.handle)
    stream.synchronize()
    step_ms = []
    L.synchronize()  # the timed steps start and end on an idle device
    for _ in range(steps):
        e0.record(stream)
        prover.prove_device(sid, b.h_fft, b.poly, b.rows, stream=stream.handle)
        e1.record(stream)
        step_ms.append(e0.elapsed_ms(e1))
    stream.synchronize()
    L.synchronize()
    launches = {}
    prover.set_launch_timing(True)
    for _ in range(steps):
        prover.prove_device(sid, b.h_fft, b.poly, b.rows, stream=stream.handle)
        for kind, ms in prover.launch_times():
            launches.setdefault(LAUNCH_KINDS[kind], []).append(ms)
    stream.synchronize()
    prover.set_launch_timing(False)
    # the same sequence replayed from a hipGraph
    prover.set_graph(True)
    prover.prove_device(sid, b.h_fft, b.poly, b.rows, stream=stream.handle)  # capture
    stream.synchronize()
    graph_ms = []
    for _ in range(steps):
        e0.record(stream)
        prover.prove_device(sid, b.h_fft, b.poly, b.rows, stream=stream.handle)
        e1.record(stream)
        graph_ms.append(e0.elapsed_ms(e1))
    prover.set_graph(False)
    stream.synchronize()
    e0.destroy()
    e1.destroy()
    stream.destroy()
    return step_ms, launches, graph_ms


# ---------------------------------------------------------------- P2 / P1: plain path

def time_plain(b, steps):
    stream = L.Stream()
    tc = L.DeviceBuffer(b.rows * 16 * b.n2 * 32)
    h = L.DeviceBuffer(b.rows * b.n2 * 144)
    e0, e1 = L.Event(), L.Event()
    out = L.DeviceBuffer(b.rows * b.n2 * 144)

    def step():
        s = stream.handle
        L.call("fk20_poly2toeplitz_coefficients", tc.ptr, b.poly.ptr, b.rows, b.N, s)
        L.call("fr_fft", tc.ptr, tc.ptr, b.rows * 16, b.n2, s)
        L.call("fk20_msm", h.ptr, tc.ptr, b.xext.ptr, b.rows, b.n2, s)
        L.call("g1p_ift", h.ptr, h.ptr, b.rows, b.n2, s)
        L.call("fk20_hext2h", h.ptr, b.rows, b.n2, s)
        L.call("g1p_fft", out.ptr, h.ptr, b.rows, b.n2, s)

    step()
    stream.synchronize()
    ms = []
    for _ in range(steps):
        e0.record(stream)
        step()
        e1.record(stream)
        ms.append(e0.elapsed_ms(e1))
    for x in (tc, h, out):
        x.free()
    e0.destroy()
    e1.destroy()
    stream.destroy()
    return ms


def time_p1(b, sample_rows):
    """Per-row plain launch sequence: fk20_poly2h_fft with rows = 1, row after row."""
    stream = L.Stream()
    e0, e1 = L.Event(), L.Event()
    out = L.DeviceBuffer(sample_rows * b.n2 * 144)
    L.call("fk20_poly2h_fft", out.ptr, b.poly.ptr, b.xext.ptr, 1, b.N, stream.handle)
    stream.synchronize()
    lat = []
    for r in range(sample_rows):
        e0.record(stream)
        L.call("fk20_poly2h_fft", out.ptr + r * b.n2 * 144, b.poly.ptr + r * b.N * 32, b.xext.ptr, 1, b.N,
               stream.handle)
        e1.record(stream)
        lat.append(e0.elapsed_ms(e1))
    stream.synchronize()
    out.free()
    e0.destroy()
    e1.destroy()
    stream.destroy()
    return lat


def time_f1(b, prover, sid, reps):
    """Single-row latency of the batched path (C1 on the GPU): fk20_prove_rows with rows = 1."""
    stream = L.Stream()
    e0, e1 = L.Event(), L.Event()
    prover.prove_device(sid, b.h_fft, b.poly, 1, stream=stream.handle)
    stream.synchronize()
    lat = []
    for _ in range(reps):
        e0.record(stream)
        prover.prove_device(sid, b.h_fft, b.poly, 1, stream=stream.handle)
        e1.record(stream)
        lat.append(e0.elapsed_ms(e1))
    stream.synchronize()
    e0.destroy()
    e1.destroy()
    stream.destroy()
    return lat


# ---------------------------------------------------------------- C4: uneven batch

def time_uneven(prover, rows_each, window, steps, seed, keep=False):
    """512 rows each of N = 1024 / 2048 / 4096 (three setups) in one fk20_prove_batch pass;
    each setup uses its autotuned window (tuned.json) or `window`.  Returns the pass times,
    per group (N, secret, host polynomials, host h_fft of the last timed pass) for the check
    against the oracle in the CPU leg, and (keep=True) the setups left loaded, by N."""
    from fk20_amd.autotune import tuned_window

    items, bufs, groups = [], [], []
    for k, N in enumerate((1024, 2048, 4096)):
        tau = S.random_secret(seed + 100 + k)
        xext = device_setup(N, tau)
        sid = prover.load_setup(xext, N, tuned_window(N, window))
        xext.free()
        poly = S.random_polys(rows_each, N, seed + 200 + k)
        pb = L.DeviceBuffer.from_array(poly)
        hb = L.DeviceBuffer(rows_each * (N // 8) * 144)
        bufs += [pb, hb]
        items.append((sid, rows_each, pb, hb))
        groups.append([N, tau, poly, hb])
    stream = L.Stream()
    e0, e1 = L.Event(), L.Event()
    prover.prove_batch_device(items, stream=stream.handle)
    stream.synchronize()
    ms = []
    for _ in range(steps):
        e0.record(stream)
        prover.prove_batch_device(items, stream=stream.handle)
        e1.record(stream)
        ms.append(e0.elapsed_ms(e1))
    stream.synchronize()
    for g in groups:
        g[3] = g[3].download((rows_each, g[0] // 8, 18))
    kept = {}
    for (sid, _, _, _), g in zip(items, groups):
        if keep:
            kept[g[0]] = sid
        else:
            prover.unload_setup(sid)
    for x in bufs:
        x.free()
    e0.destroy()
    e1.destroy()
    stream.destroy()
    return ms, groups, kept


# ---------------------------------------------------------------- C5: serving loop

# ---------------------------------------------------------------- C3: subkernel microbench

def random_points(count, seed):
    """count points k*G (k uniform mod r) with a random Z (non-normalised), on the device."""
    gens = L.DeviceBuffer.from_array(np.tile(G1_GEN, (count, 1)))
    ks = L.DeviceBuffer.from_array(S.random_fr(count, seed))
    pts = L.DeviceBuffer(count * 144)
    L.call("fk20_g1p_op", 3, pts.ptr, gens.ptr, None, ks.ptr, count, None)
    z = L.DeviceBuffer.from_array(np.repeat(S.random_fp_nonzero(count, seed + 1), 3, axis=0))
    L.call("fk20_fp_op", 2, pts.ptr, pts.ptr, z.ptr, 3 * count, None)  # (X, Y, Z) * z
    L.synchronize()
    for b in (gens, ks, z):
        b.free()
    return pts


def time_calls(fn, steps, stream):
    e0, e1 = L.Event(), L.Event()
    fn()
    stream.synchronize()
    ms = []
    for _ in range(steps):
        e0.record(stream)
        fn()
        e1.record(stream)
        ms.append(e0.elapsed_ms(e1))
    stream.synchronize()
    e0.destroy()
    e1.destroy()
    return ms


def bench_micro(args):
    """C3: the drop-in subkernels at their FK20 shapes -- rows*16 Fr NTTs of size n2 on
    uniform Fr, rows G1 FFT/IFFTs of n2 points k*G with random Z, and fk20_msm over n2
    columns x 16 terms (uniform scalars, random projective points) at `rows` rows (throughput)
    and at one row (latency)."""
    rows, n2 = args.rows, args.N // 8
    lg = int(np.log2(n2))
    stream = L.Stream()
    s = stream.handle
    out = {}
    log("micro: %d x 16 Fr NTT-%d" % (rows, n2))
    fr = L.DeviceBuffer.from_array(S.random_fr(rows * 16 * n2, args.seed))
    for name in ("fr_fft", "fr_ift"):
        ms = statistics.median(time_calls(lambda: L.call(name, fr.ptr, fr.ptr, rows * 16, n2, s), args.steps, stream))
        fr_muls = rows * 16 * (n2 // 2) * lg
        out[name] = {"count": rows * 16, "n": n2, "ms_p50": round(ms, 4),
                     "transforms_per_s": round(rows * 16 / (ms / 1e3), 1),
                     "fr_mul_per_s": round(fr_muls / (ms / 1e3), 1)}
    log("micro: %d G1 FFT/IFFT-%d on k*G with random Z" % (rows, n2))
    pts = random_points(rows * n2, args.seed + 2)
    res = L.DeviceBuffer(rows * n2 * 144)
    for name in ("g1p_fft", "g1p_ift"):
        ms = statistics.median(time_calls(lambda: L.call(name, res.ptr, pts.ptr, rows, n2, s), args.steps, stream))
        canon = rows * (WL.stage_fp_muls(args.N)["fft" if name == "g1p_fft" else "ifft"]) * WL.FP_MUL32
        out[name] = {"count": rows, "n": n2, "ms_p50": round(ms, 3),
                     "transforms_per_s": round(rows / (ms / 1e3), 2),
                     "butterflies_per_s": round(rows * (n2 // 2) * lg / (ms / 1e3), 1),
                     "canonical_Tmul32_per_s": round(canon / (ms / 1e3) / 1e12, 3)}
    pts.free()
    log("micro: fk20_msm %d columns x 16 terms, rows %d and 1" % (n2, rows))
    tc = L.DeviceBuffer.from_array(S.random_fr(rows * 16 * n2, args.seed + 3))
    bases = random_points(16 * n2, args.seed + 4)
    for r in (rows, 1):
        ms = statistics.median(time_calls(lambda: L.call("fk20_msm", res.ptr, tc.ptr, bases.ptr, r, n2, s),
                                          max(2, args.steps // 2) if r > 1 else args.steps, stream))
        canon = r * WL.stage_fp_muls(args.N)["msm"] * WL.FP_MUL32
        out["fk20_msm_rows%d" % r] = {"rows": r, "columns": n2, "terms": 16, "ms_p50": round(ms, 3),
                                      "terms_per_s": round(r * 16 * n2 / (ms / 1e3), 1),
                                      "canonical_Tmul32_per_s": round(canon / (ms / 1e3) / 1e12, 3)}
    for b in (fr, res, tc, bases):
        b.free()
    stream.destroy()
    print(json.dumps({"metric": "FK20 subkernel microbench (drop-in entry points)", "unit": "ms",
                      "n_gpus": 1, "steps": args.steps, "higher_is_better": False,
                      "data": "synthetic: uniform Fr, points k*G with random Z (seed %d)" % args.seed,
                      "config": {"workload": "fk20_subkernels", "rows": rows, "N": args.N, "fft_size": n2},
                      "kernels": out}))


# ---------------------------------------------------------------- field multiplication rate

def bench_madrate(args):
    """Issue rate of v_mad_u64_u32 through fk20_mad_probe: 8 blocks of 256 lanes per CU
    (8 waves per SIMD), 8 independent chains per lane."""
    blocks, iters = 256 * 8, 20000
    out = L.DeviceBuffer(blocks * 256 * 8)
    stream = L.Stream()
    L.call("fk20_mad_probe", out.ptr, blocks, 100, stream.handle)
    stream.synchronize()
    ms = statistics.median(time_calls(lambda: L.call("fk20_mad_probe", out.ptr, blocks, iters, stream.handle),
                                      args.steps, stream))
    out.free()
    stream.destroy()
    mads = blocks * 256 * 8 * iters
    rate = mads / (ms / 1e3)
    print(json.dumps({"metric": "v_mad_u64_u32 issue rate (fk20_mad_probe)", "value": round(rate, 1),
                      "unit": "mad/s", "n_gpus": 1, "steps": args.steps, "higher_is_better": True,
                      "dtype": "u32", "data": "synthetic",
                      "config": {"workload": "mad_rate", "blocks": blocks, "lanes_per_block": 256,
                                 "chains_per_lane": 8, "iters": iters},
                      "ms_p50": round(ms, 3), "peak_assumed": WL.PEAK_MUL32_PER_S,
                      "frac_of_assumed_peak": round(rate / WL.PEAK_MUL32_PER_S, 4)}))


def bench_limbs(args):
    """The Fp multiplication in the library's 14 x 29-bit product-scanning layout against a
    12 x 32-bit CIOS one (fk20_fp_mul_probe): 8 blocks of 256 lanes per CU, two independent
    chains of multiplications per lane; both results checked against Python integers on a
    sample of lanes."""
    P = 0x1A0111EA397FE69A4B1BA7B6434BACD764774B84F38512BF6730D2A0F6B0F6241EABFFFEB153FFFFB9FEFFFFFFFFAAAB
    rng = np.random.default_rng(args.seed)
    xs = [int.from_bytes(rng.bytes(48), "little") % P for _ in range(1024)]
    yv = int.from_bytes(rng.bytes(48), "little") % P

    def words(v):
        return np.array([(v >> (32 * i)) & 0xFFFFFFFF for i in range(12)], dtype=np.uint32)

    xb = L.DeviceBuffer.from_array(np.stack([words(x) for x in xs]))
    yb = L.DeviceBuffer.from_array(words(yv))
    blocks, iters = 256 * 8, 2000
    out = L.DeviceBuffer(blocks * 256 * 48)
    stream = L.Stream()
    res = {}
    for limbs, rbits in ((29, 406), (32, 384)):
        L.call("fk20_fp_mul_probe", limbs, out.ptr, xb.ptr, yb.ptr, blocks, 10, stream.handle)
        stream.synchronize()
        ms = statistics.median(time_calls(lambda: L.call("fk20_fp_mul_probe", limbs, out.ptr, xb.ptr, yb.ptr, blocks,
                                                         iters, stream.handle), args.steps, stream))
        got = out.download((blocks * 256, 12), np.uint32)
        rinv = pow(2, -rbits, P)
        ok = True
        for lane in range(0, blocks * 256, blocks * 256 // 64):  # 64 sample lanes
            a = xs[lane & 1023]
            # chains: a y^iters and (a y) y^iters, each product times R^-1, then their product
            c1 = a * pow(yv * rinv, iters, P) % P
            c2 = a * yv * rinv % P * pow(yv * rinv, iters, P) % P
            v = sum(int(w) << (32 * i) for i, w in enumerate(got[lane]))
            ok = ok and v % P == c1 * c2 * rinv % P and v < 2 * P
        muls = blocks * 256 * (2 * iters + 2)
        res[limbs] = {"ms_p50": round(ms, 3), "fp_mul_per_s": round(muls / (ms / 1e3), 1), "exact": ok}
        log("limbs %d: %.3f ms, %.1f G Fp mul/s, exact %s" % (limbs, ms, muls / (ms / 1e3) / 1e9, ok))
    for b in (xb, yb, out):
        b.free()
    stream.destroy()
    print(json.dumps({"metric": "Fp Montgomery multiplication rate by limb layout (fk20_fp_mul_probe)",
                      "value": res[29]["fp_mul_per_s"], "unit": "Fp mul/s", "n_gpus": 1, "steps": args.steps,
                      "higher_is_better": True, "dtype": "u32", "data": "synthetic",
                      "config": {"workload": "fp_limbs", "blocks": blocks, "lanes_per_block": 256,
                                 "chains_per_lane": 2, "iters": iters},
                      "layouts": {"14x29_product_scanning": res[29], "12x32_cios": res[32]},
                      "ratio_29_over_32": round(res[29]["fp_mul_per_s"] / res[32]["fp_mul_per_s"], 3)}))
    if not (res[29]["exact"] and res[32]["exact"]):
        sys.exit("limb probe results differ from Python integers")


def bench_fieldmul(args):
    """Throughput of the Fp Montgomery multiplication itself (14 x 29-bit limbs), the unit
    every G1 kernel is built from: Fermat inversions a^(p-2) through fk20_fp_op (one lane per
    element, 384 squarings + one multiplication per set bit of p - 2, no memory traffic in
    the loop), on `fieldmul_count` uniform non-zero elements; a sample is checked on the host."""
    count = args.fieldmul_count
    p_mod = 0x1A0111EA397FE69A4B1BA7B6434BACD764774B84F38512BF6730D2A0F6B0F6241EABFFFEB153FFFFB9FEFFFFFFFFAAAB
    ops_per = 384 + bin(p_mod - 2).count("1")
    a = S.random_fp_nonzero(count, args.seed + 7)
    A = L.DeviceBuffer.from_array(a)
    Z = L.DeviceBuffer(count * 48)
    stream = L.Stream()
    L.call("fk20_fp_op", 5, Z.ptr, A.ptr, None, count, stream.handle)
    stream.synchronize()
    ms = statistics.median(time_calls(lambda: L.call("fk20_fp_op", 5, Z.ptr, A.ptr, None, count, stream.handle),
                                      args.steps, stream))
    z = Z.download((count, 6))
    for i in range(0, count, max(1, count // 16)):
        x = sum(int(w) << (64 * k) for k, w in enumerate(a[i].tolist()))
        y = sum(int(w) << (64 * k) for k, w in enumerate(z[i].tolist()))
        assert x * y % p_mod == 1, i
    for b_ in (A, Z):
        b_.free()
    stream.destroy()
    rate = count * ops_per / (ms / 1e3)
    print(json.dumps({"metric": "Fp multiplications per second (14x29-bit Montgomery, in Fermat inversions)",
                      "value": round(rate, 1), "unit": "Fp mul/s", "n_gpus": 1, "steps": args.steps,
                      "higher_is_better": True, "dtype": "u32", "data": "synthetic: uniform non-zero Fp",
                      "config": {"workload": "fp_mul_rate", "elements": count, "muls_per_element": ops_per},
                      "ms_p50": round(ms, 3), "mul32_per_s": round(rate * WL.FP_MUL32, 1),
                      "note": "squarings counted as multiplications; mul32 = the 300-per-Fp-mul convention"}))


# ---------------------------------------------------------------- commitments (the step before the path)

def bench_commit(args):
    """KZG commitments of `rows` polynomials of N coefficients (FK20Py commit_to_poly) over
    window tables of the SRS, timed like the batched path; a bounded sample checked against
    the closed form [f(tau)]G on the host."""
    rows, N = args.rows, args.N
    log("commit: %d rows, N=%d, window %d" % (rows, N, args.window))
    tau = S.random_secret(args.seed + 1)
    pw, acc = [], 1
    for _ in range(N):
        pw.append(S.int_to_fr(acc))
        acc = acc * tau % R_MOD
    gens = L.DeviceBuffer.from_array(np.tile(G1_GEN, (N, 1)))
    ks = L.DeviceBuffer.from_array(np.array(pw))
    srs = L.DeviceBuffer(N * 144)
    L.call("fk20_g1p_op", 3, srs.ptr, gens.ptr, None, ks.ptr, N, None)
    L.synchronize()
    prover = Prover()
    t0 = time.perf_counter()
    sid = prover.load_commit_setup(srs, N, args.window)
    setup_s = time.perf_counter() - t0
    for b in (gens, ks, srs):
        b.free()
    poly_host = S.random_polys(rows, N, args.seed)
    poly = L.DeviceBuffer.from_array(poly_host)
    out = L.DeviceBuffer(rows * 144)
    stream = L.Stream()
    ms = time_calls(lambda: prover.commit_device(sid, out, poly, rows, stream=stream.handle), args.steps, stream)
    got = out.download((rows, 18))
    exact = None
    if not args.no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle as O  # the parity reference
        exact = all(np.array_equal(O.to_affine(got[r]), O.to_affine(O.commit_closed_form(poly_host[r], tau)))
                    for r in range(min(rows, 8)))
    p50 = statistics.median(ms)
    exe = N * ((256 + args.window - 1) // args.window) * WL.MIXED_ADD * rows
    value = rows / (p50 / 1e3)
    print(json.dumps({
        "metric": "KZG commitments per second (N=%d, %d rows)" % (N, rows), "value": round(value, 2),
        "unit": "rows/s", "n_gpus": 1, "steps": args.steps, "ms_per_step": round(p50, 3),
        "higher_is_better": True, "dtype": "u32",
        "data": "synthetic: uniform Fr coefficients (PCG64 seed %d), SRS of a seeded secret" % args.seed,
        "config": {"workload": "kzg_commit", "rows": rows, "N": N, "window_bits": args.window},
        "setup_table_build_s": round(setup_s, 3),
        "roofline": {"bound": "valu", "achieved": round(exe * WL.FP_MUL32 / (p50 / 1e3) / 1e12, 4),
                     "peak": round(WL.PEAK_MUL32_PER_S / 1e12, 2), "unit": "Tmul32/s",
                     "frac": round(exe * WL.FP_MUL32 / (p50 / 1e3) / WL.PEAK_MUL32_PER_S, 4),
                     "work": "executed mixed additions (11 Fp mul each) x 300 mul32; reductions excluded"},
        "bit_exact_rows_checked": 0 if exact is None else min(rows, 8), "bit_exact": exact}))
    prover.close()
    poly.free()
    out.free()
    stream.destroy()


def serve_levels(prover, sids, rates, duration, budget_bytes, slots, max_batch_rows, seed, stop_past_saturation=True,
                 warmup_s=1.0):
    """Poisson arrivals of one-row blobs, N = 1024/2048/4096 weighted 1:2:5, into the serving
    loop (fk20_server_*) over `prover`'s resident setups sids[N], at each offered load in
    rows/s for `duration` seconds; latency = completion - submission of each request.  An
    untimed level of `warmup_s` seconds at the highest load runs first (the first level of a
    process ran 20-50 % slower at p50 than the same level repeated)."""
    import threading

    from fk20_amd.lib import PinnedArray
    from fk20_amd.server import Server

    polys, outs = {}, {}
    pool = 2048
    for k, N in enumerate((1024, 2048, 4096)):
        polys[N] = PinnedArray((64, N, 4))
        polys[N].array[:] = S.random_polys(64, N, seed + 300 + k)
        outs[N] = PinnedArray((pool, N // 8, 18))
    sizes = np.array([1024] * 1 + [2048] * 2 + [4096] * 5)
    levels = []
    plan = ([(max(rates), warmup_s, False)] if warmup_s > 0 else []) + [(r, duration, True) for r in rates]
    for rate, secs, timed in plan:
        srv = Server(prover, budget_bytes=budget_bytes, slots=slots, max_batch_rows=max_batch_rows)
        rng = np.random.default_rng(seed + int(rate) + (0 if timed else 7))
        n = max(64, int(rate * secs))
        gaps = rng.exponential(1.0 / rate, size=n)
        which = sizes[rng.integers(0, len(sizes), size=n)]
        ids = [None] * n
        lat = [None] * n
        done_at = [None] * n
        submitted = threading.Semaphore(0)

        def waiter():
            for i in range(n):
                submitted.acquire()
                lat[i] = srv.wait(ids[i])
                done_at[i] = time.perf_counter()

        wt = threading.Thread(target=waiter)
        wt.start()
        t0 = time.perf_counter()
        t_next = t0
        for i in range(n):
            t_next += gaps[i]
            d = t_next - time.perf_counter()
            if d > 0:
                time.sleep(d)
            N = int(which[i])
            slot = i % pool
            ids[i] = srv.submit(sids[N], polys[N].ptr + (i % 64) * N * 32, 1,
                                outs[N].ptr + slot * (N // 8) * 144)
            submitted.release()
        wt.join()
        t_end = max(done_at)
        st = srv.stats()
        srv.close()
        lat_ms = np.array(lat) * 1e3
        if not timed:
            log("serve warm-up: %.0f rows/s offered for %.1f s, p50 %.2f ms (not reported)"
                % (rate, secs, float(np.percentile(lat_ms, 50))))
            continue
        levels.append({
            "offered_rows_per_s": rate,
            "sustained_rows_per_s": round(n / (t_end - t0), 1),
            "p50_ms": round(float(np.percentile(lat_ms, 50)), 3),
            "p99_ms": round(float(np.percentile(lat_ms, 99)), 3),
            "requests": n,
            "batches": st["batches"],
            "max_inflight_rows": st["max_inflight_rows"],
        })
        log("serve %.0f rows/s offered: sustained %.0f, p50 %.2f ms, p99 %.2f ms, %d batches"
            % (rate, levels[-1]["sustained_rows_per_s"], levels[-1]["p50_ms"], levels[-1]["p99_ms"], st["batches"]))
        if stop_past_saturation and levels[-1]["sustained_rows_per_s"] < 0.8 * rate:
            break  # past saturation
    for a in list(polys.values()) + list(outs.values()):
        a.free()
    return levels


def serving_setups(prover, window, seed):
    """The three resident setups (N = 1024 / 2048 / 4096) of the serving workload."""
    from fk20_amd.autotune import tuned_window

    sids = {}
    for k, N in enumerate((1024, 2048, 4096)):
        xext = device_setup(N, S.random_secret(seed + 100 + k))
        sids[N] = prover.load_setup(xext, N, tuned_window(N, window))
        xext.free()
    return sids


def bench_serve(args):
    """Poisson arrivals of one-row blobs, N = 1024/2048/4096 weighted 1:2:5, offered load swept."""
    prover = Prover()
    sids = serving_setups(prover, args.window, args.seed)
    rates = [float(r) for r in args.rates.split(",")]
    levels = serve_levels(prover, sids, rates, args.duration, args.budget_gb << 30, args.slots,
                          args.max_batch_rows, args.seed)
    prover.close()
    best = max(levels, key=lambda x: x["sustained_rows_per_s"])
    out = {
        "metric": "FK20 serving: sustained rows/s and p50/p99 latency under Poisson load",
        "value": best["sustained_rows_per_s"],
        "unit": "rows/s",
        "n_gpus": 1,
        "higher_is_better": True,
        "dtype": "u32",
        "data": "synthetic one-row blobs, N=1024/2048/4096 weighted 1:2:5, Poisson arrivals",
        "config": {"workload": "fk20_serving", "budget_bytes": args.budget_gb << 30, "slots": args.slots,
                   "max_batch_rows": args.max_batch_rows, "msm_window_bits": args.window,
                   "duration_s_per_level": args.duration, "warmup_s": 1.0},
        "levels": levels,
    }
    print(json.dumps(out), flush=True)


# ---------------------------------------------------------------- CPU baseline

def host_cpus():
    """(CPUs this process may run on, CPUs the machine reports, CPU model name)."""
    try:
        avail = len(os.sched_getaffinity(0))
    except AttributeError:
        avail = os.cpu_count() or 1
    model = None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return avail, os.cpu_count(), model


def cpu_baseline(b, threads, runs, gpu_out, uneven_groups):
    """The CPU leg: the oracle's restatement of FK20Py fk20_multi (double-and-add G1, one row
    per thread) timed on the host cores, `runs` runs of `threads` rows each (median run), and
    the bit-exactness checks of this run's GPU outputs: every row of the timed batch and of
    the uneven batch against the scalar-shadow closed form, and the CPU rows against the same
    GPU rows."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O

    O.lib()
    avail, total, model = host_cpus()
    if threads <= 0:
        threads = max(1, min(16, avail))  # at most 16 host threads per GPU
    runs = max(1, min(runs, b.rows // threads))
    times, cpu_rows = [], []
    for k in range(runs):
        lo = k * threads
        t0 = time.perf_counter()
        cpu_rows.append(O.poly2h_fft(b.poly_host[lo:lo + threads], b.xext_host, nthreads=threads))
        times.append(time.perf_counter() - t0)
    el = statistics.median(times)
    checked = runs * threads
    cpu_same = O.match_affine(gpu_out[:checked], O.to_affine(np.concatenate(cpu_rows).reshape(-1, 18)))[0] == 0
    bad, first = O.match_affine(gpu_out, O.fk20_closed_form_affine(b.poly_host, b.tau))
    exact = {"batch_rows": b.rows, "batch_points_differing": bad, "cpu_rows_equal_gpu": cpu_same}
    if uneven_groups:
        exact["uneven_rows"] = 0
        exact["uneven_points_differing"] = 0
        for N, tau, poly, out in uneven_groups:
            exact["uneven_rows"] += poly.shape[0]
            exact["uneven_points_differing"] += O.match_affine(out, O.fk20_closed_form_affine(poly, tau))[0]
    ok = bad == 0 and cpu_same and not exact.get("uneven_points_differing")
    return {
        "value": round(threads / el, 4),
        "unit": "rows/s",
        "cores": threads,
        "kind": "port",
        "sample": "%d runs x %d rows of N=%d (%d of the %d timed rows, one oracle row per thread), "
                  "median run %.1f s" % (runs, threads, b.N, checked, b.rows, el),
        "run_s": [round(t, 2) for t in times],
        "host_cpus_available": avail,
        "host_cpus_total": total,
        "cpu_model": model,
    }, exact, ok


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--rows", type=int, default=512)
    ap.add_argument("--N", type=int, default=4096)
    ap.add_argument("--window", type=int, default=0,
                    help="MSM window bits (8 to 13); 0: the autotuned choice (tuned.json) or 8")
    ap.add_argument("--seed", type=int, default=2024)
    ap.add_argument("--plain-steps", type=int, default=2, help="steps of the plain P2 form (0: skip)")
    ap.add_argument("--p1-rows", type=int, default=4, help="rows timed for the per-row form P1 (0: skip)")
    ap.add_argument("--cpu-runs", type=int, default=3, help="CPU baseline runs (one row per thread each)")
    ap.add_argument("--cpu-threads", type=int, default=0,
                    help="CPU baseline threads (0: the CPUs this process may use, at most 16)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--streams", type=int, default=0,
                    help="launch sequences per batch: 2 halves on two streams or 1 (0: the tuned choice)")
    ap.add_argument("--chain", action="store_true",
                    help="transform stages as one persistent dependency-driven launch (default: the tuned choice)")
    ap.add_argument("--schedule", default="tuned",
                    help="'tuned' (tuned.json for N and rows) or 'wnaf,radix,waves_per_simd', e.g. 5,4,2")
    ap.add_argument("--uneven-rows", type=int, default=512,
                    help="rows per size in the uneven N=1024/2048/4096 batch (0: skip)")
    ap.add_argument("--workload", default="batch",
                    choices=["batch", "serve", "micro", "commit", "fieldmul", "madrate", "limbs"])
    ap.add_argument("--fieldmul-count", type=int, default=1 << 21, help="fieldmul: elements inverted per call")
    ap.add_argument("--calibrate-copy-mb", type=int, default=0,
                    help="before timing, one device-to-device copy of this many MiB (the known-size "
                         "kernel that scales FETCH_SIZE / WRITE_SIZE counter runs)")
    ap.add_argument("--rates", default="1000,2000,4000,6000,8000,10000,14000",
                    help="serve: offered loads in rows/s")
    ap.add_argument("--duration", type=float, default=2.0, help="serve: seconds per load level")
    ap.add_argument("--serve-rates", default="2000,8000",
                    help="batch workload: offered loads (rows/s) of the serving leg after the batch legs "
                         "(empty: skip); the line's serving p50/p99 are those of the last load")
    ap.add_argument("--budget-gb", type=int, default=8, help="serve: device budget for in-flight rows")
    ap.add_argument("--slots", type=int, default=2, help="serve: batches in flight")
    ap.add_argument("--max-batch-rows", type=int, default=1024)
    args = ap.parse_args()

    if args.gpus != 1:
        print("bench.py measures one GPU only: FK20 rows are independent, and this build runs "
              "and reports the single-GPU path (--gpus %d requested)" % args.gpus, file=sys.stderr)
        sys.exit(2)
    if L.device_count() < 1:
        sys.exit("no HIP device visible")

    if not args.window:
        from fk20_amd.autotune import tuned_window
        args.window = tuned_window(args.N)
    if args.workload == "serve":
        bench_serve(args)
        return
    if args.workload == "micro":
        bench_micro(args)
        return
    if args.workload == "commit":
        bench_commit(args)
        return
    if args.workload == "fieldmul":
        bench_fieldmul(args)
        return
    if args.workload == "madrate":
        bench_madrate(args)
        return
    if args.workload == "limbs":
        bench_limbs(args)
        return

    if args.calibrate_copy_mb:
        nb = args.calibrate_copy_mb << 20
        src, dst = L.DeviceBuffer(nb), L.DeviceBuffer(nb)
        src.fill(1)
        L.call("fk20_memcpy_d2d", dst.ptr, src.ptr, nb)
        L.synchronize()
        src.free()
        dst.free()
        log("calibration copy: %d MiB device to device" % args.calibrate_copy_mb)

    log("generating %d rows, N=%d, setup on device" % (args.rows, args.N))
    b = Batch(args.rows, args.N, args.seed)
    prover = Prover()
    # the transform schedule: the autotuned one for (N, rows) (tuned.json), unless given here
    from fk20_amd.autotune import apply_schedule, tuned_schedule
    sch = tuned_schedule(args.N, args.rows) if args.schedule == "tuned" else dict(
        zip(("wnaf", "radix", "waves_per_simd"), (int(x) for x in args.schedule.split(","))), streams=2, chain=0)
    if args.streams:
        sch["streams"] = args.streams
    if args.chain:
        sch["chain"] = 1
    apply_schedule(prover, sch)
    log("transform schedule: %s" % json.dumps(sch))
    t0 = time.perf_counter()
    sid = prover.load_setup(b.xext, args.N, args.window)
    setup_s = time.perf_counter() - t0
    prover.reserve(args.rows, args.N)
    log("setup tables built in %.2f s (window %d bits)" % (setup_s, args.window))

    log("timing F: %d warmup + %d steps" % (args.warmup, args.steps))
    step_ms, launches, graph_ms = time_fused(b, prover, sid, args.steps, args.warmup)
    ms = statistics.median(step_ms)
    value = args.rows / (ms / 1e3)
    log("F: %.2f ms/step p50, %.1f rows/s; graph replay %.2f ms" % (ms, value, statistics.median(graph_ms)))

    # dominant kernel: the one with the largest total time per step
    per_step = {k: sum(v) / args.steps for k, v in launches.items()}
    dom = max(per_step, key=per_step.get)
    n_launch = int(round(len(launches[dom]) / args.steps))
    avg_ms = statistics.mean(launches[dom])
    lg = int(np.log2(args.N // 8))
    if dom == "k_dit_stage":
        stages = list(range(lg - 1))
    elif dom == "k_dif_stage":
        stages = list(range(1, lg))
    elif dom == "k_twiddle_layer":
        stages = WL.r22_layers(args.N)
    else:
        stages = [None]
    # executed work of this kernel's algorithm per launch (mean over its launches in a step)
    wl_kw = {"wnaf": sch["wnaf"], "r22": sch["radix"] == 4}
    exe_fp = (sum(WL.executed_launch_fp_muls(args.N, dom, st, args.window, **wl_kw) for st in stages)
              / len(stages) * args.rows)
    work = exe_fp * WL.FP_MUL32
    canon = WL.launch_mul32(args.N, dom, args.rows)
    achieved = work / (avg_ms / 1e3) / 1e12
    peak = WL.PEAK_MUL32_PER_S / 1e12
    roofline = {
        "bound": "valu",
        "kernel": dom,
        "launches_per_step": n_launch,
        "avg_launch_ms": round(avg_ms, 4),
        "achieved": round(achieved, 4),
        "peak": round(peak, 2),
        "unit": "Tmul32/s",
        "frac": round(achieved / peak, 4),
        "traffic": None,
        "traffic_detail": None,
        "valu_counters": None,
        "work": "executed Fp muls of the kernel's algorithm x 300 mul32 (12x32-bit CIOS convention)",
        "timing": "per-launch HIP events on the launch stream, one launch sequence of all rows "
                  "(the --streams 1 layout: each launch has the GPU to itself)",
        "fp_mul_per_s": round(exe_fp / (avg_ms / 1e3), 1),
        "achieved_canonical": round(canon / (avg_ms / 1e3) / 1e12, 4),
        # the whole step in the benchmarked layout: executed work of every kernel / step time
        "achieved_step": round(WL.executed_row_fp_muls(args.N, args.window, chain=bool(sch["chain"]), **wl_kw)
                               * args.rows * WL.FP_MUL32 / (ms / 1e3) / 1e12, 4),
    }

    roofline["valu_counters"] = measured_valu(dom, args.N, args.rows, args.window, avg_ms / 1e3)
    try:  # the library's own Fp multiplication rate (bench.py --workload fieldmul)
        with open(FPMUL_JSON) as f:
            ceil = json.load(f)["value"]
        roofline["fp_mul_ceiling"] = {"fp_mul_per_s": ceil, "frac": round(exe_fp / (avg_ms / 1e3) / ceil, 4),
                                      "source": os.path.relpath(FPMUL_JSON, ROOT)}
    except (OSError, KeyError, ValueError):
        roofline["fp_mul_ceiling"] = None
    try:  # the measured v_mad_u64_u32 issue rate (bench.py --workload madrate)
        with open(MAD_JSON) as f:
            mad = json.load(f)["value"]
        roofline["mad_probe"] = {"mad_per_s": mad, "peak_frac": round(mad / WL.PEAK_MUL32_PER_S, 4),
                                 "achieved_frac": round(achieved * 1e12 / mad, 4),
                                 "source": os.path.relpath(MAD_JSON, ROOT)}
    except (OSError, KeyError, ValueError):
        roofline["mad_probe"] = None
    tr = measured_traffic(dom, args.N, args.rows, args.window)
    if tr:
        roofline["traffic"] = tr["bytes"]
        # algorithmic bytes: a stage launch reads and writes every (position, row) point once
        n2 = args.N // 8
        alg = {"k_dit_stage": 2, "k_dif_stage": 2, "k_middle": 2,
               "k_transform_chain": 2 * (2 * lg - 1)}.get(dom, 0) * n2 * args.rows * 144
        if dom == "k_twiddle_layer":  # each twiddled point read and written once
            layers = WL.r22_layers(args.N)
            alg = round(sum(len(WL.r22_layer(sd, lg)) for sd, _ in layers) / len(layers) * args.rows * 288)
        tr["algorithmic_bytes"] = alg or None
        roofline["traffic_detail"] = tr

    forms = {"F_rows_per_s": round(value, 3),
             "F_graph_rows_per_s": round(args.rows / (statistics.median(graph_ms) / 1e3), 3),
             "F_step_ms_p50": round(ms, 3), "F_step_ms_p99": round(pct(step_ms, 99), 3)}
    if args.plain_steps:
        log("timing P2 plain batched grid: %d steps" % args.plain_steps)
        p2 = time_plain(b, args.plain_steps)
        forms["P2_rows_per_s"] = round(args.rows / (statistics.median(p2) / 1e3), 3)
        forms["P2_step_ms_p50"] = round(statistics.median(p2), 3)
    if args.p1_rows:
        log("timing P1 per-row launches on %d rows" % args.p1_rows)
        p1 = time_p1(b, args.p1_rows)
        forms["P1_rows_per_s"] = round(1e3 / statistics.median(p1), 3)
        forms["P1_row_latency_ms_p50"] = round(statistics.median(p1), 3)
        forms["P1_sample_rows"] = args.p1_rows
        forms["F_over_P1"] = round(value / forms["P1_rows_per_s"], 1)
    if "P2_rows_per_s" in forms:
        forms["F_over_P2"] = round(value / forms["P2_rows_per_s"], 2)
    if args.p1_rows:
        f1 = time_f1(b, prover, sid, max(3, args.p1_rows))
        forms["F_1row_latency_ms_p50"] = round(statistics.median(f1), 3)
    # the F output of the last timed step (the graph replays rewrote it)
    gpu_out = b.h_fft.download((b.rows, b.n2, 18))
    # the uneven and serving legs load their own three setups: free this one's tables first
    prover.unload_setup(sid)
    uneven, uneven_groups, serve_sids = None, None, {}
    if args.uneven_rows:
        log("timing uneven batch: %d rows each of N=1024/2048/4096 in one pass" % args.uneven_rows)
        # uneven batches (and the serving loop's mixed batches after them) run the schedule
        # tuned for the mix
        msch = tuned_schedule("mixed", 3 * args.uneven_rows) if args.schedule == "tuned" else sch
        apply_schedule(prover, msch)
        log("uneven/serving schedule: %s" % json.dumps(msch))
        um, uneven_groups, serve_sids = time_uneven(prover, args.uneven_rows, args.window,
                                                    max(2, args.steps // 2), args.seed, keep=True)
        uneven = {"rows": 3 * args.uneven_rows, "ms_p50": round(statistics.median(um), 3),
                  "rows_per_s": round(3 * args.uneven_rows / (statistics.median(um) / 1e3), 1),
                  "schedule": msch}
    serving = None
    if args.serve_rates:
        # the serving loop over the same three resident setups: one-row blobs, N = 1024 / 2048 /
        # 4096 weighted 1:2:5, Poisson arrivals at each offered load
        rates = [float(r) for r in args.serve_rates.split(",")]
        log("serving: %s rows/s offered, %.1f s each, %d GiB budget" % (args.serve_rates, args.duration, args.budget_gb))
        if not serve_sids:
            serve_sids = serving_setups(prover, args.window, args.seed)
        levels = serve_levels(prover, serve_sids, rates, args.duration, args.budget_gb << 30, args.slots,
                              args.max_batch_rows, args.seed, stop_past_saturation=False)
        serving = {"budget_bytes": args.budget_gb << 30, "slots": args.slots,
                   "max_batch_rows": args.max_batch_rows, "duration_s_per_level": args.duration, "warmup_s": 1.0,
                   "blobs": "one row each, N=1024/2048/4096 weighted 1:2:5, Poisson arrivals",
                   "p50_ms": levels[-1]["p50_ms"], "p99_ms": levels[-1]["p99_ms"],
                   "offered_rows_per_s": levels[-1]["offered_rows_per_s"],
                   "sustained_rows_per_s": levels[-1]["sustained_rows_per_s"], "levels": levels}
    for sid in serve_sids.values():
        prover.unload_setup(sid)

    cpu, exact, ok = None, None, None
    if not args.no_cpu:
        log("cpu leg: oracle baseline (%d runs) and bit-exact checks of every GPU row" % args.cpu_runs)
        cpu, exact, ok = cpu_baseline(b, args.cpu_threads, args.cpu_runs, gpu_out, uneven_groups)

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
        "config": {"workload": "fk20_batch", "rows": args.rows, "N": args.N, "l": 16, "fft_size": args.N // 8,
                   "form": "F batched (fk20_prove_rows)", "msm_window_bits": args.window,
                   "transform": "persistent chain" if sch["chain"] else "per-stage launches",
                   "streams": sch["streams"], "schedule": sch},
        "latency_ms": {"step_p50": round(ms, 3), "step_p99": round(pct(step_ms, 99), 3)},
        "forms": forms,
        "uneven_batch": uneven,
        "serving": serving,
        "launch_ms_per_step": {k: round(v, 3) for k, v in sorted(per_step.items(), key=lambda x: -x[1])},
        "setup_table_build_s": round(setup_s, 3),
        "roofline": roofline,
        "cpu_baseline": cpu,
        "bit_exact": ok,
        "bit_exact_detail": exact,
    }
    print(json.dumps(out), flush=True)
    prover.close()
    if ok is False:
        sys.exit("GPU rows differ from the oracle")


if __name__ == "__main__":
    main()
