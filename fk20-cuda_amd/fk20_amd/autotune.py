"""Autotuner for the batched path: picks the MSM window width per polynomial size.

For each N it builds the setup tables of every candidate window (8, 10, 12, 13 bits; larger
windows mean fewer mixed additions per term but twice the table memory per bit), times
fk20_prove_rows over a batch with HIP events, and keeps the fastest whose tables fit the
memory cap.  Results persist as JSON keyed by device name and N (tuned.json next to the
library), and bench.py / callers read them with `tuned_window`.

    python3 -m fk20_amd.autotune [--rows 512] [--sizes 1024,2048,4096] [--max-table-gb 72]
"""
import argparse
import json
import os
import statistics
import sys

import numpy as np

from . import lib as L
from . import synthetic as S
from .prover import Prover, setup_table_bytes

TUNED = os.path.join(L.PKG_DIR, "tuned.json")
CANDIDATE_WINDOWS = (8, 10, 12, 13)


def device_name():
    # libfk20 is built for gfx950 only, so the target name is the key
    return "gfx950"


def load():
    if os.path.exists(TUNED):
        with open(TUNED) as f:
            return json.load(f)
    return {}


def tuned_window(N, default=8):
    entry = load().get(device_name(), {}).get(str(N))
    return int(entry["window"]) if entry else default


def _time_window(prover, xext_dev, N, window, poly_dev, out_dev, rows, reps):
    sid = prover.load_setup(xext_dev, N, window)
    stream = L.Stream()
    e0, e1 = L.Event(), L.Event()
    prover.prove_device(sid, out_dev, poly_dev, rows, stream=stream.handle)  # warm-up
    stream.synchronize()
    ms = []
    for _ in range(reps):
        e0.record(stream)
        prover.prove_device(sid, out_dev, poly_dev, rows, stream=stream.handle)
        e1.record(stream)
        ms.append(e0.elapsed_ms(e1))
    prover.unload_setup(sid)
    for x in (e0, e1):
        x.destroy()
    stream.destroy()
    return statistics.median(ms)


def tune(sizes=(1024, 2048, 4096), rows=512, max_table_gb=72.0, reps=3, seed=11, xext_for=None):
    """xext_for(N) -> DeviceBuffer of xext_fft[16][N/8] (a synthetic setup by default)."""
    if xext_for is None:
        sys.path.insert(0, os.path.dirname(L.PKG_DIR))
        from bench import device_setup  # the GPU setup builder of the benchmark
        xext_for = lambda N: device_setup(N, S.random_secret(seed + N))
    results = load()
    dev = results.setdefault(device_name(), {})
    prover = Prover()
    for N in sizes:
        xext = xext_for(N)
        poly = L.DeviceBuffer.from_array(S.random_polys(rows, N, seed))
        out = L.DeviceBuffer(rows * (N // 8) * 144)
        timings = {}
        for w in CANDIDATE_WINDOWS:
            if setup_table_bytes(N, w) > max_table_gb * (1 << 30):
                continue
            timings[w] = _time_window(prover, xext, N, w, poly, out, rows, reps)
            print("autotune N=%d window=%d: %.3f ms / %d rows" % (N, w, timings[w], rows), file=sys.stderr, flush=True)
        best = min(timings, key=timings.get)
        dev[str(N)] = {"window": best, "rows": rows, "ms": {str(k): round(v, 3) for k, v in timings.items()}}
        for b in (xext, poly, out):
            b.free()
    prover.close()
    with open(TUNED, "w") as f:
        json.dump(results, f, indent=1, sort_keys=True)
    return results


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--rows", type=int, default=512)
    ap.add_argument("--sizes", default="1024,2048,4096")
    ap.add_argument("--max-table-gb", type=float, default=72.0)
    args = ap.parse_args()
    res = tune(tuple(int(x) for x in args.sizes.split(",")), args.rows, args.max_table_gb)
    print(json.dumps(res))


if __name__ == "__main__":
    main()
