"""Autotuner of the batched path: MSM window and transform schedule per polynomial size and
batch size.

For each N it first builds the setup tables of every candidate window (8, 10, 12, 13 bits;
larger windows mean fewer mixed additions per term but twice the table memory per bit),
times fk20_prove_rows over a batch with HIP events, and keeps the fastest whose tables fit
the memory cap.  With that window it then times, for each batch size, every transform
schedule the library builds -- wNAF width 3/4/5 of the twiddle programs, radix-2^2 or radix-2
butterflies, kernels compiled for two or three waves per SIMD -- each with the rows split in
two halves on two streams or in one launch sequence, and the persistent single-launch
transform at each width, and keeps the fastest.  Results persist as JSON keyed by device name,
N and batch size (tuned.json next to the library); bench.py and callers read them with
`tuned_window` and `tuned_schedule` / `apply_tuned`.

The reference fixes the same kinds of choice at compile time (CUDAGRAPH, fk20benchmark.cu:5;
FFT_Sharedmem, fk20benchmark.cu:302; SHAREDMEMROOTS, fk20_hext_fft2h_fft.cu:4).

    python3 -m fk20_amd.autotune [--rows 512,64,1] [--sizes 1024,2048,4096] [--max-table-gb 96]
                                 [--mixed-rows 512]
"""
import argparse
import json
import os
import statistics
import sys

from . import lib as L
from . import synthetic as S
from .prover import Prover, setup_table_bytes

TUNED = os.path.join(L.PKG_DIR, "tuned.json")
CANDIDATE_WINDOWS = (8, 10, 12, 13)
WNAF_WIDTHS = (3, 4, 5)
RADICES = (4, 2)
WAVES = (2, 3)
DEFAULT_SCHEDULE = {"wnaf": 5, "radix": 4, "waves_per_simd": 2, "streams": 2, "chain": 0}


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


def tuned_schedule(N, rows):
    """The tuned schedule for N (or "mixed": batches of N = 1024/2048/4096 in one pass) at the
    largest tuned batch size not above `rows` (the smallest tuned size for smaller batches),
    or the library default."""
    entry = load().get(device_name(), {}).get(str(N), {}).get("schedules", {})
    if not entry:
        return dict(DEFAULT_SCHEDULE)
    sizes = sorted(int(r) for r in entry)
    fit = [r for r in sizes if r <= rows]
    pick = fit[-1] if fit else sizes[0]
    e = entry[str(pick)]
    return {k: int(e[k]) for k in DEFAULT_SCHEDULE}


def apply_schedule(prover, sch):
    prover.set_schedule(sch["wnaf"], sch["radix"], sch["waves_per_simd"])
    prover.set_streams(sch["streams"])
    prover.set_chain(bool(sch["chain"]))


def apply_tuned(prover, N, rows):
    sch = tuned_schedule(N, rows)
    apply_schedule(prover, sch)
    return sch


def schedule_key(sch):
    return "w%d/r%d/o%d/s%d%s" % (sch["wnaf"], sch["radix"], sch["waves_per_simd"], sch["streams"],
                                  "/chain" if sch["chain"] else "")


def candidate_schedules():
    out = []
    for w in WNAF_WIDTHS:
        for r in RADICES:
            for o in WAVES:
                for s in (2, 1):
                    out.append({"wnaf": w, "radix": r, "waves_per_simd": o, "streams": s, "chain": 0})
        out.append({"wnaf": w, "radix": 2, "waves_per_simd": 2, "streams": 1, "chain": 1})
    return out


def _time(prover, sid, poly_dev, out_dev, rows, reps):
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
    stream.synchronize()
    for x in (e0, e1):
        x.destroy()
    stream.destroy()
    return statistics.median(ms)


def tune(sizes=(1024, 2048, 4096), rows_list=(512, 64, 1), max_table_gb=96.0, reps=3, seed=11, xext_for=None,
         windows=CANDIDATE_WINDOWS, schedules=None):
    """xext_for(N) -> DeviceBuffer of xext_fft[16][N/8] (a synthetic setup by default)."""
    if xext_for is None:
        xext_for = lambda N: S.device_setup(N, S.random_secret(seed + N))
    schedules = schedules or candidate_schedules()
    results = load()
    dev = results.setdefault(device_name(), {})
    prover = Prover()
    rmax = max(rows_list)
    for N in sizes:
        xext = xext_for(N)
        poly = L.DeviceBuffer.from_array(S.random_polys(rmax, N, seed))
        out = L.DeviceBuffer(rmax * (N // 8) * 144)
        # 1. the MSM window, default schedule, largest batch
        apply_schedule(prover, DEFAULT_SCHEDULE)
        timings = {}
        for w in windows:
            if setup_table_bytes(N, w) > max_table_gb * (1 << 30):
                continue
            sid = prover.load_setup(xext, N, w)
            timings[w] = _time(prover, sid, poly, out, rmax, reps)
            prover.unload_setup(sid)
            print("autotune N=%d window=%d: %.3f ms / %d rows" % (N, w, timings[w], rmax), file=sys.stderr, flush=True)
        best_w = min(timings, key=timings.get)
        entry = {"window": best_w, "rows": rmax, "ms": {str(k): round(v, 3) for k, v in timings.items()},
                 "schedules": {}}
        # 2. the transform schedule per batch size, at that window
        sid = prover.load_setup(xext, N, best_w)
        for rows in rows_list:
            st = {}
            for sch in schedules:
                apply_schedule(prover, sch)
                st[schedule_key(sch)] = (_time(prover, sid, poly, out, rows, reps), sch)
            best_k = min(st, key=lambda k: st[k][0])
            e = dict(st[best_k][1])
            e["ms"] = {k: round(v[0], 3) for k, v in sorted(st.items(), key=lambda kv: kv[1][0])}
            entry["schedules"][str(rows)] = e
            print("autotune N=%d rows=%d: %s %.3f ms (default %s %.3f ms)" % (
                N, rows, best_k, st[best_k][0], schedule_key(DEFAULT_SCHEDULE),
                st[schedule_key(DEFAULT_SCHEDULE)][0]), file=sys.stderr, flush=True)
        prover.unload_setup(sid)
        apply_schedule(prover, DEFAULT_SCHEDULE)
        dev[str(N)] = entry
        for b in (xext, poly, out):
            b.free()
    prover.close()
    with open(TUNED, "w") as f:
        json.dump(results, f, indent=1, sort_keys=True)
    return results


def tune_mixed(rows_each=512, sizes=(1024, 2048, 4096), reps=3, seed=12, schedules=None):
    """The transform schedule for uneven batches: rows_each rows of every size in one
    fk20_prove_batch pass, each setup at its tuned window; stored under "mixed"."""
    schedules = schedules or candidate_schedules()
    results = load()
    dev = results.setdefault(device_name(), {})
    prover = Prover()
    items, bufs = [], []
    for N in sizes:
        xext = S.device_setup(N, S.random_secret(seed + N))
        sid = prover.load_setup(xext, N, tuned_window(N, 13))
        xext.free()
        pb = L.DeviceBuffer.from_array(S.random_polys(rows_each, N, seed))
        hb = L.DeviceBuffer(rows_each * (N // 8) * 144)
        bufs += [pb, hb]
        items.append((sid, rows_each, pb, hb))
    stream = L.Stream()
    e0, e1 = L.Event(), L.Event()
    st = {}
    for sch in schedules:
        apply_schedule(prover, sch)
        prover.prove_batch_device(items, stream=stream.handle)
        stream.synchronize()
        ms = []
        for _ in range(reps):
            e0.record(stream)
            prover.prove_batch_device(items, stream=stream.handle)
            e1.record(stream)
            ms.append(e0.elapsed_ms(e1))
        st[schedule_key(sch)] = (statistics.median(ms), sch)
    best_k = min(st, key=lambda k: st[k][0])
    e = dict(st[best_k][1])
    e["ms"] = {k: round(v[0], 3) for k, v in st.items()}
    total = rows_each * len(sizes)
    dev["mixed"] = {"sizes": list(sizes), "rows_each": rows_each, "schedules": {str(total): e}}
    print("autotune mixed %d rows: %s %.3f ms (default %s %.3f ms)" % (
        total, best_k, st[best_k][0], schedule_key(DEFAULT_SCHEDULE), st[schedule_key(DEFAULT_SCHEDULE)][0]),
        file=sys.stderr, flush=True)
    for x in (e0, e1):
        x.destroy()
    stream.destroy()
    for b in bufs:
        b.free()
    prover.close()
    with open(TUNED, "w") as f:
        json.dump(results, f, indent=1, sort_keys=True)
    return results


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--rows", default="512,64,1", help="batch sizes to tune the schedule for")
    ap.add_argument("--sizes", default="1024,2048,4096")
    ap.add_argument("--max-table-gb", type=float, default=96.0)
    ap.add_argument("--reps", type=int, default=3)
    ap.add_argument("--mixed-rows", type=int, default=512,
                    help="rows per size of the uneven batch whose schedule is tuned last (0: skip)")
    args = ap.parse_args()
    res = tune(tuple(int(x) for x in args.sizes.split(",")), tuple(int(x) for x in args.rows.split(",")),
               args.max_table_gb, args.reps)
    if args.mixed_rows:
        res = tune_mixed(args.mixed_rows, reps=args.reps)
    print(json.dumps(res))


if __name__ == "__main__":
    main()
