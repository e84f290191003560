"""Plain HBM copy through the library's kCopy kernel, for calibrating
hardware byte counters: a copy of B bytes reads B and writes B bytes, so
rocprofv3's FETCH_SIZE / WRITE_SIZE for these dispatches divided by B gives
the factor that turns those counters into HBM bytes for the other kernels.

    python -m kripke_amd.hbm_copy [--gib 4] [--reps 5]

Prints one JSON line: bytes per copy, dispatches, and the achieved copy
bandwidth (read + write) from the host clock.
"""
import argparse
import ctypes
import json
import time

from .capi import Context, check, lib


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gib", type=float, default=4.0, help="bytes per buffer, GiB")
    ap.add_argument("--reps", type=int, default=5)
    args = ap.parse_args()
    n = int(args.gib * (1 << 30)) // 8
    ctx = Context(0)
    L = lib()
    a, b = ctypes.c_void_p(), ctypes.c_void_p()
    check(L.kr_malloc(ctx.h, n * 8, ctypes.byref(a)), "kr_malloc")
    check(L.kr_malloc(ctx.h, n * 8, ctypes.byref(b)), "kr_malloc")
    try:
        check(L.kr_kconst(ctx.h, a, n, 1.0), "kr_kconst")
        check(L.kr_kcopy(ctx.h, b, a, n), "kr_kcopy")  # warm-up dispatch
        ctx.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.reps):
            check(L.kr_kcopy(ctx.h, b, a, n), "kr_kcopy")
        ctx.synchronize()
        dt = (time.perf_counter() - t0) / args.reps
        print(json.dumps({"kernel": "k_copy", "bytes_read_per_dispatch": n * 8, "bytes_written_per_dispatch": n * 8,
                          "dispatches": args.reps + 1, "copy_GBps": 2 * n * 8 / dt / 1e9}))
    finally:
        L.kr_free(ctx.h, a)
        L.kr_free(ctx.h, b)
        ctx.close()


if __name__ == "__main__":
    main()
