"""Benchmark: FP64 point synthesis + analysis round trip on a Gauss-Legendre
grid at nmax = 2160 (2161 x 4322 points), Kaula-like seeded coefficients,
GM = 3.986004415e14, R = 6378136.3 m, r = R, inputs resident in HBM.

One step = one SHS (coefficients -> grid) plus one SHA (grid ->
coefficients) through the device-resident C ABI.  The metric counts
(nmax+1)^2 * nlat units per transform, i.e. 2 (nmax+1)^2 nlat per step.

Prints one JSON line (see the keys below).  --gpus N > 1 is not supported:
this build runs on one GPU.
"""
import argparse
import ctypes as ct
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_FP64_TFLOPS = 78.6   # MI355X vector FP64 (also its FP64 matrix peak)
PEAK_HBM_GBS = 8000.0


def kernel_ms(lib, name):
    tot, cnt = ct.c_double(), ct.c_long()
    lib.charm_amd_prof_read.argtypes = [ct.c_char_p, ct.POINTER(ct.c_double), ct.POINTER(ct.c_long)]
    lib.charm_amd_prof_read(name.encode(), ct.byref(tot), ct.byref(cnt))
    return tot.value, cnt.value


def cpu_baseline(N, C, S, lat, w, lon, R, GM, rows, threads):
    """The oracle (scalar C port of CHarm's algorithm) on a bounded sample of
    the same workload: SHS of `rows` GL rows at degree N on the host."""
    os.environ["OMP_NUM_THREADS"] = str(threads)
    from oracle.oracle import Oracle
    import numpy as np

    orc = Oracle("d")
    ring_idx = np.linspace(0, len(lat) - 1, rows).round().astype(np.int64)
    r = np.full(len(lat), R)
    t0 = time.perf_counter()
    orc.shs_grid(N, C, S, GM, R, lat, r, lon, gl=True, rings=ring_idx)
    dt = time.perf_counter() - t0
    return rows * (N + 1) ** 2 / dt, dt


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--nmax", type=int, default=2160)
    ap.add_argument("--cpu-rows", type=int, default=256)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--copy-mb", type=int, default=0,
                    help="before timing, copy this many MiB device-to-device once "
                         "(a reference kernel of known traffic for counter scaling)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this build runs on one GPU only (--gpus %d requested); "
              "multi-GPU sharding is described in DESIGN.md, not implemented." % args.gpus,
              file=sys.stderr)
        sys.exit(2)

    import numpy as np

    import charm_amd
    from charm_amd.device import DevArray, DeviceGrid, Stream
    from charm_amd.workloads import GM, R_EARTH, kaula_coeffs

    lib = charm_amd.lib()
    N = args.nmax
    C, S = kaula_coeffs(N, 2160)
    grid = DeviceGrid(N, R_EARTH)
    nlat, nlon = grid.nlat, grid.nlon
    cs = DevArray.from_numpy(np.stack([C, S], axis=1))
    f = DevArray((nlat, nlon))
    out = DevArray(cs.shape)
    stream = Stream()

    if args.copy_mb > 0:
        from charm_amd.device import hip, synchronize

        nb = args.copy_mb << 20
        src, dst = DevArray(nb // 8), DevArray(nb // 8)
        hip().hipMemcpy(dst.ptr, src.ptr, nb, 3)
        synchronize()
        del src, dst

    for _ in range(args.warmup):
        grid.shs(cs, N, GM, R_EARTH, f, stream)
        grid.sha(f, N, GM, R_EARTH, out, stream)
    stream.synchronize()
    o = out.numpy()
    rt_err = float(max(np.abs(o[:, 0] - C).max(), np.abs(o[:, 1] - S).max()) / np.abs(C).max())

    lib.charm_amd_prof_reset()
    lib.charm_amd_prof_enable(1)
    stream.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        grid.shs(cs, N, GM, R_EARTH, f, stream)
        grid.sha(f, N, GM, R_EARTH, out, stream)
    stream.synchronize()
    t1 = time.perf_counter()
    lib.charm_amd_prof_enable(0)

    elapsed = t1 - t0
    units_per_transform = float(N + 1) ** 2 * nlat
    value = 2.0 * units_per_transform * args.steps / elapsed
    kt = {k: kernel_ms(lib, k) for k in ("shs_grid_kernel", "fft_c2r", "sha_kernel",
                                          "sha_reduce_kernel", "fft_r2c")}
    avg = {k: (v[0] / v[1] if v[1] else float("nan")) for k, v in kt.items()}
    shs_ms = avg["shs_grid_kernel"] + avg["fft_c2r"]
    sha_ms = avg["fft_r2c"] + avg["sha_kernel"] + avg["sha_reduce_kernel"]

    # dominant kernel: FP64 recurrence; 8 flops per (n, m, ring pair)
    P = (N + 1) * (N + 2) // 2
    ring_pairs = (nlat + 1) // 2
    flops = 8.0 * P * ring_pairs
    dom = max(("shs_grid_kernel", "sha_kernel"), key=lambda k: avg[k])
    achieved = flops / (avg[dom] * 1e-3) / 1e12

    line = {
        "metric": "(nmax+1)^2*nlat/s FP64 SHS+SHA round trip (GL grid, device-resident)",
        "value": value,
        "unit": "(nmax+1)^2*nlat/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * elapsed / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded Kaula-like coefficients)",
        "config": {"workload": "GL SHS+SHA round trip nmax=%d, grid %dx%d, r=R" % (N, nlat, nlon),
                   "nmax": N, "nlat": nlat, "nlon": nlon, "parallelism": "single GPU"},
        "shs_units_per_s": units_per_transform / (shs_ms * 1e-3),
        "sha_units_per_s": units_per_transform / (sha_ms * 1e-3),
        "kernel_avg_ms": avg,
        "roundtrip_rel_err": rt_err,
        "roofline": {"bound": "mfma", "compute": "vector FP64 (VALU)", "kernel": dom,
                     "achieved": achieved, "peak": PEAK_FP64_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved / PEAK_FP64_TFLOPS, "traffic": None,
                     "flops_per_launch": flops},
    }
    if not args.no_cpu:
        threads = min(16, os.cpu_count() or 1)
        from charm_amd.crd import PointGridGL

        g = PointGridGL(N, R_EARTH)
        v, dt = cpu_baseline(N, C, S, g.lat, g.w, g.lon, R_EARTH, GM, args.cpu_rows, threads)
        line["cpu_baseline"] = {
            "value": v, "unit": "(nmax+1)^2*nlat/s", "cores": threads, "kind": "port",
            "sample": "SHS of %d of the %d GL rows at nmax=%d (Legendre sums + direct DFT per "
                      "row), oracle in double, %.1f s wall" % (args.cpu_rows, nlat, N, dt)}
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
