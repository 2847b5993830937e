"""Benchmarks of CHarm-style point synthesis (SHS) and analysis (SHA) in FP64
on one MI355X, on seeded synthetic inputs (GM = 3.986004415e14,
R = 6378136.3 m, Kaula-like coefficients: C_00 = 1, degree 1 zero,
C_nm, S_nm ~ N(0, (1e-5/n^2)^2)).

Workloads (--workload):
  gl2160  (default) SHS + SHA round trip on the Gauss-Legendre grid of
          nmax = 2160 (2161 x 4322 points), r = R, inputs resident in HBM.  One
          step = one SHS + one SHA; value = 2 (nmax+1)^2 nlat / step time.
          Also times SHS at 1e6 scattered points once (lat = asin U(-1,1),
          lon = U(0, 2 pi), r = R + U(0, 9000 m)) and the CPU baseline.
  gl10800 the same round trip at nmax = 10800 (10801 x 21602 points).
  batch   64 requests, nmax ~ U{90..2700}, each with its own coefficients and
          GL grid, half SHS and half SHA: per-call vs one batched launch.
          Beside it (key sctr_batch): 32 scattered SHS requests, nmax ~
          U{90..2700}, 1e3..1e5 points each, per-call vs batched.
  multi   256 coefficient sets at nmax = 720 on one GL grid (721 x 1442):
          synthesis by the FP64 matrix-core contraction vs one call per set.
  grad    second-order gradient (fxx..fzz, charm_shs_point_grad2) on the GL
          grid of nmax = 2160, device-resident; one step = one grad2 call;
          value = (nmax+1)^2 nlat / step time.  Beside it: grad1 and the
          plain synthesis on the same grid, grad1 at 2e5 scattered points,
          and the CPU baseline (the oracle's gradient at scattered points).
  io      model files of degree 2190 (the serving model): write and read
          tbl, dov, mtx and bin (charm_shc_write_*/read_*), then upload to
          HBM; coefficient pairs per second, host-side parsing threads.
  gfm     global gravity forward modelling (charm_gfm_global_density_const):
          a degree-540 shape, topography powers 1..5 (GL grids of degree
          540..2700, one SHS + power + SHA each), potential to degree 540;
          value = sum over powers of 2 (540+1)^2 nlat_p / wall time; CPU
          baseline: the oracle on a degree-360 shape, same powers.
  serve   Poisson arrivals against a resident nmax = 2190 model (60% scattered
          SHS of 1e3..1e6 points, 30% GL SHS on grids nmax ~ U{180..5400}
          with the model truncated to the grid, 10% GL SHA nmax ~ U{180..2160});
          p50/p99 latency.

Prints one JSON line.  --gpus N > 1 exits: this build runs on one GPU.
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
UNIT = "(nmax+1)^2*nlat/s"
# HBM bytes per launch at nmax = 2160 for the ring pairs per lane K the
# built-in rule picks (synthesis 3, analysis 6; analysis at K = 5 from an
# earlier run), measured with rocprofv3 --pmc FETCH_SIZE and WRITE_SIZE in
# separate runs and scaled by a 256 MiB device copy measured in the same runs
# (HBM bytes = 2.0 FETCH_SIZE + 1.0 WRITE_SIZE).  Analysis writes one partial
# triangle per ring-tile slot.
TRAFFIC_GL2160 = {("sha_kernel", 6): 345.7e6, ("sha_kernel", 5): 367.3e6,
                  ("shs_grid_kernel", 3): 189.2e6}


def last_ring_k(lib, name):
    """Ring pairs per lane of the last launch of a grid kernel."""
    return lib.charm_amd_last_ring_k(0 if name == "shs_grid_kernel" else 1)


def kernel_ms(lib, name):
    """(total ms, launches) recorded by the library's event timers."""
    tot, cnt = ct.c_double(), ct.c_long()
    lib.charm_amd_prof_read.argtypes = [ct.c_char_p, ct.POINTER(ct.c_double), ct.POINTER(ct.c_long)]
    lib.charm_amd_prof_read(name.encode(), ct.byref(tot), ct.byref(cnt))
    return tot.value, cnt.value


def base_line(metric, value, steps, warmup, ms, config):
    return {"metric": metric, "value": value, "unit": UNIT, "n_gpus": 1, "steps": steps,
            "warmup": warmup, "ms_per_step": ms, "higher_is_better": True, "scaling": "weak",
            "vs_baseline": None, "dtype": "f64",
            "data": "synthetic (seeded Kaula-like coefficients)", "config": config}


def cpu_threads():
    """Host threads for the CPU baseline: OMP_NUM_THREADS when set (the host's
    share of cores), else the physical cores this process may use."""
    from oracle.cpu_gl import physical_cores

    env = os.environ.get("OMP_NUM_THREADS", "")
    return int(env) if env.isdigit() and int(env) > 0 else physical_cores()


def cpu_baseline(N, C, S, GM, R, reps):
    """The repo's CPU restatement of CHarm's GL synthesis and analysis
    (oracle/cpu_gl.c: symmetry fold, 32-ring blocks in AVX-512 or AVX2
    loops, OpenMP over the host's cores, scipy.fft along the parallels) on
    the full round trip at degree N: median of `reps` runs after a warm-up."""
    import numpy as np
    from oracle.cpu_gl import CpuGL, cpu_model, physical_cores
    from oracle.oracle import Oracle

    threads = cpu_threads()
    cg = CpuGL(threads)
    lat, w, lon = Oracle("d").gl_grid(N)
    t_shs, t_sha = [], []
    for rep in range(reps + 1):
        t0 = time.perf_counter()
        f = cg.shs(N, C, S, GM, R, R, lat, len(lon))
        t1 = time.perf_counter()
        Co, So = cg.sha(N, f, lat, w, R, GM, R)
        t2 = time.perf_counter()
        if rep:
            t_shs.append(t1 - t0)
            t_sha.append(t2 - t1)
    err = float(max(np.abs(Co - C).max(), np.abs(So - S).max()) / np.abs(C).max())
    ts, ta = float(np.median(t_shs)), float(np.median(t_sha))
    units = float(N + 1) ** 2 * len(lat)
    return {"value": 2.0 * units / (ts + ta), "unit": UNIT, "cores": threads, "kind": "port",
            "label": "repo CPU restatement, not CHarm",
            "sample": "full SHS+SHA round trip at nmax=%d on the GL grid (%dx%d), median of %d "
                      "after a warm-up: oracle/cpu_gl.c (ring pairs folded, 32-ring blocks in "
                      "%s loops, OpenMP, subnormals flushed as CHarm's -ffast-math build) + "
                      "scipy.fft rows, %d threads" % (N, len(lat), len(lon), reps,
                                                       cg.isa.upper(), threads),
            "cpu_model": cpu_model(), "physical_cores_visible": physical_cores(),
            "shs_s": ts, "sha_s": ta, "shs_units_per_s": units / ts,
            "sha_units_per_s": units / ta, "roundtrip_rel_err": err}


# CHarm's published round-trip errors (docs/src/benchmarks.rst:60-65, read
# off its plot, about +-15%; 2160 and 10800 interpolated log-log in N)
PUBLISHED_EPS = {1000: (8.5e-13, 5.3e-14), 2160: (4.0e-12, 1.38e-13), 10800: (4.9e-11, 9.1e-13)}


def accuracy_protocol(N):
    """CHarm's accuracy protocol (bench/bench.c:192-300): C, S ~ U[-1, 1),
    mu = R = r = 1, GL synthesis and analysis at degree N; eps_max and
    eps_rms of the recovered coefficients."""
    import numpy as np

    from charm_amd.device import DevArray, DeviceGrid, Stream
    from charm_amd.workloads import uniform_coeffs

    C, S = uniform_coeffs(N, N)
    grid = DeviceGrid(N, 1.0)
    st = Stream()
    cs = DevArray.from_numpy(np.stack([C, S], axis=1))
    f = DevArray((grid.nlat, grid.nlon))
    out = DevArray(cs.shape)
    grid.shs(cs, N, 1.0, 1.0, f, st)
    grid.sha(f, N, 1.0, 1.0, out, st)
    st.synchronize()
    o = out.numpy()
    d = np.concatenate([o[:, 0] - C, o[:, 1] - S])
    res = {"nmax": N, "eps_max": float(np.abs(d).max()),
           "eps_rms": float(np.sqrt(2.0 * np.sum(d * d) / ((N + 1) * (N + 2)))),
           "coefficients": "U[-1,1) seed %d, mu=R=r=1" % N}
    if N in PUBLISHED_EPS:
        res["charm_published_eps_max"], res["charm_published_eps_rms"] = PUBLISHED_EPS[N]
    return res


def gl_roundtrip(args, N, with_sctr, with_cpu):
    import numpy as np

    import charm_amd
    from charm_amd.device import DevArray, DeviceGrid, Stream
    from charm_amd.workloads import GM, R_EARTH, kaula_coeffs

    lib = charm_amd.lib()
    C, S = kaula_coeffs(N, N)
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
    scale = np.abs(C).max()
    rt_err = float(max(np.abs(o[:, 0] - C).max(), np.abs(o[:, 1] - S).max()) / scale)
    print("# warm-up done (%d steps), round-trip rel err %.2e" % (args.warmup, rt_err),
          file=sys.stderr, flush=True)

    # timed region: HIP events bracket the two recurrence kernels only (each
    # event pair adds a few microseconds between dependent kernels)
    lib.charm_amd_prof_reset()
    lib.charm_amd_prof_enable(2)
    stream.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        grid.shs(cs, N, GM, R_EARTH, f, stream)
        grid.sha(f, N, GM, R_EARTH, out, stream)
    stream.synchronize()
    t1 = time.perf_counter()
    lib.charm_amd_prof_enable(0)
    kt = {k: kernel_ms(lib, k) for k in ("shs_grid_kernel", "sha_kernel")}
    # the other kernels of the round trip: a few more steps, not timed
    lib.charm_amd_prof_reset()
    lib.charm_amd_prof_enable(1)
    for _ in range(min(args.steps, 5)):
        grid.shs(cs, N, GM, R_EARTH, f, stream)
        grid.sha(f, N, GM, R_EARTH, out, stream)
    stream.synchronize()
    lib.charm_amd_prof_enable(0)
    kt.update({k: kernel_ms(lib, k) for k in ("fft_c2r", "sha_reduce_kernel", "fft_r2c")})

    elapsed = t1 - t0
    units = float(N + 1) ** 2 * nlat
    # (no sha_reduce_kernel launches when the analysis kernel reduces its slots)
    avg = {k: v[0] / v[1] for k, v in kt.items() if v[1]}
    shs_ms = avg["shs_grid_kernel"] + avg["fft_c2r"]
    sha_ms = avg["fft_r2c"] + avg["sha_kernel"] + avg.get("sha_reduce_kernel", 0.0)
    # dominant kernel: FP64 recurrence, 8 flops per (n, m, ring pair)
    P = (N + 1) * (N + 2) // 2
    flops = 8.0 * P * ((nlat + 1) // 2)
    dom = max(("shs_grid_kernel", "sha_kernel"), key=lambda k: avg[k])
    achieved = flops / (avg[dom] * 1e-3) / 1e12

    line = base_line("(nmax+1)^2*nlat/s FP64 SHS+SHA round trip (GL grid, device-resident)",
                     2.0 * units * args.steps / elapsed, args.steps, args.warmup,
                     1e3 * elapsed / args.steps,
                     {"workload": "GL SHS+SHA round trip nmax=%d, grid %dx%d, r=R" % (N, nlat, nlon),
                      "nmax": N, "nlat": nlat, "nlon": nlon, "parallelism": "single GPU"})
    line.update({
        "shs_units_per_s": units / (shs_ms * 1e-3),
        "sha_units_per_s": units / (sha_ms * 1e-3),
        "kernel_avg_ms": avg,
        "roundtrip_rel_err": rt_err,
        "accuracy_protocol": accuracy_protocol(N),
        "roofline": {"bound": "valu", "compute": "vector FP64 (VALU)", "kernel": dom,
                     "achieved": achieved, "peak": PEAK_FP64_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved / PEAK_FP64_TFLOPS,
                     "ring_pairs_per_lane": last_ring_k(lib, dom),
                     "traffic": args.traffic if args.traffic is not None else
                     (TRAFFIC_GL2160.get((dom, last_ring_k(lib, dom))) if N == 2160 else None),
                     "flops_per_launch": flops,
                     "frac_by_kernel": {k: flops / (avg[k] * 1e-3) / 1e12 / PEAK_FP64_TFLOPS
                                        for k in ("shs_grid_kernel", "sha_kernel")}},
    })
    del f, out
    if with_sctr:
        line.update(sctr_measure(N, cs, args.sctr_points, stream))
    if with_cpu:
        line["cpu_baseline"] = cpu_baseline(N, C, S, GM, R_EARTH, args.cpu_reps)
    return line


def sctr_measure(N, cs, npoint, stream):
    """SHS at `npoint` scattered points (device arrays), timed once after a
    warm-up call."""
    import charm_amd
    from charm_amd.device import DevArray, shs_sctr
    from charm_amd.workloads import GM, R_EARTH, scattered_points

    lib = charm_amd.lib()
    lat, lon, r = scattered_points(npoint, 12345)
    d = [DevArray.from_numpy(a) for a in (lat, lon, r)]
    out = DevArray(npoint)
    shs_sctr(*d, cs, N, GM, R_EARTH, out, stream)
    stream.synchronize()
    lib.charm_amd_prof_reset()
    lib.charm_amd_prof_enable(1)
    t0 = time.perf_counter()
    shs_sctr(*d, cs, N, GM, R_EARTH, out, stream)
    stream.synchronize()
    dt = time.perf_counter() - t0
    lib.charm_amd_prof_enable(0)
    kms, _ = kernel_ms(lib, "sctr_kernel")
    P = (N + 1) * (N + 2) // 2
    # per (n, m, point): 3 flops recurrence + 4 flops C/S accumulation
    return {"sctr_npoint": npoint, "sctr_nmax": N, "sctr_s": dt,
            "sctr_units_per_s": (N + 1) ** 2 * npoint / dt,
            "sctr_kernel_ms": kms,
            "sctr_kernel_frac_fp64_peak": 7.0 * P * npoint / (kms * 1e-3) / 1e12 / PEAK_FP64_TFLOPS}


def grad_bench(args):
    import numpy as np

    import charm_amd
    from charm_amd.device import (DERIV_GRAD1, DERIV_GRAD2, DevArray, DeviceGrid, Stream,
                                  shs_sctr_deriv)
    from charm_amd.workloads import GM, R_EARTH, kaula_coeffs, scattered_points

    lib = charm_amd.lib()
    N = 2160
    C, S = kaula_coeffs(N, N)
    grid = DeviceGrid(N, R_EARTH)
    nlat, nlon = grid.nlat, grid.nlon
    cs = DevArray.from_numpy(np.stack([C, S], axis=1))
    f6 = DevArray((6, nlat, nlon))
    f1 = DevArray((nlat, nlon))
    stream = Stream()
    for _ in range(args.warmup):
        grid.shs_deriv(cs, N, GM, R_EARTH, DERIV_GRAD2, f6, stream)
    stream.synchronize()
    t6 = f6.numpy()
    trace_err = float(np.abs(t6[0] + t6[3] + t6[5]).max() / np.abs(t6).max())
    print("# warm-up done, Laplace trace |fxx+fyy+fzz|/max %.2e" % trace_err, file=sys.stderr,
          flush=True)

    def timed(fn, steps):
        lib.charm_amd_prof_reset()
        lib.charm_amd_prof_enable(1)
        stream.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            fn()
        stream.synchronize()
        dt = time.perf_counter() - t0
        lib.charm_amd_prof_enable(0)
        return dt / steps

    g2 = timed(lambda: grid.shs_deriv(cs, N, GM, R_EARTH, DERIV_GRAD2, f6, stream), args.steps)
    k_shs = kernel_ms(lib, "shs_grid_kernel")
    k_fft = kernel_ms(lib, "fft_c2r")
    k_cmb = kernel_ms(lib, "grad_combine_kernel")
    ntab = k_shs[1] // args.steps
    g1 = timed(lambda: grid.shs_deriv(cs, N, GM, R_EARTH, DERIV_GRAD1, f6, stream),
               max(1, args.steps // 2))
    s0 = timed(lambda: grid.shs(cs, N, GM, R_EARTH, f1, stream), max(1, args.steps // 2))
    units = float(N + 1) ** 2 * nlat
    P = (N + 1) * (N + 2) // 2
    flops = 8.0 * P * ((nlat + 1) // 2)
    shs_avg = k_shs[0] / k_shs[1]
    achieved = flops / (shs_avg * 1e-3) / 1e12
    cmb_avg = k_cmb[0] / max(k_cmb[1], 1)
    cmb_bytes = (ntab + 6) * 8.0 * nlat * nlon + 3 * 8.0 * nlat
    line = base_line("(nmax+1)^2*nlat/s FP64 second-order gradient (6 outputs, GL grid, "
                     "device-resident)", units / g2, args.steps, args.warmup, 1e3 * g2,
                     {"workload": "charm_shs_point_grad2 on the GL grid nmax=%d, %dx%d, r=R"
                                  % (N, nlat, nlon),
                      "nmax": N, "nlat": nlat, "nlon": nlon, "parallelism": "single GPU"})
    line.update({
        "grad2_ms": 1e3 * g2, "grad1_ms": 1e3 * g1, "shs_ms": 1e3 * s0,
        "grad2_over_shs": g2 / s0, "syntheses_per_grad2": ntab,
        "kernel_avg_ms": {"shs_grid_kernel": shs_avg, "fft_c2r": k_fft[0] / max(k_fft[1], 1),
                          "grad_combine_kernel": cmb_avg},
        "laplace_trace_rel": trace_err,
        "combine_GBps": cmb_bytes / (cmb_avg * 1e-3) / 1e9,
        "roofline": {"bound": "valu", "compute": "vector FP64 (VALU)", "kernel": "shs_grid_kernel",
                     "achieved": achieved, "peak": PEAK_FP64_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved / PEAK_FP64_TFLOPS,
                     "traffic": TRAFFIC_GL2160.get(("shs_grid_kernel", last_ring_k(lib, "shs_grid_kernel"))),
                     "flops_per_launch": flops},
    })
    del f6, f1
    # grad1 at scattered points
    npt = args.grad_points
    lat, lon, r = scattered_points(npt, 777)
    d = [DevArray.from_numpy(a) for a in (lat, lon, r)]
    out = DevArray((3, npt))
    shs_sctr_deriv(*d, cs, N, GM, R_EARTH, DERIV_GRAD1, out, stream)
    stream.synchronize()
    gs = timed(lambda: shs_sctr_deriv(*d, cs, N, GM, R_EARTH, DERIV_GRAD1, out, stream), 1)
    line.update({"sctr_grad1_npoint": npt, "sctr_grad1_s": gs,
                 "sctr_grad1_units_per_s": (N + 1) ** 2 * npt / gs})
    if not args.no_cpu:
        threads = cpu_threads()
        os.environ["OMP_NUM_THREADS"] = str(threads)
        from oracle.oracle import Oracle

        orc = Oracle("d")
        ns = args.cpu_grad_points
        t0 = time.perf_counter()
        orc.shs_deriv_sctr("grad1", N, C, S, GM, R_EARTH, lat[:ns], lon[:ns], r[:ns])
        dt = time.perf_counter() - t0
        line["cpu_baseline"] = {
            "value": (N + 1) ** 2 * ns / dt, "unit": UNIT, "cores": threads, "kind": "port",
            "label": "repo CPU restatement (the scalar oracle), not CHarm",
            "sample": "grad1 at %d of the scattered points at nmax=%d (CHarm's per-term "
                      "derivative recurrences), oracle in double, %.1f s wall; compare "
                      "sctr_grad1_units_per_s" % (ns, N, dt)}
    return line


def io_bench(args):
    import tempfile

    import numpy as np

    from charm_amd import shc
    from charm_amd.device import DevArray, synchronize
    from charm_amd.workloads import GM, R_EARTH, kaula_coeffs

    N = 2190
    C, S = kaula_coeffs(N, 5)
    model = shc.Shc.from_arrays(N, C, S, GM, R_EARTH)
    pairs = (N + 1) * (N + 2) // 2
    res = {}
    with tempfile.TemporaryDirectory(dir=os.environ.get("TMPDIR", "/tmp")) as d:
        for ft in ("tbl", "dov", "mtx", "bin"):
            path = os.path.join(d, "model." + ft)
            t0 = time.perf_counter()
            model.to_file(ft, path)
            tw = time.perf_counter() - t0
            size = os.path.getsize(path)
            t0 = time.perf_counter()
            back = shc.Shc.from_file(ft, path)
            tr = time.perf_counter() - t0
            assert np.array_equal(back.c, C) and np.array_equal(back.s, S), ft
            t0 = time.perf_counter()
            dev = DevArray.from_numpy(np.stack([back.c, back.s], axis=1))
            synchronize()
            tu = time.perf_counter() - t0
            res[ft] = {"bytes": size, "write_s": tw, "read_s": tr, "upload_s": tu,
                       "read_MBps": size / tr / 1e6, "pairs_per_s": pairs / tr}
            del dev, back
            print("# %s: %.1f MB read in %.3f s" % (ft, size / 1e6, tr), file=sys.stderr,
                  flush=True)
    best = min(v["read_s"] + v["upload_s"] for k, v in res.items() if k != "bin")
    line = base_line("coefficient pairs/s, text model file read + upload to HBM (best text "
                     "format)", pairs / best, 1, 0, 1e3 * best,
                     {"workload": "degree-%d model (%d coefficient pairs) through tbl/dov/mtx/bin"
                                  % (N, pairs), "nmax": N, "parallelism": "host threads + 1 GPU"})
    line["unit"] = "pairs/s"
    line["dtype"] = "f64"
    line.update({"formats": res, "host_threads": min(16, os.cpu_count() or 1),
                 "note": "host-side parsing (strtod) over up to 16 threads; exact round trips"})
    return line


def gfm_bench(args):
    import numpy as np

    from charm_amd import gfm, shc
    from charm_amd.workloads import kaula_coeffs

    def shape_model(N, seed):
        C, S = kaula_coeffs(N, seed)
        C, S = C * 2.0e5, S * 2.0e5
        C[0] = 1.74e6
        return C, S

    def units(N, pmax):
        # per power: SHS of the degree-N shape and SHA to degree N on the
        # (pN+1)-ring grid
        return sum(2.0 * float(N + 1) ** 2 * (p * N + 1) for p in range(1, pmax + 1))

    N, pmax, R, rho, G, M = 540, 5, 1.73e6, 2500.0, 6.6743e-11, 7.346e22
    C, S = shape_model(N, 7)
    shape = shc.Shc.from_arrays(N, C, S)
    gfm.global_density_const(shape, N, R, rho, G, M, 1, pmax, N)   # warm-up (tables, plans)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        pot = gfm.global_density_const(shape, N, R, rho, G, M, 1, pmax, N)
    dt = (time.perf_counter() - t0) / args.steps
    line = base_line("(nmax+1)^2*nlat/s FP64 SHS+SHA inside global gravity forward modelling",
                     units(N, pmax) / dt, args.steps, 1, 1e3 * dt,
                     {"workload": "charm_gfm_global_density_const, shape nmax=%d, powers 1..%d "
                                  "(GL grids of degree %d..%d), potential nmax=%d"
                                  % (N, pmax, N, pmax * N, N),
                      "nmax": N, "parallelism": "single GPU"})
    line["c00"] = float(pot.c[0])
    if not args.no_cpu:
        threads = cpu_threads()
        os.environ["OMP_NUM_THREADS"] = str(threads)
        from oracle.oracle import Oracle

        n2 = 360
        C2, S2 = shape_model(n2, 7)
        t0 = time.perf_counter()
        Oracle("d").gfm_global((C2, S2), n2, R, rho, G, M, 1, pmax, n2)
        dt2 = time.perf_counter() - t0
        line["cpu_baseline"] = {
            "value": units(n2, pmax) / dt2, "unit": UNIT, "cores": threads, "kind": "port",
            "label": "repo CPU restatement (the scalar oracle), not CHarm",
            "sample": "the same modelling for a degree-%d shape (grids up to degree %d), "
                      "oracle SHS/SHA in double plus numpy powers, %.2f s wall" % (n2, pmax * n2, dt2)}
    return line


def _timed(fn, reps):
    fn()
    t0 = time.perf_counter()
    for _ in range(reps):
        fn()
    return (time.perf_counter() - t0) / reps


def grid_batch(mix, reps, label):
    """One mix of GL requests [(is_synthesis, N, seed)], each with its own
    Kaula coefficients and grid: device-resident per-call, one batch (eager
    and replayed from a HIP graph), and the plain host-array API per call
    (charm_shs_point / charm_sha_point) with grids, models and analysis inputs
    built before the timer, as CHarm's own benchmark times one call
    (bench/bench.c:225-279)."""
    import numpy as np

    from charm_amd import crd, sha, shc, shs
    from charm_amd.device import SHA_GRID, SHS_GRID, Batch, DevArray, DeviceGrid, Stream
    from charm_amd.workloads import GM, R_EARTH, kaula_coeffs

    st = Stream()
    reqs, host = [], []
    units = 0.0
    for is_shs, N, seed in mix:
        kind = SHS_GRID if is_shs else SHA_GRID
        C, S = kaula_coeffs(N, seed)
        g = DeviceGrid(N, R_EARTH)
        cs = DevArray.from_numpy(np.stack([C, S], axis=1))
        f = DevArray((g.nlat, g.nlon))
        if kind == SHA_GRID:
            g.shs(cs, N, GM, R_EARTH, f, st)
            cs = DevArray(cs.shape)
        reqs.append((kind, N, GM, R_EARTH, g, cs, f))
        units += float(N + 1) ** 2 * g.nlat
    st.synchronize()
    for (kind, N, _, _, g, _, f), (is_shs, _, seed) in zip(reqs, mix):
        pnt = crd.PointGridGL(N, R_EARTH)
        C, S = kaula_coeffs(N, seed)
        host.append((kind, N, pnt, shc.Shc.from_arrays(N, C, S, GM, R_EARTH),
                     None if kind == SHS_GRID else f.numpy()))
    print("# %s: %d requests set up" % (label, len(reqs)), file=sys.stderr, flush=True)

    def per_call():
        for kind, N, mu, R, g, cs, f in reqs:
            if kind == SHS_GRID:
                g.shs(cs, N, mu, R, f, st)
            else:
                g.sha(f, N, mu, R, cs, st)
        st.synchronize()

    def outputs():
        return [(r[6].numpy() if r[0] == SHS_GRID else r[5].numpy()) for r in reqs]

    def host_api():
        for kind, N, pnt, m, fh in host:
            if kind == SHS_GRID:
                shs.point(pnt, m, N)
            else:
                sha.point(pnt, fh, N, GM, R_EARTH)

    t_call = _timed(per_call, reps)
    ref = outputs()
    b_eager = Batch(reqs, use_graph=False)
    t_batch = _timed(lambda: (b_eager.run(st), st.synchronize()), reps)
    b_graph = Batch(reqs, use_graph=True)
    t_graph = _timed(lambda: (b_graph.run(st), st.synchronize()), reps)
    # synthesis results are bitwise those of the calls alone; analysis sums its
    # ring tiles in another grouping (the batch's ring pairs per lane)
    got = outputs()
    same = all(np.array_equal(x, y) for x, y, r in zip(ref, got, reqs) if r[0] == SHS_GRID)
    sha_diff = max([float(np.abs(y - x).max() / np.abs(x).max())
                    for x, y, r in zip(ref, got, reqs) if r[0] == SHA_GRID] + [0.0])
    t_host = _timed(host_api, 1)
    print("# %s: timed per-call, batched and host API" % label, file=sys.stderr, flush=True)
    best = min(t_batch, t_graph)
    return units, best, {
        "requests": len(reqs), "units": units,
        "per_call_dev_units_per_s": units / t_call, "batch_units_per_s": units / t_batch,
        "batch_graph_units_per_s": units / t_graph, "graph_used": b_graph.uses_graph,
        "speedup_vs_per_call_dev": t_call / best,
        "host_api_per_call_s": t_host, "speedup_vs_host_api_per_call": t_host / best,
        "shs_bitwise_equal_to_alone": bool(same), "sha_max_rel_diff_vs_alone": sha_diff}


def sctr_batch(degs, npts, seed0, reps, label):
    """Scattered synthesis requests (degree degs[i], npts[i] points with
    r = R + U(0, 9000 m), own coefficients): device-resident per-call, one
    batch replayed from a HIP graph (the batch sorts each point set once, at
    creation), and the plain host-array API per call with the point
    structures and models built before the timer."""
    import numpy as np

    from charm_amd import crd, shc, shs
    from charm_amd.device import SHS_SCTR, Batch, DevArray, Stream, shs_sctr
    from charm_amd.workloads import GM, R_EARTH, kaula_coeffs, scattered_points

    st = Stream()
    reqs, host = [], []
    units = 0.0
    for i, (N, P) in enumerate(zip(degs, npts)):
        N, P = int(N), int(P)
        C, S = kaula_coeffs(N, seed0 + i)
        pts = scattered_points(P, seed0 + 1000 + i)
        reqs.append((SHS_SCTR, N, GM, R_EARTH, tuple(DevArray.from_numpy(a) for a in pts),
                     DevArray.from_numpy(np.stack([C, S], axis=1)), DevArray(P)))
        host.append((N, crd.PointSctr(*pts), shc.Shc.from_arrays(N, C, S, GM, R_EARTH)))
        units += float(N + 1) ** 2 * P

    def per_call():
        for kind, N, mu, R, pts, cs, f in reqs:
            shs_sctr(*pts, cs, N, mu, R, f, st)
        st.synchronize()

    def host_api():
        for N, pnt, m in host:
            shs.point(pnt, m, N)

    t_call = _timed(per_call, reps)
    ref = [r[6].numpy() for r in reqs]
    b = Batch(reqs, use_graph=True)
    t_batch = _timed(lambda: (b.run(st), st.synchronize()), reps)
    same = all(np.array_equal(x, r[6].numpy()) for x, r in zip(ref, reqs))
    t_host = _timed(host_api, 1)
    print("# %s: timed scattered per-call, batched and host API" % label, file=sys.stderr,
          flush=True)
    return {"requests": len(reqs), "points": int(np.sum(npts)),
            "per_call_dev_units_per_s": units / t_call,
            "batch_graph_units_per_s": units / t_batch,
            "speedup_vs_per_call_dev": t_call / t_batch,
            "host_api_per_call_s": t_host, "speedup_vs_host_api_per_call": t_host / t_batch,
            "graph_used": b.uses_graph, "bitwise_equal_to_alone": bool(same)}


def batch_bench(args):
    """C4a: the 64-request mix, nmax ~ U{90..2700}; beside it a mix of 256
    small grids (nmax ~ U{16..256}), 32 scattered sets (1e3..1e5 points,
    nmax ~ U{90..2700}) and 256 small scattered sets (100..2000 points, nmax
    ~ U{16..360})."""
    import numpy as np

    from charm_amd.workloads import batch_mix

    units, best, res = grid_batch(batch_mix(), args.steps, "batch")
    line = base_line("(nmax+1)^2*nlat/s FP64 batched SHS/SHA, 64 requests nmax~U{90..2700}",
                     units / best, args.steps, 1, 1e3 * best,
                     {"workload": "64 GL requests (32 SHS, 32 SHA), nmax~U{90..2700} seed 64, "
                                  "own coefficients and grid each", "requests": 64,
                      "parallelism": "single GPU"})
    line.update(res)
    rng = np.random.default_rng(256)
    small = [(i % 2 == 0, int(N), 5000 + i) for i, N in enumerate(rng.integers(16, 257, 256))]
    line["small_grids"] = grid_batch(small, args.steps, "small grids")[2]
    line["small_grids"]["mix"] = "256 GL requests (128 SHS, 128 SHA), nmax~U{16..256} seed 256"
    rng = np.random.default_rng(65)
    degs = rng.integers(90, 2701, 32)
    npts = np.round(10.0 ** rng.uniform(3, 5, 32)).astype(int)
    line["sctr_batch"] = sctr_batch(degs, npts, 2000, args.steps, "sctr batch")
    line["sctr_batch"]["mix"] = "32 scattered sets, nmax~U{90..2700}, 1e3..1e5 points seed 65"
    rng = np.random.default_rng(257)
    degs = rng.integers(16, 361, 256)
    npts = rng.integers(100, 2001, 256)
    line["small_sctr_batch"] = sctr_batch(degs, npts, 7000, args.steps, "small sctr batch")
    line["small_sctr_batch"]["mix"] = ("256 scattered sets, nmax~U{16..360}, 100..2000 points "
                                       "seed 257")
    return line


def multi_bench(args):
    """C4b: nset coefficient sets at N on one shared GL grid, synthesis by the
    FP64 matrix-core contraction vs one synthesis call per set."""
    import numpy as np

    import charm_amd
    from charm_amd.device import DevArray, DeviceGrid, Stream
    from charm_amd.workloads import GM, R_EARTH, kaula_coeffs

    lib = charm_amd.lib()
    N, nset = args.multi_nmax, args.multi_sets
    grid = DeviceGrid(N, R_EARTH)
    nlat, nlon = grid.nlat, grid.nlon
    P = (N + 1) * (N + 2) // 2
    rng = np.random.default_rng(720)
    sig = np.zeros(P)
    from charm_amd.workloads import degree_order
    n_of, m_of = degree_order(N)
    sig[n_of >= 2] = 1e-5 / n_of[n_of >= 2] ** 2.0
    cs = rng.standard_normal((nset, P, 2)) * sig[None, :, None]
    cs[:, :, 1][:, m_of == 0] = 0.0
    cs[:, 0, 0] = 1.0
    dcs = DevArray.from_numpy(cs)
    out = DevArray((nset, nlat, nlon))
    st = Stream()
    for _ in range(args.warmup):
        grid.shs_multi(dcs, N, nset, GM, R_EARTH, out, st)
    st.synchronize()
    lib.charm_amd_prof_reset()
    lib.charm_amd_prof_enable(1)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        grid.shs_multi(dcs, N, nset, GM, R_EARTH, out, st)
    st.synchronize()
    t_multi = (time.perf_counter() - t0) / args.steps
    lib.charm_amd_prof_enable(0)
    kms, kcnt = kernel_ms(lib, "shs_multi_kernel")
    fms, fcnt = kernel_ms(lib, "fft_c2r_multi")
    kavg, favg = kms / max(kcnt, 1), fms / max(fcnt, 1)
    # one call per set on the same grid (device-resident)
    views = [DevArray.from_numpy(cs[i]) for i in range(nset)]
    f1 = DevArray((nlat, nlon))
    grid.shs(views[0], N, GM, R_EARTH, f1, st)
    st.synchronize()
    t0 = time.perf_counter()
    for i in range(nset):
        grid.shs(views[i], N, GM, R_EARTH, f1, st)
    st.synchronize()
    t_single = time.perf_counter() - t0
    got = out.numpy()[nset - 1]
    ref = f1.numpy()
    err = float(np.abs(got - ref).max() / np.abs(ref).max())
    units = float(nset) * (N + 1) ** 2 * nlat
    flops = 4.0 * nset * P * ((nlat + 1) // 2)
    achieved = flops / (kavg * 1e-3) / 1e12
    line = base_line("(nmax+1)^2*nlat/s FP64 SHS of %d sets on one GL grid (matrix cores)" % nset,
                     units / t_multi, args.steps, args.warmup, 1e3 * t_multi,
                     {"workload": "%d coefficient sets, nmax=%d, one GL grid %dx%d, r=R" % (nset, N, nlat, nlon),
                      "sets": nset, "nmax": N, "parallelism": "single GPU"})
    line.update({"kernel_avg_ms": {"shs_multi_kernel": kavg, "fft_c2r_multi": favg},
                 "per_set_calls_units_per_s": units / t_single,
                 "speedup_vs_per_set_calls": t_single / t_multi,
                 "last_set_rel_diff_vs_single": err,
                 "roofline": {"bound": "mfma", "compute": "FP64 matrix cores (v_mfma_f64_16x16x4f64)",
                              "kernel": "shs_multi_kernel", "achieved": achieved, "peak": PEAK_FP64_TFLOPS,
                              "unit": "TFLOP/s", "frac": achieved / PEAK_FP64_TFLOPS, "traffic": None,
                              "flops_per_launch": flops}})
    return line


def serve_bench(args):
    import numpy as np

    from charm_amd import shc
    from charm_amd.serve import Server
    from charm_amd.workloads import GM, R_EARTH, kaula_coeffs, scattered_points

    NM = 2190
    C, S = kaula_coeffs(NM, 2190)
    srv = Server(200 * 2**30, args.inflight)
    mid = srv.add_model(shc.Shc.from_arrays(NM, C, S, GM, R_EARTH))
    rng = np.random.default_rng(2190)
    nreq = args.requests
    reqs, units = [], 0.0
    for i in range(nreq):
        u = rng.random()
        if u < 0.6:
            npt = int(round(10 ** rng.uniform(3, 6)))
            reqs.append(("sctr", NM, scattered_points(npt, 5000 + i)))
            units += (NM + 1) ** 2 * npt
        elif u < 0.9:
            Ng = int(rng.integers(180, 5401))
            reqs.append(("shs", min(Ng, NM), Ng))
            units += (min(Ng, NM) + 1) ** 2 * (Ng + 1)
        else:
            N = int(rng.integers(180, 2161))
            reqs.append(("sha", N, N))
            units += (N + 1) ** 2 * (N + 1)
    sha_in = {i: srv.shs_grid(mid, n, a, R_EARTH).wait().copy()
              for i, (kind, n, a) in enumerate(reqs) if kind == "sha"}
    srv.shs_sctr(mid, NM, *scattered_points(1000, 1)).wait()
    srv.shs_grid(mid, 180, 180, R_EARTH).wait()
    print("# serve: %d requests prepared" % nreq, file=sys.stderr, flush=True)
    gaps = rng.exponential(1.0 / args.rate, nreq)
    tickets = []
    t0 = time.perf_counter()
    t_next = t0
    for i, (kind, n, a) in enumerate(reqs):
        t_next += gaps[i]
        while time.perf_counter() < t_next:
            time.sleep(0.0002)
        if kind == "sctr":
            tickets.append(srv.shs_sctr(mid, n, *a))
        elif kind == "shs":
            tickets.append(srv.shs_grid(mid, n, a, R_EARTH))
        else:
            tickets.append(srv.sha_grid(sha_in[i], n, R_EARTH, GM, R_EARTH))
    outs = [t.wait().copy() for t in tickets]
    wall = time.perf_counter() - t0
    lat = np.array([t.latency_ms for t in tickets])
    st_poisson = srv.stats()
    print("# serve: Poisson run done", file=sys.stderr, flush=True)

    def submit(i):
        kind, n, a = reqs[i]
        if kind == "sctr":
            return srv.shs_sctr(mid, n, *a)
        if kind == "shs":
            return srv.shs_grid(mid, n, a, R_EARTH)
        return srv.sha_grid(sha_in[i], n, R_EARTH, GM, R_EARTH)

    # every request again, one at a time: its result must equal the served one
    t0 = time.perf_counter()
    alone = [submit(i).wait().copy() for i in range(nreq)]
    t_seq = time.perf_counter() - t0
    same = all(np.array_equal(x, y) for x, y in zip(outs, alone))
    # closed loop: everything submitted at once
    t0 = time.perf_counter()
    ts = [submit(i) for i in range(nreq)]
    sat = [t.wait().copy() for t in ts]
    t_sat = time.perf_counter() - t0
    same = same and all(np.array_equal(x, y) for x, y in zip(sat, alone))
    st_all = srv.stats()
    srv.close()
    # the plain host-array API per call, grids and point structures built
    # before the timer
    from charm_amd import crd, sha, shs

    model = shc.Shc.from_arrays(NM, C, S, GM, R_EARTH)
    calls = []
    for i, (kind, n, a) in enumerate(reqs):
        if kind == "sctr":
            calls.append((shs.point, (crd.PointSctr(*a), model, n)))
        elif kind == "shs":
            calls.append((shs.point, (crd.PointGridGL(a, R_EARTH), model, n)))
        else:
            calls.append((sha.point, (crd.PointGridGL(a, R_EARTH), sha_in[i], n, GM, R_EARTH)))
    calls[0][0](*calls[0][1])
    t0 = time.perf_counter()
    for fn, a in calls:
        fn(*a)
    t_host = time.perf_counter() - t0
    print("# serve: sequential, closed-loop and host API runs done", file=sys.stderr, flush=True)
    line = base_line("(nmax+1)^2*nlat/s FP64 served SHS/SHA mix, p50/p99 latency", units / wall,
                     nreq, 2, 1e3 * wall / nreq,
                     {"workload": "Poisson arrivals (rate %.1f/s) on a resident nmax=%d model: 60%% "
                                  "scattered SHS 1e3..1e6 points (log-uniform), 30%% GL SHS grid "
                                  "nmax~U{180..5400} (model truncated to min(grid, %d)), 10%% GL "
                                  "SHA nmax~U{180..2160}" % (args.rate, NM, NM),
                      "requests": nreq, "hbm_budget_GiB": 200, "max_inflight": args.inflight,
                      "parallelism": "single GPU"})
    line.update({"p50_ms": float(np.percentile(lat, 50)), "p99_ms": float(np.percentile(lat, 99)),
                 "mean_ms": float(lat.mean()), "wall_s": wall,
                 "served_equal_to_alone": bool(same),
                 "closed_loop": {"wall_s": t_sat, "units_per_s": units / t_sat,
                                 "speedup_vs_one_at_a_time_served": t_seq / t_sat,
                                 "speedup_vs_host_api_per_call": t_host / t_sat},
                 "one_at_a_time_served_s": t_seq, "host_api_per_call_s": t_host,
                 "server_stats_poisson": st_poisson, "server_stats_total": st_all,
                 "note": "units of scattered requests count (nmax+1)^2 per point; results "
                         "copied back to host memory inside the latency; the Poisson value is "
                         "bounded by the arrival rate, the closed loop is the server's "
                         "throughput"})
    return line


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--workload", default="gl2160",
                    choices=("gl2160", "gl10800", "batch", "multi", "grad", "io", "gfm",
                             "serve"))
    ap.add_argument("--cpu-reps", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-sctr", action="store_true")
    ap.add_argument("--sctr-points", type=int, default=1000000)
    ap.add_argument("--grad-points", type=int, default=200000)
    ap.add_argument("--cpu-grad-points", type=int, default=3000)
    ap.add_argument("--multi-sets", type=int, default=256)
    ap.add_argument("--multi-nmax", type=int, default=720)
    ap.add_argument("--requests", type=int, default=40)
    ap.add_argument("--rate", type=float, default=20.0)
    ap.add_argument("--inflight", type=int, default=4)
    ap.add_argument("--traffic", type=float, default=None,
                    help="HBM bytes per launch of the dominant kernel (rocprofv3 FETCH_SIZE and "
                         "WRITE_SIZE, scaled by a copy kernel) to report in the roofline object")
    ap.add_argument("--copy-mb", type=int, default=0,
                    help="before timing, copy this many MiB device-to-device once "
                         "(a reference kernel of known traffic for counter scaling)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this build runs on one GPU only (--gpus %d requested); "
              "sharding across GPUs is designed, not implemented." % args.gpus,
              file=sys.stderr)
        sys.exit(2)
    if args.workload == "gl2160":
        line = gl_roundtrip(args, 2160, not args.no_sctr, not args.no_cpu)
    elif args.workload == "gl10800":
        line = gl_roundtrip(args, 10800, False, False)
    elif args.workload == "batch":
        line = batch_bench(args)
    elif args.workload == "multi":
        line = multi_bench(args)
    elif args.workload == "grad":
        line = grad_bench(args)
    elif args.workload == "io":
        line = io_bench(args)
    elif args.workload == "gfm":
        line = gfm_bench(args)
    else:
        line = serve_bench(args)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
