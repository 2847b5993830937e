"""Ring-tile autotuner (offline): times the grid kernels at every number of
ring pairs per lane K on device-resident Gauss-Legendre grids of a range of
degrees and fits the K rule by ring count.

    python -m charm_amd.tune [nmax ...]

For each degree it prints the kernel time (HIP events around the launches,
fastest of two rounds) of synthesis at K = 1..4 and analysis at K = 2, 3, 4, 5, 6, 8
(the analysis kernel plus the reduction of its partial slots, whose number
depends on K), and
at the end the fitted rule: for each kernel the ring-pair counts up to which
each K is fastest.  The library's built-in rule (point_api.hip, fitted_k) was
fitted this way on MI355X; CHARM_AMD_TUNE=1 refines it at run time per
ring-count bucket, and CHARM_AMD_SHS_K / CHARM_AMD_SHA_K force a value.
"""
import ctypes as _ct
import os
import sys

import numpy as np

DEFAULT_DEGREES = (64, 128, 200, 300, 450, 600, 800, 1000, 1300, 1700, 2160, 2700, 3500, 4500,
                   5400, 7000, 8200, 10800)
KINDS = ((0, (1, 2, 3, 4), "CHARM_AMD_SHS_K", ("shs_grid_kernel",)),
         (1, (2, 3, 4, 5, 6, 8), "CHARM_AMD_SHA_K", ("sha_kernel", "sha_reduce_kernel")))


def _kernel_ms(lib, name, calls):
    """Average time per transform of the named kernel (0 if it did not run)."""
    tot, cnt = _ct.c_double(), _ct.c_long()
    lib.charm_amd_prof_read.argtypes = [_ct.c_char_p, _ct.POINTER(_ct.c_double),
                                        _ct.POINTER(_ct.c_long)]
    lib.charm_amd_prof_read(name.encode(), _ct.byref(tot), _ct.byref(cnt))
    return tot.value / calls


def sweep(degrees=DEFAULT_DEGREES, out=sys.stdout):
    """{(kind, nmax): {K: ms}} for the given degrees."""
    from . import lib
    from .device import DevArray, DeviceGrid, Stream
    from .workloads import GM, R_EARTH, kaula_coeffs

    L = lib()
    L.charm_amd_prof_enable.argtypes = [_ct.c_int]
    st = Stream()
    res = {}
    for N in degrees:
        C, S = kaula_coeffs(N, 5)
        g = DeviceGrid(N, R_EARTH)
        cs = DevArray.from_numpy(np.stack([C, S], axis=1))
        f = DevArray((g.nlat, g.nlon))
        co = DevArray(cs.shape)
        reps = 3 if N > 5000 else 8

        def run(kind):
            if kind == 0:
                g.shs(cs, N, GM, R_EARTH, f, st)
            else:
                g.sha(f, N, GM, R_EARTH, co, st)

        for kind, ks, env, names in KINDS:
            t = res.setdefault((kind, N), {})
            old = os.environ.get(env)
            try:
                for rnd in range(2):
                    for k in ks:
                        os.environ[env] = str(k)
                        run(kind)
                        run(kind)
                        st.synchronize()
                        L.charm_amd_prof_reset()
                        L.charm_amd_prof_enable(1)
                        for _ in range(reps):
                            run(kind)
                        st.synchronize()
                        L.charm_amd_prof_enable(0)
                        ms = sum(_kernel_ms(L, nm, reps) for nm in names)
                        t[k] = min(t.get(k, 1e30), ms)
            finally:
                if old is None:
                    os.environ.pop(env, None)
                else:
                    os.environ[env] = old
        print("nmax %5d  synthesis %s  analysis %s" % (
            N, " ".join("K%d %.4f" % kv for kv in sorted(res[(0, N)].items())),
            " ".join("K%d %.4f" % kv for kv in sorted(res[(1, N)].items()))), file=out, flush=True)
        del g, cs, f, co
    return res


def fit(res):
    """Rule per kind: [(largest ring-pair count, K)] from the fastest K at
    each degree (GL grids have nmax // 2 + 1 ring pairs)."""
    rules = {}
    for kind, _, _, _ in KINDS:
        pts = sorted((N // 2 + 1, min(t, key=t.get)) for (kd, N), t in res.items() if kd == kind)
        rule = []
        for nring, k in pts:
            if rule and rule[-1][1] == k:
                rule[-1] = (nring, k)
            else:
                rule.append((nring, k))
        rules[kind] = rule
    return rules


def main(argv=None):
    argv = sys.argv[1:] if argv is None else argv
    degrees = [int(a) for a in argv] or list(DEFAULT_DEGREES)
    res = sweep(degrees)
    for kind, rule in fit(res).items():
        print("%s: %s" % ("synthesis" if kind == 0 else "analysis",
                          ", ".join("K=%d up to %d ring pairs" % (k, n) for n, k in rule)))


if __name__ == "__main__":
    main()
