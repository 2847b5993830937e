"""Serving: requests admitted while others run give the same results as the
same requests run alone (to 1e-14 of max|result|; synthesis is deterministic
per point, analysis reduces rings in a fixed order), under a tight HBM budget
that forces queueing."""
import threading
import time

import numpy as np
import pytest

import charm_amd
from charm_amd import shc
from charm_amd.serve import Server
from charm_amd.workloads import GM, R_EARTH, kaula_coeffs, scattered_points

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def need_gpu():
    if charm_amd.lib().charm_amd_device_count() < 1:
        pytest.fail("no HIP device visible: the GPU tests need an MI355X")


def close(a, b, tol=1e-14):
    return np.abs(a - b).max() <= tol * max(np.abs(b).max(), 1e-300)


def test_serve_alone_vs_mix():
    NM = 300
    C, S = kaula_coeffs(NM, 31)
    model = shc.Shc.from_arrays(NM, C, S, GM, R_EARTH)
    mix = []
    rng = np.random.default_rng(5)
    for i in range(24):
        k = i % 3
        if k == 0:
            lat, lon, r = scattered_points(int(rng.integers(1000, 20000)), 100 + i)
            mix.append(("sctr", int(rng.integers(50, NM + 1)), (lat, lon, r)))
        elif k == 1:
            N = int(rng.integers(30, 400))
            mix.append(("shs", min(N, NM), N))
        else:
            N = int(rng.integers(30, 250))
            mix.append(("sha", N, N))
    # inputs of the analyses: synthesized fields
    fields = {}
    alone = []
    srv = Server(8 * 2**30, 4)
    mid = srv.add_model(model)
    for j, (kind, n, arg) in enumerate(mix):
        if kind == "sha":
            fields[j] = srv.shs_grid(mid, n, n, R_EARTH).wait().copy()
    # alone: one at a time
    for j, (kind, n, arg) in enumerate(mix):
        if kind == "sctr":
            t = srv.shs_sctr(mid, n, *arg)
        elif kind == "shs":
            t = srv.shs_grid(mid, n, arg, R_EARTH)
        else:
            t = srv.sha_grid(fields[j], n, R_EARTH, GM, R_EARTH)
        alone.append(t.wait().copy())
    # in the mix: all submitted from two threads, interleaved, then waited
    tickets = [None] * len(mix)

    def submit(idx):
        for j in idx:
            kind, n, arg = mix[j]
            if kind == "sctr":
                tickets[j] = srv.shs_sctr(mid, n, *arg)
            elif kind == "shs":
                tickets[j] = srv.shs_grid(mid, n, arg, R_EARTH)
            else:
                tickets[j] = srv.sha_grid(fields[j], n, R_EARTH, GM, R_EARTH)
            time.sleep(0.0005)

    th = [threading.Thread(target=submit, args=(list(range(p, len(mix), 2)),)) for p in (0, 1)]
    for x in th:
        x.start()
    for x in th:
        x.join()
    for j, t in enumerate(tickets):
        got = t.wait()
        assert close(got, alone[j]), (j, mix[j][0], mix[j][1])
        assert t.latency_ms > 0
    srv.close()


def test_serve_sha_roundtrip_and_budget_queueing():
    """A budget that holds about one large request at a time: requests queue
    and still complete correctly."""
    N = 512
    C, S = kaula_coeffs(N, 41)
    model = shc.Shc.from_arrays(N, C, S, GM, R_EARTH)
    srv = Server(64 * 2**20, 4)
    mid = srv.add_model(model)
    f = srv.shs_grid(mid, N, N, R_EARTH).wait().copy()
    ts = [srv.sha_grid(f, N, R_EARTH, GM, R_EARTH) for _ in range(4)]
    scale = np.abs(C).max()
    for t in ts:
        out = t.wait()
        assert np.abs(out[:, 0] - C).max() <= 1e-12 * scale
        assert np.abs(out[:, 1] - S).max() <= 1e-12 * scale
    srv.close()
