"""Serving: requests admitted while others run give the same results as the
same requests run alone (to 1e-14 of max|result|; synthesis is deterministic
per point, analysis reduces rings in a fixed order), under a tight HBM budget
that forces queueing; small scattered requests run as group launches with
results equal to the requests run alone; cancellation and host-injected
device errors report their error codes and release what they held."""
import threading
import time

import numpy as np
import pytest

import charm_amd
from charm_amd import shc
from charm_amd.serve import Server
from charm_amd.workloads import GM, R_EARTH, kaula_coeffs, scattered_points

CHARM_EMEM, CHARM_EFUNCARG = 1, 2

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


def test_serve_groups_small_scattered_requests():
    """Small scattered requests submitted together run as group launches;
    each grouped result equals the request run alone bit for bit (the same
    per-point arithmetic and the same order chunks), also for requests at
    different degrees and with points off the reference sphere."""
    NM = 360
    C, S = kaula_coeffs(NM, 71)
    model = shc.Shc.from_arrays(NM, C, S, GM, R_EARTH)
    srv = Server(8 * 2**30, 2)
    mid = srv.add_model(model)
    rng = np.random.default_rng(7)
    reqs = []
    for i in range(40):
        lat, lon, r = scattered_points(int(rng.integers(1, 30000)), 300 + i)
        if i % 4 == 0:
            r = np.full_like(r, R_EARTH)          # all on the sphere: the unit-ratio path
        reqs.append((int(rng.integers(2, NM + 1)), (lat, lon, r)))
    alone = [srv.shs_sctr(mid, n, *a).wait().copy() for n, a in reqs]
    st0 = srv.stats()
    assert st0["grouped"] == 0
    tickets = [srv.shs_sctr(mid, n, *a) for n, a in reqs]
    for t, ref, (n, _) in zip(tickets, alone, reqs):
        assert np.array_equal(t.wait(), ref), n
    st = srv.stats()
    print("serve stats", st, flush=True)
    assert st["completed"] == 2 * len(reqs) and st["failed"] == 0
    assert st["grouped"] > 0 and st["max_group"] >= 2
    assert st["launches"] - st0["launches"] < len(reqs)
    srv.close()


def test_serve_cancel_and_injected_errors():
    """Cancelled requests report CHARM_EFUNCARG, injected device errors their
    own code (CHARM_EMEM), and the server serves correctly afterwards; every
    byte charged to requests is released."""
    NM = 200
    C, S = kaula_coeffs(NM, 81)
    model = shc.Shc.from_arrays(NM, C, S, GM, R_EARTH)
    srv = Server(8 * 2**30, 2)
    mid = srv.add_model(model)
    base = srv.in_use()
    lat, lon, r = scattered_points(5000, 82)
    ref = srv.shs_sctr(mid, NM, lat, lon, r).wait().copy()
    ref_grid = srv.shs_grid(mid, NM, 400, R_EARTH).wait().copy()
    resident = srv.in_use()          # the model and the nmax-400 grid
    assert resident > base
    # a stream of large grid requests keeps the queue busy; cancel the tail
    ts = [srv.shs_grid(mid, NM, 400, R_EARTH) for _ in range(12)]
    cancelled = [t.cancel() for t in ts[6:]]
    n_cancel = 0
    for t, c in zip(ts, [False] * 6 + cancelled):
        try:
            assert np.array_equal(t.wait(), ref_grid)
        except charm_amd.CHarmError as e:
            assert c and e.code == CHARM_EFUNCARG
            n_cancel += 1
    assert n_cancel == sum(cancelled)
    # injected device errors: the next two launches fail with CHARM_EMEM
    srv.inject_error(2, CHARM_EMEM)
    codes = []
    for _ in range(3):
        t = srv.shs_sctr(mid, NM, lat, lon, r)
        try:
            assert np.array_equal(t.wait(), ref)
            codes.append(0)
        except charm_amd.CHarmError as e:
            codes.append(e.code)
    assert codes == [CHARM_EMEM, CHARM_EMEM, 0]
    st = srv.stats()
    assert st["failed"] == 2 and st["cancelled"] == n_cancel
    assert srv.in_use() == resident
    srv.close()
