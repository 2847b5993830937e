"""The benchmark configurations at their full size, checked against the
oracle (CHarm's algorithm in double, and in __float128 where stated) on
sampled rings, points or sets, and batched/served results against the same
requests run alone:

* C2 scattered: 1e6 points at nmax 2160 (four points per lane from 2^18 points
  on), and the four-point kernel forced on a small case;
* C4a: the 64-request batched mix (nmax ~ U{90..2700}, seed 64);
* C4b: 256 coefficient sets at nmax 720 on one grid (FP64 matrix cores);
* C5: a resident nmax-2190 model serving 1e6-point scattered requests and grid
  requests on grids up to nmax 5400 (truncated at the model's 2190).

Bound: max|x - x_ref| <= 1e-12 max|x_ref|; batched/served = alone bit for bit
for synthesis, to 1e-14 for analysis (its ring reduction is grouped
differently)."""
import numpy as np
import pytest

import charm_amd
from charm_amd import shc
from charm_amd.device import (SHA_GRID, SHS_GRID, Batch, DevArray, DeviceGrid, Stream, shs_sctr)
from charm_amd.serve import Server
from charm_amd.workloads import GM, R_EARTH, batch_mix, kaula_coeffs, scattered_points
from oracle.oracle import Oracle, tri_index

pytestmark = pytest.mark.gpu

TOL = 1e-12


def relmax(x, ref):
    return float(np.max(np.abs(np.asarray(x) - ref)) / np.max(np.abs(ref)))


@pytest.fixture(scope="module")
def orcd():
    return Oracle("d")


@pytest.fixture(scope="module")
def orcq():
    return Oracle("q")


def sctr(lat, lon, r, C, S, N, st):
    d = [DevArray.from_numpy(a) for a in (lat, lon, r)]
    cs = DevArray.from_numpy(np.stack([C, S], axis=1))
    out = DevArray(lat.size)
    shs_sctr(*d, cs, N, GM, R_EARTH, out, st)
    st.synchronize()
    return out.numpy()


def test_c2_scattered_1e6_points(orcd, orcq):
    N = 2160
    C, S = kaula_coeffs(N, N)
    lat, lon, r = scattered_points(1000000, 2160)
    print("C2: 1e6 scattered points, nmax 2160", flush=True)
    f = sctr(lat, lon, r, C, S, N, Stream())
    idx = np.random.default_rng(8).choice(lat.size, 32, replace=False)
    fd = orcd.shs_sctr(N, C, S, GM, R_EARTH, lat[idx], lon[idx], r[idx])
    fq = orcq.shs_sctr(N, C, S, GM, R_EARTH, lat[idx], lon[idx], r[idx])
    e_d, e_q = relmax(f[idx], fd), relmax(f[idx], fq)
    print("C2 1e6 points (32 sampled): vs double %.2e, vs quad %.2e" % (e_d, e_q))
    assert max(e_d, e_q) <= TOL


@pytest.mark.parametrize("k", [1, 2, 4])
def test_scattered_points_per_lane_forced(orcd, monkeypatch, k):
    """Each points-per-lane variant of the scattered kernel, forced through
    CHARM_AMD_SCTR_K on a small case, against the oracle at every point."""
    N = 300
    C, S = kaula_coeffs(N, 77)
    lat, lon, r = scattered_points(3000, 78)
    monkeypatch.setenv("CHARM_AMD_SCTR_K", str(k))
    f = sctr(lat, lon, r, C, S, N, Stream())
    fd = orcd.shs_sctr(N, C, S, GM, R_EARTH, lat, lon, r)
    assert relmax(f, fd) <= TOL


def test_c4a_batched_mix(orcd):
    mix = batch_mix()
    st = Stream()
    reqs, alone, inputs = [], [], []
    print("C4a: 64 requests, nmax %d..%d" % (min(q[1] for q in mix), max(q[1] for q in mix)),
          flush=True)
    for is_shs, N, seed in mix:
        C, S = kaula_coeffs(N, seed)
        g = DeviceGrid(N, R_EARTH)
        cs = DevArray.from_numpy(np.stack([C, S], axis=1))
        f = DevArray((g.nlat, g.nlon))
        g.shs(cs, N, GM, R_EARTH, f, st)
        if is_shs:
            st.synchronize()
            alone.append(f.numpy())
            reqs.append((SHS_GRID, N, GM, R_EARTH, g, cs, DevArray((g.nlat, g.nlon))))
        else:
            out = DevArray(cs.shape)
            g.sha(f, N, GM, R_EARTH, out, st)
            st.synchronize()
            alone.append(out.numpy())
            reqs.append((SHA_GRID, N, GM, R_EARTH, g, DevArray(cs.shape), f))
        inputs.append((C, S))
    b = Batch(reqs, use_graph=True)
    for rep in range(2):
        b.run(st)
        st.synchronize()
        for (kind, N, mu, R, g, cs, f), ref in zip(reqs, alone):
            if kind == SHS_GRID:
                assert np.array_equal(f.numpy(), ref), N
            else:
                assert np.abs(cs.numpy() - ref).max() <= 1e-14 * np.abs(ref).max(), N
    # sampled rings of the largest and smallest synthesis and analysis against the oracle
    shs_i = [i for i, q in enumerate(mix) if q[0]]
    sha_i = [i for i, q in enumerate(mix) if not q[0]]
    for i in (max(shs_i, key=lambda j: mix[j][1]), min(shs_i, key=lambda j: mix[j][1])):
        kind, N, mu, R, g, cs, f = reqs[i]
        C, S = inputs[i]
        rings = np.array([0, N // 3, N // 2, N], dtype=np.int64)
        p = g.pnt
        ref = orcd.shs_grid(N, C, S, GM, R_EARTH, p.lat, p.r, p.lon, gl=True, rings=rings)
        e = relmax(f.numpy()[rings], ref)
        print("C4a synthesis nmax %d, 4 rings vs oracle: %.2e" % (N, e))
        assert e <= TOL
    i = min(sha_i, key=lambda j: mix[j][1])
    kind, N, mu, R, g, cs, f = reqs[i]
    fn = f.numpy()
    p = g.pnt
    Co, So = orcd.sha_gl(N, fn, p.lat, p.w, R_EARTH, GM, R_EARTH)
    got = cs.numpy()
    scale = max(np.abs(Co).max(), np.abs(So).max())
    e = max(np.abs(got[:, 0] - Co).max(), np.abs(got[:, 1] - So).max()) / scale
    print("C4a analysis nmax %d vs oracle: %.2e" % (N, e))
    assert e <= TOL


def test_c4b_256_sets_nmax720(orcd):
    N, nset = 720, 256
    grid = DeviceGrid(N, R_EARTH)
    st = Stream()
    rng = np.random.default_rng(720)
    seeds = rng.integers(0, 2**31, nset)
    cs = np.stack([np.stack(kaula_coeffs(N, int(s)), axis=1) for s in seeds])
    out = DevArray((nset, grid.nlat, grid.nlon))
    print("C4b: 256 sets at nmax 720", flush=True)
    grid.shs_multi(DevArray.from_numpy(cs), N, nset, GM, R_EARTH, out, st)
    st.synchronize()
    got = out.numpy()
    p = grid.pnt
    rings = np.array([0, 101, 360, 720], dtype=np.int64)
    for i in (0, 77, 255):
        ref = orcd.shs_grid(N, cs[i, :, 0], cs[i, :, 1], GM, R_EARTH, p.lat, p.r, p.lon, gl=True,
                            rings=rings)
        e = relmax(got[i][rings], ref)
        print("C4b set %d, 4 rings vs oracle: %.2e" % (i, e))
        assert e <= TOL


def test_c5_served_nmax2190_model(orcd):
    NM = 2190
    C, S = kaula_coeffs(NM, NM)
    model = shc.Shc.from_arrays(NM, C, S, GM, R_EARTH)
    srv = Server(200 * 2**30, 4)
    mid = srv.add_model(model)
    rng = np.random.default_rng(2190)
    reqs = [("sctr", NM, scattered_points(1000000, 51)),
            ("sctr", 1500, scattered_points(200000, 52)),
            ("shs", NM, 5400),
            ("shs", 800, 900),
            ("sctr", NM, scattered_points(1000, 53))]
    print("C5: nmax-2190 model, 1e6-point and nmax-5400 grid requests", flush=True)

    def submit(kind, n, arg):
        if kind == "sctr":
            return srv.shs_sctr(mid, n, *arg)
        return srv.shs_grid(mid, n, arg, R_EARTH)

    alone = [submit(*q).wait().copy() for q in reqs]
    tickets = [submit(*q) for q in reqs]
    for t, ref, q in zip(tickets, alone, reqs):
        assert np.array_equal(t.wait(), ref), q[:2]
    # samples against the oracle (coefficients truncated to the request's degree)
    for (kind, n, arg), f in zip(reqs, alone):
        Cn = np.concatenate([C[tri_index(NM, m, m):tri_index(NM, n, m) + 1] for m in range(n + 1)])
        Sn = np.concatenate([S[tri_index(NM, m, m):tri_index(NM, n, m) + 1] for m in range(n + 1)])
        if kind == "sctr":
            lat, lon, r = arg
            idx = rng.choice(lat.size, 12, replace=False)
            ref = orcd.shs_sctr(n, Cn, Sn, GM, R_EARTH, lat[idx], lon[idx], r[idx])
            e = relmax(f[idx], ref)
        else:
            from charm_amd import crd
            g = crd.PointGridGL(arg, R_EARTH)
            rings = np.array([1, arg // 2, arg - 3], dtype=np.int64)
            jl = np.linspace(0, len(g.lon) - 1, 16).round().astype(np.int64)
            ref = orcd.shs_grid(n, Cn, Sn, GM, R_EARTH, g.lat, g.r, g.lon[jl], gl=True,
                                rings=rings)
            e = relmax(f[np.ix_(rings, jl)], ref)
        print("C5 %s nmax %d: sampled vs oracle %.2e" % (kind, n, e))
        assert e <= TOL
    srv.close()
