"""Batched execution: one persistent launch per transform kind over requests
of uneven degree and shape (grids and scattered point sets), and HIP-graph
replays of it.  Each request's result must equal the same request run alone
(synthesis on grids and at scattered points: bit for bit, since every ring or
point is summed in the same order; analysis: to 1e-14 of max|C|, since the
cross-ring reduction is grouped differently)."""
import numpy as np
import pytest

import charm_amd
from charm_amd.device import (SHA_GRID, SHS_GRID, SHS_SCTR, Batch, DevArray, DeviceGrid, Stream,
                              shs_sctr)
from charm_amd.workloads import GM, R_EARTH, kaula_coeffs, scattered_points

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def need_gpu():
    if charm_amd.lib().charm_amd_device_count() < 1:
        pytest.fail("no HIP device visible: the GPU tests need an MI355X")


def make_requests(degrees, seed0):
    reqs, expect = [], []
    st = Stream()
    for i, (kind, N, Ngrid) in enumerate(degrees):
        C, S = kaula_coeffs(N, seed0 + i)
        cs = DevArray.from_numpy(np.stack([C, S], axis=1))
        if kind == SHS_SCTR:
            # Ngrid is the number of points; every other set lies on the sphere r = R
            lat, lon, r = scattered_points(Ngrid, seed0 + i)
            if i % 2:
                r = np.full_like(lat, R_EARTH)
            pts = tuple(DevArray.from_numpy(a) for a in (lat, lon, r))
            f = DevArray(Ngrid)
            shs_sctr(*pts, cs, N, GM, R_EARTH, f, st)      # alone
            st.synchronize()
            expect.append(f.numpy())
            reqs.append((SHS_SCTR, N, GM, R_EARTH, pts, cs, DevArray(Ngrid)))
            continue
        g = DeviceGrid(Ngrid, R_EARTH + (0.0 if i % 3 else 500.0))
        f = DevArray((g.nlat, g.nlon))
        if kind == SHS_GRID:
            g.shs(cs, N, GM, R_EARTH, f, st)       # alone
            st.synchronize()
            expect.append(f.numpy())
            out_f = DevArray((g.nlat, g.nlon))
            reqs.append((SHS_GRID, N, GM, R_EARTH, g, cs, out_f))
        else:
            g.shs(cs, N, GM, R_EARTH, f, st)       # input field
            out = DevArray(cs.shape)
            g.sha(f, N, GM, R_EARTH, out, st)      # alone
            st.synchronize()
            expect.append(out.numpy())
            reqs.append((SHA_GRID, N, GM, R_EARTH, g, DevArray(cs.shape), f))
    return reqs, expect


def check(reqs, expect):
    for (kind, N, mu, R, g, cs, f), ref in zip(reqs, expect):
        got = cs.numpy() if kind == SHA_GRID else f.numpy()
        if kind != SHA_GRID:
            assert np.array_equal(got, ref), N
        else:
            assert np.abs(got - ref).max() <= 1e-14 * np.abs(ref).max(), N


DEGREES = [(SHS_GRID, 90, 90), (SHA_GRID, 133, 133), (SHS_GRID, 300, 310), (SHA_GRID, 257, 257),
           (SHS_GRID, 7, 12), (SHA_GRID, 64, 80), (SHS_GRID, 511, 511), (SHA_GRID, 400, 400)]


def test_batch_matches_alone():
    reqs, expect = make_requests(DEGREES, 100)
    b = Batch(reqs, use_graph=False)
    st = Stream()
    b.run(st)
    st.synchronize()
    check(reqs, expect)


def test_batch_graph_replays():
    """The first run is eager and records a graph; replays (with the inputs
    changed in between by other work) must each reproduce the alone results."""
    reqs, expect = make_requests(DEGREES, 200)
    b = Batch(reqs, use_graph=True)
    st = Stream()
    b.run(st)
    st.synchronize()
    check(reqs, expect)
    for rep in range(3):
        # scribble over the outputs with unrelated eager work, then replay
        for kind, N, mu, R, g, cs, f in reqs:
            if kind == SHS_GRID:
                z = DevArray.from_numpy(np.zeros(cs.shape))
                g.shs(z, N, mu, R, f, st)
            elif kind == SHS_SCTR:
                f.copy_from_numpy(np.zeros(f.shape))
        st.synchronize()
        b.run(st)
        st.synchronize()
        check(reqs, expect)
    assert b.uses_graph


MIXED = [(SHS_SCTR, 360, 100000), (SHS_GRID, 200, 200), (SHS_SCTR, 90, 37), (SHA_GRID, 150, 150),
         (SHS_SCTR, 600, 3001), (SHS_SCTR, 1, 700), (SHS_SCTR, 250, 20000)]


def test_batch_scattered_matches_alone():
    """Scattered point sets of very different sizes (37 to 1e5 points, so
    from 1 to 64 order chunks) mixed with grid requests."""
    reqs, expect = make_requests(MIXED, 300)
    b = Batch(reqs, use_graph=False)
    st = Stream()
    b.run(st)
    st.synchronize()
    check(reqs, expect)
    b.run(st)      # a second run over the same batch gives the same values
    st.synchronize()
    check(reqs, expect)


def test_batch_scattered_graph_replays():
    reqs, expect = make_requests(MIXED, 400)
    b = Batch(reqs, use_graph=True)
    st = Stream()
    for rep in range(3):
        for kind, N, mu, R, g, cs, f in reqs:
            if kind == SHS_SCTR:
                f.copy_from_numpy(np.zeros(f.shape))
        st.synchronize()
        b.run(st)
        st.synchronize()
        check(reqs, expect)
    assert b.uses_graph
