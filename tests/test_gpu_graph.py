"""The device-resident entry points under HIP stream capture: a grid
synthesis and analysis (at the grid radius and at another radius, so that
the radius powers and the rescale table are made inside the graph), a
scattered synthesis and a gradient are captured once and replayed; every
replay, with unrelated eager work and synchronisations in between, must
reproduce the eager results bit for bit.  Nothing a captured call records
may point at host memory that is gone by the time of the replay."""
import ctypes as ct

import numpy as np
import pytest

from charm_amd.device import (DERIV_GRAD1, DevArray, DeviceGrid, Stream, hip, shs_sctr)
from charm_amd.workloads import GM, R_EARTH, kaula_coeffs, scattered_points

pytestmark = pytest.mark.gpu

_vp = ct.c_void_p


def _graph_api():
    h = hip()
    h.hipStreamBeginCapture.argtypes = [_vp, ct.c_int]
    h.hipStreamEndCapture.argtypes = [_vp, ct.POINTER(_vp)]
    h.hipGraphInstantiate.argtypes = [ct.POINTER(_vp), _vp, _vp, _vp, ct.c_size_t]
    h.hipGraphLaunch.argtypes = [_vp, _vp]
    h.hipGraphExecDestroy.argtypes = [_vp]
    h.hipGraphDestroy.argtypes = [_vp]
    return h


def test_device_calls_replay_from_a_graph():
    N = 300
    C, S = kaula_coeffs(N, 61)
    grid = DeviceGrid(N, R_EARTH + 2000.0)
    st = Stream()
    cs = DevArray.from_numpy(np.stack([C, S], axis=1))
    lat, lon, r = scattered_points(50000, 62)
    pts = [DevArray.from_numpy(a) for a in (lat, lon, r)]
    f1 = DevArray((grid.nlat, grid.nlon))       # synthesis at the grid radius
    f2 = DevArray((grid.nlat, grid.nlon))       # synthesis at R (radius powers)
    c1 = DevArray(cs.shape)                     # analysis with rescale to R
    fs = DevArray(lat.size)
    gr = DevArray((3, grid.nlat, grid.nlon))

    def work():
        grid.shs(cs, N, GM, R_EARTH + 2000.0, f1, st)
        grid.shs(cs, N, GM, R_EARTH, f2, st)
        grid.sha(f2, N, GM, R_EARTH, c1, st)
        shs_sctr(*pts, cs, N, GM, R_EARTH, fs, st)
        grid.shs_deriv(cs, N, GM, R_EARTH, DERIV_GRAD1, gr, st)

    # eager (twice: the first call at a degree builds tables and plans)
    work()
    work()
    st.synchronize()
    outs = (f1, f2, c1, fs, gr)
    ref = [o.numpy().copy() for o in outs]
    h = _graph_api()
    assert h.hipStreamBeginCapture(st.handle, 2) == 0      # relaxed mode
    work()
    graph = _vp()
    assert h.hipStreamEndCapture(st.handle, ct.byref(graph)) == 0
    ex = _vp()
    assert h.hipGraphInstantiate(ct.byref(ex), graph, None, None, 0) == 0
    try:
        for rep in range(4):
            # unrelated eager work that overwrites the outputs and churns the
            # library's buffers, then a replay
            for o in outs:
                o.copy_from_numpy(np.full(o.shape, np.nan))
            grid.shs(DevArray.from_numpy(np.zeros(cs.shape)), N, GM, R_EARTH, f2, st)
            st.synchronize()
            assert h.hipGraphLaunch(ex, st.handle) == 0
            st.synchronize()
            for o, x in zip(outs, ref):
                assert np.array_equal(o.numpy(), x), rep
    finally:
        h.hipGraphExecDestroy(ex)
        h.hipGraphDestroy(graph)
