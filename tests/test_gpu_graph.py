"""The device-resident entry points under HIP stream capture.

They take per-call scratch from the stream-ordered allocator, which under
capture would become graph memory-allocation nodes.  A replay of such a graph
left the GPU faulted on MI355X, with no mistake found in the library that
explains it, so the entry points refuse a capturing stream: the refusal is
tested here, and the replay test is kept but skipped."""
import ctypes as ct

import numpy as np
import pytest

from charm_amd._err import CHARM_EFUNCARG
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


def _device_calls(grid, cs, pts, outs, N, st):
    f1, f2, c1, fs, gr = outs
    return [
        lambda: grid.shs(cs, N, GM, R_EARTH + 2000.0, f1, st),
        lambda: grid.shs(cs, N, GM, R_EARTH, f2, st),
        lambda: grid.sha(f2, N, GM, R_EARTH, c1, st),
        lambda: shs_sctr(*pts, cs, N, GM, R_EARTH, fs, st),
        lambda: grid.shs_deriv(cs, N, GM, R_EARTH, DERIV_GRAD1, gr, st),
    ]


def test_device_calls_refuse_stream_capture():
    """Each device-resident call on a capturing stream returns CHarm's
    CHARM_EFUNCARG before it enqueues anything: the captured graph is empty."""
    from charm_amd._err import CHarmError

    N = 40
    C, S = kaula_coeffs(N, 61)
    grid = DeviceGrid(N, R_EARTH + 2000.0)
    st = Stream()
    cs = DevArray.from_numpy(np.stack([C, S], axis=1))
    lat, lon, r = scattered_points(500, 62)
    pts = [DevArray.from_numpy(a) for a in (lat, lon, r)]
    outs = (DevArray((grid.nlat, grid.nlon)), DevArray((grid.nlat, grid.nlon)), DevArray(cs.shape),
            DevArray(lat.size), DevArray((3, grid.nlat, grid.nlon)))
    h = _graph_api()
    h.hipGraphGetNodes.argtypes = [_vp, _vp, ct.POINTER(ct.c_size_t)]
    assert h.hipStreamBeginCapture(st.handle, 2) == 0
    errors = []
    try:
        for call in _device_calls(grid, cs, pts, outs, N, st):
            try:
                call()
            except CHarmError as e:
                errors.append((e.code, str(e)))
    finally:
        graph = _vp()
        rc = h.hipStreamEndCapture(st.handle, ct.byref(graph))
    assert rc == 0
    try:
        n = ct.c_size_t(0)
        assert h.hipGraphGetNodes(graph, None, ct.byref(n)) == 0
        assert n.value == 0
    finally:
        h.hipGraphDestroy(graph)
    assert len(errors) == 5
    assert all(code == CHARM_EFUNCARG and "stream capture" in msg for code, msg in errors)


@pytest.mark.skip(reason="a replay of captured device calls (graph memory-allocation nodes) "
                         "left the GPU faulted on MI355X; the entry points now refuse capture")
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
