"""Many coefficient sets on one device-resident GL grid (FP64 matrix-core
contraction): every set's grid equals the single-set synthesis of that set
(CHarm evaluates each set by its own charm_shs_point call,
src/shs/shs_point.c:26-129) within 1e-12 of max|f|."""
import numpy as np
import pytest

from charm_amd.device import DevArray, DeviceGrid, Stream
from charm_amd.workloads import GM, R_EARTH, kaula_coeffs

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("nmax,nset", [(0, 3), (7, 17), (160, 40), (720, 16)])
def test_shs_multi_matches_single(nmax, nset):
    grid = DeviceGrid(nmax, R_EARTH)
    st = Stream()
    sets = [kaula_coeffs(nmax, 500 + i) for i in range(nset)]
    cs = np.stack([np.stack([C, S], axis=1) for C, S in sets])
    dcs = DevArray.from_numpy(cs)
    out = DevArray((nset, grid.nlat, grid.nlon))
    grid.shs_multi(dcs, nmax, nset, GM, R_EARTH, out, st)
    st.synchronize()
    got = out.numpy()
    for i in range(nset):
        one = DevArray.from_numpy(cs[i])
        f = DevArray((grid.nlat, grid.nlon))
        grid.shs(one, nmax, GM, R_EARTH, f, st)
        st.synchronize()
        ref = f.numpy()
        assert np.abs(got[i] - ref).max() <= 1e-12 * np.abs(ref).max(), i


def test_shs_multi_rejects_other_radius():
    from charm_amd import CHarmError

    grid = DeviceGrid(10, R_EARTH)
    cs = DevArray.from_numpy(np.zeros((2, 66, 2)))
    out = DevArray((2, grid.nlat, grid.nlon))
    with pytest.raises(CHarmError):
        grid.shs_multi(cs, 10, 2, GM, R_EARTH + 1000.0, out)
