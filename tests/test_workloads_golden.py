"""The seeded input generators against their committed first outputs
(tests/golden/workloads_first_outputs.json, written by
tests/golden/make_workloads_fixture.py), plus the properties the workloads
are defined by."""
import json
import os

import numpy as np

from charm_amd.workloads import (R_EARTH, batch_mix, degree_order, kaula_coeffs,
                                 scattered_points, uniform_coeffs)

HERE = os.path.dirname(os.path.abspath(__file__))


def test_generators_match_golden_first_outputs():
    import importlib.util
    spec = importlib.util.spec_from_file_location(
        "mkfix", os.path.join(HERE, "golden", "make_workloads_fixture.py"))
    mk = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mk)
    with open(os.path.join(HERE, "golden", "workloads_first_outputs.json")) as fh:
        want = json.load(fh)
    got = json.loads(json.dumps(mk.outputs()))
    assert got == want


def test_generator_definitions():
    N = 300
    C, S = kaula_coeffs(N, 3)
    n, m = degree_order(N)
    assert C[0] == 1.0 and np.all(C[n == 1] == 0) and np.all(S[n == 1] == 0)
    assert np.all(S[m == 0] == 0)
    # N(0, (1e-5/n^2)^2) for n >= 2: normalised values are standard normal
    z = np.concatenate([C[n >= 2] * n[n >= 2] ** 2 / 1e-5, S[(n >= 2) & (m > 0)] *
                        n[(n >= 2) & (m > 0)] ** 2 / 1e-5])
    assert abs(z.mean()) < 0.02 and abs(z.std() - 1.0) < 0.02
    lat, lon, r = scattered_points(200000, 5)
    assert np.all(np.abs(lat) <= np.pi / 2) and abs(np.sin(lat).mean()) < 0.01
    assert np.all((lon >= 0) & (lon < 2 * np.pi)) and np.all((r >= R_EARTH) & (r < R_EARTH + 9000))
    Cu, Su = uniform_coeffs(200, 9)
    assert np.all((Cu >= -1) & (Cu < 1)) and np.all(Su[:201] == 0)
    mix = batch_mix()
    assert len(mix) == 64 and all(90 <= q[1] <= 2700 for q in mix)
    assert sum(q[0] for q in mix) == 32
