"""Writes workloads_first_outputs.json: the first outputs of the seeded input
generators (charm_amd/workloads.py), so that a change of generator or of
numpy's bit stream shows up as a failing test rather than as silently
different workloads.  Run from the repository root:
    python tests/golden/make_workloads_fixture.py"""
import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from charm_amd.workloads import batch_mix, kaula_coeffs, scattered_points, uniform_coeffs  # noqa: E402


def outputs():
    C, S = kaula_coeffs(2160, 7)
    lat, lon, r = scattered_points(1000000, 11)
    Cu, Su = uniform_coeffs(1000, 1000)
    return {
        "kaula_2160_seed7": {"C_first": C[:8].tolist(), "S_first": S[2161:2169].tolist(),
                             "C_last": C[-3:].tolist(), "C_sum": float(C.sum())},
        "scattered_1e6_seed11": {"lat_first": lat[:6].tolist(), "lon_first": lon[:6].tolist(),
                                 "r_first": r[:6].tolist(), "lat_last": lat[-2:].tolist()},
        "uniform_1000_seed1000": {"C_first": Cu[:6].tolist(), "S_first": Su[1001:1007].tolist()},
        "batch_mix_64_seed64": [list(q) for q in batch_mix()],
    }


if __name__ == "__main__":
    with open(os.path.join(HERE, "workloads_first_outputs.json"), "w") as fh:
        json.dump(outputs(), fh, indent=1)
