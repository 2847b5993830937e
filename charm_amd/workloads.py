"""Seeded synthetic inputs of the benchmark workloads.

* Kaula-like coefficients: C_00 = 1, degree 1 zero, S_n0 = 0 and, for n >= 2,
  C_nm, S_nm ~ N(0, (1e-5 / n^2)^2).
* Scattered points: lat = asin(U(-1, 1)), lon = U(0, 2 pi), r = R + U(0, 9000 m).
* Earth constants GM = 3.986004415e14 m^3 s^-2, R = 6378136.3 m.
"""
import numpy as np

GM = 3.986004415e14
R_EARTH = 6378136.3


def kaula_coeffs(nmax, seed):
    """(C, S) m-major triangles of degree nmax (CHarm's layout)."""
    rng = np.random.default_rng(seed)
    ncs = (nmax + 1) * (nmax + 2) // 2
    m_of = np.repeat(np.arange(nmax + 1), np.arange(nmax + 1, 0, -1))
    start = np.repeat(np.arange(nmax + 1) * (nmax + 1) - (np.arange(nmax + 1) * (np.arange(nmax + 1) - 1)) // 2,
                      np.arange(nmax + 1, 0, -1))
    n_of = m_of + (np.arange(ncs) - start)
    sig = np.zeros(ncs)
    big = n_of >= 2
    sig[big] = 1e-5 / n_of[big].astype(np.float64) ** 2
    C = rng.standard_normal(ncs) * sig
    S = rng.standard_normal(ncs) * sig
    S[m_of == 0] = 0.0
    C[0] = 1.0
    return C, S


def scattered_points(npoint, seed, R=R_EARTH):
    rng = np.random.default_rng(seed)
    lat = np.arcsin(rng.uniform(-1.0, 1.0, npoint))
    lon = rng.uniform(0.0, 2.0 * np.pi, npoint)
    r = R + rng.uniform(0.0, 9000.0, npoint)
    return lat, lon, r


def degree_order(nmax):
    """Arrays n, m for every position of the m-major triangle."""
    ncs = (nmax + 1) * (nmax + 2) // 2
    m_of = np.repeat(np.arange(nmax + 1), np.arange(nmax + 1, 0, -1))
    start = np.repeat(np.arange(nmax + 1) * (nmax + 1) - (np.arange(nmax + 1) * (np.arange(nmax + 1) - 1)) // 2,
                      np.arange(nmax + 1, 0, -1))
    n_of = m_of + (np.arange(ncs) - start)
    return n_of, m_of


def uniform_coeffs(nmax, seed):
    """CHarm's accuracy protocol (bench/bench.c:192-225): C_nm, S_nm from
    U[-1, 1), S_n0 = 0, m-major triangles of degree nmax."""
    rng = np.random.default_rng(seed)
    ncs = (nmax + 1) * (nmax + 2) // 2
    C = rng.uniform(-1.0, 1.0, ncs)
    S = rng.uniform(-1.0, 1.0, ncs)
    S[: nmax + 1] = 0.0
    return C, S


def batch_mix(count=64, seed=64):
    """The mixed-degree request stream of the batched workload: request i has
    degree N_i ~ U{90..2700}, synthesis for even i and analysis for odd i, and
    Kaula coefficients of seed 1000 + i.  Returns [(is_synthesis, N_i, seed_i)]."""
    degs = np.random.default_rng(seed).integers(90, 2701, count)
    return [(i % 2 == 0, int(N), 1000 + i) for i, N in enumerate(degs)]
