"""The temperature/brown oracle (oracle/thermal.py).

* Philox4x32-10 against the known-answer vectors published with the
  Random123 library (kat_vectors: counters/keys all zero, all ones, and the
  digits of pi);
* the normals: mean 0, variance 1, uncorrelated components and draws;
* the scaling of temperature.cu:11-52 on hand-computed cells (maps, a zero
  Msat cell);
* physics: 2048 independent macrospins (no exchange, no demag) in B = 50 mT
  at 300 K, stepped by solver/heun with this field, reach the Boltzmann
  magnetisation <m_B> = L(x) = coth x - 1/x, x = Ms V B / (kB T).  A noise
  variance off by a factor 2 (the factor 2 of 2 alpha kB T, or dt) moves
  <m_B> to 0.36 or 0.10 against 0.20.
"""
import numpy as np
import pytest

from oracle import oracle as O
from oracle import thermal as T

KAT = [((0, 0, 0, 0), (0, 0), (0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8)),
       ((0xFFFFFFFF,) * 4, (0xFFFFFFFF,) * 2, (0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD)),
       ((0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344), (0xA4093822, 0x299F31D0),
        (0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1))]


@pytest.mark.parametrize("ctr,key,want", KAT)
def test_philox_known_answers(ctr, key, want):
    got = T.philox4x32_10(*[np.array([c], np.uint64) for c in ctr], *key)
    assert tuple(int(g[0]) for g in got) == want


def test_normals_moments():
    n = T.normals(200000, 5, 3)
    assert np.abs(n.mean(axis=1)).max() < 0.01
    assert np.abs(n.var(axis=1) - 1.0).max() < 0.015
    c = np.corrcoef(np.concatenate([n, T.normals(200000, 5, 4)]))
    assert np.abs(c - np.eye(6)).max() < 0.01           # components and draws
    assert not np.array_equal(T.normals(16, 5, 0), T.normals(16, 6, 0))


def test_scaling_by_hand():
    n, cell = (1, 2, 2), (2e-9, 3e-9, 5e-9)
    tm = np.array([1.0, 0.5, 2.0, 1.0], np.float32)
    am = np.array([1.0, 2.0, 1.0, 1.0], np.float32)
    mm = np.array([1.0, 1.0, 0.0, 0.25], np.float32)
    temp, alpha, ms, dt = 300.0, 0.1, 8e5, 1e-13
    h = T.htherm(n, cell, 11, 2, temp, alpha, ms, O.GAMMA0, dt, temp_map=tm, alpha_map=am,
                 msat_map=mm).reshape(3, 4)
    nz = T.normals(4, 11, 2)
    V = 2e-9 * 3e-9 * 5e-9
    for i in range(4):
        if mm[i] == 0:
            assert np.all(h[:, i] == 0)
            continue
        sd = np.sqrt(alpha * am[i] * 2 * T.KB * temp * tm[i] /
                     (T.MU0 * V * O.GAMMA0 * dt * ms * mm[i]))
        assert np.allclose(h[:, i], nz[:, i] * sd, rtol=2e-6)


def test_macrospin_ensemble_reaches_boltzmann():
    n, cell = (1, 32, 64), (4e-9, 4e-9, 4e-9)
    B, temp, alpha, ms, dt = 0.05, 300.0, 0.5, 8e5, 2e-13
    prob = O.Problem(n, cell, msat=ms, aex=0.0, alpha=alpha, demag=False, bext=(B, 0.0, 0.0))
    m = O.normalize(np.random.default_rng(3).normal(size=(3,) + n))    # isotropic start
    mean = np.zeros(3)
    ns = 0
    for k in range(2500 + 4000):                # 2500 steps ~ 5.7 rotational times
        ht = T.htherm(n, cell, 7, k, temp, alpha, ms, O.GAMMA0, dt).astype(np.float64)
        m = T.heun_step(prob, m, dt, ht, O.GAMMA0)
        if k >= 2500 and k % 10 == 0:
            mean += m.reshape(3, -1).mean(axis=1)
            ns += 1
    mean /= ns
    x = ms * cell[0] * cell[1] * cell[2] * B / (T.KB * temp)
    L = 1.0 / np.tanh(x) - 1.0 / x
    # B along user z = internal component 0; Heun's weak-order bias at this
    # step is ~ +0.009 (+0.006 at half the step)
    assert abs(mean[0] - L) < 0.02, (mean, L)
    assert abs(mean[1]) < 0.02 and abs(mean[2]) < 0.02
