"""Oracle of the temperature/brown field (test infrastructure only).

* scaling: TempBrownUpdater.Update (modules/mod_temperature.go:71-134) and
  temperature_scaleKern (libmumax2/temperature.cu:11-52):
  H_c(i) = n_c(i) * sqrt((alpha_i T_i a) / (b Ms_i)) in float32, 0 where the
  Msat map is 0, a = float32(alpha 2 kB T), b = float32(mu0 V gamma dt Msat);
* noise: the reference's cuRAND stream is not reproduced; the build's
  counter-based generator is Philox4x32-10 (Salmon, Moraes, Dror, Shaw,
  "Parallel random numbers: as easy as 1, 2, 3", SC 2011), restated here from
  the paper: counter (cell lo, cell hi, draw lo, draw hi), key (seed lo, seed
  hi); uniforms (x + 1/2) 2^-32; Box-Muller in FP64 on words (0, 1) for
  components 0 and 1 and (2, 3) for component 2.

Pinned in tests/test_oracle_thermal.py: Philox against the known-answer
vectors published with the Random123 library, the normals' moments, and the
physics -- an ensemble of independent macrospins driven by this field reaches
the Boltzmann (Langevin-function) magnetisation.
"""
import numpy as np

KB = 1.380650424e-23          # common/constants.go:21
MU0 = 4.0 * np.pi * 1e-7

_M0, _M1 = np.uint64(0xD2511F53), np.uint64(0xCD9E8D57)
_W0, _W1 = 0x9E3779B9, 0xBB67AE85
_MASK = np.uint64(0xFFFFFFFF)


def philox4x32_10(c0, c1, c2, c3, k0, k1):
    """Vectorised Philox4x32-10: counters as uint32 arrays, key as ints."""
    c = [np.asarray(x, dtype=np.uint64) & _MASK for x in (c0, c1, c2, c3)]
    k0, k1 = int(k0) & 0xFFFFFFFF, int(k1) & 0xFFFFFFFF
    for r in range(10):
        if r > 0:
            k0 = (k0 + _W0) & 0xFFFFFFFF
            k1 = (k1 + _W1) & 0xFFFFFFFF
        p0 = _M0 * c[0]
        p1 = _M1 * c[2]
        hi0, lo0 = p0 >> np.uint64(32), p0 & _MASK
        hi1, lo1 = p1 >> np.uint64(32), p1 & _MASK
        c = [hi1 ^ c[1] ^ np.uint64(k0), lo1, hi0 ^ c[3] ^ np.uint64(k1), lo0]
    return [x.astype(np.uint32) for x in c]


def normals(N, seed, draw):
    """(3, N) float32 standard normals of draw `draw` of `seed`."""
    i = np.arange(N, dtype=np.uint64)
    seed = int(seed) & 0xFFFFFFFFFFFFFFFF
    draw = int(draw)
    w = philox4x32_10(i & _MASK, i >> np.uint64(32), np.full(N, draw & 0xFFFFFFFF),
                      np.full(N, (draw >> 32) & 0xFFFFFFFF), seed & 0xFFFFFFFF, seed >> 32)
    u = [(x.astype(np.float64) + 0.5) * 2.0 ** -32 for x in w]
    ra = np.sqrt(-2.0 * np.log(u[0]))
    rb = np.sqrt(-2.0 * np.log(u[2]))
    ta = 2.0 * np.pi * u[1]
    tb = 2.0 * np.pi * u[3]
    return np.stack([ra * np.cos(ta), ra * np.sin(ta), rb * np.cos(tb)]).astype(np.float32)


def htherm(n, cell, seed, draw, temp, alpha, msat, gamma, dt, temp_map=None, alpha_map=None,
           msat_map=None):
    """H_therm (3, n0, n1, n2) float32 as the build draws it."""
    N = int(np.prod(n))
    nz = normals(N, seed, draw)
    V = cell[0] * cell[1] * cell[2]
    a = np.float32(alpha * KB * 2 * temp)
    b = np.float32(MU0 * V * gamma * dt * msat)
    one = np.ones(N, np.float32)
    am = one if alpha_map is None else np.asarray(alpha_map, np.float32).ravel()
    tm = one if temp_map is None else np.asarray(temp_map, np.float32).ravel()
    mm = one if msat_map is None else np.asarray(msat_map, np.float32).ravel()
    with np.errstate(divide="ignore", invalid="ignore"):
        s = np.sqrt((am * tm * a) / (b * mm)).astype(np.float32)
    h = np.where(mm != 0, nz * s, np.float32(0.0)).astype(np.float32)
    return h.reshape((3,) + tuple(n))


def heun_step(prob, m, dt, ht, gamma):
    """One solver/heun step (solve_heun.go:27-79, the corrector form of
    solve_rk12.go:124-158) with the field ht held fixed over the step:
    predictor m1 = normalize(m + h tau(m)), corrector
    normalize(m1 + h/2 (tau(m1) - tau(m))), h = float32(dt gamma)."""
    from oracle import oracle as O
    h = float(np.float32(dt * gamma))
    m0 = O.normalize(m)
    t0 = prob.torque(m0, prob.heff(m0) + ht)
    mp = O.normalize(m0 + h * t0)
    t1 = prob.torque(mp, prob.heff(mp) + ht)
    return O.normalize(mp + 0.5 * h * (t1 - t0))
