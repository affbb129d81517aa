"""Seeded synthetic inputs for the benchmark configurations.

Random numbers come from SplitMix64 (Steele, Lea, Flood 2014): element i of
stream `seed` is mix(seed + (i+1)*0x9E3779B97F4A7C15), turned into a double
in [0,1) from its top 53 bits; Gaussians by Box-Muller.  Everything is numpy,
so the same float32 arrays feed the GPU and the FP64 oracle.

Sizes here are in *user* order (x, y, z); `to_internal` swaps them into the
reference's internal (z, y, x) order (src/mumax/gpu/doc.go:22-35).
"""
import numpy as np

MU0 = 4.0 * 3.14159265358979323846264338327950288 * 1e-7
GAMMA0 = 2.211e5

_GOLD = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)


def splitmix64(seed, n, offset=0):
    """n uint64 outputs of SplitMix64 stream `seed`, starting at element offset."""
    with np.errstate(over="ignore"):
        i = np.arange(offset + 1, offset + n + 1, dtype=np.uint64)
        z = np.uint64(seed) + i * _GOLD
        z = (z ^ (z >> np.uint64(30))) * _M1
        z = (z ^ (z >> np.uint64(27))) * _M2
        z = z ^ (z >> np.uint64(31))
    return z


def uniform(seed, n, offset=0):
    return (splitmix64(seed, n, offset) >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)


def gaussian(seed, n, offset=0):
    m = (n + 1) // 2
    u = uniform(seed, 2 * m, offset)
    u1 = 1.0 - u[0::2]           # (0, 1]
    u2 = u[1::2]
    r = np.sqrt(-2.0 * np.log(u1))
    g = np.concatenate([r * np.cos(2 * np.pi * u2), r * np.sin(2 * np.pi * u2)])
    return g[:n]


def random_unit_m(seed, n_internal):
    """Normalised Gaussian 3-vectors, float32 (3, n0, n1, n2), internal order."""
    N = int(np.prod(n_internal))
    g = gaussian(seed, 3 * N).reshape(3, N)
    g /= np.sqrt((g * g).sum(axis=0))
    return g.reshape((3,) + tuple(n_internal)).astype(np.float32)


def smooth_texture_m(seed, n_internal, cell_internal, nmodes=16):
    """Sum of `nmodes` random-phase sinusoids per component, normalised."""
    n0, n1, n2 = n_internal
    L = [n_internal[i] * cell_internal[i] for i in range(3)]
    u = uniform(seed, 3 * nmodes * 5)
    z = (np.arange(n0) + 0.5) * cell_internal[0]
    y = (np.arange(n1) + 0.5) * cell_internal[1]
    x = (np.arange(n2) + 0.5) * cell_internal[2]
    m = np.zeros((3, n0, n1, n2))
    t = 0
    for c in range(3):
        for q in range(nmodes):
            kz = 2 * np.pi * np.floor(1 + 2 * u[t]) / L[0] if n0 > 1 else 0.0
            ky = 2 * np.pi * np.floor(1 + 4 * u[t + 1]) / L[1]
            kx = 2 * np.pi * np.floor(1 + 4 * u[t + 2]) / L[2]
            ph = 2 * np.pi * u[t + 3]
            amp = 0.5 + u[t + 4]
            t += 5
            m[c] += amp * np.sin(kz * z[:, None, None] + ky * y[None, :, None]
                                 + kx * x[None, None, :] + ph)
    m /= np.sqrt((m * m).sum(axis=0))
    return m.astype(np.float32)


def vortex_m(seed, n_internal, cell_internal):
    """A vortex per layer stack: seeded centre, chirality, polarity and core
    radius; in-plane curl around the centre, out-of-plane core."""
    n0, n1, n2 = n_internal
    u = uniform(seed, 6, offset=7 << 32)
    y = (np.arange(n1) + 0.5) * cell_internal[1]
    x = (np.arange(n2) + 0.5) * cell_internal[2]
    yc = (0.2 + 0.6 * u[0]) * n1 * cell_internal[1]
    xc = (0.2 + 0.6 * u[1]) * n2 * cell_internal[2]
    chi = 1.0 if u[2] < 0.5 else -1.0
    pol = 1.0 if u[3] < 0.5 else -1.0
    rc = (5.0 + 15.0 * u[4]) * 1e-9
    dy = y[:, None] - yc
    dx = x[None, :] - xc
    r = np.sqrt(dx * dx + dy * dy) + 1e-30
    m = np.zeros((3, n0, n1, n2))
    m[2] = (-chi * dy / r)[None]                 # user x
    m[1] = (chi * dx / r)[None]                  # user y
    m[0] = (pol * np.exp(-(r * r) / (rc * rc)))[None]
    m /= np.sqrt((m * m).sum(axis=0))
    return m.astype(np.float32)


def domain_wall_m(seed, n_internal, cell_internal):
    """A straight Neel or Bloch wall: seeded normal direction, position, width
    and domain axis (in-plane)."""
    n0, n1, n2 = n_internal
    u = uniform(seed, 6, offset=9 << 32)
    y = (np.arange(n1) + 0.5) * cell_internal[1]
    x = (np.arange(n2) + 0.5) * cell_internal[2]
    th = 2 * np.pi * u[0]
    ny_, nx_ = np.sin(th), np.cos(th)
    pos = (0.3 + 0.4 * u[1]) * (abs(nx_) * n2 * cell_internal[2] + abs(ny_) * n1 * cell_internal[1])
    width = (5.0 + 20.0 * u[2]) * 1e-9
    s = nx_ * x[None, :] + ny_ * y[:, None]
    s = s - s.min() - pos
    ph = 2 * np.pi * u[3]
    ax, ay = np.cos(ph), np.sin(ph)              # domain axis (user x, y)
    m = np.zeros((3, n0, n1, n2))
    t, sh = np.tanh(s / width), 1.0 / np.cosh(s / width)
    m[2] = (ax * t)[None]
    m[1] = (ay * t)[None]
    if u[4] < 0.5:                               # Neel: rotates in-plane
        m[2] += (-ay * sh)[None]
        m[1] += (ax * sh)[None]
    else:                                        # Bloch: rotates out of plane
        m[0] = sh[None] * np.ones((n0, 1, 1))
    m /= np.sqrt((m * m).sum(axis=0))
    return m.astype(np.float32)


def training_texture(k, n_internal, cell_internal):
    """Texture number k of the surrogate's training set: vortices, domain
    walls, smooth and rough random states in turn."""
    kind = k % 4
    if kind == 0:
        return vortex_m(1000 + k, n_internal, cell_internal)
    if kind == 1:
        return domain_wall_m(2000 + k, n_internal, cell_internal)
    if kind == 2:
        return smooth_texture_m(3000 + k, n_internal, cell_internal, nmodes=8)
    return random_unit_m(4000 + k, n_internal)


def to_internal(v):
    """User (x, y, z) triple -> internal (z, y, x)."""
    return (v[2], v[1], v[0])


def config(k):
    """The benchmark configurations as concrete inputs (internal order)."""
    if k == 1:
        n = to_internal((128, 32, 1))
        cell = to_internal((3.90625e-9, 3.90625e-9, 3e-9))
        ph = 2 * np.pi * uniform(1, 1, offset=1 << 40)[0]
        b_user = (0.010 * np.cos(ph), 0.010 * np.sin(ph), 0.0)
        return dict(name="film 128x32x1", n=n, cell=cell, msat=8.0e5, aex=1.3e-11, alpha=0.02,
                    gamma=GAMMA0, bext=to_internal(b_user), ku=None, maxerr=1e-5,
                    m=random_unit_m(1, n), run=1e-9)
    if k == 2:
        n = to_internal((256, 256, 64))
        cell = (4e-9, 4e-9, 4e-9)
        return dict(name="bulk 256x256x64", n=n, cell=cell, msat=8.0e5, aex=1.3e-11, alpha=0.1,
                    gamma=GAMMA0, bext=(0.0, 0.0, 0.0), ku=5e5, anis_u=to_internal((0, 0, 1)),
                    maxerr=1e-5, m=random_unit_m(2, n))
    if k == 3:
        n = to_internal((1024, 1024, 32))
        cell = (2e-9, 2e-9, 2e-9)
        return dict(name="film 1024x1024x32", n=n, cell=cell, msat=1.0e6, aex=1.5e-11,
                    alpha=0.05, gamma=GAMMA0, bext=(0.0, 0.0, 0.0), ku=None, maxerr=1e-5,
                    m=smooth_texture_m(3, n, cell))
    if k == 5:
        n = to_internal((256, 256, 4))
        cell = to_internal((4e-9, 4e-9, 3e-9))
        return dict(name="film 256x256x4", n=n, cell=cell, msat=8.0e5, aex=1.3e-11, alpha=0.02,
                    gamma=GAMMA0, bext=(0.0, 0.0, 0.0), ku=None, maxerr=1e-5,
                    m=random_unit_m(5, n))
    raise ValueError("unknown config %r" % k)


def batch_configs(count=64, seed=4):
    """Configuration 4: uneven film sizes 32x32x1 .. 256x128x4, cell 4x4x3 nm."""
    out = []
    for i in range(count):
        u = uniform(seed + i, 5)
        nx = 32 * (1 + int(u[0] * 8))
        ny = 32 * (1 + int(u[1] * 4))
        nz = 1 + int(u[2] * 4)
        n = to_internal((nx, ny, nz))
        cell = to_internal((4e-9, 4e-9, 3e-9))
        out.append(dict(name="batch %d" % i, n=n, cell=cell, msat=6e5 + 6e5 * u[3],
                        aex=1.3e-11, alpha=0.005 + 0.095 * u[4], gamma=GAMMA0,
                        bext=(0.0, 0.0, 0.0), ku=None, maxerr=1e-5, j=to_internal((1e12, 0, 0)),
                        pol=0.56, xi=0.05, m=random_unit_m(seed + i + 1000, n)))
    return out


def setup_sim(sim, cfg):
    """Applies a configuration dict to an mmx.Sim."""
    sim.setv("Msat", cfg["msat"])
    sim.setv("Aex", cfg["aex"])
    sim.setv("alpha", cfg["alpha"])
    sim.setv("gamma", cfg["gamma"])
    sim.setv("B_ext", cfg["bext"])
    sim.setv("m_maxerror", cfg["maxerr"])
    if cfg.get("ku") is not None:
        sim.load("anisotropy/uniaxial")
        sim.setv("Ku", cfg["ku"])
        sim.setv("anisU", cfg["anis_u"])
    if cfg.get("j") is not None:
        sim.load("zhang-li")
        sim.setv("j", cfg["j"])
        sim.setv("polarisation", cfg["pol"])
        sim.setv("xi", cfg["xi"])
    sim.setarray("m", cfg["m"])
