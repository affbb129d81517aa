"""The demagexch oracle (oracle/demagexch.py), pinned without itself:

* the exchange kernel's values by hand (kernel_exchange.go:102-124), with the
  reference's overwrites on a length-1 axis and a periodic length-2 axis;
* its transform against the closed form sum_d (2 cos(2 pi k_d / p_d) - 2) / h_d^2
  (a transposed cell size, a dropped axis or a sign flip fails);
* the convolution with it against the exchange6 oracle (orc_exchange6,
  exchange6.cu:12-163): zero padding versus Neumann edges differ by exactly
  m_i (centre + sum of the present neighbours' 1/h^2), parallel to m;
* dex = Ms K_d + 2 Aex / (mu0 Ms) K_ex by linearity of the transform.
"""
import numpy as np
import pytest

from oracle import demagexch as D
from oracle import oracle as O

CELL = (2e-9, 3e-9, 5e-9)


def _h2():
    return [c * c for c in CELL]


def test_exchange_kernel_by_hand():
    h = _h2()
    K = D.exch6_kernel((4, 6, 8), CELL)
    centre = np.float32(-2 / h[0] - 2 / h[1] - 2 / h[2])
    want = {(0, 0, 0): centre, (1, 0, 0): np.float32(1 / h[0]), (3, 0, 0): np.float32(1 / h[0]),
            (0, 1, 0): np.float32(1 / h[1]), (0, 5, 0): np.float32(1 / h[1]),
            (0, 0, 1): np.float32(1 / h[2]), (0, 0, 7): np.float32(1 / h[2])}
    for s in range(3):
        nz = {tuple(int(x) for x in i): K[s][tuple(i)] for i in np.argwhere(K[s] != 0)}
        assert nz == want
    assert not np.any(K[3:])
    # length-1 axis: both x neighbours wrap onto the centre and replace it
    K1 = D.exch6_kernel((1, 6, 8), CELL)
    assert K1[0][0, 0, 0] == np.float32(1 / h[0])
    assert np.count_nonzero(K1[0]) == 5
    # periodic length-2 axis: +1 and -1 are the same position, assigned, not summed
    K2 = D.exch6_kernel((4, 2, 8), CELL)
    assert K2[1][0, 1, 0] == np.float32(1 / h[1])
    assert K2[1][0, 0, 0] == centre
    assert np.count_nonzero(K2[1]) == 6


@pytest.mark.parametrize("p", [(4, 6, 8), (1, 8, 10)])
def test_exchange_kernel_transform_closed_form(p):
    h = _h2()
    kh, bad = O.kernel_fft(p, K9=D.exch6_kernel(p, CELL))
    assert bad < 1e-12
    k0, k1, k2 = np.meshgrid(np.arange(p[0]), np.arange(p[1]), np.arange(p[2] // 2 + 1),
                             indexing="ij")
    # neighbours: 2 cos(2 pi k_d / p_d) / h_d^2 on each axis of length > 1
    lam = 2 * np.cos(2 * np.pi * k1 / p[1]) / h[1] + 2 * np.cos(2 * np.pi * k2 / p[2]) / h[2]
    if p[0] > 1:
        lam = lam + 2 * np.cos(2 * np.pi * k0 / p[0]) / h[0] - 2 / h[0] - 2 / h[1] - 2 / h[2]
    else:
        # the centre is 1/h_x^2, not -2/h_x^2 - 2/h_y^2 - 2/h_z^2
        lam = lam + 1 / h[0]
    lam = lam / np.prod(p)
    for s in range(3):
        assert np.abs(kh[s] - lam).max() <= 1e-6 * np.abs(lam).max()
    assert not np.any(kh[3:])


@pytest.mark.parametrize("n,pbc", [((3, 5, 6), (0, 0, 0)), ((1, 6, 7), (0, 0, 0)),
                                   ((3, 4, 6), (0, 1, 0))])
def test_exchange_convolution_against_exchange6(n, pbc):
    p = O.pad_size(n, pbc)
    kh, _ = O.kernel_fft(p, K9=D.exch6_kernel(p, CELL))
    m = O.normalize(np.random.default_rng(4).normal(size=(3,) + n))
    conv = O.demag_conv(n, p, kh, m)
    ex = O.exchange6(n, pbc, CELL, 1.0, m)
    # centre of the kernel and the neighbours the zero-padded grid has
    hx = [c * c for c in CELL]
    centre = 1 / hx[0] if p[0] == 1 else -2 / hx[0] - 2 / hx[1] - 2 / hx[2]
    present = np.zeros(n)
    idx = np.indices(n)
    for d in range(3):
        if n[d] == 1:
            continue
        for sgn in (1, -1):
            inside = (idx[d] + sgn >= 0) & (idx[d] + sgn < n[d])
            present += np.where(inside | (pbc[d] == 1), 1 / hx[d], 0.0)
    want = ex + m * (centre + present)
    assert np.abs(conv - want).max() <= 1e-6 * np.abs(ex).max()
    # the difference is parallel to m: the torque is that of exchange6
    assert np.abs(np.cross(conv - ex, m, axis=0)).max() <= 1e-6 * np.abs(ex).max()


def test_dex_kernel_is_the_linear_combination():
    n, pbc, ms, aex = (2, 6, 8), (0, 0, 0), 8e5, 1.3e-11
    p = O.pad_size(n, pbc)
    kd, _ = O.kernel_fft(p, K9=O.kernel_arne(p, CELL, pbc, 8))
    ke, _ = O.kernel_fft(p, K9=D.exch6_kernel(p, CELL))
    want = ms * kd + (2 * aex / (D.MU0 * ms)) * ke
    got = D.dex_khat(p, CELL, pbc, ms, aex)
    assert np.abs(got - want).max() <= 2e-7 * np.abs(want).max()
    # the field: Ms H_demag-per-Ms plus the exchange kernel's field
    m = O.normalize(np.random.default_rng(1).normal(size=(3,) + n))
    hd = D.h_dex(n, pbc, CELL, ms, aex, m)
    want_h = ms * O.demag_conv(n, p, kd, m) + (2 * aex / (D.MU0 * ms)) * O.demag_conv(n, p, ke, m)
    assert np.abs(hd - want_h).max() <= 1e-6 * np.abs(want_h).max()
    # folding: Ms conv(m, folded) = H_demag(accuracy 6) + H_dex
    fold = D.folded_khat(n, pbc, CELL, ms, aex)
    k6, _ = O.kernel_fft(p, K9=O.kernel_arne(p, CELL, pbc, 6))
    both = ms * O.demag_conv(n, p, k6, m) + hd
    assert np.abs(ms * O.demag_conv(n, p, fold, m) - both).max() <= 1e-9 * np.abs(both).max()


def test_accuracy_zero_is_one_point_per_axis():
    """demag_acc is a Uint (verify.go:42-47): 0 is allowed.  Kernel_Arne then
    takes maxSize = d/0 = inf and one integration point per axis
    (kernel_magnetostatic.go:107-119), which for cubic cells is also what
    accuracy 1 gives (no displacement is shorter than the cell); for
    non-cubic cells the self cell's point count differs."""
    p = (4, 8, 8)
    k0 = O.kernel_arne(p, (3e-9, 3e-9, 3e-9), (0, 0, 0), 0)
    k1 = O.kernel_arne(p, (3e-9, 3e-9, 3e-9), (0, 0, 0), 1)
    assert np.array_equal(k0, k1)
    kn0 = O.kernel_arne(p, (2e-9, 3e-9, 5e-9), (0, 0, 0), 0)
    kn1 = O.kernel_arne(p, (2e-9, 3e-9, 5e-9), (0, 0, 0), 1)
    assert not np.array_equal(kn0, kn1)
    assert np.isfinite(kn0).all()
