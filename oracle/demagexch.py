"""Oracle of the demagexch module (test infrastructure only).

* Exch6NgbrKernel (modules/kernel_exchange.go:102-124): the 6-neighbour
  Laplacian on the padded grid, float32, the same for the three diagonal
  components (FullTensorIdx[s][s], common/index.go:41-44).  Its seven values
  are assigned one after another at wrapped positions (kernel.go:19-27), so a
  later one replaces an earlier one at the same position: on a length-1 axis
  the centre ends up 1/h_x^2, on a periodic length-2 axis each neighbour holds
  1/h^2 once.
* dexKernUpdater.Update (mod_demagexchange.go:134-149): dex =
  float32(Ms K_d + 2 Aex / (mu0 Ms) K_ex), K_d the float32 Kernel_Arne at
  demag_acc (default 8, :36-37), Ms the Msat multiplier (a mask is not used).
* H_dex = conv(m, dex) (ConvPlan.Convolve, convplan.go:160-175): m itself,
  not Ms m, convolved with the kernel's transform (real parts, 1/Np).

Pinned in tests/test_oracle_demagexch.py: the kernel's values by hand (the
overwrites included), its transform against the closed form
sum_d (2 cos(2 pi k_d / p_d) - 2) / h_d^2, the convolution with it against the
exchange6 oracle (they differ by exactly -m / h^2 per missing neighbour at the
faces), and the combination formula by linearity of the transform.
"""
import numpy as np

from oracle import oracle as O

MU0 = 4.0 * np.pi * 1e-7       # common/constants.go:18


def exch6_kernel(p, cell):
    """(9, p0, p1, p2) float32: the exchange kernel on the padded grid p."""
    h = [c * c for c in cell]
    seq = [((0, 0, 0), -2.0 / h[0] - 2.0 / h[1] - 2.0 / h[2]),
           ((1, 0, 0), 1.0 / h[0]), ((-1, 0, 0), 1.0 / h[0]),
           ((0, 1, 0), 1.0 / h[1]), ((0, -1, 0), 1.0 / h[1]),
           ((0, 0, 1), 1.0 / h[2]), ((0, 0, -1), 1.0 / h[2])]
    K = np.zeros((9,) + tuple(p), np.float32)
    for s in range(3):
        for d, v in seq:
            K[s][d[0] % p[0], d[1] % p[1], d[2] % p[2]] = np.float32(v)
    return K


def dex_kernel(p, cell, pbc, msat, aex, accuracy=8):
    """(9, p0, p1, p2) float32 real-space demag + exchange kernel."""
    kd = O.kernel_arne(p, cell, pbc, accuracy).astype(np.float64)
    ke = exch6_kernel(p, cell).astype(np.float64)
    return (msat * kd + ((2.0 * aex) / (MU0 * msat)) * ke).astype(np.float32)


def dex_khat(p, cell, pbc, msat, aex, accuracy=8):
    """Transformed dex kernel (6, p0, p1, p2//2+1) float64, 1/Np included."""
    kh, bad = O.kernel_fft(p, K9=dex_kernel(p, cell, pbc, msat, aex, accuracy))
    return kh


def h_dex(n, pbc, cell, msat, aex, m, accuracy=8):
    """H_dex (3, n0, n1, n2) in A/m at magnetisation m."""
    p = O.pad_size(n, pbc)
    return O.demag_conv(n, p, dex_khat(p, cell, pbc, msat, aex, accuracy), m)


def folded_khat(n, pbc, cell, msat, aex, accuracy=6, dex_accuracy=8):
    """The transform whose Ms-scaled convolution is H_demag + H_dex (no Msat
    mask): khat(K_d at accuracy) + khat(dex) / Ms, for oracle.Problem(kernel=)."""
    p = O.pad_size(n, pbc)
    kd, _ = O.kernel_fft(p, K9=O.kernel_arne(p, cell, pbc, accuracy))
    return kd + dex_khat(p, cell, pbc, msat, aex, dex_accuracy) / msat
