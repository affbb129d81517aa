"""Pins of the surrogate oracle (oracle/surrogate.py): the convolution against
a direct per-pixel sum, its backward passes against finite differences of the
FP64 loss (no bf16 rounding), bf16 rounding against known bit patterns, the
coarsen / upsample pair, and the residual target's closed form."""
import numpy as np
import pytest

from oracle import surrogate as S


def test_bf16_rounding():
    x = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -9, -2.5, 1e-3, 65504.0])
    r = S.bf16(x)
    assert r[0] == 1.0
    assert r[1] == 1.0                      # tie -> even
    assert r[2] == 1.0 + 2 ** -7            # above the tie -> up
    assert r[3] == -2.5
    assert abs(r[4] - 1e-3) < 1e-3 * 2 ** -8
    assert np.all(np.abs(r - x) <= np.abs(x) * 2 ** -8)


def test_conv_against_direct_sum():
    rng = np.random.default_rng(0)
    x = rng.normal(size=(5, 7, 4))
    w = rng.normal(size=(3, 3, 4, 3))
    b = rng.normal(size=3)
    out = S.conv3x3(x, w, b)
    for (y, xx) in [(0, 0), (2, 3), (4, 6), (0, 6)]:
        want = b.copy()
        for dy in range(3):
            for dx in range(3):
                yy, xs = y + dy - 1, xx + dx - 1
                if 0 <= yy < 5 and 0 <= xs < 7:
                    want += x[yy, xs] @ w[dy, dx]
        np.testing.assert_allclose(out[y, xx], want, rtol=1e-12)


def test_backward_against_finite_differences():
    """d loss / d W and d loss / d b of the FP64 net (no rounding) by central
    differences at random coordinates of every layer."""
    rng = np.random.default_rng(1)
    n0 = 2
    Ws, bs = S.init_params(seed=3, n0=n0)
    Ws[3] = Ws[3] * 10                       # not tiny, so the loss is sensitive
    x0 = rng.normal(size=(6, 8, 32))
    x0[:, :, 6 * n0 + 2:] = 0
    tgt = rng.normal(size=(6, 8, 3 * n0)) * 0.1

    def loss_of(Ws_, bs_):
        out, _ = S.forward(Ws_, bs_, x0, rnd=None)
        return S.loss_and_grad_out(out, tgt, 3 * n0, rnd=None)[0]

    out, acts = S.forward(Ws, bs, x0, rnd=None)
    _, g = S.loss_and_grad_out(out, tgt, 3 * n0, rnd=None)
    dWs, dbs = S.backward(Ws, acts, g, rnd=None)
    eps = 1e-6
    for l in range(4):
        for _ in range(6):
            idx = tuple(rng.integers(0, s) for s in Ws[l].shape)
            if l == 0 and idx[2] >= 6 * n0 + 2:
                continue
            Wp = [w.copy() for w in Ws]
            Wm = [w.copy() for w in Ws]
            Wp[l][idx] += eps
            Wm[l][idx] -= eps
            fd = (loss_of(Wp, bs) - loss_of(Wm, bs)) / (2 * eps)
            assert abs(fd - dWs[l][idx]) <= 1e-6 * max(1e-3, abs(fd)) + 1e-9, (l, idx, fd, dWs[l][idx])
        j = int(rng.integers(0, bs[l].shape[0]))
        if l == 3 and j >= 3 * n0:
            j = 0
        bp = [b.copy() for b in bs]
        bm = [b.copy() for b in bs]
        bp[l][j] += eps
        bm[l][j] -= eps
        fd = (loss_of(Ws, bp) - loss_of(Ws, bm)) / (2 * eps)
        assert abs(fd - dbs[l][j]) <= 1e-6 * max(1e-3, abs(fd)) + 1e-9


def test_coarsen_upsample_and_target():
    rng = np.random.default_rng(2)
    n = (4, 6, 8)
    m = rng.normal(size=(3,) + n)
    mc = S.coarsen(m, n)
    assert mc.shape == (3, 4, 3, 4)
    assert np.isclose(mc[1, 1, 2, 3], m[1, 1, 4:6, 6:8].mean())
    up = S.upsample(mc, n)
    assert up.shape == m.shape and np.allclose(S.coarsen(up, n), mc)
    film = (1, 4, 4)
    mf = rng.normal(size=(3,) + film)
    assert S.coarsen(mf, film).shape == (3, 1, 2, 2)
    x = S.pack_input(m, up, 8e5)
    assert x[0, 0, 6 * 4] == -1 and x[1, 0, 6 * 4] == 1 and x[0, 1, 6 * 4 + 1] == 1
    # target + upsampled coarse field = fine field
    bf = rng.normal(size=(3,) + n)
    t = S.residual_target(bf, up, 8e5)
    np.testing.assert_allclose(S.predict_b(np.concatenate([t, np.zeros(t.shape[:2] + (4,))], 2), up, 8e5),
                               bf, rtol=1e-12, atol=1e-15)
