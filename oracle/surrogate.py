"""FP64 oracle of the learned demag correction (test infrastructure only).

No reference counterpart: the reference has no learned component.  This is
the CPU restatement of the build's surrogate, used by tests/ to check the GPU
kernels (forward, backward, gradients) and pinned by its own finite-difference
gradient check and closed-form cases (tests/test_oracle_surrogate.py).

Model (internal z,y,x order, film of n0 layers on an n1 x n2 in-plane grid):

* coarse grid: the film plane halved (in-plane cell sizes doubled), the n0
  layers kept; coarse m = mean of the 2x2 fine cells of a layer;
* net input per in-plane pixel (y, x): 32 channels, channel c*n0 + i = m_c at
  layer i, 3*n0 + c*n0 + i = upsampled coarse B_c / (mu0 Ms) at layer i
  (piecewise constant), 6*n0 and 6*n0 + 1 = the position inside the coarse
  cell (+1 / -1 for odd / even y and x), the rest 0;
* 4 conv layers 3x3, zero padding: 32 -> 32 -> 32 -> 32 -> 16, ReLU after the
  first three; output channel c*n0 + i = residual (B_fine - B_coarse_up)_c
  / (mu0 Ms) at layer i (channels 3*n0.. unused);
* bf16 rounding where the GPU rounds: the packed input, the weights, the
  ReLU outputs, the output gradient and the back-propagated gradients; all
  sums in FP32 on the GPU, FP64 here.
* loss = sum (out - target)^2 / sum target^2 over pixels and the 3*n0 used
  channels (relative squared error of the sample).

Weights: W[l] shape (3, 3, Cin, Cout), out[y, x] = b + sum_{dy,dx} in[y+dy-1,
x+dx-1] @ W[dy, dx].
"""
import numpy as np

MU0 = 4.0 * np.pi * 1e-7
CH = (32, 32, 32, 32, 16)


def bf16(x):
    """Round to bfloat16 (nearest even) and return as float64."""
    f = np.asarray(x, dtype=np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def coarse_shape(n):
    return (n[0], n[1] // 2, n[2] // 2)


def coarsen(m, n):
    """(3, n0, n1, n2) -> (3, n0, n1/2, n2/2), means of 2x2 in-plane blocks."""
    n0, n1, n2 = n
    return m.reshape(3, n0, n1 // 2, 2, n2 // 2, 2).mean(axis=(3, 5))


def upsample(bc, n):
    return np.repeat(np.repeat(bc, 2, axis=2), 2, axis=3)


def pack_input(m, bc_up, msat):
    """(3,n0,n1,n2) fields -> (n1, n2, 32) net input (bf16-rounded)."""
    _, n0, n1, n2 = m.shape
    x = np.zeros((n1, n2, 32))
    for c in range(3):
        for i in range(n0):
            x[:, :, c * n0 + i] = m[c, i]
            x[:, :, 3 * n0 + c * n0 + i] = bc_up[c, i] / (MU0 * msat)
    x[:, :, 6 * n0] = np.where(np.arange(n1) % 2 == 1, 1.0, -1.0)[:, None]
    x[:, :, 6 * n0 + 1] = np.where(np.arange(n2) % 2 == 1, 1.0, -1.0)[None, :]
    return bf16(x)


def conv3x3(x, w, b=None):
    H, W, _ = x.shape
    xp = np.pad(x, ((1, 1), (1, 1), (0, 0)))
    out = np.zeros((H, W, w.shape[3]))
    for dy in range(3):
        for dx in range(3):
            out += xp[dy:dy + H, dx:dx + W, :] @ w[dy, dx]
    if b is not None:
        out += b
    return out


def conv3x3_backward_data(dz, w):
    """d(in) for out = conv3x3(in, w): the transposed convolution."""
    H, W, _ = dz.shape
    dzp = np.pad(dz, ((1, 1), (1, 1), (0, 0)))
    dx_ = np.zeros((H, W, w.shape[2]))
    for dy in range(3):
        for dx in range(3):
            # in[y + dy - 1] contributed to out[y]: in[q] gets dz[q - dy + 1]
            dx_ += dzp[2 - dy:2 - dy + H, 2 - dx:2 - dx + W, :] @ w[dy, dx].T
    return dx_


def conv3x3_backward_weight(x, dz):
    H, W, _ = x.shape
    xp = np.pad(x, ((1, 1), (1, 1), (0, 0)))
    dw = np.zeros((3, 3, x.shape[2], dz.shape[2]))
    for dy in range(3):
        for dx in range(3):
            dw[dy, dx] = xp[dy:dy + H, dx:dx + W, :].reshape(-1, x.shape[2]).T @ \
                dz.reshape(-1, dz.shape[2])
    return dw, dz.sum(axis=(0, 1))


def init_params(seed=0, n0=4):
    """He-initialised weights (first layer over its 6*n0 + 2 used inputs), a small
    last layer (initial residual ~ 0), zero biases."""
    rng = np.random.default_rng(seed)
    Ws, bs = [], []
    for l in range(4):
        cin, cout = CH[l], CH[l + 1]
        fan = 9 * (6 * n0 + 2 if l == 0 else cin)
        w = rng.normal(size=(3, 3, cin, cout)) * np.sqrt(2.0 / fan)
        if l == 3:
            w *= 0.1
        Ws.append(w)
        bs.append(np.zeros(cout))
    return Ws, bs


def _ident(x):
    return np.asarray(x, dtype=np.float64)


def forward(Ws, bs, x0, rnd=bf16):
    """Returns (out FP64 (H, W, 16), activations [x0, a1, a2, a3]) with the
    GPU's bf16 roundings (rnd=None: none, for finite-difference checks)."""
    rnd = rnd or _ident
    acts = [x0]
    a = x0
    for l in range(3):
        z = conv3x3(a, rnd(Ws[l]), bs[l])
        a = rnd(np.maximum(z, 0.0))
        acts.append(a)
    out = conv3x3(a, rnd(Ws[3]), bs[3])
    return out, acts


def loss_and_grad_out(out, target, nused, rnd=bf16):
    """Relative squared error over the nused channels; d loss / d out
    rounded to bf16."""
    rnd = rnd or _ident
    d = out[:, :, :nused] - target
    t2 = float((np.asarray(target, np.float64) ** 2).sum())
    loss = float((d * d).sum()) / t2
    g = np.zeros_like(out)
    g[:, :, :nused] = 2.0 * d / t2
    return loss, rnd(g)


def backward(Ws, acts, gout, rnd=bf16):
    """Gradients (dW list, db list) given bf16 d loss / d out."""
    rnd = rnd or _ident
    dWs, dbs = [None] * 4, [None] * 4
    dz = gout
    for l in range(3, -1, -1):
        dWs[l], dbs[l] = conv3x3_backward_weight(acts[l], dz)
        if l > 0:
            dx_ = conv3x3_backward_data(dz, rnd(Ws[l]))
            dz = rnd(dx_ * (acts[l] > 0))
    return dWs, dbs


def residual_target(b_fine, bc_up, msat):
    """(3,n0,n1,n2) B fields -> (n1, n2, 3*n0) target channels."""
    _, n0, n1, n2 = b_fine.shape
    t = np.zeros((n1, n2, 3 * n0))
    for c in range(3):
        for i in range(n0):
            t[:, :, c * n0 + i] = (b_fine[c, i] - bc_up[c, i]) / (MU0 * msat)
    return t


def predict_b(out, bc_up, msat):
    """Net output -> fine B estimate (3, n0, n1, n2)."""
    _, n0, n1, n2 = bc_up.shape
    b = bc_up.copy()
    for c in range(3):
        for i in range(n0):
            b[c, i] += MU0 * msat * out[:, :, c * n0 + i]
    return b
