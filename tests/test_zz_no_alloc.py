"""mmx_reduce_noalloc (the allocation-free reduction) against the reductor.

The entry point keeps its partials in a device array built into the library
instead of a reductor's hipMalloc'd buffer; its results must be the
reductor's, bit for bit, for every operation.  Two calls in a row, the second
on a smaller input, show that what the first call left in the shared array
does not reach the second result.  Sizes beyond the stated maximum are refused
before anything runs.
"""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

OPS = ("sum", "dot", "max", "min", "maxabs", "maxdiff", "maxnorm")


def _say(*a):
    print("[case]", *a, flush=True)


@pytest.fixture(scope="module")
def torch_gpu():
    import torch
    assert torch.cuda.is_available(), "GPU tests need a GPU"
    import mmx
    mmx.check(mmx.lib().mmx_init(0))
    return torch


def _noalloc(op, a, b, c, n):
    import mmx
    out = C.c_double(np.nan)
    code = OPS.index(op)
    mmx.check(mmx.lib().mmx_reduce_noalloc(code, a, b, c, n, None, C.byref(out)))
    return out.value


def _reductor(r, op, a, b, c, n):
    return {"sum": lambda: r.sum(a, n), "dot": lambda: r.dot(a, b, n),
            "max": lambda: r.max(a, n), "min": lambda: r.min(a, n),
            "maxabs": lambda: r.maxabs(a, n), "maxdiff": lambda: r.maxdiff(a, b, n),
            "maxnorm": lambda: r.maxnorm(a, b, c, n)}[op]()


@pytest.mark.parametrize("sizes", [(3000001, 777), (524288 * 4 + 5, 1025), (1 << 22, 1)])
def test_noalloc_matches_reductor_two_calls(torch_gpu, sizes):
    import mmx
    torch = torch_gpu
    rng = np.random.default_rng(sizes[0] % 1000)
    for n in sizes:  # the second call is smaller: fewer partials than the first left behind
        _say("noalloc n", n)
        host = [rng.normal(size=n).astype(np.float32) for _ in range(3)]
        a, b, c = [torch.from_numpy(h).cuda() for h in host]
        pa, pb, pc = a.data_ptr(), b.data_ptr(), c.data_ptr()
        torch.cuda.synchronize()
        r = mmx.Reductor(n)
        for op in OPS:
            got = _noalloc(op, pa, pb, pc, n)
            want = float(_reductor(r, op, pa, pb, pc, n))
            assert got == want, (op, n, got, want)
        # and against numpy (FP64 sums; exact extrema)
        h64 = [h.astype(np.float64) for h in host]
        assert abs(_noalloc("sum", pa, pb, pc, n) - h64[0].sum()) <= 1e-9 * max(1.0, np.abs(h64[0]).sum())
        assert _noalloc("maxdiff", pa, pb, pc, n) == float(np.abs(host[0] - host[1]).max())
        assert _noalloc("max", pa, pb, pc, n) == float(host[0].max())
        del r


def test_noalloc_refuses_beyond_maximum(torch_gpu):
    import mmx
    torch = torch_gpu
    a = torch.zeros(16, device="cuda")
    out = C.c_double(0.0)
    rc = mmx.lib().mmx_reduce_noalloc(0, a.data_ptr(), None, None, 2147483648, None, C.byref(out))
    assert rc != 0
    assert b"no workspace" in mmx.lib().mmx_last_error()
    rc = mmx.lib().mmx_reduce_noalloc(1, a.data_ptr(), None, None, 16, None, C.byref(out))
    assert rc != 0  # dot without b
    assert _noalloc("sum", a.data_ptr(), None, None, 16) == 0.0
