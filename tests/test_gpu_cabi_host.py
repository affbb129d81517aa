"""The C ABI driven from compiled host code, as the cgo wrappers would.

tests/cabi/engine_rk12.c (plain C, built by __graft_entry__.build() against
include/mmx.h, include/mmx_libmumax2.h and libmmx.so) allocates its own device
buffers and composes the Go engine's H_eff (demag B, exchange6, Zeeman,
SumUpdater) and rk12 tries from the libmumax2-shaped entry points (float**
arrays, hipStream_t* streams: normalizeAsync, torqueAsync, exchange6Async,
vecMaddAsync, maddAsync, madd2Async, partialMaxDiffAsync finished on the
host) and the demag plan, and runs the same tries through the fused
simulation.  Both are replayed by the FP64 oracle: m to 1e-5, the error
estimates to 1e-3 relative.
"""
import os
import struct
import subprocess

import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
EXE = os.path.join(HERE, "cabi", "engine_rk12")


@pytest.mark.parametrize("n,cell", [((2, 16, 32), (3e-9, 3.5e-9, 4e-9)),
                                    ((1, 32, 64), (3e-9, 3.90625e-9, 3.90625e-9))])
def test_engine_sequence_through_the_c_abi(tmp_path, n, cell):
    assert os.access(EXE, os.X_OK), "tests/cabi/engine_rk12 is not built: run __graft_entry__.build()"
    print("[case] C host engine rk12", n, flush=True)
    ms, aex, alpha, bext, dt, ntries = 8e5, 1.3e-11, 0.05, (0.0, 0.004, 0.009), 1e-14, 12
    m = O.normalize(np.random.default_rng(11).normal(size=(3,) + n)).astype(np.float32)
    inp, outp = tmp_path / "in.bin", tmp_path / "out.bin"
    with open(inp, "wb") as f:
        f.write(struct.pack("<3i", *n))
        f.write(struct.pack("<3d", *cell))
        f.write(struct.pack("<6d", ms, aex, alpha, *bext))
        f.write(struct.pack("<d", dt))
        f.write(struct.pack("<i", ntries))
        f.write(m.tobytes())
    r = subprocess.run([EXE, str(inp), str(outp)], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr, flush=True)
    assert r.returncode == 0
    N = int(np.prod(n))
    raw = outp.read_bytes()
    errs = np.frombuffer(raw[:16 * ntries], np.float64).reshape(ntries, 2)
    got = np.frombuffer(raw[16 * ntries:], np.float32).reshape(2, 3, *n)
    prob = O.Problem(n, cell, msat=ms, aex=aex, alpha=alpha, bext=bext)
    mo = m.astype(np.float64).copy()
    for k in range(ntries):
        eo = prob.try_step(mo, dt * (1 + 0.1 * k))
        assert np.all(np.abs(errs[k] - eo) <= 1e-3 * eo + 1e-12), (k, errs[k], eo)
    want = prob.normalized(mo)
    d_engine = np.abs(got[0] - want).max()
    d_sim = np.abs(got[1] - want).max()
    print("[case]  max |dm| engine %.2e fused %.2e" % (d_engine, d_sim), flush=True)
    assert d_engine < 1e-5 and d_sim < 1e-5
