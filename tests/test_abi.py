"""The C-ABI library builds, loads without a GPU and exports every entry point
that include/mmx.h declares (no compute calls here)."""
import ctypes
import glob
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADERS = sorted(glob.glob(os.path.join(ROOT, "include", "*.h")))
LIB = os.path.join(ROOT, "2_amd", "mmx", "libmmx.so")


def _declared():
    names = set()
    for h in HEADERS:
        text = re.sub(r"/\*.*?\*/", "", open(h).read(), flags=re.S)
        names.update(re.findall(r"\b(mmx_[A-Za-z0-9_]+)\s*\(", text))
    return sorted(names)


@pytest.fixture(scope="module")
def built():
    if not os.path.exists(LIB):
        subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "2_amd", "csrc")], check=True)
    return LIB


def test_header_declares_entry_points():
    names = _declared()
    assert len(names) >= 40
    for must in ("mmx_exchange6", "mmx_torque", "mmx_normalize", "mmx_zhangli",
                 "mmx_uniaxial_anisotropy", "mmx_demag_update_b", "mmx_sim_step",
                 "mmx_reductor_maxdiff", "mmx_madd2", "mmx_vec_madd",
                 # libmumax2's own signatures (include/mmx_libmumax2.h)
                 "mmx_torqueAsync", "mmx_exchange6Async", "mmx_kernelMulMicromag3DAsync",
                 "mmx_partialMaxDiffAsync", "mmx_setUsedGPUs", "mmx_convplan_convolve"):
        assert must in names


def test_library_exports_every_declared_symbol(built):
    out = subprocess.run(["nm", "-D", "--defined-only", built], capture_output=True, text=True,
                         check=True).stdout
    exported = set(re.findall(r"\s[TW]\s+(mmx_[A-Za-z0-9_]+)$", out, flags=re.M))
    missing = [n for n in _declared() if n not in exported]
    assert not missing, missing


def test_python_binding_covers_header(built):
    import mmx
    assert sorted(mmx.EXPORTED_SYMBOLS) == _declared()


def test_library_loads_without_gpu(built):
    lib = ctypes.CDLL(built)
    lib.mmx_version.restype = ctypes.c_char_p
    lib.mmx_last_error.restype = ctypes.c_char_p
    assert b"gfx950" in lib.mmx_version()
    assert lib.mmx_last_error() is not None


def test_product_does_not_touch_oracle():
    """The product sources never import, load or link the oracle."""
    for d in ("2_amd", "include"):
        for dirpath, _, files in os.walk(os.path.join(ROOT, d)):
            for f in files:
                if f.endswith((".py", ".hip", ".cpp", ".hpp", ".h", "Makefile")):
                    txt = open(os.path.join(dirpath, f)).read()
                    for pat in ("import oracle", "from oracle", "libmmx_oracle", "orc_",
                                "mmx_oracle"):
                        assert pat not in txt, (os.path.join(dirpath, f), pat)


def test_libmumax2_entry_points_abort_on_bad_arguments(built):
    """The libmumax2-shaped calls return nothing and abort on an error, as the
    reference's gpu_safe does (gpu_safe.h:36-41); checked before any HIP call."""
    code = ("import ctypes as C; L = C.CDLL(%r); "
            "L.mmx_setUsedGPUs((C.c_int * 2)(0, 1), 2)" % built)
    r = subprocess.run(["python", "-c", code], capture_output=True, text=True)
    assert r.returncode == -6, (r.returncode, r.stderr)
    assert "setUsedGPUs" in r.stderr and "one GPU" in r.stderr
