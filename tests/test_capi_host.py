"""CPU: the C-ABI library loads, exports every entry point include/pisco.h
declares, and its host-side paths (argument and configuration checks,
geometry helpers) behave as documented -- without touching a GPU."""
import ctypes
import os
import re

import pytest

from pisco4cmb_amd import _lib, pypisco

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    text = open(os.path.join(ROOT, "include", "pisco.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(pisco_\w+)\s*\(", text)))


def test_every_declared_symbol_is_exported_and_bound():
    names = declared_functions()
    assert len(names) >= 29
    raw = ctypes.CDLL(_lib.LIB_PATH)
    for n in names:
        assert hasattr(raw, n), n
        assert n in _lib.SIGNATURES, n


def test_no_cpu_fallback_in_product():
    """The product never imports the oracle."""
    pkg = os.path.join(ROOT, "pisco4cmb_amd")
    for dirpath, _, files in os.walk(pkg):
        for f in files:
            if f.endswith((".py", ".hip", ".hpp", ".h", ".cpp")):
                src = open(os.path.join(dirpath, f)).read()
                assert "oracle_bind" not in src and "liboracle" not in src, f


def test_create_rejects_invalid_config_before_device_use():
    """GPUConvolver's memory check (gpuconvolver.cu:41-64) and ptgPerConv."""
    lib = _lib.lib
    ctx = ctypes.c_void_p()
    cfg = pypisco.default_run_config(ptgPerConv=1000, pixelsPerDisc=1000, maxMemUsage=1000)
    assert lib.pisco_create(ctypes.byref(cfg), ctypes.byref(ctx)) == _lib.PISCO_ERR_MEMORY
    assert b"invalid configuration" in lib.pisco_last_error()
    cfg = pypisco.default_run_config(ptgPerConv=0)
    assert lib.pisco_create(ctypes.byref(cfg), ctypes.byref(ctx)) == _lib.PISCO_ERR_INVALID
    assert lib.pisco_create(None, ctypes.byref(ctx)) == _lib.PISCO_ERR_INVALID
    conv = pypisco.GPUConvolver(pypisco.default_run_config(ptgPerConv=10, pixelsPerDisc=10,
                                                           maxMemUsage=10))
    with pytest.raises(ValueError):
        conv.configure_execution()


def test_null_context_and_flags_are_errors():
    lib = _lib.lib
    assert lib.pisco_exec_convolution(None, b"p", None, None) == _lib.PISCO_ERR_INVALID
    assert lib.pisco_update_sky(None, 64, None, None, None, None) == _lib.PISCO_ERR_INVALID
    assert lib.pisco_set_tuning(None, b"adjmap_chunk", 1) == _lib.PISCO_ERR_INVALID
    lib.pisco_destroy(None)


def test_geometry_helpers():
    lib = _lib.lib
    assert lib.pisco_beam_rho_max(0, 10) == -1.0
    assert lib.pisco_beam_rho_max(16, 12 * 16 * 16) == -1.0
    assert lib.pisco_ang2pix(1, 0.5 * 3.141592653589793, 0.0) == 4
    th, ph = pypisco.pix2ang(1, 0)
    assert abs(th - 0.8410686705679303) < 1e-15 and abs(ph - 0.7853981633974483) < 1e-15
    # fields2beams.py:102 sizing: 5 deg at NSIDE 4096
    assert pypisco.beam_npix_for_radius(4096, 5 * 3.141592653589793 / 180) == 382812
    with pytest.raises(ValueError):
        pypisco.PolBeam(64, 10, 0.0, "x")
