"""The GW-step library (liteqt_amd/libgw_hip.so) loads, exports every function
include/gw.h declares, and rejects invalid sizes before any GPU call (runs without a GPU)."""
import ctypes
import math
import os
import re

import pytest

from liteqt_amd import gw

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_functions():
    src = open(os.path.join(ROOT, "include", "gw.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(gw_[a-z_]+)\s*\(", src)))


def test_library_exports_every_declared_symbol():
    if not os.path.exists(gw.LIB_PATH):
        pytest.fail("libgw_hip.so not built (run __graft_entry__.build())")
    names = header_functions()
    assert names == ["gw_last_error", "gw_pairs_polarization", "gw_pairs_selfenergy", "gw_params_reference",
                     "gw_polarization", "gw_polarization_fft", "gw_release", "gw_selfenergy", "gw_selfenergy_fft",
                     "gw_w_full"]
    L = ctypes.CDLL(gw.LIB_PATH)
    for name in names:
        assert hasattr(L, name), name


def test_reference_parameters_of_the_one_dimensional_driver():
    """src/green_rgf.f90:3192-3193 and 3227-3229: nopmax = nen/2 - 10,
    dE_P = (0, -dE/2/pi) * spindeg, dE_S = (0, dE/2/pi)."""
    p = gw.params_reference(128, 18, 2000, 1, 0.01, 2.0, 0.3)
    assert (p.nm, p.nx, p.nen, p.ndiag, p.nopmax) == (128, 18, 2000, 1, 990)
    assert p.dE_P[0] == 0.0 and p.dE_P[1] == -1.0 * 0.01 / 2.0 / math.pi * 2.0
    assert p.dE_S[0] == 0.0 and p.dE_S[1] == 0.01 / 2.0 / math.pi and p.alpha_mix == 0.3


def test_pair_counts():
    L = gw.lib()
    # P: sum over |nop| <= nopmax of (nen - |nop|)
    assert L.gw_pairs_polarization(10, 3) == 10 + 2 * (9 + 8 + 7)
    # Sigma: ie = 2..nen, nop in [max(-nopmax, ie - nen + 1), min(nopmax, ie - 1)] (1-based)
    nen, nopmax = 12, 4
    n = sum(len([nop for nop in range(-nopmax, nopmax + 1) if ie > max(nop, 1) and ie < nen + nop])
            for ie in range(1, nen + 1))
    assert L.gw_pairs_selfenergy(nen, nopmax) == n


@pytest.mark.parametrize("fn,kw,msg", [
    ("gw_polarization", dict(nopmax=1000), "nopmax"),
    ("gw_polarization", dict(ndiag=-1), "ndiag"),
    ("gw_selfenergy", dict(nx=4), "nx >= 5"),
    ("gw_polarization", dict(), "null array"),
    # above the direct forms' limits the FFT form takes over (no size error), up to
    # nen + nopmax = 2^24 points per series
    ("gw_polarization", dict(nen=3074), "null array"),
    ("gw_selfenergy", dict(nen=2042), "null array"),
    ("gw_polarization", dict(nen=1 << 24, nopmax=2), "2^24"),
    ("gw_selfenergy", dict(nen=(1 << 24) - 2, nopmax=3), "2^24"),
    ("gw_polarization_fft", dict(nen=(1 << 24) - 8, nopmax=8), "null array"),
    ("gw_polarization_fft", dict(nen=(1 << 24) - 8, nopmax=9), "2^24"),
    ("gw_selfenergy_fft", dict(nx=4), "nx >= 5"),
    ("gw_selfenergy_fft", dict(), "null array"),
])
def test_invalid_arguments_fail_before_any_launch(fn, kw, msg):
    args = dict(nm=4, nx=6, nen=2000, ndiag=1, nopmax=990)
    args.update(kw)
    p = gw.params(args["nm"], args["nx"], args["nen"], args["ndiag"], args["nopmax"], -1j, 1j, 0.5)
    n = 7 if fn.startswith("gw_polarization") else 10
    rc = getattr(gw.lib(), fn)(ctypes.byref(p), *([None] * n))
    assert rc != 0 and msg in gw.lib().gw_last_error().decode()
