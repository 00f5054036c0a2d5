"""The C-ABI library loads and exports every function include/rgf.h declares
(no compute call: runs without a GPU)."""
import ctypes
import os
import re

import pytest

from liteqt_amd import rgf

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_functions():
    src = open(os.path.join(ROOT, "include", "rgf.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(rgf_[a-z_]+)\s*\(", src)))


def test_header_declares_expected_entry_points():
    assert header_functions() == sorted(rgf.EXPORTS)


def test_library_exports_every_declared_symbol():
    if not os.path.exists(rgf.LIB_PATH):
        pytest.fail("librgf_hip.so not built (run __graft_entry__.build())")
    L = ctypes.CDLL(rgf.LIB_PATH)
    for name in header_functions():
        assert hasattr(L, name), name


def test_reference_parameters():
    p = rgf.params()
    assert p.eta_dev == 0.0 and p.eta_lead == 1e-3 and p.dec_tol == 1e-100
    assert p.dec_nmax == 100 and p.boltz == 8.61734e-05 and p.mode == rgf.MODE_COMPAT


def test_workspace_bytes():
    # (2*nx + temporaries) blocks of nm^2 complex128 per energy slot
    b1 = rgf.lib().rgf_workspace_bytes(256, 200, 1)
    assert b1 % (256 * 256 * 16) == 0 and b1 // (256 * 256 * 16) >= 400
    assert rgf.lib().rgf_workspace_bytes(256, 200, 10) == 10 * b1


def test_tensor_checks_reject_wrong_dtype_layout_and_size():
    """Raw pointers cross the C ABI: the binding refuses tensors whose dtype,
    layout or size the engine would misread (no GPU call is made)."""
    import torch
    from liteqt_amd import rgf
    cpu = torch.device("cpu")
    ok = torch.zeros(4, dtype=torch.float64)
    assert rgf._req(ok, "w", torch.float64, 4, cpu) == ok.data_ptr()
    assert rgf._req(None, "w", torch.float64, 4, cpu) is None
    for bad, why in ((torch.zeros(4, dtype=torch.float32), "expected"),
                     (torch.zeros(4, 2, dtype=torch.float64).t(), "contiguous"),
                     (torch.zeros(3, dtype=torch.float64), "elements")):
        with pytest.raises(rgf.RgfError, match=why):
            rgf._req(bad, "w", torch.float64, 4, cpu)
    # a host tensor for a device output: refused unless the output may live in pinned
    # host memory and the tensor is pinned (an ordinary host tensor is not)
    gpu = torch.device("cuda:0")
    with pytest.raises(rgf.RgfError, match="expected cuda"):
        rgf._req(ok, "w", torch.float64, 4, gpu)
    with pytest.raises(rgf.RgfError, match="expected cuda"):
        rgf._req(torch.zeros(8, dtype=torch.int32), "info", torch.int32, 8, gpu, host_ok=True)


def test_offdiag_and_scba_argument_errors_before_any_device_call():
    """Argument errors of the block-tridiagonal-Sigma and SCBA entry points are
    reported before the device is touched (dummy pointers are never read)."""
    from liteqt_amd import rgf as r
    L = r.lib()
    dummy = ctypes.c_void_p(16)
    b = r.Batch()
    b.nm, b.nx, b.nE = 32, 4, 2
    b.E = b.Hii = b.H1i = dummy.value
    b.sigR_lo = dummy.value
    p = r.params(r.MODE_COMPAT)
    assert L.rgf_solve_batched(ctypes.byref(p), ctypes.byref(b), None) == 1  # RGF_EINVAL: needs physical
    assert b"PHYSICAL" in L.rgf_last_error()
    p = r.params(r.MODE_PHYSICAL)
    sc = r.Scba(1e-3, 0.5)
    b.sigR_lo = None
    assert L.rgf_scba_step(ctypes.byref(p), ctypes.byref(b), ctypes.byref(sc), None) == 1  # sigR/sigL missing
    b.sigR = b.sigL = dummy.value
    b.sig_stride = 0
    assert L.rgf_scba_step(ctypes.byref(p), ctypes.byref(b), ctypes.byref(sc), None) == 1  # shared Sigma
    assert b"per energy" in L.rgf_last_error()
    h = ctypes.c_void_p()
    assert L.rgf_graph_create(ctypes.byref(p), ctypes.byref(b), None, None, ctypes.byref(h)) == 1
    assert b"stream" in L.rgf_last_error() and not h.value


def test_solve_devices_rejects_bad_devices_before_any_device_call():
    from liteqt_amd import rgf as r
    L = r.lib()
    arr = (r.Batch * 2)()
    for b in arr:
        b.nm, b.nx, b.nE = 64, 3, 1
        b.E = b.Hii = b.H1i = 16
    arr[1].nm = 40
    p = r.params(r.MODE_PHYSICAL)
    assert L.rgf_solve_devices(ctypes.byref(p), arr, 2, 0, None, None, None) == 4  # RGF_EUNSUP
    assert b"device 1" in L.rgf_last_error()
    arr[1].nm = 64
    arr[1].nx = 1
    assert L.rgf_solve_devices(ctypes.byref(p), arr, 2, 0, None, None, None) == 1
    arr[1].nx = 3
    arr[0].sigL_up = 16
    assert L.rgf_solve_devices(ctypes.byref(p), arr, 2, 0, None, None, None) == 4
    h = ctypes.c_void_p()
    assert L.rgf_graph_create_devices(ctypes.byref(p), arr, 2, 0, None, ctypes.byref(h)) == 1  # NULL stream


def test_overlap_solve_is_switched_off_before_any_device_call():
    """The overlap (green_RGF_CMS) GPU solve reports RGF_EUNSUP; the check runs
    before the device is touched."""
    from liteqt_amd import rgf as r
    L = r.lib()
    b = r.Batch()
    b.nm, b.nx, b.nE = 32, 3, 1
    b.E = b.Hii = b.H1i = b.Sii = 16
    p = r.params(r.MODE_COMPAT)
    assert L.rgf_solve_batched(ctypes.byref(p), ctypes.byref(b), None) == 4
    assert b"switched off" in L.rgf_last_error()


def test_w_solve_argument_errors_before_any_device_call():
    """rgf_w_solve_batched: open-boundary blocks are unsupported, nx >= 2, every array
    required; all reported before the device is touched."""
    from liteqt_amd import rgf as r
    L = r.lib()
    d = ctypes.c_void_p(16).value
    b = r.WBatch(32, 4, 2, 1, d, d, d, d, d, d, d, d, None, None)
    assert L.rgf_w_solve_batched(ctypes.byref(b), None) == 4  # RGF_EUNSUP
    assert b"NBC" in L.rgf_last_error()
    b.nbc, b.nx = 0, 1
    assert L.rgf_w_solve_batched(ctypes.byref(b), None) == 1
    b.nx, b.PG = 4, None
    assert L.rgf_w_solve_batched(ctypes.byref(b), None) == 1
    assert b"required" in L.rgf_last_error()
    b.PG, b.nm = d, 600
    assert L.rgf_w_solve_batched(ctypes.byref(b), None) == 4
