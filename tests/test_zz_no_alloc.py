"""GPU: rgf_solve_batched_noalloc (the allocation-free entry point, workspace in a
fixed device array) through the C ABI against rgf_solve_batched on the same
inputs, two calls in a row, the second with a smaller batch; and the rejection
of a batch whose single energy slot exceeds the fixed workspace."""
import ctypes

import numpy as np
import pytest

from liteqt_amd import configs

pytestmark = pytest.mark.gpu

OUTS = ("ldos", "ndens", "pdens", "GRi1", "GLi1", "GGi1", "cur", "tr", "tre", "dens_n", "dens_p", "info")


def _rgf():
    import torch
    if not torch.cuda.is_available():
        pytest.fail("no HIP device visible")
    from liteqt_amd import rgf
    return rgf, torch


def _outputs(torch, nE, nx, nm, dev):
    c = lambda: torch.full((nE, nx, nm, nm), float("nan"), dtype=torch.complex128, device=dev)
    return {"ldos": c(), "ndens": c(), "pdens": c(), "GRi1": c(), "GLi1": c(), "GGi1": c(),
            "cur": torch.full((nE, nx, nm, nm), float("nan"), dtype=torch.float64, device=dev),
            "tr": torch.zeros(nE, dtype=torch.float64, device=dev),
            "tre": torch.zeros(nE, dtype=torch.float64, device=dev),
            "dens_n": torch.zeros(nE, nx, dtype=torch.float64, device=dev),
            "dens_p": torch.zeros(nE, nx, dtype=torch.float64, device=dev),
            "info": torch.zeros(nE, 4, dtype=torch.int32, device=dev)}


def _solve(rgf, torch, fn, p, E, temp, mu, Hd, Cd, sL, sR, out):
    b = rgf._batch(torch, E, temp, mu, Hd, Cd, sL, sR, 0)
    rgf._out_batch(torch, b, out, E.device)
    stream = torch.cuda.current_stream()
    rgf._check(fn(ctypes.byref(p), ctypes.byref(b), ctypes.c_void_p(stream.cuda_stream)))
    torch.cuda.synchronize()


def _case(rgf, torch, nm, nx, nE, mode, seed):
    dev = torch.device("cuda:0")
    Hii, H1i = configs.random_device(nm, nx, seed)
    E = configs.energy_grid(-1.5, 1.5, nE)
    sL, sG, sR = configs.scattering_self_energies(nm, nx, seed + 1, E=0.0, scale=0.01)
    p = rgf.params(mode, eta_dev=1e-4, eta_lead=1e-4, dec_tol=1e-12)
    args = (p, torch.tensor(E, device=dev), (300.0, 250.0), (0.1, -0.1), rgf.to_device_blocks(Hii, dev),
            rgf.to_device_blocks(H1i, dev), rgf.to_device_blocks(sL, dev), rgf.to_device_blocks(sR, dev))
    ref = _outputs(torch, nE, nx, nm, dev)
    got = _outputs(torch, nE, nx, nm, dev)
    L = rgf.lib()
    _solve(rgf, torch, L.rgf_solve_batched, *args, ref)
    print(f"no-alloc: nm={nm} nx={nx} nE={nE} mode={mode}", flush=True)
    _solve(rgf, torch, L.rgf_solve_batched_noalloc, *args, got)
    for k in OUTS:
        a, r = got[k], ref[k]
        if k in ("GRi1", "GLi1", "GGi1", "cur"):  # block nx is not written by either entry point
            a, r = a[:, : nx - 1], r[:, : nx - 1]
        if k != "info":
            assert torch.isfinite(r).all(), k
        # the same kernels on the same inputs: identical to the last bit
        assert torch.equal(a, r), (k, nm, nx, nE)
    return ref


@pytest.mark.parametrize("mode", [0, 1])
def test_noalloc_equals_allocating_entry_two_calls(mode):
    rgf, torch = _rgf()
    _case(rgf, torch, 64, 6, 40, mode, 11)
    # second call, smaller batch: the fixed workspace keeps the first call's contents
    _case(rgf, torch, 32, 3, 7, mode, 12)


def test_noalloc_rejects_slot_beyond_fixed_workspace():
    rgf, torch = _rgf()
    dev = torch.device("cuda:0")
    nm, nx, nE = 256, 600, 2  # one slot: (2*600 + 30) MiB > 1 GiB
    p = rgf.params(rgf.MODE_COMPAT)
    E = torch.zeros(nE, dtype=torch.float64, device=dev)
    H = torch.zeros(nx, nm, nm, dtype=torch.complex128, device=dev)
    b = rgf._batch(torch, E, (300.0, 300.0), (0.0, 0.0), H, H, None, None, 0)
    rc = rgf.lib().rgf_solve_batched_noalloc(ctypes.byref(p), ctypes.byref(b), None)
    assert rc == 4, rc  # RGF_EUNSUP, nothing launched
    msg = rgf.lib().rgf_last_error().decode()
    assert "fixed workspace" in msg, msg
