"""ctypes binding of the GW self-energy step (include/gw.h, liteqt_amd/libgw_hip.so).

Arrays are torch complex128 tensors on the GPU in the reference's layout: a Fortran
(nm, nm, nx, nen) array is a C-contiguous tensor of shape (nen, nx, nm, nm)
(`to_device` / `to_host` convert numpy Fortran arrays).  There is no CPU path: the
functions raise when the library is missing.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libgw_hip.so")
_lib = None


class GwError(RuntimeError):
    pass


class Params(ctypes.Structure):
    _fields_ = [("nm", ctypes.c_int), ("nx", ctypes.c_int), ("nen", ctypes.c_int), ("ndiag", ctypes.c_int),
                ("nopmax", ctypes.c_int), ("dE_P", ctypes.c_double * 2), ("dE_S", ctypes.c_double * 2),
                ("alpha_mix", ctypes.c_double)]


def lib():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise GwError(f"GW library missing: {LIB_PATH} (run __graft_entry__.build())")
    L = ctypes.CDLL(LIB_PATH)
    vp = ctypes.c_void_p
    P = ctypes.POINTER(Params)
    L.gw_params_reference.argtypes = [P, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_double,
                                      ctypes.c_double, ctypes.c_double]
    L.gw_polarization.argtypes = [P] + [vp] * 7
    L.gw_selfenergy.argtypes = [P] + [vp] * 10
    L.gw_polarization_fft.argtypes = [P] + [vp] * 7
    L.gw_selfenergy_fft.argtypes = [P] + [vp] * 10
    L.gw_w_full.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int] + [vp] * 9
    L.gw_pairs_polarization.argtypes = [ctypes.c_int, ctypes.c_int]
    L.gw_pairs_polarization.restype = ctypes.c_longlong
    L.gw_pairs_selfenergy.argtypes = [ctypes.c_int, ctypes.c_int]
    L.gw_pairs_selfenergy.restype = ctypes.c_longlong
    L.gw_last_error.restype = ctypes.c_char_p
    L.gw_release.argtypes = []
    L.gw_release.restype = None
    _lib = L
    return L


def params_reference(nm, nx, nen, ndiag, dE, spindeg=1.0, alpha_mix=0.5):
    """gw_params of green_rgf_solve_gw_1D (nopmax = nen/2 - 10, dE_P, dE_S)."""
    p = Params()
    lib().gw_params_reference(ctypes.byref(p), nm, nx, nen, ndiag, dE, spindeg, alpha_mix)
    return p


def params(nm, nx, nen, ndiag, nopmax, dE_P, dE_S, alpha_mix):
    p = Params()
    p.nm, p.nx, p.nen, p.ndiag, p.nopmax = nm, nx, nen, ndiag, nopmax
    p.dE_P[0], p.dE_P[1] = complex(dE_P).real, complex(dE_P).imag
    p.dE_S[0], p.dE_S[1] = complex(dE_S).real, complex(dE_S).imag
    p.alpha_mix = alpha_mix
    return p


def _check(rc):
    if rc != 0:
        raise GwError(f"gw error {rc}: {lib().gw_last_error().decode()}")


def _arr(t, p, name):
    import torch
    if t is None or t.dtype != torch.complex128 or not t.is_cuda or not t.is_contiguous():
        raise GwError(f"{name}: need a contiguous complex128 CUDA tensor")
    if tuple(t.shape) != (p.nen, p.nx, p.nm, p.nm):
        raise GwError(f"{name}: shape {tuple(t.shape)} != (nen, nx, nm, nm) = {(p.nen, p.nx, p.nm, p.nm)}")
    return ctypes.c_void_p(t.data_ptr())


def _stream():
    import torch
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def polarization(p, Gl, Gg, Gr, Pl, Pg, Pr, method="auto"):
    """P^<, P^>, P^r into Pl, Pg, Pr (src/green_rgf.f90:3186-3221).  method "auto" (the
    direct form up to nen = 3072, else the FFT form) or "fft" (the reference's
    corr1d2(..., method='fft'), src/fft_mod.f90:56-90)."""
    a = [_arr(t, p, n) for t, n in ((Gl, "Gl"), (Gg, "Gg"), (Gr, "Gr"), (Pl, "Pl"), (Pg, "Pg"), (Pr, "Pr"))]
    fn = {"auto": lib().gw_polarization, "fft": lib().gw_polarization_fft}[method]
    _check(fn(ctypes.byref(p), *a, _stream()))


def selfenergy(p, Gl, Gg, Gr, Wl, Wg, Wr, Sl, Sg, Sr, method="auto"):
    """Sigma^<, Sigma^>, Sigma^r updated in place (src/green_rgf.f90:3233-3290).  method
    "auto" (the direct form up to nen = 2040, else the FFT form) or "fft" (conv1d2 with
    method='fft', src/fft_mod.f90:182-212)."""
    a = [_arr(t, p, n) for t, n in ((Gl, "Gl"), (Gg, "Gg"), (Gr, "Gr"), (Wl, "Wl"), (Wg, "Wg"), (Wr, "Wr"),
                                    (Sl, "Sl"), (Sg, "Sg"), (Sr, "Sr"))]
    fn = {"auto": lib().gw_selfenergy, "fft": lib().gw_selfenergy_fft}[method]
    _check(fn(ctypes.byref(p), *a, _stream()))


def w_full(Vii, V1i, PL, PG, PR, WL=None, WG=None, WR=None):
    """Dense screened interaction of the one-dimensional driver, green_calc_w_full
    (src/green_rgf.f90:3347-3396), for every energy of the (nen, nx, nm, nm) device
    tensors PL, PG, PR; Vii, V1i are (nx, nm, nm) device tensors (the memory of Fortran
    (nm, nm, nx) arrays).  Returns (WL, WG, WR) diagonal blocks, shaped like P."""
    import torch
    nen, nx, nm = int(PR.shape[0]), int(PR.shape[1]), int(PR.shape[2])
    for t, name in ((PL, "PL"), (PG, "PG"), (PR, "PR")):
        if t.dtype != torch.complex128 or not t.is_cuda or not t.is_contiguous() or tuple(t.shape) != (nen, nx, nm, nm):
            raise GwError(f"{name}: need a contiguous complex128 CUDA tensor of shape {(nen, nx, nm, nm)}")
    for t, name in ((Vii, "Vii"), (V1i, "V1i")):
        if t.dtype != torch.complex128 or not t.is_cuda or not t.is_contiguous() or tuple(t.shape) != (nx, nm, nm):
            raise GwError(f"{name}: need a contiguous complex128 CUDA tensor of shape {(nx, nm, nm)}")
    WL = torch.empty_like(PR) if WL is None else WL
    WG = torch.empty_like(PR) if WG is None else WG
    WR = torch.empty_like(PR) if WR is None else WR
    ptr = lambda t: ctypes.c_void_p(t.data_ptr())
    _check(lib().gw_w_full(nm, nx, nen, ptr(Vii), ptr(V1i), ptr(PL), ptr(PG), ptr(PR), ptr(WL), ptr(WG), ptr(WR),
                           _stream()))
    return WL, WG, WR


def to_device(a, device="cuda:0"):
    """numpy (nm, nm, nx, nen) array -> device tensor (nen, nx, nm, nm) of the same memory order."""
    import torch
    return torch.from_numpy(np.ascontiguousarray(np.asarray(a, np.complex128).T)).to(device)


def to_host(t):
    """Inverse of to_device: a numpy (nm, nm, nx, nen) array."""
    return t.cpu().numpy().T
