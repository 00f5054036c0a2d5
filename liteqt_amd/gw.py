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


def w_blocks(Vii, V1i, PL, PG, PR, WL=None, WG=None, WR=None, chunk=256):
    """The same diagonal blocks as w_full, W^r = (I - V P^r)^-1 V and W^{<,>} = W^r P^{<,>}
    W^r^H, at O(nx nm^3) per energy instead of O((nx nm)^3): M = I - V P^r is block-
    tridiagonal and S = V P^< V^H block-pentadiagonal, so on super-blocks of two slabs
    (2 nm orbitals, max(2, ceil(nx/2)) of them, padded with decoupled zero slabs) both are
    block-tridiagonal and the engine's block-tridiagonal-Sigma solve gives X = M^-1 on its
    diagonal and first off-diagonal super-blocks and X S X^H on its diagonal ones: it runs
    with H = 0, z = 0, Sigma^R = -M (diagonal and off-diagonal super-blocks, so A = M and
    the zero edge couplings leave the contacts out) and Sigma^< = S.  Then W^r_ii =
    X_{i,i-1} V_{i-1,i} + X_ii V_ii + X_{i,i+1} V_{i+1,i}, W^<_ii = (X S X^H)_ii, and W^>
    by a second solve with S = V P^> V^H.  The solve takes Sigma^<_{l,l-1} =
    -(Sigma^<_{l-1,l})^H, so P^< and P^> must be anti-Hermitian, as the polarization step
    makes them (then S is too); P^r and V are general.  Needs 2 nm a multiple of 32,
    at most 256.  Tensors as in w_full; `chunk` energies per solve."""
    import torch
    from . import rgf
    nen, nx, nm = int(PR.shape[0]), int(PR.shape[1]), int(PR.shape[2])
    for t, name in ((PL, "PL"), (PG, "PG"), (PR, "PR")):
        if t.dtype != torch.complex128 or not t.is_cuda or not t.is_contiguous() or tuple(t.shape) != (nen, nx, nm, nm):
            raise GwError(f"{name}: need a contiguous complex128 CUDA tensor of shape {(nen, nx, nm, nm)}")
    for t, name in ((Vii, "Vii"), (V1i, "V1i")):
        if t.dtype != torch.complex128 or not t.is_cuda or not t.is_contiguous() or tuple(t.shape) != (nx, nm, nm):
            raise GwError(f"{name}: need a contiguous complex128 CUDA tensor of shape {(nx, nm, nm)}")
    n2 = 2 * nm
    if n2 % 32 or n2 > 256 or nx < 1:
        raise GwError("w_blocks: 2 nm must be a multiple of 32 and at most 256 (use w_full)")
    WL = torch.empty_like(PR) if WL is None else WL
    WG = torch.empty_like(PR) if WG is None else WG
    WR = torch.empty_like(PR) if WR is None else WR
    dev = PR.device
    tm = lambda t: t.transpose(-1, -2)  # the device blocks hold A^T: these are the matrices
    H_ = lambda t: t.conj().transpose(-1, -2)
    nxp = max(2, (nx + 1) // 2)  # super-blocks (the solve needs two)
    nx2 = 2 * nxp
    Vd = torch.zeros((nx2, nm, nm), dtype=torch.complex128, device=dev)
    V1 = torch.zeros_like(Vd)  # V1[i] = V_{i+1,i}
    Vd[:nx] = tm(Vii)
    V1[:nx - 1] = tm(V1i)[:nx - 1]
    eye = torch.eye(nm, dtype=torch.complex128, device=dev)
    p = rgf.params(rgf.MODE_PHYSICAL, eta_dev=0.0, eta_lead=1e-3)
    zsup = torch.zeros((nxp, n2, n2), dtype=torch.complex128, device=dev)  # H = 0

    def pad(P):  # (ce, nx, nm, nm) device blocks -> (ce, nx2, nm, nm) matrices
        out = torch.zeros((P.shape[0], nx2, nm, nm), dtype=torch.complex128, device=dev)
        out[:, :nx] = tm(P)
        return out

    def lesser_source(P):  # super-blocks of S = V P V^H: diagonal and (s-1, s)
        ce = P.shape[0]
        Vu = H_(V1)  # Vu[i] = V_{i,i+1}
        Sii = Vd @ P @ H_(Vd)
        Sii[:, 1:] += V1[:-1] @ P[:, :-1] @ H_(V1[:-1])
        Sii[:, :-1] += Vu[:-1] @ P[:, 1:] @ V1[:-1]
        Sup1 = Vd[:-1] @ P[:, :-1] @ H_(V1[:-1]) + Vu[:-1] @ P[:, 1:] @ H_(Vd[1:])  # S_{i,i+1}
        Slo1 = V1[:-1] @ P[:, :-1] @ H_(Vd[:-1]) + Vd[1:] @ P[:, 1:] @ V1[:-1]     # S_{i+1,i}
        Sup2 = Vu[:-2] @ P[:, 1:-1] @ H_(V1[1:-1]) if nx2 > 2 else None           # S_{i,i+2}
        Sd = torch.zeros((ce, nxp, n2, n2), dtype=torch.complex128, device=dev)
        Sd[..., :nm, :nm] = Sii[:, 0::2]
        Sd[..., nm:, nm:] = Sii[:, 1::2]
        Sd[..., :nm, nm:] = Sup1[:, 0::2]
        Sd[..., nm:, :nm] = Slo1[:, 0::2]
        Su = torch.zeros_like(Sd)
        if nxp > 1:
            Su[:, 1:, :nm, :nm] = Sup2[:, 0::2]   # S_{2s-2,2s}
            Su[:, 1:, nm:, :nm] = Sup1[:, 1::2]   # S_{2s-1,2s}
            Su[:, 1:, nm:, nm:] = Sup2[:, 1::2]   # S_{2s-1,2s+1}
        return tm(Sd).contiguous(), tm(Su).contiguous()

    for e0 in range(0, nen, chunk):
        e1 = min(nen, e0 + chunk)
        ce = e1 - e0
        Pr = pad(PR[e0:e1])
        Mii = eye - Vd @ Pr
        Mlo = -(V1[:-1] @ Pr[:, :-1])       # M_{i+1,i}
        Mup = -(H_(V1[:-1]) @ Pr[:, 1:])    # M_{i,i+1}
        A = torch.zeros((ce, nxp, n2, n2), dtype=torch.complex128, device=dev)
        A[..., :nm, :nm] = Mii[:, 0::2]
        A[..., nm:, nm:] = Mii[:, 1::2]
        A[..., :nm, nm:] = Mup[:, 0::2]
        A[..., nm:, :nm] = Mlo[:, 0::2]
        Lo = torch.zeros_like(A)
        Up = torch.zeros_like(A)
        if nxp > 1:
            Lo[:, 1:, :nm, nm:] = Mlo[:, 1::2]  # M_{2s,2s-1}
            Up[:, 1:, nm:, :nm] = Mup[:, 1::2]  # M_{2s-1,2s}
        sig = {"sigR": tm(-A).contiguous(), "sigR_lo": tm(-Lo).contiguous(), "sigR_up": tm(-Up).contiguous()}
        del Pr, Mii, Mlo, Mup, A, Lo, Up
        E = torch.zeros(ce, dtype=torch.float64, device=dev)
        shape = (ce, nxp, n2, n2)
        X = torch.empty(shape, dtype=torch.complex128, device=dev)
        Xu = torch.zeros(shape, dtype=torch.complex128, device=dev)
        Xl = torch.zeros(shape, dtype=torch.complex128, device=dev)
        for P, Wout, first in ((PL, WL, True), (PG, WG, False)):
            Sd, Su = lesser_source(pad(P[e0:e1]))
            Gn = torch.empty(shape, dtype=torch.complex128, device=dev)
            out = {"ndens": Gn}
            if first:
                out.update(ldos=X, GRi1=Xu, GRi1_lo=Xl)
            rgf.solve_batched(p, E, (300.0, 300.0), (0.0, 0.0), zsup, zsup, sigL=Sd, sigR=sig["sigR"],
                              sigR_lo=sig["sigR_lo"], sigR_up=sig["sigR_up"], sigL_up=Su, out=out)
            Gt = tm(Gn)
            Wd = torch.stack([Gt[..., :nm, :nm], Gt[..., nm:, nm:]], dim=2).reshape(ce, nx2, nm, nm)
            Wout[e0:e1] = tm(Wd[:, :nx])
            del Sd, Su, Gn, Gt, Wd
        Xt, Xut, Xlt = tm(X), tm(Xu), tm(Xl)
        Wr = torch.zeros((ce, nx2, nm, nm), dtype=torch.complex128, device=dev)
        Vu = H_(V1)
        # even slabs 2s: X_{2s,2s-1} V_{2s-1,2s} + X_{2s,2s} V_{2s,2s} + X_{2s,2s+1} V_{2s+1,2s}
        Wr[:, 0::2] = Xt[..., :nm, :nm] @ Vd[0::2] + Xt[..., :nm, nm:] @ V1[0::2]
        if nxp > 1:
            Wr[:, 2::2] += Xlt[:, :-1, :nm, nm:] @ Vu[1:-1:2]
        # odd slabs 2s+1: X_{2s+1,2s} V_{2s,2s+1} + X_{2s+1,2s+1} V_{2s+1,2s+1} + X_{2s+1,2s+2} V_{2s+2,2s+1}
        Wr[:, 1::2] = Xt[..., nm:, :nm] @ Vu[0::2] + Xt[..., nm:, nm:] @ Vd[1::2]
        if nxp > 1:
            Wr[:, 1:-1:2] += Xut[:, :-1, nm:, :nm] @ V1[1:-1:2]
        WR[e0:e1] = tm(Wr[:, :nx])
        del X, Xu, Xl, Xt, Xut, Xlt, Wr, sig
    return WL, WG, WR


def to_device(a, device="cuda:0"):
    """numpy (nm, nm, nx, nen) array -> device tensor (nen, nx, nm, nm) of the same memory order."""
    import torch
    return torch.from_numpy(np.ascontiguousarray(np.asarray(a, np.complex128).T)).to(device)


def to_host(t):
    """Inverse of to_device: a numpy (nm, nm, nx, nen) array."""
    return t.cpu().numpy().T
