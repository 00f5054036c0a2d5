"""ctypes wrapper of the CPU oracle (liboracle_rgf.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, always as the checker or the timed CPU baseline,
never as a product path.

Arrays follow the reference layout: complex128 Fortran-ordered (nm, nm, nx)
blocks (green_rgf.f90:3950-3958), optionally with a trailing energy axis.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle_rgf.so")

COMPAT = 0
PHYSICAL = 1


class OrcParams(ctypes.Structure):
    _fields_ = [("eta_dev", ctypes.c_double), ("eta_lead", ctypes.c_double),
                ("dec_tol", ctypes.c_double), ("dec_nmax", ctypes.c_int),
                ("boltz", ctypes.c_double), ("mode", ctypes.c_int)]


class OrcDiag(ctypes.Structure):
    _fields_ = [("iters_left", ctypes.c_int), ("iters_right", ctypes.c_int),
                ("nmax_hit", ctypes.c_int), ("inv_failures", ctypes.c_int)]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            raise RuntimeError("oracle library missing: run `make -C oracle` (or __graft_entry__.build())")
        L = ctypes.CDLL(_LIB_PATH)
        dp = ctypes.POINTER(ctypes.c_double)
        L.orc_zinv.argtypes = [ctypes.c_int, dp]
        L.orc_zinv.restype = ctypes.c_int
        L.orc_zgemm.argtypes = [ctypes.c_char, ctypes.c_char, ctypes.c_int, dp, dp, dp]
        L.orc_sancho.argtypes = [ctypes.c_int, ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                 ctypes.c_int, dp, dp, dp, dp, ctypes.POINTER(ctypes.c_int),
                                 ctypes.POINTER(ctypes.c_int)]
        L.orc_sancho.restype = ctypes.c_int
        L.orc_ferm.argtypes = [ctypes.c_double]
        L.orc_ferm.restype = ctypes.c_double
        L.orc_params_reference.argtypes = [ctypes.POINTER(OrcParams)]
        L.orc_rgf.argtypes = [ctypes.POINTER(OrcParams), dp, ctypes.c_int, ctypes.c_int, ctypes.c_double,
                              dp, dp, dp, dp, dp, dp, dp, dp, dp, dp, dp, dp, dp, dp, dp,
                              ctypes.POINTER(OrcDiag)]
        L.orc_rgf_energies.argtypes = [ctypes.POINTER(OrcParams), dp, ctypes.c_int, ctypes.c_int,
                                       ctypes.c_int, dp, dp, dp, dp, dp, dp, dp, ctypes.c_longlong,
                                       dp, dp, dp, dp, dp, dp, dp, dp, dp,
                                       ctypes.POINTER(OrcDiag), ctypes.c_int]
        _lib = L
    return _lib


def _p(a):
    if a is None:
        return None
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def _cf(a):
    return np.asfortranarray(a, dtype=np.complex128)


def params(mode=COMPAT, eta_dev=None, eta_lead=None, dec_tol=None, dec_nmax=None, boltz=None):
    """Reference defaults (green_rgf.f90:3968, 3974, 4457, 4473, 4474) unless overridden."""
    p = OrcParams()
    lib().orc_params_reference(ctypes.byref(p))
    p.mode = mode
    if eta_dev is not None:
        p.eta_dev = eta_dev
    if eta_lead is not None:
        p.eta_lead = eta_lead
    if dec_tol is not None:
        p.dec_tol = dec_tol
    if dec_nmax is not None:
        p.dec_nmax = dec_nmax
    if boltz is not None:
        p.boltz = boltz
    return p


def zinv(A):
    A = _cf(A).copy(order="F")
    info = lib().orc_zinv(A.shape[0], _p(A))
    return A, info


def zgemm(ta, tb, A, B):
    A = _cf(A)
    B = _cf(B)
    C = np.zeros_like(A, order="F")
    lib().orc_zgemm(ta.encode(), tb.encode(), A.shape[0], _p(A), _p(B), _p(C))
    return C


def ferm(a):
    return lib().orc_ferm(a)


def sancho(E, H00, H10, eta_lead=1e-3, tol=1e-100, nmax=100):
    H00 = _cf(H00)
    H10 = _cf(H10)
    n = H00.shape[0]
    G00 = np.zeros((n, n), np.complex128, order="F")
    GBB = np.zeros((n, n), np.complex128, order="F")
    hit = ctypes.c_int(0)
    fails = ctypes.c_int(0)
    it = lib().orc_sancho(n, E, eta_lead, tol, nmax, _p(H00), _p(H10), _p(G00), _p(GBB),
                          ctypes.byref(hit), ctypes.byref(fails))
    return G00, GBB, it, hit.value, fails.value


def rgf(p, temp, E, mu, Hii, H1i, sigL=None, sigG=None, sigR=None, offdiag=True):
    """One energy, green_RGF_RS semantics.  Returns a dict of outputs."""
    Hii = _cf(Hii)
    H1i = _cf(H1i)
    nm, _, nx = Hii.shape
    sigL = None if sigL is None else _cf(sigL)
    sigG = None if sigG is None else _cf(sigG)
    sigR = None if sigR is None else _cf(sigR)
    shp = (nm, nm, nx)
    out = {k: np.zeros(shp, np.complex128, order="F") for k in ("ndens", "pdens", "ldos")}
    out["cur"] = np.zeros(shp, np.float64, order="F")
    if offdiag:
        for k in ("GRi1", "GLi1", "GGi1"):
            out[k] = np.zeros(shp, np.complex128, order="F")
    tr = ctypes.c_double(0.0)
    tre = ctypes.c_double(0.0)
    d = OrcDiag()
    t = np.asarray(temp, np.float64)
    m = np.asarray(mu, np.float64)
    dp = ctypes.POINTER(ctypes.c_double)
    lib().orc_rgf(ctypes.byref(p), _p(t), nm, nx, E, _p(m), _p(Hii), _p(H1i), _p(sigL), _p(sigG),
                  _p(sigR), _p(out["ndens"]), _p(out["pdens"]), _p(out["ldos"]),
                  ctypes.cast(ctypes.byref(tr), dp), ctypes.cast(ctypes.byref(tre), dp),
                  _p(out["cur"]), _p(out.get("GRi1")), _p(out.get("GLi1")), _p(out.get("GGi1")),
                  ctypes.byref(d))
    out["tr"] = tr.value
    out["tre"] = tre.value
    out["diag"] = (d.iters_left, d.iters_right, d.nmax_hit, d.inv_failures)
    return out


def rgf_energies(p, temp, E, mu, Hii, H1i, sigL=None, sigG=None, sigR=None,
                 outputs=("ndens", "pdens", "ldos", "cur", "GRi1", "GLi1", "GGi1"), nthreads=0):
    """Energy loop (OpenMP over energies, one per thread).  sig* may be (nm,nm,nx)
    shared or (nm,nm,nx,nE) per energy.  Returns dict with trailing energy axis."""
    Hii = _cf(Hii)
    H1i = _cf(H1i)
    nm, _, nx = Hii.shape
    E = np.ascontiguousarray(E, np.float64)
    nE = E.shape[0]
    stride = 0
    sigs = []
    for s in (sigL, sigG, sigR):
        if s is None:
            sigs.append(None)
            continue
        s = _cf(s)
        if s.ndim == 4:
            stride = nm * nm * nx
        sigs.append(s)
    shp = (nm, nm, nx, nE)
    out = {}
    for k in ("ndens", "pdens", "ldos", "GRi1", "GLi1", "GGi1"):
        out[k] = np.zeros(shp, np.complex128, order="F") if k in outputs else None
    out["cur"] = np.zeros(shp, np.float64, order="F") if "cur" in outputs else None
    tr = np.zeros(nE)
    tre = np.zeros(nE)
    diag = (OrcDiag * nE)()
    t = np.asarray(temp, np.float64)
    m = np.asarray(mu, np.float64)
    lib().orc_rgf_energies(ctypes.byref(p), _p(t), nm, nx, nE, _p(E), _p(m), _p(Hii), _p(H1i),
                           _p(sigs[0]), _p(sigs[1]), _p(sigs[2]), stride,
                           _p(out["ndens"]), _p(out["pdens"]), _p(out["ldos"]), _p(tr), _p(tre),
                           _p(out["cur"]), _p(out["GRi1"]), _p(out["GLi1"]), _p(out["GGi1"]),
                           diag, nthreads)
    out["tr"] = tr
    out["tre"] = tre
    out["diag"] = [(d.iters_left, d.iters_right, d.nmax_hit, d.inv_failures) for d in diag]
    return {k: v for k, v in out.items() if v is not None}


def rgf_fixed_leads(p, temp, E, mu, Hii, H1i, sigL=None, sigR=None, leadL_H00=None, leadR_H00=None):
    """orc_rgf_fixed_leads: the lead decimations use the given onsite blocks
    (default: the unperturbed edge blocks Hii[...,0] + sigR[...,0], Hii[...,-1] + sigR[...,-1])."""
    L = lib()
    if not hasattr(L, "_fl"):
        dp = ctypes.POINTER(ctypes.c_double)
        L.orc_rgf_fixed_leads.argtypes = [ctypes.POINTER(OrcParams), dp, ctypes.c_int, ctypes.c_int,
                                          ctypes.c_double, dp, dp, dp, dp, dp, dp, dp, dp, dp, dp, dp,
                                          ctypes.POINTER(OrcDiag)]
        L._fl = True
    Hii, H1i = _cf(Hii), _cf(H1i)
    nm, _, nx = Hii.shape
    sigL = None if sigL is None else _cf(sigL)
    sigR = None if sigR is None else _cf(sigR)
    if leadL_H00 is None:
        leadL_H00 = Hii[:, :, 0] + (sigR[:, :, 0] if sigR is not None else 0)
    if leadR_H00 is None:
        leadR_H00 = Hii[:, :, nx - 1] + (sigR[:, :, nx - 1] if sigR is not None else 0)
    leadL_H00, leadR_H00 = _cf(leadL_H00), _cf(leadR_H00)
    ndens = np.zeros((nm, nm, nx), np.complex128, order="F")
    ldos = np.zeros((nm, nm, nx), np.complex128, order="F")
    tr, tre = ctypes.c_double(0.0), ctypes.c_double(0.0)
    t = np.asarray(temp, np.float64)
    m = np.asarray(mu, np.float64)
    dp = ctypes.POINTER(ctypes.c_double)
    L.orc_rgf_fixed_leads(ctypes.byref(p), _p(t), nm, nx, float(E), _p(m), _p(Hii), _p(H1i), _p(sigL), _p(sigR),
                          _p(leadL_H00), _p(leadR_H00), _p(ndens), _p(ldos), ctypes.cast(ctypes.byref(tr), dp),
                          ctypes.cast(ctypes.byref(tre), dp), None)
    return {"ndens": ndens, "ldos": ldos, "tr": tr.value, "tre": tre.value}


def rgf_offdiag(p, temp, E, mu, Hii, H1i, sigL=None, sigR=None, sigR_lo=None, sigR_up=None, sigL_up=None):
    """orc_rgf_offdiag: physical-mode RGF with block-tridiagonal Sigma (one energy).
    sigR_lo[..., l] = Sigma^R_{l,l-1}, sigR_up[..., l] = Sigma^R_{l-1,l},
    sigL_up[..., l] = Sigma^<_{l-1,l} (l >= 1).  Returns the physical-mode outputs
    of rgf() plus GRi1_lo[..., l] = G^R_{l+1,l}."""
    L = lib()
    if not hasattr(L, "_od"):
        dp = ctypes.POINTER(ctypes.c_double)
        L.orc_rgf_offdiag.argtypes = ([ctypes.POINTER(OrcParams), dp, ctypes.c_int, ctypes.c_int, ctypes.c_double, dp]
                                      + [dp] * 17 + [ctypes.POINTER(OrcDiag)])
        L._od = True
    Hii, H1i = _cf(Hii), _cf(H1i)
    nm, _, nx = Hii.shape
    sigs = [None if s is None else _cf(s) for s in (sigL, sigR, sigR_lo, sigR_up, sigL_up)]
    shp = (nm, nm, nx)
    out = {k: np.zeros(shp, np.complex128, order="F") for k in ("ndens", "pdens", "ldos", "GRi1", "GLi1", "GGi1",
                                                               "GRi1_lo")}
    out["cur"] = np.zeros(shp, np.float64, order="F")
    tr, tre = ctypes.c_double(0.0), ctypes.c_double(0.0)
    d = OrcDiag()
    dp = ctypes.POINTER(ctypes.c_double)
    t = np.asarray(temp, np.float64)
    m = np.asarray(mu, np.float64)
    L.orc_rgf_offdiag(ctypes.byref(p), _p(t), nm, nx, float(E), _p(m), _p(Hii), _p(H1i), *[_p(s) for s in sigs],
                      _p(out["ndens"]), _p(out["pdens"]), _p(out["ldos"]), ctypes.cast(ctypes.byref(tr), dp),
                      ctypes.cast(ctypes.byref(tre), dp), _p(out["cur"]), _p(out["GRi1"]), _p(out["GLi1"]),
                      _p(out["GGi1"]), _p(out["GRi1_lo"]), ctypes.byref(d))
    out["tr"], out["tre"] = tr.value, tre.value
    out["diag"] = (d.iters_left, d.iters_right, d.nmax_hit, d.inv_failures)
    return out


def sancho_s(E, S00, H00, H10, eta_lead=1e-3, tol=1e-100, nmax=100):
    """orc_sancho_s: the decimation with the overlap S00 of the lead cells."""
    L = lib()
    if not hasattr(L, "_ss"):
        dp = ctypes.POINTER(ctypes.c_double)
        L.orc_sancho_s.argtypes = [ctypes.c_int, ctypes.c_double, ctypes.c_double, ctypes.c_double, ctypes.c_int,
                                   dp, dp, dp, dp, dp, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int)]
        L.orc_sancho_s.restype = ctypes.c_int
        L._ss = True
    S00, H00, H10 = _cf(S00), _cf(H00), _cf(H10)
    n = H00.shape[0]
    G00 = np.zeros((n, n), np.complex128, order="F")
    GBB = np.zeros((n, n), np.complex128, order="F")
    hit, fails = ctypes.c_int(0), ctypes.c_int(0)
    it = L.orc_sancho_s(n, E, eta_lead, tol, nmax, _p(S00), _p(H00), _p(H10), _p(G00), _p(GBB),
                        ctypes.byref(hit), ctypes.byref(fails))
    return G00, GBB, it, hit.value, fails.value


def rgf_cms(p, temp, E, mu, Sii, Hii, H1i, sigL=None, sigG=None, sigR=None):
    """orc_rgf_cms: green_RGF_CMS semantics (one energy, overlap Sii)."""
    L = lib()
    if not hasattr(L, "_cms"):
        dp = ctypes.POINTER(ctypes.c_double)
        L.orc_rgf_cms.argtypes = ([ctypes.POINTER(OrcParams), dp, ctypes.c_int, ctypes.c_int, ctypes.c_double, dp]
                                  + [dp] * 15 + [ctypes.POINTER(OrcDiag)])
        L._cms = True
    Sii, Hii, H1i = _cf(Sii), _cf(Hii), _cf(H1i)
    nm, _, nx = Hii.shape
    sigs = [None if s is None else _cf(s) for s in (sigL, sigG, sigR)]
    shp = (nm, nm, nx)
    out = {k: np.zeros(shp, np.complex128, order="F") for k in ("ndens", "pdens", "ldos", "GRi1", "GLi1", "GGi1")}
    out["cur"] = np.zeros(shp, np.float64, order="F")
    tr, tre = ctypes.c_double(0.0), ctypes.c_double(0.0)
    d = OrcDiag()
    dp = ctypes.POINTER(ctypes.c_double)
    t, m = np.asarray(temp, np.float64), np.asarray(mu, np.float64)
    L.orc_rgf_cms(ctypes.byref(p), _p(t), nm, nx, float(E), _p(m), _p(Sii), _p(Hii), _p(H1i), *[_p(s) for s in sigs],
                  _p(out["ndens"]), _p(out["pdens"]), _p(out["ldos"]), ctypes.cast(ctypes.byref(tr), dp),
                  ctypes.cast(ctypes.byref(tre), dp), _p(out["cur"]), _p(out["GRi1"]), _p(out["GLi1"]),
                  _p(out["GGi1"]), ctypes.byref(d))
    out["tr"], out["tre"] = tr.value, tre.value
    out["diag"] = (d.iters_left, d.iters_right, d.nmax_hit, d.inv_failures)
    return out


def rgf_offdiag_fixed_leads(p, temp, E, mu, Hii, H1i, sigL=None, sigR=None, sigR_lo=None, sigR_up=None,
                            sigL_up=None, leadL_H00=None, leadR_H00=None):
    """orc_rgf_offdiag_fixed_leads: ndens and tr with the lead onsite blocks given
    (default: the unperturbed edge blocks Hii + sigR)."""
    L = lib()
    if not hasattr(L, "_odfl"):
        dp = ctypes.POINTER(ctypes.c_double)
        L.orc_rgf_offdiag_fixed_leads.argtypes = ([ctypes.POINTER(OrcParams), dp, ctypes.c_int, ctypes.c_int,
                                                   ctypes.c_double, dp] + [dp] * 11)
        L._odfl = True
    Hii, H1i = _cf(Hii), _cf(H1i)
    nm, _, nx = Hii.shape
    sigs = [None if s is None else _cf(s) for s in (sigL, sigR, sigR_lo, sigR_up, sigL_up)]
    sR = sigs[1]
    if leadL_H00 is None:
        leadL_H00 = Hii[:, :, 0] + (sR[:, :, 0] if sR is not None else 0)
    if leadR_H00 is None:
        leadR_H00 = Hii[:, :, nx - 1] + (sR[:, :, nx - 1] if sR is not None else 0)
    leadL_H00, leadR_H00 = _cf(leadL_H00), _cf(leadR_H00)
    ndens = np.zeros((nm, nm, nx), np.complex128, order="F")
    tr = ctypes.c_double(0.0)
    t, m = np.asarray(temp, np.float64), np.asarray(mu, np.float64)
    dp = ctypes.POINTER(ctypes.c_double)
    L.orc_rgf_offdiag_fixed_leads(ctypes.byref(p), _p(t), nm, nx, float(E), _p(m), _p(Hii), _p(H1i),
                                  *[_p(s) for s in sigs], _p(leadL_H00), _p(leadR_H00), _p(ndens),
                                  ctypes.cast(ctypes.byref(tr), dp))
    return {"ndens": ndens, "tr": tr.value}
