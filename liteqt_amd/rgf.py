"""Python binding of the MI355X RGF engine (librgf_hip.so, C ABI of include/rgf.h).

Mirrors the reference's operator for this path, green_RGF_RS
(src/green_rgf.f90:3950-3958): same argument order and meaning, complex128
arrays in Fortran layout (nm, nm, nx), optional off-diagonal outputs.  Three
levels:
  * rgf_rs(...)            one energy, host (numpy) arrays, drop-in semantics;
  * rgf_rs_energies(...)   the drivers' energy loop (green_rgf.f90:3174-3180)
                           as one batched GPU solve, host arrays;
  * solve_batched(...)     device-resident torch tensors, asynchronous on the
                           current HIP stream (the benchmark / SCBA path).
There is no CPU fallback: every call runs the HIP kernels and raises when the
library or a GPU is missing.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "librgf_hip.so")
PLAIN_LIB_PATH = os.path.join(_HERE, "librgf_plain.so")

MODE_COMPAT = 0
MODE_PHYSICAL = 1

ERRORS = {1: "EINVAL", 2: "EHIP", 3: "ENOMEM", 4: "EUNSUP"}


class RgfError(RuntimeError):
    pass


class Params(ctypes.Structure):
    _fields_ = [("eta_dev", ctypes.c_double), ("eta_lead", ctypes.c_double),
                ("dec_tol", ctypes.c_double), ("dec_nmax", ctypes.c_int),
                ("boltz", ctypes.c_double), ("mode", ctypes.c_int)]


class Batch(ctypes.Structure):
    _fields_ = [("nm", ctypes.c_int), ("nx", ctypes.c_int), ("nE", ctypes.c_int),
                ("max_slots", ctypes.c_int),
                ("E", ctypes.c_void_p), ("temp", ctypes.c_double * 2), ("mu", ctypes.c_double * 2),
                ("Hii", ctypes.c_void_p), ("H1i", ctypes.c_void_p),
                ("sigL", ctypes.c_void_p), ("sigG", ctypes.c_void_p), ("sigR", ctypes.c_void_p),
                ("sig_stride", ctypes.c_longlong),
                ("ldos", ctypes.c_void_p), ("ndens", ctypes.c_void_p), ("pdens", ctypes.c_void_p),
                ("GRi1", ctypes.c_void_p), ("GLi1", ctypes.c_void_p), ("GGi1", ctypes.c_void_p),
                ("cur", ctypes.c_void_p), ("tr", ctypes.c_void_p), ("tre", ctypes.c_void_p),
                ("dens_n", ctypes.c_void_p), ("dens_p", ctypes.c_void_p), ("info", ctypes.c_void_p),
                ("sigR_lo", ctypes.c_void_p), ("sigR_up", ctypes.c_void_p), ("sigL_up", ctypes.c_void_p),
                ("Sii", ctypes.c_void_p), ("ckpt", ctypes.c_int), ("GRi1_lo", ctypes.c_void_p)]


class MultiReport(ctypes.Structure):
    _fields_ = [("slots", ctypes.c_int), ("tuned", ctypes.c_int), ("items", ctypes.c_longlong),
                ("ws_bytes", ctypes.c_longlong)]


class Scba(ctypes.Structure):
    _fields_ = [("D", ctypes.c_double), ("mix", ctypes.c_double)]


class WBatch(ctypes.Structure):
    _fields_ = [("nm", ctypes.c_int), ("nx", ctypes.c_int), ("nE", ctypes.c_int), ("nbc", ctypes.c_int),
                ("Vii", ctypes.c_void_p), ("V1i", ctypes.c_void_p), ("PL", ctypes.c_void_p),
                ("PG", ctypes.c_void_p), ("PR", ctypes.c_void_p), ("WL", ctypes.c_void_p),
                ("WG", ctypes.c_void_p), ("WR", ctypes.c_void_p), ("WRi1", ctypes.c_void_p),
                ("info", ctypes.c_void_p)]


# every function declared in include/rgf.h
EXPORTS = ("rgf_params_reference", "rgf_last_error", "rgf_version", "rgf_rs", "rgf_rs_ex",
           "rgf_rs_energies", "rgf_solve_batched", "rgf_workspace_bytes", "rgf_release",
           "rgf_zgemm_batched", "rgf_zinv_batched", "rgf_adjoint_batched", "rgf_adjoint_workspace_bytes",
           "rgf_scba_step", "rgf_graph_create", "rgf_graph_launch", "rgf_graph_destroy", "rgf_solve_devices",
           "rgf_graph_create_devices", "rgf_cms", "rgf_cms_ex", "rgf_cms_energies", "rgf_solve_batched_noalloc",
           "rgf_integrate", "rgf_w_solve_batched")

_lib = None


def lib():
    """Load librgf_hip.so (built by __graft_entry__.build()); raise if it is missing."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RgfError(f"HIP engine library missing: {LIB_PATH} (run __graft_entry__.build())")
    L = ctypes.CDLL(LIB_PATH)
    vp = ctypes.c_void_p
    L.rgf_params_reference.argtypes = [ctypes.POINTER(Params)]
    L.rgf_last_error.restype = ctypes.c_char_p
    L.rgf_version.restype = ctypes.c_char_p
    L.rgf_rs.argtypes = [vp, ctypes.c_int, ctypes.c_int, ctypes.c_double, vp] + [vp] * 14
    L.rgf_rs_ex.argtypes = [ctypes.POINTER(Params), vp, ctypes.c_int, ctypes.c_int, ctypes.c_double, vp] + [vp] * 14
    L.rgf_rs_energies.argtypes = ([ctypes.POINTER(Params), vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, vp, vp,
                                   vp, vp, vp, vp, vp, ctypes.c_int] + [vp] * 10)
    L.rgf_solve_batched.argtypes = [ctypes.POINTER(Params), ctypes.POINTER(Batch), vp]
    L.rgf_solve_batched_noalloc.argtypes = [ctypes.POINTER(Params), ctypes.POINTER(Batch), vp]
    L.rgf_integrate.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, vp, vp, vp, vp, vp, ctypes.c_double,
                                vp, vp, vp, vp]
    L.rgf_cms_ex.argtypes = [ctypes.POINTER(Params), vp, ctypes.c_int, ctypes.c_int, ctypes.c_double, vp] + [vp] * 15
    L.rgf_cms_energies.argtypes = ([ctypes.POINTER(Params), vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, vp, vp,
                                    vp, vp, vp, vp, vp, vp, ctypes.c_int] + [vp] * 10)
    L.rgf_scba_step.argtypes = [ctypes.POINTER(Params), ctypes.POINTER(Batch), ctypes.POINTER(Scba), vp]
    L.rgf_graph_create.argtypes = [ctypes.POINTER(Params), ctypes.POINTER(Batch), ctypes.POINTER(Scba), vp,
                                   ctypes.POINTER(vp)]
    L.rgf_graph_launch.argtypes = [vp, ctypes.c_int, vp]
    L.rgf_solve_devices.argtypes = [ctypes.POINTER(Params), ctypes.POINTER(Batch), ctypes.c_int, ctypes.c_int, vp,
                                    ctypes.POINTER(MultiReport), vp]
    L.rgf_graph_create_devices.argtypes = [ctypes.POINTER(Params), ctypes.POINTER(Batch), ctypes.c_int,
                                           ctypes.c_int, vp, ctypes.POINTER(vp)]
    L.rgf_graph_destroy.argtypes = [vp]
    L.rgf_graph_destroy.restype = None
    L.rgf_zgemm_batched.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, vp, ctypes.c_int, vp, vp,
                                    ctypes.c_int, vp]
    L.rgf_zinv_batched.argtypes = [ctypes.c_int, ctypes.c_int, vp, vp, vp, vp]
    L.rgf_adjoint_batched.argtypes = [ctypes.POINTER(Params), ctypes.POINTER(Batch), vp, vp, vp, vp, vp, vp, vp]
    L.rgf_adjoint_workspace_bytes.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int]
    L.rgf_adjoint_workspace_bytes.restype = ctypes.c_longlong
    L.rgf_w_solve_batched.argtypes = [ctypes.POINTER(WBatch), vp]
    L.rgf_workspace_bytes.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int]
    L.rgf_workspace_bytes.restype = ctypes.c_longlong
    _lib = L
    return L


def _check(rc):
    if rc != 0:
        msg = lib().rgf_last_error().decode(errors="replace")
        raise RgfError(f"rgf error {ERRORS.get(rc, rc)}: {msg}")


def params(mode=MODE_COMPAT, eta_dev=None, eta_lead=None, dec_tol=None, dec_nmax=None, boltz=None):
    """Reference parameters (green_rgf.f90:3968, 3974, 4457, 4473, 4474) unless overridden."""
    p = Params()
    lib().rgf_params_reference(ctypes.byref(p))
    p.mode = mode
    for k, v in (("eta_dev", eta_dev), ("eta_lead", eta_lead), ("dec_tol", dec_tol),
                 ("dec_nmax", dec_nmax), ("boltz", boltz)):
        if v is not None:
            setattr(p, k, v)
    return p


def _cf(a):
    return None if a is None else np.asfortranarray(a, dtype=np.complex128)


def _ptr(a):
    return None if a is None else a.ctypes.data


def rgf_rs(temp, E, mu, Hii, H1i, sigL=None, sigG=None, sigR=None, p=None, offdiag=True):
    """One energy with green_RGF_RS semantics on the GPU.  Returns a dict with
    ndens, pdens, ldos, cur, tr, tre (and GRi1, GLi1, GGi1 when offdiag)."""
    Hii, H1i = _cf(Hii), _cf(H1i)
    sigL, sigG, sigR = _cf(sigL), _cf(sigG), _cf(sigR)
    nm, _, nx = Hii.shape
    shp = (nm, nm, nx)
    out = {k: np.zeros(shp, np.complex128, order="F") for k in ("ndens", "pdens", "ldos")}
    out["cur"] = np.zeros(shp, np.float64, order="F")
    if offdiag:
        for k in ("GRi1", "GLi1", "GGi1"):
            out[k] = np.zeros(shp, np.complex128, order="F")
    tr = np.zeros(1)
    tre = np.zeros(1)
    t = np.ascontiguousarray(temp, np.float64)
    m = np.ascontiguousarray(mu, np.float64)
    if p is None:
        p = params()
    rc = lib().rgf_rs_ex(ctypes.byref(p), _ptr(t), nm, nx, float(E), _ptr(m), _ptr(Hii), _ptr(H1i), _ptr(sigL),
                         _ptr(sigG), _ptr(sigR), _ptr(out["ndens"]), _ptr(out["pdens"]), _ptr(out["ldos"]),
                         _ptr(tr), _ptr(tre), _ptr(out["cur"]), _ptr(out.get("GRi1")), _ptr(out.get("GLi1")),
                         _ptr(out.get("GGi1")))
    _check(rc)
    out["tr"] = float(tr[0])
    out["tre"] = float(tre[0])
    return out


def rgf_rs_energies(p, temp, E, mu, Hii, H1i, sigL=None, sigG=None, sigR=None,
                    outputs=("ndens", "pdens", "ldos", "cur", "GRi1", "GLi1", "GGi1")):
    """All energies of E in one batched GPU solve (host arrays).  sig* either
    (nm,nm,nx) shared or (nm,nm,nx,nE) per energy.  Returns a dict with a
    trailing energy axis, plus tr, tre and info (nE, 4)."""
    Hii, H1i = _cf(Hii), _cf(H1i)
    nm, _, nx = Hii.shape
    E = np.ascontiguousarray(E, np.float64)
    nE = E.shape[0]
    sigs = [_cf(s) for s in (sigL, sigG, sigR)]
    per_e = any(s is not None and s.ndim == 4 for s in sigs)
    if per_e:
        sigs = [None if s is None else (s if s.ndim == 4 else
                np.asfortranarray(np.repeat(s[..., None], nE, axis=3))) for s in sigs]
    shp = (nm, nm, nx, nE)
    out = {}
    for k in ("ndens", "pdens", "ldos", "GRi1", "GLi1", "GGi1"):
        out[k] = np.zeros(shp, np.complex128, order="F") if k in outputs else None
    out["cur"] = np.zeros(shp, np.float64, order="F") if "cur" in outputs else None
    tr = np.zeros(nE)
    tre = np.zeros(nE)
    info = np.zeros((nE, 4), np.int32)
    t = np.ascontiguousarray(temp, np.float64)
    m = np.ascontiguousarray(mu, np.float64)
    rc = lib().rgf_rs_energies(ctypes.byref(p), _ptr(t), nm, nx, nE, _ptr(E), _ptr(m), _ptr(Hii), _ptr(H1i),
                               _ptr(sigs[0]), _ptr(sigs[1]), _ptr(sigs[2]), int(per_e),
                               _ptr(out["ndens"]), _ptr(out["pdens"]), _ptr(out["ldos"]), _ptr(tr), _ptr(tre),
                               _ptr(out["cur"]), _ptr(out["GRi1"]), _ptr(out["GLi1"]), _ptr(out["GGi1"]),
                               _ptr(info))
    _check(rc)
    out["tr"] = tr
    out["tre"] = tre
    out["info"] = info
    return {k: v for k, v in out.items() if v is not None}


def rgf_cms(temp, E, mu, Sii, Hii, H1i, sigL=None, sigG=None, sigR=None, p=None):
    """One energy with green_RGF_CMS semantics (overlap Sii) on the GPU; the
    outputs of rgf_rs."""
    Sii, Hii, H1i = _cf(Sii), _cf(Hii), _cf(H1i)
    sigL, sigG, sigR = _cf(sigL), _cf(sigG), _cf(sigR)
    nm, _, nx = Hii.shape
    shp = (nm, nm, nx)
    out = {k: np.zeros(shp, np.complex128, order="F") for k in ("ndens", "pdens", "ldos", "GRi1", "GLi1", "GGi1")}
    out["cur"] = np.zeros(shp, np.float64, order="F")
    tr, tre = np.zeros(1), np.zeros(1)
    t, m = np.ascontiguousarray(temp, np.float64), np.ascontiguousarray(mu, np.float64)
    if p is None:
        p = params()
    _check(lib().rgf_cms_ex(ctypes.byref(p), _ptr(t), nm, nx, float(E), _ptr(m), _ptr(Sii), _ptr(Hii), _ptr(H1i),
                            _ptr(sigL), _ptr(sigG), _ptr(sigR), _ptr(out["ndens"]), _ptr(out["pdens"]),
                            _ptr(out["ldos"]), _ptr(tr), _ptr(tre), _ptr(out["cur"]), _ptr(out["GRi1"]),
                            _ptr(out["GLi1"]), _ptr(out["GGi1"])))
    out["tr"], out["tre"] = float(tr[0]), float(tre[0])
    return out


def _tp(t):
    return None if t is None else t.data_ptr()


def _req(t, name, dtype, numel, device, host_ok=False):
    """Validate a tensor handed to the C ABI as a raw pointer: right dtype,
    contiguous, on `device` (or, with host_ok, in pinned host memory, which the
    device addresses directly), with at least `numel` elements.  None passes."""
    if t is None:
        return None
    if t.dtype != dtype:
        raise RgfError(f"{name}: expected {dtype}, got {t.dtype}")
    if not t.is_contiguous():
        raise RgfError(f"{name}: tensor must be contiguous")
    if t.device != device and not (host_ok and t.device.type == "cpu" and (t.numel() == 0 or t.is_pinned())):
        raise RgfError(f"{name}: tensor on {t.device}, expected {device}")
    if t.numel() < numel:
        raise RgfError(f"{name}: {t.numel()} elements, need {numel}")
    return t.data_ptr()


_OUT_KINDS = {"ldos": "c", "ndens": "c", "pdens": "c", "GRi1": "c", "GLi1": "c", "GGi1": "c", "GRi1_lo": "c", "cur": "r",
              "tr": "e", "tre": "e", "dens_n": "x", "dens_p": "x", "info": "i"}


_plain = None


def plain_lib():
    """The plain per-energy comparison path (rocBLAS/rocSOLVER call chain)."""
    global _plain
    if _plain is None:
        if not os.path.exists(PLAIN_LIB_PATH):
            raise RgfError(f"plain-path library missing: {PLAIN_LIB_PATH} (run __graft_entry__.build())")
        L = ctypes.CDLL(PLAIN_LIB_PATH)
        L.rgf_plain_solve.argtypes = [ctypes.POINTER(Params), ctypes.POINTER(Batch), ctypes.c_void_p]
        L.rgf_plain_last_error.restype = ctypes.c_char_p
        _plain = L
    return _plain


def solve_batched(p, E, temp, mu, Hii, H1i, sigL=None, sigR=None, out=None, max_slots=0, stream=None,
                  plain=False, sigR_lo=None, sigR_up=None, sigL_up=None, scba=None, ckpt=0):
    """Device-resident batched solve (plain=True: the per-energy rocBLAS/rocSOLVER
    comparison path with the same arguments).  E, Hii, H1i, sig*, out[...] are torch
    tensors on the HIP device: complex blocks as complex128 tensors of shape
    (nE, nx, nm, nm) holding column-major blocks (i.e. the memory of a Fortran
    (nm,nm,nx,nE) array; use to_device_blocks()).  out maps names in
    {ldos, ndens, pdens, GRi1, GLi1, GGi1, cur, tr, tre, dens_n, dens_p, info}
    to preallocated tensors.  sigR_lo / sigR_up / sigL_up: first off-diagonal
    Sigma blocks (physical mode).  scba=(D, mix): one SCBA iteration that also
    refreshes the given Sigma tensors in place (rgf_scba_step).  ckpt: storage of
    the left-connected blocks (rgf_batch.ckpt: -1 every slab, K > 0 checkpoints every
    K slabs with segment recompute, 0 automatic).  Asynchronous on `stream`
    (default: torch's current stream)."""
    import torch
    out = out or {}
    b = _batch(torch, E, temp, mu, Hii, H1i, sigL, sigR, max_slots, sigR_lo, sigR_up, sigL_up)
    b.ckpt = int(ckpt)
    _out_batch(torch, b, out, E.device)
    if stream is None:
        stream = torch.cuda.current_stream()
    if plain:
        rc = plain_lib().rgf_plain_solve(ctypes.byref(p), ctypes.byref(b), ctypes.c_void_p(stream.cuda_stream))
        if rc != 0:
            raise RgfError(f"plain path error {rc}: {plain_lib().rgf_plain_last_error().decode()}")
        return out
    if scba is not None:
        sc = Scba(float(scba[0]), float(scba[1]))
        rc = lib().rgf_scba_step(ctypes.byref(p), ctypes.byref(b), ctypes.byref(sc), ctypes.c_void_p(stream.cuda_stream))
    else:
        rc = lib().rgf_solve_batched(ctypes.byref(p), ctypes.byref(b), ctypes.c_void_p(stream.cuda_stream))
    _check(rc)
    return out


def _out_batch(torch, b, out, dev):
    nE, nx, nm = b.nE, b.nx, b.nm
    blk = nE * nx * nm * nm
    sizes = {"c": (torch.complex128, blk), "r": (torch.float64, blk), "e": (torch.float64, nE),
             "x": (torch.float64, nE * nx), "i": (torch.int32, 4 * nE)}
    for k, kind in _OUT_KINDS.items():
        dt, cnt = sizes[kind]
        # info may sit in pinned host memory: the kernels write each energy's four
        # counters there as it finishes, so the host can watch a long launch without
        # queueing work on the device
        setattr(b, k, _req(out.get(k), k, dt, cnt, dev, host_ok=(kind == "i")))


class Graph:
    """A captured batched solve (rgf_graph_create) replayed with launch(n).
    With scba=(D, mix) each replay is one SCBA iteration that refreshes the
    Sigma tensors in place.  The tensors passed here must stay alive and in
    place for the graph's lifetime; `stream` must be a created stream
    (torch.cuda.Stream())."""

    def __init__(self, p, E, temp, mu, Hii, H1i, stream, sigL=None, sigR=None, out=None, max_slots=0,
                 sigR_lo=None, sigR_up=None, sigL_up=None, scba=None):
        import torch
        self._keep = (E, Hii, H1i, sigL, sigR, sigR_lo, sigR_up, sigL_up, out)
        b = _batch(torch, E, temp, mu, Hii, H1i, sigL, sigR, max_slots, sigR_lo, sigR_up, sigL_up)
        _out_batch(torch, b, out or {}, E.device)
        self._stream = stream
        h = ctypes.c_void_p()
        sc = None if scba is None else Scba(float(scba[0]), float(scba[1]))
        _check(lib().rgf_graph_create(ctypes.byref(p), ctypes.byref(b), None if sc is None else ctypes.byref(sc),
                                      ctypes.c_void_p(stream.cuda_stream), ctypes.byref(h)))
        self._h = h

    def launch(self, replays=1):
        _check(lib().rgf_graph_launch(self._h, int(replays), ctypes.c_void_p(self._stream.cuda_stream)))

    def close(self):
        if self._h:
            lib().rgf_graph_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _batch(torch, E, temp, mu, Hii, H1i, sigL, sigR, max_slots, sigR_lo=None, sigR_up=None, sigL_up=None):
    """rgf_batch with the inputs checked: E float64 (nE,), Hii/H1i complex128
    (nx, nm, nm) column-major blocks, sig* (nx, nm, nm) shared or (nE, nx, nm, nm)."""
    dev = E.device
    nE = int(E.shape[0])
    nx, nm = int(Hii.shape[-3]), int(Hii.shape[-1])
    b = Batch()
    b.nm, b.nx, b.nE, b.max_slots = nm, nx, nE, int(max_slots)
    b.E = _req(E, "E", torch.float64, nE, dev)
    b.temp[0], b.temp[1] = float(temp[0]), float(temp[1])
    b.mu[0], b.mu[1] = float(mu[0]), float(mu[1])
    blk = nx * nm * nm
    b.Hii = _req(Hii, "Hii", torch.complex128, blk, dev)
    b.H1i = _req(H1i, "H1i", torch.complex128, blk, dev)
    stride = 0
    sigs = {"sigL": sigL, "sigR": sigR, "sigR_lo": sigR_lo, "sigR_up": sigR_up, "sigL_up": sigL_up}
    for k, t in sigs.items():
        if t is not None and t.dim() == 4:
            stride = blk
    for k, t in sigs.items():
        if t is not None and stride and t.dim() != 4:
            raise RgfError(f"{k}: all Sigma tensors must be per energy (nE, nx, nm, nm) when one is")
        setattr(b, k, _req(t, k, torch.complex128, blk + stride * (nE - 1), dev))
    b.sig_stride = stride
    return b


def to_device_blocks(a, device="cuda"):
    """numpy Fortran (nm,nm,nx[,nE]) complex array -> torch complex128 tensor whose
    memory is the same Fortran layout: shape (nx, nm, nm) / (nE, nx, nm, nm) with
    the last two axes (col, row)."""
    import torch
    a = np.asfortranarray(a, dtype=np.complex128)
    t = np.ascontiguousarray(a.transpose(tuple(range(a.ndim - 1, -1, -1))))
    return torch.from_numpy(t).to(device)


def from_device_blocks(t):
    """Inverse of to_device_blocks."""
    a = t.detach().cpu().numpy()
    return np.asfortranarray(a.transpose(tuple(range(a.ndim - 1, -1, -1))))


def zgemm_batched(A, B, opA=0, opB=0, rowmajor=False, stream=None):
    """C[b] = op(A[b]) op(B[b]) for torch complex128 tensors (nb, n, n) holding
    column-major matrices (see to_device_blocks)."""
    import torch
    C = torch.empty_like(A)
    s = stream or torch.cuda.current_stream()
    _check(lib().rgf_zgemm_batched(int(A.shape[-1]), int(A.shape[0]), int(opA), A.data_ptr(), int(opB),
                                   B.data_ptr(), C.data_ptr(), int(bool(rowmajor)), ctypes.c_void_p(s.cuda_stream)))
    return C


def zinv_batched(A, stream=None):
    """(A[b]^-1, info) for torch complex128 tensors (nb, n, n) of column-major matrices."""
    import torch
    C = torch.empty_like(A)
    info = torch.zeros(A.shape[0], dtype=torch.int32, device=A.device)
    s = stream or torch.cuda.current_stream()
    _check(lib().rgf_zinv_batched(int(A.shape[-1]), int(A.shape[0]), A.data_ptr(), C.data_ptr(), info.data_ptr(),
                                  ctypes.c_void_p(s.cuda_stream)))
    return C, info


def adjoint_batched(p, E, temp, mu, Hii, H1i, sigL=None, sigR=None, w_tr=None, w_n=None, max_slots=0,
                    stream=None, want_dH=True, tr=None, info=None, lead_H00=None, sigR_lo=None, sigR_up=None,
                    sigL_up=None, ckpt=0):
    """Adjoint RGF (physical mode).  Tensors as in solve_batched; w_tr float64
    (nE,), w_n float64 (nE, nx).  Without lead_H00 the contacts repeat the edge
    slabs (the reference's device) and are differentiated with them; lead_H00
    complex128 (2, nm, nm), the left and right lead onsite blocks, holds them
    fixed.  ckpt: the forward state kept for every slab (-1),
    at checkpoints every ckpt >= 2 slabs with segments recomputed, or chosen by
    device memory (0).  Returns (J[nE], dU[nE, nx], dH1i[nx, nm, nm]
    column-major blocks or None)."""
    import torch
    b = _batch(torch, E, temp, mu, Hii, H1i, sigL, sigR, max_slots, sigR_lo, sigR_up, sigL_up)
    b.ckpt = int(ckpt)
    nE, nx, nm, dev = b.nE, b.nx, b.nm, E.device
    J = torch.zeros(nE, dtype=torch.float64, device=dev)
    dU = torch.zeros(nE, nx, dtype=torch.float64, device=dev)
    dH = torch.zeros(nx, nm, nm, dtype=torch.complex128, device=dev) if want_dH else None
    b.tr = _req(tr, "tr", torch.float64, nE, dev)
    b.info = _req(info, "info", torch.int32, 4 * nE, dev)
    pw = _req(w_tr, "w_tr", torch.float64, nE, dev)
    pn = _req(w_n, "w_n", torch.float64, nE * nx, dev)
    pl = _req(lead_H00, "lead_H00", torch.complex128, 2 * nm * nm, dev)
    if stream is None:
        stream = torch.cuda.current_stream()
    _check(lib().rgf_adjoint_batched(ctypes.byref(p), ctypes.byref(b), pw, pn, pl, _tp(J), _tp(dU),
                                     _tp(dH), ctypes.c_void_p(stream.cuda_stream)))
    return J, dU, dH


# constants of the reference's current sum (src/green_rgf.f90:18, 22, 24)
E0_REF = 1.6022e-19
HBAR_REF = 1.0546e-34
TPI_REF = 3.14159265359 * 2.0


def w_solve_batched(Vii, V1i, PL, PG, PR, want_i1=False, stream=None):
    """Screened interaction W of green_rgf_calc_w (src/green_rgf.f90:3399-3831, NBC = 0)
    for a batch of energies on the GPU.  Device tensors in the reference layout (see
    to_device_blocks): Vii, V1i (nx, nm, nm); PL, PG, PR (nE, nx, nm, nm).  Returns
    dict WL, WG, WR (and WRi1) of shape (nE, nx, nm, nm) and info (nE,), asynchronous
    on the current stream."""
    import torch
    nE, nx, nm = int(PR.shape[0]), int(PR.shape[1]), int(PR.shape[2])
    dev = PR.device
    c = torch.complex128
    for t, name, numel in ((Vii, "Vii", nx * nm * nm), (V1i, "V1i", nx * nm * nm), (PL, "PL", nE * nx * nm * nm),
                           (PG, "PG", nE * nx * nm * nm), (PR, "PR", nE * nx * nm * nm)):
        _req(t, name, c, numel, dev)
    out = {k: torch.empty_like(PR) for k in ("WL", "WG", "WR")}
    if want_i1:
        out["WRi1"] = torch.empty_like(PR)
    out["info"] = torch.zeros(nE, dtype=torch.int32, device=dev)
    b = WBatch(nm, nx, nE, 0, Vii.data_ptr(), V1i.data_ptr(), PL.data_ptr(), PG.data_ptr(), PR.data_ptr(),
               out["WL"].data_ptr(), out["WG"].data_ptr(), out["WR"].data_ptr(),
               out["WRi1"].data_ptr() if want_i1 else None, out["info"].data_ptr())
    s = stream or torch.cuda.current_stream()
    _check(lib().rgf_w_solve_batched(ctypes.byref(b), ctypes.c_void_p(s.cuda_stream)))
    return out


def current_scale(spindeg=2.0, nk=1):
    """scale of rgf_integrate's current for the reference's formula
    sum(tr) * dE * e0 / tpi / hbar * e0 * spindeg / nk (src/green_rgf.f90:3189, 788): amperes."""
    return E0_REF / TPI_REF / HBAR_REF * E0_REF * float(spindeg) / float(nk)


def integrate(E, midgap=None, dens_n=None, dens_p=None, tr=None, scale=None, stream=None):
    """Energy/k integrals of per-energy outputs on the GPU (rgf_integrate):
    calc_n_electron's nelec/pelec (src/green_rgf.f90:3920-3945) and the current.
    E float64 (nE,); dens_n / dens_p float64 (nk, nE, nx) or (nE, nx); tr (nk, nE)
    or (nE,); midgap float64 (nx,).  scale defaults to current_scale(2, nk).
    Returns (nelec, pelec, current) tensors (None where not computed)."""
    import torch
    dev = E.device
    nE = int(E.shape[0])
    ref = dens_n if dens_n is not None else dens_p
    if ref is not None:
        nx = int(ref.shape[-1])
        nk = int(ref.numel() // (nE * nx))
    elif tr is not None:
        nx = 1
        nk = int(tr.numel() // nE)
    else:
        raise RgfError("integrate: nothing to integrate")
    if scale is None:
        scale = current_scale(2.0, nk)
    nelec = pelec = cur = None
    if ref is not None:
        nelec = torch.zeros(nx, dtype=torch.float64, device=dev)
        pelec = torch.zeros(nx, dtype=torch.float64, device=dev)
    if tr is not None:
        cur = torch.zeros(1, dtype=torch.float64, device=dev)
    if stream is None:
        stream = torch.cuda.current_stream()
    _check(lib().rgf_integrate(nk, nE, nx, _req(E, "E", torch.float64, nE, dev),
                               _req(midgap, "midgap", torch.float64, nx, dev) if ref is not None else None,
                               _req(dens_n, "dens_n", torch.float64, nk * nE * nx, dev),
                               _req(dens_p, "dens_p", torch.float64, nk * nE * nx, dev),
                               _req(tr, "tr", torch.float64, nk * nE, dev), float(scale), _tp(nelec), _tp(pelec),
                               _tp(cur), ctypes.c_void_p(stream.cuda_stream)))
    return nelec, pelec, cur


def _device_batches(torch, devices):
    """rgf_batch array for rgf_solve_devices from a list of dicts with keys
    E, temp, mu, Hii, H1i and optional sigL, sigR, out (as solve_batched)."""
    arr = (Batch * len(devices))()
    for j, d in enumerate(devices):
        b = _batch(torch, d["E"], d["temp"], d["mu"], d["Hii"], d["H1i"], d.get("sigL"), d.get("sigR"), 0)
        _out_batch(torch, b, d.get("out") or {}, d["E"].device)
        arr[j] = b
    return arr


def solve_devices(p, devices, autotune=False, stream=None):
    """Many independent devices in one persistent launch (rgf_solve_devices).
    Returns (tile T per device, report dict).  Asynchronous on `stream`."""
    import torch
    arr = _device_batches(torch, devices)
    tiles = (ctypes.c_int * len(devices))()
    rep = MultiReport()
    if stream is None:
        stream = torch.cuda.current_stream()
    _check(lib().rgf_solve_devices(ctypes.byref(p), arr, len(devices), int(bool(autotune)), tiles, ctypes.byref(rep),
                                   ctypes.c_void_p(stream.cuda_stream)))
    return list(tiles), {"slots": rep.slots, "tuned": rep.tuned, "items": rep.items, "ws_bytes": rep.ws_bytes}


class DevicesGraph(Graph):
    """rgf_graph_create_devices: the many-device launch captured as one HIP graph."""

    def __init__(self, p, devices, stream, autotune=False):
        import torch
        self._keep = devices
        arr = _device_batches(torch, devices)
        self._stream = stream
        h = ctypes.c_void_p()
        _check(lib().rgf_graph_create_devices(ctypes.byref(p), arr, len(devices), int(bool(autotune)),
                                              ctypes.c_void_p(stream.cuda_stream), ctypes.byref(h)))
        self._h = h
