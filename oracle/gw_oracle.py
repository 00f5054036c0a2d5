"""CPU oracle of the GW self-energy step of green_rgf_solve_gw_1D (test
infrastructure only: imported by tests, never by the product path).

Restates src/green_rgf.f90 (all line numbers below are in that file):

* 3186-3188: the real parts of G^r, G^<, G^> are zeroed before the step
  (dcmplx(0*dble(g), aimag(g))).
* polarization, 3190-3221: for every slab ix, row i and j in max(i-ndiag,1) ..
  min(nm,i+ndiag), nop in -nopmax..nopmax, iop = nop + nen/2 and
  ie in max(nop+1,1) .. min(nen,nen+nop):
      P^<(i,j,ix,iop) += G^<(i,j,ix,ie) * G^>(j,i,ix,ie-nop)
      P^>(i,j,ix,iop) += G^>(i,j,ix,ie) * G^<(j,i,ix,ie-nop)
      P^r(i,j,ix,iop) += G^<(i,j,ix,ie) * conj(G^r(i,j,ix,ie-nop)) + G^r(i,j,ix,ie) * G^<(j,i,ix,ie-nop)
  then P *= dE_P (zero elsewhere).
* self-energy, 3233-3276: for ie in 1..nen, nop in -nopmax..nopmax with
  ie > max(nop,1) and ie < nen+nop, iop = nop + nen/2:
      Sigma^<_new(i,j,ix,ie) += G^<(i,j,ix,ie-nop) * W^<(i,j,ix,iop)
      Sigma^>_new(i,j,ix,ie) += G^>(i,j,ix,ie-nop) * W^>(i,j,ix,iop)
      Sigma^r_new(i,j,ix,ie) += G^<(ie-nop) W^r(iop) + G^r(ie-nop) W^<(iop) + G^r(ie-nop) W^r(iop)
  then Sigma_new *= dE_S; Sigma^r_new = dcmplx(Re Sigma^r_new, Im(Sigma^>_new - Sigma^<_new)/2).
* 3278-3290: Sigma_gw += alpha_mix (Sigma_new - Sigma_gw); slabs 1, 2 take slab 3 and
  slabs nx, nx-1 take slab nx-2 (after the first copy).

Arrays are numpy complex128 of shape (nm, nm, nx, nen) (index order of the
reference).  Loops over nop are kept; the sums over ie, rows and columns are
vectorised, so the sum order within one (i, j, nop) differs from the reference's
sequential one only by rounding.
"""
import math

import numpy as np


def reference_params(nen, dE, spindeg=1.0):
    """nopmax and the two prefactors of the one-dimensional driver
    (src/green_rgf.f90:3192-3193, 3227-3229)."""
    nopmax = nen // 2 - 10
    dE_P = complex(0.0, -1.0 * dE / 2.0 / math.pi) * spindeg
    dE_S = complex(0.0, dE / 2.0 / math.pi)
    return nopmax, dE_P, dE_S


def imag_only(G):
    """dcmplx(0*dble(G), aimag(G)) (src/green_rgf.f90:3186-3188)."""
    return np.asarray(G, np.complex128).imag * 1j


def band_mask(nm, ndiag):
    i = np.arange(nm)[:, None]
    j = np.arange(nm)[None, :]
    return np.abs(i - j) <= ndiag


def polarization(Gl, Gg, Gr, ndiag, nopmax, dE_P):
    """P^<, P^>, P^r (src/green_rgf.f90:3190-3221)."""
    Gl, Gg, Gr = imag_only(Gl), imag_only(Gg), imag_only(Gr)
    nm, _, nx, nen = Gl.shape
    GlT = Gl.transpose(1, 0, 2, 3)
    GgT = Gg.transpose(1, 0, 2, 3)
    Pl = np.zeros_like(Gl)
    Pg = np.zeros_like(Gl)
    Pr = np.zeros_like(Gl)
    for nop in range(-nopmax, nopmax + 1):
        iop = nop + nen // 2 - 1                    # 0-based
        lo, hi = max(nop + 1, 1) - 1, min(nen, nen + nop)  # 0-based ie range [lo, hi)
        a = slice(lo, hi)
        b = slice(lo - nop, hi - nop)
        Pl[..., iop] = np.sum(Gl[..., a] * GgT[..., b], axis=-1)
        Pg[..., iop] = np.sum(Gg[..., a] * GlT[..., b], axis=-1)
        Pr[..., iop] = np.sum(Gl[..., a] * np.conj(Gr[..., b]) + Gr[..., a] * GlT[..., b], axis=-1)
    m = band_mask(nm, ndiag)[:, :, None, None]
    return Pl * m * dE_P, Pg * m * dE_P, Pr * m * dE_P


def selfenergy_new(Gl, Gg, Gr, Wl, Wg, Wr, ndiag, nopmax, dE_S):
    """Sigma^<_new, Sigma^>_new, Sigma^r_new (src/green_rgf.f90:3233-3277)."""
    Gl, Gg, Gr = imag_only(Gl), imag_only(Gg), imag_only(Gr)
    nm, _, nx, nen = Gl.shape
    Sl = np.zeros_like(Gl)
    Sg = np.zeros_like(Gl)
    Sr = np.zeros_like(Gl)
    ie = np.arange(1, nen + 1)
    for nop in range(-nopmax, nopmax + 1):
        iop = nop + nen // 2 - 1
        ok = (ie > max(nop, 1)) & (ie < nen + nop)
        e = ie[ok] - 1                 # 0-based ie
        k = ie[ok] - nop - 1           # 0-based ie - nop
        Sl[..., e] += Gl[..., k] * Wl[..., iop, None]
        Sg[..., e] += Gg[..., k] * Wg[..., iop, None]
        Sr[..., e] += Gl[..., k] * Wr[..., iop, None] + Gr[..., k] * Wl[..., iop, None] + Gr[..., k] * Wr[..., iop, None]
    m = band_mask(nm, ndiag)[:, :, None, None]
    Sl, Sg, Sr = Sl * m * dE_S, Sg * m * dE_S, Sr * m * dE_S
    Sr = Sr.real + 1j * ((Sg - Sl).imag / 2.0)
    return Sl, Sg, Sr


def mix_and_copy_edges(S_gw, S_new, alpha):
    """Sigma_gw + alpha (Sigma_new - Sigma_gw), then the edge copy (src/green_rgf.f90:3278-3290)."""
    S = S_gw + alpha * (S_new - S_gw)
    nx = S.shape[2]
    for ix in (1, 2):
        S[:, :, ix - 1, :] = S[:, :, 2, :]
    for ix in (1, 2):
        S[:, :, nx - ix, :] = S[:, :, nx - 3, :]
    return S


def selfenergy_step(Gl, Gg, Gr, Wl, Wg, Wr, Sl_gw, Sg_gw, Sr_gw, ndiag, nopmax, dE_S, alpha):
    """The updated running self-energies (Sigma^<, Sigma^>, Sigma^r)."""
    nl, ng, nr = selfenergy_new(Gl, Gg, Gr, Wl, Wg, Wr, ndiag, nopmax, dE_S)
    return (mix_and_copy_edges(Sl_gw, nl, alpha), mix_and_copy_edges(Sg_gw, ng, alpha),
            mix_and_copy_edges(Sr_gw, nr, alpha))
