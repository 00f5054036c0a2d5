"""Output files of the reference's RGF drivers (src/green_rgf.f90:4541-4923), for the
arrays the engine returns (numpy, the reference's Fortran index order).

* fortran_e: Fortran `Ew.d` editing (scale factor 0): `0.dddd` times a power of ten,
  right-justified in w columns, `E+xx` for two-digit exponents and `+xxx` (no `E`) for
  three, nearest-even rounding of the exact binary value (what ROCm's flang prints).
* write_spectrum (src/green_rgf.f90:4541-4571): `<dataset><iiii>.dat`, per energy and
  slab j and cell k of the slab one line `(j-1) Lx NS + (k-1) Lx, E, Re tr * coeff(1),
  Im tr * coeff(2)` with tr the trace of the cell's nb x nb diagonal block, format
  `(4E18.4)`, an empty line after each energy.
* write_current_spectrum (4681-4702): `<dataset>_<iiii>.dat`, per energy and bond j one
  line `(j-1) Lx, E, sum_ij cur(i, j, bond, E)`, `(3E18.4)`, an empty line per energy.
* write_transmission_spectrum (4847-4858): `<dataset>_<iiii>.dat`, `E, tr(E)` per line,
  `(2E18.4)`.
* calc_block_current (4602-4640): the bond current density between neighbouring cells
  of a slab, B(io, jo) = H(i, j) G^<(j, i) - H(j, i) G^<(i, j) (i in cell ib, j in cell
  ib + 1) times dE e0 / (2 pi hbar) e0 spindeg, with the reference's constants
  (e0 = 1.6022e-19 C, hbar = 1.0546e-34 J s, 2 pi = 6.28318530718); the last bond of a
  slab (to the next slab) repeats the slab's last intra-slab value, as in the reference.

`iiii` is the reference's STRING(i): the iteration number as four digits.
"""
import os

import numpy as np

E0 = 1.6022e-19
HBAR = 1.0546e-34
TPI = 6.28318530718


def string4(i):
    """STRING(i) of the reference (src/green_rgf.f90:4927-4940): four zero-padded digits."""
    if i > 9999 or i < 0:
        raise ValueError("iteration number must be in 0..9999")
    return "%04d" % i


def fortran_e(x, w=18, d=4):
    x = float(x)
    if x == 0.0:
        s = ("-" if np.signbit(x) else "") + "0." + "0" * d + "E+00"
    else:
        m, e = ("%.*e" % (d - 1, x)).split("e")
        neg = m.startswith("-")
        digits = m.lstrip("-").replace(".", "")
        ex = int(e) + 1
        s = ("-" if neg else "") + "0." + digits
        s += ("E%+03d" % ex) if abs(ex) <= 99 else ("%+04d" % ex)
    return s.rjust(w)


def parse_fortran_e(field):
    """Inverse of fortran_e for reading the files back (`0.1234-101` has no `E`)."""
    t = field.strip()
    k = max(t.rfind("+"), t.rfind("-"))
    if k > 0 and t[k - 1] not in "Ee":
        t = t[:k] + "E" + t[k:]
    return float(t)


def _line(vals, w=18, d=4):
    return "".join(fortran_e(v, w, d) for v in vals) + "\n"


def spectrum_traces(G, nb, ns):
    """Traces of the nb x nb diagonal blocks of each cell: complex (ns, length, nen) from
    G (nb*ns, nb*ns, length, nen)."""
    G = np.asarray(G)
    nm, _, length, nen = G.shape
    if nm != nb * ns:
        raise ValueError("G's block size must be nb*ns")
    return np.einsum("iijk->ijk", G).reshape(ns, nb, length, nen).sum(axis=1)


def write_spectrum(dataset, i, G, en, nb, ns, lx, coeff, append=False):
    """G: complex (nb*ns, nb*ns, length, nen)."""
    write_spectrum_traces(dataset, i, spectrum_traces(G, nb, ns), en, lx, coeff, append)


def write_spectrum_traces(dataset, i, tr, en, lx, coeff, append=False):
    """write_spectrum from the cell traces (ns, length, nen) of spectrum_traces."""
    ns, length, nen = tr.shape
    with open(f"{dataset}{string4(i)}.dat", "a" if append else "w") as f:
        for ie in range(nen):
            for j in range(length):
                for k in range(ns):
                    t = tr[k, j, ie]
                    f.write(_line((j * lx * ns + k * lx, en[ie], t.real * coeff[0], t.imag * coeff[1])))
            f.write("\n")


def write_current_spectrum(dataset, i, cur, en, lx):
    """cur: real (nb, nb, length, nen) bond current densities."""
    write_current_totals(dataset, i, np.asarray(cur).sum(axis=(0, 1)), en, lx)


def write_current_totals(dataset, i, tot, en, lx):
    """write_current_spectrum from the per-bond sums tot (length, nen)."""
    length, nen = tot.shape
    with open(f"{dataset}_{string4(i)}.dat", "w") as f:
        for ie in range(nen):
            for j in range(length):
                f.write(_line((float(j) * lx, en[ie], tot[j, ie])))
            f.write("\n")


def write_transmission_spectrum(dataset, i, tr, en):
    with open(f"{dataset}_{string4(i)}.dat", "w") as f:
        for ie in range(len(en)):
            f.write(_line((en[ie], float(np.real(tr[ie])))))


def calc_block_current(H, G_lesser, en, spindeg, nb, ns, dE=None):
    """(tot_cur, tot_ecur, jdens): real (nb, nb, nx*ns), (nb, nb, nx*ns), (nb, nb, nx*ns, nen)
    from H (nm, nm, nx) and G^< (nm, nm, nx, nen), nm = nb*ns, ns >= 2.  dE: the grid
    spacing (default en[1] - en[0]; give it when en is a chunk of the grid)."""
    H = np.asarray(H)
    Gl = np.asarray(G_lesser)
    nm, _, nx, nen = Gl.shape
    if nm != nb * ns or ns < 2:
        raise ValueError("need nm = nb*ns and ns >= 2 (the reference leaves the bond undefined for ns = 1)")
    scale = ((en[1] - en[0]) if dE is None else dE) * E0 / TPI / HBAR * E0 * float(spindeg)
    jdens = np.zeros((nb, nb, nx * ns, nen))
    for ix in range(nx):
        for ib in range(ns - 1):
            a = slice(ib * nb, (ib + 1) * nb)
            b = slice((ib + 1) * nb, (ib + 2) * nb)
            # B(io, jo) = H(i, j) G<(j, i) - H(j, i) G<(i, j), i = io in cell ib, j = jo in cell ib+1
            Bm = (H[a, b, ix][:, :, None] * np.swapaxes(Gl[b, a, ix, :], 0, 1)
                  - np.swapaxes(H[b, a, ix], 0, 1)[:, :, None] * Gl[a, b, ix, :]) * scale
            jdens[:, :, ix * ns + ib, :] = Bm.real
        jdens[:, :, ix * ns + ns - 1, :] = jdens[:, :, ix * ns + ns - 2, :]
    tot_cur = jdens.sum(axis=3)
    tot_ecur = (jdens * np.asarray(en)[None, None, None, :]).sum(axis=3)
    return tot_cur, tot_ecur, jdens
