"""Wannier tight-binding input of the reference, for the RGF engine.

* load_ham_dat: the `ham_dat` reader (src/wannierHam.f90:68-192, w90_load_from_file):
  `nvb spin_deg`, a comment line, the cell vectors alpha, beta, gamma (one per line),
  a comment line, the number of hopping lines and then `x y z i j Re Im` per line
  (H_ij of the cell offset x*alpha + y*beta), a comment line and the nb Wannier centres.
  Offsets along gamma are not used (the reference keys its table by x and y only).
* mat_def: the slab blocks of a device along alpha (src/wannierHam.f90:303-330,
  w90_MAT_DEF): a slab of ns cells, Hii[(i,k)] = sum_y Hr(i-k, y) e^{-i r.k},
  H1i[(i,k)] = sum_y Hr(i-k+ns, y) e^{-i r.k} (the coupling of a slab to the previous
  one), r = (i-k) alpha + y beta, k = kx xhat + ky yhat; offsets outside
  [xmin, xmax] are absent.
* device_blocks: the RGF driver's device (src/main.f90:500-530): `length` identical
  slabs, the potential pot[l] * potscale added to the first nb diagonal entries of
  slab l (the reference's loop runs over nb, not over the nb*ns orbitals of a slab).

The transverse wave vector is an argument (the reference's driver passes kt_CBM from a
band-structure scan, which is 0 for a one-dimensional system such as the (8,0) tube of
examples/cnt_8-0).
"""
import numpy as np


class W90:
    """Real-space Wannier Hamiltonian: Hr[:, :, x - xmin, y - ymin] = H(x alpha + y beta)."""

    def __init__(self, nvb, spin_deg, cell, Hr, xmin, ymin, centers):
        self.nvb = nvb
        self.spin_deg = spin_deg
        self.cell = cell                  # columns: alpha, beta, gamma (Angstrom)
        self.Hr = Hr
        self.xmin, self.ymin = xmin, ymin
        self.nb = Hr.shape[0]
        self.xmax = xmin + Hr.shape[2] - 1
        self.ymax = ymin + Hr.shape[3] - 1
        self.centers = centers            # (3, nb)
        alpha, beta, gamm = cell[:, 0], cell[:, 1], cell[:, 2]
        self.xhat = alpha / np.linalg.norm(alpha)
        yhat = -np.cross(self.xhat, gamm)
        self.yhat = yhat / np.linalg.norm(yhat)
        self.Lx = abs(float(alpha @ self.xhat))
        self.Ly = abs(float(beta @ self.yhat))
        self.Lz = float(np.linalg.norm(gamm))


def _tokens(lines, i):
    return lines[i].split()


def load_ham_dat(path, axis=None):
    """Read a ham_dat file (src/wannierHam.f90:68-192).  axis: optional permutation of the
    three coordinates (the reference's lreorder_axis / axis arguments, 1-based)."""
    with open(path) as f:
        lines = [ln for ln in f.read().splitlines() if ln.strip()]
    nvb, spin_deg = (int(float(t)) for t in _tokens(lines, 0)[:2])
    cell = np.array([[float(t) for t in _tokens(lines, 2 + c)[:3]] for c in range(3)]).T  # columns
    if axis is not None:
        cell = cell[:, [a - 1 for a in axis]]
    n = int(_tokens(lines, 6)[0])
    ham = np.array([[float(t) for t in _tokens(lines, 7 + r)[:7]] for r in range(n)])
    if axis is not None:
        ham[:, 0:3] = ham[:, [a - 1 for a in axis]]
    x = np.floor(ham[:, 0]).astype(int)
    y = np.floor(ham[:, 1]).astype(int)
    ib = np.floor(ham[:, 3]).astype(int) - 1
    jb = np.floor(ham[:, 4]).astype(int) - 1
    xmin, xmax, ymin, ymax = x.min(), x.max(), y.min(), y.max()
    nb = int(ham[:, 3].max())
    Hr = np.zeros((nb, nb, xmax - xmin + 1, ymax - ymin + 1), np.complex128)
    Hr[ib, jb, x - xmin, y - ymin] = ham[:, 5] + 1j * ham[:, 6]  # later lines overwrite, as in the reference
    first = 7 + n + 1  # after the centre comment line
    centers = np.array([[float(t) for t in _tokens(lines, first + b)[:3]] for b in range(nb)]).T
    if axis is not None:
        centers = centers[[a - 1 for a in axis], :]
    w = W90(nvb, spin_deg, cell, Hr, xmin, ymin, centers)
    # centres into the first cell (orthogonal cells), z centred on 0 (src/wannierHam.f90:176-186)
    alpha, beta, gamm = cell[:, 0], cell[:, 1], cell[:, 2]
    for b in range(nb):
        c = w.centers[:, b]
        if c[0] < 0:
            c += alpha
        if c[0] > w.Lx:
            c -= alpha
        if c[1] < 0:
            c += beta
        if c[1] > w.Ly:
            c -= beta
        if c[2] < 0:
            c += gamm
        if c[2] > w.Lz:
            c -= gamm
    w.centers[2, :] -= w.centers[2, :].mean()
    return w


def mat_def(w, kx, ky, ns):
    """(Hii, H1i) of a slab of ns cells along alpha (src/wannierHam.f90:303-330)."""
    nb = w.nb
    nm = nb * ns
    Hii = np.zeros((nm, nm), np.complex128)
    H1i = np.zeros((nm, nm), np.complex128)
    alpha, beta = w.cell[:, 0], w.cell[:, 1]
    kv = kx * w.xhat + ky * w.yhat
    for i in range(1, ns + 1):
        for k in range(1, ns + 1):
            for j in range(w.ymin, w.ymax + 1):
                r = (i - k) * alpha + j * beta
                ph = np.exp(-1j * float(r @ kv))
                bi, bk = slice((i - 1) * nb, i * nb), slice((k - 1) * nb, k * nb)
                if w.xmin <= i - k <= w.xmax:
                    Hii[bi, bk] += w.Hr[:, :, i - k - w.xmin, j - w.ymin] * ph
                if w.xmin <= i - k + ns <= w.xmax:
                    H1i[bi, bk] += w.Hr[:, :, i - k + ns - w.xmin, j - w.ymin] * ph
    return Hii, H1i


def read_pot(path, length):
    """First `length` values of a pot_dat file (one value per slab, src/main.f90:517-522)."""
    with open(path) as f:
        vals = [float(ln.split()[0]) for ln in f if ln.strip()]
    return np.array(vals[:length])


def device_blocks(w, ns, length, pot=None, potscale=1.0, kx=0.0, ky=0.0):
    """Hii, H1i (Fortran (nm, nm, length), nm = nb*ns) of the RGF driver's device
    (src/main.f90:500-530): identical slabs from mat_def, pot[l]*potscale on the first
    nb diagonal entries of slab l."""
    h00, h10 = mat_def(w, kx, ky, ns)
    nm = h00.shape[0]
    Hii = np.empty((nm, nm, length), np.complex128, order="F")
    H1i = np.empty((nm, nm, length), np.complex128, order="F")
    Hii[:, :, :] = h00[:, :, None]
    H1i[:, :, :] = h10[:, :, None]
    if pot is not None:
        idx = np.arange(w.nb)
        Hii[idx, idx, :] += (np.asarray(pot[:length]) * potscale)[None, :]
    return Hii, H1i
