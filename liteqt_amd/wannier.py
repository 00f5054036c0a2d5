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

* bare_coulomb / coulomb_blocks / device_coulomb: the bare Coulomb interaction between
  Wannier centres (src/wannierHam.f90:464-488), its slab blocks (w90_bare_coulomb_blocks,
  368-397) and the RGF driver's device copies (src/main.f90:531-537), the V of the GW loop.

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


# constants of bare_coulomb (src/wannierHam.f90:468-471); epsilon0 is a default-real
# (single precision) literal there, so its value is that of the float32 rounding
_PI_COUL = 3.14159265359
_E_COUL = 1.6e-19
_EPS0_COUL = float(np.float32(8.85e-12))


def bare_coulomb(w, a1, a2, eps, r0, ldiag):
    """(nb, nb) real bare Coulomb matrix between the centres of cell 0 and the cell at
    a1 alpha + a2 beta, in eV (src/wannierHam.f90:464-488): e / (4 pi eps0 eps |r| 1e-10)
    tanh(|r| / r0) with r = a1 alpha + a2 beta + c_i - c_j (Angstrom); at r = 0 the
    self-interaction e / (4 pi eps0 eps 1e-10) / r0 when ldiag, else 0.1."""
    alpha, beta = w.cell[:, 0], w.cell[:, 1]
    r = (a1 * alpha + a2 * beta)[:, None, None] + w.centers[:, :, None] - w.centers[:, None, :]
    normr = np.sqrt((r * r).sum(axis=0))
    with np.errstate(divide="ignore", invalid="ignore"):
        v = _E_COUL / (4.0 * _PI_COUL * _EPS0_COUL * eps * normr * 1.0e-10) * np.tanh(normr / r0)
    self_v = _E_COUL / (4.0 * _PI_COUL * _EPS0_COUL * eps * 1.0e-10) * (1.0 / r0) if ldiag else 0.1
    return np.where(normr > 0.0, v, self_v)


def coulomb_blocks(w, ns, eps, r0, ldiag):
    """(Vii, V1i) of a slab of ns cells (w90_bare_coulomb_blocks at kx = ky = 0,
    src/wannierHam.f90:368-397): block (i, k) of Vii sums bare_coulomb(i-k, j) over the
    Hamiltonian's y offsets j when i-k lies in its x range; block (i, k) of V1i sums the
    same bare_coulomb(i-k, j) when i-k+ns lies in that range (the reference passes i-k,
    not i-k+ns, there)."""
    nb = w.nb
    nm = nb * ns
    Vii = np.zeros((nm, nm), np.complex128)
    V1i = np.zeros((nm, nm), np.complex128)
    for i in range(1, ns + 1):
        for k in range(1, ns + 1):
            bi, bk = slice((i - 1) * nb, i * nb), slice((k - 1) * nb, k * nb)
            for j in range(w.ymin, w.ymax + 1):
                if w.xmin <= i - k <= w.xmax:
                    Vii[bi, bk] += bare_coulomb(w, i - k, j, eps, r0, ldiag)
                if w.xmin <= i - k + ns <= w.xmax:
                    V1i[bi, bk] += bare_coulomb(w, i - k, j, eps, r0, ldiag)
    return Vii, V1i


def device_coulomb(w, ns, length, eps, r0, ldiag):
    """Vii, V1i (Fortran (nm, nm, length)) of the RGF driver: the slab blocks copied to
    every slab (src/main.f90:531-537)."""
    v00, v10 = coulomb_blocks(w, ns, eps, r0, ldiag)
    nm = v00.shape[0]
    Vii = np.empty((nm, nm, length), np.complex128, order="F")
    V1i = np.empty((nm, nm, length), np.complex128, order="F")
    Vii[:, :, :] = v00[:, :, None]
    V1i[:, :, :] = v10[:, :, None]
    return Vii, V1i
