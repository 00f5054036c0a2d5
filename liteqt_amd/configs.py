"""Synthetic workloads for the RGF engine (device Hamiltonians, energy grids, parameters).

The device is block-tridiagonal with nx slabs of nm orbitals.  Blocks use the
reference layout (complex128, Fortran order (nm, nm, nx)): Hii[:, :, l] is the
onsite block of slab l and H1i[:, :, l] the coupling H_{l,l-1} between slab l-1
and slab l, H1i[:, :, 0] also coupling slab 0 to the left lead
(green_rgf.f90:3992-4042).  Leads repeat the edge slabs.

Workloads:
  * cnt_80_ballistic  -- (8,0) zigzag carbon nanotube, pz nearest-neighbour
    tight binding, 32 orbitals per 4.26 A cell, linear potential drop.
  * random_device     -- random Hermitian onsite blocks plus per-slab random
    couplings, identical edge (lead) blocks.
  * energy_grid       -- E_i = emin + (i/nE)(emax-emin), i = 1..nE
    (src/main.f90:524, src/main3d.f90:658).
"""
import numpy as np

BOLTZ_REF = 8.61734e-05  # eV/K, green_rgf.f90:3968


def energy_grid(emin, emax, nE):
    """Energy grid of the reference drivers, src/main.f90:524."""
    i = np.arange(1, nE + 1, dtype=np.float64)
    return i / float(nE) * (emax - emin) + emin


def cnt_zigzag_cell(n=8, t=-2.7, acc=1.42):
    """(n,0) zigzag nanotube unit cell: 4n pz orbitals in 4 rings, cell length 3*acc.

    Returns (H00, H10) with H10 = H_{l,l-1}, the coupling from the previous cell.
    Bonds are found from the rolled-up coordinates with a 1.2*acc cutoff, so the
    connectivity is checked geometrically (each atom must have 3 neighbours).
    """
    a = np.sqrt(3.0) * acc
    R = n * a / (2 * np.pi)
    T = 3.0 * acc
    # ring z positions and angular offsets (in units of 2*pi/n)
    rings = [(0.0, 0.0), (0.5 * acc, 0.5), (1.5 * acc, 0.5), (2.0 * acc, 0.0)]
    xyz = []
    for z, off in rings:
        for j in range(n):
            ph = 2 * np.pi * (j + off) / n
            xyz.append((R * np.cos(ph), R * np.sin(ph), z))
    xyz = np.array(xyz)
    N = len(xyz)
    cut = 1.2 * acc
    H00 = np.zeros((N, N), np.complex128)
    H10 = np.zeros((N, N), np.complex128)
    nn = np.zeros(N, int)
    for i in range(N):
        for j in range(N):
            if i != j and np.linalg.norm(xyz[i] - xyz[j]) < cut:
                H00[i, j] = t
                nn[i] += 1
            # i in cell l, j in cell l-1 (shifted by -T)
            if np.linalg.norm(xyz[i] - (xyz[j] - np.array([0, 0, T]))) < cut:
                H10[i, j] = t
                nn[i] += 1
            if np.linalg.norm(xyz[i] - (xyz[j] + np.array([0, 0, T]))) < cut:
                nn[i] += 1
    if not np.all(nn == 3):
        raise RuntimeError("nanotube construction: every atom must have 3 neighbours")
    return H00, H10


def cnt_80_ballistic(L=100, drop=-0.2, t=-2.7):
    """(8,0) CNT device: L cells, onsite potential U_l = drop*(l-1)/(L-1) on the diagonal."""
    H00, H10 = cnt_zigzag_cell(8, t)
    nm = H00.shape[0]
    Hii = np.empty((nm, nm, L), np.complex128, order="F")
    H1i = np.empty((nm, nm, L), np.complex128, order="F")
    for l in range(L):
        U = drop * l / (L - 1) if L > 1 else 0.0
        Hii[:, :, l] = H00 + U * np.eye(nm)
        H1i[:, :, l] = H10
    return Hii, H1i


def random_device(nm, L, seed, u_range=0.3, coupling_sigma=0.5, identical_leads=True):
    """Random device: H_ll = (A + A^H)/(2 sqrt(nm)) + U_l I, A_ij ~ CN(0,1),
    U_l ~ U(-u_range, u_range); couplings T_l i.i.d. CN(0, (coupling_sigma/sqrt(nm))^2).
    With identical_leads the last slab repeats the first (onsite and coupling),
    so both leads are the same material."""
    rng = np.random.default_rng(seed)
    s = 1.0 / np.sqrt(2.0)
    A = (rng.standard_normal((nm, nm, L)) + 1j * rng.standard_normal((nm, nm, L))) * s
    Hii = np.asfortranarray((A + np.conj(np.transpose(A, (1, 0, 2)))) / (2.0 * np.sqrt(nm)))
    U = rng.uniform(-u_range, u_range, L)
    idx = np.arange(nm)
    Hii[idx, idx, :] += U[None, :]
    sig = coupling_sigma / np.sqrt(nm)
    H1i = np.asfortranarray((rng.standard_normal((nm, nm, L)) + 1j * rng.standard_normal((nm, nm, L))) * s * sig)
    if identical_leads and L > 1:
        Hii[:, :, L - 1] = Hii[:, :, 0]
        H1i[:, :, L - 1] = H1i[:, :, 0]
    return np.asfortranarray(Hii), np.asfortranarray(H1i)


def uniform_coupling(H1i):
    """Copy of H1i with every slab's coupling equal to slab 0's."""
    out = np.empty_like(H1i, order="F")
    out[:, :, :] = H1i[:, :, :1]
    return out


def scattering_self_energies(nm, L, seed, E=0.0, mu=0.0, kT=0.0259, scale=1e-3, nE=None):
    """Diagonal-block scattering self-energies consistent with
    Sigma^> - Sigma^< = Sigma^R - Sigma^A:  Sigma^R = -i Gamma/2 (Gamma PSD,
    |Gamma| ~ scale), Sigma^< = i f Gamma, Sigma^> = -i (1-f) Gamma.
    Returns (sigL, sigG, sigR) of shape (nm, nm, L) (or with a trailing energy
    axis when E is an array)."""
    rng = np.random.default_rng(seed)
    s = 1.0 / np.sqrt(2.0)
    X = (rng.standard_normal((nm, nm, L)) + 1j * rng.standard_normal((nm, nm, L))) * s
    Gam = np.einsum("ikl,jkl->ijl", X, np.conj(X)) * (scale / nm)
    Es = np.atleast_1d(np.asarray(E, np.float64))
    f = 1.0 / (1.0 + np.exp((Es - mu) / kT))
    sigR = np.asfortranarray(np.repeat((-0.5j * Gam)[..., None], len(Es), axis=3))
    sigL = np.asfortranarray(1j * Gam[..., None] * f[None, None, None, :])
    sigG = np.asfortranarray(-1j * Gam[..., None] * (1 - f)[None, None, None, :])
    if np.ndim(E) == 0:
        return sigL[..., 0].copy(order="F"), sigG[..., 0].copy(order="F"), sigR[..., 0].copy(order="F")
    return sigL, sigG, sigR


def scattering_blocks(nm, nx, seed, scale=1e-3):
    """Block-tridiagonal scattering matrices of a synthetic SCBA step: a
    Hermitian positive semidefinite Gamma = B B^H (B block lower bidiagonal,
    random) and a Hermitian Lambda, both of Frobenius size ~scale per block.
    Returns (Gam_d, Gam_lo, Lam_d, Lam_lo), (nm, nm, nx) each, with
    Gam_lo[..., l] = Gamma_{l,l-1} (block 0 zero)."""
    rng = np.random.default_rng(seed)

    def cn(*shape):
        return (rng.standard_normal(shape) + 1j * rng.standard_normal(shape)) / np.sqrt(2.0)

    Bd = cn(nx, nm, nm) / np.sqrt(2 * nm)
    Bl = cn(nx, nm, nm) / np.sqrt(2 * nm)
    Bl[0] = 0
    H = lambda a: np.conj(np.swapaxes(a, -1, -2))
    Gd = Bd @ H(Bd) + Bl @ H(Bl)
    Glo = np.zeros_like(Gd)
    Glo[1:] = Bl[1:] @ H(Bd[:-1])  # (B B^H)_{l,l-1} = B_{l,l-1} B_{l-1,l-1}^H
    A = cn(nx, nm, nm) / np.sqrt(nm)
    Ld = (A + H(A)) / 2
    Llo = cn(nx, nm, nm) / np.sqrt(nm)
    Llo[0] = 0
    tof = lambda a: np.asfortranarray(np.moveaxis(a, 0, -1)) * scale
    return tof(Gd), tof(Glo), tof(Ld), tof(Llo)


def offdiag_self_energies(blocks, f):
    """(sigL, sigR, sigR_lo, sigR_up, sigL_up) from scattering_blocks at
    occupation f: Sigma^R = Lambda - i Gamma / 2, Sigma^< = i f Gamma (so that
    Sigma^> - Sigma^< = Sigma^R - Sigma^A and Sigma^< = -f (Sigma^R - Sigma^A))."""
    Gd, Glo, Ld, Llo = blocks
    H = lambda a: np.conj(np.transpose(a, (1, 0, 2)))
    sigR = Ld - 0.5j * Gd
    sigR_lo = Llo - 0.5j * Glo
    sigR_up = H(Llo) - 0.5j * H(Glo)
    sigL = 1j * f * Gd
    sigL_up = 1j * f * H(Glo)
    c = lambda a: np.asfortranarray(a, dtype=np.complex128)
    return c(sigL), c(sigR), c(sigR_lo), c(sigR_up), c(sigL_up)
