"""Dense brute-force Green's functions of an open block-tridiagonal device.

An independent formulation of what the RGF recursion computes
(the dense path of the reference, src/green.f90:2410-2587): assemble the whole
(nm*nx)^2 matrix z - H - Sigma, invert it with LAPACK, and read off the blocks.
Test infrastructure only.
"""
import numpy as np


def dag(a):
    return np.conj(np.swapaxes(a, -1, -2))


def fermi(E, mu, boltz, T):
    return 1.0 / (1.0 + np.exp((E - mu) / (boltz * T)))


def lead_self_energies(sancho, E, Hii, H1i, sigR, eta_lead, tol, nmax):
    """Sigma_L = H1i_1 G00_L H1i_1^H, Sigma_R = H1i_nx^H G00_R H1i_nx with the
    surface functions of leads repeating the edge slabs (green_rgf.f90:3992-4003,
    4032-4038).  `sancho` is the (separately pinned) decimation routine."""
    nx = Hii.shape[2]
    H00L = Hii[:, :, 0] + (sigR[:, :, 0] if sigR is not None else 0)
    H00R = Hii[:, :, nx - 1] + (sigR[:, :, nx - 1] if sigR is not None else 0)
    tL = H1i[:, :, 0]
    tR = H1i[:, :, nx - 1]
    GL = sancho(E, H00L, dag(tL), eta_lead, tol, nmax)[0]
    GR = sancho(E, H00R, tR, eta_lead, tol, nmax)[0]
    return tL @ GL @ dag(tL), dag(tR) @ GR @ tR


def dense_rgf(E, eta_dev, Hii, H1i, SigL, SigR, fL, fR, sigL=None, sigR=None, couplings=None, sig_off=None):
    """All blocks of G^R, G^<, G^> for the open device.

    couplings[l] (l = 1..nx-1) is H_{l,l-1}; default H1i[:, :, l] (the
    physically consistent indexing).  sig_off (optional) holds first
    off-diagonal scattering blocks: 'R_lo'[..., l] = Sigma^R_{l,l-1},
    'R_up'[..., l] = Sigma^R_{l-1,l}, 'L_up'[..., l] = Sigma^<_{l-1,l}
    (Sigma^<_{l,l-1} = -(Sigma^<_{l-1,l})^H).  Returns a dict with the
    reference's output names plus 'GRi1_lo' (G^R_{l+1,l}), 'G', 'Gl', 'Gg'
    (full matrices) and 'tr', 'tre'."""
    nm, _, nx = Hii.shape
    N = nm * nx
    if couplings is None:
        couplings = [H1i[:, :, l] for l in range(nx)]
    H = np.zeros((N, N), np.complex128)
    for l in range(nx):
        s = slice(l * nm, (l + 1) * nm)
        H[s, s] = Hii[:, :, l] + (sigR[:, :, l] if sigR is not None else 0)
        if l > 0:
            sp = slice((l - 1) * nm, l * nm)
            H[s, sp] = couplings[l]
            H[sp, s] = dag(couplings[l])
            if sig_off is not None and sig_off.get("R_lo") is not None:
                H[s, sp] += sig_off["R_lo"][:, :, l]
            if sig_off is not None and sig_off.get("R_up") is not None:
                H[sp, s] += sig_off["R_up"][:, :, l]
    A = (E + 1j * eta_dev) * np.eye(N) - H
    A[:nm, :nm] -= SigL
    A[N - nm:, N - nm:] -= SigR
    G = np.linalg.inv(A)
    Sl = np.zeros((N, N), np.complex128)
    if sigL is not None:
        for l in range(nx):
            s = slice(l * nm, (l + 1) * nm)
            Sl[s, s] = sigL[:, :, l]
    if sig_off is not None and sig_off.get("L_up") is not None:
        for l in range(1, nx):
            s, sp = slice(l * nm, (l + 1) * nm), slice((l - 1) * nm, l * nm)
            Sl[sp, s] = sig_off["L_up"][:, :, l]
            Sl[s, sp] = -dag(sig_off["L_up"][:, :, l])
    GamL = -(SigL - dag(SigL))  # i*Gamma_L
    GamR = -(SigR - dag(SigR))
    Sl[:nm, :nm] += GamL * fL
    Sl[N - nm:, N - nm:] += GamR * fR
    Gl = G @ Sl @ dag(G)
    Gg = Gl + G - dag(G)

    def b(M, i, j):
        return M[i * nm:(i + 1) * nm, j * nm:(j + 1) * nm]

    out = {"G": G, "Gl": Gl, "Gg": Gg}
    shp = (nm, nm, nx)
    for key, M in (("ldos", G), ("ndens", Gl), ("pdens", Gg)):
        o = np.zeros(shp, np.complex128, order="F")
        for l in range(nx):
            o[:, :, l] = b(M, l, l)
        out[key] = o
    for key, M in (("GRi1", G), ("GLi1", Gl), ("GGi1", Gg)):
        o = np.zeros(shp, np.complex128, order="F")
        for l in range(nx - 1):
            o[:, :, l] = b(M, l, l + 1)
        out[key] = o
    o = np.zeros(shp, np.complex128, order="F")
    for l in range(nx - 1):
        o[:, :, l] = b(G, l + 1, l)
    out["GRi1_lo"] = o
    cur = np.zeros(shp, np.float64, order="F")
    for l in range(nx - 1):
        Hc = couplings[l + 1]
        cur[:, :, l] = np.real(dag(Hc) @ b(Gl, l + 1, l)) - np.real(Hc @ b(Gl, l, l + 1))
    out["cur"] = cur
    # Meir-Wingreen spectra at the right / left contacts (green_rgf.f90:4058-4066, 4151-4160)
    GlR, GgR = b(Gl, nx - 1, nx - 1), b(Gg, nx - 1, nx - 1)
    GlL, GgL = b(Gl, 0, 0), b(Gg, 0, 0)
    out["tr"] = -np.real(np.trace(GamR * fR @ GgR - GamR * (fR - 1.0) @ GlR))
    out["tre"] = np.real(np.trace(GamL * fL @ GgL - GamL * (fL - 1.0) @ GlL))
    out["GamL"] = GamL
    out["GamR"] = GamR
    return out


def relerr(a, b):
    """Relative Frobenius error |a-b|/|b| (absolute when b == 0)."""
    nb = np.linalg.norm(b)
    d = np.linalg.norm(a - b)
    return d / nb if nb > 0 else d


def blockwise_relerr(a, b):
    """Max over blocks (last axis = slab) of the relative Frobenius error."""
    return max(relerr(a[:, :, l], b[:, :, l]) for l in range(a.shape[2]))
