"""CPU oracle (test infrastructure) for the energy integrals after the RGF solve.

Only tests, smoke() and bench.py's CPU leg use it, as the checker.

calc_n_electron(G_lesser, G_greater, E, midgap) restates the reference's
subroutine calc_n_electron (src/green_rgf.f90:3920-3945) on its full arrays
(nm, nm, nx, nen, nk):
    dE = (E(2) - E(1)) / (2 pi),  dK = 2 pi / nk
    for every energy i, k-point ik, slab ix, orbital j:
      E(i) > midgap(ix):  nelec(ix) += Im G^<(j,j,ix,i,ik) dE dK
      otherwise:          pelec(ix) -= Im G^>(j,j,ix,i,ik) dE dK
with pi = 3.14159265359 (the module constant, src/green_rgf.f90:23).

current(tr, E, spindeg, nk) restates the SCBA drivers' terminal current
sum(tr) (E(2) - E(1)) e0 / tpi / hbar e0 spindeg / nk (src/green_rgf.f90:3189,
788) with the module's e0 = 1.6022e-19, hbar = 1.0546e-34 (lines 18, 22, 24).
"""
import numpy as np

PI_REF = 3.14159265359
TPI_REF = PI_REF * 2.0
E0_REF = 1.6022e-19
HBAR_REF = 1.0546e-34


def calc_n_electron(G_lesser, G_greater, E, midgap):
    """Vectorised restatement; G arrays (nm, nm, nx, nen) or (nm, nm, nx, nen, nk)."""
    Gl = np.asarray(G_lesser)
    Gg = np.asarray(G_greater)
    if Gl.ndim == 4:
        Gl, Gg = Gl[..., None], Gg[..., None]
    nm, _, nx, nen, nk = Gl.shape
    E = np.asarray(E, dtype=np.float64)
    dE = (E[1] - E[0]) / (2.0 * PI_REF)
    dK = 2.0 * PI_REF / nk
    # Im of the diagonal elements summed over orbitals: (nx, nen, nk)
    tl = np.imag(np.einsum("jjxek->xek", Gl))
    tg = np.imag(np.einsum("jjxek->xek", Gg))
    above = E[None, :] > np.asarray(midgap, dtype=np.float64)[:, None]  # (nx, nen)
    nelec = (tl * above[..., None]).sum(axis=(1, 2)) * dE * dK
    pelec = -(tg * (~above)[..., None]).sum(axis=(1, 2)) * dE * dK
    return nelec, pelec


def calc_n_electron_loops(G_lesser, G_greater, E, midgap):
    """Literal 1-based loop transcription of src/green_rgf.f90:3920-3945 (tiny inputs only)."""
    Gl = np.asarray(G_lesser)
    Gg = np.asarray(G_greater)
    if Gl.ndim == 4:
        Gl, Gg = Gl[..., None], Gg[..., None]
    nm, _, nx, nen, nk = Gl.shape
    nelec = [0.0] * (nx + 1)
    pelec = [0.0] * (nx + 1)
    dE = (E[2 - 1] - E[1 - 1]) / (2.0 * PI_REF)
    dK = 2.0 * PI_REF / float(nk)
    for i in range(1, nen + 1):
        for ik in range(1, nk + 1):
            for ix in range(1, nx + 1):
                for j in range(1, nm + 1):
                    if E[i - 1] > midgap[ix - 1]:
                        nelec[ix] = nelec[ix] + Gl[j - 1, j - 1, ix - 1, i - 1, ik - 1].imag * dE * dK
                    else:
                        pelec[ix] = pelec[ix] - Gg[j - 1, j - 1, ix - 1, i - 1, ik - 1].imag * dE * dK
    return np.array(nelec[1:]), np.array(pelec[1:])


def current(tr, E, spindeg=2.0, nk=1):
    """Terminal current in amperes: sum(tr) (E2 - E1) e0 / tpi / hbar e0 spindeg / nk."""
    tr = np.asarray(tr, dtype=np.float64)
    return float(np.sum(tr) * (E[1] - E[0]) * E0_REF / TPI_REF / HBAR_REF * E0_REF * float(spindeg) / float(nk))
