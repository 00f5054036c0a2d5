"""green_RGF_RS (src/green_rgf.f90:3950-4164) in numpy over a batch of energies, each
block product a BLAS zgemm (numpy's OpenBLAS, multi-threaded) and each inverse LAPACK's.

Test infrastructure, not the product: it is the BLAS-backed CPU baseline that bench.py
times (a host with a tuned BLAS runs the reference's loop this way), and
tests/test_np_rgf.py holds it equal to the C oracle (oracle/rgf_oracle.c, rgf_impl) to
1e-10.  The step sequence is the C oracle's (the reference's, with its compat quirks):
  sancho             src/green_rgf.f90:4451-4537 (the decimation, per energy until its
                     own convergence; the nmax fallback to the bare surface)
  left contact       3992-4012          forward sweep      4013-4030
  right contact, tr  4032-4066          backward sweep     4069-4149 (cur, GLi1, GGi1,
                                                           GRi1, ldos, ndens, pdens)
  tre                4151-4160
Arrays: (nE, n, n) stacks of matrices (C order, the matrix itself, not Fortran memory).
"""
import numpy as np

COMPAT, PHYSICAL = 0, 1


def _h(a):
    return np.conj(np.swapaxes(a, -1, -2))


def _inv(a):
    return np.linalg.inv(a)


def ferm(a):
    return 1.0 / (1.0 + np.exp(a))


def sancho(E, eta, H00, H10, tol, nmax):
    """Surface function G00 per energy (E: (nE,)), H00, H10: (n, n) shared; returns
    (G00 (nE, n, n), iterations (nE,), nmax hit (nE,))."""
    nE, n = len(E), H00.shape[0]
    z = (E + 1j * eta)[:, None, None] * np.eye(n)
    HBB = np.broadcast_to(H00, (nE, n, n)).copy()
    HSS = HBB.copy()
    h10 = np.broadcast_to(H10, (nE, n, n)).copy()
    h01 = np.broadcast_to(_h(H10), (nE, n, n)).copy()
    it_run = np.zeros(nE, int)
    hit = np.zeros(nE, bool)
    act = np.ones(nE, bool)
    for it in range(1, nmax + 1):
        ix = np.nonzero(act)[0]
        if len(ix) == 0:
            break
        it_run[ix] = it
        A = _inv(z[ix] - HBB[ix])
        B = A @ h10[ix]
        C = h01[ix] @ B
        HSS[ix] += C
        HBB[ix] += C
        C2 = h10[ix] @ B
        B2 = A @ h01[ix]
        HBB[ix] += h10[ix] @ B2
        C3 = h01[ix] @ B2
        h10[ix] = C2
        h01[ix] = C3
        err = np.sum(np.abs(C3), axis=(1, 2))
        done = err < tol
        if it == nmax:
            nd = ix[~done]
            hit[nd] = True
            HSS[nd] = H00
            HBB[nd] = H00
        act[ix[done]] = False
    return _inv(z - HSS), it_run, hit


def rgf(mode, eta_dev, eta_lead, tol, nmax, boltz, temp, E, mu, Hii, H1i, sigL=None, sigR=None):
    """One batch of energies E (nE,) on the device Hii, H1i (nx, n, n) with diagonal
    scattering blocks sigL, sigR (nx, n, n) shared by the energies, or None.  Returns a
    dict of (nE, nx, n, n) arrays ldos, ndens, pdens, GRi1, GLi1, GGi1, cur (real) and
    (nE,) tr, tre, and the decimation counts."""
    E = np.asarray(E, np.float64)
    nE, nx, n = len(E), Hii.shape[0], Hii.shape[1]
    phys = mode == PHYSICAL
    I = np.eye(n)
    z = (E + 1j * eta_dev)[:, None, None] * I
    templ, tempr = temp
    temp_right = tempr if phys else templ
    temp_left_tre = templ if phys else tempr
    sR = (lambda l: sigR[l]) if sigR is not None else (lambda l: 0.0)
    sL = (lambda l: sigL[l]) if sigL is not None else (lambda l: 0.0)
    out = {k: np.zeros((nE, nx, n, n), complex) for k in ("ldos", "ndens", "pdens", "GRi1", "GLi1", "GGi1")}
    out["cur"] = np.zeros((nE, nx, n, n))
    Gl = np.zeros((nE, nx, n, n), complex)
    Gln = np.zeros_like(Gl)
    # left contact
    H00 = Hii[0] + sR(0)
    H10 = H1i[0]
    G00, itL, hitL = sancho(E, eta_lead, H00, _h(H10), tol, nmax)
    sigmal = H10 @ G00 @ _h(H10)
    fL = ferm((E - mu[0]) / (boltz * templ))[:, None, None]
    sig = -(sigmal - _h(sigmal)) * fL + sL(0)
    A = _inv(z - H00 - sigmal)
    Gl[:, 0] = A
    Gln[:, 0] = A @ sig @ _h(A)
    # forward sweep
    for l in range(1, nx - 1):
        H00 = Hii[l] + sR(l)
        H10 = H1i[l]
        A = _inv(z - H00 - H10 @ Gl[:, l - 1] @ _h(H10))
        Gl[:, l] = A
        Gln[:, l] = A @ (H10 @ Gln[:, l - 1] @ _h(H10) + sL(l)) @ _h(A)
    # right contact
    H00 = Hii[nx - 1] + sR(nx - 1)
    H10 = H1i[nx - 1]
    GR00, itR, hitR = sancho(E, eta_lead, H00, H10, tol, nmax)
    sigmar = _h(H10) @ GR00 @ H10
    G = _inv(z - H00 - sigmar - H10 @ Gl[:, nx - 2] @ _h(H10))
    out["ldos"][:, nx - 1] = G
    fR = ferm((E - mu[1]) / (boltz * temp_right))[:, None, None]
    T1 = sigmar - _h(sigmar)
    sig = -T1 * fR + H10 @ Gln[:, nx - 2] @ _h(H10) + sL(nx - 1)
    Gn = G @ sig @ _h(G)
    out["ndens"][:, nx - 1] = Gn
    Gp = Gn + G - _h(G)
    out["pdens"][:, nx - 1] = Gp
    tr = -np.real(np.trace((-T1 * fR) @ Gp - (-T1 * (fR - 1.0)) @ Gn, axis1=1, axis2=2))
    # backward sweep
    for l in range(nx - 2, -1, -1):
        H10 = H1i[l + 1] if phys else H1i[l]
        gl, gln = Gl[:, l], Gln[:, l]
        B = Gn @ H10 @ _h(gl) + G @ H10 @ gln
        cur = np.real(_h(H10) @ B)
        Glt = gl @ _h(H10) @ Gn + gln @ _h(H10) @ _h(G)  # G<_{l,l+1}
        out["GLi1"][:, l] = Glt
        cur = cur - np.real(H10 @ Glt)
        out["cur"][:, l] = cur
        if not phys:
            out["GGi1"][:, l] = _h(gl) @ _h(H10) @ Gp
        else:
            Glo = G @ H10 @ gl  # G^R_{l+1,l} with the pre-update G
        GN0 = gl @ _h(H10) @ G
        out["GRi1"][:, l] = GN0
        if phys:
            out["GGi1"][:, l] = Glt + GN0 - _h(Glo)
        Gnew = gl + GN0 @ H10 @ gl
        Gn = gln + gl @ _h(H10) @ Gn @ H10 @ _h(gl) + GN0 @ H10 @ gln + gln @ _h(H10) @ _h(GN0)
        G = Gnew
        out["ldos"][:, l] = G
        out["ndens"][:, l] = Gn
        out["pdens"][:, l] = Gn + G - _h(G)
    # tre
    Gp = Gn + G - _h(G)
    fLt = ferm((E - mu[0]) / (boltz * temp_left_tre))[:, None, None]
    T1 = sigmal - _h(sigmal)
    tre = np.real(np.trace((-T1 * fLt) @ Gp - (-T1 * (fLt - 1.0)) @ Gn, axis1=1, axis2=2))
    out["tr"], out["tre"] = tr, tre
    out["diag"] = np.stack([itL, itR, hitL.astype(int) | (hitR.astype(int) << 1)], axis=1)
    return out
