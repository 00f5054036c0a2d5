"""Reverse-mode (adjoint) RGF in numpy: the algorithm the GPU adjoint kernel
runs, written out step by step for testing.

Physical-mode device (H1i[..., l] = H_{l,l-1}), lead self-energies held fixed.
Objective per energy:  J = w_tr * tr + sum_l w_n[l] * Im Tr G^<_ll, with tr the
Meir-Wingreen current spectrum into the right contact (green_rgf.f90:4058-4066).
Gradients: dJ/dU_l for a potential U_l on the diagonal of slab l, and
Hbar_l = dJ/dRe(H1i_l) + i dJ/dIm(H1i_l) for the device-internal couplings
l = 1..nx-2 (0-based; the edge couplings also define the leads and are not
differentiated).  For a real J and complex matrix M the adjoint convention is
Mbar = dJ/dRe M + i dJ/dIm M, so dJ = Re Tr(Mbar^H dM).
Test infrastructure only.
"""
import numpy as np


def H(a):
    return np.conj(a.T)


def forward(E, eta, Hii, H1i, sigL, sigR, SigL, SigR, fL, fR):
    nm, _, N = Hii.shape
    I = np.eye(nm)
    z = E + 1j * eta
    Z = lambda a, l: a[..., l] if a is not None else 0.0
    g, gl, s = [None] * N, [None] * N, [None] * N
    KL = -(SigL - H(SigL))
    KR = -(SigR - H(SigR))
    g[0] = np.linalg.inv(z * I - Hii[..., 0] - Z(sigR, 0) - SigL)
    s[0] = Z(sigL, 0) + KL * fL
    gl[0] = g[0] @ s[0] @ H(g[0])
    for l in range(1, N - 1):
        T = H1i[..., l]
        g[l] = np.linalg.inv(z * I - Hii[..., l] - Z(sigR, l) - T @ g[l - 1] @ H(T))
        s[l] = Z(sigL, l) + T @ gl[l - 1] @ H(T)
        gl[l] = g[l] @ s[l] @ H(g[l])
    T = H1i[..., N - 1]
    G, Gn = [None] * N, [None] * N
    G[N - 1] = np.linalg.inv(z * I - Hii[..., N - 1] - Z(sigR, N - 1) - SigR - T @ g[N - 2] @ H(T))
    s[N - 1] = Z(sigL, N - 1) + KR * fR + T @ gl[N - 2] @ H(T)
    Gn[N - 1] = G[N - 1] @ s[N - 1] @ H(G[N - 1])
    for l in range(N - 2, -1, -1):
        T = H1i[..., l + 1]
        X = g[l] @ H(T)
        P = X @ G[l + 1]
        R = P @ T
        G[l] = g[l] + R @ g[l]
        Gn[l] = gl[l] + (X @ Gn[l + 1]) @ H(X) + R @ gl[l] + (gl[l] @ H(T)) @ H(P)
    tr = -np.real(np.trace(KR @ Gn[N - 1])) - fR * np.real(np.trace(KR @ (G[N - 1] - H(G[N - 1]))))
    return dict(g=g, gl=gl, s=s, G=G, Gn=Gn, tr=tr, KR=KR)


def adjoint(E, eta, Hii, H1i, sigL, sigR, SigL, SigR, fL, fR, w_tr, w_n, edges=False):
    """Returns (J, dU[nx], Hbar[nm, nm, nx]); with edges, also the adjoints of the
    contact self-energies (SigLbar, SigRbar) and, in Hbar[..., nx-1], the device's own
    use of the last coupling (the edge blocks of Hbar are zero otherwise)."""
    nm, _, N = Hii.shape
    f = forward(E, eta, Hii, H1i, sigL, sigR, SigL, SigR, fL, fR)
    g, gl, s, G, Gn, KR = f["g"], f["gl"], f["s"], f["G"], f["Gn"], f["KR"]
    I = np.eye(nm)
    J = w_tr * f["tr"] + sum(w_n[l] * np.imag(np.trace(Gn[l])) for l in range(N))
    gb = [np.zeros((nm, nm), complex) for _ in range(N)]
    glb = [np.zeros((nm, nm), complex) for _ in range(N)]
    Tb = np.zeros((nm, nm, N), complex)
    dU = np.zeros(N)
    # reverse of the backward sweep, l = 0 .. N-2
    Gb = np.zeros((nm, nm), complex)
    Lb = 1j * w_n[0] * I
    for l in range(N - 1):
        T = H1i[..., l + 1]
        X = g[l] @ H(T)
        P = X @ G[l + 1]
        R = P @ T
        Q = X @ Gn[l + 1]
        W = gl[l] @ H(T)
        gb[l] += Gb + H(R) @ Gb
        Rb = Gb @ H(g[l]) + Lb @ H(gl[l])
        glb[l] += Lb + H(R) @ Lb
        Qb = Lb @ X
        Xb = H(Lb) @ Q
        Wb = Lb @ P
        Pb = H(Lb) @ W
        glb[l] += Wb @ T
        Tb[..., l + 1] += H(Wb) @ gl[l]
        Xb += Qb @ H(Gn[l + 1])
        Lb_next = H(X) @ Qb
        Pb += Rb @ H(T)
        Tb[..., l + 1] += H(P) @ Rb
        Xb += Pb @ H(G[l + 1])
        Gb_next = H(X) @ Pb
        gb[l] += Xb @ T
        Tb[..., l + 1] += H(Xb) @ g[l]
        Gb, Lb = Gb_next, Lb_next + 1j * w_n[l + 1] * I
    # objective terms on the last slab (tr)
    Lb = Lb + w_tr * (-H(KR))
    Gb = Gb + w_tr * (-fR) * (H(KR) - KR)
    # reverse of the right boundary
    Gl_, s_ = G[N - 1], s[N - 1]
    M = Gl_ @ s_
    Mb = Lb @ Gl_
    Gb = Gb + H(Lb) @ M + Mb @ H(s_)
    sb = H(Gl_) @ Mb
    Ab = -H(Gl_) @ Gb @ H(Gl_)
    dU[N - 1] = -np.real(np.trace(Ab))
    T = H1i[..., N - 1]
    gb[N - 2] += H(T) @ (-Ab @ T)
    glb[N - 2] += H(T) @ (sb @ T)
    # contact self-energy Sigma_R: -Sigma_R in A, fR KR in s, and KR in tr itself
    Gn_ = Gn[N - 1]
    SRb = -Ab + fR * (H(sb) - sb) + w_tr * (H(Gn_) - Gn_) + 2 * w_tr * fR * (H(Gl_) - Gl_)
    # the device's own use of the last coupling: A = ... - T g T^H, s = ... + T g< T^H
    TbN = -(Ab @ T @ H(g[N - 2]) + H(Ab) @ T @ g[N - 2]) + sb @ T @ H(gl[N - 2]) + H(sb) @ T @ gl[N - 2]
    # reverse of the forward sweep, l = N-2 .. 0
    for l in range(N - 2, -1, -1):
        M = g[l] @ s[l]
        Mb = glb[l] @ g[l]
        gb[l] += H(glb[l]) @ M + Mb @ H(s[l])
        sb = H(g[l]) @ Mb
        Ab = -H(g[l]) @ gb[l] @ H(g[l])
        dU[l] = -np.real(np.trace(Ab))
        if l >= 1:
            T = H1i[..., l]
            K = T @ g[l - 1]
            Kb = -Ab @ T
            Tb[..., l] += -H(Ab) @ K + Kb @ H(g[l - 1])
            gb[l - 1] += H(T) @ Kb
            Nn = T @ gl[l - 1]
            Nb = sb @ T
            Tb[..., l] += H(sb) @ Nn + Nb @ H(gl[l - 1])
            glb[l - 1] += H(T) @ Nb
    Tb[..., 0] = 0
    if not edges:
        Tb[..., N - 1] = 0
        return J, dU, Tb
    Tb[..., N - 1] += TbN
    # Sigma_L: -Sigma_L in A_0, fL KL in s_0 (Ab, sb of slab 0 from the last step above)
    SLb = -Ab + fL * (H(sb) - sb)
    return J, dU, Tb, SLb, SRb


def stein(Q, P, P2, nmax=200, tol=1e-8):
    """X = Q + P X P2 (the sum of P^j Q P2^j) by doubling: X += P X P2, P <- P P,
    P2 <- P2 P2, until the sum of |P_ij| is at most tol (the next term is then below
    tol^2 relative to X) or nmax steps.  Returns (X, steps)."""
    X = Q.copy()
    for k in range(1, nmax + 1):
        X = X + P @ X @ P2
        P = P @ P
        P2 = P2 @ P2
        if np.abs(P).sum() <= tol:
            return X, k
    return X, nmax


def lead_left(C, g, Sb):
    """Adjoints of a left contact Sigma_L = C g C^H whose surface function solves
    g = (z - H00 - C g C^H)^-1 (the lead repeats the edge slab, C = H1i_1): given
    Sigma_L-bar, returns (H00-bar, C-bar)."""
    gb = H(C) @ Sb @ C
    Cb = Sb @ C @ H(g) + H(Sb) @ C @ g
    X, _ = stein(gb, H(C) @ H(g), H(g) @ C)  # total g-bar through the fixed point
    Hb00 = H(g) @ X @ H(g)                   # -Abar of A = z - H00 - C g C^H
    Cb = Cb + Hb00 @ C @ H(g) + H(Hb00) @ C @ g
    return Hb00, Cb


def lead_right(C, g, Sb):
    """The right contact: Sigma_R = C^H g C, g = (z - H00 - C^H g C)^-1 (C = H1i_nx)."""
    gb = C @ Sb @ H(C)
    Cb = H(g) @ C @ Sb + g @ C @ H(Sb)
    X, _ = stein(gb, C @ H(g), H(g) @ H(C))
    Hb00 = H(g) @ X @ H(g)
    Cb = Cb + H(g) @ C @ Hb00 + g @ C @ H(Hb00)
    return Hb00, Cb


def adjoint_with_leads(E, eta_dev, eta_lead, Hii, H1i, sigL, sigR, fL, fR, w_tr, w_n, sancho, tol, nmax):
    """The device whose contacts repeat its edge slabs (the reference's device:
    src/green_rgf.f90:3992-4003, 4032-4038): the gradients include the contacts'
    dependence on the edge blocks, through the decimation's fixed point.  A potential
    U_0 (U_nx) shifts the left (right) lead's onsite block with its edge slab; H1i_1
    (H1i_nx) is the left (right) contact coupling and the lead's own cell coupling.
    `sancho` is the surface-function routine (E, H00, H10, eta, tol, nmax)."""
    N = Hii.shape[2]
    Z = lambda a, l: a[..., l] if a is not None else 0.0
    tL, tR = H1i[..., 0], H1i[..., N - 1]
    gL = sancho(E, Hii[..., 0] + Z(sigR, 0), H(tL), eta_lead, tol, nmax)[0]
    gR = sancho(E, Hii[..., N - 1] + Z(sigR, N - 1), tR, eta_lead, tol, nmax)[0]
    SigL, SigR = tL @ gL @ H(tL), H(tR) @ gR @ tR
    J, dU, Tb, SLb, SRb = adjoint(E, eta_dev, Hii, H1i, sigL, sigR, SigL, SigR, fL, fR, w_tr, w_n, edges=True)
    HbL, CbL = lead_left(tL, gL, SLb)
    HbR, CbR = lead_right(tR, gR, SRb)
    dU[0] += np.real(np.trace(HbL))
    dU[N - 1] += np.real(np.trace(HbR))
    Tb[..., 0] = CbL
    Tb[..., N - 1] += CbR
    return J, dU, Tb
