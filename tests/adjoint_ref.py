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


def adjoint(E, eta, Hii, H1i, sigL, sigR, SigL, SigR, fL, fR, w_tr, w_n):
    """Returns (J, dU[nx], Hbar[nm, nm, nx])."""
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
    Tb[..., N - 1] = 0
    return J, dU, Tb
