"""Reverse-mode (adjoint) RGF with a block-tridiagonal scattering self-energy,
in numpy: the algorithm the GPU adjoint kernel runs, step by step.

Physical device (H1i[..., l] = H_{l,l-1}) with fixed contacts and fixed Sigma:
diagonal sigR, sigL and first off-diagonal sigR_lo[l] = Sigma^R_{l,l-1},
sigR_up[l] = Sigma^R_{l-1,l}, sigL_up[l] = Sigma^<_{l-1,l} (any may be None).
Objective per energy J = w_tr * tr + sum_l w_n[l] * Im Tr G^<_ll; gradients
dJ/dU_l (potential on slab l's diagonal) and Hbar_l = dJ/dRe H1i_l +
i dJ/dIm H1i_l for l = 1..nx-2 (0-based).  With no off-diagonal blocks this is
the recursion of tests/adjoint_ref.py written with d = H1i_l, u = H1i_l^H.
Forward (d_l = H1i_l + Sigma^R_{l,l-1}, u_l = H1i_l^H + Sigma^R_{l-1,l}, s_l = Sigma^<_{l-1,l}):
  g_l  = (z - H_l - Sigma^R_l - d_l g_{l-1} u_l)^-1
  g<_l = g_l [Sigma^<_l + d_l g<_{l-1} d_l^H + V - V^H] g_l^H,  V = d_l g_{l-1} s_l
Backward (pair l, l+1):
  Q = g_l u, R = Q G_{l+1}, G_l = g_l + R d g_l,
  G<_l = g<_l + Q G<_{l+1} Q^H + W - W^H,  W = g_l s R^H + R d g<_l.
Test infrastructure only.
"""
import numpy as np


def H(a):
    return np.conj(a.T)


def adjoint(E, eta, Hii, H1i, sigL, sigR, sigR_lo, sigR_up, sigL_up, SigL, SigR, fL, fR, w_tr, w_n):
    nm, _, N = Hii.shape
    I = np.eye(nm)
    z = E + 1j * eta
    Z0 = np.zeros((nm, nm), complex)
    blk = lambda a, l: a[..., l] if a is not None else Z0
    KL = -(SigL - H(SigL))
    KR = -(SigR - H(SigR))
    d = [None] + [H1i[..., l] + blk(sigR_lo, l) for l in range(1, N)]
    u = [None] + [H(H1i[..., l]) + blk(sigR_up, l) for l in range(1, N)]
    s = [None] + [blk(sigL_up, l) for l in range(1, N)]
    g, gl, S, B = [None] * N, [None] * N, [None] * N, [None] * N
    # forward (the last slab's g, gl are G_N, G<_N)
    for l in range(N):
        A = z * I - Hii[..., l] - blk(sigR, l)
        Sl = blk(sigL, l).copy()
        if l == 0:
            A = A - SigL
            Sl = Sl + fL * KL
        if l == N - 1:
            A = A - SigR
            Sl = Sl + fR * KR
        if l > 0:
            B[l] = d[l] @ g[l - 1]
            A = A - B[l] @ u[l]
            V = B[l] @ s[l]
            Sl = Sl + d[l] @ gl[l - 1] @ H(d[l]) + V - H(V)
        g[l] = np.linalg.inv(A)
        S[l] = Sl
        gl[l] = g[l] @ Sl @ H(g[l])
    G, Gn = [None] * N, [None] * N
    G[N - 1], Gn[N - 1] = g[N - 1], gl[N - 1]
    for l in range(N - 2, -1, -1):
        Q = g[l] @ u[l + 1]
        R = Q @ G[l + 1]
        G[l] = g[l] + R @ d[l + 1] @ g[l]
        W = g[l] @ s[l + 1] @ H(R) + R @ d[l + 1] @ gl[l]
        Gn[l] = gl[l] + Q @ Gn[l + 1] @ H(Q) + W - H(W)
    tr = -np.real(np.trace(KR @ Gn[N - 1])) - fR * np.real(np.trace(KR @ (G[N - 1] - H(G[N - 1]))))
    J = w_tr * tr + sum(w_n[l] * np.imag(np.trace(Gn[l])) for l in range(N))

    gb = [np.zeros((nm, nm), complex) for _ in range(N)]
    glb = [np.zeros((nm, nm), complex) for _ in range(N)]
    db = [np.zeros((nm, nm), complex) for _ in range(N)]
    ub = [np.zeros((nm, nm), complex) for _ in range(N)]
    dU = np.zeros(N)
    # reverse of the backward sweep, l = 0 .. N-2; (Gb, Lb) are the adjoints of (G_l, G<_l)
    Gb = np.zeros((nm, nm), complex)
    Lb = 1j * w_n[0] * I
    for l in range(N - 1):
        dd, uu, ss = d[l + 1], u[l + 1], s[l + 1]
        Q = g[l] @ uu
        R = Q @ G[l + 1]
        Rd = R @ dd
        QGn = Q @ Gn[l + 1]
        gs = g[l] @ ss
        # G<_l = T0 + W - W^H, T0 = QGn Q^H + g<_l, W = gs R^H + Rd g<_l
        Wb = Lb - H(Lb)
        gsb = Wb @ R
        Rb = H(Wb) @ gs
        Rdb = Wb @ H(gl[l])
        glb[l] += H(Rd) @ Wb + Lb
        QGnb = Lb @ Q
        Qb = H(Lb) @ QGn
        # G_l = g + Rd g
        gb[l] += Gb + H(Rd) @ Gb
        Rdb += Gb @ H(g[l])
        # Rd = R d
        Rb += Rdb @ H(dd)
        db[l + 1] += H(R) @ Rdb
        # gs = g s  (Sigma fixed)
        gb[l] += gsb @ H(ss)
        # QGn = Q G<_{l+1}, R = Q G_{l+1}, Q = g u
        Qb += QGnb @ H(Gn[l + 1]) + Rb @ H(G[l + 1])
        Lb_next = H(Q) @ QGnb
        Gb_next = H(Q) @ Rb
        gb[l] += Qb @ H(uu)
        ub[l + 1] += H(g[l]) @ Qb
        Gb, Lb = Gb_next, Lb_next + 1j * w_n[l + 1] * I
    # objective terms on the last slab (tr)
    Lb = Lb + w_tr * (-H(KR))
    Gb = Gb + w_tr * (-fR) * (H(KR) - KR)
    gb[N - 1] += Gb
    glb[N - 1] += Lb
    # reverse of the forward sweep, l = N-1 .. 0
    for l in range(N - 1, -1, -1):
        M = g[l] @ S[l]
        Mb = glb[l] @ g[l]
        gb[l] += H(glb[l]) @ M + Mb @ H(S[l])
        Sb = H(g[l]) @ Mb
        Ab = -H(g[l]) @ gb[l] @ H(g[l])
        dU[l] = -np.real(np.trace(Ab))
        if l >= 1:
            # A = ... - B u, B = d g_{l-1};  S = ... + Y d^H + V - V^H, Y = d g<_{l-1}, V = B s
            Bb = -Ab @ H(u[l])
            ub[l] += -H(B[l]) @ Ab
            Y = d[l] @ gl[l - 1]
            Yb = Sb @ d[l]
            db[l] += H(Sb) @ Y + Yb @ H(gl[l - 1])
            glb[l - 1] += H(d[l]) @ Yb
            Vb = Sb - H(Sb)
            Bb += Vb @ H(s[l])
            db[l] += Bb @ H(g[l - 1])
            gb[l - 1] += H(d[l]) @ Bb
    Tb = np.zeros((nm, nm, N), complex)
    for l in range(1, N - 1):
        Tb[..., l] = db[l] + H(ub[l])
    return J, dU, Tb
