"""CPU oracle of the screened interaction W (test infrastructure only: imported by
tests, never by the product path).

* rgf_w: green_rgf_calc_w (src/green_rgf.f90:3399-3831) with NBC = 0 and no
  off-diagonal P blocks, the inputs of every call site (846, 1551, 3008), restated
  step by step:
    3454-3492  S = V P^r: S_ii = V_ii P_ii, S_{i+1,i} = V1i_i P_ii, S_{i,i+1} = V1i_i^H P_{i+1,i+1},
               the last slab's off-diagonal blocks copied from slab nx-1; M = I - S
    3700-3738  right-to-left: xlr_nx = M_nx^-1, Wlr_nx = xlr_nx V_nx;
               B = M_{i,i+1} xlr_{i+1}, xlr_i = (M_ii - B M_{i+1,i})^-1,
               Wlr_i = xlr_i (V_ii - B V1i_i)
    3739-3791  left-to-right: W_1 = Wlr_1, W_{1,2} = (V1i_1^T - W_1 M_{2,1}^T) xlr_2^T;
               W_i = Wlr_i - xlr_i M_{i,i-1} W_{i-1,i}, W_{i,i+1} likewise
    3793-3826  X = i Im W_{i,i+1};  W^<_i = W_i P^<_i W_i^H + X_a P^<_b (-X_a)^T
               + X_m^T P^<_m (-X_m), a = i, b = i+1 (a = nx-1, b = nx at i = nx),
               m = max(i-1, 1); W^> likewise
  The left-connected W^</W^> blocks, LL / LG and XR of the reference feed no output
  when NBC = 0 and are not formed.  The reference never writes WRi1 at slab nx
  (returned here as zero).
* w_full: green_calc_w_full (src/green_rgf.f90:3347-3396), the dense form of the 1D
  driver: W^r = (I - V P^r)^-1 V, W^{<,>} = W^r P^{<,>} W^r^H, diagonal blocks.

Arrays: numpy complex128, V blocks (nm, nm, nx), P / W blocks (nm, nm, nx) of one
energy.
"""
import numpy as np


def _inv(a):
    return np.linalg.inv(a)


def rgf_w(Vii, V1i, PL, PG, PR):
    """(WL, WG, WR, WRi1) of green_rgf_calc_w for one energy."""
    nm, _, nx = Vii.shape
    assert nx >= 2
    I = np.eye(nm)
    M = np.empty_like(PR)
    M1i = np.empty_like(PR)  # M_{i+1,i}
    Mi1 = np.empty_like(PR)  # M_{i,i+1}
    for ix in range(nx):
        M[:, :, ix] = I - Vii[:, :, ix] @ PR[:, :, ix]
    for i in range(nx - 1):
        M1i[:, :, i] = -(V1i[:, :, i] @ PR[:, :, i])
        Mi1[:, :, i] = -(V1i[:, :, i].conj().T @ PR[:, :, i + 1])
    M1i[:, :, nx - 1] = M1i[:, :, nx - 2]
    Mi1[:, :, nx - 1] = Mi1[:, :, nx - 2]
    xlr = np.empty_like(PR)
    wlr = np.empty_like(PR)
    A = _inv(M[:, :, nx - 1])
    xlr[:, :, nx - 1] = A
    wlr[:, :, nx - 1] = A @ Vii[:, :, nx - 1]
    for ix in range(nx - 2, -1, -1):
        B = Mi1[:, :, ix] @ xlr[:, :, ix + 1]
        C = B @ M1i[:, :, ix]
        A = _inv(M[:, :, ix] - C)
        xlr[:, :, ix] = A
        C = B @ V1i[:, :, ix]
        D = Vii[:, :, ix] - C
        wlr[:, :, ix] = A @ D
    WR = np.empty_like(PR)
    WRi1 = np.zeros_like(PR)
    WR[:, :, 0] = wlr[:, :, 0]
    B = WR[:, :, 0] @ M1i[:, :, 0].T
    C = V1i[:, :, 0].T - B
    WRi1[:, :, 0] = C @ xlr[:, :, 1].T
    for ix in range(1, nx):
        B = xlr[:, :, ix] @ M1i[:, :, ix - 1]
        C = B @ WRi1[:, :, ix - 1]
        WR[:, :, ix] = wlr[:, :, ix] - C
        if ix < nx - 1:
            B = WR[:, :, ix] @ M1i[:, :, ix].T
            C = V1i[:, :, ix].T - B
            WRi1[:, :, ix] = C @ xlr[:, :, ix + 1].T
    X = 1j * WRi1.imag
    out = []
    for P in (PL, PG):
        Wx = np.empty_like(PR)
        for ix in range(nx):
            Bx = WR[:, :, ix] @ P[:, :, ix] @ WR[:, :, ix].conj().T
            if ix < nx - 1:
                Bx = Bx + (X[:, :, ix] @ P[:, :, ix + 1]) @ (-X[:, :, ix]).T
            else:
                Bx = Bx + (X[:, :, ix - 1] @ P[:, :, nx - 1]) @ (-X[:, :, ix - 1]).T
            m = max(ix - 1, 0)
            Bx = Bx + (X[:, :, m].T @ P[:, :, m]) @ (-X[:, :, m])
            Wx[:, :, ix] = Bx
        out.append(Wx)
    return out[0], out[1], WR, WRi1


def block2full(blocks, nm, nx, offdiag=None):
    """Dense NT x NT matrix of diagonal blocks (and, with offdiag, V1i below the
    diagonal and its conjugate transpose above: src/green_rgf.f90:3326-3345)."""
    F = np.zeros((nm * nx, nm * nx), complex)
    for i in range(nx):
        F[i * nm:(i + 1) * nm, i * nm:(i + 1) * nm] = blocks[:, :, i]
    if offdiag is not None:
        for i in range(nx - 1):
            F[(i + 1) * nm:(i + 2) * nm, i * nm:(i + 1) * nm] = offdiag[:, :, i]
            F[i * nm:(i + 1) * nm, (i + 1) * nm:(i + 2) * nm] = offdiag[:, :, i].conj().T
    return F


def dense_w(Vii, V1i, PL, PG, PR):
    """Dense W^r = (I - V P^r)^-1 V and W^{<,>} = W^r P^{<,>} W^r^H (full matrices)."""
    nm, _, nx = Vii.shape
    V = block2full(Vii, nm, nx, V1i)
    Pr, Pl, Pg = (block2full(p, nm, nx) for p in (PR, PL, PG))
    M = np.eye(nm * nx) - V @ Pr
    WR = np.linalg.solve(M, V)
    return WR @ Pl @ WR.conj().T, WR @ Pg @ WR.conj().T, WR


def w_full(Vii, V1i, PL, PG, PR):
    """Diagonal blocks (WL, WG, WR) of green_calc_w_full."""
    nm, _, nx = Vii.shape
    full = dense_w(Vii, V1i, PL, PG, PR)
    return tuple(np.stack([F[i * nm:(i + 1) * nm, i * nm:(i + 1) * nm] for i in range(nx)], axis=2) for F in full)
