"""The W oracle (oracle/w_oracle.py) pinned without the reference build: the block
recursion of green_rgf_calc_w against a dense inverse, its W^< / W^> formula against
an independent neighbour sum over dense W blocks, and the dense form's own identities."""
import numpy as np
import pytest

from oracle import w_oracle as wo


def _rand(rng, *shape):
    return rng.standard_normal(shape) + 1j * rng.standard_normal(shape)


def _inputs(rng, nm, nx, symmetric, scale=0.3):
    """V Hermitian block-tridiagonal (real, hence symmetric, when symmetric); P^r
    block-diagonal (complex symmetric when symmetric); P^<, P^> anti-Hermitian."""
    if symmetric:
        Vii = rng.standard_normal((nm, nm, nx))
        Vii = Vii + Vii.transpose(1, 0, 2)
        V1i = rng.standard_normal((nm, nm, nx))
        PR = _rand(rng, nm, nm, nx)
        PR = scale / nm * (PR + PR.transpose(1, 0, 2))
    else:
        Vii = _rand(rng, nm, nm, nx)
        Vii = Vii + Vii.conj().transpose(1, 0, 2)
        V1i = _rand(rng, nm, nm, nx)
        PR = scale / nm * _rand(rng, nm, nm, nx)
    PL = _rand(rng, nm, nm, nx)
    PL = scale / nm * (PL - PL.conj().transpose(1, 0, 2))
    PG = _rand(rng, nm, nm, nx)
    PG = scale / nm * (PG - PG.conj().transpose(1, 0, 2))
    return Vii, V1i, PL, PG, PR


def _blk(F, nm, i, j):
    return F[i * nm:(i + 1) * nm, j * nm:(j + 1) * nm]


def _rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


@pytest.mark.parametrize("nm,nx", [(1, 2), (3, 2), (4, 3), (5, 6), (8, 9)])
def test_recursion_gives_dense_blocks_for_symmetric_inputs(nm, nx):
    """W^r_ii and W^r_{i,i+1} equal the blocks of (I - V P^r)^-1 V (symmetric V, P^r)."""
    rng = np.random.default_rng(10 * nm + nx)
    Vii, V1i, PL, PG, PR = _inputs(rng, nm, nx, True)
    WL, WG, WR, WRi1 = wo.rgf_w(Vii, V1i, PL, PG, PR)
    _, _, D = wo.dense_w(Vii, V1i, PL, PG, PR)
    assert np.allclose(D, D.T, atol=1e-12 * np.abs(D).max())  # the symmetry the recursion relies on
    for i in range(nx):
        assert _rel(WR[:, :, i], _blk(D, nm, i, i)) < 1e-12
    for i in range(nx - 1):
        assert _rel(WRi1[:, :, i], _blk(D, nm, i, i + 1)) < 1e-12
    assert np.all(WRi1[:, :, nx - 1] == 0)


@pytest.mark.parametrize("nm,nx", [(2, 2), (3, 4), (6, 7)])
def test_lesser_is_the_neighbour_sum_of_imaginary_off_diagonal_blocks(nm, nx):
    """W^<_i = sum over k = i-1, i, i+1 of Y_ik P^<_k Y_ik^H with Y_ii = W_ii and Y_ik =
    i Im W_ik of the dense W (symmetric inputs); at the first slab the missing left
    neighbour is replaced by (W_21, P_1) and at the last the missing right one by
    (W_{nx-1,nx}, P_nx), as the reference's index clamps do."""
    rng = np.random.default_rng(7 * nm + nx)
    Vii, V1i, PL, PG, PR = _inputs(rng, nm, nx, True)
    WL, WG, WR, _ = wo.rgf_w(Vii, V1i, PL, PG, PR)
    _, _, D = wo.dense_w(Vii, V1i, PL, PG, PR)
    Y = lambda i, j: 1j * _blk(D, nm, i, j).imag
    for P, Wx in ((PL, WL), (PG, WG)):
        for i in range(nx):
            W = _blk(D, nm, i, i)
            want = W @ P[:, :, i] @ W.conj().T
            right = (Y(i, i + 1), P[:, :, i + 1]) if i < nx - 1 else (Y(nx - 2, nx - 1), P[:, :, nx - 1])
            left = (Y(i, i - 1), P[:, :, i - 1]) if i > 0 else (Y(1, 0), P[:, :, 0])
            for Yk, Pk in (right, left):
                want = want + Yk @ Pk @ Yk.conj().T
            assert _rel(Wx[:, :, i], want) < 1e-11, i


def test_transposes_are_the_references_for_hermitian_inputs():
    """With a complex Hermitian V the reference's plain transposes no longer give the
    dense blocks: the oracle keeps them (it restates the reference, not the ideal)."""
    rng = np.random.default_rng(3)
    Vii, V1i, PL, PG, PR = _inputs(rng, 4, 4, False)
    _, _, WR, _ = wo.rgf_w(Vii, V1i, PL, PG, PR)
    _, _, D = wo.dense_w(Vii, V1i, PL, PG, PR)
    assert _rel(WR[:, :, 2], _blk(D, 4, 2, 2)) > 1e-3
    assert _rel(WR[:, :, 3], _blk(D, 4, 3, 3)) > 1e-3


def test_dense_form_identities():
    """w_full: W^r solves (I - V P^r) W^r = V; W^< is anti-Hermitian for anti-Hermitian
    P^<; with P = 0, W^r = V."""
    rng = np.random.default_rng(4)
    nm, nx = 3, 4
    Vii, V1i, PL, PG, PR = _inputs(rng, nm, nx, False)
    WLf, WGf, WRf = wo.dense_w(Vii, V1i, PL, PG, PR)
    V = wo.block2full(Vii, nm, nx, V1i)
    Pr = wo.block2full(PR, nm, nx)
    assert np.allclose((np.eye(nm * nx) - V @ Pr) @ WRf, V, atol=1e-12 * np.abs(V).max())
    assert np.allclose(WLf, -WLf.conj().T, atol=1e-12 * np.abs(WLf).max())
    z = np.zeros_like(PR)
    _, _, W0 = wo.w_full(Vii, V1i, z, z, z)
    for i in range(nx):
        assert np.array_equal(W0[:, :, i], Vii[:, :, i] + 0j)
    # green_rgf_calc_w with P = 0: W^r = V on the diagonal, W_{i,i+1} = V1i_i^T, W^< = 0
    WL0, _, WR0, WRi10 = wo.rgf_w(Vii, V1i, z, z, z)
    assert np.allclose(WR0, Vii) and np.all(WL0 == 0)
    for i in range(nx - 1):
        assert np.allclose(WRi10[:, :, i], V1i[:, :, i].T)
