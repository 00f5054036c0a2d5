"""GPU unit tests of the engine's building blocks (complex GEMM on FP64 MFMA,
Gauss-Jordan inverse) against numpy/LAPACK on the same inputs."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

SIZES = [32, 64, 96, 128, 160, 192, 256]
# blocks above 256: several inverse panel rows per thread, 128 x 64 (n % 128 == 0) or 64 x 64 tiles
BIG = [320, 384, 448, 512]


def _env():
    import torch
    if not torch.cuda.is_available():
        pytest.fail("no HIP device visible")
    from liteqt_amd import rgf
    return torch, rgf


def _colmajor(torch, M):
    # numpy (nb, n, n) matrices -> device tensor whose memory is column-major per matrix
    return torch.from_numpy(np.ascontiguousarray(np.swapaxes(M, -1, -2))).cuda()


def _back(t):
    return np.swapaxes(t.cpu().numpy(), -1, -2)


@pytest.mark.parametrize("n", SIZES + [320, 512])
@pytest.mark.parametrize("opA,opB", [(0, 0), (0, 1), (1, 0), (1, 1)])
def test_zgemm_batched(n, opA, opB):
    torch, rgf = _env()
    rng = np.random.default_rng(n + 10 * opA + opB)
    nb = 3
    A = rng.standard_normal((nb, n, n)) + 1j * rng.standard_normal((nb, n, n))
    B = rng.standard_normal((nb, n, n)) + 1j * rng.standard_normal((nb, n, n))
    op = lambda X, o: np.conj(np.swapaxes(X, -1, -2)) if o else X
    ref = op(A, opA) @ op(B, opB)
    for rowmajor in (False, True):
        C = rgf.zgemm_batched(_colmajor(torch, A), _colmajor(torch, B), opA, opB, rowmajor=rowmajor)
        torch.cuda.synchronize()
        got = C.cpu().numpy() if rowmajor else _back(C)
        err = np.linalg.norm(got - ref) / np.linalg.norm(ref)
        assert err < 1e-14, (n, opA, opB, rowmajor, err)


@pytest.mark.parametrize("n", SIZES + BIG)
def test_zinv_batched(n):
    torch, rgf = _env()
    rng = np.random.default_rng(n)
    nb = 4
    A = rng.standard_normal((nb, n, n)) + 1j * rng.standard_normal((nb, n, n))
    A[1, 0, 0] = 0.0  # forces a row exchange in the first column
    Ainv, info = rgf.zinv_batched(_colmajor(torch, A))
    torch.cuda.synchronize()
    got = _back(Ainv)
    assert np.all(info.cpu().numpy() == 0)
    for b in range(nb):
        ref = np.linalg.inv(A[b])
        err = np.linalg.norm(got[b] - ref) / np.linalg.norm(ref)
        assert err < 1e-12 * max(1.0, np.linalg.cond(A[b]) / 1e3), (n, b, err)


def test_zinv_singular_zero_and_info():
    torch, rgf = _env()
    n = 32
    A = np.zeros((1, n, n), np.complex128)
    A[0, :, :] = np.eye(n)
    A[0, 5, :] = 0.0  # zero row -> no pivot for some column
    Ainv, info = rgf.zinv_batched(_colmajor(torch, A))
    torch.cuda.synchronize()
    assert info.cpu().numpy()[0] > 0
    assert np.all(Ainv.cpu().numpy() == 0)


@pytest.mark.parametrize("n", [96, 128, 192, 256, 384, 512])
def test_zinv_singular_after_strip_updates(n):
    """A zero row is never a usable pivot, so the elimination runs through every
    strip update and reports the last column; the output is zeroed."""
    torch, rgf = _env()
    rng = np.random.default_rng(7 + n)
    A = rng.standard_normal((2, n, n)) + 1j * rng.standard_normal((2, n, n))
    A[0, n // 2 + 3, :] = 0.0
    Ainv, info = rgf.zinv_batched(_colmajor(torch, A))
    torch.cuda.synchronize()
    info = info.cpu().numpy()
    assert info[0] == n and info[1] == 0
    assert np.all(Ainv[0].cpu().numpy() == 0)
    ref = np.linalg.inv(A[1])
    assert np.linalg.norm(_back(Ainv)[1] - ref) / np.linalg.norm(ref) < 1e-11


@pytest.mark.parametrize("n", [320, 512])
def test_zinv_pivots_from_rows_above_the_workgroup(n):
    """Blocks above 256 rows hold rows tid and tid + 256 per thread: a matrix whose
    column maxima sit in the upper rows (a row-reversed diagonally dominant
    matrix) takes every pivot from the second row set, and a zero leading block
    forces exchanges across it."""
    torch, rgf = _env()
    rng = np.random.default_rng(n + 1)
    A = rng.standard_normal((2, n, n)) + 1j * rng.standard_normal((2, n, n))
    A[0] += 8.0 * np.sqrt(n) * np.eye(n)
    A[0] = A[0][::-1].copy()  # the dominant entry of column j is in row n - 1 - j
    A[1, :16, :16] = 0.0
    Ainv, info = rgf.zinv_batched(_colmajor(torch, A))
    torch.cuda.synchronize()
    assert np.all(info.cpu().numpy() == 0)
    got = _back(Ainv)
    for b in range(2):
        ref = np.linalg.inv(A[b])
        assert np.linalg.norm(got[b] - ref) / np.linalg.norm(ref) < 1e-11, (n, b)
