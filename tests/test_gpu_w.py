"""GPU: the screened interaction W (rgf_w_solve_batched, green_rgf_calc_w) against the
CPU oracle (oracle/w_oracle.py, pinned in tests/test_w_oracle.py).  Bar: relative
Frobenius error 1e-10 per block of W^r, W^<, W^> and W^r_{i,i+1}, for symmetric inputs
(where the recursion gives the blocks of (I - V P^r)^-1 V) and Hermitian ones (where
it follows the reference's transposes)."""
import numpy as np
import pytest

from oracle import w_oracle as wo

pytestmark = pytest.mark.gpu
TOL = 1e-10


def _rgf():
    import torch
    if not torch.cuda.is_available():
        pytest.fail("no HIP device visible")
    from liteqt_amd import rgf
    return rgf


def _rand(rng, *shape):
    return rng.standard_normal(shape) + 1j * rng.standard_normal(shape)


def _inputs(rng, nm, nx, nE, symmetric, scale=0.3):
    if symmetric:
        Vii = rng.standard_normal((nm, nm, nx))
        Vii = (Vii + Vii.transpose(1, 0, 2)) / np.sqrt(nm)
        V1i = rng.standard_normal((nm, nm, nx)) / np.sqrt(nm)
    else:
        Vii = _rand(rng, nm, nm, nx)
        Vii = (Vii + Vii.conj().transpose(1, 0, 2)) / np.sqrt(nm)
        V1i = _rand(rng, nm, nm, nx) / np.sqrt(nm)
    P = []
    for _ in range(3):
        x = _rand(rng, nm, nm, nx, nE) * scale / nm
        P.append(x)
    PL = P[0] - P[0].conj().transpose(1, 0, 2, 3)
    PG = P[1] - P[1].conj().transpose(1, 0, 2, 3)
    PR = P[2] + P[2].transpose(1, 0, 2, 3) if symmetric else P[2]
    return Vii + 0j, V1i + 0j, PL, PG, PR


@pytest.mark.parametrize("nm,nx,nE,sym", [(32, 2, 3, True), (64, 5, 4, False), (40, 4, 3, True), (96, 3, 2, False),
                                          (128, 3, 2, True), (256, 2, 1, False), (300, 2, 1, True), (17, 7, 5, False)])
def test_w_matches_oracle(nm, nx, nE, sym):
    import torch
    rgf = _rgf()
    print(f"W-RGF nm={nm} nx={nx} nE={nE} symmetric={sym}", flush=True)
    rng = np.random.default_rng(nm + 100 * nx + nE)
    Vii, V1i, PL, PG, PR = _inputs(rng, nm, nx, nE, sym)
    dev = torch.device("cuda:0")
    tb = lambda a: rgf.to_device_blocks(a, dev)
    out = rgf.w_solve_batched(tb(Vii), tb(V1i), tb(PL), tb(PG), tb(PR), want_i1=True)
    torch.cuda.synchronize()
    got = {k: rgf.from_device_blocks(v) for k, v in out.items() if k != "info"}
    assert int(out["info"].abs().sum().item()) == 0
    for e in range(nE):
        WL, WG, WR, WRi1 = wo.rgf_w(Vii, V1i, PL[..., e], PG[..., e], PR[..., e])
        for name, ref in (("WL", WL), ("WG", WG), ("WR", WR), ("WRi1", WRi1)):
            for i in range(nx):
                g, r = got[name][:, :, i, e], ref[:, :, i]
                if name == "WRi1" and i == nx - 1:
                    assert np.all(g == 0)
                    continue
                err = np.linalg.norm(g - r)
                assert err <= TOL * np.linalg.norm(r), (name, e, i, err / np.linalg.norm(r))


def test_w_symmetric_inputs_give_the_dense_blocks():
    """nm = 64, nx = 6: the GPU's W^r against the blocks of the dense (I - V P^r)^-1 V."""
    import torch
    rgf = _rgf()
    rng = np.random.default_rng(5)
    nm, nx, nE = 64, 6, 2
    Vii, V1i, PL, PG, PR = _inputs(rng, nm, nx, nE, True)
    dev = torch.device("cuda:0")
    tb = lambda a: rgf.to_device_blocks(a, dev)
    out = rgf.w_solve_batched(tb(Vii), tb(V1i), tb(PL), tb(PG), tb(PR))
    WR = rgf.from_device_blocks(out["WR"])
    for e in range(nE):
        _, _, D = wo.dense_w(Vii, V1i, PL[..., e], PG[..., e], PR[..., e])
        for i in range(nx):
            r = D[i * nm:(i + 1) * nm, i * nm:(i + 1) * nm]
            assert np.linalg.norm(WR[:, :, i, e] - r) <= TOL * np.linalg.norm(r)


@pytest.mark.parametrize("nm,nx,nE", [(8, 4, 3), (20, 3, 5), (64, 6, 2), (128, 4, 2)])
def test_w_full_matches_dense_oracle(nm, nx, nE):
    """green_calc_w_full on the GPU (gw_w_full: dense LU solve per energy) against the
    oracle's dense numpy solve, Hermitian V and random P, 1e-10 per block."""
    import torch
    from liteqt_amd import gw
    _rgf()
    print(f"W full nm={nm} nx={nx} nE={nE}", flush=True)
    rng = np.random.default_rng(3 * nm + nx)
    Vii, V1i, PL, PG, PR = _inputs(rng, nm, nx, nE, False)
    dev = torch.device("cuda:0")
    from liteqt_amd import rgf
    tb = lambda a: rgf.to_device_blocks(a, dev)
    WL, WG, WR = gw.w_full(tb(Vii), tb(V1i), tb(PL), tb(PG), tb(PR))
    torch.cuda.synchronize()
    got = [rgf.from_device_blocks(x) for x in (WL, WG, WR)]
    for e in range(nE):
        ref = wo.w_full(Vii, V1i, PL[..., e], PG[..., e], PR[..., e])
        for name, g, r in zip(("WL", "WG", "WR"), got, ref):
            for i in range(nx):
                err = np.linalg.norm(g[:, :, i, e] - r[:, :, i])
                assert err <= TOL * np.linalg.norm(r[:, :, i]), (name, e, i, err / np.linalg.norm(r[:, :, i]))


def test_w_full_symmetric_inputs_equal_the_block_recursion():
    """For symmetric V and P^r the dense form (green_calc_w_full) and the block recursion
    (green_rgf_calc_w) give the same W^r diagonal blocks."""
    import torch
    from liteqt_amd import gw
    rgf = _rgf()
    rng = np.random.default_rng(8)
    nm, nx, nE = 32, 5, 3
    Vii, V1i, PL, PG, PR = _inputs(rng, nm, nx, nE, True)
    dev = torch.device("cuda:0")
    tb = lambda a: rgf.to_device_blocks(a, dev)
    _, _, WRf = gw.w_full(tb(Vii), tb(V1i), tb(PL), tb(PG), tb(PR))
    WRr = rgf.w_solve_batched(tb(Vii), tb(V1i), tb(PL), tb(PG), tb(PR))["WR"]
    torch.cuda.synchronize()
    assert torch.linalg.vector_norm(WRf - WRr).item() <= TOL * torch.linalg.vector_norm(WRf).item()


@pytest.mark.parametrize("nm,nx,nE,chunk", [(16, 5, 3, 256), (32, 6, 5, 2), (64, 3, 2, 256), (128, 4, 2, 1),
                                            (48, 7, 2, 256), (16, 1, 2, 256), (32, 2, 3, 256)])
def test_w_blocks_match_dense_oracle(nm, nx, nE, chunk):
    """The super-block form of the dense W (gw.w_blocks: the block-tridiagonal-Sigma solve
    on two-slab super-blocks, odd nx padded) against the oracle's dense solve and the
    GPU's dense LU form (gw.w_full): Hermitian V, general P^r, anti-Hermitian P^<, P^>,
    1e-10 per block; energies in chunks smaller than nE included."""
    import torch
    from liteqt_amd import gw
    rgf = _rgf()
    print(f"W blocks nm={nm} nx={nx} nE={nE} chunk={chunk}", flush=True)
    rng = np.random.default_rng(5 * nm + nx)
    Vii, V1i, PL, PG, PR = _inputs(rng, nm, nx, nE, False)
    dev = torch.device("cuda:0")
    tb = lambda a: rgf.to_device_blocks(a, dev)
    Wb = gw.w_blocks(tb(Vii), tb(V1i), tb(PL), tb(PG), tb(PR), chunk=chunk)
    Wf = gw.w_full(tb(Vii), tb(V1i), tb(PL), tb(PG), tb(PR))
    torch.cuda.synchronize()
    got = [rgf.from_device_blocks(x) for x in Wb]
    for e in range(nE):
        ref = wo.w_full(Vii, V1i, PL[..., e], PG[..., e], PR[..., e])
        for name, g, r in zip(("WL", "WG", "WR"), got, ref):
            for i in range(nx):
                err = np.linalg.norm(g[:, :, i, e] - r[:, :, i])
                assert err <= TOL * np.linalg.norm(r[:, :, i]), (name, e, i, err / np.linalg.norm(r[:, :, i]))
    for name, a, b in zip(("WL", "WG", "WR"), Wb, Wf):
        assert torch.linalg.vector_norm(a - b).item() <= TOL * torch.linalg.vector_norm(b).item(), name
