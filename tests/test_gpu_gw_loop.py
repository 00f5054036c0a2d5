"""GPU: the one-dimensional GW loop (liteqt_amd/example_gw.py, green_rgf_solve_gw_1D)
against the same loop composed of the CPU oracles (pyoracle's green_RGF_RS, the GW-step
oracle, the dense W oracle) on a small device, and the example's loop run end to end
at a reduced energy count.  Bar: relative Frobenius error 1e-9 per slab of the running
self-energy and of the last G (the chain G -> P -> W -> Sigma -> G compounds the
rounding of every stage)."""
import os

import numpy as np
import pytest

from oracle import gw_oracle as go
from oracle import pyoracle as orc
from oracle import w_oracle as wo

pytestmark = pytest.mark.gpu
TOL = 1e-9
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _env():
    import torch
    if not torch.cuda.is_available():
        pytest.fail("no HIP device visible")
    from liteqt_amd import example_gw, rgf
    return torch, example_gw, rgf


def _device(nm, nx, seed):
    rng = np.random.default_rng(seed)
    c = lambda *s: rng.standard_normal(s) + 1j * rng.standard_normal(s)
    h = c(nm, nm)
    Hii = np.asfortranarray(np.repeat(((h + h.conj().T) / (2 * np.sqrt(nm)))[:, :, None], nx, axis=2))
    H1i = np.asfortranarray(np.repeat((0.4 * c(nm, nm) / np.sqrt(nm))[:, :, None], nx, axis=2))
    v = rng.standard_normal((nm, nm))
    Vii = np.asfortranarray(np.repeat(((v + v.T) / nm + 0.5 * np.eye(nm))[:, :, None], nx, axis=2) + 0j)
    V1i = np.asfortranarray(np.repeat((0.1 * rng.standard_normal((nm, nm)) / nm)[:, :, None], nx, axis=2) + 0j)
    return Hii, H1i, Vii, V1i


def cpu_loop(Hii, H1i, Vii, V1i, en, temp, mu, ndiag, spindeg, alpha, niter):
    nm, _, nx = Hii.shape
    nE = len(en)
    dE = en[1] - en[0]
    nopmax, dE_P, dE_S = go.reference_params(nE, dE, spindeg)
    po = orc.params(orc.COMPAT)
    z = lambda: np.zeros((nm, nm, nx, nE), complex)
    SL, SG, SR = z(), z(), z()
    for _ in range(niter + 1):
        r = orc.rgf_energies(po, temp, en, mu, Hii, H1i, SL, SG, SR, outputs=("ndens", "pdens", "ldos"))
        GL, GG, GR = r["ndens"], r["pdens"], r["ldos"]
        PL, PG, PR = go.polarization(GL, GG, GR, ndiag, nopmax, dE_P)
        WL, WG, WR = z(), z(), z()
        for e in range(nE):
            WL[..., e], WG[..., e], WR[..., e] = wo.w_full(Vii, V1i, PL[..., e], PG[..., e], PR[..., e])
        SL, SG, SR = go.selfenergy_step(GL, GG, GR, WL, WG, WR, SL, SG, SR, ndiag, nopmax, dE_S, alpha)
    return {"GL": GL, "GR": GR, "SL": SL, "SG": SG, "SR": SR, "WR": WR, "PR": PR}


@pytest.mark.parametrize("ndiag", [1, 31])
def test_gw_loop_matches_the_oracle_loop(ndiag):
    torch, example_gw, rgf = _env()
    nm, nx, nE, niter = 32, 5, 48, 2
    Hii, H1i, Vii, V1i = _device(nm, nx, 11 + ndiag)
    en = np.linspace(-2.0, 2.0, nE)
    temp, mu = (300.0, 300.0), (0.2, -0.2)
    print(f"gw loop nm={nm} nx={nx} nen={nE} ndiag={ndiag} niter={niter}", flush=True)
    st = example_gw.run_loop(Hii, H1i, Vii, V1i, en, temp, mu, ndiag, 2.0, 0.5, niter)
    torch.cuda.synchronize()
    ref = cpu_loop(Hii, H1i, Vii, V1i, en, temp, mu, ndiag, 2.0, 0.5, niter)
    for key in ("SL", "SG", "SR", "GL", "GR", "WR", "PR"):
        got = rgf.from_device_blocks(st[key])
        r = ref[key]
        for ix in range(nx):
            err = np.linalg.norm(got[:, :, ix] - r[:, :, ix])
            scale = np.linalg.norm(r[:, :, ix])
            assert err <= TOL * max(scale, 1e-300), (key, ix, err, scale)


def test_example_gw_loop_end_to_end(tmp_path):
    """examples/cnt_8-0 (fixture tests/golden/cnt_8-0): two iterations of the loop at 100
    energies, every file of the reference written, finite, the iteration table grown by
    one line per iteration, and the running Sigma nonzero after the first iteration."""
    torch, example_gw, rgf = _env()
    from liteqt_amd.writers import parse_fortran_e
    num = lambda f: np.array([parse_fortran_e(x) for ln in open(f) for x in ln.split()])
    ex = os.path.join(ROOT, "tests", "golden", "cnt_8-0")
    res = example_gw.run(ex, str(tmp_path), niter=1, nen=100)
    names = ["gw_ldos", "gw_ndos", "gw_pdos", "gw_PR", "gw_PL", "gw_PG", "gw_WR", "gw_WL", "gw_WG",
             "gw_SigR", "gw_SigL", "gw_SigG"]
    for it in (0, 1):
        for n in names:
            f = tmp_path / f"{n}{it:04d}.dat"
            assert f.exists(), f
            vals = num(f)
            assert vals.size > 0 and np.all(np.isfinite(vals)), f
        for n in ("gw_Jdens", "gw_trR", "gw_trL"):
            assert (tmp_path / f"{n}_{it:04d}.dat").exists()
    lines = open(tmp_path / "gw_Id_iteration.dat").read().splitlines()
    assert len(lines) == 2 and lines[0].split()[0] == "0" and lines[1].split()[0] == "1"
    sig = num(tmp_path / "gw_SigR0000.dat")
    assert np.abs(sig.reshape(-1, 4)[:, 2:]).max() > 0
    torch.cuda.empty_cache()
