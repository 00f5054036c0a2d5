"""GPU: the overlap solve (green_RGF_CMS semantics, rgf_cms / rgf_cms_ex and the
batched device entry with rgf_batch.Sii) against the CPU oracle orc_rgf_cms
(pinned in tests/test_oracle_cms.py)."""
import numpy as np
import pytest

from liteqt_amd import configs
from oracle import pyoracle as orc
from tests import parity
from tests.test_oracle_cms import _overlap

pytestmark = [pytest.mark.gpu,
              pytest.mark.skip(reason="the overlap GPU solve is switched off: its first run at nm=64 ended in a GPU "
                                      "memory fault with no cause found yet (the engine returns RGF_EUNSUP)")]


def _rgf():
    import torch
    if not torch.cuda.is_available():
        pytest.fail("no HIP device visible")
    from liteqt_amd import rgf
    return rgf


def _check(g, r, H1i, tag):
    """parity.check with the G^> scales taken from the identity value (CMS reports
    its recursion's G^> in pdens; that block is checked on its own)."""
    nx = r["ldos"].shape[2]
    scales = dict(r)
    gp_id = np.zeros_like(r["pdens"])
    for l in range(nx):
        G = r["ldos"][:, :, l]
        gp_id[:, :, l] = r["ndens"][:, :, l] + G - np.conj(G.T)
    scales["pdens"] = gp_id
    amp = parity.amp_gp(scales, nx - 1)
    gg = {k: v for k, v in g.items() if k not in ("pdens", "GGi1")}
    parity.check(gg, scales, H1i, "compat", tag)
    for l in range(nx):
        err = np.linalg.norm(g["pdens"][:, :, l] - r["pdens"][:, :, l])
        assert err <= parity.TOL * amp * np.linalg.norm(r["pdens"][:, :, l]), (tag, "pdens", l)
    for l in range(nx - 1):
        err = np.linalg.norm(g["GGi1"][:, :, l] - r["GGi1"][:, :, l])
        assert err <= parity.TOL * amp * np.linalg.norm(r["GGi1"][:, :, l]), (tag, "GGi1", l)


@pytest.mark.parametrize("nm,nx", [(32, 4), (64, 3), (128, 3), (20, 3), (1, 2)])
def test_cms_matches_oracle(nm, nx):
    rgf = _rgf()
    Hii, H1i = configs.random_device(nm, nx, 70 + nm, identical_leads=False)
    sL, sG, sR = configs.scattering_self_energies(nm, nx, 71 + nm, E=0.1, mu=0.0, scale=0.02)
    S = _overlap(nm, nx, 72 + nm)
    temp, mu, E = (300.0, 250.0), (0.1, -0.1), 0.13
    p = rgf.params(rgf.MODE_COMPAT, eta_dev=1e-4, eta_lead=1e-4, dec_tol=1e-12)
    po = orc.params(orc.COMPAT, eta_dev=1e-4, eta_lead=1e-4, dec_tol=1e-12)
    g = rgf.rgf_cms(temp, E, mu, S, Hii, H1i, sL, sG, sR, p=p)
    r = orc.rgf_cms(po, temp, E, mu, S, Hii, H1i, sL, sG, sR)
    _check(g, r, H1i, f"nm={nm} nx={nx}")


def test_cms_default_entry_is_the_reference_configuration():
    rgf = _rgf()
    nm, nx = 32, 3
    Hii, H1i = configs.random_device(nm, nx, 5)
    S = _overlap(nm, nx, 6)
    temp, mu, E = (300.0, 300.0), (0.1, -0.1), 0.31
    g = rgf.rgf_cms(temp, E, mu, S, Hii, H1i)
    r = orc.rgf_cms(orc.params(orc.COMPAT), temp, E, mu, S, Hii, H1i)
    _check(g, r, H1i, "reference parameters")


def test_cms_device_batched_with_per_energy_sigma():
    rgf = _rgf()
    import torch
    dev = torch.device("cuda:0")
    nm, nx, nE = 64, 4, 5
    Hii, H1i = configs.random_device(nm, nx, 15)
    S = _overlap(nm, nx, 16)
    E = np.linspace(-0.3, 0.3, nE)
    sig = [configs.scattering_self_energies(nm, nx, 20 + e, E=E[e], mu=0.0, scale=0.02) for e in range(nE)]
    sL = np.stack([s[0] for s in sig], axis=-1)
    sR = np.stack([s[2] for s in sig], axis=-1)
    temp, mu = (300.0, 300.0), (0.1, -0.1)
    p = rgf.params(rgf.MODE_COMPAT, eta_dev=1e-4, eta_lead=1e-4, dec_tol=1e-12)
    po = orc.params(orc.COMPAT, eta_dev=1e-4, eta_lead=1e-4, dec_tol=1e-12)
    out = {k: torch.zeros(nE, nx, nm, nm, dtype=torch.complex128, device=dev) for k in ("ldos", "ndens", "pdens")}
    out["tr"] = torch.zeros(nE, dtype=torch.float64, device=dev)
    out["dens_p"] = torch.zeros(nE, nx, dtype=torch.float64, device=dev)
    b = rgf._batch(torch, torch.tensor(E, device=dev), temp, mu, rgf.to_device_blocks(Hii, dev),
                   rgf.to_device_blocks(H1i, dev), rgf.to_device_blocks(sL, dev), rgf.to_device_blocks(sR, dev), 3)
    Sd = rgf.to_device_blocks(S, dev)
    b.Sii = Sd.data_ptr()
    rgf._out_batch(torch, b, out, dev)
    import ctypes
    rgf._check(rgf.lib().rgf_solve_batched(ctypes.byref(p), ctypes.byref(b),
                                           ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)))
    torch.cuda.synchronize()
    h = {k: rgf.from_device_blocks(out[k]) for k in ("ldos", "ndens", "pdens")}
    for e in range(nE):
        r = orc.rgf_cms(po, temp, E[e], mu, S, Hii, H1i, sL[..., e], None, sR[..., e])
        for l in range(nx):
            assert np.linalg.norm(h["ldos"][..., l, e] - r["ldos"][..., l]) <= 1e-10 * np.linalg.norm(r["ldos"][..., l])
            assert np.linalg.norm(h["ndens"][..., l, e] - r["ndens"][..., l]) <= 1e-10 * np.linalg.norm(r["ndens"][..., l])
        dp = -np.imag(np.trace(r["pdens"], axis1=0, axis2=1))
        assert np.allclose(out["dens_p"][e].cpu().numpy(), dp, rtol=1e-8, atol=1e-12)
        assert abs(out["tr"][e].item() - r["tr"]) <= 1e-9 * max(1.0, abs(r["tr"]))


def test_cms_rejects_physical_mode_and_offdiagonal_sigma():
    rgf = _rgf()
    nm, nx = 32, 2
    Hii, H1i = configs.random_device(nm, nx, 1)
    S = _overlap(nm, nx, 2)
    with pytest.raises(rgf.RgfError):
        rgf.rgf_cms((300.0, 300.0), 0.1, (0.0, 0.0), S, Hii, H1i, p=rgf.params(rgf.MODE_PHYSICAL))
