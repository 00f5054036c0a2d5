"""GPU adjoint RGF against (1) the numpy reverse pass of tests/adjoint_ref.py
(itself pinned to finite differences of the oracle), (2) central finite
differences of the CPU oracle with h = 1e-6 eV on random slabs, (3)
torch.autograd.gradcheck through liteqt_amd.autograd.RGFObjective."""
import numpy as np
import pytest

from liteqt_amd import configs
from oracle import pyoracle as orc
from tests import adjoint_ref
from tests.dense_ref import fermi, lead_self_energies

pytestmark = pytest.mark.gpu


def _env():
    import torch
    if not torch.cuda.is_available():
        pytest.fail("no HIP device visible")
    from liteqt_amd import rgf
    return torch, rgf


def _case(nm, nx, seed, scatter=True):
    Hii, H1i = configs.random_device(nm, nx, seed, identical_leads=False)
    if scatter:
        sL, _, sR = configs.scattering_self_energies(nm, nx, seed + 1, E=0.1, mu=0.0, scale=0.02)
    else:
        sL = sR = None
    return Hii, H1i, sL, sR


def _fixed_leads(rgf, Hii, sR, dev):
    """lead_H00 of the unperturbed device (edge slabs plus sigR there): contacts held fixed."""
    nx = Hii.shape[2]
    L = Hii[:, :, 0] + (sR[:, :, 0] if sR is not None else 0)
    R = Hii[:, :, nx - 1] + (sR[:, :, nx - 1] if sR is not None else 0)
    return rgf.to_device_blocks(np.stack([L, R], axis=2), dev)


@pytest.mark.parametrize("nm,nx,scatter,leads", [(32, 4, True, "follow"), (64, 5, False, "follow"),
                                                 (32, 2, True, "follow"), (128, 3, True, "follow"),
                                                 (32, 4, True, "fixed"), (64, 5, False, "fixed")])
def test_adjoint_kernel_matches_reference_reverse_pass(nm, nx, scatter, leads):
    """Contacts that repeat the edge slabs (lead_H00 absent: their derivatives through
    the decimation included, tests/adjoint_ref.adjoint_with_leads) and contacts held
    fixed (lead_H00 given, tests/adjoint_ref.adjoint)."""
    torch, rgf = _env()
    dev = torch.device("cuda:0")
    Hii, H1i, sL, sR = _case(nm, nx, 40 + nm + nx, scatter)
    kw = dict(eta_dev=1e-4, eta_lead=1e-4, dec_tol=1e-12)
    p = rgf.params(rgf.MODE_PHYSICAL, **kw)
    po = orc.params(orc.PHYSICAL, **kw)
    E = np.array([-0.3, 0.05, 0.4])
    temp, mu = (300.0, 250.0), (0.1, -0.1)
    rng = np.random.default_rng(nm + nx)
    w_tr = rng.standard_normal(len(E))
    w_n = rng.standard_normal((len(E), nx))
    Jg, dUg, dHg = rgf.adjoint_batched(
        p, torch.tensor(E, device=dev), temp, mu, rgf.to_device_blocks(Hii, dev), rgf.to_device_blocks(H1i, dev),
        sigL=None if sL is None else rgf.to_device_blocks(sL, dev),
        sigR=None if sR is None else rgf.to_device_blocks(sR, dev),
        w_tr=torch.tensor(w_tr, device=dev), w_n=torch.tensor(w_n, device=dev),
        lead_H00=_fixed_leads(rgf, Hii, sR, dev) if leads == "fixed" else None)
    torch.cuda.synchronize()
    Jg, dUg = Jg.cpu().numpy(), dUg.cpu().numpy()
    dHg = rgf.from_device_blocks(dHg)
    dH_ref = np.zeros((nm, nm, nx), complex)
    for e, En in enumerate(E):
        fL, fR = fermi(En, mu[0], po.boltz, temp[0]), fermi(En, mu[1], po.boltz, temp[1])
        if leads == "fixed":
            SigL, SigR = lead_self_energies(orc.sancho, En, Hii, H1i, sR, po.eta_lead, po.dec_tol, po.dec_nmax)
            J, dU, dH = adjoint_ref.adjoint(En, po.eta_dev, Hii, H1i, sL, sR, SigL, SigR, fL, fR, w_tr[e], w_n[e])
        else:
            J, dU, dH = adjoint_ref.adjoint_with_leads(En, po.eta_dev, po.eta_lead, Hii, H1i, sL, sR, fL, fR, w_tr[e],
                                                       w_n[e], orc.sancho, po.dec_tol, po.dec_nmax)
        assert abs(Jg[e] - J) <= 1e-10 * max(1.0, abs(J)), (e, Jg[e], J)
        assert np.linalg.norm(dUg[e] - dU) <= 1e-9 * max(1.0, np.linalg.norm(dU)), (e, dUg[e], dU)
        dH_ref += dH
    assert np.linalg.norm(dHg - dH_ref) <= 1e-9 * max(1.0, np.linalg.norm(dH_ref))


def test_adjoint_kernel_matches_oracle_central_fd():
    """dJ/dU_l and directional dJ/dH1i_l vs central differences (h = 1e-6) of the
    CPU oracle's own device, whose contacts repeat the edge slabs: every slab's
    potential (the edge ones move the leads too) and every coupling, the contact
    couplings included."""
    torch, rgf = _env()
    dev = torch.device("cuda:0")
    nm, nx = 32, 12
    Hii, H1i, sL, sR = _case(nm, nx, 77, scatter=False)
    kw = dict(eta_dev=1e-4, eta_lead=1e-4, dec_tol=1e-12)
    p, po = rgf.params(rgf.MODE_PHYSICAL, **kw), orc.params(orc.PHYSICAL, **kw)
    E0 = 0.21
    temp, mu = (300.0, 300.0), (0.1, -0.1)
    w_n = np.linspace(0.5, 1.5, nx)
    J, dU, dH = rgf.adjoint_batched(p, torch.tensor([E0], dtype=torch.float64, device=dev), temp, mu, rgf.to_device_blocks(Hii, dev),
                                    rgf.to_device_blocks(H1i, dev),
                                    w_tr=torch.tensor([1.0], dtype=torch.float64, device=dev),
                                    w_n=torch.tensor(w_n[None, :], device=dev))
    torch.cuda.synchronize()
    dU = dU.cpu().numpy()[0]
    dH = rgf.from_device_blocks(dH)

    def Jo(Hm, Cm):
        o = orc.rgf(po, temp, E0, mu, Hm, Cm, offdiag=False)
        return o["tr"] + sum(w_n[l] * np.imag(np.trace(o["ndens"][:, :, l])) for l in range(nx))

    h = 1e-6
    rng = np.random.default_rng(5)
    for l in [0, nx - 1] + list(rng.choice(np.arange(1, nx - 1), 6, replace=False)):
        Hp, Hm = Hii.copy(order="F"), Hii.copy(order="F")
        Hp[:, :, l] += h * np.eye(nm)
        Hm[:, :, l] -= h * np.eye(nm)
        fd = (Jo(Hp, H1i) - Jo(Hm, H1i)) / (2 * h)
        assert abs(fd - dU[l]) <= 1e-6 * max(1.0, abs(fd)), (l, fd, dU[l])
    def fd_coupling(l, D, step):
        Cp, Cm = H1i.copy(order="F"), H1i.copy(order="F")
        Cp[:, :, l] += step * D
        Cm[:, :, l] -= step * D
        return (Jo(Hii, Cp) - Jo(Hii, Cm)) / (2 * step)

    for l in [0, nx - 1] + list(rng.choice(np.arange(1, nx - 1), 3, replace=False)):
        D = rng.standard_normal((nm, nm)) + 1j * rng.standard_normal((nm, nm))
        # Richardson: J is strongly curved along the contact couplings at eta = 1e-4 (the
        # O(h^2) term of one central difference reaches 1e-5 relative at h = 1e-6)
        fd = (4.0 * fd_coupling(l, D, h) - fd_coupling(l, D, 2 * h)) / 3.0
        ad = np.real(np.vdot(dH[:, :, l], D))
        assert abs(fd - ad) <= 1e-6 * max(1.0, abs(fd)), (l, fd, ad)


def test_autograd_function_gradcheck():
    torch, rgf = _env()
    from liteqt_amd.autograd import RGFObjective
    dev = torch.device("cuda:0")
    nm, nx = 32, 4
    Hii, H1i, sL, sR = _case(nm, nx, 91, scatter=True)
    p = rgf.params(rgf.MODE_PHYSICAL, eta_dev=1e-3, eta_lead=1e-3, dec_tol=1e-12)
    E = torch.tensor([0.0, 0.3], dtype=torch.float64, device=dev)
    Hd, Cd = rgf.to_device_blocks(Hii, dev), rgf.to_device_blocks(H1i, dev)
    sLd, sRd = rgf.to_device_blocks(sL, dev), rgf.to_device_blocks(sR, dev)
    w_tr = torch.tensor([1.0, 0.5], dtype=torch.float64, device=dev)
    w_n = torch.ones(2, nx, dtype=torch.float64, device=dev) * 0.3
    U = torch.zeros(nx, dtype=torch.float64, device=dev, requires_grad=True)
    Tint = Cd[1:nx - 1].clone().requires_grad_(True)
    f = lambda U, Tint: RGFObjective.apply(U, Tint, p, E, (300.0, 300.0), (0.1, -0.1), Hd, Cd, sLd, sRd, w_tr, w_n)
    assert torch.autograd.gradcheck(f, (U, Tint), eps=1e-6, atol=1e-6, rtol=1e-5, fast_mode=True)
    # leads that repeat the edge slabs: every potential and every coupling, the contact ones included
    Tall = Cd.clone().requires_grad_(True)
    g = lambda U, T: RGFObjective.apply(U, T, p, E, (300.0, 300.0), (0.1, -0.1), Hd, Cd, sLd, sRd, w_tr, w_n, None,
                                        True)
    assert torch.autograd.gradcheck(g, (U, Tall), eps=1e-6, atol=1e-6, rtol=1e-5, fast_mode=True)


def _od_case(nm, nx, seed, E0):
    Hii, H1i = configs.random_device(nm, nx, seed, identical_leads=False)
    blocks = configs.scattering_blocks(nm, nx, seed + 1, scale=0.02)
    po = orc.params(orc.PHYSICAL, eta_dev=1e-4, eta_lead=1e-4, dec_tol=1e-12)
    return Hii, H1i, configs.offdiag_self_energies(blocks, fermi(E0, 0.0, po.boltz, 300.0))


@pytest.mark.parametrize("nm,nx", [(32, 5), (64, 4)])
def test_offdiag_adjoint_kernel_matches_reference_reverse_pass(nm, nx):
    """Block-tridiagonal Sigma: the kernel against tests/adjoint_ref_od.py."""
    from tests import adjoint_ref_od
    torch, rgf = _env()
    dev = torch.device("cuda:0")
    E = np.array([-0.2, 0.15])
    Hii, H1i, sig = _od_case(nm, nx, 60 + nm, 0.0)
    kw = dict(eta_dev=1e-4, eta_lead=1e-4, dec_tol=1e-12)
    p, po = rgf.params(rgf.MODE_PHYSICAL, **kw), orc.params(orc.PHYSICAL, **kw)
    temp, mu = (300.0, 250.0), (0.1, -0.1)
    rng = np.random.default_rng(nm)
    w_tr, w_n = rng.standard_normal(len(E)), rng.standard_normal((len(E), nx))
    names = ("sigL", "sigR", "sigR_lo", "sigR_up", "sigL_up")
    sd = {k: rgf.to_device_blocks(v, dev) for k, v in zip(names, sig)}
    Jg, dUg, dHg = rgf.adjoint_batched(p, torch.tensor(E, dtype=torch.float64, device=dev), temp, mu,
                                       rgf.to_device_blocks(Hii, dev), rgf.to_device_blocks(H1i, dev),
                                       w_tr=torch.tensor(w_tr, device=dev), w_n=torch.tensor(w_n, device=dev),
                                       lead_H00=_fixed_leads(rgf, Hii, sig[1], dev), **sd)
    torch.cuda.synchronize()
    Jg, dUg, dHg = Jg.cpu().numpy(), dUg.cpu().numpy(), rgf.from_device_blocks(dHg)
    dH_ref = np.zeros((nm, nm, nx), complex)
    for e, En in enumerate(E):
        SigL, SigR = lead_self_energies(orc.sancho, En, Hii, H1i, sig[1], po.eta_lead, po.dec_tol, po.dec_nmax)
        fL, fR = fermi(En, mu[0], po.boltz, temp[0]), fermi(En, mu[1], po.boltz, temp[1])
        J, dU, dH = adjoint_ref_od.adjoint(En, po.eta_dev, Hii, H1i, *sig, SigL, SigR, fL, fR, w_tr[e], w_n[e])
        assert abs(Jg[e] - J) <= 1e-10 * max(1.0, abs(J)), (e, Jg[e], J)
        assert np.linalg.norm(dUg[e] - dU) <= 1e-9 * max(1.0, np.linalg.norm(dU)), (e, dUg[e], dU)
        dH_ref += dH
    assert np.linalg.norm(dHg - dH_ref) <= 1e-9 * max(1.0, np.linalg.norm(dH_ref))


def test_offdiag_adjoint_density_gradient_matches_oracle_fd():
    """dn_l/dU_m (the density adjoint) with block-tridiagonal Sigma against
    central differences (h = 1e-6) of the oracle with contacts and Sigma fixed."""
    torch, rgf = _env()
    dev = torch.device("cuda:0")
    nm, nx, E0 = 32, 8, 0.07
    Hii, H1i, sig = _od_case(nm, nx, 99, E0)
    kw = dict(eta_dev=1e-4, eta_lead=1e-4, dec_tol=1e-12)
    p, po = rgf.params(rgf.MODE_PHYSICAL, **kw), orc.params(orc.PHYSICAL, **kw)
    temp, mu = (300.0, 300.0), (0.1, -0.1)
    w_n = np.zeros(nx)
    w_n[3] = 1.0  # J = Im Tr G<_33: the density of slab 3
    names = ("sigL", "sigR", "sigR_lo", "sigR_up", "sigL_up")
    sd = {k: rgf.to_device_blocks(v, dev) for k, v in zip(names, sig)}
    _, dU, _ = rgf.adjoint_batched(p, torch.tensor([E0], dtype=torch.float64, device=dev), temp, mu,
                                   rgf.to_device_blocks(Hii, dev), rgf.to_device_blocks(H1i, dev),
                                   w_n=torch.tensor(w_n[None, :], device=dev),
                                   lead_H00=_fixed_leads(rgf, Hii, sig[1], dev), **sd)
    torch.cuda.synchronize()
    dU = dU.cpu().numpy()[0]

    def Jo(Hm):
        o = orc.rgf_offdiag_fixed_leads(po, temp, E0, mu, Hm, H1i, *sig,
                                        leadL_H00=Hii[:, :, 0] + sig[1][:, :, 0],
                                        leadR_H00=Hii[:, :, nx - 1] + sig[1][:, :, nx - 1])
        return np.imag(np.trace(o["ndens"][:, :, 3]))

    h = 1e-6
    for l in range(nx):
        Hp, Hm = Hii.copy(order="F"), Hii.copy(order="F")
        Hp[:, :, l] += h * np.eye(nm)
        Hm[:, :, l] -= h * np.eye(nm)
        fd = (Jo(Hp) - Jo(Hm)) / (2 * h)
        assert abs(fd - dU[l]) <= 1e-6 * max(1.0, abs(fd)), (l, fd, dU[l])


@pytest.mark.parametrize("offdiag", [False, True])
def test_adjoint_recompute_slot_is_bit_identical(offdiag, monkeypatch):
    """The small workspace slot (products of the sweeps recomputed in the reverse
    passes, the path taken under memory pressure) gives the same bits as the
    default slot that keeps them."""
    torch, rgf = _env()
    dev = torch.device("cuda:0")
    nm, nx = 64, 6
    E = np.array([-0.25, 0.05, 0.3])
    if offdiag:
        Hii, H1i, sig = _od_case(nm, nx, 91, 0.0)
        names = ("sigL", "sigR", "sigR_lo", "sigR_up", "sigL_up")
        sd = {k: rgf.to_device_blocks(v, dev) for k, v in zip(names, sig)}
    else:
        Hii, H1i, sL, sR = _case(nm, nx, 93, scatter=True)
        sd = {"sigL": rgf.to_device_blocks(sL, dev), "sigR": rgf.to_device_blocks(sR, dev)}
    p = rgf.params(rgf.MODE_PHYSICAL, eta_dev=1e-4, eta_lead=1e-4, dec_tol=1e-12)
    rng = np.random.default_rng(5)
    w_tr = torch.tensor(rng.standard_normal(len(E)), device=dev)
    w_n = torch.tensor(rng.standard_normal((len(E), nx)), device=dev)
    args = (p, torch.tensor(E, dtype=torch.float64, device=dev), (300.0, 250.0), (0.1, -0.1),
            rgf.to_device_blocks(Hii, dev), rgf.to_device_blocks(H1i, dev))
    outs = []
    for recompute in ("0", "1"):
        monkeypatch.setenv("LITEQT_ADJ_RECOMPUTE", recompute)
        J, dU, dH = rgf.adjoint_batched(*args, w_tr=w_tr, w_n=w_n, **sd)
        torch.cuda.synchronize()
        outs.append((J.clone(), dU.clone(), dH.clone()))
    for a, b in zip(*outs):
        assert torch.equal(a, b)


@pytest.mark.parametrize("offdiag,nx,K", [(False, 11, 3), (True, 9, 4), (False, 7, 2), (True, 13, 5)])
def test_adjoint_checkpointed_forward_state_is_bit_identical(offdiag, nx, K):
    """The forward state (g, g^<, S) kept only at checkpoints every K slabs, refilled a
    segment at a time by the backward sweep and both reverse passes: J, dU and dH equal
    the keep-every-slab adjoint bit for bit."""
    torch, rgf = _env()
    dev = torch.device("cuda:0")
    nm = 64
    E = np.array([-0.25, 0.05, 0.3, 0.45])
    print(f"adjoint checkpoints offdiag={offdiag} nx={nx} K={K}", flush=True)
    if offdiag:
        Hii, H1i, sig = _od_case(nm, nx, 95, 0.0)
        names = ("sigL", "sigR", "sigR_lo", "sigR_up", "sigL_up")
        sd = {k: rgf.to_device_blocks(v, dev) for k, v in zip(names, sig)}
    else:
        Hii, H1i, sL, sR = _case(nm, nx, 97, scatter=True)
        sd = {"sigL": rgf.to_device_blocks(sL, dev), "sigR": rgf.to_device_blocks(sR, dev)}
    p = rgf.params(rgf.MODE_PHYSICAL, eta_dev=1e-4, eta_lead=1e-4, dec_tol=1e-12)
    rng = np.random.default_rng(7)
    w_tr = torch.tensor(rng.standard_normal(len(E)), device=dev)
    w_n = torch.tensor(rng.standard_normal((len(E), nx)), device=dev)
    args = (p, torch.tensor(E, dtype=torch.float64, device=dev), (300.0, 250.0), (0.1, -0.1),
            rgf.to_device_blocks(Hii, dev), rgf.to_device_blocks(H1i, dev))
    outs = []
    for ck in (-1, K):
        rgf.lib().rgf_release()  # the slot size follows ckpt
        J, dU, dH = rgf.adjoint_batched(*args, w_tr=w_tr, w_n=w_n, ckpt=ck, **sd)
        torch.cuda.synchronize()
        outs.append((J.clone(), dU.clone(), dH.clone()))
    for a, b in zip(*outs):
        assert torch.equal(a, b)
