"""Pins of the block-tridiagonal-Sigma reverse pass (tests/adjoint_ref_od.py):
central finite differences of the oracle (orc_rgf_offdiag_fixed_leads, contacts
held fixed) for dJ/dU_l on every slab and directional dJ/dH1i_l, and agreement
with tests/adjoint_ref.py when the off-diagonal blocks are absent."""
import numpy as np
import pytest

from liteqt_amd import configs
from oracle import pyoracle as orc
from tests import adjoint_ref, adjoint_ref_od
from tests.dense_ref import fermi, lead_self_energies


def _setup(nm, nx, seed, scale):
    Hii, H1i = configs.random_device(nm, nx, seed, coupling_sigma=1.0, identical_leads=False)
    blocks = configs.scattering_blocks(nm, nx, seed + 1, scale=scale)
    return Hii, H1i, blocks


def _J(p, E, temp, mu, Hii, H1i, sig, leadL, leadR, w_tr, w_n):
    o = orc.rgf_offdiag_fixed_leads(p, temp, E, mu, Hii, H1i, *sig, leadL_H00=leadL, leadR_H00=leadR)
    return w_tr * o["tr"] + sum(w_n[l] * np.imag(np.trace(o["ndens"][:, :, l])) for l in range(Hii.shape[2]))


@pytest.mark.parametrize("nm,nx,scale", [(3, 5, 0.05), (4, 4, 0.2)])
def test_offdiag_adjoint_matches_oracle_fd(nm, nx, scale):
    E = 0.12
    Hii, H1i, blocks = _setup(nm, nx, 9 + nm, scale)
    p = orc.params(orc.PHYSICAL, eta_dev=1e-3, eta_lead=1e-3, dec_tol=1e-14, dec_nmax=200)
    temp, mu = (300.0, 250.0), (0.2, -0.1)
    sig = configs.offdiag_self_energies(blocks, fermi(E, 0.0, p.boltz, 300.0))
    sigL, sigR, R_lo, R_up, L_up = sig
    rng = np.random.default_rng(4)
    w_tr, w_n = 0.7, rng.standard_normal(nx)
    leadL = Hii[:, :, 0] + sigR[:, :, 0]
    leadR = Hii[:, :, nx - 1] + sigR[:, :, nx - 1]
    SigL, SigR = lead_self_energies(orc.sancho, E, Hii, H1i, sigR, p.eta_lead, p.dec_tol, p.dec_nmax)
    fL, fR = fermi(E, mu[0], p.boltz, temp[0]), fermi(E, mu[1], p.boltz, temp[1])
    J, dU, Hb = adjoint_ref_od.adjoint(E, p.eta_dev, Hii, H1i, sigL, sigR, R_lo, R_up, L_up, SigL, SigR, fL, fR,
                                       w_tr, w_n)
    J0 = _J(p, E, temp, mu, Hii, H1i, sig, leadL, leadR, w_tr, w_n)
    assert abs(J - J0) < 1e-10 * max(1, abs(J0))
    h = 1e-5
    for l in range(nx):
        Hp, Hm = Hii.copy(order="F"), Hii.copy(order="F")
        Hp[:, :, l] += h * np.eye(nm)
        Hm[:, :, l] -= h * np.eye(nm)
        fd = (_J(p, E, temp, mu, Hp, H1i, sig, leadL, leadR, w_tr, w_n)
              - _J(p, E, temp, mu, Hm, H1i, sig, leadL, leadR, w_tr, w_n)) / (2 * h)
        assert abs(fd - dU[l]) < 1e-6 * max(1, abs(fd)), (l, fd, dU[l])
    for l in range(1, nx - 1):
        D = rng.standard_normal((nm, nm)) + 1j * rng.standard_normal((nm, nm))
        Cp, Cm = H1i.copy(order="F"), H1i.copy(order="F")
        Cp[:, :, l] += h * D
        Cm[:, :, l] -= h * D
        fd = (_J(p, E, temp, mu, Hii, Cp, sig, leadL, leadR, w_tr, w_n)
              - _J(p, E, temp, mu, Hii, Cm, sig, leadL, leadR, w_tr, w_n)) / (2 * h)
        ad = np.real(np.vdot(Hb[:, :, l], D))
        assert abs(fd - ad) < 1e-6 * max(1, abs(fd)), (l, fd, ad)


def test_offdiag_adjoint_without_offdiag_blocks_is_the_diagonal_adjoint():
    nm, nx, E = 4, 5, -0.2
    Hii, H1i, blocks = _setup(nm, nx, 31, 0.05)
    p = orc.params(orc.PHYSICAL, eta_dev=1e-3, eta_lead=1e-3, dec_tol=1e-14, dec_nmax=200)
    sigL, sigR, _, _, _ = configs.offdiag_self_energies(blocks, 0.4)
    SigL, SigR = lead_self_energies(orc.sancho, E, Hii, H1i, sigR, p.eta_lead, p.dec_tol, p.dec_nmax)
    w_n = np.linspace(-1, 1, nx)
    a = adjoint_ref_od.adjoint(E, p.eta_dev, Hii, H1i, sigL, sigR, None, None, None, SigL, SigR, 0.3, 0.1, 0.9, w_n)
    b = adjoint_ref.adjoint(E, p.eta_dev, Hii, H1i, sigL, sigR, SigL, SigR, 0.3, 0.1, 0.9, w_n)
    assert abs(a[0] - b[0]) < 1e-11 * max(1, abs(b[0]))
    assert np.allclose(a[1], b[1], rtol=1e-9, atol=1e-10)
    assert np.allclose(a[2], b[2], rtol=1e-9, atol=1e-10)
