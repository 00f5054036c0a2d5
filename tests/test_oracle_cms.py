"""Pins of the oracle's green_RGF_CMS restatement (orc_rgf_cms) and of the
decimation with an overlap (orc_sancho_s):

  * S = I reproduces green_RGF_RS in compat mode to rounding (every output but
    the G^> blocks, which CMS takes from its own recursion; z S - H is summed in
    another order than z - H);
  * with a general overlap and uniform couplings, G^R, G^< (diagonal and first
    off-diagonal), cur, tr and tre equal the dense inverse of z S - H - Sigma^R S
    - Sigma_contacts with the lesser source Sigma^< S + contacts;
  * the G^> recursion with the never-filled left-connected G^> (Glp = 0):
    pdens_l = g_l H^H pdens_{l+1} H g_l^H and GGi1_l = g_l^H H^H pdens_{l+1},
    with g_l from a dense left-connected inverse;
  * the decimation with overlap: closed form of a 1D chain with onsite overlap
    and the Dyson fixed point of a matrix lead.
"""
import numpy as np
import pytest

from liteqt_amd import configs
from oracle import pyoracle as orc
from tests.dense_ref import dag, fermi, relerr


def _overlap(nm, nx, seed, strength=0.08):
    rng = np.random.default_rng(seed)
    X = (rng.standard_normal((nm, nm, nx)) + 1j * rng.standard_normal((nm, nm, nx))) / np.sqrt(2 * nm)
    S = np.zeros((nm, nm, nx), complex, order="F")
    for l in range(nx):
        S[:, :, l] = np.eye(nm) + strength * (X[:, :, l] @ dag(X[:, :, l]))
    return S


def _case(nm, nx, seed):
    Hii, H1i = configs.random_device(nm, nx, seed, coupling_sigma=1.0, identical_leads=False)
    H1i = configs.uniform_coupling(H1i)
    sigL, sigG, sigR = configs.scattering_self_energies(nm, nx, seed + 1, E=0.2, mu=0.05, scale=0.05)
    return Hii, H1i, sigL, sigG, sigR


def _leads(p, E, S, Hii, H1i, sigR):
    nx = Hii.shape[2]
    T0, TN = H1i[:, :, 0], H1i[:, :, nx - 1]
    H0 = Hii[:, :, 0] + sigR[:, :, 0] @ S[:, :, 0]
    HN = Hii[:, :, nx - 1] + sigR[:, :, nx - 1] @ S[:, :, nx - 1]
    GL = orc.sancho_s(E, S[:, :, 0], H0, dag(T0), p.eta_lead, p.dec_tol, p.dec_nmax)[0]
    GR = orc.sancho_s(E, S[:, :, nx - 1], HN, TN, p.eta_lead, p.dec_tol, p.dec_nmax)[0]
    return T0 @ GL @ dag(T0), dag(TN) @ GR @ TN


def _dense_overlap(E, eta, S, Hii, H1i, sigL, sigR, SigL, SigR, fL, fR, upto=None):
    """Blocks of G = (z S - H - Sigma^R S - contacts)^-1 and G^< = G (Sigma^< S + contacts) G^H
    for slabs 0..upto (default all; with upto < nx-1 the right contact is absent)."""
    nm, _, nx = Hii.shape
    L = nx if upto is None else upto + 1
    N = nm * L
    A = np.zeros((N, N), complex)
    Sl = np.zeros((N, N), complex)
    z = E + 1j * eta
    for l in range(L):
        s = slice(l * nm, (l + 1) * nm)
        A[s, s] = z * S[:, :, l] - Hii[:, :, l] - sigR[:, :, l] @ S[:, :, l]
        Sl[s, s] = sigL[:, :, l] @ S[:, :, l]
        if l > 0:
            sp = slice((l - 1) * nm, l * nm)
            A[s, sp] = -H1i[:, :, l]
            A[sp, s] = -dag(H1i[:, :, l])
    A[:nm, :nm] -= SigL
    GamL = -(SigL - dag(SigL))
    Sl[:nm, :nm] += GamL * fL
    GamR = -(SigR - dag(SigR))
    if upto is None:
        A[N - nm:, N - nm:] -= SigR
        Sl[N - nm:, N - nm:] += GamR * fR
    G = np.linalg.inv(A)
    Gl = G @ Sl @ dag(G)
    b = lambda M, i, j: M[i * nm:(i + 1) * nm, j * nm:(j + 1) * nm]
    return G, Gl, b, GamL, GamR


def test_cms_with_identity_overlap_is_rs_compat():
    nm, nx = 4, 5
    Hii, H1i, sigL, sigG, sigR = _case(nm, nx, 3)
    H1i = configs.random_device(nm, nx, 9, coupling_sigma=1.0)[1]  # non-uniform: compat indexing exercised
    S = np.zeros((nm, nm, nx), complex, order="F")
    for l in range(nx):
        S[:, :, l] = np.eye(nm)
    p = orc.params(orc.COMPAT)
    temp, mu, E = (300.0, 150.0), (0.1, -0.1), 0.13
    a = orc.rgf_cms(p, temp, E, mu, S, Hii, H1i, sigL, sigG, sigR)
    r = orc.rgf(p, temp, E, mu, Hii, H1i, sigL, sigG, sigR)
    for k in ("ldos", "ndens", "GRi1", "GLi1", "cur"):
        assert np.max(np.abs(a[k] - r[k])) <= 1e-13 * max(1.0, np.max(np.abs(r[k]))), k
    assert abs(a["tr"] - r["tr"]) <= 1e-12 * max(1.0, abs(r["tr"])) and a["diag"] == r["diag"]
    assert abs(a["tre"] - r["tre"]) <= 1e-12 * max(1.0, abs(r["tre"]))
    # the last slab's G^> is the identity value in both
    assert relerr(a["pdens"][:, :, nx - 1], r["pdens"][:, :, nx - 1]) < 1e-12
    # below it CMS follows its own recursion
    assert relerr(a["pdens"][:, :, 0], r["pdens"][:, :, 0]) > 1e-6


def test_cms_matches_dense_overlap_device():
    nm, nx = 3, 5
    Hii, H1i, sigL, sigG, sigR = _case(nm, nx, 11)
    S = _overlap(nm, nx, 12)
    p = orc.params(orc.COMPAT, eta_dev=1e-3, eta_lead=1e-3, dec_tol=1e-14, dec_nmax=200)
    temp, mu, E = (300.0, 300.0), (0.1, -0.1), 0.2
    o = orc.rgf_cms(p, temp, E, mu, S, Hii, H1i, sigL, sigG, sigR)
    SigL, SigR = _leads(p, E, S, Hii, H1i, sigR)
    fL, fR = fermi(E, mu[0], p.boltz, temp[0]), fermi(E, mu[1], p.boltz, temp[0])
    G, Gl, b, GamL, GamR = _dense_overlap(E, p.eta_dev, S, Hii, H1i, sigL, sigR, SigL, SigR, fL, fR)
    for l in range(nx):
        assert relerr(o["ldos"][:, :, l], b(G, l, l)) < 1e-10, l
        assert relerr(o["ndens"][:, :, l], b(Gl, l, l)) < 1e-10, l
    for l in range(nx - 1):
        T = H1i[:, :, l + 1]
        assert relerr(o["GRi1"][:, :, l], b(G, l, l + 1)) < 1e-10, l
        assert relerr(o["GLi1"][:, :, l], b(Gl, l, l + 1)) < 1e-10, l
        cur = np.real(dag(T) @ b(Gl, l + 1, l)) - np.real(T @ b(Gl, l, l + 1))
        assert np.max(np.abs(o["cur"][:, :, l] - cur)) < 1e-10 * max(1.0, np.max(np.abs(cur))), l
    Gg = Gl + G - dag(G)
    tr = -np.real(np.trace(GamR * fR @ b(Gg, nx - 1, nx - 1) - GamR * (fR - 1) @ b(Gl, nx - 1, nx - 1)))
    tre = np.real(np.trace(GamL * fL @ b(Gg, 0, 0) - GamL * (fL - 1) @ b(Gl, 0, 0)))
    assert abs(o["tr"] - tr) < 1e-10 * max(1.0, abs(tr))
    assert abs(o["tre"] - tre) < 1e-10 * max(1.0, abs(tre))


def test_cms_gp_recursion_with_unset_glp():
    nm, nx = 3, 5
    Hii, H1i, sigL, sigG, sigR = _case(nm, nx, 21)
    S = _overlap(nm, nx, 22)
    p = orc.params(orc.COMPAT, eta_dev=1e-3, eta_lead=1e-3, dec_tol=1e-14, dec_nmax=200)
    temp, mu, E = (300.0, 300.0), (0.1, -0.1), -0.1
    o = orc.rgf_cms(p, temp, E, mu, S, Hii, H1i, sigL, sigG, sigR)
    SigL, SigR = _leads(p, E, S, Hii, H1i, sigR)
    fL, fR = fermi(E, mu[0], p.boltz, temp[0]), fermi(E, mu[1], p.boltz, temp[0])
    G, Gl, b, _, _ = _dense_overlap(E, p.eta_dev, S, Hii, H1i, sigL, sigR, SigL, SigR, fL, fR)
    gp = b(Gl, nx - 1, nx - 1) + b(G, nx - 1, nx - 1) - dag(b(G, nx - 1, nx - 1))
    assert relerr(o["pdens"][:, :, nx - 1], gp) < 1e-9
    for l in range(nx - 2, -1, -1):
        Gs, _, bs, _, _ = _dense_overlap(E, p.eta_dev, S, Hii, H1i, sigL, sigR, SigL, SigR, fL, fR, upto=l)
        g = bs(Gs, l, l)  # left-connected g_l
        H = H1i[:, :, l]  # compat indexing (uniform couplings here)
        assert relerr(o["GGi1"][:, :, l], dag(g) @ dag(H) @ gp) < 1e-9, l
        gp = g @ dag(H) @ gp @ H @ dag(g)
        assert relerr(o["pdens"][:, :, l], gp) < 1e-9, l


@pytest.mark.parametrize("E", [-1.8, 0.0, 0.9])
def test_sancho_with_onsite_overlap_closed_form(E):
    """A chain with onsite overlap s: z -> z s in the closed-form surface function."""
    eps, t, s, eta = 0.1, -1.0 + 0.3j, 1.3, 1e-3
    G00 = orc.sancho_s(E, np.array([[s]]), np.array([[eps]]), np.array([[t]]), eta, 1e-100, 100)[0]
    w = (E + 1j * eta) * s - eps
    r = np.sqrt(w * w - 4 * abs(t) ** 2 + 0j)
    g1, g2 = (w - r) / (2 * abs(t) ** 2), (w + r) / (2 * abs(t) ** 2)
    g = g1 if g1.imag < 0 else g2
    assert abs(G00[0, 0] - g) < 1e-10 * abs(g)


def test_sancho_with_overlap_fixed_point():
    n = 4
    rng = np.random.default_rng(5)
    X = (rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))) / np.sqrt(2)
    H00 = (X + dag(X)) / 2
    H10 = 0.4 * (rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n)))
    S00 = _overlap(n, 1, 6, 0.2)[:, :, 0]
    E, eta = 0.3, 1e-3
    G00, _, it, hit, _ = orc.sancho_s(E, S00, H00, H10, eta, 1e-100, 100)
    assert hit == 0
    z = E + 1j * eta
    rhs = np.linalg.inv(z * S00 - H00 - dag(H10) @ G00 @ H10)
    assert relerr(G00, rhs) < 1e-11
