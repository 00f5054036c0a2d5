"""CPU pins of oracle/integrals.py (calc_n_electron, src/green_rgf.f90:3920-3945,
and the SCBA drivers' current, src/green_rgf.f90:3189, 788)."""
import numpy as np

from liteqt_amd import configs
from oracle import integrals as itg
from oracle import pyoracle as orc


def _rand(rng, *shape):
    return rng.standard_normal(shape) + 1j * rng.standard_normal(shape)


def test_vectorised_equals_literal_loops():
    rng = np.random.default_rng(3)
    nm, nx, nen, nk = 3, 4, 9, 2
    Gl, Gg = _rand(rng, nm, nm, nx, nen, nk), _rand(rng, nm, nm, nx, nen, nk)
    E = np.sort(rng.uniform(-1, 1, nen))
    midgap = np.array([-2.0, E[4], 0.1, 2.0])  # all electrons / a tie at E[4] / split / all holes
    n1, p1 = itg.calc_n_electron(Gl, Gg, E, midgap)
    n2, p2 = itg.calc_n_electron_loops(Gl, Gg, E, midgap)
    assert np.allclose(n1, n2, rtol=1e-13, atol=1e-15)
    assert np.allclose(p1, p2, rtol=1e-13, atol=1e-15)
    # every energy goes to exactly one of the two sums; the tie E == midgap is a hole
    assert p1[0] == 0.0 and n1[3] == 0.0
    above = (E > midgap[1])
    assert above.sum() == nen - 5


def test_closed_form_constant_diagonal_ignores_offdiagonal_and_real_parts():
    rng = np.random.default_rng(4)
    nm, nx, nen, nk = 4, 3, 11, 3
    a, b = 0.37, -0.81
    Gl, Gg = _rand(rng, nm, nm, nx, nen, nk), _rand(rng, nm, nm, nx, nen, nk)
    j = np.arange(nm)
    Gl[j, j] = rng.standard_normal((nm, nx, nen, nk)) + 1j * a
    Gg[j, j] = rng.standard_normal((nm, nx, nen, nk)) + 1j * b
    E = np.linspace(-1.0, 1.0, nen)
    midgap = np.array([-0.35, 0.0, 0.95])
    n, p = itg.calc_n_electron(Gl, Gg, E, midgap)
    dE = E[1] - E[0]
    for ix in range(nx):
        cnt = int(np.sum(E > midgap[ix]))
        # dE/(2 pi) * (2 pi / nk) * nk = dE: k-points and 2 pi cancel
        assert np.isclose(n[ix], a * nm * cnt * dE, rtol=1e-12)
        assert np.isclose(p[ix], -b * nm * (nen - cnt) * dE, rtol=1e-12)


def test_current_formula_and_landauer_quantum():
    E = np.linspace(-1.0, 1.0, 201)
    tr = np.ones_like(E)
    # one transmitting mode over a window of width W carries W e^2/h per spin, e/h in amperes per eV
    I = itg.current(tr, E, spindeg=2.0, nk=1)
    e_over_h = itg.E0_REF * itg.E0_REF / (itg.TPI_REF * itg.HBAR_REF)  # A per eV
    assert np.isclose(I, 2.0 * e_over_h * 201 * (E[1] - E[0]), rtol=1e-14)
    assert np.isclose(itg.current(np.stack([tr, tr]), E, 2.0, nk=2), I, rtol=1e-14)


def test_sum_rule_through_the_oracle_solve():
    """The whole chain: a periodic device (every slab and coupling the same, so every
    state is a band state of the leads) with eta_dev = 0 is in equilibrium
    G^< = i f A and G^> = -i (1 - f) A exactly; with mu above the band and every
    energy above the midgap, nelec counts int Tr A dE / 2pi = nm states per slab,
    and with mu below the band and every energy below it pelec does the same.
    calc_n_electron weights an energy by dE/(2 pi) * dK with dK = 2 pi / nk, so
    at nk = 1 both equal 2 pi nm.  The lead eta smooths the band-edge
    singularities; grid and tails leave ~1e-5 (measured 2e-5 of nm)."""
    nm, nx = 2, 3
    rng = np.random.default_rng(9)
    A = rng.standard_normal((nm, nm)) + 1j * rng.standard_normal((nm, nm))
    h0 = (A + A.conj().T) / 4.0
    t0 = (rng.standard_normal((nm, nm)) + 1j * rng.standard_normal((nm, nm))) * 0.5
    Hii = np.asfortranarray(np.repeat(h0[:, :, None], nx, axis=2))
    H1i = np.asfortranarray(np.repeat(t0[:, :, None], nx, axis=2))
    E = np.linspace(-5.0, 5.0, 20001)
    p = orc.params(orc.PHYSICAL, eta_dev=0.0, eta_lead=2e-3, dec_tol=1e-12, dec_nmax=200)
    for mu, mg, kind in ((30.0, -1e9, "n"), (-30.0, 1e9, "p")):
        r = orc.rgf_energies(p, (300.0, 300.0), E, (mu, mu), Hii, H1i, outputs=("ndens", "pdens"), nthreads=8)
        n, pe = itg.calc_n_electron(r["ndens"], r["pdens"], E, np.full(nx, mg))
        got = n if kind == "n" else pe
        other = pe if kind == "n" else n
        assert np.all(np.abs(got / (2 * np.pi) - nm) < 1e-3 * nm), (kind, got / (2 * np.pi))
        assert np.all(other == 0.0)
