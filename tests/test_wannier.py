"""Wannier input (liteqt_amd/wannier.py) on the reference's example examples/cnt_8-0
(files copied unchanged to tests/golden/cnt_8-0): the ham_dat reader against the file's
own header and lines, the slab-block builder against exact band folding, and the
device assembly's potential rule."""
import os

import numpy as np

from liteqt_amd import wannier

HERE = os.path.dirname(os.path.abspath(__file__))
EX = os.path.join(HERE, "golden", "cnt_8-0")


def _hermitian(w):
    """Copy of w whose hoppings satisfy H(-R) = H(R)^H exactly (the file's values are
    rounded to 1e-6, so they do only to that precision)."""
    Hr = w.Hr.copy()
    for x in range(1, w.xmax + 1):
        Hr[:, :, -x - w.xmin, :] = np.conj(np.swapaxes(Hr[:, :, x - w.xmin, :], 0, 1))
    h0 = Hr[:, :, -w.xmin, :]
    Hr[:, :, -w.xmin, :] = 0.5 * (h0 + np.conj(np.swapaxes(h0, 0, 1)))
    return wannier.W90(w.nvb, w.spin_deg, w.cell, Hr, w.xmin, w.ymin, w.centers)


def test_ham_dat_reader_on_the_reference_example():
    w = wannier.load_ham_dat(os.path.join(EX, "ham_dat"))
    assert (w.nvb, w.spin_deg, w.nb) == (23, 2, 32)
    assert (w.xmin, w.xmax, w.ymin, w.ymax) == (-10, 10, 0, 0)
    assert np.allclose(w.cell[:, 0], [4.2761526, 0, 0]) and np.allclose(w.cell[:, 2], [0, 0, 24.6881123])
    # first and last hopping lines of the file
    assert w.Hr[0, 0, 0, 0] == -0.000150 and w.Hr[31, 31, 20, 0] == -0.000013
    assert np.count_nonzero(w.Hr) > 0.9 * 21504
    assert w.centers.shape == (3, 32) and abs(w.centers[2].mean()) < 1e-12
    # rounded to six decimals, the file is Hermitian to that precision
    for x in range(0, 11):
        d = w.Hr[:, :, x - w.xmin, 0] - np.conj(w.Hr[:, :, -x - w.xmin, 0].T)
        assert np.abs(d).max() < 2e-6


def _unit_cell_bloch(w, k, rmax):
    return sum(w.Hr[:, :, x - w.xmin, 0] * np.exp(-1j * k * x) for x in range(-rmax, rmax + 1))


def _truncated(w, rmax):
    Hr = w.Hr.copy()
    for x in range(w.xmin, w.xmax + 1):
        if abs(x) > rmax:
            Hr[:, :, x - w.xmin, :] = 0
    return wannier.W90(w.nvb, w.spin_deg, w.cell, Hr, w.xmin, w.ymin, w.centers)


def test_slab_blocks_fold_the_unit_cell_bands_exactly():
    """Slab of ns cells, hoppings up to ns cells (every cell pair at such an offset lies
    in one slab or in neighbouring slabs): eig(Hii + H1i e^{-iK} + H1i^H e^{iK}) is the
    union over m of eig(h((K + 2 pi m)/ns)), h(k) = sum_{|R| <= ns} H(R) e^{-ikR}."""
    w0 = _hermitian(wannier.load_ham_dat(os.path.join(EX, "ham_dat")))
    for ns in (4, 2, 3):
        w = _truncated(w0, ns)
        Hii, H1i = wannier.mat_def(w, 0.0, 0.0, ns)
        assert np.abs(Hii - np.conj(Hii.T)).max() < 1e-14
        for K in (0.0, 0.7, -2.1, np.pi):
            Hs = Hii + H1i * np.exp(-1j * K) + np.conj(H1i.T) * np.exp(1j * K)
            got = np.sort(np.linalg.eigvalsh(Hs))
            ref = np.sort(np.concatenate([np.linalg.eigvalsh(_unit_cell_bloch(w, (K + 2 * np.pi * m) / ns, ns))
                                          for m in range(ns)]))
            assert np.abs(got - ref).max() < 1e-10 * np.abs(ref).max(), (ns, K)


def test_slab_blocks_take_every_offset_of_neighbouring_slabs():
    """Block (i, k) of Hii is H(i - k) and of H1i is H(i - k + ns), i.e. the coupling to
    the previous slab keeps offsets 1 .. 2ns-1 (and drops what lies two slabs apart)."""
    w = wannier.load_ham_dat(os.path.join(EX, "ham_dat"))
    ns, nb = 4, w.nb
    Hii, H1i = wannier.mat_def(w, 0.0, 0.0, ns)
    H = lambda x: w.Hr[:, :, x - w.xmin, 0]
    for i in range(ns):
        for k in range(ns):
            bi, bk = slice(i * nb, (i + 1) * nb), slice(k * nb, (k + 1) * nb)
            assert np.array_equal(Hii[bi, bk], H(i - k))
            assert np.array_equal(H1i[bi, bk], H(i - k + ns))


def test_device_blocks_follow_the_driver():
    """Identical slabs; the potential of slab l on the first nb diagonal entries only
    (src/main.f90:524-529 loops over nb)."""
    w = wannier.load_ham_dat(os.path.join(EX, "ham_dat"))
    pot = wannier.read_pot(os.path.join(EX, "pot_dat"), 18)
    assert pot.shape == (18,) and abs(pot[0] - 0.305410421179578) < 1e-15
    Hii, H1i = wannier.device_blocks(w, 4, 18, pot, potscale=2.0)
    h00, h10 = wannier.mat_def(w, 0.0, 0.0, 4)
    assert Hii.shape == (128, 128, 18) and Hii.flags.f_contiguous
    for l in (0, 7, 17):
        d = Hii[:, :, l] - h00
        assert np.allclose(np.diag(d)[:32], 2.0 * pot[l]) and np.abs(np.diag(d)[32:]).max() == 0
        assert np.abs(d - np.diag(np.diag(d))).max() == 0
        assert np.array_equal(H1i[:, :, l], h10)


def test_oracle_transmission_of_the_pristine_tube_counts_modes():
    """The builder's blocks through the CPU oracle: with Hermitian hoppings up to ns
    cells, no potential and f_L = 1, f_R = 0, T(E) is the number of propagating modes
    of the unit-cell bands (7, 2 and 0 modes at these energies; 0 in a gap)."""
    from oracle import pyoracle as orc
    from tests.test_gpu_wannier import _modes
    ns = 4
    w = _truncated(_hermitian(wannier.load_ham_dat(os.path.join(EX, "ham_dat"))), ns)
    Hii, H1i = wannier.device_blocks(w, ns, 4)
    po = orc.params(orc.PHYSICAL, eta_dev=0.0, eta_lead=1e-7, dec_tol=1e-14, dec_nmax=200)
    for E, expect in ((-9.3, 7), (-4.4, 2), (1.2, 0)):
        M = _modes(w, E, ns, nk=1001)
        assert M == expect
        o = orc.rgf(po, (300.0, 300.0), E, (100.0, -100.0), Hii, H1i, offdiag=False)
        assert abs(o["tr"] - M) < 1e-6 and abs(o["tre"] - M) < 1e-6, (E, o["tr"], M)


def test_coulomb_blocks_of_the_example():
    """w90_bare_coulomb_blocks on examples/cnt_8-0 (ns = 4, epsilon = 1, r0 = 3 A, only
    diag): real symmetric blocks; the self-interaction e/(4 pi eps0 r0) and the distant
    pairs' V |r| / tanh(|r|/r0) equal the Coulomb constant (14.40 eV A) to the reference's
    rounded constants (e = 1.6e-19 C, eps0 = 8.85e-12: 14.387); V scales as 1/epsilon;
    the 0.1 eV self-term without ldiag; and V1i takes bare_coulomb(i-k) as the reference
    writes it (so here, with every offset of a 4-cell slab in range, V1i = Vii)."""
    w = wannier.load_ham_dat(os.path.join(EX, "ham_dat"))
    Vii, V1i = wannier.coulomb_blocks(w, 4, 1.0, 3.0, True)
    assert Vii.shape == (128, 128) and np.all(Vii.imag == 0)
    assert np.allclose(Vii, Vii.T, rtol=0, atol=1e-13)
    k = 14.3996  # e^2 / (4 pi eps0) in eV A
    assert abs(Vii[0, 0].real * 3.0 - k) / k < 1e-3
    alpha = w.cell[:, 0]
    bc = wannier.bare_coulomb(w, 4, 0, 1.0, 3.0, True)
    r = 4 * alpha[:, None, None] + w.centers[:, :, None] - w.centers[:, None, :]
    nr = np.sqrt((r * r).sum(0))
    assert np.all(np.abs(bc * nr / np.tanh(nr / 3.0) - k) / k < 1e-3)
    V2, _ = wannier.coulomb_blocks(w, 4, 2.0, 3.0, True)
    assert np.allclose(V2, Vii / 2.0, rtol=1e-15, atol=0)
    Vn, _ = wannier.coulomb_blocks(w, 4, 1.0, 3.0, False)
    assert Vn[0, 0] == 0.1 and np.array_equal(Vn[0, 1], Vii[0, 1])
    assert np.array_equal(V1i, Vii)
    Vd, V1d = wannier.device_coulomb(w, 4, 18, 1.0, 3.0, True)
    assert Vd.shape == (128, 128, 18) and np.array_equal(Vd[:, :, 17], Vii) and np.array_equal(V1d[:, :, 0], V1i)
