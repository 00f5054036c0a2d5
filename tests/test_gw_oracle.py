"""Pins of the GW-step oracle (oracle/gw_oracle.py) by what the mathematics fixes:
correlations and convolutions of Gaussians in closed form (sign of w, the nen/2
frequency offset, which element is transposed, dE), constant inputs (triangle
counts), the correlation Z(i + n/2) = sum_ie X(ie) Y(ie - i) of
src/fft_mod.f90:64-70 computed by FFT (the method of its 'fft' branch: zero padding
to 2n - 1, reversed second series), a literal 1-based transcription of the self-energy loop on a
tiny device, and the mixing / edge-copy limits."""
import math

import numpy as np

from oracle import gw_oracle as gw

SQ = math.sqrt(math.pi / 2.0)


def _gauss(E, c, s):
    return np.exp(-((E - c) ** 2) / s ** 2)


def _grid(nen, lo=-2.0, hi=2.0):
    E = np.linspace(lo, hi, nen)
    return E, E[1] - E[0]


def test_polarization_of_gaussians():
    nen, s = 400, 0.15
    E, dE = _grid(nen)
    nopmax, dE_P, _ = gw.reference_params(nen, dE, spindeg=2.0)
    rng = np.random.default_rng(1)
    G = [rng.standard_normal((2, 2, 1, nen)) + 0j for _ in range(3)]  # real parts must be ignored
    Gl, Gg, Gr = G
    (E1, A1), (E3, A3), (E2, A2), (E4, A4), (E5, A5) = (0.3, 1.3), (-0.4, 0.7), (-0.2, 0.9), (0.1, 1.1), (0.05, 0.6)
    Gl[0, 1, 0] += 1j * A1 * _gauss(E, E1, s)
    Gl[1, 0, 0] += 1j * A3 * _gauss(E, E3, s)
    Gg[1, 0, 0] += 1j * A2 * _gauss(E, E2, s)
    Gg[0, 1, 0] += 1j * A4 * _gauss(E, E4, s)
    Gr[0, 1, 0] += 1j * A5 * _gauss(E, E5, s)
    Pl, Pg, Pr = gw.polarization(Gl, Gg, Gr, 1, nopmax, dE_P)
    nop = np.arange(-nopmax, nopmax + 1)
    w = nop * dE
    iop = nop + nen // 2 - 1

    def corr(c1, c2):  # sum_E g(E - c1) g(E - w - c2) ~ (1/dE) int
        return SQ * s / dE * np.exp(-((c1 - c2 - w) ** 2) / (2 * s ** 2))

    # (i A g)(i B g) = -A B g g
    want_l = dE_P * (-A1 * A2) * corr(E1, E2)
    want_g = dE_P * (-A4 * A3) * corr(E4, E3)
    want_r = dE_P * (A1 * A5 * corr(E1, E5) - A5 * A3 * corr(E5, E3))
    for got, want in ((Pl, want_l), (Pg, want_g), (Pr, want_r)):
        g = got[0, 1, 0, iop]
        assert np.abs(g - want).max() < 1e-12 * np.abs(want).max(), np.abs(g - want).max()
        # nothing outside the frequency window
        mask = np.ones(nen, bool)
        mask[iop] = False
        assert np.all(got[0, 1, 0, mask] == 0)


def test_polarization_of_constants_counts_terms():
    nen, nm = 64, 3
    nopmax = nen // 2 - 10
    dE_P = complex(0.0, -0.25)
    G = np.full((nm, nm, 2, nen), 5.0 + 1j)  # imaginary part 1
    Pl, Pg, Pr = gw.polarization(G, G, G, 1, nopmax, dE_P)
    nop = np.arange(-nopmax, nopmax + 1)
    cnt = nen - np.abs(nop)
    assert np.array_equal(Pl[0, 1, 1, nop + nen // 2 - 1], dE_P * -cnt)
    assert np.array_equal(Pr[0, 1, 1, nop + nen // 2 - 1], dE_P * (cnt - cnt))  # a f - c e = 1 - 1
    assert np.all(Pl[0, 2] == 0) and np.all(Pg[2, 0] == 0)  # |i - j| = 2 > ndiag


def test_polarization_equals_the_fft_correlation():
    """sum_ie X(ie) Y(ie - nop) by FFT: with y reversed, z[q] = sum_p x[p] y[n-1-q+p]
    (0-based cyclic convolution of length 2n - 1 >= the support), so the sum is z[n-1+nop]."""
    rng = np.random.default_rng(2)
    nm, nx, nen, ndiag = 3, 2, 50, 2
    Gl, Gg, Gr = (rng.standard_normal((nm, nm, nx, nen)) + 1j * rng.standard_normal((nm, nm, nx, nen)) for _ in range(3))
    nopmax = nen // 2 - 1
    dE_P = complex(0.0, -0.01)
    Pl, Pg, Pr = gw.polarization(Gl, Gg, Gr, ndiag, nopmax, dE_P)

    nop = np.arange(-nopmax, nopmax + 1)
    iop = nop + nen // 2 - 1

    def corr_fft(x, y):
        n = x.shape[-1]
        z = np.fft.ifft(np.fft.fft(x, 2 * n - 1) * np.fft.fft(y[::-1], 2 * n - 1))
        out = np.zeros(n, complex)
        out[iop] = z[n - 1 + nop]
        return out

    a, b, c = 1j * Gl.imag, 1j * Gg.imag, 1j * Gr.imag
    for ix in range(nx):
        for i in range(nm):
            for j in range(nm):
                zl = corr_fft(a[i, j, ix], b[j, i, ix]) * dE_P
                zr = (corr_fft(a[i, j, ix], np.conj(c[i, j, ix])) + corr_fft(c[i, j, ix], a[j, i, ix])) * dE_P
                if abs(i - j) > ndiag:
                    zl = zr = 0 * zl
                assert np.abs(Pl[i, j, ix] - zl).max() < 1e-12, (i, j, ix)
                assert np.abs(Pr[i, j, ix] - zr).max() < 1e-12, (i, j, ix)


def test_selfenergy_of_gaussians():
    nen, s = 400, 0.12
    E, dE = _grid(nen)
    nopmax, _, dE_S = gw.reference_params(nen, dE)
    nm, nx = 1, 5
    z = np.zeros((nm, nm, nx, nen), complex)
    Gl, Gg, Gr = z.copy(), z.copy(), z.copy()
    Wl, Wg, Wr = z.copy(), z.copy(), z.copy()
    iop = np.arange(-nopmax, nopmax + 1) + nen // 2 - 1
    w = np.arange(-nopmax, nopmax + 1) * dE
    (cl, Al), (cg, Ag), (cr, Ar) = (0.2, 0.8), (-0.3, 1.2), (0.1, 0.5)
    (ol, Bl), (og, Bg), (orr, Br) = (0.4, 0.7 - 0.2j), (-0.35, 1.1 + 0.3j), (0.15, -0.4 + 0.9j)
    for G, c, A in ((Gl, cl, Al), (Gg, cg, Ag), (Gr, cr, Ar)):
        G[0, 0, 2] = 3.0 + 1j * A * _gauss(E, c, s)  # real part ignored
    for W, o, B in ((Wl, ol, Bl), (Wg, og, Bg), (Wr, orr, Br)):
        W[0, 0, 2, iop] = B * _gauss(w, o, s)
    Sl, Sg, Sr = gw.selfenergy_new(Gl, Gg, Gr, Wl, Wg, Wr, 0, nopmax, dE_S)

    def conv(c, o):  # sum_w g(E - w - c) g(w - o) ~ (1/dE) int
        return SQ * s / dE * np.exp(-((E - c - o) ** 2) / (2 * s ** 2))

    want_l = dE_S * 1j * Al * Bl * conv(cl, ol)
    want_g = dE_S * 1j * Ag * Bg * conv(cg, og)
    want_r = dE_S * 1j * (Al * Br * conv(cl, orr) + Ar * Bl * conv(cr, ol) + Ar * Br * conv(cr, orr))
    assert np.abs(Sl[0, 0, 2] - want_l).max() < 1e-12 * np.abs(want_l).max()
    assert np.abs(Sg[0, 0, 2] - want_g).max() < 1e-12 * np.abs(want_g).max()
    assert np.abs(Sr[0, 0, 2].real - want_r.real).max() < 1e-12 * np.abs(want_r).max()
    assert np.array_equal(Sr[0, 0, 2].imag, (Sg[0, 0, 2] - Sl[0, 0, 2]).imag / 2.0)


def test_selfenergy_literal_loop_on_a_tiny_device():
    """Literal 1-based transcription of src/green_rgf.f90:3240-3277 (tiny sizes):
    includes the index conditions (ie = 1 empty, the last energy of G never used)."""
    rng = np.random.default_rng(3)
    nm, nx, nen, ndiag = 3, 5, 26, 1
    c = lambda: rng.standard_normal((nm, nm, nx, nen)) + 1j * rng.standard_normal((nm, nm, nx, nen))
    Gl, Gg, Gr, Wl, Wg, Wr = c(), c(), c(), c(), c(), c()
    for nopmax in (nen // 2 - 10, nen // 2 - 1):
        dE_S = complex(0.0, 0.03)
        got = gw.selfenergy_new(Gl, Gg, Gr, Wl, Wg, Wr, ndiag, nopmax, dE_S)
        gl, gg, gr = (1j * G.imag for G in (Gl, Gg, Gr))
        Sl = np.zeros_like(Gl)
        Sg = np.zeros_like(Gl)
        Sr = np.zeros_like(Gl)
        for ix in range(1, nx + 1):
            for i in range(1, nm + 1):
                lo, hi = max(i - ndiag, 1), min(nm, i + ndiag)
                for ie in range(1, nen + 1):
                    for nop in range(-nopmax, nopmax + 1):
                        if ie > max(nop, 1) and ie < nen + nop:
                            iop = nop + nen // 2
                            for j in range(lo, hi + 1):
                                I, J, X, K, O = i - 1, j - 1, ix - 1, ie - nop - 1, iop - 1
                                Sl[I, J, X, ie - 1] += gl[I, J, X, K] * Wl[I, J, X, O]
                                Sg[I, J, X, ie - 1] += gg[I, J, X, K] * Wg[I, J, X, O]
                                Sr[I, J, X, ie - 1] += (gl[I, J, X, K] * Wr[I, J, X, O] + gr[I, J, X, K] * Wl[I, J, X, O]
                                                       + gr[I, J, X, K] * Wr[I, J, X, O])
        Sl, Sg, Sr = Sl * dE_S, Sg * dE_S, Sr * dE_S
        Sr = Sr.real + 1j * ((Sg - Sl).imag / 2.0)
        for a, b in zip(got, (Sl, Sg, Sr)):
            assert np.abs(a - b).max() < 1e-13
        assert np.all(got[0][..., 0] == 0)


def test_mixing_and_edge_copy_limits():
    rng = np.random.default_rng(4)
    S_gw = rng.standard_normal((2, 2, 7, 5)) + 1j * rng.standard_normal((2, 2, 7, 5))
    S_new = rng.standard_normal((2, 2, 7, 5)) + 1j * rng.standard_normal((2, 2, 7, 5))
    keep = gw.mix_and_copy_edges(S_gw, S_new, 0.0)
    take = gw.mix_and_copy_edges(S_gw, S_new, 1.0)
    for out, src in ((keep, S_gw), (take, S_new)):  # alpha = 1: S_gw + (S_new - S_gw), one rounding
        for ix in (0, 1, 2):
            assert np.abs(out[:, :, ix] - src[:, :, 2]).max() < 1e-15
        for ix in (4, 5, 6):
            assert np.abs(out[:, :, ix] - src[:, :, 4]).max() < 1e-15
        assert np.abs(out[:, :, 3] - src[:, :, 3]).max() < 1e-15
    assert np.array_equal(keep[:, :, 0], S_gw[:, :, 2]) and np.array_equal(keep[:, :, 6], S_gw[:, :, 4])
    half = gw.mix_and_copy_edges(S_gw, S_new, 0.5)
    assert np.allclose(half[:, :, 3], 0.5 * (S_gw[:, :, 3] + S_new[:, :, 3]), rtol=0, atol=1e-15)
