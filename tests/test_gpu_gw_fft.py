"""GPU: the FFT form of the GW step (gw_polarization_fft / gw_selfenergy_fft, and the
plain entry points above the direct forms' limits) against the CPU oracle of the direct
sums (oracle/gw_oracle.py, pinned in tests/test_gw_oracle.py) and against the GPU's
direct form on the same inputs.

Bar: relative Frobenius error 1e-10 per slab, the scale being the largest norm among
the step's outputs on that slab (an FFT's rounding is relative to the norms of the
series it correlates, not to each output value); exact zeros outside the band and the
frequency window, as in the direct form."""
import numpy as np
import pytest

from oracle import gw_oracle as orc

pytestmark = pytest.mark.gpu
TOL = 1e-10


def _gw():
    import torch
    if not torch.cuda.is_available():
        pytest.fail("no HIP device visible")
    from liteqt_amd import gw
    return gw


def _rand(rng, shape):
    return rng.standard_normal(shape) + 1j * rng.standard_normal(shape)


def _close(got, ref, tag, refs):
    for ix in range(ref.shape[2]):
        err = np.linalg.norm(got[:, :, ix] - ref[:, :, ix])
        scale = max(np.linalg.norm(r[:, :, ix]) for r in refs)
        assert err <= TOL * scale, (tag, ix, err, scale)


def _nopmax(nen, wide):
    if wide is True:
        return nen // 2 - 1
    if wide is False:
        return nen // 2 - 10
    return wide


# (nm, nx, nen, ndiag, wide, method): N = 64 .. 8192 in one LDS pass, 16384 and 32768
# in the four-step split; 72 rows with a full band spans two 64-row blocks
POL_CASES = [(6, 3, 64, 1, False, "fft"), (5, 2, 301, 9, True, "fft"), (4, 2, 33, 0, 0, "fft"),
             (3, 2, 2, 1, 0, "fft"), (5, 1, 7, 1, True, "fft"), (12, 2, 200, 11, False, "fft"),
             (72, 1, 130, 71, True, "fft"), (72, 1, 130, 40, False, "fft"), (3, 2, 2048, 2, True, "fft"),
             (2, 2, 3100, 1, False, "auto"), (3, 2, 4100, 1, True, "auto"), (2, 1, 8200, 1, False, "auto"),
             (1, 1, 12000, 0, False, "auto")]


@pytest.mark.parametrize("nm,nx,nen,ndiag,wide,method", POL_CASES)
def test_polarization_fft_matches_oracle(nm, nx, nen, ndiag, wide, method):
    import torch
    gw = _gw()
    print(f"polarization fft nm={nm} nx={nx} nen={nen} ndiag={ndiag} method={method}", flush=True)
    rng = np.random.default_rng(nen + 7 * nm)
    G = [_rand(rng, (nm, nm, nx, nen)) for _ in range(3)]
    nopmax = _nopmax(nen, wide)
    dE_P = complex(0.0, -0.013)
    p = gw.params(nm, nx, nen, ndiag, nopmax, dE_P, 0.0, 0.0)
    Gd = [gw.to_device(g) for g in G]
    Pd = [torch.full_like(Gd[0], complex(7.0, -7.0)) for _ in range(3)]  # every entry must be written
    gw.polarization(p, *Gd, *Pd, method=method)
    torch.cuda.synchronize()
    ref = orc.polarization(*G, ndiag, nopmax, dE_P)
    m = orc.band_mask(nm, ndiag)
    win = np.zeros(nen, bool)
    win[np.arange(-nopmax, nopmax + 1) + nen // 2 - 1] = True
    for name, d, r in zip(("P<", "P>", "Pr"), Pd, ref):
        got = gw.to_host(d)
        _close(got, r, name, ref)
        assert np.all(got[~m] == 0), name
        assert np.all(got[..., ~win] == 0), name


SIG_CASES = [(6, 5, 64, 1, False, 0.3, "fft"), (5, 7, 301, 9, True, 0.7, "fft"), (3, 5, 33, 0, 0, 0.5, "fft"),
             (2, 5, 4, 1, 1, 0.5, "fft"), (4, 9, 9, 2, True, 0.0, "fft"), (10, 5, 100, 9, False, 0.5, "fft"),
             (13, 6, 40, 8, True, 0.3, "fft"), (70, 5, 60, 69, True, 0.5, "fft"), (2, 5, 2040, 1, True, 0.25, "fft"),
             (2, 5, 2100, 1, False, 0.5, "auto"), (2, 6, 4100, 1, True, 0.5, "auto"), (1, 5, 8200, 0, False, 1.0, "auto"),
             (1, 5, 12000, 0, True, 0.5, "auto")]


@pytest.mark.parametrize("nm,nx,nen,ndiag,wide,alpha,method", SIG_CASES)
def test_selfenergy_fft_matches_oracle(nm, nx, nen, ndiag, wide, alpha, method):
    import torch
    gw = _gw()
    print(f"selfenergy fft nm={nm} nx={nx} nen={nen} ndiag={ndiag} method={method}", flush=True)
    rng = np.random.default_rng(5 * nen + nm)
    G = [_rand(rng, (nm, nm, nx, nen)) for _ in range(3)]
    W = [_rand(rng, (nm, nm, nx, nen)) for _ in range(3)]
    S = [_rand(rng, (nm, nm, nx, nen)) for _ in range(3)]
    nopmax = _nopmax(nen, wide)
    dE_S = complex(0.0, 0.021)
    p = gw.params(nm, nx, nen, ndiag, nopmax, 0.0, dE_S, alpha)
    Sd = [gw.to_device(s) for s in S]
    gw.selfenergy(p, *[gw.to_device(g) for g in G], *[gw.to_device(w) for w in W], *Sd, method=method)
    torch.cuda.synchronize()
    ref = orc.selfenergy_step(*G, *W, *S, ndiag, nopmax, dE_S, alpha)
    for name, d, r in zip(("S<", "S>", "Sr"), Sd, ref):
        _close(gw.to_host(d), r, name, ref)


def test_fft_and_direct_forms_agree_at_example_size():
    """The reference example's step size (nm = 128, nen = 2000, ndiag = 1, nopmax =
    nen/2 - 10) on 6 slabs: both forms on the GPU over every element."""
    import torch
    gw = _gw()
    nm, nx, nen, ndiag = 128, 6, 2000, 1
    p = gw.params_reference(nm, nx, nen, ndiag, 20.0 / (nen - 1), 2.0, 0.5)
    g = torch.Generator(device="cuda:0").manual_seed(11)

    def rnd():
        return torch.randn((nen, nx, nm, nm), dtype=torch.complex128, device="cuda:0", generator=g)

    G = [rnd() for _ in range(3)]
    Pd = [torch.empty_like(G[0]) for _ in range(3)]
    Pf = [torch.empty_like(G[0]) for _ in range(3)]
    print("example-size polarization: direct and fft", flush=True)
    gw.polarization(p, *G, *Pd)
    gw.polarization(p, *G, *Pf, method="fft")
    torch.cuda.synchronize()
    scale = max(torch.linalg.vector_norm(x).item() for x in Pd)
    lo = nen // 2 - 1 - p.nopmax  # first energy index of the frequency window
    for a, b in zip(Pd, Pf):
        assert torch.linalg.vector_norm(a - b).item() <= TOL * scale
        assert torch.count_nonzero(b[:lo]).item() == 0 and torch.count_nonzero(b[lo + 2 * p.nopmax + 1:]).item() == 0
        assert torch.count_nonzero(b[:, :, 0, 2:]).item() == 0  # column 0 beyond the band
    del Pd, Pf
    W = [rnd() for _ in range(3)]
    S = [rnd() for _ in range(3)]
    Sd = [s.clone() for s in S]
    print("example-size self-energy: direct and fft", flush=True)
    gw.selfenergy(p, *G, *W, *Sd)
    gw.selfenergy(p, *G, *W, *S, method="fft")
    torch.cuda.synchronize()
    scale = max(torch.linalg.vector_norm(x).item() for x in Sd)
    for a, b in zip(Sd, S):
        assert torch.linalg.vector_norm(a - b).item() <= TOL * scale
    del G, W, S, Sd
    torch.cuda.empty_cache()
