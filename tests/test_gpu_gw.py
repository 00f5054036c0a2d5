"""GPU: the GW self-energy step (include/gw.h through liteqt_amd/gw.py) against the CPU
oracle (oracle/gw_oracle.py, pinned in tests/test_gw_oracle.py).  Bar: relative
Frobenius error 1e-10 per slab of every output (the sums are of O(nen) terms without
systematic cancellation), exact zeros outside the band and the frequency window."""
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


def _close(got, ref, tag, refs=None):
    """Per slab err <= TOL * (largest norm among the step's outputs on that slab): P^r
    vanishes identically on the diagonal at nop = 0 (a c - c a per term), where the GPU's
    two FMAs leave rounding noise of that scale."""
    refs = [ref] if refs is None else refs
    for ix in range(ref.shape[2]):
        err = np.linalg.norm(got[:, :, ix] - ref[:, :, ix])
        scale = max(np.linalg.norm(r[:, :, ix]) for r in refs)
        assert err <= TOL * scale, (tag, ix, err)


def _nopmax(nen, wide):
    """nen/2 - 10 (the 1D driver), nen/2 - 1 (the widest window) or a given value."""
    if wide is True:
        return nen // 2 - 1
    if wide is False:
        return nen // 2 - 10
    return wide


@pytest.mark.parametrize("nm,nx,nen,ndiag,wide", [(6, 3, 64, 1, False), (5, 2, 301, 9, True), (4, 2, 1100, 1, False),
                                                  (3, 1, 2048, 2, True), (2, 1, 3072, 1, False),
                                                  (4, 2, 33, 0, 0), (3, 2, 2, 1, 0), (5, 1, 7, 1, True),
                                                  (12, 2, 200, 11, False), (20, 1, 64, 9, True),
                                                  (72, 1, 130, 71, True), (72, 1, 130, 40, False)])
def test_polarization_matches_oracle(nm, nx, nen, ndiag, wide):
    import torch
    gw = _gw()
    print(f"polarization nm={nm} nx={nx} nen={nen} ndiag={ndiag}", flush=True)
    rng = np.random.default_rng(nen + nm)
    G = [_rand(rng, (nm, nm, nx, nen)) for _ in range(3)]
    nopmax = _nopmax(nen, wide)
    dE_P = complex(0.0, -0.013)
    p = gw.params(nm, nx, nen, ndiag, nopmax, dE_P, 0.0, 0.0)
    Gd = [gw.to_device(g) for g in G]
    Pd = [torch.full_like(Gd[0], complex(7.0, -7.0)) for _ in range(3)]  # every entry must be written
    gw.polarization(p, *Gd, *Pd)
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


@pytest.mark.parametrize("nm,nx,nen,ndiag,wide,alpha", [(6, 5, 64, 1, False, 0.3), (5, 7, 301, 9, True, 0.7),
                                                        (3, 6, 1100, 1, False, 1.0), (2, 5, 2040, 1, True, 0.25),
                                                        (3, 5, 33, 0, 0, 0.5), (2, 5, 4, 1, 1, 0.5),
                                                        (4, 9, 9, 2, True, 0.0), (10, 5, 100, 9, False, 0.5),
                                                        (13, 6, 40, 8, True, 0.3)])
def test_selfenergy_matches_oracle(nm, nx, nen, ndiag, wide, alpha):
    gw = _gw()
    print(f"selfenergy nm={nm} nx={nx} nen={nen} ndiag={ndiag}", flush=True)
    import torch
    rng = np.random.default_rng(3 * nen + nm)
    G = [_rand(rng, (nm, nm, nx, nen)) for _ in range(3)]
    W = [_rand(rng, (nm, nm, nx, nen)) for _ in range(3)]
    S = [_rand(rng, (nm, nm, nx, nen)) for _ in range(3)]
    nopmax = _nopmax(nen, wide)
    dE_S = complex(0.0, 0.021)
    p = gw.params(nm, nx, nen, ndiag, nopmax, 0.0, dE_S, alpha)
    Sd = [gw.to_device(s) for s in S]
    gw.selfenergy(p, *[gw.to_device(g) for g in G], *[gw.to_device(w) for w in W], *Sd)
    torch.cuda.synchronize()
    ref = orc.selfenergy_step(*G, *W, *S, ndiag, nopmax, dE_S, alpha)
    for name, d, r in zip(("S<", "S>", "Sr"), Sd, ref):
        _close(gw.to_host(d), r, name, ref)


def test_example_sized_step_on_sampled_elements():
    """The reference example's GW step size (nm = 128, nx = 18, nen = 2000, ndiag = 1,
    nopmax = nen/2 - 10): every band element is computed on the GPU; sampled elements
    (the step is separable per element: P_ij needs the series of ij and ji) against the
    oracle, an out-of-band element's mixing and the edge copies exactly."""
    import torch
    gw = _gw()
    nm, nx, nen, ndiag = 128, 18, 2000, 1
    dE = 20.0 / (nen - 1)
    p = gw.params_reference(nm, nx, nen, ndiag, dE, 2.0, 0.5)
    nopmax, dE_P, dE_S = orc.reference_params(nen, dE, 2.0)
    dE_S = complex(p.dE_S[0], p.dE_S[1])
    assert p.nopmax == nopmax and complex(p.dE_P[0], p.dE_P[1]) == dE_P
    g = torch.Generator(device="cuda:0").manual_seed(5)

    def rnd():
        return torch.randn((nen, nx, nm, nm), dtype=torch.complex128, device="cuda:0", generator=g)

    G = [rnd() for _ in range(3)]
    P = [torch.empty_like(G[0]) for _ in range(3)]
    print("example-size polarization", flush=True)
    gw.polarization(p, *G, *P)
    torch.cuda.synchronize()
    series = lambda t, i, j, ix: t[:, ix, j, i].cpu().numpy()  # (nen,) of element (i, j, slab ix)
    for (i, j, ix) in ((0, 0, 0), (0, 1, 5), (77, 76, 17), (127, 126, 9), (64, 64, 3)):
        sub = [np.zeros((2, 2, 1, nen), complex) for _ in range(3)]
        for s, t in zip(sub, G):
            s[0, 1, 0], s[1, 0, 0], s[0, 0, 0] = series(t, i, j, ix), series(t, j, i, ix), series(t, i, j, ix)
        k = (0, 1) if i != j else (0, 0)
        ref = orc.polarization(*sub, 1, nopmax, dE_P)
        for name, d, r in zip(("P<", "P>", "Pr"), P, ref):
            got = series(d, i, j, ix)
            want = r[k[0], k[1], 0]
            assert np.linalg.norm(got - want) <= TOL * np.linalg.norm(want), (name, i, j, ix)
    assert torch.count_nonzero(P[0][:, 4, 0, 2]).item() == 0  # |i - j| = 2: outside the band
    del P
    W = [rnd() for _ in range(3)]
    S = [rnd() for _ in range(3)]
    S0 = [series(t, 0, 5, 2) for t in S] + [series(t, 40, 41, 7) for t in S]
    print("example-size self-energy", flush=True)
    gw.selfenergy(p, *G, *W, *S)
    torch.cuda.synchronize()
    # out of band: Sigma_gw + alpha (0 - Sigma_gw), copied to slabs 1-2
    for k in range(3):
        want = S0[k] + 0.5 * (0.0 - S0[k])
        assert np.array_equal(series(S[k], 0, 5, 2), want)
        assert np.array_equal(series(S[k], 0, 5, 0), want) and np.array_equal(series(S[k], 0, 5, 1), want)
    # band element of an interior slab
    i, j, ix = 40, 41, 7
    sub = lambda ts: [series(t, i, j, ix).reshape(1, 1, 1, nen) for t in ts]
    z = np.zeros((1, 1, 5, nen), complex)

    def slab5(x):  # the element in slab 2 of a 5-slab device (the step's first interior slab)
        y = z.copy()
        y[0, 0, 2] = x[0, 0, 0]
        return y

    ref = orc.selfenergy_step(*[slab5(x) for x in sub(G)], *[slab5(x) for x in sub(W)],
                              *[slab5(s.reshape(1, 1, 1, nen)) for s in S0[3:]], 1, nopmax, dE_S, 0.5)
    for k in range(3):
        got = series(S[k], i, j, ix)
        want = ref[k][0, 0, 2]
        assert np.linalg.norm(got - want) <= TOL * np.linalg.norm(want), k
    # edge copy of the last interior slab
    assert torch.equal(S[0][:, 17], S[0][:, 15]) and torch.equal(S[2][:, 16], S[2][:, 15])
    del G, W, S
    torch.cuda.empty_cache()  # ~85 GB back to the device for the tests that follow
