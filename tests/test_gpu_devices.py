"""GPU: many independent devices (different nm, nx and energy grids) in one
persistent launch (rgf_solve_devices), with and without the autotuner, and as
a HIP graph; every device's outputs against the CPU oracle."""
import numpy as np
import pytest

from liteqt_amd import configs
from oracle import pyoracle as orc
from tests import parity

pytestmark = pytest.mark.gpu

KW = dict(eta_dev=1e-4, eta_lead=1e-4, dec_tol=1e-12)
TEMP, MU = (300.0, 250.0), (0.1, -0.1)
SPECS = [(32, 4, 3), (64, 3, 2), (96, 2, 1), (128, 3, 2), (64, 5, 4), (32, 2, 1), (192, 2, 1), (256, 2, 1), (128, 2, 0)]


def _env():
    import torch
    if not torch.cuda.is_available():
        pytest.fail("no HIP device visible")
    from liteqt_amd import rgf
    return torch, rgf


def _devices(torch, rgf, dev, scatter_every=1):
    host, devs = [], []
    for j, (nm, nx, nE) in enumerate(SPECS):
        Hii, H1i = configs.random_device(nm, nx, 300 + j, identical_leads=False)
        sL = sR = None
        if j % scatter_every == 0:
            sL, _, sR = configs.scattering_self_energies(nm, nx, 400 + j, E=0.1, mu=0.0, scale=0.02)
        E = np.linspace(-0.5, 0.5, nE) + 0.01 * j
        out = {k: torch.zeros(nE, nx, nm, nm, dtype=torch.complex128, device=dev) for k in ("ldos", "ndens", "GRi1",
                                                                                              "GLi1")}
        for k in ("tr", "tre"):
            out[k] = torch.zeros(nE, dtype=torch.float64, device=dev)
        out["dens_n"] = torch.zeros(nE, nx, dtype=torch.float64, device=dev)
        out["info"] = torch.zeros(nE, 4, dtype=torch.int32, device=dev)
        d = {"E": torch.tensor(E, dtype=torch.float64, device=dev), "temp": TEMP, "mu": MU,
             "Hii": rgf.to_device_blocks(Hii, dev), "H1i": rgf.to_device_blocks(H1i, dev), "out": out}
        if sL is not None:
            d["sigL"], d["sigR"] = rgf.to_device_blocks(sL, dev), rgf.to_device_blocks(sR, dev)
        devs.append(d)
        host.append((Hii, H1i, sL, sR, E))
    return host, devs


def _check(rgf, po, mode, host, devs):
    for j, ((Hii, H1i, sL, sR, E), d) in enumerate(zip(host, devs)):
        if len(E) == 0:
            continue
        r = orc.rgf_energies(po, TEMP, E, MU, Hii, H1i, sL, None, sR, outputs=("ldos", "ndens", "GRi1", "GLi1"))
        g = {k: rgf.from_device_blocks(d["out"][k]) for k in ("ldos", "ndens", "GRi1", "GLi1")}
        tr, tre = d["out"]["tr"].cpu().numpy(), d["out"]["tre"].cpu().numpy()
        for e in range(len(E)):
            ge = {k: g[k][..., e] for k in g}
            ge["tr"], ge["tre"] = tr[e], tre[e]
            re_ = {k: r[k][..., e] for k in ("ldos", "ndens", "GRi1", "GLi1")}
            re_["tr"], re_["tre"] = r["tr"][e], r["tre"][e]
            # tr/tre tolerance uses G^> scales: take them from a one-energy oracle call
            full = orc.rgf(po, TEMP, E[e], MU, Hii, H1i, sL, None, sR)
            re_["pdens"] = full["pdens"]
            parity.check(ge, re_, H1i, mode, f"device {j} e={e}")
        dn = np.imag(np.trace(r["ndens"], axis1=0, axis2=1)).T
        assert np.allclose(d["out"]["dens_n"].cpu().numpy(), dn, rtol=1e-10, atol=1e-12), j
        assert (d["out"]["info"].cpu().numpy()[:, :2] == np.array([x[0:2] for x in r["diag"]])).all(), j


def _solve_each(rgf, p, devs):
    """The same devices, one rgf_solve_batched launch each (fresh output tensors)."""
    outs = []
    for d in devs:
        o = {k: v.clone().zero_() for k, v in d["out"].items()}
        rgf.solve_batched(p, d["E"], d["temp"], d["mu"], d["Hii"], d["H1i"], sigL=d.get("sigL"), sigR=d.get("sigR"),
                          out=o)
        outs.append(o)
    return outs


@pytest.mark.parametrize("mode", ["compat", "physical"])
def test_devices_one_launch_equals_per_device_launches(mode):
    """Without the autotuner every device runs the tile variant rgf_solve_batched
    picks, so the one-launch results are bit-identical to per-device launches;
    devices without scattering included."""
    torch, rgf = _env()
    dev = torch.device("cuda:0")
    host, devs = _devices(torch, rgf, dev, scatter_every=2)
    p = rgf.params(rgf.MODE_COMPAT if mode == "compat" else rgf.MODE_PHYSICAL, **KW)
    tiles, rep = rgf.solve_devices(p, devs, autotune=False)
    torch.cuda.synchronize()
    assert rep["items"] == sum(s[2] for s in SPECS) and rep["slots"] >= 1 and rep["tuned"] == 0
    ref = _solve_each(rgf, p, devs)
    torch.cuda.synchronize()
    for j, (d, r) in enumerate(zip(devs, ref)):
        for k in r:
            assert torch.equal(d["out"][k], r[k]), (j, k)


def test_devices_autotuned_match_oracle():
    """With the autotuner (tile variants may differ from rgf_solve_batched's)
    every device's outputs against the CPU oracle at the parity bar; the devices
    carry scattering self-energies (broadened states)."""
    torch, rgf = _env()
    dev = torch.device("cuda:0")
    host, devs = _devices(torch, rgf, dev, scatter_every=1)
    p = rgf.params(rgf.MODE_PHYSICAL, **KW)
    po = orc.params(orc.PHYSICAL, **KW)
    tiles, rep = rgf.solve_devices(p, devs, autotune=True)
    torch.cuda.synchronize()
    assert rep["items"] == sum(s[2] for s in SPECS) and rep["slots"] >= 1
    for (nm, _, _), t in zip(SPECS, tiles):
        assert nm % (32 * t) == 0 or (t == 4 and nm % 128 == 0), (nm, t)
    _check(rgf, po, "physical", host, devs)


def test_devices_graph_replays_match_eager():
    torch, rgf = _env()
    dev = torch.device("cuda:0")
    host, devs = _devices(torch, rgf, dev)
    p = rgf.params(rgf.MODE_PHYSICAL, **KW)
    rgf.solve_devices(p, devs)
    torch.cuda.synchronize()
    want = [{k: v.clone() for k, v in d["out"].items()} for d in devs]
    for d in devs:
        for v in d["out"].values():
            v.zero_()
    stream = torch.cuda.Stream()
    g = rgf.DevicesGraph(p, devs, stream)
    try:
        for rep in range(2):
            g.launch(1)
            stream.synchronize()
            for j, (d, w) in enumerate(zip(devs, want)):
                for k in w:
                    assert torch.equal(d["out"][k], w[k]), (rep, j, k)
            for d in devs:
                d["out"]["tr"].zero_()
            torch.cuda.synchronize()
    finally:
        g.close()


def test_devices_info_in_pinned_host_memory():
    """info may be a pinned host tensor (the kernels write it through the device's
    mapping of host memory): the counters equal those of a device-resident info, and
    the other outputs are unchanged by where info lives."""
    torch, rgf = _env()
    dev = torch.device("cuda:0")
    host, devs = _devices(torch, rgf, dev, scatter_every=2)
    p = rgf.params(rgf.MODE_COMPAT, **KW)
    rgf.solve_devices(p, devs, autotune=False)
    torch.cuda.synchronize()
    want = [{k: v.clone() for k, v in d["out"].items()} for d in devs]
    for d in devs:
        for v in d["out"].values():
            v.zero_()
        d["out"]["info"] = torch.zeros(d["out"]["info"].shape, dtype=torch.int32, pin_memory=True)
    torch.cuda.synchronize()
    rgf.solve_devices(p, devs, autotune=False)
    torch.cuda.synchronize()
    for j, (d, w) in enumerate(zip(devs, want)):
        assert torch.equal(d["out"]["info"], w["info"].cpu()), j
        for k in w:
            if k != "info":
                assert torch.equal(d["out"][k], w[k]), (j, k)
    # the batched entry point takes it as well
    d, w = devs[3], want[3]
    o = {k: v.clone().zero_() for k, v in d["out"].items() if k != "info"}
    o["info"] = torch.zeros(d["out"]["info"].shape, dtype=torch.int32, pin_memory=True)
    rgf.solve_batched(p, d["E"], d["temp"], d["mu"], d["Hii"], d["H1i"], sigL=d.get("sigL"), sigR=d.get("sigR"), out=o)
    torch.cuda.synchronize()
    assert torch.equal(o["info"], w["info"].cpu())
