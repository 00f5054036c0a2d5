"""GPU: the per-energy drop-in (rgf_rs / green_RGF_RS_gpu) called concurrently, as
the reference's OpenMP energy loops call green_RGF_RS from every thread
(src/green_rgf.f90:1301-1318).  Concurrent calls are gathered into one batched
launch; each caller must get exactly the result of its own energy."""
import os
import subprocess
import threading
import time

import numpy as np
import pytest

from liteqt_amd import configs
from oracle import pyoracle as orc
from tests import parity

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
OMP_DRIVER = os.path.join(ROOT, "tests", "fortran", "rgf_omp_driver")


def _rgf():
    import torch
    if not torch.cuda.is_available():
        pytest.fail("no HIP device visible")
    from liteqt_amd import rgf
    return rgf


def _threads(fn, n):
    start = threading.Barrier(n)
    out = [None] * n

    def run(k):
        start.wait()
        out[k] = fn(k)

    ts = [threading.Thread(target=run, args=(k,)) for k in range(n)]
    for t in ts:
        t.start()
    for t in ts:
        t.join()
    return out


def test_concurrent_calls_equal_their_own_energies():
    """8 threads, each its own energy and its own scattering self-energy: every
    output bit-identical to the batched energy loop on the same inputs, which
    itself meets the oracle."""
    rgf = _rgf()
    nm, nx, n = 64, 5, 8
    Hii, H1i = configs.random_device(nm, nx, 61)
    E = configs.energy_grid(-1.0, 1.0, n)
    sL, sG, sR = configs.scattering_self_energies(nm, nx, 62, E=E, scale=0.02)
    p = rgf.params(rgf.MODE_COMPAT, eta_dev=1e-4, eta_lead=1e-4, dec_tol=1e-12)
    temp, mu = (300.0, 300.0), (0.1, -0.1)
    got = _threads(lambda k: rgf.rgf_rs(temp, E[k], mu, Hii, H1i, sL[..., k], sG[..., k], sR[..., k], p=p), n)
    ref = rgf.rgf_rs_energies(p, temp, E, mu, Hii, H1i, sL, sG, sR)
    for k in range(n):
        for key in ("ldos", "ndens", "pdens", "cur", "GRi1", "GLi1", "GGi1"):
            assert np.array_equal(got[k][key], ref[key][..., k]), (k, key)
        assert got[k]["tr"] == ref["tr"][k] and got[k]["tre"] == ref["tre"][k]
    po = orc.params(orc.COMPAT, eta_dev=1e-4, eta_lead=1e-4, dec_tol=1e-12)
    for k in (0, n - 1):
        r = orc.rgf(po, temp, E[k], mu, Hii, H1i, sL[..., k], sG[..., k], sR[..., k])
        parity.check(got[k], r, H1i, "compat", f"concurrent E[{k}]")


def test_concurrent_calls_run_as_one_batch():
    """8 concurrent calls of an energy that takes ~0.1 s alone finish in well under
    8 times that: they ran as one launch, not one after another."""
    rgf = _rgf()
    nm, nx, n = 256, 12, 8
    Hii, H1i = configs.random_device(nm, nx, 63)
    E = configs.energy_grid(-1.0, 1.0, n)
    temp, mu = (300.0, 300.0), (0.1, -0.1)
    p = rgf.params(rgf.MODE_COMPAT, eta_dev=1e-5, eta_lead=1e-5, dec_tol=1e-12)
    rgf.rgf_rs(temp, E[0], mu, Hii, H1i, p=p, offdiag=False)  # workspace and staging sized
    t0 = time.perf_counter()
    one = rgf.rgf_rs(temp, E[0], mu, Hii, H1i, p=p, offdiag=False)
    t_one = time.perf_counter() - t0
    t0 = time.perf_counter()
    got = _threads(lambda k: rgf.rgf_rs(temp, E[k], mu, Hii, H1i, p=p, offdiag=False), n)
    t_all = time.perf_counter() - t0
    print(f"\none call {t_one:.3f} s, {n} concurrent calls {t_all:.3f} s", flush=True)
    assert np.array_equal(got[0]["ldos"], one["ldos"])
    assert t_all < 0.5 * n * t_one, (t_one, t_all)


def test_fortran_openmp_loop_matches_serial_loop_and_oracle(tmp_path):
    """The Fortran driver's OpenMP energy loop (8 threads calling green_RGF_RS_gpu)
    gives the bits of its serial loop (1 thread), and the oracle's values."""
    if not os.path.exists(OMP_DRIVER):
        pytest.fail("OpenMP Fortran driver not built (run __graft_entry__.build())")
    nm, nx, nE = 40, 4, 12  # nm not a multiple of 32: the padded path
    Hii, H1i = configs.random_device(nm, nx, 64)
    E = configs.energy_grid(-0.8, 0.8, nE)
    temp, mu = np.array([300.0, 300.0]), np.array([0.1, -0.1])
    with open(tmp_path / "in.bin", "wb") as f:
        np.array([nm, nx, nE], np.int32).tofile(f)
        for a in (temp, mu, E):
            np.asarray(a, np.float64).tofile(f)
        for a in (Hii, H1i):
            np.asfortranarray(a, np.complex128).ravel(order="F").tofile(f)
    res = {}
    for nt in (1, 8):
        env = dict(os.environ, OMP_NUM_THREADS=str(nt))
        r = subprocess.run([OMP_DRIVER], cwd=tmp_path, capture_output=True, text=True, timeout=120, env=env)
        assert r.returncode == 0, r.stdout + r.stderr
        print(r.stdout.strip(), flush=True)
        raw = np.fromfile(tmp_path / "out.bin", dtype=np.uint8)
        tr = raw[: 8 * nE].view(np.float64)
        tre = raw[8 * nE: 16 * nE].view(np.float64)
        ldos = raw[16 * nE:].view(np.complex128).reshape((nm, nm, nx, nE), order="F")
        res[nt] = (tr.copy(), tre.copy(), ldos.copy())
    for a, b in zip(res[1], res[8]):
        assert np.array_equal(a, b)
    ref = orc.rgf_energies(orc.params(orc.COMPAT), tuple(temp), E, tuple(mu), Hii, H1i, outputs=("ldos",))
    tr, tre, ldos = res[8]
    for e in range(nE):
        for l in range(nx):
            assert np.linalg.norm(ldos[..., l, e] - ref["ldos"][..., l, e]) <= 1e-10 * np.linalg.norm(ref["ldos"][..., l, e])
    assert np.all(np.abs(tr - ref["tr"]) <= 1e-9 * np.maximum(1.0, np.abs(ref["tr"])))
