"""CPU: the sweep kernels' block-level program (solve_energy of
liteqt_amd/csrc/rgf_sweep.hpp, compiled with g++ against the host stand-in
tests/hostsweep/rgf_blocks_host.hpp) run on the CPU.  Every block pointer the
program hands to a block operation is checked against the arrays it may touch
(inputs, outputs, the workspace slot) and against the operand aliasing the
device functions allow; the results are compared with the oracle.  Covers the
plain solve (green_RGF_RS) and the overlap solve (green_RGF_CMS), the latter
at the shape of its GPU memory fault (nm = 64, nx = 3)."""
import ctypes
import os
import shutil
import subprocess

import numpy as np
import pytest

from liteqt_amd import configs
from oracle import pyoracle as orc
from tests import parity
from tests.test_oracle_cms import _overlap

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HERE = os.path.join(ROOT, "tests", "hostsweep")


@pytest.fixture(scope="module")
def hostlib(tmp_path_factory):
    d = tmp_path_factory.mktemp("hostsweep")
    shutil.copy(os.path.join(ROOT, "liteqt_amd", "csrc", "rgf_sweep.hpp"), d / "rgf_sweep.hpp")
    shutil.copy(os.path.join(ROOT, "liteqt_amd", "csrc", "rgf_wsweep.hpp"), d / "rgf_wsweep.hpp")
    shutil.copy(os.path.join(ROOT, "liteqt_amd", "csrc", "rgf_adjoint.hpp"), d / "rgf_adjoint.hpp")
    shutil.copy(os.path.join(HERE, "rgf_blocks_host.hpp"), d / "rgf_blocks.hpp")
    shutil.copy(os.path.join(HERE, "host_sweep.cpp"), d / "host_sweep.cpp")
    so = d / "libhostsweep.so"
    subprocess.run(["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-Wall", "-Wno-unused-variable",
                    "-Wno-unused-but-set-variable", "-o", str(so), str(d / "host_sweep.cpp")], check=True, cwd=d)
    L = ctypes.CDLL(str(so))
    vp, ci, cd = ctypes.c_void_p, ctypes.c_int, ctypes.c_double
    L.host_sweep.argtypes = ([ci, ci, ci, ci, ci, vp, vp, vp, cd, cd, cd, ci, cd] + [vp] * 5 + [ctypes.c_longlong]
                             + [vp] * 13 + [ctypes.c_char_p, ci, ci])
    L.host_w.argtypes = [ci, ci, ci] + [vp] * 12 + [ctypes.c_char_p, ci]
    L.host_adjoint.argtypes = ([ci, ci, ci, vp, vp, vp, cd, cd, cd, ci, cd] + [vp] * 10 + [ci, ci, vp,
                               ctypes.c_char_p, ci])
    return L


def relerr(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


def _p(a):
    return None if a is None else a.ctypes.data


def run(L, cms, mode, p, temp, E, mu, Hii, H1i, Sii=None, sigL=None, sigR=None, ckpt=0):
    nm, _, nx = Hii.shape
    E = np.ascontiguousarray(np.atleast_1d(E), np.float64)
    nE = len(E)
    shp = (nm, nm, nx, nE)
    out = {k: np.zeros(shp, np.complex128, order="F") for k in ("ldos", "ndens", "pdens", "GRi1", "GLi1", "GGi1")}
    out["cur"] = np.zeros(shp, np.float64, order="F")
    for k in ("tr", "tre"):
        out[k] = np.zeros(nE)
    out["dens_n"], out["dens_p"] = np.zeros((nE, nx)), np.zeros((nE, nx))
    out["info"] = np.zeros((nE, 4), np.int32)
    f = lambda a: None if a is None else np.asfortranarray(a, np.complex128)
    Hii, H1i, Sii, sigL, sigR = f(Hii), f(H1i), f(Sii), f(sigL), f(sigR)
    log = ctypes.create_string_buffer(1 << 16)
    nops = ctypes.c_longlong(0)
    t, m = np.asarray(temp, np.float64), np.asarray(mu, np.float64)  # kept alive across the call
    nv = L.host_sweep(int(cms), mode, nm, nx, nE, _p(E), _p(t), _p(m), p.eta_dev, p.eta_lead, p.dec_tol, p.dec_nmax, p.boltz,
                      _p(Hii), _p(H1i), _p(Sii), _p(sigL), _p(sigR), 0,
                      *[_p(out[k]) for k in ("ldos", "ndens", "pdens", "GRi1", "GLi1", "GGi1", "cur", "tr", "tre",
                                             "dens_n", "dens_p", "info")], ctypes.byref(nops), log, len(log), ckpt)
    return nv, log.value.decode(), nops.value, out


@pytest.mark.parametrize("nm,nx", [(64, 3), (32, 4), (20, 3), (8, 5)])
def test_overlap_solve_program_in_bounds_and_equal_to_oracle(hostlib, nm, nx):
    Hii, H1i = configs.random_device(nm, nx, 70 + nm, identical_leads=False)
    sL, sG, sR = configs.scattering_self_energies(nm, nx, 71 + nm, E=0.1, mu=0.0, scale=0.02)
    S = _overlap(nm, nx, 72 + nm)
    temp, mu, E = (300.0, 250.0), (0.1, -0.1), 0.13
    po = orc.params(orc.COMPAT, eta_dev=1e-4, eta_lead=1e-4, dec_tol=1e-12)
    nv, log, nops, g = run(hostlib, 1, 0, po, temp, E, mu, Hii, H1i, S, sL, sR)
    assert nv == 0, log
    assert nops > 10 * nx
    r = orc.rgf_cms(po, temp, E, mu, S, Hii, H1i, sL, sG, sR)
    for k in ("ldos", "ndens", "GRi1", "GLi1"):
        nb = nx if k in ("ldos", "ndens") else nx - 1
        for l in range(nb):
            assert relerr(g[k][..., l, 0], r[k][..., l]) <= 1e-10, (k, l)
    for l in range(nx):
        assert np.linalg.norm(g["pdens"][..., l, 0] - r["pdens"][..., l]) <= 1e-9 * np.linalg.norm(r["pdens"][..., l])
    assert abs(g["tr"][0] - r["tr"]) <= 1e-9 * max(1.0, abs(r["tr"]))


@pytest.mark.parametrize("mode", [0, 1])
@pytest.mark.parametrize("scat", [True, False])
def test_plain_solve_program_in_bounds_and_equal_to_oracle(hostlib, mode, scat):
    """Without scattering Sigma the program takes its anti-Hermitian path (g^< projected,
    the backward G^< update with one product fewer); with it, the reference's sequence."""
    nm, nx = 16, 5
    Hii, H1i = configs.random_device(nm, nx, 5)
    sL, sG, sR = configs.scattering_self_energies(nm, nx, 6, E=0.0, scale=0.02)
    if not scat:
        sL = sG = sR = None
    E = np.array([-0.4, 0.05, 0.7])
    temp, mu = (300.0, 250.0), (0.1, -0.1)
    po = orc.params(mode, eta_dev=1e-4, eta_lead=1e-4, dec_tol=1e-12)
    nv, log, nops, g = run(hostlib, 0, mode, po, temp, E, mu, Hii, H1i, None, sL, sR)
    assert nv == 0, log
    r = orc.rgf_energies(po, temp, E, mu, Hii, H1i, sL, sG, sR)
    for e in range(len(E)):
        for k in ("ldos", "ndens", "GRi1", "GLi1"):
            nb = nx if k in ("ldos", "ndens") else nx - 1
            for l in range(nb):
                assert relerr(g[k][..., l, e], r[k][..., l, e]) <= 1e-10, (k, l, e)
        for l in range(nx):
            assert np.linalg.norm(g["pdens"][..., l, e] - r["pdens"][..., l, e]) <= \
                1e-9 * np.linalg.norm(r["pdens"][..., l, e])
        assert abs(g["tr"][e] - r["tr"][e]) <= 1e-9 * max(1.0, abs(r["tr"][e]))
        assert abs(g["tre"][e] - r["tre"][e]) <= 1e-9 * max(1.0, abs(r["tre"][e]))


@pytest.mark.parametrize("nx,K", [(11, 3), (12, 5), (9, 4), (7, 1), (6, 6), (10, 8)])
@pytest.mark.parametrize("scat", [True, False])
def test_checkpointed_program_in_bounds_and_equal_to_keep_all(hostlib, nx, K, scat):
    """Segment recompute: with the left-connected blocks kept only every K slabs and
    in one K-slab segment (a workspace slot of 2 ((nx-2)/K + 1) + 2K + 30 blocks),
    every block pointer stays inside that smaller slot and every output equals the
    keep-all program's bit for bit."""
    nm = 8
    Hii, H1i = configs.random_device(nm, nx, 90 + nx)
    sL, sG, sR = configs.scattering_self_energies(nm, nx, 91 + nx, E=0.0, scale=0.02)
    if not scat:
        sL = sR = None
    E = np.array([-0.3, 0.2])
    temp, mu = (300.0, 250.0), (0.1, -0.1)
    for mode in (0, 1):
        po = orc.params(mode, eta_dev=1e-4, eta_lead=1e-4, dec_tol=1e-12)
        nv0, log0, n0, g0 = run(hostlib, 0, mode, po, temp, E, mu, Hii, H1i, None, sL, sR, ckpt=0)
        nv, log, n1, g = run(hostlib, 0, mode, po, temp, E, mu, Hii, H1i, None, sL, sR, ckpt=K)
        assert nv0 == 0 and nv == 0, log
        if K > 1 and nx - 2 > K:
            assert n1 > n0  # segments were recomputed
        for k in ("ldos", "ndens", "pdens", "GRi1", "GLi1", "GGi1", "cur", "tr", "tre", "info"):
            assert np.array_equal(g[k], g0[k]), (k, nx, K, mode)


@pytest.mark.parametrize("nm,nx,nE,sym", [(3, 2, 2, True), (4, 5, 3, False), (5, 7, 2, True)])
def test_w_program_in_bounds_and_equal_to_oracle(hostlib, nm, nx, nE, sym):
    """The screened-interaction program of rgf_wsweep.hpp: every block in bounds, the
    allowed aliasing only, results equal to the W oracle."""
    from oracle import w_oracle as wo
    from tests.test_gpu_w import _inputs
    rng = np.random.default_rng(nm + nx)
    Vii, V1i, PL, PG, PR = (np.asfortranarray(x, np.complex128) for x in _inputs(rng, nm, nx, nE, sym))
    Vc = np.asfortranarray(V1i.conj())
    W = {k: np.zeros((nm, nm, nx, nE), np.complex128, order="F") for k in ("WL", "WG", "WR", "WRi1")}
    info = np.zeros(nE, np.int32)
    nops = ctypes.c_longlong(0)
    log = ctypes.create_string_buffer(1 << 16)
    nv = hostlib.host_w(nm, nx, nE, _p(Vii), _p(V1i), _p(Vc), _p(PL), _p(PG), _p(PR), _p(W["WL"]), _p(W["WG"]),
                        _p(W["WR"]), _p(W["WRi1"]), _p(info), ctypes.byref(nops), log, len(log))
    assert nv == 0, log.value.decode()
    assert np.all(info == 0) and nops.value > 0
    for e in range(nE):
        ref = dict(zip(("WL", "WG", "WR", "WRi1"), wo.rgf_w(Vii, V1i, PL[..., e], PG[..., e], PR[..., e])))
        for k, r in ref.items():
            assert relerr(W[k][..., e], r) < 1e-12, (k, e)


@pytest.mark.parametrize("nx,K", [(7, 2), (9, 3), (10, 4), (6, 4)])
def test_adjoint_checkpointed_program_in_bounds_and_equal_to_keep_all(hostlib, nx, K):
    """The adjoint with its forward state kept only at checkpoints every K slabs (the
    backward sweep and both reverse passes refill one segment at a time): every block
    in bounds, and J, dU, dH bit-identical to the keep-all program (small and large
    slot)."""
    nm, nE = 4, 2
    Hii, H1i = configs.random_device(nm, nx, 30 + nx)
    sL, sG, sR = configs.scattering_self_energies(nm, nx, 31, E=np.array([-0.3, 0.2]), scale=0.05)
    E = np.array([-0.3, 0.2])
    temp, mu = np.array([300.0, 250.0]), np.array([0.1, -0.1])
    w_tr = np.array([1.0, 0.5])
    w_n = np.random.default_rng(1).standard_normal((nE, nx))
    outs = {}
    for keep, ck in ((1, 0), (0, 0), (0, K)):
        J, tr, dU = np.zeros(nE), np.zeros(nE), np.zeros((nE, nx))
        dH = np.zeros((nm, nm, nx), np.complex128, order="F")
        nops = ctypes.c_longlong(0)
        log = ctypes.create_string_buffer(1 << 16)
        nv = hostlib.host_adjoint(nm, nx, nE, _p(E), _p(temp), _p(mu), 1e-4, 1e-2, 1e-10, 100, 8.61734e-5,
                                  _p(Hii), _p(H1i), _p(sL), _p(sR), _p(w_tr), _p(w_n), _p(J), _p(tr), _p(dU), _p(dH),
                                  keep, ck, ctypes.byref(nops), log, len(log))
        assert nv == 0, log.value.decode()
        outs[(keep, ck)] = (J, tr, dU, dH, nops.value)
    base = outs[(0, 0)]
    for key in ((1, 0), (0, K)):
        for a, b in zip(outs[key][:4], base[:4]):
            assert np.array_equal(a, b), key
    if (nx - 3) // K >= 1:  # more than one segment: the refills ran
        assert outs[(0, K)][4] > base[4]
    else:  # one segment stays in the buffer from the forward sweep on
        assert outs[(0, K)][4] == base[4]


def test_adjoint_with_contacts_program_equal_to_reference(hostlib):
    """The adjoint program with contacts that repeat the edge slabs (no fixed lead
    blocks): the contacts' derivatives through the decimation's fixed point (edge dU,
    edge couplings) equal tests/adjoint_ref.adjoint_with_leads, which central
    differences of the oracle pin (tests/test_adjoint_ref.py)."""
    from tests import adjoint_ref
    from tests.dense_ref import fermi
    nm, nx, nE = 4, 6, 2
    Hii, H1i = configs.random_device(nm, nx, 41, coupling_sigma=1.0, identical_leads=False)
    E = np.array([-0.3, 0.2])
    sL, sG, sR = configs.scattering_self_energies(nm, nx, 42, E=E, scale=0.05)
    temp, mu = np.array([300.0, 250.0]), np.array([0.1, -0.1])
    w_tr = np.array([1.0, 0.5])
    w_n = np.random.default_rng(2).standard_normal((nE, nx))
    eta_dev, eta_lead, tol, nmax, boltz = 1e-4, 1e-2, 1e-12, 100, 8.61734e-5
    J, tr, dU = np.zeros(nE), np.zeros(nE), np.zeros((nE, nx))
    dH = np.zeros((nm, nm, nx), np.complex128, order="F")
    nops = ctypes.c_longlong(0)
    log = ctypes.create_string_buffer(1 << 16)
    f = lambda a: np.asfortranarray(a, np.complex128)
    Hf, Cf, sLf, sRf = f(Hii), f(H1i), f(sL), f(sR)
    nv = hostlib.host_adjoint(nm, nx, nE, _p(E), _p(temp), _p(mu), eta_dev, eta_lead, tol, nmax, boltz,
                              _p(Hf), _p(Cf), _p(sLf), _p(sRf), _p(w_tr), _p(w_n), _p(J), _p(tr), _p(dU), _p(dH),
                              0, 0, ctypes.byref(nops), log, len(log))
    assert nv == 0, log.value.decode()
    dH_ref = np.zeros_like(dH)
    for e in range(nE):
        fL, fR = fermi(E[e], mu[0], boltz, temp[0]), fermi(E[e], mu[1], boltz, temp[1])
        Jr, dUr, Hr = adjoint_ref.adjoint_with_leads(E[e], eta_dev, eta_lead, Hii, H1i, sL[..., e], sR[..., e], fL, fR,
                                                     w_tr[e], w_n[e], orc.sancho, tol, nmax)
        assert abs(J[e] - Jr) <= 1e-10 * max(1.0, abs(Jr)), e
        assert np.abs(dU[e] - dUr).max() <= 1e-9 * max(1.0, np.abs(dUr).max()), (e, dU[e], dUr)
        dH_ref += Hr
    for l in range(nx):
        assert relerr(dH[:, :, l], dH_ref[:, :, l]) <= 1e-9, l
