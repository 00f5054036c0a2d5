"""Parity metric between the HIP engine and the CPU oracle.

The bar is a relative Frobenius error of 1e-10 per block.  Outputs that the
reference forms by subtracting nearly equal terms are measured against the
scale of those terms (the error a faithful evaluation of the same formula can
make), not against their possibly tiny result:

  * G^>_ll = G^<_ll + G_ll - G_ll^H (green_rgf.f90:4056, 4148).  Below the Fermi
    level 1 - f -> 0 and |G^>| is ~(1-f) times the operands; in a band gap
    G_ll is nearly Hermitian (|G - G^H| can fall to 1e-15 |G| in the middle of
    a device), so the scale is that of the operands, |G^<_ll| + 2 |G_ll| (any
    G_ll within the ldos bar gives a G^>_ll within this one).
  * GGi1 (compat) = g^H H^H G^>_{nx,nx} inherits the amplification
    amp = (|G^<_nn| + |G_nn - G_nn^H|) / |G^>_nn| of its G^> factor; GGi1
    (physical) = G^<_{l,l+1} + G_{l,l+1} - G_{l+1,l}^H has the scale
    |G^<_{l,l+1}| + 2|G_{l,l+1}|.
  * cur = Re(H^H G^<_{l+1,l}) - Re(H G^<_{l,l+1}) cancels in equilibrium; its
    scale is |cur| + 2 |H| |G^<_{l,l+1}|.
  * tr / tre are traces with the G^> blocks above: tolerance
    max(1, |value|) * amp of the contact block.
"""
import numpy as np

TOL = 1e-10


def _n(a):
    return float(np.linalg.norm(a))


def amp_gp(ref, l):
    """Cancellation amplification of G^>_ll = G^<_ll + G_ll - G_ll^H."""
    gp = _n(ref["pdens"][..., l])
    scale = _n(ref["ndens"][..., l]) + 2 * _n(ref["ldos"][..., l])
    return max(1.0, scale / gp) if gp > 0 else float("inf")


def block_errors(gpu, ref, H1i, mode):
    """Yield (name, slab, err, allowed) for every compared quantity of one energy."""
    nx = ref["ldos"].shape[2]
    for k in ("ldos", "ndens"):
        if k in gpu:
            for l in range(nx):
                yield k, l, _n(gpu[k][..., l] - ref[k][..., l]), TOL * _n(ref[k][..., l])
    if "pdens" in gpu:
        for l in range(nx):
            scale = _n(ref["ndens"][..., l]) + 2 * _n(ref["ldos"][..., l])
            yield "pdens", l, _n(gpu["pdens"][..., l] - ref["pdens"][..., l]), TOL * scale
    for k in ("GRi1", "GLi1"):
        if k in gpu:
            for l in range(nx - 1):
                yield k, l, _n(gpu[k][..., l] - ref[k][..., l]), TOL * _n(ref[k][..., l])
    if "GGi1" in gpu:
        for l in range(nx - 1):
            if mode == "compat":
                scale = _n(ref["GGi1"][..., l]) * amp_gp(ref, nx - 1)
            else:
                scale = _n(ref["GLi1"][..., l]) + 2 * _n(ref["GRi1"][..., l])
            yield "GGi1", l, _n(gpu["GGi1"][..., l] - ref["GGi1"][..., l]), TOL * scale
    if "cur" in gpu:
        for l in range(nx - 1):
            Hc = H1i[..., l if mode == "compat" else l + 1]
            scale = _n(ref["cur"][..., l]) + 2 * _n(Hc) * _n(ref["GLi1"][..., l])
            yield "cur", l, _n(gpu["cur"][..., l] - ref["cur"][..., l]), TOL * scale
    if "tr" in gpu:
        yield "tr", nx - 1, abs(gpu["tr"] - ref["tr"]), TOL * max(1.0, abs(ref["tr"])) * amp_gp(ref, nx - 1)
    if "tre" in gpu:
        yield "tre", 0, abs(gpu["tre"] - ref["tre"]), TOL * max(1.0, abs(ref["tre"])) * amp_gp(ref, 0)


def check(gpu, ref, H1i, mode, tag=""):
    """Assert parity; return the worst err/allowed ratio."""
    worst = 0.0
    for name, l, err, allowed in block_errors(gpu, ref, H1i, mode):
        ratio = err / allowed if allowed > 0 else (0.0 if err == 0 else float("inf"))
        worst = max(worst, ratio)
        assert err <= allowed, f"{tag} {name} slab {l}: err {err:.3e} > allowed {allowed:.3e}"
    return worst


def energy_slice(d, e):
    """Per-energy view of a dict of (..., nE) arrays / (nE,) vectors."""
    out = {}
    for k, v in d.items():
        if k in ("info", "diag"):
            continue
        v = np.asarray(v)
        out[k] = v[..., e] if v.ndim == 4 else v[e]
    return out
