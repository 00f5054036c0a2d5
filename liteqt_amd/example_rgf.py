"""The reference's example run on the engine: iteration 0 of green_rgf_solve_gw_1D
(src/green_rgf.f90:3170-3186, no self-energy yet) for a Wannier device, writing the
driver's files (gw_ldos0000.dat, gw_ndos0000.dat, gw_pdos0000.dat, gw_Jdens_0000.dat,
gw_trR_0000.dat, gw_trL_0000.dat, gw_Id_iteration.dat).

The device is the RGF driver's (src/main.f90:500-530, liteqt_amd/wannier.py): `length`
slabs of ns cells from ham_dat, the pot_dat potential on the first nb orbitals of each
slab; energies En(i) = emin + i (emax - emin) / nen (src/main.f90:524).  The example's
input file (examples/cnt_8-0/input) has two lines more than src/main.f90:55-63 reads
(`scba tol`, `epsilon`), so read_example_input takes its values by their labels.
Energies run through the engine in chunks; each chunk is reduced at once to the cell
traces of G^r, G^<, G^> and to the bond currents, so the host holds no (nm, nm, nx, nen)
array.
"""
import os
import re

import numpy as np

from . import configs, rgf, wannier, writers


def read_example_input(path):
    """Values of the example input file by their labels."""
    out = {}
    labels = {"num cell in sancho": "ns", "length num cells": "length", "emin emax num_en": "grid",
              "read potential from pot_dat": "lreadpot", "mu_s mu_d": "mu", "temperature": "temp",
              "num iter": "niter", "epsilon": "eps", "r0": "r0", "alpha mix": "alpha", "only diag": "ldiag"}
    lines = open(path).read().splitlines()
    for k, ln in enumerate(lines):
        if "!" not in ln:
            continue
        val, lab = ln.split("!", 1)
        lab = lab.strip()
        for key, name in labels.items():
            if lab.startswith(key):
                out[name] = val.split()
                if name == "lreadpot" and val.strip().upper().startswith("T"):
                    out["potscale"] = float(lines[k + 1].split()[0])  # the line after (src/main.f90:55-57)
    return {"ns": int(out["ns"][0]), "length": int(out["length"][0]),
            "emin": float(out["grid"][0]), "emax": float(out["grid"][1]), "nen": int(out["grid"][2]),
            "lreadpot": out["lreadpot"][0].upper().startswith("T"), "potscale": out.get("potscale", 1.0),
            "mu": (float(out["mu"][0]), float(out["mu"][1])), "temp": (float(out["temp"][0]), float(out["temp"][1])),
            # the GW loop's values (src/main.f90:58-66)
            "niter": int(out["niter"][0]) if "niter" in out else 0,
            "eps": float(out["eps"][0]) if "eps" in out else 1.0,
            "r0": float(out["r0"][0]) if "r0" in out else 1.0,
            "alpha": float(out["alpha"][0]) if "alpha" in out else 0.5,
            "ldiag": out["ldiag"][0].upper().startswith("T") if "ldiag" in out else True}


def run(example_dir, out_dir, nen=None, chunk=250):
    """Run iteration 0 of the 1D driver on the example in example_dir; files go to out_dir.
    nen: override the number of energies.  Returns the transmissions (tr, tre) and energies."""
    cfg = read_example_input(os.path.join(example_dir, "input"))
    if nen:
        cfg["nen"] = nen
    w = wannier.load_ham_dat(os.path.join(example_dir, "ham_dat"))
    ns, length = cfg["ns"], cfg["length"]
    pot = wannier.read_pot(os.path.join(example_dir, "pot_dat"), length) if cfg["lreadpot"] else None
    Hii, H1i = wannier.device_blocks(w, ns, length, pot, cfg["potscale"])
    en = configs.energy_grid(cfg["emin"], cfg["emax"], cfg["nen"])
    nE, nb, spindeg = len(en), w.nb, float(w.spin_deg)
    dE = en[1] - en[0]
    p = rgf.params(rgf.MODE_COMPAT)  # green_RGF_RS's own parameters
    tr_r = np.zeros((ns, length, nE), complex)
    tr_l = np.zeros((ns, length, nE), complex)
    tr_g = np.zeros((ns, length, nE), complex)
    jtot = np.zeros((length * ns, nE))
    tr = np.zeros(nE)
    tre = np.zeros(nE)
    for e0 in range(0, nE, chunk):
        sl = slice(e0, min(nE, e0 + chunk))
        g = rgf.rgf_rs_energies(p, cfg["temp"], en[sl], cfg["mu"], Hii, H1i, outputs=("ndens", "pdens", "ldos"))
        # the driver writes these files before it zeroes the real parts of G
        # (src/green_rgf.f90:3181-3188), so they see the full G
        tr_r[..., sl] = writers.spectrum_traces(g["ldos"], nb, ns)
        tr_l[..., sl] = writers.spectrum_traces(g["ndens"], nb, ns)
        tr_g[..., sl] = writers.spectrum_traces(g["pdens"], nb, ns)
        _, _, jd = writers.calc_block_current(Hii, g["ndens"], en[sl], spindeg, nb, ns, dE=dE)
        jtot[:, sl] = jd.sum(axis=(0, 1))
        tr[sl], tre[sl] = g["tr"], g["tre"]
    os.makedirs(out_dir, exist_ok=True)
    d = lambda name: os.path.join(out_dir, name)
    writers.write_spectrum_traces(d("gw_ldos"), 0, tr_r, en, w.Lx, (1.0, -2.0))
    writers.write_spectrum_traces(d("gw_ndos"), 0, tr_l, en, w.Lx, (1.0, 1.0))
    writers.write_spectrum_traces(d("gw_pdos"), 0, tr_g, en, w.Lx, (1.0, -1.0))
    writers.write_current_totals(d("gw_Jdens"), 0, jtot, en, w.Lx)
    writers.write_transmission_spectrum(d("gw_trR"), 0, tr * spindeg, en)
    writers.write_transmission_spectrum(d("gw_trL"), 0, tre * spindeg, en)
    cur = lambda t: np.sum(t) * dE * writers.E0 / writers.TPI / writers.HBAR * writers.E0 * spindeg
    with open(d("gw_Id_iteration.dat"), "a") as f:  # write(101,'(I4,2E16.6)') iter, I_L, I_R
        f.write("%4d" % 0 + writers.fortran_e(cur(tre), 16, 6) + writers.fortran_e(cur(tr), 16, 6) + "\n")
    return {"en": en, "tr": tr, "tre": tre, "jtot": jtot, "cfg": cfg}
