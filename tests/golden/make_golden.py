"""Writes the golden fixtures of tests/golden/ (run in a container that has the
reference tree, after `make -C oracle`).

Sources:
  * oracle/_ref (DollarStoreSCF's own GaussianIntegral.cpp/GTO.cpp/
    BasisReader.cpp compiled unmodified): Boys values, ERIs, S/T/V;
  * mpmath at 50 digits: accurate Boys values;
  * the oracle's SCF restatement (oracle/oracle.py) over the reference-mode
    integrals: converged energies (the reference's SCF layer does not build
    here: it needs LAPACK and source fixes).

Usage:  python tests/golden/make_golden.py
"""
from __future__ import annotations

import json
import math
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from dollarstorescf_amd import geometry as G  # noqa: E402
from oracle import oracle as O  # noqa: E402
from oracle import ref as R  # noqa: E402

BASIS = os.path.join(ROOT, "dollarstorescf_amd", "basis")
XS = [0.0, 1e-6, 1e-3, 0.1, 0.5, 0.838, 1.0, 1.525, 2.0, 5.0, 10.0, 15.0, 20.0, 24.9, 25.0, 26.0, 27.2, 28.3,
      30.0, 33.8, 35.0, 40.0, 45.0, 60.0, 100.0]


def dump(name, obj):
    with open(os.path.join(HERE, name), "w") as f:
        json.dump(obj, f, indent=0)
        f.write("\n")


def boys():
    import mpmath as mp

    mp.mp.dps = 50
    ref = [[R.boys(x, m) for m in range(9)] for x in XS]
    acc = []
    for x in XS:
        row = []
        for m in range(15):
            if x == 0.0:
                v = mp.mpf(1) / (2 * m + 1)
            else:
                X = mp.mpf(x)
                v = mp.gammainc(m + mp.mpf(1) / 2, 0, X) / (2 * X ** (m + mp.mpf(1) / 2))
            row.append(float(v))
        acc.append(row)
    dump("boys.json", {"x": XS, "reference_series": ref, "mpmath": acc,
                       "note": "reference_series: boysFunctionExact (GaussianIntegral.cpp:209-223) from oracle/_ref; "
                               "mpmath: gamma(m+1/2,0,x)/(2 x^(m+1/2)) at 50 digits"})


def eris():
    out = {}
    cases = {
        "water_sto3g": ("STO-3G", G.water()),
        "water_631g": ("6-31G", G.water()),
        "water_631gs": ("6-31GP", G.water()),
        "water_def2svp": ("def2-SVP", G.water()),
        "water2_631g": ("6-31G", G.water_box(2, seed=3)),
    }
    rng = np.random.default_rng(11)
    for key, (basis, (Z, xyz)) in cases.items():
        rm = R.RefMolecule(os.path.join(BASIS, basis), Z, xyz)
        n = rm.n
        if key == "water_sto3g":
            q = np.array([[i, j, k, l] for i in range(n) for j in range(i + 1) for k in range(n) for l in range(k + 1)
                          if i * (i + 1) // 2 + j >= k * (k + 1) // 2 + l], np.int32)
        else:
            q = rng.integers(0, n, size=(400, 4)).astype(np.int32)
        v = rm.eri_list(q)
        S, T, V = rm.one_electron() if key in ("water_sto3g", "water_def2svp") else (None, None, None)
        out[key] = {"basis": basis, "Z": Z.tolist(), "xyz_bohr": xyz.tolist(), "n_ao": n,
                    "quartets": q.tolist(), "eri": v.tolist()}
        if S is not None:
            out[key].update(S=S.tolist(), T=T.tolist(), V=V.tolist())
    dump("reference_integrals.json", out)


def energies():
    res = {}
    cases = {
        "he_sto3g": ("STO-3G", np.array([2]), np.zeros((1, 3))),
        "h2_sto3g_1.4": ("STO-3G", np.array([1, 1]), np.array([[0, 0, 0], [0, 0, 1.4]])),
        "water_sto3g": ("STO-3G",) + G.water(),
        "water_631g": ("6-31G",) + G.water(),
        "water_lattice8_631g": ("6-31G",) + G.water_lattice(1),
    }
    for key, (basis, Z, xyz) in cases.items():
        res[key] = {"basis": basis}
        for mode, name in ((0, "reference"), (1, "accurate")):
            r = O.scf_energy(os.path.join(BASIS, basis), Z, xyz, mode=mode, diis_size=8, conv=1e-10)
            res[key][name] = {"energy": r.energy, "iterations": r.iterations, "components": r.components}
            print(key, name, r.energy, r.iterations, flush=True)
    dump("energies.json", {"note": "oracle RHF (DIIS 8, |dE_elec| < 1e-10) over reference-mode or accurate-mode "
                                   "integrals; reference Boys mode = the reference's own arithmetic",
                           "cases": res})


if __name__ == "__main__":
    boys()
    eris()
    energies()
