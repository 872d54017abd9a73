"""Writes tests/golden/gradient.json: the reference's RHF gradient terms
(scf::getEnergyGradient's loops over its own derivative integrals,
oracle/_ref, reference Boys series) for seeded synthetic densities on s/p and
d-shell systems.  The densities are random (Pa = C_occ C_occ^T of a random C, W
from random orbital energies): the fixture pins the formula, not an SCF.
Systems: water/STO-3G and two waters/6-31G (s, p), water/def2-SVP,
water/6-31G* and two waters/def2-SVP (d shells, up to (dd|dd)).

Usage (in a container with the reference tree, after `make -C oracle`):
    python tests/golden/make_gradient_golden.py
"""
from __future__ import annotations

import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

import dollarstorescf_amd as ds  # noqa: E402
from dollarstorescf_amd import geometry as G  # noqa: E402
from oracle import oracle as O  # noqa: E402
from oracle import ref as R  # noqa: E402


def main():
    out = {}
    Zw, xw = G.water()
    Z2, x2 = G.water_box(2, seed=3)
    cases = [("water_sto3g", "STO-3G", Zw, xw, 5), ("water2_631g", "6-31G", Z2, x2, 10),
             # d shells: one water (d on O), and two waters whose O d shells meet in (dd|dd)
             ("water_def2svp", "def2-SVP", Zw, xw, 5), ("water_631gs", "6-31GP", Zw, xw, 5),
             ("water2_def2svp", "def2-SVP", Z2, x2, 10)]
    for name, basis, Z, xyz, nocc in cases:
        rm = R.RefMolecule(ds.basis_path(basis), Z, xyz)
        n = rm.n
        rng = np.random.default_rng(17)
        C = rng.standard_normal((n, nocc)) * 0.3
        eps = -np.sort(rng.random(nocc) * 3.0)
        Pa = O.pack(C @ C.T)
        W = O.pack(2.0 * (C * eps) @ C.T)
        parts = {k: rm.gradient(2 * Pa, Pa, W, part).tolist() for k, part in (("one_electron", 1),
                                                                               ("two_electron", 2),
                                                                               ("nuclear", 4))}
        out[name] = {"basis": basis, "Z": [int(z) for z in Z], "xyz_bohr": np.asarray(xyz).reshape(-1).tolist(),
                     "Pa": Pa.tolist(), "W": W.tolist(), **parts}
    with open(os.path.join(HERE, "gradient.json"), "w") as f:
        json.dump(out, f, indent=0)
        f.write("\n")


if __name__ == "__main__":
    main()
