"""Writes tests/golden/molden.json: Molden files produced by the reference's
own writer logic and FormattedStream (oracle/_ref, built from the reference
tree by oracle/Makefile) for seeded orbitals on three molecule/basis cases.
The orbitals are synthetic (random, with rounding-boundary values planted in
the first orbital): the file format, not the SCF, is what the fixture pins.

Usage (in a container with the reference tree, after `make -C oracle`):
    python tests/golden/make_molden_golden.py
"""
from __future__ import annotations

import json
import os
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

import dollarstorescf_amd as ds  # noqa: E402
from dollarstorescf_amd import geometry as G  # noqa: E402
from oracle import ref as R  # noqa: E402

# values at or near the formatter's rounding and notation boundaries
SPECIAL = [0.0, 1.0, -1.0, 0.5, -0.5, 1e-5, -1e-5, 9.99999999e-1, 9.999999995e-1, 0.999999994999, 123.456789,
           -3.14159265358979, 1e-10, 5e-9, 1e-4, 1.00000001e-4, 9.99999995e-5, 10000.0, 99999.9999, 1e-300,
           2.5e-8, 0.12345678949, 0.1234567895, -7.77777777777]


def cases():
    Zw, xw = G.water()
    Z2, x2 = G.water_box(2, seed=3)
    Za, xa = G.alkane(2)
    return [("water_sto3g", "STO-3G", Zw, xw, True), ("water2_def2svp", "def2-SVP", Z2, x2, False),
            ("ethane_631gp", "6-31GP", Za, xa, False)]


def main():
    out = {}
    rng = np.random.default_rng(11)
    for name, basis, Z, xyz, virt in cases():
        rm = R.RefMolecule(ds.basis_path(basis), Z, xyz)
        n = rm.n
        nocc = (int(np.sum(Z))) // 2
        k = n if virt else nocc
        C = np.zeros((n, n))
        C[:, :k] = rng.standard_normal((n, k)) * rng.choice([1e-3, 1e-1, 1.0, 3.0], size=(n, k))
        m = min(n, len(SPECIAL))
        C[:m, 0] = SPECIAL[:m]
        eps = np.sort(rng.standard_normal(n) * 5.0)
        with tempfile.TemporaryDirectory() as d:
            text = rm.molden(os.path.join(d, "ref.molden"), C, eps, nocc, virt)
        out[name] = {"basis": basis, "Z": [int(z) for z in Z], "xyz_bohr": np.asarray(xyz).reshape(-1).tolist(),
                     "nocc": nocc, "write_virtual": virt, "C_columns": C[:, :k].tolist(), "eps": eps.tolist(),
                     "molden": text}
    with open(os.path.join(HERE, "molden.json"), "w") as f:
        json.dump(out, f, indent=0)
        f.write("\n")


if __name__ == "__main__":
    main()
