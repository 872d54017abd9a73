"""ctypes view of oracle/_ref/libdsscf_ref.so (DollarStoreSCF's own integral
code built from its sources by oracle/Makefile).  TEST INFRASTRUCTURE ONLY:
used to pin the oracle and as the timed reference CPU baseline."""
from __future__ import annotations

import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
PATH = os.path.join(HERE, "_ref", "libdsscf_ref.so")

_LIB = None


def available() -> bool:
    return os.path.exists(PATH)


def lib():
    global _LIB
    if _LIB is None:
        L = ctypes.CDLL(PATH)
        d, i, p = ctypes.c_double, ctypes.c_int, ctypes.c_void_p
        L.ref_boys.restype = d
        L.ref_boys.argtypes = [d, ctypes.c_uint]
        L.ref_load_basis.restype = i
        L.ref_load_basis.argtypes = [ctypes.c_char_p]
        L.ref_build.restype = i
        L.ref_build.argtypes = [i, p, p]
        L.ref_ao.argtypes = [i, p, p, p, p, p, p, p, p, i]
        L.ref_eri.restype = d
        L.ref_eri.argtypes = [i, i, i, i]
        L.ref_eri_list.argtypes = [p, ctypes.c_long, p]
        L.ref_time_eri_list.restype = d
        L.ref_time_eri_list.argtypes = [p, ctypes.c_long, p]
        L.ref_one_electron.argtypes = [p, p, p]
        L.ref_gradient.argtypes = [p, p, p, i, p]
        L.ref_eri_gradient.argtypes = [i, i, i, i, p]
        L.ref_time_eri_gradient.restype = d
        L.ref_time_eri_gradient.argtypes = [p, ctypes.c_long, p]
        L.ref_molden.restype = i
        L.ref_molden.argtypes = [ctypes.c_char_p, p, p, i, i]
        _LIB = L
    return _LIB


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


class RefMolecule:
    """Reference AOs for one molecule/basis (single-threaded; the reference's
    integral code keeps global scratch)."""

    def __init__(self, basis_path: str, Z, xyz_bohr):
        L = lib()
        L.ref_load_basis(basis_path.encode())
        self.Z = np.ascontiguousarray(Z, dtype=np.int32)
        self.xyz = np.ascontiguousarray(xyz_bohr, dtype=np.float64).reshape(-1)
        self.n = L.ref_build(len(self.Z), _p(self.Z), _p(self.xyz))
        if self.n < 0:
            raise ValueError("missing basis for an element")

    def rebuild(self):
        lib().ref_build(len(self.Z), _p(self.Z), _p(self.xyz))

    def ao(self, i: int):
        cap = 64
        l = np.zeros(1, np.int32)
        c = np.zeros(3)
        npr = np.zeros(1, np.int32)
        z = np.zeros(cap)
        w = np.zeros(cap)
        nt = np.zeros(1, np.int32)
        tw = np.zeros(cap)
        tx = np.zeros(3 * cap, np.int32)
        lib().ref_ao(i, _p(l), _p(c), _p(npr), _p(z), _p(w), _p(nt), _p(tw), _p(tx), cap)
        return dict(l=int(l[0]), center=c, zeta=z[:npr[0]].copy(), weight=w[:npr[0]].copy(),
                    terms=[(tw[t], *tx[3 * t:3 * t + 3]) for t in range(nt[0])])

    def eri_list(self, quartets) -> np.ndarray:
        q = np.ascontiguousarray(quartets, dtype=np.int32).reshape(-1, 4)
        out = np.empty(len(q))
        lib().ref_eri_list(_p(q), len(q), _p(out))
        return out

    def time_eri_list(self, quartets):
        q = np.ascontiguousarray(quartets, dtype=np.int32).reshape(-1, 4)
        out = np.empty(len(q))
        t = lib().ref_time_eri_list(_p(q), len(q), _p(out))
        return t, out

    def one_electron(self):
        D = self.n * (self.n + 1) // 2
        S, T, V = np.empty(D), np.empty(D), np.empty(D)
        lib().ref_one_electron(_p(S), _p(T), _p(V))
        return S, T, V


    def gradient(self, Pt, Pa, W, part: int = 7):
        """getEnergyGradient restated over the reference's derivative integrals:
        part 1 one-electron, 2 two-electron, 4 nuclear repulsion (sums allowed)."""
        self.rebuild()
        g = np.empty(3 * len(self.Z))
        arrs = [np.ascontiguousarray(a, dtype=np.float64) for a in (Pt, Pa, W)]
        lib().ref_gradient(_p(arrs[0]), _p(arrs[1]), _p(arrs[2]), part, _p(g))
        return g.reshape(-1, 3)

    def eri_gradient(self, i, j, k, l):
        out = np.empty(12)
        lib().ref_eri_gradient(i, j, k, l, _p(out))
        return out.reshape(4, 3)

    def time_eri_gradient(self, quartets) -> float:
        q = np.ascontiguousarray(quartets, dtype=np.int32).reshape(-1, 4)
        sink = np.zeros(1)
        self.rebuild()
        return lib().ref_time_eri_gradient(_p(q), len(q), _p(sink))

    def molden(self, path: str, C, eps, nocc: int, write_virtual: bool):
        """The reference's writeOrbitalsMolden text for orbitals C (n x n,
        column i = orbital i) and energies eps, through its own FormattedStream."""
        C = np.ascontiguousarray(C, dtype=np.float64)
        eps = np.ascontiguousarray(eps, dtype=np.float64)
        self.rebuild()
        if lib().ref_molden(path.encode(), _p(C), _p(eps), nocc, 1 if write_virtual else 0) != 0:
            raise OSError("ref_molden: cannot write " + path)
        with open(path) as f:
            return f.read()


def boys(x: float, m: int) -> float:
    return lib().ref_boys(x, m)
