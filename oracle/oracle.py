"""Oracle: CPU restatement of DollarStoreSCF's SCF hot path (test infrastructure).

TEST INFRASTRUCTURE ONLY.  Only ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` may import this module, and only as the
checker.  The product (``dollarstorescf_amd``) never imports it.

Contents, each restating the cited reference code:

* ``read_basis``           -- Gaussian::readBasis / readAtomicBasis
                              (src/chem/BasisReader.cpp:9-113); Fortran ``D``
                              exponents as TextReader::readFloat (src/util/TextReader.cpp:57-82)
* ``read_xyz``             -- readXYZFile (src/chem/XYZReader.cpp:6-35), A_to_a0
                              (src/chem/Constants.hpp:15)
* ``solid_harmonic_terms`` -- ContractedGaussian ctor (src/chem/GTO.cpp:15-88)
* ``normalize_primitives`` -- GTO.cpp:90-93
* ``build_aos``            -- scf::constructBasis AO order (src/chem/scf/SCFBasis.cpp:120-140)
* ``OracleBasis``          -- ctypes view onto oracle.c (ERI, 1e integrals, tensor, J/K)
* ``DIIS``                 -- flo::DIISSolver (src/lalib/DIIS.ipp:5-100) with the
                              Pulay error of scf::diis (src/chem/scf/SCFDIIS.cpp:19-52)
* ``run_scf``              -- scf::solveMOs / runSCFIteration / calculateEnergies
                              (src/chem/scf/SCFSolver.cpp:45-237, src/chem/scf/Energy.cpp:21-78)
* ``run_uhf``              -- the same for multiplicity > 1 (the unrestricted branches of
                              SCFSolver.cpp:77-172, SCFDIIS.cpp:11-52, Energy.cpp:21-50)

Pinning is in tests/test_oracle.py (reference build under oracle/_ref, golden
fixtures under tests/golden, closed forms and invariants).
"""
from __future__ import annotations

import ctypes
import math
import os
from dataclasses import dataclass, field

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
PI = 3.14159265358979          # src/chem/Constants.hpp:5
A_TO_A0 = 1.889726126          # src/chem/Constants.hpp:15

ELEMENTS = ["H", "He", "Li", "Be", "B", "C", "N", "O", "F", "Ne",
            "Na", "Mg", "Al", "Si", "P", "S", "Cl", "Ar"]


def element_number(sym: str) -> int:
    """getElement (src/chem/Element.cpp:178-185) restricted to Z <= 18."""
    if 0 < len(sym) < 3 and sym in ELEMENTS:
        return ELEMENTS.index(sym) + 1
    return 0


# ------------------------------------------------------------------ parsing --
def _read_float(tok: str) -> float:
    return float(tok.replace("D", "E").replace("d", "E").replace("e", "E"))


def read_basis(path: str) -> dict:
    """Return {Z: [(l, [(zeta, unnormalized_weight), ...]), ...]}.

    BasisReader.cpp:98-113: a line ``El 0`` opens an element; shells follow as
    ``S|P|D|...|SP nprim scale`` with nprim primitive lines; ``****`` ends it.
    ``SP`` becomes an S shell followed by a P shell sharing exponents
    (BasisReader.cpp:84-93).  ECP blocks are skipped (:13)."""
    lines = open(path).read().splitlines()
    out: dict = {}
    am = {"S": 0, "P": 1, "D": 2, "F": 3, "G": 4, "H": 5, "I": 6, "J": 7}
    r = 0
    while r < len(lines):
        tok = lines[r].split()
        z = element_number(tok[0]) if tok else 0
        if z and len(tok) > 1 and tok[1].startswith("0"):
            r += 1
            first = lines[r].split() if r < len(lines) else []
            if first and first[0].endswith("ECP"):
                r += 1
                continue
            shells = []
            while r < len(lines):
                t = lines[r].split()
                if not t or t[0].startswith("*"):
                    break
                typ, nprim = t[0], int(t[1])
                if not (1 <= nprim <= 100):
                    raise ValueError("primitive count outside 1-100")
                s0, s1 = [], []
                for p in range(nprim):
                    r += 1
                    v = lines[r].split()
                    zeta = _read_float(v[0])
                    if zeta <= 0.0:
                        raise ValueError("non-positive exponent")
                    s0.append((zeta, _read_float(v[1])))
                    if typ == "SP":
                        s1.append((zeta, _read_float(v[2])))
                r += 1
                if typ == "SP":
                    shells.append((0, s0))
                    shells.append((1, s1))
                else:
                    shells.append((am[typ], s0))
            out[z] = shells
        r += 1
    return out


def read_xyz(path: str):
    """readXYZFile: count line, comment line, then ``El x y z`` in Angstrom."""
    lines = open(path).read().splitlines()
    n = int(lines[0].split()[0])
    Z, xyz = [], []
    for ln in lines[2:2 + n]:
        t = ln.split()
        Z.append(element_number(t[0]))
        xyz.append([float(t[1]) * A_TO_A0, float(t[2]) * A_TO_A0, float(t[3]) * A_TO_A0])
    return np.array(Z, dtype=np.int32), np.array(xyz, dtype=np.float64)


# ---------------------------------------------------------- GTO construction --
def _factorial(n: int) -> float:
    r = 1
    for i in range(n, 0, -1):
        r *= i
    return float(r)


def _double_factorial_f(x: float) -> float:
    r = 1.0
    i = x
    while i > 0:
        r *= i
        i -= 2.0
    return r


def _binomial(n: float, k: float) -> float:
    r = 1.0
    i = k
    while i > 0:
        r /= i
        i -= 1.0
    i = n
    while i > n - k:
        r *= i
        i -= 1.0
    return r


def solid_harmonic_terms(l: int, m: int):
    """Cartesian expansion [w, x, y, z] of the reference's real solid harmonic
    (l, m) including its normalization (GTO.cpp:15-88)."""
    am = abs(m)
    terms = []
    i, x = 0, (am if m <= 0 else am - 1)
    while x >= 0:
        terms.append([(-1.0 if i & 1 else 1.0) * _binomial(float(am), float(x)), x, am - x, 0])
        x -= 2
        i += 1
    leg = _double_factorial_f(2.0 * am - 1.0)
    if m < 0:
        leg *= _factorial(l - am) / _factorial(l + am)
    for t in terms:
        t[0] *= leg
    ynorm = math.sqrt(float(2 * l + 1) * _factorial(l - m) / (4.0 * PI * _factorial(l + m)))
    if m:
        ynorm *= math.sqrt(2.0)
    if l > am:
        prev = [list(t) for t in terms]
        cur = terms
        for t in cur:
            t[0] *= float(2 * am + 1)
            t[3] += 1
        for ii in range(am + 1, l):
            sf = float(ii - am + 1)
            nxt = []
            for t in cur:
                nxt.append([t[0] * (float(2 * ii + 1) / sf), t[1], t[2], t[3] + 1])
            for t in prev:
                f = -float(ii + am) / sf
                nxt.append([t[0] * f, t[1] + 2, t[2], t[3]])
                nxt.append([t[0] * f, t[1], t[2] + 2, t[3]])
                nxt.append([t[0] * f, t[1], t[2], t[3] + 2])
            j = 0
            while j < len(nxt):
                k = j + 1
                while k < len(nxt):
                    if nxt[j][1:] == nxt[k][1:]:
                        nxt[j][0] += nxt[k][0]
                        del nxt[k]
                    else:
                        k += 1
                j += 1
            prev, cur = cur, nxt
        terms = cur
    for t in terms:
        t[0] *= ynorm
    return terms


def normalize_primitives(l: int, prims):
    """GTO.cpp:90-93: w = c * sqrt((2 zeta)^(l+3/2) 2^(l+2) / (sqrt(pi) (2l+1)!!))."""
    out = []
    for zeta, c in prims:
        nf = (math.pow(2.0 * zeta, l + 1.5) * math.pow(2.0, l + 2)) / (math.sqrt(PI) * _double_factorial_f(2.0 * l + 1.0))
        out.append((zeta, c * math.sqrt(nf)))
    return out


@dataclass
class AO:
    l: int
    m: int
    center: tuple
    prims: list               # [(zeta, normalized weight)]
    terms: list               # [[w, x, y, z]]


def cartesian_components(l: int):
    """(a, b, c) with a + b + c = l in x-major order (xx, xy, xz, yy, yz, zz for d)."""
    return [(a, b, l - a - b) for a in range(l, -1, -1) for b in range(l - a, -1, -1)]


def cartesian_norm(l: int, a: int, b: int, c: int) -> float:
    """Weight of x^a y^b z^c that normalizes one primitive carrying the
    reference's primitive weight (GTO.cpp:90-93, which normalizes r^l times a
    unit real solid harmonic): sqrt((2l+1)!! / (4 pi (2a-1)!! (2b-1)!! (2c-1)!!)).
    For l <= 1 it is the real solid harmonic's own weight."""
    den = _double_factorial_f(2.0 * a - 1.0) * _double_factorial_f(2.0 * b - 1.0) * _double_factorial_f(2.0 * c - 1.0)
    return math.sqrt(_double_factorial_f(2.0 * l + 1.0) / (4.0 * PI * den))


def build_aos(basis: dict, Z, xyz, cartesian: bool = False):
    """AO list in atom -> shell (file order) -> m=-l..l order (SCFBasis.cpp:120-140).

    cartesian: shells with l >= 2 become ncart(l) Cartesian functions x^a y^b z^c
    (x-major order, m = component index), each primitive normalized
    (cartesian_norm), the 6D convention of 6-31G*; s and p shells are the
    reference's own.  No reference counterpart (the reference builds real
    solid harmonics only)."""
    aos = []
    for a in range(len(Z)):
        shells = basis.get(int(Z[a]))
        if not shells:
            raise ValueError(f"no basis for Z={Z[a]}")
        for l, prims in shells:
            np_ = normalize_primitives(l, prims)
            c = tuple(float(v) for v in xyz[a])
            if cartesian and l >= 2:
                for k, (x, y, z) in enumerate(cartesian_components(l)):
                    aos.append(AO(l, k, c, np_, [[cartesian_norm(l, x, y, z), x, y, z]]))
                continue
            for m in range(-l, l + 1):
                aos.append(AO(l, m, c, np_, solid_harmonic_terms(l, m)))
    return aos


# ------------------------------------------------------------ C restatement --
class _OrcBasis(ctypes.Structure):
    _fields_ = [("n_ao", ctypes.c_int),
                ("center", ctypes.c_void_p), ("l", ctypes.c_void_p),
                ("prim_off", ctypes.c_void_p), ("nprim", ctypes.c_void_p),
                ("zeta", ctypes.c_void_p), ("weight", ctypes.c_void_p),
                ("term_off", ctypes.c_void_p), ("nterm", ctypes.c_void_p),
                ("term_w", ctypes.c_void_p), ("term_xyz", ctypes.c_void_p)]


_LIB = None


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(HERE, "_build", "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle not built: run `make -C oracle`")
        L = ctypes.CDLL(path)
        d, i, p = ctypes.c_double, ctypes.c_int, ctypes.c_void_p
        L.orc_boys_ref.restype = d
        L.orc_boys_ref.argtypes = [d, ctypes.c_uint]
        L.orc_boys_accurate.restype = d
        L.orc_boys_accurate.argtypes = [d, ctypes.c_uint]
        L.orc_eri.restype = d
        L.orc_eri.argtypes = [p, i, i, i, i, i]
        L.orc_eri_list.argtypes = [p, p, ctypes.c_long, i, p]
        L.orc_build_tensor.argtypes = [p, i, p]
        L.orc_coulomb.argtypes = [p, i, p, p]
        L.orc_exchange.argtypes = [p, i, p, p]
        L.orc_one_electron.argtypes = [p, i, p, p, i, p, p, p]
        L.orc_set_threads.argtypes = [i]
        L.orc_set_threads.restype = None
        L.orc_max_threads.restype = i
        _LIB = L
    return _LIB


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(ctypes.c_void_p)


class OracleBasis:
    """Flat AO description handed to oracle.c."""

    def __init__(self, aos, Z=None, xyz=None):
        self.aos = aos
        self.n = len(aos)
        self.Z = np.ascontiguousarray(Z, dtype=np.int32) if Z is not None else None
        self.xyz = np.ascontiguousarray(xyz, dtype=np.float64) if xyz is not None else None
        self.center = np.array([a.center for a in aos], dtype=np.float64).reshape(-1)
        self.l = np.array([a.l for a in aos], dtype=np.int32)
        self.nprim = np.array([len(a.prims) for a in aos], dtype=np.int32)
        self.prim_off = np.concatenate([[0], np.cumsum(self.nprim)[:-1]]).astype(np.int32)
        self.zeta = np.array([p[0] for a in aos for p in a.prims], dtype=np.float64)
        self.weight = np.array([p[1] for a in aos for p in a.prims], dtype=np.float64)
        self.nterm = np.array([len(a.terms) for a in aos], dtype=np.int32)
        self.term_off = np.concatenate([[0], np.cumsum(self.nterm)[:-1]]).astype(np.int32)
        self.term_w = np.array([t[0] for a in aos for t in a.terms], dtype=np.float64)
        self.term_xyz = np.array([t[1:] for a in aos for t in a.terms], dtype=np.int32).reshape(-1)
        self._s = _OrcBasis(self.n, *[ctypes.cast(_ptr(x), ctypes.c_void_p) for x in (
            self.center, self.l, self.prim_off, self.nprim, self.zeta, self.weight,
            self.term_off, self.nterm, self.term_w, self.term_xyz)])

    @property
    def ref(self):
        return ctypes.byref(self._s)

    def eri(self, i, j, k, l, mode=0) -> float:
        """(ij|kl); mode 0 = reference Boys series, 1 = accurate Boys."""
        return lib().orc_eri(self.ref, i, j, k, l, mode)

    def eri_list(self, quartets, mode=0) -> np.ndarray:
        q = np.ascontiguousarray(quartets, dtype=np.int32).reshape(-1, 4)
        out = np.empty(len(q), dtype=np.float64)
        lib().orc_eri_list(self.ref, _ptr(q), len(q), mode, _ptr(out))
        return out

    def tensor(self, mode=0) -> np.ndarray:
        D = self.n * (self.n + 1) // 2
        t = np.empty(D * (D + 1) // 2, dtype=np.float64)
        lib().orc_build_tensor(self.ref, mode, _ptr(t))
        return t

    def one_electron(self, mode=0):
        D = self.n * (self.n + 1) // 2
        S, T, V = (np.empty(D) for _ in range(3))
        lib().orc_one_electron(self.ref, len(self.Z), _ptr(self.Z), _ptr(self.xyz), mode, _ptr(S), _ptr(T), _ptr(V))
        return S, T, V


def set_threads(n: int) -> None:
    """OpenMP threads of the tensor fill and the J/K loops (1 = the reference's
    single-threaded path)."""
    lib().orc_set_threads(int(n))


def max_threads() -> int:
    return lib().orc_max_threads()


def coulomb(tensor, n, Ptot_packed):
    J = np.empty(n * (n + 1) // 2)
    lib().orc_coulomb(_ptr(tensor), n, _ptr(np.ascontiguousarray(Ptot_packed)), _ptr(J))
    return J


def exchange(tensor, n, Pa_packed):
    K = np.empty(n * (n + 1) // 2)
    lib().orc_exchange(_ptr(tensor), n, _ptr(np.ascontiguousarray(Pa_packed)), _ptr(K))
    return K


# ------------------------------------------------------------ packed helpers --
def unpack(p: np.ndarray, n: int) -> np.ndarray:
    """Packed-U (element (y,x), y<=x at x(x+1)/2+y; src/lalib/Matrix.ipp:332-351) -> full."""
    M = np.zeros((n, n))
    iu = np.triu_indices(n)
    # column-major upper packing: order by column x then row y
    ys, xs = iu[0], iu[1]
    idx = xs * (xs + 1) // 2 + ys
    M[ys, xs] = p[idx]
    M[xs, ys] = p[idx]
    return M


def pack(M: np.ndarray) -> np.ndarray:
    n = M.shape[0]
    ys, xs = np.triu_indices(n)
    p = np.empty(n * (n + 1) // 2)
    p[xs * (xs + 1) // 2 + ys] = M[ys, xs]
    return p


# ------------------------------------------------------------------- SCF ----
class DIIS:
    """flo::DIISSolver (DIIS.ipp:5-100): at most `size` vectors, the one with
    the largest B_ii replaced when full, B bordered by ones, B c = (0..0,1)."""

    def __init__(self, size: int):
        self.size = size
        self.err: list = []
        self.res: list = []
        self.B = np.zeros((size + 1, size + 1))
        self.rms = 0.0
        self.c = None

    def add(self, F: np.ndarray, e: np.ndarray):
        new_rms = float(np.sum(e * e))
        self.rms = new_rms
        if len(self.err) < self.size:
            index = len(self.err)
            self.err.append(None)
            self.res.append(None)
            vc = len(self.err)
            self.B[:vc, vc] = 1.0
            self.B[vc, :vc] = 1.0
            self.B[vc, vc] = 0.0
        else:
            vc = len(self.err)
            index, mx = 0, 0.0
            for i in range(self.size):
                if self.B[i, i] > mx:
                    mx, index = self.B[i, i], i
        self.err[index] = e
        self.res[index] = F
        self._last = index
        self.B[index, index] = new_rms
        for i in range(vc):
            if i != index:
                v = float(np.sum(self.err[i] * e))
                self.B[i, index] = self.B[index, i] = v

    def solve(self):
        vc = len(self.err)
        rhs = np.zeros(vc + 1)
        rhs[vc] = 1.0
        try:
            self.c = np.linalg.solve(self.B[:vc + 1, :vc + 1], rhs)
        except np.linalg.LinAlgError:
            # Singular B (identical error vectors, e.g. a one-AO system): the
            # reference's dspsv_ reports info>0 and leaves no usable solution;
            # here the newest Fock matrix is kept instead.
            self.c = np.zeros(vc + 1)
            self.c[self._last] = 1.0

    def result(self):
        if len(self.res) <= 1:
            return self.res[0]
        out = self.c[0] * self.res[0]
        for i in range(1, len(self.res)):
            out = out + self.c[i] * self.res[i]
        return out


@dataclass
class SCFResult:
    energy: float
    electronic_energy: float
    iterations: int
    history: list = field(default_factory=list)
    components: dict = field(default_factory=dict)
    Pa: np.ndarray | None = None


def run_scf(S, T, V, jk, n_alpha: int, e_nuc: float, diis_size=8, conv=1e-8, max_iter=500):
    """RHF as scf::solveMOs (SCFSolver.cpp:196-237).

    ``jk(Ptot, Pa) -> (J, K)`` full matrices.  Core guess P=0
    (useCoreGuess, :179-194); F = H + J[2Pa] - K[Pa] (:29-62); DIIS from
    iteration 2 with error S Pa F - (S Pa F)^T (SCFDIIS.cpp:19-25); Loewdin
    X = U s^-1/2 U^T (SCFBasis.cpp:85-100); E_elec = 2 sum eps_occ - tr(Pa G)
    with G from the previous density (:118-143); stop when |dE_elec| < conv
    (:174-177).  Final energy as calculateEnergies (Energy.cpp:21-50)."""
    H = T + V
    s, U = np.linalg.eigh(S)
    X = (U * (1.0 / np.sqrt(s))) @ U.T
    n = S.shape[0]
    Pa = np.zeros((n, n))
    diis = DIIS(diis_size)
    it, E, Eprev = 0, 0.0, 0.0
    hist = []
    while not (it and abs(E - Eprev) < conv):
        if it >= max_iter:
            raise RuntimeError("SCF did not converge")
        it += 1
        J, K = jk(2.0 * Pa, Pa)
        G = J - K
        F0 = H + G
        if it == 1:
            F = F0
        else:
            spf = S @ (Pa @ F0)
            diis.add(F0, spf - spf.T)
            diis.solve()
            F = diis.result()
        Fp = X @ F @ X
        eps, Cp = np.linalg.eigh(Fp)
        C = X @ Cp
        Pa = C[:, :n_alpha] @ C[:, :n_alpha].T
        Eprev, E = E, 2.0 * float(np.sum(eps[:n_alpha])) - float(np.sum(Pa * G))
        hist.append(E)
    Ptot = 2.0 * Pa
    J, K = jk(Ptot, Pa)
    comps = dict(kinetic=float(np.sum(Ptot * T)), nuclear_attraction=float(np.sum(Ptot * V)),
                 coulomb=float(np.sum(Pa * J)), exchange=-float(np.sum(Pa * K)), nuclear=e_nuc)
    return SCFResult(sum(comps.values()), E, it, hist, comps, Pa)


def run_uhf(S, T, V, jk_u, n_alpha: int, n_beta: int, e_nuc: float, diis_size=8, conv=1e-8, max_iter=500):
    """UHF as scf::solveMOs with spin_treatment unrestricted (SCFSolver.cpp:
    145-237; multiplicity > 1 selects it, SCFBasis.cpp:29-34).

    ``jk_u(Ptot, Pa, Pb) -> (J, Ka, Kb)`` full matrices.  Core guess Pa = Pb =
    0 (:179-194).  Per iteration, from the densities at its start: G_s =
    J[Pa + Pb] - K[P_s] for s = alpha, beta (computeCoulombTerms /
    computeExchangeTerms per spin, :29-43, :145-159), F0_s = H + G_s, DIIS
    from iteration 2 with one solver per spin and error S P_s F0_s - (S P_s
    F0_s)^T (SCFDIIS.cpp:11-52); then both Roothaan-Hall solutions and
    densities (:160-166); a spin without electrons keeps its density
    (:77-89).  E_elec = sum eps_a,occ + sum eps_b,occ - (tr(Pa Ga) +
    tr(Pb Gb)) / 2 (:130-141).  Final energies as calculateEnergies
    (Energy.cpp:21-50): Coulomb tr(Ptot J)/2, exchange -(tr(Pa Ka) +
    tr(Pb Kb))/2."""
    H = T + V
    s, U = np.linalg.eigh(S)
    X = (U * (1.0 / np.sqrt(s))) @ U.T
    n = S.shape[0]
    P = [np.zeros((n, n)), np.zeros((n, n))]
    occ = [n_alpha, n_beta]
    diis = [DIIS(diis_size), DIIS(diis_size)]
    it, E, Eprev = 0, 0.0, 0.0
    hist = []
    while not (it and abs(E - Eprev) < conv):
        if it >= max_iter:
            raise RuntimeError("SCF did not converge")
        it += 1
        J, Ka, Kb = jk_u(P[0] + P[1], P[0], P[1])
        G = [J - Ka, J - Kb]
        F = []
        for sp in range(2):
            F0 = H + G[sp]
            if it == 1:
                F.append(F0)
            else:
                spf = S @ (P[sp] @ F0)
                diis[sp].add(F0, spf - spf.T)
                diis[sp].solve()
                F.append(diis[sp].result())
        eps = []
        for sp in range(2):
            e_s, Cp = np.linalg.eigh(X @ F[sp] @ X)
            C = X @ Cp
            eps.append(e_s)
            if occ[sp]:
                P[sp] = C[:, :occ[sp]] @ C[:, :occ[sp]].T
        Eprev = E
        E = (float(np.sum(eps[0][:n_alpha])) + float(np.sum(eps[1][:n_beta]))
             - 0.5 * (float(np.sum(P[0] * G[0])) + float(np.sum(P[1] * G[1]))))
        hist.append(E)
    Ptot = P[0] + P[1]
    J, Ka, Kb = jk_u(Ptot, P[0], P[1])
    comps = dict(kinetic=float(np.sum(Ptot * T)), nuclear_attraction=float(np.sum(Ptot * V)),
                 coulomb=0.5 * float(np.sum(Ptot * J)),
                 exchange=-0.5 * (float(np.sum(P[0] * Ka)) + float(np.sum(P[1] * Kb))), nuclear=e_nuc)
    r = SCFResult(sum(comps.values()), E, it, hist, comps, P[0])
    r.Pb = P[1]
    return r


def nuclear_repulsion(Z, xyz) -> float:
    """Energy.cpp:43-48 (element numbers as charges)."""
    e = 0.0
    for i in range(len(Z)):
        for j in range(i + 1, len(Z)):
            e += int(Z[i]) * int(Z[j]) / float(np.linalg.norm(np.asarray(xyz[i]) - np.asarray(xyz[j])))
    return e


def tensor_jk(tensor, n):
    """J/K from the packed tensor with the reference's loops (ExactCoulomb.cpp:103-123)."""
    def jk(Ptot, Pa):
        return unpack(coulomb(tensor, n, pack(Ptot)), n), unpack(exchange(tensor, n, pack(Pa)), n)
    return jk


def tensor_jk_u(tensor, n):
    """Unrestricted J/K from the packed tensor (the same loops per spin)."""
    def jk(Ptot, Pa, Pb):
        return (unpack(coulomb(tensor, n, pack(Ptot)), n), unpack(exchange(tensor, n, pack(Pa)), n),
                unpack(exchange(tensor, n, pack(Pb)), n))
    return jk


def scf_energy(basis_path: str, Z, xyz, charge=0, mult=1, mode=0, diis_size=8, conv=1e-8, cartesian=False):
    """Whole oracle single point: basis, AOs, 1e, packed tensor, RHF
    (multiplicity 1) or UHF (SCFBasis.cpp:17-37: n_alpha = (N + mult) / 2,
    n_beta = (N + 1 - mult) / 2)."""
    bas = read_basis(basis_path)
    aos = build_aos(bas, Z, xyz, cartesian)
    ob = OracleBasis(aos, Z, xyz)
    S, T, V = (unpack(x, ob.n) for x in ob.one_electron(mode))
    t = ob.tensor(mode)
    ne = int(np.sum(Z)) - charge
    n_alpha = (ne + mult) // 2
    if mult > 1:
        return run_uhf(S, T, V, tensor_jk_u(t, ob.n), n_alpha, (ne + 1 - mult) // 2, nuclear_repulsion(Z, xyz),
                       diis_size, conv)
    return run_scf(S, T, V, tensor_jk(t, ob.n), n_alpha, nuclear_repulsion(Z, xyz), diis_size, conv)
