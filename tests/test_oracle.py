"""Pins the oracle (CPU restatement) before anything is checked against it.

Each check is chosen so that a plausible mistake fails one of them:
reference-build values (oracle/_ref and the golden fixtures it wrote),
closed forms, mpmath, symmetry/invariance of the ERIs, literature energies.
"""
import json
import math
import os

import numpy as np
import pytest

from dollarstorescf_amd import geometry as G
from oracle import oracle as O
from oracle import ref as R

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = os.path.join(HERE, "golden")
BASIS = os.path.join(os.path.dirname(HERE), "dollarstorescf_amd", "basis")


def gold(name):
    with open(os.path.join(GOLD, name)) as f:
        return json.load(f)


def oracle_basis(basis, Z, xyz):
    Z = np.asarray(Z)
    xyz = np.asarray(xyz, float)
    return O.OracleBasis(O.build_aos(O.read_basis(os.path.join(BASIS, basis)), Z, xyz), Z, xyz)


# --------------------------------------------------------------- Boys ----
def test_boys_reference_series_matches_reference_build():
    g = gold("boys.json")
    for x, row in zip(g["x"], g["reference_series"]):
        for m, v in enumerate(row):
            assert O.lib().orc_boys_ref(x, m) == v  # bit-for-bit


def test_boys_accurate_vs_mpmath():
    g = gold("boys.json")
    for x, row in zip(g["x"], g["mpmath"]):
        for m, v in enumerate(row):
            assert abs(O.lib().orc_boys_accurate(x, m) - v) <= 2e-15 * max(v, 1e-300) + 1e-300


def test_boys_closed_forms():
    L = O.lib()
    for m in range(9):
        assert L.orc_boys_ref(0.0, m) == pytest.approx(1.0 / (2 * m + 1), abs=1e-15)
    for x in (0.1, 0.7, 3.0, 12.0, 22.0):
        f0 = math.sqrt(math.pi) * math.erf(math.sqrt(x)) / (2 * math.sqrt(x))
        assert L.orc_boys_accurate(x, 0) == pytest.approx(f0, rel=1e-14)
        # downward recursion F_m = (2x F_{m+1} + e^-x)/(2m+1)
        for m in range(6):
            lhs = L.orc_boys_accurate(x, m)
            rhs = (2 * x * L.orc_boys_accurate(x, m + 1) + math.exp(-x)) / (2 * m + 1)
            assert lhs == pytest.approx(rhs, rel=1e-13)


def test_boys_reference_error_band():
    """The reference's truncated series undershoots (all-positive terms
    dropped, up to rounding); its absolute error stays below 3e-11."""
    g = gold("boys.json")
    worst = 0.0
    for x, rs, acc in zip(g["x"], g["reference_series"], g["mpmath"]):
        for m in range(9):
            if x <= 1.1 * m + 25.0:
                assert rs[m] <= acc[m] * (1 + 1e-13)
            worst = max(worst, abs(rs[m] - acc[m]))
    assert worst < 3e-11


# ---------------------------------------------------------------- ERI ----
@pytest.mark.parametrize("key", ["water_sto3g", "water_631g", "water_631gs", "water_def2svp", "water2_631g"])
def test_eri_matches_reference_fixture(key):
    c = gold("reference_integrals.json")[key]
    ob = oracle_basis(c["basis"], c["Z"], c["xyz_bohr"])
    assert ob.n == c["n_ao"]
    got = ob.eri_list(np.array(c["quartets"], np.int32), mode=0)
    assert np.max(np.abs(got - np.array(c["eri"]))) < 1e-13


def test_eri_live_reference_build(ref_available):
    if not ref_available:
        pytest.skip("oracle/_ref not built (no reference tree)")
    Z, xyz = G.water_box(2, seed=5)
    ob = oracle_basis("def2-SVP", Z, xyz)
    rm = R.RefMolecule(os.path.join(BASIS, "def2-SVP"), Z, xyz)
    q = np.random.default_rng(4).integers(0, ob.n, size=(300, 4)).astype(np.int32)
    assert np.max(np.abs(ob.eri_list(q, 0) - rm.eri_list(q))) < 1e-13


def test_eri_ssss_closed_form():
    """(ss|ss) over one primitive each: 2 pi^2.5/(pq sqrt(p+q)) K_AB K_CD F0(T) * norms."""
    prim = {1: [(0, [(1.3, 1.0)])]}
    Z = np.array([1, 1])
    xyz = np.array([[0.0, 0.0, 0.0], [0.4, -0.3, 1.1]])
    aos = O.build_aos(prim, Z, xyz)
    ob = O.OracleBasis(aos, Z, xyz)
    a = 1.3
    w = aos[0].prims[0][1] * aos[0].terms[0][0]
    # (00|11): bra both on A, ket both on B
    p = q = 2 * a
    R2 = float(np.sum((xyz[0] - xyz[1]) ** 2))
    T = p * q / (p + q) * R2
    F0 = math.sqrt(math.pi) * math.erf(math.sqrt(T)) / (2 * math.sqrt(T))
    want = 2 * O.PI ** 2.5 / (p * q * math.sqrt(p + q)) * F0 * w ** 4
    assert ob.eri(0, 0, 1, 1, mode=1) == pytest.approx(want, rel=1e-13)


def test_eri_eightfold_symmetry_and_invariance():
    Z, xyz = G.water()
    ob = oracle_basis("6-31GP", Z, xyz)
    rng = np.random.default_rng(9)
    shift = np.array([1.7, -2.2, 0.4])
    th = 0.7
    Rz = np.array([[math.cos(th), -math.sin(th), 0], [math.sin(th), math.cos(th), 0], [0, 0, 1]])
    ob_t = oracle_basis("6-31GP", Z, xyz + shift)
    for _ in range(60):
        i, j, k, l = (int(v) for v in rng.integers(0, ob.n, 4))
        v = ob.eri(i, j, k, l, 1)
        for (a, b, c, d) in [(j, i, k, l), (i, j, l, k), (k, l, i, j), (l, k, j, i)]:
            assert ob.eri(a, b, c, d, 1) == pytest.approx(v, abs=1e-14)
        assert ob_t.eri(i, j, k, l, 1) == pytest.approx(v, abs=1e-13)
    # rotation: the Coulomb energy of a density built only from s functions is invariant
    s_idx = [i for i, a in enumerate(ob.aos) if a.l == 0]
    ob_r = oracle_basis("6-31GP", Z, xyz @ Rz.T)
    for _ in range(20):
        i, j, k, l = (int(v) for v in rng.choice(s_idx, 4))
        assert ob_r.eri(i, j, k, l, 1) == pytest.approx(ob.eri(i, j, k, l, 1), abs=1e-13)
    # Schwarz inequality
    for _ in range(40):
        i, j, k, l = (int(v) for v in rng.integers(0, ob.n, 4))
        assert abs(ob.eri(i, j, k, l, 1)) <= math.sqrt(ob.eri(i, j, i, j, 1) * ob.eri(k, l, k, l, 1)) + 1e-14


# ---------------------------------------------------------- 1e terms -----
@pytest.mark.parametrize("key", ["water_sto3g", "water_def2svp"])
def test_one_electron_matches_reference_fixture(key):
    c = gold("reference_integrals.json")[key]
    ob = oracle_basis(c["basis"], c["Z"], c["xyz_bohr"])
    S, T, V = ob.one_electron(0)
    for got, want in ((S, c["S"]), (T, c["T"]), (V, c["V"])):
        assert np.max(np.abs(got - np.array(want))) < 1e-12


# -------------------------------------------------------------- J/K ------
def test_tensor_jk_matches_full_contraction():
    Z, xyz = G.water()
    ob = oracle_basis("6-31G", Z, xyz)
    n = ob.n
    t = ob.tensor(1)
    full = np.zeros((n, n, n, n))
    idx = [(i, j, k, l) for i in range(n) for j in range(n) for k in range(n) for l in range(n)]
    vals = ob.eri_list(np.array(idx, np.int32), 1)
    full[tuple(np.array(idx).T)] = vals
    C = np.random.default_rng(1).standard_normal((n, 4))
    Pa = C @ C.T
    J = O.unpack(O.coulomb(t, n, O.pack(2 * Pa)), n)
    K = O.unpack(O.exchange(t, n, O.pack(Pa)), n)
    assert np.allclose(J, np.einsum("mnls,ls->mn", full, 2 * Pa), atol=1e-12)
    assert np.allclose(K, np.einsum("mlns,ls->mn", full, Pa), atol=1e-12)


# --------------------------------------------------------------- SCF -----
def test_scf_literature_values():
    e = O.scf_energy(os.path.join(BASIS, "STO-3G"), np.array([2]), np.zeros((1, 3)), conv=1e-10).energy
    assert e == pytest.approx(-2.8077839575, abs=2e-8)   # He/STO-3G RHF
    e = O.scf_energy(os.path.join(BASIS, "STO-3G"), np.array([1, 1]), np.array([[0, 0, 0], [0, 0, 1.4]]), conv=1e-10).energy
    assert e == pytest.approx(-1.1167, abs=1e-4)          # Szabo-Ostlund H2/STO-3G at 1.4 a0


def test_scf_invariants_and_fixture():
    Z, xyz = G.water()
    r = O.scf_energy(os.path.join(BASIS, "STO-3G"), Z, xyz, conv=1e-10)
    assert r.energy == pytest.approx(gold("energies.json")["cases"]["water_sto3g"]["reference"]["energy"], abs=1e-10)
    ob = oracle_basis("STO-3G", Z, xyz)
    S = O.unpack(ob.one_electron(0)[0], ob.n)
    # idempotency P S P = P and electron count tr(Ptot S) = 10
    assert np.allclose(r.Pa @ S @ r.Pa, r.Pa, atol=1e-7)
    assert np.trace(2 * r.Pa @ S) == pytest.approx(10.0, abs=1e-9)


# --------------------------------------------------------- UHF (mult > 1) --
def _full_tensor(ob, mode=0):
    n = ob.n
    idx = np.array([(i, j, k, l) for i in range(n) for j in range(n) for k in range(n) for l in range(n)], np.int32)
    full = np.zeros((n, n, n, n))
    full[tuple(idx.T)] = ob.eri_list(idx, mode)
    return full


@pytest.mark.parametrize("basis", ["STO-3G", "6-31G"])
def test_uhf_one_electron_is_the_core_eigenproblem(basis):
    """One electron (H atom, doublet): no two-electron energy survives
    (J[Pa] = K[Pa] cancel in G_a and the beta spin is empty), so the UHF
    energy is the lowest generalized eigenvalue of (H, S); STO-3G gives the
    textbook -0.46658 Eh.  Fails a wrong sign of K, a beta density built
    from zero electrons, or a J/K from the wrong spin."""
    Z, xyz = np.array([1]), np.zeros((1, 3))
    r = O.scf_energy(os.path.join(BASIS, basis), Z, xyz, mult=2, conv=1e-12)
    ob = oracle_basis(basis, Z, xyz)
    S, T, V = (O.unpack(x, ob.n) for x in ob.one_electron(0))
    s, U = np.linalg.eigh(S)
    X = (U / np.sqrt(s)) @ U.T
    e0 = np.linalg.eigvalsh(X @ (T + V) @ X)[0]
    assert r.energy == pytest.approx(e0, abs=1e-11)
    assert np.max(np.abs(r.Pb)) == 0.0
    assert abs(r.components["coulomb"] + r.components["exchange"]) < 1e-12
    if basis == "STO-3G":
        assert r.energy == pytest.approx(-0.46658185, abs=1e-7)


def test_uhf_full_alpha_shell_closed_form():
    """Triplet H2/STO-3G (the reference's test/hydrogen.xyz geometry): both
    electrons are alpha and fill both orbitals, so Pa = S^-1 whatever the
    SCF does, and E = tr(Pa H) + 1/2 sum Pa_mn Pa_ls [(mn|ls) - (ml|ns)] +
    E_nuc from the full four-index tensor."""
    Z = np.array([1, 1])
    xyz = np.array([[0.37, 0, 0], [-0.37, 0, 0]]) * O.A_TO_A0
    r = O.scf_energy(os.path.join(BASIS, "STO-3G"), Z, xyz, mult=3, conv=1e-12)
    ob = oracle_basis("STO-3G", Z, xyz)
    S, T, V = (O.unpack(x, ob.n) for x in ob.one_electron(0))
    full = _full_tensor(ob)
    Pa = np.linalg.inv(S)
    e = (np.sum(Pa * (T + V)) + 0.5 * np.einsum("mn,ls,mnls->", Pa, Pa, full)
         - 0.5 * np.einsum("mn,ls,mlns->", Pa, Pa, full) + O.nuclear_repulsion(Z, xyz))
    assert np.allclose(r.Pa, Pa, atol=1e-12)
    assert r.energy == pytest.approx(e, abs=1e-11)


def test_uhf_closed_shell_equals_rhf():
    """With n_alpha = n_beta the unrestricted iteration (two DIIS histories,
    per-spin Fock matrices, the UHF energy expression) must retrace RHF:
    water/6-31G, same iterations and energy to 1e-10 Eh."""
    Z, xyz = G.water()
    ob = oracle_basis("6-31G", Z, xyz)
    S, T, V = (O.unpack(x, ob.n) for x in ob.one_electron(0))
    t = ob.tensor(0)
    en = O.nuclear_repulsion(Z, xyz)
    rr = O.run_scf(S, T, V, O.tensor_jk(t, ob.n), 5, en, conv=1e-10)
    ru = O.run_uhf(S, T, V, O.tensor_jk_u(t, ob.n), 5, 5, en, conv=1e-10)
    assert ru.energy == pytest.approx(rr.energy, abs=1e-10)
    assert ru.iterations == rr.iterations
    assert np.allclose(ru.Pa, ru.Pb, atol=1e-10)


def test_uhf_literature_and_invariants():
    """N atom quartet, STO-3G: 5 alpha electrons fill all five functions
    and the UHF energy is the textbook -53.71901 Eh.  NH2 radical (doublet,
    N-H 1.024 A, 103.4 deg), 6-31G: tr(Pa S) = 5, tr(Pb S) = 4, each density
    idempotent in the S metric, and <S^2> = S(S+1) + n_beta - tr(Pa S Pb S)
    (the reference's spin-contamination estimate, SCFSolver.cpp:96-116) a
    little above 3/4.  (The reference's iteration, DIIS from the second
    step of a core guess, does not settle for radicals whose unpaired
    electron sits in a degenerate pair, such as OH; NH2 is bent.)"""
    e = O.scf_energy(os.path.join(BASIS, "STO-3G"), np.array([7]), np.zeros((1, 3)), mult=4, conv=1e-10).energy
    assert e == pytest.approx(-53.71901, abs=1e-5)
    th = math.radians(103.4)
    Z = np.array([7, 1, 1])
    xyz = np.array([[0, 0, 0], [1.024, 0, 0], [1.024 * math.cos(th), 1.024 * math.sin(th), 0]]) * O.A_TO_A0
    r = O.scf_energy(os.path.join(BASIS, "6-31G"), Z, xyz, mult=2, conv=1e-10)
    S = O.unpack(oracle_basis("6-31G", Z, xyz).one_electron(0)[0], len(r.Pa))
    assert np.trace(r.Pa @ S) == pytest.approx(5.0, abs=1e-9)
    assert np.trace(r.Pb @ S) == pytest.approx(4.0, abs=1e-9)
    for P in (r.Pa, r.Pb):
        assert np.allclose(P @ S @ P, P, atol=1e-7)
    s2 = 0.75 + 4 - np.trace(r.Pa @ S @ r.Pb @ S)
    assert 0.75 < s2 < 0.77


# ------------------------------------------------- Cartesian d (6D) ----
def test_cartesian_d_normalization_and_overlap_closed_forms():
    """One-primitive Cartesian d shell: every component has unit self-overlap
    (the normalization closed form), <xx|yy> = 1/3 and <xx|xy> = 0 on one
    centre, and s/p AOs equal the spherical build's."""
    basis = {8: [(0, [(1.3, 1.0)]), (1, [(0.9, 1.0)]), (2, [(0.8, 1.0)])]}
    Z, xyz = np.array([8], np.int32), np.zeros((1, 3))
    aos = O.build_aos(basis, Z, xyz, cartesian=True)
    assert [(a.l, a.m) for a in aos] == [(0, 0), (1, -1), (1, 0), (1, 1)] + [(2, k) for k in range(6)]
    assert [tuple(a.terms[0][1:]) for a in aos[4:]] == [(2, 0, 0), (1, 1, 0), (1, 0, 1), (0, 2, 0), (0, 1, 1), (0, 0, 2)]
    S = O.unpack(O.OracleBasis(aos, Z, xyz).one_electron(0)[0], len(aos))
    assert np.max(np.abs(np.diag(S) - 1.0)) < 1e-14
    xx, xy, yy = 4, 5, 7
    assert abs(S[xx, yy] - 1.0 / 3.0) < 1e-14
    assert abs(S[xx, xy]) < 1e-15
    sph = O.build_aos(basis, Z, xyz)
    for a, b in zip(aos[:4], sph[:4]):
        assert (a.l, a.m, a.terms) == (b.l, b.m, b.terms)


def test_cartesian_d_spans_spherical_eris():
    """Every spherical d AO is a fixed combination of the Cartesian ones,
    Y_2m = sum_t w_t x^a y^b z^c = sum_t (w_t / N_abc) cart_abc, so each
    spherical ERI (water 6-31G*, d AOs in all four positions) equals the
    same combination of Cartesian ERIs: a wrong normalization, component
    order or transposed index fails it."""
    Z, xyz = G.water_box(2, seed=5)
    bas = O.read_basis(os.path.join(BASIS, "6-31GP"))
    sph = O.build_aos(bas, Z, xyz)
    car = O.build_aos(bas, Z, xyz, cartesian=True)
    obs, obc = O.OracleBasis(sph, Z, xyz), O.OracleBasis(car, Z, xyz)
    # each spherical AO as {cartesian AO index: coefficient}
    comb, ic = [], 0
    i = 0
    while i < len(sph):
        a = sph[i]
        if a.l < 2:
            comb.append({ic: 1.0})
            i += 1
            ic += 1
            continue
        comps = O.cartesian_components(2)
        for k in range(5):
            t = sph[i + k].terms
            comb.append({ic + comps.index(tuple(x[1:])): x[0] / O.cartesian_norm(2, *x[1:]) for x in t})
        i += 5
        ic += 6
    assert ic == len(car)
    rng = np.random.default_rng(2)
    d_sph = [k for k, a in enumerate(sph) if a.l == 2]
    q = rng.integers(0, len(sph), size=(60, 4))
    q[:20, 0] = rng.choice(d_sph, 20)
    q[20:40, 1:3] = rng.choice(d_sph, (20, 2))
    q[40:, :] = rng.choice(d_sph, (20, 4))
    for i, j, k, l in q:
        want = obs.eri(int(i), int(j), int(k), int(l), 1)
        got = 0.0
        for a, ca in comb[i].items():
            for b, cb in comb[j].items():
                for c, cc in comb[k].items():
                    for d, cd in comb[l].items():
                        got += ca * cb * cc * cd * obc.eri(a, b, c, d, 1)
        assert abs(got - want) < 1e-12 * max(1.0, abs(want))
