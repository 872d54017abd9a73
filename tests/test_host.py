"""CPU tests of the product's host side: the C-ABI library loads and exports
every declared symbol, the host chemistry layer (basis parsing, AO
construction, one-electron terms) agrees with the oracle, error paths report
codes instead of aborting, the FLOP model reproduces its reference values."""
import os
import re
import subprocess
import sys

import numpy as np
import pytest

import dollarstorescf_amd as ds
from dollarstorescf_amd import flops
from dollarstorescf_amd import geometry as G
from oracle import oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    names = set()
    for h in ("dsscf.h", "dsscf_scf.h"):
        text = open(os.path.join(ROOT, "include", h)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        names |= set(re.findall(r"\b(dsscf_\w+)\s*\(", text))
    return names


def test_library_exports_every_declared_symbol():
    lib = ds.lib()
    missing = [n for n in sorted(declared_symbols()) if not hasattr(lib, n)]
    assert not missing, missing
    assert len(declared_symbols()) >= 30


def test_library_facts():
    assert ds.lib().dsscf_max_l() >= 1
    cls = ds.class_list()
    assert cls[0] == (0, 0, 0, 0)
    assert len(cls) == ds.lib().dsscf_class_count()


@pytest.mark.parametrize("basis", ["STO-3G", "6-31G", "6-31GP", "6-31GPP", "def2-SVP"])
def test_host_aos_and_one_electron_match_oracle(basis):
    Z, xyz = G.water_box(2, seed=5)
    job = ds.Job().set_options(boys_mode=ds.BOYS_REFERENCE).set_basis(basis).set_molecule(Z, xyz).construct_host()
    aos = O.build_aos(O.read_basis(ds.basis_path(basis)), Z, xyz)
    assert job.n_ao == len(aos)
    for i, a in enumerate(aos):
        h = job.ao(i)
        assert (h["l"], h["m"]) == (a.l, a.m)
        assert [t[1:] for t in h["terms"]] == [tuple(t[1:]) for t in a.terms]
        assert np.allclose([t[0] for t in h["terms"]], [t[0] for t in a.terms], rtol=1e-15, atol=0)
        assert np.allclose(h["weight"], [p[1] for p in a.prims], rtol=1e-15, atol=0)
    ob = O.OracleBasis(aos, Z, xyz)
    S, T, V = ob.one_electron(0)
    assert np.max(np.abs(job.get("S") - S)) < 1e-13
    assert np.max(np.abs(job.get("T") - T)) < 1e-12
    assert np.max(np.abs(job.get("V") - V)) < 1e-11


@pytest.mark.parametrize("basis", ["6-31GP", "def2-SVP"])
def test_host_cartesian_aos_and_one_electron_match_oracle(basis):
    """Cartesian-d jobs: the host AO list (6 x-major components per d shell,
    each with its normalizing weight) and S, T, V equal the oracle's
    Cartesian construction (pinned in test_oracle.py)."""
    Z, xyz = G.water_box(2, seed=5)
    job = ds.Job().set_options(boys_mode=ds.BOYS_REFERENCE, cartesian=True).set_basis(basis).set_molecule(Z, xyz)
    job.construct_host()
    aos = O.build_aos(O.read_basis(ds.basis_path(basis)), Z, xyz, cartesian=True)
    assert job.n_ao == len(aos) and sum(a.l == 2 for a in aos) == 12
    for i, a in enumerate(aos):
        h = job.ao(i)
        assert (h["l"], h["m"]) == (a.l, a.m)
        assert [t[1:] for t in h["terms"]] == [tuple(t[1:]) for t in a.terms]
        assert np.allclose([t[0] for t in h["terms"]], [t[0] for t in a.terms], rtol=1e-15, atol=0)
    S, T, V = O.OracleBasis(aos, Z, xyz).one_electron(0)
    assert np.max(np.abs(job.get("S") - S)) < 1e-13
    assert np.max(np.abs(job.get("T") - T)) < 1e-12
    assert np.max(np.abs(job.get("V") - V)) < 1e-11


def test_host_one_electron_accurate_mode_shift_small():
    """V with the accurate Boys differs from the reference series by < 1e-9."""
    Z, xyz = G.water()
    a = ds.Job().set_options(boys_mode=ds.BOYS_ACCURATE).set_basis("6-31G").set_molecule(Z, xyz).construct_host()
    b = ds.Job().set_options(boys_mode=ds.BOYS_REFERENCE).set_basis("6-31G").set_molecule(Z, xyz).construct_host()
    d = np.max(np.abs(a.get("V") - b.get("V")))
    assert 0 < d < 1e-9


def test_basis_parser_errors(tmp_path):
    bad = tmp_path / "bad"
    bad.write_text("H     0\nS    3   1.00\n  0.34D+01  0.15D+00\n  -1.0  0.5\n  0.1 0.4\n****\n")
    with pytest.raises(ds.DsscfError) as e:
        ds.Job().set_basis(str(bad))
    assert "positive" in str(e.value)
    with pytest.raises(ds.DsscfError):
        ds.Job().set_basis(str(tmp_path / "missing"))
    sp = tmp_path / "sp"
    sp.write_text("O     0\nSP   2   1.00\n  1.0  0.5  0.6\n  0.3  0.5  0.4\n****\n")
    job = ds.Job().set_basis(str(sp)).set_molecule([8], np.zeros((1, 3)))
    # an SP line becomes an S shell followed by a P shell: 4 AOs, l = 0,1,1,1
    job.construct_host()
    assert job.n_ao == 4 and [job.ao(i)["l"] for i in range(4)] == [0, 1, 1, 1]


def test_multiplicity_mismatch_is_reported():
    with pytest.raises(ds.DsscfError):
        ds.Job().set_basis("STO-3G").set_molecule([8, 1], np.array([[0, 0, 0], [0, 0, 1.8]]), multiplicity=1)


def test_xyz_reader(tmp_path):
    f = tmp_path / "w.xyz"
    f.write_text("3\nwater\nO 0 0 0\nH 0.7576 0.5870 0\nH -0.7576 0.5870 0\n")
    job = ds.Job().set_basis("STO-3G").load_xyz(str(f)).construct_host()
    assert job.n_ao == 7
    Z, xyz = O.read_xyz(str(f))
    ob = O.OracleBasis(O.build_aos(O.read_basis(ds.basis_path("STO-3G")), Z, xyz), Z, xyz)
    assert np.max(np.abs(job.get("S") - ob.one_electron(0)[0])) < 1e-14


def test_flop_model_reference_values():
    want = {(0, 0, 0, 0): 69, (1, 0, 0, 0): 106, (1, 1, 0, 0): 211, (1, 0, 1, 0): 301, (1, 1, 1, 0): 790,
            (1, 1, 1, 1): 2191, (2, 0, 0, 0): 205, (2, 2, 2, 2): 38673}
    for c, v in want.items():
        assert flops.f_prim(*c) == v
    assert flops.n_vrr(1, 1, 1, 1) == 195 and flops.n_vrr(2, 2, 2, 2) == 3666


def test_geometry_generators():
    Z, xyz = G.water_lattice(1)
    assert len(Z) == 24 and list(Z[:3]) == [8, 1, 1]
    O_ = xyz[Z == 8] / G.A_TO_A0
    d = np.linalg.norm(O_[:, None] - O_[None], axis=-1)
    assert d[d > 0].min() > 2.3
    for k in range(8):
        w = xyz[3 * k:3 * k + 3] / G.A_TO_A0
        assert np.linalg.norm(w[1] - w[0]) == pytest.approx(0.9572)
        assert np.linalg.norm(w[2] - w[0]) == pytest.approx(0.9572)
    Z, xyz = G.water_box(6, seed=2)
    O_ = xyz[Z == 8] / G.A_TO_A0
    d = np.linalg.norm(O_[:, None] - O_[None], axis=-1)
    assert d[d > 0].min() >= 2.6 - 1e-9


def test_bench_refuses_multi_gpu():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True, text=True)
    assert r.returncode == 2 and "one GPU" in r.stderr


@pytest.mark.parametrize("n", [1, 2, 3, 7, 40, 130])
def test_host_eigensolver_vs_numpy(n):
    """The SCF layer's symmetric eigensolver (dspev_ role): eigenvalues equal
    numpy's eigvalsh, eigenvectors orthonormal and A V = V diag(w)."""
    rng = np.random.default_rng(n)
    A = rng.standard_normal((n, n))
    A = A + A.T
    w, V = ds.syev(A)
    assert np.all(np.diff(w) >= 0.0)
    assert np.max(np.abs(w - np.linalg.eigvalsh(A))) < 1e-12 * max(1.0, np.abs(w).max())
    assert np.max(np.abs(V.T @ V - np.eye(n))) < 1e-12
    assert np.max(np.abs(A @ V - V * w)) < 1e-12 * max(1.0, np.abs(w).max())


def test_host_eigensolver_degenerate_and_split():
    """Degenerate spectra and matrices that split into blocks (zero couplings)."""
    A = np.eye(30)
    A[0, 0] = 3.0
    w, V = ds.syev(A)
    assert np.allclose(w, np.sort(np.diag(A)), atol=1e-14)
    assert np.max(np.abs(A @ V - V * w)) < 1e-13
    rng = np.random.default_rng(3)
    B = rng.standard_normal((12, 12))
    B = B + B.T
    B[5, :] = B[:, 5] = 0.0
    B[5, 5] = -2.0
    w, V = ds.syev(B)
    assert np.max(np.abs(w - np.linalg.eigvalsh(B))) < 1e-12
    assert np.max(np.abs(B @ V - V * w)) < 1e-12


@pytest.mark.timeout(60)
def test_server_rounds_return_every_job():
    """The serving loop's host side (worker construction, admission in
    submission order, retirement, poll) hands back every submitted job.  In
    a container without a GPU each job's device preparation fails with
    DSSCF_ERR_HIP; on a GPU host the jobs converge."""
    srv = ds.Server(0, budget_bytes=1e9, max_active=2, workers=2)
    for k, (basis, (Z, xyz)) in enumerate([("STO-3G", G.water()), ("6-31G", G.water_box(2, seed=1)),
                                           ("STO-3G", G.alkane(1))]):
        job = ds.Job(0).set_options(max_iterations=50).set_basis(basis).set_molecule(Z, xyz)
        srv.submit(job, k)
    done = srv.run_until_idle()
    srv.close()
    assert sorted(d["tag"] for d in done) == [0, 1, 2]
    st = {d["status"] for d in done}
    assert st in ({0}, {2})
    for d in done:
        assert d["t_submit"] <= d["t_done"]


def test_server_refuses_duplicate_tag_second_submit_and_early_destroy():
    """A job the server still holds can be neither submitted again (a second
    request would construct the same Job on two workers) nor destroyed (its
    context may be in a fused build), and Python refuses a tag still in
    flight (it keeps the job alive through that tag).  Once polled, the job
    may be destroyed."""
    srv = ds.Server(0, budget_bytes=1e9, max_active=1, workers=1)
    Z, xyz = G.water()
    a = ds.Job(0).set_options(max_iterations=50).set_basis("STO-3G").set_molecule(Z, xyz)
    b = ds.Job(0).set_options(max_iterations=50).set_basis("STO-3G").set_molecule(Z, xyz)
    srv.submit(a, 0)
    with pytest.raises(ds.DsscfError):
        srv.submit(b, 0)  # tag 0 in flight
    rc = ds.lib().dsscf_server_submit(srv._h, a._h, 1)
    assert rc == 3  # DSSCF_ERR_STATE: already held
    assert ds.lib().dsscf_job_destroy(a._h) == 3
    done = srv.run_until_idle()
    assert [d["tag"] for d in done] == [0]
    assert ds.lib().dsscf_job_destroy(a._h) == 0
    a._h = None
    srv.close()
    b.close()


def test_server_refuses_open_shell_job():
    """An unrestricted job (multiplicity 2) is solved alone: the server
    refuses it at submit with DSSCF_ERR_UNSUPPORTED (4) and a message."""
    srv = ds.Server(0, budget_bytes=1e9, max_active=2, workers=1)
    job = ds.Job(0).set_basis("STO-3G").set_molecule(np.array([1], np.int32), np.zeros((1, 3)), multiplicity=2)
    with pytest.raises(ds.DsscfError) as ei:
        srv.submit(job, 3)
    assert ei.value.code == 4
    assert "open-shell" in str(ei.value)
    srv.close()
    job.close()


def test_server_refuses_job_that_can_never_fit():
    """Forced failure: a job whose device-footprint estimate exceeds the whole
    budget is retired at admission with DSSCF_ERR_NOMEM (5) and a message,
    before any device work, instead of waiting for memory forever."""
    srv = ds.Server(0, budget_bytes=1e4, max_active=4, workers=1)
    Z, xyz = G.water()
    job = ds.Job(0).set_basis("STO-3G").set_molecule(Z, xyz)
    srv.submit(job, 7)
    done = srv.run_until_idle()
    srv.close()
    assert [d["tag"] for d in done] == [7]
    assert done[0]["status"] == 5
    assert "budget" in job._err()


def _molden_cases():
    import json
    with open(os.path.join(ROOT, "tests", "golden", "molden.json")) as f:
        return json.load(f)


@pytest.mark.parametrize("case", ["water_sto3g", "water2_def2svp", "ethane_631gp"])
def test_molden_matches_reference_writer(case, tmp_path):
    """dsscf_job_write_molden reproduces, byte for byte, the file that the
    reference's writeOrbitalsMolden logic and its own FormattedStream produce
    for the same orbitals (fixture from oracle/_ref; orbitals synthetic, with
    values at the number formatter's rounding and notation boundaries)."""
    g = _molden_cases()[case]
    Z = np.array(g["Z"], np.int32)
    xyz = np.array(g["xyz_bohr"]).reshape(-1, 3)
    job = ds.Job(0).set_basis(g["basis"]).set_molecule(Z, xyz).construct_host()
    n = job.n_ao
    Ck = np.array(g["C_columns"])
    C = np.zeros((n, n))
    C[:, :Ck.shape[1]] = Ck
    job.set_orbitals(C, np.array(g["eps"]))
    path = str(tmp_path / "out.molden")
    job.write_molden(path, g["write_virtual"])
    with open(path) as f:
        got = f.read()
    assert got == g["molden"]


def test_molden_needs_orbitals(tmp_path):
    Z, xyz = G.water()
    job = ds.Job(0).set_basis("STO-3G").set_molecule(Z, xyz).construct_host()
    with pytest.raises(ds.DsscfError):
        job.write_molden(str(tmp_path / "x.molden"))
    with pytest.raises(ValueError):
        job.set_orbitals(np.zeros((3, 3)), np.zeros(3))


@pytest.mark.parametrize("case", ["water_sto3g", "water2_631g", "water_def2svp", "water_631gs", "water2_def2svp"])
def test_gradient_host_terms_match_reference(case):
    """One-electron (kinetic, overlap, nuclear attraction) and nuclear-repulsion
    gradient terms (Gradient.cpp:61-132) for fixed synthetic densities vs the
    reference's own derivative integrals and loops (fixture from oracle/_ref)."""
    import json
    with open(os.path.join(ROOT, "tests", "golden", "gradient.json")) as f:
        g = json.load(f)[case]
    Z = np.array(g["Z"], np.int32)
    xyz = np.array(g["xyz_bohr"]).reshape(-1, 3)
    job = ds.Job(0).set_options(boys_mode=ds.BOYS_REFERENCE).set_basis(g["basis"]).set_molecule(Z, xyz)
    job.construct_host()
    Pa, W = np.array(g["Pa"]), np.array(g["W"])
    g1, gn = job.gradient_terms(2 * Pa, W)
    r1, rn = np.array(g["one_electron"]), np.array(g["nuclear"])
    assert np.max(np.abs(g1 - r1)) < 1e-12 * max(1.0, np.max(np.abs(r1)))
    assert np.max(np.abs(gn - rn)) < 1e-13 * max(1.0, np.max(np.abs(rn)))
    # translational invariance of the sum over atoms
    assert np.max(np.abs((g1 + gn).sum(axis=0))) < 1e-10


def test_gradient_terms_reject_cartesian_d():
    """The gradient follows the reference's AO set: a Cartesian-d job is refused."""
    Z, xyz = G.water()
    job = ds.Job(0).set_options(cartesian=True).set_basis("6-31GP").set_molecule(Z, xyz).construct_host()
    n = job.n_ao
    with pytest.raises(ds.DsscfError) as ei:
        job.gradient_terms(np.zeros(n * (n + 1) // 2), np.zeros(n * (n + 1) // 2))
    assert ei.value.code == 4


def test_host_layer_under_address_and_ub_sanitizers(tmp_path):
    """The host chemistry, linear algebra, one-electron gradient and Molden
    number code built with -fsanitize=address,undefined (g++, host code only)
    runs tests/host_sanitize/driver.cpp -- basis and XYZ parsing with their
    error paths, spherical and Cartesian AOs, one-electron matrices in both
    Boys modes, gradient terms, the eigensolver, the linear solve, the number
    text -- to its end without a sanitizer report."""
    host = os.path.join(ROOT, "dollarstorescf_amd", "csrc", "host")
    flags = ["-std=c++17", "-O1", "-g", "-fsanitize=address,undefined", "-fno-sanitize-recover=all",
             "-ffunction-sections", "-fdata-sections", "-I" + host, "-I" + os.path.join(ROOT, "include")]
    units = [(os.path.join(ROOT, "tests", "host_sanitize", "driver.cpp"), []),
             (os.path.join(host, "chem.cpp"), []), (os.path.join(host, "linalg.cpp"), []),
             (os.path.join(host, "linalg.cpp"), ["-DLA_AVX2", "-mavx2", "-mfma"]),
             (os.path.join(host, "molden.cpp"), []), (os.path.join(host, "gradient.cpp"), [])]
    objs = []
    for k, (src, extra) in enumerate(units):
        o = str(tmp_path / ("u%d.o" % k))
        subprocess.run(["g++", *flags, *extra, "-c", src, "-o", o], check=True, capture_output=True)
        objs.append(o)
    exe = str(tmp_path / "driver")
    # the gradient unit's device-side callers are never reached: drop them
    subprocess.run(["g++", "-fsanitize=address,undefined", "-Wl,--gc-sections", *objs, "-o", exe], check=True,
                   capture_output=True)
    work = tmp_path / "work"
    work.mkdir()
    r = subprocess.run([exe, os.path.join(ROOT, "dollarstorescf_amd", "basis"), str(work)], capture_output=True,
                       text=True, env={**os.environ, "ASAN_OPTIONS": "detect_leaks=1"})
    assert r.returncode == 0, r.stdout + r.stderr
    assert "host layer OK" in r.stdout
