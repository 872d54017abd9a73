"""The SCF iteration's host part on the device (DIIS, X F X, the density by
purification, energy terms; rocSOLVER dsyevd for the final orbitals) against
the host step and the oracle, through the C-ABI (dsscf_job_set_step)."""
import numpy as np
import pytest

import dollarstorescf_amd as ds
from dollarstorescf_amd import geometry as G
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _solve(basis, Z, xyz, device_step, diis=8, conv=1e-10, boys=ds.BOYS_ACCURATE, schwarz=1e-12):
    job = ds.Job(0).set_options(diis_size=diis, energy_threshold=conv, max_iterations=80, boys_mode=boys,
                                schwarz_threshold=schwarz, device_step=device_step)
    job.set_basis(basis).set_molecule(Z, xyz).construct()
    e, it = job.solve()
    return job, e, it


def test_device_step_water_sto3g_vs_oracle():
    """Config 0 (water/STO-3G, DIIS 6, converged to 1e-10) with the device step:
    the total energy within 1e-8 Eh of the oracle's RHF in the reference Boys mode."""
    Z, xyz = G.water()
    _, e, _ = _solve("STO-3G", Z, xyz, True, diis=6, boys=ds.BOYS_REFERENCE, schwarz=0.0)
    ref = O.scf_energy(ds.basis_path("STO-3G"), Z, xyz, mode=ds.BOYS_REFERENCE, diis_size=8, conv=1e-10)
    assert abs(e - ref.energy) < 1e-8


@pytest.mark.parametrize("case", ["water_sto3g", "water8_631g", "butane_def2svp", "water2_631gp"])
def test_device_step_matches_host_step(case):
    """Same job with the iteration's host part on the CPU and on the device:
    the same number of iterations (within one), total energies within 1e-9 Eh,
    densities within 1e-5 (both stop at |dE| < 1e-10), and the same
    per-iteration energies while they are still far from convergence (the two
    steps differ only in rounding)."""
    Z, xyz, basis = {
        "water_sto3g": (*G.water(), "STO-3G"),
        "water8_631g": (*G.water_lattice(1), "6-31G"),
        "butane_def2svp": (*G.alkane(4), "def2-SVP"),
        "water2_631gp": (*G.water_box(2, seed=3), "6-31GP"),
    }[case]
    jh, eh, ith = _solve(basis, Z, xyz, False)
    jd, ed, itd = _solve(basis, Z, xyz, True)
    steps = jd.get("device_steps")  # device iterations, purification steps, eigensolver fallbacks
    print(f"{case}: {ith} host / {itd} device iterations, {steps[1] / steps[0]:.1f} purification steps per "
          f"iteration, {steps[2]:.0f} fallbacks, dE {ed - eh:.2e}", flush=True)
    assert steps[0] == itd and steps[2] == 0
    assert abs(ith - itd) <= 1, (ith, itd)
    assert abs(eh - ed) < 1e-9, (eh, ed)
    assert np.max(np.abs(jh.get("Pa") - jd.get("Pa"))) < 1e-5
    hh, hd = jh.history(), jd.history()
    assert np.max(np.abs(hh[:3] - hd[:3])) < 1e-8
    # the energy terms come from the same final J/K path
    for k, v in jh.energies().items():
        assert abs(v - jd.energies()[k]) < 1e-5, k


@pytest.mark.parametrize("case", ["water8_631g", "butane_def2svp"])
def test_device_step_purification_matches_eigensolver_route(case):
    """The device step's two routes to the density: purification of
    F' = X F X (device_step=1) and rocSOLVER eigenvectors (device_step=2, the
    route an unfinished purification falls back to) give the same
    iterations, per-iteration energies within 1e-9 Eh and final densities
    within 1e-8; the counts show which route ran."""
    Z, xyz, basis = {"water8_631g": (*G.water_lattice(1), "6-31G"), "butane_def2svp": (*G.alkane(4), "def2-SVP")}[case]
    jp, ep, itp = _solve(basis, Z, xyz, 1)
    je, ee, ite = _solve(basis, Z, xyz, 2)
    cp, ce = jp.get("device_steps"), je.get("device_steps")
    assert cp[0] == itp and cp[2] == 0 and cp[1] >= 10 * itp
    assert ce[0] == ite and ce[1] == 0 and ce[2] == ite
    assert itp == ite
    assert np.max(np.abs(jp.history() - je.history())) < 1e-9
    assert abs(ep - ee) < 1e-10
    assert np.max(np.abs(jp.get("Pa") - je.get("Pa"))) < 1e-8


def test_device_step_orbitals():
    """Orbitals of the device step solve the Roothaan-Hall equations
    F C = S C eps with C^T S C = 1 (F = H + G at the converged density), and
    Pa = C_occ C_occ^T."""
    Z, xyz = G.water_box(2, seed=3)
    job, _, _ = _solve("6-31G", Z, xyz, True)
    n = job.n_ao
    S = O.unpack(job.get("S"), n)
    F = O.unpack(job.get("H") + job.get("J") - job.get("K"), n)
    C = job.get("C")
    eps = job.get("eps")
    assert np.all(np.diff(eps) >= -1e-12)
    assert np.max(np.abs(C.T @ S @ C - np.eye(n))) < 1e-9
    # F is built from the final density while C diagonalized the previous
    # (DIIS-extrapolated) Fock matrix: equal to the convergence level only
    assert np.max(np.abs(F @ C - S @ C * eps)) < 1e-4
    nocc = 10
    assert np.max(np.abs(O.unpack(job.get("Pa"), n) - C[:, :nocc] @ C[:, :nocc].T)) < 1e-12


def test_device_step_rejects_long_history():
    job = ds.Job(0)
    with pytest.raises(ds.DsscfError):
        job.set_options(diis_size=20, device_step=True)


@pytest.mark.parametrize("device_step", [False, True])
def test_molden_of_solved_job(device_step, tmp_path):
    """A solved job's Molden file (dsscf_job_write_molden) carries its orbital
    energies and coefficients: parsed back, in the reference's AO order
    (Molden p order x, y, z from m = -1, 0, 1 via WriteOrbitals.cpp:25-31),
    they match the job's C and eps to the printed 8 significant digits."""
    Z, xyz = G.water_box(2, seed=3)
    job, _, _ = _solve("def2-SVP", Z, xyz, device_step)
    n = job.n_ao
    path = str(tmp_path / "w2.molden")
    job.write_molden(path)
    text = open(path).read().split("[MO]\n")[1].splitlines()
    C = job.get("C")
    eps = job.get("eps")
    # AO position in the file -> the reference's AO index (perm per shell)
    order, k = [], 0
    ao_l = [job.ao(i)["l"] for i in range(n)]
    perm = {0: [0], 1: [0, 2, 1], 2: [2, 1, 3, 0, 4]}
    while k < n:
        l = ao_l[k]
        order += [k + p for p in perm[l]]
        k += 2 * l + 1
    mo, row = -1, 0
    for line in text:
        if line.startswith("Sym="):
            mo += 1
            row = 0
        elif line.startswith("Ene="):
            assert abs(float(line.split()[1]) - eps[mo]) <= 1e-5 * max(1.0, abs(eps[mo]))
        elif line.startswith("Occup="):
            assert int(line.split()[1]) == (2 if mo < 10 else 0)  # 20 electrons
        elif line.strip() and not line.startswith("Spin"):
            idx, val = line.split()
            assert int(idx) == row + 1
            want = C[order[row], mo]
            assert abs(float(val) - want) <= 1e-7 * max(abs(want), 1e-3), (mo, row)
            row += 1
    assert mo == n - 1
