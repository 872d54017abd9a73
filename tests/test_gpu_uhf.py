"""GPU parity of unrestricted (open-shell) jobs: K[Pa] and K[Pb] from one
pass over the integrals (dsscf_jk_uhf) against the reference's stored-tensor
loops, and UHF energies against the oracle's restatement of the reference's
open-shell iteration (SCFSolver.cpp:145-237)."""
import math

import numpy as np
import pytest

import dollarstorescf_amd as ds
from dollarstorescf_amd import geometry as G
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _nh2():
    th = math.radians(103.4)
    xyz = np.array([[0, 0, 0], [1.024, 0, 0], [1.024 * math.cos(th), 1.024 * math.sin(th), 0]]) * O.A_TO_A0
    return np.array([7, 1, 1], np.int32), xyz


def _ch2():
    th = math.radians(133.9)
    xyz = np.array([[0, 0, 0], [1.078, 0, 0], [1.078 * math.cos(th), 1.078 * math.sin(th), 0]]) * O.A_TO_A0
    return np.array([6, 1, 1], np.int32), xyz


@pytest.mark.parametrize("basis,mol", [("6-31GP", "nh2"), ("def2-SVP", "ethane")])
def test_uhf_jk_vs_tensor(basis, mol):
    """J[Pa + Pb], K[Pa], K[Pb] of random densities against the reference's
    loops (ExactCoulomb.cpp:103-123, oracle) over the unscreened tensor,
    reference Boys mode: 1e-11 of max |J|.  Ethane/def2-SVP runs the
    wave-cooperative (dd|dp), (dd|dd) classes too.  J and K[Pa] also equal
    the restricted build's at (Pa + Pb, Pa), which shares the kernels."""
    Z, xyz = _nh2() if mol == "nh2" else G.alkane(2)
    job = ds.Job(0).set_options(boys_mode=ds.BOYS_REFERENCE, schwarz_threshold=0.0)
    job.set_basis(basis).set_molecule(Z, xyz, multiplicity=2 if mol == "nh2" else 1).construct()
    n = job.n_ao
    rng = np.random.default_rng(8)
    Ca, Cb = rng.standard_normal((n, 6)) * 0.3, rng.standard_normal((n, 5)) * 0.3
    Pa, Pb = O.pack(Ca @ Ca.T), O.pack(Cb @ Cb.T)
    J, Ka, Kb = job.context.jk_uhf(Pa + Pb, Pa, Pb)
    ob = O.OracleBasis(O.build_aos(O.read_basis(ds.basis_path(basis)), Z, xyz), Z, xyz)
    t = ob.tensor(0)
    Jr, Kar, Kbr = O.coulomb(t, n, Pa + Pb), O.exchange(t, n, Pa), O.exchange(t, n, Pb)
    scale = max(1.0, np.max(np.abs(Jr)))
    print(f"{mol}/{basis}: n_ao={n} |dJ|={np.max(np.abs(J - Jr)):.2e} |dKa|={np.max(np.abs(Ka - Kar)):.2e} "
          f"|dKb|={np.max(np.abs(Kb - Kbr)):.2e}", flush=True)
    assert np.max(np.abs(J - Jr)) < 1e-11 * scale
    assert np.max(np.abs(Ka - Kar)) < 1e-11 * scale
    assert np.max(np.abs(Kb - Kbr)) < 1e-11 * scale
    J2, K2 = job.context.jk(Pa + Pb, Pa)
    assert np.max(np.abs(J2 - J)) < 1e-12 * scale
    assert np.max(np.abs(K2 - Ka)) < 1e-12 * scale


@pytest.mark.parametrize("case", ["N/STO-3G", "NH2/6-31G", "CH2/6-31G", "NH2/def2-SVP"])
@pytest.mark.parametrize("boys", [ds.BOYS_REFERENCE, ds.BOYS_ACCURATE])
def test_uhf_energy_vs_oracle(case, boys):
    """Converged UHF energies (|dE_elec| < 1e-10, DIIS 8, core guess)
    within 1e-8 Eh of the oracle in the same Boys mode, with every energy
    term within 1e-7 Eh: the N atom quartet (textbook -53.71901 Eh), the NH2
    radical doublet and the CH2 triplet."""
    mol, basis = case.split("/")
    if mol == "N":
        Z, xyz, mult = np.array([7], np.int32), np.zeros((1, 3)), 4
    elif mol == "NH2":
        (Z, xyz), mult = _nh2(), 2
    else:
        (Z, xyz), mult = _ch2(), 3
    job = ds.Job(0).set_options(diis_size=8, energy_threshold=1e-10, boys_mode=boys, schwarz_threshold=0.0,
                                max_iterations=300)
    job.set_basis(basis).set_molecule(Z, xyz, multiplicity=mult).construct()
    e, it = job.solve()
    ref = O.scf_energy(ds.basis_path(basis), Z, xyz, mult=mult, mode=boys, diis_size=8, conv=1e-10)
    print(f"{case} boys={boys}: E={e:.10f} (oracle {ref.energy:.10f}), iterations {it} (oracle {ref.iterations})",
          flush=True)
    assert abs(e - ref.energy) < 1e-8
    got = job.energies()
    for k in ("kinetic", "nuclear_attraction", "coulomb", "exchange", "nuclear"):
        assert abs(got[k] - ref.components[k]) < 1e-7, k
    n = job.n_ao
    Pa, Pb = job.get("Pa"), job.get("Pb")
    S = O.unpack(job.get("S"), n)
    ne = int(np.sum(Z))
    assert np.trace(O.unpack(Pa, n) @ S) == pytest.approx((ne + mult) // 2, abs=1e-8)
    assert np.trace(O.unpack(Pb, n) @ S) == pytest.approx((ne + 1 - mult) // 2, abs=1e-8)
    if mol == "N" and basis == "STO-3G":
        assert e == pytest.approx(-53.71901, abs=1e-5)


def test_uhf_jobs_are_solved_alone():
    """Open-shell jobs are refused by the batched solver, the device
    iteration step and the gradient (DSSCF_ERR_UNSUPPORTED, 4), and still
    solve alone afterwards."""
    Z, xyz = _nh2()
    job = ds.Job(0).set_options(energy_threshold=1e-8).set_basis("6-31G").set_molecule(Z, xyz, multiplicity=2)
    job.construct()
    with pytest.raises(ds.DsscfError) as ei:
        ds.solve_jobs([job])
    assert ei.value.code == 4
    dev = ds.Job(0).set_options(device_step=True).set_basis("6-31G").set_molecule(Z, xyz, multiplicity=2).construct()
    with pytest.raises(ds.DsscfError) as ei:
        dev.solve()
    assert ei.value.code == 4
    e, _ = job.solve()
    with pytest.raises(ds.DsscfError) as ei:
        job.gradient()
    assert ei.value.code == 4
    assert e < -55.0
