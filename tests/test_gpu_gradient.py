"""RHF nuclear gradients on the device (dsscf_eri_gradient, dsscf_job_gradient)
against the reference's derivative integrals (fixtures from oracle/_ref) and
against finite differences of the SCF energy, for s, p and d shells."""
import json
import os

import numpy as np
import pytest

import dollarstorescf_amd as ds
from dollarstorescf_amd import geometry as G

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))


def _fixture(case):
    with open(os.path.join(HERE, "golden", "gradient.json")) as f:
        return json.load(f)[case]


def _ao_atom(job, xyz):
    return np.array([int(np.argmin(np.linalg.norm(xyz - job.ao(i)["center"], axis=1))) for i in range(job.n_ao)],
                    np.int32)


@pytest.mark.parametrize("case", ["water_sto3g", "water2_631g", "water_def2svp", "water_631gs", "water2_def2svp"])
@pytest.mark.parametrize("boys", [ds.BOYS_REFERENCE, ds.BOYS_ACCURATE])
def test_eri_gradient_matches_reference(case, boys):
    """Two-electron gradient part (ExactCoulomb.cpp:125-195 over
    electronRepulsionGradient) for a fixed synthetic density, no screening:
    1e-11 relative to the largest component in the reference Boys mode, 1e-9
    in the accurate mode (the two Boys forms differ by up to 3e-11).  The
    d-shell systems run every class with a d shell through the
    wave-cooperative kernel, two waters/def2-SVP up to (dd|dd)."""
    g = _fixture(case)
    Z = np.array(g["Z"], np.int32)
    xyz = np.array(g["xyz_bohr"]).reshape(-1, 3)
    job = ds.Job(0).set_options(boys_mode=boys, schwarz_threshold=0.0).set_basis(g["basis"]).set_molecule(Z, xyz)
    job.construct()
    got = job.context.eri_gradient(np.array(g["Pa"]), _ao_atom(job, xyz), len(Z))
    want = np.array(g["two_electron"])
    tol = (1e-11 if boys == ds.BOYS_REFERENCE else 1e-9) * np.max(np.abs(want))
    print(f"{case} boys={boys}: max |d| = {np.max(np.abs(got - want)):.2e} of max {np.max(np.abs(want)):.2e}", flush=True)
    assert np.max(np.abs(got - want)) < tol, np.max(np.abs(got - want))


@pytest.mark.parametrize("case", ["water_sto3g", "water2_631g", "water_def2svp", "water_631gs", "water2_def2svp"])
@pytest.mark.parametrize("boys", [ds.BOYS_REFERENCE, ds.BOYS_ACCURATE])
def test_one_electron_gradient_device_matches_reference(case, boys):
    """Kinetic, overlap and nuclear-attraction gradient terms on the device
    (dsscf_one_electron_gradient, Gradient.cpp:61-116) for the fixture's
    synthetic densities: 1e-11 relative to the reference build's terms in the
    reference Boys mode (the accurate mode moves the attraction terms by the
    Boys difference, below 1e-9 relative), and equal to the host
    restatement (dsscf_job_gradient_terms) in the same mode."""
    g = _fixture(case)
    Z = np.array(g["Z"], np.int32)
    xyz = np.array(g["xyz_bohr"]).reshape(-1, 3)
    job = ds.Job(0).set_options(boys_mode=boys, schwarz_threshold=0.0).set_basis(g["basis"]).set_molecule(Z, xyz)
    job.construct()
    Pa, W = np.array(g["Pa"]), np.array(g["W"])
    got = job.context.one_electron_gradient(Z, xyz, _ao_atom(job, xyz), 2 * Pa, W)
    want = np.array(g["one_electron"])
    host, _ = job.gradient_terms(2 * Pa, W)
    scale = np.max(np.abs(want))
    print(f"{case} boys={boys}: device vs reference {np.max(np.abs(got - want)):.2e}, vs host "
          f"{np.max(np.abs(got - host)):.2e} of max {scale:.2e}", flush=True)
    assert np.max(np.abs(got - host)) < 1e-11 * scale
    assert np.max(np.abs(got - want)) < (1e-11 if boys == ds.BOYS_REFERENCE else 1e-9) * scale


def test_eri_gradient_d_classes_with_sub_pairs_and_lane_groups():
    """The same d-shell fixture with every shell pair split into sub-pairs of
    at most 5 primitive pairs and lane groups forced on (the work units the
    gradient shares with the Fock build): the sub-quartets add up to the
    same derivatives (1e-11 relative, reference Boys)."""
    g = _fixture("water2_def2svp")
    Z = np.array(g["Z"], np.int32)
    xyz = np.array(g["xyz_bohr"]).reshape(-1, 3)
    job = ds.Job(0).set_options(boys_mode=ds.BOYS_REFERENCE, schwarz_threshold=0.0)
    job.set_basis(g["basis"]).set_molecule(Z, xyz).construct()
    ctx = job.context
    ctx.set_options(boys_mode=ds.BOYS_REFERENCE, schwarz_threshold=0.0, work_item_flops=1.0, group_prims=1.0)
    ctx.prepare()
    assert ctx.stats().n_work_pairs > ctx.stats().n_pairs
    got = ctx.eri_gradient(np.array(g["Pa"]), _ao_atom(job, xyz), len(Z))
    want = np.array(g["two_electron"])
    assert np.max(np.abs(got - want)) < 1e-11 * np.max(np.abs(want))


def _energy(basis, Z, xyz):
    job = ds.Job(0).set_options(diis_size=8, energy_threshold=1e-11, max_iterations=200, schwarz_threshold=0.0)
    job.set_basis(basis).set_molecule(Z, xyz).construct()
    e, _ = job.solve()
    return e, job


@pytest.mark.parametrize("basis,system", [("STO-3G", "water"), ("6-31G", "water2"), ("def2-SVP", "water"),
                                          ("6-31GP", "water2")])
def test_gradient_matches_finite_differences(basis, system):
    """dsscf_job_gradient at the converged density vs central differences of
    the total SCF energy (h = 1e-3 bohr, SCF converged to 1e-11): 2e-6 Eh/bohr;
    the gradient sums to zero over atoms (translational invariance)."""
    Z, xyz = G.water() if system == "water" else G.water_box(2, seed=3)
    xyz = np.array(xyz, float).reshape(-1, 3)
    e0, job = _energy(basis, Z, xyz)
    grad, parts = job.gradient(parts=True)
    assert np.max(np.abs(grad.sum(axis=0))) < 1e-9
    h = 1e-3
    rng = np.random.default_rng(0)
    coords = [(a, d) for a in range(len(Z)) for d in range(3)]
    if len(coords) > 9:  # a sample of the coordinates keeps the test short
        coords = [coords[i] for i in rng.choice(len(coords), 9, replace=False)]
    for a, d in coords:
        xp, xm = xyz.copy(), xyz.copy()
        xp[a, d] += h
        xm[a, d] -= h
        fd = (_energy(basis, Z, xp)[0] - _energy(basis, Z, xm)[0]) / (2 * h)
        assert abs(fd - grad[a, d]) < 2e-6, (a, d, fd, grad[a, d])


def test_gradient_rejects_cartesian_d():
    Z, xyz = G.water()
    job = ds.Job(0).set_options(cartesian=True).set_basis("6-31GP").set_molecule(Z, xyz).construct()
    job.solve()
    with pytest.raises(ds.DsscfError) as ei:
        job.gradient()
    assert ei.value.code == 4
