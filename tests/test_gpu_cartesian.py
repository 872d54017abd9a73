"""GPU parity of the Cartesian-d form (d shells as the six functions xx, xy,
xz, yy, yz, zz, each primitive normalized): ERIs, one-electron matrices,
J/K and RHF energies against the oracle's Cartesian construction (pinned in
tests/test_oracle.py by closed forms and by the spherical ERIs it spans)."""
import numpy as np
import pytest

import dollarstorescf_amd as ds
from dollarstorescf_amd import geometry as G
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _oracle(basis, Z, xyz):
    return O.OracleBasis(O.build_aos(O.read_basis(ds.basis_path(basis)), Z, xyz, cartesian=True), Z, xyz)


def _job(basis, Z, xyz, boys, schwarz=0.0):
    j = ds.Job(0).set_options(boys_mode=boys, schwarz_threshold=schwarz, cartesian=True)
    return j.set_basis(basis).set_molecule(Z, xyz).construct()


@pytest.mark.parametrize("boys", [ds.BOYS_REFERENCE, ds.BOYS_ACCURATE])
@pytest.mark.parametrize("basis", ["6-31GP", "def2-SVP"])
def test_cartesian_eri_all_unique_water(boys, basis):
    """Every unique (ij|kl) of water in Cartesian 6-31G* (19 AOs) and
    def2-SVP (25 AOs, up to (dd|dd), register and wave-cooperative kernels):
    1e-12 absolute against the oracle in the same Boys mode."""
    Z, xyz = G.water()
    job = _job(basis, Z, xyz, boys)
    n = job.n_ao
    assert n == {"6-31GP": 19, "def2-SVP": 25}[basis]
    q = np.array([[i, j, k, l] for i in range(n) for j in range(i + 1) for k in range(n) for l in range(k + 1)
                  if i * (i + 1) // 2 + j >= k * (k + 1) // 2 + l], np.int32)
    got = job.context.eri(q)
    want = _oracle(basis, Z, xyz).eri_list(q, boys)
    assert np.max(np.abs(got - want)) < 1e-12


@pytest.mark.parametrize("basis", ["6-31GP", "def2-SVP"])
def test_cartesian_one_electron_device_vs_oracle(basis):
    """S, T, V from the device kernel (Job.construct) against the oracle's
    Cartesian one-electron integrals, reference Boys mode: 1e-12."""
    Z, xyz = G.water_box(2, seed=5)
    job = _job(basis, Z, xyz, ds.BOYS_REFERENCE)
    S, T, V = _oracle(basis, Z, xyz).one_electron(0)
    for name, want in (("S", S), ("T", T), ("V", V)):
        assert np.max(np.abs(job.get(name) - want)) < 1e-12 * max(1.0, np.abs(want).max()), name


@pytest.mark.parametrize("basis", ["6-31GP", "def2-SVP"])
def test_cartesian_jk_vs_tensor(basis):
    """J[Ptot], K[Pa] of two waters (Cartesian d) against the reference's
    stored-tensor loops over the oracle's Cartesian tensor: 1e-11 of max |J|."""
    Z, xyz = G.water_box(2, seed=3)
    job = _job(basis, Z, xyz, ds.BOYS_REFERENCE)
    n = job.n_ao
    C = np.random.default_rng(4).standard_normal((n, 6))
    Pa = O.pack(C @ C.T * 0.1)
    job.set_density(Pa)
    J, K = job.fock_build()
    t = _oracle(basis, Z, xyz).tensor(0)
    Jr, Kr = O.coulomb(t, n, 2 * Pa), O.exchange(t, n, Pa)
    scale = max(1.0, np.max(np.abs(Jr)))
    assert np.max(np.abs(J - Jr)) < 1e-11 * scale
    assert np.max(np.abs(K - Kr)) < 1e-11 * scale


@pytest.mark.parametrize("boys", [ds.BOYS_REFERENCE, ds.BOYS_ACCURATE])
def test_cartesian_scf_water_631gs_vs_oracle(boys):
    """RHF/6-31G* (6D) of water converged to 1e-10: within 1e-8 Eh of the
    oracle's RHF over its stored Cartesian tensor, and below the spherical
    (5D) energy, whose basis the Cartesian one contains."""
    Z, xyz = G.water()
    job = ds.Job(0).set_options(energy_threshold=1e-10, boys_mode=boys, schwarz_threshold=0.0, cartesian=True)
    job.set_basis("6-31GP").set_molecule(Z, xyz).construct()
    e, _ = job.solve()
    ref = O.scf_energy(ds.basis_path("6-31GP"), Z, xyz, mode=boys, conv=1e-10, cartesian=True)
    assert abs(e - ref.energy) < 1e-8
    sph = ds.Job(0).set_options(energy_threshold=1e-10, boys_mode=boys, schwarz_threshold=0.0)
    e5, _ = sph.set_basis("6-31GP").set_molecule(Z, xyz).construct().solve()
    assert e < e5 - 1e-6


def test_cartesian_jobs_fused_equal_alone_and_molden_refused(tmp_path):
    """Cartesian-d jobs of several sizes solved together with fused builds
    converge to their own energies (1e-9 Eh); a batch refuses to mix
    spherical and Cartesian contexts; the Molden writer (the reference's 5D
    layout) refuses a Cartesian job."""
    specs = [G.water(), G.water_box(2, seed=3), G.water_box(4, seed=8), G.alkane(2)]

    def make(Z, xyz, cart=True):
        j = ds.Job(0).set_options(diis_size=8, energy_threshold=1e-9, max_iterations=80, schwarz_threshold=1e-12,
                                  cartesian=cart)
        return j.set_basis("6-31GP").set_molecule(Z, xyz)

    alone = [make(Z, xyz).construct().solve()[0] for Z, xyz in specs]
    jobs = ds.construct_jobs([make(Z, xyz) for Z, xyz in specs])
    e, it, st = ds.solve_jobs(jobs, use_graph=True, autotune=True)
    assert list(st) == [0] * len(specs)
    assert np.max(np.abs(np.array(alone) - e)) < 1e-9
    sph = make(*specs[0], cart=False).construct()
    b = ds.Batch(0)
    with pytest.raises(ds.DsscfError):
        b.set([jobs[0].context, sph.context], graph=False)
    b.close()
    with pytest.raises(ds.DsscfError) as ei:
        jobs[0].write_molden(str(tmp_path / "c.molden"))
    assert ei.value.code == 4
