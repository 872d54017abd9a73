"""GPU parity of the Fock-build hot path against the oracle (through the C-ABI)."""
import json
import os

import numpy as np
import pytest

import dollarstorescf_amd as ds
from dollarstorescf_amd import geometry as G
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _oracle_basis(basis, Z, xyz):
    aos = O.build_aos(O.read_basis(ds.basis_path(basis)), Z, xyz)
    return O.OracleBasis(aos, Z, xyz)


def _job(basis, Z, xyz, boys, schwarz=0.0):
    j = ds.Job(0).set_options(boys_mode=boys, schwarz_threshold=schwarz)
    return j.set_basis(basis).set_molecule(Z, xyz).construct()


@pytest.mark.parametrize("boys", [ds.BOYS_REFERENCE, ds.BOYS_ACCURATE])
def test_eri_all_unique_water_sto3g(boys):
    """Every unique (ij|kl) of cfg0 water/STO-3G: 1e-12 absolute vs the oracle in the same Boys mode."""
    Z, xyz = G.water()
    job = _job("STO-3G", Z, xyz, boys)
    n = job.n_ao
    q = np.array([[i, j, k, l] for i in range(n) for j in range(i + 1) for k in range(n) for l in range(k + 1)
                  if i * (i + 1) // 2 + j >= k * (k + 1) // 2 + l], np.int32)
    assert len(q) == 406
    got = job.context.eri(q)
    want = _oracle_basis("STO-3G", Z, xyz).eri_list(q, boys)
    assert np.max(np.abs(got - want)) < 1e-12


@pytest.mark.parametrize("boys", [ds.BOYS_REFERENCE, ds.BOYS_ACCURATE])
def test_eri_sample_water_lattice_631g(boys):
    Z, xyz = G.water_lattice(1)
    job = _job("6-31G", Z, xyz, boys)
    n = job.n_ao
    rng = np.random.default_rng(1)
    q = rng.integers(0, n, size=(4000, 4)).astype(np.int32)
    got = job.context.eri(q)
    want = _oracle_basis("6-31G", Z, xyz).eri_list(q, boys)
    assert np.max(np.abs(got - want)) < 1e-12


@pytest.mark.parametrize("basis,nw", [("STO-3G", 1), ("6-31G", 1), ("6-31G", 2)])
def test_jk_vs_tensor(basis, nw):
    """J[Ptot], K[Pa] of a random density vs the reference's stored-tensor loops (oracle)."""
    Z, xyz = G.water() if nw == 1 else G.water_box(nw, seed=3)
    job = _job(basis, Z, xyz, ds.BOYS_REFERENCE)
    n = job.n_ao
    rng = np.random.default_rng(2)
    C = rng.standard_normal((n, 5))
    Pa = O.pack(C @ C.T * 0.1)
    job.set_density(Pa)
    J, K = job.fock_build()
    t = _oracle_basis(basis, Z, xyz).tensor(0)
    Jr, Kr = O.coulomb(t, n, 2 * Pa), O.exchange(t, n, Pa)
    scale = max(1.0, np.max(np.abs(Jr)))
    assert np.max(np.abs(J - Jr)) < 1e-11 * scale
    assert np.max(np.abs(K - Kr)) < 1e-11 * scale


@pytest.mark.parametrize("boys", [ds.BOYS_REFERENCE, ds.BOYS_ACCURATE])
def test_scf_water_sto3g(boys):
    """cfg0: converged RHF energy within 1e-8 Eh of the oracle (DIIS 6, 1e-10 convergence)."""
    Z, xyz = G.water()
    job = ds.Job(0).set_options(diis_size=6, energy_threshold=1e-10, boys_mode=boys, schwarz_threshold=0.0)
    job.set_basis("STO-3G").set_molecule(Z, xyz).construct()
    e, it = job.solve()
    ref = O.scf_energy(ds.basis_path("STO-3G"), Z, xyz, mode=boys, diis_size=8, conv=1e-10)
    assert abs(e - ref.energy) < 1e-8


@pytest.mark.parametrize("boys", [ds.BOYS_REFERENCE, ds.BOYS_ACCURATE])
def test_eri_all_unique_water_def2svp(boys):
    """Every unique (ij|kl) of water/def2-SVP (spherical d): all 21 classes, register and LDS kernels."""
    Z, xyz = G.water()
    job = _job("def2-SVP", Z, xyz, boys)
    n = job.n_ao
    q = np.array([[i, j, k, l] for i in range(n) for j in range(i + 1) for k in range(n) for l in range(k + 1)
                  if i * (i + 1) // 2 + j >= k * (k + 1) // 2 + l], np.int32)
    got = job.context.eri(q)
    want = _oracle_basis("def2-SVP", Z, xyz).eri_list(q, boys)
    assert np.max(np.abs(got - want)) < 1e-12


@pytest.mark.parametrize("basis", ["6-31GP", "def2-SVP"])
def test_eri_sample_d_two_waters(basis):
    Z, xyz = G.water_box(2, seed=5)
    job = _job(basis, Z, xyz, ds.BOYS_REFERENCE)
    rng = np.random.default_rng(7)
    q = rng.integers(0, job.n_ao, size=(5000, 4)).astype(np.int32)
    want = _oracle_basis(basis, Z, xyz).eri_list(q, 0)
    assert np.max(np.abs(job.context.eri(q) - want)) < 1e-12


@pytest.mark.parametrize("basis", ["def2-SVP", "6-31GP"])
def test_jk_vs_tensor_d(basis):
    Z, xyz = G.water()
    job = _job(basis, Z, xyz, ds.BOYS_REFERENCE)
    n = job.n_ao
    C = np.random.default_rng(3).standard_normal((n, 5))
    Pa = O.pack(C @ C.T * 0.1)
    job.set_density(Pa)
    J, K = job.fock_build()
    t = _oracle_basis(basis, Z, xyz).tensor(0)
    Jr, Kr = O.coulomb(t, n, 2 * Pa), O.exchange(t, n, Pa)
    scale = max(1.0, np.max(np.abs(Jr)))
    assert np.max(np.abs(J - Jr)) < 1e-11 * scale
    assert np.max(np.abs(K - Kr)) < 1e-11 * scale


def test_scf_water_def2svp():
    Z, xyz = G.water()
    job = ds.Job(0).set_options(energy_threshold=1e-10, boys_mode=ds.BOYS_REFERENCE, schwarz_threshold=0.0)
    job.set_basis("def2-SVP").set_molecule(Z, xyz).construct()
    e, _ = job.solve()
    ref = O.scf_energy(ds.basis_path("def2-SVP"), Z, xyz, mode=0, conv=1e-10)
    assert abs(e - ref.energy) < 1e-8


@pytest.mark.parametrize("graph", [False, True])
def test_batch_matches_single_context(graph):
    """Fused batch of mixed molecules/bases == each context's own Fock build;
    several replays (graph mode replays the captured launch sequence)."""
    specs = [("STO-3G", G.water()), ("6-31G", G.water_box(2, seed=3)), ("def2-SVP", G.water()),
             ("6-31GP", G.water_box(3, seed=4))]
    jobs, dens, singles = [], [], []
    rng = np.random.default_rng(5)
    for basis, (Z, xyz) in specs:
        j = _job(basis, Z, xyz, ds.BOYS_ACCURATE, schwarz=1e-12)
        n = j.n_ao
        C = rng.standard_normal((n, 4))
        Pa = O.pack(C @ C.T * 0.1)
        j.set_density(Pa)
        singles.append(j.fock_build())
        jobs.append(j)
        dens.append(Pa)
    b = ds.Batch(0)
    b.set([j.context for j in jobs], graph=graph)
    for rep in range(3):
        J, K = b.jk([2 * p for p in dens], dens)
        for (Js, Ks), Jb, Kb in zip(singles, J, K):
            assert np.max(np.abs(Jb - Js)) < 1e-11
            assert np.max(np.abs(Kb - Ks)) < 1e-11
    chosen = b.autotune()
    assert set(int(c) for c in chosen) <= {64, 128, 256}
    J, K = b.jk([2 * p for p in dens], dens)
    for (Js, Ks), Jb, Kb in zip(singles, J, K):
        assert np.max(np.abs(Jb - Js)) < 1e-11


class _DeviceBuffers:
    """Plain device allocations through the HIP runtime the library itself
    links (memory plumbing for tests of the device-pointer entry points)."""

    def __init__(self):
        import ctypes

        self.ct = ctypes
        self.hip = ctypes.CDLL("/opt/rocm/lib/libamdhip64.so")
        self.ptrs = []

    def upload(self, a):
        a = np.ascontiguousarray(a, np.float64)
        p = self.ct.c_void_p()
        assert self.hip.hipMalloc(self.ct.byref(p), self.ct.c_size_t(max(a.nbytes, 8))) == 0
        self.ptrs.append(p)
        assert self.hip.hipMemcpy(p, a.ctypes.data_as(self.ct.c_void_p), self.ct.c_size_t(a.nbytes), 1) == 0
        return p.value

    def download(self, ptr, n):
        out = np.empty(n)
        assert self.hip.hipMemcpy(out.ctypes.data_as(self.ct.c_void_p), self.ct.c_void_p(ptr),
                                  self.ct.c_size_t(out.nbytes), 2) == 0
        return out

    def write(self, ptr, a):
        a = np.ascontiguousarray(a, np.float64)
        assert self.hip.hipMemcpy(self.ct.c_void_p(ptr), a.ctypes.data_as(self.ct.c_void_p),
                                  self.ct.c_size_t(a.nbytes), 1) == 0

    def close(self):
        for p in self.ptrs:
            self.hip.hipFree(p)
        self.ptrs = []


@pytest.mark.parametrize("graph", [False, True])
def test_batch_device_buffers_match_host_io(graph):
    """dsscf_batch_set_device / dsscf_batch_jk_device: the fused plan reads
    the members' densities from caller-owned device buffers and writes J/K
    there in place; equal (1e-11) to each context's own Fock build, over
    several replays with the densities changed between them on the device."""
    specs = [("STO-3G", G.water()), ("6-31G", G.water_box(2, seed=3)), ("def2-SVP", G.alkane(2))]
    jobs, dens = [], []
    rng = np.random.default_rng(11)
    for basis, (Z, xyz) in specs:
        j = _job(basis, Z, xyz, ds.BOYS_ACCURATE, schwarz=1e-12)
        C = rng.standard_normal((j.n_ao, 4))
        jobs.append(j)
        dens.append(O.pack(C @ C.T * 0.1))
    mem = _DeviceBuffers()
    d_pa = [mem.upload(p) for p in dens]
    d_pt = [mem.upload(2 * p) for p in dens]
    d_j = [mem.upload(np.zeros_like(p)) for p in dens]
    d_k = [mem.upload(np.zeros_like(p)) for p in dens]
    b = ds.Batch(0)
    b.set_device([j.context for j in jobs], d_pt, d_pa, d_j, d_k, graph=graph)
    with pytest.raises(ds.DsscfError):
        b.jk([2 * p for p in dens], dens)  # host I/O refused in device mode
    for rep in range(3):
        scale = 1.0 + 0.5 * rep
        for x, y, p in zip(d_pa, d_pt, dens):
            mem.write(x, scale * p)
            mem.write(y, 2 * scale * p)
        b.jk_device()
        for j, p, jd, kd in zip(jobs, dens, d_j, d_k):
            j.set_density(scale * p)
            Js, Ks = j.fock_build()
            assert np.max(np.abs(mem.download(jd, len(p)) - Js)) < 1e-11
            assert np.max(np.abs(mem.download(kd, len(p)) - Ks)) < 1e-11
    b.close()
    mem.close()


@pytest.mark.parametrize("basis,mol", [("6-31G", "water8"), ("def2-SVP", "butane")])
def test_jcd_slot_records_match_atomics(basis, mol):
    """J_cd through slot records and the k_jcd reduction (the default) equals
    J_cd through one atomic add per quartet and element (jcd_max_bytes < 0):
    J within 1e-12 of max|J|, K unchanged, on the 104-AO benchmark system and
    butane/def2-SVP (register and wave-cooperative d classes)."""
    Z, xyz = G.water_lattice(1) if mol == "water8" else G.alkane(4)
    job = _job(basis, Z, xyz, ds.BOYS_ACCURATE, schwarz=1e-12)
    ctx = job.context
    n = job.n_ao
    C = np.random.default_rng(21).standard_normal((n, 8))
    Pa = O.pack(C @ C.T * 0.05)
    out = []
    for cap in (0.0, -1.0):
        ctx.set_options(boys_mode=ds.BOYS_ACCURATE, schwarz_threshold=1e-12, jcd_max_bytes=cap)
        ctx.prepare()
        out.append(ctx.jk(2 * Pa, Pa))
    (J0, K0), (J1, K1) = out
    assert np.max(np.abs(J0 - J1)) < 1e-12 * np.max(np.abs(J1))
    assert np.max(np.abs(K0 - K1)) < 1e-12 * np.max(np.abs(K1))


@pytest.mark.parametrize("basis,mol", [("6-31G", "water8"), ("def2-SVP", "octane")])
def test_ket_order_by_distance_matches_shuffled(basis, mol):
    """Kets ordered by distance from the bra (the default from 1e8 unscreened
    quartets on) and shuffled (the default below) give the same J and K to
    rounding (1e-12 of max |J|, |K|), also through the sub-pairs and wave
    classes of octane/def2-SVP."""
    Z, xyz = G.water_lattice(1) if mol == "water8" else G.alkane(8)
    job = _job(basis, Z, xyz, ds.BOYS_ACCURATE, schwarz=1e-12)
    ctx = job.context
    n = job.n_ao
    C = np.random.default_rng(31).standard_normal((n, 10))
    Pa = O.pack(C @ C.T * 0.05)
    out = []
    for order in (1, 2):
        ctx.set_options(boys_mode=ds.BOYS_ACCURATE, schwarz_threshold=1e-12, ket_order=order)
        ctx.prepare()
        out.append(ctx.jk(2 * Pa, Pa))
    (J1, K1), (J2, K2) = out
    assert np.max(np.abs(J1 - J2)) < 1e-12 * np.max(np.abs(J1))
    assert np.max(np.abs(K1 - K2)) < 1e-12 * np.max(np.abs(K1))


def test_jk_primitive_screening_accurate():
    """Default accurate mode screens primitive quartets (bound 1e-17): J/K stay
    within 1e-11 (relative to max|J|) of the unscreened accurate-mode oracle."""
    Z, xyz = G.water_box(3, seed=4)
    job = ds.Job(0).set_options(boys_mode=ds.BOYS_ACCURATE, schwarz_threshold=0.0)
    job.set_basis("6-31G").set_molecule(Z, xyz).construct()
    st = job.context.stats()
    n = job.n_ao
    assert st.n_prim_quartets < sum(st.class_prim_quartets) + 1  # evaluated count reported
    C = np.random.default_rng(8).standard_normal((n, 6))
    Pa = O.pack(C @ C.T * 0.1)
    job.set_density(Pa)
    J, K = job.fock_build()
    t = _oracle_basis("6-31G", Z, xyz).tensor(1)
    Jr, Kr = O.coulomb(t, n, 2 * Pa), O.exchange(t, n, Pa)
    scale = max(1.0, np.max(np.abs(Jr)))
    assert np.max(np.abs(J - Jr)) < 1e-11 * scale
    assert np.max(np.abs(K - Kr)) < 1e-11 * scale


@pytest.mark.parametrize("basis", ["6-31G", "def2-SVP"])
def test_graph_replays_match_eager_builds(basis):
    """A context's Fock build replays its captured launch sequence (classes
    forked over several streams); the event-timed build makes the same forked
    launches directly; per-class timing launches every kernel on one stream
    in turn.  Several replays, with direct builds and host synchronizations
    between them, agree with the one-stream result (to the summation order of
    the atomic adds)."""
    Z, xyz = G.water_box(3, seed=11)
    job = _job(basis, Z, xyz, ds.BOYS_ACCURATE, schwarz=1e-12)
    ctx, n = job.context, job.n_ao
    C = np.random.default_rng(12).standard_normal((n, 6))
    Pa = O.pack(C @ C.T * 0.1)
    ctx.set_density(2 * Pa, Pa)
    ctx.time_classes(iters=1)  # eager
    J0, K0 = ctx.get_jk(n)
    sj, sk = np.abs(J0).max(), np.abs(K0).max()
    kms = np.zeros(ds.lib().dsscf_class_count())
    for rep in range(4):
        ctx.jk_resident()
        J, K = ctx.get_jk(n)
        assert np.max(np.abs(J - J0)) < 1e-12 * sj, rep
        assert np.max(np.abs(K - K0)) < 1e-12 * sk, rep
        ctx.jk_resident(kms)
        J, K = ctx.get_jk(n)
        assert np.max(np.abs(J - J0)) < 1e-12 * sj, rep
        assert np.max(np.abs(K - K0)) < 1e-12 * sk, rep
        ctx.time_classes(iters=1)
        J, K = ctx.get_jk(n)
        assert np.max(np.abs(J - J0)) < 1e-12 * sj, rep
    assert np.all(kms >= 0.0) and kms.sum() > 0.0


def test_mixed_precision_alkane_def2svp():
    """Quartets with Schwarz bound below 1e-5 run their setup, small-T Boys and
    VRR in FP32 (C4H10/def2-SVP, 106 AOs): J and K at a fixed density stay within
    1e-7 of the all-FP64 build, and the converged RHF energy within 1e-6 Eh."""
    Z, xyz = G.alkane(4)
    res = {}
    for thr in (0.0, 1e-5):
        job = ds.Job(0).set_options(diis_size=8, energy_threshold=1e-9, max_iterations=60, schwarz_threshold=1e-12,
                                    fp32_threshold=thr)
        job.set_basis("def2-SVP").set_molecule(Z, xyz).construct()
        st = job.context.stats()
        n = job.n_ao
        C = np.random.default_rng(21).standard_normal((n, 17)) * 0.2
        Pa = O.pack(C @ C.T)
        job.set_density(Pa)
        J, K = job.fock_build()
        e, it = job.solve()
        res[thr] = (J, K, e, st.n_quartets_fp32, st.n_quartets)
    J64, K64, e64, nf64, nq = res[0.0]
    J32, K32, e32, nf32, nq32 = res[1e-5]
    assert n == 106 and nf64 == 0 and nq32 == nq
    assert 0.2 * nq < nf32 < nq  # a large share of the quartets ran in FP32
    assert np.max(np.abs(J32 - J64)) < 1e-7
    assert np.max(np.abs(K32 - K64)) < 1e-7
    assert abs(e32 - e64) < 1e-6


@pytest.mark.parametrize("use_graph,copies", [(False, 1), (True, 1), (True, 3)])
def test_jobs_solved_together_match_alone(use_graph, copies):
    """Uneven jobs (several bases and sizes) solved in lock step with fused
    Fock builds (and the block-size autotuner) converge to the energies each
    reaches alone, in the same number of iterations.  Three copies of the six
    jobs (18 jobs) run as two alternating groups (one group's fused build
    beside the other's host steps)."""
    specs = [("STO-3G", G.water()), ("6-31G", G.water_box(2, seed=3)), ("6-31GP", G.water_box(3, seed=4)),
             ("def2-SVP", G.water()), ("6-31G", G.alkane(2)), ("STO-3G", G.water_box(4, seed=9))]

    def make(basis, Z, xyz):
        job = ds.Job(0).set_options(diis_size=8, energy_threshold=1e-9, max_iterations=80, schwarz_threshold=1e-12)
        return job.set_basis(basis).set_molecule(Z, xyz)

    alone = []
    for basis, (Z, xyz) in specs:
        j = make(basis, Z, xyz)
        j.construct()
        alone.append(j.solve())
    jobs = ds.construct_jobs([make(b, Z, xyz) for b, (Z, xyz) in specs * copies])
    e, it, st = ds.solve_jobs(jobs, use_graph=use_graph, autotune=True)
    assert list(st) == [0] * len(specs) * copies
    for (ea, ia), eb, ib in zip(alone * copies, e, it):
        assert abs(ea - eb) < 1e-9
        assert abs(ia - ib) <= 1


def _serve_job(basis, Z, xyz, fp32=0.0):
    job = ds.Job(0).set_options(diis_size=8, energy_threshold=1e-9, max_iterations=100, schwarz_threshold=1e-12,
                                fp32_threshold=fp32)
    return job.set_basis(basis).set_molecule(Z, xyz)


@pytest.mark.parametrize("fp32,device_steps", [(0.0, False), (1e-5, False), (0.0, True), (1e-5, True)])
def test_server_result_same_alone_and_in_mix(fp32, device_steps):
    """A job served alone and the same job served inside a mix of other
    molecules and bases (arriving while others iterate, with a footprint
    budget that forces queueing) converge to the same energy, within 1e-9 Eh;
    every job in the mix equals its own Job.solve energy within 1e-9 Eh
    (FP64) and, with FP32 below Schwarz 1e-5, within 1e-6 Eh of it.  With
    device steps the densities, J/K, Fock matrices and DIIS histories stay
    in HBM, the fused builds read and write them in place and each round's
    iterations run as one grouped launch sequence (density by purification);
    the reference energies are still the host-step Job.solve ones."""
    probe = ("6-31G", G.water_box(3, seed=4))
    others = [("STO-3G", G.water()), ("def2-SVP", G.water()), ("6-31G", G.alkane(3)),
              ("STO-3G", G.water_box(5, seed=2)), ("6-31GP", G.water_box(2, seed=7)), ("def2-SVP", G.alkane(1))]
    srv = ds.Server(0, budget_bytes=8e9, max_active=8, workers=3, device_steps=device_steps)
    srv.submit(_serve_job(probe[0], *probe[1], fp32=fp32), 0)
    alone = srv.run_until_idle()
    assert len(alone) == 1 and alone[0]["status"] == 0
    srv.close()
    # the mix: a small budget and four slots force admission in several waves
    srv = ds.Server(0, budget_bytes=2.5e8, max_active=4, workers=3, device_steps=device_steps)
    specs = others[:3] + [probe] + others[3:]
    srv.submit(_serve_job(specs[0][0], *specs[0][1], fp32=fp32), 0)
    srv.submit(_serve_job(specs[1][0], *specs[1][1], fp32=fp32), 1)
    done = []
    for _ in range(3):
        srv.step()
        done += srv.poll()
    for k in range(2, len(specs)):
        srv.submit(_serve_job(specs[k][0], *specs[k][1], fp32=fp32), k)
        srv.step()
        done += srv.poll()
    done += srv.run_until_idle()
    srv.close()
    assert sorted(d["tag"] for d in done) == list(range(len(specs)))
    by_tag = {d["tag"]: d for d in done}
    assert all(d["status"] == 0 for d in done)
    assert abs(by_tag[3]["energy"] - alone[0]["energy"]) < 1e-9
    for k, (basis, (Z, xyz)) in enumerate(specs):
        j = _serve_job(basis, Z, xyz)  # plain FP64, alone
        j.construct()
        e, _ = j.solve()
        assert abs(by_tag[k]["energy"] - e) < (1e-9 if fp32 == 0.0 else 1e-6), (k, basis)


@pytest.mark.parametrize("key", ["water_sto3g", "water_def2svp"])
def test_one_electron_device_matches_reference_fixture(key):
    """S, T, V computed on the device (reference Boys series) against the
    matrices the reference's own integral routines produced (1e-12)."""
    c = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "reference_integrals.json")))[key]
    job = ds.Job(0).set_options(boys_mode=ds.BOYS_REFERENCE, schwarz_threshold=0.0)
    job.set_basis(c["basis"]).set_molecule(np.array(c["Z"], np.int32), np.array(c["xyz_bohr"])).construct()
    for name in ("S", "T", "V"):
        assert np.max(np.abs(job.get(name) - np.array(c[name]))) < 1e-12, name


@pytest.mark.parametrize("basis,mol", [("6-31GP", "water3"), ("def2-SVP", "alkane3"), ("STO-3G", "alkane1")])
def test_one_electron_device_matches_host(basis, mol):
    """Device S, T, V (Job.construct) equal the host restatement
    (Job.construct_host) in the accurate Boys mode, d shells included."""
    Z, xyz = G.water_box(3, seed=5) if mol == "water3" else G.alkane(int(mol[-1]))
    dev = ds.Job(0).set_basis(basis).set_molecule(Z, xyz).construct()
    host = ds.Job(0).set_basis(basis).set_molecule(Z, xyz).construct_host()
    for name in ("S", "T", "V", "H"):
        a, b = dev.get(name), host.get(name)
        assert np.max(np.abs(a - b)) < 1e-12 * max(1.0, np.abs(b).max()), name


@pytest.mark.parametrize("mode", [1, 2])
def test_device_boys_vs_mpmath(mode):
    """The device's accurate Boys forms (1: 801-row Taylor table, 2: interval
    table in LDS, the FP64 J/K kernels' form) for F_0..F_L, L <= 8, against
    50-digit mpmath values: 2e-15 relative, plus 1e-18 absolute for the
    asymptotic form at T >= 40, which drops terms of order e^-T / 2T."""
    g = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "boys.json")))
    x = np.array(g["x"])
    mp = np.array(g["mpmath"])
    ctx = ds.Context(0)
    for L in range(9):
        F = ctx.boys(x, L, mode)
        ref = mp[:, :L + 1]
        assert np.all(np.abs(F - ref) <= 2e-15 * ref + 1e-18), L


def test_device_boys_interval_matches_taylor_table():
    """Interval table vs 801-row table on a dense grid across every interval
    edge and the T = 40 switch: the two accurate forms agree to 4e-15 relative."""
    T = np.concatenate([np.linspace(0.0, 45.0, 90001), np.arange(1, 81) * 0.5 - 1e-13, np.arange(0, 81) * 0.5])
    ctx = ds.Context(0)
    for L in (0, 1, 4, 8):
        a = ctx.boys(T, L, 1)
        b = ctx.boys(T, L, 2)
        assert np.all(np.abs(a - b) <= 4e-15 * np.abs(a)), L


@pytest.mark.parametrize("boys,device_step", [(ds.BOYS_REFERENCE, False), (ds.BOYS_ACCURATE, False),
                                              (ds.BOYS_ACCURATE, True)])
def test_scf_config1_water_lattice_vs_oracle(boys, device_step):
    """Config 1 (8-water 2x2x2 lattice, RHF/6-31G, 104 AOs, Schwarz 1e-12,
    converged to 1e-10): total energy within 1e-8 Eh of the oracle's RHF over
    the full stored tensor in the same Boys mode (the reference's series in
    reference mode), and each energy term within 1e-7 Eh."""
    g = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "energies.json")))
    ref = g["cases"]["water_lattice8_631g"]["reference" if boys == ds.BOYS_REFERENCE else "accurate"]
    Z, xyz = G.water_lattice(1)
    job = ds.Job(0).set_options(diis_size=8, energy_threshold=1e-10, max_iterations=80, boys_mode=boys,
                                schwarz_threshold=1e-12, device_step=device_step)
    job.set_basis("6-31G").set_molecule(Z, xyz).construct()
    assert job.n_ao == 104
    e, _ = job.solve()
    assert abs(e - ref["energy"]) < 1e-8, (e, ref["energy"])
    got = job.energies()
    for k, name in (("kinetic", "kinetic"), ("nuclear_attraction", "nuclear_attraction"), ("coulomb", "coulomb"),
                    ("exchange", "exchange"), ("nuclear", "nuclear")):
        assert abs(got[k] - ref["components"][name]) < 1e-7, k
