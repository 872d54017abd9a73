"""Fock-build benchmark on one MI355X.

Workload: 8 rigid waters on a seeded 2x2x2 lattice (3.0 A spacing, +-0.3 A
jitter, random orientations), RHF/6-31G, 104 AOs, s/p shells only, Schwarz
threshold 1e-12, FP64 throughout.  A step is one Fock build: every surviving
unique shell quartet evaluated (Boys, OS/HGP VRR, HRR) and digested into J
and K, with the density resident in HBM.  The density is the converged SCF density of the same
job (at most 30 iterations).

Prints one JSON line (metric, value = Fock builds/s, roofline of the dominant
class kernel, CPU baseline of the reference's own ERI routine).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "Fock builds/s, ERI quartets/s; SCF/s + p99 vs FP64 plain at |dE|<=1e-6 Eh"
FP64_PEAK_TFLOPS = 78.6


def cpu_model() -> str:
    """The host CPU's model name (as lscpu prints it), from /proc/cpuinfo."""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def host_cores() -> int:
    try:
        return len(os.sched_getaffinity(0))
    except AttributeError:
        return os.cpu_count() or 1


def packed_quartets(n: int) -> np.ndarray:
    """Every unique AO quartet (ij|kl) in the order of the reference's packed
    tensor (ExactCoulomb.cpp:58-61): pair index m = j + i(i+1)/2 <= n = l + k(k+1)/2,
    m-major."""
    ii, jj = np.tril_indices(n)
    order = np.argsort(jj + ii * (ii + 1) // 2, kind="stable")
    pi, pj = ii[order].astype(np.int32), jj[order].astype(np.int32)
    D = len(pi)
    m = np.repeat(np.arange(D, dtype=np.int64), D - np.arange(D))
    n_idx = np.concatenate([np.arange(k, D, dtype=np.int64) for k in range(D)])
    return np.stack([pi[m], pj[m], pi[n_idx], pj[n_idx]], axis=1)


def cpu_baseline(ctx, Z, xyz, n_ao, Pa, J_gpu, K_gpu, sample, jk_iters=3, seed=7):
    """The reference's own Fock build on the host, in its two parts.

    * Per iteration (the reported value): the stored-tensor J and K loops of
      getExactCoulombMatrix / getExactExchangeMatrix (ExactCoulomb.cpp:103-123,
      as computeCoulombTerms / computeExchangeTerms call them every iteration,
      SCFSolver.cpp:29-43), restated in oracle/oracle.c, on 1 core over the real
      packed tensor of this 104-AO system (14.9 M doubles, filled with this
      engine's integrals: the loops' time does not depend on the values, and
      their J/K are checked against the GPU's).  Fock builds/s = 1 / (J + K).
    * One-off: the tensor build (assignExactRepulsionTensor, ExactCoulomb.cpp:
      41-72), i.e. the reference's electronRepulsionIntegral (oracle/_ref, built
      from its sources) over every unique AO quartet, extrapolated from a
      random sample of them."""
    from oracle import oracle as O
    from oracle import ref

    q = packed_quartets(n_ao)
    t0 = time.perf_counter()
    tensor = ctx.eri(q)
    t_fill = time.perf_counter() - t0
    del q
    O.set_threads(1)
    Ptot = 2.0 * Pa
    t_j = t_k = 0.0
    for _ in range(jk_iters):
        t0 = time.perf_counter()
        J = O.coulomb(tensor, n_ao, Ptot)
        t1 = time.perf_counter()
        K = O.exchange(tensor, n_ao, Pa)
        t_j += t1 - t0
        t_k += time.perf_counter() - t1
    t_j /= jk_iters
    t_k /= jk_iters
    P = host_cores()
    O.set_threads(P)
    t0 = time.perf_counter()
    for _ in range(jk_iters):
        O.coulomb(tensor, n_ao, Ptot)
        O.exchange(tensor, n_ao, Pa)
    t_jk_p = (time.perf_counter() - t0) / jk_iters
    O.set_threads(1)
    n_unique = len(tensor)
    tensor_s = None
    if ref.available():
        rm = ref.RefMolecule(os.path.join(ROOT, "dollarstorescf_amd", "basis", "6-31G"), Z, xyz)
        rng = np.random.default_rng(seed)
        qs = packed_quartets(n_ao)[rng.integers(0, n_unique, size=sample)]
        t, _ = rm.time_eri_list(np.ascontiguousarray(qs, np.int32))
        tensor_s = t / sample * n_unique
    scale = max(1.0, float(np.max(np.abs(J_gpu))))
    return {"value": 1.0 / (t_j + t_k), "unit": "Fock builds/s", "cores": 1, "kind": "port",
            "cpu_model": cpu_model(), "host_cores": P,
            "per_iteration_s": {"J": t_j, "K": t_k, "J+K": t_j + t_k},
            "per_iteration_s_all_cores": t_jk_p,
            "tensor_build_s_extrapolated": tensor_s,
            "jk_vs_gpu_max_abs": {"J": float(np.max(np.abs(J - J_gpu))), "K": float(np.max(np.abs(K - K_gpu))),
                                  "scale": scale},
            "sample": f"per-iteration J+K: the reference's stored-tensor loops (ExactCoulomb.cpp:103-123, restated "
                      f"in oracle/oracle.c) over the {n_unique}-element packed tensor of this system at the "
                      f"converged density, mean of {jk_iters} on 1 core ({P}-thread figure beside it); one-off "
                      f"tensor build: the reference's electronRepulsionIntegral (GaussianIntegral.cpp:561, "
                      f"oracle/_ref) on {sample} random unique AO quartets, extrapolated to all {n_unique} "
                      f"(labelled extrapolated); tensor filled from this engine's ERIs in {t_fill:.1f} s"}


def traffic_of(kernel: str):
    """HBM bytes per dispatch of `kernel` on this workload from the committed
    rocprofv3 FETCH_SIZE/WRITE_SIZE summary (each counter scaled by what it reads
    for a device copy of known size), or None when the summary lacks it."""
    path = os.path.join(ROOT, "profiles", "fock_water8_631g_traffic.json")
    try:
        with open(path) as f:
            rec = json.load(f)["kernels"].get(kernel)
    except (OSError, ValueError, KeyError):
        return None
    return None if rec is None else float(rec["hbm_bytes"])


def copy_calibration(mb: int, reps: int = 5):
    """A plain device copy of known size, for scaling the FETCH_SIZE/WRITE_SIZE
    counters (torch is used here as memory plumbing only)."""
    import torch

    n = mb * (1 << 20) // 8
    x = torch.ones(n, dtype=torch.float64, device="cuda")
    y = torch.empty_like(x)
    for _ in range(reps):
        y.copy_(x)
    torch.cuda.synchronize()
    print(json.dumps({"copy_mib": mb, "reps": reps, "bytes_read_per_copy": n * 8, "bytes_written_per_copy": n * 8}))


def bench_alkane(a):
    """All-trans C20H42 in def2-SVP (490 AOs, classes up to (dd|dd)), Schwarz
    1e-12: the all-FP64 path against FP32 setup/Boys/VRR for quartets whose
    Schwarz bound is below 1e-5.  Both SCF runs start from the core guess; the
    line reports Fock builds/s of both at their converged densities, the
    per-iteration energy differences and the final |dE|."""
    import dollarstorescf_amd as ds
    from dollarstorescf_amd import flops, geometry

    Z, xyz = geometry.alkane(20)
    classes = ds.class_list()
    runs = {}
    Pa_ref = None
    for mode, thr in (("fp64", 0.0), ("mixed", 1e-5)):
        job = ds.Job(0).set_options(diis_size=8, energy_threshold=1e-8, max_iterations=60, schwarz_threshold=1e-12,
                                    fp32_threshold=thr)
        job.set_basis("def2-SVP").set_molecule(Z, xyz)
        t0 = time.perf_counter()
        job.construct()
        t_construct = time.perf_counter() - t0
        t0 = time.perf_counter()
        energy, iters = job.solve()
        t_scf = time.perf_counter() - t0
        ctx = job.context
        st = ctx.stats()
        ctx.autotune()  # block sizes for the repeated builds below (not timed)
        Pa = job.get("Pa")
        ctx.set_density(2.0 * Pa, Pa)
        for _ in range(a.warmup):
            ctx.jk_resident()
        t0 = time.perf_counter()
        for _ in range(a.steps):
            ctx.jk_resident()
        ms = (time.perf_counter() - t0) / a.steps * 1e3
        # per class, its kernels one after another on one stream (FP64 and FP32 lists summed)
        per_cls = ctx.time_classes(iters=max(3, a.steps // 4))[:len(classes)]
        # J/K of both precisions at one density: the all-FP64 run's converged one
        if mode == "fp64":
            Pa_ref = Pa
        ctx.set_density(2.0 * Pa_ref, Pa_ref)
        ctx.jk_resident()
        J, K = ctx.get_jk(job.n_ao)
        cq = list(st.class_quartets)[:len(classes)]
        cp = list(st.class_prim_quartets)[:len(classes)]
        cq32 = list(st.class_quartets_fp32)[:len(classes)]
        runs[mode] = {"fock_builds_per_s": 1e3 / ms, "ms_per_build": ms, "energy": energy, "iterations": iters,
                      "scf_seconds": t_scf, "construct_seconds": t_construct, "n_ao": job.n_ao,
                      "shell_quartets": int(st.n_quartets), "fp32_shell_quartets": int(st.n_quartets_fp32),
                      "prim_quartets": int(st.n_prim_quartets),
                      "fock_build_tflops": flops.fock_build_flops(classes, cq, cp) / (ms * 1e-3) / 1e12,
                      "history": [float(x) for x in job.history()],
                      "_cls": per_cls, "_q": cq, "_q32": cq32, "_p": cp, "_jk": (J, K)}
    h0, h1 = runs["fp64"].pop("history"), runs["mixed"].pop("history")
    c64, c32 = runs["fp64"].pop("_cls"), runs["mixed"].pop("_cls")
    q64, q32 = runs["fp64"].pop("_q"), runs["mixed"].pop("_q32")
    runs["mixed"].pop("_q")
    runs["fp64"].pop("_q32")
    p64 = runs["fp64"].pop("_p")
    runs["mixed"].pop("_p")
    (J64, K64), (J32, K32) = runs["fp64"].pop("_jk"), runs["mixed"].pop("_jk")
    # per class, one stream, classes in turn; FP64 roofline: algorithmic FLOPs
    # (dollarstorescf_amd/flops.py) / the class's FP64 duration / 78.6 TFLOP/s
    per_class = []
    for c, q, pq, f, m64, m32 in zip(classes, q64, p64, q32, c64, c32):
        if q <= 0:
            continue
        tf = flops.class_flops(c, q, pq) / (m64 * 1e-3) / 1e12 if m64 > 0 else 0.0
        per_class.append({"class": "".join("spdf"[l] for l in c), "quartets": int(q), "prim_quartets": int(pq),
                          "fp32_quartets": int(f), "fp64_ms": float(m64), "mixed_ms": float(m32),
                          "fp64_tflops": tf, "fp64_frac": tf / FP64_PEAK_TFLOPS})
    out = {
        "metric": METRIC, "value": runs["mixed"]["fock_builds_per_s"], "unit": "Fock builds/s", "n_gpus": 1,
        "steps": a.steps, "warmup": a.warmup, "ms_per_step": runs["mixed"]["ms_per_build"],
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64+f32",
        "data": "synthetic (all-trans alkane geometry)",
        "config": {"workload": "RHF/def2-SVP Fock build, all-trans C20H42 (C-C 1.54 A, C-H 1.09 A, tetrahedral), "
                               "490 AOs, Schwarz 1e-12, FP32 below Schwarz 1e-5", "parallelism": "single GPU"},
        "fp64": runs["fp64"], "mixed": runs["mixed"],
        "final_dE": runs["mixed"]["energy"] - runs["fp64"]["energy"],
        "per_iteration_dE": [b - c for b, c in zip(h1, h0)],
        "jk_max_abs_diff_at_fp64_density": {"J": float(np.max(np.abs(J32 - J64))),
                                            "K": float(np.max(np.abs(K32 - K64))),
                                            "max_abs_J": float(np.max(np.abs(J64)))},
        "classes": per_class,
    }
    print(json.dumps(out))


def bench_scf(a):
    """Whole SCF solves from the core guess with the iteration's host part
    (DIIS, X F X, diagonalization, density) on the CPU and on the device, on
    the 8-water/6-31G lattice (104 AOs) and all-trans C20H42/def2-SVP (490 AOs).
    Construction is not timed; the Fock builds are the same device path in
    both modes."""
    import dollarstorescf_amd as ds
    from dollarstorescf_amd import geometry

    Zw, xw = geometry.water()
    for step in (False, True):  # load rocBLAS/rocSOLVER kernels and warm the device
        ds.Job(0).set_options(device_step=step).set_basis("STO-3G").set_molecule(Zw, xw).construct().solve()
    systems = [("water8_631g", "6-31G", geometry.water_lattice(1)), ("c20h42_def2svp", "def2-SVP", geometry.alkane(20))]
    rows = {}
    for name, basis, (Z, xyz) in systems:
        row = {}
        for mode, step in (("host_step", False), ("device_step", True)):
            job = ds.Job(0).set_options(diis_size=8, energy_threshold=1e-8, max_iterations=80,
                                        schwarz_threshold=1e-12, device_step=step)
            job.set_basis(basis).set_molecule(Z, xyz).construct()
            t0 = time.perf_counter()
            e, it = job.solve()
            dt = time.perf_counter() - t0
            row[mode] = {"seconds": dt, "iterations": it, "energy": e, "ms_per_iteration": dt / (it + 1) * 1e3,
                         "n_ao": job.n_ao}
        row["speedup"] = row["host_step"]["seconds"] / row["device_step"]["seconds"]
        row["dE"] = row["device_step"]["energy"] - row["host_step"]["energy"]
        rows[name] = row
    big = rows["c20h42_def2svp"]["device_step"]
    out = {"metric": METRIC, "value": 1.0 / big["seconds"], "unit": "SCF solves/s", "n_gpus": 1, "steps": 1,
           "warmup": 1, "ms_per_step": big["seconds"] * 1e3, "higher_is_better": True, "scaling": "weak",
           "vs_baseline": None, "dtype": "f64", "data": "synthetic (seeded water lattice, all-trans alkane)",
           "config": {"workload": "RHF SCF from the core guess to |dE| < 1e-8, device vs host iteration step: "
                                  "8-water/6-31G (104 AOs) and C20H42/def2-SVP (490 AOs), Schwarz 1e-12, FP64",
                      "parallelism": "single GPU"},
           "systems": rows}
    print(json.dumps(out))


def bench_gradient(a):
    """RHF nuclear gradient of the 8-water lattice (6-31G: 104 AOs, s/p
    shells; def2-SVP: 200 AOs with d shells) at its converged density: the
    two-electron part on the device (every surviving shell quartet, Schwarz
    1e-12), the one-electron and nuclear-repulsion parts on the host.  CPU
    baseline: the reference's electronRepulsionGradient (oracle/_ref, 1
    core) on a sample of unique AO quartets, extrapolated to all of them."""
    import dollarstorescf_amd as ds
    from dollarstorescf_amd import geometry

    Z, xyz = geometry.water_lattice(1)
    xyz = np.asarray(xyz, float).reshape(-1, 3)
    job = ds.Job(0).set_options(diis_size=8, energy_threshold=1e-10, max_iterations=60, schwarz_threshold=1e-12)
    job.set_basis(a.basis).set_molecule(Z, xyz).construct()
    energy, iters = job.solve()
    n = job.n_ao
    ao_atom = np.array([int(np.argmin(np.linalg.norm(xyz - job.ao(i)["center"], axis=1))) for i in range(n)], np.int32)
    Pa = job.get("Pa")
    ctx = job.context
    for _ in range(a.warmup):
        ctx.eri_gradient(Pa, ao_atom, len(Z))
    t0 = time.perf_counter()
    for _ in range(a.steps):
        ctx.eri_gradient(Pa, ao_atom, len(Z))
    ms_eri = (time.perf_counter() - t0) / a.steps * 1e3
    t0 = time.perf_counter()
    for _ in range(a.steps):
        g = job.gradient()
    ms_total = (time.perf_counter() - t0) / a.steps * 1e3
    cpu = None
    if not a.no_cpu:
        from oracle import ref
        if ref.available():
            rm = ref.RefMolecule(ds.basis_path(a.basis), Z, xyz.reshape(-1))
            rng = np.random.default_rng(5)
            sample = max(1000, a.cpu_sample // 10)
            q = rng.integers(0, n, size=(sample, 4)).astype(np.int32)
            t = rm.time_eri_gradient(q)
            n_unique = (n * (n + 1) // 2) * (n * (n + 1) // 2 + 1) // 2
            cpu = {"value": 1.0 / (t / sample * n_unique), "unit": "ERI gradients/s", "cores": 1, "kind": "reference",
                   "sample": f"{sample} random AO quartets of the same system through the reference's "
                             f"electronRepulsionGradient (GaussianIntegral.cpp:1052), {t:.1f} s on 1 core, "
                             f"extrapolated to {n_unique} unique AO quartets"}
    out = {"metric": METRIC, "value": 1e3 / ms_eri, "unit": "ERI gradients/s", "n_gpus": 1, "steps": a.steps,
           "warmup": a.warmup, "ms_per_step": ms_eri, "higher_is_better": True, "scaling": "weak",
           "vs_baseline": None, "dtype": "f64", "data": "synthetic (seeded water lattice geometry)",
           "config": {"workload": f"RHF/{a.basis} nuclear gradient, 8-water 2x2x2 lattice, {n} AOs, Schwarz 1e-12, "
                                  "converged density", "parallelism": "single GPU"},
           "ms_full_gradient": ms_total, "scf_energy": energy, "scf_iterations": iters,
           "max_abs_gradient": float(np.max(np.abs(g))), "sum_over_atoms": np.abs(g.sum(axis=0)).max(),
           "cpu_baseline": cpu}
    print(json.dumps(out))


def batch_specs(n_jobs: int, seed: int = 2024):
    """n_jobs molecules of 1-6 rigid waters (uniform), each placed by rejection
    sampling in a 6 A box with minimum O-O distance 2.6 A and random
    orientations; every job has its own seed drawn from one generator."""
    from dollarstorescf_amd import geometry

    rng = np.random.default_rng(seed)
    sizes = rng.integers(1, 7, size=n_jobs)
    seeds = rng.integers(0, 2**31 - 1, size=n_jobs)
    return [geometry.water_box(int(k), seed=int(s)) for k, s in zip(sizes, seeds)]


def bench_batch(a):
    """512 independent RHF/6-31G* jobs with Cartesian d (6D; 19-114 AOs).
    Fused: the jobs are constructed together (host parts in parallel) and
    solved in lock step, each round's Fock builds of all unconverged jobs one
    launch per quartet class (hipGraph replays, block sizes from the
    autotuner).  Loop: the same jobs solved one after another, each with its
    own context.  Both time the SCF solves (construction reported apart);
    value = fused SCF/s."""
    import dollarstorescf_amd as ds

    specs = batch_specs(a.jobs)

    def make(Z, xyz, schedule=None):
        job = ds.Job(0).set_options(diis_size=8, energy_threshold=1e-8, max_iterations=100, schwarz_threshold=1e-12,
                                    cartesian=True)
        return job.set_basis("6-31GP").set_molecule(Z, xyz).set_schedule(schedule)

    # the project's schedule autotuner on a sample of the same workload (every
    # 16th job, 1-6 waters): per class sub-pair size, lane groups, kets per
    # unit and block size, then applied to all jobs (set-up, not timed)
    t0 = time.perf_counter()
    sample = ds.construct_jobs([make(Z, xyz) for Z, xyz in specs[::16]])
    tb = ds.Batch(0)
    tb.set([j.context for j in sample], graph=False)
    schedule = tb.autotune_schedule()
    tb.close()
    for j in sample:
        j.close()
    t_tune = time.perf_counter() - t0

    t0 = time.perf_counter()
    jobs = ds.construct_jobs([make(Z, xyz, schedule) for Z, xyz in specs])
    t_construct_fused = time.perf_counter() - t0
    n_ao = [j.n_ao for j in jobs]
    t0 = time.perf_counter()
    e_fused, it_fused, st = ds.solve_jobs(jobs, use_graph=True, autotune=True)
    t_fused = time.perf_counter() - t0
    for j in jobs:
        j.close()
    del jobs

    e_loop, it_loop, t_loop, t_construct_loop = [], [], 0.0, 0.0
    for Z, xyz in specs:
        j = make(Z, xyz)
        t0 = time.perf_counter()
        j.construct()
        t1 = time.perf_counter()
        e, it = j.solve()
        t_loop += time.perf_counter() - t1
        t_construct_loop += t1 - t0
        e_loop.append(e)
        it_loop.append(it)
        j.close()
    de = np.abs(np.array(e_loop) - e_fused)
    out = {
        "metric": METRIC, "value": len(specs) / t_fused, "unit": "SCF solves/s", "n_gpus": 1,
        "steps": len(specs), "warmup": 0, "ms_per_step": t_fused / len(specs) * 1e3, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic (seeded water boxes)",
        "config": {"workload": f"{len(specs)} RHF/6-31G* jobs (Cartesian d, 6D), 1-6 waters in a 6 A box, min O-O "
                               "2.6 A, seed 2024; Schwarz 1e-12; converge |dE| < 1e-8",
                   "n_ao_min": int(min(n_ao)), "n_ao_max": int(max(n_ao)), "parallelism": "single GPU"},
        "schedule": {"tune_seconds": t_tune, "sample_jobs": len(specs[::16]),
                     "per_class": {"".join("spdf"[l] for l in c): {k: (float(v[i]) if k != "kets_per_unit" and
                                                                     k != "block_size" else int(v[i]))
                                                                 for k, v in schedule.items()}
                                   for i, c in enumerate(ds.class_list())}},
        "fused": {"solves_per_s": len(specs) / t_fused, "seconds": t_fused, "construct_seconds": t_construct_fused,
                  "iterations_total": int(np.sum(it_fused)), "failed": int(np.count_nonzero(st))},
        "loop": {"solves_per_s": len(specs) / t_loop, "seconds": t_loop, "construct_seconds": t_construct_loop,
                 "iterations_total": int(np.sum(it_loop))},
        "speedup_fused_vs_loop": t_loop / t_fused,
        "max_abs_dE_fused_vs_loop": float(de.max()),
    }
    print(json.dumps(out))


def serve_requests(n_req: int, rate: float, seed: int = 7):
    """Seeded Poisson stream: arrival times (s) and (basis, Z, xyz) per request.
    Basis uniform over STO-3G, 6-31G, def2-SVP; molecule a water cluster of
    1-12 waters (6 A box for up to 6, grown at constant density beyond) or an
    all-trans alkane CnH2n+2, n = 1-12, with equal odds."""
    from dollarstorescf_amd import geometry

    rng = np.random.default_rng(seed)
    arrivals = np.cumsum(rng.exponential(1.0 / rate, size=n_req))
    out = []
    for _ in range(n_req):
        basis = ["STO-3G", "6-31G", "def2-SVP"][int(rng.integers(0, 3))]
        k = int(rng.integers(1, 13))
        if rng.random() < 0.5:
            Z, xyz = geometry.water_box(k, seed=int(rng.integers(0, 2**31 - 1)), box=6.0 * max(1.0, k / 6.0) ** (1 / 3))
        else:
            Z, xyz = geometry.alkane(k)
        out.append((basis, Z, xyz))
    return arrivals, out


def serve_stream(ds, arrivals, reqs, fp32: float, max_active: int, budget: float, device_steps: bool = False,
                 conv: float = 1e-8):
    """Replay the stream in real time against one server; returns per-request
    latency (finish - scheduled arrival), energies, status and the wall time
    from the first arrival to the last finish.  conv: the jobs' |dE_elec|
    stopping threshold."""
    jobs = []
    for basis, Z, xyz in reqs:
        job = ds.Job(0).set_options(diis_size=8, energy_threshold=conv, max_iterations=100, schwarz_threshold=1e-12,
                                    fp32_threshold=fp32)
        jobs.append(job.set_basis(basis).set_molecule(Z, xyz))
    srv = ds.Server(0, budget_bytes=budget, max_active=max_active, workers=12, device_steps=device_steps)
    n = len(reqs)
    t0 = srv.clock()
    i, done = 0, {}
    while len(done) < n:
        now = srv.clock() - t0
        while i < n and arrivals[i] <= now:
            srv.submit(jobs[i], i)
            i += 1
        if srv.idle():
            if i < n:
                time.sleep(max(0.0, arrivals[i] - (srv.clock() - t0)))
            continue
        srv.step()
        for d in srv.poll():
            done[d["tag"]] = d
    stats = srv.stats()
    srv.close()
    lat = np.array([done[k]["t_done"] - t0 - arrivals[k] for k in range(n)])
    e = np.array([done[k]["energy"] for k in range(n)])
    st = np.array([done[k]["status"] for k in range(n)])
    wall = max(d["t_done"] for d in done.values()) - t0 - arrivals[0]
    for j in jobs:
        j.close()
    return lat, e, st, wall, stats


def bench_serve(a):
    """Serving: a seeded Poisson stream (mean a.rate requests/s) of RHF
    requests, replayed in real time against the continuous-batching server
    (jobs constructed by worker threads, admitted by device-memory budget,
    one fused Fock build per round over every job on the device).  The
    stream runs with FP32 below Schwarz 1e-5 (value) and again with every
    quartet in FP64; reported: p50/p99 latency, solves/s, |dE| per request.
    Every job's densities, J/K, Fock matrices and DIIS history stay in HBM and
    each round runs the iterations of all its jobs as one grouped launch
    sequence (density by purification); --host-steps runs the iterations'
    linear algebra on host worker threads instead."""
    import dollarstorescf_amd as ds

    arrivals, reqs = serve_requests(a.requests, a.rate)
    runs = {}
    for mode, thr in (("mixed", 1e-5), ("fp64", 0.0)):
        lat, e, st, wall, stats = serve_stream(ds, arrivals, reqs, thr, max_active=64, budget=32e9,
                                               device_steps=not a.host_steps)
        runs[mode] = {"p50_latency_s": float(np.percentile(lat, 50)), "p99_latency_s": float(np.percentile(lat, 99)),
                      "mean_latency_s": float(lat.mean()), "solves_per_s": len(reqs) / wall, "wall_s": wall,
                      "failed": int(np.count_nonzero(st)), "server": stats, "_e": e}
    de = np.abs(runs["mixed"].pop("_e") - runs["fp64"].pop("_e"))
    n_ao = []
    for basis, Z, xyz in reqs:
        j = ds.Job(0).set_basis(basis).set_molecule(Z, xyz)
        j.construct_host()
        n_ao.append(j.n_ao)
        j.close()
    out = {
        "metric": METRIC, "value": runs["mixed"]["solves_per_s"], "unit": "SCF solves/s", "n_gpus": 1,
        "steps": len(reqs), "warmup": 0, "ms_per_step": runs["mixed"]["wall_s"] / len(reqs) * 1e3,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64+f32",
        "data": "synthetic (seeded Poisson stream of water clusters and alkanes)",
        "config": {"workload": f"{len(reqs)} RHF requests, Poisson arrivals at {a.rate:g}/s (seed 7), STO-3G / "
                               "6-31G / def2-SVP, 1-12 waters or C1-C12 alkanes; server: 64 slots, 32 GB budget, "
                               "12 host workers",
                   "iterations": "device-resident (grouped device steps, density by purification)"
                   if not a.host_steps
                   else "host steps on the workers",
                   "n_ao_min": int(min(n_ao)), "n_ao_max": int(max(n_ao)), "parallelism": "single GPU"},
        "mixed": runs["mixed"], "fp64": runs["fp64"],
        "p50_latency_s": runs["mixed"]["p50_latency_s"], "p99_latency_s": runs["mixed"]["p99_latency_s"],
        "max_abs_dE_mixed_vs_fp64": float(de.max()),
    }
    print(json.dumps(out))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--cpu-sample", type=int, default=2_000_000)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--boys", choices=["accurate", "reference"], default="accurate")
    ap.add_argument("--workload", choices=["water8", "alkane", "batch", "serve", "scf", "gradient"], default="water8",
                    help="water8: the Fock-build benchmark (default); alkane: C20H42/def2-SVP, FP64 vs "
                         "FP32 below Schwarz 1e-5; batch: 512 RHF/6-31G* jobs of 1-6 waters, fused vs one "
                         "job at a time; serve: Poisson stream of SCF requests, mixed precision vs FP64; scf: whole "
                         "SCF solves with the iteration step on the device vs the host; gradient: RHF nuclear "
                         "gradient of the water8 system")
    ap.add_argument("--basis", default="6-31G", help="basis of the gradient workload (6-31G: s/p; def2-SVP or "
                                                     "6-31GP: with d shells)")
    ap.add_argument("--jobs", type=int, default=512, help="jobs of the batch workload")
    ap.add_argument("--requests", type=int, default=200, help="requests of the serve workload")
    ap.add_argument("--rate", type=float, default=200.0, help="mean arrival rate of the serve workload (1/s)")
    ap.add_argument("--host-steps", action="store_true",
                    help="serve workload: run the iterations' linear algebra on host worker threads instead of "
                         "the grouped device steps")
    ap.add_argument("--copy-mb", type=int, default=0,
                    help="only run a device-to-device copy of this many MiB (counter calibration) and exit")
    a = ap.parse_args()
    if a.copy_mb:
        copy_calibration(a.copy_mb)
        return
    if a.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for {a.gpus}); the multi-GPU layout is a design, "
              f"not a measured path.", file=sys.stderr)
        sys.exit(2)

    if a.workload == "alkane":
        bench_alkane(a)
        return
    if a.workload == "batch":
        bench_batch(a)
        return
    if a.workload == "serve":
        bench_serve(a)
        return
    if a.workload == "scf":
        bench_scf(a)
        return
    if a.workload == "gradient":
        bench_gradient(a)
        return

    import dollarstorescf_amd as ds
    from dollarstorescf_amd import flops, geometry

    Z, xyz = geometry.water_lattice(1)
    boys = ds.BOYS_ACCURATE if a.boys == "accurate" else ds.BOYS_REFERENCE
    job = ds.Job(0).set_options(diis_size=8, energy_threshold=1e-8, max_iterations=30, boys_mode=boys,
                                schwarz_threshold=1e-12)
    job.set_basis("6-31G").set_molecule(Z, xyz)
    t0 = time.perf_counter()
    job.construct()
    t_construct = time.perf_counter() - t0
    t0 = time.perf_counter()
    try:
        energy, iters = job.solve()
        converged = True
    except ds.DsscfError as e:  # 20-iteration cap
        if e.code != 6:
            raise
        energy, iters, converged = float(job.history()[-1]), 30, False
    t_scf = time.perf_counter() - t0
    n = job.n_ao
    ctx = job.context
    st = ctx.stats()
    # threads per block of each class from the project's autotuner, timed on
    # this context's own quartet lists (set-up of the repeated builds, not timed)
    t0 = time.perf_counter()
    tuned = ctx.autotune()
    t_tune = time.perf_counter() - t0
    Pa = job.get("Pa")
    ctx.set_density(2.0 * Pa, Pa)

    classes = ds.class_list()
    kernel_ms = np.zeros(len(classes))
    for _ in range(a.warmup):
        ctx.jk_resident()
    t0 = time.perf_counter()
    for _ in range(a.steps):
        ctx.jk_resident()  # each call ends with a stream synchronize
    dt = time.perf_counter() - t0
    ms = dt / a.steps * 1e3
    builds_per_s = a.steps / dt
    # the same steps again with every class kernel (they run concurrently on the
    # context's streams) bracketed by HIP events on its own stream
    t0 = time.perf_counter()
    for _ in range(a.steps):
        ctx.jk_resident(kernel_ms)
    ms_instrumented = (time.perf_counter() - t0) / a.steps * 1e3
    kernel_ms /= a.steps

    # dominant kernel and its roofline: the same builds once more with the class
    # kernels launched one after another on one stream, each bracketed by HIP
    # events (concurrent kernels share the CUs, so their event spans include
    # waiting for one another)
    serial_ms = ctx.time_classes(iters=a.steps)[:len(classes)]
    dom = int(np.argmax(serial_ms))
    ms_dom = float(serial_ms[dom])
    cq = list(st.class_quartets)[:len(classes)]
    cp = list(st.class_prim_quartets)[:len(classes)]
    dom_flops = flops.class_flops(classes[dom], cq[dom], cp[dom])
    total_flops = flops.fock_build_flops(classes, cq, cp)
    achieved = dom_flops / (ms_dom * 1e-3) / 1e12
    name = "".join("spdf"[l] for l in classes[dom])

    def class_rows(ms_of):
        return [{"class": "".join("spdf"[l] for l in c), "quartets": int(q), "prim_quartets": int(pq),
                 "ms": float(m), "tflops": (flops.class_flops(c, q, pq) / (m * 1e-3) / 1e12) if m > 0 else 0.0}
                for c, q, pq, m in zip(classes, cq, cp, ms_of) if q > 0]

    per_class = class_rows(kernel_ms)
    per_class_serial = class_rows(serial_ms)
    cpu = None
    if not a.no_cpu:
        J_gpu, K_gpu = ctx.get_jk(n)
        cpu = cpu_baseline(ctx, Z, xyz, n, Pa, J_gpu, K_gpu, a.cpu_sample)

    out = {
        "metric": METRIC,
        "value": builds_per_s,
        "unit": "Fock builds/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": ms,
        "ms_per_step_event_timed": ms_instrumented,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded water lattice geometry)",
        "config": {"workload": "RHF/6-31G Fock build, 8-water 2x2x2 lattice (3.0 A, +-0.3 A jitter, seed 1), "
                               "104 AOs, Schwarz 1e-12, primitive-pair screening 1e-17 (accurate Boys), FP64",
                   "boys": a.boys, "prim_threshold": ds.DEFAULT_PRIM_THRESHOLD if a.boys == "accurate" else 0.0,
                   "n_ao": n, "shell_quartets": int(st.n_quartets),
                   "shell_quartets_unscreened": int(st.n_quartets_unscreened),
                   "prim_quartets": int(st.n_prim_quartets), "parallelism": "single GPU",
                   "block_sizes": {"".join("spd"[l] for l in c): int(b) for c, b, q in zip(classes, tuned, cq)
                                   if q > 0},
                   "autotune_seconds": t_tune},
        "eri_shell_quartets_per_s": st.n_quartets * builds_per_s,
        "eri_prim_quartets_per_s": st.n_prim_quartets * builds_per_s,
        "fock_build_tflops": total_flops / (ms * 1e-3) / 1e12,
        "roofline": {"bound": "fp64_valu", "achieved": achieved, "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved / FP64_PEAK_TFLOPS, "traffic": traffic_of(f"k_jk<{name}>"),
                     "kernel": f"k_jk<{name}>", "kernel_ms": ms_dom,
                     "note": "FP64 vector-ALU bound; peak is MI355X FP64 78.6 TFLOP/s; achieved = algorithmic "
                             "FLOPs of the class (dollarstorescf_amd/flops.py) / its kernel's average duration "
                             "(HIP events around it over --steps builds whose class kernels run one after "
                             "another on one stream, right after the timed steps); traffic = HBM bytes per "
                             "dispatch from profiles/fock_water8_631g_traffic.json"},
        "classes": per_class,  # event spans with the classes running concurrently (4 streams)
        "classes_serial": per_class_serial,  # one stream, classes in turn (the roofline's timing)
        "cpu_baseline": cpu,
        "scf": {"energy": energy, "iterations": iters, "converged": converged, "seconds": t_scf,
                "construct_seconds": t_construct},
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
