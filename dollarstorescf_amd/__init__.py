"""dollarstorescf_amd -- MI355X-native direct-SCF Fock-build engine.

Python view of the in-tree C-ABI library ``libdsscf.so`` (include/dsscf.h and
include/dsscf_scf.h).  Everything numeric runs in the library: HIP kernels
for ERIs/Boys/screening/J-K on the GPU, the host C++ SCF layer for the rest.
There is no fallback: if the library or a GPU is missing, calls raise.
"""
from __future__ import annotations

import ctypes
import time
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libdsscf.so")
BASIS_DIR = os.path.join(HERE, "basis")

BOYS_REFERENCE = 0
BOYS_ACCURATE = 1
DEFAULT_PRIM_THRESHOLD = 1e-17

ERRORS = {1: "bad argument", 2: "HIP error", 3: "call out of order", 4: "unsupported",
          5: "device allocation failed", 6: "did not converge"}


class DsscfError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"[{ERRORS.get(code, code)}] {msg}")
        self.code = code


class Options(ctypes.Structure):
    _fields_ = [("boys_mode", ctypes.c_int), ("schwarz_threshold", ctypes.c_double),
                ("fp32_threshold", ctypes.c_double), ("prim_threshold", ctypes.c_double),
                ("work_item_flops", ctypes.c_double), ("group_prims", ctypes.c_double),
                ("jcd_max_bytes", ctypes.c_double), ("ket_order", ctypes.c_int), ("cartesian", ctypes.c_int)]


class Stats(ctypes.Structure):
    _fields_ = [("n_pairs", ctypes.c_longlong), ("n_prim_pairs", ctypes.c_longlong),
                ("n_quartets", ctypes.c_longlong), ("n_prim_quartets", ctypes.c_longlong),
                ("n_quartets_unscreened", ctypes.c_longlong), ("n_class", ctypes.c_int),
                ("class_quartets", ctypes.c_longlong * 21), ("class_prim_quartets", ctypes.c_longlong * 21),
                ("n_quartets_fp32", ctypes.c_longlong), ("class_quartets_fp32", ctypes.c_longlong * 21),
                ("class_prim_quartets_fp32", ctypes.c_longlong * 21), ("n_work_pairs", ctypes.c_longlong),
                ("n_work_items", ctypes.c_longlong), ("class_lanes", ctypes.c_longlong * 21),
                ("class_lane_slots", ctypes.c_longlong * 21), ("class_lane_trips", ctypes.c_longlong * 21),
                ("class_wave_trips", ctypes.c_longlong * 21)]


_LIB = None


def lib():
    """Load libdsscf.so (raises if it was not built)."""
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} not built: run `python -m dollarstorescf_amd.build`")
        L = ctypes.CDLL(LIB_PATH)
        i, d, p, c = ctypes.c_int, ctypes.c_double, ctypes.c_void_p, ctypes.c_char_p
        L.dsscf_last_error.restype = c
        L.dsscf_job_error.restype = c
        L.dsscf_job_error.argtypes = [p]
        L.dsscf_job_context.restype = p
        L.dsscf_job_context.argtypes = [p]
        sigs = {
            "dsscf_ctx_create": [p, i], "dsscf_ctx_destroy": [p], "dsscf_set_options": [p, p],
            "dsscf_set_basis": [p, i, p, p, p, p, p, p], "dsscf_prepare": [p],
            "dsscf_jk": [p, p, p, p, p], "dsscf_jk_device": [p, p, p, p, p],
            "dsscf_eri_batch": [p, p, ctypes.c_longlong, p], "dsscf_schwarz": [p, p],
            "dsscf_get_stats": [p, p], "dsscf_set_density": [p, p, p], "dsscf_jk_resident": [p],
            "dsscf_jk_resident_timed": [p, p],
            "dsscf_get_jk": [p, p, p], "dsscf_time_jk_classes": [p, i, p], "dsscf_time_jk": [p, i, i, p, p, p], "dsscf_class_l": [i, p],
            "dsscf_job_create": [p, i], "dsscf_job_destroy": [p],
            "dsscf_job_set_options": [p, i, d, i, i, d, i], "dsscf_job_set_precision": [p, d, d],
            "dsscf_job_set_step": [p, i], "dsscf_job_set_cartesian": [p, i], "dsscf_eri_batch_noalloc": [p, p, ctypes.c_longlong, p], "dsscf_job_set_orbitals": [p, p, p], "dsscf_job_write_molden": [p, c, i],
            "dsscf_job_gradient": [p, p, p], "dsscf_eri_gradient": [p, p, i, p, p], "dsscf_job_n_atom": [p],
            "dsscf_job_gradient_terms": [p, p, p, p, p],
            "dsscf_job_set_molecule": [p, i, p, p, i, i], "dsscf_job_load_xyz": [p, c, i, i],
            "dsscf_job_set_basis_file": [p, c], "dsscf_job_construct": [p], "dsscf_job_construct_host": [p],
            "dsscf_job_ao": [p, i, p, p, p, p, p, p, p, p, p, i], "dsscf_job_n_ao": [p],
            "dsscf_job_get": [p, c, p], "dsscf_job_set_density": [p, p], "dsscf_job_fock_build": [p],
            "dsscf_job_solve": [p, p, p], "dsscf_job_energies": [p, p], "dsscf_job_history": [p, p, i],
            "dsscf_syev": [i, p, p, p], "dsscf_jobs_construct": [p, i], "dsscf_jobs_solve": [p, i, i, i, p, p, p],
            "dsscf_server_create": [p, i, d, i, i], "dsscf_server_destroy": [p], "dsscf_server_submit": [p, p, ctypes.c_longlong],
            "dsscf_server_step": [p, p, p], "dsscf_server_set_device_steps": [p, i], "dsscf_server_poll": [p, i, p, p, p, p, p, p, p, p],
            "dsscf_server_idle": [p], "dsscf_server_stats": [p, p], "dsscf_one_electron": [p, i, p, p, p, p, p],
            "dsscf_boys_eval": [p, i, i, i, p, p],
        }
        for name, args in sigs.items():
            f = getattr(L, name)
            f.argtypes = args
            f.restype = i
        for name, args in {"dsscf_batch_create": [p, i], "dsscf_batch_destroy": [p], "dsscf_batch_set": [p, p, i],
                           "dsscf_batch_set_graph": [p, i], "dsscf_batch_jk": [p, p, p, p, p],
                           "dsscf_batch_set_device": [p, p, i, p, p, p, p], "dsscf_batch_jk_device": [p],
                           "dsscf_batch_time": [p, i, i, p], "dsscf_batch_autotune": [p, p],
                           "dsscf_batch_autotune_schedule": [p, p, p, p, p], "dsscf_set_class_schedule": [p, p, p, p],
                           "dsscf_job_set_schedule": [p, p, p, p], "dsscf_job_set_guess": [p, p, ctypes.c_longlong],
                           "dsscf_jk_uhf": [p, p, p, p, p, p, p],
                           "dsscf_one_electron_gradient": [p, i, p, p, p, p, p, p]}.items():
            f = getattr(L, name)
            f.argtypes = args
            f.restype = i
        L.dsscf_server_clock.restype = d
        L.dsscf_server_clock.argtypes = [p]
        L.dsscf_ctx_device_bytes.restype = ctypes.c_longlong
        L.dsscf_ctx_device_bytes.argtypes = [p]
        for name in ("dsscf_max_l", "dsscf_class_count", "dsscf_xform_size"):
            getattr(L, name).restype = i
            getattr(L, name).argtypes = []
        _LIB = L
    return _LIB


def _p(a: np.ndarray):
    return a.ctypes.data_as(ctypes.c_void_p)


def _check(rc, msg_fn=None):
    if rc != 0:
        msg = msg_fn() if msg_fn else lib().dsscf_last_error().decode()
        raise DsscfError(rc, msg)


def basis_path(name: str) -> str:
    return os.path.join(BASIS_DIR, name)


def construct_jobs(jobs):
    """Job.construct for several jobs, the host parts in parallel."""
    arr = (ctypes.c_void_p * len(jobs))(*[j._h.value for j in jobs])
    rc = lib().dsscf_jobs_construct(arr, len(jobs))
    if rc:
        bad = [j._err() for j in jobs if j._err()]
        raise DsscfError(rc, bad[0] if bad else "dsscf_jobs_construct failed")
    return jobs


def solve_jobs(jobs, use_graph=True, autotune=True):
    """Job.solve for several constructed jobs in lock step, their Fock builds
    fused into one launch per quartet class.  Returns (energies, iterations,
    status) arrays; status 0 = converged, else the job's error code."""
    n = len(jobs)
    arr = (ctypes.c_void_p * n)(*[j._h.value for j in jobs])
    e, it, st = np.zeros(n), np.zeros(n, np.int32), np.zeros(n, np.int32)
    rc = lib().dsscf_jobs_solve(arr, n, 1 if use_graph else 0, 1 if autotune else 0, _p(e), _p(it), _p(st))
    if rc and not np.any(st):  # refused before any job ran (a bad or unsupported job)
        bad = [j._err() for j in jobs if j._err()]
        raise DsscfError(rc, bad[0] if bad else "dsscf_jobs_solve failed")
    return e, it, st


class Server:
    """Continuous-batching SCF service on one device (dsscf_server_*).  Submit
    unconstructed jobs; step() runs one round; poll() returns finished ones as
    dicts (tag, job, status, energy, iterations, t_submit, t_admit, t_done on
    the server clock)."""

    def __init__(self, device=0, budget_bytes=8e9, max_active=64, workers=4, device_steps=False):
        self._h = ctypes.c_void_p()
        _check(lib().dsscf_server_create(ctypes.byref(self._h), device, float(budget_bytes), max_active, workers))
        self._jobs = {}
        if device_steps:
            _check(lib().dsscf_server_set_device_steps(self._h, 1))

    def close(self):
        if getattr(self, "_h", None):
            lib().dsscf_server_destroy(self._h)
        self._h = None

    __del__ = close

    def submit(self, job, tag: int):
        """Queue ``job`` under ``tag``.  A tag still in flight is refused (the
        server keeps the job alive through it), and so is a job the server
        already holds (DSSCF_ERR_STATE from the library)."""
        if tag in self._jobs:
            raise DsscfError(1, f"Server.submit: tag {tag} is still in flight")
        self._jobs[tag] = job
        rc = lib().dsscf_server_submit(self._h, job._h, tag)
        if rc:
            del self._jobs[tag]
            raise DsscfError(rc, job._err())

    def step(self):
        a, w = ctypes.c_int(), ctypes.c_int()
        _check(lib().dsscf_server_step(self._h, ctypes.byref(a), ctypes.byref(w)))
        return a.value, w.value

    def poll(self, cap=1024):
        tags = np.zeros(cap, np.int64)
        st, it = np.zeros(cap, np.int32), np.zeros(cap, np.int32)
        e, ts, ta, td = np.zeros(cap), np.zeros(cap), np.zeros(cap), np.zeros(cap)
        jobs = (ctypes.c_void_p * cap)()
        n = lib().dsscf_server_poll(self._h, cap, _p(tags), jobs, _p(st), _p(e), _p(it), _p(ts), _p(ta), _p(td))
        out = []
        for k in range(n):
            tag = int(tags[k])
            out.append({"tag": tag, "job": self._jobs.pop(tag), "status": int(st[k]), "energy": float(e[k]),
                        "iterations": int(it[k]), "t_submit": float(ts[k]), "t_admit": float(ta[k]),
                        "t_done": float(td[k])})
        return out

    def clock(self) -> float:
        return lib().dsscf_server_clock(self._h)

    def idle(self) -> bool:
        return lib().dsscf_server_idle(self._h) == 1

    def stats(self) -> dict:
        out = np.zeros(5)
        _check(lib().dsscf_server_stats(self._h, _p(out)))
        return dict(zip(["rounds", "set_seconds", "fock_seconds", "host_seconds", "jobs_per_round"],
                        [out[0], out[1], out[2], out[3], out[4] / max(out[0], 1.0)]))

    def run_until_idle(self):
        """Step until every submitted job is back; returns them."""
        out = []
        while not self.idle():
            a, w = self.step()
            out += self.poll()
            if a == 0 and w > 0:
                time.sleep(0.0005)  # jobs still constructing on the workers (step waits itself otherwise)
        return out + self.poll()


def syev(A):
    """Eigenvalues (ascending) and eigenvectors (columns) of a symmetric matrix
    with the SCF layer's host eigensolver."""
    A = np.ascontiguousarray(A, np.float64)
    n = A.shape[0]
    w, V = np.empty(n), np.empty((n, n))
    _check(lib().dsscf_syev(n, _p(A), _p(w), _p(V)))
    return w, V


def class_list():
    L = lib()
    out = []
    for c in range(L.dsscf_class_count()):
        l4 = np.zeros(4, np.int32)
        _check(L.dsscf_class_l(c, _p(l4)))
        out.append(tuple(int(v) for v in l4))
    return out


class Context:
    """One dsscf_ctx (device buffers + stream) over explicit shell data."""

    def __init__(self, device: int = 0):
        self._h = ctypes.c_void_p()
        _check(lib().dsscf_ctx_create(ctypes.byref(self._h), device))
        self._own = True
        self._device = device

    @classmethod
    def borrow(cls, handle, device: int = 0):
        obj = cls.__new__(cls)
        obj._h = ctypes.c_void_p(handle)
        obj._own = False
        obj._device = device
        return obj

    def close(self):
        if getattr(self, "_own", False) and self._h:
            lib().dsscf_ctx_destroy(self._h)
        self._h = None

    __del__ = close

    def set_options(self, boys_mode=BOYS_ACCURATE, schwarz_threshold=1e-12, fp32_threshold=0.0, prim_threshold=0.0,
                    work_item_flops=0.0, group_prims=0.0, jcd_max_bytes=0.0, ket_order=0, cartesian=0):
        o = Options(boys_mode, schwarz_threshold, fp32_threshold, prim_threshold, work_item_flops, group_prims,
                    jcd_max_bytes, ket_order, int(cartesian))
        _check(lib().dsscf_set_options(self._h, ctypes.byref(o)))

    def set_basis(self, l, nprim, centers, exps, coefs, xform):
        arrs = [np.ascontiguousarray(l, np.int32), np.ascontiguousarray(nprim, np.int32),
                np.ascontiguousarray(centers, np.float64).reshape(-1), np.ascontiguousarray(exps, np.float64),
                np.ascontiguousarray(coefs, np.float64), np.ascontiguousarray(xform, np.float64)]
        _check(lib().dsscf_set_basis(self._h, len(arrs[0]), *[_p(a) for a in arrs]))

    def prepare(self):
        _check(lib().dsscf_prepare(self._h))

    def jk(self, Ptot_packed, Pa_packed):
        Pt = np.ascontiguousarray(Ptot_packed, np.float64)
        Pa = np.ascontiguousarray(Pa_packed, np.float64)
        J, K = np.empty_like(Pt), np.empty_like(Pt)
        _check(lib().dsscf_jk(self._h, _p(Pt), _p(Pa), _p(J), _p(K)))
        return J, K

    def jk_uhf(self, Ptot_packed, Pa_packed, Pb_packed):
        """Unrestricted Fock build (dsscf_jk_uhf): J[Ptot], K[Pa], K[Pb] from
        one pass over the integrals."""
        Pt = np.ascontiguousarray(Ptot_packed, np.float64)
        Pa = np.ascontiguousarray(Pa_packed, np.float64)
        Pb = np.ascontiguousarray(Pb_packed, np.float64)
        J, Ka, Kb = np.empty_like(Pt), np.empty_like(Pt), np.empty_like(Pt)
        _check(lib().dsscf_jk_uhf(self._h, _p(Pt), _p(Pa), _p(Pb), _p(J), _p(Ka), _p(Kb)))
        return J, Ka, Kb

    def set_density(self, Ptot_packed, Pa_packed):
        self._Pt = np.ascontiguousarray(Ptot_packed, np.float64)
        self._Pa = np.ascontiguousarray(Pa_packed, np.float64)
        _check(lib().dsscf_set_density(self._h, _p(self._Pt), _p(self._Pa)))

    def jk_resident(self, kernel_ms=None):
        """One Fock build on the resident density; with kernel_ms (float64 array of
        class_count() entries) each class kernel's event-timed duration is added to it."""
        if kernel_ms is None:
            _check(lib().dsscf_jk_resident(self._h))
        else:
            assert kernel_ms.dtype == np.float64 and kernel_ms.size >= lib().dsscf_class_count()
            _check(lib().dsscf_jk_resident_timed(self._h, _p(kernel_ms)))

    def get_jk(self, n_ao):
        J, K = np.empty(n_ao * (n_ao + 1) // 2), np.empty(n_ao * (n_ao + 1) // 2)
        _check(lib().dsscf_get_jk(self._h, _p(J), _p(K)))
        return J, K

    def eri(self, quartets):
        q = np.ascontiguousarray(quartets, np.int32).reshape(-1, 4)
        out = np.empty(len(q))
        _check(lib().dsscf_eri_batch(self._h, _p(q), len(q), _p(out)))
        return out

    def one_electron(self, Z, xyz_bohr, n_ao: int):
        """S, T, V (packed-U) of this context's n_ao AOs and the given nuclei, on the device."""
        Z = np.ascontiguousarray(Z, np.int32)
        xyz = np.ascontiguousarray(xyz_bohr, np.float64).reshape(-1)
        npk = n_ao * (n_ao + 1) // 2
        S, T, V = np.empty(npk), np.empty(npk), np.empty(npk)
        _check(lib().dsscf_one_electron(self._h, len(Z), _p(Z), _p(xyz), _p(S), _p(T), _p(V)))
        return S, T, V

    def one_electron_gradient(self, Z, xyz_bohr, ao_atom, Pt_packed, W_packed):
        """Kinetic, overlap and nuclear-attraction gradient terms on the device
        (dsscf_one_electron_gradient): (natom, 3)."""
        Z = np.ascontiguousarray(Z, np.int32)
        xyz = np.ascontiguousarray(xyz_bohr, np.float64).reshape(-1)
        at = np.ascontiguousarray(ao_atom, np.int32)
        Pt = np.ascontiguousarray(Pt_packed, np.float64)
        W = np.ascontiguousarray(W_packed, np.float64)
        g = np.empty(3 * len(Z))
        _check(lib().dsscf_one_electron_gradient(self._h, len(Z), _p(Z), _p(xyz), _p(at), _p(Pt), _p(W), _p(g)))
        return g.reshape(-1, 3)

    def eri_gradient(self, Pa_packed, ao_atom, natom: int):
        """Two-electron part of the RHF gradient at alpha density Pa (packed-U):
        (natom, 3), Eh/bohr; ao_atom gives each AO's atom."""
        Pa = np.ascontiguousarray(Pa_packed, np.float64)
        at = np.ascontiguousarray(ao_atom, np.int32)
        g = np.empty(3 * natom)
        _check(lib().dsscf_eri_gradient(self._h, _p(Pa), natom, _p(at), _p(g)))
        return g.reshape(natom, 3)

    def boys(self, T, L: int, mode: int = 2):
        """F_0..F_L at the arguments T on the device (mode 0 reference series, 1 accurate
        Taylor table, 2 accurate interval table in LDS); shape (len(T), L+1)."""
        T = np.ascontiguousarray(T, np.float64)
        out = np.empty((len(T), L + 1))
        _check(lib().dsscf_boys_eval(self._h, mode, L, len(T), _p(T), _p(out)))
        return out

    def stats(self) -> Stats:
        st = Stats()
        _check(lib().dsscf_get_stats(self._h, ctypes.byref(st)))
        return st

    def schwarz(self, n_pairs):
        Q = np.empty(n_pairs)
        _check(lib().dsscf_schwarz(self._h, _p(Q)))
        return Q

    def autotune(self):
        """Threads per block of each thread-per-quartet class, chosen by timing
        64/128/256 on this context's own quartet lists (dsscf_batch_autotune
        over a batch of this context alone, which keeps the choice in the
        context).  Pays off for many builds of one geometry."""
        b = Batch(self._device)
        try:
            b.set([self], graph=False)
            return b.autotune()
        finally:
            b.close()

    def time_classes(self, iters=10):
        out = np.zeros(lib().dsscf_class_count())
        _check(lib().dsscf_time_jk_classes(self._h, iters, _p(out)))
        return out

    def time_jk(self, warmup=2, iters=10):
        ms, dom = ctypes.c_double(), ctypes.c_double()
        cls = ctypes.c_int()
        _check(lib().dsscf_time_jk(self._h, warmup, iters, ctypes.byref(ms), ctypes.byref(dom), ctypes.byref(cls)))
        return ms.value, dom.value, cls.value


class Batch:
    """Fused Fock builds over several prepared contexts (dsscf_batch_*)."""

    def __init__(self, device: int = 0):
        self._h = ctypes.c_void_p()
        _check(lib().dsscf_batch_create(ctypes.byref(self._h), device))
        self._members = []

    def close(self):
        if getattr(self, "_h", None):
            lib().dsscf_batch_destroy(self._h)
        self._h = None

    __del__ = close

    def set(self, contexts, graph=True):
        self._members = list(contexts)
        arr = (ctypes.c_void_p * len(contexts))(*[c._h.value for c in contexts])
        _check(lib().dsscf_batch_set(self._h, arr, len(contexts)))
        _check(lib().dsscf_batch_set_graph(self._h, 1 if graph else 0))

    def jk(self, Ptots, Pas):
        n = len(self._members)
        Pt = [np.ascontiguousarray(x, np.float64) for x in Ptots]
        Pa = [np.ascontiguousarray(x, np.float64) for x in Pas]
        J = [np.empty_like(x) for x in Pt]
        K = [np.empty_like(x) for x in Pt]
        mk = lambda xs: (ctypes.c_void_p * n)(*[x.ctypes.data for x in xs])
        _check(lib().dsscf_batch_jk(self._h, mk(Pt), mk(Pa), mk(J), mk(K)))
        return J, K

    def set_device(self, contexts, d_Ptot, d_Pa, d_J, d_K, graph=True):
        """Members whose packed-U densities and J/K live in caller-owned device
        buffers, given as device addresses (one per member)."""
        self._members = list(contexts)
        n = len(contexts)
        arr = (ctypes.c_void_p * n)(*[c._h.value for c in contexts])
        mk = lambda xs: (ctypes.c_void_p * n)(*[int(x) for x in xs])
        _check(lib().dsscf_batch_set_device(self._h, arr, n, mk(d_Ptot), mk(d_Pa), mk(d_J), mk(d_K)))
        _check(lib().dsscf_batch_set_graph(self._h, 1 if graph else 0))

    def jk_device(self):
        _check(lib().dsscf_batch_jk_device(self._h))

    def time(self, warmup=2, iters=10):
        ms = ctypes.c_double()
        _check(lib().dsscf_batch_time(self._h, warmup, iters, ctypes.byref(ms)))
        return ms.value

    def autotune(self):
        out = np.zeros(lib().dsscf_class_count(), np.int32)
        _check(lib().dsscf_batch_autotune(self._h, _p(out)))
        return out

    def autotune_schedule(self):
        """The schedule autotuner (dsscf_batch_autotune_schedule): per class the
        sub-pair FLOP size, lane-group target, kets per work unit and threads
        per block that run the members' own lists fastest (0 = default).  The
        members end prepared with it; returns a dict of per-class arrays."""
        n = lib().dsscf_class_count()
        wif, gp = np.zeros(n), np.zeros(n)
        ch, bs = np.zeros(n, np.int32), np.zeros(n, np.int32)
        _check(lib().dsscf_batch_autotune_schedule(self._h, _p(wif), _p(gp), _p(ch), _p(bs)))
        return {"work_item_flops": wif, "group_prims": gp, "kets_per_unit": ch, "block_size": bs}


class Job:
    """One SCF job: molecule + basis -> AOs, 1e terms, device Fock builds, RHF."""

    def __init__(self, device: int = 0):
        self._h = ctypes.c_void_p()
        _check(lib().dsscf_job_create(ctypes.byref(self._h), device))
        self._device = device

    def _err(self):
        return lib().dsscf_job_error(self._h).decode()

    def close(self):
        if getattr(self, "_h", None):
            lib().dsscf_job_destroy(self._h)
        self._h = None

    __del__ = close

    def set_options(self, diis_size=8, energy_threshold=1e-8, max_iterations=500, boys_mode=BOYS_ACCURATE,
                    schwarz_threshold=1e-12, verbose=0, fp32_threshold=0.0, prim_threshold=None, device_step=False,
                    cartesian=False):
        """prim_threshold defaults to 1e-17 in the accurate Boys mode and to 0 (every
        primitive quartet, as the reference) in the reference mode.  device_step runs
        each iteration's DIIS, X F X and density on the device (True or 1: the
        density by purification; 2: by rocSOLVER eigenvectors every iteration).
        cartesian: d shells as six Cartesian functions (6D) instead of the
        reference's five real solid harmonics."""
        if prim_threshold is None:
            prim_threshold = DEFAULT_PRIM_THRESHOLD if boys_mode == BOYS_ACCURATE else 0.0
        _check(lib().dsscf_job_set_options(self._h, diis_size, energy_threshold, max_iterations, boys_mode,
                                           schwarz_threshold, verbose), self._err)
        _check(lib().dsscf_job_set_precision(self._h, fp32_threshold, prim_threshold), self._err)
        _check(lib().dsscf_job_set_step(self._h, int(device_step)), self._err)
        _check(lib().dsscf_job_set_cartesian(self._h, 1 if cartesian else 0), self._err)
        return self

    def set_molecule(self, Z, xyz_bohr, charge=0, multiplicity=1):
        Z = np.ascontiguousarray(Z, np.int32)
        xyz = np.ascontiguousarray(xyz_bohr, np.float64).reshape(-1)
        _check(lib().dsscf_job_set_molecule(self._h, len(Z), _p(Z), _p(xyz), charge, multiplicity), self._err)
        return self

    def load_xyz(self, path, charge=0, multiplicity=1):
        _check(lib().dsscf_job_load_xyz(self._h, path.encode(), charge, multiplicity), self._err)
        return self

    def set_basis(self, name_or_path: str):
        path = name_or_path if os.path.sep in name_or_path else basis_path(name_or_path)
        _check(lib().dsscf_job_set_basis_file(self._h, path.encode()), self._err)
        return self

    def construct(self):
        _check(lib().dsscf_job_construct(self._h), self._err)
        return self

    def construct_host(self):
        """AO list and one-electron terms only (no GPU needed)."""
        _check(lib().dsscf_job_construct_host(self._h), self._err)
        return self

    def ao(self, i: int) -> dict:
        cap = 128
        l, m, npr, nt = (np.zeros(1, np.int32) for _ in range(4))
        c = np.zeros(3)
        z, w, tw = np.zeros(cap), np.zeros(cap), np.zeros(cap)
        tx = np.zeros(3 * cap, np.int32)
        _check(lib().dsscf_job_ao(self._h, i, _p(l), _p(m), _p(c), _p(npr), _p(z), _p(w), _p(nt), _p(tw), _p(tx), cap),
               self._err)
        return dict(l=int(l[0]), m=int(m[0]), center=c, zeta=z[:npr[0]].copy(), weight=w[:npr[0]].copy(),
                    terms=[(float(tw[t]), *map(int, tx[3 * t:3 * t + 3])) for t in range(nt[0])])

    @property
    def n_ao(self) -> int:
        return lib().dsscf_job_n_ao(self._h)

    def get(self, name: str) -> np.ndarray:
        n = self.n_ao
        size = {"eps": n, "epsb": n, "C": n * n, "Cb": n * n, "device_steps": 3}.get(name, n * (n + 1) // 2)
        out = np.empty(size)
        _check(lib().dsscf_job_get(self._h, name.encode(), _p(out)), self._err)
        return out.reshape(n, n) if name in ("C", "Cb") else out

    def set_orbitals(self, C, eps):
        """Replace the orbitals (C: n x n, column i = orbital i; eps: n energies)."""
        C = np.ascontiguousarray(C, np.float64)
        eps = np.ascontiguousarray(eps, np.float64)
        n = self.n_ao
        if C.shape != (n, n) or eps.shape != (n,):
            raise ValueError("set_orbitals: C must be n x n and eps n long")
        _check(lib().dsscf_job_set_orbitals(self._h, _p(C), _p(eps)), self._err)

    def gradient(self, parts: bool = False):
        """dE/dR of the solved job, (natom, 3) in Eh/bohr; with parts=True also the
        one-electron, two-electron and nuclear-repulsion parts."""
        na = lib().dsscf_job_n_atom(self._h)
        g = np.empty(3 * na)
        pt = np.empty(9 * na)
        _check(lib().dsscf_job_gradient(self._h, _p(g), _p(pt)), self._err)
        g = g.reshape(na, 3)
        if parts:
            return g, pt.reshape(3, na, 3)
        return g

    def gradient_terms(self, Pt_packed, W_packed):
        """Host terms of the gradient for given densities: (one-electron, nuclear
        repulsion), each (natom, 3)."""
        na = lib().dsscf_job_n_atom(self._h)
        Pt = np.ascontiguousarray(Pt_packed, np.float64)
        W = np.ascontiguousarray(W_packed, np.float64)
        g1, gn = np.empty(3 * na), np.empty(3 * na)
        _check(lib().dsscf_job_gradient_terms(self._h, _p(Pt), _p(W), _p(g1), _p(gn)), self._err)
        return g1.reshape(na, 3), gn.reshape(na, 3)

    def set_schedule(self, schedule):
        """Per-class list schedule for this job's context (a dict as
        Batch.autotune_schedule returns; None = defaults).  Before construct."""
        if schedule is None:
            _check(lib().dsscf_job_set_schedule(self._h, None, None, None), self._err)
            return self
        wif = np.ascontiguousarray(schedule["work_item_flops"], np.float64)
        gp = np.ascontiguousarray(schedule["group_prims"], np.float64)
        ch = np.ascontiguousarray(schedule["kets_per_unit"], np.int32)
        self._sched = (wif, gp, ch)
        _check(lib().dsscf_job_set_schedule(self._h, _p(wif), _p(gp), _p(ch)), self._err)
        return self

    def set_guess(self, Pa_packed):
        """Warm start: the packed alpha density the SCF starts from (None = the
        core guess), e.g. a neighbouring geometry's ``get("Pa")``."""
        if Pa_packed is None:
            _check(lib().dsscf_job_set_guess(self._h, None, 0), self._err)
            return self
        self._guess = np.ascontiguousarray(Pa_packed, np.float64)
        _check(lib().dsscf_job_set_guess(self._h, _p(self._guess), len(self._guess)), self._err)
        return self

    def write_molden(self, path: str, write_virtual: bool = True):
        """Molden file of the current orbitals (scf::writeOrbitalsMolden)."""
        _check(lib().dsscf_job_write_molden(self._h, path.encode(), 1 if write_virtual else 0), self._err)

    def set_density(self, Pa_packed):
        Pa = np.ascontiguousarray(Pa_packed, np.float64)
        _check(lib().dsscf_job_set_density(self._h, _p(Pa)), self._err)

    def fock_build(self):
        _check(lib().dsscf_job_fock_build(self._h), self._err)
        return self.get("J"), self.get("K")

    def solve(self):
        e = ctypes.c_double()
        it = ctypes.c_int()
        _check(lib().dsscf_job_solve(self._h, ctypes.byref(e), ctypes.byref(it)), self._err)
        return e.value, it.value

    def energies(self) -> dict:
        out = np.empty(5)
        _check(lib().dsscf_job_energies(self._h, _p(out)), self._err)
        return dict(zip(["kinetic", "nuclear_attraction", "coulomb", "exchange", "nuclear"], out.tolist()))

    def history(self):
        n = lib().dsscf_job_history(self._h, None, 0)
        out = np.empty(max(n, 1))
        lib().dsscf_job_history(self._h, _p(out), n)
        return out[:n]

    @property
    def context(self) -> Context:
        return Context.borrow(lib().dsscf_job_context(self._h), self._device)


def single_point(Z, xyz_bohr, basis: str, charge=0, multiplicity=1, device=0, **opts):
    """Convenience: build and solve one RHF job; returns (energy, iterations, job)."""
    job = Job(device).set_options(**opts).set_basis(basis).set_molecule(Z, xyz_bohr, charge, multiplicity).construct()
    e, it = job.solve()
    return e, it, job
