"""dsscf_eri_batch_noalloc (no device allocation, fixed per-GPU scratch)
against dsscf_eri_batch, through the C-ABI: two calls in a row on one
context, the second with a smaller input, both bit-identical to the
allocating entry point (the same kernels on the same shell quartets), and a
request whose staging exceeds the scratch refused before any work."""
import ctypes

import numpy as np
import pytest

import dollarstorescf_amd as ds
from dollarstorescf_amd import geometry as G

pytestmark = pytest.mark.gpu


def _noalloc(ctx, q):
    L = ds.lib()
    f = L.dsscf_eri_batch_noalloc
    f.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_longlong, ctypes.c_void_p]
    f.restype = ctypes.c_int
    q = np.ascontiguousarray(q, np.int32).reshape(-1, 4)
    out = np.full(len(q), np.nan)
    rc = f(ctx._h, q.ctypes.data_as(ctypes.c_void_p), len(q), out.ctypes.data_as(ctypes.c_void_p))
    return rc, out


@pytest.mark.parametrize("basis,boys", [("6-31G", ds.BOYS_ACCURATE), ("def2-SVP", ds.BOYS_REFERENCE)])
def test_eri_batch_noalloc_matches_eri_batch(basis, boys):
    Z, xyz = G.water_box(2, seed=5)
    job = ds.Job(0).set_options(boys_mode=boys, schwarz_threshold=0.0)
    job.set_basis(basis).set_molecule(Z, xyz).construct()
    ctx, n = job.context, job.n_ao
    rng = np.random.default_rng(11)
    for size in (20000, 3000):  # the second call smaller: stale scratch contents must not leak in
        q = rng.integers(0, n, size=(size, 4)).astype(np.int32)
        print(f"noalloc: {basis} n_ao={n} requests={size}", flush=True)
        rc, got = _noalloc(ctx, q)
        assert rc == 0, ds.lib().dsscf_last_error().decode()
        want = ctx.eri(q)
        assert np.array_equal(got, want)


def test_eri_batch_noalloc_refuses_oversized_request():
    """A class whose staging would pass the 16 MiB scratch is refused with
    DSSCF_ERR_ARG and a message, and the output is left untouched."""
    Z, xyz = G.alkane(14)
    job = ds.Job(0).set_options(schwarz_threshold=0.0)
    job.set_basis("def2-SVP").set_molecule(Z, xyz).construct()
    n = job.n_ao
    # 14 d shells: 105 d-shell pairs, 5565 unique (dd|dd) shell quartets of 625
    # integrals (27.8 MB of staging); random requests over the d AOs touch most
    d_aos = [i for i in range(n) if job.ao(i)["l"] == 2]
    rng = np.random.default_rng(3)
    q = rng.choice(d_aos, size=(200000, 4)).astype(np.int32)
    rc, got = _noalloc(job.context, q)
    assert rc == 1
    assert "scratch" in ds.lib().dsscf_last_error().decode()
    assert np.all(np.isnan(got))
