"""omc_lookup_batch_noalloc (the allocation-free lookup batch) against
omc_lookup_batch on the same inputs: grid indices and cross sections bit for
bit and the same probe totals, on two calls in a row, the second smaller, for
the plain and the learned search.

The caller's device buffers and stream come from the HIP runtime that
libopenmc.so itself is linked against (libamdhip64.so.7, already loaded by
the library), through ctypes, so the test holds one HIP runtime in the
process whatever ran before it."""
from ctypes import CDLL, POINTER, byref, c_int, c_size_t, c_uint64, c_void_p

import numpy as np
import pytest

from helpers import chdir, prepare

pytestmark = pytest.mark.gpu

H2D, D2H = 1, 2  # hipMemcpyHostToDevice, hipMemcpyDeviceToHost


def _hip():
    from openmc_amd import lib  # noqa: F401  (loads libopenmc.so and its HIP runtime)
    h = CDLL("libamdhip64.so.7")
    h.hipMalloc.argtypes = [POINTER(c_void_p), c_size_t]
    h.hipFree.argtypes = [c_void_p]
    h.hipMemcpy.argtypes = [c_void_p, c_void_p, c_size_t, c_int]
    h.hipStreamCreate.argtypes = [POINTER(c_void_p)]
    h.hipStreamDestroy.argtypes = [c_void_p]
    h.hipStreamSynchronize.argtypes = [c_void_p]
    for f in ("hipMalloc", "hipFree", "hipMemcpy", "hipStreamCreate", "hipStreamDestroy", "hipStreamSynchronize"):
        getattr(h, f).restype = c_int
    return h


def _check(rc, what):
    assert rc == 0, "%s returned %d" % (what, rc)


def _noalloc(lib, hip, i_nuc, E, learned, stream):
    """One call with caller-owned device buffers (allocated and freed here,
    outside the library) on the given stream (None: the library's)."""
    n = E.size
    bufs = []
    try:
        ptrs = []
        for size in (max(8 * n, 8), max(4 * n, 4), max(16 * n, 16)):
            p = c_void_p()
            _check(hip.hipMalloc(byref(p), size), "hipMalloc")
            bufs.append(p)
            ptrs.append(p)
        e_d, ig_d, xs_d = ptrs
        if n:
            _check(hip.hipMemcpy(e_d, E.ctypes.data, 8 * n, H2D), "hipMemcpy")
        pr = c_uint64(0)
        # (a non-zero return raises, through the binding's error check)
        lib._dll.omc_lookup_batch_noalloc(i_nuc, e_d.value, n, 1 if learned else 0, ig_d.value, xs_d.value,
                                          pr, stream.value if stream is not None else None)
        if stream is not None:
            _check(hip.hipStreamSynchronize(stream), "hipStreamSynchronize")
        ig = np.zeros(n, dtype=np.int32)
        xs = np.zeros((n, 2), dtype=np.float64)
        if n:
            _check(hip.hipMemcpy(ig.ctypes.data, ig_d, 4 * n, D2H), "hipMemcpy")
            _check(hip.hipMemcpy(xs.ctypes.data, xs_d, 16 * n, D2H), "hipMemcpy")
        return ig, xs, pr.value
    finally:
        for p in bufs:
            hip.hipFree(p)


def test_noalloc_lookup_matches_lookup_batch():
    from openmc_amd import lib
    hip = _hip()
    print("case: no-alloc lookup batch, cfg0 nuclides, n 50000 then 7000", flush=True)
    d, model, nucs = prepare(0, n_particles=100, n_batches=1)
    rng = np.random.default_rng(21)
    stream = c_void_p()
    _check(hip.hipStreamCreate(byref(stream)), "hipStreamCreate")
    with chdir(d):
        lib.init(output=False)
        try:
            lib.learned_build(256, 100)
            e_min, e_max, _ = lib.energy_grid()
            for learned in (False, True):
                for i_nuc, n, st in ((0, 50000, stream), (3, 7000, None)):
                    pts = nucs[i_nuc].energy
                    E = np.concatenate([np.exp(rng.uniform(np.log(e_min * 0.5), np.log(e_max * 1.5), n // 2)),
                                        rng.choice(pts, n - n // 2)])
                    ig_r, xs_r, pr_r = lib.lookup_batch(i_nuc, E, learned=learned)
                    ig_n, xs_n, pr_n = _noalloc(lib, hip, i_nuc, E, learned, st)
                    assert np.array_equal(ig_r, ig_n), (learned, n)
                    assert np.array_equal(xs_r.view(np.uint64), xs_n.view(np.uint64)), (learned, n)
                    assert pr_r == pr_n, (learned, n, pr_r, pr_n)
            # an empty batch, and a batch beyond the scratch array's limit
            ig_n, xs_n, pr_n = _noalloc(lib, hip, 0, np.zeros(0), False, None)
            assert pr_n == 0
            with pytest.raises(lib.OpenMCError, match="scratch limit"):
                lib._dll.omc_lookup_batch_noalloc(0, None, 4194305, 0, None, None, None, None)
        finally:
            lib.finalize()
            hip.hipStreamDestroy(stream)
