"""GPU parity: the HIP path through the C-ABI against the oracle.

Bit-exact for everything integer- or history-shaped (grid indices, per-
particle history hashes that fold in every event's type, surface, nuclide,
grid index, distance, energy and direction bits); FP64 tallies within 1e-12
relative per bin (only the order of the atomic additions differs).
"""
import numpy as np
import pytest

from helpers import assert_tallies_close, gpu_run, oracle_run, prepare

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def lib():
    from openmc_amd import lib as _lib
    return _lib


def _oracle_lib():
    from oracle.oracle import load_library
    return load_library()


# ---------------------------------------------------------------- primitives

def test_math_primitives_bitwise(lib):
    rng = np.random.default_rng(7)
    ol = _oracle_lib()
    n = 200000
    x_log = np.concatenate([rng.random(n // 4), np.exp(rng.uniform(-40, 40, n // 4)),
                            1.0 + (rng.random(n // 4) - 0.5) * 1e-6,
                            np.ldexp(rng.integers(1, 2**53, n // 4).astype(np.float64), -53)])
    x_ang = np.concatenate([rng.random(n // 2) * 2 * np.pi, rng.random(n // 2) * np.pi / 2])
    for fn, xs, ref in [(0, x_log, ol.oracle_log), (1, x_ang, ol.oracle_sin), (2, x_ang, ol.oracle_cos)]:
        g = lib.math_batch(fn, xs)
        h = np.array([ref(float(v)) for v in xs])
        assert np.array_equal(g.view(np.uint64), h.view(np.uint64)), "fn %d differs" % fn
    xs = rng.random(n) * 1e6
    ys = rng.random(n) * 3 + 1e-3
    assert np.array_equal(lib.math_batch(3, xs).view(np.uint64), np.sqrt(xs).view(np.uint64))
    assert np.array_equal(lib.math_batch(4, xs, ys).view(np.uint64), (xs / ys).view(np.uint64))
    seeds = rng.integers(0, 2**63, 5000, dtype=np.int64).view(np.float64)
    g = lib.math_batch(5, seeds)
    from ctypes import c_uint64
    h = []
    for s in seeds.view(np.uint64):
        c = c_uint64(int(s))
        h.append(ol.oracle_prn(c))
    assert np.array_equal(g, np.array(h))
    big = rng.integers(0, 2**63, 5000, dtype=np.int64).view(np.uint64)
    big[::2] |= np.uint64(1) << np.uint64(63)
    g = lib.math_batch(6, big.view(np.float64))
    assert np.array_equal(g, big.astype(np.float64))


# ---------------------------------------------------------------- lookups

def test_lookup_plain_and_learned_bitwise(lib):
    d, model, nucs = prepare(0, n_particles=100, n_batches=1)
    from oracle.oracle import Oracle
    import os
    o = Oracle(os.path.join(d, "oracle_problem.txt"))
    from helpers import chdir
    rng = np.random.default_rng(11)
    with chdir(d):
        lib.init(output=False)
        try:
            max_err = lib.learned_build(256, 200)
            assert max_err <= 64
            for i_nuc in range(len(nucs)):
                pts = nucs[i_nuc].energy
                e_min, e_max, spacing = o.energy_grid()
                E = np.concatenate([np.exp(rng.uniform(np.log(e_min), np.log(e_max), 20000)), pts[1:-1],
                                    0.5 * (pts[:-1] + pts[1:]), e_min * np.exp(np.arange(8000) * spacing),
                                    [pts[0] * 0.5, pts[-1] * 1.5]])
                ig_o, xs_o = o.lookup(i_nuc, E)
                ig_g, xs_g, pr_plain = lib.lookup_batch(i_nuc, E, learned=False)
                assert np.array_equal(ig_o, ig_g)
                assert np.array_equal(xs_o[:, 1:3].view(np.uint64), xs_g.view(np.uint64))
                ig_l, xs_l, pr_learned = lib.lookup_batch(i_nuc, E, learned=True)
                assert np.array_equal(ig_o, ig_l)
                assert np.array_equal(xs_g.view(np.uint64), xs_l.view(np.uint64))
        finally:
            lib.finalize()


# ---------------------------------------------------------------- transport

@pytest.mark.parametrize("mode", ["history", "event"])
def test_transport_cfg0_histories_and_tallies(mode):
    d, model, nucs = prepare(0, n_particles=3000, n_batches=2)
    o = oracle_run(d, event_based=(mode == "event"))
    g = gpu_run(d, options={"trace_hash": 1})
    assert g["n_real"] == 2
    # last batch histories, bit for bit
    h_o = o.hashes(3000)
    assert np.array_equal(g["hashes"], h_o), "%d of 3000 histories differ" % np.sum(g["hashes"] != h_o)
    assert_tallies_close(g["results"], o.results())
    c = o.counters()
    s = g["stats"]
    assert s["lost"] == 0 and c["lost"] == 0
    assert s["collisions"] == c["collisions"]
    assert s["segments"] == c["segments"]
    assert s["surf_events"] == c["surf_events"]
    # the elastic cross section is evaluated at every scatter (A12); the
    # inelastic fallthrough needs a cutoff within rounding of sigma_s
    assert s["scatter_fallthrough"] == c["scatter_fallthrough"] == 0


@pytest.mark.parametrize("filters", ["none", "energy"])
def test_transport_tally_without_mesh_filter(filters):
    """Tallies with no filter (one bin) or an energy filter alone, as the
    reference's fixed_source regression model uses: histories bit-identical,
    tallies within 1e-12 per bin."""
    print("case: cfg0 tally filters %s" % filters, flush=True)
    kw = dict(n_particles=2000, n_batches=2, mesh=None)
    if filters == "energy":
        kw["energy_bins"] = [1e-5, 1.0, 1e3, 1e5, 2e7]
    d, model, nucs = prepare(0, **kw)
    o = oracle_run(d)
    g = gpu_run(d, options={"trace_hash": 1})
    assert g["results"].shape[0] == (1 if filters == "none" else 4)
    assert np.array_equal(g["hashes"], o.hashes(2000))
    assert_tallies_close(g["results"], o.results())
    assert g["stats"]["segments"] == o.counters()["segments"]


def test_transport_cfg1_lattice_energy_filter():
    d, model, nucs = prepare(1, n_particles=2000, n_batches=2, grid_lo=2000, grid_hi=6000, n_nuclides=12,
                             lattice_n=5, mesh=(10, 10, 20), half_height=20.0,
                             energy_bins=[1e-5, 1.0, 1e3, 1e5, 2e7])
    o = oracle_run(d)
    g = gpu_run(d, options={"trace_hash": 1})
    h_o = o.hashes(2000)
    assert np.array_equal(g["hashes"], h_o), "%d of 2000 histories differ" % np.sum(g["hashes"] != h_o)
    assert_tallies_close(g["results"], o.results())
    assert g["stats"]["surf_events"] == o.counters()["surf_events"]
    assert g["stats"]["surf_events"] > 0


def test_transport_graph_and_direct_launch_identical():
    d, model, nucs = prepare(0, n_particles=2000, n_batches=1)
    a = gpu_run(d, options={"trace_hash": 1, "use_graph": 1})
    b = gpu_run(d, options={"trace_hash": 1, "use_graph": 0, "cycles_per_check": 3})
    c = gpu_run(d, options={"trace_hash": 1, "max_in_flight": 700})
    assert np.array_equal(a["hashes"], b["hashes"])
    assert np.array_equal(a["hashes"], c["hashes"])
    assert_tallies_close(a["results"], b["results"])
    assert_tallies_close(a["results"], c["results"])


@pytest.mark.parametrize("opts", [{"sort": 0, "xs_lanes": 1}, {"sort": 1, "xs_lanes": 2},
                                  {"sort": 1, "xs_lanes": 64}, {"sort": 0, "xs_lanes": 8},
                                  {"sort": 1, "xs_lanes": 1, "sort_blocks": 3, "blocks": 5},
                                  {"sort": 1, "xs_lanes": 1, "sort_blocks": 256},
                                  {"sort": 1, "xs_lanes": 1, "xs_group": 1},
                                  {"sort": 0, "xs_lanes": 1, "xs_group": 2},
                                  {"sort": 1, "xs_lanes": 1, "xs_group": 8},
                                  {"block": 64, "blocks": 7}, {"block": 128},
                                  {"sort_every": 4, "cycles_per_check": 8}, {"sort_every": 3, "sort_bits": 0},
                                  {"wave_search": 1}, {"wave_search": 1, "xs_group": 2}, {"tally_copies": 8},
                                  {"sort": 0, "xs_group": 8}, {"history_below": 0},
                                  {"history_below": 1 << 30, "history_steps": 3, "cycles_per_check": 1},
                                  {"history_below": 200, "history_steps": 1}, {"fuse_advance": 0},
                                  {"fuse_advance": 1, "wave_search": 1, "history_below": 0},
                                  {"recompute_micro": 1}, {"recompute_micro": 1, "fuse_advance": 1},
                                  {"recompute_micro": 1, "history_below": 1 << 30, "history_steps": 3},
                                  {"recompute_micro": 0}, {"recompute_micro": 0, "history_below": 0},
                                  {"fuse_advance": 1, "xs_group": 2}, {"fuse_advance": 1, "xs_group": 2,
                                                                       "recompute_micro": 0}])
def test_transport_launch_shapes_identical(opts):
    """Queue sorting and lanes-per-particle change the order of work, never a history."""
    d, model, nucs = prepare(1, n_particles=1500, n_batches=1, grid_lo=2000, grid_hi=6000, n_nuclides=12,
                             lattice_n=5, mesh=(10, 10, 20), half_height=20.0)
    o = oracle_run(d)
    g = gpu_run(d, options=dict(trace_hash=1, **opts))
    assert np.array_equal(g["hashes"], o.hashes(1500))
    assert_tallies_close(g["results"], o.results())


@pytest.mark.parametrize("group", [1, 4, 8])
def test_transport_learned_index_same_histories(group):
    d, model, nucs = prepare(0, n_particles=2000, n_batches=1)
    a = gpu_run(d, options={"trace_hash": 1})
    b = gpu_run(d, options={"trace_hash": 1, "xs_group": group}, learned=(256, 200))
    assert np.array_equal(a["hashes"], b["hashes"])
    assert_tallies_close(a["results"], b["results"])


def test_learned_loss_and_gradient_match_cpu(lib):
    """Forward/backward kernels of the learned index vs a numpy restatement."""
    d, model, nucs = prepare(0, n_particles=100, n_batches=1)
    from helpers import chdir
    ol = _oracle_lib()
    n_seg = 64
    with chdir(d):
        lib.init(output=False)
        try:
            e_min, e_max, _ = lib.energy_grid()
            for i_nuc in range(2):
                E = nucs[i_nuc].energy
                x = np.array([ol.oracle_log(float(v / e_min)) for v in E])
                u_lo = ol.oracle_log(float(E[0] / e_min))
                u_hi = ol.oracle_log(float(E[-1] / e_min))
                inv_w = n_seg / (u_hi - u_lo)
                uu = (x - u_lo) * inv_w
                seg = np.clip(np.where(uu < 0, 0, uu.astype(np.int64)), 0, n_seg - 1)
                rng = np.random.default_rng(i_nuc)
                params = np.stack([rng.uniform(100, 300, n_seg), rng.uniform(0, E.size, n_seg)], 1)
                params = params.astype(np.float32)
                f = (uu - seg).astype(np.float32)
                a = params[seg, 0]
                b = params[seg, 1]
                pred = (a * f).astype(np.float32) + b
                r = pred.astype(np.float64) - np.arange(E.size)
                loss_ref = np.bincount(seg, weights=r * r, minlength=n_seg)
                ga_ref = np.bincount(seg, weights=2 * r * f.astype(np.float64), minlength=n_seg)
                gb_ref = np.bincount(seg, weights=2 * r, minlength=n_seg)
                loss, grad = lib.learned_loss_grad(i_nuc, params)
                assert np.allclose(loss, loss_ref, rtol=1e-10, atol=1e-6)
                assert np.allclose(grad[:, 0], ga_ref, rtol=1e-10, atol=1e-6)
                assert np.allclose(grad[:, 1], gb_ref, rtol=1e-10, atol=1e-6)
                # a plausible mistake (dropping the factor f) must not pass
                assert not np.allclose(grad[:, 0], gb_ref, rtol=1e-3)
        finally:
            lib.finalize()


def test_learned_gradient_matches_finite_differences(lib):
    """The backward kernel's gradient equals central differences of the forward
    kernel's loss.  The loss of one segment is quadratic in (a, b), so a
    central difference with a unit step is exact up to the FP32 rounding of
    the prediction; this pins the backward kernel without restating it."""
    print("case: learned gradient vs finite differences, cfg0 nuclides", flush=True)
    d, model, nucs = prepare(0, n_particles=100, n_batches=1)
    from helpers import chdir
    n_seg = 32
    with chdir(d):
        lib.init(output=False)
        try:
            for i_nuc in range(2):
                rng = np.random.default_rng(100 + i_nuc)
                n = nucs[i_nuc].energy.size
                params = np.stack([rng.uniform(0, 600, n_seg), rng.uniform(0, n, n_seg)], 1).astype(np.float32)
                loss, grad = lib.learned_loss_grad(i_nuc, params)
                for col in range(2):
                    up = params.copy()
                    dn = params.copy()
                    up[:, col] += 1.0
                    dn[:, col] -= 1.0
                    fd = (lib.learned_loss_grad(i_nuc, up)[0] - lib.learned_loss_grad(i_nuc, dn)[0]) / 2.0
                    scale = np.abs(grad[:, col]).max()
                    assert np.allclose(grad[:, col], fd, rtol=1e-4, atol=1e-4 * scale), (i_nuc, col)
                    # the other partial does not match (a swapped column would pass otherwise)
                    assert not np.allclose(grad[:, 1 - col], fd, rtol=1e-2, atol=1e-3 * scale)
        finally:
            lib.finalize()


def test_learned_error_bound_enforced_by_doubling(lib):
    """With a tight bound the build trains again, with twice the segments, the
    nuclides whose recorded error exceeds it; every recorded error is then
    within the bound and every lookup still equals the oracle's index.  (The
    bound must exceed half the largest cluster of points that no segment can
    split: the 40 points of one resonance span ~1e-8 in log E, so about 20.)"""
    print("case: learned error bound 32, cfg0 nuclides", flush=True)
    d, model, nucs = prepare(0, n_particles=100, n_batches=1)
    from oracle.oracle import Oracle
    import os
    o = Oracle(os.path.join(d, "oracle_problem.txt"))
    from helpers import chdir
    rng = np.random.default_rng(5)
    with chdir(d):
        lib.init(output=False)
        try:
            lib.set_option("learned_err_bound", 32)
            assert lib.learned_build(16, 100) <= 32
            st = lib.learned_stats()
            assert st["rounds"] > 1 and st["err_bound"] == 32 and st["max_err"] <= 32
            assert st["max_segments"] > 16 and st["min_segments"] >= 16
            sizes = []
            for i_nuc in range(len(nucs)):
                p, e = lib.learned_model(i_nuc)
                assert e.max() <= 32
                n_s = p.shape[0]
                assert n_s % 16 == 0 and (n_s // 16) & (n_s // 16 - 1) == 0
                sizes.append(n_s)
                pts = nucs[i_nuc].energy
                e_min, e_max, spacing = o.energy_grid()
                E = np.concatenate([np.exp(rng.uniform(np.log(e_min), np.log(e_max), 20000)), pts,
                                    0.5 * (pts[:-1] + pts[1:]), e_min * np.exp(np.arange(8000) * spacing)])
                ig_o, xs_o = o.lookup(i_nuc, E)
                ig_l, xs_l, probes = lib.lookup_batch(i_nuc, E, learned=True)
                assert np.array_equal(ig_o, ig_l)
                assert np.array_equal(xs_o[:, 1:3].view(np.uint64), xs_l.view(np.uint64))
                # the recorded bound holds for every query: |pred - lower bound| <= err
                # of its segment, with the prediction in FP32 as the lookup makes it
                Ein = E[(E >= pts[0]) & (E <= pts[-1])]
                lb = np.searchsorted(pts, Ein, side="left")
                ol = _oracle_lib()
                x = np.array([ol.oracle_log(float(v / e_min)) for v in Ein])
                u_lo = ol.oracle_log(float(pts[0] / e_min))
                inv_w = n_s / (ol.oracle_log(float(pts[-1] / e_min)) - u_lo)
                uu = (x - u_lo) * inv_w
                seg = np.clip(np.where(uu < 0, 0, np.floor(uu)).astype(np.int64), 0, n_s - 1)
                f = (uu - seg).astype(np.float32)
                pred = (p[seg, 0] * f).astype(np.float32) + p[seg, 1]
                assert np.all(np.abs(pred.astype(np.float64) - lb) <= e[seg] + 1e-9)
            assert sum(sizes) == st["total_segments"]
            with pytest.raises(lib.OpenMCError, match="learned_err_bound"):
                lib.set_option("learned_err_bound", 65)
        finally:
            lib.finalize()


def test_learned_training_reduces_loss_and_bounds_hold(lib):
    d, model, nucs = prepare(0, n_particles=100, n_batches=1)
    from helpers import chdir
    with chdir(d):
        lib.init(output=False)
        try:
            lib.learned_build(256, 200)
            st = lib.learned_stats()
            p1, e1 = lib.learned_model(0, 256)
            assert e1.max() <= 64 and st["max_err"] <= 64
            # the Adam steps on the forward/backward kernels lower the squared error
            assert st["loss_trained"] < st["loss_init"]
        finally:
            lib.finalize()


# ---------------------------------------------------------------- weight windows

WW_SLAB = dict(n_particles=600, n_batches=2, n_grid=2000, n_layers=3, layer=4.0, per_layer=3, half_width=5.0,
               density=0.05, ww_mesh=(12, 1, 1), ww_length=2.0, tally_mesh=(12, 1, 1), energy_groups=2)


def test_weight_windows_bank_overflow_is_an_error():
    """A bank too small for the splitting stops the run with an error (the
    kernels drop the extra sites inside their bounds and flag it)."""
    d, model, nucs = prepare(4, **WW_SLAB)
    from openmc_amd import lib
    with pytest.raises(lib.OpenMCError, match="secondary bank overflowed"):
        gpu_run(d, options={"bank_cap": 2})


@pytest.mark.parametrize("bank_cap,tail,recompute,sets", [(64, 65536, -1, 1), (64, 0, -1, 1), (64, 65536, 0, 1),
                                                          (64, 65536, -1, 2)],
                         ids=["history-tail", "event-only", "micro-rows", "two-sets"])
def test_transport_weight_windows_histories(bank_cap, tail, recompute, sets):
    """Splitting, roulette and LIFO revival from the per-particle bank give the
    oracle's histories bit for bit (3-layer slab, 12-bin windows), with the
    history-based tail and with event kernels only; with two window sets a
    particle takes the first set with a valid window at its position."""
    print("case: weight windows bank %d tail %d recompute %d sets %d" % (bank_cap, tail, recompute, sets),
          flush=True)
    d, model, nucs = prepare(4, ww_split_sets=(sets == 2), **WW_SLAB)
    o = oracle_run(d, event_based=True)
    g = gpu_run(d, options={"trace_hash": 1, "bank_cap": bank_cap, "history_below": tail,
                            "recompute_micro": recompute})
    h_o = o.hashes(600)
    assert np.array_equal(g["hashes"], h_o), "%d of 600 histories differ" % np.sum(g["hashes"] != h_o)
    assert_tallies_close(g["results"], o.results())
    c = o.counters()
    assert g["stats"]["collisions"] == c["collisions"]
    assert g["stats"]["surf_events"] == c["surf_events"]
    lk = o.leakage()
    # global tallies [K_COLLISION, K_ABSORPTION, K_TRACKLENGTH, LEAKAGE] x [VALUE, SUM, SUM_SQ]
    assert np.asarray(g["global"]).reshape(-1, 3)[3, 1] == pytest.approx(lk["leak_sum"], rel=1e-12)
