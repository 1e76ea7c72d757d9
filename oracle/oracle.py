"""ORACLE -- test infrastructure only: ctypes wrapper of oracle/_build/liboracle.so.

Used by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg as
the checker and the CPU baseline; never by the product.
"""
import os
from ctypes import CDLL, POINTER, c_char_p, c_double, c_int, c_int64, c_uint64, c_void_p

import numpy as np

from . import build as _build
from .problem import write_oracle_problem

_libs = {}


def load_library(glibc_math=False):
    path = _build.LIB_GLIBC if glibc_math else _build.LIB
    if path in _libs:
        return _libs[path]
    if not os.path.exists(path):
        _build.build()
    lib = CDLL(path)
    lib.oracle_load.argtypes = [c_char_p]
    lib.oracle_load.restype = c_void_p
    lib.oracle_last_error.restype = c_char_p
    lib.oracle_free.argtypes = [c_void_p]
    lib.oracle_run.argtypes = [c_void_p, c_int, c_int, c_int64, c_int, c_int]
    lib.oracle_run.restype = c_int
    lib.oracle_n_tally_bins.argtypes = [c_void_p]
    lib.oracle_n_tally_bins.restype = c_int64
    lib.oracle_results.argtypes = [c_void_p]
    lib.oracle_results.restype = POINTER(c_double)
    lib.oracle_batch_values.argtypes = [c_void_p]
    lib.oracle_batch_values.restype = POINTER(c_double)
    lib.oracle_hashes.argtypes = [c_void_p]
    lib.oracle_hashes.restype = POINTER(c_uint64)
    lib.oracle_dhashes.argtypes = [c_void_p]
    lib.oracle_dhashes.restype = POINTER(c_uint64)
    lib.oracle_n_realizations.argtypes = [c_void_p]
    lib.oracle_counters.argtypes = [c_void_p, POINTER(c_int64)]
    lib.oracle_leakage.argtypes = [c_void_p, POINTER(c_double)]
    lib.oracle_energy_grid.argtypes = [c_void_p, POINTER(c_double)]
    lib.oracle_grid_index.argtypes = [c_void_p, c_int]
    lib.oracle_grid_index.restype = POINTER(c_int)
    lib.oracle_lookup.argtypes = [c_void_p, c_int, c_double, POINTER(c_int), POINTER(c_double)]
    lib.oracle_prn.argtypes = [POINTER(c_uint64)]
    lib.oracle_prn.restype = c_double
    lib.oracle_future_seed.argtypes = [c_uint64, c_uint64]
    lib.oracle_future_seed.restype = c_uint64
    lib.oracle_maxwell.argtypes = [c_double, POINTER(c_uint64)]
    lib.oracle_maxwell.restype = c_double
    lib.oracle_watt.argtypes = [c_double, c_double, POINTER(c_uint64)]
    lib.oracle_watt.restype = c_double
    lib.oracle_rotate_angle.argtypes = [POINTER(c_double), c_double, POINTER(c_double), POINTER(c_uint64),
                                        POINTER(c_double)]
    lib.oracle_ww_apply.argtypes = [c_void_p, POINTER(c_double), POINTER(c_uint64), POINTER(c_double),
                                    POINTER(c_double), c_int]
    lib.oracle_ww_apply.restype = c_int
    lib.oracle_surface_distance.argtypes = [c_int, POINTER(c_double), POINTER(c_double), POINTER(c_double), c_int]
    lib.oracle_surface_distance.restype = c_double
    lib.oracle_lattice_distance.argtypes = [c_void_p, c_int, POINTER(c_double), POINTER(c_double), POINTER(c_int)]
    lib.oracle_lattice_distance.restype = c_double
    lib.oracle_lattice_indices.argtypes = [c_void_p, c_int, POINTER(c_double), POINTER(c_double), POINTER(c_int)]
    lib.oracle_locate.argtypes = [c_void_p, POINTER(c_double), POINTER(c_double), POINTER(c_int)]
    lib.oracle_locate.restype = c_double
    lib.oracle_target_velocity.argtypes = [c_double, c_double, POINTER(c_double), c_double, POINTER(c_uint64),
                                           POINTER(c_double)]
    lib.oracle_elastic.argtypes = [c_double, c_double, c_double, c_double, POINTER(c_double), POINTER(c_uint64),
                                   POINTER(c_double)]
    lib.oracle_raytrace.argtypes = [c_void_p, POINTER(c_double), POINTER(c_double), POINTER(c_double),
                                    POINTER(c_int), POINTER(c_double), c_int]
    lib.oracle_raytrace.restype = c_int
    lib.oracle_in_source_domain.argtypes = [c_void_p, POINTER(c_double)]
    lib.oracle_sample_source.argtypes = [c_void_p, POINTER(c_uint64), POINTER(c_double)]
    for fn in ("oracle_log", "oracle_sin", "oracle_cos"):
        getattr(lib, fn).argtypes = [c_double]
        getattr(lib, fn).restype = c_double
    _libs[path] = lib
    return lib


def surface_distance(stype, coeffs, r, u, coincident=False):
    """Surface::distance of one surface (types: 1..3 x/y/z-plane, 4 plane,
    5 z-cylinder, 6 x-cylinder, 7 y-cylinder, 8 sphere)."""
    lib = load_library()
    c = (c_double * 4)(*(list(coeffs) + [0.0] * (4 - len(coeffs))))
    return lib.oracle_surface_distance(stype, c, (c_double * 3)(*r), (c_double * 3)(*u), 1 if coincident else 0)


def elastic(awr, kT, E, u, seed, free_gas_threshold=400.0):
    """elastic_scatter off a nuclide: (E', u', seed after)."""
    lib = load_library()
    s = c_uint64(seed)
    out = (c_double * 4)()
    lib.oracle_elastic(awr, kT, free_gas_threshold, E, (c_double * 3)(*u), s, out)
    return out[0], (out[1], out[2], out[3]), s.value


def target_velocity(awr, E, u, kT, seed):
    lib = load_library()
    s = c_uint64(seed)
    out = (c_double * 3)()
    lib.oracle_target_velocity(awr, E, (c_double * 3)(*u), kT, s, out)
    return (out[0], out[1], out[2]), s.value


class Oracle:
    """One loaded problem of the CPU restatement."""

    def __init__(self, problem_file, glibc_math=False):
        self.lib = load_library(glibc_math)
        h = self.lib.oracle_load(problem_file.encode())
        if not h:
            raise RuntimeError("oracle: " + self.lib.oracle_last_error().decode())
        self.h = h

    @classmethod
    def from_model(cls, model, nuclides, directory, glibc_math=False):
        path = write_oracle_problem(model, nuclides, directory)
        return cls(path, glibc_math)

    def close(self):
        if self.h:
            self.lib.oracle_free(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def run(self, event_based=False, threads=1, n_particles=0, n_batches=0, trace_hash=True):
        r = self.lib.oracle_run(self.h, 1 if event_based else 0, threads, n_particles, n_batches,
                                1 if trace_hash else 0)
        if r != 0:
            raise RuntimeError("oracle: " + self.lib.oracle_last_error().decode())

    @property
    def n_bins(self):
        return self.lib.oracle_n_tally_bins(self.h)

    def results(self):
        n = self.n_bins
        return np.ctypeslib.as_array(self.lib.oracle_results(self.h), (n, 1, 3)).copy()

    def hashes(self, n):
        return np.ctypeslib.as_array(self.lib.oracle_hashes(self.h), (n,)).copy()

    def discrete_hashes(self, n):
        """Hashes of the histories' discrete events only (event kinds, cells,
        surfaces, lattice moves, nuclides, grid indices, absorption), without
        any floating-point bits."""
        return np.ctypeslib.as_array(self.lib.oracle_dhashes(self.h), (n,)).copy()

    def n_realizations(self):
        return self.lib.oracle_n_realizations(self.h)

    def counters(self):
        out = (c_int64 * 9)()
        self.lib.oracle_counters(self.h, out)
        keys = ["lookups", "collisions", "segments", "lost", "xs_events", "adv_events", "surf_events",
                "coll_events", "scatter_fallthrough"]
        return dict(zip(keys, list(out)))

    def leakage(self):
        out = (c_double * 5)()
        self.lib.oracle_leakage(self.h, out)
        return dict(zip(["leak_sum", "leak_sum_sq", "leak_batch", "total_weight", "tracklength"], list(out)))

    def energy_grid(self):
        out = (c_double * 3)()
        self.lib.oracle_energy_grid(self.h, out)
        return tuple(out)

    def grid_index(self, i_nuc, n_log_bins=8000):
        return np.ctypeslib.as_array(self.lib.oracle_grid_index(self.h, i_nuc), (n_log_bins + 1,)).copy()

    def ww_apply(self, r, E, wgt, u=(1.0, 0.0, 0.0), wgt_born=1.0, wgt_ww_born=-1.0, ww_factor=0.0, n_split=0,
                 seed=1):
        """apply_weight_windows on one particle: returns (state, bank weights, seed after),
        state = dict(wgt, wgt_ww_born, ww_factor, n_split, n_bank)."""
        vals = list(r) + list(u) + [E, wgt, wgt_born, wgt_ww_born, ww_factor, float(n_split)]
        inp = (c_double * 12)(*vals)
        out = (c_double * 5)()
        bank = (c_double * 64)()
        s = c_uint64(seed)
        n = self.lib.oracle_ww_apply(self.h, inp, s, out, bank, 64)
        state = dict(zip(["wgt", "wgt_ww_born", "ww_factor", "n_split", "n_bank"], list(out)))
        return state, list(bank)[:min(n, 64)], s.value

    # elementary pieces, for the oracle's own pins
    def lattice_distance(self, lat, r, u):
        t = (c_int * 3)()
        d = self.lib.oracle_lattice_distance(self.h, lat, (c_double * 3)(*r), (c_double * 3)(*u), t)
        return d, tuple(t)

    def lattice_indices(self, lat, r, u):
        ijk = (c_int * 3)()
        self.lib.oracle_lattice_indices(self.h, lat, (c_double * 3)(*r), (c_double * 3)(*u), ijk)
        return tuple(ijk)

    def locate(self, r, u):
        """find_cell + distance_to_boundary at (r, u): (distance, info dict)."""
        out = (c_int * 15)()
        d = self.lib.oracle_locate(self.h, (c_double * 3)(*r), (c_double * 3)(*u), out)
        o = list(out)
        return d, {"found": o[0] == 1, "n_coord": o[1], "cells": o[2:2 + max(o[1], 0)],
                   "lattice_index": tuple(o[6:9]), "surface": o[9], "coord_level": o[10],
                   "translation": tuple(o[11:14])}

    def raytrace(self, r0, r1, u, cap=4096):
        bins = (c_int * cap)()
        fr = (c_double * cap)()
        n = self.lib.oracle_raytrace(self.h, (c_double * 3)(*r0), (c_double * 3)(*r1), (c_double * 3)(*u), bins,
                                     fr, cap)
        return list(bins)[:min(n, cap)], list(fr)[:min(n, cap)]

    def in_source_domain(self, r):
        return self.lib.oracle_in_source_domain(self.h, (c_double * 3)(*r)) == 1

    def sample_source(self, seed):
        s = c_uint64(seed)
        out = (c_double * 7)()
        self.lib.oracle_sample_source(self.h, s, out)
        return list(out), s.value

    def lookup(self, i_nuc, energies):
        E = np.asarray(energies, dtype=np.float64)
        ig = np.zeros(E.size, dtype=np.int32)
        xs = np.zeros((E.size, 4))
        g = c_int()
        o = (c_double * 4)()
        for k, e in enumerate(E):
            self.lib.oracle_lookup(self.h, i_nuc, float(e), g, o)
            ig[k] = g.value
            xs[k] = list(o)
        return ig, xs
