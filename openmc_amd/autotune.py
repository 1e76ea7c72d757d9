"""Launch-option autotuner for the event loop.

The options below change the order and shape of the GPU work, never a
history (every one of them is covered by the launch-shape parity tests):

  blocks            workgroups of the event kernels (0 = from the capacity)
  block             threads per workgroup of the event kernels (64, 128, 256)
  sort_bits         energy buckets per binade of the calculate-xs sort key
  sort_every        full (material, energy) sort every n cycles, material-only between
  cycles_per_check  event cycles per host check of the queue counts (and per hipGraph)
  xs_group          nuclide lookups interleaved per lane in calculate-xs
  history_below     the queue threshold below which a batch finishes history-based
  fuse_advance      advance in the calculate-xs launch (xs_group 2 or 4; -1: when
                    the collisions repeat the lookups)
  recompute_micro   collisions repeat the nuclide lookups instead of reading
                    the micro-total rows (-1: chosen by material size)

``tune`` times whole batches of the problem (or problem set) that the
library holds and picks, option by option, the fastest value (coordinate
descent from the current settings; a change must win by more than
``min_gain`` to be kept).  Each candidate replays batch 1 after a hard reset,
so every candidate transports the same particles.  The library is left with
the chosen options set, the batch counter at 0 and the tallies cleared.

``tune_cached`` stores the result in a JSON file keyed by a signature of the
model files and the library, and reuses it while the signature matches.
"""
import hashlib
import json
import os
import time

from . import lib

DEFAULT_SPACE = [
    ("blocks", [0, 4096, 8192, 16384, 32768]),
    ("block", [256, 128, 64]),
    ("sort_bits", [-1, 8, 6, 5, 3, 1]),
    ("sort_every", [1, 2, 4]),
    ("cycles_per_check", [16, 8, 32, 64]),
    ("xs_group", [-1, 2, 4, 8]),
    ("history_below", [65536, 0, 16384, 262144, 1048576]),
    ("fuse_advance", [-1, 0, 1]),
    ("recompute_micro", [-1, 0, 1]),
]
DEFAULTS = {"blocks": 0, "block": 256, "sort_bits": -1, "sort_every": 1, "cycles_per_check": 16, "xs_group": -1,
            "history_below": 65536, "fuse_advance": -1, "recompute_micro": -1}


def _time_batch():
    """Wall time of batch 1 of the loaded problem (after a hard reset, which
    also restores the default seed and stride: the caller's are put back)."""
    seed, stride = lib.settings.seed, lib.settings.stride
    lib.hard_reset()
    lib.settings.seed, lib.settings.stride = seed, stride
    t = time.perf_counter()
    lib.next_batch()
    return time.perf_counter() - t


def tune(space=None, start=None, repeats=1, min_gain=0.01, log=None):
    """Coordinate descent over `space` (list of (option, values)) on the
    problem the library holds.  Returns (best options, list of trials)."""
    space = space or DEFAULT_SPACE
    best = dict(DEFAULTS)
    best.update(start or {})
    for k, v in best.items():
        lib.set_option(k, v)
    lib.simulation_init()
    _time_batch()  # first touch of the particle state, graph capture
    best_t = min(_time_batch() for _ in range(repeats))
    trials = [{"options": dict(best), "seconds": best_t}]
    if log:
        log("autotune start %s: %.4f s" % (best, best_t))
    for name, values in space:
        for v in values:
            if v == best.get(name):
                continue
            lib.set_option(name, v)
            t = min(_time_batch() for _ in range(repeats))
            cand = dict(best, **{name: v})
            trials.append({"options": cand, "seconds": t})
            if log:
                log("autotune %s=%s: %.4f s" % (name, v, t))
            if t < best_t * (1.0 - min_gain):
                best, best_t = cand, t
        lib.set_option(name, best[name])
    seed, stride = lib.settings.seed, lib.settings.stride
    lib.hard_reset()
    lib.settings.seed, lib.settings.stride = seed, stride
    lib.reset()
    return best, trials


def signature(directories):
    """Hash of the model files of a problem (set) and of the built library."""
    h = hashlib.sha1()
    for d in directories:
        with open(os.path.join(d, "model.xml"), "rb") as f:
            h.update(f.read())
    lib_path = lib._filename
    st = os.stat(lib_path)
    h.update(("%d:%d" % (st.st_size, int(st.st_mtime))).encode())
    return h.hexdigest()


def tune_cached(directories, path, **kw):
    """tune(), reusing the options stored at `path` when the problem files and
    the library are unchanged.  Returns (options, record)."""
    sig = signature(directories)
    if os.path.exists(path):
        try:
            with open(path) as f:
                rec = json.load(f)
            if rec.get("signature") == sig:
                for k, v in rec["options"].items():
                    lib.set_option(k, v)
                return rec["options"], dict(rec, cached=True)
        except (OSError, ValueError, KeyError):
            pass
    best, trials = tune(**kw)
    rec = {"signature": sig, "options": best, "trials": trials}
    try:
        with open(path, "w") as f:
            json.dump(rec, f, indent=1)
    except OSError:
        pass
    return best, rec
