"""Reading a statepoint file back (a subset of openmc.StatePoint,
openmc/statepoint.py), through the library's libhdf5: no h5py needed.

The file layout is the one openmc_statepoint_write produces
(src/state_point.cpp:36-360); results hold (sum, sum_sq) per filter bin and
score, and mean / std_dev follow Tally.mean and Tally.std_dev
(openmc/tallies.py): mean = sum / n, std_dev = sqrt((sum_sq / n - mean^2) / (n - 1)).
"""
import os

import numpy as np

from . import lib

GLOBAL_TALLY_NAMES = ["k-collision", "k-absorption", "k-tracklength", "leakage"]


class MeshInfo:
    def __init__(self, path, mesh_id):
        g = "tallies/meshes/mesh %d" % mesh_id
        self.id = mesh_id
        self.type = lib.h5_read(path, g + "/type", kind="str")[0]
        self.dimension = lib.h5_read(path, g + "/dimension", kind="i8")
        self.lower_left = lib.h5_read(path, g + "/lower_left")
        self.upper_right = lib.h5_read(path, g + "/upper_right")
        self.width = lib.h5_read(path, g + "/width")


class TallyInfo:
    def __init__(self, path, tally_id, present):
        g = "tallies/tally %d" % tally_id
        self.id = tally_id
        self.estimator = lib.h5_read(path, g + "/estimator", kind="str")[0]
        self.num_realizations = int(lib.h5_read(path, g + "/n_realizations", kind="i8")[0])
        n_filters = int(lib.h5_read(path, g + "/n_filters", kind="i8")[0])
        self.filter_ids = list(lib.h5_read(path, g + "/filters", kind="i8")) if n_filters else []
        self.nuclides = lib.h5_read(path, g + "/nuclides", kind="str")
        self.scores = lib.h5_read(path, g + "/score_bins", kind="str")
        self.sum = self.sum_sq = None
        if present:
            r = lib.h5_read(path, g + "/results").reshape(-1, len(self.scores), 2)
            self.sum = r[:, :, 0]
            self.sum_sq = r[:, :, 1]

    @property
    def mean(self):
        return self.sum / self.num_realizations

    @property
    def std_dev(self):
        n = self.num_realizations
        mean = self.mean
        if n < 2:
            return np.zeros_like(mean)
        return np.sqrt(np.maximum(self.sum_sq / n - mean * mean, 0.0) / (n - 1))


class StatePoint:
    """The header, global tallies, meshes, filters and tallies of a statepoint."""

    def __init__(self, path):
        self.path = os.path.abspath(os.fspath(path))
        p = self.path
        self.filetype = lib.h5_read(p, "", "filetype", kind="str")[0]
        if self.filetype != "statepoint":
            raise IOError("%s is not a statepoint file" % p)
        self.version = tuple(int(v) for v in lib.h5_read(p, "", "version", kind="i8"))
        self.seed = int(lib.h5_read(p, "seed", kind="i8")[0])
        self.stride = int(lib.h5_read(p, "stride", kind="i8")[0])
        self.run_mode = lib.h5_read(p, "run_mode", kind="str")[0]
        self.energy_mode = lib.h5_read(p, "energy_mode", kind="str")[0]
        self.n_particles = int(lib.h5_read(p, "n_particles", kind="i8")[0])
        self.n_batches = int(lib.h5_read(p, "n_batches", kind="i8")[0])
        self.current_batch = int(lib.h5_read(p, "current_batch", kind="i8")[0])
        self.n_realizations = int(lib.h5_read(p, "n_realizations", kind="i8")[0])
        self.source_present = bool(lib.h5_read(p, "", "source_present", kind="i8")[0])
        self.tallies_present = bool(lib.h5_read(p, "", "tallies_present", kind="i8")[0])
        gt = lib.h5_read(p, "global_tallies").reshape(-1, 3)
        n = self.n_realizations
        self.global_tallies = {}
        for name, row in zip(GLOBAL_TALLY_NAMES, gt):
            mean = row[1] / n if n else 0.0
            std = np.sqrt(max(row[2] / n - mean * mean, 0.0) / (n - 1)) if n > 1 else 0.0
            self.global_tallies[name] = {"sum": row[1], "sum_sq": row[2], "mean": mean, "std_dev": std}
        n_meshes = int(lib.h5_read(p, "tallies/meshes", "n_meshes", kind="i8")[0])
        ids = lib.h5_read(p, "tallies/meshes", "ids", kind="i8") if n_meshes else []
        self.meshes = {int(i): MeshInfo(p, int(i)) for i in ids}
        n_filters = int(lib.h5_read(p, "tallies/filters", "n_filters", kind="i8")[0])
        self.filters = {}
        for i in (lib.h5_read(p, "tallies/filters", "ids", kind="i8") if n_filters else []):
            g = "tallies/filters/filter %d" % i
            ftype = lib.h5_read(p, g + "/type", kind="str")[0]
            bins = lib.h5_read(p, g + "/bins", kind="i8" if ftype == "mesh" else "f8")
            self.filters[int(i)] = {"type": ftype, "n_bins": int(lib.h5_read(p, g + "/n_bins", kind="i8")[0]),
                                    "bins": bins}
        n_tallies = int(lib.h5_read(p, "tallies", "n_tallies", kind="i8")[0])
        ids = lib.h5_read(p, "tallies", "ids", kind="i8") if n_tallies else []
        self.tallies = {int(i): TallyInfo(p, int(i), self.tallies_present) for i in ids}
        self.runtime = {k: float(lib.h5_read(p, "runtime/" + k)[0])
                        for k in ("total initialization", "simulation", "active batches", "total",
                                  "writing statepoints")}

    def source_sites(self):
        """Number of sites in the source-bank dataset (0 when absent)."""
        if not self.source_present:
            return 0
        return lib.h5_count(self.path, "source_bank")
