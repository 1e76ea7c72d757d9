"""ctypes bindings to this build's libopenmc.so, mirroring ``openmc.lib``.

Same entry points, argument meanings and error behaviour as the reference's
bindings (openmc/lib/__init__.py, core.py, tally.py, settings.py, error.py):
``init``/``run``/``next_batch``/``finalize``, ``settings.event_based`` and
friends as data descriptors over the exported globals, ``tallies[id].mean`` /
``.std_dev`` computed from the zero-copy results array, and error codes mapped
to exceptions.  Extra functions specific to the GPU path (options, run
statistics, history hashes, batched lookups, learned index) are prefixed
nothing special but documented as build-specific.

Importing this module loads libopenmc.so and fails loudly if it is missing;
nothing here falls back to a CPU implementation.
"""
import os
from contextlib import contextmanager
from ctypes import (CDLL, POINTER, byref, c_bool, c_char, c_char_p, c_double, c_float, c_int, c_int32,
                    c_int64, c_size_t, c_uint64, c_void_p, cast, create_string_buffer)

import numpy as np

_filename = os.path.join(os.path.dirname(os.path.abspath(__file__)), "libopenmc.so")
if not os.path.exists(_filename):
    raise ImportError("libopenmc.so has not been built (run __graft_entry__.build())")
_dll = CDLL(_filename)


class OpenMCError(Exception):
    pass


class AllocationError(OpenMCError):
    pass


class OutOfBoundsError(OpenMCError):
    pass


class InvalidArgumentError(OpenMCError):
    pass


class InvalidTypeError(OpenMCError):
    pass


class InvalidIDError(OpenMCError):
    pass


class GeometryError(OpenMCError):
    pass


class DataError(OpenMCError):
    pass


class PhysicsError(OpenMCError):
    pass


def _errcode(s):
    return c_int.in_dll(_dll, s).value


def _error_handler(err, func, args):
    """Raise the exception matching the error code (openmc/lib/error.py:8-41)."""
    msg = (c_char * 256).in_dll(_dll, "openmc_err_msg").value.decode()
    table = [("OPENMC_E_ALLOCATE", AllocationError), ("OPENMC_E_OUT_OF_BOUNDS", OutOfBoundsError),
             ("OPENMC_E_INVALID_ARGUMENT", InvalidArgumentError), ("OPENMC_E_INVALID_TYPE", InvalidTypeError),
             ("OPENMC_E_INVALID_ID", InvalidIDError), ("OPENMC_E_GEOMETRY", GeometryError),
             ("OPENMC_E_DATA", DataError), ("OPENMC_E_PHYSICS", PhysicsError)]
    for name, exc in table:
        if err == _errcode(name):
            raise exc(msg)
    if err < 0:
        raise OpenMCError(msg or "Unknown error encountered (code %d)." % err)
    return err


def _bind(name, argtypes, restype=c_int, check=True):
    f = getattr(_dll, name)
    f.argtypes = argtypes
    f.restype = restype
    if check and restype is c_int:
        f.errcheck = _error_handler
    return f


_bind("openmc_init", [c_int, POINTER(POINTER(c_char)), c_void_p])
_bind("openmc_run", [])
_bind("openmc_simulation_init", [])
_bind("openmc_simulation_finalize", [])
_bind("openmc_next_batch", [POINTER(c_int)])
_bind("openmc_finalize", [])
_bind("openmc_reset", [])
_bind("openmc_hard_reset", [])
_bind("openmc_reset_timers", [])
_bind("openmc_get_tally_index", [c_int32, POINTER(c_int32)])
_bind("openmc_tally_get_id", [c_int32, POINTER(c_int32)])
_bind("openmc_tally_get_n_realizations", [c_int32, POINTER(c_int32)])
_bind("openmc_tally_results", [c_int32, POINTER(POINTER(c_double)), POINTER(c_size_t)])
_bind("openmc_global_tallies", [POINTER(POINTER(c_double))])
_bind("tallies_size", [], c_size_t, check=False)
_bind("openmc_get_seed", [], c_int64, check=False)
_bind("openmc_set_seed", [c_int64], None, check=False)
_bind("openmc_get_stride", [], c_uint64, check=False)
_bind("openmc_set_stride", [c_uint64], None, check=False)
# model queries (openmc/lib/{tally,filter,mesh,material,cell,nuclide}.py)
_bind("openmc_tally_get_filters", [c_int32, POINTER(POINTER(c_int32)), POINTER(c_size_t)])
_bind("openmc_tally_get_scores", [c_int32, POINTER(POINTER(c_int)), POINTER(c_int)])
_bind("openmc_tally_get_nuclides", [c_int32, POINTER(POINTER(c_int)), POINTER(c_int)])
_bind("openmc_tally_get_estimator", [c_int32, POINTER(c_int)])
_bind("openmc_tally_get_type", [c_int32, POINTER(c_int32)])
_bind("openmc_tally_get_active", [c_int32, POINTER(c_bool)])
_bind("openmc_tally_get_writable", [c_int32, POINTER(c_bool)])
_bind("openmc_tally_reset", [c_int32])
_bind("tally_filters_size", [], c_size_t, check=False)
_bind("openmc_get_filter_index", [c_int32, POINTER(c_int32)])
_bind("openmc_filter_get_id", [c_int32, POINTER(c_int32)])
_bind("openmc_filter_get_type", [c_int32, c_char_p])
_bind("openmc_filter_get_num_bins", [c_int32, POINTER(c_int)])
_bind("openmc_mesh_filter_get_mesh", [c_int32, POINTER(c_int32)])
_bind("openmc_energy_filter_get_bins", [c_int32, POINTER(POINTER(c_double)), POINTER(c_size_t)])
_bind("n_meshes", [], c_int, check=False)
_bind("openmc_get_mesh_index", [c_int32, POINTER(c_int32)])
_bind("openmc_mesh_get_id", [c_int32, POINTER(c_int32)])
_bind("openmc_mesh_get_type", [c_int32, c_char_p])
_bind("openmc_mesh_get_n_elements", [c_int32, POINTER(c_size_t)])
_bind("openmc_regular_mesh_get_dimension", [c_int32, POINTER(POINTER(c_int)), POINTER(c_int)])
_bind("openmc_regular_mesh_get_params", [c_int32, POINTER(POINTER(c_double)), POINTER(POINTER(c_double)),
                                         POINTER(POINTER(c_double)), POINTER(c_int)])
_bind("n_materials", [], c_size_t, check=False)
_bind("openmc_get_material_index", [c_int32, POINTER(c_int32)])
_bind("openmc_material_get_id", [c_int32, POINTER(c_int32)])
_bind("openmc_material_get_name", [c_int32, POINTER(c_char_p)])
_bind("openmc_material_get_densities", [c_int32, POINTER(POINTER(c_int)), POINTER(POINTER(c_double)),
                                        POINTER(c_int)])
_bind("openmc_material_get_density", [c_int32, POINTER(c_double)])
_bind("cells_size", [], c_int, check=False)
_bind("openmc_get_cell_index", [c_int32, POINTER(c_int32)])
_bind("openmc_cell_get_id", [c_int32, POINTER(c_int32)])
_bind("openmc_cell_get_name", [c_int32, POINTER(c_char_p)])
_bind("openmc_cell_get_fill", [c_int32, POINTER(c_int), POINTER(POINTER(c_int32)), POINTER(c_int32)])
_bind("openmc_cell_get_num_instances", [c_int32, POINTER(c_int32)])
_bind("nuclides_size", [], c_size_t, check=False)
_bind("openmc_get_nuclide_index", [c_char_p, POINTER(c_int)])
_bind("openmc_nuclide_name", [c_int, POINTER(c_char_p)])
_bind("openmc_weight_windows_size", [], c_size_t, check=False)
_bind("openmc_weight_windows_get_bounds", [c_int32, POINTER(POINTER(c_double)), POINTER(POINTER(c_double)),
                                           POINTER(c_size_t)])
_bind("openmc_weight_windows_get_energy_bounds", [c_int32, POINTER(POINTER(c_double)), POINTER(c_size_t)])
_bind("openmc_weight_windows_get_mesh", [c_int32, POINTER(c_int32)])
_bind("calc_zn", [c_int, c_double, c_double, POINTER(c_double)], None, check=False)
_bind("calc_zn_rad", [c_int, c_double, POINTER(c_double)], None, check=False)
_bind("omc_set_option", [c_char_p, c_double])
_bind("omc_get_stats", [POINTER(c_double), c_int], c_int, check=False)
_bind("omc_stat_names", [], c_char_p, check=False)
_bind("omc_particle_hashes", [POINTER(POINTER(c_uint64)), POINTER(c_int64)])
_bind("omc_energy_grid", [POINTER(c_double)])
_bind("omc_n_nuclides", [], c_int, check=False)
_bind("omc_grid_index", [c_int], POINTER(c_int), check=False)
_bind("omc_lookup_batch", [c_int, POINTER(c_double), c_int64, c_int, POINTER(c_int), POINTER(c_double),
                           POINTER(c_uint64)])
_bind("omc_lookup_batch_noalloc", [c_int, c_void_p, c_int64, c_int, c_void_p, c_void_p, POINTER(c_uint64),
                                   c_void_p])
_bind("omc_macro_batch", [c_int, POINTER(c_double), c_int64, c_int, c_int, POINTER(c_double),
                          POINTER(c_double)])
_bind("omc_math_batch", [c_int, POINTER(c_double), POINTER(c_double), c_int64, POINTER(c_double)])
_bind("omc_copy_kernel", [c_int64, c_int, POINTER(c_double)])
_bind("omc_check_model", [c_char_p, POINTER(c_double), c_int, POINTER(c_int), c_int64], c_int, check=False)
_bind("omc_init_problems", [c_int, POINTER(c_char_p)])
_bind("omc_load_model_host", [c_char_p])
_bind("omc_n_problems", [], c_int, check=False)
_bind("omc_problem_info", [c_int, POINTER(c_int64), POINTER(c_int64), POINTER(c_int32)])
_bind("omc_problem_global_tallies", [c_int, POINTER(POINTER(c_double))])
_bind("omc_check_problem_set", [c_int, POINTER(c_char_p), POINTER(c_double), c_int], c_int, check=False)
_bind("omc_write_nuclide_hdf5", [c_char_p, c_char_p, c_double, c_double, c_int, POINTER(c_double),
                                 POINTER(c_double), POINTER(c_double), c_int64])
_bind("omc_read_nuclide", [c_char_p, c_char_p, c_double, c_int64, POINTER(c_double), POINTER(c_double),
                           POINTER(c_double), POINTER(c_double), POINTER(c_double)], c_int, check=False)
_bind("openmc_statepoint_write", [c_char_p, POINTER(c_bool)])
_bind("openmc_is_statepoint_batch", [], c_bool, check=False)
_bind("omc_problem_statepoint_write", [c_int, c_char_p, POINTER(c_bool)])
_bind("omc_h5_read", [c_char_p, c_char_p, c_char_p, c_int, c_void_p, c_size_t, POINTER(c_size_t)])
_bind("omc_learned_build", [c_int, c_int, POINTER(c_int)])
_bind("omc_learned_loss_grad", [c_int, c_int, POINTER(c_float), POINTER(c_double), POINTER(c_double)])
_bind("omc_learned_model", [c_int, POINTER(c_float), POINTER(c_int)], c_int, check=False)
_bind("omc_learned_stats", [POINTER(c_double)], c_int, check=False)
_bind("omc_learned_stats2", [POINTER(c_double)], c_int, check=False)

is_initialized = False


class _DLLGlobal:
    """Data descriptor over an exported global (openmc/lib/core.py:724-734)."""

    def __init__(self, ctype, name):
        self.ctype = ctype
        self.name = name

    def __get__(self, instance, owner):
        return self.ctype.in_dll(_dll, self.name).value

    def __set__(self, instance, value):
        self.ctype.in_dll(_dll, self.name).value = value


class _Settings:
    event_based = _DLLGlobal(c_bool, "event_based")
    particles = _DLLGlobal(c_int64, "n_particles")
    inactive = _DLLGlobal(c_int32, "n_inactive")
    generations_per_batch = _DLLGlobal(c_int32, "gen_per_batch")
    run_CE = _DLLGlobal(c_bool, "run_CE")
    verbosity = _DLLGlobal(c_int, "verbosity")
    cmfd_run = _DLLGlobal(c_bool, "cmfd_run")
    entropy_on = _DLLGlobal(c_bool, "entropy_on")
    max_lost_particles = _DLLGlobal(c_int32, "max_lost_particles")
    need_depletion_rx = _DLLGlobal(c_bool, "need_depletion_rx")
    output_summary = _DLLGlobal(c_bool, "output_summary")
    rel_max_lost_particles = _DLLGlobal(c_double, "rel_max_lost_particles")
    restart_run = _DLLGlobal(c_bool, "restart_run")
    weight_windows_on = _DLLGlobal(c_bool, "weight_windows_on")

    @property
    def path_statepoint(self):
        path = c_char_p.in_dll(_dll, "path_statepoint_c").value
        return path.decode() if path is not None else None

    @property
    def run_mode(self):
        return {1: "fixed source"}.get(c_int.in_dll(_dll, "run_mode").value)

    @property
    def seed(self):
        return _dll.openmc_get_seed()

    @seed.setter
    def seed(self, seed):
        _dll.openmc_set_seed(seed)

    @property
    def stride(self):
        return _dll.openmc_get_stride()

    @stride.setter
    def stride(self, stride):
        _dll.openmc_set_stride(stride)


settings = _Settings()


def load_model_host(directory="."):
    """Read a model directory as init() does, without a GPU: the model
    queries (cells(), materials(), meshes(), filters(), tallies) work,
    transport does not; finalize() releases it."""
    _dll.omc_load_model_host(str(directory).encode())


def init(args=None, intracomm=None, output=True):
    """Initialize from model.xml in the current directory (openmc/lib/core.py:309-350)."""
    global is_initialized
    args = ["openmc"] + (list(args) if args is not None else [])
    argv = (POINTER(c_char) * (len(args) + 1))()
    bufs = [create_string_buffer(a.encode()) for a in args]
    for i, b in enumerate(bufs):
        argv[i] = b
    if not output:
        c_int.in_dll(_dll, "verbosity").value = 1
    _dll.openmc_init(len(args), argv, None)
    is_initialized = True


def run(output=True):
    _dll.openmc_run()


def simulation_init():
    _dll.openmc_simulation_init()


def simulation_finalize():
    _dll.openmc_simulation_finalize()


def next_batch():
    status = c_int()
    _dll.openmc_next_batch(status)
    return status.value


def finalize():
    global is_initialized
    _dll.openmc_finalize()
    is_initialized = False


def reset():
    _dll.openmc_reset()


def hard_reset():
    _dll.openmc_hard_reset()


def reset_timers():
    _dll.openmc_reset_timers()


def current_batch():
    return c_int.in_dll(_dll, "current_batch").value


def num_realizations():
    return c_int32.in_dll(_dll, "n_realizations").value


@contextmanager
def run_in_memory(**kwargs):
    init(**kwargs)
    try:
        yield
    finally:
        finalize()


class Tally:
    """Read access to one tally's results (openmc/lib/tally.py)."""

    def __init__(self, tally_id):
        index = c_int32()
        _dll.openmc_get_tally_index(tally_id, index)
        self._index = index.value
        self.id = tally_id

    @classmethod
    def from_index(cls, index):
        """The tally at an index of the tallies array (ids repeat across the
        problems of a problem set; indices do not)."""
        t = cls.__new__(cls)
        t._index = int(index)
        tid = c_int32()
        _dll.openmc_tally_get_id(t._index, tid)
        t.id = tid.value
        return t

    @property
    def num_realizations(self):
        n = c_int32()
        _dll.openmc_tally_get_n_realizations(self._index, n)
        return n.value

    @property
    def filters(self):
        ptr = POINTER(c_int32)()
        n = c_size_t()
        _dll.openmc_tally_get_filters(self._index, ptr, n)
        return [_filter_from_index(ptr[i]) for i in range(n.value)]

    @property
    def scores(self):
        ptr = POINTER(c_int)()
        n = c_int()
        _dll.openmc_tally_get_scores(self._index, ptr, n)
        return [_SCORES.get(ptr[i], ptr[i]) for i in range(n.value)]

    @property
    def nuclides(self):
        ptr = POINTER(c_int)()
        n = c_int()
        _dll.openmc_tally_get_nuclides(self._index, ptr, n)
        return ["total" if ptr[i] == -1 else Nuclide(ptr[i]).name for i in range(n.value)]

    @property
    def estimator(self):
        e = c_int()
        _dll.openmc_tally_get_estimator(self._index, e)
        return {0: "analog", 1: "tracklength", 2: "collision"}[e.value]

    @property
    def type(self):
        t = c_int32()
        _dll.openmc_tally_get_type(self._index, t)
        return {0: "volume", 1: "mesh-surface", 2: "surface", 3: "pulse-height"}[t.value]

    @property
    def active(self):
        a = c_bool()
        _dll.openmc_tally_get_active(self._index, a)
        return a.value

    def reset(self):
        _dll.openmc_tally_reset(self._index)

    @property
    def results(self):
        data = POINTER(c_double)()
        shape = (c_size_t * 3)()
        _dll.openmc_tally_results(self._index, data, shape)
        return np.ctypeslib.as_array(data, tuple(shape))

    @property
    def mean(self):
        n = self.num_realizations
        sum_ = self.results[:, :, 1]
        return sum_ / n if n > 0 else sum_.copy()

    @property
    def std_dev(self):
        results = self.results
        std_dev = np.empty(results.shape[:2])
        std_dev.fill(np.inf)
        n = self.num_realizations
        if n > 1:
            sum_ = results[:, :, 1]
            sum_sq = results[:, :, 2]
            mean = sum_ / n
            nonzero = np.abs(mean) > 0
            std_dev[nonzero] = np.sqrt((sum_sq[nonzero] / n - mean[nonzero] ** 2) / (n - 1))
        return std_dev


class _TallyMapping:
    def __getitem__(self, key):
        return Tally(key)

    def __iter__(self):
        for i in range(len(self)):
            tid = c_int32()
            _dll.openmc_tally_get_id(i, tid)
            yield tid.value

    def __len__(self):
        return _dll.tallies_size()

    def keys(self):
        return list(iter(self))


tallies = _TallyMapping()

_SCORES = {-1: "flux"}


class Nuclide:
    """A loaded nuclide (openmc/lib/nuclide.py)."""

    def __init__(self, index):
        self._index = int(index)

    @property
    def name(self):
        name = c_char_p()
        _dll.openmc_nuclide_name(self._index, name)
        return name.value.decode()


def nuclides():
    return {Nuclide(i).name: Nuclide(i) for i in range(_dll.nuclides_size())}


class RegularMesh:
    """A regular mesh of the model (openmc/lib/mesh.py)."""

    def __init__(self, index):
        self._index = int(index)
        mid = c_int32()
        _dll.openmc_mesh_get_id(self._index, mid)
        self.id = mid.value

    def _params(self):
        ll, ur, w = POINTER(c_double)(), POINTER(c_double)(), POINTER(c_double)()
        n = c_int()
        _dll.openmc_regular_mesh_get_params(self._index, ll, ur, w, n)
        return [np.ctypeslib.as_array(p, (n.value,)).copy() for p in (ll, ur, w)]

    @property
    def dimension(self):
        dims = POINTER(c_int)()
        n = c_int()
        _dll.openmc_regular_mesh_get_dimension(self._index, dims, n)
        return tuple(dims[i] for i in range(n.value))

    @property
    def lower_left(self):
        return self._params()[0]

    @property
    def upper_right(self):
        return self._params()[1]

    @property
    def width(self):
        return self._params()[2]

    @property
    def n_elements(self):
        n = c_size_t()
        _dll.openmc_mesh_get_n_elements(self._index, n)
        return n.value


def meshes():
    out = {}
    for i in range(_dll.n_meshes()):
        m = RegularMesh(i)
        out[m.id] = m
    return out


class Filter:
    """A tally filter of the model (openmc/lib/filter.py)."""

    def __init__(self, index):
        self._index = int(index)
        fid = c_int32()
        _dll.openmc_filter_get_id(self._index, fid)
        self.id = fid.value

    @property
    def n_bins(self):
        n = c_int()
        _dll.openmc_filter_get_num_bins(self._index, n)
        return n.value


class MeshFilter(Filter):
    filter_type = "mesh"

    @property
    def mesh(self):
        idx = c_int32()
        _dll.openmc_mesh_filter_get_mesh(self._index, idx)
        return RegularMesh(idx.value)


class EnergyFilter(Filter):
    filter_type = "energy"

    @property
    def bins(self):
        ptr = POINTER(c_double)()
        n = c_size_t()
        _dll.openmc_energy_filter_get_bins(self._index, ptr, n)
        e = np.ctypeslib.as_array(ptr, (n.value,)).copy()
        return np.column_stack([e[:-1], e[1:]])


def _filter_from_index(index):
    buf = create_string_buffer(20)
    _dll.openmc_filter_get_type(index, buf)
    return {"mesh": MeshFilter, "energy": EnergyFilter}[buf.value.decode()](index)


def filters():
    out = {}
    for i in range(_dll.tally_filters_size()):
        f = _filter_from_index(i)
        out[f.id] = f
    return out


class Material:
    """A material of the model (openmc/lib/material.py)."""

    def __init__(self, index):
        self._index = int(index)
        mid = c_int32()
        _dll.openmc_material_get_id(self._index, mid)
        self.id = mid.value

    @property
    def name(self):
        name = c_char_p()
        _dll.openmc_material_get_name(self._index, name)
        return name.value.decode()

    @property
    def nuclides(self):
        return self._densities()[0]

    @property
    def densities(self):
        return self._densities()[1]

    def _densities(self):
        nuc, dens = POINTER(c_int)(), POINTER(c_double)()
        n = c_int()
        _dll.openmc_material_get_densities(self._index, nuc, dens, n)
        names = [Nuclide(nuc[i]).name for i in range(n.value)]
        return names, np.ctypeslib.as_array(dens, (n.value,)).copy()

    def get_density(self, units="atom/b-cm"):
        if units == "atom/b-cm":
            return float(np.sum(self.densities))
        rho = c_double()
        _dll.openmc_material_get_density(self._index, rho)
        return rho.value


def materials():
    out = {}
    for i in range(_dll.n_materials()):
        m = Material(i)
        out[m.id] = m
    return out


class Cell:
    """A cell of the model (openmc/lib/cell.py)."""

    def __init__(self, index):
        self._index = int(index)
        cid = c_int32()
        _dll.openmc_cell_get_id(self._index, cid)
        self.id = cid.value

    @property
    def name(self):
        name = c_char_p()
        _dll.openmc_cell_get_name(self._index, name)
        return name.value.decode()

    @property
    def fill(self):
        """(type, index): type 'material' (index -1 = void), 'universe' or 'lattice'."""
        t = c_int()
        ptr = POINTER(c_int32)()
        n = c_int32()
        _dll.openmc_cell_get_fill(self._index, t, ptr, n)
        return {0: "material", 1: "universe", 2: "lattice"}[t.value], ptr[0]

    @property
    def num_instances(self):
        n = c_int32()
        _dll.openmc_cell_get_num_instances(self._index, n)
        return n.value


def cells():
    out = {}
    for i in range(_dll.cells_size()):
        c = Cell(i)
        out[c.id] = c
    return out


def calc_zn(n, rho, phi):
    """Zernike moments in the order of openmc.lib.math.calc_zn."""
    zn = np.zeros(((n + 1) * (n + 2)) // 2)
    _dll.calc_zn(n, rho, phi, zn.ctypes.data_as(POINTER(c_double)))
    return zn


def calc_zn_rad(n, rho):
    zn = np.zeros(n // 2 + 1)
    _dll.calc_zn_rad(n, rho, zn.ctypes.data_as(POINTER(c_double)))
    return zn


def global_tallies():
    ptr = POINTER(c_double)()
    _dll.openmc_global_tallies(ptr)
    return np.ctypeslib.as_array(ptr, (4, 3)).copy()


# --- build-specific ---------------------------------------------------------

def set_option(name, value):
    _dll.omc_set_option(name.encode(), float(value))


def stats():
    names = _dll.omc_stat_names().decode().split(",")
    out = (c_double * len(names))()
    _dll.omc_get_stats(out, len(names))
    return dict(zip(names, list(out)))


def particle_hashes():
    ptr = POINTER(c_uint64)()
    n = c_int64()
    _dll.omc_particle_hashes(ptr, n)
    if n.value == 0:
        return np.zeros(0, dtype=np.uint64)
    return np.ctypeslib.as_array(ptr, (n.value,)).copy()


def energy_grid():
    out = (c_double * 3)()
    _dll.omc_energy_grid(out)
    return tuple(out)


def grid_index(i_nuc, n_log_bins):
    ptr = _dll.omc_grid_index(i_nuc)
    return np.ctypeslib.as_array(ptr, (n_log_bins + 1,)).copy()


def lookup_batch(i_nuc, energies, learned=False):
    E = np.ascontiguousarray(energies, dtype=np.float64)
    ig = np.zeros(E.size, dtype=np.int32)
    xs = np.zeros((E.size, 2), dtype=np.float64)
    probes = c_uint64()
    _dll.omc_lookup_batch(i_nuc, E.ctypes.data_as(POINTER(c_double)), E.size, int(learned),
                          ig.ctypes.data_as(POINTER(c_int)), xs.ctypes.data_as(POINTER(c_double)), probes)
    return ig, xs, probes.value


def macro_nuclides(material):
    """Number of nuclides of a material (by index)."""
    nuc = POINTER(c_int)()
    dens = POINTER(c_double)()
    n = c_int()
    _dll.openmc_material_get_densities(material, nuc, dens, n)
    return n.value


def macro_batch(material, energies, learned=False, repeat=1):
    E = np.ascontiguousarray(energies, dtype=np.float64)
    out = np.zeros((E.size, 2), dtype=np.float64)
    ms = c_double()
    _dll.omc_macro_batch(material, E.ctypes.data_as(POINTER(c_double)), E.size, int(learned), int(repeat),
                         out.ctypes.data_as(POINTER(c_double)), ms)
    return out, ms.value


def copy_kernel(n_bytes, reps=10):
    """Time a plain device copy of n_bytes; returns ms per copy."""
    ms = c_double()
    _dll.omc_copy_kernel(int(n_bytes), int(reps), ms)
    return ms.value


def init_problems(directories, output=True):
    """Build-specific: initialize a problem set, several independent models
    (one directory each) whose particles share one event loop.  run() and
    next_batch() then advance every problem; each keeps its own ids, seeds,
    source, tally and leakage."""
    global is_initialized
    arr = (c_char_p * len(directories))(*[os.fspath(d).encode() for d in directories])
    if not output:
        c_int.in_dll(_dll, "verbosity").value = 1
    _dll.omc_init_problems(len(directories), arr)
    is_initialized = True


def n_problems():
    return _dll.omc_n_problems()


def problem_info(problem):
    """Particles per batch, first population index and tally index (-1 if
    none) of one problem of the set."""
    n, first, tally = c_int64(), c_int64(), c_int32()
    _dll.omc_problem_info(int(problem), n, first, tally)
    return {"n_particles": n.value, "first": first.value, "tally_index": tally.value}


def problem_tally(problem):
    """The Tally of one problem of the set (None without one)."""
    idx = problem_info(problem)["tally_index"]
    return None if idx < 0 else Tally.from_index(idx)


def problem_global_tallies(problem):
    """Global tallies [4, 3] (sum, sum of squares in columns 1, 2) of one problem."""
    ptr = POINTER(c_double)()
    _dll.omc_problem_global_tallies(int(problem), ptr)
    return np.ctypeslib.as_array(ptr, (4, 3)).copy()


def check_problem_set(directories):
    """Host-only reading of a problem set (no GPU): totals and per-problem offsets."""
    arr = (c_char_p * len(directories))(*[os.fspath(d).encode() for d in directories])
    out = (c_double * (16 + 8 * len(directories)))()
    n = _dll.omc_check_problem_set(len(directories), arr, out, len(out))
    if n < 0:
        _error_handler(n, None, None)
    vals = list(out)[:n]
    keys = ["n_nuclides", "n_materials", "n_cells", "n_surfaces", "n_universes", "n_lattices", "n_levels",
            "n_particles", "n_tallies", "energy_min", "energy_max"]
    head = dict(zip(keys, vals[:len(keys)]))
    pk = ["first", "n_particles", "root_universe", "mat_begin", "surf_begin", "cell_begin", "tally", "tally_off"]
    rest = vals[len(keys):]
    parts = [dict(zip(pk, rest[8 * k:8 * k + 8])) for k in range(len(rest) // 8)]
    return head, parts


def write_nuclide_hdf5(path, name, awr, kT, temperature, energy, elastic, capture):
    """Build-specific: one nuclide (MT 2 and MT 102 at one temperature [K]) in
    the reference's HDF5 layout."""
    E = np.ascontiguousarray(energy, dtype=np.float64)
    el = np.ascontiguousarray(elastic, dtype=np.float64)
    cap = np.ascontiguousarray(capture, dtype=np.float64)
    if not (E.size == el.size == cap.size):
        raise ValueError("energy, elastic and capture need the same length")
    _dll.omc_write_nuclide_hdf5(os.fspath(path).encode(), name.encode(), float(awr), float(kT), int(temperature),
                                E.ctypes.data_as(POINTER(c_double)), el.ctypes.data_as(POINTER(c_double)),
                                cap.ctypes.data_as(POINTER(c_double)), E.size)


def statepoint_write(filename=None, write_source=True):
    """Write a statepoint file (openmc/lib/core.py:566-580).

    Parameters
    ----------
    filename : str or None
        Path to the statepoint to write. If None is passed, a default name that
        contains the current batch will be written.
    write_source : bool
        Whether or not to include the source bank in the statepoint.
    """
    if filename is not None:
        filename = c_char_p(os.fspath(filename).encode())
    _dll.openmc_statepoint_write(filename, c_bool(write_source))


def is_statepoint_batch():
    """Whether the current batch is one of the statepoint batches."""
    return bool(_dll.openmc_is_statepoint_batch())


def problem_statepoint_write(problem, filename=None, write_source=True):
    """Build-specific: the statepoint of one problem of a problem set."""
    if filename is not None:
        filename = c_char_p(os.fspath(filename).encode())
    _dll.omc_problem_statepoint_write(int(problem), filename, c_bool(write_source))


def h5_read(path, obj, attribute=None, kind="f8"):
    """Build-specific, host only: a dataset (or an attribute of `obj`) of an
    HDF5 file through the library's libhdf5, without h5py.  kind "f8" and
    "i8" give numpy arrays (flattened), "str" a list of strings."""
    k = {"f8": 0, "i8": 1, "str": 2}[kind]
    n = c_size_t(0)
    args = [os.fspath(path).encode(), obj.encode(), attribute.encode() if attribute else None, k]
    _dll.omc_h5_read(*args, None, 0, byref(n))
    if k == 2:
        buf = (c_char * max(n.value, 1))()
        _dll.omc_h5_read(*args, cast(buf, c_void_p), n.value, byref(n))
        return [w.decode() for w in bytes(buf.raw[:n.value]).split(b"\0") if w]
    out = np.zeros(n.value, dtype=np.float64 if k == 0 else np.int64)
    _dll.omc_h5_read(*args, out.ctypes.data_as(c_void_p), out.size, byref(n))
    return out


def h5_count(path, obj, attribute=None):
    """Build-specific, host only: the number of elements of a dataset or
    attribute (of any type, compound included)."""
    n = c_size_t(0)
    _dll.omc_h5_read(os.fspath(path).encode(), obj.encode(), attribute.encode() if attribute else None, 1, None, 0,
                     byref(n))
    return n.value


def read_nuclide(path, name, temperature=293.6, n_max=1 << 22):
    """Build-specific, host only: a nuclide from an HDF5 file as the library
    reads it: dict(energy, total, absorption, elastic, awr, kT)."""
    bufs = [np.zeros(n_max) for _ in range(4)]
    out2 = (c_double * 2)()
    n = _dll.omc_read_nuclide(os.fspath(path).encode(), name.encode(), float(temperature), n_max,
                              *[b.ctypes.data_as(POINTER(c_double)) for b in bufs], out2)
    if n < 0:
        _error_handler(n, None, None)
    keys = ["energy", "total", "absorption", "elastic"]
    out = {k: b[:n].copy() for k, b in zip(keys, bufs)}
    out["awr"], out["kT"] = out2[0], out2[1]
    return out


def check_model(directory, n_gi=0):
    """Parse a model and initialize its data on the host only (no GPU).

    Returns (summary dict, normalized atom densities, first n_gi grid-index entries).
    """
    out = (c_double * 4096)()
    gi = np.zeros(max(n_gi, 1), dtype=np.int32)
    n = _dll.omc_check_model(directory.encode(), out, 4096, gi.ctypes.data_as(POINTER(c_int)), n_gi)
    if n < 0:
        _error_handler(n, None, None)
    keys = ["n_nuclides", "n_materials", "n_cells", "n_surfaces", "n_lattices", "n_levels", "energy_min",
            "energy_max", "log_spacing", "n_particles", "n_batches", "n_tally_bins", "root_universe"]
    vals = list(out)[:n]
    return dict(zip(keys, vals[:len(keys)])), np.array(vals[len(keys):]), gi[:n_gi]


def math_batch(fn, x, y=None):
    """fn: 0 log, 1 sin, 2 cos, 3 sqrt, 4 x/y, 5 prn(seed bits of x), 6 u64->f64 of bits of x."""
    X = np.ascontiguousarray(x, dtype=np.float64)
    Y = None if y is None else np.ascontiguousarray(y, dtype=np.float64)
    out = np.zeros(X.size)
    _dll.omc_math_batch(fn, X.ctypes.data_as(POINTER(c_double)),
                        None if Y is None else Y.ctypes.data_as(POINTER(c_double)), X.size,
                        out.ctypes.data_as(POINTER(c_double)))
    return out


def learned_build(n_seg=256, epochs=200):
    e = c_int()
    _dll.omc_learned_build(n_seg, epochs, e)
    return e.value


def learned_loss_grad(i_nuc, params):
    p = np.ascontiguousarray(params, dtype=np.float32).reshape(-1, 2)
    n_seg = p.shape[0]
    loss = np.zeros(n_seg)
    grad = np.zeros((n_seg, 2))
    _dll.omc_learned_loss_grad(i_nuc, n_seg, p.ctypes.data_as(POINTER(c_float)),
                               loss.ctypes.data_as(POINTER(c_double)), grad.ctypes.data_as(POINTER(c_double)))
    return loss, grad


def learned_stats():
    out = (c_double * 6)()
    if _dll.omc_learned_stats(out) < 0:
        raise AllocationError("learned index has not been built")
    keys = ["n_seg", "max_err", "mean_err", "loss_init", "loss_trained", "loss_final"]
    st = dict(zip(keys, list(out)))
    out2 = (c_double * 5)()
    _dll.omc_learned_stats2(out2)
    st.update(zip(["rounds", "err_bound", "total_segments", "min_segments", "max_segments"], list(out2)))
    return st


def learned_model(i_nuc, n_seg=None):
    """(params[n][2], err[n]) of one nuclide's model; n is its own segment
    count (the base count, or a multiple where the error bound needed it)."""
    n = _dll.omc_learned_model(i_nuc, None, None)
    if n < 0:
        raise AllocationError("learned index has not been built")
    p = np.zeros((n, 2), dtype=np.float32)
    e = np.zeros(n, dtype=np.int32)
    _dll.omc_learned_model(i_nuc, p.ctypes.data_as(POINTER(c_float)), e.ctypes.data_as(POINTER(c_int)))
    return p, e
