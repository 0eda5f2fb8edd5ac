"""MI355X (gfx950) implementation of the ElasticScattering hot path.

Python mirror of the reference's operator API over the C ABI in
``include/es_abi.h`` (ctypes, no torch types):

* :class:`Simulation` — ``class Simulation`` (src/sim/Simulation.h:56-147):
  ``init_sample`` / ``compute_sigmas`` like ``InitSample`` / ``ComputeSigmas``;
  its ``set_impurities`` / ``generate_impurities`` are the two ``Grid``
  constructors (src/sim/Grid.h:93-94), binned on the GPU and kept there.
* :class:`Batch` — the sample loop of SimulationRunner::RunSimulation
  (src/SimulationRunner.cpp:88-98) as one launch per stage over many samples
  of uneven size, optionally replayed from a captured hipGraph, with a tuner
  for the trace launch shape.

The HIP library ``libes_hip.so`` is built in-tree by
``python -m elasticscattering_amd.build``; importing the API without it raises.
"""
import ctypes as C
import os

import numpy as np

__all__ = [
    "EsError", "GridInfo", "GridDesc", "Settings", "Sigma", "Metrics", "SolverDesc",
    "BatchSample", "BatchShape", "Simulation", "Batch", "load_library", "parse_config", "run", "LIB_PATH",
    "EXPORTED_SYMBOLS",
]

PKG_DIR = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG_DIR, "libes_hip.so")

ES_OK = 0
ES_ERR_INVALID_ARG = -1
ES_ERR_HIP = -2
ES_ERR_NO_GRID = -3
ES_ERR_NO_SAMPLE = -4
ES_ERR_OOM = -5
ES_ERR_OUT_OF_GRID = -6
ES_ERR_NO_DEVICE = -7
ES_NOALLOC_MAX_IMPURITIES = 131072  # es_grid_bin_noalloc scratch limits (include/es_abi.h)
ES_NOALLOC_MAX_CELLS = 65536
ES_PRECISION_FP64 = 0
ES_PRECISION_MIXED = 1
_PRECISIONS = {"fp64": ES_PRECISION_FP64, "mixed": ES_PRECISION_MIXED}


class EsError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"es error {code}: {msg}")
        self.code = code


class GridInfo(C.Structure):
    _fields_ = [
        ("region_size", C.c_double),
        ("region_extends", C.c_double),
        ("cell_size", C.c_double),
        ("indexed_impurity_count", C.c_int64),
        ("unique_impurity_count", C.c_int64),
        ("index_size", C.c_int32),
        ("cells_per_row", C.c_int32),
    ]


class GridDesc(C.Structure):
    _fields_ = [
        ("region_size", C.c_double),
        ("region_extends", C.c_double),
        ("impurity_radius", C.c_double),
        ("impurity_density", C.c_double),
        ("target_cell_population", C.c_int32),
        ("cells_per_row", C.c_int32),
    ]


class Settings(C.Structure):
    _fields_ = [
        ("tau", C.c_double),
        ("alpha", C.c_double),
        ("is_clockwise", C.c_int32),
        ("particle_speed", C.c_double),
        ("incoherent_tau", C.c_double),
    ]


class Sigma(C.Structure):
    _fields_ = [("xx", C.c_double), ("xy", C.c_double)]


class Metrics(C.Structure):
    _fields_ = [
        ("cells_passed", C.c_int64),
        ("impurities_tested", C.c_int64),
        ("particles_inside_impurity", C.c_int64),
        ("particles_escaped", C.c_int64),
        ("particles_at_bound", C.c_int64),
        ("max_impurities_tested", C.c_int64),
        ("max_cells_passed", C.c_int64),
        ("trace_cap_hits", C.c_int64),
        ("real_particles", C.c_int64),
        ("avg_particle_lifetime", C.c_double),
        ("time_elapsed_lifetimes", C.c_double),
        ("time_elapsed_temperatures", C.c_double),
        ("traced_particles", C.c_int64),
        ("edge_hits", C.c_int64),
        ("cross_calls", C.c_int64),
        ("sigma_evals", C.c_int64),
        ("cross_exact", C.c_int64),
        ("fp64_recomputed", C.c_int64),
    ]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


_METRICS_DTYPE = np.ctypeslib.as_array((Metrics * 1)()).dtype


class SolverDesc(C.Structure):
    _fields_ = [
        ("positions_per_row", C.c_int32),
        ("particles_per_quadrant", C.c_int32),
        ("grid_info", GridInfo),
        ("precision", C.c_int32),
        ("device", C.c_int32),
    ]


class BatchSample(C.Structure):
    _fields_ = [
        ("seed", C.c_uint64),
        ("impurity_count", C.c_int64),
        ("grid", GridDesc),
        ("grid_cell_size", C.c_double),
        ("positions_per_row", C.c_int32),
        ("particles_per_quadrant", C.c_int32),
        ("settings", Settings),
        ("coherent", C.c_int32),
        ("share_field_of", C.c_int32),
        ("fields", C.POINTER(C.c_double)),
        ("n_fields", C.c_int32),
        ("temperatures", C.POINTER(C.c_double)),
        ("n_temperatures", C.c_int32),
    ]


class BatchShape(C.Structure):
    _fields_ = [("waves_per_eu", C.c_int32), ("chunk", C.c_int32), ("grid_mult", C.c_int32),
                ("cooperative", C.c_int32)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


_PD = C.POINTER(C.c_double)
_P = C.c_void_p

# name -> (restype, argtypes); every symbol include/es_abi.h declares.
_SIGNATURES = {
    "es_solver_create": (C.c_int, [C.POINTER(SolverDesc), C.POINTER(_P)]),
    "es_solver_destroy": (None, [_P]),
    "es_last_error": (C.c_char_p, [_P]),
    "es_grid_set_impurities": (C.c_int, [_P, _PD, _PD, C.c_int64, C.POINTER(GridDesc), C.POINTER(GridInfo)]),
    "es_grid_generate": (C.c_int, [_P, C.c_uint64, C.c_int64, C.POINTER(GridDesc), C.POINTER(GridInfo)]),
    "es_grid_read": (C.c_int, [_P, _PD, C.POINTER(C.c_int32)]),
    "es_grid_set_binned": (C.c_int, [_P, _PD, C.POINTER(C.c_int32), C.c_int32, C.c_double, C.c_double, C.c_double,
                                     C.POINTER(GridInfo)]),
    "es_solver_init_sample": (C.c_int, [_P, C.POINTER(Settings), C.c_int32]),
    "es_solver_compute_sigmas": (C.c_int, [_P, C.c_double, _PD, C.c_int32, C.POINTER(Sigma), C.POINTER(Metrics), _PD]),
    "es_solver_compute_sigmas_sweep": (C.c_int, [_P, _PD, C.c_int32, _PD, C.c_int32, C.POINTER(Sigma),
                                                 C.POINTER(Metrics)]),
    "es_solver_compute_sigmas_with_images": (C.c_int, [_P, C.c_double, _PD, C.c_int32, C.POINTER(Sigma),
                                                       C.POINTER(Metrics), _PD]),
    "es_batch_create": (C.c_int, [C.c_int32, C.POINTER(_P)]),
    "es_batch_destroy": (None, [_P]),
    "es_batch_last_error": (C.c_char_p, [_P]),
    "es_batch_add_sample": (C.c_int, [_P, C.POINTER(BatchSample), C.POINTER(C.c_int32)]),
    "es_batch_clear": (C.c_int, [_P]),
    "es_batch_reset": (C.c_int, [_P]),
    "es_batch_set_shape": (C.c_int, [_P, C.POINTER(BatchShape)]),
    "es_batch_get_shape": (C.c_int, [_P, C.POINTER(BatchShape)]),
    "es_batch_set_capture": (C.c_int, [_P, C.c_int32]),
    "es_batch_set_precision": (C.c_int, [_P, C.c_int32]),
    "es_batch_run": (C.c_int, [_P, C.POINTER(Sigma), C.POINTER(Metrics), _PD]),
    "es_batch_prepare": (C.c_int, [_P]),
    "es_batch_launch": (C.c_int, [_P]),
    "es_batch_collect": (C.c_int, [_P, C.POINTER(Sigma), C.POINTER(Metrics), _PD]),
    "es_batch_counts": (C.c_int, [_P, C.POINTER(C.c_int32), C.POINTER(C.c_int32), C.POINTER(C.c_int32),
                                  C.POINTER(C.c_int64)]),
    "es_batch_tune": (C.c_int, [_P, C.POINTER(BatchShape), _PD]),
    "es_batch_launches": (C.c_int, [_P, C.POINTER(C.c_int32), C.POINTER(C.c_int32), C.POINTER(C.c_int32)]),
    "es_version": (C.c_char_p, []),
    "es_math_eval": (C.c_int, [C.c_int32, C.c_int32, _PD, _PD, C.c_int64]),
    "es_grid_bin_noalloc": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int64, C.POINTER(GridDesc), C.c_void_p, C.c_void_p,
                                      C.c_void_p, C.POINTER(GridInfo), C.c_void_p]),
    # include/es_surrogate.h
    "es_mlp_create": (C.c_int, [C.c_void_p, C.POINTER(_P)]),
    "es_mlp_destroy": (None, [_P]),
    "es_mlp_last_error": (C.c_char_p, [_P]),
    "es_mlp_members": (C.c_int, [_P]),
    "es_mlp_get_params": (C.c_int, [_P, C.POINTER(C.c_float)]),
    "es_mlp_set_params": (C.c_int, [_P, C.POINTER(C.c_float)]),
    "es_mlp_set_data": (C.c_int, [_P, C.POINTER(C.c_float), C.POINTER(C.c_float), C.c_int64]),
    "es_mlp_gradients": (C.c_int, [_P, C.POINTER(C.c_int32), C.c_int32, C.POINTER(C.c_float),
                                   C.POINTER(C.c_float), C.POINTER(C.c_float)]),
    "es_mlp_train": (C.c_int, [_P, C.c_void_p, C.POINTER(C.c_float)]),
    "es_mlp_predict": (C.c_int, [_P, C.POINTER(C.c_float), C.c_int64, C.POINTER(C.c_float)]),
    "es_mlp_last_train_time": (C.c_int, [_P, C.POINTER(C.c_double), C.POINTER(C.c_int64)]),
    # include/es_runner.h
    "es_config_parse": (C.c_int, [C.c_char_p, C.c_char_p, C.c_int]),
    "es_run": (C.c_int, [C.c_char_p, C.c_int, C.c_char_p, C.c_int]),
    "es_log_images": (C.c_int, [C.c_char_p, C.c_int, _PD, _PD, _PD]),
}
EXPORTED_SYMBOLS = tuple(_SIGNATURES)

_lib = None


def load_library(path=None):
    """Loads libes_hip.so (raises if it has not been built); ES_LIBRARY overrides the path."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = path or os.environ.get("ES_LIBRARY") or LIB_PATH
    if not os.path.exists(p):
        raise ImportError(f"{p} is missing: build it with `python -m elasticscattering_amd.build`")
    lib = C.CDLL(p)
    for name, (res, args) in _SIGNATURES.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    if path is None:
        _lib = lib
    return lib


def parse_config(path):
    """The .config file as the reference parses it (SimulationConfiguration.h:77-229), as a dict."""
    import json

    lib = load_library()
    buf = C.create_string_buffer(1 << 16)
    rc = lib.es_config_parse(str(path).encode(), buf, len(buf))
    d = json.loads(buf.value.decode())
    if rc != ES_OK:
        raise EsError(rc, d.get("error", "parse failed"))
    return d


def run(config_path, verbose=False):
    """SimulationRunner::Run (SimulationRunner.cpp:32-52) on GPU 0; returns the output directory."""
    lib = load_library()
    buf = C.create_string_buffer(4096)
    rc = lib.es_run(str(config_path).encode(), int(bool(verbose)), buf, len(buf))
    if rc != ES_OK:
        raise EsError(rc, "es_run failed (see stderr with verbose=True)")
    return buf.value.decode()


def log_images(path, lifetimes, sigma_xx, sigma_xy):
    """Logger::LogImages (Logger.cpp:320-334): three dim x dim images side by side as an RGB PNG."""
    lt, xx, xy = (np.ascontiguousarray(a, dtype=np.float64) for a in (lifetimes, sigma_xx, sigma_xy))
    dim = lt.shape[0]
    rc = load_library().es_log_images(str(path).encode(), dim, _dptr(lt), _dptr(xx), _dptr(xy))
    if rc != ES_OK:
        raise EsError(rc, f"cannot write {path}")


MATH_FUNCTIONS = {"sin": 0, "cos": 1, "asin": 2, "exp": 3}  # ES_MATH_* of include/es_abi.h


def math_eval(fn, x, device=0):
    """The device's correctly rounded sin / cos / asin / exp (es_math.h) applied to x."""
    x = np.ascontiguousarray(x, dtype=np.float64)
    y = np.empty_like(x)
    lib = load_library()
    rc = lib.es_math_eval(int(device), MATH_FUNCTIONS[fn], _dptr(x), _dptr(y), x.size)
    if rc != ES_OK:
        raise EsError(rc, lib.es_last_error(None).decode())
    return y


def _dptr(a):
    return a.ctypes.data_as(_PD) if a is not None else None


class Simulation:
    """``class Simulation`` (Simulation.h:56-147) on one GPU.

    ``Simulation(positions_per_row, particles_per_quadrant, grid_info)`` as in
    Simulation.h:131; ``grid_info`` is the GridInformation of sample 0's grid
    (region_size, region_extends, cell_size); when omitted it is taken from the
    first grid set on the solver, as SimulationRunner.cpp:75-84 does.
    """

    def __init__(self, positions_per_row, particles_per_quadrant, grid_info=None, device=0, precision="fp64"):
        self.lib = load_library()
        self.ppr = int(positions_per_row)
        self.ppq = int(particles_per_quadrant)
        if grid_info is None:
            gi = GridInfo()  # taken from the first grid set on this solver
        elif isinstance(grid_info, GridInfo):
            gi = grid_info
        else:
            gi = GridInfo(**grid_info)
        desc = SolverDesc(self.ppr, self.ppq, gi, _PRECISIONS.get(precision, precision), int(device))
        h = _P()
        rc = self.lib.es_solver_create(C.byref(desc), C.byref(h))
        if rc != ES_OK:
            raise EsError(rc, self.lib.es_last_error(None).decode())
        self.h = h
        self.grid_info = gi

    @property
    def total_particles(self):
        return self.ppr * self.ppr * 4 * self.ppq

    def _check(self, rc):
        if rc != ES_OK:
            raise EsError(rc, self.lib.es_last_error(self.h).decode())

    # Grid ------------------------------------------------------------
    def set_impurities(self, x, y, region_size, region_extends, impurity_radius, target_cell_population=15,
                       cells_per_row=0):
        """Grid(std::vector<v2>, ...) (Grid.cpp:48-59), binned on the GPU."""
        x = np.ascontiguousarray(x, dtype=np.float64)
        y = np.ascontiguousarray(y, dtype=np.float64)
        d = GridDesc(region_size, region_extends, impurity_radius, 0.0, int(target_cell_population),
                     int(cells_per_row))
        info = GridInfo()
        self._check(self.lib.es_grid_set_impurities(self.h, _dptr(x), _dptr(y), len(x), C.byref(d), C.byref(info)))
        return info

    def generate_impurities(self, seed, region_size, region_extends, impurity_radius, density=0.0, count=0,
                            target_cell_population=15, cells_per_row=0):
        """Grid(seed, ...) (Grid.cpp:28-46) with the es_rng.h generator, on the GPU."""
        d = GridDesc(region_size, region_extends, impurity_radius, density, int(target_cell_population),
                     int(cells_per_row))
        info = GridInfo()
        self._check(self.lib.es_grid_generate(self.h, int(seed), int(count), C.byref(d), C.byref(info)))
        return info

    def set_binned(self, ordered_xy, imp_index, cells_per_row, spawn_region_start, spawn_region_size,
                   impurity_radius):
        """The reference Grid's own CSR as given: GetImpurities() [Ni, 2] (cell-ordered, duplicates),
        GetIndex() [cpr^2] inclusive ends and GetSettings() (Grid.h:74-91); nothing is binned again."""
        xy = np.ascontiguousarray(ordered_xy, dtype=np.float64).reshape(-1)
        idx = np.ascontiguousarray(imp_index, dtype=np.int32)
        info = GridInfo()
        self._check(self.lib.es_grid_set_binned(self.h, _dptr(xy) if xy.size else None,
                                                idx.ctypes.data_as(C.POINTER(C.c_int32)), int(cells_per_row),
                                                float(spawn_region_start), float(spawn_region_size),
                                                float(impurity_radius), C.byref(info)))
        return info

    def read_grid(self, info):
        """(ordered impurities [Ni, 2], inclusive end index [cpr^2]) like Grid::GetImpurities/GetIndex."""
        xy = np.empty(2 * max(info.indexed_impurity_count, 1), dtype=np.float64)
        idx = np.empty(info.cells_per_row * info.cells_per_row, dtype=np.int32)
        self._check(self.lib.es_grid_read(self.h, _dptr(xy), idx.ctypes.data_as(C.POINTER(C.c_int32))))
        return xy[: 2 * info.indexed_impurity_count].reshape(-1, 2), idx

    # Simulation ------------------------------------------------------
    def init_sample(self, tau, alpha, is_clockwise, particle_speed, coherent, incoherent_tau=0.0):
        """Simulation::InitSample (Simulation.h:111-126)."""
        s = Settings(tau, alpha, int(bool(is_clockwise)), particle_speed, incoherent_tau)
        self._check(self.lib.es_solver_init_sample(self.h, C.byref(s), int(bool(coherent))))

    def compute_sigmas(self, B, temperatures, want_lifetimes=False):
        """Simulation::ComputeSigmas (Simulation.h:128): ([(xx, xy)] per T, metrics dict, lifetimes|None)."""
        T = np.ascontiguousarray(temperatures, dtype=np.float64)
        out = (Sigma * max(len(T), 1))()
        m = Metrics()
        lts = np.empty(self.total_particles, dtype=np.float64) if want_lifetimes else None
        self._check(self.lib.es_solver_compute_sigmas(self.h, float(B), _dptr(T), len(T), out, C.byref(m),
                                                      _dptr(lts)))
        return [(out[i].xx, out[i].xy) for i in range(len(T))], m.as_dict(), lts

    def compute_sigmas_with_images(self, B, temperatures):
        """Simulation::ComputeSigmasWithImages (Simulation.h:129): ([(xx, xy)] per T, metrics dict,
        images [nT][3][ppr][ppr] = phi-integrated lifetime, sigma_xx, sigma_xy per position)."""
        T = np.ascontiguousarray(temperatures, dtype=np.float64)
        out = (Sigma * max(len(T), 1))()
        m = Metrics()
        img = np.empty((len(T), 3, self.ppr, self.ppr), dtype=np.float64)
        self._check(self.lib.es_solver_compute_sigmas_with_images(self.h, float(B), _dptr(T), len(T), out,
                                                                  C.byref(m), _dptr(img)))
        return [(out[i].xx, out[i].xy) for i in range(len(T))], m.as_dict(), img

    def compute_sweep(self, fields, temperatures):
        """All fields in one launch: (sigma [nB, nT, 2], metrics per field).

        The metrics are a numpy record array with the es_metrics fields
        (``ms[b]["cells_passed"]``, or a column ``ms["cells_passed"]``)."""
        Bs = np.ascontiguousarray(fields, dtype=np.float64)
        T = np.ascontiguousarray(temperatures, dtype=np.float64)
        out = (Sigma * max(len(Bs) * len(T), 1))()
        ms = (Metrics * len(Bs))()
        self._check(self.lib.es_solver_compute_sigmas_sweep(self.h, _dptr(Bs), len(Bs), _dptr(T), len(T), out, ms))
        sig = np.frombuffer(out, dtype=np.float64, count=2 * len(Bs) * len(T)).reshape(len(Bs), len(T), 2).copy()
        return sig, np.frombuffer(ms, dtype=_METRICS_DTYPE).copy()

    def close(self):
        if getattr(self, "h", None):
            self.lib.es_solver_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


class Batch:
    """Many samples of uneven size traced in one launch per stage.

    Each :meth:`add` is one pass (coherent or incoherent) of one field: its
    impurities are drawn with the es_rng.h generator and binned on the GPU
    (Grid.cpp:28-59), or taken from an earlier entry (``share_field_of``).
    :meth:`run` returns, per entry, sigma [n_fields, n_temperatures, 2] — the
    values ``Simulation.compute_sweep`` gives for the same sample.
    """

    def __init__(self, device=0):
        self.lib = load_library()
        h = _P()
        rc = self.lib.es_batch_create(int(device), C.byref(h))
        if rc != ES_OK:
            raise EsError(rc, self.lib.es_batch_last_error(None).decode())
        self.h = h
        self.entries = []  # (n_fields, n_temperatures)

    def _check(self, rc):
        if rc != ES_OK:
            raise EsError(rc, self.lib.es_batch_last_error(self.h).decode())

    def add(self, seed, region_size, region_extends, impurity_radius, positions_per_row, particles_per_quadrant,
            fields, temperatures, tau, alpha, is_clockwise, particle_speed, coherent, incoherent_tau=0.0,
            density=0.0, count=0, target_cell_population=15, cells_per_row=0, grid_cell_size=0.0,
            share_field_of=-1):
        """Adds one sample pass; returns its entry index."""
        Bs = np.ascontiguousarray(fields, dtype=np.float64)
        T = np.ascontiguousarray(temperatures, dtype=np.float64)
        s = BatchSample()
        s.seed = int(seed)
        s.impurity_count = int(count)
        s.grid = GridDesc(region_size, region_extends, impurity_radius, density, int(target_cell_population),
                          int(cells_per_row))
        s.grid_cell_size = float(grid_cell_size)
        s.positions_per_row = int(positions_per_row)
        s.particles_per_quadrant = int(particles_per_quadrant)
        s.settings = Settings(tau, alpha, int(bool(is_clockwise)), particle_speed, incoherent_tau)
        s.coherent = int(bool(coherent))
        s.share_field_of = int(share_field_of)
        s.fields = _dptr(Bs)
        s.n_fields = len(Bs)
        s.temperatures = _dptr(T) if len(T) else None
        s.n_temperatures = len(T)
        idx = C.c_int32(-1)
        self._check(self.lib.es_batch_add_sample(self.h, C.byref(s), C.byref(idx)))
        self.entries.append((len(Bs), len(T)))
        return idx.value

    def clear(self):
        """Removes every entry; keeps the drawn fields for later entries with the same recipe
        (seed, count, grid description), which then skip drawing and binning."""
        self._check(self.lib.es_batch_clear(self.h))
        self.entries = []

    def reset(self):
        """Removes every entry and forgets the drawn fields: later entries draw and bin
        their fields again (into the memory the batch already holds)."""
        self._check(self.lib.es_batch_reset(self.h))
        self.entries = []

    def counts(self):
        ns, npr, no = C.c_int32(), C.c_int32(), C.c_int32()
        ni = C.c_int64()
        self._check(self.lib.es_batch_counts(self.h, C.byref(ns), C.byref(npr), C.byref(no), C.byref(ni)))
        return {"samples": ns.value, "pairs": npr.value, "outputs": no.value, "traced_items": ni.value}

    @property
    def shape(self):
        sh = BatchShape()
        self._check(self.lib.es_batch_get_shape(self.h, C.byref(sh)))
        return sh.as_dict()

    def set_shape(self, waves_per_eu=3, chunk=256, grid_mult=1, cooperative=-1):
        """Trace launch shape.  cooperative: 0 crossing disks solved per lane, 1 shared by the
        wave, 2 shared and screened, -1 (default) 2 or 1 by the fields' impurities per cell."""
        sh = BatchShape(int(waves_per_eu), int(chunk), int(grid_mult), int(cooperative))
        self._check(self.lib.es_batch_set_shape(self.h, C.byref(sh)))

    def set_capture(self, enable):
        self._check(self.lib.es_batch_set_capture(self.h, int(bool(enable))))

    def set_precision(self, precision):
        """"fp64" (default) or "mixed" (FP32 trace with FP64 recompute of undecidable cases)."""
        self._check(self.lib.es_batch_set_precision(self.h, _PRECISIONS.get(precision, precision)))

    def tune(self):
        """Times every trace launch shape on this batch; keeps and returns the fastest (shape, seconds)."""
        sh = BatchShape()
        t = C.c_double()
        self._check(self.lib.es_batch_tune(self.h, C.byref(sh), C.byref(t)))
        return sh.as_dict(), t.value

    def launches(self):
        """{"segments", "kernels" (kernel launches per run), "graph_launches" (1 when captured)}."""
        sg, kn, gl = C.c_int32(), C.c_int32(), C.c_int32()
        self._check(self.lib.es_batch_launches(self.h, C.byref(sg), C.byref(kn), C.byref(gl)))
        return {"segments": sg.value, "kernels": kn.value, "graph_launches": gl.value}

    def run(self, metrics=True):
        """([sigma [nB, nT, 2] per entry], [metrics record array [nB] per entry] | None, device seconds).

        The metrics of an entry are a numpy record array with the es_metrics fields
        (``mets[e][b]["cells_passed"]``)."""
        return self._results(self.lib.es_batch_run, metrics)

    def prepare(self):
        """Enqueues the binning of the fields added since the last run, and their read-backs,
        without waiting (es_batch_prepare): launch() then only waits for the read-backs."""
        self._check(self.lib.es_batch_prepare(self.h))

    def launch(self):
        """Enqueues the run and returns; :meth:`collect` waits for it (the host may prepare
        another batch meanwhile)."""
        self._check(self.lib.es_batch_launch(self.h))

    def collect(self, metrics=True):
        """The launched run's results, as :meth:`run` returns them."""
        return self._results(self.lib.es_batch_collect, metrics)

    def _results(self, fn, metrics):
        nout = sum(b * t for b, t in self.entries)
        npairs = sum(b for b, _ in self.entries)
        out = (Sigma * max(nout, 1))()
        ms = (Metrics * max(npairs, 1))() if metrics else None
        secs = C.c_double()
        self._check(fn(self.h, out, ms, C.byref(secs)))
        flat = np.frombuffer(out, dtype=np.float64, count=2 * nout) if nout else np.empty(0)
        recs = np.frombuffer(ms, dtype=_METRICS_DTYPE, count=npairs).copy() if metrics else None
        sig, mets, o, p = [], [] if metrics else None, 0, 0
        for nB, nT in self.entries:
            sig.append(flat[2 * o: 2 * (o + nB * nT)].reshape(nB, nT, 2).copy())
            o += nB * nT
            if metrics:
                mets.append(recs[p:p + nB])
            p += nB
        return sig, mets, secs.value

    def close(self):
        if getattr(self, "h", None):
            self.lib.es_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()
