"""ctypes access to the CPU oracle (oracle/es_oracle.c) for the tests.

Test infrastructure: this module and the oracle it loads are the checker, never
the thing measured or shipped.  Five libraries:

* ``esm``  — oracle built on the project's correctly rounded es_math.h (the
  GPU's math);
* ``libm`` — the same restatement on glibc's sin/cos/asin/exp;
* ``esm_libm_asin`` — es_math.h except glibc's asin (attributes the
  differences between the two builds above to one function);
* ``esm_sel`` — the esm build that also checks the GPU trace's choice of
  which crossing points need their angle (orc_selection_stats);
* ``ref``  — the reference's own kernel headers compiled from the reference
  tree (oracle/_ref/libesref.so, only where that tree was present at build).
"""
import ctypes as C
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_DIR = os.path.join(ROOT, "oracle")

_D = C.c_double
_I = C.c_int
_I64 = C.c_int64
_PD = C.POINTER(C.c_double)
_PI = C.POINTER(C.c_int)
_PI32 = C.POINTER(C.c_int32)


class OrcMetrics(C.Structure):
    _fields_ = [
        ("cells_passed", _I64),
        ("impurities_tested", _I64),
        ("particles_inside_impurity", _I64),
        ("particles_escaped", _I64),
        ("particles_at_bound", _I64),
        ("max_impurities_tested", _I64),
        ("max_cells_passed", _I64),
        ("trace_cap_hits", _I64),
        ("edge_hits", _I64),
        ("cross_calls", _I64),
    ]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class OrcRunMetrics(C.Structure):
    _fields_ = [("pm", OrcMetrics), ("real_particles", _I64), ("avg_particle_lifetime", _D)]


class OrcSimParams(C.Structure):
    _fields_ = [
        ("positions_per_row", C.c_int32),
        ("particles_per_quadrant", C.c_int32),
        ("region_size", _D),
        ("region_extends", _D),
        ("cell_size", _D),
        ("tau", _D),
        ("alpha", _D),
        ("is_clockwise", C.c_int32),
        ("particle_speed", _D),
        ("incoherent_tau", _D),
        ("nthreads", C.c_int32),
    ]


class OrcSigma(C.Structure):
    _fields_ = [("xx_naive", _D), ("xy_naive", _D), ("xx_comp", _D), ("xy_comp", _D)]


_SIGS = {
    "smod": (_D, [_D, _D]),
    "get_angle": (_D, [_D, _D, _D, _D, _D, _I]),
    "angle_in_range": (_I, [_D, _D, _D, _I]),
    "cross_angle": (_D, [_D, _D, _I]),
    "circles_cross": (_I, [_D, _D, _D, _D, _D, _D]),
    "cross_points": (None, [_D, _D, _D, _D, _D, _D, _PD]),
    "first_cross_time": (_D, [_D, _D, _D, _I, _D, _D, _D, _D, _D, _D, _D]),
    "inside_impurity": (_I, [_D, _D, _D, _D, _D]),
    "bound_time": (_D, [_D, _D, _D, _I, _I, _I]),
    "bound_angle": (_D, [_D, _D, _I]),
    "orbit_center": (None, [_D, _D, _D, _D, _D, _D, _I, _PD]),
    "get_cell": (None, [_D, _D, _D, _D, _I, _PI]),
    "to_world": (None, [_I, _I, _D, _D, _PD]),
    "different_point": (_I, [_D, _D, _D, _D, _D]),
    "within_bounds": (_I, [_I, _I, _I]),
    "get_index": (_I, [_I, _I, _I]),
    "next_cell": (_I, [_D, _D, _D, _I, _D, _D, _D, _D, _I, _PD, _PI, _PD, _PI]),
    "simpson_weight": (_D, [_I, _I]),
    "get_sigma": (_D, [_D, _D, _D, _D]),
}

_ORC_ONLY = {
    "m_sin": (_D, [_D]),
    "m_cos": (_D, [_D]),
    "m_asin": (_D, [_D]),
    "m_exp": (_D, [_D]),
    "grid_create": (C.c_void_p, [_PD, _PD, _I64, _D, _D, _D, _I, _I]),
    "grid_generate": (C.c_void_p, [C.c_uint64, _D, _D, _D, _I64, _D, _I, _I]),
    "grid_free": (None, [C.c_void_p]),
    "grid_info": (None, [C.c_void_p, C.POINTER(_I64), _PD]),
    "grid_copy": (None, [C.c_void_p, _PD, _PI32]),
    "generate_impurities": (None, [C.c_uint64, _I64, _D, _D, _PD, _PD]),
    "compute_sigmas": (_I, [C.c_void_p, C.POINTER(OrcSimParams), _I, _D, _PD, _I,
                            C.POINTER(OrcSigma), C.POINTER(OrcRunMetrics), _PD]),
    "compute_lifetimes": (_I, [C.c_void_p, C.POINTER(OrcSimParams), _I, _D, _PD, C.POINTER(OrcMetrics)]),
    "derive_temperatures": (_I, [C.c_void_p, C.POINTER(OrcSimParams), _I, _D, _PD, _PD, _I,
                                 C.POINTER(OrcSigma)]),
    "derive_images": (_I, [C.c_void_p, C.POINTER(OrcSimParams), _I, _D, _PD, _D, _PD]),
    "particle_lifetime": (_D, [C.c_void_p, _I, _I, _D, _D, _D, _D, _D, _I, _I, _D, _D,
                               C.POINTER(OrcMetrics)]),
}

_cache = {}


def lib_path(kind):
    return {
        "esm": os.path.join(ORACLE_DIR, "liboracle_esm.so"),
        "libm": os.path.join(ORACLE_DIR, "liboracle_libm.so"),
        "esm_libm_asin": os.path.join(ORACLE_DIR, "liboracle_esm_libm_asin.so"),
        "esm_sel": os.path.join(ORACLE_DIR, "liboracle_esm_sel.so"),
        "ref": os.path.join(ORACLE_DIR, "_ref", "libesref.so"),
    }[kind]


def load(kind="esm"):
    """Load one oracle build; returns an object with the flat functions."""
    if kind in _cache:
        return _cache[kind]
    path = lib_path(kind)
    if not os.path.exists(path):
        return None
    lib = C.CDLL(path)
    prefix = "ref_" if kind == "ref" else "orc_"
    ns = type("OracleLib", (), {})()
    ns.kind = kind
    ns.lib = lib
    sigs = dict(_SIGS)
    if kind != "ref":
        sigs.update(_ORC_ONLY)
    for name, (res, args) in sigs.items():
        f = getattr(lib, prefix + name)
        f.restype = res
        f.argtypes = args
        setattr(ns, name, f)
    _cache[kind] = ns
    return ns


def dptr(a):
    return a.ctypes.data_as(_PD)


class OracleGrid:
    """An impurity grid binned by the oracle's restatement of Grid.cpp."""

    def __init__(self, lib, handle):
        self.lib = lib
        self.h = handle
        ints = (_I64 * 3)()
        dbl = (_D * 3)()
        lib.grid_info(handle, ints, dbl)
        self.cells_per_row = int(ints[0])
        self.unique_impurity_count = int(ints[1])
        self.total_indexed = int(ints[2])
        self.spawn_lo, self.spawn_hi, self.cell_size = float(dbl[0]), float(dbl[1]), float(dbl[2])

    @classmethod
    def from_points(cls, lib, x, y, region_size, region_extends, radius, target=15, cells_per_row=0):
        x = np.ascontiguousarray(x, dtype=np.float64)
        y = np.ascontiguousarray(y, dtype=np.float64)
        h = lib.grid_create(dptr(x), dptr(y), len(x), region_size, region_extends, radius, int(target),
                            int(cells_per_row))
        if not h:
            raise MemoryError("oracle grid_create failed")
        return cls(lib, h)

    @classmethod
    def generate(cls, lib, seed, region_size, region_extends, density, radius, target=15, cells_per_row=0, n=0):
        h = lib.grid_generate(seed, region_size, region_extends, density, int(n), radius, int(target),
                              int(cells_per_row))
        if not h:
            raise ValueError("oracle grid_generate rejected the settings")
        return cls(lib, h)

    def csr(self):
        xy = np.empty(2 * max(self.total_indexed, 1), dtype=np.float64)
        idx = np.empty(self.cells_per_row * self.cells_per_row, dtype=np.int32)
        self.lib.grid_copy(self.h, dptr(xy), idx.ctypes.data_as(_PI32))
        return xy[: 2 * self.total_indexed].reshape(-1, 2), idx

    def close(self):
        if self.h:
            self.lib.grid_free(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def sim_params(ppr, ppq, region_size, region_extends, cell_size, tau, alpha, clockwise, speed,
               incoherent_tau=0.0, nthreads=1):
    return OrcSimParams(ppr, ppq, region_size, region_extends, cell_size, tau, alpha, int(clockwise), speed,
                        incoherent_tau, nthreads)


def compute_sigmas(lib, grid, params, coherent, B, temperatures, want_lifetimes=False):
    """SimulationCPU::ComputeSigmas for one pass; returns (sigmas, metrics, lifetimes|None)."""
    T = np.ascontiguousarray(temperatures, dtype=np.float64)
    out = (OrcSigma * max(len(T), 1))()
    m = OrcRunMetrics()
    P = params.positions_per_row ** 2 * 4 * params.particles_per_quadrant
    lts = np.empty(P, dtype=np.float64) if want_lifetimes else None
    rc = lib.compute_sigmas(grid.h, C.byref(params), int(coherent), B, dptr(T), len(T), out, C.byref(m),
                            dptr(lts) if lts is not None else None)
    if rc != 0:
        raise RuntimeError("oracle compute_sigmas failed")
    return [out[i] for i in range(len(T))], m, lts


def compute_lifetimes(lib, grid, params, coherent, B):
    P = params.positions_per_row ** 2 * 4 * params.particles_per_quadrant
    lts = np.empty(P, dtype=np.float64)
    m = OrcMetrics()
    lib.compute_lifetimes(grid.h, C.byref(params), int(coherent), B, dptr(lts), C.byref(m))
    return lts, m


def read_points(path):
    """Reads an impurity file the way TestUtils.h:13-45 does (x y per line; '#', ':/' comments)."""
    xs, ys = [], []
    with open(path) as f:
        for line in f:
            line = line.rstrip("\n")
            if not line or line[0] == "#" or (len(line) > 1 and line[0] == ":" and line[1] == "/"):
                continue
            parts = line.split(" ", 1)
            xs.append(float(parts[0]))
            ys.append(float(parts[1]) if len(parts) > 1 else 0.0)
    return np.array(xs), np.array(ys)


def derive_images(lib, grid, params, coherent, B, raw_lifetimes, temperature, ppr):
    """IterationResult images of one temperature: [3][ppr][ppr] (lifetime, sigma_xx, sigma_xy)."""
    lt = np.ascontiguousarray(raw_lifetimes, dtype=np.float64)
    out = np.empty(3 * ppr * ppr, dtype=np.float64)
    rc = lib.derive_images(grid.h, C.byref(params), int(coherent), float(B), dptr(lt), float(temperature), dptr(out))
    assert rc == 0
    return out.reshape(3, ppr, ppr)


# Cases on which the glibc and correctly rounded builds (and the GPU) are compared.
AGREEMENT_FIELDS = [0.01, 0.1, 1.0, 20.0, 40.0]


def agreement_case(lib, case):
    """(grid, params, temperatures) of the three fields the libm/esm agreement is checked on:
    the reference's test_impurities.dat (SigmaIntegrationTest.h:17-19 settings), the small
    FP64 workload (2,000 disks, 64^2 positions, 17 phi) and the production field (60,000
    disks, 128^2 cells) at 64^2 positions, 65 phi, T in {2, 40, 100}."""
    from elasticscattering_amd import workloads as W

    if case == "file1":
        x, y = read_points(os.path.join(ROOT, "tests", "golden", "test_impurities.dat"))
        g = OracleGrid.from_points(lib, x, y, 1e-7, 1e-7, 2e-9, 10)
        return g, sim_params(31, 7, 1e-7, 1e-7, g.cell_size, 1.0, 0.3, 1, 1.67834e5, nthreads=8), [1.0, 10.0]
    w = W.SMALL_FP64 if case == "small" else W.PRODUCTION_FP64
    rs, ext = W.region(w)
    g = OracleGrid.generate(lib, w["seeds"][0], rs, ext, 0.0, w["impurity_radius"], 15, w["cells_per_row"],
                            n=w["impurity_count"])
    if case == "small":
        return g, sim_params(64, 17, rs, ext, g.cell_size, 1e-12, w["alpha"], 1, 7e5, incoherent_tau=1e-12,
                             nthreads=8), [0.0]
    return g, sim_params(64, 65, rs, ext, g.cell_size, 1e-12, w["alpha"], 1, 7e5, nthreads=8), [2.0, 40.0, 100.0]
