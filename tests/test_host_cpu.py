"""Host-side checks that need no GPU: the C ABI library loads and exports every
symbol its headers declare; the .config parser follows the reference rules
(src/SimulationConfiguration.h:77-229); solver creation fails cleanly without
a device."""
import ctypes as C
import os
import re

import pytest

import elasticscattering_amd as es

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")


def _declared_functions():
    names = []
    for h in ("es_abi.h", "es_runner.h", "es_surrogate.h"):
        text = open(os.path.join(ROOT, "include", h)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        names += re.findall(r"^\s*(?:const\s+)?[a-z_0-9]+\s*\*?\s*(es_[a-z_0-9]+)\s*\(", text, flags=re.M)
    return sorted(set(names))


def test_library_exports_every_declared_symbol():
    lib = es.load_library()
    declared = _declared_functions()
    assert len(declared) >= 12
    for name in declared:
        assert hasattr(lib, name), name
    assert sorted(es.EXPORTED_SYMBOLS) == declared


def test_version_string():
    assert b"gfx950" in es.load_library().es_version()


def _has_gpu():
    try:
        import torch
        return torch.cuda.device_count() > 0
    except Exception:
        return False


@pytest.mark.skipif(_has_gpu(), reason="checks the no-device path")
def test_solver_create_without_device_reports_no_device():
    lib = es.load_library()
    h = C.c_void_p()
    d = es.SolverDesc(31, 7, es.GridInfo(), es.ES_PRECISION_FP64, 0)
    assert lib.es_solver_create(C.byref(d), C.byref(h)) == es.ES_ERR_NO_DEVICE
    assert b"device" in lib.es_last_error(None)


def test_solver_create_rejects_bad_sizes():
    lib = es.load_library()
    h = C.c_void_p()
    for ppr, ppq in ((1, 7), (31, 1), (0, 0)):
        d = es.SolverDesc(ppr, ppq, es.GridInfo(), es.ES_PRECISION_FP64, 0)
        assert lib.es_solver_create(C.byref(d), C.byref(h)) == es.ES_ERR_INVALID_ARG
    d = es.SolverDesc(31, 7, es.GridInfo(), 99, 0)
    assert lib.es_solver_create(C.byref(d), C.byref(h)) == es.ES_ERR_INVALID_ARG


def test_parse_reference_configs():
    c = es.parse_config(os.path.join(GOLDEN, "real_case.config"))
    assert c["positions_per_row"] == 127 and c["particles_per_quadrant"] == 7 and c["num_samples"] == 1
    assert c["magnetic_fields"] == [0.01, 0.01 + 1 * (5 - 0.01) / 2, 5.0]
    assert c["temperatures"] == [0.01, 0.1, 1, 10, 100]
    assert c["output_type"] == "all" and c["target_cell_population"] == 10
    assert c["region_size"] == 1e-7 and c["impurity_radius"] == 1.11e-9 and c["is_clockwise"] == 1
    c = es.parse_config(os.path.join(GOLDEN, "log_with_images.config"))
    assert c["output_type"] == "nothing" and c["output_directory"] == ""
    assert c["magnetic_fields"][-1] == 0.01 + 5 * ((60 - 0.01) / 5) and len(c["magnetic_fields"]) == 6


def _write(tmp_path, text, name="t.config"):
    p = tmp_path / name
    p.write_text(text)
    return str(p)


BASE = """num_samples 3
dimension 21
phi_steps 5
temperatures 0 -1 2  40
magnetic_field_min 1
magnetic_field_max 3
magnetic_field_n 3
particle_speed 7e5
tau 1e-12
alpha 0.3
clockwise 1
region_size 1e-6
region_extends 5e-7
impurity_density 1e15
impurity_radius 1.1e-8
start_seed 13
"""


def test_parser_rules(tmp_path):
    c = es.parse_config(_write(tmp_path, "# comment\n:/ also comment\n" + BASE + "dimension 23\n"))
    assert c["positions_per_row"] == 21                  # first occurrence wins
    assert c["temperatures"] == [2.0, 40.0]              # T <= 0 dropped, empty tokens ignored
    assert c["magnetic_fields"] == [1.0, 2.0, 3.0]
    assert c["sample_seeds"] == [113, 226, 339]          # (seed + 100) * (i + 1)
    assert c["target_cell_population"] == 15             # default
    assert c["digits_in_sample_num"] == 1
    assert c["output_type"] == "all" and c["base_output_directory"] == "ESLogs"


def test_parser_errors(tmp_path):
    with pytest.raises(es.EsError, match="Dimension must be odd"):
        es.parse_config(_write(tmp_path, BASE.replace("dimension 21", "dimension 22")))
    with pytest.raises(es.EsError, match="Phi steps must be odd"):
        es.parse_config(_write(tmp_path, BASE.replace("phi_steps 5", "phi_steps 4")))
    with pytest.raises(es.EsError, match="n<2"):
        es.parse_config(_write(tmp_path, BASE.replace("magnetic_field_n 3", "magnetic_field_n 1")))
    with pytest.raises(es.EsError, match="Samples must be greater than 0"):
        es.parse_config(_write(tmp_path, BASE.replace("num_samples 3", "num_samples 0")))
    with pytest.raises(es.EsError, match="missing key 'tau'"):
        es.parse_config(_write(tmp_path, BASE.replace("tau 1e-12\n", "")))
    with pytest.raises(es.EsError):
        es.parse_config(str(tmp_path / "does_not_exist.config"))


def test_parser_extension_keys(tmp_path):
    c = es.parse_config(_write(tmp_path, BASE.replace("dimension 21", "dimension 64").replace("phi_steps 5", "phi_steps 17")
                               + "allow_even 1\ncells_per_row 32\nincoherent_tau 1e-12\nimpurity_count 2000\n"
                               + "magnetic_field_spacing log\ndump_lifetimes 1\n"))
    assert c["positions_per_row"] == 64 and c["particles_per_quadrant"] == 17
    assert c["cells_per_row"] == 32 and c["incoherent_tau"] == 1e-12 and c["impurity_count"] == 2000
    assert c["temperatures"] == [0.0, 2.0, 40.0]         # T = 0 kept when the incoherent tau is explicit
    assert c["magnetic_fields"][0] == pytest.approx(1.0) and c["magnetic_fields"][1] == pytest.approx(3 ** 0.5)
    assert c["dump_lifetimes"] == 1


def test_parser_windows_line_endings(tmp_path):
    c = es.parse_config(_write(tmp_path, BASE.replace("\n", "\r\n") + "output_type nothing\r\n"))
    assert c["output_type"] == "nothing" and c["positions_per_row"] == 21
