"""AddressSanitizer + UndefinedBehaviorSanitizer builds of the host code, on
the CPU: the runner's C++ (es_config.cpp, es_runner.cpp, es_png.cpp) and the
CPU oracle (both libm builds), each built with g++/gcc twice from the drivers
under tests/sanitize/ — with the sanitizers and without — and run on the same
inputs.  Bars: exit 0, no sanitizer report (leaks included), and the same
output text (and PNG bytes) from both builds.

The reference's own convention on bad input is to print and exit(0)
(ErrorMacros.h:22-38), and it reads outside the cell grid for start positions
off the grid (lifetime.h:88-91); these builds show that the host side here
has no such read on the inputs it is given.
"""
import os
import re
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SAN = ["-g", "-fsanitize=address,undefined", "-fno-sanitize-recover=all", "-fno-omit-frame-pointer"]
ENV = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0:halt_on_error=1",
           UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1")
REPORT = re.compile(r"runtime error|AddressSanitizer|LeakSanitizer|UndefinedBehaviorSanitizer")
HOST = os.path.join(ROOT, "elasticscattering_amd", "csrc", "host")
LIBDIR = os.path.join(ROOT, "elasticscattering_amd")


def _compile(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0 and ("asan" in r.stderr or "ubsan" in r.stderr) and "cannot find" in r.stderr:
        pytest.skip("the sanitizer runtimes are not installed: " + r.stderr.strip().splitlines()[-1])
    assert r.returncode == 0, r.stderr


def _run(cmd, cwd):
    r = subprocess.run(cmd, cwd=cwd, capture_output=True, text=True, env=ENV, timeout=300)
    assert not REPORT.search(r.stderr), r.stderr[-4000:]
    assert r.returncode == 0, r.stderr[-4000:]
    return r.stdout


@pytest.mark.parametrize("math", ["esm", "libm"])
def test_oracle_under_asan_ubsan(tmp_path, math):
    if shutil.which("gcc") is None:
        pytest.skip("gcc missing")
    src = os.path.join(ROOT, "tests", "sanitize", "oracle_driver.c")
    flags = ["-std=gnu11", "-O2", "-march=x86-64-v3", "-ffp-contract=off", "-fno-fast-math", "-Wall"]
    if math == "esm":
        flags.append("-DORACLE_USE_ESM")
    out = {}
    for kind, extra in (("san", SAN), ("plain", [])):
        exe = str(tmp_path / f"oracle_{kind}")
        _compile(["gcc", *flags, *extra, "-o", exe, src, "-lm", "-lpthread"])
        out[kind] = _run([exe], str(tmp_path))
    assert out["san"].strip().endswith("done 0")
    assert "nan" not in out["san"]
    assert out["san"] == out["plain"]


def _mask_seeds(text):
    # seeds drawn from the clock when a config names none (SimulationConfiguration.h:148-155)
    text = re.sub(r'"sample_seeds": \[[^\]]*\]', '"sample_seeds": _', text)
    return re.sub(r'"start_seed": \d+', '"start_seed": _', text)


def test_host_runner_under_asan_ubsan(tmp_path):
    if shutil.which("g++") is None:
        pytest.skip("g++ missing")
    lib = os.path.join(LIBDIR, "libes_hip.so")
    if not os.path.exists(lib):
        pytest.skip("libes_hip.so not built")
    golden = os.path.join(ROOT, "tests", "golden")
    bad = tmp_path / "bad.config"
    bad.write_text("num_samples 1\ndimension 22\nphi_steps 5\n")
    configs = [os.path.join(golden, n) for n in ("real_case.config", "log_with_images.config",
                                                  "test_smaller_case.config")]
    configs += [str(bad), str(tmp_path / "missing.config")]
    src = [os.path.join(ROOT, "tests", "sanitize", "host_driver.cpp"),
           *(os.path.join(HOST, f) for f in ("es_config.cpp", "es_runner.cpp", "es_png.cpp"))]
    out, png = {}, {}
    for kind, extra in (("san", SAN + ["-O1"]), ("plain", ["-O2"])):
        exe = str(tmp_path / f"host_{kind}")
        _compile(["g++", "-std=c++17", "-ffp-contract=off", "-Wall", *extra, "-o", exe, *src, f"-L{LIBDIR}", "-les_hip",
                  f"-Wl,-rpath,{LIBDIR}"])
        work = tmp_path / kind
        work.mkdir()
        # es_run is called only where no GPU is present: it then stops at the
        # solver's no-device error after parsing and creating the output folder.
        args = [exe, str(work)] + configs
        if os.path.exists("/dev/kfd"):
            args = [exe, "--no-run", str(work)] + configs
        out[kind] = _mask_seeds(_run(args, str(work)))
        png[kind] = (work / "img.png").read_bytes()
    s = out["san"]
    assert s == out["plain"]
    assert png["san"] == png["plain"] and png["san"][:8] == b"\x89PNG\r\n\x1a\n"
    lines = s.splitlines()
    parse = [ln for ln in lines if ln.startswith("parse ")]
    assert [int(re.search(r"rc=(-?\d+)", ln).group(1)) for ln in parse] == [0, 0, 0, -1, -1]
    assert "Dimension must be odd" in parse[3] and "Could not load" in parse[4]
    assert all(ln == "small rc=-1 untouched=1" for ln in lines if ln.startswith("small "))
    assert "png rc=0" in lines
    if not os.path.exists("/dev/kfd"):
        assert "run rc=-7" in lines  # ES_ERR_NO_DEVICE
