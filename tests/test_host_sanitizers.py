"""The C++ host mirror (teufel_amd/host) under AddressSanitizer and
UndefinedBehaviorSanitizer, on the CPU.

The host sources and tests/cpp/host_api_test.cpp are compiled with g++
-fsanitize=address,undefined against tests/cpp/tg_stub.cpp, a device-less
stand-in of the C ABI (the sanitizers instrument host code only).  The
program then reads and writes screen files: the reference's
DiffractionScreen.h5, the MeshedScreen writer's output, the ScreenObserver
HDF5 file, and truncated and bit-flipped copies of them.  Every run must end
either normally or with the reference's IOexception (the program's exit
status 110), with no sanitizer report: a malformed file is an error, never an
out-of-bounds read or undefined behaviour.
"""
import os
import shutil
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HOST = os.path.join(ROOT, "teufel_amd", "host")
BIN = os.path.join(ROOT, "tests", "cpp", "_build", "host_api_test_asan")
H5 = os.path.join(ROOT, "tests", "golden", "DiffractionScreen.h5")
ENV = dict(os.environ, ASAN_OPTIONS="allocator_may_return_null=1:detect_leaks=1:exitcode=66",
           UBSAN_OPTIONS="halt_on_error=1:print_stacktrace=1:exitcode=67")
IOEXCEPTION = 110


@pytest.fixture(scope="module")
def program():
    cxx = shutil.which(os.environ.get("CXX", "g++"))
    if not cxx:
        pytest.skip("no C++ compiler")
    os.makedirs(os.path.dirname(BIN), exist_ok=True)
    srcs = sorted(os.path.join(HOST, f) for f in os.listdir(HOST) if f.endswith(".cpp"))
    tmp = "%s.%d" % (BIN, os.getpid())
    cmd = [cxx, "-std=c++14", "-O1", "-g", "-fno-omit-frame-pointer", "-fsanitize=address,undefined",
           "-fno-sanitize-recover=undefined", "-I" + HOST, "-o", tmp] + srcs + [
        os.path.join(ROOT, "tests", "cpp", "host_api_test.cpp"), os.path.join(ROOT, "tests", "cpp", "tg_stub.cpp")]
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stderr[-4000:]
    os.replace(tmp, BIN)   # whole: parallel test workers may build it at once
    return BIN


def run(program, *args):
    r = subprocess.run([program] + list(args), capture_output=True, text=True, timeout=120, env=ENV, cwd="/tmp")
    report = "AddressSanitizer" in r.stderr or "runtime error:" in r.stderr or "LeakSanitizer" in r.stderr
    return r.returncode, report, r.stdout + r.stderr


def test_valid_files_run_clean(program, tmp_path):
    rc, rep, out = run(program, "h5info", H5)
    assert rc == 0 and not rep, out[-3000:]
    rc, rep, out = run(program, "h5write", H5, str(tmp_path / "w.h5"))
    assert rc == 0 and not rep, out[-3000:]
    rc, rep, out = run(program, "h5info", str(tmp_path / "w.h5"))
    assert rc == 0 and not rep, out[-3000:]
    rc, rep, out = run(program, "screenh5", str(tmp_path / "s.h5"))
    assert rc == 0 and not rep, out[-3000:]
    rc, rep, out = run(program, "screenread", str(tmp_path / "s.h5"), str(tmp_path / "s.dump"))
    assert rc == 0 and not rep, out[-3000:]


def test_no_device_is_an_ioexception(program):
    """Without a device the adapter raises IOexception (tg_create fails)."""
    rc, rep, out = run(program, "kat", "fp64")
    assert rc == IOEXCEPTION and not rep and "tg_create" in out, out[-3000:]


def test_concurrent_calls_without_device(program, tmp_path):
    """16 threads calling integrateFieldTrace at once: with no device every
    combined launch fails, and every caller gets the IOexception (exit status
    110), with no sanitizer report and no hang."""
    from test_host_api import _perpixel_input
    inp = str(tmp_path / "in.bin")
    _perpixel_input(inp, n_particles=2, n_steps=50, npx=8, n_bins=16)
    rc, rep, out = run(program, "perpixel", inp, "16", "fp64")
    assert rc == IOEXCEPTION and not rep and "tg_create" in out, out[-3000:]


def test_thread_sanitizer_on_combined_calls(tmp_path):
    """The same concurrent calls under ThreadSanitizer (host code and the
    device-less stand-in of the C ABI): no data race in the adapter's request
    combining."""
    cxx = shutil.which(os.environ.get("CXX", "g++"))
    if not cxx:
        pytest.skip("no C++ compiler")
    exe = os.path.join(ROOT, "tests", "cpp", "_build", "host_api_test_tsan")
    srcs = sorted(os.path.join(HOST, f) for f in os.listdir(HOST) if f.endswith(".cpp"))
    tmp = "%s.%d" % (exe, os.getpid())
    r = subprocess.run([cxx, "-std=c++14", "-O1", "-g", "-fsanitize=thread", "-I" + HOST, "-o", tmp] + srcs + [
        os.path.join(ROOT, "tests", "cpp", "host_api_test.cpp"), os.path.join(ROOT, "tests", "cpp", "tg_stub.cpp")],
        capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stderr[-4000:]
    os.replace(tmp, exe)
    from test_host_api import _perpixel_input
    inp = str(tmp_path / "in.bin")
    _perpixel_input(inp, n_particles=2, n_steps=50, npx=8, n_bins=16)
    env = dict(os.environ, TSAN_OPTIONS="halt_on_error=1:exitcode=68")
    r = subprocess.run([exe, "perpixel", inp, "16", "fp64"], capture_output=True, text=True, timeout=120, env=env,
                       cwd="/tmp")
    assert r.returncode == IOEXCEPTION and "ThreadSanitizer" not in r.stderr, (r.stdout + r.stderr)[-3000:]


def _corrupt_runs(program, tmp_path, src, mode, n_trunc, n_flip, seed):
    data = open(src, "rb").read()
    rng = np.random.default_rng(seed)
    cases = [("trunc", n) for n in sorted(set([0, 1, 8, 64, 96, 512, len(data) - 1] +
                                              list(rng.integers(0, len(data), n_trunc))))]
    cases += [("flip", int(s)) for s in range(n_flip)]
    outcomes = {0: 0, IOEXCEPTION: 0}
    for kind, v in cases:
        if kind == "trunc":
            blob = data[:v]
        else:
            b = bytearray(data)
            r = np.random.default_rng(seed * 1000 + v)
            # metadata (superblock, heaps, B-trees, object headers) sits in the
            # first few KB; flip 1-4 bits there or anywhere in the file
            hi = min(len(b), 8192) if v % 2 == 0 else len(b)
            for _ in range(int(r.integers(1, 5))):
                i = int(r.integers(0, hi))
                b[i] ^= 1 << int(r.integers(0, 8))
            blob = bytes(b)
        path = str(tmp_path / "bad.h5")
        with open(path, "wb") as f:
            f.write(blob)
        args = [mode, path] + ([str(tmp_path / "bad.out")] if mode in ("screenread", "h5write") else [])
        rc, rep, out = run(program, *args)
        assert not rep and rc in (0, IOEXCEPTION), (kind, v, rc, out[-3000:])
        outcomes[rc] += 1
        if kind == "trunc" and v < len(data) // 2:
            assert rc == IOEXCEPTION, (kind, v, out[-2000:])
    return outcomes


def test_corrupt_meshed_screen_files(program, tmp_path):
    outcomes = _corrupt_runs(program, tmp_path, H5, "h5info", 40, 120, 7)
    print("h5info on corrupt files:", outcomes)
    assert outcomes[IOEXCEPTION] > 0


def test_corrupt_screen_observer_files(program, tmp_path):
    good = str(tmp_path / "good.h5")
    rc, rep, out = run(program, "screenh5", good)
    assert rc == 0 and not rep, out
    src = str(tmp_path / "src.h5")
    shutil.copy(good, src)
    outcomes = _corrupt_runs(program, tmp_path, src, "screenread", 30, 80, 11)
    print("screenread on corrupt files:", outcomes)
    assert outcomes[IOEXCEPTION] > 0


def _read_sdds_ascii(path):
    """A reader of version-1 ASCII SDDS files, written from the protocol (the
    "SDDS1" line, &parameter / &column / &data namelists up to &end, then per
    page: one line per parameter value, the row count, one line per row;
    lines starting with '!' are comments).  Returns the definitions, the
    parameter values and the columns of the first page."""
    import shlex
    lines = open(path).read().splitlines()
    assert lines[0] == "SDDS1"
    params, cols, data = [], [], None
    i = 1
    while data is None:
        ln = lines[i].strip()
        i += 1
        assert ln.startswith("&") and ln.endswith("&end"), ln
        kind, body = ln[1:].split(None, 1)
        fields = {}
        for item in shlex.split(body.replace(",", " ")):
            if item == "&end":
                continue
            k, _, v = item.partition("=")
            fields[k] = v
        {"parameter": params, "column": cols}.get(kind, []).append(fields)
        if kind == "data":
            data = fields
    assert data["mode"] == "ascii"
    body = [ln for ln in lines[i:] if not ln.startswith("!")]
    pvals = {p["name"]: (int(v) if p["type"] == "long" else float(v)) for p, v in zip(params, body)}
    nrows = int(body[len(params)])
    rows = np.array([[float(x) for x in ln.split()] for ln in body[len(params) + 1:len(params) + 1 + nrows]])
    assert len(body) == len(params) + 1 + nrows
    return params, cols, pvals, {c["name"]: rows[:, j] for j, c in enumerate(cols)}


def test_point_observer_sdds_file(program, tmp_path):
    """PointObserver::generateOutput writes the reference's ASCII SDDS file
    (src/point_observer.cpp:120-182, read by scripts/plot_PointObservation.py):
    its parameters, columns (names, units, types) and every value, exactly."""
    out = str(tmp_path / "point.sdds")
    rc, rep, log = run(program, "pointsdds", out)
    assert rc == 0 and not rep, log[-3000:]
    params, cols, pv, cv = _read_sdds_ascii(out)
    assert [(p["name"], p["type"]) for p in params] == [("NumberTimeSteps", "long"), ("t0", "double"),
                                                        ("dt", "double")]
    assert [(c["name"], c["units"], c["type"]) for c in cols] == [
        ("t", "s", "double"), ("Ex", "V/m", "double"), ("Ey", "V/m", "double"), ("Ez", "V/m", "double"),
        ("Bx", "T", "double"), ("By", "T", "double"), ("Bz", "T", "double")]
    nots, t0, dt = 37, 3.3e-8, 2.5e-15
    assert pv == {"NumberTimeSteps": nots, "t0": t0, "dt": dt}
    it = np.arange(nots)
    assert np.array_equal(cv["t"], t0 + it * dt)
    for c, name in enumerate(("Ex", "Ey", "Ez", "Bx", "By", "Bz")):
        assert np.array_equal(cv[name], it * 1000.0 + c * 0.001 - 1.0 / 3.0)
