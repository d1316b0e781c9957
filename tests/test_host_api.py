"""The C++ host API (teufel_amd/host: ChargedParticle, Bunch, Beam, FieldTrace,
MeshedScreen, ScreenObserver, PointObserver, the HDF5 screen reader) driven by
tests/cpp/host_api_test.cpp.

CPU: the host library and the test program build and link; the HDF5 reader
reads the reference's screen file (tests/golden/DiffractionScreen.h5, a copy
of the data file of tests/teufel.DiffractionScreen.cpp).
GPU: the program runs the reference's own deposit known answers
(tests/teufel.integrate_field.cpp), loop check (tests/teufel.loop.cpp) and
diffraction-screen energy (tests/teufel.DiffractionScreen.cpp, 0.463 +- 0.020
uJ) through the host classes, and integrates a MeshedScreen<Beam> and a
ScreenObserver<Bunch> whose dumps are compared with the oracle and the
reference fixture (FP64: eps_E <= 1e-12; mixed: eps_W <= 1e-5).
"""
import math
import os
import re
import shutil
import struct
import subprocess

import numpy as np
import pytest

import oracle
from teufel_amd import host_build
from teufel_amd import workloads as W
from teufel_amd.metrics import energy_rel_error, trace_rel_error

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "tests", "cpp", "host_api_test.cpp")
BIN = os.path.join(ROOT, "tests", "cpp", "_build", "host_api_test")


def build_test_program():
    """Host library and test program (g++ only; the HIP library must exist)."""
    os.makedirs(os.path.dirname(BIN), exist_ok=True)
    host_build.build()
    return host_build.build_program(SRC, BIN)


def _program():
    if not os.path.exists(BIN) or os.path.getmtime(BIN) < os.path.getmtime(SRC):
        build_test_program()
    return BIN


def run(*args, timeout=120):
    print("\n[case] host_api_test " + " ".join(args), flush=True)
    r = subprocess.run([_program()] + list(args), capture_output=True, text=True, timeout=timeout)
    print(r.stdout[-4000:], r.stderr[-2000:], flush=True)
    return r


def test_host_library_builds_and_links():
    from teufel_amd import build as B
    if not os.path.exists(os.path.join(ROOT, "teufel_amd", "lib", "libteufel_gpu.so")):
        B.build()
    assert os.path.exists(build_test_program())


H5 = os.path.join(ROOT, "tests", "golden", "DiffractionScreen.h5")


def _h5dump():
    for cand in (shutil.which("h5dump"), "/opt/conda/bin/h5dump"):
        if cand and os.path.exists(cand):
            return cand
    return None


def _h5dump_values(path, obj, kind="-d", extra=()):
    """Numbers of a dataset (kind "-d") or attribute ("-a") printed by the HDF5
    library's own h5dump (an independent reader of the same file)."""
    r = subprocess.run([_h5dump(), "-m", "%.17g", kind, obj] + list(extra) + [path], capture_output=True, text=True,
                       timeout=120, cwd="/tmp")
    assert r.returncode == 0, r.stderr
    body = r.stdout.split("DATA {", 1)[1].split("}", 1)[0]
    body = re.sub(r"\([0-9,]+\):", " ", body)
    return np.array([float(x) for x in body.replace(",", " ").split()])


def test_screen_file_reader():
    """The HDF5 reader on the reference's screen file (tests/DiffractionScreen.h5,
    kept as a fixture): sizes and attributes as the file declares them, every
    observation position at its triangle's centroid, a flat mesh facing -z,
    and the values the HDF5 library's h5dump reads from the same file."""
    _program()
    r = subprocess.run([BIN, "h5info", H5], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stdout + r.stderr
    v = r.stdout.split()
    info = dict(zip(v[0::2], v[1::2]))
    if _h5dump():
        t0 = _h5dump_values(H5, "/ObservationTime")
        total = 0.0
        for x in t0:
            total += x
        assert float(info["t0sum"]) == total
        cp = _h5dump_values(H5, "/MeshCornerPoints")
        tail = r.stdout.split("cp_last")[1].split()[:3]
        assert [float(x) for x in tail] == list(cp[-3:])
        tri = _h5dump_values(H5, "/MeshTriangles")
        assert float(info["tri_last"]) == tri[-1]
    assert info["Ncp"] == "777" and info["Np"] == "1462" and info["Nt"] == "500"
    assert float(info["dt"]) == 2e-13 and float(info["t0_0"]) == -5e-11
    assert float(info["centroid_dev"]) < 1e-15
    assert abs(float(info["area"]) - 1.23984687103686e-3) < 1e-15
    assert r.stdout.rstrip().endswith("0.000000 0.000000 -1.000000")


H5_BACK = os.path.join(ROOT, "tests", "golden", "BackwardDiffractionScreen.h5")


def test_backward_screen_file_reader():
    """The reader on the reference's second screen file
    (tests/BackwardDiffractionScreen.h5, kept as a fixture): a flat mesh of
    1338 cells facing +z, 500 bins of 0.2 ps, every observation position at
    its triangle's centroid, and the values the HDF5 library's h5dump reads
    from the same file."""
    _program()
    r = subprocess.run([BIN, "h5info", H5_BACK], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stdout + r.stderr
    v = r.stdout.split()
    info = dict(zip(v[0::2], v[1::2]))
    if _h5dump():
        t0 = _h5dump_values(H5_BACK, "/ObservationTime")
        total = 0.0
        for x in t0:
            total += x
        assert float(info["t0sum"]) == total
        assert float(info["t0_0"]) == t0[0]
        cp = _h5dump_values(H5_BACK, "/MeshCornerPoints")
        tail = r.stdout.split("cp_last")[1].split()[:3]
        assert [float(x) for x in tail] == list(cp[-3:])
        tri = _h5dump_values(H5_BACK, "/MeshTriangles")
        assert float(info["tri_last"]) == tri[-1]
    assert info["Ncp"] == "712" and info["Np"] == "1338" and info["Nt"] == "500"
    assert float(info["dt"]) == 2e-13
    assert float(info["centroid_dev"]) < 1e-15
    assert r.stdout.rstrip().endswith("0.000000 0.000000 1.000000")


def test_screen_file_writer(tmp_path):
    """MeshedScreen::writeFile (src/mesh.cpp:427-736 layout): the written file
    reads back identically, and the HDF5 library's h5dump reads its
    structure, attributes and field values."""
    _program()
    out = str(tmp_path / "written.h5")
    r = subprocess.run([BIN, "h5write", H5, out], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stdout + r.stderr
    if not _h5dump():
        pytest.skip("h5dump not available for the independent check")
    hdr = subprocess.run([_h5dump(), "-H", out], capture_output=True, text=True, timeout=60, cwd="/tmp")
    assert hdr.returncode == 0, hdr.stderr
    for name, shape in (("ElMagField", "( 1462, 500, 6 )"), ("MeshCornerPoints", "( 777, 3 )"),
                        ("MeshTriangles", "( 1462, 3 )"), ("ObservationPosition", "( 1462, 3 )"),
                        ("ObservationTime", "( 1462 )")):
        block = hdr.stdout.split('DATASET "%s"' % name)[1].split("DATASET")[0]
        assert shape in block
    assert _h5dump_values(out, "/ObservationTime/Nt", "-a").tolist() == [500]
    assert _h5dump_values(out, "/ObservationTime/dt", "-a").tolist() == [2e-13]
    assert _h5dump_values(out, "/MeshCornerPoints/Ncp", "-a").tolist() == [777]
    f = _h5dump_values(out, "/ElMagField", "-d", ["-s", "3,1,0", "-c", "1,1,6"])
    assert np.array_equal(f, 3 + 1e-3 * np.arange(6, 12))
    assert np.array_equal(_h5dump_values(out, "/ObservationTime"), _h5dump_values(H5, "/ObservationTime"))


def test_screen_observer_file(tmp_path):
    """ScreenObserver::generateOutput writes the HDF5 file of
    src/screen.cpp:297-558 (datasets Screen, ObservationPosition, ElMagField
    with their attributes, the layout scripts/teufel.py:23-75 reads); it reads
    back identically, and the HDF5 library's h5dump reads the same structure,
    attributes and values."""
    _program()
    out = str(tmp_path / "screen.h5")
    r = subprocess.run([BIN, "screenh5", out], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stdout + r.stderr
    if not _h5dump():
        pytest.skip("h5dump not available for the independent check")
    hdr = subprocess.run([_h5dump(), "-H", out], capture_output=True, text=True, timeout=60, cwd="/tmp")
    assert hdr.returncode == 0, hdr.stderr
    for name, shape in (("ElMagField", "( 9, 7, 16, 6 )"), ("ObservationPosition", "( 9, 7, 3 )"),
                        ("Screen", "( 4, 3 )")):
        block = hdr.stdout.split('DATASET "%s"' % name)[1].split("DATASET")[0]
        assert shape in block and "H5T_IEEE_F64LE" in block
    for attr, want in (("Nx", 9), ("Ny", 7), ("NOTS", 16), ("t0", 3.3e-8), ("dt", 2.5e-15), ("dtx", 1e-15),
                       ("dty", -2e-15)):
        assert _h5dump_values(out, "/Screen/" + attr, "-a").tolist() == [want], attr
    assert _h5dump_values(out, "/ElMagField/NOTS", "-a").tolist() == [16]
    assert _h5dump_values(out, "/ElMagField/dt", "-a").tolist() == [2.5e-15]
    assert _h5dump_values(out, "/ObservationPosition/Nx", "-a").tolist() == [9]
    g = _h5dump_values(out, "/Screen").reshape(4, 3)
    dx, dy = np.array([0.004, 0.0, 0.0005]), np.array([0.0, 0.005, 0.0])
    n = -np.cross(dx, dy)
    assert np.array_equal(g[0], [0.01, -0.02, 10.0]) and np.array_equal(g[2], dx) and np.array_equal(g[3], dy)
    assert np.allclose(g[1], n / np.linalg.norm(n), rtol=0, atol=1e-16)
    # cell (ix=4, iy=2): origin + dX (4 - 9//2) + dY (2 - 7//2)
    p = _h5dump_values(out, "/ObservationPosition", "-d", ["-s", "4,2,0", "-c", "1,1,3"])
    assert np.allclose(p, [0.01, -0.025, 10.0], rtol=0, atol=1e-15)
    f = _h5dump_values(out, "/ElMagField", "-d", ["-s", "8,6,15,0", "-c", "1,1,1,6"])
    assert np.array_equal(f, 8000.0 + 60.0 + 15e-3 + 1e-6 * np.arange(6))


def test_snapshot_observer_file(tmp_path):
    """SnapshotObserver::generateOutput writes the HDF5 file of
    src/snapshot.cpp:226-392 (ObservationPosition [Nx][Ny][3] with Nx, Ny;
    ElMagField [Nx][Ny][6] with t0); it reads back identically, and h5dump
    reads the same shapes, attributes and values."""
    _program()
    out = str(tmp_path / "snap.h5")
    r = subprocess.run([BIN, "snaph5", out], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stdout + r.stderr
    if not _h5dump():
        pytest.skip("h5dump not available for the independent check")
    hdr = subprocess.run([_h5dump(), "-H", out], capture_output=True, text=True, timeout=60, cwd="/tmp")
    assert hdr.returncode == 0, hdr.stderr
    for name, shape in (("ElMagField", "( 6, 5, 6 )"), ("ObservationPosition", "( 6, 5, 3 )")):
        block = hdr.stdout.split('DATASET "%s"' % name)[1].split("DATASET")[0]
        assert shape in block and "H5T_IEEE_F64LE" in block
    assert _h5dump_values(out, "/ObservationPosition/Nx", "-a").tolist() == [6]
    assert _h5dump_values(out, "/ObservationPosition/Ny", "-a").tolist() == [5]
    assert _h5dump_values(out, "/ElMagField/t0", "-a").tolist() == [2.7e-8]
    # cell (ix=5, iy=0): origin + dX (5 - 5/2) + dY (0 - 2) (src/snapshot.cpp:405-408)
    p = _h5dump_values(out, "/ObservationPosition", "-d", ["-s", "5,0,0", "-c", "1,1,3"])
    assert np.allclose(p, [-0.02 + 0.003 * 2.5, 0.01 - 0.004 * 2.0, 7.5 + 0.001 * 2.5], rtol=0, atol=1e-15)
    f = _h5dump_values(out, "/ElMagField", "-d", ["-s", "5,4,0", "-c", "1,1,6"])
    assert np.array_equal(f, 500.0 + 4.0 + 1e-3 * np.arange(6) - 0.5)


def test_screen_file_rejects_malformed(tmp_path):
    """The ScreenObserver file reader raises IOexception (exit status 110 of
    the test program) for a MeshedScreen file, which lacks the Screen dataset."""
    _program()
    r = subprocess.run([BIN, "screenread", H5, str(tmp_path / "x.dump")], capture_output=True, text=True, timeout=60)
    assert r.returncode == 110 and "no dataset Screen" in r.stdout, r.stdout + r.stderr


def diffraction_trajectory():
    """tests/teufel.DiffractionScreen.cpp:73-103: 100 pC at gamma 20 on a
    straight line from z = -10 m to +10 m (an empty lattice), passing z = 0 at
    t = 0, 20000 steps."""
    c = 2.99792458e8
    gamma = 20.0
    beta = math.sqrt(1.0 - 1.0 / (gamma * gamma))
    bg = math.sqrt(gamma * gamma - 1.0)
    n = 20000
    t_start = -10.0 / (beta * c)
    dt = (20.0 / (beta * c)) / n
    k = np.arange(n + 1)
    z = np.zeros((1, n + 1))
    return oracle.Trajectories(t=(t_start + k * dt)[None], x=z, y=z, z=(-10.0 + beta * c * (k * dt))[None],
                               ux=z, uy=z, uz=np.full((1, n + 1), bg), ax=z, ay=z, az=z,
                               charge=np.array([-100e-12 / 1.6021766208e-19]))


@pytest.mark.gpu
@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_diffraction_screen_end_to_end(tmp_path, precision):
    tr = diffraction_trajectory()
    traj = str(tmp_path / "traj.bin")
    with open(traj, "wb") as f:
        f.write(struct.pack("<ii", 1, tr.n_points))
        for name in oracle.Trajectories.NAMES:
            f.write(np.ascontiguousarray(getattr(tr, name), "<f8").tobytes())
        f.write(np.ascontiguousarray(tr.charge, "<f8").tobytes())
    dump = str(tmp_path / "diffraction.dump")
    r = run("diffraction", H5, traj, dump, precision)
    assert r.returncode == 0
    energy = float(r.stdout.split("radiation energy =")[1].split()[0])
    assert abs(energy - 0.463) <= 0.020          # the reference test's bound
    pts, t0, dt, traces = read_dump(dump)
    assert traces.shape == (1462, 500, 6) and dt == 2e-13
    # MeshedScreen::writeReport (src/mesh.cpp:783-821)
    rep = r.stdout.split("Meshed Screen")[1]
    assert "Np=1462  Nt=500" in rep
    assert abs(float(rep.split("Energy incident on screen =")[1].split()[0]) / energy - 1) < 1e-5
    S = np.array([oracle.trace_poynting(t, dt) for t in traces])
    peak = int(np.argmax(np.linalg.norm(S, axis=1)))
    assert "peak pos[%d] =" % peak in rep
    assert abs(float(rep.split("peak energy density =")[1].split()[0]) / np.linalg.norm(S[peak]) - 1) < 1e-5
    assert "dt  = 0.0002 ns" in rep
    pick = np.arange(0, 1462, 7)
    ref = oracle.integrate_screen(tr, pts[pick], t0[pick], dt, 500)
    if precision == "fp64":
        assert trace_rel_error(traces[pick], ref) <= 1e-12
    else:
        n = np.array([0.0, 0.0, -1.0])
        assert energy_rel_error(pixel_energy(traces[pick], dt, n), pixel_energy(ref, dt, n))[0] <= 1e-5


@pytest.mark.gpu
@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_backward_diffraction_screen(tmp_path, precision):
    """The reference's BackwardDiffractionScreen.h5 (a screen facing +z)
    read into a MeshedScreen<Bunch> with the DiffractionScreen test's charge
    (100 pC at gamma 20 over 20 m), integrated on the GPU and compared on
    every 5th pixel with the oracle: FP64 within 1e-12, mixed within eps_W
    1e-5.  The straight trajectory sweeps its field past this screen, so the
    energy through it is ~0 (the traces are not)."""
    tr = diffraction_trajectory()
    traj = str(tmp_path / "traj.bin")
    with open(traj, "wb") as f:
        f.write(struct.pack("<ii", 1, tr.n_points))
        for name in oracle.Trajectories.NAMES:
            f.write(np.ascontiguousarray(getattr(tr, name), "<f8").tobytes())
        f.write(np.ascontiguousarray(tr.charge, "<f8").tobytes())
    dump = str(tmp_path / "backward.dump")
    r = run("meshscreen", H5_BACK, traj, dump, precision)
    assert r.returncode == 0, r.stdout + r.stderr
    print(r.stdout.splitlines()[0])
    pts, t0, dt, traces = read_dump(dump)
    assert traces.shape == (1338, 500, 6) and dt == 2e-13
    pick = np.arange(0, 1338, 5)
    ref = oracle.integrate_screen(tr, pts[pick], t0[pick], dt, 500)
    assert np.abs(ref).max() > 0
    if precision == "fp64":
        assert trace_rel_error(traces[pick], ref) <= 1e-12
    else:
        n = np.array([0.0, 0.0, 1.0])
        assert energy_rel_error(pixel_energy(traces[pick], dt, n), pixel_energy(ref, dt, n))[0] <= 1e-5


def pixel_energy(traces, dt, normal):
    """W_p = -S_p . n_p with S_p = dt sum_i E_i x B_i / mu0."""
    S = np.array([oracle.trace_poynting(t, dt) for t in traces])
    return W.energy_density(S, np.broadcast_to(normal, S.shape))


def read_dump(path):
    with open(path, "rb") as f:
        raw = f.read()
    assert raw[:8] == b"TGDUMP1\0"
    n_pt, n_s = struct.unpack_from("<ii", raw, 8)
    (dt,) = struct.unpack_from("<d", raw, 16)
    off = 24
    head = np.empty((n_pt, 4))
    traces = np.empty((n_pt, n_s, 6))
    for i in range(n_pt):
        head[i] = np.frombuffer(raw, "<f8", 4, off)
        off += 32
        traces[i] = np.frombuffer(raw, "<f8", 6 * n_s, off).reshape(n_s, 6)
        off += 48 * n_s
    assert off == len(raw)
    return head[:, :3], head[:, 3], dt, traces


def write_input(path, tr, obs, t0, dt, nt, area, normal, scr):
    np_, npts = tr.t.shape
    with open(path, "wb") as f:
        f.write(struct.pack("<ii", np_, npts))
        for name in oracle.Trajectories.NAMES:
            f.write(np.ascontiguousarray(getattr(tr, name), "<f8").tobytes())
        f.write(np.ascontiguousarray(tr.charge, "<f8").tobytes())
        f.write(struct.pack("<iid", obs.shape[0], nt, dt))
        for a in (obs, t0, area, normal):
            f.write(np.ascontiguousarray(a, "<f8").tobytes())
        f.write(np.asarray(scr["origin"] + scr["dx"] + scr["dy"], "<f8").tobytes())
        f.write(struct.pack("<ii", scr["nx"], scr["ny"]))
        f.write(struct.pack("<dddd", scr["t0"], scr["dt"], scr["dtx"], scr["dty"]))
        f.write(struct.pack("<i", scr["nots"]))


@pytest.mark.gpu
@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_host_kats(precision):
    r = run("kat", precision)
    assert r.returncode == 0


@pytest.mark.gpu
def test_host_loop_point_observer():
    r = run("loop", "fp64")
    assert r.returncode == 0


@pytest.mark.gpu
@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_host_screens_vs_oracle(tmp_path, precision):
    g = np.load(os.path.join(ROOT, "tests", "golden", "bunch_ref.npz"))
    tr = oracle.Trajectories(**{n: g[n] for n in oracle.Trajectories.NAMES}, charge=g["charge"])
    # MeshedScreen: the fixture's 4 x 4 pixels (0.05 m pitch at 10 m, per-pixel t0)
    obs, t0, dt, nt = g["obs"], g["t0"], float(g["dt"]), int(g["n_bins"])
    area = np.full(obs.shape[0], 0.05 * 0.05)
    normal = np.tile([0.0, 0.0, -1.0], (obs.shape[0], 1))
    # ScreenObserver: 9 x 7 cells around the axis at 10 m, linear time offsets
    axis = int(np.flatnonzero((obs[:, 0] == 0) & (obs[:, 1] == 0))[0])
    scr = dict(origin=[0.0, 0.0, 10.0], dx=[0.004, 0.0, 0.0], dy=[0.0, 0.005, 0.0], nx=9, ny=7,
               t0=float(t0[axis]), dt=dt, dtx=1e-15, dty=-2e-15, nots=nt)
    inp = str(tmp_path / "in.bin")
    mesh_dump, scr_h5, scr_dump = str(tmp_path / "mesh.dump"), str(tmp_path / "screen.h5"), str(tmp_path / "screen.dump")
    write_input(inp, tr, obs, t0, dt, nt, area, normal, scr)
    r = run("screens", inp, mesh_dump, scr_h5, precision)
    assert r.returncode == 0
    # the ScreenObserver output is its HDF5 file; read it back through the
    # file reader into a dump
    assert run("screenread", scr_h5, scr_dump).returncode == 0

    pts, pt0, pdt, mesh = read_dump(mesh_dump)
    assert np.array_equal(pts, obs) and np.array_equal(pt0, t0) and pdt == dt
    ref = oracle.integrate_screen(tr, obs, t0, dt, nt)
    if precision == "fp64":
        assert trace_rel_error(mesh, ref) <= 1e-12
        assert trace_rel_error(mesh, g["traces"]) <= 1e-12
    else:
        assert energy_rel_error(pixel_energy(mesh, dt, normal), pixel_energy(ref, dt, normal))[0] <= 1e-5
    e_line = [ln for ln in r.stdout.splitlines() if "totalEnergy" in ln][0]
    energy = float(e_line.split("=")[1].split()[0])
    e_ref = oracle.total_energy(ref, dt, area, normal)
    assert abs(energy - e_ref) <= (1e-12 if precision == "fp64" else 1e-5) * abs(e_ref)

    spts, st0, sdt, straces = read_dump(scr_dump)
    assert straces.shape == (63, scr["nots"], 6) and sdt == scr["dt"]
    ix, iy = np.meshgrid(np.arange(9) - 9 // 2, np.arange(7) - 7 // 2, indexing="ij")
    assert np.allclose(spts[:, 0], 0.004 * ix.ravel(), rtol=0, atol=1e-15)
    assert np.allclose(spts[:, 1], 0.005 * iy.ravel(), rtol=0, atol=1e-15)
    assert np.allclose(st0, scr["t0"] + 1e-15 * ix.ravel() - 2e-15 * iy.ravel(), rtol=0, atol=1e-22)
    sref = oracle.integrate_screen(tr, spts, st0, scr["dt"], scr["nots"])
    if precision == "fp64":
        assert trace_rel_error(straces, sref) <= 1e-12
    else:
        n = np.array([0.0, 0.0, -1.0])
        assert energy_rel_error(pixel_energy(straces, sdt, n), pixel_energy(sref, sdt, n))[0] <= 1e-5


def _perpixel_input(path, n_particles=8, n_steps=2000, npx=16, n_bins=128):
    w = W.make("c2", n_particles=n_particles, n_steps=n_steps, seed=41)
    u = w.und
    tr = oracle.track_undulator(w.bunch.x0, w.bunch.u0, w.dt, w.n_steps, u.bpeak, u.lambda_u, u.n_periods,
                                u.origin)
    obs, nrm, area = W.grid_screen(npx, 0.2, 10.0)
    t_c, x_c = W._centre(u, 60.0)
    t0, dt = W.set_windows(obs, n_bins, u, 60.0, t_c, x_c, 0.3e-3)
    scr = dict(origin=[0.0, 0.0, 10.0], dx=[0.01, 0.0, 0.0], dy=[0.0, 0.01, 0.0], nx=2, ny=2, t0=float(t0[0]),
               dt=dt, dtx=0.0, dty=0.0, nots=n_bins)
    write_input(path, tr, obs, t0, dt, n_bins, area, nrm, scr)
    return obs.shape[0]


@pytest.mark.gpu
@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_host_per_pixel_calls_are_combined(tmp_path, precision):
    """The reference's per-pixel loop, 16 threads calling
    Bunch::integrateFieldTrace(point, trace) at once for 256 pixels, gives the
    batched MeshedScreen::integrate() result (FP64 1e-12, mixed 1e-5 of each
    pixel's largest field), and the adapter combines the calls into far fewer
    launches than pixels."""
    inp = str(tmp_path / "in.bin")
    n_pix = _perpixel_input(inp)
    r = run("perpixel", inp, "16", precision)
    assert r.returncode == 0
    line = [ln for ln in r.stdout.splitlines() if ln.startswith("perpixel:")][0]
    launches = int(re.search(r"launches (\d+)", line).group(1))
    assert launches <= 1 + n_pix // 2, line
