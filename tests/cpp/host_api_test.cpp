// host_api_test.cpp - the reference's own radiation tests, written against the
// C++ host API (teufel_amd/host) and run on the GPU.
//
//   host_api_test kat [fp64|mixed]
//       the five deposit cases of tests/teufel.integrate_field.cpp:91-191
//       (pseudo-particles with injected retarded times), same expected values
//       and the same 1e-6 tolerance
//   host_api_test loop [fp64|mixed]
//       tests/teufel.loop.cpp:160-190: the field at the centre of a circular
//       orbit seen by a PointObserver<Bunch> and by the snapshot
//       Bunch::RetardedField(tau, centre) equals mu0 I / (2R) within 1e-3; a
//       SnapshotObserver<Bunch> centred there gives the same field.  The orbit
//       (one revolution, 200 steps) is the exact circle instead of the
//       reference's Vay-tracked one.
//   host_api_test h5info <screen.h5>
//       reads a TEUFEL screen file with the HDF5 reader and prints its sizes,
//       total area and the largest distance between an observation position
//       and its triangle's centroid (no GPU needed)
//   host_api_test h5write <screen.h5> <out.h5>
//       reads a screen file, sets known trace values, writes it with
//       MeshedScreen::writeFile and reads it back (no GPU needed)
//   host_api_test diffraction <screen.h5> <trajectory> <out.dump> [fp64|mixed]
//       tests/teufel.DiffractionScreen.cpp:130-160: the screen file read into a
//       MeshedScreen<Bunch>, one macro-particle's trajectory (from
//       <trajectory>, written by tests/test_host_api.py), integrate, print
//       totalEnergy(); the Python test checks 0.463 +- 0.020 uJ (the reference
//       test's own bound) and the traces against the oracle
//   host_api_test meshscreen <screen.h5> <trajectory> <out.dump> [fp64|mixed]
//       the same for any screen file, without the 0.463 uJ check (the
//       reference's BackwardDiffractionScreen.h5)
//   host_api_test screens <input> <mesh.dump> <screen.h5> [fp64|mixed]
//       trajectories and a mesh definition from <input> (written by
//       tests/test_host_api.py), integrated by MeshedScreen<Beam> (two
//       bunches) and ScreenObserver<Bunch> (whose output is the HDF5 file of
//       src/screen.cpp:297-558); both are compared against the oracle by the
//       Python test.  Prints the mesh's totalEnergy().
//   host_api_test perpixel <input> <threads> [fp64|mixed]
//       the reference's per-pixel loop: <threads> threads call
//       Bunch::integrateFieldTrace(point, trace) for their pixels of a
//       MeshedScreen<Bunch> at once; the result must equal one
//       MeshedScreen::integrate() of the same screen (the adapter combines
//       the concurrent calls into batched launches); prints both times
//   host_api_test screenh5 <out.h5>
//       a ScreenObserver<Bunch> with known field values written by
//       generateOutput() and read back with readScreenFile: geometry, cell
//       positions, sampling and every value must survive (no GPU needed)
//   host_api_test screenread <screen.h5> <out.dump>
//       reads a ScreenObserver HDF5 file and writes its cells and traces as a
//       plain dump (positions, cell t0, samples) for the Python tests
//   host_api_test pointsdds <out.sdds>
//       a PointObserver<Bunch> with known field values written by
//       generateOutput(): the ASCII SDDS file of src/point_observer.cpp:120-182,
//       which the Python test parses and checks value for value (no GPU
//       needed)
//   host_api_test snaph5 <out.h5>
//       a SnapshotObserver<Bunch> with known field values written by
//       generateOutput() (the HDF5 file of src/snapshot.cpp:226-392) and read
//       back with the HDF5 reader: dataset shapes, attributes, cell positions
//       and every value must survive (no GPU needed)
//
// Exit status: number of failed checks (0 = pass), 100+ on usage/IO errors.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <iostream>
#include <memory>
#include <string>
#include <thread>
#include <vector>

#include "bunch.h"
#include "global.h"
#include "gpu_radiation.h"
#include "h5lite.h"
#include "screens.h"
#include "snapshot.h"

namespace {

int set_precision(const char* arg)
{
    if (!arg || std::strcmp(arg, "fp64") == 0) {
        teufel_gpu::set_precision(0);
        return 0;
    }
    if (std::strcmp(arg, "mixed") == 0) {
        teufel_gpu::set_precision(1);
        return 0;
    }
    std::fprintf(stderr, "unknown precision %s\n", arg);
    return 100;
}

// ---- deposit known answers ---------------------------------------------------

struct Pseudo {
    double t1, t2, e1, e2;
};

int kat_case(const char* name, const std::vector<Pseudo>& ps, double t0, double dt, int n,
             const std::vector<double>& expected)
{
    std::vector<double> ts, f;
    for (const Pseudo& p : ps) {
        ts.push_back(p.t1);
        ts.push_back(p.t2);
        double f1[6] = {p.e1, 0, 0, 0, 0, 0}, f2[6] = {p.e2, 0, 0, 0, 0, 0};
        f.insert(f.end(), f1, f1 + 6);
        f.insert(f.end(), f2, f2 + 6);
    }
    FieldTrace ft(t0, dt, (std::size_t)n);
    teufel_gpu::deposit_only((int)ps.size(), 2, ts.data(), f.data(), &ft);
    int errors = 0;
    std::printf("%s:\n", name);
    for (int i = 0; i < n; i++) {
        double ex = ft.get_field((std::size_t)i).E().x;
        bool bad = std::fabs(ex - expected[i]) > 1e-6;
        errors += bad;
        std::printf("  t=%6.2f : Ex=%9.6f%s\n", ft.get_time((std::size_t)i), ex, bad ? "  <-- expected" : "");
    }
    return errors;
}

int run_kat()
{
    int errors = 0;
    errors += kat_case("aligned", {{0.0, 1.0, 1.5, 0.5}}, 0.0, 0.2, 8, {0.725, 1.3, 1.1, 0.9, 0.7, 0.275, 0.0, 0.0});
    errors += kat_case("shifted", {{0.0, 1.0, 1.5, 0.5}}, -0.1, 0.2, 8, {0.0, 1.4, 1.2, 1.0, 0.8, 0.6, 0.0, 0.0});
    errors += kat_case("spike a", {{0.05, 0.15, 1.5, 0.5}}, -0.2, 0.2, 4, {0.0, 0.3125, 0.1875, 0.0});
    errors += kat_case("spike b", {{0.05, 0.15, 1.5, 0.5}}, -0.1, 0.2, 4, {0.0, 0.5, 0.0, 0.0});
    errors += kat_case("five overlapping",
                       {{1.0, 2.0, 1.0, 0.0}, {1.0, 2.0, 0.0, 1.0}, {2.0, 2.2, 1.0, 1.0}, {2.2, 2.4, 1.0, 0.0},
                        {2.4, 3.0, 0.0, 1.0}},
                       0.0, 0.5, 8, {0.0, 0.0, 0.5, 1.0, 0.9875, 0.316667, 0.395833, 0.0});
    return errors;
}

// ---- circular loop -------------------------------------------------------------

int run_loop()
{
    const double B0 = 0.5, gamma = 10.0;
    const double bg = std::sqrt(gamma * gamma - 1.0);
    const double omega = B0 / (gamma * mecsquared) * SpeedOfLight * SpeedOfLight;
    const double radius = bg * mecsquared / (B0 * SpeedOfLight);
    const double tau = 2 * Pi / omega;
    const double charge = 1.0e-9;
    const double ne = charge / ElementaryCharge;
    const double expected = MuNull * charge / tau / (2.0 * radius);
    // positive charge moving along +z in B = +y: circle about (-R, 0, 0) in x-z
    ChargedParticle* e = new ChargedParticle(ne, ne);
    const int steps = 200;
    for (int k = 0; k <= steps; k++) {
        double t = tau * k / steps, ph = omega * t;
        Vector X(-radius + radius * std::cos(ph), 0.0, radius * std::sin(ph));
        Vector P(-bg * std::sin(ph), 0.0, bg * std::cos(ph));
        Vector A(-bg * omega * std::cos(ph), 0.0, -bg * omega * std::sin(ph));
        if (k == 0)
            e->initTrajectory(t, X, P, A);
        else
            e->storeTrajectoryPoint(t, X, P, A);
    }
    Bunch* bunch = new Bunch();
    bunch->Add(e);
    PointObserver<Bunch> obs("", Vector(-radius, 0.0, 0.0), tau - 1.0e-15, 2.0e-15, 2);
    obs.setSource(bunch);
    obs.integrate();
    double b = obs.getField(0).B().norm();
    int errors = std::fabs(b - expected) / expected > 1e-3;
    std::printf("loop: R = %.6g m, tau = %.6g s, |B| = %.6g T, expected %.6g T%s\n", radius, tau, b, expected,
                errors ? "  <-- failed" : "");
    // the snapshot of tests/teufel.loop.cpp:166-172: RetardedField(tau, centre)
    ElMagField snap = bunch->RetardedField(tau, Vector(-radius, 0.0, 0.0));
    if (std::fabs(snap.B().norm() - expected) / expected > 1e-3) {
        std::printf("loop: snapshot |B| = %.6g T  <-- failed\n", snap.B().norm());
        errors++;
    } else {
        std::printf("loop: snapshot |B| = %.6g T\n", snap.B().norm());
    }
    SnapshotObserver<Bunch> so("", Vector(-radius, 0.0, 0.0), Vector(1e-3, 0, 0), Vector(0, 1e-3, 0), 3, 3, tau);
    so.setSource(bunch);
    so.integrate();
    if (so.getField(1, 1).B().x != snap.B().x || so.getField(1, 1).E().z != snap.E().z) {
        std::printf("loop: SnapshotObserver centre cell differs from RetardedField\n");
        errors++;
    }
    // buffer exchange round trip (src/observer.h getBuffer / fromBuffer)
    std::size_t n = obs.getBufferSize();
    std::unique_ptr<double[]> buf(obs.getBuffer());
    PointObserver<Bunch> copy("", Vector(-radius, 0.0, 0.0), tau - 1.0e-15, 2.0e-15, 2);
    copy.fromBuffer(buf.get(), n);
    if (copy.getField(0).B().norm() != b) {
        std::printf("loop: getBuffer/fromBuffer round trip differs\n");
        errors++;
    }
    delete bunch;
    return errors;
}

// ---- screens from an input file -----------------------------------------------

struct Reader {
    FILE* f;
    bool ok = true;
    template <class T>
    void get(T* p, size_t n)
    {
        ok = ok && std::fread(p, sizeof(T), n, f) == n;
    }
    template <class T>
    T one()
    {
        T v{};
        get(&v, 1);
        return v;
    }
    Vector vec()
    {
        double v[3] = {0, 0, 0};
        get(v, 3);
        return Vector(v[0], v[1], v[2]);
    }
};

// the input of the screens and perpixel modes (written by tests/test_host_api.py)
struct Input {
    int32_t np = 0, npts = 0, npix = 0, nt = 0, nx = 0, ny = 0, nots = 0;
    std::vector<double> a, q, t0, area;
    double dt = 0, st0 = 0, sdt = 0, dtx = 0, dty = 0;
    std::vector<Vector> pts, nrm;
    Vector origin, dx, dy;
};

int read_input(const char* input, Input& in)
{
    FILE* f = std::fopen(input, "rb");
    if (!f) {
        std::fprintf(stderr, "cannot open %s\n", input);
        return 101;
    }
    Reader r{f};
    in.np = r.one<int32_t>();
    in.npts = r.one<int32_t>();
    if (!r.ok || in.np < 1 || in.npts < 1) {
        std::fclose(f);
        return 102;
    }
    size_t n = (size_t)in.np * in.npts;
    in.a.resize(10 * n);
    in.q.resize(in.np);
    r.get(in.a.data(), in.a.size());
    r.get(in.q.data(), in.q.size());
    in.npix = r.one<int32_t>();
    in.nt = r.one<int32_t>();
    in.dt = r.one<double>();
    if (!r.ok || in.npix < 1 || in.nt < 1) {
        std::fclose(f);
        return 103;
    }
    in.t0.resize(in.npix);
    in.area.resize(in.npix);
    for (int i = 0; i < in.npix; i++) in.pts.push_back(r.vec());
    r.get(in.t0.data(), in.npix);
    r.get(in.area.data(), in.npix);
    for (int i = 0; i < in.npix; i++) in.nrm.push_back(r.vec());
    in.origin = r.vec();
    in.dx = r.vec();
    in.dy = r.vec();
    in.nx = r.one<int32_t>();
    in.ny = r.one<int32_t>();
    in.st0 = r.one<double>();
    in.sdt = r.one<double>();
    in.dtx = r.one<double>();
    in.dty = r.one<double>();
    in.nots = r.one<int32_t>();
    std::fclose(f);
    return r.ok ? 0 : 104;
}

// particle i of the input: charge in e, trajectory rows [quantity][particle][point]
ChargedParticle* make_particle(const Input& in, int i)
{
    const size_t n = (size_t)in.np * in.npts;
    const double* t = in.a.data();
    ChargedParticle* p = new ChargedParticle(in.q[i], 1.0);
    for (int k = 0; k < in.npts; k++) {
        size_t g = (size_t)i * in.npts + k;
        Vector X(t[n + g], t[2 * n + g], t[3 * n + g]);
        Vector P(t[4 * n + g], t[5 * n + g], t[6 * n + g]);
        Vector A(t[7 * n + g], t[8 * n + g], t[9 * n + g]);
        if (k == 0)
            p->initTrajectory(t[g], X, P, A);
        else
            p->storeTrajectoryPoint(t[g], X, P, A);
    }
    return p;
}

// The reference's per-pixel loop (src/mesh.cpp:296-301: many threads call
// integrateFieldTrace(point, trace) of one source, each on its own pixels)
// against one MeshedScreen<Bunch>::integrate() of the same pixels: the
// adapter combines the concurrent calls into batched launches, and both must
// agree (FP64: 1e-12 of each pixel's largest field; mixed: 1e-5).  Prints the
// time of each and how many launches the threads' calls took.
int run_perpixel(const char* input, int n_threads, bool mixed)
{
    Input in;
    if (int rc = read_input(input, in)) return rc;
    std::unique_ptr<Bunch> all(new Bunch());
    for (int i = 0; i < in.np; i++) all->Add(make_particle(in, i));
    MeshedScreen<Bunch> mesh(in.pts, in.t0, in.dt, in.nt, in.area, in.nrm);
    mesh.setSource(all.get());
    auto t_a = std::chrono::steady_clock::now();
    mesh.integrate();
    const double s_batched = std::chrono::duration<double>(std::chrono::steady_clock::now() - t_a).count();
    std::unique_ptr<double[]> batched(mesh.getBuffer());
    mesh.zero();
    t_a = std::chrono::steady_clock::now();
    std::vector<std::thread> pool;
    std::vector<std::string> errs(n_threads);
    for (int t = 0; t < n_threads; t++)
        pool.emplace_back([&, t]() {
            try {
                for (int ip = t; ip < mesh.get_Np(); ip += n_threads)
                    all->integrateFieldTrace(mesh.get_point(ip), mesh.get_trace(ip));
            } catch (const std::exception& e) {
                errs[t] = e.what();
            }
        });
    for (std::thread& th : pool) th.join();
    const double s_pixels = std::chrono::duration<double>(std::chrono::steady_clock::now() - t_a).count();
    for (const std::string& e : errs)
        if (!e.empty()) throw IOexception(e);
    std::unique_ptr<double[]> per(mesh.getBuffer());
    int errors = 0;
    double worst = 0.0;
    const size_t per_pix = (size_t)6 * in.nt;
    for (int ip = 0; ip < in.npix; ip++) {
        double m = 0.0, d = 0.0;
        for (size_t k = 0; k < per_pix; k++) {
            m = std::max(m, std::fabs(batched[ip * per_pix + k]));
            d = std::max(d, std::fabs(per[ip * per_pix + k] - batched[ip * per_pix + k]));
        }
        if (m > 0) worst = std::max(worst, d / m);
    }
    if (!(worst <= (mixed ? 1e-5 : 1e-12))) errors++;
    std::printf("perpixel: %d pixels, %d threads: batched %.4f s, per-pixel calls %.4f s, launches %llu, "
                "max rel diff %.3g\n",
                in.npix, n_threads, s_batched, s_pixels, (unsigned long long)teufel_gpu::launch_count(), worst);
    return errors;
}

int run_screens(const char* input, const char* mesh_out, const char* screen_out)
{
    Input in;
    if (int rc = read_input(input, in)) return rc;
    const int32_t np = in.np, npix = in.npix, nt = in.nt, nx = in.nx, ny = in.ny, nots = in.nots;
    const double dt = in.dt, st0 = in.st0, sdt = in.sdt, dtx = in.dtx, dty = in.dty;
    const std::vector<Vector>& pts = in.pts;
    const std::vector<Vector>& nrm = in.nrm;
    const std::vector<double>& t0 = in.t0;
    const std::vector<double>& area = in.area;
    const Vector origin = in.origin, dx = in.dx, dy = in.dy;
    auto make = [&](int i) { return make_particle(in, i); };
    // a Beam of two bunches (first half, second half) and one Bunch of all
    Beam* beam = new Beam();
    Bunch* b1 = new Bunch();
    Bunch* b2 = new Bunch();
    Bunch* all = new Bunch();
    for (int i = 0; i < np; i++) {
        (i < (np + 1) / 2 ? b1 : b2)->Add(make(i));
        all->Add(make(i));
    }
    beam->Add(b1);
    beam->Add(b2);

    MeshedScreen<Beam> mesh(pts, t0, dt, nt, area, nrm, mesh_out);
    mesh.setSource(beam);
    mesh.integrate();
    mesh.generateOutput();
    std::printf("mesh: %d points x %d samples, totalEnergy = %.17g J\n", npix, nt, mesh.totalEnergy());

    ScreenObserver<Bunch> screen(screen_out, origin, dx, dy, (unsigned)nx, (unsigned)ny, st0, sdt, dtx, dty,
                                 (unsigned)nots);
    screen.setSource(all);
    screen.integrate();
    screen.generateOutput();
    std::printf("screen: %d x %d cells x %d samples\n", nx, ny, nots);

    int errors = 0;
    // integrate() accumulates: a second pass doubles the traces
    std::unique_ptr<double[]> once(screen.getBuffer());
    screen.integrate();
    std::unique_ptr<double[]> twice(screen.getBuffer());
    for (size_t i = 0; i < screen.getBufferSize(); i++)
        if (std::fabs(twice[i] - 2 * once[i]) > 1e-12 * std::fabs(2 * once[i]) + 1e-300) {
            errors++;
            std::printf("screen: second integrate() does not accumulate at %zu\n", i);
            break;
        }
    screen.fromBuffer(once.get(), screen.getBufferSize());
    if (screen.getField(0, 0, 0).E().x != once[0]) errors++;
    delete beam;
    delete all;
    return errors;
}

// ---- screen files ------------------------------------------------------------

int run_h5info(const char* path)
{
    MeshedScreen<Bunch> mesh(path);
    mesh.init();
    H5LiteFile f(path);
    const std::vector<double> cp = f.read_double("MeshCornerPoints");
    const std::vector<int64_t> tri = f.read_int("MeshTriangles");
    double worst = 0.0;
    for (int ip = 0; ip < mesh.get_Np(); ip++) {
        Vector c;
        for (int k = 0; k < 3; k++) {
            const int64_t v = tri[3 * ip + k];
            c += Vector(cp[3 * v], cp[3 * v + 1], cp[3 * v + 2]) / 3.0;
        }
        worst = std::max(worst, (c - mesh.get_point(ip)).norm());
    }
    double t0sum = 0.0;
    for (int ip = 0; ip < mesh.get_Np(); ip++) t0sum += mesh.get_t0(ip);
    const size_t last = cp.size() - 3;
    std::printf("t0sum %.17g cp_last %.17g %.17g %.17g tri_last %lld\n", t0sum, cp[last], cp[last + 1], cp[last + 2],
                (long long)tri.back());
    std::printf("Ncp %d Np %d Nt %d dt %.17g t0_0 %.17g area %.17g centroid_dev %.3g normal_0 %.6f %.6f %.6f\n",
                mesh.get_Ncp(), mesh.get_Np(), mesh.get_Nt(), mesh.get_dt(), mesh.get_t0(0), mesh.get_total_area(),
                worst, mesh.get_normal(0).x, mesh.get_normal(0).y, mesh.get_normal(0).z);
    return 0;
}

// read a screen file, give every trace sample a known value, write it with
// writeFile and read it back
int run_h5write(const char* in, const char* out)
{
    MeshedScreen<Bunch> a(in);
    for (int ip = 0; ip < a.get_Np(); ip++) {
        double* d = a.get_trace(ip)->data();
        for (int k = 0; k < 6 * a.get_Nt(); k++) d[k] = ip + 1e-3 * k;
    }
    a.writeFile(out);
    MeshedScreen<Bunch> b(out);
    int errors = 0;
    if (b.get_Np() != a.get_Np() || b.get_Nt() != a.get_Nt() || b.get_dt() != a.get_dt() || b.get_Ncp() != a.get_Ncp())
        errors++;
    for (int ip = 0; ip < a.get_Np() && !errors; ip++) {
        if (!(b.get_point(ip).x == a.get_point(ip).x && b.get_t0(ip) == a.get_t0(ip))) errors++;
        const double *x = a.get_trace(ip)->data(), *y = b.get_trace(ip)->data();
        for (int k = 0; k < 6 * a.get_Nt(); k++)
            if (x[k] != y[k]) {
                errors++;
                break;
            }
    }
    a.init();
    b.init();
    if (a.get_total_area() != b.get_total_area()) errors++;
    std::printf("h5write: %d points x %d samples written and read back, errors %d\n", b.get_Np(), b.get_Nt(), errors);
    return errors;
}

int run_diffraction(const char* h5, const char* traj, const char* out, bool check_energy = true)
{
    FILE* f = std::fopen(traj, "rb");
    if (!f) return 101;
    Reader r{f};
    int32_t np = r.one<int32_t>(), npts = r.one<int32_t>();
    if (!r.ok || np != 1 || npts < 2) {
        std::fclose(f);
        return 102;
    }
    std::vector<double> a(10 * (size_t)npts);
    r.get(a.data(), a.size());
    double q = r.one<double>();
    std::fclose(f);
    if (!r.ok) return 104;
    ChargedParticle* e = new ChargedParticle(q, -q);
    const size_t n = (size_t)npts;
    for (size_t k = 0; k < n; k++) {
        Vector X(a[n + k], a[2 * n + k], a[3 * n + k]), P(a[4 * n + k], a[5 * n + k], a[6 * n + k]),
            A(a[7 * n + k], a[8 * n + k], a[9 * n + k]);
        if (k == 0)
            e->initTrajectory(a[k], X, P, A);
        else
            e->storeTrajectoryPoint(a[k], X, P, A);
    }
    Bunch bunch;
    bunch.Add(e);
    MeshedScreen<Bunch> screen(h5);
    screen.init();
    screen.zero();
    screen.setSource(&bunch);
    screen.integrate();
    const double EuJ = screen.totalEnergy() * 1.0e6;
    std::printf("radiation energy = %.17g uJ%s\n", EuJ, check_energy ? " (reference test: 0.463 +- 0.020 uJ)" : "");
    std::fflush(stdout);
    screen.writeReport(&std::cout);
    // the dump takes the screen's own file name; write it where asked
    MeshedScreen<Bunch> copy(h5);
    std::unique_ptr<double[]> buf(screen.getBuffer());
    copy.fromBuffer(buf.get(), screen.getBufferSize());
    std::vector<Vector> pts;
    std::vector<FieldTrace*> tr;
    for (int ip = 0; ip < copy.get_Np(); ip++) {
        pts.push_back(copy.get_point(ip));
        tr.push_back(copy.get_trace(ip));
    }
    writeDump(out, pts, tr);
    return check_energy && std::fabs(EuJ - 0.463) > 0.020;
}

// ScreenObserver HDF5 file written by generateOutput and read back
int run_screenh5(const char* out)
{
    const Vector origin(0.01, -0.02, 10.0), dx(0.004, 0.0, 0.0005), dy(0.0, 0.005, 0.0);
    const unsigned nx = 9, ny = 7, nots = 16;
    const double t0 = 3.3e-8, dt = 2.5e-15, dtx = 1e-15, dty = -2e-15;
    ScreenObserver<Bunch> scr(out, origin, dx, dy, nx, ny, t0, dt, dtx, dty, nots);
    auto value = [&](unsigned ix, unsigned iy, unsigned it, int c) { return ix * 1000.0 + iy * 10.0 + it * 1e-3 + c * 1e-6; };
    for (unsigned ix = 0; ix < nx; ix++)
        for (unsigned iy = 0; iy < ny; iy++)
            for (unsigned it = 0; it < nots; it++)
                scr.setField(ix, iy, it,
                             ElMagField(Vector(value(ix, iy, it, 0), value(ix, iy, it, 1), value(ix, iy, it, 2)),
                                        Vector(value(ix, iy, it, 3), value(ix, iy, it, 4), value(ix, iy, it, 5))));
    scr.generateOutput();
    const ScreenFile f = readScreenFile(out);
    int errors = 0;
    auto same = [](const Vector& a, const Vector& b) { return a.x == b.x && a.y == b.y && a.z == b.z; };
    if (f.Nx != nx || f.Ny != ny || f.NOTS != nots) errors++;
    if (!same(f.origin, origin) || !same(f.dX, dx) || !same(f.dY, dy) || !same(f.normal, scr.getNormal())) errors++;
    if (f.t0 != t0 || f.dt != dt || f.dtx != dtx || f.dty != dty) errors++;
    // normal = -(dX x dY) / |dX x dY| (src/screen.cpp:51-52)
    const Vector n = cross(dx, dy) * -1.0;
    if (std::fabs(f.normal.x - n.x / n.norm()) > 1e-16 || std::fabs(f.normal.z - n.z / n.norm()) > 1e-16) errors++;
    for (unsigned ix = 0; ix < nx && errors == 0; ix++)
        for (unsigned iy = 0; iy < ny; iy++) {
            if (!same(f.position[ix * ny + iy], scr.CellPosition(ix, iy))) errors++;
            for (unsigned it = 0; it < nots; it++)
                for (int c = 0; c < 6; c++)
                    if (f.field[((size_t)(ix * ny + iy) * nots + it) * 6 + c] != value(ix, iy, it, c)) errors++;
        }
    std::printf("screenh5: %u x %u cells x %u samples written and read back, errors %d\n", nx, ny, nots, errors);
    return errors;
}

// PointObserver SDDS file written by generateOutput; the field of sample it,
// component c, is it * 1000 + c * 0.001 - 1/3 (1/3 rounded: a value that
// needs all 17 digits)
int run_pointsdds(const char* out)
{
    const unsigned nots = 37;
    PointObserver<Bunch> obs(out, Vector(0.01, -0.02, 10.0), 3.3e-8, 2.5e-15, nots);
    auto value = [](unsigned it, int c) { return it * 1000.0 + c * 0.001 - 1.0 / 3.0; };
    for (unsigned it = 0; it < nots; it++)
        obs.setField(it, ElMagField(Vector(value(it, 0), value(it, 1), value(it, 2)),
                                    Vector(value(it, 3), value(it, 4), value(it, 5))));
    obs.generateOutput();
    std::printf("pointsdds: %u samples written\n", nots);
    return 0;
}

// SnapshotObserver HDF5 file written by generateOutput and read back
int run_snaph5(const char* out)
{
    const Vector origin(-0.02, 0.01, 7.5), dx(0.003, 0.0, 0.001), dy(0.0, 0.004, 0.0);
    const unsigned nx = 6, ny = 5;
    const double t = 2.7e-8;
    SnapshotObserver<Bunch> snap(out, origin, dx, dy, nx, ny, t);
    auto value = [](unsigned ix, unsigned iy, int c) { return ix * 100.0 + iy + c * 1e-3 - 0.5; };
    for (unsigned ix = 0; ix < nx; ix++)
        for (unsigned iy = 0; iy < ny; iy++)
            snap.setField(ix, iy, ElMagField(Vector(value(ix, iy, 0), value(ix, iy, 1), value(ix, iy, 2)),
                                             Vector(value(ix, iy, 3), value(ix, iy, 4), value(ix, iy, 5))));
    snap.generateOutput();
    const H5LiteFile f(out);
    int errors = 0;
    const std::vector<uint64_t> pd = f.dims("ObservationPosition"), fd = f.dims("ElMagField");
    if (pd != std::vector<uint64_t>{nx, ny, 3} || fd != std::vector<uint64_t>{nx, ny, 6}) errors++;
    if (f.attr_int("ObservationPosition", "Nx") != nx || f.attr_int("ObservationPosition", "Ny") != ny) errors++;
    if (f.attr_double("ElMagField", "t0") != t) errors++;
    if (f.exists("ParticleCoordinates")) errors++;
    const std::vector<double> pos = f.read_double("ObservationPosition"), fld = f.read_double("ElMagField");
    for (unsigned ix = 0; ix < nx && errors == 0; ix++)
        for (unsigned iy = 0; iy < ny; iy++) {
            const Vector p = snap.CellPosition((int)ix, (int)iy);
            const double* q = &pos[(ix * ny + iy) * 3];
            if (q[0] != p.x || q[1] != p.y || q[2] != p.z) errors++;
            for (int c = 0; c < 6; c++)
                if (fld[(ix * ny + iy) * 6 + c] != value(ix, iy, c)) errors++;
        }
    std::printf("snaph5: %u x %u cells written and read back, errors %d\n", nx, ny, errors);
    return errors;
}

// a ScreenObserver file as a plain dump: positions, per-cell t0
// (src/screen.cpp:106-110), samples
int run_screenread(const char* in, const char* out)
{
    const ScreenFile f = readScreenFile(in);
    std::vector<Vector> pts;
    std::vector<FieldTrace*> tr;
    for (unsigned ix = 0; ix < f.Nx; ix++)
        for (unsigned iy = 0; iy < f.Ny; iy++) {
            const double t0 = f.t0 + f.dtx * (double)((int)ix - (int)f.Nx / 2) + f.dty * (double)((int)iy - (int)f.Ny / 2);
            FieldTrace* t = new FieldTrace(t0, f.dt, f.NOTS);
            std::memcpy(t->data(), &f.field[(size_t)(ix * f.Ny + iy) * f.NOTS * 6], sizeof(double) * 6 * f.NOTS);
            pts.push_back(f.position[ix * f.Ny + iy]);
            tr.push_back(t);
        }
    writeDump(out, pts, tr);
    for (FieldTrace* t : tr) delete t;
    std::printf("screenread: %u x %u cells x %u samples\n", f.Nx, f.Ny, f.NOTS);
    return 0;
}

}  // namespace

int main(int argc, char** argv)
{
    if (argc < 2) {
        std::fprintf(stderr, "usage: %s kat|loop|screens ...\n", argv[0]);
        return 100;
    }
    const std::string mode = argv[1];
    try {
        int errors;
        if (mode == "kat") {
            if (set_precision(argc > 2 ? argv[2] : nullptr)) return 100;
            errors = run_kat();
        } else if (mode == "loop") {
            if (set_precision(argc > 2 ? argv[2] : nullptr)) return 100;
            errors = run_loop();
        } else if (mode == "h5info" && argc >= 3) {
            return run_h5info(argv[2]);
        } else if (mode == "h5write" && argc >= 4) {
            return run_h5write(argv[2], argv[3]);
        } else if (mode == "screenh5" && argc >= 3) {
            return run_screenh5(argv[2]);
        } else if (mode == "screenread" && argc >= 4) {
            return run_screenread(argv[2], argv[3]);
        } else if (mode == "pointsdds" && argc >= 3) {
            return run_pointsdds(argv[2]);
        } else if (mode == "snaph5" && argc >= 3) {
            return run_snaph5(argv[2]);
        } else if (mode == "diffraction" && argc >= 5) {
            if (set_precision(argc > 5 ? argv[5] : nullptr)) return 100;
            errors = run_diffraction(argv[2], argv[3], argv[4]);
        } else if (mode == "meshscreen" && argc >= 5) {
            if (set_precision(argc > 5 ? argv[5] : nullptr)) return 100;
            errors = run_diffraction(argv[2], argv[3], argv[4], false);
        } else if (mode == "perpixel" && argc >= 4) {
            if (set_precision(argc > 4 ? argv[4] : nullptr)) return 100;
            errors = run_perpixel(argv[2], std::atoi(argv[3]), argc > 4 && std::strcmp(argv[4], "mixed") == 0);
        } else if (mode == "screens" && argc >= 5) {
            if (set_precision(argc > 5 ? argv[5] : nullptr)) return 100;
            errors = run_screens(argv[2], argv[3], argv[4]);
        } else {
            std::fprintf(stderr, "bad arguments\n");
            return 100;
        }
        teufel_gpu::shutdown();
        std::printf("%s: errors %d - %s\n", mode.c_str(), errors, errors ? "test failed" : "OK");
        return errors;
    } catch (const std::exception& e) {
        std::printf("%s: exception: %s\n", mode.c_str(), e.what());
        return 110;
    }
}
