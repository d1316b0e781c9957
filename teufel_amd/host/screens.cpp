// screens.cpp - MeshedScreen, ScreenObserver and PointObserver for Beam and Bunch.
#include "screens.h"

#include <ostream>

#include <cstdint>
#include <cstdio>
#include <cstring>

#include "gpu_radiation.h"
#include "h5lite.h"

void writeDump(const std::string& filename, const std::vector<Vector>& points, const std::vector<FieldTrace*>& traces)
{
    if (filename.empty()) return;
    FILE* f = std::fopen(filename.c_str(), "wb");
    if (!f) throw IOexception("writeDump - cannot open " + filename);
    const char magic[8] = {'T', 'G', 'D', 'U', 'M', 'P', '1', 0};
    int32_t np = (int32_t)points.size(), ns = traces.empty() ? 0 : (int32_t)traces[0]->get_N();
    double dt = traces.empty() ? 0.0 : traces[0]->get_dt();
    bool ok = std::fwrite(magic, 1, 8, f) == 8 && std::fwrite(&np, 4, 1, f) == 1 && std::fwrite(&ns, 4, 1, f) == 1 &&
              std::fwrite(&dt, 8, 1, f) == 1;
    for (size_t i = 0; ok && i < points.size(); i++) {
        double head[4] = {points[i].x, points[i].y, points[i].z, traces[i]->get_t0()};
        ok = std::fwrite(head, 8, 4, f) == 4 &&
             std::fwrite(traces[i]->data(), 8, 6 * traces[i]->get_N(), f) == 6 * traces[i]->get_N();
    }
    std::fclose(f);
    if (!ok) throw IOexception("writeDump - write error on " + filename);
}

ScreenFile readScreenFile(const std::string& filename)
{
    H5LiteFile f(filename);
    const char* who = "readScreenFile - ";
    for (const char* d : {"Screen", "ObservationPosition", "ElMagField"})
        if (!f.exists(d)) throw IOexception(std::string(who) + filename + " has no dataset " + d);
    ScreenFile s;
    const int64_t nx = f.attr_int("Screen", "Nx"), ny = f.attr_int("Screen", "Ny"), nots = f.attr_int("Screen", "NOTS");
    if (nx < 0 || ny < 0 || nots < 0 || nx > (1 << 20) || ny > (1 << 20) || nots > (1 << 30))
        throw IOexception(std::string(who) + "implausible screen size in " + filename);
    s.Nx = (unsigned)nx;
    s.Ny = (unsigned)ny;
    s.NOTS = (unsigned)nots;
    if (f.dims("Screen") != std::vector<uint64_t>{4, 3} ||
        f.dims("ObservationPosition") != std::vector<uint64_t>{s.Nx, s.Ny, 3} ||
        f.dims("ElMagField") != std::vector<uint64_t>{s.Nx, s.Ny, s.NOTS, 6})
        throw IOexception(std::string(who) + "dataset shapes do not match Nx, Ny, NOTS in " + filename);
    const std::vector<double> g = f.read_double("Screen");
    s.origin = Vector(g[0], g[1], g[2]);
    s.normal = Vector(g[3], g[4], g[5]);
    s.dX = Vector(g[6], g[7], g[8]);
    s.dY = Vector(g[9], g[10], g[11]);
    s.t0 = f.attr_double("Screen", "t0");
    s.dt = f.attr_double("Screen", "dt");
    s.dtx = f.attr_double("Screen", "dtx");
    s.dty = f.attr_double("Screen", "dty");
    const std::vector<double> p = f.read_double("ObservationPosition");
    for (size_t i = 0; i + 2 < p.size(); i += 3) s.position.emplace_back(p[i], p[i + 1], p[i + 2]);
    s.field = f.read_double("ElMagField");
    return s;
}

namespace {
double* traces_to_buffer(const std::vector<FieldTrace*>& tr, std::size_t n)
{
    double* buffer = new double[n];
    double* bp = buffer;
    for (const FieldTrace* t : tr) {
        std::memcpy(bp, t->data(), 6 * t->get_N() * sizeof(double));
        bp += 6 * t->get_N();
    }
    return buffer;
}

void buffer_to_traces(const double* buffer, std::size_t size, std::size_t expect, const std::vector<FieldTrace*>& tr,
                      const char* who)
{
    if (size != expect) throw IOexception(std::string(who) + "::fromBuffer() - buffer size mismatch.");
    const double* bp = buffer;
    for (FieldTrace* t : tr) {
        std::memcpy(t->data(), bp, 6 * t->get_N() * sizeof(double));
        bp += 6 * t->get_N();
    }
}

template <typename SourceT>
void integrate_points(SourceT* source, const std::vector<Vector>& points, const std::vector<FieldTrace*>& traces)
{
    if (!source) throw RANGEexception("Observer::integrate() - no source set.");
    std::vector<ChargedParticle*> ps;
    collect_particles(source, ps);
    teufel_gpu::integrate(ps, points, traces);
}
}  // namespace

// ---- MeshedScreen ----------------------------------------------------------

template <typename SourceT>
MeshedScreen<SourceT>::MeshedScreen(const std::vector<Vector>& pts, const std::vector<double>& t0, double dt_,
                                    int Nt_, const std::vector<double>& area_, const std::vector<Vector>& normal_,
                                    const std::string& filename)
    : Np((int)pts.size()), Nt(Nt_), dt(dt_), points(pts), area(area_), normal(normal_), FileName(filename)
{
    if (t0.size() != pts.size()) throw IOexception("MeshedScreen - points/t0 size mismatch.");
    if (Nt < 0) throw IOexception("MeshedScreen - negative trace length.");
    for (int i = 0; i < Np; i++) A.push_back(new FieldTrace(t0[i], dt, (std::size_t)Nt));
}

template <typename SourceT>
MeshedScreen<SourceT>::MeshedScreen(const std::string& filename) : Np(0), Nt(0), dt(0.0), FileName(filename)
{
    H5LiteFile f(filename);
    const int64_t ncp = f.attr_int("MeshCornerPoints", "Ncp");
    const std::vector<double> cp = f.read_double("MeshCornerPoints");
    if (ncp < 0 || cp.size() != 3 * (size_t)ncp) throw IOexception("MeshedScreen - MeshCornerPoints size mismatch.");
    for (int64_t i = 0; i < ncp; i++) corners.push_back(Vector(cp[3 * i], cp[3 * i + 1], cp[3 * i + 2]));
    const int64_t np = f.attr_int("ObservationPosition", "Np");
    const std::vector<double> op = f.read_double("ObservationPosition");
    if (np < 0 || op.size() != 3 * (size_t)np) throw IOexception("MeshedScreen - ObservationPosition size mismatch.");
    Np = (int)np;
    for (int i = 0; i < Np; i++) points.push_back(Vector(op[3 * i], op[3 * i + 1], op[3 * i + 2]));
    const std::vector<int64_t> tri = f.read_int("MeshTriangles");
    if (tri.size() != 3 * (size_t)Np) throw IOexception("MeshedScreen - MeshTriangles size mismatch.");
    for (int64_t v : tri) {
        if (v < 0 || v >= ncp) throw IOexception("MeshedScreen - triangle corner index out of range.");
        triangles.push_back((int)v);
    }
    std::vector<double> t0(Np, 0.0);
    if (f.exists("ObservationTime")) {
        const int64_t nt = f.attr_int("ObservationTime", "Nt");
        if (nt < 0) throw IOexception("MeshedScreen - negative Nt.");
        Nt = (int)nt;
        dt = f.attr_double("ObservationTime", "dt");
        t0 = f.read_double("ObservationTime");
        if (t0.size() != (size_t)Np) throw IOexception("MeshedScreen - ObservationTime size mismatch.");
        // the traces are allocated from Np and the Nt attribute: a file
        // declaring more than 64 GiB of them is taken as corrupt rather than
        // exhausting the host's memory
        if ((double)Np * (double)Nt * 48.0 > 68719476736.0)
            throw IOexception("MeshedScreen - the file declares " + std::to_string((double)Np * Nt * 48.0 / 1e9) +
                              " GB of traces (Np x Nt), more than 64 GiB.");
        if (f.exists("ElMagField") && f.dims("ElMagField") != std::vector<uint64_t>{(uint64_t)Np, (uint64_t)Nt, 6})
            throw IOexception("MeshedScreen - ElMagField size mismatch.");
    }
    for (int i = 0; i < Np; i++) A.push_back(new FieldTrace(t0[i], dt, (std::size_t)Nt));
    if (f.exists("ElMagField")) {
        const std::vector<double> fld = f.read_double("ElMagField");
        if (fld.size() != (size_t)Np * Nt * 6) throw IOexception("MeshedScreen - ElMagField size mismatch.");
        for (int i = 0; i < Np; i++)
            std::memcpy(A[i]->data(), &fld[(size_t)i * Nt * 6], (size_t)Nt * 6 * sizeof(double));
    }
}

template <typename SourceT>
void MeshedScreen<SourceT>::init()
{
    if (triangles.empty()) return;
    area.assign(Np, 0.0);
    normal.assign(Np, Vector());
    for (int ip = 0; ip < Np; ip++) {
        const Vector p1 = corners[triangles[3 * ip]], p2 = corners[triangles[3 * ip + 1]],
                     p3 = corners[triangles[3 * ip + 2]];
        Vector n = cross(p3 - p1, p2 - p1);
        area[ip] = n.norm() * 0.5;
        normal[ip] = n / n.norm();
    }
}

template <typename SourceT>
void MeshedScreen<SourceT>::writeFile(const std::string& filename) const
{
    if (triangles.empty()) throw IOexception("MeshedScreen::writeFile - the screen has no mesh (not read from a file).");
    H5LiteWriter w;
    std::vector<double> cp, op, t0, fld;
    for (const Vector& v : corners) cp.insert(cp.end(), {v.x, v.y, v.z});
    for (const Vector& v : points) op.insert(op.end(), {v.x, v.y, v.z});
    for (const FieldTrace* t : A) t0.push_back(t->get_t0());
    w.add_double("MeshCornerPoints", {corners.size(), 3}, cp);
    w.attr_int32("MeshCornerPoints", "Ncp", (int32_t)corners.size());
    w.add_double("ObservationPosition", {(uint64_t)Np, 3}, op);
    w.attr_int32("ObservationPosition", "Np", Np);
    w.add_int32("MeshTriangles", {(uint64_t)Np, 3}, std::vector<int32_t>(triangles.begin(), triangles.end()));
    w.attr_int32("MeshTriangles", "Ntri", Np);
    w.add_double("ObservationTime", {(uint64_t)Np}, t0);
    w.attr_int32("ObservationTime", "Nt", Nt);
    w.attr_double("ObservationTime", "dt", dt);
    if (Nt > 0) {
        fld.reserve((size_t)Np * Nt * 6);
        for (const FieldTrace* t : A) fld.insert(fld.end(), t->data(), t->data() + 6 * (size_t)Nt);
        w.add_double("ElMagField", {(uint64_t)Np, (uint64_t)Nt, 6}, fld);
    }
    w.write(filename);
}

template <typename SourceT>
void MeshedScreen<SourceT>::generateOutput()
{
    if (!triangles.empty())
        writeFile(FileName);
    else
        writeDump(FileName, points, A);
}

template <typename SourceT>
double MeshedScreen<SourceT>::get_total_area() const
{
    double total = 0.0;
    for (double a : area) total += a;
    return total;
}

template <typename SourceT>
MeshedScreen<SourceT>::~MeshedScreen()
{
    for (FieldTrace* t : A) delete t;
}

template <typename SourceT>
void MeshedScreen<SourceT>::zero()
{
    for (FieldTrace* t : A) t->zero();
}

template <typename SourceT>
double MeshedScreen<SourceT>::totalEnergy() const
{
    if ((int)area.size() != Np || (int)normal.size() != Np)
        throw IOexception("MeshedScreen::totalEnergy - cell areas and normals not set.");
    double total = 0.0;
    for (int ip = 0; ip < Np; ip++) total -= area[ip] * dot(A[ip]->Poynting(), normal[ip]);
    return total;
}

template <typename SourceT>
void MeshedScreen<SourceT>::writeReport(std::ostream* st) const
{
    std::ostream& o = *st;
    const std::streamsize keep = o.precision();
    o << "\n   Meshed Screen\n----------------------\n";
    o << "Np=" << Np << "  Nt=" << Nt << "\n";
    o.precision(6);
    o << "total area = " << get_total_area() * 1.0e4 << " cm²\n";
    o << "Energy incident on screen = " << totalEnergy() * 1.0e6 << " µJ\n";
    int peak_ip = -1;
    double peak = 0.0;
    Vector S_peak;
    for (int ip = 0; ip < Np; ip++) {
        const Vector S = A[ip]->Poynting();
        if (S.norm() > peak) {
            peak = S.norm();
            peak_ip = ip;
            S_peak = S;
        }
    }
    if (peak_ip >= 0) {
        const Vector& p = points[peak_ip];
        o << "peak pos[" << peak_ip << "] =   (" << p.x << ", " << p.y << ", " << p.z << ") m   \n";
        o << "peak energy density = " << peak << " J/m²\n";
        o << "Poynting vector S = (" << S_peak.x << ", " << S_peak.y << ", " << S_peak.z << ") J/m²\n";
        o << "t0  = " << A[peak_ip]->get_t0() * 1e9 << " ns  \n";
        o << "dt  = " << dt * 1.0e9 << " ns\n";
    }
    o << std::endl;
    o.precision(keep);
}

template <typename SourceT>
void MeshedScreen<SourceT>::integrate()
{
    integrate_points(this->source, points, A);
}

template <typename SourceT>
double* MeshedScreen<SourceT>::getBuffer()
{
    return traces_to_buffer(A, getBufferSize());
}

template <typename SourceT>
void MeshedScreen<SourceT>::fromBuffer(double* buffer, std::size_t size)
{
    buffer_to_traces(buffer, size, getBufferSize(), A, "MeshedScreen");
}

// ---- ScreenObserver --------------------------------------------------------

template <typename SourceT>
ScreenObserver<SourceT>::ScreenObserver(std::string filename, Vector position, Vector dx, Vector dy, unsigned int nx,
                                        unsigned int ny, double t0, double dt, double dtx, double dty,
                                        unsigned int nots)
    : FileName(filename), Origin(position), dX(dx), dY(dy), Nx(nx), Ny(ny), NOTS(nots), t0_obs(t0), dt_obs(dt),
      dtx_obs(dtx), dty_obs(dty)
{
    normal = cross(dX, dY) * -1.0;
    normal.normalize();
    for (unsigned int ix = 0; ix < Nx; ix++)
        for (unsigned int iy = 0; iy < Ny; iy++) Traces.push_back(new FieldTrace(CellTimeZero(ix, iy), dt_obs, NOTS));
}

template <typename SourceT>
ScreenObserver<SourceT>::~ScreenObserver()
{
    for (FieldTrace* t : Traces) delete t;
}

template <typename SourceT>
Vector ScreenObserver<SourceT>::CellPosition(unsigned int ix, unsigned int iy) const
{
    return Origin + dX * (double)((int)ix - (int)Nx / 2) + dY * (double)((int)iy - (int)Ny / 2);
}

template <typename SourceT>
double ScreenObserver<SourceT>::CellTimeZero(unsigned int ix, unsigned int iy) const
{
    return t0_obs + dtx_obs * (double)((int)ix - (int)Nx / 2) + dty_obs * (double)((int)iy - (int)Ny / 2);
}

template <typename SourceT>
ElMagField ScreenObserver<SourceT>::getField(unsigned int ix, unsigned int iy, unsigned int it)
{
    if (ix >= Nx || iy >= Ny || it >= NOTS) throw IOexception("ScreenObserver::getField() - index out of range.");
    return Traces[ix * Ny + iy]->get_field(it);
}

template <typename SourceT>
void ScreenObserver<SourceT>::setField(unsigned int ix, unsigned int iy, unsigned int it, ElMagField field)
{
    if (ix >= Nx || iy >= Ny || it >= NOTS) throw IOexception("ScreenObserver::setField() - index out of range.");
    Traces[ix * Ny + iy]->set_field(it, field);
}

template <typename SourceT>
void ScreenObserver<SourceT>::integrate()
{
    std::vector<Vector> pts;
    for (unsigned int ix = 0; ix < Nx; ix++)
        for (unsigned int iy = 0; iy < Ny; iy++) pts.push_back(CellPosition(ix, iy));
    integrate_points(this->source, pts, Traces);
}

template <typename SourceT>
double* ScreenObserver<SourceT>::getBuffer()
{
    return traces_to_buffer(Traces, getBufferSize());
}

template <typename SourceT>
void ScreenObserver<SourceT>::fromBuffer(double* buffer, std::size_t size)
{
    buffer_to_traces(buffer, size, getBufferSize(), Traces, "ScreenObserver");
}

template <typename SourceT>
void ScreenObserver<SourceT>::WriteTimeDomainFieldHDF5()
{
    H5LiteWriter w;
    const int32_t nx = (int32_t)Nx, ny = (int32_t)Ny, nots = (int32_t)NOTS;
    std::vector<double> pos;
    pos.reserve((size_t)Nx * Ny * 3);
    for (unsigned int ix = 0; ix < Nx; ix++)
        for (unsigned int iy = 0; iy < Ny; iy++) {
            const Vector p = CellPosition(ix, iy);
            pos.insert(pos.end(), {p.x, p.y, p.z});
        }
    w.add_double("ObservationPosition", {Nx, Ny, 3}, pos);
    w.attr_int32("ObservationPosition", "Nx", nx);
    w.attr_int32("ObservationPosition", "Ny", ny);
    w.add_double("Screen", {4, 3},
                 {Origin.x, Origin.y, Origin.z, normal.x, normal.y, normal.z, dX.x, dX.y, dX.z, dY.x, dY.y, dY.z});
    w.attr_int32("Screen", "Nx", nx);
    w.attr_int32("Screen", "Ny", ny);
    w.attr_int32("Screen", "NOTS", nots);
    w.attr_double("Screen", "t0", t0_obs);
    w.attr_double("Screen", "dt", dt_obs);
    w.attr_double("Screen", "dtx", dtx_obs);
    w.attr_double("Screen", "dty", dty_obs);
    std::vector<double> fld;
    fld.reserve(getBufferSize());
    for (const FieldTrace* t : Traces) fld.insert(fld.end(), t->data(), t->data() + 6 * (size_t)NOTS);
    w.add_double("ElMagField", {Nx, Ny, NOTS, 6}, fld);
    w.attr_double("ElMagField", "t0", t0_obs);
    w.attr_double("ElMagField", "dt", dt_obs);
    w.attr_int32("ElMagField", "NOTS", nots);
    w.write(FileName);
}

// ---- PointObserver ---------------------------------------------------------

template <typename SourceT>
PointObserver<SourceT>::PointObserver(std::string filename, Vector position, double t0, double dt, std::size_t nots)
    : FileName(filename), Pos(position), trace(new FieldTrace(t0, dt, nots))
{
}

template <typename SourceT>
PointObserver<SourceT>::~PointObserver()
{
    delete trace;
}

template <typename SourceT>
void PointObserver<SourceT>::integrate()
{
    integrate_points(this->source, std::vector<Vector>(1, Pos), std::vector<FieldTrace*>(1, trace));
}

template <typename SourceT>
double* PointObserver<SourceT>::getBuffer()
{
    return traces_to_buffer(std::vector<FieldTrace*>(1, trace), getBufferSize());
}

template <typename SourceT>
void PointObserver<SourceT>::fromBuffer(double* buffer, std::size_t size)
{
    buffer_to_traces(buffer, size, getBufferSize(), std::vector<FieldTrace*>(1, trace), "PointObserver");
}

// ASCII SDDS, version 1 of the protocol (M. Borland, "A Self-Describing File
// Protocol for Simulation Integration and Shared Postprocessors", PAC 1995):
// the "SDDS1" line, the &parameter and &column namelists, "&data
// mode=ascii", then one page: each parameter's value on a line of its own,
// the row count, and one line per row.  The layout is the one
// src/point_observer.cpp:120-182 makes with the SDDS library (names,
// symbols, units, descriptions and types of its definitions); doubles are
// written with 17 significant digits, so every value reads back exactly.
template <typename SourceT>
void PointObserver<SourceT>::WriteTimeDomainFieldSDDS()
{
    if (FileName.empty()) return;
    FILE* f = std::fopen(FileName.c_str(), "w");
    if (!f) throw IOexception("PointObserver::WriteTimeDomainFieldSDDS - cannot open " + FileName);
    const unsigned int nots = (unsigned int)trace->get_N();
    const double t0 = trace->get_t0(), dt = trace->get_dt();
    bool ok = std::fprintf(f, "SDDS1\n") > 0;
    ok = ok && std::fprintf(f, "&parameter name=NumberTimeSteps, type=long, &end\n") > 0;
    ok = ok && std::fprintf(f, "&parameter name=t0, type=double, &end\n") > 0;
    ok = ok && std::fprintf(f, "&parameter name=dt, type=double, &end\n") > 0;
    const char* col[7][3] = {{"t", "s", "time"},
                             {"Ex", "V/m", "electric field"},
                             {"Ey", "V/m", "electric field"},
                             {"Ez", "V/m", "electric field"},
                             {"Bx", "T", "magnetic field"},
                             {"By", "T", "magnetic field"},
                             {"Bz", "T", "magnetic field"}};
    for (const auto& c : col)
        ok = ok && std::fprintf(f, "&column name=%s, symbol=%s, units=%s, description=\"%s\", type=double, &end\n", c[0],
                                c[0], c[1], c[2]) > 0;
    ok = ok && std::fprintf(f, "&data mode=ascii, &end\n! page number 1\n%u\n%.17g\n%.17g\n%u\n", nots, t0, dt, nots) > 0;
    for (unsigned int i = 0; ok && i < nots; i++) {
        const ElMagField F = trace->get_field((std::size_t)i);
        const Vector E = F.E(), B = F.B();
        ok = std::fprintf(f, "%.17g %.17g %.17g %.17g %.17g %.17g %.17g\n", t0 + (double)i * dt, E.x, E.y, E.z, B.x, B.y,
                          B.z) > 0;
    }
    ok = (std::fclose(f) == 0) && ok;
    if (!ok) throw IOexception("PointObserver::WriteTimeDomainFieldSDDS - write error on " + FileName);
}

template class MeshedScreen<Beam>;
template class MeshedScreen<Bunch>;
template class ScreenObserver<Beam>;
template class ScreenObserver<Bunch>;
template class PointObserver<Beam>;
template class PointObserver<Bunch>;
