// screens.h - the three observers of the radiation step (interfaces of
// MeshedScreen<SourceT> src/mesh.h:99-215, ScreenObserver<SourceT>
// src/screen.h:100-160 and PointObserver<SourceT> src/point_observer.h:68-110).
// integrate() hands all pixels of the observer to the GPU in one batched
// launch instead of the reference's OpenMP pixel loop (src/mesh.cpp:288-342,
// src/screen.cpp:112-180, src/point_observer.cpp:62-72).  SourceT is Beam or
// Bunch.  generateOutput() writes the reference's file layouts: HDF5 for
// MeshedScreen (read from a file) and ScreenObserver, ASCII SDDS for
// PointObserver; a MeshedScreen built from points writes a plain binary dump
// (documented in writeDump).
#pragma once

#include <ostream>
#include <string>
#include <vector>

#include "observer.h"

// write "TGDUMP1\0", int32 n_points, int32 n_samples, f64 dt, then per point
// f64 x,y,z,t0 and the samples [n_samples][6] (little endian)
void writeDump(const std::string& filename, const std::vector<Vector>& points,
               const std::vector<FieldTrace*>& traces);

template <typename SourceT>
class MeshedScreen : public Observer<SourceT> {
public:
    // Observation points (cell centres), per-point t0, common dt and Nt; the
    // cell areas and normals (towards the source) are needed only by
    // totalEnergy().  The reference reads these from an HDF5 file
    // (src/mesh.cpp:30-218).
    MeshedScreen(const std::vector<Vector>& points, const std::vector<double>& t0, double dt, int Nt,
                 const std::vector<double>& area = std::vector<double>(),
                 const std::vector<Vector>& normal = std::vector<Vector>(),
                 const std::string& filename = std::string());
    // Read a TEUFEL screen file (HDF5, src/mesh.cpp:30-218): datasets
    // MeshCornerPoints [Ncp][3], MeshTriangles [Np][3] (corner indices),
    // ObservationPosition [Np][3], optional ObservationTime [Np] (attributes
    // Nt, dt; without it the traces are empty) and ElMagField [Np][Nt][6].
    explicit MeshedScreen(const std::string& filename);
    ~MeshedScreen() override;

    // cell areas and unit normals n = (p3-p1) x (p2-p1) / |...| of the
    // triangles of a screen read from file (src/mesh.cpp:227-247); a screen
    // built from points keeps the areas and normals it was given
    void init();
    int get_Ncp() const { return (int)corners.size(); }
    double get_area(int ip) const { return area.at(ip); }
    Vector get_normal(int ip) const { return normal.at(ip); }
    double get_total_area() const;
    void zero();
    int get_Np() const { return Np; }
    int get_Nt() const { return Nt; }
    double get_dt() const { return dt; }
    double get_t0(int ip) const { return A[ip]->get_t0(); }
    Vector get_point(int ip) const { return points[ip]; }
    FieldTrace* get_trace(int ip) { return A[ip]; }
    Vector Poynting(int ip) const { return A[ip]->Poynting(); }
    // -sum_p area_p S_p . normal_p [J] (src/mesh.cpp:744-754)
    double totalEnergy() const;
    // Summary in the reference's layout (src/mesh.cpp:783-821): Np and Nt,
    // total area [cm²], incident energy [µJ], and the cell of the largest
    // |Poynting| (first one on ties) with its position, energy density, S
    // vector, t0 and dt [ns]
    void writeReport(std::ostream* st) const;

    void integrate() override;
    std::size_t getBufferSize() override { return (std::size_t)Np * Nt * 6; }
    double* getBuffer() override;
    void fromBuffer(double* buffer, std::size_t size) override;
    // HDF5 file with the datasets the file constructor reads plus ElMagField
    // [Np][Nt][6] (src/mesh.cpp:427-736); needs the mesh triangles, so only
    // for screens read from a file
    void writeFile(const std::string& filename) const;
    // writeFile(filename of the constructor) for screens read from a file
    // (src/mesh.cpp:823-826), the plain dump otherwise
    void generateOutput() override;

private:
    int Np, Nt;
    double dt;
    std::vector<Vector> points;
    std::vector<double> area;
    std::vector<Vector> normal;
    std::vector<FieldTrace*> A;
    std::string FileName;
    std::vector<Vector> corners;             // MeshCornerPoints (file screens)
    std::vector<int> triangles;              // MeshTriangles, 3 per cell
};

// The contents of a ScreenObserver HDF5 file (the layout of
// ScreenObserver::WriteTimeDomainFieldHDF5, src/screen.cpp:297-558, as
// scripts/teufel.py:23-75 reads it): dataset Screen [4][3] = origin, normal,
// dX, dY with attributes Nx, Ny, NOTS, t0, dt, dtx, dty; ObservationPosition
// [Nx][Ny][3] (attributes Nx, Ny); ElMagField [Nx][Ny][NOTS][6] (attributes
// t0, dt, NOTS).
struct ScreenFile {
    unsigned int Nx = 0, Ny = 0, NOTS = 0;
    Vector origin, normal, dX, dY;
    double t0 = 0.0, dt = 0.0, dtx = 0.0, dty = 0.0;
    std::vector<Vector> position;   // [ix * Ny + iy]
    std::vector<double> field;      // [ix][iy][it][6]
};
// Read such a file; IOexception for a file that does not hold that layout.
ScreenFile readScreenFile(const std::string& filename);

template <typename SourceT>
class ScreenObserver : public Observer<SourceT> {
public:
    // nx x ny cells at position + dx (ix - nx/2) + dy (iy - ny/2), trace start
    // t0 + dtx (ix - nx/2) + dty (iy - ny/2) (integer division, src/screen.cpp:98-110)
    ScreenObserver(std::string filename, Vector position, Vector dx, Vector dy, unsigned int nx, unsigned int ny,
                   double t0, double dt, double dtx, double dty, unsigned int nots);
    ~ScreenObserver() override;

    Vector CellPosition(unsigned int ix, unsigned int iy) const;
    double CellTimeZero(unsigned int ix, unsigned int iy) const;
    ElMagField getField(unsigned int ix, unsigned int iy, unsigned int it);
    void setField(unsigned int ix, unsigned int iy, unsigned int it, ElMagField field);

    Vector getNormal() const { return normal; }

    void integrate() override;
    std::size_t getBufferSize() override { return (std::size_t)Nx * Ny * NOTS * 6; }
    double* getBuffer() override;
    void fromBuffer(double* buffer, std::size_t size) override;
    // the HDF5 file of src/screen.cpp:297-558 (layout: ScreenFile)
    void WriteTimeDomainFieldHDF5();
    void generateOutput() override { WriteTimeDomainFieldHDF5(); }

private:
    std::string FileName;
    Vector Origin, dX, dY;
    Vector normal;   // -(dX x dY), normalised (src/screen.cpp:51-52)
    unsigned int Nx, Ny, NOTS;
    double t0_obs, dt_obs, dtx_obs, dty_obs;
    std::vector<FieldTrace*> Traces;   // [ix * Ny + iy]
};

template <typename SourceT>
class PointObserver : public Observer<SourceT> {
public:
    PointObserver(std::string filename, Vector position, double t0, double dt, std::size_t nots);
    ~PointObserver() override;

    void integrate() override;
    ElMagField getField(std::size_t idx) { return trace->get_field(idx); }
    void setField(std::size_t idx, ElMagField f) { trace->set_field(idx, f); }
    FieldTrace* getTrace() { return trace; }
    std::size_t getBufferSize() override { return 6 * trace->get_N(); }
    double* getBuffer() override;
    void fromBuffer(double* buffer, std::size_t size) override;
    // The reference's ASCII SDDS file (src/point_observer.cpp:120-182, read by
    // scripts/plot_PointObservation.py): parameters NumberTimeSteps, t0, dt;
    // columns t, Ex, Ey, Ez (V/m), Bx, By, Bz (T), one row per sample.
    void WriteTimeDomainFieldSDDS();
    void generateOutput() override { WriteTimeDomainFieldSDDS(); }

private:
    std::string FileName;
    Vector Pos;
    FieldTrace* trace;
};
