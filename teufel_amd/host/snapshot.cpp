// snapshot.cpp - SnapshotObserver for Beam and Bunch.
#include "snapshot.h"

#include "fieldtrace.h"
#include "gpu_radiation.h"
#include "h5lite.h"
#include "screens.h"

template <typename SourceT>
SnapshotObserver<SourceT>::SnapshotObserver(std::string filename, Vector position, Vector dx, Vector dy,
                                            unsigned int nx, unsigned int ny, double t)
    : FileName(filename), O(position), dX(dx), dY(dy), Nx(nx), Ny(ny), t_obs(t),
      FieldArray((std::size_t)nx * ny)
{
}

template <typename SourceT>
Vector SnapshotObserver<SourceT>::CellPosition(int ix, int iy) const
{
    return O + dX * ((double)ix - 0.5 * ((double)Nx - 1.0)) + dY * ((double)iy - 0.5 * ((double)Ny - 1.0));
}

template <typename SourceT>
ElMagField SnapshotObserver<SourceT>::getField(unsigned int ix, unsigned int iy) const
{
    if (ix >= Nx || iy >= Ny) throw IOexception("SnapshotObserver::getField() - requested index out of range.");
    return FieldArray[(std::size_t)ix * Ny + iy];
}

template <typename SourceT>
void SnapshotObserver<SourceT>::setField(unsigned int ix, unsigned int iy, ElMagField field)
{
    if (ix >= Nx || iy >= Ny) throw IOexception("SnapshotObserver::setField() - requested index out of range.");
    FieldArray[(std::size_t)ix * Ny + iy] = field;
}

template <typename SourceT>
void SnapshotObserver<SourceT>::integrate()
{
    if (!this->source) throw RANGEexception("SnapshotObserver::integrate() - no source set.");
    std::vector<ChargedParticle*> ps;
    collect_particles(this->source, ps);
    std::vector<Vector> pts;
    for (unsigned int ix = 0; ix < Nx; ix++)
        for (unsigned int iy = 0; iy < Ny; iy++) pts.push_back(CellPosition((int)ix, (int)iy));
    teufel_gpu::fields_at(ps, std::vector<double>(pts.size(), t_obs), pts, FieldArray);
}

template <typename SourceT>
double* SnapshotObserver<SourceT>::getBuffer()
{
    double* buffer = new double[getBufferSize()];
    double* bp = buffer;
    for (const ElMagField& f : FieldArray) {
        const Vector E = f.E(), B = f.B();
        *bp++ = E.x; *bp++ = E.y; *bp++ = E.z;
        *bp++ = B.x; *bp++ = B.y; *bp++ = B.z;
    }
    return buffer;
}

template <typename SourceT>
void SnapshotObserver<SourceT>::fromBuffer(double* buffer, std::size_t size)
{
    if (size != getBufferSize()) throw IOexception("SnapshotObserver::fromBuffer() - buffer size mismatch.");
    const double* bp = buffer;
    for (ElMagField& f : FieldArray) {
        f = ElMagField(Vector(bp[0], bp[1], bp[2]), Vector(bp[3], bp[4], bp[5]));
        bp += 6;
    }
}

template <typename SourceT>
void SnapshotObserver<SourceT>::WriteFieldHDF5()
{
    H5LiteWriter w;
    std::vector<double> pos;
    pos.reserve((std::size_t)Nx * Ny * 3);
    for (unsigned int ix = 0; ix < Nx; ix++)
        for (unsigned int iy = 0; iy < Ny; iy++) {
            const Vector p = CellPosition((int)ix, (int)iy);
            pos.insert(pos.end(), {p.x, p.y, p.z});
        }
    w.add_double("ObservationPosition", {Nx, Ny, 3}, pos);
    w.attr_int32("ObservationPosition", "Nx", (int32_t)Nx);
    w.attr_int32("ObservationPosition", "Ny", (int32_t)Ny);
    double* buffer = getBuffer();
    const std::vector<double> fld(buffer, buffer + getBufferSize());
    delete[] buffer;
    w.add_double("ElMagField", {Nx, Ny, 6}, fld);
    w.attr_double("ElMagField", "t0", t_obs);
    w.write(FileName);
}

template <typename SourceT>
void SnapshotObserver<SourceT>::generateOutput()
{
    WriteFieldHDF5();
}

template class SnapshotObserver<Beam>;
template class SnapshotObserver<Bunch>;
