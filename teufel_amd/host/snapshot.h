// snapshot.h - SnapshotObserver<SourceT> (interface of src/snapshot.h:64-218):
// the field on an Nx x Ny grid of points at one time, from each particle's
// trajectory point on the observer's past light cone (RetardedField(t, x)).
// integrate() is one GPU call for all cells instead of the reference's
// OpenMP cell loop (src/snapshot.cpp:72-91).  SourceT is Beam or Bunch.
#pragma once

#include <string>
#include <vector>

#include "observer.h"

template <typename SourceT>
class SnapshotObserver : public Observer<SourceT> {
public:
    // cells at position + dx (ix - (nx-1)/2) + dy (iy - (ny-1)/2)
    // (src/snapshot.cpp:405-408), observed at time t
    SnapshotObserver(std::string filename, Vector position, Vector dx, Vector dy, unsigned int nx, unsigned int ny,
                     double t);
    ~SnapshotObserver() override {}

    Vector CellPosition(int ix, int iy) const;
    double getTime() const { return t_obs; }
    ElMagField getField(unsigned int ix, unsigned int iy) const;
    void setField(unsigned int ix, unsigned int iy, ElMagField field);

    void integrate() override;
    // [ix][iy][6] doubles (src/snapshot.cpp:178-223)
    std::size_t getBufferSize() override { return (std::size_t)Nx * Ny * 6; }
    double* getBuffer() override;
    void fromBuffer(double* buffer, std::size_t size) override;
    // the HDF5 file of src/snapshot.cpp:226-392: ObservationPosition
    // [Nx][Ny][3] (int attributes Nx, Ny) and ElMagField [Nx][Ny][6] (double
    // attribute t0, the observation time).  The reference's optional
    // ParticleCoordinates dataset holds particle positions the tracking
    // stored at that time; there is no tracking here, so it is not written.
    void WriteFieldHDF5();
    // WriteFieldHDF5 (src/snapshot.cpp:398-402)
    void generateOutput() override;

private:
    std::string FileName;
    Vector O, dX, dY;
    unsigned int Nx, Ny;
    double t_obs;
    std::vector<ElMagField> FieldArray;  // [ix * Ny + iy]
};
