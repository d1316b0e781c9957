// fieldtrace.h - time trace of the field at one observation point (interface
// of FieldTrace, src/fieldtrace.h:38-173, for the radiation path): N samples
// from t0 with step dt; sample i is the mean field over the bucket
// [t0 + (i-1/2) dt, t0 + (i+1/2) dt).  Storage is contiguous
// (E_x, E_y, E_z, B_x, B_y, B_z) per sample, the layout the device returns.
#pragma once

#include <cstddef>
#include <vector>

#include "fields.h"

class FieldTrace {
public:
    FieldTrace(double t0, double dt, std::size_t N);
    explicit FieldTrace(const FieldTrace* original);

    double get_t0() const { return t0; }
    double get_dt() const { return dt; }
    std::size_t get_N() const { return N; }
    double get_time(std::size_t index) const;
    double get_center_time() const { return t0 + 0.5 * dt * (double)N; }

    void zero();
    ElMagField get_field(std::size_t index) const;
    void set_field(std::size_t index, ElMagField f);
    // accumulate into sample `index` (throws IOexception when out of range)
    void add(std::size_t index, ElMagField f);

    void get_buffer(ElMagField* buffer, std::size_t Nb) const;
    void set_buffer(const ElMagField* buffer, std::size_t Nb);

    // dt * sum_i E_i x B_i / mu0 [J/m^2]
    Vector Poynting() const;

    // raw samples, 6 doubles each
    double* data() { return v.data(); }
    const double* data() const { return v.data(); }

private:
    double t0, dt;
    std::size_t N;
    std::vector<double> v;
};
