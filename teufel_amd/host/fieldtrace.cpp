// fieldtrace.cpp - FieldTrace (interface of src/fieldtrace.h for the radiation path).
#include "fieldtrace.h"

#include <algorithm>
#include <cstring>

FieldTrace::FieldTrace(double t0_, double dt_, std::size_t N_) : t0(t0_), dt(dt_), N(N_), v(6 * N_, 0.0) {}

FieldTrace::FieldTrace(const FieldTrace* o) : t0(o->t0), dt(o->dt), N(o->N), v(o->v) {}

double FieldTrace::get_time(std::size_t index) const
{
    if (index >= N) throw IOexception("FieldTrace::get_time - index out of range.");
    return t0 + index * dt;
}

void FieldTrace::zero() { std::fill(v.begin(), v.end(), 0.0); }

ElMagField FieldTrace::get_field(std::size_t index) const
{
    if (index >= N) throw IOexception("FieldTrace::get_field - index out of range.");
    const double* s = &v[6 * index];
    return ElMagField(Vector(s[0], s[1], s[2]), Vector(s[3], s[4], s[5]));
}

void FieldTrace::set_field(std::size_t index, ElMagField f)
{
    if (index >= N) throw IOexception("FieldTrace::set_field - index out of range.");
    double* s = &v[6 * index];
    Vector E = f.E(), B = f.B();
    s[0] = E.x; s[1] = E.y; s[2] = E.z;
    s[3] = B.x; s[4] = B.y; s[5] = B.z;
}

void FieldTrace::add(std::size_t index, ElMagField f)
{
    if (index >= N) throw IOexception("FieldTrace::add - index out of range.");
    double* s = &v[6 * index];
    Vector E = f.E(), B = f.B();
    s[0] += E.x; s[1] += E.y; s[2] += E.z;
    s[3] += B.x; s[4] += B.y; s[5] += B.z;
}

void FieldTrace::get_buffer(ElMagField* buffer, std::size_t Nb) const
{
    if (Nb != N) throw IOexception("FieldTrace::get_buffer - size mismatch.");
    for (std::size_t i = 0; i < N; i++) buffer[i] = get_field(i);
}

void FieldTrace::set_buffer(const ElMagField* buffer, std::size_t Nb)
{
    if (Nb != N) throw IOexception("FieldTrace::set_buffer - size mismatch.");
    for (std::size_t i = 0; i < N; i++) set_field(i, buffer[i]);
}

Vector FieldTrace::Poynting() const
{
    Vector sum;
    for (std::size_t i = 0; i < N; i++) sum += get_field(i).Poynting();
    return sum * dt;
}
