// bunch.cpp - Bunch and Beam; their field integration is one device call over
// all particles in summation order.
#include "bunch.h"

#include "gpu_radiation.h"

Bunch::~Bunch()
{
    for (ChargedParticle* p : P) delete p;
}

ChargedParticle* Bunch::getParticle(int i)
{
    if (i < 0 || i >= (int)P.size()) throw IOexception("Bunch::getParticle - index out of range.");
    return P[i];
}

void Bunch::integrateFieldTrace(Vector ObservationPoint, FieldTrace* trace)
{
    std::vector<Vector> pos(1, ObservationPoint);
    std::vector<FieldTrace*> tr(1, trace);
    teufel_gpu::integrate(P, pos, tr);
}

Beam::~Beam()
{
    for (Bunch* b : B) delete b;
}

Bunch* Beam::getBunch(int i)
{
    if (i < 0 || i >= (int)B.size()) throw IOexception("Beam::getBunch - index out of range.");
    return B[i];
}

int Beam::getNOP() const
{
    int n = 0;
    for (const Bunch* b : B) n += b->getNOP();
    return n;
}

void Beam::integrateFieldTrace(Vector ObservationPoint, FieldTrace* trace)
{
    std::vector<ChargedParticle*> all;
    collect(all);
    std::vector<Vector> pos(1, ObservationPoint);
    std::vector<FieldTrace*> tr(1, trace);
    teufel_gpu::integrate(all, pos, tr);
}

ElMagField Bunch::RetardedField(double obs_time, Vector obs_pos)
{
    std::vector<ElMagField> out;
    teufel_gpu::fields_at(P, std::vector<double>(1, obs_time), std::vector<Vector>(1, obs_pos), out);
    return out[0];
}

ElMagField Beam::RetardedField(double obs_time, Vector obs_pos)
{
    std::vector<ChargedParticle*> all;
    collect(all);
    std::vector<ElMagField> out;
    teufel_gpu::fields_at(all, std::vector<double>(1, obs_time), std::vector<Vector>(1, obs_pos), out);
    return out[0];
}
