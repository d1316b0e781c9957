// particle.cpp - ChargedParticle trajectory store; the field integration goes
// to the GPU through gpu_radiation.
#include "particle.h"

#include "gpu_radiation.h"

namespace {
unsigned long long next_version()
{
    static unsigned long long v = 0;
    return ++v;
}
}  // namespace

ChargedParticle::ChargedParticle() : NP(0), Charge(-1.0), Mass(1.0), Version(next_version()) {}

ChargedParticle::ChargedParticle(double charge, double mass)
    : NP(0), Charge(charge), Mass(mass), Version(next_version())
{
}

ChargedParticle::ChargedParticle(const ChargedParticle* o)
    : NP(o->NP), Charge(o->Charge), Mass(o->Mass), Time(o->Time), X(o->X), P(o->P), A(o->A),
      Version(next_version())
{
}

void ChargedParticle::initTrajectory(double t, Vector x, Vector p, Vector a)
{
    Time.assign(1, t);
    X.assign(1, x);
    P.assign(1, p);
    A.assign(1, a);
    NP = 1;
    Version = next_version();
}

void ChargedParticle::storeTrajectoryPoint(double time, Vector pos, Vector mom, Vector acc)
{
    Time.push_back(time);
    X.push_back(pos);
    P.push_back(mom);
    A.push_back(acc);
    NP++;
    Version = next_version();
}

double ChargedParticle::TrajTime(int step) const
{
    if (step < 0 || step >= NP) throw IOexception("ChargedParticle::TrajTime - index out of range.");
    return Time[step];
}

Vector ChargedParticle::TrajPoint(int step) const
{
    if (step < 0 || step >= NP) throw IOexception("ChargedParticle::TrajPoint - index out of range.");
    return X[step];
}

Vector ChargedParticle::TrajMomentum(int step) const
{
    if (step < 0 || step >= NP) throw IOexception("ChargedParticle::TrajMomentum - index out of range.");
    return P[step];
}

Vector ChargedParticle::TrajAccel(int step) const
{
    if (step < 0 || step >= NP) throw IOexception("ChargedParticle::TrajAccel - index out of range.");
    return A[step];
}

void ChargedParticle::integrateFieldTrace(Vector ObservationPoint, FieldTrace* trace)
{
    std::vector<ChargedParticle*> one(1, this);
    std::vector<Vector> pos(1, ObservationPoint);
    std::vector<FieldTrace*> tr(1, trace);
    teufel_gpu::integrate(one, pos, tr);
}

ElMagField ChargedParticle::RetardedField(double obs_time, Vector obs_pos)
{
    std::vector<ElMagField> out;
    teufel_gpu::fields_at(std::vector<ChargedParticle*>(1, this), std::vector<double>(1, obs_time),
                          std::vector<Vector>(1, obs_pos), out);
    return out[0];
}
