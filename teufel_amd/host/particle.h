// particle.h - a tracked particle and its stored trajectory (interface of
// ChargedParticle, src/particle.h, for the radiation path).
//
// The trajectory holds, per point, the time t [s], position X [m], momentum
// P = beta*gamma and A = d(beta*gamma)/dt [1/s], exactly as the reference
// stores them (src/particle.h:417-427).  integrateFieldTrace evaluates the
// Lienard-Wiechert field of the whole trajectory at one observation point on
// the GPU and accumulates it into the trace (src/particle.cpp:771-867).
#pragma once

#include <vector>

#include "fieldtrace.h"
#include "fields.h"
#include "global.h"
#include "vector.h"

class ChargedParticle {
public:
    // electron: charge -1 e, mass 1 m_e (src/particle.cpp:29-31)
    ChargedParticle();
    // charge in units of e, mass in units of m_e
    ChargedParticle(double charge, double mass);
    explicit ChargedParticle(const ChargedParticle* origin);
    virtual ~ChargedParticle() {}

    int getNP() const { return NP; }
    double getCharge() const { return Charge; }
    double getMass() const { return Mass; }

    // Start a trajectory with one point; storeTrajectoryPoint appends one.
    void initTrajectory(double t, Vector x, Vector p, Vector a);
    void storeTrajectoryPoint(double time, Vector pos, Vector mom, Vector acc);

    // Trajectory access; out-of-range indices throw IOexception.
    double TrajTime(int step) const;
    Vector TrajPoint(int step) const;
    Vector TrajMomentum(int step) const;
    Vector TrajAccel(int step) const;

    // Accumulate the field observed at ObservationPoint into trace.
    void integrateFieldTrace(Vector ObservationPoint, FieldTrace* trace);

    // The field at obs_pos at time obs_time from the trajectory point on the
    // observer's past light cone (src/particle.cpp:612-769), on the GPU.
    // Throws RANGEexception without trajectory points.
    ElMagField RetardedField(double obs_time, Vector obs_pos);

    // Incremented by every change of the trajectory (the device cache key).
    unsigned long long version() const { return Version; }

protected:
    int NP;

private:
    double Charge;
    double Mass;
    std::vector<double> Time;
    std::vector<Vector> X, P, A;
    unsigned long long Version;
};
