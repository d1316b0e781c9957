// bunch.h / Beam - ordered particle containers (interfaces of Bunch,
// src/bunch.h:206-252, and Beam, src/beam.h:60-95, for the radiation path).
// Fields are summed over particles in index order, bunches in order
// (src/bunch.cpp:739-745, src/beam.cpp:443-450).  A Bunch owns its
// particles, a Beam its bunches (src/bunch.cpp:351-355, src/beam.cpp:39-43).
#pragma once

#include <vector>

#include "particle.h"

class Bunch {
public:
    Bunch() {}
    ~Bunch();
    void Add(ChargedParticle* part) { P.push_back(part); }
    int getNOP() const { return (int)P.size(); }
    ChargedParticle* getParticle(int i);
    void integrateFieldTrace(Vector ObservationPoint, FieldTrace* trace);
    // sum of the particles' RetardedField(t, x) (src/bunch.cpp:729-737)
    ElMagField RetardedField(double obs_time, Vector obs_pos);
    // all particles in summation order
    void collect(std::vector<ChargedParticle*>& out) const { out.insert(out.end(), P.begin(), P.end()); }

private:
    std::vector<ChargedParticle*> P;
};

class Beam {
public:
    Beam() {}
    ~Beam();
    void Add(Bunch* bunch) { B.push_back(bunch); }
    int getNOB() const { return (int)B.size(); }
    Bunch* getBunch(int i);
    int getNOP() const;
    void integrateFieldTrace(Vector ObservationPoint, FieldTrace* trace);
    // sum over bunches of their RetardedField(t, x) (src/beam.cpp:433-441)
    ElMagField RetardedField(double obs_time, Vector obs_pos);
    void collect(std::vector<ChargedParticle*>& out) const
    {
        for (const Bunch* b : B) b->collect(out);
    }

private:
    std::vector<Bunch*> B;
};
