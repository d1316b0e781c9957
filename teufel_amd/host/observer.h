// observer.h - observer interface (interface of BaseObserver / Observer<SourceT>,
// src/observer.h:43-146).  integrate() adds the fields of the source to
// every trace of the observer; getBuffer/fromBuffer exchange all traces as
// one [points][samples][6] array of doubles.
#pragma once

#include <cstddef>
#include <vector>

#include "bunch.h"

enum RadSourceType { LatticeObservation, BeamObservation };

class BaseObserver {
public:
    virtual ~BaseObserver() {}
    virtual void integrate() = 0;
    void setSourceType(RadSourceType s) { source_type = s; }
    RadSourceType getSourceType() { return source_type; }
    // number of doubles of getBuffer()
    virtual std::size_t getBufferSize() = 0;
    // all traces in a new[] buffer the caller deletes
    virtual double* getBuffer() = 0;
    virtual void fromBuffer(double* buffer, std::size_t size) = 0;
    virtual void generateOutput() = 0;

protected:
    RadSourceType source_type = BeamObservation;
};

template <typename SourceT>
class Observer : public BaseObserver {
public:
    Observer() { source_type = BeamObservation; }
    virtual ~Observer() {}
    virtual void setSource(SourceT* src) { source = src; }
    virtual void integrate() = 0;

public:
    SourceT* source = nullptr;
};

// particles of a source in the summation order of its integrateFieldTrace
inline void collect_particles(Beam* b, std::vector<ChargedParticle*>& out) { b->collect(out); }
inline void collect_particles(Bunch* b, std::vector<ChargedParticle*>& out) { b->collect(out); }
