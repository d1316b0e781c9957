// gpu_radiation.h - adapter from the host classes to the C ABI
// (include/teufel_gpu.h).  One device context per process, created on first
// use.  integrate() may be called from many threads at once, as the
// reference's per-pixel OpenMP loop calls integrateFieldTrace
// (src/mesh.cpp:296-301): concurrent calls are combined, one caller launching
// all requests queued so far, one batched launch per particle list, while the
// others wait, and every call returns once its own traces hold its fields.
// Observer::integrate() hands all pixels over in one call.  Non-zero status
// codes of the C ABI become IOexception.
#pragma once

#include <vector>

#include "fieldtrace.h"
#include "particle.h"

namespace teufel_gpu {

// TG_FP64 (0, default) or TG_MIXED (1); the environment variable
// TEUFEL_GPU_PRECISION=mixed selects mixed at first use.
void set_precision(int precision);
int precision();

// Device ordinal used when the context is created (default 0, or the
// environment variable TEUFEL_GPU_DEVICE).
void set_device(int device);

// Accumulate the fields of `particles` (summed in this order) at `points`
// into `traces` (one per point; each has its own t0, and traces sharing
// dt and N are integrated in one batched launch).  Particles with fewer
// trajectory points are padded by repeating their last point, which adds
// exact zeros, as the reference's step loop would.
void integrate(const std::vector<ChargedParticle*>& particles, const std::vector<Vector>& points,
               const std::vector<FieldTrace*>& traces);

// The field of `particles` (summed in this order) at points[i] at times[i]
// (RetardedField(double, Vector), src/particle.cpp:612-769), one device call
// for all observations; out is resized to points.size().  A particle without
// trajectory points raises RANGEexception, as the reference does.
void fields_at(const std::vector<ChargedParticle*>& particles, const std::vector<double>& times,
               const std::vector<Vector>& points, std::vector<ElMagField>& out);

// The device bucket deposit driven by injected retarded times ts[n_pseudo][n_points]
// and fields fields[n_pseudo][n_points][6], accumulated into trace.
void deposit_only(int n_pseudo, int n_points, const double* ts, const double* fields, FieldTrace* trace);

// Number of batched launches integrate() has made in this process (combined
// concurrent calls count once), for tests and reports.
unsigned long long launch_count();

// Release the device context (also done at process exit).
void shutdown();

}  // namespace teufel_gpu
