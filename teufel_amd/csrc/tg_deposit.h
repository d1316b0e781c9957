// tg_deposit.h - the bucket deposit of one trajectory step (host and device).
//
// Restates the deposit of ChargedParticle::integrateFieldTrace
// (src/particle.cpp:788-865): a step observed from ts1 to ts2 with end fields
// f1, f2 is integrated, linearly interpolated, over trace buckets of width dt
// centred at t0 + i dt, bucket index round((t - t0)/dt).  Three rules of the
// reference cover every step:
//   single   the step lies in one bucket:           (f1+f2)/2 * dts/dt
//   partial  first / last partially covered bucket: mean of the interpolated
//            end values times the covered time / dt
//   full     fully covered bucket:                   f(bucket centre)
// Every contribution is a*f1 + b*f2, added through acc(j, a, b).  Time values
// (bucket edges and centres) are formed exactly as the reference forms them,
// so only amplitude weights differ from the reference, by rounding; the
// caller compiles this with FMA contraction off.
//
// TG_HD marks the functions for host and device; this header is included by
// the HIP kernels and, with TG_HD empty, by a host test of the rule.
#pragma once

#ifndef TG_HD
#define TG_HD __host__ __device__
#endif

namespace tg {

// One bucket j of a step, any rule.  first: the step starts in bucket j (and
// not before the trace); last: the step ends in bucket j (j < n_bins).
template <class Acc>
TG_HD inline void deposit_bucket(const Acc& acc, int j, bool first, bool last, double ts1, double ts2,
                                 double idts, double t0, double dt, double inv_dt)
{
#pragma clang fp contract(off)
    if (!(first || last)) {
        const double tc = t0 + j * dt;                           // particle.cpp:813,849
        acc(j, (ts2 - tc) * idts, (tc - ts1) * idts);
        return;
    }
    const double left = first ? ts1 : t0 + (j - 0.5) * dt;       // particle.cpp:819,857
    const double right = last ? ts2 : t0 + (j + 0.5) * dt;       // particle.cpp:841
    const double h = 0.5 * ((right - left) * inv_dt);
    acc(j, h * ((ts2 - left) * idts + (ts2 - right) * idts), h * ((left - ts1) * idts + (right - ts1) * idts));
}

// The whole step, restricted to buckets [b0, b1).  idx1/idx2 are the bucket
// indices of ts1/ts2 (round half away from zero); tmin/tmax the trace span
// t0 - dt/2, t0 + (n_bins - 1/2) dt; rcp(x) returns 1/x.
template <class Acc, class Rcp>
TG_HD inline void deposit_step(const Acc& acc, const Rcp& rcp, bool active, double ts1, int idx1, double ts2,
                               int idx2, double t0, double dt, double inv_dt, double tmin, double tmax,
                               int n_bins, int b0, int b1)
{
#pragma clang fp contract(off)
    if (!(active && ts2 > tmin && ts1 < tmax)) return;          // particle.cpp:793
    const bool caseA = ts1 <= tmin;                             // starts before the trace
    const int i2 = idx2 > n_bins ? n_bins : idx2;
    const double dts = ts2 - ts1;
    if (!caseA && idx1 == i2) {
        // single bucket (particle.cpp:835-839)
        if (idx1 >= b0 && idx1 < b1) {
            const double w = 0.5 * (dts * inv_dt);
            acc(idx1, w, w);
        }
        return;
    }
    const double idts = rcp(dts);
    if (!caseA && i2 == idx1 + 1) {
        // two buckets split at the edge t_end = t0 + (idx1 + 1/2) dt
        if (idx1 >= b0 && idx1 < b1) deposit_bucket(acc, idx1, true, false, ts1, ts2, idts, t0, dt, inv_dt);
        if (i2 < n_bins && i2 >= b0 && i2 < b1) deposit_bucket(acc, i2, false, true, ts1, ts2, idts, t0, dt, inv_dt);
        return;
    }
    // general: a step over three or more buckets, or one that starts before the trace
    const int lo = caseA ? 0 : idx1;
    const int hi = i2 < n_bins - 1 ? i2 : n_bins - 1;
    const int jlo = lo > b0 ? lo : b0;
    const int jhi = hi < b1 - 1 ? hi : b1 - 1;
    for (int j = jlo; j <= jhi; ++j)
        deposit_bucket(acc, j, !caseA && j == idx1, j == i2, ts1, ts2, idts, t0, dt, inv_dt);
}

}  // namespace tg
