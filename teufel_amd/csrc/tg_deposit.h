// tg_deposit.h - the bucket deposit of one trajectory step (host and device).
//
// Restates the deposit of ChargedParticle::integrateFieldTrace
// (src/particle.cpp:788-865): a step observed from ts1 to ts2 with end fields
// f1, f2 is integrated, linearly interpolated, over trace buckets of width dt
// centred at t0 + i dt, bucket index round((t - t0)/dt).  The reference's
// rules:
//   single   the step lies in one bucket:           (f1+f2)/2 * dts/dt
//   partial  first / last partially covered bucket: mean of the interpolated
//            end values times the covered time / dt
//   full     fully covered bucket:                   f(bucket centre)
// Every contribution is a*f1 + b*f2.  Time values (bucket edges and centres)
// are formed exactly as the reference forms them, so only amplitude weights
// differ from the reference, by rounding; callers compile this with FMA
// contraction off.
//
// A step that neither starts before the trace nor spans three or more
// buckets - nearly every step - touches one bucket or two adjacent ones split
// at the edge e = t0 + (idx1 + 1/2) dt; step_common gives both targets and
// weights without branches.  step_general covers the rest.  deposit_step
// composes the two with the reference's semantics.
//
// TG_HD marks the functions for host and device; this header is included by
// the HIP kernels and, with TG_HD empty, by a host test of the rule.
#pragma once

#ifndef TG_HD
#define TG_HD __host__ __device__
#endif

namespace tg {

// The two buckets of a common step: window-relative indices (the caller's
// `none_a` / `none_b` when absent or outside [b0, b1)) and weights.
template <class F>
struct TwoBuckets {
    int ja, jb;
    F aa, ba, ab, bb;
};

// True when the step needs step_general: it starts before the trace
// (ts1 <= tmin) or spans three or more buckets.
TG_HD inline bool step_is_general(bool case_a, int idx1, int i2) { return case_a || i2 > idx1 + 1; }

// Common step (valid, not general) in weight type F; rcp(x) = 1/x in F.
//   bucket A = idx1: h_A (f1 + f_e),  bucket B = idx1+1: h_B (f_e + f2),
//   f_e = f1 (ts2-e)/dts + f2 (e-ts1)/dts,  h_A = (e-ts1)/(2dt), h_B = (ts2-e)/(2dt);
// a single-bucket step is the case e = ts2.
template <class F, class Rcp>
TG_HD inline TwoBuckets<F> step_common(const Rcp& rcp, bool common, double ts1, int idx1, double ts2, int i2,
                                       double t0, double dt, double inv_dt, int n_bins, int b0, int b1,
                                       int none_a, int none_b)
{
#pragma clang fp contract(off)
    TwoBuckets<F> o;
    const bool single = i2 == idx1;
    const double e = single ? ts2 : t0 + (idx1 + 0.5) * dt;     // particle.cpp:841 t_end
    const F da = (F)(e - ts1), db = (F)(ts2 - e);               // exact time differences
    // a single-bucket step has la = 1, lb = 0 exactly (also for a zero-length
    // step, which adds exact zeros as in the reference)
    const F idts = rcp(single ? (F)1 : da + db);
    const F la = single ? (F)1 : da * idts, lb = single ? (F)0 : db * idts;
    const F fid = (F)inv_dt;
    const F ha = (F)0.5 * (da * fid), hb = (F)0.5 * (db * fid);
    o.aa = ha * ((F)1 + lb);
    o.ba = ha * la;
    o.ab = hb * lb;
    o.bb = hb * (la + (F)1);
    o.ja = (common && idx1 >= b0 && idx1 < b1) ? idx1 - b0 : none_a;
    o.jb = (common && !single && i2 < n_bins && i2 >= b0 && i2 < b1) ? i2 - b0 : none_b;
    return o;
}

// step_common for the FP64 radiation kernel: twice the weights (its window
// holds twice the bucket sums and halves them when it is flushed: a power of
// two, so every rounding is the same) and no selects on the weights.  A
// single-bucket step has its "edge" at e = ts2, so db = 0 and lb = 0 exactly
// and la = dts RN(1/dts), within an ulp of the 1 of step_common; a
// zero-length step adds zeros (its length is raised to 1e-300 before the
// reciprocal; retarded times differ by at least an ulp, ~1e-24 s, or not at
// all).  A step that is not common gets zero weights (its time differences
// are taken as zero), as lane_step requires.  [b0, b1) lies within the
// trace.
template <class F, class Rcp>
TG_HD inline TwoBuckets<F> step_common_x2(const Rcp& rcp, bool common, double ts1, int idx1, double ts2, int i2,
                                          double t0, double dt, double inv_dt, int b0, int b1, int none_a,
                                          int none_b)
{
#pragma clang fp contract(off)
    TwoBuckets<F> o;
    const bool single = i2 == idx1;
    const double e = single ? ts2 : t0 + (idx1 + 0.5) * dt;     // particle.cpp:841 t_end
    const F da = common ? (F)(e - ts1) : (F)0, db = common ? (F)(ts2 - e) : (F)0;   // exact time differences
    const F len = da + db;
    const F idts = rcp(len > (F)1e-300 ? len : (F)1e-300);
    const F la = da * idts, lb = db * idts;
    const F fid = (F)inv_dt;
    const F ha = da * fid, hb = db * fid;
    o.aa = ha * ((F)1 + lb);
    o.ba = ha * la;
    o.ab = hb * lb;
    o.bb = hb * (la + (F)1);
    const unsigned span = (unsigned)(b1 - b0);
    o.ja = (common && (unsigned)(idx1 - b0) < span) ? idx1 - b0 : none_a;
    o.jb = (common && !single && (unsigned)(i2 - b0) < span) ? i2 - b0 : none_b;
    return o;
}

// step_common in bucket units, for the mixed kernel: a point sits at bucket
// position q = (ts - t0)/dt = idx + u, u in [-1/2, 1/2] its offset from the
// centre of its bucket idx.  Positions are taken relative to the centre of
// bucket idx1: the step runs from u1 to x2 = (idx2 - idx1) + u2, the edge
// between buckets idx1 and idx1 + 1 lies at 1/2 (a single-bucket step has
// its "edge" at x2), so
//   (e - ts1)/dt = e - u1,  (ts2 - e)/dt = x2 - e,  dts/dt = x2 - u1,
// and no weight needs a select: a single-bucket step gets lb = 0 from
// x2 - e = 0.  A zero-length step adds zeros (its length is raised to
// 1e-30 before the reciprocal), and so does a step that is not common (its
// differences are taken as zero), as lane_step requires.  idx2 is the
// unclamped index of ts2 (i2 = min(idx2, n_bins)); [b0, b1) lies within the
// trace.  The weights are
// twice those of step_common (the mixed kernel's window holds twice the
// bucket sums and halves them when it is flushed: a power of two, so every
// rounding is the same, and the weights need no factor 1/2).
template <class F, class Rcp>
TG_HD inline TwoBuckets<F> step_common_units(const Rcp& rcp, bool common, int idx1, F u1, int idx2, F u2, int i2,
                                             int b0, int b1, int none_a, int none_b)
{
    TwoBuckets<F> o;
    const bool single = i2 == idx1;
    const F x2 = (F)(idx2 - idx1) + u2;
    const F e = single ? x2 : (F)0.5;
    const F da = common ? e - u1 : (F)0, db = common ? x2 - e : (F)0;
    const F len = da + db;
    const F idts = rcp(len > (F)1e-30 ? len : (F)1e-30);
    const F la = da * idts, lb = db * idts;
    o.aa = da * ((F)1 + lb);
    o.ba = da * la;
    o.ab = db * lb;
    o.bb = db * (la + (F)1);
    const unsigned span = (unsigned)(b1 - b0);
    o.ja = (common && (unsigned)(idx1 - b0) < span) ? idx1 - b0 : none_a;
    o.jb = (common && !single && (unsigned)(i2 - b0) < span) ? i2 - b0 : none_b;
    return o;
}

// a*b + c with one rounding, in F.
TG_HD inline double fma_f(double a, double b, double c) { return __builtin_fma(a, b, c); }
TG_HD inline float fma_f(float a, float b, float c) { return __builtin_fmaf(a, b, c); }

// Point-weight accumulation of one lane's walk over consecutive common
// steps (the per-lane walk of the radiation kernel).  A walk's steps arrive
// in order and their buckets never decrease, so the lane keeps its sum for
// the bucket of its current step's start point in acc and adds it to the
// window once, when it leaves that bucket.  A common step adds aa f1 + ba f2
// to its bucket idx1 and, crossing, ab f1 + bb f2 to idx1 + 1 (step_common);
// the share of f2 in the bucket the next step starts in (ba, or bb after a
// crossing) is held back in wc and added with the next step's own share of
// that point, as (wc + aa) f1.  So a point's field is multiplied once per
// bucket its two steps give it, and a bucket receives one add per lane and
// particle: 6 products a step, 12 more and one window add per bucket
// crossed.  The bucket sums are those of deposit_step, grouped by point.
template <class F>
struct LaneAcc {
    F acc[6];
    F wc;
};

template <class F>
TG_HD inline void lane_begin(LaneAcc<F>& L)
{
    for (int c = 0; c < 6; ++c) L.acc[c] = (F)0;
    L.wc = (F)0;
}

// Step p1 -> p2 of a lane that walks it.  common: valid and not general,
// with weights t (step_common_x2 or step_common_units, which give any other
// step zero weights); any other step adds nothing here (a general one goes
// to step_general), and its end point starts the next bucket sum from zero.
// leaves: the step does not end in bucket idx1 (crossing or general); ja: the
// window-relative index of bucket idx1, in_win: whether it lies in the
// window.  add(j, c, v) adds v to component c of window bucket j.
template <class F, class Add>
TG_HD inline void lane_step(LaneAcc<F>& L, const Add& add, bool common, bool leaves, const TwoBuckets<F>& t, int ja,
                            bool in_win, const F* f1, const F* f2)
{
    (void)common;
    const F wa = L.wc + t.aa;
    for (int c = 0; c < 6; ++c) L.acc[c] = fma_f(wa, f1[c], L.acc[c]);
    if (leaves) {
        if (in_win)
            for (int c = 0; c < 6; ++c) add(ja, c, fma_f(t.ba, f2[c], L.acc[c]));
        for (int c = 0; c < 6; ++c) L.acc[c] = t.ab * f1[c];
        L.wc = t.bb;
    } else {
        L.wc = t.ba;
    }
}

// The lane stops past the end point p2 of its last step: the held-back share
// of p2 goes into acc.
template <class F>
TG_HD inline void lane_fold(LaneAcc<F>& L, const F* f2)
{
    for (int c = 0; c < 6; ++c) L.acc[c] = fma_f(L.wc, f2[c], L.acc[c]);
    L.wc = (F)0;
}

// acc into window bucket j, the bucket of the lane's last point (after
// lane_fold).
template <class F, class Add>
TG_HD inline void lane_close(const LaneAcc<F>& L, const Add& add, int j)
{
    for (int c = 0; c < 6; ++c) add(j, c, L.acc[c]);
}

// One bucket j of a general step.  first: the step starts in bucket j (and not
// before the trace); last: the step ends in bucket j (j < n_bins).
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

// General step (valid): every bucket of its range within [b0, b1).
template <class Acc>
TG_HD inline void step_general(const Acc& acc, double idts, bool case_a, double ts1, int idx1, double ts2, int i2,
                               double t0, double dt, double inv_dt, int n_bins, int b0, int b1)
{
    const int lo = case_a ? 0 : idx1;
    const int hi = i2 < n_bins - 1 ? i2 : n_bins - 1;
    const int jlo = lo > b0 ? lo : b0;
    const int jhi = hi < b1 - 1 ? hi : b1 - 1;
    for (int j = jlo; j <= jhi; ++j)
        deposit_bucket(acc, j, !case_a && j == idx1, j == i2, ts1, ts2, idts, t0, dt, inv_dt);
}

// The whole step with the reference's semantics, restricted to buckets
// [b0, b1); acc(j, a, b) adds a*f1 + b*f2 into bucket j.  idx1/idx2 are the
// bucket indices of ts1/ts2, tmin/tmax the trace span.
template <class Acc, class Rcp>
TG_HD inline void deposit_step(const Acc& acc, const Rcp& rcp, bool active, double ts1, int idx1, double ts2,
                               int idx2, double t0, double dt, double inv_dt, double tmin, double tmax,
                               int n_bins, int b0, int b1)
{
    if (!(active && ts2 > tmin && ts1 < tmax)) return;          // particle.cpp:793
    const bool case_a = ts1 <= tmin;                            // particle.cpp:796
    const int i2 = idx2 > n_bins ? n_bins : idx2;               // particle.cpp:806,834
    if (step_is_general(case_a, idx1, i2)) {
        step_general(acc, rcp(ts2 - ts1), case_a, ts1, idx1, ts2, i2, t0, dt, inv_dt, n_bins, b0, b1);
        return;
    }
    const TwoBuckets<double> t =
        step_common<double>(rcp, true, ts1, idx1, ts2, i2, t0, dt, inv_dt, n_bins, b0, b1, -1, -1);
    if (t.ja >= 0) acc(t.ja + b0, t.aa, t.ba);
    if (t.jb >= 0) acc(t.jb + b0, t.ab, t.bb);
}

// deposit_step in bucket units (the mixed kernel's form): points at
// q = idx + u (step_common_units); the trace spans q in (-1/2, n_bins - 1/2),
// so the window test and case A become index tests, and a general step runs
// step_general with t0 = 0, dt = 1.  rcp: reciprocal in F, rcpd in double.
// acc receives the rule's own weights (step_common_units' halved).
template <class F, class Acc, class Rcp, class RcpD>
TG_HD inline void deposit_step_units(const Acc& acc, const Rcp& rcp, const RcpD& rcpd, bool active, int idx1, F u1,
                                     int idx2, F u2, int n_bins, int b0, int b1)
{
    if (!(active && idx2 >= 0 && idx1 < n_bins)) return;         // particle.cpp:793
    const bool case_a = idx1 < 0;                               // particle.cpp:796
    const int i2 = idx2 > n_bins ? n_bins : idx2;
    if (step_is_general(case_a, idx1, i2)) {
        const double q1 = (double)idx1 + (double)u1, q2 = (double)idx2 + (double)u2;
        step_general(acc, rcpd(q2 - q1), case_a, q1, idx1, q2, i2, 0.0, 1.0, 1.0, n_bins, b0, b1);
        return;
    }
    const TwoBuckets<F> t = step_common_units<F>(rcp, true, idx1, u1, idx2, u2, i2, b0, b1, -1, -1);
    if (t.ja >= 0) acc(t.ja + b0, (F)0.5 * t.aa, (F)0.5 * t.ba);
    if (t.jb >= 0) acc(t.jb + b0, (F)0.5 * t.ab, (F)0.5 * t.bb);
}

}  // namespace tg
