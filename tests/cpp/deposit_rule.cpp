// Host check of the product's bucket-deposit rule (teufel_amd/csrc/tg_deposit.h)
// against the oracle's restatement of src/particle.cpp:788-865
// (or_deposit_step), on random steps, including steps that start or end on
// bucket edges, start before the trace or run past its end, and every split of
// the trace into windows [b0, b1).  Prints the largest difference relative to
// the largest trace value and exits non-zero above 1e-14.  The mixed kernel's
// bucket-unit form (deposit_step_units: index and FP32 offset from the bucket
// centre, FP32 weights) is checked on the same steps in absolute terms: its
// weights are bucket fractions good to ~1e-7 (FP32 offsets), so every bucket
// must agree within 1e-6 of the larger end field; a wrong index, sign or
// weight moves a step's share of a bucket, far above that for all but the
// shortest steps drawn.
//
// Lane walks: the per-lane walk's deposit (the lane accumulator of lane_step,
// lane_fold and lane_close, with step_general for general steps) run over
// whole random trajectories - short, long and zero-length steps, steps that
// start before the trace or run past its end - through windows of W buckets
// as the radiation kernel walks them: each window from the step that ends in
// it, to the step that ends at or beyond its end, with and without the spill
// bucket (a step ending in the next window's first bucket is deposited whole,
// its share of that bucket carried over).  The traces must equal the
// oracle's deposit of every step within 1e-14 of the largest trace value
// (FP64 weights: the FP64 kernel's doubled, select-free step_common_x2), or within 2e-6 of the largest end field per step touching a
// bucket (bucket units, FP32).  A dropped held-back share, a bucket sum added
// twice or to the wrong bucket, or a lost spill bucket fails by the size of a
// step's contribution.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <vector>

#define TG_HD
#include "../../teufel_amd/csrc/tg_deposit.h"

extern "C" int or_deposit_step(double, double, const double*, const double*, double, double, int, double*);

struct Acc {
    double* tr;
    const double* f1;
    const double* f2;
    void operator()(int j, double a, double b) const
    {
        for (int c = 0; c < 6; c++) tr[6 * j + c] += a * f1[c] + b * f2[c];
    }
};
struct Rcp {
    double operator()(double x) const { return 1.0 / x; }
};
struct RcpF {
    float operator()(float x) const { return 1.0f / x; }
};

static double uni() { return rand() / (RAND_MAX + 1.0); }

// Window cells of one lane: cells[6 j + c], j in [0, span).
template <class F>
struct CellAdd {
    F* cells;
    void operator()(int j, int c, F v) const { cells[6 * j + c] += v; }
};
template <class F>
struct CellAcc {
    F* cells;
    int b0;
    const F* f1;
    const F* f2;
    double scale;   // 2: the kernels' windows hold twice the sums
    void operator()(int j, double a, double b) const
    {
        for (int c = 0; c < 6; c++) cells[6 * (j - b0) + c] += (F)(scale * a) * f1[c] + (F)(scale * b) * f2[c];
    }
};

// One lane's walk of a trajectory (NP points: bucket index idx, bucket-unit
// offset u or retarded time ts, fields f) through the trace in windows of W
// buckets, as the radiation kernel's per-lane walk deposits it.  units: the
// mixed kernel's bucket units in FP32; else FP64 with retarded times.
template <class F>
static void lane_walk(int NP, const int* idx, const double* ts, const float* u, const F* f, double t0, double dt,
                      int nb, int W, bool spill_on, bool units, std::vector<double>& trace)
{
    const double inv = 1.0 / dt, tmin = t0 - dt * 0.5, tmax = t0 + dt * ((double)nb - 0.5);
    std::vector<F> cells(6 * (W + 1));
    int k = 0;
    bool carry = false;
    for (int b0 = 0; b0 < nb; b0 += W) {
        const int b1 = b0 + W < nb ? b0 + W : nb, nw = b1 - b0;
        const bool spill = spill_on && b1 < nb;
        if (carry)
            for (int c = 0; c < 6; c++) cells[c] = cells[6 * W + c];
        for (int i = carry ? 6 : 0; i < 6 * (W + 1); i++) cells[i] = (F)0;
        // skip the steps that end before b0 (with carry, also one that ends
        // in bucket b0 from before it: deposited whole in the spill bucket)
        while (k < NP - 1 && (idx[k + 1] < b0 || (carry && idx[k + 1] == b0 && idx[k] < b0))) k++;
        tg::LaneAcc<F> L;
        tg::lane_begin(L);
        CellAdd<F> add{cells.data()};
        bool walking = k < NP - 1 && idx[k] < b1;
        while (walking) {
            const F* f1 = f + 6 * k;
            const F* f2 = f + 6 * (k + 1);
            const int i2 = idx[k + 1] > nb ? nb : idx[k + 1];
            const bool valid = units ? (idx[k + 1] >= 0 && idx[k] < nb) : (ts[k + 1] > tmin && ts[k] < tmax);
            const bool case_a = units ? idx[k] < 0 : ts[k] <= tmin;
            const bool gen = valid && tg::step_is_general(case_a, idx[k], i2);
            const bool common = valid && !gen;
            tg::TwoBuckets<F> t;
            if (units)
                t = tg::step_common_units<F>(RcpF{}, common, idx[k], (F)u[k], idx[k + 1], (F)u[k + 1], i2, b0, b1,
                                             -1, -1);
            else
                t = tg::step_common_x2<F>(Rcp{}, common, ts[k], idx[k], ts[k + 1], i2, t0, dt, inv, b0, b1, -1,
                                          -1);
            const int ja = idx[k] - b0;
            tg::lane_step(L, add, common, gen || i2 != idx[k], t, ja, (unsigned)ja < (unsigned)nw, f1, f2);
            if (gen) {
                CellAcc<F> acc{cells.data(), b0, f1, f2, 2.0};
                const int hi = b1 + ((spill && i2 == b1) ? 1 : 0);
                if (units) {
                    const double q1 = (double)idx[k] + (double)u[k], q2 = (double)idx[k + 1] + (double)u[k + 1];
                    tg::step_general(acc, 1.0 / (q2 - q1), case_a, q1, idx[k], q2, i2, 0.0, 1.0, 1.0, nb, b0, hi);
                } else {
                    tg::step_general(acc, 1.0 / (ts[k + 1] - ts[k]), case_a, ts[k], idx[k], ts[k + 1], i2, t0, dt,
                                     inv, nb, b0, hi);
                }
            }
            if (idx[k + 1] >= b1) {
                walking = false;
                if (spill && idx[k + 1] == b1) {
                    k++;
                    tg::lane_fold(L, f2);
                    tg::lane_close(L, add, nw);
                }
            } else {
                k++;
                walking = k < NP - 1;
                if (!walking) {
                    tg::lane_fold(L, f2);
                    if ((unsigned)(idx[k] - b0) < (unsigned)nw) tg::lane_close(L, add, idx[k] - b0);
                }
            }
        }
        for (int j = 0; j < nw; j++)
            for (int c = 0; c < 6; c++) trace[6 * (b0 + j) + c] += 0.5 * (double)cells[6 * j + c];
        carry = spill;
    }
}

// Random trajectories through the lane walk against the oracle's deposit.
static int lane_walks(long cases)
{
    double worst = 0.0, worst_units = 0.0;
    for (long it = 0; it < cases; it++) {
        const int nb = 2 + rand() % 40, NP = 2 + rand() % 80;
        const double dt = 0.05 + uni(), t0 = (uni() - 0.5) * 7.0;
        const double mean = (nb + 3.0) * dt / NP * (0.3 + 2.0 * uni());   // steps per bucket from ~0.3 to ~3
        std::vector<double> ts(NP), f(6 * NP);
        ts[0] = t0 + (uni() * 0.6 - 0.55) * nb * dt;
        for (int k = 1; k < NP; k++) {
            const int kind = rand() % 16;
            double d = mean * (0.2 + 1.6 * uni());
            if (kind == 0) d = 0.0;                         // zero-length step
            if (kind == 1) d = dt * (2.0 + 4.0 * uni());    // general step (3+ buckets)
            ts[k] = ts[k - 1] + d;
            if (kind == 2) {                                // end on a bucket edge
                const double e = t0 + (ceil((ts[k - 1] - t0) / dt + 0.5) - 0.5) * dt;
                if (e >= ts[k - 1]) ts[k] = e;
            }
        }
        for (double& v : f) v = uni() - 0.5;
        std::vector<double> ref(6 * nb, 0.0);
        std::vector<int> touch(nb, 0);
        for (int k = 0; k + 1 < NP; k++) {
            or_deposit_step(ts[k], ts[k + 1], &f[6 * k], &f[6 * k + 6], t0, dt, nb, ref.data());
            const int lo = (int)round((ts[k] - t0) / dt), hi = (int)round((ts[k + 1] - t0) / dt);
            for (int j = lo < 0 ? 0 : lo; j <= hi && j < nb; j++) touch[j]++;
        }
        std::vector<int> idx(NP), idx_u(NP);
        std::vector<float> u(NP), ff(6 * NP);
        for (int k = 0; k < NP; k++) {
            idx[k] = (int)round((ts[k] - t0) / dt);
            const double q = (ts[k] - t0) / dt;
            idx_u[k] = (int)nearbyint(q);
            u[k] = (float)(q - idx_u[k]);
        }
        for (int i = 0; i < 6 * NP; i++) ff[i] = (float)f[i];
        const int W = 1 + rand() % 8;
        for (int spill = 0; spill < 2; spill++) {
            std::vector<double> got(6 * nb, 0.0), got_u(6 * nb, 0.0);
            lane_walk<double>(NP, idx.data(), ts.data(), nullptr, f.data(), t0, dt, nb, W, spill, false, got);
            lane_walk<float>(NP, idx_u.data(), ts.data(), u.data(), ff.data(), t0, dt, nb, W, spill, true, got_u);
            double m = 0;
            for (double v : ref) m = fmax(m, fabs(v));
            for (int i = 0; i < 6 * nb; i++) {
                worst = fmax(worst, fabs(got[i] - ref[i]) / (m > 0 ? m : 1.0));
                worst_units = fmax(worst_units, fabs(got_u[i] - ref[i]) / (2e-6 * (touch[i / 6] + 1)) * 2e-6);
            }
        }
    }
    printf("lane walks %ld: worst relative difference %.3e, bucket units %.3e (per step touching a bucket, "
           "of the end fields)\n", cases, worst, worst_units);
    return worst <= 1e-14 && worst_units <= 2e-6 ? 0 : 1;
}

int main(int argc, char** argv)
{
    long cases = argc > 1 ? atol(argv[1]) : 200000;
    srand(12345);
    double worst = 0.0, worst_units = 0.0;
    for (long it = 0; it < cases; it++) {
        int nb = 2 + rand() % 24;
        double dt = 0.05 + uni();
        double t0 = (uni() - 0.5) * 7.0;
        double ts1 = t0 + (uni() * 1.4 - 0.3) * nb * dt;
        if (rand() % 8 == 0) ts1 = t0 + (rand() % (nb + 2) - 1.5) * dt;  // on an edge
        double ts2 = ts1 + dt * (rand() % 4 == 0 ? uni() * 6.0 : uni() * 1.2) + 1e-9;
        if (rand() % 8 == 0) {
            double e = t0 + (rand() % (nb + 2) - 0.5) * dt;
            if (e > ts1) ts2 = e;
        }
        double f1[6], f2[6];
        for (int c = 0; c < 6; c++) {
            f1[c] = uni() - 0.5;
            f2[c] = uni() - 0.5;
        }
        std::vector<double> ref(6 * nb, 0.0), got(6 * nb, 0.0);
        or_deposit_step(ts1, ts2, f1, f2, t0, dt, nb, ref.data());
        double inv = 1.0 / dt, tmin = t0 - dt * 0.5, tmax = t0 + dt * ((double)nb - 0.5);
        int i1 = (int)round((ts1 - t0) / dt), i2 = (int)round((ts2 - t0) / dt);
        int W = 1 + rand() % 6;
        Acc acc{got.data(), f1, f2};
        for (int b0 = 0; b0 < nb; b0 += W) {
            int b1 = b0 + W < nb ? b0 + W : nb;
            tg::deposit_step(acc, Rcp{}, true, ts1, i1, ts2, i2, t0, dt, inv, tmin, tmax, nb, b0, b1);
        }
        // bucket units: q = (ts - t0)/dt = idx + u, idx the nearest integer
        // (half to even, as the kernel's rint), u the FP32 offset
        std::vector<double> units(6 * nb, 0.0);
        Acc acc_u{units.data(), f1, f2};
        const double q1 = (ts1 - t0) / dt, q2 = (ts2 - t0) / dt;
        const int j1 = (int)nearbyint(q1), j2 = (int)nearbyint(q2);
        const float u1 = (float)(q1 - j1), u2 = (float)(q2 - j2);
        for (int b0 = 0; b0 < nb; b0 += W) {
            int b1 = b0 + W < nb ? b0 + W : nb;
            tg::deposit_step_units<float>(acc_u, RcpF{}, Rcp{}, true, j1, u1, j2, u2, nb, b0, b1);
        }
        double m = 0, fm = 0;
        for (double v : ref) m = fmax(m, fabs(v));
        for (int c = 0; c < 6; c++) fm = fmax(fm, fmax(fabs(f1[c]), fabs(f2[c])));
        for (int i = 0; i < 6 * nb; i++) {
            worst = fmax(worst, fabs(got[i] - ref[i]) / (m > 0 ? m : 1.0));
            worst_units = fmax(worst_units, fabs(units[i] - ref[i]) / fm);
        }
    }
    printf("cases %ld worst relative difference %.3e, bucket units (FP32 weights) %.3e of the end fields\n",
           cases, worst, worst_units);
    const int walks = lane_walks(cases / 10);
    return worst <= 1e-14 && worst_units <= 1e-6 && walks == 0 ? 0 : 1;
}
