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
    return worst <= 1e-14 && worst_units <= 1e-6 ? 0 : 1;
}
