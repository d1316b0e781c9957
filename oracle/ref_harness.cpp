/*
 * ref_harness.cpp - C entry points around the reference's own, unmodified
 * ChargedParticle / FieldTrace / ElMagField / Vector code.
 *
 * TEST INFRASTRUCTURE ONLY.  oracle/Makefile compiles this file together with
 * src/{vector,fields,fieldtrace,particle,global}.cpp read in place from the
 * reference tree into oracle/_ref/libteufel_ref.so.  The reference's Bunch,
 * Beam, MeshedScreen and PlanarUndulator sources need HDF5, SDDS or muParser,
 * which this image lacks, so the few lines they contribute on this path are
 * restated here:
 *   - the particle loop of Bunch::integrateFieldTrace (src/bunch.cpp:739-745),
 *   - the OpenMP pixel loop of MeshedScreen::integrate (src/mesh.cpp:288-342),
 *   - PlanarUndulator::LocalField (src/undulator.cpp:107-126) as a
 *     GeneralField for the reference's own InitVay/StepVay.
 * The bucket deposit, the Lienard-Wiechert field, the retarded time and the
 * Vay pusher that run here are the reference's compiled code.
 */
#include <math.h>
#include <string.h>
#include <vector>
#include <omp.h>

#include "global.h"
#include "vector.h"
#include "fields.h"
#include "fieldtrace.h"
#include "particle.h"

namespace {

/* A beam: the reference's particles built from SoA trajectory arrays. */
struct RefBeam {
    std::vector<ChargedParticle*> particles;
    ~RefBeam() { for (auto* p : particles) delete p; }
};

/* Test particle whose retarded times/fields are injected, as the Special class
 * of tests/teufel.integrate_field.cpp:51-82 does through the virtual
 * RetardedTime/RetardedField overrides. */
class InjectedParticle : public ChargedParticle {
public:
    InjectedParticle(int n, const double* ts, const double* f) : ts_(ts), f_(f) { NP = n; }
    double RetardedTime(int index, Vector) override { return ts_[index]; }
    ElMagField RetardedField(int index, Vector) override
    {
        const double* g = f_ + 6 * index;
        return ElMagField(Vector(g[0], g[1], g[2]), Vector(g[3], g[4], g[5]));
    }
private:
    const double* ts_;
    const double* f_;
};

/* PlanarUndulator::LocalField (src/undulator.cpp:107-126) placed at `origin`
 * (LocalizedField::Field, src/fields.cpp:180-183). */
class UndulatorField : public GeneralField {
public:
    UndulatorField(double b, double lambda, int n, Vector origin)
        : BPeak(b), LambdaU(lambda), NPeriods(n), Origin(origin), kz(2.0 * Pi / lambda) {}
    bool is_external() override { return true; }
    ElMagField Field(double t, Vector Xg) override
    {
        (void)t;
        Vector X = Xg - Origin;
        Vector E = Vector(0.0, 0.0, 0.0);
        Vector B = Vector(0.0, 0.0, 0.0);
        double z1 = -(NPeriods + 1) * LambdaU / 2.0;
        double z2 = (NPeriods + 1) * LambdaU / 2.0;
        if (z1 <= X.z && X.z < z2) {
            B.x = 0.0;
            B.y = BPeak * sin(kz * X.z) * cosh(kz * X.y);
            B.z = BPeak * cos(kz * X.z) * sinh(kz * X.y);
            if (X.z < z1 + LambdaU) B *= (X.z - z1) / LambdaU;
            if (z2 - LambdaU <= X.z) B *= (z2 - X.z) / LambdaU;
        }
        ElMagField sum;               // Lattice::Field sums its elements onto zero
        sum += ElMagField(E, B);
        return sum;
    }
private:
    double BPeak, LambdaU;
    int NPeriods;
    Vector Origin;
    double kz;
};

void unpack_trace(FieldTrace* tr, double* out)
{
    for (std::size_t i = 0; i < tr->get_N(); i++) {
        ElMagField f = tr->get_field(i);
        out[6 * i + 0] = f.E().x; out[6 * i + 1] = f.E().y; out[6 * i + 2] = f.E().z;
        out[6 * i + 3] = f.B().x; out[6 * i + 4] = f.B().y; out[6 * i + 5] = f.B().z;
    }
}

void pack_trace(FieldTrace* tr, const double* in)
{
    for (std::size_t i = 0; i < tr->get_N(); i++) {
        const double* g = in + 6 * i;
        tr->set_field(i, ElMagField(Vector(g[0], g[1], g[2]), Vector(g[3], g[4], g[5])));
    }
}

} // namespace

extern "C" {

/* Build the reference particles from SoA arrays (particle-major). */
void* ref_beam_create(int n_particles, int n_points, const double* t, const double* x,
                      const double* y, const double* z, const double* ux, const double* uy,
                      const double* uz, const double* ax, const double* ay, const double* az,
                      const double* charge)
{
    RefBeam* b = new RefBeam;
    for (int p = 0; p < n_particles; p++) {
        ChargedParticle* q = new ChargedParticle(charge[p], 1.0);
        std::size_t g = (std::size_t)p * n_points;
        if (n_points > 0) {
            q->initTrajectory(t[g], Vector(x[g], y[g], z[g]), Vector(ux[g], uy[g], uz[g]),
                              Vector(ax[g], ay[g], az[g]));
            for (int k = 1; k < n_points; k++) {
                std::size_t h = g + k;
                q->storeTrajectoryPoint(t[h], Vector(x[h], y[h], z[h]), Vector(ux[h], uy[h], uz[h]),
                                        Vector(ax[h], ay[h], az[h]));
            }
        }
        b->particles.push_back(q);
    }
    return b;
}

void ref_beam_destroy(void* beam) { delete static_cast<RefBeam*>(beam); }

/* MeshedScreen<Bunch>::integrate over the given pixels: traces[n_pix][n_bins][6]
 * are accumulated (+=).  Particle subset [p_begin, p_end).  Returns 0, or 1 if
 * the reference threw. */
int ref_beam_integrate(void* beam, int p_begin, int p_end, int n_pix, const double* obs,
                       const double* t0, double dt, int n_bins, double* traces, int n_threads)
{
    RefBeam* b = static_cast<RefBeam*>(beam);
    if (n_threads <= 0) n_threads = omp_get_max_threads();
    int err = 0;
#pragma omp parallel for schedule(static) num_threads(n_threads) reduction(| : err)
    for (int i = 0; i < n_pix; i++) {
        try {
            FieldTrace tr(t0[i], dt, (std::size_t)(n_bins < 0 ? 0 : n_bins));
            double* out = traces + (std::size_t)i * (std::size_t)n_bins * 6;
            pack_trace(&tr, out);
            Vector P(obs[3 * i], obs[3 * i + 1], obs[3 * i + 2]);
            for (int p = p_begin; p < p_end; p++) b->particles[p]->integrateFieldTrace(P, &tr);
            unpack_trace(&tr, out);
        } catch (...) {
            err |= 1;
        }
    }
    return err;
}

/* The reference deposit driven by injected retarded times and fields
 * (n_pseudo pseudo-particles of n_points each; ts[n_pseudo][n_points],
 * f[n_pseudo][n_points][6]); trace[n_bins][6] accumulated. */
int ref_pseudo_integrate(int n_pseudo, int n_points, const double* ts, const double* f,
                         double t0, double dt, int n_bins, double* trace)
{
    try {
        FieldTrace tr(t0, dt, (std::size_t)n_bins);
        pack_trace(&tr, trace);
        for (int p = 0; p < n_pseudo; p++) {
            InjectedParticle q(n_points, ts + (std::size_t)p * n_points, f + (std::size_t)p * n_points * 6);
            q.integrateFieldTrace(Vector(0.0, 0.0, 0.0), &tr);
        }
        unpack_trace(&tr, trace);
    } catch (...) {
        return 1;
    }
    return 0;
}

/* Reference InitVay + StepVay through the undulator; outputs n_steps+1 points
 * per particle, SoA particle-major. */
void ref_track_undulator(int n_particles, const double* x0, const double* u0, double t_start,
                         double charge, double mass, double dt, int n_steps, double bpeak,
                         double lambda_u, int n_periods, const double* origin, double* t,
                         double* x, double* y, double* z, double* ux, double* uy, double* uz,
                         double* ax, double* ay, double* az)
{
    UndulatorField field(bpeak, lambda_u, n_periods, Vector(origin[0], origin[1], origin[2]));
    for (int p = 0; p < n_particles; p++) {
        ChargedParticle q(charge, mass);
        q.initTrajectory(t_start, Vector(x0[3 * p], x0[3 * p + 1], x0[3 * p + 2]),
                         Vector(u0[3 * p], u0[3 * p + 1], u0[3 * p + 2]), Vector(0.0, 0.0, 0.0));
        q.InitVay(dt, &field);
        for (int n = 0; n < n_steps; n++) q.StepVay(&field);
        std::size_t g = (std::size_t)p * (n_steps + 1);
        for (int k = 0; k <= n_steps; k++) {
            Vector X = q.TrajPoint(k), P = q.TrajMomentum(k), A = q.TrajAccel(k);
            t[g + k] = q.TrajTime(k);
            x[g + k] = X.x; y[g + k] = X.y; z[g + k] = X.z;
            ux[g + k] = P.x; uy[g + k] = P.y; uz[g + k] = P.z;
            ax[g + k] = A.x; ay[g + k] = A.y; az[g + k] = A.z;
        }
    }
}

/* Reference InitVay + StepVay in a homogeneous E/B field (HomogeneousField,
 * src/fields.cpp), one particle; used by the circular-loop checks. */
void ref_track_homogeneous(const double* E, const double* B, const double* x0, const double* u0,
                           double charge, double mass, double dt, int n_steps, double* t,
                           double* x, double* y, double* z, double* ux, double* uy, double* uz,
                           double* ax, double* ay, double* az)
{
    HomogeneousField field(Vector(E[0], E[1], E[2]), Vector(B[0], B[1], B[2]));
    ChargedParticle q(charge, mass);
    q.initTrajectory(0.0, Vector(x0[0], x0[1], x0[2]), Vector(u0[0], u0[1], u0[2]), Vector(0.0, 0.0, 0.0));
    q.InitVay(dt, &field);
    for (int n = 0; n < n_steps; n++) q.StepVay(&field);
    for (int k = 0; k <= n_steps; k++) {
        Vector X = q.TrajPoint(k), P = q.TrajMomentum(k), A = q.TrajAccel(k);
        t[k] = q.TrajTime(k);
        x[k] = X.x; y[k] = X.y; z[k] = X.z;
        ux[k] = P.x; uy[k] = P.y; uz[k] = P.z;
        ax[k] = A.x; ay[k] = A.y; az[k] = A.z;
    }
}

/* Reference InitVay + StepVay through the reference's Lattice (src/fields.cpp:
 * 194-239) holding the undulator (has_und) and a HomogeneousField E0, B0
 * (has_hom), added in that order; the Lattice owns and deletes both. */
void ref_track_lattice(int n_particles, const double* x0, const double* u0, double t_start, double charge,
                       double mass, double dt, int n_steps, int has_und, double bpeak, double lambda_u,
                       int n_periods, const double* origin, int has_hom, const double* E0, const double* B0,
                       double* t, double* x, double* y, double* z, double* ux, double* uy, double* uz,
                       double* ax, double* ay, double* az)
{
    Lattice lattice;
    if (has_und) lattice.addElement(new UndulatorField(bpeak, lambda_u, n_periods, Vector(origin[0], origin[1], origin[2])));
    if (has_hom) lattice.addElement(new HomogeneousField(Vector(E0[0], E0[1], E0[2]), Vector(B0[0], B0[1], B0[2])));
    for (int p = 0; p < n_particles; p++) {
        ChargedParticle q(charge, mass);
        q.initTrajectory(t_start, Vector(x0[3 * p], x0[3 * p + 1], x0[3 * p + 2]),
                         Vector(u0[3 * p], u0[3 * p + 1], u0[3 * p + 2]), Vector(0.0, 0.0, 0.0));
        q.InitVay(dt, &lattice);
        for (int n = 0; n < n_steps; n++) q.StepVay(&lattice);
        std::size_t g = (std::size_t)p * (n_steps + 1);
        for (int k = 0; k <= n_steps; k++) {
            Vector X = q.TrajPoint(k), P = q.TrajMomentum(k), A = q.TrajAccel(k);
            t[g + k] = q.TrajTime(k);
            x[g + k] = X.x; y[g + k] = X.y; z[g + k] = X.z;
            ux[g + k] = P.x; uy[g + k] = P.y; uz[g + k] = P.z;
            ax[g + k] = A.x; ay[g + k] = A.y; az[g + k] = A.z;
        }
    }
}

/* ChargedParticle::RetardedField(double t, Vector x) (src/particle.cpp:612-769)
 * of particle p of a beam: the single-time light-cone path. */
void ref_retarded_field_at(void* beam, int p, double obs_time, const double* obs, double* f)
{
    RefBeam* b = static_cast<RefBeam*>(beam);
    ElMagField EB = b->particles[p]->RetardedField(obs_time, Vector(obs[0], obs[1], obs[2]));
    f[0] = EB.E().x; f[1] = EB.E().y; f[2] = EB.E().z;
    f[3] = EB.B().x; f[4] = EB.B().y; f[5] = EB.B().z;
}

} // extern "C"
