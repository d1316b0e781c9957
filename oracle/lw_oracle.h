/*
 * lw_oracle.h - FP64 CPU restatement of TEUFEL's post-tracking radiation path.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the teufel_amd product links, loads or
 * calls this library; only tests/, __graft_entry__.smoke() and the cpu_baseline
 * leg of bench.py use it, and only as the checker.
 *
 * Every function restates one reference function (file:line cited on each).
 * Arithmetic follows the reference operation by operation (same association
 * order, divisions where the reference divides, no FMA contraction: the file is
 * compiled with -ffp-contract=off), so that on x86-64 it reproduces the
 * reference's doubles bit for bit.  Which oracle functions are pinned by which
 * check is listed in the header of each test file under tests/.
 *
 * Trajectory layout: structure of arrays, particle-major, n_points per particle:
 *   t[p*n_points + k], x/y/z (position, m), ux/uy/uz (beta*gamma),
 *   ax/ay/az (d(beta*gamma)/dt, 1/s), charge[p] (units of e).
 * Trace layout: [n_pix][n_bins][6] doubles, (Ex,Ey,Ez,Bx,By,Bz) per sample,
 * the layout of MeshedScreen::getBuffer (src/mesh.cpp:375-399).
 */
#ifndef LW_ORACLE_H
#define LW_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* Physical constants of src/global.h:37-44 (CODATA-2014 e and eps0). */
#define OR_ELEMENTARY_CHARGE 1.6021766208e-19
#define OR_MEC2 0.5109989461e6
#define OR_SPEED_OF_LIGHT 2.99792458e8
#define OR_PI 3.1415926535897932384626433832795029
#define OR_EPS0 8.854187817e-12
#define OR_MU0 (4 * OR_PI * 1e-7)
#define OR_INV_REST_MASS 1.758820025e11

typedef struct {
    int n_particles;
    int n_points;
    const double *t, *x, *y, *z;
    const double *ux, *uy, *uz;
    const double *ax, *ay, *az;
    const double *charge;
} or_trajectories;

/* ChargedParticle::RetardedTime(int, Vector)   src/particle.cpp:560-568 */
double or_retarded_time(const or_trajectories* tr, int p, int k, const double obs[3]);

/* ChargedParticle::RetardedField(int, Vector)  src/particle.cpp:570-596
 * f = (Ex,Ey,Ez,Bx,By,Bz). */
void or_retarded_field(const or_trajectories* tr, int p, int k, const double obs[3], double f[6]);

/* ChargedParticle::RetardedField(double, Vector)  src/particle.cpp:612-769:
 * the field of particle p at obs at time obs_time, from the point on its
 * trajectory on the observer's past light cone (bisection to one step,
 * secant refinement, linear interpolation; back-extrapolation with the start
 * velocity before the trajectory; zero after its end).  Returns 1 for an
 * empty trajectory (the reference throws). */
int or_field_at(const or_trajectories* tr, int p, double obs_time, const double obs[3], double f[6]);

/* The per-step bucket deposit of ChargedParticle::integrateFieldTrace,
 * src/particle.cpp:788-865, for one trajectory step whose end points are
 * observed at ts1, ts2 with fields f1, f2.  trace = [n_bins][6]. */
int or_deposit_step(double ts1, double ts2, const double f1[6], const double f2[6],
                    double t0, double dt, int n_bins, double* trace);

/* ChargedParticle::integrateFieldTrace          src/particle.cpp:771-867
 * Accumulates (+=) the field of particle p observed at obs into trace. */
int or_integrate_particle(const or_trajectories* tr, int p, const double obs[3],
                          double t0, double dt, int n_bins, double* trace);

/* integrateFieldTrace driven by precomputed per-point retarded times and
 * fields, as the Special test particle of tests/teufel.integrate_field.cpp:51-82
 * injects them through the virtual RetardedTime/RetardedField overrides.
 * ts[n_points], f[n_points][6]. */
int or_integrate_pseudo(int n_points, const double* ts, const double* f,
                        double t0, double dt, int n_bins, double* trace);

/* Bunch::integrateFieldTrace (src/bunch.cpp:739-745) and
 * Beam::integrateFieldTrace (src/beam.cpp:443-450): particles in index order
 * (bunches concatenated in order) into one trace. */
int or_integrate_bunch(const or_trajectories* tr, const double obs[3],
                       double t0, double dt, int n_bins, double* trace);

/* MeshedScreen<Beam|Bunch>::integrate            src/mesh.cpp:288-342
 * OpenMP parallel loop over pixels; obs[n_pix][3], t0[n_pix];
 * traces[n_pix][n_bins][6] accumulated.  n_threads <= 0: OpenMP default. */
int or_integrate_screen(const or_trajectories* tr, int n_pix, const double* obs,
                        const double* t0, double dt, int n_bins, double* traces,
                        int n_threads);

/* Number of (pixel, particle step) pairs that pass the window test of
 * ChargedParticle::integrateFieldTrace                 src/particle.cpp:778-793
 * (ts2 > tmin && ts1 < tmax), summed over pixels and particles. */
long long or_count_in_window(const or_trajectories* tr, int n_pix, const double* obs, const double* t0,
                             double dt, int n_bins);

/* ScreenObserver::CellPosition / CellTimeZero     src/screen.cpp:98-110
 * Fills obs[nx*ny][3] and t0[nx*ny] in the getBuffer order of
 * src/screen.cpp:248-271 (ix outer, iy inner). */
void or_screen_cells(const double origin[3], const double dX[3], const double dY[3],
                     int nx, int ny, double t0_obs, double dtx, double dty,
                     double* obs, double* t0);

/* FieldTrace::Poynting                            src/fieldtrace.cpp:192-198
 * S = dt * sum_i E_i x B_i / mu0 (J/m^2). */
void or_trace_poynting(const double* trace, int n_bins, double dt, double S[3]);

/* MeshedScreen::totalEnergy                       src/mesh.cpp:744-754
 * -sum_p area_p * S_p . normal_p */
double or_total_energy(int n_pix, const double* traces, int n_bins, double dt,
                       const double* area, const double* normal);

/* Planar undulator field, PlanarUndulator::Setup/LocalField
 * src/undulator.cpp:66-80,107-126, placed at `origin` as LocalizedField::Field
 * (src/fields.cpp:180-183).  Returns B (E is zero). */
void or_undulator_field(double bpeak, double lambda_u, int n_periods,
                        const double origin[3], const double X[3], double B[3]);

/* Vay pusher through a planar undulator: ChargedParticle::initTrajectory +
 * InitVay + n_steps x StepVay (src/particle.cpp:326-341,407-442,479-512),
 * with charge/mass in units of e / m_e.  Writes n_steps+1 points per particle
 * (SoA, particle-major) into the output arrays.  Initial acceleration zero as
 * the reference tests pass it (tests/teufel.undulator.cpp:113-117). */
void or_track_undulator(int n_particles, const double* x0, const double* u0, double t_start,
                        double charge, double mass, double dt, int n_steps,
                        double bpeak, double lambda_u, int n_periods, const double origin[3],
                        double* t, double* x, double* y, double* z,
                        double* ux, double* uy, double* uz,
                        double* ax, double* ay, double* az, int n_threads);

/* The same pusher through a lattice (Lattice::Field, src/fields.cpp:229-239)
 * of the undulator (has_und) and a homogeneous field E0 [V/m], B0 [T]
 * (has_hom; HomogeneousField, src/fields.cpp:155-163), summed in that order.
 * or_track_undulator is the case has_und = 1, has_hom = 0. */
void or_track_lattice(int n_particles, const double* x0, const double* u0, double t_start,
                      double charge, double mass, double dt, int n_steps,
                      int has_und, double bpeak, double lambda_u, int n_periods, const double origin[3],
                      int has_hom, const double E0[3], const double B0[3],
                      double* t, double* x, double* y, double* z,
                      double* ux, double* uy, double* uz,
                      double* ax, double* ay, double* az, int n_threads);

#ifdef __cplusplus
}
#endif
#endif
