/*
 * teufel_gpu.h - C ABI of the MI355X (gfx950) Lienard-Wiechert radiation
 * library: the post-tracking radiation step of TEUFEL on one GPU.
 *
 * Plain C: opaque context, plain pointers and sizes, int status codes.  No C++
 * exceptions and no framework types cross this boundary.  All entry points are
 * callable from one host thread per context (the reference calls
 * Observer::integrate() from one thread, src/teufel_hybrid.cpp:708); distinct
 * contexts may be used concurrently.
 *
 * Units and layouts follow the reference:
 *   trajectories  SoA, particle-major, n_points per particle:
 *                 t [s], x,y,z [m], ux,uy,uz = beta*gamma, ax,ay,az =
 *                 d(beta*gamma)/dt [1/s] exactly as ChargedParticle stores them
 *                 (src/particle.h:417-427); charge per particle in units of e.
 *   traces        [n_pix][n_bins][6] doubles (Ex,Ey,Ez [V/m], Bx,By,Bz [T]),
 *                 the getBuffer layout of src/mesh.cpp:375-399.  Sample i of a
 *                 pixel is the average field over [t0+(i-1/2)dt, t0+(i+1/2)dt).
 *
 * Semantics kept from the reference:
 *   - integration accumulates (+=) into the device trace store; tg_zero_screen
 *     clears it (MeshedScreen::zero, src/mesh.cpp:281-286);
 *   - a screen with n_bins < 2 and trajectories with n_points < 2 are silent
 *     no-ops (src/particle.cpp:778,783);
 *   - bucket indices are round-half-away-from-zero of (t_ret - t0)/dt and a
 *     step contributes only if t_ret2 > tmin && t_ret1 < tmax
 *     (src/particle.cpp:793,805,829-830).
 */
#ifndef TEUFEL_GPU_H
#define TEUFEL_GPU_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* status codes (0 = success).  The host adapter maps non-zero codes onto the
 * reference's IOexception (src/global.h:78-96). */
#define TG_OK 0
#define TG_ERR_ARG 1   /* invalid argument (null pointer, bad size, dt <= 0, non-finite) */
#define TG_ERR_OOM 2   /* device or host allocation failed */
#define TG_ERR_HIP 3   /* a HIP runtime call failed */
#define TG_ERR_RANGE 4 /* index out of range (screen id, particle range) */
#define TG_ERR_STATE 5 /* call sequence error (e.g. integrate before trajectories) */

/* precision of tg_integrate */
#define TG_FP64 0  /* plain FP64; timing arithmetic bit-compatible with the reference */
#define TG_MIXED 1 /* FP64 retarded time and 1-n.beta, FP32 field amplitudes, FP64 traces */

typedef struct tg_context tg_context;

/* Launch shape of the radiation kernel; zero fields mean "autotuned/default". */
typedef struct {
    int bins_per_window;     /* trace bins held in LDS per wave at a time (>= 1, capped by the LDS
                                size; 0: the autotuned or default window) */
    int particles_per_chunk; /* particles per work item (partial traces summed in order) */
    int use_graph;           /* 1: replay the launch sequence from a captured hipGraph; one capture
                                serves every particle range of the same size (tg_integrate_range
                                over equal batches replays it, shifted to each batch) */
    int window_groups;       /* windows of a (tile, chunk) split into this many work items */
    int walk;                /* 1: the wave walks each particle's steps together (lanes share a
                                step); 2: every lane (pixel) walks its own steps; 0: default (2) */
    int spill;               /* per-lane walk: 1 = each window holds one more bucket that takes the
                                steps crossing into the next window whole, 2 = without it (such steps
                                are walked again in the next window); 0: autotuned or default (1) */
} tg_tuning;

/* Context on one device; one trace store and one trajectory set per context.
 * Replaces the per-rank observer loop + MPI_Reduce of
 * src/teufel_hybrid.cpp:691-763 with an on-device store. */
int tg_create(int device, tg_context** ctx);
int tg_destroy(tg_context* ctx);
/* Text of the last error on this context ("" if none); valid until the next call. */
const char* tg_last_error(const tg_context* ctx);
/* Device name and compute unit count, for reports. */
int tg_device_info(const tg_context* ctx, char* name, int name_len, int* n_cu);

/* Trajectories of all particles (Beam order: bunches, then particles; the
 * summation order of Beam/Bunch::integrateFieldTrace, src/beam.cpp:443-450,
 * src/bunch.cpp:739-745).  Host arrays of n_particles*n_points doubles,
 * copied; the caller keeps ownership.  Replaces the particle trajectory store
 * read by ChargedParticle::integrateFieldTrace (src/particle.cpp:771-867). */
int tg_set_trajectories(tg_context* ctx, int n_particles, int n_points,
                        const double* t, const double* x, const double* y, const double* z,
                        const double* ux, const double* uy, const double* uz,
                        const double* ax, const double* ay, const double* az,
                        const double* charge);

/* Same from device-resident arrays (e.g. produced by tg_track_undulator or by
 * the caller's own kernels); copied, the caller keeps ownership. */
int tg_set_trajectories_device(tg_context* ctx, int n_particles, int n_points,
                               const double* t, const double* x, const double* y, const double* z,
                               const double* ux, const double* uy, const double* uz,
                               const double* ax, const double* ay, const double* az,
                               const double* charge);

/* A screen: n_pix observation points obs[n_pix][3] with per-pixel t0[n_pix],
 * common dt and n_bins.  Mirrors the FieldTrace(t0,dt,N) per pixel of
 * MeshedScreen (src/mesh.cpp:169-175), ScreenObserver (src/screen.cpp:74-88)
 * and PointObserver (src/point_observer.cpp:36-45).  The trace store starts
 * zeroed. */
int tg_add_screen(tg_context* ctx, int n_pix, const double* obs, const double* t0,
                  double dt, int n_bins, int* screen_id);
int tg_zero_screen(tg_context* ctx, int screen_id);
int tg_remove_screens(tg_context* ctx);
int tg_screen_count(const tg_context* ctx, int* n_screens);

/* Observer::integrate() (src/observer.h:140) for every screen of the context,
 * in one batched launch: adds the fields of all particles to every pixel's
 * trace.  precision: TG_FP64 or TG_MIXED.  tuning may be NULL.  Returns after
 * the work is queued on the context's stream; tg_synchronize waits for it.
 * TG_ERR_ARG when a trajectory step moves faster than light or back in time
 * (the retarded time must increase along every trajectory, as it does for
 * any particle slower than light), and for TG_MIXED when the beam is far off
 * the z axis (|beta_x|+|beta_y| > 0.25); both are checked once per
 * trajectory set and precision. */
int tg_integrate(tg_context* ctx, int precision, const tg_tuning* tuning);
int tg_synchronize(tg_context* ctx);

/* As tg_integrate for a particle subset [p_begin, p_end) and one screen
 * (screen_id >= 0) or all screens (screen_id < 0). */
int tg_integrate_range(tg_context* ctx, int precision, const tg_tuning* tuning,
                       int screen_id, int p_begin, int p_end);

/* No counterpart in the reference: for callers that cannot allocate device
 * memory while they run (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  As tg_integrate_range, with all launch
 * scratch (the work-list counter, the screen descriptors, the work list and
 * the per-wave cursor slots) in a static device array of
 * TG_STATIC_SCRATCH_BYTES per device instead of buffers the library
 * allocates, and the particle range as one chunk, so that the fields are added
 * straight into the trace store (no partial traces).  It calls no hipMalloc or
 * hipFree and no allocating library routine (the work list is planned in host
 * memory).  Set-up comes first: the trajectory records of the precision must
 * be packed, and the screens tiled, by an earlier tg_integrate,
 * tg_integrate_range or tg_autotune of that precision (TG_ERR_STATE
 * otherwise).  A launch whose scratch would exceed the array fails with
 * TG_ERR_RANGE, naming how many particles per call fit.  All contexts on one
 * device share that device's array, which keeps the last call's contents:
 * calls that overlap on one GPU race.  Synchronizes the context's stream. */
#define TG_STATIC_SCRATCH_BYTES (256u << 20)
int tg_integrate_range_static(tg_context* ctx, int precision, const tg_tuning* tuning,
                              int screen_id, int p_begin, int p_end);

/* Copy a screen's traces to host, out[n_pix][n_bins][6]; accumulate = 1 adds
 * into out (FieldTrace += semantics, src/fieldtrace.cpp:185-190), 0 overwrites.
 * Replaces getBuffer + MPI_Reduce + fromBuffer (src/mesh.cpp:375-425,
 * src/teufel_hybrid.cpp:726-763).  Synchronizes the context's stream. */
int tg_read_traces(tg_context* ctx, int screen_id, double* out, int accumulate);
/* Device pointer to a screen's trace store ([n_pix][n_bins][6] doubles). */
int tg_screen_traces_device(tg_context* ctx, int screen_id, double** traces);

/* Per-pixel radiated energy density on device: S_p = dt * sum_i E_i x B_i / mu0
 * (FieldTrace::Poynting, src/fieldtrace.cpp:192-198), out[n_pix][3] on host. */
int tg_screen_poynting(tg_context* ctx, int screen_id, double* out);

/* Spectral power density of the two in-plane polarisations of every pixel
 * [J/(m² Hz)], computed on device (rocFFT): with EH = E.ex, EV = E.ey,
 * BH = B.ex, BV = B.ey, dens_h[p][i] = |F(EH) conj(F(BV))|_i / mu0 * 2 dt^2
 * and dens_v likewise from EV and BH, for i < n_bins/2; freq[i] =
 * i (1/dt) / (n_bins - 1).  Replaces spectralDensityH / spectralDensityV of
 * the reference's post-processing (scripts/teufel.py:198-226, the same
 * frequency axis, np.linspace(0, 1/dt, n_bins)[:n_bins//2]), with mu0 of
 * src/global.h:37-44.  ex, ey: unit vectors of the screen plane (3 doubles
 * each); freq may be null; dens_h, dens_v: [n_pix][n_bins/2] on host. */
int tg_screen_spectra(tg_context* ctx, int screen_id, const double* ex, const double* ey, double* freq,
                      double* dens_h, double* dens_v);

/* Number of (pixel, particle step) pairs that pass the window test of
 * ChargedParticle::integrateFieldTrace, ts2 > tmin && ts1 < tmax
 * (src/particle.cpp:778-793), summed over the particles and the pixels of
 * screen screen_id (-1: all screens).  The measure of how much of the
 * N_pairs = n_pix x n_particles x (n_points - 1) an integrate deposits; the
 * retarded times are those of the FP64 kernel.  No counterpart in the
 * reference, which has no such count. */
int tg_count_in_window(tg_context* ctx, int screen_id, long long* n_inwin);

/* The bucket deposit alone, driven by injected per-point retarded times and
 * fields, as the Special particle of tests/teufel.integrate_field.cpp:51-82:
 * ts[n_pseudo][n_points], fields[n_pseudo][n_points][6] (host), trace
 * [n_bins][6] (host, accumulated).  Runs the device deposit code of
 * tg_integrate.  Replaces ChargedParticle::integrateFieldTrace with
 * overridden RetardedTime/RetardedField (src/particle.cpp:771-867). */
int tg_deposit_only(tg_context* ctx, int n_pseudo, int n_points, const double* ts,
                    const double* fields, double t0, double dt, int n_bins, double* trace);

/* The field at given times and points: ChargedParticle::RetardedField(double
 * t, Vector x) (src/particle.cpp:612-769) summed over particles [p_begin,
 * p_end) in order, as Bunch/Beam::RetardedField (src/bunch.cpp:729-737,
 * src/beam.cpp:433-441) sum them; SnapshotObserver::integrate
 * (src/snapshot.cpp:72-91) is one call with one time for all cells.
 * obs_time[n_obs], obs_pos[n_obs][3], fields[n_obs][6] (host; accumulate = 1
 * adds into fields).  A point before a particle's first trajectory point is
 * back-extrapolated with its start velocity, one after its last sees nothing
 * from it (the reference's rules).  TG_ERR_STATE without trajectories. */
int tg_fields_at(tg_context* ctx, int n_obs, const double* obs_time, const double* obs_pos,
                 int p_begin, int p_end, double* fields, int accumulate);

/* Pick tuning for the context's current screens and trajectories by timing
 * candidate launch shapes on the device; the result is used by later
 * tg_integrate calls with tuning == NULL and written to `out` if non-NULL. */
int tg_autotune(tg_context* ctx, int precision, tg_tuning* out);

/* Elapsed device time (ms) of the last tg_integrate's radiation kernel and of
 * the whole launch sequence, measured with HIP events on the context's stream. */
int tg_last_timing(tg_context* ctx, float* kernel_ms, float* total_ms);

/* Diagnostic with no counterpart in the reference: when the environment has
 * TG_STATS=1 at tg_create, the last tg_integrate's radiation kernel counts the
 * trajectory steps each wavefront evaluated (wave_steps, 64 pixels each), the
 * particle walks (one per work item, window and contributing particle) and
 * the work items; otherwise all three read 0. */
int tg_last_stats(tg_context* ctx, long long* wave_steps, long long* walks, long long* work_items);

/* Trajectory generation on the device (the tracking step that precedes the
 * radiation step): Vay pusher (J.-L. Vay 2008) through a planar undulator,
 * ChargedParticle::InitVay/StepVay (src/particle.cpp:407-512) with
 * PlanarUndulator::LocalField (src/undulator.cpp:107-126) placed at origin[3].
 * x0[n][3], u0[n][3] initial positions/momenta (host), all particles start at
 * t_start; charge/mass in units of e / m_e; n_steps+1 points per particle
 * become the context's trajectories (replacing any previous set). */
int tg_track_undulator(tg_context* ctx, int n_particles, const double* x0, const double* u0,
                       double t_start, double charge, double mass, double dt, int n_steps,
                       double bpeak, double lambda_u, int n_periods, const double* origin);

/* The same pusher through a lattice of the planar undulator and a
 * homogeneous field: Lattice::Field (src/fields.cpp:229-239) summing the
 * undulator's field and then HomogeneousField(E0, B0) (src/fields.cpp:
 * 155-163).  origin == NULL: no undulator (bpeak, lambda_u, n_periods
 * unused).  E0 [V/m], B0 [T] (3 doubles each): the homogeneous field; a NULL
 * one is zero, both NULL: no homogeneous element.  tg_track_undulator is the
 * case E0 = B0 = NULL. */
int tg_track_lattice(tg_context* ctx, int n_particles, const double* x0, const double* u0,
                     double t_start, double charge, double mass, double dt, int n_steps,
                     double bpeak, double lambda_u, int n_periods, const double* origin,
                     const double* E0, const double* B0);

/* Copy particles [p_begin, p_end) of the context's trajectories to host SoA
 * arrays ((p_end-p_begin)*n_points each; any pointer may be NULL). */
int tg_get_trajectories(tg_context* ctx, int p_begin, int p_end,
                        double* t, double* x, double* y, double* z,
                        double* ux, double* uy, double* uz,
                        double* ax, double* ay, double* az, double* charge);
int tg_trajectory_shape(const tg_context* ctx, int* n_particles, int* n_points);

#ifdef __cplusplus
}
#endif
#endif
