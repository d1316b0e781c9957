/*
 * lw_oracle.c - FP64 CPU restatement of TEUFEL's post-tracking radiation path.
 *
 * TEST INFRASTRUCTURE ONLY (the checker, never the product).  See lw_oracle.h
 * for the layout conventions.  Each function cites the reference code it
 * restates; arithmetic keeps the reference's operation order (compile with
 * -ffp-contract=off) so results match the reference bit for bit on x86-64.
 */
#include "lw_oracle.h"

#include <math.h>
#include <stddef.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* ---- 3-vector helpers in the operation order of src/vector.cpp ---------- */

typedef struct { double x, y, z; } vec3;

static vec3 v_make(double x, double y, double z) { vec3 r = {x, y, z}; return r; }
static vec3 v_add(vec3 a, vec3 b) { return v_make(a.x + b.x, a.y + b.y, a.z + b.z); }      /* vector.cpp:36-43 */
static vec3 v_sub(vec3 a, vec3 b) { return v_make(a.x - b.x, a.y - b.y, a.z - b.z); }
static vec3 v_mul(vec3 a, double f) { return v_make(a.x * f, a.y * f, a.z * f); }          /* vector.cpp operator* */
static vec3 v_div(vec3 a, double f) { return v_make(a.x / f, a.y / f, a.z / f); }          /* vector.cpp:111-118 */
static double v_abs2nd(vec3 a) { return a.x * a.x + a.y * a.y + a.z * a.z; }               /* vector.cpp:132-135 */
static double v_norm(vec3 a) { return sqrt(a.x * a.x + a.y * a.y + a.z * a.z); }           /* vector.cpp:127-130 */
static double v_dot(vec3 a, vec3 b)                                                         /* vector.cpp:153-160 */
{
    double t = a.x * b.x;
    t += a.y * b.y;
    t += a.z * b.z;
    return t;
}
static vec3 v_cross(vec3 a, vec3 b)                                                         /* vector.cpp:162-169 */
{
    return v_make(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}

/* ---- ElMagField helpers (src/fields.cpp:50-108) ------------------------- */

static void f_scale(const double a[6], double s, double r[6]) { for (int i = 0; i < 6; i++) r[i] = a[i] * s; }
static void f_sum(const double a[6], const double b[6], double r[6]) { for (int i = 0; i < 6; i++) r[i] = a[i] + b[i]; }

/* f1*wa + f2*wb as ElMagField expressions (products first, then the sum) */
static void f_lerp(const double f1[6], double wa, const double f2[6], double wb, double r[6])
{
    double a[6], b[6];
    f_scale(f1, wa, a);
    f_scale(f2, wb, b);
    f_sum(a, b, r);
}

/* FieldTrace::add  src/fieldtrace.cpp:185-190 (index is a size_t there, so a
 * negative index is out of range as well) */
static int trace_add(double* trace, int n_bins, long idx, const double f[6])
{
    if (idx < 0 || idx >= n_bins) return 1;
    double* s = trace + 6 * (size_t)idx;
    for (int i = 0; i < 6; i++) s[i] += f[i];
    return 0;
}

/* ---- trajectory access -------------------------------------------------- */

static vec3 traj_x(const or_trajectories* tr, size_t g) { return v_make(tr->x[g], tr->y[g], tr->z[g]); }
static vec3 traj_u(const or_trajectories* tr, size_t g) { return v_make(tr->ux[g], tr->uy[g], tr->uz[g]); }
static vec3 traj_a(const or_trajectories* tr, size_t g) { return v_make(tr->ax[g], tr->ay[g], tr->az[g]); }

double or_retarded_time(const or_trajectories* tr, int p, int k, const double obs[3])
{
    /* particle.cpp:560-568 */
    size_t g = (size_t)p * (size_t)tr->n_points + (size_t)k;
    double t = tr->t[g];
    vec3 RVec = v_sub(v_make(obs[0], obs[1], obs[2]), traj_x(tr, g));
    double R = v_norm(RVec);
    return t + R / OR_SPEED_OF_LIGHT;
}

void or_retarded_field(const or_trajectories* tr, int p, int k, const double obs[3], double f[6])
{
    /* particle.cpp:570-596 */
    size_t g = (size_t)p * (size_t)tr->n_points + (size_t)k;
    vec3 SourceX = traj_x(tr, g);
    vec3 SourceBeta = traj_u(tr, g);
    vec3 SourceBetaPrime = traj_a(tr, g);
    double betagamma2 = v_abs2nd(SourceBeta);
    double gamma2 = betagamma2 + 1.0;
    double gamma = sqrt(gamma2);
    SourceBeta = v_div(SourceBeta, gamma);
    SourceBetaPrime = v_div(SourceBetaPrime, gamma);
    vec3 RVec = v_sub(v_make(obs[0], obs[1], obs[2]), SourceX);
    double R = v_norm(RVec);
    vec3 N = v_div(RVec, v_norm(RVec));                       /* Vector::normalize vector.cpp:137-143 */
    double scale = tr->charge[p] * OR_ELEMENTARY_CHARGE / (4.0 * OR_PI * OR_EPS0);
    double bn3rd = pow(1.0 - v_dot(SourceBeta, N), 3.0);
    vec3 NmB = v_sub(N, SourceBeta);
    vec3 EField = v_div(NmB, R * R * bn3rd * gamma2);
    EField = v_add(EField, v_div(v_cross(N, v_cross(NmB, SourceBetaPrime)), R * bn3rd * OR_SPEED_OF_LIGHT));
    EField = v_mul(EField, scale);
    vec3 BField = v_div(v_cross(N, EField), OR_SPEED_OF_LIGHT);
    f[0] = EField.x; f[1] = EField.y; f[2] = EField.z;
    f[3] = BField.x; f[4] = BField.y; f[5] = BField.z;
}

int or_field_at(const or_trajectories* tr, int p, double obs_time, const double obs[3], double f[6])
{
    /* ChargedParticle::RetardedField(double, Vector)  particle.cpp:612-769 */
    const int NP = tr->n_points;
    const size_t base = (size_t)p * (size_t)NP;
    for (int i = 0; i < 6; i++) f[i] = 0.0;
    if (NP < 1) return 1;                                       /* particle.cpp:615 RANGEexception */
    vec3 O = v_make(obs[0], obs[1], obs[2]);
    int i1 = 0;
    double R1 = v_norm(v_sub(O, traj_x(tr, base + i1)));
    double t1 = obs_time - (tr->t[base + i1] + R1 / OR_SPEED_OF_LIGHT);
    int i2 = NP - 1;
    double R2 = v_norm(v_sub(O, traj_x(tr, base + i2)));
    double t2 = obs_time - (tr->t[base + i2] + R2 / OR_SPEED_OF_LIGHT);
    if (!(t2 <= 0.0)) return 0;                                 /* particle.cpp:633: no signal yet */
    vec3 SourceX, SourceBeta, SourceBetaPrime;
    double gamma2;
    if (t1 <= 0.0) {
        /* particle.cpp:641-660: back-extrapolation with the start velocity */
        vec3 Rc = v_div(v_sub(O, traj_x(tr, base)), OR_SPEED_OF_LIGHT);
        SourceBeta = traj_u(tr, base);
        double betagamma2 = v_abs2nd(SourceBeta);
        gamma2 = betagamma2 + 1.0;
        double gamma = sqrt(gamma2);
        SourceBeta = v_div(SourceBeta, gamma);
        double t0 = obs_time;
        double bR = v_dot(SourceBeta, Rc);
        double bb = v_dot(SourceBeta, SourceBeta);
        double RR = v_dot(Rc, Rc);
        double SourceT = (t0 - bR - sqrt((t0 - bR) * (t0 - bR) - (1 - bb) * (t0 * t0 - RR))) / (1 - bb);
        SourceX = v_add(traj_x(tr, base), v_mul(v_mul(SourceBeta, OR_SPEED_OF_LIGHT), SourceT));
        SourceBetaPrime = v_make(0.0, 0.0, 0.0);
    } else {
        /* particle.cpp:662-745: bisection to one step, then secant refinement */
        while (i2 - i1 > 1) {
            int iMid = (i2 + i1) / 2;
            double RMid = v_norm(v_sub(O, traj_x(tr, base + iMid)));
            double tMid = obs_time - (tr->t[base + iMid] + RMid / OR_SPEED_OF_LIGHT);
            if (tMid > 0.0) {
                i1 = iMid;
                t1 = tMid;
            } else {
                i2 = iMid;
                t2 = tMid;
            }
        }
        double frac = t1 / (t1 - t2);
        double T1 = tr->t[base + i1], T2 = tr->t[base + i2];
        vec3 X1 = traj_x(tr, base + i1), X2 = traj_x(tr, base + i2);
        double SourceT = T1 * (1.0 - frac) + T2 * frac;
        SourceX = v_add(v_mul(X1, 1.0 - frac), v_mul(X2, frac));
        double R = v_norm(v_sub(O, SourceX));
        double dt = SourceT + R / OR_SPEED_OF_LIGHT - obs_time;
        double r1 = v_norm(v_sub(O, X1));
        double r2 = v_norm(v_sub(O, X2));
        double dtdfrac = (T2 - T1) + (r2 - r1) / OR_SPEED_OF_LIGHT;
        double dfrac = dt / dtdfrac;
        int number_ref = 1;
        while (fabs(dfrac) > 1e-5 && number_ref < 10) {
            double Lastdt = dt;
            frac -= dfrac;
            SourceT = T1 * (1.0 - frac) + T2 * frac;
            SourceX = v_add(v_mul(X1, 1.0 - frac), v_mul(X2, frac));
            R = v_norm(v_sub(O, SourceX));
            dt = SourceT + R / OR_SPEED_OF_LIGHT - obs_time;
            dtdfrac = -(dt - Lastdt) / dfrac;
            dfrac = dt / dtdfrac;
            number_ref++;
        }
        frac -= dfrac;
        SourceX = v_add(v_mul(X1, 1.0 - frac), v_mul(X2, frac));
        SourceBeta = v_add(v_mul(traj_u(tr, base + i1), 1.0 - frac), v_mul(traj_u(tr, base + i2), frac));
        double betagamma2 = v_abs2nd(SourceBeta);
        gamma2 = betagamma2 + 1.0;
        double gamma = sqrt(gamma2);
        SourceBeta = v_div(SourceBeta, gamma);
        SourceBetaPrime = v_add(v_mul(traj_a(tr, base + i1), 1.0 - frac), v_mul(traj_a(tr, base + i2), frac));
        SourceBetaPrime = v_div(SourceBetaPrime, gamma);
    }
    /* particle.cpp:747-766 */
    vec3 RVec = v_sub(O, SourceX);
    double R = v_norm(RVec);
    vec3 N = v_div(RVec, v_norm(RVec));
    double scale = tr->charge[p] * OR_ELEMENTARY_CHARGE / (4.0 * OR_PI * OR_EPS0);
    double bn3rd = pow(1.0 - v_dot(SourceBeta, N), 3.0);
    vec3 NmB = v_sub(N, SourceBeta);
    vec3 EField = v_div(NmB, R * R * bn3rd * gamma2);
    EField = v_add(EField, v_div(v_cross(N, v_cross(NmB, SourceBetaPrime)), R * bn3rd * OR_SPEED_OF_LIGHT));
    EField = v_mul(EField, scale);
    vec3 BField = v_div(v_cross(N, EField), OR_SPEED_OF_LIGHT);
    f[0] = EField.x; f[1] = EField.y; f[2] = EField.z;
    f[3] = BField.x; f[4] = BField.y; f[5] = BField.z;
    return 0;
}

int or_deposit_step(double ts1, double ts2, const double f1[6], const double f2[6],
                    double t0, double dt, int nots, double* trace)
{
    /* particle.cpp:788-865, body of the step loop after RetardedTime */
    int err = 0;
    double tmin = t0 - dt * 0.5;
    double tmax = t0 + dt * ((double)nots - 0.5);
    double dts = ts2 - ts1;
    double field[6];
    if ((ts2 > tmin) && (ts1 < tmax)) {
        if (ts1 <= tmin) {
            /* step starts before the window (and may cover all of it) */
            int idx2 = (int)round((ts2 - t0) / dt);
            if (idx2 > nots) idx2 = nots;
            for (int idx = 0; idx < idx2; idx++) {
                double t_center = t0 + idx * dt;
                f_lerp(f1, (ts2 - t_center) / dts, f2, (t_center - ts1) / dts, field);
                err |= trace_add(trace, nots, idx, field);
            }
            if (idx2 < nots) {
                double t_start = t0 + (idx2 - 0.5) * dt;
                double f_start[6], s[6], h[6];
                f_lerp(f1, (ts2 - t_start) / dts, f2, (t_start - ts1) / dts, f_start);
                f_sum(f_start, f2, s);
                f_scale(s, 0.5, h);
                f_scale(h, (ts2 - t_start) / dt, field);
                err |= trace_add(trace, nots, idx2, field);
            }
        } else {
            int idx1 = (int)round((ts1 - t0) / dt);
            int idx2 = (int)round((ts2 - t0) / dt);
            if (idx2 > nots) idx2 = nots;
            if (idx1 == idx2) {
                double s[6], h[6];
                f_sum(f1, f2, s);
                f_scale(s, 0.5, h);
                f_scale(h, dts / dt, field);
                err |= trace_add(trace, nots, idx1, field);
            } else {
                double t_end = t0 + (idx1 + 0.5) * dt;
                double f_end[6], s[6], h[6];
                f_lerp(f1, (ts2 - t_end) / dts, f2, (t_end - ts1) / dts, f_end);
                f_sum(f1, f_end, s);
                f_scale(s, 0.5, h);
                f_scale(h, (t_end - ts1) / dt, field);
                err |= trace_add(trace, nots, idx1, field);
                for (int idx = idx1 + 1; idx < idx2; idx++) {
                    double t_center = t0 + idx * dt;
                    f_lerp(f1, (ts2 - t_center) / dts, f2, (t_center - ts1) / dts, field);
                    err |= trace_add(trace, nots, idx, field);
                }
                if (idx2 < nots) {
                    double t_start = t0 + (idx2 - 0.5) * dt;
                    double f_start[6];
                    f_lerp(f1, (ts2 - t_start) / dts, f2, (t_start - ts1) / dts, f_start);
                    f_sum(f_start, f2, s);
                    f_scale(s, 0.5, h);
                    f_scale(h, (ts2 - t_start) / dt, field);
                    err |= trace_add(trace, nots, idx2, field);
                }
            }
        }
    }
    return err;
}

int or_integrate_particle(const or_trajectories* tr, int p, const double obs[3],
                          double t0, double dt, int nots, double* trace)
{
    /* particle.cpp:771-867: every trajectory point's retarded time and field
     * are evaluated twice (once as step end, once as step start), as in the
     * reference; both evaluations give identical doubles. */
    int err = 0;
    if (nots < 2) return 0;
    double tmin = t0 - dt * 0.5;
    double tmax = t0 + dt * ((double)nots - 0.5);
    int NP = tr->n_points;
    if (NP < 2) return 0;
    for (int step = 0; step < NP - 1; step++) {
        double ts1 = or_retarded_time(tr, p, step, obs);
        double ts2 = or_retarded_time(tr, p, step + 1, obs);
        if ((ts2 > tmin) && (ts1 < tmax)) {
            double f1[6], f2[6];
            or_retarded_field(tr, p, step, obs, f1);
            or_retarded_field(tr, p, step + 1, obs, f2);
            err |= or_deposit_step(ts1, ts2, f1, f2, t0, dt, nots, trace);
        }
    }
    return err;
}

int or_integrate_pseudo(int n_points, const double* ts, const double* f,
                        double t0, double dt, int nots, double* trace)
{
    int err = 0;
    if (nots < 2 || n_points < 2) return 0;
    for (int step = 0; step < n_points - 1; step++)
        err |= or_deposit_step(ts[step], ts[step + 1], f + 6 * step, f + 6 * (step + 1), t0, dt, nots, trace);
    return err;
}

int or_integrate_bunch(const or_trajectories* tr, const double obs[3],
                       double t0, double dt, int nots, double* trace)
{
    /* bunch.cpp:739-745 / beam.cpp:443-450 */
    int err = 0;
    for (int p = 0; p < tr->n_particles; p++)
        err |= or_integrate_particle(tr, p, obs, t0, dt, nots, trace);
    return err;
}

int or_integrate_screen(const or_trajectories* tr, int n_pix, const double* obs,
                        const double* t0, double dt, int nots, double* traces,
                        int n_threads)
{
    /* mesh.cpp:288-342: parallel for over pixels, each trace has one writer */
    int err = 0;
#ifdef _OPENMP
    if (n_threads <= 0) n_threads = omp_get_max_threads();
#pragma omp parallel for schedule(static) num_threads(n_threads) reduction(| : err)
#endif
    for (int i = 0; i < n_pix; i++)
        err |= or_integrate_bunch(tr, obs + 3 * (size_t)i, t0[i], dt, nots,
                                  traces + (size_t)i * (size_t)nots * 6);
    (void)n_threads;
    return err;
}

long long or_count_in_window(const or_trajectories* tr, int n_pix, const double* obs, const double* t0,
                             double dt, int nots)
{
    /* particle.cpp:778-793: the steps of every (pixel, particle) pair that
     * pass the window test ts2 > tmin && ts1 < tmax (the silent no-ops for
     * nots < 2 and NP < 2 count nothing) */
    long long n = 0;
    if (nots < 2 || tr->n_points < 2) return 0;
    for (int i = 0; i < n_pix; i++) {
        double tmin = t0[i] - dt * 0.5;
        double tmax = t0[i] + dt * ((double)nots - 0.5);
        for (int p = 0; p < tr->n_particles; p++)
            for (int step = 0; step < tr->n_points - 1; step++) {
                double ts1 = or_retarded_time(tr, p, step, obs + 3 * (size_t)i);
                double ts2 = or_retarded_time(tr, p, step + 1, obs + 3 * (size_t)i);
                if ((ts2 > tmin) && (ts1 < tmax)) n++;
            }
    }
    return n;
}

void or_screen_cells(const double origin[3], const double dX[3], const double dY[3],
                     int nx, int ny, double t0_obs, double dtx, double dty,
                     double* obs, double* t0)
{
    /* screen.cpp:98-110 (integer division for the centre index) */
    vec3 O = v_make(origin[0], origin[1], origin[2]);
    vec3 DX = v_make(dX[0], dX[1], dX[2]);
    vec3 DY = v_make(dY[0], dY[1], dY[2]);
    for (int ix = 0; ix < nx; ix++)
        for (int iy = 0; iy < ny; iy++) {
            size_t i = (size_t)ix * (size_t)ny + (size_t)iy;
            vec3 P = v_add(v_add(O, v_mul(DX, (double)(ix - nx / 2))), v_mul(DY, (double)(iy - ny / 2)));
            obs[3 * i + 0] = P.x; obs[3 * i + 1] = P.y; obs[3 * i + 2] = P.z;
            t0[i] = t0_obs + dtx * (double)(ix - nx / 2) + dty * (double)(iy - ny / 2);
        }
}

void or_trace_poynting(const double* trace, int n_bins, double dt, double S[3])
{
    /* fieldtrace.cpp:192-198 with ElMagField::Poynting fields.cpp:48 */
    vec3 sum = v_make(0.0, 0.0, 0.0);
    for (int i = 0; i < n_bins; i++) {
        const double* f = trace + 6 * (size_t)i;
        vec3 s = v_div(v_cross(v_make(f[0], f[1], f[2]), v_make(f[3], f[4], f[5])), OR_MU0);
        sum = v_add(sum, s);
    }
    sum = v_mul(sum, dt);
    S[0] = sum.x; S[1] = sum.y; S[2] = sum.z;
}

double or_total_energy(int n_pix, const double* traces, int n_bins, double dt,
                       const double* area, const double* normal)
{
    /* mesh.cpp:744-754 */
    double total = 0.0;
    for (int ip = 0; ip < n_pix; ip++) {
        double S[3];
        or_trace_poynting(traces + (size_t)ip * (size_t)n_bins * 6, n_bins, dt, S);
        total -= area[ip] * v_dot(v_make(S[0], S[1], S[2]), v_make(normal[3 * ip], normal[3 * ip + 1], normal[3 * ip + 2]));
    }
    return total;
}

/* ---- tracking: planar undulator + Vay pusher ---------------------------- */

void or_undulator_field(double bpeak, double lambda_u, int n_periods,
                        const double origin[3], const double Xg[3], double Bout[3])
{
    /* undulator.cpp:66-80 (ky = kz = 2 pi / lambda) and 107-126 */
    double kz = 2.0 * OR_PI / lambda_u;
    vec3 X = v_sub(v_make(Xg[0], Xg[1], Xg[2]), v_make(origin[0], origin[1], origin[2]));
    vec3 B = v_make(0.0, 0.0, 0.0);
    double z1 = -(n_periods + 1) * lambda_u / 2.0;
    double z2 = (n_periods + 1) * lambda_u / 2.0;
    if (z1 <= X.z && X.z < z2) {
        B.x = 0.0;
        B.y = bpeak * sin(kz * X.z) * cosh(kz * X.y);
        B.z = bpeak * cos(kz * X.z) * sinh(kz * X.y);
        if (X.z < z1 + lambda_u) B = v_mul(B, (X.z - z1) / lambda_u);
        if (z2 - lambda_u <= X.z) B = v_mul(B, (z2 - X.z) / lambda_u);
    }
    Bout[0] = B.x; Bout[1] = B.y; Bout[2] = B.z;
}

typedef struct {
    double qm, qmt2, dt;
    double t;
    vec3 X, P, A;
    vec3 p_i1;
    double gamma_i1;
} vay_state;

static void vay_second_half(vay_state* s, vec3 E, vec3 B)
{
    /* particle.cpp:428-438 == 500-510 */
    vec3 p_prime = v_add(s->P, v_mul(v_div(E, OR_SPEED_OF_LIGHT), s->qmt2));
    double gamma2_prime = v_abs2nd(p_prime) + 1.0;
    vec3 tau = v_mul(B, s->qmt2);
    double u_star = v_dot(p_prime, tau);
    double tau2 = v_abs2nd(tau);
    double sigma = gamma2_prime - tau2;
    s->gamma_i1 = sqrt(0.5 * (sigma + sqrt(sigma * sigma + 4.0 * (tau2 + u_star * u_star))));
    vec3 tt = v_div(tau, s->gamma_i1);
    s->p_i1 = v_div(v_add(v_add(p_prime, v_mul(tt, v_dot(p_prime, tt))), v_cross(p_prime, tt)), 1 + v_abs2nd(tt));
}

/* Lattice::Field (fields.cpp:229-239): the elements' fields summed in order,
 * the undulator (has_und) and then a HomogeneousField E0, B0 (has_hom,
 * fields.cpp:155-163).  Without the homogeneous element the undulator's B is
 * taken as it is and E is zero. */
static void lattice_field(int has_und, double bpeak, double lambda_u, int n_periods, const double origin[3],
                          int has_hom, const double E0[3], const double B0[3], vec3 X, vec3* E, vec3* B)
{
    double Xa[3] = {X.x, X.y, X.z}, Ba[3] = {0.0, 0.0, 0.0};
    if (has_und) or_undulator_field(bpeak, lambda_u, n_periods, origin, Xa, Ba);
    *E = v_make(0.0, 0.0, 0.0);
    *B = v_make(Ba[0], Ba[1], Ba[2]);
    if (has_hom) {
        *E = v_add(*E, v_make(E0[0], E0[1], E0[2]));
        *B = v_add(*B, v_make(B0[0], B0[1], B0[2]));
    }
}

/* InitVay + StepVay (particle.cpp:407-442, 479-512) through the lattice of
 * lattice_field; n_steps+1 points per particle, SoA particle-major. */
void or_track_lattice(int n_particles, const double* x0, const double* u0, double t_start,
                      double charge, double mass, double dt, int n_steps,
                      int has_und, double bpeak, double lambda_u, int n_periods, const double origin[3],
                      int has_hom, const double E0[3], const double B0[3],
                      double* t, double* x, double* y, double* z,
                      double* ux, double* uy, double* uz,
                      double* ax, double* ay, double* az, int n_threads)
{
    size_t NP = (size_t)n_steps + 1;
#ifdef _OPENMP
    if (n_threads <= 0) n_threads = omp_get_max_threads();
#pragma omp parallel for schedule(static) num_threads(n_threads)
#endif
    for (int p = 0; p < n_particles; p++) {
        size_t g = (size_t)p * NP;
        vay_state s;
        /* initTrajectory particle.cpp:326-341 */
        s.t = t_start;
        s.X = v_make(x0[3 * p], x0[3 * p + 1], x0[3 * p + 2]);
        s.P = v_make(u0[3 * p], u0[3 * p + 1], u0[3 * p + 2]);
        /* InitVay particle.cpp:407-442 */
        s.dt = dt;
        s.qm = charge / mass * OR_INV_REST_MASS;
        s.qmt2 = s.qm * 0.5 * dt;
        vec3 E, B;
        lattice_field(has_und, bpeak, lambda_u, n_periods, origin, has_hom, E0, B0, s.X, &E, &B);
        vay_second_half(&s, E, B);
        double gamma_h = sqrt(v_abs2nd(s.P) + 1.0);
        vec3 beta_h = v_div(s.P, gamma_h);
        s.A = v_mul(v_add(v_cross(beta_h, B), v_div(E, OR_SPEED_OF_LIGHT)), s.qm);
        t[g] = s.t; x[g] = s.X.x; y[g] = s.X.y; z[g] = s.X.z;
        ux[g] = s.P.x; uy[g] = s.P.y; uz[g] = s.P.z;
        ax[g] = s.A.x; ay[g] = s.A.y; az[g] = s.A.z;
        for (int n = 1; n <= n_steps; n++) {
            /* StepVay particle.cpp:479-512 */
            vec3 p_i = s.p_i1;
            vec3 beta_i = v_div(p_i, s.gamma_i1);
            s.t += s.dt;
            s.X = v_add(s.X, v_mul(v_mul(beta_i, OR_SPEED_OF_LIGHT), s.dt));
            lattice_field(has_und, bpeak, lambda_u, n_periods, origin, has_hom, E0, B0, s.X, &E, &B);
            s.P = v_add(s.p_i1, v_mul(v_add(v_div(E, OR_SPEED_OF_LIGHT), v_cross(beta_i, B)), s.qmt2));
            gamma_h = sqrt(v_abs2nd(s.P) + 1.0);
            beta_h = v_div(s.P, gamma_h);
            s.A = v_mul(v_add(v_cross(beta_h, B), v_div(E, OR_SPEED_OF_LIGHT)), s.qm);
            vay_second_half(&s, E, B);
            size_t q = g + (size_t)n;
            t[q] = s.t; x[q] = s.X.x; y[q] = s.X.y; z[q] = s.X.z;
            ux[q] = s.P.x; uy[q] = s.P.y; uz[q] = s.P.z;
            ax[q] = s.A.x; ay[q] = s.A.y; az[q] = s.A.z;
        }
    }
    (void)n_threads;
}

void or_track_undulator(int n_particles, const double* x0, const double* u0, double t_start,
                        double charge, double mass, double dt, int n_steps,
                        double bpeak, double lambda_u, int n_periods, const double origin[3],
                        double* t, double* x, double* y, double* z,
                        double* ux, double* uy, double* uz,
                        double* ax, double* ay, double* az, int n_threads)
{
    const double zero[3] = {0.0, 0.0, 0.0};
    or_track_lattice(n_particles, x0, u0, t_start, charge, mass, dt, n_steps, 1, bpeak, lambda_u, n_periods,
                     origin, 0, zero, zero, t, x, y, z, ux, uy, uz, ax, ay, az, n_threads);
}
