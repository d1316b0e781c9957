// vay_kernel.hip - trajectory generation on the device: the Vay pusher
// (J.-L. Vay, Phys. Plasmas 15, 056701 (2008)) through a lattice of a planar
// undulator and a homogeneous field, one lane per particle.  Restates
// ChargedParticle::InitVay / StepVay (src/particle.cpp:407-442, 479-512),
// PlanarUndulator::LocalField (src/undulator.cpp:107-126), HomogeneousField
// (src/fields.cpp:155-163) and Lattice::Field (src/fields.cpp:229-239) in
// the reference's operation order (FMA contraction off).  The undulator's
// sin/cos/cosh/sinh come from the device math library, so trajectories
// through it agree with the host to rounding; in a homogeneous field alone
// every operation is correctly rounded and they agree bitwise.
#include "tg_device.h"

namespace tg {

struct V3 {
    double x, y, z;
};
// Vector arithmetic of src/vector.cpp, one rounding per operation: the
// pragma keeps the compiler from fusing a product into a following sum once
// these are inlined (the reference is built without FMA).
__device__ __forceinline__ V3 v3(double x, double y, double z) { return V3{x, y, z}; }
__device__ __forceinline__ V3 operator+(V3 a, V3 b)
{
#pragma clang fp contract(off)
    return v3(a.x + b.x, a.y + b.y, a.z + b.z);
}
__device__ __forceinline__ V3 operator-(V3 a, V3 b)
{
#pragma clang fp contract(off)
    return v3(a.x - b.x, a.y - b.y, a.z - b.z);
}
__device__ __forceinline__ V3 operator*(V3 a, double f)
{
#pragma clang fp contract(off)
    return v3(a.x * f, a.y * f, a.z * f);
}
__device__ __forceinline__ V3 operator/(V3 a, double f)
{
#pragma clang fp contract(off)
    return v3(a.x / f, a.y / f, a.z / f);
}
__device__ __forceinline__ double abs2nd(V3 a)
{
#pragma clang fp contract(off)
    return a.x * a.x + a.y * a.y + a.z * a.z;
}
__device__ __forceinline__ double dot(V3 a, V3 b)
{
#pragma clang fp contract(off)
    double t = a.x * b.x;
    t += a.y * b.y;
    t += a.z * b.z;
    return t;
}
__device__ __forceinline__ V3 cross(V3 a, V3 b)
{
#pragma clang fp contract(off)
    return v3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}

struct Undulator {
    double bpeak, lambda_u, kz, z1, z2;
    V3 origin;
    __device__ V3 field(V3 Xg) const
    {
#pragma clang fp contract(off)
        V3 X = Xg - origin;
        V3 B = v3(0.0, 0.0, 0.0);
        if (z1 <= X.z && X.z < z2) {
            B.x = 0.0;
            B.y = bpeak * sin(kz * X.z) * cosh(kz * X.y);
            B.z = bpeak * cos(kz * X.z) * sinh(kz * X.y);
            if (X.z < z1 + lambda_u) B = B * ((X.z - z1) / lambda_u);
            if (z2 - lambda_u <= X.z) B = B * ((z2 - X.z) / lambda_u);
        }
        return B;
    }
};

// The lattice's field at X (Lattice::Field, src/fields.cpp:229-239): the
// elements' fields summed in order, the undulator's (when present) and then
// the homogeneous E0, B0 (when present).  Without the homogeneous element the
// undulator's B is taken as it is and E is zero.
struct LatticeField {
    Undulator und;
    V3 E0, B0;
    int has_und, has_hom;
    __device__ void field(V3 X, V3& E, V3& B) const
    {
#pragma clang fp contract(off)
        E = v3(0.0, 0.0, 0.0);
        B = has_und ? und.field(X) : v3(0.0, 0.0, 0.0);
        if (has_hom) {
            E = E + E0;
            B = B + B0;
        }
    }
};

// Second half step of particle.cpp:428-438 / 500-510.
__device__ __forceinline__ void vay_second_half(V3 P, V3 E, V3 B, double qmt2, V3& p_i1, double& gamma_i1)
{
#pragma clang fp contract(off)
    V3 p_prime = P + (E / kSpeedOfLight) * qmt2;
    double gamma2_prime = abs2nd(p_prime) + 1.0;
    V3 tau = B * qmt2;
    double u_star = dot(p_prime, tau);
    double tau2 = abs2nd(tau);
    double sigma = gamma2_prime - tau2;
    gamma_i1 = sqrt(0.5 * (sigma + sqrt(sigma * sigma + 4.0 * (tau2 + u_star * u_star))));
    V3 tt = tau / gamma_i1;
    p_i1 = ((p_prime + tt * dot(p_prime, tt)) + cross(p_prime, tt)) / (1 + abs2nd(tt));
}

__global__ void vay_lattice_kernel(int n, const double* __restrict__ x0, const double* __restrict__ u0,
                                   double t_start, double qm, double qmt2, double dt, int n_steps,
                                   LatticeField lat, double* __restrict__ t, double* __restrict__ x,
                                     double* __restrict__ y, double* __restrict__ z,
                                     double* __restrict__ ux, double* __restrict__ uy,
                                     double* __restrict__ uz, double* __restrict__ ax,
                                     double* __restrict__ ay, double* __restrict__ az)
{
#pragma clang fp contract(off)
    int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    const size_t NP = (size_t)n_steps + 1;
    size_t g = (size_t)p * NP;
    double tc = t_start;
    V3 X = v3(x0[3 * p], x0[3 * p + 1], x0[3 * p + 2]);
    V3 P = v3(u0[3 * p], u0[3 * p + 1], u0[3 * p + 2]);
    V3 p_i1;
    double gamma_i1;
    // InitVay
    V3 E, B;
    lat.field(X, E, B);
    vay_second_half(P, E, B, qmt2, p_i1, gamma_i1);
    double gamma_h = sqrt(abs2nd(P) + 1.0);
    V3 beta_h = P / gamma_h;
    V3 A = (cross(beta_h, B) + E / kSpeedOfLight) * qm;
    t[g] = tc; x[g] = X.x; y[g] = X.y; z[g] = X.z;
    ux[g] = P.x; uy[g] = P.y; uz[g] = P.z;
    ax[g] = A.x; ay[g] = A.y; az[g] = A.z;
    for (int s = 1; s <= n_steps; ++s) {
        // StepVay
        V3 beta_i = p_i1 / gamma_i1;
        tc += dt;
        X = X + (beta_i * kSpeedOfLight) * dt;
        lat.field(X, E, B);
        P = p_i1 + (E / kSpeedOfLight + cross(beta_i, B)) * qmt2;
        gamma_h = sqrt(abs2nd(P) + 1.0);
        beta_h = P / gamma_h;
        A = (cross(beta_h, B) + E / kSpeedOfLight) * qm;
        vay_second_half(P, E, B, qmt2, p_i1, gamma_i1);
        size_t q = g + (size_t)s;
        t[q] = tc; x[q] = X.x; y[q] = X.y; z[q] = X.z;
        ux[q] = P.x; uy[q] = P.y; uz[q] = P.z;
        ax[q] = A.x; ay[q] = A.y; az[q] = A.z;
    }
}

}  // namespace tg
