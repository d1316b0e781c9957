// tg_device.h - device-side data layout and arithmetic of the gfx950
// radiation kernels.
//
// Retarded time and bucket index are computed bit-for-bit as the reference
// computes them on the host (src/particle.cpp:560-568, 805, 829-830): one ulp
// of t_ret (~7e-24 s at 33 ns) moves ~1e-10 of a bucket's content to its
// neighbour, so FP64 parity below 1e-12 needs identical doubles there.  That
// chain therefore uses only IEEE-rounded operations in the reference's order:
// FMA contraction off, correctly rounded sqrt, and divisions either as IEEE
// division or as Markstein's division by a constant (q = a*y, r = fma(-q,b,a),
// q' = fma(r,y,q) with y = RN(1/b); P. Markstein, IBM J. Res. Dev. 34 (1990)
// 111-119), which returns RN(a/b).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace tg {

// src/global.h:37-44
constexpr double kElementaryCharge = 1.6021766208e-19;
constexpr double kSpeedOfLight = 2.99792458e8;
constexpr double kPi = 3.1415926535897932384626433832795029;
constexpr double kEpsNull = 8.854187817e-12;
constexpr double kMuNull = 4 * kPi * 1e-7;
constexpr double kInvRestMass = 1.758820025e11;
constexpr double kInvSpeedOfLight = 1.0 / kSpeedOfLight;  // RN(1/c), folded by the compiler

constexpr int kLanes = 64;  // one wavefront = one pixel tile

// Packed trajectory point for the FP64 kernel: 11 doubles in 96 B, read per
// lane (vector loads) by the per-lane walk, or staged through LDS for the
// wave-uniform walk.  beta = u/gamma and beta' = a/gamma are hoisted out of
// the pixel loop, computed exactly as src/particle.cpp:575-582 computes them;
// the charge scale q e/(4 pi eps0) of src/particle.cpp:588 is folded into
// the two amplitude factors of the field (eval_f64).
struct alignas(32) RecF64 {
    double t, x, y, z;
    double bx, by, bz;
    double apx, apy, apz;   // beta' * scale / c
    double sv;              // scale / gamma^2
    double pad;
};

// Packed trajectory point for the mixed kernel: 64 B, read per lane with
// vector loads in the per-lane walk, one scalar load in the wave-uniform walk.  FP64
// where the retarded time and the dominant term of 1 - n.beta need it (t, X,
// beta_z); FP32 for the transverse beta, beta' pre-multiplied by the charge
// scale / c, and the velocity-term scale / gamma^2.  The FP32 transverse
// beta assumes the beam travels close to the z axis (|beta_x|+|beta_y| small),
// as in TEUFEL's undulator lattices; the pack kernel flags other beams.
struct alignas(64) RecMixed {
    double t, x, y, z;
    double bz;
    float bx, by;
    float apx, apy, apz;   // beta' * scale / c
    float sv;              // scale / gamma^2
};
static_assert(sizeof(RecF64) == 96, "RecF64 layout");
static_assert(sizeof(RecMixed) == 64, "RecMixed layout");

// Layout of the packed records in HBM: piece-major per particle.  Piece i
// (bytes [16 i, 16 i + 16) of the record) of all points of a particle lies
// together, 16 B per point, so piece i of point k sits at
// i * rec_piece_stride(NP) + 16 k of the particle's records.  The per-lane
// walk reads one piece of a different point in every lane: lanes at nearby
// points share cache lines (8 points per 128-B line, against 2 for an array
// of mixed records and 1.3 of FP64 ones), and the address of a point is one
// shift, with the piece offsets wave-uniform (the buffer loads' scalar
// offset).  Each piece holds NP + 2 points rounded up to 16 (the walk reads
// up to two points ahead; the padding is zero), so pieces start on 256-B
// boundaries.
template <class Rec>
constexpr int kRecPieces = (int)(sizeof(Rec) / 16);
__host__ __device__ constexpr unsigned rec_piece_stride(int NP)
{
    return (unsigned)((NP + 2 + 15) / 16) * 256u;
}
template <class Rec>
__host__ __device__ constexpr size_t rec_particle_bytes(int NP)
{
    return (size_t)kRecPieces<Rec> * rec_piece_stride(NP);
}
// Byte offset of piece i of point k in its particle's records.
__host__ __device__ constexpr unsigned rec_piece_offset(int k, int i, unsigned stride)
{
    return (unsigned)i * stride + 16u * (unsigned)k;
}

// One screen as the kernel sees it.  Particle chunk 0 of a launch adds its
// fields straight into the trace store; chunks 1.. write partial traces that
// a second kernel adds, in chunk order, after the launch.
struct ScreenDev {
    const double* obs;    // [n_pix][3]
    const double* t0;     // [n_pix]
    const int* perm;      // [n_tiles*64] pixel index per lane, -1 = padding
    double* traces;       // [n_pix][n_bins][6], the trace store (chunk 0)
    double* partial;      // [n_chunks - 1][n_pix][n_bins][6] (chunks 1..)
    double dt;
    double inv_dt;        // RN(1/dt)
    int n_pix;
    int n_bins;
    int n_tiles;
    int window;           // trace bins per LDS window
    int spill;            // per-lane walk: windows hold a spill bucket (run_item)
};

// One wavefront's work: a pixel tile of one screen and a chunk of particles.
struct WorkItem {
    int screen;
    int tile;
    int chunk;
    int p0, p1;           // particle range
    int bin_lo, bin_hi;   // bucket range [bin_lo, bin_hi) of this item (a group of windows)
};

// -------- IEEE-exact helpers (no contraction) ------------------------------

// RN(a/b) for b with y = RN(1/b) precomputed (Markstein).
__device__ __forceinline__ double div_by_const(double a, double b, double y)
{
    double q = a * y;
    double r = __builtin_fma(-q, b, a);
    return __builtin_fma(r, y, q);
}

// RN(1/b) from the hardware estimate and two Newton steps with FMA: the last
// step rounds correctly (Markstein) for the normal, finite b used here.
__device__ __forceinline__ double rcp_rn(double b)
{
    double y = __builtin_amdgcn_rcp(b);
    double e = __builtin_fma(-b, y, 1.0);
    y = __builtin_fma(y, e, y);
    e = __builtin_fma(-b, y, 1.0);
    return __builtin_fma(y, e, y);
}

// 1/b to ~2^-46 relative (the hardware estimate and one Newton step with
// FMA), for reciprocals that scale amplitudes only (the deposit weights, the
// 1/(R kappa^3) of the field), where the trace tolerance (1e-12 of each
// pixel's largest field) leaves ~1e4 of room.  Timing and n = R/|R|, whose
// rounding must match the reference's, use rcp_rn.
__device__ __forceinline__ double rcp_amp(double b)
{
    const double y = __builtin_amdgcn_rcp(b);
    return __builtin_fma(y, __builtin_fma(-b, y, 1.0), y);
}

// RN(sqrt(x)) for normal x > 0 far from the denormal range: Goldschmidt
// iteration from the hardware 1/sqrt estimate, then two FMA residual
// corrections; the last one rounds correctly (the sequence compilers use
// for IEEE sqrt, without their range scaling).
__device__ __forceinline__ double sqrt_rn(double x)
{
    double y = __builtin_amdgcn_rsq(x);
    double g = x * y, h = 0.5 * y;
    double r = __builtin_fma(-h, g, 0.5);
    h = __builtin_fma(h, r, h);
    g = __builtin_fma(g, r, g);
    double d = __builtin_fma(-g, g, x);
    g = __builtin_fma(d, h, g);
    d = __builtin_fma(-g, g, x);
    return __builtin_fma(d, h, g);
}

// v_cvt_i32_f64: saturates out-of-range values (infinities included) to the
// int range, NaN to 0 (CDNA ISA, float-to-int conversions).  Written as the
// instruction because a C++ conversion of an out-of-range value is undefined.
__device__ __forceinline__ int cvt_sat_i32(double r)
{
    int i;
    asm("v_cvt_i32_f64 %0, %1" : "=v"(i) : "v"(r));
    return i;
}

// round() of the reference (half away from zero), clamped to +-1e9 (one
// v_med3_i32 after the saturating conversion); the clamp only touches values
// far outside every trace, where the reference never evaluates the index.
__device__ __forceinline__ int round_index(double q)
{
    const int i = cvt_sat_i32(__builtin_round(q));
    return min(max(i, -1000000000), 1000000000);
}

// Nearest integer (half to even) of q, clamped into int range: the mixed
// kernel's bucket index.
__device__ __forceinline__ int rint_index(double q)
{
    double r = __builtin_rint(q);
    r = __builtin_fmin(__builtin_fmax(r, -1.0e9), 1.0e9);
    return (int)r;
}

}  // namespace tg
