// lw_kernels.hip - gfx950 kernels of the post-tracking radiation step.
//
// Reference path: ChargedParticle::integrateFieldTrace (src/particle.cpp:771-867)
// called per pixel by the OpenMP observer loops (src/mesh.cpp:288-342,
// src/screen.cpp:112-180) and per particle by Bunch/Beam
// (src/bunch.cpp:739-745, src/beam.cpp:443-450).
//
// Decomposition.  One wavefront = 64 pixels (one per lane, a "tile") x one
// chunk of particles x a group of bucket windows (a work item); persistent
// waves take items from a work list.  The tile's traces are produced window
// by window: a window of W bins x 6 components x 64 lanes lives in LDS (FP64
// cells for the FP64 kernel, FP32 cells flushed into FP64 for the mixed one).
// For every particle of the chunk each lane walks its own pixel's trajectory
// steps whose retarded-time interval touches the window, evaluating each
// trajectory point once (the reference evaluates each point twice) and
// depositing with the reference's 3-case bucket rule.  A lane only ever
// writes its own pixel's LDS cells, in particle order and step order, so the
// deposit needs no atomics and is bitwise reproducible.  Per particle a
// cursor (the wave's earliest stopping step, plus one-byte lane offsets from
// it) carries each lane's position into the next window, so every step is
// walked about once per work item.  Records are read per lane one step ahead
// of use.  Particle chunks write partial traces that a second kernel sums in
// fixed chunk order into the screen's trace store.  The wave-uniform walk
// (all lanes on one step, records read with scalar loads or staged through
// LDS) is kept as an alternative.
#include "tg_device.h"
#include "tg_deposit.h"

#include <climits>

namespace tg {

// ---------------------------------------------------------------------------
// Packing: raw ChargedParticle trajectory (t, X, u=beta*gamma, a) -> records.
// ---------------------------------------------------------------------------

// Trajectory checks of the pack kernels, bits of their flags word.
constexpr int kFlagOffAxis = 1;       // |beta_x| + |beta_y| > 0.25 (the mixed record)
constexpr int kFlagSuperluminal = 2;  // a step with |dX| > c dt or dt < 0

// Flags step (point g -> g + 1) of a particle if it is faster than light or
// goes back in time.  For any particle slower than light the retarded time
// t + |P - X|/c increases along the trajectory, seen from every point P (its
// increment is at least dt - |dX|/c); the walks rely on that.  Repeated
// points (dt = 0, dX = 0) are allowed: they make zero-length steps.
__device__ __forceinline__ void check_step(size_t g, const double* __restrict__ t, const double* __restrict__ x,
                                           const double* __restrict__ y, const double* __restrict__ z,
                                           int* __restrict__ flags)
{
    const double dT = t[g + 1] - t[g];
    const double ex = x[g + 1] - x[g], ey = y[g + 1] - y[g], ez = z[g + 1] - z[g];
    const double cdt = kSpeedOfLight * dT;
    if (dT < 0.0 || ex * ex + ey * ey + ez * ez > cdt * cdt) atomicOr(flags, kFlagSuperluminal);
}

// Record of point k of particle p into the piece-major layout (rec_piece_offset).
template <class Rec>
__device__ __forceinline__ void store_rec(char* out, int n_points, size_t p, size_t k, const Rec& r)
{
    char* dst = out + p * rec_particle_bytes<Rec>(n_points);
    const unsigned stride = rec_piece_stride(n_points);
    const uint4* src = reinterpret_cast<const uint4*>(&r);
#pragma unroll
    for (int i = 0; i < kRecPieces<Rec>; ++i)
        *reinterpret_cast<uint4*>(dst + rec_piece_offset((int)k, i, stride)) = src[i];
}

__global__ void pack_f64_kernel(size_t n, int n_points, const double* __restrict__ t,
                                const double* __restrict__ x, const double* __restrict__ y,
                                const double* __restrict__ z, const double* __restrict__ ux,
                                const double* __restrict__ uy, const double* __restrict__ uz,
                                const double* __restrict__ ax, const double* __restrict__ ay,
                                const double* __restrict__ az, const double* __restrict__ charge,
                                char* __restrict__ out, int* __restrict__ flags)
{
#pragma clang fp contract(off)
    size_t g = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= n) return;
    size_t p = g / (size_t)n_points;
    if (g - p * (size_t)n_points + 1 < (size_t)n_points) check_step(g, t, x, y, z, flags);
    // src/particle.cpp:575-582: gamma2 = |u|^2 + 1, beta = u/gamma, beta' = a/gamma
    double bg2 = ux[g] * ux[g] + uy[g] * uy[g] + uz[g] * uz[g];
    double gamma2 = bg2 + 1.0;
    double gamma = __builtin_sqrt(gamma2);
    const double scale = charge[p] * kElementaryCharge / (4.0 * kPi * kEpsNull);  // particle.cpp:588
    const double sa = scale / kSpeedOfLight;
    RecF64 r;
    r.t = t[g];
    r.x = x[g];
    r.y = y[g];
    r.z = z[g];
    r.bx = ux[g] / gamma;
    r.by = uy[g] / gamma;
    r.bz = uz[g] / gamma;
    r.apx = ax[g] / gamma * sa;
    r.apy = ay[g] / gamma * sa;
    r.apz = az[g] / gamma * sa;
    r.sv = scale / gamma2;
    r.pad = 0.0;
    store_rec(out, n_points, p, g - p * (size_t)n_points, r);
}

__global__ void pack_mixed_kernel(size_t n, int n_points, const double* __restrict__ t,
                                  const double* __restrict__ x, const double* __restrict__ y,
                                  const double* __restrict__ z, const double* __restrict__ ux,
                                  const double* __restrict__ uy, const double* __restrict__ uz,
                                  const double* __restrict__ ax, const double* __restrict__ ay,
                                  const double* __restrict__ az, const double* __restrict__ charge,
                                  char* __restrict__ out, int* __restrict__ flags)
{
#pragma clang fp contract(off)
    size_t g = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= n) return;
    size_t p = g / (size_t)n_points;
    if (g - p * (size_t)n_points + 1 < (size_t)n_points) check_step(g, t, x, y, z, flags);
    double bg2 = ux[g] * ux[g] + uy[g] * uy[g] + uz[g] * uz[g];
    double gamma2 = bg2 + 1.0;
    double gamma = __builtin_sqrt(gamma2);
    double scale = charge[p] * kElementaryCharge / (4.0 * kPi * kEpsNull);
    RecMixed r;
    r.t = t[g];
    r.x = x[g];
    r.y = y[g];
    r.z = z[g];
    r.bz = uz[g] / gamma;
    double bx = ux[g] / gamma, by = uy[g] / gamma;
    r.bx = (float)bx;
    r.by = (float)by;
    double sa = scale / kSpeedOfLight;
    r.apx = (float)(ax[g] / gamma * sa);
    r.apy = (float)(ay[g] / gamma * sa);
    r.apz = (float)(az[g] / gamma * sa);
    r.sv = (float)(scale / gamma2);
    store_rec(out, n_points, p, g - p * (size_t)n_points, r);
    // the FP32 transverse beta needs |beta_x| + |beta_y| << 1 (see RecMixed)
    if (__builtin_fabs(bx) + __builtin_fabs(by) > 0.25) atomicOr(flags, kFlagOffAxis);
}

// ---------------------------------------------------------------------------
// One trajectory point seen from one pixel.
// ---------------------------------------------------------------------------

template <class F>
struct PointT {
    double ts;     // retarded (arrival) time
    int idx;       // bucket index round((ts - t0)/dt)
    F f[6];        // Ex, Ey, Ez, Bx, By, Bz
};
using Point = PointT<double>;

// A trajectory point as the mixed kernel keeps it: the bucket position
// q = (ts - t0)/dt as its bucket index idx = rint(q) and the FP32 offset
// u = q - idx in [-1/2, 1/2] from that bucket's centre (the deposit needs
// only these, step_common_units), and the amplitudes E and c B in FP32 (the
// 1/c of B is applied when the LDS window is flushed into the FP64 trace).
struct PointM {
    int idx;
    float u;
    float f[6];    // Ex, Ey, Ez, c Bx, c By, c Bz
};

// Retarded time t + |P - X|/c and bucket index: src/particle.cpp:560-568,
// 805, 829-830, in the reference's operation order.
__device__ __forceinline__ void timing_exact(double t, double dx, double dy, double dz, double t0,
                                             double dt, double inv_dt, double& R, double& ts, int& idx)
{
#pragma clang fp contract(off)
    double r2 = dx * dx + dy * dy + dz * dz;
    R = sqrt_rn(r2);
    ts = t + div_by_const(R, kSpeedOfLight, kInvSpeedOfLight);
    idx = round_index(div_by_const(ts - t0, dt, inv_dt));
}

// FP64 Lienard-Wiechert field (src/particle.cpp:570-596).  n = RVec/R is
// formed bit-identically (Markstein with RN(1/R)) and 1 - n.beta in the
// reference's summation order, because that difference cancels ~4 digits.
// The acceleration term uses a x (b x c) = b (a.c) - c (a.b), so with
// ap = beta' scale/c and sv = scale/gamma^2 from the record
//   E = [(n - beta)(sv/R + n.ap) - ap n.(n - beta)] / (R kappa^3).
// n.(n - beta) is kept as computed rather than replaced by kappa: |n|^2
// differs from 1 by an ulp, which is ~1e-12 of kappa, and the reference's
// cross products see that n.  The point returns c B = n x E (the window's
// flush applies 1/c).
__device__ __forceinline__ void eval_f64(const RecF64& r, double Px, double Py, double Pz, double t0,
                                         double dt, double inv_dt, Point& o)
{
    double dx, dy, dz, R, nx, ny, nz, kappa, invR;
    {
#pragma clang fp contract(off)
        dx = Px - r.x;
        dy = Py - r.y;
        dz = Pz - r.z;
        timing_exact(r.t, dx, dy, dz, t0, dt, inv_dt, R, o.ts, o.idx);
        invR = rcp_rn(R);
        nx = div_by_const(dx, R, invR);
        ny = div_by_const(dy, R, invR);
        nz = div_by_const(dz, R, invR);
        double nb = r.bx * nx + r.by * ny + r.bz * nz;
        kappa = 1.0 - nb;
    }
    const double mx = nx - r.bx, my = ny - r.by, mz = nz - r.bz;   // n - beta
    const double k3 = kappa * kappa * kappa;
    const double irk3 = rcp_amp(R * k3);                          // 1/(R kappa^3)
    const double nap = nx * r.apx + ny * r.apy + nz * r.apz;      // n . beta' scale/c
    const double nm = nx * mx + ny * my + nz * mz;                // n . (n - beta)
    const double c1 = irk3 * __builtin_fma(r.sv, invR, nap);
    const double c2 = irk3 * nm;
    const double ex = mx * c1 - r.apx * c2;
    const double ey = my * c1 - r.apy * c2;
    const double ez = mz * c1 - r.apz * c2;
    o.f[0] = ex;
    o.f[1] = ey;
    o.f[2] = ez;
    o.f[3] = ny * ez - nz * ey;   // c B = n x E
    o.f[4] = nz * ex - nx * ez;
    o.f[5] = nx * ey - ny * ex;
}

// Retarded time and bucket position of the mixed kernel, also returning 1/R:
// 1/sqrt from the FP32 hardware estimate fy (~1 ulp of FP32, which the FP32
// amplitudes use as it is) and one FP64 Newton step, y (~2^-46 relative).
// Every FMA is explicit, so every caller gets the same doubles.  The position
// q = (ts - t0)/dt is returned as its nearest bucket idx (half to even, which
// differs from the reference's half-away rounding only for a retarded time
// exactly on a bucket edge; clamped to +-1e9) and the offset u = q - rint(q),
// in [-1/2, 1/2].  Beyond the clamp (|q| > 1e9 buckets, a point a billion
// buckets away from the trace) idx + u is not q.
__device__ __forceinline__ void timing_mixed(double t, double dx, double dy, double dz, double t0,
                                             double inv_dt, double& y, float& fy, int& idx, float& u)
{
#pragma clang fp contract(off)
    const double r2 = __builtin_fma(dx, dx, __builtin_fma(dy, dy, dz * dz));
    fy = __builtin_amdgcn_rsqf((float)r2);
    y = (double)fy;
    const double h = r2 * y;
    const double e = __builtin_fma(-h, y, 1.0);
    y = __builtin_fma(0.5 * y, e, y);
    const double R = r2 * y;
    // q = (t + R/c - t0)/dt as R (1/(c dt)) + (t/dt - t0/dt): both terms and
    // q keep ~1e-9 of a bucket (t/dt ~ 1e7)
    const double q = __builtin_fma(R, kInvSpeedOfLight * inv_dt, __builtin_fma(t, inv_dt, -(t0 * inv_dt)));
    const double r = __builtin_rint(q);
    idx = min(max(cvt_sat_i32(r), -1000000000), 1000000000);
    u = (float)(q - r);
}

// Mixed precision: FP64 distance, retarded time and the n_z beta_z part of
// 1 - n.beta (the cancelling term), FP32 for everything else.
__device__ __forceinline__ void eval_mixed(const RecMixed& r, double Px, double Py, double Pz,
                                           double t0, double dt, double inv_dt, PointM& o)
{
    const double dx = Px - r.x, dy = Py - r.y, dz = Pz - r.z;
    double y;
    float fy;
    timing_mixed(r.t, dx, dy, dz, t0, inv_dt, y, fy, o.idx, o.u);
    const double nz = dz * y;
    const float nx = (float)dx * fy, ny = (float)dy * fy, fnz = (float)nz;
    const float kappa = (float)__builtin_fma(-nz, r.bz, 1.0) - (nx * r.bx + ny * r.by);
    const float mx = nx - r.bx, my = ny - r.by, mz = (float)(nz - r.bz);
    const float irk3 = fy * __builtin_amdgcn_rcpf(kappa * kappa * kappa);   // 1/(R kappa^3)
    // scale/(R^2 kappa^3 gamma^2) + n.beta' scale/(R kappa^3 c)
    const float c1 = irk3 * __builtin_fmaf(r.sv, fy, nx * r.apx + ny * r.apy + fnz * r.apz);
    const float c2 = irk3 * kappa;
    const float ex = mx * c1 - r.apx * c2;
    const float ey = my * c1 - r.apy * c2;
    const float ez = mz * c1 - r.apz * c2;
    o.f[0] = ex;
    o.f[1] = ey;
    o.f[2] = ez;
    o.f[3] = ny * ez - fnz * ey;   // c B = n x E
    o.f[4] = fnz * ex - nx * ez;
    o.f[5] = nx * ey - ny * ex;
}

// Per precision: trajectory record, amplitude type (also the LDS window type)
// and evaluated point.  FP64 accumulates the window in FP64 LDS; the mixed
// kernel accumulates an FP32 LDS window that is flushed into the FP64 partial
// trace before its cells hold much more than 2048 terms each.
template <int PREC>
struct RecOf;
template <>
struct RecOf<0> { using type = RecF64; using amp = double; using point = Point; };
template <>
struct RecOf<1> { using type = RecMixed; using amp = float; using point = PointM; };

template <int PREC>
__device__ __forceinline__ void eval_point(const typename RecOf<PREC>::type& r, double Px, double Py,
                                           double Pz, double t0, double dt, double inv_dt,
                                           typename RecOf<PREC>::point& o)
{
    if constexpr (PREC == 0)
        eval_f64(r, Px, Py, Pz, t0, dt, inv_dt, o);
    else
        eval_mixed(r, Px, Py, Pz, t0, dt, inv_dt, o);
}

// Reciprocals of the deposit weights (amplitude-level, not timing).
struct RcpRn {
    __device__ __forceinline__ double operator()(double x) const { return rcp_rn(x); }
};
struct RcpAmp {
    __device__ __forceinline__ double operator()(double x) const { return rcp_amp(x); }
};
struct RcpF32 {
    __device__ __forceinline__ float operator()(float x) const { return __builtin_amdgcn_rcpf(x); }
};

// LDS trace window of one wave: cell (bucket j, component c, lane) at
// ((j*6)+c)*64 + lane, W buckets.  A step's absent targets (negative) mask
// their lanes off the adds, so no cell outside the window is touched.  Every
// cell has exactly one writer (its own lane) and a wave's LDS accesses
// execute in order, so plain read-add-write is race-free and deterministic.
template <class F>
struct Window {
    using LdsF = __attribute__((address_space(3))) F;
    LdsF* lane_base;  // lds + lane
    // j in [0, W): a 24-bit product (one full-rate v_mad_u32_u24 for the address)
    __device__ __forceinline__ LdsF* cell(int j) const { return lane_base + __mul24(j, 6 * kLanes); }
    // *p += v.  FP64 cells add in the LDS unit (ds_add_f64): each lane owns
    // its column, so no two lanes meet at one address and the adds of a lane
    // happen in program order; the VALU does not read the cell back.  FP32
    // cells are read, added and written by the VALU (ds_add_f32 measured 14
    // times slower than that on gfx950).
    static __device__ __forceinline__ void add(LdsF* p, F v)
    {
        if constexpr (sizeof(F) == 8)
            __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WAVEFRONT);
        else
            *p += v;
    }
    // bucket ja += aa f1 + ba f2 and bucket jb += ab f1 + bb f2 (ja != jb);
    // a negative (absent) target masks the lane off its adds.
    __device__ __forceinline__ void add2(int ja, F aa, F ba, int jb, F ab, F bb, const F* f1, const F* f2) const
    {
        if constexpr (sizeof(F) == 8) {
            if (ja >= 0) {
                LdsF* ca = cell(ja);
#pragma unroll
                for (int c = 0; c < 6; ++c) add(ca + c * kLanes, aa * f1[c] + ba * f2[c]);
            }
            if (jb >= 0) {
                LdsF* cb = cell(jb);
#pragma unroll
                for (int c = 0; c < 6; ++c) add(cb + c * kLanes, ab * f1[c] + bb * f2[c]);
            }
        } else {
            // cell + (a f1 + b f2): one rounding at the cell's magnitude per
            // step (folding the add into the FMAs would make it two, and
            // the window's FP32 sums measurably worse)
            if (ja >= 0) {
                LdsF* ca = cell(ja);
#pragma unroll
                for (int c = 0; c < 6; ++c) ca[c * kLanes] = ca[c * kLanes] + (aa * f1[c] + ba * f2[c]);
            }
            if (jb >= 0) {
                LdsF* cb = cell(jb);
#pragma unroll
                for (int c = 0; c < 6; ++c) cb[c * kLanes] = cb[c * kLanes] + (ab * f1[c] + bb * f2[c]);
            }
        }
    }
};

// Both windows hold twice the bucket sums (the common weights come doubled
// from step_common_units and step_common_x2), halved when they are flushed.
template <class F>
constexpr double kCellScale = 2.0;

// Accumulator for step_general: absolute bucket j -> window cell j - b0.
template <class F>
struct WindowAcc {
    Window<F> w;
    int b0;
    const F* f1;
    const F* f2;
    __device__ __forceinline__ void operator()(int j, double a, double b) const
    {
        auto* c0 = w.cell(j - b0);
        const F fa = (F)(kCellScale<F> * a), fb = (F)(kCellScale<F> * b);
#pragma unroll
        for (int c = 0; c < 6; ++c) Window<F>::add(c0 + c * kLanes, fa * f1[c] + fb * f2[c]);
    }
};

// Accumulator for step_general in the per-lane walk: bucket j -> window cell
// j - b0, and, when the window has a spill bucket (nw = b1 - b0 >= 0), bucket
// b1 -> the lane's spill registers sp (run_item).  nw < 0: no spill bucket.
template <class F>
struct WindowAccSpill {
    Window<F> w;
    int b0, nw;
    const F* f1;
    const F* f2;
    F* sp;
    __device__ __forceinline__ void operator()(int j, double a, double b) const
    {
        const F fa = (F)(kCellScale<F> * a), fb = (F)(kCellScale<F> * b);
        if (j - b0 == nw) {
#pragma unroll
            for (int c = 0; c < 6; ++c) sp[c] = sp[c] + (fa * f1[c] + fb * f2[c]);
        } else {
            auto* c0 = w.cell(j - b0);
#pragma unroll
            for (int c = 0; c < 6; ++c) Window<F>::add(c0 + c * kLanes, fa * f1[c] + fb * f2[c]);
        }
    }
};

// s_waitcnt vmcnt(0) with expcnt and lgkmcnt left open (gfx9 encoding:
// vmcnt[3:0] in bits 3:0 and [5:4] in bits 15:14, expcnt in 6:4, lgkmcnt in 11:8).
constexpr int kWaitVmcnt0 = 0x0F70;

// The records of one particle as a buffer resource (wave-uniform): the
// per-lane walk reads a record as whole 16-B buffer loads (4 for the mixed
// record, 6 for FP64) and takes the fields by bit casts.  The gathers of the
// per-lane walk are bound by the vector-memory address path, which handles
// every load instruction separately, and a field-typed copy of the record
// compiles to more, narrower loads.  The resource covers the particle's
// records (piece-major, rec_piece_offset): the point's offset 16 k is the
// per-lane offset, the piece's offset the wave-uniform scalar offset.  The
// walk's read-ahead near the end of a particle reads the zero padding of each
// piece (or, past it, zeros from the resource), which only stopped lanes
// evaluate.
typedef unsigned u32x4 __attribute__((ext_vector_type(4)));
typedef unsigned u32x2 __attribute__((ext_vector_type(2)));
constexpr int kBufferWord3 = 0x00020000;   // raw 32-bit data format, gfx9 buffer resource word 3

__device__ __forceinline__ double dbl(unsigned lo, unsigned hi)
{
    return __builtin_bit_cast(double, ((unsigned long long)hi << 32) | lo);
}
__device__ __forceinline__ float flt(unsigned v) { return __builtin_bit_cast(float, v); }

template <class Rec>
struct RecBuffer {
    __amdgpu_buffer_rsrc_t rsrc;
    int stride;   // bytes between the pieces of a point (wave-uniform)
    __device__ __forceinline__ RecBuffer(const char* rp, int NP)
        : rsrc(__builtin_amdgcn_make_buffer_rsrc(const_cast<char*>(rp), (short)0, (int)rec_particle_bytes<Rec>(NP),
                                                 kBufferWord3)),
          stride((int)rec_piece_stride(NP))
    {
    }
    __device__ __forceinline__ u32x4 part(int k, int i) const
    {
        return __builtin_amdgcn_raw_buffer_load_b128(rsrc, 16 * k, stride * i, 0);
    }
    __device__ __forceinline__ Rec load(int k) const;
};

template <>
__device__ __forceinline__ RecMixed RecBuffer<RecMixed>::load(int k) const
{
    const u32x4 a = part(k, 0), b = part(k, 1), c = part(k, 2), d = part(k, 3);
    RecMixed r;
    r.t = dbl(a.x, a.y);
    r.x = dbl(a.z, a.w);
    r.y = dbl(b.x, b.y);
    r.z = dbl(b.z, b.w);
    r.bz = dbl(c.x, c.y);
    r.bx = flt(c.z);
    r.by = flt(c.w);
    r.apx = flt(d.x);
    r.apy = flt(d.y);
    r.apz = flt(d.z);
    r.sv = flt(d.w);
    return r;
}
template <>
__device__ __forceinline__ RecF64 RecBuffer<RecF64>::load(int k) const
{
    const u32x4 a = part(k, 0), b = part(k, 1), c = part(k, 2), d = part(k, 3), e = part(k, 4);
    // the last piece holds sv and the padding: its first 8 B only
    const u32x2 f = __builtin_amdgcn_raw_buffer_load_b64(rsrc, 16 * k, stride * 5, 0);
    RecF64 r;
    r.t = dbl(a.x, a.y);
    r.x = dbl(a.z, a.w);
    r.y = dbl(b.x, b.y);
    r.z = dbl(b.z, b.w);
    r.bx = dbl(c.x, c.y);
    r.by = dbl(c.z, c.w);
    r.bz = dbl(d.x, d.y);
    r.apx = dbl(d.z, d.w);
    r.apy = dbl(e.x, e.y);
    r.apz = dbl(e.z, e.w);
    r.sv = dbl(f.x, f.y);
    r.pad = 0.0;
    return r;
}

// The minimum over the wave's active lanes (all lanes where it is called),
// by the device library's cross-lane reduction (DPP row operations and a
// readlane: no LDS round trips).
__device__ __forceinline__ int wave_min(int v) { return __ockl_wfred_min_i32(v); }

// Deposit of step (p1 -> p2) into the window [b0, b1) (the wave-uniform
// walk), FP64: retarded times and bucket edges as the reference forms them.
// A step that ends at or beyond b1 deposits only its buckets below b1 here;
// the next window walks it again for the rest.
__device__ __forceinline__ void deposit_window(const Window<double>& win, bool active, const Point& p1,
                                               const Point& p2, double t0, double dt, double inv_dt, double tmin,
                                               double tmax, int n_bins, int b0, int b1)
{
    const bool valid = active && p2.ts > tmin && p1.ts < tmax;    // particle.cpp:793
    const bool case_a = p1.ts <= tmin;
    const int i2 = p2.idx > n_bins ? n_bins : p2.idx;
    const bool gen = valid && step_is_general(case_a, p1.idx, i2);
    const TwoBuckets<double> t = step_common_x2<double>(RcpAmp{}, valid && !gen, p1.ts, p1.idx, p2.ts, i2, t0, dt,
                                                        inv_dt, b0, b1, -1, -1);
    win.add2(t.ja, t.aa, t.ba, t.jb, t.ab, t.bb, p1.f, p2.f);
    if (__ballot(gen)) {
        if (gen) {
            WindowAcc<double> acc{win, b0, p1.f, p2.f};
            step_general(acc, rcp_rn(p2.ts - p1.ts), case_a, p1.ts, p1.idx, p2.ts, i2, t0, dt, inv_dt, n_bins,
                         b0, b1);
        }
    }
}

// Mixed: the same rule in bucket units (deposit_step_units), the common
// step branch-free with FP32 weights, the general step under a ballot.
__device__ __forceinline__ void deposit_window(const Window<float>& win, bool active, const PointM& p1,
                                               const PointM& p2, double, double, double, double, double,
                                               int n_bins, int b0, int b1)
{
    const bool valid = active && p2.idx >= 0 && p1.idx < n_bins;  // particle.cpp:793
    const bool case_a = p1.idx < 0;                               // particle.cpp:796
    const int i2 = p2.idx > n_bins ? n_bins : p2.idx;
    const bool gen = valid && step_is_general(case_a, p1.idx, i2);
    const TwoBuckets<float> t = step_common_units<float>(RcpF32{}, valid && !gen, p1.idx, p1.u, p2.idx, p2.u, i2,
                                                         b0, b1, -1, -1);
    win.add2(t.ja, t.aa, t.ba, t.jb, t.ab, t.bb, p1.f, p2.f);
    if (__ballot(gen)) {
        if (gen) {
            WindowAcc<float> acc{win, b0, p1.f, p2.f};
            const double q1 = (double)p1.idx + (double)p1.u, q2 = (double)p2.idx + (double)p2.u;
            step_general(acc, rcp_rn(q2 - q1), case_a, q1, p1.idx, q2, i2, 0.0, 1.0, 1.0, n_bins, b0, b1);
        }
    }
}

// Window cell adder of the lane accumulator (lane_step, lane_close).
template <class F>
struct WindowAdd {
    Window<F> w;
    __device__ __forceinline__ void operator()(int j, int c, F v) const { Window<F>::add(w.cell(j) + c * kLanes, v); }
};

// The per-lane walk's deposit of step (p1 -> p2): the rule of
// deposit_window (same weights and general steps), with the common steps
// summed per bucket in the lane accumulator L (lane_step) instead of added to
// the window step by step.  The accumulator of a lane that does not walk the
// step takes garbage (it is closed when the lane stops, and the next walk
// starts it afresh), so its update needs no mask.  With spill, the window has
// a spill bucket b1 held in the lane's registers sp (run_item): a step that
// ends in bucket b1 (the lane's last step in the window) is deposited whole,
// its b1 part into sp, and is not walked again in the next window, which
// takes the spill bucket over as its first; a step that ends beyond b1
// deposits only its buckets below b1 here (the next window walks it again for
// the rest).  Walked steps start before b1, so a common step reaches at most
// b1, whose share stays in L until the lane stops.  FP64.
__device__ __forceinline__ void deposit_lane(const Window<double>& win, LaneAcc<double>& L, bool walking,
                                             const Point& p1, const Point& p2, double t0, double dt, double inv_dt,
                                             double tmin, double tmax, int n_bins, int b0, int b1, bool spill,
                                             double* sp)
{
    const bool valid = walking && p2.ts > tmin && p1.ts < tmax;    // particle.cpp:793
    const bool case_a = p1.ts <= tmin;
    const int i2 = p2.idx > n_bins ? n_bins : p2.idx;
    const bool gen = valid && step_is_general(case_a, p1.idx, i2);
    const bool common = valid && !gen;
    const TwoBuckets<double> t = step_common_x2<double>(RcpAmp{}, common, p1.ts, p1.idx, p2.ts, i2, t0, dt, inv_dt,
                                                        b0, b1, -1, -1);
    const int ja = p1.idx - b0;
    lane_step(L, WindowAdd<double>{win}, common, walking && (gen || i2 != p1.idx), t, ja,
              (unsigned)ja < (unsigned)(b1 - b0), p1.f, p2.f);
    if (__ballot(gen)) {
        if (gen) {
            WindowAccSpill<double> acc{win, b0, spill ? b1 - b0 : -1, p1.f, p2.f, sp};
            step_general(acc, rcp_rn(p2.ts - p1.ts), case_a, p1.ts, p1.idx, p2.ts, i2, t0, dt, inv_dt, n_bins,
                         b0, b1 + ((spill && i2 == b1) ? 1 : 0));
        }
    }
}

// Mixed: the same in bucket units (deposit_window for the mixed kernel).
__device__ __forceinline__ void deposit_lane(const Window<float>& win, LaneAcc<float>& L, bool walking,
                                             const PointM& p1, const PointM& p2, double, double, double, double,
                                             double, int n_bins, int b0, int b1, bool spill, float* sp)
{
    const bool valid = walking && p2.idx >= 0 && p1.idx < n_bins;  // particle.cpp:793
    const bool case_a = p1.idx < 0;                               // particle.cpp:796
    const int i2 = p2.idx > n_bins ? n_bins : p2.idx;
    const bool gen = valid && step_is_general(case_a, p1.idx, i2);
    const bool common = valid && !gen;
    const TwoBuckets<float> t = step_common_units<float>(RcpF32{}, common, p1.idx, p1.u, p2.idx, p2.u, i2, b0, b1,
                                                         -1, -1);
    const int ja = p1.idx - b0;
    lane_step(L, WindowAdd<float>{win}, common, walking && (gen || i2 != p1.idx), t, ja,
              (unsigned)ja < (unsigned)(b1 - b0), p1.f, p2.f);
    if (__ballot(gen)) {
        if (gen) {
            WindowAccSpill<float> acc{win, b0, spill ? b1 - b0 : -1, p1.f, p2.f, sp};
            const double q1 = (double)p1.idx + (double)p1.u, q2 = (double)p2.idx + (double)p2.u;
            step_general(acc, rcp_rn(q2 - q1), case_a, q1, p1.idx, q2, i2, 0.0, 1.0, 1.0, n_bins, b0,
                         b1 + ((spill && i2 == b1) ? 1 : 0));
        }
    }
}

// Two ways to feed the wave the (uniform) trajectory record of each step.
//
// LdsStager (FP64 kernel, 96-B records): the particle's records are staged
// through LDS in chunks of 16 (1536 B); each lane loads 24 B of a chunk with
// coalesced vector loads one chunk ahead of use, and the wave reads the
// current record from LDS.  The wait for a chunk falls once per 16 steps.
//
// ScalarPrefetch (mixed kernel, 64-B records): the record's four 16-B pieces
// (piece-major layout) read with scalar loads, requested one step ahead of use.
constexpr int kChunkRecs = 16;
constexpr int kChunkDoubles = kChunkRecs * 12;   // 192 doubles = 64 lanes x 3

template <class Rec>
struct LdsStager {
    const char* __restrict__ rp;      // the particle's records (piece-major)
    double* slots;                    // LDS: 2 chunks, each [piece][kChunkRecs points][2 doubles]
    int lane;
    int last_chunk;                   // last chunk of kChunkRecs points within each piece
    unsigned stride;                  // bytes between pieces
    int next_fill;                    // chunk held in (q0,q1,q2)
    double q0, q1, q2;

    static constexpr int kPieceDoubles = 2 * kChunkRecs;   // one piece of a chunk: 32 doubles
    // the three doubles lane `lane` loads of chunk `chunk`: element g of the
    // chunk's [piece][point][2] doubles, g = 3 lane + e
    __device__ __forceinline__ double elem(int chunk, int e) const
    {
        const int g = 3 * lane + e;
        const int piece = g / kPieceDoubles, w = g % kPieceDoubles;
        const double* base = reinterpret_cast<const double*>(rp + (size_t)piece * stride);
        return base[kPieceDoubles * (chunk < last_chunk ? chunk : last_chunk) + w];
    }
    __device__ __forceinline__ void load(int chunk, double& a, double& b, double& c) const
    {
        a = elem(chunk, 0);
        b = elem(chunk, 1);
        c = elem(chunk, 2);
    }
    __device__ __forceinline__ void store(int chunk, double a, double b, double c) const
    {
        double* d = slots + (chunk & 1) * kChunkDoubles + 3 * lane;
        d[0] = a;
        d[1] = b;
        d[2] = c;
    }
    // Stage the chunks holding records k and k+1 and request the next one.
    __device__ __forceinline__ void start(int k)
    {
        const int c0 = k / kChunkRecs;
        double a, b, c;
        load(c0, a, b, c);
        double d, e, f;
        load(c0 + 1, d, e, f);
        load(c0 + 2, q0, q1, q2);
        store(c0, a, b, c);
        store(c0 + 1, d, e, f);
        next_fill = c0 + 2;
    }
    // Record r of a staged chunk, gathered from its pieces.
    __device__ __forceinline__ Rec at(int r) const
    {
        static_assert(kPieces * kPieceDoubles == kChunkDoubles, "a chunk is the pieces of kChunkRecs points");
        const int j = r % kChunkRecs;
        const double* b = slots + ((r / kChunkRecs) & 1) * kChunkDoubles + 2 * j;
        Rec out;
        double* o = reinterpret_cast<double*>(&out);
#pragma unroll
        for (int i = 0; i < kPieces; ++i) {
            o[2 * i] = b[kPieceDoubles * i];
            o[2 * i + 1] = b[kPieceDoubles * i + 1];
        }
        return out;
    }
    static constexpr int kPieces = kRecPieces<Rec>;
    // Record r, r advancing by one per call after start.
    __device__ __forceinline__ Rec next(int r)
    {
        if (r / kChunkRecs == next_fill) {
            store(next_fill, q0, q1, q2);
            load(next_fill + 1, q0, q1, q2);
            ++next_fill;
        }
        return at(r);
    }
    __device__ __forceinline__ Rec first(int k) const { return at(k); }
};

template <class Rec>
struct ScalarPrefetch {
    const char* __restrict__ rp;   // the particle's records (piece-major)
    int NP;
    Rec rn;   // record of the next point
    // record k (uniform), gathered from its pieces
    __device__ __forceinline__ Rec at(int k) const
    {
        Rec r;
        uint4* o = reinterpret_cast<uint4*>(&r);
        const unsigned stride = rec_piece_stride(NP);
#pragma unroll
        for (int i = 0; i < kRecPieces<Rec>; ++i)
            o[i] = *reinterpret_cast<const uint4*>(rp + rec_piece_offset(k, i, stride));
        return r;
    }
    __device__ __forceinline__ void start(int k) { rn = at(k + 1); }
    __device__ __forceinline__ Rec first(int k) const { return at(k); }
    __device__ __forceinline__ Rec next(int r)
    {
        const Rec cur = rn;
        rn = at(r + 1 < NP ? r + 1 : NP - 1);
        return cur;
    }
};

template <int PREC>
struct SourceOf;
template <>
struct SourceOf<0> { using type = LdsStager<RecF64>; static constexpr int lds_doubles = 2 * kChunkDoubles; };
template <>
struct SourceOf<1> { using type = ScalarPrefetch<RecMixed>; static constexpr int lds_doubles = 0; };

template <int PREC>
__device__ __forceinline__ typename SourceOf<PREC>::type make_source(const char* rp, int NP, double* stage, int lane)
{
    if constexpr (PREC == 0)
        return LdsStager<RecF64>{rp, stage, lane, (int)(rec_piece_stride(NP) / (16 * kChunkRecs)) - 1,
                                 rec_piece_stride(NP), 0, 0.0, 0.0, 0.0};
    else
        return ScalarPrefetch<RecMixed>{rp, NP, RecMixed{}};
}

// Walk the steps of one particle starting at step k (uniform) for the window
// [b0, b1).  Returns (through nk/nmin) the first step that reaches buckets
// >= b1 for some lane and the smallest start bucket of that step over the
// lanes.  The trajectory record of the next point is requested by a scalar
// load one step ahead of its use.  With verify, k is a searched guess: step
// back until no lane starts at or beyond b0 at point k.
template <int PREC>
__device__ __forceinline__ int walk_particle(const char* __restrict__ rp, int k,
                                             int NP, bool active, double Px, double Py, double Pz,
                                             double t0, double dt, double inv_dt, double tmin,
                                             double tmax, int n_bins, int b0, int b1,
                                             const Window<typename RecOf<PREC>::amp>& win, double* stage,
                                             int lane, bool verify, int& nk, int& nmin)
{
    using Pt = typename RecOf<PREC>::point;
    auto src = make_source<PREC>(rp, NP, stage, lane);
    Pt p0, p1;
    src.start(k);
    eval_point<PREC>(src.first(k), Px, Py, Pz, t0, dt, inv_dt, p0);
    if (verify) {
        int back = 32;
        while (k > 0 && __ballot(active && p0.idx >= b0)) {
            k = k > back ? k - back : 0;
            back *= 2;
            src.start(k);
            eval_point<PREC>(src.first(k), Px, Py, Pz, t0, dt, inv_dt, p0);
        }
    }
    int next_k = -1, next_min = INT_MAX;
    int steps = 0;
    while (true) {
        if (!__ballot(active && p0.idx < b1)) {
            // every lane is past the window: this and later steps belong to later windows
            if (next_k < 0) {
                next_k = k;
                next_min = wave_min(active ? p0.idx : INT_MAX);
            }
            break;
        }
        eval_point<PREC>(src.next(k + 1), Px, Py, Pz, t0, dt, inv_dt, p1);
        deposit_window(win, active, p0, p1, t0, dt, inv_dt, tmin, tmax, n_bins, b0, b1);
        ++steps;
        if (next_k < 0 && __ballot(active && p1.idx >= b1)) {
            next_k = k;
            next_min = wave_min(active ? p0.idx : INT_MAX);
        }
        ++k;
        if (k >= NP - 1) {
            if (next_k < 0) {
                next_k = NP - 1;
                next_min = INT_MAX;
            }
            break;
        }
        p0 = p1;
    }
    nk = next_k;
    nmin = next_min;
    return steps;
}

// Bucket index of trajectory point k seen from P, computed exactly as the
// full evaluation of that precision computes it (t, x, y, z lead both record
// types).
template <int PREC>
__device__ __forceinline__ int point_index(const RecBuffer<typename RecOf<PREC>::type>& rb, int k, double Px,
                                           double Py, double Pz, double t0, double dt, double inv_dt)
{
    const u32x4 a = rb.part(k, 0), b = rb.part(k, 1);
    const double t = dbl(a.x, a.y), x = dbl(a.z, a.w), y = dbl(b.x, b.y), z = dbl(b.z, b.w);
    int idx;
    if constexpr (PREC == 0) {
        double R, ts;
        timing_exact(t, Px - x, Py - y, Pz - z, t0, dt, inv_dt, R, ts, idx);
    } else {
        double yr;
        float fy, u;
        timing_mixed(t, Px - x, Py - y, Pz - z, t0, inv_dt, yr, fy, idx, u);
    }
    return idx;
}

// Walk of one particle for the window [b0, b1) in which every lane (pixel)
// follows its own step: from its start each lane first skips, with the
// retarded time alone, the steps that end before b0 (they deposit nothing in
// the window), then evaluates and deposits its steps until one ends at or
// beyond b1 (that crossing step is deposited here and again in the next
// window).  Records are read per lane, one step ahead.  Lanes thus spend full
// evaluations only on their own steps; the wave runs as long as its longest
// lane.  Returns the main-loop trip count; nk = the smallest stopping step
// over the lanes (the next window's start), nmin = the smallest bucket at
// those steps.  A lane starts at the wave-uniform k, or, with search, at the
// point before the first point (from 1 on) whose bucket is >= b0, found by a
// binary search over the retarded time (monotone in k; the skip from 0 would
// stop at the same step), or, with lc_in (not null), at k plus its offset
// from k to the step it stopped at in the previous window, saturated at 255
// (every point between lies before b0, so the skip would have passed it).
// lc_out (when not null) receives the offsets of this walk's stopping steps
// from nk.  A searched start, and a stored offset below the saturation, is
// exactly where the skip would stop (the point after it lies at or beyond
// b0), so only lanes starting at the wave's cursor or at a saturated offset
// skip.  With spill (deposit_lane), a lane whose last step ends in bucket
// b1 has deposited it whole, its b1 part into the spill registers sp, and
// stops at the step after it; carry says that the previous window of the
// item spilled, so that the skip passes such a step.
template <int PREC>
__device__ __forceinline__ int walk_lanes(const char* __restrict__ rp, int k, int NP,
                                          bool active, double Px, double Py, double Pz, double t0, double dt,
                                          double inv_dt, double tmin, double tmax, int n_bins, int b0, int b1,
                                          const Window<typename RecOf<PREC>::amp>& win, bool search,
                                          const unsigned char* __restrict__ lc_in, unsigned char* __restrict__ lc_out,
                                          bool spill, bool carry, typename RecOf<PREC>::amp* sp, int& nk, int& nmin)
{
    using Rec = typename RecOf<PREC>::type;
    const RecBuffer<Rec> rbuf(rp, NP);
    int kl = k;
    bool exact = search;
    if (search) {
        int lo = 1, hi = NP;  // first point in [lo, hi) with bucket >= b0; NP: none
        bool open = active && lo < hi;
        while (__ballot(open)) {
            if (open) {
                const int mid = (lo + hi) >> 1;
                if (point_index<PREC>(rbuf, mid, Px, Py, Pz, t0, dt, inv_dt) >= b0)
                    hi = mid;
                else
                    lo = mid + 1;
                open = lo < hi;
            }
        }
        kl = min(lo - 1, NP - 1);
    } else if (lc_in) {
        const int off = lc_in[threadIdx.x];
        kl = min(k + off, NP - 1);
        exact = off < 255;
    }
    // With carry (the previous window spilled), a step that starts before b0
    // and ends in bucket b0 was deposited whole there: skipped too.
    bool skip = active && kl < NP - 1 && !exact;
    int prev = (skip && carry) ? point_index<PREC>(rbuf, kl, Px, Py, Pz, t0, dt, inv_dt) : INT_MIN;
    while (__ballot(skip)) {
        if (skip) {
            const int nxt = point_index<PREC>(rbuf, kl + 1, Px, Py, Pz, t0, dt, inv_dt);
            if (nxt < b0 || (carry && nxt == b0 && prev < b0)) {
                ++kl;
                prev = nxt;
                skip = kl < NP - 1;
            } else {
                skip = false;
            }
        }
    }
    using Pt = typename RecOf<PREC>::point;
    Pt p0;
    eval_point<PREC>(rbuf.load(kl), Px, Py, Pz, t0, dt, inv_dt, p0);
    bool walking = active && kl < NP - 1 && p0.idx < b1;
    int steps = 0;
    // Each step first waits for the record read ahead in the previous step,
    // then issues the next one, so the loads stay in flight for a whole step.
    // (Left to itself the scheduler may issue the next loads first; the wait
    // for the old record then also waits for them, which exposes their full
    // latency every step: 25 % slower on c2.)  idx0 is the bucket of the
    // lane's current step start, which a stopped lane reports for the next
    // window.
    int idx0 = p0.idx;
    // The read-ahead of record kl + 2 is not clamped to the particle: past
    // its end the buffer resource returns zeros (kl <= NP - 1).
    Rec ra = rbuf.load(kl + 1);
    // Two steps per loop trip, with the roles of the two point buffers and
    // the two record buffers swapped between them, so the step's end point is
    // not copied into the next step's start point, nor the read-ahead record
    // into the current one (on c2 6 % faster for FP64, 1 % for mixed).
    Pt pb;
    Rec rb;
    // The lane's bucket sum (deposit_lane), added to the window as the lane
    // stops after a step that took it into bucket idx0: the spill bucket
    // (the registers sp), or the bucket of the particle's last point.
    using F = typename RecOf<PREC>::amp;
    LaneAcc<F> L;
    lane_begin(L);
    const WindowAdd<F> wadd{win};
    auto step = [&](const Rec& rc, Rec& rnext, const Pt& ps, Pt& pe) {
        __builtin_amdgcn_s_waitcnt(kWaitVmcnt0);
        __builtin_amdgcn_sched_barrier(0);
        rnext = rbuf.load(kl + 2);
        eval_point<PREC>(rc, Px, Py, Pz, t0, dt, inv_dt, pe);
        deposit_lane(win, L, walking, ps, pe, t0, dt, inv_dt, tmin, tmax, n_bins, b0, b1, spill, sp);
        ++steps;
        // The lane moves past the step if it ends inside the window, or, with
        // spill, in bucket b1 (deposited whole, the rest of it into the spill
        // bucket); it walks on from inside the window.  A lane that stops
        // past its last step adds its bucket sum to the window: the spill
        // bucket, or the bucket of the particle's last point.
        const bool past = walking && (pe.idx < b1 || (spill && pe.idx == b1));
        kl += past ? 1 : 0;
        idx0 = past ? pe.idx : idx0;
        walking = past && pe.idx < b1 && kl < NP - 1;
        if (past && !walking) {
            lane_fold(L, pe.f);
            const int j = idx0 - b0;
            if (spill && j == b1 - b0) {
#pragma unroll
                for (int c = 0; c < 6; ++c) sp[c] = sp[c] + L.acc[c];
            } else if ((unsigned)j < (unsigned)(b1 - b0)) {
                lane_close(L, wadd, j);
            }
        }
    };
    for (;;) {
        if (!__ballot(walking)) break;
        step(ra, rb, p0, pb);
        if (!__ballot(walking)) break;
        step(rb, ra, pb, p0);
    }
    const bool stopped = active && kl < NP - 1;
    const int k_next = wave_min(stopped ? kl : INT_MAX);
    nk = k_next == INT_MAX ? NP - 1 : k_next;
    nmin = wave_min(stopped ? idx0 : INT_MAX);
    if (lc_out) lc_out[threadIdx.x] = (unsigned char)min(max(kl - nk, 0), 255);
    return steps;
}

// The records of particle p (piece-major layout, rec_piece_offset).
template <int PREC>
__device__ __forceinline__ const char* particle_records(const typename RecOf<PREC>::type* recs, int p, int NP)
{
    return reinterpret_cast<const char*>(recs) + (size_t)p * rec_particle_bytes<typename RecOf<PREC>::type>(NP);
}

// First trajectory point of each particle of a batch whose bucket for one
// representative pixel is >= b: lanes = particles, binary search over the
// monotone retarded time.  The result seeds a cursor that the walk verifies.
template <int PREC>
__device__ __forceinline__ int search_first_point(const char* __restrict__ rp_lane, int NP, double Px, double Py,
                                                  double Pz, double t0, double dt, double inv_dt, int b)
{
    int lo = 0, hi = NP - 1;
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        // t, x, y, z: pieces 0 and 1 of both record types
        const unsigned stride = rec_piece_stride(NP);
        const uint4 a = *reinterpret_cast<const uint4*>(rp_lane + rec_piece_offset(mid, 0, stride)),
                    c = *reinterpret_cast<const uint4*>(rp_lane + rec_piece_offset(mid, 1, stride));
        const double t = dbl(a.x, a.y), x = dbl(a.z, a.w), y = dbl(c.x, c.y), z = dbl(c.z, c.w);
        double R, ts;
        int idx;
        timing_exact(t, Px - x, Py - y, Pz - z, t0, dt, inv_dt, R, ts, idx);
        if (idx >= b)
            hi = mid;
        else
            lo = mid + 1;
    }
    return lo;
}

// One work item (tile, particle chunk, window group) on one wave.  cur_base
// holds the item's cursors (p1-p0 int2), lane_cur (when not null) its
// per-lane offsets (kLanes bytes per particle); both are private to the wave.
template <int PREC, bool LANES>
__device__ __forceinline__ void run_item(const WorkItem& it, const ScreenDev* __restrict__ screens,
                                         const typename RecOf<PREC>::type* __restrict__ recs, int n_points,
                                         int2* __restrict__ cur_base, unsigned char* __restrict__ lane_cur,
                                         int lds_window, unsigned long long& n_steps, unsigned long long& n_walks)
{
    using F = typename RecOf<PREC>::amp;
    // Mixed: the FP32 window is flushed into the FP64 partial trace once its
    // cells may hold about kFlushTerms summed terms, so an FP32 sum never
    // runs much beyond 2048 terms (~2^-24 sqrt(2048) relative).  A per-lane
    // walk adds a lane's bucket sum once to each cell it reaches, and a
    // general step adds to a cell directly, counted as 3 terms per walk; a
    // wave-uniform walk of T wave-steps adds about T/W terms to each of its W
    // cells, counted as T/W + 1.
    constexpr unsigned kFlushTerms = 2048;
    extern __shared__ double lds[];
    const int lane = threadIdx.x;
    const ScreenDev S = screens[it.screen];
    const int pix = S.perm[it.tile * kLanes + lane];
    const bool active = pix >= 0;
    const int pl = active ? pix : S.perm[it.tile * kLanes];  // lane 0 of a tile is a real pixel
    const double Px = S.obs[3 * pl], Py = S.obs[3 * pl + 1], Pz = S.obs[3 * pl + 2];
    const double t0 = S.t0[pl];
    const double dt = S.dt, inv_dt = S.inv_dt;
    const int nb = S.n_bins, W = S.window, NP = n_points;
    double tmin, tmax;
    {
#pragma clang fp contract(off)
        tmin = t0 - dt * 0.5;                         // particle.cpp:780-781
        tmax = t0 + dt * ((double)nb - 0.5);
    }
    const Window<F> win{(typename Window<F>::LdsF*)(reinterpret_cast<F*>(lds)) + lane};
    // FP64 wave-uniform walk: two record chunks after the window
    double* stage = reinterpret_cast<double*>(reinterpret_cast<F*>(lds) + (size_t)lds_window * 6 * kLanes);
    // wave-uniform walk: seed the cursors of a group that does not start at
    // bucket 0 (the per-lane walk searches each lane's start itself)
    const bool seeded = !LANES && it.bin_lo > 0;
    if (seeded) {
        const int l0 = S.perm[it.tile * kLanes];
        const double Rx = S.obs[3 * l0], Ry = S.obs[3 * l0 + 1], Rz = S.obs[3 * l0 + 2], Rt0 = S.t0[l0];
        for (int pb = it.p0; pb < it.p1; pb += kLanes) {
            const int p = pb + lane;
            if (p < it.p1) {
                int j = search_first_point<PREC>(particle_records<PREC>(recs, p, NP), NP, Rx, Ry, Rz, Rt0, dt, inv_dt,
                                                 it.bin_lo);
                int k = j > 0 ? j - 1 : 0;
                cur_base[p - it.p0] = make_int2(k < NP - 1 ? k : NP - 1, INT_MIN);
            }
        }
    }
    // Spill bucket (per-lane walk): a window that is not the item's last has
    // one more bucket, its b1, which receives the walks' last steps that end
    // there (deposit_lane); the next window starts with it as its bucket 0.
    // Each lane adds only to its own pixel's spill bucket, so it lives in the
    // lane's registers sp, not in LDS: the window's LDS, which sets how many
    // waves fit a CU, holds its W buckets only.  carry: bucket 0 of this
    // window holds the previous window's spill bucket; carry_stored: part of
    // that bucket is already in the partial trace (the mixed kernel flushes
    // the spill bucket with the window); carry_terms: the terms it may hold
    // since.
    F sp[6];
#pragma unroll
    for (int c = 0; c < 6; ++c) sp[c] = (F)0;
    bool carry = false, carry_stored = false;
    unsigned carry_terms = 0;
    for (int b0 = it.bin_lo; b0 < it.bin_hi; b0 += W) {
        const int b1 = b0 + W < it.bin_hi ? b0 + W : it.bin_hi;
        const int nw = b1 - b0;
        const bool spill = LANES && S.spill && b1 < it.bin_hi;
        const int ncell = nw * 6;
        // chunk 0 adds into the trace store, the others store partial traces
        const size_t px = (size_t)(active ? pix : 0);
        double* out = it.chunk == 0 ? S.traces + (px * nb + b0) * 6
                                    : S.partial + (((size_t)(it.chunk - 1) * S.n_pix + px) * nb + b0) * 6;
        if (carry) {
            // the previous window's spill bucket becomes bucket 0
#pragma unroll
            for (int c = 0; c < 6; ++c) win.lane_base[c * kLanes] = sp[c];
            for (int j = 6; j < ncell; ++j) win.lane_base[j * kLanes] = (F)0;
        } else {
            for (int j = 0; j < ncell; ++j) win.lane_base[j * kLanes] = (F)0;
        }
#pragma unroll
        for (int c = 0; c < 6; ++c) sp[c] = (F)0;
        bool first_flush = true;
        bool b0_stored = carry && carry_stored;   // out of bucket 0 already holds part of it
        bool spill_stored = false;                // out of the spill bucket (bucket nw) holds part of it
        bool dirty = carry;    // the window holds terms not yet in the partial trace
        // estimated terms per cell since the last flush, times nw
        unsigned terms = carry ? carry_terms : 0;
        // component c of a window cell as an FP64 trace value (cells hold
        // twice the sums, and c B for the magnetic components)
        auto value = [&](F raw, int c) { return (double)raw * (c >= 3 ? 0.5 * kInvSpeedOfLight : 0.5); };
        // bucket j < nw (LDS), or j = nw (the spill registers), into out
        auto store = [&](int j, bool add) {
            add = add || it.chunk == 0;
#pragma unroll
            for (int c = 0; c < 6; ++c) {
                const double v = value(j == nw ? sp[c] : (F)win.lane_base[(6 * j + c) * kLanes], c);
                out[6 * j + c] = add ? out[6 * j + c] + v : v;
            }
        };
        auto flush = [&]() {
            if (active) {
                for (int j = 0; j < nw; ++j) store(j, !first_flush || (j == 0 && b0_stored));
                if (spill) store(nw, spill_stored);
            }
            for (int j = 0; j < ncell; ++j) win.lane_base[j * kLanes] = (F)0;
#pragma unroll
            for (int c = 0; c < 6; ++c) sp[c] = (F)0;
            first_flush = false;
            spill_stored = spill_stored || spill;
            dirty = false;
            terms = 0;
        };
        for (int pb = it.p0; pb < it.p1; pb += kLanes) {
            const int npb = it.p1 - pb < kLanes ? it.p1 - pb : kLanes;
            int2 cur = make_int2(0, INT_MIN);
            if ((seeded || b0 > it.bin_lo) && lane < npb) cur = cur_base[pb - it.p0 + lane];
            unsigned long long todo = __ballot(lane < npb && cur.x < NP - 1 && cur.y < b1);
            int2 ncur = cur;
            while (todo) {
                const int j = __ffsll((long long)todo) - 1;
                todo &= todo - 1;
                const int k = __builtin_amdgcn_readlane(cur.x, j);
                const bool first = __builtin_amdgcn_readlane(cur.y, j) == INT_MIN;  // first walk in this item
                const int p = pb + j;
                int nk, nmin, ws;
                if constexpr (LANES) {
                    // per-lane offsets: read when a previous window of this
                    // item stored them, written when a next window follows
                    unsigned char* lc = lane_cur ? lane_cur + (size_t)(p - it.p0) * kLanes : nullptr;
                    ws = walk_lanes<PREC>(particle_records<PREC>(recs, p, NP), k, NP, active, Px, Py, Pz, t0, dt,
                                          inv_dt, tmin, tmax, nb, b0, b1, win, first, first ? nullptr : lc,
                                          b1 < it.bin_hi ? lc : nullptr, spill, carry, sp, nk, nmin);
                } else {
                    const bool verify = first && k > 0;
                    ws = walk_particle<PREC>(particle_records<PREC>(recs, p, NP), k, NP, active, Px, Py, Pz, t0, dt, inv_dt, tmin,
                                             tmax, nb, b0, b1, win, stage, lane, verify, nk, nmin);
                }
                n_steps += ws;
                ++n_walks;
                if (lane == j) ncur = make_int2(nk, nmin);
                if constexpr (PREC == 1) {
                    // terms counted in units of 1/nw
                    dirty = true;
                    terms += LANES ? 3u * (unsigned)nw : (unsigned)ws + (unsigned)nw;
                    if (terms >= kFlushTerms * (unsigned)nw) flush();
                }
            }
            if (b1 < it.bin_hi && lane < npb) cur_base[pb - it.p0 + lane] = ncur;
        }
        // the window's buckets into the partial trace; the spill bucket stays
        // in the registers for the next window
        if (active) {
            if (PREC == 0 || first_flush) {
                for (int j = 0; j < nw; ++j) store(j, j == 0 && b0_stored);
            } else if (dirty) {
                for (int j = 0; j < nw; ++j) store(j, true);
            }
        }
        carry = spill;
        carry_stored = spill_stored;
        carry_terms = terms;
    }
}

// Persistent waves: each wave takes work items in list order from the
// counter next_item (zero at launch) until the list is exhausted, and keeps
// the cursors of its current item in its own slot of slot_store (slot_bytes
// per wave: slot_particles int2 cursors, then, when lane_offsets, kLanes
// bytes per particle).  The grid is sized to the resident waves, so the
// cursor store scales with the machine, not with the work list.  p_shift
// (when not null) is read at launch and added to every item's particle
// range, so that a captured launch replays over another particle batch of
// the same size.
template <int PREC, bool LANES>
__global__ __launch_bounds__(64) void radiation_kernel(const WorkItem* __restrict__ items, int n_items,
                                                        int* __restrict__ next_item,
                                                        const ScreenDev* __restrict__ screens,
                                                        const typename RecOf<PREC>::type* __restrict__ recs,
                                                        int n_points, unsigned char* __restrict__ slot_store,
                                                        size_t slot_bytes, int slot_particles, int lane_offsets,
                                                        unsigned long long* __restrict__ stats, int lds_window,
                                                        const int* __restrict__ p_shift)
{
    const int shift = p_shift ? __builtin_amdgcn_readfirstlane(*p_shift) : 0;
    unsigned char* slot = slot_store + (size_t)blockIdx.x * slot_bytes;
    int2* cur_base = reinterpret_cast<int2*>(slot);
    unsigned char* lane_cur = lane_offsets ? slot + (size_t)slot_particles * sizeof(int2) : nullptr;
    for (;;) {
        int i = 0;
        if (threadIdx.x == 0) i = atomicAdd(next_item, 1);
        i = __builtin_amdgcn_readfirstlane(i);
        if (i >= n_items) break;
        WorkItem it = items[i];
        it.p0 += shift;
        it.p1 += shift;
        unsigned long long n_steps = 0, n_walks = 0;
        run_item<PREC, LANES>(it, screens, recs, n_points, cur_base, lane_cur, lds_window, n_steps, n_walks);
        if (stats && threadIdx.x == 0) {
            stats[2 * (size_t)i] = n_steps;
            stats[2 * (size_t)i + 1] = n_walks;
        }
    }
}

template __global__ void radiation_kernel<0, false>(const WorkItem*, int, int*, const ScreenDev*, const RecF64*, int,
                                                    unsigned char*, size_t, int, int, unsigned long long*, int,
                                                    const int*);
template __global__ void radiation_kernel<1, false>(const WorkItem*, int, int*, const ScreenDev*, const RecMixed*, int,
                                                    unsigned char*, size_t, int, int, unsigned long long*, int,
                                                    const int*);
template __global__ void radiation_kernel<0, true>(const WorkItem*, int, int*, const ScreenDev*, const RecF64*, int,
                                                    unsigned char*, size_t, int, int, unsigned long long*, int,
                                                    const int*);
template __global__ void radiation_kernel<1, true>(const WorkItem*, int, int*, const ScreenDev*, const RecMixed*, int,
                                                    unsigned char*, size_t, int, int, unsigned long long*, int,
                                                    const int*);

// ---------------------------------------------------------------------------
// Add the partial traces of chunks 1.. into the trace store (which chunk 0
// added into), chunks in order.
// ---------------------------------------------------------------------------
__global__ void reduce_partials_kernel(const double* __restrict__ partial, double* __restrict__ out,
                                       size_t n, int n_chunks)
{
    size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        double s = out[i];
        for (int c = 0; c < n_chunks; ++c) s += partial[(size_t)c * n + i];
        out[i] = s;
    }
}

__global__ void set_int_kernel(int* __restrict__ p, int v)
{
    if (threadIdx.x == 0 && blockIdx.x == 0) *p = v;
}

__global__ void fill_kernel(double* __restrict__ p, size_t n, double v)
{
    size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) p[i] = v;
}

// ---------------------------------------------------------------------------
// Deposit only: injected (t_ret, field) per point of pseudo-particles into one
// trace in global memory, with the deposit code of the radiation kernel.
// ---------------------------------------------------------------------------
struct GlobalAcc {
    double* trace;
    const double* f1;
    const double* f2;
    __device__ __forceinline__ void operator()(int j, double a, double b) const
    {
        double* cell = trace + 6 * j;
        for (int c = 0; c < 6; ++c) cell[c] += a * f1[c] + b * f2[c];
    }
};

__global__ void deposit_only_kernel(int n_pseudo, int n_points, const double* __restrict__ ts,
                                    const double* __restrict__ f, double t0, double dt, double inv_dt,
                                    int n_bins, double* __restrict__ trace)
{
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    double tmin, tmax;
    {
#pragma clang fp contract(off)
        tmin = t0 - dt * 0.5;
        tmax = t0 + dt * ((double)n_bins - 0.5);
    }
    for (int p = 0; p < n_pseudo; ++p) {
        const double* tp = ts + (size_t)p * n_points;
        const double* fp = f + (size_t)p * n_points * 6;
        for (int k = 0; k + 1 < n_points; ++k) {
            double ts1 = tp[k], ts2 = tp[k + 1];
            int i1 = round_index(div_by_const(ts1 - t0, dt, inv_dt));
            int i2 = round_index(div_by_const(ts2 - t0, dt, inv_dt));
            GlobalAcc acc{trace, fp + 6 * k, fp + 6 * (k + 1)};
            deposit_step(acc, RcpRn{}, true, ts1, i1, ts2, i2, t0, dt, inv_dt, tmin, tmax, n_bins, 0, n_bins);
        }
    }
}

// ---------------------------------------------------------------------------
// Field at given times and points: ChargedParticle::RetardedField(double,
// Vector) (src/particle.cpp:612-769) summed over particles in order
// (src/bunch.cpp:729-737, src/beam.cpp:433-441), one lane per observation.
// The light-cone search (times, distances, bisection decisions, secant
// refinement and its stopping test) is computed with correctly rounded
// operations in the reference's order, so it takes the reference's path and
// lands on the reference's interpolation fraction; the field amplitudes then
// agree to rounding.
// ---------------------------------------------------------------------------
namespace snap {

struct V {
    double x, y, z;
};
__device__ __forceinline__ V mk(double x, double y, double z) { return V{x, y, z}; }

__device__ __forceinline__ double div_rn(double a, double b) { return div_by_const(a, b, rcp_rn(b)); }

// correctly rounded sqrt for any non-negative input
__device__ __forceinline__ double sqrt_any(double x)
{
    return (x > 1e-300 && x < 1e300) ? sqrt_rn(x) : __builtin_sqrt(x);
}

struct Traj {
    const double *t, *x, *y, *z, *ux, *uy, *uz, *ax, *ay, *az;
};

}  // namespace snap

__global__ void field_at_kernel(int n_obs, const double* __restrict__ obs_time, const double* __restrict__ obs_pos,
                                snap::Traj tr, const double* __restrict__ charge, int NP, int p0, int p1,
                                double* __restrict__ out, int accumulate)
{
#pragma clang fp contract(off)
    using snap::V;
    using snap::mk;
    using snap::div_rn;
    using snap::sqrt_any;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n_obs) return;
    const double T = obs_time[i];
    const V O = mk(obs_pos[3 * i], obs_pos[3 * i + 1], obs_pos[3 * i + 2]);
    auto sub = [](V a, V b) { return mk(a.x - b.x, a.y - b.y, a.z - b.z); };
    auto add = [](V a, V b) { return mk(a.x + b.x, a.y + b.y, a.z + b.z); };
    auto mul = [](V a, double f) { return mk(a.x * f, a.y * f, a.z * f); };
    auto divv = [](V a, double f) {
        const double y = rcp_rn(f);
        return mk(div_by_const(a.x, f, y), div_by_const(a.y, f, y), div_by_const(a.z, f, y));
    };
    auto norm = [&](V a) { return sqrt_any(a.x * a.x + a.y * a.y + a.z * a.z); };
    auto dot = [](V a, V b) {
        double t = a.x * b.x;
        t += a.y * b.y;
        t += a.z * b.z;
        return t;
    };
    auto cross = [](V a, V b) { return mk(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x); };
    double acc[6] = {0, 0, 0, 0, 0, 0};
    for (int p = p0; p < p1; ++p) {
        const size_t base = (size_t)p * NP;
        auto X = [&](int k) { return mk(tr.x[base + k], tr.y[base + k], tr.z[base + k]); };
        auto U = [&](int k) { return mk(tr.ux[base + k], tr.uy[base + k], tr.uz[base + k]); };
        auto A = [&](int k) { return mk(tr.ax[base + k], tr.ay[base + k], tr.az[base + k]); };
        auto tret = [&](int k) { return T - (tr.t[base + k] + div_by_const(norm(sub(O, X(k))), kSpeedOfLight, kInvSpeedOfLight)); };
        int i1 = 0, i2 = NP - 1;
        double t1 = tret(i1), t2 = tret(i2);
        if (!(t2 <= 0.0)) continue;                              // particle.cpp:633: no signal yet
        V SX, SB, SBP;
        double gamma2;
        if (t1 <= 0.0) {
            // particle.cpp:641-660: back-extrapolation with the start velocity
            const V Rc = divv(sub(O, X(0)), kSpeedOfLight);
            SB = U(0);
            gamma2 = (SB.x * SB.x + SB.y * SB.y + SB.z * SB.z) + 1.0;
            const double gamma = sqrt_any(gamma2);
            SB = divv(SB, gamma);
            const double bR = dot(SB, Rc), bb = dot(SB, SB), RR = dot(Rc, Rc);
            const double ST = div_rn(T - bR - sqrt_any((T - bR) * (T - bR) - (1 - bb) * (T * T - RR)), 1 - bb);
            SX = add(X(0), mul(mul(SB, kSpeedOfLight), ST));
            SBP = mk(0.0, 0.0, 0.0);
        } else {
            // particle.cpp:662-745: bisection to one step, secant refinement
            while (i2 - i1 > 1) {
                const int im = (i2 + i1) / 2;
                const double tm = tret(im);
                if (tm > 0.0) {
                    i1 = im;
                    t1 = tm;
                } else {
                    i2 = im;
                    t2 = tm;
                }
            }
            double frac = div_rn(t1, t1 - t2);
            const double T1 = tr.t[base + i1], T2 = tr.t[base + i2];
            const V X1 = X(i1), X2 = X(i2);
            double ST = T1 * (1.0 - frac) + T2 * frac;
            SX = add(mul(X1, 1.0 - frac), mul(X2, frac));
            double R = norm(sub(O, SX));
            double dtv = ST + div_by_const(R, kSpeedOfLight, kInvSpeedOfLight) - T;
            const double r1 = norm(sub(O, X1)), r2 = norm(sub(O, X2));
            double dtdfrac = (T2 - T1) + div_by_const(r2 - r1, kSpeedOfLight, kInvSpeedOfLight);
            double dfrac = div_rn(dtv, dtdfrac);
            int number_ref = 1;
            while (__builtin_fabs(dfrac) > 1e-5 && number_ref < 10) {
                const double last = dtv;
                frac -= dfrac;
                ST = T1 * (1.0 - frac) + T2 * frac;
                SX = add(mul(X1, 1.0 - frac), mul(X2, frac));
                R = norm(sub(O, SX));
                dtv = ST + div_by_const(R, kSpeedOfLight, kInvSpeedOfLight) - T;
                dtdfrac = div_rn(-(dtv - last), dfrac);
                dfrac = div_rn(dtv, dtdfrac);
                number_ref++;
            }
            frac -= dfrac;
            SX = add(mul(X1, 1.0 - frac), mul(X2, frac));
            SB = add(mul(U(i1), 1.0 - frac), mul(U(i2), frac));
            gamma2 = (SB.x * SB.x + SB.y * SB.y + SB.z * SB.z) + 1.0;
            const double gamma = sqrt_any(gamma2);
            SB = divv(SB, gamma);
            SBP = divv(add(mul(A(i1), 1.0 - frac), mul(A(i2), frac)), gamma);
        }
        // particle.cpp:747-766
        const V RV = sub(O, SX);
        const double R = norm(RV);
        const V N = divv(RV, R);
        const double scale = charge[p] * kElementaryCharge / (4.0 * kPi * kEpsNull);
        const double k1 = 1.0 - dot(SB, N);
        const double bn3 = k1 * k1 * k1;
        const V NmB = sub(N, SB);
        V E = divv(NmB, R * R * bn3 * gamma2);
        E = add(E, divv(cross(N, cross(NmB, SBP)), R * bn3 * kSpeedOfLight));
        E = mul(E, scale);
        const V B = divv(cross(N, E), kSpeedOfLight);
        acc[0] += E.x;
        acc[1] += E.y;
        acc[2] += E.z;
        acc[3] += B.x;
        acc[4] += B.y;
        acc[5] += B.z;
    }
    for (int c = 0; c < 6; ++c) out[6 * (size_t)i + c] = accumulate ? out[6 * (size_t)i + c] + acc[c] : acc[c];
}

// ---------------------------------------------------------------------------
// Pairs (pixel, particle step) that pass the window test of
// src/particle.cpp:778-793, ts2 > tmin && ts1 < tmax, with the retarded times
// computed exactly as the FP64 kernel (and the reference) computes them from
// the raw trajectory arrays (t, x, y, z).  One thread per (pixel, particle),
// pixels fastest; the retarded time is monotone
// along a trajectory, so the first point after tmin and the last before tmax
// are found by binary search and the passing steps lie between them.  Each
// block adds its sum to *total.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void inwin_count_kernel(const double* __restrict__ obs,
                                                          const double* __restrict__ t0v, int n_pix, double dt,
                                                          double inv_dt, int n_bins,
                                                          const double* __restrict__ rt,
                                                          const double* __restrict__ rx,
                                                          const double* __restrict__ ry,
                                                          const double* __restrict__ rz, int n_particles, int NP,
                                                          unsigned long long* __restrict__ total)
{
    __shared__ unsigned long long part[256 / kLanes];
    const size_t gid = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    unsigned long long cnt = 0;
    if (gid < (size_t)n_pix * n_particles) {
        const int i = (int)(gid % n_pix), p = (int)(gid / n_pix);
        const double Px = obs[3 * i], Py = obs[3 * i + 1], Pz = obs[3 * i + 2], t0 = t0v[i];
        double tmin, tmax;
        {
#pragma clang fp contract(off)
            tmin = t0 - dt * 0.5;                      // particle.cpp:780-781
            tmax = t0 + dt * ((double)n_bins - 0.5);
        }
        const size_t base = (size_t)p * NP;
        auto ts_of = [&](int k) {
            double R, ts;
            int idx;
            const size_t g = base + k;
            timing_exact(rt[g], Px - rx[g], Py - ry[g], Pz - rz[g], t0, dt, inv_dt, R, ts, idx);
            return ts;
        };
        int lo = 0, hi = NP;  // a = first point with ts > tmin (NP: none)
        while (lo < hi) {
            const int mid = (lo + hi) >> 1;
            if (ts_of(mid) > tmin) hi = mid; else lo = mid + 1;
        }
        const int a = lo;
        lo = 0;
        hi = NP;              // first point with ts >= tmax; b = the one before it
        while (lo < hi) {
            const int mid = (lo + hi) >> 1;
            if (ts_of(mid) >= tmax) hi = mid; else lo = mid + 1;
        }
        const int b = lo - 1;
        // step k passes when point k+1 is after tmin (k >= a-1) and point k before tmax (k <= b)
        const int k0 = a - 1 > 0 ? a - 1 : 0, k1 = b < NP - 2 ? b : NP - 2;
        cnt = k1 >= k0 ? (unsigned long long)(k1 - k0 + 1) : 0ull;
    }
    for (int off = kLanes / 2; off > 0; off >>= 1) cnt += __shfl_xor(cnt, off);
    if ((threadIdx.x & (kLanes - 1)) == 0) part[threadIdx.x / kLanes] = cnt;
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned long long s = 0;
        for (int w = 0; w < (int)(blockDim.x / kLanes); ++w) s += part[w];
        atomicAdd(total, s);
    }
}

// ---------------------------------------------------------------------------
// Spectra of the in-plane fields (scripts/teufel.py:167-226): per pixel and
// sample, EH = E.ex, EV = E.ey, BH = B.ex, BV = B.ey; sig holds four signals
// of n_bins samples per pixel, [EH | BV | EV | BH][n_pix][n_bins].
// ---------------------------------------------------------------------------
__global__ void inplane_kernel(const double* __restrict__ traces, size_t n, double ex0, double ex1, double ex2,
                               double ey0, double ey1, double ey2, double* __restrict__ sig)
{
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t j = (size_t)blockIdx.x * blockDim.x + threadIdx.x; j < n; j += stride) {
        const double* f = traces + 6 * j;
        sig[j] = f[0] * ex0 + f[1] * ex1 + f[2] * ex2;              // EH
        sig[n + j] = f[3] * ey0 + f[4] * ey1 + f[5] * ey2;          // BV
        sig[2 * n + j] = f[0] * ey0 + f[1] * ey1 + f[2] * ey2;      // EV
        sig[3 * n + j] = f[3] * ex0 + f[4] * ex1 + f[5] * ex2;      // BH
    }
}

// |F(EH) conj(F(BV))| / mu0 * 2 dt / (df nf) with df = 1/(nf dt), i.e. * 2 dt^2,
// for the first n_bins/2 frequencies (scripts/teufel.py:198-211); the same
// for EV, BH (:213-226).  spec: the four signals' one-sided transforms,
// n_bins/2 + 1 complex values each.
__global__ void spectral_density_kernel(const double2* __restrict__ spec, int n_pix, int n_bins, double scale,
                                        double* __restrict__ dens_h, double* __restrict__ dens_v)
{
    const int nh = n_bins / 2, nc = n_bins / 2 + 1;
    const size_t n = (size_t)n_pix * nh, sig = (size_t)n_pix * nc;
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t j = (size_t)blockIdx.x * blockDim.x + threadIdx.x; j < n; j += stride) {
        const size_t p = j / nh, i = j % nh, k = p * nc + i;
        const double2 eh = spec[k], bv = spec[sig + k], ev = spec[2 * sig + k], bh = spec[3 * sig + k];
        // |a conj(b)| = |a| |b|
        dens_h[j] = sqrt(eh.x * eh.x + eh.y * eh.y) * sqrt(bv.x * bv.x + bv.y * bv.y) * scale;
        dens_v[j] = sqrt(ev.x * ev.x + ev.y * ev.y) * sqrt(bh.x * bh.x + bh.y * bh.y) * scale;
    }
}

// ---------------------------------------------------------------------------
// Per-pixel Poynting integral S = dt * sum E x B / mu0 (src/fieldtrace.cpp:192-198).
// ---------------------------------------------------------------------------
__global__ void poynting_kernel(const double* __restrict__ traces, int n_pix, int n_bins, double dt,
                                double* __restrict__ out)
{
    int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n_pix) return;
    const double* tr = traces + (size_t)p * n_bins * 6;
    double sx = 0.0, sy = 0.0, sz = 0.0;
    for (int i = 0; i < n_bins; ++i) {
        const double* f = tr + 6 * i;
        sx += (f[1] * f[5] - f[2] * f[4]) / kMuNull;
        sy += (f[2] * f[3] - f[0] * f[5]) / kMuNull;
        sz += (f[0] * f[4] - f[1] * f[3]) / kMuNull;
    }
    out[3 * p] = sx * dt;
    out[3 * p + 1] = sy * dt;
    out[3 * p + 2] = sz * dt;
}

}  // namespace tg
