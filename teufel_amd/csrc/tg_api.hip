// tg_api.hip - host side of the C ABI declared in include/teufel_gpu.h.
//
// One tg_context = one device, one stream, one trajectory set and a list of
// screens whose traces live in HBM (the on-device trace store that replaces
// the per-rank getBuffer / MPI_Reduce / fromBuffer of
// src/teufel_hybrid.cpp:726-763).  Kernels are in lw_kernels.hip and
// vay_kernel.hip, compiled into this translation unit.
#include "lw_kernels.hip"
#include "vay_kernel.hip"

#include "../../include/teufel_gpu.h"

#include <rocfft/rocfft.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

using namespace tg;

namespace {

constexpr int kNumRaw = 10;  // t, x, y, z, ux, uy, uz, ax, ay, az

struct ScreenHost {
    int n_pix = 0, n_bins = 0;
    double dt = 0.0;
    std::vector<double> obs, t0;
    double* d_obs = nullptr;
    double* d_t0 = nullptr;
    double* d_traces = nullptr;
    int* d_perm = nullptr;
    int n_tiles = 0;
    bool perm_valid = false;
    int window[2] = {0, 0};  // autotuned bins per window per precision (0 = default)
    // per tile: bucket position of its first pixel at the profile steps
    // (kProfile samples of the middle particle), for work estimates
    std::vector<double> tile_prof;
};

constexpr int kProfile = 9;

struct LaunchPlan {
    int n_chunks = 0;
    int chunk = 0;  // particles per chunk
    std::vector<WorkItem> items;
    std::vector<ScreenDev> sdev;
    std::vector<int> screen_ids;
    size_t partial_doubles = 0;
    int slots = 0;              // persistent waves (grid size)
    size_t slot_bytes = 0;      // cursor store per wave
    int max_window = 0;
    int groups = 1;
    bool lanes = true;  // per-lane walk (tg_tuning.walk == 2)
    bool spill = true;  // windows with a spill bucket (per-lane walk)
    bool lane_cursors = false;  // per-lane cursor offsets kept between windows
    size_t lds = 0;             // dynamic LDS per wave
};

}  // namespace

struct tg_context {
    int device = 0;
    hipStream_t stream = nullptr;
    std::string err;
    char name[256] = {0};
    int n_cu = 0;
    size_t max_lds = 65536;
    // trajectories
    int n_particles = 0, n_points = 0;
    bool traj_set = false;
    double* d_raw = nullptr;     // kNumRaw arrays of n_particles*n_points, then charge[n_particles]
    void* d_rec[2] = {nullptr, nullptr};
    int* d_flag = nullptr;
    bool rec_valid[2] = {false, false};
    // screens
    std::vector<ScreenHost> screens;
    // launch scratch
    WorkItem* d_items = nullptr;
    size_t cap_items = 0;
    ScreenDev* d_sdev = nullptr;
    size_t cap_sdev = 0;
    unsigned char* d_slots = nullptr;  // per-wave cursor slots of the persistent radiation kernel
    size_t cap_slots = 0;
    int* d_next_item = nullptr;        // work-list counter of the persistent radiation kernel
    size_t cap_next_item = 0;
    int* d_shift = nullptr;            // particle shift a captured launch reads (graph replays)
    size_t cap_shift = 0;
    double* d_partial = nullptr;
    size_t cap_partial = 0;
    // walk statistics (diagnostic; on when TG_STATS=1 in the environment)
    bool stats_on = false;
    int default_walk = 2;
    bool lane_cursors_on = true;  // TG_LANE_CURSORS=0 turns the per-lane offsets off
    int max_slots = 0;            // TG_SLOTS=n caps the persistent grid at n waves (0: resident waves)
    int fail_trial = 0;           // TG_AUTOTUNE_FAIL_TRIAL=n fails tg_autotune's n-th trial (error-path tests)
    unsigned long long* d_stats = nullptr;
    size_t cap_stats = 0;
    long long last_steps = 0, last_walks = 0, last_items = 0;
    // timing
    hipEvent_t ev[3] = {nullptr, nullptr, nullptr};
    float kernel_ms = 0.f, total_ms = 0.f;
    // tuning
    tg_tuning tuned[2] = {};
    bool has_tuned[2] = {false, false};
    // this device's static scratch array (tg_integrate_range_static)
    void* static_scratch = nullptr;
    std::string static_scratch_err;
    // graph cache: the captured launch of key, over particles from graph_p0;
    // a replay over another range of the same size shifts them (d_shift)
    hipGraphExec_t gexec = nullptr;
    hipGraph_t graph = nullptr;
    std::string graph_key;
    int graph_p0 = 0;
};

namespace {

int fail(tg_context* c, int code, const std::string& msg)
{
    if (c) c->err = msg;
    return code;
}

#define TG_HIP(c, call)                                                                       \
    do {                                                                                      \
        hipError_t e_ = (call);                                                               \
        if (e_ != hipSuccess)                                                                 \
            return fail((c), e_ == hipErrorOutOfMemory ? TG_ERR_OOM : TG_ERR_HIP,             \
                        std::string(#call) + ": " + hipGetErrorString(e_));                   \
    } while (0)

template <class T>
int ensure_cap(tg_context* c, T*& ptr, size_t& cap, size_t n)
{
    if (n <= cap && ptr) return TG_OK;
    if (ptr) {
        hipError_t e = hipFree(ptr);
        (void)e;
        ptr = nullptr;
        cap = 0;
    }
    size_t want = std::max<size_t>(n, 1);
    TG_HIP(c, hipMalloc((void**)&ptr, want * sizeof(T)));
    cap = want;
    return TG_OK;
}

void free_graph(tg_context* c)
{
    if (c->gexec) (void)hipGraphExecDestroy(c->gexec);
    if (c->graph) (void)hipGraphDestroy(c->graph);
    c->gexec = nullptr;
    c->graph = nullptr;
    c->graph_key.clear();
}

void invalidate_traj(tg_context* c)
{
    c->rec_valid[0] = c->rec_valid[1] = false;
    for (auto& s : c->screens) s.perm_valid = false;
    free_graph(c);
}

int free_traj(tg_context* c)
{
    if (c->d_raw) TG_HIP(c, hipFree(c->d_raw));
    for (int i = 0; i < 2; ++i)
        if (c->d_rec[i]) TG_HIP(c, hipFree(c->d_rec[i]));
    c->d_raw = nullptr;
    c->d_rec[0] = c->d_rec[1] = nullptr;
    c->n_particles = c->n_points = 0;
    invalidate_traj(c);
    return TG_OK;
}

double* raw(tg_context* c, int i) { return c->d_raw + (size_t)i * c->n_particles * c->n_points; }
double* raw_charge(tg_context* c) { return c->d_raw + (size_t)kNumRaw * c->n_particles * c->n_points; }

int alloc_traj(tg_context* c, int n_particles, int n_points)
{
    // the radiation kernel addresses one particle's packed records (96 B per
    // point in FP64) with 32-bit signed byte offsets
    if ((long long)rec_particle_bytes<RecF64>(n_points) > (long long)INT_MAX)
        return fail(c, TG_ERR_ARG, "more than 22 million points per particle");
    int rc = free_traj(c);
    if (rc) return rc;
    c->traj_set = false;   // until the new set is allocated
    size_t n = (size_t)n_particles * (size_t)n_points;
    TG_HIP(c, hipMalloc((void**)&c->d_raw, std::max<size_t>(1, kNumRaw * n + (size_t)n_particles) * sizeof(double)));
    c->n_particles = n_particles;
    c->n_points = n_points;
    c->traj_set = true;
    // the tuned chunk and group counts were timed on the previous set
    for (tg_tuning& t : c->tuned) t.particles_per_chunk = t.window_groups = t.spill = 0;
    return TG_OK;
}

int ensure_records(tg_context* c, int prec)
{
    if (c->rec_valid[prec]) return TG_OK;
    size_t n = (size_t)c->n_particles * c->n_points;
    if (!c->d_rec[prec]) {
        // piece-major layout (rec_piece_offset): the padding points of every
        // piece zero (the pack kernels write only real points)
        const size_t bytes = std::max<size_t>(1, (size_t)c->n_particles * (prec == TG_FP64
                                                                               ? rec_particle_bytes<RecF64>(c->n_points)
                                                                               : rec_particle_bytes<RecMixed>(c->n_points)));
        TG_HIP(c, hipMalloc(&c->d_rec[prec], bytes));
        TG_HIP(c, hipMemsetAsync(c->d_rec[prec], 0, bytes, c->stream));
    }
    if (n) {
        unsigned blocks = (unsigned)((n + 255) / 256);
        if (!c->d_flag) TG_HIP(c, hipMalloc((void**)&c->d_flag, sizeof(int)));
        TG_HIP(c, hipMemsetAsync(c->d_flag, 0, sizeof(int), c->stream));
        if (prec == TG_FP64)
            pack_f64_kernel<<<blocks, 256, 0, c->stream>>>(n, c->n_points, raw(c, 0), raw(c, 1), raw(c, 2),
                                                         raw(c, 3), raw(c, 4), raw(c, 5), raw(c, 6), raw(c, 7),
                                                         raw(c, 8), raw(c, 9), raw_charge(c),
                                                         (char*)c->d_rec[prec], c->d_flag);
        else
            pack_mixed_kernel<<<blocks, 256, 0, c->stream>>>(n, c->n_points, raw(c, 0), raw(c, 1), raw(c, 2),
                                                           raw(c, 3), raw(c, 4), raw(c, 5), raw(c, 6), raw(c, 7),
                                                           raw(c, 8), raw(c, 9), raw_charge(c),
                                                           (char*)c->d_rec[prec], c->d_flag);
        TG_HIP(c, hipGetLastError());
        int flags = 0;
        TG_HIP(c, hipMemcpyAsync(&flags, c->d_flag, sizeof(int), hipMemcpyDeviceToHost, c->stream));
        TG_HIP(c, hipStreamSynchronize(c->stream));
        if (flags & kFlagSuperluminal)
            return fail(c, TG_ERR_ARG,
                        "a trajectory step moves faster than light (|dX| > c dt) or back in time: the retarded time "
                        "must increase along every trajectory");
        if (prec == TG_MIXED && (flags & kFlagOffAxis))
            return fail(c, TG_ERR_ARG,
                        "TG_MIXED needs beams travelling close to the z axis (|beta_x|+|beta_y| <= 0.25); "
                        "use TG_FP64");
    }
    c->rec_valid[prec] = true;
    return TG_OK;
}

bool finite_all(const double* a, size_t n)
{
    for (size_t i = 0; i < n; ++i)
        if (!std::isfinite(a[i])) return false;
    return true;
}

// Group pixels into tiles of 64 whose arrival-time curves are close, so the
// wave's lanes walk nearly the same trajectory steps for each window.  The
// profile of a pixel is the bucket position (t_ret - t0)/dt of the middle
// particle at 9 points along its trajectory; tiles are the leaves of a k-d
// split (largest spread first) cut at multiples of 64.  Two tilings are
// built, one on the profile and one on its slopes (differences of
// consecutive samples), and the one with the fewer modelled wave-steps is
// kept (tiling_cost): on c2 and c3 the slopes give ~3 % fewer, on c5 the
// profile ~13 % fewer.
void kd_split(std::vector<int>& idx, const std::vector<double>& feat, int nf, int begin, int end)
{
    int n = end - begin;
    if (n <= kLanes) return;
    int tiles = (n + kLanes - 1) / kLanes;
    int mid = begin + (tiles / 2) * kLanes;
    int best = 0;
    double spread = -1.0;
    for (int d = 0; d < nf; ++d) {
        double lo = 1e300, hi = -1e300;
        for (int i = begin; i < end; ++i) {
            double v = feat[(size_t)idx[i] * nf + d];
            lo = std::min(lo, v);
            hi = std::max(hi, v);
        }
        if (hi - lo > spread) {
            spread = hi - lo;
            best = d;
        }
    }
    std::nth_element(idx.begin() + begin, idx.begin() + mid, idx.begin() + end, [&](int a, int b) {
        double fa = feat[(size_t)a * nf + best], fb = feat[(size_t)b * nf + best];
        return fa < fb || (fa == fb && a < b);
    });
    kd_split(idx, feat, nf, begin, mid);
    kd_split(idx, feat, nf, mid, end);
}

// Modelled wave-steps of a tiling for one particle's trajectory (host arrays
// t, x, y, z of np points, every stride-th used): per pixel, the number of
// steps that end in each window of kModelWindow buckets; per tile and
// window, the largest count over its pixels (the wave runs as long as its
// longest lane); summed.
constexpr int kModelWindow = 12;
double tiling_cost(const ScreenHost& s, const std::vector<int>& perm, const std::vector<double>& traj, int np,
                   int stride)
{
    const int nwin = (s.n_bins + kModelWindow - 1) / kModelWindow;
    std::vector<int> cnt((size_t)kLanes * nwin);
    double cost = 0.0;
    for (size_t t0i = 0; t0i < perm.size(); t0i += kLanes) {
        std::fill(cnt.begin(), cnt.end(), 0);
        for (int l = 0; l < kLanes && t0i + l < perm.size(); ++l) {
            const int i = perm[t0i + l];
            if (i < 0) continue;
            for (int k = stride; k < np; k += stride) {
                const double dx = s.obs[3 * i] - traj[4 * k + 1], dy = s.obs[3 * i + 1] - traj[4 * k + 2],
                             dz = s.obs[3 * i + 2] - traj[4 * k + 3];
                const double q = (traj[4 * k] + std::sqrt(dx * dx + dy * dy + dz * dz) / kSpeedOfLight - s.t0[i]) / s.dt;
                const double w = std::floor((q + 0.5) / kModelWindow);
                if (w >= 0.0 && w < (double)nwin) ++cnt[(size_t)l * nwin + (size_t)w];
            }
        }
        for (int w = 0; w < nwin; ++w) {
            int m = 0;
            for (int l = 0; l < kLanes; ++l) m = std::max(m, cnt[(size_t)l * nwin + w]);
            cost += m;
        }
    }
    return cost;
}

int ensure_tiles(tg_context* c, ScreenHost& s)
{
    if (s.perm_valid) return TG_OK;
    const int nf = kProfile;
    std::vector<double> samp(4 * nf, 0.0);
    if (c->n_particles > 0 && c->n_points > 0) {
        int p = c->n_particles / 2;
        for (int j = 0; j < nf; ++j) {
            size_t k = (size_t)((double)(c->n_points - 1) * j / (nf - 1));
            size_t g = (size_t)p * c->n_points + k;
            for (int q = 0; q < 4; ++q)
                TG_HIP(c, hipMemcpy(&samp[4 * j + q], raw(c, q) + g, sizeof(double), hipMemcpyDeviceToHost));
        }
    }
    std::vector<double> feat((size_t)s.n_pix * nf);
    for (int i = 0; i < s.n_pix; ++i)
        for (int j = 0; j < nf; ++j) {
            double dx = s.obs[3 * i] - samp[4 * j + 1], dy = s.obs[3 * i + 1] - samp[4 * j + 2],
                   dz = s.obs[3 * i + 2] - samp[4 * j + 3];
            double ts = samp[4 * j] + std::sqrt(dx * dx + dy * dy + dz * dz) / kSpeedOfLight;
            feat[(size_t)i * nf + j] = (ts - s.t0[i]) / s.dt;
        }
    std::vector<int> idx(s.n_pix);
    for (int i = 0; i < s.n_pix; ++i) idx[i] = i;
    kd_split(idx, feat, nf, 0, s.n_pix);
    if (c->n_particles > 0 && c->n_points > 2 && s.n_pix > kLanes) {
        std::vector<double> slope((size_t)s.n_pix * (nf - 1));
        for (int i = 0; i < s.n_pix; ++i)
            for (int j = 0; j + 1 < nf; ++j)
                slope[(size_t)i * (nf - 1) + j] = feat[(size_t)i * nf + j + 1] - feat[(size_t)i * nf + j];
        std::vector<int> idx2(s.n_pix);
        for (int i = 0; i < s.n_pix; ++i) idx2[i] = i;
        kd_split(idx2, slope, nf - 1, 0, s.n_pix);
        // the middle particle's trajectory, about 1000 points of it
        const int np = c->n_points, p = c->n_particles / 2;
        std::vector<double> traj((size_t)4 * np), col(np);
        for (int q = 0; q < 4; ++q) {
            TG_HIP(c, hipMemcpy(col.data(), raw(c, q) + (size_t)p * np, np * sizeof(double), hipMemcpyDeviceToHost));
            for (int k = 0; k < np; ++k) traj[(size_t)4 * k + q] = col[k];
        }
        const int stride = std::max(1, np / 1000);
        if (tiling_cost(s, idx2, traj, np, stride) < tiling_cost(s, idx, traj, np, stride)) idx.swap(idx2);
    }
    s.n_tiles = (s.n_pix + kLanes - 1) / kLanes;
    std::vector<int> perm((size_t)s.n_tiles * kLanes, -1);
    for (int i = 0; i < s.n_pix; ++i) perm[i] = idx[i];
    s.tile_prof.assign((size_t)s.n_tiles * nf, 0.0);
    for (int t = 0; t < s.n_tiles; ++t)
        for (int j = 0; j < nf; ++j) s.tile_prof[(size_t)t * nf + j] = feat[(size_t)perm[(size_t)t * kLanes] * nf + j];
    if (s.d_perm) TG_HIP(c, hipFree(s.d_perm));
    s.d_perm = nullptr;
    TG_HIP(c, hipMalloc((void**)&s.d_perm, perm.size() * sizeof(int)));
    TG_HIP(c, hipMemcpy(s.d_perm, perm.data(), perm.size() * sizeof(int), hipMemcpyHostToDevice));
    s.perm_valid = true;
    return TG_OK;
}

int default_window(const tg_context* c, int n_bins)
{
    (void)c;
    return std::max(1, std::min(n_bins, 8));
}

size_t cell_bytes(int prec) { return prec == TG_FP64 ? sizeof(double) : sizeof(float); }

// LDS of one wave: window x 6 components x 64 lanes, plus the FP64
// wave-uniform walk's record chunks (the per-lane walk's spill bucket is held
// in registers, run_item).
size_t stage_bytes(int prec, bool lanes)
{
    if (lanes) return 0;  // the per-lane walk reads records straight from global memory
    return (prec == TG_FP64 ? SourceOf<0>::lds_doubles : SourceOf<1>::lds_doubles) * sizeof(double);
}

size_t lds_per_wave(int prec, bool lanes, int window)
{
    return (size_t)window * 6 * kLanes * cell_bytes(prec) + stage_bytes(prec, lanes);
}

int max_window_for_lds(const tg_context* c, int prec, bool lanes)
{
    return (int)((c->max_lds - stage_bytes(prec, lanes)) / (6 * kLanes * cell_bytes(prec)));
}


// tg_tuning.walk: 1 = wave-uniform walk, 2 = per-lane walk; 0 picks the
// context default (2, or the TG_WALK environment variable at tg_create).
bool lane_walk(const tg_context* c, const tg_tuning* tun)
{
    int w = (tun && tun->walk > 0) ? tun->walk : c->default_walk;
    return w != 1;
}

// The spill bucket applies to the per-lane walk; tg_tuning.spill 2 turns it off.
bool spill_on(const tg_context* c, const tg_tuning* tun)
{
    return lane_walk(c, tun) && !(tun && tun->spill == 2);
}

// Estimated trajectory steps of one particle whose arrival falls into
// buckets [lo, hi) for a tile: its profile is piecewise linear between the
// kProfile sample steps.
double steps_in_buckets(const tg_context* c, const ScreenHost& s, int tile, int lo, int hi)
{
    const double* b = s.tile_prof.data() + (size_t)tile * kProfile;
    double steps = 0.0;
    for (int j = 0; j + 1 < kProfile; ++j) {
        const double k0 = (double)(c->n_points - 1) * j / (kProfile - 1);
        const double k1 = (double)(c->n_points - 1) * (j + 1) / (kProfile - 1);
        const double a = std::min(b[j], b[j + 1]), e = std::max(b[j], b[j + 1]);
        if (e > a) {
            const double ov = std::min<double>(hi, e) - std::max<double>(lo, a);
            if (ov > 0) steps += (k1 - k0) * ov / (e - a);
        } else if (a >= lo && a < hi) {
            steps += k1 - k0;
        }
    }
    return steps;
}

const void* kernel_fn(int prec, bool lanes)
{
    if (prec == TG_FP64) return lanes ? (const void*)radiation_kernel<0, true> : (const void*)radiation_kernel<0, false>;
    return lanes ? (const void*)radiation_kernel<1, true> : (const void*)radiation_kernel<1, false>;
}

// Build the work list for screens `ids`, particles [p0,p1).
int make_plan(tg_context* c, int prec, const std::vector<int>& ids, int p0, int p1, const tg_tuning* tun,
              LaunchPlan& plan)
{
    plan = LaunchPlan();
    int np = p1 - p0;
    long long total_tiles = 0;
    for (int id : ids) total_tiles += c->screens[id].n_tiles;
    if (np <= 0 || total_tiles == 0) return TG_OK;
    int n_chunks;
    if (tun && tun->particles_per_chunk > 0) {
        n_chunks = (np + tun->particles_per_chunk - 1) / tun->particles_per_chunk;
    } else {
        long long target = 64LL * std::max(c->n_cu, 1);
        n_chunks = (int)std::max<long long>(1, (target + total_tiles - 1) / total_tiles);
        n_chunks = std::min(n_chunks, std::max(1, np / 32));
    }
    // cap the partial-trace scratch (chunks 1..; chunk 0 adds into the trace
    // store) at 16 GB (work items get finer through window groups, which cost
    // no memory)
    size_t per_chunk = 0;
    for (int id : ids) per_chunk += (size_t)c->screens[id].n_pix * c->screens[id].n_bins * 6;
    const size_t cap = (size_t)16 << 30;
    while (n_chunks > 1 && (size_t)(n_chunks - 1) * per_chunk * sizeof(double) > cap) --n_chunks;
    n_chunks = std::max(1, std::min(n_chunks, np));
    int chunk = (np + n_chunks - 1) / n_chunks;
    n_chunks = (np + chunk - 1) / chunk;
    plan.n_chunks = n_chunks;
    plan.chunk = chunk;
    plan.screen_ids = ids;
    plan.lanes = lane_walk(c, tun);
    plan.spill = spill_on(c, tun);
    int wmax_lds = max_window_for_lds(c, prec, plan.lanes);
    size_t part_off = 0;
    for (int id : ids) {
        ScreenHost& s = c->screens[id];
        ScreenDev d;
        d.obs = s.d_obs;
        d.t0 = s.d_t0;
        d.perm = s.d_perm;
        d.traces = s.d_traces;
        d.partial = nullptr;  // patched once the scratch is allocated
        d.dt = s.dt;
        d.inv_dt = 1.0 / s.dt;
        d.n_pix = s.n_pix;
        d.n_bins = s.n_bins;
        d.n_tiles = s.n_tiles;
        int w = (tun && tun->bins_per_window > 0) ? tun->bins_per_window
                                                   : (s.window[prec] > 0 ? s.window[prec] : default_window(c, s.n_bins));
        w = std::max(1, std::min({w, s.n_bins, wmax_lds}));
        d.window = w;
        d.spill = plan.spill ? 1 : 0;
        plan.max_window = std::max(plan.max_window, w);
        plan.sdev.push_back(d);
        part_off += (size_t)(n_chunks - 1) * s.n_pix * s.n_bins * 6;
    }
    plan.partial_doubles = part_off;
    // Window groups: each (tile, chunk) is split into groups of consecutive
    // windows so that the work items fill the resident wave slots (limited
    // by the LDS window) about 16 times over, which keeps the tail of the
    // launch short; a group seeds its cursors by search instead of walking
    // from step 0.
    long long base_items = total_tiles * n_chunks;
    const size_t lds_wave = lds_per_wave(prec, plan.lanes, plan.max_window);
    long long slots = (long long)std::max(c->n_cu, 1) * (long long)std::max<size_t>(1, c->max_lds / lds_wave);
    int groups = (int)std::max<long long>(1, std::min<long long>(64, (16 * slots + base_items - 1) / base_items));
    if (tun && tun->window_groups > 0) groups = tun->window_groups;
    plan.groups = groups;
    for (int ch = 0; ch < n_chunks; ++ch) {
        int a = p0 + ch * chunk, b = std::min(p1, a + chunk);
        for (size_t si = 0; si < ids.size(); ++si) {
            const ScreenHost& s = c->screens[ids[si]];
            const int W = plan.sdev[si].window;
            int nw = (s.n_bins + W - 1) / W;
            int g_n = std::min(groups, nw);
            for (int g = 0; g < g_n; ++g) {
                int w_lo = (int)((long long)nw * g / g_n), w_hi = (int)((long long)nw * (g + 1) / g_n);
                for (int t = 0; t < s.n_tiles; ++t) {
                    WorkItem w;
                    w.screen = (int)si;
                    w.tile = t;
                    w.chunk = ch;
                    w.p0 = a;
                    w.p1 = b;
                    w.bin_lo = w_lo * W;
                    w.bin_hi = std::min(s.n_bins, w_hi * W);
                    plan.items.push_back(w);
                }
            }
        }
    }
    // Persistent grid: one wave per resident slot (LDS, registers), each
    // with a cursor slot for one item: chunk int2 cursors and, for the
    // per-lane walk, kLanes bytes per particle holding each lane's stopping
    // step between windows (kept while the store stays under 16 GB; above,
    // lanes find their start from the wave's by skipping).
    plan.lds = lds_wave;
    int per_cu = 0;
    const void* kfn = kernel_fn(prec, plan.lanes);
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kfn, kLanes, lds_wave) != hipSuccess || per_cu < 1) {
        (void)hipGetLastError();
        return fail(c, TG_ERR_HIP, "radiation kernel does not fit a compute unit");
    }
    plan.slots = (int)std::min<long long>((long long)plan.items.size(), (long long)std::max(c->n_cu, 1) * per_cu);
    if (c->max_slots > 0) plan.slots = std::min(plan.slots, c->max_slots);
    auto slot_size = [&](bool offsets) {
        size_t b = (size_t)chunk * sizeof(int2) + (offsets ? (size_t)chunk * kLanes : 0);
        return (b + 255) / 256 * 256;
    };
    plan.lane_cursors = plan.lanes && c->lane_cursors_on && (size_t)plan.slots * slot_size(true) <= ((size_t)16 << 30);
    plan.slot_bytes = slot_size(plan.lane_cursors);
    // Dispatch the heaviest items first (longest-processing-time order), so
    // that the launch does not end on a few long waves.  The order of the
    // items changes no result: each writes its own partial-trace region.
    std::vector<double> cost(plan.items.size());
    for (size_t i = 0; i < plan.items.size(); ++i) {
        const WorkItem& w = plan.items[i];
        const ScreenHost& s = c->screens[ids[w.screen]];
        cost[i] = (w.p1 - w.p0) * (1.0 + steps_in_buckets(c, s, w.tile, w.bin_lo, w.bin_hi));
    }
    std::vector<size_t> order(plan.items.size());
    for (size_t i = 0; i < order.size(); ++i) order[i] = i;
    std::stable_sort(order.begin(), order.end(), [&](size_t a, size_t b) { return cost[a] > cost[b]; });
    std::vector<WorkItem> sorted(plan.items.size());
    for (size_t i = 0; i < order.size(); ++i) sorted[i] = plan.items[order[i]];
    plan.items.swap(sorted);
    return TG_OK;
}

// Allocate scratch and upload the plan; fills plan.sdev[].partial.  A
// captured graph refers to the scratch buffers and the uploaded work list, so
// any upload (which may reallocate them or overwrite the list with another
// plan's) invalidates it.
int upload_plan(tg_context* c, LaunchPlan& plan)
{
    free_graph(c);
    int rc;
    if ((rc = ensure_cap(c, c->d_partial, c->cap_partial, plan.partial_doubles))) return rc;
    if ((rc = ensure_cap(c, c->d_slots, c->cap_slots, (size_t)plan.slots * plan.slot_bytes))) return rc;
    if ((rc = ensure_cap(c, c->d_next_item, c->cap_next_item, 1))) return rc;
    if ((rc = ensure_cap(c, c->d_shift, c->cap_shift, 1))) return rc;
    if ((rc = ensure_cap(c, c->d_items, c->cap_items, plan.items.size()))) return rc;
    if ((rc = ensure_cap(c, c->d_sdev, c->cap_sdev, plan.sdev.size()))) return rc;
    if (c->stats_on && (rc = ensure_cap(c, c->d_stats, c->cap_stats, 2 * plan.items.size()))) return rc;
    size_t off = 0;
    for (size_t i = 0; i < plan.sdev.size(); ++i) {
        plan.sdev[i].partial = c->d_partial + off;
        off += (size_t)(plan.n_chunks - 1) * plan.sdev[i].n_pix * plan.sdev[i].n_bins * 6;
    }
    TG_HIP(c, hipMemcpy(c->d_items, plan.items.data(), plan.items.size() * sizeof(WorkItem), hipMemcpyHostToDevice));
    TG_HIP(c, hipMemcpy(c->d_sdev, plan.sdev.data(), plan.sdev.size() * sizeof(ScreenDev), hipMemcpyHostToDevice));
    return TG_OK;
}

// The buffers one launch of the radiation kernel works in: the work list,
// the work-list counter, the screen descriptors and the per-wave cursor slots.
struct LaunchBuffers {
    const WorkItem* items;
    int* next_item;
    const ScreenDev* sdev;
    unsigned char* slots;
    size_t slot_bytes;
    int lane_offsets;
    unsigned long long* stats;
    const int* p_shift;   // added to the items' particle ranges at launch (null: none)
};

// Queue the radiation kernel of plan on the context's stream, the work-list
// counter zeroed by a kernel first (no memset, so the sequence can be
// captured); events[0..1] around the radiation kernel when events.
int launch_radiation(tg_context* c, int prec, const LaunchPlan& plan, const LaunchBuffers& b, bool events)
{
    set_int_kernel<<<1, 64, 0, c->stream>>>(b.next_item, 0);
    TG_HIP(c, hipGetLastError());
    if (events) TG_HIP(c, hipEventRecord(c->ev[0], c->stream));
    const unsigned grid = (unsigned)plan.slots;
    const int n_items = (int)plan.items.size();
    const RecF64* r64 = (const RecF64*)c->d_rec[0];
    const RecMixed* rmx = (const RecMixed*)c->d_rec[1];
#define TG_LAUNCH(P, L, R)                                                                                        \
    radiation_kernel<P, L><<<grid, kLanes, plan.lds, c->stream>>>(b.items, n_items, b.next_item, b.sdev, R,        \
                                                                  c->n_points, b.slots, b.slot_bytes, plan.chunk, \
                                                                  b.lane_offsets, b.stats, plan.max_window,     \
                                                                  b.p_shift)
    if (prec == TG_FP64 && plan.lanes)
        TG_LAUNCH(0, true, r64);
    else if (prec == TG_FP64)
        TG_LAUNCH(0, false, r64);
    else if (plan.lanes)
        TG_LAUNCH(1, true, rmx);
    else
        TG_LAUNCH(1, false, rmx);
#undef TG_LAUNCH
    TG_HIP(c, hipGetLastError());
    if (events) TG_HIP(c, hipEventRecord(c->ev[1], c->stream));
    return TG_OK;
}

// Queue radiation kernel + chunk reductions (capturable: kernels only).
int enqueue_plan(tg_context* c, int prec, const LaunchPlan& plan, bool events, const int* p_shift = nullptr)
{
    if (plan.items.empty()) return TG_OK;
    const LaunchBuffers b{c->d_items, c->d_next_item, c->d_sdev, c->d_slots, plan.slot_bytes,
                          plan.lane_cursors ? 1 : 0, c->stats_on ? c->d_stats : nullptr, p_shift};
    int rc;
    if ((rc = launch_radiation(c, prec, plan, b, events))) return rc;
    for (size_t i = 0; i < plan.sdev.size() && plan.n_chunks > 1; ++i) {
        const ScreenHost& s = c->screens[plan.screen_ids[i]];
        size_t n = (size_t)s.n_pix * s.n_bins * 6;
        unsigned blocks = (unsigned)std::min<size_t>((n + 255) / 256, 16384);
        reduce_partials_kernel<<<blocks, 256, 0, c->stream>>>(plan.sdev[i].partial, s.d_traces, n, plan.n_chunks - 1);
        TG_HIP(c, hipGetLastError());
    }
    if (events) TG_HIP(c, hipEventRecord(c->ev[2], c->stream));
    return TG_OK;
}

int check_ctx(const tg_context* c) { return c ? TG_OK : TG_ERR_ARG; }

// Argument checks of an integrate call; fills ids with the screens that take
// part (empty: a silent no-op of the reference, nots < 2 or NP < 2,
// particle.cpp:778,783).
int integrate_args(tg_context* c, int prec, int screen_id, int p0, int p1, std::vector<int>& ids)
{
    ids.clear();
    if (prec != TG_FP64 && prec != TG_MIXED) return fail(c, TG_ERR_ARG, "precision must be TG_FP64 or TG_MIXED");
    if (!c->traj_set && !c->screens.empty()) return fail(c, TG_ERR_STATE, "tg_integrate: no trajectories set");
    if (screen_id >= (int)c->screens.size()) return fail(c, TG_ERR_RANGE, "screen id out of range");
    if (p0 < 0 || p1 > c->n_particles || p0 > p1) return fail(c, TG_ERR_RANGE, "particle range out of range");
    c->kernel_ms = c->total_ms = 0.f;
    if (c->n_points < 2 || p0 == p1) return TG_OK;
    for (int i = 0; i < (int)c->screens.size(); ++i)
        if ((screen_id < 0 || i == screen_id) && c->screens[i].n_bins >= 2 && c->screens[i].n_pix > 0)
            ids.push_back(i);
    return TG_OK;
}

// The autotuned shape, with every non-zero field of the caller's tuning
// taking precedence.
tg_tuning effective_tuning(const tg_context* c, int prec, const tg_tuning* tun)
{
    tg_tuning eff = c->has_tuned[prec] ? c->tuned[prec] : tg_tuning{};
    if (tun) {
        if (tun->bins_per_window > 0) eff.bins_per_window = tun->bins_per_window;
        if (tun->particles_per_chunk > 0) eff.particles_per_chunk = tun->particles_per_chunk;
        if (tun->window_groups > 0) eff.window_groups = tun->window_groups;
        if (tun->walk > 0) eff.walk = tun->walk;
        if (tun->spill > 0) eff.spill = tun->spill;
        eff.use_graph = tun->use_graph;
    }
    return eff;
}

int integrate_impl(tg_context* c, int prec, const tg_tuning* tun, int screen_id, int p0, int p1)
{
    std::vector<int> ids;
    int rc;
    if ((rc = integrate_args(c, prec, screen_id, p0, p1, ids))) return rc;
    if (ids.empty()) return TG_OK;
    if ((rc = ensure_records(c, prec))) return rc;
    for (int id : ids)
        if ((rc = ensure_tiles(c, c->screens[id]))) return rc;
    tg_tuning eff = effective_tuning(c, prec, tun);
    bool use_graph = eff.use_graph != 0;
    char keybuf[256];
    // the key holds the size of the particle range, not its start: a captured
    // launch replays over any range of that size (the kernel adds d_shift to
    // its items' particle ranges), as a caller streaming equal batches needs
    snprintf(keybuf, sizeof keybuf, "%d/%d/%d/%d/%d/%d/%d/%d/%zu", prec, eff.bins_per_window,
             eff.particles_per_chunk, eff.window_groups, eff.walk, eff.spill, screen_id, p1 - p0, c->screens.size());
    std::string key(keybuf);
    if (use_graph && c->gexec && c->graph_key == key) {
        set_int_kernel<<<1, 64, 0, c->stream>>>(c->d_shift, p0 - c->graph_p0);
        TG_HIP(c, hipGetLastError());
        TG_HIP(c, hipEventRecord(c->ev[0], c->stream));
        TG_HIP(c, hipGraphLaunch(c->gexec, c->stream));
        TG_HIP(c, hipEventRecord(c->ev[2], c->stream));
        TG_HIP(c, hipEventSynchronize(c->ev[2]));
        float ms = 0.f;
        TG_HIP(c, hipEventElapsedTime(&ms, c->ev[0], c->ev[2]));
        c->kernel_ms = c->total_ms = ms;
        return TG_OK;
    }
    LaunchPlan plan;
    if ((rc = make_plan(c, prec, ids, p0, p1, &eff, plan))) return rc;
    if ((rc = upload_plan(c, plan))) return rc;
    if (use_graph) {
        free_graph(c);
        set_int_kernel<<<1, 64, 0, c->stream>>>(c->d_shift, 0);
        TG_HIP(c, hipGetLastError());
        TG_HIP(c, hipStreamSynchronize(c->stream));
        TG_HIP(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
        rc = enqueue_plan(c, prec, plan, false, c->d_shift);
        hipGraph_t g = nullptr;
        hipError_t e = hipStreamEndCapture(c->stream, &g);
        if (rc) {
            if (g) (void)hipGraphDestroy(g);
            return rc;
        }
        if (e != hipSuccess) return fail(c, TG_ERR_HIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
        c->graph = g;
        TG_HIP(c, hipGraphInstantiate(&c->gexec, c->graph, nullptr, nullptr, 0));
        c->graph_key = key;
        c->graph_p0 = p0;
        TG_HIP(c, hipEventRecord(c->ev[0], c->stream));
        TG_HIP(c, hipGraphLaunch(c->gexec, c->stream));
        TG_HIP(c, hipEventRecord(c->ev[2], c->stream));
        TG_HIP(c, hipEventSynchronize(c->ev[2]));
        float ms = 0.f;
        TG_HIP(c, hipEventElapsedTime(&ms, c->ev[0], c->ev[2]));
        c->kernel_ms = c->total_ms = ms;
        return TG_OK;
    }
    if ((rc = enqueue_plan(c, prec, plan, true))) return rc;
    TG_HIP(c, hipEventSynchronize(c->ev[2]));
    TG_HIP(c, hipEventElapsedTime(&c->kernel_ms, c->ev[0], c->ev[1]));
    TG_HIP(c, hipEventElapsedTime(&c->total_ms, c->ev[0], c->ev[2]));
    if (c->stats_on) {
        std::vector<unsigned long long> st(2 * plan.items.size());
        TG_HIP(c, hipMemcpy(st.data(), c->d_stats, st.size() * sizeof(unsigned long long), hipMemcpyDeviceToHost));
        c->last_steps = c->last_walks = 0;
        for (size_t i = 0; i < plan.items.size(); ++i) {
            c->last_steps += (long long)st[2 * i];
            c->last_walks += (long long)st[2 * i + 1];
        }
        c->last_items = (long long)plan.items.size();
    }
    return TG_OK;
}

}  // namespace

// ---------------------------------------------------------------------------
// Launch scratch without allocation (tg_integrate_range_static).  All of that
// entry point's launch scratch (the work-list counter, the screen
// descriptors, the work list and the per-wave cursor slots) lives in one
// file-scope device array, one copy per device, held until the process ends.
// Host code never hands the array's name to a kernel or a copy: its device
// address on each device comes from hipGetSymbolAddress (static_scratch_of).
// ---------------------------------------------------------------------------
static __device__ __align__(256) unsigned char scratch[TG_STATIC_SCRATCH_BYTES];

namespace {

std::mutex g_scratch_mu;
std::vector<void*> g_scratch_ptr;      // device address of `scratch`, per device
std::vector<std::string> g_scratch_err;

// This device's copy of the array, looked up once per device with the device
// current (tg_create's set-up); nullptr with the reason in err if the lookup
// failed.
void* static_scratch_of(int device, std::string& err)
{
    std::lock_guard<std::mutex> g(g_scratch_mu);
    if ((int)g_scratch_ptr.size() <= device) {
        g_scratch_ptr.resize(device + 1, nullptr);
        g_scratch_err.resize(device + 1);
    }
    if (!g_scratch_ptr[device] && g_scratch_err[device].empty()) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            (void)hipGetLastError();
            g_scratch_err[device] = std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e);
        } else {
            g_scratch_ptr[device] = p;
        }
    }
    err = g_scratch_err[device];
    return g_scratch_ptr[device];
}

size_t align256(size_t n) { return (n + 255) / 256 * 256; }

// tg_integrate_range_static: one particle chunk, all scratch in the array.
int integrate_static(tg_context* c, int prec, const tg_tuning* tun, int screen_id, int p0, int p1)
{
    const char* who = "tg_integrate_range_static: ";
    std::vector<int> ids;
    int rc;
    if ((rc = integrate_args(c, prec, screen_id, p0, p1, ids))) return rc;
    if (ids.empty()) return TG_OK;
    if (!c->static_scratch) return fail(c, TG_ERR_HIP, who + c->static_scratch_err);
    // set-up allocates; this entry point only uses what set-up made
    if (!c->rec_valid[prec])
        return fail(c, TG_ERR_STATE,
                    std::string(who) + "the trajectory records of this precision are not packed yet (the first "
                                       "tg_integrate, tg_integrate_range or tg_autotune of the precision packs them)");
    for (int id : ids)
        if (!c->screens[id].perm_valid)
            return fail(c, TG_ERR_STATE,
                        std::string(who) + "screen " + std::to_string(id) +
                            " is not tiled yet (the first tg_integrate, tg_integrate_range or tg_autotune tiles it)");
    tg_tuning eff = effective_tuning(c, prec, tun);
    eff.particles_per_chunk = p1 - p0;   // one chunk: fields straight into the trace store
    eff.use_graph = 0;
    LaunchPlan plan;
    if ((rc = make_plan(c, prec, ids, p0, p1, &eff, plan))) return rc;
    if (plan.items.empty()) return TG_OK;
    if (plan.n_chunks != 1) return fail(c, TG_ERR_STATE, std::string(who) + "plan with more than one chunk");
    // layout: counter | screen descriptors | work list | cursor slots
    const size_t off_sdev = 256;
    const size_t off_items = align256(off_sdev + plan.sdev.size() * sizeof(ScreenDev));
    const size_t off_slots = align256(off_items + plan.items.size() * sizeof(WorkItem));
    auto slot_size = [&](bool offsets) {
        return align256((size_t)plan.chunk * sizeof(int2) + (offsets ? (size_t)plan.chunk * kLanes : 0));
    };
    bool offsets = plan.lanes && c->lane_cursors_on;
    if (off_slots + (size_t)plan.slots * slot_size(offsets) > (size_t)TG_STATIC_SCRATCH_BYTES) offsets = false;
    const size_t need = off_slots + (size_t)plan.slots * slot_size(offsets);
    if (need > (size_t)TG_STATIC_SCRATCH_BYTES) {
        const size_t room = (size_t)TG_STATIC_SCRATCH_BYTES > off_slots ? (size_t)TG_STATIC_SCRATCH_BYTES - off_slots : 0;
        const long long fit = (long long)(room / std::max(1, plan.slots) / sizeof(int2));
        return fail(c, TG_ERR_RANGE,
                    std::string(who) + "the launch needs " + std::to_string(need) + " bytes of scratch, the static " +
                        "array holds " + std::to_string((size_t)TG_STATIC_SCRATCH_BYTES) + "; integrate at most about " +
                        std::to_string(std::max(0LL, fit - 8)) + " particles per call");
    }
    unsigned char* base = static_cast<unsigned char*>(c->static_scratch);
    for (ScreenDev& d : plan.sdev) d.partial = nullptr;   // chunk 0 only: no partial traces
    TG_HIP(c, hipMemcpyAsync(base + off_sdev, plan.sdev.data(), plan.sdev.size() * sizeof(ScreenDev),
                             hipMemcpyHostToDevice, c->stream));
    TG_HIP(c, hipMemcpyAsync(base + off_items, plan.items.data(), plan.items.size() * sizeof(WorkItem),
                             hipMemcpyHostToDevice, c->stream));
    const LaunchBuffers b{reinterpret_cast<const WorkItem*>(base + off_items), reinterpret_cast<int*>(base),
                          reinterpret_cast<const ScreenDev*>(base + off_sdev), base + off_slots, slot_size(offsets),
                          offsets ? 1 : 0, nullptr, nullptr};
    if ((rc = launch_radiation(c, prec, plan, b, true))) return rc;
    TG_HIP(c, hipEventRecord(c->ev[2], c->stream));
    TG_HIP(c, hipEventSynchronize(c->ev[2]));   // the host copies' sources live until here
    TG_HIP(c, hipEventElapsedTime(&c->kernel_ms, c->ev[0], c->ev[1]));
    TG_HIP(c, hipEventElapsedTime(&c->total_ms, c->ev[0], c->ev[2]));
    return TG_OK;
}

int set_traj_common(tg_context* c, int n_particles, int n_points, const double* const* src, const double* charge,
                    hipMemcpyKind kind)
{
    if (n_particles < 0 || n_points < 0) return fail(c, TG_ERR_ARG, "negative trajectory size");
    if ((long long)rec_particle_bytes<RecF64>(n_points) > (long long)INT_MAX)
        return fail(c, TG_ERR_ARG, "more than 22 million points per particle");
    for (int i = 0; i < kNumRaw; ++i)
        if (!src[i] && (size_t)n_particles * n_points > 0) return fail(c, TG_ERR_ARG, "null trajectory array");
    if (!charge && n_particles > 0) return fail(c, TG_ERR_ARG, "null charge array");
    size_t n = (size_t)n_particles * n_points;
    if (kind == hipMemcpyHostToDevice) {
        for (int i = 0; i < kNumRaw; ++i)
            if (n && !finite_all(src[i], n)) return fail(c, TG_ERR_ARG, "non-finite trajectory value");
        if (n_particles && !finite_all(charge, n_particles)) return fail(c, TG_ERR_ARG, "non-finite charge");
    }
    int rc = alloc_traj(c, n_particles, n_points);
    if (rc) return rc;
    for (int i = 0; i < kNumRaw; ++i)
        if (n) TG_HIP(c, hipMemcpy(raw(c, i), src[i], n * sizeof(double), kind));
    if (n_particles) TG_HIP(c, hipMemcpy(raw_charge(c), charge, n_particles * sizeof(double), kind));
    return TG_OK;
}

}  // namespace

// ============================== C ABI ======================================

extern "C" {

int tg_create(int device, tg_context** out)
{
    if (!out) return TG_ERR_ARG;
    *out = nullptr;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n == 0) return TG_ERR_HIP;
    if (device < 0 || device >= n) return TG_ERR_ARG;
    tg_context* c = new (std::nothrow) tg_context;
    if (!c) return TG_ERR_OOM;
    c->device = device;
    auto bad = [&](int rc) {
        tg_destroy(c);
        return rc;
    };
    if (hipSetDevice(device) != hipSuccess) return bad(TG_ERR_HIP);
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) return bad(TG_ERR_HIP);
    for (auto& e : c->ev)
        if (hipEventCreate(&e) != hipSuccess) return bad(TG_ERR_HIP);
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) != hipSuccess) return bad(TG_ERR_HIP);
    snprintf(c->name, sizeof c->name, "%s %s", prop.name, prop.gcnArchName);
    c->n_cu = prop.multiProcessorCount;
    c->max_lds = prop.sharedMemPerBlock;
    const char* st = getenv("TG_STATS");
    c->stats_on = st && st[0] == '1';
    const char* wk = getenv("TG_WALK");
    c->default_walk = (wk && wk[0] == '1') ? 1 : 2;
    const char* lcv = getenv("TG_LANE_CURSORS");
    c->lane_cursors_on = !(lcv && lcv[0] == '0');
    const char* sl = getenv("TG_SLOTS");
    c->max_slots = sl ? std::max(0, atoi(sl)) : 0;
    const char* ft = getenv("TG_AUTOTUNE_FAIL_TRIAL");
    c->fail_trial = ft ? std::max(0, atoi(ft)) : 0;
    const void* kernels[] = {(const void*)radiation_kernel<0, false>, (const void*)radiation_kernel<1, false>,
                             (const void*)radiation_kernel<0, true>, (const void*)radiation_kernel<1, true>};
    for (const void* k : kernels)
        if (hipFuncSetAttribute(k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)c->max_lds) != hipSuccess)
            (void)hipGetLastError();
    c->static_scratch = static_scratch_of(device, c->static_scratch_err);
    *out = c;
    return TG_OK;
}

int tg_destroy(tg_context* c)
{
    if (!c) return TG_ERR_ARG;
    (void)hipSetDevice(c->device);
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    free_graph(c);
    (void)free_traj(c);
    for (auto& s : c->screens) {
        (void)hipFree(s.d_obs);
        (void)hipFree(s.d_t0);
        (void)hipFree(s.d_traces);
        (void)hipFree(s.d_perm);
    }
    (void)hipFree(c->d_items);
    (void)hipFree(c->d_sdev);
    (void)hipFree(c->d_slots);
    (void)hipFree(c->d_next_item);
    (void)hipFree(c->d_shift);
    (void)hipFree(c->d_partial);
    (void)hipFree(c->d_stats);
    (void)hipFree(c->d_flag);
    for (auto& e : c->ev)
        if (e) (void)hipEventDestroy(e);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
    return TG_OK;
}

const char* tg_last_error(const tg_context* c) { return c ? c->err.c_str() : "null context"; }

int tg_device_info(const tg_context* c, char* name, int name_len, int* n_cu)
{
    if (!c) return TG_ERR_ARG;
    if (name && name_len > 0) snprintf(name, (size_t)name_len, "%s", c->name);
    if (n_cu) *n_cu = c->n_cu;
    return TG_OK;
}

int tg_set_trajectories(tg_context* c, int n_particles, int n_points, const double* t, const double* x,
                        const double* y, const double* z, const double* ux, const double* uy, const double* uz,
                        const double* ax, const double* ay, const double* az, const double* charge)
{
    if (check_ctx(c)) return TG_ERR_ARG;
    TG_HIP(c, hipSetDevice(c->device));
    const double* src[kNumRaw] = {t, x, y, z, ux, uy, uz, ax, ay, az};
    return set_traj_common(c, n_particles, n_points, src, charge, hipMemcpyHostToDevice);
}

int tg_set_trajectories_device(tg_context* c, int n_particles, int n_points, const double* t, const double* x,
                               const double* y, const double* z, const double* ux, const double* uy,
                               const double* uz, const double* ax, const double* ay, const double* az,
                               const double* charge)
{
    if (check_ctx(c)) return TG_ERR_ARG;
    TG_HIP(c, hipSetDevice(c->device));
    const double* src[kNumRaw] = {t, x, y, z, ux, uy, uz, ax, ay, az};
    return set_traj_common(c, n_particles, n_points, src, charge, hipMemcpyDeviceToDevice);
}

int tg_trajectory_shape(const tg_context* c, int* n_particles, int* n_points)
{
    if (!c) return TG_ERR_ARG;
    if (n_particles) *n_particles = c->n_particles;
    if (n_points) *n_points = c->n_points;
    return TG_OK;
}

int tg_get_trajectories(tg_context* c, int p_begin, int p_end, double* t, double* x, double* y, double* z,
                        double* ux, double* uy, double* uz, double* ax, double* ay, double* az, double* charge)
{
    if (check_ctx(c)) return TG_ERR_ARG;
    if (p_begin < 0 || p_end > c->n_particles || p_begin > p_end) return fail(c, TG_ERR_RANGE, "particle range");
    TG_HIP(c, hipSetDevice(c->device));
    TG_HIP(c, hipStreamSynchronize(c->stream));
    double* dst[kNumRaw] = {t, x, y, z, ux, uy, uz, ax, ay, az};
    size_t off = (size_t)p_begin * c->n_points, n = (size_t)(p_end - p_begin) * c->n_points;
    for (int i = 0; i < kNumRaw; ++i)
        if (dst[i] && n) TG_HIP(c, hipMemcpy(dst[i], raw(c, i) + off, n * sizeof(double), hipMemcpyDeviceToHost));
    if (charge && p_end > p_begin)
        TG_HIP(c, hipMemcpy(charge, raw_charge(c) + p_begin, (p_end - p_begin) * sizeof(double),
                            hipMemcpyDeviceToHost));
    return TG_OK;
}

int tg_add_screen(tg_context* c, int n_pix, const double* obs, const double* t0, double dt, int n_bins,
                  int* screen_id)
{
    if (check_ctx(c)) return TG_ERR_ARG;
    if (n_pix < 0 || n_bins < 0) return fail(c, TG_ERR_ARG, "negative screen size");
    if (n_pix > 0 && (!obs || !t0)) return fail(c, TG_ERR_ARG, "null screen array");
    if (!(dt > 0.0) || !std::isfinite(dt)) return fail(c, TG_ERR_ARG, "dt must be positive and finite");
    if (n_pix > 0 && (!finite_all(obs, (size_t)3 * n_pix) || !finite_all(t0, (size_t)n_pix)))
        return fail(c, TG_ERR_ARG, "non-finite screen geometry");
    TG_HIP(c, hipSetDevice(c->device));
    ScreenHost s;
    s.n_pix = n_pix;
    s.n_bins = n_bins;
    s.dt = dt;
    s.obs.assign(obs, obs + (size_t)3 * n_pix);
    s.t0.assign(t0, t0 + (size_t)n_pix);
    size_t nt = (size_t)n_pix * n_bins * 6;
    // the screen's device arrays, released again if any step of the set-up fails
    auto setup = [&]() -> int {
        TG_HIP(c, hipMalloc((void**)&s.d_obs, std::max<size_t>(1, (size_t)3 * n_pix) * sizeof(double)));
        TG_HIP(c, hipMalloc((void**)&s.d_t0, std::max<size_t>(1, (size_t)n_pix) * sizeof(double)));
        TG_HIP(c, hipMalloc((void**)&s.d_traces, std::max<size_t>(1, nt) * sizeof(double)));
        if (n_pix) {
            TG_HIP(c, hipMemcpy(s.d_obs, obs, (size_t)3 * n_pix * sizeof(double), hipMemcpyHostToDevice));
            TG_HIP(c, hipMemcpy(s.d_t0, t0, (size_t)n_pix * sizeof(double), hipMemcpyHostToDevice));
        }
        if (nt) {
            fill_kernel<<<(unsigned)std::min<size_t>((nt + 255) / 256, 16384), 256, 0, c->stream>>>(s.d_traces, nt,
                                                                                                      0.0);
            TG_HIP(c, hipGetLastError());
            TG_HIP(c, hipStreamSynchronize(c->stream));
        }
        return TG_OK;
    };
    if (const int rc = setup()) {
        for (double* d : {s.d_obs, s.d_t0, s.d_traces})
            if (d) (void)hipFree(d);
        return rc;
    }
    c->screens.push_back(std::move(s));
    free_graph(c);
    if (screen_id) *screen_id = (int)c->screens.size() - 1;
    return TG_OK;
}

int tg_zero_screen(tg_context* c, int id)
{
    if (check_ctx(c)) return TG_ERR_ARG;
    if (id < 0 || id >= (int)c->screens.size()) return fail(c, TG_ERR_RANGE, "screen id out of range");
    TG_HIP(c, hipSetDevice(c->device));
    ScreenHost& s = c->screens[id];
    size_t nt = (size_t)s.n_pix * s.n_bins * 6;
    if (nt) {
        fill_kernel<<<(unsigned)std::min<size_t>((nt + 255) / 256, 16384), 256, 0, c->stream>>>(s.d_traces, nt, 0.0);
        TG_HIP(c, hipGetLastError());
    }
    return TG_OK;
}

int tg_remove_screens(tg_context* c)
{
    if (check_ctx(c)) return TG_ERR_ARG;
    TG_HIP(c, hipSetDevice(c->device));
    TG_HIP(c, hipStreamSynchronize(c->stream));
    for (auto& s : c->screens) {
        TG_HIP(c, hipFree(s.d_obs));
        TG_HIP(c, hipFree(s.d_t0));
        TG_HIP(c, hipFree(s.d_traces));
        if (s.d_perm) TG_HIP(c, hipFree(s.d_perm));
    }
    c->screens.clear();
    free_graph(c);
    return TG_OK;
}

int tg_screen_count(const tg_context* c, int* n)
{
    if (!c || !n) return TG_ERR_ARG;
    *n = (int)c->screens.size();
    return TG_OK;
}

int tg_integrate(tg_context* c, int precision, const tg_tuning* tuning)
{
    if (check_ctx(c)) return TG_ERR_ARG;
    TG_HIP(c, hipSetDevice(c->device));
    return integrate_impl(c, precision, tuning, -1, 0, c->n_particles);
}

int tg_integrate_range(tg_context* c, int precision, const tg_tuning* tuning, int screen_id, int p_begin, int p_end)
{
    if (check_ctx(c)) return TG_ERR_ARG;
    TG_HIP(c, hipSetDevice(c->device));
    return integrate_impl(c, precision, tuning, screen_id, p_begin, p_end);
}

int tg_integrate_range_static(tg_context* c, int precision, const tg_tuning* tuning, int screen_id, int p_begin,
                              int p_end)
{
    if (check_ctx(c)) return TG_ERR_ARG;
    TG_HIP(c, hipSetDevice(c->device));
    return integrate_static(c, precision, tuning, screen_id, p_begin, p_end);
}

int tg_synchronize(tg_context* c)
{
    if (check_ctx(c)) return TG_ERR_ARG;
    TG_HIP(c, hipSetDevice(c->device));
    TG_HIP(c, hipStreamSynchronize(c->stream));
    return TG_OK;
}

int tg_read_traces(tg_context* c, int id, double* out, int accumulate)
{
    if (check_ctx(c)) return TG_ERR_ARG;
    if (id < 0 || id >= (int)c->screens.size()) return fail(c, TG_ERR_RANGE, "screen id out of range");
    ScreenHost& s = c->screens[id];
    size_t nt = (size_t)s.n_pix * s.n_bins * 6;
    if (nt && !out) return fail(c, TG_ERR_ARG, "null output buffer");
    TG_HIP(c, hipSetDevice(c->device));
    TG_HIP(c, hipStreamSynchronize(c->stream));
    if (!nt) return TG_OK;
    if (!accumulate) {
        TG_HIP(c, hipMemcpy(out, s.d_traces, nt * sizeof(double), hipMemcpyDeviceToHost));
        return TG_OK;
    }
    std::vector<double> tmp;
    try {
        tmp.resize(nt);
    } catch (...) {
        return fail(c, TG_ERR_OOM, "host buffer for accumulate");
    }
    TG_HIP(c, hipMemcpy(tmp.data(), s.d_traces, nt * sizeof(double), hipMemcpyDeviceToHost));
    for (size_t i = 0; i < nt; ++i) out[i] += tmp[i];
    return TG_OK;
}

int tg_screen_traces_device(tg_context* c, int id, double** traces)
{
    if (check_ctx(c) || !traces) return TG_ERR_ARG;
    if (id < 0 || id >= (int)c->screens.size()) return fail(c, TG_ERR_RANGE, "screen id out of range");
    *traces = c->screens[id].d_traces;
    return TG_OK;
}

int tg_screen_poynting(tg_context* c, int id, double* out)
{
    if (check_ctx(c)) return TG_ERR_ARG;
    if (id < 0 || id >= (int)c->screens.size()) return fail(c, TG_ERR_RANGE, "screen id out of range");
    ScreenHost& s = c->screens[id];
    if (s.n_pix == 0) return TG_OK;
    if (!out) return fail(c, TG_ERR_ARG, "null output buffer");
    TG_HIP(c, hipSetDevice(c->device));
    double* d = nullptr;
    TG_HIP(c, hipMalloc((void**)&d, (size_t)3 * s.n_pix * sizeof(double)));
    poynting_kernel<<<(s.n_pix + 63) / 64, 64, 0, c->stream>>>(s.d_traces, s.n_pix, s.n_bins, s.dt, d);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipMemcpyAsync(out, d, (size_t)3 * s.n_pix * sizeof(double), hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    (void)hipFree(d);
    if (e != hipSuccess) return fail(c, TG_ERR_HIP, std::string("tg_screen_poynting: ") + hipGetErrorString(e));
    return TG_OK;
}

int tg_count_in_window(tg_context* c, int screen_id, long long* n_inwin)
{
    if (check_ctx(c)) return TG_ERR_ARG;
    if (!n_inwin) return fail(c, TG_ERR_ARG, "null output");
    if (screen_id >= (int)c->screens.size()) return fail(c, TG_ERR_RANGE, "screen id out of range");
    *n_inwin = 0;
    if (c->n_points < 2 || c->n_particles == 0) return TG_OK;  // particle.cpp:783
    if (!c->traj_set) return fail(c, TG_ERR_STATE, "tg_count_in_window: no trajectories set");
    TG_HIP(c, hipSetDevice(c->device));
    unsigned long long* d = nullptr;
    TG_HIP(c, hipMalloc((void**)&d, sizeof(unsigned long long)));
    hipError_t e = hipMemsetAsync(d, 0, sizeof(unsigned long long), c->stream);
    for (int id = 0; id < (int)c->screens.size() && e == hipSuccess; ++id) {
        const ScreenHost& s = c->screens[id];
        if ((screen_id >= 0 && id != screen_id) || s.n_bins < 2 || s.n_pix == 0) continue;  // particle.cpp:778
        const size_t n = (size_t)s.n_pix * c->n_particles;
        inwin_count_kernel<<<(unsigned)((n + 255) / 256), 256, 0, c->stream>>>(
            s.d_obs, s.d_t0, s.n_pix, s.dt, 1.0 / s.dt, s.n_bins, raw(c, 0), raw(c, 1), raw(c, 2), raw(c, 3),
            c->n_particles, c->n_points, d);
        e = hipGetLastError();
    }
    unsigned long long h = 0;
    if (e == hipSuccess) e = hipMemcpyAsync(&h, d, sizeof h, hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    (void)hipFree(d);
    if (e != hipSuccess) return fail(c, TG_ERR_HIP, std::string("tg_count_in_window: ") + hipGetErrorString(e));
    *n_inwin = (long long)h;
    return TG_OK;
}

int tg_screen_spectra(tg_context* c, int id, const double* ex, const double* ey, double* freq, double* dens_h,
                      double* dens_v)
{
    if (check_ctx(c)) return TG_ERR_ARG;
    if (id < 0 || id >= (int)c->screens.size()) return fail(c, TG_ERR_RANGE, "screen id out of range");
    if (!ex || !ey) return fail(c, TG_ERR_ARG, "null screen axis");
    const ScreenHost& s = c->screens[id];
    const int nb = s.n_bins, nh = s.n_bins / 2;
    if (freq)  // np.linspace(0, 1/dt, nb)[:nb//2] (scripts/teufel.py:204-206)
        for (int i = 0; i < nh; ++i) freq[i] = i * ((1.0 / s.dt) / (nb - 1));
    if (s.n_pix == 0 || nh == 0) return TG_OK;
    if (!dens_h || !dens_v) return fail(c, TG_ERR_ARG, "null output buffer");
    TG_HIP(c, hipSetDevice(c->device));
    static std::once_flag fft_once;
    std::call_once(fft_once, [] { rocfft_setup(); });
    const size_t n = (size_t)s.n_pix * nb, nc = (size_t)nb / 2 + 1, nout = (size_t)s.n_pix * nh;
    double *sig = nullptr, *dh = nullptr, *dv = nullptr;
    double2* spec = nullptr;
    void* work = nullptr;
    rocfft_plan plan = nullptr;
    rocfft_execution_info info = nullptr;
    std::string what;
    hipError_t e = hipMalloc((void**)&sig, 4 * n * sizeof(double));
    if (e == hipSuccess) e = hipMalloc((void**)&spec, 4 * (size_t)s.n_pix * nc * sizeof(double2));
    if (e == hipSuccess) e = hipMalloc((void**)&dh, nout * sizeof(double));
    if (e == hipSuccess) e = hipMalloc((void**)&dv, nout * sizeof(double));
    if (e == hipSuccess) {
        const unsigned blocks = (unsigned)std::min<size_t>((n + 255) / 256, 16384);
        inplane_kernel<<<blocks, 256, 0, c->stream>>>(s.d_traces, n, ex[0], ex[1], ex[2], ey[0], ey[1], ey[2], sig);
        e = hipGetLastError();
    }
    // one-sided transforms of the 4 n_pix real signals (batched, out of place)
    const size_t len = (size_t)nb;
    if (e == hipSuccess) {
        size_t wsz = 0;
        if (rocfft_plan_create(&plan, rocfft_placement_notinplace, rocfft_transform_type_real_forward,
                               rocfft_precision_double, 1, &len, 4 * (size_t)s.n_pix, nullptr) != rocfft_status_success ||
            rocfft_plan_get_work_buffer_size(plan, &wsz) != rocfft_status_success ||
            rocfft_execution_info_create(&info) != rocfft_status_success ||
            rocfft_execution_info_set_stream(info, c->stream) != rocfft_status_success)
            what = "rocFFT plan";
        else if (wsz && (e = hipMalloc(&work, wsz)) == hipSuccess &&
                 rocfft_execution_info_set_work_buffer(info, work, wsz) != rocfft_status_success)
            what = "rocFFT work buffer";
        if (what.empty() && e == hipSuccess) {
            void* in[1] = {sig};
            void* out[1] = {spec};
            if (rocfft_execute(plan, in, out, info) != rocfft_status_success) what = "rocFFT execute";
        }
    }
    if (e == hipSuccess && what.empty()) {
        const double scale = 2.0 * s.dt * s.dt / kMuNull;
        const unsigned blocks = (unsigned)std::min<size_t>((nout + 255) / 256, 16384);
        spectral_density_kernel<<<blocks, 256, 0, c->stream>>>(spec, s.n_pix, nb, scale, dh, dv);
        e = hipGetLastError();
        if (e == hipSuccess) e = hipMemcpyAsync(dens_h, dh, nout * sizeof(double), hipMemcpyDeviceToHost, c->stream);
        if (e == hipSuccess) e = hipMemcpyAsync(dens_v, dv, nout * sizeof(double), hipMemcpyDeviceToHost, c->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    }
    if (info) rocfft_execution_info_destroy(info);
    if (plan) rocfft_plan_destroy(plan);
    (void)hipFree(work);
    (void)hipFree(sig);
    (void)hipFree(spec);
    (void)hipFree(dh);
    (void)hipFree(dv);
    if (!what.empty()) return fail(c, TG_ERR_HIP, "tg_screen_spectra: " + what + " failed");
    if (e != hipSuccess) return fail(c, TG_ERR_HIP, std::string("tg_screen_spectra: ") + hipGetErrorString(e));
    return TG_OK;
}

int tg_deposit_only(tg_context* c, int n_pseudo, int n_points, const double* ts, const double* fields, double t0,
                    double dt, int n_bins, double* trace)
{
    if (check_ctx(c)) return TG_ERR_ARG;
    if (n_pseudo < 0 || n_points < 0 || n_bins < 0) return fail(c, TG_ERR_ARG, "negative size");
    if (!(dt > 0.0) || !std::isfinite(dt) || !std::isfinite(t0)) return fail(c, TG_ERR_ARG, "bad t0/dt");
    if (n_bins < 2 || n_points < 2 || n_pseudo == 0) return TG_OK;  // particle.cpp:778,783
    if (!ts || !fields || !trace) return fail(c, TG_ERR_ARG, "null pointer");
    TG_HIP(c, hipSetDevice(c->device));
    size_t nts = (size_t)n_pseudo * n_points, nf = nts * 6, ntr = (size_t)n_bins * 6;
    double *dts = nullptr, *df = nullptr, *dtr = nullptr;
    hipError_t e = hipMalloc((void**)&dts, nts * sizeof(double));
    if (e == hipSuccess) e = hipMalloc((void**)&df, nf * sizeof(double));
    if (e == hipSuccess) e = hipMalloc((void**)&dtr, ntr * sizeof(double));
    if (e == hipSuccess) e = hipMemcpy(dts, ts, nts * sizeof(double), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(df, fields, nf * sizeof(double), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(dtr, trace, ntr * sizeof(double), hipMemcpyHostToDevice);
    if (e == hipSuccess) {
        deposit_only_kernel<<<1, 64, 0, c->stream>>>(n_pseudo, n_points, dts, df, t0, dt, 1.0 / dt, n_bins, dtr);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    if (e == hipSuccess) e = hipMemcpy(trace, dtr, ntr * sizeof(double), hipMemcpyDeviceToHost);
    (void)hipFree(dts);
    (void)hipFree(df);
    (void)hipFree(dtr);
    if (e != hipSuccess)
        return fail(c, e == hipErrorOutOfMemory ? TG_ERR_OOM : TG_ERR_HIP,
                    std::string("tg_deposit_only: ") + hipGetErrorString(e));
    return TG_OK;
}

int tg_fields_at(tg_context* c, int n_obs, const double* obs_time, const double* obs_pos, int p_begin, int p_end,
                 double* fields, int accumulate)
{
    if (check_ctx(c)) return TG_ERR_ARG;
    if (n_obs < 0) return fail(c, TG_ERR_ARG, "tg_fields_at: negative n_obs");
    if (n_obs > 0 && (!obs_time || !obs_pos || !fields)) return fail(c, TG_ERR_ARG, "tg_fields_at: null array");
    if (!c->traj_set || c->n_points < 1)
        return fail(c, TG_ERR_STATE, "tg_fields_at: no trajectory (RetardedField needs at least one point)");
    if (p_begin < 0 || p_end > c->n_particles || p_begin > p_end) return fail(c, TG_ERR_RANGE, "particle range");
    if (n_obs == 0) return TG_OK;
    if (!finite_all(obs_time, (size_t)n_obs) || !finite_all(obs_pos, 3 * (size_t)n_obs))
        return fail(c, TG_ERR_ARG, "tg_fields_at: non-finite observation");
    TG_HIP(c, hipSetDevice(c->device));
    double *d_t = nullptr, *d_x = nullptr, *d_f = nullptr;
    auto cleanup = [&]() {
        if (d_t) (void)hipFree(d_t);
        if (d_x) (void)hipFree(d_x);
        if (d_f) (void)hipFree(d_f);
    };
    const size_t n = (size_t)n_obs;
    if (hipMalloc((void**)&d_t, n * sizeof(double)) != hipSuccess ||
        hipMalloc((void**)&d_x, 3 * n * sizeof(double)) != hipSuccess ||
        hipMalloc((void**)&d_f, 6 * n * sizeof(double)) != hipSuccess) {
        cleanup();
        return fail(c, TG_ERR_OOM, "tg_fields_at: device allocation failed");
    }
    hipError_t e = hipMemcpy(d_t, obs_time, n * sizeof(double), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(d_x, obs_pos, 3 * n * sizeof(double), hipMemcpyHostToDevice);
    if (e == hipSuccess && accumulate) e = hipMemcpy(d_f, fields, 6 * n * sizeof(double), hipMemcpyHostToDevice);
    if (e == hipSuccess) {
        snap::Traj tr{raw(c, 0), raw(c, 1), raw(c, 2), raw(c, 3), raw(c, 4),
                      raw(c, 5), raw(c, 6), raw(c, 7), raw(c, 8), raw(c, 9)};
        field_at_kernel<<<(unsigned)((n + 63) / 64), 64, 0, c->stream>>>(
            n_obs, d_t, d_x, tr, raw_charge(c), c->n_points, p_begin, p_end, d_f, accumulate);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    if (e == hipSuccess) e = hipMemcpy(fields, d_f, 6 * n * sizeof(double), hipMemcpyDeviceToHost);
    cleanup();
    if (e != hipSuccess) return fail(c, TG_ERR_HIP, std::string("tg_fields_at: ") + hipGetErrorString(e));
    return TG_OK;
}

int tg_last_timing(tg_context* c, float* kernel_ms, float* total_ms)
{
    if (check_ctx(c)) return TG_ERR_ARG;
    if (kernel_ms) *kernel_ms = c->kernel_ms;
    if (total_ms) *total_ms = c->total_ms;
    return TG_OK;
}

int tg_last_stats(tg_context* c, long long* wave_steps, long long* walks, long long* work_items)
{
    if (check_ctx(c)) return TG_ERR_ARG;
    if (wave_steps) *wave_steps = c->last_steps;
    if (walks) *walks = c->last_walks;
    if (work_items) *work_items = c->last_items;
    return TG_OK;
}

int tg_autotune(tg_context* c, int precision, tg_tuning* out)
{
    if (check_ctx(c)) return TG_ERR_ARG;
    if (precision != TG_FP64 && precision != TG_MIXED) return fail(c, TG_ERR_ARG, "precision");
    TG_HIP(c, hipSetDevice(c->device));
    tg_tuning best = {};
    if (c->n_particles == 0 || c->n_points < 2 || c->screens.empty()) {
        c->tuned[precision] = best;
        c->has_tuned[precision] = true;
        if (out) *out = best;
        return TG_OK;
    }
    // Time each window size on a particle sample of all screens together, as
    // they are launched: every wave of a launch gets the LDS of the largest
    // window, so one window for the whole batch is what occupancy depends on.
    // The sample is sized to about 1e11 (pixel, step) pairs per candidate
    // (under a second).  It runs with the window groups of the full launch
    // and is cut into as many particle chunks as the full launch, so it has
    // the full launch's number of work items per resident wave (which sets
    // how long the launch's tail is); only the items' particle count is
    // smaller.
    const int cands[] = {2, 3, 4, 5, 6, 7, 8, 9, 10, 12, 16};
    int wmax = max_window_for_lds(c, precision, lane_walk(c, nullptr));
    double pairs_per_particle = 0;
    for (const ScreenHost& s : c->screens) pairs_per_particle += (double)s.n_pix * (c->n_points - 1);
    const double want = 1e11 / std::max(1.0, pairs_per_particle);
    int sample = (int)std::min<double>(c->n_particles, std::max(64.0, want));
    // particles per chunk of the sample when the full launch has n_ch chunks
    auto sample_chunk = [&](long long n_ch) { return (int)std::max<long long>(1, (sample + n_ch - 1) / n_ch); };
    std::vector<int> ids;
    for (int i = 0; i < (int)c->screens.size(); ++i)
        if (c->screens[i].n_bins >= 2 && c->screens[i].n_pix > 0) ids.push_back(i);
    int rc;
    if ((rc = ensure_records(c, precision))) return rc;
    for (int id : ids)
        if ((rc = ensure_tiles(c, c->screens[id]))) return rc;
    int max_bins = 0;
    std::vector<std::vector<double>> saved(c->screens.size());
    for (size_t id = 0; id < c->screens.size(); ++id) {
        ScreenHost& s = c->screens[id];
        max_bins = std::max(max_bins, s.n_bins);
        if (s.n_bins < 2 || s.n_pix == 0) continue;
        size_t nt = (size_t)s.n_pix * s.n_bins * 6;
        saved[id].resize(nt);
        TG_HIP(c, hipMemcpy(saved[id].data(), s.d_traces, nt * sizeof(double), hipMemcpyDeviceToHost));
    }
    // TG_AUTOTUNE_LOG=1 prints every trial's kernel time to stderr
    const char* log_env = std::getenv("TG_AUTOTUNE_LOG");
    const bool log = log_env && log_env[0] == '1';
    float best_ms = 1e30f;
    int best_w = default_window(c, max_bins);
    LaunchPlan best_full;
    int n_trials = 0;
    // after a trial launch has added the sample into the traces
    auto injected = [&]() {
        return ++n_trials == c->fail_trial
                   ? fail(c, TG_ERR_STATE, "tg_autotune: trial " + std::to_string(n_trials) +
                                               " failed (TG_AUTOTUNE_FAIL_TRIAL)")
                   : TG_OK;
    };
    // One timed launch of the sample with launch shape t (its chunk given as
    // the full launch's chunk count n_ch); the kernel time in ms, < 0 on error.
    auto trial = [&](tg_tuning t, long long n_ch, const char* what) -> float {
        t.particles_per_chunk = sample_chunk(n_ch);
        if ((rc = integrate_impl(c, precision, &t, -1, 0, sample)) || (rc = injected())) return -1.f;
        if (log)
            std::fprintf(stderr, "[tg_autotune] %s W=%d groups=%d chunk(full)=%lld spill=%d: %.2f ms (%d particles)\n",
                         what, t.bins_per_window, t.window_groups, (long long)((c->n_particles + n_ch - 1) / n_ch),
                         t.spill, c->kernel_ms, sample);
        return c->kernel_ms;
    };
    auto trials = [&]() -> int {
        // First stage: the window, with the default spill bucket.  Each
        // window is timed once; the two fastest are timed twice more and
        // compared by their fastest trial (neighbouring windows lie within
        // 1-2 % of each other, about the spread of one trial).
        struct Cand {
            int w;
            float ms;
            LaunchPlan full;
            tg_tuning t;
        };
        std::vector<Cand> cs;
        for (int w : cands) {
            if (w > wmax || w > max_bins) continue;
            tg_tuning t = {};
            t.bins_per_window = w;
            LaunchPlan full;
            if ((rc = make_plan(c, precision, ids, 0, c->n_particles, &t, full))) return rc;
            t.window_groups = full.groups;
            const float ms = trial(t, full.n_chunks, "window");
            if (ms < 0.f) return rc;
            cs.push_back(Cand{w, ms, full, t});
        }
        std::sort(cs.begin(), cs.end(), [](const Cand& a, const Cand& b) { return a.ms < b.ms; });
        for (size_t i = 0; i < cs.size() && i < 2; ++i)
            for (int r = 0; r < 2; ++r) {
                const float ms = trial(cs[i].t, cs[i].full.n_chunks, "window again");
                if (ms < 0.f) return rc;
                cs[i].ms = std::min(cs[i].ms, ms);
            }
        for (const Cand& cd : cs)
            if (cd.ms < best_ms) {
                best_ms = cd.ms;
                best_w = cd.w;
                best_full = cd.full;
            }
        // Second stage, at the chosen window: the number of window groups (item
        // granularity against the search each group costs), the particle
        // chunk (items per tile against partial traces to reduce), each halved
        // and doubled around the default, and the window without its spill
        // bucket (six registers per lane fewer, against walking every
        // window's crossing steps twice), at the same window and one wider.  The fastest shape and the default are then timed once
        // more each, and the shape replaces the default only when its fastest
        // trial is more than 1 % faster, so timing noise does not move it.
        if (best_full.chunk > 0) {
            const int g0 = best_full.groups, ch0 = best_full.chunk;
            tg_tuning pick = {};
            pick.bins_per_window = best_w;
            const float base_ms = best_ms;
            float shape_ms = base_ms;
            size_t per_chunk = 0;
            for (int id : ids) per_chunk += (size_t)c->screens[id].n_pix * c->screens[id].n_bins * 6;
            const int spill_w1 = std::min(best_w + 1, std::min(max_bins, max_window_for_lds(c, precision, true)));
            const int shapes[6][4] = {{best_w, g0 / 2, ch0, 0}, {best_w, g0 * 2, ch0, 0}, {best_w, g0, ch0 * 2, 0},
                                      {best_w, g0, ch0 / 2, 0}, {best_w, g0, ch0, 2},     {spill_w1, g0, ch0, 2}};
            long long pick_nch = 0;
            for (const auto& sh : shapes) {
                const int w = sh[0], g = sh[1], ch = sh[2], sp = sh[3];
                if (g < 1 || g > 64 || ch < 1 || ch > c->n_particles) continue;
                if (sp == 2 && !lane_walk(c, nullptr)) continue;
                const size_t n_ch = ((size_t)c->n_particles + ch - 1) / ch;
                if ((n_ch - 1) * per_chunk * sizeof(double) > ((size_t)16 << 30)) continue;
                tg_tuning t = {};
                t.bins_per_window = w;
                t.window_groups = g;
                t.spill = sp;
                const float ms = trial(t, (long long)n_ch, "shape");
                if (ms < 0.f) return rc;
                if (ms < shape_ms) {
                    shape_ms = ms;
                    pick = t;
                    pick.particles_per_chunk = ch;
                    pick_nch = (long long)n_ch;
                }
            }
            float again_ms = base_ms;
            if (pick_nch > 0) {
                tg_tuning t = pick;
                const float ms = trial(t, pick_nch, "shape again");
                if (ms < 0.f) return rc;
                shape_ms = std::min(shape_ms, ms);
                tg_tuning b = {};
                b.bins_per_window = best_w;
                b.window_groups = g0;
                const float mb = trial(b, best_full.n_chunks, "default again");
                if (mb < 0.f) return rc;
                again_ms = std::min(base_ms, mb);
            }
            if (shape_ms < 0.99f * again_ms) {
                best_w = pick.bins_per_window;
                best.spill = pick.spill;
                if (pick.window_groups != g0 || pick.particles_per_chunk != ch0) {
                    best.window_groups = pick.window_groups;
                    best.particles_per_chunk = pick.particles_per_chunk;
                }
            }
        }
        return TG_OK;
    };
    // the trials add the sample's fields into the screens: restore every
    // screen's traces whether or not the trials succeeded
    const int trial_rc = trials();
    int restore_rc = TG_OK;
    for (size_t id = 0; id < c->screens.size(); ++id) {
        if (saved[id].empty()) continue;
        ScreenHost& s = c->screens[id];
        const hipError_t e = hipMemcpy(s.d_traces, saved[id].data(), saved[id].size() * sizeof(double),
                                       hipMemcpyHostToDevice);
        if (e != hipSuccess && restore_rc == TG_OK)
            restore_rc = fail(c, TG_ERR_HIP, std::string("tg_autotune: restoring traces: ") + hipGetErrorString(e));
    }
    free_graph(c);
    if (trial_rc) return trial_rc;
    if (restore_rc) return restore_rc;
    for (ScreenHost& s : c->screens) s.window[precision] = best_w;
    best.bins_per_window = 0;  // per-screen windows stored on the screens
    c->tuned[precision] = best;
    c->has_tuned[precision] = true;
    if (out) {
        *out = best;
        int w = 0;
        for (auto& s : c->screens) w = std::max(w, s.window[precision]);
        out->bins_per_window = w;
    }
    return TG_OK;
}

int tg_track_lattice(tg_context* c, int n_particles, const double* x0, const double* u0, double t_start,
                     double charge, double mass, double dt, int n_steps, double bpeak, double lambda_u,
                     int n_periods, const double* origin, const double* E0, const double* B0)
{
    if (check_ctx(c)) return TG_ERR_ARG;
    if (n_particles < 0 || n_steps < 0) return fail(c, TG_ERR_ARG, "negative size");
    if (n_particles > 0 && (!x0 || !u0)) return fail(c, TG_ERR_ARG, "null initial conditions");
    if (!(dt > 0.0) || !(mass > 0.0)) return fail(c, TG_ERR_ARG, "dt and mass must be > 0");
    if (origin && !(lambda_u > 0.0)) return fail(c, TG_ERR_ARG, "lambda_u must be > 0");
    TG_HIP(c, hipSetDevice(c->device));
    int rc = alloc_traj(c, n_particles, n_steps + 1);
    if (rc) return rc;
    if (n_particles == 0) return TG_OK;
    double* d0 = nullptr;
    TG_HIP(c, hipMalloc((void**)&d0, (size_t)6 * n_particles * sizeof(double)));
    hipError_t e = hipMemcpy(d0, x0, (size_t)3 * n_particles * sizeof(double), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(d0 + 3 * (size_t)n_particles, u0, (size_t)3 * n_particles * sizeof(double),
                                       hipMemcpyHostToDevice);
    // charge array
    std::vector<double> q((size_t)n_particles, charge);
    if (e == hipSuccess) e = hipMemcpy(raw_charge(c), q.data(), q.size() * sizeof(double), hipMemcpyHostToDevice);
    if (e == hipSuccess) {
        LatticeField lat{};
        lat.has_und = origin != nullptr;
        if (origin) {
            lat.und.bpeak = bpeak;
            lat.und.lambda_u = lambda_u;
            lat.und.kz = 2.0 * kPi / lambda_u;
            lat.und.z1 = -(n_periods + 1) * lambda_u / 2.0;
            lat.und.z2 = (n_periods + 1) * lambda_u / 2.0;
            lat.und.origin = V3{origin[0], origin[1], origin[2]};
        }
        lat.has_hom = E0 != nullptr || B0 != nullptr;
        lat.E0 = E0 ? V3{E0[0], E0[1], E0[2]} : V3{0.0, 0.0, 0.0};
        lat.B0 = B0 ? V3{B0[0], B0[1], B0[2]} : V3{0.0, 0.0, 0.0};
        double qm = charge / mass * kInvRestMass;
        double qmt2 = qm * 0.5 * dt;
        vay_lattice_kernel<<<(n_particles + 63) / 64, 64, 0, c->stream>>>(
            n_particles, d0, d0 + 3 * (size_t)n_particles, t_start, qm, qmt2, dt, n_steps, lat, raw(c, 0), raw(c, 1),
            raw(c, 2), raw(c, 3), raw(c, 4), raw(c, 5), raw(c, 6), raw(c, 7), raw(c, 8), raw(c, 9));
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    (void)hipFree(d0);
    if (e != hipSuccess) return fail(c, TG_ERR_HIP, std::string("tg_track_lattice: ") + hipGetErrorString(e));
    return TG_OK;
}

int tg_track_undulator(tg_context* c, int n_particles, const double* x0, const double* u0, double t_start,
                       double charge, double mass, double dt, int n_steps, double bpeak, double lambda_u,
                       int n_periods, const double* origin)
{
    if (check_ctx(c)) return TG_ERR_ARG;
    if (!origin) return fail(c, TG_ERR_ARG, "null origin");
    return tg_track_lattice(c, n_particles, x0, u0, t_start, charge, mass, dt, n_steps, bpeak, lambda_u, n_periods,
                            origin, nullptr, nullptr);
}

}  // extern "C"
