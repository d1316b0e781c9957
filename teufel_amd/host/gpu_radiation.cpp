// gpu_radiation.cpp - host adapter over the C ABI.
#include "gpu_radiation.h"

#include <condition_variable>
#include <cstdlib>
#include <cstring>
#include <exception>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <utility>

#include "../../include/teufel_gpu.h"

namespace teufel_gpu {
namespace {

// One integrate() call waiting to be launched.
struct Request {
    const std::vector<ChargedParticle*>* particles;
    const std::vector<Vector>* points;
    const std::vector<FieldTrace*>* traces;
    bool done = false;
    std::exception_ptr error;
};

struct State {
    std::mutex mu;
    // integrate() requests from concurrent callers (the reference's
    // per-pixel OpenMP loop calls integrateFieldTrace from many threads,
    // src/mesh.cpp:296-301): one caller at a time launches everything queued
    // so far, one batched launch per particle list, while the others wait.
    std::mutex qmu;
    std::condition_variable qcv;
    std::vector<Request*> queue;
    bool launching = false;
    unsigned long long launches = 0;   // batched launches made (under mu)
    tg_context* ctx = nullptr;
    int device = -1;
    int precision = -1;
    // trajectory cache: (particle, version) list of the last upload
    std::vector<std::pair<const ChargedParticle*, unsigned long long>> key;
    ~State()
    {
        if (ctx) tg_destroy(ctx);
    }
};

State& state()
{
    static State s;
    return s;
}

void check(tg_context* ctx, int rc, const char* what)
{
    if (rc != TG_OK) {
        std::string m = std::string("teufel_gpu: ") + what + " failed (status " + std::to_string(rc) + ")";
        if (ctx) m += ": " + std::string(tg_last_error(ctx));
        throw IOexception(m);
    }
}

tg_context* context(State& s)
{
    if (!s.ctx) {
        if (s.device < 0) {
            const char* d = std::getenv("TEUFEL_GPU_DEVICE");
            s.device = d ? std::atoi(d) : 0;
        }
        if (s.precision < 0) {
            const char* p = std::getenv("TEUFEL_GPU_PRECISION");
            s.precision = (p && std::strcmp(p, "mixed") == 0) ? TG_MIXED : TG_FP64;
        }
        check(nullptr, tg_create(s.device, &s.ctx), "tg_create");
    }
    return s.ctx;
}

void upload(State& s, const std::vector<ChargedParticle*>& ps)
{
    std::vector<std::pair<const ChargedParticle*, unsigned long long>> key;
    key.reserve(ps.size());
    for (const ChargedParticle* p : ps) key.emplace_back(p, p->version());
    if (key == s.key) return;
    int np = (int)ps.size(), npts = 0;
    for (const ChargedParticle* p : ps) npts = std::max(npts, p->getNP());
    size_t n = (size_t)np * npts;
    std::vector<double> a(10 * n), q(np);
    double* t = a.data();
    double *x = t + n, *y = x + n, *z = y + n, *ux = z + n, *uy = ux + n, *uz = uy + n, *ax = uz + n, *ay = ax + n,
           *az = ay + n;
    for (int i = 0; i < np; i++) {
        const ChargedParticle* p = ps[i];
        q[i] = p->getCharge();
        for (int k = 0; k < npts; k++) {
            int kk = p->getNP() > 0 ? std::min(k, p->getNP() - 1) : -1;
            size_t g = (size_t)i * npts + k;
            if (kk < 0) {
                t[g] = x[g] = y[g] = z[g] = ux[g] = uy[g] = uz[g] = ax[g] = ay[g] = az[g] = 0.0;
                continue;
            }
            Vector X = p->TrajPoint(kk), P = p->TrajMomentum(kk), A = p->TrajAccel(kk);
            t[g] = p->TrajTime(kk);
            x[g] = X.x; y[g] = X.y; z[g] = X.z;
            ux[g] = P.x; uy[g] = P.y; uz[g] = P.z;
            ax[g] = A.x; ay[g] = A.y; az[g] = A.z;
        }
    }
    check(s.ctx, tg_set_trajectories(s.ctx, np, npts, t, x, y, z, ux, uy, uz, ax, ay, az, q.data()),
          "tg_set_trajectories");
    s.key.swap(key);
}

}  // namespace

void set_precision(int p)
{
    if (p != TG_FP64 && p != TG_MIXED) throw IOexception("teufel_gpu::set_precision - TG_FP64 or TG_MIXED");
    std::lock_guard<std::mutex> g(state().mu);
    state().precision = p;
}

int precision()
{
    State& s = state();
    std::lock_guard<std::mutex> g(s.mu);
    context(s);
    return s.precision;
}

void set_device(int device)
{
    State& s = state();
    std::lock_guard<std::mutex> g(s.mu);
    if (s.ctx && device != s.device) {
        tg_destroy(s.ctx);
        s.ctx = nullptr;
        s.key.clear();
    }
    s.device = device;
}

namespace {

// One batched launch for points/traces over particles (state mutex held).
void integrate_locked(State& s, const std::vector<ChargedParticle*>& particles, const std::vector<Vector>& points,
                      const std::vector<FieldTrace*>& traces)
{
    tg_context* ctx = context(s);
    upload(s, particles);
    check(ctx, tg_remove_screens(ctx), "tg_remove_screens");
    // one screen per (dt, N) class of traces
    std::map<std::pair<double, size_t>, std::vector<size_t>> classes;
    for (size_t i = 0; i < traces.size(); i++) {
        if (!traces[i]) throw IOexception("teufel_gpu::integrate - null FieldTrace.");
        classes[std::make_pair(traces[i]->get_dt(), traces[i]->get_N())].push_back(i);
    }
    std::vector<std::pair<int, const std::vector<size_t>*>> screens;
    for (const auto& c : classes) {
        const std::vector<size_t>& idx = c.second;
        std::vector<double> obs(3 * idx.size()), t0(idx.size());
        for (size_t j = 0; j < idx.size(); j++) {
            const Vector& P = points[idx[j]];
            obs[3 * j] = P.x; obs[3 * j + 1] = P.y; obs[3 * j + 2] = P.z;
            t0[j] = traces[idx[j]]->get_t0();
        }
        int sid = -1;
        check(ctx, tg_add_screen(ctx, (int)idx.size(), obs.data(), t0.data(), c.first.first, (int)c.first.second, &sid),
              "tg_add_screen");
        screens.emplace_back(sid, &c.second);
    }
    check(ctx, tg_integrate(ctx, s.precision, nullptr), "tg_integrate");
    ++s.launches;
    for (const auto& sc : screens) {
        const std::vector<size_t>& idx = *sc.second;
        size_t nb = traces[idx[0]]->get_N();
        std::vector<double> buf(idx.size() * nb * 6);
        check(ctx, tg_read_traces(ctx, sc.first, buf.data(), 0), "tg_read_traces");
        for (size_t j = 0; j < idx.size(); j++) {
            double* d = traces[idx[j]]->data();
            const double* src = &buf[j * nb * 6];
            for (size_t k = 0; k < nb * 6; k++) d[k] += src[k];
        }
    }
    check(ctx, tg_remove_screens(ctx), "tg_remove_screens");
}

// Launch a batch of queued requests: requests over the same particle list
// (the common case, every pixel of an observer loop over one source) share
// one launch; each request's traces receive its fields, and a request whose
// launch failed gets that launch's exception.
void launch_batch(State& s, const std::vector<Request*>& batch)
{
    std::lock_guard<std::mutex> g(s.mu);
    std::vector<bool> taken(batch.size(), false);
    for (size_t i = 0; i < batch.size(); i++) {
        if (taken[i]) continue;
        std::vector<size_t> group;
        for (size_t j = i; j < batch.size(); j++)
            if (!taken[j] && (j == i || *batch[j]->particles == *batch[i]->particles)) {
                taken[j] = true;
                group.push_back(j);
            }
        std::vector<Vector> pts;
        std::vector<FieldTrace*> trs;
        for (size_t j : group) {
            pts.insert(pts.end(), batch[j]->points->begin(), batch[j]->points->end());
            trs.insert(trs.end(), batch[j]->traces->begin(), batch[j]->traces->end());
        }
        std::exception_ptr err;
        try {
            integrate_locked(s, *batch[i]->particles, pts, trs);
        } catch (...) {
            err = std::current_exception();
        }
        for (size_t j : group) batch[j]->error = err;
    }
}

}  // namespace

void integrate(const std::vector<ChargedParticle*>& particles, const std::vector<Vector>& points,
               const std::vector<FieldTrace*>& traces)
{
    if (points.size() != traces.size()) throw IOexception("teufel_gpu::integrate - points/traces size mismatch.");
    for (const FieldTrace* t : traces)
        if (!t) throw IOexception("teufel_gpu::integrate - null FieldTrace.");
    State& s = state();
    Request r{&particles, &points, &traces};
    std::unique_lock<std::mutex> lk(s.qmu);
    s.queue.push_back(&r);
    while (!r.done) {
        if (s.launching) {
            s.qcv.wait(lk);
            continue;
        }
        // this caller launches everything queued until now, its own request
        // included; requests queued meanwhile go into the next batch
        s.launching = true;
        std::vector<Request*> batch;
        batch.swap(s.queue);
        lk.unlock();
        try {
            launch_batch(s, batch);
        } catch (...) {
            // (host allocation failures outside a launch) every request of
            // the batch gets the error, and the next caller may launch
            const std::exception_ptr err = std::current_exception();
            for (Request* q : batch)
                if (!q->error) q->error = err;
        }
        lk.lock();
        for (Request* q : batch) q->done = true;
        s.launching = false;
        s.qcv.notify_all();
    }
    lk.unlock();
    if (r.error) std::rethrow_exception(r.error);
}

void fields_at(const std::vector<ChargedParticle*>& particles, const std::vector<double>& times,
               const std::vector<Vector>& points, std::vector<ElMagField>& out)
{
    if (times.size() != points.size()) throw IOexception("teufel_gpu::fields_at - times/points size mismatch.");
    for (const ChargedParticle* p : particles)
        if (p->getNP() < 1)
            throw RANGEexception("ChargedParticle::RetardedField() - must initalize trajectory before calling");
    out.assign(points.size(), ElMagField());
    if (particles.empty() || points.empty()) return;
    State& s = state();
    std::lock_guard<std::mutex> g(s.mu);
    tg_context* ctx = context(s);
    upload(s, particles);
    std::vector<double> pos(3 * points.size()), f(6 * points.size(), 0.0);
    for (size_t i = 0; i < points.size(); i++) {
        pos[3 * i] = points[i].x;
        pos[3 * i + 1] = points[i].y;
        pos[3 * i + 2] = points[i].z;
    }
    check(ctx, tg_fields_at(ctx, (int)points.size(), times.data(), pos.data(), 0, (int)particles.size(), f.data(), 0),
          "tg_fields_at");
    for (size_t i = 0; i < points.size(); i++)
        out[i] = ElMagField(Vector(f[6 * i], f[6 * i + 1], f[6 * i + 2]), Vector(f[6 * i + 3], f[6 * i + 4], f[6 * i + 5]));
}

void deposit_only(int n_pseudo, int n_points, const double* ts, const double* fields, FieldTrace* trace)
{
    State& s = state();
    std::lock_guard<std::mutex> g(s.mu);
    tg_context* ctx = context(s);
    check(ctx,
          tg_deposit_only(ctx, n_pseudo, n_points, ts, fields, trace->get_t0(), trace->get_dt(), (int)trace->get_N(),
                          trace->data()),
          "tg_deposit_only");
}

unsigned long long launch_count()
{
    State& s = state();
    std::lock_guard<std::mutex> g(s.mu);
    return s.launches;
}

void shutdown()
{
    State& s = state();
    std::lock_guard<std::mutex> g(s.mu);
    if (s.ctx) tg_destroy(s.ctx);
    s.ctx = nullptr;
    s.key.clear();
}

}  // namespace teufel_gpu
