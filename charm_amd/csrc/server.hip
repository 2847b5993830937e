// Request-serving engine: continuous admission of synthesis/analysis
// requests against models and grids resident in HBM.
//
// A worker thread owns all device work.  Requests are admitted in arrival
// order while the HBM budget and the in-flight limit allow; each admitted
// request runs on its own stream from a pool, so requests that arrive while
// others are on the GPU overlap with them.  Completion is detected with
// events; results are copied out through pinned staging buffers, device
// memory is returned to the stream-ordered pool, and the next requests are
// admitted.  Results do not depend on what else is in flight: every kernel is
// deterministic and requests share only read-only data.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <mutex>
#include <thread>
#include <vector>

#include "internal.h"

using namespace charm_amd;

namespace {

double now_ms()
{
    using namespace std::chrono;
    return duration<double, std::milli>(steady_clock::now().time_since_epoch()).count();
}

struct Grid;

struct SrvReq {
    long ticket = 0;
    charm_amd_srv_request r{};
    size_t bytes = 0;
    int state = 0;  // 0 queued, 1 in flight, 2 done, 3 failed
    double t_submit = 0, t_start = 0, t_done = 0;
    int stream = -1;
    hipEvent_t ev = nullptr;
    std::vector<void *> dev;
    void *pinned = nullptr;
    size_t pinned_bytes = 0, out_bytes = 0;
    char msg[256] = {0};
    int code = 0;
    struct Grid *grid = nullptr;  // resident grid in use (not evictable meanwhile)
};

struct Model {
    double *cs = nullptr;
    unsigned long nmax = 0;
    double mu = 0, R = 0;
};

struct GridKey {
    unsigned long nmax;
    double r;
    bool operator<(const GridKey &o) const { return nmax < o.nmax || (nmax == o.nmax && r < o.r); }
};

struct Grid {
    charm_amd_grid *g = nullptr;
    size_t nlat = 0, nlon = 0;
    size_t bytes = 0;
    int users = 0;
    double last = 0;
};

}  // namespace

struct charm_amd_server {
    size_t budget = 0, in_use = 0;
    int max_inflight = 4;
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::thread worker;
    bool stop = false;
    long next_ticket = 1;
    std::vector<Model> models;
    std::map<GridKey, Grid> grids;
    std::deque<SrvReq *> pending;
    std::vector<SrvReq *> inflight;
    std::map<long, SrvReq *> all;
    std::vector<hipStream_t> streams;
    std::vector<bool> stream_busy;
    struct Pinned {
        void *p;
        size_t n;
        bool busy;
    };
    std::vector<Pinned> pinned;
};

namespace {

size_t gl_bytes(unsigned long nmax)
{
    const size_t nlat = nmax + 1, nlon = 2 * (nmax + 1), nring = (nlat + 1) / 2;
    // latitude tables + seed table (8 B mantissa + 4 B exponent per (m, ring))
    return nring * 32 + nlat * 16 + nlon * 8 + (size_t)std::max<unsigned long>(nmax, 1) * nring * 12;
}

// device bytes a request needs while in flight (grid tables counted apart)
size_t footprint(const charm_amd_srv_request &r)
{
    const size_t P = tri_len(r.nmax);
    switch (r.kind) {
    case CHARM_AMD_SRV_SHS_SCTR: {
        const size_t nw = (r.npoint + 63) / 64;
        const size_t nchunk = std::min<size_t>(64, std::max<size_t>(1, 4096 / std::max<size_t>(nw, 1)));
        return r.npoint * 8 * (6 + nchunk) + P * 16;
    }
    case CHARM_AMD_SRV_SHS_GRID: {
        const size_t nlat = r.grid_nmax + 1, nlon = 2 * nlat;
        return nlat * nlon * 8 + nlat * (nlon / 2 + 1) * 16 + P * 16;
    }
    default: {
        const size_t nlat = r.grid_nmax + 1, nlon = 2 * nlat, nring = (nlat + 1) / 2;
        const size_t ntile = (nring + 191) / 192 + 1;
        return nlat * nlon * 8 + nlat * (nlon / 2 + 1) * 16 + ntile * P * 16 + P * 16;
    }
    }
}

void *pinned_get(charm_amd_server *s, size_t n)
{
    for (auto &p : s->pinned)
        if (!p.busy && p.n >= n) {
            p.busy = true;
            return p.p;
        }
    void *p = nullptr;
    if (hipHostMalloc(&p, std::max<size_t>(n, 64), hipHostMallocDefault) != hipSuccess) return nullptr;
    s->pinned.push_back({p, std::max<size_t>(n, 64), true});
    return p;
}

void pinned_put(charm_amd_server *s, void *q)
{
    for (auto &p : s->pinned)
        if (p.p == q) p.busy = false;
}

// grid cache: returns a resident grid (creating it if needed), or nullptr
Grid *get_grid(charm_amd_server *s, unsigned long nmax, double r, charm_err *err)
{
    GridKey k{nmax, r};
    auto it = s->grids.find(k);
    if (it != s->grids.end()) {
        it->second.last = now_ms();
        return &it->second;
    }
    // evict idle grids (least recently used first) until the new one fits
    const size_t need = gl_bytes(nmax);
    while (s->in_use + need > s->budget) {
        auto victim = s->grids.end();
        for (auto g = s->grids.begin(); g != s->grids.end(); ++g)
            if (g->second.users == 0 && (victim == s->grids.end() || g->second.last < victim->second.last))
                victim = g;
        if (victim == s->grids.end()) return nullptr;
        s->in_use -= victim->second.bytes;
        charm_amd_grid_free(victim->second.g);
        s->grids.erase(victim);
    }
    charm_point *pnt = charm_crd_point_gl(nmax, r);
    if (!pnt) {
        CHARM_AMD_ERR(err, CHARM_EMEM, "Failed to build the Gauss--Legendre grid.");
        return nullptr;
    }
    Grid g;
    g.g = charm_amd_grid_create(pnt, err);
    g.nlat = pnt->nlat;
    g.nlon = pnt->nlon;
    charm_crd_point_free(pnt);
    if (!g.g) return nullptr;
    g.bytes = need;
    g.last = now_ms();
    s->in_use += need;
    return &(s->grids[k] = g);
}

bool dalloc(SrvReq *q, hipStream_t st, size_t bytes, void **p)
{
    if (hipMallocAsync(p, std::max<size_t>(bytes, 16), st) != hipSuccess) return false;
    q->dev.push_back(*p);
    return true;
}

void fail(SrvReq *q, int code, const char *msg)
{
    q->code = code;
    std::snprintf(q->msg, sizeof q->msg, "%s", msg);
    q->state = 3;
}

// launch a request's work on its stream (called by the worker, lock held)
void start(charm_amd_server *s, SrvReq *q)
{
    const charm_amd_srv_request &r = q->r;
    hipStream_t st = s->streams[q->stream];
    charm_err *err = charm_err_init();
    q->t_start = now_ms();
    const Model *md = (r.model >= 0 && r.model < (int)s->models.size()) ? &s->models[r.model] : nullptr;
    const size_t P = tri_len(r.nmax);
    bool ok = true;
    if (r.kind == CHARM_AMD_SRV_SHS_SCTR) {
        const size_t n = r.npoint;
        void *lat, *lon, *rr, *f, *cs;
        ok = md && r.nmax <= md->nmax && dalloc(q, st, n * 8, &lat) && dalloc(q, st, n * 8, &lon) &&
             dalloc(q, st, n * 8, &rr) && dalloc(q, st, n * 8, &f) && dalloc(q, st, P * 16, &cs);
        if (ok) {
            q->out_bytes = n * 8;
            q->pinned = pinned_get(s, 3 * n * 8);
            ok = q->pinned != nullptr;
        }
        if (ok) {
            double *pin = (double *)q->pinned;
            std::memcpy(pin, r.lat, n * 8);
            std::memcpy(pin + n, r.lon, n * 8);
            std::memcpy(pin + 2 * n, r.r, n * 8);
            ok = hipMemcpyAsync(lat, pin, n * 8, hipMemcpyHostToDevice, st) == hipSuccess &&
                 hipMemcpyAsync(lon, pin + n, n * 8, hipMemcpyHostToDevice, st) == hipSuccess &&
                 hipMemcpyAsync(rr, pin + 2 * n, n * 8, hipMemcpyHostToDevice, st) == hipSuccess;
            charm_amd_cs_truncate(md->cs, md->nmax, (double *)cs, r.nmax, st, err);
            charm_amd_shs_sctr_dev(n, (double *)lat, (double *)lon, (double *)rr, r.nmax, md->mu, md->R,
                                   (double *)cs, (double *)f, st, err);
            ok = ok && charm_err_isempty(err) &&
                 hipMemcpyAsync(pin, f, n * 8, hipMemcpyDeviceToHost, st) == hipSuccess;
        }
    } else if (r.kind == CHARM_AMD_SRV_SHS_GRID) {
        Grid *g = md ? get_grid(s, r.grid_nmax, r.r_grid, err) : nullptr;
        void *f, *cs;
        ok = g && md && r.nmax <= md->nmax && r.nmax <= r.grid_nmax;
        if (g) {
            g->users++;
            q->grid = g;
        }
        if (ok) ok = dalloc(q, st, g->nlat * g->nlon * 8, &f) && dalloc(q, st, P * 16, &cs);
        if (ok) {
            q->out_bytes = g->nlat * g->nlon * 8;
            q->pinned = pinned_get(s, q->out_bytes);
            ok = q->pinned != nullptr;
        }
        if (ok) {
            charm_amd_cs_truncate(md->cs, md->nmax, (double *)cs, r.nmax, st, err);
            charm_amd_shs_grid_dev(g->g, r.nmax, md->mu, md->R, (double *)cs, (double *)f, st, err);
            ok = charm_err_isempty(err) &&
                 hipMemcpyAsync(q->pinned, f, q->out_bytes, hipMemcpyDeviceToHost, st) == hipSuccess;
        }
    } else {
        Grid *g = get_grid(s, r.grid_nmax, r.r_grid, err);
        void *f, *cs;
        ok = g && r.nmax <= r.grid_nmax;
        if (g) {
            g->users++;
            q->grid = g;
        }
        if (ok) ok = dalloc(q, st, g->nlat * g->nlon * 8, &f) && dalloc(q, st, P * 16, &cs);
        if (ok) {
            q->out_bytes = P * 16;
            q->pinned = pinned_get(s, std::max(q->out_bytes, g->nlat * g->nlon * 8));
            ok = q->pinned != nullptr;
        }
        if (ok) {
            std::memcpy(q->pinned, r.f_in, g->nlat * g->nlon * 8);
            ok = hipMemcpyAsync(f, q->pinned, g->nlat * g->nlon * 8, hipMemcpyHostToDevice, st) ==
                 hipSuccess;
            charm_amd_sha_grid_dev(g->g, r.nmax, r.mu, r.R, (double *)f, (double *)cs, st, err);
            ok = ok && charm_err_isempty(err) &&
                 hipMemcpyAsync(q->pinned, cs, q->out_bytes, hipMemcpyDeviceToHost, st) == hipSuccess;
        }
    }
    if (ok) ok = hipEventRecord(q->ev, st) == hipSuccess;
    if (!ok) {
        (void)hipStreamSynchronize(st);
        fail(q, charm_err_isempty(err) ? CHARM_EMEM : err->code,
             charm_err_isempty(err) ? "Failed to admit the request on the device." : err->msg);
    } else {
        q->state = 1;
    }
    charm_err_free(err);
}

void finish(charm_amd_server *s, SrvReq *q)
{
    hipStream_t st = s->streams[q->stream];
    if (q->state == 1 && q->pinned && q->r.out) std::memcpy(q->r.out, q->pinned, q->out_bytes);
    for (void *p : q->dev) (void)hipFreeAsync(p, st);
    q->dev.clear();
    if (q->pinned) pinned_put(s, q->pinned);
    q->pinned = nullptr;
    if (q->grid) q->grid->users--;
    q->grid = nullptr;
    s->in_use -= q->bytes;
    s->stream_busy[q->stream] = false;
    if (q->state == 1) q->state = 2;
    q->t_done = now_ms();
}

void worker_loop(charm_amd_server *s)
{
    set_thread_tuning(false);  // never time kernels against requests in flight
    std::unique_lock<std::mutex> lk(s->mu);
    while (true) {
        bool progress = false;
        // retire finished requests
        for (size_t i = 0; i < s->inflight.size();) {
            SrvReq *q = s->inflight[i];
            if (q->state == 3 || hipEventQuery(q->ev) == hipSuccess) {
                finish(s, q);
                s->inflight.erase(s->inflight.begin() + i);
                progress = true;
            } else {
                i++;
            }
        }
        if (progress) s->cv_done.notify_all();
        // admit in arrival order while the budget and a stream allow
        while (!s->pending.empty() && (int)s->inflight.size() < s->max_inflight) {
            SrvReq *q = s->pending.front();
            if (s->in_use + q->bytes > s->budget && !s->inflight.empty()) break;
            int k = -1;
            for (int j = 0; j < (int)s->streams.size(); j++)
                if (!s->stream_busy[j]) {
                    k = j;
                    break;
                }
            if (k < 0) break;
            s->pending.pop_front();
            s->stream_busy[k] = true;
            q->stream = k;
            s->in_use += q->bytes;
            s->inflight.push_back(q);
            start(s, q);
            progress = true;
        }
        if (s->stop && s->pending.empty() && s->inflight.empty()) break;
        if (!progress) {
            if (s->inflight.empty())
                s->cv_work.wait(lk, [s] { return s->stop || !s->pending.empty(); });
            else
                s->cv_work.wait_for(lk, std::chrono::microseconds(50));
        }
    }
}

}  // namespace

extern "C" {

charm_amd_server *charm_amd_server_create(size_t hbm_budget_bytes, int max_inflight, charm_err *err)
{
    charm_amd_server *s = new charm_amd_server();
    s->budget = hbm_budget_bytes;
    s->max_inflight = std::max(1, max_inflight);
    for (int i = 0; i < s->max_inflight; i++) {
        hipStream_t st;
        if (hipStreamCreateWithFlags(&st, hipStreamNonBlocking) != hipSuccess) {
            CHARM_AMD_ERR(err, CHARM_EMEM, "Failed to create the server streams.");
            charm_amd_server_free(s);
            return nullptr;
        }
        s->streams.push_back(st);
        s->stream_busy.push_back(false);
    }
    s->worker = std::thread(worker_loop, s);
    return s;
}

int charm_amd_server_add_model(charm_amd_server *s, const charm_shc *shcs, unsigned long nmax,
                               charm_err *err)
{
    if (!s || !shcs || nmax > shcs->nmax) {
        CHARM_AMD_ERR(err, CHARM_EFUNCARG, "Invalid model.");
        return -1;
    }
    const size_t P = tri_len(nmax);
    std::vector<double> cs(2 * P);
    size_t q = 0;
    for (unsigned long m = 0; m <= nmax; m++)
        for (unsigned long n = m; n <= nmax; n++, q++) {
            cs[2 * q] = shcs->c[m][n - m];
            cs[2 * q + 1] = (m == 0) ? 0.0 : shcs->s[m][n - m];
        }
    std::lock_guard<std::mutex> lk(s->mu);
    if (s->in_use + P * 16 > s->budget) {
        CHARM_AMD_ERR(err, CHARM_EMEM, "The model does not fit the HBM budget.");
        return -1;
    }
    Model md;
    md.nmax = nmax;
    md.mu = shcs->mu;
    md.R = shcs->r;
    if (hipMalloc((void **)&md.cs, P * 16) != hipSuccess ||
        hipMemcpy(md.cs, cs.data(), P * 16, hipMemcpyHostToDevice) != hipSuccess) {
        CHARM_AMD_ERR(err, CHARM_EMEM, "Failed to make the model resident.");
        return -1;
    }
    s->in_use += P * 16;
    s->models.push_back(md);
    return (int)s->models.size() - 1;
}

long charm_amd_server_submit(charm_amd_server *s, const charm_amd_srv_request *r, charm_err *err)
{
    if (!s || !r || !r->out ||
        (r->kind != CHARM_AMD_SRV_SHS_SCTR && r->kind != CHARM_AMD_SRV_SHS_GRID &&
         r->kind != CHARM_AMD_SRV_SHA_GRID)) {
        CHARM_AMD_ERR(err, CHARM_EFUNCARG, "Invalid request.");
        return -1;
    }
    SrvReq *q = new SrvReq();
    q->r = *r;
    q->bytes = footprint(*r);
    if (hipEventCreateWithFlags(&q->ev, hipEventDisableTiming) != hipSuccess) {
        delete q;
        CHARM_AMD_ERR(err, CHARM_EMEM, "Failed to create an event.");
        return -1;
    }
    std::lock_guard<std::mutex> lk(s->mu);
    if (q->bytes > s->budget) {
        (void)hipEventDestroy(q->ev);
        delete q;
        CHARM_AMD_ERR(err, CHARM_EMEM, "The request exceeds the HBM budget.");
        return -1;
    }
    q->ticket = s->next_ticket++;
    q->t_submit = now_ms();
    s->all[q->ticket] = q;
    s->pending.push_back(q);
    s->cv_work.notify_one();
    return q->ticket;
}

int charm_amd_server_wait(charm_amd_server *s, long ticket, double *latency_ms, charm_err *err)
{
    std::unique_lock<std::mutex> lk(s->mu);
    auto it = s->all.find(ticket);
    if (it == s->all.end()) {
        CHARM_AMD_ERR(err, CHARM_EFUNCARG, "Unknown ticket.");
        return -1;
    }
    SrvReq *q = it->second;
    s->cv_done.wait(lk, [q] { return q->state >= 2 && q->t_done > 0; });
    if (latency_ms) *latency_ms = q->t_done - q->t_submit;
    int rc = 0;
    if (q->state == 3) {
        CHARM_AMD_ERR(err, q->code ? q->code : CHARM_EFUNCARG, q->msg);
        rc = -1;
    }
    (void)hipEventDestroy(q->ev);
    s->all.erase(it);
    delete q;
    return rc;
}

size_t charm_amd_server_in_use(charm_amd_server *s)
{
    std::lock_guard<std::mutex> lk(s->mu);
    return s->in_use;
}

void charm_amd_server_free(charm_amd_server *s)
{
    if (!s) return;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->stop = true;
        s->cv_work.notify_all();
    }
    if (s->worker.joinable()) s->worker.join();
    (void)hipDeviceSynchronize();
    for (auto &kv : s->all) {
        (void)hipEventDestroy(kv.second->ev);
        delete kv.second;
    }
    for (auto &g : s->grids) charm_amd_grid_free(g.second.g);
    for (auto &m : s->models) (void)hipFree(m.cs);
    for (auto &p : s->pinned) (void)hipHostFree(p.p);
    for (auto st : s->streams) (void)hipStreamDestroy(st);
    delete s;
}

}  // extern "C"
