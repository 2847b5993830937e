// Request-serving engine: continuous admission of synthesis/analysis
// requests against models and grids resident in HBM.
//
// A submitting thread does every piece of host work a request needs before it
// is queued: it copies the request's input arrays into pinned staging
// buffers and makes its Gauss-Legendre grid resident (host nodes and device
// tables, kept in a cache under its own lock).  A worker thread owns the
// device work: it admits queued requests while the HBM budget and the
// in-flight limit allow -- in arrival order, a request that does not fit yet
// letting later ones pass at most BYPASS_MAX times -- and runs each admitted
// request on its own stream, so that requests arriving while others are on the
// GPU overlap with them.  Small scattered-point requests admitted together
// run as one group: one persistent launch over all their (point block, order
// chunk) items (run_sctr_jobs), with each result equal to the request run
// alone.  Completion is detected with events; charm_amd_server_wait copies
// the result out of pinned staging in the waiting thread.  The worker never
// builds grids, copies host arrays or times kernels.  Results do not depend
// on what else is in flight: every kernel is deterministic and requests share
// only read-only data.
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

constexpr int BYPASS_MAX = 8;                  // later requests passing one that does not fit
constexpr size_t GROUP_MAX_POINTS = 1 << 17;   // scattered requests grouped up to this size
constexpr int GROUP_MAX = 32;                  // requests per group launch
constexpr size_t PINNED_CAP = size_t(4) << 30;  // pinned staging kept for reuse

double now_ms()
{
    using namespace std::chrono;
    return duration<double, std::milli>(steady_clock::now().time_since_epoch()).count();
}

struct Pinned {
    void *p = nullptr;
    size_t n = 0;
    bool busy = false;
};

struct Grid {
    charm_amd_grid *g = nullptr;
    size_t nlat = 0, nlon = 0, nring = 0;
    size_t bytes = 0;
    int users = 0;  // queued or in-flight requests using it (under the server lock)
    double last = 0;
};

struct SrvReq {
    long ticket = 0;
    charm_amd_srv_request r{};
    size_t bytes = 0;
    int state = 0;  // 0 queued, 1 in flight, 2 done, 3 failed, 4 cancelled
    double t_submit = 0, t_start = 0, t_done = 0;
    int stream = -1;
    hipEvent_t ev = nullptr;  // completion of the request's launch unit (its group's lead)
    std::vector<void *> dev;
    Pinned *in = nullptr, *out = nullptr;
    size_t out_bytes = 0;
    Grid *grid = nullptr;
    int bypassed = 0;
    bool cancel = false;
    char msg[256] = {0};
    int code = 0;
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

}  // namespace

struct charm_amd_server {
    size_t budget = 0, in_use = 0;
    int max_inflight = 4;
    std::mutex mu;  // queue, accounting, streams, requests
    std::condition_variable cv_work, cv_done;
    std::thread worker;
    bool stop = false;
    long next_ticket = 1;
    std::vector<Model> models;
    std::deque<SrvReq *> pending;
    std::vector<std::vector<SrvReq *>> inflight;  // launch units (a request or a group)
    std::map<long, SrvReq *> all;
    std::vector<hipStream_t> streams;
    std::vector<bool> stream_busy;
    int fail_next = 0;  // simulated device errors still to inject
    int fail_code = 0;
    charm_amd_srv_stats stats{};
    std::mutex grid_mu;  // grid cache (lock order: grid_mu, then mu)
    std::map<GridKey, Grid *> grids;
    std::mutex pin_mu;  // pinned staging pool
    std::vector<Pinned *> pinned;
};

namespace {

size_t gl_bytes(unsigned long nmax)
{
    const size_t nlat = nmax + 1, nlon = 2 * (nmax + 1), nring = (nlat + 1) / 2;
    // latitude tables + seed table (8 B mantissa + 4 B exponent per (m, ring))
    return nring * 32 + nlat * 24 + nlon * 8 + (size_t)std::max<unsigned long>(nmax, 1) * nring * 12;
}

// device bytes a request holds while in flight: its inputs and outputs in
// HBM, the model truncated to its degree, and what the library's call
// allocates internally (grid tables are counted with the grid)
size_t footprint(const charm_amd_srv_request &r)
{
    const size_t P = tri_len(r.nmax);
    if (r.kind == CHARM_AMD_SRV_SHS_SCTR)
        return r.npoint * 4 * sizeof(double) + P * 16 + sctr_call_scratch(r.npoint, r.nmax);
    const size_t nlat = r.grid_nmax + 1, nlon = 2 * nlat, nring = (nlat + 1) / 2;
    return nlat * nlon * sizeof(double) + P * 16 +
           grid_call_scratch(r.kind == CHARM_AMD_SRV_SHA_GRID, nring, nlat, nlon, r.nmax);
}

// A pinned staging buffer of at least n bytes from the pool (reused when an
// idle one is at most twice as large).  Idle buffers are released, largest
// first, while a new one would take the pool beyond PINNED_CAP bytes.
Pinned *pinned_get(charm_amd_server *s, size_t n)
{
    n = std::max<size_t>(n, 64);
    std::lock_guard<std::mutex> lk(s->pin_mu);
    for (Pinned *p : s->pinned)
        if (!p->busy && p->n >= n && p->n <= 2 * n) {
            p->busy = true;
            return p;
        }
    size_t total = 0;
    for (Pinned *p : s->pinned) total += p->n;
    while (total + n > PINNED_CAP) {
        auto big = s->pinned.end();
        for (auto it = s->pinned.begin(); it != s->pinned.end(); ++it)
            if (!(*it)->busy && (big == s->pinned.end() || (*it)->n > (*big)->n)) big = it;
        if (big == s->pinned.end()) break;  // all busy: the pool grows
        total -= (*big)->n;
        (void)hipHostFree((*big)->p);
        delete *big;
        s->pinned.erase(big);
    }
    Pinned *p = new Pinned();
    if (hipHostMalloc(&p->p, n, hipHostMallocDefault) != hipSuccess) {
        (void)hipGetLastError();
        delete p;
        return nullptr;
    }
    p->n = n;
    p->busy = true;
    s->pinned.push_back(p);
    return p;
}

void pinned_put(charm_amd_server *s, Pinned *p)
{
    if (!p) return;
    std::lock_guard<std::mutex> lk(s->pin_mu);
    p->busy = false;
}

// A resident grid for (nmax, r), built if needed (in the calling thread, a
// submitter, never the worker).  Idle grids are evicted, least recently used
// first, until the new one fits the budget.  The returned grid has one more
// user (released when its request finishes).
Grid *acquire_grid(charm_amd_server *s, unsigned long nmax, double r, charm_err *err)
{
    std::lock_guard<std::mutex> glk(s->grid_mu);
    GridKey k{nmax, r};
    auto it = s->grids.find(k);
    if (it != s->grids.end()) {
        std::lock_guard<std::mutex> lk(s->mu);
        it->second->users++;
        it->second->last = now_ms();
        return it->second;
    }
    const size_t need = gl_bytes(nmax);
    for (;;) {
        Grid *victim = nullptr;
        GridKey vk{0, 0.0};
        {
            std::lock_guard<std::mutex> lk(s->mu);
            if (s->in_use + need <= s->budget) break;
            for (auto &kv : s->grids)
                if (kv.second->users == 0 && (!victim || kv.second->last < victim->last)) {
                    victim = kv.second;
                    vk = kv.first;
                }
            if (!victim) break;  // built anyway: requests queue until memory frees
            s->in_use -= victim->bytes;
        }
        s->grids.erase(vk);
        charm_amd_grid_free(victim->g);
        delete victim;
    }
    charm_point *pnt = charm_crd_point_gl(nmax, r);
    if (!pnt) {
        CHARM_AMD_ERR(err, CHARM_EMEM, "Failed to build the Gauss--Legendre grid.");
        return nullptr;
    }
    Grid *g = new Grid();
    g->g = charm_amd_grid_create(pnt, err);
    g->nlat = pnt->nlat;
    g->nlon = pnt->nlon;
    g->nring = (pnt->nlat + 1) / 2;
    charm_crd_point_free(pnt);
    if (!g->g) {
        delete g;
        return nullptr;
    }
    g->bytes = need;
    g->last = now_ms();
    std::lock_guard<std::mutex> lk(s->mu);
    s->in_use += need;
    g->users = 1;
    s->grids[k] = g;
    return g;
}

void fail(SrvReq *q, int code, const char *msg)
{
    q->code = code;
    std::snprintf(q->msg, sizeof q->msg, "%s", msg);
    q->state = 3;
}

bool dalloc(SrvReq *q, hipStream_t st, size_t bytes, void **p)
{
    if (hipMallocAsync(p, std::max<size_t>(bytes, 16), st) != hipSuccess) {
        (void)hipGetLastError();
        return false;
    }
    q->dev.push_back(*p);
    return true;
}

bool groupable(const SrvReq *q)
{
    return q->r.kind == CHARM_AMD_SRV_SHS_SCTR && q->r.npoint > 0 &&
           q->r.npoint <= GROUP_MAX_POINTS;
}

// Enqueue a launch unit (one request, or a group of small scattered requests)
// on stream k.  Called by the worker without the server lock: the requests
// are out of the queue and belong to the worker.
void start(charm_amd_server *s, std::vector<SrvReq *> &unit, int k)
{
    hipStream_t st = s->streams[k];
    charm_err *err = charm_err_init();
    std::vector<SctrJob> jobs;
    std::vector<SrvReq *> sct;
    std::vector<Model> mds;  // copies: models may be added concurrently
    int inject = 0;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        for (SrvReq *q : unit) {
            q->stream = k;
            q->t_start = now_ms();
            const int mi = q->r.model;
            mds.push_back(mi >= 0 && mi < (int)s->models.size() ? s->models[mi] : Model());
        }
        if (s->fail_next > 0) {
            s->fail_next--;
            inject = s->fail_code ? s->fail_code : CHARM_EFUNCARG;
        }
    }
    bool ok = inject == 0;
    for (size_t u = 0; ok && u < unit.size(); u++) {
        SrvReq *q = unit[u];
        const charm_amd_srv_request &r = q->r;
        const Model *md = &mds[u];
        const size_t P = tri_len(r.nmax);
        if (r.kind == CHARM_AMD_SRV_SHS_SCTR) {
            const size_t n = r.npoint;
            void *lat, *lon, *rr, *f, *cs;
            ok = dalloc(q, st, n * 8, &lat) && dalloc(q, st, n * 8, &lon) &&
                 dalloc(q, st, n * 8, &rr) && dalloc(q, st, n * 8, &f) && dalloc(q, st, P * 16, &cs);
            if (!ok) break;
            const double *pin = (const double *)q->in->p;
            ok = hipMemcpyAsync(lat, pin, n * 8, hipMemcpyHostToDevice, st) == hipSuccess &&
                 hipMemcpyAsync(lon, pin + n, n * 8, hipMemcpyHostToDevice, st) == hipSuccess &&
                 hipMemcpyAsync(rr, pin + 2 * n, n * 8, hipMemcpyHostToDevice, st) == hipSuccess;
            charm_amd_cs_truncate(md->cs, md->nmax, (double *)cs, r.nmax, st, err);
            ok = ok && charm_err_isempty(err);
            jobs.push_back({n, (const double *)lat, (const double *)lon, (const double *)rr, r.nmax,
                            md->mu, md->R, (const double *)cs, (double *)f});
            sct.push_back(q);
            q->out_bytes = n * 8;
        } else if (r.kind == CHARM_AMD_SRV_SHS_GRID) {
            Grid *g = q->grid;
            void *f, *cs;
            ok = dalloc(q, st, g->nlat * g->nlon * 8, &f) && dalloc(q, st, P * 16, &cs);
            if (!ok) break;
            charm_amd_cs_truncate(md->cs, md->nmax, (double *)cs, r.nmax, st, err);
            charm_amd_shs_grid_dev(g->g, r.nmax, md->mu, md->R, (double *)cs, (double *)f, st, err);
            q->out_bytes = g->nlat * g->nlon * 8;
            ok = charm_err_isempty(err) &&
                 hipMemcpyAsync(q->out->p, f, q->out_bytes, hipMemcpyDeviceToHost, st) == hipSuccess;
        } else {
            Grid *g = q->grid;
            void *f, *cs;
            ok = dalloc(q, st, g->nlat * g->nlon * 8, &f) && dalloc(q, st, P * 16, &cs);
            if (!ok) break;
            ok = hipMemcpyAsync(f, q->in->p, g->nlat * g->nlon * 8, hipMemcpyHostToDevice, st) ==
                 hipSuccess;
            charm_amd_sha_grid_dev(g->g, r.nmax, r.mu, r.R, (double *)f, (double *)cs, st, err);
            q->out_bytes = P * 16;
            ok = ok && charm_err_isempty(err) &&
                 hipMemcpyAsync(q->out->p, cs, q->out_bytes, hipMemcpyDeviceToHost, st) == hipSuccess;
        }
    }
    if (ok && !jobs.empty()) {
        // one request: the plain device call; several: one persistent launch
        if (jobs.size() == 1) {
            const SctrJob &j = jobs[0];
            charm_amd_shs_sctr_dev(j.npoint, j.lat, j.lon, j.r, j.nmax, j.mu, j.R, j.cs, j.f, st,
                                   err);
            ok = charm_err_isempty(err);
        } else {
            ok = run_sctr_jobs(jobs.data(), (int)jobs.size(), st) == hipSuccess;
        }
        for (size_t j = 0; ok && j < sct.size(); j++)
            ok = hipMemcpyAsync(sct[j]->out->p, jobs[j].f, sct[j]->out_bytes,
                                hipMemcpyDeviceToHost, st) == hipSuccess;
    }
    if (ok) ok = hipEventRecord(unit.front()->ev, st) == hipSuccess;
    if (!ok) (void)hipStreamSynchronize(st);
    std::lock_guard<std::mutex> lk(s->mu);
    for (SrvReq *q : unit) {
        if (ok)
            q->state = 1;
        else if (inject)
            fail(q, inject, "Simulated device error (charm_amd_server_inject_error).");
        else
            fail(q, charm_err_isempty(err) ? CHARM_EMEM : err->code,
                 charm_err_isempty(err) ? "Failed to run the request on the device." : err->msg);
    }
    s->stats.launches++;
    if (unit.size() > 1) {
        s->stats.grouped += unit.size();
        s->stats.max_group = std::max<unsigned long>(s->stats.max_group, unit.size());
    }
    charm_err_free(err);
}

// Completion of a launch unit, with the server lock held
void finish(charm_amd_server *s, std::vector<SrvReq *> &unit)
{
    hipStream_t st = s->streams[unit.front()->stream];
    for (SrvReq *q : unit) {
        for (void *p : q->dev) (void)hipFreeAsync(p, st);
        q->dev.clear();
        pinned_put(s, q->in);
        q->in = nullptr;
        if (q->grid) q->grid->users--;
        q->grid = nullptr;
        s->in_use -= q->bytes;
        if (q->state == 1) {
            if (q->cancel) {
                q->state = 4;
                s->stats.cancelled++;
            } else {
                q->state = 2;
                s->stats.completed++;
            }
        } else {
            s->stats.failed++;
        }
        if (q->state != 2) {
            pinned_put(s, q->out);
            q->out = nullptr;
        }
        q->t_done = now_ms();
    }
    s->stream_busy[unit.front()->stream] = false;
}

// Admission test of a request of `bytes`, with the server lock held.  The
// library's cached scratch counts against the budget: when it would not fit
// beside the admitted requests, *trim_to receives the size to trim the cache
// to, which the worker does after releasing the lock (trimming waits for the
// buffers' last uses and frees device memory)
bool fits(charm_amd_server *s, size_t bytes, size_t *trim_to)
{
    if (s->in_use + bytes > s->budget) return false;
    const size_t room = s->budget - s->in_use - bytes;
    if (scratch_cached_bytes() > room) *trim_to = std::min(*trim_to, room);
    return true;
}

void worker_loop(charm_amd_server *s)
{
    set_thread_tuning(false);  // never time kernels against requests in flight
    std::unique_lock<std::mutex> lk(s->mu);
    while (true) {
        bool progress = false;
        // retire finished launch units
        for (size_t i = 0; i < s->inflight.size();) {
            auto &unit = s->inflight[i];
            SrvReq *lead = unit.front();
            if (lead->state == 3 || hipEventQuery(lead->ev) == hipSuccess) {
                finish(s, unit);
                s->inflight.erase(s->inflight.begin() + i);
                progress = true;
            } else {
                i++;
            }
        }
        if (progress) s->cv_done.notify_all();
        // admit: in arrival order; a request that does not fit yet lets later
        // ones pass, at most BYPASS_MAX of them; small scattered requests that
        // fit together form one group launch (one group per pass)
        std::vector<std::vector<SrvReq *>> launch;
        std::vector<int> launch_stream;
        std::vector<SrvReq *> skipped;
        size_t trim_to = SIZE_MAX;
        int nfree = 0;
        for (bool b : s->stream_busy) nfree += !b;
        for (auto it = s->pending.begin(); it != s->pending.end() && nfree > 0;) {
            SrvReq *q = *it;
            bool starving = false;
            for (SrvReq *p : skipped) starving = starving || p->bypassed >= BYPASS_MAX;
            if (starving) break;
            // with nothing in flight a request runs whatever the budget says
            // (its own bytes were checked against it at submission)
            if (!fits(s, q->bytes, &trim_to) && !(s->inflight.empty() && launch.empty())) {
                skipped.push_back(q);
                ++it;
                continue;
            }
            std::vector<SrvReq *> unit{q};
            s->in_use += q->bytes;
            q->state = 1;  // out of the queue: cancelling now discards the result
            it = s->pending.erase(it);
            const bool group = groupable(q);
            if (group)
                for (auto jt = it; jt != s->pending.end() && (int)unit.size() < GROUP_MAX;) {
                    SrvReq *p = *jt;
                    if (groupable(p) && fits(s, p->bytes, &trim_to)) {
                        s->in_use += p->bytes;
                        p->state = 1;
                        unit.push_back(p);
                        jt = s->pending.erase(jt);
                    } else {
                        ++jt;
                    }
                }
            for (SrvReq *p : skipped) p->bypassed += (int)unit.size();
            int k = 0;
            while (s->stream_busy[k]) k++;
            s->stream_busy[k] = true;
            nfree--;
            launch.push_back(unit);
            launch_stream.push_back(k);
            if (group) break;  // the group scan may have erased what `it` points at
        }
        if (!launch.empty()) {
            lk.unlock();
            if (trim_to != SIZE_MAX) scratch_trim(trim_to);
            for (size_t u = 0; u < launch.size(); u++) {
                try {
                    start(s, launch[u], launch_stream[u]);
                } catch (...) {  // a failed host allocation: the unit fails, the worker goes on
                    std::lock_guard<std::mutex> g(s->mu);
                    for (SrvReq *q : launch[u]) {
                        q->stream = launch_stream[u];
                        fail(q, CHARM_EMEM, "Failed to start the request on the host.");
                    }
                }
            }
            lk.lock();
            for (auto &unit : launch) s->inflight.push_back(unit);
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

void free_req(SrvReq *q)
{
    if (q->ev) (void)hipEventDestroy(q->ev);
    delete q;
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

static int server_add_model(charm_amd_server *s, const charm_shc *shcs, unsigned long nmax,
                            charm_err *err);

int charm_amd_server_add_model(charm_amd_server *s, const charm_shc *shcs, unsigned long nmax,
                               charm_err *err)
{
    try {
        return server_add_model(s, shcs, nmax, err);
    } catch (...) {
        CHARM_AMD_ERR(err, CHARM_EMEM, "Failed to add the model (host memory).");
        return -1;
    }
}

static int server_add_model(charm_amd_server *s, const charm_shc *shcs, unsigned long nmax,
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
    Model md;
    md.nmax = nmax;
    md.mu = shcs->mu;
    md.R = shcs->r;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        if (s->in_use + P * 16 > s->budget) {
            CHARM_AMD_ERR(err, CHARM_EMEM, "The model does not fit the HBM budget.");
            return -1;
        }
        s->in_use += P * 16;
    }
    if (hipMalloc((void **)&md.cs, P * 16) != hipSuccess ||
        hipMemcpy(md.cs, cs.data(), P * 16, hipMemcpyHostToDevice) != hipSuccess) {
        (void)hipGetLastError();
        std::lock_guard<std::mutex> lk(s->mu);
        s->in_use -= P * 16;
        CHARM_AMD_ERR(err, CHARM_EMEM, "Failed to make the model resident.");
        return -1;
    }
    std::lock_guard<std::mutex> lk(s->mu);
    s->models.push_back(md);
    return (int)s->models.size() - 1;
}

static long server_submit(charm_amd_server *s, const charm_amd_srv_request *r, charm_err *err);

long charm_amd_server_submit(charm_amd_server *s, const charm_amd_srv_request *r, charm_err *err)
{
    try {
        return server_submit(s, r, err);
    } catch (...) {
        CHARM_AMD_ERR(err, CHARM_EMEM, "Failed to queue the request (host memory).");
        return -1;
    }
}

static long server_submit(charm_amd_server *s, const charm_amd_srv_request *r, charm_err *err)
{
    if (!s || !r || !r->out ||
        (r->kind != CHARM_AMD_SRV_SHS_SCTR && r->kind != CHARM_AMD_SRV_SHS_GRID &&
         r->kind != CHARM_AMD_SRV_SHA_GRID)) {
        CHARM_AMD_ERR(err, CHARM_EFUNCARG, "Invalid request.");
        return -1;
    }
    {
        std::lock_guard<std::mutex> lk(s->mu);
        const bool shs = r->kind != CHARM_AMD_SRV_SHA_GRID;
        const Model *md = (r->model >= 0 && r->model < (int)s->models.size()) ? &s->models[r->model]
                                                                              : nullptr;
        const bool bad = (shs && (!md || r->nmax > md->nmax)) ||
                         (r->kind == CHARM_AMD_SRV_SHS_SCTR &&
                          (!r->lat || !r->lon || !r->r || r->npoint >= (size_t)INT32_MAX)) ||
                         (r->kind != CHARM_AMD_SRV_SHS_SCTR && r->nmax > r->grid_nmax) ||
                         (r->kind == CHARM_AMD_SRV_SHA_GRID && !r->f_in);
        if (bad) {
            CHARM_AMD_ERR(err, CHARM_EFUNCARG, "Invalid request.");
            return -1;
        }
    }
    SrvReq *q = new SrvReq();
    q->r = *r;
    q->t_submit = now_ms();
    q->bytes = footprint(*r);
    if (q->bytes > s->budget) {
        delete q;
        CHARM_AMD_ERR(err, CHARM_EMEM, "The request exceeds the HBM budget.");
        return -1;
    }
    if (hipEventCreateWithFlags(&q->ev, hipEventDisableTiming) != hipSuccess) {
        (void)hipGetLastError();
        q->ev = nullptr;
        free_req(q);
        CHARM_AMD_ERR(err, CHARM_EMEM, "Failed to create an event.");
        return -1;
    }
    // host work in the submitting thread: inputs into pinned staging, output
    // staging, the resident grid
    bool ok = true;
    if (r->kind == CHARM_AMD_SRV_SHS_SCTR) {
        const size_t n = r->npoint;
        q->in = pinned_get(s, 3 * n * sizeof(double));
        q->out = pinned_get(s, n * sizeof(double));
        ok = q->in && q->out;
        if (ok) {
            double *pin = (double *)q->in->p;
            std::memcpy(pin, r->lat, n * 8);
            std::memcpy(pin + n, r->lon, n * 8);
            std::memcpy(pin + 2 * n, r->r, n * 8);
        }
    } else {
        q->grid = acquire_grid(s, r->grid_nmax, r->r_grid, err);
        ok = q->grid != nullptr;
        if (ok) {
            const size_t gb = q->grid->nlat * q->grid->nlon * sizeof(double);
            if (r->kind == CHARM_AMD_SRV_SHA_GRID) {
                q->in = pinned_get(s, gb);
                q->out = pinned_get(s, tri_len(r->nmax) * 16);
                ok = q->in && q->out;
                if (ok) std::memcpy(q->in->p, r->f_in, gb);
            } else {
                q->out = pinned_get(s, gb);
                ok = q->out != nullptr;
            }
        }
    }
    if (!ok) {
        if (charm_err_isempty(err))
            CHARM_AMD_ERR(err, CHARM_EMEM, "Failed to stage the request in host memory.");
        pinned_put(s, q->in);
        pinned_put(s, q->out);
        if (q->grid) {
            std::lock_guard<std::mutex> lk(s->mu);
            q->grid->users--;
        }
        free_req(q);
        return -1;
    }
    std::lock_guard<std::mutex> lk(s->mu);
    q->ticket = s->next_ticket++;
    s->all[q->ticket] = q;
    s->pending.push_back(q);
    s->stats.submitted++;
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
    s->all.erase(it);
    lk.unlock();
    if (latency_ms) *latency_ms = q->t_done - q->t_submit;
    int rc = 0;
    if (q->state == 2) {
        // the result leaves pinned staging in the waiting thread
        std::memcpy(q->r.out, q->out->p, q->out_bytes);
        pinned_put(s, q->out);
        q->out = nullptr;
    } else if (q->state == 4) {
        CHARM_AMD_ERR(err, CHARM_EFUNCARG, "The request was cancelled.");
        rc = -1;
    } else {
        CHARM_AMD_ERR(err, q->code ? q->code : CHARM_EFUNCARG, q->msg);
        rc = -1;
    }
    free_req(q);
    return rc;
}

int charm_amd_server_cancel(charm_amd_server *s, long ticket, charm_err *err)
{
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->all.find(ticket);
    if (it == s->all.end()) {
        CHARM_AMD_ERR(err, CHARM_EFUNCARG, "Unknown ticket.");
        return -1;
    }
    SrvReq *q = it->second;
    if (q->state >= 2) return 1;  // already completed
    if (q->state == 0) {
        auto pq = std::find(s->pending.begin(), s->pending.end(), q);
        if (pq != s->pending.end()) s->pending.erase(pq);
        pinned_put(s, q->in);
        pinned_put(s, q->out);
        q->in = q->out = nullptr;
        if (q->grid) q->grid->users--;
        q->grid = nullptr;
        q->state = 4;
        q->t_done = now_ms();
        s->stats.cancelled++;
        s->cv_done.notify_all();
        return 0;
    }
    q->cancel = true;  // in flight: the result is discarded when it completes
    return 0;
}

void charm_amd_server_inject_error(charm_amd_server *s, int count, int code)
{
    std::lock_guard<std::mutex> lk(s->mu);
    s->fail_next = std::max(0, count);
    s->fail_code = code;
}

void charm_amd_server_get_stats(charm_amd_server *s, charm_amd_srv_stats *st)
{
    std::lock_guard<std::mutex> lk(s->mu);
    *st = s->stats;
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
    for (auto &kv : s->all) free_req(kv.second);
    for (auto &g : s->grids) {
        charm_amd_grid_free(g.second->g);
        delete g.second;
    }
    for (auto &m : s->models) (void)hipFree(m.cs);
    for (Pinned *p : s->pinned) {
        (void)hipHostFree(p->p);
        delete p;
    }
    for (auto st : s->streams) (void)hipStreamDestroy(st);
    delete s;
}

}  // extern "C"
