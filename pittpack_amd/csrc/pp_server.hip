// pp_server.hip -- serving form: a request loop that admits solves of mixed sizes, boundary
// conditions and precisions by HBM footprint and runs them as batches (pp_batch_core.h)
// while earlier batches are still on the GPU.
//
// No counterpart in the reference (one grid per PencilDcmp object, src/pencilDcmp.cpp).
//
// Design: callers submit requests from any thread; one scheduler thread owns every HIP call.
// Each request is planned alone at submission, which gives its exact device footprint (a
// batch's region is at most the sum of its members' single-member regions, since every part
// of the plan is a sum over members and each part is rounded up once).  The scheduler forms
// a batch per precision from the queued requests, in arrival order, whose footprints fit the
// free part of one device pool (first fit), plans and binds it, uploads its tables and
// right-hand sides, launches it on a free stream of its own and records an event; it retires
// finished batches by polling their events.  A request that does not fit may be overtaken
// by later, smaller ones only a bounded number of times (PP_OPT_MAX_OVERTAKE); after that
// nothing behind it is admitted until it runs, so a large request cannot starve under a
// stream of small ones.  Twiddle tables stay resident for the server's lifetime;
// per-request eigenvalue and boundary tables are planned on the host and uploaded with the
// batch.  Solutions go device -> pinned staging on the batch's stream; the waiting caller
// copies them into its own buffer, so the scheduler never blocks on a host copy.  A member's
// arithmetic does not depend on its batch, so a result is the same alone and in any mix.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <list>
#include <map>
#include <mutex>
#include <new>
#include <thread>
#include <vector>

#include "pittpack_hip.h"
#include "pp_batch_core.h"
#include "pp_launch.h"

using namespace pp;

namespace {

double now_ms()
{
    using namespace std::chrono;
    return duration<double, std::milli>(steady_clock::now().time_since_epoch()).count();
}

// first-fit allocator over [0, cap) in 256-byte units
class RangePool {
  public:
    explicit RangePool(size_t cap = 0) { reset(cap); }
    void reset(size_t cap)
    {
        free_.clear();
        cap_ = cap / 256 * 256;
        if (cap_) free_[0] = cap_;
        used_ = 0;
    }
    bool alloc(size_t bytes, size_t &off)
    {
        bytes = (bytes + 255) / 256 * 256;
        if (bytes == 0) bytes = 256;
        for (auto it = free_.begin(); it != free_.end(); ++it) {
            if (it->second >= bytes) {
                off = it->first;
                const size_t rest = it->second - bytes;
                free_.erase(it);
                if (rest) free_[off + bytes] = rest;
                sizes_[off] = bytes;
                used_ += bytes;
                return true;
            }
        }
        return false;
    }
    void release(size_t off)
    {
        auto s = sizes_.find(off);
        if (s == sizes_.end()) return;
        size_t bytes = s->second;
        sizes_.erase(s);
        used_ -= bytes;
        auto next = free_.lower_bound(off);
        if (next != free_.end() && off + bytes == next->first) {
            bytes += next->second;
            next = free_.erase(next);
        }
        if (next != free_.begin()) {
            auto prev = std::prev(next);
            if (prev->first + prev->second == off) {
                prev->second += bytes;
                return;
            }
        }
        free_[off] = bytes;
    }
    size_t used() const { return used_; }
    size_t cap() const { return cap_; }

  private:
    std::map<size_t, size_t> free_;   // offset -> length
    std::map<size_t, size_t> sizes_;  // allocated offset -> length
    size_t cap_ = 0, used_ = 0;
};

struct Request {
    long long ticket = 0;
    int nx = 0, ny = 0, nz = 0;
    char bc[6] = {0, 0, 0, 0, 0, 0};
    double X[6] = {0, 1, 0, 1, 0, 1};
    int precision = 0;
    const double *rhs = nullptr;  // host right-hand side (null: counter generator with `seed`)
    unsigned long long seed = 0;
    bool want = false;            // copy the solution back to the host
    size_t need = 0;              // device footprint: the region of this request planned alone
    int overtaken = 0;            // batches admitted past this request while it did not fit
    long long batch_seq = -1;     // launch number of the batch that ran it
    double t_submit = 0, t_launch = 0, t_done = 0;
    int status = 1;               // 1 queued, 2 running, 0 done, >1 error code (negated below)
    int rc = PP_SUCCESS;
    size_t stage_off = 0;         // pinned staging of the solution
    bool staged = false;
    long long points() const { return (long long)nx * ny * nz; }
};

struct InFlight {
    BatchCore core;
    size_t off = 0;
    int slot = -1;
    hipEvent_t done = nullptr;
    std::vector<Request *> reqs;
};

}  // namespace

struct pp_server_s {
    char *dev = nullptr;       // device pool
    char *stage = nullptr;     // pinned host staging for solutions
    RangePool dpool, spool;
    TableCache tables;
    std::vector<hipStream_t> streams;
    std::vector<bool> slot_busy;
    std::vector<hipEvent_t> events;
    int max_members = 256;
    int refine = 2;
    int max_overtake = 4;
    int group_cap = 1 << 30;  // halved when a group does not fit the empty pool
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::deque<Request *> queue;
    std::map<long long, Request *> reqs;
    std::list<InFlight *> inflight;
    long long next_ticket = 1;
    bool stop = false;
    std::thread worker;
    // statistics
    long long completed = 0, points_done = 0, batches = 0;
    double first_submit = 0, last_done = 0;
    int max_concurrent = 0;
};

namespace {

// device footprint of one request: its region planned as a batch of one (the scheduler's
// batches use the same refinement setting; a batch's region never exceeds the sum of its
// members' regions).  Returns a PittPack code for a request that cannot be planned.
int request_bytes(const Request *r, int refine, size_t &bytes)
{
    BatchCore c;
    c.precision = r->precision;
    c.refine = refine;
    int id = -1;
    int rc = c.add(r->nx, r->ny, r->nz, r->bc, r->X, &id);
    if (!rc) rc = c.plan();
    bytes = rc ? 0 : c.region_bytes;
    return rc;
}

void fail_request(pp_server_s *s, Request *r, int rc)
{
    r->rc = rc;
    r->status = 0;
    r->t_done = now_ms();
    s->cv_done.notify_all();
}

// retire finished batches (caller holds no lock)
bool retire(pp_server_s *s)
{
    bool any = false;
    for (auto it = s->inflight.begin(); it != s->inflight.end();) {
        InFlight *f = *it;
        const hipError_t q = hipEventQuery(f->done);
        if (q == hipErrorNotReady) {
            ++it;
            continue;
        }
        const double t = now_ms();
        {
            std::lock_guard<std::mutex> lk(s->mu);
            for (Request *r : f->reqs) {
                r->rc = (q == hipSuccess) ? PP_SUCCESS : PP_HIP_ERROR;
                r->status = 0;
                r->t_done = t;
                s->completed++;
                s->points_done += r->points();
            }
            s->dpool.release(f->off);
            s->slot_busy[f->slot] = false;
            s->last_done = t;
        }
        s->cv_done.notify_all();
        delete f;
        it = s->inflight.erase(it);
        any = true;
    }
    return any;
}

// launch one batch of requests of one precision; returns false when the pools cannot take it
// (the requests are then left for a later attempt)
bool launch(pp_server_s *s, std::vector<Request *> &group, int slot)
{
    InFlight *f = new (std::nothrow) InFlight();
    if (!f) return false;
    f->core.precision = group[0]->precision;
    f->core.refine = s->refine;
    std::vector<Request *> ok;
    for (Request *r : group) {
        int id = -1;
        const int rc = f->core.add(r->nx, r->ny, r->nz, r->bc, r->X, &id);
        if (rc) {
            std::lock_guard<std::mutex> lk(s->mu);
            fail_request(s, r, rc);
        } else {
            ok.push_back(r);
        }
    }
    if (ok.empty()) {
        delete f;
        return true;
    }
    int rc = f->core.plan();
    size_t off = 0;
    bool fits;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        fits = !rc && s->dpool.alloc(f->core.region_bytes, off);
    }
    if (rc || !fits) {
        delete f;
        if (rc) {
            std::lock_guard<std::mutex> lk(s->mu);
            for (Request *r : ok) fail_request(s, r, rc);
            return true;
        }
        return false;
    }
    f->off = off;
    f->slot = slot;
    f->done = s->events[slot];
    f->reqs = ok;
    hipStream_t st = s->streams[slot];
    rc = f->core.bind(s->dev + off, &s->tables);
    if (!rc) rc = f->core.upload(st);
    {
        // callers read a request's status under the lock
        std::lock_guard<std::mutex> lk(s->mu);
        const double t = now_ms();
        for (Request *r : ok) {
            r->t_launch = t;
            r->status = 2;
            r->batch_seq = s->batches;
        }
    }
    for (size_t i = 0; !rc && i < ok.size(); i++) {
        Request *r = ok[i];
        double *d = f->core.d_rhs + f->core.mem[i].offset;
        const size_t bytes = (size_t)r->points() * sizeof(double);
        if (r->rhs) rc = hipMemcpyAsync(d, r->rhs, bytes, hipMemcpyHostToDevice, st) == hipSuccess ? 0 : PP_HIP_ERROR;
        else rc = launch_fill_uniform(d, (long)r->points(), r->seed, 0, st) == hipSuccess ? 0 : PP_HIP_ERROR;
    }
    if (!rc) rc = f->core.run(st, nullptr);
    for (size_t i = 0; !rc && i < ok.size(); i++) {
        Request *r = ok[i];
        if (!r->staged) continue;
        const double *d = f->core.d_sol + f->core.mem[i].offset;
        rc = hipMemcpyAsync(s->stage + r->stage_off, d, (size_t)r->points() * sizeof(double), hipMemcpyDeviceToHost,
                            st) == hipSuccess
                 ? 0
                 : PP_HIP_ERROR;
    }
    if (!rc) rc = hipEventRecord(f->done, st) == hipSuccess ? 0 : PP_HIP_ERROR;
    if (rc) {
        (void)hipStreamSynchronize(st);
        std::lock_guard<std::mutex> lk(s->mu);
        for (Request *r : ok) fail_request(s, r, rc);
        s->dpool.release(off);
        delete f;
        return true;
    }
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->slot_busy[slot] = true;
        s->batches++;
    }
    s->inflight.push_back(f);
    if ((int)s->inflight.size() > s->max_concurrent) s->max_concurrent = (int)s->inflight.size();
    return true;
}

void worker_loop(pp_server_s *s)
{
    (void)hipSetDevice(0);
    for (;;) {
        bool progressed = retire(s);
        // admission: the queued requests (in arrival order) whose footprint fits, one batch per
        // precision, each on a free stream
        std::vector<Request *> groups[2];
        int slots[2] = {-1, -1};
        {
            std::unique_lock<std::mutex> lk(s->mu);
            if (s->queue.empty() && s->inflight.empty()) {
                if (s->stop) return;
                s->cv_work.wait(lk, [&] { return s->stop || !s->queue.empty(); });
                if (s->stop && s->queue.empty()) return;
            }
            size_t budget = s->dpool.cap() - s->dpool.used();
            // requests left behind this round because they did not fit; once one of them has
            // been overtaken max_overtake times, nothing behind it is admitted
            std::vector<Request *> skipped;
            bool blocked = false;
            for (int p = 0; p < 2 && !blocked; p++) {
                int slot = -1;
                for (int k = 0; k < (int)s->slot_busy.size(); k++)
                    if (!s->slot_busy[k] && k != slots[0]) {
                        slot = k;
                        break;
                    }
                if (slot < 0) break;
                const int cap = std::min(s->max_members, s->group_cap);
                for (auto it = s->queue.begin(); it != s->queue.end() && (int)groups[p].size() < cap;) {
                    Request *r = *it;
                    if (r->precision != p) {
                        ++it;
                        continue;
                    }
                    bool fits = r->need <= budget;
                    // solution staging (pinned host) for callers that want the result
                    if (fits && r->want && !r->staged) {
                        if (s->spool.alloc((size_t)r->points() * sizeof(double), r->stage_off)) r->staged = true;
                        else fits = false;
                    }
                    if (!fits) {
                        if (r->overtaken >= s->max_overtake) {
                            blocked = true;
                            break;
                        }
                        skipped.push_back(r);
                        ++it;
                        continue;
                    }
                    budget -= r->need;
                    groups[p].push_back(r);
                    it = s->queue.erase(it);
                }
                if (!groups[p].empty()) slots[p] = slot;
            }
            // each admitted batch holding a later request overtakes a skipped one once
            for (Request *r : skipped)
                for (int p = 0; p < 2; p++)
                    if (!groups[p].empty() && groups[p].back()->ticket > r->ticket) r->overtaken++;
        }
        for (int p = 0; p < 2; p++) {
            if (groups[p].empty()) continue;
            if (launch(s, groups[p], slots[p])) {
                progressed = true;
                std::lock_guard<std::mutex> lk(s->mu);
                s->group_cap = 1 << 30;
                continue;
            }
            std::lock_guard<std::mutex> lk(s->mu);
            if (s->inflight.empty() && groups[p].size() == 1) {
                // alone in an empty pool and still not placed: it can never run
                fail_request(s, groups[p][0], PP_ALLOCATION_FAIL);
                progressed = true;
                continue;
            }
            // the free space is fragmented by batches in flight (retry after they retire), or
            // -- which the footprint bound rules out -- a group is too large for the empty
            // pool (the next groups are then halved until they fit)
            if (s->inflight.empty()) s->group_cap = std::max(1, (int)groups[p].size() / 2);
            for (auto it = groups[p].rbegin(); it != groups[p].rend(); ++it) s->queue.push_front(*it);
        }
        if (!progressed) std::this_thread::sleep_for(std::chrono::microseconds(20));
    }
}

}  // namespace

extern "C" {

int pp_server_create(long long hbm_bytes, long long staging_bytes, int streams, pp_server *out)
{
    if (!out || hbm_bytes <= 0 || staging_bytes < 0 || streams < 1 || streams > 16) return PP_INPUT_ARGS;
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return PP_NO_DEVICE;
    pp_server_s *s = new (std::nothrow) pp_server_s();
    if (!s) return PP_ALLOCATION_FAIL;
    auto fail = [&](int rc) {
        if (s->dev) (void)hipFree(s->dev);
        if (s->stage) (void)hipHostFree(s->stage);
        for (auto st : s->streams) (void)hipStreamDestroy(st);
        for (auto e : s->events) (void)hipEventDestroy(e);
        delete s;
        return rc;
    };
    if (hipMalloc((void **)&s->dev, (size_t)hbm_bytes) != hipSuccess) {
        s->dev = nullptr;
        return fail(PP_ALLOCATION_FAIL);
    }
    s->dpool.reset((size_t)hbm_bytes);
    if (staging_bytes > 0) {
        if (hipHostMalloc((void **)&s->stage, (size_t)staging_bytes, hipHostMallocDefault) != hipSuccess) {
            s->stage = nullptr;
            return fail(PP_ALLOCATION_FAIL);
        }
    }
    s->spool.reset((size_t)staging_bytes);
    for (int k = 0; k < streams; k++) {
        hipStream_t st;
        hipEvent_t e;
        if (hipStreamCreateWithFlags(&st, hipStreamNonBlocking) != hipSuccess) return fail(PP_HIP_ERROR);
        s->streams.push_back(st);
        if (hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess) return fail(PP_HIP_ERROR);
        s->events.push_back(e);
    }
    s->slot_busy.assign(streams, false);
    // resident twiddle tables for every compiled line length (the arbitrary-length tables are
    // made on first use and stay resident too)
    for (int n = 16; n <= 1024; n += 2)
        if (line_compiled(n)) {
            AxisDev t;
            const int rc = s->tables.get(n, t);
            if (rc) return fail(rc);
        }
    s->worker = std::thread(worker_loop, s);
    *out = s;
    return PP_SUCCESS;
}

int pp_server_set_option(pp_server s, int key, long long value)
{
    if (!s) return PP_INVALID_HANDLE;
    std::lock_guard<std::mutex> lk(s->mu);
    if (key == PP_OPT_REFINE) {
        if (value < 0 || value > 16) return PP_INPUT_ARGS;
        s->refine = (int)value;
        return PP_SUCCESS;
    }
    if (key == PP_OPT_BATCH_MEMBERS) {
        if (value < 1 || value > 1 << 20) return PP_INPUT_ARGS;
        s->max_members = (int)value;
        return PP_SUCCESS;
    }
    if (key == PP_OPT_MAX_OVERTAKE) {
        if (value < 0 || value > 1 << 20) return PP_INPUT_ARGS;
        s->max_overtake = (int)value;
        return PP_SUCCESS;
    }
    return PP_INPUT_ARGS;
}

int pp_server_submit(pp_server s, int nx, int ny, int nz, const char *bc, const double *X, int precision,
                     const double *rhs, unsigned long long seed, int want_solution, long long *ticket)
{
    if (!s) return PP_INVALID_HANDLE;
    if (!bc || !X || !ticket || (precision != 0 && precision != 1)) return PP_INPUT_ARGS;
    Request *r = new (std::nothrow) Request();
    if (!r) return PP_ALLOCATION_FAIL;
    r->nx = nx;
    r->ny = ny;
    r->nz = nz;
    for (int i = 0; i < 6; i++) {
        r->bc[i] = bc[i];
        r->X[i] = X[i];
    }
    r->precision = precision;
    r->rhs = rhs;
    r->seed = seed;
    r->want = want_solution != 0;
    int refine;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        refine = s->refine;
    }
    // plan the request alone: unsupported sizes or boundary conditions are returned here, and
    // what can never be admitted is rejected
    const int prc = request_bytes(r, refine, r->need);
    if (prc) {
        delete r;
        return prc;
    }
    if (r->need > s->dpool.cap() || (r->want && (size_t)r->points() * sizeof(double) > s->spool.cap())) {
        delete r;
        return PP_ALLOCATION_FAIL;
    }
    {
        std::lock_guard<std::mutex> lk(s->mu);
        r->ticket = s->next_ticket++;
        r->t_submit = now_ms();
        if (s->completed == 0 && s->reqs.empty()) s->first_submit = r->t_submit;
        s->reqs[r->ticket] = r;
        s->queue.push_back(r);
        *ticket = r->ticket;
    }
    s->cv_work.notify_one();
    return PP_SUCCESS;
}

int pp_server_footprint(pp_server s, int nx, int ny, int nz, const char *bc, const double *X, int precision,
                        long long *bytes)
{
    if (!s) return PP_INVALID_HANDLE;
    if (!bc || !X || !bytes || (precision != 0 && precision != 1)) return PP_INPUT_ARGS;
    Request r;
    r.nx = nx;
    r.ny = ny;
    r.nz = nz;
    for (int i = 0; i < 6; i++) {
        r.bc[i] = bc[i];
        r.X[i] = X[i];
    }
    r.precision = precision;
    int refine;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        refine = s->refine;
    }
    size_t b = 0;
    const int rc = request_bytes(&r, refine, b);
    *bytes = (long long)b;
    return rc;
}

int pp_server_wait(pp_server s, long long ticket, double *u, double *latency_ms)
{
    return pp_server_wait_ex(s, ticket, u, latency_ms, nullptr);
}

// blocks until the request is done; copies the solution (nx*ny*nz doubles) into u when it
// was requested; latency_ms = submit -> done (scheduler clock); batch_seq = launch number of
// the batch that ran it (-1 if it never ran)
int pp_server_wait_ex(pp_server s, long long ticket, double *u, double *latency_ms, long long *batch_seq)
{
    if (!s) return PP_INVALID_HANDLE;
    Request *r;
    {
        std::unique_lock<std::mutex> lk(s->mu);
        auto it = s->reqs.find(ticket);
        if (it == s->reqs.end()) return PP_INPUT_ARGS;
        r = it->second;
        s->cv_done.wait(lk, [&] { return r->status == 0; });
        s->reqs.erase(it);
    }
    const int rc = r->rc;
    if (!rc && r->staged && u) std::memcpy(u, s->stage + r->stage_off, (size_t)r->points() * sizeof(double));
    if (latency_ms) *latency_ms = r->t_done - r->t_submit;
    if (batch_seq) *batch_seq = r->batch_seq;
    if (r->staged) {
        std::lock_guard<std::mutex> lk(s->mu);
        s->spool.release(r->stage_off);
    }
    delete r;
    return rc;
}

int pp_server_stats(pp_server s, long long *completed, long long *points, long long *batches, double *span_ms,
                    int *max_concurrent)
{
    if (!s) return PP_INVALID_HANDLE;
    std::lock_guard<std::mutex> lk(s->mu);
    if (completed) *completed = s->completed;
    if (points) *points = s->points_done;
    if (batches) *batches = s->batches;
    if (span_ms) *span_ms = s->last_done - s->first_submit;
    if (max_concurrent) *max_concurrent = s->max_concurrent;
    return PP_SUCCESS;
}

int pp_server_destroy(pp_server s)
{
    if (!s) return PP_INVALID_HANDLE;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->stop = true;
    }
    s->cv_work.notify_all();
    if (s->worker.joinable()) s->worker.join();
    for (auto st : s->streams) {
        (void)hipStreamSynchronize(st);
        (void)hipStreamDestroy(st);
    }
    for (auto e : s->events) (void)hipEventDestroy(e);
    for (auto &kv : s->reqs) delete kv.second;
    if (s->dev) (void)hipFree(s->dev);
    if (s->stage) (void)hipHostFree(s->stage);
    delete s;
    return PP_SUCCESS;
}

}  // extern "C"
