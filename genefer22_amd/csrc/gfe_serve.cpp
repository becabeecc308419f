// gfe_serve.cpp -- the serving loop (include/gfe_serve.h): continuous batching of candidate
// requests of mixed n over pools of gfe_batch, admission by device-memory footprint, and a
// lone-request path on warm single-candidate engines.
//
// Host code only: it drives the batch and single-candidate engines through their C ABI and owns no
// device memory itself (hipMemGetInfo for the default budget is its only HIP call).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cstring>
#include <deque>
#include <memory>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "../../include/gfe.h"
#include "../../include/gfe_batch.h"
#include "../../include/gfe_serve.h"
#include "gfe_engine.h"
#include "genefer.h"

namespace {

double steady_s() {
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

// Default slots of a pool of n: about 2^26 digits per pool (enough candidates to fill the device
// with one pool), between 16 and 1024 slots.
uint32_t default_slots(uint32_t n, uint32_t depth = 0) {
    const uint64_t s = (n >= 26) ? 1 : (uint64_t(1) << (26 - n));
    const uint32_t q = (uint32_t)std::min<uint64_t>(1024, std::max<uint64_t>(16, s));
    // proof pools: a slot holds 3 + 2^depth registers instead of 3; keep the pool's registers near
    // a quick-test pool's (at least 4 slots)
    return depth == 0 ? q : std::max<uint32_t>(4, q * 3 / (3 + (1u << depth)));
}

struct Request {
    uint64_t tag;
    uint32_t n, b;
    uint64_t iters;
    uint32_t depth;   // Pietrzak-Li proof depth (0: quick test)
    double t_arrival;
};

struct Admitted {
    Request rq;
    double t_admit;
};

// A warm single-candidate engine of the lone-request path, for one (n, prime set computed in) and
// any b of that set (gfe_rebase); one created with the limb transform's tables serves both the
// bases that use it and those that do not.
struct Warm {
    gfe_engine *e = nullptr;
    uint32_t n = 0;
    int rns = 0, is32 = 0;
    bool limb = false;
    uint64_t bytes = 0;
};

// The lone-request path: a quick or truncated test that arrives while no pool has work runs alone
// on a warm engine, on a host thread, as gfe_run_quick / gfe_run_prp run it (the same residues as
// in a pool).  SOLO_WORKERS such runs may go side by side: a single candidate below n = 18 is
// bound by launch latency and leaves most of the device idle.
constexpr int SOLO_WORKERS = 2;

struct Solo {
    std::thread th;
    std::atomic<bool> finished{false};
    bool running = false;
    Admitted job;
    gfe_engine *e = nullptr;   // the warm engine in use, or null: the worker creates `created`
    Warm created;
    gfe_serve_result out;
    int rc = GFE_OK;
    std::string err;
    double t_done = 0;
};

struct Pool {
    gfe_batch *bt = nullptr;
    uint32_t n = 0, cap = 0;
    uint32_t depth = 0;   // proof pools hold 3 + 2^depth registers per slot
    bool mixed = false;   // a mixed pool (n = 13 allocation) takes the quick tests of n = 12 and 13
    uint64_t bytes = 0;
    uint32_t busy = 0;
};

}  // namespace

// A lone run of at most this many squarings is short enough for light-load requests to wait for.
constexpr uint64_t SOLO_SHORT = 65536;

struct gfe_server {
    int device = 0;
    double t0 = 0;
    gfe_server_stats st;
    std::deque<Request> queue;
    std::vector<Pool> pools;
    uint64_t seq = 0;                                  // the batch tag of each admitted request
    std::unordered_map<uint64_t, Admitted> inflight;  // by batch tag
    std::deque<gfe_serve_result> done;
    std::vector<gfe_batch_result> buf;
    int solo_rate = 8;            // the lone-request path while at most this many arrivals in the last second
    std::deque<double> recent;    // arrival times of the last second
    Solo solos[SOLO_WORKERS];
    std::vector<Warm> warm;

    bool in_use(const gfe_engine *e) const {
        for (const Solo &S : solos)
            if (S.running && S.e == e) return true;
        return false;
    }
    int solos_running() const {
        int r = 0;
        for (const Solo &S : solos) r += S.running ? 1 : 0;
        return r;
    }

    // Light load: few enough arrivals in the last second that a lone run does not hold up the pools.
    bool light() {
        const double now = clock();
        while (!recent.empty() && recent.front() < now - 1.0) recent.pop_front();
        return solo_rate > 0 && (int)recent.size() <= solo_rate;
    }

    double clock() const { return steady_s() - t0; }

    static Warm warm_kind(uint32_t b, uint32_t n) {
        Warm w;
        w.n = n;
        gfe::select_primes(b, n, &w.rns, &w.is32);
        w.is32 = gfe::transform_is32(w.rns, w.is32);
        w.limb = gfe::limb_ok(b, n) && gfe::default_limb_plan(n).limb;
        w.bytes = gfe::engine_bytes(w.rns, size_t(1) << n, 3, w.limb);
        return w;
    }
    // A warm engine that can take kind k: same n and prime set, and the limb transform's tables
    // when k uses them (an engine with them serves the other bases of its set too).
    long find_warm(const Warm &k) const {
        for (size_t i = 0; i < warm.size(); ++i)
            if (warm[i].n == k.n && warm[i].rns == k.rns && warm[i].is32 == k.is32 && (warm[i].limb || !k.limb) &&
                !in_use(warm[i].e))
                return (long)i;
        return -1;
    }
    // Frees warm engines not in use until `need` more bytes fit in the budget.
    bool make_room(uint64_t need) {
        for (size_t i = warm.size(); i-- > 0 && st.bytes_reserved + need > st.bytes_budget;) {
            if (in_use(warm[i].e)) continue;
            gfe_destroy(warm[i].e);
            st.bytes_reserved -= warm[i].bytes;
            warm.erase(warm.begin() + (long)i);
        }
        return st.bytes_reserved + need <= st.bytes_budget;
    }
    bool idle() const {
        for (const Pool &P : pools)
            if (P.busy) return false;
        return inflight.empty();
    }

    // Runs on a lone-request thread: touches only its Solo and its engine.
    void solo_work(Solo *Sp) {
        Solo &S = *Sp;
        const Request rq = S.job.rq;
        try {
            gfe_engine *e = S.e;
            if (!e) {
                if (gfe_create(rq.b, rq.n, 3, device, &e) != GFE_OK) throw std::runtime_error(gfe_last_error());
                S.created.e = e;
            } else if (gfe_rebase(e, rq.b) != GFE_OK) {
                throw std::runtime_error(gfe_last_error());
            }
            // a truncated run needs only E's top bits (exponent_top), not E
            std::unique_ptr<gfe::genefer> Gp(rq.iters ? new gfe::genefer(e, rq.b, rq.n, 3, rq.iters)
                                                      : new gfe::genefer(e, rq.b, rq.n, 3));
            gfe::genefer &G = *Gp;
            const size_t es = G.esize();
            const bool full = rq.iters == 0 || rq.iters >= es;
            const int B_GL = full ? gfe::genefer::B_GerbiczLi(es) : 0;
            G.prp(B_GL, 0, false, full ? 0 : (long)(es - rq.iters));
            G.t().getInt(G.gi());
            uint64_t r = 0, o = 0;
            const bool one = G.gi().isOne(r, o);
            memset(&S.out, 0, sizeof(S.out));
            S.out.tag = rq.tag; S.out.n = rq.n; S.out.b = rq.b;
            S.out.res64 = r; S.out.old64 = o; S.out.hash64 = one ? 0 : G.gi().gethash64();
            S.out.esize = es; S.out.iters = full ? es : rq.iters; S.out.is_prp = one ? 1 : 0;
            S.out.gl_ok = full ? (G.GL(B_GL) ? 1 : 0) : -1;
        } catch (const std::exception &ex) {
            S.rc = GFE_EHIP;
            S.err = ex.what();
        }
        S.t_done = clock();
        S.finished.store(true, std::memory_order_release);
    }

    // Starts `rq` alone on a warm engine of its kind (one is created on the thread when none is
    // warm and it fits the budget); false: the request takes the pool path.
    bool start_solo(Solo &solo, const Request &rq) {
        const Warm k = warm_kind(rq.b, rq.n);
        const long wi = find_warm(k);
        if (wi < 0) {
            if (!make_room(k.bytes)) return false;
            st.bytes_reserved += k.bytes;
            st.bytes_peak = std::max(st.bytes_peak, st.bytes_reserved);
        }
        solo.job = Admitted{rq, clock()};
        solo.e = wi >= 0 ? warm[(size_t)wi].e : nullptr;
        solo.created = k;
        solo.rc = GFE_OK;
        solo.err.clear();
        solo.finished.store(false);
        try {
            solo.th = std::thread(&gfe_server::solo_work, this, &solo);
        } catch (const std::exception &) {
            if (wi < 0) st.bytes_reserved -= k.bytes;
            return false;
        }
        solo.running = true;
        return true;
    }

    int collect_solo(Solo &solo) {
        if (!solo.running || !solo.finished.load(std::memory_order_acquire)) return GFE_OK;
        solo.th.join();
        solo.running = false;
        if (!solo.e) {  // the worker was to create the engine
            if (solo.created.e && solo.rc == GFE_OK) {
                warm.push_back(solo.created);
            } else {
                if (solo.created.e) gfe_destroy(solo.created.e);
                st.bytes_reserved -= solo.created.bytes;
            }
        }
        if (solo.rc != GFE_OK) return gfe::set_error(solo.rc, "serve: lone run: " + solo.err);
        gfe_serve_result o = solo.out;
        o.t_arrival = solo.job.rq.t_arrival; o.t_admit = solo.job.t_admit; o.t_done = solo.t_done;
        done.push_back(o);
        st.completed++;
        st.solo++;
        return GFE_OK;
    }

    // A warm engine for the quick tests of n ahead of the first request: the kind of the NTT3
    // candidates of n (with the limb transform's tables where some b of n uses them).
    int warm_up(uint32_t n) {
        uint32_t pick = 0;
        for (int pass = 0; pass < 2 && !pick; ++pass)
            for (double x = 8388608.0; x < 2.0e9 && !pick; x *= 1.01) {
                const uint32_t b = (uint32_t)x & ~1u;
                const Warm k = warm_kind(b, n);
                if (k.rns == 3 && (k.limb || pass == 1)) pick = b;
            }
        if (!pick) return GFE_OK;
        const Warm k = warm_kind(pick, n);
        if (find_warm(k) >= 0 || !make_room(k.bytes)) return GFE_OK;
        Warm w = k;
        const int rc = gfe_create(pick, n, 3, device, &w.e);
        if (rc != GFE_OK) return rc;
        warm.push_back(w);
        st.bytes_reserved += w.bytes;
        st.bytes_peak = std::max(st.bytes_peak, st.bytes_reserved);
        return GFE_OK;
    }

    void release(size_t i) {
        gfe_batch_destroy(pools[i].bt);
        st.bytes_reserved -= pools[i].bytes;
        pools.erase(pools.begin() + (long)i);
        st.pools_released++;
        st.pools = (uint32_t)pools.size();
    }

    // Creates a pool of n (proof depth `depth`) with up to `want` slots within the budget, releasing
    // idle pools of other kinds when needed.  *idx receives its index, or -1 when the budget has no
    // room.
    // Quick tests of n <= 13 share one kind of pool: a mixed pool, whose squaring blocks are one
    // launch for both n (gfe_batch_create_mixed).
    static bool mixed_kind(uint32_t n, uint32_t depth) { return depth == 0 && n <= 13; }
    static bool pool_takes(const Pool &P, uint32_t n, uint32_t depth) {
        return mixed_kind(n, depth) ? P.mixed : (!P.mixed && P.n == n && P.depth == depth);
    }

    int create_pool(uint32_t n, uint32_t depth, uint32_t want, long *idx) {
        *idx = -1;
        const bool mixed = mixed_kind(n, depth);
        if (mixed) n = 13;
        auto fits = [&](uint32_t c) {
            return st.bytes_reserved + gfe_batch_proof_device_bytes(n, c, depth) <= st.bytes_budget;
        };
        if (!fits(1)) make_room(gfe_batch_proof_device_bytes(n, 1, depth));
        if (!fits(1)) {
            for (size_t i = pools.size(); i-- > 0;)
                if (pools[i].busy == 0 && (pools[i].n != n || pools[i].depth != depth || pools[i].mixed != mixed)) {
                    release(i);
                    if (fits(1)) break;
                }
        }
        if (!fits(1)) return GFE_OK;
        uint32_t cap = want;
        while (cap > 1 && !fits(cap)) cap = std::max<uint32_t>(1, cap / 2);
        Pool P;
        P.n = n;
        P.depth = depth;
        P.mixed = mixed;
        P.cap = cap;
        P.bytes = gfe_batch_proof_device_bytes(n, cap, depth);
        const int rc = mixed ? gfe_batch_create_mixed(cap, device, &P.bt) : gfe_batch_create_proof(n, cap, depth, device, &P.bt);
        if (rc != GFE_OK) return rc;
        pools.push_back(P);
        st.bytes_reserved += P.bytes;
        st.bytes_peak = std::max(st.bytes_peak, st.bytes_reserved);
        st.pools_created++;
        st.pools = (uint32_t)pools.size();
        *idx = (long)pools.size() - 1;
        return GFE_OK;
    }

    int admit() {
        for (auto it = queue.begin(); it != queue.end();) {
            const Request &rq = *it;
            if (rq.depth == 0 && idle() && light()) {
                Solo *free_slot = nullptr;
                bool short_run = false;
                for (Solo &S : solos) {
                    if (!S.running) free_slot = free_slot ? free_slot : &S;
                    else if (S.job.rq.iters != 0 && S.job.rq.iters <= SOLO_SHORT) short_run = true;
                }
                if (free_slot && start_solo(*free_slot, rq)) {
                    st.admitted++;
                    st.queued--;
                    it = queue.erase(it);
                    continue;
                }
                // at light load a request waits for a short lone run rather than share the device
                // with it from a pool: the wait is shorter than the slowdown of both
                if (!free_slot && short_run) {
                    ++it;
                    continue;
                }
            }
            long pi = -1;
            for (size_t i = 0; i < pools.size(); ++i)
                if (pool_takes(pools[i], rq.n, rq.depth) && pools[i].busy < pools[i].cap) { pi = (long)i; break; }
            if (pi < 0) {
                const int rc = create_pool(rq.n, rq.depth, default_slots(rq.n, rq.depth), &pi);
                if (rc != GFE_OK) return rc;
            }
            if (pi < 0) { ++it; continue; }  // waits for memory
            Pool &P = pools[(size_t)pi];
            int32_t slot = -1;
            const uint64_t id = seq++;
            const int rc = gfe_batch_submit_n(P.bt, rq.n, rq.b, rq.iters, id, &slot);
            if (rc == GFE_EFULL) { ++it; continue; }
            if (rc != GFE_OK) return rc;
            inflight[id] = Admitted{rq, clock()};
            P.busy++;
            st.admitted++;
            st.queued--;
            it = queue.erase(it);
        }
        return GFE_OK;
    }

    int collect() {
        for (Solo &S : solos) {
            const int rs = collect_solo(S);
            if (rs != GFE_OK) return rs;
        }
        for (Pool &P : pools) {
            if (P.busy == 0) continue;
            if (buf.size() < P.cap) buf.resize(P.cap);
            uint32_t got = 0;
            const int rc = gfe_batch_poll(P.bt, buf.data(), (uint32_t)buf.size(), &got);
            if (rc != GFE_OK) return rc;
            const double t = clock();
            for (uint32_t k = 0; k < got; ++k) {
                const gfe_batch_result &r = buf[k];
                const auto f = inflight.find(r.tag);
                if (f == inflight.end()) return gfe::set_error(GFE_EINVAL, "serve: result of an unknown request");
                const Admitted a = f->second;
                inflight.erase(f);
                gfe_serve_result o;
                memset(&o, 0, sizeof(o));
                o.tag = a.rq.tag; o.n = r.n; o.b = r.b; o.res64 = r.res64; o.old64 = r.old64; o.hash64 = r.hash64;
                o.esize = r.esize; o.iters = r.iters; o.is_prp = r.is_prp; o.gl_ok = r.gl_ok;
                o.depth = r.depth; o.pkey = r.pkey;
                memcpy(o.mu_hash64, r.mu_hash64, sizeof(o.mu_hash64));
                o.t_arrival = a.rq.t_arrival; o.t_admit = a.t_admit; o.t_done = t;
                done.push_back(o);
                P.busy--;
                st.completed++;
            }
        }
        return GFE_OK;
    }

    int pump(int32_t *busy_out) {
        st.pumps++;
        int rc = collect();
        if (rc == GFE_OK) rc = admit();
        if (rc != GFE_OK) return rc;
        bool stepped = false;
        for (Pool &P : pools) {
            if (P.busy == 0) continue;
            const int bl = gfe_batch_backlog(P.bt);
            if (bl < 0) return bl;
            if (bl >= 2) continue;
            int32_t active = 0;
            rc = gfe_batch_step(P.bt, &active);
            if (rc != GFE_OK) return rc;
            st.blocks++;
            stepped = true;
        }
        rc = collect();
        if (rc != GFE_OK) return rc;
        uint32_t busy = (uint32_t)solos_running();
        for (const Pool &P : pools) busy += P.busy;
        st.busy_slots = busy;
        if (!stepped && busy > 0) std::this_thread::sleep_for(std::chrono::microseconds(50));
        if (busy_out) *busy_out = (int32_t)busy;
        return GFE_OK;
    }
};

extern "C" {

int gfe_server_create(int device, uint64_t hbm_budget, gfe_server **out) {
    if (!out) return gfe::set_error(GFE_EINVAL, "null output");
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return gfe::set_error(GFE_ENODEV, "no HIP device");
    if (device < 0 || device >= ndev) return gfe::set_error(GFE_ENODEV, "device index out of range");
    if (hbm_budget == 0) {
        size_t free_b = 0, total_b = 0;
        if (hipSetDevice(device) != hipSuccess || hipMemGetInfo(&free_b, &total_b) != hipSuccess)
            return gfe::set_error(GFE_EHIP, "hipMemGetInfo");
        hbm_budget = (uint64_t)free_b / 4 * 3;
    }
    gfe_server *s = new gfe_server;
    s->device = device;
    s->t0 = steady_s();
    memset(&s->st, 0, sizeof(s->st));
    s->st.bytes_budget = hbm_budget;
    *out = s;
    return GFE_OK;
}

int gfe_server_destroy(gfe_server *s) {
    if (!s) return GFE_OK;
    for (Solo &S : s->solos)
        if (S.running) {
            S.th.join();
            if (!S.e && S.created.e) gfe_destroy(S.created.e);
        }
    for (Warm &w : s->warm) gfe_destroy(w.e);
    for (Pool &P : s->pools) gfe_batch_destroy(P.bt);
    delete s;
    return GFE_OK;
}

int gfe_server_set_solo(gfe_server *s, int max_rate) {
    if (!s) return gfe::set_error(GFE_EINVAL, "null server");
    if (max_rate < 0) return gfe::set_error(GFE_EINVAL, "serve: negative arrival rate");
    s->solo_rate = max_rate;
    return GFE_OK;
}

double gfe_server_clock(const gfe_server *s) { return s ? s->clock() : 0.0; }

namespace {

int prepare_pool(gfe_server *s, uint32_t n, uint32_t capacity) {
    for (const Pool &P : s->pools)
        if (gfe_server::pool_takes(P, n, 0)) return GFE_OK;  // n <= 13: the mixed pool already serves n
    long pi = -1;
    const int rc = s->create_pool(n, 0, capacity ? capacity : default_slots(n), &pi);
    if (rc != GFE_OK) return rc;
    if (pi < 0) return gfe::set_error(GFE_ENOMEM, "serve: pool does not fit in the budget");
    return GFE_OK;
}

}  // namespace

int gfe_server_prepare(gfe_server *s, uint32_t n, uint32_t capacity) {
    if (!s) return gfe::set_error(GFE_EINVAL, "null server");
    if (n < 12 || n > 23) return gfe::set_error(GFE_EINVAL, "n must be in [12, 23]");
    const int rc = prepare_pool(s, n, capacity);
    if (rc != GFE_OK || s->solo_rate == 0) return rc;
    return s->warm_up(n);
}

int gfe_server_submit(gfe_server *s, uint32_t n, uint32_t b, uint64_t iters, uint64_t tag) {
    return gfe_server_submit_proof(s, n, b, iters, 0, tag);
}

int gfe_server_submit_proof(gfe_server *s, uint32_t n, uint32_t b, uint64_t iters, uint32_t depth, uint64_t tag) {
    if (!s) return gfe::set_error(GFE_EINVAL, "null server");
    if (n < 12 || n > 23) return gfe::set_error(GFE_EINVAL, "n must be in [12, 23]");
    if (b < 1024 || b > 2000000000u) return gfe::set_error(GFE_EINVAL, "b must be in [1024, 2000000000]");
    if (depth > GFE_BATCH_MAX_DEPTH) return gfe::set_error(GFE_EINVAL, "serve: proof depth out of range");
    if (gfe_batch_proof_device_bytes(n, 1, depth) > s->st.bytes_budget)
        return gfe::set_error(GFE_ENOMEM, "serve: one candidate of this n exceeds the budget");
    s->queue.push_back(Request{tag, n, b, depth ? 0 : iters, depth, s->clock()});
    s->recent.push_back(s->queue.back().t_arrival);
    s->st.submitted++;
    s->st.queued++;
    return GFE_OK;
}

int gfe_server_pump(gfe_server *s, int32_t *busy) {
    if (!s) return gfe::set_error(GFE_EINVAL, "null server");
    return s->pump(busy);
}

int gfe_server_poll(gfe_server *s, gfe_serve_result *out, uint32_t max, uint32_t *count) {
    if (!s || !count) return gfe::set_error(GFE_EINVAL, "poll");
    const int rc = s->collect();
    if (rc != GFE_OK) return rc;
    uint32_t k = 0;
    while (k < max && out && !s->done.empty()) {
        out[k++] = s->done.front();
        s->done.pop_front();
    }
    *count = k;
    return GFE_OK;
}

int gfe_server_get_stats(const gfe_server *s, gfe_server_stats *out) {
    if (!s || !out) return gfe::set_error(GFE_EINVAL, "stats");
    *out = s->st;
    return GFE_OK;
}

int gfe_serve_replay(const uint32_t *n, const uint32_t *b, const uint64_t *iters, const uint32_t *depth,
                     const double *arrival, uint32_t count, int device, uint64_t hbm_budget, const uint32_t *prepare_n,
                     uint32_t prepare_count, int solo_rate, gfe_serve_result *results, gfe_server_stats *stats) {
    if (!n || !b || !iters || !arrival || !results || count == 0) return gfe::set_error(GFE_EINVAL, "replay arguments");
    for (uint32_t k = 1; k < count; ++k)
        if (arrival[k] < arrival[k - 1]) return gfe::set_error(GFE_EINVAL, "replay: arrivals must be non-decreasing");
    gfe_server *s = nullptr;
    int rc = gfe_server_create(device, hbm_budget, &s);
    if (rc != GFE_OK) return rc;
    if (solo_rate >= 0) s->solo_rate = solo_rate;
    // every pool first, then the warm engines: streams share the device's few hardware queues in
    // creation order, and the pools, which run together, should each get one of their own
    for (uint32_t i = 0; i < prepare_count && rc == GFE_OK; ++i) {
        if (prepare_n[i] < 12 || prepare_n[i] > 23) rc = gfe::set_error(GFE_EINVAL, "n must be in [12, 23]");
        else rc = prepare_pool(s, prepare_n[i], 0);
    }
    for (uint32_t i = 0; i < prepare_count && rc == GFE_OK; ++i) rc = s->warm_up(prepare_n[i]);
    const double start = s->clock();
    uint32_t next = 0, finished = 0;
    std::vector<gfe_serve_result> got(256);
    while (rc == GFE_OK && finished < count) {
        const double now = s->clock() - start;
        while (next < count && arrival[next] <= now && rc == GFE_OK) {
            rc = gfe_server_submit_proof(s, n[next], b[next], iters[next], depth ? depth[next] : 0, next);
            ++next;
        }
        if (rc != GFE_OK) break;
        int32_t busy = 0;
        rc = gfe_server_pump(s, &busy);
        if (rc != GFE_OK) break;
        uint32_t c = 0;
        rc = gfe_server_poll(s, got.data(), (uint32_t)got.size(), &c);
        for (uint32_t k = 0; k < c; ++k) {
            gfe_serve_result r = got[k];
            r.t_arrival -= start; r.t_admit -= start; r.t_done -= start;
            results[r.tag] = r;
        }
        finished += c;
        if (busy == 0 && s->queue.empty() && next < count) {
            const double wait = arrival[next] - (s->clock() - start);
            if (wait > 0) std::this_thread::sleep_for(std::chrono::duration<double>(std::min(wait, 0.01)));
        }
    }
    if (stats) *stats = s->st;
    gfe_server_destroy(s);
    return rc;
}

}  // extern "C"
