// qk_server.hip -- continuous-batching request server (qk_server_* in include/qk.h).
//
// A scheduler thread admits pending requests (American requests first, the
// rest in arrival order) while their device footprint fits a first-fit arena
// carved from one allocation, packs them into a batch (work items = request x
// tile range; a batch's work is capped at batch_items_per_wave items per
// resident wavefront), and launches it on a free stream.  Up to max_inflight
// batches run at once (beyond the first only while the queue holds
// overlap_min_backlog batches' work, if set), so requests arriving while
// others are on the GPU start as soon as a stream and memory free up.  An
// express lane (one more stream) batches small requests ahead of arrival
// order beside the running batch, at raised issue priority.
// American requests generate their path store inside a batch launch; once it
// is ready they join the Longstaff-Schwartz pump, whose every launch advances
// all joined requests by one exercise date.  Completion is observed by polling
// events; results are kept until the client polls or waits.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <mutex>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "qk_capi_internal.hpp"
#include "qk_launch.hpp"

using namespace qk;

namespace qk {
namespace {

using clk = std::chrono::steady_clock;
constexpr size_t ALIGN = 256;
constexpr int PUMP_DEPTH = 8;          // pump steps queued ahead on the pump stream
constexpr size_t PUMP_SLOT_BYTES = 512; // pinned staging per pump slot: MergeReqD, then qk_stats at +256
size_t up(size_t x) { return (x + ALIGN - 1) / ALIGN * ALIGN; }

// First-fit allocator over one device allocation, with coalescing frees.
struct Arena {
    char *base = nullptr;
    size_t size = 0, used = 0, peak = 0;
    std::map<size_t, size_t> free_;   // offset -> bytes
    hipError_t init(size_t bytes) {
        size = up(bytes);
        hipError_t e = hipMalloc((void **)&base, size);
        if (e == hipSuccess) free_[0] = size;
        return e;
    }
    size_t alloc(size_t bytes) {
        bytes = up(std::max<size_t>(bytes, ALIGN));
        for (auto it = free_.begin(); it != free_.end(); ++it) {
            if (it->second >= bytes) {
                const size_t off = it->first, rest = it->second - bytes;
                free_.erase(it);
                if (rest) free_[off + bytes] = rest;
                used += bytes;
                peak = std::max(peak, used);
                return off;
            }
        }
        return SIZE_MAX;
    }
    void release(size_t off, size_t bytes) {
        bytes = up(std::max<size_t>(bytes, ALIGN));
        used -= bytes;
        auto it = free_.emplace(off, bytes).first;
        auto nx = std::next(it);
        if (nx != free_.end() && it->first + it->second == nx->first) {
            it->second += nx->second;
            free_.erase(nx);
        }
        if (it != free_.begin()) {
            auto pv = std::prev(it);
            if (pv->first + pv->second == it->first) {
                pv->second += it->second;
                free_.erase(it);
            }
        }
    }
};

// Byte layout of one request's arena block.
struct Layout {
    size_t psum, pm2, pitm, scratch, store, cf, coeffs, bpart, mreq, creq, stats, total;
};

struct SReq {
    uint64_t ticket = 0;
    qk_contract c{};
    qk_payoff p{};
    bool american = false;
    int n_basis = 4;               // American: Laguerre functions (3, 4, 5 = P02, P03, P04)
    bool weighted = false;         // American: times exp(-X/2)
    clk::time_point t_submit, t_admit, t_done;
    bool admitted = false;
    qk_status status = QK_OK;
    qk_stats stats{};
    std::string err;
    size_t off = SIZE_MAX;
    Layout lay{};
    int64_t n_tiles = 0, n_chunks = 0, chunk = 0;
    int nb = 0;
    bool done = false;
    // American: pump state
    int t = -1;                    // next date to process
    int pslot = -1;                // pinned staging slot while in the pump
    int pump = -1;                 // index of the pump chain that advances it
    hipEvent_t ev = nullptr;       // backward sweep and merge finished
    hipEvent_t ev_fwd = nullptr;   // path store generated
    char *base = nullptr;
    char *dev(size_t o) const { return base + off + o; }
};

// One Longstaff-Schwartz pump chain: its own stream, on which every step
// advances each of its requests by one date.  Two chains run side by side so
// that one chain's solve launch and launch gaps overlap the other's path work.
constexpr int N_PUMPS = 4;
// Pending requests examined behind one that waits for memory.
constexpr int ADMIT_SCAN = 512;
struct PumpChain {
    hipStream_t stream = nullptr;
    std::vector<SReq *> active;            // in this chain (t >= 0)
    std::deque<hipEvent_t> steps;          // events of queued steps
};

struct Slot {
    hipStream_t stream = nullptr;
    hipEvent_t ev_done = nullptr;
    bool busy = false;
    bool express = false;          // the express lane: small requests only, raised issue priority
    std::vector<SReq *> reqs, merged;
    DevBuf dbuf;
    char *hbuf = nullptr;
    size_t hcap = 0;
    qk_stats *hstats = nullptr;
    size_t hstats_cap = 0;
    // Outgrown staging buffers.  hipFree and hipHostFree wait for the whole
    // device, which would stall the scheduler thread (and with it every pump
    // chain) behind the batch running on the other slot, so they are released
    // only when the server is destroyed.
    std::vector<void *> retired_host, retired_dev;
};

}  // namespace
}  // namespace qk

struct qk_server {
    qk_ctx *ctx = nullptr;
    qk_server_config cfg{};
    int n_cu = 256;
    Arena arena;
    std::vector<Slot> slots;
    // Longstaff-Schwartz pump
    qk::PumpChain pumps[qk::N_PUMPS];
    char *pump_hbuf = nullptr;                 // [N_PUMPS * PUMP_MAX_REQ][PUMP_SLOT_BYTES] pinned
    std::vector<bool> pslot_used;
    std::deque<SReq *> lsm_waiting;            // forward launched, not yet in a pump
    std::vector<SReq *> lsm_finishing;         // merge enqueued, completion pending
    std::vector<hipEvent_t> event_pool;
    std::thread th;
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::deque<SReq *> pending;
    // Path-steps of the pending requests, in double: a request may hold up to
    // 2^56, and a queue of them would overflow 64-bit integers.
    double pending_ps = 0.0;
    std::unordered_map<uint64_t, SReq *> reqs;
    uint64_t next_ticket = 1;
    bool stop = false;
    int64_t n_submitted = 0, n_completed = 0, n_batches = 0, path_steps = 0;
    double busy_s = 0.0;
    clk::time_point busy_since{};
    int n_busy = 0;
};

namespace qk {
namespace {

Layout layout_of(const qk_contract &c, bool american, int64_t n_tiles, int64_t n_chunks, int nb) {
    const size_t rs = c.precision == QK_FP32 ? 4 : 8;
    Layout L{};
    size_t o = 0;
    auto take = [&](size_t bytes) {
        const size_t at = o;
        o += up(std::max<size_t>(bytes, 1));
        return at;
    };
    L.psum = take(n_tiles * 8);
    L.pm2 = take(n_tiles * 8);
    L.pitm = take(n_tiles * 4);
    L.scratch = take(n_chunks * MERGE_MOM_BYTES);
    if (american) {
        L.store = take((size_t)c.n_paths * c.n_steps * rs);
        L.cf = take((size_t)c.n_paths * rs);
        L.coeffs = take((size_t)c.n_steps * LSM_MAX_BASIS * 8);
        L.bpart = take((size_t)2 * nb * LSM_STRIDE * 8);
        L.mreq = take(sizeof(MergeReqD));
        L.creq = take(n_chunks * 4);
        L.stats = take(sizeof(qk_stats));
    }
    L.total = o;
    return L;
}

void busy_begin(qk_server *s) {
    if (s->n_busy++ == 0) s->busy_since = clk::now();
}
void busy_end(qk_server *s) {
    if (--s->n_busy == 0) s->busy_s += std::chrono::duration<double>(clk::now() - s->busy_since).count();
}

void complete(qk_server *s, SReq *r, qk_status st) {
    if (r->off != SIZE_MAX) {
        s->arena.release(r->off, r->lay.total);
        r->off = SIZE_MAX;
    }
    if (r->pslot >= 0) {
        s->pslot_used[r->pslot] = false;
        r->pslot = -1;
    }
    r->status = st;
    r->t_done = clk::now();
    r->done = true;
    ++s->n_completed;
    if (st == QK_OK) s->path_steps += r->c.n_paths * (int64_t)r->c.n_steps;
}

template <typename Real>
ReqD<Real> make_req(const SReq &r, double eps) {
    ReqD<Real> d{};
    d.lc = make_lane<Real>(r.c);
    d.pay = PayD<Real>{r.p.type, r.p.right, (Real)r.p.K, (Real)r.p.B};
    d.model = r.c.model;
    d.scheme = r.c.scheme;
    d.n_steps = r.c.n_steps;
    d.flags = r.american ? (REQ_STORE | REQ_PATHDEP) : (path_dependent(r.p.type) ? REQ_PATHDEP : 0);
    d.n_paths = r.c.n_paths;
    d.n_tiles = r.n_tiles;
    d.key0 = (uint32_t)r.c.seed;
    d.key1 = (uint32_t)(r.c.seed >> 32);
    d.itm_eps = (Real)eps;
    d.psum = (double *)r.dev(r.lay.psum);
    d.pm2 = (double *)r.dev(r.lay.pm2);
    d.pitm = (uint32_t *)r.dev(r.lay.pitm);
    d.store = r.american ? (Real *)r.dev(r.lay.store) : nullptr;
    return d;
}

void add_items(std::vector<ItemD> &items, int req_index, const SReq &r, int64_t item_path_steps) {
    const int64_t per_tile = (int64_t)TILE * r.c.n_steps;
    const int64_t tpi = std::clamp<int64_t>(item_path_steps / std::max<int64_t>(per_tile, 1), 1, 1024);
    for (int64_t t0 = 0; t0 < r.n_tiles; t0 += tpi)
        items.push_back(ItemD{req_index, (int32_t)std::min(tpi, r.n_tiles - t0), t0});
}

MergeReqD merge_desc(const SReq &r, int64_t first_chunk, double factor) {
    MergeReqD m{};
    m.psum = (const double *)r.dev(r.lay.psum);
    m.pm2 = (const double *)r.dev(r.lay.pm2);
    m.pitm = (const uint32_t *)r.dev(r.lay.pitm);
    m.scratch = r.dev(r.lay.scratch);
    m.n_tiles = r.n_tiles;
    m.n_paths = r.c.n_paths;
    m.first_chunk = first_chunk;
    m.n_chunks = r.n_chunks;
    m.factor = factor;
    return m;
}

// Appends the bytes of a vector to the staging buffer at 256-byte alignment.
template <typename T>
size_t stage(std::vector<char> &buf, const std::vector<T> &v) {
    const size_t at = up(buf.size());
    buf.resize(at + v.size() * sizeof(T));
    if (!v.empty()) std::memcpy(buf.data() + at, v.data(), v.size() * sizeof(T));
    return at;
}

hipEvent_t take_event(qk_server *srv) {
    if (!srv->event_pool.empty()) {
        hipEvent_t e = srv->event_pool.back();
        srv->event_pool.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    if (hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess) return nullptr;
    return e;
}

// Packs `batch` into the slot's staging buffer and launches it.
hipError_t launch_slot(qk_server *srv, Slot &slot, std::vector<SReq *> &batch) {
    std::vector<ReqD<float>> r32;
    std::vector<ReqD<double>> r64;
    // Work items of American requests (path stores for the pump) and of the
    // rest, launched in that order so that the pump can start on the
    // American requests while the rest of the batch still runs.
    std::vector<ItemD> i32, i64, a32, a64;
    std::vector<MergeReqD> mreq;
    std::vector<int32_t> chunk_req;
    slot.merged.clear();
    for (SReq *r : batch) {
        const double eps = r->c.precision == QK_FP32 ? 1e-6 : 1e-12;
        if (r->c.precision == QK_FP32) {
            add_items(r->american ? a32 : i32, (int)r32.size(), *r, srv->cfg.item_path_steps);
            r32.push_back(make_req<float>(*r, eps));
        } else {
            add_items(r->american ? a64 : i64, (int)r64.size(), *r, srv->cfg.item_path_steps);
            r64.push_back(make_req<double>(*r, eps));
        }
        if (!r->american) {
            const int idx = (int)mreq.size();
            mreq.push_back(merge_desc(*r, (int64_t)chunk_req.size(),
                                      std::exp(-r->c.r * r->c.T) *
                                          (r->p.type == QK_BINARY_CASH ? 1.0 : r->c.price_scale)));
            chunk_req.insert(chunk_req.end(), (size_t)r->n_chunks, idx);
            slot.merged.push_back(r);
        }
    }
    std::vector<char> buf;
    buf.reserve(4096);
    const size_t o_r32 = stage(buf, r32), o_r64 = stage(buf, r64), o_i32 = stage(buf, i32), o_i64 = stage(buf, i64);
    const size_t o_a32 = stage(buf, a32), o_a64 = stage(buf, a64);
    const size_t o_m = stage(buf, mreq), o_c = stage(buf, chunk_req);
    const size_t o_cnt = up(buf.size());
    buf.resize(o_cnt + 256, 0);                              // four batch counters, zeroed by the copy
    const size_t o_st = up(buf.size());
    buf.resize(o_st + mreq.size() * sizeof(qk_stats) + 8);
    hipError_t e;
    if (slot.hcap < buf.size()) {
        if (slot.hbuf) slot.retired_host.push_back(slot.hbuf);
        slot.hcap = buf.size() * 2;
        if ((e = hipHostMalloc((void **)&slot.hbuf, slot.hcap, hipHostMallocDefault)) != hipSuccess) {
            slot.hbuf = nullptr;
            slot.hcap = 0;
            return e;
        }
    }
    if (slot.dbuf.cap < buf.size()) {
        if (slot.dbuf.p) slot.retired_dev.push_back(slot.dbuf.p);
        slot.dbuf.p = nullptr;
        slot.dbuf.cap = 0;
        if (slot.dbuf.reserve(srv->ctx, buf.size() * 2) != QK_OK) return hipErrorOutOfMemory;
    }
    if (slot.hstats_cap < mreq.size() + 1) {
        if (slot.hstats) slot.retired_host.push_back(slot.hstats);
        slot.hstats_cap = (mreq.size() + 1) * 2;
        if ((e = hipHostMalloc((void **)&slot.hstats, slot.hstats_cap * sizeof(qk_stats), hipHostMallocDefault)) !=
            hipSuccess) {
            slot.hstats = nullptr;
            slot.hstats_cap = 0;
            return e;
        }
    }
    std::memcpy(slot.hbuf, buf.data(), buf.size());
    char *d = (char *)slot.dbuf.p;
    hipStream_t s = slot.stream;
    const int blocks = srv->n_cu * srv->cfg.blocks_per_cu;
    if ((e = hipMemcpyAsync(d, slot.hbuf, o_st, hipMemcpyHostToDevice, s)) != hipSuccess) return e;
    auto nblk = [&](size_t n_items) {
        return (int)std::min<int64_t>(blocks, ((int64_t)n_items + WAVES_PER_BLOCK - 1) / WAVES_PER_BLOCK);
    };
    // each launch has its own item counter (64 bytes apart)
    if ((e = launch_batch<float>((const ReqD<float> *)(d + o_r32), (const ItemD *)(d + o_a32), (int32_t)a32.size(),
                                 (unsigned int *)(d + o_cnt), nblk(a32.size()), slot.express, s)) != hipSuccess)
        return e;
    if ((e = launch_batch<double>((const ReqD<double> *)(d + o_r64), (const ItemD *)(d + o_a64), (int32_t)a64.size(),
                                  (unsigned int *)(d + o_cnt + 64), nblk(a64.size()), slot.express, s)) != hipSuccess)
        return e;
    for (SReq *r : batch)
        if (r->american && (e = hipEventRecord(r->ev_fwd, s)) != hipSuccess) return e;
    if ((e = launch_batch<float>((const ReqD<float> *)(d + o_r32), (const ItemD *)(d + o_i32), (int32_t)i32.size(),
                                 (unsigned int *)(d + o_cnt + 128), nblk(i32.size()), slot.express, s)) != hipSuccess)
        return e;
    if ((e = launch_batch<double>((const ReqD<double> *)(d + o_r64), (const ItemD *)(d + o_i64), (int32_t)i64.size(),
                                  (unsigned int *)(d + o_cnt + 192), nblk(i64.size()), slot.express, s)) != hipSuccess)
        return e;
    if (!mreq.empty()) {
        if ((e = launch_batch_merge((const MergeReqD *)(d + o_m), (int)mreq.size(), (const int32_t *)(d + o_c),
                                    (int64_t)chunk_req.size(), (qk_stats *)(d + o_st), s)) != hipSuccess)
            return e;
        if ((e = hipMemcpyAsync(slot.hstats, d + o_st, mreq.size() * sizeof(qk_stats), hipMemcpyDeviceToHost, s)) !=
            hipSuccess)
            return e;
    }
    if ((e = hipEventRecord(slot.ev_done, s)) != hipSuccess) return e;
    slot.reqs = batch;
    slot.busy = true;
    ++srv->n_batches;
    busy_begin(srv);
    return hipSuccess;
}

template <typename Real>
void fill_pump_req(PumpReq<Real> &q, const SReq *r) {
    q.store = (const Real *)r->dev(r->lay.store);
    q.cf = (Real *)r->dev(r->lay.cf);
    q.bpart = (double *)r->dev(r->lay.bpart);
    q.factor = nullptr;   // the pump factorises the Gram sums it accumulates
    q.coeffs = (double *)r->dev(r->lay.coeffs);
    q.psum = (double *)r->dev(r->lay.psum);
    q.pm2 = (double *)r->dev(r->lay.pm2);
    q.pitm = (uint32_t *)r->dev(r->lay.pitm);
    q.n_paths = r->c.n_paths;
    q.chunk = r->chunk;
    q.n_steps = r->c.n_steps;
    q.right = r->p.right;
    q.nb = r->nb;
    q.t = r->t;
    q.K = (Real)r->p.K;
    q.disc = (Real)std::exp(-r->c.r * (r->c.T / r->c.n_steps));   // ForwardBackwardMC.hpp:209
    q.itm_eps = (Real)(r->c.precision == QK_FP32 ? 1e-6 : 1e-12);
}

// One launch per (precision, regression basis) present, advancing every pump
// request of that kind by one date (the basis is a template parameter of the
// date kernel).
template <typename Real>
hipError_t launch_pump_step(PumpChain &pc, bool fp32) {
    for (int nbasis = 3; nbasis <= LSM_MAX_BASIS; ++nbasis) {
        for (int w = 0; w < 2; ++w) {
            PumpArgs<Real> a{};
            int blocks = 0;
            for (SReq *r : pc.active) {
                if ((r->c.precision == QK_FP32) != fp32 || r->n_basis != nbasis || (int)r->weighted != w) continue;
                fill_pump_req<Real>(a.r[a.n_req], r);
                a.block_start[a.n_req] = blocks;
                blocks += r->nb;
                ++a.n_req;
            }
            if (a.n_req == 0) continue;
            a.block_start[a.n_req] = blocks;
            const hipError_t e = launch_pump_two_phase<Real>(a, nbasis, w != 0, pc.stream);
            if (e != hipSuccess) return e;
        }
    }
    return hipSuccess;
}

// Enqueues the merge of a request whose date-0 step was just queued.
hipError_t enqueue_lsm_merge(qk_server *srv, SReq *r) {
    hipStream_t s = srv->pumps[r->pump].stream;
    char *h = srv->pump_hbuf + (size_t)r->pslot * PUMP_SLOT_BYTES;
    const MergeReqD m = merge_desc(*r, 0, r->c.price_scale);     // no e^{-rT}: OptionPricer.hpp:211-216
    std::memcpy(h, &m, sizeof(m));
    hipError_t e;
    if ((e = hipMemcpyAsync(r->dev(r->lay.mreq), h, sizeof(m), hipMemcpyHostToDevice, s)) != hipSuccess) return e;
    if ((e = hipMemsetAsync(r->dev(r->lay.creq), 0, r->n_chunks * 4, s)) != hipSuccess) return e;
    if ((e = launch_batch_merge((const MergeReqD *)r->dev(r->lay.mreq), 1, (const int32_t *)r->dev(r->lay.creq),
                                r->n_chunks, (qk_stats *)r->dev(r->lay.stats), s)) != hipSuccess)
        return e;
    if ((e = hipMemcpyAsync(h + 256, r->dev(r->lay.stats), sizeof(qk_stats), hipMemcpyDeviceToHost, s)) != hipSuccess)
        return e;
    return hipEventRecord(r->ev, s);
}

void finish_american(SReq *r, const qk_stats &st) {
    r->stats = st;
    // Immediate exercise at t0 (LSMAmerican.hpp:409-418).
    const bool fp32 = r->c.precision == QK_FP32;
    double S0 = r->c.S0, K = r->p.K;
    if (fp32) {
        S0 = (float)S0;
        K = (float)K;
    }
    const double today = r->p.right == QK_CALL ? std::max(S0 - K, 0.0) : std::max(K - S0, 0.0);
    if (today > r->stats.mean) {
        r->stats.mean = today;
        r->stats.m2 = 0.0;
        r->stats.price = today * r->c.price_scale;
        r->stats.se = 0.0;
        r->stats.itm = today > (fp32 ? 1e-6 : 1e-12) ? r->stats.n : 0;
        r->stats.prob_itm = (double)r->stats.itm / (double)r->stats.n;
    }
}

// Advances the pumps: admits requests whose path store is ready to the chain
// with the fewest paths in flight, queues one step on each chain with fewer
// than PUMP_DEPTH steps in flight, and hands requests that reached date 0 to
// the merge.  Returns true if anything was queued.
bool pump(qk_server *srv) {
    bool progress = false;
    int n_active = 0;
    for (PumpChain &pc : srv->pumps) {
        while (!pc.steps.empty() && hipEventQuery(pc.steps.front()) != hipErrorNotReady) {
            srv->event_pool.push_back(pc.steps.front());
            pc.steps.pop_front();
        }
        n_active += (int)pc.active.size();
    }
    for (auto it = srv->lsm_waiting.begin(); it != srv->lsm_waiting.end();) {
        // A request keeps its staging slot until its merge has completed, so
        // the slots bound the pumps and the finishing list together.
        if (n_active >= srv->cfg.max_lsm_streams ||
            n_active + (int)srv->lsm_finishing.size() >= (int)srv->pslot_used.size())
            break;
        SReq *r = *it;
        const hipError_t q = hipEventQuery(r->ev_fwd);
        if (q == hipErrorNotReady) {
            ++it;
            continue;
        }
        it = srv->lsm_waiting.erase(it);
        if (q != hipSuccess) {
            complete(srv, r, QK_EDEVICE);
            busy_end(srv);
            continue;
        }
        int best = -1;
        int64_t best_paths = 0;
        for (int k = 0; k < N_PUMPS; ++k) {
            PumpChain &pc = srv->pumps[k];
            if ((int)pc.active.size() >= PUMP_MAX_REQ) continue;
            int64_t paths = 0;
            for (const SReq *x : pc.active) paths += x->c.n_paths;
            if (best < 0 || paths < best_paths) {
                best = k;
                best_paths = paths;
            }
        }
        if (best < 0) {                                // every chain full: wait
            srv->lsm_waiting.push_front(r);
            break;
        }
        for (size_t k = 0; k < srv->pslot_used.size(); ++k) {
            if (!srv->pslot_used[k]) {
                srv->pslot_used[k] = true;
                r->pslot = (int)k;
                break;
            }
        }
        r->t = r->c.n_steps - 1;
        r->pump = best;
        srv->pumps[best].active.push_back(r);
        ++n_active;
        progress = true;
    }
    for (PumpChain &pc : srv->pumps) {
        if (pc.active.empty() || (int)pc.steps.size() >= PUMP_DEPTH) continue;
        hipError_t e = launch_pump_step<float>(pc, true);
        if (e == hipSuccess) e = launch_pump_step<double>(pc, false);
        std::vector<SReq *> keep;
        for (SReq *r : pc.active) {
            if (e == hipSuccess && r->t == 0) {
                e = enqueue_lsm_merge(srv, r);
                if (e == hipSuccess) {
                    srv->lsm_finishing.push_back(r);
                    continue;
                }
            }
            if (e != hipSuccess) {
                r->err = hipGetErrorString(e);
                complete(srv, r, QK_EDEVICE);
                busy_end(srv);
                continue;
            }
            --r->t;
            keep.push_back(r);
        }
        pc.active.swap(keep);
        if (hipEvent_t ev = take_event(srv)) {
            if (hipEventRecord(ev, pc.stream) == hipSuccess) pc.steps.push_back(ev);
            else srv->event_pool.push_back(ev);
        }
        progress = true;
    }
    return progress;
}

// Admits the pending request at `it` into an arena block at `off` and
// removes it from the queue.
std::deque<SReq *>::iterator take_pending(qk_server *srv, std::deque<SReq *>::iterator it, size_t off) {
    SReq *r = *it;
    r->off = off;
    r->base = srv->arena.base;
    r->t_admit = clk::now();
    r->admitted = true;
    srv->pending_ps -= (double)r->c.n_paths * (double)r->c.n_steps;
    return srv->pending.erase(it);
}

// Launches an admitted batch on a slot; its American requests then wait for
// their path stores to join the pump.  On failure every request completes
// with QK_EDEVICE.
void launch_admitted(qk_server *srv, Slot &slot, std::vector<SReq *> &batch) {
    const hipError_t e = launch_slot(srv, slot, batch);
    if (e == hipSuccess) {
        for (SReq *r : batch) {
            if (!r->american) continue;
            srv->lsm_waiting.push_back(r);
            busy_begin(srv);
        }
    } else {
        for (SReq *r : batch) {
            r->err = hipGetErrorString(e);
            complete(srv, r, QK_EDEVICE);
        }
        srv->cv_done.notify_all();
    }
}

// The express lane.  A small request that arrives while a large batch runs
// would wait for that batch and then run as long as the next one: the
// median request (~5e7 path-steps, well under a millisecond of the GPU)
// would take two batch lengths (measured p50 15.5 ms on the 4,096-request
// stream; 2.2-2.6 ms with the lane at 2^27 path-steps, p99 and burst
// throughput unchanged).  While the express slot is free, pending requests
// of at most express_max_path_steps path-steps are taken out of arrival
// order into a batch of their own, of at most EXPRESS_BATCH_FACTOR times
// that size in all, launched beside the large batch at raised issue
// priority; an American request's path store is generated there and its
// backward sweep then joins the pump (American p50 16-18 -> 8-9 ms).
// Returns true if it launched a batch.
constexpr int64_t EXPRESS_BATCH_FACTOR = 4;
bool admit_express(qk_server *srv) {
    const int64_t lim = srv->cfg.express_max_path_steps;
    if (lim <= 0 || srv->pending.empty()) return false;
    Slot *slot = nullptr;
    for (Slot &sl : srv->slots)
        if (sl.express && !sl.busy) {
            slot = &sl;
            break;
        }
    if (!slot) return false;
    std::vector<SReq *> batch;
    int64_t batch_ps = 0;
    int scanned = 0;
    for (auto it = srv->pending.begin(); it != srv->pending.end() && scanned < ADMIT_SCAN; ++scanned) {
        SReq *r = *it;
        const int64_t ps = r->c.n_paths * (int64_t)r->c.n_steps;
        if (ps > lim || r->lay.total > srv->arena.size / 64) {
            ++it;
            continue;
        }
        if (batch_ps + ps > EXPRESS_BATCH_FACTOR * lim) break;
        const size_t off = srv->arena.alloc(r->lay.total);
        if (off == SIZE_MAX) break;
        batch.push_back(r);
        batch_ps += ps;
        it = take_pending(srv, it, off);
    }
    if (batch.empty()) return false;
    launch_admitted(srv, *slot, batch);
    return true;
}

// One pass of the scheduler: retire finished work, run the pump, then admit
// and launch.  Returns true if anything changed.  Called with srv->mu held.
bool schedule_once(qk_server *srv) {
    bool progress = false;
    for (Slot &slot : srv->slots) {
        if (!slot.busy) continue;
        const hipError_t q = hipEventQuery(slot.ev_done);
        if (q == hipErrorNotReady) continue;
        for (size_t i = 0; i < slot.merged.size(); ++i) {
            SReq *r = slot.merged[i];
            r->stats = slot.hstats[i];
            complete(srv, r, q == hipSuccess ? QK_OK : QK_EDEVICE);
        }
        slot.busy = false;
        slot.reqs.clear();
        slot.merged.clear();
        busy_end(srv);
        progress = true;
    }
    for (size_t i = 0; i < srv->lsm_finishing.size();) {
        SReq *r = srv->lsm_finishing[i];
        const hipError_t q = hipEventQuery(r->ev);
        if (q == hipErrorNotReady) {
            ++i;
            continue;
        }
        qk_stats st;
        std::memcpy(&st, srv->pump_hbuf + (size_t)r->pslot * PUMP_SLOT_BYTES + 256, sizeof(st));
        if (q == hipSuccess) finish_american(r, st);
        complete(srv, r, q == hipSuccess ? QK_OK : QK_EDEVICE);
        srv->lsm_finishing.erase(srv->lsm_finishing.begin() + i);
        busy_end(srv);
        progress = true;
    }
    if (pump(srv)) progress = true;
    if (progress) srv->cv_done.notify_all();
    if (admit_express(srv)) progress = true;
    // Admission while memory and a batch slot are free: American requests
    // first (below), then the rest in arrival order.  A request that does
    // not fit the free memory stops admission of requests of its size, so
    // large requests are not starved by smaller ones; small
    // requests behind it (no path store, at most 1/64 of the arena: tile
    // moments only, released when their batch ends) may still go, so the
    // GPU keeps computing while the blocked request waits for memory that
    // the Longstaff-Schwartz pump is freeing.
    Slot *free_slot = nullptr;
    int n_busy_slots = 0;
    for (Slot &slot : srv->slots) {
        if (slot.express) continue;
        if (slot.busy) ++n_busy_slots;
        else if (!free_slot) free_slot = &slot;
    }
    if (!free_slot || srv->pending.empty()) return progress;
    // A batch takes requests until its work reaches batch_items_per_wave
    // items per resident wavefront: batches that end often let admission
    // (and with it the American requests waiting for arena memory) proceed
    // while the GPU computes, instead of one launch holding most of a burst.
    const int64_t cap = (int64_t)srv->cfg.batch_items_per_wave * srv->n_cu * srv->cfg.blocks_per_cu *
                        WAVES_PER_BLOCK * srv->cfg.item_path_steps;
    // Beside a batch in flight, a second one fills the first one's tail (its
    // last work items leave CUs idle until the host launches the next batch),
    // but the two share the GPU and each takes longer.  With a backlog gate
    // the overlap is used only while the queue holds that many batches' work
    // (a burst), and light traffic keeps one batch at a time.
    if (n_busy_slots > 0 && srv->cfg.overlap_min_backlog > 0 &&
        srv->pending_ps < (double)srv->cfg.overlap_min_backlog * (double)cap)
        return progress;
    std::vector<SReq *> batch;
    bool blocked = false;
    int scanned = 0;
    int64_t batch_ps = 0;
    auto admit = [&](std::deque<SReq *>::iterator it, size_t off) {
        batch.push_back(*it);
        batch_ps += (*it)->c.n_paths * (int64_t)(*it)->c.n_steps;
        return take_pending(srv, it, off);
    };
    // American requests first, in arrival order while they fit: their path
    // stores hold arena memory from the forward launch to the end of the
    // backward sweep, so they are the stream's bottleneck, and the batch
    // computes the rest while their sweeps run (the pump's memory-bound
    // date kernels overlap the batches' arithmetic).
    for (auto it = srv->pending.begin(); it != srv->pending.end();) {
        SReq *r = *it;
        if (!r->american || r->lay.total > srv->arena.size) {
            ++it;
            continue;
        }
        const size_t off = srv->arena.alloc(r->lay.total);
        if (off == SIZE_MAX) break;
        it = admit(it, off);
    }
    for (auto it = srv->pending.begin(); it != srv->pending.end() && scanned < ADMIT_SCAN;) {
        SReq *r = *it;
        if (!batch.empty() && batch_ps >= cap) break;
        if (r->lay.total > srv->arena.size) {
            srv->pending_ps -= (double)r->c.n_paths * (double)r->c.n_steps;
            it = srv->pending.erase(it);
            r->err = "request footprint exceeds the server arena";
            complete(srv, r, QK_ENOMEM);
            srv->cv_done.notify_all();
            progress = true;
            continue;
        }
        if (blocked) {
            ++scanned;
            if (r->american || r->lay.total > srv->arena.size / 64) {
                ++it;
                continue;
            }
        }
        const size_t off = srv->arena.alloc(r->lay.total);
        if (off == SIZE_MAX) {
            blocked = true;
            ++it;
            continue;
        }
        it = admit(it, off);
    }
    if (batch.empty()) return progress;
    launch_admitted(srv, *free_slot, batch);
    return true;
}

void scheduler(qk_server *srv) {
    (void)hipSetDevice(srv->ctx->device);
    std::unique_lock<std::mutex> lk(srv->mu);
    for (;;) {
        const bool progress = schedule_once(srv);
        const bool inflight = srv->n_busy > 0;
        if (srv->stop && !inflight && srv->pending.empty()) break;
        if (!progress) {
            if (inflight) srv->cv_work.wait_for(lk, std::chrono::microseconds(20));
            else srv->cv_work.wait_for(lk, std::chrono::milliseconds(5));
        }
    }
}

void fill_result(const SReq *r, qk_result *out) {
    out->stats = r->stats;
    out->status = r->status;
    out->latency_s = std::chrono::duration<double>(r->t_done - r->t_submit).count();
    out->queue_s = r->admitted ? std::chrono::duration<double>(r->t_admit - r->t_submit).count() : out->latency_s;
}

}  // namespace
}  // namespace qk

extern "C" qk_status qk_server_create(qk_ctx *ctx, const qk_server_config *cfg, qk_server **out) {
    if (!ctx || !out) return QK_EINVAL;
    *out = nullptr;
    if (hipSetDevice(ctx->device) != hipSuccess) return fail(ctx, QK_EDEVICE, "hipSetDevice");
    qk_server *s = new qk_server();
    s->ctx = ctx;
    if (cfg) s->cfg = *cfg;
    size_t free_b = 0, total_b = 0;
    if (hipMemGetInfo(&free_b, &total_b) != hipSuccess) {
        delete s;
        return fail(ctx, QK_EDEVICE, "hipMemGetInfo");
    }
    if (!s->cfg.mem_budget_bytes) s->cfg.mem_budget_bytes = (size_t)(0.6 * (double)free_b);
    if (s->cfg.max_inflight <= 0) s->cfg.max_inflight = 2;
    if (s->cfg.max_lsm_streams <= 0 || s->cfg.max_lsm_streams > N_PUMPS * PUMP_MAX_REQ)
        s->cfg.max_lsm_streams = N_PUMPS * PUMP_MAX_REQ;
    if (s->cfg.item_path_steps <= 0) s->cfg.item_path_steps = 131072;
    if (s->cfg.blocks_per_cu <= 0) s->cfg.blocks_per_cu = 4;
    if (s->cfg.batch_items_per_wave <= 0) s->cfg.batch_items_per_wave = 32;
    if (s->cfg.overlap_min_backlog < 0) s->cfg.overlap_min_backlog = 0;
    int n_cu = 0;
    if (hipDeviceGetAttribute(&n_cu, hipDeviceAttributeMultiprocessorCount, ctx->device) == hipSuccess && n_cu > 0)
        s->n_cu = n_cu;
    if (s->cfg.mem_budget_bytes > free_b || s->arena.init(s->cfg.mem_budget_bytes) != hipSuccess) {
        delete s;
        return fail(ctx, QK_ENOMEM, "server arena allocation failed");
    }
    if (s->cfg.express_max_path_steps == 0) s->cfg.express_max_path_steps = (int64_t)1 << 27;
    s->slots.resize(s->cfg.max_inflight + (s->cfg.express_max_path_steps > 0 ? 1 : 0));
    if (s->cfg.express_max_path_steps > 0) s->slots.back().express = true;
    for (Slot &sl : s->slots) {
        if (hipStreamCreateWithFlags(&sl.stream, hipStreamNonBlocking) != hipSuccess ||
            hipEventCreateWithFlags(&sl.ev_done, hipEventDisableTiming) != hipSuccess) {
            qk_server_destroy(s);
            return fail(ctx, QK_EDEVICE, "stream creation failed");
        }
    }
    s->pslot_used.assign(N_PUMPS * PUMP_MAX_REQ, false);
    bool pump_ok = hipHostMalloc((void **)&s->pump_hbuf, N_PUMPS * PUMP_MAX_REQ * PUMP_SLOT_BYTES,
                                 hipHostMallocDefault) == hipSuccess;
    for (PumpChain &pc : s->pumps)
        pump_ok = pump_ok && hipStreamCreateWithFlags(&pc.stream, hipStreamNonBlocking) == hipSuccess;
    if (!pump_ok) {
        qk_server_destroy(s);
        return fail(ctx, QK_EDEVICE, "pump stream creation failed");
    }
    s->th = std::thread(scheduler, s);
    *out = s;
    return QK_OK;
}

extern "C" qk_status qk_server_destroy(qk_server *s) {
    if (!s) return QK_EINVAL;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->stop = true;
    }
    s->cv_work.notify_all();
    if (s->th.joinable()) s->th.join();
    (void)hipSetDevice(s->ctx->device);
    for (Slot &sl : s->slots) {
        if (sl.stream) (void)hipStreamSynchronize(sl.stream);
        if (sl.ev_done) (void)hipEventDestroy(sl.ev_done);
        if (sl.stream) (void)hipStreamDestroy(sl.stream);
        if (sl.hbuf) (void)hipHostFree(sl.hbuf);
        if (sl.hstats) (void)hipHostFree(sl.hstats);
        for (void *h : sl.retired_host) (void)hipHostFree(h);
        for (void *d : sl.retired_dev) (void)hipFree(d);
        sl.dbuf.release();
    }
    for (PumpChain &pc : s->pumps) {
        if (pc.stream) {
            (void)hipStreamSynchronize(pc.stream);
            (void)hipStreamDestroy(pc.stream);
        }
        for (hipEvent_t e : pc.steps) (void)hipEventDestroy(e);
    }
    for (hipEvent_t e : s->event_pool) (void)hipEventDestroy(e);
    if (s->pump_hbuf) (void)hipHostFree(s->pump_hbuf);
    for (auto &kv : s->reqs) {
        if (kv.second->ev) (void)hipEventDestroy(kv.second->ev);
        if (kv.second->ev_fwd) (void)hipEventDestroy(kv.second->ev_fwd);
        delete kv.second;
    }
    if (s->arena.base) (void)hipFree(s->arena.base);
    delete s;
    return QK_OK;
}

extern "C" qk_status qk_server_submit(qk_server *s, const qk_contract *c, const qk_payoff *p, uint64_t *ticket) {
    return qk_server_submit_ex(s, c, p, nullptr, ticket);
}

extern "C" qk_status qk_server_submit_ex(qk_server *s, const qk_contract *c, const qk_payoff *p,
                                         const qk_lsm_options *lsm, uint64_t *ticket) {
    if (!s || !c || !p || !ticket) return QK_EINVAL;
    if (lsm && (lsm->basis < QK_LAGUERRE_P02 || lsm->basis > QK_LAGUERRE_P04))
        return fail(s->ctx, QK_EINVAL, "lsm basis must be LaguerreP02, P03 or P04");
    qk_ctx *ctx = s->ctx;
    std::lock_guard<std::mutex> lk(s->mu);
    qk_status st = validate_contract(ctx, c);
    if (st != QK_OK) return st;
    const bool american = p->type == QK_AMERICAN;
    if (american) {
        if (p->right != QK_CALL && p->right != QK_PUT) return fail(ctx, QK_EINVAL, "payoff.right out of range");
        if (!(p->K > 0.0)) return fail(ctx, QK_EINVAL, "payoff.K must be positive");
    } else if ((st = validate_payoff(ctx, p)) != QK_OK) {
        return st;
    }
    SReq *r = new SReq();
    r->ticket = s->next_ticket++;
    r->c = *c;
    r->p = *p;
    r->american = american;
    if (lsm) {
        r->n_basis = 3 + lsm->basis;
        r->weighted = lsm->weighted != 0;
    }
    r->n_tiles = (c->n_paths + TILE - 1) / TILE;
    r->n_chunks = (r->n_tiles + MERGE_CHUNK - 1) / MERGE_CHUNK;
    if (american)
        r->nb = server_pump_blocks(c->n_paths, &r->chunk);
    r->lay = layout_of(*c, american, r->n_tiles, r->n_chunks, r->nb);
    r->t_submit = clk::now();
    if (american && (hipEventCreateWithFlags(&r->ev, hipEventDisableTiming) != hipSuccess ||
                     hipEventCreateWithFlags(&r->ev_fwd, hipEventDisableTiming) != hipSuccess)) {
        if (r->ev) (void)hipEventDestroy(r->ev);
        delete r;
        return fail(ctx, QK_EDEVICE, "hipEventCreate");
    }
    s->reqs[r->ticket] = r;
    s->pending.push_back(r);
    s->pending_ps += (double)c->n_paths * (double)c->n_steps;
    ++s->n_submitted;
    *ticket = r->ticket;
    s->cv_work.notify_one();
    return QK_OK;
}

static qk_status take_result(qk_server *s, uint64_t ticket, qk_result *out, bool block) {
    std::unique_lock<std::mutex> lk(s->mu);
    auto it = s->reqs.find(ticket);
    if (it == s->reqs.end()) return QK_ENOTFOUND;
    SReq *r = it->second;
    if (!r->done) {
        if (!block) return QK_ENOTREADY;
        s->cv_done.wait(lk, [&] { return r->done; });
    }
    fill_result(r, out);
    if (r->ev) (void)hipEventDestroy(r->ev);
    if (r->ev_fwd) (void)hipEventDestroy(r->ev_fwd);
    s->reqs.erase(it);
    delete r;
    return QK_OK;
}

extern "C" qk_status qk_server_poll(qk_server *s, uint64_t ticket, qk_result *out) {
    if (!s || !out) return QK_EINVAL;
    return take_result(s, ticket, out, false);
}

extern "C" qk_status qk_server_wait(qk_server *s, uint64_t ticket, qk_result *out) {
    if (!s || !out) return QK_EINVAL;
    return take_result(s, ticket, out, true);
}

extern "C" qk_status qk_server_stats_get(qk_server *s, qk_server_stats *out) {
    if (!s || !out) return QK_EINVAL;
    std::lock_guard<std::mutex> lk(s->mu);
    out->n_submitted = s->n_submitted;
    out->n_completed = s->n_completed;
    out->n_batches = s->n_batches;
    out->path_steps = s->path_steps;
    out->arena_bytes = (int64_t)s->arena.size;
    out->arena_peak_bytes = (int64_t)s->arena.peak;
    out->busy_s = s->busy_s + (s->n_busy ? std::chrono::duration<double>(clk::now() - s->busy_since).count() : 0.0);
    return QK_OK;
}

extern "C" qk_status qk_server_replay(qk_server *s, const qk_contract *cs, const qk_payoff *ps, const double *arrival_s,
                                      int32_t n, qk_result *out) {
    if (!s || !cs || !ps || !arrival_s || !out || n < 0) return QK_EINVAL;
    std::vector<uint64_t> tickets((size_t)n, 0);
    const auto t0 = clk::now();
    qk_status first_err = QK_OK;
    for (int32_t i = 0; i < n; ++i) {
        const auto due = t0 + std::chrono::duration_cast<clk::duration>(std::chrono::duration<double>(arrival_s[i]));
        while (clk::now() < due) {
            const auto left = due - clk::now();
            if (left > std::chrono::microseconds(200)) std::this_thread::sleep_for(left - std::chrono::microseconds(100));
        }
        const qk_status st = qk_server_submit(s, &cs[i], &ps[i], &tickets[i]);
        if (st != QK_OK) {
            out[i] = qk_result{};
            out[i].status = st;
            if (first_err == QK_OK) first_err = st;
        }
    }
    for (int32_t i = 0; i < n; ++i) {
        if (!tickets[i]) continue;
        const qk_status st = qk_server_wait(s, tickets[i], &out[i]);
        if (st != QK_OK && first_err == QK_OK) first_err = st;
    }
    return first_err;
}
