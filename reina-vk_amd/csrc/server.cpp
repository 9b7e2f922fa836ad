// server.cpp — request loop of the batching / serving mode (reina_rt_server_*,
// reina_rt_submit / poll / wait).
//
// The worker thread alternates, at graph boundaries (graph_iters bounce
// iterations, a multiple of 6):
//   1. admission: waiting requests (in submission order, or least estimated
//      work first with an age bound) take a request index (queue words carry
//      it in their top 8 bits) and a contiguous run of slot tiles (first
//      fit); their JobRec is copied to the device table and k_init_job
//      appends their camera rays to the current queue;
//   2. one replay of the captured graph: graph_iters x (k_trace MULTI,
//      k_shade_multi);
//   3. a read-back of the per-request finished-pixel counters; complete
//      requests get k_finish_job (image write) and release their slots.
//   4. when at most tail_paths paths are live, the tail launch takes them
//      over: k_tail_gather moves them to the tail queue (the shared queues
//      are then empty) and k_tail MULTI runs on the tail stream, while the
//      loop admits new requests and runs their graphs beside it (its grid is
//      a share of the GPU's resident blocks when requests are waiting at its
//      launch, else the whole GPU); the tail's requests complete once its event
//      has fired. One tail launch is in flight at a time; meanwhile the
//      shared queues' paths keep running in graphs whatever their number.
// Iteration i uses control slot i % 3 and queue i & 1; every boundary is a
// multiple of 6, so admission always appends to control slot 0 / queue 0.
#include <hip/hip_runtime.h>

#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <mutex>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "api_internal.h"

using namespace rt;

namespace {


struct Request {
    ReinaScene* scene = nullptr;
    TraceParams P{};
    float4* img = nullptr;
    uint32_t npx = 0;
    uint32_t state = REINA_RT_REQ_QUEUED;
    double submit_s = 0, admit_s = 0, done_s = 0;
    double work = 0;   // estimated cost: path slots x samples x max bounces
    double chain = 0;  // per-pixel sample chain: samples x max bounces
    uint32_t job = 0, tile0 = 0, ntiles = 0;
};

}  // namespace

struct ReinaServer {
    int device = 0;
    bool quantized = false;
    bool nee = false;            // holds NEE state and shadow queues (REINA_RT_SERVER_NEE)
    Tuning tune{};
    uint32_t tile_log2 = 10;     // slot tile: 1 << tile_log2 pixels
    uint32_t tiles = 0;          // slot tiles (multiple of kParts)
    uint32_t graph_iters = 6;    // bounce iterations per graph replay
    uint32_t admit_order = REINA_RT_ADMIT_FIFO;
    double max_wait_s = 0.5;     // REINA_RT_ADMIT_SHORTEST: age that takes precedence
    uint32_t max_pending = 0;
    uint32_t max_requests = 0;   // queued + running bound (0: none)
    uint32_t running = 0;        // admitted, not yet complete (mu)
    uint64_t admit_paths = 0;    // admission target (UINT64_MAX: none)
    uint64_t live_paths = 0;     // live paths after the last graph (worker thread only)
    // device memory
    void* dbuf = nullptr;
    PathState S{};
    Queues Q{};
    Ctrl* ctrl = nullptr;
    unsigned long long* stats = nullptr;
    uint32_t* done_d = nullptr;
    JobRec* jobs_d = nullptr;
    // pinned host memory
    JobRec* jobs_h = nullptr;
    uint32_t* done_h = nullptr;
    unsigned long long* stats_h = nullptr;
    uint32_t* count_h = nullptr;         // ctrl[0].count after a graph: the live paths per partition
    hipStream_t st = nullptr;
    hipGraphExec_t graph = nullptr;
    // tail launch beside the loop (not REINA_RT_SERVER_TAIL_INLINE)
    bool tail_async = false;
    hipStream_t st_tail = nullptr;
    uint32_t tail_max_blocks = 0;             // grid cap of the concurrent tail launch
    bool tail_share_fixed = false;            // tail_cu_percent given: the cap holds even with nothing waiting
    hipEvent_t ev_gathered = nullptr, ev_tail = nullptr;
    Queues TQ{};                              // tail queue: ray[0], pix[0], cap entries per partition
    Ctrl* tctrl = nullptr;
    void* tbuf = nullptr;
    // bookkeeping (mu)
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::deque<uint64_t> pending;
    std::unordered_map<uint64_t, Request> reqs;
    std::vector<uint8_t> tile_used;
    std::vector<uint64_t> job_ticket;   // per request index: ticket, 0 = free
    uint64_t next_ticket = 1;
    bool stop = false;
    bool failed = false;
    std::string error;
    ReinaServerStats counts{};
    std::chrono::steady_clock::time_point t0;
    std::thread worker;

    double now() const {
        return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    }
};

static int first_fit(const std::vector<uint8_t>& used, uint32_t n) {
    uint32_t run = 0;
    for (uint32_t t = 0; t < used.size(); ++t) {
        run = used[t] ? 0 : run + 1;
        if (run == n) return (int)(t + 1 - n);
    }
    return -1;
}

// Records the graph of kGraphIters bounce iterations over the shared queues.
static hipError_t capture_graph(ReinaServer* s) {
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamBeginCapture(s->st, hipStreamCaptureModeThreadLocal);
    if (e != hipSuccess) return e;
    for (uint32_t i = 0; i < s->graph_iters; ++i) {
        const uint32_t slot = i % 3, in_q = i & 1u;
        TraceInput tin{&s->ctrl[slot].count[0][0], s->Q.cap, 0, &s->ctrl[slot].fetch[0][0], &s->ctrl[slot].done[0]};
        launch_trace_multi(s->quantized, s->Q.ray[in_q], s->Q.hit, tin, s->stats, s->tune, s->Q.pix[in_q],
                           s->jobs_d, s->st);
        launch_shade_multi(s->jobs_d, s->S, s->Q, in_q, s->ctrl, slot, s->done_d, s->tune.shade_grid, s->nee, s->st);
        if (s->nee)
            launch_connect_multi(s->quantized, s->jobs_d, s->Q, s->S, s->ctrl, slot, s->Q.pix[in_q], s->stats, s->tune,
                                 s->st);
    }
    e = hipStreamEndCapture(s->st, &g);
    if (e != hipSuccess) return e;
    e = hipGraphInstantiate(&s->graph, g, nullptr, nullptr, 0);
    hipGraphDestroy(g);
    return e;
}

#define W_TRY(expr)                                                                  \
    do {                                                                             \
        hipError_t e_ = (expr);                                                      \
        if (e_ != hipSuccess) {                                                      \
            fail_worker(s, std::string(#expr) + ": " + hipGetErrorString(e_));      \
            return;                                                                  \
        }                                                                            \
    } while (0)

// A device error ends the server: every queued and running request fails.
static void fail_worker(ReinaServer* s, const std::string& msg) {
    std::lock_guard<std::mutex> lk(s->mu);
    s->failed = true;
    s->error = msg;
    for (auto& kv : s->reqs)
        if (kv.second.state != REINA_RT_REQ_DONE) kv.second.state = REINA_RT_REQ_FAILED;
    s->pending.clear();
    s->cv_done.notify_all();
}

// Position in s->pending (submission order) of the next request to admit.
// FIFO: the oldest. SHORTEST / LONGEST: the oldest if it has waited
// max_wait_s or longer, else the least estimated work / the longest sample
// chain, then the most work (earliest on ties).
static size_t next_pending(ReinaServer* s, double now) {
    if (s->admit_order == REINA_RT_ADMIT_FIFO) return 0;
    if (now - s->reqs[s->pending.front()].submit_s >= s->max_wait_s) return 0;
    size_t best = 0;
    const Request* b = &s->reqs[s->pending[0]];
    for (size_t i = 1; i < s->pending.size(); ++i) {
        const Request* r = &s->reqs[s->pending[i]];
        const bool better = s->admit_order == REINA_RT_ADMIT_SHORTEST
                                ? r->work < b->work
                                : (r->chain > b->chain || (r->chain == b->chain && r->work > b->work));
        if (better) { best = i; b = r; }
    }
    return best;
}

// Writes the images of finished requests (k_finish_job on the worker's
// stream), then releases their slots and request indices and wakes waiters;
// they are removed from `list`.
static bool finish_requests(ReinaServer* s, const std::vector<uint64_t>& fin, std::vector<uint64_t>& list) {
    if (fin.empty()) return true;
    for (uint64_t tk : fin) {
        Request r;
        {
            std::lock_guard<std::mutex> lk(s->mu);
            r = s->reqs[tk];
        }
        const hipError_t e = launch_finish_job(s->jobs_d, r.job, r.npx, s->S, s->st);
        if (e != hipSuccess) {
            fail_worker(s, std::string("launch_finish_job: ") + hipGetErrorString(e));
            return false;
        }
    }
    const hipError_t e = hipStreamSynchronize(s->st);
    if (e != hipSuccess) {
        fail_worker(s, std::string("hipStreamSynchronize: ") + hipGetErrorString(e));
        return false;
    }
    std::lock_guard<std::mutex> lk(s->mu);
    const double t = s->now();
    for (uint64_t tk : fin) {
        Request& r = s->reqs[tk];
        r.state = REINA_RT_REQ_DONE;
        r.done_s = t;
        for (uint32_t k = 0; k < r.ntiles; ++k) s->tile_used[r.tile0 + k] = 0;
        s->counts.slot_tiles_used -= r.ntiles;
        s->job_ticket[r.job] = 0;
        s->running -= 1;
        s->counts.completed += 1;
        for (size_t k = 0; k < list.size(); ++k)
            if (list[k] == tk) { list.erase(list.begin() + (std::ptrdiff_t)k); break; }
    }
    s->cv_done.notify_all();
    return true;
}

// Reads back the finished-pixel counters and ray counts (worker's stream).
static bool read_counters(ReinaServer* s) {
    hipError_t e = hipMemcpyAsync(s->done_h, s->done_d, kMaxJobs * sizeof(uint32_t), hipMemcpyDeviceToHost, s->st);
    if (e == hipSuccess)
        e = hipMemcpyAsync(s->stats_h, s->stats, 8 * sizeof(unsigned long long), hipMemcpyDeviceToHost, s->st);
    if (e == hipSuccess) e = hipStreamSynchronize(s->st);
    if (e != hipSuccess) {
        fail_worker(s, std::string("counter read-back: ") + hipGetErrorString(e));
        return false;
    }
    std::lock_guard<std::mutex> lk(s->mu);
    s->counts.extension_rays = s->stats_h[0];
    s->counts.tail_rays = s->stats_h[4];
    return true;
}

// The tail launch in flight has ended: every request it held is complete.
static bool end_tail(ReinaServer* s, std::vector<uint64_t>& tailing) {
    if (!read_counters(s)) return false;
    bool all = true;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        for (uint64_t tk : tailing) {
            const Request& r = s->reqs[tk];
            all = all && s->done_h[r.job] >= r.npx;
        }
    }
    if (!all) {   // the tail launch runs every path it holds to the end
        fail_worker(s, "a request left the tail launch unfinished");
        return false;
    }
    const std::vector<uint64_t> fin = tailing;
    return finish_requests(s, fin, tailing);
}

static void worker_main(ReinaServer* s) {
    if (hipSetDevice(s->device) != hipSuccess) {
        fail_worker(s, "hipSetDevice failed in the server thread");
        return;
    }
    std::vector<uint64_t> running;   // admitted requests whose paths are in the shared queues
    std::vector<uint64_t> tailing;   // requests the tail launch in flight is finishing
    bool tail_busy = false;
    auto busy_from = std::chrono::steady_clock::now();
    auto add_busy = [&]() {
        std::lock_guard<std::mutex> lk(s->mu);
        s->counts.busy_s += std::chrono::duration<double>(std::chrono::steady_clock::now() - busy_from).count();
    };
    for (;;) {
        // a tail launch that has ended: its requests complete
        if (tail_busy) {
            const hipError_t q = hipEventQuery(s->ev_tail);
            if (q == hipSuccess) {
                if (!end_tail(s, tailing)) return;
                tail_busy = false;
                if (running.empty()) add_busy();
            } else if (q != hipErrorNotReady) {
                fail_worker(s, std::string("tail launch: ") + hipGetErrorString(q));
                return;
            }
        }
        std::vector<Request> admitted;
        {
            std::unique_lock<std::mutex> lk(s->mu);
            if (running.empty() && s->pending.empty()) {
                if (s->stop && !tail_busy) break;
                s->cv_work.wait_for(lk, std::chrono::milliseconds(tail_busy ? 1 : 20));
                continue;
            }
            if (running.empty() && !tail_busy) busy_from = std::chrono::steady_clock::now();
            // paths in flight: the shared queues' (a tail launch holds its own)
            uint64_t inflight = running.empty() ? 0 : s->live_paths;
            const double tnow = s->now();
            while (!s->pending.empty()) {
                const size_t pi = next_pending(s, tnow);
                const uint64_t tk = s->pending[pi];
                Request& r = s->reqs[tk];
                if (!running.empty() && inflight + r.npx > s->admit_paths) break;
                int job = -1;
                for (uint32_t j = 0; j < kMaxJobs; ++j)
                    if (s->job_ticket[j] == 0) { job = (int)j; break; }
                if (job < 0) break;
                const int t0 = first_fit(s->tile_used, r.ntiles);
                if (t0 < 0) break;
                for (uint32_t t = 0; t < r.ntiles; ++t) s->tile_used[t0 + t] = 1;
                s->counts.slot_tiles_used += r.ntiles;
                r.job = (uint32_t)job;
                r.tile0 = (uint32_t)t0;
                r.state = REINA_RT_REQ_RUNNING;
                r.admit_s = s->now();
                s->job_ticket[job] = tk;
                s->running += 1;
                running.push_back(tk);
                admitted.push_back(r);
                inflight += r.npx;
                s->pending.erase(s->pending.begin() + (std::ptrdiff_t)pi);
            }
        }
        for (const Request& r : admitted) {
            JobRec& jr = s->jobs_h[r.job];
            jr.sv = r.scene->sv;
            jr.P = r.P;
            jr.img = (float4 RT_GLOBAL*)r.img;
            jr.px_base = r.tile0 << s->tile_log2;
            jr.npx = r.npx;
            W_TRY(hipMemcpyAsync(s->jobs_d + r.job, &jr, sizeof(JobRec), hipMemcpyHostToDevice, s->st));
            W_TRY(hipMemsetAsync(s->done_d + r.job, 0, sizeof(uint32_t), s->st));
            W_TRY(launch_init_job(s->jobs_d, r.job, r.npx, s->S, s->Q, s->ctrl, 0, 0, s->done_d, s->tile_log2,
                                  s->st));
        }
        if (running.empty()) continue;
        W_TRY(hipGraphLaunch(s->graph, s->st));
        W_TRY(hipMemcpyAsync(s->count_h, s->ctrl[0].count, sizeof(Ctrl::count), hipMemcpyDeviceToHost, s->st));
        if (!read_counters(s)) return;
        {
            std::lock_guard<std::mutex> lk(s->mu);
            s->counts.iterations += s->graph_iters;
            s->counts.graph_replays += 1;
        }
        uint64_t live = 0;
        uint32_t maxc = 0;
        for (uint32_t p = 0; p < kParts; ++p) {
            const uint32_t c = s->count_h[p * kLine];
            live += c;
            maxc = c > maxc ? c : maxc;
        }
        s->live_paths = live;
        bool to_tail = false;
        if (s->tune.tail_paths != 0 && live != 0 && live <= s->tune.tail_paths) {
            if (!s->tail_async) {
                // Few paths left: one tail launch finishes every running request (the
                // graph leaves the next iteration's rays in queue 0, control slot 0).
                W_TRY(launch_tail_multi(s->quantized, s->nee, s->jobs_d, s->S, s->Q, 0, s->ctrl, 0, s->stats,
                                        s->done_d, s->tune, (uint32_t)live, s->st));
                W_TRY(hipMemsetAsync(&s->ctrl[0], 0, sizeof(Ctrl), s->st));
                if (!read_counters(s)) return;
                s->live_paths = 0;
                std::lock_guard<std::mutex> lk(s->mu);
                s->counts.tail_launches += 1;
            } else if (!tail_busy) {
                // The live paths move to the tail queue and finish on the tail stream; the
                // shared queues are empty again and take new requests at once. (While a
                // tail launch is in flight, the few live paths keep running in graphs.)
                // with nothing waiting to run beside it, the tail launch takes the whole GPU
                // (unless tail_cu_percent fixes its share)
                bool waiting = s->tail_share_fixed;
                {
                    std::lock_guard<std::mutex> lk(s->mu);
                    waiting = waiting || !s->pending.empty();
                }
                W_TRY(launch_tail_gather(s->Q, s->ctrl, s->TQ, s->tctrl, maxc, s->st));
                W_TRY(hipMemsetAsync(&s->ctrl[0], 0, sizeof(Ctrl), s->st));
                W_TRY(hipEventRecord(s->ev_gathered, s->st));
                W_TRY(hipStreamWaitEvent(s->st_tail, s->ev_gathered, 0));
                W_TRY(launch_tail_multi(s->quantized, s->nee, s->jobs_d, s->S, s->TQ, 0, s->tctrl, 0, s->stats,
                                        s->done_d, s->tune, (uint32_t)live, s->st_tail,
                                        waiting ? s->tail_max_blocks : 0u));
                W_TRY(hipEventRecord(s->ev_tail, s->st_tail));
                tail_busy = true;
                to_tail = true;
                s->live_paths = 0;
                std::lock_guard<std::mutex> lk(s->mu);
                s->counts.tail_launches += 1;
            }
        }
        std::vector<uint64_t> finished;
        {
            std::lock_guard<std::mutex> lk(s->mu);
            for (uint64_t tk : running) {
                const Request& r = s->reqs[tk];
                if (s->done_h[r.job] >= r.npx) finished.push_back(tk);
            }
        }
        if (!finish_requests(s, finished, running)) return;
        if (to_tail) {
            tailing.insert(tailing.end(), running.begin(), running.end());
            running.clear();
        }
        if (running.empty() && !tail_busy) add_busy();
    }
}

static void free_server(ReinaServer* s) {
    if (s->graph) hipGraphExecDestroy(s->graph);
    if (s->st) hipStreamDestroy(s->st);
    if (s->st_tail) hipStreamDestroy(s->st_tail);
    if (s->ev_gathered) hipEventDestroy(s->ev_gathered);
    if (s->ev_tail) hipEventDestroy(s->ev_tail);
    if (s->tbuf) hipFree(s->tbuf);
    if (s->tctrl) hipFree(s->tctrl);
    if (s->dbuf) hipFree(s->dbuf);
    if (s->ctrl) hipFree(s->ctrl);
    if (s->jobs_h) hipHostFree(s->jobs_h);
    if (s->done_h) hipHostFree(s->done_h);
    if (s->stats_h) hipHostFree(s->stats_h);
    if (s->count_h) hipHostFree(s->count_h);
    delete s;
}

extern "C" {

int reina_rt_server_create(const ReinaServerDesc* d, ReinaServer** out) {
    if (!d || !out) return fail(REINA_RT_ERR_INVALID_ARG, "null argument");
    *out = nullptr;
    if (d->struct_size != sizeof(ReinaServerDesc))
        return fail(REINA_RT_ERR_ABI, "ReinaServerDesc.struct_size mismatch (expected " +
                                          std::to_string(sizeof(ReinaServerDesc)) + ")");
    if (d->flags & ~(REINA_RT_SERVER_NEE | REINA_RT_SERVER_TAIL_INLINE))
        return fail(REINA_RT_ERR_INVALID_ARG, "unknown server flag bits");
    if (d->tail_cu_percent > 100) return fail(REINA_RT_ERR_INVALID_ARG, "tail_cu_percent must be 0..100");
    const bool nee = (d->flags & REINA_RT_SERVER_NEE) != 0;
    const uint64_t bpp = nee ? REINA_RT_SERVER_BYTES_PER_PIXEL_NEE : REINA_RT_SERVER_BYTES_PER_PIXEL;
    const uint32_t tile = d->slot_tile ? d->slot_tile : 1024u;
    if (tile != 256 && tile != 512 && tile != 1024 && tile != 2048 && tile != 4096)
        return fail(REINA_RT_ERR_INVALID_ARG, "slot_tile must be 0, 256, 512, 1024, 2048 or 4096");
    const uint32_t tile_log2 = (uint32_t)__builtin_ctz(tile);
    const uint32_t graph_iters = d->graph_iters ? d->graph_iters : 6u;
    if (graph_iters % 6 != 0 || graph_iters > 24)
        return fail(REINA_RT_ERR_INVALID_ARG, "graph_iters must be 0, 6, 12, 18 or 24");
    if (d->admit_order != REINA_RT_ADMIT_FIFO && d->admit_order != REINA_RT_ADMIT_SHORTEST &&
        d->admit_order != REINA_RT_ADMIT_LONGEST)
        return fail(REINA_RT_ERR_INVALID_ARG, "admit_order must be REINA_RT_ADMIT_FIFO, _SHORTEST or _LONGEST");
    const uint64_t per_tile = bpp * tile;
    uint64_t tiles = d->path_budget_bytes / per_tile;
    tiles -= tiles % kParts;
    if (tiles < kParts) return fail(REINA_RT_ERR_INVALID_ARG, "path_budget_bytes below 8 slot tiles");
    if (tiles * tile > (1ull << kSlotBits)) tiles = (1ull << kSlotBits) / tile;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return fail(REINA_RT_ERR_DEVICE, "no HIP device available");
    if ((int)d->device >= ndev) return fail(REINA_RT_ERR_INVALID_ARG, "device ordinal out of range");
    Tuning tu{d->tuning.stack_depth, d->tuning.refill_min, d->tuning.grid_blocks, d->tuning.shade_blocks,
              d->tuning.waves_per_simd, d->tuning.tail_paths};
    if (tu.tail_paths == 0) tu.tail_paths = kDefaultTailPaths;
    else if (tu.tail_paths == REINA_RT_TAIL_OFF) tu.tail_paths = 0;   // internally 0 = never
    if (tu.stack_depth == 0) tu.stack_depth = 16;
    if (tu.stack_depth != 8 && tu.stack_depth != 16 && tu.stack_depth != 24 && tu.stack_depth != 32)
        return fail(REINA_RT_ERR_INVALID_ARG, "stack_depth must be 8, 16, 24 or 32");
    if (tu.refill_min == 0) tu.refill_min = 32;
    if (tu.refill_min > 64) return fail(REINA_RT_ERR_INVALID_ARG, "refill_min must be 1..64");
    if (tu.waves_per_simd != 0 && tu.waves_per_simd != 4 && tu.waves_per_simd != 5 && tu.waves_per_simd != 6 && tu.waves_per_simd != 8)
        return fail(REINA_RT_ERR_INVALID_ARG, "waves_per_simd must be 0, 4, 5, 6 or 8");
    tu.claim = d->tuning.claim_rays ? d->tuning.claim_rays : 128u;
    if (tu.claim != 64 && tu.claim != 128 && tu.claim != 256 && tu.claim != 512)
        return fail(REINA_RT_ERR_INVALID_ARG, "claim_rays must be 0, 64, 128, 256 or 512");
    ReinaServer* s = new ReinaServer();
    s->device = (int)d->device;
    s->quantized = d->quantized != 0;
    s->nee = nee;
    s->tiles = (uint32_t)tiles;
    s->tile_log2 = tile_log2;
    s->graph_iters = graph_iters;
    s->admit_order = d->admit_order;
    s->max_wait_s = (d->max_wait_ms ? d->max_wait_ms : 500u) * 1e-3;
    s->max_pending = d->max_pending ? d->max_pending : 1024;
    s->max_requests = d->max_requests;
    s->admit_paths = (d->admit_paths == 0 || d->admit_paths == REINA_RT_ADMIT_ALL) ? UINT64_MAX : d->admit_paths;
    s->t0 = std::chrono::steady_clock::now();
    auto bail = [&](int code, const std::string& msg) {
        free_server(s);
        return fail(code, msg);
    };
    if (hipSetDevice(s->device) != hipSuccess) return bail(REINA_RT_ERR_DEVICE, "hipSetDevice failed");
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, s->device) != hipSuccess) return bail(REINA_RT_ERR_DEVICE, "device query failed");
    const uint32_t cus = (uint32_t)prop.multiProcessorCount;
    if (tu.grid == 0) tu.grid = cus * (tu.stack_depth <= 8 ? 8 : 4);
    if (tu.shade_grid == 0) tu.shade_grid = cus * 16;
    s->tune = tu;
    const uint64_t slots = tiles * tile;
    const uint64_t bytes = slots * bpp;
    if (hipMalloc(&s->dbuf, bytes) != hipSuccess)
        return bail(REINA_RT_ERR_OUT_OF_MEMORY, "hipMalloc of " + std::to_string(bytes) + " bytes failed");
    char* p = (char*)s->dbuf;
    s->S.st0 = (uint4*)p; p += 16 * slots;
    s->S.st1 = (float4*)p; p += 16 * slots;
    s->S.st2 = (float4*)p; p += 16 * slots;
    s->S.st3 = (float4*)p; p += 16 * slots;
    s->Q.ray[0] = (float4*)p; p += 32 * slots;
    s->Q.ray[1] = (float4*)p; p += 32 * slots;
    s->Q.hit = (float4*)p; p += 16 * slots;
    s->Q.pix[0] = (uint32_t*)p; p += 4 * slots;
    s->Q.pix[1] = (uint32_t*)p; p += 4 * slots;
    if (nee) {
        s->S.st4 = (float4*)p; p += 16 * slots;
        s->Q.sray = (float4*)p; p += 32 * slots;
        s->Q.scon = (float4*)p; p += 16 * slots;
    }
    s->Q.cap = (uint32_t)(slots / kParts);
    const size_t cbytes = 3 * sizeof(Ctrl) + 256 + kMaxJobs * sizeof(uint32_t) + kMaxJobs * sizeof(JobRec);
    if (hipMalloc(&s->ctrl, cbytes) != hipSuccess) return bail(REINA_RT_ERR_OUT_OF_MEMORY, "hipMalloc failed");
    s->stats = (unsigned long long*)((char*)s->ctrl + 3 * sizeof(Ctrl));
    s->done_d = (uint32_t*)((char*)s->ctrl + 3 * sizeof(Ctrl) + 256);
    s->jobs_d = (JobRec*)((char*)s->done_d + kMaxJobs * sizeof(uint32_t));
    if (hipMemset(s->ctrl, 0, cbytes) != hipSuccess) return bail(REINA_RT_ERR_DEVICE, "hipMemset failed");
    if (hipHostMalloc((void**)&s->jobs_h, kMaxJobs * sizeof(JobRec), 0) != hipSuccess ||
        hipHostMalloc((void**)&s->done_h, kMaxJobs * sizeof(uint32_t), 0) != hipSuccess ||
        hipHostMalloc((void**)&s->stats_h, 8 * sizeof(unsigned long long), 0) != hipSuccess ||
        hipHostMalloc((void**)&s->count_h, sizeof(Ctrl::count), 0) != hipSuccess)
        return bail(REINA_RT_ERR_OUT_OF_MEMORY, "hipHostMalloc failed");
    if (hipStreamCreate(&s->st) != hipSuccess) return bail(REINA_RT_ERR_DEVICE, "hipStreamCreate failed");
    s->tail_async = !(d->flags & REINA_RT_SERVER_TAIL_INLINE) && tu.tail_paths != 0;
    if (s->tail_async) {
        // tail queue: every live path of a tail launch keeps its partition and position
        const uint64_t tcap = tu.tail_paths < slots / kParts ? tu.tail_paths : slots / kParts;
        s->TQ.cap = (uint32_t)tcap;
        const uint64_t tbytes = kParts * tcap * (32 + 4);
        if (hipMalloc(&s->tbuf, tbytes) != hipSuccess)
            return bail(REINA_RT_ERR_OUT_OF_MEMORY, "hipMalloc of the " + std::to_string(tbytes) + "-byte tail queue failed");
        s->TQ.ray[0] = (float4*)s->tbuf;
        s->TQ.pix[0] = (uint32_t*)((char*)s->tbuf + kParts * tcap * 32);
        if (hipMalloc(&s->tctrl, sizeof(Ctrl)) != hipSuccess || hipMemset(s->tctrl, 0, sizeof(Ctrl)) != hipSuccess)
            return bail(REINA_RT_ERR_OUT_OF_MEMORY, "hipMalloc failed");
        // The tail launch takes tail_cu_percent of the blocks the tail kernel can keep
        // resident (3 per CU at its 3-wave register budget, 2 with NEE); the graphs'
        // kernels fill the rest of the GPU.
        const uint32_t pct = d->tail_cu_percent ? d->tail_cu_percent : 50u;
        const uint64_t full = (uint64_t)cus * (nee ? 2u : 3u);
        s->tail_max_blocks = (uint32_t)(full * pct / 100u);
        if (s->tail_max_blocks < kParts) s->tail_max_blocks = kParts;
        s->tail_share_fixed = d->tail_cu_percent != 0;
        const hipError_t e = hipStreamCreate(&s->st_tail);
        if (e != hipSuccess) return bail(REINA_RT_ERR_DEVICE, std::string("tail stream: ") + hipGetErrorString(e));
        if (hipEventCreateWithFlags(&s->ev_gathered, hipEventDisableTiming) != hipSuccess ||
            hipEventCreateWithFlags(&s->ev_tail, hipEventDisableTiming) != hipSuccess)
            return bail(REINA_RT_ERR_DEVICE, "hipEventCreate failed");
    }
    hipError_t e = capture_graph(s);
    if (e != hipSuccess) return bail(REINA_RT_ERR_DEVICE, std::string("graph capture: ") + hipGetErrorString(e));
    s->tile_used.assign(s->tiles, 0);
    s->job_ticket.assign(kMaxJobs, 0);
    s->counts.slot_tiles_total = s->tiles;
    s->worker = std::thread(worker_main, s);
    *out = s;
    return REINA_RT_OK;
}

int reina_rt_server_destroy(ReinaServer* s) {
    if (!s) return fail(REINA_RT_ERR_INVALID_ARG, "server is null");
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->stop = true;
    }
    s->cv_work.notify_all();
    if (s->worker.joinable()) s->worker.join();
    hipSetDevice(s->device);
    hipStreamSynchronize(s->st);
    if (s->st_tail) hipStreamSynchronize(s->st_tail);
    free_server(s);
    return REINA_RT_OK;
}

int reina_rt_submit(ReinaServer* s, ReinaScene* scene, const RtPushConsts* pc, float* image, uint32_t W, uint32_t H,
                    uint64_t* ticket) {
    if (!s || !scene || !pc || !image || !ticket) return fail(REINA_RT_ERR_INVALID_ARG, "null argument");
    if (W == 0 || H == 0) return fail(REINA_RT_ERR_INVALID_ARG, "zero image size");
    if ((uint64_t)W * H >= (1ull << 31)) return fail(REINA_RT_ERR_INVALID_ARG, "image too large");
    if (scene->device != s->device) return fail(REINA_RT_ERR_INVALID_ARG, "scene lives on another device");
    if (scene->quantized != s->quantized)
        return fail(REINA_RT_ERR_UNSUPPORTED, "scene node format differs from the server's");
    const uint32_t groups = 1u << ((scene->options & REINA_RT_TRACE_SAMPLE_GROUPS_MASK) >> 8);
    if (pc->samplesPerPixel % groups != 0u)
        return fail(REINA_RT_ERR_INVALID_ARG, "samplesPerPixel " + std::to_string(pc->samplesPerPixel) +
                                                  " is not a multiple of the " + std::to_string(groups) +
                                                  " sample groups");
    if ((uint64_t)W * H * groups >= (1ull << 31)) return fail(REINA_RT_ERR_INVALID_ARG, "image x sample groups too large");
    const bool nee = (scene->options & REINA_RT_TRACE_NEE) != 0;
    if (nee && !s->nee)
        return fail(REINA_RT_ERR_UNSUPPORTED, "the scene asks for next-event estimation; the server was created "
                                              "without REINA_RT_SERVER_NEE");
    if (nee && scene->emi.empty())
        return fail(REINA_RT_ERR_NO_EMISSIVE, "next-event estimation needs at least one emissive instance");
    if (scene->info.max_depth > s->tune.stack_depth + 1)
        return fail(REINA_RT_ERR_UNSUPPORTED, "scene BVH depth " + std::to_string(scene->info.max_depth) +
                                                  " exceeds the server's stack depth " +
                                                  std::to_string(s->tune.stack_depth));
    const uint32_t npx = W * H * groups;   // path slots: one per pixel and sample group
    const uint32_t ntiles = (uint32_t)(((uint64_t)npx + (1u << s->tile_log2) - 1u) >> s->tile_log2);
    std::lock_guard<std::mutex> lk(s->mu);
    if (s->failed) return fail(REINA_RT_ERR_DEVICE, "server stopped after a device error: " + s->error);
    if (s->stop) return fail(REINA_RT_ERR_INVALID_ARG, "server is shutting down");
    s->counts.submitted += 1;
    if (ntiles > s->tiles) {
        s->counts.rejected_budget += 1;
        return fail(REINA_RT_ERR_OVER_BUDGET, "request needs " + std::to_string(ntiles) + " slot tiles, server has " +
                                                  std::to_string(s->tiles));
    }
    if (s->pending.size() >= s->max_pending) {
        s->counts.rejected_busy += 1;
        return fail(REINA_RT_ERR_BUSY, "request queue full");
    }
    if (s->max_requests && s->pending.size() + s->running >= s->max_requests) {
        s->counts.rejected_busy += 1;
        return fail(REINA_RT_ERR_BUSY, "server holds max_requests requests");
    }
    const uint64_t tk = s->next_ticket++;
    Request r;
    r.scene = scene;
    r.P = make_trace_params(pc, W, H);
    r.P.nee = nee ? 1u : 0u;
    r.P.fixes = scene->options & REINA_RT_TRACE_FIXES;
    r.P.groups = groups;
    r.P.spp = pc->samplesPerPixel / groups;
    r.img = (float4*)image;
    r.npx = npx;
    r.ntiles = ntiles;
    r.chain = (double)r.P.spp * (r.P.maxBounces ? r.P.maxBounces : 1u);
    r.work = (double)npx * r.chain;
    r.submit_s = s->now();
    s->reqs[tk] = r;
    s->pending.push_back(tk);
    *ticket = tk;
    s->cv_work.notify_one();
    return REINA_RT_OK;
}

static void fill_status(const Request& r, ReinaRequestStatus* o) {
    o->state = r.state;
    o->npx = r.P.W * r.P.H;
    o->submit_s = r.submit_s;
    o->admit_s = r.admit_s;
    o->done_s = r.done_s;
}

int reina_rt_poll(ReinaServer* s, uint64_t ticket, ReinaRequestStatus* o) {
    if (!s || !o) return fail(REINA_RT_ERR_INVALID_ARG, "null argument");
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->reqs.find(ticket);
    if (it == s->reqs.end()) return fail(REINA_RT_ERR_INVALID_ARG, "unknown or released ticket");
    fill_status(it->second, o);
    if (it->second.state == REINA_RT_REQ_DONE || it->second.state == REINA_RT_REQ_FAILED) s->reqs.erase(it);
    if (o->state == REINA_RT_REQ_FAILED) return fail(REINA_RT_ERR_DEVICE, s->error);
    return REINA_RT_OK;
}

int reina_rt_wait(ReinaServer* s, uint64_t ticket, uint32_t timeout_ms, ReinaRequestStatus* o) {
    if (!s || !o) return fail(REINA_RT_ERR_INVALID_ARG, "null argument");
    std::unique_lock<std::mutex> lk(s->mu);
    auto done = [&]() {
        auto it = s->reqs.find(ticket);
        return it == s->reqs.end() || it->second.state == REINA_RT_REQ_DONE || it->second.state == REINA_RT_REQ_FAILED;
    };
    if (!s->cv_done.wait_for(lk, std::chrono::milliseconds(timeout_ms), done))
        return fail(REINA_RT_ERR_BUSY, "timed out waiting for the request");
    auto it = s->reqs.find(ticket);
    if (it == s->reqs.end()) return fail(REINA_RT_ERR_INVALID_ARG, "unknown or released ticket");
    fill_status(it->second, o);
    s->reqs.erase(it);
    if (o->state == REINA_RT_REQ_FAILED) return fail(REINA_RT_ERR_DEVICE, s->error);
    return REINA_RT_OK;
}

int reina_rt_server_stats(ReinaServer* s, ReinaServerStats* o) {
    if (!s || !o) return fail(REINA_RT_ERR_INVALID_ARG, "null argument");
    std::lock_guard<std::mutex> lk(s->mu);
    *o = s->counts;
    o->pending = (uint32_t)s->pending.size();
    uint32_t running = 0;
    for (uint64_t t : s->job_ticket) running += t != 0;
    o->running = running;
    return REINA_RT_OK;
}

}  // extern "C"
