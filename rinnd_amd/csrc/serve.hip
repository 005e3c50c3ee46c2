// serve.hip - the query service: asynchronous batch submission over a resident index (include/rinnd_hip.h).
//
// The reference answers queries on the caller's thread (NNDescentIndex::query, core/index.rs:406-439) and
// serialises single queries on a Mutex<SearchWorkspace> (crates/rinnd/src/lib.rs:571-586).  Here requests from any
// host thread are queued; a dispatcher thread gathers the pending ones into one launch of the greedy-search kernel
// (requests of different k and epsilon share the launch through per-query parameters), and a completion thread
// retires the launches in order.  Memory is fixed at creation: max_inflight launch slots (each with its own stream,
// its query rows, parameter block and output staging for max_batch queries of up to max_k results, and an equal
// share of the pool of visited sets, whose size fixes the number of concurrent searches).  A launch closes once it
// holds target_batch queries, so launches stay short and the slots' launches overlap on the device.
#include <algorithm>
#include <condition_variable>
#include <deque>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "index.h"

using namespace rinnd;

namespace {

constexpr int64_t kServeSeedCap = 1 << 16;  // frontier entries per query beyond the LDS part

struct Request {
    const float* q = nullptr;
    int64_t nq = 0, k = 0, width = 0, rerank = 0;
    float eps = 0.f;
    int32_t* out_i = nullptr;
    float* out_d = nullptr;
    bool dev_out = false;
    int64_t obase = 0;  // host-output requests: first element of its results in the slot's output staging
    bool done = false;
    rinnd_status status = RINND_OK;
    std::string error;
};

// Device parameter block of a launch: per query its id, output pointers, search width, k, rerank width, epsilon.
struct ParamLayout {
    size_t qid, oi, od, k, kout, rerank, eps, bytes;
    explicit ParamLayout(int64_t B) {
        size_t o = 0;
        auto take = [&](size_t b) {
            const size_t r = o;
            o += (b + 255) & ~(size_t)255;
            return r;
        };
        qid = take(8 * B);
        oi = take(8 * B);
        od = take(8 * B);
        k = take(4 * B);
        kout = take(4 * B);
        rerank = take(4 * B);
        eps = take(4 * B);
        bytes = o;
    }
};

struct LaunchSlot {
    hipStream_t stream = nullptr;  // launches of different slots run concurrently
    int64_t vis0 = 0;              // first visited set of this slot's share of the pool
    DBuf<char> params;
    DBuf<float> Q;   // max_batch rows of ld floats
    DBuf<int32_t> oi;
    DBuf<float> od;  // max_batch x max_k staged results of host-output requests
    DBuf<int32_t> status;
    void* pinned = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    std::vector<std::shared_ptr<Request>> reqs;
    bool busy = false;
};

}  // namespace

struct rinnd_server {
    rinnd_index* ix = nullptr;
    rinnd_server_params p{};
    int device = 0;
    hipStream_t stream = nullptr;
    int visit_slots = 1;  // visited sets per launch slot
    int64_t seed_cap = 1;
    DBuf<uint8_t> visited;
    DBuf<int32_t> epochs;
    DBuf<float> spill_d;
    DBuf<int32_t> spill_i;
    std::vector<LaunchSlot> slots;
    mutable std::mutex mu;
    std::condition_variable cv_work, cv_inflight, cv_done;
    std::deque<std::shared_ptr<Request>> pending;
    std::deque<int> inflight;  // launch slots in launch order
    std::unordered_map<uint64_t, std::shared_ptr<Request>> live;
    uint64_t next_ticket = 1;
    bool stop = false, dispatch_finished = false;
    std::thread dispatcher, completer;
    int64_t target_batch = 1024;
    rinnd_server_stats stats{};
    mutable std::string last_error;
};

namespace {

void fail_batch(rinnd_server* S, std::vector<std::shared_ptr<Request>>& batch, rinnd_status st, const std::string& msg) {
    for (auto& r : batch) {
        r->done = true;
        r->status = st;
        r->error = msg;
    }
}

// Stages a batch into launch slot L and launches the search; throws Error on a HIP failure.
void launch_batch(rinnd_server* S, LaunchSlot& L, std::vector<std::shared_ptr<Request>>& batch) {
    const rinnd_index& ix = *S->ix;
    hipStream_t s = L.stream;
    const ParamLayout P(S->p.max_batch);
    char* hp = (char*)L.pinned;
    int64_t* qid = (int64_t*)(hp + P.qid);
    int32_t** poi = (int32_t**)(hp + P.oi);
    float** pod = (float**)(hp + P.od);
    int32_t* qk = (int32_t*)(hp + P.k);
    int32_t* qko = (int32_t*)(hp + P.kout);
    int32_t* qrr = (int32_t*)(hp + P.rerank);
    float* qe = (float*)(hp + P.eps);
    int64_t row = 0, oofs = 0, maxw = 1;
    for (auto& r : batch) {
        HIP_TRY(hipMemcpy2DAsync(L.Q.p + row * ix.ld, ix.ld * 4, r->q, ix.dim * 4, ix.dim * 4, r->nq, hipMemcpyDefault,
                                 s));
        if (!r->dev_out) r->obase = oofs;
        for (int64_t j = 0; j < r->nq; ++j) {
            qid[row + j] = j;  // the request's own query ids, as rinnd_index_query numbers them
            poi[row + j] = r->dev_out ? r->out_i + j * r->k : L.oi.p + oofs + j * r->k;
            pod[row + j] = r->dev_out ? r->out_d + j * r->k : L.od.p + oofs + j * r->k;
            qk[row + j] = (int32_t)r->width;
            qko[row + j] = (int32_t)r->k;
            qrr[row + j] = (int32_t)r->rerank;
            qe[row + j] = r->eps;
        }
        if (!r->dev_out) oofs += r->nq * r->k;
        row += r->nq;
        maxw = std::max(maxw, r->width);
    }
    HIP_TRY(hipMemcpyAsync(L.params.p, L.pinned, P.bytes, hipMemcpyHostToDevice, s));
    HIP_TRY(hipMemsetAsync(L.status.p, 0, 4, s));
    SearchArgs a = index_search_args(ix, L.Q.p, row, 0, 0.f, 0, maxw, 0,
                                     S->p.quantized == 2 ? kQueryI8 : S->p.quantized ? kQueryFP8 : kQueryF32, nullptr,
                                     nullptr);
    a.q_qid = (const int64_t*)(L.params.p + P.qid);
    a.q_oi = (int32_t* const*)(L.params.p + P.oi);
    a.q_od = (float* const*)(L.params.p + P.od);
    a.q_k = (const int32_t*)(L.params.p + P.k);
    a.q_kout = (const int32_t*)(L.params.p + P.kout);
    a.q_rerank = (const int32_t*)(L.params.p + P.rerank);
    a.q_eps = (const float*)(L.params.p + P.eps);
    a.visited = S->visited.p + L.vis0 * std::max<int64_t>(ix.n, 1);
    a.epochs = S->epochs.p + L.vis0;
    a.seed_spill_d = S->spill_d.p + L.vis0 * S->seed_cap;
    a.seed_spill_i = S->spill_i.p + L.vis0 * S->seed_cap;
    a.seed_cap = S->seed_cap;
    a.status = L.status.p;
    HIP_TRY(hipEventRecord(L.ev0, s));
    launch_search(s, a, (int)std::min<int64_t>(S->visit_slots, std::max<int64_t>(row, 1)));
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipEventRecord(L.ev1, s));
}

void dispatch_loop(rinnd_server* S) {
    (void)hipSetDevice(S->device);
    for (;;) {
        int slot = -1;
        std::vector<std::shared_ptr<Request>> batch;
        {
            std::unique_lock<std::mutex> lk(S->mu);
            auto free_slot = [&]() {
                for (size_t i = 0; i < S->slots.size(); ++i)
                    if (!S->slots[i].busy) return (int)i;
                return -1;
            };
            S->cv_work.wait(lk, [&] { return (S->stop && S->pending.empty()) || (!S->pending.empty() && free_slot() >= 0); });
            if (S->pending.empty()) break;  // stopping, nothing left
            slot = free_slot();
            // arrival batching: the pending requests in order, until the launch holds target_batch queries
            int64_t total = 0;
            while (!S->pending.empty() && total < S->target_batch && total + S->pending.front()->nq <= S->p.max_batch) {
                total += S->pending.front()->nq;
                batch.push_back(std::move(S->pending.front()));
                S->pending.pop_front();
            }
            S->slots[slot].busy = true;
        }
        LaunchSlot& L = S->slots[slot];
        try {
            launch_batch(S, L, batch);
        } catch (const Error& e) {
            (void)hipGetLastError();
            std::lock_guard<std::mutex> lk(S->mu);
            fail_batch(S, batch, e.code, e.msg);
            L.busy = false;
            S->cv_done.notify_all();
            S->cv_work.notify_all();
            continue;
        }
        std::lock_guard<std::mutex> lk(S->mu);
        L.reqs = std::move(batch);
        S->inflight.push_back(slot);
        S->cv_inflight.notify_one();
    }
    std::lock_guard<std::mutex> lk(S->mu);
    S->dispatch_finished = true;
    S->cv_inflight.notify_all();
}

void complete_loop(rinnd_server* S) {
    (void)hipSetDevice(S->device);
    for (;;) {
        int slot = -1;
        {
            std::unique_lock<std::mutex> lk(S->mu);
            S->cv_inflight.wait(lk, [&] { return !S->inflight.empty() || S->dispatch_finished; });
            if (S->inflight.empty()) break;
            slot = S->inflight.front();
            S->inflight.pop_front();
        }
        LaunchSlot& L = S->slots[slot];
        rinnd_status st = RINND_OK;
        std::string msg;
        float ms = 0.f;
        int32_t kst = 0;
        if (hipEventSynchronize(L.ev1) != hipSuccess || hipMemcpy(&kst, L.status.p, 4, hipMemcpyDeviceToHost) != hipSuccess) {
            st = RINND_ERR_HIP;
            msg = std::string("search launch failed: ") + hipGetErrorString(hipGetLastError());
        } else {
            (void)hipEventElapsedTime(&ms, L.ev0, L.ev1);
            if (kst == 1) {
                st = RINND_ERR_UNSUPPORTED;
                msg = "a query's frontier exceeded the server's limit of 65,536 entries";
            } else if (kst != 0) {
                st = RINND_ERR_INTERNAL;
                msg = "search workspace capacity exceeded";
            }
        }
        int64_t nq = 0;
        for (auto& r : L.reqs) {
            nq += r->nq;
            if (st == RINND_OK && !r->dev_out) {
                const size_t cnt = (size_t)r->nq * r->k;
                if (hipMemcpy(r->out_i, L.oi.p + r->obase, cnt * 4, hipMemcpyDeviceToHost) != hipSuccess ||
                    hipMemcpy(r->out_d, L.od.p + r->obase, cnt * 4, hipMemcpyDeviceToHost) != hipSuccess) {
                    r->status = RINND_ERR_HIP;
                    r->error = "copying the results to host memory failed";
                }
            }
        }
        std::lock_guard<std::mutex> lk(S->mu);
        for (auto& r : L.reqs) {
            if (r->status == RINND_OK && st != RINND_OK) {
                r->status = st;
                r->error = msg;
            }
            r->done = true;
        }
        S->stats.launches += 1;
        S->stats.queries += nq;
        S->stats.max_launch_queries = std::max(S->stats.max_launch_queries, nq);
        S->stats.gpu_ms += ms;
        L.reqs.clear();
        L.busy = false;
        S->cv_done.notify_all();
        S->cv_work.notify_all();
    }
}

rinnd_status finish(const rinnd_server* S, const std::shared_ptr<Request>& r) {
    if (r->status != RINND_OK) S->last_error = r->error;
    return r->status;
}

}  // namespace

extern "C" {

rinnd_status rinnd_server_create(rinnd_index* ix, const rinnd_server_params* params, rinnd_server** out) {
    if (!ix || !out) return RINND_ERR_INVALID_ARG;
    *out = nullptr;
    auto S = std::make_unique<rinnd_server>();
    S->ix = ix;
    if (params) S->p = *params;
    rinnd_server_params& p = S->p;
    if (p.max_batch <= 0) p.max_batch = 4096;
    if (p.max_inflight <= 0) p.max_inflight = 3;
    if (p.max_k <= 0) p.max_k = 128;
    if (p.visited_bytes <= 0) p.visited_bytes = (int64_t)8 << 30;
    S->device = ix->ctx->device;
    try {
        if (p.max_batch > (1 << 20) || p.max_inflight > 64 || p.max_k > 512)
            throw Error{RINND_ERR_INVALID_ARG, "server limits: max_batch <= 2^20, max_inflight <= 64, max_k <= 512"};
        if (p.candidate_width < 0 || p.rerank_width < 0 || p.min_width < 0)
            throw Error{RINND_ERR_INVALID_ARG, "negative width"};
        if (p.min_width > p.max_k) throw Error{RINND_ERR_INVALID_ARG, "min_width exceeds max_k"};
        if (p.quantized && p.min_width > 0)
            throw Error{RINND_ERR_INVALID_ARG, "min_width applies to FP32 servers (quantized servers take widths)"};
        if (p.quantized < 0 || p.quantized > 2)
            throw Error{RINND_ERR_INVALID_ARG, "quantized must be 0 (FP32), 1 (FP8) or 2 (int8 codes)"};
        if (p.quantized == 2 && !ix->i8c.p)
            throw Error{RINND_ERR_INVALID_ARG, "the index has no int8 codes (rinnd_index_encode_i8)"};
        if (p.quantized != 2 && ix->data_released)
            throw Error{RINND_ERR_INVALID_ARG, "FP32 vector data was released: only the int8 search is available"};
        HIP_TRY(hipSetDevice(S->device));
        if (p.quantized == 1) {
            std::lock_guard<std::mutex> lk(ix->ctx->mu);  // the codes are built on the index's stream
            ensure_fp8_codes(*ix);
            HIP_TRY(hipStreamSynchronize(ix->ctx->stream));
        }
        HIP_TRY(hipStreamCreateWithFlags(&S->stream, hipStreamNonBlocking));
        S->seed_cap = std::max<int64_t>(1, std::min<int64_t>(ix->n, kServeSeedCap));
        const int64_t per_slot = ix->n + 8 * S->seed_cap + 4;
        // each launch slot owns an equal share of the visited sets (concurrent launches never share one)
        const int64_t total_sets = std::max<int64_t>(p.max_inflight, std::min<int64_t>(16384, p.visited_bytes / per_slot));
        S->visit_slots = (int)(total_sets / p.max_inflight);
        const int64_t sets = (int64_t)S->visit_slots * p.max_inflight;
        S->visited.alloc((size_t)sets * std::max<int64_t>(ix->n, 1));
        S->epochs.alloc(sets);
        S->spill_d.alloc((size_t)sets * S->seed_cap);
        S->spill_i.alloc((size_t)sets * S->seed_cap);
        S->target_batch = p.target_batch > 0 ? std::min(p.target_batch, p.max_batch) : std::min<int64_t>(1024, p.max_batch);
        HIP_TRY(hipMemsetAsync(S->visited.p, 0, S->visited.bytes(), S->stream));
        HIP_TRY(hipMemsetAsync(S->epochs.p, 0, S->epochs.bytes(), S->stream));
        const ParamLayout P(p.max_batch);
        S->slots.resize(p.max_inflight);
        for (size_t li = 0; li < S->slots.size(); ++li) {
            LaunchSlot& L = S->slots[li];
            L.vis0 = (int64_t)li * S->visit_slots;
            HIP_TRY(hipStreamCreateWithFlags(&L.stream, hipStreamNonBlocking));
            L.params.alloc(P.bytes);
            L.Q.alloc((size_t)p.max_batch * ix->ld);
            L.oi.alloc((size_t)p.max_batch * p.max_k);
            L.od.alloc((size_t)p.max_batch * p.max_k);
            L.status.alloc(1);
            HIP_TRY(hipHostMalloc(&L.pinned, P.bytes, hipHostMallocDefault));
            HIP_TRY(hipEventCreate(&L.ev0));
            HIP_TRY(hipEventCreate(&L.ev1));
        }
        HIP_TRY(hipStreamSynchronize(S->stream));
    } catch (const Error& e) {
        ix->ctx->last_error = e.msg;
        (void)hipGetLastError();
        for (auto& L : S->slots) {
            if (L.pinned) (void)hipHostFree(L.pinned);
            if (L.ev0) (void)hipEventDestroy(L.ev0);
            if (L.ev1) (void)hipEventDestroy(L.ev1);
            if (L.stream) (void)hipStreamDestroy(L.stream);
        }
        if (S->stream) (void)hipStreamDestroy(S->stream);
        return e.code;
    }
    rinnd_server* raw = S.release();
    raw->dispatcher = std::thread(dispatch_loop, raw);
    raw->completer = std::thread(complete_loop, raw);
    *out = raw;
    return RINND_OK;
}

static rinnd_status submit(rinnd_server* S, const float* queries, int64_t nq, int64_t k, float epsilon, int64_t cw,
                           int64_t rw, int32_t* out_idx, float* out_dist, uint64_t* ticket) {
    if (!S || !ticket) return RINND_ERR_INVALID_ARG;
    auto fail = [&](rinnd_status st, const std::string& m) {
        std::lock_guard<std::mutex> lk(S->mu);
        S->last_error = m;
        return st;
    };
    if (nq < 0 || nq > S->p.max_batch) return fail(RINND_ERR_INVALID_ARG, "nq must be in [0, max_batch]");
    if (k < 1) return fail(RINND_ERR_INVALID_ARG, "k must be positive");
    if (nq > 0 && (!queries || !out_idx || !out_dist)) return fail(RINND_ERR_INVALID_ARG, "null buffer");
    auto r = std::make_shared<Request>();
    r->q = queries;
    r->nq = nq;
    r->k = k;
    r->eps = epsilon;
    r->out_i = out_idx;
    r->out_d = out_dist;
    r->width = std::max<int64_t>(k, S->p.min_width);  // search-width floor of an FP32 server
    if (S->p.quantized) {  // resolve_quantized_candidate_width / should_rerank_quantized (core/index.rs:55-90)
        if (cw < 0 || rw < 0) return fail(RINND_ERR_INVALID_ARG, "negative width");
        r->width = cw == 0 ? std::max(k, rw) : cw;
        if (r->width < k) return fail(RINND_ERR_INVALID_ARG, "quantized candidate width must be at least k");
        if (rw > r->width) return fail(RINND_ERR_INVALID_ARG, "quantized rerank width must not exceed candidate width");
        if (cw != 0 && rw != 0 && rw < k)
            return fail(RINND_ERR_INVALID_ARG, "quantized rerank width must be zero or at least k");
        r->rerank = cw == 0 ? (rw > k ? rw : 0) : rw;
        if (r->rerank > 0 && S->ix->data_released)
            return fail(RINND_ERR_INVALID_ARG, "exact reranking is unavailable because FP32 vector data was released");
    }
    if (r->width > S->p.max_k)
        return fail(RINND_ERR_UNSUPPORTED, "the request's search width exceeds the server's beam memory (max_k)");
    const bool di = device_pointer(out_idx), dd = device_pointer(out_dist);
    if (nq > 0 && di != dd) return fail(RINND_ERR_INVALID_ARG, "out_idx and out_dist must both be host or both device");
    r->dev_out = di;
    std::lock_guard<std::mutex> lk(S->mu);
    if (S->stop) {
        S->last_error = "the server is stopping";
        return RINND_ERR_INVALID_ARG;
    }
    *ticket = S->next_ticket++;
    S->live[*ticket] = r;
    S->stats.requests += 1;
    if (nq == 0) {
        r->done = true;
    } else {
        S->pending.push_back(r);
        S->cv_work.notify_one();
    }
    return RINND_OK;
}

rinnd_status rinnd_server_submit(rinnd_server* S, const float* queries, int64_t nq, int64_t k, float epsilon,
                                 int32_t* out_idx, float* out_dist, uint64_t* ticket) {
    if (!S) return RINND_ERR_INVALID_ARG;
    return submit(S, queries, nq, k, epsilon, S->p.candidate_width, S->p.rerank_width, out_idx, out_dist, ticket);
}

rinnd_status rinnd_server_submit_quantized(rinnd_server* S, const float* queries, int64_t nq, int64_t k, float epsilon,
                                           int64_t candidate_width, int64_t rerank_width, int32_t* out_idx,
                                           float* out_dist, uint64_t* ticket) {
    if (!S) return RINND_ERR_INVALID_ARG;
    if (!S->p.quantized) {
        std::lock_guard<std::mutex> lk(S->mu);
        S->last_error = "quantized widths require a quantized server";
        return RINND_ERR_INVALID_ARG;
    }
    return submit(S, queries, nq, k, epsilon, candidate_width, rerank_width, out_idx, out_dist, ticket);
}

rinnd_status rinnd_server_poll(rinnd_server* S, uint64_t ticket, int32_t* done) {
    if (!S || !done) return RINND_ERR_INVALID_ARG;
    std::lock_guard<std::mutex> lk(S->mu);
    auto it = S->live.find(ticket);
    if (it == S->live.end()) {
        S->last_error = "unknown ticket";
        return RINND_ERR_INVALID_ARG;
    }
    *done = it->second->done ? 1 : 0;
    if (!it->second->done) return RINND_OK;
    auto r = it->second;
    S->live.erase(it);
    return finish(S, r);
}

rinnd_status rinnd_server_wait(rinnd_server* S, uint64_t ticket) {
    if (!S) return RINND_ERR_INVALID_ARG;
    std::unique_lock<std::mutex> lk(S->mu);
    auto it = S->live.find(ticket);
    if (it == S->live.end()) {
        S->last_error = "unknown ticket";
        return RINND_ERR_INVALID_ARG;
    }
    auto r = it->second;
    S->cv_done.wait(lk, [&] { return r->done; });
    S->live.erase(ticket);
    return finish(S, r);
}

rinnd_status rinnd_server_get_stats(const rinnd_server* S, rinnd_server_stats* out) {
    if (!S || !out) return RINND_ERR_INVALID_ARG;
    std::lock_guard<std::mutex> lk(S->mu);
    *out = S->stats;
    return RINND_OK;
}

const char* rinnd_server_last_error(const rinnd_server* S) { return S ? S->last_error.c_str() : "null server"; }

void rinnd_server_destroy(rinnd_server* S) {
    if (!S) return;
    {
        std::lock_guard<std::mutex> lk(S->mu);
        S->stop = true;
    }
    S->cv_work.notify_all();
    if (S->dispatcher.joinable()) S->dispatcher.join();
    if (S->completer.joinable()) S->completer.join();
    (void)hipSetDevice(S->device);
    (void)hipStreamSynchronize(S->stream);
    for (auto& L : S->slots) {
        if (L.stream) {
            (void)hipStreamSynchronize(L.stream);
            (void)hipStreamDestroy(L.stream);
        }
        if (L.pinned) (void)hipHostFree(L.pinned);
        if (L.ev0) (void)hipEventDestroy(L.ev0);
        if (L.ev1) (void)hipEventDestroy(L.ev1);
    }
    S->slots.clear();
    S->visited.release();
    S->epochs.release();
    S->spill_d.release();
    S->spill_i.release();
    (void)hipStreamDestroy(S->stream);
    delete S;
}

}  // extern "C"
