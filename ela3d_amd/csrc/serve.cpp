// serve.cpp -- continuous batching of forward shots on one GPU.
//
// One device thread owns every admitted shot.  Each shot runs on one of
// max_active stream slots and keeps up to two rounds of steps_per_round graph
// replays in flight; after each round it records an event.  The thread polls
// those events: a shot whose last round completed is retired (results copied to
// host, its device memory returned to the pool) and the admission step then
// starts queued shots in arrival order while the device-memory budget and the
// active-shot cap allow.  Shots of any grid size and step count therefore
// overlap on the device, and arrivals join while earlier shots are still running.
//
// Nothing per shot is allocated or freed through the HIP runtime in steady
// state: shots are placed in a server-owned device pool (a few large hipMalloc
// chunks, sub-allocated first-fit), and the streams, events and pinned flag
// words are made once per slot.  hipFree waits for the whole device, so a
// retirement that freed memory would drain every other shot's queued rounds.
#include <algorithm>
#include <atomic>
#include <iterator>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <mutex>
#include <thread>

#include "shot.hpp"

namespace ela3d {

// Server-owned device memory: chunks of at least kChunk bytes from hipMalloc,
// sub-allocated first-fit in 256-byte units, freed ranges coalesced.  Chunks
// are released only when the server is destroyed.
class DevicePool {
public:
    static constexpr size_t kChunk = size_t(16) << 30;
    explicit DevicePool(size_t limit) : limit_(limit) {}
    ~DevicePool()
    {
        for (auto& c : chunks_) device_free(c.base, c.size);
    }
    // a 256-byte aligned range of `bytes`, or throws ELA3D_E_OOM
    void* alloc(size_t bytes)
    {
        bytes = (bytes + 255) & ~size_t(255);
        for (size_t q = 0; q < chunks_.size(); ++q)
            if (void* p = take(q, bytes)) return p;
        size_t want = std::max(bytes, std::min(kChunk, limit_));
        void* base = nullptr;
        try {
            base = device_alloc(want);
        } catch (const Error&) {
            want = bytes;  // no room for a whole chunk: just this shot
            base = device_alloc(want);
        }
        chunks_.push_back(Chunk{static_cast<char*>(base), want, {}});
        chunks_.back().free[0] = want;
        reserved_ += want;
        return take(chunks_.size() - 1, bytes);
    }
    void release(void* p)
    {
        auto it = live_.find(p);
        if (it == live_.end()) return;
        Chunk& c = chunks_[it->second.first];
        size_t off = size_t(static_cast<char*>(p) - c.base), len = it->second.second;
        live_.erase(it);
        auto nx = c.free.lower_bound(off);
        if (nx != c.free.end() && off + len == nx->first) {  // merge with the range after
            len += nx->second;
            nx = c.free.erase(nx);
        }
        if (nx != c.free.begin()) {  // and with the range before
            auto pv = std::prev(nx);
            if (pv->first + pv->second == off) {
                pv->second += len;
                return;
            }
        }
        c.free[off] = len;
    }
    size_t reserved() const { return reserved_; }

private:
    struct Chunk {
        char* base;
        size_t size;
        std::map<size_t, size_t> free;  // offset -> length
    };
    void* take(size_t q, size_t bytes)
    {
        Chunk& c = chunks_[q];
        for (auto it = c.free.begin(); it != c.free.end(); ++it) {
            if (it->second < bytes) continue;
            const size_t off = it->first, len = it->second;
            c.free.erase(it);
            if (len > bytes) c.free[off + bytes] = len - bytes;
            void* p = c.base + off;
            live_[p] = {int(q), bytes};
            return p;
        }
        return nullptr;
    }
    size_t limit_;
    size_t reserved_ = 0;
    std::vector<Chunk> chunks_;
    std::map<void*, std::pair<int, size_t>> live_;
};

// One stream slot: its stream, the two round events and a pinned flag word.
struct Slot {
    hipStream_t stream = nullptr;
    hipEvent_t ev[2] = {nullptr, nullptr};
    int* flag = nullptr;
    bool busy = false;
};

struct Request {
    uint64_t ticket = 0;
    ela3d_config cfg{};
    std::vector<int32_t> rec;
    std::vector<float> cp, cs, rho;
    int64_t bytes = 0;           // device bytes charged to the budget
    std::atomic<int> state{ELA3D_SHOT_QUEUED};
    std::atomic<int> status{ELA3D_OK};
    double t_submit = 0, t_admit = 0, t_done = 0;
    int nsamples = 0, nenergy = 0;
    std::vector<double> seis, energy;
    // device-thread only
    std::unique_ptr<Shot> shot;
    void* mem = nullptr;         // its range of the device pool
    int slot = -1;
    bool batched = false;        // steps in the server's batch launches
    int inflight = 0;            // batch rounds in flight that include it
    hipEvent_t ev[2] = {nullptr, nullptr};
    int ev_head = 0, ev_count = 0;
    int enqueued = 0;
};

class Server {
public:
    explicit Server(const ela3d_server_opts& o) : opts_(o), pool_(size_t(o.hbm_budget_bytes))
    {
        if (opts_.plain_loop) opts_.max_active = 1;
        HIP_CHECK(hipSetDevice(opts_.device));
        slots_.resize(size_t(opts_.max_active));
        try {
            for (Slot& sl : slots_) {  // default priority, like every stream of the library
                HIP_CHECK(hipStreamCreateWithFlags(&sl.stream, hipStreamNonBlocking));
                for (auto& e : sl.ev) HIP_CHECK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
                HIP_CHECK(hipHostMalloc(reinterpret_cast<void**>(&sl.flag), sizeof(int), hipHostMallocDefault));
            }
            if (batching()) {
                Slot& b = batch_.slot;
                HIP_CHECK(hipStreamCreateWithFlags(&b.stream, hipStreamNonBlocking));
                for (auto& e : b.ev) HIP_CHECK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
                HIP_CHECK(hipHostMalloc(reinterpret_cast<void**>(&batch_.staging), 3 * kArgsBytes,
                                        hipHostMallocDefault));
                batch_.dargs = pool_.alloc(kArgsBytes);
            }
        } catch (...) {
            free_slots();
            throw;
        }
        t0_ = std::chrono::steady_clock::now();
        th_ = std::thread([this] { loop(); });
    }

    ~Server()
    {
        {
            std::lock_guard<std::mutex> g(mu_);
            stop_ = true;
        }
        cv_.notify_all();
        if (th_.joinable()) th_.join();
        free_slots();
    }

    bool batching() const { return !opts_.plain_loop && opts_.batch_cells > 0 && opts_.max_active > 1; }

    void free_slots()
    {
        (void)hipSetDevice(opts_.device);
        if (batch_.slot.stream) (void)hipStreamSynchronize(batch_.slot.stream);
        for (auto& e : batch_.slot.ev)
            if (e) (void)hipEventDestroy(e);
        if (batch_.slot.stream) (void)hipStreamDestroy(batch_.slot.stream);
        if (batch_.staging) (void)hipHostFree(batch_.staging);
        batch_.slot = Slot();
        batch_.staging = nullptr;
        for (Slot& sl : slots_) {
            if (sl.stream) (void)hipStreamSynchronize(sl.stream);
            for (auto& e : sl.ev)
                if (e) (void)hipEventDestroy(e);
            if (sl.stream) (void)hipStreamDestroy(sl.stream);
            if (sl.flag) (void)hipHostFree(sl.flag);
            sl = Slot();
        }
    }

    double now() const
    {
        return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0_).count();
    }

    uint64_t submit(const ela3d_config& cfg, const float* cp, const float* cs, const float* rho)
    {
        ShotPlan plan = make_plan(cfg);  // validates; throws Error
        check_model(plan.cells, cp, cs, rho);
        auto r = std::unique_ptr<Request>(new Request);
        r->cfg = plan.cfg;
        r->rec = plan.rec;
        const size_t n = size_t(plan.cells);
        r->cp.assign(cp, cp + n);
        r->cs.assign(cs, cs + n);
        r->rho.assign(rho, rho + n);
        r->bytes = device_bytes_estimate(plan.cfg, &plan);
        r->nsamples = plan.nsamples;
        r->nenergy = plan.nenergy;
        if (r->bytes > opts_.hbm_budget_bytes)
            throw Error{ELA3D_E_BUDGET, "shot needs " + std::to_string(r->bytes) +
                                            " device bytes, more than the server budget"};
        std::lock_guard<std::mutex> g(mu_);
        r->ticket = ++next_ticket_;
        r->t_submit = now();
        const uint64_t t = r->ticket;
        queue_.push_back(r.get());
        all_[t] = std::move(r);
        ++submitted_;
        cv_.notify_all();
        return t;
    }

    Request* find(uint64_t t)
    {
        auto it = all_.find(t);
        if (it == all_.end()) throw Error{ELA3D_E_NOTFOUND, "unknown ticket"};
        return it->second.get();
    }

    int poll(uint64_t t)
    {
        std::lock_guard<std::mutex> g(mu_);
        Request* r = find(t);
        const int st = r->status.load();
        return st != ELA3D_OK ? st : r->state.load();
    }

    int wait(uint64_t t, double timeout_s, ela3d_result& res)
    {
        std::unique_lock<std::mutex> g(mu_);
        Request* r = find(t);
        auto done = [&] { return r->state == ELA3D_SHOT_DONE; };
        if (timeout_s < 0)
            cv_done_.wait(g, done);
        else if (!cv_done_.wait_for(g, std::chrono::duration<double>(timeout_s), done))
            return ELA3D_E_TIMEOUT;
        std::memset(&res, 0, sizeof(res));
        res.ticket = t;
        res.status = r->status;
        res.nsamples = r->nsamples;
        res.nrec = r->cfg.nrec;
        res.nenergy = r->nenergy;
        res.t_submit = r->t_submit;
        res.t_admit = r->t_admit;
        res.t_done = r->t_done;
        res.device_bytes = r->bytes;
        return ELA3D_OK;
    }

    void results(uint64_t t, double* out, size_t cap, bool seis)
    {
        std::lock_guard<std::mutex> g(mu_);
        Request* r = find(t);
        if (r->state != ELA3D_SHOT_DONE) throw Error{ELA3D_E_STATE, "shot not finished"};
        const std::vector<double>& v = seis ? r->seis : r->energy;
        if (cap < v.size() || (!v.empty() && !out)) throw Error{ELA3D_E_RANGE, "buffer too small"};
        if (!v.empty()) std::memcpy(out, v.data(), v.size() * sizeof(double));
    }

    void release(uint64_t t)
    {
        std::lock_guard<std::mutex> g(mu_);
        Request* r = find(t);
        if (r->state != ELA3D_SHOT_DONE) throw Error{ELA3D_E_STATE, "shot not finished"};
        all_.erase(t);
    }

    void stats(int64_t* out)
    {
        std::lock_guard<std::mutex> g(mu_);
        out[0] = submitted_;
        out[1] = done_;
        out[2] = failed_;
        out[3] = int64_t(active_.size());
        out[4] = int64_t(queue_.size());
        out[5] = in_use_;
        out[6] = peak_;
        out[7] = rounds_.load();
        out[8] = batched_.load();
        out[9] = batch_rounds_.load();
    }

private:
    // ---- device thread ----
    void loop()
    {
        (void)hipSetDevice(opts_.device);
        for (;;) {
            std::vector<Request*> admitted;
            {
                std::unique_lock<std::mutex> g(mu_);
                if (stop_) {  // shutting down: cancel queued shots, stop running ones
                    for (Request* r : queue_) {
                        r->status = ELA3D_E_STATE;
                        r->state = ELA3D_SHOT_DONE;
                        r->t_done = now();
                        ++done_;
                        ++failed_;
                    }
                    queue_.clear();
                    cv_done_.notify_all();
                    for (Request* r : active_)
                        if (r->enqueued < r->cfg.nstep) {
                            r->enqueued = r->cfg.nstep;
                            if (r->status == ELA3D_OK) r->status = ELA3D_E_STATE;
                        }
                    if (active_.empty()) break;
                }
                if (active_.empty() && queue_.empty()) {
                    cv_.wait_for(g, std::chrono::milliseconds(20));
                    continue;
                }
                // admission in arrival order while the budget and the cap allow
                while (!queue_.empty() && int(active_.size() + admitted.size()) < opts_.max_active &&
                       in_use_ + queue_.front()->bytes <= opts_.hbm_budget_bytes) {
                    Request* r = queue_.front();
                    queue_.pop_front();
                    r->state = ELA3D_SHOT_RUNNING;
                    r->t_admit = now();
                    in_use_ += r->bytes;
                    peak_ = std::max(peak_, in_use_);
                    admitted.push_back(r);
                }
            }
            for (Request* r : admitted) start(r);
            bool progressed = false;
            for (size_t q = 0; q < active_.size(); ++q)
                if (!active_[q]->batched) progressed |= advance(active_[q]);
            if (batching()) progressed |= advance_batch();
            // retire finished shots
            for (size_t q = 0; q < active_.size();) {
                Request* r = active_[q];
                const bool over = r->batched ? (r->enqueued >= r->cfg.nstep && r->inflight == 0)
                                             : (r->enqueued >= r->cfg.nstep && r->ev_count == 0);
                if (r->shot == nullptr || over) {
                    finish(r);
                    std::lock_guard<std::mutex> g(mu_);
                    active_.erase(active_.begin() + q);
                    progressed = true;
                } else {
                    ++q;
                }
            }
            if (progressed)
                ++rounds_;
            else
                std::this_thread::sleep_for(std::chrono::microseconds(50));
        }
    }

    void start(Request* r)
    {
        try {
            ShotPlan plan = make_plan([&] {
                ela3d_config c = r->cfg;
                c.rec = r->rec.data();
                c.use_graph = opts_.use_graph;
                c.device = opts_.device;
                return c;
            }());
            const size_t need = size_t(shot_device_bytes(plan));
            for (size_t q = 0; q < slots_.size(); ++q)
                if (!slots_[q].busy) {
                    r->slot = int(q);
                    break;
                }
            if (r->slot < 0) throw Error{ELA3D_E_STATE, "no free stream slot"};
            Slot& sl = slots_[size_t(r->slot)];
            sl.busy = true;
            r->mem = pool_.alloc(need);
            ShotMem mem;
            mem.arena = r->mem;
            mem.bytes = need;
            mem.host_flag = sl.flag;
            r->shot = make_shot(std::move(plan), r->cp.data(), r->cs.data(), r->rho.data(), sl.stream, false, mem);
            const int64_t actual = r->shot->device_bytes();
            {
                std::lock_guard<std::mutex> g(mu_);
                in_use_ += actual - r->bytes;
                r->bytes = actual;
                peak_ = std::max(peak_, in_use_);
            }
            r->ev[0] = sl.ev[0];
            r->ev[1] = sl.ev[1];
            if (batching() && !opts_.plain_loop && r->shot->plan().cells <= int64_t(opts_.batch_cells) &&
                int(batch_.members.size()) < kMaxBatch) {
                BatchMember probe;
                if (r->shot->batch_member(kBatchBlock, probe)) {
                    r->batched = true;
                    batch_.members.push_back(r);
                    ++batched_;
                }
            }
            std::vector<float>().swap(r->cp);  // the model now lives on the device
            std::vector<float>().swap(r->cs);
            std::vector<float>().swap(r->rho);
        } catch (const Error& e) {
            r->status = e.code;
            r->shot.reset();
        } catch (...) {
            r->status = ELA3D_E_INVALID;
            r->shot.reset();
        }
        std::lock_guard<std::mutex> g(mu_);
        active_.push_back(r);
    }

    // keep up to two rounds in flight; pop completed rounds.  true if anything moved
    bool advance(Request* r)
    {
        if (!r->shot) return false;
        bool moved = false;
        try {
            while (r->ev_count > 0) {
                hipEvent_t e = r->ev[r->ev_head];
                const hipError_t q = hipEventQuery(e);
                if (q == hipErrorNotReady) break;
                HIP_CHECK(q);
                r->ev_head ^= 1;
                --r->ev_count;
                moved = true;
                if (r->shot->unstable_seen()) {  // blown-up shot: stop it early
                    r->status = ELA3D_E_UNSTABLE;
                    r->enqueued = r->cfg.nstep;
                    r->ev_count = 0;
                    return true;
                }
            }
            if (opts_.plain_loop) {
                if (r->enqueued < r->cfg.nstep) {
                    r->shot->step_plain(r->cfg.nstep - r->enqueued);
                    r->enqueued = r->cfg.nstep;
                    moved = true;
                }
                return moved;
            }
            while (r->ev_count < 2 && r->enqueued < r->cfg.nstep) {
                const int n = std::min(opts_.steps_per_round, r->cfg.nstep - r->enqueued);
                r->shot->enqueue_steps(n);
                r->shot->enqueue_flag_readback();
                r->enqueued += n;
                const int slot = (r->ev_head + r->ev_count) & 1;
                HIP_CHECK(hipEventRecord(r->ev[slot], r->shot->stream()));
                ++r->ev_count;
                moved = true;
            }
        } catch (const Error& e) {
            r->status = e.code;
            r->enqueued = r->cfg.nstep;
            r->ev_count = 0;
        }
        return moved;
    }

    // Batch rounds: pop completed ones (crediting their members), then keep up to
    // two in flight.  A round steps every member that still has steps by the
    // smallest number left among them (at most steps_per_round), so no member
    // passes its nstep; members that finish leave at the next round.
    bool advance_batch()
    {
        Batch& B = batch_;
        bool moved = false;
        try {
            while (B.ev_count > 0) {
                const hipError_t q = hipEventQuery(B.slot.ev[B.ev_head]);
                if (q == hipErrorNotReady) break;
                HIP_CHECK(q);
                Round& R = B.rounds[B.ev_head];
                for (Request* r : R.who) {
                    --r->inflight;
                    if (r->shot && r->shot->unstable_seen() && r->status == ELA3D_OK) {
                        r->status = ELA3D_E_UNSTABLE;  // blown up: no further rounds
                        r->enqueued = r->cfg.nstep;
                    }
                }
                R.who.clear();
                B.ev_head ^= 1;
                --B.ev_count;
                moved = true;
            }
            // members still stepping
            std::vector<Request*> step;
            int n = opts_.steps_per_round;
            for (Request* r : B.members)
                if (r->shot && r->enqueued < r->cfg.nstep) {
                    step.push_back(r);
                    n = std::min(n, r->cfg.nstep - r->enqueued);
                }
            B.members = step;
            if (step.empty() || B.ev_count >= 2) return moved;
            hipStream_t st = B.slot.stream;
            std::vector<uint64_t> tickets;
            for (Request* r : step) tickets.push_back(r->ticket);
            if (tickets != B.joined) {  // new membership: upload the members' arguments
                char* h = B.staging + size_t(B.stage) * kArgsBytes;
                B.stage = (B.stage + 1) % 3;  // a slot is reused only after its copy has run
                auto* sa = reinterpret_cast<StepArgs<float>*>(h);
                auto* va = sa + kMaxBatch;
                auto* ta = reinterpret_cast<TailArgs<float>*>(va + kMaxBatch);
                auto* blk = reinterpret_cast<int*>(ta + kMaxBatch);
                int nb = 0;
                for (size_t q = 0; q < step.size(); ++q) {
                    BatchMember m;
                    if (!step[q]->shot->batch_member(kBatchBlock, m)) throw Error{ELA3D_E_STATE, "shot left the batch"};
                    sa[q] = m.stress;
                    va[q] = m.velocity;
                    ta[q] = m.tail;
                    blk[q] = nb;
                    nb += m.nblocks;
                }
                blk[step.size()] = nb;
                B.nblocks = nb;
                HIP_CHECK(hipMemcpyAsync(B.dargs, h, kArgsBytes, hipMemcpyHostToDevice, st));
                B.joined = tickets;
            }
            const auto* sa = static_cast<const StepArgs<float>*>(B.dargs);
            const auto* va = sa + kMaxBatch;
            const auto* ta = reinterpret_cast<const TailArgs<float>*>(va + kMaxBatch);
            const auto* blk = reinterpret_cast<const int*>(ta + kMaxBatch);
            const int ns = int(step.size());
            for (int it = 0; it < n; ++it) {
                launch_stress_multi<float>(sa, blk, ns, B.nblocks, kBatchBlock, st);
                launch_velocity_multi<float>(va, blk, ns, B.nblocks, kBatchBlock, st);
                launch_tail_multi<float>(ta, ns, st);
            }
            HIP_CHECK(hipGetLastError());
            for (Request* r : step) {
                r->shot->enqueue_flag_readback(st);
                r->shot->set_steps_done(r->enqueued + n);
                r->enqueued += n;
                ++r->inflight;
            }
            const int slot = (B.ev_head + B.ev_count) & 1;
            B.rounds[slot].who = step;
            B.rounds[slot].steps = n;
            HIP_CHECK(hipEventRecord(B.slot.ev[slot], st));
            ++B.ev_count;
            ++batch_rounds_;
            moved = true;
        } catch (const Error& e) {  // the batch cannot go on: fail its members
            (void)hipStreamSynchronize(B.slot.stream);
            for (Request* r : B.members) {
                if (r->status == ELA3D_OK) r->status = e.code;
                r->enqueued = r->cfg.nstep;
                r->inflight = 0;
            }
            for (auto& R : B.rounds) R.who.clear();
            B.members.clear();
            B.joined.clear();
            B.ev_count = 0;
            moved = true;
        }
        return moved;
    }

    void finish(Request* r)
    {
        if (r->shot) {
            try {
                const int rc = r->shot->sync();
                if (rc != ELA3D_OK && r->status == ELA3D_OK) r->status = rc;
                r->seis.resize(size_t(r->nsamples) * r->cfg.nrec * 3);
                r->energy.resize(size_t(r->nenergy) * 2);
                if (!r->seis.empty()) r->shot->seismograms(r->seis.data());
                if (!r->energy.empty()) r->shot->energy(r->energy.data());
            } catch (const Error& e) {
                if (r->status == ELA3D_OK) r->status = e.code;
            }
            r->shot.reset();
        }
        if (r->mem) {  // back to the pool: the shot's stream has finished with it
            pool_.release(r->mem);
            r->mem = nullptr;
        }
        if (r->slot >= 0) {
            slots_[size_t(r->slot)].busy = false;
            r->slot = -1;
        }
        std::lock_guard<std::mutex> g(mu_);
        in_use_ -= r->bytes;
        r->state = ELA3D_SHOT_DONE;
        r->t_done = now();
        ++done_;
        if (r->status != ELA3D_OK) ++failed_;
        cv_done_.notify_all();
    }

    ela3d_server_opts opts_;
    DevicePool pool_;             // device thread only (and the destructor)
    std::vector<Slot> slots_;
    // Batch of small shots stepped together: per step one stress, one velocity and
    // one tail launch over all members' boxes (args from device memory), in
    // rounds on the batch's own stream with up to two rounds in flight.
    struct Round {
        std::vector<Request*> who;
        int steps = 0;
    };
    struct Batch {
        Slot slot;
        std::vector<Request*> members;  // stepping, in argument order
        std::vector<uint64_t> joined;   // tickets of the members as of the last argument upload
        Round rounds[2];
        int ev_head = 0, ev_count = 0;
        void* dargs = nullptr;          // device: stress[kMaxBatch], velocity[kMaxBatch], tail[kMaxBatch], blk0
        char* staging = nullptr;        // pinned: three slots of the same layout
        int stage = 0;
        int nblocks = 0;
    } batch_;
    static constexpr size_t kArgsBytes =
        ((2 * kMaxBatch * sizeof(StepArgs<float>) + kMaxBatch * sizeof(TailArgs<float>) + (kMaxBatch + 1) * sizeof(int)) +
         255) & ~size_t(255);
    static constexpr int kBatchBlock = 128;
    std::chrono::steady_clock::time_point t0_;
    std::thread th_;
    std::mutex mu_;
    std::condition_variable cv_, cv_done_;
    std::deque<Request*> queue_;
    std::vector<Request*> active_;  // changed by the device thread under mu_
    std::map<uint64_t, std::unique_ptr<Request>> all_;
    uint64_t next_ticket_ = 0;
    int64_t in_use_ = 0, peak_ = 0;
    int64_t submitted_ = 0, done_ = 0, failed_ = 0;
    std::atomic<int64_t> rounds_{0};  // device-thread passes that moved any shot
    std::atomic<int64_t> batched_{0}, batch_rounds_{0};  // shots that joined the batch; batch rounds run
    bool stop_ = false;
};

}  // namespace ela3d

struct ela3d_server {
    std::unique_ptr<ela3d::Server> impl;
};

using ela3d::Error;

template <typename F>
static int sguard(F fn)
{
    try {
        return fn();
    } catch (const Error& e) {
        ela3d::set_last_error(e.msg);
        return e.code;
    } catch (const std::exception& e) {
        ela3d::set_last_error(e.what());
        return ELA3D_E_INVALID;
    }
}

extern "C" {

void ela3d_server_opts_default(ela3d_server_opts* o)
{
    if (!o) return;
    std::memset(o, 0, sizeof(*o));
    o->struct_size = sizeof(ela3d_server_opts);
    o->device = 0;
    o->hbm_budget_bytes = int64_t(200) * 1000 * 1000 * 1000;
    o->max_active = 16;
    o->steps_per_round = 50;
    o->use_graph = 1;
    o->plain_loop = 0;
    o->batch_cells = 128 * 128 * 128;
}

int ela3d_server_create(const ela3d_server_opts* o, ela3d_server** out)
{
    return sguard([&] {
        if (!o || !out || o->struct_size != sizeof(ela3d_server_opts)) throw Error{ELA3D_E_INVALID, "bad options"};
        if (o->max_active < 1 || o->steps_per_round < 1 || o->hbm_budget_bytes <= 0)
            throw Error{ELA3D_E_INVALID, "max_active, steps_per_round and the budget must be positive"};
        int n = 0;
        if (hipGetDeviceCount(&n) != hipSuccess || n <= o->device || o->device < 0)
            throw Error{ELA3D_E_NODEVICE, "no HIP device"};
        auto s = new ela3d_server;
        s->impl.reset(new ela3d::Server(*o));
        *out = s;
        return ELA3D_OK;
    });
}

int ela3d_submit(ela3d_server* s, const ela3d_config* cfg, const float* cp, const float* cs, const float* rho,
                 uint64_t* ticket)
{
    return sguard([&] {
        if (!s || !cfg || !cp || !cs || !rho || !ticket) throw Error{ELA3D_E_INVALID, "null argument"};
        *ticket = s->impl->submit(*cfg, cp, cs, rho);
        return ELA3D_OK;
    });
}

int ela3d_poll(ela3d_server* s, uint64_t ticket, int32_t* state)
{
    return sguard([&] {
        if (!s || !state) throw Error{ELA3D_E_INVALID, "null argument"};
        *state = s->impl->poll(ticket);
        return ELA3D_OK;
    });
}

int ela3d_wait(ela3d_server* s, uint64_t ticket, double timeout_s, ela3d_result* res)
{
    return sguard([&] {
        if (!s || !res) throw Error{ELA3D_E_INVALID, "null argument"};
        return s->impl->wait(ticket, timeout_s, *res);
    });
}

int ela3d_result_seismograms(ela3d_server* s, uint64_t ticket, double* out, size_t cap)
{
    return sguard([&] {
        if (!s) throw Error{ELA3D_E_INVALID, "null server"};
        s->impl->results(ticket, out, cap, true);
        return ELA3D_OK;
    });
}

int ela3d_result_energy(ela3d_server* s, uint64_t ticket, double* out, size_t cap)
{
    return sguard([&] {
        if (!s) throw Error{ELA3D_E_INVALID, "null server"};
        s->impl->results(ticket, out, cap, false);
        return ELA3D_OK;
    });
}

int ela3d_result_release(ela3d_server* s, uint64_t ticket)
{
    return sguard([&] {
        if (!s) throw Error{ELA3D_E_INVALID, "null server"};
        s->impl->release(ticket);
        return ELA3D_OK;
    });
}

int ela3d_server_stats(ela3d_server* s, int64_t* out)
{
    return sguard([&] {
        if (!s || !out) throw Error{ELA3D_E_INVALID, "null argument"};
        s->impl->stats(out);
        return ELA3D_OK;
    });
}

void ela3d_server_destroy(ela3d_server* s) { delete s; }

}  // extern "C"
