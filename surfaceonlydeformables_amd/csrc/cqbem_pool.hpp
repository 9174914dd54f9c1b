// Device memory cache of a context.  hipMalloc of the multi-GB spectra costs time on every job, and hipFree
// waits for the whole device, so concurrent jobs on other streams would hold each other up at every release.
// A job takes its buffers from the context's pool through a JobArena and gives them back when its stream has
// finished; blocks are reused by later jobs of similar sizes and released to HIP only when an allocation
// fails (then a larger idle block is taken if there is one, else the pool is emptied and the allocation retried) or
// the context is destroyed.
#pragma once
#include <hip/hip_runtime.h>

#include <cstddef>
#include <map>
#include <mutex>
#include <utility>
#include <vector>

namespace cqbem {

class DevicePool {
  public:
    DevicePool() = default;
    DevicePool(const DevicePool &) = delete;
    DevicePool &operator=(const DevicePool &) = delete;
    ~DevicePool() { trim(); }

    // A block of at least n bytes (its size in *got); nullptr if HIP cannot provide one.
    void *get(size_t n, size_t *got) {
        {
            std::lock_guard<std::mutex> lk(mu_);
            auto it = free_.lower_bound(n);
            if (it != free_.end() && it->first <= n + n / 8 + (2u << 20)) {  // little waste: reuse
                void *p = it->second;
                *got = it->first;
                cached_ -= it->first;
                free_.erase(it);
                return p;
            }
        }
        void *p = nullptr;
        if (hipMalloc(&p, n) != hipSuccess) {
            (void)hipGetLastError();
            {
                // the device is full: a larger idle block serves before the cache is emptied (trimming and
                // allocating again costs a device-wide free and the clearing of the new pages)
                std::lock_guard<std::mutex> lk(mu_);
                auto it = free_.lower_bound(n);
                if (it != free_.end()) {
                    p = it->second;
                    *got = it->first;
                    cached_ -= it->first;
                    free_.erase(it);
                    return p;
                }
            }
            trim();
            if (hipMalloc(&p, n) != hipSuccess) {
                (void)hipGetLastError();
                return nullptr;
            }
        }
        *got = n;
        return p;
    }
    // Returns an idle block (no work in flight uses it).
    void put(void *p, size_t n) {
        std::lock_guard<std::mutex> lk(mu_);
        free_.emplace(n, p);
        cached_ += n;
    }
    void trim() {
        std::lock_guard<std::mutex> lk(mu_);
        for (auto &kv : free_) (void)hipFree(kv.second);
        free_.clear();
        cached_ = 0;
    }
    size_t cached_bytes() {
        std::lock_guard<std::mutex> lk(mu_);
        return cached_;
    }

  private:
    std::mutex mu_;
    std::multimap<size_t, void *> free_;
    size_t cached_ = 0;
};

// The buffers of one job on one stream.  Destroyed after the job: synchronises the stream, then returns
// every block to the pool.
class JobArena {
  public:
    JobArena(DevicePool *pool, hipStream_t st) : pool_(pool), st_(st) {}
    // Fixed-region mode: every buffer is carved, 256-byte aligned, out of [base, base + cap) in request order
    // and nothing is allocated or freed; a request beyond the region fails (nullptr).
    JobArena(void *base, size_t cap, hipStream_t st) : pool_(nullptr), st_(st), base_((char *)base), cap_(cap) {}
    JobArena(const JobArena &) = delete;
    JobArena &operator=(const JobArena &) = delete;
    ~JobArena() {
        (void)hipStreamSynchronize(st_);
        for (auto &b : held_) pool_->put(b.first, b.second);
    }
    void *get(size_t n) {
        if (!pool_) {
            const size_t off = (used_ + 255) & ~(size_t)255;
            if (off > cap_ || n > cap_ - off) return nullptr;
            used_ = off + n;
            return base_ + off;
        }
        size_t got = 0;
        void *p = pool_->get(n, &got);
        if (p) held_.emplace_back(p, got);
        return p;
    }
    size_t used() const { return used_; }

  private:
    DevicePool *pool_;
    hipStream_t st_;
    std::vector<std::pair<void *, size_t>> held_;
    char *base_ = nullptr;
    size_t cap_ = 0, used_ = 0;
};

// Whether the job running on this thread times its phases with HIP events (off where the caller allows no
// resource creation beyond the launches themselves).
inline bool &phase_timers_on() {
    static thread_local bool on = true;
    return on;
}

// The arena of the job running on this thread (set for the duration of a job), or null.
inline JobArena *&current_arena() {
    static thread_local JobArena *a = nullptr;
    return a;
}

// Device allocation for job-scoped buffers: from the current arena when a job is running on this thread
// (released with the arena), otherwise hipMalloc (released with hipFree).  *pooled tells the owner which.
inline hipError_t dev_alloc(void **p, size_t n, bool *pooled) {
    if (JobArena *a = current_arena()) {
        *p = a->get(n);
        *pooled = true;
        return *p ? hipSuccess : hipErrorOutOfMemory;
    }
    *pooled = false;
    return hipMalloc(p, n);
}
inline void dev_free(void *p, bool pooled) {
    if (p && !pooled) (void)hipFree(p);
}

// Sets the current arena for a scope.
struct ArenaScope {
    JobArena *prev;
    explicit ArenaScope(JobArena *a) : prev(current_arena()) { current_arena() = a; }
    ~ArenaScope() { current_arena() = prev; }
};

}  // namespace cqbem
