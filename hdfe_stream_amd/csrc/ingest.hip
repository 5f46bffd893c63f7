// Host -> HBM ingest of row batches for gfx950 hosts: the panel's columns,
// held in ordinary (pageable) host memory, are streamed to device arrays in
// batches of rows through a ring of page-locked staging slots.
//
// The reference reads its Parquet input in row batches (core/utils.py:26-60,
// iter_group_chunks; core/passes.py:397-531) and works on each batch on the
// CPU.  Here every batch goes to HBM, where the whole fit runs: the work per
// batch is a host pack (the batch's rows of every column copied into one
// page-locked slot by a pool of host threads) and a DMA (one asynchronous
// copy per column, on a copy stream).  With two or more slots the pack of
// batch b + 1 overlaps the DMA of batch b, so the rate is the slower of the
// two: PCIe on the boxes measured, not the host's memory copy.  A slot is
// packed again only once the event recorded after its previous DMA has
// completed.
#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <functional>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "devbuf.h"

namespace hdfe {

namespace {

// A fixed set of host threads that run one job at a time: run(f) calls
// f(t, T) on threads t = 0 .. T-1 (the caller is thread 0) and returns when
// all have finished.
class HostPool {
   public:
    explicit HostPool(int nthreads) : n_(std::max(1, nthreads)) {
        for (int t = 1; t < n_; ++t) th_.emplace_back([this, t] { loop(t); });
    }
    ~HostPool() {
        {
            std::lock_guard<std::mutex> lk(mu_);
            stop_ = true;
            ++gen_;
        }
        cv_.notify_all();
        for (auto &t : th_) t.join();
    }
    int size() const { return n_; }
    void run(const std::function<void(int, int)> &f) {
        {
            std::lock_guard<std::mutex> lk(mu_);
            job_ = &f;
            pending_ = n_ - 1;
            ++gen_;
        }
        cv_.notify_all();
        f(0, n_);
        std::unique_lock<std::mutex> lk(mu_);
        done_.wait(lk, [this] { return pending_ == 0; });
        job_ = nullptr;
    }

   private:
    void loop(int t) {
        uint64_t seen = 0;
        for (;;) {
            const std::function<void(int, int)> *job;
            {
                std::unique_lock<std::mutex> lk(mu_);
                cv_.wait(lk, [&] { return gen_ != seen; });
                seen = gen_;
                if (stop_) return;
                job = job_;
            }
            (*job)(t, n_);
            {
                std::lock_guard<std::mutex> lk(mu_);
                if (--pending_ == 0) done_.notify_one();
            }
        }
    }
    int n_;
    std::vector<std::thread> th_;
    std::mutex mu_;
    std::condition_variable cv_, done_;
    const std::function<void(int, int)> *job_ = nullptr;
    uint64_t gen_ = 0;
    int pending_ = 0;
    bool stop_ = false;
};

struct Stager {
    int ncols = 0;
    std::vector<int> eb;          // bytes per element, per column
    std::vector<int64_t> coff;    // a column's byte offset inside a slot
    int64_t batch_rows = 0, slot_bytes = 0;
    int depth = 0, next = 0;      // next: the slot the next batch takes
    std::vector<char *> slot;     // page-locked
    std::vector<hipEvent_t> ev;   // the slot's last DMA
    std::vector<bool> used;
    hipStream_t stream = nullptr;
    HostPool *pool = nullptr;
    int device = 0;

    ~Stager() {
        if (stream) (void)hipStreamSynchronize(stream);
        for (auto e : ev)
            if (e) (void)hipEventDestroy(e);
        for (auto p : slot)
            if (p) (void)hipHostFree(p);
        if (stream) stream_put(stream, device);
        delete pool;
    }
};

double now_s() {
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch())
        .count();
}

}  // namespace

extern "C" {

// A stager for `ncols` columns of elem_bytes[c] bytes per element: `depth`
// page-locked slots of `batch_rows` rows each (every column), `threads` host
// threads for the packs.
int hdfe_stage_create(int ncols, const int *elem_bytes, int64_t batch_rows, int depth,
                      int threads, void **out) {
    ARG_CHECK(ncols >= 1 && ncols <= 64 && elem_bytes && batch_rows >= 1 && depth >= 1 &&
                  depth <= 8 && threads >= 1 && threads <= 256 && out,
              "hdfe_stage_create: bad argument");
    for (int c = 0; c < ncols; ++c)
        ARG_CHECK(elem_bytes[c] >= 1 && elem_bytes[c] <= 16, "hdfe_stage_create: element size");
    auto *s = new Stager();
    s->ncols = ncols;
    s->eb.assign(elem_bytes, elem_bytes + ncols);
    s->batch_rows = batch_rows;
    s->depth = depth;
    int64_t off = 0;
    for (int c = 0; c < ncols; ++c) {
        s->coff.push_back(off);
        off += ((batch_rows * elem_bytes[c] + 4095) / 4096) * 4096;
    }
    s->slot_bytes = off;
    int st = HDFE_STATUS_OK;
    auto fail = [&](hipError_t e, const char *what) {
        set_error(std::string("hdfe_stage_create: ") + what + ": " + hipGetErrorString(e));
        st = e == hipErrorOutOfMemory ? HDFE_STATUS_NOMEM : HDFE_STATUS_HIP;
    };
    hipError_t e = hipGetDevice(&s->device);
    if (e != hipSuccess) fail(e, "hipGetDevice");
    for (int i = 0; i < depth && st == HDFE_STATUS_OK; ++i) {
        void *p = nullptr;
        e = hipHostMalloc(&p, (size_t)s->slot_bytes, hipHostMallocDefault);
        if (e != hipSuccess) {
            fail(e, "hipHostMalloc");
            break;
        }
        s->slot.push_back(static_cast<char *>(p));
        hipEvent_t ev = nullptr;
        e = hipEventCreateWithFlags(&ev, hipEventDisableTiming);
        if (e != hipSuccess) {
            fail(e, "hipEventCreate");
            break;
        }
        s->ev.push_back(ev);
        s->used.push_back(false);
    }
    if (st == HDFE_STATUS_OK) {
        e = stream_get(&s->stream);
        if (e != hipSuccess) fail(e, "stream");
    }
    if (st != HDFE_STATUS_OK) {
        delete s;
        return st;
    }
    s->pool = new HostPool(threads);
    *out = s;
    return HDFE_STATUS_OK;
}

// Rows [0, n_rows) of the host columns host_cols[c] into the device columns
// dev_cols[c] at row dev_row (so that successive calls can fill one array
// batch by batch), through the slots.  The host columns are free for reuse
// on return (every batch has been packed); with `wait` the call also waits
// for the copies to land, else hdfe_stage_wait does.  stats (or NULL): [0]
// seconds in all, [1] seconds the packs took, [2] bytes moved, [3] batches.
int hdfe_stage_columns(void *h, int64_t n_rows, const void *const *host_cols,
                       void *const *dev_cols, int64_t dev_row, int wait, double *stats) {
    auto *s = static_cast<Stager *>(h);
    ARG_CHECK(s && n_rows >= 0 && host_cols && dev_cols && dev_row >= 0,
              "hdfe_stage_columns: bad argument");
    for (int c = 0; c < s->ncols; ++c)
        ARG_CHECK(host_cols[c] && dev_cols[c], "hdfe_stage_columns: null column");
    const double t0 = now_s();
    double pack = 0.0;
    int64_t bytes = 0, nb = 0;
    for (int64_t r0 = 0; r0 < n_rows; r0 += s->batch_rows, ++nb) {
        const int64_t rows = std::min<int64_t>(s->batch_rows, n_rows - r0);
        const int k = s->next;
        s->next = (s->next + 1) % s->depth;
        if (s->used[k]) HIP_TRY(hipEventSynchronize(s->ev[k]));
        char *slot = s->slot[k];
        // the pack: the batch's bytes of every column split evenly over the
        // threads, each thread copying its share of one or more columns
        int64_t total = 0;
        for (int c = 0; c < s->ncols; ++c) total += rows * s->eb[c];
        const double p0 = now_s();
        s->pool->run([&](int t, int T) {
            const int64_t a = total * t / T, b = total * (t + 1) / T;
            int64_t base = 0;
            for (int c = 0; c < s->ncols && base < b; ++c) {
                const int64_t len = rows * s->eb[c];
                const int64_t lo = std::max(a, base), hi = std::min(b, base + len);
                if (lo < hi) {
                    const char *src = static_cast<const char *>(host_cols[c]) + r0 * s->eb[c];
                    std::memcpy(slot + s->coff[c] + (lo - base), src + (lo - base),
                                (size_t)(hi - lo));
                }
                base += len;
            }
        });
        pack += now_s() - p0;
        for (int c = 0; c < s->ncols; ++c) {
            char *dst = static_cast<char *>(dev_cols[c]) + (dev_row + r0) * s->eb[c];
            HIP_TRY(hipMemcpyAsync(dst, slot + s->coff[c], (size_t)(rows * s->eb[c]),
                                   hipMemcpyHostToDevice, s->stream));
        }
        HIP_TRY(hipEventRecord(s->ev[k], s->stream));
        s->used[k] = true;
        bytes += total;
    }
    if (wait) HIP_TRY(hipStreamSynchronize(s->stream));
    if (stats) {
        stats[0] = now_s() - t0;
        stats[1] = pack;
        stats[2] = (double)bytes;
        stats[3] = (double)nb;
    }
    return HDFE_STATUS_OK;
}

// Waits for every copy the stager has enqueued.
int hdfe_stage_wait(void *h) {
    auto *s = static_cast<Stager *>(h);
    ARG_CHECK(s, "hdfe_stage_wait: bad argument");
    HIP_TRY(hipStreamSynchronize(s->stream));
    return HDFE_STATUS_OK;
}

void hdfe_stage_destroy(void *h) { delete static_cast<Stager *>(h); }

}  // extern "C"

}  // namespace hdfe
