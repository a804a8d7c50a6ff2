// CSR-Adaptive row-block metadata.
//
// The row-block array must be bit-identical to the reference's for the same
// CSR row pointer, so the binning below follows the reference's rules step
// for step (src/library/internal/data-types/csr-meta.hpp L31-269), including
//   - the 32-bit "consecutive long rows" counter with its -1 sentinel,
//   - the reduction-thread count OR'ed into the *previous* entry,
//   - the tail entry's thread count taken from (n + 1 - last_i), and
//   - the size-only pass that always counts the tail entry (so
//     clsparseCsrMetaSize can exceed the created rowBlockSize by 2).
// What this file adds: the table of long rows (rows split over several
// entries) that the deterministic long-row combine kernel needs, and the
// per-entry partial-sum slots it writes.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <atomic>
#include <condition_variable>
#include <functional>
#include <memory>
#include <mutex>
#include <new>
#include <thread>
#include <vector>

#include <unistd.h>

#include "common.hpp"

namespace {

// numThreadsForReduction (csr-meta.hpp L52-68, GCC/Clang branch):
// 256 >> (32 - clz32(r - 1)). The shift is taken modulo 32, as x86 executes
// it, for the (unreachable with 32-bit row counts) case r - 1 >= 2^31.
inline uint64_t threads_for_reduction(uint64_t num_rows)
{
    const uint32_t r1 = static_cast<uint32_t>(num_rows - 1);
    const int shift = (32 - __builtin_clz(r1)) & 31;
    return static_cast<uint64_t>(256 >> shift);
}

// A small pool of host threads kept for the life of the process (starting
// threads costs about as much as the binning itself), used for one parallel
// pass at a time. A child of fork() inherits none of the threads: it runs its
// passes on the calling thread alone.
class HostPool
{
public:
    static HostPool& get()
    {
        // On the heap, and deleted at exit only by the process that started
        // its threads: in a child of fork() the condition variables may still
        // count the parent's waiting workers, and destroying them would wait
        // for ever.
        static HostPool* pool = new HostPool;
        static const Reaper reaper{pool};
        return *pool;
    }
    // Threads the pass may use (the calling thread included).
    static unsigned width()
    {
        const unsigned hw = std::thread::hardware_concurrency();
        return std::max(1u, std::min(16u, hw == 0 ? 1u : hw));
    }
    // Runs fn(t) for t = 0 .. T-1 on T threads at once (the caller takes
    // t = 0), so fn may wait for the others; returns when all have finished.
    // Falls back to fewer workers if threads cannot be started: returns the
    // count actually used, and fn must then be called with that T.
    unsigned start(unsigned T)
    {
        std::lock_guard<std::mutex> lk(mu_);
        if (getpid() != owner_) return 1;  // a child of fork(): the workers are the parent's
        try {
            while (workers_.size() + 1 < T) {
                const unsigned idx = static_cast<unsigned>(workers_.size()) + 1;
                workers_.emplace_back([this, idx] { loop(idx); });
            }
        } catch (...) {
        }
        return std::min<unsigned>(T, static_cast<unsigned>(workers_.size()) + 1);
    }
    void run(unsigned T, const std::function<void(unsigned)>& fn)
    {
        std::lock_guard<std::mutex> call(call_mu_);
        if (T > 1) {
            std::unique_lock<std::mutex> lk(mu_);
            job_ = &fn;
            job_T_ = T;
            pending_ = T - 1;
            ++gen_;
            cv_.notify_all();
        }
        fn(0);
        if (T > 1) {
            std::unique_lock<std::mutex> lk(mu_);
            done_.wait(lk, [this] { return pending_ == 0; });
            job_ = nullptr;
        }
    }
    ~HostPool()
    {
        {
            std::lock_guard<std::mutex> lk(mu_);
            stop_ = true;
            cv_.notify_all();
        }
        for (auto& w : workers_) w.join();
    }

private:
    struct Reaper
    {
        HostPool* pool;
        ~Reaper()
        {
            if (getpid() == pool->owner_) delete pool;
        }
    };
    void loop(unsigned idx)
    {
        uint64_t seen = 0;
        for (;;) {
            const std::function<void(unsigned)>* job;
            {
                std::unique_lock<std::mutex> lk(mu_);
                cv_.wait(lk, [&] { return stop_ || gen_ != seen; });
                if (stop_) return;
                seen = gen_;
                if (idx >= job_T_) continue;
                job = job_;
            }
            (*job)(idx);
            std::lock_guard<std::mutex> lk(mu_);
            if (--pending_ == 0) done_.notify_all();
        }
    }
    std::mutex call_mu_, mu_;
    std::condition_variable cv_, done_;
    std::vector<std::thread> workers_;
    const std::function<void(unsigned)>* job_ = nullptr;
    unsigned job_T_ = 0, pending_ = 0;
    uint64_t gen_ = 0;
    bool stop_ = false;
    const pid_t owner_ = getpid();  // the process whose threads workers_ lists
};

// The binning state between rows: the current block starts after row
// last_i, holds sum nonzeros, and long_run counts the consecutive rows of a
// long-row region (UINT32_MAX: leaving one).
struct BinState
{
    uint64_t last_i = 0, sum = 0;
    uint32_t long_run = 0;
};

// Processes rows [i_begin, i_end] of the binning (the loop of csr-meta.hpp
// L93-238) from state st. Entries go to out (when not null) after its last
// element, which is the entry a block closed here ORs its reduction-thread
// count into; total counts them. longs: the rows longer than 128 nonzeros,
// from which the rows of a block below the budget are skipped in one search
// (needs a non-decreasing row pointer). Returns the row after the last one
// processed (nRows + 1 when the rows ran out inside a block).
uint64_t bin_rows(const uint32_t* rp, uint32_t nRows, uint64_t i_begin, uint64_t i_end, BinState& st,
                  std::vector<uint64_t>* out, uint64_t& total, const std::vector<uint32_t>& longs)
{
    const uint64_t blk = kBlkSize;
    uint64_t last_i = st.last_i, sum = st.sum;
    uint32_t long_run = st.long_run;
    size_t lp = std::lower_bound(longs.begin(), longs.end(), static_cast<uint32_t>(i_begin)) - longs.begin();
    uint64_t skip_hint = 0;  // rows the last skip passed over

    auto close_block = [&](uint64_t end_row) {
        // New entry starts at end_row; the finished block [last_i, end_row)
        // gets its reduction-thread count when it has more than one row.
        if (end_row - last_i > static_cast<uint64_t>(kRowsForVector)) out->back() |= threads_for_reduction(end_row - last_i);
        out->push_back(end_row << (64 - kRowBits));
    };

    uint64_t i;
    for (i = i_begin; i <= i_end; ++i) {
        // Outside a long-row region a row of at most 128 nonzeros only adds
        // to the block's sum until the sum reaches the block size (at every
        // iteration boundary sum == rp[i - 1] - rp[last_i]): skip such rows
        // and take up the rules again at the first row that can change the
        // state -- the next long row, or the first row whose end reaches the
        // budget (a galloping search; the row pointer is non-decreasing).
        if (long_run == 0 && sum == static_cast<uint64_t>(rp[i - 1] - rp[last_i])) {
            const uint32_t base = rp[last_i];
            while (lp < longs.size() && longs[lp] < i) ++lp;
            const uint64_t lim = lp < longs.size() ? longs[lp] : static_cast<uint64_t>(nRows) + 1;
            uint64_t r = i;
            if (r < lim && rp[r] - base < blk) {
                // lo: a row whose end stays under the budget; hi: the first
                // row whose end reaches it, or lim. The search starts from the
                // previous skip's length (a regular matrix fills every block
                // with about as many rows), galloping down or up from there.
                uint64_t lo = r, hi = lim;
                const uint64_t g = r + skip_hint;
                if (skip_hint > 1 && g < lim) {
                    if (rp[g] - base < blk) {
                        lo = g;
                    } else {
                        hi = g;
                        uint64_t step = 1;
                        while (hi - lo > step && rp[hi - step] - base >= blk) {
                            hi -= step;
                            step *= 2;
                        }
                        if (hi - lo > step) lo = hi - step;  // rp[hi - step] - base < blk
                    }
                }
                if (hi == lim) {
                    uint64_t step = 1;
                    while (lo + step < lim && rp[lo + step] - base < blk) {
                        lo += step;
                        step *= 2;
                    }
                    hi = std::min(lo + step, lim);
                }
                while (hi - lo > 1) {
                    const uint64_t mid = lo + (hi - lo) / 2;
                    if (rp[mid] - base < blk)
                        lo = mid;
                    else
                        hi = mid;
                }
                r = hi;
            }
            skip_hint = r - i;
            if (r > nRows) {
                sum = rp[nRows] - base;
                i = static_cast<uint64_t>(nRows) + 1;
                break;
            }
            i = r;
            sum = rp[i - 1] - base;
        }
        const uint32_t len = rp[i] - rp[i - 1];
        sum += len;

        if (len > 128) {
            ++long_run;
        } else if (long_run > 0) {
            if (len < 32)
                long_run = UINT32_MAX;
            else
                ++long_run;
        }

        if (long_run == 1) {
            // Entering a long-row region: cut off the pending short rows.
            if (i - last_i > 1) {
                if (out) close_block(i - 1);
                ++total;
                last_i = i - 1;
                sum = len;
            }
        } else if (long_run == UINT32_MAX) {
            // First short row after a long-row region.
            if (out) close_block(i - 1);
            ++total;
            last_i = i - 1;
            sum = len;
            long_run = 0;
        }

        if (i - last_i == 1 && sum > blk) {
            // One row over budget: CSR-Vector (one entry) or CSR-LongRows
            // (ceil(len / 3072) entries, capped at 2^24).
            int k = static_cast<int>(std::ceil(static_cast<double>(len) / kLongSeg));
            k = k < (1 << kWgBits) ? k : (1 << kWgBits);
            if (out) {
                for (int sg = 1; sg < k; ++sg) out->push_back(((i - 1) << (64 - kRowBits)) | static_cast<uint64_t>(sg));
                out->push_back(i << (64 - kRowBits));
            }
            total += static_cast<uint64_t>(k);
            last_i = i;
            sum = 0;
            long_run = 0;
        } else if (i - last_i > 1 && sum > blk) {
            // Several rows over budget: back off the row that did not fit.
            --i;
            if (out) close_block(i);
            ++total;
            last_i = i;
            sum = 0;
            long_run = 0;
        } else if (sum == blk) {
            if (out) close_block(i);
            ++total;
            last_i = i;
            sum = 0;
            long_run = 0;
        }
    }
    st.last_i = last_i;
    st.sum = sum;
    st.long_run = long_run;
    return i;
}

// One pass over the row pointer. With out == nullptr it only counts entries
// (ComputeRowBlocksSize); otherwise it writes them. Returns the reference's
// rowBlockSize: 2 * entries written, or 2 * entries counted; a writing pass
// also stores the counting pass's value in *counted when given. Returns 0
// (no entries) for a row pointer that decreases when `reject_decreasing`.
//
// Work split: a row longer than 128 nonzeros that follows a row shorter than
// 32 leaves the same state behind whatever came before it (a long-row region
// cannot be running at the short row, so the long row starts one and the
// block before it is cut at it). The rows are split into T chunks; thread t
// lists its chunk's long rows and first such "sync" row S_t, then (after a
// barrier) bins the rows (S_t, S_{t+1}] from the state that S_t leaves
// (thread 0 from the start), where S_{t+1} is the next published sync row.
// The ranges' entries are joined in order (a range's first cut ORs its
// thread count into the previous range's last entry). The search of bin_rows
// needs a non-decreasing row pointer; a decreasing one is binned by the
// reference's sequential pass instead (clsparseCsrMetaSize counts it).
uint64_t compute_row_blocks(const uint32_t* rp, uint32_t nRows, uint64_t* out, uint64_t* counted = nullptr,
                            bool reject_decreasing = false, bool* failed = nullptr)
{
    const bool write = out != nullptr;
    HostPool& pool = HostPool::get();
    unsigned T = std::min(HostPool::width(), std::max(1u, nRows / (1u << 17)));
    T = pool.start(T);

    struct Part
    {
        std::vector<uint32_t> longs;  // rows of the chunk longer than 128 nonzeros
        uint64_t sync = 0;            // first sync row of the chunk, 0 if none
        bool decreasing = false;
        std::vector<uint64_t> ent;    // ent[0]: the seam (ORed into the previous range's last entry)
        uint64_t total = 0, stop = 0;
        BinState st;
        bool ran = false;
        bool failed = false;  // out of host memory
    };
    std::vector<Part> part(T);
    std::atomic<unsigned> arrived{0};
    pool.run(T, [&](unsigned t) {
        Part& pt = part[t];
        const uint64_t a = 1 + static_cast<uint64_t>(nRows) * t / T, b = 1 + static_cast<uint64_t>(nRows) * (t + 1) / T;
        bool dec = false;
        try {
            for (uint64_t r = a; r < b; ++r) {
                dec |= rp[r] < rp[r - 1];
                if (rp[r] - rp[r - 1] > 128u) {
                    pt.longs.push_back(static_cast<uint32_t>(r));
                    if (pt.sync == 0 && r >= 2 && rp[r - 1] - rp[r - 2] < 32u) pt.sync = r;
                }
            }
        } catch (...) {
            pt.failed = true;
        }
        pt.decreasing = dec;
        arrived.fetch_add(1);  // every thread arrives, or the others would wait for ever
        while (arrived.load() < T) std::this_thread::yield();
        bool stop = false;
        for (const Part& q : part) stop |= q.decreasing || q.failed;
        if (stop) return;
        try {
            // Range (S_t, S_{t+1}]; thread 0 starts at the first row.
            const uint64_t lo = t == 0 ? 0 : pt.sync;
            if (t > 0 && lo == 0) return;  // no sync row in this chunk: the previous range covers it
            uint64_t hi = nRows;
            unsigned u = t + 1;
            while (u < T && part[u].sync == 0) ++u;
            if (u < T) hi = part[u].sync;
            std::vector<uint32_t> longs;
            for (unsigned k = t; k < T && (k == t || k <= u); ++k)
                for (uint32_t r : part[k].longs)
                    if (r > lo && r <= hi) longs.push_back(r);
            pt.ent.push_back(0);
            if (write) pt.ent.reserve(64 + 2 * (hi - lo) / 16);
            if (t > 0) {
                // The state after the sync row lo, which the previous range processed.
                const uint32_t len = rp[lo] - rp[lo - 1];
                if (len >= static_cast<uint32_t>(kBlkSize))  // a row of its own (CSR-Vector / long row)
                    pt.st = BinState{lo, 0, 0};
                else
                    pt.st = BinState{lo - 1, len, 1};
            }
            pt.stop = bin_rows(rp, nRows, lo + 1, hi, pt.st, write ? &pt.ent : nullptr, pt.total, longs);
            pt.ran = true;
        } catch (...) {
            pt.failed = true;
        }
    });
    for (const Part& q : part)
        if (q.failed) {
            if (failed) *failed = true;
            return 0;
        }

    bool decreasing = false;
    for (const Part& q : part) decreasing |= q.decreasing;
    if (decreasing) {
        if (reject_decreasing) return 0;
        // Every row may stop the skip: bin_rows is then the sequential pass.
        try {
            part.assign(1, Part{});
            std::vector<uint32_t> all;
            all.reserve(nRows);
            for (uint32_t r = 1; r <= nRows; ++r) all.push_back(r);
            part[0].ent.push_back(0);
            part[0].stop =
                bin_rows(rp, nRows, 1, nRows, part[0].st, write ? &part[0].ent : nullptr, part[0].total, all);
            part[0].ran = true;
        } catch (...) {
            if (failed) *failed = true;
            return 0;
        }
    }

    uint64_t total = 1;  // rowBlocks[0]
    uint64_t* w = out;
    if (write) *w++ = 0;
    const Part* last = nullptr;
    for (const Part& q : part) {
        if (!q.ran) continue;
        total += q.total;
        if (write) {
            *(w - 1) |= q.ent[0];
            std::copy(q.ent.begin() + 1, q.ent.end(), w);
            w += q.ent.size() - 1;
        }
        last = &q;
    }

    // Tail: the thread count uses i == nRows + 1 here, one more row than the
    // block holds (csr-meta.hpp L241-248).
    const uint64_t i = last->stop;
    const uint64_t last_i = last->st.last_i;
    if (write && (*(w - 1) >> (64 - kRowBits)) != static_cast<uint64_t>(nRows)) {
        *w = static_cast<uint64_t>(nRows) << (64 - kRowBits);
        if (static_cast<uint64_t>(nRows) - last_i > static_cast<uint64_t>(kRowsForVector))
            *(w - 1) |= threads_for_reduction(i - last_i);
        ++w;
    }
    ++total;

    if (counted) *counted = 2 * total;
    if (write) return 2 * static_cast<uint64_t>(w - out);
    return 2 * total;
}

// Upper bound on the entries of compute_row_blocks: the first, at most one
// per row that ends a block or is shorter than a segment, one more per
// 3072-nonzero segment, and the tail.
uint64_t row_blocks_bound(uint32_t nRows, uint64_t nnz)
{
    return 2 * static_cast<uint64_t>(nRows) + nnz / kLongSeg + 4;
}

clsparseStatus read_row_pointer(const clsparseCsrMatrix* A, clsparseControl control,
                                std::vector<uint32_t>& rp)
{
    rp.resize(static_cast<size_t>(A->num_rows) + 1);
    if (A->row_pointer == nullptr) return clsparseInvalidMemObject;
    CLS_HIP_TRY(hipMemcpyAsync(rp.data(), dev_ptr<uint32_t>(A->row_pointer, A->off_row_pointer),
                               rp.size() * sizeof(uint32_t), hipMemcpyDeviceToHost,
                               control->stream));
    CLS_HIP_TRY(hipStreamSynchronize(control->stream));
    return clsparseSuccess;
}

void free_meta(matrix_meta* m)
{
    if (m == nullptr) return;
    int cur = 0;
    (void)hipGetDevice(&cur);
    (void)hipSetDevice(m->device);
    if (m->d_rowBlocks) (void)hipFree(m->d_rowBlocks);
    if (m->d_desc) (void)hipFree(m->d_desc);
    if (m->d_long) (void)hipFree(m->d_long);
    if (m->d_partials) (void)hipFree(m->d_partials);
    (void)hipSetDevice(cur);
    m->magic = 0;
    delete m;
}

}  // namespace

// Host-side entry used by tests through the C ABI as well: the row-block
// array for a host row pointer (no device needed).
extern "C" CLSPARSE_EXPORT uint64_t clsparseAmdComputeRowBlocksHost(const uint32_t* rp,
                                                                    uint32_t nRows,
                                                                    uint64_t* out)
{
    return compute_row_blocks(rp, nRows, out);
}

// src/library/internal/data-types/csr-meta.cpp L23-42
extern "C" clsparseMetaSizeResult clsparseCsrMetaSize(clsparseCsrMatrix* csrMatx,
                                                      clsparseControl control)
{
    clsparseMetaSizeResult r{clsparseSuccess, 0};
    if (csrMatx == nullptr) {
        r.status = clsparseInvalidValue;
        return r;
    }
    if (csrMatx->meta) {
        r.metaSize = static_cast<matrix_meta*>(csrMatx->meta)->rowBlockSize;
        return r;
    }
    if (control == nullptr) {
        r.status = clsparseInvalidControlObject;
        return r;
    }
    std::vector<uint32_t> rp;
    r.status = read_row_pointer(csrMatx, control, rp);
    if (r.status != clsparseSuccess) return r;
    bool failed = false;
    r.metaSize = static_cast<clsparseIdx_t>(compute_row_blocks(rp.data(), csrMatx->num_rows, nullptr, nullptr,
                                                               false, &failed));
    if (failed) r.status = clsparseOutOfHostMemory;
    return r;
}

// src/library/internal/data-types/csr-meta.cpp L44-83
extern "C" clsparseStatus clsparseCsrMetaCreate(clsparseCsrMatrix* csrMatx,
                                                clsparseControl control)
{
    if (csrMatx == nullptr) return clsparseInvalidValue;
    if (control == nullptr) return clsparseInvalidControlObject;

    // The row pointer comes to the host through the control's pinned buffer.
    const uint32_t n = csrMatx->num_rows;
    if (csrMatx->row_pointer == nullptr) return clsparseInvalidMemObject;
    void* pinned = nullptr;
    CLS_TRY(control_pinned(control, (static_cast<size_t>(n) + 1) * sizeof(uint32_t), &pinned));
    const uint32_t* rp = static_cast<const uint32_t*>(pinned);
    CLS_HIP_TRY(hipMemcpyAsync(pinned, dev_ptr<uint32_t>(csrMatx->row_pointer, csrMatx->off_row_pointer),
                               (static_cast<size_t>(n) + 1) * sizeof(uint32_t), hipMemcpyDeviceToHost,
                               control->stream));
    CLS_HIP_TRY(hipStreamSynchronize(control->stream));

    // The kernels index vals/cols through this row pointer; a malformed one
    // would send them out of bounds, so check it here on the host: the end
    // here, a decrease in the binning's pass (which then writes nothing).
    if (rp[n] > csrMatx->num_nonzeros) return clsparseInvalidValue;

    // One writing pass into a buffer of twice the bound on the entries: the
    // device array has the counting pass's size, two words per entry (the
    // second half zero, as the reference's).
    const uint64_t bound = 2 * row_blocks_bound(n, rp[n] >= rp[0] ? rp[n] - rp[0] : 0);
    std::unique_ptr<uint64_t[]> host(new (std::nothrow) uint64_t[bound]);
    if (!host) return clsparseOutOfHostMemory;
    uint64_t size_words = 0;
    bool failed = false;
    const uint64_t created = compute_row_blocks(rp, n, host.get(), &size_words, true, &failed);
    if (failed) return clsparseOutOfHostMemory;
    if (created == 0) return clsparseInvalidValue;  // the row pointer decreases
    if (size_words > bound) return clsparseOutOfHostMemory;

    if (csrMatx->meta) {
        free_meta(static_cast<matrix_meta*>(csrMatx->meta));
        csrMatx->meta = nullptr;
    }

    matrix_meta* m = new (std::nothrow) matrix_meta;
    if (m == nullptr) return clsparseOutOfHostMemory;
    m->device = control->device;
    m->num_rows = n;
    m->num_nonzeros = csrMatx->num_nonzeros;
    std::fill(host.get() + created / 2, host.get() + size_words, uint64_t(0));
    m->rowBlockSize = static_cast<clsparseIdx_t>(created);
    m->entries = static_cast<uint32_t>(created / 2);
    m->h_rowBlocks.assign(host.get(), host.get() + m->entries);

    // Long rows: entry f with an empty row range starts a row split into
    // segments f, f+1, ... whose low bits count 1, 2, ...; the last segment
    // is the entry whose successor starts at the next row.
    std::vector<LongRow> longs;
    for (uint32_t e = 0; e + 1 < m->entries; ++e) {
        const uint32_t row = static_cast<uint32_t>(host[e] >> 32);
        const uint32_t stop = static_cast<uint32_t>(host[e + 1] >> 32);
        if (stop == row) {
            uint32_t k = 1;
            while (e + k + 1 < m->entries && static_cast<uint32_t>(host[e + k + 1] >> 32) == row) ++k;
            ++k;  // the final segment (its successor starts at row + 1)
            longs.push_back(LongRow{row, e, k, 0});
            e += k - 1;
        }
    }
    m->num_long_rows = static_cast<uint32_t>(longs.size());

    // Kernel descriptors: the entry classification of csrmv_adaptive.cl
    // L349-359, L382, L515, L564 resolved once, with each block's nnz range.
    std::vector<BlockDesc> desc(m->entries > 0 ? m->entries - 1 : 0);
    for (uint32_t e = 0; e + 1 < m->entries; ++e) {
        const uint32_t row = static_cast<uint32_t>(host[e] >> 32);
        const uint32_t stop = static_cast<uint32_t>(host[e + 1] >> 32);
        const uint32_t seg = static_cast<uint32_t>(host[e] & ((1u << kWgBits) - 1u));
        const uint32_t nrows = stop - row;
        BlockDesc d{row, nrows, rp[row], 0};
        if (nrows > static_cast<uint32_t>(kRowsForVector)) {
            d.cnt = rp[stop] - rp[row];
            if (d.cnt > static_cast<uint32_t>(kBlkSize)) {  // impossible for a valid binning
                delete m;
                return clsparseStructInvalid;
            }
        } else if (nrows == 1 && seg == 0) {
            d.cnt = rp[row + 1] - rp[row];
        } else {
            const uint64_t rs = rp[row], re = rp[row + 1];
            const uint64_t vs = std::min<uint64_t>(rs + static_cast<uint64_t>(seg) * kLongSeg, re);
            const uint64_t ve = std::min<uint64_t>(vs + kLongSeg, re);
            d.nrows = 0;
            d.base = static_cast<uint32_t>(vs);
            d.cnt = static_cast<uint32_t>(ve - vs);
        }
        desc[e] = d;
        if (d.nrows > 64 && d.nrows < 512) m->rowwise_fraction += 1.0;
    }
    if (!desc.empty()) m->rowwise_fraction /= static_cast<double>(desc.size());

    hipError_t err = hipMalloc(&m->d_rowBlocks, size_words * sizeof(uint64_t));
    if (err == hipSuccess)
        err = hipMemcpyAsync(m->d_rowBlocks, host.get(), size_words * sizeof(uint64_t),
                             hipMemcpyHostToDevice, control->stream);
    if (err == hipSuccess && !desc.empty()) {
        err = hipMalloc(&m->d_desc, desc.size() * sizeof(BlockDesc));
        if (err == hipSuccess)
            err = hipMemcpyAsync(m->d_desc, desc.data(), desc.size() * sizeof(BlockDesc),
                                 hipMemcpyHostToDevice, control->stream);
    }
    if (err == hipSuccess && !longs.empty()) {
        err = hipMalloc(&m->d_long, longs.size() * sizeof(LongRow));
        if (err == hipSuccess)
            err = hipMemcpyAsync(m->d_long, longs.data(), longs.size() * sizeof(LongRow),
                                 hipMemcpyHostToDevice, control->stream);
        // Two doubles per entry: segment sum and (extended mode) its error.
        if (err == hipSuccess)
            err = hipMalloc(&m->d_partials, 2 * static_cast<size_t>(m->entries) * sizeof(double));
    }
    if (err == hipSuccess) err = hipStreamSynchronize(control->stream);
    if (err != hipSuccess) {
        std::fprintf(stderr, "clSPARSE(hip) meta create failed: %s\n", hipGetErrorString(err));
        free_meta(m);
        return clsparseOutOfResources;
    }
    csrMatx->meta = m;
    return clsparseSuccess;
}

extern "C" clsparseStatus clsparseCsrMetaCompute(clsparseCsrMatrix* csrMatx,
                                                 clsparseControl control)
{
    return clsparseCsrMetaCreate(csrMatx, control);
}

// src/library/internal/data-types/csr-meta.cpp L85-100
extern "C" clsparseStatus clsparseCsrMetaDelete(clsparseCsrMatrix* csrMatx)
{
    if (csrMatx == nullptr) return clsparseInvalidValue;
    if (csrMatx->meta == nullptr) return clsparseSuccess;
    free_meta(static_cast<matrix_meta*>(csrMatx->meta));
    csrMatx->meta = nullptr;
    return clsparseSuccess;
}

extern "C" clsparseStatus clsparseAmdCsrMetaGetRowBlocks(const clsparseCsrMatrix* csrMatx,
                                                         cl_ulong* host_out,
                                                         clsparseIdx_t capacity,
                                                         clsparseIdx_t* entries)
{
    if (csrMatx == nullptr || csrMatx->meta == nullptr) return clsparseStructInvalid;
    const matrix_meta* m = get_meta(csrMatx);
    if (entries) *entries = m->entries;
    if (host_out) {
        const size_t k = capacity < m->entries ? capacity : m->entries;
        std::memcpy(host_out, m->h_rowBlocks.data(), k * sizeof(uint64_t));
    }
    return clsparseSuccess;
}
