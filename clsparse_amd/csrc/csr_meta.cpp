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
#include <memory>
#include <new>

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

// One pass over the row pointer. With out == nullptr it only counts entries
// (ComputeRowBlocksSize); otherwise it writes them. Returns the reference's
// rowBlockSize: 2 * entries written, or 2 * entries counted; a writing pass
// also stores the counting pass's value in *counted when given.
uint64_t compute_row_blocks(const uint32_t* rp, uint32_t nRows, uint64_t* out, uint64_t* counted = nullptr)
{
    const bool write = out != nullptr;
    uint64_t* w = out;
    uint64_t total = 1;  // rowBlocks[0]
    if (write) *w++ = 0;

    uint64_t sum = 0;
    uint64_t i = 0, last_i = 0;
    uint32_t long_run = 0;  // consecutive rows longer than 128 nnz; UINT32_MAX = leaving
    const uint64_t blk = kBlkSize;

    auto close_block = [&](uint64_t end_row) {
        // New entry starts at end_row; the finished block [last_i, end_row)
        // gets its reduction-thread count when it has more than one row.
        *w = end_row << (64 - kRowBits);
        if (end_row - last_i > static_cast<uint64_t>(kRowsForVector))
            *(w - 1) |= threads_for_reduction(end_row - last_i);
        ++w;
    };

    for (i = 1; i <= nRows; ++i) {
        // Outside a long-row region a row of at most 128 nonzeros only adds
        // to the block's sum until the sum reaches the block size (at every
        // iteration boundary sum == rp[i - 1] - rp[last_i]): skip such rows
        // and take up the rules again at the first row that can change the
        // state.
        if (long_run == 0 && sum == static_cast<uint64_t>(rp[i - 1] - rp[last_i])) {
            const uint32_t base = rp[last_i];
            uint64_t r = i;
            while (r <= nRows && rp[r] - rp[r - 1] <= 128u && rp[r] - base < blk) ++r;
            if (r > nRows) {
                i = nRows + 1;
                sum = rp[nRows] - base;
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
                if (write) close_block(i - 1);
                ++total;
                last_i = i - 1;
                sum = len;
            }
        } else if (long_run == UINT32_MAX) {
            // First short row after a long-row region.
            if (write) close_block(i - 1);
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
            if (write) {
                for (int s = 1; s < k; ++s)
                    *w++ = ((i - 1) << (64 - kRowBits)) | static_cast<uint64_t>(s);
                *w++ = i << (64 - kRowBits);
            }
            total += static_cast<uint64_t>(k);
            last_i = i;
            sum = 0;
            long_run = 0;
        } else if (i - last_i > 1 && sum > blk) {
            // Several rows over budget: back off the row that did not fit.
            --i;
            if (write) close_block(i);
            ++total;
            last_i = i;
            sum = 0;
            long_run = 0;
        } else if (sum == blk) {
            if (write) close_block(i);
            ++total;
            last_i = i;
            sum = 0;
            long_run = 0;
        }
    }

    // Tail: the thread count uses i == nRows + 1 here, one more row than the
    // block holds (csr-meta.hpp L241-248).
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
    r.metaSize = static_cast<clsparseIdx_t>(compute_row_blocks(rp.data(), csrMatx->num_rows, nullptr));
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
    // would send them out of bounds, so check it once here on the host.
    bool decreasing = false;
    for (uint32_t r = 0; r < n; ++r) decreasing |= rp[r + 1] < rp[r];
    if (decreasing) return clsparseInvalidValue;
    if (rp[n] > csrMatx->num_nonzeros) return clsparseInvalidValue;

    if (csrMatx->meta) {
        free_meta(static_cast<matrix_meta*>(csrMatx->meta));
        csrMatx->meta = nullptr;
    }

    matrix_meta* m = new (std::nothrow) matrix_meta;
    if (m == nullptr) return clsparseOutOfHostMemory;
    m->device = control->device;
    m->num_rows = n;
    m->num_nonzeros = csrMatx->num_nonzeros;

    // One writing pass into a buffer of twice the bound on the entries: the
    // device array has the counting pass's size, two words per entry (the
    // second half zero, as the reference's).
    const uint64_t bound = 2 * row_blocks_bound(n, rp[n] - rp[0]);
    std::unique_ptr<uint64_t[]> host(new (std::nothrow) uint64_t[bound]);
    if (!host) {
        delete m;
        return clsparseOutOfHostMemory;
    }
    uint64_t size_words = 0;
    const uint64_t created = compute_row_blocks(rp, n, host.get(), &size_words);
    if (size_words > bound) {
        delete m;
        return clsparseOutOfHostMemory;
    }
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
