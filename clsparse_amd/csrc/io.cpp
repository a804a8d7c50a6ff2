// Matrix Market (.mtx) readers: the reference's file entry points
// (src/include/clSPARSE.h L508-611, src/library/io/mm-reader.cpp L441-760).
//
// Behaviour kept from the reference:
//   * the path must end in ".mtx" (else clsparseInvalidFileFormat);
//   * coordinate storage only, real / integer / pattern data (complex and
//     array files are rejected with clsparseInvalidFile);
//   * the header's nnz is doubled for symmetric files (an upper bound);
//   * symmetric files are expanded (every off-diagonal entry mirrored);
//   * explicit zeroes are dropped unless read_explicit_zeroes is set;
//   * pattern files get values MAX_RAND_VAL * rand() / (RAND_MAX + 1.0), with
//     MAX_RAND_VAL = 5 (mmio.h L100);
//   * entries are sorted by (row, column) before the upload;
//   * the caller has allocated the device arrays; num_rows / num_cols /
//     num_nonzeros of the structure are overwritten.
// Differences (supersets): duplicates keep their file order (stable sort);
// an index outside the declared size, or fewer entries than the header
// announces, is an error (clsparseInvalidFile) instead of undefined
// behaviour; the file is parsed from one read, not per-token fscanf: in
// parallel over newline-aligned chunks (one thread per chunk) when every entry
// sits inside its chunk, else sequentially; the (row, column) order comes from
// a stable counting sort by row and a stable sort of each row by column.
#include <algorithm>
#include <cctype>
#include <cerrno>
#include <charconv>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <new>
#include <string>
#include <string_view>
#include <thread>
#include <vector>

#include <fcntl.h>
#include <sys/mman.h>
#include <sys/stat.h>
#include <unistd.h>

#include "common.hpp"
#include "host_threads.hpp"

namespace {

constexpr double kMaxRandVal = 5.0;

enum class MmData { Real, Integer, Pattern };

struct MmHeader
{
    MmData data = MmData::Real;
    bool symmetric = false;
    uint64_t rows = 0, cols = 0, nnz = 0;  // nnz as declared in the file
    size_t body = 0;                       // offset of the first entry
};

bool has_mtx_extension(const char* path)
{
    if (path == nullptr) return false;
    const std::string s(path);
    const size_t dot = s.find_last_of('.');
    return dot != std::string::npos && s.substr(dot + 1) == "mtx";
}

// The file mapped read-only (no copy of the text).
struct MappedFile
{
    const char* data = nullptr;
    size_t size = 0;
    explicit MappedFile(const char* path)
    {
        const int fd = ::open(path, O_RDONLY);
        if (fd < 0) return;
        struct stat st;
        if (::fstat(fd, &st) == 0 && st.st_size > 0) {
            void* m = ::mmap(nullptr, static_cast<size_t>(st.st_size), PROT_READ, MAP_PRIVATE, fd, 0);
            if (m != MAP_FAILED) {
                data = static_cast<const char*>(m);
                size = static_cast<size_t>(st.st_size);
            }
        }
        ::close(fd);
    }
    ~MappedFile()
    {
        if (data) ::munmap(const_cast<char*>(data), size);
    }
    MappedFile(const MappedFile&) = delete;
    MappedFile& operator=(const MappedFile&) = delete;
};

std::string lower(std::string s)
{
    for (char& c : s) c = static_cast<char>(std::tolower(static_cast<unsigned char>(c)));
    return s;
}

// Banner "%%MatrixMarket matrix coordinate <data> <storage>", comment lines,
// then the size line "rows cols nnz" (blank lines before it are skipped).
bool parse_header(std::string_view text, MmHeader& h)
{
    size_t pos = text.find('\n');
    const std::string first(text.substr(0, pos));
    char banner[64] = {0}, mtx[64] = {0}, crd[64] = {0}, data[64] = {0}, storage[64] = {0};
    if (std::sscanf(first.c_str(), "%63s %63s %63s %63s %63s", banner, mtx, crd, data, storage) != 5)
        return false;
    if (std::strncmp(banner, "%%MatrixMarket", 14) != 0) return false;
    if (lower(mtx) != "matrix" || lower(crd) != "coordinate") return false;
    const std::string d = lower(data), st = lower(storage);
    if (d == "real")
        h.data = MmData::Real;
    else if (d == "integer")
        h.data = MmData::Integer;
    else if (d == "pattern")
        h.data = MmData::Pattern;
    else
        return false;  // complex or unknown
    if (st == "symmetric")
        h.symmetric = true;
    else if (st != "general" && st != "skew-symmetric" && st != "hermitian")
        return false;
    // comments, blank lines, then the size line
    while (pos != std::string::npos) {
        const size_t start = pos + 1;
        pos = text.find('\n', start);
        const std::string line(text.substr(start, pos == std::string::npos ? std::string::npos : pos - start));
        if (!line.empty() && line[0] == '%') continue;
        unsigned long long r = 0, c = 0, z = 0;
        if (std::sscanf(line.c_str(), "%llu %llu %llu", &r, &c, &z) == 3) {
            h.rows = r;
            h.cols = c;
            h.nnz = z;
            h.body = (pos == std::string::npos) ? text.size() : pos + 1;
            return true;
        }
        // blank or partial line: keep scanning
    }
    return false;
}

template <typename T>
struct Entry
{
    uint32_t r, c;
    T v;
};

bool is_space(char ch)
{
    return ch == ' ' || ch == '\t' || ch == '\n' || ch == '\v' || ch == '\f' || ch == '\r';
}

// One unsigned integer token as strtoull reads it (leading white space, an
// optional '+'); a '-' sign makes the token invalid here (strtoull would
// wrap it past any matrix size, which the range check rejects).
bool parse_index(const char*& p, const char* end, uint64_t& out)
{
    while (p < end && is_space(*p)) ++p;
    if (p < end && *p == '+') ++p;
    const auto r = std::from_chars(p, end, out);
    if (r.ec != std::errc() || r.ptr == p) return false;
    p = r.ptr;
    return true;
}

// One value token: from_chars on the common decimal forms, strtod / strtof
// (which the reference's fscanf follows) for anything else.
template <typename T>
bool parse_number(const char*& p, const char* end, T& out)
{
    while (p < end && is_space(*p)) ++p;
    const char* q = p;
    if (q < end && *q == '+' && q + 1 < end && q[1] != '-' && q[1] != '+') ++q;
    const auto r = std::from_chars(q, end, out);
    if (r.ec == std::errc() && r.ptr != q && (r.ptr == end || is_space(*r.ptr))) {
        p = r.ptr;
        return true;
    }
    // Slow path on a NUL-terminated copy of the token.
    const char* t = p;
    while (t < end && !is_space(*t)) ++t;
    const std::string tok(p, t);
    char* e = nullptr;
    if constexpr (sizeof(T) == 4)
        out = std::strtof(tok.c_str(), &e);
    else
        out = std::strtod(tok.c_str(), &e);
    if (e == tok.c_str()) return false;
    p += (e - tok.c_str());
    return true;
}

template <typename T>
struct Raw
{
    uint64_t r, c;
    T v;
};

// Capacity to reserve for the entries of the text [p, end): at most `want`,
// and no more than the text can hold (an entry takes at least four bytes,
// "1 1" and a line end), whatever the header claims.
size_t entry_bound(uint64_t want, const char* p, const char* end)
{
    return static_cast<size_t>(std::min<uint64_t>(want, static_cast<uint64_t>(end - p) / 4 + 1));
}

// Entries of [p, end) until `want` are read or the text ends; *crossed is set
// when an entry runs past `end` (it continues beyond this chunk). Returns
// false on a malformed token.
template <typename T>
bool parse_range(const char* p, const char* end, const char* hard_end, bool pattern, uint64_t want,
                 std::vector<Raw<T>>& out, bool* crossed)
{
    *crossed = false;
    while (out.size() < want) {
        while (p < end && is_space(*p)) ++p;
        if (p >= end) return true;
        Raw<T> e{0, 0, T(0)};
        if (!parse_index(p, hard_end, e.r) || !parse_index(p, hard_end, e.c)) return false;
        if (!pattern && !parse_number<T>(p, hard_end, e.v)) return false;
        if (p > end) {
            *crossed = true;
            return true;
        }
        out.push_back(e);
    }
    return true;
}

unsigned reader_threads()
{
    const unsigned t = std::thread::hardware_concurrency();
    return t == 0 ? 1u : std::min(t, 16u);
}

// Raw entries [0, take) -> 0-based entries: pattern values are drawn by the
// caller beforehand; range check, zero filtering, symmetric expansion.
template <typename T>
bool filter_entries(const Raw<T>* raw, size_t take, const MmHeader& h, bool explicit_zeroes,
                    std::vector<Entry<T>>& e)
{
    e.reserve(h.symmetric ? 2 * take : take);
    for (size_t k = 0; k < take; ++k) {
        const Raw<T>& x = raw[k];
        if (x.r < 1 || x.c < 1 || x.r > h.rows || x.c > h.cols) return false;
        if (!explicit_zeroes && x.v == T(0)) continue;
        const uint32_t r = static_cast<uint32_t>(x.r - 1), c = static_cast<uint32_t>(x.c - 1);
        e.push_back({r, c, x.v});
        if (h.symmetric && r != c) e.push_back({c, r, x.v});
    }
    return true;
}

// Entries (0-based, symmetric expansion, zero filtering), sorted by (row, col).
template <typename T>
clsparseStatus read_entries(const char* path, bool explicit_zeroes, MmHeader& h, std::vector<Entry<T>>& out)
{
    const MappedFile file(path);
    if (!file.data) return clsparseInvalidFile;
    if (!parse_header(std::string_view(file.data, file.size), h)) return clsparseInvalidFile;
    if (h.rows > 0xFFFFFFFFull || h.cols > 0xFFFFFFFFull || h.nnz > 0xFFFFFFFFull)
        return clsparseInvalidFile;
    const bool pattern = h.data == MmData::Pattern;
    const char* body = file.data + h.body;
    const char* text_end = file.data + file.size;

    // Raw entries per chunk; take[k] of chunk k are among the first nnz.
    const size_t len = static_cast<size_t>(text_end - body);
    unsigned nt = len >= (1u << 20) ? reader_threads() : 1u;
    std::vector<std::vector<Raw<T>>> part;
    std::vector<size_t> take;
    bool done = false;
    if (nt > 1) {
        std::vector<const char*> cut(nt + 1, text_end);
        cut[0] = body;
        for (unsigned k = 1; k < nt; ++k) {
            const char* q = std::max(body + len * k / nt, cut[k - 1]);
            const void* nl = std::memchr(q, '\n', static_cast<size_t>(text_end - q));
            cut[k] = nl ? static_cast<const char*>(nl) + 1 : text_end;
        }
        part.assign(nt, {});
        std::vector<char> ok(nt, 1), crossed(nt, 0);
        if (!run_on_threads(nt, [&](unsigned k) {
                bool cr = false;
                part[k].reserve(entry_bound(h.nnz / nt + 16, cut[k], cut[k + 1]));
                ok[k] = parse_range<T>(cut[k], cut[k + 1], text_end, pattern, h.nnz, part[k], &cr);
                crossed[k] = cr;
            }))
            throw std::bad_alloc();
        // Usable when every chunk up to the nnz-th entry parsed cleanly and no
        // entry ran across a chunk boundary; otherwise parse sequentially.
        bool regular = true;
        uint64_t have = 0;
        take.assign(nt, 0);
        for (unsigned k = 0; k < nt && have < h.nnz; ++k) {
            if (!ok[k] || crossed[k]) {
                regular = false;
                break;
            }
            take[k] = static_cast<size_t>(std::min<uint64_t>(part[k].size(), h.nnz - have));
            have += take[k];
        }
        done = regular && have >= h.nnz;
    }
    if (!done) {
        nt = 1;
        part.assign(1, {});
        part[0].reserve(entry_bound(h.nnz, body, text_end));
        bool cr = false;
        if (!parse_range<T>(body, text_end, text_end, pattern, h.nnz, part[0], &cr)) return clsparseInvalidFile;
        if (part[0].size() < h.nnz) return clsparseInvalidFile;
        take.assign(1, part[0].size());
    }

    // Pattern values are drawn in entry order; then each chunk is filtered on
    // its own thread.
    if (pattern)
        for (unsigned k = 0; k < nt; ++k)
            for (size_t i = 0; i < take[k]; ++i)
                part[k][i].v = static_cast<T>(kMaxRandVal * (std::rand() / (RAND_MAX + 1.0)));
    std::vector<std::vector<Entry<T>>> ent(nt);
    {
        std::vector<char> ok(nt, 1);
        if (!run_on_threads(nt, [&](unsigned k) {
                ok[k] = filter_entries<T>(part[k].data(), take[k], h, explicit_zeroes, ent[k]);
            }))
            throw std::bad_alloc();
        for (unsigned k = 0; k < nt; ++k)
            if (!ok[k]) return clsparseInvalidFile;
    }
    std::vector<std::vector<Raw<T>>>().swap(part);

    // Stable (row, column) order: counting sort by row over the chunks in
    // file order, then each row by column.
    std::vector<uint64_t> start(static_cast<size_t>(h.rows) + 1, 0);
    size_t total = 0;
    for (const auto& e : ent) {
        total += e.size();
        for (const auto& x : e) ++start[x.r + 1];
    }
    if (total > 0xFFFFFFFFull) return clsparseInvalidFile;  // the symmetric expansion passed 32-bit counts
    for (uint64_t r = 0; r < h.rows; ++r) start[r + 1] += start[r];
    out.resize(total);
    {
        std::vector<uint64_t> fill(start.begin(), start.end() - 1);
        for (const auto& e : ent)
            for (const auto& x : e) out[fill[x.r]++] = x;
    }
    std::vector<std::vector<Entry<T>>>().swap(ent);
    for (uint64_t r = 0; r < h.rows; ++r) {
        auto b = out.begin() + static_cast<std::ptrdiff_t>(start[r]);
        auto en = out.begin() + static_cast<std::ptrdiff_t>(start[r + 1]);
        if (!std::is_sorted(b, en, [](const Entry<T>& x, const Entry<T>& y) { return x.c < y.c; }))
            std::stable_sort(b, en, [](const Entry<T>& x, const Entry<T>& y) { return x.c < y.c; });
    }
    return clsparseSuccess;
}

// CSR row pointer of sorted entries.
template <typename T>
void row_pointer(const std::vector<Entry<T>>& e, uint32_t rows, std::vector<uint32_t>& rp)
{
    rp.assign(static_cast<size_t>(rows) + 1, 0);
    for (const auto& x : e) ++rp[x.r + 1];
    for (uint32_t r = 0; r < rows; ++r) rp[r + 1] += rp[r];
}

template <typename T>
clsparseStatus coo_from_file_impl(clsparseCooMatrix* coo, const char* path, clsparseControl control,
                                  cl_bool explicit_zeroes)
{
    CLS_TRY(check_init_control(control));
    if (coo == nullptr) return clsparseInvalidValue;
    if (!has_mtx_extension(path)) return clsparseInvalidFileFormat;
    MmHeader h;
    std::vector<Entry<T>> e;
    CLS_TRY(read_entries<T>(path, explicit_zeroes != 0, h, e));
    const size_t nnz = e.size();
    if (nnz > 0 && (coo->values == nullptr || coo->col_indices == nullptr || coo->row_indices == nullptr))
        return clsparseInvalidMemObj;
    std::vector<T> v(nnz);
    std::vector<uint32_t> r(nnz), c(nnz);
    for (size_t i = 0; i < nnz; ++i) {
        r[i] = e[i].r;
        c[i] = e[i].c;
        v[i] = e[i].v;
    }
    coo->num_rows = static_cast<clsparseIdx_t>(h.rows);
    coo->num_cols = static_cast<clsparseIdx_t>(h.cols);
    coo->num_nonzeros = static_cast<clsparseIdx_t>(nnz);
    hipStream_t s = control->stream;
    if (nnz) {
        CLS_HIP_TRY(hipMemcpyAsync(dev_ptr<T>(coo->values, coo->off_values), v.data(), nnz * sizeof(T),
                                   hipMemcpyHostToDevice, s));
        CLS_HIP_TRY(hipMemcpyAsync(dev_ptr<uint32_t>(coo->col_indices, coo->off_col_indices), c.data(),
                                   nnz * sizeof(uint32_t), hipMemcpyHostToDevice, s));
        CLS_HIP_TRY(hipMemcpyAsync(dev_ptr<uint32_t>(coo->row_indices, coo->off_row_indices), r.data(),
                                   nnz * sizeof(uint32_t), hipMemcpyHostToDevice, s));
    }
    CLS_HIP_TRY(hipStreamSynchronize(s));
    return clsparseSuccess;
}

template <typename T>
clsparseStatus csr_from_file_impl(clsparseCsrMatrix* csr, const char* path, clsparseControl control,
                                  cl_bool explicit_zeroes)
{
    CLS_TRY(check_init_control(control));
    if (csr == nullptr) return clsparseInvalidValue;
    if (!has_mtx_extension(path)) return clsparseInvalidFileFormat;
    MmHeader h;
    std::vector<Entry<T>> e;
    CLS_TRY(read_entries<T>(path, explicit_zeroes != 0, h, e));
    const size_t nnz = e.size();
    if (csr->row_pointer == nullptr || (nnz > 0 && (csr->values == nullptr || csr->col_indices == nullptr)))
        return clsparseInvalidMemObj;
    std::vector<uint32_t> rp, c(nnz);
    std::vector<T> v(nnz);
    row_pointer(e, static_cast<uint32_t>(h.rows), rp);
    for (size_t i = 0; i < nnz; ++i) {
        c[i] = e[i].c;
        v[i] = e[i].v;
    }
    csr->num_rows = static_cast<clsparseIdx_t>(h.rows);
    csr->num_cols = static_cast<clsparseIdx_t>(h.cols);
    csr->num_nonzeros = static_cast<clsparseIdx_t>(nnz);
    hipStream_t s = control->stream;
    CLS_HIP_TRY(hipMemcpyAsync(dev_ptr<uint32_t>(csr->row_pointer, csr->off_row_pointer), rp.data(),
                               rp.size() * sizeof(uint32_t), hipMemcpyHostToDevice, s));
    if (nnz) {
        CLS_HIP_TRY(hipMemcpyAsync(dev_ptr<T>(csr->values, csr->off_values), v.data(), nnz * sizeof(T),
                                   hipMemcpyHostToDevice, s));
        CLS_HIP_TRY(hipMemcpyAsync(dev_ptr<uint32_t>(csr->col_indices, csr->off_col_indices), c.data(),
                                   nnz * sizeof(uint32_t), hipMemcpyHostToDevice, s));
    }
    CLS_HIP_TRY(hipStreamSynchronize(s));
    return clsparseSuccess;
}

// The C entry points' bodies: host memory running out (or a thread that
// could not finish its part) gives clsparseOutOfHostMemory, never an
// exception through the C ABI.
template <typename F>
clsparseStatus guarded(F&& body)
{
    try {
        return body();
    } catch (const std::bad_alloc&) {
        return clsparseOutOfHostMemory;
    } catch (...) {
        return clsparseInvalidFile;
    }
}

template <typename T>
clsparseStatus coo_from_file(clsparseCooMatrix* coo, const char* path, clsparseControl control,
                             cl_bool explicit_zeroes)
{
    return guarded([&] { return coo_from_file_impl<T>(coo, path, control, explicit_zeroes); });
}

template <typename T>
clsparseStatus csr_from_file(clsparseCsrMatrix* csr, const char* path, clsparseControl control,
                             cl_bool explicit_zeroes)
{
    return guarded([&] { return csr_from_file_impl<T>(csr, path, control, explicit_zeroes); });
}

clsparseStatus header_from_file(clsparseIdx_t* nnz, clsparseIdx_t* row, clsparseIdx_t* col, const char* filePath)
{
    // The header only: read up to the size line.
    FILE* f = std::fopen(filePath, "rb");
    if (!f) return clsparseInvalidFile;
    std::string text;
    char line[4096];
    bool sized = false;
    while (std::fgets(line, sizeof(line), f)) {
        text += line;
        unsigned long long a, b, c;
        if (line[0] != '%' && std::sscanf(line, "%llu %llu %llu", &a, &b, &c) == 3) {
            sized = true;
            break;
        }
    }
    std::fclose(f);
    MmHeader h;
    if (!sized || !parse_header(text, h)) return clsparseInvalidFile;
    const uint64_t n = h.symmetric ? 2 * h.nnz : h.nnz;
    if (n > 0xFFFFFFFFull || h.rows > 0xFFFFFFFFull || h.cols > 0xFFFFFFFFull) return clsparseInvalidFile;
    *nnz = static_cast<clsparseIdx_t>(n);
    *row = static_cast<clsparseIdx_t>(h.rows);
    *col = static_cast<clsparseIdx_t>(h.cols);
    return clsparseSuccess;
}

clsparseStatus csr_from_file_host(const char* filePath, cl_bool read_explicit_zeroes, clsparseAmdHostCsr* out)
{
    MmHeader h;
    std::vector<Entry<double>> e;
    CLS_TRY(read_entries<double>(filePath, read_explicit_zeroes != 0, h, e));
    std::vector<uint32_t> rp;
    row_pointer(e, static_cast<uint32_t>(h.rows), rp);
    const size_t nnz = e.size();
    out->num_rows = static_cast<clsparseIdx_t>(h.rows);
    out->num_cols = static_cast<clsparseIdx_t>(h.cols);
    out->num_nonzeros = static_cast<clsparseIdx_t>(nnz);
    out->row_pointer = static_cast<cl_uint*>(std::malloc(rp.size() * sizeof(cl_uint)));
    out->col_indices = static_cast<cl_uint*>(std::malloc(std::max<size_t>(nnz, 1) * sizeof(cl_uint)));
    out->values = static_cast<cl_double*>(std::malloc(std::max<size_t>(nnz, 1) * sizeof(cl_double)));
    if (!out->row_pointer || !out->col_indices || !out->values) {
        clsparseAmdHostCsrFree(out);
        return clsparseOutOfHostMemory;
    }
    std::memcpy(out->row_pointer, rp.data(), rp.size() * sizeof(cl_uint));
    for (size_t i = 0; i < nnz; ++i) {
        out->col_indices[i] = e[i].c;
        out->values[i] = e[i].v;
    }
    return clsparseSuccess;
}

}  // namespace

// src/library/io/mm-reader.cpp L441-466
extern "C" clsparseStatus clsparseHeaderfromFile(clsparseIdx_t* nnz, clsparseIdx_t* row,
                                                 clsparseIdx_t* col, const char* filePath)
{
    if (nnz == nullptr || row == nullptr || col == nullptr) return clsparseInvalidValue;
    if (!has_mtx_extension(filePath)) return clsparseInvalidFileFormat;
    return guarded([&] { return header_from_file(nnz, row, col, filePath); });
}

// mm-reader.cpp L471-519
extern "C" clsparseStatus clsparseSCooMatrixfromFile(clsparseCooMatrix* cooMatx, const char* filePath,
                                                     clsparseControl control, cl_bool read_explicit_zeroes)
{
    return coo_from_file<float>(cooMatx, filePath, control, read_explicit_zeroes);
}

// mm-reader.cpp L521-567
extern "C" clsparseStatus clsparseDCooMatrixfromFile(clsparseCooMatrix* cooMatx, const char* filePath,
                                                     clsparseControl control, cl_bool read_explicit_zeroes)
{
    return coo_from_file<double>(cooMatx, filePath, control, read_explicit_zeroes);
}

// mm-reader.cpp L569-644
extern "C" clsparseStatus clsparseSCsrMatrixfromFile(clsparseCsrMatrix* csrMatx, const char* filePath,
                                                     clsparseControl control, cl_bool read_explicit_zeroes)
{
    return csr_from_file<float>(csrMatx, filePath, control, read_explicit_zeroes);
}

// mm-reader.cpp L646-736
extern "C" clsparseStatus clsparseDCsrMatrixfromFile(clsparseCsrMatrix* csrMatx, const char* filePath,
                                                     clsparseControl control, cl_bool read_explicit_zeroes)
{
    return csr_from_file<double>(csrMatx, filePath, control, read_explicit_zeroes);
}

// Host-memory variant of the FP64 CSR reader (no device): same parsing,
// expansion, filtering and ordering.
extern "C" clsparseStatus clsparseAmdCsrMatrixfromFileHost(const char* filePath, cl_bool read_explicit_zeroes,
                                                           clsparseAmdHostCsr* out)
{
    if (out == nullptr) return clsparseInvalidValue;
    if (!has_mtx_extension(filePath)) return clsparseInvalidFileFormat;
    return guarded([&] { return csr_from_file_host(filePath, read_explicit_zeroes, out); });
}
