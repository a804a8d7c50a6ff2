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
// behaviour; the file is parsed from one buffered read, not per-token fscanf.
#include <algorithm>
#include <cctype>
#include <cerrno>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "common.hpp"

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

bool read_file(const char* path, std::string& out)
{
    FILE* f = std::fopen(path, "rb");
    if (!f) return false;
    char buf[1 << 16];
    size_t got;
    while ((got = std::fread(buf, 1, sizeof(buf), f)) > 0) out.append(buf, got);
    const bool ok = !std::ferror(f);
    std::fclose(f);
    return ok;
}

std::string lower(std::string s)
{
    for (char& c : s) c = static_cast<char>(std::tolower(static_cast<unsigned char>(c)));
    return s;
}

// Banner "%%MatrixMarket matrix coordinate <data> <storage>", comment lines,
// then the size line "rows cols nnz" (blank lines before it are skipped).
bool parse_header(const std::string& text, MmHeader& h)
{
    size_t pos = text.find('\n');
    const std::string first = text.substr(0, pos);
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
        const std::string line = text.substr(start, pos == std::string::npos ? std::string::npos : pos - start);
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

template <typename T>
T parse_value(const char* p, char** end);
template <>
float parse_value<float>(const char* p, char** end)
{
    return std::strtof(p, end);
}
template <>
double parse_value<double>(const char* p, char** end)
{
    return std::strtod(p, end);
}

// Entries (0-based, symmetric expansion, zero filtering), sorted by (row, col).
template <typename T>
clsparseStatus read_entries(const char* path, bool explicit_zeroes, MmHeader& h, std::vector<Entry<T>>& out)
{
    std::string text;
    if (!read_file(path, text)) return clsparseInvalidFile;
    if (!parse_header(text, h)) return clsparseInvalidFile;
    if (h.rows > 0xFFFFFFFFull || h.cols > 0xFFFFFFFFull || h.nnz > 0xFFFFFFFFull)
        return clsparseInvalidFile;
    out.clear();
    out.reserve(h.symmetric ? 2 * h.nnz : h.nnz);
    const char* p = text.c_str() + h.body;
    for (uint64_t i = 0; i < h.nnz; ++i) {
        char* e = nullptr;
        errno = 0;
        const unsigned long long ir = std::strtoull(p, &e, 10);
        if (e == p || errno) return clsparseInvalidFile;
        p = e;
        const unsigned long long ic = std::strtoull(p, &e, 10);
        if (e == p || errno) return clsparseInvalidFile;
        p = e;
        T v;
        if (h.data == MmData::Pattern) {
            v = static_cast<T>(kMaxRandVal * (std::rand() / (RAND_MAX + 1.0)));
        } else {
            v = parse_value<T>(p, &e);
            if (e == p) return clsparseInvalidFile;
            p = e;
        }
        if (ir < 1 || ic < 1 || ir > h.rows || ic > h.cols) return clsparseInvalidFile;
        if (!explicit_zeroes && v == T(0)) continue;
        const uint32_t r = static_cast<uint32_t>(ir - 1), c = static_cast<uint32_t>(ic - 1);
        out.push_back({r, c, v});
        if (h.symmetric && r != c) out.push_back({c, r, v});
    }
    std::stable_sort(out.begin(), out.end(), [](const Entry<T>& a, const Entry<T>& b) {
        return a.r != b.r ? a.r < b.r : a.c < b.c;
    });
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
clsparseStatus coo_from_file(clsparseCooMatrix* coo, const char* path, clsparseControl control,
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
clsparseStatus csr_from_file(clsparseCsrMatrix* csr, const char* path, clsparseControl control,
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

}  // namespace

// src/library/io/mm-reader.cpp L441-466
extern "C" clsparseStatus clsparseHeaderfromFile(clsparseIdx_t* nnz, clsparseIdx_t* row,
                                                 clsparseIdx_t* col, const char* filePath)
{
    if (nnz == nullptr || row == nullptr || col == nullptr) return clsparseInvalidValue;
    if (!has_mtx_extension(filePath)) return clsparseInvalidFileFormat;
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
