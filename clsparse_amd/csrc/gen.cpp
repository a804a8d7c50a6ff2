// Seeded synthetic matrices and vectors (host memory).
//
// Every random draw comes from a counter-based stream: SplitMix64 applied to
// (seed, row or index, draw number), so results do not depend on how many
// threads fill the arrays. Rows are generated in parallel with std::thread.
#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <thread>
#include <new>
#include <vector>

#include "common.hpp"
#include "host_threads.hpp"

namespace {

inline uint64_t splitmix64(uint64_t x)
{
    x += 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
}

// Sequential stream for one (seed, key) pair.
struct Rng
{
    uint64_t state;
    Rng(uint64_t seed, uint64_t key) : state(splitmix64(seed ^ splitmix64(key + 0x1234567ull))) {}
    uint64_t next()
    {
        state += 0x9E3779B97F4A7C15ull;
        uint64_t z = state;
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    }
    double uniform() { return static_cast<double>(next() >> 11) * 0x1.0p-53; }  // [0, 1)
    double normal()
    {
        // Box-Muller (first variate only).
        double u1 = uniform();
        const double u2 = uniform();
        if (u1 < 1e-300) u1 = 1e-300;
        return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
    }
};

unsigned num_threads()
{
    unsigned t = std::thread::hardware_concurrency();
    if (const char* e = std::getenv("OMP_NUM_THREADS")) {
        const int v = std::atoi(e);
        if (v > 0) t = static_cast<unsigned>(v);
    }
    if (t == 0) t = 1;
    if (t > 64) t = 64;
    return t;
}

// f(i) for i = 0 .. n-1 in t contiguous chunks, one thread each (f must not
// throw: it only fills preallocated arrays).
template <typename F>
void parallel_rows(uint64_t n, F&& f)
{
    const unsigned t = num_threads();
    const uint64_t chunk = (n + t - 1) / t;
    run_on_threads(t, [&](unsigned k) {
        const uint64_t lo = k * chunk;
        const uint64_t hi = std::min<uint64_t>(n, lo + chunk);
        for (uint64_t i = lo; i < hi; ++i) f(i);
    });
}

bool alloc_csr(clsparseAmdHostCsr* out, uint32_t nrows, uint32_t ncols, uint64_t nnz)
{
    if (nnz > 0xFFFFFFFFull) return false;
    out->num_rows = nrows;
    out->num_cols = ncols;
    out->num_nonzeros = static_cast<uint32_t>(nnz);
    out->row_pointer = static_cast<cl_uint*>(std::malloc((static_cast<size_t>(nrows) + 1) * sizeof(cl_uint)));
    out->col_indices = static_cast<cl_uint*>(std::malloc(std::max<uint64_t>(nnz, 1) * sizeof(cl_uint)));
    out->values = static_cast<cl_double*>(std::malloc(std::max<uint64_t>(nnz, 1) * sizeof(cl_double)));
    if (!out->row_pointer || !out->col_indices || !out->values) {
        clsparseAmdHostCsrFree(out);
        return false;
    }
    return true;
}

}  // namespace

extern "C" clsparseStatus clsparseAmdHostCsrFree(clsparseAmdHostCsr* m)
{
    if (m == nullptr) return clsparseInvalidValue;
    std::free(m->row_pointer);
    std::free(m->col_indices);
    std::free(m->values);
    m->row_pointer = nullptr;
    m->col_indices = nullptr;
    m->values = nullptr;
    m->num_rows = m->num_cols = m->num_nonzeros = 0;
    return clsparseSuccess;
}

extern "C" clsparseStatus clsparseAmdGenUniform(cl_double* out, size_t n, cl_double lo,
                                                cl_double hi, cl_ulong seed)
{
    if (out == nullptr && n > 0) return clsparseInvalidValue;
    parallel_rows(n, [&](uint64_t i) {
        const double u = static_cast<double>(splitmix64(seed * 0x100000001B3ull ^ splitmix64(i)) >> 11) *
                         0x1.0p-53;
        out[i] = lo + (hi - lo) * u;
    });
    return clsparseSuccess;
}

extern "C" clsparseStatus clsparseAmdGenPoisson2D(cl_uint nx, cl_uint ny, clsparseAmdHostCsr* out)
{
    if (out == nullptr || nx == 0 || ny == 0) return clsparseInvalidValue;
    const uint64_t n = static_cast<uint64_t>(nx) * ny;
    if (n > 0xFFFFFFFFull) return clsparseInvalidSize;
    const uint64_t nnz = 5 * n - 2ull * nx - 2ull * ny;
    if (!alloc_csr(out, static_cast<uint32_t>(n), static_cast<uint32_t>(n), nnz))
        return clsparseOutOfHostMemory;
    // Row pointer: interior rows have 5 entries, minus one per missing neighbour.
    uint64_t k = 0;
    out->row_pointer[0] = 0;
    for (uint32_t iy = 0; iy < ny; ++iy)
        for (uint32_t ix = 0; ix < nx; ++ix) {
            k += 5 - (iy == 0) - (ix == 0) - (ix + 1 == nx) - (iy + 1 == ny);
            out->row_pointer[static_cast<uint64_t>(iy) * nx + ix + 1] = static_cast<uint32_t>(k);
        }
    parallel_rows(n, [&](uint64_t i) {
        const uint32_t ix = static_cast<uint32_t>(i % nx), iy = static_cast<uint32_t>(i / nx);
        uint64_t j = out->row_pointer[i];
        auto put = [&](uint64_t c, double v) {
            out->col_indices[j] = static_cast<uint32_t>(c);
            out->values[j] = v;
            ++j;
        };
        if (iy > 0) put(i - nx, -1.0);
        if (ix > 0) put(i - 1, -1.0);
        put(i, 4.0);
        if (ix + 1 < nx) put(i + 1, -1.0);
        if (iy + 1 < ny) put(i + nx, -1.0);
    });
    return clsparseSuccess;
}

static clsparseStatus gen_power_law(cl_uint n, cl_double mean_len, cl_double sigma, cl_uint max_len,
                                    cl_uint n_dense, cl_uint dense_len, cl_ulong seed, clsparseAmdHostCsr* out)
{
    if (out == nullptr || n == 0 || mean_len <= 0 || max_len == 0) return clsparseInvalidValue;
    const uint32_t cap = std::min<uint32_t>(max_len, n);
    std::vector<uint32_t> len(n);
    // Lengths: LogNormal with E[len] = mean_len before rounding and clipping.
    const double mu = std::log(mean_len) - 0.5 * sigma * sigma;
    parallel_rows(n, [&](uint64_t i) {
        Rng r(seed, 2 * i);
        const double v = std::exp(mu + sigma * r.normal());
        double l = std::nearbyint(v);
        if (!(l >= 1.0)) l = 1.0;
        if (l > cap) l = cap;
        len[i] = static_cast<uint32_t>(l);
    });
    // Dense rows at distinct seeded positions.
    {
        Rng r(seed ^ 0xD1B54A32D192ED03ull, 0);
        std::vector<uint32_t> picked;
        while (picked.size() < std::min<uint32_t>(n_dense, n)) {
            const uint32_t p = static_cast<uint32_t>(r.next() % n);
            if (std::find(picked.begin(), picked.end(), p) == picked.end()) picked.push_back(p);
        }
        for (uint32_t p : picked) len[p] = std::min<uint32_t>(dense_len, n);
    }
    uint64_t nnz = 0;
    for (uint32_t i = 0; i < n; ++i) nnz += len[i];
    if (!alloc_csr(out, n, n, nnz)) return clsparseOutOfHostMemory;
    out->row_pointer[0] = 0;
    {
        uint64_t k = 0;
        for (uint32_t i = 0; i < n; ++i) {
            k += len[i];
            out->row_pointer[i + 1] = static_cast<uint32_t>(k);
        }
    }
    // Columns: uniform without replacement, sorted; values U[-1, 1].
    parallel_rows(n, [&](uint64_t i) {
        Rng r(seed, 2 * i + 1);
        const uint32_t L = len[i];
        uint32_t* c = out->col_indices + out->row_pointer[i];
        double* v = out->values + out->row_pointer[i];
        if (static_cast<uint64_t>(L) * 2 >= n) {
            // Dense-ish row: selection sampling over all columns (ordered).
            uint32_t need = L, k = 0;
            for (uint32_t col = 0; col < n && need > 0; ++col) {
                if (r.uniform() * (n - col) < need) {
                    c[k++] = col;
                    --need;
                }
            }
        } else {
            uint32_t have = 0;
            while (have < L) {
                for (uint32_t k = have; k < L; ++k) c[k] = static_cast<uint32_t>(r.next() % n);
                std::sort(c, c + L);
                have = static_cast<uint32_t>(std::unique(c, c + L) - c);
            }
        }
        for (uint32_t k = 0; k < L; ++k) v[k] = 2.0 * r.uniform() - 1.0;
    });
    return clsparseSuccess;
}

extern "C" clsparseStatus clsparseAmdGenPowerLaw(cl_uint n, cl_double mean_len, cl_double sigma,
                                                 cl_uint max_len, cl_uint n_dense,
                                                 cl_uint dense_len, cl_ulong seed,
                                                 clsparseAmdHostCsr* out)
{
    try {
        return gen_power_law(n, mean_len, sigma, max_len, n_dense, dense_len, seed, out);
    } catch (const std::bad_alloc&) {
        return clsparseOutOfHostMemory;  // out is untouched: its arrays come after every throwing allocation
    }
}

// Q1 (bilinear) element stiffness on the unit square for a constant tensor
// K = [[k11, k12], [k12, k22]]; local nodes (0,0), (1,0), (0,1), (1,1).
// Exact integrals of grad(phi_a)^T K grad(phi_b): SPSD by construction.
static void q1_stiffness(double k11, double k12, double k22, double ke[4][4])
{
    static const double Sxx[4][4] = {{1 / 3., -1 / 3., 1 / 6., -1 / 6.},
                                     {-1 / 3., 1 / 3., -1 / 6., 1 / 6.},
                                     {1 / 6., -1 / 6., 1 / 3., -1 / 3.},
                                     {-1 / 6., 1 / 6., -1 / 3., 1 / 3.}};
    static const double Syy[4][4] = {{1 / 3., 1 / 6., -1 / 3., -1 / 6.},
                                     {1 / 6., 1 / 3., -1 / 6., -1 / 3.},
                                     {-1 / 3., -1 / 6., 1 / 3., 1 / 6.},
                                     {-1 / 6., -1 / 3., 1 / 6., 1 / 3.}};
    // d(phi_a)/dx depends on y only and d(phi_b)/dy on x only, so
    // int dx(phi_a) dy(phi_b) = (int dx phi_a dy) (int dy phi_b dx).
    static const double cx[4] = {-0.5, 0.5, -0.5, 0.5};
    static const double cy[4] = {-0.5, -0.5, 0.5, 0.5};
    for (int a = 0; a < 4; ++a)
        for (int b = 0; b < 4; ++b)
            ke[a][b] = k11 * Sxx[a][b] + k22 * Syy[a][b] + k12 * (cx[a] * cy[b] + cx[b] * cy[a]);
}

static clsparseStatus gen_aniso_diffusion2d(cl_uint nx, cl_uint ny, cl_double corr_len, cl_double sigma,
                                            cl_double ratio, cl_double theta, cl_ulong seed,
                                            clsparseAmdHostCsr* out)
{
    if (out == nullptr || nx == 0 || ny == 0 || ratio <= 0 || corr_len <= 0)
        return clsparseInvalidValue;
    const uint64_t n = static_cast<uint64_t>(nx) * ny;
    if (n > 0xFFFFFFFFull) return clsparseInvalidSize;
    // Log-coefficient: Gaussian random field with covariance
    // sigma^2 exp(-r^2 / (2 corr_len^2)), sampled by M random Fourier modes.
    constexpr int M = 128;
    double kxm[M], kym[M], phm[M];
    {
        Rng r(seed, 0xF00D);
        for (int m = 0; m < M; ++m) {
            kxm[m] = r.normal() / corr_len;
            kym[m] = r.normal() / corr_len;
            phm[m] = 6.283185307179586 * r.uniform();
        }
    }
    const double amp = sigma * std::sqrt(2.0 / M);
    const uint32_t ex = nx + 1, ey = ny + 1;  // elements; nodes 0 and n+1 are Dirichlet
    std::vector<double> kappa(static_cast<size_t>(ex) * ey);
    parallel_rows(static_cast<uint64_t>(ex) * ey, [&](uint64_t e) {
        const double cxp = static_cast<double>(e % ex) + 0.5, cyp = static_cast<double>(e / ex) + 0.5;
        double s = 0.0;
        for (int m = 0; m < M; ++m) s += std::cos(kxm[m] * cxp + kym[m] * cyp + phm[m]);
        kappa[e] = std::exp(amp * s);
    });
    // Tensor R(theta) diag(1, 1/ratio) R(theta)^T.
    const double c = std::cos(theta), sn = std::sin(theta), eps = 1.0 / ratio;
    const double t11 = c * c + eps * sn * sn, t22 = sn * sn + eps * c * c, t12 = (1.0 - eps) * c * sn;

    // Row pointer: interior node couples with its 3x3 neighbourhood.
    uint64_t nnz = 0;
    std::vector<uint32_t> rlen(n);
    for (uint64_t i = 0; i < n; ++i) {
        const uint32_t ix = static_cast<uint32_t>(i % nx), iy = static_cast<uint32_t>(i / nx);
        const uint32_t wx = 1 + (ix > 0) + (ix + 1 < nx), wy = 1 + (iy > 0) + (iy + 1 < ny);
        rlen[i] = wx * wy;
        nnz += rlen[i];
    }
    if (!alloc_csr(out, static_cast<uint32_t>(n), static_cast<uint32_t>(n), nnz))
        return clsparseOutOfHostMemory;
    out->row_pointer[0] = 0;
    for (uint64_t i = 0; i < n; ++i) out->row_pointer[i + 1] = out->row_pointer[i] + rlen[i];
    parallel_rows(n, [&](uint64_t i) {
        const int ix = static_cast<int>(i % nx), iy = static_cast<int>(i / nx);
        double acc[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};  // [dy+1][dx+1]
        // Node (ix+1, iy+1) in node numbering; the 4 elements around it.
        for (int ey0 = 0; ey0 < 2; ++ey0)
            for (int ex0 = 0; ex0 < 2; ++ex0) {
                const int ecx = ix + ex0, ecy = iy + ey0;  // element index
                const double k = kappa[static_cast<size_t>(ecy) * ex + ecx];
                double ke[4][4];
                q1_stiffness(k * t11, k * t12, k * t22, ke);
                const int me = (1 - ex0) + 2 * (1 - ey0);  // local index of this node
                for (int b = 0; b < 4; ++b) {
                    const int dx = (b & 1) - (1 - ex0), dy = (b >> 1) - (1 - ey0);
                    acc[dy + 1][dx + 1] += ke[me][b];
                }
            }
        uint64_t j = out->row_pointer[i];
        for (int dy = -1; dy <= 1; ++dy)
            for (int dx = -1; dx <= 1; ++dx) {
                const int jx = ix + dx, jy = iy + dy;
                if (jx < 0 || jy < 0 || jx >= static_cast<int>(nx) || jy >= static_cast<int>(ny))
                    continue;
                out->col_indices[j] = static_cast<uint32_t>(jy) * nx + static_cast<uint32_t>(jx);
                out->values[j] = acc[dy + 1][dx + 1];
                ++j;
            }
    });
    return clsparseSuccess;
}

extern "C" clsparseStatus clsparseAmdGenAnisoDiffusion2D(cl_uint nx, cl_uint ny, cl_double corr_len,
                                                         cl_double sigma, cl_double ratio,
                                                         cl_double theta, cl_ulong seed,
                                                         clsparseAmdHostCsr* out)
{
    try {
        return gen_aniso_diffusion2d(nx, ny, corr_len, sigma, ratio, theta, seed, out);
    } catch (const std::bad_alloc&) {
        return clsparseOutOfHostMemory;  // out is untouched: its arrays come after every throwing allocation
    }
}

static clsparseStatus gen_diffusion3d(cl_uint nx, cl_uint ny, cl_uint nz, cl_double sigma, cl_double ax,
                                      cl_double ay, cl_double az, cl_ulong seed, clsparseAmdHostCsr* out)
{
    if (out == nullptr || nx == 0 || ny == 0 || nz == 0) return clsparseInvalidValue;
    const uint64_t n = static_cast<uint64_t>(nx) * ny * nz;
    if (n > 0xFFFFFFFFull) return clsparseInvalidSize;
    const uint64_t nnz = 7 * n - 2 * (static_cast<uint64_t>(nx) * ny + static_cast<uint64_t>(ny) * nz +
                                      static_cast<uint64_t>(nx) * nz);
    std::vector<double> kappa(n);
    parallel_rows(n, [&](uint64_t i) {
        Rng r(seed, i);
        kappa[i] = std::exp(sigma * r.normal());
    });
    if (!alloc_csr(out, static_cast<uint32_t>(n), static_cast<uint32_t>(n), nnz))
        return clsparseOutOfHostMemory;
    const uint64_t sxy = static_cast<uint64_t>(nx) * ny;
    // Row pointer from the neighbour count of each cell.
    {
        uint64_t k = 0;
        out->row_pointer[0] = 0;
        for (uint64_t i = 0; i < n; ++i) {
            const uint32_t ix = static_cast<uint32_t>(i % nx);
            const uint32_t iy = static_cast<uint32_t>((i / nx) % ny);
            const uint32_t iz = static_cast<uint32_t>(i / sxy);
            k += 7 - (ix == 0) - (ix + 1 == nx) - (iy == 0) - (iy + 1 == ny) - (iz == 0) -
                 (iz + 1 == nz);
            out->row_pointer[i + 1] = static_cast<uint32_t>(k);
        }
    }
    parallel_rows(n, [&](uint64_t i) {
        const uint32_t ix = static_cast<uint32_t>(i % nx);
        const uint32_t iy = static_cast<uint32_t>((i / nx) % ny);
        const uint32_t iz = static_cast<uint32_t>(i / sxy);
        const double ki = kappa[i];
        // Harmonic mean of the two cells, symmetric in (i, j) bit for bit.
        auto face = [&](uint64_t j, double a) { return a * 2.0 * (ki * kappa[j]) / (ki + kappa[j]); };
        double diag = 0.0;
        double off[6];
        uint64_t col[6];
        int m = 0;
        // Neighbours in increasing column order: -z, -y, -x, +x, +y, +z.
        if (iz > 0) { col[m] = i - sxy; off[m++] = face(i - sxy, az); } else diag += az * ki;
        if (iy > 0) { col[m] = i - nx; off[m++] = face(i - nx, ay); } else diag += ay * ki;
        if (ix > 0) { col[m] = i - 1; off[m++] = face(i - 1, ax); } else diag += ax * ki;
        const int lower = m;
        if (ix + 1 < nx) { col[m] = i + 1; off[m++] = face(i + 1, ax); } else diag += ax * ki;
        if (iy + 1 < ny) { col[m] = i + nx; off[m++] = face(i + nx, ay); } else diag += ay * ki;
        if (iz + 1 < nz) { col[m] = i + sxy; off[m++] = face(i + sxy, az); } else diag += az * ki;
        for (int q = 0; q < m; ++q) diag += off[q];
        uint64_t j = out->row_pointer[i];
        for (int q = 0; q < lower; ++q) {
            out->col_indices[j] = static_cast<uint32_t>(col[q]);
            out->values[j++] = -off[q];
        }
        out->col_indices[j] = static_cast<uint32_t>(i);
        out->values[j++] = diag;
        for (int q = lower; q < m; ++q) {
            out->col_indices[j] = static_cast<uint32_t>(col[q]);
            out->values[j++] = -off[q];
        }
    });
    return clsparseSuccess;
}

extern "C" clsparseStatus clsparseAmdGenDiffusion3D(cl_uint nx, cl_uint ny, cl_uint nz,
                                                    cl_double sigma, cl_double ax, cl_double ay,
                                                    cl_double az, cl_ulong seed,
                                                    clsparseAmdHostCsr* out)
{
    try {
        return gen_diffusion3d(nx, ny, nz, sigma, ax, ay, az, seed, out);
    } catch (const std::bad_alloc&) {
        return clsparseOutOfHostMemory;  // out is untouched: its arrays come after every throwing allocation
    }
}
