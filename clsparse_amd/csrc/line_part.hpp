// Row partition of the line-order diagonal-storage kernels, and the size
// predicate of their 32-bit buffer resources. Plain C++ (host and device):
// the host tests compile this header with g++ and walk the partition.
#pragma once

#include <cstdint>

#if defined(__HIPCC__)
#define CLS_HD __host__ __device__
#else
#define CLS_HD
#endif

// Rows i of an n-row matrix form `line` lines {c, c + line, c + 2 line, ...}
// (c < line, m = ceil(n / line) rows per line). Each line is cut into `segs`
// stretches of `sl` consecutive line rows; work item q < items is stretch
// q / line of line q % line. A thread of a `threads`-thread grid takes items
// t, t + threads, t + 2 threads, ..., so every row is covered exactly once
// whether the grid holds more threads than a line has columns (several
// stretches per line, at most one item per thread) or fewer (one stretch per
// line, several lines per thread).
struct LinePart
{
    uint32_t line, segs, m, sl, items;
};

CLS_HD inline LinePart line_part(uint32_t n, uint32_t line, uint32_t threads)
{
    LinePart p;
    p.line = line;
    p.segs = threads / line > 1u ? threads / line : 1u;
    p.m = (n + line - 1) / line;
    p.sl = (p.m + p.segs - 1) / p.segs;
    p.items = line * p.segs;  // <= max(line, threads)
    return p;
}

// Item q: line column c and line rows [k_beg, k_end) (empty when the
// stretch starts past the line's end).
CLS_HD inline void line_item(const LinePart& p, uint32_t q, uint32_t& c, uint32_t& k_beg, uint32_t& k_end)
{
    c = q % p.line;
    const uint32_t sg = q / p.line;
    k_beg = sg * p.sl;
    const uint32_t e = (sg + 1) * p.sl;
    k_end = e < p.m ? e : p.m;
}

// A buffer resource over K arrays of n elements of S bytes takes its byte
// range and every offset into it in 32 bits (with 16 bytes of slack for the
// widest load).
CLS_HD inline bool dia_buf_fits(int K, uint64_t n, uint32_t S)
{
    return static_cast<uint64_t>(K) * n * S < (uint64_t(1) << 32) - 16;
}
