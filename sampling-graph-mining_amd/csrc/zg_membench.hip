// zg_membench.hip -- random-gather bandwidth microbenchmark.  Measures the
// rate at which this GPU serves independent random 8-byte reads from a buffer
// much larger than the last-level cache: the achievable random-access HBM
// bandwidth that bounds the estimator kernels (each of their reads is a
// random 8..32-byte access inside one 128-byte L2 line).  Every lane runs
// `chains` independent dependent-load chains (the next address hashes the
// loaded word), the shape of an estimator's read sequence.  Measurement
// helper only; no reference counterpart.
#include <algorithm>

#include "zg_internal.h"

namespace zg {
namespace {

__global__ void k_mb_fill(uint64_t *buf, uint64_t n)
{
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        buf[i] = hmix(i + 0x51ED27ull);
}

template <int CH>
__global__ __launch_bounds__(256) void k_mb_gather(const uint64_t *buf, uint64_t n, uint32_t rounds, uint64_t *sink)
{
    const uint64_t gid = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    uint64_t x[CH];
#pragma unroll
    for (int j = 0; j < CH; ++j) x[j] = hmix(gid * CH + j + 1);
    for (uint32_t r = 0; r < rounds; ++r) {
#pragma unroll
        for (int j = 0; j < CH; ++j) {
            const uint64_t idx = __umul64hi(x[j], n);   // uniform in [0, n)
            x[j] = hmix(x[j] ^ buf[idx]);
        }
    }
    uint64_t acc = 0;
#pragma unroll
    for (int j = 0; j < CH; ++j) acc ^= x[j];
    if (acc == 0x0123456789ABCDEFull) sink[0] = acc;   // keeps the loads live
}

// L loads of 16 bytes per lane from its own random 128-byte line
template <int L>
__global__ __launch_bounds__(256) void k_mb_lines(const uint4 *buf, uint64_t nlines, uint32_t rounds, uint64_t *sink)
{
    uint64_t x = hmix((uint64_t)blockIdx.x * blockDim.x + threadIdx.x + 1);
    for (uint32_t r = 0; r < rounds; ++r) {
        const uint4 *p = buf + __umul64hi(x, nlines) * 8;
        uint4 v[L];
#pragma unroll
        for (int i = 0; i < L; ++i) v[i] = p[i];
        uint64_t a = 0;
#pragma unroll
        for (int i = 0; i < L; ++i) a ^= ((uint64_t)v[i].x << 32 | v[i].y) ^ v[i].z ^ ((uint64_t)v[i].w << 17);
        x = hmix(x ^ a);
    }
    if (x == 0x0123456789ABCDEFull) sink[0] = x;
}

// groups of 8 lanes read one random line, one 16-byte piece each; the group
// combines the pieces so the next line depends on the whole line
__global__ __launch_bounds__(256) void k_mb_lines_coop(const uint4 *buf, uint64_t nlines, uint32_t rounds,
                                                       uint64_t *sink)
{
    const uint32_t piece = threadIdx.x & 7;
    uint64_t x = hmix(((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) / 8 + 1);
    for (uint32_t r = 0; r < rounds; ++r) {
        const uint4 v = buf[__umul64hi(x, nlines) * 8 + piece];
        uint64_t a = ((uint64_t)v.x << 32 | v.y) ^ v.z ^ ((uint64_t)v.w << 17);
        a ^= __shfl_xor(a, 1);
        a ^= __shfl_xor(a, 2);
        a ^= __shfl_xor(a, 4);
        x = hmix(x ^ a);
    }
    if (x == 0x0123456789ABCDEFull) sink[0] = x;
}

// One random 16-byte read per 128-byte line, each address from the previous
// line, confined to regions of region_lines lines: workgroup b reads region
// (sel + 8 * phase) mod nreg, phase = round / rounds_per_phase, with sel =
// b mod 8 (XCD_MATCHED: the workgroups of one XCD -- consecutive workgroups go
// to the eight XCDs in turn -- share a region, so it can live in that XCD's
// L2) or sel = (b / 8) mod 8 (the workgroups of one XCD spread over eight
// regions).  Shows what the L2, the memory-side cache and the DRAM each serve
// for random lines.
template <bool XCD_MATCHED>
__global__ __launch_bounds__(256) void k_mb_region(const uint4 *buf, uint64_t region_lines, uint64_t nreg,
                                                   uint32_t rounds, uint32_t rounds_per_phase, uint64_t *sink)
{
    uint64_t x = hmix((uint64_t)blockIdx.x * blockDim.x + threadIdx.x + 1);
    const uint32_t sel = XCD_MATCHED ? (blockIdx.x & 7u) : ((blockIdx.x >> 3) & 7u);
    for (uint32_t r = 0; r < rounds; ++r) {
        const uint64_t reg = (sel + 8ull * (r / rounds_per_phase)) % nreg;
        const uint4 v = buf[(reg * region_lines + __umul64hi(x, region_lines)) * 8];
        x = hmix(x ^ (((uint64_t)v.x << 32) | v.y) ^ v.z);
    }
    if (x == 0x0123456789ABCDEFull) sink[0] = x;
}

// One estimator's access chain without the estimator: a random 32-byte record
// (two 16-byte loads), then an 8-byte entry of a second array at an index
// taken from the record, then the 16-byte header of a 128-byte bucket of a
// third array at an index taken from the entry -- the three dependent random
// reads of a triangle estimator (edge record, sampled list entry, hash
// bucket), over arrays of the graph's sizes.  Grid-stride over estimator
// indices, as the estimator kernels are.
__global__ __launch_bounds__(256) void k_mb_chain(const uint4 *__restrict__ rec, uint64_t nrec,
                                                  const uint64_t *__restrict__ ent, uint64_t nent,
                                                  const uint4 *__restrict__ bkt, uint64_t nbkt, uint64_t count,
                                                  uint64_t *sink)
{
    uint64_t acc = 0;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < count;
         i += (uint64_t)gridDim.x * blockDim.x) {
        const uint64_t x = hmix(i + 1);
        const uint4 *r = rec + 2 * __umul64hi(x, nrec);
        const uint4 a = r[0], b = r[1];
        const uint64_t e = ent[__umul64hi(hmix(x ^ a.x ^ ((uint64_t)b.w << 32)), nent)];
        const uint4 h = bkt[8 * __umul64hi(hmix(x ^ e), nbkt)];
        // a full 64-bit fold: the sentinel compare below must stay possible, or
        // the compiler proves it false and drops every load
        // (every field is used, so the record and header are read with 16-byte loads, as the
        // estimator kernels read them)
        acc += (((uint64_t)h.w << 32) | h.x) ^ (((uint64_t)h.y << 32) | h.z) ^ e ^
               (((uint64_t)a.w << 32) | a.z) ^ (((uint64_t)b.x << 32) | b.y) ^ (((uint64_t)b.w << 32) | a.y) ^ b.z;
    }
    if (acc == 0x0123456789ABCDEFull) sink[0] = acc;   // keeps the loads live
}

}  // namespace
}  // namespace zg

using namespace zg;

extern "C" int zg_random_chain_bw(uint64_t m, uint64_t estimators, int block_size, int blocks_per_cu, double *gbs,
                                  double *ms)
{
    int rc = ensure_init();
    if (rc) return rc;
    if (!gbs || m < 64 || estimators == 0) return set_error(ZG_EINVAL, "bad chain benchmark arguments");
    const int bs = block_size > 0 ? block_size : 256;
    if (bs > 256 || bs % 64) return set_error(ZG_EINVAL, "block size must be a multiple of 64 up to 256");
    const int bpc = blocks_per_cu > 0 ? blocks_per_cu : 16;
    const uint64_t grid = std::min<uint64_t>((uint64_t)device_cus() * bpc, (estimators + bs - 1) / bs);
    // the graph's arrays: m 32-byte records, 2m 8-byte entries, m/4 128-byte buckets
    const uint64_t nrec = m, nent = 2 * m, nbkt = (m + 3) / 4;
    uint4 *rec = nullptr, *bkt = nullptr;
    uint64_t *ent = nullptr, *sink = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    auto run = [&]() -> int {
        if (hipMalloc(&rec, 32 * nrec) != hipSuccess || hipMalloc(&ent, 8 * nent) != hipSuccess ||
            hipMalloc(&bkt, 128 * nbkt) != hipSuccess)
            return set_error(ZG_ENOMEM, "chain benchmark buffers: allocation failed");
        ZG_HIP(hipMalloc(&sink, 8));
        ZG_HIP(hipEventCreate(&e0));
        ZG_HIP(hipEventCreate(&e1));
        hipLaunchKernelGGL(k_mb_fill, dim3(device_cus() * 64), dim3(256), 0, 0, reinterpret_cast<uint64_t *>(rec),
                           4 * nrec);
        hipLaunchKernelGGL(k_mb_fill, dim3(device_cus() * 64), dim3(256), 0, 0, ent, nent);
        hipLaunchKernelGGL(k_mb_fill, dim3(device_cus() * 64), dim3(256), 0, 0, reinterpret_cast<uint64_t *>(bkt),
                           16 * nbkt);
        ZG_HIP(hipGetLastError());
        auto launch = [&]() {
            hipLaunchKernelGGL(k_mb_chain, dim3(grid), dim3(bs), 0, 0, rec, nrec, ent, nent, bkt, nbkt, estimators,
                               sink);
        };
        launch();   // warm-up
        ZG_HIP(hipGetLastError());
        ZG_HIP(hipEventRecord(e0, 0));
        launch();
        ZG_HIP(hipGetLastError());
        ZG_HIP(hipEventRecord(e1, 0));
        ZG_HIP(hipEventSynchronize(e1));
        float t = 0;
        ZG_HIP(hipEventElapsedTime(&t, e0, e1));
        *gbs = (double)estimators * 3 * 128.0 / (t * 1e-3) / 1e9;
        if (ms) *ms = t;
        return ZG_OK;
    };
    rc = run();
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (sink) (void)hipFree(sink);
    if (rec) (void)hipFree(rec);
    if (ent) (void)hipFree(ent);
    if (bkt) (void)hipFree(bkt);
    return rc;
}

extern "C" int zg_random_gather_bw(uint64_t buffer_bytes, int chains, uint64_t loads, int block_size,
                                   int blocks_per_cu, double *gbs, double *ms)
{
    int rc = ensure_init();
    if (rc) return rc;
    if (!gbs || buffer_bytes < 4096 || loads == 0) return set_error(ZG_EINVAL, "bad gather benchmark arguments");
    if (chains != 1 && chains != 2 && chains != 4) return set_error(ZG_EINVAL, "chains must be 1, 2 or 4");
    const int bs = block_size > 0 ? block_size : 256;
    if (bs > 256 || bs % 64) return set_error(ZG_EINVAL, "block size must be a multiple of 64 up to 256");
    const int bpc = blocks_per_cu > 0 ? blocks_per_cu : 16;
    const uint64_t grid = (uint64_t)device_cus() * bpc;
    const uint64_t lanes = grid * bs;
    const uint64_t per_lane = (loads + lanes * chains - 1) / (lanes * chains);
    if (per_lane > 0xffffffffull) return set_error(ZG_EOVERFLOW, "too many loads");
    const uint32_t rounds = (uint32_t)per_lane;
    const uint64_t n = buffer_bytes / 8;
    uint64_t *buf = nullptr, *sink = nullptr;
    if (hipMalloc(&buf, n * 8) != hipSuccess) return set_error(ZG_ENOMEM, "gather buffer allocation failed");
    hipEvent_t e0 = nullptr, e1 = nullptr;
    auto run = [&]() -> int {
        ZG_HIP(hipMalloc(&sink, 8));
        ZG_HIP(hipEventCreate(&e0));
        ZG_HIP(hipEventCreate(&e1));
        hipLaunchKernelGGL(k_mb_fill, dim3(device_cus() * 64), dim3(256), 0, 0, buf, n);
        ZG_HIP(hipGetLastError());
        auto launch = [&]() {
            if (chains == 1) hipLaunchKernelGGL(k_mb_gather<1>, dim3(grid), dim3(bs), 0, 0, buf, n, rounds, sink);
            else if (chains == 2) hipLaunchKernelGGL(k_mb_gather<2>, dim3(grid), dim3(bs), 0, 0, buf, n, rounds, sink);
            else hipLaunchKernelGGL(k_mb_gather<4>, dim3(grid), dim3(bs), 0, 0, buf, n, rounds, sink);
        };
        launch();   // warm-up
        ZG_HIP(hipGetLastError());
        ZG_HIP(hipEventRecord(e0, 0));
        launch();
        ZG_HIP(hipGetLastError());
        ZG_HIP(hipEventRecord(e1, 0));
        ZG_HIP(hipEventSynchronize(e1));
        float t = 0;
        ZG_HIP(hipEventElapsedTime(&t, e0, e1));
        const double done = (double)lanes * chains * rounds;
        *gbs = done * 128.0 / (t * 1e-3) / 1e9;
        if (ms) *ms = t;
        return ZG_OK;
    };
    rc = run();
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (sink) (void)hipFree(sink);
    (void)hipFree(buf);
    return rc;
}

extern "C" int zg_random_line_bw(uint64_t buffer_bytes, int loads_per_line, int cooperative, uint64_t lines,
                                 int block_size, int blocks_per_cu, double *gbs, double *ms)
{
    int rc = ensure_init();
    if (rc) return rc;
    if (!gbs || buffer_bytes < 4096 || lines == 0) return set_error(ZG_EINVAL, "bad line benchmark arguments");
    if (!cooperative && (loads_per_line < 1 || loads_per_line > 8))
        return set_error(ZG_EINVAL, "loads_per_line must be in 1..8");
    const int bs = block_size > 0 ? block_size : 256;
    if (bs > 256 || bs % 64) return set_error(ZG_EINVAL, "block size must be a multiple of 64 up to 256");
    const int bpc = blocks_per_cu > 0 ? blocks_per_cu : 16;
    const uint64_t grid = (uint64_t)device_cus() * bpc;
    const uint64_t readers = cooperative ? grid * bs / 8 : grid * bs;
    const uint64_t per_reader = (lines + readers - 1) / readers;
    if (per_reader > 0xffffffffull) return set_error(ZG_EOVERFLOW, "too many lines");
    const uint32_t rounds = (uint32_t)per_reader;
    const uint64_t nlines = buffer_bytes / 128;
    uint4 *buf = nullptr;
    uint64_t *sink = nullptr;
    if (hipMalloc(&buf, nlines * 128) != hipSuccess) return set_error(ZG_ENOMEM, "line buffer allocation failed");
    hipEvent_t e0 = nullptr, e1 = nullptr;
    auto run = [&]() -> int {
        ZG_HIP(hipMalloc(&sink, 8));
        ZG_HIP(hipEventCreate(&e0));
        ZG_HIP(hipEventCreate(&e1));
        hipLaunchKernelGGL(k_mb_fill, dim3(device_cus() * 64), dim3(256), 0, 0, reinterpret_cast<uint64_t *>(buf),
                           nlines * 16);
        ZG_HIP(hipGetLastError());
        auto launch = [&]() {
            if (cooperative) {
                hipLaunchKernelGGL(k_mb_lines_coop, dim3(grid), dim3(bs), 0, 0, buf, nlines, rounds, sink);
                return;
            }
            switch (loads_per_line) {
            case 1: hipLaunchKernelGGL(k_mb_lines<1>, dim3(grid), dim3(bs), 0, 0, buf, nlines, rounds, sink); break;
            case 2: hipLaunchKernelGGL(k_mb_lines<2>, dim3(grid), dim3(bs), 0, 0, buf, nlines, rounds, sink); break;
            case 3: hipLaunchKernelGGL(k_mb_lines<3>, dim3(grid), dim3(bs), 0, 0, buf, nlines, rounds, sink); break;
            case 4: hipLaunchKernelGGL(k_mb_lines<4>, dim3(grid), dim3(bs), 0, 0, buf, nlines, rounds, sink); break;
            case 5: hipLaunchKernelGGL(k_mb_lines<5>, dim3(grid), dim3(bs), 0, 0, buf, nlines, rounds, sink); break;
            case 6: hipLaunchKernelGGL(k_mb_lines<6>, dim3(grid), dim3(bs), 0, 0, buf, nlines, rounds, sink); break;
            case 7: hipLaunchKernelGGL(k_mb_lines<7>, dim3(grid), dim3(bs), 0, 0, buf, nlines, rounds, sink); break;
            default: hipLaunchKernelGGL(k_mb_lines<8>, dim3(grid), dim3(bs), 0, 0, buf, nlines, rounds, sink); break;
            }
        };
        launch();   // warm-up
        ZG_HIP(hipGetLastError());
        ZG_HIP(hipEventRecord(e0, 0));
        launch();
        ZG_HIP(hipGetLastError());
        ZG_HIP(hipEventRecord(e1, 0));
        ZG_HIP(hipEventSynchronize(e1));
        float t = 0;
        ZG_HIP(hipEventElapsedTime(&t, e0, e1));
        const double done = (double)readers * rounds;
        *gbs = done * 128.0 / (t * 1e-3) / 1e9;
        if (ms) *ms = t;
        return ZG_OK;
    };
    rc = run();
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (sink) (void)hipFree(sink);
    (void)hipFree(buf);
    return rc;
}

extern "C" int zg_random_region_bw(uint64_t buffer_bytes, uint64_t region_bytes, int xcd_matched,
                                   uint32_t rounds_per_phase, uint64_t lines, int block_size, int blocks_per_cu,
                                   double *gbs, double *ms)
{
    int rc = ensure_init();
    if (rc) return rc;
    if (!gbs || lines == 0 || rounds_per_phase == 0 || region_bytes < 128 || buffer_bytes < region_bytes)
        return set_error(ZG_EINVAL, "bad region benchmark arguments");
    const int bs = block_size > 0 ? block_size : 256;
    if (bs > 256 || bs % 64) return set_error(ZG_EINVAL, "block size must be a multiple of 64 up to 256");
    const int bpc = blocks_per_cu > 0 ? blocks_per_cu : 16;
    const uint64_t grid = (uint64_t)device_cus() * bpc;
    const uint64_t readers = grid * bs;
    const uint64_t per_reader = (lines + readers - 1) / readers;
    if (per_reader > 0xffffffffull) return set_error(ZG_EOVERFLOW, "too many lines");
    const uint32_t rounds = (uint32_t)per_reader;
    const uint64_t region_lines = region_bytes / 128, nreg = buffer_bytes / 128 / region_lines;
    uint4 *buf = nullptr;
    uint64_t *sink = nullptr;
    if (hipMalloc(&buf, nreg * region_lines * 128) != hipSuccess)
        return set_error(ZG_ENOMEM, "region buffer allocation failed");
    hipEvent_t e0 = nullptr, e1 = nullptr;
    auto run = [&]() -> int {
        ZG_HIP(hipMalloc(&sink, 8));
        ZG_HIP(hipEventCreate(&e0));
        ZG_HIP(hipEventCreate(&e1));
        hipLaunchKernelGGL(k_mb_fill, dim3(device_cus() * 64), dim3(256), 0, 0, reinterpret_cast<uint64_t *>(buf),
                           nreg * region_lines * 16);
        ZG_HIP(hipGetLastError());
        auto launch = [&]() {
            if (xcd_matched)
                hipLaunchKernelGGL(k_mb_region<true>, dim3(grid), dim3(bs), 0, 0, buf, region_lines, nreg, rounds,
                                   rounds_per_phase, sink);
            else
                hipLaunchKernelGGL(k_mb_region<false>, dim3(grid), dim3(bs), 0, 0, buf, region_lines, nreg, rounds,
                                   rounds_per_phase, sink);
        };
        launch();   // warm-up
        ZG_HIP(hipGetLastError());
        ZG_HIP(hipEventRecord(e0, 0));
        launch();
        ZG_HIP(hipGetLastError());
        ZG_HIP(hipEventRecord(e1, 0));
        ZG_HIP(hipEventSynchronize(e1));
        float t = 0;
        ZG_HIP(hipEventElapsedTime(&t, e0, e1));
        *gbs = (double)readers * rounds * 128.0 / (t * 1e-3) / 1e9;
        if (ms) *ms = t;
        return ZG_OK;
    };
    rc = run();
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (sink) (void)hipFree(sink);
    (void)hipFree(buf);
    return rc;
}
