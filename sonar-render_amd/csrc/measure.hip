// Attainable HBM bandwidth on this GPU: a plain streaming read (16-byte loads,
// grid-stride over the buffer, one XOR per load so nothing is dropped) and a
// plain streaming write, the ceilings against which the memory-bound kernels'
// rates are read.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "launch.h"
#include "measure.h"

namespace sr {

__global__ __launch_bounds__(256) void k_stream_read(const uint4 *__restrict__ src, long n, unsigned *__restrict__ out) {
    const long stride = (long)gridDim.x * blockDim.x;
    unsigned acc = 0;
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        const uint4 v = src[i];
        acc ^= v.x ^ v.y ^ v.z ^ v.w;
    }
    out[(long)blockIdx.x * blockDim.x + threadIdx.x] = acc;
}

// A plain streaming write: 16-byte stores, grid-stride, non-temporal (past the
// caches, as the S build writes a matrix larger than them) or ordinary.
template <bool NT>
__global__ __launch_bounds__(256) void k_stream_write(uint4 *__restrict__ dst, long n, unsigned v) {
    typedef unsigned int u4 __attribute__((ext_vector_type(4)));
    const long stride = (long)gridDim.x * blockDim.x;
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        const unsigned w = v ^ (unsigned)i;
        if (NT) __builtin_nontemporal_store(u4{w, w, w, w}, reinterpret_cast<u4 *>(dst + i));
        else dst[i] = make_uint4(w, w, w, w);
    }
}

// Streams `bytes` (rounded down to 16) `reps` times after two warm-up passes on its
// own stream and buffers (allocated and freed here); gbs = bytes / mean time.
// kind 0: read, 1: non-temporal write, 2: ordinary write.
static int measure_stream(int kind, long bytes, int reps, double *gbs) {
    if (bytes < (1L << 20) || reps < 1 || !gbs) { set_error("stream measurement: at least 1 MiB and one repetition"); return -2; }
    const long n = bytes / 16;
    int cus = 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, 0) != hipSuccess || cus <= 0) cus = 256;
    const int grid = 4 * cus;  // four 256-thread workgroups per CU: 16 waves each
    uint4 *buf = nullptr;
    unsigned *out = nullptr;
    hipStream_t st = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    auto pass = [&]() {
        if (kind == 0) k_stream_read<<<grid, 256, 0, st>>>(buf, n, out);
        else if (kind == 1) k_stream_write<true><<<grid, 256, 0, st>>>(buf, n, 0x3c003c00u);
        else k_stream_write<false><<<grid, 256, 0, st>>>(buf, n, 0x3c003c00u);
    };
    auto run = [&]() -> int {
        SR_HIP(hipMalloc((void **)&buf, (size_t)n * 16));
        SR_HIP(hipMalloc((void **)&out, (size_t)grid * 256 * sizeof(unsigned)));
        SR_HIP(hipStreamCreate(&st));
        SR_HIP(hipEventCreate(&e0));
        SR_HIP(hipEventCreate(&e1));
        SR_HIP(hipMemsetAsync(buf, 1, (size_t)n * 16, st));
        for (int r = 0; r < 2; r++) pass();
        SR_HIP(hipGetLastError());
        SR_HIP(hipEventRecord(e0, st));
        for (int r = 0; r < reps; r++) pass();
        SR_HIP(hipGetLastError());
        SR_HIP(hipEventRecord(e1, st));
        SR_HIP(hipEventSynchronize(e1));
        float ms = 0.f;
        SR_HIP(hipEventElapsedTime(&ms, e0, e1));
        *gbs = (double)n * 16.0 / ((double)ms / reps * 1e-3) / 1e9;
        return 0;
    };
    const int rc = run();
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (st) (void)hipStreamDestroy(st);
    if (out) (void)hipFree(out);
    if (buf) (void)hipFree(buf);
    return rc;
}

// Attainable vector-instruction issue rate: every SIMD full of waves (eight per
// SIMD), each running 16 independent single-cycle-class integer vector
// instructions per iteration (eight add chains, eight xor chains; written as
// inline asm so each is one instruction) around a scalar loop counter.
constexpr int kIssueIters = 4096;
constexpr int kIssuePerIter = 16;
__global__ __launch_bounds__(256) void k_valu_issue(unsigned *__restrict__ out, unsigned seed) {
    unsigned a[8];
#pragma unroll
    for (int c = 0; c < 8; c++) a[c] = seed ^ (threadIdx.x * 8u + c);
    const unsigned b = seed * 3u + 1u;
    for (int i = 0; i < kIssueIters; i++) {
#pragma unroll
        for (int c = 0; c < 8; c++) asm volatile("v_add_u32 %0, %0, %1" : "+v"(a[c]) : "v"(b));
#pragma unroll
        for (int c = 0; c < 8; c++) asm volatile("v_xor_b32 %0, %0, %1" : "+v"(a[c]) : "v"(b));
    }
    unsigned r = 0;
#pragma unroll
    for (int c = 0; c < 8; c++) r ^= a[c];
    out[blockIdx.x * blockDim.x + threadIdx.x] = r;
}

int measure_valu_issue(int reps, double *ginstr) {
    if (reps < 1 || !ginstr) { set_error("vector issue measurement: at least one repetition"); return -2; }
    int cus = 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, 0) != hipSuccess || cus <= 0) cus = 256;
    const int grid = cus * 8;  // 8 workgroups of 4 waves per CU: 8 waves per SIMD
    unsigned *out = nullptr;
    hipStream_t st = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    auto run = [&]() -> int {
        SR_HIP(hipMalloc((void **)&out, (size_t)grid * 256 * sizeof(unsigned)));
        SR_HIP(hipStreamCreate(&st));
        SR_HIP(hipEventCreate(&e0));
        SR_HIP(hipEventCreate(&e1));
        k_valu_issue<<<grid, 256, 0, st>>>(out, 1u);
        SR_HIP(hipGetLastError());
        SR_HIP(hipEventRecord(e0, st));
        for (int r = 0; r < reps; r++) k_valu_issue<<<grid, 256, 0, st>>>(out, (unsigned)r + 2u);
        SR_HIP(hipGetLastError());
        SR_HIP(hipEventRecord(e1, st));
        SR_HIP(hipEventSynchronize(e1));
        float ms = 0.f;
        SR_HIP(hipEventElapsedTime(&ms, e0, e1));
        const double instr = (double)grid * 4 * kIssueIters * kIssuePerIter;  // wave-instructions per launch
        *ginstr = instr / ((double)ms / reps * 1e-3) / 1e9;
        return 0;
    };
    const int rc = run();
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (st) (void)hipStreamDestroy(st);
    if (out) (void)hipFree(out);
    return rc;
}

int measure_stream_read(long bytes, int reps, double *gbs) { return measure_stream(0, bytes, reps, gbs); }

int measure_stream_write(long bytes, int reps, int nontemporal, double *gbs) {
    return measure_stream(nontemporal ? 1 : 2, bytes, reps, gbs);
}

}  // namespace sr
