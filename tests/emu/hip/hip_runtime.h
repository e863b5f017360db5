// CPU build of the library's HIP kernels, for bounds and barrier checks under
// the host sanitizers (test infrastructure, never part of the library).
//
// Every GPU thread of a workgroup is a host thread.  __syncthreads() is a
// barrier over the workgroup's threads, and the wavefront operations
// (__shfl_xor, readfirstlane) exchange values through a per-wavefront buffer
// between two barriers over its 64 lanes, so every access the kernel makes is
// ordered exactly as the kernel's own barriers order it.  ThreadSanitizer then
// reports any pair of accesses of different threads that no barrier orders
// (an LDS race), and AddressSanitizer any access past the dynamic LDS size or
// a global buffer (the LDS of a workgroup is a heap block of exactly the
// launch's size).  LDS starts filled with NaN bytes, so a read of LDS that
// nothing wrote shows up in the results.  Workgroups run one after another.
#pragma once
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <atomic>
#include <chrono>
#include <thread>
#include <vector>

#define __global__
#define __device__
#define __host__
#define __forceinline__ inline
#define __launch_bounds__(...)
// a kernel's static LDS array is one array for all threads of the running workgroup
// (workgroups run one after another); the kernels' dynamic-LDS declaration is replaced
// by the test that builds them
#define __shared__ static
#define __align__(n) alignas(n)

struct dim3 {
    unsigned x, y, z;
    dim3(unsigned a = 1, unsigned b = 1, unsigned c = 1) : x(a), y(b), z(c) {}
};
struct double2 {
    double x, y;
};
struct double4 {
    double x, y, z, w;
};
struct float4 {
    float x, y, z, w;
};
struct int2 {
    int x, y;
};
inline double4 make_double4(double a, double b, double c, double d) { return double4{a, b, c, d}; }
inline float4 make_float4(float a, float b, float c, float d) { return float4{a, b, c, d}; }
inline int2 make_int2(int a, int b) { return int2{a, b}; }
inline unsigned __brev(unsigned v) {
    unsigned r = 0;
    for (int i = 0; i < 32; i++) r |= ((v >> i) & 1u) << (31 - i);
    return r;
}
struct float2 {
    float x, y;
};
inline double2 make_double2(double a, double b) { return double2{a, b}; }
inline float2 make_float2(float a, float b) { return float2{a, b}; }

typedef int hipError_t;
enum { hipSuccess = 0, hipErrorInvalidValue = 1, hipErrorNotSupported = 801, hipErrorLaunchFailure = 719 };
typedef void *hipStream_t;
enum hipFuncAttribute { hipFuncAttributeMaxDynamicSharedMemorySize = 8 };
inline hipError_t hipFuncSetAttribute(const void *, hipFuncAttribute, int) { return hipSuccess; }
// a device of two CUs: the persistent kernels run two workgroups that each walk
// several work items
enum hipDeviceAttribute_t { hipDeviceAttributeMultiprocessorCount = 0 };
inline hipError_t hipGetDevice(int *d) { *d = 0; return hipSuccess; }
// device memory is host memory (new buffers hold NaN bytes, as the library's own
// checks of unwritten memory assume nothing)
inline hipError_t hipMalloc(void **p, size_t n) {
    *p = aligned_alloc(256, (n + 255) & ~(size_t)255);
    if (!*p) return hipErrorInvalidValue;
    memset(*p, 0xff, n);
    return hipSuccess;
}
template <class T>
inline hipError_t hipMalloc(T **p, size_t n) {
    return hipMalloc((void **)p, n);
}
inline hipError_t hipFree(void *p) { free(p); return hipSuccess; }
inline hipError_t hipMemset(void *p, int v, size_t n) { memset(p, v, n); return hipSuccess; }
inline hipError_t hipDeviceSynchronize() { return hipSuccess; }
inline const char *hipGetErrorString(hipError_t) { return "emulated HIP error"; }
// streams and events: launches run to completion when made, so every stream is
// always idle; copies are plain memcpy
enum hipMemcpyKind { hipMemcpyHostToHost = 0, hipMemcpyHostToDevice = 1, hipMemcpyDeviceToHost = 2,
                     hipMemcpyDeviceToDevice = 3, hipMemcpyDefault = 4 };
typedef void *hipEvent_t;
constexpr unsigned hipStreamNonBlocking = 1;
inline hipError_t hipSetDevice(int) { return hipSuccess; }
inline hipError_t hipGetDeviceCount(int *n) { *n = 1; return hipSuccess; }
inline hipError_t hipStreamCreateWithFlags(hipStream_t *s, unsigned) {
    static char tag;
    *s = &tag;
    return hipSuccess;
}
inline hipError_t hipStreamDestroy(hipStream_t) { return hipSuccess; }
inline hipError_t hipStreamSynchronize(hipStream_t) { return hipSuccess; }
inline hipError_t hipMemcpy(void *d, const void *src, size_t n, hipMemcpyKind) { memcpy(d, src, n); return hipSuccess; }
inline hipError_t hipMemcpyAsync(void *d, const void *src, size_t n, hipMemcpyKind, hipStream_t) {
    memcpy(d, src, n);
    return hipSuccess;
}
inline hipError_t hipEventCreate(hipEvent_t *e) {
    static char tag;
    *e = &tag;
    return hipSuccess;
}
inline hipError_t hipEventDestroy(hipEvent_t) { return hipSuccess; }
inline hipError_t hipEventRecord(hipEvent_t, hipStream_t) { return hipSuccess; }
inline hipError_t hipEventElapsedTime(float *ms, hipEvent_t, hipEvent_t) { *ms = 0.f; return hipSuccess; }
// a file-scope __device__ array is a host array: its address is the symbol's own
#define HIP_SYMBOL(x) (x)
template <class T>
inline hipError_t hipGetSymbolAddress(void **p, const T &sym) {
    *p = (void *)&sym;
    return hipSuccess;
}
inline hipError_t hipDeviceGetAttribute(int *v, hipDeviceAttribute_t, int) { *v = 2; return hipSuccess; }

inline int min(int a, int b) { return a < b ? a : b; }
inline int max(int a, int b) { return a > b ? a : b; }
inline long long min(long long a, long long b) { return a < b ? a : b; }
inline long long max(long long a, long long b) { return a > b ? a : b; }
inline unsigned min(unsigned a, unsigned b) { return a < b ? a : b; }
inline unsigned max(unsigned a, unsigned b) { return a > b ? a : b; }

inline thread_local dim3 emu_tid, emu_bid, emu_bdim, emu_gdim;

namespace emu {

// generation barrier on atomics (acquire/release, which ThreadSanitizer models
// exactly); a wait longer than 20 s means the threads do not all reach the same
// barrier (a divergent __syncthreads), and aborts
struct Barrier {
    std::atomic<int> count{0};
    std::atomic<unsigned> gen{0};
    int n = 0;
    void wait() {
        const unsigned g = gen.load(std::memory_order_acquire);
        if (count.fetch_add(1, std::memory_order_acq_rel) + 1 == n) {
            count.store(0, std::memory_order_relaxed);
            gen.fetch_add(1, std::memory_order_acq_rel);
            return;
        }
        const auto t0 = std::chrono::steady_clock::now();
        while (gen.load(std::memory_order_acquire) == g) {
            std::this_thread::yield();
            if (std::chrono::steady_clock::now() - t0 > std::chrono::seconds(20)) {
                fprintf(stderr, "emu: barrier timeout (threads do not all reach one barrier)\n");
                abort();
            }
        }
    }
};

struct Wave {
    Barrier bar;
    uint64_t buf[64];
};

struct Block {
    Barrier bar;
    std::vector<Wave> waves;
    char *lds = nullptr;
};

inline thread_local Block *cur = nullptr;
inline hipError_t last_error = hipSuccess;

inline char *lds() { return cur->lds; }

template <class T>
T exchange(T v, int src_of_lane_xor, bool first) {
    static_assert(sizeof(T) <= 8, "wave exchange of at most 8 bytes");
    const int t = (int)emu_tid.x, lane = t & 63;
    Wave &w = cur->waves[t >> 6];
    uint64_t b = 0;
    memcpy(&b, &v, sizeof(T));
    w.buf[lane] = b;
    w.bar.wait();
    const uint64_t r = w.buf[first ? 0 : (lane ^ src_of_lane_xor)];
    w.bar.wait();
    T out;
    memcpy(&out, &r, sizeof(T));
    return out;
}

// every lane's value of v, in lane order
template <class T>
void gather64(T v, T (&out)[64]) {
    static_assert(sizeof(T) <= 8, "wave exchange of at most 8 bytes");
    const int t = (int)emu_tid.x, lane = t & 63;
    Wave &w = cur->waves[t >> 6];
    uint64_t b = 0;
    memcpy(&b, &v, sizeof(T));
    w.buf[lane] = b;
    w.bar.wait();
    for (int i = 0; i < 64; i++) memcpy(&out[i], &w.buf[i], sizeof(T));
    w.bar.wait();
}

// DPP move (the controls the kernels use): row_shr:1..15 within rows of 16 lanes,
// row_bcast:15 (lane 15 of each row to the next row) and row_bcast:31 (lane 31 to
// rows 2 and 3); a lane whose row or bank is masked off, or whose source is
// outside the row with bound_ctrl off, keeps `old`
inline int update_dpp(int old, int src, int ctrl, int row_mask, int bank_mask, bool bound_ctrl) {
    int all[64];
    gather64(src, all);
    const int l = (int)emu_tid.x & 63, row = l >> 4, l16 = l & 15;
    if (!((row_mask >> row) & 1) || !((bank_mask >> (l16 >> 2)) & 1)) return old;
    int from = -1;
    if (ctrl >= 0x111 && ctrl <= 0x11f) {
        if (l16 >= ctrl - 0x110) from = l - (ctrl - 0x110);
    } else if (ctrl == 0x142) {
        if (row >= 1) from = 16 * row - 1;
    } else if (ctrl == 0x143) {
        if (row >= 2) from = 31;
    } else {
        fprintf(stderr, "emu: DPP control 0x%x not emulated\n", ctrl);
        abort();
    }
    if (from < 0) return bound_ctrl ? 0 : old;
    return all[from];
}

inline unsigned long long ballot(bool pred) {
    int all[64];
    gather64((int)pred, all);
    unsigned long long m = 0;
    for (int i = 0; i < 64; i++)
        if (all[i]) m |= 1ull << i;
    return m;
}

// v_mfma_f64_16x16x4f64: D (16 x 16) = C + A (16 x 4) B (4 x 16), lane l holding
// A[l % 16][l / 16] and B[l / 16][l % 16]; register q of lane l is D[l / 16 + 4 q][l % 16]
// v_mfma_f64_16x16x4f64 / v_mfma_f32_16x16x4f32, one 16 x 16 x 4 block: lane l holds
// A(l % 16, l / 16) and B(l / 16, l % 16); register q of the accumulator holds
// D(l / 16 + 4 q, l % 16) in the FP64 form and D(4 (l / 16) + q, l % 16) in the FP32
// form.  V is the kernel's own 4-vector type.
template <class V, class R>
inline V mfma_16x16x4(R a, R b, V c, bool f64) {
    R A[64], B[64];
    gather64(a, A);
    gather64(b, B);
    const int l = (int)emu_tid.x & 63;
    V d = c;
    for (int q = 0; q < 4; q++) {
        const int row = f64 ? (l >> 4) + 4 * q : 4 * (l >> 4) + q, col = l & 15;
        R s = 0;
        for (int k = 0; k < 4; k++) s += A[16 * k + row] * B[16 * k + col];
        d[q] = c[q] + s;
    }
    return d;
}

template <class T>
T exchange_from(T v, int src) {
    const int t = (int)emu_tid.x;
    return exchange(v, (t & 63) ^ (src & 63), false);
}

}  // namespace emu

#define threadIdx emu_tid
#define blockIdx emu_bid
#define blockDim emu_bdim
#define gridDim emu_gdim

inline void __syncthreads() { emu::cur->bar.wait(); }
template <class T>
inline T __shfl_xor(T v, int o, int width = 64) {
    (void)width;
    return emu::exchange(v, o, false);
}
#define __builtin_amdgcn_readfirstlane(v) emu::exchange((v), 0, true)
// readlane(v, l): lane l's v (l uniform); as an exchange with lane (lane ^ lane ^ l)
#define __builtin_amdgcn_readlane(v, l) emu::exchange_from((v), (l))
// permlane16/32_swap(x, x): x of this lane and of lane ^ 16 (32).  The kernels
// use only the unordered pair, so the emulation returns {own, partner}.
namespace emu {
struct U2 {
    unsigned v[2];
    unsigned operator[](int i) const { return v[i]; }
};
}  // namespace emu
#define __builtin_amdgcn_permlane16_swap(a, b, fi, bc) emu::U2{{(unsigned)(a), emu::exchange((unsigned)(b), 16, false)}}
#define __builtin_amdgcn_permlane32_swap(a, b, fi, bc) emu::U2{{(unsigned)(a), emu::exchange((unsigned)(b), 32, false)}}
inline int __double2loint(double d) {
    uint64_t u;
    memcpy(&u, &d, 8);
    return (int)(uint32_t)u;
}
inline int __double2hiint(double d) {
    uint64_t u;
    memcpy(&u, &d, 8);
    return (int)(uint32_t)(u >> 32);
}
inline double __hiloint2double(int hi, int lo) {
    const uint64_t u = ((uint64_t)(uint32_t)hi << 32) | (uint32_t)lo;
    double d;
    memcpy(&d, &u, 8);
    return d;
}
#define __builtin_amdgcn_s_setprio(p) ((void)0)
#define __builtin_amdgcn_update_dpp(old, src, ctrl, rmask, bmask, bc) emu::update_dpp((old), (src), (ctrl), (rmask), (bmask), (bc))
#define __ballot(p) emu::ballot(p)
#define __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, cbsz, abid, blgp) emu::mfma_16x16x4((double)(a), (double)(b), (c), true)
#define __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, cbsz, abid, blgp) emu::mfma_16x16x4((float)(a), (float)(b), (c), false)
// a wavefront-scope fence: the lanes of a wavefront meet, which orders their
// LDS accesses as the hardware's in-order LDS queue does
#define __builtin_amdgcn_fence(order, scope) emu::wave_fence(scope)
inline int __ffsll(unsigned long long x) { return __builtin_ffsll((long long)x); }
inline long long __double_as_longlong(double d) {
    long long r;
    memcpy(&r, &d, 8);
    return r;
}
inline double __longlong_as_double(long long v) {
    double r;
    memcpy(&r, &v, 8);
    return r;
}
namespace emu {
inline void wave_fence(const char *scope) {
    if (strcmp(scope, "wavefront") != 0) {
        fprintf(stderr, "emu: fence scope %s not emulated\n", scope);
        abort();
    }
    cur->waves[(int)emu_tid.x >> 6].bar.wait();
}
}  // namespace emu
// the barrier that follows orders global memory as well (host threads share it)
inline void __threadfence_block() { std::atomic_thread_fence(std::memory_order_seq_cst); }
inline int atomicOr(int *p, int v) { return __atomic_fetch_or(p, v, __ATOMIC_SEQ_CST); }
inline hipError_t hipGetLastError() {
    const hipError_t e = emu::last_error;
    emu::last_error = hipSuccess;
    return e;
}

namespace emu {
template <class... P, class... A>
hipError_t launch(void (*k)(P...), dim3 grid, dim3 block, size_t lds_bytes, hipStream_t, A... args) {
    const int nt = (int)(block.x * block.y * block.z);
    if (nt % 64 || lds_bytes > 160 * 1024) return last_error = hipErrorInvalidValue;
    // one host thread per GPU thread for the whole launch; the workgroups run one
    // after another (x fastest), each on freshly NaN-filled LDS, between two
    // barriers over all threads
    Block blk;
    blk.bar.n = nt;
    blk.waves = std::vector<Wave>(nt / 64);
    for (auto &w : blk.waves) w.bar.n = 64;
    const size_t sz = (lds_bytes + 15) & ~(size_t)15;
    blk.lds = (char *)aligned_alloc(16, sz ? sz : 16);
    Barrier between;
    between.n = nt;
    const unsigned nblocks = grid.x * grid.y * grid.z;
    std::vector<std::thread> th;
    th.reserve(nt);
    for (int t = 0; t < nt; t++)
        th.emplace_back([&, t] {
            cur = &blk;
            emu_tid = dim3(t);
            emu_bdim = block;
            emu_gdim = grid;
            for (unsigned b = 0; b < nblocks; b++) {
                if (t == 0) memset(blk.lds, 0xff, sz);
                between.wait();
                emu_bid = dim3(b % grid.x, (b / grid.x) % grid.y, b / (grid.x * grid.y));
                k(args...);
                between.wait();
            }
        });
    for (auto &t : th) t.join();
    free(blk.lds);
    return hipSuccess;
}
}  // namespace emu

#define hipLaunchKernelGGL(k, g, b, lds, s, ...) emu::launch(k, dim3(g), dim3(b), (size_t)(lds), s, __VA_ARGS__)
