// CPU emulation of the HIP execution model, for host builds of the 1D kernel source under
// AddressSanitizer (test infrastructure only; the product is built by hipcc for gfx950).
//
// Every workgroup runs on the calling CPU thread: its work-items are fibers (a small x86-64
// stack switch, no signal-mask system calls) that a scheduler runs in turn until each reaches
// a barrier.  Supported:
//   * __syncthreads / __syncthreads_or (workgroup barrier; every live work-item must reach a
//     barrier at the same source line, or the emulator aborts with a divergence report);
//   * __shfl / __shfl_xor / __shfl_up / __shfl_down over 64-lane wavefronts (a wave-level
//     barrier on each side of the exchange; every lane of the wave must take part);
//   * atomics on global or shared memory (work-items never run concurrently);
//   * kernel launches, rewritten from the triple-chevron form by tests/emu/emu_build.py into
//     emu::launch(grid, block, body): blocks run one after the other;
//   * the HIP runtime calls the host code uses: copies are memcpy, streams are ignored,
//     every call succeeds.
// Work-item order inside each barrier interval is a seeded permutation (EMU_ORDER_SEED):
// results that change with the seed expose a missing barrier (a race on shared memory).
// "Device" buffers are ordinary host allocations, so AddressSanitizer checks every access
// of a kernel against the caller's allocation, and every shared (LDS) array against its
// declared size.
#pragma once
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <cmath>
#include <functional>
#include <random>
#include <vector>

#if defined(__SANITIZE_ADDRESS__)
#include <sanitizer/common_interface_defs.h>
#define EMU_ASAN 1
#else
#define EMU_ASAN 0
#endif

// ---- qualifiers --------------------------------------------------------------------------
#define __global__
#define __device__
#define __host__
#define __forceinline__ inline
#define __noinline__ __attribute__((noinline))
#define __shared__ static
#define __launch_bounds__(...)
#define __align__(n) __attribute__((aligned(n)))

// ---- vector types ------------------------------------------------------------------------
struct __attribute__((aligned(8))) float2 {
    float x, y;
};
struct __attribute__((aligned(16))) float4 {
    float x, y, z, w;
};
struct __attribute__((aligned(16))) double2 {
    double x, y;
};
struct __attribute__((aligned(8))) int2 {
    int x, y;
};
inline float2 make_float2(float x, float y) { return float2{x, y}; }
inline float4 make_float4(float x, float y, float z, float w) { return float4{x, y, z, w}; }
inline double2 make_double2(double x, double y) { return double2{x, y}; }
inline int2 make_int2(int x, int y) { return int2{x, y}; }

struct dim3 {
    unsigned x, y, z;
    dim3(unsigned a = 1, unsigned b = 1, unsigned c = 1) : x(a), y(b), z(c) {}
};

// ---- runtime API -------------------------------------------------------------------------
typedef int hipError_t;
typedef void *hipStream_t;
typedef void *hipEvent_t;
enum { hipSuccess = 0, hipErrorInvalidValue = 1 };
enum hipMemcpyKind {
    hipMemcpyHostToHost = 0,
    hipMemcpyHostToDevice = 1,
    hipMemcpyDeviceToHost = 2,
    hipMemcpyDeviceToDevice = 3,
    hipMemcpyDefault = 4
};
inline const char *hipGetErrorString(hipError_t) { return "emulated HIP error"; }
inline hipError_t hipGetLastError() { return hipSuccess; }
inline hipError_t hipPeekAtLastError() { return hipSuccess; }
inline hipError_t hipMemcpyAsync(void *d, const void *s, size_t n, hipMemcpyKind, hipStream_t) {
    if (n) memmove(d, s, n);
    return hipSuccess;
}
inline hipError_t hipMemcpy(void *d, const void *s, size_t n, hipMemcpyKind) {
    if (n) memmove(d, s, n);
    return hipSuccess;
}
inline hipError_t hipMemsetAsync(void *d, int v, size_t n, hipStream_t) {
    memset(d, v, n);
    return hipSuccess;
}
inline hipError_t hipMemset(void *d, int v, size_t n) {
    memset(d, v, n);
    return hipSuccess;
}
inline hipError_t hipStreamSynchronize(hipStream_t) { return hipSuccess; }
inline hipError_t hipGetDevice(int *d) {
    *d = 0;
    return hipSuccess;
}
// a device symbol is an ordinary host object here
#define HIP_SYMBOL(x) ((const void *)&(x))
inline hipError_t hipGetSymbolAddress(void **p, const void *sym) {
    *p = const_cast<void *>(sym);
    return hipSuccess;
}
inline hipError_t hipDeviceSynchronize() { return hipSuccess; }

// Stack switch: saves the callee-saved registers on the current stack, stores the stack
// pointer in *from, loads `to` and restores the registers saved there.
extern "C" void emu_switch(void **from, void *to);
asm(R"(
    .text
    .globl emu_switch
    .type emu_switch, @function
emu_switch:
    pushq %rbp
    pushq %rbx
    pushq %r12
    pushq %r13
    pushq %r14
    pushq %r15
    movq %rsp, (%rdi)
    movq %rsi, %rsp
    popq %r15
    popq %r14
    popq %r13
    popq %r12
    popq %rbx
    popq %rbp
    ret
    .size emu_switch, .-emu_switch
)");

namespace emu {

constexpr int WAVE = 64;

enum State { RUN = 0, AT_BAR = 1, AT_WAVE = 2, DONE = 3 };

struct Fiber {
    void *sp = nullptr;
    char *stack = nullptr;
    size_t stack_size = 0;
    int tid = 0;
    dim3 idx;
    State state = RUN;
    int site = 0;      // source line of the barrier it waits at
    int orv = 0;       // __syncthreads_or operand
    uint64_t shfl = 0; // value offered to the wave exchange
};

struct Block {
    std::vector<Fiber> f;
    void *sched_sp = nullptr;
    dim3 bidx, bdim, gdim;
    const std::function<void()> *body = nullptr;
    int cur = 0;
    int or_result = 0;
    void *sched_fake = nullptr;
    const void *main_bottom = nullptr;
    size_t main_size = 0;
};

inline Block *&blk() {
    static Block *b = nullptr;
    return b;
}
inline Fiber &me() { return blk()->f[blk()->cur]; }

inline size_t &stack_bytes() {
    static size_t s = 1 << 20;
    return s;
}

[[noreturn]] inline void die(const char *what, int a = 0, int b = 0) {
    fprintf(stderr, "EMU ERROR: %s (%d, %d)\n", what, a, b);
    fflush(stderr);
    abort();
}

// fiber <-> scheduler switches, annotated for AddressSanitizer
inline void to_sched() {
    Block *B = blk();
    Fiber &F = me();
#if EMU_ASAN
    void *fake = nullptr;
    __sanitizer_start_switch_fiber(&fake, B->main_bottom, B->main_size);
    emu_switch(&F.sp, B->sched_sp);
    __sanitizer_finish_switch_fiber(fake, nullptr, nullptr);
#else
    emu_switch(&F.sp, B->sched_sp);
#endif
}

inline void fiber_entry() {
    Block *B = blk();
#if EMU_ASAN
    __sanitizer_finish_switch_fiber(nullptr, &B->main_bottom, &B->main_size);
#endif
    (*B->body)();
    Fiber &F = me();
    F.state = DONE;
#if EMU_ASAN
    __sanitizer_start_switch_fiber(nullptr, B->main_bottom, B->main_size);
#endif
    void *dead;
    emu_switch(&dead, B->sched_sp);
    abort();   // a finished fiber is never resumed
}

inline void wait(State s, int site) {
    Fiber &F = me();
    F.state = s;
    F.site = site;
    to_sched();
}

inline unsigned &order_seed() {
    static unsigned s = [] {
        const char *e = getenv("EMU_ORDER_SEED");
        return e ? (unsigned)strtoul(e, nullptr, 10) : 0u;
    }();
    return s;
}

inline void run_block(Block &B) {
    const int n = (int)B.f.size();
    std::vector<int> order(n);
    for (int i = 0; i < n; ++i) order[i] = i;
    std::mt19937 rng(order_seed() * 7919u + B.bidx.x);
    for (int i = 0; i < n; ++i) {
        // initial frame: six saved registers, then the entry as return address and a zero
        // return address for the entry itself (the entry sees a call-aligned stack)
        Fiber &F = B.f[i];
        uintptr_t top = ((uintptr_t)F.stack + F.stack_size) & ~(uintptr_t)15;
        void **sp = (void **)(top - 64);
        for (int r = 0; r < 6; ++r) sp[r] = nullptr;
        sp[6] = (void *)&fiber_entry;
        sp[7] = nullptr;
        F.sp = sp;
    }
    for (;;) {
        if (order_seed()) std::shuffle(order.begin(), order.end(), rng);
        for (int k = 0; k < n; ++k) {
            const int i = order[k];
            if (B.f[i].state != RUN) continue;
            B.cur = i;
#if EMU_ASAN
            __sanitizer_start_switch_fiber(&B.sched_fake, B.f[i].stack, B.f[i].stack_size);
            emu_switch(&B.sched_sp, B.f[i].sp);
            __sanitizer_finish_switch_fiber(B.sched_fake, nullptr, nullptr);
#else
            emu_switch(&B.sched_sp, B.f[i].sp);
#endif
        }
        // every fiber now waits at a barrier or has finished
        int done = 0, at_bar = 0;
        for (auto &F : B.f) {
            done += F.state == DONE;
            at_bar += F.state == AT_BAR;
        }
        if (done == n) return;
        bool released = false;
        // wave exchanges: the whole wave must be there
        for (int w0 = 0; w0 < n; w0 += WAVE) {
            const int w1 = std::min(n, w0 + WAVE);
            int cnt = 0, site = -1;
            for (int i = w0; i < w1; ++i)
                if (B.f[i].state == AT_WAVE) {
                    ++cnt;
                    if (site < 0) site = B.f[i].site;
                    else if (B.f[i].site != site) die("wave exchange at different lines", site, B.f[i].site);
                }
            if (cnt == 0) continue;
            if (cnt != w1 - w0) die("wave exchange with part of the wave (divergence)", site, cnt);
            for (int i = w0; i < w1; ++i) B.f[i].state = RUN;
            released = true;
        }
        if (released) continue;
        if (at_bar + done == n && at_bar > 0) {
            if (done) die("workgroup barrier after some work-items finished", B.f[0].site, done);
            const int site = B.f[0].site;
            int orr = 0;
            for (auto &F : B.f) {
                if (F.site != site) die("workgroup barrier at different lines", site, F.site);
                orr |= F.orv;
            }
            B.or_result = orr;
            for (auto &F : B.f) F.state = RUN;
            continue;
        }
        die("deadlock: no barrier can be released", at_bar, done);
    }
}

inline void launch(dim3 grid, dim3 block, const std::function<void()> &body) {
    const int nthreads = (int)(block.x * block.y * block.z);
    Block B;
    B.f.resize(nthreads);
    B.bdim = block;
    B.gdim = grid;
    B.body = &body;
    for (int i = 0; i < nthreads; ++i) {
        Fiber &F = B.f[i];
        F.stack_size = stack_bytes();
        F.stack = (char *)malloc(F.stack_size);
        F.tid = i;
        F.idx = dim3(i % block.x, (i / block.x) % block.y, i / (block.x * block.y));
    }
    Block *prev = blk();
    blk() = &B;
    for (unsigned bz = 0; bz < grid.z; ++bz)
        for (unsigned by = 0; by < grid.y; ++by)
            for (unsigned bx = 0; bx < grid.x; ++bx) {
                B.bidx = dim3(bx, by, bz);
                for (auto &F : B.f) {
                    F.state = RUN;
                    F.orv = 0;
                }
                run_block(B);
            }
    blk() = prev;
    for (auto &F : B.f) free(F.stack);
}

template <class T>
inline T exchange(T v, int src_lane, int site) {
    static_assert(sizeof(T) <= 8, "exchange of at most 8 bytes");
    Fiber &F = me();
    uint64_t bits = 0;
    memcpy(&bits, &v, sizeof(T));
    F.shfl = bits;
    wait(AT_WAVE, site);
    const int base = F.tid - F.tid % WAVE;
    const int lane = F.tid % WAVE;
    const int src = (src_lane < 0 || src_lane >= WAVE) ? lane : src_lane;
    T out;
    memcpy(&out, &blk()->f[base + src].shfl, sizeof(T));
    wait(AT_WAVE, -site);   // everyone has read before the next exchange overwrites
    return out;
}

inline int syncthreads_or(int v, int site) {
    me().orv = v ? 1 : 0;
    wait(AT_BAR, site);
    me().orv = 0;
    return blk()->or_result;
}

}  // namespace emu

#define threadIdx (emu::me().idx)
#define blockIdx (emu::blk()->bidx)
#define blockDim (emu::blk()->bdim)
#define gridDim (emu::blk()->gdim)
#define __syncthreads() emu::wait(emu::AT_BAR, __LINE__)
#define __syncthreads_or(v) emu::syncthreads_or((v), __LINE__)
#define __shfl_xor(v, m, w) emu::exchange((v), (int)(emu::me().tid % emu::WAVE) ^ (m), __LINE__)
#define __shfl_up(v, d, w)                                                                  \
    emu::exchange((v), (int)(emu::me().tid % emu::WAVE) - (int)(d), __LINE__)
#define __shfl_down(v, d, w)                                                                \
    emu::exchange((v), (int)(emu::me().tid % emu::WAVE) + (int)(d), __LINE__)
#define __shfl(v, s, w) emu::exchange((v), (int)(s), __LINE__)

template <class T>
inline T atomicAdd(T *p, T v) {
    T o = *p;
    *p = o + v;
    return o;
}
inline int atomicOr(int *p, int v) {
    int o = *p;
    *p = o | v;
    return o;
}
inline unsigned atomicOr(unsigned *p, unsigned v) {
    unsigned o = *p;
    *p = o | v;
    return o;
}
template <class T>
inline T atomicMax(T *p, T v) {
    T o = *p;
    *p = o > v ? o : v;
    return o;
}
