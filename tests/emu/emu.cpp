// emu.cpp - CPU implementation of the HIP subset in tests/emu/hip/hip_runtime.h.
//
// TEST INFRASTRUCTURE ONLY (see the header).  Allocations are malloc'ed with their exact
// size, so AddressSanitizer reports any access outside them.  A kernel launch runs its
// workgroups one after the other; the work-items of a workgroup are fibers (a minimal
// x86-64 context switch, with the ASan fiber annotations) scheduled round robin, and
// every cross-lane operation or barrier is a rendezvous of the active work-items of its
// wavefront (of 64) or workgroup.
#include "hip/hip_runtime.h"

#include <sys/mman.h>

#include <chrono>
#include <vector>

#ifdef __SANITIZE_ADDRESS__
#include <sanitizer/common_interface_defs.h>
#define LFEMU_ASAN 1
#else
#define LFEMU_ASAN 0
#endif

// ---------------------------------------------------------------------------- runtime
struct lfemu_stream { int id; };
struct lfemu_event { double t; };

static double now_ms() {
    using namespace std::chrono;
    return duration<double, std::milli>(steady_clock::now().time_since_epoch()).count();
}

hipError_t lfemu_malloc(void **p, size_t n) {
    *p = std::malloc(n ? n : 1);
    return *p ? hipSuccess : hipErrorOutOfMemory;
}
hipError_t hipFree(void *p) { std::free(p); return hipSuccess; }
hipError_t lfemu_host_malloc(void **p, size_t n, unsigned) { return lfemu_malloc(p, n); }
hipError_t hipHostFree(void *p) { std::free(p); return hipSuccess; }
hipError_t hipMemcpy(void *d, const void *s, size_t n, hipMemcpyKind) {
    if (n) std::memmove(d, s, n);
    return hipSuccess;
}
hipError_t hipMemcpyAsync(void *d, const void *s, size_t n, hipMemcpyKind k, hipStream_t) {
    return hipMemcpy(d, s, n, k);
}
hipError_t hipMemset(void *p, int v, size_t n) { if (n) std::memset(p, v, n); return hipSuccess; }
hipError_t hipMemsetAsync(void *p, int v, size_t n, hipStream_t) { return hipMemset(p, v, n); }
hipError_t hipStreamCreate(hipStream_t *s) { *s = new lfemu_stream{0}; return hipSuccess; }
hipError_t hipStreamDestroy(hipStream_t s) { delete s; return hipSuccess; }
hipError_t hipStreamSynchronize(hipStream_t) { return hipSuccess; }
hipError_t hipStreamWaitEvent(hipStream_t, hipEvent_t, unsigned) { return hipSuccess; }
hipError_t hipEventCreate(hipEvent_t *e) { *e = new lfemu_event{now_ms()}; return hipSuccess; }
hipError_t hipEventCreateWithFlags(hipEvent_t *e, unsigned) { return hipEventCreate(e); }
hipError_t hipEventDestroy(hipEvent_t e) { delete e; return hipSuccess; }
hipError_t hipEventRecord(hipEvent_t e, hipStream_t) { e->t = now_ms(); return hipSuccess; }
hipError_t hipEventSynchronize(hipEvent_t) { return hipSuccess; }
hipError_t hipEventElapsedTime(float *ms, hipEvent_t a, hipEvent_t b) {
    *ms = (float)(b->t - a->t);
    return hipSuccess;
}
hipError_t hipDeviceSynchronize() { return hipSuccess; }
hipError_t hipGetLastError() { return hipSuccess; }
const char *hipGetErrorString(hipError_t e) { return e == hipSuccess ? "no error" : "lfemu error"; }
hipError_t hipGetDeviceCount(int *n) { *n = 1; return hipSuccess; }
hipError_t hipGetDevice(int *d) { *d = 0; return hipSuccess; }
hipError_t hipSetDevice(int) { return hipSuccess; }
hipError_t hipGetDeviceProperties(hipDeviceProp_t *p, int) {
    std::memset(p, 0, sizeof(*p));
    std::snprintf(p->name, sizeof(p->name), "CPU emulation");
    std::snprintf(p->gcnArchName, sizeof(p->gcnArchName), "emu");
    p->multiProcessorCount = 8;
    p->totalGlobalMem = (size_t)64 << 30;
    return hipSuccess;
}
hipError_t lfemu_symbol_address(void **p, const void *sym) { *p = (void *)sym; return hipSuccess; }

// ---------------------------------------------------------------------------- fibers
// void lfemu_switch(void **save_sp, void *new_sp): saves the callee-saved registers on
// the current stack, stores its stack pointer, switches to new_sp and restores there
asm(R"(
    .text
    .globl lfemu_switch
    .type lfemu_switch,@function
lfemu_switch:
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
    .size lfemu_switch, .-lfemu_switch
)");
extern "C" void lfemu_switch(void **save_sp, void *new_sp);

namespace lfemu {

thread_local Idx g_block{0, 0, 0}, g_bdim{1, 1, 1}, g_gdim{1, 1, 1};

namespace {

constexpr size_t kStack = 512 * 1024;
constexpr int kMaxItems = 1024;

struct Fiber {
    Item item;
    void *sp = nullptr;
    char *stack = nullptr;          // lowest address of the usable stack
    bool done = false;
    void *fake = nullptr;           // ASan fake stack of this fiber while switched out
};

struct Wave {
    int nactive = 0;
    unsigned gen = 0;
    int arrive[2] = {0, 0};
    int val[2][64];
    bool posted[2][64];
};

struct Group {
    int n = 0;                       // work-items
    int nwave = 0;
    Fiber fib[kMaxItems];
    Wave wave[kMaxItems / 64];
    int bar_active = 0, bar_arrive = 0;
    unsigned bar_gen = 0;
    int current = -1;
    void *sched_sp = nullptr;
    const void *sched_bottom = nullptr;
    size_t sched_size = 0;
    const std::function<void()> *body = nullptr;
    bool progress = false;
};

Group *g_grp = nullptr;
thread_local Item g_host_item{{0, 0, 0}};

char *stack_pool(int i) {
    static char *pool[kMaxItems] = {};
    if (!pool[i]) {
        void *m = mmap(nullptr, kStack + 4096, PROT_READ | PROT_WRITE, MAP_PRIVATE | MAP_ANONYMOUS, -1, 0);
        if (m == MAP_FAILED) { std::perror("lfemu: mmap"); std::abort(); }
        mprotect(m, 4096, PROT_NONE);                     // guard page below the stack
        pool[i] = (char *)m + 4096;
    }
    return pool[i];
}

void to_sched(Fiber &f, bool finishing) {
    Group &G = *g_grp;
#if LFEMU_ASAN
    __sanitizer_start_switch_fiber(finishing ? nullptr : &f.fake, G.sched_bottom, G.sched_size);
#endif
    lfemu_switch(&f.sp, G.sched_sp);
#if LFEMU_ASAN
    __sanitizer_finish_switch_fiber(f.fake, &G.sched_bottom, &G.sched_size);
#endif
}

void yield_wait() {
    Group &G = *g_grp;
    to_sched(G.fib[G.current], false);
}

void fiber_main() {
    Group &G = *g_grp;
    Fiber &f = G.fib[G.current];
#if LFEMU_ASAN
    __sanitizer_finish_switch_fiber(nullptr, &G.sched_bottom, &G.sched_size);
#endif
    (*G.body)();
    // leave the wave and the workgroup: rendezvous waiting for this item complete
    Fiber &me = G.fib[G.current];
    const int w = G.current / 64;
    Wave &W = G.wave[w];
    W.nactive--;
    const int s = W.gen & 1;
    if (W.arrive[s] > 0 && W.arrive[s] == W.nactive) {
        W.gen++;
        W.arrive[s ^ 1] = 0;
        std::memset(W.posted[s ^ 1], 0, sizeof(W.posted[s ^ 1]));
    }
    G.bar_active--;
    if (G.bar_arrive > 0 && G.bar_arrive == G.bar_active) { G.bar_arrive = 0; G.bar_gen++; }
    me.done = true;
    G.progress = true;
    to_sched(me, true);
    std::abort();                                        // never resumed
}

// wave rendezvous: post v, wait until every active item of the wave posted; returns the
// slot whose values are now complete
int wave_post(int v) {
    Group &G = *g_grp;
    const int me = G.current;
    Wave &W = G.wave[me / 64];
    const unsigned g = W.gen;
    const int s = g & 1;
    W.val[s][me % 64] = v;
    W.posted[s][me % 64] = true;
    W.arrive[s]++;
    G.progress = true;
    if (W.arrive[s] == W.nactive) {
        W.gen++;
        W.arrive[s ^ 1] = 0;
        std::memset(W.posted[s ^ 1], 0, sizeof(W.posted[s ^ 1]));
    }
    while (W.gen == g) yield_wait();
    return s;
}

}  // namespace

Item &cur() {
    if (!g_grp || g_grp->current < 0) return g_host_item;
    return g_grp->fib[g_grp->current].item;
}

int xlane_shift(int old, int v, int delta, bool zero_fill) {
    Group &G = *g_grp;
    const int lane = G.current % 64;
    const int s = wave_post(v);
    const Wave &W = G.wave[G.current / 64];
    const int src = lane + delta;
    if (src < 0 || src > 63 || !W.posted[s][src]) return zero_fill ? 0 : old;
    return W.val[s][src];
}

int xlane_read(int v, int lane) {
    Group &G = *g_grp;
    const int s = wave_post(v);
    const Wave &W = G.wave[G.current / 64];
    lane &= 63;                                          // v_readlane uses bits [5:0] of the lane select
    return W.posted[s][lane] ? W.val[s][lane] : 0;
}

int xlane_first(int v) {
    Group &G = *g_grp;
    const int s = wave_post(v);
    const Wave &W = G.wave[G.current / 64];
    for (int l = 0; l < 64; ++l)
        if (W.posted[s][l]) return W.val[s][l];
    return v;
}

int xlane_perm(int v, int mode, int arg) {
    Group &G = *g_grp;
    const int lane = G.current % 64;
    const int s = wave_post(v);
    const Wave &W = G.wave[G.current / 64];
    int src = mode == 0 ? lane + arg : (mode == 1 ? lane - arg : (lane ^ arg));
    if (src < 0 || src > 63) return v;                   // out of the wave: own value
    return W.posted[s][src] ? W.val[s][src] : 0;
}

void barrier() {
    Group &G = *g_grp;
    const unsigned g = G.bar_gen;
    G.bar_arrive++;
    G.progress = true;
    if (G.bar_arrive == G.bar_active) { G.bar_arrive = 0; G.bar_gen++; }
    while (G.bar_gen == g) yield_wait();
}

void launch(dim3 grid, dim3 block, size_t, hipStream_t, const std::function<void()> &f) {
    const int n = (int)(block.x * block.y * block.z);
    if (n <= 0 || n > kMaxItems) { std::fprintf(stderr, "lfemu: workgroup of %d items\n", n); std::abort(); }
    static Group *grp = new Group();
    Group &G = *grp;
    Group *prev = g_grp;
    g_grp = &G;
    g_bdim = Idx{block.x, block.y, block.z};
    g_gdim = Idx{grid.x, grid.y, grid.z};
    for (unsigned bz = 0; bz < grid.z; ++bz)
        for (unsigned by = 0; by < grid.y; ++by)
            for (unsigned bx = 0; bx < grid.x; ++bx) {
                g_block = Idx{bx, by, bz};
                G.n = n;
                G.nwave = (n + 63) / 64;
                G.body = &f;
                G.bar_active = n; G.bar_arrive = 0;
                for (int w = 0; w < G.nwave; ++w) {
                    Wave &W = G.wave[w];
                    W.nactive = std::min(64, n - w * 64);
                    W.arrive[0] = W.arrive[1] = 0;
                    std::memset(W.posted, 0, sizeof(W.posted));
                }
                for (int i = 0; i < n; ++i) {
                    Fiber &F = G.fib[i];
                    F.item.tid = Idx{(unsigned)(i % block.x), (unsigned)((i / block.x) % block.y),
                                     (unsigned)(i / (block.x * block.y))};
                    F.done = false;
                    F.fake = nullptr;
                    F.stack = stack_pool(i);
                    // initial frame: six callee-saved registers, the entry, a fake return
                    uintptr_t top = ((uintptr_t)(F.stack + kStack)) & ~(uintptr_t)15;
                    void **sp = (void **)(top - 8);
                    *sp = nullptr;                                 // fake return address
                    *--sp = (void *)&fiber_main;
                    for (int r = 0; r < 6; ++r) *--sp = nullptr;
                    F.sp = sp;
                }
                int left = n;
                while (left > 0) {
                    G.progress = false;
                    for (int i = 0; i < n; ++i) {
                        Fiber &F = G.fib[i];
                        if (F.done) continue;
                        G.current = i;
                        void *fake = nullptr;
#if LFEMU_ASAN
                        __sanitizer_start_switch_fiber(&fake, F.stack, kStack);
#endif
                        lfemu_switch(&G.sched_sp, F.sp);
#if LFEMU_ASAN
                        __sanitizer_finish_switch_fiber(fake, nullptr, nullptr);
#endif
                        G.current = -1;
                        if (F.done) --left;
                    }
                    if (!G.progress && left > 0) {
                        std::fprintf(stderr, "lfemu: workgroup (%u,%u,%u) deadlocked: work-items wait "
                                     "at different cross-lane operations or barriers\n", bx, by, bz);
                        std::abort();
                    }
                }
            }
    g_grp = prev;
}

}  // namespace lfemu
