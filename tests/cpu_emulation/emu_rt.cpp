// Fiber-based SIMT emulation of HIP kernels on the CPU (lanes as coroutines,
// wave intrinsics as rendezvous), plus a malloc-backed HIP runtime.  Built with
// AddressSanitizer so that every kernel access outside an allocation is caught.
#include "hip/hip_runtime.h"
#include <csetjmp>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <ucontext.h>
#include <vector>
#include <mutex>
#include <atomic>
#include <condition_variable>
#include <thread>
#include <unordered_map>
#include <chrono>
#include <string>
#include <deque>
#include <map>
#include <set>
#include <sanitizer/asan_interface.h>
#include <sanitizer/common_interface_defs.h>

// AddressSanitizer is told about every switch between lane stacks; a build
// without it (the ThreadSanitizer build, which instruments the host runtime
// only) switches without the annotations
#if defined(__SANITIZE_ADDRESS__)
#define EMU_FIBER_START(fake, bottom, size) __sanitizer_start_switch_fiber(fake, bottom, size)
#define EMU_FIBER_FINISH(fake, bottom, size) __sanitizer_finish_switch_fiber(fake, bottom, size)
#else
#define EMU_FIBER_START(fake, bottom, size) ((void)(fake), (void)(bottom), (void)(size))
#define EMU_FIBER_FINISH(fake, bottom, size) ((void)(fake), (void)(bottom), (void)(size))
#endif

namespace emu {

// A lane is a fiber that lives for the whole process: started once (the only
// ucontext switch, which costs signal-mask system calls), it runs the kernel
// body for a workgroup, parks at `idle`, and is restarted there for the next
// workgroup with a plain jump (jump_to).
struct fiber {
    ucontext_t uctx;
    void *jb[5], *idle[5]; // __builtin_setjmp buffers
    bool alive = false;   // started once; parked at `idle` between workgroups
    bool started = false, done = false;
    idx3 tid;
    int wave = 0, lane = 0;
    char *stack = nullptr;
    void *fake = nullptr;
};

struct rendezvous {
    int live = 0, arrived = 0, site = 0;
    const char *what = "";
    uint64_t gen = 0;
    uint64_t slots[64], result[64];
    bool exited[64];
};

// A workgroup runs on one host thread; with EMU_THREADS=N a kernel's
// workgroups are spread over N host threads that run at the same time, so all
// scheduling state is per thread.  The kernel being run (body, arguments,
// shape) is shared and read-only while its workgroups run.
static const size_t kStack = 512 * 1024;
static thread_local fiber *fibers = nullptr; // 1024 lanes, allocated per thread
static thread_local fiber *cur = nullptr;
static thread_local void *sched_jb[5];

// Context switches between the scheduler and the lanes' stacks use GCC's
// __builtin_setjmp / __builtin_longjmp: AddressSanitizer intercepts the
// library longjmp and unpoisons the current stack from there on, which here
// is a live lane stack, not frames being unwound (and costs a shadow fill
// per switch).  The longjmp sits in a function of its own, as GCC requires.
[[gnu::noinline]] static void jump_to(void **buf) { __builtin_longjmp(buf, 1); }
static thread_local void *sched_fake = nullptr;
static thread_local const void *sched_bottom = nullptr;
static thread_local size_t sched_size = 0;
static thread_local rendezvous waves[16];
static thread_local rendezvous block;
static thread_local idx3 g_block;
static idx3 g_bdim, g_gdim;
static void (*g_body)(void *) = nullptr;
static void *g_ctx = nullptr;
static const char *g_name = "";
static thread_local unsigned long long progress = 0;
static thread_local idx3 host_idx = {0, 0, 0};
static thread_local unsigned long long workgroups_started = 0;

static bool env_on(const char *name) {
    const char *v = getenv(name);
    return v && *v && *v != '0';
}

bool redzones() {
    static const bool on = env_on("EMU_REDZONES");
    return on;
}

static bool poison_buffers() {
    static const bool on = env_on("EMU_POISON");
    return on;
}

void poison(void *p, size_t n) { ASAN_POISON_MEMORY_REGION(p, n); }
void unpoison(void *p, size_t n) { ASAN_UNPOISON_MEMORY_REGION(p, n); }

// A __shared__ object (function-static here) refilled with 0xA5 the first time
// a lane of each workgroup reaches its declaration.
void shared_region(void *p, size_t n) {
    static thread_local std::unordered_map<void *, unsigned long long> filled;
    if (!poison_buffers())
        return;
    unsigned long long &at = filled[p];
    if (at != workgroups_started) {
        at = workgroups_started;
        memset(p, 0xA5, n);
    }
}

idx3 &thread_idx() { return cur ? cur->tid : host_idx; }
idx3 &block_idx() { return g_block; }
idx3 &block_dim() { return g_bdim; }
idx3 &grid_dim() { return g_gdim; }
int lane() { return cur->lane; }

void fail(const char *msg) {
    fprintf(stderr, "EMU FAIL in %s block (%u,%u,%u) thread %u: %s\n", g_name, g_block.x, g_block.y,
            g_block.z, cur ? cur->tid.x : 0u, msg);
    abort();
}

static void to_scheduler() {
    fiber *f = cur;
    if (!__builtin_setjmp(f->jb)) {
        EMU_FIBER_START(&f->fake, sched_bottom, sched_size);
        jump_to(sched_jb);
    }
    EMU_FIBER_FINISH(f->fake, nullptr, nullptr);
}

static void complete(rendezvous &r) {
    for (int l = 0; l < 64; ++l)
        r.result[l] = r.exited[l] ? 0 : r.slots[l];
    r.arrived = 0;
    r.gen++;
    progress++;
}

const uint64_t *wave_exchange(uint64_t v, int site, const char *what) {
    fiber *f = cur;
    rendezvous &w = waves[f->wave];
    if (w.arrived == 0) {
        w.site = site;
        w.what = what;
    } else if (w.site != site) {
        char buf[256];
        snprintf(buf, sizeof buf, "divergent wave op: line %d (%s) vs line %d (%s)", site, what,
                 w.site, w.what);
        fail(buf);
    }
    w.slots[f->lane] = v;
    w.arrived++;
    progress++;
    uint64_t my = w.gen;
    if (w.arrived == w.live)
        complete(w);
    while (w.gen == my)
        to_scheduler();
    return w.result;
}

void wave_barrier(int site) { wave_exchange(0, site, "wave_sync"); }

void block_barrier(int site) {
    if (block.arrived == 0)
        block.site = site;
    else if (block.site != site) {
        char buf[128];
        snprintf(buf, sizeof buf, "divergent __syncthreads: line %d vs line %d", site, block.site);
        fail(buf);
    }
    block.arrived++;
    progress++;
    uint64_t my = block.gen;
    if (block.arrived == block.live)
        complete(block);
    while (block.gen == my)
        to_scheduler();
}

static void fiber_entry() {
    EMU_FIBER_FINISH(nullptr, &sched_bottom, &sched_size);
    for (;;) {
        fiber *f = cur;
        g_body(g_ctx);
        f->done = true;
        progress++;
        rendezvous &w = waves[f->wave];
        w.live--;
        w.exited[f->lane] = true;
        if (w.arrived && w.arrived == w.live)
            complete(w);
        block.live--;
        if (block.arrived && block.arrived == block.live)
            complete(block);
        if (!__builtin_setjmp(f->idle)) { // park until the next workgroup
            EMU_FIBER_START(&f->fake, sched_bottom, sched_size);
            jump_to(sched_jb);
        }
        EMU_FIBER_FINISH(cur->fake, &sched_bottom, &sched_size);
    }
}

static void run_block(unsigned bx, unsigned by, unsigned bz) {
    if (!fibers)
        fibers = new fiber[1024];
    g_block = {bx, by, bz};
    workgroups_started++;
    unsigned n = g_bdim.x * g_bdim.y * g_bdim.z;
    int nwaves = (int)((n + 63) / 64);
    for (int k = 0; k < nwaves; ++k) {
        waves[k] = rendezvous();
        waves[k].live = (int)(n - 64 * k < 64 ? n - 64 * k : 64);
        for (int l = 0; l < 64; ++l)
            waves[k].exited[l] = l >= waves[k].live;
    }
    block = rendezvous();
    block.live = (int)n;
    for (int l = 0; l < 64; ++l)
        block.exited[l] = false;
    for (unsigned t = 0; t < n; ++t) {
        fiber &f = fibers[t];
        if (!f.stack)
            f.stack = (char *)malloc(kStack);
        f.started = false;
        f.done = false;
        f.tid = {t % g_bdim.x, (t / g_bdim.x) % g_bdim.y, t / (g_bdim.x * g_bdim.y)};
        f.wave = (int)(t / 64);
        f.lane = (int)(t % 64);
        f.fake = nullptr;
    }
    unsigned remaining = n;
    while (remaining) {
        unsigned long long before = progress;
        for (unsigned t = 0; t < n; ++t) {
            fiber &f = fibers[t];
            if (f.done)
                continue;
            cur = &f;
            if (!__builtin_setjmp(sched_jb)) {
                if (!f.started && f.alive) { // restart a parked lane
                    f.started = true;
                    EMU_FIBER_START(&sched_fake, f.stack, kStack);
                    jump_to(f.idle);
                } else if (!f.started) {
                    f.started = true;
                    f.alive = true;
                    getcontext(&f.uctx);
                    f.uctx.uc_stack.ss_sp = f.stack;
                    f.uctx.uc_stack.ss_size = kStack;
                    f.uctx.uc_link = nullptr;
                    makecontext(&f.uctx, fiber_entry, 0);
                    EMU_FIBER_START(&sched_fake, f.stack, kStack);
                    ucontext_t dummy;
                    swapcontext(&dummy, &f.uctx);
                } else {
                    EMU_FIBER_START(&sched_fake, f.stack, kStack);
                    jump_to(f.jb);
                }
            }
            EMU_FIBER_FINISH(sched_fake, nullptr, nullptr);
            cur = nullptr;
            if (f.done)
                remaining--;
        }
        if (progress == before && remaining) {
            char buf[160];
            snprintf(buf, sizeof buf, "deadlock: %u lanes blocked (block barrier %d arrived, wave0 %d arrived at line %d)",
                     remaining, block.arrived, waves[0].arrived, waves[0].site);
            fail(buf);
        }
    }
}

unsigned long long kernels_run = 0;

static std::recursive_mutex run_mutex; // server threads launch kernels too

// The workgroups of the kernel being run, in the order they are handed out.
static std::vector<size_t> g_order;
static std::atomic<size_t> g_next{0};

static void run_order_item(size_t b) {
    run_block((unsigned)(b % g_gdim.x), (unsigned)((b / g_gdim.x) % g_gdim.y),
              (unsigned)(b / ((size_t)g_gdim.x * g_gdim.y)));
}

static void drain() {
    for (;;) {
        size_t i = g_next.fetch_add(1);
        if (i >= g_order.size())
            return;
        run_order_item(g_order[i]);
    }
}

// EMU_THREADS=N: N - 1 pool threads and the launching thread take
// workgroups from one counter, so workgroups of a kernel run concurrently.
struct pool {
    std::mutex m;
    std::condition_variable go, done;
    unsigned long long generation = 0;
    int busy = 0;
    std::vector<std::thread> threads;
    explicit pool(int n) {
        for (int k = 0; k < n; ++k)
            threads.emplace_back([this] {
                unsigned long long seen = 0;
                for (;;) {
                    {
                        std::unique_lock<std::mutex> lock(m);
                        go.wait(lock, [&] { return generation != seen; });
                        seen = generation;
                    }
                    drain();
                    std::lock_guard<std::mutex> lock(m);
                    if (--busy == 0)
                        done.notify_all();
                }
            });
        for (std::thread &t : threads)
            t.detach();
    }
    void run() {
        {
            std::lock_guard<std::mutex> lock(m);
            busy = (int)threads.size();
            generation++;
        }
        go.notify_all();
        drain();
        std::unique_lock<std::mutex> lock(m);
        done.wait(lock, [&] { return busy == 0; });
    }
};

static int emu_threads() {
    static const int n = [] {
        const char *v = getenv("EMU_THREADS");
        int k = v ? atoi(v) : 1;
        return k < 1 ? 1 : k > 64 ? 64 : k;
    }();
    return n;
}

// EMU_PROFILE: launches, workgroups and seconds per kernel, printed at exit.
struct kernel_profile {
    unsigned long long launches = 0, workgroups = 0;
    double seconds = 0;
};
static std::map<std::string, kernel_profile> *profiles;

static void print_profiles() {
    for (const auto &p : *profiles)
        fprintf(stderr, "[profile] %-40s %8llu launches %12llu workgroups %9.2f s\n",
                p.first.c_str(), p.second.launches, p.second.workgroups, p.second.seconds);
}

void run_kernel(const char *name, dim3 grid, dim3 blk, void (*body)(void *), void *ctx) {
    std::lock_guard<std::recursive_mutex> lock(run_mutex);
    static const bool profiling = env_on("EMU_PROFILE");
    auto started = std::chrono::steady_clock::now();
    if (profiling && !profiles) {
        profiles = new std::map<std::string, kernel_profile>;
        atexit(print_profiles);
    }
    g_name = name;
    g_body = body;
    g_ctx = ctx;
    g_bdim = {blk.x, blk.y, blk.z};
    g_gdim = {grid.x, grid.y, grid.z};
    if (blk.x * blk.y * blk.z > 1024)
        fail("block too large");
    size_t total = (size_t)grid.x * grid.y * grid.z;
    g_order.resize(total);
    for (size_t i = 0; i < total; ++i)
        g_order[i] = i;
    // EMU_SHUFFLE=<seed>: workgroups in a seeded random order (a GPU runs
    // them concurrently, in no fixed order), else in index order
    static const char *shuffle = getenv("EMU_SHUFFLE");
    if (shuffle && *shuffle) {
        unsigned long long st = strtoull(shuffle, nullptr, 0) ^ (kernels_run * 0x9E3779B97F4A7C15ull);
        for (size_t i = total; i > 1; --i) { // Fisher-Yates with SplitMix64
            unsigned long long z = (st += 0x9E3779B97F4A7C15ull);
            z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
            z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
            z ^= z >> 31;
            std::swap(g_order[i - 1], g_order[(size_t)(z % i)]);
        }
    }
    g_next = 0;
    int n = emu_threads();
    if (n > 1) {
        // never destroyed: its detached threads wait on its condition
        // variable until the process ends
        static pool *workers = new pool(n - 1);
        workers->run();
    } else {
        drain();
    }
    kernels_run++;
    if (profiling) {
        kernel_profile &p = (*profiles)[name];
        p.launches++;
        p.workgroups += total;
        p.seconds += std::chrono::duration<double>(std::chrono::steady_clock::now() - started).count();
    }
}

} // namespace emu

// ---- runtime --------------------------------------------------------------------
// EMU_ASYNC_STREAMS: a stream's kernels, device-to-device copies, copies from
// or to pinned host memory and fills are queued and run only when the host
// waits for them: hipStreamSynchronize, hipEventSynchronize, hipEventQuery of
// an event recorded after them (the first query of each recording reports
// hipErrorNotReady and runs nothing), hipStreamDestroy, or a device-wide wait
// (hipFree and hipHostFree synchronise the device, as HIP documents).  A copy
// between pageable host memory and the device first runs the stream's earlier
// work and then copies at once, as HIP documents for unpinned memory.  Host
// code that reads a result before waiting for it, or rewrites pinned staging
// that a queued copy has yet to read, then gets wrong pixels; a buffer freed
// while queued work still uses it is a use-after-free.  Without the mode,
// work runs when it is issued.
struct emu_stream {
    std::mutex m;
    std::deque<std::function<void()>> q;
    uint64_t queued = 0, done = 0;
};
struct emu_event {
    emu_stream *stream = nullptr;
    uint64_t at = 0; // the stream's queued count when recorded
    bool queried = false;
    double seconds = 0; // wall clock when the stream reached the event
};
static int g_err = hipSuccess;
static std::mutex streams_mutex;
static std::set<emu_stream *> live_streams;
static std::recursive_mutex device_mutex; // queued work runs one op at a time
static std::mutex pinned_mutex;
static std::map<uintptr_t, size_t> pinned; // hipHostMalloc blocks

static bool async_streams() {
    static const bool on = emu::env_on("EMU_ASYNC_STREAMS");
    return on;
}

static bool is_pinned(const void *p) {
    std::lock_guard<std::mutex> lock(pinned_mutex);
    uintptr_t a = (uintptr_t)p;
    auto it = pinned.upper_bound(a);
    if (it == pinned.begin())
        return false;
    --it;
    return a < it->first + it->second;
}

// Runs stream st's queued work until `upto` items of it have run.
static void drain_stream(emu_stream *st, uint64_t upto) {
    std::lock_guard<std::recursive_mutex> device(device_mutex);
    for (;;) {
        std::function<void()> op;
        {
            std::lock_guard<std::mutex> lock(st->m);
            if (st->done >= upto || st->q.empty())
                return;
            op = std::move(st->q.front());
            st->q.pop_front();
        }
        op();
        std::lock_guard<std::mutex> lock(st->m);
        st->done++;
    }
}

static void drain_all() {
    std::lock_guard<std::mutex> lock(streams_mutex);
    for (emu_stream *st : live_streams)
        drain_stream(st, UINT64_MAX);
}

void emu::launch(hipStream_t st, std::function<void()> op) {
    if (!async_streams() || !st) {
        op();
        return;
    }
    std::lock_guard<std::mutex> lock(st->m);
    st->q.push_back(std::move(op));
    st->queued++;
}

// EMU_STRICT_STREAMS: host-to-device copies on the NULL stream are held back
// until the NULL stream or the device is synchronised (or another NULL-stream
// copy, hipFree or hipHostFree orders after them), with their destination
// poisoned meanwhile.  The library's streams are non-blocking: work on them
// does not wait for the NULL stream, so a kernel that reads such a
// destination first aborts under AddressSanitizer.
struct held_copy {
    char *dst;
    std::vector<char> data;
};
static std::mutex held_mutex;
static std::vector<held_copy> held;

static bool strict_streams() {
    static const bool on = emu::env_on("EMU_STRICT_STREAMS");
    return on;
}

static void land_held_copies() {
    std::lock_guard<std::mutex> lock(held_mutex);
    for (held_copy &c : held) {
        ASAN_UNPOISON_MEMORY_REGION(c.dst, c.data.size());
        memmove(c.dst, c.data.data(), c.data.size());
    }
    held.clear();
}

hipError_t hipMalloc(void **p, size_t n) {
    *p = malloc(n ? n : 1);
    if (*p && emu::poison_buffers())
        memset(*p, 0xA5, n ? n : 1);
    return *p ? hipSuccess : hipErrorOutOfMemory;
}
hipError_t hipFree(void *p) {
    drain_all();
    land_held_copies();
    free(p);
    return hipSuccess;
}
hipError_t hipHostMalloc(void **p, size_t n, unsigned) {
    *p = malloc(n ? n : 1);
    if (*p) {
        std::lock_guard<std::mutex> lock(pinned_mutex);
        pinned[(uintptr_t)*p] = n ? n : 1;
    }
    return *p ? hipSuccess : hipErrorOutOfMemory;
}
hipError_t hipHostFree(void *p) {
    drain_all();
    land_held_copies();
    {
        std::lock_guard<std::mutex> lock(pinned_mutex);
        pinned.erase((uintptr_t)p);
    }
    free(p);
    return hipSuccess;
}
hipError_t hipMemcpy(void *d, const void *s, size_t n, hipMemcpyKind k) {
    land_held_copies(); // the NULL stream runs its copies in order
    if (!n)
        return hipSuccess;
    if (strict_streams() && k == hipMemcpyHostToDevice) {
        std::lock_guard<std::mutex> lock(held_mutex);
        held.push_back(held_copy{(char *)d, std::vector<char>((const char *)s, (const char *)s + n)});
        ASAN_POISON_MEMORY_REGION(d, n);
        return hipSuccess;
    }
    memmove(d, s, n);
    return hipSuccess;
}
hipError_t hipMemcpyAsync(void *d, const void *s, size_t n, hipMemcpyKind k, hipStream_t st) {
    if (!st)
        return hipMemcpy(d, s, n, k);
    if (!n)
        return hipSuccess;
    if (async_streams()) {
        bool queued = k == hipMemcpyDeviceToDevice ||
                      (k == hipMemcpyHostToDevice && is_pinned(s)) ||
                      (k == hipMemcpyDeviceToHost && is_pinned(d));
        if (queued) { // a pinned source is read when the copy runs
            emu::launch(st, [d, s, n]() { memmove(d, s, n); });
            return hipSuccess;
        }
        drain_stream(st, UINT64_MAX); // pageable memory: a synchronous copy
    }
    memmove(d, s, n);
    return hipSuccess;
}
hipError_t hipMemsetAsync(void *d, int v, size_t n, hipStream_t st) {
    emu::launch(st, [d, v, n]() { memset(d, v, n); });
    return hipSuccess;
}
hipError_t hipStreamSynchronize(hipStream_t st) {
    if (!st)
        land_held_copies();
    else
        drain_stream(st, UINT64_MAX);
    return hipSuccess;
}
hipError_t hipStreamCreate(hipStream_t *s) {
    *s = new emu_stream;
    std::lock_guard<std::mutex> lock(streams_mutex);
    live_streams.insert(*s);
    return hipSuccess;
}
hipError_t hipStreamCreateWithFlags(hipStream_t *s, unsigned) { return hipStreamCreate(s); }
hipError_t hipStreamDestroy(hipStream_t s) {
    drain_stream(s, UINT64_MAX);
    {
        std::lock_guard<std::mutex> lock(streams_mutex);
        live_streams.erase(s);
    }
    delete s;
    return hipSuccess;
}
hipError_t hipGetLastError() { int e = g_err; g_err = hipSuccess; return e; }
const char *hipGetErrorString(hipError_t) { return "emulated error"; }
hipError_t hipSetDevice(int) { return hipSuccess; }
hipError_t hipGetDeviceCount(int *n) { *n = 1; return hipSuccess; }
hipError_t hipEventCreate(hipEvent_t *e) { *e = new emu_event; return hipSuccess; }
hipError_t hipEventCreateWithFlags(hipEvent_t *e, unsigned) { return hipEventCreate(e); }
hipError_t hipEventDestroy(hipEvent_t e) { delete e; return hipSuccess; }
hipError_t hipEventRecord(hipEvent_t e, hipStream_t st) {
    e->stream = st;
    e->queried = false;
    // the time is taken in stream order (at once without EMU_ASYNC_STREAMS)
    emu::launch(st, [e]() {
        e->seconds = std::chrono::duration<double>(
                         std::chrono::steady_clock::now().time_since_epoch()).count();
    });
    if (st) {
        std::lock_guard<std::mutex> lock(st->m);
        e->at = st->queued;
    }
    return hipSuccess;
}
hipError_t hipEventQuery(hipEvent_t e) {
    if (!async_streams() || !e->stream)
        return hipSuccess;
    if (!e->queried) {
        e->queried = true;
        return hipErrorNotReady;
    }
    drain_stream(e->stream, e->at);
    return hipSuccess;
}
hipError_t hipEventSynchronize(hipEvent_t e) {
    if (e->stream)
        drain_stream(e->stream, e->at);
    return hipSuccess;
}
hipError_t hipEventElapsedTime(float *ms, hipEvent_t start, hipEvent_t stop) {
    *ms = (float)((stop->seconds - start->seconds) * 1000.0);
    return hipSuccess;
}
hipError_t hipStreamBeginCapture(hipStream_t, hipStreamCaptureMode) { return hipErrorNotSupported; }
hipError_t hipStreamEndCapture(hipStream_t, hipGraph_t *) { return hipErrorNotSupported; }
hipError_t hipGraphInstantiate(hipGraphExec_t *, hipGraph_t, void *, char *, size_t) { return hipErrorNotSupported; }
hipError_t hipGraphLaunch(hipGraphExec_t, hipStream_t) { return hipErrorNotSupported; }
hipError_t hipGraphExecDestroy(hipGraphExec_t) { return hipSuccess; }
hipError_t hipGraphDestroy(hipGraph_t) { return hipSuccess; }
