// gm_runtime.hip: device selection, stream, caching HBM arena, errors,
// per-kernel event timing and the device exclusive scan.
//
// Every host thread that calls the library gets its own non-blocking stream
// at default priority (the initialising thread's is created by gm_init), so
// concurrent callers -- the request server's workers -- overlap on the device.
// A caller may put its own stream in place of the thread's library stream
// (gm_set_stream), so the library's work is ordered with the caller's.
// The arena keeps freed blocks in size-keyed free lists (HBM is large;
// hipMalloc/hipFree are slow and synchronising), tagged with the stream that
// freed them and an event recorded there at the free: a block is reused at
// once on its own stream (stream order makes that safe), and on another
// stream after that stream is made to wait for the event on the device (no
// host synchronisation, so concurrent callers do not serialise).  Blocks go
// back to the driver on allocation failure or gm_arena_trim().

#include "gm_internal.h"
#include "gm_block.h"

#include <mutex>
#include <map>
#include <unordered_map>
#include <vector>
#include <string>
#include <cstdio>
#include <cstdarg>
#include <cstring>

namespace gm
{

static std::mutex g_mu;
static int g_init_rc = 1;             // 1: not yet initialised
static int g_device = -1;
static hipStream_t g_stream = nullptr;               // the initialising thread's stream
static thread_local hipStream_t t_stream = nullptr;  // this thread's library stream
static thread_local hipStream_t t_user = nullptr;    // a caller's stream (gm_set_stream)
static thread_local bool t_user_set = false;
static std::vector<hipStream_t> g_thread_streams;    // created for other threads
static int g_cu = 0;
static int64_t g_hbm = 0;
static thread_local char g_err[512] = "";

int set_error(int code, const char *fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof(g_err), fmt, ap);
    va_end(ap);
    return code;
}

int check_hip(hipError_t e, const char *what)
{
    if (e == hipSuccess) return GM_SUCCESS;
    if (e == hipErrorOutOfMemory || e == hipErrorMemoryAllocation)
        return set_error(GM_OUT_OF_MEMORY, "%s: %s", what, hipGetErrorString(e));
    return set_error(GM_PANIC, "%s: %s", what, hipGetErrorString(e));
}

hipStream_t stream()
{
    if (t_user_set) return t_user;
    if (t_stream) return t_stream;
    hipStream_t s = nullptr;
    if (hipStreamCreateWithFlags(&s, hipStreamNonBlocking) != hipSuccess) return g_stream;
    {
        std::lock_guard<std::mutex> lk(g_mu);
        g_thread_streams.push_back(s);
    }
    t_stream = s;
    return s;
}
int cu_count() { return g_cu; }
int64_t device_hbm_bytes() { return g_hbm; }

int ensure_init()
{
    if (g_init_rc == GM_SUCCESS) return GM_SUCCESS;
    return gm_init();
}

bool is_initialised() { return g_init_rc == GM_SUCCESS; }

size_t type_size(int32_t type)
{
    switch (type)
    {
        case GM_BOOL: case GM_INT8: case GM_UINT8: return 1;
        case GM_INT16: case GM_UINT16: return 2;
        case GM_INT32: case GM_UINT32: case GM_FP32: return 4;
        case GM_INT64: case GM_UINT64: case GM_FP64: return 8;
        default: return 0;
    }
}

bool semiring_supported(const gm_semiring *sr)
{
    if (!sr) return false;
#define GM_SR_MATCH(ADD, MULT, ZC, Z) \
    if (sr->add_op == ADD && sr->mult_op == MULT && sr->ztype == ZC) return true;
    GM_SEMIRING_LIST(GM_SR_MATCH)
#undef GM_SR_MATCH
    return false;
}

bool operand_ok(const gm_semiring *sr, const gm_csr *X)
{
    if (sr->mult_op == GM_MULT_PAIR || sr->mult_op == GM_MULT_K) return true;
    return X->x != nullptr && X->type == sr->ztype;
}

bool csr_ok(const gm_csr *X)
{
    return X && X->p && X->nrows >= 0 && X->ncols >= 0 && (X->j || X->nrows == 0 || true);
}

//------------------------------------------------------------------------------
// arena
//------------------------------------------------------------------------------

struct FreeBlock { void *p; hipStream_t s; hipEvent_t ev; };
static std::multimap<size_t, FreeBlock> g_free;     // size -> block, freeing stream
static std::unordered_map<void *, size_t> g_live;   // block -> size
static size_t g_cached_bytes = 0, g_live_bytes = 0;
static int64_t g_fail_after = -1;

static size_t round_size(size_t b)
{
    if (b < 256) return 256;
    if (b < ((size_t) 1 << 20))
    {
        size_t r = 256;
        while (r < b) r <<= 1;
        return r;
    }
    const size_t g = (size_t) 2 << 20;
    return (b + g - 1) / g * g;
}

static std::vector<hipEvent_t> g_free_events;        // recycled free-block events

static void trim_locked()
{
    for (auto &kv : g_free)
    {
        (void) hipFree(kv.second.p);
        if (kv.second.ev) g_free_events.push_back(kv.second.ev);
    }
    g_free.clear();
    g_cached_bytes = 0;
}

//------------------------------------------------------------------------------
// profiling
//------------------------------------------------------------------------------

static bool g_prof = false;
struct ProfEv { int name_id; hipEvent_t a, b; };
static std::vector<ProfEv> g_prof_evs;
static std::vector<std::string> g_prof_names;
static std::vector<hipEvent_t> g_event_pool;

static hipEvent_t get_event()
{
    if (!g_event_pool.empty())
    {
        hipEvent_t e = g_event_pool.back();
        g_event_pool.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    (void) hipEventCreate(&e);
    return e;
}

ProfScope::ProfScope(const char *name) : slot(-1)
{
    if (!g_prof) return;
    hipStream_t st = stream();      // before the lock: it may create the stream
    std::lock_guard<std::mutex> lk(g_mu);
    int id = -1;
    for (size_t t = 0; t < g_prof_names.size(); t++)
        if (g_prof_names[t] == name) { id = (int) t; break; }
    if (id < 0) { g_prof_names.emplace_back(name); id = (int) g_prof_names.size() - 1; }
    ProfEv ev{id, get_event(), get_event()};
    (void) hipEventRecord(ev.a, st);
    g_prof_evs.push_back(ev);
    slot = (int) g_prof_evs.size() - 1;
}

ProfScope::~ProfScope()
{
    if (slot < 0) return;
    hipStream_t st = stream();
    std::lock_guard<std::mutex> lk(g_mu);
    (void) hipEventRecord(g_prof_evs[slot].b, st);
}

static thread_local void *g_pinned = nullptr;
static thread_local size_t g_pinned_bytes = 0;

void *pinned_scratch(size_t bytes)
{
    if (bytes <= g_pinned_bytes) return g_pinned;
    if (g_pinned) (void) hipHostFree(g_pinned);
    g_pinned = nullptr;
    g_pinned_bytes = 0;
    size_t want = bytes < 65536 ? 65536 : bytes;
    if (hipHostMalloc(&g_pinned, want, hipHostMallocDefault) != hipSuccess) return nullptr;
    g_pinned_bytes = want;
    return g_pinned;
}

//------------------------------------------------------------------------------
// exclusive scan (int64): reduce tiles -> scan tile sums -> rescan tiles
//------------------------------------------------------------------------------

constexpr int SCAN_NT = 256;
constexpr int SCAN_IPT = 8;
constexpr int SCAN_TILE = SCAN_NT * SCAN_IPT;

__global__ __launch_bounds__(SCAN_NT) void k_scan_reduce(const int64_t *__restrict__ in,
                                                         int64_t n, int64_t *__restrict__ sums)
{
    __shared__ int64_t lds[SCAN_NT / 64 + 1];
    const int64_t base = (int64_t) blockIdx.x * SCAN_TILE;
    int64_t s = 0;
    #pragma unroll
    for (int t = 0; t < SCAN_IPT; t++)
    {
        int64_t q = base + (int64_t) t * SCAN_NT + threadIdx.x;
        if (q < n) s += in[q];
    }
    int64_t tot = block_sum<int64_t, SCAN_NT>(s, lds);
    if (threadIdx.x == 0) sums[blockIdx.x] = tot;
}

__global__ __launch_bounds__(SCAN_NT) void k_scan_apply(const int64_t *in, int64_t n,
                                                        const int64_t *__restrict__ offs,
                                                        int64_t *out)
{
    __shared__ int64_t lds[SCAN_NT / 64 + 1];
    const int64_t base = (int64_t) blockIdx.x * SCAN_TILE + (int64_t) threadIdx.x * SCAN_IPT;
    int64_t v[SCAN_IPT];
    int64_t s = 0;
    #pragma unroll
    for (int t = 0; t < SCAN_IPT; t++)
    {
        int64_t q = base + t;
        v[t] = (q < n) ? in[q] : 0;
        s += v[t];
    }
    int64_t tot;
    int64_t ex = block_exclusive_scan<int64_t, SCAN_NT>(s, lds, &tot);
    int64_t run = ex + (offs ? offs[blockIdx.x] : 0);
    #pragma unroll
    for (int t = 0; t < SCAN_IPT; t++)
    {
        int64_t q = base + t;
        if (q < n) out[q] = run;
        run += v[t];
    }
    if (blockIdx.x == gridDim.x - 1 && threadIdx.x == SCAN_NT - 1) out[n] = run;
}

int scan_exclusive_i64(const int64_t *d_in, int64_t *d_out, int64_t n)
{
    if (n <= 0)
    {
        GM_HIP(hipMemsetAsync(d_out, 0, sizeof(int64_t), stream()), "scan memset");
        return GM_SUCCESS;
    }
    int64_t nt = (n + SCAN_TILE - 1) / SCAN_TILE;
    if (nt == 1)
    {
        ProfScope ps("scan_apply");
        k_scan_apply<<<1, SCAN_NT, 0, stream()>>>(d_in, n, nullptr, d_out);
        return check_hip(hipGetLastError(), "k_scan_apply");
    }
    DevBuf sums;
    if (!sums.alloc(sizeof(int64_t) * (nt + 1)))
        return set_error(GM_OUT_OF_MEMORY, "scan workspace");
    {
        ProfScope ps("scan_reduce");
        k_scan_reduce<<<(unsigned) nt, SCAN_NT, 0, stream()>>>(d_in, n, sums.as<int64_t>());
        GM_HIP(hipGetLastError(), "k_scan_reduce");
    }
    GM_TRY(scan_exclusive_i64(sums.as<int64_t>(), sums.as<int64_t>(), nt));
    {
        ProfScope ps("scan_apply");
        k_scan_apply<<<(unsigned) nt, SCAN_NT, 0, stream()>>>(d_in, n, sums.as<int64_t>(), d_out);
        GM_HIP(hipGetLastError(), "k_scan_apply");
    }
    return GM_SUCCESS;
}

} // namespace gm

using namespace gm;

//------------------------------------------------------------------------------
// C ABI
//------------------------------------------------------------------------------

namespace gm
{
static gm_dot3_params g_d3_tuned = {0, 0, 0, 0, 0, 0};
static gm_saxpy3_params g_sx_tuned = {0, 0, 0, 0};

gm_dot3_params dot3_tuned()
{
    std::lock_guard<std::mutex> lk(g_mu);
    return g_d3_tuned;
}

gm_saxpy3_params saxpy3_tuned()
{
    std::lock_guard<std::mutex> lk(g_mu);
    return g_sx_tuned;
}
} // namespace gm

extern "C" {

int gm_dot3_set_defaults(const gm_dot3_params *p)
{
    std::lock_guard<std::mutex> lk(g_mu);
    g_d3_tuned = p ? *p : gm_dot3_params{0, 0, 0, 0, 0, 0};
    return GM_SUCCESS;
}

int gm_saxpy3_set_defaults(const gm_saxpy3_params *p)
{
    std::lock_guard<std::mutex> lk(g_mu);
    g_sx_tuned = p ? *p : gm_saxpy3_params{0, 0, 0, 0};
    return GM_SUCCESS;
}

int gm_load_tuning(const char *path)
{
    if (!path) return set_error(GM_INVALID_VALUE, "gm_load_tuning: NULL path");
    FILE *f = fopen(path, "r");
    if (!f) return set_error(GM_INVALID_VALUE, "gm_load_tuning: cannot open %s", path);
    gm_dot3_params d = dot3_tuned();
    gm_saxpy3_params x = saxpy3_tuned();
    char line[256];
    int rc = GM_SUCCESS;
    while (fgets(line, sizeof(line), f))
    {
        char key[128];
        long long v;
        if (line[0] == '#' || line[0] == '\n') continue;
        if (sscanf(line, "%127s %lld", key, &v) != 2)
        {
            rc = set_error(GM_INVALID_VALUE, "gm_load_tuning: bad line in %s: %s", path, line);
            break;
        }
        const std::string k(key);
        if (k == "dot3.thread_max") d.thread_max = v;
        else if (k == "dot3.team16_max") d.team16_max = v;
        else if (k == "dot3.wave_max") d.wave_max = v;
        else if (k == "dot3.bitmap_min_len") d.bitmap_min_len = v;
        else if (k == "dot3.order") d.order = v;
        else if (k == "dot3.row_min") d.row_min = v;
        else if (k == "saxpy3.wave_small_max") x.wave_small_max = v;
        else if (k == "saxpy3.wave_large_max") x.wave_large_max = v;
        else if (k == "saxpy3.block_max") x.block_max = v;
        else if (k == "saxpy3.mode") x.mode = v;
        else
        {
            rc = set_error(GM_INVALID_VALUE, "gm_load_tuning: unknown key %s in %s", key, path);
            break;
        }
    }
    fclose(f);
    if (rc != GM_SUCCESS) return rc;
    gm_dot3_set_defaults(&d);
    gm_saxpy3_set_defaults(&x);
    return GM_SUCCESS;
}

static int gm_init_device(void)
{
    std::lock_guard<std::mutex> lk(g_mu);
    if (g_init_rc == GM_SUCCESS) return GM_SUCCESS;
    int ndev = 0;
    hipError_t e = hipGetDeviceCount(&ndev);
    if (e != hipSuccess || ndev <= 0)
    {
        g_init_rc = set_error(GM_PANIC, "gm_init: no HIP device available (%s)",
                              hipGetErrorString(e));
        return g_init_rc;
    }
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) dev = 0;
    if (hipSetDevice(dev) != hipSuccess)
    {
        g_init_rc = set_error(GM_PANIC, "gm_init: hipSetDevice(%d) failed", dev);
        return g_init_rc;
    }
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, dev) != hipSuccess)
    {
        g_init_rc = set_error(GM_PANIC, "gm_init: hipGetDeviceProperties failed");
        return g_init_rc;
    }
    g_cu = prop.multiProcessorCount;
    g_hbm = (int64_t) prop.totalGlobalMem;
    if (hipStreamCreateWithFlags(&g_stream, hipStreamNonBlocking) != hipSuccess)
    {
        g_init_rc = set_error(GM_PANIC, "gm_init: stream creation failed");
        return g_init_rc;
    }
    // the allocation-free entry points' scratch: its device address, looked
    // up once with this device current
    const int src = reduce_scratch_setup();
    if (src != GM_SUCCESS)
    {
        (void) hipStreamDestroy(g_stream);
        g_stream = nullptr;
        g_init_rc = src;
        return g_init_rc;
    }
    g_device = dev;
    t_stream = g_stream;
    g_init_rc = GM_SUCCESS;
    return GM_SUCCESS;
}

// device set-up, then (once per process) the tuning file named by the
// environment variable GM_TUNING_FILE, if set
int gm_init(void)
{
    const int rc = gm_init_device();
    if (rc == GM_SUCCESS)
    {
        static std::once_flag once;
        std::call_once(once, [] {
            const char *path = getenv("GM_TUNING_FILE");
            if (path && *path) (void) gm_load_tuning(path);
        });
    }
    return rc;
}

void gm_finalize(void)
{
    std::lock_guard<std::mutex> lk(g_mu);
    if (g_init_rc != GM_SUCCESS) return;
    (void) hipStreamSynchronize(g_stream);
    trim_locked();
    for (auto e : g_free_events) (void) hipEventDestroy(e);
    g_free_events.clear();
    for (auto &kv : g_live) (void) hipFree(kv.first);
    g_live.clear();
    g_live_bytes = 0;
    for (auto &ev : g_prof_evs) { (void) hipEventDestroy(ev.a); (void) hipEventDestroy(ev.b); }
    g_prof_evs.clear();
    for (auto e : g_event_pool) (void) hipEventDestroy(e);
    g_event_pool.clear();
    if (g_pinned) (void) hipHostFree(g_pinned);
    g_pinned = nullptr;
    g_pinned_bytes = 0;
    for (auto s2 : g_thread_streams) { (void) hipStreamSynchronize(s2); (void) hipStreamDestroy(s2); }
    g_thread_streams.clear();
    (void) hipStreamDestroy(g_stream);
    g_stream = nullptr;
    t_stream = nullptr;
    g_init_rc = 1;
}

const char *gm_last_error(void) { return g_err; }

int gm_set_stream(void *s)
{
    GM_TRY(ensure_init());
    if (s == nullptr)
    {
        t_user = nullptr;
        t_user_set = false;
        return GM_SUCCESS;
    }
    hipStream_t hs = (hipStream_t) s;
    // reject a handle the runtime does not know (query returns its state)
    const hipError_t q = hipStreamQuery(hs);
    if (q != hipSuccess && q != hipErrorNotReady)
    {
        (void) hipGetLastError();
        return set_error(GM_INVALID_VALUE, "gm_set_stream: not a stream of this device (%s)",
                         hipGetErrorString(q));
    }
    t_user = hs;
    t_user_set = true;
    return GM_SUCCESS;
}

void *gm_get_stream(void) { return (void *) stream(); }

int gm_device_count(void)
{
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return 0;
    return n;
}

int gm_get_device_info(gm_device_info *info)
{
    GM_TRY(ensure_init());
    if (!info) return GM_INVALID_VALUE;
    hipDeviceProp_t prop;
    GM_HIP(hipGetDeviceProperties(&prop, g_device), "hipGetDeviceProperties");
    memset(info, 0, sizeof(*info));
    snprintf(info->name, sizeof(info->name), "%s", prop.name);
    snprintf(info->arch, sizeof(info->arch), "%s", prop.gcnArchName);
    info->cu_count = prop.multiProcessorCount;
    info->lds_per_block = (int32_t) prop.sharedMemPerBlock;
    info->hbm_bytes = (int64_t) prop.totalGlobalMem;
    // memoryClockRate is in kHz; double data rate
    info->hbm_peak_gbs = 2.0 * (double) prop.memoryClockRate * 1e3 *
                         ((double) prop.memoryBusWidth / 8.0) / 1e9;
    return GM_SUCCESS;
}

void *gm_malloc(size_t bytes)
{
    if (ensure_init() != GM_SUCCESS) return nullptr;
    std::lock_guard<std::mutex> lk(g_mu);
    if (g_fail_after >= 0)
    {
        if (g_fail_after == 0)
        {
            g_fail_after = -1;
            set_error(GM_OUT_OF_MEMORY, "gm_malloc: injected allocation failure");
            return nullptr;
        }
        g_fail_after--;
    }
    size_t r = round_size(bytes);
    hipStream_t me = stream();
    auto best = g_free.end();
    for (auto it = g_free.lower_bound(r); it != g_free.end() && it->first <= 2 * r; ++it)
    {
        if (it->second.s == me) { best = it; break; }
        if (best == g_free.end()) best = it;
    }
    if (best != g_free.end())
    {
        void *p = best->second.p;
        const size_t sz = best->first;
        // a block freed on another stream may still be in use there: this
        // stream waits on the device for the work queued before the free
        if (best->second.s != me && best->second.ev) (void) hipStreamWaitEvent(me, best->second.ev, 0);
        if (best->second.ev) g_free_events.push_back(best->second.ev);
        g_free.erase(best);
        g_cached_bytes -= sz;
        g_live[p] = sz;
        g_live_bytes += sz;
        return p;
    }
    void *p = nullptr;
    hipError_t e = hipMalloc(&p, r);
    if (e != hipSuccess)
    {
        (void) hipGetLastError();
        (void) hipDeviceSynchronize();
        trim_locked();
        e = hipMalloc(&p, r);
        if (e != hipSuccess)
        {
            (void) hipGetLastError();
            set_error(GM_OUT_OF_MEMORY, "gm_malloc: %zu bytes: %s", r, hipGetErrorString(e));
            return nullptr;
        }
    }
    g_live[p] = r;
    g_live_bytes += r;
    return p;
}

void gm_free(void *p)
{
    if (!p) return;
    std::lock_guard<std::mutex> lk(g_mu);
    auto it = g_live.find(p);
    if (it == g_live.end()) return;   // not ours: ignore
    size_t sz = it->second;
    g_live.erase(it);
    g_live_bytes -= sz;
    hipStream_t st = stream();
    hipEvent_t ev = nullptr;
    if (!g_free_events.empty()) { ev = g_free_events.back(); g_free_events.pop_back(); }
    else if (hipEventCreateWithFlags(&ev, hipEventDisableTiming) != hipSuccess) ev = nullptr;
    if (ev && hipEventRecord(ev, st) != hipSuccess)
    {
        (void) hipGetLastError();
        (void) hipEventDestroy(ev);
        ev = nullptr;
    }
    if (!ev) (void) hipStreamSynchronize(st);     // no event: the block is idle once st drains
    g_free.emplace(sz, FreeBlock{p, st, ev});
    g_cached_bytes += sz;
}

int gm_arena_trim(void)
{
    if (g_init_rc != GM_SUCCESS) return GM_SUCCESS;
    std::lock_guard<std::mutex> lk(g_mu);
    (void) hipDeviceSynchronize();
    trim_locked();
    return GM_SUCCESS;
}

int64_t gm_arena_bytes_in_use(void) { return (int64_t) g_live_bytes; }

int64_t gm_arena_bytes_cached(void)
{
    std::lock_guard<std::mutex> lk(g_mu);
    return (int64_t) g_cached_bytes;
}

void gm_debug_fail_after(int64_t n)
{
    std::lock_guard<std::mutex> lk(g_mu);
    g_fail_after = n;
}

int gm_memcpy_h2d(void *dst, const void *src, size_t bytes)
{
    GM_TRY(ensure_init());
    if (bytes == 0) return GM_SUCCESS;
    GM_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, stream()), "h2d");
    GM_HIP(hipStreamSynchronize(stream()), "h2d sync");
    return GM_SUCCESS;
}

int gm_memcpy_d2h(void *dst, const void *src, size_t bytes)
{
    GM_TRY(ensure_init());
    if (bytes == 0) return GM_SUCCESS;
    GM_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, stream()), "d2h");
    GM_HIP(hipStreamSynchronize(stream()), "d2h sync");
    return GM_SUCCESS;
}

int gm_memcpy_d2d(void *dst, const void *src, size_t bytes)
{
    GM_TRY(ensure_init());
    if (bytes == 0) return GM_SUCCESS;
    GM_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, stream()), "d2d");
    return GM_SUCCESS;
}

int gm_memset(void *dst, int value, size_t bytes)
{
    GM_TRY(ensure_init());
    if (bytes == 0) return GM_SUCCESS;
    GM_HIP(hipMemsetAsync(dst, value, bytes, stream()), "memset");
    return GM_SUCCESS;
}

int gm_synchronize(void)
{
    GM_TRY(ensure_init());
    GM_HIP(hipStreamSynchronize(stream()), "synchronize");
    return GM_SUCCESS;
}

void gm_prof_enable(int on) { g_prof = (on != 0); }

void gm_prof_reset(void)
{
    if (g_init_rc == GM_SUCCESS) (void) hipDeviceSynchronize();
    std::lock_guard<std::mutex> lk(g_mu);
    for (auto &ev : g_prof_evs) { g_event_pool.push_back(ev.a); g_event_pool.push_back(ev.b); }
    g_prof_evs.clear();
    g_prof_names.clear();
}

int gm_prof_get(gm_prof_record *rec, int cap)
{
    if (g_init_rc == GM_SUCCESS) (void) hipDeviceSynchronize();
    std::lock_guard<std::mutex> lk(g_mu);
    std::vector<double> tot(g_prof_names.size(), 0.0);
    std::vector<int64_t> cnt(g_prof_names.size(), 0);
    for (auto &ev : g_prof_evs)
    {
        float ms = 0.f;
        if (hipEventElapsedTime(&ms, ev.a, ev.b) == hipSuccess)
        {
            tot[ev.name_id] += ms;
            cnt[ev.name_id]++;
        }
    }
    int n = (int) g_prof_names.size();
    for (int t = 0; t < n && t < cap; t++)
    {
        memset(&rec[t], 0, sizeof(rec[t]));
        snprintf(rec[t].name, sizeof(rec[t].name), "%s", g_prof_names[t].c_str());
        rec[t].launches = cnt[t];
        rec[t].total_ms = tot[t];
    }
    return n;
}

} // extern "C"
