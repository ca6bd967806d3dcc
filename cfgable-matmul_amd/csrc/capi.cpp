// capi.cpp -- the extern "C" boundary of libcfgmm (declared in include/cfgmm.h).
//
// Host-pointer entry points mirror the reference's whole-problem functions
// (src/impl_matmul.rs:155-248, src/impl_matmul_non0tab.rs:210-326,
// src/impl_aodmao2rho.rs:171-261): same argument meaning, accumulate into the
// output, synchronous return.  Where the reference panics (slice index out of
// bounds, `unimplemented!` for transb with a mask, `assert!` on BLKSIZE) the
// C ABI returns a status code and records a message for cfgmm_last_error().
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <thread>
#include <utility>
#include <vector>

#include "internal.hpp"

namespace cfgmm {

// ---------------------------------------------------------------------------
// errors
// ---------------------------------------------------------------------------
static thread_local std::string g_last_error;

void set_error(const std::string &msg) { g_last_error = msg; }
void clear_error() { g_last_error.clear(); }
int fail(int status, const std::string &msg) {
    set_error(msg);
    return status;
}

int hip_status(hipError_t e, const char *what) {
    if (e == hipSuccess) return CFGMM_OK;
    std::string msg = std::string(what) + ": " + hipGetErrorString(e);
    if (e == hipErrorOutOfMemory) return fail(CFGMM_EOOM, msg);
    if (e == hipErrorNoDevice || e == hipErrorNoBinaryForGpu || e == hipErrorInvalidDevice)
        return fail(CFGMM_ENODEV, msg);
    return fail(CFGMM_EHIP, msg);
}

// ---------------------------------------------------------------------------
// tracing
// ---------------------------------------------------------------------------
struct TraceRec {
    std::string name;
    hipEvent_t start, end;
};
static std::atomic<bool> g_trace_on{false};
static std::mutex g_trace_mutex;
static std::vector<TraceRec> g_trace;
static thread_local int t_trace_off = 0;

TraceOff::TraceOff() { ++t_trace_off; }
TraceOff::~TraceOff() { --t_trace_off; }

hipEvent_t trace_start(hipStream_t stream) {
    if (t_trace_off || !g_trace_on.load(std::memory_order_relaxed)) return nullptr;
    hipEvent_t e = nullptr;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    if (hipEventRecord(e, stream) != hipSuccess) {
        (void)hipEventDestroy(e);
        return nullptr;
    }
    return e;
}

void trace_stop(hipStream_t stream, const char *name, hipEvent_t start) {
    if (!start) return;
    hipEvent_t e = nullptr;
    if (hipEventCreate(&e) != hipSuccess || hipEventRecord(e, stream) != hipSuccess) {
        (void)hipEventDestroy(start);
        if (e) (void)hipEventDestroy(e);
        return;
    }
    std::lock_guard<std::mutex> lock(g_trace_mutex);
    g_trace.push_back(TraceRec{name, start, e});
}

// ---------------------------------------------------------------------------
// tunables
// ---------------------------------------------------------------------------
static std::atomic<int> g_tune[CFGMM_TUNE_COUNT] = {{2}, {1}, {4}, {1}, {0}, {0}, {1}, {0}, {0}, {0}, {0}, {0}, {0}};
static const int g_tune_lo[CFGMM_TUNE_COUNT] = {1, 1, 1, 1, 0, 0, 1, 0, 0, 0, 0, 0, 0};
static const int g_tune_hi[CFGMM_TUNE_COUNT] = {8, 4, 64, 4, 128, 32, 2, 64, 1, 1, 2, 2, 1};

int tune_get(int param) { return g_tune[param].load(std::memory_order_relaxed); }

// ---------------------------------------------------------------------------
// device checks and properties
// ---------------------------------------------------------------------------
static std::mutex g_dev_mutex;
static std::map<int, std::pair<int, int>> g_dev_info;  // device -> (ok, cu count)

static int check_device(int *cus_out = nullptr) {
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return hip_status(e, "hipGetDevice");
    std::lock_guard<std::mutex> lock(g_dev_mutex);
    auto it = g_dev_info.find(dev);
    if (it == g_dev_info.end()) {
        hipDeviceProp_t prop;
        e = hipGetDeviceProperties(&prop, dev);
        if (e != hipSuccess) return hip_status(e, "hipGetDeviceProperties");
        const bool ok = std::strncmp(prop.gcnArchName, "gfx950", 6) == 0;
        it = g_dev_info.emplace(dev, std::make_pair(ok ? 1 : 0, prop.multiProcessorCount)).first;
        if (!ok) set_error(std::string("device is ") + prop.gcnArchName + ", libcfgmm is built for gfx950 only");
    }
    if (!it->second.first)
        return fail(CFGMM_ENODEV, "libcfgmm is built for gfx950 only; the current device is not gfx950");
    if (cus_out) *cus_out = it->second.second;
    return CFGMM_OK;
}

int device_cu_count() {
    int cus = 256;
    check_device(&cus);
    return cus;
}

// ---------------------------------------------------------------------------
// per-stream workspace
// ---------------------------------------------------------------------------
// g_ws_mutex guards the map only (its nodes never move); each Workspace's own mutex is held while
// that stream is synchronized or its buffers change, so a call waiting on one stream's work does
// not hold up calls on other streams.
static std::mutex g_ws_mutex;
static std::map<std::pair<int, hipStream_t>, Workspace> g_ws;

static Workspace &workspace_of(int dev, hipStream_t stream) {
    std::lock_guard<std::mutex> lock(g_ws_mutex);
    return g_ws[std::make_pair(dev, stream)];
}

// The allocation holds kFlagBytes of flag words, then the workspace proper.
constexpr size_t kFlagBytes = 256;

int workspace_get(hipStream_t stream, size_t bytes, void **out) {
    int dev = 0;
    CFGMM_HIP_TRY(hipGetDevice(&dev));
    Workspace &ws = workspace_of(dev, stream);
    std::lock_guard<std::mutex> lock(ws.mu);
    if (ws.bytes < bytes) {
        if (ws.ptr) {
            // kernels already enqueued on this stream may still read the old buffer
            CFGMM_HIP_TRY(hipStreamSynchronize(stream));
            CFGMM_HIP_TRY(hipFree(ws.ptr));
            ws.ptr = nullptr;
            ws.bytes = 0;
        }
        const size_t want = std::max(bytes + bytes / 4, (size_t)1 << 20);
        void *p = nullptr;
        CFGMM_HIP_TRY(hipMalloc(&p, want + kFlagBytes));
        // fresh memory may hold anything, an earlier call's epoch included: the flag words start at 0
        const hipError_t e = hipMemsetAsync(p, 0, kFlagBytes, stream);
        if (e != hipSuccess) {
            (void)hipFree(p);
            return hip_status(e, "workspace flags");
        }
        ws.ptr = p;
        ws.bytes = want;
    }
    *out = static_cast<char *>(ws.ptr) + kFlagBytes;
    return CFGMM_OK;
}

int *workspace_flag(hipStream_t stream) {
    int dev = 0;
    (void)hipGetDevice(&dev);
    Workspace &ws = workspace_of(dev, stream);
    std::lock_guard<std::mutex> lock(ws.mu);
    return static_cast<int *>(ws.ptr);
}

static std::atomic<unsigned> g_epoch{0};

unsigned workspace_epoch() {
    unsigned e;
    do e = ++g_epoch;
    while (e == 0);
    return e;
}

int host_staging_get(hipStream_t stream, size_t bytes, void **out) {
    int dev = 0;
    CFGMM_HIP_TRY(hipGetDevice(&dev));
    Workspace &ws = workspace_of(dev, stream);
    std::lock_guard<std::mutex> lock(ws.mu);
    if (ws.host_done) CFGMM_HIP_TRY(hipEventSynchronize(ws.host_done));  // earlier copies have read it
    if (ws.host_bytes < bytes) {
        if (ws.host) {
            CFGMM_HIP_TRY(hipStreamSynchronize(stream));
            CFGMM_HIP_TRY(hipHostFree(ws.host));
            ws.host = nullptr;
            ws.host_bytes = 0;
        }
        const size_t want = std::max(bytes + bytes / 4, (size_t)1 << 16);
        void *p = nullptr;
        CFGMM_HIP_TRY(hipHostMalloc(&p, want, hipHostMallocDefault));
        ws.host = p;
        ws.host_bytes = want;
    }
    *out = ws.host;
    return CFGMM_OK;
}

int host_staging_release(hipStream_t stream) {
    int dev = 0;
    CFGMM_HIP_TRY(hipGetDevice(&dev));
    Workspace &ws = workspace_of(dev, stream);
    std::lock_guard<std::mutex> lock(ws.mu);
    if (!ws.host_done) CFGMM_HIP_TRY(hipEventCreateWithFlags(&ws.host_done, hipEventDisableTiming));
    return hip_status(hipEventRecord(ws.host_done, stream), "host staging: event record");
}

// ---------------------------------------------------------------------------
// argument validation (extents the reference would index)
// ---------------------------------------------------------------------------
static bool mul_add_overflows(size_t a, size_t b, size_t c, size_t *out) {
    size_t t;
    if (__builtin_mul_overflow(a, b, &t)) return true;
    return __builtin_add_overflow(t, c, out);
}

// Elements spanned by `rows` rows of `cols` entries at leading dimension `ld`.
static int extent(size_t rows, size_t cols, size_t ld, size_t *out) {
    if (rows == 0 || cols == 0) {
        *out = 0;
        return CFGMM_OK;
    }
    if (mul_add_overflows(rows - 1, ld, cols, out)) return fail(CFGMM_EINVAL_DIMS, "extent overflows size_t");
    return CFGMM_OK;
}

static int check_len(const char *name, const void *ptr, size_t len, size_t need) {
    if (need == 0) return CFGMM_OK;
    if (!ptr) return fail(CFGMM_EINVAL, std::string(name) + " is null");
    if (len < need)
        return fail(CFGMM_EBOUNDS, std::string(name) + ": slice length " + std::to_string(len) + " < required " +
                                       std::to_string(need) + " (index out of bounds)");
    return CFGMM_OK;
}

static int check_ld(const char *name, size_t ld, size_t need, bool used) {
    if (used && ld < need)
        return fail(CFGMM_EINVAL_DIMS, std::string(name) + " = " + std::to_string(ld) + " < row length " +
                                           std::to_string(need));
    return CFGMM_OK;
}

#define CFGMM_TRY(expr)                 \
    do {                                \
        int _st = (expr);               \
        if (_st != CFGMM_OK) return _st; \
    } while (0)

struct GemmExtents {
    size_t a, b, c;
};

static int gemm_extents(size_t m, size_t n, size_t k, size_t lda, size_t ldb, size_t ldc, bool ta, bool tb,
                        GemmExtents *ext) {
    const bool any = m && n && k;
    CFGMM_TRY(check_ld("lda", lda, ta ? m : k, any));
    CFGMM_TRY(check_ld("ldb", ldb, tb ? k : n, any));
    CFGMM_TRY(check_ld("ldc", ldc, n, m && n));
    ext->a = ext->b = 0;
    if (any) {
        CFGMM_TRY(ta ? extent(k, m, lda, &ext->a) : extent(m, k, lda, &ext->a));
        CFGMM_TRY(tb ? extent(n, k, ldb, &ext->b) : extent(k, n, ldb, &ext->b));
    }
    CFGMM_TRY(extent(m, n, ldc, &ext->c));
    return CFGMM_OK;
}

static int check_blksize(size_t blksize) {
    if (blksize == 0 || blksize % 16 != 0)
        return fail(CFGMM_EINVAL, "blksize must be a positive multiple of 16 (BLKSIZE % (NR_LANE*LANE) == 0), got " +
                                      std::to_string(blksize));
    return CFGMM_OK;
}

struct RhoExtents {
    size_t rhos, dms, ao, tab;
};

static int rho_extents(size_t stride_ao, size_t stride_grid, size_t nao, size_t ngrid, size_t nset, size_t ldtab,
                       size_t blksize, RhoExtents *ext) {
    CFGMM_TRY(check_blksize(blksize));
    const bool any = nao && ngrid && nset;
    const size_t nblk = ngrid ? div_up(ngrid, blksize) : 0;
    CFGMM_TRY(check_ld("stride_ao", stride_ao, nao, any));
    CFGMM_TRY(check_ld("stride_grid", stride_grid, ngrid, any));
    CFGMM_TRY(check_ld("ldtab", ldtab, nblk, any));
    *ext = RhoExtents{0, 0, 0, 0};
    if (!any) return CFGMM_OK;
    size_t sa2;
    if (__builtin_mul_overflow(stride_ao, stride_ao, &sa2)) return fail(CFGMM_EINVAL_DIMS, "stride_ao^2 overflows");
    size_t dm_one;
    CFGMM_TRY(extent(nao, nao, stride_ao, &dm_one));
    if (mul_add_overflows(nset - 1, sa2, dm_one, &ext->dms)) return fail(CFGMM_EINVAL_DIMS, "dms extent overflows");
    CFGMM_TRY(extent(nset, ngrid, stride_grid, &ext->rhos));
    CFGMM_TRY(extent(nao, ngrid, stride_grid, &ext->ao));
    CFGMM_TRY(extent(nao, nblk, ldtab, &ext->tab));
    return CFGMM_OK;
}

// ---------------------------------------------------------------------------
// host-pointer staging: one stream and a set of grow-only device buffers per
// calling thread (kept until the process ends).
// ---------------------------------------------------------------------------
// Transfers of kPinMin bytes and up go through two pinned chunks of kPinChunk bytes: host threads
// copy a chunk between the caller's (pageable) memory and one pinned chunk while the DMA engine
// moves the other (about 52 GB/s on the host side measured, against 30 GB/s for pageable DMA).
constexpr size_t kPinChunk = (size_t)32 << 20, kPinMin = (size_t)8 << 20;
struct HostStage {
    int device = -1;
    hipStream_t stream = nullptr;
    void *buf[5] = {nullptr, nullptr, nullptr, nullptr, nullptr};
    size_t cap[5] = {0, 0, 0, 0, 0};
    void *pin[2] = {nullptr, nullptr};          // pinned chunks (hipHostMalloc)
    hipEvent_t pin_done[2] = {nullptr, nullptr};  // recorded after the last DMA touching each chunk
    bool pin_pending[2] = {false, false};
    bool pin_failed = false;                    // no pinned memory: pageable transfers only
};
static thread_local HostStage t_stage;

// memcpy split over up to 8 threads (the caller is one of them)
static void parallel_memcpy(void *dst, const void *src, size_t n) {
    const unsigned hw = std::max(1u, std::thread::hardware_concurrency());
    const unsigned nt = n < ((size_t)2 << 20) ? 1u : std::min(8u, hw);
    std::vector<std::thread> th;
    th.reserve(nt - 1);
    for (unsigned t = 1; t < nt; ++t)
        th.emplace_back([=] {
            const size_t a = n * t / nt, b = n * (t + 1) / nt;
            std::memcpy(static_cast<char *>(dst) + a, static_cast<const char *>(src) + a, b - a);
        });
    std::memcpy(dst, src, n / nt);
    for (auto &x : th) x.join();
}

// true when the pinned chunks exist (allocated on first use)
static bool pin_ready() {
    if (t_stage.pin_failed) return false;
    if (t_stage.pin[0]) return true;
    for (int i = 0; i < 2; ++i) {
        if (hipHostMalloc(&t_stage.pin[i], kPinChunk, hipHostMallocDefault) != hipSuccess ||
            hipEventCreateWithFlags(&t_stage.pin_done[i], hipEventDisableTiming) != hipSuccess) {
            t_stage.pin[i] = nullptr;
            t_stage.pin_failed = true;
            (void)hipGetLastError();
            return false;
        }
    }
    return true;
}

// waits until the DMA that last used pinned chunk i is done
static int pin_wait(int i) {
    if (t_stage.pin_pending[i]) CFGMM_HIP_TRY(hipEventSynchronize(t_stage.pin_done[i]));
    t_stage.pin_pending[i] = false;
    return CFGMM_OK;
}

static int copy_h2d(void *dev, const void *host, size_t bytes) {
    if (bytes < kPinMin || !pin_ready())
        return hip_status(hipMemcpyAsync(dev, host, bytes, hipMemcpyHostToDevice, t_stage.stream), "hipMemcpyAsync H2D");
    for (size_t off = 0, i = 0; off < bytes; off += kPinChunk, ++i) {
        const int c = (int)(i & 1);
        const size_t n = std::min(kPinChunk, bytes - off);
        CFGMM_TRY(pin_wait(c));
        parallel_memcpy(t_stage.pin[c], static_cast<const char *>(host) + off, n);
        CFGMM_HIP_TRY(hipMemcpyAsync(static_cast<char *>(dev) + off, t_stage.pin[c], n, hipMemcpyHostToDevice,
                                     t_stage.stream));
        CFGMM_HIP_TRY(hipEventRecord(t_stage.pin_done[c], t_stage.stream));
        t_stage.pin_pending[c] = true;
    }
    return CFGMM_OK;
}

// device-to-host, synchronous: every byte is in `host` on return
static int copy_d2h(void *host, const void *dev, size_t bytes) {
    if (bytes < kPinMin || !pin_ready()) {
        CFGMM_HIP_TRY(hipMemcpyAsync(host, dev, bytes, hipMemcpyDeviceToHost, t_stage.stream));
        return hip_status(hipStreamSynchronize(t_stage.stream), "hipStreamSynchronize");
    }
    // chunk i lands in pinned chunk i % 2 while the host copies chunk i - 1 out of the other
    const size_t nchunks = div_up(bytes, kPinChunk);
    for (size_t i = 0; i <= nchunks; ++i) {
        if (i < nchunks) {
            const int c = (int)(i & 1);
            const size_t off = i * kPinChunk, n = std::min(kPinChunk, bytes - off);
            CFGMM_TRY(pin_wait(c));  // an earlier DMA into or out of this chunk
            CFGMM_HIP_TRY(hipMemcpyAsync(t_stage.pin[c], static_cast<const char *>(dev) + off, n,
                                         hipMemcpyDeviceToHost, t_stage.stream));
            CFGMM_HIP_TRY(hipEventRecord(t_stage.pin_done[c], t_stage.stream));
            t_stage.pin_pending[c] = true;
        }
        if (i > 0) {
            const int c = (int)((i - 1) & 1);
            const size_t off = (i - 1) * kPinChunk, n = std::min(kPinChunk, bytes - off);
            CFGMM_TRY(pin_wait(c));
            parallel_memcpy(static_cast<char *>(host) + off, t_stage.pin[c], n);
        }
    }
    return hip_status(hipStreamSynchronize(t_stage.stream), "hipStreamSynchronize");
}

static int stage_begin() {
    CFGMM_TRY(check_device());
    int dev = 0;
    CFGMM_HIP_TRY(hipGetDevice(&dev));
    if (t_stage.device != dev) {
        // buffers of another device are abandoned, not freed (their device is not current)
        t_stage = HostStage();
        t_stage.device = dev;
        CFGMM_HIP_TRY(hipStreamCreateWithFlags(&t_stage.stream, hipStreamNonBlocking));
    }
    return CFGMM_OK;
}

static int stage_buf(int slot, size_t bytes, void **out) {
    if (t_stage.cap[slot] < bytes) {
        if (t_stage.buf[slot]) {
            CFGMM_HIP_TRY(hipStreamSynchronize(t_stage.stream));
            CFGMM_HIP_TRY(hipFree(t_stage.buf[slot]));
            t_stage.buf[slot] = nullptr;
            t_stage.cap[slot] = 0;
        }
        CFGMM_HIP_TRY(hipMalloc(&t_stage.buf[slot], bytes));
        t_stage.cap[slot] = bytes;
    }
    *out = t_stage.buf[slot];
    return CFGMM_OK;
}

template <typename T>
static int stage_in(int slot, const T *host, size_t count, T **dev) {
    void *p = nullptr;
    CFGMM_TRY(stage_buf(slot, std::max<size_t>(count * sizeof(T), 256), &p));
    if (count) CFGMM_TRY(copy_h2d(p, host, count * sizeof(T)));
    *dev = static_cast<T *>(p);
    return CFGMM_OK;
}

static int stage_out(double *host, const double *dev, size_t count) {
    if (count) return copy_d2h(host, dev, count * sizeof(double));
    return hip_status(hipStreamSynchronize(t_stage.stream), "hipStreamSynchronize");
}

// ---------------------------------------------------------------------------
// shared bodies
// ---------------------------------------------------------------------------
static int gemm_host(double *c, size_t c_len, const double *a, size_t a_len, const double *b, size_t b_len, size_t m,
                     size_t n, size_t k, size_t lda, size_t ldb, size_t ldc, int transa, int transb,
                     const uint8_t *tab, size_t tab_len, size_t ldtab, size_t blksize) {
    clear_error();
    const bool masked = tab != nullptr || blksize != 0;
    if (masked && transb)
        return fail(CFGMM_EINVAL_TRANSB_MASK,
                    "transb=true is not implemented in matmul_loop_parallel_mnk_pack_a_non0tab");
    if (masked) CFGMM_TRY(check_blksize(blksize));
    GemmExtents ext;
    CFGMM_TRY(gemm_extents(m, n, k, lda, ldb, ldc, transa, transb, &ext));
    size_t tab_ext = 0;
    if (masked && m && n && k) {
        const size_t nblk = div_up(n, blksize);
        CFGMM_TRY(check_ld("ldtab", ldtab, nblk, true));
        CFGMM_TRY(extent(k, nblk, ldtab, &tab_ext));
        CFGMM_TRY(check_len("tab", tab, tab_len, tab_ext));
    }
    CFGMM_TRY(check_len("c", c, c_len, ext.c));
    CFGMM_TRY(check_len("a", a, a_len, ext.a));
    CFGMM_TRY(check_len("b", b, b_len, ext.b));
    if (!m || !n || !k) return CFGMM_OK;
    CFGMM_TRY(stage_begin());
    double *dc, *da, *db;
    CFGMM_TRY(stage_in(0, c, ext.c, &dc));
    CFGMM_TRY(stage_in(1, a, ext.a, &da));
    CFGMM_TRY(stage_in(2, b, ext.b, &db));
    if (masked) {
        uint8_t *dt;
        CFGMM_TRY(stage_in(3, tab, tab_ext, &dt));
        CFGMM_TRY(launch_dgemm_non0tab(t_stage.stream, dc, da, db, m, n, k, lda, ldb, ldc, transa != 0, false, dt,
                                       ldtab, blksize));
    } else {
        CFGMM_TRY(launch_dgemm(t_stage.stream, dc, da, db, m, n, k, lda, ldb, ldc, transa != 0, transb != 0));
    }
    return stage_out(c, dc, ext.c);
}

static int rho_host(double *rhos, size_t rhos_len, const double *dms, size_t dms_len, const double *ao_lhs,
                    size_t ao_lhs_len, const double *ao_rhs, size_t ao_rhs_len, size_t stride_ao, size_t stride_grid,
                    size_t nao, size_t ngrid, size_t nset, const uint8_t *tab, size_t tab_len, size_t ldtab,
                    size_t blksize) {
    clear_error();
    RhoExtents ext;
    CFGMM_TRY(rho_extents(stride_ao, stride_grid, nao, ngrid, nset, ldtab, blksize, &ext));
    CFGMM_TRY(check_len("rhos", rhos, rhos_len, ext.rhos));
    CFGMM_TRY(check_len("dms", dms, dms_len, ext.dms));
    CFGMM_TRY(check_len("ao_lhs", ao_lhs, ao_lhs_len, ext.ao));
    CFGMM_TRY(check_len("ao_rhs", ao_rhs, ao_rhs_len, ext.ao));
    CFGMM_TRY(check_len("tab", tab, tab_len, ext.tab));
    if (!nao || !ngrid || !nset) return CFGMM_OK;
    CFGMM_TRY(stage_begin());
    double *dr, *dd, *dl, *dright;
    uint8_t *dt;
    CFGMM_TRY(stage_in(0, rhos, ext.rhos, &dr));
    CFGMM_TRY(stage_in(1, dms, ext.dms, &dd));
    CFGMM_TRY(stage_in(2, ao_lhs, ext.ao, &dl));
    if (ao_rhs == ao_lhs) {
        dright = dl;
    } else {
        CFGMM_TRY(stage_in(4, ao_rhs, ext.ao, &dright));
    }
    CFGMM_TRY(stage_in(3, tab, ext.tab, &dt));
    CFGMM_TRY(launch_rho(t_stage.stream, dr, dd, dl, dright, stride_ao, stride_grid, nao, ngrid, nset, dt, ldtab,
                         blksize));
    return stage_out(rhos, dr, ext.rhos);
}

// Arguments of a device-pointer GEMM; the device is checked by the caller after every argument.
static int check_dev_gemm(const double *c, const double *a, const double *b, size_t m, size_t n, size_t k,
                          size_t lda, size_t ldb, size_t ldc, int transa, int transb) {
    GemmExtents ext;
    CFGMM_TRY(gemm_extents(m, n, k, lda, ldb, ldc, transa, transb, &ext));
    if (m && n && k && (!c || !a || !b)) return fail(CFGMM_EINVAL, "null device pointer");
    return CFGMM_OK;
}

static int check_precision(int precision, int nslices) {
    if (precision == CFGMM_PREC_FP64) return CFGMM_OK;
    if (precision != CFGMM_PREC_OZAKI_INT8) return fail(CFGMM_EINVAL, "unknown precision mode " + std::to_string(precision));
    if (nslices < 2 || nslices > 6) return fail(CFGMM_EINVAL, "nslices must be in [2, 6]");
    return CFGMM_OK;
}

}  // namespace cfgmm

using namespace cfgmm;

extern "C" {

const char *cfgmm_last_error(void) { return g_last_error.c_str(); }
int cfgmm_version(void) { return 10000; }

// ---- GEMM ------------------------------------------------------------------
int cfgmm_matmul_loop_parallel_mnk_pack_a(double *c, size_t c_len, const double *a, size_t a_len, const double *b,
                                          size_t b_len, size_t m, size_t n, size_t k, size_t lda, size_t ldb,
                                          size_t ldc, int transa, int transb) {
    return gemm_host(c, c_len, a, a_len, b, b_len, m, n, k, lda, ldb, ldc, transa, transb, nullptr, 0, 0, 0);
}

int cfgmm_matmul_loop_macro_mb(double *c, size_t c_len, const double *a, size_t a_len, const double *b, size_t b_len,
                               size_t m, size_t n, size_t k, size_t lda, size_t ldb, size_t ldc, int transa,
                               int transb) {
    return gemm_host(c, c_len, a, a_len, b, b_len, m, n, k, lda, ldb, ldc, transa, transb, nullptr, 0, 0, 0);
}

int cfgmm_matmul_anyway_full(double *c, size_t c_len, const double *a, size_t a_len, const double *b, size_t b_len,
                             size_t m, size_t n, size_t k, size_t lda, size_t ldb, size_t ldc, int transa,
                             int transb) {
    return gemm_host(c, c_len, a, a_len, b, b_len, m, n, k, lda, ldb, ldc, transa, transb, nullptr, 0, 0, 0);
}

int cfgmm_dgemm_dev(void *stream, double *c, const double *a, const double *b, size_t m, size_t n, size_t k,
                    size_t lda, size_t ldb, size_t ldc, int transa, int transb) {
    clear_error();
    CFGMM_TRY(check_dev_gemm(c, a, b, m, n, k, lda, ldb, ldc, transa, transb));
    CFGMM_TRY(check_device());
    return launch_dgemm(static_cast<hipStream_t>(stream), c, a, b, m, n, k, lda, ldb, ldc, transa != 0, transb != 0);
}

int cfgmm_dgemm_ex_dev(void *stream, double *c, const double *a, const double *b, size_t m, size_t n, size_t k,
                       size_t lda, size_t ldb, size_t ldc, int transa, int transb, int precision, int nslices) {
    clear_error();
    CFGMM_TRY(check_dev_gemm(c, a, b, m, n, k, lda, ldb, ldc, transa, transb));
    CFGMM_TRY(check_precision(precision, nslices));
    CFGMM_TRY(check_device());
    hipStream_t s = static_cast<hipStream_t>(stream);
    if (precision == CFGMM_PREC_FP64)
        return launch_dgemm(s, c, a, b, m, n, k, lda, ldb, ldc, transa != 0, transb != 0);
    return launch_dgemm_ozaki(s, c, a, b, m, n, k, lda, ldb, ldc, transa != 0, transb != 0, nslices);
}

// ---- masked GEMM -------------------------------------------------------------
int cfgmm_matmul_loop_parallel_mnk_pack_a_non0tab(double *c, size_t c_len, const double *a, size_t a_len,
                                                  const double *b, size_t b_len, size_t m, size_t n, size_t k,
                                                  size_t lda, size_t ldb, size_t ldc, int transa, int transb,
                                                  const uint8_t *tab, size_t tab_len, size_t ldtab, size_t blksize) {
    if (blksize == 0) {
        clear_error();
        return fail(CFGMM_EINVAL, "blksize must be a positive multiple of 16, got 0");
    }
    return gemm_host(c, c_len, a, a_len, b, b_len, m, n, k, lda, ldb, ldc, transa, transb, tab, tab_len, ldtab,
                     blksize);
}

int cfgmm_matmul_anyway_full_non0tab(double *c, size_t c_len, const double *a, size_t a_len, const double *b,
                                     size_t b_len, size_t m, size_t n, size_t k, size_t lda, size_t ldb, size_t ldc,
                                     int transa, int transb, const uint8_t *tab, size_t tab_len, size_t ldtab) {
    return gemm_host(c, c_len, a, a_len, b, b_len, m, n, k, lda, ldb, ldc, transa, transb, tab, tab_len, ldtab, 48);
}

static int masked_dev_common(void *stream, double *c, const double *a, const double *b, size_t m, size_t n, size_t k,
                             size_t lda, size_t ldb, size_t ldc, int transa, int transb, const uint8_t *tab,
                             size_t ldtab, size_t blksize) {
    CFGMM_TRY(check_blksize(blksize));
    CFGMM_TRY(check_dev_gemm(c, a, b, m, n, k, lda, ldb, ldc, transa, transb));
    if (m && n && k) {
        CFGMM_TRY(check_ld("ldtab", ldtab, div_up(n, blksize), true));
        if (!tab) return fail(CFGMM_EINVAL, "tab is null");
    }
    CFGMM_TRY(check_device());
    return launch_dgemm_non0tab(static_cast<hipStream_t>(stream), c, a, b, m, n, k, lda, ldb, ldc, transa != 0,
                                transb != 0, tab, ldtab, blksize);
}

int cfgmm_dgemm_non0tab_dev(void *stream, double *c, const double *a, const double *b, size_t m, size_t n, size_t k,
                            size_t lda, size_t ldb, size_t ldc, int transa, int transb, const uint8_t *tab,
                            size_t ldtab, size_t blksize) {
    clear_error();
    if (transb)
        return fail(CFGMM_EINVAL_TRANSB_MASK,
                    "transb=true is not implemented in matmul_loop_parallel_mnk_pack_a_non0tab");
    return masked_dev_common(stream, c, a, b, m, n, k, lda, ldb, ldc, transa, 0, tab, ldtab, blksize);
}

int cfgmm_dgemm_non0tab_ex_dev(void *stream, double *c, const double *a, const double *b, size_t m, size_t n,
                               size_t k, size_t lda, size_t ldb, size_t ldc, int transa, int transb,
                               const uint8_t *tab, size_t ldtab, size_t blksize) {
    clear_error();
    return masked_dev_common(stream, c, a, b, m, n, k, lda, ldb, ldc, transa, transb, tab, ldtab, blksize);
}

int cfgmm_sgemm_non0tab_dev(void *stream, float *c, const float *a, const float *b, size_t m, size_t n, size_t k,
                            size_t lda, size_t ldb, size_t ldc, int transa, int transb, const uint8_t *tab,
                            size_t ldtab, size_t blksize) {
    clear_error();
    CFGMM_TRY(check_blksize(blksize));
    // extents and null checks do not depend on the element type
    CFGMM_TRY(check_dev_gemm(reinterpret_cast<const double *>(c), reinterpret_cast<const double *>(a),
                             reinterpret_cast<const double *>(b), m, n, k, lda, ldb, ldc, transa, transb));
    if (m && n && k) {
        CFGMM_TRY(check_ld("ldtab", ldtab, div_up(n, blksize), true));
        if (!tab) return fail(CFGMM_EINVAL, "tab is null");
    }
    CFGMM_TRY(check_device());
    return launch_sgemm_non0tab(static_cast<hipStream_t>(stream), c, a, b, m, n, k, lda, ldb, ldc, transa != 0,
                                transb != 0, tab, ldtab, blksize);
}

// ---- rho ------------------------------------------------------------------------
int cfgmm_aodmao2rho_loop_parallel(double *rhos, size_t rhos_len, const double *dms, size_t dms_len,
                                   const double *ao_lhs, size_t ao_lhs_len, const double *ao_rhs, size_t ao_rhs_len,
                                   size_t stride_ao, size_t stride_grid, size_t nao, size_t ngrid, size_t nset,
                                   const uint8_t *tab, size_t tab_len, size_t ldtab, size_t blksize) {
    return rho_host(rhos, rhos_len, dms, dms_len, ao_lhs, ao_lhs_len, ao_rhs, ao_rhs_len, stride_ao, stride_grid, nao,
                    ngrid, nset, tab, tab_len, ldtab, blksize);
}

int cfgmm_aodmao2rho_anyway(double *rhos, size_t rhos_len, const double *dms, size_t dms_len, const double *ao_lhs,
                            size_t ao_lhs_len, const double *ao_rhs, size_t ao_rhs_len, size_t stride_ao,
                            size_t stride_grid, size_t nao, size_t ngrid, size_t nset, const uint8_t *tab,
                            size_t tab_len, size_t ldtab) {
    return rho_host(rhos, rhos_len, dms, dms_len, ao_lhs, ao_lhs_len, ao_rhs, ao_rhs_len, stride_ao, stride_grid, nao,
                    ngrid, nset, tab, tab_len, ldtab, 48);
}

static int rho_dev_common(void *stream, double *rhos, const double *dms, const double *ao_lhs, const double *ao_rhs,
                          size_t stride_ao, size_t stride_grid, size_t nao, size_t ngrid, size_t nset,
                          const uint8_t *tab, size_t ldtab, size_t blksize, int precision, int nslices) {
    clear_error();
    RhoExtents ext;
    CFGMM_TRY(rho_extents(stride_ao, stride_grid, nao, ngrid, nset, ldtab, blksize, &ext));
    if (nao && ngrid && nset && (!rhos || !dms || !ao_lhs || !ao_rhs || !tab))
        return fail(CFGMM_EINVAL, "null device pointer");
    CFGMM_TRY(check_precision(precision, nslices));
    CFGMM_TRY(check_device());
    hipStream_t s = static_cast<hipStream_t>(stream);
    if (precision == CFGMM_PREC_FP64)
        return launch_rho(s, rhos, dms, ao_lhs, ao_rhs, stride_ao, stride_grid, nao, ngrid, nset, tab, ldtab, blksize);
    return launch_rho_ozaki(s, rhos, dms, ao_lhs, ao_rhs, stride_ao, stride_grid, nao, ngrid, nset, tab, ldtab,
                            blksize, nslices);
}

int cfgmm_aodmao2rho_dev(void *stream, double *rhos, const double *dms, const double *ao_lhs, const double *ao_rhs,
                         size_t stride_ao, size_t stride_grid, size_t nao, size_t ngrid, size_t nset,
                         const uint8_t *tab, size_t ldtab, size_t blksize) {
    return rho_dev_common(stream, rhos, dms, ao_lhs, ao_rhs, stride_ao, stride_grid, nao, ngrid, nset, tab, ldtab,
                          blksize, CFGMM_PREC_FP64, 0);
}

int cfgmm_aodmao2rho_dev_noalloc(void *stream, double *rhos, const double *dms, const double *ao_lhs,
                                 const double *ao_rhs, size_t stride_ao, size_t stride_grid, size_t nao, size_t ngrid,
                                 size_t nset, const uint8_t *tab, size_t ldtab, size_t blksize) {
    clear_error();
    RhoExtents ext;
    CFGMM_TRY(rho_extents(stride_ao, stride_grid, nao, ngrid, nset, ldtab, blksize, &ext));
    if (nao && ngrid && nset && (!rhos || !dms || !ao_lhs || !ao_rhs || !tab))
        return fail(CFGMM_EINVAL, "null device pointer");
    CFGMM_TRY(check_device());
    return launch_rho_noalloc(static_cast<hipStream_t>(stream), rhos, dms, ao_lhs, ao_rhs, stride_ao, stride_grid,
                              nao, ngrid, nset, tab, ldtab, blksize);
}

int cfgmm_aodmao2rho_multi_dev(void *stream, double *rhos, size_t rho_comp_stride, const double *dms,
                               const double *ao_lhs, size_t lhs_comp_stride, size_t ncomp, const double *ao_rhs,
                               size_t stride_ao, size_t stride_grid, size_t nao, size_t ngrid, size_t nset,
                               const uint8_t *tab, size_t ldtab, size_t blksize) {
    clear_error();
    RhoExtents ext;
    CFGMM_TRY(rho_extents(stride_ao, stride_grid, nao, ngrid, nset, ldtab, blksize, &ext));
    if (ncomp > 16) return fail(CFGMM_EINVAL, "at most 16 left-hand components");
    const bool any = nao && ngrid && nset && ncomp;
    if (any && (!rhos || !dms || !ao_lhs || !ao_rhs || !tab)) return fail(CFGMM_EINVAL, "null device pointer");
    if (any && ncomp > 1 && rho_comp_stride < (nset - 1) * stride_grid + ngrid)
        return fail(CFGMM_EINVAL_DIMS, "rho_comp_stride: the components' rho ranges overlap");
    CFGMM_TRY(check_device());
    return launch_rho_multi(static_cast<hipStream_t>(stream), rhos, rho_comp_stride, dms, ao_lhs, lhs_comp_stride,
                            ncomp, ao_rhs, stride_ao, stride_grid, nao, ngrid, nset, tab, ldtab, blksize);
}

int cfgmm_aogridao2mat_dev(void *stream, double *V, size_t ldv, const double *ao_lhs, const double *ao_rhs,
                           const double *w, size_t stride_grid, size_t nao, size_t ngrid, const uint8_t *tab,
                           size_t ldtab, size_t blksize) {
    clear_error();
    CFGMM_TRY(check_blksize(blksize));
    const bool any = nao && ngrid;
    CFGMM_TRY(check_ld("ldv", ldv, nao, any));
    CFGMM_TRY(check_ld("stride_grid", stride_grid, ngrid, any));
    CFGMM_TRY(check_ld("ldtab", ldtab, ngrid ? div_up(ngrid, blksize) : 0, any));
    if (any && (!V || !ao_lhs || !ao_rhs || !tab)) return fail(CFGMM_EINVAL, "null device pointer");
    CFGMM_TRY(check_device());
    return launch_vxc(static_cast<hipStream_t>(stream), V, ldv, ao_lhs, ao_rhs, w, stride_grid, nao, ngrid, tab,
                      ldtab, blksize);
}

int cfgmm_eval_gto_dev(void *stream, double *phi, size_t stride_grid, uint8_t *tab, size_t ldtab,
                       const cfgmm_gto_shell *shells, size_t nshell, const double *exps, const double *coefs,
                       size_t nprim_total, const double *coords, size_t ngrid, size_t nao, size_t blksize,
                       double cutoff) {
    clear_error();
    CFGMM_TRY(check_blksize(blksize));
    const bool any = nshell && ngrid;
    CFGMM_TRY(check_ld("stride_grid", stride_grid, ngrid, any));
    CFGMM_TRY(check_ld("ldtab", ldtab, ngrid ? div_up(ngrid, blksize) : 0, any));
    if (!(cutoff >= 0.0)) return fail(CFGMM_EINVAL, "cutoff must be >= 0");
    if (nshell && !shells) return fail(CFGMM_EINVAL, "shells is null");
    if (nprim_total && (!exps || !coefs)) return fail(CFGMM_EINVAL, "exps / coefs are null");
    if (nao > INT32_MAX || ngrid > INT32_MAX || nshell > INT32_MAX || nprim_total > INT32_MAX)
        return fail(CFGMM_EINVAL, "gto: size out of range");
    for (size_t i = 0; i < nshell; ++i) {
        const cfgmm_gto_shell &sh = shells[i];
        const std::string who = "shell " + std::to_string(i) + ": ";
        if (sh.l < 0 || sh.l > 4) return fail(CFGMM_EINVAL, who + "l must be 0..4");
        if (sh.nprim < 1 || sh.prim_off < 0 || (size_t)sh.prim_off + sh.nprim > nprim_total)
            return fail(CFGMM_EINVAL, who + "primitives outside exps / coefs");
        const int nf = (sh.l + 1) * (sh.l + 2) / 2;
        if (sh.ao_off < 0 || (size_t)sh.ao_off + nf > nao) return fail(CFGMM_EINVAL, who + "AOs outside nao");
    }
    if (any && (!phi || !tab || !coords)) return fail(CFGMM_EINVAL, "null device pointer");
    CFGMM_TRY(check_device());
    return launch_gto(static_cast<hipStream_t>(stream), phi, stride_grid, tab, ldtab, shells, nshell, exps, coefs,
                      nprim_total, coords, ngrid, blksize, cutoff);
}

int cfgmm_aodmao2rho_batched_dev(void *stream, const cfgmm_rho_problem *problems, size_t nproblems, int precision,
                                 int nslices) {
    clear_error();
    if (nproblems && !problems) return fail(CFGMM_EINVAL, "problems is null");
    for (size_t i = 0; i < nproblems; ++i) {
        const cfgmm_rho_problem &q = problems[i];
        RhoExtents ext;
        int st = rho_extents(q.stride_ao, q.stride_grid, q.nao, q.ngrid, q.nset, q.ldtab, q.blksize, &ext);
        if (st != CFGMM_OK) return fail(st, "problem " + std::to_string(i) + ": " + g_last_error);
        if (q.nao && q.ngrid && q.nset && (!q.rhos || !q.dms || !q.ao_lhs || !q.ao_rhs || !q.tab))
            return fail(CFGMM_EINVAL, "problem " + std::to_string(i) + ": null device pointer");
    }
    CFGMM_TRY(check_precision(precision, nslices));
    CFGMM_TRY(check_device());
    hipStream_t s = static_cast<hipStream_t>(stream);
    if (precision == CFGMM_PREC_FP64) return launch_rho_batched(s, problems, nproblems);
    {
        for (size_t i = 0; i < nproblems; ++i) {
            const cfgmm_rho_problem &q = problems[i];
            CFGMM_TRY(launch_rho_ozaki(s, q.rhos, q.dms, q.ao_lhs, q.ao_rhs, q.stride_ao, q.stride_grid, q.nao,
                                       q.ngrid, q.nset, q.tab, q.ldtab, q.blksize, nslices));
        }
        return CFGMM_OK;
    }
}

// ---- resident rho operands -------------------------------------------------------
static ResidentRho *impl(const cfgmm_rho_resident *h) {
    return reinterpret_cast<ResidentRho *>(const_cast<cfgmm_rho_resident *>(h));
}

int cfgmm_rho_resident_create(void *stream, const double *dms, const double *ao_lhs, const double *ao_rhs,
                              size_t stride_ao, size_t stride_grid, size_t nao, size_t ngrid, size_t nset,
                              const uint8_t *tab, size_t ldtab, size_t blksize, int precision, int nslices,
                              cfgmm_rho_resident **out) {
    clear_error();
    if (!out) return fail(CFGMM_EINVAL, "out is null");
    *out = nullptr;
    RhoExtents ext;
    CFGMM_TRY(rho_extents(stride_ao, stride_grid, nao, ngrid, nset, ldtab, blksize, &ext));
    if (nao && ngrid && nset && (!dms || !ao_lhs || !ao_rhs || !tab))
        return fail(CFGMM_EINVAL, "null device pointer");
    if (precision != CFGMM_PREC_OZAKI_INT8)
        return fail(CFGMM_EINVAL, "resident rho operands exist in split precision (CFGMM_PREC_OZAKI_INT8) only");
    CFGMM_TRY(check_device());
    ResidentRho *h = nullptr;
    CFGMM_TRY(resident_create(static_cast<hipStream_t>(stream), dms, ao_lhs, ao_rhs, stride_ao, stride_grid, nao,
                              ngrid, nset, tab, ldtab, blksize, nslices, &h));
    *out = reinterpret_cast<cfgmm_rho_resident *>(h);
    return CFGMM_OK;
}

static int check_resident_eval(const cfgmm_rho_resident *h, const double *rhos, size_t stride_rho, size_t ngrid_used,
                               size_t nset_used) {
    if (!h) return fail(CFGMM_EINVAL, "resident handle is null");
    const ResidentRho *r = impl(h);
    if (ngrid_used > resident_ngrid(r) || nset_used > resident_nset(r))
        return fail(CFGMM_EINVAL_DIMS, "resident rho: ngrid_used / nset_used exceed the resident problem");
    if (ngrid_used != resident_ngrid(r) && ngrid_used % resident_blksize(r) != 0)
        return fail(CFGMM_EINVAL_DIMS, "resident rho: ngrid_used must be the full grid or a multiple of blksize");
    if (ngrid_used && nset_used > 1 && stride_rho < ngrid_used)
        return fail(CFGMM_EINVAL_DIMS, "resident rho: stride_rho < ngrid_used");
    if (ngrid_used && nset_used && !rhos) return fail(CFGMM_EINVAL, "rhos is null");
    int dev = 0;
    CFGMM_HIP_TRY(hipGetDevice(&dev));
    if (dev != resident_device(r)) return fail(CFGMM_EINVAL, "resident rho: handle belongs to another device");
    return CFGMM_OK;
}

int cfgmm_rho_resident_eval(void *stream, const cfgmm_rho_resident *h, double *rhos, size_t stride_rho,
                            size_t ngrid_used, size_t nset_used) {
    clear_error();
    CFGMM_TRY(check_resident_eval(h, rhos, stride_rho, ngrid_used, nset_used));
    return resident_eval(static_cast<hipStream_t>(stream), impl(h), rhos, stride_rho, ngrid_used, nset_used);
}

int cfgmm_rho_resident_eval_batched(void *stream, const cfgmm_rho_resident_job *evals, size_t n) {
    clear_error();
    if (n && !evals) return fail(CFGMM_EINVAL, "evals is null");
    std::vector<ResidentEvalReq> reqs(n);
    for (size_t i = 0; i < n; ++i) {
        const cfgmm_rho_resident_job &e = evals[i];
        const int st = check_resident_eval(e.h, e.rhos, e.stride_rho, e.ngrid_used, e.nset_used);
        if (st != CFGMM_OK) return fail(st, "evaluation " + std::to_string(i) + ": " + g_last_error);
        reqs[i] = ResidentEvalReq{impl(e.h), e.rhos, e.stride_rho, e.ngrid_used, e.nset_used};
    }
    return resident_eval_batched(static_cast<hipStream_t>(stream), reqs.data(), n);
}

int cfgmm_rho_resident_bytes(const cfgmm_rho_resident *h, size_t *bytes) {
    clear_error();
    if (!h || !bytes) return fail(CFGMM_EINVAL, "null argument");
    *bytes = resident_bytes(impl(h));
    return CFGMM_OK;
}

int cfgmm_rho_resident_destroy(cfgmm_rho_resident *h) {
    clear_error();
    return resident_destroy(impl(h));
}

// ---- serving loop -------------------------------------------------------------------
static Server *srv(const cfgmm_server *s) { return reinterpret_cast<Server *>(const_cast<cfgmm_server *>(s)); }

int cfgmm_server_create(const cfgmm_server_config *cfg, cfgmm_server **out) {
    clear_error();
    if (!cfg || !out) return fail(CFGMM_EINVAL, "null argument");
    *out = nullptr;
    CFGMM_TRY(check_device());
    Server *s = nullptr;
    CFGMM_TRY(server_create(cfg, &s));
    *out = reinterpret_cast<cfgmm_server *>(s);
    return CFGMM_OK;
}

int cfgmm_server_add_molecule(cfgmm_server *s, int id, const double *dms, const double *phi, const uint8_t *tab,
                              size_t nao, size_t ngrid, size_t nset, size_t blksize) {
    clear_error();
    if (!s) return fail(CFGMM_EINVAL, "server is null");
    return server_add_molecule(srv(s), id, dms, phi, tab, nao, ngrid, nset, blksize);
}

int cfgmm_server_warm(cfgmm_server *s, const int *ids, size_t n) {
    clear_error();
    if (!s || (n && !ids)) return fail(CFGMM_EINVAL, "null argument");
    return server_warm(srv(s), ids, n);
}

int cfgmm_server_run(cfgmm_server *s, const cfgmm_request *reqs, size_t n, cfgmm_request_result *results) {
    clear_error();
    if (!s || (n && (!reqs || !results))) return fail(CFGMM_EINVAL, "null argument");
    return server_run(srv(s), reqs, n, results);
}

int cfgmm_server_get_stats(const cfgmm_server *s, cfgmm_server_stats *out) {
    clear_error();
    if (!s || !out) return fail(CFGMM_EINVAL, "null argument");
    return server_stats(srv(s), out);
}

int cfgmm_server_destroy(cfgmm_server *s) {
    clear_error();
    if (!s) return CFGMM_OK;
    return server_destroy(srv(s));
}

int cfgmm_tune_set(int param, int value) {
    clear_error();
    if (param < 0 || param >= CFGMM_TUNE_COUNT) return fail(CFGMM_EINVAL, "unknown tunable " + std::to_string(param));
    if (param == CFGMM_TUNE_RESIDENT_COLS && value % 16 != 0)
        return fail(CFGMM_EINVAL, "resident columns must be 0, 16, 32, 48 or 64");
    if (param == CFGMM_TUNE_DGEMM_SMALL_TILE && value != 0 && value != 64 && value != 128)
        return fail(CFGMM_EINVAL, "DGEMM small tile must be 0, 64 or 128");
    if (value < g_tune_lo[param] || value > g_tune_hi[param])
        return fail(CFGMM_EINVAL, "tunable " + std::to_string(param) + " out of range: " + std::to_string(value));
    g_tune[param].store(value);
    return CFGMM_OK;
}

int cfgmm_tune_get(int param, int *value) {
    clear_error();
    if (param < 0 || param >= CFGMM_TUNE_COUNT) return fail(CFGMM_EINVAL, "unknown tunable " + std::to_string(param));
    *value = tune_get(param);
    return CFGMM_OK;
}

int cfgmm_aodmao2rho_ex_dev(void *stream, double *rhos, const double *dms, const double *ao_lhs,
                            const double *ao_rhs, size_t stride_ao, size_t stride_grid, size_t nao, size_t ngrid,
                            size_t nset, const uint8_t *tab, size_t ldtab, size_t blksize, int precision,
                            int nslices) {
    return rho_dev_common(stream, rhos, dms, ao_lhs, ao_rhs, stride_ao, stride_grid, nao, ngrid, nset, tab, ldtab,
                          blksize, precision, nslices);
}

// ---- plumbing -------------------------------------------------------------------
int cfgmm_device_count(int *count) {
    clear_error();
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e == hipErrorNoDevice) n = 0, e = hipSuccess;
    CFGMM_HIP_TRY(e);
    *count = n;
    return CFGMM_OK;
}
int cfgmm_device_malloc(void **ptr, size_t bytes) {
    clear_error();
    CFGMM_TRY(check_device());
    return hip_status(hipMalloc(ptr, std::max<size_t>(bytes, 1)), "hipMalloc");
}
int cfgmm_device_free(void *ptr) { return hip_status(hipFree(ptr), "hipFree"); }
int cfgmm_memcpy_htod(void *dst, const void *src, size_t bytes) {
    return hip_status(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice), "hipMemcpy H2D");
}
int cfgmm_memcpy_dtoh(void *dst, const void *src, size_t bytes) {
    return hip_status(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost), "hipMemcpy D2H");
}
int cfgmm_memcpy_htod_async(void *dst, const void *src, size_t bytes, void *stream) {
    return hip_status(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, static_cast<hipStream_t>(stream)),
                      "hipMemcpyAsync H2D");
}
int cfgmm_memcpy_dtoh_async(void *dst, const void *src, size_t bytes, void *stream) {
    return hip_status(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, static_cast<hipStream_t>(stream)),
                      "hipMemcpyAsync D2H");
}
int cfgmm_memcpy_dtod_async(void *dst, const void *src, size_t bytes, void *stream) {
    return hip_status(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, static_cast<hipStream_t>(stream)),
                      "hipMemcpyAsync D2D");
}
int cfgmm_memset_dev(void *dst, int value, size_t bytes) {
    return hip_status(hipMemset(dst, value, bytes), "hipMemset");
}
int cfgmm_mem_get_info(size_t *free_bytes, size_t *total_bytes) {
    return hip_status(hipMemGetInfo(free_bytes, total_bytes), "hipMemGetInfo");
}
int cfgmm_stream_create(void **stream) {
    hipStream_t s = nullptr;
    CFGMM_HIP_TRY(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    *stream = s;
    return CFGMM_OK;
}
int cfgmm_stream_destroy(void *stream) {
    {
        int dev = 0;
        if (hipGetDevice(&dev) == hipSuccess) {
            std::lock_guard<std::mutex> lock(g_ws_mutex);
            auto it = g_ws.find(std::make_pair(dev, static_cast<hipStream_t>(stream)));
            if (it != g_ws.end()) {
                // a stream being destroyed has no calls in flight on other threads
                (void)hipStreamSynchronize(static_cast<hipStream_t>(stream));
                if (it->second.ptr) (void)hipFree(it->second.ptr);
                if (it->second.host) (void)hipHostFree(it->second.host);
                if (it->second.host_done) (void)hipEventDestroy(it->second.host_done);
                g_ws.erase(it);
            }
        }
    }
    return hip_status(hipStreamDestroy(static_cast<hipStream_t>(stream)), "hipStreamDestroy");
}
int cfgmm_stream_synchronize(void *stream) {
    return hip_status(hipStreamSynchronize(static_cast<hipStream_t>(stream)), "hipStreamSynchronize");
}
int cfgmm_device_synchronize(void) { return hip_status(hipDeviceSynchronize(), "hipDeviceSynchronize"); }
int cfgmm_event_create(void **event) {
    hipEvent_t e = nullptr;
    CFGMM_HIP_TRY(hipEventCreate(&e));
    *event = e;
    return CFGMM_OK;
}
int cfgmm_event_destroy(void *event) {
    return hip_status(hipEventDestroy(static_cast<hipEvent_t>(event)), "hipEventDestroy");
}
int cfgmm_event_record(void *event, void *stream) {
    return hip_status(hipEventRecord(static_cast<hipEvent_t>(event), static_cast<hipStream_t>(stream)),
                      "hipEventRecord");
}
int cfgmm_event_synchronize(void *event) {
    return hip_status(hipEventSynchronize(static_cast<hipEvent_t>(event)), "hipEventSynchronize");
}
int cfgmm_event_query(void *event) {
    hipError_t e = hipEventQuery(static_cast<hipEvent_t>(event));
    if (e == hipErrorNotReady) return 1;
    return hip_status(e, "hipEventQuery");
}
int cfgmm_event_elapsed_ms(float *ms, void *start, void *end) {
    return hip_status(
        hipEventElapsedTime(ms, static_cast<hipEvent_t>(start), static_cast<hipEvent_t>(end)),
        "hipEventElapsedTime");
}

int cfgmm_trace_enable(int on) {
    g_trace_on.store(on != 0);
    return CFGMM_OK;
}

int cfgmm_trace_reset(void) {
    std::lock_guard<std::mutex> lock(g_trace_mutex);
    for (auto &r : g_trace) {
        (void)hipEventSynchronize(r.end);
        (void)hipEventDestroy(r.start);
        (void)hipEventDestroy(r.end);
    }
    g_trace.clear();
    return CFGMM_OK;
}

int cfgmm_trace_read(const char *name, double *total_ms, long long *count) {
    clear_error();
    std::lock_guard<std::mutex> lock(g_trace_mutex);
    double tot = 0.0;
    long long n = 0;
    for (auto &r : g_trace) {
        if (r.name != name) continue;
        CFGMM_HIP_TRY(hipEventSynchronize(r.end));
        float ms = 0.f;
        CFGMM_HIP_TRY(hipEventElapsedTime(&ms, r.start, r.end));
        tot += ms;
        ++n;
    }
    *total_ms = tot;
    *count = n;
    return CFGMM_OK;
}

}  // extern "C"
