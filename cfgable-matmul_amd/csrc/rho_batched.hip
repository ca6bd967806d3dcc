// rho_batched.hip -- many independent masked rho problems in one persistent launch.
//
// No counterpart in the reference (aodmao2rho_loop_parallel handles one problem,
// src/impl_aodmao2rho.rs:171-244); each problem keeps exactly that contract.
//
//   1. batch_build_lists: the compacted AO lists of every (problem, grid block).
//   2. batch_count_chunks / batch_scan_offsets / batch_fill_tasks: one task per
//      (problem, block, 64-point subtile, 64-row chunk of the block's list, group
//      of density matrices) in a device queue; problems are numbered
//      largest-first by the host, each problem's tasks are contiguous and ordered
//      by the AO range of the chunk rows.
//   3. batch_rho_kernel:  a persistent grid (CUs x workgroups-per-CU, both
//      tunable) pulls tasks from per-XCD ranges of the array (atomic heads,
//      stealing from the other ranges at the end) until all are drained and
//      runs the same chunk body as the per-problem kernel (rho_body.hpp).
//   4. batch_reduce:      rho += sum of the chunk partials, in a fixed order.
// A problem whose lhs and rhs are one array and whose D_s are all symmetric (batch_symmetry, on
// the device) takes the upper-triangle form of the chunk body, as the per-problem call does.
// Every task writes its own partial slot, so the result does not depend on
// which workgroup ran which task: it is bitwise the same as a per-problem call.
#include "internal.hpp"
#include "reduce_waves.hpp"
#include "rho_body.hpp"

#include <algorithm>
#include <map>
#include <numeric>
#include <tuple>
#include <vector>

namespace cfgmm {
namespace {

using rho::kThreads;
using rho::kUC;

struct BatchProblem {
    double *rhos;
    const double *dms, *ao_lhs, *ao_rhs;
    const uint8_t *tab;
    long stride_ao, stride_grid, ldtab;
    int nao, ngrid, nset, blksize, nblk, nsub, nt;
    long lists_off, counts_off, partial_off;
    int sym;  // index + 1 of the problem's (D, stride, nao, nset) in the symmetry table; 0: lhs is not rhs
    int wave;  // reduction pass: problems whose rho ranges overlap add into rho in call order
    int call;  // index of the problem in the caller's array
};

// One distinct (D, stride_ao, nao, nset) of the batch whose problems have lhs == rhs; flag starts
// at 1 and batch_symmetry clears it unless every D_s is symmetric.
struct SymCheck {
    const double *dms;
    long stride_ao;
    int nao, nset;
    int flag;
};

// Workgroup (k, entry): the k-th pair (ti <= tj) of 64x64 tiles of the entry's D_s, every s;
// clears the entry's flag at a difference.
__global__ __launch_bounds__(256) void batch_symmetry(SymCheck *__restrict__ checks) {
    SymCheck &p = checks[blockIdx.y];
    // read once for the workgroup (another workgroup may clear it meanwhile): all threads leave
    // together or none does -- the comparison below has barriers
    __shared__ int go;
    if (threadIdx.x == 0) go = __hip_atomic_load(&p.flag, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __syncthreads();
    if (!go) return;
    const long k = blockIdx.x;
    // upper-triangle index k = tj (tj + 1) / 2 + ti, 0 <= ti <= tj
    long tj = (long)((sqrt(8.0 * (double)k + 1.0) - 1.0) / 2.0);
    while (tj * (tj + 1) / 2 > k) --tj;
    while ((tj + 1) * (tj + 2) / 2 <= k) ++tj;
    const int ti = (int)(k - tj * (tj + 1) / 2);
    if (tj * 64 >= p.nao) return;
    __shared__ unsigned long long t[64][65];
    for (int s = 0; s < p.nset; ++s) {
        if (rho::tile_pair_differs<256>(p.dms + (long)s * p.stride_ao * p.stride_ao, p.stride_ao, p.nao, ti, (int)tj, t)) {
            if (threadIdx.x == 0) __hip_atomic_store(&p.flag, 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            return;  // the answer is the workgroup's: every thread leaves together
        }
    }
}

__global__ __launch_bounds__(256) void batch_build_lists(const BatchProblem *__restrict__ probs, int *__restrict__ lists,
                                                         int *__restrict__ counts) {
    const BatchProblem p = probs[blockIdx.y];
    const int b = blockIdx.x;
    if (b >= p.nblk) return;
    __shared__ int wave_cnt[4];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    int *__restrict__ out = lists + p.lists_off + (long)b * p.nao;
    int base = 0;
    for (int x0 = 0; x0 < p.nao; x0 += 256) {
        const int x = x0 + threadIdx.x;
        const bool on = x < p.nao && p.tab[(long)x * p.ldtab + b] != 0;
        const unsigned long long bal = __ballot(on);
        const int prefix = __popcll(bal & ((1ull << lane) - 1ull));
        if (lane == 0) wave_cnt[wave] = __popcll(bal);
        __syncthreads();
        int off = base;
        for (int w = 0; w < wave; ++w) off += wave_cnt[w];
        if (on) out[off + prefix] = x;
        base += wave_cnt[0] + wave_cnt[1] + wave_cnt[2] + wave_cnt[3];
        __syncthreads();
    }
    if (threadIdx.x == 0) counts[p.counts_off + b] = base;
}

// Task word: x = problem, y = block, z = subtile << 16 | chunk, w = first density matrix.
//
// The queue holds the problems one after another (largest first); inside a problem the tasks
// are ordered (set group, chunk in AO order, subtile): the chunks are sorted by the AO range of
// their rows (bucketed midpoint of the first and last row), so that tasks sharing rows of D
// run together on one XCD and find them in its L2 (the per-problem kernel's plan, rho.hip).
// Three steps: chunk counts per problem, an exclusive scan into queue offsets, the fill.  The
// order inside an AO bucket follows atomics; no result depends on it.
constexpr int kFillThreads = 256;
constexpr int kFillBuckets = 256;

// cnum[pi] = number of non-empty chunks of problem pi
__global__ __launch_bounds__(kFillThreads) void batch_count_chunks(const BatchProblem *__restrict__ probs,
                                                                   const int *__restrict__ counts,
                                                                   int *__restrict__ cnum) {
    const BatchProblem p = probs[blockIdx.x];
    __shared__ int tot;
    if (threadIdx.x == 0) tot = 0;
    __syncthreads();
    int mine = 0;
    for (int b = threadIdx.x; b < p.nblk; b += kFillThreads) mine += (counts[p.counts_off + b] + kUC - 1) / kUC;
    atomicAdd(&tot, mine);
    __syncthreads();
    if (threadIdx.x == 0) cnum[blockIdx.x] = tot;
}

// toff[pi] = first queue slot of problem pi (tasks: groups x chunks x subtiles); ctr[0] = total
__global__ __launch_bounds__(1024) void batch_scan_offsets(const BatchProblem *__restrict__ probs, int np, int nsmax,
                                                           const int *__restrict__ cnum, int *__restrict__ toff,
                                                           int *__restrict__ ctr) {
    __shared__ long part[1024];
    const int tid = threadIdx.x;
    // each thread owns a contiguous run of problems
    const int per = (np + 1023) / 1024, lo = min(np, tid * per), hi = min(np, lo + per);
    long sum = 0;
    for (int i = lo; i < hi; ++i) {
        const BatchProblem &p = probs[i];
        sum += (long)cnum[i] * ((p.nset + nsmax - 1) / nsmax) * p.nsub;
    }
    part[tid] = sum;
    __syncthreads();
    for (int d = 1; d < 1024; d <<= 1) {
        const long add = tid >= d ? part[tid - d] : 0;
        __syncthreads();
        part[tid] += add;
        __syncthreads();
    }
    long run = part[tid] - sum;
    for (int i = lo; i < hi; ++i) {
        const BatchProblem &p = probs[i];
        toff[i] = (int)run;
        run += (long)cnum[i] * ((p.nset + nsmax - 1) / nsmax) * p.nsub;
    }
    if (tid == 1023) ctr[0] = (int)part[1023];
}

__global__ __launch_bounds__(kFillThreads) void batch_fill_tasks(const BatchProblem *__restrict__ probs,
                                                                 const int *__restrict__ lists,
                                                                 const int *__restrict__ counts, int nsmax,
                                                                 const int *__restrict__ cnum,
                                                                 const int *__restrict__ toff,
                                                                 int4 *__restrict__ tasks) {
    const int pi = blockIdx.x;
    const BatchProblem p = probs[pi];
    __shared__ int hist[kFillBuckets];
    const int tid = threadIdx.x;
    const int maxch = (p.nao + kUC - 1) / kUC;
    const long nslot = (long)p.nblk * maxch;
    auto key = [&](long sl) -> int {  // AO bucket of chunk slot sl, -1 if empty
        const int b = (int)(sl / maxch), c = (int)(sl % maxch);
        const int n = counts[p.counts_off + b];
        if (c * kUC >= n) return -1;
        const int *__restrict__ lb = lists + p.lists_off + (long)b * p.nao;
        const long mid = ((long)lb[c * kUC] + lb[min(c * kUC + kUC, n) - 1]) / 2;
        return (int)(mid * kFillBuckets / p.nao);
    };
    static_assert(kFillBuckets == kFillThreads, "one bucket per thread");
    hist[tid] = 0;
    __syncthreads();
    for (long sl = tid; sl < nslot; sl += kFillThreads) {
        const int k = key(sl);
        if (k >= 0) atomicAdd(&hist[k], 1);
    }
    __syncthreads();
    const int mine = hist[tid];
    int incl = mine;
    for (int d = 1; d < kFillBuckets; d <<= 1) {
        __syncthreads();
        const int add = tid >= d ? hist[tid - d] : 0;
        __syncthreads();
        incl += add;
        hist[tid] = incl;
    }
    __syncthreads();
    hist[tid] = incl - mine;
    __syncthreads();
    const int nchunks = cnum[pi], ngrp = (p.nset + nsmax - 1) / nsmax;
    int4 *__restrict__ out = tasks + toff[pi];
    for (long sl = tid; sl < nslot; sl += kFillThreads) {
        const int k = key(sl);
        if (k < 0) continue;
        const int pos = atomicAdd(&hist[k], 1);
        const int b = (int)(sl / maxch), c = (int)(sl % maxch);
        for (int g = 0; g < ngrp; ++g)
            for (int sub = 0; sub < p.nsub; ++sub)
                out[((long)g * nchunks + pos) * p.nsub + sub] = int4{pi, b, (sub << 16) | c, g * nsmax};
    }
}

__global__ void zero_int_kernel(int *p, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0;
}

template <int NS, bool SYM>
__device__ inline void run_chunk_nt(int nt, const rho::RhoChunk &t, unsigned char *smem) {
    switch (nt) {
        case 1: rho::rho_chunk_body<NS, 1, SYM>(t, smem); break;
        case 2: rho::rho_chunk_body<NS, 2, SYM>(t, smem); break;
        case 3: rho::rho_chunk_body<NS, 3, SYM>(t, smem); break;
        default: rho::rho_chunk_body<NS, 4, SYM>(t, smem); break;
    }
}
template <int NS>
__device__ inline void run_chunk_ns(int nt, bool sym, const rho::RhoChunk &t, unsigned char *smem) {
    if (sym)
        run_chunk_nt<NS, true>(nt, t, smem);
    else
        run_chunk_nt<NS, false>(nt, t, smem);
}

template <int NSMAX>
__global__ __launch_bounds__(kThreads) void batch_rho_kernel(const BatchProblem *__restrict__ probs,
                                                             const int4 *__restrict__ tasks,
                                                             const int *__restrict__ ntasks, int *__restrict__ head,
                                                             const int *__restrict__ lists,
                                                             const int *__restrict__ counts, double *__restrict__ partial,
                                                             const SymCheck *__restrict__ checks) {
    __shared__ __attribute__((aligned(16))) unsigned char smem[rho::chunk_lds_bytes<NSMAX, 4>()];
    __shared__ int cur;
    const int n = *ntasks;
    // The queue is cut into 8 contiguous ranges, one per XCD (workgroup i runs on XCD i % 8):
    // a block's run of tasks stays on one XCD and its L2.  A workgroup drains its own range,
    // then the others in turn, and leaves after at most 8 empty draws.
    int range = blockIdx.x % 8, empty = 0;
    for (;;) {
        if (threadIdx.x == 0) {
            int ti = n;
            while (empty < 8) {
                const int lo = (int)((long)n * range / 8), hi = (int)((long)n * (range + 1) / 8);
                const int c = lo + atomicAdd(&head[range], 1);
                if (c < hi) {
                    ti = c;
                    break;
                }
                ++empty;
                range = (range + 1) % 8;
            }
            cur = ti;
        }
        __syncthreads();
        const int ti = cur;
        __syncthreads();  // every thread has read `cur` before thread 0 overwrites it
        if (ti >= n) break;
        const int4 w = tasks[ti];
        const BatchProblem &p = probs[w.x];
        rho::RhoChunk t;
        t.partial = partial + p.partial_off;
        t.dms = p.dms;
        t.ao_lhs = p.ao_lhs;
        t.ao_rhs = p.ao_rhs;
        t.stride_ao = p.stride_ao;
        t.stride_grid = p.stride_grid;
        t.ngrid = p.ngrid;
        t.nset = p.nset;
        t.set0 = w.w;
        t.list = lists + p.lists_off + (long)w.y * p.nao;
        t.cnt = counts[p.counts_off + w.y];
        t.chunk = w.z & 0xffff;
        t.g0 = w.y * p.blksize + (w.z >> 16) * p.nt * 16;
        const bool sym = p.sym > 0 && checks[p.sym - 1].flag != 0;
        if (t.g0 >= p.ngrid) continue;  // subtile past the end of a partial last block (uniform)
        const int ns = min(NSMAX, p.nset - w.w);
        if (NSMAX >= 4 && ns == 4)
            run_chunk_ns<(NSMAX >= 4 ? 4 : 1)>(p.nt, sym, t, smem);
        else if (NSMAX >= 3 && ns == 3)
            run_chunk_ns<(NSMAX >= 3 ? 3 : 1)>(p.nt, sym, t, smem);
        else if (NSMAX >= 2 && ns == 2)
            run_chunk_ns<(NSMAX >= 2 ? 2 : 1)>(p.nt, sym, t, smem);
        else
            run_chunk_ns<1>(p.nt, sym, t, smem);
    }
}

__global__ void batch_reduce(const BatchProblem *__restrict__ probs, const int *__restrict__ counts,
                             const double *__restrict__ partial, int wave) {
    const BatchProblem p = probs[blockIdx.y];
    if (p.wave != wave) return;
    const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (long)p.nset * p.ngrid) return;
    const int s = (int)(idx / p.ngrid), g = (int)(idx % p.ngrid);
    const int nch = (counts[p.counts_off + g / p.blksize] + kUC - 1) / kUC;
    const double *__restrict__ part = partial + p.partial_off + (long)s * p.ngrid + g;
    const long step = (long)p.nset * p.ngrid;
    // eight loads in flight at a time (clamped indices, no branch around a load), summed in chunk order
    double v = 0.0;
    for (int c0 = 0; c0 < nch; c0 += 8) {
        double x[8];
#pragma unroll
        for (int i = 0; i < 8; ++i) x[i] = part[min(c0 + i, nch - 1) * step];
#pragma unroll
        for (int i = 0; i < 8; ++i) v = c0 + i < nch ? v + x[i] : v;
    }
    p.rhos[(long)s * p.stride_grid + g] += v;
}

int assign_reduce_waves(std::vector<BatchProblem> &bp) {
    // bp is in launch order (largest first); rank[i] = position of bp[i] in call order
    const size_t n = bp.size();
    std::vector<size_t> by_call(n), rank(n);
    std::iota(by_call.begin(), by_call.end(), 0);
    std::sort(by_call.begin(), by_call.end(), [&](size_t x, size_t y) { return bp[x].call < bp[y].call; });
    for (size_t r = 0; r < n; ++r) rank[by_call[r]] = r;
    std::vector<RhoSpan> spans(n);
    for (size_t i = 0; i < n; ++i) {
        const BatchProblem &p = bp[i];
        const uintptr_t lo = reinterpret_cast<uintptr_t>(p.rhos);
        spans[i] = RhoSpan{lo, lo + sizeof(double) * ((uintptr_t)(p.nset - 1) * p.stride_grid + p.ngrid),
                           (int)rank[i]};
    }
    std::vector<int> wave_of;
    const int nwaves = assign_reduce_waves_spans(spans, wave_of);
    for (size_t i = 0; i < n; ++i) bp[i].wave = wave_of[rank[i]];
    return nwaves;
}

int tile_nt(size_t blksize) {
    const size_t q = blksize / 16;
    for (int nt = 4; nt >= 1; --nt)
        if (q % nt == 0) return nt;
    return 1;
}

}  // namespace

int launch_rho_batched(hipStream_t stream, const cfgmm_rho_problem *problems, size_t nprob) {
    if (nprob == 0) return CFGMM_OK;
    if (nprob > 65535) return fail(CFGMM_EINVAL, "batched rho: at most 65535 problems per call");
    // sets per task: 1, 2 or 4 -- the task builder and the kernel must agree
    const int nsmax_req = tune_get(CFGMM_TUNE_BATCH_MAX_SETS);
    const int nsmax = nsmax_req >= 4 ? 4 : (nsmax_req >= 2 ? 2 : 1);
    const int wgs_per_cu = tune_get(CFGMM_TUNE_BATCH_WGS_PER_CU);
    // largest problems first, so the persistent grid drains big tasks early
    std::vector<size_t> order(nprob);
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(), [&](size_t x, size_t y) {
        const auto &a = problems[x], &b = problems[y];
        return (double)a.nao * a.nao * a.ngrid * a.nset > (double)b.nao * b.nao * b.ngrid * b.nset;
    });
    std::vector<BatchProblem> bp;
    bp.reserve(nprob);
    long lists_n = 0, counts_n = 0, partial_n = 0, tasks_max = 0, max_pairs = 0;
    std::vector<SymCheck> checks;
    std::map<std::tuple<const double *, size_t, size_t, size_t>, int> sym_index;
    int max_nblk = 0;
    long max_elems = 0;
    for (size_t oi : order) {
        const cfgmm_rho_problem &q = problems[oi];
        if (q.nao == 0 || q.ngrid == 0 || q.nset == 0) continue;
        if (q.nao > INT32_MAX || q.ngrid > INT32_MAX || q.nset > 65535)
            return fail(CFGMM_EINVAL, "batched rho: problem size out of range");
        BatchProblem p{};
        p.call = (int)oi;
        p.rhos = q.rhos;
        p.dms = q.dms;
        p.ao_lhs = q.ao_lhs;
        p.ao_rhs = q.ao_rhs;
        p.tab = q.tab;
        p.stride_ao = (long)q.stride_ao;
        p.stride_grid = (long)q.stride_grid;
        p.ldtab = (long)q.ldtab;
        p.nao = (int)q.nao;
        p.ngrid = (int)q.ngrid;
        p.nset = (int)q.nset;
        p.blksize = (int)q.blksize;
        p.nblk = (int)div_up(q.ngrid, q.blksize);
        p.nt = tile_nt(q.blksize);
        p.nsub = (int)(q.blksize / (16 * p.nt));
        p.lists_off = lists_n;
        p.counts_off = counts_n;
        p.partial_off = partial_n;
        if (q.ao_lhs == q.ao_rhs && tune_get(CFGMM_TUNE_RHO_SYMMETRIC) == 0) {
            // problems sharing D (and its used extent) share one test
            const auto key = std::make_tuple(q.dms, q.stride_ao, q.nao, q.nset);
            auto it = sym_index.find(key);
            if (it == sym_index.end()) {
                it = sym_index.emplace(key, (int)checks.size()).first;
                checks.push_back(SymCheck{q.dms, (long)q.stride_ao, (int)q.nao, (int)q.nset, 1});
                max_pairs = std::max(max_pairs, (long)div_up(q.nao, 64) * ((long)div_up(q.nao, 64) + 1) / 2);
            }
            p.sym = it->second + 1;
        }
        const long maxchunks = (long)div_up(q.nao, kUC);
        if (maxchunks > 0xffff || p.nsub > 0x7fff) return fail(CFGMM_EINVAL, "batched rho: problem too large");
        lists_n += (long)p.nblk * p.nao;
        counts_n += p.nblk;
        partial_n += maxchunks * p.nset * p.ngrid;
        tasks_max += (long)p.nblk * p.nsub * maxchunks * (long)div_up(q.nset, nsmax);
        max_nblk = std::max(max_nblk, p.nblk);
        max_elems = std::max(max_elems, (long)p.nset * p.ngrid);
        bp.push_back(p);
    }
    if (bp.empty()) return CFGMM_OK;
    if (tasks_max > INT32_MAX) return fail(CFGMM_EINVAL, "batched rho: too many tasks for one launch");
    const int nwaves = assign_reduce_waves(bp);
    const int np = (int)bp.size();

    auto al = [](size_t x) { return div_up(x, 256) * 256; };
    const size_t o_lists = al(sizeof(BatchProblem) * np);
    const size_t o_counts = o_lists + al(sizeof(int) * lists_n);
    const size_t o_part = o_counts + al(sizeof(int) * counts_n);
    const size_t o_tasks = o_part + al(sizeof(double) * partial_n);
    const size_t o_cnum = o_tasks + al(sizeof(int4) * tasks_max);
    const size_t o_toff = o_cnum + al(sizeof(int) * np);
    const size_t o_ctr = o_toff + al(sizeof(int) * np);
    void *ws = nullptr;
    const size_t o_chk = o_ctr + al(9 * sizeof(int));
    int st = workspace_get(stream, o_chk + sizeof(SymCheck) * std::max<size_t>(checks.size(), 1), &ws);
    if (st != CFGMM_OK) return st;
    char *base = static_cast<char *>(ws);
    BatchProblem *d_probs = reinterpret_cast<BatchProblem *>(base);
    int *lists = reinterpret_cast<int *>(base + o_lists);
    int *counts = reinterpret_cast<int *>(base + o_counts);
    double *partial = reinterpret_cast<double *>(base + o_part);
    int4 *tasks = reinterpret_cast<int4 *>(base + o_tasks);
    int *cnum = reinterpret_cast<int *>(base + o_cnum);
    int *toff = reinterpret_cast<int *>(base + o_toff);
    int *ctr = reinterpret_cast<int *>(base + o_ctr);  // [0] = task count, [1..8] = heads of the XCD ranges

    // Pageable source: the copy is complete when hipMemcpyAsync returns, so `bp` may go out of scope.
    CFGMM_HIP_TRY(hipMemcpyAsync(d_probs, bp.data(), sizeof(BatchProblem) * np, hipMemcpyHostToDevice, stream));
    SymCheck *d_checks = reinterpret_cast<SymCheck *>(base + o_chk);
    if (!checks.empty())
        CFGMM_HIP_TRY(hipMemcpyAsync(d_checks, checks.data(), sizeof(SymCheck) * checks.size(), hipMemcpyHostToDevice,
                                     stream));
    hipEvent_t t0 = trace_start(stream);
    hipLaunchKernelGGL(zero_int_kernel, dim3(1), dim3(64), 0, stream, ctr, 9);
    hipLaunchKernelGGL(batch_build_lists, dim3((unsigned)max_nblk, (unsigned)np), dim3(256), 0, stream, d_probs,
                       lists, counts);
    hipLaunchKernelGGL(batch_count_chunks, dim3((unsigned)np), dim3(kFillThreads), 0, stream, d_probs, counts, cnum);
    hipLaunchKernelGGL(batch_scan_offsets, dim3(1), dim3(1024), 0, stream, d_probs, np, nsmax, cnum, toff, ctr);
    hipLaunchKernelGGL(batch_fill_tasks, dim3((unsigned)np), dim3(kFillThreads), 0, stream, d_probs, lists, counts,
                       nsmax, cnum, toff, tasks);
    if (!checks.empty())  // the symmetric form where lhs is rhs and every D_s is symmetric
        hipLaunchKernelGGL(batch_symmetry, dim3((unsigned)max_pairs, (unsigned)checks.size()), dim3(256), 0, stream,
                           d_checks);
    trace_stop(stream, "batch_prepare", t0);
    st = hip_status(hipGetLastError(), "batched rho prepare launch");
    if (st != CFGMM_OK) return st;

    const int grid = std::max(1, device_cu_count() * wgs_per_cu);
    t0 = trace_start(stream);
    if (nsmax >= 4)
        hipLaunchKernelGGL((batch_rho_kernel<4>), dim3(grid), dim3(kThreads), 0, stream, d_probs, tasks, ctr, ctr + 1,
                           lists, counts, partial, d_checks);
    else if (nsmax == 2)
        hipLaunchKernelGGL((batch_rho_kernel<2>), dim3(grid), dim3(kThreads), 0, stream, d_probs, tasks, ctr, ctr + 1,
                           lists, counts, partial, d_checks);
    else
        hipLaunchKernelGGL((batch_rho_kernel<1>), dim3(grid), dim3(kThreads), 0, stream, d_probs, tasks, ctr, ctr + 1,
                           lists, counts, partial, d_checks);
    trace_stop(stream, "batch_rho_kernel", t0);
    st = hip_status(hipGetLastError(), "batch_rho_kernel launch");
    if (st != CFGMM_OK) return st;
    // one pass per wave, in wave order: problems that add into the same rho elements do so in the
    // order of the call, as separate calls would
    t0 = trace_start(stream);
    for (int w = 0; w < nwaves; ++w)
        hipLaunchKernelGGL(batch_reduce, dim3((unsigned)div_up(max_elems, 256), (unsigned)np), dim3(256), 0, stream,
                           d_probs, counts, partial, w);
    trace_stop(stream, "batch_reduce", t0);
    return hip_status(hipGetLastError(), "batch_reduce launch");
}

}  // namespace cfgmm
