// server.cpp -- a continuously batching request loop for masked rho and DGEMM jobs on one GPU.
//
// Requests (rho over a catalogued molecule, or a DGEMM whose operands are
// generated on the device from a seed) arrive at given times.  The loop admits
// a request when a slot -- a stream of its own -- is free and the device
// memory it needs fits the HBM budget next to what is already held: resident
// molecules (D, phi and the mask, and in split precision their per-block
// tiles) and in-flight requests.  Molecules that no in-flight request uses are
// evicted least-recently-used first when memory is short.  At every dispatch
// the pending FP64 rho requests are fused into one batched persistent launch
// (up to max_batch of them); split-precision rho requests run over their
// molecules' resident tiles, one evaluation each, several per slot.  Work on different slots overlaps on the device.  Each result is
// hashed on the device (cfgmm checksum: order-independent over the result
// bits), so a caller can compare a request's result alone and in the mix
// exactly; full results are copied back on request after the run.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <string>
#include <thread>
#include <vector>

#include "internal.hpp"

namespace cfgmm {

namespace {

using Clock = std::chrono::steady_clock;

double seconds_since(Clock::time_point t0) {
    return std::chrono::duration<double>(Clock::now() - t0).count();
}

size_t round_alloc(size_t bytes) { return div_up(std::max<size_t>(bytes, 1), 4096) * 4096; }

}  // namespace

struct ServerMolecule {
    int id = 0;
    const double *h_dms = nullptr, *h_phi = nullptr;
    const uint8_t *h_tab = nullptr;
    size_t nao = 0, ngrid = 0, nset = 0, blksize = 0, nblk = 0;
    // split precision: rho over resident tiles; false in FP64, and for a molecule whose tiles could
    // take more than the configured cap (its requests then run in FP64 from D, phi and the mask)
    bool split = false;
    // residency
    bool resident = false;
    void *d_dms = nullptr, *d_phi = nullptr, *d_tab = nullptr;
    ResidentRho *tiles = nullptr;
    size_t bytes = 0;
    int inflight = 0;
    unsigned long long last_use = 0;
};

struct ServerSlot {
    hipStream_t stream = nullptr;
    hipEvent_t done = nullptr;
    bool busy = false;
    int batch_id = -1;
    std::vector<size_t> members;        // request indices
    std::vector<void *> allocs;         // device blocks released at completion
    size_t bytes = 0;                   // budget charged for this batch
};

struct Server {
    cfgmm_server_config cfg{};
    int device = 0;
    std::map<int, ServerMolecule> mols;
    std::vector<ServerSlot> slots;
    size_t held = 0;        // bytes charged: resident molecules + in-flight batches
    size_t peak = 0;
    unsigned long long tick = 0;
    unsigned long long *h_sums = nullptr;  // pinned, one checksum per request of the current run
    unsigned long long *d_sums = nullptr;
    size_t sums_cap = 0;
    cfgmm_server_stats stats{};
    std::vector<std::pair<size_t, void *>> keep;  // (request, device result) copied back after the run

    ~Server() {
        for (auto &sl : slots) {
            if (sl.stream) (void)hipStreamSynchronize(sl.stream);
        }
        for (auto &kv : mols) evict(kv.second);
        for (auto &sl : slots) {
            if (sl.done) (void)hipEventDestroy(sl.done);
            if (sl.stream) (void)hipStreamDestroy(sl.stream);
        }
        if (h_sums) (void)hipHostFree(h_sums);
        if (d_sums) (void)hipFree(d_sums);
    }

    void evict(ServerMolecule &m) {
        if (!m.resident) return;
        if (m.tiles) (void)resident_destroy(m.tiles);
        if (m.d_dms) (void)hipFree(m.d_dms);
        if (m.d_phi) (void)hipFree(m.d_phi);
        if (m.d_tab) (void)hipFree(m.d_tab);
        m.tiles = nullptr;
        m.d_dms = m.d_phi = m.d_tab = nullptr;
        held -= m.bytes;
        m.bytes = 0;
        m.resident = false;
    }

    // Frees idle resident molecules, least recently used first, until `need` more bytes fit.
    bool make_room(size_t need, int keep_id) {
        while (held + need > cfg.hbm_budget_bytes) {
            ServerMolecule *victim = nullptr;
            for (auto &kv : mols) {
                ServerMolecule &m = kv.second;
                if (!m.resident || m.inflight > 0 || m.id == keep_id) continue;
                if (!victim || m.last_use < victim->last_use) victim = &m;
            }
            if (!victim) return false;
            evict(*victim);
            stats.evictions++;
        }
        return true;
    }

    size_t molecule_bytes(const ServerMolecule &m) const {
        size_t b = round_alloc(sizeof(double) * m.nset * m.nao * m.nao) + round_alloc(sizeof(double) * m.nao * m.ngrid) +
                   round_alloc(m.nao * m.nblk);
        return b;
    }

    // Frees whatever device memory a molecule that is not (yet) resident holds, after the work
    // on `stream` that may still use it; leaves the molecule clean for a later upload.
    void discard(ServerMolecule &m, hipStream_t stream) {
        (void)hipStreamSynchronize(stream);
        if (m.tiles) (void)resident_destroy(m.tiles);
        if (m.d_dms) (void)hipFree(m.d_dms);
        if (m.d_phi) (void)hipFree(m.d_phi);
        if (m.d_tab) (void)hipFree(m.d_tab);
        m.tiles = nullptr;
        m.d_dms = m.d_phi = m.d_tab = nullptr;
        m.bytes = 0;
        m.resident = false;
    }

    // Allocations, copies and (split precision) tiles of a molecule, into m's fields.
    int upload(ServerMolecule &m, hipStream_t stream) {
        CFGMM_HIP_TRY(hipMalloc(&m.d_dms, sizeof(double) * m.nset * m.nao * m.nao));
        CFGMM_HIP_TRY(hipMalloc(&m.d_phi, sizeof(double) * m.nao * m.ngrid));
        CFGMM_HIP_TRY(hipMalloc(&m.d_tab, std::max<size_t>(m.nao * m.nblk, 1)));
        CFGMM_HIP_TRY(hipMemcpyAsync(m.d_dms, m.h_dms, sizeof(double) * m.nset * m.nao * m.nao, hipMemcpyHostToDevice,
                                     stream));
        CFGMM_HIP_TRY(
            hipMemcpyAsync(m.d_phi, m.h_phi, sizeof(double) * m.nao * m.ngrid, hipMemcpyHostToDevice, stream));
        CFGMM_HIP_TRY(hipMemcpyAsync(m.d_tab, m.h_tab, m.nao * m.nblk, hipMemcpyHostToDevice, stream));
        if (m.split) {
            ResidentRho *tiles = nullptr;
            const int st = resident_create(stream, static_cast<double *>(m.d_dms), static_cast<double *>(m.d_phi),
                                           static_cast<double *>(m.d_phi), m.nao, m.ngrid, m.nao, m.ngrid, m.nset,
                                           static_cast<uint8_t *>(m.d_tab), m.nblk, m.blksize, cfg.rho_nslices,
                                           &tiles);
            if (st != CFGMM_OK) return st;
            m.tiles = tiles;
        }
        return hip_status(hipStreamSynchronize(stream), "server: molecule upload");
    }

    // Uploads a molecule (and, in split precision, builds its tiles) on `stream`; synchronous.
    // The molecule is marked resident and charged to the budget only once everything succeeded;
    // on any failure what was allocated is freed and the molecule stays as it was.
    int make_resident(ServerMolecule &m, hipStream_t stream) {
        if (m.resident) return CFGMM_OK;
        const size_t base = molecule_bytes(m);
        if (!make_room(base, m.id)) return fail(CFGMM_EOOM, "server: molecule does not fit the HBM budget");
        const int st = upload(m, stream);
        if (st != CFGMM_OK) {
            const std::string msg = cfgmm_last_error();  // discard() may overwrite it
            discard(m, stream);
            return fail(st, msg);
        }
        // the tiles replace D for evaluations; D and phi (small next to the tiles) stay for re-creation
        m.bytes = base + (m.tiles ? resident_bytes(m.tiles) : 0);
        m.resident = true;
        held += m.bytes;
        if (held > cfg.hbm_budget_bytes && !make_room(0, m.id)) {
            evict(m);
            return fail(CFGMM_EOOM, "server: molecule tiles do not fit the HBM budget");
        }
        stats.uploads++;
        peak = std::max(peak, held);
        return CFGMM_OK;
    }

    // Row stride of a rho result on the device: the FP64 kernels index rho[s*stride_grid + g]
    // with the molecule's own grid as stride_grid (as the phi rows); the resident tiles take
    // any stride, so split-precision results are compact.
    size_t rho_ld(const cfgmm_request &r) const {
        auto it = mols.find(r.molecule);
        return it == mols.end() || it->second.split ? r.ngrid : it->second.ngrid;
    }
    size_t rho_elems(const cfgmm_request &r) const { return (r.nset - 1) * rho_ld(r) + r.ngrid; }

    size_t request_bytes(const cfgmm_request &r) const {
        if (r.kind == CFGMM_REQ_GEMM) return round_alloc(sizeof(double) * (r.m * r.k + r.k * r.n + r.m * r.n));
        auto it = mols.find(r.molecule);
        const size_t nao = it == mols.end() ? 0 : it->second.nao;
        // result + chunk partials of the rho kernels (per-stream workspace)
        return round_alloc(sizeof(double) * rho_elems(r)) +
               round_alloc(sizeof(double) * div_up(nao, 64) * r.nset * r.ngrid) +
               round_alloc(sizeof(int) * (nao + 1) * div_up(r.ngrid, 64));
    }
};

namespace {

int validate_request(const Server &sv, const cfgmm_request &r) {
    if (r.kind == CFGMM_REQ_GEMM) {
        if (r.m == 0 || r.n == 0 || r.k == 0) return fail(CFGMM_EINVAL, "server: empty GEMM request");
        return CFGMM_OK;
    }
    if (r.kind != CFGMM_REQ_RHO) return fail(CFGMM_EINVAL, "server: unknown request kind");
    auto it = sv.mols.find(r.molecule);
    if (it == sv.mols.end()) return fail(CFGMM_EINVAL, "server: unknown molecule " + std::to_string(r.molecule));
    const ServerMolecule &m = it->second;
    if (r.ngrid == 0 || r.nset == 0 || r.ngrid > m.ngrid || r.nset > m.nset)
        return fail(CFGMM_EINVAL_DIMS, "server: rho request exceeds its molecule");
    if (r.ngrid != m.ngrid && r.ngrid % m.blksize != 0)
        return fail(CFGMM_EINVAL_DIMS, "server: rho ngrid must be the molecule's or a multiple of blksize");
    return CFGMM_OK;
}

}  // namespace

int server_create(const cfgmm_server_config *cfg, Server **out) {
    *out = nullptr;
    if (cfg->slots < 1 || cfg->slots > 16) return fail(CFGMM_EINVAL, "server: slots must be in [1, 16]");
    if (cfg->max_batch < 1) return fail(CFGMM_EINVAL, "server: max_batch must be >= 1");
    if (cfg->rho_precision != CFGMM_PREC_FP64 && cfg->rho_precision != CFGMM_PREC_OZAKI_INT8)
        return fail(CFGMM_EINVAL, "server: unknown rho precision");
    if (cfg->rho_precision == CFGMM_PREC_OZAKI_INT8 && (cfg->rho_nslices < 2 || cfg->rho_nslices > 6))
        return fail(CFGMM_EINVAL, "server: rho_nslices must be in [2, 6]");
    auto sv = std::make_unique<Server>();
    sv->cfg = *cfg;
    CFGMM_HIP_TRY(hipGetDevice(&sv->device));
    sv->slots.resize(cfg->slots);
    for (auto &sl : sv->slots) {
        CFGMM_HIP_TRY(hipStreamCreateWithFlags(&sl.stream, hipStreamNonBlocking));
        CFGMM_HIP_TRY(hipEventCreateWithFlags(&sl.done, hipEventDisableTiming));
    }
    *out = sv.release();
    return CFGMM_OK;
}

int server_destroy(Server *sv) {
    delete sv;
    return CFGMM_OK;
}

int server_add_molecule(Server *sv, int id, const double *dms, const double *phi, const uint8_t *tab, size_t nao,
                        size_t ngrid, size_t nset, size_t blksize) {
    if (sv->mols.count(id)) return fail(CFGMM_EINVAL, "server: molecule id already registered");
    if (!dms || !phi || !tab || nao == 0 || ngrid == 0 || nset == 0)
        return fail(CFGMM_EINVAL, "server: empty molecule or null pointer");
    if (blksize == 0 || blksize % 16 != 0) return fail(CFGMM_EINVAL, "server: blksize must be a multiple of 16");
    ServerMolecule m;
    m.id = id;
    m.h_dms = dms;
    m.h_phi = phi;
    m.h_tab = tab;
    m.nao = nao;
    m.ngrid = ngrid;
    m.nset = nset;
    m.blksize = blksize;
    m.nblk = div_up(ngrid, blksize);
    if (sv->cfg.rho_precision == CFGMM_PREC_OZAKI_INT8) {
        // the tiles' size: the upper-triangle form when every D_s is symmetric (phi is on both
        // sides), by the rule resident_create applies on the device: each pair of entries equal
        // bitwise or within 2^-50 of the larger magnitude
        bool sym = true;
        for (size_t s = 0; s < nset && sym; ++s) {
            const double *d = dms + s * nao * nao;
            for (size_t i = 0; i < nao && sym; ++i)
                for (size_t j = i + 1; j < nao; ++j) {
                    const double x = d[i * nao + j], y = d[j * nao + i];
                    const bool same = std::memcmp(&x, &y, sizeof(double)) == 0 ||
                                      std::fabs(x - y) <= 0x1p-50 * std::fmax(std::fabs(x), std::fabs(y));
                    if (!same) {
                        sym = false;
                        break;
                    }
                }
        }
        const size_t cap = sv->cfg.split_tile_cap_bytes ? sv->cfg.split_tile_cap_bytes : sv->cfg.hbm_budget_bytes / 4;
        m.split = resident_bytes_bound(tab, m.nblk, nao, m.nblk, nset, blksize, sv->cfg.rho_nslices, sym) <= cap;
    }
    sv->mols[id] = m;
    return CFGMM_OK;
}

int server_warm(Server *sv, const int *ids, size_t n) {
    for (size_t i = 0; i < n; ++i) {
        auto it = sv->mols.find(ids[i]);
        if (it == sv->mols.end()) return fail(CFGMM_EINVAL, "server: unknown molecule");
        ServerMolecule &m = it->second;
        if (m.resident) continue;
        size_t need = sv->molecule_bytes(m);
        if (sv->held + need > sv->cfg.hbm_budget_bytes) break;  // warm only what fits without evicting
        int st = sv->make_resident(m, sv->slots[0].stream);
        if (st != CFGMM_OK) return st;
        m.last_use = ++sv->tick;
    }
    return CFGMM_OK;
}

int server_run(Server *sv, const cfgmm_request *reqs, size_t n, cfgmm_request_result *res) {
    for (size_t i = 0; i < n; ++i) {
        int st = validate_request(*sv, reqs[i]);
        if (st != CFGMM_OK) return fail(st, "request " + std::to_string(i) + ": " + std::string(cfgmm_last_error()));
    }
    if (sv->sums_cap < n) {
        if (sv->h_sums) (void)hipHostFree(sv->h_sums);
        if (sv->d_sums) (void)hipFree(sv->d_sums);
        sv->h_sums = nullptr;
        sv->d_sums = nullptr;
        CFGMM_HIP_TRY(hipHostMalloc(reinterpret_cast<void **>(&sv->h_sums), sizeof(unsigned long long) * std::max<size_t>(n, 1)));
        CFGMM_HIP_TRY(hipMalloc(&sv->d_sums, sizeof(unsigned long long) * std::max<size_t>(n, 1)));
        sv->sums_cap = n;
    }
    for (size_t i = 0; i < n; ++i) {
        res[i] = cfgmm_request_result{};
        res[i].status = -1;
        res[i].batch = -1;
    }
    std::vector<size_t> order(n);
    for (size_t i = 0; i < n; ++i) order[i] = i;
    std::stable_sort(order.begin(), order.end(),
                     [&](size_t a, size_t b) { return reqs[a].arrival_s < reqs[b].arrival_s; });
    std::deque<size_t> pending;
    size_t next = 0, done = 0;
    int batches = 0;
    sv->keep.clear();
    const auto t0 = Clock::now();

    // Per-request buffers come from the stream-ordered pool (hipMallocAsync / hipFreeAsync):
    // no device-wide synchronization, memory reused across requests.
    auto release = [&](ServerSlot &sl) {
        for (void *p : sl.allocs) (void)hipFreeAsync(p, sl.stream);
        sl.allocs.clear();
        sv->held -= sl.bytes;
        sl.bytes = 0;
    };

    int status = CFGMM_OK;
    while (done < n && status == CFGMM_OK) {
        bool progress = false;
        const double now = seconds_since(t0);
        while (next < n && reqs[order[next]].arrival_s <= now) pending.push_back(order[next++]);

        // completions
        for (auto &sl : sv->slots) {
            if (!sl.busy) continue;
            const hipError_t q = hipEventQuery(sl.done);
            if (q == hipErrorNotReady) continue;
            if (q != hipSuccess) {
                status = hip_status(q, "server: completion");
                break;
            }
            const double t = seconds_since(t0);
            for (size_t r : sl.members) {
                res[r].done_s = t;
                res[r].checksum = sv->h_sums[r];
                res[r].status = CFGMM_OK;
                if (reqs[r].kind == CFGMM_REQ_RHO) sv->mols[reqs[r].molecule].inflight--;
            }
            done += sl.members.size();
            sl.members.clear();
            release(sl);
            sl.busy = false;
            progress = true;
        }
        if (status != CFGMM_OK) break;

        // dispatch: one batch per free slot
        for (auto &sl : sv->slots) {
            if (sl.busy || pending.empty()) continue;
            std::vector<size_t> batch;
            size_t batch_bytes = 0;     // request buffers, charged to the slot
            size_t upload_bytes = 0;    // molecules this batch makes resident (charged when uploaded)
            std::vector<int> uploads;
            int batch_kind = -1;
            for (auto it = pending.begin(); it != pending.end();) {
                const cfgmm_request &r = reqs[*it];
                if (batch_kind >= 0 && (r.kind != batch_kind || r.kind == CFGMM_REQ_GEMM)) {
                    ++it;
                    continue;
                }
                if ((int)batch.size() >= sv->cfg.max_batch) break;
                size_t need = sv->request_bytes(r);
                const bool upload = r.kind == CFGMM_REQ_RHO && !sv->mols[r.molecule].resident &&
                                    std::find(uploads.begin(), uploads.end(), r.molecule) == uploads.end();
                if (upload) need += sv->molecule_bytes(sv->mols[r.molecule]);
                const size_t want = batch_bytes + upload_bytes + need;
                if (sv->held + want > sv->cfg.hbm_budget_bytes &&
                    !sv->make_room(want, r.kind == CFGMM_REQ_RHO ? r.molecule : -1)) {
                    ++it;  // does not fit now: later, smaller requests may
                    continue;
                }
                if (upload) {
                    uploads.push_back(r.molecule);
                    upload_bytes += sv->molecule_bytes(sv->mols[r.molecule]);
                }
                batch.push_back(*it);
                batch_bytes += sv->request_bytes(r);
                batch_kind = r.kind;
                if (r.kind == CFGMM_REQ_RHO) {
                    // pin the molecule now, so that admitting the rest of the batch cannot evict it
                    ServerMolecule &m = sv->mols[r.molecule];
                    m.inflight++;
                    m.last_use = ++sv->tick;
                }
                it = pending.erase(it);
                if (r.kind == CFGMM_REQ_GEMM) break;
            }
            if (batch.empty()) continue;
            const double t_admit = seconds_since(t0);
            sl.batch_id = batches++;
            sl.members = batch;
            sl.bytes = batch_bytes;
            sv->held += batch_bytes;
            sv->peak = std::max(sv->peak, sv->held);
            hipStream_t s = sl.stream;
            int st = CFGMM_OK;
            if (batch_kind == CFGMM_REQ_GEMM) {
                const size_t r = batch[0];
                const cfgmm_request &q = reqs[r];
                double *a = nullptr, *b = nullptr, *c = nullptr;
                st = hip_status(hipMallocAsync(reinterpret_cast<void **>(&a), sizeof(double) * q.m * q.k, s),
                                "server: hipMallocAsync");
                if (st == CFGMM_OK)
                    st = hip_status(hipMallocAsync(reinterpret_cast<void **>(&b), sizeof(double) * q.k * q.n, s),
                                    "server: hipMallocAsync");
                if (st == CFGMM_OK)
                    st = hip_status(hipMallocAsync(reinterpret_cast<void **>(&c), sizeof(double) * q.m * q.n, s),
                                    "server: hipMallocAsync");
                if (a) sl.allocs.push_back(a);
                if (b) sl.allocs.push_back(b);
                if (st == CFGMM_OK) st = launch_fill_uniform(s, a, q.m * q.k, 2 * q.seed);
                if (st == CFGMM_OK) st = launch_fill_uniform(s, b, q.k * q.n, 2 * q.seed + 1);
                if (st == CFGMM_OK) st = launch_fill_zero(s, c, sizeof(double) * q.m * q.n);
                if (st == CFGMM_OK) st = launch_dgemm(s, c, a, b, q.m, q.n, q.k, q.k, q.n, q.n, false, false);
                if (st == CFGMM_OK) st = launch_checksum(s, c, 1, q.m * q.n, q.m * q.n, sv->d_sums + r);
                if (c) {
                    if (q.out)
                        sv->keep.emplace_back(r, c);
                    else
                        sl.allocs.push_back(c);
                }
                sv->stats.gemm_requests++;
            } else {
                std::vector<cfgmm_rho_problem> probs;
                std::vector<double *> outs;
                for (size_t r : batch) {
                    const cfgmm_request &q = reqs[r];
                    ServerMolecule &m = sv->mols[q.molecule];
                    if (!m.resident) {
                        sv->stats.misses++;
                        st = sv->make_resident(m, s);
                        if (st != CFGMM_OK) break;
                    } else {
                        sv->stats.hits++;
                    }
                    double *o = nullptr;
                    st = hip_status(hipMallocAsync(reinterpret_cast<void **>(&o), sizeof(double) * sv->rho_elems(q), s),
                                    "server: hipMallocAsync");
                    if (st != CFGMM_OK) break;
                    outs.push_back(o);
                    if (q.out)
                        sv->keep.emplace_back(r, o);
                    else
                        sl.allocs.push_back(o);
                    st = launch_fill_zero(s, o, sizeof(double) * sv->rho_elems(q));
                    if (st != CFGMM_OK) break;
                    if (!m.split) {
                        sv->stats.rho_fp64_requests++;
                        cfgmm_rho_problem p{};
                        p.rhos = o;
                        p.dms = static_cast<double *>(m.d_dms);
                        p.ao_lhs = p.ao_rhs = static_cast<double *>(m.d_phi);
                        p.stride_ao = m.nao;
                        p.stride_grid = m.ngrid;
                        p.nao = m.nao;
                        p.ngrid = q.ngrid;
                        p.nset = q.nset;
                        p.tab = static_cast<uint8_t *>(m.d_tab);
                        p.ldtab = m.nblk;
                        p.blksize = m.blksize;
                        probs.push_back(p);
                    } else {
                        // one evaluation per request: measured faster on the serving mix than one
                        // fused launch per dispatch (work of different slots overlaps instead)
                        st = resident_eval(s, m.tiles, o, sv->rho_ld(q), q.ngrid, q.nset);
                        if (st != CFGMM_OK) break;
                    }
                }
                if (st == CFGMM_OK && !probs.empty()) {
                    st = probs.size() == 1 ? launch_rho(s, const_cast<double *>(probs[0].rhos), probs[0].dms,
                                                        probs[0].ao_lhs, probs[0].ao_rhs, probs[0].stride_ao,
                                                        probs[0].stride_grid, probs[0].nao, probs[0].ngrid,
                                                        probs[0].nset, probs[0].tab, probs[0].ldtab, probs[0].blksize)
                                           : launch_rho_batched(s, probs.data(), probs.size());
                }
                for (size_t i = 0; st == CFGMM_OK && i < batch.size(); ++i) {
                    const cfgmm_request &q = reqs[batch[i]];
                    st = launch_checksum(s, outs[i], q.nset, q.ngrid, sv->rho_ld(q), sv->d_sums + batch[i]);
                }
                sv->stats.rho_requests += batch.size();
            }
            // only this batch's checksums: other slots write theirs concurrently
            for (size_t i = 0; st == CFGMM_OK && i < batch.size(); ++i)
                st = hip_status(hipMemcpyAsync(sv->h_sums + batch[i], sv->d_sums + batch[i], sizeof(unsigned long long),
                                               hipMemcpyDeviceToHost, s),
                                "server: checksum D2H");
            if (st == CFGMM_OK) st = hip_status(hipEventRecord(sl.done, s), "server: event record");
            if (st != CFGMM_OK) {
                status = st;
                release(sl);  // the failed batch's buffers and its budget charge
                break;
            }
            for (size_t r : batch) {
                res[r].admit_s = t_admit;
                res[r].batch = sl.batch_id;
                res[r].batch_size = (int)batch.size();
            }
            sv->stats.batches++;
            sl.busy = true;
            progress = true;
        }
        if (status != CFGMM_OK) break;
        if (!progress && !pending.empty()) {
            bool any_busy = false;
            for (auto &sl : sv->slots) any_busy = any_busy || sl.busy;
            if (!any_busy) {  // nothing will free memory: the head request can never be admitted
                status = fail(CFGMM_EOOM, "server: request " + std::to_string(pending.front()) +
                                              " does not fit the HBM budget");
                break;
            }
        }
        if (!progress) {
            // nothing to do until a completion or the next arrival: sleep only while no slot is
            // busy (a completion is then seen at once) and the next arrival is far off
            const double wait = next < n ? reqs[order[next]].arrival_s - seconds_since(t0) : 0.0;
            bool any_busy = false;
            for (auto &sl : sv->slots) any_busy = any_busy || sl.busy;
            if (!any_busy && wait > 200e-6 && pending.empty())
                std::this_thread::sleep_for(std::chrono::microseconds(100));
            else
                std::this_thread::yield();
        }
    }
    // drain (on error too), then copy back the kept results
    for (auto &sl : sv->slots) (void)hipStreamSynchronize(sl.stream);
    for (auto &sl : sv->slots) {
        if (sl.busy) {
            release(sl);
            sl.busy = false;
        }
    }
    for (auto &kv : sv->keep) {
        const cfgmm_request &q = reqs[kv.first];
        const size_t rows = q.kind == CFGMM_REQ_GEMM ? 1 : q.nset;
        const size_t cols = q.kind == CFGMM_REQ_GEMM ? q.m * q.n : q.ngrid;
        const size_t ld = q.kind == CFGMM_REQ_GEMM ? cols : sv->rho_ld(q);
        if (status == CFGMM_OK)
            status = hip_status(hipMemcpy2D(q.out, sizeof(double) * cols, kv.second, sizeof(double) * ld,
                                            sizeof(double) * cols, rows, hipMemcpyDeviceToHost),
                                "server: result D2H");
        (void)hipFreeAsync(kv.second, sv->slots[0].stream);
    }
    (void)hipStreamSynchronize(sv->slots[0].stream);
    sv->keep.clear();
    for (auto &kv : sv->mols) kv.second.inflight = 0;
    sv->stats.peak_bytes = sv->peak;
    sv->stats.held_bytes = sv->held;
    return status;
}

int server_stats(const Server *sv, cfgmm_server_stats *out) {
    *out = sv->stats;
    out->peak_bytes = sv->peak;
    out->held_bytes = sv->held;
    return CFGMM_OK;
}

}  // namespace cfgmm
