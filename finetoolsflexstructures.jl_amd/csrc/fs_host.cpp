// fs_host.cpp -- C ABI of libflexshell: contexts, models, CSR pattern and
// scatter-plan construction, and the launch sequences of the hot path.
//
// Nothing here computes element or matrix values on the host: the host only
// builds integer structure (node adjacency, CSR pattern, pair tables) and
// small per-material factors; all floating-point work of stiffness, mass,
// normals and the time loop runs in the kernels of fs_kernels.hip.
#include <hip/hip_runtime.h>

#include <sched.h>

#include <algorithm>
#include <array>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstdlib>
#include <functional>
#include <map>
#include <mutex>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <new>
#include <string>
#include <thread>
#include <type_traits>
#include <vector>

#include "../../include/flexshell.h"
#include "fs_device.h"

namespace fs {
hipError_t launch_element_blocks(const ElemParams &p, const double *x, const double *y, const double *z,
                                 const int32_t *conn, const double *nx, const double *ny, const double *nz,
                                 const uint8_t *valid, double *Kb, int *status, hipStream_t st, int variant);
int element_variants(int nen);
void element_variant_shape(int nen, int variant, int *ne, int *wg);
hipError_t launch_scatter(const ScatterPlan &sp, const double *Kb, double *vals, hipStream_t st);
int fused_tile_elems(int nen);
hipError_t launch_element_fused(const ElemParams &p, const double *x, const double *y, const double *z,
                                const int32_t *conn, const double *nx, const double *ny, const double *nz,
                                const uint8_t *valid, int *status, const ScatterPlan &sp, const TilePlan &tp,
                                double *vals, hipStream_t st);
hipError_t launch_pair_finalize(const ScatterPlan &sp, const TilePlan &tp, double *vals, hipStream_t st);
hipError_t launch_normals(int nen, int64_t N, int E, const double *x, const double *y, const double *z,
                          const int32_t *conn, const int32_t *ne_ptr, const int32_t *ne_list, double ntol,
                          const double *gnrm, const int32_t *egroup, double *nx, double *ny, double *nz,
                          uint8_t *valid, hipStream_t st);
hipError_t launch_number_dofs(int64_t n, const uint8_t *fixed, int64_t *tile_free, int64_t *total, int64_t *dof,
                              hipStream_t st);
int64_t number_dofs_tiles(int64_t n);
hipError_t launch_surface_load(int nen, int npts, int64_t N, int E, const double *x, const double *y,
                               const double *z, const int32_t *conn, const int32_t *ne_ptr, const int32_t *ne_list,
                               int per_point, const double *force, const int64_t *node_dof, int64_t nrows, double *F,
                               hipStream_t st);
hipError_t launch_load_points(int nen, int npts, int E, const double *x, const double *y, const double *z,
                              const int32_t *conn, double *out, hipStream_t st);
hipError_t launch_resultants(int nen, int composite, const ResParams &rp, const double *x, const double *y,
                             const double *z, const int32_t *conn, const double *nx, const double *ny,
                             const double *nz, const uint8_t *valid, hipStream_t st);
hipError_t launch_mass(int nen, int64_t N, int E, int composite, const double *g_rho_t, const double *g_inertia,
                       const int32_t *elem_group, const double *x, const double *y, const double *z,
                       const int32_t *conn, const int32_t *ne_ptr, const int32_t *ne_list, const int64_t *node_dof,
                       int64_t nrows, double *diag, hipStream_t st);
hipError_t launch_cd_init(int64_t n, double dt, double g0, const double *invMC, const double *Fmag, const double *U0,
                          const double *V, double *A, double *U1, const int32_t *row_group, const double *group_dt,
                          const double *gtab, hipStream_t st);
hipError_t launch_cd_step(int tpr, const RunRows &rows, const double *kval,
                          const double *Ucur, double *Unext, double *V, double *A, const double *invMC,
                          const double *Cd, const double *Fmag, double gstep, double dt, int damped,
                          const int32_t *probe_slot, double *probe_dst, const int32_t *row_group,
                          const double *group_dt, const double *gtab, hipStream_t st);
hipError_t launch_spmv(const RunRows &rows, const double *kval, const double *x, double *y, hipStream_t st);
hipError_t launch_spmv_node(int nl, int pj, const NodeRows &nrw, const int2 *slotrec, const double *kval,
                            const double *x, double *y, hipStream_t st);
hipError_t launch_cd_node_step(int nl, int pj, const NodeRows &nrw, const int2 *slotrec, const double *kval,
                               const double *Ucur, double *Unext, double *V, double *A, const double *invMC,
                               const double *Cd, const double *Fmag, double gstep, double dt, int damped,
                               const int32_t *probe_slot, double *probe_dst, const int32_t *row_group,
                               const double *group_dt, const double *gtab, hipStream_t st);
hipError_t launch_dot(int64_t n, int mode, const double *x, const double *y, const double *d, double *part,
                      double *out, hipStream_t st);
hipError_t launch_scale(int64_t n, const double *ss, const double *d, const double *x, double *y, hipStream_t st);
hipError_t launch_pwr_init(int64_t n, const double *mdiag, double *inv, double *ss, hipStream_t st);
hipError_t noalloc_workspace(int device, double **out);
hipError_t launch_mixed_build(const ScatterPlan &sp, const double *Kb, const int64_t *goff, float2 *kmb, short2 *kml,
                              hipStream_t st);
hipError_t launch_mixed_drill(int nen, int64_t N, int E, const int32_t *ne_ptr, const int32_t *ne_list,
                              const int32_t *conn, const double *dre, double *dr, hipStream_t st);
hipError_t launch_coo_reduce(int64_t nnz, const int64_t *cptr, const int64_t *clist, const double *mats,
                             double *out, hipStream_t st);
hipError_t launch_cd_mixed_step(int nen, int nsr, const MixedOp &op, const double *Ucur, double *Unext,
                                double *V, double *A, const double *invMC, const double *Cd, const double *Fmag,
                                double gstep, double dt, int damped, int forced, const int32_t *probe_slot,
                                double *probe_dst, const int32_t *row_group, const double *group_dt,
                                const double *gtab, hipStream_t st);
}  // namespace fs

namespace {
// Device buffers released by the models of a context are kept for reuse by
// the same context instead of being returned with hipFree, which waits for
// the whole device (every other context's work included).  Reuse is safe
// because every ABI call ends with its stream synchronised, so a released
// buffer has no pending work.  Shared by the context and every live buffer
// allocated through it; outlives the context until the last buffer is gone.
struct BufCache {
    std::mutex mu;
    std::multimap<size_t, void *> spare;  // bytes -> block
    size_t spare_bytes = 0;
    size_t cap = size_t(4) << 30;         // bytes kept per context at most
    int refs = 1;                         // the context + live buffers
    bool alive = true;                    // false once the context is destroyed
};
}  // namespace

struct fs_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    // the caller's handle plus one per live model: fs_ctx_destroy drops the
    // caller's reference, the last release tears the context down, so a
    // model may outlive fs_ctx_destroy (it may then only be destroyed)
    std::atomic<int> refs{1};
    std::string err;
    BufCache *cache = nullptr;
    // element-kernel launch shape per (family T3/Q4, shear formulation):
    // -1 = not tuned yet (the autotuner times every shape on the first
    // stiffness evaluation of that kind and keeps the fastest)
    int elem_variant[2][2] = {{-1, -1}, {-1, -1}};
    bool autotune = true;
    // assembly pipeline: the element kernels run in asm_chunks element
    // ranges on `stream`; the scatter of the node pairs completed by each
    // range runs on `stream2` (default priority) while the next range computes
    hipStream_t stream2 = nullptr;
    int asm_chunks = 1;
    std::vector<hipEvent_t> ev;  // ordering events of the pipeline (no timing)
    // pinned staging buffers of the large device-to-host copies (allocated
    // on first use, FS_STAGE_BYTES each) and their completion events
    void *stage[2] = {nullptr, nullptr};
    hipEvent_t stage_ev[2] = {nullptr, nullptr};
    // row products of the FP64 explicit step: 0 = lanes per row
    // (cd_step_kernel), 8/16 = lanes per node (cd_node_step_kernel) with
    // cd_pj positions per lane (0: from the block's longest row)
    int cd_lanes = 8;
    int cd_pj = 0;
    // the device's workspace of the allocation-free entry points (address of
    // its copy of the file-scope array, taken at context creation), and the
    // status of that lookup
    double *scratch = nullptr;
    hipError_t scratch_status = hipSuccess;
    // stiffness assembly: 1 = fused element tiles (element kernel writes the
    // CSR values of the pairs it completes, boundary pairs finalised after);
    // 0 = element blocks to HBM, then the pair scatter
    int assembly = 1;
};

namespace {

struct DevBuf {
    void *p = nullptr;
    size_t bytes = 0;          // size of the block (may exceed the request)
    BufCache *cache = nullptr; // owner's cache, or nullptr
    DevBuf() = default;
    DevBuf(const DevBuf &) = delete;
    DevBuf &operator=(const DevBuf &) = delete;
    ~DevBuf() { release(); }
    void release() {
        if (p) {
            if (cache) {
                int left;
                {
                    std::lock_guard<std::mutex> lk(cache->mu);
                    if (cache->alive && cache->spare_bytes + bytes <= cache->cap) {
                        cache->spare.insert({bytes, p});
                        cache->spare_bytes += bytes;
                    } else {
                        (void)hipFree(p);
                    }
                    left = --cache->refs;
                }
                if (left == 0) delete cache;
            } else {
                (void)hipFree(p);
            }
        }
        p = nullptr;
        bytes = 0;
        cache = nullptr;
    }
    template <class T>
    T *as() const { return static_cast<T *>(p); }
};

// Fused-assembly plan of a block (fs::TilePlan): the element tiles, their
// pair entries and the boundary pairs; built on the first fused stiffness.
struct TileStore {
    int ne = 0;  // elements per tile it was built for (0: not built)
    int64_t ntiles = 0, nent = 0, nbnd = 0, nslots = 0;
    DevBuf tile_elem, tile_conn, tp_ptr, trec, tcontrib, tslot, brec, bp_ptr, partial;
    void reset() {
        ne = 0;
        ntiles = nent = nbnd = nslots = 0;
        for (DevBuf *b : {&tile_elem, &tile_conn, &tp_ptr, &trec, &tcontrib, &tslot, &brec, &bp_ptr, &partial})
            b->release();
    }
};

struct Block {
    bool built = false;
    int64_t nrows = 0;
    int64_t nnz = 0;
    int64_t nnodes = 0;
    std::vector<int64_t> rowptr;   // nrows + 1
    // column lists, kept as per-node structure and expanded on request
    // (fs_model_pattern_get): code[J*6+c] offset of comp c within J's free
    // (0..5) or fixed (8..13) run, -1 not in the block; off_free[k] /
    // off_fix[k] offset of neighbour slot k's free / fixed run in the row
    std::vector<int8_t> code;
    std::vector<int32_t> off_free, off_fix;
    // scatter plan (device)
    int64_t npairs = 0;
    // running maximum over the pairs, in pair order, of the last element
    // contributing to a pair: the pairs [0, k) with pair_emax[k-1] < e are
    // complete once the elements [0, e) are (host)
    std::vector<int32_t> pair_emax;
    DevBuf nbr_ptr, nbr_node, nbr_pair, nbr_off, node_code, node_rowstart, contrib_ptr, contrib, pair_node,
        pair_slot;
    // device values (CSR order) and, for the free-free block, the run records
    // of the row products of the time loop (fs::RunRows)
    DevBuf rowrec, slotrec, kval, mdiag;
    // free-free block: node records of the node-grouped step (fs::NodeRows)
    DevBuf noderec, nodeinfo;
    int64_t nrownodes = 0;
    int32_t max_rowlen = 0;
    bool k_ready = false, m_ready = false;
    // host copies for the fused-assembly planner: neighbour slot -> (pair <<
    // 1 | side), pair nodes and slots, contributions per pair (counts)
    std::vector<int32_t> h_nb_pair, h_pnode, h_pslot, h_cptr;
    // per node {first free row's CSR position, first fixed row's (-1), row
    // length, 4-bit comp codes}; per neighbour slot the free / fixed run offsets
    std::vector<int32_t> h_node_rec, h_nb_off;
    TileStore tiles;
    void reset() {
        built = k_ready = m_ready = false;
        h_nb_pair.clear();
        h_pnode.clear();
        h_pslot.clear();
        h_cptr.clear();
        h_node_rec.clear();
        h_nb_off.clear();
        tiles.reset();
        nrows = nnz = npairs = nrownodes = 0;
        max_rowlen = 0;
        rowptr.clear();
        code.clear();
        off_free.clear();
        off_fix.clear();
        pair_emax.clear();
        for (DevBuf *b : {&nbr_ptr, &nbr_node, &nbr_pair, &nbr_off, &node_code, &node_rowstart, &contrib_ptr,
                          &contrib, &pair_node, &pair_slot, &rowrec, &slotrec, &kval, &mdiag, &noderec, &nodeinfo})
            b->release();
    }
};

}  // namespace

// Mixed-precision operator resident for the explicit loop (see fs::MixedOp).
struct MixedStore {
    bool built = false;
    uint64_t kgen = 0;   // element-parameter generation it was built from
    bool shear_fp32 = false;  // transverse shear in the FP32 slices (no factor rows)
    int nsr = 0;
    int64_t nslices = 0;
    int64_t kmb_pairs = 0;  // float2 values
    DevBuf goff, kmb, kml, cs, dr, ne_pack;
    float build_ms = 0.f;
    void reset() {
        built = false;
        for (DevBuf *b : {&goff, &kmb, &kml, &cs, &dr, &ne_pack}) b->release();
    }
};

// Device state of the explicit loop, kept with its step graph between
// calls: a call whose launch parameters (sizes, scalars, load factors,
// buffer addresses and the operator's records) equal those of the graph
// uploads its initial state into the same buffers and replays the graph,
// instead of capturing and instantiating its nsteps launches again.
struct StepCache {
    DevBuf dU[2], dV, dA, dF, dInv, dC, dSlot, dProbe, dRG, dGdt, dG;
    std::vector<uint64_t> key;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    int64_t replays = 0;  // launches of a graph taken from the cache
    void drop_graph() {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
        exec = nullptr;
        graph = nullptr;
        key.clear();
    }
    ~StepCache() { drop_graph(); }
};

struct fs_model {
    fs_ctx *ctx = nullptr;
    int nen = 0;
    int64_t N = 0, E = 0;
    std::vector<int32_t> conn;       // E x nen row-major, 0-based (host)
    std::vector<int32_t> ne_ptr, ne_list;  // node -> elements (ascending)
    std::vector<int32_t> nbr_ptr, nbr_list;  // node -> neighbour nodes (sorted)
    DevBuf x, y, z, dconn, dne_ptr, dne_list, nx, ny, nz, valid, node_dof, Kb, status;
    bool normals_ready = false;
    bool dofs_ready = false;
    std::vector<int64_t> dof;  // N x 6 row-major, 0-based
    int64_t nall = 0, nfree = 0;
    Block blk[2];
    // group data for the last stiffness/mass call
    DevBuf g_Labd, g_Lh, g_csys, g_t, g_group, g_inertia, m_group, g_homog, g_rho_t;
    double timing[7] = {0, 0, 0, 0, 0, 0, 0};
    // parameters of the last element-stiffness launch (device pointers into
    // the group buffers above) and their generation
    fs::ElemParams last_ep{};
    uint64_t kgen = 0;
    MixedStore mixed;
    StepCache step;
};

// Buffered element matrices of a FEMM with no element kernel here, reduced
// into CSR on the device (fs_coo_*): the pattern and, per stored entry, the
// positions of its contributions in the element-matrix buffer.
struct fs_coo {
    fs_ctx *ctx = nullptr;
    int64_t nrows = 0, ncols = 0, nmat = 0;
    int32_t kr = 0, kc = 0;
    int64_t nnz = 0;
    std::vector<int64_t> rowptr, colval;  // 0-based
    DevBuf cptr, clist, mats, vals;
};

namespace {

// Host threads for the parallel loops: the CPUs this process may run on,
// capped by OMP_NUM_THREADS when that is set.
int64_t host_threads() {
    static const int64_t nt = [] {
        int64_t n = std::max<int64_t>(1, std::thread::hardware_concurrency());
        cpu_set_t set;
        if (sched_getaffinity(0, sizeof set, &set) == 0) n = std::max(1, CPU_COUNT(&set));
        if (const char *e = std::getenv("OMP_NUM_THREADS")) {
            const long v = std::strtol(e, nullptr, 10);
            if (v > 0) n = std::min<int64_t>(n, v);
        }
        return n;
    }();
    return nt;
}

// Runs fn(begin, end) over [0, n) in up to host_threads() chunks of at
// least `grain` iterations on threads of their own; for host loops whose
// iterations write disjoint data.
template <class F>
void parallel_chunks(int64_t n, F fn, int64_t grain = 65536) {
    const int64_t nt = std::max<int64_t>(1, std::min<int64_t>(host_threads(), n / grain));
    if (nt <= 1) {
        fn((int64_t)0, n);
        return;
    }
    std::vector<std::thread> th;
    th.reserve(nt);
    for (int64_t t = 0; t < nt; ++t) th.emplace_back(fn, n * t / nt, n * (t + 1) / nt);
    for (auto &x : th) x.join();
}

// Records the message of a failing call and waits for the work the call
// had already queued on the context's streams, so that the device buffers
// the call releases on its way out (returned to the context's spare list)
// have no pending kernels or copies.  Every ABI call therefore ends with its
// streams synchronised, on the error paths too.
int fail(fs_ctx *c, int code, const char *fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    if (c) {
        c->err = buf;
        if (c->stream) (void)hipStreamSynchronize(c->stream);
        if (c->stream2) (void)hipStreamSynchronize(c->stream2);
    }
    // the error is reported: clear the runtime's last-error state, which the
    // launch checks (hipGetLastError) of later calls would otherwise read
    (void)hipGetLastError();
    return code;
}

#define FS_HIP(ctx, call)                                                                           \
    do {                                                                                            \
        hipError_t e_ = (call);                                                                     \
        if (e_ != hipSuccess) return fail((ctx), FS_ERR_DEVICE, "%s: %s", #call, hipGetErrorString(e_)); \
    } while (0)

int dalloc(fs_ctx *c, DevBuf &b, size_t bytes) {
    b.release();
    if (bytes == 0) bytes = 8;
    BufCache *cache = c->cache;
    {
        // best fit among the spare blocks, at most 25 % + 1 MiB larger
        std::lock_guard<std::mutex> lk(cache->mu);
        auto it = cache->spare.lower_bound(bytes);
        if (it != cache->spare.end() && it->first <= bytes + bytes / 4 + (size_t(1) << 20)) {
            b.p = it->second;
            b.bytes = it->first;
            cache->spare_bytes -= it->first;
            cache->spare.erase(it);
            b.cache = cache;
            ++cache->refs;
            return FS_OK;
        }
    }
    hipError_t e = hipMalloc(&b.p, bytes);
    if (e == hipErrorOutOfMemory) {
        // give the spare blocks back and try once more
        {
            std::lock_guard<std::mutex> lk(cache->mu);
            for (auto &kv : cache->spare) (void)hipFree(kv.second);
            cache->spare.clear();
            cache->spare_bytes = 0;
        }
        e = hipMalloc(&b.p, bytes);
    }
    if (e != hipSuccess) {
        b.p = nullptr;
        return fail(c, FS_ERR_NOMEM, "hipMalloc(%zu bytes): %s", bytes, hipGetErrorString(e));
    }
    b.bytes = bytes;
    std::lock_guard<std::mutex> lk(cache->mu);
    b.cache = cache;
    ++cache->refs;
    return FS_OK;
}

template <class T>
int upload(fs_ctx *c, DevBuf &b, const T *src, size_t n) {
    int s = dalloc(c, b, n * sizeof(T));
    if (s) return s;
    if (n) FS_HIP(c, hipMemcpyAsync(b.p, src, n * sizeof(T), hipMemcpyHostToDevice, c->stream));
    return FS_OK;
}

// dalloc that keeps the buffer when it is already large enough (its
// address then stays the same from call to call).
int ensure(fs_ctx *c, DevBuf &b, size_t bytes) {
    if (b.p && b.bytes >= std::max<size_t>(bytes, 8)) return FS_OK;
    return dalloc(c, b, bytes);
}

template <class T>
int put(fs_ctx *c, DevBuf &b, const T *src, size_t n) {
    int s = ensure(c, b, n * sizeof(T));
    if (s) return s;
    if (n) FS_HIP(c, hipMemcpyAsync(b.p, src, n * sizeof(T), hipMemcpyHostToDevice, c->stream));
    return FS_OK;
}

// Appends the bytes of a trivially copyable value to a cache key.
template <class T>
void key_add(std::vector<uint64_t> &k, const T &v) {
    static_assert(std::is_trivially_copyable<T>::value, "key values are plain data");
    uint64_t w[(sizeof(T) + 7) / 8] = {};
    std::memcpy(w, &v, sizeof(T));
    k.insert(k.end(), w, w + (sizeof(T) + 7) / 8);
}

constexpr size_t FS_STAGE_BYTES = size_t(32) << 20;

// Device-to-host copy ordered on the context's own stream, so that contexts
// sharing a device never serialise on the null stream.  Copies of more than
// two staging buffers go through the context's pinned buffers in chunks: the
// DMA of chunk i+1 runs while host threads copy chunk i into `dst` (which
// also spreads the first-touch page faults of a fresh destination over the
// threads).
int d2h(fs_ctx *c, void *dst, const void *src, size_t bytes) {
    if (bytes == 0) return FS_OK;
    if (bytes <= 2 * FS_STAGE_BYTES) {
        FS_HIP(c, hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, c->stream));
        FS_HIP(c, hipStreamSynchronize(c->stream));
        return FS_OK;
    }
    for (int i = 0; i < 2; ++i) {
        if (!c->stage[i]) FS_HIP(c, hipHostMalloc(&c->stage[i], FS_STAGE_BYTES, hipHostMallocDefault));
        if (!c->stage_ev[i]) FS_HIP(c, hipEventCreateWithFlags(&c->stage_ev[i], hipEventDisableTiming));
    }
    const size_t nch = (bytes + FS_STAGE_BYTES - 1) / FS_STAGE_BYTES;
    auto chunk = [&](size_t i) { return std::min(FS_STAGE_BYTES, bytes - i * FS_STAGE_BYTES); };
    auto enqueue = [&](size_t i) -> hipError_t {
        hipError_t e = hipMemcpyAsync(c->stage[i % 2], (const char *)src + i * FS_STAGE_BYTES, chunk(i),
                                      hipMemcpyDeviceToHost, c->stream);
        return e != hipSuccess ? e : hipEventRecord(c->stage_ev[i % 2], c->stream);
    };
    FS_HIP(c, enqueue(0));
    for (size_t i = 0; i < nch; ++i) {
        // buffer (i+1) % 2 was last read by the host copy of chunk i-1, which has returned
        if (i + 1 < nch) FS_HIP(c, enqueue(i + 1));
        FS_HIP(c, hipEventSynchronize(c->stage_ev[i % 2]));
        const char *from = (const char *)c->stage[i % 2];
        char *to = (char *)dst + i * FS_STAGE_BYTES;
        parallel_chunks((int64_t)chunk(i), [&](int64_t b0, int64_t b1) { std::memcpy(to + b0, from + b0, b1 - b0); },
                        int64_t(1) << 20);
    }
    return FS_OK;
}

// Upper Cholesky factor of a symmetric n x n matrix (row-major): M = L' L.
bool chol_upper(int n, const double *M, double *L) {
    for (int i = 0; i < n * n; ++i) L[i] = 0.0;
    for (int i = 0; i < n; ++i) {
        double d = M[i * n + i];
        for (int k = 0; k < i; ++k) d -= L[k * n + i] * L[k * n + i];
        if (!(d > 0.0)) return false;
        const double lii = std::sqrt(d);
        L[i * n + i] = lii;
        for (int j = i + 1; j < n; ++j) {
            double s = M[i * n + j];
            for (int k = 0; k < i; ++k) s -= L[k * n + i] * L[k * n + j];
            L[i * n + j] = s / lii;
        }
    }
    return true;
}

// column-major n x n -> row-major
void cm2rm(int n, const double *cm, double *rm) {
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j) rm[i * n + j] = cm[j * n + i];
}

int build_block(fs_model *m, int which) {
    fs_ctx *c = m->ctx;
    Block &B = m->blk[which];
    if (B.built) return FS_OK;
    if (!m->dofs_ready) return fail(c, FS_ERR_STATE, "pattern: equation numbers not set (fs_model_set_dofs)");
    const int64_t N = m->N;
    const int64_t nrows = which == FS_BLOCK_FF ? m->nfree : m->nall;
    const int64_t nfree = m->nfree;
    // per-node codes: offset of each comp within the node's free / fixed run
    std::vector<int8_t> &code = B.code;
    code.assign(N * 6, -1);
    std::vector<int8_t> nfr(N, 0), nfx(N, 0);
    std::vector<int64_t> minfr(N, INT64_MAX), minfx(N, INT64_MAX);
    // lowest node (2 I + 0: free, 2 I + 1: fixed) whose dofs are not consecutive
    std::atomic<int64_t> bad{INT64_MAX};
    auto flag = [&](int64_t v) {
        int64_t cur = bad.load();
        while (v < cur && !bad.compare_exchange_weak(cur, v)) {
        }
    };
    parallel_chunks(N, [&](int64_t I0, int64_t I1) {
        for (int64_t I = I0; I < I1; ++I) {
            int64_t fr[6], fx[6];
            int a = 0, b = 0;
            for (int cc = 0; cc < 6; ++cc) {
                const int64_t d = m->dof[I * 6 + cc];
                if (d >= nrows) continue;
                if (d < nfree)
                    fr[a++] = d;
                else
                    fx[b++] = d;
            }
            std::sort(fr, fr + a);
            std::sort(fx, fx + b);
            if (a && fr[a - 1] - fr[0] + 1 != a) {
                flag(2 * I);
                continue;
            }
            if (b && fx[b - 1] - fx[0] + 1 != b) {
                flag(2 * I + 1);
                continue;
            }
            for (int cc = 0; cc < 6; ++cc) {
                const int64_t d = m->dof[I * 6 + cc];
                if (d >= nrows) continue;
                if (d < nfree)
                    code[I * 6 + cc] = (int8_t)(d - fr[0]);
                else
                    code[I * 6 + cc] = (int8_t)(8 + (d - fx[0]));
            }
            nfr[I] = (int8_t)a;
            nfx[I] = (int8_t)b;
            if (a) minfr[I] = fr[0];
            if (b) minfx[I] = fx[0];
        }
    });
    if (bad.load() != INT64_MAX)
        return fail(c, FS_ERR_UNSUPPORTED, "node %lld: %s dofs are not consecutive numbers",
                    (long long)(bad.load() / 2), bad.load() % 2 ? "fixed" : "free");
    // per node: sorted column list = runs of the neighbours ordered by first dof
    // (runs are contiguous integer ranges, so sorting runs sorts the columns)
    const size_t nnb = m->nbr_list.size();
    std::vector<int32_t> rowlen(N, 0);
    std::vector<int32_t> &nb_offfree = B.off_free, &nb_offfix = B.off_fix;
    nb_offfree.assign(nnb, -1);
    nb_offfix.assign(nnb, -1);
    parallel_chunks(N, [&](int64_t I0, int64_t I1) {
        std::vector<std::pair<int64_t, int64_t>> runs;  // (first dof, nbr slot*2 + kind)
        for (int64_t I = I0; I < I1; ++I) {
            runs.clear();
            for (int k = m->nbr_ptr[I]; k < m->nbr_ptr[I + 1]; ++k) {
                const int J = m->nbr_list[k];
                if (nfr[J]) runs.push_back({minfr[J], (int64_t)k * 2 + 0});
                if (nfx[J]) runs.push_back({minfx[J], (int64_t)k * 2 + 1});
            }
            std::sort(runs.begin(), runs.end());
            int32_t pos = 0;
            for (auto &r : runs) {
                const int64_t k = r.second / 2;
                const int J = m->nbr_list[k];
                if (r.second % 2 == 0) {
                    nb_offfree[k] = pos;
                    pos += nfr[J];
                } else {
                    nb_offfix[k] = pos;
                    pos += nfx[J];
                }
            }
            rowlen[I] = pos;
        }
    });
    // row owners
    std::vector<int64_t> owner(nrows, -1);
    for (int64_t I = 0; I < N; ++I)
        for (int cc = 0; cc < 6; ++cc) {
            const int64_t d = m->dof[I * 6 + cc];
            if (d < nrows) owner[d] = I;
        }
    B.rowptr.assign(nrows + 1, 0);
    for (int64_t d = 0; d < nrows; ++d) {
        if (owner[d] < 0) return fail(c, FS_ERR_ARG, "dof %lld is not used by any node", (long long)d);
        B.rowptr[d + 1] = B.rowptr[d] + rowlen[owner[d]];
    }
    const int64_t nnz = B.rowptr[nrows];
    if (nnz >= (int64_t(1) << 31))
        return fail(c, FS_ERR_UNSUPPORTED, "block with %lld stored entries exceeds 32-bit positions", (long long)nnz);
    // upper node pairs: the neighbours J >= I of each node, in order
    std::vector<int64_t> pair_base(N + 1, 0);
    parallel_chunks(N, [&](int64_t I0, int64_t I1) {
        for (int64_t I = I0; I < I1; ++I) {
            int cnt = 0;
            for (int k = m->nbr_ptr[I]; k < m->nbr_ptr[I + 1]; ++k)
                if (m->nbr_list[k] >= I) ++cnt;
            pair_base[I + 1] = cnt;
        }
    });
    for (int64_t I = 0; I < N; ++I) pair_base[I + 1] += pair_base[I];
    const int64_t npairs = pair_base[N];
    if (npairs >= (int64_t(1) << 27)) return fail(c, FS_ERR_UNSUPPORTED, "too many node pairs (%lld)", (long long)npairs);
    auto slot_of = [&](int64_t I, int64_t J) -> int {
        const int32_t *b0 = m->nbr_list.data() + m->nbr_ptr[I];
        const int32_t *e0 = m->nbr_list.data() + m->nbr_ptr[I + 1];
        return (int)(std::lower_bound(b0, e0, (int32_t)J) - m->nbr_list.data());
    };
    // slot of each node in its own list: the upper pairs of I are its slots
    // from diag[I] on, numbered from pair_base[I]
    std::vector<int32_t> diag(N);
    parallel_chunks(N, [&](int64_t I0, int64_t I1) {
        for (int64_t I = I0; I < I1; ++I) diag[I] = slot_of(I, I);
    });
    // neighbour tables: pair, side, run offsets; node row starts (each node
    // writes its own slots, its own upper pairs and its own rows)
    std::vector<int32_t> nb_pair(nnb), nb_off(2 * nnb), rowstart(N * 6, -1);
    std::vector<int32_t> pnode(2 * npairs), pslot(2 * npairs);
    parallel_chunks(N, [&](int64_t I0, int64_t I1) {
        for (int64_t I = I0; I < I1; ++I) {
            for (int k = m->nbr_ptr[I]; k < m->nbr_ptr[I + 1]; ++k) {
                const int J = m->nbr_list[k];
                const int64_t p = I <= J ? pair_base[I] + (k - diag[I]) : pair_base[J] + (slot_of(J, I) - diag[J]);
                nb_pair[k] = (int32_t)((p << 1) | (I > J ? 1 : 0));
                nb_off[2 * k] = nb_offfree[k];
                nb_off[2 * k + 1] = nb_offfix[k];
                if (I <= J) {
                    pnode[2 * p] = (int32_t)I;
                    pnode[2 * p + 1] = J;
                    pslot[2 * p] = k;
                    pslot[2 * p + 1] = slot_of(J, I);
                }
            }
            for (int cc = 0; cc < 6; ++cc) {
                const int64_t d = m->dof[I * 6 + cc];
                if (d < nrows) rowstart[I * 6 + cc] = (int32_t)B.rowptr[d];
            }
        }
    });
    // contributions, element order within each pair
    const int nen = m->nen;
    const int nblk = nen * (nen + 1) / 2;
    if ((int64_t)m->E * nblk >= (int64_t(1) << 30))
        return fail(c, FS_ERR_UNSUPPORTED, "too many element blocks (%lld)", (long long)(m->E * nblk));
    // The contributions of the upper pair (I, J) are the elements incident to
    // both nodes (an intersection of the two ascending incidence lists, so in
    // element order), each with the local block (a <= b) that holds the pair
    // and whether the block's row node is J (transposed).  Two passes over
    // the pairs: counts, then the lists.
    auto pair_contribs = [&](int64_t p, int32_t *outl) -> int {
        const int I = pnode[2 * p], J = pnode[2 * p + 1];
        int ki = m->ne_ptr[I], kj = m->ne_ptr[J], cnt = 0;
        const int ei = m->ne_ptr[I + 1], ej = m->ne_ptr[J + 1];
        while (ki < ei && kj < ej) {
            const int32_t e1 = m->ne_list[ki], e2 = m->ne_list[kj];
            if (e1 < e2) {
                ++ki;
            } else if (e2 < e1) {
                ++kj;
            } else {
                if (outl) {
                    const int32_t *ce = m->conn.data() + (int64_t)e1 * nen;
                    int a = 0, b = 0;
                    for (int q = 0; q < nen; ++q) {
                        if (ce[q] == I) a = q;
                        if (ce[q] == J) b = q;
                    }
                    const int lo = std::min(a, b), hi = std::max(a, b);
                    const int bl = lo * nen - lo * (lo - 1) / 2 + (hi - lo);  // upper pairs in row order
                    const int tr = ce[lo] > ce[hi] ? 1 : 0;
                    outl[cnt] = (int32_t)((((int64_t)e1 * nblk + bl) << 1) | tr);
                }
                ++cnt;
                ++ki;
                ++kj;
            }
        }
        return cnt;
    };
    std::vector<int32_t> cptr(npairs + 1, 0);
    parallel_chunks(npairs, [&](int64_t p0, int64_t p1) {
        for (int64_t p = p0; p < p1; ++p) cptr[p + 1] = pair_contribs(p, nullptr);
    });
    for (int64_t p = 0; p < npairs; ++p) cptr[p + 1] += cptr[p];
    std::vector<int32_t> clist(cptr[npairs]);
    B.pair_emax.assign(npairs, -1);
    parallel_chunks(npairs, [&](int64_t p0, int64_t p1) {
        for (int64_t p = p0; p < p1; ++p) {
            const int n = pair_contribs(p, clist.data() + cptr[p]);
            if (n) B.pair_emax[p] = (clist[cptr[p] + n - 1] >> 1) / nblk;  // last (largest) element
        }
    });
    for (int64_t p = 1; p < npairs; ++p) B.pair_emax[p] = std::max(B.pair_emax[p], B.pair_emax[p - 1]);
    // run records of the free-free block (fs::RunRows): every row of node I
    // is the runs of I's neighbours in slot order, each run the neighbour's
    // free dofs (consecutive columns starting at its first free dof)
    std::vector<int32_t> rrec, srec, nrec, ninfo;
    if (which == FS_BLOCK_FF) {
        // node records: the nodes with free dofs, in node order
        for (int64_t I = 0; I < N; ++I)
            if (nfr[I]) {
                nrec.push_back((int32_t)B.rowptr[minfr[I]]);
                nrec.push_back((int32_t)minfr[I]);
                nrec.push_back(m->nbr_ptr[I]);
                nrec.push_back(m->nbr_ptr[I + 1]);
                ninfo.push_back((rowlen[I] << 3) | nfr[I]);
                B.max_rowlen = std::max(B.max_rowlen, rowlen[I]);
            }
        rrec.assign(4 * nrows, 0);
        parallel_chunks(nrows, [&](int64_t d0, int64_t d1) {
            for (int64_t d = d0; d < d1; ++d) {
                const int64_t I = owner[d];
                rrec[4 * d + 0] = (int32_t)B.rowptr[d];
                rrec[4 * d + 1] = m->nbr_ptr[I];
                rrec[4 * d + 2] = m->nbr_ptr[I + 1];
            }
        });
        for (int64_t I = 0; I < N; ++I)
            if (rowlen[I] >= (1 << 27)) return fail(c, FS_ERR_UNSUPPORTED, "row of more than 2^27 entries");
        srec.assign(2 * nnb, 0);
        parallel_chunks((int64_t)nnb, [&](int64_t k0, int64_t k1) {
            for (int64_t k = k0; k < k1; ++k) {
                const int J = m->nbr_list[k];
                const int nf = nfr[J];
                if (nf == 0) continue;  // no free dofs: an empty run
                srec[2 * k] = (nb_offfree[k] << 4) | nf;
                srec[2 * k + 1] = (int32_t)minfr[J];
            }
        });
    }
    // upload
    int s;
    if ((s = upload(c, B.nbr_ptr, m->nbr_ptr.data(), m->nbr_ptr.size()))) return s;
    if ((s = upload(c, B.nbr_node, m->nbr_list.data(), m->nbr_list.size()))) return s;
    if ((s = upload(c, B.nbr_pair, nb_pair.data(), nb_pair.size()))) return s;
    if ((s = upload(c, B.nbr_off, nb_off.data(), nb_off.size()))) return s;
    if ((s = upload(c, B.node_code, code.data(), code.size()))) return s;
    if ((s = upload(c, B.node_rowstart, rowstart.data(), rowstart.size()))) return s;
    if ((s = upload(c, B.contrib_ptr, cptr.data(), cptr.size()))) return s;
    if ((s = upload(c, B.contrib, clist.data(), clist.size()))) return s;
    if ((s = upload(c, B.pair_node, pnode.data(), pnode.size()))) return s;
    if ((s = upload(c, B.pair_slot, pslot.data(), pslot.size()))) return s;
    if (which == FS_BLOCK_FF) {
        if ((s = upload(c, B.rowrec, rrec.data(), rrec.size()))) return s;
        if ((s = upload(c, B.slotrec, srec.data(), srec.size()))) return s;
        if ((s = upload(c, B.noderec, nrec.data(), nrec.size()))) return s;
        if ((s = upload(c, B.nodeinfo, ninfo.data(), ninfo.size()))) return s;
        B.nrownodes = (int64_t)ninfo.size();
    }
    if ((s = dalloc(c, B.kval, (size_t)nnz * sizeof(double)))) return s;
    if ((s = dalloc(c, B.mdiag, (size_t)nrows * sizeof(double)))) return s;
    FS_HIP(c, hipStreamSynchronize(c->stream));
    B.h_node_rec.assign(4 * N, -1);
    for (int64_t I = 0; I < N; ++I) {
        int32_t codes = 0;
        for (int cc = 0; cc < 6; ++cc) codes |= (code[I * 6 + cc] < 0 ? 15 : (int)code[I * 6 + cc]) << (4 * cc);
        B.h_node_rec[4 * I] = nfr[I] ? (int32_t)B.rowptr[minfr[I]] : -1;
        B.h_node_rec[4 * I + 1] = nfx[I] ? (int32_t)B.rowptr[minfx[I]] : -1;
        B.h_node_rec[4 * I + 2] = rowlen[I];
        B.h_node_rec[4 * I + 3] = codes;
    }
    B.h_nb_off = std::move(nb_off);
    B.h_nb_pair = std::move(nb_pair);
    B.h_pnode = std::move(pnode);
    B.h_pslot = std::move(pslot);
    B.h_cptr = std::move(cptr);
    B.npairs = npairs;
    B.nnodes = N;
    B.nrows = nrows;
    B.nnz = nnz;
    B.built = true;
    B.k_ready = B.m_ready = false;
    return FS_OK;
}

fs::RunRows runrows_of(const Block &B) {
    fs::RunRows r;
    r.n = B.nrows;
    r.rowrec = B.rowrec.as<int4>();
    r.slotrec = B.slotrec.as<int2>();
    return r;
}

fs::NodeRows noderows_of(const Block &B) {
    fs::NodeRows r;
    r.n = B.nrownodes;
    r.rec = B.noderec.as<int4>();
    r.info = B.nodeinfo.as<int32_t>();
    return r;
}

// Positions per lane of the node-grouped row products: the context's setting,
// or enough for one pass over the longest row of the block where the
// register budget allows (8 lanes: 4..8, 16 lanes: 2..4).
int node_pj(const fs_ctx *c, const Block &B) {
    if (c->cd_lanes == 0 || c->cd_pj != 0) return c->cd_pj;
    const int lo = c->cd_lanes == 8 ? 4 : 2, hi = c->cd_lanes == 8 ? 8 : 4;
    return std::min(hi, std::max(lo, (B.max_rowlen + c->cd_lanes - 1) / c->cd_lanes));
}

fs::ScatterPlan plan_of(const Block &B) {
    fs::ScatterPlan sp;
    sp.nnodes = B.nnodes;
    sp.nbr_ptr = B.nbr_ptr.as<int32_t>();
    sp.nbr_node = B.nbr_node.as<int32_t>();
    sp.nbr_pair = B.nbr_pair.as<int32_t>();
    sp.nbr_off = B.nbr_off.as<int32_t>();
    sp.node_code = B.node_code.as<int8_t>();
    sp.node_rowstart = B.node_rowstart.as<int32_t>();
    sp.contrib_ptr = B.contrib_ptr.as<int32_t>();
    sp.contrib = B.contrib.as<int32_t>();
    sp.npairs = B.npairs;
    sp.pbegin = 0;
    sp.pend = B.npairs;
    sp.pair_node = B.pair_node.as<int32_t>();
    sp.pair_slot = B.pair_slot.as<int32_t>();
    return sp;
}

fs::TilePlan tileplan_of(const Block &B) {
    fs::TilePlan tp;
    tp.ntiles = B.tiles.ntiles;
    tp.nbnd = B.tiles.nbnd;
    tp.tp_ptr = B.tiles.tp_ptr.as<int32_t>();
    tp.trec = B.tiles.trec.as<int4>();
    tp.tcontrib = B.tiles.tcontrib.as<uint64_t>();
    tp.tslot = B.tiles.tslot.as<int32_t>();
    tp.brec = B.tiles.brec.as<int4>();
    tp.bp_ptr = B.tiles.bp_ptr.as<int32_t>();
    tp.partial = B.tiles.partial.as<double>();
    return tp;
}

// 3 x 21-bit Morton code of a point scaled into [0, 1)^3.
uint64_t morton3(double x, double y, double z) {
    auto spread = [](uint64_t v) {
        v &= 0x1fffff;
        v = (v | v << 32) & 0x1f00000000ffffULL;
        v = (v | v << 16) & 0x1f0000ff0000ffULL;
        v = (v | v << 8) & 0x100f00f00f00f00fULL;
        v = (v | v << 4) & 0x10c30c30c30c30c3ULL;
        v = (v | v << 2) & 0x1249249249249249ULL;
        return v;
    };
    auto q = [](double t) {
        const double u = t * 2097152.0;
        return (uint64_t)(u < 0.0 ? 0.0 : (u > 2097151.0 ? 2097151.0 : u));
    };
    return spread(q(x)) | spread(q(y)) << 1 | spread(q(z)) << 2;
}

// Fused-assembly plan of a block.  Elements are grouped into tiles of NE:
// per connected component (components in the order of their first
// element), elements sorted along a Morton curve of their centroids scaled
// by the component's bounding box, then cut into consecutive runs of NE, so
// a tile is a compact patch and never spans two components (a model gives
// the same tiles alone and inside a batch).  A tile's elements are kept in
// ascending order.  For every upper node pair a tile touches, its
// contributions within the tile are listed in element order; the pair is
// completed in the tile when they are all its contributions, otherwise (or
// when more than six fall in one tile, in runs of six) it leaves partial
// sums, added up afterwards in tile order.
int build_tiles(fs_model *m, int which, int NE) {
    fs_ctx *c = m->ctx;
    Block &B = m->blk[which];
    TileStore &T = B.tiles;
    T.reset();
    const int nen = m->nen;
    const int nblk = nen * (nen + 1) / 2;
    const int64_t E = m->E, N = m->N;
    if (NE > 32 || nblk > 16) return fail(c, FS_ERR_UNSUPPORTED, "fused assembly: tile of %d elements", NE);
    // components: union-find over the nodes of each element
    std::vector<int32_t> par(N);
    for (int64_t i = 0; i < N; ++i) par[i] = (int32_t)i;
    auto find = [&](int32_t x) {
        while (par[x] != x) {
            par[x] = par[par[x]];
            x = par[x];
        }
        return x;
    };
    for (int64_t e = 0; e < E; ++e)
        for (int a = 1; a < nen; ++a) {
            const int32_t r0 = find(m->conn[e * nen]), r1 = find(m->conn[e * nen + a]);
            if (r0 != r1) par[std::max(r0, r1)] = std::min(r0, r1);
        }
    std::vector<int32_t> comp(E), cid(N, -1);
    int32_t ncomp = 0;
    for (int64_t e = 0; e < E; ++e) {
        const int32_t r = find(m->conn[e * nen]);
        if (cid[r] < 0) cid[r] = ncomp++;
        comp[e] = cid[r];
    }
    // centroids, component boxes, Morton keys
    std::vector<double> cen(3 * E), lo(3 * ncomp, 1e300), hi(3 * ncomp, -1e300);
    std::vector<double> X(3 * N);
    {
        std::vector<double> x(N), y(N), z(N);
        FS_HIP(c, hipMemcpyAsync(x.data(), m->x.p, N * 8, hipMemcpyDeviceToHost, c->stream));
        FS_HIP(c, hipMemcpyAsync(y.data(), m->y.p, N * 8, hipMemcpyDeviceToHost, c->stream));
        FS_HIP(c, hipMemcpyAsync(z.data(), m->z.p, N * 8, hipMemcpyDeviceToHost, c->stream));
        FS_HIP(c, hipStreamSynchronize(c->stream));
        for (int64_t i = 0; i < N; ++i) {
            X[3 * i] = x[i];
            X[3 * i + 1] = y[i];
            X[3 * i + 2] = z[i];
        }
    }
    for (int64_t e = 0; e < E; ++e) {
        for (int k = 0; k < 3; ++k) {
            double v = 0.0;
            for (int a = 0; a < nen; ++a) v += X[3 * (int64_t)m->conn[e * nen + a] + k];
            v /= nen;
            cen[3 * e + k] = v;
            lo[3 * comp[e] + k] = std::min(lo[3 * comp[e] + k], v);
            hi[3 * comp[e] + k] = std::max(hi[3 * comp[e] + k], v);
        }
    }
    std::vector<uint64_t> key(E);
    parallel_chunks(E, [&](int64_t e0, int64_t e1) {
        for (int64_t e = e0; e < e1; ++e) {
            const int g = comp[e];
            double ext = 0.0;
            for (int k = 0; k < 3; ++k) ext = std::max(ext, hi[3 * g + k] - lo[3 * g + k]);
            const double sc = ext > 0.0 ? 1.0 / (ext * (1.0 + 1e-9)) : 0.0;
            key[e] = morton3((cen[3 * e] - lo[3 * g]) * sc, (cen[3 * e + 1] - lo[3 * g + 1]) * sc,
                             (cen[3 * e + 2] - lo[3 * g + 2]) * sc);
        }
    });
    std::vector<int32_t> order(E);
    for (int64_t e = 0; e < E; ++e) order[e] = (int32_t)e;
    std::sort(order.begin(), order.end(), [&](int32_t a, int32_t b) {
        if (comp[a] != comp[b]) return comp[a] < comp[b];
        if (key[a] != key[b]) return key[a] < key[b];
        return a < b;
    });
    // tiles by recursive coordinate bisection of each component's element
    // centroids: a set of n > NE elements is cut across its longest extent
    // into ceil(k / 2) NE elements and the rest (k = ceil(n / NE), ties by
    // element number), so that every tile but the last of a set holds NE
    // elements and tiles are compact patches visited in a space-filling
    // order.  Against runs of NE elements along the Morton curve of the
    // centroids this cuts the node pairs that straddle tiles from 421,805 to
    // 389,058 and their partial slots from 912,990 to 823,835 on the 512 x 512
    // Q4 cylinder (12-element tiles); the T3 plate (32) is unchanged.
    std::vector<int64_t> tstart;
    {
        // the cuts are in place: the tiles are consecutive ranges of `order`
        // (in bisection order), and halves of more than 2^16 elements are
        // cut on threads of their own
        std::mutex mu;
        std::function<void(int64_t, int64_t)> bisect = [&](int64_t b, int64_t n) {
            if (n <= NE) {
                std::lock_guard<std::mutex> lk(mu);
                tstart.push_back(b);
                return;
            }
            int32_t *p = order.data() + b;
            const int64_t k = (n + NE - 1) / NE, n1 = (k / 2) * NE;
            double lo3[3] = {1e300, 1e300, 1e300}, hi3[3] = {-1e300, -1e300, -1e300};
            for (int64_t i = 0; i < n; ++i)
                for (int d = 0; d < 3; ++d) {
                    lo3[d] = std::min(lo3[d], cen[3 * (int64_t)p[i] + d]);
                    hi3[d] = std::max(hi3[d], cen[3 * (int64_t)p[i] + d]);
                }
            int ax = 0;
            for (int d = 1; d < 3; ++d)
                if (hi3[d] - lo3[d] > hi3[ax] - lo3[ax]) ax = d;
            std::nth_element(p, p + n1, p + n, [&](int32_t u, int32_t v) {
                const double cu = cen[3 * (int64_t)u + ax], cv = cen[3 * (int64_t)v + ax];
                return cu != cv ? cu < cv : u < v;
            });
            if (n > (int64_t(1) << 16)) {
                std::thread th(bisect, b + n1, n - n1);
                bisect(b, n1);
                th.join();
            } else {
                bisect(b, n1);
                bisect(b + n1, n - n1);
            }
        };
        for (int64_t b0 = 0; b0 < E;) {
            int64_t b1 = b0;
            while (b1 < E && comp[order[b1]] == comp[order[b0]]) ++b1;
            bisect(b0, b1 - b0);
            b0 = b1;
        }
        std::sort(tstart.begin(), tstart.end());
    }
    const int64_t ntiles = (int64_t)tstart.size();
    tstart.push_back(E);
    if (ntiles >= (int64_t(1) << 31) / 64) return fail(c, FS_ERR_UNSUPPORTED, "fused assembly: too many tiles");
    std::vector<int32_t> telem((size_t)ntiles * NE);
    for (int64_t t = 0; t < ntiles; ++t) {
        std::sort(order.begin() + tstart[t], order.begin() + tstart[t + 1]);
        const int64_t n = tstart[t + 1] - tstart[t];
        for (int le = 0; le < NE; ++le) telem[t * NE + le] = order[tstart[t] + (le < n ? le : 0)];
    }
    // pair entries per tile
    auto slot_of = [&](int64_t I, int64_t J) -> int32_t {
        const int32_t *b0 = m->nbr_list.data() + m->nbr_ptr[I];
        const int32_t *e0 = m->nbr_list.data() + m->nbr_ptr[I + 1];
        return (int32_t)(std::lower_bound(b0, e0, (int32_t)J) - m->nbr_list.data());
    };
    struct Ent {
        int32_t p;
        uint64_t cw;
        int32_t boundary;  // 1: leaves a partial sum
    };
    std::vector<std::vector<Ent>> tent(ntiles);
    std::atomic<int> bad{0};
    parallel_chunks(ntiles, [&](int64_t t0, int64_t t1) {
        std::vector<std::array<int32_t, 4>> tup;  // (pair, element, code, 0)
        for (int64_t t = t0; t < t1; ++t) {
            tup.clear();
            const int64_t n = tstart[t + 1] - tstart[t];
            for (int le = 0; le < n; ++le) {
                const int32_t e = telem[t * NE + le];
                const int32_t *ce = m->conn.data() + (int64_t)e * nen;
                int bl = 0;
                for (int a = 0; a < nen; ++a)
                    for (int b = a; b < nen; ++b, ++bl) {
                        const int32_t na = ce[a], nb = ce[b];
                        const int32_t I = std::min(na, nb), J = std::max(na, nb);
                        const int32_t k = slot_of(I, J);
                        const int32_t pp = B.h_nb_pair[k] >> 1;
                        const int32_t code = le | (bl << 5) | ((na > nb ? 1 : 0) << 9);
                        tup.push_back({pp, e, code, 0});
                    }
            }
            std::sort(tup.begin(), tup.end(), [](const std::array<int32_t, 4> &x, const std::array<int32_t, 4> &y) {
                return x[0] != y[0] ? x[0] < y[0] : x[1] < y[1];
            });
            std::vector<Ent> &out = tent[t];
            for (size_t i = 0; i < tup.size();) {
                size_t j = i;
                while (j < tup.size() && tup[j][0] == tup[i][0]) ++j;
                const int32_t pp = tup[i][0];
                const int cnt = (int)(j - i);
                const int gcnt = B.h_cptr[pp + 1] - B.h_cptr[pp];
                if (cnt > gcnt) bad = 1;
                const bool complete = cnt == gcnt && cnt <= 6;
                for (size_t q = i; q < j; q += 6) {
                    const int nq = (int)std::min<size_t>(6, j - q);
                    uint64_t cw = (uint64_t)nq;
                    for (int r = 0; r < nq; ++r) cw |= (uint64_t)tup[q + r][2] << (3 + 10 * r);
                    out.push_back({pp, cw, complete ? 0 : 1});
                }
                i = j;
            }
            // entries of one kind next to each other, so that the waves of the
            // output phase take one path: completed pairs of two nodes whose
            // dofs are all in the block, completed diagonal pairs, other
            // completed pairs, boundary pairs.  The order within a kind (and
            // of the entries of one pair) stays, so the sums do not change.
            auto kind = [&](const Ent &en) {
                if (en.boundary) return 3;
                const int64_t I = B.h_pnode[2 * (int64_t)en.p], J = B.h_pnode[2 * (int64_t)en.p + 1];
                if (I == J) return 1;
                const bool all = (B.h_node_rec[4 * I + 3] & 0xffffff) == 0x543210 &&
                                 (B.h_node_rec[4 * J + 3] & 0xffffff) == 0x543210;
                return all ? 0 : 2;
            };
            std::stable_sort(out.begin(), out.end(),
                             [&](const Ent &x, const Ent &y) { return kind(x) < kind(y); });
        }
    }, 256);
    if (bad) return fail(c, FS_ERR_STATE, "fused assembly: tile contributions exceed the pair's");
    // flatten; partial slots in tile order; boundary pairs and their slots
    std::vector<int32_t> tp_ptr(ntiles + 1, 0);
    for (int64_t t = 0; t < ntiles; ++t) tp_ptr[t + 1] = tp_ptr[t] + (int32_t)tent[t].size();
    const int64_t nent = tp_ptr[ntiles];
    // placement record of upper pair pp (fs::TilePlan trec)
    auto place = [&](int64_t pp, int32_t *rec) {
        const int64_t I = B.h_pnode[2 * pp], J = B.h_pnode[2 * pp + 1];
        const int64_t kI = B.h_pslot[2 * pp], kJ = B.h_pslot[2 * pp + 1];
        for (int k = 0; k < 4; ++k) {
            rec[k] = B.h_node_rec[4 * I + k];
            rec[4 + k] = B.h_node_rec[4 * J + k];
        }
        if (I == J) rec[3] |= 1 << 30;
        rec[8] = B.h_nb_off[2 * kI];
        rec[9] = B.h_nb_off[2 * kI + 1];
        rec[10] = B.h_nb_off[2 * kJ];
        rec[11] = B.h_nb_off[2 * kJ + 1];
    };
    std::vector<int32_t> trec(12 * nent), tslot(nent);
    std::vector<uint64_t> tcw(nent);
    // partial slots: the slots of one boundary pair are consecutive (its
    // tiles in tile order), so the finalize reads each pair's partials as one
    // stretch
    std::vector<int32_t> pcount(B.npairs + 1, 0);
    for (int64_t t = 0; t < ntiles; ++t)
        for (const Ent &en : tent[t])
            if (en.boundary) pcount[en.p + 1]++;
    std::vector<int32_t> bidx(B.npairs, -1), brec, bp_ptr(1, 0);
    for (int64_t pp = 0; pp < B.npairs; ++pp)
        if (pcount[pp + 1]) {
            bidx[pp] = (int32_t)(bp_ptr.size() - 1);
            brec.resize(brec.size() + 12);
            place(pp, &brec[brec.size() - 12]);
            bp_ptr.push_back(bp_ptr.back() + pcount[pp + 1]);
        }
    const int64_t nbnd = (int64_t)bp_ptr.size() - 1;
    const int64_t nslots = bp_ptr.back();
    std::vector<int32_t> bfill(bp_ptr.begin(), bp_ptr.end() - 1);
    for (int64_t t = 0; t < ntiles; ++t)
        for (size_t i = 0; i < tent[t].size(); ++i) {
            const Ent &en = tent[t][i];
            const int64_t q = tp_ptr[t] + (int64_t)i;
            place(en.p, &trec[12 * q]);
            tcw[q] = en.cw;
            tslot[q] = en.boundary ? bfill[bidx[en.p]]++ : -1;
        }
    int st;
    // the tiles' own connectivity (node a of tile element le at (t NE + le) nen + a)
    std::vector<int32_t> tconn(telem.size() * nen);
    for (size_t i = 0; i < telem.size(); ++i)
        for (int a = 0; a < nen; ++a) tconn[i * nen + a] = m->conn[(size_t)telem[i] * nen + a];
    if ((st = upload(c, T.tile_elem, telem.data(), telem.size())) || (st = upload(c, T.tile_conn, tconn.data(), tconn.size())) ||
        (st = upload(c, T.tp_ptr, tp_ptr.data(), tp_ptr.size())) ||
        (st = upload(c, T.trec, trec.data(), trec.size())) || (st = upload(c, T.tcontrib, tcw.data(), tcw.size())) ||
        (st = upload(c, T.tslot, tslot.data(), tslot.size())) ||
        (st = upload(c, T.brec, brec.data(), brec.size())) ||
        (st = upload(c, T.bp_ptr, bp_ptr.data(), bp_ptr.size())) ||
        (st = dalloc(c, T.partial, (size_t)std::max<int64_t>(nslots, 1) * 36 * 8)))
        return st;
    FS_HIP(c, hipStreamSynchronize(c->stream));
    T.ne = NE;
    T.ntiles = ntiles;
    T.nent = nent;
    T.nbnd = nbnd;
    T.nslots = nslots;
    return FS_OK;
}

// Element-kernel launch on the model's device arrays.
hipError_t launch_elements(fs_model *m, const fs::ElemParams &ep, int variant, hipStream_t st) {
    return fs::launch_element_blocks(ep, m->x.as<double>(), m->y.as<double>(), m->z.as<double>(),
                                     m->dconn.as<int32_t>(), m->nx.as<double>(), m->ny.as<double>(),
                                     m->nz.as<double>(), m->valid.as<uint8_t>(), m->Kb.as<double>(),
                                     m->status.as<int>(), st, variant);
}

// Validate the stiffness parameters, upload the per-group material records,
// size m->Kb, clear the status flag and pick the element-kernel launch shape
// (the autotuner times every shape on the first evaluation of a kind).
int prepare_elements(fs_model *m, const fs_shell_params *p, fs::ElemParams &ep, int *variant_out, bool blocks = true) {
    fs_ctx *c = m->ctx;
    if (!p) return fail(c, FS_ERR_ARG, "stiffness: NULL parameters");
    if (!m->normals_ready) return fail(c, FS_ERR_STATE, "stiffness: normals not computed (associategeometry!)");
    if (p->formulation != FS_SHEAR_AVERAGE_B && p->formulation != FS_SHEAR_AVERAGE_K)
        return fail(c, FS_ERR_ARG, "stiffness: unknown transverse shear formulation %d", p->formulation);
    if (m->nen == FS_Q4 && p->formulation != FS_SHEAR_AVERAGE_B)
        return fail(c, FS_ERR_ARG, "stiffness: the Q4 family has a single (MITC) shear formulation");
    std::memset(&ep, 0, sizeof ep);
    ep.nen = m->nen;
    ep.composite = p->composite ? 1 : 0;
    ep.formulation = p->formulation;
    ep.nelem = (int32_t)m->E;
    ep.ebegin = 0;
    ep.eend = (int32_t)m->E;
    ep.nnodes = m->N;
    ep.thickness = p->thickness;
    ep.stab_c = p->stab_c;
    ep.drill_scale = p->drilling_stiffness_scale;
    int s = 0;
    if (!p->composite) {
        // one material record per group: chol(Dps), chol(5/6 Dt), t, stab_c, drilling scale
        const bool grouped = p->group_homog != nullptr;
        const int ng = grouped ? p->ngroups : 1;
        if (grouped && ng < 1) return fail(c, FS_ERR_ARG, "stiffness: group_homog needs ngroups >= 1");
        std::vector<double> rec(16 * (size_t)ng);
        for (int g = 0; g < ng; ++g) {
            const double *src = grouped ? p->group_homog + 16 * g : nullptr;
            double Dps[9], Dt[4];
            cm2rm(3, grouped ? src : p->Dps, Dps);
            cm2rm(2, grouped ? src + 9 : p->Dt, Dt);
            for (int i = 0; i < 4; ++i) Dt[i] *= 5.0 / 6.0;  // FEMMShellT3FFModule.jl:658-659
            double *r = &rec[16 * g];
            if (!chol_upper(3, Dps, r)) return fail(c, FS_ERR_NUMERIC, "stiffness: Dps of group %d is not positive definite", g);
            if (!chol_upper(2, Dt, r + 9)) return fail(c, FS_ERR_NUMERIC, "stiffness: Dt of group %d is not positive definite", g);
            r[13] = grouped ? src[13] : p->thickness;
            r[14] = grouped ? src[14] : p->stab_c;
            r[15] = grouped ? src[15] : p->drilling_stiffness_scale;
            if (!(r[13] > 0.0)) return fail(c, FS_ERR_ARG, "stiffness: thickness of group %d must be positive", g);
        }
        if ((s = upload(c, m->g_homog, rec.data(), rec.size()))) return s;
        ep.g_homog = m->g_homog.as<double>();
        if (grouped && p->elem_group) {
            for (int64_t e = 0; e < m->E; ++e)
                if (p->elem_group[e] < 0 || p->elem_group[e] >= ng)
                    return fail(c, FS_ERR_ARG, "stiffness: element %lld has group %d", (long long)e, p->elem_group[e]);
            // one group: no group array, so the kernels read one material record
            if (ng > 1) {
                if ((s = upload(c, m->g_group, p->elem_group, (size_t)m->E))) return s;
                ep.elem_group = m->g_group.as<int32_t>();
            }
        }
    } else {
        const int ng = p->ngroups;
        if (ng < 1 || !p->group_ABDH || !p->group_csys || !p->group_thickness)
            return fail(c, FS_ERR_ARG, "stiffness: layered shell needs ngroups >= 1 and group arrays");
        std::vector<double> Labd(36 * ng), Lh(4 * ng), csys(9 * ng), tg(ng);
        for (int g = 0; g < ng; ++g) {
            const double *src = p->group_ABDH + 31 * g;
            double A[9], Bm[9], D[9], H[4], ABD[36];
            cm2rm(3, src, A);
            cm2rm(3, src + 9, Bm);
            cm2rm(3, src + 18, D);
            cm2rm(2, src + 27, H);
            for (int i = 0; i < 3; ++i)
                for (int j = 0; j < 3; ++j) {
                    ABD[i * 6 + j] = A[i * 3 + j];
                    ABD[i * 6 + 3 + j] = Bm[i * 3 + j];
                    ABD[(3 + i) * 6 + j] = Bm[j * 3 + i];
                    ABD[(3 + i) * 6 + 3 + j] = D[i * 3 + j];
                }
            if (!chol_upper(6, ABD, &Labd[36 * g]))
                return fail(c, FS_ERR_NUMERIC, "stiffness: laminate [A B; B D] of group %d is not positive definite", g);
            if (!chol_upper(2, H, &Lh[4 * g]))
                return fail(c, FS_ERR_NUMERIC, "stiffness: laminate H of group %d is not positive definite", g);
            cm2rm(3, p->group_csys + 9 * g, &csys[9 * g]);
            tg[g] = p->group_thickness[g];
        }
        if ((s = upload(c, m->g_Labd, Labd.data(), Labd.size()))) return s;
        if ((s = upload(c, m->g_Lh, Lh.data(), Lh.size()))) return s;
        if ((s = upload(c, m->g_csys, csys.data(), csys.size()))) return s;
        if ((s = upload(c, m->g_t, tg.data(), tg.size()))) return s;
        ep.g_Labd = m->g_Labd.as<double>();
        ep.g_Lh = m->g_Lh.as<double>();
        ep.g_csys = m->g_csys.as<double>();
        ep.g_t = m->g_t.as<double>();
        if (p->elem_group) {
            for (int64_t e = 0; e < m->E; ++e)
                if (p->elem_group[e] < 0 || p->elem_group[e] >= ng)
                    return fail(c, FS_ERR_ARG, "stiffness: element %lld has group %d", (long long)e, p->elem_group[e]);
            if (ng > 1) {
                if ((s = upload(c, m->g_group, p->elem_group, (size_t)m->E))) return s;
                ep.elem_group = m->g_group.as<int32_t>();
            }
        }
    }
    FS_HIP(c, hipMemsetAsync(m->status.p, 0, sizeof(int), c->stream));
    if (!blocks) {
        *variant_out = 0;  // the fused path has one launch shape per family
        return FS_OK;
    }
    const int nblk = m->nen * (m->nen + 1) / 2;
    const size_t kbytes = (size_t)m->E * nblk * 36 * sizeof(double);
    if (m->Kb.bytes < kbytes && (s = dalloc(c, m->Kb, kbytes))) return s;
    int &variant = c->elem_variant[m->nen == 4 ? 1 : 0][ep.formulation == FS_SHEAR_AVERAGE_K ? 1 : 0];
    if (variant < 0 && !c->autotune) variant = 0;
    if (variant < 0) {
        // autotuner: one warm launch and one timed launch per shape on
        // this model; every shape computes bitwise the same blocks
        hipEvent_t e0, e1;
        FS_HIP(c, hipEventCreate(&e0));
        FS_HIP(c, hipEventCreate(&e1));
        float best = 0.f;
        int bestv = 0;
        for (int v = 0; v < fs::element_variants(m->nen); ++v) {
            float t = 0.f;
            for (int rep = 0; rep < 2; ++rep) {
                FS_HIP(c, hipEventRecord(e0, c->stream));
                FS_HIP(c, launch_elements(m, ep, v, c->stream));
                FS_HIP(c, hipEventRecord(e1, c->stream));
                FS_HIP(c, hipEventSynchronize(e1));
                FS_HIP(c, hipEventElapsedTime(&t, e0, e1));
            }
            if (v == 0 || t < best) {
                best = t;
                bestv = v;
            }
        }
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        variant = bestv;
    }
    *variant_out = variant;
    return FS_OK;
}

// After the stream has been synchronised: the singular-metric flag of the
// element kernels, and the bookkeeping of a successful evaluation.
int finish_elements(fs_model *m, const fs::ElemParams &ep, int status) {
    if (status)
        return fail(m->ctx, FS_ERR_NUMERIC, "Singular metric matrix in _gradN_e! (degenerate quadrilateral)");
    m->last_ep = ep;
    ++m->kgen;
    return FS_OK;
}

// Compute element node-pair blocks into m->Kb (device), timed with events.
int element_blocks(fs_model *m, const fs_shell_params *p, float *elem_ms) {
    fs_ctx *c = m->ctx;
    fs::ElemParams ep;
    int variant = 0;
    int s = prepare_elements(m, p, ep, &variant);
    if (s) return s;
    hipEvent_t e0, e1;
    FS_HIP(c, hipEventCreate(&e0));
    FS_HIP(c, hipEventCreate(&e1));
    FS_HIP(c, hipEventRecord(e0, c->stream));
    FS_HIP(c, launch_elements(m, ep, variant, c->stream));
    FS_HIP(c, hipEventRecord(e1, c->stream));
    int st = 0;
    FS_HIP(c, hipMemcpyAsync(&st, m->status.p, sizeof(int), hipMemcpyDeviceToHost, c->stream));
    FS_HIP(c, hipStreamSynchronize(c->stream));
    FS_HIP(c, hipEventElapsedTime(elem_ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return finish_elements(m, ep, st);
}

// Build the mixed-precision operator from the parameters of the last
// stiffness evaluation: split element kernel (membrane + bending blocks,
// shear factor rows, drilling blocks), then the FP32 slices of K_mb and the
// per-node drilling blocks.
int build_mixed(fs_model *m, bool shear_fp32) {
    fs_ctx *c = m->ctx;
    Block &B = m->blk[FS_BLOCK_FF];
    MixedStore &X = m->mixed;
    X.reset();
    const int nen = m->nen;
    const int nblk = nen * (nen + 1) / 2;
    const int K = 6 * nen;
    const int nsr = shear_fp32 ? 0 : nen == 3 ? (m->last_ep.formulation == FS_SHEAR_AVERAGE_K ? 6 : 2) : 8;
    const int64_t N = m->N, E = m->E, nrows = 6 * N;
    // slices of 10 nodes (60 rows): slot count = largest neighbour count of the slice's nodes
    (void)nrows;
    const int64_t ns = (N + 9) / 10;
    std::vector<int64_t> goff(ns + 1, 0);
    for (int64_t g = 0; g < ns; ++g) {
        const int64_t i0 = 10 * g, i1 = std::min<int64_t>(N - 1, 10 * g + 9);
        int mx = 0;
        for (int64_t i = i0; i <= i1; ++i) mx = std::max(mx, m->nbr_ptr[i + 1] - m->nbr_ptr[i]);
        goff[g + 1] = goff[g] + (int64_t)mx * 3 * 60;
    }
    int s;
    DevBuf kbm, dre;
    if ((s = dalloc(c, kbm, (size_t)E * nblk * 36 * 8)) || (s = dalloc(c, dre, (size_t)E * nen * 9 * 8)) ||
        (nsr > 0 && (s = dalloc(c, X.cs, (size_t)E * nsr * K * 8))) || (s = dalloc(c, X.dr, (size_t)N * 6 * 8)) ||
        (s = upload(c, X.goff, goff.data(), goff.size())) ||
        (s = dalloc(c, X.kmb, (size_t)goff[ns] * sizeof(float2))) ||
        (s = dalloc(c, X.kml, (size_t)goff[ns] * sizeof(short2))))
        return s;
    if (nsr > 0) {
        // incidence records: element and its nodes, in each node's ascending element order
        const int ps = nen == 3 ? 4 : 8;
        std::vector<int32_t> pack((size_t)m->ne_list.size() * ps, 0);
        for (size_t k = 0; k < m->ne_list.size(); ++k) {
            const int32_t e = m->ne_list[k];
            pack[k * ps] = e;
            for (int a = 0; a < nen; ++a) pack[k * ps + 1 + a] = m->conn[(size_t)e * nen + a];
        }
        if ((s = upload(c, X.ne_pack, pack.data(), pack.size()))) return s;
    }
    fs::ElemParams ep = m->last_ep;
    ep.split_cs = nsr > 0 ? X.cs.as<double>() : nullptr;  // nullptr: shear stays in the blocks
    ep.split_dr = dre.as<double>();
    hipEvent_t e0, e1;
    FS_HIP(c, hipEventCreate(&e0));
    FS_HIP(c, hipEventCreate(&e1));
    FS_HIP(c, hipEventRecord(e0, c->stream));
    FS_HIP(c, hipMemsetAsync(X.kmb.p, 0, (size_t)goff[ns] * sizeof(float2), c->stream));
    FS_HIP(c, hipMemsetAsync(X.kml.p, 0, (size_t)goff[ns] * sizeof(short2), c->stream));
    FS_HIP(c, fs::launch_element_blocks(ep, m->x.as<double>(), m->y.as<double>(), m->z.as<double>(),
                                        m->dconn.as<int32_t>(), m->nx.as<double>(), m->ny.as<double>(),
                                        m->nz.as<double>(), m->valid.as<uint8_t>(), kbm.as<double>(),
                                        m->status.as<int>(), c->stream, 0));
    FS_HIP(c, fs::launch_mixed_build(plan_of(B), kbm.as<double>(), X.goff.as<int64_t>(), X.kmb.as<float2>(), X.kml.as<short2>(),
                                     c->stream));
    FS_HIP(c, fs::launch_mixed_drill(nen, N, (int)E, m->dne_ptr.as<int32_t>(), m->dne_list.as<int32_t>(),
                                     m->dconn.as<int32_t>(), dre.as<double>(), X.dr.as<double>(), c->stream));
    FS_HIP(c, hipEventRecord(e1, c->stream));
    FS_HIP(c, hipStreamSynchronize(c->stream));
    FS_HIP(c, hipEventElapsedTime(&X.build_ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    X.nsr = nsr;
    X.shear_fp32 = shear_fp32;
    X.nslices = ns;
    X.kmb_pairs = goff[ns];
    X.kgen = m->kgen;
    X.built = true;
    return FS_OK;
}

fs::MixedOp mixed_op(const fs_model *m) {
    const Block &B = m->blk[FS_BLOCK_FF];
    fs::MixedOp op;
    op.nrows = 6 * m->N;
    op.nnodes = (int32_t)m->N;
    op.nelem = (int32_t)m->E;
    op.goff = m->mixed.goff.as<int64_t>();
    op.kmb = m->mixed.kmb.as<float2>();
    op.kml = m->mixed.kml.as<short2>();
    op.nbr_ptr = B.nbr_ptr.as<int32_t>();
    op.nbr_node = B.nbr_node.as<int32_t>();
    op.dr = m->mixed.dr.as<double>();
    op.ne_ptr = m->dne_ptr.as<int32_t>();
    op.ne_pack = m->mixed.ne_pack.as<int32_t>();
    op.cs = m->mixed.cs.as<double>();
    return op;
}

}  // namespace

extern "C" {

const char *fs_version(void) { return "flexshell 0.1.0 (gfx950)"; }

int fs_ctx_create(int device, fs_ctx **out) {
    if (!out) return FS_ERR_ARG;
    *out = nullptr;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) return FS_ERR_DEVICE;
    if (device < 0 || device >= n) return FS_ERR_ARG;
    fs_ctx *c = new (std::nothrow) fs_ctx;
    if (!c) return FS_ERR_NOMEM;
    c->cache = new (std::nothrow) BufCache;
    if (!c->cache) {
        delete c;
        return FS_ERR_NOMEM;
    }
    c->device = device;
    if (hipSetDevice(device) != hipSuccess || hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
        delete c->cache;
        delete c;
        return FS_ERR_DEVICE;
    }
    // with `device` current: the device's copy of the allocation-free workspace
    c->scratch_status = fs::noalloc_workspace(device, &c->scratch);
    if (c->scratch_status != hipSuccess) (void)hipGetLastError();
    *out = c;
    return FS_OK;
}

}  // extern "C"

namespace {
// Drops one reference; the last one (the caller's handle or the last live
// model) destroys the streams, events, staging buffers and the context.
void ctx_release(fs_ctx *ctx) {
    if (--ctx->refs > 0) return;
    (void)hipSetDevice(ctx->device);
    if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
    if (ctx->stream2) {
        (void)hipStreamSynchronize(ctx->stream2);
        (void)hipStreamDestroy(ctx->stream2);
    }
    for (hipEvent_t ev : ctx->ev) (void)hipEventDestroy(ev);
    for (int i = 0; i < 2; ++i) {
        if (ctx->stage[i]) (void)hipHostFree(ctx->stage[i]);
        if (ctx->stage_ev[i]) (void)hipEventDestroy(ctx->stage_ev[i]);
    }
    if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
    delete ctx;
}
}  // namespace

extern "C" {

void fs_ctx_destroy(fs_ctx *ctx) {
    if (!ctx) return;
    (void)hipSetDevice(ctx->device);
    if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
    if (ctx->stream2) (void)hipStreamSynchronize(ctx->stream2);
    int left;
    {
        std::lock_guard<std::mutex> lk(ctx->cache->mu);
        for (auto &kv : ctx->cache->spare) (void)hipFree(kv.second);
        ctx->cache->spare.clear();
        ctx->cache->spare_bytes = 0;
        ctx->cache->alive = false;  // buffers of models still alive are freed when released
        left = --ctx->cache->refs;
    }
    if (left == 0) delete ctx->cache;
    ctx->cache = nullptr;
    ctx_release(ctx);
}

const char *fs_last_error(const fs_ctx *ctx) { return ctx ? ctx->err.c_str() : "no context"; }

int fs_model_create(fs_ctx *ctx, int nen, int64_t nnodes, const double *xyz, int64_t nelem, const int64_t *conn,
                    int index_base, fs_model **out) {
    if (!ctx || !out) return FS_ERR_ARG;
    *out = nullptr;
    ctx->err.clear();
    if (nen != FS_T3 && nen != FS_Q4) return fail(ctx, FS_ERR_ARG, "nen must be 3 or 4, got %d", nen);
    if (nnodes <= 0 || nelem <= 0 || !xyz || !conn) return fail(ctx, FS_ERR_ARG, "empty mesh or NULL arrays");
    if (nnodes >= (int64_t(1) << 31) || nelem >= (int64_t(1) << 31))
        return fail(ctx, FS_ERR_UNSUPPORTED, "mesh too large for 32-bit node/element indices");
    FS_HIP(ctx, hipSetDevice(ctx->device));
    fs_model *m = new (std::nothrow) fs_model;
    if (!m) return fail(ctx, FS_ERR_NOMEM, "model allocation");
    m->ctx = ctx;
    m->nen = nen;
    m->N = nnodes;
    m->E = nelem;
    m->conn.resize((size_t)nelem * nen);
    for (int64_t i = 0; i < nelem * nen; ++i) {
        const int64_t v = conn[i] - index_base;
        if (v < 0 || v >= nnodes) {
            delete m;
            return fail(ctx, FS_ERR_ARG, "connectivity entry %lld out of range", (long long)i);
        }
        m->conn[i] = (int32_t)v;
    }
    for (int64_t e = 0; e < nelem; ++e)
        for (int a = 0; a < nen; ++a)
            for (int b = a + 1; b < nen; ++b)
                if (m->conn[e * nen + a] == m->conn[e * nen + b]) {
                    delete m;
                    return fail(ctx, FS_ERR_ARG, "element %lld repeats a node", (long long)e);
                }
    // node -> element lists in ascending element order
    m->ne_ptr.assign(nnodes + 1, 0);
    for (int64_t i = 0; i < nelem * nen; ++i) m->ne_ptr[m->conn[i] + 1]++;
    for (int64_t n = 0; n < nnodes; ++n) m->ne_ptr[n + 1] += m->ne_ptr[n];
    m->ne_list.resize(m->ne_ptr[nnodes]);
    {
        std::vector<int32_t> fill(m->ne_ptr.begin(), m->ne_ptr.end() - 1);
        for (int64_t e = 0; e < nelem; ++e)
            for (int a = 0; a < nen; ++a) m->ne_list[fill[m->conn[e * nen + a]]++] = (int32_t)e;
    }
    // node -> neighbour lists (sorted, including the node itself); a node in
    // no element has no neighbours: its rows stay empty, as in the COO
    // assembly.  Two passes over the nodes (sizes, then lists).
    m->nbr_ptr.assign(nnodes + 1, 0);
    auto neighbours = [&](int64_t n, std::vector<int32_t> &tmp) {
        tmp.clear();
        for (int k = m->ne_ptr[n]; k < m->ne_ptr[n + 1]; ++k) {
            const int e = m->ne_list[k];
            for (int a = 0; a < nen; ++a) tmp.push_back(m->conn[(int64_t)e * nen + a]);
        }
        std::sort(tmp.begin(), tmp.end());
        tmp.erase(std::unique(tmp.begin(), tmp.end()), tmp.end());
    };
    parallel_chunks(nnodes, [&](int64_t n0, int64_t n1) {
        std::vector<int32_t> tmp;
        for (int64_t n = n0; n < n1; ++n) {
            neighbours(n, tmp);
            m->nbr_ptr[n + 1] = (int32_t)tmp.size();
        }
    });
    for (int64_t n = 0; n < nnodes; ++n) m->nbr_ptr[n + 1] += m->nbr_ptr[n];
    m->nbr_list.resize(m->nbr_ptr[nnodes]);
    parallel_chunks(nnodes, [&](int64_t n0, int64_t n1) {
        std::vector<int32_t> tmp;
        for (int64_t n = n0; n < n1; ++n) {
            neighbours(n, tmp);
            std::copy(tmp.begin(), tmp.end(), m->nbr_list.begin() + m->nbr_ptr[n]);
        }
    });
    // device geometry (SoA) and connectivity planes
    int s;
    std::vector<int32_t> planes((size_t)nelem * nen);
    parallel_chunks(nelem, [&](int64_t e0, int64_t e1) {
        for (int64_t e = e0; e < e1; ++e)
            for (int a = 0; a < nen; ++a) planes[(size_t)a * nelem + e] = m->conn[e * nen + a];
    });
    if ((s = upload(ctx, m->x, xyz, nnodes)) || (s = upload(ctx, m->y, xyz + nnodes, nnodes)) ||
        (s = upload(ctx, m->z, xyz + 2 * nnodes, nnodes)) || (s = upload(ctx, m->dconn, planes.data(), planes.size())) ||
        (s = upload(ctx, m->dne_ptr, m->ne_ptr.data(), m->ne_ptr.size())) ||
        (s = upload(ctx, m->dne_list, m->ne_list.data(), m->ne_list.size())) ||
        (s = dalloc(ctx, m->nx, nnodes * 8)) || (s = dalloc(ctx, m->ny, nnodes * 8)) ||
        (s = dalloc(ctx, m->nz, nnodes * 8)) || (s = dalloc(ctx, m->valid, nnodes)) ||
        (s = dalloc(ctx, m->status, sizeof(int)))) {
        delete m;
        return s;
    }
    if (hipStreamSynchronize(ctx->stream) != hipSuccess) {
        delete m;
        return fail(ctx, FS_ERR_DEVICE, "upload of the mesh failed");
    }
    ++ctx->refs;
    *out = m;
    return FS_OK;
}

void fs_model_destroy(fs_model *m) {
    if (!m) return;
    fs_ctx *c = m->ctx;
    (void)hipSetDevice(c->device);
    (void)hipStreamSynchronize(c->stream);
    if (c->stream2) (void)hipStreamSynchronize(c->stream2);
    delete m;  // its buffers go back to the context's spare list, or to hipFree once the context is destroyed
    ctx_release(c);
}

int fs_model_set_dofs(fs_model *m, const int64_t *dofnums, int64_t nalldofs, int64_t nfreedofs, int index_base) {
    if (!m) return FS_ERR_ARG;
    fs_ctx *c = m->ctx;
    c->err.clear();
    if (!dofnums || nalldofs <= 0 || nfreedofs < 0 || nfreedofs > nalldofs)
        return fail(c, FS_ERR_ARG, "set_dofs: bad sizes (nalldofs %lld, nfreedofs %lld)", (long long)nalldofs,
                    (long long)nfreedofs);
    const int64_t N = m->N;
    std::vector<int64_t> dof(N * 6);
    std::vector<uint8_t> seen(nalldofs, 0);
    for (int64_t I = 0; I < N; ++I)
        for (int cc = 0; cc < 6; ++cc) {
            const int64_t d = dofnums[cc * N + I] - index_base;  // column-major N x 6
            if (d < 0 || d >= nalldofs) return fail(c, FS_ERR_ARG, "dof number out of range at node %lld", (long long)I);
            if (seen[d]) return fail(c, FS_ERR_ARG, "dof number %lld used twice", (long long)(d + index_base));
            seen[d] = 1;
            dof[I * 6 + cc] = d;
        }
    if ((int64_t)N * 6 != nalldofs) return fail(c, FS_ERR_ARG, "nalldofs must equal 6 * nnodes");
    FS_HIP(c, hipSetDevice(c->device));
    m->dof.swap(dof);
    m->nall = nalldofs;
    m->nfree = nfreedofs;
    m->blk[0].reset();
    m->blk[1].reset();
    m->mixed.reset();
    int s = upload(c, m->node_dof, m->dof.data(), m->dof.size());
    if (s) return s;
    FS_HIP(c, hipStreamSynchronize(c->stream));
    m->dofs_ready = true;
    return FS_OK;
}

int fs_model_number_dofs(fs_model *m, const uint8_t *is_fixed, const int64_t *perm, int64_t *dofnums_out,
                         int64_t *nfree_out) {
    if (!m) return FS_ERR_ARG;
    fs_ctx *c = m->ctx;
    c->err.clear();
    if (!is_fixed) return fail(c, FS_ERR_ARG, "number_dofs: NULL flags");
    const int64_t N = m->N, n = N * 6;
    std::vector<uint8_t> pf;
    if (perm) {
        // flags in numbering order: node perm[k] is the k-th node numbered
        std::vector<uint8_t> seen((size_t)N, 0);
        pf.resize((size_t)n);
        for (int64_t k = 0; k < N; ++k) {
            const int64_t I = perm[k];
            if (I < 0 || I >= N || seen[I]) return fail(c, FS_ERR_ARG, "number_dofs: perm is not a permutation");
            seen[I] = 1;
            for (int cc = 0; cc < 6; ++cc) pf[k * 6 + cc] = is_fixed[I * 6 + cc];
        }
    }
    FS_HIP(c, hipSetDevice(c->device));
    DevBuf dfix, dtile, dtot;
    int s;
    if ((s = upload(c, dfix, perm ? pf.data() : is_fixed, (size_t)n)) ||
        (s = dalloc(c, dtile, fs::number_dofs_tiles(n) * 8 + 8)) || (s = dalloc(c, dtot, 8)) ||
        (s = dalloc(c, m->node_dof, (size_t)n * 8)))
        return s;
    FS_HIP(c, hipMemsetAsync(dtot.p, 0, 8, c->stream));
    FS_HIP(c, fs::launch_number_dofs(n, dfix.as<uint8_t>(), dtile.as<int64_t>(), dtot.as<int64_t>(),
                                     m->node_dof.as<int64_t>(), c->stream));
    std::vector<int64_t> dof((size_t)n);
    int64_t nfree = 0;
    FS_HIP(c, hipMemcpyAsync(&nfree, dtot.p, 8, hipMemcpyDeviceToHost, c->stream));
    FS_HIP(c, hipMemcpyAsync(dof.data(), m->node_dof.p, (size_t)n * 8, hipMemcpyDeviceToHost, c->stream));
    FS_HIP(c, hipStreamSynchronize(c->stream));
    if (perm) {
        // back to node order on the host, and the device copy replaced
        std::vector<int64_t> byn((size_t)n);
        for (int64_t k = 0; k < N; ++k)
            for (int cc = 0; cc < 6; ++cc) byn[perm[k] * 6 + cc] = dof[k * 6 + cc];
        dof.swap(byn);
        if ((s = upload(c, m->node_dof, dof.data(), dof.size()))) return s;
        FS_HIP(c, hipStreamSynchronize(c->stream));
    }
    // the host copy feeds the pattern planner
    m->dof.swap(dof);
    m->nall = n;
    m->nfree = nfree;
    m->blk[0].reset();
    m->blk[1].reset();
    m->mixed.reset();
    m->dofs_ready = true;
    if (dofnums_out)
        for (int64_t I = 0; I < m->N; ++I)
            for (int cc = 0; cc < 6; ++cc) dofnums_out[cc * m->N + I] = m->dof[I * 6 + cc];
    if (nfree_out) *nfree_out = nfree;
    return FS_OK;
}

static bool load_rule_ok(int nen, int npts) { return nen == 3 ? (npts == 1 || npts == 3 || npts == 6) : npts == 4; }

int fs_model_surface_load(fs_model *m, int32_t npts, int32_t kind, const double *force, int block, double *F_out) {
    if (!m) return FS_ERR_ARG;
    fs_ctx *c = m->ctx;
    c->err.clear();
    if (!force || !F_out) return fail(c, FS_ERR_ARG, "surface_load: NULL arguments");
    if (kind != FS_LOAD_UNIFORM && kind != FS_LOAD_PER_POINT) return fail(c, FS_ERR_ARG, "surface_load: kind %d", kind);
    if (block != FS_BLOCK_ALL && block != FS_BLOCK_FF) return fail(c, FS_ERR_ARG, "unknown block %d", block);
    if (!m->dofs_ready) return fail(c, FS_ERR_STATE, "surface_load: equation numbers not set");
    if (!load_rule_ok(m->nen, npts)) return fail(c, FS_ERR_ARG, "surface_load: %d-point rule", npts);
    FS_HIP(c, hipSetDevice(c->device));
    const int64_t nrows = block == FS_BLOCK_FF ? m->nfree : m->nall;
    DevBuf dforce, dF;
    int s;
    if ((s = upload(c, dforce, force, kind == FS_LOAD_UNIFORM ? 6 : (size_t)m->E * npts * 6)) ||
        (s = dalloc(c, dF, (size_t)nrows * 8)))
        return s;
    FS_HIP(c, fs::launch_surface_load(m->nen, npts, m->N, (int)m->E, m->x.as<double>(), m->y.as<double>(),
                                      m->z.as<double>(), m->dconn.as<int32_t>(), m->dne_ptr.as<int32_t>(),
                                      m->dne_list.as<int32_t>(), kind == FS_LOAD_PER_POINT ? 1 : 0,
                                      dforce.as<double>(), m->node_dof.as<int64_t>(), nrows, dF.as<double>(),
                                      c->stream));
    return d2h(c, F_out, dF.p, (size_t)nrows * 8);
}

int fs_model_load_points(fs_model *m, int32_t npts, double *xyz_out) {
    if (!m) return FS_ERR_ARG;
    fs_ctx *c = m->ctx;
    c->err.clear();
    if (!xyz_out) return fail(c, FS_ERR_ARG, "load_points: NULL output");
    if (!load_rule_ok(m->nen, npts)) return fail(c, FS_ERR_ARG, "load_points: %d-point rule", npts);
    FS_HIP(c, hipSetDevice(c->device));
    const size_t n = (size_t)m->E * npts * 3;
    DevBuf d;
    int s = dalloc(c, d, n * 8);
    if (s) return s;
    FS_HIP(c, fs::launch_load_points(m->nen, npts, (int)m->E, m->x.as<double>(), m->y.as<double>(), m->z.as<double>(),
                                     m->dconn.as<int32_t>(), d.as<double>(), c->stream));
    return d2h(c, xyz_out, d.p, n * 8);
}

// Normals kernel in either form (gnrm null: isoparametric) and the copies out.
static int normals_common(fs_model *m, double threshold_angle_deg, const double *gnrm, const int32_t *egroup,
                          double *normals_out, uint8_t *valid_out) {
    fs_ctx *c = m->ctx;
    const double s = std::sin(threshold_angle_deg / 180 * M_PI);
    const double ntol = 1 - std::sqrt(1 - s * s);  // FEMMShellT3FFModule.jl:600
    FS_HIP(c, fs::launch_normals(m->nen, m->N, (int)m->E, m->x.as<double>(), m->y.as<double>(), m->z.as<double>(),
                                 m->dconn.as<int32_t>(), m->dne_ptr.as<int32_t>(), m->dne_list.as<int32_t>(), ntol,
                                 gnrm, egroup, m->nx.as<double>(), m->ny.as<double>(), m->nz.as<double>(),
                                 m->valid.as<uint8_t>(), c->stream));
    if (normals_out) {
        FS_HIP(c, hipMemcpyAsync(normals_out, m->nx.p, m->N * 8, hipMemcpyDeviceToHost, c->stream));
        FS_HIP(c, hipMemcpyAsync(normals_out + m->N, m->ny.p, m->N * 8, hipMemcpyDeviceToHost, c->stream));
        FS_HIP(c, hipMemcpyAsync(normals_out + 2 * m->N, m->nz.p, m->N * 8, hipMemcpyDeviceToHost, c->stream));
    }
    if (valid_out) FS_HIP(c, hipMemcpyAsync(valid_out, m->valid.p, m->N, hipMemcpyDeviceToHost, c->stream));
    FS_HIP(c, hipStreamSynchronize(c->stream));
    m->normals_ready = true;
    return FS_OK;
}

int fs_model_normals_isoparametric(fs_model *m, double threshold_angle_deg, double *normals_out, uint8_t *valid_out) {
    if (!m) return FS_ERR_ARG;
    fs_ctx *c = m->ctx;
    c->err.clear();
    FS_HIP(c, hipSetDevice(c->device));
    return normals_common(m, threshold_angle_deg, nullptr, nullptr, normals_out, valid_out);
}

int fs_model_normals_layered(fs_model *m, double threshold_angle_deg, int32_t ngroups, const double *group_normal,
                             const int32_t *elem_group, double *normals_out, uint8_t *valid_out) {
    if (!m) return FS_ERR_ARG;
    fs_ctx *c = m->ctx;
    c->err.clear();
    if (ngroups < 1 || !group_normal) return fail(c, FS_ERR_ARG, "normals: layered form needs ngroups >= 1 normals");
    if (elem_group)
        for (int64_t e = 0; e < m->E; ++e)
            if (elem_group[e] < 0 || elem_group[e] >= ngroups)
                return fail(c, FS_ERR_ARG, "normals: element %lld has group %d", (long long)e, elem_group[e]);
    FS_HIP(c, hipSetDevice(c->device));
    DevBuf dg, de;
    int s;
    if ((s = upload(c, dg, group_normal, (size_t)ngroups * 3))) return s;
    if (elem_group && (s = upload(c, de, elem_group, (size_t)m->E))) return s;
    return normals_common(m, threshold_angle_deg, dg.as<double>(), elem_group ? de.as<int32_t>() : nullptr,
                          normals_out, valid_out);
}

int fs_model_normals_upload(fs_model *m, const double *normals, const uint8_t *valid) {
    if (!m) return FS_ERR_ARG;
    fs_ctx *c = m->ctx;
    c->err.clear();
    if (!normals || !valid) return fail(c, FS_ERR_ARG, "normals_upload: NULL arrays");
    FS_HIP(c, hipSetDevice(c->device));
    FS_HIP(c, hipMemcpyAsync(m->nx.p, normals, m->N * 8, hipMemcpyHostToDevice, c->stream));
    FS_HIP(c, hipMemcpyAsync(m->ny.p, normals + m->N, m->N * 8, hipMemcpyHostToDevice, c->stream));
    FS_HIP(c, hipMemcpyAsync(m->nz.p, normals + 2 * m->N, m->N * 8, hipMemcpyHostToDevice, c->stream));
    FS_HIP(c, hipMemcpyAsync(m->valid.p, valid, m->N, hipMemcpyHostToDevice, c->stream));
    FS_HIP(c, hipStreamSynchronize(c->stream));
    m->normals_ready = true;
    return FS_OK;
}

int fs_model_pattern(fs_model *m, int block, int64_t *nrows, int64_t *nnz) {
    if (!m) return FS_ERR_ARG;
    fs_ctx *c = m->ctx;
    c->err.clear();
    if (block != FS_BLOCK_ALL && block != FS_BLOCK_FF) return fail(c, FS_ERR_ARG, "unknown block %d", block);
    FS_HIP(c, hipSetDevice(c->device));
    int s = build_block(m, block);
    if (s) return s;
    if (nrows) *nrows = m->blk[block].nrows;
    if (nnz) *nnz = m->blk[block].nnz;
    return FS_OK;
}

int fs_model_pattern_get(fs_model *m, int block, int64_t *rowptr, int64_t *colval, int index_base) {
    if (!m) return FS_ERR_ARG;
    int s = fs_model_pattern(m, block, nullptr, nullptr);
    if (s) return s;
    const Block &B = m->blk[block];
    if (rowptr)
        parallel_chunks(B.nrows + 1, [&](int64_t i0, int64_t i1) {
            for (int64_t i = i0; i < i1; ++i) rowptr[i] = B.rowptr[i] + index_base;
        });
    if (colval) {
        // every row of node I holds the same columns: the runs of I's
        // neighbours, each at its offset within the row
        const int64_t nrows = B.nrows;
        parallel_chunks(m->N, [&](int64_t I0, int64_t I1) {
            std::vector<int64_t> cols;
            for (int64_t I = I0; I < I1; ++I) {
                int64_t r0 = -1;
                for (int cc = 0; cc < 6 && r0 < 0; ++cc)
                    if (m->dof[I * 6 + cc] < nrows) r0 = m->dof[I * 6 + cc];
                if (r0 < 0) continue;  // no row of this node in the block
                cols.assign(B.rowptr[r0 + 1] - B.rowptr[r0], 0);
                for (int k = m->nbr_ptr[I]; k < m->nbr_ptr[I + 1]; ++k) {
                    const int J = m->nbr_list[k];
                    for (int cc = 0; cc < 6; ++cc) {
                        const int cd = B.code[(int64_t)J * 6 + cc];
                        if (cd < 0) continue;
                        const int pos = cd < 8 ? B.off_free[k] + cd : B.off_fix[k] + cd - 8;
                        cols[pos] = m->dof[(int64_t)J * 6 + cc] + index_base;
                    }
                }
                for (int cc = 0; cc < 6; ++cc) {
                    const int64_t d = m->dof[I * 6 + cc];
                    if (d < nrows) std::memcpy(colval + B.rowptr[d], cols.data(), cols.size() * sizeof(int64_t));
                }
            }
        });
    }
    return FS_OK;
}

int fs_model_element_blocks(fs_model *m, const fs_shell_params *p, double *blocks_out) {
    if (!m) return FS_ERR_ARG;
    fs_ctx *c = m->ctx;
    c->err.clear();
    FS_HIP(c, hipSetDevice(c->device));
    float ms = 0.f;
    int s = element_blocks(m, p, &ms);
    if (s) return s;
    m->timing[0] = ms;
    if (blocks_out) {
        const int nblk = m->nen * (m->nen + 1) / 2;
        return d2h(c, blocks_out, m->Kb.p, (size_t)m->E * nblk * 36 * 8);
    }
    return FS_OK;
}

int fs_model_stiffness(fs_model *m, const fs_shell_params *p, int block, double *nzval_out) {
    if (!m) return FS_ERR_ARG;
    fs_ctx *c = m->ctx;
    c->err.clear();
    if (block != FS_BLOCK_ALL && block != FS_BLOCK_FF) return fail(c, FS_ERR_ARG, "unknown block %d", block);
    FS_HIP(c, hipSetDevice(c->device));
    int s = build_block(m, block);
    if (s) return s;
    Block &B = m->blk[block];
    fs::ElemParams ep;
    int variant = 0;
    if (c->assembly == 1) {
        // fused element tiles: one element kernel writes the CSR values of the
        // pairs its tile completes and the partial sums of the others; the
        // finalize kernel adds the partials of the boundary pairs
        const int NE = fs::fused_tile_elems(m->nen);
        if (B.tiles.ne != NE && (s = build_tiles(m, block, NE))) return s;
        if ((s = prepare_elements(m, p, ep, &variant, false))) return s;
        ep.ebegin = 0;
        ep.eend = (int32_t)(B.tiles.ntiles * NE);
        ep.tile_elem = B.tiles.tile_elem.as<int32_t>();
        ep.tile_conn = B.tiles.tile_conn.as<int32_t>();
        const fs::ScatterPlan sp = plan_of(B);
        const fs::TilePlan tp = tileplan_of(B);
        hipEvent_t t0, t1, t2;
        FS_HIP(c, hipEventCreate(&t0));
        FS_HIP(c, hipEventCreate(&t1));
        FS_HIP(c, hipEventCreate(&t2));
        FS_HIP(c, hipEventRecord(t0, c->stream));
        FS_HIP(c, fs::launch_element_fused(ep, m->x.as<double>(), m->y.as<double>(), m->z.as<double>(),
                                           m->dconn.as<int32_t>(), m->nx.as<double>(), m->ny.as<double>(),
                                           m->nz.as<double>(), m->valid.as<uint8_t>(), m->status.as<int>(), sp, tp,
                                           B.kval.as<double>(), c->stream));
        FS_HIP(c, hipEventRecord(t1, c->stream));
        FS_HIP(c, fs::launch_pair_finalize(sp, tp, B.kval.as<double>(), c->stream));
        FS_HIP(c, hipEventRecord(t2, c->stream));
        int st = 0;
        FS_HIP(c, hipMemcpyAsync(&st, m->status.p, sizeof(int), hipMemcpyDeviceToHost, c->stream));
        FS_HIP(c, hipStreamSynchronize(c->stream));
        float ms_elem = 0.f, ms_all = 0.f;
        FS_HIP(c, hipEventElapsedTime(&ms_elem, t0, t1));
        FS_HIP(c, hipEventElapsedTime(&ms_all, t0, t2));
        (void)hipEventDestroy(t0);
        (void)hipEventDestroy(t1);
        (void)hipEventDestroy(t2);
        ep.tile_elem = nullptr;
        ep.tile_conn = nullptr;
        ep.eend = (int32_t)m->E;
        if ((s = finish_elements(m, ep, st))) return s;
        m->timing[0] = ms_elem;
        m->timing[1] = ms_all - ms_elem;
        m->timing[2] = ms_all;
        B.k_ready = true;
        if (nzval_out) return d2h(c, nzval_out, B.kval.p, (size_t)B.nnz * 8);
        return FS_OK;
    }
    if ((s = prepare_elements(m, p, ep, &variant))) return s;
    // Element ranges of whole workgroups.  After range k the pairs whose
    // contributing elements all lie in ranges 0..k are complete (pair_emax),
    // so their scatter is queued on stream2 behind an event and overlaps the
    // element kernels of the next ranges.  Each pair is still summed by one
    // thread in ascending element order: the values are bitwise those of the
    // unpipelined sequence.
    int ne = 1, wg = 0;
    fs::element_variant_shape(m->nen, variant, &ne, &wg);
    const int64_t E = m->E;
    const int nch = std::max(1, c->asm_chunks);
    int64_t per = (E + nch - 1) / nch;
    per = std::max<int64_t>(ne, (per + ne - 1) / ne * ne);
    const int64_t nranges = E == 0 ? 0 : (E + per - 1) / per;
    while ((int64_t)c->ev.size() < nranges + 1) {
        hipEvent_t ev;
        FS_HIP(c, hipEventCreateWithFlags(&ev, hipEventDisableTiming));
        c->ev.push_back(ev);
    }
    if (!c->stream2) FS_HIP(c, hipStreamCreateWithFlags(&c->stream2, hipStreamNonBlocking));
    const bool piped = nranges > 1;
    hipEvent_t t0, t1, t2;
    FS_HIP(c, hipEventCreate(&t0));
    FS_HIP(c, hipEventCreate(&t1));
    FS_HIP(c, hipEventCreate(&t2));
    FS_HIP(c, hipEventRecord(t0, c->stream));
    fs::ScatterPlan sp = plan_of(B);
    int64_t pdone = 0;
    for (int64_t k = 0; k < nranges; ++k) {
        ep.ebegin = (int32_t)(k * per);
        ep.eend = (int32_t)std::min(E, (k + 1) * per);
        FS_HIP(c, launch_elements(m, ep, variant, c->stream));
        if (!piped) break;
        int64_t pend = B.npairs;
        if (ep.eend < E)
            pend = std::upper_bound(B.pair_emax.begin(), B.pair_emax.end(), ep.eend - 1) - B.pair_emax.begin();
        if (pend > pdone) {
            FS_HIP(c, hipEventRecord(c->ev[k], c->stream));
            FS_HIP(c, hipStreamWaitEvent(c->stream2, c->ev[k], 0));
            sp.pbegin = pdone;
            sp.pend = pend;
            FS_HIP(c, fs::launch_scatter(sp, m->Kb.as<double>(), B.kval.as<double>(), c->stream2));
            pdone = pend;
        }
    }
    FS_HIP(c, hipEventRecord(t1, c->stream));
    if (piped) {
        FS_HIP(c, hipEventRecord(c->ev[nranges], c->stream2));
        FS_HIP(c, hipStreamWaitEvent(c->stream, c->ev[nranges], 0));
    } else {
        FS_HIP(c, fs::launch_scatter(plan_of(B), m->Kb.as<double>(), B.kval.as<double>(), c->stream));
    }
    FS_HIP(c, hipEventRecord(t2, c->stream));
    int st = 0;
    FS_HIP(c, hipMemcpyAsync(&st, m->status.p, sizeof(int), hipMemcpyDeviceToHost, c->stream));
    FS_HIP(c, hipStreamSynchronize(c->stream));
    float ms_elem = 0.f, ms_all = 0.f;
    FS_HIP(c, hipEventElapsedTime(&ms_elem, t0, t1));
    FS_HIP(c, hipEventElapsedTime(&ms_all, t0, t2));
    (void)hipEventDestroy(t0);
    (void)hipEventDestroy(t1);
    (void)hipEventDestroy(t2);
    if ((s = finish_elements(m, ep, st))) return s;
    m->last_ep.ebegin = 0;
    m->last_ep.eend = (int32_t)E;
    m->timing[0] = ms_elem;
    m->timing[1] = ms_all - ms_elem;
    m->timing[2] = ms_all;
    B.k_ready = true;
    if (nzval_out) return d2h(c, nzval_out, B.kval.p, (size_t)B.nnz * 8);
    return FS_OK;
}

int fs_model_mass(fs_model *m, const fs_mass_params *p, int block, double *diag_out) {
    if (!m || !p) return FS_ERR_ARG;
    fs_ctx *c = m->ctx;
    c->err.clear();
    if (block != FS_BLOCK_ALL && block != FS_BLOCK_FF) return fail(c, FS_ERR_ARG, "unknown block %d", block);
    FS_HIP(c, hipSetDevice(c->device));
    int s = build_block(m, block);
    if (s) return s;
    Block &B = m->blk[block];
    const double *ginert = nullptr;
    const int32_t *ggrp = nullptr;
    const double *grt = nullptr;
    if (!p->composite) {
        const bool grouped = p->group_rho_t != nullptr;
        const int ng = grouped ? p->ngroups : 1;
        if (grouped && ng < 1) return fail(c, FS_ERR_ARG, "mass: group_rho_t needs ngroups >= 1");
        std::vector<double> rt(2 * (size_t)ng);
        for (int g = 0; g < ng; ++g) {
            rt[2 * g] = grouped ? p->group_rho_t[2 * g] : p->rho;
            rt[2 * g + 1] = grouped ? p->group_rho_t[2 * g + 1] : p->thickness;
        }
        if ((s = upload(c, m->g_rho_t, rt.data(), rt.size()))) return s;
        grt = m->g_rho_t.as<double>();
        if (grouped && p->elem_group) {
            for (int64_t e = 0; e < m->E; ++e)
                if (p->elem_group[e] < 0 || p->elem_group[e] >= ng)
                    return fail(c, FS_ERR_ARG, "mass: element %lld has group %d", (long long)e, p->elem_group[e]);
            if ((s = upload(c, m->m_group, p->elem_group, (size_t)m->E))) return s;
            ggrp = m->m_group.as<int32_t>();
        }
    } else {
        if (p->ngroups < 1 || !p->group_inertia) return fail(c, FS_ERR_ARG, "mass: layered shell needs group inertia");
        if ((s = upload(c, m->g_inertia, p->group_inertia, (size_t)2 * p->ngroups))) return s;
        ginert = m->g_inertia.as<double>();
        if (p->elem_group) {
            for (int64_t e = 0; e < m->E; ++e)
                if (p->elem_group[e] < 0 || p->elem_group[e] >= p->ngroups)
                    return fail(c, FS_ERR_ARG, "mass: element %lld has group %d", (long long)e, p->elem_group[e]);
            if ((s = upload(c, m->m_group, p->elem_group, (size_t)m->E))) return s;
            ggrp = m->m_group.as<int32_t>();
        }
    }
    FS_HIP(c, fs::launch_mass(m->nen, m->N, (int)m->E, p->composite ? 1 : 0, grt, ginert, ggrp,
                              m->x.as<double>(), m->y.as<double>(), m->z.as<double>(), m->dconn.as<int32_t>(),
                              m->dne_ptr.as<int32_t>(), m->dne_list.as<int32_t>(), m->node_dof.as<int64_t>(),
                              B.nrows, B.mdiag.as<double>(), c->stream));
    if (diag_out)
        FS_HIP(c, hipMemcpyAsync(diag_out, B.mdiag.p, (size_t)B.nrows * 8, hipMemcpyDeviceToHost, c->stream));
    FS_HIP(c, hipStreamSynchronize(c->stream));
    B.m_ready = true;
    return FS_OK;
}

int fs_model_resultants(fs_model *m, const fs_shell_params *p, const double *u, int64_t nlist,
                        const int32_t *felist, int32_t ocs_kind, const double *ocs, int32_t flags, double *out,
                        int32_t *mismatch) {
    if (!m) return FS_ERR_ARG;
    fs_ctx *c = m->ctx;
    c->err.clear();
    if (!p || !u || !out || nlist < 0) return fail(c, FS_ERR_ARG, "resultants: bad arguments");
    if (!m->normals_ready) return fail(c, FS_ERR_STATE, "resultants: normals not computed (associategeometry!)");
    if (ocs_kind < FS_OCSYS_ISOPARAMETRIC || ocs_kind > FS_OCSYS_PER_POINT || (ocs_kind != 0 && !ocs))
        return fail(c, FS_ERR_ARG, "resultants: output csys kind %d", ocs_kind);
    if (flags & ~FS_RES_LAYERED_Q4_SINGLE_T) return fail(c, FS_ERR_ARG, "resultants: unknown flags %d", flags);
    if (nlist > (int64_t)INT32_MAX / 4) return fail(c, FS_ERR_UNSUPPORTED, "resultants: %lld elements", (long long)nlist);
    const int npts = m->nen == 3 ? 1 : 4;
    if (felist)
        for (int64_t l = 0; l < nlist; ++l)
            if (felist[l] < 0 || felist[l] >= m->E)
                return fail(c, FS_ERR_ARG, "resultants: element %d out of range", felist[l]);
    if (!felist && nlist > m->E) return fail(c, FS_ERR_ARG, "resultants: nlist exceeds the element count");
    FS_HIP(c, hipSetDevice(c->device));
    fs::ResParams rp{};
    rp.nlist = (int32_t)nlist;
    rp.nelem = (int32_t)m->E;
    rp.ocs_kind = ocs_kind;
    rp.twice = (p->composite && m->nen == 4 && !(flags & FS_RES_LAYERED_Q4_SINGLE_T)) ? 1 : 0;
    rp.stab_c = p->stab_c;
    int s = 0;
    DevBuf dmat, dgrp, du, dlist, docs, dout, dst;
    std::vector<double> rec;
    int ng = 1;
    const int32_t *egroup = nullptr;
    if (!p->composite) {
        // per group: Dps (row-major), 5/6 Dt, thickness, stab_c
        const bool grouped = p->group_homog != nullptr;
        ng = grouped ? p->ngroups : 1;
        if (ng < 1) return fail(c, FS_ERR_ARG, "resultants: group_homog needs ngroups >= 1");
        rec.assign(16 * (size_t)ng, 0.0);
        for (int g = 0; g < ng; ++g) {
            const double *src = grouped ? p->group_homog + 16 * g : nullptr;
            double *r = &rec[16 * g];
            cm2rm(3, grouped ? src : p->Dps, r);
            cm2rm(2, grouped ? src + 9 : p->Dt, r + 9);
            for (int i = 0; i < 4; ++i) r[9 + i] *= 5.0 / 6.0;  // FEMMShellT3FFModule.jl:875-876, Q4RS :1091-1092
            r[13] = grouped ? src[13] : p->thickness;
            r[14] = grouped ? src[14] : p->stab_c;
            if (!(r[13] > 0.0)) return fail(c, FS_ERR_ARG, "resultants: thickness of group %d must be positive", g);
        }
        if (grouped) egroup = p->elem_group;
    } else {
        ng = p->ngroups;
        if (ng < 1 || !p->group_ABDH || !p->group_csys || !p->group_thickness)
            return fail(c, FS_ERR_ARG, "resultants: layered shell needs ngroups >= 1 and group arrays");
        rec.assign(41 * (size_t)ng, 0.0);
        for (int g = 0; g < ng; ++g) {
            const double *src = p->group_ABDH + 31 * g;
            double *r = &rec[41 * g];
            cm2rm(3, src, r);
            cm2rm(3, src + 9, r + 9);
            cm2rm(3, src + 18, r + 18);
            cm2rm(2, src + 27, r + 27);
            cm2rm(3, p->group_csys + 9 * g, r + 31);
            r[40] = p->group_thickness[g];
        }
        egroup = p->elem_group;
    }
    if (egroup) {
        for (int64_t e = 0; e < m->E; ++e)
            if (egroup[e] < 0 || egroup[e] >= ng)
                return fail(c, FS_ERR_ARG, "resultants: element %lld has group %d", (long long)e, egroup[e]);
        if ((s = upload(c, dgrp, egroup, (size_t)m->E))) return s;
        rp.elem_group = dgrp.as<int32_t>();
    }
    if ((s = upload(c, dmat, rec.data(), rec.size()))) return s;
    rp.g_mat = dmat.as<double>();
    if ((s = upload(c, du, u, (size_t)m->N * 6))) return s;
    rp.u = du.as<double>();
    if (felist) {
        if ((s = upload(c, dlist, felist, (size_t)nlist))) return s;
        rp.felist = dlist.as<int32_t>();
    }
    if (ocs_kind != FS_OCSYS_ISOPARAMETRIC) {
        const size_t nm = ocs_kind == FS_OCSYS_MATRIX ? 1 : (size_t)nlist * npts;
        std::vector<double> rm(9 * nm);
        for (size_t k = 0; k < nm; ++k) cm2rm(3, ocs + 9 * k, &rm[9 * k]);
        if ((s = upload(c, docs, rm.data(), rm.size()))) return s;
        rp.ocs = docs.as<double>();
    }
    if ((s = dalloc(c, dout, (size_t)nlist * npts * 8 * sizeof(double)))) return s;
    if ((s = dalloc(c, dst, 2 * sizeof(int)))) return s;
    rp.out = dout.as<double>();
    rp.status = dst.as<int>();
    FS_HIP(c, hipMemsetAsync(dst.p, 0, 2 * sizeof(int), c->stream));
    hipEvent_t e0, e1;
    FS_HIP(c, hipEventCreate(&e0));
    FS_HIP(c, hipEventCreate(&e1));
    FS_HIP(c, hipEventRecord(e0, c->stream));
    FS_HIP(c, fs::launch_resultants(m->nen, p->composite ? 1 : 0, rp, m->x.as<double>(), m->y.as<double>(),
                                    m->z.as<double>(), m->dconn.as<int32_t>(), m->nx.as<double>(),
                                    m->ny.as<double>(), m->nz.as<double>(), m->valid.as<uint8_t>(), c->stream));
    FS_HIP(c, hipEventRecord(e1, c->stream));
    int st[2] = {0, 0};
    FS_HIP(c, hipMemcpyAsync(st, dst.p, sizeof st, hipMemcpyDeviceToHost, c->stream));
    FS_HIP(c, hipMemcpyAsync(out, dout.p, (size_t)nlist * npts * 8 * sizeof(double), hipMemcpyDeviceToHost,
                             c->stream));
    FS_HIP(c, hipStreamSynchronize(c->stream));
    float ms = 0.f;
    FS_HIP(c, hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    m->timing[5] = ms;
    if (st[0]) return fail(c, FS_ERR_NUMERIC, "resultants: singular metric matrix in _gradN_e!");
    if (mismatch) *mismatch = st[1] ? 1 : 0;
    return FS_OK;
}

int fs_model_explicit(fs_model *m, const fs_cd_params *p, const double *U0, const double *V0, const double *Fmag,
                      const double *g, const int64_t *probe_dofs, int32_t nprobes, double *probe_out, double *U_out,
                      double *V_out) {
    if (!m || !p) return FS_ERR_ARG;
    fs_ctx *c = m->ctx;
    c->err.clear();
    Block &B = m->blk[FS_BLOCK_FF];
    if (!B.k_ready || !B.m_ready)
        return fail(c, FS_ERR_STATE, "explicit: FF-block stiffness and mass must be computed first");
    if (p->precision != FS_PREC_FP64 && p->precision != FS_PREC_MIXED && p->precision != FS_PREC_MIXED_F32)
        return fail(c, FS_ERR_UNSUPPORTED, "explicit: precision %d", p->precision);
    const bool mixed = p->precision != FS_PREC_FP64;
    const bool batch = p->ngroups > 0;
    if (!U0 || !V0 || !Fmag || !g || p->nsteps < 0 || (!batch && !(p->dt > 0.0)))
        return fail(c, FS_ERR_ARG, "explicit: bad arguments");
    if (batch && (!p->row_group || !p->group_dt || !p->group_cdamp))
        return fail(c, FS_ERR_ARG, "explicit: a batch needs row_group, group_dt and group_cdamp");
    const int stride = p->probe_stride > 0 ? p->probe_stride : 1;
    if (nprobes < 0 || (nprobes > 0 && (!probe_dofs || !probe_out))) return fail(c, FS_ERR_ARG, "explicit: probes");
    const int64_t n = B.nrows;
    for (int k = 0; k < nprobes; ++k)
        if (probe_dofs[k] < 0 || probe_dofs[k] >= n) return fail(c, FS_ERR_ARG, "explicit: probe dof out of range");
    const int ng = batch ? p->ngroups : 1;
    if (batch) {
        for (int64_t i = 0; i < n; ++i)
            if (p->row_group[i] < 0 || p->row_group[i] >= ng) return fail(c, FS_ERR_ARG, "explicit: row_group");
        for (int gi = 0; gi < ng; ++gi)
            if (!(p->group_dt[gi] > 0.0)) return fail(c, FS_ERR_ARG, "explicit: group_dt must be positive");
    }
    FS_HIP(c, hipSetDevice(c->device));
    int s;
    double build_ms = 0.0;
    if (mixed) {
        if (m->N > (int64_t)INT32_MAX / 6) return fail(c, FS_ERR_UNSUPPORTED, "explicit: model too large for mixed");
        const bool shear_fp32 = p->precision == FS_PREC_MIXED_F32;
        if (!m->mixed.built || m->mixed.kgen != m->kgen || m->mixed.shear_fp32 != shear_fp32) {
            if ((s = build_mixed(m, shear_fp32))) return s;
            build_ms = m->mixed.build_ms;
        }
    }
    // State vectors: the free dofs themselves (FP64 CSR path), or node-major
    // dof positions 6 i + c (mixed path; fixed positions stay zero).
    const int64_t nv = mixed ? 6 * m->N : n;
    std::vector<int64_t> pos(n);
    if (mixed) {
        for (int64_t I = 0; I < m->N; ++I)
            for (int cc = 0; cc < 6; ++cc) {
                const int64_t d = m->dof[I * 6 + cc];
                if (d < n) pos[d] = 6 * I + cc;
            }
    } else {
        for (int64_t d = 0; d < n; ++d) pos[d] = d;
    }
    std::vector<double> M(n), U0v(nv, 0.0), V0v(nv, 0.0), Fv(nv, 0.0), Cd(nv, 0.0), invMC(nv, 0.0);
    std::vector<int32_t> rgv(batch ? nv : 0, 0);
    if ((s = d2h(c, M.data(), B.mdiag.p, n * 8))) return s;
    for (int64_t i = 0; i < n; ++i) {
        const double dt = batch ? p->group_dt[p->row_group[i]] : (p->dt_mass > 0.0 ? p->dt_mass : p->dt);
        const double cd = batch ? p->group_cdamp[p->row_group[i]] : p->cdamp;
        const int64_t q = pos[i];
        Cd[q] = cd * M[i];
        invMC[q] = 1.0 / (M[i] + (dt / 2) * Cd[q]);  // plate_expl_examples.jl:68-71, before the dt adjustment
        U0v[q] = U0[i];
        V0v[q] = V0[i];
        Fv[q] = Fmag[i];
        if (batch) rgv[q] = p->row_group[i];
    }
    bool damped = false;
    for (int gi = 0; gi < ng; ++gi) damped |= (batch ? p->group_cdamp[gi] : p->cdamp) != 0.0;
    StepCache &sc = m->step;
    DevBuf(&dU)[2] = sc.dU;
    DevBuf &dV = sc.dV, &dA = sc.dA, &dF = sc.dF, &dInv = sc.dInv, &dC = sc.dC, &dSlot = sc.dSlot,
           &dProbe = sc.dProbe, &dRG = sc.dRG, &dGdt = sc.dGdt, &dG = sc.dG;
    if ((s = put(c, dU[0], U0v.data(), nv)) || (s = ensure(c, dU[1], nv * 8)) || (s = put(c, dV, V0v.data(), nv)) ||
        (s = ensure(c, dA, nv * 8)) || (s = put(c, dF, Fv.data(), nv)) || (s = put(c, dInv, invMC.data(), nv)) ||
        (s = put(c, dC, Cd.data(), nv))) {
        sc.drop_graph();
        return s;
    }
    if (batch) {
        if ((s = put(c, dRG, rgv.data(), nv)) || (s = put(c, dGdt, p->group_dt, ng)) ||
            (s = put(c, dG, g, (size_t)(p->nsteps + 1) * ng))) {
            sc.drop_graph();
            return s;
        }
    }
    const int64_t nrec = p->nsteps / stride + 1;
    if (nprobes > 0) {
        std::vector<int32_t> slot(nv, -1);
        for (int k = 0; k < nprobes; ++k) slot[pos[probe_dofs[k]]] = k;
        if ((s = put(c, dSlot, slot.data(), nv)) || (s = ensure(c, dProbe, (size_t)nrec * nprobes * 8))) {
            sc.drop_graph();
            return s;
        }
        for (int k = 0; k < nprobes; ++k) probe_out[k] = U0[probe_dofs[k]];  // peek(0, U0)
    }
    // steps whose load factor is zero for every model skip the Fmag read (mixed path)
    std::vector<uint8_t> forced(p->nsteps + 1, 0);
    for (int64_t step = 0; step <= p->nsteps; ++step)
        for (int gi = 0; gi < ng && !forced[step]; ++gi) forced[step] = g[step * ng + gi] != 0.0;
    // rows of the shell stencils hold 7 (T3) to 9 (Q4) neighbour runs: 4 lanes
    // per row, up to three runs per lane loaded together (measured fastest
    // of 2, 4, 8 and 16 lanes per row on both families)
    const int tpr = 4;
    const int cd_pj = node_pj(c, B);
    const int32_t *rg = batch ? dRG.as<int32_t>() : nullptr;
    const double *gdt = batch ? dGdt.as<double>() : nullptr;
    const fs::MixedOp op = mixed ? mixed_op(m) : fs::MixedOp{};
    auto enqueue = [&](hipStream_t st) -> hipError_t {
        hipError_t e = fs::launch_cd_init(nv, p->dt, batch ? 0.0 : g[0], dInv.as<double>(), dF.as<double>(),
                                          dU[0].as<double>(), dV.as<double>(), dA.as<double>(), dU[1].as<double>(),
                                          rg, gdt, batch ? dG.as<double>() : nullptr, st);
        if (e != hipSuccess) return e;
        int cur = 1;
        for (int64_t step = 1; step <= p->nsteps; ++step) {
            const bool rec = nprobes > 0 && (step % stride == 0);
            int32_t *slot = rec ? dSlot.as<int32_t>() : nullptr;
            double *pdst = rec ? dProbe.as<double>() + (step / stride) * nprobes : nullptr;
            const double *gt = batch ? dG.as<double>() + step * ng : nullptr;
            if (mixed)
                e = fs::launch_cd_mixed_step(m->nen, m->mixed.nsr, op, dU[cur].as<double>(),
                                             dU[1 - cur].as<double>(), dV.as<double>(), dA.as<double>(),
                                             dInv.as<double>(), dC.as<double>(), dF.as<double>(),
                                             batch ? 0.0 : g[step], p->dt, damped ? 1 : 0, forced[step] ? 1 : 0,
                                             slot, pdst, rg, gdt, gt, st);
            else if (c->cd_lanes > 0)
                e = fs::launch_cd_node_step(c->cd_lanes, cd_pj, noderows_of(B), B.slotrec.as<int2>(), B.kval.as<double>(),
                                            dU[cur].as<double>(), dU[1 - cur].as<double>(), dV.as<double>(),
                                            dA.as<double>(), dInv.as<double>(), dC.as<double>(), dF.as<double>(),
                                            batch ? 0.0 : g[step], p->dt, damped ? 1 : 0, slot, pdst, rg, gdt, gt, st);
            else
                e = fs::launch_cd_step(tpr, runrows_of(B), B.kval.as<double>(), dU[cur].as<double>(), dU[1 - cur].as<double>(),
                                       dV.as<double>(), dA.as<double>(), dInv.as<double>(), dC.as<double>(),
                                       dF.as<double>(), batch ? 0.0 : g[step], p->dt, damped ? 1 : 0, slot, pdst, rg,
                                       gdt, gt, st);
            if (e != hipSuccess) return e;
            cur = 1 - cur;
        }
        return hipSuccess;
    };
    hipEvent_t e0, e1;
    FS_HIP(c, hipEventCreate(&e0));
    FS_HIP(c, hipEventCreate(&e1));
    m->timing[6] = 0.0;
    if (p->use_graph) {
        // everything the captured launches take as arguments
        std::vector<uint64_t> key;
        key.reserve(64 + (size_t)p->nsteps + 1);
        for (const int64_t v : {nv, n, (int64_t)p->nsteps, (int64_t)p->precision, (int64_t)stride, (int64_t)nprobes,
                                (int64_t)ng, (int64_t)batch, (int64_t)damped, (int64_t)c->cd_lanes, (int64_t)cd_pj,
                                (int64_t)tpr, (int64_t)m->nen, (int64_t)m->mixed.nsr})
            key_add(key, v);
        key_add(key, p->dt);
        for (const DevBuf *b : {&dU[0], &dU[1], &dV, &dA, &dF, &dInv, &dC, &dSlot, &dProbe, &dRG, &dGdt, &dG})
            key_add(key, b->p);
        key_add(key, B.kval.p);
        key_add(key, B.slotrec.p);
        if (mixed) key_add(key, op);
        else if (c->cd_lanes > 0) key_add(key, noderows_of(B));
        else key_add(key, runrows_of(B));
        for (int64_t step = 0; step <= p->nsteps; ++step) {
            key.push_back(forced[step]);
            if (!batch) key_add(key, g[step]);
        }
        if (!sc.exec || key != sc.key) {
            const auto tc0 = std::chrono::steady_clock::now();
            sc.drop_graph();
            FS_HIP(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
            hipError_t ce = enqueue(c->stream);
            hipError_t ee = hipStreamEndCapture(c->stream, &sc.graph);
            if (ce != hipSuccess || ee != hipSuccess) {
                sc.drop_graph();
                return fail(c, FS_ERR_DEVICE, "explicit: graph capture failed: %s",
                            hipGetErrorString(ce != hipSuccess ? ce : ee));
            }
            hipError_t ie = hipGraphInstantiate(&sc.exec, sc.graph, nullptr, nullptr, 0);
            if (ie != hipSuccess) {
                sc.exec = nullptr;
                sc.drop_graph();
                return fail(c, FS_ERR_DEVICE, "explicit: graph instantiation failed: %s", hipGetErrorString(ie));
            }
            sc.key.swap(key);
            m->timing[6] = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - tc0).count();
        } else {
            ++sc.replays;
        }
        FS_HIP(c, hipEventRecord(e0, c->stream));
        FS_HIP(c, hipGraphLaunch(sc.exec, c->stream));
        FS_HIP(c, hipEventRecord(e1, c->stream));
        FS_HIP(c, hipStreamSynchronize(c->stream));
    } else {
        FS_HIP(c, hipEventRecord(e0, c->stream));
        FS_HIP(c, enqueue(c->stream));
        FS_HIP(c, hipEventRecord(e1, c->stream));
        FS_HIP(c, hipStreamSynchronize(c->stream));
    }
    float ms = 0.f;
    FS_HIP(c, hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    m->timing[3] = ms;
    m->timing[4] = build_ms;
    // after the last step the buffer not indexed by the loop's final `cur` holds U_nsteps
    const int fin = (p->nsteps % 2 == 1) ? 1 : 0;
    if (U_out || V_out) {
        std::vector<double> Uv(nv), Vv(nv);
        if ((s = d2h(c, Uv.data(), dU[p->nsteps == 0 ? 0 : fin].p, nv * 8)) || (s = d2h(c, Vv.data(), dV.p, nv * 8)))
            return s;
        for (int64_t i = 0; i < n; ++i) {
            if (U_out) U_out[i] = Uv[pos[i]];
            if (V_out) V_out[i] = Vv[pos[i]];
        }
    }
    if (nprobes > 0 && nrec > 1 &&
        (s = d2h(c, probe_out + nprobes, dProbe.as<double>() + nprobes, (size_t)(nrec - 1) * nprobes * 8)))
        return s;
    return FS_OK;
}

}  // extern "C"

namespace {
// Power iteration of pwr (clamp_cyl_transient_examples.jl:45-60) on the
// resident FF-block K and M, from v (device, overwritten): w, inv (n each),
// part (512) and ss (4) are device work arrays.
int power_iteration(fs_model *m, double *v, double *w, double *inv, double *part, double *ss, int32_t maxit,
                    double *omega_max) {
    fs_ctx *c = m->ctx;
    Block &B = m->blk[FS_BLOCK_FF];
    const int64_t n = B.nrows;
    const fs::RunRows rows = runrows_of(B);
    const fs::NodeRows nrw = noderows_of(B);
    const double *kv = B.kval.as<double>();
    const int pj = node_pj(c, B);
    // w = K v in the step's form (node-grouped unless the row form is set)
    auto spmv = [&]() {
        return c->cd_lanes > 0 ? fs::launch_spmv_node(c->cd_lanes, pj, nrw, B.slotrec.as<int2>(), kv, v, w, c->stream)
                               : fs::launch_spmv(rows, kv, v, w, c->stream);
    };
    FS_HIP(c, fs::launch_pwr_init(n, B.mdiag.as<double>(), inv, ss, c->stream));  // invM, scale factors 1
    for (int it = 0; it < maxit; ++it) {
        FS_HIP(c, spmv());                                                            // w = K v
        FS_HIP(c, fs::launch_dot(n, 0, w, nullptr, nullptr, part, ss, c->stream));
        FS_HIP(c, fs::launch_scale(n, ss, nullptr, w, w, c->stream));                 // w /= |w|
        FS_HIP(c, fs::launch_scale(n, ss + 3, inv, w, v, c->stream));                 // v = invM .* w
        FS_HIP(c, fs::launch_dot(n, 0, v, nullptr, nullptr, part, ss + 1, c->stream));
        FS_HIP(c, fs::launch_scale(n, ss + 1, nullptr, v, v, c->stream));             // v /= |v|
    }
    FS_HIP(c, spmv());
    FS_HIP(c, fs::launch_dot(n, 1, w, v, nullptr, part, ss, c->stream));
    FS_HIP(c, fs::launch_dot(n, 2, v, nullptr, B.mdiag.as<double>(), part, ss + 1, c->stream));
    double h[2];
    FS_HIP(c, hipMemcpyAsync(h, ss, 16, hipMemcpyDeviceToHost, c->stream));
    FS_HIP(c, hipStreamSynchronize(c->stream));
    *omega_max = std::sqrt(h[0] / h[1]);
    return FS_OK;
}
}  // namespace

extern "C" {

int fs_model_omega_max(fs_model *m, const double *v0, int32_t maxit, double *omega_max) {
    if (!m || !v0 || !omega_max || maxit < 0) return FS_ERR_ARG;
    fs_ctx *c = m->ctx;
    c->err.clear();
    Block &B = m->blk[FS_BLOCK_FF];
    if (!B.k_ready || !B.m_ready) return fail(c, FS_ERR_STATE, "omega_max: FF-block K and M must be computed first");
    FS_HIP(c, hipSetDevice(c->device));
    const int64_t n = B.nrows;
    DevBuf dv, dw, dinv, dpart, dss;
    int s;
    if ((s = upload(c, dv, v0, n)) || (s = dalloc(c, dw, n * 8)) || (s = dalloc(c, dinv, n * 8)) ||
        (s = dalloc(c, dpart, 512 * 8)) || (s = dalloc(c, dss, 4 * 8)))
        return s;
    return power_iteration(m, dv.as<double>(), dw.as<double>(), dinv.as<double>(), dpart.as<double>(),
                           dss.as<double>(), maxit, omega_max);
}

int fs_model_omega_max_noalloc(fs_model *m, const double *v0, int32_t maxit, double *omega_max) {
    if (!m || !v0 || !omega_max || maxit < 0) return FS_ERR_ARG;
    fs_ctx *c = m->ctx;
    c->err.clear();
    Block &B = m->blk[FS_BLOCK_FF];
    if (!B.k_ready || !B.m_ready) return fail(c, FS_ERR_STATE, "omega_max: FF-block K and M must be computed first");
    const int64_t n = B.nrows;
    if (n > fs::NOALLOC_MAX_ROWS)
        return fail(c, FS_ERR_UNSUPPORTED,
                    "omega_max without allocation: %lld free dofs exceed the workspace's %lld (use fs_model_omega_max)",
                    (long long)n, (long long)fs::NOALLOC_MAX_ROWS);
    if (!c->scratch)
        return fail(c, FS_ERR_DEVICE, "omega_max without allocation: no workspace on device %d (%s)", c->device,
                    hipGetErrorString(c->scratch_status));
    FS_HIP(c, hipSetDevice(c->device));
    double *v = c->scratch, *w = v + fs::NOALLOC_MAX_ROWS, *inv = w + fs::NOALLOC_MAX_ROWS;
    double *part = inv + fs::NOALLOC_MAX_ROWS, *ss = part + 512;
    if (n) FS_HIP(c, hipMemcpyAsync(v, v0, (size_t)n * 8, hipMemcpyHostToDevice, c->stream));
    return power_iteration(m, v, w, inv, part, ss, maxit, omega_max);
}

int fs_ctx_tuning(const fs_ctx *ctx, int nen, int formulation, int32_t *shape) {
    if (!ctx || !shape || (nen != FS_T3 && nen != FS_Q4) ||
        (formulation != FS_SHEAR_AVERAGE_B && formulation != FS_SHEAR_AVERAGE_K))
        return FS_ERR_ARG;
    const int v = ctx->elem_variant[nen == FS_Q4 ? 1 : 0][formulation];
    shape[0] = v;
    int ne = 0, wg = 0;
    if (v >= 0) fs::element_variant_shape(nen, v, &ne, &wg);
    shape[1] = v >= 0 ? ne : 0;
    shape[2] = v >= 0 ? wg : 0;
    return FS_OK;
}

int fs_ctx_set_tuning(fs_ctx *ctx, int nen, int formulation, int32_t variant) {
    if (!ctx || (nen != FS_T3 && nen != FS_Q4) ||
        (formulation != FS_SHEAR_AVERAGE_B && formulation != FS_SHEAR_AVERAGE_K) || variant < -2 ||
        variant >= fs::element_variants(nen))
        return FS_ERR_ARG;
    if (variant == -2) {
        ctx->autotune = false;
        ctx->elem_variant[nen == FS_Q4 ? 1 : 0][formulation] = -1;
        return FS_OK;
    }
    ctx->autotune = true;
    ctx->elem_variant[nen == FS_Q4 ? 1 : 0][formulation] = variant;
    return FS_OK;
}

int fs_ctx_set_pipeline(fs_ctx *ctx, int ranges) {
    if (!ctx || ranges < 1 || ranges > 64) return FS_ERR_ARG;
    ctx->asm_chunks = ranges;
    return FS_OK;
}

int fs_ctx_set_assembly(fs_ctx *ctx, int mode) {
    if (!ctx) return FS_ERR_ARG;
    if (mode != 0 && mode != 1) return fail(ctx, FS_ERR_ARG, "assembly mode %d", mode);
    ctx->assembly = mode;
    return FS_OK;
}

int fs_ctx_set_cd_lanes(fs_ctx *ctx, int lanes_per_node, int positions_per_lane) {
    if (!ctx) return FS_ERR_ARG;
    const int pj = positions_per_lane;
    const bool ok = lanes_per_node == 0 || (lanes_per_node == 8 && (pj == 0 || (pj >= 4 && pj <= 8))) ||
                    (lanes_per_node == 16 && (pj == 0 || (pj >= 2 && pj <= 4)));
    if (!ok) return fail(ctx, FS_ERR_ARG, "cd lanes: %d lanes per node with %d positions per lane", lanes_per_node, pj);
    ctx->cd_lanes = lanes_per_node;
    ctx->cd_pj = lanes_per_node ? pj : 0;
    return FS_OK;
}

int fs_coo_create(fs_ctx *ctx, int64_t nrows, int64_t ncols, int32_t kr, int32_t kc, int64_t nmat,
                  const int64_t *dofnums_row, const int64_t *dofnums_col, int index_base, fs_coo **out) {
    if (!ctx || !out) return FS_ERR_ARG;
    *out = nullptr;
    ctx->err.clear();
    if (nrows < 0 || ncols < 0 || kr < 1 || kc < 1 || kr > 4096 || kc > 4096 || nmat < 0 ||
        (nmat > 0 && (!dofnums_row || !dofnums_col)))
        return fail(ctx, FS_ERR_ARG, "coo: bad sizes (nrows %lld, ncols %lld, k %d x %d, %lld matrices)",
                    (long long)nrows, (long long)ncols, kr, kc, (long long)nmat);
    const int64_t ke = (int64_t)kr * kc;
    for (int64_t i = 0; i < nmat * kr; ++i)
        if (dofnums_row[i] - index_base < 0) return fail(ctx, FS_ERR_ARG, "coo: row dof number below the index base");
    for (int64_t i = 0; i < nmat * kc; ++i)
        if (dofnums_col[i] - index_base < 0) return fail(ctx, FS_ERR_ARG, "coo: column dof number below the index base");
    FS_HIP(ctx, hipSetDevice(ctx->device));
    fs_coo *h = new (std::nothrow) fs_coo;
    if (!h) return fail(ctx, FS_ERR_NOMEM, "coo handle");
    h->ctx = ctx;
    h->nrows = nrows;
    h->ncols = ncols;
    h->nmat = nmat;
    h->kr = kr;
    h->kc = kc;
    // The COO stream of assemble! (element order; within an element
    // column-major: for j in columns, for i in rows) is bucketed by row by
    // threads that each own a row range and scan the stream in order, so a
    // row receives its entries in stream order.
    std::vector<int64_t> cnt(nrows + 1, 0);
    auto rowpass = [&](std::vector<int64_t> *fill, int64_t *bcol, int64_t *bpos) {
        const int64_t nt = std::max<int64_t>(1, std::min<int64_t>(host_threads(), nrows / 4096 + 1));
        std::vector<std::thread> th;
        for (int64_t t = 0; t < nt; ++t)
            th.emplace_back([&, t] {
                const int64_t r0 = nrows * t / nt, r1 = nrows * (t + 1) / nt;
                for (int64_t m = 0; m < nmat; ++m) {
                    const int64_t *dr = dofnums_row + m * kr, *dc = dofnums_col + m * kc;
                    for (int j = 0; j < kc; ++j) {
                        const int64_t c = dc[j] - index_base;
                        if (c >= ncols) continue;
                        for (int i = 0; i < kr; ++i) {
                            const int64_t r = dr[i] - index_base;
                            if (r < r0 || r >= r1) continue;
                            if (!fill) {
                                cnt[r + 1]++;
                            } else {
                                const int64_t q = (*fill)[r]++;
                                bcol[q] = c;
                                bpos[q] = m * ke + (int64_t)j * kr + i;
                            }
                        }
                    }
                }
            });
        for (auto &x : th) x.join();
    };
    rowpass(nullptr, nullptr, nullptr);
    for (int64_t r = 0; r < nrows; ++r) cnt[r + 1] += cnt[r];
    const int64_t M = cnt[nrows];
    std::vector<int64_t> bcol(M), bpos(M), fill(cnt.begin(), cnt.end());
    rowpass(&fill, bcol.data(), bpos.data());
    // per row: a stable sort by column keeps stream order among duplicates;
    // each run of equal columns is one stored entry (zeros retained)
    std::vector<int64_t> rlen(nrows, 0), perm(M);
    parallel_chunks(nrows, [&](int64_t r0, int64_t r1) {
        for (int64_t r = r0; r < r1; ++r) {
            const int64_t b = cnt[r], e = cnt[r + 1];
            for (int64_t q = b; q < e; ++q) perm[q] = q;
            std::stable_sort(perm.begin() + b, perm.begin() + e, [&](int64_t x, int64_t y) { return bcol[x] < bcol[y]; });
            int64_t u = 0;
            for (int64_t q = b; q < e; ++q)
                if (q == b || bcol[perm[q]] != bcol[perm[q - 1]]) ++u;
            rlen[r] = u;
        }
    }, 4096);
    h->rowptr.assign(nrows + 1, 0);
    for (int64_t r = 0; r < nrows; ++r) h->rowptr[r + 1] = h->rowptr[r] + rlen[r];
    const int64_t nnz = h->rowptr[nrows];
    h->nnz = nnz;
    h->colval.resize(nnz);
    std::vector<int64_t> cptr(nnz + 1), clist(M);
    parallel_chunks(nrows, [&](int64_t r0, int64_t r1) {
        for (int64_t r = r0; r < r1; ++r) {
            int64_t s = h->rowptr[r] - 1;
            for (int64_t q = cnt[r]; q < cnt[r + 1]; ++q) {
                if (q == cnt[r] || bcol[perm[q]] != bcol[perm[q - 1]]) {
                    ++s;
                    h->colval[s] = bcol[perm[q]];
                    cptr[s] = q;
                }
                clist[q] = bpos[perm[q]];
            }
        }
    }, 4096);
    cptr[nnz] = M;
    int st;
    if ((st = upload(ctx, h->cptr, cptr.data(), cptr.size())) || (st = upload(ctx, h->clist, clist.data(), clist.size())) ||
        (st = dalloc(ctx, h->vals, (size_t)nnz * 8))) {
        delete h;
        return st;
    }
    if (hipStreamSynchronize(ctx->stream) != hipSuccess) {
        delete h;
        return fail(ctx, FS_ERR_DEVICE, "coo: upload of the reduction plan failed");
    }
    ++ctx->refs;
    *out = h;
    return FS_OK;
}

int fs_coo_pattern(const fs_coo *h, int64_t *nnz) {
    if (!h || !nnz) return FS_ERR_ARG;
    *nnz = h->nnz;
    return FS_OK;
}

int fs_coo_pattern_get(const fs_coo *h, int64_t *rowptr, int64_t *colval, int index_base) {
    if (!h || !rowptr || !colval) return FS_ERR_ARG;
    for (int64_t r = 0; r <= h->nrows; ++r) rowptr[r] = h->rowptr[r] + index_base;
    parallel_chunks(h->nnz, [&](int64_t a, int64_t b) {
        for (int64_t q = a; q < b; ++q) colval[q] = h->colval[q] + index_base;
    });
    return FS_OK;
}

int fs_coo_values(fs_coo *h, const double *elmats, double *nzval) {
    if (!h) return FS_ERR_ARG;
    fs_ctx *c = h->ctx;
    c->err.clear();
    const size_t nvals = (size_t)h->nmat * h->kr * h->kc;
    if (nvals > 0 && !elmats) return fail(c, FS_ERR_ARG, "coo: NULL element matrices");
    FS_HIP(c, hipSetDevice(c->device));
    int s;
    if ((s = upload(c, h->mats, elmats, nvals))) return s;
    FS_HIP(c, fs::launch_coo_reduce(h->nnz, h->cptr.as<int64_t>(), h->clist.as<int64_t>(), h->mats.as<double>(),
                                    h->vals.as<double>(), c->stream));
    if (nzval) return d2h(c, nzval, h->vals.p, (size_t)h->nnz * 8);
    FS_HIP(c, hipStreamSynchronize(c->stream));
    return FS_OK;
}

void fs_coo_destroy(fs_coo *h) {
    if (!h) return;
    fs_ctx *c = h->ctx;
    (void)hipSetDevice(c->device);
    (void)hipStreamSynchronize(c->stream);
    delete h;
    ctx_release(c);
}

int fs_model_mixed_info(const fs_model *m, int64_t *info) {
    if (!m || !info) return FS_ERR_ARG;
    const MixedStore &X = m->mixed;
    info[0] = X.built ? X.kmb_pairs : 0;
    info[1] = X.built ? X.nsr : 0;
    info[2] = (int64_t)m->nbr_list.size();
    info[3] = X.built ? X.nslices : 0;
    return FS_OK;
}

int fs_model_last_timing(const fs_model *m, int which, double *ms) {
    if (!m || !ms || which < 0 || which > 6) return FS_ERR_ARG;
    *ms = m->timing[which];
    return FS_OK;
}

}  // extern "C"
