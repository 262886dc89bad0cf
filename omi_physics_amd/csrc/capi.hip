// capi.hip — host side of libomi_hip.so: batch memory, world placement, the
// per-step launch sequence (optionally captured in a HIP graph) and the C ABI
// declared in include/omi_hip.h.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <mutex>
#include <cstdio>
#include <cstring>
#include <string>
#include <utility>
#include <vector>

#include "../../include/omi_hip.h"
#include "omi_batch.hpp"
#include "omi_math.hpp"
#include "omi_gjk.hpp"

namespace omi {
void launch_begin(const Dev& d, hipStream_t s);
void launch_pre(const Dev& d, int precision, int64_t nhw, double dt, double margin, hipStream_t s, bool collide_only);
void launch_stage_forces(int32_t n, const int32_t* mt, const uint8_t* awake, const double* g, const double* gf,
                         const double* ld, const double* ad, const double* qd, double dlin, double dang, double dt,
                         double* lv, double* av, hipStream_t s);
void launch_stage_positions(int32_t n, const int32_t* mt, const uint8_t* awake, const double* lv, const double* av,
                            double dt, double* pos, double* quat, hipStream_t s);
void launch_cache_seed(const Dev& d, int64_t base, int32_t n, hipStream_t s);
void launch_refit(const Dev& d, int64_t nhw, double margin, hipStream_t s);
void launch_sap_sort(const Dev& d, int c, const int32_t* wl, const int32_t* wn, int32_t np, int32_t block,
                     int grid, hipStream_t s);
void launch_sap_pairs(const Dev& d, int32_t record, int32_t grid, hipStream_t s);
void launch_narrow(const Dev& d, int precision, hipStream_t s);
void launch_compact(const Dev& d, int precision, int grid, hipStream_t s);
void launch_cache_plan(const Dev& d, hipStream_t s);
void launch_cache_write(const Dev& d, int precision, hipStream_t s);
void launch_post(const Dev& d, int64_t nhw, double dt, hipStream_t s);
void launch_joints(const Dev& d, int precision, double dt, int32_t lds_bytes, hipStream_t s);
void launch_sleep(const Dev& d, int64_t nhw, double dt, hipStream_t s, double lin, double ang, double tsleep);
void launch_sleep_union(const Dev& d, int64_t nhw, hipStream_t s);
void launch_sleep_decide(const Dev& d, int64_t nhw, double dt, hipStream_t s, double lin, double ang, double tsleep);
void launch_post_sleep(const Dev& d, int64_t nhw, double dt, hipStream_t s, double lin, double ang, double tsleep);
void launch_toggle(const Dev& d, hipStream_t s);
void launch_copy_segments(const CopySeg* segs, int32_t nseg, hipStream_t s);
void launch_trigger(const Dev& d, hipStream_t s);
void launch_capsule_triangles(int64_t n, const double* c0, const double* c1, double radius, const double* verts,
                              uint8_t* hit, double* points, double* normals, double* depths, hipStream_t s);
void launch_snapshot(const Dev& d, const int32_t* wl, const int64_t* woff, int32_t nw, double* pos, double* aa,
                     uint8_t* awake, uint8_t* dyn, hipStream_t s);
void launch_stage_boxbox(int precision, int32_t np, const double* ca, const double* Ra, const double* ha,
                         const double* cb, const double* Rb, const double* hb, int32_t* cnt, double* nrm,
                         double* pts, double* dep, hipStream_t s);
void launch_stage_spheres(int precision, int32_t np, int32_t mode, const double* ca, const double* ra,
                          const double* cb, const double* Rb, const double* hb, const double* rb, int32_t* cnt,
                          double* nrm, double* pt, double* dep, hipStream_t s);
void launch_prepare(const Dev& d, int precision, SolveIn in, bool stage, hipStream_t s, int32_t skip_above);
void launch_prepare_worlds(const Dev& d, int precision, SolveIn in, bool stage, const int32_t* wl,
                           const int32_t* wn, int grid, hipStream_t s);
void launch_solve_class(const Dev& d, int precision, int block, SolveIn in, const int32_t* wl, const int32_t* wn,
                        int32_t lds_bytes, int grid, hipStream_t s, int32_t isl);
void launch_islands(const Dev& d, const int32_t* wl, const int32_t* wn, int32_t lds_bytes, int grid, hipStream_t s);
bool islands_fit(int32_t n, int32_t lds_bytes);
size_t epa_work_bytes();
bool allow_dynamic_lds(int bytes);
bool allow_joint_lds(int bytes);
int32_t allow_convex_lds();
void launch_stage_convex(int32_t np, const int32_t* ka, const double* ca, const double* Ra, const double* ha,
                         const int32_t* oa, const int32_t* na, const int32_t* kb, const double* cb, const double* Rb,
                         const double* hb, const int32_t* ob, const int32_t* nb, const double* hull, int64_t nhull,
                         void* work, int32_t slots, int32_t* cnt, double* nrm, double* pt, double* dep, hipStream_t s);
}  // namespace omi

using namespace omi;

static thread_local std::string g_err;

static int32_t fail(int32_t code, const std::string& msg) {
    g_err = msg;
    return code;
}

static int32_t capacity_error(uint64_t ov);

#define HIPCHK(x)                                                                                    \
    do {                                                                                             \
        hipError_t e_ = (x);                                                                         \
        if (e_ != hipSuccess) return fail(OMI_ERR_HIP, std::string(#x) + ": " + hipGetErrorString(e_)); \
    } while (0)

// First-fit allocator over [0, cap) used for body ranges and table pools.
struct Ranges {
    int64_t cap = 0;
    std::vector<std::pair<int64_t, int64_t>> free_;   // (offset, length), sorted
    void init(int64_t c) { cap = c; free_.clear(); if (c > 0) free_.push_back({0, c}); }
    int64_t alloc(int64_t n) {
        if (n == 0) return 0;
        for (size_t i = 0; i < free_.size(); i++)
            if (free_[i].second >= n) {
                int64_t o = free_[i].first;
                free_[i].first += n; free_[i].second -= n;
                if (free_[i].second == 0) free_.erase(free_.begin() + i);
                return o;
            }
        return -1;
    }
    void release(int64_t o, int64_t n) {
        if (n == 0) return;
        free_.push_back({o, n});
        std::sort(free_.begin(), free_.end());
        std::vector<std::pair<int64_t, int64_t>> m;
        for (auto& r : free_) {
            if (!m.empty() && m.back().first + m.back().second == r.first) m.back().second += r.second;
            else m.push_back(r);
        }
        free_.swap(m);
    }
    int64_t high_water() const {
        // end of the last allocated byte: cap minus a free tail
        if (!free_.empty() && free_.back().first + free_.back().second == cap) return free_.back().first;
        return cap;
    }
};

struct WorldSlot {
    bool live = false;
    int64_t off = 0, n = 0;
    int64_t mat_off = 0, mat_n = 0, filt_off = 0, filt_n = 0, hull_off = 0, hull_n = 0, j_off = 0, j_n = 0;
    int32_t flags = 0;
};

// Timed stages of a step; solve_top is the solver launch of the largest size
// class alone on the batch stream, solve_rest the wait for the other classes.
static const char* kKernelNames =
    "begin,pre,sap,narrow,compact,prepare,solve_top,solve_rest,cache,joints,post,sleep,toggle";
static const int kNumKernels = 13;

struct omi_batch {
    omi_batch_config cfg{};
    Dev d{};
    hipStream_t stream = nullptr;
    hipStream_t aux[3] = {nullptr, nullptr, nullptr};   // size-class branches of the solve (default priority)
    hipEvent_t fork = nullptr, join[3] = {nullptr, nullptr, nullptr};
    hipEvent_t rows = nullptr;                     // the smaller size classes' solver rows done
    hipEvent_t next = nullptr;                     // the second largest class's solver rows done
    hipEvent_t solved = nullptr, cached = nullptr;  // the solve done; the warm-start cache rebuilt
    std::vector<void*> allocs;
    int64_t bytes = 0;
    Ranges bodies, mats, filts, hulls, joints;
    bool any_joints = false;
    bool any_triggers = false;
    std::vector<WorldSlot> slots;
    int32_t* d_wl = nullptr;               // [4][max_worlds] world slots per size class
    int32_t* d_wn = nullptr;               // [4] class counts
    int32_t np_class[4] = {256, 1024, 4096, 0};
    int32_t lds_bytes[4] = {0, 0, 0, 0};    // solver: dynamic LDS per world of each size class
    int32_t joint_lds_max = 0;              // dynamic LDS a joint launch may take
    int32_t joint_lds = 0;                  // joints: LDS for the joint bodies of the largest joint world
    int32_t live[4] = {0, 0, 0, 0};        // grid of each class's per-world launches (0: none)
    int32_t sap_block[4] = {64, 128, 1024, 1024};
    int32_t solve_block[4] = {64, 128, 256, 256};
    // size classes solved island by island (few worlds, each fitting the
    // island solve's LDS): their worlds' islands spread over the chip
    int32_t isl[4] = {0, 0, 0, 0};
    int32_t cus = 256;                     // compute units of the device
    int32_t np_lds_max = 4096;
    int32_t dyn_lds = 0;                   // dynamic LDS a launch may take
    // entries of each class's broadphase sort in LDS (0: global scratch);
    // classes 0-2 their size bound, class 3 its largest world when it fits
    int32_t sort_np[4] = {256, 1024, 4096, 0};
    int32_t sap_grid = 2048;               // blocks of the chip-wide broadphase pair passes
    bool lists_dirty = true;
    int64_t nhw = 0;
    bool use_graph = true;
    hipGraphExec_t graph = nullptr;
    double graph_dt = 0.0;
    bool timing = false;
    std::vector<hipEvent_t> ev;
    std::vector<double> kms;
    int solver_block = 256;
    // pose snapshots: packed device rows, their world list, pinned host copy
    // (grown on demand, outside any captured step)
    unsigned char* snap_dev = nullptr;
    unsigned char* snap_host = nullptr;
    int64_t snap_cap = 0;                  // bodies
    int32_t* snap_wl = nullptr;
    int64_t* snap_woff = nullptr;
    int32_t snap_wcap = 0;
    // Staged device writes (world loads and removals, world lists): pending
    // segments and their data on the host, then at a flush one transfer from
    // a pinned arena on the upload stream (it overlaps work queued on the
    // batch stream) into device staging, and one k_copy_segments launch
    // ordered on the batch stream.  Arena space is reused once the last
    // launch that read it has run.
    std::vector<CopySeg> seg;              // pending
    std::vector<uint8_t> seg_rel;          // 1: src is an offset into seg_data
    size_t seg_group = 0;                  // first segment of the open group
    std::vector<int32_t> seg_cuts;         // group ends: one launch per group, in order
    std::vector<unsigned char> seg_data;
    unsigned char* stage_host = nullptr;   // pinned
    unsigned char* stage_dev = nullptr;
    int64_t stage_cap = 0, stage_used = 0;
    hipStream_t up_stream = nullptr;
    hipEvent_t ev_up = nullptr, ev_staged = nullptr;
    // Readback of worlds (omi_batch_fetch_world): gathered on the device into
    // staging, one transfer into a pinned arena; tickets until collected.
    // Each fetch holds a buffer pair (pinned host, device) from a pool until
    // its ticket is collected.
    struct FetchBuf { int64_t cap = 0; unsigned char* host = nullptr; unsigned char* dev = nullptr; };
    struct Fetch { int64_t n = 0; int32_t buf = -1; hipEvent_t ev = nullptr; bool live = false; };
    std::vector<FetchBuf> fbufs;
    std::vector<uint8_t> fbuf_busy;
    std::vector<Fetch> fetches;
    // step markers (omi_batch_mark): an event and the overflow flags copied
    // at that point into a pinned slot
    struct Mark { hipEvent_t ev = nullptr; uint64_t* flags = nullptr; bool live = false; };
    std::vector<Mark> marks;

    template <typename T>
    int32_t alloc(T** p, int64_t count) {
        size_t sz = sizeof(T) * (size_t)(count > 0 ? count : 1);
        void* q = nullptr;
        HIPCHK(hipMalloc(&q, sz));
        HIPCHK(hipMemset(q, 0, sz));
        allocs.push_back(q);
        bytes += (int64_t)sz;
        *p = (T*)q;
        return OMI_OK;
    }
};

extern "C" const char* omi_last_error(void) { return g_err.c_str(); }
extern "C" int32_t omi_abi_version(void) { return 3; }
extern "C" int32_t omi_device_count(void) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return 0;
    return n;
}

extern "C" int32_t omi_device_memory(int64_t* free_bytes, int64_t* total_bytes) {
    if (!free_bytes || !total_bytes) return fail(OMI_ERR_ARG, "null argument");
    size_t f = 0, t = 0;
    HIPCHK(hipMemGetInfo(&f, &t));
    *free_bytes = (int64_t)f;
    *total_bytes = (int64_t)t;
    return OMI_OK;
}

static int32_t check_device() {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n == 0)
        return fail(OMI_ERR_HIP, "no HIP device is visible: the contact pipeline needs an MI355X (gfx950)");
    return OMI_OK;
}

static void free_all(omi_batch* b) {
    if (b->graph) hipGraphExecDestroy(b->graph);
    if (b->up_stream) hipStreamSynchronize(b->up_stream);
    if (b->stage_host) hipHostFree(b->stage_host);
    if (b->stage_dev) hipFree(b->stage_dev);
    b->stage_host = b->stage_dev = nullptr;
    for (auto& f : b->fbufs) {
        if (f.host) hipHostFree(f.host);
        if (f.dev) hipFree(f.dev);
    }
    b->fbufs.clear();
    for (auto& f : b->fetches) if (f.ev) hipEventDestroy(f.ev);
    b->fetches.clear();
    for (auto& m : b->marks) {
        if (m.ev) hipEventDestroy(m.ev);
        if (m.flags) hipHostFree(m.flags);
    }
    b->marks.clear();
    if (b->ev_up) hipEventDestroy(b->ev_up);
    if (b->ev_staged) hipEventDestroy(b->ev_staged);
    if (b->up_stream) hipStreamDestroy(b->up_stream);
    b->ev_up = b->ev_staged = nullptr;
    b->up_stream = nullptr;
    if (b->snap_dev) hipFree(b->snap_dev);
    if (b->snap_host) hipHostFree(b->snap_host);
    if (b->snap_wl) hipFree(b->snap_wl);
    if (b->snap_woff) hipFree(b->snap_woff);
    b->snap_dev = b->snap_host = nullptr;
    b->snap_wl = nullptr;
    b->snap_woff = nullptr;
    for (auto e : b->ev) hipEventDestroy(e);
    for (void* p : b->allocs) hipFree(p);
    if (b->stream) hipStreamDestroy(b->stream);
    for (int q = 0; q < 3; q++) {
        if (b->aux[q]) hipStreamDestroy(b->aux[q]);
        if (b->join[q]) hipEventDestroy(b->join[q]);
    }
    if (b->fork) hipEventDestroy(b->fork);
    for (hipEvent_t e : {b->rows, b->next, b->solved, b->cached})
        if (e) hipEventDestroy(e);
    b->allocs.clear();
}

extern "C" int32_t omi_batch_create(const omi_batch_config* cfg, omi_batch** out) {
    if (!cfg || !out) return fail(OMI_ERR_ARG, "null argument");
    if (cfg->precision != OMI_FP64 && cfg->precision != OMI_FP32)
        return fail(OMI_ERR_ARG, "precision must be OMI_FP64 (64) or OMI_FP32 (32)");
    if (cfg->max_worlds <= 0 || cfg->max_bodies <= 0 || cfg->max_pairs <= 0 || cfg->max_contacts <= 0)
        return fail(OMI_ERR_ARG, "capacities must be positive");
    if (cfg->max_pairs > (1ll << 30) || cfg->max_contacts > (1ll << 24) || cfg->max_bodies > (1ll << 30))
        return fail(OMI_ERR_ARG, "capacity exceeds the index range (pairs 2^30, contacts 2^24, bodies 2^30)");
    if (cfg->max_solver_iterations < 0 || cfg->max_solver_iterations > 30)
        return fail(OMI_ERR_ARG, "max_solver_iterations must be 0 (default 16) .. 30");
    if (cfg->precision == OMI_FP64)
        for (int c = 0; c < 4; c++) {
            const int32_t blk = cfg->solve_block[c] > 0 ? cfg->solve_block[c] : cfg->solver_block;
            if (blk > FP64_SOLVE_BLOCK_MAX)
                return fail(OMI_ERR_ARG, "FP64 solver blocks are 64, 128 or 256 threads");
        }
    int32_t rc = check_device();
    if (rc) return rc;
    omi_batch* b = new omi_batch();
    b->cfg = *cfg;
    Dev& d = b->d;
    const int64_t N = cfg->max_bodies, P = cfg->max_pairs, K = cfg->max_contacts;
    const int32_t W = cfg->max_worlds;
    int64_t H = 16;
    while (H < 2 * K) H <<= 1;
    d.ncap = N; d.pcap = P; d.kcap = K; d.hcap = H; d.wcap = W;
    d.clip16_all = cfg->clip16_all ? 1 : 0;
    const size_t rs = cfg->precision == OMI_FP64 ? 8 : 4;
    int dev = 0;
    hipGetDevice(&dev);
    int smem = 65536;
    hipDeviceGetAttribute(&smem, hipDeviceAttributeMaxSharedMemoryPerBlock, dev);
    int cus = 256;
    hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    b->sap_grid = std::max(8, cus) * 8;
    b->cus = std::max(1, cus);
    // dynamic LDS per launch: the device's LDS less 1 KB for the kernels'
    // static LDS, or 64 KB less that if the runtime refuses the larger size
    int dyn_lds = smem - 1024;
    if (dyn_lds > 65536 - 1024 && !allow_dynamic_lds(dyn_lds)) dyn_lds = 65536 - 1024;
    b->np_lds_max = dyn_lds >= 4096 * 16 ? 4096 : 2048;
    b->np_class[2] = b->np_lds_max;
    b->dyn_lds = dyn_lds;
    for (int c = 0; c < 3; c++) b->sort_np[c] = b->np_class[c];
    // the solver's LDS per world of a size class: small worlds take less so
    // that more of them share a CU (a world that does not fit keeps its
    // solver state in global memory)
    {
        const int32_t full = (dyn_lds - 1024) & ~15;
        const bool f64 = cfg->precision == OMI_FP64;
        const int32_t want[4] = {f64 ? 32 << 10 : 24 << 10, f64 ? 80 << 10 : 64 << 10, full, full};
        for (int c = 0; c < 4; c++) b->lds_bytes[c] = std::min(want[c], full);
        // the joint sweeps: the LDS less 8 KB for their static LDS
        int32_t jl = std::max(0, full - 8192);
        if (jl > 65536 - 8192 && !allow_joint_lds(jl)) jl = 65536 - 8192;
        b->joint_lds_max = jl;
    }
    b->solver_block = cfg->solver_block > 0 ? cfg->solver_block : 256;
    auto okb = [](int32_t x) { return x == 64 || x == 128 || x == 256 || x == 512 || x == 1024; };
    for (int c = 0; c < 4; c++) {
        if (okb(cfg->sap_block[c])) b->sap_block[c] = cfg->sap_block[c];
        if (okb(cfg->solve_block[c])) b->solve_block[c] = cfg->solve_block[c];
        else if (cfg->solver_block > 0 && okb(cfg->solver_block)) b->solve_block[c] = cfg->solver_block;
    }
#define A(p, n) do { if ((rc = b->alloc(&(p), (n))) != OMI_OK) { free_all(b); delete b; return rc; } } while (0)
    A(d.pos, 3 * N); A(d.quat, 4 * N); A(d.ppos, 3 * N); A(d.pquat, 4 * N);
    A(d.lv, 6 * N); d.av = d.lv + 3 * N;
    A(d.lo, 3 * N); A(d.hi, 3 * N); A(d.half, 3 * N); A(d.inv_mass, N); A(d.inv_inertia, 3 * N);
    A(d.gf, N); A(d.ld, N); A(d.ad, N); A(d.qd, N); A(d.timer, N);
    unsigned char *rot, *iiw, *svel;
    A(rot, 9 * N * (int64_t)rs); A(iiw, IIW_STRIDE * N * (int64_t)rs);
    A(svel, 6 * N * (int64_t)rs);
    d.rot = rot; d.iiw = iiw; d.svel = svel;
    A(d.world, N); A(d.mat, N); A(d.filt, N); A(d.hoff, N); A(d.hcnt, N);
    A(d.mt, N); A(d.kind, N); A(d.bflags, N); A(d.awake, N); A(d.awake0, N); A(d.contacted, N); A(d.notready, N);
    A(d.parent, N); A(d.last, N); A(d.lmask, N);
    A(d.skey, 2 * N); A(d.sidx, 2 * N); A(d.soff, 2 * N);
    A(d.sbox, 4 * N); A(d.sinfo, N); A(d.sorig, N);
    A(d.hull, 3 * std::max<int64_t>(cfg->max_hull_pts, 1));
    A(d.w_off, W); A(d.w_cnt, W); A(d.w_flags, W); A(d.w_iters, W);
    A(d.w_mat_off, W); A(d.w_nmat, W); A(d.w_filt_off, W); A(d.w_nfilt, W);
    A(d.w_grav, 3 * W); A(d.w_damp, 2 * W);
    A(d.w_ps, W); A(d.w_pn, W); A(d.w_bps, W); A(d.w_bpn, W); A(d.w_cs, W); A(d.w_cn, W); A(d.w_mn, W); A(d.w_ms, W); A(d.w_levels, W);
    A(d.w_cache_s, 2 * W); A(d.w_cache_n, 2 * W);
    A(d.w_x, W); A(d.w_t0, W); A(d.w_tn, W);
    {
        // broadphase candidate tiles: max_candidates x-window candidates per step
        const int64_t xc = cfg->max_candidates > 0 ? cfg->max_candidates : 32 * P + (1 << 20);
        d.tcap = std::min<int64_t>(xc / SAP_TILE + W + 1, (int64_t)1 << 30);
    }
    A(d.tworld, d.tcap); A(d.tcnt, 2 * (d.tcap + 1));
    const int64_t TAB = std::max<int64_t>(64 * (int64_t)W, 1024);
    A(d.mat_fric, TAB); A(d.mat_rest, TAB); A(d.filt_tab, TAB);
    A(d.pa, P); A(d.pb, P);
    d.tpcap = cfg->max_trigger_pairs > 0 ? cfg->max_trigger_pairs : std::max<int64_t>(1024, P / 16);
    A(d.tpa, d.tpcap); A(d.tpb, d.tpcap); A(d.tva, d.tpcap); A(d.tvb, d.tpcap);
    A(d.bpa, cfg->record_broadphase ? P : 1); A(d.bpb, cfg->record_broadphase ? P : 1);
    A(d.mcnt, P); A(d.mpts, 3 * 4 * P); A(d.cvx, P);
    d.pchunk_cap = (P + CPT_CHUNK - 1) / CPT_CHUNK + 1;
    A(d.pfx, 2 * (P + 1)); A(d.pchunk, 2 * d.pchunk_cap);
    if (cfg->max_hull_pts > 1) {
        // GJK/EPA workspaces: one per thread of the convex narrowphase
        // launch, the pair capacity clamped to [4096, 16384] (16384 take
        // about 380 MB), so that a large world's queued pairs spread over
        // every CU
        unsigned char* ew;
        d.epa_slots = (int32_t)std::min<int64_t>(16384, std::max<int64_t>(4096, (P + 63) / 64 * 64));
        A(ew, (int64_t)d.epa_slots * (int64_t)epa_work_bytes());
        d.epa_work = ew;
        d.cvx_lds = allow_convex_lds();
    } else {
        d.epa_slots = 0;
        d.epa_work = nullptr;
    }
    unsigned char *mn, *md, *cn, *cd, *ci, *cs;
    A(mn, 3 * P * (int64_t)rs); A(md, 4 * P * (int64_t)rs);
    d.mnrm = mn; d.mdep = md;
    A(d.ca, K); A(d.cb, K); A(d.cpt, 3 * K);
    A(cn, 3 * K * (int64_t)rs); A(cd, K * (int64_t)rs); A(ci, 3 * K * (int64_t)rs); A(cs, SOLVE_ROW_BYTES * K);
    A(d.mrec, K);
    d.cnrm = cn; d.cdep = cd; d.cimp = ci; d.csolv = cs;
    d.slmax = 1 + (cfg->max_solver_iterations > 0 ? cfg->max_solver_iterations : 16);
    A(d.clevel, K * d.slmax); A(d.corder, K * d.slmax); A(d.lcount, K * d.slmax);
    A(d.cord_a, K * d.slmax); A(d.cord_b, K * d.slmax); A(d.cord_c, K * d.slmax);
    A(d.isl_lab, K); A(d.isl_desc, K * ISL_DESC); A(d.isl_mrec, K); A(d.isl_cidx, K);
    for (int q = 0; q < 2; q++) {
        unsigned char* ki;
        A(d.ka[q], K); A(d.kb[q], K); A(d.kpt[q], 3 * K); A(ki, 3 * K * (int64_t)rs); d.kimp[q] = ki;
        A(d.hkey[q], H); A(d.hval[q], H); A(d.hslot[q], K);
        HIPCHK(hipMemset(d.hkey[q], 0xff, sizeof(uint64_t) * H));
    }
    const int64_t JC = std::max<int64_t>(cfg->max_joints, 1);
    d.jcap = JC;
    A(d.ja, JC); A(d.jb, JC); A(d.janc, 3 * JC); A(d.jla, 3 * JC); A(d.jlb, 3 * JC); A(d.jhas, JC);
    unsigned char* jr;
    A(jr, (rs == 8 ? 160 : 96) * JC);     // JointItem<R>
    d.jrow = jr;
    A(d.jlevel, JC); A(d.jorder, JC); A(d.jcount, JC); A(d.jend, JC); A(d.jbody, 2 * JC);
    A(d.w_js, W); A(d.w_jn, W); A(d.w_jiters, W); A(d.w_jsched, W); A(d.w_jns, W);
    A(d.sel, 1); A(d.cnt, C_COUNT); A(d.hslot_n, 2);
    A(b->d_wl, 4 * W); A(b->d_wn, 4);
#undef A
    // every body slot starts free (world = -1)
    {
        std::vector<int32_t> m1((size_t)N, -1);
        HIPCHK(hipMemcpy(d.world, m1.data(), sizeof(int32_t) * N, hipMemcpyHostToDevice));
    }
    HIPCHK(hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking));
    HIPCHK(hipStreamCreateWithFlags(&b->up_stream, hipStreamNonBlocking));
    HIPCHK(hipEventCreateWithFlags(&b->ev_up, hipEventDisableTiming));
    HIPCHK(hipEventCreateWithFlags(&b->ev_staged, hipEventDisableTiming));
    // the size classes' branch streams are created once a second class holds
    // worlds: a batch of one class (a member of a group, say) then holds a
    // single stream, and so a single hardware queue
    for (int q = 0; q < 3; q++) HIPCHK(hipEventCreateWithFlags(&b->join[q], hipEventDisableTiming));
    HIPCHK(hipEventCreateWithFlags(&b->fork, hipEventDisableTiming));
    for (hipEvent_t* e : {&b->rows, &b->next, &b->solved, &b->cached}) HIPCHK(hipEventCreateWithFlags(e, hipEventDisableTiming));
    b->bodies.init(N);
    b->mats.init(TAB);
    b->filts.init(TAB);
    b->hulls.init(std::max<int64_t>(cfg->max_hull_pts, 1));
    b->joints.init(JC);
    b->slots.resize(W);
    *out = b;
    return OMI_OK;
}

extern "C" int32_t omi_batch_destroy(omi_batch* b) {
    if (!b) return OMI_OK;
    hipStreamSynchronize(b->stream);
    free_all(b);
    delete b;
    return OMI_OK;
}

extern "C" int64_t omi_batch_device_bytes(const omi_batch* b) { return b ? b->bytes : 0; }

// host [n][k] rows -> device component-major rows at body offset `off`
template <typename T>
static int32_t upload_soa(omi_batch* b, T* dst, int64_t ldd, int64_t off, const T* src, int64_t n, int k) {
    std::vector<T> tmp((size_t)n);
    for (int c = 0; c < k; c++) {
        for (int64_t i = 0; i < n; i++) tmp[(size_t)i] = src[i * k + c];
        HIPCHK(hipMemcpyAsync(dst + (size_t)c * ldd + off, tmp.data(), sizeof(T) * n, hipMemcpyHostToDevice, b->stream));
        HIPCHK(hipStreamSynchronize(b->stream));
    }
    return OMI_OK;
}

template <typename T>
static int32_t download_soa(omi_batch* b, T* dst, const T* src, int64_t lds, int64_t off, int64_t n, int k) {
    std::vector<T> tmp((size_t)n);
    for (int c = 0; c < k; c++) {
        HIPCHK(hipMemcpyAsync(tmp.data(), src + (size_t)c * lds + off, sizeof(T) * n, hipMemcpyDeviceToHost, b->stream));
        HIPCHK(hipStreamSynchronize(b->stream));
        for (int64_t i = 0; i < n; i++) dst[i * k + c] = tmp[(size_t)i];
    }
    return OMI_OK;
}

template <typename T>
static int32_t put(omi_batch* b, T* dst, const T* src, int64_t n) {
    if (n <= 0) return OMI_OK;
    HIPCHK(hipMemcpyAsync(dst, src, sizeof(T) * n, hipMemcpyHostToDevice, b->stream));
    HIPCHK(hipStreamSynchronize(b->stream));
    return OMI_OK;
}

template <typename T>
static int32_t get(omi_batch* b, T* dst, const T* src, int64_t n) {
    if (n <= 0) return OMI_OK;
    HIPCHK(hipMemcpyAsync(dst, src, sizeof(T) * n, hipMemcpyDeviceToHost, b->stream));
    HIPCHK(hipStreamSynchronize(b->stream));
    return OMI_OK;
}

// ---- staged device writes --------------------------------------------------

static int64_t al16h(int64_t x) { return (x + 15) & ~(int64_t)15; }

// The segments of one launch are copied by concurrent workgroups in no order,
// so a write that overlaps an earlier pending write of the open group closes
// the group: groups are launched one after another.
static void st_order(omi_batch* b, uint64_t dst, int64_t bytes) {
    for (size_t i = b->seg_group; i < b->seg.size(); i++) {
        const CopySeg& g = b->seg[i];
        const uint64_t e = g.dst + (uint64_t)(g.n * g.es);
        if (dst < e && g.dst < dst + (uint64_t)bytes) {
            b->seg_cuts.push_back((int32_t)b->seg.size());
            b->seg_group = b->seg.size();
            return;
        }
    }
}

// n elements of src to dst (a device address inside the batch)
template <typename T>
static void st_copy(omi_batch* b, T* dst, const T* src, int64_t n) {
    if (n <= 0) return;
    st_order(b, (uint64_t)(uintptr_t)dst, (int64_t)sizeof(T) * n);
    const int64_t off = al16h((int64_t)b->seg_data.size());
    b->seg_data.resize((size_t)(off + (int64_t)sizeof(T) * n));
    std::memcpy(b->seg_data.data() + off, src, sizeof(T) * (size_t)n);
    b->seg.push_back(CopySeg{(uint64_t)(uintptr_t)dst, (uint64_t)off, n, (int32_t)sizeof(T), 0});
    b->seg_rel.push_back(1);
}

// n elements from one device address of the batch to another
template <typename T>
static void st_dcopy(omi_batch* b, T* dst, const T* src, int64_t n) {
    if (n <= 0) return;
    st_order(b, (uint64_t)(uintptr_t)dst, (int64_t)sizeof(T) * n);
    b->seg.push_back(CopySeg{(uint64_t)(uintptr_t)dst, (uint64_t)(uintptr_t)src, n, (int32_t)sizeof(T), 0});
    b->seg_rel.push_back(0);
}

// host [n][k] rows -> device component-major rows at body offset `off`
template <typename T>
static void st_soa(omi_batch* b, T* dst, int64_t ldd, int64_t off, const T* src, int64_t n, int k) {
    if (n <= 0) return;
    std::vector<T> tmp((size_t)n);
    for (int c = 0; c < k; c++) {
        for (int64_t i = 0; i < n; i++) tmp[(size_t)i] = src[i * k + c];
        st_copy(b, dst + (size_t)c * ldd + off, tmp.data(), n);
    }
}

template <typename T>
static void st_fill(omi_batch* b, T* dst, T v, int64_t n) {
    if (n <= 0) return;
    st_order(b, (uint64_t)(uintptr_t)dst, (int64_t)sizeof(T) * n);
    uint64_t bits = 0;
    std::memcpy(&bits, &v, sizeof(T));
    b->seg.push_back(CopySeg{(uint64_t)(uintptr_t)dst, bits, n, (int32_t)sizeof(T), 1});
    b->seg_rel.push_back(0);
}

// Enqueue the pending segments: one host-to-device transfer on the upload
// stream, then one copy launch on the batch stream after it (so after every
// step queued before and before every step queued after).  Returns without
// waiting.
static int32_t st_flush(omi_batch* b) {
    if (b->seg.empty()) return OMI_OK;
    const int64_t nseg = (int64_t)b->seg.size();
    const int64_t tab = al16h(nseg * (int64_t)sizeof(CopySeg));
    const int64_t need = tab + al16h((int64_t)b->seg_data.size());
    // arena space is free again once the last copy launch has run
    if (b->stage_used > 0 && hipEventQuery(b->ev_staged) == hipSuccess) b->stage_used = 0;
    if (b->stage_used + need > b->stage_cap) {
        HIPCHK(hipEventSynchronize(b->ev_staged));
        b->stage_used = 0;
    }
    if (need > b->stage_cap) {
        HIPCHK(hipStreamSynchronize(b->up_stream));
        int64_t cap = std::max<int64_t>(1 << 20, b->stage_cap);
        while (cap < need) cap <<= 1;
        if (b->stage_host) HIPCHK(hipHostFree(b->stage_host));
        if (b->stage_dev) HIPCHK(hipFree(b->stage_dev));
        b->stage_host = b->stage_dev = nullptr;
        b->stage_cap = 0;
        HIPCHK(hipHostMalloc((void**)&b->stage_host, (size_t)cap, hipHostMallocDefault));
        HIPCHK(hipMalloc((void**)&b->stage_dev, (size_t)cap));
        b->stage_cap = cap;
    }
    unsigned char* h = b->stage_host + b->stage_used;
    unsigned char* dv = b->stage_dev + b->stage_used;
    for (size_t i = 0; i < b->seg.size(); i++)
        if (b->seg_rel[i]) b->seg[i].src = (uint64_t)(uintptr_t)(dv + tab + (int64_t)b->seg[i].src);
    std::memcpy(h, b->seg.data(), sizeof(CopySeg) * (size_t)nseg);
    if (!b->seg_data.empty()) std::memcpy(h + tab, b->seg_data.data(), b->seg_data.size());
    std::vector<int32_t> cuts;
    cuts.swap(b->seg_cuts);
    cuts.push_back((int32_t)nseg);
    b->seg.clear();
    b->seg_rel.clear();
    b->seg_data.clear();
    b->seg_group = 0;
    HIPCHK(hipMemcpyAsync(dv, h, (size_t)need, hipMemcpyHostToDevice, b->up_stream));
    HIPCHK(hipEventRecord(b->ev_up, b->up_stream));
    HIPCHK(hipStreamWaitEvent(b->stream, b->ev_up, 0));
    int32_t first = 0;
    for (int32_t cut : cuts) {
        launch_copy_segments((const CopySeg*)dv + first, cut - first, b->stream);
        first = cut;
    }
    HIPCHK(hipGetLastError());
    HIPCHK(hipEventRecord(b->ev_staged, b->stream));
    b->stage_used += need;
    return OMI_OK;
}

// The executable graph is destroyed only after its launches on the batch
// stream have finished (a group steps its members without synchronising, and
// rebuilding the lists between group steps re-instantiates the graph).
static void invalidate_graph(omi_batch* b) {
    if (b->graph) {
        hipStreamSynchronize(b->stream);
        hipGraphExecDestroy(b->graph);
        b->graph = nullptr;
    }
}

// A world's removal as staged writes (its bodies' world ids to -1, its flags
// and cache counts to zero) with its host ranges released; the caller flushes.
static void remove_staged(omi_batch* b, int32_t slot) {
    WorldSlot& s = b->slots[slot];
    if (!s.live) return;
    const int32_t W = b->cfg.max_worlds;
    st_fill<int32_t>(b, b->d.world + s.off, -1, s.n);
    st_fill<int32_t>(b, b->d.w_flags + slot, 0, 1);
    st_fill<int32_t>(b, b->d.w_cache_n + slot, 0, 1);
    st_fill<int32_t>(b, b->d.w_cache_n + W + slot, 0, 1);
    b->bodies.release(s.off, s.n);
    b->mats.release(s.mat_off, s.mat_n);
    b->filts.release(s.filt_off, s.filt_n);
    b->hulls.release(s.hull_off, s.hull_n);
    b->joints.release(s.j_off, s.j_n);
    s = WorldSlot();
    b->lists_dirty = true;
}

// Removal and load are ordered on the batch stream (staged writes, see
// st_flush): neither waits for queued steps, and a step queued after either
// call sees the change.
extern "C" int32_t omi_batch_remove_world(omi_batch* b, int32_t slot) {
    if (!b || slot < 0 || slot >= b->cfg.max_worlds) return fail(OMI_ERR_ARG, "bad world slot");
    if (!b->slots[slot].live) return OMI_OK;
    remove_staged(b, slot);
    return st_flush(b);
}

extern "C" int32_t omi_batch_load_world(omi_batch* b, int32_t slot, const omi_world_desc* w) {
    if (!b || !w) return fail(OMI_ERR_ARG, "null argument");
    if (slot < 0 || slot >= b->cfg.max_worlds) return fail(OMI_ERR_ARG, "world slot out of range");
    if (w->n < 0) return fail(OMI_ERR_ARG, "negative body count");
    if (w->n_mat < 1 || w->n_filt < 1) return fail(OMI_ERR_ARG, "material and filter tables need >= 1 slot");
    if (w->velocity_iterations < 0 || w->velocity_iterations > b->d.slmax - 1)
        return fail(OMI_ERR_ARG, "velocity_iterations exceeds the batch's max_solver_iterations");
    const int64_t n = w->n;
    for (int64_t i = 0; i < n; i++) {
        int k = w->kind[i], m = w->motion_type[i];
        if (k < 0 || k > 3 || m < 0 || m > 2) return fail(OMI_ERR_ARG, "bad kind or motion type");
        if (w->mat_slot[i] < 0 || w->mat_slot[i] >= w->n_mat || w->filt_slot[i] < 0 || w->filt_slot[i] >= w->n_filt)
            return fail(OMI_ERR_ARG, "material/filter slot out of range");
        if (k == 3 && (w->hull_off[i] < 0 || w->hull_cnt[i] < 1 || w->hull_off[i] + w->hull_cnt[i] > w->n_hull_pts))
            return fail(OMI_ERR_ARG, "convex body without valid hull points");
    }
    if (w->n_joints < 0) return fail(OMI_ERR_ARG, "negative joint count");
    for (int32_t j = 0; j < w->n_joints; j++) {
        int32_t a = w->joint_a[j], c = w->joint_b[j];
        if (a < -1 || a >= w->n || c < -1 || c >= w->n || (a < 0 && c < 0) || a == c)
            return fail(OMI_ERR_ARG, "joint body index out of range");
    }
    remove_staged(b, slot);
    WorldSlot s;
    s.off = b->bodies.alloc(n);
    if (s.off < 0) {
        st_flush(b);                          // the previous world's removal
        return fail(OMI_ERR_CAPACITY, "not enough free body capacity in the batch");
    }
    s.n = n;
    s.mat_n = (int64_t)w->n_mat * w->n_mat;
    s.filt_n = (int64_t)w->n_filt * w->n_filt;
    s.hull_n = w->n_hull_pts;
    s.mat_off = b->mats.alloc(s.mat_n);
    s.filt_off = b->filts.alloc(s.filt_n);
    s.hull_off = b->hulls.alloc(s.hull_n);
    s.j_n = w->n_joints;
    s.j_off = b->joints.alloc(s.j_n);
    if (s.j_off < 0) {
        b->bodies.release(s.off, s.n);
        if (s.mat_off >= 0) b->mats.release(s.mat_off, s.mat_n);
        if (s.filt_off >= 0) b->filts.release(s.filt_off, s.filt_n);
        if (s.hull_off >= 0) b->hulls.release(s.hull_off, s.hull_n);
        st_flush(b);
        return fail(OMI_ERR_CAPACITY, "joint capacity of the batch is full");
    }
    if (s.mat_off < 0 || s.filt_off < 0 || s.hull_off < 0) {
        b->joints.release(s.j_off, s.j_n);
        b->bodies.release(s.off, s.n);
        if (s.mat_off >= 0) b->mats.release(s.mat_off, s.mat_n);
        if (s.filt_off >= 0) b->filts.release(s.filt_off, s.filt_n);
        if (s.hull_off >= 0) b->hulls.release(s.hull_off, s.hull_n);
        st_flush(b);
        return fail(OMI_ERR_CAPACITY, "table or hull pool of the batch is full");
    }
    Dev& d = b->d;
    const int64_t N = d.ncap;
    st_soa(b, d.pos, N, s.off, w->position, n, 3);
    st_soa(b, d.quat, N, s.off, w->orientation, n, 4);
    st_soa(b, d.ppos, N, s.off, w->prev_position ? w->prev_position : w->position, n, 3);
    st_soa(b, d.pquat, N, s.off, w->prev_orientation ? w->prev_orientation : w->orientation, n, 4);
    st_soa(b, d.lv, N, s.off, w->linear_velocity, n, 3);
    st_soa(b, d.av, N, s.off, w->angular_velocity, n, 3);
    st_soa(b, d.half, N, s.off, w->half, n, 3);
    st_soa(b, d.inv_inertia, N, s.off, w->inv_inertia, n, 3);
    st_copy(b, d.inv_mass + s.off, w->inv_mass, n);
    st_copy(b, d.gf + s.off, w->gravity_factor, n);
    st_copy(b, d.ld + s.off, w->linear_damping, n);
    st_copy(b, d.ad + s.off, w->angular_damping, n);
    st_copy(b, d.qd + s.off, w->quadratic_drag, n);
    st_copy(b, d.timer + s.off, w->sleep_timer, n);
    st_copy(b, d.awake + s.off, w->awake, n);
    std::vector<uint8_t> mt((size_t)n), kind((size_t)n), bf((size_t)n);
    std::vector<int32_t> wid((size_t)n, slot), hoff((size_t)n), hcnt((size_t)n);
    for (int64_t i = 0; i < n; i++) {
        mt[i] = (uint8_t)w->motion_type[i];
        kind[i] = (uint8_t)w->kind[i];
        bf[i] = (uint8_t)((w->collider_shape[i] >= 0 ? BF_COLLIDER : 0) | (w->trigger_shape[i] >= 0 ? BF_TRIGGER : 0));
        if (w->trigger_shape[i] >= 0 && w->has_collision && !b->any_triggers) { b->any_triggers = true; invalidate_graph(b); }
        hoff[i] = w->kind[i] == 3 ? (int32_t)(s.hull_off + w->hull_off[i]) : 0;
        hcnt[i] = w->kind[i] == 3 ? w->hull_cnt[i] : 0;
    }
    st_copy(b, d.mt + s.off, mt.data(), n);
    st_copy(b, d.kind + s.off, kind.data(), n);
    st_copy(b, d.bflags + s.off, bf.data(), n);
    st_copy(b, d.mat + s.off, w->mat_slot, n);
    st_copy(b, d.filt + s.off, w->filt_slot, n);
    st_copy(b, d.hoff + s.off, hoff.data(), n);
    st_copy(b, d.hcnt + s.off, hcnt.data(), n);
    st_copy(b, d.mat_fric + s.mat_off, w->friction_table, s.mat_n);
    st_copy(b, d.mat_rest + s.mat_off, w->restitution_table, s.mat_n);
    st_copy(b, d.filt_tab + s.filt_off, w->filter_table, s.filt_n);
    if (s.hull_n) st_copy(b, d.hull + 3 * s.hull_off, w->hull_pts, 3 * s.hull_n);
    const int32_t W = b->cfg.max_worlds;
    int32_t i32;
    if (s.j_n > 0) {
        std::vector<int32_t> ja((size_t)s.j_n), jb((size_t)s.j_n);
        std::vector<uint8_t> jh((size_t)s.j_n, 0);
        std::vector<double> zl((size_t)(3 * s.j_n), 0.0);
        for (int64_t j = 0; j < s.j_n; j++) {
            ja[j] = w->joint_a[j] >= 0 ? (int32_t)(s.off + w->joint_a[j]) : -1;
            jb[j] = w->joint_b[j] >= 0 ? (int32_t)(s.off + w->joint_b[j]) : -1;
            if (w->joint_fixed) jh[j] = (uint8_t)(w->joint_fixed[j] & 3);
        }
        st_copy(b, d.ja + s.j_off, ja.data(), s.j_n);
        st_copy(b, d.jb + s.j_off, jb.data(), s.j_n);
        st_copy(b, d.jhas + s.j_off, jh.data(), s.j_n);
        st_soa(b, d.janc, d.jcap, s.j_off, w->joint_anchor, s.j_n, 3);
        st_soa(b, d.jla, d.jcap, s.j_off, w->joint_local_a ? w->joint_local_a : zl.data(), s.j_n, 3);
        st_soa(b, d.jlb, d.jcap, s.j_off, w->joint_local_b ? w->joint_local_b : zl.data(), s.j_n, 3);
        if (!b->any_joints) { b->any_joints = true; invalidate_graph(b); }
        // LDS for the records of this world's joint bodies during the sweeps
        std::vector<int32_t> ends;
        ends.reserve(2 * (size_t)s.j_n);
        for (int64_t j = 0; j < s.j_n; j++) {
            if (ja[j] >= 0) ends.push_back(ja[j]);
            if (jb[j] >= 0) ends.push_back(jb[j]);
        }
        std::sort(ends.begin(), ends.end());
        const int64_t nb = std::unique(ends.begin(), ends.end()) - ends.begin();
        const int64_t per_body = b->cfg.precision == OMI_FP64 ? 132 : 100;   // JointBody<R>::lds
        const int64_t want = std::min<int64_t>((nb * per_body + 15) & ~15, (int64_t)b->joint_lds_max);
        if (want > b->joint_lds) { b->joint_lds = (int32_t)want; invalidate_graph(b); }
    }
    i32 = (int32_t)s.j_off; st_copy(b, d.w_js + slot, &i32, 1);
    i32 = (int32_t)s.j_n; st_copy(b, d.w_jn + slot, &i32, 1);
    i32 = w->joint_iterations; st_copy(b, d.w_jiters + slot, &i32, 1);
    i32 = 0; st_copy(b, d.w_jsched + slot, &i32, 1);   // joint schedule rebuilt
    s.flags = WF_ACTIVE | (w->has_collision ? WF_COLLISION : 0) | (w->sleep_enabled ? WF_SLEEP : 0) |
              (s.j_n > 0 ? WF_JOINTS : 0);
    i32 = (int32_t)s.off; st_copy(b, d.w_off + slot, &i32, 1);
    i32 = (int32_t)n; st_copy(b, d.w_cnt + slot, &i32, 1);
    i32 = w->velocity_iterations; st_copy(b, d.w_iters + slot, &i32, 1);
    i32 = (int32_t)s.mat_off; st_copy(b, d.w_mat_off + slot, &i32, 1);
    i32 = w->n_mat; st_copy(b, d.w_nmat + slot, &i32, 1);
    i32 = (int32_t)s.filt_off; st_copy(b, d.w_filt_off + slot, &i32, 1);
    i32 = w->n_filt; st_copy(b, d.w_nfilt + slot, &i32, 1);
    i32 = 0;
    st_copy(b, d.w_mn + slot, &i32, 1);
    st_copy(b, d.w_ms + slot, &i32, 1);
    st_copy(b, d.w_cache_n + slot, &i32, 1);
    st_copy(b, d.w_cache_n + W + slot, &i32, 1);
    st_copy(b, d.w_pn + slot, &i32, 1);
    st_copy(b, d.w_bpn + slot, &i32, 1);
    st_copy(b, d.w_cn + slot, &i32, 1);
    for (int r = 0; r < 3; r++)
        st_copy(b, d.w_grav + (size_t)r * W + slot, &w->gravity[r], 1);
    st_copy(b, d.w_damp + slot, &w->default_linear_damping, 1);
    st_copy(b, d.w_damp + W + slot, &w->default_angular_damping, 1);
    st_copy(b, d.w_flags + slot, &s.flags, 1);
    st_copy(b, d.world + s.off, wid.data(), n);
    s.live = true;
    b->slots[slot] = s;
    b->lists_dirty = true;
    return st_flush(b);
}

extern "C" int32_t omi_batch_set_state(omi_batch* b, int32_t slot, const double* pos, const double* quat,
                                       const double* lv, const double* av, const uint8_t* awake,
                                       const double* timer) {
    if (!b || slot < 0 || slot >= b->cfg.max_worlds || !b->slots[slot].live) return fail(OMI_ERR_ARG, "bad world slot");
    const WorldSlot& s = b->slots[slot];
    Dev& d = b->d;
    int32_t rc;
    HIPCHK(hipStreamSynchronize(b->stream));
    if (pos && (rc = upload_soa(b, d.pos, d.ncap, s.off, pos, s.n, 3))) return rc;
    if (quat && (rc = upload_soa(b, d.quat, d.ncap, s.off, quat, s.n, 4))) return rc;
    if (lv && (rc = upload_soa(b, d.lv, d.ncap, s.off, lv, s.n, 3))) return rc;
    if (av && (rc = upload_soa(b, d.av, d.ncap, s.off, av, s.n, 3))) return rc;
    if (awake && (rc = put(b, d.awake + s.off, awake, s.n))) return rc;
    if (timer && (rc = put(b, d.timer + s.off, timer, s.n))) return rc;
    return OMI_OK;
}

extern "C" int32_t omi_batch_read_joint_anchors(omi_batch* b, int32_t slot, double* la, double* lb, uint8_t* fixed) {
    if (!b || slot < 0 || slot >= b->cfg.max_worlds || !b->slots[slot].live) return fail(OMI_ERR_ARG, "bad world slot");
    const WorldSlot& s = b->slots[slot];
    const Dev& d = b->d;
    HIPCHK(hipStreamSynchronize(b->stream));
    int32_t rc;
    if (s.j_n == 0) return OMI_OK;
    if (la && (rc = download_soa(b, la, d.jla, d.jcap, s.j_off, s.j_n, 3))) return rc;
    if (lb && (rc = download_soa(b, lb, d.jlb, d.jcap, s.j_off, s.j_n, 3))) return rc;
    if (fixed && (rc = get(b, fixed, d.jhas + s.j_off, s.j_n))) return rc;
    return OMI_OK;
}

extern "C" int32_t omi_batch_set_aabbs(omi_batch* b, int32_t slot, const double* aabb_min, const double* aabb_max) {
    if (!b || slot < 0 || slot >= b->cfg.max_worlds || !b->slots[slot].live) return fail(OMI_ERR_ARG, "bad world slot");
    const WorldSlot& s = b->slots[slot];
    int32_t rc;
    HIPCHK(hipStreamSynchronize(b->stream));
    if (aabb_min && (rc = upload_soa(b, b->d.lo, b->d.ncap, s.off, aabb_min, s.n, 3))) return rc;
    if (aabb_max && (rc = upload_soa(b, b->d.hi, b->d.ncap, s.off, aabb_max, s.n, 3))) return rc;
    return OMI_OK;
}

extern "C" int32_t omi_batch_set_params(omi_batch* b, int32_t slot, const double* inv_mass, const double* inv_inertia,
                                        const double* gravity_factor, const double* linear_damping,
                                        const double* angular_damping, const double* quadratic_drag,
                                        const double* gravity, const double* default_damping, int32_t sleep_enabled,
                                        int32_t velocity_iterations) {
    if (!b || slot < 0 || slot >= b->cfg.max_worlds || !b->slots[slot].live) return fail(OMI_ERR_ARG, "bad world slot");
    if (velocity_iterations > b->d.slmax - 1)
        return fail(OMI_ERR_ARG, "velocity_iterations exceeds the batch's max_solver_iterations");
    WorldSlot& s = b->slots[slot];
    Dev& d = b->d;
    const int32_t W = b->cfg.max_worlds;
    int32_t rc;
    HIPCHK(hipStreamSynchronize(b->stream));
    if (inv_mass && (rc = put(b, d.inv_mass + s.off, inv_mass, s.n))) return rc;
    if (inv_inertia && (rc = upload_soa(b, d.inv_inertia, d.ncap, s.off, inv_inertia, s.n, 3))) return rc;
    if (gravity_factor && (rc = put(b, d.gf + s.off, gravity_factor, s.n))) return rc;
    if (linear_damping && (rc = put(b, d.ld + s.off, linear_damping, s.n))) return rc;
    if (angular_damping && (rc = put(b, d.ad + s.off, angular_damping, s.n))) return rc;
    if (quadratic_drag && (rc = put(b, d.qd + s.off, quadratic_drag, s.n))) return rc;
    if (gravity)
        for (int r = 0; r < 3; r++)
            if ((rc = put(b, d.w_grav + (size_t)r * W + slot, &gravity[r], 1))) return rc;
    if (default_damping) {
        if ((rc = put(b, d.w_damp + slot, &default_damping[0], 1))) return rc;
        if ((rc = put(b, d.w_damp + W + slot, &default_damping[1], 1))) return rc;
    }
    if (velocity_iterations >= 0 && (rc = put(b, d.w_iters + slot, &velocity_iterations, 1))) return rc;
    if (sleep_enabled >= 0) {
        s.flags = sleep_enabled ? (s.flags | WF_SLEEP) : (s.flags & ~WF_SLEEP);
        if ((rc = put(b, d.w_flags + slot, &s.flags, 1))) return rc;
    }
    return OMI_OK;
}

// The warm-start cache of one world: the current buffer's entries of the slot,
// world-local, impulses as doubles.
extern "C" int32_t omi_batch_read_cache(omi_batch* b, int32_t slot, int64_t cap, int64_t* count, int32_t* a,
                                        int32_t* bb, double* point, double* nimp, double* timp) {
    if (!b || slot < 0 || slot >= b->cfg.max_worlds || !b->slots[slot].live || !count)
        return fail(OMI_ERR_ARG, "bad arguments");
    const WorldSlot& s = b->slots[slot];
    const Dev& d = b->d;
    const int32_t W = b->cfg.max_worlds;
    HIPCHK(hipStreamSynchronize(b->stream));
    int32_t sel = 0, rs = 0, rn = 0, rc;
    if ((rc = get(b, &sel, d.sel, 1))) return rc;
    if ((rc = get(b, &rs, d.w_cache_s + (size_t)sel * W + slot, 1))) return rc;
    if ((rc = get(b, &rn, d.w_cache_n + (size_t)sel * W + slot, 1))) return rc;
    *count = rn;
    const int64_t m = std::min<int64_t>(rn, cap);
    if (m <= 0) return OMI_OK;
    std::vector<int32_t> ia((size_t)m), ib((size_t)m);
    if ((rc = get(b, ia.data(), d.ka[sel] + rs, m))) return rc;
    if ((rc = get(b, ib.data(), d.kb[sel] + rs, m))) return rc;
    for (int64_t k = 0; k < m; k++) {
        if (a) a[k] = (int32_t)(ia[k] - s.off);
        if (bb) bb[k] = (int32_t)(ib[k] - s.off);
    }
    const size_t R = b->cfg.precision == OMI_FP64 ? 8 : 4;
    std::vector<double> tmp((size_t)m);
    std::vector<unsigned char> raw((size_t)m * R);
    for (int c = 0; c < 3; c++) {
        if (point) {
            if ((rc = get(b, tmp.data(), d.kpt[sel] + (size_t)c * d.kcap + rs, m))) return rc;
            for (int64_t k = 0; k < m; k++) point[3 * k + c] = tmp[k];
        }
        HIPCHK(hipMemcpy(raw.data(), (const unsigned char*)d.kimp[sel] + ((size_t)c * d.kcap + rs) * R, m * R,
                         hipMemcpyDeviceToHost));
        for (int64_t k = 0; k < m; k++) {
            const double v = R == 8 ? ((const double*)raw.data())[k] : (double)((const float*)raw.data())[k];
            if (c == 0 && nimp) nimp[k] = v;
            if (c > 0 && timp) timp[2 * k + c - 1] = v;
        }
    }
    return OMI_OK;
}

// Replace one world's warm-start cache: the entries are appended to the
// current buffer (after every world's entries of its last fill) and their
// group heads hashed, so the next step seeds from them as from its own cache.
// Entries of one (a, b) must be consecutive, as omi_batch_read_cache returns them.
extern "C" int32_t omi_batch_write_cache(omi_batch* b, int32_t slot, int64_t n, const int32_t* a, const int32_t* bb,
                                         const double* point, const double* nimp, const double* timp) {
    if (!b || slot < 0 || slot >= b->cfg.max_worlds || !b->slots[slot].live || n < 0 || (n > 0 && (!a || !bb || !point || !nimp || !timp)))
        return fail(OMI_ERR_ARG, "bad arguments");
    const WorldSlot& s = b->slots[slot];
    Dev& d = b->d;
    const int32_t W = b->cfg.max_worlds;
    std::vector<std::pair<int32_t, int32_t>> heads;
    for (int64_t k = 0; k < n; k++) {
        if (a[k] < 0 || a[k] >= s.n || bb[k] < 0 || bb[k] >= s.n) return fail(OMI_ERR_ARG, "cache body index out of range");
        if (k == 0 || a[k] != a[k - 1] || bb[k] != bb[k - 1]) heads.push_back({a[k], bb[k]});
    }
    std::sort(heads.begin(), heads.end());
    if (std::adjacent_find(heads.begin(), heads.end()) != heads.end())
        return fail(OMI_ERR_ARG, "cache entries of one body pair must be consecutive");
    HIPCHK(hipStreamSynchronize(b->stream));
    int32_t sel = 0, hn[2] = {0, 0}, rc;
    if ((rc = get(b, &sel, d.sel, 1))) return rc;
    if ((rc = get(b, hn, d.hslot_n, 2))) return rc;
    const int64_t base = hn[sel];
    if (base + n > d.kcap) return fail(OMI_ERR_CAPACITY, "warm-start cache capacity of the batch is full");
    int32_t zero = 0, cnt = (int32_t)n, b32 = (int32_t)base;
    if (n > 0) {
        std::vector<int32_t> ga((size_t)n), gb((size_t)n);
        for (int64_t k = 0; k < n; k++) { ga[k] = (int32_t)(s.off + a[k]); gb[k] = (int32_t)(s.off + bb[k]); }
        if ((rc = put(b, d.ka[sel] + base, ga.data(), n))) return rc;
        if ((rc = put(b, d.kb[sel] + base, gb.data(), n))) return rc;
        if ((rc = upload_soa(b, d.kpt[sel], d.kcap, base, point, n, 3))) return rc;
        const size_t R = b->cfg.precision == OMI_FP64 ? 8 : 4;
        std::vector<unsigned char> raw((size_t)n * R);
        for (int c = 0; c < 3; c++) {
            for (int64_t k = 0; k < n; k++) {
                const double v = c == 0 ? nimp[k] : timp[2 * k + c - 1];
                if (R == 8) ((double*)raw.data())[k] = v; else ((float*)raw.data())[k] = (float)v;
            }
            HIPCHK(hipMemcpy((unsigned char*)d.kimp[sel] + ((size_t)c * d.kcap + base) * R, raw.data(), n * R,
                             hipMemcpyHostToDevice));
        }
        hn[sel] = (int32_t)(base + n);
        if ((rc = put(b, d.hslot_n, hn, 2))) return rc;
    }
    if ((rc = put(b, d.w_cache_s + (size_t)sel * W + slot, n > 0 ? &b32 : &zero, 1))) return rc;
    if ((rc = put(b, d.w_cache_n + (size_t)sel * W + slot, &cnt, 1))) return rc;
    launch_cache_seed(d, base, (int32_t)n, b->stream);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(b->stream));
    return OMI_OK;
}

// World lists per size class.  Lists live in device memory and every class
// that has ever held a world keeps its launch, so loading or removing worlds
// does not invalidate a captured step graph unless a new class appears.
static int32_t rebuild_lists(omi_batch* b) {
    std::vector<int32_t> lists[4];
    for (int32_t w = 0; w < b->cfg.max_worlds; w++) {
        const WorldSlot& s = b->slots[w];
        if (!s.live) continue;
        int32_t NP = 1;
        while (NP < s.n) NP <<= 1;
        int c = NP <= b->np_class[0] ? 0 : NP <= b->np_class[1] ? 1 : NP <= b->np_class[2] ? 2 : 3;
        lists[c].push_back(w);
    }
    // largest worlds first within a class: their blocks are dispatched first,
    // so the longest solver (and sort) workgroups start before the short ones
    // rather than queueing behind them for CU space
    for (int c = 0; c < 4; c++)
        std::stable_sort(lists[c].begin(), lists[c].end(),
                         [&](int32_t x, int32_t y) { return b->slots[x].n > b->slots[y].n; });
    int32_t rc, counts[4];
    bool grew = false;
    // the largest world of class 3 sorts in LDS when its 16 bytes per body fit
    {
        int32_t big = 0;
        for (int32_t w : lists[3]) big = std::max(big, (int32_t)b->slots[w].n);
        big = (big + 63) & ~63;
        const int32_t np3 = big > 0 && (int64_t)big * 16 <= b->dyn_lds ? big : 0;
        if (np3 != b->sort_np[3]) { b->sort_np[3] = np3; grew = true; }
    }
    // A class is solved island by island when its worlds are too few to fill
    // the chip (fewer than one per 32 CUs) and every one of them fits the
    // island solve's LDS; else one block per world.  One class at most (the
    // largest that qualifies): the island table and its count are shared.
    int32_t isl_class = -1;
    for (int c = 3; c >= 0 && isl_class < 0 && b->cfg.island_solve >= 0; c--) {
        bool ok = !lists[c].empty() && (int64_t)lists[c].size() * 32 <= b->cus;
        for (int32_t w : lists[c]) ok = ok && islands_fit(b->slots[w].n, b->lds_bytes[c]);
        if (ok) isl_class = c;
    }
    for (int c = 0; c < 4; c++) {
        const int32_t isl = c == isl_class;
        if (isl != b->isl[c]) { b->isl[c] = isl; grew = true; }
    }
    for (int c = 0; c < 4; c++) {
        counts[c] = (int32_t)lists[c].size();
        st_copy(b, b->d_wl + (size_t)c * b->cfg.max_worlds, lists[c].data(), counts[c]);
        // a class's launches take a grid of its worlds rounded up to a power of
        // two, kept as a high-water mark so the captured graph changes rarely;
        // blocks past the class's count exit at once, but each still needs a
        // CU with the launch's LDS free
        int32_t g = 1;
        while (g < counts[c]) g <<= 1;
        if (counts[c] > 0 && g > b->live[c]) { b->live[c] = g; grew = true; }
    }
    st_copy(b, b->d_wn, counts, 4);
    if ((rc = st_flush(b))) return rc;
    int32_t nlive = 0;
    for (int c = 0; c < 4; c++) nlive += b->live[c] > 0;
    for (int q = 0; q < 3 && nlive >= 2; q++)
        if (!b->aux[q]) HIPCHK(hipStreamCreateWithFlags(&b->aux[q], hipStreamNonBlocking));
    b->nhw = b->d.ncap;
    b->lists_dirty = false;
    if (grew) invalidate_graph(b);
    return OMI_OK;
}

// Launches of the size classes c in `live`, largest first: the largest on the
// batch stream, the others on branch streams forked and joined with events,
// so a captured graph runs them side by side.
template <typename F, typename G>
static void fork_classes(omi_batch* b, F launch_class, G after_top) {
    hipStream_t s = b->stream;
    int top = -1;
    for (int c = 3; c >= 0; c--)
        if (b->live[c]) { top = c; break; }
    if (top < 0) return;
    int used[3] = {0, 0, 0}, q = 0;
    hipEventRecord(b->fork, s);
    for (int c = top; c >= 0; c--) {
        if (!b->live[c]) continue;
        hipStream_t sc = s;
        if (c != top) {
            sc = b->aux[q];
            hipStreamWaitEvent(sc, b->fork, 0);
        }
        launch_class(c, sc);
        if (c == top) after_top();
        if (c != top) { hipEventRecord(b->join[q], sc); used[q] = 1; q++; }
    }
    for (int k = 0; k < 3; k++)
        if (used[k]) hipStreamWaitEvent(s, b->join[k], 0);
}

// Broadphase: the size classes' sorts side by side, then the chip-wide passes.
static void enqueue_sap(omi_batch* b, int32_t record) {
    const Dev& d = b->d;
    fork_classes(b, [&](int c, hipStream_t sc) {
        launch_sap_sort(d, c, b->d_wl + (size_t)c * d.wcap, b->d_wn + c, b->sort_np[c], b->sap_block[c],
                        b->live[c], sc);
    }, [] {});
    launch_sap_pairs(d, record, b->sap_grid, b->stream);
}

// The solver rows (with in.wake, the wake pass rides on them) and the size
// classes' solve launches, largest class first.  The largest class's worlds
// set the step's critical path, so its solve is launched first, on the batch
// stream, and takes its CUs before the smaller classes' blocks; its blocks
// write their own worlds' rows while their first wavefront assigns the levels
// (a class of 64-thread blocks, or solved island by island, gets its rows from
// a launch of its own first).  The next class runs beside it on a first
// branch stream, after a launch that writes that class's rows (a few tens of
// microseconds: the largest class's blocks, each taking a whole CU, are all
// placed before the next class's blocks spread over the CUs).  After that
// launch, the remaining classes' rows, then their solves, run on a second and
// a third branch (so each class's blocks are placed before a smaller class's),
// and `early` (given the stream it runs on) on the third, before it waits for
// those rows; returns whether `early` was launched.  With stage timing on (or
// one class) everything runs in stage order, each stage timed alone: the other
// classes' rows, the largest class's solve with its rows (marked), the others.
template <typename Early, typename Mark>
static bool enqueue_solve(omi_batch* b, SolveIn in, bool stage, Early early, Mark mark) {
    const Dev& d = b->d;
    hipStream_t s = b->stream;
    const int prec = b->cfg.precision;
    auto solve_class = [&](int c, hipStream_t sc, bool rows) {
        const int32_t* wl = b->d_wl + (size_t)c * d.wcap;
        if (b->isl[c]) {
            // the class's islands, then two blocks per CU taking them in turn
            if (rows) launch_prepare_worlds(d, prec, in, stage, wl, b->d_wn + c, b->live[c], sc);
            launch_islands(d, wl, b->d_wn + c, b->lds_bytes[c], b->live[c], sc);
            launch_solve_class(d, prec, b->solve_block[c], in, wl, b->d_wn + c, b->lds_bytes[c],
                               std::max(b->live[c], 2 * b->cus), sc, 1);
        } else if (rows && b->solve_block[c] == 64) {
            launch_prepare_worlds(d, prec, in, stage, wl, b->d_wn + c, b->live[c], sc);
            launch_solve_class(d, prec, b->solve_block[c], in, wl, b->d_wn + c, b->lds_bytes[c], b->live[c], sc, 0);
        } else {
            launch_solve_class(d, prec, b->solve_block[c], in, wl, b->d_wn + c, b->lds_bytes[c], b->live[c], sc,
                               rows ? 2 : 0);
        }
    };
    int top = -1, nclass = 0, rest[4], nrest = 0;
    for (int c = 3; c >= 0; c--) {
        if (!b->live[c]) continue;
        if (top < 0) top = c;
        else rest[nrest++] = c;
        nclass++;
    }
    if (b->timing || nclass < 2) {
        // (the largest class's blocks still write their own rows, so its
        // solve launch is the one the step graph runs)
        const bool own = nclass >= 2;
        // class c's worlds are those of more bodies than the next smaller
        // class's power-of-two bound
        launch_prepare(d, prec, in, stage, s, own ? b->np_class[rest[0]] : 0);
        mark();
        fork_classes(b, [&](int c, hipStream_t sc) { solve_class(c, sc, own && c == top); }, mark);
        return false;
    }
    // (the batch stream's launches are captured first: a graph runs the first
    // branch of a fork on its launch stream and the later ones on streams of
    // their own)
    hipEventRecord(b->fork, s);
    solve_class(top, s, true);
    mark();
    hipStreamWaitEvent(b->aux[0], b->fork, 0);
    launch_prepare_worlds(d, prec, in, stage, b->d_wl + (size_t)rest[0] * d.wcap, b->d_wn + rest[0],
                          b->live[rest[0]], b->aux[0]);
    hipEventRecord(b->next, b->aux[0]);
    solve_class(rest[0], b->aux[0], false);
    hipEventRecord(b->join[0], b->aux[0]);
    bool used[3] = {true, false, false};
    if (nrest >= 2) {
        hipStreamWaitEvent(b->aux[1], b->next, 0);
        launch_prepare(d, prec, in, stage, b->aux[1], b->np_class[rest[1]]);
        hipEventRecord(b->rows, b->aux[1]);
        solve_class(rest[1], b->aux[1], false);
        hipEventRecord(b->join[1], b->aux[1]);
        used[1] = true;
    }
    hipStreamWaitEvent(b->aux[2], b->next, 0);
    early(b->aux[2]);
    if (nrest >= 3) {
        hipStreamWaitEvent(b->aux[2], b->rows, 0);
        solve_class(rest[2], b->aux[2], false);
    }
    hipEventRecord(b->join[2], b->aux[2]);
    used[2] = true;
    for (int k = 0; k < 3; k++)
        if (used[k]) hipStreamWaitEvent(s, b->join[k], 0);
    return true;
}

// The warm-start cache rebuild, then joints, positions and sleep; without
// stage timing the cache rebuild runs on a branch stream beside the others
// (it reads the contacts and their impulses, they the velocities) and is
// joined before the cache buffers toggle.  `early`: the cache plan and the
// sleep pass's island union already ran beside the solve.
template <typename Mark>
static void enqueue_after_solve(omi_batch* b, double dt, bool early, Mark mark) {
    const Dev& d = b->d;
    hipStream_t s = b->stream;
    const int prec = b->cfg.precision;
    const bool side = !b->timing && b->aux[0];
    hipStream_t cs = side ? b->aux[0] : s;
    if (side) {
        hipEventRecord(b->solved, s);
        hipStreamWaitEvent(cs, b->solved, 0);
    }
    if (!early) launch_cache_plan(d, cs);
    launch_cache_write(d, prec, cs);
    if (side) hipEventRecord(b->cached, cs);
    mark();
    if (b->any_triggers) launch_trigger(d, s);
    if (b->any_joints) launch_joints(d, prec, dt, b->joint_lds, s);
    mark();
    if (b->timing) {
        launch_post(d, b->nhw, dt, s); mark();
        if (!early) launch_sleep_union(d, b->nhw, s);
        launch_sleep_decide(d, b->nhw, dt, s, 0.08, 0.25, 0.4); mark();
    } else {
        // one launch for positions and the speed tests (the stage times
        // keep them apart)
        if (!early) launch_sleep_union(d, b->nhw, s);
        launch_post_sleep(d, b->nhw, dt, s, 0.08, 0.25, 0.4);
    }
    if (side) hipStreamWaitEvent(s, b->cached, 0);
    launch_toggle(d, s); mark();
}

static void enqueue_step(omi_batch* b, double dt) {
    const Dev& d = b->d;
    hipStream_t s = b->stream;
    const int prec = b->cfg.precision;
    const bool tm = b->timing && !b->ev.empty();
    int e = 0;
    auto mark = [&]() { if (tm) hipEventRecord(b->ev[e++], s); };
    mark();
    launch_begin(d, s); mark();
    launch_pre(d, prec, b->nhw, dt, 0.05, s, false); mark();
    enqueue_sap(b, b->cfg.record_broadphase);
    mark();
    launch_narrow(d, prec, s); mark();
    launch_compact(d, prec, b->sap_grid, s); mark();
    SolveIn in{nullptr, nullptr, 0.5, 0.8, 0.005, 1, 1};
    // the cache plan and the sleep pass's island union need only the step's
    // contacts: they run beside the solve when a branch is free
    const bool early = enqueue_solve(b, in, false,
                                     [&](hipStream_t es) {
                                         launch_sleep_union(d, b->nhw, es);
                                         launch_cache_plan(d, es);
                                     }, mark);
    mark();
    enqueue_after_solve(b, dt, early, mark);
}

// The collision stage alone (_CollisionStages.run, world.py:657-676): the host
// has integrated forces and refitted the AABBs and will solve joints,
// integrate positions and update sleep itself.
static void enqueue_collide(omi_batch* b, double dt) {
    const Dev& d = b->d;
    hipStream_t s = b->stream;
    const int prec = b->cfg.precision;
    launch_begin(d, s);
    launch_pre(d, prec, b->nhw, dt, 0.05, s, true);
    enqueue_sap(b, b->cfg.record_broadphase);
    launch_narrow(d, prec, s);
    launch_compact(d, prec, b->sap_grid, s);
    SolveIn in{nullptr, nullptr, 0.5, 0.8, 0.005, 1, 1};
    // (the cache plan needs only the step's contacts)
    const bool early = enqueue_solve(b, in, false, [&](hipStream_t es) { launch_cache_plan(d, es); }, [] {});
    if (!early) launch_cache_plan(d, s);
    launch_cache_write(d, prec, s);
    if (b->any_triggers) launch_trigger(d, s);
    launch_toggle(d, s);
}

extern "C" int32_t omi_batch_set_timing(omi_batch* b, int32_t on) {
    if (!b) return fail(OMI_ERR_ARG, "null batch");
    b->timing = on != 0;
    invalidate_graph(b);
    if (b->timing && b->ev.empty()) {
        b->ev.resize(kNumKernels + 1);
        for (auto& e : b->ev) HIPCHK(hipEventCreate(&e));
    }
    return OMI_OK;
}

extern "C" const char* omi_batch_kernel_names(void) { return kKernelNames; }

extern "C" int32_t omi_batch_kernel_times(omi_batch* b, int32_t n, double* ms) {
    if (!b || !ms) return fail(OMI_ERR_ARG, "null argument");
    if (b->ev.empty()) return fail(OMI_ERR_STATE, "timing is off");
    HIPCHK(hipStreamSynchronize(b->stream));
    for (int i = 0; i < n && i < kNumKernels; i++) {
        float t = 0.f;
        HIPCHK(hipEventElapsedTime(&t, b->ev[i], b->ev[i + 1]));
        ms[i] = t;
    }
    return OMI_OK;
}

extern "C" int32_t omi_batch_world_stats(omi_batch* b, int32_t n, int32_t* bodies, int32_t* contacts,
                                         int32_t* manifolds, int32_t* levels, int32_t* size_class) {
    if (!b || n < 0 || n > b->cfg.max_worlds) return fail(OMI_ERR_ARG, "bad arguments");
    HIPCHK(hipStreamSynchronize(b->stream));
    const Dev& d = b->d;
    std::vector<int32_t> fl((size_t)std::max(n, 1));
    int32_t rc;
    if (n == 0) return OMI_OK;
    if ((rc = get(b, fl.data(), d.w_flags, n))) return rc;
    if (bodies && (rc = get(b, bodies, d.w_cnt, n))) return rc;
    if (contacts && (rc = get(b, contacts, d.w_cn, n))) return rc;
    if (manifolds && (rc = get(b, manifolds, d.w_mn, n))) return rc;
    if (levels && (rc = get(b, levels, d.w_levels, n))) return rc;
    for (int32_t w = 0; w < n; w++) {
        const bool live = b->slots[w].live;
        if (!live) {
            if (bodies) bodies[w] = 0;
            if (contacts) contacts[w] = 0;
            if (manifolds) manifolds[w] = 0;
            if (levels) levels[w] = 0;
        }
        if (size_class) {
            int32_t NP = 1;
            while (NP < b->slots[w].n) NP <<= 1;
            size_class[w] = !live ? -1 : NP <= b->np_class[0] ? 0 : NP <= b->np_class[1] ? 1 : NP <= b->np_class[2] ? 2 : 3;
        }
    }
    return OMI_OK;
}

extern "C" int32_t omi_batch_set_blocks(omi_batch* b, const int32_t* sap_block, const int32_t* solve_block) {
    if (!b || !sap_block || !solve_block) return fail(OMI_ERR_ARG, "null argument");
    auto okb = [](int32_t x) { return x == 0 || x == 64 || x == 128 || x == 256 || x == 512 || x == 1024; };
    for (int c = 0; c < 4; c++) {
        if (!okb(sap_block[c]) || !okb(solve_block[c]))
            return fail(OMI_ERR_ARG, "block sizes must be 0, 64, 128, 256, 512 or 1024");
        if (b->cfg.precision == OMI_FP64 && solve_block[c] > FP64_SOLVE_BLOCK_MAX)
            return fail(OMI_ERR_ARG, "FP64 solver blocks are 64, 128 or 256 threads");
    }
    for (int c = 0; c < 4; c++) {
        if (sap_block[c]) b->sap_block[c] = sap_block[c];
        if (solve_block[c]) b->solve_block[c] = solve_block[c];
    }
    invalidate_graph(b);
    return OMI_OK;
}

extern "C" int32_t omi_batch_use_graph(omi_batch* b, int32_t on) {
    if (!b) return fail(OMI_ERR_ARG, "null batch");
    b->use_graph = on != 0;
    if (!b->use_graph) invalidate_graph(b);
    return OMI_OK;
}

// Capture the step (with its timing events when timing is on) in a HIP graph.
static int32_t ensure_graph(omi_batch* b, double dt) {
    if (b->graph && b->graph_dt == dt) return OMI_OK;
    invalidate_graph(b);
    // capture starts on an idle stream (staged writes may be queued on it)
    HIPCHK(hipStreamSynchronize(b->stream));
    hipGraph_t g;
    HIPCHK(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
    enqueue_step(b, dt);
    HIPCHK(hipStreamEndCapture(b->stream, &g));
    hipError_t e = hipGraphInstantiate(&b->graph, g, nullptr, nullptr, 0);
    hipGraphDestroy(g);
    if (e != hipSuccess) { b->graph = nullptr; return fail(OMI_ERR_HIP, std::string("graph instantiate: ") + hipGetErrorString(e)); }
    b->graph_dt = dt;
    return OMI_OK;
}

extern "C" int32_t omi_batch_prepare(omi_batch* b, double dt) {
    if (!b) return fail(OMI_ERR_ARG, "null batch");
    if (!(dt > 0.0)) return fail(OMI_ERR_ARG, "dt must be positive");
    int32_t rc;
    if (b->lists_dirty && (rc = rebuild_lists(b))) return rc;
    if (b->use_graph && !b->timing) return ensure_graph(b, dt);
    return OMI_OK;
}

extern "C" int32_t omi_batch_step(omi_batch* b, double dt, int32_t nsteps) {
    if (!b) return fail(OMI_ERR_ARG, "null batch");
    if (!(dt > 0.0)) return fail(OMI_ERR_ARG, "dt must be positive");
    int32_t rc;
    if (b->lists_dirty && (rc = rebuild_lists(b))) return rc;
    for (int32_t k = 0; k < nsteps; k++) {
        // per-kernel timing events cannot be timed inside a captured graph: timed
        // steps launch eagerly
        if (b->use_graph && !b->timing) {
            if ((rc = ensure_graph(b, dt))) return rc;
            HIPCHK(hipGraphLaunch(b->graph, b->stream));
        } else {
            enqueue_step(b, dt);
            HIPCHK(hipGetLastError());
        }
    }
    return OMI_OK;
}

extern "C" int32_t omi_batch_collide(omi_batch* b, double dt) {
    if (!b) return fail(OMI_ERR_ARG, "null batch");
    if (!(dt > 0.0)) return fail(OMI_ERR_ARG, "dt must be positive");
    int32_t rc;
    if (b->lists_dirty && (rc = rebuild_lists(b))) return rc;
    enqueue_collide(b, dt);
    HIPCHK(hipGetLastError());
    return OMI_OK;
}

// ---- groups: batches stepped together ---------------------------------
//
// Each member's step runs on the member's own stream (its captured step graph,
// or its kernels launched eagerly when it times them), forked from the group's
// stream and joined back at the end of the step.  The members' streams sit on
// different hardware queues, so their kernels overlap on the device: a member
// holding a few large worlds runs its long solver launch while another
// member's chip-wide stages run over many small worlds, instead of every world
// waiting at the same stage boundaries.

struct omi_group {
    std::vector<omi_batch*> m;
    std::vector<hipEvent_t> join;          // per member
    hipStream_t stream = nullptr;
    hipEvent_t fork = nullptr;
};

static void group_free(omi_group* g) {
    for (auto e : g->join) if (e) hipEventDestroy(e);
    if (g->fork) hipEventDestroy(g->fork);
    if (g->stream) hipStreamDestroy(g->stream);
}

extern "C" int32_t omi_group_create(omi_batch* const* members, int32_t n, omi_group** out) {
    if (!members || !out || n <= 0) return fail(OMI_ERR_ARG, "a group needs at least one batch");
    for (int32_t i = 0; i < n; i++) {
        if (!members[i]) return fail(OMI_ERR_ARG, "null member batch");
        for (int32_t j = 0; j < i; j++)
            if (members[j] == members[i]) return fail(OMI_ERR_ARG, "a batch appears twice in the group");
    }
    omi_group* g = new omi_group();
    g->m.assign(members, members + n);
    g->join.assign((size_t)n, nullptr);
    hipError_t e = hipStreamCreateWithFlags(&g->stream, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&g->fork, hipEventDisableTiming);
    for (int32_t i = 0; i < n && e == hipSuccess; i++) e = hipEventCreateWithFlags(&g->join[i], hipEventDisableTiming);
    if (e != hipSuccess) {
        group_free(g);
        delete g;
        return fail(OMI_ERR_HIP, std::string("group streams/events: ") + hipGetErrorString(e));
    }
    *out = g;
    return OMI_OK;
}

extern "C" int32_t omi_group_destroy(omi_group* g) {
    if (!g) return OMI_OK;
    hipStreamSynchronize(g->stream);
    group_free(g);
    delete g;
    return OMI_OK;
}

static int32_t group_ready(omi_group* g, double dt) {
    int32_t rc;
    for (omi_batch* b : g->m) {
        if (b->lists_dirty && (rc = rebuild_lists(b))) return rc;
        if (b->use_graph && !b->timing && (rc = ensure_graph(b, dt))) return rc;
    }
    return OMI_OK;
}

extern "C" int32_t omi_group_prepare(omi_group* g, double dt) {
    if (!g) return fail(OMI_ERR_ARG, "null group");
    if (!(dt > 0.0)) return fail(OMI_ERR_ARG, "dt must be positive");
    return group_ready(g, dt);
}

extern "C" int32_t omi_group_step(omi_group* g, double dt, int32_t nsteps) {
    if (!g) return fail(OMI_ERR_ARG, "null group");
    if (!(dt > 0.0)) return fail(OMI_ERR_ARG, "dt must be positive");
    int32_t rc;
    if ((rc = group_ready(g, dt))) return rc;
    // the steps start after work already queued on the members' streams
    // (uploads, state changes)
    for (size_t i = 0; i < g->m.size(); i++) {
        HIPCHK(hipEventRecord(g->join[i], g->m[i]->stream));
        HIPCHK(hipStreamWaitEvent(g->stream, g->join[i], 0));
    }
    for (int32_t k = 0; k < nsteps; k++) {
        HIPCHK(hipEventRecord(g->fork, g->stream));
        for (size_t i = 0; i < g->m.size(); i++) {
            omi_batch* b = g->m[i];
            HIPCHK(hipStreamWaitEvent(b->stream, g->fork, 0));
            if (b->use_graph && !b->timing) {
                HIPCHK(hipGraphLaunch(b->graph, b->stream));
            } else {
                enqueue_step(b, dt);
                HIPCHK(hipGetLastError());
            }
            HIPCHK(hipEventRecord(g->join[i], b->stream));
            HIPCHK(hipStreamWaitEvent(g->stream, g->join[i], 0));
        }
    }
    return OMI_OK;
}

extern "C" int32_t omi_group_sync(omi_group* g) {
    if (!g) return fail(OMI_ERR_ARG, "null group");
    HIPCHK(hipStreamSynchronize(g->stream));
    // every member's capacity flags, reporting the first member that overflowed
    int32_t first = OMI_OK;
    std::string msg;
    for (omi_batch* b : g->m) {
        int32_t rc = omi_batch_sync(b);
        if (rc && first == OMI_OK) { first = rc; msg = g_err; }
    }
    if (first) return fail(first, msg);
    return OMI_OK;
}

extern "C" int32_t omi_batch_refit(omi_batch* b, double margin) {
    if (!b) return fail(OMI_ERR_ARG, "null batch");
    int32_t rc;
    if (b->lists_dirty && (rc = rebuild_lists(b))) return rc;
    launch_refit(b->d, b->nhw, margin, b->stream);
    HIPCHK(hipGetLastError());
    return OMI_OK;
}

extern "C" int32_t omi_batch_sync(omi_batch* b) {
    if (!b) return fail(OMI_ERR_ARG, "null batch");
    HIPCHK(hipStreamSynchronize(b->stream));
    uint64_t ov = 0;
    HIPCHK(hipMemcpy(&ov, b->d.cnt + C_OVERFLOW, sizeof(uint64_t), hipMemcpyDeviceToHost));
    if (ov) {
        // the flags are sticky over the steps since the last sync; report and clear them
        HIPCHK(hipMemset(b->d.cnt + C_OVERFLOW, 0, sizeof(uint64_t)));
        return capacity_error(ov);
    }
    return OMI_OK;
}

extern "C" int32_t omi_batch_islands(omi_batch* b, int32_t* size_class, int64_t* islands) {
    if (!b || !size_class || !islands) return fail(OMI_ERR_ARG, "null argument");
    uint64_t c = 0;
    HIPCHK(hipStreamSynchronize(b->stream));
    HIPCHK(hipMemcpy(&c, b->d.cnt + C_ISLANDS, sizeof c, hipMemcpyDeviceToHost));
    *size_class = -1;
    for (int k = 0; k < 4; k++)
        if (b->isl[k]) *size_class = k;
    *islands = *size_class >= 0 ? (int64_t)c : 0;
    return OMI_OK;
}

extern "C" int32_t omi_batch_counters(omi_batch* b, int64_t* out8) {
    if (!b || !out8) return fail(OMI_ERR_ARG, "null argument");
    uint64_t c[C_COUNT];
    HIPCHK(hipStreamSynchronize(b->stream));
    HIPCHK(hipMemcpy(c, b->d.cnt, sizeof c, hipMemcpyDeviceToHost));
    int64_t bodies = 0;
    for (auto& s : b->slots) if (s.live) bodies += s.n;
    out8[0] = bodies; out8[1] = (int64_t)c[C_XCAND]; out8[2] = (int64_t)c[C_BPTOTAL]; out8[3] = (int64_t)c[C_PAIRS];
    out8[4] = (int64_t)c[C_MANIFOLDS]; out8[5] = (int64_t)c[C_CONTACTS]; out8[6] = (int64_t)c[C_MAXLEVEL];
    out8[7] = (int64_t)c[C_OVERFLOW];
    return OMI_OK;
}

extern "C" int32_t omi_batch_read_world(omi_batch* b, int32_t slot, double* pos, double* quat, double* lv,
                                        double* av, uint8_t* awake, double* timer, double* lo, double* hi,
                                        double* ppos, double* pquat) {
    if (!b || slot < 0 || slot >= b->cfg.max_worlds || !b->slots[slot].live) return fail(OMI_ERR_ARG, "bad world slot");
    const WorldSlot& s = b->slots[slot];
    const Dev& d = b->d;
    HIPCHK(hipStreamSynchronize(b->stream));
    int32_t rc;
    if (pos && (rc = download_soa(b, pos, d.pos, d.ncap, s.off, s.n, 3))) return rc;
    if (quat && (rc = download_soa(b, quat, d.quat, d.ncap, s.off, s.n, 4))) return rc;
    if (lv && (rc = download_soa(b, lv, d.lv, d.ncap, s.off, s.n, 3))) return rc;
    if (av && (rc = download_soa(b, av, d.av, d.ncap, s.off, s.n, 3))) return rc;
    if (lo && (rc = download_soa(b, lo, d.lo, d.ncap, s.off, s.n, 3))) return rc;
    if (hi && (rc = download_soa(b, hi, d.hi, d.ncap, s.off, s.n, 3))) return rc;
    if (ppos && (rc = download_soa(b, ppos, d.ppos, d.ncap, s.off, s.n, 3))) return rc;
    if (pquat && (rc = download_soa(b, pquat, d.pquat, d.ncap, s.off, s.n, 4))) return rc;
    if (awake && (rc = get(b, awake, d.awake + s.off, s.n))) return rc;
    if (timer && (rc = get(b, timer, d.timer + s.off, s.n))) return rc;
    return OMI_OK;
}

// ---- asynchronous readback -------------------------------------------------
//
// A fetch gathers one world's state on the device (staged copy segments) into
// a buffer of its own and copies it to pinned host memory, all ordered on the
// batch stream after the queued steps; the host collects it later.  Buffer
// layout: 27 doubles per body, component-major (pos 3, quat 4, lv 3, av 3,
// timer, aabb lo 3, hi 3, prev pos 3, prev quat 4), then the awake bytes.

static const int kFetchDoubles = 27;

extern "C" int32_t omi_batch_fetch_world(omi_batch* b, int32_t slot, int64_t* ticket) {
    if (!b || !ticket) return fail(OMI_ERR_ARG, "null argument");
    if (slot < 0 || slot >= b->cfg.max_worlds || !b->slots[slot].live) return fail(OMI_ERR_ARG, "bad world slot");
    const WorldSlot& s = b->slots[slot];
    const Dev& d = b->d;
    const int64_t n = s.n;
    const int64_t need = al16h(8 * kFetchDoubles * std::max<int64_t>(n, 1)) + al16h(std::max<int64_t>(n, 1));
    // a free pooled buffer that fits, else a new one
    int32_t bi = -1;
    for (size_t i = 0; i < b->fbufs.size(); i++)
        if (!b->fbuf_busy[i] && b->fbufs[i].cap >= need && (bi < 0 || b->fbufs[i].cap < b->fbufs[bi].cap))
            bi = (int32_t)i;
    if (bi < 0) {
        {
            omi_batch::FetchBuf f;
            f.cap = std::max<int64_t>(need, 1 << 16);
            HIPCHK(hipHostMalloc((void**)&f.host, (size_t)f.cap, hipHostMallocDefault));
            if (hipMalloc((void**)&f.dev, (size_t)f.cap) != hipSuccess) {
                hipHostFree(f.host);
                return fail(OMI_ERR_HIP, "fetch buffer: device allocation failed");
            }
            b->fbufs.push_back(f);
            b->fbuf_busy.push_back(0);
            bi = (int32_t)b->fbufs.size() - 1;
        }
    }
    int32_t fi = -1;
    for (size_t i = 0; i < b->fetches.size(); i++)
        if (!b->fetches[i].live) { fi = (int32_t)i; break; }
    if (fi < 0) {
        omi_batch::Fetch f;
        HIPCHK(hipEventCreateWithFlags(&f.ev, hipEventDisableTiming));
        b->fetches.push_back(f);
        fi = (int32_t)b->fetches.size() - 1;
    }
    double* out = (double*)b->fbufs[bi].dev;
    const double* srcs[9] = {d.pos, d.quat, d.lv, d.av, d.timer, d.lo, d.hi, d.ppos, d.pquat};
    const int comps[9] = {3, 4, 3, 3, 1, 3, 3, 3, 4};
    int64_t q = 0;
    for (int a = 0; a < 9; a++)
        for (int c = 0; c < comps[a]; c++, q++) st_dcopy(b, out + q * n, srcs[a] + (size_t)c * d.ncap + s.off, n);
    st_dcopy(b, (uint8_t*)(out + kFetchDoubles * n), d.awake + s.off, n);
    int32_t rc;
    if ((rc = st_flush(b))) return rc;
    HIPCHK(hipMemcpyAsync(b->fbufs[bi].host, b->fbufs[bi].dev, (size_t)need, hipMemcpyDeviceToHost, b->stream));
    HIPCHK(hipEventRecord(b->fetches[fi].ev, b->stream));
    b->fetches[fi].n = n;
    b->fetches[fi].buf = bi;
    b->fetches[fi].live = true;
    b->fbuf_busy[bi] = 1;
    *ticket = fi;
    return OMI_OK;
}

extern "C" int32_t omi_batch_fetch_ready(omi_batch* b, int64_t ticket) {
    if (!b || ticket < 0 || ticket >= (int64_t)b->fetches.size() || !b->fetches[ticket].live)
        return fail(OMI_ERR_ARG, "bad fetch ticket");
    const hipError_t e = hipEventQuery(b->fetches[ticket].ev);
    if (e == hipSuccess) return 1;
    if (e == hipErrorNotReady) return 0;
    return fail(OMI_ERR_HIP, std::string("fetch: ") + hipGetErrorString(e));
}

extern "C" int32_t omi_batch_fetch_wait(omi_batch* b, int64_t ticket, double* pos, double* quat, double* lv,
                                        double* av, uint8_t* awake, double* timer, double* lo, double* hi,
                                        double* ppos, double* pquat) {
    if (!b || ticket < 0 || ticket >= (int64_t)b->fetches.size() || !b->fetches[ticket].live)
        return fail(OMI_ERR_ARG, "bad fetch ticket");
    omi_batch::Fetch& f = b->fetches[ticket];
    HIPCHK(hipEventSynchronize(f.ev));
    const int64_t n = f.n;
    const double* src = (const double*)b->fbufs[f.buf].host;
    double* dsts[9] = {pos, quat, lv, av, timer, lo, hi, ppos, pquat};
    const int comps[9] = {3, 4, 3, 3, 1, 3, 3, 3, 4};
    int64_t q = 0;
    for (int a = 0; a < 9; a++) {
        for (int c = 0; c < comps[a]; c++, q++) {
            if (!dsts[a]) continue;
            const double* col = src + q * n;
            for (int64_t i = 0; i < n; i++) dsts[a][i * comps[a] + c] = col[i];
        }
    }
    if (awake) std::memcpy(awake, src + kFetchDoubles * n, (size_t)n);
    b->fbuf_busy[f.buf] = 0;
    f.live = false;
    f.buf = -1;
    return OMI_OK;
}

// ---- markers -----------------------------------------------------------------

static int32_t capacity_error(uint64_t ov) {
    char msg[320];
    snprintf(msg, sizeof msg,
             "batch capacity exceeded in a step since the last sync (flags 0x%llx: 1 pairs, 2 contacts, "
             "4 cache, 8 broadphase record, 16 broadphase candidates, 32 trigger pairs, 64 island LDS); results "
             "of that step are invalid — recreate the batch with larger max_pairs/max_contacts/max_candidates",
             (unsigned long long)ov);
    return fail(OMI_ERR_CAPACITY, msg);
}

extern "C" int32_t omi_batch_mark(omi_batch* b, int64_t* ticket) {
    if (!b || !ticket) return fail(OMI_ERR_ARG, "null argument");
    int32_t mi = -1;
    for (size_t i = 0; i < b->marks.size(); i++)
        if (!b->marks[i].live) { mi = (int32_t)i; break; }
    if (mi < 0) {
        omi_batch::Mark m;
        HIPCHK(hipEventCreateWithFlags(&m.ev, hipEventDisableTiming));
        if (hipHostMalloc((void**)&m.flags, sizeof(uint64_t), hipHostMallocDefault) != hipSuccess) {
            hipEventDestroy(m.ev);
            return fail(OMI_ERR_HIP, "marker: pinned allocation failed");
        }
        b->marks.push_back(m);
        mi = (int32_t)b->marks.size() - 1;
    }
    omi_batch::Mark& m = b->marks[mi];
    HIPCHK(hipMemcpyAsync(m.flags, b->d.cnt + C_OVERFLOW, sizeof(uint64_t), hipMemcpyDeviceToHost, b->stream));
    HIPCHK(hipEventRecord(m.ev, b->stream));
    m.live = true;
    *ticket = mi;
    return OMI_OK;
}

extern "C" int32_t omi_batch_mark_wait(omi_batch* b, int64_t ticket) {
    if (!b || ticket < 0 || ticket >= (int64_t)b->marks.size() || !b->marks[ticket].live)
        return fail(OMI_ERR_ARG, "bad marker ticket");
    omi_batch::Mark& m = b->marks[ticket];
    HIPCHK(hipEventSynchronize(m.ev));
    m.live = false;
    const uint64_t ov = *m.flags;
    return ov ? capacity_error(ov) : OMI_OK;
}

extern "C" int32_t omi_batch_read_trigger_overlaps(omi_batch* b, int32_t slot, int64_t cap, int64_t* count,
                                                   int32_t* trigger, int32_t* other) {
    if (!b || slot < 0 || slot >= b->cfg.max_worlds || !b->slots[slot].live || !count)
        return fail(OMI_ERR_ARG, "bad arguments");
    const WorldSlot& s = b->slots[slot];
    const Dev& d = b->d;
    HIPCHK(hipStreamSynchronize(b->stream));
    uint64_t tot = 0;
    int32_t rc;
    if ((rc = get(b, &tot, d.cnt + C_TOVER, 1))) return rc;
    const int64_t T = tot < (uint64_t)d.tpcap ? (int64_t)tot : d.tpcap;
    std::vector<int32_t> ta((size_t)T), tb((size_t)T);
    if ((rc = get(b, ta.data(), d.tva, T))) return rc;
    if ((rc = get(b, tb.data(), d.tvb, T))) return rc;
    std::vector<std::pair<int32_t, int32_t>> mine;
    for (int64_t k = 0; k < T; k++)
        if (ta[k] >= s.off && ta[k] < s.off + s.n)
            mine.push_back({ta[k] - (int32_t)s.off, tb[k] - (int32_t)s.off});
    std::sort(mine.begin(), mine.end());
    *count = (int64_t)mine.size();
    for (int64_t k = 0; k < cap && k < (int64_t)mine.size(); k++) {
        if (trigger) trigger[k] = mine[k].first;
        if (other) other[k] = mine[k].second;
    }
    return OMI_OK;
}

// Bytes per snapshot row: position 3 and axis-angle 4 doubles, awake and dynamic.
static const int64_t kSnapRow = 7 * 8 + 2;

extern "C" int32_t omi_batch_snapshot(omi_batch* b, int32_t nslots, const int32_t* slots, int64_t* offsets,
                                      double* pos, double* axis_angle, uint8_t* awake, uint8_t* dynamic) {
    if (!b || nslots < 0 || (nslots > 0 && (!slots || !offsets))) return fail(OMI_ERR_ARG, "bad arguments");
    int64_t total = 0;
    for (int32_t k = 0; k < nslots; k++) {
        const int32_t w = slots[k];
        if (w < 0 || w >= b->cfg.max_worlds || !b->slots[w].live) return fail(OMI_ERR_ARG, "bad world slot");
        offsets[k] = total;
        total += b->slots[w].n;
    }
    if (offsets) offsets[nslots] = total;
    if (!pos && !axis_angle && !awake && !dynamic) return OMI_OK;
    if (total == 0) return OMI_OK;
    HIPCHK(hipStreamSynchronize(b->stream));
    if (total > b->snap_cap) {
        if (b->snap_dev) { HIPCHK(hipFree(b->snap_dev)); b->snap_dev = nullptr; }
        if (b->snap_host) { HIPCHK(hipHostFree(b->snap_host)); b->snap_host = nullptr; }
        b->snap_cap = 0;
        const int64_t cap = std::max<int64_t>(total, 1024);
        HIPCHK(hipMalloc((void**)&b->snap_dev, (size_t)(cap * kSnapRow)));
        HIPCHK(hipHostMalloc((void**)&b->snap_host, (size_t)(cap * kSnapRow), hipHostMallocDefault));
        b->snap_cap = cap;
    }
    if (nslots > b->snap_wcap) {
        if (b->snap_wl) { HIPCHK(hipFree(b->snap_wl)); b->snap_wl = nullptr; }
        if (b->snap_woff) { HIPCHK(hipFree(b->snap_woff)); b->snap_woff = nullptr; }
        b->snap_wcap = 0;
        HIPCHK(hipMalloc((void**)&b->snap_wl, sizeof(int32_t) * (size_t)nslots));
        HIPCHK(hipMalloc((void**)&b->snap_woff, sizeof(int64_t) * (size_t)nslots));
        b->snap_wcap = nslots;
    }
    int32_t rc;
    if ((rc = put(b, b->snap_wl, slots, nslots))) return rc;
    if ((rc = put(b, b->snap_woff, offsets, nslots))) return rc;
    // packed rows: [total][3] position, [total][4] axis-angle, [total] awake, [total] dynamic
    double* dpos = (double*)b->snap_dev;
    double* daa = dpos + 3 * total;
    uint8_t* daw = (uint8_t*)(daa + 4 * total);
    uint8_t* ddy = daw + total;
    launch_snapshot(b->d, b->snap_wl, b->snap_woff, nslots, dpos, daa, daw, ddy, b->stream);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(b->snap_host, b->snap_dev, (size_t)(total * kSnapRow), hipMemcpyDeviceToHost, b->stream));
    HIPCHK(hipStreamSynchronize(b->stream));
    const unsigned char* h = b->snap_host;
    if (pos) std::memcpy(pos, h, sizeof(double) * 3 * total);
    if (axis_angle) std::memcpy(axis_angle, h + sizeof(double) * 3 * total, sizeof(double) * 4 * total);
    if (awake) std::memcpy(awake, h + sizeof(double) * 7 * total, (size_t)total);
    if (dynamic) std::memcpy(dynamic, h + sizeof(double) * 7 * total + total, (size_t)total);
    return OMI_OK;
}

template <typename R>
static void to_double(const std::vector<unsigned char>& raw, double* out, int64_t n) {
    const R* p = (const R*)raw.data();
    for (int64_t i = 0; i < n; i++) out[i] = (double)p[i];
}

extern "C" int32_t omi_batch_read_contacts(omi_batch* b, int32_t slot, int64_t cap, int64_t* count, int32_t* a,
                                           int32_t* bb, double* point, double* normal, double* depth,
                                           double* nimp, double* timp) {
    if (!b || slot < 0 || slot >= b->cfg.max_worlds || !b->slots[slot].live) return fail(OMI_ERR_ARG, "bad world slot");
    const WorldSlot& s = b->slots[slot];
    const Dev& d = b->d;
    int32_t rc, cs = 0, cn = 0;
    HIPCHK(hipStreamSynchronize(b->stream));
    if ((rc = get(b, &cs, d.w_cs + slot, 1))) return rc;
    if ((rc = get(b, &cn, d.w_cn + slot, 1))) return rc;
    if (count) *count = cn;
    const int64_t m = std::min<int64_t>(cn, cap);
    if (m <= 0) return OMI_OK;
    const size_t rs = b->cfg.precision == OMI_FP64 ? 8 : 4;
    std::vector<int32_t> ia((size_t)m), ib((size_t)m);
    if ((rc = get(b, ia.data(), d.ca + cs, m))) return rc;
    if ((rc = get(b, ib.data(), d.cb + cs, m))) return rc;
    for (int64_t k = 0; k < m; k++) {
        if (a) a[k] = (int32_t)(ia[k] - s.off);
        if (bb) bb[k] = (int32_t)(ib[k] - s.off);
    }
    std::vector<double> tmp((size_t)m);
    std::vector<unsigned char> raw((size_t)m * rs);
    for (int c = 0; c < 3; c++) {
        if (point) {
            if ((rc = get(b, tmp.data(), d.cpt + (size_t)c * d.kcap + cs, m))) return rc;
            for (int64_t k = 0; k < m; k++) point[3 * k + c] = tmp[k];
        }
        if (normal) {
            HIPCHK(hipMemcpy(raw.data(), (const unsigned char*)d.cnrm + ((size_t)c * d.kcap + cs) * rs, m * rs, hipMemcpyDeviceToHost));
            if (rs == 8) to_double<double>(raw, tmp.data(), m); else to_double<float>(raw, tmp.data(), m);
            for (int64_t k = 0; k < m; k++) normal[3 * k + c] = tmp[k];
        }
        if (nimp || timp) {
            HIPCHK(hipMemcpy(raw.data(), (const unsigned char*)d.cimp + ((size_t)c * d.kcap + cs) * rs, m * rs, hipMemcpyDeviceToHost));
            if (rs == 8) to_double<double>(raw, tmp.data(), m); else to_double<float>(raw, tmp.data(), m);
            for (int64_t k = 0; k < m; k++) {
                if (c == 0 && nimp) nimp[k] = tmp[k];
                if (c > 0 && timp) timp[2 * k + c - 1] = tmp[k];
            }
        }
    }
    if (depth) {
        HIPCHK(hipMemcpy(raw.data(), (const unsigned char*)d.cdep + (size_t)cs * rs, m * rs, hipMemcpyDeviceToHost));
        if (rs == 8) to_double<double>(raw, depth, m); else to_double<float>(raw, depth, m);
    }
    return OMI_OK;
}

extern "C" int32_t omi_batch_read_pairs(omi_batch* b, int32_t slot, int32_t solid, int64_t cap, int64_t* count,
                                        int32_t* pairs) {
    if (!b || slot < 0 || slot >= b->cfg.max_worlds || !b->slots[slot].live) return fail(OMI_ERR_ARG, "bad world slot");
    if (!solid && !b->cfg.record_broadphase) return fail(OMI_ERR_STATE, "broadphase pairs are recorded only with record_broadphase");
    const WorldSlot& s = b->slots[slot];
    const Dev& d = b->d;
    int32_t ps = 0, pn = 0, rc;
    HIPCHK(hipStreamSynchronize(b->stream));
    if ((rc = get(b, &ps, (solid ? d.w_ps : d.w_bps) + slot, 1))) return rc;
    if ((rc = get(b, &pn, (solid ? d.w_pn : d.w_bpn) + slot, 1))) return rc;
    if (count) *count = pn;
    const int64_t m = std::min<int64_t>(pn, cap);
    if (m <= 0 || !pairs) return OMI_OK;
    std::vector<int32_t> pa((size_t)m), pb((size_t)m);
    if ((rc = get(b, pa.data(), (solid ? d.pa : d.bpa) + ps, m))) return rc;
    if ((rc = get(b, pb.data(), (solid ? d.pb : d.bpb) + ps, m))) return rc;
    for (int64_t k = 0; k < m; k++) { pairs[2 * k] = (int32_t)(pa[k] - s.off); pairs[2 * k + 1] = (int32_t)(pb[k] - s.off); }
    return OMI_OK;
}

// ---------------------------------------------------------------- stages --

struct DevBuf {
    std::vector<void*> p;
    ~DevBuf() { for (void* q : p) hipFree(q); }
    template <typename T>
    T* up(const T* src, int64_t n) {
        void* q = nullptr;
        if (hipMalloc(&q, sizeof(T) * (size_t)std::max<int64_t>(n, 1)) != hipSuccess) return nullptr;
        p.push_back(q);
        if (src && n > 0) hipMemcpy(q, src, sizeof(T) * n, hipMemcpyHostToDevice);
        return (T*)q;
    }
};

extern "C" int32_t omi_capsule_triangles(int64_t n, const double* p0, const double* p1, double radius,
                                         const double* verts, uint8_t* hit, double* points, double* normals,
                                         double* depths) {
    if (n < 0 || (n > 0 && (!p0 || !p1 || !verts || !hit || !points || !normals || !depths)))
        return fail(OMI_ERR_ARG, "bad arguments");
    int32_t rc = check_device();
    if (rc) return rc;
    if (n == 0) return OMI_OK;
    DevBuf m;
    double *dc0 = m.up(p0, 3), *dc1 = m.up(p1, 3), *dv = m.up(verts, 9 * n);
    uint8_t* dh = m.up<uint8_t>(nullptr, n);
    double *dp = m.up<double>(nullptr, 3 * n), *dn = m.up<double>(nullptr, 3 * n), *dd = m.up<double>(nullptr, n);
    if (!dd) return fail(OMI_ERR_HIP, "device allocation failed");
    launch_capsule_triangles(n, dc0, dc1, radius, dv, dh, dp, dn, dd, nullptr);
    HIPCHK(hipGetLastError());
    HIPCHK(hipDeviceSynchronize());
    HIPCHK(hipMemcpy(hit, dh, (size_t)n, hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(points, dp, sizeof(double) * 3 * n, hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(normals, dn, sizeof(double) * 3 * n, hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(depths, dd, sizeof(double) * n, hipMemcpyDeviceToHost));
    return OMI_OK;
}

extern "C" int32_t omi_box_box(int32_t np, int32_t precision, const double* ca, const double* Ra, const double* ha,
                               const double* cb, const double* Rb, const double* hb, int32_t* count, double* normal,
                               double* points, double* depths) {
    if (np < 0 || (precision != 64 && precision != 32)) return fail(OMI_ERR_ARG, "bad arguments");
    int32_t rc = check_device();
    if (rc) return rc;
    if (np == 0) return OMI_OK;
    DevBuf m;
    double *dca = m.up(ca, 3 * np), *dRa = m.up(Ra, 9 * np), *dha = m.up(ha, 3 * np);
    double *dcb = m.up(cb, 3 * np), *dRb = m.up(Rb, 9 * np), *dhb = m.up(hb, 3 * np);
    int32_t* dc = m.up<int32_t>(nullptr, np);
    double *dn = m.up<double>(nullptr, 3 * np), *dp = m.up<double>(nullptr, 12 * np), *dd = m.up<double>(nullptr, 4 * np);
    if (!dd) return fail(OMI_ERR_HIP, "device allocation failed");
    launch_stage_boxbox(precision, np, dca, dRa, dha, dcb, dRb, dhb, dc, dn, dp, dd, nullptr);
    HIPCHK(hipGetLastError());
    HIPCHK(hipDeviceSynchronize());
    HIPCHK(hipMemcpy(count, dc, sizeof(int32_t) * np, hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(normal, dn, sizeof(double) * 3 * np, hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(points, dp, sizeof(double) * 12 * np, hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(depths, dd, sizeof(double) * 4 * np, hipMemcpyDeviceToHost));
    return OMI_OK;
}

static int32_t stage_spheres(int32_t np, int32_t precision, int32_t mode, const double* ca, const double* ra,
                             const double* cb, const double* Rb, const double* hb, const double* rb, int32_t* count,
                             double* normal, double* point, double* depth) {
    if (np < 0 || (precision != 64 && precision != 32)) return fail(OMI_ERR_ARG, "bad arguments");
    int32_t rc = check_device();
    if (rc) return rc;
    if (np == 0) return OMI_OK;
    DevBuf m;
    double *dca = m.up(ca, 3 * np), *dra = m.up(ra, np), *dcb = m.up(cb, 3 * np);
    double *dRb = mode ? m.up(Rb, 9 * np) : nullptr, *dhb = mode ? m.up(hb, 3 * np) : nullptr;
    double* drb = mode ? nullptr : m.up(rb, np);
    int32_t* dc = m.up<int32_t>(nullptr, np);
    double *dn = m.up<double>(nullptr, 3 * np), *dp = m.up<double>(nullptr, 3 * np), *dd = m.up<double>(nullptr, np);
    if (!dd) return fail(OMI_ERR_HIP, "device allocation failed");
    launch_stage_spheres(precision, np, mode, dca, dra, dcb, dRb, dhb, drb, dc, dn, dp, dd, nullptr);
    HIPCHK(hipGetLastError());
    HIPCHK(hipDeviceSynchronize());
    HIPCHK(hipMemcpy(count, dc, sizeof(int32_t) * np, hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(normal, dn, sizeof(double) * 3 * np, hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(point, dp, sizeof(double) * 3 * np, hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(depth, dd, sizeof(double) * np, hipMemcpyDeviceToHost));
    return OMI_OK;
}

extern "C" int32_t omi_sphere_sphere(int32_t np, int32_t precision, const double* ca, const double* ra,
                                     const double* cb, const double* rb, int32_t* count, double* normal,
                                     double* point, double* depth) {
    return stage_spheres(np, precision, 0, ca, ra, cb, nullptr, nullptr, rb, count, normal, point, depth);
}

extern "C" int32_t omi_sphere_box(int32_t np, int32_t precision, const double* cs, const double* rs,
                                  const double* cb, const double* Rb, const double* hb, int32_t* count,
                                  double* normal, double* point, double* depth) {
    return stage_spheres(np, precision, 1, cs, rs, cb, Rb, hb, nullptr, count, normal, point, depth);
}

// A one-world batch whose arrays the stage entry points fill directly.
static int32_t mini_batch(int64_t nb, int64_t npairs, int64_t ncon, int32_t precision, int32_t record,
                          omi_batch** out, int64_t ncand = 0, int32_t iters = 0) {
    omi_batch_config c{};
    c.max_candidates = ncand;
    c.max_solver_iterations = iters;
    c.max_worlds = 1;
    c.precision = precision;
    c.max_bodies = std::max<int64_t>(nb, 1);
    c.max_pairs = std::max<int64_t>(npairs, 1);
    c.max_contacts = std::max<int64_t>(ncon, 1);
    c.max_hull_pts = 1;
    c.record_broadphase = record;
    c.solver_block = 256;
    return omi_batch_create(&c, out);
}

extern "C" int32_t omi_refit(int32_t n, const double* pos, const double* quat, const int32_t* kind,
                             const double* half, const int32_t* hull_off, const int32_t* hull_cnt,
                             const double* hull_pts, int64_t n_hull_pts, double margin, double* lo, double* hi) {
    if (n < 0) return fail(OMI_ERR_ARG, "negative body count");
    if (n == 0) return OMI_OK;
    omi_batch_config c{};
    c.max_worlds = 1; c.precision = 64; c.max_bodies = n; c.max_pairs = 1; c.max_contacts = 1;
    c.max_hull_pts = std::max<int64_t>(n_hull_pts, 1);
    omi_batch* b = nullptr;
    int32_t rc = omi_batch_create(&c, &b);
    if (rc) return rc;
    Dev& d = b->d;
    std::vector<uint8_t> k((size_t)n);
    std::vector<int32_t> ho((size_t)n, 0), hc((size_t)n, 0), wz((size_t)n, 0);
    for (int32_t i = 0; i < n; i++) {
        k[i] = (uint8_t)kind[i];
        if (kind[i] == 3) {
            if (!hull_off || !hull_cnt || hull_off[i] < 0 || hull_cnt[i] < 1 || hull_off[i] + hull_cnt[i] > n_hull_pts) {
                omi_batch_destroy(b);
                return fail(OMI_ERR_ARG, "convex body without valid hull points");
            }
            ho[i] = hull_off[i]; hc[i] = hull_cnt[i];
        }
    }
    if (!(rc = upload_soa(b, d.pos, d.ncap, 0, pos, n, 3)) && !(rc = upload_soa(b, d.quat, d.ncap, 0, quat, n, 4)) &&
        !(rc = upload_soa(b, d.half, d.ncap, 0, half, n, 3)) && !(rc = put(b, d.kind, k.data(), n)) &&
        !(rc = put(b, d.hoff, ho.data(), n)) && !(rc = put(b, d.hcnt, hc.data(), n)) &&
        !(rc = put(b, d.world, wz.data(), n)) &&
        !(rc = (n_hull_pts > 0 ? put(b, d.hull, hull_pts, 3 * n_hull_pts) : OMI_OK))) {
        launch_refit(d, n, margin, b->stream);
        if (hipGetLastError() != hipSuccess) rc = fail(OMI_ERR_HIP, "refit launch failed");
        if (!rc) rc = download_soa(b, lo, d.lo, d.ncap, 0, n, 3);
        if (!rc) rc = download_soa(b, hi, d.hi, d.ncap, 0, n, 3);
    }
    omi_batch_destroy(b);
    return rc;
}

extern "C" int32_t omi_integrate_forces(int32_t n, const int32_t* motion_type, const uint8_t* awake,
                                        const double* gravity, const double* gravity_factor,
                                        const double* linear_damping, const double* angular_damping,
                                        const double* quadratic_drag, double default_linear_damping,
                                        double default_angular_damping, double dt, double* lv, double* av) {
    if (n < 0) return fail(OMI_ERR_ARG, "negative body count");
    if (n == 0) return OMI_OK;
    if (!motion_type || !awake || !gravity || !gravity_factor || !linear_damping || !angular_damping ||
        !quadratic_drag || !lv || !av)
        return fail(OMI_ERR_ARG, "null argument");
    int32_t rc = check_device();
    if (rc) return rc;
    DevBuf m;
    int32_t* dmt = m.up(motion_type, n);
    uint8_t* daw = m.up(awake, n);
    double *dg = m.up(gravity, 3 * (int64_t)n), *dgf = m.up(gravity_factor, n), *dld = m.up(linear_damping, n);
    double *dad = m.up(angular_damping, n), *dqd = m.up(quadratic_drag, n);
    double *dlv = m.up(lv, 3 * (int64_t)n), *dav = m.up(av, 3 * (int64_t)n);
    if (!dav) return fail(OMI_ERR_HIP, "device allocation failed");
    launch_stage_forces(n, dmt, daw, dg, dgf, dld, dad, dqd, default_linear_damping, default_angular_damping, dt, dlv,
                        dav, nullptr);
    HIPCHK(hipGetLastError());
    HIPCHK(hipDeviceSynchronize());
    HIPCHK(hipMemcpy(lv, dlv, sizeof(double) * 3 * n, hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(av, dav, sizeof(double) * 3 * n, hipMemcpyDeviceToHost));
    return OMI_OK;
}

extern "C" int32_t omi_integrate_positions(int32_t n, const int32_t* motion_type, const uint8_t* awake,
                                           const double* lv, const double* av, double dt, double* pos, double* quat) {
    if (n < 0) return fail(OMI_ERR_ARG, "negative body count");
    if (n == 0) return OMI_OK;
    if (!motion_type || !awake || !lv || !av || !pos || !quat) return fail(OMI_ERR_ARG, "null argument");
    int32_t rc = check_device();
    if (rc) return rc;
    DevBuf m;
    int32_t* dmt = m.up(motion_type, n);
    uint8_t* daw = m.up(awake, n);
    double *dlv = m.up(lv, 3 * (int64_t)n), *dav = m.up(av, 3 * (int64_t)n);
    double *dp = m.up(pos, 3 * (int64_t)n), *dq = m.up(quat, 4 * (int64_t)n);
    if (!dq) return fail(OMI_ERR_HIP, "device allocation failed");
    launch_stage_positions(n, dmt, daw, dlv, dav, dt, dp, dq, nullptr);
    HIPCHK(hipGetLastError());
    HIPCHK(hipDeviceSynchronize());
    HIPCHK(hipMemcpy(pos, dp, sizeof(double) * 3 * n, hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(quat, dq, sizeof(double) * 4 * n, hipMemcpyDeviceToHost));
    return OMI_OK;
}

extern "C" int32_t omi_sleep(int32_t n, int64_t K, const int32_t* ca, const int32_t* cb, const int32_t* motion_type,
                             double dt, double lin_thresh, double ang_thresh, double t_sleep, double* lv, double* av,
                             uint8_t* awake, double* sleep_timer) {
    if (n < 0 || K < 0) return fail(OMI_ERR_ARG, "negative size");
    if (n == 0) return OMI_OK;
    if (!motion_type || !lv || !av || !awake || !sleep_timer || (K > 0 && (!ca || !cb)))
        return fail(OMI_ERR_ARG, "null argument");
    for (int64_t k = 0; k < K; k++)
        if (ca[k] < 0 || ca[k] >= n || cb[k] < 0 || cb[k] >= n) return fail(OMI_ERR_ARG, "contact body index out of range");
    omi_batch* b = nullptr;
    int32_t rc = mini_batch(n, 1, std::max<int64_t>(K, 1), 64, 0, &b);
    if (rc) return rc;
    Dev& d = b->d;
    std::vector<uint8_t> mt((size_t)n);
    std::vector<int32_t> wz((size_t)n, 0), par((size_t)n);
    for (int32_t i = 0; i < n; i++) { mt[i] = (uint8_t)motion_type[i]; par[i] = i; }
    int32_t zero = 0, Ki = (int32_t)K, flags = WF_ACTIVE | WF_SLEEP;
    uint64_t kc = (uint64_t)K;
    if (!(rc = upload_soa(b, d.lv, d.ncap, 0, lv, n, 3)) && !(rc = upload_soa(b, d.av, d.ncap, 0, av, n, 3)) &&
        !(rc = put(b, d.mt, mt.data(), n)) && !(rc = put(b, d.awake, awake, n)) &&
        !(rc = put(b, d.timer, sleep_timer, n)) && !(rc = put(b, d.world, wz.data(), n)) &&
        !(rc = put(b, d.parent, par.data(), n)) && !(rc = put(b, d.ca, ca, K)) && !(rc = put(b, d.cb, cb, K)) &&
        !(rc = put(b, d.w_cs, &zero, 1)) && !(rc = put(b, d.w_cn, &Ki, 1)) && !(rc = put(b, d.w_flags, &flags, 1)) &&
        !(rc = put(b, d.cnt + C_CONTACTS, &kc, 1))) {
        launch_sleep(d, n, dt, b->stream, lin_thresh, ang_thresh, t_sleep);
        if (hipGetLastError() != hipSuccess) rc = fail(OMI_ERR_HIP, "sleep launch failed");
        if (!rc) rc = download_soa(b, lv, d.lv, d.ncap, 0, n, 3);
        if (!rc) rc = download_soa(b, av, d.av, d.ncap, 0, n, 3);
        if (!rc) rc = get(b, awake, d.awake, n);
        if (!rc) rc = get(b, sleep_timer, d.timer, n);
    }
    omi_batch_destroy(b);
    return rc;
}

extern "C" int32_t omi_sap_pairs(int32_t n, const double* lo, const double* hi, const int32_t* motion_type,
                                 const int32_t* filt_slot, int32_t n_filt, const uint8_t* filter_table,
                                 int32_t* pairs, int64_t cap, int64_t* count) {
    if (n < 0 || n_filt < 1 || !count) return fail(OMI_ERR_ARG, "bad arguments");
    *count = 0;
    if (n < 2) return OMI_OK;
    int64_t pcap = std::max<int64_t>(std::min<int64_t>((int64_t)n * 16, 1 << 26), 1024);
    int64_t xcap = std::max<int64_t>((int64_t)n * 64, 1 << 16);
    for (;;) {
        omi_batch* b = nullptr;
        int32_t rc = mini_batch(n, pcap, 1, 64, 1, &b, xcap);
        if (rc) return rc;
        Dev& d = b->d;
        std::vector<uint8_t> mt((size_t)n), bf((size_t)n, BF_COLLIDER), aw((size_t)n, 1);
        std::vector<int32_t> wz((size_t)n, 0);
        for (int32_t i = 0; i < n; i++) mt[i] = (uint8_t)motion_type[i];
        int32_t flags = WF_ACTIVE | WF_COLLISION, zero = 0, nn = n;
        int32_t nf = n_filt;
        int32_t wl = 0;
        rc = upload_soa(b, d.lo, d.ncap, 0, lo, n, 3);
        if (!rc) rc = upload_soa(b, d.hi, d.ncap, 0, hi, n, 3);
        if (!rc) rc = put(b, d.mt, mt.data(), n);
        if (!rc) rc = put(b, d.bflags, bf.data(), n);
        if (!rc) rc = put(b, d.awake, aw.data(), n);
        if (!rc) rc = put(b, d.world, wz.data(), n);
        if (!rc) rc = put(b, d.filt, filt_slot, n);
        if (!rc) rc = put(b, d.filt_tab, filter_table, (int64_t)n_filt * n_filt);
        if (!rc) rc = put(b, d.w_flags, &flags, 1);
        if (!rc) rc = put(b, d.w_off, &zero, 1);
        if (!rc) rc = put(b, d.w_cnt, &nn, 1);
        if (!rc) rc = put(b, d.w_filt_off, &zero, 1);
        if (!rc) rc = put(b, d.w_nfilt, &nf, 1);
        if (rc) { omi_batch_destroy(b); return rc; }
        b->slots[0].live = true;
        b->slots[0].n = n;
        if ((rc = rebuild_lists(b))) { omi_batch_destroy(b); return rc; }
        launch_begin(d, b->stream);
        enqueue_sap(b, 1);
        if (hipGetLastError() != hipSuccess) { omi_batch_destroy(b); return fail(OMI_ERR_HIP, "broadphase launch failed"); }
        uint64_t c[C_COUNT];
        hipStreamSynchronize(b->stream);
        hipMemcpy(c, d.cnt, sizeof c, hipMemcpyDeviceToHost);
        if (c[C_OVERFLOW] & 16ull) {
            omi_batch_destroy(b);
            if (c[C_XCAND] == 0) return fail(OMI_ERR_CAPACITY, "more than 2^31 x-window candidates in one world");
            xcap = (int64_t)c[C_XCAND] + 2 * SAP_TILE;
            continue;
        }
        if (c[C_OVERFLOW] & 8ull) {
            omi_batch_destroy(b);
            pcap = (int64_t)c[C_BPTOTAL] + 16;
            continue;
        }
        int64_t np_ = (int64_t)c[C_BPTOTAL];
        *count = np_;
        const int64_t m = std::min<int64_t>(np_, cap);
        std::vector<int32_t> pa((size_t)std::max<int64_t>(m, 1)), pb((size_t)std::max<int64_t>(m, 1));
        if (m > 0 && pairs) {
            rc = get(b, pa.data(), d.bpa, m);
            if (!rc) rc = get(b, pb.data(), d.bpb, m);
            for (int64_t k = 0; k < m && !rc; k++) { pairs[2 * k] = pa[k]; pairs[2 * k + 1] = pb[k]; }
        }
        omi_batch_destroy(b);
        return rc;
    }
}

extern "C" int32_t omi_prepare_and_solve(int64_t K, int32_t N, const int64_t* a_idx, const int64_t* b_idx,
                                         const double* point, const double* normal, const double* depth, double* pos,
                                         double* lv, double* av, const double* inv_mass, const double* invIw,
                                         const double* restitution, const double* friction, double* nImp,
                                         double* tImp, double rthr, int32_t iters, int32_t warm, double correction,
                                         double slop, int32_t precision) {
    if (K < 0 || N < 0 || (precision != 64 && precision != 32)) return fail(OMI_ERR_ARG, "bad arguments");
    if (iters < 0 || iters > 30 || (warm != 0 && warm != 1)) return fail(OMI_ERR_ARG, "iters must be 0..30, warm 0 or 1");
    if (K == 0) return OMI_OK;
    for (int64_t k = 0; k < K; k++)
        if (a_idx[k] < 0 || a_idx[k] >= N || b_idx[k] < 0 || b_idx[k] >= N || a_idx[k] == b_idx[k])
            return fail(OMI_ERR_ARG, "contact body index out of range");
    omi_batch* b = nullptr;
    int32_t rc = mini_batch(N, 1, K, precision, 0, &b, 0, std::max(iters, 1));
    if (rc) return rc;
    Dev& d = b->d;
    const bool f64 = precision == 64;
    const size_t rs = f64 ? 8 : 4;
    std::vector<int32_t> ia((size_t)K), ib((size_t)K), wz((size_t)N, 0);
    std::vector<uint8_t> mt((size_t)N, M_DYNAMIC);
    for (int64_t k = 0; k < K; k++) { ia[k] = (int32_t)a_idx[k]; ib[k] = (int32_t)b_idx[k]; }
    // manifolds: runs of contacts of one pair (every body counts as dynamic here)
    std::vector<int4> mrec;
    for (int64_t k = 0; k < K; k++) {
        if (k == 0 || ia[k] != ia[k - 1] || ib[k] != ib[k - 1])
            mrec.push_back(make_int4((int32_t)k, MR_DYN_A | MR_DYN_B, ia[k], ib[k]));
        if ((mrec.back().y & MR_COUNT) == MR_COUNT)   // a run longer than a record holds starts a new one
            mrec.push_back(make_int4((int32_t)k, MR_DYN_A | MR_DYN_B, ia[k], ib[k]));
        mrec.back().y++;
    }
    int32_t Mi = (int32_t)mrec.size();
    uint64_t kc = (uint64_t)K;
    // Real-typed uploads (component-major)
    auto upR = [&](void* dst, int64_t ld, const double* src, int64_t n, int k, int stride_k) -> int32_t {
        std::vector<unsigned char> tmp((size_t)n * rs);
        for (int c = 0; c < k; c++) {
            for (int64_t i = 0; i < n; i++) {
                double v = src[i * stride_k + c];
                if (f64) ((double*)tmp.data())[i] = v; else ((float*)tmp.data())[i] = (float)v;
            }
            if (hipMemcpy((unsigned char*)dst + (size_t)c * ld * rs, tmp.data(), n * rs, hipMemcpyHostToDevice) != hipSuccess)
                return fail(OMI_ERR_HIP, "upload failed");
        }
        return OMI_OK;
    };
    // world inverse inertia [N][9] -> one IIW_STRIDE record per body
    auto upIIW = [&](const double* src) -> int32_t {
        std::vector<unsigned char> tmp((size_t)N * IIW_STRIDE * rs, 0);
        for (int64_t i = 0; i < N; i++)
            for (int c = 0; c < 9; c++) {
                double v = src[i * 9 + c];
                if (f64) ((double*)tmp.data())[i * IIW_STRIDE + c] = v;
                else ((float*)tmp.data())[i * IIW_STRIDE + c] = (float)v;
            }
        if (hipMemcpy(d.iiw, tmp.data(), tmp.size(), hipMemcpyHostToDevice) != hipSuccess)
            return fail(OMI_ERR_HIP, "upload failed");
        return OMI_OK;
    };
    std::vector<double> imp((size_t)K * 3);
    for (int64_t k = 0; k < K; k++) { imp[3 * k] = nImp[k]; imp[3 * k + 1] = tImp[2 * k]; imp[3 * k + 2] = tImp[2 * k + 1]; }
    DevBuf tb;
    double* drest = tb.up(restitution, K);
    double* dfric = tb.up(friction, K);
    int32_t Ni = N, Ki = (int32_t)K, zero = 0, flags = WF_ACTIVE;
    if (!(rc = upload_soa(b, d.pos, d.ncap, 0, pos, N, 3)) && !(rc = upload_soa(b, d.lv, d.ncap, 0, lv, N, 3)) &&
        !(rc = upload_soa(b, d.av, d.ncap, 0, av, N, 3)) && !(rc = put(b, d.inv_mass, inv_mass, N)) &&
        !(rc = put(b, d.mt, mt.data(), N)) && !(rc = put(b, d.world, wz.data(), N)) &&
        !(rc = upIIW(invIw)) && !(rc = put(b, d.ca, ia.data(), K)) &&
        !(rc = put(b, d.cb, ib.data(), K)) && !(rc = upload_soa(b, d.cpt, d.kcap, 0, point, K, 3)) &&
        !(rc = upR(d.cnrm, d.kcap, normal, K, 3, 3)) && !(rc = upR(d.cdep, d.kcap, depth, K, 1, 1)) &&
        !(rc = upR(d.cimp, d.kcap, imp.data(), K, 3, 3)) && !(rc = put(b, d.w_off, &zero, 1)) &&
        !(rc = put(b, d.w_cnt, &Ni, 1)) && !(rc = put(b, d.w_cs, &zero, 1)) && !(rc = put(b, d.w_cn, &Ki, 1)) &&
        !(rc = put(b, d.w_iters, &iters, 1)) && !(rc = put(b, d.w_flags, &flags, 1)) &&
        !(rc = put(b, d.mrec, mrec.data(), Mi)) && !(rc = put(b, d.w_mn, &Mi, 1)) && !(rc = put(b, d.w_ms, &zero, 1)) &&
        !(rc = put(b, d.cnt + C_CONTACTS, &kc, 1))) {
        SolveIn in{drest, dfric, rthr, correction, slop, warm};
        b->slots[0].live = true;
        b->slots[0].n = N;
        if (!(rc = rebuild_lists(b)))
            enqueue_solve(b, in, true, [](hipStream_t) {}, [] {});
        if (!rc && hipGetLastError() != hipSuccess) rc = fail(OMI_ERR_HIP, "solver launch failed");
        if (!rc) rc = download_soa(b, pos, d.pos, d.ncap, 0, N, 3);
        if (!rc) rc = download_soa(b, lv, d.lv, d.ncap, 0, N, 3);
        if (!rc) rc = download_soa(b, av, d.av, d.ncap, 0, N, 3);
        if (!rc) {
            std::vector<unsigned char> raw((size_t)K * rs);
            for (int c = 0; c < 3 && !rc; c++) {
                if (hipMemcpy(raw.data(), (unsigned char*)d.cimp + (size_t)c * d.kcap * rs, K * rs, hipMemcpyDeviceToHost) != hipSuccess) {
                    rc = fail(OMI_ERR_HIP, "download failed");
                    break;
                }
                for (int64_t k = 0; k < K; k++) {
                    double v = f64 ? ((double*)raw.data())[k] : (double)((float*)raw.data())[k];
                    if (c == 0) nImp[k] = v; else tImp[2 * k + c - 1] = v;
                }
            }
        }
    }
    omi_batch_destroy(b);
    return rc;
}

extern "C" int32_t omi_collide_convex(int32_t np, int32_t precision, const int32_t* kind_a, const double* ca,
                                      const double* Ra, const double* ha, const int32_t* off_a, const int32_t* cnt_a,
                                      const int32_t* kind_b, const double* cb, const double* Rb, const double* hb,
                                      const int32_t* off_b, const int32_t* cnt_b, const double* hull_pts,
                                      int64_t n_hull_pts, int32_t* count, double* normal, double* point,
                                      double* depth) {
    if (np < 0 || (precision != 64 && precision != 32)) return fail(OMI_ERR_ARG, "bad arguments");
    int32_t rc = check_device();
    if (rc) return rc;
    if (np == 0) return OMI_OK;
    for (int32_t i = 0; i < np; i++) {
        for (int s2 = 0; s2 < 2; s2++) {
            int k = s2 ? kind_b[i] : kind_a[i];
            if (k != 1 && k != 2 && k != 3) return fail(OMI_ERR_ARG, "kind must be sphere, box or convex");
            if (k == 3) {
                int32_t o = s2 ? off_b[i] : off_a[i], c = s2 ? cnt_b[i] : cnt_a[i];
                if (o < 0 || c < 1 || (int64_t)o + c > n_hull_pts) return fail(OMI_ERR_ARG, "hull range out of bounds");
            }
        }
    }
    DevBuf m;
    std::vector<int32_t> zeros((size_t)np, 0);
    int32_t *dka = m.up(kind_a, np), *dkb = m.up(kind_b, np);
    int32_t *doa = m.up(off_a ? off_a : zeros.data(), np), *dna = m.up(cnt_a ? cnt_a : zeros.data(), np);
    int32_t *dob = m.up(off_b ? off_b : zeros.data(), np), *dnb = m.up(cnt_b ? cnt_b : zeros.data(), np);
    double *dca = m.up(ca, 3 * np), *dRa = m.up(Ra, 9 * np), *dha = m.up(ha, 3 * np);
    double *dcb = m.up(cb, 3 * np), *dRb = m.up(Rb, 9 * np), *dhb = m.up(hb, 3 * np);
    double* dh = m.up(hull_pts, 3 * std::max<int64_t>(n_hull_pts, 1));
    const int32_t slots = std::min<int32_t>(np, 4096);
    unsigned char* work = m.up<unsigned char>(nullptr, (int64_t)slots * (int64_t)epa_work_bytes());
    int32_t* dc = m.up<int32_t>(nullptr, np);
    double *dn = m.up<double>(nullptr, 3 * np), *dp = m.up<double>(nullptr, 3 * np), *dd = m.up<double>(nullptr, np);
    if (!dd || !work) return fail(OMI_ERR_HIP, "device allocation failed");
    launch_stage_convex(np, dka, dca, dRa, dha, doa, dna, dkb, dcb, dRb, dhb, dob, dnb, dh, n_hull_pts, work, slots, dc,
                        dn, dp, dd, nullptr);
    HIPCHK(hipGetLastError());
    HIPCHK(hipDeviceSynchronize());
    HIPCHK(hipMemcpy(count, dc, sizeof(int32_t) * np, hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(normal, dn, sizeof(double) * 3 * np, hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(point, dp, sizeof(double) * 3 * np, hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(depth, dd, sizeof(double) * np, hipMemcpyDeviceToHost));
    return OMI_OK;
}

// ------------------------------------------------ no-allocation entry point --
//
// omi_collide_convex_noalloc runs the convex stage kernel on caller-owned
// device buffers with its GJK/EPA workspaces taken from `scratch`, a
// file-scope device array of OMI_NOALLOC_EPA_SLOTS workspaces: it calls no
// hipMalloc or hipFree.  Each device that runs it holds its own copy of the
// array (11.9 MB) until the process ends.  In host code `scratch` names a
// host-side stand-in, so its device address is looked up once per device with
// hipGetSymbolAddress and kept in g_scratch_dev; only that address reaches a
// kernel.  The workspaces carry nothing from one pair to the next, so the
// array is never cleared; two calls that overlap on one device share it and
// race.

static_assert(OMI_NOALLOC_EPA_SLOTS % 64 == 0, "whole wavefronts of workspace slots");
static __device__ __align__(256) unsigned char scratch[(size_t)OMI_NOALLOC_EPA_SLOTS * sizeof(EpaWork)];

static constexpr int kMaxDevices = 64;
static std::mutex g_scratch_mu;
static void* g_scratch_dev[kMaxDevices] = {};

// The current device's address of `scratch`, looked up on its first use there.
static int32_t scratch_address(void** out) {
    int dev = -1;
    HIPCHK(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) return fail(OMI_ERR_ARG, "device ordinal beyond the scratch table");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_dev[dev]) {
        void* p = nullptr;
        hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return fail(OMI_ERR_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
        g_scratch_dev[dev] = p;
    }
    *out = g_scratch_dev[dev];
    return OMI_OK;
}

extern "C" int32_t omi_collide_convex_noalloc(int32_t np, const int32_t* kind_a, const double* ca, const double* Ra,
                                              const double* ha, const int32_t* off_a, const int32_t* cnt_a,
                                              const int32_t* kind_b, const double* cb, const double* Rb,
                                              const double* hb, const int32_t* off_b, const int32_t* cnt_b,
                                              const double* hull_pts, int64_t n_hull_pts, int32_t* count,
                                              double* normal, double* point, double* depth, void* stream) {
    if (np < 0 || n_hull_pts < 0) return fail(OMI_ERR_ARG, "npairs and n_hull_pts must be >= 0");
    if (np == 0) return OMI_OK;
    if (!kind_a || !ca || !Ra || !ha || !off_a || !cnt_a || !kind_b || !cb || !Rb || !hb || !off_b || !cnt_b ||
        !hull_pts || !count || !normal || !point || !depth)
        return fail(OMI_ERR_ARG, "every buffer must be a device pointer (hull offsets and counts included)");
    void* work = nullptr;
    int32_t rc = scratch_address(&work);
    if (rc) return rc;
    const int32_t slots = std::min<int32_t>((np + 63) / 64 * 64, OMI_NOALLOC_EPA_SLOTS);
    launch_stage_convex(np, kind_a, ca, Ra, ha, off_a, cnt_a, kind_b, cb, Rb, hb, off_b, cnt_b, hull_pts, n_hull_pts,
                        work, slots, count, normal, point, depth, (hipStream_t)stream);
    HIPCHK(hipGetLastError());
    return OMI_OK;
}
