// k_step.hip — the per-step kernels of the contact pipeline on gfx950.
//
// Stage map (reference -> kernel):
//   world.step prev-pose copy, integrate_forces, refit  -> k_pre        (1 thread / body)
//   BroadPhase.pairs + solid filter                     -> k_sap        (1 block / world)
//   NarrowPhase.generate (box/sphere)                   -> k_narrow     (1 thread / pair)
//   contact list in pair-then-manifold order            -> k_compact    (1 block / world)
//   SequentialImpulseSolver._wake_sleepers              -> k_wake       (1 block / world)
//   _solve_native_full + prepare_and_solve              -> k_solve      (1 block / world)
//   warm-start cache rebuild                            -> k_cache      (1 block / world)
//   integrate_positions                                 -> k_post       (1 thread / body)
//   _update_sleep + _contact_islands                    -> k_sleep_*    (contacts, bodies)
#include "omi_math.hpp"
#include "omi_batch.hpp"
#include "omi_gjk.hpp"

namespace omi {

#define BODY(a, c, i) (a)[(size_t)(c) * d.ncap + (size_t)(i)]
#define CON(a, c, k) (a)[(size_t)(c) * d.kcap + (size_t)(k)]

__device__ __forceinline__ int64_t gtid() { return (int64_t)blockIdx.x * blockDim.x + threadIdx.x; }
__device__ __forceinline__ int64_t gstride() { return (int64_t)gridDim.x * blockDim.x; }

// Orders doubles as unsigned integers; -0.0 and +0.0 map to one key, like
// numpy's comparison-based argsort sees them.
__device__ __forceinline__ uint64_t order_key(double x) {
    if (x == 0.0) x = 0.0;
    uint64_t u = (uint64_t)__double_as_longlong(x);
    return (u >> 63) ? ~u : (u | 0x8000000000000000ull);
}

__device__ __forceinline__ uint64_t hash_mix(uint64_t x) {
    x ^= x >> 33; x *= 0xff51afd7ed558ccdull; x ^= x >> 33; x *= 0xc4ceb9fe1a85ec53ull; x ^= x >> 33;
    return x;
}

// ---------------------------------------------------------- block scans --
// Exclusive scan of a[0..n) in place (LDS or global), BLOCK threads; returns
// the total to every thread.  scratch: BLOCK/64 + 2 ints of LDS.
template <int BLOCK>
__device__ int32_t block_exscan(int32_t* a, int32_t n, int32_t* scratch) {
    constexpr int NW = BLOCK / 64;
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    int32_t running = 0;
    for (int32_t base = 0; base < n; base += BLOCK) {
        int32_t v = base + tid < n ? a[base + tid] : 0, x = v;
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
            int32_t y = __shfl_up(x, o, 64);
            if (lane >= o) x += y;
        }
        if (lane == 63) scratch[wv] = x;
        __syncthreads();
        int32_t wbase = 0, tot = 0;
#pragma unroll
        for (int q = 0; q < NW; q++) { int32_t s = scratch[q]; if (q < wv) wbase += s; tot += s; }
        if (base + tid < n) a[base + tid] = running + wbase + x - v;
        running += tot;
        __syncthreads();
    }
    return running;
}

// Exclusive prefix of a 0/1 flag over the block (one chunk); returns the
// prefix, sets `total`.
template <int BLOCK>
__device__ __forceinline__ int32_t block_flag_prefix(bool f, int32_t* scratch, int32_t& total) {
    constexpr int NW = BLOCK / 64;
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    uint64_t m = __ballot(f);
    int32_t pre = __popcll(m & ((1ull << lane) - 1ull));
    if (lane == 0) scratch[wv] = __popcll(m);
    __syncthreads();
    int32_t wbase = 0, tot = 0;
#pragma unroll
    for (int q = 0; q < NW; q++) { int32_t s = scratch[q]; if (q < wv) wbase += s; tot += s; }
    __syncthreads();
    total = tot;
    return wbase + pre;
}

template <int BLOCK>
__device__ __forceinline__ int64_t block_sum(int64_t v, int64_t* scratch) {
    constexpr int NW = BLOCK / 64;
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
    if (lane == 0) scratch[wv] = v;
    __syncthreads();
    int64_t t = 0;
#pragma unroll
    for (int q = 0; q < NW; q++) t += scratch[q];
    __syncthreads();
    return t;
}

// ------------------------------------------------------------ per step --

__global__ void k_begin(Dev d) {
    int64_t i = gtid();
    if (i < C_COUNT && i != C_OVERFLOW) d.cnt[i] = 0;   // overflow flags stay until the host reads them
    for (int64_t w = i; w < d.wcap; w += gstride()) d.w_flags[w] &= ~WF_SEQWAKE;
}

// Refit of body i from pose (world.py:471-508; body.py:115-123 for hulls).
__device__ __forceinline__ void refit_body(const Dev& d, int64_t i, const double Rm[9], double margin) {
    const int kind = d.kind[i];
    double p[3] = {BODY(d.pos, 0, i), BODY(d.pos, 1, i), BODY(d.pos, 2, i)};
    double h[3] = {BODY(d.half, 0, i), BODY(d.half, 1, i), BODY(d.half, 2, i)};
    double lo[3], hi[3];
    if (kind == K_NONE) {
        for (int r = 0; r < 3; r++) { lo[r] = p[r]; hi[r] = p[r]; }
    } else if (kind == K_CONVEX) {
        const int32_t off = d.hoff[i], m = d.hcnt[i];
        for (int r = 0; r < 3; r++) { lo[r] = 0.0; hi[r] = 0.0; }
        for (int32_t k = 0; k < m; k++) {
            const double* l = d.hull + 3 * (size_t)(off + k);
            for (int r = 0; r < 3; r++) {
                double v = ((l[0] * Rm[3 * r] + l[1] * Rm[3 * r + 1]) + l[2] * Rm[3 * r + 2]) + p[r];
                if (k == 0 || v < lo[r]) lo[r] = v;
                if (k == 0 || v > hi[r]) hi[r] = v;
            }
        }
        for (int r = 0; r < 3; r++) { lo[r] = lo[r] - margin; hi[r] = hi[r] + margin; }
    } else {
        for (int r = 0; r < 3; r++) {
            double hw = kind == K_SPHERE ? h[r] : box_hw(Rm, h, r);
            lo[r] = (p[r] - hw) - margin;
            hi[r] = (p[r] + hw) + margin;
        }
    }
    for (int r = 0; r < 3; r++) { BODY(d.lo, r, i) = lo[r]; BODY(d.hi, r, i) = hi[r]; }
}

// Start of step for every live body: previous pose, sleep scratch reset,
// integrate_forces (backend.py:26-40) for awake dynamic bodies, then the
// collision-side per-body data: AABB, rotation, world inverse inertia
// (solver.py:215-220, einsum order (p0 + p2) + p1).
template <typename R>
__global__ void k_pre(Dev d, int64_t nhw, double dt, double margin) {
    for (int64_t i = gtid(); i < nhw; i += gstride()) {
        const int32_t w = d.world[i];
        if (w < 0) continue;
        const int32_t wf = d.w_flags[w];
        double q[4];
        for (int c = 0; c < 3; c++) BODY(d.ppos, c, i) = BODY(d.pos, c, i);
        for (int c = 0; c < 4; c++) { q[c] = BODY(d.quat, c, i); BODY(d.pquat, c, i) = q[c]; }
        const int mt = d.mt[i];
        const uint8_t aw = d.awake[i];
        d.awake0[i] = aw;
        d.parent[i] = (int32_t)i;
        d.contacted[i] = 0;
        d.notready[i] = 0;
        double v[3] = {BODY(d.lv, 0, i), BODY(d.lv, 1, i), BODY(d.lv, 2, i)};
        if (mt == M_DYNAMIC && aw) {
            const double gf = d.gf[i], dt_ = dt;
            for (int r = 0; r < 3; r++) v[r] = v[r] + (d.w_grav[(size_t)r * d.wcap + w] * gf) * dt_;
            double fl = maxd(1.0 - (d.ld[i] + d.w_damp[w]) * dt_, 0.0);
            for (int r = 0; r < 3; r++) v[r] = v[r] * fl;
            double sp = sqrt((v[0] * v[0] + v[1] * v[1]) + v[2] * v[2]);
            const double qd = d.qd[i];
            for (int r = 0; r < 3; r++) v[r] = v[r] - ((qd * sp) * v[r]) * dt_;
            double fa = maxd(1.0 - (d.ad[i] + d.w_damp[d.wcap + w]) * dt_, 0.0);
            for (int r = 0; r < 3; r++) BODY(d.lv, r, i) = v[r];
            for (int r = 0; r < 3; r++) BODY(d.av, r, i) = BODY(d.av, r, i) * fa;
        } else if (mt == M_DYNAMIC && (wf & WF_COLLISION)) {
            // A sleeping body with a wake-grade speed makes the wake order matter.
            double v2 = (v[0] * v[0] + v[1] * v[1]) + v[2] * v[2];
            if (v2 > 0.15 * 0.15) atomicOr(&d.w_flags[w], WF_SEQWAKE);
        }
        if (wf & (WF_COLLISION | WF_JOINTS)) {
            double Rm[9];
            quat_to_matrix(q, Rm);
            if (wf & WF_COLLISION) refit_body(d, i, Rm, margin);
            R* rot = (R*)d.rot;
            R* iiw = (R*)d.iiw;
            const double dg[3] = {BODY(d.inv_inertia, 0, i), BODY(d.inv_inertia, 1, i), BODY(d.inv_inertia, 2, i)};
            for (int r = 0; r < 9; r++) BODY(rot, r, i) = (R)Rm[r];
            for (int r = 0; r < 3; r++)
                for (int c = 0; c < 3; c++) {
                    double p0 = (Rm[3 * r] * dg[0]) * Rm[3 * c], p1 = (Rm[3 * r + 1] * dg[1]) * Rm[3 * c + 1];
                    double p2 = (Rm[3 * r + 2] * dg[2]) * Rm[3 * c + 2];
                    BODY(iiw, 3 * r + c, i) = (R)((p0 + p2) + p1);
                }
        }
    }
}

// Refit only (PhysicsWorld.refit_aabbs called outside a step).
__global__ void k_refit(Dev d, int64_t nhw, double margin) {
    for (int64_t i = gtid(); i < nhw; i += gstride()) {
        const int32_t w = d.world[i];
        if (w < 0) continue;
        double q[4], Rm[9];
        for (int c = 0; c < 4; c++) q[c] = BODY(d.quat, c, i);
        quat_to_matrix(q, Rm);
        refit_body(d, i, Rm, margin);
    }
}

// ------------------------------------------------------------ broadphase --

// BroadPhase.pairs (broadphase.py:243-294), in three passes.
//
// k_sap_sort, one block per world: bitonic sort of (lo.x key, index) — the
// stable argsort order — in LDS (global scratch for large worlds); the
// candidate window of sorted body k is [k+1, searchsorted(lo_sx, hi_sx[k],
// 'right')); so[k] is the exclusive prefix of the window lengths, so
// candidate c of the world is (k, k + 1 + c - so[k]) for the last k with
// so[k] <= c.  The y/z bounds, motion/filter flags and original index are
// written in sorted order, and the world reserves ceil(X / SAP_TILE) tiles of
// its X candidates.
//
// k_sap_pairs<COUNT>, the whole chip: each wavefront walks a contiguous run of
// tiles, 64 candidates at a time, and counts the pairs that pass the y/z
// test, the motion/filter test and the solid filter (world.py:663-670).
// k_sap_scan turns the per-tile counts into output offsets and per-world
// pair ranges; k_sap_pairs<WRITE> walks again and writes the pairs in
// candidate order, i.e. the reference's order.
enum : uint32_t { SI_DYN = 1, SI_MOVING = 4, SI_COLLIDER = 8, SI_TRIGGER = 16 };   // filter slot << 8

template <int BLOCK, bool GLOBAL>
__global__ __launch_bounds__(BLOCK) void k_sap_sort(Dev d, const int32_t* wlist, const int32_t* wcount,
                                                    int32_t np_max) {
    extern __shared__ unsigned char smem[];
    __shared__ int32_t s_scan[BLOCK / 64 + 2];
    __shared__ int64_t s_sum[BLOCK / 64 + 2];
    __shared__ int32_t s_t0;
    if ((int32_t)blockIdx.x >= *wcount) return;
    const int32_t w = wlist[blockIdx.x];
    const int tid = threadIdx.x;
    const int32_t flags = d.w_flags[w], n = d.w_cnt[w];
    const int64_t off = d.w_off[w];
    if (!(flags & WF_ACTIVE) || !(flags & WF_COLLISION) || n < 2) {
        if (tid == 0) { d.w_x[w] = 0; d.w_tn[w] = 0; }
        return;
    }
    int32_t NP = 1;
    while (NP < n) NP <<= 1;
    uint64_t* key;
    int32_t *idx, *coff;
    if (GLOBAL) {
        key = d.skey + 2 * off; idx = d.sidx + 2 * off; coff = d.soff + 2 * off;
    } else {
        key = (uint64_t*)smem; idx = (int32_t*)(key + np_max); coff = idx + np_max;
    }
    for (int32_t t = tid; t < NP; t += BLOCK) {
        if (t < n) { key[t] = order_key(BODY(d.lo, 0, off + t)); idx[t] = t; }
        else { key[t] = ~0ull; idx[t] = 0x7fffffff; }
    }
    __syncthreads();
    for (int32_t k = 2; k <= NP; k <<= 1)
        for (int32_t j = k >> 1; j > 0; j >>= 1) {
            for (int32_t t = tid; t < (NP >> 1); t += BLOCK) {
                int32_t i = 2 * t - (t & (j - 1));
                int32_t l = i + j;
                uint64_t ki = key[i], kl = key[l];
                int32_t ii = idx[i], il = idx[l];
                bool gt = ki > kl || (ki == kl && ii > il);
                bool up = (i & k) == 0;
                if (gt == up) { key[i] = kl; key[l] = ki; idx[i] = il; idx[l] = ii; }
            }
            __syncthreads();
        }
    int64_t part = 0;
    for (int32_t k = tid; k < n; k += BLOCK) {
        const int64_t g = off + idx[k];
        const uint64_t hk = order_key(BODY(d.hi, 0, g));
        int32_t lo_i = 0, hi_i = n;
        while (lo_i < hi_i) {
            int32_t mid = (lo_i + hi_i) >> 1;
            if (key[mid] <= hk) lo_i = mid + 1; else hi_i = mid;
        }
        const int32_t c = lo_i - k - 1 > 0 ? lo_i - k - 1 : 0;
        coff[k] = c;
        part += c;
        // the sorted record of body g
        const int mt = d.mt[g];
        const uint8_t bf = d.bflags[g];
        uint32_t info = (uint32_t)d.filt[g] << 8;
        if (mt == M_DYNAMIC) info |= SI_DYN;
        if ((mt == M_DYNAMIC && d.awake[g]) || mt == M_KINEMATIC) info |= SI_MOVING;
        if (bf & BF_COLLIDER) info |= SI_COLLIDER;
        if (bf & BF_TRIGGER) info |= SI_TRIGGER;
        const int64_t s = off + k;
        d.sbox[s] = BODY(d.lo, 1, g);
        d.sbox[d.ncap + s] = BODY(d.hi, 1, g);
        d.sbox[2 * d.ncap + s] = BODY(d.lo, 2, g);
        d.sbox[3 * d.ncap + s] = BODY(d.hi, 2, g);
        d.sinfo[s] = info;
        d.sorig[s] = idx[k];
    }
    const int64_t X = block_sum<BLOCK>(part, s_sum);
    if (X > 0x7fffffffll - SAP_TILE) {
        // more x-window candidates than one world's 32-bit candidate index holds
        if (tid == 0) { atomicOr((unsigned long long*)&d.cnt[C_OVERFLOW], 16ull); d.w_x[w] = 0; d.w_tn[w] = 0; }
        return;
    }
    __syncthreads();
    block_exscan<BLOCK>(coff, n, s_scan);
    int32_t* so = d.soff + 2 * off;
    if (!GLOBAL)
        for (int32_t k = tid; k < n; k += BLOCK) so[k] = coff[k];
    const int32_t T = (int32_t)((X + SAP_TILE - 1) / SAP_TILE);
    if (tid == 0) {
        so[n] = (int32_t)X;
        d.w_x[w] = X;
        atomicAdd((unsigned long long*)&d.cnt[C_XCAND], (unsigned long long)X);
        const uint64_t t0 = T ? atomicAdd((unsigned long long*)&d.cnt[C_TILES], (unsigned long long)T) : 0ull;
        const bool ok = t0 + (uint64_t)T <= (uint64_t)d.tcap;
        if (!ok) atomicOr((unsigned long long*)&d.cnt[C_OVERFLOW], 16ull);
        d.w_t0[w] = (int32_t)(ok ? t0 : 0);
        d.w_tn[w] = ok ? T : 0;
        s_t0 = ok ? (int32_t)t0 : -1 - (int32_t)(t0 < (uint64_t)d.tcap ? t0 : (uint64_t)d.tcap);
    }
    __syncthreads();
    // tile -> world map; tiles of a world that did not fit are marked empty
    const int32_t t0 = s_t0 >= 0 ? s_t0 : -1 - s_t0;
    const int32_t tend = s_t0 >= 0 ? t0 + T : (t0 + T < d.tcap ? t0 + T : (int32_t)d.tcap);
    for (int32_t t = t0 + tid; t < tend; t += BLOCK) d.tworld[t] = s_t0 >= 0 ? w : -1;
}

struct SapRec {
    double ylo, yhi, zlo, zhi;
    uint32_t info;
    int32_t orig;
};

__device__ __forceinline__ SapRec sap_rec(const Dev& d, int64_t s) {
    SapRec r;
    r.ylo = d.sbox[s]; r.yhi = d.sbox[d.ncap + s]; r.zlo = d.sbox[2 * d.ncap + s]; r.zhi = d.sbox[3 * d.ncap + s];
    r.info = d.sinfo[s]; r.orig = d.sorig[s];
    return r;
}

// Pair test of the reference broadphase (broadphase.py:280-294) and the solid
// filter (world.py:663-670) on two sorted records.
__device__ __forceinline__ void sap_test(const Dev& d, const SapRec& A, const SapRec& B, int32_t fo, int32_t nf,
                                         bool& bp, bool& solid) {
    bool keep = (A.yhi >= B.ylo) && (B.yhi >= A.ylo) && (A.zhi >= B.zlo) && (B.zhi >= A.zlo);
    keep = keep && ((A.info | B.info) & SI_DYN);
    if (keep) {
        const bool af = A.orig < B.orig;
        const uint32_t fp = (af ? A.info : B.info) >> 8, fq = (af ? B.info : A.info) >> 8;
        keep = d.filt_tab[fo + (int32_t)fp * nf + (int32_t)fq] != 0;
    }
    bp = keep;
    solid = keep && !((A.info | B.info) & SI_TRIGGER) && (A.info & B.info & SI_COLLIDER) &&
            ((A.info | B.info) & SI_MOVING);
}

template <int BLOCK, bool WRITE>
__global__ __launch_bounds__(BLOCK) void k_sap_pairs(Dev d, int32_t record) {
    constexpr int NW = BLOCK / 64;
    const int lane = threadIdx.x & 63;
    const uint64_t tt = d.cnt[C_TILES];
    const int64_t T = tt < (uint64_t)d.tcap ? (int64_t)tt : d.tcap;
    // Blocks b and b + 8 run on the same XCD: give each XCD's blocks adjacent
    // tile ranges so a world's sorted records stay in one L2.
    const int32_t G = gridDim.x, b = blockIdx.x;
    const int32_t bx = (G % 8 == 0) ? (b % 8) * (G / 8) + b / 8 : b;
    const int64_t waves = (int64_t)G * NW, wv = (int64_t)bx * NW + (threadIdx.x >> 6);
    const int64_t t_lo = T * wv / waves, t_hi = T * (wv + 1) / waves;
    for (int64_t t = t_lo; t < t_hi; t++) {
        const int32_t w = d.tworld[t];
        if (w < 0) {
            if (!WRITE && lane == 0) { d.tcnt[t] = 0; d.tcnt[d.tcap + 1 + t] = 0; }
            continue;
        }
        const int64_t off = d.w_off[w];
        const int32_t n = d.w_cnt[w];
        const int32_t fo = d.w_filt_off[w], nf = d.w_nfilt[w];
        const int32_t* so = d.soff + 2 * off;
        const int32_t X = so[n];
        const int32_t c0 = (int32_t)(t - d.w_t0[w]) * SAP_TILE;
        const int32_t c1 = X - c0 < SAP_TILE ? X : c0 + SAP_TILE;
        // k of candidate c0: the last k with so[k] <= c0
        int32_t kw;
        {
            int32_t lo_i = 0, hi_i = n;
            while (lo_i < hi_i) {
                int32_t mid = (lo_i + hi_i) >> 1;
                if (so[mid] <= c0) lo_i = mid + 1; else hi_i = mid;
            }
            kw = lo_i - 1;
        }
        int64_t out_s = 0, out_b = 0;
        if (WRITE) { out_s = d.tcnt[t]; out_b = d.tcnt[d.tcap + 1 + t]; }
        int32_t cnt_s = 0, cnt_b = 0;
        for (int32_t cb = c0; cb < c1; cb += 64) {
            const int32_t c = cb + lane;
            const bool valid = c < c1;
            // k of this lane's candidate: with v_l = so[kb + 1 + l] - cb (ascending
            // over the lanes), k = kb + #{l : v_l <= lane}; rounds of 64 entries
            // until every valid lane has found its k.
            int32_t k = -1;
            for (int32_t kb = kw;; kb += 64) {
                const int32_t e = kb + 1 + lane;
                const int32_t v = e <= n ? so[e] - cb : 0x7fffffff;
                int32_t m = 0;
#pragma unroll
                for (int step = 32; step > 0; step >>= 1) {
                    const int32_t probe = __shfl(v, m + step - 1, 64);
                    if (probe <= lane) m += step;
                }
                if (m == 63 && __shfl(v, 63, 64) <= lane) m = 64;
                if (k < 0 && m < 64) k = kb + m;
                if (!__any(valid && k < 0)) break;
            }
            bool bp = false, sol = false;
            int64_t p = 0, q = 0;
            if (valid) {
                const int32_t j = k + 1 + (c - so[k]);
                const SapRec A = sap_rec(d, off + k), B = sap_rec(d, off + j);
                sap_test(d, A, B, fo, nf, bp, sol);
                p = off + (A.orig < B.orig ? A.orig : B.orig);
                q = off + (A.orig < B.orig ? B.orig : A.orig);
            }
            const uint64_t ms = __ballot(sol), mb = __ballot(bp);
            const uint64_t below = (1ull << lane) - 1ull;
            if (WRITE) {
                if (sol) {
                    const int64_t o = out_s + cnt_s + __popcll(ms & below);
                    if (o < d.pcap) { d.pa[o] = (int32_t)p; d.pb[o] = (int32_t)q; }
                }
                if (record && bp) {
                    const int64_t o = out_b + cnt_b + __popcll(mb & below);
                    if (o < d.pcap) { d.bpa[o] = (int32_t)p; d.bpb[o] = (int32_t)q; }
                }
            }
            cnt_s += __popcll(ms);
            cnt_b += __popcll(mb);
            kw = __builtin_amdgcn_readlane(k, 63);
        }
        if (!WRITE && lane == 0) { d.tcnt[t] = cnt_s; d.tcnt[d.tcap + 1 + t] = cnt_b; }
    }
}

// Per-tile counts -> exclusive offsets (in place, one block), then every
// world's solid and broadphase pair ranges.
template <int BLOCK>
__global__ __launch_bounds__(BLOCK) void k_sap_scan(Dev d, int32_t record) {
    constexpr int NW = BLOCK / 64, ITEMS = 8;
    __shared__ int64_t s_w[2][NW];
    __shared__ int64_t s_run[2];
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const uint64_t tt = d.cnt[C_TILES];
    const int64_t T = tt < (uint64_t)d.tcap ? (int64_t)tt : d.tcap;
    int64_t* arr[2] = {d.tcnt, d.tcnt + d.tcap + 1};
    if (tid < 2) s_run[tid] = 0;
    __syncthreads();
    for (int64_t base = 0; base < T; base += (int64_t)BLOCK * ITEMS) {
        const int64_t i0 = base + (int64_t)tid * ITEMS;
        int64_t v[2][ITEMS], sum[2] = {0, 0};
        for (int a = 0; a < 2; a++)
            for (int r = 0; r < ITEMS; r++) {
                v[a][r] = i0 + r < T ? arr[a][i0 + r] : 0;
                sum[a] += v[a][r];
            }
        int64_t x[2] = {sum[0], sum[1]};
        for (int a = 0; a < 2; a++) {
#pragma unroll
            for (int o = 1; o < 64; o <<= 1) {
                int64_t y = __shfl_up(x[a], o, 64);
                if (lane >= o) x[a] += y;
            }
            if (lane == 63) s_w[a][wv] = x[a];
        }
        __syncthreads();
        for (int a = 0; a < 2; a++) {
            int64_t wbase = 0, tot = 0;
            for (int q = 0; q < NW; q++) { int64_t s = s_w[a][q]; if (q < wv) wbase += s; tot += s; }
            int64_t run = s_run[a] + wbase + x[a] - sum[a];
            for (int r = 0; r < ITEMS; r++) {
                if (i0 + r < T) arr[a][i0 + r] = run;
                run += v[a][r];
            }
            __syncthreads();
            if (tid == 0) s_run[a] += tot;
        }
        __syncthreads();
    }
    if (tid == 0) {
        arr[0][T] = s_run[0];
        arr[1][T] = s_run[1];
        d.cnt[C_PAIRS] = (uint64_t)s_run[0];
        d.cnt[C_BPTOTAL] = (uint64_t)s_run[1];
        if (record) d.cnt[C_BPAIRS] = (uint64_t)s_run[1];
        if (s_run[0] > d.pcap) atomicOr((unsigned long long*)&d.cnt[C_OVERFLOW], 1ull);
        if (record && s_run[1] > d.pcap) atomicOr((unsigned long long*)&d.cnt[C_OVERFLOW], 8ull);
    }
    __syncthreads();
    for (int32_t w = tid; w < d.wcap; w += BLOCK) {
        const int32_t f = d.w_flags[w];
        const bool live = (f & WF_ACTIVE) && (f & WF_COLLISION) && d.w_tn[w] > 0;
        int64_t ps = 0, pe = 0, bs = 0, be = 0;
        if (live) {
            const int64_t t0 = d.w_t0[w], t1 = t0 + d.w_tn[w];
            ps = arr[0][t0]; pe = arr[0][t1]; bs = arr[1][t0]; be = arr[1][t1];
        }
        d.w_ps[w] = (int32_t)(ps < d.pcap ? ps : d.pcap);
        d.w_pn[w] = pe <= d.pcap ? (int32_t)(pe - ps) : 0;
        d.w_bps[w] = (int32_t)(bs < d.pcap ? bs : d.pcap);
        d.w_bpn[w] = (record && be <= d.pcap) ? (int32_t)(be - bs) : 0;
    }
}

// ----------------------------------------------------------- narrowphase --

// NarrowPhase.generate dispatch (narrowphase.py:78-100) for one solid pair.
// FP64 runs in world coordinates exactly like the reference; FP32 first moves
// the pair into a frame at body a's centre (offset taken in FP64).
template <typename R>
__device__ __forceinline__ int narrow_pair(const Dev& d, int64_t a, int64_t b, R* nrm, R* pts, R* dep,
                                           double* org) {
    const R* rot = (const R*)d.rot;
    const int ka = d.kind[a], kb = d.kind[b];
    R Ra[9], Rb[9], ha[3], hb[3], ca[3], cb[3];
    for (int r = 0; r < 9; r++) { Ra[r] = BODY(rot, r, a); Rb[r] = BODY(rot, r, b); }
    for (int r = 0; r < 3; r++) { ha[r] = (R)BODY(d.half, r, a); hb[r] = (R)BODY(d.half, r, b); }
    for (int r = 0; r < 3; r++) {
        double pa = BODY(d.pos, r, a), pb = BODY(d.pos, r, b);
        if (sizeof(R) == sizeof(double)) { ca[r] = (R)pa; cb[r] = (R)pb; org[r] = 0.0; }
        else { ca[r] = R(0); cb[r] = (R)(pb - pa); org[r] = pa; }
    }
    int cnt = 0;
    if (ka == K_BOX && kb == K_BOX) cnt = box_box(ca, Ra, ha, cb, Rb, hb, nrm, pts, dep);
    else if (ka == K_SPHERE && kb == K_SPHERE) cnt = sphere_sphere(ca, ha[0], cb, hb[0], nrm, pts, dep);
    else if (ka == K_SPHERE && kb == K_BOX) cnt = sphere_box(ca, ha[0], cb, Rb, hb, nrm, pts, dep);
    else if (ka == K_BOX && kb == K_SPHERE) {
        cnt = sphere_box(cb, hb[0], ca, Ra, ha, nrm, pts, dep);
        nrm[0] = -nrm[0]; nrm[1] = -nrm[1]; nrm[2] = -nrm[2];
    }
    return cnt;
}

template <typename R>
__global__ void k_narrow(Dev d) {
    const uint64_t tot = d.cnt[C_PAIRS];
    const int64_t P = tot < (uint64_t)d.pcap ? (int64_t)tot : d.pcap;
    R* mn = (R*)d.mnrm;
    R* md = (R*)d.mdep;
    for (int64_t p = gtid(); p < P; p += gstride()) {
        const int64_t a = d.pa[p], b = d.pb[p];
        if (d.kind[a] == K_CONVEX || d.kind[b] == K_CONVEX) {
            // queued for the GJK/EPA kernel, which writes this pair's manifold
            d.mcnt[p] = 0;
            uint64_t slot = atomicAdd((unsigned long long*)&d.cnt[C_CONVEX], 1ull);
            d.cvx[slot] = (int32_t)p;
            continue;
        }
        R nrm[3] = {R(0), R(0), R(0)}, pts[12], dep[4];
        double org[3];
        int cnt = narrow_pair<R>(d, a, b, nrm, pts, dep, org);
        d.mcnt[p] = (int8_t)cnt;
        for (int r = 0; r < 3; r++) mn[(size_t)r * d.pcap + p] = nrm[r];
        for (int m = 0; m < cnt; m++) {
            for (int r = 0; r < 3; r++) d.mpts[(size_t)r * 4 * d.pcap + 4 * p + m] = org[r] + (double)pts[3 * m + r];
            md[4 * p + m] = dep[m];
        }
    }
}

__device__ __forceinline__ void make_proxy(const Dev& d, int64_t i, Proxy& P) {
    P.kind = d.kind[i];
    for (int r = 0; r < 3; r++) { P.c[r] = BODY(d.pos, r, i); P.h[r] = BODY(d.half, r, i); }
    double q[4] = {BODY(d.quat, 0, i), BODY(d.quat, 1, i), BODY(d.quat, 2, i), BODY(d.quat, 3, i)};
    quat_to_matrix(q, P.R);
    P.local = d.hull + 3 * (size_t)d.hoff[i];
    P.m = P.kind == K_CONVEX ? d.hcnt[i] : 0;
}

// collide.collide fallback to gjk.collide_convex (collide.py:251-269) for the
// pairs k_narrow queued; FP64 in both precisions, one thread per pair, each
// with its own EPA workspace.
template <typename R>
__global__ void k_narrow_convex(Dev d) {
    const uint64_t tot = d.cnt[C_CONVEX];
    const int64_t n = tot < (uint64_t)d.pcap ? (int64_t)tot : d.pcap;
    const int64_t slot = gtid();
    if (slot >= d.epa_slots) return;
    EpaWork* W = (EpaWork*)d.epa_work + slot;
    R* mn = (R*)d.mnrm;
    R* md = (R*)d.mdep;
    for (int64_t q = slot; q < n; q += d.epa_slots) {
        const int64_t p = d.cvx[q];
        Proxy A, B;
        make_proxy(d, d.pa[p], A);
        make_proxy(d, d.pb[p], B);
        double nrm[3] = {0.0, 0.0, 0.0}, pt[3], dep = 0.0;
        int cnt = collide_convex(A, B, W, nrm, pt, dep);
        d.mcnt[p] = (int8_t)cnt;
        for (int r = 0; r < 3; r++) mn[(size_t)r * d.pcap + p] = (R)nrm[r];
        if (cnt) {
            for (int r = 0; r < 3; r++) d.mpts[(size_t)r * 4 * d.pcap + 4 * p] = pt[r];
            md[4 * p] = (R)dep;
        }
    }
}

__global__ void k_stage_convex(int32_t np, const int32_t* ka, const double* ca, const double* Ra, const double* ha,
                               const int32_t* oa, const int32_t* na, const int32_t* kb, const double* cb,
                               const double* Rb, const double* hb, const int32_t* ob, const int32_t* nb,
                               const double* hull, EpaWork* work, int32_t slots, int32_t* cnt, double* nrm,
                               double* pt, double* dep) {
    const int64_t slot = gtid();
    if (slot >= slots) return;
    for (int64_t i = slot; i < np; i += slots) {
        Proxy A, B;
        A.kind = ka[i]; B.kind = kb[i];
        for (int r = 0; r < 3; r++) { A.c[r] = ca[3 * i + r]; A.h[r] = ha[3 * i + r]; B.c[r] = cb[3 * i + r]; B.h[r] = hb[3 * i + r]; }
        for (int r = 0; r < 9; r++) { A.R[r] = Ra[9 * i + r]; B.R[r] = Rb[9 * i + r]; }
        A.local = hull + 3 * (size_t)(A.kind == K_CONVEX ? oa[i] : 0); A.m = A.kind == K_CONVEX ? na[i] : 0;
        B.local = hull + 3 * (size_t)(B.kind == K_CONVEX ? ob[i] : 0); B.m = B.kind == K_CONVEX ? nb[i] : 0;
        double n3[3] = {0.0, 0.0, 0.0}, p3[3] = {0.0, 0.0, 0.0}, dd = 0.0;
        int c = collide_convex(A, B, work + slot, n3, p3, dd);
        cnt[i] = c;
        for (int r = 0; r < 3; r++) { nrm[3 * i + r] = n3[r]; pt[3 * i + r] = p3[r]; }
        dep[i] = dd;
    }
}

// Contacts in the reference's order (pair order, then manifold order) packed
// into one contiguous slice per world.
template <typename R, int BLOCK>
__global__ __launch_bounds__(BLOCK) void k_compact(Dev d, int32_t record) {
    __shared__ int32_t s_scan[BLOCK / 64 + 2];
    __shared__ int64_t s_sum[BLOCK / 64 + 2];
    __shared__ int32_t s_base[2];
    const int32_t w = blockIdx.x;
    const int tid = threadIdx.x;
    if (w >= d.wcap) return;
    const int32_t pn = (d.w_flags[w] & WF_ACTIVE) ? d.w_pn[w] : 0, ps = d.w_ps[w];
    if (pn == 0) { if (tid == 0) d.w_cn[w] = 0; return; }
    int64_t kc = 0, mf = 0;
    for (int32_t i = tid; i < pn; i += BLOCK) { int c = d.mcnt[ps + i]; kc += c; mf += c > 0; }
    const int64_t K = block_sum<BLOCK>(kc, s_sum);
    const int64_t M = block_sum<BLOCK>(mf, s_sum);
    if (tid == 0) {
        uint64_t base = atomicAdd((unsigned long long*)&d.cnt[C_CONTACTS], (unsigned long long)K);
        int32_t ok = base + K <= (uint64_t)d.kcap;
        if (!ok) atomicOr((unsigned long long*)&d.cnt[C_OVERFLOW], 2ull);
        s_base[0] = (int32_t)base; s_base[1] = ok;
        d.w_cs[w] = (int32_t)base; d.w_cn[w] = ok ? (int32_t)K : 0;
        atomicAdd((unsigned long long*)&d.cnt[C_MANIFOLDS], (unsigned long long)M);
    }
    __syncthreads();
    if (!s_base[1]) return;
    int32_t run = s_base[0];
    const R* mn = (const R*)d.mnrm;
    const R* md = (const R*)d.mdep;
    R* cn = (R*)d.cnrm;
    R* cdp = (R*)d.cdep;
    R* imp = (R*)d.cimp;
    for (int32_t cb = 0; cb < pn; cb += BLOCK) {
        const int32_t i = cb + tid;
        const int32_t p = ps + i;
        const int32_t c = i < pn ? d.mcnt[p] : 0;
        // exclusive prefix of the counts over the chunk (counts are 0..4)
        int32_t x = c;
        const int lane = tid & 63, wv = tid >> 6;
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) { int32_t y = __shfl_up(x, o, 64); if (lane >= o) x += y; }
        if (lane == 63) s_scan[wv] = x;
        __syncthreads();
        int32_t wbase = 0, tot = 0;
        for (int q = 0; q < BLOCK / 64; q++) { int32_t s = s_scan[q]; if (q < wv) wbase += s; tot += s; }
        const int32_t dst = run + wbase + x - c;
        for (int m = 0; m < c; m++) {
            const int64_t k = dst + m;
            d.ca[k] = d.pa[p]; d.cb[k] = d.pb[p];
            for (int r = 0; r < 3; r++) {
                CON(d.cpt, r, k) = d.mpts[(size_t)r * 4 * d.pcap + 4 * (size_t)p + m];
                CON(cn, r, k) = mn[(size_t)r * d.pcap + p];
                CON(imp, r, k) = R(0);
            }
            cdp[k] = md[4 * (size_t)p + m];
        }
        run += tot;
        __syncthreads();
    }
}

// ------------------------------------------------------------------ wake --

// SequentialImpulseSolver._wake_sleepers (solver.py:321-337).  Waking is
// idempotent and, unless a sleeping dynamic body carries a wake-grade speed
// (flag WF_SEQWAKE, set by k_pre), independent of contact order, so the
// contacts are processed in parallel; otherwise one thread replays the
// reference's sequential loop.
template <int BLOCK>
__global__ __launch_bounds__(BLOCK) void k_wake(Dev d) {
    const int32_t w = blockIdx.x;
    if (w >= d.wcap || !(d.w_flags[w] & WF_ACTIVE)) return;
    const int32_t K = d.w_cn[w], cs = d.w_cs[w];
    if (K == 0) return;
    const double th = 0.15 * 0.15;
    auto v2 = [&](int64_t i) {
        double x = BODY(d.lv, 0, i), y = BODY(d.lv, 1, i), z = BODY(d.lv, 2, i);
        return (x * x + y * y) + z * z;
    };
    if (d.w_flags[w] & WF_SEQWAKE) {
        if (threadIdx.x != 0) return;
        for (int32_t k = 0; k < K; k++) {
            const int64_t a = d.ca[cs + k], b = d.cb[cs + k];
            const int ma = d.mt[a], mb = d.mt[b];
            bool wa = ma == M_KINEMATIC || (ma == M_DYNAMIC && d.awake[a] && v2(a) > th);
            if (!d.awake[b] && wa) { d.awake[b] = 1; d.timer[b] = 0.0; }
            bool wb = mb == M_KINEMATIC || (mb == M_DYNAMIC && d.awake[b] && v2(b) > th);
            if (!d.awake[a] && wb) { d.awake[a] = 1; d.timer[a] = 0.0; }
        }
        return;
    }
    for (int32_t k = threadIdx.x; k < K; k += BLOCK) {
        const int64_t a = d.ca[cs + k], b = d.cb[cs + k];
        const int ma = d.mt[a], mb = d.mt[b];
        const bool wa = ma == M_KINEMATIC || (ma == M_DYNAMIC && d.awake0[a] && v2(a) > th);
        const bool wb = mb == M_KINEMATIC || (mb == M_DYNAMIC && d.awake0[b] && v2(b) > th);
        if (!d.awake0[b] && wa) { d.awake[b] = 1; d.timer[b] = 0.0; }
        if (!d.awake0[a] && wb) { d.awake[a] = 1; d.timer[a] = 0.0; }
    }
}

// ---------------------------------------------------------------- solver --

// Solver row layout per contact (R [SR_N][kcap]).
enum { SR_RA = 0, SR_RB = 3, SR_T1 = 6, SR_T2 = 9, SR_NM = 12, SR_T1M = 13, SR_T2M = 14, SR_VB = 15,
       SR_FR = 16, SR_IMA = 17, SR_IMB = 18, SR_NX = 19, SR_IMP = 22, SR_DEP = 25, SR_N = 26 };

// _solver_native._effmass (_solver_native.pyx:154-176)
template <typename R>
__device__ __forceinline__ R effmass(const R* ra, const R* rb, R dx, R dy, R dz, const R* Ia, const R* Ib,
                                     R ima, R imb) {
    R ax = ra[1] * dz - ra[2] * dy, ay = ra[2] * dx - ra[0] * dz, az = ra[0] * dy - ra[1] * dx;
    R bx = rb[1] * dz - rb[2] * dy, by = rb[2] * dx - rb[0] * dz, bz = rb[0] * dy - rb[1] * dx;
    R kk = ima + imb, wx, wy, wz;
    wx = Ia[0] * ax + Ia[1] * ay + Ia[2] * az;
    wy = Ia[3] * ax + Ia[4] * ay + Ia[5] * az;
    wz = Ia[6] * ax + Ia[7] * ay + Ia[8] * az;
    kk += ax * wx + ay * wy + az * wz;
    wx = Ib[0] * bx + Ib[1] * by + Ib[2] * bz;
    wy = Ib[3] * bx + Ib[4] * by + Ib[5] * bz;
    wz = Ib[6] * bx + Ib[7] * by + Ib[8] * bz;
    kk += bx * wx + by * wy + bz * wz;
    return kk > R(1e-12) ? R(1) / kk : R(0);
}

// _apply_w (_solver_native.pyx:132-151) on register copies of the two bodies.
template <typename R>
__device__ __forceinline__ void apply_w(R* va, R* vb, R ima, R imb, const R* Ia, const R* Ib, const R* rA,
                                        const R* rB, R Px, R Py, R Pz) {
    R crx, cry, crz;
    va[0] -= ima * Px; va[1] -= ima * Py; va[2] -= ima * Pz;
    vb[0] += imb * Px; vb[1] += imb * Py; vb[2] += imb * Pz;
    crx = rA[1] * Pz - rA[2] * Py; cry = rA[2] * Px - rA[0] * Pz; crz = rA[0] * Py - rA[1] * Px;
    va[3] -= Ia[0] * crx + Ia[1] * cry + Ia[2] * crz;
    va[4] -= Ia[3] * crx + Ia[4] * cry + Ia[5] * crz;
    va[5] -= Ia[6] * crx + Ia[7] * cry + Ia[8] * crz;
    crx = rB[1] * Pz - rB[2] * Py; cry = rB[2] * Px - rB[0] * Pz; crz = rB[0] * Py - rB[1] * Px;
    vb[3] += Ib[0] * crx + Ib[1] * cry + Ib[2] * crz;
    vb[4] += Ib[3] * crx + Ib[4] * cry + Ib[5] * crz;
    vb[5] += Ib[6] * crx + Ib[7] * cry + Ib[8] * crz;
}

template <typename R>
__device__ __forceinline__ void relvel(const R* va, const R* vb, const R* rA, const R* rB, R& x, R& y, R& z) {
    x = (vb[0] + vb[4] * rB[2] - vb[5] * rB[1]) - (va[0] + va[4] * rA[2] - va[5] * rA[1]);
    y = (vb[1] + vb[5] * rB[0] - vb[3] * rB[2]) - (va[1] + va[5] * rA[0] - va[3] * rA[2]);
    z = (vb[2] + vb[3] * rB[1] - vb[4] * rB[0]) - (va[2] + va[3] * rA[1] - va[4] * rA[0]);
}

// Cached (nImp, t0, t1) for contact k, solver._seed_pair (solver.py:222-233):
// nearest cached point of the same (a, b) with d² < 0.04, first strict min.
template <typename R>
__device__ __forceinline__ void seed_lookup(const Dev& d, int32_t w, int32_t old, int64_t ci, int64_t a, int64_t b,
                                            R& n0, R& t0, R& t1) {
    const uint64_t key = ((uint64_t)(uint32_t)a << 32) | (uint32_t)b;
    const int64_t hm = d.hcap - 1;
    int64_t h = (int64_t)(hash_mix(key) & (uint64_t)hm);
    const int32_t rs = d.w_cache_s[(size_t)old * d.wcap + w], rn = d.w_cache_n[(size_t)old * d.wcap + w];
    for (int probe = 0; probe < 64; probe++) {
        const uint64_t kk = d.hkey[old][h];
        if (kk == ~0ull) return;
        if (kk == key) {
            int32_t v = d.hval[old][h];
            if (v < rs || v >= rs + rn) return;
            const R* kimp = (const R*)d.kimp[old];
            double best = 0.04;
            for (int32_t q = v; q < rs + rn && d.ka[old][q] == a && d.kb[old][q] == b; q++) {
                double dx = CON(d.cpt, 0, ci) - CON(d.kpt[old], 0, q);
                double dy = CON(d.cpt, 1, ci) - CON(d.kpt[old], 1, q);
                double dz = CON(d.cpt, 2, ci) - CON(d.kpt[old], 2, q);
                double dd = dx * dx + dy * dy + dz * dz;
                if (dd < best) { best = dd; n0 = CON(kimp, 0, q); t0 = CON(kimp, 1, q); t1 = CON(kimp, 2, q); }
            }
            return;
        }
        h = (h + 1) & hm;
    }
}

// One world's contacts solved exactly in the reference's Gauss-Seidel order
// (_solver_native.pyx:179-312).  Contact k depends only on the previous contact
// (in sequence, across sweeps) that touched one of its dynamic bodies; level
// L(k) = 1 + max(L of those) puts every dependency on an earlier level, so the
// contacts of one level touch disjoint dynamic bodies and run concurrently
// while every contact still reads exactly the velocities the sequential loop
// would.  Static/kinematic bodies are read-only (their inverse mass and
// inertia are zero, so the reference's writes to them leave them unchanged).
template <typename R, int BLOCK, bool STAGE>
__global__ __launch_bounds__(BLOCK) void k_solve(Dev d, SolveIn in, const int32_t* wlist, const int32_t* wcount,
                                                 int32_t lds_last, int32_t lds_state) {
    extern __shared__ __align__(16) unsigned char s_dyn[];
    __shared__ int32_t s_scan[BLOCK / 64 + 2];
    __shared__ int32_t s_D;
    if ((int32_t)blockIdx.x >= *wcount) return;
    const int32_t w = wlist[blockIdx.x];
    const int tid = threadIdx.x;
    if (w >= d.wcap || !(d.w_flags[w] & WF_ACTIVE)) return;
    const int32_t K = d.w_cn[w], cs = d.w_cs[w], n = d.w_cnt[w];
    const int64_t off = d.w_off[w];
    if (K == 0) { if (tid == 0) d.w_levels[w] = 0; return; }
    const bool fp64 = sizeof(R) == sizeof(double);
    // Where the bodies' velocities and world inverse inertia live while the
    // world is solved: in LDS when the world fits the launch's budget, else the
    // FP64 state itself (FP64) or an FP32 copy (FP32) in HBM.
    const bool state_lds = lds_state >= n;
    int32_t* s_last = (int32_t*)s_dyn;
    R* s_state = (R*)(s_dyn + (((size_t)lds_last * 4 + 15) & ~(size_t)15));
    R* V;
    const R* I;
    int64_t vs, vb;
    const R* iiw = (const R*)d.iiw;
    if (state_lds) { V = s_state; I = s_state + 6 * (size_t)n; vs = n; vb = off; }
    else { V = fp64 ? (R*)d.lv : (R*)d.svel; I = iiw; vs = d.ncap; vb = 0; }
#define VEL(c, i) V[(size_t)(c) * vs + (size_t)((i) - vb)]
#define IIW(c, i) I[(size_t)(c) * vs + (size_t)((i) - vb)]
    if (state_lds || !fp64)
        for (int32_t i = tid; i < n; i += BLOCK) {
            for (int c = 0; c < 6; c++) VEL(c, off + i) = (R)BODY(d.lv, c, off + i);
            if (state_lds)
                for (int c = 0; c < 9; c++) s_state[(size_t)(6 + c) * n + i] = BODY(iiw, c, off + i);
        }
    R* S = (R*)d.csolv;
    R* imp = (R*)d.cimp;
    const R* cn = (const R*)d.cnrm;
    const R* cdep = (const R*)d.cdep;
    int32_t* sa = d.sab;
    int32_t* sb = d.sab + d.kcap;

    // Level schedule.  Wave 0 walks the contacts 64 at a time: lane i starts
    // from the last levels recorded for its dynamic bodies, then takes the
    // final levels of the earlier lanes in this group (broadcast in order) and
    // raises its own past any that share a dynamic body.  `last` (per body,
    // LDS for worlds that fit) keeps the running maximum with atomicMax.
    int32_t* last = lds_last >= n ? s_last : d.last + off;
    for (int32_t i = tid; i < n; i += BLOCK) last[i] = 0;
    for (int32_t k = tid; k < K; k += BLOCK) d.lcount[cs + k] = 0;
    __syncthreads();
    if (tid < 64) {
        int32_t D = 0;
        for (int32_t base = 0; base < K; base += 64) {
            const int32_t k = base + tid;
            const bool valid = k < K;
            int32_t al = -1, bl = -1;
            if (valid) {
                const int64_t a = d.ca[cs + k], b = d.cb[cs + k];
                if (d.mt[a] == M_DYNAMIC) al = (int32_t)(a - off);
                if (d.mt[b] == M_DYNAMIC) bl = (int32_t)(b - off);
            }
            // atomic loads: a global-memory `last` must not be read from a stale L1 line
            const int32_t la = al >= 0 ? __hip_atomic_load(&last[al], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) : 0;
            const int32_t lb = bl >= 0 ? __hip_atomic_load(&last[bl], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) : 0;
            int32_t L = (la > lb ? la : lb) + 1;
            const int32_t m = K - base < 64 ? K - base : 64;
            for (int32_t j = 0; j < m - 1; j++) {
                const int32_t aj = __builtin_amdgcn_readlane(al, j);
                const int32_t bj = __builtin_amdgcn_readlane(bl, j);
                const int32_t Lj = __builtin_amdgcn_readlane(L, j);
                const bool share = (al >= 0 && (al == aj || al == bj)) || (bl >= 0 && (bl == aj || bl == bj));
                if (tid > j && share && Lj + 1 > L) L = Lj + 1;
            }
            if (valid) {
                if (al >= 0) atomicMax(&last[al], L);
                if (bl >= 0) atomicMax(&last[bl], L);
                d.clevel[cs + k] = L;
                D = L > D ? L : D;
            }
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_s_waitcnt(0);
        }
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) { int32_t x = __shfl_xor(D, o, 64); D = x > D ? x : D; }
        if (tid == 0) {
            s_D = D;
            d.w_levels[w] = D;
            atomicMax((unsigned long long*)&d.cnt[C_MAXLEVEL], (unsigned long long)D);
        }
    }
    __syncthreads();
    for (int32_t k = tid; k < K; k += BLOCK) atomicAdd(&d.lcount[cs + d.clevel[cs + k] - 1], 1);
    __syncthreads();
    const int32_t D = s_D;
    block_exscan<BLOCK>(d.lcount + cs, D, s_scan);
    for (int32_t k = tid; k < K; k += BLOCK) {
        const int32_t L = d.clevel[cs + k];
        const int32_t slot = atomicAdd(&d.lcount[cs + L - 1], 1);
        d.corder[cs + slot] = k;
    }
    __syncthreads();
    // after the scatter lcount[l] is the end of level l

    // Prepare every contact from the pre-solve velocities, writing its row at
    // its level-sorted slot so each level's rows are contiguous.
    const int32_t old = *d.sel;
    const int32_t mo = d.w_mat_off[w], nm = d.w_nmat[w];
    for (int32_t t = tid; t < K; t += BLOCK) {
        const int32_t k = d.corder[cs + t];
        const int64_t ci = cs + k, si = cs + t, a = d.ca[ci], b = d.cb[ci];
        R rest, fric, n0 = R(0), u0 = R(0), w0 = R(0);
        if (STAGE) {
            rest = (R)in.st_rest[k]; fric = (R)in.st_fric[k];
            n0 = CON(imp, 0, ci); u0 = CON(imp, 1, ci); w0 = CON(imp, 2, ci);
        } else {
            const int32_t tb = mo + d.mat[a] * nm + d.mat[b];
            rest = (R)d.mat_rest[tb]; fric = (R)d.mat_fric[tb];
            if (in.warm) seed_lookup<R>(d, w, old, ci, a, b, n0, u0, w0);
        }
        const R ima = (R)d.inv_mass[a], imb = (R)d.inv_mass[b];
        R rA[3], rB[3], Ia[9], Ib[9], va[6], vb_[6];
        for (int r = 0; r < 3; r++) {
            double p = CON(d.cpt, r, ci);
            rA[r] = (R)(p - BODY(d.pos, r, a));
            rB[r] = (R)(p - BODY(d.pos, r, b));
        }
        for (int r = 0; r < 9; r++) { Ia[r] = IIW(r, a); Ib[r] = IIW(r, b); }
        for (int c = 0; c < 6; c++) { va[c] = VEL(c, a); vb_[c] = VEL(c, b); }
        const R nx = CON(cn, 0, ci), ny = CON(cn, 1, ci), nz = CON(cn, 2, ci);
        R axx, ayy, azz;
        if ((nx >= R(0) ? nx : -nx) < R(0.9)) { axx = R(1); ayy = R(0); azz = R(0); }
        else { axx = R(0); ayy = R(1); azz = R(0); }
        R tx = ayy * nz - azz * ny, ty = azz * nx - axx * nz, tz = axx * ny - ayy * nx;
        R ln = sqrt(tx * tx + ty * ty + tz * tz);
        R inv = ln > R(1e-12) ? R(1) / ln : R(0);
        tx *= inv; ty *= inv; tz *= inv;
        const R ux = ny * tz - nz * ty, uy = nz * tx - nx * tz, uz = nx * ty - ny * tx;
        R dvx, dvy, dvz;
        relvel(va, vb_, rA, rB, dvx, dvy, dvz);
        const R vn0 = dvx * nx + dvy * ny + dvz * nz;
        for (int r = 0; r < 3; r++) { CON(S, SR_RA + r, si) = rA[r]; CON(S, SR_RB + r, si) = rB[r]; }
        CON(S, SR_T1, si) = tx; CON(S, SR_T1 + 1, si) = ty; CON(S, SR_T1 + 2, si) = tz;
        CON(S, SR_T2, si) = ux; CON(S, SR_T2 + 1, si) = uy; CON(S, SR_T2 + 2, si) = uz;
        CON(S, SR_NM, si) = effmass(rA, rB, nx, ny, nz, Ia, Ib, ima, imb);
        CON(S, SR_T1M, si) = effmass(rA, rB, tx, ty, tz, Ia, Ib, ima, imb);
        CON(S, SR_T2M, si) = effmass(rA, rB, ux, uy, uz, Ia, Ib, ima, imb);
        CON(S, SR_VB, si) = vn0 < -(R)in.rthr ? -rest * vn0 : R(0);
        CON(S, SR_FR, si) = fric;
        CON(S, SR_IMA, si) = ima; CON(S, SR_IMB, si) = imb;
        CON(S, SR_NX, si) = nx; CON(S, SR_NX + 1, si) = ny; CON(S, SR_NX + 2, si) = nz;
        CON(S, SR_IMP, si) = n0; CON(S, SR_IMP + 1, si) = u0; CON(S, SR_IMP + 2, si) = w0;
        CON(S, SR_DEP, si) = cdep[ci];
        sa[si] = (int32_t)a; sb[si] = (int32_t)b;
    }
    __syncthreads();
    auto level_begin = [&](int32_t l) { return l == 0 ? 0 : d.lcount[cs + l - 1]; };
    auto body_store = [&](int64_t a, int64_t b, const R* va, const R* vb2) {
        if (d.mt[a] == M_DYNAMIC) for (int c = 0; c < 6; c++) VEL(c, a) = va[c];
        if (d.mt[b] == M_DYNAMIC) for (int c = 0; c < 6; c++) VEL(c, b) = vb2[c];
    };
    // warm start
    if (in.warm) {
        for (int32_t l = 0; l < D; l++) {
            const int32_t e = d.lcount[cs + l];
            for (int32_t t = level_begin(l) + tid; t < e; t += BLOCK) {
                const int64_t si = cs + t, a = sa[si], b = sb[si];
                R va[6], vb2[6], rA[3], rB[3], Ia[9], Ib[9];
                for (int c = 0; c < 6; c++) { va[c] = VEL(c, a); vb2[c] = VEL(c, b); }
                for (int r = 0; r < 3; r++) { rA[r] = CON(S, SR_RA + r, si); rB[r] = CON(S, SR_RB + r, si); }
                for (int r = 0; r < 9; r++) { Ia[r] = IIW(r, a); Ib[r] = IIW(r, b); }
                const R n0 = CON(S, SR_IMP, si), u0 = CON(S, SR_IMP + 1, si), w0 = CON(S, SR_IMP + 2, si);
                const R P0 = n0 * CON(S, SR_NX, si) + u0 * CON(S, SR_T1, si) + w0 * CON(S, SR_T2, si);
                const R P1 = n0 * CON(S, SR_NX + 1, si) + u0 * CON(S, SR_T1 + 1, si) + w0 * CON(S, SR_T2 + 1, si);
                const R P2 = n0 * CON(S, SR_NX + 2, si) + u0 * CON(S, SR_T1 + 2, si) + w0 * CON(S, SR_T2 + 2, si);
                apply_w(va, vb2, CON(S, SR_IMA, si), CON(S, SR_IMB, si), Ia, Ib, rA, rB, P0, P1, P2);
                body_store(a, b, va, vb2);
            }
            __syncthreads();
        }
    }
    const int32_t iters = d.w_iters[w];
    for (int32_t it = 0; it < iters; it++) {
        for (int32_t l = 0; l < D; l++) {
            const int32_t e = d.lcount[cs + l];
            for (int32_t t = level_begin(l) + tid; t < e; t += BLOCK) {
                const int64_t si = cs + t, a = sa[si], b = sb[si];
                R va[6], vb2[6], rA[3], rB[3], Ia[9], Ib[9], t1[3], t2[3];
                for (int c = 0; c < 6; c++) { va[c] = VEL(c, a); vb2[c] = VEL(c, b); }
                for (int r = 0; r < 3; r++) {
                    rA[r] = CON(S, SR_RA + r, si); rB[r] = CON(S, SR_RB + r, si);
                    t1[r] = CON(S, SR_T1 + r, si); t2[r] = CON(S, SR_T2 + r, si);
                }
                for (int r = 0; r < 9; r++) { Ia[r] = IIW(r, a); Ib[r] = IIW(r, b); }
                const R ima = CON(S, SR_IMA, si), imb = CON(S, SR_IMB, si);
                const R nx = CON(S, SR_NX, si), ny = CON(S, SR_NX + 1, si), nz = CON(S, SR_NX + 2, si);
                R nI = CON(S, SR_IMP, si), tI0 = CON(S, SR_IMP + 1, si), tI1 = CON(S, SR_IMP + 2, si);
                R dvx, dvy, dvz;
                relvel(va, vb2, rA, rB, dvx, dvy, dvz);
                R vn = dvx * nx + dvy * ny + dvz * nz;
                R nw = nI - CON(S, SR_NM, si) * (vn - CON(S, SR_VB, si));
                if (nw < R(0)) nw = R(0);
                R dL = nw - nI;
                nI = nw;
                apply_w(va, vb2, ima, imb, Ia, Ib, rA, rB, dL * nx, dL * ny, dL * nz);
                const R maxf = CON(S, SR_FR, si) * nI;
                relvel(va, vb2, rA, rB, dvx, dvy, dvz);
                R vt = dvx * t1[0] + dvy * t1[1] + dvz * t1[2];
                R nt = tI0 - CON(S, SR_T1M, si) * vt;
                if (nt > maxf) nt = maxf; else if (nt < -maxf) nt = -maxf;
                dL = nt - tI0; tI0 = nt;
                apply_w(va, vb2, ima, imb, Ia, Ib, rA, rB, dL * t1[0], dL * t1[1], dL * t1[2]);
                vt = dvx * t2[0] + dvy * t2[1] + dvz * t2[2];
                nt = tI1 - CON(S, SR_T2M, si) * vt;
                if (nt > maxf) nt = maxf; else if (nt < -maxf) nt = -maxf;
                dL = nt - tI1; tI1 = nt;
                apply_w(va, vb2, ima, imb, Ia, Ib, rA, rB, dL * t2[0], dL * t2[1], dL * t2[2]);
                CON(S, SR_IMP, si) = nI; CON(S, SR_IMP + 1, si) = tI0; CON(S, SR_IMP + 2, si) = tI1;
                body_store(a, b, va, vb2);
            }
            __syncthreads();
        }
    }
    // split-impulse position pass, same order
    for (int32_t l = 0; l < D; l++) {
        const int32_t e = d.lcount[cs + l];
        for (int32_t t = level_begin(l) + tid; t < e; t += BLOCK) {
            const int64_t si = cs + t;
            const R vn = CON(S, SR_DEP, si) - (R)in.slop;
            if (vn <= R(0)) continue;
            const R ia = CON(S, SR_IMA, si), ib = CON(S, SR_IMB, si);
            if (ia + ib <= R(1e-12)) continue;
            const R dL = (R)in.correction * vn / (ia + ib);
            const int64_t a = sa[si], b = sb[si];
            if (d.mt[a] == M_DYNAMIC)
                for (int r = 0; r < 3; r++) {
                    if (fp64) BODY(d.pos, r, a) -= (double)(ia * dL * CON(S, SR_NX + r, si));
                    else BODY(d.pos, r, a) = BODY(d.pos, r, a) - (double)(ia * dL * CON(S, SR_NX + r, si));
                }
            if (d.mt[b] == M_DYNAMIC)
                for (int r = 0; r < 3; r++) {
                    if (fp64) BODY(d.pos, r, b) += (double)(ib * dL * CON(S, SR_NX + r, si));
                    else BODY(d.pos, r, b) = BODY(d.pos, r, b) + (double)(ib * dL * CON(S, SR_NX + r, si));
                }
        }
        __syncthreads();
    }
    // impulses back to contact order (warm-start cache and readback)
    for (int32_t t = tid; t < K; t += BLOCK) {
        const int64_t ci = cs + d.corder[cs + t], si = cs + t;
        for (int r = 0; r < 3; r++) CON(imp, r, ci) = CON(S, SR_IMP + r, si);
    }
    // velocities back to the FP64 state: copied from LDS (FP64), or the FP32
    // solver's change folded in
    if (state_lds || !fp64)
        for (int32_t i = tid; i < n; i += BLOCK)
            if (d.mt[off + i] == M_DYNAMIC)
                for (int c = 0; c < 6; c++) {
                    if (fp64) {
                        BODY(d.lv, c, off + i) = (double)VEL(c, off + i);
                    } else {
                        double v0 = BODY(d.lv, c, off + i);
                        R dv = VEL(c, off + i) - (R)v0;
                        BODY(d.lv, c, off + i) = v0 + (double)dv;
                    }
                }
#undef VEL
#undef IIW
}


// ----------------------------------------------------------------- joints --

// Level schedule shared shape (see k_solve): wave 0 assigns levels 64 items at
// a time, then every thread counts, scans and scatters the items by level.
// get(k, al, bl) gives item k's world-local dynamic body indices (-1 = none).
template <int BLOCK, typename Get>
__device__ int32_t assign_levels(int32_t K, Get get, int32_t* last, int32_t* level, int32_t* lcount,
                                 int32_t* order, int32_t* s_scan, int32_t* s_D) {
    const int tid = threadIdx.x;
    for (int32_t k = tid; k < K; k += BLOCK) lcount[k] = 0;
    __syncthreads();
    if (tid < 64) {
        int32_t D = 0;
        for (int32_t base = 0; base < K; base += 64) {
            const int32_t k = base + tid;
            const bool valid = k < K;
            int32_t al = -1, bl = -1;
            if (valid) get(k, al, bl);
            const int32_t la = al >= 0 ? __hip_atomic_load(&last[al], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) : 0;
            const int32_t lb = bl >= 0 ? __hip_atomic_load(&last[bl], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) : 0;
            int32_t L = (la > lb ? la : lb) + 1;
            const int32_t m = K - base < 64 ? K - base : 64;
            for (int32_t j = 0; j < m - 1; j++) {
                const int32_t aj = __builtin_amdgcn_readlane(al, j);
                const int32_t bj = __builtin_amdgcn_readlane(bl, j);
                const int32_t Lj = __builtin_amdgcn_readlane(L, j);
                const bool share = (al >= 0 && (al == aj || al == bj)) || (bl >= 0 && (bl == aj || bl == bj));
                if (tid > j && share && Lj + 1 > L) L = Lj + 1;
            }
            if (valid) {
                if (al >= 0) atomicMax(&last[al], L);
                if (bl >= 0) atomicMax(&last[bl], L);
                level[k] = L;
                D = L > D ? L : D;
            }
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_s_waitcnt(0);
        }
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) { int32_t x = __shfl_xor(D, o, 64); D = x > D ? x : D; }
        if (tid == 0) *s_D = D;
    }
    __syncthreads();
    const int32_t D = *s_D;
    for (int32_t k = tid; k < K; k += BLOCK) atomicAdd(&lcount[level[k] - 1], 1);
    __syncthreads();
    block_exscan<BLOCK>(lcount, D, s_scan);
    for (int32_t k = tid; k < K; k += BLOCK) order[atomicAdd(&lcount[level[k] - 1], 1)] = k;
    __syncthreads();
    return D;
}

// 3x3 solve by LU with partial pivoting, LAPACK dgetf2 + dgetrs order
// (np.linalg.solve in PointConstraint.solve, joints.py:119).
template <typename R>
__device__ __forceinline__ void solve3(const R* K, const R* rhs, R* x) {
    R A[3][3], b[3];
    for (int r = 0; r < 3; r++) { b[r] = rhs[r]; for (int c = 0; c < 3; c++) A[r][c] = K[3 * r + c]; }
    int piv[3];
#pragma unroll
    for (int j = 0; j < 3; j++) {
        int p = j;
        R m = abs_(A[j][j]);
        for (int i = j + 1; i < 3; i++) if (abs_(A[i][j]) > m) { m = abs_(A[i][j]); p = i; }
        piv[j] = p;
        if (A[p][j] != R(0)) {
            if (p != j) for (int c = 0; c < 3; c++) { R t = A[j][c]; A[j][c] = A[p][c]; A[p][c] = t; }
            R inv = R(1) / A[j][j];
            for (int i = j + 1; i < 3; i++) A[i][j] *= inv;
        }
        for (int i = j + 1; i < 3; i++)
            for (int c = j + 1; c < 3; c++) A[i][c] -= A[i][j] * A[j][c];
    }
    for (int j = 0; j < 3; j++) if (piv[j] != j) { R t = b[j]; b[j] = b[piv[j]]; b[piv[j]] = t; }
    for (int i = 1; i < 3; i++) for (int k = 0; k < i; k++) b[i] -= A[i][k] * b[k];
    for (int i = 2; i >= 0; i--) {
        for (int k = i + 1; k < 3; k++) b[i] -= A[i][k] * b[k];
        b[i] /= A[i][i];
    }
    x[0] = b[0]; x[1] = b[1]; x[2] = b[2];
}

// skew(r) @ I @ skew(r) (joints.py:20-23, 96-97)
template <typename R>
__device__ __forceinline__ void skew_sandwich(const R* r, const R* I, R* o) {
    const R S[9] = {R(0), -r[2], r[1], r[2], R(0), -r[0], -r[1], r[0], R(0)};
    R M[9];
    for (int i = 0; i < 3; i++)
        for (int c = 0; c < 3; c++) M[3 * i + c] = (S[3 * i] * I[c] + S[3 * i + 1] * I[3 + c]) + S[3 * i + 2] * I[6 + c];
    for (int i = 0; i < 3; i++)
        for (int c = 0; c < 3; c++) o[3 * i + c] = (M[3 * i] * S[c] + M[3 * i + 1] * S[3 + c]) + M[3 * i + 2] * S[6 + c];
}

enum { JR_K = 0, JR_RA = 9, JR_RB = 12, JR_BIAS = 15, JR_N = 18 };

// PhysicsWorld._solve_joints (world.py:599-611): wake the joint bodies,
// PointConstraint.prepare every joint (joints.py:84-101), then
// joint_iterations Gauss-Seidel sweeps of PointConstraint.solve
// (joints.py:111-121) in joint order, run level by level like the contacts.
template <typename R, int BLOCK>
__global__ __launch_bounds__(BLOCK) void k_joints(Dev d, double dt) {
    __shared__ int32_t s_scan[BLOCK / 64 + 2];
    __shared__ int32_t s_D;
    const int32_t w = blockIdx.x;
    const int tid = threadIdx.x;
    if (w >= d.wcap || (d.w_flags[w] & (WF_ACTIVE | WF_JOINTS)) != (WF_ACTIVE | WF_JOINTS)) return;
    const int32_t J = d.w_jn[w], js = d.w_js[w], n = d.w_cnt[w];
    const int64_t off = d.w_off[w];
    if (J == 0) return;
    const bool fp64 = sizeof(R) == sizeof(double);
    R* row = (R*)d.jrow;
    const R* iiw = (const R*)d.iiw;
#define JROW(c, j) row[(size_t)(c) * d.jcap + (size_t)(j)]
    for (int32_t t = tid; t < J; t += BLOCK) {
        const int64_t j = js + t;
        const int32_t a = d.ja[j], b = d.jb[j];
        if (a >= 0) { d.awake[a] = 1; d.timer[a] = 0.0; }
        if (b >= 0) { d.awake[b] = 1; d.timer[b] = 0.0; }
        R Ia[9], Ib[9], rA[3], rB[3];
        double rAd[3], rBd[3];
        uint8_t has = d.jhas[j];
        for (int s2 = 0; s2 < 2; s2++) {
            const int32_t i = s2 ? b : a;
            R* I = s2 ? Ib : Ia;
            double* rr = s2 ? rBd : rAd;
            double* loc = s2 ? d.jlb : d.jla;
            if (i < 0) {
                for (int r = 0; r < 9; r++) I[r] = R(0);
                for (int r = 0; r < 3; r++) rr[r] = 0.0;
                continue;
            }
            for (int r = 0; r < 9; r++) I[r] = BODY(iiw, r, i);
            double q[4] = {BODY(d.quat, 0, i), BODY(d.quat, 1, i), BODY(d.quat, 2, i), BODY(d.quat, 3, i)}, Rm[9];
            quat_to_matrix(q, Rm);
            const double p[3] = {BODY(d.pos, 0, i), BODY(d.pos, 1, i), BODY(d.pos, 2, i)};
            if (!(has & (1 << s2))) {
                const double dd[3] = {d.janc[j] - p[0], d.janc[d.jcap + j] - p[1], d.janc[2 * d.jcap + j] - p[2]};
                for (int c = 0; c < 3; c++)
                    loc[(size_t)c * d.jcap + j] = (Rm[c] * dd[0] + Rm[3 + c] * dd[1]) + Rm[6 + c] * dd[2];
                has |= (uint8_t)(1 << s2);
            }
            const double l[3] = {loc[j], loc[d.jcap + j], loc[2 * d.jcap + j]};
            for (int r = 0; r < 3; r++) rr[r] = (p[r] + ((Rm[3 * r] * l[0] + Rm[3 * r + 1] * l[1]) + Rm[3 * r + 2] * l[2])) - p[r];
        }
        d.jhas[j] = has;
        for (int r = 0; r < 3; r++) { rA[r] = (R)rAd[r]; rB[r] = (R)rBd[r]; }
        const R ima = a < 0 ? R(0) : (R)d.inv_mass[a], imb = b < 0 ? R(0) : (R)d.inv_mass[b];
        R SA[9], SB[9];
        skew_sandwich(rA, Ia, SA);
        skew_sandwich(rB, Ib, SB);
        for (int r = 0; r < 3; r++)
            for (int c = 0; c < 3; c++) {
                R k = (ima + imb) * (r == c ? R(1) : R(0));
                k -= SA[3 * r + c];
                k -= SB[3 * r + c];
                JROW(JR_K + 3 * r + c, j) = k;
            }
        for (int r = 0; r < 3; r++) {
            const double anc = d.janc[(size_t)r * d.jcap + j];
            const double pa = (a < 0 ? anc : BODY(d.pos, r, a)) + rAd[r];
            const double pb = (b < 0 ? anc : BODY(d.pos, r, b)) + rBd[r];
            JROW(JR_RA + r, j) = rA[r]; JROW(JR_RB + r, j) = rB[r];
            JROW(JR_BIAS + r, j) = (R)(0.2 * (pb - pa));
        }
    }
    int32_t* last = d.last + off;
    for (int32_t i = tid; i < n; i += BLOCK) last[i] = 0;
    __syncthreads();
    auto get = [&](int32_t t, int32_t& al, int32_t& bl) {
        const int32_t a = d.ja[js + t], b = d.jb[js + t];
        al = (a >= 0 && d.mt[a] == M_DYNAMIC) ? (int32_t)(a - off) : -1;
        bl = (b >= 0 && d.mt[b] == M_DYNAMIC) ? (int32_t)(b - off) : -1;
    };
    const int32_t D = assign_levels<BLOCK>(J, get, last, d.jlevel + js, d.jcount + js, d.jorder + js, s_scan, &s_D);
    const int32_t iters = d.w_jiters[w];
    const R rdt = (R)dt;
    for (int32_t it = 0; it < iters; it++) {
        for (int32_t l = 0; l < D; l++) {
            const int32_t e = d.jcount[js + l], bg = l == 0 ? 0 : d.jcount[js + l - 1];
            for (int32_t t = bg + tid; t < e; t += BLOCK) {
                const int64_t j = js + d.jorder[js + t];
                const int32_t a = d.ja[j], b = d.jb[j];
                R va[6], vb[6], rA[3], rB[3], Kj[9], rhs[3], P[3];
                for (int r = 0; r < 3; r++) { rA[r] = JROW(JR_RA + r, j); rB[r] = JROW(JR_RB + r, j); }
                for (int r = 0; r < 9; r++) Kj[r] = JROW(JR_K + r, j);
                for (int c = 0; c < 6; c++) { va[c] = a >= 0 ? (R)BODY(d.lv, c, a) : R(0); vb[c] = b >= 0 ? (R)BODY(d.lv, c, b) : R(0); }
                R pva[3] = {R(0), R(0), R(0)}, pvb[3] = {R(0), R(0), R(0)};
                if (a >= 0) {
                    pva[0] = va[0] + (va[4] * rA[2] - va[5] * rA[1]);
                    pva[1] = va[1] + (va[5] * rA[0] - va[3] * rA[2]);
                    pva[2] = va[2] + (va[3] * rA[1] - va[4] * rA[0]);
                }
                if (b >= 0) {
                    pvb[0] = vb[0] + (vb[4] * rB[2] - vb[5] * rB[1]);
                    pvb[1] = vb[1] + (vb[5] * rB[0] - vb[3] * rB[2]);
                    pvb[2] = vb[2] + (vb[3] * rB[1] - vb[4] * rB[0]);
                }
                for (int r = 0; r < 3; r++) rhs[r] = -((pvb[r] - pva[r]) + JROW(JR_BIAS + r, j) / rdt);
                solve3(Kj, rhs, P);
                for (int s2 = 0; s2 < 2; s2++) {
                    const int32_t i = s2 ? b : a;
                    if (i < 0 || d.mt[i] != M_DYNAMIC) continue;
                    R* v = s2 ? vb : va;
                    const R* rr = s2 ? rB : rA;
                    const R Q[3] = {s2 ? P[0] : -P[0], s2 ? P[1] : -P[1], s2 ? P[2] : -P[2]};
                    const R im = (R)d.inv_mass[i];
                    R nv[6];
                    for (int r = 0; r < 3; r++) nv[r] = v[r] + im * Q[r];
                    const R cx = rr[1] * Q[2] - rr[2] * Q[1], cy = rr[2] * Q[0] - rr[0] * Q[2], cz = rr[0] * Q[1] - rr[1] * Q[0];
                    for (int r = 0; r < 3; r++)
                        nv[3 + r] = v[3 + r] + ((BODY(iiw, 3 * r, i) * cx + BODY(iiw, 3 * r + 1, i) * cy) + BODY(iiw, 3 * r + 2, i) * cz);
                    for (int c = 0; c < 6; c++) {
                        if (fp64) BODY(d.lv, c, i) = (double)nv[c];
                        else BODY(d.lv, c, i) = BODY(d.lv, c, i) + (double)(nv[c] - v[c]);
                    }
                }
            }
            __syncthreads();
        }
    }
#undef JROW
}

// ------------------------------------------------------- warm-start cache --

__global__ void k_hash_clear(Dev d) {
    const int32_t nw = *d.sel ^ 1;
    for (int64_t i = gtid(); i < d.hcap; i += gstride()) d.hkey[nw][i] = ~0ull;
}

// The cache after this step (solver.py:271-278): this step's contacts with
// their impulses, or — when the world had no contacts — last step's cache
// unchanged (solver.py:197-198).  Group heads go into the hash.
template <typename R, int BLOCK>
__global__ __launch_bounds__(BLOCK) void k_cache(Dev d) {
    __shared__ int32_t s_base[2];
    const int32_t w = blockIdx.x, tid = threadIdx.x;
    if (w >= d.wcap) return;
    const int32_t old = *d.sel, nw = old ^ 1;
    const size_t os = (size_t)old * d.wcap + w, ns = (size_t)nw * d.wcap + w;
    if (!(d.w_flags[w] & WF_ACTIVE)) { if (tid == 0) d.w_cache_n[ns] = 0; return; }
    const int32_t K = d.w_cn[w];
    const int32_t cnt = K > 0 ? K : d.w_cache_n[os];
    if (cnt == 0) { if (tid == 0) d.w_cache_n[ns] = 0; return; }
    if (tid == 0) {
        uint64_t base = atomicAdd((unsigned long long*)&d.cnt[C_CACHE], (unsigned long long)cnt);
        int32_t ok = base + cnt <= (uint64_t)d.kcap;
        if (!ok) atomicOr((unsigned long long*)&d.cnt[C_OVERFLOW], 4ull);
        s_base[0] = (int32_t)base; s_base[1] = ok;
        d.w_cache_s[ns] = (int32_t)base; d.w_cache_n[ns] = ok ? cnt : 0;
    }
    __syncthreads();
    if (!s_base[1]) return;
    const int32_t base = s_base[0];
    const int32_t src = K > 0 ? d.w_cs[w] : d.w_cache_s[os];
    const int32_t* sa = K > 0 ? d.ca : d.ka[old];
    const int32_t* sb = K > 0 ? d.cb : d.kb[old];
    const double* sp = K > 0 ? d.cpt : d.kpt[old];
    const R* si = (const R*)(K > 0 ? d.cimp : d.kimp[old]);
    R* di = (R*)d.kimp[nw];
    for (int32_t k = tid; k < cnt; k += BLOCK) {
        const int64_t s = src + k, t = base + k;
        const int32_t a = sa[s], b = sb[s];
        d.ka[nw][t] = a; d.kb[nw][t] = b;
        for (int r = 0; r < 3; r++) { CON(d.kpt[nw], r, t) = CON(sp, r, s); CON(di, r, t) = CON(si, r, s); }
        if (k == 0 || sa[s - 1] != a || sb[s - 1] != b) {
            const uint64_t key = ((uint64_t)(uint32_t)a << 32) | (uint32_t)b;
            const int64_t hm = d.hcap - 1;
            int64_t h = (int64_t)(hash_mix(key) & (uint64_t)hm);
            for (int64_t probe = 0; probe < d.hcap; probe++) {
                unsigned long long prev = atomicCAS((unsigned long long*)&d.hkey[nw][h], ~0ull, (unsigned long long)key);
                if (prev == ~0ull || prev == key) { d.hval[nw][h] = (int32_t)t; break; }
                h = (h + 1) & hm;
            }
        }
    }
}

__global__ void k_toggle(Dev d) {
    if (gtid() == 0) *d.sel ^= 1;
}

// ------------------------------------------------------------- integrate --

// NumpyBackend.integrate_positions (backend.py:42-49) with quat_integrate
// (mathutil.py:108-113) for awake dynamic and kinematic bodies.
__global__ void k_post(Dev d, int64_t nhw, double dt) {
    for (int64_t i = gtid(); i < nhw; i += gstride()) {
        if (d.world[i] < 0) continue;
        const int mt = d.mt[i];
        if (!((mt == M_DYNAMIC || mt == M_KINEMATIC) && d.awake[i])) continue;
        for (int r = 0; r < 3; r++) BODY(d.pos, r, i) = BODY(d.pos, r, i) + BODY(d.lv, r, i) * dt;
        const double wx = BODY(d.av, 0, i), wy = BODY(d.av, 1, i), wz = BODY(d.av, 2, i);
        const double bx = BODY(d.quat, 0, i), by = BODY(d.quat, 1, i), bz = BODY(d.quat, 2, i),
                     bw = BODY(d.quat, 3, i);
        const double mx = ((0.0 * bx + wx * bw) + wy * bz) - wz * by;
        const double my = ((0.0 * by - wx * bz) + wy * bw) + wz * bx;
        const double mz = ((0.0 * bz + wx * by) - wy * bx) + wz * bw;
        const double mw = ((0.0 * bw - wx * bx) - wy * by) - wz * bz;
        const double s0 = bx + (0.5 * mx) * dt, s1 = by + (0.5 * my) * dt;
        const double s2 = bz + (0.5 * mz) * dt, s3 = bw + (0.5 * mw) * dt;
        const double nn = sqrt(((s0 * s0 + s1 * s1) + s2 * s2) + s3 * s3);
        const double dd = maxd(nn, 1e-12);
        BODY(d.quat, 0, i) = s0 / dd; BODY(d.quat, 1, i) = s1 / dd;
        BODY(d.quat, 2, i) = s2 / dd; BODY(d.quat, 3, i) = s3 / dd;
    }
}

// ----------------------------------------------------------------- sleep --

__device__ __forceinline__ int32_t uf_find(int32_t* par, int32_t x) {
    int32_t p = __atomic_load_n(&par[x], __ATOMIC_RELAXED);
    while (p != x) {
        x = p;
        p = __atomic_load_n(&par[x], __ATOMIC_RELAXED);
    }
    return x;
}

// _contact_islands (world.py:540-567): union of contacts whose two ends are
// both dynamic; every dynamic end counts as contacted.
__global__ void k_sleep_union(Dev d) {
    const uint64_t tot = d.cnt[C_CONTACTS];
    const int64_t K = tot < (uint64_t)d.kcap ? (int64_t)tot : d.kcap;
    for (int64_t k = gtid(); k < K; k += gstride()) {
        const int32_t a = d.ca[k], b = d.cb[k];
        const int32_t w = d.world[a];
        if (w < 0 || !(d.w_flags[w] & WF_SLEEP)) continue;
        const bool da = d.mt[a] == M_DYNAMIC, db = d.mt[b] == M_DYNAMIC;
        if (da) d.contacted[a] = 1;
        if (db) d.contacted[b] = 1;
        if (!(da && db)) continue;
        int32_t x = a, y = b;
        for (;;) {
            x = uf_find(d.parent, x);
            y = uf_find(d.parent, y);
            if (x == y) break;
            if (x < y) { int32_t t = x; x = y; y = t; }
            if (atomicCAS(&d.parent[x], x, y) == x) break;
        }
    }
}

// _update_sleep (world.py:510-538), first half: speed tests, timers, and an
// island is marked not-ready if any member is not ready.
__global__ void k_sleep_body(Dev d, int64_t nhw, double dt) {
    for (int64_t i = gtid(); i < nhw; i += gstride()) {
        const int32_t w = d.world[i];
        if (w < 0 || !(d.w_flags[w] & WF_SLEEP)) continue;
        const bool dyn = d.mt[i] == M_DYNAMIC;
        const double x = BODY(d.lv, 0, i), y = BODY(d.lv, 1, i), z = BODY(d.lv, 2, i);
        const double p = BODY(d.av, 0, i), q = BODY(d.av, 1, i), r = BODY(d.av, 2, i);
        const double s2 = (x * x + y * y) + z * z, w2 = (p * p + q * q) + r * r;
        const bool slow = dyn && s2 < 0.08 * 0.08 && w2 < 0.25 * 0.25;
        double tm = slow ? d.timer[i] + dt : 0.0;
        d.timer[i] = tm;
        const bool ready = slow && tm > 0.4;
        d.awake0[i] = (uint8_t)((slow ? 1 : 0) | (ready ? 2 : 0));   // reuse as per-body flags
        if (d.contacted[i] && !ready) d.notready[uf_find(d.parent, (int32_t)i)] = 1;
    }
}

__global__ void k_sleep_final(Dev d, int64_t nhw) {
    for (int64_t i = gtid(); i < nhw; i += gstride()) {
        const int32_t w = d.world[i];
        if (w < 0 || !(d.w_flags[w] & WF_SLEEP)) continue;
        const bool dyn = d.mt[i] == M_DYNAMIC;
        const uint8_t f = d.awake0[i];
        const bool slow = f & 1, ready = f & 2;
        const bool to_sleep = d.contacted[i] ? !d.notready[uf_find(d.parent, (int32_t)i)] : (dyn && ready);
        uint8_t aw = d.awake[i];
        if (to_sleep) aw = 0;
        if (dyn && !slow) aw = 1;
        d.awake[i] = aw;
        if (!aw && dyn)
            for (int c = 0; c < 6; c++) BODY(d.lv, c, i) = 0.0;
    }
}

// ------------------------------------------------------ stage-level kernels --

template <typename R>
__global__ void k_stage_boxbox(int32_t np, const double* ca, const double* Ra, const double* ha, const double* cb,
                               const double* Rb, const double* hb, int32_t* cnt, double* nrm, double* pts,
                               double* dep) {
    for (int64_t i = gtid(); i < np; i += gstride()) {
        R A[9], B[9], h1[3], h2[3], c1[3], c2[3], n[3] = {R(0), R(0), R(0)}, p[12], dd[4];
        for (int r = 0; r < 9; r++) { A[r] = (R)Ra[9 * i + r]; B[r] = (R)Rb[9 * i + r]; }
        double org[3];
        for (int r = 0; r < 3; r++) {
            h1[r] = (R)ha[3 * i + r]; h2[r] = (R)hb[3 * i + r];
            if (sizeof(R) == sizeof(double)) { c1[r] = (R)ca[3 * i + r]; c2[r] = (R)cb[3 * i + r]; org[r] = 0.0; }
            else { c1[r] = R(0); c2[r] = (R)(cb[3 * i + r] - ca[3 * i + r]); org[r] = ca[3 * i + r]; }
        }
        int c = box_box(c1, A, h1, c2, B, h2, n, p, dd);
        cnt[i] = c;
        for (int r = 0; r < 3; r++) nrm[3 * i + r] = n[r];
        for (int m = 0; m < 4; m++) {
            for (int r = 0; r < 3; r++) pts[12 * i + 3 * m + r] = m < c ? org[r] + (double)p[3 * m + r] : 0.0;
            dep[4 * i + m] = m < c ? (double)dd[m] : 0.0;
        }
    }
}

template <typename R>
__global__ void k_stage_spheres(int32_t np, int32_t mode, const double* ca, const double* ra, const double* cb,
                                const double* Rb, const double* hb, const double* rb, int32_t* cnt, double* nrm,
                                double* pt, double* dep) {
    for (int64_t i = gtid(); i < np; i += gstride()) {
        R c1[3], c2[3], n[3] = {R(0), R(0), R(0)}, p[3] = {R(0), R(0), R(0)}, dd = R(0);
        double org[3];
        for (int r = 0; r < 3; r++) {
            if (sizeof(R) == sizeof(double)) { c1[r] = (R)ca[3 * i + r]; c2[r] = (R)cb[3 * i + r]; org[r] = 0.0; }
            else { c1[r] = R(0); c2[r] = (R)(cb[3 * i + r] - ca[3 * i + r]); org[r] = ca[3 * i + r]; }
        }
        int c;
        if (mode == 0) {
            c = sphere_sphere(c1, (R)ra[i], c2, (R)rb[i], n, p, &dd);
        } else {
            R M[9], h[3];
            for (int r = 0; r < 9; r++) M[r] = (R)Rb[9 * i + r];
            for (int r = 0; r < 3; r++) h[r] = (R)hb[3 * i + r];
            c = sphere_box(c1, (R)ra[i], c2, M, h, n, p, &dd);
        }
        cnt[i] = c;
        for (int r = 0; r < 3; r++) { nrm[3 * i + r] = n[r]; pt[3 * i + r] = c ? org[r] + (double)p[r] : 0.0; }
        dep[i] = c ? (double)dd : 0.0;
    }
}

// ----------------------------------------------------------- launchers --

#define GRID_FOR(n, b) ((int)(((n) + (b)-1) / (b) < 8192 ? ((n) + (b)-1) / (b) : 8192))

struct Launch {
    static void begin(const Dev& d, hipStream_t s) {
        int g = GRID_FOR(d.wcap > C_COUNT ? d.wcap : C_COUNT, 256);
        hipLaunchKernelGGL(k_begin, dim3(g), dim3(256), 0, s, d);
    }
};

}  // namespace omi

// host-side entry points into the kernels, used by capi.hip
namespace omi {
void launch_begin(const Dev& d, hipStream_t s) { Launch::begin(d, s); }

void launch_pre(const Dev& d, int precision, int64_t nhw, double dt, double margin, hipStream_t s) {
    if (nhw <= 0) return;
    int g = GRID_FOR(nhw, 256);
    if (precision == 64) hipLaunchKernelGGL(k_pre<double>, dim3(g), dim3(256), 0, s, d, nhw, dt, margin);
    else hipLaunchKernelGGL(k_pre<float>, dim3(g), dim3(256), 0, s, d, nhw, dt, margin);
}

void launch_refit(const Dev& d, int64_t nhw, double margin, hipStream_t s) {
    if (nhw <= 0) return;
    hipLaunchKernelGGL(k_refit, dim3(GRID_FOR(nhw, 256)), dim3(256), 0, s, d, nhw, margin);
}

// Per-world launches go by size class: c = 0..2 hold worlds whose padded size
// fits np_class[c] bodies in LDS, c = 3 the larger ones (global-memory
// scratch).  Each class's world list and count live in device memory, so a
// captured graph stays valid as worlds come and go.
template <int B>
static void sap_sort_one(const Dev& d, const int32_t* wl, const int32_t* wn, int32_t np, bool global, hipStream_t s) {
    if (global) hipLaunchKernelGGL((k_sap_sort<B, true>), dim3(d.wcap), dim3(B), 0, s, d, wl, wn, 0);
    else hipLaunchKernelGGL((k_sap_sort<B, false>), dim3(d.wcap), dim3(B), (size_t)np * 16, s, d, wl, wn, np);
}

void launch_sap(const Dev& d, const int32_t* wl, const int32_t* wn, const int32_t* np_class, const int32_t* blocks,
                const int32_t* live, int32_t record, int32_t grid, hipStream_t s) {
    for (int c = 0; c < 4; c++) {
        if (!live[c]) continue;
        const bool global = c == 3;
        const int32_t* l = wl + (size_t)c * d.wcap;
        const int32_t* n = wn + c;
        switch (blocks[c]) {
            case 64: sap_sort_one<64>(d, l, n, np_class[c], global, s); break;
            case 128: sap_sort_one<128>(d, l, n, np_class[c], global, s); break;
            case 512: sap_sort_one<512>(d, l, n, np_class[c], global, s); break;
            case 1024: sap_sort_one<1024>(d, l, n, np_class[c], global, s); break;
            default: sap_sort_one<256>(d, l, n, np_class[c], global, s); break;
        }
    }
    hipLaunchKernelGGL((k_sap_pairs<256, false>), dim3(grid), dim3(256), 0, s, d, record);
    hipLaunchKernelGGL((k_sap_scan<1024>), dim3(1), dim3(1024), 0, s, d, record);
    hipLaunchKernelGGL((k_sap_pairs<256, true>), dim3(grid), dim3(256), 0, s, d, record);
}

void launch_narrow(const Dev& d, int precision, hipStream_t s) {
    int g = GRID_FOR(d.pcap, 256);
    if (precision == 64) hipLaunchKernelGGL(k_narrow<double>, dim3(g), dim3(256), 0, s, d);
    else hipLaunchKernelGGL(k_narrow<float>, dim3(g), dim3(256), 0, s, d);
    if (d.epa_slots > 0) {
        int ge = (d.epa_slots + 63) / 64;
        if (precision == 64) hipLaunchKernelGGL(k_narrow_convex<double>, dim3(ge), dim3(64), 0, s, d);
        else hipLaunchKernelGGL(k_narrow_convex<float>, dim3(ge), dim3(64), 0, s, d);
    }
}

size_t epa_work_bytes() { return sizeof(EpaWork); }

void launch_stage_convex(int32_t np, const int32_t* ka, const double* ca, const double* Ra, const double* ha,
                         const int32_t* oa, const int32_t* na, const int32_t* kb, const double* cb, const double* Rb,
                         const double* hb, const int32_t* ob, const int32_t* nb, const double* hull, void* work,
                         int32_t slots, int32_t* cnt, double* nrm, double* pt, double* dep, hipStream_t s) {
    hipLaunchKernelGGL(k_stage_convex, dim3((slots + 63) / 64), dim3(64), 0, s, np, ka, ca, Ra, ha, oa, na, kb, cb,
                       Rb, hb, ob, nb, hull, (EpaWork*)work, slots, cnt, nrm, pt, dep);
}

void launch_compact(const Dev& d, int precision, int32_t record, hipStream_t s) {
    if (precision == 64) hipLaunchKernelGGL((k_compact<double, 256>), dim3(d.wcap), dim3(256), 0, s, d, record);
    else hipLaunchKernelGGL((k_compact<float, 256>), dim3(d.wcap), dim3(256), 0, s, d, record);
}

void launch_wake(const Dev& d, hipStream_t s) {
    hipLaunchKernelGGL((k_wake<256>), dim3(d.wcap), dim3(256), 0, s, d);
}

template <typename R, int B>
static void solve_one(const Dev& d, SolveIn in, const int32_t* wl, const int32_t* wn, int32_t lds_last,
                      int32_t lds_state, bool stage, hipStream_t s) {
    size_t lds = (((size_t)lds_last * 4 + 15) & ~(size_t)15) + (size_t)lds_state * 15 * sizeof(R);
    if (stage)
        hipLaunchKernelGGL((k_solve<R, B, true>), dim3(d.wcap), dim3(B), lds, s, d, in, wl, wn, lds_last, lds_state);
    else
        hipLaunchKernelGGL((k_solve<R, B, false>), dim3(d.wcap), dim3(B), lds, s, d, in, wl, wn, lds_last, lds_state);
}

template <typename R>
static void solve_block(const Dev& d, int block, SolveIn in, const int32_t* wl, const int32_t* wn,
                        int32_t lds_last, int32_t lds_state, bool stage, hipStream_t s) {
    switch (block) {
        case 64: solve_one<R, 64>(d, in, wl, wn, lds_last, lds_state, stage, s); break;
        case 128: solve_one<R, 128>(d, in, wl, wn, lds_last, lds_state, stage, s); break;
        case 512: solve_one<R, 512>(d, in, wl, wn, lds_last, lds_state, stage, s); break;
        case 1024: solve_one<R, 1024>(d, in, wl, wn, lds_last, lds_state, stage, s); break;
        default: solve_one<R, 256>(d, in, wl, wn, lds_last, lds_state, stage, s); break;
    }
}

// For size class c, lds_last[c] bodies keep their level counters in LDS and
// lds_state[c] bodies their velocities and inverse inertia; larger worlds of
// the class use global memory for them.
void launch_solve(const Dev& d, int precision, const int32_t* blocks, SolveIn in, const int32_t* wl,
                  const int32_t* wn, const int32_t* lds_last, const int32_t* lds_state, const int32_t* live,
                  bool stage, hipStream_t s) {
    for (int c = 0; c < 4; c++) {
        if (!live[c]) continue;
        const int32_t* l = wl + (size_t)c * d.wcap;
        if (precision == 64) solve_block<double>(d, blocks[c], in, l, wn + c, lds_last[c], lds_state[c], stage, s);
        else solve_block<float>(d, blocks[c], in, l, wn + c, lds_last[c], lds_state[c], stage, s);
    }
}

void launch_cache(const Dev& d, int precision, hipStream_t s) {
    hipLaunchKernelGGL(k_hash_clear, dim3(GRID_FOR(d.hcap, 256)), dim3(256), 0, s, d);
    if (precision == 64) hipLaunchKernelGGL((k_cache<double, 256>), dim3(d.wcap), dim3(256), 0, s, d);
    else hipLaunchKernelGGL((k_cache<float, 256>), dim3(d.wcap), dim3(256), 0, s, d);
}

void launch_joints(const Dev& d, int precision, double dt, hipStream_t s) {
    if (precision == 64) hipLaunchKernelGGL((k_joints<double, 256>), dim3(d.wcap), dim3(256), 0, s, d, dt);
    else hipLaunchKernelGGL((k_joints<float, 256>), dim3(d.wcap), dim3(256), 0, s, d, dt);
}

void launch_post(const Dev& d, int64_t nhw, double dt, hipStream_t s) {
    if (nhw <= 0) return;
    hipLaunchKernelGGL(k_post, dim3(GRID_FOR(nhw, 256)), dim3(256), 0, s, d, nhw, dt);
}

void launch_sleep(const Dev& d, int64_t nhw, double dt, hipStream_t s) {
    if (nhw <= 0) return;
    hipLaunchKernelGGL(k_sleep_union, dim3(GRID_FOR(d.kcap, 256)), dim3(256), 0, s, d);
    hipLaunchKernelGGL(k_sleep_body, dim3(GRID_FOR(nhw, 256)), dim3(256), 0, s, d, nhw, dt);
    hipLaunchKernelGGL(k_sleep_final, dim3(GRID_FOR(nhw, 256)), dim3(256), 0, s, d, nhw);
}

void launch_toggle(const Dev& d, hipStream_t s) { hipLaunchKernelGGL(k_toggle, dim3(1), dim3(64), 0, s, d); }

void launch_stage_boxbox(int precision, int32_t np, const double* ca, const double* Ra, const double* ha,
                         const double* cb, const double* Rb, const double* hb, int32_t* cnt, double* nrm,
                         double* pts, double* dep, hipStream_t s) {
    int g = GRID_FOR(np, 256);
    if (precision == 64)
        hipLaunchKernelGGL(k_stage_boxbox<double>, dim3(g), dim3(256), 0, s, np, ca, Ra, ha, cb, Rb, hb, cnt, nrm, pts, dep);
    else
        hipLaunchKernelGGL(k_stage_boxbox<float>, dim3(g), dim3(256), 0, s, np, ca, Ra, ha, cb, Rb, hb, cnt, nrm, pts, dep);
}

void launch_stage_spheres(int precision, int32_t np, int32_t mode, const double* ca, const double* ra,
                          const double* cb, const double* Rb, const double* hb, const double* rb, int32_t* cnt,
                          double* nrm, double* pt, double* dep, hipStream_t s) {
    int g = GRID_FOR(np, 256);
    if (precision == 64)
        hipLaunchKernelGGL(k_stage_spheres<double>, dim3(g), dim3(256), 0, s, np, mode, ca, ra, cb, Rb, hb, rb, cnt, nrm, pt, dep);
    else
        hipLaunchKernelGGL(k_stage_spheres<float>, dim3(g), dim3(256), 0, s, np, mode, ca, ra, cb, Rb, hb, rb, cnt, nrm, pt, dep);
}
}  // namespace omi
