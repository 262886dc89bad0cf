// k_step.hip — the per-step kernels of the contact pipeline on gfx950.
//
// Stage map (reference -> kernel):
//   world.step prev-pose copy, integrate_forces, refit  -> k_pre        (1 thread / body)
//   BroadPhase.pairs + solid filter                     -> k_sap        (1 block / world)
//   NarrowPhase.generate (box/sphere)                   -> k_narrow     (1 thread / pair)
//   contact list in pair-then-manifold order            -> k_pair_*     (scan over all pairs)
//   SequentialImpulseSolver._wake_sleepers              -> with the rows (wake_pair / wake_seq)
//   _solve_native_full + prepare_and_solve              -> k_solve      (1 block / world)
//   warm-start cache rebuild                            -> k_cache_all  (1 thread / entry)
//   integrate_positions                                 -> k_post       (1 thread / body)
//   _update_sleep + _contact_islands                    -> k_sleep_*    (contacts, bodies)
#include <type_traits>

#include "omi_math.hpp"
#include "omi_batch.hpp"
#include "omi_gjk.hpp"
#include "omi_capsule.hpp"

namespace omi {

#define BODY(a, c, i) (a)[(size_t)(c) * d.ncap + (size_t)(i)]
#define CON(a, c, k) (a)[(size_t)(c) * d.kcap + (size_t)(k)]
// world inverse inertia of body i, row-major 3x3 in a 12-slot record
#define IIW(a, i, c) (a)[(size_t)(i) * IIW_STRIDE + (size_t)(c)]
// field f of contact ci's solver row (one record of RowLayout<R>::stride slots)
#define ROW(a, ci, f) (a)[(size_t)(ci) * RowLayout<R>::stride + (size_t)(f)]

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
__device__ __forceinline__ int32_t block_exscan(int32_t* a, int32_t n, int32_t* scratch) {
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

// NumpyBackend.integrate_forces (backend.py:26-40) for one awake dynamic body:
// gravity, linear damping, quadratic drag on the damped speed, angular damping,
// in the reference's operation order.
__device__ __forceinline__ void forces_body(double* v, double* w, const double* g, double gf, double ld, double ad,
                                            double qd, double dlin, double dang, double dt) {
    for (int r = 0; r < 3; r++) v[r] = v[r] + (g[r] * gf) * dt;
    const double fl = maxd(1.0 - (ld + dlin) * dt, 0.0);
    for (int r = 0; r < 3; r++) v[r] = v[r] * fl;
    const double sp = sqrt((v[0] * v[0] + v[1] * v[1]) + v[2] * v[2]);
    for (int r = 0; r < 3; r++) v[r] = v[r] - ((qd * sp) * v[r]) * dt;
    const double fa = maxd(1.0 - (ad + dang) * dt, 0.0);
    for (int r = 0; r < 3; r++) w[r] = w[r] * fa;
}

// NumpyBackend.integrate_positions (backend.py:42-49) with quat_integrate
// (mathutil.py:108-113) for one moving body: x += v dt, then
// q <- normalize(q + dt/2 (w, 0) q) with quat_mul's term order.
__device__ __forceinline__ void positions_body(double* p, double* q, const double* v, const double* w, double dt) {
    for (int r = 0; r < 3; r++) p[r] = p[r] + v[r] * dt;
    const double wx = w[0], wy = w[1], wz = w[2];
    const double bx = q[0], by = q[1], bz = q[2], bw = q[3];
    const double mx = ((0.0 * bx + wx * bw) + wy * bz) - wz * by;
    const double my = ((0.0 * by - wx * bz) + wy * bw) + wz * bx;
    const double mz = ((0.0 * bz + wx * by) - wy * bx) + wz * bw;
    const double mw = ((0.0 * bw - wx * bx) - wy * by) - wz * bz;
    const double s0 = bx + (0.5 * mx) * dt, s1 = by + (0.5 * my) * dt;
    const double s2 = bz + (0.5 * mz) * dt, s3 = bw + (0.5 * mw) * dt;
    const double nn = sqrt(((s0 * s0 + s1 * s1) + s2 * s2) + s3 * s3);
    const double dd = maxd(nn, 1e-12);
    q[0] = s0 / dd; q[1] = s1 / dd; q[2] = s2 / dd; q[3] = s3 / dd;
}

// k_pre modes: the whole step's first kernel, or the first kernel of a
// collision stage alone (the host has integrated forces and refitted already).
enum : int { PRE_STEP = 0, PRE_COLLIDE = 1 };

// Start of step for every live body: previous pose, sleep scratch reset,
// integrate_forces for awake dynamic bodies, then the collision-side per-body
// data: AABB, rotation, world inverse inertia (solver.py:215-220, einsum order
// (p0 + p2) + p1).  PRE_COLLIDE keeps pose, velocity and AABB as they are and
// only resets the scratch and writes rotation and world inverse inertia.
template <typename R, int MODE>
__global__ void k_pre(Dev d, int64_t nhw, double dt, double margin) {
    for (int64_t i = gtid(); i < nhw; i += gstride()) {
        const int32_t w = d.world[i];
        if (w < 0) continue;
        const int32_t wf = d.w_flags[w];
        double q[4];
        for (int c = 0; c < 4; c++) q[c] = BODY(d.quat, c, i);
        if (MODE == PRE_STEP) {
            for (int c = 0; c < 3; c++) BODY(d.ppos, c, i) = BODY(d.pos, c, i);
            for (int c = 0; c < 4; c++) BODY(d.pquat, c, i) = q[c];
        }
        const int mt = d.mt[i];
        const uint8_t aw = d.awake[i];
        d.awake0[i] = aw;
        d.parent[i] = (int32_t)i;
        d.contacted[i] = 0;
        d.notready[i] = 0;
        double v[3] = {BODY(d.lv, 0, i), BODY(d.lv, 1, i), BODY(d.lv, 2, i)};
        if (mt == M_DYNAMIC && aw) {
            if (MODE == PRE_STEP) {
                double wv[3] = {BODY(d.av, 0, i), BODY(d.av, 1, i), BODY(d.av, 2, i)};
                const double g[3] = {d.w_grav[w], d.w_grav[(size_t)d.wcap + w], d.w_grav[2 * (size_t)d.wcap + w]};
                forces_body(v, wv, g, d.gf[i], d.ld[i], d.ad[i], d.qd[i], d.w_damp[w], d.w_damp[d.wcap + w], dt);
                for (int r = 0; r < 3; r++) { BODY(d.lv, r, i) = v[r]; BODY(d.av, r, i) = wv[r]; }
            }
        } else if (mt == M_DYNAMIC && (wf & WF_COLLISION)) {
            // A sleeping body with a wake-grade speed makes the wake order matter.
            double v2 = (v[0] * v[0] + v[1] * v[1]) + v[2] * v[2];
            if (v2 > 0.15 * 0.15) atomicOr(&d.w_flags[w], WF_SEQWAKE);
        }
        if (wf & (WF_COLLISION | WF_JOINTS)) {
            double Rm[9];
            quat_to_matrix(q, Rm);
            if (MODE == PRE_STEP && (wf & WF_COLLISION)) refit_body(d, i, Rm, margin);
            R* rot = (R*)d.rot;
            R* iiw = (R*)d.iiw;
            const double dg[3] = {BODY(d.inv_inertia, 0, i), BODY(d.inv_inertia, 1, i), BODY(d.inv_inertia, 2, i)};
            for (int r = 0; r < 9; r++) BODY(rot, r, i) = (R)Rm[r];
            for (int r = 0; r < 3; r++)
                for (int c = 0; c < 3; c++) {
                    double p0 = (Rm[3 * r] * dg[0]) * Rm[3 * c], p1 = (Rm[3 * r + 1] * dg[1]) * Rm[3 * c + 1];
                    double p2 = (Rm[3 * r + 2] * dg[2]) * Rm[3 * c + 2];
                    IIW(iiw, i, 3 * r + c) = (R)((p0 + p2) + p1);
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
    for (int32_t t = tid; t < n; t += BLOCK) { key[t] = order_key(BODY(d.lo, 0, off + t)); idx[t] = t; }
    __syncthreads();
    // Bitonic sort of (key, index) with every comparator ascending: each
    // merge of size k opens with a mirror stage (i against its mirror in the
    // block) and continues with half-cleaners.  Positions past n stand for
    // +inf keys, which no ascending comparator moves, so the comparators that
    // reach them are skipped and only n entries are held.
    auto cmpx = [&](int32_t i, int32_t l) {
        if (l >= n) return;
        const uint64_t ki = key[i], kl = key[l];
        const int32_t ii = idx[i], il = idx[l];
        if (ki > kl || (ki == kl && ii > il)) { key[i] = kl; key[l] = ki; idx[i] = il; idx[l] = ii; }
    };
    for (int32_t k = 2; k <= NP; k <<= 1) {
        const int32_t h = k >> 1;
        for (int32_t t = tid; t < (NP >> 1); t += BLOCK) {
            const int32_t b0 = (t / h) * k, o = t % h;
            cmpx(b0 + o, b0 + k - 1 - o);
        }
        __syncthreads();
        for (int32_t j = k >> 2; j > 0; j >>= 1) {
            for (int32_t t = tid; t < (NP >> 1); t += BLOCK) {
                const int32_t i = 2 * t - (t & (j - 1));
                cmpx(i, i + j);
            }
            __syncthreads();
        }
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
            bool bp = false, sol = false, trig = false;
            int64_t p = 0, q = 0;
            if (valid) {
                const int32_t j = k + 1 + (c - so[k]);
                const SapRec A = sap_rec(d, off + k), B = sap_rec(d, off + j);
                sap_test(d, A, B, fo, nf, bp, sol);
                trig = bp && ((A.info | B.info) & SI_TRIGGER);
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
                // pairs with a trigger body (the sensor pairs of the step; an
                // unordered list, they are few)
                if (trig) {
                    const uint64_t o = atomicAdd((unsigned long long*)&d.cnt[C_TPAIRS], 1ull);
                    if (o < (uint64_t)d.tpcap) { d.tpa[o] = (int32_t)p; d.tpb[o] = (int32_t)q; }
                }
            }
            cnt_s += __popcll(ms);
            cnt_b += __popcll(mb);
            kw = __builtin_amdgcn_readlane(k, 63);
        }
        if (!WRITE && lane == 0) { d.tcnt[t] = cnt_s; d.tcnt[d.tcap + 1 + t] = cnt_b; }
    }
}

// Exclusive scans of a0[0..T) and a1[0..T) in place by one block; the totals
// land in totals[0], totals[1] (LDS) for every thread.
template <int BLOCK>
__device__ void block_exscan2(int64_t* a0, int64_t* a1, int64_t T, int64_t* totals) {
    constexpr int NW = BLOCK / 64, ITEMS = 8;
    __shared__ int64_t s_w[2][NW];
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    int64_t* arr[2] = {a0, a1};
    int64_t* s_run = totals;
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
}

// Per-tile counts -> exclusive offsets (in place, one block), then every
// world's solid and broadphase pair ranges.
template <int BLOCK>
__global__ __launch_bounds__(BLOCK) void k_sap_scan(Dev d, int32_t record) {
    __shared__ int64_t s_run[2];
    const int tid = threadIdx.x;
    const uint64_t tt = d.cnt[C_TILES];
    const int64_t T = tt < (uint64_t)d.tcap ? (int64_t)tt : d.tcap;
    int64_t* arr[2] = {d.tcnt, d.tcnt + d.tcap + 1};
    block_exscan2<BLOCK>(arr[0], arr[1], T, s_run);
    if (tid == 0) {
        arr[0][T] = s_run[0];
        arr[1][T] = s_run[1];
        d.cnt[C_PAIRS] = (uint64_t)s_run[0];
        d.cnt[C_BPTOTAL] = (uint64_t)s_run[1];
        // pairs of the worlds whose lists fit: all of them, or those below
        // the world that crosses the capacity (set in the loop below)
        d.cnt[C_PVALID] = (uint64_t)(s_run[0] < d.pcap ? s_run[0] : d.pcap);
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
        // pair ranges are disjoint and ascending in tile order: at most one
        // world crosses the capacity
        if (live && ps < d.pcap && pe > d.pcap) d.cnt[C_PVALID] = (uint64_t)ps;
        d.w_ps[w] = (int32_t)(ps < d.pcap ? ps : d.pcap);
        d.w_pn[w] = pe <= d.pcap ? (int32_t)(pe - ps) : 0;
        d.w_bps[w] = (int32_t)(bs < d.pcap ? bs : d.pcap);
        d.w_bpn[w] = (record && be <= d.pcap) ? (int32_t)(be - bs) : 0;
    }
}


// ----------------------------------------------------------- narrowphase --

// NarrowPhase.generate dispatch (narrowphase.py:78-100) for one solid pair.
// FP64 runs in world coordinates exactly like the reference; FP32 first moves
// the pair into a frame at body a's centre (offset taken in FP64).  Without
// GENERAL a box pair whose clip outgrows 8 vertices returns -1.
template <typename R, bool GENERAL>
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
    if (ka == K_BOX && kb == K_BOX) cnt = box_box<R, GENERAL>(ca, Ra, ha, cb, Rb, hb, nrm, pts, dep);
    else if (ka == K_SPHERE && kb == K_SPHERE) cnt = sphere_sphere(ca, ha[0], cb, hb[0], nrm, pts, dep);
    else if (ka == K_SPHERE && kb == K_BOX) cnt = sphere_box(ca, ha[0], cb, Rb, hb, nrm, pts, dep);
    else if (ka == K_BOX && kb == K_SPHERE) {
        cnt = sphere_box(cb, hb[0], ca, Ra, ha, nrm, pts, dep);
        nrm[0] = -nrm[0]; nrm[1] = -nrm[1]; nrm[2] = -nrm[2];
    }
    return cnt;
}

// One pair's manifold into its slot (count, normal, points, depths).
template <typename R>
__device__ __forceinline__ void put_manifold(const Dev& d, int64_t p, int cnt, const R* nrm, const R* pts,
                                             const R* dep, const double* org) {
    R* mn = (R*)d.mnrm;
    R* md = (R*)d.mdep;
    d.mcnt[p] = (int8_t)cnt;
    for (int r = 0; r < 3; r++) mn[(size_t)r * d.pcap + p] = nrm[r];
#pragma unroll
    for (int m = 0; m < 4; m++) {
        if (m >= cnt) break;
        for (int r = 0; r < 3; r++) d.mpts[(size_t)r * 4 * d.pcap + 4 * p + m] = org[r] + (double)pts[3 * m + r];
        md[4 * p + m] = dep[m];
    }
}

// Compiled for at least two wavefronts per SIMD: the FP64 box-box path then
// fits the 256 vector registers and two wavefronts hide each other's latency.
// A block takes NARROW_BLOCK pairs at a time and hands them to its threads box
// pairs first, so the costly box-box path runs in as few wavefronts as the
// box pairs fill (a wavefront runs every path one of its pairs takes); each
// pair's manifold goes to its own slot, so the order changes nothing.  The
// kernel holds no private memory: a box pair whose clip polygon outgrows the 8
// register slots (near-degenerate input, rare) is queued from the top of the
// convex queue's array and clipped by k_narrow_clip16.
template <typename R>
__global__ __launch_bounds__(NARROW_BLOCK, 2) void k_narrow(Dev d) {
    constexpr int NW = NARROW_BLOCK / 64;
    __shared__ int32_t s_cnt[NW];
    __shared__ int32_t s_pair[NARROW_BLOCK];
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const uint64_t tot = d.cnt[C_PAIRS];
    const int64_t P = tot < (uint64_t)d.pcap ? (int64_t)tot : d.pcap;
    for (int64_t base = (int64_t)blockIdx.x * NARROW_BLOCK; base < P; base += (int64_t)gridDim.x * NARROW_BLOCK) {
        const int64_t p = base + tid;
        bool boxes = false, other = false;
        if (p < P) {
            const int64_t a = d.pa[p], b = d.pb[p];
            const int ka = d.kind[a], kb = d.kind[b];
            if (ka == K_CONVEX || kb == K_CONVEX) {
                // queued for the GJK/EPA kernel, which writes this pair's manifold
                d.mcnt[p] = 0;
                uint64_t slot = atomicAdd((unsigned long long*)&d.cnt[C_CONVEX], 1ull);
                d.cvx[slot] = (int32_t)p;
            } else if (ka == K_BOX && kb == K_BOX) {
                boxes = true;
            } else {
                other = true;
            }
        }
        const uint64_t mb = __ballot(boxes), mo = __ballot(other), lt = (1ull << lane) - 1ull;
        if (lane == 0) s_cnt[wv] = __popcll(mb) | (__popcll(mo) << 16);
        __syncthreads();
        int32_t nb = 0, nbb = 0, nob = 0, no = 0;
#pragma unroll
        for (int q = 0; q < NW; q++) {
            const int32_t c = s_cnt[q];
            nb += c & 0xffff;
            no += c >> 16;
            if (q < wv) { nbb += c & 0xffff; nob += c >> 16; }
        }
        if (boxes) s_pair[nbb + __popcll(mb & lt)] = (int32_t)p;
        if (other) s_pair[nb + nob + __popcll(mo & lt)] = (int32_t)p;
        __syncthreads();
        if (tid < nb + no) {
            const int64_t q = s_pair[tid];
            const int64_t a = d.pa[q], b = d.pb[q];
            R nrm[3] = {R(0), R(0), R(0)}, pts[12] = {}, dep[4] = {};
            double org[3];
            const int cnt = narrow_pair<R, false>(d, a, b, nrm, pts, dep, org);
            if (cnt < 0 || d.clip16_all) {
                // the convex queue fills cvx from 0 and this one from P - 1: a
                // pair is in at most one of them, so they never meet
                d.mcnt[q] = 0;
                const uint64_t k = atomicAdd((unsigned long long*)&d.cnt[C_CLIP16], 1ull);
                d.cvx[P - 1 - (int64_t)k] = (int32_t)q;
            } else {
                put_manifold<R>(d, q, cnt, nrm, pts, dep, org);
            }
        }
        __syncthreads();
    }
}

// The queued box pairs of k_narrow, clipped on 16-vertex polygons in private
// memory (_collide_native._clip in full).
template <typename R>
__global__ __launch_bounds__(64) void k_narrow_clip16(Dev d) {
    const uint64_t tot = d.cnt[C_PAIRS];
    const int64_t P = tot < (uint64_t)d.pcap ? (int64_t)tot : d.pcap;
    const uint64_t nq = d.cnt[C_CLIP16];
    const int64_t n = nq < (uint64_t)P ? (int64_t)nq : P;
    for (int64_t q = gtid(); q < n; q += gstride()) {
        const int64_t p = d.cvx[P - 1 - q];
        R nrm[3] = {R(0), R(0), R(0)}, pts[12] = {}, dep[4] = {};
        double org[3];
        const int cnt = narrow_pair<R, true>(d, d.pa[p], d.pb[p], nrm, pts, dep, org);
        put_manifold<R>(d, p, cnt, nrm, pts, dep, org);
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
// with its own EPA workspace.  With wp_lds bytes of dynamic LDS
// (2 * 3 * HULL_STAGE_PTS doubles per thread) each thread first stages its
// two hulls' world points there, element-major so that the lanes of the
// wavefront read consecutive words, and the support loops read LDS.
template <typename R>
__global__ __launch_bounds__(64) void k_narrow_convex(Dev d, int32_t wp_lds) {
    extern __shared__ __align__(16) unsigned char s_dyn[];
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
        if (wp_lds) {
            double* s = (double*)s_dyn + threadIdx.x;
            stage_world_pts(A, s, 64);
            stage_world_pts(B, s + 3 * HULL_STAGE_PTS * 64, 64);
        }
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

// gjk.collide_convex over independent pairs, one thread per workspace slot;
// pairs beyond `slots` take the slots in turn (a slot's workspace is rewritten
// by every pair).  A pair whose kinds are not sphere/box/convex or whose hull
// range leaves [0, nhull) gets count -1 and zero outputs, without reading its
// hull.
__global__ __launch_bounds__(64) void k_stage_convex(int32_t np, const int32_t* ka, const double* ca, const double* Ra, const double* ha,
                               const int32_t* oa, const int32_t* na, const int32_t* kb, const double* cb,
                               const double* Rb, const double* hb, const int32_t* ob, const int32_t* nb,
                               const double* hull, int64_t nhull, EpaWork* work, int32_t slots, int32_t* cnt,
                               double* nrm, double* pt, double* dep) {
    const int64_t slot = gtid();
    if (slot >= slots) return;
    auto valid = [&](int32_t k, const int32_t* off, const int32_t* num, int64_t i) {
        if (k != K_SPHERE && k != K_BOX && k != K_CONVEX) return false;
        return k != K_CONVEX || (off[i] >= 0 && num[i] >= 1 && (int64_t)off[i] + num[i] <= nhull);
    };
    for (int64_t i = slot; i < np; i += slots) {
        if (!valid(ka[i], oa, na, i) || !valid(kb[i], ob, nb, i)) {
            cnt[i] = -1;
            for (int r = 0; r < 3; r++) { nrm[3 * i + r] = 0.0; pt[3 * i + r] = 0.0; }
            dep[i] = 0.0;
            continue;
        }
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

// TriggerSystem.update's overlap test (triggers.py:24-70) over the step's
// broadphase pairs with a trigger body, after the contact solve (the poses
// _CollisionStages.run hands it, world.py:657-676).  _sensor_pair orders a
// pair (i < j) as (trigger, other): i when i carries a trigger shape, else j;
// an other body with neither collider nor trigger shape is skipped.
// _overlapping collides the two shapes at their poses (collide.collide:
// box-box, sphere-sphere, sphere-box either way round, GJK/EPA otherwise,
// always FP64) and the pair overlaps when that yields a contact.  A body's
// collider and trigger shapes are one shape here (the loader enforces it).
// Overlaps go to an unordered list (trigger body, other body); the enter /
// stay / exit diff against the previous step runs on the host.  One thread
// per pair, each with its own EPA workspace slot when the batch holds hulls.
__global__ __launch_bounds__(64) void k_trigger(Dev d, int32_t threads) {
    const uint64_t tot = d.cnt[C_TPAIRS];
    const int64_t T = tot < (uint64_t)d.tpcap ? (int64_t)tot : d.tpcap;
    const int64_t slot = gtid();
    if (slot == 0 && tot > (uint64_t)d.tpcap) atomicOr((unsigned long long*)&d.cnt[C_OVERFLOW], 32ull);
    if (slot >= threads) return;
    for (int64_t t = slot; t < T; t += threads) {
        const int64_t i = d.tpa[t], j = d.tpb[t];
        const bool ti = d.bflags[i] & BF_TRIGGER;
        const int64_t trg = ti ? i : j, oth = ti ? j : i;
        if (d.kind[oth] == K_NONE) continue;
        Proxy A, B;
        make_proxy(d, trg, A);
        make_proxy(d, oth, B);
        double nrm[3] = {0.0, 0.0, 0.0}, pts[12], dep[4];
        int cnt;
        if (A.kind == K_BOX && B.kind == K_BOX) cnt = box_box<double>(A.c, A.R, A.h, B.c, B.R, B.h, nrm, pts, dep);
        else if (A.kind == K_SPHERE && B.kind == K_SPHERE) cnt = sphere_sphere<double>(A.c, A.h[0], B.c, B.h[0], nrm, pts, dep);
        else if (A.kind == K_SPHERE && B.kind == K_BOX) cnt = sphere_box<double>(A.c, A.h[0], B.c, B.R, B.h, nrm, pts, dep);
        else if (A.kind == K_BOX && B.kind == K_SPHERE) cnt = sphere_box<double>(B.c, B.h[0], A.c, A.R, A.h, nrm, pts, dep);
        else if (d.epa_work) cnt = collide_convex(A, B, (EpaWork*)d.epa_work + slot, nrm, pts, dep[0]);
        else cnt = 0;
        if (cnt > 0) {
            const uint64_t o = atomicAdd((unsigned long long*)&d.cnt[C_TOVER], 1ull);
            if (o < (uint64_t)d.tpcap) { d.tva[o] = (int32_t)trg; d.tvb[o] = (int32_t)oth; }
        }
    }
}

// Contacts in the reference's order (pair order, then manifold order) packed
// into one contiguous slice per world, with one manifold record per pair
// that produced contacts: {first contact, contacts | a dynamic << 8 | b
// dynamic << 9, body a, body b}; first contact is a batch index, the bodies
// are world-local.  Pair lists are world-contiguous and ascending, so one
// exclusive scan of the contact counts over all valid pairs places every
// world's contacts: chunk sums (k_pair_count), a scan of the chunk sums
// (k_pair_scan), then every pair writes its contacts and record at its own
// prefix (k_pair_write) and every world takes its slice from the prefixes
// at its first and one-past-last pair (k_world_slices).  A world whose
// slice ends past the contact capacity gets no contacts (overflow flag 2),
// and so do all worlds after it.
__device__ __forceinline__ int64_t pair_chunks(const Dev& d, int64_t& P) {
    P = (int64_t)d.cnt[C_PVALID];
    return (P + CPT_CHUNK - 1) / CPT_CHUNK;
}

__global__ __launch_bounds__(CPT_BLOCK) void k_pair_count(Dev d) {
    __shared__ int64_t s_sum[CPT_BLOCK / 64 + 2];
    int64_t P;
    const int64_t nch = pair_chunks(d, P);
    for (int64_t ch = blockIdx.x; ch < nch; ch += gridDim.x) {
        int64_t c = 0, m = 0;
#pragma unroll
        for (int j = 0; j < CPT_ITEMS; j++) {
            const int64_t p = ch * CPT_CHUNK + j * CPT_BLOCK + threadIdx.x;
            const int v = p < P ? d.mcnt[p] : 0;
            c += v;
            m += v > 0;
        }
        // contacts in the high half, manifolds in the low (both < 2^31 per chunk)
        const int64_t t = block_sum<CPT_BLOCK>((c << 32) | m, s_sum);
        if (threadIdx.x == 0) { d.pchunk[ch] = t >> 32; d.pchunk[d.pchunk_cap + ch] = t & 0xffffffffll; }
    }
}

template <int BLOCK>
__global__ __launch_bounds__(BLOCK) void k_pair_scan(Dev d) {
    __shared__ int64_t s_run[2];
    int64_t P;
    const int64_t nch = pair_chunks(d, P);
    block_exscan2<BLOCK>(d.pchunk, d.pchunk + d.pchunk_cap, nch, s_run);
    if (threadIdx.x == 0) {
        const int64_t K = s_run[0], M = s_run[1];
        d.cnt[C_CONTACTS] = (uint64_t)K;
        d.cnt[C_MANIFOLDS] = (uint64_t)M;
        if (K > d.kcap) atomicOr((unsigned long long*)&d.cnt[C_OVERFLOW], 2ull);
        d.pfx[P] = (int32_t)(K < INT32_MAX ? K : INT32_MAX);
        d.pfx[d.pcap + 1 + P] = (int32_t)(M < INT32_MAX ? M : INT32_MAX);
    }
}

template <typename R>
__global__ __launch_bounds__(CPT_BLOCK) void k_pair_write(Dev d) {
    constexpr int NW = CPT_BLOCK / 64;
    __shared__ int32_t s_w[2][CPT_ITEMS][NW];
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    int64_t P;
    const int64_t nch = pair_chunks(d, P);
    const R* mn = (const R*)d.mnrm;
    const R* md = (const R*)d.mdep;
    R* cn = (R*)d.cnrm;
    R* cdp = (R*)d.cdep;
    R* imp = (R*)d.cimp;
    int32_t* pfx_m = d.pfx + d.pcap + 1;
    for (int64_t ch = blockIdx.x; ch < nch; ch += gridDim.x) {
        int32_t c[CPT_ITEMS], xc[CPT_ITEMS], xm[CPT_ITEMS];
#pragma unroll
        for (int j = 0; j < CPT_ITEMS; j++) {
            const int64_t p = ch * CPT_CHUNK + j * CPT_BLOCK + tid;
            c[j] = p < P ? d.mcnt[p] : 0;
            xc[j] = c[j];
            xm[j] = c[j] > 0;
        }
        // inclusive prefixes within each wavefront row
#pragma unroll
        for (int o = 1; o < 64; o <<= 1)
#pragma unroll
            for (int j = 0; j < CPT_ITEMS; j++) {
                const int32_t u = __shfl_up(xc[j], o, 64), v = __shfl_up(xm[j], o, 64);
                if (lane >= o) { xc[j] += u; xm[j] += v; }
            }
        if (lane == 63)
#pragma unroll
            for (int j = 0; j < CPT_ITEMS; j++) { s_w[0][j][wv] = xc[j]; s_w[1][j][wv] = xm[j]; }
        __syncthreads();
        // pair order within the chunk is row j, then thread: rows before j,
        // then the wavefronts before this one in row j
        int64_t bc = d.pchunk[ch], bm = d.pchunk[d.pchunk_cap + ch];
#pragma unroll
        for (int j = 0; j < CPT_ITEMS; j++) {
            int32_t rc = 0, rm = 0, wc = 0, wm = 0;
#pragma unroll
            for (int q = 0; q < NW; q++) {
                const int32_t sc = s_w[0][j][q], sm = s_w[1][j][q];
                rc += sc; rm += sm;
                if (q < wv) { wc += sc; wm += sm; }
            }
            const int64_t p = ch * CPT_CHUNK + j * CPT_BLOCK + tid;
            const int64_t k0 = bc + wc + xc[j] - c[j];          // first contact of the pair (batch index)
            const int64_t m0 = bm + wm + xm[j] - (c[j] > 0);    // its manifold record
            if (p < P) {
                d.pfx[p] = (int32_t)(k0 < INT32_MAX ? k0 : INT32_MAX);
                pfx_m[p] = (int32_t)(m0 < INT32_MAX ? m0 : INT32_MAX);
            }
            if (c[j] > 0 && k0 + c[j] <= d.kcap) {
                const int64_t a = d.pa[p], b = d.pb[p];
                const int64_t off = d.w_off[d.world[a]];
                const int32_t fl = (d.mt[a] == M_DYNAMIC ? MR_DYN_A : 0) | (d.mt[b] == M_DYNAMIC ? MR_DYN_B : 0);
                d.mrec[m0] = make_int4((int32_t)k0, c[j] | fl, (int32_t)(a - off), (int32_t)(b - off));
                for (int m = 0; m < c[j]; m++) {
                    const int64_t k = k0 + m;
                    d.ca[k] = (int32_t)a; d.cb[k] = (int32_t)b;
                    for (int r = 0; r < 3; r++) {
                        CON(d.cpt, r, k) = d.mpts[(size_t)r * 4 * d.pcap + 4 * (size_t)p + m];
                        CON(cn, r, k) = mn[(size_t)r * d.pcap + p];
                        CON(imp, r, k) = R(0);
                    }
                    cdp[k] = md[4 * (size_t)p + m];
                }
            }
            bc += rc; bm += rm;
        }
        __syncthreads();
    }
}

// Every world's contact slice and manifold records from the pair prefixes.
__global__ void k_world_slices(Dev d) {
    const int32_t* pfx_m = d.pfx + d.pcap + 1;
    for (int64_t w = gtid(); w < d.wcap; w += gstride()) {
        const int32_t pn = (d.w_flags[w] & WF_ACTIVE) ? d.w_pn[w] : 0, ps = d.w_ps[w];
        int32_t cs = 0, cn = 0, ms = 0, mn = 0;
        if (pn > 0) {
            cs = d.pfx[ps];
            const int32_t ce = d.pfx[ps + pn];
            ms = pfx_m[ps];
            if (ce <= d.kcap) { cn = ce - cs; mn = pfx_m[ps + pn] - ms; }
            else cs = 0;
        }
        d.w_cs[w] = cs; d.w_cn[w] = cn; d.w_ms[w] = ms; d.w_mn[w] = mn;
    }
}

// ------------------------------------------------------------------ wake --

// SequentialImpulseSolver._wake_sleepers (solver.py:321-337), run by the
// kernels that write a world's solver rows, for the same contacts (it reads
// the velocities before the solve changes them).  Waking is idempotent and,
// unless a sleeping dynamic body carries a wake-grade speed (flag WF_SEQWAKE,
// set by k_pre), independent of contact order, so each contact's thread wakes
// its pair (wake_pair); otherwise one thread replays the reference's
// sequential loop over the world's contacts (wake_seq) and wake_pair does
// nothing.
__device__ __forceinline__ double wake_speed2(const Dev& d, int64_t i) {
    const double x = BODY(d.lv, 0, i), y = BODY(d.lv, 1, i), z = BODY(d.lv, 2, i);
    return (x * x + y * y) + z * z;
}

__device__ __forceinline__ void wake_pair(const Dev& d, int32_t w, int64_t a, int64_t b) {
    if ((d.w_flags[w] & (WF_ACTIVE | WF_SEQWAKE)) != WF_ACTIVE) return;
    const double th = 0.15 * 0.15;
    const int ma = d.mt[a], mb = d.mt[b];
    const bool wa = ma == M_KINEMATIC || (ma == M_DYNAMIC && d.awake0[a] && wake_speed2(d, a) > th);
    const bool wb = mb == M_KINEMATIC || (mb == M_DYNAMIC && d.awake0[b] && wake_speed2(d, b) > th);
    if (!d.awake0[b] && wa) { d.awake[b] = 1; d.timer[b] = 0.0; }
    if (!d.awake0[a] && wb) { d.awake[a] = 1; d.timer[a] = 0.0; }
}

__device__ __forceinline__ void wake_seq(const Dev& d, int32_t w) {
    if ((d.w_flags[w] & (WF_ACTIVE | WF_SEQWAKE)) != (WF_ACTIVE | WF_SEQWAKE)) return;
    const double th = 0.15 * 0.15;
    const int32_t K = d.w_cn[w], cs = d.w_cs[w];
    for (int32_t k = 0; k < K; k++) {
        const int64_t a = d.ca[cs + k], b = d.cb[cs + k];
        const int ma = d.mt[a], mb = d.mt[b];
        const bool wa = ma == M_KINEMATIC || (ma == M_DYNAMIC && d.awake[a] && wake_speed2(d, a) > th);
        if (!d.awake[b] && wa) { d.awake[b] = 1; d.timer[b] = 0.0; }
        const bool wb = mb == M_KINEMATIC || (mb == M_DYNAMIC && d.awake[b] && wake_speed2(d, b) > th);
        if (!d.awake[a] && wb) { d.awake[a] = 1; d.timer[a] = 0.0; }
    }
}

// ---------------------------------------------------------------- solver --

// Solver row layout per contact (R [kcap][stride]).
// FP64 row: the reference's per-contact quantities; the sweep recomputes the
// angular terms in the reference's operation order. A sweep item loads only
// the first SR_HOT values, so fewer registers hold a prefetched item: not the
// seeds (the level loop reads the impulse state) nor the second tangent
// (recomputed as n x t1, the expression that made it).  SR_TA / SR_TB: the
// split-impulse position terms ima dL and imb dL of the contact (SR_TA = -1:
// no correction).
enum { SR_RA = 0, SR_RB = 3, SR_T1 = 6, SR_NX = 9, SR_NM = 12, SR_T1M = 13, SR_T2M = 14, SR_VB = 15,
       SR_FR = 16, SR_IMA = 17, SR_IMB = 18, SR_DEP = 19, SR_T2 = 20, SR_IMP = 23, SR_TA = 26, SR_TB = 27,
       SR_HOT = 20 };
// FP32 row: directions d = n, t1, t2 with the Jacobian terms folded in,
// cA_d = rA x d, cB_d = rB x d, kA_d = IA (rA x d), kB_d = IB (rB x d), so a
// sweep is dot products and scaled adds (FMA) on the same quantities.
enum { FR_DIR = 0, FR_CA = 9, FR_CB = 18, FR_KA = 27, FR_KB = 36, FR_NM = 45, FR_T1M = 46, FR_T2M = 47,
       FR_VB = 48, FR_FR = 49, FR_IMA = 50, FR_IMB = 51, FR_DEP = 52, FR_IMP = 53, FR_STRIDE = 56 };
template <typename R> struct RowLayout;
template <> struct RowLayout<double> {
    static constexpr int stride = SR_STRIDE, hot = SR_HOT, imp = SR_IMP, dep = SR_DEP, ima = SR_IMA, imb = SR_IMB,
                         nrm = SR_NX;
};
template <> struct RowLayout<float> {
    static constexpr int stride = FR_STRIDE, hot = FR_STRIDE, imp = FR_IMP, dep = FR_DEP, ima = FR_IMA, imb = FR_IMB,
                         nrm = FR_DIR;
};
static_assert(SR_DEP < SR_HOT && SR_NX + 3 <= SR_HOT && SR_HOT % 2 == 0, "the FP64 sweep's fields lie in the hot prefix");
static_assert(SR_NX + 3 <= SR_DEP && SR_IMA < SR_DEP && SR_IMB < SR_DEP && SR_FR < SR_DEP && SR_VB < SR_DEP &&
                  FR_FR < FR_DEP && FR_IMA < FR_DEP && FR_IMB < FR_DEP && FR_VB < FR_DEP && FR_KB + 9 <= FR_DEP,
              "the velocity passes read only the values before the depth");

// n values of type R from 16-byte aligned memory as 16-byte loads
// (a last partial chunk is loaded value by value)
template <typename R, int N>
__device__ __forceinline__ void load_vec(const R* src, R* dst) {
    constexpr int P = 16 / sizeof(R);
#pragma unroll
    for (int q = 0; q < N / P; q++) {
        const uint4 t = ((const uint4*)src)[q];
        __builtin_memcpy(dst + q * P, &t, 16);
    }
#pragma unroll
    for (int q = (N / P) * P; q < N; q++) dst[q] = src[q];
}

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
    for (int probe = 0; probe < HASH_PROBES; probe++) {
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
// (_solver_native.pyx:179-312).  The warm start and the `iters` sweeps are one
// sequence of velocity updates over the contacts in narrowphase order; item
// (s, k) is contact k in pass s.  An item reads and writes only its two bodies,
// and static/kinematic bodies are read-only (their inverse mass and inertia
// are zero, so the reference's writes to them leave them unchanged), so item i
// depends only on the most recent earlier item that touched one of its dynamic
// bodies.  Level L(i) = 1 + max(L of those) puts every dependency on an earlier
// level: the items of one level touch disjoint dynamic bodies and run
// concurrently, every item reads exactly the velocities the sequential loop
// would, and later passes start on bodies whose previous pass is done.  The
// split-impulse position pass reads no velocity and writes only positions, so
// its update of contact k runs with item (0, k), which keeps each body's
// position updates in contact order.
//
// Levels are assigned per manifold (the run of contacts of one pair, which
// chain through both bodies and so take consecutive levels) by one wavefront,
// 64 manifolds at a time.  Each lane ORs its bit into a 64-bit mask per dynamic
// body; the highest lower bit in its bodies' masks is its predecessor inside
// the group.  Levels start from the per-body table `last` and are relaxed over
// those predecessors until no lane changes (rounds = the longest chain inside
// the group); the last lane to touch a body writes its level back.
template <bool G>
struct LevelTable {
    int32_t* last;
    uint64_t* mask;
    __device__ __forceinline__ int32_t get_last(int32_t i) const {
        if (G) return __hip_atomic_load(&last[i], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        return last[i];
    }
    __device__ __forceinline__ uint64_t get_mask(int32_t i) const {
        if (G) return __hip_atomic_load(&mask[i], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        return mask[i];
    }
    __device__ __forceinline__ void set_last(int32_t i, int32_t v) const {
        if (G) __hip_atomic_store(&last[i], v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        else last[i] = v;
    }
    __device__ __forceinline__ void or_mask(int32_t i, uint64_t m) const {
        __hip_atomic_fetch_or(&mask[i], m, __ATOMIC_RELAXED,
                              G ? __HIP_MEMORY_SCOPE_AGENT : __HIP_MEMORY_SCOPE_WORKGROUP);
    }
    __device__ __forceinline__ void clear_mask(int32_t i) const {
        if (G) __hip_atomic_store(&mask[i], (uint64_t)0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        else mask[i] = 0;
    }
    // orders this wavefront's table accesses.  LDS executes one wavefront's
    // instructions in issue order, so there a compiler barrier suffices and no
    // wait is spent; global memory takes a real fence.
    __device__ __forceinline__ void fence() const {
        if (G) __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "agent");
        else asm volatile("" ::: "memory");
    }
};

// Levels of manifold items [0, NI) (item i is manifold i mod M of pass
// i / M) into level[] (the first contact's level); with `lcnt` (LDS, zeroed)
// also the number of contacts on each level below `cap` (`over` is set when a
// level reaches cap).  Returns the depth; d0 gets the deepest level holding a
// pass-0 contact.  Called by the 64 threads of one wavefront; the table starts
// at zero.
//
// With `prev` (M + cap ints of LDS) the groups keep to one pass each, and after
// each pass s >= 1 the largest shift of a manifold's level from pass s - 1,
// Delta = max_m (L_s(m) - L_{s-1}(m)), is known.  Every later pass may then be
// pass s shifted by Delta, L_{s+j}(m) = L_s(m) + j Delta: a shift keeps every
// dependency inside a pass, and a pass-(s+1) item's dependency on the last
// pass-s item q of one of its bodies holds because it held one pass earlier
// (L_s(f) >= L_{s-1}(q) + c_q) and q moved by at most Delta.  Such a schedule
// is as valid as the longest-path one (every item still runs after the items
// it depends on), only deeper when Delta exceeds the longest path's growth per
// pass, so the rest of the passes are shifted once that excess, times the
// passes left, costs fewer level steps than assigning them would (about half
// a level step per group of 64 manifolds).
template <bool G>
__device__ __forceinline__ int32_t assign_item_levels(const LevelTable<G>& T, int32_t M, int32_t NI, const int4* mrec,
                                      int32_t* level, int32_t* lcnt, int32_t cap, bool& over, int32_t& d0,
                                      int32_t* prev = nullptr) {
    const int lane = threadIdx.x & 63;
    int32_t D = 0, Dp = 0;
    const bool aligned = prev != nullptr;
    const int32_t SLp = (NI + M - 1) / M;                  // passes
    const int32_t GP = (M + 63) / 64;                      // groups per pass when aligned
    const int32_t NG = aligned ? SLp * GP : (NI + 63) / 64;
    // this lane's item of the next group: pass sn, group gn within it (aligned),
    // item in, manifold mn
    int32_t sn = 0, gn = 0, in_ = lane, mn = lane;
    if (!aligned) while (mn >= M) mn -= M;
    auto next_valid = [&]() { return aligned ? gn * 64 + lane < M : in_ < NI; };
    bool v_n = next_valid();
    int4 rec = make_int4(0, 1, 0, 0);
    if (v_n) rec = mrec[mn];
    int32_t dmax = 0, dpass = 0, dprev = 0;   // this pass: largest shift, depth; the last pass's depth
    for (int32_t g = 0; g < NG; g++) {
        const bool valid = v_n;
        const int32_t i = in_, m = mn, s = sn;
        const bool pass_end = aligned && gn + 1 == GP;
        const int4 cur = rec;
        // the next group's manifold records load while this group resolves
        if (aligned) {
            if (++gn == GP) { gn = 0; sn++; }
            mn = gn * 64 + lane;
            in_ = sn * M + mn;
        } else {
            in_ += 64;
            mn += 64;
            while (mn >= M) mn -= M;
        }
        v_n = g + 1 < NG && next_valid();
        if (v_n) rec = mrec[mn];
        // the manifold's level one pass earlier (read early: it does not
        // depend on this group)
        const int32_t pl = (aligned && valid && i >= M) ? prev[m] : 0;
        const int32_t al = (valid && (cur.y & MR_DYN_A)) ? cur.z : -1;
        const int32_t bl = (valid && (cur.y & MR_DYN_B)) ? cur.w : -1;
        const int32_t c = valid ? (cur.y & MR_COUNT) : 1;
        const uint64_t bit = 1ull << lane;
        if (al >= 0) T.or_mask(al, bit);
        if (bl >= 0) T.or_mask(bl, bit);
        T.fence();
        const uint64_t below = bit - 1ull, above = ~((bit << 1) - 1ull);
        // both sides' entries read at once (a missing side reads entry 0 and
        // drops it), so the wavefront waits for the table once
        const uint64_t ma0 = T.get_mask(al >= 0 ? al : 0), mb0 = T.get_mask(bl >= 0 ? bl : 0);
        const int32_t la0 = T.get_last(al >= 0 ? al : 0), lb0 = T.get_last(bl >= 0 ? bl : 0);
        const uint64_t ma = al >= 0 ? ma0 : 0ull, mb = bl >= 0 ? mb0 : 0ull;
        int32_t L = 1;
        if (al >= 0 && la0 + 1 > L) L = la0 + 1;
        if (bl >= 0 && lb0 + 1 > L) L = lb0 + 1;
        const int32_t pa = (ma & below) ? 63 - __clzll(ma & below) : lane;
        const int32_t pb = (mb & below) ? 63 - __clzll(mb & below) : lane;
        // relaxation: a predecessor's manifold ends on level L + c - 1 of its
        // lane, so each round exchanges L + c
        for (;;) {
            const int32_t Lc = L + c;
            const int32_t Ea = __shfl(Lc, pa, 64), Eb = __shfl(Lc, pb, 64);
            int32_t nL = L;
            if (pa != lane && Ea > nL) nL = Ea;
            if (pb != lane && Eb > nL) nL = Eb;
            const bool ch = nL != L;
            L = nL;
            if (!__any(ch)) break;
        }
        T.fence();
        const int32_t Le = L + c - 1;   // level of the manifold's last contact
        if (al >= 0 && !(ma & above)) T.set_last(al, Le);
        if (bl >= 0 && !(mb & above)) T.set_last(bl, Le);
        if (al >= 0) T.clear_mask(al);
        if (bl >= 0) T.clear_mask(bl);
        T.fence();
        if (valid) {
            level[i] = L;
            D = Le > D ? Le : D;
            if (i < M) Dp = Le > Dp ? Le : Dp;
            if (lcnt) {
                if (Le <= cap) for (int32_t j = 0; j < c; j++) atomicAdd(&lcnt[L - 1 + j], 1);
                else over = true;
            }
            if (aligned) {
                if (i >= M) { const int32_t dl = L - pl; dmax = dl > dmax ? dl : dmax; }
                prev[m] = L;
                dpass = Le > dpass ? Le : dpass;
            }
        }
        if (pass_end) {
            // pass s done
#pragma unroll
            for (int o = 32; o > 0; o >>= 1) {
                int32_t x = __shfl_xor(dmax, o, 64); dmax = x > dmax ? x : dmax;
                x = __shfl_xor(dpass, o, 64); dpass = x > dpass ? x : dpass;
            }
            // (hs, cap ints after `prev`: the contacts per level after the
            // previous pass, on levels below dprev)
            int32_t* hs = prev + M;
            if (s >= 1 && s + 1 < SLp && 2 * (dmax - (dpass - dprev)) <= GP) {
                // the remaining passes: pass s shifted by dmax per pass; their
                // contacts per level are pass s's, shifted
                const int32_t last = dpass + (SLp - 1 - s) * dmax;
                D = last > D ? last : D;
                for (int32_t mm = lane; mm < M; mm += 64) {
                    const int32_t L0 = prev[mm];
                    for (int32_t j = s + 1; j < SLp; j++) level[j * M + mm] = L0 + (j - s) * dmax;
                }
                if (lcnt && !over && last <= cap) {
                    // pass s's contacts per level: the counts now less those
                    // after the previous pass (all read before any is added to)
                    for (int32_t l = lane; l < dpass; l += 64) hs[l] = lcnt[l] - (l < dprev ? hs[l] : 0);
                    T.fence();
                    for (int32_t l = lane; l < dpass; l += 64) {
                        const int32_t h = hs[l];
                        if (h)
                            for (int32_t j = 1; j < SLp - s; j++) atomicAdd(&lcnt[l + j * dmax], h);
                    }
                } else if (lcnt) {
                    over = true;
                }
                break;
            }
            if (lcnt && dpass <= cap) {
                T.fence();
                for (int32_t l = lane; l < dpass; l += 64) hs[l] = lcnt[l];
                T.fence();
            }
            dprev = dpass;
            dmax = 0;
            dpass = 0;
        }
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        int32_t x = __shfl_xor(D, o, 64); D = x > D ? x : D;
        x = __shfl_xor(Dp, o, 64); Dp = x > Dp ? x : Dp;
    }
    over = __any(over);
    d0 = Dp;
    return D;
}

// Schedule item code: contact (bits 31..7), side a dynamic (6), side b
// dynamic (5), pass (4..0).
__device__ __forceinline__ int32_t item_code(int32_t k, bool da, bool db, int32_t s) {
    return (k << 7) | ((int32_t)da << 6) | ((int32_t)db << 5) | s;
}

// One scheduled item's immutable data: the contact's solver row and the
// bodies' world inverse inertia.
template <typename R>
struct SolveItem {
    int32_t ci, a, b;               // contact, bodies (batch indices)
    int32_t k;                      // contact within the world or island (its impulse state)
    int32_t s, cida, cidb;          // pass; compact ids of the sides (LDS state)
    bool da, db;
    R row[RowLayout<R>::hot];
    R Ia[sizeof(R) == 8 ? 9 : 1], Ib[sizeof(R) == 8 ? 9 : 1];   // FP64 rows only
};
static_assert(SR_STRIDE * sizeof(double) == SOLVE_ROW_BYTES && FR_STRIDE * sizeof(float) == SOLVE_ROW_BYTES,
              "solver rows fill SOLVE_ROW_BYTES in both precisions");

// The split-impulse position terms of one contact (_solver_native.pyx:297-305):
// dL = correction (depth - slop) / (ima + imb), ta = ima dL, tb = imb dL;
// false (no correction) unless depth - slop > 0 and ima + imb > 1e-12.
template <typename R>
__device__ __forceinline__ bool pos_terms(R ima, R imb, R dep, const SolveIn& in, R& ta, R& tb) {
    const R vn = dep - (R)in.slop;
    if (!(vn > R(0) && ima + imb > R(1e-12))) return false;
    const R dL = (R)in.correction * vn / (ima + imb);
    ta = ima * dL;
    tb = imb * dL;
    return true;
}

// The solver row of contact ci (bodies a, b) of world w, whose contacts start
// at cs (_solver_native.pyx:179-256): lever arms, tangent basis, effective
// masses, restitution bias, friction, seeds (warm-start cache,
// solver._seed_pair, or the stage call's arrays), from the pre-solve
// velocities.
template <typename R, bool STAGE>
__device__ __forceinline__ void prepare_contact(const Dev& d, const SolveIn& in, int64_t ci, int64_t a, int64_t b,
                                                int32_t w, int32_t cs, int32_t old) {
    R* S_ = (R*)d.csolv;
    const R* imp = (const R*)d.cimp;
    const R* cn = (const R*)d.cnrm;
    const R* cdep = (const R*)d.cdep;
    const R* iiw = (const R*)d.iiw;
    {
        R rest, fric, n0 = R(0), u0 = R(0), w0 = R(0);
        if (STAGE) {
            rest = (R)in.st_rest[ci - cs]; fric = (R)in.st_fric[ci - cs];
            n0 = CON(imp, 0, ci); u0 = CON(imp, 1, ci); w0 = CON(imp, 2, ci);
        } else {
            const int32_t tb = d.w_mat_off[w] + d.mat[a] * d.w_nmat[w] + d.mat[b];
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
        for (int r = 0; r < 9; r++) { Ia[r] = IIW(iiw, a, r); Ib[r] = IIW(iiw, b, r); }
        for (int c = 0; c < 6; c++) { va[c] = (R)BODY(d.lv, c, a); vb_[c] = (R)BODY(d.lv, c, b); }
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
        constexpr int RS = RowLayout<R>::stride;
        R row[RS];
        if (sizeof(R) == sizeof(double)) {
            for (int r = 0; r < 3; r++) { row[SR_RA + r] = rA[r]; row[SR_RB + r] = rB[r]; }
            row[SR_T1] = tx; row[SR_T1 + 1] = ty; row[SR_T1 + 2] = tz;
            row[SR_T2] = ux; row[SR_T2 + 1] = uy; row[SR_T2 + 2] = uz;
            row[SR_NX] = nx; row[SR_NX + 1] = ny; row[SR_NX + 2] = nz;
            R ta = R(-1), tb = R(0);
            pos_terms<R>(ima, imb, (R)cdep[ci], in, ta, tb);
            row[SR_TA] = ta; row[SR_TB] = tb;
        } else {
            const R dir[9] = {nx, ny, nz, tx, ty, tz, ux, uy, uz};
            for (int q = 0; q < 9; q++) row[FR_DIR + q] = dir[q];
            for (int dd = 0; dd < 3; dd++) {
                const R* e = dir + 3 * dd;
                const R ca[3] = {rA[1] * e[2] - rA[2] * e[1], rA[2] * e[0] - rA[0] * e[2], rA[0] * e[1] - rA[1] * e[0]};
                const R cb[3] = {rB[1] * e[2] - rB[2] * e[1], rB[2] * e[0] - rB[0] * e[2], rB[0] * e[1] - rB[1] * e[0]};
                for (int r = 0; r < 3; r++) {
                    row[FR_CA + 3 * dd + r] = ca[r];
                    row[FR_CB + 3 * dd + r] = cb[r];
                    row[FR_KA + 3 * dd + r] = Ia[3 * r] * ca[0] + Ia[3 * r + 1] * ca[1] + Ia[3 * r + 2] * ca[2];
                    row[FR_KB + 3 * dd + r] = Ib[3 * r] * cb[0] + Ib[3 * r + 1] * cb[1] + Ib[3 * r + 2] * cb[2];
                }
            }
        }
        const int NM_ = sizeof(R) == 8 ? (int)SR_NM : (int)FR_NM, T1M_ = sizeof(R) == 8 ? (int)SR_T1M : (int)FR_T1M;
        const int T2M_ = sizeof(R) == 8 ? (int)SR_T2M : (int)FR_T2M, VB_ = sizeof(R) == 8 ? (int)SR_VB : (int)FR_VB;
        const int FRI_ = sizeof(R) == 8 ? (int)SR_FR : (int)FR_FR;
        row[NM_] = effmass(rA, rB, nx, ny, nz, Ia, Ib, ima, imb);
        row[T1M_] = effmass(rA, rB, tx, ty, tz, Ia, Ib, ima, imb);
        row[T2M_] = effmass(rA, rB, ux, uy, uz, Ia, Ib, ima, imb);
        row[VB_] = vn0 < -(R)in.rthr ? -rest * vn0 : R(0);
        row[FRI_] = fric;
        row[RowLayout<R>::ima] = ima; row[RowLayout<R>::imb] = imb;
        row[RowLayout<R>::imp] = n0; row[RowLayout<R>::imp + 1] = u0; row[RowLayout<R>::imp + 2] = w0;
        row[RowLayout<R>::dep] = cdep[ci];
        R* dst = &ROW(S_, ci, 0);
        constexpr int P = 16 / sizeof(R);
#pragma unroll
        for (int q = 0; q < RS / P; q++) {
            uint4 t;
            __builtin_memcpy(&t, row + q * P, 16);
            ((uint4*)dst)[q] = t;
        }
    }
}

// Rows of the step's contacts, one thread per contact; with skip_above > 0
// only those of worlds of at most skip_above bodies (the larger worlds'
// rows come from k_prepare_worlds).
template <typename R, bool STAGE>
__global__ __launch_bounds__(256) void k_prepare(Dev d, SolveIn in, int32_t skip_above) {
    const uint64_t tc = d.cnt[C_CONTACTS];
    const int64_t T = tc < (uint64_t)d.kcap ? (int64_t)tc : d.kcap;
    const int32_t old = *d.sel;
    for (int64_t ci = gtid(); ci < T; ci += gstride()) {
        const int64_t a = d.ca[ci], b = d.cb[ci];
        const int32_t w = d.world[a];
        if (w < 0) continue;
        if (skip_above > 0 && d.w_cnt[w] > skip_above) continue;
        const int32_t cs = d.w_cs[w];
        if (ci < cs || ci >= (int64_t)cs + d.w_cn[w]) continue;   // a world that overflowed
        if (in.wake) {
            if (ci == cs) wake_seq(d, w);
            wake_pair(d, w, a, b);
        }
        prepare_contact<R, STAGE>(d, in, ci, a, b, w, cs, old);
    }
}

// Rows of the contacts of the listed worlds (list wl, count *wn): G blocks
// per world, their threads striding over its contacts.
template <typename R, bool STAGE>
__global__ __launch_bounds__(256) void k_prepare_worlds(Dev d, SolveIn in, const int32_t* wl, const int32_t* wn,
                                                        int32_t G) {
    const int32_t j = blockIdx.x / G, part = blockIdx.x % G;
    if (j >= *wn) return;
    const int32_t w = wl[j];
    if (w >= d.wcap || !(d.w_flags[w] & WF_ACTIVE)) return;
    const int32_t cs = d.w_cs[w], K = d.w_cn[w];
    const int32_t old = *d.sel;
    if (in.wake && part == 0 && threadIdx.x == 0) wake_seq(d, w);
    for (int32_t k = part * 256 + (int32_t)threadIdx.x; k < K; k += G * 256) {
        const int64_t ci = (int64_t)cs + k;
        const int64_t a = d.ca[ci], b = d.cb[ci];
        if (in.wake) wake_pair(d, w, a, b);
        prepare_contact<R, STAGE>(d, in, ci, a, b, w, cs, old);
    }
}

// Level bounds kept in LDS per world (deeper schedules use global memory).
constexpr int32_t SOLVE_DCAP = 1024;

__host__ __device__ __forceinline__ size_t al16(size_t x) { return (x + 15) & ~(size_t)15; }

// Where one world's solver state lives during the level loop.  In LDS
// (byte offsets into the launch's dynamic LDS): the compact id of every body in
// contact (map, world-local index -> id), their velocities (one record of
// VelRec<R>::stride R per id: 6 doubles, or 6 floats padded to 8, so a body's
// velocity is three or two 16-byte LDS accesses), the contacts' impulses (3 R
// per contact, contact-major), and positions (FP64) or position corrections
// (FP32) of the bodies in contact (3 R per id); otherwise the velocities,
// impulses and positions of the batch in global memory.
template <typename R> struct VelRec { static constexpr int stride = sizeof(R) == 8 ? 6 : 8; };
template <typename R>
struct SolveState {
    size_t map, vel, imp, pos;     // LDS offsets
    int32_t nd;                    // bodies in contact (compact ids)
    bool pos_lds;
    R* V; int64_t vs, vi;          // global velocities: V[c * vs + i * vi]
    R* P; int64_t ps, pc, pb;      // global impulses: P[(ci - pb) * ps + c * pc]
    int64_t off;
};

// Bucket the contact items by level: bound[l] is the start of level l on
// entry (after the scan) and its end on exit.  A thread takes a manifold and
// loads its levels of PB passes at once before it places their items (the
// item stores would otherwise order every level load after the previous
// pass's stores).  The order of the items inside a level is free: they share
// no dynamic body.  cord_c (the item's batch contact) is written for island
// solves only (gmap).
template <int BLOCK>
__device__ __forceinline__ void scatter_items(int32_t* bound, int32_t SL, int32_t M, const int4* mrec, const int32_t* level,
                              int32_t cl, const int32_t* gmap, int64_t off, int32_t* corder, int32_t* cord_a,
                              int32_t* cord_b, int32_t* cord_c) {
    constexpr int PB = 8;
    for (int32_t m = threadIdx.x; m < M; m += BLOCK) {
        const int4 r = mrec[m];
        const int32_t c = r.y & MR_COUNT;
        const bool da = r.y & MR_DYN_A, db = r.y & MR_DYN_B;
        const int32_t ka = (int32_t)off + r.z, kb = (int32_t)off + r.w;
        for (int32_t s0 = 0; s0 < SL; s0 += PB) {
            int32_t Ls[PB];
#pragma unroll
            for (int q = 0; q < PB; q++) Ls[q] = s0 + q < SL ? level[(size_t)(s0 + q) * M + m] : 0;
#pragma unroll
            for (int q = 0; q < PB; q++) {
                if (s0 + q >= SL) break;
                for (int32_t j = 0; j < c; j++) {
                    const int32_t slot = atomicAdd(&bound[Ls[q] - 1 + j], 1);
                    const int32_t k = r.x - cl + j;
                    corder[slot] = item_code(k, da, db, s0 + q);
                    if (gmap) cord_c[slot] = gmap[k];
                    cord_a[slot] = ka;
                    cord_b[slot] = kb;
                }
            }
        }
    }
}

// Contacts per level into bound[] (zeroed, global), PB passes' levels loaded
// at once as in scatter_items.
template <int BLOCK>
__device__ __forceinline__ void count_items(int32_t* bound, int32_t SL, int32_t M, const int4* mrec,
                                            const int32_t* level) {
    constexpr int PB = 8;
    for (int32_t m = threadIdx.x; m < M; m += BLOCK) {
        const int32_t c = mrec[m].y & MR_COUNT;
        for (int32_t s0 = 0; s0 < SL; s0 += PB) {
            int32_t Ls[PB];
#pragma unroll
            for (int q = 0; q < PB; q++) Ls[q] = s0 + q < SL ? level[(size_t)(s0 + q) * M + m] : 0;
#pragma unroll
            for (int q = 0; q < PB; q++) {
                if (s0 + q >= SL) break;
                for (int32_t j = 0; j < c; j++) atomicAdd(&bound[Ls[q] - 1 + j], 1);
            }
        }
    }
}

// The level loop of one world, in steps of at most BLOCK items of one level.
// Each thread's item of step k + 2 (code and bodies) and of step k + 1 (row
// and inverse inertia) load while step k runs.  POS = false: the velocity passes (warm start and sweeps) over all D
// levels; POS = true: the split-impulse position pass, which reads only the
// contacts' depth, masses and normal and so runs on its own after the
// velocity passes, over the levels holding pass-0 items (D is then that
// count).  LIGHT: the state the mode updates (velocities and impulses, or
// positions) is in LDS and a level ends on an LDS barrier; otherwise on a
// full barrier.  LB: the level bounds are in LDS, else in global memory.  LB
// is a template argument, not a flag, so that no global load feeds the bound
// registers the loop carries: the compiler's wait for such a load, placed at
// every level, would also wait for the next level's item loads and expose
// their latency.
template <typename R, int BLOCK, bool LIGHT, bool POS, bool LB, bool ISL>
__device__ __forceinline__ void run_levels(const Dev& d, const SolveIn& in, const SolveState<R>& st, size_t bound_off,
                           int32_t K, int32_t D, int32_t NC, int32_t S, const int32_t* gbound,
                           const int32_t* corder, const int32_t* cord_a, const int32_t* cord_b,
                           const int32_t* cord_c, int32_t cs) {
    extern __shared__ __align__(16) unsigned char s_dyn[];
    const int tid = threadIdx.x;
    const bool fp64 = sizeof(R) == sizeof(double);
    // LIGHT: LDS pointers built here from the shared array so the compiler
    // addresses them as LDS
    const int32_t* lmap = (const int32_t*)(s_dyn + st.map);
    R* lV = (R*)(s_dyn + st.vel);
    R* lP = (R*)(s_dyn + st.imp);
    R* lpos = (R*)(s_dyn + st.pos);
    const int32_t* lbound = (const int32_t*)(s_dyn + bound_off);
    const R* S_ = (const R*)d.csolv;
    const R* iiw = (const R*)d.iiw;
    // velocity c of side body i (global index) with compact id ci_
#define IMP(c, ci) (LIGHT ? lP[(size_t)it.k * 3 + (size_t)(c)] \
                          : st.P[(size_t)((ci) - st.pb) * st.ps + (size_t)(c) * st.pc])
    // an island's item carries its batch contact; a world's is cs + its contact
    struct IdRec { int32_t code, a, b, c; };
    auto fetch_id = [&](int32_t slot, IdRec& r) {
        if (slot < NC) {
            r.code = corder[slot]; r.a = cord_a[slot]; r.b = cord_b[slot];
            if constexpr (ISL) r.c = cord_c[slot];
        }
    };
    auto fetch_item = [&](const IdRec& r, SolveItem<R>& it) {
        it.s = r.code & 31;
        it.da = (r.code >> 6) & 1;
        it.db = (r.code >> 5) & 1;
        it.k = r.code >> 7;
        if constexpr (ISL) it.ci = r.c;
        else it.ci = cs + it.k;
        it.a = r.a; it.b = r.b;
        if (LIGHT) { it.cida = lmap[r.a - st.off]; it.cidb = lmap[r.b - st.off]; }
        // exactly the values the mode reads (the velocity passes stop short of
        // the depth): a loaded value nothing reads lets the compiler reuse its
        // register, which then waits for the load in flight
        load_vec<R, POS ? RowLayout<R>::dep + 1 : RowLayout<R>::dep>(&ROW(S_, it.ci, 0), it.row);
        if constexpr (sizeof(R) == 8 && !POS) {
            load_vec<R, 9>(&IIW(iiw, it.a, 0), it.Ia);
            load_vec<R, 9>(&IIW(iiw, it.b, 0), it.Ib);
        }
    };
    // a body's six velocity components: 16-byte LDS accesses in LIGHT mode
    auto vel_load = [&](int64_t i, int32_t cid, R* v) {
        if constexpr (LIGHT) {
            const R* p = lV + (size_t)cid * VelRec<R>::stride;
            if constexpr (sizeof(R) == 8) {
#pragma unroll
                for (int q = 0; q < 3; q++) { const double2 t = ((const double2*)p)[q]; v[2 * q] = t.x; v[2 * q + 1] = t.y; }
            } else {
                const float4 t0 = ((const float4*)p)[0];
                const float2 t1 = ((const float2*)p)[2];
                v[0] = t0.x; v[1] = t0.y; v[2] = t0.z; v[3] = t0.w; v[4] = t1.x; v[5] = t1.y;
            }
        } else {
            for (int c = 0; c < 6; c++) v[c] = st.V[(size_t)c * st.vs + (size_t)i * st.vi];
        }
    };
    auto vel_store = [&](int64_t i, int32_t cid, const R* v) {
        if constexpr (LIGHT) {
            R* p = lV + (size_t)cid * VelRec<R>::stride;
            if constexpr (sizeof(R) == 8) {
#pragma unroll
                for (int q = 0; q < 3; q++) ((double2*)p)[q] = make_double2(v[2 * q], v[2 * q + 1]);
            } else {
                ((float4*)p)[0] = make_float4(v[0], v[1], v[2], v[3]);
                ((float2*)p)[2] = make_float2(v[4], v[5]);
            }
        } else {
            for (int c = 0; c < 6; c++) st.V[(size_t)c * st.vs + (size_t)i * st.vi] = v[c];
        }
    };
    auto pos_add = [&](int64_t i, int32_t cid, int r, R term, bool neg, bool dyn) {
        // split-impulse correction of body i (global index, compact id cid),
        // component r.  FP64 positions are the state itself; FP32 accumulates
        // the correction from zero and adds it to the FP64 position after.
        R* pp = LIGHT ? &lpos[(size_t)r * st.nd + (size_t)cid] : nullptr;
        if (!fp64) {
            if (!dyn) return;
            if (LIGHT) { if (neg) *pp -= term; else *pp += term; }
            else BODY(d.pos, r, i) = neg ? BODY(d.pos, r, i) - (double)term : BODY(d.pos, r, i) + (double)term;
            return;
        }
        const double p = LIGHT ? (double)*pp : BODY(d.pos, r, i);
        const double q = neg ? p - (double)term : p + (double)term;
        if (dyn) {
            if (LIGHT) *pp = (R)q; else BODY(d.pos, r, i) = q;
        } else if (__double_as_longlong(q) != __double_as_longlong(p)) {
            // a static or kinematic side (inverse mass 0, term +-0): the only
            // change is -0.0 -> +0.0, written idempotently since several
            // items of one level may touch the same such body
            if (LIGHT) *pp = (R)q; else BODY(d.pos, r, i) = q;
        }
    };
    auto process = [&](const SolveItem<R>& it) {
        const int64_t a = it.a, b = it.b, ci = it.ci;
        const R ima = it.row[RowLayout<R>::ima], imb = it.row[RowLayout<R>::imb];
        const R* nv = it.row + RowLayout<R>::nrm;
        if constexpr (POS) {
            // split-impulse position pass (_solver_native.pyx:297-305)
            if (it.s != 0) return;
            const R vn = it.row[RowLayout<R>::dep] - (R)in.slop;
            if (vn > R(0) && ima + imb > R(1e-12)) {
                const R dL = (R)in.correction * vn / (ima + imb);
                for (int r = 0; r < 3; r++) pos_add(a, it.cida, r, ima * dL * nv[r], true, it.da);
                for (int r = 0; r < 3; r++) pos_add(b, it.cidb, r, imb * dL * nv[r], false, it.db);
            }
            return;
        }
        if (it.s >= S) return;
        if constexpr (sizeof(R) == 8) {
            // the reference's operation order (bit-exact FP64)
            const R* rA = it.row + SR_RA;
            const R* rB = it.row + SR_RB;
            const R* t1 = it.row + SR_T1;
            // t2 = n x t1 exactly as the row's preparation formed it
            const R t2[3] = {nv[1] * t1[2] - nv[2] * t1[1], nv[2] * t1[0] - nv[0] * t1[2],
                             nv[0] * t1[1] - nv[1] * t1[0]};
            const R* Ia = it.Ia;
            const R* Ib = it.Ib;
            R va[6], vb2[6];
            vel_load(a, it.cida, va);
            vel_load(b, it.cidb, vb2);
            // A wavefront's lanes can hold warm-start and sweep items at once:
            // both impulses go through one apply_w, so such a wavefront runs
            // that update once instead of once per kind.
            const bool warm = it.s < in.warm;
            R nI = IMP(0, ci), tI0 = IMP(1, ci), tI1 = IMP(2, ci);
            R Px, Py, Pz;
            if (warm) {
                // warm start (_solver_native.pyx:258-266): the seeds' impulse
                Px = nI * nv[0] + tI0 * t1[0] + tI1 * t2[0];
                Py = nI * nv[1] + tI0 * t1[1] + tI1 * t2[1];
                Pz = nI * nv[2] + tI0 * t1[2] + tI1 * t2[2];
            } else {
                // one sweep over the contact (_solver_native.pyx:268-295),
                // normal impulse first
                R dvx, dvy, dvz;
                relvel(va, vb2, rA, rB, dvx, dvy, dvz);
                R vn = dvx * nv[0] + dvy * nv[1] + dvz * nv[2];
                R nw = nI - it.row[SR_NM] * (vn - it.row[SR_VB]);
                if (nw < R(0)) nw = R(0);
                const R dL = nw - nI;
                nI = nw;
                Px = dL * nv[0]; Py = dL * nv[1]; Pz = dL * nv[2];
            }
            apply_w(va, vb2, ima, imb, Ia, Ib, rA, rB, Px, Py, Pz);
            if (!warm) {
                R dvx, dvy, dvz, dL;
                const R maxf = it.row[SR_FR] * nI;
                relvel(va, vb2, rA, rB, dvx, dvy, dvz);
                R vt = dvx * t1[0] + dvy * t1[1] + dvz * t1[2];
                R nt = tI0 - it.row[SR_T1M] * vt;
                if (nt > maxf) nt = maxf; else if (nt < -maxf) nt = -maxf;
                dL = nt - tI0; tI0 = nt;
                apply_w(va, vb2, ima, imb, Ia, Ib, rA, rB, dL * t1[0], dL * t1[1], dL * t1[2]);
                vt = dvx * t2[0] + dvy * t2[1] + dvz * t2[2];
                nt = tI1 - it.row[SR_T2M] * vt;
                if (nt > maxf) nt = maxf; else if (nt < -maxf) nt = -maxf;
                dL = nt - tI1; tI1 = nt;
                apply_w(va, vb2, ima, imb, Ia, Ib, rA, rB, dL * t2[0], dL * t2[1], dL * t2[2]);
                IMP(0, ci) = nI; IMP(1, ci) = tI0; IMP(2, ci) = tI1;
            }
            if (it.da) vel_store(a, it.cida, va);
            if (it.db) vel_store(b, it.cidb, vb2);
        } else {
            // FP32: the same updates on the folded Jacobian terms, with FMA
            const R* row = it.row;
            R va[6], vb2[6];
            vel_load(a, it.cida, va);
            vel_load(b, it.cidb, vb2);
            // J_d . v = (vb - va) . d + wb . (rB x d) - wa . (rA x d)
            auto jv = [&](int dd) -> R {
                const R* e = row + FR_DIR + 3 * dd;
                const R* ca = row + FR_CA + 3 * dd;
                const R* cb = row + FR_CB + 3 * dd;
                R x = (vb2[0] - va[0]) * e[0];
                x = __builtin_fmaf(vb2[1] - va[1], e[1], x);
                x = __builtin_fmaf(vb2[2] - va[2], e[2], x);
                x = __builtin_fmaf(vb2[3], cb[0], x);
                x = __builtin_fmaf(vb2[4], cb[1], x);
                x = __builtin_fmaf(vb2[5], cb[2], x);
                x = __builtin_fmaf(-va[3], ca[0], x);
                x = __builtin_fmaf(-va[4], ca[1], x);
                return __builtin_fmaf(-va[5], ca[2], x);
            };
            // impulse lam along direction dd
            auto apply = [&](int dd, R lam) {
                const R* e = row + FR_DIR + 3 * dd;
                const R* ka = row + FR_KA + 3 * dd;
                const R* kb = row + FR_KB + 3 * dd;
                const R la = ima * lam, lb = imb * lam;
                for (int r = 0; r < 3; r++) {
                    va[r] = __builtin_fmaf(-la, e[r], va[r]);
                    vb2[r] = __builtin_fmaf(lb, e[r], vb2[r]);
                    va[3 + r] = __builtin_fmaf(-lam, ka[r], va[3 + r]);
                    vb2[3 + r] = __builtin_fmaf(lam, kb[r], vb2[3 + r]);
                }
            };
            // warm start (_solver_native.pyx:258-266): the seeds along n, t1,
            // t2; a sweep (_solver_native.pyx:268-295): the normal impulse,
            // then both tangent ones from the velocity after it.  The three
            // updates are shared by both kinds of item.
            const bool warm = it.s < in.warm;
            R nI = IMP(0, ci), tI0 = IMP(1, ci), tI1 = IMP(2, ci);
            R lam0 = nI, lam1 = tI0, lam2 = tI1;
            if (!warm) {
                R nw = nI - row[FR_NM] * (jv(0) - row[FR_VB]);
                if (nw < R(0)) nw = R(0);
                lam0 = nw - nI;
                nI = nw;
            }
            apply(0, lam0);
            if (!warm) {
                const R maxf = row[FR_FR] * nI;
                const R vt1 = jv(1), vt2 = jv(2);
                R nt = tI0 - row[FR_T1M] * vt1;
                if (nt > maxf) nt = maxf; else if (nt < -maxf) nt = -maxf;
                lam1 = nt - tI0;
                tI0 = nt;
                nt = tI1 - row[FR_T2M] * vt2;
                if (nt > maxf) nt = maxf; else if (nt < -maxf) nt = -maxf;
                lam2 = nt - tI1;
                tI1 = nt;
            }
            apply(1, lam1);
            apply(2, lam2);
            if (!warm) { IMP(0, ci) = nI; IMP(1, ci) = tI0; IMP(2, ci) = tI1; }
            if (it.da) vel_store(a, it.cida, va);
            if (it.db) vel_store(b, it.cidb, vb2);
        }
    };
    // end of level l (l < D)
    auto bound_raw = [&](int32_t l) -> int32_t {
        if constexpr (LB) return lbound[l];
        // past L1: written by this block's atomics
        return __hip_atomic_load(&gbound[l], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    };
    // the items of levels [0, D): all NC for the velocity passes, those up
    // to the last pass-0 level for the position pass
    int32_t NE = D > 0 ? bound_raw(D - 1) : 0;
    NE = NE < NC ? NE : NC;
    // Steps: each level's items in chunks of at most BLOCK, so every thread
    // runs at most one item per step (the items of a level share no body, so
    // any split of a level is a valid order).  sl/sp: the level and start of
    // the next step to define; all uniform.
    int32_t sl = 0, sp = 0;
    auto step_end = [&]() -> int32_t {
        if (sp >= NE) return NE;
        const int32_t le = sl < D ? bound_raw(sl) : NE;
        const int32_t e = sp + BLOCK < le ? sp + BLOCK : le;
        sp = e;
        if (e == le) sl++;
        return e;
    };
    // The loop is unrolled by two so the two item buffers swap roles instead
    // of being copied (a copy would wait for the loads in flight).  Each step
    // loads the ids of step k + 2 and the items of step k + 1 while step k
    // runs.
    auto level_barrier = [&]() {
        if (LIGHT) {
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
            if (BLOCK > 64) __builtin_amdgcn_s_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
        } else {
            __syncthreads();
        }
    };
    // One step of prefetch: an item's loads are bound by the vector memory
    // path (each lane's 16-byte piece of its own row is a separate cache line),
    // not by latency, so a deeper prefetch gains nothing.
    int32_t e_prev = 0;
    int32_t e_cur = step_end();
    int32_t e_next = step_end();
    IdRec idA{0, 0, 0, 0}, idB{0, 0, 0, 0};
    SolveItem<R> itA, itB;
    fetch_id(tid, idA);
    if (tid < e_cur) fetch_item(idA, itA);
    fetch_id(e_cur + tid, idB);     // step 1
    // Every step starts with nothing in flight: the loads a step issues
    // for the next one are waited for at its end, after its own items
    // ran.  The compiler then has no pending load to place a wait for
    // inside an item (where, as it cannot tell which loads a lane
    // skipped, it would wait for the newest ones).
    __builtin_amdgcn_s_waitcnt(0);
    auto do_step = [&](SolveItem<R>& cur, SolveItem<R>& nxt, const IdRec& idn, IdRec& id2) {
        const int32_t e_next2 = step_end();
        fetch_id(e_next + tid, id2);                               // step k + 2
        if (e_cur + tid < e_next) fetch_item(idn, nxt);           // step k + 1
        if (e_prev + tid < e_cur) process(cur);
        __builtin_amdgcn_s_waitcnt(0);
        level_barrier();
        e_prev = e_cur; e_cur = e_next; e_next = e_next2;
    };
    while (e_prev < NE) {
        do_step(itA, itB, idB, idA);
        if (e_prev < NE) do_step(itB, itA, idA, idB);
    }
#undef IMP
}

// run_levels with the bounds' place chosen at run time
template <typename R, int BLOCK, bool LIGHT, bool POS, bool ISL, typename... A>
__device__ __forceinline__ void run_levels_lb(bool lds_bounds, A&&... a) {
    if (lds_bounds) run_levels<R, BLOCK, LIGHT, POS, true, ISL>(a...);
    else run_levels<R, BLOCK, LIGHT, POS, false, ISL>(a...);
}

// LDS of the per-body position pass for a world of n bodies, K contacts and M
// manifolds: list starts (n + 1), entries (2K), active contacts per manifold
// (M), active flags (K bits) and, with rb = sizeof(R), the contacts' terms
// and normals (5K of R).
__host__ __device__ __forceinline__ size_t split_pos_bytes(int32_t K, int32_t M, int32_t n, size_t rb) {
    return al16((size_t)(n + 1) * 4) + al16((size_t)K * 8) + al16((size_t)M) + al16((size_t)(K + 31) / 32 * 4) +
           al16((size_t)K * 5 * rb);
}

// The split-impulse position pass (_solver_native.pyx:297-305) of one world,
// by the block, in `lds` (split_pos_bytes(K, M, n, 0) bytes at least; the
// contacts' terms and normals are staged there too when `staged`, else each
// entry reads them from its row).  The pass reads only contact depth, masses
// and normal, and bodies do not interact in it: each dynamic body takes the
// corrections of its contacts in contact order.  So the block builds per-body
// entry lists in that order, then every body folds its list in registers: as
// many steps as the longest list, where a level schedule of the pass-0 items
// would take the depth of the velocity pass.  The lists fill a manifold at a
// time (its contacts are consecutive and share the pair): a wavefront of
// manifolds per step, each lane ranked among the earlier lanes naming the
// same body, weighted by their active contacts.  A static or kinematic side
// (inverse mass 0, terms +-0) can only turn a -0.0 coordinate into +0.0; that
// is written directly, idempotently.  FP32 accumulates the corrections from
// zero and adds them to the FP64 position.
template <typename R, int BLOCK>
__device__ __forceinline__ void split_positions(const Dev& d, const SolveIn& in, int32_t cl, const int32_t* gmap, int32_t K,
                                int32_t M, int32_t n, int64_t off, const int4* mrec, unsigned char* lds, bool staged,
                                int32_t* s_scan) {
    constexpr int NW = BLOCK / 64;
    constexpr bool fp64 = sizeof(R) == sizeof(double);
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    int32_t* cur = (int32_t*)lds;                                          // [n + 1]
    unsigned char* q = lds + al16((size_t)(n + 1) * 4);
    int32_t* ent = (int32_t*)q;                                            // [2K]: contact << 1 | side b
    q += al16((size_t)K * 8);
    uint8_t* nact = (uint8_t*)q;                                           // [M]
    q += al16((size_t)M);
    uint32_t* act = (uint32_t*)q;                                          // [K / 32]
    q += al16((size_t)(K + 31) / 32 * 4);
    R* T = (R*)q;                                                          // [5][K]: ima dL, imb dL, normal
    const R* S_ = (const R*)d.csolv;
    for (int32_t i = tid; i <= n; i += BLOCK) cur[i] = 0;
    for (int32_t i = tid; i < (K + 31) / 32; i += BLOCK) act[i] = 0;
    __syncthreads();
    // contact k's terms (FP64: computed by k_prepare) and normal
    auto terms = [&](int32_t k, R& ta, R& tb, R* nv) -> bool {
        const int64_t ci = gmap ? (int64_t)gmap[k] : (int64_t)cl + k;
        for (int r = 0; r < 3; r++) nv[r] = ROW(S_, ci, RowLayout<R>::nrm + r);
        if constexpr (fp64) {
            ta = ROW(S_, ci, SR_TA);
            tb = ROW(S_, ci, SR_TB);
            return ta >= R(0);
        } else {
            return pos_terms<R>(ROW(S_, ci, RowLayout<R>::ima), ROW(S_, ci, RowLayout<R>::imb),
                                ROW(S_, ci, RowLayout<R>::dep), in, ta, tb);
        }
    };
    // per contact: active flag, and its terms and normal when staged
#pragma unroll 4
    for (int32_t k = tid; k < K; k += BLOCK) {
        R ta = R(0), tb = R(0), nv[3];
        if (!terms(k, ta, tb, nv)) continue;
        atomicOr(&act[k >> 5], 1u << (k & 31));
        if (staged) {
            T[k] = ta; T[K + k] = tb;
            for (int r = 0; r < 3; r++) T[(size_t)(2 + r) * K + k] = nv[r];
        }
    }
    __syncthreads();
    auto active = [&](int32_t k) { return (act[k >> 5] >> (k & 31) & 1u) != 0; };
    // per manifold: active contacts; entries per dynamic body; the -0.0 rule
    // on the other sides
    for (int32_t m = tid; m < M; m += BLOCK) {
        const int4 r4 = mrec[m];
        const int32_t c = r4.y & MR_COUNT, k0 = r4.x - cl;
        int32_t na = 0;
        for (int32_t j = 0; j < c; j++) {
            const int32_t k = k0 + j;
            if (!active(k)) continue;
            na++;
            if (fp64 && (r4.y & (MR_DYN_A | MR_DYN_B)) != (MR_DYN_A | MR_DYN_B)) {
                R ta, tb, nv[3];
                if (staged) {
                    ta = T[k]; tb = T[K + k];
                    for (int r = 0; r < 3; r++) nv[r] = T[(size_t)(2 + r) * K + k];
                } else {
                    terms(k, ta, tb, nv);
                }
                for (int side = 0; side < 2; side++) {
                    if (r4.y & (side ? MR_DYN_B : MR_DYN_A)) continue;
                    const int64_t x = off + (side ? r4.w : r4.z);
                    const R t = side ? tb : ta;
                    for (int r = 0; r < 3; r++) {
                        const double p = BODY(d.pos, r, x);
                        const double qv = side ? p + (double)(t * nv[r]) : p - (double)(t * nv[r]);
                        if (__double_as_longlong(qv) != __double_as_longlong(p)) BODY(d.pos, r, x) = qv;
                    }
                }
            }
        }
        nact[m] = (uint8_t)na;
        if (na) {
            if (r4.y & MR_DYN_A) atomicAdd(&cur[r4.z], na);
            if (r4.y & MR_DYN_B) atomicAdd(&cur[r4.w], na);
        }
    }
    __syncthreads();
    const int32_t tot = block_exscan<BLOCK>(cur, n, s_scan);
    if (tid == 0) cur[n] = tot;
    __syncthreads();
    // fill in contact order: a block of manifolds at a time.  Every
    // wavefront ranks its lanes at once: a lane's peers (lanes naming its
    // body on either side) come from one ballot per bit of the body index,
    // its rank is their active contacts in the lanes before it (the counts
    // 1..4 bit-sliced into three ballots).  Then the wavefronts in turn
    // place their entries after the body's cursor and move it.
    int32_t nbits = 1;
    while ((1 << nbits) <= n) nbits++;
    const uint64_t lt = (1ull << lane) - 1ull;
    int4 nxt = tid < M ? mrec[tid] : make_int4(cl, 0, 0, 0);
    for (int32_t base = 0; base < M; base += BLOCK) {
        const int32_t m = base + tid;
        const int4 r4 = nxt;
        nxt = m + BLOCK < M ? mrec[m + BLOCK] : make_int4(cl, 0, 0, 0);   // the next block's record
        int32_t A = n, B = n, na = 0, k0 = 0, c = 0;
        if (m < M) {
            na = nact[m];
            k0 = r4.x - cl;
            c = r4.y & MR_COUNT;
            if (na && (r4.y & MR_DYN_A)) A = r4.z;
            if (na && (r4.y & MR_DYN_B)) B = r4.w;
        }
        uint64_t pAA = ~0ull, pAB = ~0ull, pBA = ~0ull, pBB = ~0ull;
        for (int b = 0; b < nbits; b++) {
            const bool xa = (A >> b) & 1, xb = (B >> b) & 1;
            const uint64_t ba = __ballot(xa), bb = __ballot(xb);
            pAA &= xa ? ba : ~ba; pAB &= xa ? bb : ~bb;
            pBA &= xb ? ba : ~ba; pBB &= xb ? bb : ~bb;
        }
        const uint64_t w1 = __ballot(na & 1), w2 = __ballot(na & 2), w4 = __ballot(na & 4);
        auto wsum = [&](uint64_t x) { return __popcll(x & w1) + 2 * __popcll(x & w2) + 4 * __popcll(x & w4); };
        const int32_t ra = wsum(pAA & lt) + wsum(pAB & lt), ta = wsum(pAA) + wsum(pAB);
        const int32_t rb = wsum(pBA & lt) + wsum(pBB & lt), tb = wsum(pBA) + wsum(pBB);
        const bool ha = A < n, hb = B < n;
        for (int w = 0; w < NW; w++) {
            if (wv == w) {
                const int32_t ca0 = ha ? cur[A] + ra : 0, cb0 = hb ? cur[B] + rb : 0;
                int32_t jj = 0;
                for (int32_t j = 0; j < c; j++) {
                    const int32_t k = k0 + j;
                    if (!active(k)) continue;
                    if (ha) ent[ca0 + jj] = k << 1;
                    if (hb) ent[cb0 + jj] = (k << 1) | 1;
                    jj++;
                }
                // the lane with a body's last entries in the wavefront moves
                // its cursor
                __builtin_amdgcn_wave_barrier();
                if (ha && ra + na == ta) cur[A] = ca0 + na;
                if (hb && rb + na == tb) cur[B] = cb0 + na;
            }
            __syncthreads();
        }
    }
    // cur[i] is now the end of body i's list, cur[i - 1] its start
    for (int32_t i = tid; i < n; i += BLOCK) {
        const int32_t e0 = i ? cur[i - 1] : 0, e1 = cur[i];
        if (e0 == e1) continue;
        const int64_t x = off + i;
        double p[3] = {BODY(d.pos, 0, x), BODY(d.pos, 1, x), BODY(d.pos, 2, x)};
        R acc[3] = {R(0), R(0), R(0)};
        // each entry's term and normal, the next entry's loaded before this
        // one is folded
        auto fetch = [&](int32_t e, R& t, R* nv, bool& sb) {
            const int32_t v = ent[e], k = v >> 1;
            sb = v & 1;
            if (staged) {
                t = sb ? T[K + k] : T[k];
                for (int r = 0; r < 3; r++) nv[r] = T[(size_t)(2 + r) * K + k];
            } else {
                R ta = R(0), tb = R(0);
                terms(k, ta, tb, nv);
                t = sb ? tb : ta;
            }
        };
        R tn, nn[3];
        bool sn;
        fetch(e0, tn, nn, sn);
        for (int32_t e = e0; e < e1; e++) {
            const R t = tn, nv[3] = {nn[0], nn[1], nn[2]};
            const bool sb = sn;
            if (e + 1 < e1) fetch(e + 1, tn, nn, sn);
            for (int r = 0; r < 3; r++) {
                const R term = t * nv[r];
                if (fp64) p[r] = sb ? p[r] + (double)term : p[r] - (double)term;
                else acc[r] = sb ? acc[r] + term : acc[r] - term;
            }
        }
        for (int r = 0; r < 3; r++) BODY(d.pos, r, x) = fp64 ? p[r] : p[r] + (double)acc[r];
    }
}

// ---------------------------------------------------------------- islands --

// The islands of one world (SequentialImpulseSolver.build_islands,
// solver.py:412-445): bodies joined by manifolds whose two sides are dynamic,
// each manifold in the island of its dynamic side.  Islands share no body
// whose velocity the solver changes, so sweeping each island's contacts in
// their order in the world gives the world's Gauss-Seidel result exactly.
// One block per world: union-find over the manifolds (the root is the
// smallest body index), manifolds and contacts per root, then one record per
// island in the class's island table and the island's manifolds in world
// order with island-local first contacts, and the batch index of each island
// contact.  Parent, manifold and contact counts per body are in LDS (a world
// solved island by island fits them; else overflow bit 6 is set).
template <int BLOCK>
__global__ __launch_bounds__(BLOCK) void k_islands(Dev d, const int32_t* wlist, const int32_t* wcount,
                                                   int32_t lds_bytes) {
    extern __shared__ __align__(16) unsigned char s_dyn[];
    __shared__ int32_t s_scan[BLOCK / 64 + 2];
    __shared__ int32_t s_base;
    if ((int32_t)blockIdx.x >= *wcount) return;
    const int32_t w = wlist[blockIdx.x];
    const int tid = threadIdx.x;
    if (w >= d.wcap || !(d.w_flags[w] & WF_ACTIVE)) return;
    if (tid == 0) d.w_levels[w] = 0;
    const int32_t K = d.w_cn[w], cs = d.w_cs[w], n = d.w_cnt[w], M = d.w_mn[w], ms = d.w_ms[w];
    if (K == 0 || M == 0) return;
    const size_t arr = al16((size_t)(n + 1) * 4);
    if (3 * arr > (size_t)lds_bytes) {
        if (tid == 0) atomicOr((unsigned long long*)&d.cnt[C_OVERFLOW], 64ull);
        return;
    }
    const int4* mrec = d.mrec + ms;
    int32_t* P = (int32_t*)s_dyn;                 // parent, then island index of each root
    int32_t* MC = (int32_t*)(s_dyn + arr);        // manifolds per root, then cursor
    int32_t* KC = (int32_t*)(s_dyn + 2 * arr);    // contacts per root, then cursor
    for (int32_t i = tid; i <= n; i += BLOCK) { P[i] = i; MC[i] = 0; KC[i] = 0; }
    __syncthreads();
    auto ld = [&](int32_t i) { return __hip_atomic_load(&P[i], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP); };
    // root with path halving (halving writes only move a body closer to its root)
    auto find = [&](int32_t x) {
        for (;;) {
            const int32_t p = ld(x);
            if (p == x) return x;
            const int32_t g = ld(p);
            if (g == p) return p;
            __hip_atomic_store(&P[x], g, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
            x = g;
        }
    };
    for (int32_t m = tid; m < M; m += BLOCK) {
        const int4 r = mrec[m];
        if ((r.y & (MR_DYN_A | MR_DYN_B)) != (MR_DYN_A | MR_DYN_B)) continue;
        int32_t x = r.z, y = r.w;
        for (;;) {
            x = find(x);
            y = find(y);
            if (x == y) break;
            if (x < y) { const int32_t t = x; x = y; y = t; }
            if (atomicCAS(&P[x], x, y) == x) break;
        }
    }
    __syncthreads();
    // each manifold's root: that of its dynamic side (of side a when neither
    // is dynamic: such a manifold moves no body and joins no island)
    for (int32_t m = tid; m < M; m += BLOCK) {
        const int4 r = mrec[m];
        const int32_t rt = find((r.y & MR_DYN_A) || !(r.y & MR_DYN_B) ? r.z : r.w);
        d.isl_lab[ms + m] = rt;
        atomicAdd(&MC[rt], 1);
        atomicAdd(&KC[rt], r.y & MR_COUNT);
    }
    __syncthreads();
    for (int32_t i = tid; i < n; i += BLOCK) P[i] = MC[i] > 0;
    __syncthreads();
    const int32_t NI = block_exscan<BLOCK>(P, n, s_scan);
    if (tid == 0) {
        P[n] = NI;
        s_base = (int32_t)atomicAdd((unsigned long long*)&d.cnt[C_ISLANDS], (unsigned long long)NI);
    }
    __syncthreads();
    const int32_t base = s_base;
    auto is_root = [&](int32_t i) { return P[i + 1] > P[i]; };
    for (int32_t i = tid; i < n; i += BLOCK) {
        if (!is_root(i)) continue;
        int32_t* ds = d.isl_desc + (size_t)(base + P[i]) * ISL_DESC;
        ds[0] = w; ds[1] = MC[i]; ds[2] = KC[i];
    }
    __syncthreads();
    block_exscan<BLOCK>(MC, n, s_scan);
    block_exscan<BLOCK>(KC, n, s_scan);
    for (int32_t i = tid; i < n; i += BLOCK) {
        if (!is_root(i)) continue;
        int32_t* ds = d.isl_desc + (size_t)(base + P[i]) * ISL_DESC;
        ds[3] = ms + MC[i]; ds[4] = cs + KC[i];
    }
    __syncthreads();
    // from here MC / KC are each island's cursors (manifold and contact
    // positions in the world, from its bases) and P its contact base
    for (int32_t i = tid; i < n; i += BLOCK) P[i] = KC[i];
    __syncthreads();
    // the island records of each root's manifolds, in world order: one
    // wavefront, 64 manifolds at a time (the next group's records load
    // meanwhile); a lane's peers (lanes with its root) from a ballot per bit
    // of the root, its rank and contact offset from the peers before it, the
    // cursors moved by the last lane of each root
    if (tid < 64) {
        const int lane = tid;
        const uint64_t lt = (1ull << lane) - 1ull;
        int32_t nbits = 1;
        while ((1 << nbits) <= n) nbits++;
        int4 nr = lane < M ? mrec[lane] : make_int4(0, 0, 0, 0);
        int32_t nl = lane < M ? d.isl_lab[ms + lane] : 0;
        for (int32_t g = 0; g < M; g += 64) {
            const int32_t m = g + lane;
            const bool valid = m < M;
            const int4 r = nr;
            const int32_t rt = nl;
            if (m + 64 < M) { nr = mrec[m + 64]; nl = d.isl_lab[ms + m + 64]; }
            const int32_t c = valid ? (r.y & MR_COUNT) : 0;
            uint64_t peers = __ballot(valid);
            for (int b = 0; b < nbits; b++) {
                const bool x = (rt >> b) & 1;
                const uint64_t bx = __ballot(x);
                peers &= x ? bx : ~bx;
            }
            const int32_t rank = __popcll(peers & lt);
            // contacts per manifold: MR_COUNT bits, three unless a count needs more
            const int cb = __ballot(c > 7) ? 8 : 3;
            int32_t krank = 0;
            for (int b = 0; b < cb; b++) krank += __popcll(peers & lt & __ballot((c >> b) & 1)) << b;
            if (valid) {
                const int32_t pos = MC[rt] + rank, kpos = KC[rt] + krank;
                const int32_t kl = kpos - P[rt];
                d.isl_mrec[ms + pos] = make_int4(kl, r.y, r.z, r.w);
                for (int32_t j = 0; j < c; j++) d.isl_cidx[cs + kpos + j] = r.x + j;
                if (!(peers >> lane >> 1)) { MC[rt] = pos + 1; KC[rt] = kpos + c; }
            }
            // LDS runs this wavefront's accesses in issue order
            asm volatile("" ::: "memory");
        }
    }
}

// One world's (or island's) solve.  With self_rows (a world, BLOCK > 64)
// the block also writes its contacts' solver rows: the threads past the first
// wavefront do it while that wavefront assigns the levels.
template <typename R, int BLOCK, bool ISL>
__device__ __forceinline__ void solve_unit(const Dev& d, const SolveIn& in, int32_t lds_bytes, int32_t w, int32_t K,
                                           int32_t M, const int4* mrec, int32_t cl, const int32_t* gmap, int64_t kb,
                                           bool self_rows = false) {
    constexpr bool isl = ISL;
    extern __shared__ __align__(16) unsigned char s_dyn[];
    __shared__ int32_t s_scan[BLOCK / 64 + 2];
    __shared__ int32_t s_D, s_D0, s_nd;
    __shared__ bool s_over;
    const int tid = threadIdx.x;
    const int32_t n = d.w_cnt[w];
    const int64_t off = d.w_off[w];
    auto gci = [&](int32_t k) -> int64_t { return gmap ? (int64_t)gmap[k] : (int64_t)cl + k; };
    const bool fp64 = sizeof(R) == sizeof(double);
    constexpr size_t RB = sizeof(R);
    // The launch's dynamic LDS: the last 4 * SOLVE_DCAP bytes hold the level
    // bounds; the rest holds the level table during the assignment, then the
    // level-loop state, each when it fits (global memory otherwise).
    const size_t avail = (size_t)lds_bytes - 4 * (size_t)SOLVE_DCAP;
    int32_t* s_bound = (int32_t*)(s_dyn + avail);
    R* S_ = (R*)d.csolv;
    const int32_t iters = d.w_iters[w];
    const int32_t S = in.warm + iters;          // velocity passes
    const int32_t SL = S > 0 ? S : 1;           // passes in the schedule (the position pass rides on pass 0)
    const int32_t NI = SL * M;                  // manifold items
    const int32_t NC = SL * K;                  // contact items
    const int64_t ib = kb * d.slmax;           // this unit's item slice
    int32_t* level = d.clevel + ib;
    int32_t* corder = d.corder + ib;
    int32_t* cord_a = d.cord_a + ib;
    int32_t* cord_b = d.cord_b + ib;
    int32_t* cord_c = d.cord_c + ib;
    int32_t* g_bound = d.lcount + ib;

    // ---- level assignment: table (body masks 8 B, last levels 4 B per body)
    // and, when they fit beside it, the manifold records, so the assignment
    // loop issues no global loads (on this GPU a global load's wait also waits
    // for the loop's earlier level stores)
    const bool table_lds = al16((size_t)n * 8) + (size_t)n * 4 <= avail;
    const size_t mrec_at = al16((size_t)n * 8) + al16((size_t)n * 4);
    const bool mrec_lds = table_lds && mrec_at + (size_t)M * 16 <= avail;
    // the previous pass's level per manifold, for shifting the last passes
    // (worlds of at least four groups of manifolds per pass)
    const size_t prev_at = al16(mrec_at + (size_t)M * 16);
    const bool prev_lds = mrec_lds && M >= 256 && prev_at + (size_t)M * 4 + 4 * (size_t)SOLVE_DCAP <= avail;
    // an island shares the world's body range with the world's other islands,
    // so its tables must be its block's own
    if (isl && !(table_lds && al16((size_t)n * 4) <= avail)) {
        if (tid == 0) atomicOr((unsigned long long*)&d.cnt[C_OVERFLOW], 64ull);
        return;
    }
    if (table_lds) {
        uint64_t* sm = (uint64_t*)s_dyn;
        int32_t* sl = (int32_t*)(s_dyn + al16((size_t)n * 8));
        for (int32_t i = tid; i < n; i += BLOCK) { sl[i] = 0; sm[i] = 0; }
        if (mrec_lds) {
            int4* lm = (int4*)(s_dyn + mrec_at);
            for (int32_t m = tid; m < M; m += BLOCK) lm[m] = mrec[m];
        }
    } else {
        for (int32_t i = tid; i < n; i += BLOCK) { d.last[off + i] = 0; d.lmask[off + i] = 0; }
    }
    for (int32_t l = tid; l < SOLVE_DCAP; l += BLOCK) s_bound[l] = 0;
    __syncthreads();
    if (tid < 64) {
        int32_t D, D0;
        bool over = false;
        if (table_lds) {
            uint64_t* sm = (uint64_t*)s_dyn;
            int32_t* sl = (int32_t*)(s_dyn + al16((size_t)n * 8));
            // two calls, so each knows its records' address space
            if (mrec_lds)
                D = assign_item_levels<false>(LevelTable<false>{sl, sm}, M, NI, (const int4*)(s_dyn + mrec_at), level,
                                              s_bound, SOLVE_DCAP, over, D0,
                                              prev_lds ? (int32_t*)(s_dyn + prev_at) : nullptr);
            else
                D = assign_item_levels<false>(LevelTable<false>{sl, sm}, M, NI, mrec, level, s_bound, SOLVE_DCAP,
                                              over, D0);
        } else {
            D = assign_item_levels<true>(LevelTable<true>{d.last + off, d.lmask + off}, M, NI, mrec, level, s_bound,
                                         SOLVE_DCAP, over, D0);
        }
        if (tid == 0) {
            s_D = D; s_D0 = D0; s_over = over;
            if (isl) atomicMax(&d.w_levels[w], D);
            else d.w_levels[w] = D;
            atomicMax((unsigned long long*)&d.cnt[C_MAXLEVEL], (unsigned long long)D);
        }
    } else if (!ISL && BLOCK > 64 && self_rows) {
        const int32_t old = *d.sel;
        if (in.wake && tid == 64) wake_seq(d, w);
        for (int32_t k = tid - 64; k < K; k += BLOCK - 64) {
            const int64_t ci = (int64_t)cl + k;
            const int64_t a = d.ca[ci], b = d.cb[ci];
            if (in.wake) wake_pair(d, w, a, b);
            prepare_contact<R, false>(d, in, ci, a, b, w, cl, old);
        }
    }
    __syncthreads();
    const int32_t D = s_D, D0 = s_D0;
    const bool lds_bounds = !s_over;

    // ---- compact ids of the bodies in contact (the table is no longer needed)
    // (-1: not in contact).  The first lane to name a body claims it and
    // takes the next id, one counter update per wavefront; which body gets
    // which id does not change any result.
    const bool map_lds = al16((size_t)n * 4) <= avail;
    int32_t* cmap = map_lds ? (int32_t*)s_dyn : d.last + off;
    for (int32_t i = tid; i < n; i += BLOCK) cmap[i] = -1;
    if (tid == 0) s_nd = 0;
    __syncthreads();
    {
        const int lane = tid & 63;
        const uint64_t lt = (1ull << lane) - 1ull;
        for (int32_t base = 0; base < M; base += BLOCK) {
            const int32_t m = base + tid;
            const bool v = m < M;
            const int4 r = v ? mrec[m] : make_int4(0, 0, 0, 0);
            for (int side = 0; side < 2; side++) {
                const int32_t x = side ? r.w : r.z;
                const bool own = v && atomicCAS(&cmap[x], -1, -2) == -1;
                const uint64_t bal = __ballot(own);
                int32_t wb = 0;
                if (lane == 0 && bal) wb = atomicAdd(&s_nd, (int32_t)__popcll(bal));
                wb = __shfl(wb, 0, 64);
                if (own) cmap[x] = wb + (int32_t)__popcll(bal & lt);
            }
        }
    }
    __syncthreads();
    const int32_t nd = s_nd;
    auto touched = [&](int32_t i) { return cmap[i] >= 0; };
    SolveState<R> st;
    st.map = 0;
    st.vel = al16((size_t)n * 4);
    st.imp = st.vel + al16((size_t)nd * VelRec<R>::stride * RB);
    st.pos = st.imp + al16((size_t)K * 3 * RB);
    st.nd = nd;
    st.off = off;
    // velocities and impulses (positions come later, in the same space)
    const bool state_lds = map_lds && st.pos <= avail;
    st.pos_lds = false;
    // global velocity state (a world whose state does not fit the LDS): FP64
    // one 48-byte record per body (a body's six components in one or two
    // lines rather than six), FP32 component-major
    st.V = (R*)d.svel;
    st.vs = fp64 ? 1 : d.ncap;
    st.vi = fp64 ? 6 : 1;
    st.P = S_ + RowLayout<R>::imp; st.ps = RowLayout<R>::stride; st.pc = 1; st.pb = 0;

    // ---- contacts per level -> level starts -> items scattered by level
    // (before the velocities are staged: the manifold records' LDS copy made
    // for the assignment is still intact)
    const int4* mrec_b = mrec_lds ? (const int4*)(s_dyn + mrec_at) : mrec;
    if (lds_bounds) {
        block_exscan<BLOCK>(s_bound, D, s_scan);
        scatter_items<BLOCK>(s_bound, SL, M, mrec_b, level, cl, gmap, off, corder, cord_a, cord_b, cord_c);
    } else {
        for (int32_t l = tid; l < D; l += BLOCK) g_bound[l] = 0;
        __syncthreads();
        count_items<BLOCK>(g_bound, SL, M, mrec_b, level);
        __syncthreads();
        block_exscan<BLOCK>(g_bound, D, s_scan);
        scatter_items<BLOCK>(g_bound, SL, M, mrec_b, level, cl, gmap, off, corder, cord_a, cord_b, cord_c);
    }
    __syncthreads();

    // ---- stage velocities, impulses
    if (state_lds) {
        R* V = (R*)(s_dyn + st.vel);
        for (int32_t i = tid; i < n; i += BLOCK) {
            if (!touched(i)) continue;
            const int32_t c = cmap[i];
            for (int q = 0; q < 6; q++) V[(size_t)c * VelRec<R>::stride + q] = (R)BODY(d.lv, q, off + i);
        }
        R* I = (R*)(s_dyn + st.imp);
        for (int32_t k = tid; k < K; k += BLOCK)
            for (int q = 0; q < 3; q++) I[(size_t)k * 3 + q] = ROW(S_, gci(k), RowLayout<R>::imp + q);
    } else if (!fp64) {
        R* V = (R*)d.svel;
        for (int32_t i = tid; i < n; i += BLOCK) {
            if (!touched(i)) continue;
            for (int q = 0; q < 6; q++) V[(size_t)q * d.ncap + off + i] = (R)BODY(d.lv, q, off + i);
        }
    } else {
        R* V = (R*)d.svel;
        for (int32_t i = tid; i < n; i += BLOCK) {
            if (!touched(i)) continue;
            for (int q = 0; q < 6; q++) V[(size_t)(off + i) * 6 + q] = BODY(d.lv, q, off + i);
        }
    }
    __syncthreads();
    // bound[l] is now the end of level l
    if (S > 0) {
        if (state_lds)
            run_levels_lb<R, BLOCK, true, false, ISL>(lds_bounds, d, in, st, avail, K, D, NC, S, g_bound, corder,
                                                      cord_a, cord_b, cord_c, cl);
        else
            run_levels_lb<R, BLOCK, false, false, ISL>(lds_bounds, d, in, st, avail, K, D, NC, S, g_bound, corder,
                                                       cord_a, cord_b, cord_c, cl);
    }
    __syncthreads();

    // ---- results back: impulses to the contacts (warm-start cache and
    // readback); velocities to the FP64 state (copied in FP64, the FP32
    // solver's change folded in for FP32)
    R* imp = (R*)d.cimp;
    for (int32_t k = tid; k < K; k += BLOCK) {
        const int64_t ci = gci(k);
        for (int r = 0; r < 3; r++)
            CON(imp, r, ci) = state_lds ? ((R*)(s_dyn + st.imp))[(size_t)k * 3 + r] : ROW(S_, ci, RowLayout<R>::imp + r);
    }
    for (int32_t i = tid; i < n; i += BLOCK) {
        if (d.mt[off + i] != M_DYNAMIC) continue;
        if (state_lds) {
            if (!touched(i)) continue;
            const int32_t c = cmap[i];
            const R* V = (const R*)(s_dyn + st.vel);
            for (int q = 0; q < 6; q++) {
                if (fp64) {
                    BODY(d.lv, q, off + i) = (double)V[(size_t)c * VelRec<R>::stride + q];
                } else {
                    const double v0 = BODY(d.lv, q, off + i);
                    const R dv = V[(size_t)c * VelRec<R>::stride + q] - (R)v0;
                    BODY(d.lv, q, off + i) = v0 + (double)dv;
                }
            }
        } else if (!fp64) {
            if (!touched(i)) continue;
            for (int q = 0; q < 6; q++) {
                const double v0 = BODY(d.lv, q, off + i);
                const R dv = ((const R*)d.svel)[(size_t)q * d.ncap + off + i] - (R)v0;
                BODY(d.lv, q, off + i) = v0 + (double)dv;
            }
        } else {
            if (!touched(i)) continue;
            for (int q = 0; q < 6; q++) BODY(d.lv, q, off + i) = ((const R*)d.svel)[(size_t)(off + i) * 6 + q];
        }
    }
    // ---- the split-impulse position pass: per body, in the LDS the velocity
    // state used, when its lists fit; else over the levels holding pass-0
    // items, positions of the bodies in contact in LDS when they fit (FP64
    // holds the positions themselves, FP32 the corrections from zero)
    if (D0 > 0) {
        __syncthreads();
        if (split_pos_bytes(K, M, n, 0) <= avail) {
            split_positions<R, BLOCK>(d, in, cl, gmap, K, M, n, off, mrec, s_dyn,
                                      split_pos_bytes(K, M, n, RB) <= avail, s_scan);
        } else {
            st.pos = st.vel;
            st.pos_lds = map_lds && st.pos + al16((size_t)nd * 3 * RB) <= avail;
            if (st.pos_lds) {
                R* P = (R*)(s_dyn + st.pos);
                for (int32_t i = tid; i < n; i += BLOCK) {
                    if (!touched(i)) continue;
                    const int32_t c = cmap[i];
                    for (int q = 0; q < 3; q++) P[(size_t)q * nd + c] = fp64 ? (R)BODY(d.pos, q, off + i) : R(0);
                }
                __syncthreads();
                run_levels_lb<R, BLOCK, true, true, ISL>(lds_bounds, d, in, st, avail, K, D0, NC, S, g_bound, corder,
                                                         cord_a, cord_b, cord_c, cl);
                __syncthreads();
                for (int32_t i = tid; i < n; i += BLOCK) {
                    if (!touched(i)) continue;
                    const int32_t c = cmap[i];
                    for (int q = 0; q < 3; q++) {
                        if (fp64 && d.mt[off + i] == M_DYNAMIC) {
                            BODY(d.pos, q, off + i) = (double)P[(size_t)q * nd + c];
                        } else if (fp64) {
                            // a static or kinematic side: only -0.0 -> +0.0 can
                            // change, written only if it did (another island
                            // may share the body)
                            const double p0 = BODY(d.pos, q, off + i), p1 = (double)P[(size_t)q * nd + c];
                            if (__double_as_longlong(p1) != __double_as_longlong(p0)) BODY(d.pos, q, off + i) = p1;
                        } else if (d.mt[off + i] == M_DYNAMIC)
                            BODY(d.pos, q, off + i) = BODY(d.pos, q, off + i) + (double)P[(size_t)q * nd + c];
                    }
                }
            } else {
                run_levels_lb<R, BLOCK, false, true, ISL>(lds_bounds, d, in, st, avail, K, D0, NC, S, g_bound, corder,
                                                          cord_a, cord_b, cord_c, cl);
            }
        }
    }
}


// The solver of a size class: one world per block, or (ISL) the class's
// islands, each block taking islands in turn from the island table.  Two
// instantiations, so the per-world kernel holds no register for the islands.
template <typename R, int BLOCK, bool ISL>
__global__ __launch_bounds__(BLOCK) void k_solve(Dev d, SolveIn in, const int32_t* wlist, const int32_t* wcount,
                                                 int32_t lds_bytes, int32_t self_rows) {
    if constexpr (!ISL) {
        if ((int32_t)blockIdx.x >= *wcount) return;
        const int32_t w = wlist[blockIdx.x];
        if (w >= d.wcap || !(d.w_flags[w] & WF_ACTIVE)) return;
        const int32_t K = d.w_cn[w], cs = d.w_cs[w];
        if (K == 0) { if (threadIdx.x == 0) d.w_levels[w] = 0; return; }
        solve_unit<R, BLOCK, false>(d, in, lds_bytes, w, K, d.w_mn[w], d.mrec + d.w_ms[w], cs, nullptr, cs,
                                    self_rows != 0);
    } else {
    const uint64_t ni = d.cnt[C_ISLANDS];
    const int32_t NIs = (int32_t)(ni < (uint64_t)d.kcap ? ni : (uint64_t)d.kcap);
    for (int32_t ii = blockIdx.x; ii < NIs; ii += gridDim.x) {
        const int32_t* ds = d.isl_desc + (size_t)ii * ISL_DESC;
        const int32_t w = ds[0], M = ds[1], K = ds[2], mb = ds[3], kbase = ds[4];
        solve_unit<R, BLOCK, true>(d, in, lds_bytes, w, K, M, d.isl_mrec + mb, 0, d.isl_cidx + kbase, kbase);
        __syncthreads();
    }
    }
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

// 3x3 LU factorisation with partial pivoting in LAPACK dgetf2 order; with
// lu3_solve (dgetrs order) it is np.linalg.solve in PointConstraint.solve
// (joints.py:119).  K does not change between the solve sweeps of a step, so
// the factors are computed once per step and give the same result as
// factoring K at every call.
template <typename R>
__device__ __forceinline__ void lu3(const R* K, R* LU, int* piv) {
    R A[3][3];
    for (int r = 0; r < 3; r++) for (int c = 0; c < 3; c++) A[r][c] = K[3 * r + c];
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
    for (int r = 0; r < 3; r++) for (int c = 0; c < 3; c++) LU[3 * r + c] = A[r][c];
}

template <typename R>
__device__ __forceinline__ void lu3_solve(const R* LU, int pv, const R* rhs, R* x) {
    const int piv[3] = {pv & 3, (pv >> 2) & 3, 2};
    R b[3] = {rhs[0], rhs[1], rhs[2]};
    for (int j = 0; j < 2; j++) if (piv[j] != j) { R t = b[j]; b[j] = b[piv[j]]; b[piv[j]] = t; }
    for (int i = 1; i < 3; i++) for (int k = 0; k < i; k++) b[i] -= LU[3 * i + k] * b[k];
    for (int i = 2; i >= 0; i--) {
        for (int k = i + 1; k < 3; k++) b[i] -= LU[3 * i + k] * b[k];
        b[i] /= LU[3 * i + i];
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

// One joint's solver row, stored at its position in the level order:
// LU factors of K, the pivots (p0 + 4 p1), rA, rB, bias / dt.
enum { JR_LU = 0, JR_PIV = 9, JR_RA = 10, JR_RB = 13, JR_BDT = 16, JR_N = 19 };

template <typename R>
struct alignas(16) JointItem {
    R v[JR_N];
    int32_t ea, eb;   // end records: -1 = the world, else (index << 1) | dynamic
};
static_assert(sizeof(JointItem<double>) == 160 && sizeof(JointItem<float>) == 96, "joint row sizes the host allocates");

// Per body a joint solve stages in LDS: a record of six FP64 velocities,
// the world inverse inertia (9 R) and inverse mass (R), padded to 16 bytes,
// and the body's world-local index.
template <typename R>
struct JointBody {
    static constexpr int vel = 48;
    static constexpr int rec = 48 + (10 * (int)sizeof(R) + 15) / 16 * 16;
    static constexpr int lds = rec + 4;
};
constexpr int JOINT_LEVELS = 1024;

// PhysicsWorld._solve_joints (world.py:599-611): wake the joint bodies,
// PointConstraint.prepare every joint (joints.py:84-101), then
// joint_iterations Gauss-Seidel sweeps of PointConstraint.solve
// (joints.py:111-121) in joint order, run level by level like the contacts.
// The level schedule depends only on the joint list and the motion types,
// which are fixed while the world is loaded: it is built on the world's first
// step and kept (w_jsched = 2 levels + LDS form, 0 = none).  The velocities of
// the bodies the joints touch live in LDS during the sweeps when they fit in
// lds_bytes and the schedule has at most JOINT_LEVELS levels (else the sweeps
// update the body arrays); each thread loads its item of the next level while
// it solves the current one.
template <typename R, int BLOCK>
__global__ __launch_bounds__(BLOCK) void k_joints(Dev d, double dt, int32_t lds_bytes) {
    extern __shared__ __align__(16) unsigned char s_dyn[];
    __shared__ int32_t s_scan[BLOCK / 64 + 2];
    __shared__ int32_t s_D, s_ns;
    __shared__ int32_t s_lend[JOINT_LEVELS];
    const int32_t w = blockIdx.x;
    const int tid = threadIdx.x;
    if (w >= d.wcap || (d.w_flags[w] & (WF_ACTIVE | WF_JOINTS)) != (WF_ACTIVE | WF_JOINTS)) return;
    const int32_t J = d.w_jn[w], js = d.w_js[w], n = d.w_cnt[w];
    const int64_t off = d.w_off[w];
    if (J == 0) return;
    const bool fp64 = sizeof(R) == sizeof(double);
    JointItem<R>* rows = (JointItem<R>*)d.jrow;
    const R* iiw = (const R*)d.iiw;
    using JB = JointBody<R>;
    const int32_t cap = lds_bytes / JB::lds;
    unsigned char* s_rec = s_dyn;
    int32_t* s_body = (int32_t*)(s_dyn + (size_t)cap * JB::rec);
    int32_t sched = d.w_jsched[w];
    {
        const int32_t ns0 = d.w_jns[w], D0 = sched >> 1;
        if (sched != 0 && (sched & 1) != (ns0 <= cap && D0 <= JOINT_LEVELS)) sched = 0;   // LDS budget changed
    }
    if (sched == 0) {
        // ---- build the schedule: levels, joint order, the joint bodies
        int32_t* last = d.last + off;
        for (int32_t i = tid; i < n; i += BLOCK) last[i] = 0;
        if (tid == 0) s_ns = 0;
        __syncthreads();
        auto get = [&](int32_t t, int32_t& al, int32_t& bl) {
            const int32_t a = d.ja[js + t], b = d.jb[js + t];
            al = (a >= 0 && d.mt[a] == M_DYNAMIC) ? (int32_t)(a - off) : -1;
            bl = (b >= 0 && d.mt[b] == M_DYNAMIC) ? (int32_t)(b - off) : -1;
        };
        const int32_t D = assign_levels<BLOCK>(J, get, last, d.jlevel + js, d.jcount + js, d.jorder + js, s_scan, &s_D);
        // number the bodies the joints touch (static and kinematic ends
        // included: their velocities are read)
        for (int32_t i = tid; i < n; i += BLOCK) last[i] = -1;
        __syncthreads();
        for (int32_t t = tid; t < J; t += BLOCK) {
            const int32_t a = d.ja[js + t], b = d.jb[js + t];
            if (a >= 0) last[a - off] = 0;
            if (b >= 0) last[b - off] = 0;
        }
        __syncthreads();
        for (int32_t i = tid; i < n; i += BLOCK)
            if (last[i] == 0) {
                const int32_t s2 = atomicAdd(&s_ns, 1);
                last[i] = s2;
                d.jbody[2 * (int64_t)js + s2] = i;
            }
        __syncthreads();
        const int32_t ns = s_ns;
        const bool lds = ns <= cap && D <= JOINT_LEVELS;
        auto end = [&](int32_t i) -> int32_t {
            if (i < 0) return -1;
            const int32_t x = lds ? last[i - off] : i;
            return (x << 1) | (d.mt[i] == M_DYNAMIC ? 1 : 0);
        };
        for (int32_t t = tid; t < J; t += BLOCK) {
            const int64_t j = js + d.jorder[js + t];
            d.jend[js + t] = make_int2(end(d.ja[j]), end(d.jb[j]));
        }
        sched = 2 * D + (lds ? 1 : 0);
        if (tid == 0) { d.w_jsched[w] = sched; d.w_jns[w] = ns; }
        __syncthreads();
    }
    const int32_t D = sched >> 1, ns = d.w_jns[w];
    const bool in_lds = sched & 1;
    if (in_lds)
        for (int32_t s = tid; s < ns; s += BLOCK) {
            const int32_t bl = d.jbody[2 * (int64_t)js + s];
            s_body[s] = bl;
            const int64_t i = off + bl;
            unsigned char* rec = s_rec + (size_t)s * JB::rec;
            double2* v = (double2*)rec;
            v[0] = make_double2(BODY(d.lv, 0, i), BODY(d.lv, 1, i));
            v[1] = make_double2(BODY(d.lv, 2, i), BODY(d.lv, 3, i));
            v[2] = make_double2(BODY(d.lv, 4, i), BODY(d.lv, 5, i));
            R* I = (R*)(rec + JB::vel);
            for (int r = 0; r < 9; r++) I[r] = IIW(iiw, i, r);
            I[9] = (R)d.inv_mass[i];
        }
    const R rdt = (R)dt;
    for (int32_t t = tid; t < J; t += BLOCK) {
        const int32_t jo = d.jorder[js + t];
        const int64_t j = js + jo, p = js + t;
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
            for (int r = 0; r < 9; r++) I[r] = IIW(iiw, i, r);
            double q[4] = {BODY(d.quat, 0, i), BODY(d.quat, 1, i), BODY(d.quat, 2, i), BODY(d.quat, 3, i)}, Rm[9];
            quat_to_matrix(q, Rm);
            const double pp[3] = {BODY(d.pos, 0, i), BODY(d.pos, 1, i), BODY(d.pos, 2, i)};
            if (!(has & (1 << s2))) {
                const double dd[3] = {d.janc[j] - pp[0], d.janc[d.jcap + j] - pp[1], d.janc[2 * d.jcap + j] - pp[2]};
                for (int c = 0; c < 3; c++)
                    loc[(size_t)c * d.jcap + j] = (Rm[c] * dd[0] + Rm[3 + c] * dd[1]) + Rm[6 + c] * dd[2];
                has |= (uint8_t)(1 << s2);
            }
            const double l[3] = {loc[j], loc[d.jcap + j], loc[2 * d.jcap + j]};
            for (int r = 0; r < 3; r++) rr[r] = (pp[r] + ((Rm[3 * r] * l[0] + Rm[3 * r + 1] * l[1]) + Rm[3 * r + 2] * l[2])) - pp[r];
        }
        d.jhas[j] = has;
        for (int r = 0; r < 3; r++) { rA[r] = (R)rAd[r]; rB[r] = (R)rBd[r]; }
        const R ima = a < 0 ? R(0) : (R)d.inv_mass[a], imb = b < 0 ? R(0) : (R)d.inv_mass[b];
        R SA[9], SB[9], K[9];
        skew_sandwich(rA, Ia, SA);
        skew_sandwich(rB, Ib, SB);
        for (int r = 0; r < 3; r++)
            for (int c = 0; c < 3; c++) {
                R k = (ima + imb) * (r == c ? R(1) : R(0));
                k -= SA[3 * r + c];
                k -= SB[3 * r + c];
                K[3 * r + c] = k;
            }
        int piv[3];
        JointItem<R> it;
        lu3(K, it.v + JR_LU, piv);
        it.v[JR_PIV] = (R)(piv[0] + 4 * piv[1]);
        for (int r = 0; r < 3; r++) {
            const double anc = d.janc[(size_t)r * d.jcap + j];
            const double pa = (a < 0 ? anc : BODY(d.pos, r, a)) + rAd[r];
            const double pb = (b < 0 ? anc : BODY(d.pos, r, b)) + rBd[r];
            it.v[JR_RA + r] = rA[r]; it.v[JR_RB + r] = rB[r];
            it.v[JR_BDT + r] = (R)(0.2 * (pb - pa)) / rdt;
        }
        const int2 ends = d.jend[p];
        it.ea = ends.x; it.eb = ends.y;
        rows[p] = it;
    }
    __syncthreads();
    auto load_item = [&](int64_t p, JointItem<R>& it) { it = rows[p]; };
    // LDS and body-array forms of the sweeps are compiled apart (L), so that
    // the LDS form never waits for the item loads in flight.
    auto vel_get = [&](auto L, int32_t e, double* v) {
        const int32_t x = e >> 1;
        if constexpr (decltype(L)::value) {
            const double2* q = (const double2*)(s_rec + (size_t)x * JB::rec);
            const double2 q0 = q[0], q1 = q[1], q2 = q[2];
            v[0] = q0.x; v[1] = q0.y; v[2] = q1.x; v[3] = q1.y; v[4] = q2.x; v[5] = q2.y;
        } else {
            for (int c = 0; c < 6; c++) v[c] = BODY(d.lv, c, x);
        }
    };
    auto vel_put = [&](auto L, int32_t e, const double* v) {
        const int32_t x = e >> 1;
        if constexpr (decltype(L)::value) {
            double2* q = (double2*)(s_rec + (size_t)x * JB::rec);
            q[0] = make_double2(v[0], v[1]); q[1] = make_double2(v[2], v[3]); q[2] = make_double2(v[4], v[5]);
        } else {
            for (int c = 0; c < 6; c++) BODY(d.lv, c, x) = v[c];
        }
    };
    // Two lanes per joint: lane 2q + h takes end h (0: a, 1: b) of joint q of
    // the pass — its velocity, its anchor's point velocity and its share of
    // the impulse — and both lanes solve the joint's 3x3 system from the same
    // right-hand side.  A joint whose two ends are one body keeps only end
    // b's update, as the sequential loop's second write does.
    const int32_t half = tid & 1, jq = tid >> 1;
    constexpr int32_t JPB = BLOCK / 2;   // joints per pass of a level
    auto solve_item = [&](auto L, const JointItem<R>& it) {
        const int32_t e = half ? it.eb : it.ea;
        const R* rr = it.v + (half ? JR_RB : JR_RA);
        double wv[6] = {0, 0, 0, 0, 0, 0};
        R v[6], rhs[3], P[3], pv[3] = {R(0), R(0), R(0)}, po[3];
        if (e >= 0) vel_get(L, e, wv);
        for (int c = 0; c < 6; c++) v[c] = (R)wv[c];
        if (e >= 0) {
            pv[0] = v[0] + (v[4] * rr[2] - v[5] * rr[1]);
            pv[1] = v[1] + (v[5] * rr[0] - v[3] * rr[2]);
            pv[2] = v[2] + (v[3] * rr[1] - v[4] * rr[0]);
        }
        for (int r = 0; r < 3; r++) po[r] = __shfl_xor(pv[r], 1, 64);
        const R* pva = half ? po : pv;
        const R* pvb = half ? pv : po;
        for (int r = 0; r < 3; r++) rhs[r] = -((pvb[r] - pva[r]) + it.v[JR_BDT + r]);
        lu3_solve(it.v + JR_LU, (int)it.v[JR_PIV], rhs, P);
        const bool same = it.ea >= 0 && it.eb >= 0 && (it.ea >> 1) == (it.eb >> 1);
        if (e < 0 || !(e & 1) || (same && !half)) return;
        R I[10];   // world inverse inertia, inverse mass
        if constexpr (decltype(L)::value) {
            const R* q = (const R*)(s_rec + (size_t)(e >> 1) * JB::rec + JB::vel);
            for (int r = 0; r < 10; r++) I[r] = q[r];
        } else {
            for (int r = 0; r < 9; r++) I[r] = IIW(iiw, e >> 1, r);
            I[9] = (R)d.inv_mass[e >> 1];
        }
        const R Q[3] = {half ? P[0] : -P[0], half ? P[1] : -P[1], half ? P[2] : -P[2]};
        const R im = I[9];
        R nv[6];
        for (int r = 0; r < 3; r++) nv[r] = v[r] + im * Q[r];
        const R cx = rr[1] * Q[2] - rr[2] * Q[1], cy = rr[2] * Q[0] - rr[0] * Q[2], cz = rr[0] * Q[1] - rr[1] * Q[0];
        for (int r = 0; r < 3; r++) nv[3 + r] = v[3 + r] + ((I[3 * r] * cx + I[3 * r + 1] * cy) + I[3 * r + 2] * cz);
        for (int c = 0; c < 6; c++) wv[c] = fp64 ? (double)nv[c] : wv[c] + (double)(nv[c] - v[c]);
        vel_put(L, e, wv);
    };
    const int32_t iters = d.w_jiters[w];
    // end of each level in the order (an LDS copy in the LDS form)
    if (in_lds)
        for (int32_t l = tid; l < D; l += BLOCK) s_lend[l] = d.jcount[js + l];
    __syncthreads();
    // One step per (sweep, level).  A step loads this lane pair's item of the
    // next step while it solves its item of this one; with the velocities in
    // LDS the barrier between steps orders LDS only, so those loads stay in
    // flight across it.  Unrolled by two so the item buffers swap roles
    // instead of being copied (a copy would wait for the loads).
    const int32_t G = iters * D;
    auto jstep = [&](auto L, int32_t l, bool more, JointItem<R>& cur, JointItem<R>& nxt) {
        auto lend = [&](int32_t k) -> int32_t {
            if constexpr (decltype(L)::value) return s_lend[k];
            else return d.jcount[js + k];
        };
        const int32_t bg = l == 0 ? 0 : lend(l - 1), e = lend(l);
        if (more) {
            const int32_t l2 = l + 1 == D ? 0 : l + 1;
            const int32_t bg2 = l2 == 0 ? 0 : lend(l2 - 1), e2 = lend(l2);
            if (bg2 + jq < e2) load_item(js + bg2 + jq, nxt);
        }
        if (bg + jq < e) solve_item(L, cur);
        for (int32_t t = bg + jq + JPB; t < e; t += JPB) {
            JointItem<R> x;
            load_item(js + t, x);
            solve_item(L, x);
        }
        if constexpr (decltype(L)::value) {
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
            __builtin_amdgcn_s_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
        } else {
            __syncthreads();
        }
    };
    auto sweeps = [&](auto L) {
        JointItem<R> itA, itB;
        if (G > 0 && jq < d.jcount[js]) load_item(js + jq, itA);
        int32_t l = 0;
        for (int32_t g = 0; g < G; g += 2) {
            jstep(L, l, g + 1 < G, itA, itB);
            l = l + 1 == D ? 0 : l + 1;
            if (g + 1 < G) {
                jstep(L, l, g + 2 < G, itB, itA);
                l = l + 1 == D ? 0 : l + 1;
            }
        }
    };
    if (in_lds) sweeps(std::true_type{});
    else sweeps(std::false_type{});
    __syncthreads();
    if (in_lds)
        for (int32_t s = tid; s < ns; s += BLOCK) {
            const int64_t i = off + s_body[s];
            if (d.mt[i] != M_DYNAMIC) continue;
            const double* v = (const double*)(s_rec + (size_t)s * JB::rec);
            for (int c = 0; c < 6; c++) BODY(d.lv, c, i) = v[c];
        }
}

// ------------------------------------------------------- warm-start cache --

// Empty the slots of the buffer this step writes: each cache entry of the
// step that last filled it (two steps ago) recorded its slot, or -1.
__global__ void k_hash_clear(Dev d) {
    const int32_t nw = *d.sel ^ 1;
    const int32_t n = d.hslot_n[nw];
    for (int64_t i = gtid(); i < n; i += gstride()) {
        const int32_t h = d.hslot[nw][i];
        if (h >= 0) d.hkey[nw][h] = ~0ull;
    }
}

// Where each world's cache entries go this step: its contacts, or — when it
// has none — last step's entries, placed by an exclusive scan over the worlds
// in slot order; a world whose entries pass the capacity keeps none (and the
// overflow is reported).  C_CACHE gets the total asked for, at most twice
// the capacity (this step's contact slices and last step's entries each fit
// in it), so the int32 scan cannot overflow.
template <int BLOCK>
__global__ __launch_bounds__(BLOCK) void k_cache_plan(Dev d) {
    __shared__ int32_t s_scan[BLOCK / 64 + 2];
    const int32_t old = *d.sel, nw = old ^ 1;
    int32_t* start = d.w_cache_s + (size_t)nw * d.wcap;
    int32_t* count = d.w_cache_n + (size_t)nw * d.wcap;
    for (int32_t w = threadIdx.x; w < d.wcap; w += BLOCK) {
        int32_t c = 0;
        if (d.w_flags[w] & WF_ACTIVE) {
            const int32_t K = d.w_cn[w];
            c = K > 0 ? K : d.w_cache_n[(size_t)old * d.wcap + w];
        }
        start[w] = c;
        count[w] = c;
    }
    __syncthreads();
    const int32_t tot = block_exscan<BLOCK>(start, d.wcap, s_scan);
    bool over = false;
    for (int32_t w = threadIdx.x; w < d.wcap; w += BLOCK)
        if ((int64_t)start[w] + count[w] > d.kcap) { count[w] = 0; over = true; }
    if (over) atomicOr((unsigned long long*)&d.cnt[C_OVERFLOW], 4ull);
    if (threadIdx.x == 0) d.cnt[C_CACHE] = (uint64_t)tot;
}

// Insert the head of the group of cache entries at t of buffer `buf` (the
// first entry of a run of one (a, b)) into that buffer's hash; returns the
// hash slot it newly filled, or -1.  A key already present (a stale entry of
// an earlier fill, which lookups reject by range) is repointed.  No slot within
// HASH_PROBES probes is a cache overflow: the group then starts cold, as
// lookups give up after as many probes.
__device__ __forceinline__ int32_t cache_insert(const Dev& d, int buf, int64_t t, int32_t a, int32_t b) {
    const uint64_t key = ((uint64_t)(uint32_t)a << 32) | (uint32_t)b;
    const int64_t hm = d.hcap - 1;
    int64_t h = (int64_t)(hash_mix(key) & (uint64_t)hm);
    for (int probe = 0; probe < HASH_PROBES; probe++) {
        unsigned long long prev = atomicCAS((unsigned long long*)&d.hkey[buf][h], ~0ull, (unsigned long long)key);
        if (prev == ~0ull || prev == key) {
            d.hval[buf][h] = (int32_t)t;
            return prev == ~0ull ? (int32_t)h : -1;
        }
        h = (h + 1) & hm;
    }
    atomicOr((unsigned long long*)&d.cnt[C_OVERFLOW], 4ull);
    return -1;
}

// The cache after this step (solver.py:271-278): this step's contacts with
// their impulses, or — when the world had no contacts — last step's cache
// unchanged (solver.py:197-198).  One thread per entry over the whole batch:
// this step's contacts of every world that has contacts, then last step's
// entries of every world that has none, each copied to the place k_cache_plan
// gave its world (an entry is its world's when it lies in the world's source
// range, which also rejects entries left by removed worlds).  Group heads go
// into the hash; the order of the inserts only decides which probe slot a key
// takes, and lookups probe until they find the key.
template <typename R>
__global__ __launch_bounds__(256) void k_cache_all(Dev d) {
    const int32_t old = *d.sel, nw = old ^ 1;
    const uint64_t tc = d.cnt[C_CONTACTS];
    const int64_t T = tc < (uint64_t)d.kcap ? (int64_t)tc : d.kcap;
    const int64_t To = d.hslot_n[old];
    R* di = (R*)d.kimp[nw];
    for (int64_t e = gtid(); e < T + To; e += gstride()) {
        const bool cur = e < T;
        const int64_t s = cur ? e : e - T;
        const int32_t* sa = cur ? d.ca : d.ka[old];
        const int32_t* sb = cur ? d.cb : d.kb[old];
        const int32_t a = sa[s];
        if (a < 0 || a >= d.ncap) continue;
        const int32_t w = d.world[a];
        if (w < 0 || !(d.w_flags[w] & WF_ACTIVE)) continue;
        const size_t os = (size_t)old * d.wcap + w, ns = (size_t)nw * d.wcap + w;
        const int32_t K = d.w_cn[w];
        if (cur != (K > 0)) continue;
        const int32_t cnt = cur ? K : d.w_cache_n[os];
        const int64_t src = cur ? d.w_cs[w] : d.w_cache_s[os];
        if (s < src || s >= src + cnt || d.w_cache_n[ns] != cnt) continue;
        const int64_t k = s - src, t = d.w_cache_s[ns] + k;
        const int32_t b = sb[s];
        const double* sp = cur ? d.cpt : d.kpt[old];
        const R* si = (const R*)(cur ? d.cimp : d.kimp[old]);
        d.ka[nw][t] = a; d.kb[nw][t] = b;
        for (int r = 0; r < 3; r++) { CON(d.kpt[nw], r, t) = CON(sp, r, s); CON(di, r, t) = CON(si, r, s); }
        int32_t slot = -1;
        if (k == 0 || sa[s - 1] != a || sb[s - 1] != b) slot = cache_insert(d, nw, t, a, b);
        d.hslot[nw][t] = slot;
    }
}

// A world's warm-start cache written from the host (omi_batch_write_cache):
// the entries [base, base + n) of the current buffer are in place; their group
// heads go into its hash.
__global__ void k_cache_seed(Dev d, int64_t base, int32_t n) {
    const int32_t cur = *d.sel;
    for (int64_t k = gtid(); k < n; k += gstride()) {
        const int64_t t = base + k;
        const int32_t a = d.ka[cur][t], b = d.kb[cur][t];
        int32_t slot = -1;
        if (k == 0 || d.ka[cur][t - 1] != a || d.kb[cur][t - 1] != b) slot = cache_insert(d, cur, t, a, b);
        d.hslot[cur][t] = slot;
    }
}

// End of step: the cache buffer just written keeps its entry count (for
// emptying its hash slots before its next use) and becomes the current one.
__global__ void k_toggle(Dev d) {
    if (gtid() == 0) {
        const int32_t nw = *d.sel ^ 1;
        const uint64_t c = d.cnt[C_CACHE];
        d.hslot_n[nw] = (int32_t)(c < (uint64_t)d.kcap ? c : (uint64_t)d.kcap);
        *d.sel = nw;
    }
}

// Staged device writes: block b copies segment b (elements strided over the
// block's threads).  The grid is exactly the segment count.
__global__ __launch_bounds__(256) void k_copy_segments(const CopySeg* segs) {
    const CopySeg sg = segs[blockIdx.x];
    for (int64_t i = threadIdx.x; i < sg.n; i += blockDim.x) {
        if (sg.es == 8) {
            ((uint64_t*)sg.dst)[i] = sg.fill ? sg.src : ((const uint64_t*)sg.src)[i];
        } else if (sg.es == 4) {
            ((uint32_t*)sg.dst)[i] = sg.fill ? (uint32_t)sg.src : ((const uint32_t*)sg.src)[i];
        } else {
            ((uint8_t*)sg.dst)[i] = sg.fill ? (uint8_t)sg.src : ((const uint8_t*)sg.src)[i];
        }
    }
}

// ------------------------------------------------------------- integrate --

// NumpyBackend.integrate_positions (backend.py:42-49) with quat_integrate
// (mathutil.py:108-113) for awake dynamic and kinematic bodies.
__device__ __forceinline__ void post_body(const Dev& d, int64_t i, double dt) {
    if (d.world[i] < 0) return;
    const int mt = d.mt[i];
    if (!((mt == M_DYNAMIC || mt == M_KINEMATIC) && d.awake[i])) return;
    double p[3], q[4], v[3], w[3];
    for (int r = 0; r < 3; r++) { p[r] = BODY(d.pos, r, i); v[r] = BODY(d.lv, r, i); w[r] = BODY(d.av, r, i); }
    for (int r = 0; r < 4; r++) q[r] = BODY(d.quat, r, i);
    positions_body(p, q, v, w, dt);
    for (int r = 0; r < 3; r++) BODY(d.pos, r, i) = p[r];
    for (int r = 0; r < 4; r++) BODY(d.quat, r, i) = q[r];
}

__global__ void k_post(Dev d, int64_t nhw, double dt) {
    for (int64_t i = gtid(); i < nhw; i += gstride()) post_body(d, i, dt);
}

// The integrate stages on one world's host-layout rows ([n][3] velocities,
// positions, [n][4] quaternions; [n][3] per-body gravity), for the
// stage-level entry points.
__global__ void k_stage_forces(int32_t n, const int32_t* mt, const uint8_t* awake, const double* g, const double* gf,
                               const double* ld, const double* ad, const double* qd, double dlin, double dang,
                               double dt, double* lv, double* av) {
    for (int64_t i = gtid(); i < n; i += gstride()) {
        if (!(mt[i] == M_DYNAMIC && awake[i])) continue;
        double v[3] = {lv[3 * i], lv[3 * i + 1], lv[3 * i + 2]}, w[3] = {av[3 * i], av[3 * i + 1], av[3 * i + 2]};
        forces_body(v, w, g + 3 * i, gf[i], ld[i], ad[i], qd[i], dlin, dang, dt);
        for (int r = 0; r < 3; r++) { lv[3 * i + r] = v[r]; av[3 * i + r] = w[r]; }
    }
}

__global__ void k_stage_positions(int32_t n, const int32_t* mt, const uint8_t* awake, const double* lv,
                                  const double* av, double dt, double* pos, double* quat) {
    for (int64_t i = gtid(); i < n; i += gstride()) {
        if (!((mt[i] == M_DYNAMIC || mt[i] == M_KINEMATIC) && awake[i])) continue;
        double p[3] = {pos[3 * i], pos[3 * i + 1], pos[3 * i + 2]};
        double q[4] = {quat[4 * i], quat[4 * i + 1], quat[4 * i + 2], quat[4 * i + 3]};
        positions_body(p, q, lv + 3 * i, av + 3 * i, dt);
        for (int r = 0; r < 3; r++) pos[3 * i + r] = p[r];
        for (int r = 0; r < 4; r++) quat[4 * i + r] = q[r];
    }
}

// -------------------------------------------------------------- snapshot --

// ThreadedSimulation._publish (threaded.py:149-158): per body the position,
// the orientation as VRML axis-angle by quat_to_axis_angle (mathutil.py:90-105,
// with quat_normalize = q / max(|q|, 1e-12), the norm summed left to right),
// awake, and motion type == DYNAMIC.  One block per listed world; world k's
// bodies go to packed rows woff[k] .. woff[k] + n.
__global__ void k_snapshot(Dev d, const int32_t* wl, const int64_t* woff, int32_t nw, double* pos, double* aa,
                           uint8_t* awake, uint8_t* dyn) {
    const int32_t k = blockIdx.x;
    if (k >= nw) return;
    const int32_t w = wl[k];
    const int64_t off = d.w_off[w], o = woff[k];
    const int32_t n = d.w_cnt[w];
    for (int32_t j = threadIdx.x; j < n; j += blockDim.x) {
        const int64_t i = off + j, r = o + j;
        for (int c = 0; c < 3; c++) pos[3 * r + c] = BODY(d.pos, c, i);
        const double qx = BODY(d.quat, 0, i), qy = BODY(d.quat, 1, i), qz = BODY(d.quat, 2, i),
                     qw = BODY(d.quat, 3, i);
        const double nn = maxd(sqrt(((qx * qx + qy * qy) + qz * qz) + qw * qw), 1e-12);
        const double ux = qx / nn, uy = qy / nn, uz = qz / nn;
        double cw = qw / nn;
        cw = cw < -1.0 ? -1.0 : cw > 1.0 ? 1.0 : cw;
        const double sn = sqrt(maxd(1.0 - cw * cw, 0.0));
        const bool small = sn < 1e-9;
        const double ss = small ? 1.0 : sn;
        aa[4 * r + 0] = small ? 0.0 : ux / ss;
        aa[4 * r + 1] = small ? 1.0 : uy / ss;
        aa[4 * r + 2] = small ? 0.0 : uz / ss;
        aa[4 * r + 3] = small ? 0.0 : 2.0 * acos(cw);
        awake[r] = d.awake[i];
        dyn[r] = d.mt[i] == M_DYNAMIC ? 1 : 0;
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
        if (k < d.w_cs[w] || k >= (int64_t)d.w_cs[w] + d.w_cn[w]) continue;   // a world that overflowed
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
__device__ __forceinline__ void sleep_body(const Dev& d, int64_t i, double dt, double lin2, double ang2,
                                           double tsleep) {
    const int32_t w = d.world[i];
    if (w < 0 || !(d.w_flags[w] & WF_SLEEP)) return;
    const bool dyn = d.mt[i] == M_DYNAMIC;
    const double x = BODY(d.lv, 0, i), y = BODY(d.lv, 1, i), z = BODY(d.lv, 2, i);
    const double p = BODY(d.av, 0, i), q = BODY(d.av, 1, i), r = BODY(d.av, 2, i);
    const double s2 = (x * x + y * y) + z * z, w2 = (p * p + q * q) + r * r;
    const bool slow = dyn && s2 < lin2 && w2 < ang2;
    double tm = slow ? d.timer[i] + dt : 0.0;
    d.timer[i] = tm;
    const bool ready = slow && tm > tsleep;
    d.awake0[i] = (uint8_t)((slow ? 1 : 0) | (ready ? 2 : 0));   // reuse as per-body flags
    if (d.contacted[i] && !ready) d.notready[uf_find(d.parent, (int32_t)i)] = 1;
}

__global__ void k_sleep_body(Dev d, int64_t nhw, double dt, double lin2, double ang2, double tsleep) {
    for (int64_t i = gtid(); i < nhw; i += gstride()) sleep_body(d, i, dt, lin2, ang2, tsleep);
}

// Positions, then the sleep pass's speed tests, one thread per body: the
// positions write only poses, the speed tests read velocities and timers.
__global__ void k_post_sleep(Dev d, int64_t nhw, double dt, double lin2, double ang2, double tsleep) {
    for (int64_t i = gtid(); i < nhw; i += gstride()) {
        post_body(d, i, dt);
        sleep_body(d, i, dt, lin2, ang2, tsleep);
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
__global__ __launch_bounds__(NARROW_BLOCK) void k_stage_boxbox(int32_t np, const double* ca, const double* Ra, const double* ha, const double* cb,
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

void launch_pre(const Dev& d, int precision, int64_t nhw, double dt, double margin, hipStream_t s, bool collide_only) {
    if (nhw <= 0) return;
    int g = GRID_FOR(nhw, 256);
    if (collide_only) {
        if (precision == 64) hipLaunchKernelGGL((k_pre<double, PRE_COLLIDE>), dim3(g), dim3(256), 0, s, d, nhw, dt, margin);
        else hipLaunchKernelGGL((k_pre<float, PRE_COLLIDE>), dim3(g), dim3(256), 0, s, d, nhw, dt, margin);
    } else {
        if (precision == 64) hipLaunchKernelGGL((k_pre<double, PRE_STEP>), dim3(g), dim3(256), 0, s, d, nhw, dt, margin);
        else hipLaunchKernelGGL((k_pre<float, PRE_STEP>), dim3(g), dim3(256), 0, s, d, nhw, dt, margin);
    }
}

void launch_stage_forces(int32_t n, const int32_t* mt, const uint8_t* awake, const double* g, const double* gf,
                         const double* ld, const double* ad, const double* qd, double dlin, double dang, double dt,
                         double* lv, double* av, hipStream_t s) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_stage_forces, dim3(GRID_FOR(n, 256)), dim3(256), 0, s, n, mt, awake, g, gf, ld, ad, qd, dlin,
                       dang, dt, lv, av);
}

void launch_stage_positions(int32_t n, const int32_t* mt, const uint8_t* awake, const double* lv, const double* av,
                            double dt, double* pos, double* quat, hipStream_t s) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_stage_positions, dim3(GRID_FOR(n, 256)), dim3(256), 0, s, n, mt, awake, lv, av, dt, pos,
                       quat);
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
static void sap_sort_one(const Dev& d, const int32_t* wl, const int32_t* wn, int32_t np, bool global, int grid,
                         hipStream_t s) {
    if (global) hipLaunchKernelGGL((k_sap_sort<B, true>), dim3(grid), dim3(B), 0, s, d, wl, wn, 0);
    else hipLaunchKernelGGL((k_sap_sort<B, false>), dim3(grid), dim3(B), (size_t)np * 16, s, d, wl, wn, np);
}

// The broadphase sort of size class c, in LDS or (np = 0) in global scratch.
void launch_sap_sort(const Dev& d, int c, const int32_t* wl, const int32_t* wn, int32_t np, int32_t block,
                     int grid, hipStream_t s) {
    // np: entries each world's sort holds in LDS (16 bytes each), 0: the
    // worlds' sorts run in global memory
    const bool global = np <= 0;
    switch (block) {
        case 64: sap_sort_one<64>(d, wl, wn, np, global, grid, s); break;
        case 128: sap_sort_one<128>(d, wl, wn, np, global, grid, s); break;
        case 512: sap_sort_one<512>(d, wl, wn, np, global, grid, s); break;
        case 1024: sap_sort_one<1024>(d, wl, wn, np, global, grid, s); break;
        default: sap_sort_one<256>(d, wl, wn, np, global, grid, s); break;
    }
}

// The chip-wide candidate passes after every class is sorted.
void launch_sap_pairs(const Dev& d, int32_t record, int32_t grid, hipStream_t s) {
    hipLaunchKernelGGL((k_sap_pairs<256, false>), dim3(grid), dim3(256), 0, s, d, record);
    hipLaunchKernelGGL((k_sap_scan<1024>), dim3(1), dim3(1024), 0, s, d, record);
    hipLaunchKernelGGL((k_sap_pairs<256, true>), dim3(grid), dim3(256), 0, s, d, record);
}

void launch_narrow(const Dev& d, int precision, hipStream_t s) {
    int g = GRID_FOR(d.pcap, NARROW_BLOCK);
    if (precision == 64) hipLaunchKernelGGL(k_narrow<double>, dim3(g), dim3(NARROW_BLOCK), 0, s, d);
    else hipLaunchKernelGGL(k_narrow<float>, dim3(g), dim3(NARROW_BLOCK), 0, s, d);
    if (precision == 64) hipLaunchKernelGGL(k_narrow_clip16<double>, dim3(4), dim3(64), 0, s, d);
    else hipLaunchKernelGGL(k_narrow_clip16<float>, dim3(4), dim3(64), 0, s, d);
    if (d.epa_slots > 0) {
        int ge = (d.epa_slots + 63) / 64;
        if (precision == 64) hipLaunchKernelGGL(k_narrow_convex<double>, dim3(ge), dim3(64), d.cvx_lds, s, d, d.cvx_lds);
        else hipLaunchKernelGGL(k_narrow_convex<float>, dim3(ge), dim3(64), d.cvx_lds, s, d, d.cvx_lds);
    }
}

size_t epa_work_bytes() { return sizeof(EpaWork); }

void launch_stage_convex(int32_t np, const int32_t* ka, const double* ca, const double* Ra, const double* ha,
                         const int32_t* oa, const int32_t* na, const int32_t* kb, const double* cb, const double* Rb,
                         const double* hb, const int32_t* ob, const int32_t* nb, const double* hull, int64_t nhull,
                         void* work, int32_t slots, int32_t* cnt, double* nrm, double* pt, double* dep, hipStream_t s) {
    hipLaunchKernelGGL(k_stage_convex, dim3((slots + 63) / 64), dim3(64), 0, s, np, ka, ca, Ra, ha, oa, na, kb, cb,
                       Rb, hb, ob, nb, hull, nhull, (EpaWork*)work, slots, cnt, nrm, pt, dep);
}

void launch_compact(const Dev& d, int precision, int grid, hipStream_t s) {
    const int64_t nch = (d.pcap + CPT_CHUNK - 1) / CPT_CHUNK;
    const int g = (int)(nch < grid ? nch : grid);
    hipLaunchKernelGGL(k_pair_count, dim3(g), dim3(CPT_BLOCK), 0, s, d);
    hipLaunchKernelGGL((k_pair_scan<1024>), dim3(1), dim3(1024), 0, s, d);
    if (precision == 64) hipLaunchKernelGGL((k_pair_write<double>), dim3(g), dim3(CPT_BLOCK), 0, s, d);
    else hipLaunchKernelGGL((k_pair_write<float>), dim3(g), dim3(CPT_BLOCK), 0, s, d);
    hipLaunchKernelGGL(k_world_slices, dim3(GRID_FOR(d.wcap, 256)), dim3(256), 0, s, d);
}

template <typename R, int B>
static void solve_one(const Dev& d, SolveIn in, const int32_t* wl, const int32_t* wn, int32_t lds_bytes,
                      int grid, hipStream_t s, int32_t isl) {
    if (isl == 1)
        hipLaunchKernelGGL((k_solve<R, B, true>), dim3(grid), dim3(B), (size_t)lds_bytes, s, d, in, wl, wn, lds_bytes, 0);
    else
        hipLaunchKernelGGL((k_solve<R, B, false>), dim3(grid), dim3(B), (size_t)lds_bytes, s, d, in, wl, wn, lds_bytes,
                           (int32_t)(isl == 2));
}

template <typename R>
static void solve_block(const Dev& d, int block, SolveIn in, const int32_t* wl, const int32_t* wn,
                        int32_t lds_bytes, int grid, hipStream_t s, int32_t isl) {
    switch (block) {
        case 64: solve_one<R, 64>(d, in, wl, wn, lds_bytes, grid, s, isl); break;
        case 128: solve_one<R, 128>(d, in, wl, wn, lds_bytes, grid, s, isl); break;
        default: solve_one<R, 256>(d, in, wl, wn, lds_bytes, grid, s, isl); break;
        // the FP64 sweep holds two prefetched items in ~300 VGPRs per thread,
        // more than a 512-thread block leaves it: FP64 runs up to 256 threads
        // (the C-ABI rejects larger FP64 blocks)
        case 512:
            if constexpr (sizeof(R) == 4) solve_one<R, 512>(d, in, wl, wn, lds_bytes, grid, s, isl);
            break;
        case 1024:
            if constexpr (sizeof(R) == 4) solve_one<R, 1024>(d, in, wl, wn, lds_bytes, grid, s, isl);
            break;
    }
}

// Let the solver and broadphase-sort launches use `bytes` of dynamic LDS;
// returns false (and clears the error) if the runtime refuses.
template <typename F>
static bool allow_lds_fn(F f, int bytes) {
    if (hipFuncSetAttribute((const void*)f, hipFuncAttributeMaxDynamicSharedMemorySize, bytes) == hipSuccess)
        return true;
    (void)hipGetLastError();
    return false;
}

template <typename R, int B>
static bool allow_lds_one(int bytes) {
    return allow_lds_fn(k_solve<R, B, false>, bytes) && allow_lds_fn(k_solve<R, B, true>, bytes);
}

bool allow_dynamic_lds(int bytes) {
    bool ok = true;
    ok = allow_lds_one<double, 64>(bytes) && ok; ok = allow_lds_one<double, 128>(bytes) && ok;
    ok = allow_lds_one<double, 256>(bytes) && ok;
    ok = allow_lds_one<float, 64>(bytes) && ok; ok = allow_lds_one<float, 128>(bytes) && ok;
    ok = allow_lds_one<float, 256>(bytes) && ok; ok = allow_lds_one<float, 512>(bytes) && ok;
    ok = allow_lds_one<float, 1024>(bytes) && ok;
    ok = allow_lds_fn(k_sap_sort<64, false>, bytes) && ok; ok = allow_lds_fn(k_sap_sort<128, false>, bytes) && ok;
    ok = allow_lds_fn(k_sap_sort<256, false>, bytes) && ok; ok = allow_lds_fn(k_sap_sort<512, false>, bytes) && ok;
    ok = allow_lds_fn(k_sap_sort<1024, false>, bytes) && ok;
    ok = allow_lds_fn(k_islands<256>, bytes) && ok;
    return ok;
}

// Whether a world of n bodies can be solved island by island with lds_bytes
// of dynamic LDS: the level table and the body map of an island's block, and
// the island pass's three per-body arrays, all in LDS.
bool islands_fit(int32_t n, int32_t lds_bytes) {
    const size_t avail = (size_t)lds_bytes - 4 * (size_t)SOLVE_DCAP;
    return lds_bytes > 4 * SOLVE_DCAP && al16((size_t)n * 8) + (size_t)n * 4 <= avail &&
           3 * al16((size_t)(n + 1) * 4) <= (size_t)lds_bytes;
}

// The islands of size class c's worlds (one block per world) into the island
// table the class's island-by-island solve launch reads.
void launch_islands(const Dev& d, const int32_t* wl, const int32_t* wn, int32_t lds_bytes, int grid, hipStream_t s) {
    hipLaunchKernelGGL((k_islands<256>), dim3(grid), dim3(256), (size_t)lds_bytes, s, d, wl, wn, lds_bytes);
}

// The convex narrowphase's staged hull points: returns the dynamic LDS a
// launch takes, or 0 when the runtime refuses it.
int32_t allow_convex_lds() {
    constexpr int bytes = 2 * 3 * HULL_STAGE_PTS * 64 * 8;
    return allow_lds_fn(k_narrow_convex<double>, bytes) && allow_lds_fn(k_narrow_convex<float>, bytes) ? bytes : 0;
}

// The joint launches' dynamic LDS (they also hold about 4 KB of static LDS).
bool allow_joint_lds(int bytes) {
    return allow_lds_fn(k_joints<double, 256>, bytes) && allow_lds_fn(k_joints<float, 256>, bytes);
}

// Solver rows for every contact (one thread per contact).
// Rows of the step's contacts; with skip_above > 0 only those of worlds of at
// most skip_above bodies.
void launch_prepare(const Dev& d, int precision, SolveIn in, bool stage, hipStream_t s, int32_t skip_above) {
    const int g = GRID_FOR(d.kcap, 256);
    if (precision == 64) {
        if (stage) hipLaunchKernelGGL((k_prepare<double, true>), dim3(g), dim3(256), 0, s, d, in, skip_above);
        else hipLaunchKernelGGL((k_prepare<double, false>), dim3(g), dim3(256), 0, s, d, in, skip_above);
    } else {
        if (stage) hipLaunchKernelGGL((k_prepare<float, true>), dim3(g), dim3(256), 0, s, d, in, skip_above);
        else hipLaunchKernelGGL((k_prepare<float, false>), dim3(g), dim3(256), 0, s, d, in, skip_above);
    }
}

// Rows of the contacts of the listed worlds (list wl, count *wn, at most
// `grid` worlds), eight blocks of 256 threads per world.
void launch_prepare_worlds(const Dev& d, int precision, SolveIn in, bool stage, const int32_t* wl,
                           const int32_t* wn, int grid, hipStream_t s) {
    constexpr int32_t G = 8;
    const dim3 g((unsigned)grid * G);
    if (precision == 64) {
        if (stage) hipLaunchKernelGGL((k_prepare_worlds<double, true>), g, dim3(256), 0, s, d, in, wl, wn, G);
        else hipLaunchKernelGGL((k_prepare_worlds<double, false>), g, dim3(256), 0, s, d, in, wl, wn, G);
    } else {
        if (stage) hipLaunchKernelGGL((k_prepare_worlds<float, true>), g, dim3(256), 0, s, d, in, wl, wn, G);
        else hipLaunchKernelGGL((k_prepare_worlds<float, false>), g, dim3(256), 0, s, d, in, wl, wn, G);
    }
}

// One block per world of size class c (world list wl, count *wn), or with
// isl = 1 the class's islands over `grid` blocks, with lds_bytes of dynamic
// LDS (at least 4 * SOLVE_DCAP + 16, a multiple of 16); isl = 2: one block per
// world that also writes its contacts' solver rows (blocks above 64 threads).
void launch_solve_class(const Dev& d, int precision, int block, SolveIn in, const int32_t* wl, const int32_t* wn,
                        int32_t lds_bytes, int grid, hipStream_t s, int32_t isl) {
    if (precision == 64) solve_block<double>(d, block, in, wl, wn, lds_bytes, grid, s, isl);
    else solve_block<float>(d, block, in, wl, wn, lds_bytes, grid, s, isl);
}

// The warm-start cache rebuild in two parts: the hash slots of the buffer
// the step writes emptied and each world's place planned (both need only the
// step's contact counts), then the entries copied (they need the solve's
// impulses).
void launch_cache_plan(const Dev& d, hipStream_t s) {
    hipLaunchKernelGGL(k_hash_clear, dim3(GRID_FOR(d.kcap, 256)), dim3(256), 0, s, d);
    hipLaunchKernelGGL((k_cache_plan<1024>), dim3(1), dim3(1024), 0, s, d);
}

void launch_cache_write(const Dev& d, int precision, hipStream_t s) {
    const int g = GRID_FOR(d.kcap, 256);
    if (precision == 64) hipLaunchKernelGGL((k_cache_all<double>), dim3(g), dim3(256), 0, s, d);
    else hipLaunchKernelGGL((k_cache_all<float>), dim3(g), dim3(256), 0, s, d);
}

// One block per world; lds_bytes (a multiple of 16) of dynamic LDS holds
// the staged records of up to lds_bytes / 132 (FP64) or / 100 (FP32) joint
// bodies per world.
void launch_joints(const Dev& d, int precision, double dt, int32_t lds_bytes, hipStream_t s) {
    if (precision == 64)
        hipLaunchKernelGGL((k_joints<double, 256>), dim3(d.wcap), dim3(256), lds_bytes, s, d, dt, lds_bytes);
    else hipLaunchKernelGGL((k_joints<float, 256>), dim3(d.wcap), dim3(256), lds_bytes, s, d, dt, lds_bytes);
}

void launch_post(const Dev& d, int64_t nhw, double dt, hipStream_t s) {
    if (nhw <= 0) return;
    hipLaunchKernelGGL(k_post, dim3(GRID_FOR(nhw, 256)), dim3(256), 0, s, d, nhw, dt);
}

// _update_sleep's thresholds (world.py:510-511) as the step uses them.
// The sleep pass in two parts: the islands of the step's contacts (they need
// only the contact list), then the speed tests and the sleep decisions (they
// need the solved velocities).
void launch_sleep_union(const Dev& d, int64_t nhw, hipStream_t s) {
    if (nhw > 0) hipLaunchKernelGGL(k_sleep_union, dim3(GRID_FOR(d.kcap, 256)), dim3(256), 0, s, d);
}

void launch_sleep_decide(const Dev& d, int64_t nhw, double dt, hipStream_t s, double lin, double ang, double tsleep) {
    if (nhw <= 0) return;
    hipLaunchKernelGGL(k_sleep_body, dim3(GRID_FOR(nhw, 256)), dim3(256), 0, s, d, nhw, dt, lin * lin, ang * ang,
                       tsleep);
    hipLaunchKernelGGL(k_sleep_final, dim3(GRID_FOR(nhw, 256)), dim3(256), 0, s, d, nhw);
}

// Positions and the sleep decisions (after the islands of the step's contacts).
void launch_post_sleep(const Dev& d, int64_t nhw, double dt, hipStream_t s, double lin, double ang, double tsleep) {
    if (nhw <= 0) return;
    hipLaunchKernelGGL(k_post_sleep, dim3(GRID_FOR(nhw, 256)), dim3(256), 0, s, d, nhw, dt, lin * lin, ang * ang,
                       tsleep);
    hipLaunchKernelGGL(k_sleep_final, dim3(GRID_FOR(nhw, 256)), dim3(256), 0, s, d, nhw);
}

void launch_sleep(const Dev& d, int64_t nhw, double dt, hipStream_t s, double lin, double ang, double tsleep) {
    launch_sleep_union(d, nhw, s);
    launch_sleep_decide(d, nhw, dt, s, lin, ang, tsleep);
}

void launch_toggle(const Dev& d, hipStream_t s) { hipLaunchKernelGGL(k_toggle, dim3(1), dim3(64), 0, s, d); }

void launch_copy_segments(const CopySeg* segs, int32_t nseg, hipStream_t s) {
    if (nseg > 0) hipLaunchKernelGGL(k_copy_segments, dim3(nseg), dim3(256), 0, s, segs);
}

void launch_cache_seed(const Dev& d, int64_t base, int32_t n, hipStream_t s) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_cache_seed, dim3(GRID_FOR(n, 256)), dim3(256), 0, s, d, base, n);
}

// _collide_native.capsule_triangles for one capsule and n triangles
// verts[n][3][3], one thread per triangle; non-contacts give zeros.
__global__ void k_capsule_triangles(int64_t n, const double* c0, const double* c1, double radius,
                                    const double* verts, uint8_t* hit, double* points, double* normals,
                                    double* depths) {
    __shared__ double cap[6];
    if (threadIdx.x < 3) { cap[threadIdx.x] = c0[threadIdx.x]; cap[3 + threadIdx.x] = c1[threadIdx.x]; }
    __syncthreads();
    for (int64_t i = gtid(); i < n; i += gstride()) {
        const double* v = verts + 9 * i;
        double pt[3] = {0.0, 0.0, 0.0}, nr[3] = {0.0, 0.0, 0.0}, dp = 0.0;
        const int h = capsule_triangle(cap, cap + 3, radius, v, v + 3, v + 6, pt, nr, dp);
        hit[i] = (uint8_t)h;
        for (int m = 0; m < 3; m++) { points[3 * i + m] = pt[m]; normals[3 * i + m] = nr[m]; }
        depths[i] = dp;
    }
}

void launch_capsule_triangles(int64_t n, const double* c0, const double* c1, double radius, const double* verts,
                              uint8_t* hit, double* points, double* normals, double* depths, hipStream_t s) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_capsule_triangles, dim3(GRID_FOR(n, 256)), dim3(256), 0, s, n, c0, c1, radius, verts, hit,
                       points, normals, depths);
}

void launch_trigger(const Dev& d, hipStream_t s) {
    // with hulls, one thread per EPA workspace slot; else a fixed grid
    const int32_t threads = d.epa_work ? d.epa_slots : 4096;
    hipLaunchKernelGGL(k_trigger, dim3((threads + 63) / 64), dim3(64), 0, s, d, threads);
}

void launch_snapshot(const Dev& d, const int32_t* wl, const int64_t* woff, int32_t nw, double* pos, double* aa,
                     uint8_t* awake, uint8_t* dyn, hipStream_t s) {
    if (nw <= 0) return;
    hipLaunchKernelGGL(k_snapshot, dim3(nw), dim3(256), 0, s, d, wl, woff, nw, pos, aa, awake, dyn);
}

void launch_stage_boxbox(int precision, int32_t np, const double* ca, const double* Ra, const double* ha,
                         const double* cb, const double* Rb, const double* hb, int32_t* cnt, double* nrm,
                         double* pts, double* dep, hipStream_t s) {
    int g = GRID_FOR(np, NARROW_BLOCK);
    if (precision == 64)
        hipLaunchKernelGGL(k_stage_boxbox<double>, dim3(g), dim3(NARROW_BLOCK), 0, s, np, ca, Ra, ha, cb, Rb, hb, cnt, nrm, pts, dep);
    else
        hipLaunchKernelGGL(k_stage_boxbox<float>, dim3(g), dim3(NARROW_BLOCK), 0, s, np, ca, Ra, ha, cb, Rb, hb, cnt, nrm, pts, dep);
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
