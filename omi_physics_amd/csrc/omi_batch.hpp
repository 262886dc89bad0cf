// omi_batch.hpp — device-resident batch layout shared by the kernels and the host.
//
// All per-body arrays are structure-of-arrays over a global body index
// (worlds occupy contiguous ranges), component-major: component c of body i is
// at [c * ncap + i], so each component load of a wavefront is one coalesced
// 512-byte (FP64) transaction.  Per-contact arrays use the same scheme over
// the contact index.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace omi {

enum : int32_t { WF_ACTIVE = 1, WF_COLLISION = 2, WF_SLEEP = 4, WF_SEQWAKE = 8, WF_JOINTS = 16 };
enum : uint8_t { BF_COLLIDER = 1, BF_TRIGGER = 2 };
// manifold record field y: contacts | side a dynamic | side b dynamic
enum : int32_t { MR_COUNT = 0xff, MR_DYN_A = 0x100, MR_DYN_B = 0x200 };

// counters (uint64 slots)
enum : int32_t {
    C_PAIRS = 0,       // solid pairs reserved
    C_BPAIRS = 1,      // broadphase pairs reserved (record mode)
    C_CONTACTS = 2,    // contacts reserved
    C_CACHE = 3,       // warm-start cache entries reserved (new buffer)
    C_XCAND = 4,       // x-window candidates
    C_BPTOTAL = 5,     // broadphase pairs (all worlds)
    C_MANIFOLDS = 6,
    C_MAXLEVEL = 7,
    C_OVERFLOW = 8,    // bit 0 pairs, 1 contacts, 2 cache, 3 broadphase record, 4 candidate tiles, 5 trigger pairs,
                       // 6 a world solved island by island outgrew the launch's LDS
    C_CONVEX = 9,      // convex pairs queued for GJK/EPA
    C_TILES = 10,      // broadphase candidate tiles reserved
    C_PVALID = 11,     // solid pairs of the worlds whose pair lists fit (a prefix of the pair list)
    C_TPAIRS = 12,     // broadphase pairs with a trigger body (in no particular order)
    C_TOVER = 13,      // trigger overlaps found
    C_ISLANDS = 14,    // islands of the step in the size classes solved island by island
    C_CLIP16 = 15,     // box pairs whose clip outgrew 8 vertices (queued from the top of cvx)
    C_COUNT = 16
};

// One staged device write (host world loads and removals, world lists,
// readback of retiring worlds): n elements of es bytes (1, 4 or 8) to dst,
// from src, or (fill) src's low es bytes repeated.
struct CopySeg {
    uint64_t dst;
    uint64_t src;
    int64_t n;
    int32_t es;
    int32_t fill;
};
static_assert(sizeof(CopySeg) == 32, "32-byte segment records");

// ints per island record: world, manifolds, contacts, first manifold, first contact
constexpr int32_t ISL_DESC = 8;

// x-window candidates per broadphase tile (one wavefront task)
constexpr int32_t SAP_TILE = 512;
// threads per block of the box/sphere narrowphase (one pair per thread)
constexpr int32_t NARROW_BLOCK = 256;
// contact compaction: pairs per chunk (CPT_ITEMS rows of CPT_BLOCK threads)
constexpr int32_t CPT_BLOCK = 256, CPT_ITEMS = 4, CPT_CHUNK = CPT_BLOCK * CPT_ITEMS;
// slots of one contact's solver row (26 used) and of one body's world inverse
// inertia (9 used): whole 16-byte chunks in FP32 and FP64
constexpr int32_t SR_STRIDE = 28;
constexpr int32_t IIW_STRIDE = 12;
// largest solver block in FP64: two prefetched sweep items need more
// registers per thread than a 512-thread block leaves
constexpr int32_t FP64_SOLVE_BLOCK_MAX = 256;
// bytes of one contact's solver row (FP64: 28 doubles; FP32: 56 floats)
constexpr int32_t SOLVE_ROW_BYTES = 224;
// linear probes of the warm-start hash, for both insertion and lookup; an
// insertion that finds no slot within them is reported as a cache overflow
constexpr int32_t HASH_PROBES = 64;

struct Dev {
    int64_t ncap, pcap, kcap, hcap;   // bodies, pairs, contacts, hash slots (pow2)
    int32_t wcap;
    // per body
    double *pos, *quat, *ppos, *pquat;    // [3|4][ncap]
    double *lv, *av;                      // [3][ncap] FP64 state
    double *lo, *hi;                      // [3][ncap]
    double *half, *inv_mass, *inv_inertia;
    double *gf, *ld, *ad, *qd, *timer;
    void *rot;                            // R[9][ncap] narrowphase rotation
    void *iiw;                            // R[ncap][IIW_STRIDE] world inverse inertia, row-major 3x3
    void *svel;                           // FP32: R[6][ncap] solver velocity staging; FP64: R[ncap][6] (body records)
    int32_t *world;                       // world slot or -1
    int32_t *mat, *filt, *hoff, *hcnt;
    uint8_t *mt, *kind, *bflags, *awake, *awake0, *contacted, *notready;
    int32_t *parent, *last;               // union-find, level schedule scratch
    uint64_t *lmask;                      // [ncap] level schedule body masks (worlds too large for LDS)
    uint64_t *skey;                       // [2*ncap] broadphase sort keys (large worlds)
    int32_t *sidx, *soff;                 // [2*ncap] broadphase scratch; soff: window prefix per world
    double *sbox;                         // [4][ncap] y/z bounds in sorted order (lo.y, hi.y, lo.z, hi.z)
    uint32_t *sinfo;                      // [ncap] motion/collider/trigger bits | filter slot << 8, sorted
    int32_t *sorig;                       // [ncap] world-local body index of each sorted position
    double *hull;                         // [hull_cap][3] local hull points, row-major
    // per world
    int32_t *w_off, *w_cnt, *w_flags, *w_iters;
    int32_t *w_mat_off, *w_nmat, *w_filt_off, *w_nfilt;
    double *w_grav, *w_damp;              // [3][wcap], [2][wcap]
    int32_t *w_ps, *w_pn, *w_bps, *w_bpn; // solid / broadphase pair ranges
    int32_t *w_cs, *w_cn;                 // contact ranges
    int32_t *w_mn;                        // manifolds (runs of contacts of one pair) per world
    int32_t *w_ms;                        // first manifold record of each world
    int32_t *w_levels;                    // levels of the last solve
    int64_t *w_x;                         // x-window candidates of the step
    int32_t *w_t0, *w_tn;                 // candidate tiles: first, count
    int64_t tcap;                         // candidate tiles per step
    int32_t *tworld;                      // [tcap] world of each tile (-1: none)
    int64_t *tcnt;                        // [2][tcap + 1] solid / broadphase pairs per tile -> offsets
    int32_t *w_cache_s, *w_cache_n;       // [2][wcap] cache ranges per buffer
    double *mat_fric, *mat_rest;          // pooled tables
    uint8_t *filt_tab;
    // pairs
    int32_t *pa, *pb, *bpa, *bpb;
    int32_t *tpa, *tpb;                   // [tpcap] broadphase pairs with a trigger body
    int32_t *tva, *tvb;                   // [tpcap] trigger overlaps: (trigger body, other body)
    int64_t tpcap;
    int8_t *mcnt;
    int32_t *pfx;                         // [2][pcap + 1] first contact / manifold record of each pair
    int64_t *pchunk;                      // [2][pchunk_cap] contacts / manifolds per pair chunk -> offsets
    int64_t pchunk_cap;
    int32_t *cvx;                         // [pcap] pairs that need GJK/EPA
    void *epa_work;                       // EpaWork per GJK/EPA thread
    int32_t epa_slots;
    int32_t cvx_lds;                      // dynamic LDS of the convex narrowphase (0: hull points from HBM)
    int32_t clip16_all;                   // every box/sphere pair through k_narrow_clip16 (testing)
    double *mpts;                         // [3][4*pcap] absolute points
    void *mnrm, *mdep;                    // R [3][pcap], R [4*pcap]
    // contacts
    int32_t *ca, *cb;
    double *cpt;                          // [3][kcap]
    void *cnrm, *cdep, *cimp;             // R [3][kcap], R [kcap], R [3][kcap] (n, t1, t2 impulses)
    void *csolv;                          // R [kcap][SR_STRIDE] solver rows
    int4 *mrec;                           // [kcap] manifolds: first contact, contacts, dynamic a, dynamic b
    int32_t slmax;                        // passes per solve the schedule holds (iterations + warm start)
    int32_t *clevel, *corder, *lcount;    // [slmax * kcap] level schedule of the solver passes
    int32_t *cord_a, *cord_b;             // [slmax * kcap] bodies of each scheduled item
    int32_t *cord_c;                      // [slmax * kcap] contact (batch index) of each scheduled item
    // islands (size classes solved island by island): per manifold its root
    // body; per island a record (world, manifolds, contacts, first island
    // manifold, first island contact); the island's manifold records in order
    // with island-local first contacts; the batch index of each island contact
    int32_t *isl_lab;                     // [kcap]
    int32_t *isl_desc;                    // [kcap][ISL_DESC]
    int4 *isl_mrec;                       // [kcap]
    int32_t *isl_cidx;                    // [kcap]
    // warm-start cache, two buffers
    int32_t *ka[2], *kb[2];
    double *kpt[2];                       // [3][kcap]
    void *kimp[2];                        // R [3][kcap]
    uint64_t *hkey[2];
    int32_t *hval[2];
    int32_t *hslot[2];                    // [kcap] hash slot each cache entry filled, or -1
    int32_t *hslot_n;                     // [2] cache entries of each buffer's last fill
    // point joints
    int64_t jcap;
    int32_t *ja, *jb;                     // [jcap] body index or -1 (the world)
    double *janc, *jla, *jlb;             // [3][jcap] world anchor, local anchors
    uint8_t *jhas;                        // bit 0 / 1: local anchor of a / b fixed
    void *jrow;                           // [jcap] joint solver rows in level order (160 B FP64, 96 B FP32)
    int32_t *jlevel, *jorder, *jcount;    // level schedule of the joints: level, order -> joint, level ends
    int2 *jend;                           // [jcap] end records of each level-order position
    int32_t *jbody;                       // [2*jcap] world-local body of each joint-body slot
    int32_t *w_js, *w_jn, *w_jiters;      // per-world joint range, iterations
    int32_t *w_jsched, *w_jns;            // per world: 2 levels + LDS form (0: none), joint bodies
    int32_t *sel;                         // current cache buffer index (device)
    uint64_t *cnt;                        // counters
};

// Where the solver of one world gets seeds, restitution and friction: batch
// mode looks seeds up in the warm-start cache and reads the material tables;
// stage mode (omi_prepare_and_solve) takes them as arrays.
struct SolveIn {
    const double* st_rest;   // stage mode: [K] restitution, else null
    const double* st_fric;
    double rthr, correction, slop;
    int32_t warm;
    int32_t wake;            // 1: the kernels writing a world's rows also wake its sleepers
};

}  // namespace omi
