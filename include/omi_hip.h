/* omi_hip.h — C ABI of the MI355X contact pipeline (libomi_hip.so).
 *
 * Plain C: pointers, sizes and int status codes (0 = ok, < 0 = error, message
 * in omi_last_error()).  No device types cross this boundary; device memory is
 * owned by the library.  Host buffers are caller-owned, C-contiguous,
 * row-major ([n][3] vectors, [n][4] quaternions (x, y, z, w), [n][3][3]
 * matrices), exactly the numpy column layout of the reference's
 * PhysicsWorld (world.py:75-82).
 *
 * Two families of entry points:
 *   1. Stage-level calls that replace one reference function on identical
 *      inputs (synchronous; they copy in, run on the GPU, copy out).  Each
 *      declaration cites the reference interface it replaces.
 *   2. The batch: many independent worlds resident in HBM, stepped together
 *      (one world is the plain single-world form).
 */
#ifndef OMI_HIP_H
#define OMI_HIP_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define OMI_OK 0
#define OMI_ERR_ARG (-1)
#define OMI_ERR_HIP (-2)
#define OMI_ERR_CAPACITY (-3)
#define OMI_ERR_STATE (-4)

#define OMI_FP64 64
#define OMI_FP32 32

/* shape kinds (world._refit_kind codes, world.py:440-441; 3 = convex hull) */
#define OMI_KIND_NONE 0
#define OMI_KIND_SPHERE 1
#define OMI_KIND_BOX 2
#define OMI_KIND_CONVEX 3

/* Message of the last failed call on this thread. */
const char *omi_last_error(void);
/* ABI version (bumped on incompatible change). */
int32_t omi_abi_version(void);
/* Number of visible HIP devices (0 on a host without a GPU). */
int32_t omi_device_count(void);
/* Free and total device memory of the current device, in bytes (the serving
 * loop sizes its arena from it). */
int32_t omi_device_memory(int64_t *free_bytes, int64_t *total_bytes);

/* ---- 1. stage-level entry points ------------------------------------- */

/* Replaces PhysicsWorld.refit_aabbs (world.py:471-508) on one world's columns:
 * box |R|·half ± margin, sphere radius ± margin, no shape a point, convex the
 * hull's world bounds ± margin.  FP64, bit-exact with the reference's NumPy. */
int32_t omi_refit(int32_t n, const double *pos, const double *quat, const int32_t *kind,
                  const double *half, const int32_t *hull_off, const int32_t *hull_cnt,
                  const double *hull_pts, int64_t n_hull_pts, double margin,
                  double *aabb_min, double *aabb_max);

/* Replaces NumpyBackend.integrate_forces (backend.py:26-40) on one world's
 * columns: for awake dynamic bodies v += g·gravity_factor·dt, then
 * v *= max(1 - (linear_damping + default_linear_damping)·dt, 0), then
 * v -= quadratic_drag·|v|·v·dt, and w *= max(1 - (angular_damping +
 * default_angular_damping)·dt, 0).  gravity is per body [n][3] (the reference's
 * resolve_gravity, world.py:386-396); lv, av [n][3] are updated in place.
 * FP64, bit-exact with the reference's NumPy. */
int32_t omi_integrate_forces(int32_t n, const int32_t *motion_type, const uint8_t *awake,
                             const double *gravity, const double *gravity_factor,
                             const double *linear_damping, const double *angular_damping,
                             const double *quadratic_drag, double default_linear_damping,
                             double default_angular_damping, double dt, double *lv, double *av);

/* Replaces NumpyBackend.integrate_positions (backend.py:42-49) with
 * quat_integrate (mathutil.py:108-113): awake dynamic and kinematic bodies move
 * x += v·dt and q <- normalize(q + dt/2·(w, 0)⊗q); pos [n][3] and quat [n][4]
 * are updated in place.  FP64, bit-exact. */
int32_t omi_integrate_positions(int32_t n, const int32_t *motion_type, const uint8_t *awake,
                                const double *lv, const double *av, double dt, double *pos,
                                double *quat);

/* Replaces PhysicsWorld._update_sleep with _contact_islands (world.py:510-567)
 * for n bodies and the step's K contacts (body indices ca[k], cb[k]): a dynamic
 * body is slow when |v| < lin_thresh and |w| < ang_thresh; its timer grows by dt
 * while slow and resets otherwise; an island of dynamic-dynamic contacts sleeps
 * when every member has been slow for more than t_sleep, a contacted-by-nothing
 * dynamic body when it alone has; moving dynamic bodies wake; sleeping dynamic
 * bodies get zero velocity.  lv, av, awake and sleep_timer are updated in place.
 * The step uses lin_thresh 0.08, ang_thresh 0.25, t_sleep 0.4. */
int32_t omi_sleep(int32_t n, int64_t K, const int32_t *ca, const int32_t *cb, const int32_t *motion_type,
                  double dt, double lin_thresh, double ang_thresh, double t_sleep, double *lv, double *av,
                  uint8_t *awake, double *sleep_timer);

/* Replaces BroadPhase.pairs (broadphase.py:243-294): stable sweep on lo.x,
 * inclusive y/z overlap, >= 1 dynamic body, collision-filter table lookup;
 * pairs (min, max) in the reference's (rank_i, rank_j) order.  Writes up to
 * `cap` pairs to pairs[2*cap] and the full count to *count (count > cap means
 * the output was truncated). */
int32_t omi_sap_pairs(int32_t n, const double *aabb_min, const double *aabb_max,
                      const int32_t *motion_type, const int32_t *filt_slot, int32_t n_filt,
                      const uint8_t *filter_table, int32_t *pairs, int64_t cap, int64_t *count);

/* Replaces _collide_native.box_box (_collide_native.pyx:266-282), batched over
 * npairs independent pairs.  R* are [npairs][3][3] with the box axes in the
 * columns.  Outputs: count[npairs] in 0..4, normal[npairs][3] (a -> b),
 * points[npairs][4][3], depths[npairs][4].  precision is OMI_FP64 or OMI_FP32. */
int32_t omi_box_box(int32_t npairs, int32_t precision,
                    const double *ca, const double *Ra, const double *ha,
                    const double *cb, const double *Rb, const double *hb,
                    int32_t *count, double *normal, double *points, double *depths);

/* Replaces _collide_native.sphere_sphere (_collide_native.pyx:285-299), batched. */
int32_t omi_sphere_sphere(int32_t npairs, int32_t precision, const double *ca, const double *ra,
                          const double *cb, const double *rb, int32_t *count, double *normal,
                          double *point, double *depth);

/* Replaces _collide_native.sphere_box (_collide_native.pyx:302-344), batched;
 * the normal points sphere -> box like the reference. */
int32_t omi_sphere_box(int32_t npairs, int32_t precision, const double *cs, const double *rs,
                       const double *cb, const double *Rb, const double *hb, int32_t *count,
                       double *normal, double *point, double *depth);

/* Replaces _collide_native.capsule_triangles (_collide_native.pyx:454-574),
 * the character controller's query behind collide.capsule_triangle_batch
 * (collide.py:438-462): the capsule segment p0 p1 with `radius` against n
 * triangles verts[n][3][3] (triangle, vertex, axis).  Outputs for every
 * triangle: hit[n] 0/1, points[n][3] (on the triangle), normals[n][3]
 * (triangle -> capsule), depths[n]; zeros where hit is 0.  FP64, bit-exact
 * with the Cython loop. */
int32_t omi_capsule_triangles(int64_t n, const double *p0, const double *p1, double radius,
                              const double *verts, uint8_t *hit, double *points, double *normals,
                              double *depths);

/* Replaces gjk.collide_convex (gjk.py:249-264) for one batch of pairs.
 * kind_* is OMI_KIND_SPHERE/BOX/CONVEX; hulls are [*_cnt] points starting at
 * *_off in hull_pts (local frame); R* [npairs][3][3]; half for boxes, half[0]
 * the radius for spheres.  count is 0 or 1. */
int32_t omi_collide_convex(int32_t npairs, int32_t precision,
                           const int32_t *kind_a, const double *ca, const double *Ra,
                           const double *ha, const int32_t *off_a, const int32_t *cnt_a,
                           const int32_t *kind_b, const double *cb, const double *Rb,
                           const double *hb, const int32_t *off_b, const int32_t *cnt_b,
                           const double *hull_pts, int64_t n_hull_pts,
                           int32_t *count, double *normal, double *point, double *depth);

/* No counterpart in the reference: a variant of omi_collide_convex for callers
 * that cannot allocate device memory (code under HIP graph capture, tight
 * serving loops, harnesses that forbid hipMalloc).  Same arguments and results,
 * but every buffer is a caller-owned DEVICE pointer, and the call is
 * asynchronous on `stream` (a hipStream_t; NULL = the null stream): the caller
 * synchronises before reading the outputs.  The GJK/EPA workspaces come from a
 * static device array of OMI_NOALLOC_EPA_SLOTS workspaces (11.9 MB per device,
 * held until the process ends); any number of pairs runs through them in turns,
 * so no pair count is refused for lack of scratch.  The array is shared by every
 * call on a device: calls that overlap on one device race.  Host checks: npairs
 * >= 0 and non-NULL buffers (OMI_ERR_ARG); the kernel gives a pair whose kind is
 * not sphere/box/convex or whose hull range leaves [0, n_hull_pts) count -1 and
 * zero outputs. */
#define OMI_NOALLOC_EPA_SLOTS 512
int32_t omi_collide_convex_noalloc(int32_t npairs, const int32_t *kind_a, const double *ca,
                                   const double *Ra, const double *ha, const int32_t *off_a,
                                   const int32_t *cnt_a, const int32_t *kind_b, const double *cb,
                                   const double *Rb, const double *hb, const int32_t *off_b,
                                   const int32_t *cnt_b, const double *hull_pts, int64_t n_hull_pts,
                                   int32_t *count, double *normal, double *point, double *depth,
                                   void *stream);

/* Replaces _solver_native.prepare_and_solve (_solver_native.pyx:179-312) with
 * the same arguments: K contacts over N bodies; pos/lv/av [N][3], nImp [K] and
 * tImp [K][2] updated in place.  The GPU keeps the reference's Gauss-Seidel
 * contact order: contacts run concurrently only when no earlier contact in
 * the sequence touches the same dynamic body, so FP64 results are
 * bit-identical to the sequential loop. */
int32_t omi_prepare_and_solve(int64_t K, int32_t N, const int64_t *a_idx, const int64_t *b_idx,
                              const double *point, const double *normal, const double *depth,
                              double *pos, double *lv, double *av, const double *inv_mass,
                              const double *invIw, const double *restitution,
                              const double *friction, double *nImp, double *tImp,
                              double restitution_threshold, int32_t iters, int32_t warm_start,
                              double correction, double slop, int32_t precision);

/* ---- 2. batches of device-resident worlds ------------------------------ */

/* One world as flat host arrays (omi_physics_amd.world.export_world_arrays). */
typedef struct {
    int32_t n;
    const double *position, *orientation, *linear_velocity, *angular_velocity;
    const double *prev_position, *prev_orientation;
    const double *inv_mass, *inv_inertia;
    const double *gravity_factor, *linear_damping, *angular_damping, *quadratic_drag;
    const double *sleep_timer;
    const int32_t *motion_type, *collider_shape, *trigger_shape;
    const uint8_t *awake;
    const int32_t *kind;
    const double *half;
    const int32_t *hull_off, *hull_cnt;
    const double *hull_pts;
    int64_t n_hull_pts;
    const int32_t *mat_slot;
    int32_t n_mat;
    const double *friction_table, *restitution_table;
    const int32_t *filt_slot;
    int32_t n_filt;
    const uint8_t *filter_table;
    double gravity[3];
    double default_linear_damping, default_angular_damping;
    int32_t sleep_enabled, velocity_iterations, has_collision;
    /* PointConstraint joints (joints.py:60-121) solved after the contacts
     * (world.py:599-611): body indices (-1 = the world), world anchors [J][3]. */
    int32_t n_joints, joint_iterations;
    const int32_t *joint_a, *joint_b;
    const double *joint_anchor;
    /* body-frame anchors [J][3] given by the caller (PointConstraint local_a /
     * local_b, joints.py:66-73): bit 0 / 1 of joint_fixed[j] marks local a / b
     * as given; the others are taken from the pose at the joint's first step.
     * All three may be NULL (none given). */
    const double *joint_local_a, *joint_local_b;
    const uint8_t *joint_fixed;
} omi_world_desc;

typedef struct {
    int32_t max_worlds;
    int32_t precision;          /* OMI_FP64 or OMI_FP32 (narrowphase + solver arithmetic) */
    int64_t max_bodies;
    int64_t max_pairs;          /* solid pairs per step, all worlds */
    int64_t max_contacts;       /* contact points per step, all worlds */
    int64_t max_hull_pts;
    int64_t max_joints;
    int32_t record_broadphase;  /* also keep the unfiltered broadphase pair list */
    int32_t solver_block;       /* threads per world in the solver launch (64..1024, 0 = auto) */
    /* Per size class (worlds of <= 256, <= 1024, <= 4096 bodies, larger):
     * threads per world for the broadphase and the solver launches, one of
     * 64, 128, 256, 512, 1024 (FP64 solver: 64, 128, 256; larger is
     * OMI_ERR_ARG); 0 picks the default.  The autotuner sets these. */
    int32_t sap_block[4];
    int32_t solve_block[4];
    /* x-window candidates of the sweep-and-prune per step, all worlds, that
     * the broadphase tiles can hold (0 = 32 * max_pairs + 2^20). */
    int64_t max_candidates;
    /* largest velocity_iterations of a world this batch will hold, 1..30
     * (0 = 16): sizes the solver's level schedule. */
    int32_t max_solver_iterations;
    /* broadphase pairs with a trigger body per step, all worlds
     * (0 = max(1024, max_pairs / 16)). */
    int64_t max_trigger_pairs;
    /* 0: a size class holding too few worlds to fill the device (fewer than
     * one per 32 compute units), each fitting the solver's LDS, is solved
     * island by island (solver.py:412-445), the largest such class only;
     * -1: every world in one solver block.  The results are the same. */
    int32_t island_solve;
    /* 1: every box/sphere pair takes the narrowphase's 16-vertex clip pass
     * (normally only a box pair whose clip outgrows 8 vertices does); for
     * testing that pass, the results are the same. */
    int32_t clip16_all;
} omi_batch_config;

typedef struct omi_batch omi_batch;

int32_t omi_batch_create(const omi_batch_config *cfg, omi_batch **out);
int32_t omi_batch_destroy(omi_batch *b);
/* Place (or replace) a world in `slot`; resets its warm-start cache.  The
 * device writes are ordered on the batch stream after the queued steps (one
 * transfer on a side stream overlapping them, one copy launch): the call does
 * not wait for the device, the descriptor's arrays may be freed on return,
 * and steps queued afterwards include the world. */
int32_t omi_batch_load_world(omi_batch *b, int32_t slot, const omi_world_desc *d);
/* Free `slot` (ordered the same way): steps queued afterwards skip its bodies,
 * and its memory is reusable at once by a later load. */
int32_t omi_batch_remove_world(omi_batch *b, int32_t slot);
/* Overwrite the mutable state of a loaded world (same body count). */
int32_t omi_batch_set_state(omi_batch *b, int32_t slot, const double *pos, const double *quat,
                            const double *lv, const double *av, const uint8_t *awake,
                            const double *sleep_timer);
/* Overwrite one world's AABBs ([n][3] each; either may be NULL). */
int32_t omi_batch_set_aabbs(omi_batch *b, int32_t slot, const double *aabb_min, const double *aabb_max);
/* Overwrite one world's per-body parameters and world settings that a host may
 * edit between steps (the reference reads them every step, backend.py:26-49,
 * solver.py:329): any pointer may be NULL to keep the value; gravity[3],
 * default_damping[2] (linear, angular); sleep_enabled and velocity_iterations
 * -1 to keep.  Motion types, shapes, materials and filters change the world's
 * layout: reload it. */
int32_t omi_batch_set_params(omi_batch *b, int32_t slot, const double *inv_mass,
                             const double *inv_inertia, const double *gravity_factor,
                             const double *linear_damping, const double *angular_damping,
                             const double *quadratic_drag, const double *gravity,
                             const double *default_damping, int32_t sleep_enabled,
                             int32_t velocity_iterations);
/* One world's warm-start cache (SequentialImpulseSolver._cache, solver.py:91,
 * 271-278): the contacts its next step seeds from, world-local, with their
 * normal and two tangent impulses.  Count first (pass NULLs), then copy up to
 * `cap`.  Batch-layer state transfer so a host can keep the cache across a
 * world reload, as the reference keeps it across edits. */
int32_t omi_batch_read_cache(omi_batch *b, int32_t slot, int64_t cap, int64_t *count, int32_t *a,
                             int32_t *bb, double *point, double *normal_impulse,
                             double *tangent_impulse);
/* Replace one world's warm-start cache with n entries (entries of one body pair
 * consecutive, as omi_batch_read_cache returns them). */
int32_t omi_batch_write_cache(omi_batch *b, int32_t slot, int64_t n, const int32_t *a,
                              const int32_t *bb, const double *point, const double *normal_impulse,
                              const double *tangent_impulse);
/* The body-frame joint anchors of one world as the device holds them: for
 * each joint, bit 0 / 1 of fixed[j] says whether local a / b is set (given at
 * load, or taken from the pose at the joint's first step, joints.py:84-101);
 * la, lb [J][3].  Any pointer may be NULL. */
int32_t omi_batch_read_joint_anchors(omi_batch *b, int32_t slot, double *la, double *lb, uint8_t *fixed);
/* Build the world lists and capture the step graph for dt without running it,
 * so the first omi_batch_step does no set-up work. */
int32_t omi_batch_prepare(omi_batch *b, double dt);
/* Enqueue nsteps fixed steps of dt for every loaded world (asynchronous). */
int32_t omi_batch_step(omi_batch *b, double dt, int32_t nsteps);
/* Enqueue the collision stage alone for every loaded world (asynchronous):
 * _CollisionStages.run (world.py:657-676) — broadphase on the AABBs as they are,
 * solid filter, narrowphase, wake, contact solve with warm start, trigger
 * overlaps — for a host that integrates forces, refits, solves joints,
 * integrates positions and updates sleep itself (the reference's
 * PhysicsWorld.step, world.py:410-430). */
int32_t omi_batch_collide(omi_batch *b, double dt);
/* Recompute AABBs only (asynchronous). */
int32_t omi_batch_refit(omi_batch *b, double margin);
/* Wait for the batch's queued work; reports capacity overflow of any step. */
int32_t omi_batch_sync(omi_batch *b);
/* Copy one world's state out (any pointer may be NULL). */
int32_t omi_batch_read_world(omi_batch *b, int32_t slot, double *pos, double *quat, double *lv,
                             double *av, uint8_t *awake, double *sleep_timer, double *aabb_min,
                             double *aabb_max, double *prev_pos, double *prev_quat);
/* Asynchronous readback (batch layer, no counterpart in the reference): gather
 * one world's state, as omi_batch_read_world returns it, after the queued
 * work, into pinned host memory of the batch's own; returns at once with a
 * ticket.  A server retiring a world fetches it, removes it and admits the
 * next world without waiting for the device. */
int32_t omi_batch_fetch_world(omi_batch *b, int32_t slot, int64_t *ticket);
/* A marker after the queued work (asynchronous); with omi_batch_mark_wait a
 * host keeps one step queued while it waits for the one before. */
int32_t omi_batch_mark(omi_batch *b, int64_t *ticket);
/* Wait until the work queued before the marker has run; reports a capacity
 * overflow of the steps up to it as omi_batch_sync does (the flags stay set
 * until omi_batch_sync clears them); the ticket is spent. */
int32_t omi_batch_mark_wait(omi_batch *b, int64_t ticket);
/* 1 when the fetch has arrived, 0 while it is in flight (negative: error). */
int32_t omi_batch_fetch_ready(omi_batch *b, int64_t ticket);
/* Wait for the fetch and copy it out (any pointer may be NULL; the layouts of
 * omi_batch_read_world); the ticket is spent. */
int32_t omi_batch_fetch_wait(omi_batch *b, int64_t ticket, double *pos, double *quat, double *lv,
                             double *av, uint8_t *awake, double *sleep_timer, double *aabb_min,
                             double *aabb_max, double *prev_pos, double *prev_quat);
/* Pose snapshot of the worlds slots[0..nslots), packed world after world:
 * position [n][3], orientation as VRML axis-angle [n][4] (x, y, z, angle), awake
 * [n] and dynamic [n] (motion type DYNAMIC), with n the listed worlds' total
 * bodies and world slots[k] at rows offsets[k] .. offsets[k+1] (offsets holds
 * nslots + 1 entries).  Pass NULL data pointers to size the buffers from the
 * offsets first.  Replaces ThreadedSimulation._publish (threaded.py:149-158),
 * quat_to_axis_angle (mathutil.py:90-105) computed on the device, one packed
 * download for every listed world. */
int32_t omi_batch_snapshot(omi_batch *b, int32_t nslots, const int32_t *slots, int64_t *offsets,
                           double *pos, double *axis_angle, uint8_t *awake, uint8_t *dynamic);
/* Trigger overlaps of one world's last step (TriggerSystem.update's overlap
 * set, triggers.py:24-70): pairs (trigger body, other body), world-local,
 * sorted; count first (pass NULLs), then copy up to `cap`.  The enter / stay /
 * exit diff against the previous step is the caller's (omi_physics_amd
 * .triggers.TriggerSystem). */
int32_t omi_batch_read_trigger_overlaps(omi_batch *b, int32_t slot, int64_t cap, int64_t *count,
                                        int32_t *trigger, int32_t *other);
/* Contacts of one world's last step in solver order: count first (pass NULLs),
 * then copy up to `cap` of them. Body indices are world-local. */
int32_t omi_batch_read_contacts(omi_batch *b, int32_t slot, int64_t cap, int64_t *count,
                                int32_t *a, int32_t *bb, double *point, double *normal,
                                double *depth, double *normal_impulse, double *tangent_impulse);
/* Pairs of one world's last step (solid = 1: the narrowphase input; 0: the
 * broadphase output, needs record_broadphase). World-local indices. */
int32_t omi_batch_read_pairs(omi_batch *b, int32_t slot, int32_t solid, int64_t cap,
                             int64_t *count, int32_t *pairs);
/* Totals of the last step over all worlds:
 * [bodies, x_candidates, broadphase pairs, solid pairs, manifolds, contacts,
 *  max levels per world, overflow flags]. */
int32_t omi_batch_counters(omi_batch *b, int64_t *out8);
/* Per-kernel device time of the last timed step, in ms, for `n` kernels in
 * the order of omi_batch_kernel_names (needs omi_batch_set_timing(b, 1)). */
int32_t omi_batch_set_timing(omi_batch *b, int32_t on);
int32_t omi_batch_kernel_times(omi_batch *b, int32_t n, double *ms);
const char *omi_batch_kernel_names(void);
/* Capture the step in a HIP graph (1) or launch kernels eagerly (0). */
int32_t omi_batch_use_graph(omi_batch *b, int32_t on);
/* Per-world figures of the last step for slots [0, n): bodies, contacts,
 * manifolds, solver levels, size class (-1 = free slot); any pointer may be
 * NULL.  Batch-layer diagnostics with no counterpart in the reference. */
int32_t omi_batch_world_stats(omi_batch *b, int32_t n, int32_t *bodies, int32_t *contacts,
                              int32_t *manifolds, int32_t *levels, int32_t *size_class);
/* The size class solved island by island (-1: none) and the islands of its
 * worlds in the last step.  Batch-layer diagnostics with no counterpart in
 * the reference. */
int32_t omi_batch_islands(omi_batch *b, int32_t *size_class, int64_t *islands);
/* Threads per world of the broadphase sort and of the solver for each size
 * class (as in omi_batch_config; 0 keeps the current value).  Batch-layer
 * tuning knob with no counterpart in the reference; the autotuner uses it. */
int32_t omi_batch_set_blocks(omi_batch *b, const int32_t *sap_block, const int32_t *solve_block);
/* Device bytes held by the batch. */
int64_t omi_batch_device_bytes(const omi_batch *b);

/* ---- groups: batches stepped together ---------------------------------
 * Batch-layer scheduling with no counterpart in the reference (which steps one
 * world at a time, world.py:410-430).  A group steps its member batches side
 * by side: each step forks from the group's stream to every member's stream
 * (one hardware queue each), launches the member's captured step graph there
 * (its kernels eagerly when it times them) and joins back, so a member holding
 * a few large worlds runs its solver while another member's chip-wide stages
 * run over many small worlds.
 * Results are those of stepping each member alone.  The group does not own
 * its members: destroy the group before them.  Member reads, state changes
 * and loads stay on the member's own calls and are ordered after the group's
 * queued steps. */
typedef struct omi_group omi_group;
int32_t omi_group_create(omi_batch *const *members, int32_t n, omi_group **out);
int32_t omi_group_destroy(omi_group *g);
/* Capture the group's step graph for dt ahead of the first step. */
int32_t omi_group_prepare(omi_group *g, double dt);
/* Enqueue nsteps steps of dt for every member (asynchronous). */
int32_t omi_group_step(omi_group *g, double dt, int32_t nsteps);
/* Wait for the group's queued steps; reports the first member whose capacity
 * overflowed (every member's flags are read and cleared). */
int32_t omi_group_sync(omi_group *g);

#ifdef __cplusplus
}
#endif
#endif
