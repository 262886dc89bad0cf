/* oracle.h — CPU restatement of the omi_physics contact pipeline.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker that the HIP path is
 * compared against; it is loaded by tests/, by __graft_entry__.smoke() and by
 * bench.py's CPU-baseline leg, and by nothing else.  It is never part of the
 * product path.
 *
 * Everything is FP64, single-threaded per world, and follows the reference's
 * accelerated path operation by operation:
 *   - Cython loops  : _collide_native.pyx:16-344, _solver_native.pyx:132-312
 *   - NumPy glue    : world.py:410-567 (step, refit, sleep), backend.py:26-49,
 *                     broadphase.py:243-301, narrowphase.py:51-101,
 *                     solver.py:192-278,321-337, mathutil.py:21-41,59-76,108-113
 *   - GJK/EPA       : gjk.py (whole), body.py:33-132 (support proxies)
 * NumPy reduction orders are the ones this host's NumPy 2.2 uses:
 *   np.linalg.norm / np.sum over an axis of length 3 or 4 add left to right;
 *   einsum('nij,nj->ni') and einsum('nij,nkj->nik') over j=0..2 compute
 *   (p0 + p2) + p1.
 * Compile with -ffp-contract=off: the reference's x86-64 C and NumPy loops do
 * not fuse multiply-adds.
 */
#ifndef OMI_ORACLE_H
#define OMI_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* One world, flat arrays, row-major [n][3] / [n][4] (same layout as
 * omi_physics_amd.world.export_world_arrays). */
typedef struct {
    int32_t n;
    const double *position, *orientation, *linear_velocity, *angular_velocity;
    const double *inv_mass, *inv_inertia;
    const double *gravity_factor, *linear_damping, *angular_damping, *quadratic_drag;
    const double *sleep_timer;
    const int32_t *motion_type, *collider_shape, *trigger_shape;
    const uint8_t *awake;
    const int32_t *kind;             /* 0 none, 1 sphere, 2 box, 3 convex */
    const double *half;              /* [n][3] */
    const int32_t *hull_off, *hull_cnt;
    const double *hull_pts;          /* [H][3] local points */
    const int32_t *mat_slot; int32_t n_mat;
    const double *friction_table, *restitution_table;   /* [n_mat][n_mat] */
    const int32_t *filt_slot; int32_t n_filt;
    const uint8_t *filter_table;                         /* [n_filt][n_filt] */
    double gravity[3];
    double default_linear_damping, default_angular_damping;
    int32_t sleep_enabled, velocity_iterations, has_collision;
    /* point (ball) joints, joints.py:60-121: body indices (-1 = the world) and
     * the world anchor given at construction */
    int32_t n_joints, joint_iterations;
    const int32_t *joint_a, *joint_b;
    const double *joint_anchor;
    /* body-frame anchors given at construction (joints.py:66-73), [J][3]: bit
     * 0 / 1 of joint_fixed[j] marks local a / b as given; NULL = none given */
    const double *joint_local_a, *joint_local_b;
    const uint8_t *joint_fixed;
} ow_desc;

typedef struct OWorld OWorld;

OWorld *ow_create(const ow_desc *d);
void ow_destroy(OWorld *w);
void ow_step(OWorld *w, double dt);
int32_t ow_body_count(const OWorld *w);
void ow_get_state(const OWorld *w, double *pos, double *quat, double *lv, double *av,
                  uint8_t *awake, double *sleep_timer, double *lo, double *hi);
void ow_set_params(OWorld *w, const double *inv_mass, const double *inv_inertia, const double *gf,
                   const double *ld, const double *ad, const double *qd, const double *g3, const double *damp2,
                   int32_t sleep_enabled, int32_t iters);
void ow_set_state(OWorld *w, const double *pos, const double *quat, const double *lv, const double *av,
                  const uint8_t *awake, const double *timer);
void ow_set_velocity(OWorld *w, const double *lv, const double *av);
int32_t ow_num_contacts(const OWorld *w);
void ow_get_contacts(const OWorld *w, int32_t *a, int32_t *b, double *point, double *normal,
                     double *depth, double *nimp, double *timp);
int32_t ow_num_pairs(const OWorld *w, int32_t solid);
void ow_get_pairs(const OWorld *w, int32_t solid, int32_t *pairs);
/* counters of the last step: [pairs, solid pairs, x-candidates, manifolds, contacts] */
void ow_counters(const OWorld *w, int64_t *out5);
int64_t ow_num_trigger_overlaps(const OWorld *w);
void ow_get_trigger_overlaps(const OWorld *w, int32_t *out);

/* ---- stage-level restatements (each cites its reference function) ---- */
void ow_quat_to_matrix(const double *q, double *R);
void ow_capsule_triangles(int64_t n, const double *c0, const double *c1, double radius, const double *verts,
                          uint8_t *hit, double *points, double *normals, double *depths);
void ow_snapshot(int32_t n, const double *pos, const double *quat, const uint8_t *awake,
                 const int32_t *motion_type, double *out_pos, double *axis_angle, uint8_t *out_awake,
                 uint8_t *dynamic);
void ow_refit(int32_t n, const double *pos, const double *quat, const int32_t *kind,
              const double *half, const int32_t *hull_off, const int32_t *hull_cnt,
              const double *hull_pts, double margin, double *lo, double *hi);
int64_t ow_sap_pairs(int32_t n, const double *lo, const double *hi, const int32_t *motion_type,
                     const int32_t *filt_slot, int32_t n_filt, const uint8_t *filter_table,
                     int32_t *pairs, int64_t cap, int64_t *x_candidates);
int32_t ow_box_box(const double *ca, const double *Ra, const double *ha,
                   const double *cb, const double *Rb, const double *hb,
                   double *normal, double *pts, double *depths);
int32_t ow_sphere_sphere(const double *ca, double ra, const double *cb, double rb,
                         double *normal, double *pt, double *depth);
int32_t ow_sphere_box(const double *cs, double rs, const double *cb, const double *Rb,
                      const double *hb, double *normal, double *pt, double *depth);
int32_t ow_collide_convex(int32_t ka, const double *ca, const double *Ra, const double *ha,
                          const double *pts_a, int32_t na,
                          int32_t kb, const double *cb, const double *Rb, const double *hb,
                          const double *pts_b, int32_t nb,
                          double *normal, double *pt, double *depth);
int32_t ow_gjk_intersect(int32_t ka, const double *ca, const double *Ra, const double *ha,
                         const double *pts_a, int32_t na,
                         int32_t kb, const double *cb, const double *Rb, const double *hb,
                         const double *pts_b, int32_t nb);
void ow_prepare_and_solve(int64_t K, const int64_t *a_idx, const int64_t *b_idx,
                          const double *point, const double *normal, const double *depth,
                          double *pos, double *lv, double *av, const double *inv_mass,
                          const double *invIw, const double *restitution, const double *friction,
                          double *nImp, double *tImp, double restitution_threshold,
                          int32_t iters, int32_t warm_start, double correction, double slop);
/* NumpyBackend.integrate_forces (backend.py:26-40); gravity_stride 0 applies
 * gravity[3] to every body, 3 reads per-body rows [n][3]. */
void ow_integrate_forces(int32_t n, const int32_t *motion_type, const uint8_t *awake,
                         const double *gravity, int32_t gravity_stride, const double *gravity_factor,
                         const double *linear_damping, const double *angular_damping,
                         const double *quadratic_drag, double dlin, double dang,
                         double *lv, double *av, double dt);
/* 3x3 solve with LU and partial pivoting (LAPACK dgetf2 + dgetrs order), as
 * np.linalg.solve does for PointConstraint.solve (joints.py:119). */
void ow_solve3(const double *K, const double *rhs, double *x);
void ow_integrate_positions(int32_t n, const int32_t *motion_type, const uint8_t *awake,
                            const double *lv, const double *av, double *pos, double *quat,
                            double dt);
void ow_update_sleep(int32_t n, int64_t K, const int32_t *ca, const int32_t *cb, const int32_t *motion_type,
                     double dt, double lin_thresh, double ang_thresh, double t_sleep, double *lv, double *av,
                     uint8_t *awake, double *sleep_timer);

#ifdef __cplusplus
}
#endif
#endif
