/* oracle.c — CPU restatement of the omi_physics contact pipeline (FP64).
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP path, the CPU baseline of
 * bench.py, never the product.  Every function names the reference code it
 * restates.  Build: see oracle/Makefile (gcc -O2 -ffp-contract=off).
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define EPS_COLLIDE 1e-9  /* collide.py:34, _collide_native.pyx:13 */
#define EPS_GJK 1e-10     /* gjk.py:19 */

enum { K_NONE = 0, K_SPHERE = 1, K_BOX = 2, K_CONVEX = 3 };
enum { M_STATIC = 0, M_KINEMATIC = 1, M_DYNAMIC = 2 };

static double dmax(double a, double b) { return a > b ? a : b; }

/* ---------------------------------------------------------------- math -- */

/* ---------------------------------------------------- capsule/triangles -- */

static double cdot3(const double *a, const double *b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }

/* _collide_native._closest_on_tri (_collide_native.pyx:361-402): closest point
 * on triangle abc to p by Ericson's Voronoi regions. */
static void closest_on_tri(const double *p, const double *a, const double *b, const double *c, double *out)
{
    double ab[3], ac[3], ap[3], bp[3], cp[3], bc[3];
    for (int m = 0; m < 3; m++) { ab[m] = b[m] - a[m]; ac[m] = c[m] - a[m]; ap[m] = p[m] - a[m]; }
    double d1 = cdot3(ab, ap), d2 = cdot3(ac, ap);
    if (d1 <= 0.0 && d2 <= 0.0) { memcpy(out, a, 24); return; }
    for (int m = 0; m < 3; m++) bp[m] = p[m] - b[m];
    double d3 = cdot3(ab, bp), d4 = cdot3(ac, bp);
    if (d3 >= 0.0 && d4 <= d3) { memcpy(out, b, 24); return; }
    double vc = d1 * d4 - d3 * d2;
    if (vc <= 0.0 && d1 >= 0.0 && d3 <= 0.0) {
        double v = d1 / (d1 - d3);
        for (int m = 0; m < 3; m++) out[m] = a[m] + v * ab[m];
        return;
    }
    for (int m = 0; m < 3; m++) cp[m] = p[m] - c[m];
    double d5 = cdot3(ab, cp), d6 = cdot3(ac, cp);
    if (d6 >= 0.0 && d5 <= d6) { memcpy(out, c, 24); return; }
    double vb = d5 * d2 - d1 * d6;
    if (vb <= 0.0 && d2 >= 0.0 && d6 <= 0.0) {
        double w = d2 / (d2 - d6);
        for (int m = 0; m < 3; m++) out[m] = a[m] + w * ac[m];
        return;
    }
    double va = d3 * d6 - d5 * d4;
    if (va <= 0.0 && (d4 - d3) >= 0.0 && (d5 - d6) >= 0.0) {
        double w = (d4 - d3) / ((d4 - d3) + (d5 - d6));
        for (int m = 0; m < 3; m++) { bc[m] = c[m] - b[m]; out[m] = b[m] + w * bc[m]; }
        return;
    }
    double denom = 1.0 / (va + vb + vc);
    double v = vb * denom, w = vc * denom;
    for (int m = 0; m < 3; m++) out[m] = a[m] + ab[m] * v + ac[m] * w;
}

static double clamp01(double x) { return x < 0.0 ? 0.0 : (x > 1.0 ? 1.0 : x); }

/* _collide_native._closest_seg_seg (_collide_native.pyx:405-451). */
static void closest_seg_seg(const double *p1, const double *q1, const double *p2, const double *q2,
                            double *o1, double *o2)
{
    double d1[3], d2[3], r[3], s, t;
    for (int m = 0; m < 3; m++) { d1[m] = q1[m] - p1[m]; d2[m] = q2[m] - p2[m]; r[m] = p1[m] - p2[m]; }
    double aa = cdot3(d1, d1), e = cdot3(d2, d2), f = cdot3(d2, r);
    if (aa <= EPS_COLLIDE && e <= EPS_COLLIDE) { memcpy(o1, p1, 24); memcpy(o2, p2, 24); return; }
    if (aa <= EPS_COLLIDE) {
        s = 0.0;
        t = clamp01(f / e);
    } else {
        double cc = cdot3(d1, r);
        if (e <= EPS_COLLIDE) {
            t = 0.0;
            s = clamp01(-cc / aa);
        } else {
            double bb = cdot3(d1, d2), denom = aa * e - bb * bb;
            s = denom > EPS_COLLIDE ? clamp01((bb * f - cc * e) / denom) : 0.0;
            t = (bb * s + f) / e;
            if (t < 0.0) { t = 0.0; s = clamp01(-cc / aa); }
            else if (t > 1.0) { t = 1.0; s = clamp01((bb - cc) / aa); }
        }
    }
    for (int m = 0; m < 3; m++) { o1[m] = p1[m] + d1[m] * s; o2[m] = p2[m] + d2[m] * t; }
}

/* _collide_native.capsule_triangles (_collide_native.pyx:454-574): the capsule
 * (segment c0 c1, radius) against n triangles verts[n][3][3].  Face contact
 * first -- the side is the capsule centre's, the depth unbounded: the cap
 * endpoint whose foot lies on the triangle and reaches deepest -- else the
 * nearest feature pair (endpoint to triangle, axis to each edge) inside the
 * radius, the normal from the triangle to the capsule. */
void ow_capsule_triangles(int64_t n, const double *c0, const double *c1, double radius, const double *verts,
                          uint8_t *hit, double *points, double *normals, double *depths)
{
    double centre[3];
    const double *caps[2] = {c0, c1};
    for (int m = 0; m < 3; m++) centre[m] = (c0[m] + c1[m]) * 0.5;
    for (int64_t i = 0; i < n; i++) {
        const double *v0 = verts + 9 * i, *v1 = v0 + 3, *v2 = v0 + 6;
        double e1[3], e2[3], face[3], fn[3], sp[3], tp[3], foot[3], bestpt[3], ap[3], ep[3];
        double bsp[3] = {0.0, 0.0, 0.0}, btp[3] = {0.0, 0.0, 0.0};
        for (int m = 0; m < 3; m++) { e1[m] = v1[m] - v0[m]; e2[m] = v2[m] - v0[m]; }
        face[0] = e1[1] * e2[2] - e1[2] * e2[1];
        face[1] = e1[2] * e2[0] - e1[0] * e2[2];
        face[2] = e1[0] * e2[1] - e1[1] * e2[0];
        double face_len = sqrt(cdot3(face, face));
        int face_found = 0;
        double best_reach = -1e300;
        if (face_len > EPS_COLLIDE) {
            for (int m = 0; m < 3; m++) fn[m] = face[m] / face_len;
            double gx = centre[0] - v0[0], gy = centre[1] - v0[1], gz = centre[2] - v0[2];
            if (fn[0] * gx + fn[1] * gy + fn[2] * gz < 0.0)
                for (int m = 0; m < 3; m++) fn[m] = -fn[m];
            for (int k = 0; k < 2; k++) {
                for (int m = 0; m < 3; m++) sp[m] = caps[k][m];
                gx = sp[0] - v0[0]; gy = sp[1] - v0[1]; gz = sp[2] - v0[2];
                double height = fn[0] * gx + fn[1] * gy + fn[2] * gz;
                for (int m = 0; m < 3; m++) foot[m] = sp[m] - fn[m] * height;
                closest_on_tri(sp, v0, v1, v2, tp);
                gx = foot[0] - tp[0]; gy = foot[1] - tp[1]; gz = foot[2] - tp[2];
                if (gx * gx + gy * gy + gz * gz > EPS_COLLIDE * EPS_COLLIDE) continue;
                double reach = radius - height;
                if (reach > 0.0 && reach > best_reach) {
                    best_reach = reach;
                    face_found = 1;
                    memcpy(bestpt, tp, 24);
                }
            }
        }
        if (face_found) {
            hit[i] = 1;
            depths[i] = best_reach;
            for (int m = 0; m < 3; m++) { points[3 * i + m] = bestpt[m]; normals[3 * i + m] = fn[m]; }
            continue;
        }
        double best_d2 = 1e300;
        for (int k = 0; k < 2; k++) {
            for (int m = 0; m < 3; m++) sp[m] = caps[k][m];
            closest_on_tri(sp, v0, v1, v2, tp);
            double gx = sp[0] - tp[0], gy = sp[1] - tp[1], gz = sp[2] - tp[2];
            double d2 = gx * gx + gy * gy + gz * gz;
            if (d2 < best_d2) { best_d2 = d2; memcpy(bsp, sp, 24); memcpy(btp, tp, 24); }
        }
        for (int k = 0; k < 3; k++) {
            const double *pa = k == 0 ? v0 : (k == 1 ? v1 : v2), *pb = k == 0 ? v1 : (k == 1 ? v2 : v0);
            closest_seg_seg(c0, c1, pa, pb, ap, ep);
            double gx = ap[0] - ep[0], gy = ap[1] - ep[1], gz = ap[2] - ep[2];
            double d2 = gx * gx + gy * gy + gz * gz;
            if (d2 < best_d2) { best_d2 = d2; memcpy(bsp, ap, 24); memcpy(btp, ep, 24); }
        }
        if (best_d2 >= radius * radius) { hit[i] = 0; continue; }
        double dist = sqrt(best_d2), nx, ny, nz;
        if (dist > EPS_COLLIDE) {
            nx = (bsp[0] - btp[0]) / dist; ny = (bsp[1] - btp[1]) / dist; nz = (bsp[2] - btp[2]) / dist;
        } else if (face_len > EPS_COLLIDE) {
            nx = face[0] / face_len; ny = face[1] / face_len; nz = face[2] / face_len;
        } else {
            nx = ny = nz = 0.0;
        }
        hit[i] = 1;
        depths[i] = radius - dist;
        for (int m = 0; m < 3; m++) points[3 * i + m] = btp[m];
        normals[3 * i] = nx; normals[3 * i + 1] = ny; normals[3 * i + 2] = nz;
    }
}

/* ThreadedSimulation._publish (threaded.py:149-158): the pose snapshot of n
 * bodies.  Orientation by quat_to_axis_angle (mathutil.py:90-105): normalise
 * (norm added left to right, divided by max(norm, 1e-12)), clip w to [-1, 1],
 * s = sqrt(max(1 - w w, 0)); s < 1e-9 gives (0, 1, 0, 0), else (q_xyz / s,
 * 2 acos w).  dynamic = motion type DYNAMIC. */
void ow_snapshot(int32_t n, const double *pos, const double *quat, const uint8_t *awake,
                 const int32_t *motion_type, double *out_pos, double *axis_angle, uint8_t *out_awake,
                 uint8_t *dynamic)
{
    for (int32_t i = 0; i < n; i++) {
        const double *q = quat + 4 * i;
        double nn = sqrt(((q[0] * q[0] + q[1] * q[1]) + q[2] * q[2]) + q[3] * q[3]);
        double d = dmax(nn, 1e-12);
        double x = q[0] / d, y = q[1] / d, z = q[2] / d, w = q[3] / d;
        if (w < -1.0) w = -1.0;
        if (w > 1.0) w = 1.0;
        double s = sqrt(dmax(1.0 - w * w, 0.0));
        double *o = axis_angle + 4 * i;
        if (s < 1e-9) {
            o[0] = 0.0; o[1] = 1.0; o[2] = 0.0; o[3] = 0.0;
        } else {
            o[0] = x / s; o[1] = y / s; o[2] = z / s; o[3] = 2.0 * acos(w);
        }
        for (int c = 0; c < 3; c++) out_pos[3 * i + c] = pos[3 * i + c];
        out_awake[i] = awake[i];
        dynamic[i] = motion_type[i] == M_DYNAMIC;
    }
}

/* mathutil.quat_to_matrix (mathutil.py:59-76) after quat_normalize
 * (mathutil.py:21-27): norm adds the squares left to right. R row-major,
 * columns are the body axes. */
void ow_quat_to_matrix(const double *q, double *R)
{
    double n = sqrt(((q[0] * q[0] + q[1] * q[1]) + q[2] * q[2]) + q[3] * q[3]);
    double d = dmax(n, 1e-12);
    double x = q[0] / d, y = q[1] / d, z = q[2] / d, w = q[3] / d;
    double xx = x * x, yy = y * y, zz = z * z;
    double xy = x * y, xz = x * z, yz = y * z;
    double wx = w * x, wy = w * y, wz = w * z;
    R[0] = 1.0 - 2.0 * (yy + zz);
    R[1] = 2.0 * (xy - wz);
    R[2] = 2.0 * (xz + wy);
    R[3] = 2.0 * (xy + wz);
    R[4] = 1.0 - 2.0 * (xx + zz);
    R[5] = 2.0 * (yz - wx);
    R[6] = 2.0 * (xz - wy);
    R[7] = 2.0 * (yz + wx);
    R[8] = 1.0 - 2.0 * (xx + yy);
}

static double dot3(const double *a, const double *b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }
static void cross3(const double *a, const double *b, double *o)
{
    double x = a[1] * b[2] - a[2] * b[1];
    double y = a[2] * b[0] - a[0] * b[2];
    double z = a[0] * b[1] - a[1] * b[0];
    o[0] = x; o[1] = y; o[2] = z;
}
static void sub3(const double *a, const double *b, double *o) { o[0] = a[0] - b[0]; o[1] = a[1] - b[1]; o[2] = a[2] - b[2]; }

/* ---------------------------------------------------------------- refit -- */

/* Convex proxy world points: local @ R.T + center (body.py:115-119). */
static void convex_world(const double *local, int32_t m, const double *R, const double *c, double *out)
{
    for (int32_t p = 0; p < m; p++) {
        const double *l = local + 3 * p;
        for (int r = 0; r < 3; r++)
            out[3 * p + r] = ((l[0] * R[3 * r + 0] + l[1] * R[3 * r + 1]) + l[2] * R[3 * r + 2]) + c[r];
    }
}

/* PhysicsWorld.refit_aabbs (world.py:471-508): boxes |R|·half via
 * einsum('nij,nj->ni') = (p0 + p2) + p1; spheres the radius; no shape a point
 * with no margin; convex the min/max of the world hull points (body.py:121-123). */
void ow_refit(int32_t n, const double *pos, const double *quat, const int32_t *kind,
              const double *half, const int32_t *hull_off, const int32_t *hull_cnt,
              const double *hull_pts, double margin, double *lo, double *hi)
{
    for (int32_t i = 0; i < n; i++) {
        const double *p = pos + 3 * i;
        double R[9], hw[3];
        ow_quat_to_matrix(quat + 4 * i, R);
        for (int r = 0; r < 3; r++) {
            double a0 = fabs(R[3 * r]) * half[3 * i], a1 = fabs(R[3 * r + 1]) * half[3 * i + 1];
            double a2 = fabs(R[3 * r + 2]) * half[3 * i + 2];
            hw[r] = (a0 + a2) + a1;
        }
        if (kind[i] == K_SPHERE) { hw[0] = half[3 * i]; hw[1] = half[3 * i + 1]; hw[2] = half[3 * i + 2]; }
        for (int r = 0; r < 3; r++) {
            lo[3 * i + r] = (p[r] - hw[r]) - margin;
            hi[3 * i + r] = (p[r] + hw[r]) + margin;
        }
        if (kind[i] == K_NONE) {
            for (int r = 0; r < 3; r++) { lo[3 * i + r] = p[r]; hi[3 * i + r] = p[r]; }
        } else if (kind[i] == K_CONVEX) {
            int32_t m = hull_cnt[i];
            double *wp = (double *)malloc(sizeof(double) * 3 * (m > 0 ? m : 1));
            convex_world(hull_pts + 3 * hull_off[i], m, R, p, wp);
            for (int r = 0; r < 3; r++) {
                double mn = wp[r], mx = wp[r];
                for (int32_t k = 1; k < m; k++) {
                    double v = wp[3 * k + r];
                    if (v < mn) mn = v;
                    if (v > mx) mx = v;
                }
                lo[3 * i + r] = mn - margin;
                hi[3 * i + r] = mx + margin;
            }
            free(wp);
        }
    }
}

/* ---------------------------------------------------------- broadphase -- */

/* Stable merge sort of body indices by lo.x (numpy argsort kind='stable'). */
static void stable_sort(int32_t *idx, int32_t *tmp, int32_t n, const double *lo)
{
    for (int32_t width = 1; width < n; width *= 2) {
        for (int32_t s = 0; s < n; s += 2 * width) {
            int32_t m = s + width < n ? s + width : n, e = s + 2 * width < n ? s + 2 * width : n;
            int32_t i = s, j = m, o = s;
            while (i < m && j < e) {
                /* take from the right run only when strictly smaller */
                if (lo[3 * idx[j]] < lo[3 * idx[i]]) tmp[o++] = idx[j++]; else tmp[o++] = idx[i++];
            }
            while (i < m) tmp[o++] = idx[i++];
            while (j < e) tmp[o++] = idx[j++];
        }
        memcpy(idx, tmp, sizeof(int32_t) * (size_t)n);
    }
}

/* BroadPhase.pairs (broadphase.py:243-294): stable argsort of lo.x,
 * ends = searchsorted(lo_sx, hi_sx, 'right'), windows [k+1, ends[k]), inclusive
 * y/z overlap, at least one motion_type == 2, (min, max) in (rank_i, rank_j)
 * order, then the collision-filter check (_filters_interact, :297-301), which
 * the table lookup makes equal to the reference's fast path too. */
int64_t ow_sap_pairs(int32_t n, const double *lo, const double *hi, const int32_t *mt,
                     const int32_t *filt_slot, int32_t n_filt, const uint8_t *ftab,
                     int32_t *pairs, int64_t cap, int64_t *x_candidates)
{
    int64_t np_ = 0, xc = 0;
    if (x_candidates) *x_candidates = 0;
    if (n < 2) return 0;
    int32_t *order = (int32_t *)malloc(sizeof(int32_t) * n);
    int32_t *tmp = (int32_t *)malloc(sizeof(int32_t) * n);
    for (int32_t i = 0; i < n; i++) order[i] = i;
    stable_sort(order, tmp, n, lo);
    free(tmp);
    for (int32_t k = 0; k < n; k++) {
        int32_t a = order[k];
        double hx = hi[3 * a];
        /* searchsorted side='right': first index whose lo.x > hx */
        int32_t lo_i = 0, hi_i = n;
        while (lo_i < hi_i) {
            int32_t mid = (lo_i + hi_i) >> 1;
            if (lo[3 * order[mid]] <= hx) lo_i = mid + 1; else hi_i = mid;
        }
        int32_t end = lo_i;
        for (int32_t j = k + 1; j < end; j++) {
            int32_t b = order[j];
            xc++;
            int keep = (hi[3 * a + 1] >= lo[3 * b + 1]) && (hi[3 * b + 1] >= lo[3 * a + 1]) &&
                       (hi[3 * a + 2] >= lo[3 * b + 2]) && (hi[3 * b + 2] >= lo[3 * a + 2]);
            keep = keep && (mt[a] == M_DYNAMIC || mt[b] == M_DYNAMIC);
            if (!keep) continue;
            int32_t p = a < b ? a : b, q = a < b ? b : a;
            if (!ftab[filt_slot[p] * n_filt + filt_slot[q]]) continue;
            if (np_ < cap) { pairs[2 * np_] = p; pairs[2 * np_ + 1] = q; }
            np_++;
        }
    }
    free(order);
    if (x_candidates) *x_candidates = xc;
    return np_;
}

/* --------------------------------------------------------- narrowphase -- */

/* _collide_native._face_axis (_collide_native.pyx:79-89) */
static int face_axis(const double *R, const double *dir, double *sign)
{
    double best = -1.0;
    int axis = 0;
    for (int m = 0; m < 3; m++) {
        double dv = R[m] * dir[0] + R[3 + m] * dir[1] + R[6 + m] * dir[2];
        if (fabs(dv) > best) { best = fabs(dv); axis = m; *sign = dv >= 0.0 ? 1.0 : -1.0; }
    }
    return axis;
}

/* _collide_native._face_vertices (_collide_native.pyx:92-110) */
static void face_vertices(const double *c, const double *R, const double *h, int axis, double sign, double *v)
{
    int o0 = axis == 0 ? 1 : 0, o1 = axis != 2 ? 2 : 1;
    if (axis == 1) { o0 = 0; o1 = 2; }
    double fc[3], u[3], w[3];
    for (int r = 0; r < 3; r++) {
        fc[r] = c[r] + R[3 * r + axis] * sign * h[axis];
        u[r] = R[3 * r + o0] * h[o0];
        w[r] = R[3 * r + o1] * h[o1];
    }
    for (int r = 0; r < 3; r++) {
        v[r] = fc[r] + u[r] + w[r];
        v[3 + r] = fc[r] - u[r] + w[r];
        v[6 + r] = fc[r] - u[r] - w[r];
        v[9 + r] = fc[r] + u[r] - w[r];
    }
}

/* _collide_native._reduce (_collide_native.pyx:212-263) */
static int reduce4(const double *P, const double *D, int nc, double *op, double *od)
{
    int keep[4], nkeep = 0;
    double best, dd, d0, ds;
    if (nc <= 4) {
        for (int i = 0; i < nc; i++) { for (int r = 0; r < 3; r++) op[3 * i + r] = P[3 * i + r]; od[i] = D[i]; }
        return nc;
    }
    keep[0] = 0; best = P[0];
    for (int i = 1; i < nc; i++) if (P[3 * i] > best) { best = P[3 * i]; keep[0] = i; }
    keep[1] = 0; best = -1.0;
    for (int i = 0; i < nc; i++) {
        dd = 0.0;
        for (int r = 0; r < 3; r++) { d0 = P[3 * i + r] - P[3 * keep[0] + r]; dd += d0 * d0; }
        if (dd > best) { best = dd; keep[1] = i; }
    }
    keep[2] = 0; best = -1.0;
    for (int i = 0; i < nc; i++) {
        ds = 0.0; dd = 0.0;
        for (int r = 0; r < 3; r++) {
            d0 = P[3 * i + r] - P[3 * keep[1] + r]; ds += d0 * d0;
            d0 = P[3 * i + r] - P[3 * keep[0] + r]; dd += d0 * d0;
        }
        if (sqrt(ds) + sqrt(dd) > best) { best = sqrt(ds) + sqrt(dd); keep[2] = i; }
    }
    keep[3] = 0; best = -1.0;
    for (int i = 0; i < nc; i++) {
        dd = 0.0;
        for (int r = 0; r < 3; r++) { d0 = P[3 * i + r] - P[3 * keep[2] + r]; dd += d0 * d0; }
        if (sqrt(dd) > best) { best = sqrt(dd); keep[3] = i; }
    }
    for (int i = 0; i < 4; i++) {
        int dup = 0;
        for (int j = 0; j < nkeep; j++) if (keep[i] == keep[j]) { dup = 1; break; }
        if (!dup) keep[nkeep++] = keep[i];
    }
    for (int i = 0; i < nkeep; i++) { for (int r = 0; r < 3; r++) op[3 * i + r] = P[3 * keep[i] + r]; od[i] = D[keep[i]]; }
    return nkeep;
}

/* _collide_native._clip (_collide_native.pyx:113-209) */
static int clip_manifold(const double *ca, const double *Ra, const double *ha,
                         const double *cb, const double *Rb, const double *hb,
                         const double *normal, double depth, double *op, double *od)
{
    double a_align = -1.0, b_align = -1.0, dv;
    for (int m = 0; m < 3; m++) {
        dv = fabs(Ra[m] * normal[0] + Ra[3 + m] * normal[1] + Ra[6 + m] * normal[2]);
        if (dv > a_align) a_align = dv;
        dv = fabs(Rb[m] * normal[0] + Rb[3 + m] * normal[1] + Rb[6 + m] * normal[2]);
        if (dv > b_align) b_align = dv;
    }
    const double *rc, *rR, *rh, *ic, *iR, *ih;
    int flip;
    if (a_align >= b_align) { rc = ca; rR = Ra; rh = ha; ic = cb; iR = Rb; ih = hb; flip = 0; }
    else { rc = cb; rR = Rb; rh = hb; ic = ca; iR = Ra; ih = ha; flip = 1; }
    double ref_n[3], neg[3];
    for (int r = 0; r < 3; r++) ref_n[r] = flip ? -normal[r] : normal[r];
    for (int r = 0; r < 3; r++) neg[r] = -ref_n[r];
    double isign = 1.0, rsign = 1.0;
    int iaxis = face_axis(iR, neg, &isign);
    double poly[16][3], buf[16][3], refv[4][3];
    int npoly = 4;
    face_vertices(ic, iR, ih, iaxis, isign, &poly[0][0]);
    int raxis = face_axis(rR, ref_n, &rsign);
    face_vertices(rc, rR, rh, raxis, rsign, &refv[0][0]);
    double fp[3] = {refv[0][0], refv[0][1], refv[0][2]};
    for (int other = 0; other < 3 && npoly > 0; other++) {
        if (other == raxis) continue;
        for (int si = 0; si < 2; si++) {
            double s = si == 0 ? 1.0 : -1.0, pn[3];
            pn[0] = rR[other] * s; pn[1] = rR[3 + other] * s; pn[2] = rR[6 + other] * s;
            double offset = (rc[0] * pn[0] + rc[1] * pn[1] + rc[2] * pn[2]) + rh[other];
            int nb = 0;
            for (int i = 0; i < npoly; i++) {
                const double *cu = poly[i], *nx = poly[(i + 1) % npoly];
                double dc = pn[0] * cu[0] + pn[1] * cu[1] + pn[2] * cu[2] - offset;
                double dn = pn[0] * nx[0] + pn[1] * nx[1] + pn[2] * nx[2] - offset;
                if (dc <= 0.0) { buf[nb][0] = cu[0]; buf[nb][1] = cu[1]; buf[nb][2] = cu[2]; nb++; }
                if (dc * dn < 0.0) {
                    double t = dc / (dc - dn);
                    for (int r = 0; r < 3; r++) buf[nb][r] = cu[r] + t * (nx[r] - cu[r]);
                    nb++;
                }
            }
            npoly = nb;
            memcpy(poly, buf, sizeof(double) * 3 * (size_t)npoly);
            if (npoly == 0) break;
        }
    }
    double allp[16 * 3], alld[16];
    int nc = 0;
    for (int i = 0; i < npoly; i++) {
        double pen = (fp[0] - poly[i][0]) * ref_n[0] + (fp[1] - poly[i][1]) * ref_n[1] +
                     (fp[2] - poly[i][2]) * ref_n[2];
        if (pen >= -1e-6) {
            for (int r = 0; r < 3; r++) allp[3 * nc + r] = poly[i][r] + ref_n[r] * pen;
            alld[nc] = pen > 0.0 ? pen : 0.0;
            nc++;
        }
    }
    if (nc == 0) {
        for (int r = 0; r < 3; r++) op[r] = 0.5 * (ca[r] + cb[r]);
        od[0] = depth;
        return 1;
    }
    return reduce4(allp, alld, nc, op, od);
}

/* _collide_native._box_box (_collide_native.pyx:20-76): 15-axis SAT, first
 * minimum overlap wins, signed by the projection of the centre offset. */
int32_t ow_box_box(const double *ca, const double *Ra, const double *ha,
                   const double *cb, const double *Rb, const double *hb,
                   double *normal, double *pts, double *depths)
{
    double axes[15][3];
    int nax = 0;
    for (int i = 0; i < 3; i++) { axes[nax][0] = Ra[i]; axes[nax][1] = Ra[3 + i]; axes[nax][2] = Ra[6 + i]; nax++; }
    for (int i = 0; i < 3; i++) { axes[nax][0] = Rb[i]; axes[nax][1] = Rb[3 + i]; axes[nax][2] = Rb[6 + i]; nax++; }
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) {
            double cx = Ra[3 + i] * Rb[6 + j] - Ra[6 + i] * Rb[3 + j];
            double cy = Ra[6 + i] * Rb[j] - Ra[i] * Rb[6 + j];
            double cz = Ra[i] * Rb[3 + j] - Ra[3 + i] * Rb[j];
            double l2 = cx * cx + cy * cy + cz * cz;
            if (l2 > EPS_COLLIDE * EPS_COLLIDE) {
                double inv = 1.0 / sqrt(l2);
                axes[nax][0] = cx * inv; axes[nax][1] = cy * inv; axes[nax][2] = cz * inv; nax++;
            }
        }
    double d[3] = {cb[0] - ca[0], cb[1] - ca[1], cb[2] - ca[2]};
    double best = 1e300, ba[3] = {0.0, 0.0, 0.0};
    normal[0] = normal[1] = normal[2] = 0.0;
    for (int k = 0; k < nax; k++) {
        const double *x = axes[k];
        double ra = fabs(x[0] * Ra[0] + x[1] * Ra[3] + x[2] * Ra[6]) * ha[0] +
                    fabs(x[0] * Ra[1] + x[1] * Ra[4] + x[2] * Ra[7]) * ha[1] +
                    fabs(x[0] * Ra[2] + x[1] * Ra[5] + x[2] * Ra[8]) * ha[2];
        double rb = fabs(x[0] * Rb[0] + x[1] * Rb[3] + x[2] * Rb[6]) * hb[0] +
                    fabs(x[0] * Rb[1] + x[1] * Rb[4] + x[2] * Rb[7]) * hb[1] +
                    fabs(x[0] * Rb[2] + x[1] * Rb[5] + x[2] * Rb[8]) * hb[2];
        double proj = x[0] * d[0] + x[1] * d[1] + x[2] * d[2];
        double ov = ra + rb - fabs(proj);
        if (ov < 0.0) return 0;
        if (ov < best) {
            best = ov;
            if (proj >= 0.0) { ba[0] = x[0]; ba[1] = x[1]; ba[2] = x[2]; }
            else { ba[0] = -x[0]; ba[1] = -x[1]; ba[2] = -x[2]; }
        }
    }
    normal[0] = ba[0]; normal[1] = ba[1]; normal[2] = ba[2];
    return clip_manifold(ca, Ra, ha, cb, Rb, hb, ba, best, pts, depths);
}

/* _collide_native.sphere_sphere (_collide_native.pyx:285-299) */
int32_t ow_sphere_sphere(const double *ca, double ra, const double *cb, double rb,
                         double *normal, double *pt, double *depth)
{
    double dx = cb[0] - ca[0], dy = cb[1] - ca[1], dz = cb[2] - ca[2];
    double dist = sqrt(dx * dx + dy * dy + dz * dz);
    double r = ra + rb;
    if (dist >= r) return 0;
    double nx, ny, nz;
    if (dist > EPS_COLLIDE) { nx = dx / dist; ny = dy / dist; nz = dz / dist; }
    else { nx = 0.0; ny = 1.0; nz = 0.0; }
    double dep = r - dist, f = ra - 0.5 * dep;
    normal[0] = nx; normal[1] = ny; normal[2] = nz;
    pt[0] = ca[0] + nx * f; pt[1] = ca[1] + ny * f; pt[2] = ca[2] + nz * f;
    *depth = dep;
    return 1;
}

/* _collide_native.sphere_box (_collide_native.pyx:302-344); the normal is
 * sphere -> box. */
int32_t ow_sphere_box(const double *cs, double rs, const double *cb, const double *R,
                      const double *hb, double *normal, double *pt, double *depth)
{
    double sx = cs[0] - cb[0], sy = cs[1] - cb[1], sz = cs[2] - cb[2];
    double lx = R[0] * sx + R[3] * sy + R[6] * sz;
    double ly = R[1] * sx + R[4] * sy + R[7] * sz;
    double lz = R[2] * sx + R[5] * sy + R[8] * sz;
    double cx = lx, cy = ly, cz = lz;
    if (cx > hb[0]) cx = hb[0]; else if (cx < -hb[0]) cx = -hb[0];
    if (cy > hb[1]) cy = hb[1]; else if (cy < -hb[1]) cy = -hb[1];
    if (cz > hb[2]) cz = hb[2]; else if (cz < -hb[2]) cz = -hb[2];
    double dx = lx - cx, dy = ly - cy, dz = lz - cz;
    double dist = sqrt(dx * dx + dy * dy + dz * dz);
    if (dist > rs) return 0;
    double nlx, nly, nlz, dep;
    if (dist > EPS_COLLIDE) {
        nlx = dx / dist; nly = dy / dist; nlz = dz / dist;
        dep = rs - dist;
    } else {
        double m0 = hb[0] - (lx >= 0 ? lx : -lx);
        double m1 = hb[1] - (ly >= 0 ? ly : -ly);
        double m2 = hb[2] - (lz >= 0 ? lz : -lz);
        int axis = 0; double mm = m0;
        if (m1 < mm) { axis = 1; mm = m1; }
        if (m2 < mm) { axis = 2; mm = m2; }
        nlx = nly = nlz = 0.0;
        if (axis == 0) nlx = lx >= 0 ? 1.0 : -1.0;
        else if (axis == 1) nly = ly >= 0 ? 1.0 : -1.0;
        else nlz = lz >= 0 ? 1.0 : -1.0;
        dep = rs + mm;
    }
    double nx = R[0] * nlx + R[1] * nly + R[2] * nlz;
    double ny = R[3] * nlx + R[4] * nly + R[5] * nlz;
    double nz = R[6] * nlx + R[7] * nly + R[8] * nlz;
    pt[0] = cb[0] + R[0] * cx + R[1] * cy + R[2] * cz;
    pt[1] = cb[1] + R[3] * cx + R[4] * cy + R[5] * cz;
    pt[2] = cb[2] + R[6] * cx + R[7] * cy + R[8] * cz;
    normal[0] = -nx; normal[1] = -ny; normal[2] = -nz;
    *depth = dep;
    return 1;
}

/* ------------------------------------------------------------- GJK/EPA -- */

typedef struct {
    int kind;              /* K_SPHERE, K_BOX, K_CONVEX */
    double c[3], R[9], h[3];
    const double *world;   /* convex: world points [m][3] */
    int m;
} Proxy;

/* SphereProxy/BoxProxy/ConvexProxy.support (body.py:46-48, 69-73, 125-128). */
static void support1(const Proxy *P, const double *d, double *out)
{
    if (P->kind == K_SPHERE) {
        double n = sqrt((d[0] * d[0] + d[1] * d[1]) + d[2] * d[2]);
        double s = dmax(n, 1e-12);
        double u[3] = {d[0] / s, d[1] / s, d[2] / s};
        for (int r = 0; r < 3; r++) out[r] = P->c[r] + P->h[0] * u[r];
    } else if (P->kind == K_BOX) {
        double v[3];
        for (int c = 0; c < 3; c++) {
            double l = (P->R[c] * d[0] + P->R[3 + c] * d[1]) + P->R[6 + c] * d[2];
            double sg = l > 0.0 ? 1.0 : (l < 0.0 ? -1.0 : 0.0);
            v[c] = sg * P->h[c];
        }
        for (int r = 0; r < 3; r++)
            out[r] = P->c[r] + ((P->R[3 * r] * v[0] + P->R[3 * r + 1] * v[1]) + P->R[3 * r + 2] * v[2]);
    } else {
        int best = 0;
        double bv = 0.0;
        for (int p = 0; p < P->m; p++) {
            const double *w = P->world + 3 * p;
            double v = (w[0] * d[0] + w[1] * d[1]) + w[2] * d[2];
            if (p == 0 || v > bv) { bv = v; best = p; }
        }
        for (int r = 0; r < 3; r++) out[r] = P->world[3 * best + r];
    }
}

/* center_hint (body.py:50-52, 75-77, 130-132): the vertex mean for convex. */
static void center_hint(const Proxy *P, double *out)
{
    if (P->kind != K_CONVEX) { out[0] = P->c[0]; out[1] = P->c[1]; out[2] = P->c[2]; return; }
    double s[3] = {0.0, 0.0, 0.0};
    for (int p = 0; p < P->m; p++) for (int r = 0; r < 3; r++) s[r] += P->world[3 * p + r];
    for (int r = 0; r < 3; r++) out[r] = s[r] / (double)P->m;
}

typedef struct { double v[3], a[3], b[3]; } Vtx;

/* gjk._support (gjk.py:32-36) */
static void msupport(const Proxy *A, const Proxy *B, const double *d, Vtx *o)
{
    double nd[3] = {-d[0], -d[1], -d[2]};
    support1(A, d, o->a);
    support1(B, nd, o->b);
    sub3(o->a, o->b, o->v);
}

static void any_perp(const double *v, double *o)   /* gjk.py:114-117 */
{
    double a[3] = {0.0, 0.0, 0.0};
    if (fabs(v[0]) < 0.9) a[0] = 1.0; else a[1] = 1.0;
    cross3(v, a, o);
}

/* simplex is s[0..n-1], newest last (the reference list order) */
static int gjk_line(Vtx *s, int *n, double *d)   /* gjk.py:69-79 */
{
    Vtx b = s[0], a = s[1];
    double ab[3], ao[3] = {-a.v[0], -a.v[1], -a.v[2]}, t[3];
    sub3(b.v, a.v, ab);
    if (dot3(ab, ao) > 0) {
        cross3(ab, ao, t); cross3(t, ab, d);
        if (dot3(d, d) < EPS_GJK) any_perp(ab, d);
        s[0] = b; s[1] = a; *n = 2;
        return 0;
    }
    s[0] = a; *n = 1;
    d[0] = ao[0]; d[1] = ao[1]; d[2] = ao[2];
    return 0;
}

static int gjk_triangle(Vtx *s, int *n, double *d)   /* gjk.py:82-95 */
{
    Vtx c = s[0], b = s[1], a = s[2];
    double ab[3], ac[3], ao[3] = {-a.v[0], -a.v[1], -a.v[2]}, abc[3], t[3];
    sub3(b.v, a.v, ab); sub3(c.v, a.v, ac);
    cross3(ab, ac, abc);
    cross3(abc, ac, t);
    if (dot3(t, ao) > 0) {
        if (dot3(ac, ao) > 0) {
            double u[3];
            cross3(ac, ao, u); cross3(u, ac, d);
            s[0] = c; s[1] = a; *n = 2;
            return 0;
        }
        s[0] = b; s[1] = a; *n = 2;
        return gjk_line(s, n, d);
    }
    cross3(ab, abc, t);
    if (dot3(t, ao) > 0) { s[0] = b; s[1] = a; *n = 2; return gjk_line(s, n, d); }
    if (dot3(abc, ao) > 0) { s[0] = c; s[1] = b; s[2] = a; *n = 3; d[0] = abc[0]; d[1] = abc[1]; d[2] = abc[2]; return 0; }
    s[0] = b; s[1] = c; s[2] = a; *n = 3;
    d[0] = -abc[0]; d[1] = -abc[1]; d[2] = -abc[2];
    return 0;
}

static int gjk_tetra(Vtx *s, int *n, double *d)   /* gjk.py:98-111 */
{
    Vtx dd = s[0], c = s[1], b = s[2], a = s[3];
    double ab[3], ac[3], ad[3], ao[3] = {-a.v[0], -a.v[1], -a.v[2]}, abc[3], acd[3], adb[3];
    sub3(b.v, a.v, ab); sub3(c.v, a.v, ac); sub3(dd.v, a.v, ad);
    cross3(ab, ac, abc); cross3(ac, ad, acd); cross3(ad, ab, adb);
    if (dot3(abc, ao) > 0) { s[0] = c; s[1] = b; s[2] = a; *n = 3; return gjk_triangle(s, n, d); }
    if (dot3(acd, ao) > 0) { s[0] = dd; s[1] = c; s[2] = a; *n = 3; return gjk_triangle(s, n, d); }
    if (dot3(adb, ao) > 0) { s[0] = b; s[1] = dd; s[2] = a; *n = 3; return gjk_triangle(s, n, d); }
    d[0] = d[1] = d[2] = 0.0;
    return 1;
}

/* gjk.gjk_intersect (gjk.py:39-57) */
static int gjk(const Proxy *A, const Proxy *B, Vtx *s, int *n)
{
    double z[3] = {0.0, 0.0, 0.0}, sa[3], sb[3], d[3];
    support1(B, z, sb); support1(A, z, sa);
    sub3(sb, sa, d);
    if (dot3(d, d) < EPS_GJK) { d[0] = 1.0; d[1] = 0.0; d[2] = 0.0; }
    double nd[3] = {-d[0], -d[1], -d[2]};
    msupport(A, B, nd, &s[0]);
    *n = 1;
    d[0] = -s[0].v[0]; d[1] = -s[0].v[1]; d[2] = -s[0].v[2];
    for (int it = 0; it < 32; it++) {
        if (dot3(d, d) < EPS_GJK) return 1;
        Vtx w;
        msupport(A, B, d, &w);
        if (dot3(w.v, d) < 0) return 0;
        s[(*n)++] = w;
        int contains;
        if (*n == 2) contains = gjk_line(s, n, d);
        else if (*n == 3) contains = gjk_triangle(s, n, d);
        else contains = gjk_tetra(s, n, d);
        if (contains) return 1;
    }
    return 0;
}

#define EPA_MAXV (4 + 64 + 1)
#define EPA_MAXF 256

/* gjk._closest_face (gjk.py:176-194) */
static int closest_face(const Vtx *v, const int (*f)[3], int nf, double *bn, double *bd)
{
    int best = -1;
    *bd = INFINITY;
    for (int i = 0; i < nf; i++) {
        double e1[3], e2[3], nn[3];
        sub3(v[f[i][1]].v, v[f[i][0]].v, e1); sub3(v[f[i][2]].v, v[f[i][0]].v, e2);
        cross3(e1, e2, nn);
        double ln = sqrt((nn[0] * nn[0] + nn[1] * nn[1]) + nn[2] * nn[2]);
        if (ln < EPS_GJK) continue;
        nn[0] /= ln; nn[1] /= ln; nn[2] /= ln;
        double d = dot3(nn, v[f[i][0]].v);
        if (d < 0) { nn[0] = -nn[0]; nn[1] = -nn[1]; nn[2] = -nn[2]; d = -d; }
        if (d < *bd) { *bd = d; bn[0] = nn[0]; bn[1] = nn[1]; bn[2] = nn[2]; best = i; }
    }
    return best;
}

/* gjk._rebuild (gjk.py:197-217): horizon edges keep the dict's insertion order. */
static int rebuild(const Vtx *v, int (*f)[3], int nf, int ni)
{
    int kept[EPA_MAXF][3];
    int ek[3 * EPA_MAXF][2], ealive[3 * EPA_MAXF], ne = 0, nk = 0;
    const double *w = v[ni].v;
    for (int i = 0; i < nf; i++) {
        const double *a = v[f[i][0]].v;
        double e1[3], e2[3], nn[3], wa[3];
        sub3(v[f[i][1]].v, a, e1); sub3(v[f[i][2]].v, a, e2); cross3(e1, e2, nn); sub3(w, a, wa);
        if (dot3(nn, wa) > 0) {
            int es[3][2] = {{f[i][0], f[i][1]}, {f[i][1], f[i][2]}, {f[i][2], f[i][0]}};
            for (int e = 0; e < 3; e++) {
                int found = -1;
                for (int q = 0; q < ne; q++)
                    if (ealive[q] && ek[q][0] == es[e][1] && ek[q][1] == es[e][0]) { found = q; break; }
                if (found >= 0) ealive[found] = 0;
                else { ek[ne][0] = es[e][0]; ek[ne][1] = es[e][1]; ealive[ne] = 1; ne++; }
            }
        } else {
            kept[nk][0] = f[i][0]; kept[nk][1] = f[i][1]; kept[nk][2] = f[i][2]; nk++;
        }
    }
    for (int q = 0; q < ne; q++)
        if (ealive[q]) {
            if (nk >= EPA_MAXF) return -1;
            kept[nk][0] = ek[q][0]; kept[nk][1] = ek[q][1]; kept[nk][2] = ni; nk++;
        }
    memcpy(f, kept, sizeof(int) * 3 * (size_t)nk);
    return nk;
}

/* gjk._contact_from_face + _barycentric (gjk.py:220-246) */
static void contact_from_face(const Vtx *v, const int *face, const double *nrm, double dist, double *point)
{
    const Vtx *a = &v[face[0]], *b = &v[face[1]], *c = &v[face[2]];
    double p[3] = {nrm[0] * dist, nrm[1] * dist, nrm[2] * dist};
    double v0[3], v1[3], v2[3], bary[3];
    sub3(b->v, a->v, v0); sub3(c->v, a->v, v1); sub3(p, a->v, v2);
    double d00 = dot3(v0, v0), d01 = dot3(v0, v1), d11 = dot3(v1, v1);
    double d20 = dot3(v2, v0), d21 = dot3(v2, v1);
    double den = d00 * d11 - d01 * d01;
    if (fabs(den) < EPS_GJK) { bary[0] = 1.0; bary[1] = 0.0; bary[2] = 0.0; }
    else {
        double vv = (d11 * d20 - d01 * d21) / den;
        double ww = (d00 * d21 - d01 * d20) / den;
        bary[0] = 1.0 - vv - ww; bary[1] = vv; bary[2] = ww;
    }
    for (int r = 0; r < 3; r++) {
        double pa = bary[0] * a->a[r] + bary[1] * b->a[r] + bary[2] * c->a[r];
        double pb = bary[0] * a->b[r] + bary[1] * b->b[r] + bary[2] * c->b[r];
        point[r] = 0.5 * (pa + pb);
    }
}

/* gjk.epa + _expand_to_tetra (gjk.py:120-173) */
static int epa(const Proxy *A, const Proxy *B, const Vtx *simplex, int ns,
               double *nrm, double *depth, double *point)
{
    Vtx v[EPA_MAXV];
    int f[EPA_MAXF][3];
    int nv = ns;
    for (int i = 0; i < ns; i++) v[i] = simplex[i];
    if (nv < 4) {
        static const double dirs[6][3] = {{1, 0, 0}, {-1, 0, 0}, {0, 1, 0}, {0, -1, 0}, {0, 0, 1}, {0, 0, -1}};
        for (int k = 0; k < 6 && nv < 4; k++) {
            Vtx w;
            msupport(A, B, dirs[k], &w);
            int ok = 1;
            for (int p = 0; p < nv; p++) {
                double e[3];
                sub3(w.v, v[p].v, e);
                if (!(dot3(e, e) > EPS_GJK)) { ok = 0; break; }
            }
            if (ok) v[nv++] = w;
        }
        if (nv < 4) return 0;
        double e1[3], e2[3], e3[3], c[3];
        sub3(v[1].v, v[0].v, e1); sub3(v[2].v, v[0].v, e2); sub3(v[3].v, v[0].v, e3);
        cross3(e1, e2, c);
        if (dot3(c, e3) > 0) { Vtx t = v[1]; v[1] = v[2]; v[2] = t; }
    }
    nv = 4;
    int init[4][3] = {{0, 1, 2}, {0, 2, 3}, {0, 3, 1}, {1, 3, 2}};
    memcpy(f, init, sizeof(init));
    int nf = 4;
    double bn[3], bd;
    for (int it = 0; it < 64; it++) {
        int fi = closest_face(v, (const int (*)[3])f, nf, bn, &bd);
        if (fi < 0) return 0;
        Vtx w;
        msupport(A, B, bn, &w);
        double d = dot3(w.v, bn);
        if (d - bd < 1e-6) {
            contact_from_face(v, f[fi], bn, bd, point);
            nrm[0] = bn[0]; nrm[1] = bn[1]; nrm[2] = bn[2]; *depth = bd;
            return 1;
        }
        v[nv++] = w;
        nf = rebuild(v, f, nf, nv - 1);
        if (nf <= 0) return 0;
    }
    int fi = closest_face(v, (const int (*)[3])f, nf, bn, &bd);
    if (fi < 0) return 0;
    contact_from_face(v, f[fi], bn, bd, point);
    nrm[0] = bn[0]; nrm[1] = bn[1]; nrm[2] = bn[2]; *depth = bd;
    return 1;
}

static void make_proxy(Proxy *P, int kind, const double *c, const double *R, const double *h,
                       const double *pts, int m, double *scratch)
{
    P->kind = kind;
    for (int r = 0; r < 3; r++) { P->c[r] = c[r]; P->h[r] = h ? h[r] : 0.0; }
    for (int r = 0; r < 9; r++) P->R[r] = R[r];
    P->world = NULL; P->m = m;
    if (kind == K_CONVEX) { convex_world(pts, m, R, c, scratch); P->world = scratch; }
}

int32_t ow_gjk_intersect(int32_t ka, const double *ca, const double *Ra, const double *ha,
                         const double *pts_a, int32_t na,
                         int32_t kb, const double *cb, const double *Rb, const double *hb,
                         const double *pts_b, int32_t nb)
{
    double wa[3 * 64], wb[3 * 64];
    Proxy A, B;
    Vtx s[4];
    int n;
    make_proxy(&A, ka, ca, Ra, ha, pts_a, na, wa);
    make_proxy(&B, kb, cb, Rb, hb, pts_b, nb, wb);
    return gjk(&A, &B, s, &n);
}

/* gjk.collide_convex (gjk.py:249-264): one contact, normal oriented A -> B by
 * the centre hints, dropped when depth <= 1e-9. */
int32_t ow_collide_convex(int32_t ka, const double *ca, const double *Ra, const double *ha,
                          const double *pts_a, int32_t na,
                          int32_t kb, const double *cb, const double *Rb, const double *hb,
                          const double *pts_b, int32_t nb,
                          double *normal, double *pt, double *depth)
{
    double wa[3 * 64], wb[3 * 64];
    Proxy A, B;
    Vtx s[4];
    int n;
    if (na > 64 || nb > 64) return -1;
    make_proxy(&A, ka, ca, Ra, ha, pts_a, na, wa);
    make_proxy(&B, kb, cb, Rb, hb, pts_b, nb, wb);
    if (!gjk(&A, &B, s, &n)) return 0;
    double nrm[3], dep, p[3];
    if (!epa(&A, &B, s, n, nrm, &dep, p)) return 0;
    double ha_[3], hb_[3], dh[3];
    center_hint(&A, ha_); center_hint(&B, hb_);
    sub3(hb_, ha_, dh);
    if (dot3(nrm, dh) < 0) { nrm[0] = -nrm[0]; nrm[1] = -nrm[1]; nrm[2] = -nrm[2]; }
    if (dep <= 1e-9) return 0;
    double ln = sqrt((nrm[0] * nrm[0] + nrm[1] * nrm[1]) + nrm[2] * nrm[2]);
    double s_ = dmax(ln, 1e-12);
    normal[0] = nrm[0] / s_; normal[1] = nrm[1] / s_; normal[2] = nrm[2] / s_;
    pt[0] = p[0]; pt[1] = p[1]; pt[2] = p[2];
    *depth = dep;
    return 1;
}

/* -------------------------------------------------------------- solver -- */

/* _solver_native._apply_w (_solver_native.pyx:132-151) */
static void apply_w(double *lv, double *av, int64_t a, int64_t b, double ima, double imb,
                    const double *invIw, const double *rA, const double *rB,
                    double Px, double Py, double Pz)
{
    double crx, cry, crz;
    const double *Ia = invIw + 9 * a, *Ib = invIw + 9 * b;
    lv[3 * a] -= ima * Px; lv[3 * a + 1] -= ima * Py; lv[3 * a + 2] -= ima * Pz;
    lv[3 * b] += imb * Px; lv[3 * b + 1] += imb * Py; lv[3 * b + 2] += imb * Pz;
    crx = rA[1] * Pz - rA[2] * Py; cry = rA[2] * Px - rA[0] * Pz; crz = rA[0] * Py - rA[1] * Px;
    av[3 * a] -= Ia[0] * crx + Ia[1] * cry + Ia[2] * crz;
    av[3 * a + 1] -= Ia[3] * crx + Ia[4] * cry + Ia[5] * crz;
    av[3 * a + 2] -= Ia[6] * crx + Ia[7] * cry + Ia[8] * crz;
    crx = rB[1] * Pz - rB[2] * Py; cry = rB[2] * Px - rB[0] * Pz; crz = rB[0] * Py - rB[1] * Px;
    av[3 * b] += Ib[0] * crx + Ib[1] * cry + Ib[2] * crz;
    av[3 * b + 1] += Ib[3] * crx + Ib[4] * cry + Ib[5] * crz;
    av[3 * b + 2] += Ib[6] * crx + Ib[7] * cry + Ib[8] * crz;
}

/* _solver_native._effmass (_solver_native.pyx:154-176) */
static double effmass(const double *ra, const double *rb, double dx, double dy, double dz,
                      const double *Ia, const double *Ib, double ima, double imb)
{
    double ax = ra[1] * dz - ra[2] * dy, ay = ra[2] * dx - ra[0] * dz, az = ra[0] * dy - ra[1] * dx;
    double bx = rb[1] * dz - rb[2] * dy, by = rb[2] * dx - rb[0] * dz, bz = rb[0] * dy - rb[1] * dx;
    double kk = ima + imb, wx, wy, wz;
    wx = Ia[0] * ax + Ia[1] * ay + Ia[2] * az;
    wy = Ia[3] * ax + Ia[4] * ay + Ia[5] * az;
    wz = Ia[6] * ax + Ia[7] * ay + Ia[8] * az;
    kk += ax * wx + ay * wy + az * wz;
    wx = Ib[0] * bx + Ib[1] * by + Ib[2] * bz;
    wy = Ib[3] * bx + Ib[4] * by + Ib[5] * bz;
    wz = Ib[6] * bx + Ib[7] * by + Ib[8] * bz;
    kk += bx * wx + by * wy + bz * wz;
    return kk > 1e-12 ? 1.0 / kk : 0.0;
}

#define RELVEL(lv, av, a, b, rA, rB, dvx, dvy, dvz)                                              \
    do {                                                                                         \
        dvx = (lv[3 * b] + av[3 * b + 1] * rB[2] - av[3 * b + 2] * rB[1]) -                      \
              (lv[3 * a] + av[3 * a + 1] * rA[2] - av[3 * a + 2] * rA[1]);                       \
        dvy = (lv[3 * b + 1] + av[3 * b + 2] * rB[0] - av[3 * b] * rB[2]) -                      \
              (lv[3 * a + 1] + av[3 * a + 2] * rA[0] - av[3 * a] * rA[2]);                       \
        dvz = (lv[3 * b + 2] + av[3 * b] * rB[1] - av[3 * b + 1] * rB[0]) -                      \
              (lv[3 * a + 2] + av[3 * a] * rA[1] - av[3 * a + 1] * rA[0]);                       \
    } while (0)

/* _solver_native.prepare_and_solve (_solver_native.pyx:179-312): prepare every
 * contact, warm start, `iters` Gauss-Seidel sweeps in contact order (normal,
 * then t1 and t2 sharing one relative velocity), then one split-impulse
 * position pass. */
void ow_prepare_and_solve(int64_t K, const int64_t *ai, const int64_t *bi,
                          const double *point, const double *normal, const double *depth,
                          double *pos, double *lv, double *av, const double *inv_mass,
                          const double *invIw, const double *rest, const double *fric,
                          double *nImp, double *tImp, double rthr,
                          int32_t iters, int32_t warm, double correction, double slop)
{
    double *rA = (double *)malloc(sizeof(double) * 3 * (size_t)(K ? K : 1));
    double *rB = (double *)malloc(sizeof(double) * 3 * (size_t)(K ? K : 1));
    double *t1 = (double *)malloc(sizeof(double) * 3 * (size_t)(K ? K : 1));
    double *t2 = (double *)malloc(sizeof(double) * 3 * (size_t)(K ? K : 1));
    double *sc = (double *)malloc(sizeof(double) * 6 * (size_t)(K ? K : 1));
    double *nM = sc, *t1M = sc + K, *t2M = sc + 2 * K, *vB = sc + 3 * K, *ima = sc + 4 * K, *imb = sc + 5 * K;
    for (int64_t k = 0; k < K; k++) {
        int64_t a = ai[k], b = bi[k];
        double ia = inv_mass[a], ib = inv_mass[b];
        ima[k] = ia; imb[k] = ib;
        for (int r = 0; r < 3; r++) { rA[3 * k + r] = point[3 * k + r] - pos[3 * a + r]; rB[3 * k + r] = point[3 * k + r] - pos[3 * b + r]; }
        double nx = normal[3 * k], ny = normal[3 * k + 1], nz = normal[3 * k + 2];
        double axx, ayy, azz;
        if ((nx >= 0 ? nx : -nx) < 0.9) { axx = 1.0; ayy = 0.0; azz = 0.0; }
        else { axx = 0.0; ayy = 1.0; azz = 0.0; }
        double tx = ayy * nz - azz * ny, ty = azz * nx - axx * nz, tz = axx * ny - ayy * nx;
        double ln = sqrt(tx * tx + ty * ty + tz * tz);
        double inv = ln > 1e-12 ? 1.0 / ln : 0.0;
        tx *= inv; ty *= inv; tz *= inv;
        t1[3 * k] = tx; t1[3 * k + 1] = ty; t1[3 * k + 2] = tz;
        t2[3 * k] = ny * tz - nz * ty; t2[3 * k + 1] = nz * tx - nx * tz; t2[3 * k + 2] = nx * ty - ny * tx;
        const double *Ia = invIw + 9 * a, *Ib = invIw + 9 * b;
        nM[k] = effmass(rA + 3 * k, rB + 3 * k, nx, ny, nz, Ia, Ib, ia, ib);
        t1M[k] = effmass(rA + 3 * k, rB + 3 * k, tx, ty, tz, Ia, Ib, ia, ib);
        t2M[k] = effmass(rA + 3 * k, rB + 3 * k, t2[3 * k], t2[3 * k + 1], t2[3 * k + 2], Ia, Ib, ia, ib);
        double dvx, dvy, dvz;
        const double *ra = rA + 3 * k, *rb = rB + 3 * k;
        RELVEL(lv, av, a, b, ra, rb, dvx, dvy, dvz);
        double vn0 = dvx * nx + dvy * ny + dvz * nz;
        vB[k] = vn0 < -rthr ? -rest[k] * vn0 : 0.0;
    }
    if (warm)
        for (int64_t k = 0; k < K; k++) {
            const double *n = normal + 3 * k, *u = t1 + 3 * k, *v = t2 + 3 * k;
            apply_w(lv, av, ai[k], bi[k], ima[k], imb[k], invIw, rA + 3 * k, rB + 3 * k,
                    nImp[k] * n[0] + tImp[2 * k] * u[0] + tImp[2 * k + 1] * v[0],
                    nImp[k] * n[1] + tImp[2 * k] * u[1] + tImp[2 * k + 1] * v[1],
                    nImp[k] * n[2] + tImp[2 * k] * u[2] + tImp[2 * k + 1] * v[2]);
        }
    for (int32_t it = 0; it < iters; it++)
        for (int64_t k = 0; k < K; k++) {
            int64_t a = ai[k], b = bi[k];
            const double *ra = rA + 3 * k, *rb = rB + 3 * k, *u = t1 + 3 * k, *v = t2 + 3 * k;
            double nx = normal[3 * k], ny = normal[3 * k + 1], nz = normal[3 * k + 2];
            double dvx, dvy, dvz;
            RELVEL(lv, av, a, b, ra, rb, dvx, dvy, dvz);
            double vn = dvx * nx + dvy * ny + dvz * nz;
            double nw = nImp[k] - nM[k] * (vn - vB[k]);
            if (nw < 0.0) nw = 0.0;
            double dL = nw - nImp[k];
            nImp[k] = nw;
            apply_w(lv, av, a, b, ima[k], imb[k], invIw, ra, rb, dL * nx, dL * ny, dL * nz);
            double maxf = fric[k] * nImp[k];
            RELVEL(lv, av, a, b, ra, rb, dvx, dvy, dvz);
            double vt = dvx * u[0] + dvy * u[1] + dvz * u[2];
            double nt = tImp[2 * k] - t1M[k] * vt;
            if (nt > maxf) nt = maxf; else if (nt < -maxf) nt = -maxf;
            dL = nt - tImp[2 * k]; tImp[2 * k] = nt;
            apply_w(lv, av, a, b, ima[k], imb[k], invIw, ra, rb, dL * u[0], dL * u[1], dL * u[2]);
            vt = dvx * v[0] + dvy * v[1] + dvz * v[2];
            nt = tImp[2 * k + 1] - t2M[k] * vt;
            if (nt > maxf) nt = maxf; else if (nt < -maxf) nt = -maxf;
            dL = nt - tImp[2 * k + 1]; tImp[2 * k + 1] = nt;
            apply_w(lv, av, a, b, ima[k], imb[k], invIw, ra, rb, dL * v[0], dL * v[1], dL * v[2]);
        }
    for (int64_t k = 0; k < K; k++) {
        double vn = depth[k] - slop;
        if (vn <= 0.0) continue;
        double ia = ima[k], ib = imb[k];
        if (ia + ib <= 1e-12) continue;
        double dL = correction * vn / (ia + ib);
        int64_t a = ai[k], b = bi[k];
        for (int r = 0; r < 3; r++) pos[3 * a + r] -= ia * dL * normal[3 * k + r];
        for (int r = 0; r < 3; r++) pos[3 * b + r] += ib * dL * normal[3 * k + r];
    }
    free(rA); free(rB); free(t1); free(t2); free(sc);
}

/* --------------------------------------------------------- integration -- */

/* NumpyBackend.integrate_forces (backend.py:26-40) over dyn & awake bodies. */
void ow_integrate_forces(int32_t n, const int32_t *mt, const uint8_t *awake, const double *g, int32_t g_stride,
                         const double *gf, const double *ld, const double *ad, const double *qd,
                         double dlin, double dang, double *lv, double *av, double dt)
{
    for (int32_t i = 0; i < n; i++) {
        if (!(mt[i] == M_DYNAMIC && awake[i])) continue;
        double *v = lv + 3 * i, *w = av + 3 * i;
        const double *gi = g + (size_t)g_stride * i;   /* stride 0: one vector for every body */
        for (int r = 0; r < 3; r++) v[r] = v[r] + (gi[r] * gf[i]) * dt;
        double fl = dmax(1.0 - (ld[i] + dlin) * dt, 0.0);
        for (int r = 0; r < 3; r++) v[r] = v[r] * fl;
        double sp = sqrt((v[0] * v[0] + v[1] * v[1]) + v[2] * v[2]);
        for (int r = 0; r < 3; r++) v[r] = v[r] - ((qd[i] * sp) * v[r]) * dt;
        double fa = dmax(1.0 - (ad[i] + dang) * dt, 0.0);
        for (int r = 0; r < 3; r++) w[r] = w[r] * fa;
    }
}

/* NumpyBackend.integrate_positions (backend.py:42-49) + quat_integrate
 * (mathutil.py:108-113) over (dyn | kin) & awake bodies. */
void ow_integrate_positions(int32_t n, const int32_t *mt, const uint8_t *awake,
                            const double *lv, const double *av, double *pos, double *quat, double dt)
{
    for (int32_t i = 0; i < n; i++) {
        if (!((mt[i] == M_DYNAMIC || mt[i] == M_KINEMATIC) && awake[i])) continue;
        for (int r = 0; r < 3; r++) pos[3 * i + r] = pos[3 * i + r] + lv[3 * i + r] * dt;
        double *q = quat + 4 * i;
        double wx = av[3 * i], wy = av[3 * i + 1], wz = av[3 * i + 2];
        double bx = q[0], by = q[1], bz = q[2], bw = q[3];
        /* quat_mul((w,0), q) with aw = 0 (mathutil.py:32-41) */
        double mx = ((0.0 * bx + wx * bw) + wy * bz) - wz * by;
        double my = ((0.0 * by - wx * bz) + wy * bw) + wz * bx;
        double mz = ((0.0 * bz + wx * by) - wy * bx) + wz * bw;
        double mw = ((0.0 * bw - wx * bx) - wy * by) - wz * bz;
        double s0 = q[0] + (0.5 * mx) * dt, s1 = q[1] + (0.5 * my) * dt;
        double s2 = q[2] + (0.5 * mz) * dt, s3 = q[3] + (0.5 * mw) * dt;
        double nn = sqrt(((s0 * s0 + s1 * s1) + s2 * s2) + s3 * s3);
        double d = dmax(nn, 1e-12);
        q[0] = s0 / d; q[1] = s1 / d; q[2] = s2 / d; q[3] = s3 / d;
    }
}

/* ------------------------------------------------------------ world -- */

typedef struct {
    int32_t a, b;
    double anchor[3], la[3], lb[3];
    int has_la, has_lb;
    /* prepared each step */
    double rA[3], rB[3], K[9], bias[3], ima, imb, Ia[9], Ib[9];
} PJoint;

struct OWorld {
    int32_t n;
    double *pos, *quat, *lv, *av, *inv_mass, *inv_inertia, *gf, *ld, *ad, *qd, *timer, *half;
    double *lo, *hi, *R, *invIw;
    int32_t *mt, *cs, *ts, *kind, *hull_off, *hull_cnt, *mat, *filt;
    uint8_t *awake;
    double *hull_pts;
    int32_t n_mat, n_filt;
    double *ftab_f, *rtab;
    uint8_t *ftab;
    double g[3], dlin, dang;
    int32_t sleep_enabled, iters, has_collision;
    /* last step */
    int32_t *pairs, *solid;
    int64_t npairs, nsolid, pair_cap, xcand, nmanifold;
    int64_t K, kcap;
    int32_t *ca, *cb;
    double *cp, *cn, *cd, *cni, *cti;
    /* warm-start cache: the contacts of the last step that had any */
    int64_t cK;
    int32_t *cca, *ccb;
    double *ccp, *ccn, *cct;
    int32_t *hkey_start;  /* open-addressing table over (a,b) -> first cache index */
    int64_t hcap;
    PJoint *joints;
    int32_t nj, jiters;
    /* trigger overlaps of the last step: (trigger body, other body) */
    int32_t *tov;
    int64_t ntov, tov_cap;
};

static void *dupmem(const void *p, size_t sz) { void *q = malloc(sz ? sz : 1); if (p && sz) memcpy(q, p, sz); return q; }

OWorld *ow_create(const ow_desc *d)
{
    OWorld *w = (OWorld *)calloc(1, sizeof(OWorld));
    int32_t n = d->n;
    size_t n3 = sizeof(double) * 3 * (size_t)n, n1 = sizeof(double) * (size_t)n, i1 = sizeof(int32_t) * (size_t)n;
    w->n = n;
    w->pos = dupmem(d->position, n3); w->quat = dupmem(d->orientation, 4 * n1);
    w->lv = dupmem(d->linear_velocity, n3); w->av = dupmem(d->angular_velocity, n3);
    w->inv_mass = dupmem(d->inv_mass, n1); w->inv_inertia = dupmem(d->inv_inertia, n3);
    w->gf = dupmem(d->gravity_factor, n1); w->ld = dupmem(d->linear_damping, n1);
    w->ad = dupmem(d->angular_damping, n1); w->qd = dupmem(d->quadratic_drag, n1);
    w->timer = dupmem(d->sleep_timer, n1); w->half = dupmem(d->half, n3);
    w->mt = dupmem(d->motion_type, i1); w->cs = dupmem(d->collider_shape, i1);
    w->ts = dupmem(d->trigger_shape, i1); w->kind = dupmem(d->kind, i1);
    w->hull_off = dupmem(d->hull_off, i1); w->hull_cnt = dupmem(d->hull_cnt, i1);
    w->mat = dupmem(d->mat_slot, i1); w->filt = dupmem(d->filt_slot, i1);
    w->awake = dupmem(d->awake, (size_t)n);
    int64_t H = 0;
    for (int32_t i = 0; i < n; i++) if (d->kind[i] == K_CONVEX && d->hull_off[i] + d->hull_cnt[i] > H) H = d->hull_off[i] + d->hull_cnt[i];
    w->hull_pts = dupmem(d->hull_pts, sizeof(double) * 3 * (size_t)H);
    w->n_mat = d->n_mat; w->n_filt = d->n_filt;
    w->ftab_f = dupmem(d->friction_table, sizeof(double) * (size_t)(d->n_mat * d->n_mat));
    w->rtab = dupmem(d->restitution_table, sizeof(double) * (size_t)(d->n_mat * d->n_mat));
    w->ftab = dupmem(d->filter_table, (size_t)(d->n_filt * d->n_filt));
    for (int r = 0; r < 3; r++) w->g[r] = d->gravity[r];
    w->dlin = d->default_linear_damping; w->dang = d->default_angular_damping;
    w->sleep_enabled = d->sleep_enabled; w->iters = d->velocity_iterations; w->has_collision = d->has_collision;
    w->lo = (double *)calloc((size_t)(3 * n + 1), sizeof(double));
    w->hi = (double *)calloc((size_t)(3 * n + 1), sizeof(double));
    w->R = (double *)calloc((size_t)(9 * n + 1), sizeof(double));
    w->invIw = (double *)calloc((size_t)(9 * n + 1), sizeof(double));
    w->nj = d->n_joints;
    w->jiters = d->joint_iterations;
    w->joints = (PJoint *)calloc((size_t)(w->nj > 0 ? w->nj : 1), sizeof(PJoint));
    for (int32_t j = 0; j < w->nj; j++) {
        w->joints[j].a = d->joint_a[j];
        w->joints[j].b = d->joint_b[j];
        for (int r = 0; r < 3; r++) w->joints[j].anchor[r] = d->joint_anchor[3 * j + r];
        /* PointConstraint(local_a=..., local_b=...): a given anchor is used from
         * the first prepare on (joints.py:93-96) */
        const uint8_t fx = d->joint_fixed ? d->joint_fixed[j] : 0;
        if ((fx & 1) && d->joint_local_a) {
            w->joints[j].has_la = 1;
            for (int r = 0; r < 3; r++) w->joints[j].la[r] = d->joint_local_a[3 * j + r];
        }
        if ((fx & 2) && d->joint_local_b) {
            w->joints[j].has_lb = 1;
            for (int r = 0; r < 3; r++) w->joints[j].lb[r] = d->joint_local_b[3 * j + r];
        }
    }
    w->pair_cap = 1024;
    w->pairs = (int32_t *)malloc(sizeof(int32_t) * 2 * (size_t)w->pair_cap);
    w->solid = (int32_t *)malloc(sizeof(int32_t) * 2 * (size_t)w->pair_cap);
    return w;
}

void ow_destroy(OWorld *w)
{
    if (!w) return;
    void *ps[] = {w->pos, w->quat, w->lv, w->av, w->inv_mass, w->inv_inertia, w->gf, w->ld, w->ad, w->qd,
                  w->timer, w->half, w->lo, w->hi, w->R, w->invIw, w->mt, w->cs, w->ts, w->kind,
                  w->hull_off, w->hull_cnt, w->mat, w->filt, w->awake, w->hull_pts, w->ftab_f, w->rtab,
                  w->ftab, w->pairs, w->solid, w->ca, w->cb, w->cp, w->cn, w->cd, w->cni, w->cti,
                  w->cca, w->ccb, w->ccp, w->ccn, w->cct, w->hkey_start, w->joints, w->tov};
    for (size_t i = 0; i < sizeof(ps) / sizeof(ps[0]); i++) free(ps[i]);
    free(w);
}

int32_t ow_body_count(const OWorld *w) { return w->n; }

void ow_get_state(const OWorld *w, double *pos, double *quat, double *lv, double *av,
                  uint8_t *awake, double *timer, double *lo, double *hi)
{
    size_t n = (size_t)w->n;
    if (pos) memcpy(pos, w->pos, sizeof(double) * 3 * n);
    if (quat) memcpy(quat, w->quat, sizeof(double) * 4 * n);
    if (lv) memcpy(lv, w->lv, sizeof(double) * 3 * n);
    if (av) memcpy(av, w->av, sizeof(double) * 3 * n);
    if (awake) memcpy(awake, w->awake, n);
    if (timer) memcpy(timer, w->timer, sizeof(double) * n);
    if (lo) memcpy(lo, w->lo, sizeof(double) * 3 * n);
    if (hi) memcpy(hi, w->hi, sizeof(double) * 3 * n);
}

/* Host edits between steps: any pointer NULL keeps the value; sleep_enabled
 * and iters -1 keep theirs. */
void ow_set_params(OWorld *w, const double *inv_mass, const double *inv_inertia, const double *gf,
                   const double *ld, const double *ad, const double *qd, const double *g3, const double *damp2,
                   int32_t sleep_enabled, int32_t iters)
{
    const size_t n = (size_t)w->n;
    if (inv_mass) memcpy(w->inv_mass, inv_mass, sizeof(double) * n);
    if (inv_inertia) memcpy(w->inv_inertia, inv_inertia, sizeof(double) * 3 * n);
    if (gf) memcpy(w->gf, gf, sizeof(double) * n);
    if (ld) memcpy(w->ld, ld, sizeof(double) * n);
    if (ad) memcpy(w->ad, ad, sizeof(double) * n);
    if (qd) memcpy(w->qd, qd, sizeof(double) * n);
    if (g3) for (int r = 0; r < 3; r++) w->g[r] = g3[r];
    if (damp2) { w->dlin = damp2[0]; w->dang = damp2[1]; }
    if (sleep_enabled >= 0) w->sleep_enabled = sleep_enabled;
    if (iters >= 0) w->iters = iters;
}

void ow_set_state(OWorld *w, const double *pos, const double *quat, const double *lv, const double *av,
                  const uint8_t *awake, const double *timer)
{
    const size_t n = (size_t)w->n;
    if (pos) memcpy(w->pos, pos, sizeof(double) * 3 * n);
    if (quat) memcpy(w->quat, quat, sizeof(double) * 4 * n);
    if (lv) memcpy(w->lv, lv, sizeof(double) * 3 * n);
    if (av) memcpy(w->av, av, sizeof(double) * 3 * n);
    if (awake) memcpy(w->awake, awake, n);
    if (timer) memcpy(w->timer, timer, sizeof(double) * n);
}

void ow_set_velocity(OWorld *w, const double *lv, const double *av)
{
    memcpy(w->lv, lv, sizeof(double) * 3 * (size_t)w->n);
    memcpy(w->av, av, sizeof(double) * 3 * (size_t)w->n);
}

int32_t ow_num_contacts(const OWorld *w) { return (int32_t)w->K; }

void ow_get_contacts(const OWorld *w, int32_t *a, int32_t *b, double *point, double *normal,
                     double *depth, double *nimp, double *timp)
{
    size_t K = (size_t)w->K;
    if (a) memcpy(a, w->ca, sizeof(int32_t) * K);
    if (b) memcpy(b, w->cb, sizeof(int32_t) * K);
    if (point) memcpy(point, w->cp, sizeof(double) * 3 * K);
    if (normal) memcpy(normal, w->cn, sizeof(double) * 3 * K);
    if (depth) memcpy(depth, w->cd, sizeof(double) * K);
    if (nimp) memcpy(nimp, w->cni, sizeof(double) * K);
    if (timp) memcpy(timp, w->cti, sizeof(double) * 2 * K);
}

int32_t ow_num_pairs(const OWorld *w, int32_t solid) { return (int32_t)(solid ? w->nsolid : w->npairs); }

void ow_get_pairs(const OWorld *w, int32_t solid, int32_t *pairs)
{
    size_t np_ = (size_t)(solid ? w->nsolid : w->npairs);
    memcpy(pairs, solid ? w->solid : w->pairs, sizeof(int32_t) * 2 * np_);
}

void ow_counters(const OWorld *w, int64_t *o)
{
    o[0] = w->npairs; o[1] = w->nsolid; o[2] = w->xcand; o[3] = w->nmanifold; o[4] = w->K;
}

static void push_contact(OWorld *w, int32_t a, int32_t b, const double *p, const double *nrm, double dep)
{
    if (w->K >= w->kcap) {
        int64_t c = w->kcap ? 2 * w->kcap : 1024;
        w->ca = realloc(w->ca, sizeof(int32_t) * (size_t)c); w->cb = realloc(w->cb, sizeof(int32_t) * (size_t)c);
        w->cp = realloc(w->cp, sizeof(double) * 3 * (size_t)c); w->cn = realloc(w->cn, sizeof(double) * 3 * (size_t)c);
        w->cd = realloc(w->cd, sizeof(double) * (size_t)c); w->cni = realloc(w->cni, sizeof(double) * (size_t)c);
        w->cti = realloc(w->cti, sizeof(double) * 2 * (size_t)c);
        w->kcap = c;
    }
    int64_t k = w->K++;
    w->ca[k] = a; w->cb[k] = b;
    for (int r = 0; r < 3; r++) { w->cp[3 * k + r] = p[r]; w->cn[3 * k + r] = nrm[r]; }
    w->cd[k] = dep; w->cni[k] = 0.0; w->cti[2 * k] = 0.0; w->cti[2 * k + 1] = 0.0;
}

static uint64_t hmix(int32_t a, int32_t b)
{
    uint64_t x = ((uint64_t)(uint32_t)a << 32) | (uint32_t)b;
    x ^= x >> 33; x *= 0xff51afd7ed558ccdULL; x ^= x >> 33;
    return x;
}

/* solver._seed_pair (solver.py:222-233): nearest cached point of the same
 * (a, b) with d² < 0.04, first strict minimum. */
static void seed_pair(const OWorld *w, int64_t k, double *ni, double *t0, double *t1)
{
    if (!w->hcap) return;
    int32_t a = w->ca[k], b = w->cb[k];
    uint64_t h = hmix(a, b) & (uint64_t)(w->hcap - 1);
    for (;;) {
        int32_t s = w->hkey_start[h];
        if (s < 0) return;
        if (w->cca[s] == a && w->ccb[s] == b) {
            double best_d = 0.04;
            int found = 0;
            for (int64_t q = s; q < w->cK && w->cca[q] == a && w->ccb[q] == b; q++) {
                double dx = w->cp[3 * k] - w->ccp[3 * q], dy = w->cp[3 * k + 1] - w->ccp[3 * q + 1];
                double dz = w->cp[3 * k + 2] - w->ccp[3 * q + 2];
                double dd = dx * dx + dy * dy + dz * dz;
                if (dd < best_d) { best_d = dd; *ni = w->ccn[q]; *t0 = w->cct[2 * q]; *t1 = w->cct[2 * q + 1]; found = 1; }
            }
            (void)found;
            return;
        }
        h = (h + 1) & (uint64_t)(w->hcap - 1);
    }
}

/* solver._solve_native_full cache rebuild (solver.py:271-278). */
static void store_cache(OWorld *w)
{
    int64_t K = w->K;
    free(w->cca); free(w->ccb); free(w->ccp); free(w->ccn); free(w->cct); free(w->hkey_start);
    w->cca = dupmem(w->ca, sizeof(int32_t) * (size_t)K); w->ccb = dupmem(w->cb, sizeof(int32_t) * (size_t)K);
    w->ccp = dupmem(w->cp, sizeof(double) * 3 * (size_t)K); w->ccn = dupmem(w->cni, sizeof(double) * (size_t)K);
    w->cct = dupmem(w->cti, sizeof(double) * 2 * (size_t)K);
    w->cK = K;
    int64_t cap = 16;
    while (cap < 2 * K) cap <<= 1;
    w->hcap = cap;
    w->hkey_start = (int32_t *)malloc(sizeof(int32_t) * (size_t)cap);
    for (int64_t i = 0; i < cap; i++) w->hkey_start[i] = -1;
    for (int64_t k = 0; k < K; k++) {
        if (k > 0 && w->ca[k] == w->ca[k - 1] && w->cb[k] == w->cb[k - 1]) continue;
        uint64_t h = hmix(w->ca[k], w->cb[k]) & (uint64_t)(cap - 1);
        while (w->hkey_start[h] >= 0) h = (h + 1) & (uint64_t)(cap - 1);
        w->hkey_start[h] = (int32_t)k;
    }
}

static int awake_mover(const OWorld *w, int32_t i)   /* world.py:381-384 */
{
    return (w->mt[i] == M_DYNAMIC && w->awake[i]) || w->mt[i] == M_KINEMATIC;
}

static void collide_pair(OWorld *w, int32_t a, int32_t b)   /* narrowphase.py:78-100 */
{
    int ka = w->kind[a], kb = w->kind[b];
    double nrm[3], pts[12], dep[4];
    const double *pa = w->pos + 3 * a, *pb = w->pos + 3 * b;
    const double *Ra = w->R + 9 * a, *Rb = w->R + 9 * b, *ha = w->half + 3 * a, *hb = w->half + 3 * b;
    int cnt;
    if (ka == K_BOX && kb == K_BOX) cnt = ow_box_box(pa, Ra, ha, pb, Rb, hb, nrm, pts, dep);
    else if (ka == K_SPHERE && kb == K_SPHERE) cnt = ow_sphere_sphere(pa, ha[0], pb, hb[0], nrm, pts, dep);
    else if (ka == K_SPHERE && kb == K_BOX) cnt = ow_sphere_box(pa, ha[0], pb, Rb, hb, nrm, pts, dep);
    else if (ka == K_BOX && kb == K_SPHERE) {
        cnt = ow_sphere_box(pb, hb[0], pa, Ra, ha, nrm, pts, dep);
        nrm[0] = -nrm[0]; nrm[1] = -nrm[1]; nrm[2] = -nrm[2];
    } else {
        const double *pta = ka == K_CONVEX ? w->hull_pts + 3 * w->hull_off[a] : NULL;
        const double *ptb = kb == K_CONVEX ? w->hull_pts + 3 * w->hull_off[b] : NULL;
        cnt = ow_collide_convex(ka, pa, Ra, ha, pta, ka == K_CONVEX ? w->hull_cnt[a] : 0,
                                kb, pb, Rb, hb, ptb, kb == K_CONVEX ? w->hull_cnt[b] : 0,
                                nrm, pts, dep);
        if (cnt < 0) cnt = 0;
    }
    if (cnt > 0) w->nmanifold++;
    for (int i = 0; i < cnt; i++) push_contact(w, a, b, pts + 3 * i, nrm, dep[i]);
}

/* TriggerSystem.update's overlap test (triggers.py:24-70) for one broadphase
 * pair (i < j) with a trigger body: _sensor_pair orders it as (trigger,
 * other) -- i when i carries a trigger shape, else j -- and skips an other
 * body with neither collider nor trigger shape; _overlapping collides the
 * trigger's shape at its pose with the other's shape (collide.collide:
 * box-box, sphere-sphere, sphere-box either way round, GJK/EPA otherwise)
 * and overlaps when that yields a contact.  A body's collider and trigger
 * shapes are one shape here (kind, half, hull), which the loader enforces.
 * Poses are those after the contact solve, as in _CollisionStages.run. */
static int trigger_overlap(OWorld *w, int32_t i, int32_t j, int32_t *trg, int32_t *oth)
{
    int32_t t = w->ts[i] >= 0 ? i : j, o = w->ts[i] >= 0 ? j : i;
    *trg = t; *oth = o;
    if (w->kind[o] == K_NONE) return 0;
    int kt = w->kind[t], ko = w->kind[o];
    double nrm[3], pts[12], dep[4];
    const double *pt_ = w->pos + 3 * t, *po = w->pos + 3 * o;
    const double *Rt = w->R + 9 * t, *Ro = w->R + 9 * o, *ht = w->half + 3 * t, *ho = w->half + 3 * o;
    int cnt;
    if (kt == K_BOX && ko == K_BOX) cnt = ow_box_box(pt_, Rt, ht, po, Ro, ho, nrm, pts, dep);
    else if (kt == K_SPHERE && ko == K_SPHERE) cnt = ow_sphere_sphere(pt_, ht[0], po, ho[0], nrm, pts, dep);
    else if (kt == K_SPHERE && ko == K_BOX) cnt = ow_sphere_box(pt_, ht[0], po, Ro, ho, nrm, pts, dep);
    else if (kt == K_BOX && ko == K_SPHERE) cnt = ow_sphere_box(po, ho[0], pt_, Rt, ht, nrm, pts, dep);
    else {
        const double *pa = kt == K_CONVEX ? w->hull_pts + 3 * w->hull_off[t] : NULL;
        const double *pb = ko == K_CONVEX ? w->hull_pts + 3 * w->hull_off[o] : NULL;
        cnt = ow_collide_convex(kt, pt_, Rt, ht, pa, kt == K_CONVEX ? w->hull_cnt[t] : 0,
                                ko, po, Ro, ho, pb, ko == K_CONVEX ? w->hull_cnt[o] : 0, nrm, pts, dep);
    }
    return cnt > 0;
}

static void find_trigger_overlaps(OWorld *w)
{
    w->ntov = 0;
    for (int64_t p = 0; p < w->npairs; p++) {
        int32_t i = w->pairs[2 * p], j = w->pairs[2 * p + 1], t, o;
        if (w->ts[i] < 0 && w->ts[j] < 0) continue;
        if (!trigger_overlap(w, i, j, &t, &o)) continue;
        if (w->ntov == w->tov_cap) {
            w->tov_cap = w->tov_cap ? 2 * w->tov_cap : 64;
            w->tov = realloc(w->tov, sizeof(int32_t) * 2 * (size_t)w->tov_cap);
        }
        w->tov[2 * w->ntov] = t; w->tov[2 * w->ntov + 1] = o; w->ntov++;
    }
}

int64_t ow_num_trigger_overlaps(const OWorld *w) { return w->ntov; }

void ow_get_trigger_overlaps(const OWorld *w, int32_t *out)
{
    for (int64_t k = 0; k < 2 * w->ntov; k++) out[k] = w->tov[k];
}

void ow_solve3(const double *K, const double *rhs, double *x)
{
    double A[3][3], b[3];
    for (int r = 0; r < 3; r++) { b[r] = rhs[r]; for (int c = 0; c < 3; c++) A[r][c] = K[3 * r + c]; }
    int piv[3];
    for (int j = 0; j < 3; j++) {
        int p = j;
        double m = fabs(A[j][j]);
        for (int i = j + 1; i < 3; i++) if (fabs(A[i][j]) > m) { m = fabs(A[i][j]); p = i; }
        piv[j] = p;
        if (A[p][j] != 0.0) {
            if (p != j) for (int c = 0; c < 3; c++) { double t = A[j][c]; A[j][c] = A[p][c]; A[p][c] = t; }
            double inv = 1.0 / A[j][j];
            for (int i = j + 1; i < 3; i++) A[i][j] *= inv;
        }
        for (int i = j + 1; i < 3; i++)
            for (int c = j + 1; c < 3; c++) A[i][c] -= A[i][j] * A[j][c];
    }
    for (int j = 0; j < 3; j++) if (piv[j] != j) { double t = b[j]; b[j] = b[piv[j]]; b[piv[j]] = t; }
    for (int i = 1; i < 3; i++) for (int k = 0; k < i; k++) b[i] -= A[i][k] * b[k];
    for (int i = 2; i >= 0; i--) {
        for (int k = i + 1; k < 3; k++) b[i] -= A[i][k] * b[k];
        b[i] /= A[i][i];
    }
    x[0] = b[0]; x[1] = b[1]; x[2] = b[2];
}

/* _JointBase._inv_inertia_world (joints.py:35-40): R diag Rᵀ; zero for the world. */
static void joint_inv_inertia(const OWorld *w, int32_t i, double *o)
{
    if (i < 0) { for (int r = 0; r < 9; r++) o[r] = 0.0; return; }
    double R[9];
    ow_quat_to_matrix(w->quat + 4 * i, R);
    const double *d = w->inv_inertia + 3 * i;
    for (int r = 0; r < 3; r++)
        for (int c = 0; c < 3; c++) {
            double p0 = (R[3 * r] * d[0]) * R[3 * c], p1 = (R[3 * r + 1] * d[1]) * R[3 * c + 1];
            double p2 = (R[3 * r + 2] * d[2]) * R[3 * c + 2];
            o[3 * r + c] = (p0 + p2) + p1;
        }
}

/* skew(r) @ I @ skew(r) (joints.py:20-23, 96-97) */
static void skew_sandwich(const double *r, const double *I, double *o)
{
    const double S[9] = {0.0, -r[2], r[1], r[2], 0.0, -r[0], -r[1], r[0], 0.0};
    double M[9];
    for (int i = 0; i < 3; i++)
        for (int c = 0; c < 3; c++) M[3 * i + c] = (S[3 * i] * I[c] + S[3 * i + 1] * I[3 + c]) + S[3 * i + 2] * I[6 + c];
    for (int i = 0; i < 3; i++)
        for (int c = 0; c < 3; c++) o[3 * i + c] = (M[3 * i] * S[c] + M[3 * i + 1] * S[3 + c]) + M[3 * i + 2] * S[6 + c];
}

/* PointConstraint.prepare (joints.py:84-101) */
static void joint_prepare(OWorld *w, PJoint *J)
{
    joint_inv_inertia(w, J->a, J->Ia);
    joint_inv_inertia(w, J->b, J->Ib);
    for (int s = 0; s < 2; s++) {
        int32_t i = s ? J->b : J->a;
        double *loc = s ? J->lb : J->la;
        int *has = s ? &J->has_lb : &J->has_la;
        double *rr = s ? J->rB : J->rA;
        if (i < 0) { for (int r = 0; r < 3; r++) rr[r] = J->anchor[r] - J->anchor[r]; continue; }
        double R[9];
        ow_quat_to_matrix(w->quat + 4 * i, R);
        const double *p = w->pos + 3 * i;
        if (!*has) {
            double d[3] = {J->anchor[0] - p[0], J->anchor[1] - p[1], J->anchor[2] - p[2]};
            for (int c = 0; c < 3; c++) loc[c] = (R[c] * d[0] + R[3 + c] * d[1]) + R[6 + c] * d[2];
            *has = 1;
        }
        for (int r = 0; r < 3; r++) {
            double wa = p[r] + ((R[3 * r] * loc[0] + R[3 * r + 1] * loc[1]) + R[3 * r + 2] * loc[2]);
            rr[r] = wa - p[r];
        }
    }
    J->ima = J->a < 0 ? 0.0 : w->inv_mass[J->a];
    J->imb = J->b < 0 ? 0.0 : w->inv_mass[J->b];
    double SA[9], SB[9];
    skew_sandwich(J->rA, J->Ia, SA);
    skew_sandwich(J->rB, J->Ib, SB);
    for (int r = 0; r < 3; r++)
        for (int c = 0; c < 3; c++) {
            double k = (J->ima + J->imb) * (r == c ? 1.0 : 0.0);
            k -= SA[3 * r + c];
            k -= SB[3 * r + c];
            J->K[3 * r + c] = k;
        }
    for (int r = 0; r < 3; r++) {
        double pa = (J->a < 0 ? J->anchor[r] : w->pos[3 * J->a + r]) + J->rA[r];
        double pb = (J->b < 0 ? J->anchor[r] : w->pos[3 * J->b + r]) + J->rB[r];
        J->bias[r] = 0.2 * (pb - pa);
    }
}

/* PointConstraint.solve (joints.py:111-121) */
static void joint_solve(OWorld *w, PJoint *J, double dt)
{
    double va[3] = {0.0, 0.0, 0.0}, vb[3] = {0.0, 0.0, 0.0}, t[3];
    if (J->a >= 0) {
        cross3(w->av + 3 * J->a, J->rA, t);
        for (int r = 0; r < 3; r++) va[r] = w->lv[3 * J->a + r] + t[r];
    }
    if (J->b >= 0) {
        cross3(w->av + 3 * J->b, J->rB, t);
        for (int r = 0; r < 3; r++) vb[r] = w->lv[3 * J->b + r] + t[r];
    }
    double rhs[3], P[3];
    for (int r = 0; r < 3; r++) rhs[r] = -((vb[r] - va[r]) + J->bias[r] / dt);
    ow_solve3(J->K, rhs, P);
    for (int s = 0; s < 2; s++) {
        int32_t i = s ? J->b : J->a;
        if (i < 0) continue;
        const double *rr = s ? J->rB : J->rA, *I = s ? J->Ib : J->Ia;
        double Q[3] = {s ? P[0] : -P[0], s ? P[1] : -P[1], s ? P[2] : -P[2]}, c[3];
        for (int r = 0; r < 3; r++) w->lv[3 * i + r] += w->inv_mass[i] * Q[r];
        cross3(rr, Q, c);
        for (int r = 0; r < 3; r++) w->av[3 * i + r] += (I[3 * r] * c[0] + I[3 * r + 1] * c[1]) + I[3 * r + 2] * c[2];
    }
}

/* PhysicsWorld._solve_joints (world.py:599-611) */
static void solve_joints(OWorld *w, double dt)
{
    for (int32_t j = 0; j < w->nj; j++) {
        PJoint *J = &w->joints[j];
        if (J->a >= 0) { w->awake[J->a] = 1; w->timer[J->a] = 0.0; }
        if (J->b >= 0) { w->awake[J->b] = 1; w->timer[J->b] = 0.0; }
        joint_prepare(w, J);
    }
    for (int32_t it = 0; it < w->jiters; it++)
        for (int32_t j = 0; j < w->nj; j++) joint_solve(w, &w->joints[j], dt);
}

/* PhysicsWorld.step (world.py:410-430) with _CollisionStages.run (:658-676),
 * SequentialImpulseSolver.solve (solver.py:192-213) and _update_sleep
 * (world.py:510-567). */
/* PhysicsWorld._update_sleep + _contact_islands (world.py:510-567) over n
 * bodies and the step's K contacts: slow = dynamic with |v|^2 < lin^2 and
 * |w|^2 < ang^2 (sums left to right); the timer grows by dt while slow and
 * resets otherwise; ready = slow and timer > t_sleep; dynamic-dynamic contacts
 * join islands; an island sleeps when every member is ready, a ready dynamic
 * body in no contact sleeps alone; moving dynamic bodies wake; sleeping dynamic
 * bodies get zero velocity. */
void ow_update_sleep(int32_t n, int64_t K, const int32_t *ca, const int32_t *cb, const int32_t *mt, double dt,
                     double lin, double ang, double t_sleep, double *lv, double *av, uint8_t *awake, double *timer)
{
    uint8_t *slow = (uint8_t *)calloc((size_t)n + 1, 1), *ready = (uint8_t *)calloc((size_t)n + 1, 1);
    uint8_t *contacted = (uint8_t *)calloc((size_t)n + 1, 1), *notready = (uint8_t *)calloc((size_t)n + 1, 1);
    int32_t *par = (int32_t *)malloc(sizeof(int32_t) * ((size_t)n + 1));
    for (int32_t i = 0; i < n; i++) {
        int dyn = mt[i] == M_DYNAMIC;
        double s2 = (lv[3 * i] * lv[3 * i] + lv[3 * i + 1] * lv[3 * i + 1]) + lv[3 * i + 2] * lv[3 * i + 2];
        double w2 = (av[3 * i] * av[3 * i] + av[3 * i + 1] * av[3 * i + 1]) + av[3 * i + 2] * av[3 * i + 2];
        slow[i] = dyn && s2 < lin * lin && w2 < ang * ang;
        if (slow[i]) timer[i] += dt; else timer[i] = 0.0;
        ready[i] = slow[i] && timer[i] > t_sleep;
        par[i] = i;
    }
    for (int64_t k = 0; k < K; k++) {
        int32_t a = ca[k], b = cb[k];
        int da = mt[a] == M_DYNAMIC, db = mt[b] == M_DYNAMIC;
        if (da) contacted[a] = 1;
        if (db) contacted[b] = 1;
        if (da && db) {
            int32_t x = a, y = b;
            while (par[x] != x) x = par[x];
            while (par[y] != y) y = par[y];
            if (x != y) par[x] = y;
        }
    }
    for (int32_t i = 0; i < n; i++) {
        if (!contacted[i]) continue;
        int32_t r = i;
        while (par[r] != r) r = par[r];
        par[i] = r;
    }
    for (int32_t i = 0; i < n; i++) if (contacted[i] && !ready[i]) notready[par[i]] = 1;
    for (int32_t i = 0; i < n; i++) {
        int dyn = mt[i] == M_DYNAMIC;
        int to_sleep = contacted[i] ? !notready[par[i]] : (dyn && ready[i]);
        if (to_sleep) awake[i] = 0;
        if (dyn && !slow[i]) awake[i] = 1;
        if (!awake[i] && dyn) {
            lv[3 * i] = lv[3 * i + 1] = lv[3 * i + 2] = 0.0;
            av[3 * i] = av[3 * i + 1] = av[3 * i + 2] = 0.0;
        }
    }
    free(slow); free(ready); free(contacted); free(notready); free(par);
}

void ow_step(OWorld *w, double dt)
{
    int32_t n = w->n;
    ow_integrate_forces(n, w->mt, w->awake, w->g, 0, w->gf, w->ld, w->ad, w->qd, w->dlin, w->dang, w->lv, w->av, dt);
    w->K = 0;
    w->npairs = w->nsolid = w->xcand = w->nmanifold = 0;
    if (w->has_collision) {
        ow_refit(n, w->pos, w->quat, w->kind, w->half, w->hull_off, w->hull_cnt, w->hull_pts, 0.05, w->lo, w->hi);
        int64_t np_;
        for (;;) {
            np_ = ow_sap_pairs(n, w->lo, w->hi, w->mt, w->filt, w->n_filt, w->ftab, w->pairs, w->pair_cap, &w->xcand);
            if (np_ <= w->pair_cap) break;
            w->pair_cap = 2 * np_;
            w->pairs = realloc(w->pairs, sizeof(int32_t) * 2 * (size_t)w->pair_cap);
            w->solid = realloc(w->solid, sizeof(int32_t) * 2 * (size_t)w->pair_cap);
        }
        w->npairs = np_;
        for (int64_t p = 0; p < np_; p++) {
            int32_t i = w->pairs[2 * p], j = w->pairs[2 * p + 1];
            if (w->ts[i] >= 0 || w->ts[j] >= 0) continue;
            if (!(w->cs[i] >= 0 && w->cs[j] >= 0)) continue;
            if (!(awake_mover(w, i) || awake_mover(w, j))) continue;
            w->solid[2 * w->nsolid] = i; w->solid[2 * w->nsolid + 1] = j; w->nsolid++;
        }
        for (int32_t i = 0; i < n; i++) ow_quat_to_matrix(w->quat + 4 * i, w->R + 9 * i);
        for (int64_t p = 0; p < w->nsolid; p++) collide_pair(w, w->solid[2 * p], w->solid[2 * p + 1]);
        int64_t K = w->K;
        if (K > 0) {
            /* _wake_sleepers (solver.py:321-337), sequential as written */
            double *v2 = (double *)malloc(sizeof(double) * (size_t)n);
            for (int32_t i = 0; i < n; i++) v2[i] = (w->lv[3 * i] * w->lv[3 * i] + w->lv[3 * i + 1] * w->lv[3 * i + 1]) + w->lv[3 * i + 2] * w->lv[3 * i + 2];
            for (int64_t k = 0; k < K; k++) {
                int32_t a = w->ca[k], b = w->cb[k];
                int wa = w->mt[a] == M_KINEMATIC || (w->mt[a] == M_DYNAMIC && w->awake[a] && v2[a] > 0.15 * 0.15);
                if (!w->awake[b] && wa) { w->awake[b] = 1; w->timer[b] = 0.0; }
                int wb = w->mt[b] == M_KINEMATIC || (w->mt[b] == M_DYNAMIC && w->awake[b] && v2[b] > 0.15 * 0.15);
                if (!w->awake[a] && wb) { w->awake[a] = 1; w->timer[a] = 0.0; }
            }
            free(v2);
            /* material tables, warm-start seeds, world inverse inertia (solver.py:248-264) */
            double *rest = (double *)malloc(sizeof(double) * (size_t)K), *fr = (double *)malloc(sizeof(double) * (size_t)K);
            int64_t *ai = (int64_t *)malloc(sizeof(int64_t) * (size_t)K), *bi = (int64_t *)malloc(sizeof(int64_t) * (size_t)K);
            for (int64_t k = 0; k < K; k++) {
                int32_t a = w->ca[k], b = w->cb[k];
                ai[k] = a; bi[k] = b;
                rest[k] = w->rtab[w->mat[a] * w->n_mat + w->mat[b]];
                fr[k] = w->ftab_f[w->mat[a] * w->n_mat + w->mat[b]];
                double ni = 0.0, t0 = 0.0, t1 = 0.0;
                seed_pair(w, k, &ni, &t0, &t1);
                w->cni[k] = ni; w->cti[2 * k] = t0; w->cti[2 * k + 1] = t1;
            }
            /* _inv_inertia_world_all (solver.py:215-220): einsum order (p0 + p2) + p1 */
            for (int32_t i = 0; i < n; i++) {
                const double *R = w->R + 9 * i, *d = w->inv_inertia + 3 * i;
                double *o = w->invIw + 9 * i;
                for (int r = 0; r < 3; r++)
                    for (int c = 0; c < 3; c++) {
                        double p0 = (R[3 * r] * d[0]) * R[3 * c], p1 = (R[3 * r + 1] * d[1]) * R[3 * c + 1];
                        double p2 = (R[3 * r + 2] * d[2]) * R[3 * c + 2];
                        o[3 * r + c] = (p0 + p2) + p1;
                    }
            }
            ow_prepare_and_solve(K, ai, bi, w->cp, w->cn, w->cd, w->pos, w->lv, w->av, w->inv_mass, w->invIw,
                                 rest, fr, w->cni, w->cti, 0.5, w->iters, 1, 0.8, 0.005);
            store_cache(w);
            free(rest); free(fr); free(ai); free(bi);
        }
        find_trigger_overlaps(w);
    } else {
        w->ntov = 0;
    }
    if (w->nj > 0) solve_joints(w, dt);
    ow_integrate_positions(n, w->mt, w->awake, w->lv, w->av, w->pos, w->quat, dt);
    if (w->sleep_enabled)
        ow_update_sleep(n, w->K, w->ca, w->cb, w->mt, dt, 0.08, 0.25, 0.4, w->lv, w->av, w->awake, w->timer);
}
