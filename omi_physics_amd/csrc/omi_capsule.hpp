// omi_capsule.hpp — capsule against triangles on the device (gfx950), FP64.
//
// _collide_native.capsule_triangles (_collide_native.pyx:361-574), the
// character controller's query: one capsule (segment c0 c1, radius) against a
// batch of triangles, one thread per triangle.  Operation order is the
// Cython loop's term by term (the library builds with -ffp-contract=off), so
// results are bit-identical to it.
#pragma once
#include <hip/hip_runtime.h>

namespace omi {

constexpr double CAP_EPS = 1e-9;   // _collide_native.pyx:13

__device__ __forceinline__ double cdot3(const double* a, const double* b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }
__device__ __forceinline__ void cset3(double* o, const double* a) { o[0] = a[0]; o[1] = a[1]; o[2] = a[2]; }
__device__ __forceinline__ double cclamp01(double x) { return x < 0.0 ? 0.0 : (x > 1.0 ? 1.0 : x); }

// Closest point on triangle abc to p by Voronoi regions (Ericson), _closest_on_tri.
__device__ inline void closest_on_tri(const double* p, const double* a, const double* b, const double* c, double* out) {
    double ab[3], ac[3], ap[3], bp[3], cp[3];
    for (int m = 0; m < 3; m++) { ab[m] = b[m] - a[m]; ac[m] = c[m] - a[m]; ap[m] = p[m] - a[m]; }
    const double d1 = cdot3(ab, ap), d2 = cdot3(ac, ap);
    if (d1 <= 0.0 && d2 <= 0.0) { cset3(out, a); return; }
    for (int m = 0; m < 3; m++) bp[m] = p[m] - b[m];
    const double d3 = cdot3(ab, bp), d4 = cdot3(ac, bp);
    if (d3 >= 0.0 && d4 <= d3) { cset3(out, b); return; }
    const double vc = d1 * d4 - d3 * d2;
    if (vc <= 0.0 && d1 >= 0.0 && d3 <= 0.0) {
        const double v = d1 / (d1 - d3);
        for (int m = 0; m < 3; m++) out[m] = a[m] + v * ab[m];
        return;
    }
    for (int m = 0; m < 3; m++) cp[m] = p[m] - c[m];
    const double d5 = cdot3(ab, cp), d6 = cdot3(ac, cp);
    if (d6 >= 0.0 && d5 <= d6) { cset3(out, c); return; }
    const double vb = d5 * d2 - d1 * d6;
    if (vb <= 0.0 && d2 >= 0.0 && d6 <= 0.0) {
        const double w = d2 / (d2 - d6);
        for (int m = 0; m < 3; m++) out[m] = a[m] + w * ac[m];
        return;
    }
    const double va = d3 * d6 - d5 * d4;
    if (va <= 0.0 && (d4 - d3) >= 0.0 && (d5 - d6) >= 0.0) {
        const double w = (d4 - d3) / ((d4 - d3) + (d5 - d6));
        for (int m = 0; m < 3; m++) out[m] = b[m] + w * (c[m] - b[m]);
        return;
    }
    const double denom = 1.0 / (va + vb + vc);
    const double v = vb * denom, w = vc * denom;
    for (int m = 0; m < 3; m++) out[m] = a[m] + ab[m] * v + ac[m] * w;
}

// Closest pair of points between segments p1 q1 and p2 q2, _closest_seg_seg.
__device__ inline void closest_seg_seg(const double* p1, const double* q1, const double* p2, const double* q2,
                                       double* o1, double* o2) {
    double d1[3], d2[3], r[3], s, t;
    for (int m = 0; m < 3; m++) { d1[m] = q1[m] - p1[m]; d2[m] = q2[m] - p2[m]; r[m] = p1[m] - p2[m]; }
    const double aa = cdot3(d1, d1), e = cdot3(d2, d2), f = cdot3(d2, r);
    if (aa <= CAP_EPS && e <= CAP_EPS) { cset3(o1, p1); cset3(o2, p2); return; }
    if (aa <= CAP_EPS) {
        s = 0.0;
        t = cclamp01(f / e);
    } else {
        const double cc = cdot3(d1, r);
        if (e <= CAP_EPS) {
            t = 0.0;
            s = cclamp01(-cc / aa);
        } else {
            const double bb = cdot3(d1, d2), denom = aa * e - bb * bb;
            s = denom > CAP_EPS ? cclamp01((bb * f - cc * e) / denom) : 0.0;
            t = (bb * s + f) / e;
            if (t < 0.0) { t = 0.0; s = cclamp01(-cc / aa); }
            else if (t > 1.0) { t = 1.0; s = cclamp01((bb - cc) / aa); }
        }
    }
    for (int m = 0; m < 3; m++) { o1[m] = p1[m] + d1[m] * s; o2[m] = p2[m] + d2[m] * t; }
}

// One triangle: face contact first (the side is the capsule centre's, the
// depth unbounded: the cap endpoint whose foot lies on the triangle and that
// reaches deepest), else the nearest feature pair -- endpoint to triangle,
// axis to each edge -- inside the radius, normal from the triangle to the
// capsule.  Returns 1 on contact and fills point, normal, depth.
__device__ inline int capsule_triangle(const double* c0, const double* c1, double radius, const double* v0,
                                       const double* v1, const double* v2, double* point, double* normal,
                                       double& depth) {
    double centre[3], e1[3], e2[3], face[3], fn[3], tp[3], foot[3], best[3];
    for (int m = 0; m < 3; m++) {
        centre[m] = (c0[m] + c1[m]) * 0.5;
        e1[m] = v1[m] - v0[m];
        e2[m] = v2[m] - v0[m];
    }
    face[0] = e1[1] * e2[2] - e1[2] * e2[1];
    face[1] = e1[2] * e2[0] - e1[0] * e2[2];
    face[2] = e1[0] * e2[1] - e1[1] * e2[0];
    const double face_len = sqrt(cdot3(face, face));
    bool face_found = false;
    double best_reach = -1e300;
    if (face_len > CAP_EPS) {
        for (int m = 0; m < 3; m++) fn[m] = face[m] / face_len;
        double gx = centre[0] - v0[0], gy = centre[1] - v0[1], gz = centre[2] - v0[2];
        if (fn[0] * gx + fn[1] * gy + fn[2] * gz < 0.0)
            for (int m = 0; m < 3; m++) fn[m] = -fn[m];
        for (int k = 0; k < 2; k++) {
            const double* sp = k ? c1 : c0;
            gx = sp[0] - v0[0]; gy = sp[1] - v0[1]; gz = sp[2] - v0[2];
            const double height = fn[0] * gx + fn[1] * gy + fn[2] * gz;
            for (int m = 0; m < 3; m++) foot[m] = sp[m] - fn[m] * height;
            closest_on_tri(sp, v0, v1, v2, tp);
            gx = foot[0] - tp[0]; gy = foot[1] - tp[1]; gz = foot[2] - tp[2];
            if (gx * gx + gy * gy + gz * gz > CAP_EPS * CAP_EPS) continue;
            const double reach = radius - height;
            if (reach > 0.0 && reach > best_reach) {
                best_reach = reach;
                face_found = true;
                cset3(best, tp);
            }
        }
    }
    if (face_found) {
        depth = best_reach;
        cset3(point, best);
        cset3(normal, fn);
        return 1;
    }
    double best_d2 = 1e300, bsp[3] = {0.0, 0.0, 0.0}, btp[3] = {0.0, 0.0, 0.0}, ap[3], ep[3];
    for (int k = 0; k < 2; k++) {
        const double* sp = k ? c1 : c0;
        closest_on_tri(sp, v0, v1, v2, tp);
        const double gx = sp[0] - tp[0], gy = sp[1] - tp[1], gz = sp[2] - tp[2];
        const double d2 = gx * gx + gy * gy + gz * gz;
        if (d2 < best_d2) { best_d2 = d2; cset3(bsp, sp); cset3(btp, tp); }
    }
    for (int k = 0; k < 3; k++) {
        const double* pa = k == 0 ? v0 : (k == 1 ? v1 : v2);
        const double* pb = k == 0 ? v1 : (k == 1 ? v2 : v0);
        closest_seg_seg(c0, c1, pa, pb, ap, ep);
        const double gx = ap[0] - ep[0], gy = ap[1] - ep[1], gz = ap[2] - ep[2];
        const double d2 = gx * gx + gy * gy + gz * gz;
        if (d2 < best_d2) { best_d2 = d2; cset3(bsp, ap); cset3(btp, ep); }
    }
    if (best_d2 >= radius * radius) return 0;
    const double dist = sqrt(best_d2);
    if (dist > CAP_EPS) {
        for (int m = 0; m < 3; m++) normal[m] = (bsp[m] - btp[m]) / dist;
    } else if (face_len > CAP_EPS) {
        for (int m = 0; m < 3; m++) normal[m] = face[m] / face_len;
    } else {
        normal[0] = normal[1] = normal[2] = 0.0;
    }
    depth = radius - dist;
    cset3(point, btp);
    return 1;
}

}  // namespace omi
