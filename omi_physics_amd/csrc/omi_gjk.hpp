// omi_gjk.hpp — GJK intersection + EPA penetration for convex proxies on gfx950.
//
// Follows gjk.py (whole file) and the support proxies of body.py:33-132 step
// by step, in FP64 (the mixed-precision form keeps EPA in FP64 as well).  The
// growing polytope (<= 69 vertices with their witness points, <= 256 faces)
// does not fit in registers, so each thread works in its own slice of a
// global workspace (EpaWork), sized once per batch.
#pragma once
#include "omi_math.hpp"

namespace omi {

#define GJK_EPS 1e-10
#define EPA_MAXV 69
#define EPA_MAXF 256

struct Proxy {
    int kind;                 // K_SPHERE, K_BOX, K_CONVEX
    double c[3], R[9], h[3];
    const double* local;      // convex: local points [m][3]
    int m;
    const double* wp = nullptr;   // convex: world points, element k at wp[k * ws] (else from local)
    int ws = 0;
};

// Hulls of up to this many points get their world points staged once per
// pair (the convex narrowphase keeps them in LDS).
constexpr int HULL_STAGE_PTS = 32;

struct EpaWork {              // one per thread
    double v[EPA_MAXV][3], a[EPA_MAXV][3], b[EPA_MAXV][3];
    int f[EPA_MAXF][3];
    // per face, computed once when the face is made: the unit normal
    // oriented away from the origin and its distance (+inf for a degenerate
    // face), which _closest_face recomputes from the same vertices each time
    double fn[EPA_MAXF][3], fd[EPA_MAXF];
    int ek[3 * EPA_MAXF][2];
    unsigned char ealive[3 * EPA_MAXF];
};

__device__ __forceinline__ double dot3(const double* a, const double* b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }
__device__ __forceinline__ void cross3(const double* a, const double* b, double* o) {
    double x = a[1] * b[2] - a[2] * b[1], y = a[2] * b[0] - a[0] * b[2], z = a[0] * b[1] - a[1] * b[0];
    o[0] = x; o[1] = y; o[2] = z;
}
__device__ __forceinline__ void sub3(const double* a, const double* b, double* o) {
    o[0] = a[0] - b[0]; o[1] = a[1] - b[1]; o[2] = a[2] - b[2];
}

// World point p of a convex proxy: local @ R.T + c (body.py:119)
__device__ __forceinline__ void convex_pt(const Proxy& P, int p, double* o) {
    const double* l = P.local + 3 * p;
    for (int r = 0; r < 3; r++) o[r] = ((l[0] * P.R[3 * r] + l[1] * P.R[3 * r + 1]) + l[2] * P.R[3 * r + 2]) + P.c[r];
}

// The same point from the staged world points when there are any.
__device__ __forceinline__ void convex_w(const Proxy& P, int p, double* o) {
    if (P.wp) {
        for (int r = 0; r < 3; r++) o[r] = P.wp[(size_t)(3 * p + r) * P.ws];
    } else {
        convex_pt(P, p, o);
    }
}

// Stage a convex proxy's world points at base (element k at base[k * stride])
// when it has at most HULL_STAGE_PTS of them; the values are convex_pt's.
__device__ __forceinline__ void stage_world_pts(Proxy& P, double* base, int stride) {
    if (P.kind != K_CONVEX || P.m > HULL_STAGE_PTS) return;
    for (int p = 0; p < P.m; p++) {
        double o[3];
        convex_pt(P, p, o);
        for (int r = 0; r < 3; r++) base[(size_t)(3 * p + r) * stride] = o[r];
    }
    P.wp = base;
    P.ws = stride;
}

// SphereProxy/BoxProxy/ConvexProxy.support (body.py:46-48, 69-73, 125-128)
__device__ void support1(const Proxy& P, const double* d, double* out) {
    if (P.kind == K_SPHERE) {
        double n = sqrt((d[0] * d[0] + d[1] * d[1]) + d[2] * d[2]);
        double s = maxd(n, 1e-12);
        for (int r = 0; r < 3; r++) out[r] = P.c[r] + P.h[0] * (d[r] / s);
    } else if (P.kind == K_BOX) {
        double v[3];
        for (int c = 0; c < 3; c++) {
            double l = (P.R[c] * d[0] + P.R[3 + c] * d[1]) + P.R[6 + c] * d[2];
            double sg = l > 0.0 ? 1.0 : (l < 0.0 ? -1.0 : 0.0);
            v[c] = sg * P.h[c];
        }
        for (int r = 0; r < 3; r++) out[r] = P.c[r] + ((P.R[3 * r] * v[0] + P.R[3 * r + 1] * v[1]) + P.R[3 * r + 2] * v[2]);
    } else {
        int best = 0;
        double bv = 0.0, w[3];
        for (int p = 0; p < P.m; p++) {
            convex_w(P, p, w);
            double v = (w[0] * d[0] + w[1] * d[1]) + w[2] * d[2];
            if (p == 0 || v > bv) { bv = v; best = p; }
        }
        convex_w(P, best, out);
    }
}

// center_hint (body.py:50-52, 75-77, 130-132)
__device__ void center_hint(const Proxy& P, double* out) {
    if (P.kind != K_CONVEX) { out[0] = P.c[0]; out[1] = P.c[1]; out[2] = P.c[2]; return; }
    double s[3] = {0.0, 0.0, 0.0}, w[3];
    for (int p = 0; p < P.m; p++) { convex_w(P, p, w); for (int r = 0; r < 3; r++) s[r] += w[r]; }
    for (int r = 0; r < 3; r++) out[r] = s[r] / (double)P.m;
}

struct Vtx { double v[3], a[3], b[3]; };

__device__ __forceinline__ void msupport(const Proxy& A, const Proxy& B, const double* d, Vtx& o) {
    double nd[3] = {-d[0], -d[1], -d[2]};
    support1(A, d, o.a);
    support1(B, nd, o.b);
    sub3(o.a, o.b, o.v);
}

__device__ __forceinline__ void any_perp(const double* v, double* o) {
    double a[3] = {0.0, 0.0, 0.0};
    if (fabs(v[0]) < 0.9) a[0] = 1.0; else a[1] = 1.0;
    cross3(v, a, o);
}

// gjk._line (gjk.py:69-79): s[0] older, s[1] newest
__device__ void gjk_line(Vtx* s, int& n, double* d) {
    Vtx b = s[0], a = s[1];
    double ab[3], ao[3] = {-a.v[0], -a.v[1], -a.v[2]}, t[3];
    sub3(b.v, a.v, ab);
    if (dot3(ab, ao) > 0) {
        cross3(ab, ao, t); cross3(t, ab, d);
        if (dot3(d, d) < GJK_EPS) any_perp(ab, d);
        s[0] = b; s[1] = a; n = 2;
        return;
    }
    s[0] = a; n = 1;
    d[0] = ao[0]; d[1] = ao[1]; d[2] = ao[2];
}

// gjk._triangle (gjk.py:82-95)
__device__ void gjk_triangle(Vtx* s, int& n, double* d) {
    Vtx c = s[0], b = s[1], a = s[2];
    double ab[3], ac[3], ao[3] = {-a.v[0], -a.v[1], -a.v[2]}, abc[3], t[3];
    sub3(b.v, a.v, ab); sub3(c.v, a.v, ac);
    cross3(ab, ac, abc);
    cross3(abc, ac, t);
    if (dot3(t, ao) > 0) {
        if (dot3(ac, ao) > 0) {
            double u[3];
            cross3(ac, ao, u); cross3(u, ac, d);
            s[0] = c; s[1] = a; n = 2;
            return;
        }
        s[0] = b; s[1] = a; n = 2;
        gjk_line(s, n, d);
        return;
    }
    cross3(ab, abc, t);
    if (dot3(t, ao) > 0) { s[0] = b; s[1] = a; n = 2; gjk_line(s, n, d); return; }
    if (dot3(abc, ao) > 0) { s[0] = c; s[1] = b; s[2] = a; n = 3; d[0] = abc[0]; d[1] = abc[1]; d[2] = abc[2]; return; }
    s[0] = b; s[1] = c; s[2] = a; n = 3;
    d[0] = -abc[0]; d[1] = -abc[1]; d[2] = -abc[2];
}

// gjk._tetra (gjk.py:98-111); true when the tetrahedron holds the origin
__device__ bool gjk_tetra(Vtx* s, int& n, double* d) {
    Vtx dd = s[0], c = s[1], b = s[2], a = s[3];
    double ab[3], ac[3], ad[3], ao[3] = {-a.v[0], -a.v[1], -a.v[2]}, abc[3], acd[3], adb[3];
    sub3(b.v, a.v, ab); sub3(c.v, a.v, ac); sub3(dd.v, a.v, ad);
    cross3(ab, ac, abc); cross3(ac, ad, acd); cross3(ad, ab, adb);
    if (dot3(abc, ao) > 0) { s[0] = c; s[1] = b; s[2] = a; n = 3; gjk_triangle(s, n, d); return false; }
    if (dot3(acd, ao) > 0) { s[0] = dd; s[1] = c; s[2] = a; n = 3; gjk_triangle(s, n, d); return false; }
    if (dot3(adb, ao) > 0) { s[0] = b; s[1] = dd; s[2] = a; n = 3; gjk_triangle(s, n, d); return false; }
    d[0] = d[1] = d[2] = 0.0;
    return true;
}

// gjk.gjk_intersect (gjk.py:39-57)
__device__ bool gjk(const Proxy& A, const Proxy& B, Vtx* s, int& n) {
    double z[3] = {0.0, 0.0, 0.0}, sa[3], sb[3], d[3];
    support1(B, z, sb); support1(A, z, sa);
    sub3(sb, sa, d);
    if (dot3(d, d) < GJK_EPS) { d[0] = 1.0; d[1] = 0.0; d[2] = 0.0; }
    double nd[3] = {-d[0], -d[1], -d[2]};
    msupport(A, B, nd, s[0]);
    n = 1;
    d[0] = -s[0].v[0]; d[1] = -s[0].v[1]; d[2] = -s[0].v[2];
    for (int it = 0; it < 32; it++) {
        if (dot3(d, d) < GJK_EPS) return true;
        Vtx w;
        msupport(A, B, d, w);
        if (dot3(w.v, d) < 0) return false;
        s[n++] = w;
        if (n == 2) gjk_line(s, n, d);
        else if (n == 3) gjk_triangle(s, n, d);
        else if (gjk_tetra(s, n, d)) return true;
    }
    return false;
}

__device__ __forceinline__ void wget(const EpaWork* W, int i, Vtx& o) {
    for (int r = 0; r < 3; r++) { o.v[r] = W->v[i][r]; o.a[r] = W->a[i][r]; o.b[r] = W->b[i][r]; }
}
__device__ __forceinline__ void wput(EpaWork* W, int i, const Vtx& o) {
    for (int r = 0; r < 3; r++) { W->v[i][r] = o.v[r]; W->a[i][r] = o.a[r]; W->b[i][r] = o.b[r]; }
}

// One face's term of gjk._closest_face (gjk.py:176-194): unit normal and
// distance, skipped (+inf) when the face is degenerate.
__device__ __forceinline__ void face_eval(EpaWork* W, int i) {
    const double *p0 = W->v[W->f[i][0]], *p1 = W->v[W->f[i][1]], *p2 = W->v[W->f[i][2]];
    double e1[3], e2[3], nn[3];
    sub3(p1, p0, e1); sub3(p2, p0, e2);
    cross3(e1, e2, nn);
    double ln = sqrt((nn[0] * nn[0] + nn[1] * nn[1]) + nn[2] * nn[2]);
    if (ln < GJK_EPS) { W->fd[i] = __longlong_as_double(0x7ff0000000000000ll); return; }
    nn[0] /= ln; nn[1] /= ln; nn[2] /= ln;
    double d = dot3(nn, p0);
    if (d < 0) { nn[0] = -nn[0]; nn[1] = -nn[1]; nn[2] = -nn[2]; d = -d; }
    W->fn[i][0] = nn[0]; W->fn[i][1] = nn[1]; W->fn[i][2] = nn[2]; W->fd[i] = d;
}

// gjk._closest_face (gjk.py:176-194) over the faces' stored terms: the first
// face of least distance
__device__ int closest_face(const EpaWork* W, int nf, double* bn, double& bd) {
    int best = -1;
    bd = __longlong_as_double(0x7ff0000000000000ll);
    for (int i = 0; i < nf; i++) {
        const double d = W->fd[i];
        if (d < bd) { bd = d; best = i; }
    }
    if (best >= 0) { bn[0] = W->fn[best][0]; bn[1] = W->fn[best][1]; bn[2] = W->fn[best][2]; }
    return best;
}

// gjk._rebuild (gjk.py:197-217): horizon edges in first-seen order.  The
// faces that stay keep their order and are packed in place (the write
// position never passes the face being read), then the new faces follow.
__device__ int rebuild(EpaWork* W, int nf, int ni) {
    int ne = 0, nk = 0;
    const double* w = W->v[ni];
    for (int i = 0; i < nf; i++) {
        const double* a = W->v[W->f[i][0]];
        double e1[3], e2[3], nn[3], wa[3];
        sub3(W->v[W->f[i][1]], a, e1); sub3(W->v[W->f[i][2]], a, e2); cross3(e1, e2, nn); sub3(w, a, wa);
        if (dot3(nn, wa) > 0) {
            int es[3][2] = {{W->f[i][0], W->f[i][1]}, {W->f[i][1], W->f[i][2]}, {W->f[i][2], W->f[i][0]}};
            for (int e = 0; e < 3; e++) {
                int found = -1;
                for (int q = 0; q < ne; q++)
                    if (W->ealive[q] && W->ek[q][0] == es[e][1] && W->ek[q][1] == es[e][0]) { found = q; break; }
                if (found >= 0) W->ealive[found] = 0;
                else { W->ek[ne][0] = es[e][0]; W->ek[ne][1] = es[e][1]; W->ealive[ne] = 1; ne++; }
            }
        } else {
            if (nk != i) {
                for (int r = 0; r < 3; r++) { W->f[nk][r] = W->f[i][r]; W->fn[nk][r] = W->fn[i][r]; }
                W->fd[nk] = W->fd[i];
            }
            nk++;
        }
    }
    for (int q = 0; q < ne; q++)
        if (W->ealive[q]) {
            if (nk >= EPA_MAXF) return -1;
            W->f[nk][0] = W->ek[q][0]; W->f[nk][1] = W->ek[q][1]; W->f[nk][2] = ni;
            face_eval(W, nk);
            nk++;
        }
    return nk;
}

// gjk._contact_from_face + _barycentric (gjk.py:220-246)
__device__ void contact_from_face(const EpaWork* W, int fi, const double* nrm, double dist, double* point) {
    const int ia = W->f[fi][0], ib = W->f[fi][1], ic = W->f[fi][2];
    double p[3] = {nrm[0] * dist, nrm[1] * dist, nrm[2] * dist};
    double v0[3], v1[3], v2[3], bary[3];
    sub3(W->v[ib], W->v[ia], v0); sub3(W->v[ic], W->v[ia], v1); sub3(p, W->v[ia], v2);
    double d00 = dot3(v0, v0), d01 = dot3(v0, v1), d11 = dot3(v1, v1);
    double d20 = dot3(v2, v0), d21 = dot3(v2, v1);
    double den = d00 * d11 - d01 * d01;
    if (fabs(den) < GJK_EPS) { bary[0] = 1.0; bary[1] = 0.0; bary[2] = 0.0; }
    else {
        double vv = (d11 * d20 - d01 * d21) / den;
        double ww = (d00 * d21 - d01 * d20) / den;
        bary[0] = 1.0 - vv - ww; bary[1] = vv; bary[2] = ww;
    }
    for (int r = 0; r < 3; r++) {
        double pa = bary[0] * W->a[ia][r] + bary[1] * W->a[ib][r] + bary[2] * W->a[ic][r];
        double pb = bary[0] * W->b[ia][r] + bary[1] * W->b[ib][r] + bary[2] * W->b[ic][r];
        point[r] = 0.5 * (pa + pb);
    }
}

// gjk.epa + _expand_to_tetra (gjk.py:120-173)
__device__ bool epa(const Proxy& A, const Proxy& B, const Vtx* simplex, int ns, EpaWork* W, double* nrm,
                    double& depth, double* point) {
    int nv = ns;
    for (int i = 0; i < ns; i++) wput(W, i, simplex[i]);
    if (nv < 4) {
        const double dirs[6][3] = {{1, 0, 0}, {-1, 0, 0}, {0, 1, 0}, {0, -1, 0}, {0, 0, 1}, {0, 0, -1}};
        for (int k = 0; k < 6 && nv < 4; k++) {
            Vtx w;
            msupport(A, B, dirs[k], w);
            bool ok = true;
            for (int p = 0; p < nv; p++) {
                double e[3];
                sub3(w.v, W->v[p], e);
                if (!(dot3(e, e) > GJK_EPS)) { ok = false; break; }
            }
            if (ok) wput(W, nv++, w);
        }
        if (nv < 4) return false;
        double e1[3], e2[3], e3[3], c[3];
        sub3(W->v[1], W->v[0], e1); sub3(W->v[2], W->v[0], e2); sub3(W->v[3], W->v[0], e3);
        cross3(e1, e2, c);
        if (dot3(c, e3) > 0) { Vtx t1, t2; wget(W, 1, t1); wget(W, 2, t2); wput(W, 1, t2); wput(W, 2, t1); }
    }
    nv = 4;
    const int init[4][3] = {{0, 1, 2}, {0, 2, 3}, {0, 3, 1}, {1, 3, 2}};
    for (int i = 0; i < 4; i++) {
        for (int j = 0; j < 3; j++) W->f[i][j] = init[i][j];
        face_eval(W, i);
    }
    int nf = 4;
    double bn[3], bd;
    for (int it = 0; it < 64; it++) {
        int fi = closest_face(W, nf, bn, bd);
        if (fi < 0) return false;
        Vtx w;
        msupport(A, B, bn, w);
        double d = dot3(w.v, bn);
        if (d - bd < 1e-6) {
            contact_from_face(W, fi, bn, bd, point);
            nrm[0] = bn[0]; nrm[1] = bn[1]; nrm[2] = bn[2]; depth = bd;
            return true;
        }
        wput(W, nv++, w);
        nf = rebuild(W, nf, nv - 1);
        if (nf <= 0) return false;
    }
    int fi = closest_face(W, nf, bn, bd);
    if (fi < 0) return false;
    contact_from_face(W, fi, bn, bd, point);
    nrm[0] = bn[0]; nrm[1] = bn[1]; nrm[2] = bn[2]; depth = bd;
    return true;
}

// gjk.collide_convex (gjk.py:249-264): one contact, normal A -> B by the
// centre hints, none when depth <= 1e-9.
__device__ int collide_convex(const Proxy& A, const Proxy& B, EpaWork* W, double* normal, double* pt, double& depth) {
    Vtx s[4];
    int n = 0;
    if (!gjk(A, B, s, n)) return 0;
    double nrm[3], dep, p[3];
    if (!epa(A, B, s, n, W, nrm, dep, p)) return 0;
    double ha[3], hb[3], dh[3];
    center_hint(A, ha); center_hint(B, hb);
    sub3(hb, ha, dh);
    if (dot3(nrm, dh) < 0) { nrm[0] = -nrm[0]; nrm[1] = -nrm[1]; nrm[2] = -nrm[2]; }
    if (dep <= 1e-9) return 0;
    double ln = sqrt((nrm[0] * nrm[0] + nrm[1] * nrm[1]) + nrm[2] * nrm[2]);
    double s_ = maxd(ln, 1e-12);
    normal[0] = nrm[0] / s_; normal[1] = nrm[1] / s_; normal[2] = nrm[2] / s_;
    pt[0] = p[0]; pt[1] = p[1]; pt[2] = p[2];
    depth = dep;
    return 1;
}

}  // namespace omi
