// omi_math.hpp — device math of the contact pipeline (gfx950).
//
// Every function is templated on the arithmetic type R (double or float).  With
// R = double the operation order is the reference's, term by term, so that the
// FP64 pipeline reproduces the reference bit for bit (the library is compiled
// with -ffp-contract=off: no multiply-add fusion, IEEE division and sqrt).
// With R = float the callers move the pair into a pair-local frame first (the
// offset is taken in FP64) so FP32 only ever sees O(1) coordinates.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <type_traits>

namespace omi {

enum : int32_t { K_NONE = 0, K_SPHERE = 1, K_BOX = 2, K_CONVEX = 3 };
enum : int32_t { M_STATIC = 0, M_KINEMATIC = 1, M_DYNAMIC = 2 };

template <typename R> struct Lim;
template <> struct Lim<double> { static constexpr double big = 1e300; };
template <> struct Lim<float> { static constexpr float big = 3.0e38f; };

template <typename R> __device__ __forceinline__ R rsqrt_(R x) { return sqrt(x); }
template <typename R> __device__ __forceinline__ R abs_(R x) { return fabs(x); }
__device__ __forceinline__ double maxd(double a, double b) { return a > b ? a : b; }

// mathutil.quat_to_matrix (mathutil.py:59-76) after quat_normalize (:21-27);
// the norm adds the squares left to right like np.linalg.norm over axis -1.
__device__ __forceinline__ void quat_to_matrix(const double q[4], double R[9]) {
    double n = sqrt(((q[0] * q[0] + q[1] * q[1]) + q[2] * q[2]) + q[3] * q[3]);
    double d = maxd(n, 1e-12);
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

// Box half-width along world axis r: einsum('nij,nj->ni', |R|, half) sums the
// three products as (p0 + p2) + p1 in the reference's NumPy.
__device__ __forceinline__ double box_hw(const double R[9], const double h[3], int r) {
    double a0 = fabs(R[3 * r]) * h[0], a1 = fabs(R[3 * r + 1]) * h[1], a2 = fabs(R[3 * r + 2]) * h[2];
    return (a0 + a2) + a1;
}

// ---------------------------------------------------------------- box-box --
// _collide_native._face_axis (_collide_native.pyx:79-89)
template <typename R>
__device__ __forceinline__ int face_axis(const R* M, R d0, R d1, R d2, R& sign) {
    R best = R(-1);
    int axis = 0;
#pragma unroll
    for (int m = 0; m < 3; m++) {
        R dv = M[m] * d0 + M[3 + m] * d1 + M[6 + m] * d2;
        if (abs_(dv) > best) {
            best = abs_(dv);
            axis = m;
            sign = dv >= R(0) ? R(1) : R(-1);
        }
    }
    return axis;
}

// _collide_native._face_vertices (_collide_native.pyx:92-110); v[12]
// A value the compiler must treat as computed, not loaded: a select between
// two such values stays a select, instead of becoming one load from a
// run-time address (which would pin the source array in private memory).
template <typename R>
__device__ __forceinline__ R opaque(R x) {
    asm volatile("" : "+v"(x));
    return x;
}

template <typename R>
__device__ __forceinline__ void face_vertices(const R* c, const R* M, const R* h, int axis, R sign, R* v) {
    int o0 = axis == 0 ? 1 : 0, o1 = axis != 2 ? 2 : 1;
    if (axis == 1) { o0 = 0; o1 = 2; }
    const R hx = opaque(h[0]), hy = opaque(h[1]), hz = opaque(h[2]);
    R ha = axis == 0 ? hx : (axis == 1 ? hy : hz);
    R h0 = o0 == 0 ? hx : hy;
    R h1 = o1 == 1 ? hy : hz;
#pragma unroll
    for (int r = 0; r < 3; r++) {
        const R m0 = opaque(M[3 * r]), m1 = opaque(M[3 * r + 1]), m2 = opaque(M[3 * r + 2]);
        R col = axis == 0 ? m0 : (axis == 1 ? m1 : m2);
        R c0 = o0 == 0 ? m0 : m1;
        R c1 = o1 == 1 ? m1 : m2;
        R fc = c[r] + col * sign * ha;
        R u = c0 * h0, w = c1 * h1;
        v[r] = fc + u + w;
        v[3 + r] = fc - u + w;
        v[6 + r] = fc - u - w;
        v[9 + r] = fc + u - w;
    }
}

// _collide_native._reduce (_collide_native.pyx:212-263)
template <typename R>
__device__ int reduce4(const R* P, const R* D, int nc, R* op, R* od) {
    if (nc <= 4) {
        for (int i = 0; i < nc; i++) {
            op[3 * i] = P[3 * i]; op[3 * i + 1] = P[3 * i + 1]; op[3 * i + 2] = P[3 * i + 2];
            od[i] = D[i];
        }
        return nc;
    }
    int k0 = 0, k1 = 0, k2 = 0, k3 = 0;
    R best = P[0];
    for (int i = 1; i < nc; i++) if (P[3 * i] > best) { best = P[3 * i]; k0 = i; }
    best = R(-1);
    for (int i = 0; i < nc; i++) {
        R dd = R(0), d0;
        for (int r = 0; r < 3; r++) { d0 = P[3 * i + r] - P[3 * k0 + r]; dd += d0 * d0; }
        if (dd > best) { best = dd; k1 = i; }
    }
    best = R(-1);
    for (int i = 0; i < nc; i++) {
        R ds = R(0), dd = R(0), d0;
        for (int r = 0; r < 3; r++) {
            d0 = P[3 * i + r] - P[3 * k1 + r]; ds += d0 * d0;
            d0 = P[3 * i + r] - P[3 * k0 + r]; dd += d0 * d0;
        }
        R s = sqrt(ds) + sqrt(dd);
        if (s > best) { best = s; k2 = i; }
    }
    best = R(-1);
    for (int i = 0; i < nc; i++) {
        R dd = R(0), d0;
        for (int r = 0; r < 3; r++) { d0 = P[3 * i + r] - P[3 * k2 + r]; dd += d0 * d0; }
        R s = sqrt(dd);
        if (s > best) { best = s; k3 = i; }
    }
    int keep[4] = {k0, k1, k2, k3}, nk = 0;
    for (int i = 0; i < 4; i++) {
        bool dup = false;
        for (int j = 0; j < nk; j++) dup = dup || keep[i] == keep[j];
        if (!dup) keep[nk++] = keep[i];
    }
    for (int i = 0; i < nk; i++) {
        op[3 * i] = P[3 * keep[i]]; op[3 * i + 1] = P[3 * keep[i] + 1]; op[3 * i + 2] = P[3 * keep[i] + 2];
        od[i] = D[keep[i]];
    }
    return nk;
}

#define OMI_POLY_MAX 16

// _collide_native._clip (_collide_native.pyx:113-209) on 16-vertex buffers in
// private memory: the path for the rare polygon that outgrows the 8 register
// slots of clip_manifold below.
template <typename R>
__device__ __noinline__ int clip_manifold_general(const R* ca, const R* Ra, const R* ha, const R* cb, const R* Rb,
                                                  const R* hb, const R* nrm, R depth, R* op, R* od) {
    R a_align = R(-1), b_align = R(-1);
#pragma unroll
    for (int m = 0; m < 3; m++) {
        R dv = abs_(Ra[m] * nrm[0] + Ra[3 + m] * nrm[1] + Ra[6 + m] * nrm[2]);
        if (dv > a_align) a_align = dv;
        dv = abs_(Rb[m] * nrm[0] + Rb[3 + m] * nrm[1] + Rb[6 + m] * nrm[2]);
        if (dv > b_align) b_align = dv;
    }
    const bool flip = !(a_align >= b_align);
    const R *rc = flip ? cb : ca, *rR = flip ? Rb : Ra, *rh = flip ? hb : ha;
    const R *ic = flip ? ca : cb, *iR = flip ? Ra : Rb, *ih = flip ? ha : hb;
    R rn[3];
#pragma unroll
    for (int r = 0; r < 3; r++) rn[r] = flip ? -nrm[r] : nrm[r];
    R isign = R(1), rsign = R(1);
    int iaxis = face_axis(iR, -rn[0], -rn[1], -rn[2], isign);
    R poly[OMI_POLY_MAX * 3], buf[OMI_POLY_MAX * 3], refv[12];
    int npoly = 4;
    face_vertices(ic, iR, ih, iaxis, isign, poly);
    int raxis = face_axis(rR, rn[0], rn[1], rn[2], rsign);
    face_vertices(rc, rR, rh, raxis, rsign, refv);
    for (int other = 0; other < 3 && npoly > 0; other++) {
        if (other == raxis) continue;
        for (int si = 0; si < 2; si++) {
            R s = si == 0 ? R(1) : R(-1);
            R p0 = rR[other] * s, p1 = rR[3 + other] * s, p2 = rR[6 + other] * s;
            R offset = (rc[0] * p0 + rc[1] * p1 + rc[2] * p2) + rh[other];
            int nb = 0;
            for (int i = 0; i < npoly; i++) {
                int j = i + 1 == npoly ? 0 : i + 1;
                R dc = p0 * poly[3 * i] + p1 * poly[3 * i + 1] + p2 * poly[3 * i + 2] - offset;
                R dn = p0 * poly[3 * j] + p1 * poly[3 * j + 1] + p2 * poly[3 * j + 2] - offset;
                if (dc <= R(0) && nb < OMI_POLY_MAX) {
                    buf[3 * nb] = poly[3 * i]; buf[3 * nb + 1] = poly[3 * i + 1]; buf[3 * nb + 2] = poly[3 * i + 2];
                    nb++;
                }
                if (dc * dn < R(0) && nb < OMI_POLY_MAX) {
                    R t = dc / (dc - dn);
                    for (int r = 0; r < 3; r++) buf[3 * nb + r] = poly[3 * i + r] + t * (poly[3 * j + r] - poly[3 * i + r]);
                    nb++;
                }
            }
            npoly = nb;
            for (int i = 0; i < 3 * npoly; i++) poly[i] = buf[i];
            if (npoly == 0) break;
        }
    }
    R allp[OMI_POLY_MAX * 3], alld[OMI_POLY_MAX];
    int nc = 0;
    for (int i = 0; i < npoly; i++) {
        R pen = (refv[0] - poly[3 * i]) * rn[0] + (refv[1] - poly[3 * i + 1]) * rn[1] + (refv[2] - poly[3 * i + 2]) * rn[2];
        if (pen >= R(-1e-6)) {
            for (int r = 0; r < 3; r++) allp[3 * nc + r] = poly[3 * i + r] + rn[r] * pen;
            alld[nc] = pen > R(0) ? pen : R(0);
            nc++;
        }
    }
    if (nc == 0) {
        for (int r = 0; r < 3; r++) op[r] = R(0.5) * (ca[r] + cb[r]);
        od[0] = depth;
        return 1;
    }
    return reduce4(allp, alld, nc, op, od);
}

// A clip polygon of up to 8 vertices in registers.  Slots are written through
// one-hot selects and read at indices that are static after unrolling, so no
// slot is ever addressed at run time (which would put the polygon in private
// memory).  A convex quadrilateral clipped by 4 half-planes has at most 8
// vertices; more only arise from rounding on near-degenerate input, and then
// the clip reports overflow and the caller takes the 16-vertex path.
template <typename R>
struct Poly8 {
    R x[8], y[8], z[8];
};

// f(integral_constant<int, I>) for I = B .. N-1: loop indices that are
// constant expressions
template <int B, int N, typename F>
__device__ __forceinline__ void sfor(F&& f) {
    if constexpr (B < N) {
        f(std::integral_constant<int, B>{});
        sfor<B + 1, N>(f);
    }
}

// Slot k of P with k run-time: a select over the 8 slots.  The selects act
// on opaque values and the writes below store every slot unconditionally, so
// the compiler cannot fold them into one access at a run-time index.
template <typename R>
__device__ __forceinline__ R sel8(const R (&a)[8], int k) {
    R x = opaque(a[0]);
#pragma unroll
    for (int s = 1; s < 8; s++) x = k == s ? opaque(a[s]) : x;
    return x;
}

template <typename R>
__device__ __forceinline__ void get8(const Poly8<R>& P, int k, R& x, R& y, R& z) {
    x = sel8(P.x, k); y = sel8(P.y, k); z = sel8(P.z, k);
}

// write (x, y, z) to slot k of Q; KMAX bounds the slots k can reach
template <int KMAX, typename R>
__device__ __forceinline__ void put8(Poly8<R>& Q, int k, R x, R y, R z) {
#pragma unroll
    for (int s = 0; s < (KMAX < 8 ? KMAX : 8); s++) {
        const bool h = k == s;
        Q.x[s] = h ? x : Q.x[s]; Q.y[s] = h ? y : Q.y[s]; Q.z[s] = h ? z : Q.z[s];
    }
}

// One Sutherland-Hodgman side plane (_collide_native.pyx:171-188): P (n
// vertices) into Q, returning Q's count.  The distance of vertex (i + 1) % n
// is the same expression on the same operands as vertex i + 1's own, so it is
// reused.
template <typename R>
__device__ __forceinline__ int clip_side8(const Poly8<R>& P, int n, R p0, R p1, R p2, R off, Poly8<R>& Q,
                                          bool& over) {
    R dc[8];
#pragma unroll
    for (int i = 0; i < 8; i++) dc[i] = p0 * P.x[i] + p1 * P.y[i] + p2 * P.z[i] - off;
    int nb = 0;
    sfor<0, 8>([&](auto I) {
        constexpr int i = decltype(I)::value;
        constexpr int j = i + 1 < 8 ? i + 1 : 0;
        if (i < n) {
            const bool wrap = i + 1 == n;
            const R dn = wrap ? dc[0] : dc[j];
            const R xj = wrap ? P.x[0] : P.x[j], yj = wrap ? P.y[0] : P.y[j], zj = wrap ? P.z[0] : P.z[j];
            if (dc[i] <= R(0)) {
                over = over || nb >= 8;
                put8<2 * i + 1>(Q, nb, P.x[i], P.y[i], P.z[i]);
                nb++;
            }
            if (dc[i] * dn < R(0)) {
                const R t = dc[i] / (dc[i] - dn);
                over = over || nb >= 8;
                put8<2 * i + 2>(Q, nb, P.x[i] + t * (xj - P.x[i]), P.y[i] + t * (yj - P.y[i]),
                                P.z[i] + t * (zj - P.z[i]));
                nb++;
            }
        }
    });
    return nb;
}

// _collide_native._reduce (_collide_native.pyx:212-263) on up to 8 points in
// registers; the same selection order as reduce4.
template <typename R>
__device__ __forceinline__ int reduce8(const Poly8<R>& P, const R (&D)[8], int nc, R* op, R* od) {
    if (nc <= 4) {
#pragma unroll
        for (int i = 0; i < 4; i++)
            if (i < nc) { op[3 * i] = P.x[i]; op[3 * i + 1] = P.y[i]; op[3 * i + 2] = P.z[i]; od[i] = D[i]; }
        return nc;
    }
    int k0 = 0, k1 = 0, k2 = 0, k3 = 0;
    R best = P.x[0];
#pragma unroll
    for (int i = 1; i < 8; i++) if (i < nc && P.x[i] > best) { best = P.x[i]; k0 = i; }
    R c0x, c0y, c0z, c1x, c1y, c1z, c2x, c2y, c2z;
    get8(P, k0, c0x, c0y, c0z);
    best = R(-1);
#pragma unroll
    for (int i = 0; i < 8; i++) {
        if (i >= nc) continue;
        R dd = R(0), d0;
        d0 = P.x[i] - c0x; dd += d0 * d0;
        d0 = P.y[i] - c0y; dd += d0 * d0;
        d0 = P.z[i] - c0z; dd += d0 * d0;
        if (dd > best) { best = dd; k1 = i; }
    }
    get8(P, k1, c1x, c1y, c1z);
    best = R(-1);
#pragma unroll
    for (int i = 0; i < 8; i++) {
        if (i >= nc) continue;
        R ds = R(0), dd = R(0), d0;
        d0 = P.x[i] - c1x; ds += d0 * d0;
        d0 = P.x[i] - c0x; dd += d0 * d0;
        d0 = P.y[i] - c1y; ds += d0 * d0;
        d0 = P.y[i] - c0y; dd += d0 * d0;
        d0 = P.z[i] - c1z; ds += d0 * d0;
        d0 = P.z[i] - c0z; dd += d0 * d0;
        const R sd = sqrt(ds) + sqrt(dd);
        if (sd > best) { best = sd; k2 = i; }
    }
    get8(P, k2, c2x, c2y, c2z);
    best = R(-1);
#pragma unroll
    for (int i = 0; i < 8; i++) {
        if (i >= nc) continue;
        R dd = R(0), d0;
        d0 = P.x[i] - c2x; dd += d0 * d0;
        d0 = P.y[i] - c2y; dd += d0 * d0;
        d0 = P.z[i] - c2z; dd += d0 * d0;
        const R sd = sqrt(dd);
        if (sd > best) { best = sd; k3 = i; }
    }
    // first occurrences of k0..k3, in that order
    const int keep[4] = {k0, k1, k2, k3};
    int nk = 0;
#pragma unroll
    for (int i = 0; i < 4; i++) {
        bool dup = false;
#pragma unroll
        for (int j = 0; j < i; j++) dup = dup || keep[i] == keep[j];
        if (!dup) {
            R x, y, z;
            get8(P, keep[i], x, y, z);
            const R dv = sel8(D, keep[i]);
#pragma unroll
            for (int q = 0; q <= i; q++) {
                const bool h = nk == q;
                op[3 * q] = h ? x : op[3 * q]; op[3 * q + 1] = h ? y : op[3 * q + 1];
                op[3 * q + 2] = h ? z : op[3 * q + 2]; od[q] = h ? dv : od[q];
            }
            nk++;
        }
    }
    return nk;
}

// _collide_native._clip (_collide_native.pyx:113-209): reference face = the
// box face most aligned with the normal, incident face = the other box's face
// most anti-parallel to it, clipped against the reference face's side planes;
// points below the reference face are seated on it and reduced to 4.
// Without GENERAL a polygon that outgrows the 8 slots returns -1 (the caller
// runs the pair again with it), so the kernel holds no private arrays.
template <typename R, bool GENERAL = true>
__device__ __forceinline__ int clip_manifold(const R* ca, const R* Ra, const R* ha, const R* cb, const R* Rb,
                                             const R* hb, const R* nrm, R depth, R* op, R* od) {
    R a_align = R(-1), b_align = R(-1);
#pragma unroll
    for (int m = 0; m < 3; m++) {
        R dv = abs_(Ra[m] * nrm[0] + Ra[3 + m] * nrm[1] + Ra[6 + m] * nrm[2]);
        if (dv > a_align) a_align = dv;
        dv = abs_(Rb[m] * nrm[0] + Rb[3 + m] * nrm[1] + Rb[6 + m] * nrm[2]);
        if (dv > b_align) b_align = dv;
    }
    const bool flip = !(a_align >= b_align);
    R rc[3], rR[9], rh[3], ic[3], iR[9], ih[3], rn[3];
#pragma unroll
    for (int r = 0; r < 3; r++) {
        rc[r] = flip ? cb[r] : ca[r]; rh[r] = flip ? hb[r] : ha[r];
        ic[r] = flip ? ca[r] : cb[r]; ih[r] = flip ? ha[r] : hb[r];
        rn[r] = flip ? -nrm[r] : nrm[r];
    }
#pragma unroll
    for (int r = 0; r < 9; r++) { rR[r] = flip ? Rb[r] : Ra[r]; iR[r] = flip ? Ra[r] : Rb[r]; }
    R isign = R(1), rsign = R(1);
    const int iaxis = face_axis(iR, -rn[0], -rn[1], -rn[2], isign);
    R v[12], refv[12];
    face_vertices(ic, iR, ih, iaxis, isign, v);
    const int raxis = face_axis(rR, rn[0], rn[1], rn[2], rsign);
    face_vertices(rc, rR, rh, raxis, rsign, refv);
    Poly8<R> A, B;
#pragma unroll
    for (int i = 0; i < 8; i++) {
        A.x[i] = i < 4 ? v[3 * i] : R(0); A.y[i] = i < 4 ? v[3 * i + 1] : R(0); A.z[i] = i < 4 ? v[3 * i + 2] : R(0);
        B.x[i] = R(0); B.y[i] = R(0); B.z[i] = R(0);
    }
    // side planes: the two axes other than raxis, + then - each
    const int o0 = raxis == 0 ? 1 : 0, o1 = raxis == 2 ? 1 : 2;
    bool over = false;
    int np = 4;
#pragma unroll
    for (int q = 0; q < 4; q++) {
        const int other = q < 2 ? o0 : o1;
        const R s = (q & 1) ? R(-1) : R(1);
        const R c0 = other == 0 ? rR[0] : (other == 1 ? rR[1] : rR[2]);
        const R c1 = other == 0 ? rR[3] : (other == 1 ? rR[4] : rR[5]);
        const R c2 = other == 0 ? rR[6] : (other == 1 ? rR[7] : rR[8]);
        const R hw = other == 0 ? rh[0] : (other == 1 ? rh[1] : rh[2]);
        const R p0 = c0 * s, p1 = c1 * s, p2 = c2 * s;
        const R offset = (rc[0] * p0 + rc[1] * p1 + rc[2] * p2) + hw;
        // an empty polygon stays empty (the reference stops clipping)
        if (q & 1) np = clip_side8(B, np, p0, p1, p2, offset, A, over);
        else np = clip_side8(A, np, p0, p1, p2, offset, B, over);
    }
    if (over) {
        if constexpr (!GENERAL) return -1;
        // copies that only this rare branch hands to the out-of-line path, so
        // the caller's arrays stay in registers
        R gca[3], gRa[9], gha[3], gcb[3], gRb[9], ghb[3], gn[3], gp[12], gd[4];
#pragma unroll
        for (int r = 0; r < 3; r++) { gca[r] = ca[r]; gha[r] = ha[r]; gcb[r] = cb[r]; ghb[r] = hb[r]; gn[r] = nrm[r]; }
#pragma unroll
        for (int r = 0; r < 9; r++) { gRa[r] = Ra[r]; gRb[r] = Rb[r]; }
        const int n = clip_manifold_general(gca, gRa, gha, gcb, gRb, ghb, gn, depth, gp, gd);
#pragma unroll
        for (int r = 0; r < 12; r++) op[r] = gp[r];
#pragma unroll
        for (int r = 0; r < 4; r++) od[r] = gd[r];
        return n;
    }
    // the polygon is in A (an even number of planes)
    Poly8<R> C;
    R dc[8];
    int nc = 0;
#pragma unroll
    for (int i = 0; i < 8; i++) {
        C.x[i] = R(0); C.y[i] = R(0); C.z[i] = R(0); dc[i] = R(0);
    }
    sfor<0, 8>([&](auto I) {
        constexpr int i = decltype(I)::value;
        if (i < np) {
            const R pen = (refv[0] - A.x[i]) * rn[0] + (refv[1] - A.y[i]) * rn[1] + (refv[2] - A.z[i]) * rn[2];
            if (pen >= R(-1e-6)) {
                const R d = pen > R(0) ? pen : R(0);
                put8<i + 1>(C, nc, A.x[i] + rn[0] * pen, A.y[i] + rn[1] * pen, A.z[i] + rn[2] * pen);
#pragma unroll
                for (int s = 0; s <= i; s++) dc[s] = nc == s ? d : dc[s];
                nc++;
            }
        }
    });
    if (nc == 0) {
        for (int r = 0; r < 3; r++) op[r] = R(0.5) * (ca[r] + cb[r]);
        od[0] = depth;
        return 1;
    }
    return reduce8(C, dc, nc, op, od);
}

// _collide_native._box_box (_collide_native.pyx:20-76): 15 SAT axes generated
// on the fly (A's faces, B's faces, the non-degenerate edge crosses), first
// minimum overlap wins, sign from the centre offset.  -1 without GENERAL
// when the clip outgrows its register polygon (see clip_manifold).
template <typename R, bool GENERAL = true>
__device__ __forceinline__ int box_box(const R* ca, const R* Ra, const R* ha, const R* cb, const R* Rb,
                                       const R* hb, R* normal, R* pts, R* deps) {
    const R d0 = cb[0] - ca[0], d1 = cb[1] - ca[1], d2 = cb[2] - ca[2];
    R best = Lim<R>::big, b0 = R(0), b1 = R(0), b2 = R(0);
    bool sep = false;
    auto test = [&](R x0, R x1, R x2) {
        R ra = abs_(x0 * Ra[0] + x1 * Ra[3] + x2 * Ra[6]) * ha[0] +
               abs_(x0 * Ra[1] + x1 * Ra[4] + x2 * Ra[7]) * ha[1] +
               abs_(x0 * Ra[2] + x1 * Ra[5] + x2 * Ra[8]) * ha[2];
        R rb = abs_(x0 * Rb[0] + x1 * Rb[3] + x2 * Rb[6]) * hb[0] +
               abs_(x0 * Rb[1] + x1 * Rb[4] + x2 * Rb[7]) * hb[1] +
               abs_(x0 * Rb[2] + x1 * Rb[5] + x2 * Rb[8]) * hb[2];
        R proj = x0 * d0 + x1 * d1 + x2 * d2;
        R ov = ra + rb - abs_(proj);
        if (ov < R(0)) { sep = true; return; }
        if (ov < best) {
            best = ov;
            if (proj >= R(0)) { b0 = x0; b1 = x1; b2 = x2; }
            else { b0 = -x0; b1 = -x1; b2 = -x2; }
        }
    };
#pragma unroll
    for (int i = 0; i < 3; i++) { test(Ra[i], Ra[3 + i], Ra[6 + i]); if (sep) return 0; }
#pragma unroll
    for (int i = 0; i < 3; i++) { test(Rb[i], Rb[3 + i], Rb[6 + i]); if (sep) return 0; }
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int j = 0; j < 3; j++) {
            R cx = Ra[3 + i] * Rb[6 + j] - Ra[6 + i] * Rb[3 + j];
            R cy = Ra[6 + i] * Rb[j] - Ra[i] * Rb[6 + j];
            R cz = Ra[i] * Rb[3 + j] - Ra[3 + i] * Rb[j];
            R l2 = cx * cx + cy * cy + cz * cz;
            if (l2 > R(1e-9) * R(1e-9)) {
                R inv = R(1) / sqrt(l2);
                test(cx * inv, cy * inv, cz * inv);
                if (sep) return 0;
            }
        }
    normal[0] = b0; normal[1] = b1; normal[2] = b2;
    return clip_manifold<R, GENERAL>(ca, Ra, ha, cb, Rb, hb, normal, best, pts, deps);
}

// _collide_native.sphere_sphere (_collide_native.pyx:285-299)
template <typename R>
__device__ __forceinline__ int sphere_sphere(const R* ca, R ra, const R* cb, R rb, R* n, R* pt, R* dep) {
    R dx = cb[0] - ca[0], dy = cb[1] - ca[1], dz = cb[2] - ca[2];
    R dist = sqrt(dx * dx + dy * dy + dz * dz);
    R r = ra + rb;
    if (dist >= r) return 0;
    if (dist > R(1e-9)) { n[0] = dx / dist; n[1] = dy / dist; n[2] = dz / dist; }
    else { n[0] = R(0); n[1] = R(1); n[2] = R(0); }
    R depth = r - dist, f = ra - R(0.5) * depth;
    pt[0] = ca[0] + n[0] * f; pt[1] = ca[1] + n[1] * f; pt[2] = ca[2] + n[2] * f;
    *dep = depth;
    return 1;
}

// _collide_native.sphere_box (_collide_native.pyx:302-344); normal sphere -> box
template <typename R>
__device__ __forceinline__ int sphere_box(const R* cs, R rs, const R* cb, const R* M, const R* hb, R* n, R* pt,
                                          R* dep) {
    R sx = cs[0] - cb[0], sy = cs[1] - cb[1], sz = cs[2] - cb[2];
    R lx = M[0] * sx + M[3] * sy + M[6] * sz;
    R ly = M[1] * sx + M[4] * sy + M[7] * sz;
    R lz = M[2] * sx + M[5] * sy + M[8] * sz;
    R cx = lx, cy = ly, cz = lz;
    if (cx > hb[0]) cx = hb[0]; else if (cx < -hb[0]) cx = -hb[0];
    if (cy > hb[1]) cy = hb[1]; else if (cy < -hb[1]) cy = -hb[1];
    if (cz > hb[2]) cz = hb[2]; else if (cz < -hb[2]) cz = -hb[2];
    R dx = lx - cx, dy = ly - cy, dz = lz - cz;
    R dist = sqrt(dx * dx + dy * dy + dz * dz);
    if (dist > rs) return 0;
    R nlx = R(0), nly = R(0), nlz = R(0), depth;
    if (dist > R(1e-9)) {
        nlx = dx / dist; nly = dy / dist; nlz = dz / dist;
        depth = rs - dist;
    } else {
        R m0 = hb[0] - (lx >= R(0) ? lx : -lx);
        R m1 = hb[1] - (ly >= R(0) ? ly : -ly);
        R m2 = hb[2] - (lz >= R(0) ? lz : -lz);
        int axis = 0;
        R mm = m0;
        if (m1 < mm) { axis = 1; mm = m1; }
        if (m2 < mm) { axis = 2; mm = m2; }
        if (axis == 0) nlx = lx >= R(0) ? R(1) : R(-1);
        else if (axis == 1) nly = ly >= R(0) ? R(1) : R(-1);
        else nlz = lz >= R(0) ? R(1) : R(-1);
        depth = rs + mm;
    }
    R nx = M[0] * nlx + M[1] * nly + M[2] * nlz;
    R ny = M[3] * nlx + M[4] * nly + M[5] * nlz;
    R nz = M[6] * nlx + M[7] * nly + M[8] * nlz;
    pt[0] = cb[0] + M[0] * cx + M[1] * cy + M[2] * cz;
    pt[1] = cb[1] + M[3] * cx + M[4] * cy + M[5] * cz;
    pt[2] = cb[2] + M[6] * cx + M[7] * cy + M[8] * cz;
    n[0] = -nx; n[1] = -ny; n[2] = -nz;
    *dep = depth;
    return 1;
}

}  // namespace omi
