// newt_math.hpp — scalar, vector, quaternion and spatial algebra shared by the
// host model compiler and the gfx950 kernels.
//
// Templated on the arithmetic type R (double for the FP64 step, float for the
// FP32 bulk of the mixed step).  Operation order follows newt's expression
// trees (newt/src/math.rs, newt/src/spatial.rs, newt/src/aba.rs) so the FP64
// kernel reproduces the CPU step's rounding; the kernels are compiled with
// FP contraction off for the same reason.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#define NEWT_HD __host__ __device__ __forceinline__

namespace newt {

template <class R> NEWT_HD R rsqrt_(R x);
template <> NEWT_HD double rsqrt_<double>(double x) { return __builtin_sqrt(x); }
template <> NEWT_HD float rsqrt_<float>(float x) { return __builtin_sqrtf(x); }
template <class R> NEWT_HD R nsqrt(R x) { return rsqrt_<R>(x); }
template <class R> NEWT_HD R nabs(R x) { return __builtin_fabs(x); }
template <> NEWT_HD float nabs<float>(float x) { return __builtin_fabsf(x); }
// IEEE maxNum / minNum (newt's f32::max / f32::min ignore a NaN operand).
template <class R> NEWT_HD R nmax(R a, R b) { return __builtin_fmax(a, b); }
template <class R> NEWT_HD R nmin(R a, R b) { return __builtin_fmin(a, b); }
template <> NEWT_HD float nmax<float>(float a, float b) { return __builtin_fmaxf(a, b); }
template <> NEWT_HD float nmin<float>(float a, float b) { return __builtin_fminf(a, b); }
template <class R> NEWT_HD bool nfinite(R x) { return x == x && x - x == R(0); }

// `x as i32`: truncation, saturating, NaN -> 0 (newt/src/math.rs:126-130).
template <class R> NEWT_HD int32_t sat_i32(R x) {
    if (x != x) return 0;
    if (x >= R(2147483647.0)) return 2147483647;
    if (x <= R(-2147483648.0)) return -2147483647 - 1;
    return (int32_t)x;
}

// Deterministic sin/cos: Cody-Waite reduction by pi/2 with newt's three
// constants, then degree-7 / degree-8 Taylor polynomials in Horner form
// (newt/src/math.rs:45-134).  Device libm is never used: the polynomial is
// part of the step's definition.
template <class R> NEWT_HD R poly_sin(R x) {
    R x2 = x * x;
    R a7 = R(-1.0) / R(5040.0), a5 = R(1.0) / R(120.0), a3 = R(-1.0) / R(6.0);
    return x * (R(1.0) + x2 * (a3 + x2 * (a5 + x2 * a7)));
}
template <class R> NEWT_HD R poly_cos(R x) {
    R x2 = x * x;
    R b8 = R(1.0) / R(40320.0), b6 = R(-1.0) / R(720.0), b4 = R(1.0) / R(24.0), b2 = R(-1.0) / R(2.0);
    return R(1.0) + x2 * (b2 + x2 * (b4 + x2 * (b6 + x2 * b8)));
}
template <class R> NEWT_HD R reduce_quadrant(R x, int32_t* quadrant) {
    const R two_over_pi = R(0.63661974);
    const R c1 = R(1.5703125), c2 = R(4.8370362e-4), c3 = R(7.5497895e-8);
    R scaled = x * two_over_pi;
    int32_t q = scaled >= R(0.0) ? sat_i32(scaled + R(0.5)) : sat_i32(scaled - R(0.5));
    R k = (R)q;
    *quadrant = q;
    return ((x - k * c1) - k * c2) - k * c3;
}
template <class R> NEWT_HD void nsincos(R x, R* s, R* c) {
    int32_t q;
    R r = reduce_quadrant(x, &q);
    R ps = poly_sin(r), pc = poly_cos(r);
    switch (q & 3) {
        case 0: *s = ps; *c = pc; break;
        case 1: *s = pc; *c = -ps; break;
        case 2: *s = -ps; *c = -pc; break;
        default: *s = -pc; *c = ps; break;
    }
}

template <class R> struct V3 {
    R x, y, z;
};
template <class R> NEWT_HD V3<R> mk3(R x, R y, R z) { return V3<R>{x, y, z}; }
template <class R> NEWT_HD V3<R> z3() { return V3<R>{R(0), R(0), R(0)}; }
template <class R> NEWT_HD V3<R> operator+(V3<R> a, V3<R> b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
template <class R> NEWT_HD V3<R> operator-(V3<R> a, V3<R> b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
template <class R> NEWT_HD V3<R> operator-(V3<R> a) { return {-a.x, -a.y, -a.z}; }
template <class R> NEWT_HD V3<R> operator*(V3<R> a, R s) { return {a.x * s, a.y * s, a.z * s}; }
template <class R> NEWT_HD V3<R> operator/(V3<R> a, R s) { return {a.x / s, a.y / s, a.z / s}; }
template <class R> NEWT_HD R dot3(V3<R> a, V3<R> b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
template <class R> NEWT_HD V3<R> cross3(V3<R> a, V3<R> b) {
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
template <class R> NEWT_HD R len3(V3<R> a) { return nsqrt(dot3(a, a)); }
template <class R> NEWT_HD V3<R> unit3(V3<R> a) {
    R l = len3(a);
    return l == R(0) ? z3<R>() : a / l;
}
template <class R> NEWT_HD V3<R> abs3(V3<R> a) { return {nabs(a.x), nabs(a.y), nabs(a.z)}; }
template <class R> NEWT_HD bool is_zero3(V3<R> a) { return a.x == R(0) && a.y == R(0) && a.z == R(0); }

// Quaternion (x, y, z, w), Hamilton product (newt/src/math.rs:560-683).
template <class R> struct Qt {
    R x, y, z, w;
};
template <class R> NEWT_HD Qt<R> qid() { return Qt<R>{R(0), R(0), R(0), R(1)}; }
template <class R> NEWT_HD Qt<R> qmul(Qt<R> a, Qt<R> r) {
    return {a.w * r.x + a.x * r.w + a.y * r.z - a.z * r.y, a.w * r.y - a.x * r.z + a.y * r.w + a.z * r.x,
            a.w * r.z + a.x * r.y - a.y * r.x + a.z * r.w, a.w * r.w - a.x * r.x - a.y * r.y - a.z * r.z};
}
template <class R> NEWT_HD Qt<R> qconj(Qt<R> q) { return {-q.x, -q.y, -q.z, q.w}; }
template <class R> NEWT_HD Qt<R> qaxis(V3<R> axis, R angle) {
    V3<R> a = unit3(axis);
    R s, c;
    nsincos(angle * R(0.5), &s, &c);
    return {a.x * s, a.y * s, a.z * s, c};
}
template <class R> NEWT_HD Qt<R> qnormalize(Qt<R> q) {
    R n = nsqrt(q.x * q.x + q.y * q.y + q.z * q.z + q.w * q.w);
    if (n == R(0)) return qid<R>();
    return {q.x / n, q.y / n, q.z / n, q.w / n};
}
template <class R> NEWT_HD V3<R> qrot(Qt<R> q, V3<R> v) {
    V3<R> qv{q.x, q.y, q.z};
    V3<R> t = cross3(qv, v) * R(2.0);
    return v + t * q.w + cross3(qv, t);
}
template <class R> NEWT_HD V3<R> qrot_inv(Qt<R> q, V3<R> v) { return qrot(qconj(q), v); }
// Quat::derivative (math.rs:640-643): q ⊗ (ω, 0) / 2.
template <class R> NEWT_HD Qt<R> qderiv(Qt<R> q, V3<R> w) {
    const Qt<R> p = qmul(q, Qt<R>{w.x, w.y, w.z, R(0.0)});
    return {p.x * R(0.5), p.y * R(0.5), p.z * R(0.5), p.w * R(0.5)};
}
template <class R> NEWT_HD Qt<R> qintegrate(Qt<R> q, V3<R> w, R dt) {
    return qnormalize(qmul(q, qaxis(w, dt * len3(w))));
}

// 3x3 in row-major m[r*3+c]; built from a quaternion with Quat::to_mat3's
// entry formulas (newt/src/math.rs:620-633).
template <class R> struct M33 {
    R m[9];
};
template <class R> NEWT_HD M33<R> qmat(Qt<R> q) {
    R x = q.x, y = q.y, z = q.z, w = q.w;
    M33<R> o;
    o.m[0] = R(1.0) - R(2.0) * (y * y + z * z);
    o.m[3] = R(2.0) * (x * y + z * w);
    o.m[6] = R(2.0) * (x * z - y * w);
    o.m[1] = R(2.0) * (x * y - z * w);
    o.m[4] = R(1.0) - R(2.0) * (x * x + z * z);
    o.m[7] = R(2.0) * (y * z + x * w);
    o.m[2] = R(2.0) * (x * z + y * w);
    o.m[5] = R(2.0) * (y * z - x * w);
    o.m[8] = R(1.0) - R(2.0) * (x * x + y * y);
    return o;
}
// A row-major table as a Mat3; Mat3 product and transpose (math.rs:523-537),
// sums from 0.0.
NEWT_HD inline M33<double> m33_load(const double* p) {
    M33<double> a;
    for (int k = 0; k < 9; ++k) a.m[k] = p[k];
    return a;
}
NEWT_HD inline M33<double> m33_mul(const M33<double>& a, const M33<double>& b) {
    M33<double> o;
    for (int c = 0; c < 3; ++c)
        for (int r = 0; r < 3; ++r) {
            double s = 0.0;
            for (int k = 0; k < 3; ++k) s += a.m[r * 3 + k] * b.m[k * 3 + c];
            o.m[r * 3 + c] = s;
        }
    return o;
}
NEWT_HD inline M33<double> m33_tr(const M33<double>& a) {
    M33<double> o;
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) o.m[r * 3 + c] = a.m[c * 3 + r];
    return o;
}
template <class R> NEWT_HD V3<R> mv(const M33<R>& a, V3<R> v) {
    return {a.m[0] * v.x + a.m[1] * v.y + a.m[2] * v.z, a.m[3] * v.x + a.m[4] * v.y + a.m[5] * v.z,
            a.m[6] * v.x + a.m[7] * v.y + a.m[8] * v.z};
}
template <class R> NEWT_HD V3<R> mtv(const M33<R>& a, V3<R> v) {
    return {a.m[0] * v.x + a.m[3] * v.y + a.m[6] * v.z, a.m[1] * v.x + a.m[4] * v.y + a.m[7] * v.z,
            a.m[2] * v.x + a.m[5] * v.y + a.m[8] * v.z};
}

// Spatial motion (angular, linear) and force (torque, linear).
template <class R> struct SM {
    V3<R> w, v;
};
template <class R> struct SF {
    V3<R> t, f;
};
template <class R> NEWT_HD SM<R> smz() { return {z3<R>(), z3<R>()}; }
template <class R> NEWT_HD SF<R> sfz() { return {z3<R>(), z3<R>()}; }
template <class R> NEWT_HD SM<R> operator+(SM<R> a, SM<R> b) { return {a.w + b.w, a.v + b.v}; }
template <class R> NEWT_HD SM<R> operator*(SM<R> a, R s) { return {a.w * s, a.v * s}; }
template <class R> NEWT_HD SF<R> operator+(SF<R> a, SF<R> b) { return {a.t + b.t, a.f + b.f}; }
template <class R> NEWT_HD SF<R> operator-(SF<R> a, SF<R> b) { return {a.t - b.t, a.f - b.f}; }
template <class R> NEWT_HD SF<R> operator*(SF<R> a, R s) { return {a.t * s, a.f * s}; }
template <class R> NEWT_HD SM<R> crm(SM<R> a, SM<R> b) { return {cross3(a.w, b.w), cross3(a.w, b.v) + cross3(a.v, b.w)}; }
template <class R> NEWT_HD SF<R> crf(SM<R> a, SF<R> b) { return {cross3(a.w, b.t) + cross3(a.v, b.f), cross3(a.w, b.f)}; }
template <class R> NEWT_HD R sdot(SM<R> m, SF<R> f) { return dot3(m.w, f.t) + dot3(m.v, f.f); }

// Plücker transform parent -> child: rotation (row-major) and translation.
template <class R> struct Xf {
    M33<R> E;
    V3<R> r;
};
template <class R> NEWT_HD SM<R> xmotion(const Xf<R>& x, SM<R> m) {
    V3<R> w = mv(x.E, m.w);
    return {w, mv(x.E, m.v) + cross3(x.r, w)};
}
template <class R> NEWT_HD SF<R> xtforce(const Xf<R>& x, SF<R> f) {
    V3<R> fl = mtv(x.E, f.f);
    V3<R> tq = mtv(x.E, f.t - cross3(x.r, f.f));
    return {tq, fl};
}

// 6x6 (row-major) times motion; accumulation from +0 like newt's loops.
template <class R> NEWT_HD SF<R> m6mul(const R* a, SM<R> m) {
    R v[6] = {m.w.x, m.w.y, m.w.z, m.v.x, m.v.y, m.v.z};
    R o[6];
#pragma unroll
    for (int i = 0; i < 6; ++i) {
        R s = R(0);
#pragma unroll
        for (int j = 0; j < 6; ++j) s = s + a[i * 6 + j] * v[j];
        o[i] = s;
    }
    return {{o[0], o[1], o[2]}, {o[3], o[4], o[5]}};
}

// Xᵀ A X by basis columns (newt/src/spatial.rs:457-481, aba.rs:551-613).
template <class R> NEWT_HD void m6pullback(const R* a, const Xf<R>& x, R* out) {
#pragma unroll
    for (int c = 0; c < 6; ++c) {
        SM<R> e = smz<R>();
        R* comp[6] = {&e.w.x, &e.w.y, &e.w.z, &e.v.x, &e.v.y, &e.v.z};
        *comp[c] = R(1);
        SF<R> f = xtforce(x, m6mul(a, xmotion(x, e)));
        out[0 * 6 + c] = f.t.x;
        out[1 * 6 + c] = f.t.y;
        out[2 * 6 + c] = f.t.z;
        out[3 * 6 + c] = f.f.x;
        out[4 * 6 + c] = f.f.y;
        out[5 * 6 + c] = f.f.z;
    }
}

// clamp01 (contact.rs:2324-2334): NaN passes through.
NEWT_HD inline double clamp01(double x) { return x < 0.0 ? 0.0 : (x > 1.0 ? 1.0 : x); }
// closest_point_on_segment (contact.rs:2313-2322).
NEWT_HD inline V3<double> closest_on_segment(V3<double> p, V3<double> a, V3<double> b) {
    const V3<double> ab = b - a;
    const double denom = dot3(ab, ab);
    if (denom == 0.0) return a;
    return a + ab * clamp01(dot3(p - a, ab) / denom);
}
// closest_points_on_segments (contact.rs:2342-2385): points on [a0,a1] and
// [b0,b1], clamp-and-recheck with newt's 1e-12 degeneracy threshold.
NEWT_HD inline void closest_on_segments(V3<double> a0, V3<double> a1, V3<double> b0, V3<double> b1, V3<double>* pa,
                                        V3<double>* pb) {
    const V3<double> d1 = a1 - a0, d2 = b1 - b0, r = a0 - b0;
    const double a = dot3(d1, d1), e = dot3(d2, d2), f = dot3(d2, r);
    const double eps = 1.0e-12;
    if (a <= eps && e <= eps) {
        *pa = a0;
        *pb = b0;
        return;
    }
    if (a <= eps) {
        *pa = a0;
        *pb = b0 + d2 * clamp01(f / e);
        return;
    }
    const double c = dot3(d1, r);
    if (e <= eps) {
        *pa = a0 + d1 * clamp01(-c / a);
        *pb = b0;
        return;
    }
    const double bb = dot3(d1, d2);
    const double denom = a * e - bb * bb;
    const double s = denom != 0.0 ? clamp01((bb * f - c * e) / denom) : 0.0;
    const double tn = bb * s + f;
    double t, sf;
    if (tn < 0.0) {
        t = 0.0;
        sf = clamp01(-c / a);
    } else if (tn > e) {
        t = 1.0;
        sf = clamp01((bb - c) / a);
    } else {
        t = tn / e;
        sf = s;
    }
    *pa = a0 + d1 * sf;
    *pb = b0 + d2 * t;
}
// Box-box contact manifold: box_box_full_manifold -> box_box_sat_fallback
// (contact.rs:569-583, 716-868), edge_midpoint (1173-1196),
// box_box_edge_edge_contact (872-910), box_box_face_reference_contacts
// (917-1079) with sutherland_hodgman_axis_rect (1115-1167).  15-axis SAT with
// the minimum overlap winning; an edge-edge winner gives one contact at the
// edges' closest points, a face winner clips the most anti-parallel face of
// the other box against the reference face and keeps the four deepest
// points (stable order).  Writes up to four points on b's side, their
// penetrations and the shared normal b -> a; returns the count.
struct ClipV {
    double u, v, depth;
};
NEWT_HD inline void other_two(int k, int* a, int* b) {
    *a = k == 0 ? 1 : 0;
    *b = k == 2 ? 1 : 2;
}
NEWT_HD inline V3<double> edge_midpoint(V3<double> c, const V3<double>* basis, const double* half, int ei,
                                        V3<double> dir) {
    int j, k;
    other_two(ei, &j, &k);
    const double sj = dot3(dir, basis[j]) >= 0.0 ? 1.0 : -1.0;
    const double sk = dot3(dir, basis[k]) >= 0.0 ? 1.0 : -1.0;
    return (c + basis[j] * (sj * half[j])) + basis[k] * (sk * half[k]);
}
NEWT_HD inline int box_box(V3<double> pa, Qt<double> qa, const double* ha, V3<double> pb, Qt<double> qb,
                           const double* hb, double margin, V3<double>* cpos_out, double* pen_out,
                           V3<double>* normal) {
    const V3<double> ex = mk3(1.0, 0.0, 0.0), ey = mk3(0.0, 1.0, 0.0), ez = mk3(0.0, 0.0, 1.0);
    const V3<double> ax[3] = {qrot(qa, ex), qrot(qa, ey), qrot(qa, ez)};
    const V3<double> bx[3] = {qrot(qb, ex), qrot(qb, ey), qrot(qb, ez)};
    const V3<double> delta = pa - pb;
    auto proj = [](V3<double> l, const V3<double>* u, const double* h) {
        return (h[0] * nabs(dot3(l, u[0])) + h[1] * nabs(dot3(l, u[1]))) + h[2] * nabs(dot3(l, u[2]));
    };
    double best = __builtin_inf();
    V3<double> best_axis = ez;
    int kind = 0, k1 = 0, k2 = 0;  // 0 face of a, 1 face of b, 2 edge-edge
    auto test = [&](V3<double> l, int knd, int i, int j) {
        const double sgn = dot3(delta, l);
        const double ov = (proj(l, ax, ha) + proj(l, bx, hb)) - nabs(sgn);
        if (ov < -margin) return true;
        if (ov < best) {
            best = ov;
            best_axis = sgn >= 0.0 ? l : -l;
            kind = knd;
            k1 = i;
            k2 = j;
        }
        return false;
    };
    for (int k = 0; k < 3; ++k)
        if (test(ax[k], 0, k, 0)) return 0;
    for (int k = 0; k < 3; ++k)
        if (test(bx[k], 1, k, 0)) return 0;
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            const V3<double> raw = cross3(ax[i], bx[j]);
            const double len2 = dot3(raw, raw);
            if (len2 < 1.0e-6) continue;
            if (test(raw / nsqrt(len2), 2, i, j)) return 0;
        }
    if (!__builtin_isfinite(best)) return 0;
    const double pen_shift = best + margin;
    if (pen_shift <= 0.0) return 0;
    const V3<double> nw = best_axis;
    *normal = nw;
    if (kind == 2) {
        const V3<double> ma = edge_midpoint(pa, ax, ha, k1, -nw);
        const V3<double> mb = edge_midpoint(pb, bx, hb, k2, nw);
        V3<double> sa, sb;
        closest_on_segments(ma - ax[k1] * ha[k1], ma + ax[k1] * ha[k1], mb - bx[k2] * hb[k2], mb + bx[k2] * hb[k2],
                            &sa, &sb);
        cpos_out[0] = sb;
        pen_out[0] = pen_shift;
        return 1;
    }
    const bool ref_a = kind == 0;
    const int rk = k1;
    const V3<double> rc = ref_a ? pa : pb;
    const V3<double>* rb = ref_a ? ax : bx;
    const double* rh = ref_a ? ha : hb;
    const V3<double> ic = ref_a ? pb : pa;
    const V3<double>* ibs = ref_a ? bx : ax;
    const double* ih = ref_a ? hb : ha;
    const V3<double> ron = ref_a ? -nw : nw;
    const V3<double> rfa = rb[rk];
    const double rs = dot3(ron, rfa) >= 0.0 ? 1.0 : -1.0;
    const V3<double> rfc = rc + rfa * (rs * rh[rk]);
    int ru, rv;
    other_two(rk, &ru, &rv);
    int iax = 0;
    double isg = 1.0, idot = __builtin_inf();
    for (int k = 0; k < 3; ++k)
        for (int e = 0; e < 2; ++e) {
            const double sg = e == 0 ? 1.0 : -1.0;
            const double d = dot3(ibs[k] * sg, ron);
            if (d < idot) {
                idot = d;
                iax = k;
                isg = sg;
            }
        }
    const V3<double> ifc = ic + ibs[iax] * (isg * ih[iax]);
    int iu, iv;
    other_two(iax, &iu, &iv);
    const V3<double> U = ibs[iu], Vv = ibs[iv];
    const double hu = ih[iu], hv = ih[iv];
    const V3<double> corners[4] = {(ifc - U * hu) - Vv * hv, (ifc + U * hu) - Vv * hv, (ifc + U * hu) + Vv * hv,
                                   (ifc - U * hu) + Vv * hv};
    ClipV poly[8], tmp[8];
    int np = 4;
    for (int i = 0; i < 4; ++i) {
        const V3<double> rel = corners[i] - rfc;
        poly[i] = ClipV{dot3(rel, rb[ru]), dot3(rel, rb[rv]), -dot3(rel, ron)};
    }
    const double ru_h = rh[ru], rv_h = rh[rv];
    for (int edge = 0; edge < 4; ++edge) {
        if (np == 0) break;
        for (int i = 0; i < np; ++i) tmp[i] = poly[i];
        const int n = np;
        np = 0;
        auto side = [&](const ClipV& p) {
            return edge == 0 ? p.u + ru_h : (edge == 1 ? ru_h - p.u : (edge == 2 ? p.v + rv_h : rv_h - p.v));
        };
        for (int i = 0; i < n; ++i) {
            const ClipV cur = tmp[i], prev = tmp[(i + n - 1) % n];
            const double cs = side(cur), ps = side(prev);
            const bool ci = cs >= 0.0, pi = ps >= 0.0;
            if (ci != pi) {  // the edge crosses the clip line
                const double t = ps / (ps - cs);
                poly[np++] = ClipV{prev.u + (cur.u - prev.u) * t, prev.v + (cur.v - prev.v) * t,
                                   prev.depth + (cur.depth - prev.depth) * t};
            }
            if (ci) poly[np++] = cur;
        }
    }
    double cpen[8];
    V3<double> cpos[8];
    int count = 0;
    for (int i = 0; i < np; ++i) {
        const double pen = poly[i].depth + margin;
        if (pen <= 0.0) continue;
        const V3<double> por = (rfc + rb[ru] * poly[i].u) + rb[rv] * poly[i].v;
        cpen[count] = pen;
        cpos[count] = ref_a ? por + nw * (pen - margin) : por;
        if (++count == 8) break;
    }
    int order[8] = {0, 1, 2, 3, 4, 5, 6, 7};
    for (int i = 1; i < count; ++i)
        for (int j = i; j > 0 && cpen[order[j]] > cpen[order[j - 1]]; --j) {
            const int t = order[j - 1];
            order[j - 1] = order[j];
            order[j] = t;
        }
    const int take = count > 4 ? 4 : count;
    for (int i = 0; i < take; ++i) {
        cpos_out[i] = cpos[order[i]];
        pen_out[i] = cpen[order[i]];
    }
    return take;
}

// Sphere and capsule pairs, cap_* telling capsules from spheres
// (sphere_sphere contact.rs:407-427, sphere_capsule
// 473-499, capsule_capsule 1249-1272), reduced to sphere_vs_sphere_at
// (434-467).  The pair is (a, b) in candidate order; a capsule-sphere pair
// runs as sphere-capsule and is relabelled with the normal negated
// (dispatch_narrow_phase, 2579-2588).  Returns false when the spheres do not
// touch; else the contact point on b's side, the normal b -> a and the
// penetration.
NEWT_HD inline bool round_pair(bool cap_a, V3<double> pa, Qt<double> qa, const double* size_a, bool cap_b,
                               V3<double> pb, Qt<double> qb, const double* size_b, double margin, V3<double>* point,
                               V3<double>* normal, double* pen) {
    const bool swap = cap_a && !cap_b;
    const bool cx_cap = swap ? cap_b : cap_a, cy_cap = swap ? cap_a : cap_b;
    const V3<double> px = swap ? pb : pa, py = swap ? pa : pb;
    const Qt<double> qx = swap ? qb : qa, qy = swap ? qa : qb;
    const double* zx = swap ? size_b : size_a;
    const double* zy = swap ? size_a : size_b;
    V3<double> cx = px, cy = py;
    if (!cx_cap && cy_cap) {
        const V3<double> ay = qrot(qy, mk3(0.0, 0.0, 1.0));
        cy = closest_on_segment(px, py - ay * zy[1], py + ay * zy[1]);
    } else if (cx_cap && cy_cap) {
        const V3<double> ax = qrot(qx, mk3(0.0, 0.0, 1.0)), ay = qrot(qy, mk3(0.0, 0.0, 1.0));
        closest_on_segments(px - ax * zx[1], px + ax * zx[1], py - ay * zy[1], py + ay * zy[1], &cx, &cy);
    }
    const V3<double> delta = cx - cy;
    const double dist2 = dot3(delta, delta);
    const double srm = (zx[0] + zy[0]) + margin;
    if (dist2 >= srm * srm) return false;
    const double dist = nsqrt(dist2);
    const V3<double> n = dist > 0.0 ? delta / dist : mk3(0.0, 0.0, 1.0);
    *point = cy + n * zy[0];
    *normal = swap ? -n : n;
    *pen = srm - dist;
    return true;
}

// SolImp with MuJoCo's clamps (effective_solimp, solver.rs:319-328).
struct SolImpE {
    double dmin, dmax, width, mid;
    int power;
};
NEWT_HD inline SolImpE solimp_effective(const double* si, int power) {
    const double lo = 1.0e-4, hi = 0.9999;
    auto cl = [&](double v) { return v < lo ? lo : (v > hi ? hi : v); };
    return SolImpE{cl(si[0]), cl(si[1]), si[2] > 0.0 ? si[2] : 0.0, cl(si[3]), power < 1 ? 1 : power};
}
// impedance_at_position (solver.rs:261-309) at zero margin.
NEWT_HD inline double impedance_at(const SolImpE& s, double position) {
    double x = position / s.width;
    if (x < 0.0) x = -x;
    double y;
    if (x <= 0.0) {
        y = 0.0;
    } else if (x >= 1.0) {
        y = 1.0;
    } else if (x < s.mid) {
        double b = x / s.mid, pw = 1.0;
        for (int k = 0; k < s.power; ++k) pw = pw * b;
        y = s.mid * pw;
    } else {
        double b = (1.0 - x) / (1.0 - s.mid), pw = 1.0;
        for (int k = 0; k < s.power; ++k) pw = pw * b;
        y = 1.0 - (1.0 - s.mid) * pw;
    }
    return s.dmin + (s.dmax - s.dmin) * y;
}
// reference_accel with safe_solref (solver.rs:331-384).
NEWT_HD inline double reference_accel(double position, double velocity, double tc, double dr, const SolImpE& s,
                                      double dt) {
    const bool direct = tc < 0.0;
    if (!direct && tc < 2.0 * dt) tc = 2.0 * dt;
    double b, k;
    if (direct) {
        b = -dr / s.dmax;
        k = -tc / (s.dmax * s.dmax);
    } else {
        const double d = impedance_at(s, position);
        b = 2.0 / (s.dmax * tc);
        k = d / (s.dmax * s.dmax * tc * tc * dr * dr);
    }
    return -b * velocity - k * position;
}

}  // namespace newt
