// newt_ref.hpp — CPU oracle: a scalar-templated C++ restatement of newt's
// `World::step` for articulated trees (newt/src/world.rs:1459-1537).
//
// TEST INFRASTRUCTURE ONLY.  This file is the checker for the HIP path; it is
// compiled into oracle/_build/liboracle.so and loaded only by tests/,
// __graft_entry__.smoke() and the cpu_baseline leg of bench.py.  Nothing in
// tooling_amd/ links or calls it.
//
// Instantiations:
//   float  — newt's own arithmetic.  Built with -ffp-contract=off and SSE
//            math it reproduces newt's golden trajectories byte for byte
//            (tests/golden/*.bin, pinned by tests/test_oracle_golden.py).
//   double — the FP64 oracle: the same operations, in the same order, on
//            doubles.  The GPU FP64 path is checked against it.
//
// Every function cites the reference function it restates.  Operation order
// follows the reference expression trees, including Rust's `Sum` for floats
// (fold from -0.0) and `f32::max/min` (NaN-ignoring, like fmax/fmin).
#pragma once

#include <cmath>
#include <cstdint>
#include <cstring>
#include <cstdlib>
#include <cstdio>
#include <functional>
#include <vector>
#include <set>
#include <limits>
#include <utility>
#include <algorithm>
#include <array>
#include <memory>

namespace oracle {

[[noreturn]] inline void fail(const char* msg) {
    std::fprintf(stderr, "oracle: %s\n", msg);
    std::abort();
}

// Rust `f32::max` / `f32::min` (IEEE maxNum/minNum: a NaN operand is ignored).
template <class T> inline T rmax(T a, T b) { return std::fmax(a, b); }
template <class T> inline T rmin(T a, T b) { return std::fmin(a, b); }
// Rust `f32::clamp` (newt/src/solver.rs:317 etc.): NaN passes through.
template <class T> inline T rclamp(T v, T lo, T hi) {
    if (v < lo) v = lo;
    if (v > hi) v = hi;
    return v;
}
// Rust `f32::abs` (clears the sign bit).
template <class T> inline T rabs(T x) { return std::fabs(x); }

// Rust `x as i32` for floats: truncation toward zero, saturating, NaN -> 0.
template <class T> inline int32_t as_i32(T x) {
    if (x != x) return 0;
    if (x >= T(2147483647.0)) return 2147483647;
    if (x <= T(-2147483648.0)) return (int32_t)(-2147483647 - 1);
    return (int32_t)x;
}

// ---------------------------------------------------------------------------
// deterministic sin / cos (newt/src/math.rs:45-134)
// ---------------------------------------------------------------------------

// sin_poly: degree-7 Taylor in Horner form (math.rs:84-93).
template <class T> inline T sin_poly(T x) {
    T x2 = x * x;
    T a7 = T(-1.0) / T(5040.0);
    T a5 = T(1.0) / T(120.0);
    T a3 = T(-1.0) / T(6.0);
    return x * (T(1.0) + x2 * (a3 + x2 * (a5 + x2 * a7)));
}
// cos_poly: degree-8 Taylor in Horner form (math.rs:96-104).
template <class T> inline T cos_poly(T x) {
    T x2 = x * x;
    T b8 = T(1.0) / T(40320.0);
    T b6 = T(-1.0) / T(720.0);
    T b4 = T(1.0) / T(24.0);
    T b2 = T(-1.0) / T(2.0);
    return T(1.0) + x2 * (b2 + x2 * (b4 + x2 * (b6 + x2 * b8)));
}
// reduce_pi_over_2: Cody-Waite three-part split, constants as newt writes
// them (math.rs:113-134).  The literals are typed as T.
template <class T> inline T reduce_pi_over_2(T x, int32_t* quadrant) {
    const T TWO_OVER_PI = T(0.63661974);
    const T C1 = T(1.5703125);
    const T C2 = T(4.8370362e-4);
    const T C3 = T(7.5497895e-8);
    T scaled = x * TWO_OVER_PI;
    int32_t q = scaled >= T(0.0) ? as_i32(scaled + T(0.5)) : as_i32(scaled - T(0.5));
    T k = (T)q;
    *quadrant = q;
    return ((x - k * C1) - k * C2) - k * C3;
}
template <class T> inline T nsin(T x) {  // math.rs:45-53
    int32_t q;
    T r = reduce_pi_over_2(x, &q);
    switch (q & 3) {
        case 0: return sin_poly(r);
        case 1: return cos_poly(r);
        case 2: return -sin_poly(r);
        default: return -cos_poly(r);
    }
}
template <class T> inline T ncos(T x) {  // math.rs:56-64
    int32_t q;
    T r = reduce_pi_over_2(x, &q);
    switch (q & 3) {
        case 0: return cos_poly(r);
        case 1: return -sin_poly(r);
        case 2: return -cos_poly(r);
        default: return sin_poly(r);
    }
}

// ---------------------------------------------------------------------------
// Vec3 / Mat3 / Quat (newt/src/math.rs:258-683)
// ---------------------------------------------------------------------------

template <class T> struct V3 {
    T x, y, z;
};
template <class T> inline V3<T> v3(T x, T y, T z) { return V3<T>{x, y, z}; }
template <class T> inline V3<T> vzero() { return V3<T>{T(0), T(0), T(0)}; }
template <class T> inline V3<T> operator+(V3<T> a, V3<T> b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
template <class T> inline V3<T> operator-(V3<T> a, V3<T> b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
template <class T> inline V3<T> operator-(V3<T> a) { return {-a.x, -a.y, -a.z}; }
template <class T> inline V3<T> operator*(V3<T> a, T s) { return {a.x * s, a.y * s, a.z * s}; }
template <class T> inline V3<T> operator/(V3<T> a, T s) { return {a.x / s, a.y / s, a.z / s}; }
template <class T> inline T dot(V3<T> a, V3<T> b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
template <class T> inline V3<T> cross(V3<T> a, V3<T> b) {
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
template <class T> inline T length(V3<T> a) { return std::sqrt(dot(a, a)); }
template <class T> inline V3<T> normalize(V3<T> a) {  // math.rs:316-319
    T l = length(a);
    return l == T(0) ? vzero<T>() : a / l;
}
template <class T> inline bool is_zero_vec(V3<T> a) {  // `v == Vec3::ZERO` (PartialEq)
    return a.x == T(0) && a.y == T(0) && a.z == T(0);
}
template <class T> inline V3<T> vabs(V3<T> a) { return {rabs(a.x), rabs(a.y), rabs(a.z)}; }

// Mat3: column-major, d[c*3+r] (math.rs:370-386).
template <class T> struct M3 {
    T d[9];
    T get(int r, int c) const { return d[c * 3 + r]; }
};
template <class T> inline M3<T> m3diag(T x, T y, T z) {
    return M3<T>{{x, T(0), T(0), T(0), y, T(0), T(0), T(0), z}};
}
template <class T> inline M3<T> m3identity() { return m3diag(T(1), T(1), T(1)); }
template <class T> inline V3<T> operator*(const M3<T>& m, V3<T> r) {  // math.rs:512-520
    return {m.d[0] * r.x + m.d[3] * r.y + m.d[6] * r.z,
            m.d[1] * r.x + m.d[4] * r.y + m.d[7] * r.z,
            m.d[2] * r.x + m.d[5] * r.y + m.d[8] * r.z};
}
template <class T> inline M3<T> operator*(const M3<T>& a, const M3<T>& b) {  // math.rs:523-537
    M3<T> o;
    for (int c = 0; c < 3; ++c)
        for (int r = 0; r < 3; ++r) {
            T s = T(0.0);
            for (int k = 0; k < 3; ++k) s += a.get(r, k) * b.get(k, c);
            o.d[c * 3 + r] = s;
        }
    return o;
}
template <class T> inline M3<T> operator*(const M3<T>& a, T s) {
    M3<T> o;
    for (int i = 0; i < 9; ++i) o.d[i] = a.d[i] * s;
    return o;
}
template <class T> inline M3<T> operator+(const M3<T>& a, const M3<T>& b) {
    M3<T> o;
    for (int i = 0; i < 9; ++i) o.d[i] = a.d[i] + b.d[i];
    return o;
}
template <class T> inline M3<T> transpose(const M3<T>& m) {
    return M3<T>{{m.get(0, 0), m.get(0, 1), m.get(0, 2), m.get(1, 0), m.get(1, 1), m.get(1, 2),
                  m.get(2, 0), m.get(2, 1), m.get(2, 2)}};
}
template <class T> inline M3<T> skew(V3<T> v) {  // math.rs:418-424
    return M3<T>{{T(0.0), v.z, -v.y, -v.z, T(0.0), v.x, v.y, -v.x, T(0.0)}};
}
template <class T> inline bool m3inverse(const M3<T>& m, M3<T>* out) {  // math.rs:427-453
    T a = m.get(0, 0), b = m.get(0, 1), c = m.get(0, 2);
    T d = m.get(1, 0), e = m.get(1, 1), f = m.get(1, 2);
    T g = m.get(2, 0), h = m.get(2, 1), i = m.get(2, 2);
    T det = a * (e * i - f * h) - b * (d * i - f * g) + c * (d * h - e * g);
    if (det == T(0) || !std::isfinite(det)) return false;
    T inv = T(1.0) / det;
    *out = M3<T>{{(e * i - f * h) * inv, (f * g - d * i) * inv, (d * h - e * g) * inv,
                  (c * h - b * i) * inv, (a * i - c * g) * inv, (b * g - a * h) * inv,
                  (b * f - c * e) * inv, (c * d - a * f) * inv, (a * e - b * d) * inv}};
    return true;
}

template <class T> struct Q {
    T x, y, z, w;
};
template <class T> inline Q<T> qidentity() { return Q<T>{T(0), T(0), T(0), T(1)}; }
template <class T> inline Q<T> operator*(Q<T> a, Q<T> r) {  // math.rs:659-669
    return {a.w * r.x + a.x * r.w + a.y * r.z - a.z * r.y,
            a.w * r.y - a.x * r.z + a.y * r.w + a.z * r.x,
            a.w * r.z + a.x * r.y - a.y * r.x + a.z * r.w,
            a.w * r.w - a.x * r.x - a.y * r.y - a.z * r.z};
}
template <class T> inline Q<T> qscale(Q<T> a, T s) { return {a.x * s, a.y * s, a.z * s, a.w * s}; }
template <class T> inline Q<T> conj(Q<T> q) { return {-q.x, -q.y, -q.z, q.w}; }
template <class T> inline Q<T> from_axis_angle(V3<T> axis, T angle) {  // math.rs:574-580
    V3<T> a = normalize(axis);
    T half = angle * T(0.5);
    T s = nsin(half);
    T c = ncos(half);
    return {a.x * s, a.y * s, a.z * s, c};
}
template <class T> inline Q<T> renormalize(Q<T> q) {  // math.rs:590-597
    T n = std::sqrt(q.x * q.x + q.y * q.y + q.z * q.z + q.w * q.w);
    if (n == T(0)) return qidentity<T>();
    return {q.x / n, q.y / n, q.z / n, q.w / n};
}
template <class T> inline V3<T> rotate(Q<T> q, V3<T> v) {  // math.rs:605-612
    V3<T> qv{q.x, q.y, q.z};
    V3<T> t = cross(qv, v) * T(2.0);
    return v + t * q.w + cross(qv, t);
}
template <class T> inline V3<T> inverse_rotate(Q<T> q, V3<T> v) { return rotate(conj(q), v); }
template <class T> inline M3<T> to_mat3(Q<T> q) {  // math.rs:620-633
    T x = q.x, y = q.y, z = q.z, w = q.w;
    return M3<T>{{T(1.0) - T(2.0) * (y * y + z * z), T(2.0) * (x * y + z * w),
                  T(2.0) * (x * z - y * w), T(2.0) * (x * y - z * w),
                  T(1.0) - T(2.0) * (x * x + z * z), T(2.0) * (y * z + x * w),
                  T(2.0) * (x * z + y * w), T(2.0) * (y * z - x * w),
                  T(1.0) - T(2.0) * (x * x + y * y)}};
}
template <class T> inline Q<T> qderivative(Q<T> q, V3<T> w) {  // math.rs:640-643
    Q<T> o{w.x, w.y, w.z, T(0.0)};
    return qscale(q * o, T(0.5));
}
template <class T> inline Q<T> integrate_body_angular_velocity(Q<T> q, V3<T> w, T dt) {  // math.rs:648-650
    return renormalize(q * from_axis_angle(w, dt * length(w)));
}

// Rust `Iterator::sum::<f32>()`: fold from -0.0.
template <class T> struct RSum {
    T acc = T(-0.0);
    void add(T v) { acc = acc + v; }
};

// ---------------------------------------------------------------------------
// Generic spatial algebra used by ABA (newt/src/aba.rs:193-665).  GMat3 is
// row-major data[row][col]; GMat6 rows[row][col].
// ---------------------------------------------------------------------------

template <class T> struct Motion {
    V3<T> ang, lin;
};
template <class T> struct Force {
    V3<T> tq, lin;
};
template <class T> inline Motion<T> mzero() { return {vzero<T>(), vzero<T>()}; }
template <class T> inline Force<T> fzero() { return {vzero<T>(), vzero<T>()}; }
template <class T> inline Motion<T> madd(Motion<T> a, Motion<T> b) { return {a.ang + b.ang, a.lin + b.lin}; }
template <class T> inline Motion<T> mscale(Motion<T> a, T s) { return {a.ang * s, a.lin * s}; }
template <class T> inline Force<T> fadd(Force<T> a, Force<T> b) { return {a.tq + b.tq, a.lin + b.lin}; }
template <class T> inline Force<T> fsub(Force<T> a, Force<T> b) { return {a.tq - b.tq, a.lin - b.lin}; }
template <class T> inline Force<T> fscale(Force<T> a, T s) { return {a.tq * s, a.lin * s}; }
template <class T> inline Motion<T> cross_motion(Motion<T> a, Motion<T> b) {  // aba.rs:273-281
    return {cross(a.ang, b.ang), cross(a.ang, b.lin) + cross(a.lin, b.ang)};
}
template <class T> inline Force<T> cross_force(Motion<T> a, Force<T> f) {  // aba.rs:282-290
    return {cross(a.ang, f.tq) + cross(a.lin, f.lin), cross(a.ang, f.lin)};
}
template <class T> inline T dot_force(Motion<T> m, Force<T> f) {  // aba.rs:1104-1108
    return dot(m.ang, f.tq) + dot(m.lin, f.lin);
}

template <class T> struct G3 {  // GMat3, row-major
    T d[3][3];
};
template <class T> inline G3<T> g3zero() {
    G3<T> o;
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) o.d[r][c] = T(0);
    return o;
}
template <class T> inline G3<T> g3identity() {
    G3<T> o = g3zero<T>();
    for (int i = 0; i < 3; ++i) o.d[i][i] = T(1);
    return o;
}
template <class T> inline G3<T> g3from(const M3<T>& m) {
    G3<T> o;
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) o.d[r][c] = m.get(r, c);
    return o;
}
template <class T> inline V3<T> g3mul(const G3<T>& m, V3<T> v) {  // aba.rs:78-84
    return {m.d[0][0] * v.x + m.d[0][1] * v.y + m.d[0][2] * v.z,
            m.d[1][0] * v.x + m.d[1][1] * v.y + m.d[1][2] * v.z,
            m.d[2][0] * v.x + m.d[2][1] * v.y + m.d[2][2] * v.z};
}
template <class T> inline V3<T> g3tmul(const G3<T>& m, V3<T> v) {  // aba.rs:85-91
    return {m.d[0][0] * v.x + m.d[1][0] * v.y + m.d[2][0] * v.z,
            m.d[0][1] * v.x + m.d[1][1] * v.y + m.d[2][1] * v.z,
            m.d[0][2] * v.x + m.d[1][2] * v.y + m.d[2][2] * v.z};
}
template <class T> inline G3<T> g3mm(const G3<T>& a, const G3<T>& b) {  // aba.rs:383-394
    G3<T> o = g3zero<T>();
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c)
            for (int k = 0; k < 3; ++k) o.d[r][c] = o.d[r][c] + a.d[r][k] * b.d[k][c];
    return o;
}
template <class T> inline G3<T> g3inverse(const G3<T>& m) {  // aba.rs:395-425
    const auto& a = m.d;
    T d = a[0][0] * a[1][1] - a[0][1] * a[1][0];
    T det = a[0][0] * (a[1][1] * a[2][2] - a[1][2] * a[2][1]) -
            a[0][1] * (a[1][0] * a[2][2] - a[1][2] * a[2][0]) +
            a[0][2] * (a[1][0] * a[2][1] - a[1][1] * a[2][0]);
    T inv = T(1) / det;
    G3<T> o;
    o.d[0][0] = (a[1][1] * a[2][2] - a[1][2] * a[2][1]) * inv;
    o.d[0][1] = (a[0][2] * a[2][1] - a[0][1] * a[2][2]) * inv;
    o.d[0][2] = (a[0][1] * a[1][2] - a[0][2] * a[1][1]) * inv;
    o.d[1][0] = (a[1][2] * a[2][0] - a[1][0] * a[2][2]) * inv;
    o.d[1][1] = (a[0][0] * a[2][2] - a[0][2] * a[2][0]) * inv;
    o.d[1][2] = (a[0][2] * a[1][0] - a[0][0] * a[1][2]) * inv;
    o.d[2][0] = (a[1][0] * a[2][1] - a[1][1] * a[2][0]) * inv;
    o.d[2][1] = (a[0][1] * a[2][0] - a[0][0] * a[2][1]) * inv;
    o.d[2][2] = d * inv;
    return o;
}
// quaternion_matrix for f32 goes through Quat::to_mat3 (aba.rs:69-76).
template <class T> inline G3<T> quaternion_matrix(T x, T y, T z, T w) {
    return g3from(to_mat3(Q<T>{x, y, z, w}));
}

template <class T> struct GX {  // GXform (aba.rs:430-458)
    G3<T> rot;
    V3<T> tr;
};
template <class T> inline GX<T> gxidentity() { return {g3identity<T>(), vzero<T>()}; }
template <class T> inline Motion<T> gx_motion(const GX<T>& x, Motion<T> m) {
    V3<T> ang = g3mul(x.rot, m.ang);
    return {ang, g3mul(x.rot, m.lin) + cross(x.tr, ang)};
}
template <class T> inline Force<T> gx_tforce(const GX<T>& x, Force<T> f) {
    V3<T> lin = g3tmul(x.rot, f.lin);
    V3<T> tq = g3tmul(x.rot, f.tq - cross(x.tr, f.lin));
    return {tq, lin};
}

template <class T> struct G6 {  // GMat6 (aba.rs:461-665)
    T r[6][6];
};
template <class T> inline G6<T> g6zero() {
    G6<T> o;
    for (int i = 0; i < 6; ++i)
        for (int j = 0; j < 6; ++j) o.r[i][j] = T(0);
    return o;
}
template <class T> inline G6<T> g6add(const G6<T>& a, const G6<T>& b) {
    G6<T> o = a;
    for (int i = 0; i < 6; ++i)
        for (int j = 0; j < 6; ++j) o.r[i][j] = o.r[i][j] + b.r[i][j];
    return o;
}
template <class T> inline G6<T> g6sub(const G6<T>& a, const G6<T>& b) {
    G6<T> o = a;
    for (int i = 0; i < 6; ++i)
        for (int j = 0; j < 6; ++j) o.r[i][j] = o.r[i][j] - b.r[i][j];
    return o;
}
template <class T> inline G6<T> g6scale(G6<T> a, T s) {  // aba.rs:1114-1124
    for (int i = 0; i < 6; ++i)
        for (int j = 0; j < 6; ++j) a.r[i][j] = a.r[i][j] * s;
    return a;
}
template <class T> inline Force<T> g6times(const G6<T>& a, Motion<T> m) {  // aba.rs:500-525
    T v[6] = {m.ang.x, m.ang.y, m.ang.z, m.lin.x, m.lin.y, m.lin.z};
    T o[6];
    for (int i = 0; i < 6; ++i) {
        o[i] = T(0);
        for (int j = 0; j < 6; ++j) o[i] = o[i] + a.r[i][j] * v[j];
    }
    return {{o[0], o[1], o[2]}, {o[3], o[4], o[5]}};
}
template <class T> inline G6<T> g6outer(Force<T> f, Motion<T> m) {  // aba.rs:526-550
    T rows[6] = {f.tq.x, f.tq.y, f.tq.z, f.lin.x, f.lin.y, f.lin.z};
    T cols[6] = {m.ang.x, m.ang.y, m.ang.z, m.lin.x, m.lin.y, m.lin.z};
    G6<T> o;
    for (int i = 0; i < 6; ++i)
        for (int j = 0; j < 6; ++j) o.r[i][j] = rows[i] * cols[j];
    return o;
}
template <class T> inline G6<T> g6pull_back(const G6<T>& a, const GX<T>& x) {  // aba.rs:551-613
    G6<T> o;
    for (int c = 0; c < 6; ++c) {
        Motion<T> e = mzero<T>();
        T* comp[6] = {&e.ang.x, &e.ang.y, &e.ang.z, &e.lin.x, &e.lin.y, &e.lin.z};
        *comp[c] = T(1);
        Motion<T> tm = gx_motion(x, e);
        Force<T> f = g6times(a, tm);
        Force<T> p = gx_tforce(x, f);
        T vals[6] = {p.tq.x, p.tq.y, p.tq.z, p.lin.x, p.lin.y, p.lin.z};
        for (int r = 0; r < 6; ++r) o.r[r][c] = vals[r];
    }
    return o;
}
// Total-order "greater" on |values| (f32::total_cmp on non-negative floats).
template <class T> inline bool abs_total_gt(T a, T b) {
    if constexpr (sizeof(T) == 4) {
        uint32_t ua, ub;
        T aa = rabs(a), bb = rabs(b);
        std::memcpy(&ua, &aa, 4);
        std::memcpy(&ub, &bb, 4);
        return ua > ub;
    } else {
        uint64_t ua, ub;
        T aa = rabs(a), bb = rabs(b);
        std::memcpy(&ua, &aa, 8);
        std::memcpy(&ub, &bb, 8);
        return ua > ub;
    }
}
template <class T> inline Motion<T> g6solve(const G6<T>& m, Force<T> rhs) {  // aba.rs:614-665
    T vals[6] = {rhs.tq.x, rhs.tq.y, rhs.tq.z, rhs.lin.x, rhs.lin.y, rhs.lin.z};
    T a[6][7];
    for (int i = 0; i < 6; ++i) {
        for (int j = 0; j < 6; ++j) a[i][j] = m.r[i][j];
        a[i][6] = vals[i];
    }
    for (int col = 0; col < 6; ++col) {
        // Iterator::max_by keeps the LAST of equal maxima.
        int pivot = col;
        for (int r = col + 1; r < 6; ++r)
            if (!abs_total_gt(a[pivot][col], a[r][col])) pivot = r;
        if (pivot != col)
            for (int k = 0; k < 7; ++k) std::swap(a[col][k], a[pivot][k]);
        T inv = T(1) / a[col][col];
        T prow[7];
        for (int k = 0; k < 7; ++k) prow[k] = a[col][k];
        for (int r = 0; r < 6; ++r) {
            if (r == col) continue;
            T factor = a[r][col] * inv;
            for (int k = col; k < 7; ++k) a[r][k] = a[r][k] - factor * prow[k];
        }
    }
    return {{a[0][6] / a[0][0], a[1][6] / a[1][1], a[2][6] / a[2][2]},
            {a[3][6] / a[3][3], a[4][6] / a[4][4], a[5][6] / a[5][5]}};
}

// Mat6::from_spatial_inertia with a zero COM offset (spatial.rs:355-381,
// tree.rs:173-175): link frames sit at the COM.
template <class T> inline G6<T> spatial_inertia6(T mass, const M3<T>& inertia) {
    V3<T> com = vzero<T>();
    M3<T> cx = skew(com);
    M3<T> cx_cx = cx * cx;
    M3<T> ul = inertia + cx_cx * (-mass);
    M3<T> ur = cx * mass;
    M3<T> ll = cx * (-mass);
    M3<T> lr = m3diag(mass, mass, mass);
    G6<T> o;
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) {
            o.r[r][c] = ul.get(r, c);
            o.r[r][c + 3] = ur.get(r, c);
            o.r[r + 3][c] = ll.get(r, c);
            o.r[r + 3][c + 3] = lr.get(r, c);
        }
    return o;
}

// ---------------------------------------------------------------------------
// Model: links, trees, geoms (newt/src/tree.rs:79-237, geom.rs:352-419)
// ---------------------------------------------------------------------------

enum JointKind { J_FREE = 0, J_FIXED = 1, J_HINGE = 2, J_SLIDE = 3, J_BALL = 4 };
enum ShapeKind { S_PLANE = 0, S_SPHERE = 1, S_BOX = 2, S_CAPSULE = 3, S_MESH = 4 };

// ConvexMesh (geom.rs:305-312): mesh-local vertices and outward-wound
// triangles.
template <class T> struct ConvexMesh {
    std::vector<V3<T>> v;
    std::vector<std::array<int, 3>> f;
};
enum Integrator { I_EULER = 0, I_IMPLICITFAST = 1, I_RK4 = 2 };
enum SolverMode { M_PENALTY = 0, M_PGS = 1, M_NEWTON = 2 };
enum Cone { C_PYRAMIDAL = 0, C_ELLIPTIC = 1 };

inline int joint_nq(int k) { return k == J_FREE ? 7 : k == J_FIXED ? 0 : k == J_BALL ? 4 : 1; }
inline int joint_nv(int k) { return k == J_FREE ? 6 : k == J_FIXED ? 0 : k == J_BALL ? 3 : 1; }

template <class T> struct Link {
    int parent = -1;
    int joint = J_FIXED;
    V3<T> axis = vzero<T>();
    T damping = T(0), armature = T(0), free_damping = T(0);
    bool has_range = false;
    T range_lo = T(0), range_hi = T(0);
    T limit_solref[2] = {T(0.02), T(1.0)};                    // JointLimit::solref or SolRef::DEFAULT
    T limit_solimp[4] = {T(0.9), T(0.95), T(0.001), T(0.5)};  // JointLimit::solimp or SolImp::DEFAULT
    int limit_power = 2;
    T limit_stiffness = T(1000.0), limit_damping = T(60.0);  // JointLimit penalty spring (JointLimit::DEFAULT)
    V3<T> pj_pos = vzero<T>();
    Q<T> pj_quat = qidentity<T>();
    V3<T> cj_pos = vzero<T>();
    Q<T> cj_quat = qidentity<T>();
    T mass = T(1);
    M3<T> inertia = m3identity<T>();
};

template <class T> struct Tree {
    std::vector<Link<T>> links;
    std::vector<int> q_offset, v_offset;
    std::vector<T> q, qdot, qfrc_applied;
    bool disable_penalty_limits = false;
    int nq() const { return (int)q.size(); }
    int nv() const { return (int)qdot.size(); }
    // Tree::push_link (tree.rs:275-330): default q written into the slots.
    int push_link(const Link<T>& l) {
        int idx = (int)links.size();
        if (idx == 0 && !(l.joint == J_FREE || l.joint == J_FIXED)) fail("root joint must be Free or Fixed");
        if (idx > 0 && !(l.parent >= 0 && l.parent < idx)) fail("parent index must precede link");
        q_offset.push_back((int)q.size());
        v_offset.push_back((int)qdot.size());
        switch (l.joint) {
            case J_FREE:
                q.push_back(l.pj_pos.x); q.push_back(l.pj_pos.y); q.push_back(l.pj_pos.z);
                q.push_back(l.pj_quat.x); q.push_back(l.pj_quat.y); q.push_back(l.pj_quat.z);
                q.push_back(l.pj_quat.w);
                break;
            case J_FIXED: break;
            case J_HINGE: case J_SLIDE: q.push_back(T(0)); break;
            case J_BALL: q.push_back(T(0)); q.push_back(T(0)); q.push_back(T(0)); q.push_back(T(1)); break;
        }
        for (int k = 0; k < joint_nv(l.joint); ++k) {
            qdot.push_back(T(0));
            qfrc_applied.push_back(T(0));
        }
        links.push_back(l);
        return idx;
    }
};

template <class T> struct SolRef {
    T timeconst = T(0.02), dampratio = T(1.0);  // geom.rs:99-102
    bool is_direct() const { return timeconst < T(0.0); }
};
template <class T> struct SolImp {  // solver.rs:187-193
    T dmin = T(0.9), dmax = T(0.95), width = T(0.001), midpoint = T(0.5);
    uint32_t power = 2;
};

template <class T> struct Geom {
    int shape = S_PLANE;
    int tree = -1, link = -1;  // -1: static
    int body = -1;             // free body it rides on (Geom::body), -1: none
    V3<T> size = vzero<T>();   // sphere r; box half extents; capsule (r, half_height)
    V3<T> pos = vzero<T>();
    Q<T> quat = qidentity<T>();
    T friction = T(1), margin = T(0), gap = T(0);
    SolRef<T> solref;
    SolImp<T> solimp;
    int condim = 3;
    uint32_t group = 0xffffffffu, mask = 0xffffffffu;
    int mesh = -1;  // GeomShape::Mesh::mesh_id
    std::shared_ptr<const ConvexMesh<T>> mesh_data;
    bool is_link() const { return tree >= 0; }
    bool is_body() const { return body >= 0; }
};

template <class T> struct Pose {
    V3<T> p;
    Q<T> q;
};

// geom_world_pose (geom.rs:887-898).
template <class T> inline Pose<T> geom_world_pose(const Geom<T>& g, V3<T> pp, Q<T> pq) {
    if (!g.is_link() && !g.is_body()) return {g.pos, g.quat};
    return {pp + rotate(pq, g.pos), pq * g.quat};
}

// ---------------------------------------------------------------------------
// Forward kinematics and link velocities
// ---------------------------------------------------------------------------

// forward_kinematics_into (tree.rs:771-858).
template <class T> void forward_kinematics(const Tree<T>& t, std::vector<Pose<T>>& out) {
    out.clear();
    for (size_t i = 0; i < t.links.size(); ++i) {
        const Link<T>& l = t.links[i];
        Pose<T> pose;
        switch (l.joint) {
            case J_FREE: {
                int o = t.q_offset[i];
                pose.p = v3(t.q[o], t.q[o + 1], t.q[o + 2]);
                pose.q = Q<T>{t.q[o + 3], t.q[o + 4], t.q[o + 5], t.q[o + 6]};
                break;
            }
            case J_FIXED: {
                V3<T> pp = vzero<T>();
                Q<T> po = qidentity<T>();
                if (l.parent >= 0) { pp = out[l.parent].p; po = out[l.parent].q; }
                V3<T> jp = pp + rotate(po, l.pj_pos);
                Q<T> jo = po * l.pj_quat;
                Q<T> co = jo * conj(l.cj_quat);
                pose.p = jp - rotate(co, l.cj_pos);
                pose.q = co;
                break;
            }
            case J_HINGE: {
                const Pose<T>& par = out[l.parent];
                T ang = t.q[t.q_offset[i]];
                V3<T> jp = par.p + rotate(par.q, l.pj_pos);
                Q<T> jr = from_axis_angle(l.axis, ang);
                Q<T> co = par.q * jr;
                pose.p = jp - rotate(co, l.cj_pos);
                pose.q = co;
                break;
            }
            case J_SLIDE: {
                const Pose<T>& par = out[l.parent];
                T s = t.q[t.q_offset[i]];
                Q<T> co = par.q;
                V3<T> jp = par.p + rotate(par.q, l.pj_pos);
                V3<T> sw = rotate(par.q, l.axis * s);
                pose.p = jp + sw - rotate(co, l.cj_pos);
                pose.q = co;
                break;
            }
            case J_BALL: {
                const Pose<T>& par = out[l.parent];
                int o = t.q_offset[i];
                Q<T> qb{t.q[o], t.q[o + 1], t.q[o + 2], t.q[o + 3]};
                V3<T> jp = par.p + rotate(par.q, l.pj_pos);
                Q<T> co = par.q * qb;
                pose.p = jp - rotate(co, l.cj_pos);
                pose.q = co;
                break;
            }
        }
        out.push_back(pose);
    }
}

// link_world_velocities_into (world.rs:2777-2854): (linear at COM, angular).
template <class T>
void link_world_velocities(const Tree<T>& t, const std::vector<Pose<T>>& poses,
                           std::vector<std::pair<V3<T>, V3<T>>>& out) {
    out.assign(t.links.size(), {vzero<T>(), vzero<T>()});
    for (size_t i = 0; i < t.links.size(); ++i) {
        const Link<T>& l = t.links[i];
        if (l.parent < 0) {
            if (l.joint == J_FREE) {
                Q<T> ori = poses[i].q;
                int o = t.v_offset[i];
                V3<T> ab{t.qdot[o], t.qdot[o + 1], t.qdot[o + 2]};
                V3<T> lb{t.qdot[o + 3], t.qdot[o + 4], t.qdot[o + 5]};
                out[i] = {rotate(ori, lb), rotate(ori, ab)};
            }
            continue;
        }
        V3<T> pv = out[l.parent].first, pw = out[l.parent].second;
        V3<T> cp = poses[i].p;
        V3<T> pp = poses[l.parent].p;
        Q<T> po = poses[l.parent].q;
        switch (l.joint) {
            case J_HINGE: {
                V3<T> aw = rotate(po, l.axis);
                T qd = t.qdot[t.v_offset[i]];
                V3<T> ang = pw + aw * qd;
                V3<T> jw = pp + rotate(po, l.pj_pos);
                V3<T> pac = pv + cross(pw, cp - pp);
                out[i] = {pac + cross(aw * qd, cp - jw), ang};
                break;
            }
            case J_SLIDE: {
                V3<T> aw = rotate(po, l.axis);
                T qd = t.qdot[t.v_offset[i]];
                V3<T> pac = pv + cross(pw, cp - pp);
                out[i] = {pac + aw * qd, pw};
                break;
            }
            case J_BALL: {
                int o = t.v_offset[i];
                V3<T> ca = rotate(poses[i].q, v3(t.qdot[o], t.qdot[o + 1], t.qdot[o + 2]));
                V3<T> jw = pp + rotate(po, l.pj_pos);
                V3<T> pac = pv + cross(pw, cp - pp);
                out[i] = {pac + cross(ca, cp - jw), pw + ca};
                break;
            }
            case J_FIXED:
                out[i] = {pv + cross(pw, cp - pp), pw};
                break;
            default:
                out[i] = {vzero<T>(), vzero<T>()};
        }
    }
}

// ---------------------------------------------------------------------------
// Xup transforms (tree.rs:1124-1168, aba.rs:762-800)
// ---------------------------------------------------------------------------

// xup_for_link (tree.rs:1124-1131) as a GXform.
template <class T> inline GX<T> xup_fixed(const Link<T>& l) {
    Q<T> rel = l.pj_quat * conj(l.cj_quat);
    M3<T> r = to_mat3(conj(rel));
    V3<T> tr = l.cj_pos - r * l.pj_pos;
    return {g3from(r), tr};
}

// aba.rs `transform` for hinge and slide (aba.rs:762-791).
template <class T> inline GX<T> xup_joint(const Link<T>& l, T q) {
    if (l.joint == J_HINGE) {
        V3<T> ax = normalize(l.axis);
        T half = (T(0) - q) * T(0.5);
        T s = nsin(half), c = ncos(half);
        G3<T> rot = quaternion_matrix(ax.x * s, ax.y * s, ax.z * s, c);
        return {rot, l.cj_pos - g3mul(rot, l.pj_pos)};
    }
    // slide
    return {g3identity<T>(), (l.cj_pos - l.pj_pos) - l.axis * q};
}

// ---------------------------------------------------------------------------
// ABA (aba.rs:803-1102) and its tree wrapper (tree.rs:1021-1116)
// ---------------------------------------------------------------------------

template <class T> struct AbaWs {
    std::vector<GX<T>> xup;
    std::vector<Motion<T>> s, v, c, a;
    std::vector<std::array<Motion<T>, 3>> s3;
    std::vector<G6<T>> ia;
    std::vector<Force<T>> pa, ia_s;
    std::vector<T> d, tau;
    std::vector<G3<T>> d3_inv;
    void ensure(size_t n) {
        if (xup.size() == n) return;
        xup.assign(n, gxidentity<T>());
        s.assign(n, mzero<T>()); v = s; c = s; a = s;
        s3.assign(n, {mzero<T>(), mzero<T>(), mzero<T>()});
        ia.assign(n, g6zero<T>());
        pa.assign(n, fzero<T>()); ia_s = pa;
        d.assign(n, T(0)); tau = d;
        d3_inv.assign(n, g3zero<T>());
    }
};

// aba::run.  forces_scalar[i] per link, forces_ball[i] per link, free = root wrench.
template <class T>
void aba_run(const Tree<T>& t, const std::vector<T>& q, const std::vector<T>& qdot,
             const std::vector<Force<T>>& ext_body, const std::vector<T>& fs,
             const std::vector<V3<T>>& fb, Force<T> ffree, const std::vector<T>& damping_mass,
             AbaWs<T>& w, std::vector<T>& out) {
    size_t n = t.links.size();
    w.ensure(n);
    for (size_t i = 0; i < n; ++i) {
        const Link<T>& l = t.links[i];
        switch (l.joint) {
            case J_FREE: {
                int o = t.v_offset[i];
                w.xup[i] = gxidentity<T>();
                w.v[i] = {{qdot[o], qdot[o + 1], qdot[o + 2]}, {qdot[o + 3], qdot[o + 4], qdot[o + 5]}};
                w.c[i] = mzero<T>();
                break;
            }
            case J_FIXED: {
                w.xup[i] = xup_fixed(l);
                Motion<T> pv = l.parent >= 0 ? w.v[l.parent] : mzero<T>();
                w.v[i] = gx_motion(w.xup[i], pv);
                w.c[i] = mzero<T>();
                break;
            }
            case J_HINGE:
            case J_SLIDE: {
                GX<T> x = xup_joint(l, q[t.q_offset[i]]);
                w.xup[i] = x;
                if (l.joint == J_HINGE) w.s[i] = {l.axis, cross(l.cj_pos, l.axis)};
                else w.s[i] = {vzero<T>(), l.axis};
                Motion<T> sq = mscale(w.s[i], qdot[t.v_offset[i]]);
                w.v[i] = madd(gx_motion(x, w.v[l.parent]), sq);
                w.c[i] = cross_motion(w.v[i], sq);
                break;
            }
            case J_BALL: {
                int o = t.q_offset[i];
                G3<T> rot = quaternion_matrix(T(0) - q[o], T(0) - q[o + 1], T(0) - q[o + 2], q[o + 3]);
                w.xup[i] = {rot, l.cj_pos - g3mul(rot, l.pj_pos)};
                V3<T> r = l.cj_pos;
                V3<T> ex{T(1), T(0), T(0)}, ey{T(0), T(1), T(0)}, ez{T(0), T(0), T(1)};
                w.s3[i] = {Motion<T>{ex, cross(r, ex)}, Motion<T>{ey, cross(r, ey)}, Motion<T>{ez, cross(r, ez)}};
                int vo = t.v_offset[i];
                Motion<T> sq = madd(madd(mscale(w.s3[i][0], qdot[vo]), mscale(w.s3[i][1], qdot[vo + 1])),
                                    mscale(w.s3[i][2], qdot[vo + 2]));
                w.v[i] = madd(gx_motion(w.xup[i], w.v[l.parent]), sq);
                w.c[i] = cross_motion(w.v[i], sq);
                break;
            }
        }
    }
    for (size_t i = 0; i < n; ++i) {
        G6<T> inertia = spatial_inertia6(t.links[i].mass, t.links[i].inertia);
        w.ia[i] = inertia;
        w.pa[i] = fsub(cross_force(w.v[i], g6times(inertia, w.v[i])), ext_body[i]);
    }
    for (size_t ii = n; ii-- > 1;) {
        const Link<T>& l = t.links[ii];
        int p = l.parent;
        switch (l.joint) {
            case J_FIXED:
                w.ia[p] = g6add(w.ia[p], g6pull_back(w.ia[ii], w.xup[ii]));
                w.pa[p] = fadd(w.pa[p], gx_tforce(w.xup[ii], fadd(w.pa[ii], g6times(w.ia[ii], w.c[ii]))));
                break;
            case J_HINGE:
            case J_SLIDE: {
                Force<T> ias = g6times(w.ia[ii], w.s[ii]);
                T d = (dot_force(w.s[ii], ias) + l.armature) + damping_mass[ii];
                Force<T> ps = fadd(w.pa[ii], g6times(w.ia[ii], w.c[ii]));
                T u = fs[ii] - dot_force(w.s[ii], ps);
                T qdd = u / d;
                Force<T> pa_full = fadd(ps, fscale(ias, qdd));
                T inv_d = T(1) / d;
                Motion<T> ias_m{ias.tq, ias.lin};
                G6<T> ia_full = g6sub(w.ia[ii], g6scale(g6outer(ias, ias_m), inv_d));
                w.ia_s[ii] = ias;
                w.d[ii] = d;
                w.tau[ii] = fs[ii];
                w.ia[p] = g6add(w.ia[p], g6pull_back(ia_full, w.xup[ii]));
                w.pa[p] = fadd(w.pa[p], gx_tforce(w.xup[ii], pa_full));
                break;
            }
            case J_BALL: {
                auto& s3 = w.s3[ii];
                Force<T> ias3[3] = {g6times(w.ia[ii], s3[0]), g6times(w.ia[ii], s3[1]), g6times(w.ia[ii], s3[2])};
                G3<T> D = g3zero<T>();
                for (int r = 0; r < 3; ++r)
                    for (int c = 0; c < 3; ++c) D.d[r][c] = dot_force(s3[r], ias3[c]);
                for (int k = 0; k < 3; ++k) D.d[k][k] = (D.d[k][k] + l.armature) + damping_mass[ii];
                G3<T> Di = g3inverse(D);
                Force<T> ps = fadd(w.pa[ii], g6times(w.ia[ii], w.c[ii]));
                V3<T> sp{dot_force(s3[0], ps), dot_force(s3[1], ps), dot_force(s3[2], ps)};
                V3<T> qdd = g3mul(Di, fb[ii] - sp);
                Force<T> pa_full = fadd(fadd(fadd(ps, fscale(ias3[0], qdd.x)), fscale(ias3[1], qdd.y)), fscale(ias3[2], qdd.z));
                G6<T> ia_full = w.ia[ii];
                for (int c = 0; c < 3; ++c) {
                    Force<T> acol = fzero<T>();
                    for (int r = 0; r < 3; ++r) acol = fadd(acol, fscale(ias3[r], Di.d[r][c]));
                    Motion<T> m{ias3[c].tq, ias3[c].lin};
                    ia_full = g6sub(ia_full, g6outer(acol, m));
                }
                w.d3_inv[ii] = Di;
                w.ia[p] = g6add(w.ia[p], g6pull_back(ia_full, w.xup[ii]));
                w.pa[p] = fadd(w.pa[p], gx_tforce(w.xup[ii], pa_full));
                break;
            }
            default: fail("free joint below root");
        }
    }
    std::fill(out.begin(), out.end(), T(0));
    if (t.links[0].joint == J_FREE) {
        Force<T> rhs = fsub(ffree, w.pa[0]);
        T damping = t.links[0].free_damping;
        if (damping != T(0)) {
            int o = t.v_offset[0];
            Motion<T> qr{{qdot[o], qdot[o + 1], qdot[o + 2]}, {qdot[o + 3], qdot[o + 4], qdot[o + 5]}};
            rhs = fsub(rhs, Force<T>{qr.ang * damping, qr.lin * damping});
            for (int k = 0; k < 6; ++k) w.ia[0].r[k][k] = w.ia[0].r[k][k] + damping_mass[0];
        }
        Motion<T> root = g6solve(w.ia[0], rhs);
        w.a[0] = root;
        int o = t.v_offset[0];
        out[o] = root.ang.x; out[o + 1] = root.ang.y; out[o + 2] = root.ang.z;
        out[o + 3] = root.lin.x; out[o + 4] = root.lin.y; out[o + 5] = root.lin.z;
    }
    for (size_t i = 1; i < n; ++i) {
        const Link<T>& l = t.links[i];
        int p = l.parent;
        switch (l.joint) {
            case J_FIXED:
                w.a[i] = gx_motion(w.xup[i], w.a[p]);
                break;
            case J_HINGE:
            case J_SLIDE: {
                Motion<T> prime = madd(gx_motion(w.xup[i], w.a[p]), w.c[i]);
                Force<T> inner = fadd(g6times(w.ia[i], prime), w.pa[i]);
                T qdd = (fs[i] - dot_force(w.s[i], inner)) / w.d[i];
                w.a[i] = madd(prime, mscale(w.s[i], qdd));
                out[t.v_offset[i]] = qdd;
                break;
            }
            case J_BALL: {
                Motion<T> prime = madd(gx_motion(w.xup[i], w.a[p]), w.c[i]);
                Force<T> inner = fadd(g6times(w.ia[i], prime), w.pa[i]);
                auto& s3 = w.s3[i];
                V3<T> rhs{fb[i].x - dot_force(s3[0], inner), fb[i].y - dot_force(s3[1], inner),
                          fb[i].z - dot_force(s3[2], inner)};
                V3<T> qdd = g3mul(w.d3_inv[i], rhs);
                w.a[i] = madd(madd(madd(prime, mscale(s3[0], qdd.x)), mscale(s3[1], qdd.y)), mscale(s3[2], qdd.z));
                int o = t.v_offset[i];
                out[o] = qdd.x; out[o + 1] = qdd.y; out[o + 2] = qdd.z;
                break;
            }
            default: break;
        }
    }
}

enum VelImplicit { VI_EXPLICIT = 0, VI_JOINT_DAMPING = 1, VI_JOINT_DAMPING_ACT = 2 };

// joint_limit_scalar_force (tree.rs:1170-1207): the one-sided penalty
// spring-damper outside a hinge/slide range.
template <class T> T joint_limit_scalar_force(T q, T qdot, T lo, T hi, T k, T c) {
    if (q < lo) {
        T violation = lo - q;
        T spring = k * violation;
        T damp = qdot < T(0.0) ? -c * qdot : T(0.0);
        return spring + damp;
    }
    if (q > hi) {
        T violation = q - hi;
        T spring = -k * violation;
        T damp = qdot > T(0.0) ? -c * qdot : T(0.0);
        return spring + damp;
    }
    return T(0.0);
}

// assemble_joint_forces (forces.rs:6-85) for trees without tendons or
// actuators; joint_limit_scalar_force (tree.rs:1170-1207).
template <class T>
Force<T> assemble_joint_forces(const Tree<T>& t, std::vector<T>& scalar, std::vector<V3<T>>& ball) {
    std::fill(scalar.begin(), scalar.end(), T(0));
    std::fill(ball.begin(), ball.end(), vzero<T>());
    Force<T> free = fzero<T>();
    const T tendon = T(0.0);
    for (size_t i = 0; i < t.links.size(); ++i) {
        const Link<T>& l = t.links[i];
        switch (l.joint) {
            case J_FREE:
                if (i == 0) {
                    bool ap = t.disable_penalty_limits;
                    T f[6];
                    for (int k = 0; k < 6; ++k) f[k] = tendon + (ap ? t.qfrc_applied[k] : T(0.0));
                    free = {{f[0], f[1], f[2]}, {f[3], f[4], f[5]}};
                }
                break;
            case J_FIXED: break;
            case J_HINGE:
            case J_SLIDE: {
                T qq = t.q[t.q_offset[i]];
                T qd = t.qdot[t.v_offset[i]];
                T tau_limit = T(0.0);
                if (!t.disable_penalty_limits && l.has_range)
                    tau_limit = joint_limit_scalar_force(qq, qd, l.range_lo, l.range_hi, l.limit_stiffness,
                                                         l.limit_damping);
                T tau_act = T(-0.0);  // empty Sum
                scalar[i] = (((t.qfrc_applied[t.v_offset[i]] + tendon) - l.damping * qd) + tau_limit) + tau_act;
                break;
            }
            case J_BALL: {
                int o = t.v_offset[i];
                V3<T> om{t.qdot[o], t.qdot[o + 1], t.qdot[o + 2]};
                ball[i] = {(t.qfrc_applied[o] + tendon) - l.damping * om.x,
                           (t.qfrc_applied[o + 1] + tendon) - l.damping * om.y,
                           (t.qfrc_applied[o + 2] + tendon) - l.damping * om.z};
                break;
            }
        }
    }
    return free;
}

// aba_with_velocity_implicit_workspace (tree.rs:1021-1116).  ext: per-link
// world (force, torque) at COM.
template <class T>
void aba_tree(const Tree<T>& t, const std::vector<Pose<T>>& poses, V3<T> gravity,
              const std::vector<std::pair<V3<T>, V3<T>>>& ext, int mode, T dt, AbaWs<T>& ws,
              std::vector<T>& qacc) {
    size_t n = t.links.size();
    std::vector<T> fs(n, T(0));
    std::vector<V3<T>> fb(n, vzero<T>());
    Force<T> ffree = assemble_joint_forces(t, fs, fb);
    std::vector<Force<T>> ext_body(n);
    std::vector<T> dm(n, T(0));
    for (size_t i = 0; i < n; ++i) {
        const V3<T> fa = vzero<T>(), ta = vzero<T>();  // applied_wrenches (zero on this path)
        V3<T> force_body = inverse_rotate(poses[i].q, (ext[i].first + fa) + gravity * t.links[i].mass);
        V3<T> torque_body = inverse_rotate(poses[i].q, ext[i].second + ta);
        ext_body[i] = {torque_body, force_body};
        const Link<T>& l = t.links[i];
        switch (l.joint) {
            case J_HINGE:
            case J_SLIDE:
                // implicit_mass_damping (tree.rs:976-1001): no actuators here.
                dm[i] = mode == VI_EXPLICIT ? T(0.0) : dt * (l.damping + T(0.0));
                break;
            case J_BALL:
                dm[i] = mode == VI_EXPLICIT ? T(0.0) : dt * l.damping;
                break;
            case J_FREE:
                dm[i] = mode == VI_EXPLICIT ? T(0.0) : dt * l.free_damping;
                break;
            default: dm[i] = T(0.0);
        }
    }
    qacc.assign(t.nv(), T(0));
    aba_run(t, t.q, t.qdot, ext_body, fs, fb, ffree, dm, ws, qacc);
}

// ---------------------------------------------------------------------------
// Integrators (tree.rs:1234-1475)
// ---------------------------------------------------------------------------

// integrate_tree_positions (tree.rs:1422-1475), no mocap.
template <class T> void integrate_tree_positions(Tree<T>& t, T dt) {
    for (size_t i = 0; i < t.links.size(); ++i) {
        const Link<T>& l = t.links[i];
        switch (l.joint) {
            case J_FREE: {
                int qo = t.q_offset[i], vo = t.v_offset[i];
                Q<T> ori{t.q[qo + 3], t.q[qo + 4], t.q[qo + 5], t.q[qo + 6]};
                V3<T> om{t.qdot[vo], t.qdot[vo + 1], t.qdot[vo + 2]};
                V3<T> vel{t.qdot[vo + 3], t.qdot[vo + 4], t.qdot[vo + 5]};
                V3<T> dp = rotate(ori, vel) * dt;
                t.q[qo] += dp.x; t.q[qo + 1] += dp.y; t.q[qo + 2] += dp.z;
                Q<T> nx = integrate_body_angular_velocity(ori, om, dt);
                t.q[qo + 3] = nx.x; t.q[qo + 4] = nx.y; t.q[qo + 5] = nx.z; t.q[qo + 6] = nx.w;
                break;
            }
            case J_FIXED: break;
            case J_HINGE:
            case J_SLIDE:
                t.q[t.q_offset[i]] += t.qdot[t.v_offset[i]] * dt;
                break;
            case J_BALL: {
                int qo = t.q_offset[i], vo = t.v_offset[i];
                Q<T> ori{t.q[qo], t.q[qo + 1], t.q[qo + 2], t.q[qo + 3]};
                V3<T> om{t.qdot[vo], t.qdot[vo + 1], t.qdot[vo + 2]};
                Q<T> nx = integrate_body_angular_velocity(ori, om, dt);
                t.q[qo] = nx.x; t.q[qo + 1] = nx.y; t.q[qo + 2] = nx.z; t.q[qo + 3] = nx.w;
                break;
            }
        }
    }
}

// euler_step_with_workspace (tree.rs:1393-1420) with zero external wrenches
// (the constraint-mode callback, world.rs:2027-2029).  Returns qddot.
// `wrench` (penalty contacts): the per-link world (force, torque) callback
// of step_trees / step_trees_euler, called with the tree and its FK poses.
template <class T>
using WrenchFn = std::function<std::vector<std::pair<V3<T>, V3<T>>>(const Tree<T>&, const std::vector<Pose<T>>&)>;

template <class T>
void euler_step(Tree<T>& t, V3<T> gravity, T dt, bool implicit_fast, AbaWs<T>& ws, std::vector<T>& qacc,
                const WrenchFn<T>* wrench = nullptr) {
    std::vector<Pose<T>> poses;
    forward_kinematics(t, poses);
    std::vector<std::pair<V3<T>, V3<T>>> ext(t.links.size(), {vzero<T>(), vzero<T>()});
    if (wrench) ext = (*wrench)(t, poses);
    aba_tree(t, poses, gravity, ext, implicit_fast ? VI_JOINT_DAMPING_ACT : VI_JOINT_DAMPING, dt, ws, qacc);
    for (int i = 0; i < t.nv(); ++i) t.qdot[i] += qacc[i] * dt;
    integrate_tree_positions(t, dt);
}

// tree_deriv (tree.rs:1485-1540).
template <class T> void tree_deriv(const Tree<T>& t, const std::vector<T>& qdd, std::vector<T>& dq, std::vector<T>& dv) {
    dq.assign(t.nq(), T(0.0));
    dv = qdd;
    for (size_t i = 0; i < t.links.size(); ++i) {
        const Link<T>& l = t.links[i];
        int qo = t.q_offset[i], vo = t.v_offset[i];
        switch (l.joint) {
            case J_FREE: {
                Q<T> ori{t.q[qo + 3], t.q[qo + 4], t.q[qo + 5], t.q[qo + 6]};
                V3<T> wb{t.qdot[vo], t.qdot[vo + 1], t.qdot[vo + 2]};
                V3<T> vb{t.qdot[vo + 3], t.qdot[vo + 4], t.qdot[vo + 5]};
                V3<T> vw = rotate(ori, vb);
                dq[qo] = vw.x; dq[qo + 1] = vw.y; dq[qo + 2] = vw.z;
                Q<T> dqo = qderivative(ori, wb);
                dq[qo + 3] = dqo.x; dq[qo + 4] = dqo.y; dq[qo + 5] = dqo.z; dq[qo + 6] = dqo.w;
                break;
            }
            case J_FIXED: break;
            case J_HINGE:
            case J_SLIDE: dq[qo] = t.qdot[vo]; break;
            case J_BALL: {
                Q<T> qb{t.q[qo], t.q[qo + 1], t.q[qo + 2], t.q[qo + 3]};
                V3<T> wb{t.qdot[vo], t.qdot[vo + 1], t.qdot[vo + 2]};
                Q<T> dqo = qderivative(qb, wb);
                dq[qo] = dqo.x; dq[qo + 1] = dqo.y; dq[qo + 2] = dqo.z; dq[qo + 3] = dqo.w;
                break;
            }
        }
    }
}

// rk4_step_with_workspace (tree.rs:1234-1364), zero external wrenches, no actuators.
// k1_out (test instrumentation, not in newt): the first stage's qddot.
template <class T>
void rk4_step(Tree<T>& t, V3<T> gravity, T dt, AbaWs<T>& ws, std::vector<T>* k1_out = nullptr,
              const WrenchFn<T>* wrench = nullptr) {
    std::vector<T> q0 = t.q, qd0 = t.qdot;
    std::vector<std::pair<V3<T>, V3<T>>> ext(t.links.size(), {vzero<T>(), vzero<T>()});
    std::vector<Pose<T>> poses;
    std::vector<T> k, dq[4], dv[4];
    T stage_dt[3] = {dt * T(0.5), dt * T(0.5), dt};
    for (int s = 0; s < 4; ++s) {
        forward_kinematics(t, poses);
        if (wrench) ext = (*wrench)(t, poses);
        aba_tree(t, poses, gravity, ext, VI_EXPLICIT, dt, ws, k);
        if (s == 0 && k1_out) *k1_out = k;
        tree_deriv(t, k, dq[s], dv[s]);
        if (s < 3) {  // advance_tree_state (tree.rs:1551-1573)
            T h = stage_dt[s];
            for (size_t j = 0; j < q0.size(); ++j) t.q[j] = q0[j] + dq[s][j] * h;
            for (size_t j = 0; j < qd0.size(); ++j) t.qdot[j] = qd0[j] + dv[s][j] * h;
        }
    }
    T sixth = T(1.0) / T(6.0);
    for (size_t j = 0; j < q0.size(); ++j)
        t.q[j] = q0[j] + (((dq[0][j] + T(2.0) * dq[1][j]) + T(2.0) * dq[2][j]) + dq[3][j]) * (dt * sixth);
    for (size_t j = 0; j < qd0.size(); ++j)
        t.qdot[j] = qd0[j] + (((dv[0][j] + T(2.0) * dv[1][j]) + T(2.0) * dv[2][j]) + dv[3][j]) * (dt * sixth);
    if (t.links[0].joint == J_FREE) {
        Q<T> qn = renormalize(Q<T>{t.q[3], t.q[4], t.q[5], t.q[6]});
        t.q[3] = qn.x; t.q[4] = qn.y; t.q[5] = qn.z; t.q[6] = qn.w;
    }
    for (size_t i = 0; i < t.links.size(); ++i) {
        if (t.links[i].joint != J_BALL) continue;
        int o = t.q_offset[i];
        Q<T> qn = renormalize(Q<T>{t.q[o], t.q[o + 1], t.q[o + 2], t.q[o + 3]});
        t.q[o] = qn.x; t.q[o + 1] = qn.y; t.q[o + 2] = qn.z; t.q[o + 3] = qn.w;
    }
}

// ---------------------------------------------------------------------------
// CRB mass matrix, Cholesky, RNE (dynamics.rs:151-648)
// ---------------------------------------------------------------------------

// Spatial-vector Xform (spatial.rs:245-316) over Mat3 column-major.  CRB and
// RNE use the non-generic types; the arithmetic is the same expression trees.
template <class T> struct XF {
    M3<T> rot;
    V3<T> tr;
};
template <class T> inline Motion<T> xf_motion(const XF<T>& x, Motion<T> m) {  // spatial.rs:283-287
    V3<T> w = x.rot * m.ang;
    return {w, x.rot * m.lin + cross(x.tr, w)};
}
template <class T> inline Force<T> xf_tforce(const XF<T>& x, Force<T> f) {  // spatial.rs:310-315
    M3<T> rt = transpose(x.rot);
    return {rt * (f.tq - cross(x.tr, f.lin)), rt * f.lin};
}
// Mat6 (spatial.rs:343-552), row-major rows[r][c].
template <class T> inline Force<T> m6times(const G6<T>& a, Motion<T> m) {  // spatial.rs:384-405
    T v[6] = {m.ang.x, m.ang.y, m.ang.z, m.lin.x, m.lin.y, m.lin.z};
    T o[6];
    for (int r = 0; r < 6; ++r) {
        T s = T(0.0);
        for (int c = 0; c < 6; ++c) s += a.r[r][c] * v[c];
        o[r] = s;
    }
    return {{o[0], o[1], o[2]}, {o[3], o[4], o[5]}};
}
template <class T> inline G6<T> m6pull_back(const G6<T>& a, const XF<T>& x) {  // spatial.rs:457-481
    G6<T> o;
    for (int c = 0; c < 6; ++c) {
        Motion<T> e = mzero<T>();
        T* comp[6] = {&e.ang.x, &e.ang.y, &e.ang.z, &e.lin.x, &e.lin.y, &e.lin.z};
        *comp[c] = T(1.0);
        Force<T> f = m6times(a, xf_motion(x, e));
        Force<T> p = xf_tforce(x, f);
        T vals[6] = {p.tq.x, p.tq.y, p.tq.z, p.lin.x, p.lin.y, p.lin.z};
        for (int r = 0; r < 6; ++r) o.r[r][c] = vals[r];
    }
    return o;
}
template <class T> inline G6<T> m6plus(G6<T> a, const G6<T>& b) {
    for (int r = 0; r < 6; ++r)
        for (int c = 0; c < 6; ++c) a.r[r][c] += b.r[r][c];
    return a;
}

// compute_xup (dynamics.rs:151-181) with xup_for_link_* (tree.rs:1124-1168).
template <class T> std::vector<XF<T>> compute_xup(const Tree<T>& t) {
    std::vector<XF<T>> x(t.links.size());
    for (size_t i = 0; i < t.links.size(); ++i) {
        const Link<T>& l = t.links[i];
        switch (l.joint) {
            case J_FREE: x[i] = {m3identity<T>(), vzero<T>()}; break;
            case J_FIXED: {
                Q<T> rel = l.pj_quat * conj(l.cj_quat);
                M3<T> r = to_mat3(conj(rel));
                x[i] = {r, l.cj_pos - r * l.pj_pos};
                break;
            }
            case J_HINGE: {
                M3<T> r = to_mat3(from_axis_angle(l.axis, -t.q[t.q_offset[i]]));
                x[i] = {r, l.cj_pos - r * l.pj_pos};
                break;
            }
            case J_SLIDE:
                x[i] = {m3identity<T>(), (l.cj_pos - l.pj_pos) - l.axis * t.q[t.q_offset[i]]};
                break;
            case J_BALL: {
                int o = t.q_offset[i];
                M3<T> r = to_mat3(conj(Q<T>{t.q[o], t.q[o + 1], t.q[o + 2], t.q[o + 3]}));
                x[i] = {r, l.cj_pos - r * l.pj_pos};
                break;
            }
        }
    }
    return x;
}
template <class T> inline Motion<T> subspace_single(const Link<T>& l) {  // dynamics.rs:185-194
    if (l.joint == J_HINGE) return {l.axis, cross(l.cj_pos, l.axis)};
    if (l.joint == J_SLIDE) return {vzero<T>(), l.axis};
    return mzero<T>();
}
template <class T> inline std::array<Motion<T>, 3> subspace_ball(const Link<T>& l) {  // dynamics.rs:198-205
    V3<T> r = l.cj_pos;
    V3<T> ex{T(1), T(0), T(0)}, ey{T(0), T(1), T(0)}, ez{T(0), T(0), T(1)};
    return {Motion<T>{ex, cross(r, ex)}, Motion<T>{ey, cross(r, ey)}, Motion<T>{ez, cross(r, ez)}};
}
inline double joint_armature_unused() { return 0; }

// mass_matrix (dynamics.rs:474-578), row-major nv×nv.
template <class T> std::vector<T> mass_matrix(const Tree<T>& t) {
    size_t n = t.links.size();
    int nv = t.nv();
    std::vector<T> m((size_t)nv * nv, T(0.0));
    if (nv == 0) return m;
    std::vector<XF<T>> xup = compute_xup(t);
    std::vector<G6<T>> ic(n);
    for (size_t i = 0; i < n; ++i) ic[i] = spatial_inertia6(t.links[i].mass, t.links[i].inertia);
    for (size_t i = n; i-- > 1;) {
        int p = t.links[i].parent;
        if (p >= 0) ic[p] = m6plus(ic[p], m6pull_back(ic[i], xup[i]));
    }
    if (t.links[0].joint == J_FREE) {
        int o = t.v_offset[0];
        for (int r = 0; r < 6; ++r)
            for (int c = 0; c < 6; ++c) m[(size_t)(o + r) * nv + (o + c)] = ic[0].r[r][c];
    }
    for (size_t i = 1; i < n; ++i) {
        const Link<T>& l = t.links[i];
        int voi = t.v_offset[i];
        std::vector<Motion<T>> cols;
        if (l.joint == J_FIXED) continue;
        if (l.joint == J_HINGE || l.joint == J_SLIDE) cols.push_back(subspace_single(l));
        else if (l.joint == J_BALL) { auto s3 = subspace_ball(l); cols.assign(s3.begin(), s3.end()); }
        T arm = l.armature;
        for (size_t k = 0; k < cols.size(); ++k) {
            Force<T> f = m6times(ic[i], cols[k]);
            for (size_t kp = 0; kp < cols.size(); ++kp)
                m[(size_t)(voi + kp) * nv + (voi + k)] = dot_force(cols[kp], f);
            m[(size_t)(voi + k) * nv + (voi + k)] += arm;
            int jc = (int)i;
            while (t.links[jc].parent >= 0) {
                int j = t.links[jc].parent;
                f = xf_tforce(xup[jc], f);
                jc = j;
                int voj = t.v_offset[j];
                const Link<T>& pl = t.links[j];
                switch (pl.joint) {
                    case J_FIXED: break;
                    case J_HINGE:
                    case J_SLIDE: {
                        T val = dot_force(subspace_single(pl), f);
                        m[(size_t)voj * nv + (voi + k)] = val;
                        m[(size_t)(voi + k) * nv + voj] = val;
                        break;
                    }
                    case J_BALL: {
                        auto s3 = subspace_ball(pl);
                        for (int kk = 0; kk < 3; ++kk) {
                            T val = dot_force(s3[kk], f);
                            m[(size_t)(voj + kk) * nv + (voi + k)] = val;
                            m[(size_t)(voi + k) * nv + (voj + kk)] = val;
                        }
                        break;
                    }
                    case J_FREE: {
                        T fc[6] = {f.tq.x, f.tq.y, f.tq.z, f.lin.x, f.lin.y, f.lin.z};
                        for (int kk = 0; kk < 6; ++kk) {
                            m[(size_t)(voj + kk) * nv + (voi + k)] = fc[kk];
                            m[(size_t)(voi + k) * nv + (voj + kk)] = fc[kk];
                        }
                        break;
                    }
                }
            }
        }
    }
    return m;
}

// Tree::implicit_mass_matrix (tree.rs:645-681), no actuators.
template <class T> std::vector<T> implicit_mass_matrix(const Tree<T>& t, T dt) {
    std::vector<T> m = mass_matrix(t);
    int nv = t.nv();
    if (nv == 0 || dt == T(0.0)) return m;
    for (size_t i = 0; i < t.links.size(); ++i) {
        const Link<T>& l = t.links[i];
        T d;
        switch (l.joint) {
            case J_FREE: d = l.free_damping; break;
            case J_HINGE: case J_SLIDE: d = l.damping + T(0.0); break;
            case J_BALL: d = l.damping; break;
            default: continue;
        }
        int o = t.v_offset[i];
        for (int k = 0; k < joint_nv(l.joint); ++k) m[(size_t)(o + k) * nv + (o + k)] += dt * d;
    }
    return m;
}

// cholesky (dynamics.rs:602-622): false when a pivot is not positive.
template <class T> bool cholesky(const std::vector<T>& a, int n, std::vector<T>& l) {
    l.assign((size_t)n * n, T(0.0));
    for (int i = 0; i < n; ++i)
        for (int j = 0; j <= i; ++j) {
            T sum = a[(size_t)i * n + j];
            for (int k = 0; k < j; ++k) sum -= l[(size_t)i * n + k] * l[(size_t)j * n + k];
            if (i == j) {
                if (sum <= T(0.0)) return false;
                l[(size_t)i * n + i] = std::sqrt(sum);
            } else {
                l[(size_t)i * n + j] = sum / l[(size_t)j * n + j];
            }
        }
    return true;
}
// cholesky_solve (dynamics.rs:626-648).
template <class T> std::vector<T> cholesky_solve(const std::vector<T>& l, int n, const std::vector<T>& b) {
    std::vector<T> y(n, T(0.0)), x(n, T(0.0));
    for (int i = 0; i < n; ++i) {
        T sum = b[i];
        for (int k = 0; k < i; ++k) sum -= l[(size_t)i * n + k] * y[k];
        y[i] = sum / l[(size_t)i * n + i];
    }
    for (int i = n - 1; i >= 0; --i) {
        T sum = y[i];
        for (int k = i + 1; k < n; ++k) sum -= l[(size_t)k * n + i] * x[k];
        x[i] = sum / l[(size_t)i * n + i];
    }
    return x;
}

// inverse_dynamics (dynamics.rs:232-418); bias_forces = inverse_dynamics(0).
template <class T>
std::vector<T> inverse_dynamics(const Tree<T>& t, const std::vector<T>& qddot, V3<T> gravity,
                                const std::vector<std::pair<V3<T>, V3<T>>>& ext) {
    size_t n = t.links.size();
    std::vector<Pose<T>> poses;
    forward_kinematics(t, poses);
    std::vector<XF<T>> xup(n);
    std::vector<Motion<T>> v(n, mzero<T>()), a(n, mzero<T>()), s(n, mzero<T>());
    std::vector<std::array<Motion<T>, 3>> s3(n);
    std::vector<Force<T>> f(n, fzero<T>());
    std::vector<XF<T>> xall = compute_xup(t);
    for (size_t i = 0; i < n; ++i) {
        const Link<T>& l = t.links[i];
        xup[i] = xall[i];
        switch (l.joint) {
            case J_FREE: {
                int o = t.v_offset[i];
                v[i] = {{t.qdot[o], t.qdot[o + 1], t.qdot[o + 2]}, {t.qdot[o + 3], t.qdot[o + 4], t.qdot[o + 5]}};
                a[i] = {{qddot[o], qddot[o + 1], qddot[o + 2]}, {qddot[o + 3], qddot[o + 4], qddot[o + 5]}};
                break;
            }
            case J_FIXED: {
                Motion<T> vp = l.parent >= 0 ? v[l.parent] : mzero<T>();
                Motion<T> ap = l.parent >= 0 ? a[l.parent] : mzero<T>();
                v[i] = xf_motion(xup[i], vp);
                a[i] = xf_motion(xup[i], ap);
                break;
            }
            case J_HINGE:
            case J_SLIDE: {
                s[i] = subspace_single(l);
                T qd = t.qdot[t.v_offset[i]], qdd = qddot[t.v_offset[i]];
                Motion<T> sqd = mscale(s[i], qd), sqdd = mscale(s[i], qdd);
                Motion<T> vp = xf_motion(xup[i], v[l.parent]);
                Motion<T> ap = xf_motion(xup[i], a[l.parent]);
                v[i] = madd(vp, sqd);
                Motion<T> c = cross_motion(v[i], sqd);
                a[i] = madd(madd(ap, sqdd), c);
                break;
            }
            case J_BALL: {
                s3[i] = subspace_ball(l);
                int o = t.v_offset[i];
                Motion<T> sqd = madd(madd(mscale(s3[i][0], t.qdot[o]), mscale(s3[i][1], t.qdot[o + 1])), mscale(s3[i][2], t.qdot[o + 2]));
                Motion<T> sqdd = madd(madd(mscale(s3[i][0], qddot[o]), mscale(s3[i][1], qddot[o + 1])), mscale(s3[i][2], qddot[o + 2]));
                Motion<T> vp = xf_motion(xup[i], v[l.parent]);
                Motion<T> ap = xf_motion(xup[i], a[l.parent]);
                v[i] = madd(vp, sqd);
                Motion<T> c = cross_motion(v[i], sqd);
                a[i] = madd(madd(ap, sqdd), c);
                break;
            }
        }
        G6<T> im = spatial_inertia6(l.mass, l.inertia);
        Force<T> ia = m6times(im, a[i]);
        Force<T> iv = m6times(im, v[i]);
        Force<T> bias = cross_force(v[i], iv);
        V3<T> fw = ext[i].first + gravity * l.mass;
        V3<T> fbdy = inverse_rotate(poses[i].q, fw);
        V3<T> tbdy = inverse_rotate(poses[i].q, ext[i].second);
        f[i] = fsub(fadd(ia, bias), Force<T>{tbdy, fbdy});
    }
    std::vector<T> tau(t.nv(), T(0.0));
    for (size_t i = n; i-- > 1;) {
        const Link<T>& l = t.links[i];
        int vo = t.v_offset[i];
        switch (l.joint) {
            case J_FIXED: f[l.parent] = fadd(f[l.parent], xf_tforce(xup[i], f[i])); break;
            case J_HINGE:
            case J_SLIDE:
                tau[vo] = dot_force(s[i], f[i]) + l.armature * qddot[vo];
                f[l.parent] = fadd(f[l.parent], xf_tforce(xup[i], f[i]));
                break;
            case J_BALL: {
                for (int k = 0; k < 3; ++k) tau[vo + k] = dot_force(s3[i][k], f[i]) + l.armature * qddot[vo + k];
                f[l.parent] = fadd(f[l.parent], xf_tforce(xup[i], f[i]));
                break;
            }
            default: fail("free joint below root");
        }
    }
    if (t.links[0].joint == J_FREE) {
        int o = t.v_offset[0];
        tau[o] = f[0].tq.x; tau[o + 1] = f[0].tq.y; tau[o + 2] = f[0].tq.z;
        tau[o + 3] = f[0].lin.x; tau[o + 4] = f[0].lin.y; tau[o + 5] = f[0].lin.z;
    }
    return tau;
}

// point_jacobian (jacobian.rs:44-110): translational/rotational columns.
template <class T>
void point_jacobian(const Tree<T>& t, const std::vector<Pose<T>>& poses, int link, V3<T> local,
                    std::vector<V3<T>>& jt, std::vector<V3<T>>& jr) {
    int nv = t.nv();
    jt.assign(nv, vzero<T>());
    jr.assign(nv, vzero<T>());
    V3<T> point = poses[link].p + rotate(poses[link].q, local);
    std::vector<int> chain;
    for (int i = link; i >= 0; i = t.links[i].parent) chain.push_back(i);
    std::reverse(chain.begin(), chain.end());
    const V3<T> E[3] = {{T(1), T(0), T(0)}, {T(0), T(1), T(0)}, {T(0), T(0), T(1)}};
    for (int i : chain) {
        const Link<T>& l = t.links[i];
        V3<T> com = poses[i].p;
        Q<T> ori = poses[i].q;
        switch (l.joint) {
            case J_FREE: {
                V3<T> r = point - com;
                int vo = t.v_offset[i];
                for (int k = 0; k < 3; ++k) {
                    V3<T> aw = rotate(ori, E[k]);
                    jr[vo + k] = aw;
                    jt[vo + k] = cross(aw, r);
                }
                for (int k = 0; k < 3; ++k) jt[vo + 3 + k] = rotate(ori, E[k]);
                break;
            }
            case J_FIXED: break;
            case J_HINGE: {
                const Pose<T>& par = poses[l.parent];
                V3<T> jw = par.p + rotate(par.q, l.pj_pos);
                V3<T> aw = rotate(par.q, l.axis);
                jr[t.v_offset[i]] = aw;
                jt[t.v_offset[i]] = cross(aw, point - jw);
                break;
            }
            case J_SLIDE:
                jt[t.v_offset[i]] = rotate(poses[l.parent].q, l.axis);
                break;
            case J_BALL: {
                const Pose<T>& par = poses[l.parent];
                V3<T> jw = par.p + rotate(par.q, l.pj_pos);
                for (int k = 0; k < 3; ++k) {
                    V3<T> aw = rotate(ori, E[k]);
                    int slot = t.v_offset[i] + k;
                    jr[slot] = aw;
                    jt[slot] = cross(aw, point - jw);
                }
                break;
            }
        }
    }
}

}  // namespace oracle
