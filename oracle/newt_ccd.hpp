// newt_ccd.hpp — CPU oracle: convex-mesh contacts.
//
// TEST INFRASTRUCTURE ONLY (see newt_ref.hpp).  Restates, for convex meshes
// (GeomShape::Mesh, newt/src/geom.rs:220, 305-340):
//   newt/src/contact/ccd.rs  route_pair, the mesh support functions with
//                            their tie rules, GJK (intersection and distance),
//                            EPA with its deterministic face order, the EPA
//                            fallback axis search, the polytope / distance
//                            witnesses and the mesh-mesh multi-contact
//                            clipper (ccd_convex_contacts, 2071-2092)
//   newt/src/contact.rs      mesh_plane (1424-1510), sphere_mesh (1720-1776),
//                            closest_point_on_triangle (2265-2306)
// Each function names the reference function it follows.  Results are the
// contact geometry only; the world wraps them into Contacts (geom ids, pair
// friction and gap, the slot within the pair's buffer).
#pragma once

#include "newt_ref.hpp"

namespace oracle {

template <class T> struct CcdContact {
    V3<T> pos, normal;
    T pen;
};

template <class T> inline T lsq(V3<T> a) { return dot(a, a); }

// f32::total_cmp on non-NaN values: -0.0 orders before +0.0.
template <class T> inline int total_cmp(T a, T b) {
    if (a < b) return -1;
    if (a > b) return 1;
    const bool sa = std::signbit(a), sb = std::signbit(b);
    if (sa && !sb) return -1;
    if (!sa && sb) return 1;
    return 0;
}
template <class T> inline bool lex_precedes(V3<T> a, V3<T> b) {  // contact.rs:64-74
    int c = total_cmp(a.x, b.x);
    if (c != 0) return c < 0;
    c = total_cmp(a.y, b.y);
    if (c != 0) return c < 0;
    return total_cmp(a.z, b.z) < 0;
}
template <class T> inline bool same_vec3(V3<T> a, V3<T> b) {  // contact.rs:76-78
    return a.x == b.x && a.y == b.y && a.z == b.z;
}
template <class T> inline T clamp01(T x) {  // contact.rs:2324-2333
    if (x < T(0.0)) return T(0.0);
    if (x > T(1.0)) return T(1.0);
    return x;
}
template <class T> inline V3<T> point_to_world(const Pose<T>& p, V3<T> l) { return p.p + rotate(p.q, l); }

// closest_point_on_triangle (contact.rs:2265-2306), Ericson's regions.
template <class T> V3<T> closest_point_on_triangle(V3<T> p, V3<T> a, V3<T> b, V3<T> c) {
    const V3<T> ab = b - a, ac = c - a, ap = p - a;
    const T d1 = dot(ab, ap), d2 = dot(ac, ap);
    if (d1 <= T(0.0) && d2 <= T(0.0)) return a;
    const V3<T> bp = p - b;
    const T d3 = dot(ab, bp), d4 = dot(ac, bp);
    if (d3 >= T(0.0) && d4 <= d3) return b;
    const T vc = d1 * d4 - d3 * d2;
    if (vc <= T(0.0) && d1 >= T(0.0) && d3 <= T(0.0)) {
        const T v = d1 / (d1 - d3);
        return a + ab * v;
    }
    const V3<T> cp = p - c;
    const T d5 = dot(ab, cp), d6 = dot(ac, cp);
    if (d6 >= T(0.0) && d5 <= d6) return c;
    const T vb = d5 * d2 - d1 * d6;
    if (vb <= T(0.0) && d2 >= T(0.0) && d6 <= T(0.0)) {
        const T w = d2 / (d2 - d6);
        return a + ac * w;
    }
    const T va = d3 * d6 - d5 * d4;
    if (va <= T(0.0) && (d4 - d3) >= T(0.0) && (d5 - d6) >= T(0.0)) {
        const T w = (d4 - d3) / ((d4 - d3) + (d5 - d6));
        return b + (c - b) * w;
    }
    const T denom = T(1.0) / ((va + vb) + vc);
    const T v = vb * denom, w = vc * denom;
    return (a + ab * v) + ac * w;
}

// ---------------------------------------------------------------------------
// Mesh support (ccd.rs:189-384)
// ---------------------------------------------------------------------------

template <class T> struct MeshShape {  // CcdShape::Mesh
    Pose<T> pose;
    const ConvexMesh<T>* mesh;
};

template <class T> T vertices_extent(const std::vector<V3<T>>& v) {  // ccd.rs:189-200
    V3<T> mn = v[0], mx = v[0];
    for (size_t i = 1; i < v.size(); ++i) {
        mn = {rmin(mn.x, v[i].x), rmin(mn.y, v[i].y), rmin(mn.z, v[i].z)};
        mx = {rmax(mx.x, v[i].x), rmax(mx.y, v[i].y), rmax(mx.z, v[i].z)};
    }
    return length(mx - mn);
}

// support_vertices (ccd.rs:202-234): the vertex furthest along `dir`; among
// vertices within extent·1e-7 of the best, the lexicographically first, and
// whether two or more tied.
template <class T> V3<T> support_vertices(const std::vector<V3<T>>& v, V3<T> dir, bool* tied) {
    *tied = false;
    const T dl = length(dir);
    if (dl <= T(1.0e-20)) return v[0];
    const V3<T> d = dir / dl;
    const T tie = vertices_extent(v) * T(1.0e-7);
    T best = -std::numeric_limits<T>::infinity();
    for (const V3<T>& p : v) best = rmax(best, dot(p, d));
    V3<T> sel = vzero<T>();
    int count = 0;
    for (const V3<T>& c : v)
        if (rabs(dot(c, d) - best) <= tie) {
            if (count == 0 || lex_precedes(c, sel)) sel = c;
            ++count;
        }
    *tied = count >= 2;
    return sel;
}

// support_feature_centroid (ccd.rs:236-306): the mean of the tied support
// vertices, distinct points taken in lexicographic order with their
// multiplicity, summed in f64; none when fewer than two tie.
template <class T> bool support_feature_centroid(const std::vector<V3<T>>& v, V3<T> dir, V3<T>* out) {
    const T dl = length(dir);
    if (dl <= T(1.0e-20)) return false;
    const V3<T> d = dir / dl;
    const T tie = vertices_extent(v) * T(1.0e-7);
    T best = -std::numeric_limits<T>::infinity();
    for (const V3<T>& p : v) best = rmax(best, dot(p, d));
    int tied = 0;
    for (const V3<T>& c : v)
        if (rabs(dot(c, d) - best) <= tie) ++tied;
    double sx = 0.0, sy = 0.0, sz = 0.0;
    bool have_prev = false;
    V3<T> prev = vzero<T>();
    for (int k = 0; k < tied; ++k) {
        bool found = false;
        V3<T> next = vzero<T>();
        for (const V3<T>& c : v) {  // Iterator::min_by keeps the first minimum
            if (!(rabs(dot(c, d) - best) <= tie)) continue;
            if (have_prev && !lex_precedes(prev, c)) continue;
            if (!found || lex_precedes(c, next)) {
                next = c;
                found = true;
            }
        }
        if (!found) break;
        int count = 0;
        for (const V3<T>& c : v)
            if (same_vec3(c, next)) ++count;
        sx += (double)next.x * (double)count;
        sy += (double)next.y * (double)count;
        sz += (double)next.z * (double)count;
        prev = next;
        have_prev = true;
    }
    if (tied < 2) return false;
    const double n = (double)tied;
    *out = V3<T>{T(sx / n), T(sy / n), T(sz / n)};
    return true;
}

template <class T> struct CcdVertex {  // ccd.rs:38-45
    V3<T> m, a, b;
    bool tie_a = false, tie_b = false;
};

template <class T> inline T ccd_sq() { return T(1.0e-12) * T(1.0e-12); }  // CCD_DEGENERATE_SQUARED

template <class T> V3<T> mesh_support_with_tie(const MeshShape<T>& s, V3<T> dir, bool* tied) {  // ccd.rs:125-174
    if (!(lsq(dir) > T(1.0e-20))) dir = v3(T(1.0), T(0.0), T(0.0));
    const V3<T> local = inverse_rotate(s.pose.q, dir);  // support_vertices_transformed (346-354)
    return point_to_world(s.pose, support_vertices(s.mesh->v, local, tied));
}
template <class T> V3<T> mesh_support(const MeshShape<T>& s, V3<T> dir) {
    bool t;
    return mesh_support_with_tie(s, dir, &t);
}
template <class T> bool mesh_centered_support(const MeshShape<T>& s, V3<T> dir, V3<T>* out) {  // ccd.rs:176-186
    V3<T> p;
    if (!support_feature_centroid(s.mesh->v, inverse_rotate(s.pose.q, dir), &p)) return false;
    *out = point_to_world(s.pose, p);
    return true;
}
template <class T> T mesh_extent(const MeshShape<T>& s) { return vertices_extent(s.mesh->v); }

template <class T> CcdVertex<T> ccd_support(const MeshShape<T>& a, const MeshShape<T>& b, V3<T> dir) {  // ccd.rs:356-371
    if (!(lsq(dir) > ccd_sq<T>())) dir = v3(T(1.0), T(0.0), T(0.0));
    CcdVertex<T> r;
    r.a = mesh_support_with_tie(a, dir, &r.tie_a);
    r.b = mesh_support_with_tie(b, -dir, &r.tie_b);
    r.m = r.a - r.b;
    return r;
}
template <class T>
CcdVertex<T> ccd_support_normalized(const MeshShape<T>& a, const MeshShape<T>& b, V3<T> dir, T extent) {  // ccd.rs:373-382
    CcdVertex<T> s = ccd_support(a, b, dir);
    s.m = s.m / extent;
    return s;
}
template <class T> inline V3<T> triple_product(V3<T> a, V3<T> b, V3<T> c) {  // ccd.rs:385-387
    return b * dot(a, c) - a * dot(b, c);
}

template <class T> bool ccd_vertex_precedes(const CcdVertex<T>& a, const CcdVertex<T>& b) {  // ccd.rs:322-334
    const V3<T> l[3] = {a.m, a.a, a.b}, r[3] = {b.m, b.a, b.b};
    for (int k = 0; k < 3; ++k) {
        if (same_vec3(l[k], r[k])) continue;
        return lex_precedes(l[k], r[k]);
    }
    return false;
}
template <class T> void sort_ccd_vertices(CcdVertex<T>* v, int n) {  // ccd.rs:336-344
    for (int i = 1; i < n; ++i)
        for (int p = i; p > 0 && ccd_vertex_precedes(v[p], v[p - 1]); --p) std::swap(v[p], v[p - 1]);
}

// ---------------------------------------------------------------------------
// GJK (ccd.rs:389-546)
// ---------------------------------------------------------------------------

template <class T> struct CcdSimplex {
    CcdVertex<T> p[4];
    int len = 1;
    explicit CcdSimplex(const CcdVertex<T>& v) { p[0] = p[1] = p[2] = p[3] = v; }
    void push(const CcdVertex<T>& v) { p[len++] = v; }
    void set(std::initializer_list<CcdVertex<T>> v) {
        len = 0;
        for (const auto& x : v) p[len++] = x;
    }
    void set(const CcdVertex<T>* v, int n) {
        for (int i = 0; i < n; ++i) p[i] = v[i];
        len = n;
    }
};

// ccd_simplex_step (ccd.rs:442-546): true when the simplex encloses the origin.
template <class T> bool ccd_simplex_step(CcdSimplex<T>& s, V3<T>& dir) {
    const T SQ = ccd_sq<T>();
    const CcdVertex<T> a = s.p[s.len - 1];
    const V3<T> ao = -a.m;
    if (s.len == 2) {
        const CcdVertex<T> b = s.p[0];
        const V3<T> ab = b.m - a.m;
        if (dot(ab, ao) > T(0.0)) {
            dir = triple_product(ab, ao, ab);
            if (lsq(dir) <= SQ) dir = v3(T(0.0), T(0.0), T(1.0));
        } else {
            s.set({a});
            dir = ao;
        }
    } else if (s.len == 3) {
        const CcdVertex<T> b = s.p[1], c = s.p[0];
        const V3<T> ab = b.m - a.m, ac = c.m - a.m;
        const V3<T> abc = cross(ab, ac);
        if (lsq(abc) <= SQ) {
            s.set({a, b});
            const T el = lsq(ab);
            const T t = el > SQ ? clamp01(-dot(a.m, ab) / el) : T(0.0);
            const V3<T> closest = a.m + ab * t;
            if (lsq(closest) > SQ) {
                dir = -closest;
                return false;
            }
            V3<T> axis;
            if (rabs(ab.x) <= rabs(ab.y) && rabs(ab.x) <= rabs(ab.z)) axis = v3(T(1.0), T(0.0), T(0.0));
            else if (rabs(ab.y) <= rabs(ab.z)) axis = v3(T(0.0), T(1.0), T(0.0));
            else axis = v3(T(0.0), T(0.0), T(1.0));
            dir = cross(ab, axis);
            if (lsq(dir) <= SQ) dir = ao;
            return false;
        }
        if (dot(cross(abc, ac), ao) > T(0.0)) {
            if (dot(ac, ao) > T(0.0)) {
                s.set({a, c});
                dir = triple_product(ac, ao, ac);
            } else if (dot(ab, ao) > T(0.0)) {
                s.set({a, b});
                dir = triple_product(ab, ao, ab);
            } else {
                s.set({a});
                dir = ao;
            }
        } else if (dot(cross(ab, abc), ao) > T(0.0)) {
            if (dot(ab, ao) > T(0.0)) {
                s.set({a, b});
                dir = triple_product(ab, ao, ab);
            } else {
                s.set({a});
                dir = ao;
            }
        } else if (dot(abc, ao) > T(0.0)) {
            dir = abc;
        } else {
            s.set({a, c, b});
            dir = -abc;
        }
    } else if (s.len == 4) {
        const CcdVertex<T> b = s.p[2], c = s.p[1], d = s.p[0];
        const V3<T> ab = b.m - a.m, ac = c.m - a.m, ad = d.m - a.m;
        const V3<T> abc = cross(ab, ac), acd = cross(ac, ad), adb = cross(ad, ab);
        if (dot(abc, ao) > T(0.0)) {
            s.set({a, c, b});
            dir = abc;
        } else if (dot(acd, ao) > T(0.0)) {
            s.set({a, d, c});
            dir = acd;
        } else if (dot(adb, ao) > T(0.0)) {
            s.set({a, b, d});
            dir = adb;
        } else {
            return true;
        }
    } else {
        fail("ccd_simplex_step: simplex size");
    }
    if (lsq(dir) <= SQ) dir = v3(T(1.0), T(0.0), T(0.0));
    return false;
}

// ---------------------------------------------------------------------------
// Witnesses (ccd.rs:548-694)
// ---------------------------------------------------------------------------

template <class T> void weighted_witness(const CcdVertex<T>* p, const T* w, int n, V3<T>* pa, V3<T>* pb) {
    V3<T> a = vzero<T>(), b = vzero<T>();
    for (int i = 0; i < n; ++i) {
        a = a + p[i].a * w[i];
        b = b + p[i].b * w[i];
    }
    *pa = a;
    *pb = b;
}
template <class T>
void centered_witness(const MeshShape<T>& sa, const MeshShape<T>& sb, V3<T>* pa, V3<T>* pb, V3<T> dir) {  // 558-574
    if (!(lsq(dir) > ccd_sq<T>())) dir = v3(T(1.0), T(0.0), T(0.0));
    V3<T> c;
    if (mesh_centered_support(sa, dir, &c)) *pa = c;
    if (mesh_centered_support(sb, -dir, &c)) *pb = c;
}
// ccd_feature_support (ccd.rs:617-647)
template <class T>
bool feature_support(const MeshShape<T>& s, const CcdVertex<T>* pts, V3<T> dir, bool is_a, bool* has, V3<T>* out) {
    *has = false;
    if (!(lsq(dir) > ccd_sq<T>())) return false;
    dir = normalize(dir);
    const V3<T> sup = mesh_support(s, dir);
    T fbest = -std::numeric_limits<T>::infinity();
    for (int i = 0; i < 3; ++i) fbest = rmax(fbest, dot(is_a ? pts[i].a : pts[i].b, dir));
    const T sbest = dot(sup, dir);
    const T tol = mesh_extent(s) * T(1.0e-7);
    if (rabs(fbest - sbest) > tol) return false;
    *has = mesh_centered_support(s, dir, out);
    return true;
}
// ccd_polytope_witness (ccd.rs:576-615); a mesh shape always has a face list.
template <class T>
void polytope_witness(const MeshShape<T>& sa, const MeshShape<T>& sb, const CcdVertex<T>* f, const T* w, V3<T> normal,
                      V3<T>* pa, V3<T>* pb) {
    weighted_witness(f, w, 3, pa, pb);
    const bool ua = sa.mesh->v.size() > 4 || f[0].tie_a || f[1].tie_a || f[2].tie_a;
    const bool ub = sb.mesh->v.size() > 4 || f[0].tie_b || f[1].tie_b || f[2].tie_b;
    bool ma = true, mb = true, ha = false, hb = false;
    V3<T> ca, cb;
    if (ua) ma = feature_support(sa, f, normal, true, &ha, &ca);
    if (ub) mb = feature_support(sb, f, -normal, false, &hb, &cb);
    if (ma && ha) *pa = ca;
    if (mb && hb) *pb = cb;
}
// barycentric_triangle_origin (ccd.rs:2094-2110)
template <class T> void barycentric(V3<T> a, V3<T> b, V3<T> c, V3<T> p, T* w) {
    const V3<T> v0 = b - a, v1 = c - a, v2 = p - a;
    const T d00 = dot(v0, v0), d01 = dot(v0, v1), d11 = dot(v1, v1), d20 = dot(v2, v0), d21 = dot(v2, v1);
    const T den = d00 * d11 - d01 * d01;
    if (rabs(den) <= T(1.0e-12)) {
        w[0] = T(1.0);
        w[1] = T(0.0);
        w[2] = T(0.0);
        return;
    }
    const T v = (d11 * d20 - d01 * d21) / den;
    const T ww = (d00 * d21 - d01 * d20) / den;
    w[0] = (T(1.0) - v) - ww;
    w[1] = v;
    w[2] = ww;
}
template <class T> void triangle_witness(const CcdVertex<T>& a, const CcdVertex<T>& b, const CcdVertex<T>& c, V3<T>* pa,
                                         V3<T>* pb) {  // ccd.rs:649-656
    const V3<T> near = closest_point_on_triangle(vzero<T>(), a.m, b.m, c.m);
    T w[3];
    barycentric(a.m, b.m, c.m, near, w);
    const CcdVertex<T> p[3] = {a, b, c};
    weighted_witness(p, w, 3, pa, pb);
}
// ccd_closest_witness (ccd.rs:658-694)
template <class T> bool closest_witness(const CcdSimplex<T>& s, V3<T>* pa, V3<T>* pb) {
    if (s.len == 1) {
        *pa = s.p[0].a;
        *pb = s.p[0].b;
        return true;
    }
    if (s.len == 2) {
        const V3<T> e = s.p[1].m - s.p[0].m;
        const T den = lsq(e);
        const T t = den > ccd_sq<T>() ? clamp01(-dot(s.p[0].m, e) / den) : T(0.0);
        const T w[2] = {T(1.0) - t, t};
        weighted_witness(s.p, w, 2, pa, pb);
        return true;
    }
    if (s.len == 3) {
        triangle_witness(s.p[0], s.p[1], s.p[2], pa, pb);
        return true;
    }
    if (s.len == 4) {
        const int faces[4][3] = {{0, 1, 2}, {0, 3, 1}, {0, 2, 3}, {1, 3, 2}};
        bool have = false;
        T bd = T(0.0);
        for (const auto& f : faces) {
            V3<T> a, b;
            triangle_witness(s.p[f[0]], s.p[f[1]], s.p[f[2]], &a, &b);
            const T d = lsq(a - b);
            if (!have || d < bd) {
                *pa = a;
                *pb = b;
                bd = d;
                have = true;
            }
        }
        return true;
    }
    return false;
}

// ---------------------------------------------------------------------------
// Distance GJK (ccd.rs:696-975)
// ---------------------------------------------------------------------------

template <class T> struct DistCand {
    V3<T> point;
    T w[4];
    CcdVertex<T> f[3];
    int flen;
};

template <class T> bool origin_inside_tetrahedron(const CcdSimplex<T>& s) {  // ccd.rs:704-725
    const V3<T> a = s.p[0].m, b = s.p[1].m, c = s.p[2].m, d = s.p[3].m;
    const V3<T> ab = b - a, ac = c - a, ad = d - a;
    const T den = dot(cross(ab, ac), ad);
    if (rabs(den) <= ccd_sq<T>()) return false;
    const V3<T> ao = -a;
    const T wb = dot(cross(ao, ac), ad) / den;
    const T wc = dot(cross(ab, ao), ad) / den;
    const T wd = dot(cross(ab, ac), ao) / den;
    const T wa = ((T(1.0) - wb) - wc) - wd;
    const T lo = T(-1.0e-6), hi = T(1.0) + T(1.0e-6);
    for (T w : {wa, wb, wc, wd})
        if (!(w >= lo && w <= hi)) return false;
    return true;
}
template <class T> bool distance_candidate(const CcdSimplex<T>& s, int mask, DistCand<T>* out) {  // ccd.rs:727-792
    int idx[3] = {0, 0, 0};
    int n = 0;
    for (int i = 0; i < s.len; ++i)
        if (mask & (1 << i)) {
            if (n == 3) return false;
            idx[n++] = i;
        }
    for (int i = 1; i < n; ++i)
        for (int p = i; p > 0 && ccd_vertex_precedes(s.p[idx[p]], s.p[idx[p - 1]]); --p) std::swap(idx[p], idx[p - 1]);
    T sw[3] = {T(0.0), T(0.0), T(0.0)};
    V3<T> point;
    if (n == 1) {
        point = s.p[idx[0]].m;
        sw[0] = T(1.0);
    } else if (n == 2) {
        const V3<T> a = s.p[idx[0]].m, b = s.p[idx[1]].m;
        const V3<T> e = b - a;
        const T den = lsq(e);
        const T t = den > ccd_sq<T>() ? clamp01(-dot(a, e) / den) : T(0.0);
        point = a + e * t;
        sw[0] = T(1.0) - t;
        sw[1] = t;
    } else if (n == 3) {
        const V3<T> a = s.p[idx[0]].m, b = s.p[idx[1]].m, c = s.p[idx[2]].m;
        if (lsq(cross(b - a, c - a)) <= ccd_sq<T>()) return false;
        point = closest_point_on_triangle(vzero<T>(), a, b, c);
        barycentric(a, b, c, point, sw);
    } else {
        return false;
    }
    for (int k = 0; k < 4; ++k) out->w[k] = T(0.0);
    for (int k = 0; k < n; ++k) out->w[idx[k]] = sw[k];
    for (int k = 0; k < 3; ++k) out->f[k] = s.p[idx[0]];
    for (int k = 0; k < n; ++k) out->f[k] = s.p[idx[k]];
    out->point = point;
    out->flen = n;
    return true;
}
template <class T> bool distance_candidate_precedes(const DistCand<T>& c, const DistCand<T>& cur) {  // ccd.rs:794-815
    const T cd = lsq(c.point), ud = lsq(cur.point);
    if (cd != ud) return cd < ud;
    if (c.flen != cur.flen) return c.flen < cur.flen;
    for (int i = 0; i < c.flen; ++i) {
        if (ccd_vertex_precedes(c.f[i], cur.f[i])) return true;
        if (ccd_vertex_precedes(cur.f[i], c.f[i])) return false;
    }
    return lex_precedes(c.point, cur.point);
}
// ccd_distance_simplex_step (ccd.rs:822-868): false = None.
template <class T> bool distance_simplex_step(CcdSimplex<T>& s, V3<T>* closest) {
    if (s.len == 4 && origin_inside_tetrahedron(s)) return false;
    bool have = false;
    DistCand<T> best{};
    for (int mask = 1; mask < (1 << s.len); ++mask) {
        DistCand<T> c{};
        if (!distance_candidate(s, mask, &c)) continue;
        if (!have || distance_candidate_precedes(c, best)) {
            best = c;
            have = true;
        }
    }
    if (!have) return false;
    CcdVertex<T> sel[4];
    int n = 0;
    for (int i = 0; i < s.len; ++i)
        if (best.w[i] > T(1.0e-6)) sel[n++] = s.p[i];
    if (n == 0) {
        int li = 0;
        for (int i = 1; i < s.len; ++i) {
            const T d = best.w[i] - best.w[li];
            if (d > T(1.0e-6) || (rabs(d) <= T(1.0e-6) && ccd_vertex_precedes(s.p[i], s.p[li]))) li = i;
        }
        sel[0] = s.p[li];
        n = 1;
    }
    sort_ccd_vertices(sel, n);
    s.set(sel, n);
    *closest = best.point;
    return true;
}

struct CcdConfig {  // CCD_MESH_CONFIG (ccd.rs:403-408)
    double gjk_support_epsilon = 0.0, distance_tolerance = 1.0e-10, epa_support_epsilon = 1.0e-7;
    int max_epa_iterations = 128;
};

// ccd_distance_contact (ccd.rs:870-908)
template <class T>
bool distance_contact(const MeshShape<T>& sa, const MeshShape<T>& sb, const CcdSimplex<T>& s, T margin,
                      CcdContact<T>* out) {
    V3<T> ra, rb;
    if (!closest_witness(s, &ra, &rb)) return false;
    const V3<T> sep = ra - rb;
    const T raw = length(sep);
    const T pen = margin - raw;
    if (pen <= T(0.0)) return false;
    V3<T> pa = ra, pb = rb;
    centered_witness(sa, sb, &pa, &pb, rb - ra);
    out->normal = raw > T(0.0) ? sep / raw : v3(T(1.0), T(0.0), T(0.0));
    out->pos = (pa + pb) * T(0.5);
    out->pen = pen;
    return true;
}
// ccd_distance_gjk (ccd.rs:910-975)
template <class T>
CcdSimplex<T> distance_gjk(const MeshShape<T>& sa, const MeshShape<T>& sb, CcdSimplex<T> s, const CcdConfig& cfg, T extent) {
    const T FLOOR = ccd_sq<T>();
    const T tol = T(cfg.distance_tolerance), geps = T(cfg.gjk_support_epsilon);
    CcdSimplex<T> best_s = s;
    T best = std::numeric_limits<T>::infinity();
    for (int it = 0; it < 64; ++it) {
        V3<T> closest;
        if (!distance_simplex_step(s, &closest)) return best_s;
        const T d2 = lsq(closest);
        if (d2 < best) {
            best = d2;
            best_s = s;
        }
        if (d2 <= FLOOR) return best_s;
        const V3<T> dir = -closest;
        const CcdVertex<T> p = ccd_support_normalized(sa, sb, dir, extent);
        const T progress = dot(p.m, dir) - dot(closest, dir);
        if (progress <= geps) return best_s;
        for (int i = 0; i < s.len; ++i)
            if (lsq(s.p[i].m - p.m) <= FLOOR) return best_s;
        s.push(p);
        V3<T> next;
        if (!distance_simplex_step(s, &next)) return best_s;
        const T nd2 = lsq(next);
        if (nd2 < best) {
            best = nd2;
            best_s = s;
        }
        if (d2 - nd2 <= tol * rmax(d2, FLOOR)) return best_s;
    }
    return best_s;
}

// ---------------------------------------------------------------------------
// EPA (ccd.rs:1030-1183, 1703-1918)
// ---------------------------------------------------------------------------

template <class T> struct CcdFace {
    int idx[3];
    V3<T> normal;
    T distance;
};
template <class T> bool ccd_face(const std::vector<CcdVertex<T>>& v, int i0, int i1, int i2, CcdFace<T>* out) {  // 1036-1057
    const V3<T> a = v[i0].m, b = v[i1].m, c = v[i2].m;
    const V3<T> raw = cross(b - a, c - a);
    if (lsq(raw) <= ccd_sq<T>()) return false;
    V3<T> n = normalize(raw);
    T d = dot(n, a);
    out->idx[0] = i0;
    out->idx[1] = i1;
    out->idx[2] = i2;
    if (d < T(0.0)) {
        out->idx[1] = i2;
        out->idx[2] = i1;
        n = -n;
        d = -d;
    }
    out->normal = n;
    out->distance = d;
    return true;
}
template <class T>
bool ccd_face_precedes(const std::vector<CcdVertex<T>>& v, const CcdFace<T>& c, const CcdFace<T>& cur, T eps) {  // 1059-1092
    if (c.distance + eps < cur.distance) return true;
    if (cur.distance + eps < c.distance) return false;
    CcdVertex<T> ck[3] = {v[c.idx[0]], v[c.idx[1]], v[c.idx[2]]};
    CcdVertex<T> uk[3] = {v[cur.idx[0]], v[cur.idx[1]], v[cur.idx[2]]};
    sort_ccd_vertices(ck, 3);
    sort_ccd_vertices(uk, 3);
    for (int i = 0; i < 3; ++i) {
        if (ccd_vertex_precedes(ck[i], uk[i])) return true;
        if (ccd_vertex_precedes(uk[i], ck[i])) return false;
    }
    return false;
}
// ccd_epa_fallback_contact (ccd.rs:1106-1183)
template <class T>
bool epa_fallback(const MeshShape<T>& sa, const MeshShape<T>& sb, const std::vector<CcdFace<T>>& faces, T extent,
                  T margin, CcdContact<T>* out) {
    bool have = false;
    V3<T> bax = vzero<T>();
    T bd = T(0.0);
    auto consider = [&](V3<T> axis) {
        if (lsq(axis) <= ccd_sq<T>()) return;
        axis = normalize(axis);
        for (V3<T> ax : {axis, -axis}) {
            const CcdVertex<T> s = ccd_support_normalized(sa, sb, ax, extent);
            const T d = dot(s.m, ax);
            if (d <= T(0.0) || (have && (d > bd || (d == bd && !lex_precedes(ax, bax))))) continue;
            bax = ax;
            bd = d;
            have = true;
        }
    };
    for (const CcdFace<T>& f : faces) consider(f.normal);
    for (const MeshShape<T>* s : {&sa, &sb})
        for (const auto& f : s->mesh->f) {  // ccd_mesh_face_normal (1094-1104)
            const V3<T> a = point_to_world(s->pose, s->mesh->v[f[0]]);
            const V3<T> b = point_to_world(s->pose, s->mesh->v[f[1]]);
            const V3<T> c = point_to_world(s->pose, s->mesh->v[f[2]]);
            const V3<T> n = cross(b - a, c - a);
            if (lsq(n) > T(0.0)) consider(normalize(n));
        }
    consider(sb.pose.p - sa.pose.p);
    consider(v3(T(1.0), T(0.0), T(0.0)));
    consider(v3(T(0.0), T(1.0), T(0.0)));
    consider(v3(T(0.0), T(0.0), T(1.0)));
    if (!have) return false;
    const CcdVertex<T> s = ccd_support(sa, sb, bax);
    const T pen = bd * extent + margin;
    if (!(pen > T(0.0))) return false;
    out->pos = s.b + bax * (bd * extent * T(0.5));
    out->normal = -bax;
    out->pen = pen;
    return true;
}

// ccd_convex_contact (ccd.rs:1703-1918): GJK to an enclosing tetrahedron, then
// EPA to the closest face; the separated case runs distance GJK.
template <class T>
bool ccd_convex_contact(const MeshShape<T>& sa, const MeshShape<T>& sb, const CcdConfig& cfg, T margin,
                        CcdContact<T>* out) {
    const T extent = rmax(mesh_extent(sa), mesh_extent(sb));
    if (extent <= T(0.0)) return false;
    V3<T> dir = sb.pose.p - sa.pose.p;
    if (lsq(dir) == T(0.0)) dir = v3(T(1.0), T(0.0), T(0.0));
    CcdSimplex<T> s(ccd_support_normalized(sa, sb, dir, extent));
    dir = -s.p[0].m;
    bool enclosed = false, separated = false;
    for (int it = 0; it < 32; ++it) {
        const CcdVertex<T> p = ccd_support_normalized(sa, sb, dir, extent);
        if (dot(p.m, dir) <= T(cfg.gjk_support_epsilon)) {
            separated = true;
            break;
        }
        s.push(p);
        if (ccd_simplex_step(s, dir)) {
            enclosed = true;
            break;
        }
    }
    if (!enclosed && separated) s = distance_gjk(sa, sb, s, cfg, extent);
    if (!enclosed || s.len != 4) return distance_contact(sa, sb, s, margin, out);

    std::vector<CcdVertex<T>> v(s.p, s.p + 4);
    std::vector<CcdFace<T>> faces;
    const int init[4][3] = {{0, 1, 2}, {0, 3, 1}, {0, 2, 3}, {1, 3, 2}};
    for (const auto& ix : init) {
        CcdFace<T> f;
        if (ccd_face(v, ix[0], ix[1], ix[2], &f)) faces.push_back(f);
    }
    if (faces.empty()) return distance_contact(sa, sb, s, margin, out);
    CcdFace<T> best = faces[0];
    const T feps = T(1.0e-4), eeps = T(cfg.epa_support_epsilon);
    bool converged = false;
    for (int it = 0; it < cfg.max_epa_iterations; ++it) {
        int bi = 0;
        for (int i = 0; i < (int)faces.size(); ++i)
            if (ccd_face_precedes(v, faces[i], faces[bi], feps)) bi = i;
        best = faces[bi];
        const CcdVertex<T> sup = ccd_support_normalized(sa, sb, best.normal, extent);
        const T sd = dot(sup.m, best.normal);
        if (sd - best.distance <= eeps) {
            converged = true;
            break;
        }
        if (v.size() == 128) break;
        bool dup = false;
        for (const CcdVertex<T>& x : v)
            if (lsq(x.m - sup.m) <= ccd_sq<T>()) dup = true;
        if (dup) break;
        v.push_back(sup);
        const int nv = (int)v.size() - 1;
        std::vector<CcdFace<T>> next;
        std::vector<std::array<int, 2>> edges;
        for (const CcdFace<T>& f : faces) {
            if (dot(f.normal, sup.m) > f.distance + feps) {
                const int e[3][2] = {{f.idx[0], f.idx[1]}, {f.idx[1], f.idx[2]}, {f.idx[2], f.idx[0]}};
                for (const auto& ed : e) {
                    int found = -1;
                    for (int k = 0; k < (int)edges.size(); ++k)
                        if (edges[k][0] == ed[1] && edges[k][1] == ed[0]) {
                            found = k;
                            break;
                        }
                    if (found >= 0) {
                        edges[found] = edges.back();
                        edges.pop_back();
                    } else {
                        edges.push_back({ed[0], ed[1]});
                    }
                }
            } else {
                next.push_back(f);
            }
        }
        for (const auto& ed : edges) {
            if (next.size() == 512) break;
            CcdFace<T> f;
            if (ccd_face(v, ed[0], ed[1], nv, &f)) next.push_back(f);
        }
        faces = next;
        if (faces.empty()) return false;
    }
    if (!converged) return epa_fallback(sa, sb, faces, extent, margin, out);
    const V3<T> nearest = best.normal * best.distance;
    T bary[3];
    barycentric(v[best.idx[0]].m, v[best.idx[1]].m, v[best.idx[2]].m, nearest, bary);
    const T sum = (bary[0] + bary[1]) + bary[2];
    if (sum <= T(1.0e-8)) return false;
    for (T& b : bary) b = b / sum;
    const CcdVertex<T> feat[3] = {v[best.idx[0]], v[best.idx[1]], v[best.idx[2]]};
    V3<T> pa, pb;
    polytope_witness(sa, sb, feat, bary, best.normal, &pa, &pb);
    const T pen = best.distance * extent + margin;
    if (pen <= T(0.0)) return false;
    out->normal = -best.normal;
    out->pos = (pa + pb) * T(0.5);
    out->pen = pen;
    return true;
}

// ---------------------------------------------------------------------------
// Mesh-mesh manifold (ccd.rs:1185-1701)
// ---------------------------------------------------------------------------

constexpr int kMultiCap = 16;           // MULTI_FEATURE_CAP
constexpr int kClipCap = 2 * kMultiCap;  // MULTI_CLIP_CAP

// mesh_support_ids (ccd.rs:1189-1221): false = None (overflow).
template <class T> bool mesh_support_ids(const MeshShape<T>& s, V3<T> dir, int* ids, int* len) {
    *len = 0;
    const V3<T> ld = inverse_rotate(s.pose.q, dir);
    const T dl = length(ld);
    if (dl <= T(1.0e-12)) return true;
    const V3<T> d = ld / dl;
    T best = -std::numeric_limits<T>::infinity();
    for (const V3<T>& p : s.mesh->v) best = rmax(best, dot(p, d));
    const T tol = mesh_extent(s) * T(1.0e-7);
    for (int i = 0; i < (int)s.mesh->v.size(); ++i)
        if (rabs(dot(s.mesh->v[i], d) - best) <= tol) {
            if (*len == kMultiCap) return false;
            ids[(*len)++] = i;
        }
    return true;
}
inline bool ids_contain(const int* ids, int n, int id) {
    for (int i = 0; i < n; ++i)
        if (ids[i] == id) return true;
    return false;
}
// mesh_feature_faces (ccd.rs:1228-1253)
template <class T> bool mesh_feature_faces(const ConvexMesh<T>& m, const int* ids, int n, int* faces, int* len) {
    *len = 0;
    for (int fi = 0; fi < (int)m.f.size(); ++fi) {
        const auto& f = m.f[fi];
        bool match;
        if (n >= 3) {
            match = ids_contain(ids, n, f[0]) && ids_contain(ids, n, f[1]) && ids_contain(ids, n, f[2]);
        } else {
            const int c = (int)ids_contain(ids, n, f[0]) + (int)ids_contain(ids, n, f[1]) + (int)ids_contain(ids, n, f[2]);
            match = c == n;
        }
        if (match) {
            if (*len == kMultiCap) return false;
            faces[(*len)++] = fi;
        }
    }
    return true;
}
// mesh_face_world (ccd.rs:1255-1274): the face's outward world normal.
template <class T> bool mesh_face_normal(const MeshShape<T>& s, int fi, V3<T>* n) {
    const auto& f = s.mesh->f[fi];
    const V3<T> la = s.mesh->v[f[0]], lb = s.mesh->v[f[1]], lc = s.mesh->v[f[2]];
    const V3<T> ln = cross(lb - la, lc - la);
    if (lsq(ln) <= ccd_sq<T>()) return false;
    *n = normalize(rotate(s.pose.q, normalize(ln)));
    return true;
}
// mesh_feature_normal (ccd.rs:1276-1295): false = None.
template <class T> bool mesh_feature_normal(const MeshShape<T>& s, const int* ids, int n, V3<T> expected, V3<T>* out) {
    int faces[kMultiCap], nf;
    if (!mesh_feature_faces(*s.mesh, ids, n, faces, &nf)) return false;
    bool have = false;
    for (int k = 0; k < nf; ++k) {
        V3<T> nn;
        if (!mesh_face_normal(s, faces[k], &nn)) continue;
        if (!have || dot(nn, expected) > dot(*out, expected)) {
            *out = nn;
            have = true;
        }
    }
    return have;
}
template <class T> bool polygon_start_precedes(V3<T> a, V3<T> b) {  // ccd.rs:1350-1360
    int c = total_cmp(a.x, b.x);
    if (c != 0) return c < 0;
    c = total_cmp(b.y, a.y);
    if (c != 0) return c < 0;
    return total_cmp(a.z, b.z) < 0;
}
template <class T> bool polygon_angle_precedes(V3<T> a, V3<T> b, V3<T> center, V3<T> axis, V3<T> tangent) {  // 1362-1379
    a = a - center;
    b = b - center;
    const T ax = dot(a, axis), ay = dot(a, tangent), bx = dot(b, axis), by = dot(b, tangent);
    const bool au = ay > T(0.0) || (ay == T(0.0) && ax >= T(0.0));
    const bool bu = by > T(0.0) || (by == T(0.0) && bx >= T(0.0));
    if (au != bu) return au;
    const T cr = ax * by - ay * bx;
    if (cr != T(0.0)) return cr > T(0.0);
    return lex_precedes(a, b);
}
// mesh_feature_polygon (ccd.rs:1297-1348): false = None.
template <class T>
bool mesh_feature_polygon(const MeshShape<T>& s, const int* ids, int n, V3<T> normal, V3<T>* poly, int* len) {
    *len = 0;
    for (int k = 0; k < n; ++k) {
        const V3<T> p = point_to_world(s.pose, s.mesh->v[ids[k]]);
        bool dup = false;
        for (int j = 0; j < *len; ++j)
            if (lsq(poly[j] - p) <= ccd_sq<T>()) dup = true;
        if (!dup) {
            if (*len == kClipCap) return false;
            poly[(*len)++] = p;
        }
    }
    if (*len < 3) return true;
    V3<T> c = vzero<T>();
    for (int j = 0; j < *len; ++j) c = c + poly[j];
    c = c / T(*len);
    V3<T> axis = normalize(poly[0] - c);
    if (lsq(axis) <= ccd_sq<T>()) axis = v3(T(1.0), T(0.0), T(0.0));
    const V3<T> tangent = normalize(cross(normal, axis));
    for (int i = 1; i < *len; ++i)
        for (int p = i; p > 0 && polygon_angle_precedes(poly[p], poly[p - 1], c, axis, tangent); --p)
            std::swap(poly[p], poly[p - 1]);
    return true;
}
// mesh_feature_edges (ccd.rs:1381-1430)
template <class T> int mesh_feature_edges(const MeshShape<T>& s, const int* ids, int n, V3<T> (*edges)[2]) {
    if (n >= 2) {
        edges[0][0] = point_to_world(s.pose, s.mesh->v[ids[0]]);
        edges[0][1] = point_to_world(s.pose, s.mesh->v[ids[1]]);
        return 1;
    }
    if (n == 0) return 0;
    int faces[kMultiCap], nf;
    if (!mesh_feature_faces(*s.mesh, ids, n, faces, &nf)) return 0;
    const T SQ = ccd_sq<T>();
    int len = 0;
    for (int k = 0; k < nf; ++k) {
        const auto& f = s.mesh->f[faces[k]];
        for (int o = 0; o < 3; ++o) {
            if (f[o] == ids[0]) continue;
            const V3<T> p0 = point_to_world(s.pose, s.mesh->v[ids[0]]);
            const V3<T> p1 = point_to_world(s.pose, s.mesh->v[f[o]]);
            bool dup = false;
            for (int j = 0; j < len; ++j)
                if ((lsq(edges[j][0] - p0) <= SQ && lsq(edges[j][1] - p1) <= SQ) ||
                    (lsq(edges[j][0] - p1) <= SQ && lsq(edges[j][1] - p0) <= SQ))
                    dup = true;
            if (!dup) {
                if (len == kMultiCap) return 0;
                edges[len][0] = p0;
                edges[len][1] = p1;
                ++len;
            }
        }
    }
    return len;
}
// polygon_clip (ccd.rs:1432-1482): Sutherland-Hodgman against the
// reference polygon's edges; false = None (overflow).
template <class T>
bool polygon_clip(const V3<T>* subject, int sn, const V3<T>* ref, int rn, V3<T> normal, V3<T>* out, int* on) {
    V3<T> poly[kClipCap], clip[kClipCap];
    for (int i = 0; i < kClipCap; ++i) poly[i] = i < sn ? subject[i] : vzero<T>();
    int pn = sn;
    for (int e = 0; e < rn; ++e) {
        const V3<T> p = ref[e], q = ref[(e + 1) % rn];
        const V3<T> edge = q - p;
        int cn = 0;
        for (int i = 0; i < kClipCap; ++i) clip[i] = vzero<T>();
        for (int i = 0; i < pn; ++i) {
            const V3<T> st = poly[i], en = poly[(i + 1) % pn];
            const T sd = dot(cross(edge, st - p), normal), ed = dot(cross(edge, en - p), normal);
            const bool si = sd >= T(-1.0e-6), ei = ed >= T(-1.0e-6);
            if (si != ei) {
                const T den = sd - ed;
                if (rabs(den) > T(1.0e-12)) {
                    if (cn == kClipCap) return false;
                    const T t = sd / den;
                    clip[cn++] = st + (en - st) * t;
                }
            }
            if (ei) {
                if (cn == kClipCap) return false;
                clip[cn++] = en;
            }
        }
        for (int i = 0; i < kClipCap; ++i) poly[i] = clip[i];
        pn = cn;
        if (pn == 0) break;
    }
    for (int i = 0; i < kClipCap; ++i) out[i] = poly[i];
    *on = pn;
    return true;
}
template <class T> T quad_area(const V3<T>* p, const int* ix) {  // ccd.rs:1484-1490
    const V3<T> a = p[ix[0]], b = p[ix[1]], c = p[ix[2]], d = p[ix[3]];
    return (length(cross(b - a, c - a)) + length(cross(c - a, d - a))) * T(0.5);
}
template <class T> void quad_indices(const V3<T>* p, int n, int* best) {  // ccd.rs:1492-1510
    for (int k = 0; k < 4; ++k) best[k] = k;
    T ba = quad_area(p, best);
    for (int a = 0; a < n - 3; ++a)
        for (int b = a + 1; b < n - 2; ++b)
            for (int c = b + 1; c < n - 1; ++c)
                for (int d = c + 1; d < n; ++d) {
                    const int cand[4] = {a, b, c, d};
                    const T ar = quad_area(p, cand);
                    if (ar > ba) {
                        ba = ar;
                        for (int k = 0; k < 4; ++k) best[k] = cand[k];
                    }
                }
}
template <class T> void ordered_indices(const V3<T>* p, const int* ix, int n, int* out) {  // ccd.rs:1512-1530
    int first = 0;
    for (int i = 1; i < n; ++i) {
        const bool pre = n == 2 ? lex_precedes(p[ix[i]], p[ix[first]]) : polygon_start_precedes(p[ix[i]], p[ix[first]]);
        if (pre) first = i;
    }
    for (int k = 0; k < 4; ++k) out[k] = 0;
    for (int i = 0; i < n; ++i) out[i] = ix[(first + n - i) % n];
}
// append_multi_contacts (ccd.rs:1532-1592): false = None.
template <class T>
bool append_multi(std::vector<CcdContact<T>>& out, const V3<T>* subject, int sn, const V3<T>* ref, int rn, V3<T> ref_normal,
                  V3<T> proj, bool subject_is_a, const CcdContact<T>& base, T margin) {
    V3<T> poly[kClipCap];
    int pn;
    if (!polygon_clip(subject, sn, ref, rn, ref_normal, poly, &pn)) return false;
    if (pn == 0) return true;
    int sel[4] = {0, 1, 2, 3};
    if (pn > 4) quad_indices(poly, pn, sel);
    const int n = pn < 4 ? pn : 4;
    int ord[4];
    ordered_indices(poly, sel, n, ord);
    const V3<T> rp = ref[0];
    for (int k = 0; k < n; ++k) {
        const V3<T> sp = poly[ord[k]];
        const T sg = dot(sp - rp, proj);
        const V3<T> pa = subject_is_a ? sp : sp - proj * sg;
        const V3<T> pb = subject_is_a ? sp - proj * sg : sp;
        const T pen = margin - dot(pa - pb, base.normal);
        if (pen <= T(0.0)) continue;
        const V3<T> pos = (pa + pb) * T(0.5);
        bool dup = false;
        for (const auto& c : out)
            if (lsq(c.pos - pos) <= ccd_sq<T>()) dup = true;
        if (pen > T(0.0) && !dup && out.size() < 4) out.push_back({pos, base.normal, pen});  // ContactBuf holds 4
    }
    return true;
}
// mesh_multicontact (ccd.rs:1594-1701): false = None (keep the single contact).
template <class T>
bool mesh_multicontact(const MeshShape<T>& sa, const MeshShape<T>& sb, const CcdContact<T>& base, T margin,
                       std::vector<CcdContact<T>>& out) {
    out.clear();
    if (margin > T(0.0)) return false;
    int ia[kMultiCap], ib[kMultiCap], na, nb;
    if (!mesh_support_ids(sa, -base.normal, ia, &na)) return false;
    if (!mesh_support_ids(sb, base.normal, ib, &nb)) return false;
    if (na == 0 || nb == 0) return false;
    V3<T> nA, nB;
    if (!mesh_feature_normal(sa, ia, na, -base.normal, &nA)) return false;
    if (!mesh_feature_normal(sb, ib, nb, base.normal, &nB)) return false;
    const T FACE = T(0.996), EDGE = T(0.0888);
    if (na >= 3 && nb >= 3 && dot(nA, nB) < -FACE) {
        V3<T> pa[kClipCap], pb[kClipCap];
        int la, lb;
        if (!mesh_feature_polygon(sa, ia, na, nA, pa, &la)) return false;
        if (!mesh_feature_polygon(sb, ib, nb, nB, pb, &lb)) return false;
        if (!append_multi(out, pb, lb, pa, la, nA, nA, false, base, margin)) return false;
    } else if (na < 3 && nb >= 3) {
        V3<T> edges[kMultiCap][2];
        const int ne = mesh_feature_edges(sa, ia, na, edges);
        V3<T> pb[kClipCap];
        int lb;
        if (!mesh_feature_polygon(sb, ib, nb, nB, pb, &lb)) return false;
        for (int e = 0; e < ne; ++e) {
            const V3<T> d = normalize(edges[e][1] - edges[e][0]);
            if (rabs(dot(d, nB)) < EDGE) {
                V3<T> subj[kClipCap];
                for (int i = 0; i < kClipCap; ++i) subj[i] = vzero<T>();
                subj[0] = edges[e][0];
                subj[1] = edges[e][1];
                if (!append_multi(out, subj, 2, pb, lb, nB, base.normal, true, base, margin)) return false;
                break;
            }
        }
    } else if (nb < 3 && na >= 3) {
        V3<T> edges[kMultiCap][2];
        const int ne = mesh_feature_edges(sb, ib, nb, edges);
        V3<T> pa[kClipCap];
        int la;
        if (!mesh_feature_polygon(sa, ia, na, nA, pa, &la)) return false;
        for (int e = 0; e < ne; ++e) {
            const V3<T> d = normalize(edges[e][1] - edges[e][0]);
            if (rabs(dot(d, nA)) < EDGE) {
                V3<T> subj[kClipCap];
                for (int i = 0; i < kClipCap; ++i) subj[i] = vzero<T>();
                subj[0] = edges[e][0];
                subj[1] = edges[e][1];
                if (!append_multi(out, subj, 2, pa, la, nA, base.normal, false, base, margin)) return false;
                break;
            }
        }
    }
    return out.size() > 1;
}

// ccd_convex_contacts (ccd.rs:2071-2092): the manifold, else the single
// GJK/EPA contact.  Normal from B into A.
template <class T>
std::vector<CcdContact<T>> mesh_mesh(const MeshShape<T>& sa, const MeshShape<T>& sb, T margin) {
    std::vector<CcdContact<T>> out;
    CcdContact<T> base;
    if (!ccd_convex_contact(sa, sb, CcdConfig{}, margin, &base)) return out;
    if (mesh_multicontact(sa, sb, base, margin, out)) return out;
    out.clear();
    out.push_back(base);
    return out;
}

// mesh_plane (contact.rs:1424-1510): the deepest support vertex with the
// largest coordinate along the plane's Y (lexicographic on ties), then the
// penetrating vertex with the smallest such coordinate.  Normal = the
// plane's (from the plane into the mesh).
template <class T>
std::vector<CcdContact<T>> mesh_plane(const Pose<T>& mp, const ConvexMesh<T>& mesh, T margin, const Pose<T>& pp) {
    const V3<T> n = rotate(pp.q, v3(T(0.0), T(0.0), T(1.0))), p0 = pp.p;
    const V3<T> tangent = rotate(pp.q, v3(T(0.0), T(1.0), T(0.0)));
    const T tie = vertices_extent(mesh.v) * T(1.0e-6);
    T maxpen = -std::numeric_limits<T>::infinity();
    for (const V3<T>& v : mesh.v) maxpen = rmax(maxpen, margin - dot(point_to_world(mp, v) - p0, n));
    bool hp = false, hs = false;
    T pt = T(0.0), st = T(0.0);
    V3<T> pw = vzero<T>(), sw = vzero<T>();
    for (const V3<T>& v : mesh.v) {
        const V3<T> w = point_to_world(mp, v);
        const T pen = margin - dot(w - p0, n);
        if (pen < T(0.0) || maxpen - pen > tie) continue;
        const T tv = dot(w, tangent);
        if (!hp || tv > pt || (tv == pt && lex_precedes(w, pw))) {
            hp = true;
            pt = tv;
            pw = w;
        }
    }
    if (hp)
        for (const V3<T>& v : mesh.v) {
            const V3<T> w = point_to_world(mp, v);
            const T pen = margin - dot(w - p0, n);
            if (pen < T(0.0)) continue;
            const T tv = dot(w, tangent);
            if (!same_vec3(w, pw) && (!hs || tv < st || (tv == st && lex_precedes(w, sw)))) {
                hs = true;
                st = tv;
                sw = w;
            }
        }
    std::vector<CcdContact<T>> out;
    for (int k = 0; k < 2; ++k) {
        if (!(k == 0 ? hp : hs)) continue;
        const V3<T> w = k == 0 ? pw : sw;
        const T sg = dot(w - p0, n);
        out.push_back({w - n * (sg * T(0.5)), n, margin - sg});
    }
    return out;
}

// sphere_mesh (contact.rs:1720-1776): the closest point on any face to the
// sphere centre (first strictly closest); normal from the mesh into the
// sphere.
template <class T>
std::vector<CcdContact<T>> sphere_mesh(const Pose<T>& sp, T r, const Pose<T>& mp, const ConvexMesh<T>& mesh, T margin) {
    const V3<T> pl = inverse_rotate(mp.q, sp.p - mp.p);
    T best = std::numeric_limits<T>::infinity();
    V3<T> bq = vzero<T>();
    for (const auto& f : mesh.f) {
        const V3<T> q = closest_point_on_triangle(pl, mesh.v[f[0]], mesh.v[f[1]], mesh.v[f[2]]);
        const T d2 = lsq(pl - q);
        if (d2 < best) {
            best = d2;
            bq = q;
        }
    }
    std::vector<CcdContact<T>> out;
    const T dist = std::sqrt(best);
    const T pen = (r - dist) + margin;
    if (pen <= T(0.0)) return out;
    const V3<T> nl = dist > T(1.0e-9) ? (pl - bq) / dist : v3(T(0.0), T(0.0), T(1.0));
    const V3<T> nw = rotate(mp.q, nl);
    const V3<T> ms = point_to_world(mp, bq);
    const V3<T> ss = sp.p - nw * r;
    out.push_back({(ms + ss) * T(0.5), nw, pen});
    return out;
}

}  // namespace oracle
