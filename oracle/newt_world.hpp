// newt_world.hpp — CPU oracle, world level: broadphase, narrowphase, the
// tree-contact PGS solve and World::step for articulated trees.
//
// TEST INFRASTRUCTURE ONLY (see newt_ref.hpp).  Restates:
//   newt/src/broadphase.rs   Aabb, fat-bound hysteresis, pair filter
//   newt/src/world.rs        step, active_pairs, populate_broadphase_tree,
//                            swept_geom_bound, collect_contacts_full,
//                            step_trees / step_trees_euler, tangent_basis
//   newt/src/contact.rs      dispatch_narrow_phase, sphere/box/capsule-plane,
//                            box-box SAT + face clipping manifold
//   newt/src/solver.rs       SolImp/SolRef math, solve_tree_contacts (PGS)
#pragma once

#include "newt_ccd.hpp"
#include "newt_ref.hpp"

namespace oracle {

// ---------------------------------------------------------------------------
// Aabb (broadphase.rs:22-122)
// ---------------------------------------------------------------------------

template <class T> struct Aabb {
    V3<T> mn, mx;
};
template <class T> inline Aabb<T> aabb_unbounded() {
    T inf = std::numeric_limits<T>::infinity();
    return {{-inf, -inf, -inf}, {inf, inf, inf}};
}
template <class T> inline Aabb<T> aabb_center_ext(V3<T> c, V3<T> e) { return {c - e, c + e}; }
template <class T> inline Aabb<T> expanded(Aabb<T> a, T m) {  // broadphase.rs:41-44
    V3<T> mm{m, m, m};
    return {a.mn - mm, a.mx + mm};
}
template <class T> inline Aabb<T> fatten(Aabb<T> a) {  // broadphase.rs:46-54
    const T FATNESS = T(0.1), EPS = T(1.0e-4);
    V3<T> e = a.mx - a.mn;
    V3<T> m{rmax(e.x * FATNESS, EPS), rmax(e.y * FATNESS, EPS), rmax(e.z * FATNESS, EPS)};
    return {a.mn - m, a.mx + m};
}
template <class T> inline Aabb<T> aunion(Aabb<T> a, Aabb<T> b) {  // broadphase.rs:56-69
    return {{rmin(a.mn.x, b.mn.x), rmin(a.mn.y, b.mn.y), rmin(a.mn.z, b.mn.z)},
            {rmax(a.mx.x, b.mx.x), rmax(a.mx.y, b.mx.y), rmax(a.mx.z, b.mx.z)}};
}
template <class T> inline bool overlaps(Aabb<T> a, Aabb<T> b) {  // broadphase.rs:71-78
    return a.mn.x <= b.mx.x && a.mx.x >= b.mn.x && a.mn.y <= b.mx.y && a.mx.y >= b.mn.y &&
           a.mn.z <= b.mx.z && a.mx.z >= b.mn.z;
}
template <class T> inline bool contains(Aabb<T> a, Aabb<T> b) {  // broadphase.rs:80-87
    return a.mn.x <= b.mn.x && a.mn.y <= b.mn.y && a.mn.z <= b.mn.z && a.mx.x >= b.mx.x &&
           a.mx.y >= b.mx.y && a.mx.z >= b.mx.z;
}

// geom_aabb (broadphase.rs:708-770) for the supported shapes.
template <class T> Aabb<T> geom_aabb(const Geom<T>& g, const Pose<T>& p) {
    switch (g.shape) {
        case S_PLANE: return aabb_unbounded<T>();
        case S_SPHERE: return aabb_center_ext(p.p, V3<T>{g.size.x, g.size.x, g.size.x});
        case S_BOX: {
            V3<T> x = vabs(rotate(p.q, v3(T(1), T(0), T(0)))) * g.size.x;
            V3<T> y = vabs(rotate(p.q, v3(T(0), T(1), T(0)))) * g.size.y;
            V3<T> z = vabs(rotate(p.q, v3(T(0), T(0), T(1)))) * g.size.z;
            return aabb_center_ext(p.p, (x + y) + z);
        }
        case S_CAPSULE: {
            V3<T> axis = rotate(p.q, v3(T(0), T(0), T(1)));
            T r = g.size.x;
            return aabb_center_ext(p.p, vabs(axis) * g.size.y + V3<T>{r, r, r});
        }
        case S_MESH: {  // transformed_vertices_aabb (broadphase.rs:780-791)
            const std::vector<V3<T>>& v = g.mesh_data->v;
            V3<T> first = point_to_world(p, v[0]);
            Aabb<T> b{first, first};
            for (size_t i = 1; i < v.size(); ++i) {
                V3<T> w = point_to_world(p, v[i]);
                b = aunion(b, Aabb<T>{w, w});
            }
            return b;
        }
    }
    fail("geom_aabb: unknown shape");
}

// ---------------------------------------------------------------------------
// SolRef / SolImp math (geom.rs:156-185, solver.rs:242-384)
// ---------------------------------------------------------------------------

template <class T> SolRef<T> combine_solref(SolRef<T> a, SolRef<T> b) {  // geom.rs:156-185
    if (a.is_direct() || b.is_direct()) {
        if (a.is_direct() && b.is_direct())
            return {a.timeconst <= b.timeconst ? a.timeconst : b.timeconst,
                    a.dampratio <= b.dampratio ? a.dampratio : b.dampratio};
        return a.is_direct() ? a : b;
    }
    return {a.timeconst <= b.timeconst ? a.timeconst : b.timeconst,
            a.dampratio <= b.dampratio ? a.dampratio : b.dampratio};
}
template <class T> inline T fmin_lt(T a, T b) { return a < b ? a : b; }  // solver.rs:253
template <class T> SolImp<T> combine_solimp(SolImp<T> a, SolImp<T> b) {  // solver.rs:242-250
    SolImp<T> o;
    o.dmin = fmin_lt(a.dmin, b.dmin);
    o.dmax = fmin_lt(a.dmax, b.dmax);
    o.width = fmin_lt(a.width, b.width);
    o.midpoint = a.midpoint;
    o.power = a.power;
    return o;
}
template <class T> T powi(T x, uint32_t n) {  // solver.rs:261-268
    T r = T(1.0);
    for (uint32_t i = 0; i < n; ++i) r *= x;
    return r;
}
template <class T> T sigmoid(T x, uint32_t power, T mid) {  // solver.rs:273-286
    if (x <= T(0.0)) return T(0.0);
    if (x >= T(1.0)) return T(1.0);
    if (x < mid) return mid * powi(x / mid, power);
    return T(1.0) - (T(1.0) - mid) * powi((T(1.0) - x) / (T(1.0) - mid), power);
}
template <class T> SolImp<T> effective_solimp(SolImp<T> s) {  // solver.rs:319-328
    const T MN = T(1.0e-4), MX = T(0.9999);
    SolImp<T> o;
    o.dmin = rclamp(s.dmin, MN, MX);
    o.dmax = rclamp(s.dmax, MN, MX);
    o.width = s.width > T(0.0) ? s.width : T(0.0);
    o.midpoint = rclamp(s.midpoint, MN, MX);
    o.power = s.power < 1 ? 1 : s.power;
    return o;
}
template <class T> T impedance_at_position(T position, T margin, SolImp<T> s0) {  // solver.rs:301-309
    SolImp<T> s = effective_solimp(s0);
    T x = (position - margin) / s.width;
    if (x < T(0.0)) x = -x;
    T y = sigmoid(x, s.power, s.midpoint);
    return s.dmin + (s.dmax - s.dmin) * y;
}
template <class T> SolRef<T> safe_solref(SolRef<T> r, T dt) {  // solver.rs:331-337
    if (!r.is_direct() && r.timeconst < T(2.0) * dt) return {T(2.0) * dt, r.dampratio};
    return r;
}
template <class T> void reference_coefficients(T position, SolRef<T> r, SolImp<T> s0, T* b, T* k, T* d) {  // solver.rs:363-384
    SolImp<T> s = effective_solimp(s0);
    *d = impedance_at_position(position, T(0.0), s);
    if (r.is_direct()) {
        *b = -r.dampratio / s.dmax;
        *k = -r.timeconst / (s.dmax * s.dmax);
    } else {
        T tc = r.timeconst, dmax = s.dmax;
        *b = T(2.0) / (dmax * tc);
        *k = *d / (dmax * dmax * tc * tc * r.dampratio * r.dampratio);
    }
}
template <class T> T reference_accel(T position, T velocity, SolRef<T> r, SolImp<T> s) {  // solver.rs:353-361
    T b, k, d;
    reference_coefficients(position, r, s, &b, &k, &d);
    return -b * velocity - k * position;
}
template <class T> T contact_regularization_from_diag(T base_diag, T imp, T mu, bool pyr3) {  // solver.rs:1961-1974
    T base = (T(1.0) - imp) / imp * base_diag;
    return pyr3 ? T(2.0) * mu * mu * base : base;
}

// tangent_basis (world.rs:3104-3111).
template <class T> std::pair<V3<T>, V3<T>> tangent_basis(V3<T> n) {
    bool use_z = rabs(n.z) < T(0.9);
    V3<T> ref = use_z ? v3(T(0), T(0), T(1)) : v3(T(1), T(0), T(0));
    V3<T> t1 = normalize(cross(ref, n));
    return {t1, cross(n, t1)};
}

// ---------------------------------------------------------------------------
// Contacts (contact.rs:82-214, 218-403, 716-834, 2509-2806)
// ---------------------------------------------------------------------------

template <class T> struct Contact {
    int geom_a, geom_b;
    int slot;  // index within the pair's ContactBuf (narrow-phase order)
    V3<T> pos, normal;
    T penetration, friction, gap;
};

template <class T> inline T combine_friction(T a, T b) { return a < b ? a : b; }  // contact.rs:218-220
template <class T> inline T combine_max(T a, T b) { return a > b ? a : b; }       // contact.rs:225-227

// Plane frame: outward normal = local +Z (contact.rs:260-265).
template <class T> inline void plane_world(const Pose<T>& pp, V3<T>* n, V3<T>* p0) {
    *n = rotate(pp.q, v3(T(0), T(0), T(1)));
    *p0 = pp.p;
}

template <class T>
void sphere_plane(int ia, const Pose<T>& ps, T radius, T fr, T margin, T gap, int ib, const Pose<T>& pp,
                  std::vector<Contact<T>>& out) {  // contact.rs:269-301
    V3<T> n, p0;
    plane_world(pp, &n, &p0);
    T signed_d = dot(ps.p - p0, n);
    T raw = signed_d - radius;
    if (raw <= margin) {
        V3<T> cp = ps.p - n * (radius + raw * T(0.5));
        out.push_back({ia, ib, (int)out.size(), cp, n, margin - raw, fr, gap});
    }
}

template <class T>
void box_plane(int ia, const Pose<T>& pb, V3<T> he, T fr, T margin, T gap, int ib, const Pose<T>& pp,
               std::vector<Contact<T>>& out) {  // contact.rs:311-362
    V3<T> n, p0;
    plane_world(pp, &n, &p0);
    T dist = dot(pb.p - p0, n);
    for (int i = 0; i < 8; ++i) {
        V3<T> local{(i & 1) == 0 ? -he.x : he.x, (i & 2) == 0 ? -he.y : he.y, (i & 4) == 0 ? -he.z : he.z};
        V3<T> co = rotate(pb.q, local);
        T ld = dot(n, co);
        T raw = dist + ld;
        if (raw > margin || ld > T(0.0)) continue;
        V3<T> corner = pb.p + co;
        V3<T> cp = corner - n * (raw * T(0.5));
        if (out.size() < 4) out.push_back({ia, ib, (int)out.size(), cp, n, margin - raw, fr, gap});
    }
}

template <class T>
void capsule_plane(int ia, const Pose<T>& pc, T radius, T hh, T fr, T margin, T gap, int ib,
                   const Pose<T>& pp, std::vector<Contact<T>>& out) {  // contact.rs:367-403
    V3<T> n, p0;
    plane_world(pp, &n, &p0);
    V3<T> aw = rotate(pc.q, v3(T(0), T(0), T(1)));
    V3<T> ends[2] = {pc.p + aw * hh, pc.p - aw * hh};
    for (int e = 0; e < 2; ++e) {
        T signed_d = dot(ends[e] - p0, n);
        T raw = signed_d - radius;
        if (raw <= margin) {
            V3<T> cp = ends[e] - n * (radius + raw * T(0.5));
            out.push_back({ia, ib, (int)out.size(), cp, n, margin - raw, fr, gap});
        }
    }
}

// closest_points_on_segments (contact.rs:2342-2385); clamp01 is in newt_ccd.hpp.
template <class T>
std::pair<V3<T>, V3<T>> closest_points_on_segments(V3<T> a0, V3<T> a1, V3<T> b0, V3<T> b1) {
    V3<T> d1 = a1 - a0, d2 = b1 - b0, r = a0 - b0;
    T a = dot(d1, d1), e = dot(d2, d2), f = dot(d2, r);
    const T EPS = T(1.0e-12);
    if (a <= EPS && e <= EPS) return {a0, b0};
    if (a <= EPS) return {a0, b0 + d2 * clamp01(f / e)};
    T c = dot(d1, r);
    if (e <= EPS) return {a0 + d1 * clamp01(-c / a), b0};
    T b_ = dot(d1, d2);
    T denom = a * e - b_ * b_;
    T s = denom != T(0.0) ? clamp01((b_ * f - c * e) / denom) : T(0.0);
    T tn = b_ * s + f;
    T t, sf;
    if (tn < T(0.0)) { t = T(0.0); sf = clamp01(-c / a); }
    else if (tn > e) { t = T(1.0); sf = clamp01((b_ - c) / a); }
    else { t = tn / e; sf = s; }
    return {a0 + d1 * sf, b0 + d2 * t};
}

// closest_point_on_segment (contact.rs:2313-2322).
template <class T> V3<T> closest_point_on_segment(V3<T> p, V3<T> a, V3<T> b) {
    V3<T> ab = b - a;
    T denom = dot(ab, ab);
    if (denom == T(0.0)) return a;
    T t = clamp01(dot(p - a, ab) / denom);
    return a + ab * t;
}

// sphere_vs_sphere_at (contact.rs:434-467): one contact on B's surface,
// normal B -> A (+Z when the centers coincide).
template <class T>
void sphere_vs_sphere_at(int ia, V3<T> ca, T ra, int ib, V3<T> cb, T rb, T fr, T margin, T gap,
                         std::vector<Contact<T>>& out) {
    V3<T> delta = ca - cb;
    T dist2 = dot(delta, delta);
    T sum_r_margin = (ra + rb) + margin;
    if (dist2 >= sum_r_margin * sum_r_margin) return;
    T dist = std::sqrt(dist2);
    V3<T> n = dist > T(0.0) ? delta / dist : V3<T>{T(0.0), T(0.0), T(1.0)};
    out.push_back({ia, ib, (int)out.size(), cb + n * rb, n, sum_r_margin - dist, fr, gap});
}

// Capsule segment end points: position ∓ axis·half_height, axis = local +Z.
template <class T> void capsule_segment(const Pose<T>& p, T half_h, V3<T>* a, V3<T>* b) {
    V3<T> ax = rotate(p.q, V3<T>{T(0.0), T(0.0), T(1.0)});
    *a = p.p - ax * half_h;
    *b = p.p + ax * half_h;
}

template <class T> struct FV2 {
    T u, v, depth;
};
// sutherland_hodgman_axis_rect (contact.rs:1115-1167).
template <class T> std::vector<FV2<T>> sutherland_hodgman(const FV2<T> subject[4], T hu, T hv) {
    auto sgn = [&](int edge, FV2<T> p) {
        switch (edge) {
            case 0: return p.u + hu;
            case 1: return hu - p.u;
            case 2: return p.v + hv;
            default: return hv - p.v;
        }
    };
    auto lerp = [](FV2<T> a, FV2<T> b, T t) {
        return FV2<T>{a.u + (b.u - a.u) * t, a.v + (b.v - a.v) * t, a.depth + (b.depth - a.depth) * t};
    };
    std::vector<FV2<T>> out(subject, subject + 4);
    for (int edge = 0; edge < 4; ++edge) {
        if (out.empty()) break;
        std::vector<FV2<T>> in = out;
        out.clear();
        size_t n = in.size();
        for (size_t i = 0; i < n; ++i) {
            FV2<T> cur = in[i], prev = in[(i + n - 1) % n];
            T cs = sgn(edge, cur), ps = sgn(edge, prev);
            bool ci = cs >= T(0.0), pi = ps >= T(0.0);
            if (ci) {
                if (!pi) out.push_back(lerp(prev, cur, ps / (ps - cs)));
                out.push_back(cur);
            } else if (pi) {
                out.push_back(lerp(prev, cur, ps / (ps - cs)));
            }
        }
    }
    return out;
}
inline void other_two(int k, int* a, int* b) {
    if (k == 0) { *a = 1; *b = 2; } else if (k == 1) { *a = 0; *b = 2; } else { *a = 0; *b = 1; }
}
// edge_midpoint (contact.rs:1173-1196).
template <class T> V3<T> edge_midpoint(V3<T> c, const V3<T>* basis, const T* half, int ei, V3<T> dir) {
    int j, k;
    other_two(ei, &j, &k);
    T sj = dot(dir, basis[j]) >= T(0.0) ? T(1.0) : T(-1.0);
    T sk = dot(dir, basis[k]) >= T(0.0) ? T(1.0) : T(-1.0);
    return (c + basis[j] * (sj * half[j])) + basis[k] * (sk * half[k]);
}

// box_box_full_manifold -> box_box_sat_fallback (contact.rs:569-583, 716-868),
// box_box_edge_edge_contact (872-910), box_box_face_reference_contacts (917-1079).
template <class T>
void box_box(int ia, const Pose<T>& pa, V3<T> ha3, int ib, const Pose<T>& pb, V3<T> hb3, T fr, T margin, T gap,
             std::vector<Contact<T>>& out) {
    V3<T> ax[3] = {rotate(pa.q, v3(T(1), T(0), T(0))), rotate(pa.q, v3(T(0), T(1), T(0))), rotate(pa.q, v3(T(0), T(0), T(1)))};
    V3<T> bx[3] = {rotate(pb.q, v3(T(1), T(0), T(0))), rotate(pb.q, v3(T(0), T(1), T(0))), rotate(pb.q, v3(T(0), T(0), T(1)))};
    T ha[3] = {ha3.x, ha3.y, ha3.z}, hb[3] = {hb3.x, hb3.y, hb3.z};
    V3<T> delta = pa.p - pb.p;
    auto proj = [](V3<T> l, const V3<T>* u, const T* h) {
        return (h[0] * rabs(dot(l, u[0])) + h[1] * rabs(dot(l, u[1]))) + h[2] * rabs(dot(l, u[2]));
    };
    T best = std::numeric_limits<T>::infinity();
    V3<T> best_axis = v3(T(0), T(0), T(1));
    int kind = 0, k1 = 0, k2 = 0;  // 0 face A, 1 face B, 2 edge-edge
    auto test = [&](V3<T> l, int knd, int i, int j) {
        T sgn = dot(delta, l);
        T ov = (proj(l, ax, ha) + proj(l, bx, hb)) - rabs(sgn);
        if (ov < -margin) return true;
        if (ov < best) {
            best = ov;
            best_axis = sgn >= T(0.0) ? l : -l;
            kind = knd; k1 = i; k2 = j;
        }
        return false;
    };
    for (int k = 0; k < 3; ++k) if (test(ax[k], 0, k, 0)) return;
    for (int k = 0; k < 3; ++k) if (test(bx[k], 1, k, 0)) return;
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            V3<T> raw = cross(ax[i], bx[j]);
            T len2 = dot(raw, raw);
            if (len2 < T(1.0e-6)) continue;
            if (test(raw / std::sqrt(len2), 2, i, j)) return;
        }
    if (!std::isfinite(best)) return;
    T pen_shift = best + margin;
    if (pen_shift <= T(0.0)) return;
    V3<T> nw = best_axis;
    if (kind == 2) {
        V3<T> ma = edge_midpoint(pa.p, ax, ha, k1, -nw);
        V3<T> mb = edge_midpoint(pb.p, bx, hb, k2, nw);
        V3<T> ea0 = ma - ax[k1] * ha[k1], ea1 = ma + ax[k1] * ha[k1];
        V3<T> eb0 = mb - bx[k2] * hb[k2], eb1 = mb + bx[k2] * hb[k2];
        V3<T> pbp = closest_points_on_segments(ea0, ea1, eb0, eb1).second;
        out.push_back({ia, ib, (int)out.size(), pbp, nw, pen_shift, fr, gap});
        return;
    }
    bool ref_a = kind == 0;
    int rk = k1;
    V3<T> rc = ref_a ? pa.p : pb.p;
    const V3<T>* rb = ref_a ? ax : bx;
    const T* rh = ref_a ? ha : hb;
    V3<T> ic = ref_a ? pb.p : pa.p;
    const V3<T>* ibs = ref_a ? bx : ax;
    const T* ih = ref_a ? hb : ha;
    V3<T> ron = ref_a ? -nw : nw;
    V3<T> rfa = rb[rk];
    T rs = dot(ron, rfa) >= T(0.0) ? T(1.0) : T(-1.0);
    V3<T> rfc = rc + rfa * (rs * rh[rk]);
    int ru, rv;
    other_two(rk, &ru, &rv);
    int iax = 0;
    T isg = T(1.0), idot = std::numeric_limits<T>::infinity();
    for (int k = 0; k < 3; ++k) {
        const T signs[2] = {T(1.0), T(-1.0)};
        for (T s : signs) {
            T d = dot(ibs[k] * s, ron);
            if (d < idot) { idot = d; iax = k; isg = s; }
        }
    }
    V3<T> ifc = ic + ibs[iax] * (isg * ih[iax]);
    int iu, iv;
    other_two(iax, &iu, &iv);
    V3<T> U = ibs[iu], Vv = ibs[iv];
    T hu = ih[iu], hv = ih[iv];
    V3<T> corners[4] = {(ifc - U * hu) - Vv * hv, (ifc + U * hu) - Vv * hv, (ifc + U * hu) + Vv * hv, (ifc - U * hu) + Vv * hv};
    FV2<T> subj[4];
    for (int i = 0; i < 4; ++i) {
        V3<T> rel = corners[i] - rfc;
        subj[i] = {dot(rel, rb[ru]), dot(rel, rb[rv]), -dot(rel, ron)};
    }
    std::vector<FV2<T>> clipped = sutherland_hodgman(subj, rh[ru], rh[rv]);
    T cpen[8];
    V3<T> cpos[8];
    int count = 0;
    for (const FV2<T>& cv : clipped) {
        T pen = cv.depth + margin;
        if (pen <= T(0.0)) continue;
        V3<T> por = (rfc + rb[ru] * cv.u) + rb[rv] * cv.v;
        cpen[count] = pen;
        cpos[count] = ref_a ? por + nw * (pen - margin) : por;
        if (++count == 8) break;
    }
    int order[8] = {0, 1, 2, 3, 4, 5, 6, 7};
    for (int i = 1; i < count; ++i) {
        int j = i;
        while (j > 0 && cpen[order[j]] > cpen[order[j - 1]]) {
            std::swap(order[j - 1], order[j]);
            --j;
        }
    }
    int take = count > 4 ? 4 : count;
    for (int i = 0; i < take; ++i)
        out.push_back({ia, ib, (int)out.size(), cpos[order[i]], nw, cpen[order[i]], fr, gap});
}

// face_along_direction (contact.rs:1209-1245): a vertex inside the box on
// every axis, assigned to the face along the dominant component of `dir`.
template <class T> bool face_along_direction(V3<T> lp, V3<T> h, V3<T> dir, T* pen, V3<T>* nl) {
    const T dx = h.x - rabs(lp.x), dy = h.y - rabs(lp.y), dz = h.z - rabs(lp.z);
    if (dx < T(0.0) || dy < T(0.0) || dz < T(0.0)) return false;
    const T ax = rabs(dir.x), ay = rabs(dir.y), az = rabs(dir.z);
    T p;
    V3<T> n;
    if (ax >= ay && ax >= az) {
        const T sg = dir.x >= T(0.0) ? T(1.0) : T(-1.0);
        p = h.x - sg * lp.x;
        n = v3(sg, T(0.0), T(0.0));
    } else if (ay >= az) {
        const T sg = dir.y >= T(0.0) ? T(1.0) : T(-1.0);
        p = h.y - sg * lp.y;
        n = v3(T(0.0), sg, T(0.0));
    } else {
        const T sg = dir.z >= T(0.0) ? T(1.0) : T(-1.0);
        p = h.z - sg * lp.z;
        n = v3(T(0.0), T(0.0), sg);
    }
    if (p <= T(0.0)) return false;
    *pen = p;
    *nl = n;
    return true;
}

// box_box_vertex_face (contact.rs:587-714): A's corners inside B (normal
// B's face along B->A, point pushed onto B's surface), then B's corners
// inside A (normal A's face flipped, point the corner itself); the four
// deepest, stable on ties.  The penalty pipeline's box_box (534-553) uses
// it first and falls back to the SAT manifold when it is empty.
template <class T>
void box_box_vertex_face(int ia, const Pose<T>& pa, V3<T> ha, int ib, const Pose<T>& pb, V3<T> hb, T fr, T margin,
                         T gap, std::vector<Contact<T>>& out) {
    static const int sgn[8][3] = {{-1, -1, -1}, {-1, -1, 1}, {-1, 1, -1}, {-1, 1, 1},
                                  {1, -1, -1},  {1, -1, 1},  {1, 1, -1},  {1, 1, 1}};
    const V3<T> dw = pa.p - pb.p;
    const V3<T> din_b = inverse_rotate(pb.q, dw);
    const V3<T> din_a = inverse_rotate(pa.q, -dw);
    const V3<T> ez = v3(T(0.0), T(0.0), T(1.0));
    const V3<T> dir_b = dot(din_b, din_b) > T(0.0) ? din_b : ez;
    const V3<T> dir_a = dot(din_a, din_a) > T(0.0) ? din_a : -ez;
    T cpen[16];
    V3<T> cpos[16], cnrm[16];
    int count = 0;
    for (int c = 0; c < 8; ++c) {
        const V3<T> la = v3(T(sgn[c][0]) * ha.x, T(sgn[c][1]) * ha.y, T(sgn[c][2]) * ha.z);
        const V3<T> wv = pa.p + rotate(pa.q, la);
        const V3<T> lb = inverse_rotate(pb.q, wv - pb.p);
        T pen;
        V3<T> nl;
        if (face_along_direction(lb, hb, dir_b, &pen, &nl)) {
            const V3<T> nw = rotate(pb.q, nl);
            cpen[count] = pen;
            cpos[count] = wv + nw * pen;
            cnrm[count] = nw;
            ++count;
        }
    }
    for (int c = 0; c < 8; ++c) {
        const V3<T> lb = v3(T(sgn[c][0]) * hb.x, T(sgn[c][1]) * hb.y, T(sgn[c][2]) * hb.z);
        const V3<T> wv = pb.p + rotate(pb.q, lb);
        const V3<T> la = inverse_rotate(pa.q, wv - pa.p);
        T pen;
        V3<T> nl;
        if (face_along_direction(la, ha, dir_a, &pen, &nl)) {
            cpen[count] = pen;
            cpos[count] = wv;
            cnrm[count] = -rotate(pa.q, nl);
            ++count;
        }
    }
    int order[16];
    for (int i = 0; i < 16; ++i) order[i] = i;
    for (int i = 1; i < count; ++i)
        for (int j = i; j > 0 && cpen[order[j]] > cpen[order[j - 1]]; --j) std::swap(order[j - 1], order[j]);
    const int take = count > 4 ? 4 : count;
    for (int i = 0; i < take; ++i)
        out.push_back({ia, ib, i, cpos[order[i]], cnrm[order[i]], cpen[order[i]] + margin, fr, gap});
}

// One-direction primitive (try_narrow_phase, contact.rs:2592-2806).  Returns
// false when (a, b) is not a primitive in this order.
template <class T>
bool try_narrow(int ia, const Geom<T>& ga, const Pose<T>& pa, int ib, const Geom<T>& gb, const Pose<T>& pb,
                T fr, T margin, T gap, std::vector<Contact<T>>& out, bool legacy = false) {
    if (ga.shape == S_SPHERE && gb.shape == S_PLANE) { sphere_plane(ia, pa, ga.size.x, fr, margin, gap, ib, pb, out); return true; }
    if (ga.shape == S_BOX && gb.shape == S_PLANE) { box_plane(ia, pa, ga.size, fr, margin, gap, ib, pb, out); return true; }
    if (ga.shape == S_CAPSULE && gb.shape == S_PLANE) { capsule_plane(ia, pa, ga.size.x, ga.size.y, fr, margin, gap, ib, pb, out); return true; }
    if (ga.shape == S_MESH && gb.shape == S_PLANE) {
        for (const CcdContact<T>& c : mesh_plane(pa, *ga.mesh_data, margin, pb))
            out.push_back({ia, ib, (int)out.size(), c.pos, c.normal, c.pen, fr, gap});
        return true;
    }
    if (ga.shape == S_SPHERE && gb.shape == S_MESH) {
        for (const CcdContact<T>& c : sphere_mesh(pa, ga.size.x, pb, *gb.mesh_data, margin))
            out.push_back({ia, ib, (int)out.size(), c.pos, c.normal, c.pen, fr, gap});
        return true;
    }
    // sphere_sphere (contact.rs:407-427), sphere_capsule (473-499),
    // capsule_capsule (1249-1272).
    if (ga.shape == S_SPHERE && gb.shape == S_SPHERE) {
        sphere_vs_sphere_at(ia, pa.p, ga.size.x, ib, pb.p, gb.size.x, fr, margin, gap, out);
        return true;
    }
    if (ga.shape == S_SPHERE && gb.shape == S_CAPSULE) {
        V3<T> b0, b1;
        capsule_segment(pb, gb.size.y, &b0, &b1);
        sphere_vs_sphere_at(ia, pa.p, ga.size.x, ib, closest_point_on_segment(pa.p, b0, b1), gb.size.x, fr, margin,
                            gap, out);
        return true;
    }
    if (ga.shape == S_CAPSULE && gb.shape == S_CAPSULE) {
        V3<T> a0, a1, b0, b1;
        capsule_segment(pa, ga.size.y, &a0, &a1);
        capsule_segment(pb, gb.size.y, &b0, &b1);
        std::pair<V3<T>, V3<T>> cp = closest_points_on_segments(a0, a1, b0, b1);
        sphere_vs_sphere_at(ia, cp.first, ga.size.x, ib, cp.second, gb.size.x, fr, margin, gap, out);
        return true;
    }
    if (ga.shape == S_BOX && gb.shape == S_BOX) {
        if (legacy) {  // NarrowPhaseMode::LegacyPenalty (contact.rs:2767-2769)
            box_box_vertex_face(ia, pa, ga.size, ib, pb, gb.size, fr, margin, gap, out);
            if (!out.empty()) return true;
        }
        box_box(ia, pa, ga.size, ib, pb, gb.size, fr, margin, gap, out);
        return true;
    }
    return false;
}

// dispatch_narrow_phase (contact.rs:2509-2590), FullManifold mode, or
// LegacyPenalty (the penalty pipeline's box-box) with `legacy`.
template <class T>
void narrow_phase(int ia, const Geom<T>& ga, const Pose<T>& pa, int ib, const Geom<T>& gb, const Pose<T>& pb,
                  std::vector<Contact<T>>& out, bool legacy = false) {
    out.clear();
    T fr = combine_friction(ga.friction, gb.friction);
    T margin = combine_max(ga.margin, gb.margin);
    T gap = combine_max(ga.gap, gb.gap);
    if (ga.shape == S_MESH && gb.shape == S_MESH) {  // ccd::route_pair (contact/ccd.rs:6-8)
        for (const CcdContact<T>& c : mesh_mesh(MeshShape<T>{pa, ga.mesh_data.get()}, MeshShape<T>{pb, gb.mesh_data.get()}, margin))
            out.push_back({ia, ib, (int)out.size(), c.pos, c.normal, c.pen, fr, gap});
        return;
    }
    if (try_narrow(ia, ga, pa, ib, gb, pb, fr, margin, gap, out, legacy)) return;
    if (try_narrow(ib, gb, pb, ia, ga, pa, fr, margin, gap, out, legacy)) {
        for (auto& c : out) {
            std::swap(c.geom_a, c.geom_b);
            c.normal = -c.normal;
        }
        return;
    }
    fail("unsupported shape pair");
}

// is_pair_supported (contact.rs:232-257) for the shapes restated here.
inline bool pair_supported(int a, int b) {
    auto one = [](int x, int y) {
        return (x == S_SPHERE && y == S_PLANE) || (x == S_BOX && y == S_PLANE) || (x == S_CAPSULE && y == S_PLANE) ||
               (x == S_SPHERE && y == S_SPHERE) || (x == S_SPHERE && y == S_CAPSULE) ||
               (x == S_CAPSULE && y == S_CAPSULE) || (x == S_BOX && y == S_BOX) || (x == S_MESH && y == S_PLANE) ||
               (x == S_SPHERE && y == S_MESH) || (x == S_MESH && y == S_MESH);
    };
    return one(a, b) || one(b, a);
}

// ---------------------------------------------------------------------------
// Free bodies (body.rs:22-47) and their constraint solve
// (solve_free_bodies_diag_mode, solver.rs:703-1257) and integration
// (world.rs:1705-1817, 3209-3286).
// ---------------------------------------------------------------------------

template <class T> struct Body {
    T mass = T(1.0);
    M3<T> inertia = m3identity<T>(), inv_inertia = m3identity<T>();
    V3<T> p = vzero<T>(), v = vzero<T>();
    Q<T> q = qidentity<T>();
    V3<T> w = vzero<T>();  // body frame
    bool has_max_lin = false, has_max_ang = false;
    T max_lin = T(0.0), max_ang = T(0.0);
    T gravity_scale = T(1.0);
};

// Body::new (body.rs:53-70): the inverse inertia is precomputed.
template <class T> Body<T> make_body(T mass, const M3<T>& inertia, V3<T> p, Q<T> q) {
    Body<T> b;
    b.mass = mass;
    b.inertia = inertia;
    if (!m3inverse(inertia, &b.inv_inertia)) fail("body inertia tensor must be invertible");
    b.p = p;
    b.q = q;
    return b;
}

// solid_box_inertia (geom.rs:942-950).
template <class T> M3<T> solid_box_inertia(T mass, V3<T> h) {
    T hx2 = h.x * h.x, hy2 = h.y * h.y, hz2 = h.z * h.z;
    return m3diag((mass / T(3.0)) * (hy2 + hz2), (mass / T(3.0)) * (hx2 + hz2), (mass / T(3.0)) * (hx2 + hy2));
}

template <class T> struct BodyDeriv {  // Deriv (world.rs:3209-3215)
    V3<T> dp, dv;
    Q<T> dq;
    V3<T> dw;
};

template <class T> inline V3<T> body_gravity(const Body<T>& b, V3<T> g) {
    return b.gravity_scale == T(1.0) ? g : g * b.gravity_scale;
}

// evaluate (world.rs:3223-3245).
template <class T> BodyDeriv<T> evaluate_body(const Body<T>& s, V3<T> gravity, V3<T> f, V3<T> tau) {
    BodyDeriv<T> d;
    d.dv = body_gravity(s, gravity) + f / s.mass;
    V3<T> iw = s.inertia * s.w;
    V3<T> gyro = -cross(s.w, iw);
    V3<T> tau_body = inverse_rotate(s.q, tau);
    d.dw = s.inv_inertia * (tau_body + gyro);
    d.dq = qderivative(s.q, s.w);
    d.dp = s.v;
    return d;
}

// clamp_velocity (world.rs:3273-3286).
template <class T> V3<T> clamp_velocity(V3<T> v, bool has_cap, T cap) {
    if (!has_cap) return v;
    if (cap <= T(0.0)) return vzero<T>();
    T speed = length(v);
    return speed > cap ? normalize(v) * cap : v;
}

template <class T> inline Q<T> qadd(Q<T> a, Q<T> b) { return {a.x + b.x, a.y + b.y, a.z + b.z, a.w + b.w}; }

// advance_all for one body (world.rs:3256-3271): origin + deriv * h.
template <class T> Body<T> advance_body(const Body<T>& s, const BodyDeriv<T>& d, T h) {
    Body<T> o = s;
    o.p = s.p + d.dp * h;
    o.v = s.v + d.dv * h;
    o.q = qadd(s.q, qscale(d.dq, h));
    o.w = s.w + d.dw * h;
    return o;
}

// One constraint row of the free-body solve (ConstraintRow, solver.rs:1863-1885);
// contact rows are always linear (RowGeom::Linear) under the pyramidal cone.
template <class T> struct BodyRow {
    V3<T> dir;
    int body_a, body_b;  // -1: static side
    V3<T> arm_a, arm_b;
    T reg = T(0.0), diag = T(0.0), bias = T(0.0);
};
template <class T> struct BodyDelta {  // BodyDelta: dv_lin_world, dw_ang_body
    V3<T> dv = vzero<T>(), dw = vzero<T>();
};

template <class T> struct FreeBodySolve {
    const std::vector<Body<T>>& bodies;
    std::vector<M3<T>> inv_i_world;
    std::vector<BodyRow<T>> rows;

    // row_body_diagonal (solver.rs:1892-1928), linear rows.
    T row_body_diagonal(const BodyRow<T>& r) const {
        T a = T(0.0);
        if (r.body_a >= 0) {
            V3<T> rc = cross(r.arm_a, r.dir);
            a += dot(rc, inv_i_world[r.body_a] * rc) + T(1.0) / bodies[r.body_a].mass;
        }
        if (r.body_b >= 0) {
            V3<T> rc = cross(r.arm_b, r.dir);
            a += dot(rc, inv_i_world[r.body_b] * rc) + T(1.0) / bodies[r.body_b].mass;
        }
        return a;
    }
    // row_residual (solver.rs:1980-2020), linear rows.
    T row_residual(const BodyRow<T>& r, const std::vector<BodyDelta<T>>& bd) const {
        T s = T(0.0);
        if (r.body_a >= 0) {
            V3<T> wa = rotate(bodies[r.body_a].q, bd[r.body_a].dw);
            V3<T> vp = bd[r.body_a].dv + cross(wa, r.arm_a);
            s += dot(vp, r.dir);
        }
        if (r.body_b >= 0) {
            V3<T> wb = rotate(bodies[r.body_b].q, bd[r.body_b].dw);
            V3<T> vp = bd[r.body_b].dv + cross(wb, r.arm_b);
            s -= dot(vp, r.dir);
        }
        return s;
    }
    // apply_impulse_delta (solver.rs:2022-2069), linear rows.
    void apply(const BodyRow<T>& r, T df, std::vector<BodyDelta<T>>& bd) const {
        if (df == T(0.0)) return;
        if (r.body_a >= 0) {
            int i = r.body_a;
            V3<T> f = r.dir * df;
            V3<T> tq = cross(r.arm_a, f);
            bd[i].dv = bd[i].dv + f / bodies[i].mass;
            bd[i].dw = bd[i].dw + inverse_rotate(bodies[i].q, inv_i_world[i] * tq);
        }
        if (r.body_b >= 0) {
            int i = r.body_b;
            V3<T> f = (-r.dir) * df;
            V3<T> tq = cross(r.arm_b, f);
            bd[i].dv = bd[i].dv + f / bodies[i].mass;
            bd[i].dw = bd[i].dw + inverse_rotate(bodies[i].q, inv_i_world[i] * tq);
        }
    }
    // normal_velocity_at_point (solver.rs:2131-2154).
    T current_velocity(const BodyRow<T>& r) const {
        V3<T> va = vzero<T>(), vb = vzero<T>();
        if (r.body_a >= 0) {
            const Body<T>& b = bodies[r.body_a];
            va = b.v + cross(rotate(b.q, b.w), r.arm_a);
        }
        if (r.body_b >= 0) {
            const Body<T>& b = bodies[r.body_b];
            vb = b.v + cross(rotate(b.q, b.w), r.arm_b);
        }
        return dot(va - vb, r.dir);
    }
    // free_step_velocity_contribution (solver.rs:4012-4036) with the
    // gravity-only free step: dw_body_free = 0.
    T free_step_velocity(const BodyRow<T>& r, const std::vector<V3<T>>& dv_free) const {
        T s = T(0.0);
        if (r.body_a >= 0) {
            V3<T> dw = rotate(bodies[r.body_a].q, vzero<T>());
            V3<T> vp = dv_free[r.body_a] + cross(dw, r.arm_a);
            s += dot(vp, r.dir);
        }
        if (r.body_b >= 0) {
            V3<T> dw = rotate(bodies[r.body_b].q, vzero<T>());
            V3<T> vp = dv_free[r.body_b] + cross(dw, r.arm_b);
            s -= dot(vp, r.dir);
        }
        return s;
    }
    // row_cross_response_into (solver.rs:1604-1618).
    T cross_response(const BodyRow<T>& r, const BodyRow<T>& applied) const {
        std::vector<BodyDelta<T>> scratch(bodies.size());
        apply(applied, T(1.0), scratch);
        return row_residual(r, scratch);
    }
    // pgs_step_non_negative (solver.rs:1448-1466).
    void step_non_negative(int ri, std::vector<T>& f, std::vector<BodyDelta<T>>& bd) const {
        const BodyRow<T>& r = rows[ri];
        T res = row_residual(r, bd) + r.reg * f[ri] + r.bias;
        T delta = -res / r.diag;
        T nf = f[ri] + delta;
        T pr = nf < T(0.0) ? T(0.0) : nf;
        delta = pr - f[ri];
        f[ri] = pr;
        apply(r, delta, bd);
    }
    // pgs_step_pyramidal_pair (solver.rs:1476-1520).
    void step_pair(int ri, T h01, std::vector<T>& f, std::vector<BodyDelta<T>>& bd) const {
        int rj = ri + 1;
        T res_i = row_residual(rows[ri], bd) + rows[ri].reg * f[ri] + rows[ri].bias;
        T res_j = row_residual(rows[rj], bd) + rows[rj].reg * f[rj] + rows[rj].bias;
        T h00 = rows[ri].diag, h11 = rows[rj].diag;
        T det = h00 * h11 - h01 * h01;
        if (det <= T(0.0)) {
            step_non_negative(ri, f, bd);
            step_non_negative(rj, f, bd);
            return;
        }
        T di = (-res_i * h11 + h01 * res_j) / det;
        T dj = (h01 * res_i - h00 * res_j) / det;
        T ni = f[ri] + di, nj = f[rj] + dj;
        T a, b;
        if (ni >= T(0.0) && nj >= T(0.0)) {
            a = ni;
            b = nj;
        } else if (ni < T(0.0) && nj < T(0.0)) {
            a = T(0.0);
            b = T(0.0);
        } else if (ni < T(0.0)) {
            a = T(0.0);
            b = rmax(f[rj] - res_j / h11, T(0.0));
        } else {
            a = rmax(f[ri] - res_i / h00, T(0.0));
            b = T(0.0);
        }
        T ai = a - f[ri], aj = b - f[rj];
        f[ri] = a;
        f[rj] = b;
        apply(rows[ri], ai, bd);
        apply(rows[rj], aj, bd);
    }
};

// NewtonSystem::solve (newton.rs:275-340) for the free-body system with one
// NonNegative projection per row (pyramidal condim 3): minimize
// ½ xᵀ H x + linᵀ x over x >= 0 from x = 0 with Newton directions from one
// Cholesky factor of H and the exact line search along the projected path
// (exact_pyramidal_line_search, newton.rs:435-556).  Fails on a non-PD H.
template <class T> struct NewtonNonNeg {
    const std::vector<T>& H;
    const std::vector<T>& lin;
    int n;

    std::vector<T> mat_vec(const std::vector<T>& v) const {  // newton.rs:603-612
        std::vector<T> out(n, T(0.0));
        for (int i = 0; i < n; ++i)
            for (int j = 0; j < n; ++j) out[i] += H[(size_t)i * n + j] * v[j];
        return out;
    }
    static T dotv(const std::vector<T>& a, const std::vector<T>& b) {  // Iterator::sum
        RSum<T> s;
        for (size_t i = 0; i < a.size(); ++i) s.add(a[i] * b[i]);
        return s.acc;
    }
    T cost(const std::vector<T>& x) const {  // quadratic_cost
        return T(0.5) * dotv(x, mat_vec(x)) + dotv(lin, x);
    }
    std::vector<T> gradient(const std::vector<T>& x) const {
        std::vector<T> g = mat_vec(x);
        for (int i = 0; i < n; ++i) g[i] += lin[i];
        return g;
    }
    static void project(std::vector<T>& x) {
        for (T& v : x)
            if (v < T(0.0)) v = T(0.0);
    }
    std::vector<T> affine(const std::vector<T>& x, const std::vector<T>& d, T alpha) const {
        std::vector<T> p(n, T(0.0));
        for (int i = 0; i < n; ++i) p[i] = x[i] + alpha * d[i];
        project(p);
        return p;
    }
    static void add_root(std::vector<T>& breaks, T value, T slope) {  // newton.rs:558-566
        if (slope == T(0.0)) return;
        T alpha = -value / slope;
        if (alpha > T(0.0) && alpha < T(1.0)) breaks.push_back(alpha);
    }
    std::pair<T, T> line_search(const std::vector<T>& x, const std::vector<T>& d, T current) const {
        std::vector<T> breaks{T(0.0), T(1.0)};
        for (int i = 0; i < n; ++i) add_root(breaks, x[i], d[i]);
        std::stable_sort(breaks.begin(), breaks.end(), [](T a, T b) { return a < b; });
        std::vector<T> dd;  // Vec::dedup_by: drop an element within 1e-7 of the last kept one
        for (T b : breaks)
            if (dd.empty() || !(rabs(b - dd.back()) <= T(1e-7))) dd.push_back(b);
        std::vector<T> cand = dd;
        for (size_t k = 0; k + 1 < dd.size(); ++k) {
            T lo = dd[k], hi = dd[k + 1];
            if (hi - lo <= T(1e-7)) continue;
            T mid = T(0.5) * (lo + hi);
            T eps = (hi - lo) * T(1e-4);
            T la = rmax(mid - eps, lo);
            T ra = rmin(mid + eps, hi);
            std::vector<T> L = affine(x, d, la), R = affine(x, d, ra);
            T span = ra - la;
            if (span <= T(0.0)) continue;
            T inv = T(1.0) / span;
            std::vector<T> slope(n);
            for (int i = 0; i < n; ++i) slope[i] = (R[i] - L[i]) * inv;
            std::vector<T> g = gradient(affine(x, d, mid));
            T den = dotv(slope, mat_vec(slope));
            T num = dotv(g, slope);
            cand.push_back(den > T(0.0) ? rmin(rmax(mid - num / den, lo), hi) : mid);
        }
        T best_a = T(0.0), best_c = current;
        for (T a : cand) {
            T c = cost(affine(x, d, a));
            if (c < best_c) {
                best_c = c;
                best_a = a;
            }
        }
        return {best_a, best_c};
    }
    std::vector<T> solve(uint32_t max_iterations, T tol) const {
        std::vector<T> x(n, T(0.0));
        if (max_iterations == 0) return x;
        std::vector<T> f;
        if (!cholesky(H, n, f)) fail("Newton free-body solve failed: Newton Hessian is not positive definite");
        T cur = cost(x);
        for (uint32_t it = 0; it < max_iterations; ++it) {
            std::vector<T> g = gradient(x);
            std::vector<T> ng(n);
            for (int i = 0; i < n; ++i) ng[i] = -g[i];
            std::vector<T> d = cholesky_solve(f, n, ng);
            if (dotv(g, d) >= T(0.0)) break;
            auto [alpha, next] = line_search(x, d, cur);
            if (alpha <= T(0.0) || cur - next <= tol * (T(1.0) + rabs(cur))) break;
            for (int i = 0; i < n; ++i) x[i] += alpha * d[i];
            project(x);
            cur = next;
        }
        return x;
    }
};

// solve_free_bodies (PGS, solver.rs:703-1257) and solve_free_bodies_newton
// (Newton, solver.rs:1295-1407: A from unit impulses, H = A + diag(reg)) for
// condim-3 contacts under the pyramidal cone: per-body world wrenches
// (force, torque).
template <class T>
std::vector<std::pair<V3<T>, V3<T>>> solve_free_bodies(const std::vector<Body<T>>& bodies,
                                                      const std::vector<Geom<T>>& geoms,
                                                      const std::vector<Contact<T>>& contacts, V3<T> gravity,
                                                      T dt, int cone, uint32_t iterations, bool newton = false) {
    const size_t nb = bodies.size();
    std::vector<std::pair<V3<T>, V3<T>>> wrench(nb, {vzero<T>(), vzero<T>()});
    if (contacts.empty() || dt <= T(0.0)) return wrench;
    if (cone != C_PYRAMIDAL) fail("elliptic cone is not restated for free bodies");
    std::vector<V3<T>> dv_free(nb);
    for (size_t i = 0; i < nb; ++i) dv_free[i] = body_gravity(bodies[i], gravity) * dt;
    FreeBodySolve<T> S{bodies, {}, {}};
    struct Block {
        int start;
        SolRef<T> solref;
        SolImp<T> solimp;
        T pen, mu;
        T h01 = T(0.0), h23 = T(0.0);
    };
    std::vector<Block> blocks;
    for (const Contact<T>& c : contacts) {
        const Geom<T>& ga = geoms[c.geom_a];
        const Geom<T>& gb = geoms[c.geom_b];
        int ba = ga.body, bb = gb.body;
        if (ba < 0 && bb < 0) continue;
        T pen = c.penetration - c.gap;
        if (pen <= T(0.0)) continue;
        int condim = ga.condim < gb.condim ? ga.condim : gb.condim;
        if (condim != 3) fail("condim other than 3 is not restated for free bodies");
        T mu = combine_friction(ga.friction, gb.friction);
        // contact_friction_axes without anisotropy (solver.rs:3441-3465)
        V3<T> t1 = tangent_basis(c.normal).first;
        V3<T> t2 = cross(c.normal, t1);
        V3<T> arm_a = ba >= 0 ? c.pos - bodies[ba].p : vzero<T>();
        V3<T> arm_b = bb >= 0 ? c.pos - bodies[bb].p : vzero<T>();
        int start = (int)S.rows.size();
        // contact_row_directions (solver.rs:3401-3417)
        const V3<T> dirs[4] = {c.normal + t1 * mu, c.normal - t1 * mu, c.normal + t2 * mu, c.normal - t2 * mu};
        for (int k = 0; k < 4; ++k) S.rows.push_back(BodyRow<T>{dirs[k], ba, bb, arm_a, arm_b});
        Block b;
        b.start = start;
        b.solref = combine_solref(ga.solref, gb.solref);
        b.solimp = combine_solimp(ga.solimp, gb.solimp);
        b.pen = pen;
        b.mu = mu;
        blocks.push_back(b);
    }
    const int nr = (int)S.rows.size();
    if (nr == 0) return wrench;
    S.inv_i_world.resize(nb);
    for (size_t i = 0; i < nb; ++i) {
        M3<T> r = to_mat3(bodies[i].q);
        S.inv_i_world[i] = (r * bodies[i].inv_inertia) * transpose(r);
    }
    for (Block& b : blocks) {
        for (int k = 0; k < 4; ++k) {
            BodyRow<T>& r = S.rows[b.start + k];
            T a_ii = S.row_body_diagonal(r);
            T d = impedance_at_position(-b.pen, T(0.0), b.solimp);
            // contact_regularization (solver.rs:1930-1959), pyramidal condim 3
            T reg = T(0.0);
            if (d > T(0.0)) {
                T tran = T(0.0);
                if (r.body_a >= 0) tran += T(1.0) / bodies[r.body_a].mass;
                if (r.body_b >= 0) tran += T(1.0) / bodies[r.body_b].mass;
                reg = contact_regularization_from_diag(tran * (T(1.0) + b.mu * b.mu), d, b.mu, true);
            }
            r.reg = reg;
            r.diag = a_ii + reg;
        }
        for (int k = 0; k < 4; ++k) {
            BodyRow<T>& r = S.rows[b.start + k];
            T vcur = S.current_velocity(r);
            T dvf = S.free_step_velocity(r, dv_free);
            T a_ref = reference_accel(-b.pen, vcur, safe_solref(b.solref, dt), b.solimp);
            r.bias = dvf - a_ref * dt;
        }
    }
    for (Block& b : blocks) {
        b.h01 = S.cross_response(S.rows[b.start], S.rows[b.start + 1]);
        b.h23 = S.cross_response(S.rows[b.start + 2], S.rows[b.start + 3]);
    }
    std::vector<T> f(nr, T(0.0));
    if (newton) {
        // solve_free_body_newton_impulses (solver.rs:1295-1351)
        std::vector<T> A((size_t)nr * nr, T(0.0));
        for (int j = 0; j < nr; ++j) {
            std::vector<BodyDelta<T>> resp(nb);
            S.apply(S.rows[j], T(1.0), resp);
            for (int i = 0; i < nr; ++i) A[(size_t)i * nr + j] = S.row_residual(S.rows[i], resp);
        }
        std::vector<T> H = A;
        for (int i = 0; i < nr; ++i) H[(size_t)i * nr + i] += S.rows[i].reg;
        std::vector<T> lin(nr);
        for (int i = 0; i < nr; ++i) lin[i] = S.rows[i].bias;
        NewtonNonNeg<T> sys{H, lin, nr};
        f = sys.solve(iterations < 1 ? 1u : iterations, T(1e-7));
    } else {
        std::vector<BodyDelta<T>> bd(nb);
        for (uint32_t it = 0; it < iterations; ++it) {
            for (const Block& b : blocks) {
                S.step_non_negative(b.start, f, bd);
                S.step_pair(b.start, b.h01, f, bd);
                S.step_pair(b.start + 2, b.h23, f, bd);
            }
        }
    }
    for (int ri = 0; ri < nr; ++ri) {
        T fi = f[ri];
        if (fi == T(0.0)) continue;
        const BodyRow<T>& r = S.rows[ri];
        V3<T> fa = r.dir * (fi / dt);
        if (r.body_a >= 0) {
            wrench[r.body_a].first = wrench[r.body_a].first + fa;
            wrench[r.body_a].second = wrench[r.body_a].second + cross(r.arm_a, fa);
        }
        if (r.body_b >= 0) {
            V3<T> fb = -fa;
            wrench[r.body_b].first = wrench[r.body_b].first + fb;
            wrench[r.body_b].second = wrench[r.body_b].second + cross(r.arm_b, fb);
        }
    }
    return wrench;
}

// ---------------------------------------------------------------------------
// World
// ---------------------------------------------------------------------------

template <class T> struct World {
    T dt = T(0.005);
    V3<T> gravity{T(0.0), T(0.0), T(-9.81)};
    T penetration_slop = T(0.0);
    int integrator = I_RK4;
    int mode = M_PENALTY;
    uint32_t iterations = 20;
    int cone = C_PYRAMIDAL;
    std::vector<Tree<T>> trees;
    std::vector<Body<T>> bodies;
    std::vector<Geom<T>> geoms;
    std::vector<std::shared_ptr<const ConvexMesh<T>>> meshes;
    std::vector<bool> self_collide;  // per tree; false => in disabled_self_collision
    // Broadphase proxies: fat AABB per geom (DynamicAabbTree leaf bounds).
    std::vector<Aabb<T>> fat;
    std::vector<bool> present;
    // The bound each proxy was last updated with (swept, margin included),
    // and newt's reinsert counter (world.rs:1209-1212; a first insertion
    // counts, as update_with_filter_and_tree returns true for it).
    std::vector<Aabb<T>> bound;
    uint64_t reinserts = 0;
    // Outputs of the last step.
    std::vector<std::pair<int, int>> last_pairs;
    std::vector<Contact<T>> last_contacts;
    std::vector<T> last_qacc;  // flattened over trees
    std::vector<AbaWs<T>> ws;
    // Sensors (World::sensors, world.rs:441; SensorBank, sensor.rs:380-410).
    struct SensorSpec {
        int kind, tree, link, body, geom, out;  // (tree, link) or body: the site's parent
        V3<T> off;
        Q<T> quat;
    };
    std::vector<SensorSpec> sensors;
    std::vector<T> sensordata;
    V3<T> magnetic_field{T(0.0), T(0.0), T(0.0)};  // World::magnetic_field (world.rs:119)

    int add_tree(const Tree<T>& t, bool collide) {
        trees.push_back(t);
        self_collide.push_back(collide);
        ws.emplace_back();
        return (int)trees.size() - 1;
    }
    int add_body(const Body<T>& b) {  // world.rs:950-954
        bodies.push_back(b);
        return (int)bodies.size() - 1;
    }
    int add_mesh(const ConvexMesh<T>& m) {  // World::meshes
        meshes.push_back(std::make_shared<const ConvexMesh<T>>(m));
        return (int)meshes.size() - 1;
    }
    int add_geom(Geom<T> g) {  // world.rs:974-979
        if (g.shape == S_MESH) {
            if (g.mesh < 0 || g.mesh >= (int)meshes.size()) fail("mesh geom names no mesh");
            g.mesh_data = meshes[g.mesh];
        }
        geoms.push_back(g);
        refresh_broadphase();
        return (int)geoms.size() - 1;
    }

    void tree_poses(std::vector<std::vector<Pose<T>>>& poses) const {
        poses.resize(trees.size());
        for (size_t t = 0; t < trees.size(); ++t) forward_kinematics(trees[t], poses[t]);
    }
    Pose<T> geom_pose(int gi, const std::vector<std::vector<Pose<T>>>& poses) const {
        const Geom<T>& g = geoms[gi];
        if (g.is_body()) return geom_world_pose(g, bodies[g.body].p, bodies[g.body].q);
        if (!g.is_link()) return geom_world_pose(g, vzero<T>(), qidentity<T>());
        const Pose<T>& lp = poses[g.tree][g.link];
        return geom_world_pose(g, lp.p, lp.q);
    }

    // populate_broadphase_tree + swept_geom_bound (world.rs:1254-1358).
    void populate_broadphase() {
        std::vector<std::vector<Pose<T>>> poses;
        tree_poses(poses);
        std::vector<std::vector<std::pair<V3<T>, V3<T>>>> vel(trees.size());
        for (size_t t = 0; t < trees.size(); ++t) link_world_velocities(trees[t], poses[t], vel[t]);
        fat.resize(geoms.size());
        bound.resize(geoms.size());
        present.resize(geoms.size(), false);
        for (size_t gi = 0; gi < geoms.size(); ++gi) {
            const Geom<T>& g = geoms[gi];
            Pose<T> start = geom_pose((int)gi, poses);
            Pose<T> end = start;
            T ang_speed = T(0.0);
            V3<T> pivot = start.p;
            if (g.is_link()) {
                V3<T> v = vel[g.tree][g.link].first, w = vel[g.tree][g.link].second;
                const Pose<T>& lp = poses[g.tree][g.link];
                end = geom_world_pose(g, lp.p + v * dt,
                                      integrate_body_angular_velocity(lp.q, inverse_rotate(lp.q, w), dt));
                ang_speed = length(w);
                pivot = lp.p;
            } else if (g.is_body()) {
                const Body<T>& b = bodies[g.body];
                end = geom_world_pose(g, b.p + b.v * dt, integrate_body_angular_velocity(b.q, b.w, dt));
                ang_speed = length(rotate(b.q, b.w));
                pivot = b.p;
            }
            Aabb<T> sb = geom_aabb(g, start), eb = geom_aabb(g, end);
            V3<T> bc = (sb.mn + sb.mx) * T(0.5);
            T br = length((sb.mx - sb.mn) * T(0.5));
            T pr = length(bc - pivot) + br;
            T sweep = (ang_speed > T(0.0) && std::isfinite(pr)) ? pr * ang_speed * dt : T(0.0);
            const Aabb<T> bd = expanded(expanded(aunion(sb, eb), sweep), g.margin);
            bound[gi] = bd;
            // update_with_filter_and_tree (broadphase.rs:320-346).
            if (!present[gi] || !contains(fat[gi], bd)) {
                fat[gi] = fatten(bd);
                present[gi] = true;
                ++reinserts;
            }
        }
    }
    void refresh_broadphase() { populate_broadphase(); }  // world.rs:1119-1136

    // The state-independent pair filters: different attachment, group/mask
    // (should_collide, broadphase.rs:16-18) and disabled self collision.
    bool static_pair_ok(int a, int b) const {
        const Geom<T>& ga = geoms[a];
        const Geom<T>& gb = geoms[b];
        if (!((ga.group & gb.mask) != 0 && (gb.group & ga.mask) != 0)) return false;
        if (ga.is_link() && gb.is_link() && ga.tree == gb.tree && !self_collide[ga.tree]) return false;
        // GeomAttach equality: Static, Body(i) or Link(tree, link).
        auto kind = [](const Geom<T>& g) { return g.is_link() ? 2 : g.is_body() ? 1 : 0; };
        bool same_att = kind(ga) == kind(gb) && (!ga.is_link() || (ga.tree == gb.tree && ga.link == gb.link)) &&
                        (!ga.is_body() || ga.body == gb.body);
        return !same_att;
    }

    // update_broadphase + compute_pairs (world.rs:1190-1226, broadphase.rs:384-455):
    // all (min,max) pairs whose fat bounds overlap and that pass the filters.
    std::vector<std::pair<int, int>> active_pairs() {
        populate_broadphase();
        std::vector<std::pair<int, int>> out;
        int n = (int)geoms.size();
        for (int a = 0; a < n; ++a)
            for (int b = a + 1; b < n; ++b) {
                if (!overlaps(fat[a], fat[b])) continue;
                if (!static_pair_ok(a, b)) continue;
                out.push_back({a, b});
            }
        return out;
    }

    // collect_contacts_full (world.rs:2476-2528).
    std::vector<Contact<T>> detect_contacts(const std::vector<std::pair<int, int>>& pairs) const {
        std::vector<std::vector<Pose<T>>> poses;
        tree_poses(poses);
        std::vector<Pose<T>> gp(geoms.size());
        for (size_t gi = 0; gi < geoms.size(); ++gi) gp[gi] = geom_pose((int)gi, poses);
        std::vector<Contact<T>> out, buf;
        for (auto [a, b] : pairs) {
            narrow_phase(a, geoms[a], gp[a], b, geoms[b], gp[b], buf);
            for (auto c : buf) {
                if (penetration_slop > T(0.0) && c.gap < penetration_slop) c.gap = penetration_slop;
                out.push_back(c);
            }
        }
        return out;
    }

    struct Comp {  // WorldJacobian (solver.rs:2754-2774): a tree link's or a free body's
        int tree = -1, link = -1, body = -1;
        std::vector<T> coeff;
        V3<T> wrench_linear, wrench_angular;  // a body's linear / angular Jacobian rows
    };
    struct Row {  // WorldContactRow (solver.rs:2746-2752)
        std::vector<Comp> comps;
        T reg = T(0.0), diag = T(0.0), bias = T(0.0);
    };

    // tree_body_invweight0 (solver.rs:3719-3754).
    std::pair<T, T> body_invweight0(int ti, int link, const std::vector<Pose<T>>& poses, const std::vector<T>& factor) const {
        const Tree<T>& t = trees[ti];
        int nv = t.nv();
        std::vector<V3<T>> jt, jr;
        point_jacobian(t, poses, link, vzero<T>(), jt, jr);
        T tr = T(0.0), rot = T(0.0);
        for (int axis = 0; axis < 3; ++axis) {
            std::vector<T> lin(nv), ang(nv);
            for (int s = 0; s < nv; ++s) {
                lin[s] = axis == 0 ? jt[s].x : axis == 1 ? jt[s].y : jt[s].z;
                ang[s] = axis == 0 ? jr[s].x : axis == 1 ? jr[s].y : jr[s].z;
            }
            std::vector<T> lr = cholesky_solve(factor, nv, lin), ar = cholesky_solve(factor, nv, ang);
            RSum<T> s1, s2;
            for (int s = 0; s < nv; ++s) s1.add(lin[s] * lr[s]);
            for (int s = 0; s < nv; ++s) s2.add(ang[s] * ar[s]);
            tr += s1.acc;
            rot += s2.acc;
        }
        return {tr / T(3.0), rot / T(3.0)};
    }

    // solve_tree_contacts (solver.rs:2805-3366), PGS or Newton, over the
    // contacts touching a tree; a free body on the other side enters the
    // rows as a body component (world_jacobian_for_side, 3537-3600).
    // Returns tree_qfrc per tree.
    // With `wrenches` / `normals`: also TreeContactSolution::tree_wrenches
    // (per tree and link, world force and torque at the COM) and
    // contact_normal_forces (per input contact), solver.rs:3334-3364; with
    // `body_wrenches`: TreeContactSolution::body_wrenches (per body, world
    // force and torque).
    std::vector<std::vector<T>> solve_tree_contacts(const std::vector<Contact<T>>& contacts, int tree_implicit /*-1 none,0,1*/,
                                                    std::vector<std::vector<std::pair<V3<T>, V3<T>>>>* wrenches = nullptr,
                                                    std::vector<T>* normals = nullptr,
                                                    std::vector<std::pair<V3<T>, V3<T>>>* body_wrenches = nullptr) const {
        if (body_wrenches) body_wrenches->assign(bodies.size(), {vzero<T>(), vzero<T>()});
        std::vector<std::vector<T>> qfrc(trees.size());
        for (size_t t = 0; t < trees.size(); ++t) qfrc[t].assign(trees[t].nv(), T(0.0));
        if (wrenches) {
            wrenches->assign(trees.size(), {});
            for (size_t t = 0; t < trees.size(); ++t) (*wrenches)[t].assign(trees[t].links.size(), {vzero<T>(), vzero<T>()});
        }
        if (normals) normals->assign(contacts.size(), T(0.0));
        if (contacts.empty() || dt <= T(0.0) || trees.empty()) return qfrc;
        if (cone != C_PYRAMIDAL) fail("elliptic cone is not restated");
        std::vector<std::vector<Pose<T>>> poses;
        tree_poses(poses);
        std::vector<M3<T>> inv_i_world(bodies.size());
        for (size_t i = 0; i < bodies.size(); ++i) {
            M3<T> r = to_mat3(bodies[i].q);
            inv_i_world[i] = (r * bodies[i].inv_inertia) * transpose(r);
        }
        std::vector<std::vector<T>> factors(trees.size()), dfactors(trees.size()), freev(trees.size());
        for (size_t t = 0; t < trees.size(); ++t) {
            int n = trees[t].nv();
            if (n == 0) continue;
            std::vector<T> rm = tree_implicit < 0 ? mass_matrix(trees[t]) : implicit_mass_matrix(trees[t], dt);
            if (!cholesky(rm, n, factors[t])) fail("tree contact solve failed: mass matrix is not positive definite");
            if (!cholesky(mass_matrix(trees[t]), n, dfactors[t])) fail("tree diagApprox failed");
            // tree_free_velocity_delta (solver.rs:3602-3622).
            Tree<T> ft = trees[t];
            ft.disable_penalty_limits = true;
            std::vector<Pose<T>> fp;
            forward_kinematics(ft, fp);
            std::vector<std::pair<V3<T>, V3<T>>> ext(ft.links.size(), {vzero<T>(), vzero<T>()});
            std::vector<T> qdd;
            AbaWs<T> tmp;
            aba_tree(ft, fp, gravity, ext, tree_implicit < 0 ? VI_EXPLICIT : (tree_implicit ? VI_JOINT_DAMPING_ACT : VI_JOINT_DAMPING), dt, tmp, qdd);
            freev[t].resize(n);
            for (int s = 0; s < n; ++s) freev[t][s] = qdd[s] * dt;
        }
        struct Block {
            int start;
            SolRef<T> solref;
            SolImp<T> solimp;
            T penetration, mu;
            int contact;
        };
        std::vector<Row> rows;
        std::vector<Block> blocks;
        for (size_t ci = 0; ci < contacts.size(); ++ci) {
            const Contact<T>& c = contacts[ci];
            const Geom<T>& ga = geoms[c.geom_a];
            const Geom<T>& gb = geoms[c.geom_b];
            if (!ga.is_link() && !gb.is_link()) continue;
            T pen = c.penetration - c.gap;
            if (pen <= T(0.0)) continue;
            SolRef<T> sr = combine_solref(ga.solref, gb.solref);
            SolImp<T> si = combine_solimp(ga.solimp, gb.solimp);
            int condim = ga.condim < gb.condim ? ga.condim : gb.condim;
            if (condim != 3) fail("condim other than 3 is not restated");
            T mu = combine_friction(ga.friction, gb.friction);
            // contact_friction_axes without anisotropy (solver.rs:3441-3465).
            V3<T> t1 = tangent_basis(c.normal).first;
            V3<T> t2 = cross(c.normal, t1);
            V3<T> dirs[4] = {c.normal + t1 * mu, c.normal - t1 * mu, c.normal + t2 * mu, c.normal - t2 * mu};
            int start = (int)rows.size();
            for (int k = 0; k < 4; ++k) {
                // world_jacobian_for_side (solver.rs:3537-3600), side A then
                // side B; a static side contributes no component.
                Row r;
                for (int side = 0; side < 2; ++side) {
                    const Geom<T>& g = side == 0 ? ga : gb;
                    T sign = side == 0 ? T(1.0) : T(-1.0);
                    if (g.is_body()) {
                        Comp cp;
                        cp.body = g.body;
                        cp.wrench_linear = dirs[k] * sign;
                        cp.wrench_angular = cross(c.pos - bodies[g.body].p, dirs[k]) * sign;
                        r.comps.push_back(cp);
                        continue;
                    }
                    if (!g.is_link()) continue;
                    const Pose<T>& lp = poses[g.tree][g.link];
                    V3<T> local = inverse_rotate(lp.q, c.pos - lp.p);
                    std::vector<V3<T>> jt, jr;
                    point_jacobian(trees[g.tree], poses[g.tree], g.link, local, jt, jr);
                    Comp cp;
                    cp.tree = g.tree;
                    cp.link = g.link;
                    cp.coeff.resize(jt.size());
                    for (size_t s = 0; s < jt.size(); ++s) cp.coeff[s] = dot(jt[s], dirs[k]) * sign;
                    cp.wrench_linear = dirs[k] * sign;
                    cp.wrench_angular = cross(c.pos - lp.p, dirs[k] * sign);
                    r.comps.push_back(cp);
                }
                rows.push_back(r);
            }
            blocks.push_back({start, sr, si, pen, mu, (int)ci});
        }
        if (rows.empty()) return qfrc;
        int nr = (int)rows.size();
        std::vector<T> A((size_t)nr * nr, T(0.0));
        for (int col = 0; col < nr; ++col) {
            // apply_world_impulse (solver.rs:3799-3836) with impulse 1.
            std::vector<std::vector<T>> delta(trees.size());
            for (size_t t = 0; t < trees.size(); ++t) delta[t].assign(trees[t].nv(), T(0.0));
            std::vector<std::pair<V3<T>, V3<T>>> bdelta(bodies.size(), {vzero<T>(), vzero<T>()});
            for (const Comp& cc : rows[col].comps) {
                if (cc.body >= 0) {
                    const Body<T>& bs = bodies[cc.body];
                    V3<T> force = cc.wrench_linear * T(1.0), torque = cc.wrench_angular * T(1.0);
                    V3<T> angular_velocity = inv_i_world[cc.body] * torque;
                    bdelta[cc.body].first = bdelta[cc.body].first + force / bs.mass;
                    bdelta[cc.body].second = bdelta[cc.body].second + inverse_rotate(bs.q, angular_velocity);
                    continue;
                }
                int nv = trees[cc.tree].nv();
                std::vector<T> gi(nv);
                for (int s = 0; s < nv; ++s) gi[s] = cc.coeff[s] * T(1.0);
                std::vector<T> resp = cholesky_solve(factors[cc.tree], nv, gi);
                for (int s = 0; s < nv; ++s) delta[cc.tree][s] += resp[s];
            }
            for (int ri = 0; ri < nr; ++ri) {
                // world_velocity_from_delta (solver.rs:3687-3714).
                RSum<T> outer;
                for (const Comp& rc : rows[ri].comps) {
                    if (rc.body >= 0) {
                        const Body<T>& bs = bodies[rc.body];
                        T bv = dot(rc.wrench_linear, bdelta[rc.body].first) +
                               dot(rc.wrench_angular, rotate(bs.q, bdelta[rc.body].second));
                        outer.add(bv + T(0.0));
                        continue;
                    }
                    RSum<T> s;
                    for (int k = 0; k < trees[rc.tree].nv(); ++k) s.add(rc.coeff[k] * delta[rc.tree][k]);
                    outer.add(T(0.0) + s.acc);
                }
                A[(size_t)ri * nr + col] = outer.acc;
            }
        }
        for (const Block& b : blocks) {
            for (int off = 0; off < 4; ++off) {
                int ri = b.start + off;
                Row& r = rows[ri];
                T imp = impedance_at_position(-b.penetration, T(0.0), b.solimp);
                // tree_contact_diag_approx (solver.rs:3756-3796).
                T tr = T(0.0), rot = T(0.0);
                for (const Comp& cp : r.comps) {
                    if (cp.body >= 0) {
                        if (!is_zero_vec(cp.wrench_linear)) tr += T(1.0) / bodies[cp.body].mass;
                        else rot += dot(cp.wrench_angular, inv_i_world[cp.body] * cp.wrench_angular);
                        continue;
                    }
                    auto iw = body_invweight0(cp.tree, cp.link, poses[cp.tree], dfactors[cp.tree]);
                    if (!is_zero_vec(cp.wrench_linear)) tr += iw.first; else rot += iw.second;
                }
                (void)rot;
                T diag_approx = tr * (T(1.0) + b.mu * b.mu);
                r.reg = contact_regularization_from_diag(diag_approx, imp, b.mu, true);
                r.diag = A[(size_t)ri * nr + ri] + r.reg;
                // world_current_velocity / world_free_velocity (solver.rs:3624-3685).
                RSum<T> vel, fvel;
                for (const Comp& cp : r.comps) {
                    if (cp.body >= 0) {
                        const Body<T>& bs = bodies[cp.body];
                        vel.add(dot(cp.wrench_linear, bs.v) + dot(cp.wrench_angular, rotate(bs.q, bs.w)) + T(0.0));
                        fvel.add(dot(cp.wrench_linear, body_gravity(bs, gravity) * dt) + T(0.0));
                        continue;
                    }
                    const Tree<T>& t = trees[cp.tree];
                    RSum<T> sv, sf;
                    for (int s = 0; s < t.nv(); ++s) sv.add(cp.coeff[s] * t.qdot[s]);
                    for (int s = 0; s < t.nv(); ++s) sf.add(cp.coeff[s] * freev[cp.tree][s]);
                    vel.add(T(0.0) + sv.acc);
                    fvel.add(T(0.0) + sf.acc);
                }
                T position = -b.penetration;
                T ref = reference_accel(position, vel.acc, safe_solref(b.solref, dt), b.solimp);
                r.bias = fvel.acc - ref * dt;
            }
        }
        std::vector<T> f(nr, T(0.0));
        if (mode == M_NEWTON) {
            // Newton on H = A + diag(R) with every pyramidal row projected
            // onto f >= 0 (solver.rs:3161-3241): a zero-friction set whose
            // H is only semidefinite gets a solver-only diagonal pivot.
            std::vector<T> H = A, lin(nr), tmp;
            for (int i = 0; i < nr; ++i) {
                H[(size_t)i * nr + i] += rows[i].reg;
                lin[i] = rows[i].bias;
            }
            bool frictionless = false;
            for (const Block& b : blocks) frictionless = frictionless || b.mu == T(0.0);
            if (frictionless && !cholesky(H, nr, tmp)) {
                T scale = T(0.0);
                for (T v : H) scale = rmax(scale, rabs(v));
                T pivot = rmax(scale * T(1.0e-6), T(1.0e-7));
                for (int i = 0; i < nr; ++i) H[(size_t)i * nr + i] += pivot;
            }
            NewtonNonNeg<T> ns{H, lin, nr};
            f = ns.solve(iterations > 0 ? iterations : 1, T(1e-7));
        } else {
            // PGS sweeps (solver.rs:3269-3331; world_pgs_non_negative 3851-3861).
            for (uint32_t it = 0; it < iterations; ++it)
                for (const Block& b : blocks)
                    for (int off = 0; off < 4; ++off) {
                        int i = b.start + off;
                        RSum<T> s;
                        for (int col = 0; col < nr; ++col) s.add(A[(size_t)i * nr + col] * f[col]);
                        T residual = (rows[i].bias + s.acc) + rows[i].reg * f[i];
                        f[i] = rmax(f[i] - residual / rows[i].diag, T(0.0));
                    }
        }
        // Force recovery (solver.rs:3334-3364).
        if (normals)
            for (const Block& b : blocks) {
                T sum = T(-0.0);  // Iterator::sum over the block's four rows
                for (int off = 0; off < 4; ++off) sum = sum + f[b.start + off];
                (*normals)[b.contact] = sum / dt;
            }
        for (int i = 0; i < nr; ++i) {
            T force = f[i] / dt;
            for (const Comp& cp : rows[i].comps) {
                if (cp.body >= 0) {
                    if (body_wrenches) {
                        auto& w = (*body_wrenches)[cp.body];
                        w.first = w.first + cp.wrench_linear * force;
                        w.second = w.second + cp.wrench_angular * force;
                    }
                    continue;
                }
                for (int s = 0; s < trees[cp.tree].nv(); ++s) qfrc[cp.tree][s] += cp.coeff[s] * force;
                if (wrenches) {
                    auto& w = (*wrenches)[cp.tree][cp.link];
                    w.first = w.first + cp.wrench_linear * force;
                    w.second = w.second + cp.wrench_angular * force;
                }
            }
        }
        return qfrc;
    }

    // solve_tree_limits (solver.rs:2509-2628) over the limit rows of
    // build_tree_solver_rows (solver.rs:2337-2373): one row per hinge/slide
    // outside its range, in link order (coefficient +1 below the range, -1
    // above it); A = E M⁻¹ Eᵀ from Cholesky solves; PGS with f >= 0.
    // Returns the per-dof force delta (zeros when no row is active).
    std::vector<T> solve_tree_limits(const Tree<T>& tr) const {
        const int nv = tr.nv();
        std::vector<T> qfrc(nv, T(0.0));
        if (nv == 0 || !(dt > T(0.0))) return qfrc;
        struct LRow {
            int slot;
            T coeff, violation;
            SolRef<T> solref;
            SolImp<T> solimp;
        };
        std::vector<LRow> rows;
        for (size_t li = 0; li < tr.links.size(); ++li) {
            const Link<T>& l = tr.links[li];
            if (!(l.joint == J_HINGE || l.joint == J_SLIDE) || !l.has_range) continue;
            LRow r;
            r.slot = tr.v_offset[li];
            r.solref.timeconst = l.limit_solref[0];
            r.solref.dampratio = l.limit_solref[1];
            r.solimp.dmin = l.limit_solimp[0];
            r.solimp.dmax = l.limit_solimp[1];
            r.solimp.width = l.limit_solimp[2];
            r.solimp.midpoint = l.limit_solimp[3];
            r.solimp.power = (uint32_t)l.limit_power;
            const T q = tr.q[tr.q_offset[li]];
            if (q < l.range_lo) {
                r.coeff = T(1.0);
                r.violation = l.range_lo - q;
            } else if (q > l.range_hi) {
                r.coeff = T(-1.0);
                r.violation = q - l.range_hi;
            } else {
                continue;
            }
            rows.push_back(r);
        }
        if (rows.empty()) return qfrc;
        std::vector<T> lfac;
        if (!cholesky(mass_matrix(tr), nv, lfac)) return qfrc;
        const int n = (int)rows.size();
        std::vector<std::vector<T>> dvec(n);
        for (int i = 0; i < n; ++i) {
            std::vector<T> e(nv, T(0.0));
            if (rows[i].coeff != T(0.0)) e[rows[i].slot] += rows[i].coeff;
            dvec[i] = cholesky_solve(lfac, nv, e);
        }
        // TreeRow::dot_response / dot_qdot (solver.rs:2214-2230): sums from 0.0.
        std::vector<T> a((size_t)n * n), bias(n), diag(n);
        for (int i = 0; i < n; ++i)
            for (int j = 0; j < n; ++j) a[(size_t)i * n + j] = T(0.0) + rows[i].coeff * dvec[j][rows[i].slot];
        for (int i = 0; i < n; ++i) {
            const T v_row = T(0.0) + rows[i].coeff * tr.qdot[rows[i].slot];
            const T a_ref = reference_accel(-rows[i].violation, v_row, safe_solref(rows[i].solref, dt), rows[i].solimp);
            bias[i] = -a_ref * dt;
            const T a_ii = a[(size_t)i * n + i];
            const T d = impedance_at_position(-rows[i].violation, T(0.0), rows[i].solimp);  // impedance()
            const T r = d > T(0.0) ? (T(1.0) - d) / d * a_ii : T(0.0);
            diag[i] = a_ii + r;
        }
        std::vector<T> f(n, T(0.0));
        for (uint32_t it = 0; it < iterations; ++it)
            for (int i = 0; i < n; ++i) {
                T ri = bias[i];
                for (int j = 0; j < n; ++j) ri += a[(size_t)i * n + j] * f[j];
                const T extra = (diag[i] - a[(size_t)i * n + i]) * f[i];
                ri += extra;
                const T delta = -ri / diag[i];
                const T nf = f[i] + delta;
                f[i] = nf < T(0.0) ? T(0.0) : nf;
            }
        for (int i = 0; i < n; ++i) {
            const T f_dt = f[i] / dt;
            if (rows[i].coeff != T(0.0)) qfrc[rows[i].slot] += rows[i].coeff * f_dt;
        }
        return qfrc;
    }

    // apply_contact_wrench (world.rs:2868-2949) for geoms with no friction
    // anisotropy and bodies with no rolling friction: the normal force
    // max(0, k pen_eff - c v_n) with (k, c) = solref_to_kc (geom.rs:137-145)
    // of the combined SolRef on the reduced mass; per tangent_basis axis the
    // damping force -c v_t clamped to +-mu f_n (clamp_symmetric,
    // world.rs:3195-3203); force and r x force on each body, Newton's third
    // law on the second.
    void apply_contact_wrench(std::vector<std::pair<V3<T>, V3<T>>>& ext, const std::vector<Body<T>>& st,
                              const Contact<T>& c) const {
        const Geom<T>& ga = geoms[c.geom_a];
        const Geom<T>& gb = geoms[c.geom_b];
        const V3<T> n = c.normal;
        T m_eff;
        if (ga.is_body() && gb.is_body()) {
            const T ma = st[ga.body].mass, mb = st[gb.body].mass;
            m_eff = ma * mb / (ma + mb);
        } else if (ga.is_body()) {
            m_eff = st[ga.body].mass;
        } else if (gb.is_body()) {
            m_eff = st[gb.body].mass;
        } else {
            return;
        }
        const SolRef<T> sr = combine_solref(ga.solref, gb.solref);
        T k, cd;
        if (sr.is_direct()) {
            k = -sr.timeconst;
            cd = -sr.dampratio;
        } else {
            const T tc = sr.timeconst;
            k = m_eff / (tc * tc);
            cd = T(2.0) * sr.dampratio * m_eff / tc;
        }
        // point_velocity (world.rs:3038-3048)
        auto point_velocity = [&](const Geom<T>& g, V3<T>* r) {
            if (!g.is_body()) {
                *r = vzero<T>();
                return vzero<T>();
            }
            const Body<T>& b = st[g.body];
            *r = c.pos - b.p;
            return b.v + cross(rotate(b.q, b.w), *r);
        };
        V3<T> ra, rb;
        const V3<T> va = point_velocity(ga, &ra);
        const V3<T> vb = point_velocity(gb, &rb);
        const V3<T> vrel = va - vb;
        const T vn = dot(vrel, n);
        const T pen_eff = c.penetration - c.gap;
        if (pen_eff <= T(0.0)) return;
        const T f_raw = k * pen_eff - cd * vn;
        const T fn = f_raw > T(0.0) ? f_raw : T(0.0);
        if (fn <= T(0.0)) return;
        const V3<T> t1 = tangent_basis(n).first;
        const V3<T> t2 = cross(n, t1);  // contact_friction_axes, solver.rs:3441-3465
        const V3<T> vt = vrel - n * vn;
        const T vt1 = dot(vt, t1), vt2 = dot(vt, t2);
        auto clamp_sym = [](T x, T cap) { return x > cap ? cap : (x < -cap ? -cap : x); };
        const T ft1 = clamp_sym(-cd * vt1, c.friction * fn);
        const T ft2 = clamp_sym(-cd * vt2, c.friction * fn);
        const V3<T> fa = (n * fn + t1 * ft1) + t2 * ft2;
        if (ga.is_body()) {
            auto& e = ext[ga.body];
            e.first = e.first + fa;
            e.second = e.second + cross(ra, fa);
            e.second = e.second + vzero<T>();  // rolling_drag_torque: no rolling friction
        }
        if (gb.is_body()) {
            const V3<T> fb = -fa;
            auto& e = ext[gb.body];
            e.first = e.first + fb;
            e.second = e.second + cross(rb, fb);
            e.second = e.second + vzero<T>();
        }
    }

    // tree_wrenches_from_pairs (world.rs:2570-2624) with
    // apply_tree_contact_wrench (2633-2741): penalty wrenches on the links of
    // tree `ti` at FK poses `lp`, from the step's pairs touching the tree
    // (legacy manifold); the other side contributes only its point velocity
    // and no reaction.  last_contacts keeps the contacts of the last call.
    std::vector<std::pair<V3<T>, V3<T>>> tree_penalty_wrenches(const Tree<T>& tr, int ti,
                                                               const std::vector<Pose<T>>& lp,
                                                               const std::vector<std::pair<int, int>>& tp) {
        std::vector<std::pair<V3<T>, V3<T>>> out(tr.links.size(), {vzero<T>(), vzero<T>()});
        last_contacts.clear();
        if (tp.empty()) return out;
        std::vector<std::pair<V3<T>, V3<T>>> vel;
        link_world_velocities(tr, lp, vel);
        auto pose_of = [&](const Geom<T>& g) {
            if (g.is_body()) return geom_world_pose(g, bodies[g.body].p, bodies[g.body].q);
            if (g.is_link() && g.tree == ti) return geom_world_pose(g, lp[g.link].p, lp[g.link].q);
            return geom_world_pose(g, vzero<T>(), qidentity<T>());
        };
        auto mass_of = [&](const Geom<T>& g) {
            if (g.is_link()) return tr.links[g.link].mass * (g.tree == ti ? T(1.0) : T(0.0));
            if (g.is_body()) return bodies[g.body].mass;
            return T(0.0);
        };
        auto point_velocity = [&](const Geom<T>& g, V3<T> at, V3<T>* r) {  // point_velocity_generic (2748-2775)
            *r = vzero<T>();
            if (g.is_body()) {
                const Body<T>& b = bodies[g.body];
                *r = at - b.p;
                return b.v + cross(rotate(b.q, b.w), *r);
            }
            if (g.is_link() && g.tree == ti) {
                *r = at - lp[g.link].p;
                return vel[g.link].first + cross(vel[g.link].second, *r);
            }
            return vzero<T>();
        };
        std::vector<Contact<T>> buf;
        for (auto [a, b] : tp) {
            narrow_phase(a, geoms[a], pose_of(geoms[a]), b, geoms[b], pose_of(geoms[b]), buf, true);
            for (const auto& c : buf) {
                last_contacts.push_back(c);
                const Geom<T>& ga = geoms[c.geom_a];
                const Geom<T>& gb = geoms[c.geom_b];
                const T ma = mass_of(ga), mb = mass_of(gb);
                T m_eff;
                if (ma > T(0.0) && mb > T(0.0)) m_eff = ma * mb / (ma + mb);
                else if (ma > T(0.0)) m_eff = ma;
                else if (mb > T(0.0)) m_eff = mb;
                else continue;
                const SolRef<T> sr = combine_solref(ga.solref, gb.solref);
                T k, cd;
                if (sr.is_direct()) {
                    k = -sr.timeconst;
                    cd = -sr.dampratio;
                } else {
                    const T tc = sr.timeconst;
                    k = m_eff / (tc * tc);
                    cd = T(2.0) * sr.dampratio * m_eff / tc;
                }
                V3<T> ra, rb;
                const V3<T> va = point_velocity(ga, c.pos, &ra);
                const V3<T> vb = point_velocity(gb, c.pos, &rb);
                const V3<T> n = c.normal;
                const V3<T> vrel = va - vb;
                const T vn = dot(vrel, n);
                const T pen_eff = penetration_slop > c.gap ? c.penetration - penetration_slop : c.penetration - c.gap;
                if (pen_eff <= T(0.0)) continue;
                const T f_raw = k * pen_eff - cd * vn;
                const T fn = f_raw > T(0.0) ? f_raw : T(0.0);
                if (fn <= T(0.0)) continue;
                const V3<T> t1 = tangent_basis(n).first;
                const V3<T> t2 = cross(n, t1);
                const V3<T> vt = vrel - n * vn;
                const T vt1 = dot(vt, t1), vt2 = dot(vt, t2);
                auto clamp_sym = [](T x, T cap) { return x > cap ? cap : (x < -cap ? -cap : x); };
                const T ft1 = clamp_sym(-cd * vt1, c.friction * fn);
                const T ft2 = clamp_sym(-cd * vt2, c.friction * fn);
                const V3<T> fa = (n * fn + t1 * ft1) + t2 * ft2;
                if (ga.is_link() && ga.tree == ti) {
                    out[ga.link].first = out[ga.link].first + fa;
                    out[ga.link].second = out[ga.link].second + cross(ra, fa);
                }
                if (gb.is_link() && gb.tree == ti) {
                    const V3<T> fb = -fa;
                    out[gb.link].first = out[gb.link].first + fb;
                    out[gb.link].second = out[gb.link].second + cross(rb, fb);
                }
            }
        }
        return out;
    }

    // compute_penalty_wrenches (world.rs:2281-2301) at body state `st`:
    // collect_contacts over the step's pairs with the legacy manifold
    // (world.rs:2434-2474), then apply_contact_wrench contact by contact.
    std::vector<std::pair<V3<T>, V3<T>>> penalty_wrenches(const std::vector<Body<T>>& st,
                                                          const std::vector<std::pair<int, int>>& pairs,
                                                          std::vector<Contact<T>>* keep) const {
        std::vector<Pose<T>> gp(geoms.size());
        for (size_t gi = 0; gi < geoms.size(); ++gi) {
            const Geom<T>& g = geoms[gi];
            gp[gi] = g.is_body() ? geom_world_pose(g, st[g.body].p, st[g.body].q)
                                 : geom_world_pose(g, vzero<T>(), qidentity<T>());
        }
        std::vector<std::pair<V3<T>, V3<T>>> out(st.size(), {vzero<T>(), vzero<T>()});
        std::vector<Contact<T>> buf;
        if (keep) keep->clear();
        for (auto [a, b] : pairs) {
            if (geoms[a].is_link() || geoms[b].is_link()) continue;
            narrow_phase(a, geoms[a], gp[a], b, geoms[b], gp[b], buf, true);
            for (auto c : buf) {
                if (penetration_slop > T(0.0) && c.gap < penetration_slop) c.gap = penetration_slop;
                if (keep) keep->push_back(c);
                apply_contact_wrench(out, st, c);
            }
        }
        return out;
    }

    // step_bodies / step_bodies_euler under SolverMode::Penalty
    // (world.rs:1705-1785, 1790-1817): the wrenches are sampled again at
    // every RK4 stage state.  last_contacts keeps the last stage's contacts.
    void step_bodies_penalty(const std::vector<std::pair<int, int>>& pairs) {
        const size_t n = bodies.size();
        const std::vector<Body<T>> s0 = bodies;
        std::vector<BodyDeriv<T>> k1(n), k2(n), k3(n), k4(n);
        auto stage = [&](const std::vector<Body<T>>& st, std::vector<BodyDeriv<T>>& kk, std::vector<Contact<T>>* keep) {
            const auto ext = penalty_wrenches(st, pairs, keep);
            for (size_t i = 0; i < n; ++i) kk[i] = evaluate_body(st[i], gravity, ext[i].first, ext[i].second);
        };
        stage(s0, k1, &last_contacts);
        for (size_t i = 0; i < n; ++i) {
            const V3<T> a1[2] = {k1[i].dv, k1[i].dw};
            for (int k = 0; k < 2; ++k) {
                last_qacc.push_back(a1[k].x);
                last_qacc.push_back(a1[k].y);
                last_qacc.push_back(a1[k].z);
            }
        }
        if (integrator == I_RK4) {
            std::vector<Body<T>> s(n);
            for (size_t i = 0; i < n; ++i) s[i] = advance_body(s0[i], k1[i], dt * T(0.5));
            stage(s, k2, &last_contacts);
            for (size_t i = 0; i < n; ++i) s[i] = advance_body(s0[i], k2[i], dt * T(0.5));
            stage(s, k3, &last_contacts);
            for (size_t i = 0; i < n; ++i) s[i] = advance_body(s0[i], k3[i], dt);
            stage(s, k4, &last_contacts);
            for (size_t i = 0; i < n; ++i) bodies[i] = rk4_combine(s0[i], k1[i], k2[i], k3[i], k4[i]);
        } else {
            for (size_t i = 0; i < n; ++i) bodies[i] = euler_combine(s0[i], k1[i]);
        }
    }

    // The RK4 combination and the semi-implicit Euler update of one body
    // (world.rs:1745-1784, 1806-1816).
    Body<T> rk4_combine(const Body<T>& s0, const BodyDeriv<T>& k1, const BodyDeriv<T>& k2, const BodyDeriv<T>& k3,
                        const BodyDeriv<T>& k4) const {
        Body<T> b = s0;
        const T sixth = T(1.0) / T(6.0);
        V3<T> dp = (k1.dp + k2.dp * T(2.0) + k3.dp * T(2.0) + k4.dp) * sixth;
        V3<T> dv = (k1.dv + k2.dv * T(2.0) + k3.dv * T(2.0) + k4.dv) * sixth;
        Q<T> dq = qscale(qadd(qadd(qadd(k1.dq, qscale(k2.dq, T(2.0))), qscale(k3.dq, T(2.0))), k4.dq), sixth);
        V3<T> dw = (k1.dw + k2.dw * T(2.0) + k3.dw * T(2.0) + k4.dw) * sixth;
        b.p = s0.p + dp * dt;
        b.v = clamp_velocity(s0.v + dv * dt, s0.has_max_lin, s0.max_lin);
        b.q = renormalize(qadd(s0.q, qscale(dq, dt)));
        b.w = clamp_velocity(s0.w + dw * dt, s0.has_max_ang, s0.max_ang);
        return b;
    }
    Body<T> euler_combine(const Body<T>& s0, const BodyDeriv<T>& k1) const {
        Body<T> b = s0;
        V3<T> v = s0.v + k1.dv * dt;
        V3<T> w = s0.w + k1.dw * dt;
        b.p = s0.p + v * dt;
        b.q = integrate_body_angular_velocity(s0.q, w, dt);
        b.v = clamp_velocity(v, s0.has_max_lin, s0.max_lin);
        b.w = clamp_velocity(w, s0.has_max_ang, s0.max_ang);
        return b;
    }

    // step_bodies (RK4, world.rs:1705-1785) and step_bodies_euler (Euler and
    // ImplicitFast, 1790-1817) with the solver wrenches held over the step.
    // last_qacc gets (dlinear_velocity, dangular_velocity_body) of the
    // first evaluation per body.
    void step_bodies(const std::vector<std::pair<V3<T>, V3<T>>>& ext) {
        for (size_t i = 0; i < bodies.size(); ++i) {
            Body<T>& b = bodies[i];
            const V3<T> f = ext[i].first, tq = ext[i].second;
            BodyDeriv<T> k1 = evaluate_body(b, gravity, f, tq);
            const V3<T> a1[2] = {k1.dv, k1.dw};
            for (int k = 0; k < 2; ++k) {
                last_qacc.push_back(a1[k].x);
                last_qacc.push_back(a1[k].y);
                last_qacc.push_back(a1[k].z);
            }
            if (integrator == I_RK4) {
                const Body<T> s0 = b;
                BodyDeriv<T> k2 = evaluate_body(advance_body(s0, k1, dt * T(0.5)), gravity, f, tq);
                BodyDeriv<T> k3 = evaluate_body(advance_body(s0, k2, dt * T(0.5)), gravity, f, tq);
                BodyDeriv<T> k4 = evaluate_body(advance_body(s0, k3, dt), gravity, f, tq);
                const T sixth = T(1.0) / T(6.0);
                V3<T> dp = (k1.dp + k2.dp * T(2.0) + k3.dp * T(2.0) + k4.dp) * sixth;
                V3<T> dv = (k1.dv + k2.dv * T(2.0) + k3.dv * T(2.0) + k4.dv) * sixth;
                Q<T> dq = qscale(qadd(qadd(qadd(k1.dq, qscale(k2.dq, T(2.0))), qscale(k3.dq, T(2.0))), k4.dq), sixth);
                V3<T> dw = (k1.dw + k2.dw * T(2.0) + k3.dw * T(2.0) + k4.dw) * sixth;
                b.p = s0.p + dp * dt;
                b.v = clamp_velocity(s0.v + dv * dt, s0.has_max_lin, s0.max_lin);
                b.q = renormalize(qadd(s0.q, qscale(dq, dt)));
                b.w = clamp_velocity(s0.w + dw * dt, s0.has_max_ang, s0.max_ang);
            } else {
                V3<T> v = b.v + k1.dv * dt;
                V3<T> w = b.w + k1.dw * dt;
                b.p = b.p + v * dt;
                b.q = integrate_body_angular_velocity(b.q, w, dt);
                b.v = clamp_velocity(v, b.has_max_lin, b.max_lin);
                b.w = clamp_velocity(w, b.has_max_ang, b.max_ang);
            }
        }
    }

    // World::step (world.rs:1459-1537) for trees.
    // ---------------- sensors (sensor.rs:446-702) ---------------------------
    enum { SK_JPOS = 0, SK_JVEL = 1, SK_FPOS = 4, SK_FQUAT = 5, SK_GYRO = 6, SK_ACCEL = 7, SK_TOUCH = 8, SK_FORCE = 9,
           SK_TORQUE = 10, SK_VELO = 13, SK_MAG = 14, SK_SUBCOM = 16, SK_FLINVEL = 17, SK_FANGVEL = 18 };
    static int sensor_dim(int k) {  // SensorKind::dim (sensor.rs:200-221)
        if (k == SK_JPOS || k == SK_JVEL || k == SK_TOUCH) return 1;
        if (k == SK_FQUAT) return 4;
        return 3;
    }
    void add_sensor(int kind, int tree, int link, int body, int geom, V3<T> off, Q<T> quat) {
        int o = (int)sensordata.size();
        sensors.push_back({kind, tree, link, body, geom, o, off, quat});
        for (int k = 0; k < sensor_dim(kind); ++k) sensordata.push_back(T(0.0));
    }
    // compute_link_va (sensor.rs:834-916): per link (v, a), body frame at the
    // COM; hinge/slide by propagate_single_dof (sensor.rs:935-962).
    std::vector<std::pair<Motion<T>, Motion<T>>> link_va(const Tree<T>& t, const std::vector<T>& qdd) const {
        size_t n = t.links.size();
        std::vector<std::pair<Motion<T>, Motion<T>>> out(n, {mzero<T>(), mzero<T>()});
        for (size_t i = 0; i < n; ++i) {
            const Link<T>& l = t.links[i];
            if (l.joint == J_FREE) {
                int o = t.v_offset[i];
                out[i] = {Motion<T>{{t.qdot[o], t.qdot[o + 1], t.qdot[o + 2]}, {t.qdot[o + 3], t.qdot[o + 4], t.qdot[o + 5]}},
                          Motion<T>{{qdd[o], qdd[o + 1], qdd[o + 2]}, {qdd[o + 3], qdd[o + 4], qdd[o + 5]}}};
            } else if (l.joint == J_FIXED) {
                GX<T> x = xup_fixed(l);
                auto p = l.parent >= 0 ? out[l.parent] : std::make_pair(mzero<T>(), mzero<T>());
                out[i] = {gx_motion(x, p.first), gx_motion(x, p.second)};
            } else if (l.joint == J_HINGE || l.joint == J_SLIDE) {
                GX<T> x = xup_joint(l, t.q[t.q_offset[i]]);
                Motion<T> sub = l.joint == J_HINGE ? Motion<T>{l.axis, cross(l.cj_pos, l.axis)} : Motion<T>{vzero<T>(), l.axis};
                auto p = out[l.parent];
                Motion<T> sq = mscale(sub, t.qdot[t.v_offset[i]]);
                Motion<T> sqq = mscale(sub, qdd[t.v_offset[i]]);
                Motion<T> v = madd(gx_motion(x, p.first), sq);
                Motion<T> c = cross_motion(v, sq);
                Motion<T> a = madd(madd(gx_motion(x, p.second), sqq), c);
                out[i] = {v, a};
            } else {
                fail("ball-joint sensors are not restated");
            }
        }
        return out;
    }
    // compute_link_wrenches (sensor.rs:964-1028).
    std::vector<Force<T>> link_wrenches(const Tree<T>& t, const std::vector<Pose<T>>& poses,
                                        const std::vector<std::pair<Motion<T>, Motion<T>>>& va,
                                        const std::vector<std::pair<V3<T>, V3<T>>>& ext) const {
        size_t n = t.links.size();
        std::vector<Force<T>> f(n, fzero<T>());
        for (size_t i = 0; i < n; ++i) {
            const Link<T>& l = t.links[i];
            G6<T> im = spatial_inertia6(l.mass, l.inertia);
            Force<T> ia = g6times(im, va[i].second), iv = g6times(im, va[i].first);
            Force<T> bias = cross_force(va[i].first, iv);
            V3<T> fw = ext[i].first + gravity * l.mass;
            Force<T> fe{inverse_rotate(poses[i].q, ext[i].second), inverse_rotate(poses[i].q, fw)};
            f[i] = fsub(fadd(ia, bias), fe);
        }
        for (size_t i = n; i-- > 1;) {
            const Link<T>& l = t.links[i];
            GX<T> x = l.joint == J_FREE ? gxidentity<T>() : l.joint == J_FIXED ? xup_fixed(l) : xup_joint(l, t.q[t.q_offset[i]]);
            f[l.parent] = fadd(f[l.parent], gx_tforce(x, f[i]));
        }
        return f;
    }
    // World::build_sensor_inputs (world.rs:1578-1696) + sensor::evaluate, for
    // the cases restated: the contact list empty, or (PGS) every contact on a
    // tree; the contacts are the step's start-of-step set under PGS/Newton and
    // the post-step detection under penalty (world.rs:1516-1522).
    void evaluate_sensors(const std::vector<Contact<T>>& contacts) {
        bool needs_qddot = false, needs_va = false, needs_f = false, touch = false;
        for (const SensorSpec& s : sensors) {
            touch = touch || s.kind == SK_TOUCH;
            needs_qddot = needs_qddot || s.kind == SK_ACCEL || s.kind == SK_FORCE || s.kind == SK_TORQUE;
            needs_va = needs_va || s.kind == SK_GYRO || s.kind == SK_VELO || s.kind == SK_FLINVEL || s.kind == SK_FANGVEL;
            needs_f = needs_f || s.kind == SK_FORCE || s.kind == SK_TORQUE;
        }
        needs_va = needs_va || needs_qddot;
        std::vector<std::vector<Pose<T>>> poses;
        tree_poses(poses);
        std::vector<std::vector<std::pair<V3<T>, V3<T>>>> twr(trees.size());
        for (size_t t = 0; t < trees.size(); ++t) twr[t].assign(trees[t].links.size(), {vzero<T>(), vzero<T>()});
        std::vector<T> normals(contacts.size(), T(0.0));
        // The contact solve only feeds the acceleration, force and touch
        // sensors (their inputs; the other kinds read the state alone).
        if (!contacts.empty() && (needs_qddot || touch)) {
            for (const Contact<T>& c : contacts)
                if (!(geoms[c.geom_a].is_link() || geoms[c.geom_b].is_link()) || geoms[c.geom_a].is_body() ||
                    geoms[c.geom_b].is_body())
                    fail("sensors with free-body contacts are not restated");
            if (mode != M_PGS) fail("sensors with penalty or Newton contacts are not restated");
            int ti = integrator == I_RK4 ? -1 : integrator == I_IMPLICITFAST ? 1 : 0;
            solve_tree_contacts(contacts, ti, &twr, &normals);
        }
        std::vector<std::vector<T>> qdd(trees.size());
        std::vector<std::vector<std::pair<Motion<T>, Motion<T>>>> va(trees.size());
        for (size_t t = 0; t < trees.size(); ++t) {
            if (needs_qddot) {
                AbaWs<T> tmp;
                aba_tree(trees[t], poses[t], gravity, twr[t], VI_EXPLICIT, dt, tmp, qdd[t]);
            } else {
                qdd[t].assign(trees[t].nv(), T(0.0));
            }
            if (needs_va) va[t] = link_va(trees[t], qdd[t]);
        }
        std::vector<std::vector<Force<T>>> lf(trees.size());
        if (needs_f)
            for (size_t t = 0; t < trees.size(); ++t) lf[t] = link_wrenches(trees[t], poses[t], va[t], twr[t]);
        // body accelerations from the (zero) body wrenches (sensor.rs:527-546)
        auto body_accel = [&](const Body<T>& b, V3<T>* acom, V3<T>* alpha) {
            const V3<T> fext = vzero<T>(), text = vzero<T>();
            *acom = gravity + fext / b.mass;
            V3<T> tau_body = inverse_rotate(b.q, text);
            V3<T> iw = b.inertia * b.w;
            V3<T> gyro = -cross(b.w, iw);
            *alpha = b.inv_inertia * (tau_body + gyro);
        };
        auto site_pose = [&](const SensorSpec& s, V3<T>* pp, Q<T>* pq) {  // sensor.rs:794-808
            if (s.body >= 0) { *pp = bodies[s.body].p; *pq = bodies[s.body].q; }
            else { *pp = poses[s.tree][s.link].p; *pq = poses[s.tree][s.link].q; }
        };
        auto site_vel = [&](const SensorSpec& s, V3<T>* lin, V3<T>* ang) {  // sensor.rs:705-726
            if (s.body >= 0) {
                const Body<T>& b = bodies[s.body];
                V3<T> angular = rotate(b.q, b.w);
                V3<T> offw = rotate(b.q, s.off);
                *lin = b.v + cross(angular, offw);
                *ang = angular;
            } else {
                const Motion<T>& v = va[s.tree][s.link].first;
                const Q<T>& o = poses[s.tree][s.link].q;
                *ang = rotate(o, v.ang);
                *lin = rotate(o, v.lin + cross(v.ang, s.off));
            }
        };
        auto put3 = [&](int o, V3<T> v) { sensordata[o] = v.x; sensordata[o + 1] = v.y; sensordata[o + 2] = v.z; };
        for (const SensorSpec& s : sensors) {
            const int o = s.out;
            switch (s.kind) {
                case SK_JPOS: sensordata[o] = trees[s.tree].q[trees[s.tree].q_offset[s.link]]; break;
                case SK_JVEL: sensordata[o] = trees[s.tree].qdot[trees[s.tree].v_offset[s.link]]; break;
                case SK_FPOS: { V3<T> p; Q<T> q; site_pose(s, &p, &q); put3(o, p + rotate(q, s.off)); break; }
                case SK_FQUAT: {
                    V3<T> p; Q<T> q; site_pose(s, &p, &q);
                    Q<T> r = q * s.quat;
                    sensordata[o] = r.x; sensordata[o + 1] = r.y; sensordata[o + 2] = r.z; sensordata[o + 3] = r.w;
                    break;
                }
                case SK_GYRO: {  // parent_body_velocities (sensor.rs:814-832)
                    V3<T> w = s.body >= 0 ? bodies[s.body].w : va[s.tree][s.link].first.ang;
                    put3(o, inverse_rotate(s.quat, w));
                    break;
                }
                case SK_ACCEL: {  // accelerometer_reading (sensor.rs:1030-1088)
                    V3<T> pp, omega_world, alpha_world, a_com_world;
                    Q<T> pq;
                    site_pose(s, &pp, &pq);
                    if (s.body >= 0) {
                        const Body<T>& b = bodies[s.body];
                        V3<T> ac, alpha_body;
                        body_accel(b, &ac, &alpha_body);
                        omega_world = rotate(b.q, b.w);
                        alpha_world = rotate(b.q, alpha_body);
                        a_com_world = ac;
                    } else {
                        const Motion<T>& v = va[s.tree][s.link].first;
                        const Motion<T>& a = va[s.tree][s.link].second;
                        omega_world = rotate(pq, v.ang);
                        alpha_world = rotate(pq, a.ang);
                        a_com_world = rotate(pq, a.lin + cross(v.ang, v.lin));
                    }
                    V3<T> r = rotate(pq, s.off);
                    V3<T> a_site = (a_com_world + cross(alpha_world, r)) + cross(omega_world, cross(omega_world, r));
                    V3<T> a_proper = a_site - gravity;
                    put3(o, inverse_rotate(pq * s.quat, a_proper));
                    break;
                }
                case SK_TOUCH: {  // touch_reading (sensor.rs:1103-1112)
                    T sum = T(0.0);
                    for (size_t c = 0; c < contacts.size(); ++c)
                        if (contacts[c].geom_a == s.geom || contacts[c].geom_b == s.geom) sum += rabs(normals[c]);
                    sensordata[o] = sum;
                    break;
                }
                case SK_FORCE: put3(o, lf[s.tree][s.link].lin); break;
                case SK_TORQUE: {
                    const Force<T>& f = lf[s.tree][s.link];
                    put3(o, f.tq - cross(trees[s.tree].links[s.link].cj_pos, f.lin));
                    break;
                }
                case SK_VELO: {
                    V3<T> lin, ang, pp; Q<T> pq;
                    site_vel(s, &lin, &ang);
                    site_pose(s, &pp, &pq);
                    put3(o, inverse_rotate(pq * s.quat, lin));
                    break;
                }
                case SK_MAG: { V3<T> pp; Q<T> pq; site_pose(s, &pp, &pq); put3(o, inverse_rotate(pq * s.quat, magnetic_field)); break; }
                case SK_SUBCOM: {  // subtree_com (sensor.rs:728-757)
                    const Tree<T>& t = trees[s.tree];
                    V3<T> weighted = vzero<T>();
                    T mass = T(0.0);
                    for (size_t li = 0; li < t.links.size(); ++li) {
                        int cur = (int)li;
                        bool inc = false;
                        while (cur >= 0) {
                            if (cur == s.link) { inc = true; break; }
                            cur = t.links[cur].parent;
                        }
                        if (inc) {
                            weighted = weighted + poses[s.tree][li].p * t.links[li].mass;
                            mass += t.links[li].mass;
                        }
                    }
                    put3(o, mass == T(0.0) ? vzero<T>() : weighted / mass);
                    break;
                }
                case SK_FLINVEL: case SK_FANGVEL: {
                    V3<T> lin, ang;
                    site_vel(s, &lin, &ang);
                    put3(o, s.kind == SK_FLINVEL ? lin : ang);
                    break;
                }
                default: fail("sensor kind not restated");
            }
        }
    }

    void step() {
        if (iterations == 0) fail("solver iterations must be >= 1");
        if (mode == M_NEWTON && cone == C_ELLIPTIC) fail("newton + elliptic");
        // assert_pairs_supported over auto_pairs (world.rs:874-898, 930-947, 1042-1069).
        for (size_t a = 0; a < geoms.size(); ++a)
            for (size_t b = a + 1; b < geoms.size(); ++b)
                if (static_pair_ok((int)a, (int)b) && !pair_supported(geoms[a].shape, geoms[b].shape))
                    fail("unsupported pair");
        std::vector<std::pair<int, int>> pairs = active_pairs();
        last_pairs = pairs;
        bool constraint = mode == M_PGS || mode == M_NEWTON;
        std::vector<std::vector<T>> tq;
        std::vector<std::pair<V3<T>, V3<T>>> bw(bodies.size(), {vzero<T>(), vzero<T>()});
        last_contacts.clear();
        if (constraint) {
            std::vector<Contact<T>> contacts = detect_contacts(pairs);
            // compute_tree_contact_solution (world.rs:2154-2173) takes the
            // contacts touching a tree, compute_solver_wrenches (2097-2149)
            // the rest, and adds the tree solve's body wrenches to its own.
            std::vector<Contact<T>> tc, fc;
            for (auto& c : contacts) {
                const Geom<T>& ga = geoms[c.geom_a];
                const Geom<T>& gb = geoms[c.geom_b];
                if (ga.is_link() || gb.is_link()) tc.push_back(c);
                else fc.push_back(c);
            }
            last_contacts = contacts;
            int ti = integrator == I_RK4 ? -1 : integrator == I_IMPLICITFAST ? 1 : 0;
            std::vector<std::pair<V3<T>, V3<T>>> tbw;
            tq = solve_tree_contacts(tc, ti, nullptr, nullptr, &tbw);
            bw = solve_free_bodies(bodies, geoms, fc, gravity, dt, cone, iterations, mode == M_NEWTON);
            if (!trees.empty())
                for (size_t i = 0; i < bodies.size(); ++i) {
                    bw[i].first = bw[i].first + tbw[i].first;
                    bw[i].second = bw[i].second + tbw[i].second;
                }
        } else {
            if (!pairs.empty() && !trees.empty() && !bodies.empty())
                fail("penalty contacts between trees and bodies are not restated");
        }
        last_qacc.clear();
        for (size_t t = 0; t < trees.size(); ++t) {
            Tree<T>& tr = trees[t];
            std::vector<T> delta;
            bool prior = tr.disable_penalty_limits;
            if (constraint) {
                if (mode == M_NEWTON)
                    for (auto& l : tr.links)
                        if (l.has_range) fail("Newton joint-limit rows are not restated");
                delta = solve_tree_limits(tr);
                for (int s = 0; s < tr.nv(); ++s) delta[s] += tq[t][s];
                for (int s = 0; s < tr.nv(); ++s) tr.qfrc_applied[s] += delta[s];
                tr.disable_penalty_limits = true;
            }
            std::vector<std::pair<int, int>> tp;  // step_trees: the pairs touching this tree (penalty)
            if (!constraint)
                for (auto [a, b] : pairs)
                    if ((geoms[a].is_link() && geoms[a].tree == (int)t) || (geoms[b].is_link() && geoms[b].tree == (int)t))
                        tp.push_back({a, b});
            WrenchFn<T> wf = [&](const Tree<T>& cur, const std::vector<Pose<T>>& lp) {
                return tree_penalty_wrenches(cur, (int)t, lp, tp);
            };
            const WrenchFn<T>* wfp = tp.empty() ? nullptr : &wf;
            std::vector<T> qacc;
            if (integrator == I_RK4) {
                rk4_step(tr, gravity, dt, ws[t], &qacc, wfp);
                last_qacc.insert(last_qacc.end(), qacc.begin(), qacc.end());
            } else {
                euler_step(tr, gravity, dt, integrator == I_IMPLICITFAST, ws[t], qacc, wfp);
                last_qacc.insert(last_qacc.end(), qacc.begin(), qacc.end());
            }
            for (size_t s = 0; s < delta.size(); ++s) tr.qfrc_applied[s] -= delta[s];
            tr.disable_penalty_limits = prior;
        }
        if (!constraint && !pairs.empty() && trees.empty()) step_bodies_penalty(pairs);
        else step_bodies(bw);
        // sensors on the post-step state (world.rs:1516-1522): the
        // start-of-step contacts under PGS/Newton, a fresh detection under
        // penalty.
        if (!sensors.empty()) evaluate_sensors(constraint ? last_contacts : detect_contacts(pairs));
        refresh_broadphase();
    }
};

}  // namespace oracle
