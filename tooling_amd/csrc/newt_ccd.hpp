// newt_ccd.hpp — convex-mesh contacts for the lane-per-world step, FP64 in
// newt's operation order:
//   mesh_plane   (newt/src/contact.rs:1424-1510)
//   sphere_mesh  (newt/src/contact.rs:1720-1776)
//   mesh-mesh    ccd_convex_contacts (newt/src/contact/ccd.rs:2071-2092):
//                GJK to an enclosing simplex (442-546), distance GJK on a
//                separated pair (696-975), EPA with newt's deterministic face
//                order and caps of 128 vertices / 512 faces (1703-1918), the
//                fallback axis search when EPA does not converge (1106-1183)
//                and the face / edge manifold (1185-1701).
// One lane runs one pair.  The EPA polytope and the manifold's polygons live
// in the world's FP64 / integer workspace (dw / iw, element k at k·W + world),
// so the kernel's private frame stays small.
#pragma once

#include "newt_math.hpp"
#include "newt_types.hpp"

namespace newt {

// Element i of a per-world array laid out [i][W] (world w's slice).
template <class T> struct Lane {
    T* p;
    int W, w;
    NEWT_HD T& operator[](int i) const { return p[(size_t)i * W + w]; }
};

namespace ccd {

// The pair-level routines stay out of line: one mesh pair per lane call, and
// the step kernels that never meet a mesh keep their register budget.
#define NEWT_CCD inline __host__ __device__ __noinline__

constexpr double kSq = 1.0e-12 * 1.0e-12;  // CCD_DEGENERATE_SQUARED (= CCD_DISTANCE_FLOOR)
constexpr int kMaxVerts = 128, kMaxFaces = 512, kMaxEdges = 1024;
constexpr int kMultiCap = 16, kClipCap = 32;  // MULTI_FEATURE_CAP, MULTI_CLIP_CAP

NEWT_HD double lsq(V3<double> a) { return dot3(a, a); }
NEWT_HD V3<double> ex() { return mk3(1.0, 0.0, 0.0); }

// f32::total_cmp restricted to non-NaN values: -0.0 before +0.0.
NEWT_HD int total_cmp(double a, double b) {
    if (a < b) return -1;
    if (a > b) return 1;
    const bool sa = __builtin_signbit(a), sb = __builtin_signbit(b);
    if (sa && !sb) return -1;
    if (!sa && sb) return 1;
    return 0;
}
NEWT_HD bool lex_precedes(V3<double> a, V3<double> b) {  // contact.rs:64-74
    int c = total_cmp(a.x, b.x);
    if (c != 0) return c < 0;
    c = total_cmp(a.y, b.y);
    if (c != 0) return c < 0;
    return total_cmp(a.z, b.z) < 0;
}
NEWT_HD bool same3(V3<double> a, V3<double> b) { return a.x == b.x && a.y == b.y && a.z == b.z; }
NEWT_HD double clamp01(double x) { return x < 0.0 ? 0.0 : (x > 1.0 ? 1.0 : x); }

// closest_point_on_triangle (contact.rs:2265-2306).
NEWT_HD V3<double> closest_on_triangle(V3<double> p, V3<double> a, V3<double> b, V3<double> c) {
    const V3<double> ab = b - a, ac = c - a, ap = p - a;
    const double d1 = dot3(ab, ap), d2 = dot3(ac, ap);
    if (d1 <= 0.0 && d2 <= 0.0) return a;
    const V3<double> bp = p - b;
    const double d3 = dot3(ab, bp), d4 = dot3(ac, bp);
    if (d3 >= 0.0 && d4 <= d3) return b;
    const double vc = d1 * d4 - d3 * d2;
    if (vc <= 0.0 && d1 >= 0.0 && d3 <= 0.0) return a + ab * (d1 / (d1 - d3));
    const V3<double> cp = p - c;
    const double d5 = dot3(ab, cp), d6 = dot3(ac, cp);
    if (d6 >= 0.0 && d5 <= d6) return c;
    const double vb = d5 * d2 - d1 * d6;
    if (vb <= 0.0 && d2 >= 0.0 && d6 <= 0.0) return a + ac * (d2 / (d2 - d6));
    const double va = d3 * d6 - d5 * d4;
    if (va <= 0.0 && (d4 - d3) >= 0.0 && (d5 - d6) >= 0.0) return b + (c - b) * ((d4 - d3) / ((d4 - d3) + (d5 - d6)));
    const double den = 1.0 / ((va + vb) + vc);
    return (a + ab * (vb * den)) + ac * (vc * den);
}

// A convex mesh at a world pose (CcdShape::Mesh).
struct Mesh {
    const double* v;  // xyz per vertex, mesh-local
    const int* f;     // three vertex indices per face
    int nv, nf;
    V3<double> p;
    Qt<double> q;
    NEWT_HD V3<double> local(int i) const { return mk3(v[3 * i], v[3 * i + 1], v[3 * i + 2]); }
    NEWT_HD V3<double> world(int i) const { return p + qrot(q, local(i)); }
    NEWT_HD V3<double> world_of(V3<double> l) const { return p + qrot(q, l); }
    // vertices_extent (ccd.rs:189-200)
    NEWT_HD double extent() const {
        V3<double> mn = local(0), mx = mn;
        for (int i = 1; i < nv; ++i) {
            const V3<double> x = local(i);
            mn = mk3(nmin(mn.x, x.x), nmin(mn.y, x.y), nmin(mn.z, x.z));
            mx = mk3(nmax(mx.x, x.x), nmax(mx.y, x.y), nmax(mx.z, x.z));
        }
        return len3(mx - mn);
    }
    // support_vertices (ccd.rs:202-234) on a mesh-local direction.
    NEWT_HD V3<double> support_local(V3<double> dir, bool* tied) const {
        *tied = false;
        const double dl = len3(dir);
        if (dl <= 1.0e-20) return local(0);
        const V3<double> d = dir / dl;
        const double tie = extent() * 1.0e-7;
        double best = -__builtin_inf();
        for (int i = 0; i < nv; ++i) best = nmax(best, dot3(local(i), d));
        V3<double> sel = z3<double>();
        int count = 0;
        for (int i = 0; i < nv; ++i) {
            const V3<double> c = local(i);
            if (nabs(dot3(c, d) - best) <= tie) {
                if (count == 0 || lex_precedes(c, sel)) sel = c;
                ++count;
            }
        }
        *tied = count >= 2;
        return sel;
    }
    // support_with_tie (ccd.rs:125-174) through support_vertices_transformed.
    NEWT_HD V3<double> support(V3<double> dir, bool* tied) const {
        if (!(lsq(dir) > 1.0e-20)) dir = ex();
        return p + qrot(q, support_local(qrot_inv(q, dir), tied));
    }
    // centered_support (ccd.rs:176-186) with support_feature_centroid
    // (236-306): the mean of the tied support vertices, distinct points taken
    // in lexicographic order times their multiplicity.
    NEWT_HD bool centered(V3<double> dir, V3<double>* out) const {
        const V3<double> ld = qrot_inv(q, dir);
        const double dl = len3(ld);
        if (dl <= 1.0e-20) return false;
        const V3<double> d = ld / dl;
        const double tie = extent() * 1.0e-7;
        double best = -__builtin_inf();
        for (int i = 0; i < nv; ++i) best = nmax(best, dot3(local(i), d));
        int tied = 0;
        for (int i = 0; i < nv; ++i)
            if (nabs(dot3(local(i), d) - best) <= tie) ++tied;
        double sx = 0.0, sy = 0.0, sz = 0.0;
        bool have_prev = false;
        V3<double> prev = z3<double>();
        for (int k = 0; k < tied; ++k) {
            bool found = false;
            V3<double> next = z3<double>();
            for (int i = 0; i < nv; ++i) {
                const V3<double> c = local(i);
                if (!(nabs(dot3(c, d) - best) <= tie)) continue;
                if (have_prev && !lex_precedes(prev, c)) continue;
                if (!found || lex_precedes(c, next)) {
                    next = c;
                    found = true;
                }
            }
            if (!found) break;
            int count = 0;
            for (int i = 0; i < nv; ++i)
                if (same3(local(i), next)) ++count;
            sx += next.x * (double)count;
            sy += next.y * (double)count;
            sz += next.z * (double)count;
            prev = next;
            have_prev = true;
        }
        if (tied < 2) return false;
        const double n = (double)tied;
        *out = p + qrot(q, mk3(sx / n, sy / n, sz / n));
        return true;
    }
};

struct Vtx {  // CcdVertex (ccd.rs:38-45)
    V3<double> m, a, b;
    bool ta, tb;
};

NEWT_HD bool vtx_precedes(const Vtx& a, const Vtx& b) {  // ccd.rs:322-334
    if (!same3(a.m, b.m)) return lex_precedes(a.m, b.m);
    if (!same3(a.a, b.a)) return lex_precedes(a.a, b.a);
    if (!same3(a.b, b.b)) return lex_precedes(a.b, b.b);
    return false;
}
NEWT_HD void sort_vtx(Vtx* v, int n) {  // ccd.rs:336-344
    for (int i = 1; i < n; ++i)
        for (int p = i; p > 0 && vtx_precedes(v[p], v[p - 1]); --p) {
            const Vtx t = v[p];
            v[p] = v[p - 1];
            v[p - 1] = t;
        }
}

// ccd_support (ccd.rs:356-371), normalised by the pair extent when ext > 0
// (ccd_support_normalized, 373-382).
NEWT_HD Vtx support(const Mesh& a, const Mesh& b, V3<double> dir, double ext) {
    if (!(lsq(dir) > kSq)) dir = ex();
    Vtx r;
    r.a = a.support(dir, &r.ta);
    r.b = b.support(-dir, &r.tb);
    r.m = r.a - r.b;
    if (ext > 0.0) r.m = r.m / ext;
    return r;
}
NEWT_HD V3<double> triple(V3<double> a, V3<double> b, V3<double> c) { return b * dot3(a, c) - a * dot3(b, c); }

struct Simplex {
    Vtx p[4];
    int len;
    NEWT_HD void set1(const Vtx& a) { p[0] = a; len = 1; }
    NEWT_HD void set2(const Vtx& a, const Vtx& b) { p[0] = a; p[1] = b; len = 2; }
    NEWT_HD void set3(const Vtx& a, const Vtx& b, const Vtx& c) { p[0] = a; p[1] = b; p[2] = c; len = 3; }
};

// ccd_simplex_step (ccd.rs:442-546): true when the simplex encloses the origin.
NEWT_HD bool simplex_step(Simplex& s, V3<double>& dir) {
    const Vtx a = s.p[s.len - 1];
    const V3<double> ao = -a.m;
    if (s.len == 2) {
        const Vtx b = s.p[0];
        const V3<double> ab = b.m - a.m;
        if (dot3(ab, ao) > 0.0) {
            dir = triple(ab, ao, ab);
            if (lsq(dir) <= kSq) dir = mk3(0.0, 0.0, 1.0);
        } else {
            s.set1(a);
            dir = ao;
        }
    } else if (s.len == 3) {
        const Vtx b = s.p[1], c = s.p[0];
        const V3<double> ab = b.m - a.m, ac = c.m - a.m;
        const V3<double> abc = cross3(ab, ac);
        if (lsq(abc) <= kSq) {
            s.set2(a, b);
            const double el = lsq(ab);
            const double t = el > kSq ? clamp01(-dot3(a.m, ab) / el) : 0.0;
            const V3<double> closest = a.m + ab * t;
            if (lsq(closest) > kSq) {
                dir = -closest;
                return false;
            }
            V3<double> axis;
            if (nabs(ab.x) <= nabs(ab.y) && nabs(ab.x) <= nabs(ab.z)) axis = ex();
            else if (nabs(ab.y) <= nabs(ab.z)) axis = mk3(0.0, 1.0, 0.0);
            else axis = mk3(0.0, 0.0, 1.0);
            dir = cross3(ab, axis);
            if (lsq(dir) <= kSq) dir = ao;
            return false;
        }
        if (dot3(cross3(abc, ac), ao) > 0.0) {
            if (dot3(ac, ao) > 0.0) {
                s.set2(a, c);
                dir = triple(ac, ao, ac);
            } else if (dot3(ab, ao) > 0.0) {
                s.set2(a, b);
                dir = triple(ab, ao, ab);
            } else {
                s.set1(a);
                dir = ao;
            }
        } else if (dot3(cross3(ab, abc), ao) > 0.0) {
            if (dot3(ab, ao) > 0.0) {
                s.set2(a, b);
                dir = triple(ab, ao, ab);
            } else {
                s.set1(a);
                dir = ao;
            }
        } else if (dot3(abc, ao) > 0.0) {
            dir = abc;
        } else {
            s.set3(a, c, b);
            dir = -abc;
        }
    } else {
        const Vtx b = s.p[2], c = s.p[1], d = s.p[0];
        const V3<double> ab = b.m - a.m, ac = c.m - a.m, ad = d.m - a.m;
        const V3<double> abc = cross3(ab, ac), acd = cross3(ac, ad), adb = cross3(ad, ab);
        if (dot3(abc, ao) > 0.0) {
            s.set3(a, c, b);
            dir = abc;
        } else if (dot3(acd, ao) > 0.0) {
            s.set3(a, d, c);
            dir = acd;
        } else if (dot3(adb, ao) > 0.0) {
            s.set3(a, b, d);
            dir = adb;
        } else {
            return true;
        }
    }
    if (lsq(dir) <= kSq) dir = ex();
    return false;
}

// barycentric_triangle_origin (ccd.rs:2094-2110)
NEWT_HD void barycentric(V3<double> a, V3<double> b, V3<double> c, V3<double> p, double* w) {
    const V3<double> v0 = b - a, v1 = c - a, v2 = p - a;
    const double d00 = dot3(v0, v0), d01 = dot3(v0, v1), d11 = dot3(v1, v1), d20 = dot3(v2, v0), d21 = dot3(v2, v1);
    const double den = d00 * d11 - d01 * d01;
    if (nabs(den) <= 1.0e-12) {
        w[0] = 1.0;
        w[1] = 0.0;
        w[2] = 0.0;
        return;
    }
    const double v = (d11 * d20 - d01 * d21) / den;
    const double ww = (d00 * d21 - d01 * d20) / den;
    w[0] = (1.0 - v) - ww;
    w[1] = v;
    w[2] = ww;
}
NEWT_HD void weighted(const Vtx* p, const double* w, int n, V3<double>* pa, V3<double>* pb) {  // ccd.rs:548-556
    V3<double> a = z3<double>(), b = z3<double>();
    for (int i = 0; i < n; ++i) {
        a = a + p[i].a * w[i];
        b = b + p[i].b * w[i];
    }
    *pa = a;
    *pb = b;
}
NEWT_HD void triangle_witness(const Vtx& a, const Vtx& b, const Vtx& c, V3<double>* pa, V3<double>* pb) {  // 649-656
    const V3<double> near = closest_on_triangle(z3<double>(), a.m, b.m, c.m);
    double w[3];
    barycentric(a.m, b.m, c.m, near, w);
    const Vtx p[3] = {a, b, c};
    weighted(p, w, 3, pa, pb);
}
// ccd_closest_witness (ccd.rs:658-694) of a 1..4-point simplex.
NEWT_HD void closest_witness(const Simplex& s, V3<double>* pa, V3<double>* pb) {
    if (s.len == 1) {
        *pa = s.p[0].a;
        *pb = s.p[0].b;
    } else if (s.len == 2) {
        const V3<double> e = s.p[1].m - s.p[0].m;
        const double den = lsq(e);
        const double t = den > kSq ? clamp01(-dot3(s.p[0].m, e) / den) : 0.0;
        const double w[2] = {1.0 - t, t};
        weighted(s.p, w, 2, pa, pb);
    } else if (s.len == 3) {
        triangle_witness(s.p[0], s.p[1], s.p[2], pa, pb);
    } else {
        const int fc[4][3] = {{0, 1, 2}, {0, 3, 1}, {0, 2, 3}, {1, 3, 2}};
        double bd = 0.0;
        for (int k = 0; k < 4; ++k) {
            V3<double> a, b;
            triangle_witness(s.p[fc[k][0]], s.p[fc[k][1]], s.p[fc[k][2]], &a, &b);
            const double d = lsq(a - b);
            if (k == 0 || d < bd) {
                *pa = a;
                *pb = b;
                bd = d;
            }
        }
    }
}

struct Cand {  // CcdDistanceCandidate (ccd.rs:696-701)
    V3<double> point;
    double w[4];
    Vtx f[3];
    int flen;
};
NEWT_HD bool inside_tetra(const Simplex& s) {  // ccd.rs:704-725
    const V3<double> a = s.p[0].m, b = s.p[1].m, c = s.p[2].m, d = s.p[3].m;
    const V3<double> ab = b - a, ac = c - a, ad = d - a;
    const double den = dot3(cross3(ab, ac), ad);
    if (nabs(den) <= kSq) return false;
    const V3<double> ao = -a;
    const double wb = dot3(cross3(ao, ac), ad) / den;
    const double wc = dot3(cross3(ab, ao), ad) / den;
    const double wd = dot3(cross3(ab, ac), ao) / den;
    const double wa = ((1.0 - wb) - wc) - wd;
    const double lo = -1.0e-6, hi = 1.0 + 1.0e-6;
    return wa >= lo && wa <= hi && wb >= lo && wb <= hi && wc >= lo && wc <= hi && wd >= lo && wd <= hi;
}
NEWT_HD bool candidate(const Simplex& s, int mask, Cand* out) {  // ccd.rs:727-792
    int idx[3] = {0, 0, 0};
    int n = 0;
    for (int i = 0; i < s.len; ++i)
        if (mask & (1 << i)) {
            if (n == 3) return false;
            idx[n++] = i;
        }
    for (int i = 1; i < n; ++i)
        for (int p = i; p > 0 && vtx_precedes(s.p[idx[p]], s.p[idx[p - 1]]); --p) {
            const int t = idx[p];
            idx[p] = idx[p - 1];
            idx[p - 1] = t;
        }
    double sw[3] = {0.0, 0.0, 0.0};
    V3<double> point;
    if (n == 1) {
        point = s.p[idx[0]].m;
        sw[0] = 1.0;
    } else if (n == 2) {
        const V3<double> a = s.p[idx[0]].m, e = s.p[idx[1]].m - a;
        const double den = lsq(e);
        const double t = den > kSq ? clamp01(-dot3(a, e) / den) : 0.0;
        point = a + e * t;
        sw[0] = 1.0 - t;
        sw[1] = t;
    } else if (n == 3) {
        const V3<double> a = s.p[idx[0]].m, b = s.p[idx[1]].m, c = s.p[idx[2]].m;
        if (lsq(cross3(b - a, c - a)) <= kSq) return false;
        point = closest_on_triangle(z3<double>(), a, b, c);
        barycentric(a, b, c, point, sw);
    } else {
        return false;
    }
    for (int k = 0; k < 4; ++k) out->w[k] = 0.0;
    for (int k = 0; k < n; ++k) out->w[idx[k]] = sw[k];
    for (int k = 0; k < 3; ++k) out->f[k] = s.p[idx[k < n ? k : 0]];
    out->point = point;
    out->flen = n;
    return true;
}
NEWT_HD bool cand_precedes(const Cand& c, const Cand& u) {  // ccd.rs:794-815
    const double cd = lsq(c.point), ud = lsq(u.point);
    if (cd != ud) return cd < ud;
    if (c.flen != u.flen) return c.flen < u.flen;
    for (int i = 0; i < c.flen; ++i) {
        if (vtx_precedes(c.f[i], u.f[i])) return true;
        if (vtx_precedes(u.f[i], c.f[i])) return false;
    }
    return lex_precedes(c.point, u.point);
}
// ccd_distance_simplex_step (ccd.rs:822-868): false = None.
NEWT_HD bool distance_step(Simplex& s, V3<double>* closest) {
    if (s.len == 4 && inside_tetra(s)) return false;
    bool have = false;
    Cand best;
    for (int mask = 1; mask < (1 << s.len); ++mask) {
        Cand c;
        if (!candidate(s, mask, &c)) continue;
        if (!have || cand_precedes(c, best)) {
            best = c;
            have = true;
        }
    }
    if (!have) return false;
    Vtx sel[4];
    int n = 0;
    for (int i = 0; i < s.len; ++i)
        if (best.w[i] > 1.0e-6) sel[n++] = s.p[i];
    if (n == 0) {
        int li = 0;
        for (int i = 1; i < s.len; ++i) {
            const double d = best.w[i] - best.w[li];
            if (d > 1.0e-6 || (nabs(d) <= 1.0e-6 && vtx_precedes(s.p[i], s.p[li]))) li = i;
        }
        sel[0] = s.p[li];
        n = 1;
    }
    sort_vtx(sel, n);
    for (int i = 0; i < n; ++i) s.p[i] = sel[i];
    s.len = n;
    *closest = best.point;
    return true;
}
// ccd_distance_gjk (ccd.rs:910-975), CCD_MESH_CONFIG: support epsilon 0,
// distance tolerance 1e-10.
NEWT_CCD Simplex distance_gjk(const Mesh& a, const Mesh& b, Simplex s, double ext) {
    Simplex best_s = s;
    double best = __builtin_inf();
    for (int it = 0; it < 64; ++it) {
        V3<double> closest;
        if (!distance_step(s, &closest)) return best_s;
        const double d2 = lsq(closest);
        if (d2 < best) {
            best = d2;
            best_s = s;
        }
        if (d2 <= kSq) return best_s;
        const V3<double> dir = -closest;
        const Vtx p = support(a, b, dir, ext);
        if (dot3(p.m, dir) - dot3(closest, dir) <= 0.0) return best_s;
        for (int i = 0; i < s.len; ++i)
            if (lsq(s.p[i].m - p.m) <= kSq) return best_s;
        s.p[s.len++] = p;
        V3<double> next;
        if (!distance_step(s, &next)) return best_s;
        const double nd2 = lsq(next);
        if (nd2 < best) {
            best = nd2;
            best_s = s;
        }
        if (d2 - nd2 <= 1.0e-10 * nmax(d2, kSq)) return best_s;
    }
    return best_s;
}

struct Out {  // contact geometry: position, normal from B into A, penetration
    V3<double> pos, normal;
    double pen;
};

// ccd_centered_witness (ccd.rs:558-574)
NEWT_HD void centered_witness(const Mesh& a, const Mesh& b, V3<double>* pa, V3<double>* pb, V3<double> dir) {
    if (!(lsq(dir) > kSq)) dir = ex();
    V3<double> c;
    if (a.centered(dir, &c)) *pa = c;
    if (b.centered(-dir, &c)) *pb = c;
}
// ccd_distance_contact (ccd.rs:870-908)
NEWT_HD bool distance_contact(const Mesh& a, const Mesh& b, const Simplex& s, double margin, Out* out) {
    V3<double> ra, rb;
    closest_witness(s, &ra, &rb);
    const V3<double> sep = ra - rb;
    const double raw = len3(sep);
    const double pen = margin - raw;
    if (pen <= 0.0) return false;
    V3<double> pa = ra, pb = rb;
    centered_witness(a, b, &pa, &pb, rb - ra);
    out->normal = raw > 0.0 ? sep / raw : ex();
    out->pos = (pa + pb) * 0.5;
    out->pen = pen;
    return true;
}
// ccd_feature_support (ccd.rs:617-647)
NEWT_HD bool feature_support(const Mesh& s, const Vtx* f, V3<double> dir, bool is_a, bool* has, V3<double>* out) {
    *has = false;
    if (!(lsq(dir) > kSq)) return false;
    dir = unit3(dir);
    bool t;
    const V3<double> sup = s.support(dir, &t);
    double fbest = -__builtin_inf();
    for (int i = 0; i < 3; ++i) fbest = nmax(fbest, dot3(is_a ? f[i].a : f[i].b, dir));
    if (nabs(fbest - dot3(sup, dir)) > s.extent() * 1.0e-7) return false;
    *has = s.centered(dir, out);
    return true;
}
// ccd_polytope_witness (ccd.rs:576-615)
NEWT_HD void polytope_witness(const Mesh& a, const Mesh& b, const Vtx* f, const double* w, V3<double> n, V3<double>* pa,
                              V3<double>* pb) {
    weighted(f, w, 3, pa, pb);
    const bool ua = a.nv > 4 || f[0].ta || f[1].ta || f[2].ta;
    const bool ub = b.nv > 4 || f[0].tb || f[1].tb || f[2].tb;
    bool ma = true, mb = true, ha = false, hb = false;
    V3<double> ca, cb;
    if (ua) ma = feature_support(a, f, n, true, &ha, &ca);
    if (ub) mb = feature_support(b, f, -n, false, &hb, &cb);
    if (ma && ha) *pa = ca;
    if (mb && hb) *pb = cb;
}

// The EPA polytope in the world's workspace: vertices (m, a, b: 9 doubles;
// tie flags: 2 ints), two face buffers (normal, distance: 4 doubles; indices:
// 3 ints) used in turn, the horizon edges (2 ints each), and the manifold's
// polygons (5 x 32 points).
struct Work {
    Lane<double> d;
    Lane<int> i;
    int vd, vi, fd[2], fi[2], e, poly;
    NEWT_HD Vtx vtx(int k) const {
        const int o = vd + 9 * k;
        Vtx v;
        v.m = mk3(d[o], d[o + 1], d[o + 2]);
        v.a = mk3(d[o + 3], d[o + 4], d[o + 5]);
        v.b = mk3(d[o + 6], d[o + 7], d[o + 8]);
        v.ta = i[vi + 2 * k] != 0;
        v.tb = i[vi + 2 * k + 1] != 0;
        return v;
    }
    NEWT_HD V3<double> vm(int k) const { return mk3(d[vd + 9 * k], d[vd + 9 * k + 1], d[vd + 9 * k + 2]); }
    NEWT_HD void put_vtx(int k, const Vtx& v) const {
        const int o = vd + 9 * k;
        d[o] = v.m.x; d[o + 1] = v.m.y; d[o + 2] = v.m.z;
        d[o + 3] = v.a.x; d[o + 4] = v.a.y; d[o + 5] = v.a.z;
        d[o + 6] = v.b.x; d[o + 7] = v.b.y; d[o + 8] = v.b.z;
        i[vi + 2 * k] = v.ta ? 1 : 0;
        i[vi + 2 * k + 1] = v.tb ? 1 : 0;
    }
    NEWT_HD V3<double> fn(int b, int k) const { return mk3(d[fd[b] + 4 * k], d[fd[b] + 4 * k + 1], d[fd[b] + 4 * k + 2]); }
    NEWT_HD double fdist(int b, int k) const { return d[fd[b] + 4 * k + 3]; }
    NEWT_HD int fidx(int b, int k, int c) const { return i[fi[b] + 3 * k + c]; }
    NEWT_HD void copy_face(int bs, int ks, int bd, int kd) const {
        for (int c = 0; c < 4; ++c) d[fd[bd] + 4 * kd + c] = d[fd[bs] + 4 * ks + c];
        for (int c = 0; c < 3; ++c) i[fi[bd] + 3 * kd + c] = i[fi[bs] + 3 * ks + c];
    }
    // ccd_face (ccd.rs:1036-1057) into slot k of buffer b; false = None.
    NEWT_HD bool make_face(int b, int k, int i0, int i1, int i2) const {
        const V3<double> a = vm(i0), bb = vm(i1), c = vm(i2);
        const V3<double> raw = cross3(bb - a, c - a);
        if (lsq(raw) <= kSq) return false;
        V3<double> n = unit3(raw);
        double dist = dot3(n, a);
        int j1 = i1, j2 = i2;
        if (dist < 0.0) {
            j1 = i2;
            j2 = i1;
            n = -n;
            dist = -dist;
        }
        const int o = fd[b] + 4 * k;
        d[o] = n.x; d[o + 1] = n.y; d[o + 2] = n.z; d[o + 3] = dist;
        i[fi[b] + 3 * k] = i0;
        i[fi[b] + 3 * k + 1] = j1;
        i[fi[b] + 3 * k + 2] = j2;
        return true;
    }
    // ccd_face_precedes (ccd.rs:1059-1092)
    NEWT_HD bool face_precedes(int b, int c, int u, double eps) const {
        const double cd = fdist(b, c), ud = fdist(b, u);
        if (cd + eps < ud) return true;
        if (ud + eps < cd) return false;
        Vtx ck[3] = {vtx(fidx(b, c, 0)), vtx(fidx(b, c, 1)), vtx(fidx(b, c, 2))};
        Vtx uk[3] = {vtx(fidx(b, u, 0)), vtx(fidx(b, u, 1)), vtx(fidx(b, u, 2))};
        sort_vtx(ck, 3);
        sort_vtx(uk, 3);
        for (int k = 0; k < 3; ++k) {
            if (vtx_precedes(ck[k], uk[k])) return true;
            if (vtx_precedes(uk[k], ck[k])) return false;
        }
        return false;
    }
    NEWT_HD V3<double> pt(int buf, int k) const {
        const int o = poly + 96 * buf + 3 * k;
        return mk3(d[o], d[o + 1], d[o + 2]);
    }
    NEWT_HD void set_pt(int buf, int k, V3<double> p) const {
        const int o = poly + 96 * buf + 3 * k;
        d[o] = p.x; d[o + 1] = p.y; d[o + 2] = p.z;
    }
};

// ccd_epa_fallback_contact (ccd.rs:1106-1183) over the faces of buffer fb.
NEWT_CCD bool epa_fallback(const Mesh& a, const Mesh& b, const Work& wk, int fb, int nf, double ext, double margin, Out* out) {
    bool have = false;
    V3<double> bax = z3<double>();
    double bd = 0.0;
    auto consider = [&](V3<double> axis) {
        if (lsq(axis) <= kSq) return;
        axis = unit3(axis);
        for (int sgn = 0; sgn < 2; ++sgn) {
            const V3<double> ax = sgn == 0 ? axis : -axis;
            const Vtx s = support(a, b, ax, ext);
            const double dd = dot3(s.m, ax);
            if (dd <= 0.0 || (have && (dd > bd || (dd == bd && !lex_precedes(ax, bax))))) continue;
            bax = ax;
            bd = dd;
            have = true;
        }
    };
    for (int k = 0; k < nf; ++k) consider(wk.fn(fb, k));
    for (int sh = 0; sh < 2; ++sh) {
        const Mesh& s = sh == 0 ? a : b;
        for (int k = 0; k < s.nf; ++k) {  // ccd_mesh_face_normal (1094-1104)
            const V3<double> p0 = s.world(s.f[3 * k]), p1 = s.world(s.f[3 * k + 1]), p2 = s.world(s.f[3 * k + 2]);
            const V3<double> n = cross3(p1 - p0, p2 - p0);
            if (lsq(n) > 0.0) consider(unit3(n));
        }
    }
    consider(b.p - a.p);
    consider(mk3(1.0, 0.0, 0.0));
    consider(mk3(0.0, 1.0, 0.0));
    consider(mk3(0.0, 0.0, 1.0));
    if (!have) return false;
    const Vtx s = support(a, b, bax, 0.0);
    const double pen = bd * ext + margin;
    if (!(pen > 0.0)) return false;
    out->pos = s.b + bax * (bd * ext * 0.5);
    out->normal = -bax;
    out->pen = pen;
    return true;
}

// ccd_convex_contact (ccd.rs:1703-1918), CCD_MESH_CONFIG.
NEWT_CCD bool convex_contact(const Mesh& a, const Mesh& b, const Work& wk, double margin, Out* out) {
    const double ext = nmax(a.extent(), b.extent());
    if (ext <= 0.0) return false;
    V3<double> dir = b.p - a.p;
    if (lsq(dir) == 0.0) dir = ex();
    Simplex s;
    s.set1(support(a, b, dir, ext));
    s.p[1] = s.p[2] = s.p[3] = s.p[0];
    dir = -s.p[0].m;
    bool enclosed = false, separated = false;
    for (int it = 0; it < 32; ++it) {
        const Vtx p = support(a, b, dir, ext);
        if (dot3(p.m, dir) <= 0.0) {
            separated = true;
            break;
        }
        s.p[s.len++] = p;
        if (simplex_step(s, dir)) {
            enclosed = true;
            break;
        }
    }
    if (!enclosed && separated) s = distance_gjk(a, b, s, ext);
    if (!enclosed || s.len != 4) return distance_contact(a, b, s, margin, out);

    for (int k = 0; k < 4; ++k) wk.put_vtx(k, s.p[k]);
    int nverts = 4, cur = 0, nf = 0;
    const int init[4][3] = {{0, 1, 2}, {0, 3, 1}, {0, 2, 3}, {1, 3, 2}};
    for (int k = 0; k < 4; ++k)
        if (wk.make_face(cur, nf, init[k][0], init[k][1], init[k][2])) ++nf;
    if (nf == 0) return distance_contact(a, b, s, margin, out);
    const double feps = 1.0e-4;
    int best = 0;  // index of the best face in buffer cur
    V3<double> bn = wk.fn(cur, 0);
    double bdist = wk.fdist(cur, 0);
    int bidx[3] = {wk.fidx(cur, 0, 0), wk.fidx(cur, 0, 1), wk.fidx(cur, 0, 2)};
    bool converged = false;
    for (int it = 0; it < 128; ++it) {
        best = 0;
        for (int k = 0; k < nf; ++k)
            if (wk.face_precedes(cur, k, best, feps)) best = k;
        bn = wk.fn(cur, best);
        bdist = wk.fdist(cur, best);
        for (int c = 0; c < 3; ++c) bidx[c] = wk.fidx(cur, best, c);
        const Vtx sup = support(a, b, bn, ext);
        if (dot3(sup.m, bn) - bdist <= 1.0e-7) {
            converged = true;
            break;
        }
        if (nverts == kMaxVerts) break;
        bool dup = false;
        for (int k = 0; k < nverts; ++k)
            if (lsq(wk.vm(k) - sup.m) <= kSq) dup = true;
        if (dup) break;
        wk.put_vtx(nverts, sup);
        const int nvx = nverts++;
        const int nxt = 1 - cur;
        int nn = 0, ne = 0;
        for (int k = 0; k < nf; ++k) {
            if (dot3(wk.fn(cur, k), sup.m) > wk.fdist(cur, k) + feps) {
                for (int c = 0; c < 3; ++c) {
                    const int e0 = wk.fidx(cur, k, c), e1 = wk.fidx(cur, k, (c + 1) % 3);
                    int found = -1;
                    for (int j = 0; j < ne; ++j)
                        if (wk.i[wk.e + 2 * j] == e1 && wk.i[wk.e + 2 * j + 1] == e0) {
                            found = j;
                            break;
                        }
                    if (found >= 0) {
                        wk.i[wk.e + 2 * found] = wk.i[wk.e + 2 * (ne - 1)];
                        wk.i[wk.e + 2 * found + 1] = wk.i[wk.e + 2 * (ne - 1) + 1];
                        --ne;
                    } else if (ne < kMaxEdges) {
                        wk.i[wk.e + 2 * ne] = e0;
                        wk.i[wk.e + 2 * ne + 1] = e1;
                        ++ne;
                    }
                }
            } else {
                wk.copy_face(cur, k, nxt, nn++);
            }
        }
        for (int j = 0; j < ne; ++j) {
            if (nn == kMaxFaces) break;
            if (wk.make_face(nxt, nn, wk.i[wk.e + 2 * j], wk.i[wk.e + 2 * j + 1], nvx)) ++nn;
        }
        cur = nxt;
        nf = nn;
        if (nf == 0) return false;
    }
    if (!converged) return epa_fallback(a, b, wk, cur, nf, ext, margin, out);
    const V3<double> nearest = bn * bdist;
    double w[3];
    barycentric(wk.vm(bidx[0]), wk.vm(bidx[1]), wk.vm(bidx[2]), nearest, w);
    const double sum = (w[0] + w[1]) + w[2];
    if (sum <= 1.0e-8) return false;
    for (int k = 0; k < 3; ++k) w[k] = w[k] / sum;
    const Vtx feat[3] = {wk.vtx(bidx[0]), wk.vtx(bidx[1]), wk.vtx(bidx[2])};
    V3<double> pa, pb;
    polytope_witness(a, b, feat, w, bn, &pa, &pb);
    const double pen = bdist * ext + margin;
    if (pen <= 0.0) return false;
    out->normal = -bn;
    out->pos = (pa + pb) * 0.5;
    out->pen = pen;
    return true;
}

// ---------------- manifold (ccd.rs:1185-1701) ---------------------------

// mesh_support_ids (ccd.rs:1189-1221): false = overflow (None).
NEWT_HD bool support_ids(const Mesh& s, V3<double> dir, int* ids, int* len) {
    *len = 0;
    const V3<double> ld = qrot_inv(s.q, dir);
    const double dl = len3(ld);
    if (dl <= 1.0e-12) return true;
    const V3<double> d = ld / dl;
    double best = -__builtin_inf();
    for (int k = 0; k < s.nv; ++k) best = nmax(best, dot3(s.local(k), d));
    const double tol = s.extent() * 1.0e-7;
    for (int k = 0; k < s.nv; ++k)
        if (nabs(dot3(s.local(k), d) - best) <= tol) {
            if (*len == kMultiCap) return false;
            ids[(*len)++] = k;
        }
    return true;
}
NEWT_HD bool has_id(const int* ids, int n, int id) {
    for (int k = 0; k < n; ++k)
        if (ids[k] == id) return true;
    return false;
}
// mesh_feature_faces (ccd.rs:1228-1253)
NEWT_HD bool feature_faces(const Mesh& s, const int* ids, int n, int* faces, int* len) {
    *len = 0;
    for (int k = 0; k < s.nf; ++k) {
        const int f0 = s.f[3 * k], f1 = s.f[3 * k + 1], f2 = s.f[3 * k + 2];
        bool match;
        if (n >= 3) match = has_id(ids, n, f0) && has_id(ids, n, f1) && has_id(ids, n, f2);
        else match = (int)has_id(ids, n, f0) + (int)has_id(ids, n, f1) + (int)has_id(ids, n, f2) == n;
        if (match) {
            if (*len == kMultiCap) return false;
            faces[(*len)++] = k;
        }
    }
    return true;
}
// mesh_feature_normal with mesh_face_world's normal (ccd.rs:1255-1295).
NEWT_HD bool feature_normal(const Mesh& s, const int* ids, int n, V3<double> expected, V3<double>* out) {
    int faces[kMultiCap], nf;
    if (!feature_faces(s, ids, n, faces, &nf)) return false;
    bool have = false;
    for (int k = 0; k < nf; ++k) {
        const int fk = faces[k];
        const V3<double> la = s.local(s.f[3 * fk]), lb = s.local(s.f[3 * fk + 1]), lc = s.local(s.f[3 * fk + 2]);
        const V3<double> ln = cross3(lb - la, lc - la);
        if (lsq(ln) <= kSq) continue;
        const V3<double> nn = unit3(qrot(s.q, unit3(ln)));
        if (!have || dot3(nn, expected) > dot3(*out, expected)) {
            *out = nn;
            have = true;
        }
    }
    return have;
}
NEWT_HD bool start_precedes(V3<double> a, V3<double> b) {  // polygon_start_precedes (ccd.rs:1350-1360)
    int c = total_cmp(a.x, b.x);
    if (c != 0) return c < 0;
    c = total_cmp(b.y, a.y);
    if (c != 0) return c < 0;
    return total_cmp(a.z, b.z) < 0;
}
NEWT_HD bool angle_precedes(V3<double> a, V3<double> b, V3<double> c, V3<double> axis, V3<double> tan) {  // 1362-1379
    a = a - c;
    b = b - c;
    const double ax = dot3(a, axis), ay = dot3(a, tan), bx = dot3(b, axis), by = dot3(b, tan);
    const bool au = ay > 0.0 || (ay == 0.0 && ax >= 0.0);
    const bool bu = by > 0.0 || (by == 0.0 && bx >= 0.0);
    if (au != bu) return au;
    const double cr = ax * by - ay * bx;
    if (cr != 0.0) return cr > 0.0;
    return lex_precedes(a, b);
}
// mesh_feature_polygon (ccd.rs:1297-1348) into polygon buffer pbuf.
NEWT_HD bool feature_polygon(const Mesh& s, const int* ids, int n, V3<double> normal, const Work& wk, int pbuf, int* len) {
    *len = 0;
    for (int k = 0; k < n; ++k) {
        const V3<double> p = s.world(ids[k]);
        bool dup = false;
        for (int j = 0; j < *len; ++j)
            if (lsq(wk.pt(pbuf, j) - p) <= kSq) dup = true;
        if (!dup) {
            if (*len == kClipCap) return false;
            wk.set_pt(pbuf, (*len)++, p);
        }
    }
    if (*len < 3) return true;
    V3<double> c = z3<double>();
    for (int j = 0; j < *len; ++j) c = c + wk.pt(pbuf, j);
    c = c / (double)*len;
    V3<double> axis = unit3(wk.pt(pbuf, 0) - c);
    if (lsq(axis) <= kSq) axis = ex();
    const V3<double> tan = unit3(cross3(normal, axis));
    for (int i = 1; i < *len; ++i)
        for (int p = i; p > 0 && angle_precedes(wk.pt(pbuf, p), wk.pt(pbuf, p - 1), c, axis, tan); --p) {
            const V3<double> t = wk.pt(pbuf, p);
            wk.set_pt(pbuf, p, wk.pt(pbuf, p - 1));
            wk.set_pt(pbuf, p - 1, t);
        }
    return true;
}
// mesh_feature_edges (ccd.rs:1381-1430): the first edge of the feature (two
// or more support vertices) or the distinct edges of the faces around a
// single support vertex; 0 on overflow.
NEWT_HD int feature_edges(const Mesh& s, const int* ids, int n, V3<double> (*edges)[2]) {
    if (n >= 2) {
        edges[0][0] = s.world(ids[0]);
        edges[0][1] = s.world(ids[1]);
        return 1;
    }
    if (n == 0) return 0;
    int faces[kMultiCap], nf;
    if (!feature_faces(s, ids, n, faces, &nf)) return 0;
    int len = 0;
    for (int k = 0; k < nf; ++k)
        for (int o = 0; o < 3; ++o) {
            const int other = s.f[3 * faces[k] + o];
            if (other == ids[0]) continue;
            const V3<double> p0 = s.world(ids[0]), p1 = s.world(other);
            bool dup = false;
            for (int j = 0; j < len; ++j)
                if ((lsq(edges[j][0] - p0) <= kSq && lsq(edges[j][1] - p1) <= kSq) ||
                    (lsq(edges[j][0] - p1) <= kSq && lsq(edges[j][1] - p0) <= kSq))
                    dup = true;
            if (!dup) {
                if (len == kMultiCap) return 0;
                edges[len][0] = p0;
                edges[len][1] = p1;
                ++len;
            }
        }
    return len;
}
// polygon_clip (ccd.rs:1432-1482): the subject (buffer sb) clipped by the
// reference polygon (buffer rb) into buffer 3 (buffer 4 is scratch); false
// = overflow.
NEWT_HD bool clip(const Work& wk, int sb, int sn, int rb, int rn, V3<double> normal, int* on) {
    for (int k = 0; k < sn; ++k) wk.set_pt(3, k, wk.pt(sb, k));
    int pn = sn;
    for (int e = 0; e < rn; ++e) {
        const V3<double> p = wk.pt(rb, e), q = wk.pt(rb, (e + 1) % rn);
        const V3<double> edge = q - p;
        int cn = 0;
        for (int k = 0; k < pn; ++k) {
            const V3<double> st = wk.pt(3, k), en = wk.pt(3, (k + 1) % pn);
            const double sd = dot3(cross3(edge, st - p), normal), ed = dot3(cross3(edge, en - p), normal);
            const bool si = sd >= -1.0e-6, ei = ed >= -1.0e-6;
            if (si != ei) {
                const double den = sd - ed;
                if (nabs(den) > 1.0e-12) {
                    if (cn == kClipCap) return false;
                    wk.set_pt(4, cn++, st + (en - st) * (sd / den));
                }
            }
            if (ei) {
                if (cn == kClipCap) return false;
                wk.set_pt(4, cn++, en);
            }
        }
        for (int k = 0; k < cn; ++k) wk.set_pt(3, k, wk.pt(4, k));
        pn = cn;
        if (pn == 0) break;
    }
    *on = pn;
    return true;
}
NEWT_HD double quad_area(const Work& wk, const int* ix) {  // ccd.rs:1484-1490
    const V3<double> a = wk.pt(3, ix[0]), b = wk.pt(3, ix[1]), c = wk.pt(3, ix[2]), d = wk.pt(3, ix[3]);
    return (len3(cross3(b - a, c - a)) + len3(cross3(c - a, d - a))) * 0.5;
}

struct Buf {  // ContactBuf: at most 4 contacts kept
    Out c[4];
    int n;
};

// append_multi_contacts (ccd.rs:1532-1592): false = overflow (None).
NEWT_HD bool append_multi(Buf& out, const Work& wk, int sb, int sn, int rb, int rn, V3<double> rnormal, V3<double> proj,
                          bool subject_is_a, const Out& base, double margin) {
    int pn;
    if (!clip(wk, sb, sn, rb, rn, rnormal, &pn)) return false;
    if (pn == 0) return true;
    int sel[4] = {0, 1, 2, 3};
    if (pn > 4) {  // quad_indices (1492-1510)
        double ba = quad_area(wk, sel);
        for (int a = 0; a < pn - 3; ++a)
            for (int b = a + 1; b < pn - 2; ++b)
                for (int c = b + 1; c < pn - 1; ++c)
                    for (int d = c + 1; d < pn; ++d) {
                        const int cand[4] = {a, b, c, d};
                        const double ar = quad_area(wk, cand);
                        if (ar > ba) {
                            ba = ar;
                            for (int k = 0; k < 4; ++k) sel[k] = cand[k];
                        }
                    }
    }
    const int n = pn < 4 ? pn : 4;
    int first = 0;  // ordered_indices (1512-1530)
    for (int k = 1; k < n; ++k) {
        const V3<double> pk = wk.pt(3, sel[k]), pf = wk.pt(3, sel[first]);
        if (n == 2 ? lex_precedes(pk, pf) : start_precedes(pk, pf)) first = k;
    }
    int ord[4];
    for (int k = 0; k < n; ++k) ord[k] = sel[(first + n - k) % n];
    const V3<double> rp = wk.pt(rb, 0);
    for (int k = 0; k < n; ++k) {
        const V3<double> sp = wk.pt(3, ord[k]);
        const double sg = dot3(sp - rp, proj);
        const V3<double> pa = subject_is_a ? sp : sp - proj * sg;
        const V3<double> pb = subject_is_a ? sp - proj * sg : sp;
        const double pen = margin - dot3(pa - pb, base.normal);
        if (pen <= 0.0) continue;
        const V3<double> pos = (pa + pb) * 0.5;
        bool dup = false;
        for (int j = 0; j < out.n; ++j)
            if (lsq(out.c[j].pos - pos) <= kSq) dup = true;
        if (!dup && out.n < 4) out.c[out.n++] = Out{pos, base.normal, pen};
    }
    return true;
}
// mesh_multicontact (ccd.rs:1594-1701): false = keep the single contact.
// Polygon buffers: 0 = A's feature, 1 = B's feature, 2 = an edge subject.
NEWT_CCD bool multicontact(const Mesh& a, const Mesh& b, const Out& base, double margin, const Work& wk, Buf& out) {
    out.n = 0;
    if (margin > 0.0) return false;
    int ia[kMultiCap], ib[kMultiCap], na, nb;
    if (!support_ids(a, -base.normal, ia, &na)) return false;
    if (!support_ids(b, base.normal, ib, &nb)) return false;
    if (na == 0 || nb == 0) return false;
    V3<double> nA, nB;
    if (!feature_normal(a, ia, na, -base.normal, &nA)) return false;
    if (!feature_normal(b, ib, nb, base.normal, &nB)) return false;
    if (na >= 3 && nb >= 3 && dot3(nA, nB) < -0.996) {
        int la, lb;
        if (!feature_polygon(a, ia, na, nA, wk, 0, &la)) return false;
        if (!feature_polygon(b, ib, nb, nB, wk, 1, &lb)) return false;
        if (!append_multi(out, wk, 1, lb, 0, la, nA, nA, false, base, margin)) return false;
    } else if ((na < 3 && nb >= 3) || (nb < 3 && na >= 3)) {
        const bool edge_on_a = na < 3;
        const Mesh& es = edge_on_a ? a : b;
        const Mesh& fs = edge_on_a ? b : a;
        const int* eids = edge_on_a ? ia : ib;
        const int* fids = edge_on_a ? ib : ia;
        const int en = edge_on_a ? na : nb, fnn = edge_on_a ? nb : na;
        const V3<double> fnrm = edge_on_a ? nB : nA;
        V3<double> edges[kMultiCap][2];
        const int ne = feature_edges(es, eids, en, edges);
        const int fb = edge_on_a ? 1 : 0;
        int lf;
        if (!feature_polygon(fs, fids, fnn, fnrm, wk, fb, &lf)) return false;
        for (int e = 0; e < ne; ++e) {
            const V3<double> d = unit3(edges[e][1] - edges[e][0]);
            if (nabs(dot3(d, fnrm)) < 0.0888) {
                wk.set_pt(2, 0, edges[e][0]);
                wk.set_pt(2, 1, edges[e][1]);
                if (!append_multi(out, wk, 2, 2, fb, lf, fnrm, base.normal, edge_on_a, base, margin)) return false;
                break;
            }
        }
    }
    return out.n > 1;
}
// ccd_convex_contacts (ccd.rs:2071-2092)
NEWT_CCD void mesh_mesh(const Mesh& a, const Mesh& b, double margin, const Work& wk, Buf& out) {
    out.n = 0;
    Out base;
    if (!convex_contact(a, b, wk, margin, &base)) return;
    if (multicontact(a, b, base, margin, wk, out)) return;
    out.n = 1;
    out.c[0] = base;
}

// mesh_plane (contact.rs:1424-1510): normal = the plane's.
NEWT_CCD void mesh_plane(const Mesh& m, V3<double> pp, Qt<double> pq, double margin, Buf& out) {
    out.n = 0;
    const V3<double> n = qrot(pq, mk3(0.0, 0.0, 1.0)), tan = qrot(pq, mk3(0.0, 1.0, 0.0));
    const double tie = m.extent() * 1.0e-6;
    double maxpen = -__builtin_inf();
    for (int k = 0; k < m.nv; ++k) maxpen = nmax(maxpen, margin - dot3(m.world(k) - pp, n));
    bool hp = false, hs = false;
    double pt = 0.0, st = 0.0;
    V3<double> pw = z3<double>(), sw = z3<double>();
    for (int k = 0; k < m.nv; ++k) {
        const V3<double> w = m.world(k);
        const double pen = margin - dot3(w - pp, n);
        if (pen < 0.0 || maxpen - pen > tie) continue;
        const double tv = dot3(w, tan);
        if (!hp || tv > pt || (tv == pt && lex_precedes(w, pw))) {
            hp = true;
            pt = tv;
            pw = w;
        }
    }
    if (hp)
        for (int k = 0; k < m.nv; ++k) {
            const V3<double> w = m.world(k);
            const double pen = margin - dot3(w - pp, n);
            if (pen < 0.0) continue;
            const double tv = dot3(w, tan);
            if (!same3(w, pw) && (!hs || tv < st || (tv == st && lex_precedes(w, sw)))) {
                hs = true;
                st = tv;
                sw = w;
            }
        }
    for (int k = 0; k < 2; ++k) {
        if (!(k == 0 ? hp : hs)) continue;
        const V3<double> w = k == 0 ? pw : sw;
        const double sg = dot3(w - pp, n);
        out.c[out.n++] = Out{w - n * (sg * 0.5), n, margin - sg};
    }
}

// sphere_mesh (contact.rs:1720-1776): normal from the mesh into the sphere.
NEWT_CCD void sphere_mesh(V3<double> sp, double r, const Mesh& m, double margin, Buf& out) {
    out.n = 0;
    const V3<double> pl = qrot_inv(m.q, sp - m.p);
    double best = __builtin_inf();
    V3<double> bq = z3<double>();
    for (int k = 0; k < m.nf; ++k) {
        const V3<double> q = closest_on_triangle(pl, m.local(m.f[3 * k]), m.local(m.f[3 * k + 1]), m.local(m.f[3 * k + 2]));
        const double d2 = lsq(pl - q);
        if (d2 < best) {
            best = d2;
            bq = q;
        }
    }
    const double dist = __builtin_sqrt(best);
    const double pen = (r - dist) + margin;
    if (pen <= 0.0) return;
    const V3<double> nl = dist > 1.0e-9 ? (pl - bq) / dist : mk3(0.0, 0.0, 1.0);
    const V3<double> nw = qrot(m.q, nl);
    out.c[0] = Out{(m.world_of(bq) + (sp - nw * r)) * 0.5, nw, pen};
    out.n = 1;
}

}  // namespace ccd
}  // namespace newt
