// newt_model_json.hip — newt's JSON scene format (newt/src/model.rs,
// docs/model-format.md) behind the C ABI: newt_model_from_json builds a
// newt_model from the text, and the newt_model_get_* accessors read a
// model's descriptions back.
//
// Strict like newt's loader: unknown fields are rejected at every scope and
// every error names the JSON path of the offending value.  Numbers are
// rounded to f32 as newt parses them, and derived quantities (solid-shape
// inertias, normalized axes and quaternions) are computed in f32 in newt's
// operation order, so the model holds newt's values exactly.  What the GPU
// step does not build is rejected, not dropped: trees together with
// bodies, actuators, equalities, tendons,
// heightfields, sensors, explicit pair lists, ball joints, cylinders,
// ellipsoids and rolling friction.  Sites are accepted and ignored.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <map>
#include <set>
#include <string>
#include <vector>

#include "newt_handles.hpp"
#include "newt_json.hpp"

using namespace newt;
using newt::json::Value;

namespace {

struct LoadError {
    std::string path, what;
};

[[noreturn]] void err(const std::string& path, const std::string& what) { throw LoadError{path, what}; }

const Value& obj(const Value* v, const std::string& path, std::initializer_list<const char*> allowed) {
    if (!v || v->kind != Value::Object) err(path, "expected an object");
    for (const auto& kv : v->obj) {
        bool ok = false;
        for (const char* a : allowed) ok = ok || kv.first == a;
        if (!ok) err(path + "." + kv.first, "unknown field");
    }
    return *v;
}

float num(const Value* v, const std::string& path, bool positive = false, bool nonneg = false) {
    if (!v || v->kind != Value::Number || !std::isfinite(v->num)) err(path, "expected a finite number");
    const float x = (float)v->num;
    if (positive && !(x > 0.0f)) err(path, "must be > 0");
    if (nonneg && !(x >= 0.0f)) err(path, "must be >= 0");
    return x;
}

float num_or(const Value* v, double dflt, const std::string& path, bool positive = false, bool nonneg = false) {
    if (!v) {
        Value d;
        d.kind = Value::Number;
        d.num = dflt;
        return num(&d, path, positive, nonneg);
    }
    return num(v, path, positive, nonneg);
}

std::vector<float> vec(const Value* v, size_t n, const std::string& path) {
    if (!v || v->kind != Value::Array || v->arr.size() != n)
        err(path, "expected an array of " + std::to_string(n) + " numbers");
    std::vector<float> out;
    for (size_t i = 0; i < n; ++i) out.push_back(num(&v->arr[i], path + "[" + std::to_string(i) + "]"));
    return out;
}

std::vector<float> vec_or(const Value* v, std::vector<float> dflt, const std::string& path) {
    return v ? vec(v, dflt.size(), path) : dflt;
}

bool truthy(const Value* v) {
    if (!v) return false;
    switch (v->kind) {
        case Value::Null: return false;
        case Value::Bool: return v->b;
        case Value::Number: return v->num != 0.0;
        case Value::String: return !v->str.empty();
        case Value::Array: return !v->arr.empty();
        case Value::Object: return !v->obj.empty();
    }
    return false;
}

int posint(const Value* v, int dflt, const std::string& path) {
    if (!v) return dflt;
    if (v->kind != Value::Number || !v->integral || v->num < 1 || v->num > 2147483647.0)
        err(path, "expected a positive integer");
    return (int)v->num;
}

// Quat::renormalize (math.rs:590-597) in f32.
std::vector<float> quat_normalized(std::vector<float> q) {
    const float n = std::sqrt(((q[0] * q[0] + q[1] * q[1]) + q[2] * q[2]) + q[3] * q[3]);
    if (n == 0.0f) return {0.0f, 0.0f, 0.0f, 1.0f};
    return {q[0] / n, q[1] / n, q[2] / n, q[3] / n};
}

std::vector<float> normalized(const std::vector<float>& a, const std::string& path) {
    const float n = std::sqrt((a[0] * a[0] + a[1] * a[1]) + a[2] * a[2]);
    if (n == 0.0f) err(path, "axis must be non-zero");
    return {a[0] / n, a[1] / n, a[2] / n};
}

// {"position": [..], "orientation": [x,y,z,w] | "orientation_axis_angle": {...}}.
void pose(const Value* v, const std::string& path, bool allow_orientation, std::vector<float>* pos,
          std::vector<float>* quat) {
    *pos = {0.0f, 0.0f, 0.0f};
    *quat = {0.0f, 0.0f, 0.0f, 1.0f};
    if (!v || v->kind == Value::Null) return;
    obj(v, path, {"position", "orientation", "orientation_axis_angle"});
    *pos = vec_or(v->get("position"), {0.0f, 0.0f, 0.0f}, path + ".position");
    const Value* o = v->get("orientation");
    const Value* aa = v->get("orientation_axis_angle");
    if (o && aa) err(path, "give orientation or orientation_axis_angle, not both");
    if (o) {
        *quat = quat_normalized(vec(o, 4, path + ".orientation"));
    } else if (aa) {
        const std::string p = path + ".orientation_axis_angle";
        obj(aa, p, {"axis", "angle"});
        const std::vector<float> axis = vec(aa->get("axis"), 3, p + ".axis");
        const float angle = num(aa->get("angle"), p + ".angle");
        const std::vector<float> ax = normalized(axis, p + ".axis");
        const float half = angle * 0.5f;
        const float s = (float)std::sin((double)half), c = (float)std::cos((double)half);
        *quat = {ax[0] * s, ax[1] * s, ax[2] * s, c};
    }
    if (!allow_orientation && !((*quat)[0] == 0.0f && (*quat)[1] == 0.0f && (*quat)[2] == 0.0f && (*quat)[3] == 1.0f))
        err(path, "non-root joint offsets must have identity orientation");
}

// solid_capsule_inertia (geom.rs:966-981) in f32.
void capsule_inertia(float m, float r, float h, float* ixx, float* izz) {
    const float vol_cyl = (r * r) * (2.0f * h);
    const float vol_caps = (((4.0f / 3.0f) * r) * r) * r;
    const float total = vol_cyl + vol_caps;
    const float m_c = m * (vol_cyl / total);
    const float m_s = m - m_c;
    *izz = (((0.5f * m_c) * r) * r) + ((((2.0f / 5.0f) * m_s) * r) * r);
    const float offset = h + (3.0f / 8.0f) * r;
    const float hemi = (((83.0f / 320.0f) * m_s) * r) * r;
    const float t1 = ((1.0f / 12.0f) * m_c) * (((3.0f * r) * r) + ((4.0f * h) * h));
    *ixx = (t1 + hemi) + ((m_s * offset) * offset);
}

// parse_inertia (model.rs:326-400): diag, tensor or solid shape; column-major.
void inertia(const Value* v, float mass, const std::string& path, double out[9]) {
    obj(v, path, {"kind", "values", "shape"});
    const Value* k = v->get("kind");
    const std::string kind = k && k->kind == Value::String ? k->str : "";
    for (int i = 0; i < 9; ++i) out[i] = 0.0;
    if (kind == "diag") {
        const std::vector<float> d = vec(v->get("values"), 3, path + ".values");
        if (!(d[0] > 0 && d[1] > 0 && d[2] > 0)) err(path + ".values", "diagonal moments must be > 0");
        out[0] = d[0];
        out[4] = d[1];
        out[8] = d[2];
        return;
    }
    if (kind == "tensor") {
        const std::vector<float> t = vec(v->get("values"), 6, path + ".values");
        const double ixx = t[0], iyy = t[1], izz = t[2], ixy = t[3], ixz = t[4], iyz = t[5];
        const double det = ixx * (iyy * izz - iyz * iyz) - ixy * (ixy * izz - iyz * ixz) + ixz * (ixy * iyz - iyy * ixz);
        if (det == 0.0) err(path + ".values", "inertia tensor is singular");
        const double c[9] = {ixx, ixy, ixz, ixy, iyy, iyz, ixz, iyz, izz};
        for (int i = 0; i < 9; ++i) out[i] = c[i];
        return;
    }
    if (kind == "solid") {
        const std::string sp = path + ".shape";
        const Value* sh = v->get("shape");
        obj(sh, sp, {"kind", "half_extents", "radius", "half_height", "semi_axes"});
        const Value* skv = sh->get("kind");
        const std::string sk = skv && skv->kind == Value::String ? skv->str : "";
        if (sk == "box") {
            const std::vector<float> h = vec(sh->get("half_extents"), 3, sp + ".half_extents");
            const float hx2 = h[0] * h[0], hy2 = h[1] * h[1], hz2 = h[2] * h[2];
            const float t = mass / 3.0f;
            out[0] = t * (hy2 + hz2);
            out[4] = t * (hx2 + hz2);
            out[8] = t * (hx2 + hy2);
            return;
        }
        if (sk == "sphere") {
            const float r = num(sh->get("radius"), sp + ".radius", true);
            const float i = (((2.0f / 5.0f) * mass) * r) * r;
            out[0] = out[4] = out[8] = i;
            return;
        }
        if (sk == "capsule") {
            const float r = num(sh->get("radius"), sp + ".radius", true);
            const float hh = num(sh->get("half_height"), sp + ".half_height", false, true);
            float ixx, izz;
            capsule_inertia(mass, r, hh, &ixx, &izz);
            out[0] = out[4] = ixx;
            out[8] = izz;
            return;
        }
        err(sp + ".kind", "solid '" + sk + "' inertia is not built on the GPU path");
    }
    err(path + ".kind", "expected diag, tensor or solid");
}

void solref(const Value* v, const std::string& path, double out[2]) {
    if (!v || v->kind == Value::Null) {
        out[0] = (float)0.02;
        out[1] = 1.0;
        return;
    }
    obj(v, path, {"timeconst", "dampratio"});
    out[0] = num_or(v->get("timeconst"), 0.02, path + ".timeconst");
    out[1] = num_or(v->get("dampratio"), 1.0, path + ".dampratio");
}

void solimp(const Value* v, const std::string& path, double out[4], int* power) {
    if (!v || v->kind == Value::Null) {
        out[0] = (float)0.9;
        out[1] = (float)0.95;
        out[2] = (float)0.001;
        out[3] = 0.5;
        *power = 2;
        return;
    }
    obj(v, path, {"dmin", "dmax", "width", "midpoint", "power"});
    *power = posint(v->get("power"), 2, path + ".power");
    out[0] = num_or(v->get("dmin"), 0.9, path + ".dmin");
    out[1] = num_or(v->get("dmax"), 0.95, path + ".dmax");
    out[2] = num_or(v->get("width"), 0.001, path + ".width");
    out[3] = num_or(v->get("midpoint"), 0.5, path + ".midpoint");
}

void set3(double* d, const std::vector<float>& v) {
    for (size_t i = 0; i < v.size(); ++i) d[i] = v[i];
}

struct Loaded {
    newt_options opt{};
    std::vector<std::vector<newt_link_desc>> trees;  // World::trees in order
    std::vector<int> self_collide;
    std::vector<newt_body_desc> bodies;
    std::vector<newt_geom_desc> geoms;
    std::vector<std::vector<double>> mesh_v;  // World::meshes: vertices (x, y, z) and faces (i, j, k)
    std::vector<std::vector<int>> mesh_f;
    std::map<std::string, int> mesh_index;
};

// parse_mesh_asset (model.rs:1640-1699): {name, vertices, faces}, checked as
// ConvexMesh::validate does (geom.rs:318-340).
void load_meshes(const Value* arr, Loaded* L) {
    if (!arr) return;
    if (arr->kind != Value::Array) err("$.meshes", "expected an array");
    for (size_t i = 0; i < arr->arr.size(); ++i) {
        const Value& m = arr->arr[i];
        const std::string p = "$.meshes[" + std::to_string(i) + "]";
        obj(&m, p, {"name", "vertices", "faces"});
        const Value* nm = m.get("name");
        if (!nm || nm->kind != Value::String || nm->str.empty()) err(p + ".name", "mesh name must not be empty");
        if (L->mesh_index.count(nm->str)) err(p, "duplicate mesh name \"" + nm->str + "\"");
        const Value* vs = m.get("vertices");
        if (!vs || vs->kind != Value::Array) err(p + ".vertices", "expected an array");
        std::vector<double> v;
        for (size_t k = 0; k < vs->arr.size(); ++k)
            for (float x : vec(&vs->arr[k], 3, p + ".vertices[" + std::to_string(k) + "]")) v.push_back(x);
        const Value* fs = m.get("faces");
        if (!fs || fs->kind != Value::Array) err(p + ".faces", "expected an array");
        std::vector<int> f;
        for (size_t k = 0; k < fs->arr.size(); ++k) {
            const std::string fp = p + ".faces[" + std::to_string(k) + "]";
            const Value& t = fs->arr[k];
            if (t.kind != Value::Array || t.arr.size() != 3) err(fp, "face must have exactly 3 vertex indices");
            for (size_t c = 0; c < 3; ++c) {
                const float n = num(&t.arr[c], fp + "[" + std::to_string(c) + "]");
                if (n < 0.0f || n != std::floor(n)) err(fp + "[" + std::to_string(c) + "]", "face vertex index must be a non-negative integer");
                f.push_back((int)n);
            }
        }
        const int nv = (int)v.size() / 3, nf = (int)f.size() / 3;
        if (nv < 4) err(p, "convex mesh needs >= 4 vertices");
        if (nf < 4) err(p, "convex mesh needs >= 4 triangular faces");
        for (int x : f)
            if (x >= nv) err(p, "mesh face index out of range");
        L->mesh_index[nm->str] = (int)L->mesh_v.size();
        L->mesh_v.push_back(v);
        L->mesh_f.push_back(f);
    }
}

void load_links(const Value* arr, const std::string& path, std::vector<newt_link_desc>* out,
                std::map<std::string, int>* index) {
    if (!arr || arr->kind != Value::Array || arr->arr.empty()) err(path, "expected a non-empty array of links");
    for (size_t i = 0; i < arr->arr.size(); ++i) {
        const Value& l = arr->arr[i];
        const std::string p = path + "[" + std::to_string(i) + "]";
        obj(&l, p, {"name", "parent", "joint", "joint_offset_in_parent", "joint_offset_in_child", "mass", "inertia"});
        const Value* nm = l.get("name");
        if (!nm || nm->kind != Value::String || index->count(nm->str)) err(p + ".name", "expected a unique name");
        const Value* parent = l.get("parent");
        int pi = -1;
        if (i == 0) {
            if (parent && parent->kind != Value::Null) err(p + ".parent", "the root has no parent");
        } else {
            if (!parent || parent->kind != Value::String || !index->count(parent->str))
                err(p + ".parent", "must name an earlier link");
            pi = index->at(parent->str);
        }
        const Value* j = l.get("joint");
        obj(j, p + ".joint", {"kind", "axis", "range", "damping", "armature", "limit"});
        const Value* kv = j->get("kind");
        const std::string kind = kv && kv->kind == Value::String ? kv->str : "";
        if (kind == "ball") err(p + ".joint.kind", "ball joints are not built on the GPU path");
        int jk;
        if (kind == "free") jk = NEWT_JOINT_FREE;
        else if (kind == "fixed") jk = NEWT_JOINT_FIXED;
        else if (kind == "hinge") jk = NEWT_JOINT_HINGE;
        else if (kind == "slide") jk = NEWT_JOINT_SLIDE;
        else err(p + ".joint.kind", "expected free, fixed, hinge or slide");
        if (i == 0 && kind != "free" && kind != "fixed") err(p + ".joint.kind", "the root joint must be free or fixed");
        if (i > 0 && kind == "free") err(p + ".joint.kind", "free joints only at the root");
        const bool single = kind == "hinge" || kind == "slide";
        newt_link_desc d{};
        d.parent = pi;
        d.joint = jk;
        if (single) set3(d.axis, normalized(vec(j->get("axis"), 3, p + ".joint.axis"), p + ".joint.axis"));
        if (const Value* r = j->get("range")) {
            if (!single) err(p + ".joint.range", "only hinges and slides take a range");
            const std::vector<float> lr = vec(r, 2, p + ".joint.range");
            if (!(lr[0] < lr[1])) err(p + ".joint.range", "lo < hi required");
            d.has_range = 1;
            d.range[0] = lr[0];
            d.range[1] = lr[1];
        }
        const float damping = num_or(j->get("damping"), 0.0, p + ".joint.damping", false, true);
        const float armature = num_or(j->get("armature"), 0.0, p + ".joint.armature", false, true);
        if (const Value* lim = j->get("limit")) {
            // model.rs:679-735: stiffness and damping required, solref/solimp optional
            const std::string lp = p + ".joint.limit";
            obj(lim, lp, {"stiffness", "damping", "solref", "solimp"});
            if (!lim->get("stiffness") || !lim->get("damping")) err(lp, "stiffness and damping are required");
            d.limit_custom = 1;
            d.limit_stiffness = num(lim->get("stiffness"), lp + ".stiffness", false, true);
            d.limit_damping = num(lim->get("damping"), lp + ".damping", false, true);
            d.limit_solref[0] = 0.02;
            d.limit_solref[1] = 1.0;
            const double si_dflt[4] = {0.9, 0.95, 0.001, 0.5};
            for (int k = 0; k < 4; ++k) d.limit_solimp[k] = si_dflt[k];
            d.limit_solimp_power = 2;
            if (lim->get("solref")) solref(lim->get("solref"), lp + ".solref", d.limit_solref);
            if (lim->get("solimp")) solimp(lim->get("solimp"), lp + ".solimp", d.limit_solimp, &d.limit_solimp_power);
        }
        const float mass = num(l.get("mass"), p + ".mass", true);
        inertia(l.get("inertia"), mass, p + ".inertia", d.inertia);
        std::vector<float> ppos, pquat, cpos, cquat;
        pose(l.get("joint_offset_in_parent"), p + ".joint_offset_in_parent", i == 0, &ppos, &pquat);
        pose(l.get("joint_offset_in_child"), p + ".joint_offset_in_child", false, &cpos, &cquat);
        set3(d.parent_pos, ppos);
        set3(d.parent_quat, pquat);
        set3(d.child_pos, cpos);
        d.child_quat[3] = 1.0;
        d.mass = mass;
        d.damping = kind != "free" ? damping : 0.0;
        d.free_damping = kind == "free" ? damping : 0.0;
        d.armature = armature;
        out->push_back(d);
        (*index)[nm->str] = (int)i;
    }
}

void load_bodies(const Value* arr, Loaded* L, std::map<std::string, int>* index) {
    if (!arr) return;
    if (arr->kind != Value::Array) err("$.bodies", "expected an array");
    for (size_t i = 0; i < arr->arr.size(); ++i) {
        const Value& b = arr->arr[i];
        const std::string p = "$.bodies[" + std::to_string(i) + "]";
        obj(&b, p, {"name", "mass", "inertia", "pose", "velocity", "rolling_friction"});
        const Value* nm = b.get("name");
        if (!nm || nm->kind != Value::String || nm->str.empty() || index->count(nm->str))
            err(p + ".name", "expected a unique name");
        if (!b.get("mass")) err(p + ".mass", "required");
        const float mass = num(b.get("mass"), p + ".mass", true);
        if (!b.get("inertia")) err(p + ".inertia", "required");
        newt_body_desc d{};
        d.mass = mass;
        inertia(b.get("inertia"), mass, p + ".inertia", d.inertia);
        std::vector<float> pos, quat;
        pose(b.get("pose"), p + ".pose", true, &pos, &quat);
        set3(d.position, pos);
        set3(d.orientation, quat);
        const Value* rf = b.get("rolling_friction");
        if (rf && rf->kind != Value::Null) err(p + ".rolling_friction", "rolling friction is not built on the GPU path");
        if (const Value* vel = b.get("velocity")) {
            obj(vel, p + ".velocity", {"linear", "angular_body"});
            if (vel->get("linear")) set3(d.linear_velocity, vec(vel->get("linear"), 3, p + ".velocity.linear"));
            if (vel->get("angular_body"))
                set3(d.angular_velocity_body, vec(vel->get("angular_body"), 3, p + ".velocity.angular_body"));
        }
        d.gravity_scale = 1.0;
        (*index)[nm->str] = (int)i;
        L->bodies.push_back(d);
    }
}

void load_geoms(const Value* arr, const std::map<std::string, int>& tree_index,
                const std::vector<std::map<std::string, int>>& links,
                const std::map<std::string, int>& bodies, Loaded* L) {
    if (!arr) return;
    if (arr->kind != Value::Array) err("$.geoms", "expected an array");
    for (size_t i = 0; i < arr->arr.size(); ++i) {
        const Value& g = arr->arr[i];
        const std::string p = "$.geoms[" + std::to_string(i) + "]";
        obj(&g, p, {"name", "shape", "attach", "local_offset", "local_orientation", "friction", "solref", "solimp",
                    "condim", "margin", "gap", "torsional_friction", "rolling_friction"});
        const Value* sh = g.get("shape");
        obj(sh, p + ".shape", {"kind", "radius", "half_extents", "half_height", "semi_axes", "mesh", "hfield"});
        const Value* kv = sh->get("kind");
        const std::string kind = kv && kv->kind == Value::String ? kv->str : "";
        newt_geom_desc d{};
        if (kind == "plane") {
            d.shape = NEWT_SHAPE_PLANE;
        } else if (kind == "sphere") {
            d.shape = NEWT_SHAPE_SPHERE;
            d.size[0] = num(sh->get("radius"), p + ".shape.radius", true);
        } else if (kind == "box") {
            d.shape = NEWT_SHAPE_BOX;
            set3(d.size, vec(sh->get("half_extents"), 3, p + ".shape.half_extents"));
        } else if (kind == "capsule") {
            d.shape = NEWT_SHAPE_CAPSULE;
            d.size[0] = num(sh->get("radius"), p + ".shape.radius", true);
            d.size[1] = num(sh->get("half_height"), p + ".shape.half_height", false, true);
        } else if (kind == "mesh") {
            obj(sh, p + ".shape", {"kind", "mesh"});
            const Value* mn = sh->get("mesh");
            if (!mn || mn->kind != Value::String || !L->mesh_index.count(mn->str))
                err(p + ".shape.mesh", "unknown mesh name; declare it in the top-level meshes array");
            d.shape = NEWT_SHAPE_MESH;
            d.mesh = L->mesh_index.at(mn->str);
        } else {
            err(p + ".shape.kind", "'" + kind + "' geoms are not built on the GPU path");
        }
        d.tree = d.link = d.body = -1;
        if (d.shape != NEWT_SHAPE_MESH) d.mesh = -1;
        const Value* att = g.get("attach");
        Value dflt;
        if (!att) {
            dflt.kind = Value::Object;
            Value st;
            st.kind = Value::String;
            st.str = "static";
            dflt.obj.emplace_back("kind", st);
            att = &dflt;
        }
        obj(att, p + ".attach", {"kind", "tree", "link", "body"});
        const Value* ak = att->get("kind");
        const std::string akind = ak && ak->kind == Value::String ? ak->str : "";
        if (akind == "static") {
            if (kind != "plane") err(p + ".attach", "only plane geoms may be static");
        } else if (akind == "link") {
            const Value* t = att->get("tree");
            const Value* lk = att->get("link");
            if (!t || t->kind != Value::String || !tree_index.count(t->str) || !lk || lk->kind != Value::String ||
                !links[tree_index.at(t->str)].count(lk->str))
                err(p + ".attach", "must name a tree and one of its links");
            d.tree = tree_index.at(t->str);
            d.link = links[d.tree].at(lk->str);
        } else if (akind == "body") {
            const Value* bd = att->get("body");
            if (!bd || bd->kind != Value::String || !bodies.count(bd->str))
                err(p + ".attach.body", "must name one of the bodies");
            d.body = bodies.at(bd->str);
        } else {
            err(p + ".attach.kind", "expected static, link or body");
        }
        set3(d.pos, vec_or(g.get("local_offset"), {0.0f, 0.0f, 0.0f}, p + ".local_offset"));
        set3(d.quat, quat_normalized(vec_or(g.get("local_orientation"), {0.0f, 0.0f, 0.0f, 1.0f}, p + ".local_orientation")));
        d.condim = 3;
        if (const Value* c = g.get("condim")) {
            if (c->kind != Value::Number || !(c->num == 1 || c->num == 3 || c->num == 4 || c->num == 6))
                err(p + ".condim", "expected 1, 3, 4 or 6");
            d.condim = (int)c->num;
        }
        solimp(g.get("solimp"), p + ".solimp", d.solimp, &d.solimp_power);
        d.friction = num_or(g.get("friction"), 0.5, p + ".friction", false, true);
        d.margin = num_or(g.get("margin"), 0.0, p + ".margin", false, true);
        d.gap = num_or(g.get("gap"), 0.0, p + ".gap", false, true);
        solref(g.get("solref"), p + ".solref", d.solref);
        d.group = 0xFFFFFFFFu;
        d.mask = 0xFFFFFFFFu;
        L->geoms.push_back(d);
    }
}

Loaded load(const Value& root) {
    Loaded L;
    obj(&root, "$", {"version", "gravity", "timestep", "integrator", "bodies", "trees", "meshes", "hfields", "geoms",
                     "sites", "actuators", "contact_pairs", "solver", "equality", "tendons", "sensors"});
    if (const Value* ver = root.get("version"))
        if (ver->kind != Value::String || ver->str != "1") err("$.version", "must be \"1\"");
    const std::pair<const char*, const char*> unbuilt[] = {
        {"actuators", "actuators"}, {"equality", "equality constraints"}, {"tendons", "tendons"},
        {"hfields", "heightfields"}, {"contact_pairs", "explicit contact-pair lists"},
        {"sensors", "sensors"}};
    for (const auto& u : unbuilt)
        if (truthy(root.get(u.first))) err(std::string("$.") + u.first, std::string(u.second) + " are not built on the GPU path");
    const std::vector<float> gravity = vec_or(root.get("gravity"), {0.0f, 0.0f, -9.81f}, "$.gravity");
    const float dt = num_or(root.get("timestep"), 0.005, "$.timestep", true);
    int integ = NEWT_INTEGRATOR_RK4;
    if (const Value* iv = root.get("integrator")) {
        const std::string s = iv->kind == Value::String ? iv->str : "";
        if (s == "RK4") integ = NEWT_INTEGRATOR_RK4;
        else if (s == "Euler") integ = NEWT_INTEGRATOR_EULER;
        else if (s == "implicitfast") integ = NEWT_INTEGRATOR_IMPLICITFAST;
        else err("$.integrator", "expected RK4, Euler or implicitfast");
    }
    Value empty;
    empty.kind = Value::Object;
    const Value* solver = root.get("solver") ? root.get("solver") : &empty;
    obj(solver, "$.solver", {"mode", "iterations", "cone"});
    int mode = NEWT_SOLVER_PENALTY;
    if (const Value* mv = solver->get("mode")) {
        const std::string s = mv->kind == Value::String ? mv->str : "";
        if (s == "penalty") mode = NEWT_SOLVER_PENALTY;
        else if (s == "pgs") mode = NEWT_SOLVER_PGS;
        else if (s == "newton") mode = NEWT_SOLVER_NEWTON;
        else err("$.solver.mode", "expected penalty, pgs or newton");
    }
    const int iterations = posint(solver->get("iterations"), 20, "$.solver.iterations");
    int cone = NEWT_CONE_PYRAMIDAL;
    if (const Value* cv = solver->get("cone")) {
        const std::string s = cv->kind == Value::String ? cv->str : "";
        if (s == "pyramidal") cone = NEWT_CONE_PYRAMIDAL;
        else if (s == "elliptic") cone = NEWT_CONE_ELLIPTIC;
        else err("$.solver.cone", "expected pyramidal or elliptic");
    }
    L.opt.dt = dt;
    set3(L.opt.gravity, gravity);
    L.opt.integrator = integ;
    L.opt.solver_mode = mode;
    L.opt.iterations = iterations;
    L.opt.cone = cone;
    const Value* trees = root.get("trees");
    const Value* bodies = root.get("bodies");
    const bool n_trees = trees && trees->kind == Value::Array && !trees->arr.empty();
    std::map<std::string, int> body_index, tree_index;
    std::vector<std::map<std::string, int>> link_index;
    load_bodies(bodies, &L, &body_index);
    if (n_trees || L.bodies.empty()) {
        if (!n_trees) err("$.trees", "expected a non-empty array of trees");
        for (size_t k = 0; k < trees->arr.size(); ++k) {
            const std::string tp = "$.trees[" + std::to_string(k) + "]";
            const Value& t = trees->arr[k];
            obj(&t, tp, {"name", "self_collide", "links"});
            const Value* nm = t.get("name");
            const std::string name = nm && nm->kind == Value::String ? nm->str : "tree";
            if (tree_index.count(name)) err(tp + ".name", "expected a unique name");
            tree_index[name] = (int)k;
            int collide = 0;
            if (const Value* sc = t.get("self_collide")) {
                if (sc->kind != Value::Bool) err(tp + ".self_collide", "expected a bool");
                collide = sc->b ? 1 : 0;
            }
            L.trees.emplace_back();
            link_index.emplace_back();
            L.self_collide.push_back(collide);
            load_links(t.get("links"), tp + ".links", &L.trees.back(), &link_index.back());
        }
    }
    load_meshes(root.get("meshes"), &L);
    load_geoms(root.get("geoms"), tree_index, link_index, body_index, &L);
    return L;
}

}  // namespace

extern "C" {

newt_status newt_model_from_json(const char* text, int32_t finalize, newt_model** out) {
    if (!text || !out) return fail(NEWT_ERR_INVALID_ARG, "null argument");
    Loaded L;
    try {
        json::Reader r(text);
        const Value root = r.document();
        L = load(root);
    } catch (const json::ParseError& e) {
        return fail(NEWT_ERR_INVALID_CONFIG, "<json>: " + e.what + " at byte " + std::to_string(e.at));
    } catch (const LoadError& e) {
        return fail(NEWT_ERR_INVALID_CONFIG, e.path + ": " + e.what);
    }
    newt_model* m = nullptr;
    newt_status s = newt_model_create(&L.opt, &m);
    if (s != NEWT_OK) return s;
    for (size_t t = 0; s == NEWT_OK && t < L.trees.size(); ++t)
        s = newt_model_add_tree(m, L.trees[t].data(), (int32_t)L.trees[t].size(), L.self_collide[t], nullptr);
    for (size_t i = 0; s == NEWT_OK && i < L.bodies.size(); ++i) s = newt_model_add_body(m, &L.bodies[i], nullptr);
    for (size_t i = 0; s == NEWT_OK && i < L.mesh_v.size(); ++i)
        s = newt_model_add_mesh(m, L.mesh_v[i].data(), (int32_t)L.mesh_v[i].size() / 3, L.mesh_f[i].data(),
                                (int32_t)L.mesh_f[i].size() / 3, nullptr);
    for (size_t i = 0; s == NEWT_OK && i < L.geoms.size(); ++i) s = newt_model_add_geom(m, &L.geoms[i], nullptr);
    if (s == NEWT_OK && finalize) s = newt_model_finalize(m);
    if (s != NEWT_OK) {
        newt_model_destroy(m);
        return s;
    }
    *out = m;
    return NEWT_OK;
}

newt_status newt_model_counts(const newt_model* model, int32_t* n_trees, int32_t* n_bodies, int32_t* n_geoms) {
    if (!model) return fail(NEWT_ERR_INVALID_ARG, "null model");
    if (n_trees) *n_trees = (int32_t)model->h.trees.size();
    if (n_bodies) *n_bodies = (int32_t)model->h.bodies.size();
    if (n_geoms) *n_geoms = (int32_t)model->h.geoms.size();
    return NEWT_OK;
}

newt_status newt_model_get_options(const newt_model* model, newt_options* out) {
    if (!model || !out) return fail(NEWT_ERR_INVALID_ARG, "null argument");
    *out = model->h.opt;
    return NEWT_OK;
}

newt_status newt_model_get_tree(const newt_model* model, int32_t tree, newt_link_desc* links, int32_t cap,
                                int32_t* n_links, int32_t* self_collide) {
    if (!model) return fail(NEWT_ERR_INVALID_ARG, "null model");
    if (tree < 0 || tree >= (int32_t)model->h.trees.size()) return fail(NEWT_ERR_INVALID_ARG, "no such tree");
    const auto& t = model->h.trees[tree];
    if (n_links) *n_links = (int32_t)t.size();
    if (self_collide) *self_collide = model->h.tree_collide[tree];
    for (int32_t i = 0; links && i < cap && i < (int32_t)t.size(); ++i) links[i] = t[i];
    return NEWT_OK;
}

newt_status newt_model_get_geom(const newt_model* model, int32_t index, newt_geom_desc* out) {
    if (!model || !out) return fail(NEWT_ERR_INVALID_ARG, "null argument");
    if (index < 0 || index >= (int32_t)model->h.geoms.size()) return fail(NEWT_ERR_INVALID_ARG, "no such geom");
    *out = model->h.geoms[index];
    return NEWT_OK;
}

newt_status newt_model_get_body(const newt_model* model, int32_t index, newt_body_desc* out) {
    if (!model || !out) return fail(NEWT_ERR_INVALID_ARG, "null argument");
    if (index < 0 || index >= (int32_t)model->h.bodies.size()) return fail(NEWT_ERR_INVALID_ARG, "no such body");
    *out = model->h.bodies[index];
    return NEWT_OK;
}

}  // extern "C"
