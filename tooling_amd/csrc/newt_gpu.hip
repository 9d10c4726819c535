// newt_gpu.hip — host model compiler, device state and the C ABI of
// include/newt_gpu.h, driving the lane-per-world step kernels for gfx950.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <limits>
#include <string>
#include <vector>

#include "../../include/newt_gpu.h"
#include "newt_compile.hpp"
#include "newt_step.hpp"
#include "newt_team.hpp"
#include "newt_bodies.hpp"

#include "newt_handles.hpp"

using namespace newt;

namespace newt {

namespace {
thread_local std::string g_err;
}  // namespace

newt_status fail(newt_status s, const std::string& msg) {
    g_err = msg;
    return s;
}

}  // namespace newt

namespace {

constexpr int kBlock = 64;  // one wavefront: 64 worlds
constexpr int kBodyWaveLds = 40 * 1024;  // one free-body world: four per CU
// Free-body batches up to this size run one wavefront per world (parallel
// broadphase, narrowphase and row assembly; at most four worlds per CU by
// LDS); larger batches run one lane per world, whose throughput grows with
// the number of wavefronts.  Both give the same bits.
constexpr int kBodyWaveMaxWorlds = 2048;

}  // namespace

namespace {

template <class R>
__device__ __forceinline__ World<R> world_of(const ModelT& m, const Layout& L, const DataT& d, R* ws, double* dws,
                                             int* iws, int w) {
    return World<R>{m, L, {ws, d.W, w}, {dws, d.W, w}, {iws, d.W, w}, {d.qpos, d.W, w}, {d.qvel, d.W, w},
                    {d.qacc, d.W, w}, {d.qfrc, d.W, w}, {d.fat, d.W, w}};
}

// Bit-exact single kernel: n_steps whole World::steps per lane.
template <class R>
__global__ void __launch_bounds__(kBlock) k_step(ModelT m, Layout L, DataT d, R* ws, double* dws, int* iws,
                                                 int n_steps, int init) {
    int w = blockIdx.x * blockDim.x + threadIdx.x;
    if (w >= d.W) return;
    World<R> wd = world_of<R>(m, L, d, ws, dws, iws, w);
    if (d.fault[w] != 0) return;
    const R dt = (R)m.dt;
    wd.forward_kinematics();
    wd.link_velocities();
    if (init) {
        wd.populate(dt, true);
        iws[(size_t)(L.iwf + m.n_links) * d.W + w] = 0;
        iws[(size_t)(L.iwf + m.n_links + 1) * d.W + w] = 0;
        return;
    }
    // update_broadphase at step start (world.rs:1190-1226): a no-op unless
    // the state was set since the last step.
    wd.populate(dt, false);
    for (int s = 0; s < n_steps; ++s) {
        int st = wd.step();
        if (st != 0) {
            d.fault[w] = st;
            break;
        }
    }
    d.npair[w] = iws[(size_t)(L.iwf + m.n_links) * d.W + w];
    d.ncontact[w] = iws[(size_t)(L.iwf + m.n_links + 1) * d.W + w];
}

// Models with trees and free bodies: n_steps World::steps per lane in FP64
// (every precision), bit-identical to the FP64 oracle.
__global__ void __launch_bounds__(kBlock) k_trees_bodies(ModelT m, Layout L, DataT d, double* ws, double* dws,
                                                          int* iws, int n_steps, int init) {
    int w = blockIdx.x * blockDim.x + threadIdx.x;
    if (w >= d.W) return;
    World<double> wd = world_of<double>(m, L, d, ws, dws, iws, w);
    if (d.fault[w] != 0) return;
    FreeBodies fb{wd};
    wd.forward_kinematics();
    wd.link_velocities();
    fb.set_poses();
    if (init) {
        wd.populate(m.dt, true);
        iws[(size_t)(L.iwf + m.n_links) * d.W + w] = 0;
        iws[(size_t)(L.iwf + m.n_links + 1) * d.W + w] = 0;
        return;
    }
    wd.populate(m.dt, false);
    for (int s = 0; s < n_steps; ++s) {
        const int st = fb.step_with_trees();
        if (st != 0) {
            d.fault[w] = st;
            break;
        }
    }
    d.npair[w] = iws[(size_t)(L.iwf + m.n_links) * d.W + w];
    d.ncontact[w] = iws[(size_t)(L.iwf + m.n_links + 1) * d.W + w];
}

// Free-body models: n_steps World::steps per lane in FP64 (every precision).
template <bool Penalty>
__global__ void __launch_bounds__(kBlock) k_bodies(ModelT m, Layout L, DataT d, double* dws, int* iws, int n_steps,
                                                   int init) {
    int w = blockIdx.x * blockDim.x + threadIdx.x;
    if (w >= d.W) return;
    World<double> wd = world_of<double>(m, L, d, nullptr, dws, iws, w);
    if (d.fault[w] != 0) return;
    FreeBodies fb{wd};
    fb.set_poses();
    if (init) {
        wd.populate(m.dt, true);
        iws[(size_t)L.iwf * d.W + w] = 0;
        iws[(size_t)(L.iwf + 1) * d.W + w] = 0;
        return;
    }
    wd.populate(m.dt, false);
    for (int s = 0; s < n_steps; ++s) {
        int st = fb.step<Penalty>();
        if (st != 0) {
            d.fault[w] = st;
            break;
        }
    }
    d.npair[w] = iws[(size_t)L.iwf * d.W + w];
    d.ncontact[w] = iws[(size_t)(L.iwf + 1) * d.W + w];
}

// Free-body models, one wavefront (one workgroup) per world.
__global__ void __launch_bounds__(64) k_bodies_wave(ModelT m, BodyWaveLayout B, Layout L, DataT d, double* dws,
                                                    int* iws, double* gws, int* gwi, int n_steps, int init,
                                                    int exact) {
    extern __shared__ double lds_b[];
    const int world = blockIdx.x;
    const int lane = threadIdx.x;
    if (world >= d.W) return;
    if (d.fault[world] != 0) return;
    double* sd = lds_b;
    int* si = reinterpret_cast<int*>(lds_b + B.ints_at);
    BodyWave bw{m, B, L, d, dws, iws, sd, si, gws + (size_t)world * B.g_stride_d, gwi + (size_t)world * B.g_stride_i,
                world, lane, nullptr, nullptr, nullptr, exact};
    bw.load();
    if (init) {
        bw.populate(true);
        bw.store();
        if (lane == 0) {
            d.npair[world] = 0;
            d.ncontact[world] = 0;
        }
        return;
    }
    bw.populate(false);
    int np = 0, nc = 0;
    for (int s = 0; s < n_steps; ++s) {
        const int st = bw.step(&np, &nc);
        if (st != 0) {
            if (lane == 0) d.fault[world] = st;
            break;
        }
    }
    bw.store();
    if (lane == 0) {
        d.npair[world] = np;
        d.ncontact[world] = nc;
    }
}

// Sensors on the post-step state, one lane per world (newt_sensors.hpp).
// Parked and faulted worlds keep their previous values.
template <class R>
__global__ void __launch_bounds__(kBlock) k_sensors(ModelT m, Layout L, DataT d, R* ws, double* dws, int* iws) {
    const int w = blockIdx.x * blockDim.x + threadIdx.x;
    if (w >= d.W || d.fault[w] != 0) return;
    World<R> wd = world_of<R>(m, L, d, ws, dws, iws, w);
    const int nc = m.n_bodies == 0 ? min(max(d.ncontact[w], 0), m.cap_c) : 0;
    SensorEval<R> se{wd, m, L, {d.sensordata, d.W, w}};
    se.evaluate(nc);
}

newt_status launch_sensors(newt_data* D, hipStream_t stream) {
    const newt_model* M = D->model;
    if (M->dev.n_sensors == 0) return NEWT_OK;
    dim3 grid((D->W + kBlock - 1) / kBlock), block(kBlock);
    const bool ws32 = D->precision == NEWT_PRECISION_MIXED && M->dev.n_bodies == 0 && !lane_only(M->dev);
    if (ws32)
        hipLaunchKernelGGL(k_sensors<float>, grid, block, 0, stream, M->dev, D->L, D->d, (float*)D->ws, D->dws, D->iws);
    else
        hipLaunchKernelGGL(k_sensors<double>, grid, block, 0, stream, M->dev, D->L, D->d, (double*)D->ws, D->dws,
                           D->iws);
    HIP_TRY(hipGetLastError());
    return NEWT_OK;
}

newt_status launch_persistent(newt_data* D, int n_steps, hipStream_t stream) {
    const newt_model* M = D->model;
    const int wpb = D->wpb;
    dim3 grid((D->W + wpb - 1) / wpb), block(64 * wpb);
    HIP_TRY(team_launch(D->precision == NEWT_PRECISION_MIXED, M->dev.nv, M->dev.ext != 0, grid, block, D->lds_bytes,
                        stream, D->d_job, 1, n_steps));
    return NEWT_OK;
}

newt_status launch(newt_data* D, int n_steps, hipStream_t stream, int init) {
    const newt_model* M = D->model;
    D->last_stream = stream;
    dim3 grid((D->W + kBlock - 1) / kBlock), block(kBlock);
    if (M->dev.n_bodies > 0 && M->dev.n_links > 0) {
        hipLaunchKernelGGL(k_trees_bodies, grid, block, 0, stream, M->dev, D->L, D->d, (double*)D->ws, D->dws,
                           D->iws, n_steps, init);
        HIP_TRY(hipGetLastError());
        return NEWT_OK;
    }
    if (M->dev.n_bodies > 0) {
        const bool newton = M->dev.mode == NEWT_SOLVER_NEWTON;  // the Newton solve is wavefront-only
        // Meshes run on the lane step only.
        if (!newton && (D->precision == NEWT_PRECISION_FP64_EXACT || !D->bws || D->W > kBodyWaveMaxWorlds ||
                        M->dev.has_mesh))
        {
            if (M->dev.mode == NEWT_SOLVER_PENALTY && M->dev.n_pairs > 0)
                hipLaunchKernelGGL(k_bodies<true>, grid, block, 0, stream, M->dev, D->L, D->d, D->dws, D->iws, n_steps,
                                   init);
            else
                hipLaunchKernelGGL(k_bodies<false>, grid, block, 0, stream, M->dev, D->L, D->d, D->dws, D->iws, n_steps,
                                   init);
        }
        else
            hipLaunchKernelGGL(k_bodies_wave, dim3(D->W), dim3(64), D->BL.lds_bytes, stream, M->dev, D->BL, D->L, D->d,
                               D->dws, D->iws, D->bws, D->bwi, n_steps, init,
                               D->precision == NEWT_PRECISION_FP64_EXACT ? 1 : 0);
        HIP_TRY(hipGetLastError());
        return NEWT_OK;
    }
    const bool lane_fp64 = lane_only(M->dev);
    if (!init && D->precision != NEWT_PRECISION_FP64_EXACT && !lane_fp64)
        return launch_persistent(D, n_steps, stream);
    if (D->precision == NEWT_PRECISION_MIXED && !lane_fp64)
        hipLaunchKernelGGL(k_step<float>, grid, block, 0, stream, M->dev, D->L, D->d, (float*)D->ws, D->dws, D->iws,
                           n_steps, init);
    else
        hipLaunchKernelGGL(k_step<double>, grid, block, 0, stream, M->dev, D->L, D->d, (double*)D->ws, D->dws, D->iws,
                           n_steps, init);
    HIP_TRY(hipGetLastError());
    return NEWT_OK;
}

// Host [W][n] <-> device [n][W].
void to_lanes(const double* src, std::vector<double>& dst, int W, int n) {
    dst.resize((size_t)W * n);
    for (int w = 0; w < W; ++w)
        for (int k = 0; k < n; ++k) dst[(size_t)k * W + w] = src[(size_t)w * n + k];
}
void from_lanes(const std::vector<double>& src, double* dst, int W, int n) {
    for (int w = 0; w < W; ++w)
        for (int k = 0; k < n; ++k) dst[(size_t)w * n + k] = src[(size_t)k * W + w];
}

// Device-to-host copy ordered after the work enqueued on this data's stream
// (other streams keep running).
hipError_t d2h(const newt_data* D, void* dst, const void* src, size_t bytes) {
    if (!bytes) return hipSuccess;
    hipError_t e = hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, D->last_stream);
    if (e != hipSuccess) return e;
    return hipStreamSynchronize(D->last_stream);
}

// Pinned staging for stream-ordered uploads, sized for every world's qpos,
// qvel and fat bounds; waits until the previous upload has left it.
hipError_t stage(newt_data* D) {
    if (D->staging) return hipEventSynchronize(D->staged);
    const ModelT& m = D->model->dev;
    // Every world's qpos, qvel and bounds; or a world mask and one state.
    hipError_t e = hipHostMalloc((void**)&D->staging,
                                 (size_t)D->W * (m.nq + m.nv + 6 * m.n_geoms) * 8 + D->W + 256 +
                                     (size_t)(m.nq + m.nv + 6 * m.n_geoms) * 8,
                                 hipHostMallocDefault);
    if (e != hipSuccess) return e;
    return hipEventCreateWithFlags(&D->staged, hipEventDisableTiming);
}

}  // namespace

extern "C" {

const char* newt_last_error(void) { return g_err.c_str(); }

newt_status newt_model_create(const newt_options* options, newt_model** out) {
    if (!options || !out) return fail(NEWT_ERR_INVALID_ARG, "null argument");
    if (!(options->dt > 0.0)) return fail(NEWT_ERR_INVALID_CONFIG, "dt must be positive");
    newt_model* m = new newt_model();
    m->h.opt = *options;
    *out = m;
    return NEWT_OK;
}

newt_status newt_model_add_tree(newt_model* model, const newt_link_desc* links, int32_t n_links, int32_t self_collide,
                                int32_t* tree_index) {
    if (!model || !links || n_links < 1) return fail(NEWT_ERR_INVALID_ARG, "bad tree arguments");
    if (model->finalized) return fail(NEWT_ERR_INVALID_ARG, "model is finalized");
    model->h.trees.emplace_back(links, links + n_links);
    model->h.tree_collide.push_back(self_collide);
    if (tree_index) *tree_index = (int32_t)model->h.trees.size() - 1;
    return NEWT_OK;
}

newt_status newt_model_add_body(newt_model* model, const newt_body_desc* body, int32_t* body_index) {
    if (!model || !body) return fail(NEWT_ERR_INVALID_ARG, "bad body arguments");
    if (model->finalized) return fail(NEWT_ERR_INVALID_ARG, "model is finalized");
    model->h.bodies.push_back(*body);
    if (body_index) *body_index = (int32_t)model->h.bodies.size() - 1;
    return NEWT_OK;
}

newt_status newt_model_add_geom(newt_model* model, const newt_geom_desc* geom, int32_t* geom_index) {
    if (!model || !geom) return fail(NEWT_ERR_INVALID_ARG, "bad geom arguments");
    if (model->finalized) return fail(NEWT_ERR_INVALID_ARG, "model is finalized");
    model->h.geoms.push_back(*geom);
    if (geom_index) *geom_index = (int32_t)model->h.geoms.size() - 1;
    return NEWT_OK;
}

newt_status newt_model_add_mesh(newt_model* model, const double* vertices, int32_t n_vertices, const int32_t* faces,
                                int32_t n_faces, int32_t* mesh_index) {
    if (!model || !vertices || !faces) return fail(NEWT_ERR_INVALID_ARG, "null argument");
    if (model->finalized) return fail(NEWT_ERR_INVALID_ARG, "model is finalized");
    std::string err;
    const newt_status s = check_mesh(vertices, n_vertices, faces, n_faces, &err);
    if (s != NEWT_OK) return fail(s, err);
    HostModel& H = model->h;
    MeshT t{(int)H.mesh_v.size() / 3, n_vertices, (int)H.mesh_f.size() / 3, n_faces};
    H.mesh_v.insert(H.mesh_v.end(), vertices, vertices + 3 * (size_t)n_vertices);
    H.mesh_f.insert(H.mesh_f.end(), faces, faces + 3 * (size_t)n_faces);
    H.meshes.push_back(t);
    if (mesh_index) *mesh_index = (int32_t)H.meshes.size() - 1;
    return NEWT_OK;
}

newt_status newt_model_add_sensor(newt_model* model, const newt_sensor_desc* sensor, int32_t* sensor_index) {
    if (!model || !sensor) return fail(NEWT_ERR_INVALID_ARG, "null argument");
    if (model->finalized) return fail(NEWT_ERR_INVALID_ARG, "model already finalized");
    if (sensor_index) *sensor_index = (int32_t)model->h.sensors.size();
    model->h.sensors.push_back(*sensor);
    return NEWT_OK;
}

newt_status newt_model_set_magnetic_field(newt_model* model, const double field[3]) {
    if (!model || !field) return fail(NEWT_ERR_INVALID_ARG, "null argument");
    if (model->finalized) return fail(NEWT_ERR_INVALID_ARG, "model already finalized");
    for (int i = 0; i < 3; ++i) model->h.magnetic[i] = field[i];
    return NEWT_OK;
}

newt_status newt_model_sensor_dim(const newt_model* model, int32_t* dim) {
    if (!model || !dim) return fail(NEWT_ERR_INVALID_ARG, "null argument");
    if (!model->finalized) return fail(NEWT_ERR_INVALID_ARG, "model is not finalized");
    *dim = model->dev.sensor_dim;
    return NEWT_OK;
}

newt_status newt_model_finalize(newt_model* model) {
    if (!model) return fail(NEWT_ERR_INVALID_ARG, "null model");
    if (model->finalized) return NEWT_OK;
    std::string err;
    newt_status s = compile_model(&model->h, &err);
    if (s != NEWT_OK) return fail(s, err);
    HostModel& H = model->h;
    if (H.contact_cap > 0 && H.contact_cap < H.dev.cap_c) {
        H.dev.cap_c = H.contact_cap;
        H.dev.cap_r = 4 * H.contact_cap;
    }
    s = compile_sensors(&H, &err);
    if (s != NEWT_OK) return fail(s, err);
    HIP_TRY(hipMalloc(&model->d_sensors, sizeof(SensorT) * std::max<size_t>(1, H.hsensors.size())));
    if (!H.hsensors.empty())
        HIP_TRY(hipMemcpy(model->d_sensors, H.hsensors.data(), sizeof(SensorT) * H.hsensors.size(),
                          hipMemcpyHostToDevice));
    HIP_TRY(hipMalloc(&model->d_links, sizeof(LinkT) * std::max<size_t>(1, H.links.size())));
    HIP_TRY(hipMalloc(&model->d_geoms, sizeof(GeomT) * std::max<size_t>(1, H.hgeoms.size())));
    HIP_TRY(hipMalloc(&model->d_pairs, sizeof(PairT) * std::max<size_t>(1, H.pairs.size())));
    HIP_TRY(hipMalloc(&model->d_chain, sizeof(int) * std::max<size_t>(1, H.chain.size())));
    if (!H.links.empty())
        HIP_TRY(hipMemcpy(model->d_links, H.links.data(), sizeof(LinkT) * H.links.size(), hipMemcpyHostToDevice));
    if (!H.hgeoms.empty())
        HIP_TRY(hipMemcpy(model->d_geoms, H.hgeoms.data(), sizeof(GeomT) * H.hgeoms.size(), hipMemcpyHostToDevice));
    if (!H.pairs.empty())
        HIP_TRY(hipMemcpy(model->d_pairs, H.pairs.data(), sizeof(PairT) * H.pairs.size(), hipMemcpyHostToDevice));
    if (!H.chain.empty())
        HIP_TRY(hipMemcpy(model->d_chain, H.chain.data(), sizeof(int) * H.chain.size(), hipMemcpyHostToDevice));
    HIP_TRY(hipMalloc(&model->d_level_off, sizeof(int) * std::max<size_t>(1, H.level_off.size())));
    HIP_TRY(hipMalloc(&model->d_level_links, sizeof(int) * std::max<size_t>(1, H.level_links.size())));
    if (!H.level_off.empty())
        HIP_TRY(hipMemcpy(model->d_level_off, H.level_off.data(), sizeof(int) * H.level_off.size(),
                          hipMemcpyHostToDevice));
    if (!H.level_links.empty())
        HIP_TRY(hipMemcpy(model->d_level_links, H.level_links.data(), sizeof(int) * H.level_links.size(),
                          hipMemcpyHostToDevice));
    HIP_TRY(hipMalloc(&model->d_level_nrank, sizeof(int) * std::max<size_t>(1, H.level_nrank.size())));
    HIP_TRY(hipMalloc(&model->d_dof_link, sizeof(int) * std::max<size_t>(1, H.dof_link.size())));
    if (!H.dof_link.empty())
        HIP_TRY(hipMemcpy(model->d_dof_link, H.dof_link.data(), sizeof(int) * H.dof_link.size(),
                          hipMemcpyHostToDevice));
    if (!H.level_nrank.empty())
        HIP_TRY(hipMemcpy(model->d_level_nrank, H.level_nrank.data(), sizeof(int) * H.level_nrank.size(),
                          hipMemcpyHostToDevice));
    HIP_TRY(hipMalloc(&model->d_bodies, sizeof(BodyT) * std::max<size_t>(1, H.hbodies.size())));
    if (!H.hbodies.empty())
        HIP_TRY(hipMemcpy(model->d_bodies, H.hbodies.data(), sizeof(BodyT) * H.hbodies.size(), hipMemcpyHostToDevice));
    HIP_TRY(hipMalloc(&model->d_meshes, sizeof(MeshT) * std::max<size_t>(1, H.meshes.size())));
    HIP_TRY(hipMalloc(&model->d_mesh_v, sizeof(double) * std::max<size_t>(1, H.mesh_v.size())));
    HIP_TRY(hipMalloc(&model->d_mesh_f, sizeof(int) * std::max<size_t>(1, H.mesh_f.size())));
    if (!H.meshes.empty()) {
        HIP_TRY(hipMemcpy(model->d_meshes, H.meshes.data(), sizeof(MeshT) * H.meshes.size(), hipMemcpyHostToDevice));
        HIP_TRY(hipMemcpy(model->d_mesh_v, H.mesh_v.data(), sizeof(double) * H.mesh_v.size(), hipMemcpyHostToDevice));
        HIP_TRY(hipMemcpy(model->d_mesh_f, H.mesh_f.data(), sizeof(int) * H.mesh_f.size(), hipMemcpyHostToDevice));
    }
    HIP_TRY(hipMalloc(&model->d_limits, sizeof(LimitT) * std::max<size_t>(1, H.limits.size())));
    if (!H.limits.empty())
        HIP_TRY(hipMemcpy(model->d_limits, H.limits.data(), sizeof(LimitT) * H.limits.size(), hipMemcpyHostToDevice));
    model->dev = H.dev;
    model->dev.links = model->d_links;
    model->dev.geoms = model->d_geoms;
    model->dev.pairs = model->d_pairs;
    model->dev.chain = model->d_chain;
    model->dev.level_off = model->d_level_off;
    model->dev.level_links = model->d_level_links;
    model->dev.level_nrank = model->d_level_nrank;
    model->dev.dof_link = model->d_dof_link;
    model->dev.limits = model->d_limits;
    model->dev.bodies = model->d_bodies;
    model->dev.sensors = model->d_sensors;
    model->dev.meshes = model->d_meshes;
    model->dev.mesh_v = model->d_mesh_v;
    model->dev.mesh_f = model->d_mesh_f;
    if (H.dev.n_bodies == 0) {
        const ModelT& hd = H.dev;
        CacheLayout C = make_cache_layout(hd);
        std::vector<char> blob(C.bytes, 0);
        auto put = [&](int off, const void* src, size_t bytes) {
            if (bytes) std::memcpy(blob.data() + off, src, bytes);
        };
        put(C.links, H.links.data(), sizeof(LinkT) * H.links.size());
        put(C.geoms, H.hgeoms.data(), sizeof(GeomT) * H.hgeoms.size());
        put(C.pairs, H.pairs.data(), sizeof(PairT) * H.pairs.size());
        put(C.level_off, H.level_off.data(), sizeof(int) * H.level_off.size());
        put(C.level_links, H.level_links.data(), sizeof(int) * H.level_links.size());
        put(C.level_nrank, H.level_nrank.data(), sizeof(int) * H.level_nrank.size());
        put(C.dof_link, H.dof_link.data(), sizeof(int) * H.dof_link.size());
        put(C.limits, H.limits.data(), sizeof(LimitT) * H.limits.size());
        HIP_TRY(hipMalloc(&model->d_blob, blob.size()));
        HIP_TRY(hipMemcpy(model->d_blob, blob.data(), blob.size(), hipMemcpyHostToDevice));
        model->cache = C;
    }
    model->finalized = true;
    return NEWT_OK;
}

newt_status newt_model_set_contact_capacity(newt_model* model, int32_t max_contacts) {
    if (!model || max_contacts < 0) return fail(NEWT_ERR_INVALID_ARG, "bad capacity arguments");
    if (model->finalized) return fail(NEWT_ERR_INVALID_ARG, "model is finalized");
    model->h.contact_cap = max_contacts;
    return NEWT_OK;
}

newt_status newt_model_dims(const newt_model* model, int32_t* nq, int32_t* nv, int32_t* n_links, int32_t* n_geoms) {
    if (!model || !model->finalized) return fail(NEWT_ERR_INVALID_ARG, "model is not finalized");
    if (nq) *nq = model->dev.nq;
    if (nv) *nv = model->dev.nv;
    if (n_links) *n_links = model->dev.n_links;
    if (n_geoms) *n_geoms = model->dev.n_geoms;
    return NEWT_OK;
}

void newt_model_destroy(newt_model* model) {
    if (!model) return;
    (void)hipFree(model->d_links);
    (void)hipFree(model->d_geoms);
    (void)hipFree(model->d_pairs);
    (void)hipFree(model->d_chain);
    (void)hipFree(model->d_level_off);
    (void)hipFree(model->d_level_links);
    (void)hipFree(model->d_level_nrank);
    (void)hipFree(model->d_dof_link);
    (void)hipFree(model->d_limits);
    (void)hipFree(model->d_bodies);
    (void)hipFree(model->d_blob);
    (void)hipFree(model->d_sensors);
    (void)hipFree(model->d_meshes);
    (void)hipFree(model->d_mesh_v);
    (void)hipFree(model->d_mesh_f);
    delete model;
}

newt_status newt_data_create(const newt_model* model, int32_t n_worlds, int32_t precision, newt_data** out) {
    if (!model || !out || n_worlds < 1) return fail(NEWT_ERR_INVALID_ARG, "bad data arguments");
    if (!model->finalized) return fail(NEWT_ERR_INVALID_ARG, "model is not finalized");
    if (precision != NEWT_PRECISION_FP64 && precision != NEWT_PRECISION_MIXED &&
        precision != NEWT_PRECISION_FP64_EXACT)
        return fail(NEWT_ERR_INVALID_ARG, "unknown precision");
    // Free-body models run the lane-per-world FP64 step in every precision;
    // so do models with trees and free bodies (lane_only), on the tree layout.
    const bool bodies = model->dev.n_bodies > 0 && model->dev.n_links == 0;
    const bool with_bodies = model->dev.n_bodies > 0;
    const bool team = !bodies && !lane_only(model->dev) && precision != NEWT_PRECISION_FP64_EXACT;
    if ((model->dev.nv > 64 || model->dev.n_links > 128) && team)
        return fail(NEWT_ERR_UNSUPPORTED, "the wavefront-per-world step handles nv <= 64 and <= 128 links; "
                                          "use NEWT_PRECISION_FP64_EXACT for larger trees");
    const TeamLayout TL = with_bodies ? TeamLayout{} : make_team_layout(model->dev, precision == NEWT_PRECISION_MIXED ? 4 : 8);
    const int cache_bytes = with_bodies ? 0 : model->cache.bytes;
    const int max_wpb = team_max_worlds_for(model->dev.nv);
    const int wpb = with_bodies ? 1 : std::min(max_wpb, (kMaxLds - cache_bytes) / std::max(TL.bytes, 1));
    if (team && wpb < 1)
        return fail(NEWT_ERR_UNSUPPORTED, "one world's step state exceeds the 160 KiB of LDS of a workgroup");
    newt_data* D = new newt_data();
    D->model = model;
    D->W = n_worlds;
    D->precision = precision;
    // The FP32 workspace serves the mixed tree kernels only.
    const bool ws32 = precision == NEWT_PRECISION_MIXED && !bodies && !lane_only(model->dev);
    D->L = bodies ? make_body_layout(model->dev) : make_layout(model->dev, ws32 ? 4 : 8);
    D->TL = TL;
    D->wpb = std::max(wpb, 1);
    D->lds_bytes = cache_bytes + D->wpb * TL.bytes;
    if (team && D->lds_bytes > 65536) {
        // The attribute is per kernel, shared by every newt_data: allow the
        // whole workgroup LDS so that no data narrows another's launch.
        hipError_t e = team_set_lds(precision == NEWT_PRECISION_MIXED, model->dev.nv, model->dev.ext != 0, kMaxLds);
        if (e != hipSuccess) {
            delete D;
            return fail(NEWT_ERR_DEVICE, std::string("hipFuncSetAttribute: ") + hipGetErrorString(e));
        }
    }
    const ModelT& m = model->dev;
    size_t W = (size_t)n_worlds;
    size_t rsz = ws32 ? sizeof(float) : sizeof(double);
    auto alloc = [&](void** p, size_t bytes) -> bool {
        D->device_bytes += std::max<size_t>(bytes, 8);
        return hipMalloc(p, std::max<size_t>(bytes, 8)) == hipSuccess;
    };
    bool ok = alloc((void**)&D->d.qpos, W * m.nq * 8) && alloc((void**)&D->d.qvel, W * m.nv * 8) &&
              alloc((void**)&D->d.qacc, W * m.nv * 8) && alloc((void**)&D->d.qfrc, W * m.nv * 8) &&
              alloc((void**)&D->d.fat, W * m.n_geoms * 6 * 8) && alloc((void**)&D->d.fault, W * 4) &&
              alloc((void**)&D->d.ncontact, W * 4) && alloc((void**)&D->d.npair, W * 4) &&
              alloc(&D->ws, W * D->L.total_r * rsz) && alloc((void**)&D->dws, W * D->L.total_d * 8) &&
              alloc((void**)&D->iws, W * D->L.total_i * 4) &&
              alloc((void**)&D->ovf, W * (size_t)std::max(TL.ov_stride, 16)) &&
              alloc((void**)&D->d.sensordata, W * (size_t)std::max(m.sensor_dim, 1) * 8);
    const bool newton = bodies && m.mode == NEWT_SOLVER_NEWTON;
    // Penalty contacts and meshes run on the lane step only.
    const bool penalty_contacts = m.mode == NEWT_SOLVER_PENALTY && m.n_pairs > 0;
    if (ok && bodies && !m.has_mesh &&
        (newton || (precision != NEWT_PRECISION_FP64_EXACT && n_worlds <= kBodyWaveMaxWorlds && !penalty_contacts))) {
        // The wavefront step keeps a world's bodies and fat bounds in LDS;
        // models whose bodies do not fit run the lane step (PGS only).
        D->BL = make_body_wave_layout(m, kBodyWaveLds);
        const bool fits = 8 * (34 * m.n_bodies + 6 * m.n_geoms) + 64 <= kBodyWaveLds;
        if (newton && !fits) {
            newt_data_destroy(D);
            return fail(NEWT_ERR_UNSUPPORTED, "the Newton free-body step keeps a world's bodies in 40 KiB of LDS; "
                                              "this model has too many bodies or geoms");
        }
        if (fits) {
            ok = alloc((void**)&D->bws, W * (size_t)D->BL.g_stride_d * 8) &&
                 alloc((void**)&D->bwi, W * (size_t)D->BL.g_stride_i * 4);
            if (ok && D->BL.lds_bytes > 65536)
                ok = hipFuncSetAttribute((const void*)k_bodies_wave, hipFuncAttributeMaxDynamicSharedMemorySize,
                                         D->BL.lds_bytes) == hipSuccess;
        }
    }
    if (!ok) {
        newt_data_destroy(D);
        return fail(NEWT_ERR_OOM, "device allocation failed");
    }
    D->d.W = n_worlds;
    std::vector<double> q((size_t)W * m.nq), qd((size_t)W * m.nv, 0.0);
    for (size_t w = 0; w < W; ++w) {
        for (int k = 0; k < m.nq; ++k) q[w * m.nq + k] = model->h.q_default[k];
        for (int k = 0; k < m.nv; ++k) qd[w * m.nv + k] = model->h.v_default[k];
    }
    HIP_TRY(hipMemset(D->d.qfrc, 0, W * m.nv * 8));
    HIP_TRY(hipMemset(D->d.qacc, 0, W * m.nv * 8));
    HIP_TRY(hipMemset(D->d.fault, 0, W * 4));
    HIP_TRY(hipMemset(D->d.ncontact, 0, W * 4));
    HIP_TRY(hipMemset(D->d.npair, 0, W * 4));
    HIP_TRY(hipMemset(D->iws, 0, W * D->L.total_i * 4));
    HIP_TRY(hipMemset(D->d.sensordata, 0, W * (size_t)std::max(m.sensor_dim, 1) * 8));
    *out = D;
    // Default state, then the fat bounds World::add_geom builds from it.
    newt_status s = newt_data_set_state(D, q.data(), qd.data());
    if (s != NEWT_OK) return s;
    s = launch(D, 0, nullptr, 1);
    if (s != NEWT_OK) return s;
    HIP_TRY(hipDeviceSynchronize());
    if (team) {
        const TeamJob job = team_job(D);
        HIP_TRY(hipMalloc(&D->d_job, sizeof(TeamJob)));
        D->device_bytes += sizeof(TeamJob);
        HIP_TRY(hipMemcpy(D->d_job, &job, sizeof(TeamJob), hipMemcpyHostToDevice));
    }
    D->fat_default.resize((size_t)6 * m.n_geoms);
    if (m.n_geoms)
        HIP_TRY(hipMemcpy2D(D->fat_default.data(), 8, D->d.fat, W * 8, 8, 6 * m.n_geoms, hipMemcpyDeviceToHost));
    return NEWT_OK;
}

void newt_data_destroy(newt_data* D) {
    if (!D) return;
    (void)hipFree(D->d.qpos);
    (void)hipFree(D->d.qvel);
    (void)hipFree(D->d.qacc);
    (void)hipFree(D->d.qfrc);
    (void)hipFree(D->d.fat);
    (void)hipFree(D->d.fault);
    (void)hipFree(D->d.ncontact);
    (void)hipFree(D->d.npair);
    (void)hipFree(D->d.sensordata);
    (void)hipFree(D->ws);
    (void)hipFree(D->dws);
    (void)hipFree(D->iws);
    (void)hipFree(D->ovf);
    (void)hipFree(D->bws);
    (void)hipFree(D->bwi);
    (void)hipFree(D->d_job);
    if (D->staged) {
        (void)hipEventSynchronize(D->staged);
        (void)hipEventDestroy(D->staged);
    }
    (void)hipHostFree(D->staging);
    delete D;
}

newt_status newt_data_set_state(newt_data* D, const double* qpos, const double* qvel) {
    if (!D || !qpos || !qvel) return fail(NEWT_ERR_INVALID_ARG, "null argument");
    const ModelT& m = D->model->dev;
    std::vector<double> a, b;
    to_lanes(qpos, a, D->W, m.nq);
    to_lanes(qvel, b, D->W, m.nv);
    if (m.nq) HIP_TRY(hipMemcpyAsync(D->d.qpos, a.data(), a.size() * 8, hipMemcpyHostToDevice, D->last_stream));
    if (m.nv) HIP_TRY(hipMemcpyAsync(D->d.qvel, b.data(), b.size() * 8, hipMemcpyHostToDevice, D->last_stream));
    HIP_TRY(hipMemsetAsync(D->d.fault, 0, (size_t)D->W * 4, D->last_stream));
    HIP_TRY(hipStreamSynchronize(D->last_stream));
    return NEWT_OK;
}

newt_status newt_data_set_qfrc_applied(newt_data* D, const double* qfrc) {
    if (!D || !qfrc) return fail(NEWT_ERR_INVALID_ARG, "null argument");
    const ModelT& m = D->model->dev;
    std::vector<double> a;
    to_lanes(qfrc, a, D->W, m.nv);
    if (m.nv) HIP_TRY(hipMemcpyAsync(D->d.qfrc, a.data(), a.size() * 8, hipMemcpyHostToDevice, D->last_stream));
    HIP_TRY(hipStreamSynchronize(D->last_stream));
    return NEWT_OK;
}

newt_status newt_data_get_qfrc_applied(const newt_data* D, double* qfrc) {
    if (!D || !qfrc) return fail(NEWT_ERR_INVALID_ARG, "null argument");
    const ModelT& m = D->model->dev;
    std::vector<double> a((size_t)D->W * m.nv);
    HIP_TRY(d2h(D, a.data(), D->d.qfrc, a.size() * 8));
    from_lanes(a, qfrc, D->W, m.nv);
    return NEWT_OK;
}

newt_status newt_data_set_state_async(newt_data* D, const double* qpos, const double* qvel, void* stream) {
    if (!D || !qpos || !qvel) return fail(NEWT_ERR_INVALID_ARG, "null argument");
    const ModelT& m = D->model->dev;
    const size_t nq = (size_t)D->W * m.nq, nv = (size_t)D->W * m.nv;
    HIP_TRY(stage(D));
    for (int w = 0; w < D->W; ++w) {
        for (int k = 0; k < m.nq; ++k) D->staging[(size_t)k * D->W + w] = qpos[(size_t)w * m.nq + k];
        for (int k = 0; k < m.nv; ++k) D->staging[nq + (size_t)k * D->W + w] = qvel[(size_t)w * m.nv + k];
    }
    hipStream_t s = (hipStream_t)stream;
    if (nq) HIP_TRY(hipMemcpyAsync(D->d.qpos, D->staging, nq * 8, hipMemcpyHostToDevice, s));
    if (nv) HIP_TRY(hipMemcpyAsync(D->d.qvel, D->staging + nq, nv * 8, hipMemcpyHostToDevice, s));
    HIP_TRY(hipMemsetAsync(D->d.fault, 0, (size_t)D->W * 4, s));
    HIP_TRY(hipEventRecord(D->staged, s));
    D->last_stream = s;
    return NEWT_OK;
}

newt_status newt_data_set_worlds_async(newt_data* D, int32_t first, int32_t count, const double* qpos,
                                       const double* qvel, void* stream) {
    if (!D || !qpos || !qvel) return fail(NEWT_ERR_INVALID_ARG, "null argument");
    if (first < 0 || count < 1 || first + count > D->W) return fail(NEWT_ERR_INVALID_ARG, "world range out of bounds");
    const ModelT& m = D->model->dev;
    const size_t W = (size_t)D->W, c = (size_t)count;
    const size_t nq = c * m.nq, nv = c * m.nv, nf = c * 6 * m.n_geoms;
    HIP_TRY(stage(D));
    double* sq = D->staging;
    double* sv = sq + nq;
    double* sf = sv + nv;
    for (size_t w = 0; w < c; ++w) {
        for (int k = 0; k < m.nq; ++k) sq[(size_t)k * c + w] = qpos[w * m.nq + k];
        for (int k = 0; k < m.nv; ++k) sv[(size_t)k * c + w] = qvel[w * m.nv + k];
    }
    // A new world's proxies are the ones World::add_geom inserts at the
    // default state (world.rs:974-979); the next step refits them to qpos.
    for (int k = 0; k < 6 * m.n_geoms; ++k)
        for (size_t w = 0; w < c; ++w) sf[(size_t)k * c + w] = D->fat_default[k];
    hipStream_t s = (hipStream_t)stream;
    if (m.nq) HIP_TRY(hipMemcpy2DAsync(D->d.qpos + first, W * 8, sq, c * 8, c * 8, m.nq, hipMemcpyHostToDevice, s));
    if (m.nv) HIP_TRY(hipMemcpy2DAsync(D->d.qvel + first, W * 8, sv, c * 8, c * 8, m.nv, hipMemcpyHostToDevice, s));
    if (nf)
        HIP_TRY(hipMemcpy2DAsync(D->d.fat + first, W * 8, sf, c * 8, c * 8, 6 * m.n_geoms, hipMemcpyHostToDevice, s));
    if (m.nv) {
        HIP_TRY(hipMemset2DAsync(D->d.qfrc + first, W * 8, 0, c * 8, m.nv, s));
        HIP_TRY(hipMemset2DAsync(D->d.qacc + first, W * 8, 0, c * 8, m.nv, s));
    }
    HIP_TRY(hipMemsetAsync(D->d.fault + first, 0, c * 4, s));
    HIP_TRY(hipMemsetAsync(D->d.ncontact + first, 0, c * 4, s));
    HIP_TRY(hipMemsetAsync(D->d.npair + first, 0, c * 4, s));
    HIP_TRY(hipEventRecord(D->staged, s));
    D->last_stream = s;
    return NEWT_OK;
}

newt_status newt_data_park_worlds(newt_data* D, int32_t first, int32_t count, void* stream) {
    if (!D) return fail(NEWT_ERR_INVALID_ARG, "null data");
    if (first < 0 || count < 1 || first + count > D->W) return fail(NEWT_ERR_INVALID_ARG, "world range out of bounds");
    std::vector<int32_t> parked((size_t)count, NEWT_WORLD_PARKED);
    HIP_TRY(stage(D));
    std::memcpy(D->staging, parked.data(), parked.size() * 4);
    hipStream_t s = (hipStream_t)stream;
    HIP_TRY(hipMemcpyAsync(D->d.fault + first, D->staging, (size_t)count * 4, hipMemcpyHostToDevice, s));
    HIP_TRY(hipEventRecord(D->staged, s));
    D->last_stream = s;
    return NEWT_OK;
}

newt_status newt_data_get_worlds(const newt_data* D, int32_t first, int32_t count, double* qpos, double* qvel) {
    if (!D) return fail(NEWT_ERR_INVALID_ARG, "null data");
    if (first < 0 || count < 1 || first + count > D->W) return fail(NEWT_ERR_INVALID_ARG, "world range out of bounds");
    const ModelT& m = D->model->dev;
    const size_t W = (size_t)D->W, c = (size_t)count;
    std::vector<double> a(c * m.nq), b(c * m.nv);
    if (qpos && m.nq)
        HIP_TRY(hipMemcpy2DAsync(a.data(), c * 8, D->d.qpos + first, W * 8, c * 8, m.nq, hipMemcpyDeviceToHost,
                                 D->last_stream));
    if (qvel && m.nv)
        HIP_TRY(hipMemcpy2DAsync(b.data(), c * 8, D->d.qvel + first, W * 8, c * 8, m.nv, hipMemcpyDeviceToHost,
                                 D->last_stream));
    HIP_TRY(hipStreamSynchronize(D->last_stream));
    for (size_t w = 0; w < c; ++w) {
        if (qpos)
            for (int k = 0; k < m.nq; ++k) qpos[w * m.nq + k] = a[(size_t)k * c + w];
        if (qvel)
            for (int k = 0; k < m.nv; ++k) qvel[w * m.nv + k] = b[(size_t)k * c + w];
    }
    return NEWT_OK;
}

newt_status newt_data_device_bytes(const newt_data* D, uint64_t* bytes) {
    if (!D || !bytes) return fail(NEWT_ERR_INVALID_ARG, "null argument");
    *bytes = D->device_bytes;
    return NEWT_OK;
}

newt_status newt_data_get_state(const newt_data* D, double* qpos, double* qvel) {
    if (!D) return fail(NEWT_ERR_INVALID_ARG, "null data");
    const ModelT& m = D->model->dev;
    std::vector<double> a((size_t)D->W * m.nq), b((size_t)D->W * m.nv);
    if (qpos && m.nq) {
        HIP_TRY(d2h(D, a.data(), D->d.qpos, a.size() * 8));
        from_lanes(a, qpos, D->W, m.nq);
    }
    if (qvel && m.nv) {
        HIP_TRY(d2h(D, b.data(), D->d.qvel, b.size() * 8));
        from_lanes(b, qvel, D->W, m.nv);
    }
    return NEWT_OK;
}

newt_status newt_data_get_qacc(const newt_data* D, double* qacc) {
    if (!D || !qacc) return fail(NEWT_ERR_INVALID_ARG, "null argument");
    const ModelT& m = D->model->dev;
    std::vector<double> a((size_t)D->W * m.nv);
    if (m.nv) {
        HIP_TRY(d2h(D, a.data(), D->d.qacc, a.size() * 8));
        from_lanes(a, qacc, D->W, m.nv);
    }
    return NEWT_OK;
}

newt_status newt_data_get_faults(const newt_data* D, int32_t* faults) {
    if (!D || !faults) return fail(NEWT_ERR_INVALID_ARG, "null argument");
    HIP_TRY(d2h(D, faults, D->d.fault, (size_t)D->W * 4));
    return NEWT_OK;
}

newt_status newt_data_get_sensordata(const newt_data* D, double* sensordata) {
    if (!D || !sensordata) return fail(NEWT_ERR_INVALID_ARG, "null argument");
    const int dim = D->model->dev.sensor_dim;
    if (dim == 0) return NEWT_OK;
    std::vector<double> tmp((size_t)D->W * dim);
    HIP_TRY(d2h(D, tmp.data(), D->d.sensordata, tmp.size() * 8));
    from_lanes(tmp, sensordata, D->W, dim);
    return NEWT_OK;
}

newt_status newt_data_contact_capacity(const newt_data* D, int32_t* cap) {
    if (!D || !cap) return fail(NEWT_ERR_INVALID_ARG, "null argument");
    *cap = D->model->dev.cap_c;
    return NEWT_OK;
}

newt_status newt_data_get_contacts(const newt_data* D, int32_t* counts, int32_t* ids, double* geo) {
    if (!D || !counts || !ids || !geo) return fail(NEWT_ERR_INVALID_ARG, "null argument");
    const ModelT& m = D->model->dev;
    const Layout& L = D->L;
    size_t W = (size_t)D->W;
    int cap = m.cap_c;
    HIP_TRY(d2h(D, counts, D->d.ncontact, W * 4));
    std::vector<int> ih(W * 3 * cap);
    HIP_TRY(d2h(D, ih.data(), D->iws + (size_t)L.cids * W, ih.size() * 4));
    std::vector<double> gh(W * 9 * cap);
    HIP_TRY(d2h(D, gh.data(), D->dws + (size_t)L.cgeo * W, gh.size() * 8));
    for (size_t w = 0; w < W; ++w)
        for (int k = 0; k < cap; ++k) {
            for (int j = 0; j < 3; ++j) ids[(w * cap + k) * 3 + j] = ih[(size_t)(3 * k + j) * W + w];
            for (int j = 0; j < 7; ++j) geo[(w * cap + k) * 7 + j] = gh[(size_t)(9 * k + j) * W + w];
        }
    return NEWT_OK;
}

newt_status newt_data_pair_capacity(const newt_data* D, int32_t* cap) {
    if (!D || !cap) return fail(NEWT_ERR_INVALID_ARG, "null argument");
    *cap = std::max(D->model->dev.n_pairs, 1);
    return NEWT_OK;
}

newt_status newt_data_get_pairs(const newt_data* D, int32_t* counts, int32_t* pairs) {
    if (!D || !counts || !pairs) return fail(NEWT_ERR_INVALID_ARG, "null argument");
    const newt_model* M = D->model;
    size_t W = (size_t)D->W;
    int cap = std::max(M->dev.n_pairs, 1);
    HIP_TRY(d2h(D, counts, D->d.npair, W * 4));
    std::vector<int> ih(W * cap);
    HIP_TRY(d2h(D, ih.data(), D->iws + (size_t)D->L.pairs * W, ih.size() * 4));
    for (size_t w = 0; w < W; ++w)
        for (int k = 0; k < cap; ++k) {
            int idx = k < counts[w] ? ih[(size_t)k * W + w] : 0;
            pairs[(w * cap + k) * 2] = k < counts[w] ? M->h.pairs[idx].a : -1;
            pairs[(w * cap + k) * 2 + 1] = k < counts[w] ? M->h.pairs[idx].b : -1;
        }
    return NEWT_OK;
}

// ---- reset, keyframes, snapshots -------------------------------------------

}  // extern "C"

namespace {

// Worlds with mask[w] != 0 take the state in `src` ([nq + nv + 6 n_geoms]):
// qpos, qvel and, with `bounds`, the proxy bounds; with `clear_force` their
// applied force and qacc are zeroed.  Faults and counts are cleared.  The
// mask and source sit in pinned host memory, read by the kernel directly.
__global__ void k_reset(DataT d, int nq, int nv, int ng, const unsigned char* mask, const double* src, int bounds,
                        int clear_force) {
    const int w = blockIdx.x * blockDim.x + threadIdx.x;
    if (w >= d.W || !mask[w]) return;
    const size_t W = (size_t)d.W;
    for (int k = 0; k < nq; ++k) d.qpos[k * W + w] = src[k];
    for (int k = 0; k < nv; ++k) {
        d.qvel[k * W + w] = src[nq + k];
        if (clear_force) {
            d.qfrc[k * W + w] = 0.0;
            d.qacc[k * W + w] = 0.0;
        }
    }
    if (bounds)
        for (int k = 0; k < 6 * ng; ++k) d.fat[k * W + w] = src[nq + nv + k];
    d.fault[w] = 0;
    d.ncontact[w] = 0;
    d.npair[w] = 0;
}

newt_status reset_worlds(newt_data* D, const uint8_t* mask, const double* q, const double* v, bool defaults,
                         void* stream) {
    const ModelT& m = D->model->dev;
    HIP_TRY(stage(D));
    const size_t W = (size_t)D->W;
    unsigned char* hm = reinterpret_cast<unsigned char*>(D->staging);
    double* src = D->staging + (W + 7) / 8;
    for (size_t w = 0; w < W; ++w) hm[w] = mask ? (mask[w] != 0) : 1;
    for (int k = 0; k < m.nq; ++k) src[k] = q[k];
    for (int k = 0; k < m.nv; ++k) src[m.nq + k] = v[k];
    if (defaults)
        for (int k = 0; k < 6 * m.n_geoms; ++k) src[m.nq + m.nv + k] = D->fat_default[k];
    hipStream_t s = (hipStream_t)stream;
    hipLaunchKernelGGL(k_reset, dim3((D->W + 63) / 64), dim3(64), 0, s, D->d, m.nq, m.nv, m.n_geoms, hm, src,
                       defaults ? 1 : 0, defaults ? 1 : 0);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipEventRecord(D->staged, s));
    D->last_stream = s;
    return NEWT_OK;
}

}  // namespace

struct newt_snapshot {
    const newt_model* model = nullptr;
    int W = 0;
    DataT d{};
    uint64_t bytes = 0;
};

extern "C" {

newt_status newt_data_reset(newt_data* D, const uint8_t* world_mask, void* stream) {
    if (!D) return fail(NEWT_ERR_INVALID_ARG, "null data");
    const HostModel& H = D->model->h;
    return reset_worlds(D, world_mask, H.q_default.data(), H.v_default.data(), true, stream);
}

newt_status newt_data_reset_to_keyframe(newt_data* D, const uint8_t* world_mask, const double* qpos,
                                        const double* qvel, void* stream) {
    if (!D || !qpos || !qvel) return fail(NEWT_ERR_INVALID_ARG, "null argument");
    return reset_worlds(D, world_mask, qpos, qvel, false, stream);
}

newt_status newt_data_snapshot(const newt_data* D, newt_snapshot** out) {
    if (!D || !out) return fail(NEWT_ERR_INVALID_ARG, "null argument");
    const ModelT& m = D->model->dev;
    newt_snapshot* S = new newt_snapshot();
    S->model = D->model;
    S->W = D->W;
    S->d.W = D->W;
    const size_t W = (size_t)D->W;
    auto dup = [&](void** dst, const void* src, size_t bytes) -> hipError_t {
        bytes = std::max<size_t>(bytes, 8);
        hipError_t e = hipMalloc(dst, bytes);
        if (e != hipSuccess) return e;
        S->bytes += bytes;
        return hipMemcpyAsync(*dst, src, bytes, hipMemcpyDeviceToDevice, D->last_stream);
    };
    struct Item {
        void** dst;
        const void* src;
        size_t bytes;
    };
    const Item items[] = {{(void**)&S->d.qpos, D->d.qpos, W * m.nq * 8},
                          {(void**)&S->d.qvel, D->d.qvel, W * m.nv * 8},
                          {(void**)&S->d.qacc, D->d.qacc, W * m.nv * 8},
                          {(void**)&S->d.qfrc, D->d.qfrc, W * m.nv * 8},
                          {(void**)&S->d.fat, D->d.fat, W * m.n_geoms * 6 * 8},
                          {(void**)&S->d.fault, D->d.fault, W * 4},
                          {(void**)&S->d.ncontact, D->d.ncontact, W * 4},
                          {(void**)&S->d.npair, D->d.npair, W * 4}};
    hipError_t e = hipSuccess;
    for (const Item& it : items)
        if (e == hipSuccess) e = dup(it.dst, it.src, it.bytes);
    if (e == hipSuccess) e = hipStreamSynchronize(D->last_stream);
    if (e != hipSuccess) {
        newt_snapshot_destroy(S);
        return fail(NEWT_ERR_DEVICE, std::string("snapshot: ") + hipGetErrorString(e));
    }
    *out = S;
    return NEWT_OK;
}

newt_status newt_data_restore(newt_data* D, const newt_snapshot* S, void* stream) {
    if (!D || !S) return fail(NEWT_ERR_INVALID_ARG, "null argument");
    if (S->model != D->model || S->W != D->W)
        return fail(NEWT_ERR_INVALID_ARG, "the snapshot was taken of a data of another model or size");
    const ModelT& m = D->model->dev;
    const size_t W = (size_t)D->W;
    hipStream_t s = (hipStream_t)stream;
    auto cp = [&](void* dst, const void* src, size_t bytes) {
        return bytes ? hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, s) : hipSuccess;
    };
    HIP_TRY(cp(D->d.qpos, S->d.qpos, W * m.nq * 8));
    HIP_TRY(cp(D->d.qvel, S->d.qvel, W * m.nv * 8));
    HIP_TRY(cp(D->d.qacc, S->d.qacc, W * m.nv * 8));
    HIP_TRY(cp(D->d.qfrc, S->d.qfrc, W * m.nv * 8));
    HIP_TRY(cp(D->d.fat, S->d.fat, W * m.n_geoms * 6 * 8));
    HIP_TRY(cp(D->d.fault, S->d.fault, W * 4));
    HIP_TRY(cp(D->d.ncontact, S->d.ncontact, W * 4));
    HIP_TRY(cp(D->d.npair, S->d.npair, W * 4));
    D->last_stream = s;
    return NEWT_OK;
}

void newt_snapshot_destroy(newt_snapshot* S) {
    if (!S) return;
    for (void* p : {(void*)S->d.qpos, (void*)S->d.qvel, (void*)S->d.qacc, (void*)S->d.qfrc, (void*)S->d.fat,
                    (void*)S->d.fault, (void*)S->d.ncontact, (void*)S->d.npair})
        (void)hipFree(p);
    delete S;
}

newt_status newt_step(const newt_model* model, newt_data* data, int32_t n_steps, void* stream) {
    if (!model || !data || n_steps < 0) return fail(NEWT_ERR_INVALID_ARG, "bad step arguments");
    if (data->model != model) return fail(NEWT_ERR_INVALID_ARG, "data was created for another model");
    if (n_steps == 0) return NEWT_OK;
    newt_status st = launch(data, n_steps, (hipStream_t)stream, 0);
    if (st != NEWT_OK) return st;
    return launch_sensors(data, (hipStream_t)stream);
}

newt_status newt_data_synchronize(newt_data* data) {
    if (!data) return fail(NEWT_ERR_INVALID_ARG, "null data");
    HIP_TRY(hipStreamSynchronize(data->last_stream));
    return NEWT_OK;
}

}  // extern "C"
