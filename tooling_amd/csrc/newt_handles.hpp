// newt_handles.hpp — the opaque handles of include/newt_gpu.h (the compiled
// model and the device state of a batch), shared by the C-ABI units.
#pragma once

#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "../../include/newt_gpu.h"
#include "newt_compile.hpp"
#include "newt_step.hpp"
#include "newt_team.hpp"
#include "newt_bodies.hpp"
#include "newt_sensors.hpp"

namespace newt {
// Defined in newt_team.hip (the persistent-step instantiations).
hipError_t team_launch(bool mixed, int nv, bool ext, dim3 grid, dim3 block, int lds_bytes, hipStream_t stream,
                       const TeamJob* jobs, int n_jobs, int n_steps);
hipError_t team_set_lds(bool mixed, int nv, bool ext, int bytes);
int team_max_worlds_for(int nv);

// Records `msg` as this thread's newt_last_error() and returns `s`
// (newt_gpu.hip).
newt_status fail(newt_status s, const std::string& msg);

constexpr int kMaxLds = 160 * 1024;  // LDS of one gfx950 workgroup
}  // namespace newt

#define HIP_TRY(expr)                                                                       \
    do {                                                                                    \
        hipError_t e_ = (expr);                                                             \
        if (e_ != hipSuccess)                                                               \
            return newt::fail(NEWT_ERR_DEVICE, std::string(#expr ": ") + hipGetErrorString(e_)); \
    } while (0)

struct newt_model {
    newt::HostModel h;
    bool finalized = false;
    newt::ModelT dev{};  // device pointers
    newt::LinkT* d_links = nullptr;
    newt::GeomT* d_geoms = nullptr;
    newt::PairT* d_pairs = nullptr;
    int* d_chain = nullptr;
    int* d_level_off = nullptr;
    int* d_level_links = nullptr;
    int* d_level_nrank = nullptr;
    int* d_dof_link = nullptr;
    newt::LimitT* d_limits = nullptr;
    newt::BodyT* d_bodies = nullptr;
    newt::SensorT* d_sensors = nullptr;
    newt::MeshT* d_meshes = nullptr;
    double* d_mesh_v = nullptr;
    int* d_mesh_f = nullptr;
    newt::CacheLayout cache{};  // model tables as the persistent step caches them in LDS
    char* d_blob = nullptr;     // the same tables packed in that layout
};

struct newt_data {
    const newt_model* model = nullptr;
    int W = 0;
    int precision = 0;
    newt::Layout L{};
    newt::TeamLayout TL{};      // LDS layout of one world in the persistent step
    newt::BodyWaveLayout BL{};  // free-body models: one world's LDS block in the wavefront step
    double* bws = nullptr;      // and its per-world global fallback for contact rows
    int* bwi = nullptr;
    int wpb = 1;  // worlds (wavefronts) per workgroup of the persistent step
    int lds_bytes = 0;
    newt::DataT d{};
    void* ws = nullptr;     // real workspace (double or float)
    double* dws = nullptr;  // factorization workspace
    int* iws = nullptr;     // integer workspace
    char* ovf = nullptr;    // per-world blocks for contact sets over 64 rows (persistent step)
    uint64_t device_bytes = 0;        // device memory this data allocated
    std::vector<double> fat_default;  // fat bounds of one world at the default state
    double* staging = nullptr;        // pinned host copy of the last asynchronous state upload
    hipEvent_t staged = nullptr;      // completes when that upload has left the staging buffer
    hipStream_t last_stream = nullptr;
    newt::TeamJob* d_job = nullptr;   // the persistent step's job record (device)
};

namespace newt {
// The persistent step's job record of a newt_data (all of its worlds).
inline TeamJob team_job(const newt_data* D) {
    TeamJob j{};
    j.m = D->model->dev;
    j.T = D->TL;
    j.C = D->model->cache;
    j.blob = (const int4*)D->model->d_blob;
    j.L = D->L;
    j.d = D->d;
    j.dws = D->dws;
    j.iws = D->iws;
    j.ovf = D->ovf;
    j.wpb = D->wpb;
    j.first_block = 0;
    return j;
}
}  // namespace newt
