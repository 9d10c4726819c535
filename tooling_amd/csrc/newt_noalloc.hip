// newt_noalloc.hip — newt_step_noalloc: the persistent step on caller-owned
// device state, with its scratch in a static device array, allocating
// nothing.
//
// The scratch holds what newt_data_create allocates beside the state for the
// persistent step: the launch's job record, the last step's contacts
// (position, normal, depth, friction, gap; geom ids and slot) and pair list,
// and the global block of Jacobian rows for contact sets over 64 rows.
// Nothing in it must start at zero: every call writes the job record first,
// and every step writes what it reads.  It is one array per GPU, shared by
// every call on that GPU, so calls that overlap on one GPU race.
#include <hip/hip_runtime.h>

#include <mutex>

#include "newt_handles.hpp"

using namespace newt;

namespace {

constexpr size_t kScratchBytes = size_t(160) << 20;  // 160 MiB per GPU
constexpr int kMaxDevices = 64;

__device__ __align__(256) char scratch[kScratchBytes];

std::mutex g_scratch_mu;
char* g_scratch[kMaxDevices] = {};

size_t align256(size_t b) { return (b + 255) / 256 * 256; }

// Writes the launch's job record into the scratch array in stream order:
// the record travels as this kernel's argument, so no copy is needed.
__global__ void k_put_job(TeamJob job, TeamJob* dst) {
    constexpr int n = (int)(sizeof(TeamJob) / 4);
    const int* src = reinterpret_cast<const int*>(&job);
    int* out = reinterpret_cast<int*>(dst);
    for (int k = threadIdx.x; k < n; k += blockDim.x) out[k] = src[k];
}
static_assert(sizeof(TeamJob) % 4 == 0, "job record is copied as 32-bit words");

// Byte sizes of the scratch regions for `W` worlds: the job record, contact geometry
// [9 cap_c][W] doubles, contact ids and pairs [3 cap_c + n_pairs][W] ints,
// and W blocks of the over-64-row Jacobian rows.
struct ScratchPlan {
    size_t job, geo, ids, rows, total;
};

ScratchPlan plan(const ModelT& m, const TeamLayout& T, size_t W) {
    ScratchPlan p;
    p.job = align256(sizeof(TeamJob));
    p.geo = align256(W * (size_t)std::max(9 * m.cap_c, 1) * 8);
    p.ids = align256(W * (size_t)std::max(3 * m.cap_c + m.n_pairs, 1) * 4);
    p.rows = W * (size_t)std::max(T.ov_stride, 16);
    p.total = p.job + p.geo + p.ids + p.rows;
    return p;
}

// The device address of this GPU's scratch array, looked up once per device
// with that device current.  Host code uses this address, never the
// array's name.
newt_status scratch_ptr(char** out) {
    int dev = 0;
    HIP_TRY(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) return fail(NEWT_ERR_DEVICE, "device index out of range");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch[dev]) {
        void* p = nullptr;
        HIP_TRY(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        g_scratch[dev] = static_cast<char*>(p);
    }
    *out = g_scratch[dev];
    return NEWT_OK;
}

// Shared checks; the team layout and worlds per workgroup of the launch.
newt_status check(const newt_model* model, int32_t n_worlds, int32_t precision, TeamLayout* T, int* wpb) {
    if (!model || !model->finalized) return fail(NEWT_ERR_INVALID_ARG, "model is not finalized");
    if (n_worlds < 1) return fail(NEWT_ERR_INVALID_ARG, "n_worlds must be at least 1");
    if (precision != NEWT_PRECISION_FP64 && precision != NEWT_PRECISION_MIXED)
        return fail(NEWT_ERR_UNSUPPORTED, "newt_step_noalloc runs the FP64 and mixed persistent step only");
    const ModelT& m = model->dev;
    if (m.n_bodies > 0 || lane_only(m))
        return fail(NEWT_ERR_UNSUPPORTED, "newt_step_noalloc steps tree models on the persistent step only "
                                          "(one tree, no free bodies, no penalty joint limits, PGS or penalty)");
    if (m.nv > 64 || m.n_links > 128)
        return fail(NEWT_ERR_UNSUPPORTED, "the persistent step handles nv <= 64 and <= 128 links");
    *T = make_team_layout(m, precision == NEWT_PRECISION_MIXED ? 4 : 8);
    *wpb = std::min(team_max_worlds_for(m.nv), (kMaxLds - model->cache.bytes) / std::max(T->bytes, 1));
    if (*wpb < 1) return fail(NEWT_ERR_UNSUPPORTED, "one world's step state exceeds the LDS of a workgroup");
    return NEWT_OK;
}

}  // namespace

extern "C" {

newt_status newt_noalloc_max_worlds(const newt_model* model, int32_t precision, int32_t* max_worlds) {
    if (!max_worlds) return fail(NEWT_ERR_INVALID_ARG, "null argument");
    TeamLayout T;
    int wpb = 0;
    newt_status s = check(model, 1, precision, &T, &wpb);
    if (s != NEWT_OK) return s;
    size_t lo = 0, hi = size_t(1) << 31;  // largest W whose plan fits
    while (hi - lo > 1) {
        const size_t mid = (lo + hi) / 2;
        (plan(model->dev, T, mid).total <= kScratchBytes ? lo : hi) = mid;
    }
    *max_worlds = (int32_t)std::min<size_t>(lo, 0x7fffffff);
    return NEWT_OK;
}

newt_status newt_step_noalloc(const newt_model* model, const newt_device_state* state, int32_t n_worlds,
                              int32_t precision, int32_t reset_bounds, int32_t n_steps, void* stream) {
    TeamLayout T;
    int wpb = 0;
    newt_status s = check(model, n_worlds, precision, &T, &wpb);
    if (s != NEWT_OK) return s;
    if (!state || n_steps < 0) return fail(NEWT_ERR_INVALID_ARG, "bad step arguments");
    const ModelT& m = model->dev;
    const ScratchPlan p = plan(m, T, (size_t)n_worlds);
    if (p.total > kScratchBytes)
        return fail(NEWT_ERR_CAPACITY, "newt_step_noalloc: this batch needs more than the 160 MiB scratch "
                                       "array; see newt_noalloc_max_worlds");
    if ((m.nq && !state->qpos) || (m.nv && (!state->qvel || !state->qacc || !state->qfrc)) ||
        (m.n_geoms && !state->fat) || !state->fault || !state->ncontact || !state->npair)
        return fail(NEWT_ERR_INVALID_ARG, "a state array the model needs is null");
    if (n_steps == 0 && !reset_bounds) return NEWT_OK;
    char* base = nullptr;
    s = scratch_ptr(&base);
    if (s != NEWT_OK) return s;
    const bool mixed = precision == NEWT_PRECISION_MIXED;
    const int lds = model->cache.bytes + wpb * T.bytes;
    if (lds > 65536) HIP_TRY(team_set_lds(mixed, m.nv, m.ext != 0, kMaxLds));
    DataT d{};
    d.W = n_worlds;
    d.qpos = state->qpos;
    d.qvel = state->qvel;
    d.qacc = state->qacc;
    d.qfrc = state->qfrc;
    d.fat = state->fat;
    d.fault = state->fault;
    d.ncontact = state->ncontact;
    d.npair = state->npair;
    d.init_bounds = reset_bounds != 0;
    // The persistent step reads only these three offsets of the workspace
    // layout; here they index the compact scratch regions.
    Layout L{};
    L.cgeo = 0;
    L.cids = 0;
    L.pairs = 3 * m.cap_c;
    TeamJob job{};
    job.m = m;
    job.T = T;
    job.C = model->cache;
    job.blob = (const int4*)model->d_blob;
    job.L = L;
    job.d = d;
    job.dws = reinterpret_cast<double*>(base + p.job);
    job.iws = reinterpret_cast<int*>(base + p.job + p.geo);
    job.ovf = base + p.job + p.geo + p.ids;
    job.wpb = wpb;
    job.first_block = 0;
    TeamJob* d_job = reinterpret_cast<TeamJob*>(base);
    hipLaunchKernelGGL(k_put_job, dim3(1), dim3(64), 0, (hipStream_t)stream, job, d_job);
    HIP_TRY(hipGetLastError());
    dim3 grid((n_worlds + wpb - 1) / wpb), block(64 * wpb);
    HIP_TRY(team_launch(mixed, m.nv, m.ext != 0, grid, block, lds, (hipStream_t)stream, d_job, 1, n_steps));
    return NEWT_OK;
}

}  // extern "C"
