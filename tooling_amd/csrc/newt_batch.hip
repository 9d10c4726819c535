// newt_batch.hip — one step launch over several newt_datas of different
// models (an uneven batch: trees of different dof counts and contact loads).
//
// The persistent step reads its parameters from a table of jobs, one per
// newt_data, each owning a run of workgroups (newt_team.hpp, TeamJob).  A
// batch groups its datas by kernel instantiation (precision, register width
// 24/32/64, extended step) and makes one launch per group; with use_graph the
// launches of one newt_batch_step are captured once per step count into a HIP
// graph (groups on parallel branches when `branches` is set) and replayed.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <vector>

#include "newt_handles.hpp"

using namespace newt;

namespace {

struct Group {
    bool mixed = false, ext = false;
    int width = 24;      // register width of the instantiation
    int max_wpb = 1;     // workgroup size in wavefronts
    int lds = 0;         // dynamic LDS of the launch
    int blocks = 0;      // workgroups of the launch
    std::vector<TeamJob> jobs;
    TeamJob* d_jobs = nullptr;
};

int width_of(int nv) { return nv <= 24 ? 24 : nv <= 32 ? 32 : 64; }

}  // namespace

struct newt_batch {
    std::vector<newt_data*> datas;
    newt_batch_options opt{};
    std::vector<Group> groups;
    std::vector<hipStream_t> side;   // capture branches
    std::vector<hipEvent_t> events;  // fork/join events of the capture
    std::vector<std::pair<int, hipGraphExec_t>> graphs;  // per step count
    uint64_t device_bytes = 0;
};

namespace {

newt_status launch_groups(newt_batch* B, int n_steps, hipStream_t stream, bool capture) {
    const bool branch = capture && B->opt.branches && B->groups.size() > 1;
    if (branch) HIP_TRY(hipEventRecord(B->events[0], stream));
    for (size_t g = 0; g < B->groups.size(); ++g) {
        const Group& G = B->groups[g];
        hipStream_t s = stream;
        if (branch && g > 0) {
            s = B->side[g - 1];
            HIP_TRY(hipStreamWaitEvent(s, B->events[0], 0));
        }
        HIP_TRY(team_launch(G.mixed, G.width, G.ext, dim3(G.blocks), dim3(64 * G.max_wpb), G.lds, s, G.d_jobs,
                            (int)G.jobs.size(), n_steps));
        if (branch && g > 0) {
            HIP_TRY(hipEventRecord(B->events[g], s));
            HIP_TRY(hipStreamWaitEvent(stream, B->events[g], 0));
        }
    }
    return NEWT_OK;
}

}  // namespace

extern "C" {

newt_status newt_batch_create(newt_data* const* datas, int32_t n, const newt_batch_options* options,
                              newt_batch** out) {
    if (!datas || n < 1 || !out) return fail(NEWT_ERR_INVALID_ARG, "bad batch arguments");
    newt_batch_options opt{};
    opt.use_graph = 1;
    opt.branches = 1;
    opt.order = 1;
    if (options) opt = *options;
    for (int i = 0; i < n; ++i) {
        const newt_data* D = datas[i];
        if (!D) return fail(NEWT_ERR_INVALID_ARG, "null data in batch");
        if (!D->d_job)
            return fail(NEWT_ERR_UNSUPPORTED, "a batch steps datas of the persistent step only (tree models, "
                                              "FP64 or mixed precision, no penalty joint limits)");
        for (int k = 0; k < i; ++k)
            if (datas[k] == D) return fail(NEWT_ERR_INVALID_ARG, "a data appears twice in the batch");
    }
    newt_batch* B = new newt_batch();
    B->datas.assign(datas, datas + n);
    B->opt = opt;
    std::vector<int> idx(n);
    for (int i = 0; i < n; ++i) idx[i] = i;
    if (opt.order == 1)  // the deepest trees start first: they bound the launch
        std::stable_sort(idx.begin(), idx.end(),
                         [&](int a, int b) { return datas[a]->model->dev.nv > datas[b]->model->dev.nv; });
    for (int i : idx) {
        newt_data* D = datas[i];
        const ModelT& m = D->model->dev;
        const bool mixed = D->precision == NEWT_PRECISION_MIXED;
        const int width = width_of(m.nv);
        const bool ext = m.ext != 0;
        Group* G = nullptr;
        for (auto& g : B->groups)
            if (g.mixed == mixed && g.width == width && g.ext == ext) G = &g;
        if (!G) {
            B->groups.emplace_back();
            G = &B->groups.back();
            G->mixed = mixed;
            G->width = width;
            G->ext = ext;
        }
        const int wi = width == 24 ? 0 : width == 32 ? 1 : 2;
        int wpb = D->wpb;
        if (opt.wpb_cap[wi] > 0) wpb = std::min(wpb, (int)opt.wpb_cap[wi]);
        TeamJob j = team_job(D);
        j.wpb = wpb;
        j.first_block = G->blocks;
        G->blocks += (D->W + wpb - 1) / wpb;
        G->max_wpb = std::max(G->max_wpb, wpb);
        G->lds = std::max(G->lds, D->model->cache.bytes + wpb * D->TL.bytes);
        G->jobs.push_back(j);
    }
    auto bail = [&](newt_status s) {
        newt_batch_destroy(B);
        return s;
    };
    for (auto& G : B->groups) {
        const size_t bytes = sizeof(TeamJob) * G.jobs.size();
        if (hipMalloc(&G.d_jobs, bytes) != hipSuccess) return bail(fail(NEWT_ERR_OOM, "device allocation failed"));
        B->device_bytes += bytes;
        if (hipMemcpy(G.d_jobs, G.jobs.data(), bytes, hipMemcpyHostToDevice) != hipSuccess)
            return bail(fail(NEWT_ERR_DEVICE, "hipMemcpy of the job table failed"));
        if (G.lds > 65536 && team_set_lds(G.mixed, G.width, G.ext, kMaxLds) != hipSuccess)
            return bail(fail(NEWT_ERR_DEVICE, "hipFuncSetAttribute failed"));
    }
    for (size_t g = 0; g < B->groups.size(); ++g) {
        hipEvent_t e = nullptr;
        if (hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess)
            return bail(fail(NEWT_ERR_DEVICE, "hipEventCreate failed"));
        B->events.push_back(e);
        if (g > 0) {
            hipStream_t s = nullptr;
            if (hipStreamCreateWithFlags(&s, hipStreamNonBlocking) != hipSuccess)
                return bail(fail(NEWT_ERR_DEVICE, "hipStreamCreate failed"));
            B->side.push_back(s);
        }
    }
    *out = B;
    return NEWT_OK;
}

newt_status newt_batch_step(newt_batch* B, int32_t n_steps, void* stream) {
    if (!B || n_steps < 0) return fail(NEWT_ERR_INVALID_ARG, "bad batch step arguments");
    if (n_steps == 0) return NEWT_OK;
    hipStream_t s = (hipStream_t)stream;
    for (newt_data* D : B->datas) D->last_stream = s;
    if (!B->opt.use_graph) return launch_groups(B, n_steps, s, false);
    hipGraphExec_t exec = nullptr;
    for (auto& g : B->graphs)
        if (g.first == n_steps) exec = g.second;
    if (!exec) {
        // Captured on a private stream, so the caller's stream may be the
        // legacy default stream (which cannot be captured).
        hipStream_t cs = nullptr;
        HIP_TRY(hipStreamCreateWithFlags(&cs, hipStreamNonBlocking));
        hipGraph_t graph = nullptr;
        hipError_t e = hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal);
        newt_status st = NEWT_OK;
        if (e == hipSuccess) {
            st = launch_groups(B, n_steps, cs, true);
            hipError_t e2 = hipStreamEndCapture(cs, &graph);
            if (st == NEWT_OK && e2 != hipSuccess) e = e2;
        }
        if (e == hipSuccess && st == NEWT_OK) e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
        if (graph) (void)hipGraphDestroy(graph);
        (void)hipStreamDestroy(cs);
        if (st != NEWT_OK) return st;
        if (e != hipSuccess) return fail(NEWT_ERR_DEVICE, std::string("graph capture: ") + hipGetErrorString(e));
        B->graphs.emplace_back(n_steps, exec);
    }
    HIP_TRY(hipGraphLaunch(exec, s));
    return NEWT_OK;
}

newt_status newt_batch_launches(const newt_batch* B, int32_t* n_launches) {
    if (!B || !n_launches) return fail(NEWT_ERR_INVALID_ARG, "null argument");
    *n_launches = (int32_t)B->groups.size();
    return NEWT_OK;
}

void newt_batch_destroy(newt_batch* B) {
    if (!B) return;
    for (auto& g : B->graphs) (void)hipGraphExecDestroy(g.second);
    for (auto& G : B->groups) (void)hipFree(G.d_jobs);
    for (hipStream_t s : B->side) (void)hipStreamDestroy(s);
    for (hipEvent_t e : B->events) (void)hipEventDestroy(e);
    delete B;
}

}  // extern "C"
