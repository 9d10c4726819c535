// newt_solve.hpp — wavefront-per-world contact solve: Delassus assembly and
// projected Gauss-Seidel for one world per 64-lane wavefront.
//
// Inputs (world-major hand-off block written by the lane-per-world pre
// kernel): the Cholesky factor L of the response matrix M + dt·B, the
// pyramidal Jacobian rows J (nr x nv), each row's bias and regularizer.
//
//   Y = L⁻¹ Jᵀ            lane j solves row j (FP64 forward substitution)
//   A = Yᵀ Y = J (LLᵀ)⁻¹ Jᵀ  lane j keeps row j of A in registers
//   PGS (solver.rs:3269-3331): for each sweep, for each row i in order,
//       f_i ← max(f_i − (r_i + R_i f_i)/(A_ii + R_i), 0),  r = b + A f.
//   Each lane j holds r_j and updates it by A_ji Δf_i after row i moves
//   (an incremental form of the reference's dense row dot product); row i's
//   values reach every lane through v_readlane.
//   qfrc = Jᵀ f / dt  (solver.rs:3345-3364)
//
// R is the storage type of Y and A (double for the FP64 step, float for the
// mixed step); residuals, impulses and all accumulation are FP64.
#pragma once

#include "newt_types.hpp"

namespace newt {

__device__ __forceinline__ double readlane_f64(double v, int lane) {
    int2 p = __builtin_bit_cast(int2, v);
    p.x = __builtin_amdgcn_readlane(p.x, lane);
    p.y = __builtin_amdgcn_readlane(p.y, lane);
    return __builtin_bit_cast(double, p);
}

template <class R, int NV>
__global__ void __launch_bounds__(64) k_solve(ModelT m, Layout L, const int* __restrict__ team_rows,
                                              char* __restrict__ sb, double* __restrict__ dws, int W) {
    const int w = blockIdx.x;
    const int lane = threadIdx.x;
    const int nr = team_rows[w];
    if (nr <= 0) return;
    const int nv = m.nv;
    char* base = sb + (size_t)w * L.sb_stride;
    const double* gL = (const double*)(base + L.sb_Li);
    const double* gbias = (const double*)(base + L.sb_bias);
    const double* greg = (const double*)(base + L.sb_reg);
    const R* gJ = (const R*)(base + L.sb_J);

    __shared__ double sL[NV * NV];
    __shared__ R sY[64 * NV];
    __shared__ double sf[64];
    for (int k = lane; k < nv * nv; k += 64) sL[k] = gL[k];
    for (int k = lane; k < nr * nv; k += 64) sY[k] = gJ[k];
    __syncthreads();

    // Y row of this lane: forward substitution L y = J_lane (FP64).
    double y[NV];
    const bool active = lane < nr;
    // Column-oriented substitution: every index is a compile-time constant,
    // so y stays in registers.
#pragma clang loop unroll(full)
    for (int s = 0; s < NV; ++s) y[s] = (s < nv && active) ? (double)sY[lane * nv + s] : 0.0;
#pragma clang loop unroll(full)
    for (int k = 0; k < NV; ++k) {
        if (k < nv) {
            y[k] = y[k] / sL[k * nv + k];
#pragma clang loop unroll(full)
            for (int s = k + 1; s < NV; ++s)
                if (s < nv) y[s] -= sL[s * nv + k] * y[k];
        }
    }
    __syncthreads();
#pragma clang loop unroll(full)
    for (int s = 0; s < NV; ++s)
        if (active && s < nv) sY[lane * nv + s] = (R)y[s];
    __syncthreads();
    // Row `lane` of A from the stored (R-rounded) Y, read back from LDS.
    R a[64];
    const R* yl = sY + lane * nv;
#pragma unroll
    for (int i = 0; i < 64; ++i) {
        double acc = 0.0;
        if (i < nr && active) {
            const R* yi = sY + i * nv;
            for (int s = 0; s < nv; ++s) acc += (double)yl[s] * (double)yi[s];
        }
        a[i] = (R)acc;
    }
    double diag = 0.0;
    if (active)
        for (int s = 0; s < nv; ++s) diag += (double)yl[s] * (double)yl[s];
    diag = (double)(R)diag;
    const double reg = active ? greg[lane] : 0.0;
    const double invd = active ? 1.0 / (diag + reg) : 0.0;
    double r = active ? gbias[lane] : 0.0;
    double f = 0.0;
    const int iters = m.iterations;
    for (int it = 0; it < iters; ++it) {
#pragma unroll
        for (int i = 0; i < 64; ++i) {
            if (i < nr) {
                const double ri = readlane_f64(r, i);
                const double fi = readlane_f64(f, i);
                const double Ri = readlane_f64(reg, i);
                const double di = readlane_f64(invd, i);
                double fn = fi - (ri + Ri * fi) * di;
                fn = fn > 0.0 ? fn : 0.0;
                const double dl = fn - fi;
                if (lane == i) f = fn;
                r += (double)a[i] * dl;
            }
        }
    }
    // qfrc = Jᵀ f / dt, written into the lane-interleaved delta region.
    sf[lane] = f;
    __syncthreads();
    if (lane < nv) {
        const double dt = m.dt;
        double acc = 0.0;
        for (int j = 0; j < nr; ++j) acc += (double)gJ[j * nv + lane] * (sf[j] / dt);
        dws[(size_t)(L.delta + lane) * W + w] = acc;
    }
}

}  // namespace newt
