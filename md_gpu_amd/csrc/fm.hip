// fm.hip -- adjoint of the pair force for force matching (no reference
// counterpart; the forward pass is the reference's pair kernel,
// src/pair_forces.cu:133-204, run by k_step in force-only mode with
// charges scaled by qscale).
//
// loss = (1/M) sum_frames sum_i |F_i - F*_i|^2,  M = 3 * atoms * frames.
// With R_i = F_i - F*_i and d_ij = x_i - x_j (full list, ordered pairs):
//   dloss/dtheta = (2/M) sum_i sum_j (R_i . d_ij) d ff_ij / d theta
//   d ff / d lj1[ti,tj] = r^-14,  d ff / d lj2[ti,tj] = -r^-8,
//   d ff / d qscale    = 2 qscale qfac q_i q_j C(r) / r^2   (C: Wolf / RF / damped factor)
// Per-pair terms are FP32; each atom's partial sums are promoted to FP64
// before they are accumulated across atoms, workgroups and frames.
// Algorithmic work per ordered pair: forward 65 flop + adjoint ~45 flop.
#include "md_device.h"

namespace {
__device__ __forceinline__ float minimg(float d, float L, float Linv) {
    return fmaf(-L, rintf(d * Linv), d);
}
}  // namespace

template <int G, int COUL, int NT>
__global__ __launch_bounds__(256) void k_fm_adjoint(const StepArgs a, double* gacc) {
    constexpr int APW = 64 / G, APB = 4 * APW;
    constexpr int NACC = 2 + 2 * NT * NT;
    __shared__ double s_red[4][NACC];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int sub = lane & (G - 1), grp = lane / G;
    const int u0 = (int)(((long long)blockIdx.x * a.nunits) / gridDim.x);
    const int u1 = (int)(((long long)(blockIdx.x + 1) * a.nunits) / gridDim.x);
    double acc[NACC];
#pragma unroll
    for (int k = 0; k < NACC; k++) acc[k] = 0.0;
    for (int u = u0; u < u1; u++) {
        const int slot0 = u * APB;
        const int r = __builtin_amdgcn_readfirstlane(a.tile_rep[slot0 >> 6]);
        const RepParam& p = a.rep[r];
        const float4* __restrict__ pin = a.pos[cur_buf(a.st[r])];
        const float L0 = p.L[0], L1 = p.L[1], L2 = p.L[2];
        const float Li0 = p.Linv[0], Li1 = p.Linv[1], Li2 = p.Linv[2];
        const int slot = slot0 + wave * APW + grp;
        const bool active = a.ids[slot] >= 0;
        if (!active) continue;
        const float4 xi = pin[slot];
        const int ti = a.typ[slot];
        const int nn = min(a.ncount[slot], a.cap);
        const float4 F = a.force[slot], T = a.target[slot];
        const float Rx = F.x - T.x, Ry = F.y - T.y, Rz = F.z - T.z;
        float s1[NT], s2[NT], sq = 0.0f;
#pragma unroll
        for (int t = 0; t < NT; t++) s1[t] = s2[t] = 0.0f;
        const uint32_t* row = a.nlist + (size_t)slot * a.cap;
        for (int k = sub; k < nn; k += G) {
            const uint32_t e = row[lpos(k, G)];   // rows are laid out for G lanes (the step kernel's)
            const int j = (int)(e & MD_JMASK);
            const int tj = (int)(e >> MD_JBITS);
            const float4 pj = pin[j];
            const float dx = minimg(xi.x - pj.x, L0, Li0);
            const float dy = minimg(xi.y - pj.y, L1, Li1);
            const float dz = minimg(xi.z - pj.z, L2, Li2);
            const float rsq = fmaf(dz, dz, fmaf(dy, dy, dx * dx));
            const bool inside = !(rsq > a.rc2);
            const float g = inside ? fmaf(Rz, dz, fmaf(Ry, dy, Rx * dx)) : 0.0f;
            const float ri = __builtin_amdgcn_rsqf(rsq);
            const float ri2 = ri * ri;
            const float ri8 = (ri2 * ri2) * (ri2 * ri2);
            const float t14 = g * ri8 * ri2 * ri2 * ri2;
            const float t8 = g * ri8;
#pragma unroll
            for (int t = 0; t < NT; t++) {
                s1[t] += (tj == t) ? t14 : 0.0f;
                s2[t] -= (tj == t) ? t8 : 0.0f;
            }
            float c = 0.0f;
            if (COUL == 0) {
                c = ri - a.reacf * rsq;
            } else if (COUL == 1 || COUL == 2) {
                const float rr = rsq * ri;
                const float t = __builtin_amdgcn_rcpf(fmaf(a.palfa, rr, 1.0f));
                const float poly = ((((1.061405429f * t - 1.453152027f) * t + 1.421413741f) * t - 0.284496736f) * t +
                                    0.254829592f) * t;
                c = fmaf(a.spiialfa, rr, poly) * __builtin_amdgcn_exp2f(a.alfasq * rsq) * ri;
                if (COUL == 1) c -= a.wolf;
            }
            if (COUL != 3) sq = fmaf(g * pj.w, c * ri2, sq);
        }
        // fold the atom's partials into FP64 (all lanes of the group hold parts)
        double* a1 = acc + 2;
        double* a2 = acc + 2 + NT * NT;
#pragma unroll
        for (int t = 0; t < NT; t++) {
#pragma unroll
            for (int ta = 0; ta < NT; ta++) {
                if (ta == ti) {
                    a1[ta * NT + t] += (double)s1[t];
                    a2[ta * NT + t] += (double)s2[t];
                }
            }
        }
        acc[1] += (double)sq * 2.0 * (double)a.qq_scale * (double)a.qfac * (double)xi.w;
        if (sub == 0) acc[0] += (double)Rx * Rx + (double)Ry * Ry + (double)Rz * Rz;
    }
#pragma unroll
    for (int k = 0; k < NACC; k++)
        for (int o = 32; o >= 1; o >>= 1) acc[k] += __shfl_xor(acc[k], o, 64);
    if (lane == 0)
        for (int k = 0; k < NACC; k++) s_red[wave][k] = acc[k];
    __syncthreads();
    if (tid < NACC) {
        double t = s_red[0][tid] + s_red[1][tid] + s_red[2][tid] + s_red[3][tid];
        if (t != 0.0) atomicAdd(gacc + tid, t);
    }
}

template <int G, int NT>
static void launch_fm_g(const StepArgs& a, double* gacc, int coul, int grid, hipStream_t st) {
    switch (coul) {
    case 0: hipLaunchKernelGGL((k_fm_adjoint<G, 0, NT>), dim3(grid), dim3(256), 0, st, a, gacc); break;
    case 1: hipLaunchKernelGGL((k_fm_adjoint<G, 1, NT>), dim3(grid), dim3(256), 0, st, a, gacc); break;
    case 2: hipLaunchKernelGGL((k_fm_adjoint<G, 2, NT>), dim3(grid), dim3(256), 0, st, a, gacc); break;
    default: hipLaunchKernelGGL((k_fm_adjoint<G, 3, NT>), dim3(grid), dim3(256), 0, st, a, gacc); break;
    }
}

template <int NT>
static hipError_t launch_fm_nt(const StepArgs& a, double* gacc, int lanes, int coul, int grid, hipStream_t st) {
    switch (lanes) {
    case 4: launch_fm_g<4, NT>(a, gacc, coul, grid, st); break;
    case 8: launch_fm_g<8, NT>(a, gacc, coul, grid, st); break;
    case 16: launch_fm_g<16, NT>(a, gacc, coul, grid, st); break;
    case 32: launch_fm_g<32, NT>(a, gacc, coul, grid, st); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

// ntype must be 1, 2 or 4 (the host rounds up; unused type rows stay zero)
extern "C" hipError_t md_launch_fm_adjoint(const StepArgs& a, double* gacc, int lanes, int coul, int nt, int grid,
                                           hipStream_t st) {
    if (nt == 1) return launch_fm_nt<1>(a, gacc, lanes, coul, grid, st);
    if (nt == 2) return launch_fm_nt<2>(a, gacc, lanes, coul, grid, st);
    if (nt == 4) return launch_fm_nt<4>(a, gacc, lanes, coul, grid, st);
    return hipErrorInvalidValue;
}
