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
// before they are accumulated across atoms, workgroups and frames.  The
// Coulomb factor is the force pass's cubic table (Wolf / damped).
// Work decomposition as the step kernel: one workgroup per 64-slot tile (a
// frame is a replica), the tile's neighbor set staged in LDS, G lanes per atom
// walking its row (md_rows.h).  Each workgroup stores its FP64 sums in its own
// slot of part (no atomics); k_fm_reduce adds the slots.
// Algorithmic work per ordered pair: forward 65 flop + adjoint ~45 flop.
#include "md_rows.h"

namespace {

// Sum over the G lanes of each aligned lane group (all lanes get the sum).
template <int G>
__device__ __forceinline__ float lane_group_sum(float x) {
#pragma unroll
    for (int o = G / 2; o >= 1; o >>= 1) x += __shfl_xor(x, o, 64);
    return x;
}

// Common prologue of both kernels: the tables and the tile's set in LDS.
template <int COUL>
__device__ __forceinline__ void fm_stage(const StepArgs& a, int tile, int r, float4* s_j, float4* s_lj, float4* s_ct,
                                         float4* s_sh, const float4* __restrict__ pin) {
    const int tid = threadIdx.x;
    const RepParam& p = a.rep[r];
    if (tid < 64) s_lj[tid] = a.ljtab[tid];
    if (COUL == 1 || COUL == 2) {
        if (tid < 256) s_ct[tid] = a.ctab[tid];
        if (tid == 0) s_ct[MD_CTAB] = a.ctab[MD_CTAB];
    }
    stage_shift_table(s_sh, p.L[0], p.L[1], p.L[2]);
    __syncthreads();
    stage_tile_set(s_j, s_sh, pin, a.jset + (size_t)tile * a.jcap, min(a.jcount[tile], a.jcap));
    __syncthreads();
}

// FP64 totals of the workgroup's per-thread accumulators into part[tile][.]
template <int NW, int NACC>
__device__ __forceinline__ void fm_flush(double* acc, double (*s_red)[NACC], double* part, int tile) {
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
#pragma unroll
    for (int k = 0; k < NACC; k++)
        for (int o = 32; o >= 1; o >>= 1) acc[k] += __shfl_xor(acc[k], o, 64);
    if (lane == 0)
        for (int k = 0; k < NACC; k++) s_red[wave][k] = acc[k];
    __syncthreads();
    if (tid < NACC) {
        double t = s_red[0][tid];
        for (int w = 1; w < NW; w++) t += s_red[w][tid];
        part[(size_t)tile * NACC + tid] = t;
    }
}

}  // namespace

// Adjoint walk after a force pass (the force of the trial parameters is in
// a.force).  The Coulomb factor C(r)/r^2 comes from the same cubic table as the
// force pass (Wolf / damped), so the gradient is that of the force the GPU computes.
template <int G, int COUL, int NT>
__global__ __launch_bounds__(64 * G) void k_fm_adjoint(const StepArgs a, double* part) {
    constexpr int NW = G;
    constexpr int NACC = 2 + 2 * NT * NT;
    extern __shared__ float4 s_j[];
    __shared__ double s_red[NW][NACC];
    __shared__ float4 s_lj[64];
    __shared__ float4 s_ct[(COUL == 1 || COUL == 2) ? MD_CTAB + 1 : 1];
    __shared__ float4 s_sh[27];
    const int tid = threadIdx.x;
    const int sub = tid & (G - 1), ai = tid / G;
    const int tile = xcd_block(blockIdx.x, gridDim.x);
    const int slot = tile * 64 + ai;
    const int r = __builtin_amdgcn_readfirstlane(a.tile_rep[tile]);
    const float4* __restrict__ pin = a.pos[cur_buf(a.st[r])];
    fm_stage<COUL>(a, tile, r, s_j, s_lj, s_ct, s_sh, pin);
    double acc[NACC];
#pragma unroll
    for (int k = 0; k < NACC; k++) acc[k] = 0.0;
    const int id = a.ids[slot];
    const bool active = id >= 0;
    const float4 xi = pin[slot];
    const int ti = active ? a.typ[slot] : 0;
    const int nn = active ? min(a.ncount[slot], a.cap) : 0;
    const float4 F = a.force[slot], T = a.target[slot];
    const float Rx = F.x - T.x, Ry = F.y - T.y, Rz = F.z - T.z;
    float s1[NT], s2[NT], sq = 0.0f;
#pragma unroll
    for (int t = 0; t < NT; t++) s1[t] = s2[t] = 0.0f;
    const uint4* __restrict__ row4 = reinterpret_cast<const uint4*>(a.nl16 + (size_t)slot * a.cap);
    walk_row<G>(s_j, row4, nn, sub, [&](const float4& pj) {
        const int tj = staged_type(pj);
        const float dx = xi.x - pj.x, dy = xi.y - pj.y, dz = xi.z - pj.z;
        const float rsq = fmaf(dz, dz, fmaf(dy, dy, dx * dx));
        const bool inside = !(rsq > a.rc2);
        const float g = inside ? fmaf(Rz, dz, fmaf(Ry, dy, Rx * dx)) : 0.0f;
        const float ri = __builtin_amdgcn_rsqf(rsq);
        const float ri2 = ri * ri;
        const float ri8 = (ri2 * ri2) * (ri2 * ri2);
        const float t8 = g * ri8;
        const float t14 = t8 * ri2 * ri2 * ri2;
#pragma unroll
        for (int t = 0; t < NT; t++) {
            s1[t] += (tj == t) ? t14 : 0.0f;
            s2[t] -= (tj == t) ? t8 : 0.0f;
        }
        if (COUL == 0) {
            sq = fmaf(g * pj.w, (ri - a.reacf * rsq) * ri2, sq);
        } else if (COUL == 1 || COUL == 2) {
            const float hv = ctab_eval(s_ct, (rsq * ri) * a.ct_scale);
            sq = fmaf(g * pj.w, hv * (ri2 * ri), sq);   // C(r) / r^2, as in the force pass
        }
    });
    if (active) {
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
    fm_flush<NW, NACC>(acc, s_red, part, tile);
}

// Single-pass form (systems without bonded terms): the rows are walked once for
// the force with the trial parameters and, on the same lanes, for the basis
// sums whose dot products with the residual R_i = F_i - F*_i give the gradient:
//   b14[t] = sum_{j of type t} r^-14 d_ij,   b8[t] = sum_{j of type t} r^-8 d_ij,
//   bc = sum_j q_j C(r)/r^2 d_ij,
//   dloss/dlj1[ti,t] = (2/M) sum_i R_i . b14[t],  dloss/dlj2[ti,t] = -(2/M) sum_i R_i . b8[t],
//   dloss/dqscale = (2/M) sum_i 2 qscale qfac q_i R_i . bc.
// R_i . b is linear in b, so each lane dots the residual (known to all lanes of
// the atom after the force reduction) with its own partial sums; a.qq_scale
// holds qscale.
template <int G, int COUL, int NT>
__global__ __launch_bounds__(64 * G) void k_fm_fused(const StepArgs a, double* part) {
    constexpr int NW = G;
    constexpr int NACC = 2 + 2 * NT * NT;
    extern __shared__ float4 s_j[];
    __shared__ double s_red[NW][NACC];
    __shared__ float4 s_lj[64];
    __shared__ float4 s_ct[(COUL == 1 || COUL == 2) ? MD_CTAB + 1 : 1];
    __shared__ float4 s_sh[27];
    const int tid = threadIdx.x;
    const int sub = tid & (G - 1), ai = tid / G;
    const int tile = xcd_block(blockIdx.x, gridDim.x);
    const int slot = tile * 64 + ai;
    const int r = __builtin_amdgcn_readfirstlane(a.tile_rep[tile]);
    const float4* __restrict__ pin = a.pos[cur_buf(a.st[r])];
    fm_stage<COUL>(a, tile, r, s_j, s_lj, s_ct, s_sh, pin);
    const float qs = a.qq_scale;
    double acc[NACC];
#pragma unroll
    for (int k = 0; k < NACC; k++) acc[k] = 0.0;
    const int id = a.ids[slot];
    const bool active = id >= 0;
    const float4 xi = pin[slot];
    const int ti = active ? a.typ[slot] : 0;
    const int nn = active ? min(a.ncount[slot], a.cap) : 0;
    const float qi = xi.w * a.qfac * (qs * qs);
    const float4* s_lji = s_lj + ti * 8;
    float f[3] = {0.0f, 0.0f, 0.0f}, bc[3] = {0.0f, 0.0f, 0.0f}, b14[NT][3], b8[NT][3];
#pragma unroll
    for (int t = 0; t < NT; t++)
#pragma unroll
        for (int c = 0; c < 3; c++) b14[t][c] = b8[t][c] = 0.0f;
    const uint4* __restrict__ row4 = reinterpret_cast<const uint4*>(a.nl16 + (size_t)slot * a.cap);
    walk_row<G>(s_j, row4, nn, sub, [&](const float4& pj) {
        const int tj = staged_type(pj);
        const float4 lj = s_lji[tj];
        const float d[3] = {xi.x - pj.x, xi.y - pj.y, xi.z - pj.z};
        const float rsq = fmaf(d[2], d[2], fmaf(d[1], d[1], d[0] * d[0]));
        const float ri = __builtin_amdgcn_rsqf(rsq);
        const float ri2 = ri * ri, ri3 = ri2 * ri;
        const float ri6 = ri3 * ri3, ri8 = ri6 * ri2;
        float ff = fmaf(lj.x, ri6, -lj.y) * ri8;
        float u = 0.0f;   // q_j C(r) / r^2
        if (COUL == 0) {
            u = (ri - a.reacf * rsq) * ri2 * pj.w;
        } else if (COUL == 1 || COUL == 2) {
            u = ctab_eval(s_ct, (rsq * ri) * a.ct_scale) * pj.w * ri3;
        }
        if (COUL <= 2) ff = fmaf(u, qi, ff);
        const bool use = !(rsq > a.rc2);
        ff = use ? ff : 0.0f;
        const float w14 = use ? ri6 * ri8 : 0.0f, w8 = use ? ri8 : 0.0f, wc = use ? u : 0.0f;
#pragma unroll
        for (int c = 0; c < 3; c++) {
            f[c] = fmaf(ff, d[c], f[c]);
            if (COUL <= 2) bc[c] = fmaf(wc, d[c], bc[c]);
        }
#pragma unroll
        for (int t = 0; t < NT; t++) {
            const float x14 = tj == t ? w14 : 0.0f, x8 = tj == t ? w8 : 0.0f;
#pragma unroll
            for (int c = 0; c < 3; c++) {
                b14[t][c] = fmaf(x14, d[c], b14[t][c]);
                b8[t][c] = fmaf(x8, d[c], b8[t][c]);
            }
        }
    });
    // the atom's force on all of its lanes, then the residual
#pragma unroll
    for (int c = 0; c < 3; c++) f[c] = lane_group_sum<G>(f[c]);
    if (active) {
        const float4 T = a.target[slot];
        const float R[3] = {f[0] - T.x, f[1] - T.y, f[2] - T.z};
        if (sub == 0) acc[0] += (double)R[0] * R[0] + (double)R[1] * R[1] + (double)R[2] * R[2];
        if (COUL <= 2)
            acc[1] += (double)fmaf(R[2], bc[2], fmaf(R[1], bc[1], R[0] * bc[0])) * 2.0 * (double)qs *
                      (double)a.qfac * (double)xi.w;
        double* a1 = acc + 2;
        double* a2 = acc + 2 + NT * NT;
#pragma unroll
        for (int t = 0; t < NT; t++) {
            const double g14 = (double)fmaf(R[2], b14[t][2], fmaf(R[1], b14[t][1], R[0] * b14[t][0]));
            const double g8 = (double)fmaf(R[2], b8[t][2], fmaf(R[1], b8[t][1], R[0] * b8[t][0]));
#pragma unroll
            for (int ta = 0; ta < NT; ta++) {
                if (ta == ti) {
                    a1[ta * NT + t] += g14;
                    a2[ta * NT + t] -= g8;
                }
            }
        }
    }
    fm_flush<NW, NACC>(acc, s_red, part, tile);
}

// Workgroup k: out[k] = sum over the per-tile slots of accumulator k (fixed
// order: deterministic).
__global__ __launch_bounds__(256) void k_fm_reduce(const double* part, int nslot, int nacc, double* out) {
    __shared__ double s_red[4];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int k = blockIdx.x;
    double v = 0.0;
    for (int b = threadIdx.x; b < nslot; b += blockDim.x) v += part[(size_t)b * nacc + k];
    for (int o = 32; o >= 1; o >>= 1) v += __shfl_xor(v, o, 64);
    if (lane == 0) s_red[wave] = v;
    __syncthreads();
    if (threadIdx.x == 0) out[k] = s_red[0] + s_red[1] + s_red[2] + s_red[3];
}

template <bool FUSED, int G, int NT>
static void launch_fm_g(const StepArgs& a, double* part, int coul, hipStream_t st) {
    const size_t sh = (size_t)(a.jcap + 1) * sizeof(float4);
    const dim3 grid(a.ntiles), blk(64 * G);
    switch (coul) {
    case 0: hipLaunchKernelGGL((FUSED ? k_fm_fused<G, 0, NT> : k_fm_adjoint<G, 0, NT>), grid, blk, sh, st, a, part); break;
    case 1: hipLaunchKernelGGL((FUSED ? k_fm_fused<G, 1, NT> : k_fm_adjoint<G, 1, NT>), grid, blk, sh, st, a, part); break;
    case 2: hipLaunchKernelGGL((FUSED ? k_fm_fused<G, 2, NT> : k_fm_adjoint<G, 2, NT>), grid, blk, sh, st, a, part); break;
    default: hipLaunchKernelGGL((FUSED ? k_fm_fused<G, 3, NT> : k_fm_adjoint<G, 3, NT>), grid, blk, sh, st, a, part); break;
    }
}

template <bool FUSED, int NT>
static hipError_t launch_fm_nt(const StepArgs& a, double* part, int lanes, int coul, hipStream_t st) {
    switch (lanes) {
    case 4: launch_fm_g<FUSED, 4, NT>(a, part, coul, st); break;
    case 8: launch_fm_g<FUSED, 8, NT>(a, part, coul, st); break;
    case 16: launch_fm_g<FUSED, 16, NT>(a, part, coul, st); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

// ev0 / ev1 (optional): events recorded around the per-tile kernel alone
template <bool FUSED>
static hipError_t launch_fm(const StepArgs& a, double* part, double* out, int lanes, int coul, int nt, hipStream_t st,
                            hipEvent_t ev0, hipEvent_t ev1) {
    hipError_t e = hipErrorInvalidValue;
    if (ev0 && (e = hipEventRecord(ev0, st)) != hipSuccess) return e;
    e = hipErrorInvalidValue;
    if (nt == 1) e = launch_fm_nt<FUSED, 1>(a, part, lanes, coul, st);
    if (nt == 2) e = launch_fm_nt<FUSED, 2>(a, part, lanes, coul, st);
    if (nt == 4) e = launch_fm_nt<FUSED, 4>(a, part, lanes, coul, st);
    if (e != hipSuccess) return e;
    if (ev1 && (e = hipEventRecord(ev1, st)) != hipSuccess) return e;
    hipLaunchKernelGGL(k_fm_reduce, dim3(2 + 2 * nt * nt), dim3(256), 0, st, (const double*)part, a.ntiles,
                       2 + 2 * nt * nt, out);
    return hipGetLastError();
}

// Single pass (no bonded terms): loss and gradient sums without a separate force pass.
// a.ljtab = the trial table, a.qq_scale = qscale.
// ntype must be 1, 2 or 4 (the host rounds up; unused type rows stay zero)
// part: [ntiles][2 + 2 nt^2] scratch; out: [2 + 2 nt^2] sums (loss, qscale, lj1, lj2 terms)
extern "C" hipError_t md_launch_fm_fused(const StepArgs& a, double* part, double* out, int lanes, int coul, int nt,
                                         hipStream_t st, hipEvent_t ev0, hipEvent_t ev1) {
    return launch_fm<true>(a, part, out, lanes, coul, nt, st, ev0, ev1);
}

extern "C" hipError_t md_launch_fm_adjoint(const StepArgs& a, double* part, double* out, int lanes, int coul, int nt,
                                           hipStream_t st, hipEvent_t ev0, hipEvent_t ev1) {
    return launch_fm<false>(a, part, out, lanes, coul, nt, st, ev0, ev1);
}
