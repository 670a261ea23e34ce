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
// Algorithmic work per ordered pair: forward 65 flop + adjoint ~45 flop.
#include "md_device.h"

namespace {
__device__ __forceinline__ float minimg(float d, float L, float Linv) {
    return fmaf(-L, rintf(d * Linv), d);
}
// Every list entry holds a slot index (the list is zeroed when allocated and
// builds write only slot indices), so the four gathers of a quad need no mask.
__device__ __forceinline__ void gather4(const float4* __restrict__ pin, const uint4 e4, float4* pj) {
    const uint32_t ev[4] = {e4.x, e4.y, e4.z, e4.w};
#pragma unroll
    for (int q = 0; q < 4; q++)
        pj[q] = *reinterpret_cast<const float4*>(reinterpret_cast<const char*>(pin) + ((ev[q] & MD_JMASK) << 4));
}
}  // namespace

// Row walk as in the step kernel: lane `sub` reads 16-byte quads of its atom's
// row (logical entries sub + G m in order), the next quad in flight; four
// gathers per quad.  The Coulomb factor C(r)/r^2 comes from the same cubic
// table of the force pass (Wolf / damped), so the gradient is that of the
// force the GPU computes.  Each workgroup stores its FP64 sums in its own slot
// of part (no atomics); k_fm_reduce adds the slots.
template <int G, int COUL, int NT>
__global__ __launch_bounds__(256) void k_fm_adjoint(const StepArgs a, double* part) {
    constexpr int APW = 64 / G, APB = 4 * APW;
    constexpr int NACC = 2 + 2 * NT * NT;
    __shared__ double s_red[4][NACC];
    __shared__ float4 s_ct[(COUL == 1 || COUL == 2) ? MD_CTAB + 1 : 1];
    static_assert(MD_CTAB == 256, "the table is loaded by the 256 threads of the workgroup");
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int sub = lane & (G - 1), grp = lane / G;
    if (COUL == 1 || COUL == 2) {
        s_ct[tid] = a.ctab[tid];
        if (tid == 0) s_ct[MD_CTAB] = a.ctab[MD_CTAB];
    }
    __syncthreads();
    const int lb = xcd_block(blockIdx.x, gridDim.x);   // logical workgroup: contiguous per XCD
    const int u0 = (int)(((long long)lb * a.nunits) / gridDim.x);
    const int u1 = (int)(((long long)(lb + 1) * a.nunits) / gridDim.x);
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
        const uint4* __restrict__ row4 = reinterpret_cast<const uint4*>(a.nlist + (size_t)slot * a.cap);
        const int nb = (nn + 4 * G - 1) / (4 * G);
        uint4 nxt = nb > 0 ? row4[sub] : make_uint4(0, 0, 0, 0);
        for (int b = 0; b < nb; b++) {
            const uint4 e4 = nxt;
            if (b + 1 < nb) nxt = row4[(b + 1) * G + sub];
            const uint32_t ev[4] = {e4.x, e4.y, e4.z, e4.w};
            const int base = 4 * G * b + sub;
            float4 pj[4];
#pragma unroll
            for (int q = 0; q < 4; q++) {
                const uint32_t off = ((base + G * q < nn) ? (ev[q] & MD_JMASK) : (uint32_t)slot) << 4;
                pj[q] = *reinterpret_cast<const float4*>(reinterpret_cast<const char*>(pin) + off);
            }
#pragma unroll
            for (int q = 0; q < 4; q++) {
                const bool valid = base + G * q < nn;
                const int tj = (int)(ev[q] >> MD_JBITS);
                const float dx = minimg(xi.x - pj[q].x, L0, Li0);
                const float dy = minimg(xi.y - pj[q].y, L1, Li1);
                const float dz = minimg(xi.z - pj[q].z, L2, Li2);
                // a masked entry is the atom itself: give it a distance beyond the cutoff
                const float rsq = valid ? fmaf(dz, dz, fmaf(dy, dy, dx * dx)) : 2.0f * a.rc2;
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
                    sq = fmaf(g * pj[q].w, (ri - a.reacf * rsq) * ri2, sq);
                } else if (COUL == 1 || COUL == 2) {
                    const float hv = ctab_eval(s_ct, (rsq * ri) * a.ct_scale);
                    sq = fmaf(g * pj[q].w, hv * (ri2 * ri), sq);   // C(r) / r^2, as in the force pass
                }
            }
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
    if (tid < NACC) part[(size_t)blockIdx.x * NACC + tid] = s_red[0][tid] + s_red[1][tid] + s_red[2][tid] + s_red[3][tid];
}

// Single-pass form (systems without bonded terms): the rows are walked once for
// the force with the trial parameters and, on the same lanes, for the basis
// sums whose dot products with the residual R_i = F_i - F*_i give the gradient:
//   b14[t] = sum_{j of type t} r^-14 d_ij,   b8[t] = sum_{j of type t} r^-8 d_ij,
//   bc = sum_j q_j C(r)/r^2 d_ij,
//   dloss/dlj1[ti,t] = (2/M) sum_i R_i . b14[t],  dloss/dlj2[ti,t] = -(2/M) sum_i R_i . b8[t],
//   dloss/dqscale = (2/M) sum_i 2 qscale qfac q_i R_i . bc.
// R_i . b is linear in b, so each lane dots the residual (known to all lanes of
// the atom after the force reduction) with its own partial sums.  Same row walk
// as the step kernel (quads, the next block's gathers in flight, unmasked full
// blocks); a.qq_scale holds qscale.
template <int G, int COUL, int NT, bool TAIL>
__device__ __forceinline__ void fm_quad(const StepArgs& a, const float4* s_lj, const float4* s_ct, const uint4 e4,
                                        const float4* pj, int base, int nn, const float4 xi, float qi, float L0,
                                        float L1, float L2, float Li0, float Li1, float Li2, float* f, float (*b14)[3],
                                        float (*b8)[3], float* bc) {
    const uint32_t ev[4] = {e4.x, e4.y, e4.z, e4.w};
#pragma unroll
    for (int q = 0; q < 4; q++) {
        const int tj = (int)(ev[q] >> MD_JBITS);
        const float4 lj = *reinterpret_cast<const float4*>(reinterpret_cast<const char*>(s_lj) + (ev[q] >> (MD_JBITS - 4)));
        const float d[3] = {minimg(xi.x - pj[q].x, L0, Li0), minimg(xi.y - pj[q].y, L1, Li1),
                            minimg(xi.z - pj[q].z, L2, Li2)};
        const float rsq = fmaf(d[2], d[2], fmaf(d[1], d[1], d[0] * d[0]));
        const float ri = __builtin_amdgcn_rsqf(rsq);
        const float ri2 = ri * ri, ri3 = ri2 * ri;
        const float ri6 = ri3 * ri3, ri8 = ri6 * ri2;
        float ff = fmaf(lj.x, ri6, -lj.y) * ri8;
        float u = 0.0f;   // q_j C(r) / r^2
        if (COUL == 0) {
            u = (ri - a.reacf * rsq) * ri2 * pj[q].w;
        } else if (COUL == 1 || COUL == 2) {
            u = ctab_eval(s_ct, (rsq * ri) * a.ct_scale) * pj[q].w * ri3;
        }
        if (COUL <= 2) ff = fmaf(u, qi, ff);
        const bool use = (!TAIL || base + G * q < nn) && !(rsq > a.rc2);
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
    }
}

template <int G, int COUL, int NT>
__global__ __launch_bounds__(256) void k_fm_fused(const StepArgs a, double* part) {
    constexpr int APW = 64 / G, APB = 4 * APW;
    constexpr int NACC = 2 + 2 * NT * NT;
    __shared__ double s_red[4][NACC];
    __shared__ float4 s_lj[64];
    __shared__ float4 s_ct[(COUL == 1 || COUL == 2) ? MD_CTAB + 1 : 1];
    static_assert(MD_CTAB == 256, "the table is loaded by the 256 threads of the workgroup");
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int sub = lane & (G - 1), grp = lane / G;
    if (tid < 64) s_lj[tid] = a.ljtab[tid];
    if (COUL == 1 || COUL == 2) {
        s_ct[tid] = a.ctab[tid];
        if (tid == 0) s_ct[MD_CTAB] = a.ctab[MD_CTAB];
    }
    __syncthreads();
    const int lb = xcd_block(blockIdx.x, gridDim.x);   // logical workgroup: contiguous per XCD
    const int u0 = (int)(((long long)lb * a.nunits) / gridDim.x);
    const int u1 = (int)(((long long)(lb + 1) * a.nunits) / gridDim.x);
    const float qs = a.qq_scale;
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
        const uint4* __restrict__ row4 = reinterpret_cast<const uint4*>(a.nlist + (size_t)slot * a.cap);
        const int nb = (int)(((uint32_t)nn + 4 * G - 1) / (4 * G));   // blocks; all but the last are full
        if (nb > 0) {
            uint4 e0 = row4[sub];
            float4 p0[4];
            gather4(pin, e0, p0);
            uint4 e1 = row4[min(1, nb - 1) * G + sub];
            for (int b = 0; b < nb - 1; b++) {
                float4 p1[4];
                gather4(pin, e1, p1);
                const uint4 e2 = row4[min(b + 2, nb - 1) * G + sub];
                fm_quad<G, COUL, NT, false>(a, s_lji, s_ct, e0, p0, 0, nn, xi, qi, L0, L1, L2, Li0, Li1, Li2, f, b14,
                                            b8, bc);
                e0 = e1;
                e1 = e2;
#pragma unroll
                for (int q = 0; q < 4; q++) p0[q] = p1[q];
            }
            fm_quad<G, COUL, NT, true>(a, s_lji, s_ct, e0, p0, 4 * G * (nb - 1) + sub, nn, xi, qi, L0, L1, L2, Li0,
                                       Li1, Li2, f, b14, b8, bc);
        }
        // the atom's force on all of its lanes, then the residual
#pragma unroll
        for (int o = G / 2; o >= 1; o >>= 1)
#pragma unroll
            for (int c = 0; c < 3; c++) f[c] += __shfl_xor(f[c], o, 64);
        if (!active) continue;
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
#pragma unroll
    for (int k = 0; k < NACC; k++)
        for (int o = 32; o >= 1; o >>= 1) acc[k] += __shfl_xor(acc[k], o, 64);
    if (lane == 0)
        for (int k = 0; k < NACC; k++) s_red[wave][k] = acc[k];
    __syncthreads();
    if (tid < NACC) part[(size_t)blockIdx.x * NACC + tid] = s_red[0][tid] + s_red[1][tid] + s_red[2][tid] + s_red[3][tid];
}

// One workgroup: out[k] = sum over the adjoint kernel's workgroup slots.
__global__ __launch_bounds__(256) void k_fm_reduce(const double* part, int nslot, int nacc, double* out) {
    __shared__ double s_red[4];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    for (int k = 0; k < nacc; k++) {
        double v = 0.0;
        for (int b = threadIdx.x; b < nslot; b += blockDim.x) v += part[(size_t)b * nacc + k];
        for (int o = 32; o >= 1; o >>= 1) v += __shfl_xor(v, o, 64);
        if (lane == 0) s_red[wave] = v;
        __syncthreads();
        if (threadIdx.x == 0) out[k] = s_red[0] + s_red[1] + s_red[2] + s_red[3];
        __syncthreads();
    }
}

template <int G, int NT>
static void launch_fm_g(const StepArgs& a, double* gacc, int coul, int grid, hipStream_t st) {
    // gacc: [grid][2 + 2 NT^2] workgroup slots
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

template <int G, int NT>
static void launch_fused_g(const StepArgs& a, double* gacc, int coul, int grid, hipStream_t st) {
    switch (coul) {
    case 0: hipLaunchKernelGGL((k_fm_fused<G, 0, NT>), dim3(grid), dim3(256), 0, st, a, gacc); break;
    case 1: hipLaunchKernelGGL((k_fm_fused<G, 1, NT>), dim3(grid), dim3(256), 0, st, a, gacc); break;
    case 2: hipLaunchKernelGGL((k_fm_fused<G, 2, NT>), dim3(grid), dim3(256), 0, st, a, gacc); break;
    default: hipLaunchKernelGGL((k_fm_fused<G, 3, NT>), dim3(grid), dim3(256), 0, st, a, gacc); break;
    }
}

template <int NT>
static hipError_t launch_fused_nt(const StepArgs& a, double* gacc, int lanes, int coul, int grid, hipStream_t st) {
    switch (lanes) {
    case 4: launch_fused_g<4, NT>(a, gacc, coul, grid, st); break;
    case 8: launch_fused_g<8, NT>(a, gacc, coul, grid, st); break;
    case 16: launch_fused_g<16, NT>(a, gacc, coul, grid, st); break;
    case 32: launch_fused_g<32, NT>(a, gacc, coul, grid, st); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

// Single pass (no bonded terms): loss and gradient sums without a separate force pass.
// a.ljtab = the trial table, a.qq_scale = qscale; part/out as md_launch_fm_adjoint.
extern "C" hipError_t md_launch_fm_fused(const StepArgs& a, double* part, double* out, int lanes, int coul, int nt,
                                         int grid, hipStream_t st) {
    hipError_t e = hipErrorInvalidValue;
    if (nt == 1) e = launch_fused_nt<1>(a, part, lanes, coul, grid, st);
    if (nt == 2) e = launch_fused_nt<2>(a, part, lanes, coul, grid, st);
    if (nt == 4) e = launch_fused_nt<4>(a, part, lanes, coul, grid, st);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_fm_reduce, dim3(1), dim3(256), 0, st, (const double*)part, grid, 2 + 2 * nt * nt, out);
    return hipGetLastError();
}

// ntype must be 1, 2 or 4 (the host rounds up; unused type rows stay zero)
// part: [grid][2 + 2 nt^2] scratch; out: [2 + 2 nt^2] sums (loss, qscale, lj1, lj2 terms)
extern "C" hipError_t md_launch_fm_adjoint(const StepArgs& a, double* part, double* out, int lanes, int coul, int nt,
                                           int grid, hipStream_t st) {
    hipError_t e = hipErrorInvalidValue;
    if (nt == 1) e = launch_fm_nt<1>(a, part, lanes, coul, grid, st);
    if (nt == 2) e = launch_fm_nt<2>(a, part, lanes, coul, grid, st);
    if (nt == 4) e = launch_fm_nt<4>(a, part, lanes, coul, grid, st);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_fm_reduce, dim3(1), dim3(256), 0, st, (const double*)part, grid, 2 + 2 * nt * nt, out);
    return hipGetLastError();
}
