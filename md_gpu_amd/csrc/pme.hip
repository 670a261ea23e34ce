// pme.hip -- smooth particle-mesh Ewald (Essmann, Perera, Berkowitz, Darden, Lee,
// Pedersen, J. Chem. Phys. 103, 8577 (1995)): the reciprocal-space part that the
// damped Coulomb term (lkspace 2, src/forces.cu:97-110) leaves out, for one
// replica.  The reference vendors GROMACS' PME (src/pme.c:829-1340) without
// building it; this is the method from the paper, as the oracle's md_pme.cpp.
//
// Per step, before the step kernel (which adds frec to each atom's pair force and
// the long-range energy and virial to the replica's first tile):
//   k_pme_spread  the order-p B-spline weights of each atom's scaled coordinates
//                 u = K (x / L + 1/2) (in FP64, so the weights carry only FP32
//                 rounding), p^3 64-bit fixed-point atomic adds of q th_x th_y th_z
//                 into the charge grid by a wavefront per atom: integer sums, so the
//                 grid does not depend on the order of the adds and runs repeat bit
//                 for bit
//   k_pme_fix     fixed point -> float, clearing the fixed-point grid for the next step
//   R2C FFT       (hipFFT, on the engine's stream, its work area allocated up front)
//   k_pme_conv    multiply by the influence function C(m) = qfac exp(-pi^2 m^2/b^2)
//                 B(m) / (pi V m^2), summing the energy 1/2 C |S|^2 and the traced
//                 virial 1/2 C |S|^2 (1 - 2 pi^2 m^2 / b^2) over the half spectrum
//   C2R FFT       -> dE/dQ on the grid
//   k_pme_gather  four lanes per slot: F = -q grad(sum_k dE/dQ(k) th th th), plus the
//                 correction -qq erf(b r)/r of its excluded pairs
//   k_pme_sum     one workgroup: every workgroup's energy and virial in a fixed
//                 order, plus the self term and the neutralizing-background term
//                 at the step's volume.
// Every kernel first reads the replica's state and does nothing for a step the
// step kernel will not take (the same test as the step kernel), so the chain can
// sit inside a device-gated graph segment, and reads the replica's box (pressure
// coupling rescales it after each step).
#include <hipfft/hipfft.h>

#include "md_device.h"
#include "md_reduce.h"

namespace {

constexpr int PT = 256;   // threads per workgroup

__device__ __forceinline__ bool pme_go(const RepState& s, int mode) {
    if (mode == MODE_FORCE) return !stale(s);
    if (mode == MODE_LF) return !stale(s) && s.step < s.target;
    return s.phase == 1 && !stale(s);
}

// th[j] = M_P(w + j), dth[j] = M_P'(w + j), j = 0..P-1, 0 <= w <= 1:
// M_2(x) = 1 - |x - 1|, M_{k+1}(x) = [x M_k(x) + (k + 1 - x) M_k(x - 1)] / k,
// M_k'(x) = M_{k-1}(x) - M_{k-1}(x - 1).
template <int P, bool DER>
__device__ __forceinline__ void bspline(float w, float* th, float* dth) {
    float a[P];
    a[0] = w;
    a[1] = 1.0f - w;
#pragma unroll
    for (int j = 2; j < P; j++) a[j] = 0.0f;
#pragma unroll
    for (int k = 2; k < P; k++) {
        if (DER && k == P - 1) {
#pragma unroll
            for (int j = 0; j < P; j++) dth[j] = (j < k ? a[j] : 0.0f) - (j > 0 ? a[j - 1] : 0.0f);
        }
        const float rk = 1.0f / (float)k;
        float b[P];
#pragma unroll
        for (int j = 0; j <= k; j++)
            b[j] = ((w + (float)j) * (j < k ? a[j] : 0.0f) + ((float)(k + 1 - j) - w) * (j > 0 ? a[j - 1] : 0.0f)) * rk;
#pragma unroll
        for (int j = 0; j <= k; j++) a[j] = b[j];
    }
#pragma unroll
    for (int j = 0; j < P; j++) th[j] = a[j];
}

// Grid index k0 = floor(u) mod K and the fraction w = u - floor(u) of
// u = (x + L/2) K / L, in FP64: the modulo as fl - K floor(fl / K) with one
// correction by K where the rounded quotient lands on the wrong integer (no
// emulated 64-bit integer division).
__device__ __forceinline__ void grid_coord(float x, double hl, double kl, double ik, int K, int& k0, float& w) {
    double u = ((double)x + hl) * kl;
    if (!(fabs(u) < 1e9)) u = 0.0;   // a non-finite coordinate (the step kernel reports it) stays in the grid
    const double fl = floor(u);
    w = (float)(u - fl);
    double m = fl - (double)K * floor(fl * ik);
    m = m >= (double)K ? m - (double)K : (m < 0.0 ? m + (double)K : m);
    k0 = (int)m;
}

// index of stencil point j (k0 - j, periodic)
__device__ __forceinline__ int wrapk(int k, int K) { return k < 0 ? k + K : k; }

// The replica's box as the PME kernels use it, read at each launch: pressure
// coupling rescales it after every step.
struct PmeBox {
    float L[3], Li[3];
    double hl[3], kl[3];   // L / 2 and K / L
};
__device__ __forceinline__ PmeBox pme_box(const PmeArgs& a) {
    const RepParam& p = a.rep[0];
    PmeBox b;
    const int K[3] = {a.K0, a.K1, a.K2};
#pragma unroll
    for (int d = 0; d < 3; d++) {
        b.L[d] = p.L[d];
        b.Li[d] = p.Linv[d];
        b.hl[d] = 0.5 * (double)p.L[d];
        b.kl[d] = (double)K[d] / (double)p.L[d];
    }
    return b;
}

// Lane l of a wavefront first computes the stencil of slot (wavefront's base + l);
// then the wavefront walks its 64 atoms, its lanes taking the atom's P^3 grid
// points with the z index fastest, so each atomic instruction touches P^2 runs of
// P consecutive grid words instead of 64 scattered ones (atomics are combined
// per cache line on the way to L2).
template <int P>
__global__ __launch_bounds__(PT) void k_pme_spread(const PmeArgs a) {
    const RepState s = a.st[0];
    if (!pme_go(s, a.mode)) return;   // grid-uniform
    const int lane = threadIdx.x & 63;
    const int slot = blockIdx.x * PT + threadIdx.x;
    float q = 0.0f;
    int k0 = 0, k1 = 0, k2 = 0;
    float t0[P], t1[P], t2[P];
#pragma unroll
    for (int c = 0; c < P; c++) t0[c] = t1[c] = t2[c] = 0.0f;
    if (slot < a.nslot && a.ids[slot] >= 0) {
        const PmeBox bx = pme_box(a);
        const float4 x = a.pos[cur_buf(s)][slot];
        q = x.w;
        float w0, w1, w2;
        grid_coord(x.x, bx.hl[0], bx.kl[0], a.ik0, a.K0, k0, w0);
        grid_coord(x.y, bx.hl[1], bx.kl[1], a.ik1, a.K1, k1, w1);
        grid_coord(x.z, bx.hl[2], bx.kl[2], a.ik2, a.K2, k2, w2);
        bspline<P, false>(w0, t0, nullptr);
        bspline<P, false>(w1, t1, nullptr);
        bspline<P, false>(w2, t2, nullptr);
#pragma unroll
        for (int c = 0; c < P; c++) t0[c] *= q;   // the charge rides on the x weights
    }
    unsigned long long todo = __ballot(q != 0.0f);
    while (todo) {
        const int t = __builtin_ctzll(todo);   // wavefront-uniform
        todo &= todo - 1;
        const int b0 = __builtin_amdgcn_readlane(k0, t), b1 = __builtin_amdgcn_readlane(k1, t),
                  b2 = __builtin_amdgcn_readlane(k2, t);
        float u0[P], u1[P], u2[P];
#pragma unroll
        for (int c = 0; c < P; c++) {
            u0[c] = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(t0[c]), t));
            u1[c] = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(t1[c]), t));
            u2[c] = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(t2[c]), t));
        }
#pragma unroll
        for (int p0 = 0; p0 < P * P * P; p0 += 64) {
            const int p = p0 + lane;
            if (P * P * P % 64 != 0 && p >= P * P * P) break;
            const int i = p / (P * P), j = (p / P) % P, c = p % P;
            float wx = u0[0], wy = u1[0], wz = u2[0];
#pragma unroll
            for (int m = 1; m < P; m++) {
                wx = i == m ? u0[m] : wx;
                wy = j == m ? u1[m] : wy;
                wz = c == m ? u2[m] : wz;
            }
            const size_t idx = ((size_t)wrapk(b0 - i, a.K0) * a.K1 + wrapk(b1 - j, a.K1)) * a.K2 + wrapk(b2 - c, a.K2);
            const long long v = __double2ll_rn((double)(wx * wy * wz) * MD_PME_QUNIT);
            atomicAdd(a.qacc + idx, (unsigned long long)v);
        }
    }
}

__global__ __launch_bounds__(PT) void k_pme_fix(const PmeArgs a) {
    if (!pme_go(a.st[0], a.mode)) return;
    const size_t n = (size_t)a.K0 * a.K1 * a.K2;
    for (size_t i = (size_t)blockIdx.x * PT + threadIdx.x; i < n; i += (size_t)gridDim.x * PT) {
        const long long v = (long long)a.qacc[i];
        a.qacc[i] = 0ull;
        a.qgrid[i] = (float)((double)v * (1.0 / MD_PME_QUNIT));
    }
}

// FP64 sum of (e, w) over the workgroup (fixed order) into part[k]
__device__ __forceinline__ void block_pair_sum(double e, double w, double* part, int k) {
    __shared__ double s_red[PT / 64][2];
    e = wave_sum_d(e);
    w = wave_sum_d(w);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (lane == 0) {
        s_red[wave][0] = e;
        s_red[wave][1] = w;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double te = 0.0, tw = 0.0;
        for (int v = 0; v < PT / 64; v++) {
            te += s_red[v][0];
            tw += s_red[v][1];
        }
        part[2 * (size_t)k] = te;
        part[2 * (size_t)k + 1] = tw;
    }
}

__global__ __launch_bounds__(PT) void k_pme_conv(const PmeArgs a) {
    if (!pme_go(a.st[0], a.mode)) return;   // grid-uniform
    const int H = a.K2 / 2 + 1;
    const size_t n = (size_t)a.K0 * a.K1 * H;
    const float2* t1 = a.tab + a.K0;
    const float2* t2 = a.tab + a.K0 + a.K1;
    const PmeBox bx = pme_box(a);
    const float cpref = (float)((double)a.qfac / (3.14159265358979323846 * (double)bx.L[0] * bx.L[1] * bx.L[2]));
    double e = 0.0, w = 0.0;
    for (size_t i = (size_t)blockIdx.x * PT + threadIdx.x; i < n; i += (size_t)gridDim.x * PT) {
        const int i2 = (int)(i % H);
        const size_t r = i / H;
        const int i1 = (int)(r % a.K1), i0 = (int)(r / a.K1);
        const float2 u0 = a.tab[i0], u1 = t1[i1], u2 = t2[i2];
        const float m0 = u0.x * bx.Li[0], m1 = u1.x * bx.Li[1], m2 = u2.x * bx.Li[2];   // reciprocal vector
        const float mm = fmaf(m2, m2, fmaf(m1, m1, m0 * m0));
        const float2 S = a.sgrid[i];
        float C = 0.0f;
        if (i != 0) C = cpref * expf(-a.pi2_b2 * mm) / mm * (u0.y * u1.y * u2.y);
        // the half spectrum: k2 = 0 and (K2 even) K2/2 stand for themselves, the
        // others also for their conjugate partners
        const double half = (i2 == 0 || 2 * i2 == a.K2) ? 0.5 : 1.0;
        const double em = half * (double)C * ((double)S.x * S.x + (double)S.y * S.y);
        e += em;
        w += em * (1.0 - 2.0 * (double)a.pi2_b2 * (double)mm);
        a.sgrid[i] = make_float2(S.x * C, S.y * C);
    }
    block_pair_sum(e, w, a.part, blockIdx.x);
}

// Four lanes (a quad) per atom, 64 atoms per workgroup, so the load latency is
// spread over four times the wavefronts: lanes 0..2 of the quad compute the
// B-spline weights and derivatives of one axis each (into LDS, once per atom);
// then lane c of the quad reads the stencil points with z offsets c and c + 4
// and the quad's partial gradients are summed by DPP.  Lane c also takes the
// atom's c-th excluded partner.
template <int P>
__global__ __launch_bounds__(PT) void k_pme_gather(const PmeArgs a) {
    __shared__ float s_t[PT / 4][3][P], s_d[PT / 4][3][P];
    __shared__ int s_k[PT / 4][3];
    const RepState s = a.st[0];
    if (!pme_go(s, a.mode)) return;   // grid-uniform
    const int sub = threadIdx.x & 3, la = threadIdx.x >> 2;   // lane in the quad, atom in the workgroup
    const int slot = blockIdx.x * (PT / 4) + la;
    double e = 0.0, w = 0.0;
    const int id = slot < a.nslot ? a.ids[slot] : -1;
    const float4* __restrict__ pin = a.pos[cur_buf(s)];
    const float4 x = id >= 0 ? pin[slot] : make_float4(0.f, 0.f, 0.f, 0.f);
    const PmeBox bx = pme_box(a);
    if (x.w != 0.0f && sub < 3) {   // this lane's axis
        const float xc = sub == 0 ? x.x : (sub == 1 ? x.y : x.z);
        // selected, not indexed by the lane's axis: an indexed private array lives in scratch
        const double hl = sub == 0 ? bx.hl[0] : (sub == 1 ? bx.hl[1] : bx.hl[2]);
        const double kl = sub == 0 ? bx.kl[0] : (sub == 1 ? bx.kl[1] : bx.kl[2]);
        const double ik = sub == 0 ? a.ik0 : (sub == 1 ? a.ik1 : a.ik2);
        const int K = sub == 0 ? a.K0 : (sub == 1 ? a.K1 : a.K2);
        int k0;
        float w0, t[P], d[P];
        grid_coord(xc, hl, kl, ik, K, k0, w0);
        bspline<P, true>(w0, t, d);
        s_k[la][sub] = k0;
#pragma unroll
        for (int m = 0; m < P; m++) {
            s_t[la][sub][m] = t[m];
            s_d[la][sub][m] = d[m];
        }
    }
    __syncthreads();
    float fx = 0.0f, fy = 0.0f, fz = 0.0f;
    if (x.w != 0.0f) {
        const int k0 = s_k[la][0], k1 = s_k[la][1], k2 = s_k[la][2];
        float t0[P], t1[P], d0[P], d1[P];
#pragma unroll
        for (int m = 0; m < P; m++) {
            t0[m] = s_t[la][0][m];
            d0[m] = s_d[la][0][m];
            t1[m] = s_t[la][1][m];
            d1[m] = s_d[la][1][m];
        }
        // this lane's z points c = sub and sub + 4 (weight 0 where c >= P)
        constexpr int NH = P > 4 ? 2 : 1;
        float tz[NH], dz[NH];
        int gz[NH];
#pragma unroll
        for (int h = 0; h < NH; h++) {
            const int c = sub + 4 * h;
            const bool in = c < P;
            tz[h] = in ? s_t[la][2][in ? c : 0] : 0.0f;
            dz[h] = in ? s_d[la][2][in ? c : 0] : 0.0f;
            gz[h] = in ? wrapk(k2 - c, a.K2) : 0;
        }
        float gx = 0.0f, gy = 0.0f, gzz = 0.0f;
#pragma unroll
        for (int i = 0; i < P; i++) {
            const int g0 = wrapk(k0 - i, a.K0);
#pragma unroll
            for (int j = 0; j < P; j++) {
                const float* row = a.qgrid + ((size_t)g0 * a.K1 + wrapk(k1 - j, a.K1)) * a.K2;
                float sz = 0.0f, sdz = 0.0f;
#pragma unroll
                for (int h = 0; h < NH; h++) {
                    const float ph = row[gz[h]];
                    sz = fmaf(ph, tz[h], sz);
                    sdz = fmaf(ph, dz[h], sdz);
                }
                gx = fmaf(d0[i] * t1[j], sz, gx);
                gy = fmaf(t0[i] * d1[j], sz, gy);
                gzz = fmaf(t0[i] * t1[j], sdz, gzz);
            }
        }
        fx = -x.w * gx * (float)bx.kl[0];
        fy = -x.w * gy * (float)bx.kl[1];
        fz = -x.w * gzz * (float)bx.kl[2];
        if (a.has_excl) {
            // excluded pairs: U = -qq erf(b r)/r, F_i = qq [c2 e^{-b^2 r^2}/r - erf(b r)/r^2] d/r
            const int4 ex = a.excl[id];
            const int pj = sub == 0 ? ex.x : (sub == 1 ? ex.y : (sub == 2 ? ex.z : ex.w));
            const float4 y = pj >= 0 ? pin[a.slot_of[pj]] : make_float4(0.f, 0.f, 0.f, 0.f);
            if (y.w != 0.0f) {
                const float c2 = a.beta * 1.1283791670955126f;   // 2 beta / sqrt(pi)
                const float dx = minimg(x.x - y.x, bx.L[0], bx.Li[0]), dy = minimg(x.y - y.y, bx.L[1], bx.Li[1]),
                            dzz = minimg(x.z - y.z, bx.L[2], bx.Li[2]);
                const float r2 = fmaf(dzz, dzz, fmaf(dy, dy, dx * dx));
                const float r = sqrtf(r2), ir = 1.0f / r;
                const float br = a.beta * r;
                const float erfv = erff(br);
                const float qq = a.qfac * x.w * y.w;
                const float fr = qq * (c2 * expf(-br * br) - erfv * ir) * ir * ir;
                fx = fmaf(fr, dx, fx);
                fy = fmaf(fr, dy, fy);
                fz = fmaf(fr, dzz, fz);
                // energy and virial in FP64: for neutral molecules these terms nearly
                // cancel the self term, each far larger than their sum
                const double rd = sqrt((double)r2), bd = (double)a.beta * rd;
                const double qqd = (double)a.qfac * (double)x.w * (double)y.w, erfd = erf(bd);
                e -= 0.5 * qqd * erfd / rd;   // each pair is listed from both atoms
                w += 0.5 * qqd * (1.1283791670955126 * bd * exp(-bd * bd) - erfd) / rd;
            }
        }
    }
    fx = group_sum<4>(fx);
    fy = group_sum<4>(fy);
    fz = group_sum<4>(fz);
    if (id >= 0 && sub == 0) a.frec[slot] = make_float4(fx, fy, fz, 0.0f);
    block_pair_sum(e, w, a.part, a.nconv + blockIdx.x);
}

// Every workgroup's (energy, virial) pair of the influence and gather kernels in
// a fixed order, plus the self and background constants.  One workgroup after
// the gather: a device-scope fence and a shared counter in each of the gather's
// thousands of workgroups cost more than this launch.
__global__ __launch_bounds__(PT) void k_pme_sum(const PmeArgs a) {
    if (!pme_go(a.st[0], a.mode)) return;
    const int np = a.nconv + a.ngather;
    double te = 0.0, tw = 0.0;
    for (int k = threadIdx.x; k < np; k += PT) {
        te += a.part[2 * (size_t)k];
        tw += a.part[2 * (size_t)k + 1];
    }
    block_pair_sum(te, tw, a.out, 0);
    if (threadIdx.x == 0) {
        // self term, and the neutralizing background at the current volume (its
        // virial -L dE/dL is three times itself)
        const RepParam& p = a.rep[0];
        const double V = (double)p.L[0] * p.L[1] * p.L[2], b = (double)a.beta;
        const double ebg = -(double)a.qfac * 3.14159265358979323846 * a.qsum2 / (2.0 * V * b * b);
        a.out[0] += a.e_self + ebg;
        a.out[1] += 3.0 * ebg;
    }
}

template <int P>
void launch_spread(const PmeArgs& a, hipStream_t st) {
    hipLaunchKernelGGL(k_pme_spread<P>, dim3((a.nslot + PT - 1) / PT), dim3(PT), 0, st, a);
}

template <int P>
void launch_gather(const PmeArgs& a, hipStream_t st) {
    hipLaunchKernelGGL(k_pme_gather<P>, dim3(a.ngather), dim3(PT), 0, st, a);
}

}  // namespace

// Queue the chain for one step.  Returns 0, or 1 (HIP launch error) or 2 (hipFFT error).
extern "C" int md_launch_pme(const PmeArgs& a, hipfftHandle r2c, hipfftHandle c2r, hipStream_t st) {
    switch (a.order) {
    case 4: launch_spread<4>(a, st); break;
    case 6: launch_spread<6>(a, st); break;
    case 8: launch_spread<8>(a, st); break;
    default: return 1;
    }
    const size_t n = (size_t)a.K0 * a.K1 * a.K2;
    const int nfix = (int)std::min<size_t>((n + PT - 1) / PT, 4096);
    hipLaunchKernelGGL(k_pme_fix, dim3(nfix), dim3(PT), 0, st, a);
    if (hipGetLastError() != hipSuccess) return 1;
    if (hipfftExecR2C(r2c, a.qgrid, reinterpret_cast<hipfftComplex*>(a.sgrid)) != HIPFFT_SUCCESS) return 2;
    hipLaunchKernelGGL(k_pme_conv, dim3(a.nconv), dim3(PT), 0, st, a);
    if (hipfftExecC2R(c2r, reinterpret_cast<hipfftComplex*>(a.sgrid), a.qgrid) != HIPFFT_SUCCESS) return 2;
    switch (a.order) {
    case 4: launch_gather<4>(a, st); break;
    case 6: launch_gather<6>(a, st); break;
    default: launch_gather<8>(a, st); break;
    }
    hipLaunchKernelGGL(k_pme_sum, dim3(1), dim3(PT), 0, st, a);
    return hipGetLastError() == hipSuccess ? 0 : 1;
}
