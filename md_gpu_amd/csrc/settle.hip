// settle.hip -- rigid 3-site molecules (water): SETTLE after each leapfrog step.
//
// SETTLE (Miyamoto and Kollman, J. Comput. Chem. 13, 952 (1992)), restated from
// the paper; the reference's constraint code is uncompiled pseudo-code
// (src/settle.c, src/settle.cu:1-178, src/xshake.cu:15-82).  The step kernel has
// written the unconstrained positions x1 = x0 + v dt and velocities; this kernel,
// launched between the step kernel and its finalize, moves each molecule onto its
// rigid geometry, sets the velocities to (x - x0) / dt (leapfrog), and rewrites the
// tile's kinetic-energy and skin-trip slots from the constrained state (the
// potential energy and virial slots of the step kernel stand).
//
// One workgroup per 64-slot tile, a thread per slot.  The thread of a molecule's O
// atom handles the molecule (its H atoms may sit in other tiles: their kinetic
// energy goes to the O's tile slot, and the replica's sum over tiles is the
// same); an H thread does nothing; an unconstrained atom's thread takes its
// kinetic energy and skin check from the step kernel's update.  The O -> H vectors
// are taken by minimum image and the corrections added to each atom's own
// coordinates, so no atom changes its periodic image (the neighbor sets record
// images).
#include "md_device.h"
#include "md_reduce.h"

namespace {

__device__ __forceinline__ float3 sub3(float4 a, float4 b) { return make_float3(a.x - b.x, a.y - b.y, a.z - b.z); }
__device__ __forceinline__ float dotf(float3 a, float3 b) { return fmaf(a.z, b.z, fmaf(a.y, b.y, a.x * b.x)); }
__device__ __forceinline__ float3 crossf(float3 a, float3 b) {
    return make_float3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
__device__ __forceinline__ float3 unitf(float3 a) {
    const float r = __builtin_amdgcn_rsqf(dotf(a, a));
    return make_float3(a.x * r, a.y * r, a.z * r);
}

// One molecule: B0, C0 old O -> H vectors, B1, C1 new ones (new O at the origin);
// mA, mB the O and H masses (any common unit).  Out: the corrections of O, H1, H2.
// The same steps as the oracle's settle1: the frame of the old plane's normal, the
// tilts phi and psi that keep each atom's out-of-plane coordinate, the turn theta
// that makes the in-plane corrections combinations of the old bond vectors.
__device__ void settle_one(float mA, float mB, float rc, float h, float3 B0, float3 C0, float3 B1, float3 C1,
                           float3& dA, float3& dB, float3& dC) {
    const float M = mA + 2.0f * mB;
    const float ra = 2.0f * mB * h / M, rb = h - ra;
    const float3 com = make_float3(mB * (B1.x + C1.x) / M, mB * (B1.y + C1.y) / M, mB * (B1.z + C1.z) / M);
    const float3 a1 = make_float3(-com.x, -com.y, -com.z);
    const float3 b1 = make_float3(B1.x - com.x, B1.y - com.y, B1.z - com.z);
    const float3 c1 = make_float3(C1.x - com.x, C1.y - com.y, C1.z - com.z);
    const float3 ez = unitf(crossf(B0, C0));
    const float3 ex = unitf(crossf(a1, ez));
    const float3 ey = crossf(ez, ex);
    const float xb0 = dotf(B0, ex), yb0 = dotf(B0, ey), xc0 = dotf(C0, ex), yc0 = dotf(C0, ey);
    const float za1 = dotf(a1, ez);
    const float xb1 = dotf(b1, ex), yb1 = dotf(b1, ey), zb1 = dotf(b1, ez);
    const float xc1 = dotf(c1, ex), yc1 = dotf(c1, ey), zc1 = dotf(c1, ez);
    const float sinphi = za1 / ra;
    const float cosphi = sqrtf(fmaxf(0.0f, 1.0f - sinphi * sinphi));
    const float sinpsi = (zb1 - zc1) / (2.0f * rc * cosphi);
    const float cospsi = sqrtf(fmaxf(0.0f, 1.0f - sinpsi * sinpsi));
    const float ya2 = ra * cosphi;
    const float xb2 = -rc * cospsi;
    const float t1 = -rb * cosphi, t2 = rc * sinpsi * sinphi;
    const float yb2 = t1 - t2, yc2 = t1 + t2;
    const float alpha = xb2 * (xb0 - xc0) + yb0 * yb2 + yc0 * yc2;
    const float beta = xb2 * (yc0 - yb0) + xb0 * yb2 + xc0 * yc2;
    const float gamma = xb0 * yb1 - xb1 * yb0 + xc0 * yc1 - xc1 * yc0;
    const float ab2 = alpha * alpha + beta * beta;
    const float sinth = (alpha * gamma - beta * sqrtf(fmaxf(0.0f, ab2 - gamma * gamma))) / ab2;
    const float costh = sqrtf(fmaxf(0.0f, 1.0f - sinth * sinth));
    const float3 a3 = make_float3(-ya2 * sinth, ya2 * costh, za1);
    const float3 b3 = make_float3(xb2 * costh - yb2 * sinth, xb2 * sinth + yb2 * costh, zb1);
    const float3 c3 = make_float3(-xb2 * costh - yc2 * sinth, -xb2 * sinth + yc2 * costh, zc1);
    auto world = [&](float3 q, float3 o) {
        return make_float3(q.x * ex.x + q.y * ey.x + q.z * ez.x - o.x, q.x * ex.y + q.y * ey.y + q.z * ez.y - o.y,
                           q.x * ex.z + q.y * ey.z + q.z * ez.z - o.z);
    };
    dA = world(a3, a1);
    dB = world(b3, b1);
    dC = world(c3, c1);
}

}  // namespace

__global__ __launch_bounds__(64) void k_settle(const SettleArgs a) {
    const int tile = blockIdx.x, tid = threadIdx.x;
    const int r = __builtin_amdgcn_readfirstlane(a.tile_rep[tile]);
    const RepState s = a.st[r];
    if (stale(s) || s.step >= s.target) return;   // the step kernel did not step this replica
    const RepParam& p = a.rep[r];
    const float L0 = p.L[0], L1 = p.L[1], L2 = p.L[2], Li0 = p.Linv[0], Li1 = p.Linv[1], Li2 = p.Linv[2];
    const int slot = tile * 64 + tid;
    const int id = a.ids[slot];
    const int role = id >= 0 ? a.rig_of[id] : -1;   // >= 0: O of that molecule; -2: H; -1: unconstrained
    const int cb = cur_buf(s);
    const float4* __restrict__ pin = a.pos[cb];
    float4* __restrict__ pout = a.pos[cb ^ 1];
    const float idt = 1.0f / a.dt;
    double ke2 = 0.0;
    bool trip = false, bad = false;
    auto check = [&](float4 x, float4 pb) {   // skin (src/neighborlist.cu:174) and box checks
        const float dx = minimg(x.x - pb.x, L0, Li0), dy = minimg(x.y - pb.y, L1, Li1), dz = minimg(x.z - pb.z, L2, Li2);
        trip |= fmaf(dz, dz, fmaf(dy, dy, dx * dx)) >= a.skin2q;
        bad |= !(fabsf(x.x) <= L0 && fabsf(x.y) <= L1 && fabsf(x.z) <= L2);
    };
    auto kin = [](float4 v) { return ((double)v.x * v.x + (double)v.y * v.y + (double)v.z * v.z) / (double)v.w; };
    if (role >= 0) {
        const int4 m = a.rig_mol[role];
        const int sb = a.slot_of[m.y], sc = a.slot_of[m.z];
        const float4 xa0 = pin[slot], xb0 = pin[sb], xc0 = pin[sc];
        float4 xa1 = pout[slot], xb1 = pout[sb], xc1 = pout[sc];
        float4 va = a.vel[slot], vb = a.vel[sb], vc = a.vel[sc];   // w = 1 / m_int
        auto mi3 = [&](float3 d) { return make_float3(minimg(d.x, L0, Li0), minimg(d.y, L1, Li1), minimg(d.z, L2, Li2)); };
        const float3 B0 = mi3(sub3(xb0, xa0)), C0 = mi3(sub3(xc0, xa0));
        const float3 B1 = mi3(sub3(xb1, xa1)), C1 = mi3(sub3(xc1, xa1));
        float3 dA, dB, dC;
        settle_one(1.0f / va.w, 1.0f / vb.w, a.rc_hh, a.h, B0, C0, B1, C1, dA, dB, dC);
        xa1 = make_float4(xa1.x + dA.x, xa1.y + dA.y, xa1.z + dA.z, xa1.w);
        xb1 = make_float4(xb1.x + dB.x, xb1.y + dB.y, xb1.z + dB.z, xb1.w);
        xc1 = make_float4(xc1.x + dC.x, xc1.y + dC.y, xc1.z + dC.z, xc1.w);
        va = make_float4((xa1.x - xa0.x) * idt, (xa1.y - xa0.y) * idt, (xa1.z - xa0.z) * idt, va.w);
        vb = make_float4((xb1.x - xb0.x) * idt, (xb1.y - xb0.y) * idt, (xb1.z - xb0.z) * idt, vb.w);
        vc = make_float4((xc1.x - xc0.x) * idt, (xc1.y - xc0.y) * idt, (xc1.z - xc0.z) * idt, vc.w);
        pout[slot] = xa1;
        pout[sb] = xb1;
        pout[sc] = xc1;
        a.vel[slot] = va;
        a.vel[sb] = vb;
        a.vel[sc] = vc;
        ke2 = kin(va) + kin(vb) + kin(vc);
        check(xa1, a.pos_build[slot]);
        check(xb1, a.pos_build[sb]);
        check(xc1, a.pos_build[sc]);
    } else if (role == -1 && id >= 0) {
        ke2 = kin(a.vel[slot]);
        check(pout[slot], a.pos_build[slot]);
    }
    ke2 = wave_sum_d(ke2);
    const bool any_trip = __ballot(trip) != 0, any_bad = __ballot(bad) != 0;
    if (tid == 0) {
        a.part[((size_t)a.par * MD_NSTAT + ACC_KE2) * a.ntiles + tile] = ke2;
        a.part[((size_t)a.par * MD_NSTAT + MD_NSTAT - 1) * a.ntiles + tile] = any_trip ? 1.0 : 0.0;
        if (any_bad) *a.err = 1;
    }
}

extern "C" hipError_t md_launch_settle(const SettleArgs& a, hipStream_t st) {
    hipLaunchKernelGGL(k_settle, dim3(a.ntiles), dim3(64), 0, st, a);
    return hipGetLastError();
}
