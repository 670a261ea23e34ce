// md_rows.h -- the tile neighbor set staged in LDS and the per-atom row walk,
// shared by the step kernel (step.hip) and the force-matching kernels (fm.hip).
//
// A workgroup owns one 64-slot tile.  At its start it stages the tile's
// neighbor set (md_device.h) into LDS: entry k of jset becomes local index
// k + 1 with the position of the recorded image, x_j + k L (the shift table
// s_sh holds k L per shift code), and in w the type of j packed into the three
// low mantissa bits of the staged charge (q_j changes by at most 7 ulp; the type
// selects the LJ row), or for systems without charges the type itself as a
// float.  Local index 0 is a dummy far outside every box, the padding of the
// rows' last blocks, so no entry needs a length test: its pair lies beyond the
// cutoff.  The pair displacement is then x_i - x_j(image): three subtractions,
// no minimum image.  Row entries are byte offsets into the staged set (local
// index times 16), so an entry addresses LDS without a shift.
#pragma once
#include "md_device.h"

// k L per shift code for the replica's box (threads 0..26)
__device__ __forceinline__ void stage_shift_table(float4* s_sh, float L0, float L1, float L2) {
    const int t = threadIdx.x;
    if (t < 27) {
        const int kx = t % 3 - 1, ky = (t / 3) % 3 - 1, kz = t / 9 - 1;
        s_sh[t] = make_float4((float)kx * L0, (float)ky * L1, (float)kz * L2, 0.0f);
    }
}

// w of a staged neighbor: the charge with the type in its low mantissa bits, or
// (CHARGED = false) the type as a float
template <bool CHARGED>
__device__ __forceinline__ float staged_w(float q, uint32_t e) {
    if (CHARGED) return __uint_as_float((__float_as_uint(q) & ~7u) | (e >> MD_JBITS));
    return (float)(e >> MD_JBITS);
}

// The tile's set into s_j[1 .. jn]; s_j[0] = the dummy.  pin = the replica's
// current position buffer.  s_sh must be visible (barrier) before the call;
// the caller places a barrier after it.
template <bool CHARGED = true>
__device__ __forceinline__ void stage_tile_set(float4* s_j, const float4* s_sh, const float4* __restrict__ pin,
                                               const uint32_t* __restrict__ jset, int jn) {
    const int nth = (int)blockDim.x;
    int k = (int)threadIdx.x;
    // four loads in flight per thread before the first store
    for (; k + 3 * nth < jn; k += 4 * nth) {
        uint32_t e[4];
        float4 x[4];
#pragma unroll
        for (int u = 0; u < 4; u++) e[u] = jset[k + u * nth];
#pragma unroll
        for (int u = 0; u < 4; u++) x[u] = pin[e[u] & MD_SLOTMASK];
#pragma unroll
        for (int u = 0; u < 4; u++) {
            const float4 sv = s_sh[(e[u] >> MD_SLOTBITS) & 31];
            s_j[k + u * nth + 1] = make_float4(x[u].x + sv.x, x[u].y + sv.y, x[u].z + sv.z, staged_w<CHARGED>(x[u].w, e[u]));
        }
    }
    for (; k < jn; k += nth) {
        const uint32_t e = jset[k];
        const float4 x = pin[e & MD_SLOTMASK];
        const float4 sv = s_sh[(e >> MD_SLOTBITS) & 31];
        s_j[k + 1] = make_float4(x.x + sv.x, x.y + sv.y, x.z + sv.z, staged_w<CHARGED>(x.w, e));
    }
    if (threadIdx.x == 0) s_j[0] = make_float4(MD_FAR, MD_FAR, MD_FAR, 0.0f);
}

// type of a staged neighbor
__device__ __forceinline__ int staged_type(const float4& p) { return (int)(__float_as_uint(p.w) & 7u); }

// Row walk: lane `sub` of the atom's G lanes loads 16-byte blocks of its row
// (8 local indices: logical entries 8Gb + sub + G u, u = 0..7), two blocks ahead
// of the one it computes, reads the 8 staged neighbors from LDS and calls
// f(pj) for each.  Entries of the last block past the row length nn are the
// dummy.  e0, e1: the row's blocks 0 and 1 (block 1 only read if nb > 1).
template <int G, class F>
__device__ __forceinline__ void walk_row_from(const float4* s_j, const uint4* __restrict__ row4, int nn, int sub,
                                              uint4 e0, uint4 e1, F&& f) {
    const int nb = (int)(((uint32_t)nn + 8 * G - 1) / (8 * G));
    for (int b = 0; b < nb; b++) {
        const uint4 e2 = row4[min(b + 2, nb - 1) * G + sub];
        const uint32_t w[4] = {e0.x, e0.y, e0.z, e0.w};
        float4 pj[8];
#pragma unroll
        for (int u = 0; u < 8; u++)   // entries are byte offsets of the staged neighbors
            pj[u] = *reinterpret_cast<const float4*>(reinterpret_cast<const char*>(s_j) +
                                                     ((u & 1) ? (w[u >> 1] >> 16) : (w[u >> 1] & 0xFFFFu)));
#pragma unroll
        for (int u = 0; u < 8; u++) f(pj[u]);
        e0 = e1;
        e1 = e2;
    }
}

// walk_row_from for a workgroup whose lane groups hold a tile's atoms by
// decreasing row length (the step kernel's rperm order): lane 0 of each
// wavefront holds the wavefront's longest row, so in that row's last block no
// lane of the wavefront has more entries than lane 0 has there.  The wavefront
// computes only that many of the block's 8 entries per lane (rounded up to a
// pair); the ones it leaves out are the dummy for every lane, whose pair is
// beyond the cutoff, so the sums are the same bits as walk_row_from's.
template <int G, class F>
__device__ __forceinline__ void walk_row_sorted(const float4* s_j, const uint4* __restrict__ row4, int nn, int sub,
                                                uint4 e0, uint4 e1, F&& f) {
    const int nb = (int)(((uint32_t)nn + 8 * G - 1) / (8 * G));
    const int nn0 = __builtin_amdgcn_readfirstlane(nn);
    const int nb0 = (nn0 + 8 * G - 1) / (8 * G);
    const int last0 = (nn0 - 8 * G * (nb0 - 1) + G - 1) / G;   // lane 0's entries in its last block
    for (int b = 0; b < nb; b++) {
        const uint4 e2 = row4[min(b + 2, nb - 1) * G + sub];
        const uint32_t w[4] = {e0.x, e0.y, e0.z, e0.w};
        auto at = [&](int u) {   // entries are byte offsets of the staged neighbors
            return *reinterpret_cast<const float4*>(reinterpret_cast<const char*>(s_j) +
                                                    ((u & 1) ? (w[u >> 1] >> 16) : (w[u >> 1] & 0xFFFFu)));
        };
        const int cnt = b < nb0 - 1 ? 8 : last0;   // wavefront-uniform
        if (cnt > 6) {
            float4 pj[8];
#pragma unroll
            for (int u = 0; u < 8; u++) pj[u] = at(u);
#pragma unroll
            for (int u = 0; u < 8; u++) f(pj[u]);
        } else if (cnt > 4) {
            float4 pj[6];
#pragma unroll
            for (int u = 0; u < 6; u++) pj[u] = at(u);
#pragma unroll
            for (int u = 0; u < 6; u++) f(pj[u]);
        } else if (cnt > 2) {
            float4 pj[4];
#pragma unroll
            for (int u = 0; u < 4; u++) pj[u] = at(u);
#pragma unroll
            for (int u = 0; u < 4; u++) f(pj[u]);
        } else {
            const float4 p0 = at(0), p1 = at(1);
            f(p0);
            f(p1);
        }
        e0 = e1;
        e1 = e2;
    }
}

template <int G, class F>
__device__ __forceinline__ void walk_row(const float4* s_j, const uint4* __restrict__ row4, int nn, int sub, F&& f) {
    const int nb = (int)(((uint32_t)nn + 8 * G - 1) / (8 * G));
    if (nb == 0) return;
    walk_row_from<G>(s_j, row4, nn, sub, row4[sub], row4[min(1, nb - 1) * G + sub], f);
}
