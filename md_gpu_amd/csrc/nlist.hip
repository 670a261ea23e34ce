// nlist.hip -- binned cell-list neighbor build with space-filling-curve
// ordering of the atoms (the reference's intended neighborlist_binned,
// src/neighborlist_binned.cu:20-119 / src/neighborlist_kernel_o.cu:20-261,
// which never compiled; the list semantics are those of the O(N^2) build it
// replaces, src/neighborlist.cu:231-297).
//
// A rebuild is six kernels.  Each runs only for replicas whose list is stale
// (RepState.trip, raised by the position update when some atom moved by at
// least skin/2 since the last build: the skin criterion of
// src/neighborlist.cu:157-176), so the chain can sit in every graph segment:
//   k_nl_bin      wrap every coordinate into the box, cell of every atom (cells
//                 are ranked along a Hilbert curve), arrival rank in the cell
//   k_nl_scan     exclusive scan of cell counts -> cell_start (one block per replica)
//   k_nl_scatter  atom -> provisional slot (cell order, arrival order)
//   k_nl_permute  deterministic order inside a cell (by eighth of the cell, then
//                 original index); gathers pos/vel/force/id/type into scratch in the
//                 new order, positions wrapped into the box
//   k_nl_tbuild   one workgroup per 64-slot tile: the image cells within r_list of
//                 the tile's bounding box are listed in LDS, their members streamed
//                 through an LDS ring and tested against the tile's atoms with the
//                 canonical FP32 predicate.  Candidates that some atom of the tile
//                 hits form the tile's neighbor set (in candidate order); each atom's
//                 row holds 16-bit indices into that set, in candidate order, so rows
//                 are deterministic.  Each tile stores its longest row and set size.
//   k_nl_commit   per rebuilt replica: new list epoch (or the next step kernel opens it).
// Bytes per atom per rebuild (algorithmic): read pos/vel/force/id/type 56 B,
// write them twice (scratch + copy back) 112 B, plus the rows 2*K B and the
// tile sets (4 B per set entry, about 4*K/16 B per atom).
#include "md_device.h"

namespace {

constexpr int kTab = 1024;    // region cells per table batch

__device__ __forceinline__ int cell_coord(float x, float halfL, float cinv, int nc) {
    int k = (int)floorf((x + halfL) * cinv);
    return min(max(k, 0), nc - 1);
}

// floor(a / b), b > 0, without an integer division: the estimate from the rounded
// reciprocal of b is off by at most one for |a| < 2^24, and one correction step
// makes it exact.  (Region walks never come near that range.)
__device__ __forceinline__ int floor_div(int a, int b, float inv_b) {
    int q = (int)floorf(((float)a + 0.5f) * inv_b);
    const int r = a - q * b;
    return q + (r >= b ? 1 : 0) - (r < 0 ? 1 : 0);
}

// One axis of the image-cell region around a tile's bounding box [bmin, bmax].
struct Axis {
    int lo, nr, nc;
    float cedge, halfL, bmin, bmax, inv_nc, inv_nr;
    // image index k, wrapped cell w of region coordinate c; returns the gap between
    // the cell slab and the box along this axis (0 when they overlap)
    __device__ float gap(int c, int& k, int& w) const {
        k = floor_div(c, nc, inv_nc);
        w = c - k * nc;
        const float cl = c * cedge - halfL, ch = cl + cedge;
        return fmaxf(0.0f, fmaxf(cl - bmax, bmin - ch));
    }
};

__device__ __forceinline__ Axis make_axis(float bmin, float bmax, float rlp, float halfL, float cinv,
                                          float cedge, int nc) {
    Axis x;
    // image offsets beyond one box length never hold a candidate (r_list < L/2),
    // and the bounds keep the walk finite for any input
    const int l = max((int)floorf((bmin + halfL - rlp) * cinv), -nc);
    const int h = min((int)floorf((bmax + halfL + rlp) * cinv), 2 * nc - 1);
    x.lo = l;
    x.nr = max(h - l + 1, 0);
    x.nc = nc;
    x.inv_nc = 1.0f / (float)nc;
    x.inv_nr = 1.0f / (float)max(x.nr, 1);
    x.cedge = cedge;
    x.halfL = halfL;
    x.bmin = bmin;
    x.bmax = bmax;
    return x;
}

// exclusive scan over an NT-thread block; returns this thread's offset, sets total
template <int NT>
__device__ int block_scan(int v, int* s_w, int& total) {
    constexpr int NW = NT / 64;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    int inc = v;
    for (int o = 1; o < 64; o <<= 1) {
        int t = __shfl_up(inc, o, 64);
        if (lane >= o) inc += t;
    }
    if (lane == 63) s_w[w] = inc;
    __syncthreads();
    int off = 0;
    total = 0;
    for (int k = 0; k < NW; k++) {
        const int t = s_w[k];
        if (k < w) off += t;
        total += t;
    }
    __syncthreads();
    return off + inc - v;
}

}  // namespace

// The short rebuild kernels first read what says whether their wavefront has
// work (the gate counter, the tile's replica and that replica's state: scalar
// loads, a wavefront's 64 slots being one tile of one replica), and only the
// wavefronts of replicas being rebuilt go on to their slots' data.  In a
// batched run most replicas are not rebuilt at a given check, and their slots'
// data (tens of MB over all replicas) is then never read.
__global__ __launch_bounds__(256) void k_nl_bin(BuildArgs a) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= a.nslot) return;
    const int n_stale = a.gate->n_stale;
    const int r = __builtin_amdgcn_readfirstlane(a.tile_rep[s >> 6]);
    loads_issued();
    if (n_stale == 0) return;
    const RepState st = a.st[r];
    loads_issued();
    if (!stale(st)) return;
    const int id = a.ids[s];
    const float4 x0 = a.pos[0][s], x1 = a.pos[1][s];
    const RepParam& p = a.rep[r];
    const float h0 = p.halfL[0], h1 = p.halfL[1], h2 = p.halfL[2];
    const float c0 = p.cinv[0], c1 = p.cinv[1], c2 = p.cinv[2];
    const float L0 = p.L[0], L1 = p.L[1], L2 = p.L[2], Li0 = p.Linv[0], Li1 = p.Linv[1], Li2 = p.Linv[2];
    const int n0 = p.nc[0], n1 = p.nc[1], n2 = p.nc[2], cell_off = p.cell_off, slot_off = p.slot_off;
    loads_issued();
    if (s == slot_off) {
        a.st[r].trip[(st.nb + 1) & 1] = 0;   // the buffer the next epoch will use
        a.repw[r].Lb[0] = L0;                 // the box of this build (the skin check under pressure coupling)
        a.repw[r].Lb[1] = L1;
        a.repw[r].Lb[2] = L2;
    }
    if (id < 0) return;
    const float4 xr = cur_buf(st) ? x1 : x0;
    // the wrapped position k_nl_permute stores (same operations)
    const float x = wrap_coord(xr.x, L0, Li0), y = wrap_coord(xr.y, L1, Li1), z = wrap_coord(xr.z, L2, Li2);
    int cx = cell_coord(x, h0, c0, n0);
    int cy = cell_coord(y, h1, c1, n1);
    int cz = cell_coord(z, h2, c2, n2);
    int lin = cx + n0 * (cy + n1 * cz);
    int gc = cell_off + a.cell_rank[cell_off + lin];
    // the atom's eighth of the cell, numbered along a path through adjacent eighths
    const int sx = (x + h0) * c0 - (float)cx >= 0.5f;
    const int sy = (y + h1) * c1 - (float)cy >= 0.5f;
    const int sz = (z + h2) * c2 - (float)cz >= 0.5f;
    const int sub = (0x54672310 >> (4 * (sx | (sy << 1) | (sz << 2)))) & 7;
    a.key[s] = gc;
    a.rank[s] = atomicAdd(&a.cell_count[gc], 1) | (sub << MD_SUBSHIFT);
}

__global__ __launch_bounds__(1024) void k_nl_scan(BuildArgs a) {
    __shared__ int wsum[16];
    const int r = blockIdx.x;
    const int n_stale = a.gate->n_stale;
    const RepState st = a.st[r];
    const RepParam& p = a.rep[r];
    const int nc = p.nc[0] * p.nc[1] * p.nc[2];
    const int cell_off = p.cell_off, cs_off = p.cs_off, slot_off = p.slot_off, natom = p.natom;
    loads_issued();
    if (n_stale == 0) return;
    if (!stale(st)) return;
    const int per = (nc + 1023) / 1024;
    const int b = min(nc, (int)threadIdx.x * per), e = min(nc, b + per);
    int sum = 0;
#pragma unroll 8
    for (int k = b; k < e; k++) sum += a.cell_count[cell_off + k];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    int inc = sum;
    for (int o = 1; o < 64; o <<= 1) {
        int v = __shfl_up(inc, o, 64);
        if (lane >= o) inc += v;
    }
    if (lane == 63) wsum[w] = inc;
    __syncthreads();
    if (threadIdx.x == 0) {
        int run = 0;
        for (int k = 0; k < 16; k++) { int t = wsum[k]; wsum[k] = run; run += t; }
    }
    __syncthreads();
    int run = slot_off + wsum[w] + inc - sum;
#pragma unroll 8
    for (int k = b; k < e; k++) {
        int c = a.cell_count[cell_off + k];
        a.cell_start[cs_off + k] = run;
        run += c;
        a.cell_count[cell_off + k] = 0;   // ready for the next rebuild
    }
    if (threadIdx.x == 0) a.cell_start[cs_off + nc] = slot_off + natom;
}

__global__ __launch_bounds__(256) void k_nl_scatter(BuildArgs a) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= a.nslot) return;
    const int n_stale = a.gate->n_stale;
    const int r = __builtin_amdgcn_readfirstlane(a.tile_rep[s >> 6]);
    loads_issued();
    if (n_stale == 0) return;
    const RepState st = a.st[r];   // whole (no load that waits for nb)
    loads_issued();
    if (!stale(st)) return;
    const int id = a.ids[s];
    const int gc = a.key[s];
    const int rk = a.rank[s];
    loads_issued();
    if (id < 0) return;
    int t = a.cell_start[gc + r] + (rk & ((1 << MD_SUBSHIFT) - 1));
    a.tmp[t] = s;
    a.tmpid[t] = id | ((rk >> MD_SUBSHIFT) << MD_SUBSHIFT);   // order key: eighth of the cell, then index
}

__global__ __launch_bounds__(256) void k_nl_permute(BuildArgs a) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= a.nslot) return;
    const int n_stale = a.gate->n_stale;
    const int r = __builtin_amdgcn_readfirstlane(a.tile_rep[t >> 6]);
    loads_issued();
    if (n_stale == 0) return;
    const RepState st = a.st[r];
    loads_issued();
    if (!stale(st)) return;
    // tmp holds a slot of this replica for every t below slot_off + natom
    const RepParam& p = a.rep[r];
    const int lim = p.slot_off + p.natom;
    if (t >= lim) return;
    const int o = a.tmp[t];
    const int mykey = a.tmpid[t];
    const float L0 = p.L[0], L1 = p.L[1], L2 = p.L[2], Li0 = p.Linv[0], Li1 = p.Linv[1], Li2 = p.Linv[2];
    loads_issued();
    const int gc = a.key[o];
    const float4 xr = cur_buf(st) ? a.pos[1][o] : a.pos[0][o];
    const float4 v = a.vel[o], f4 = a.force[o];
    const int ty = a.typ[o];
    const float4 x = make_float4(wrap_coord(xr.x, L0, Li0), wrap_coord(xr.y, L1, Li1), wrap_coord(xr.z, L2, Li2), xr.w);
    const int myid = mykey & ((1 << MD_SUBSHIFT) - 1);
    int cs = a.cell_start[gc + r], ce = a.cell_start[gc + r + 1];
    int cnt = 0;
#pragma unroll 4
    for (int u = cs; u < ce; u++) cnt += (a.tmpid[u] < mykey);
    int f = cs + cnt;
    a.pos_s[f] = x;
    a.vel_s[f] = v;
    a.force_s[f] = f4;
    a.ids_s[f] = myid;
    a.typ_s[f] = ty;
    a.slot_of[myid] = f;
}

// Tile neighbor-set build: one workgroup of NW wavefronts (4 or 8) per 64-slot
// tile, 64 / NW atoms per wavefront (held as wavefront-uniform values).  The set
// and rows do not depend on NW: four wavefronts (more workgroups resident per CU
// to overlap the latency-bound candidate streaming) build the same lists as eight.
//  1. The image cells within r_list of the tile's bounding box are listed in LDS
//     (region order, two-pass block scan), in batches of kTab region cells.
//  2. Their members are streamed 64 NW per chunk (a thread per candidate); those
//     whose image lies r_list or farther from the box are dropped, the rest are
//     appended to an LDS ring in candidate order (stable block compaction).
//  3. Every 64 ring entries, each wavefront tests the 64 candidates (a lane each)
//     against its 8 atoms with the canonical predicate; a candidate that hits any
//     atom of the tile gets the next local index of the tile's set (the OR of the
//     wavefronts' hit masks goes through LDS, one barrier per 64 candidates), and
//     each atom's hits are appended to its row at count + (hits in lower lanes).
// So set and rows follow candidate order and do not depend on scheduling.
template <bool EXCL, int NW>
__global__ __launch_bounds__(64 * NW, (NW == 4 && !EXCL) ? 5 : 1) void k_nl_tbuild(BuildArgs a) {
    constexpr int NT = 64 * NW, AW = 64 / NW, RING = 1024, PER = kTab / NT;
    __shared__ int s_tstart[kTab];        // first slot of the cell
    __shared__ int s_tpref[kTab + 1];     // first candidate index of each table row; [ntab] = total
    __shared__ int s_tcode[kTab];         // image shift code of the cell
    __shared__ float4 s_cx[RING];         // ring: candidate position (unshifted), w = bits of the set entry
    __shared__ int s_cid[EXCL ? RING : 1];   // original index of the candidate (exclusions)
    __shared__ float4 s_shv[27];          // -k L per shift code (canonical displacement (xi - xj) + s)
    __shared__ float4 s_ax[64];           // the tile's atoms (x, y, z, bits of the id)
    __shared__ int4 s_ex[EXCL ? 64 : 1];
    __shared__ unsigned long long s_any[2][NW];
    __shared__ float s_bb[6];
    __shared__ int s_w[NW];
    __shared__ int s_wmax[NW];
    __shared__ int s_cnt[64];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int tile = blockIdx.x;
    const int slot0 = tile * 64;
    // the gate, the replica index and the tile's atoms in one round trip
    const int n_stale = a.gate->n_stale;
    const int r = __builtin_amdgcn_readfirstlane(a.tile_rep[tile]);
    int id_pre = -1;
    float4 x_pre = make_float4(0.f, 0.f, 0.f, 0.f);
    if (tid < 64) {
        id_pre = a.ids_s[slot0 + tid];
        x_pre = a.pos_s[slot0 + tid];
    }
    loads_issued();
    if (n_stale == 0) return;
    const RepState st = a.st[r];
    loads_issued();
    if (!stale(st)) return;
    const RepParam& p = a.rep[r];
    const int cb = cur_buf(st);
    const float L0 = p.L[0], L1 = p.L[1], L2 = p.L[2];
    if (tid < 27) {
        const int kx = tid % 3 - 1, ky = (tid / 3) % 3 - 1, kz = tid / 9 - 1;
        s_shv[tid] = make_float4(-(float)kx * L0, -(float)ky * L1, -(float)kz * L2, 0.0f);
    }
    if (tid < 64) {
        const bool act = id_pre >= 0;
        s_ax[tid] = make_float4(act ? x_pre.x : MD_FAR, act ? x_pre.y : MD_FAR, act ? x_pre.z : MD_FAR,
                                __int_as_float(id_pre));
        if (EXCL) s_ex[tid] = act ? a.excl[id_pre] : make_int4(-1, -1, -1, -1);
        // tile bounding box over the active atoms
        float lx = act ? x_pre.x : 3.0e38f, ly = act ? x_pre.y : 3.0e38f, lz = act ? x_pre.z : 3.0e38f;
        float hx = act ? x_pre.x : -3.0e38f, hy = act ? x_pre.y : -3.0e38f, hz = act ? x_pre.z : -3.0e38f;
        for (int o = 32; o >= 1; o >>= 1) {
            lx = fminf(lx, __shfl_xor(lx, o, 64)); ly = fminf(ly, __shfl_xor(ly, o, 64)); lz = fminf(lz, __shfl_xor(lz, o, 64));
            hx = fmaxf(hx, __shfl_xor(hx, o, 64)); hy = fmaxf(hy, __shfl_xor(hy, o, 64)); hz = fmaxf(hz, __shfl_xor(hz, o, 64));
        }
        if (tid == 0) {
            s_bb[0] = lx; s_bb[1] = ly; s_bb[2] = lz;
            s_bb[3] = hx; s_bb[4] = hy; s_bb[5] = hz;
        }
    }
    __syncthreads();
    const float lx = s_bb[0], ly = s_bb[1], lz = s_bb[2], hx = s_bb[3], hy = s_bb[4], hz = s_bb[5];
    // this wavefront's atoms as wavefront-uniform values; an empty slot sits far
    // outside the box, so no candidate passes the distance test for it
    float axr[AW], ayr[AW], azr[AW];
    int cnt[AW];
#pragma unroll
    for (int i = 0; i < AW; i++) {
        const float4 q = s_ax[wave * AW + i];
        axr[i] = __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(q.x)));
        ayr[i] = __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(q.y)));
        azr[i] = __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(q.z)));
        cnt[i] = 0;
    }
    const int gs = __builtin_ctz(a.row_g);   // row layout: lanes per atom of the step kernel, a power of 2
    const int cap = a.cap, jcap = a.jcap;
    const float rl2 = a.rl2;
    uint16_t* const wrow = a.nl16 + (size_t)(slot0 + wave * AW) * cap;   // row of the wavefront's first atom
    uint32_t* const jset = a.jset + (size_t)tile * jcap;
    const unsigned long long below = (1ull << lane) - 1ull;
    int jn = 0;     // size of the tile's set so far (workgroup-uniform)
    int head = 0, tail = 0, nproc = 0;   // ring positions (taken mod RING), sub-chunks processed

    // test ring entries [head, head + n) against this wavefront's atoms; assign set indices
    auto process = [&](int n) {
        const int q = (head + lane) & (RING - 1);
        const bool valid = lane < n;
        const float4 cx = s_cx[q];
        const uint32_t info = __float_as_uint(cx.w);
        const float4 sv = s_shv[valid ? (info >> MD_SLOTBITS) & 31 : MD_SHIFT0];   // ring data past n is stale
        const int js = (int)(info & MD_SLOTMASK);
        const int jid = EXCL ? s_cid[q] : 0;
        unsigned long long m[AW];
        unsigned long long any = 0;
        // each condition is balloted on its own and the masks are combined as
        // scalars: a compare then writes its lane mask directly
        const unsigned long long vmask = __builtin_amdgcn_ballot_w64(valid);
#pragma unroll
        for (int i = 0; i < AW; i++) {
            // (xi - xj) - k L: the canonical minimum-image displacement for this image
            const float dx = (axr[i] - cx.x) + sv.x;
            const float dy = (ayr[i] - cx.y) + sv.y;
            const float dz = (azr[i] - cx.z) + sv.z;
            const float rsq = fmaf(dz, dz, fmaf(dy, dy, dx * dx));
            unsigned long long mi = vmask & __builtin_amdgcn_ballot_w64(rsq < rl2) &
                                    __builtin_amdgcn_ballot_w64(js != slot0 + wave * AW + i);
            if (EXCL) {
                const int4 ex = s_ex[wave * AW + i];
                mi &= __builtin_amdgcn_ballot_w64(jid != ex.x) & __builtin_amdgcn_ballot_w64(jid != ex.y) &
                      __builtin_amdgcn_ballot_w64(jid != ex.z) & __builtin_amdgcn_ballot_w64(jid != ex.w);
            }
            m[i] = mi;
            any |= mi;
        }
        const int par = nproc & 1;
        if (lane == 0) s_any[par][wave] = any;
        __syncthreads();
        any = 0;
#pragma unroll
        for (int w = 0; w < NW; w++) any |= s_any[par][w];
        const bool mine = (any >> lane) & 1ull;
        const int li = jn + 1 + (int)__builtin_amdgcn_mbcnt_hi((uint32_t)(any >> 32),
                                                               __builtin_amdgcn_mbcnt_lo((uint32_t)any, 0u));
        if (wave == 0 && mine && li - 1 < jcap) jset[li - 1] = info;
        // a set that outgrew its capacity is rebuilt after the host grows it (the run
        // is rolled back); until then its rows point past it only at the dummy.  Row
        // entries are byte offsets of the staged entry (li x 16, li <= MD_JMAX).
        const uint16_t lis = (uint16_t)(li - 1 < jcap ? li << 4 : 0);
#pragma unroll
        for (int i = 0; i < AW; i++) {
            if (m[i]) {
                const uint32_t e = __builtin_amdgcn_mbcnt_hi((uint32_t)(m[i] >> 32),
                                                             __builtin_amdgcn_mbcnt_lo((uint32_t)m[i], (uint32_t)cnt[i]));
                const int nh = __popcll(m[i]);
                // the hit lanes are the mask itself (no per-lane bit test); the
                // capacity test per lane only where this batch can cross it
                if (cnt[i] + nh <= cap) {
                    if (__builtin_amdgcn_inverse_ballot_w64(m[i])) wrow[(uint32_t)(i * cap) + lpos16_u(e, gs)] = lis;
                } else {
                    if (__builtin_amdgcn_inverse_ballot_w64(m[i]) && e < (uint32_t)cap)
                        wrow[(uint32_t)(i * cap) + lpos16_u(e, gs)] = lis;
                }
                cnt[i] += nh;
            }
        }
        jn += __popcll(any);
        head += n;
        nproc++;
    };

    if (lx <= hx) {
        const float rlp = sqrtf(rl2) * 1.0001f + 1e-4f;
        const float rlp2 = rlp * rlp;
        // image-cell region covering the tile box grown by r_list
        const Axis ax = make_axis(lx, hx, rlp, p.halfL[0], p.cinv[0], p.cedge[0], p.nc[0]);
        const Axis ay = make_axis(ly, hy, rlp, p.halfL[1], p.cinv[1], p.cedge[1], p.nc[1]);
        const Axis az = make_axis(lz, hz, rlp, p.halfL[2], p.cinv[2], p.cedge[2], p.nc[2]);
        const int nreg = ax.nr * ay.nr * az.nr;
        const int cell_off = p.cell_off, cs_off = p.cs_off, ncx = p.nc[0], ncy = p.nc[1];
        // region cell c -> its cell's linear index and image shift code; -1 when the
        // image cell lies r_list or farther from the box
        auto locate = [=](int c, int& code) -> int {
            int kx, ky, kz, wx, wy, wz;
            const int cyz = floor_div(c, ax.nr, ax.inv_nr), cz = floor_div(cyz, ay.nr, ay.inv_nr);
            const float ddx = ax.gap(ax.lo + (c - cyz * ax.nr), kx, wx);
            const float ddy = ay.gap(ay.lo + (cyz - cz * ay.nr), ky, wy);
            const float ddz = az.gap(az.lo + cz, kz, wz);
            if (ddx * ddx + ddy * ddy + ddz * ddz >= rlp2) return -1;
            code = shift_code(kx, ky, kz);
            return wx + ncx * (wy + ncy * wz);
        };
        for (int b0 = 0; b0 < nreg; b0 += kTab) {
            // ---- table of non-empty image cells within r_list of the box (region order);
            //      pass 1 counts, pass 2 writes
            const int c0 = b0 + tid * PER, c1 = min(min(c0 + PER, nreg), b0 + kTab);
            // the thread's PER region cells: geometry first, then all their rank
            // loads, then all their bounds loads (two round trips, not 2 PER)
            int pcell[PER], pcode[PER], prank[PER], pcs[PER], pn[PER];
#pragma unroll
            for (int u = 0; u < PER; u++) {
                pcode[u] = 0;
                pcell[u] = c0 + u < c1 ? locate(c0 + u, pcode[u]) : -1;
            }
#pragma unroll
            for (int u = 0; u < PER; u++) prank[u] = pcell[u] >= 0 ? a.cell_rank[cell_off + pcell[u]] : 0;
#pragma unroll
            for (int u = 0; u < PER; u++) {
                pcs[u] = 0;
                pn[u] = 0;
                if (pcell[u] >= 0) {
                    pcs[u] = a.cell_start[cs_off + prank[u]];
                    pn[u] = a.cell_start[cs_off + prank[u] + 1] - pcs[u];
                }
            }
            int n_keep = 0, csum = 0;
#pragma unroll
            for (int u = 0; u < PER; u++) {
                n_keep += pn[u] > 0;
                csum += pn[u];
            }
            int ntab, T;
            const int off = block_scan<NT>(n_keep, s_w, ntab);
            const int coff = block_scan<NT>(csum, s_w, T);
#pragma unroll
            for (int u = 0, k = off, run = coff; u < PER; u++) {
                if (pn[u] > 0) {
                    s_tstart[k] = pcs[u];
                    s_tpref[k] = run;
                    s_tcode[k] = pcode[u];
                    k++;
                    run += pn[u];
                }
            }
            if (tid == 0) s_tpref[ntab] = T;
            __syncthreads();
            // ---- stream the candidates, NT per chunk, into the ring (stable compaction).
            //      The next chunk's members are located and their loads issued before
            //      this chunk is tested, so the loads are in flight during the tests.
            auto find = [&](int q, int& code) -> int {
                int lo_ = 0, hi_ = ntab - 1;   // last table row with pref <= q
                while (lo_ < hi_) {
                    const int mid = (lo_ + hi_ + 1) >> 1;
                    if (s_tpref[mid] <= q) lo_ = mid; else hi_ = mid - 1;
                }
                code = s_tcode[lo_];
                return s_tstart[lo_] + (q - s_tpref[lo_]);
            };
            int nj = -1, ncode = 0, nty = 0, nid = 0;
            float4 nx = make_float4(0.f, 0.f, 0.f, 0.f);
            if (tid < T) {
                nj = find(tid, ncode);
                nx = a.pos_s[nj];
                nty = a.typ_s[nj];
                if (EXCL) nid = a.ids_s[nj];
            }
            for (int base = 0; base < T; base += NT) {
                const int j = nj, code = ncode, ty = nty, cid = nid;
                float4 cx = nx;
                nj = -1;
                if (base + NT + tid < T) {
                    nj = find(base + NT + tid, ncode);
                    nx = a.pos_s[nj];
                    nty = a.typ_s[nj];
                    if (EXCL) nid = a.ids_s[nj];
                }
                bool keep = false;
                if (j >= 0) {
                    const float4 sv = s_shv[code];
                    const float ix = cx.x - sv.x, iy = cx.y - sv.y, iz = cx.z - sv.z;   // image position
                    const float gx = fmaxf(0.0f, fmaxf(lx - ix, ix - hx));
                    const float gy = fmaxf(0.0f, fmaxf(ly - iy, iy - hy));
                    const float gz = fmaxf(0.0f, fmaxf(lz - iz, iz - hz));
                    keep = gx * gx + gy * gy + gz * gz < rlp2;
                    cx.w = __uint_as_float((uint32_t)j | ((uint32_t)code << MD_SLOTBITS) | ((uint32_t)ty << MD_JBITS));
                }
                int kept;
                const int at = block_scan<NT>(keep ? 1 : 0, s_w, kept);
                if (keep) {
                    const int k = (tail + at) & (RING - 1);
                    s_cx[k] = cx;
                    if (EXCL) s_cid[k] = cid;
                }
                __syncthreads();
                tail += kept;
                while (tail - head >= 64) process(64);   // at most 63 entries stay pending
            }
        }
        if (tail > head) process(tail - head);
    }
    // row lengths, row padding up to the end of the last block (the dummy, local
    // index 0), longest row of the wavefront
    const int blk = 8 << gs;
    int mx = 0;
#pragma unroll
    for (int i = 0; i < AW; i++) {
        const int c = cnt[i];
        mx = max(mx, c);
        const int endb = min((c + blk - 1) / blk * blk, cap);
        for (int e = c + lane; e < endb; e += 64) wrow[(size_t)i * cap + lpos16(e, blk / 8)] = 0;
        if (lane == i) {
            const int cc = __float_as_int(s_ax[wave * AW + i].w) >= 0 ? c : 0;
            a.ncount[slot0 + wave * AW + i] = cc;
            s_cnt[wave * AW + i] = cc;
        }
    }
    if (lane == 0) s_wmax[wave] = mx;
    __syncthreads();
    // the tile's atoms by decreasing row length (ties by slot): the step kernel's
    // wavefronts then hold atoms whose rows end in the same block
    if (tid < 64) {
        const int c = s_cnt[tid];
        int rank = 0;
        for (int k = 0; k < 64; k++) {
            const int d = s_cnt[k];
            rank += (d > c) || (d == c && k < tid);
        }
        a.rperm[slot0 + rank] = (uint8_t)tid;
    }
    // copy the tile back from scratch into the primary arrays (new order)
    if (tid < 64) {
        const int s = slot0 + tid;
        const float4 x = a.pos_s[s];
        a.pos[cb][s] = x;
        a.pos_build[s] = x;
        a.vel[s] = a.vel_s[s];
        a.force[s] = a.force_s[s];
        a.ids[s] = a.ids_s[s];
        a.typ[s] = a.typ_s[s];
    }
    if (tid == 0) {   // longest row and set size of the tile; k_nl_commit opens the new list epoch
        int m = s_wmax[0];
        for (int w = 1; w < NW; w++) m = max(m, s_wmax[w]);
        a.tilemax[tile] = max(a.tilemax[tile], m);   // running maxima since the host cleared them
        a.jmax[tile] = max(a.jmax[tile], jn);
        a.jcount[tile] = min(jn, jcap);
    }
}

// Last kernel of a rebuild chain that is not followed by a deferred-finalize
// block: every rebuilt replica (one wavefront each) opens its new list epoch
// (a deferred-finalize block's first step kernel opens it instead).  The gate
// counters change once per workgroup.  No early exit: every wavefront reaches
// the barrier.
__global__ __launch_bounds__(256) void k_nl_commit(BuildArgs a) {
    __shared__ int s_d[4][3];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int r = blockIdx.x * 4 + wave;
    int done = 0, to_run = 0, to_force = 0;
    if (r < a.nrep) {
        const RepState st = a.st[r];
        if (stale(st)) {
            if (lane == 0) {
                a.st[r].age = 0;
                a.st[r].nb = st.nb + 1;
            }
            done = 1;
            to_run = (a.lfgate || st.phase == 0) && st.step < st.target;
            to_force = !a.lfgate && st.phase == 1;
        }
    }
    if (lane == 0) {
        s_d[wave][0] = done;
        s_d[wave][1] = to_run;
        s_d[wave][2] = to_force;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        int n = 0, nr = 0, nf = 0;
        for (int w = 0; w < 4; w++) { n += s_d[w][0]; nr += s_d[w][1]; nf += s_d[w][2]; }
        if (n) atomicSub(&a.gate->n_stale, n);
        if (nr) atomicAdd(&a.gate->n_run, nr);
        if (nf) atomicAdd(&a.gate->n_force, nf);
    }
}

// Mark the lists of all replicas stale (host-requested full rebuild).
__global__ void k_force_trip(RepState* st, Gate* g, int nrep) {
    for (int r = threadIdx.x; r < nrep; r += blockDim.x) st[r].trip[st[r].nb & 1] = 1;
    if (threadIdx.x == 0) {
        g->n_stale = nrep;
        g->n_run = 0;
        g->n_force = 0;
    }
}

// ---------------------------------------------------------------- invariants
// Consistency of the sorted atom order and the lists, checked on the device
// (test support, no counterpart in the reference).  Counts into bad[]:
//   0: an active slot s with slot_of[ids[s]] != s (the order is not a bijection)
//   1: a row entry outside the tile's set (local index 0 or beyond jcount)
//   2: a row pair (i, j) whose image distance at the build positions is not
//      below r_list, or j = i
//   3: a set entry whose slot is not an active slot of the tile's replica
//   4: a row longer than the capacity or a set larger than its capacity
//   5: a row block padding entry that is not the dummy
// One thread per atom slot; the caller has just built the lists (pos_build).
__global__ __launch_bounds__(256) void k_check_lists(BuildArgs a, int row_g, int* bad) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= a.nslot) return;
    const int tile = s >> 6;
    const int r = a.tile_rep[tile];
    const RepParam& p = a.rep[r];
    const int id = a.ids[s];
    if (id < 0) return;
    if (id >= 0 && a.slot_of[id] != s) atomicAdd(&bad[0], 1);
    const int nn = a.ncount[s];
    const int jn = a.jcount[tile];
    if (nn > a.cap || a.jmax[tile] > a.jcap) atomicAdd(&bad[4], 1);
    const float4 xi = a.pos_build[s];
    const uint16_t* row = a.nl16 + (size_t)s * a.cap;
    const uint32_t* js = a.jset + (size_t)tile * a.jcap;
    const int n = min(nn, a.cap);
    for (int e = 0; e < n; e++) {
        const int off = row[lpos16(e, row_g)];
        const int li = off >> 4;
        if ((off & 15) || li < 1 || li > jn) {
            atomicAdd(&bad[1], 1);
            continue;
        }
        const uint32_t info = js[li - 1];
        const int j = (int)(info & MD_SLOTMASK), code = (int)((info >> MD_SLOTBITS) & 31);
        if (j < p.slot_off || j >= p.slot_off + p.nslot || a.ids[j] < 0) {
            atomicAdd(&bad[3], 1);
            continue;
        }
        const float4 xj = a.pos_build[j];
        const int kx = code % 3 - 1, ky = (code / 3) % 3 - 1, kz = code / 9 - 1;
        // the image shift of the build's box (a coupled box has moved on since)
        const float dx = (xi.x - xj.x) + (-(float)kx * p.Lb[0]);
        const float dy = (xi.y - xj.y) + (-(float)ky * p.Lb[1]);
        const float dz = (xi.z - xj.z) + (-(float)kz * p.Lb[2]);
        const float rsq = fmaf(dz, dz, fmaf(dy, dy, dx * dx));
        if (!(rsq < a.rl2) || j == s) atomicAdd(&bad[2], 1);
    }
    const int blk = 8 * row_g;
    const int endb = min((nn + blk - 1) / blk * blk, a.cap);
    for (int e = n; e < endb; e++)
        if (row[lpos16(e, row_g)] != 0) atomicAdd(&bad[5], 1);
}

// ---------------------------------------------------------------- launchers

// commit = 0: a deferred-finalize block follows and its first step kernel opens the new lists
extern "C" hipError_t md_launch_rebuild(const BuildArgs& a, int commit, hipStream_t st) {
    const int nb = (a.nslot + 255) / 256;
    hipLaunchKernelGGL(k_nl_bin, dim3(nb), dim3(256), 0, st, a);
    hipLaunchKernelGGL(k_nl_scan, dim3(a.nrep), dim3(1024), 0, st, a);
    hipLaunchKernelGGL(k_nl_scatter, dim3(nb), dim3(256), 0, st, a);
    hipLaunchKernelGGL(k_nl_permute, dim3(nb), dim3(256), 0, st, a);
    const dim3 ntiles(a.nslot / 64);
    if (a.build_nw == 8) {
        if (a.has_excl) hipLaunchKernelGGL((k_nl_tbuild<true, 8>), ntiles, dim3(512), 0, st, a);
        else hipLaunchKernelGGL((k_nl_tbuild<false, 8>), ntiles, dim3(512), 0, st, a);
    } else {
        if (a.has_excl) hipLaunchKernelGGL((k_nl_tbuild<true, 4>), ntiles, dim3(256), 0, st, a);
        else hipLaunchKernelGGL((k_nl_tbuild<false, 4>), ntiles, dim3(256), 0, st, a);
    }
    if (commit) hipLaunchKernelGGL(k_nl_commit, dim3((a.nrep + 3) / 4), dim3(256), 0, st, a);
    return hipGetLastError();
}

extern "C" hipError_t md_launch_force_trip(RepState* s, Gate* g, int nrep, hipStream_t st) {
    hipLaunchKernelGGL(k_force_trip, dim3(1), dim3(256), 0, st, s, g, nrep);
    return hipGetLastError();
}

extern "C" hipError_t md_launch_check_lists(const BuildArgs& a, int row_g, int* bad, hipStream_t st) {
    hipLaunchKernelGGL(k_check_lists, dim3((a.nslot + 255) / 256), dim3(256), 0, st, a, row_g, bad);
    return hipGetLastError();
}
