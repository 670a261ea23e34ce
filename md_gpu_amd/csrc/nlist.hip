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
//   k_nl_bin      cell of every atom (cells are ranked along a Hilbert curve),
//                 arrival rank in the cell by atomic count
//   k_nl_scan     exclusive scan of cell counts -> cell_start (one block per replica)
//   k_nl_scatter  atom -> provisional slot (cell order, arrival order)
//   k_nl_permute  deterministic order inside a cell (by eighth of the cell, then
//                 original index);
//                 gathers pos/vel/force/id/type into scratch in the new order
//   build         one workgroup per 64-slot tile; the image cells within r_list
//                 of a bounding box of atoms are listed in LDS and their members
//                 streamed and tested with the canonical FP32 predicate.  Three
//                 variants (build_kernel): k_nl_build (LPA lanes per atom, tile
//                 box, double-buffered block staging), k_nl_wbuild (8 lanes per
//                 atom, per-wavefront boxes, LDS hit buffers) and k_nl_jbuild
//                 (a lane per candidate, 16 or 8 atoms per wavefront, one ballot
//                 per atom and chunk).  Hits are appended in candidate order, so
//                 rows are deterministic.  Each tile stores its longest row.
//   k_nl_commit   per rebuilt replica: new list epoch (or the next step kernel opens it).
// Bytes per atom per rebuild (algorithmic): read pos/vel/force/id/type 56 B,
// write them twice (scratch + copy back) 112 B, plus the list 4*K B.
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
    const int l = max((int)floorf((bmin + halfL - rlp) * cinv), -2 * nc);
    const int h = min((int)floorf((bmax + halfL + rlp) * cinv), 3 * nc);
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

// The short rebuild kernels issue every load that does not depend on another
// (the gate counter included) in one round trip, and the loads that depend on
// the replica index in the next: an idle launch costs about the same, a working
// one saves round trips.  A wavefront's 64 slots are one tile, so the replica
// index is wavefront-uniform and its state and parameters are scalar loads.
__global__ __launch_bounds__(256) void k_nl_bin(BuildArgs a) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= a.nslot) return;
    const int n_stale = a.gate->n_stale;
    const int r = __builtin_amdgcn_readfirstlane(a.tile_rep[s >> 6]);
    const int id = a.ids[s];
    const float4 x0 = a.pos[0][s], x1 = a.pos[1][s];
    loads_issued();
    if (n_stale == 0) return;
    const RepState st = a.st[r];
    const RepParam& p = a.rep[r];
    const float h0 = p.halfL[0], h1 = p.halfL[1], h2 = p.halfL[2];
    const float c0 = p.cinv[0], c1 = p.cinv[1], c2 = p.cinv[2];
    const int n0 = p.nc[0], n1 = p.nc[1], n2 = p.nc[2], cell_off = p.cell_off, slot_off = p.slot_off;
    loads_issued();
    if (!stale(st)) return;
    if (s == slot_off) a.st[r].trip[(st.nb + 1) & 1] = 0;   // the buffer the next epoch will use
    if (id < 0) return;
    float4 x = cur_buf(st) ? x1 : x0;
    int cx = cell_coord(x.x, h0, c0, n0);
    int cy = cell_coord(x.y, h1, c1, n1);
    int cz = cell_coord(x.z, h2, c2, n2);
    int lin = cx + n0 * (cy + n1 * cz);
    int gc = cell_off + a.cell_rank[cell_off + lin];
    // the atom's eighth of the cell, numbered along a path through adjacent eighths
    const int sx = (x.x + h0) * c0 - (float)cx >= 0.5f;
    const int sy = (x.y + h1) * c1 - (float)cy >= 0.5f;
    const int sz = (x.z + h2) * c2 - (float)cz >= 0.5f;
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
    const int id = a.ids[s];
    const int gc = a.key[s];
    const int rk = a.rank[s];
    loads_issued();
    if (n_stale == 0) return;
    const RepState st = a.st[r];   // whole (no load that waits for nb)
    loads_issued();
    if (!stale(st)) return;
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
    const int o_raw = a.tmp[t];
    const int mykey = a.tmpid[t];
    loads_issued();
    if (n_stale == 0) return;
    // Second round trip: the replica state and parameters and the atom at slot o.
    // tmp holds a slot of this replica for every t below slot_off + natom; above
    // that, and for a replica that is not rebuilt, o is only clamped into range
    // and the loaded values are not used.
    const RepState st = a.st[r];
    const RepParam& p = a.rep[r];
    const int lim = p.slot_off + p.natom;
    const int o = min(max(o_raw, 0), a.nslot - 1);
    const int gc = a.key[o];
    const float4 x0 = a.pos[0][o], x1 = a.pos[1][o], v = a.vel[o], f4 = a.force[o];
    const int ty = a.typ[o];
    loads_issued();
    if (!stale(st)) return;
    if (t >= lim) return;
    const float4 x = cur_buf(st) ? x1 : x0;
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

// One workgroup per 64-slot tile, LPA lanes per atom (LPA * 64 threads).
template <bool EXCL, int LPA>
__global__ __launch_bounds__(64 * LPA) void k_nl_build(BuildArgs a) {
    constexpr int NT = 64 * LPA;          // threads = candidates staged per chunk
    constexpr int NW = NT / 64;           // waves
    constexpr int APW = 64 / LPA;         // atoms per wave
    constexpr int PER = kTab / NT;        // region cells probed per thread and batch
    __shared__ float4 s_pos[2][NT];       // (xj, yj, zj, bits of j | tj << 29)
    __shared__ float4 s_sh[2][NT];        // (-kx L, -ky L, -kz L, bits of jid): image shift of the candidate
    __shared__ int s_tstart[kTab];
    __shared__ int s_tpref[kTab + 1];
    __shared__ int s_tshift[kTab];
    __shared__ int s_w[NW];
    __shared__ int s_wmax[NW];
    __shared__ int s_n[2];               // staged candidates per buffer
    __shared__ unsigned short s_widx[NW][NT];   // per wave: staged candidates near the wave's box
    __shared__ float s_bb[NW][6];
    if (a.gate->n_stale == 0) return;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int tile = blockIdx.x;
    const int r = __builtin_amdgcn_readfirstlane(a.tile_rep[tile]);
    const RepState st = a.st[r];
    if (!stale(st)) return;
    const RepParam& p = a.rep[r];
    const int cb = cur_buf(st);
    const int ai = wave * APW + lane / LPA;     // atom of this lane within the tile
    const int sub = lane % LPA;
    const int slot = tile * 64 + ai;
    const int id = a.ids_s[slot];
    const bool active = id >= 0;
    const float4 xi = a.pos_s[slot];
    int4 ex = make_int4(-1, -1, -1, -1);
    if (EXCL && active) ex = a.excl[id];

    // tile bounding box over active atoms
    float lx = active ? xi.x : 3.0e38f, ly = active ? xi.y : 3.0e38f, lz = active ? xi.z : 3.0e38f;
    float hx = active ? xi.x : -3.0e38f, hy = active ? xi.y : -3.0e38f, hz = active ? xi.z : -3.0e38f;
    for (int o = 32; o >= LPA; o >>= 1) {
        lx = fminf(lx, __shfl_xor(lx, o, 64)); ly = fminf(ly, __shfl_xor(ly, o, 64)); lz = fminf(lz, __shfl_xor(lz, o, 64));
        hx = fmaxf(hx, __shfl_xor(hx, o, 64)); hy = fmaxf(hy, __shfl_xor(hy, o, 64)); hz = fmaxf(hz, __shfl_xor(hz, o, 64));
    }
    // the wave's own box (its APW atoms) prunes candidates once more before the pair tests
    const float wlx = lx, wly = ly, wlz = lz, whx = hx, why = hy, whz = hz;
    if (lane == 0) {
        s_bb[wave][0] = lx; s_bb[wave][1] = ly; s_bb[wave][2] = lz;
        s_bb[wave][3] = hx; s_bb[wave][4] = hy; s_bb[wave][5] = hz;
    }
    __syncthreads();
    for (int w = 0; w < NW; w++) {
        lx = fminf(lx, s_bb[w][0]); ly = fminf(ly, s_bb[w][1]); lz = fminf(lz, s_bb[w][2]);
        hx = fmaxf(hx, s_bb[w][3]); hy = fmaxf(hy, s_bb[w][4]); hz = fmaxf(hz, s_bb[w][5]);
    }
    int cnt = 0;
    uint32_t* rowp = a.nlist + (size_t)slot * a.cap;
    const int gs = __builtin_ctz(a.row_g);   // row layout: lanes per atom of the step kernel, a power of 2
    if (lx <= hx) {
        const float rlp = sqrtf(a.rl2) * 1.0001f + 1e-4f;
        const float rlp2 = rlp * rlp;
        // image-cell region covering the tile box grown by r_list.  Coordinates lie in
        // the box, so a span of more than five box lengths never occurs; the bounds
        // keep the walk finite for any input.
        const Axis ax = make_axis(lx, hx, rlp, p.halfL[0], p.cinv[0], p.cedge[0], p.nc[0]);
        const Axis ay = make_axis(ly, hy, rlp, p.halfL[1], p.cinv[1], p.cedge[1], p.nc[1]);
        const Axis az = make_axis(lz, hz, rlp, p.halfL[2], p.cinv[2], p.cedge[2], p.nc[2]);
        const int nreg = ax.nr * ay.nr * az.nr;
        const int cell_off = p.cell_off, cs_off = p.cs_off, ncx = p.nc[0], ncy = p.nc[1];
        // region cell c -> (cell start, member count, packed image shift); count 0 when pruned or empty
        auto probe = [=](int c, int& cs, int& shift) -> int {
            int kx, ky, kz, wx, wy, wz;
            const int cyz = floor_div(c, ax.nr, ax.inv_nr), cz = floor_div(cyz, ay.nr, ay.inv_nr);
            const float ddx = ax.gap(ax.lo + (c - cyz * ax.nr), kx, wx);
            const float ddy = ay.gap(ay.lo + (cyz - cz * ay.nr), ky, wy);
            const float ddz = az.gap(az.lo + cz, kz, wz);
            if (ddx * ddx + ddy * ddy + ddz * ddz >= rlp2) return 0;
            const int rank = a.cell_rank[cell_off + wx + ncx * (wy + ncy * wz)];
            cs = a.cell_start[cs_off + rank];
            shift = (kx + 8) | ((ky + 8) << 4) | ((kz + 8) << 8);
            return a.cell_start[cs_off + rank + 1] - cs;
        };
        for (int b0 = 0; b0 < nreg; b0 += kTab) {
            // ---- table of non-empty image cells within r_list of the box (region order);
            //      pass 1 counts, pass 2 writes (no per-thread arrays)
            const int c0 = b0 + tid * PER, c1 = min(min(c0 + PER, nreg), b0 + kTab);
            int n_keep = 0, csum = 0;
            for (int c = c0; c < c1; c++) {
                int cs, sh;
                const int n = probe(c, cs, sh);
                n_keep += n > 0;
                csum += n;
            }
            int ntab, T;
            const int off = block_scan<NT>(n_keep, s_w, ntab);
            const int coff = block_scan<NT>(csum, s_w, T);
            for (int c = c0, k = off, run = coff; c < c1; c++) {
                int cs, sh;
                const int n = probe(c, cs, sh);
                if (n > 0) {
                    s_tstart[k] = cs;
                    s_tpref[k] = run;
                    s_tshift[k] = sh;
                    k++;
                    run += n;
                }
            }
            if (tid == 0) s_tpref[ntab] = T;
            __syncthreads();
            // ---- stream the candidates through LDS.  A candidate whose image lies
            //      r_list or farther from the tile box cannot pass for any atom of the
            //      tile: it is dropped before staging (stable compaction, so rows keep
            //      candidate order).
            auto fetch = [&](int base, float4& rp, float4& rs) -> bool {
                const int q = base + tid;
                if (q >= T) return false;
                int lo_ = 0, hi_ = ntab - 1;   // last entry with pref <= q
                while (lo_ < hi_) {
                    const int mid = (lo_ + hi_ + 1) >> 1;
                    if (s_tpref[mid] <= q) lo_ = mid; else hi_ = mid - 1;
                }
                const int j = s_tstart[lo_] + (q - s_tpref[lo_]);
                const int sh = s_tshift[lo_];
                rp = a.pos_s[j];
                // -k L is exact for the |k| <= 1 images that can pass the predicate
                rs = make_float4(-(float)((sh & 15) - 8) * p.L[0], -(float)(((sh >> 4) & 15) - 8) * p.L[1],
                                 -(float)(((sh >> 8) & 15) - 8) * p.L[2], EXCL ? __int_as_float(a.ids_s[j]) : 0.0f);
                const float ix = rp.x - rs.x, iy = rp.y - rs.y, iz = rp.z - rs.z;   // image position
                const float gx = fmaxf(0.0f, fmaxf(lx - ix, ix - hx));
                const float gy = fmaxf(0.0f, fmaxf(ly - iy, iy - hy));
                const float gz = fmaxf(0.0f, fmaxf(lz - iz, iz - hz));
                rp.w = __int_as_float((int)((uint32_t)j | ((uint32_t)a.typ_s[j] << MD_JBITS)));
                return gx * gx + gy * gy + gz * gz < rlp2;
            };
            auto stage = [&](int buf, bool keep, const float4& rp, const float4& rs) {
                int kept;
                const int at = block_scan<NT>(keep ? 1 : 0, s_w, kept);
                if (keep) {
                    s_pos[buf][at] = rp;
                    s_sh[buf][at] = rs;
                }
                if (tid == 0) s_n[buf] = kept;
            };
            auto test = [&](int buf, int k, int wn, uint32_t& ent) -> bool {
                if (k >= wn) return false;
                const int q = s_widx[wave][k];
                const float4 pj = s_pos[buf][q];
                const float4 sj = s_sh[buf][q];
                // (xi - xj) - k L: the canonical minimum-image displacement for this image
                const float dx = (xi.x - pj.x) + sj.x;
                const float dy = (xi.y - pj.y) + sj.y;
                const float dz = (xi.z - pj.z) + sj.z;
                const float rsq = fmaf(dz, dz, fmaf(dy, dy, dx * dx));
                ent = (uint32_t)__float_as_int(pj.w);
                bool ok = active && rsq < a.rl2 && (int)(ent & MD_JMASK) != slot;
                if (EXCL) {
                    const int jid = __float_as_int(sj.w);
                    ok = ok && jid != ex.x && jid != ex.y && jid != ex.z && jid != ex.w;
                }
                return ok;
            };
            // append in candidate order: rank among this atom's LPA lanes
            auto append = [&](bool ok, uint32_t ent) {
                const unsigned long long m = __ballot(ok);
                const uint32_t gb = (uint32_t)(m >> (lane & ~(LPA - 1))) & (uint32_t)((1ull << LPA) - 1);
                const int before = __popc(gb & ((1u << sub) - 1u));
                if (ok && cnt + before < a.cap) rowp[lpos_sh(cnt + before, gs)] = ent;
                cnt += __popc(gb);
            };
            const int nch = (T + NT - 1) / NT;
            float4 rp = make_float4(0, 0, 0, 0), rs = make_float4(0, 0, 0, 0);
            if (nch > 0) {
                const bool keep = fetch(0, rp, rs);
                stage(0, keep, rp, rs);
            }
            __syncthreads();
            for (int ch = 0; ch < nch; ch++) {
                const int buf = ch & 1;
                const bool more = ch + 1 < nch;
                bool keep_next = false;
                if (more) keep_next = fetch((ch + 1) * NT, rp, rs);
                const int n = s_n[buf];
                // this wave's candidates: staged ones within r_list of the wave's box (in order)
                int wn = 0;
                for (int c0 = 0; c0 < n; c0 += 64) {
                    const int c = c0 + lane;
                    bool near = false;
                    if (c < n) {
                        const float4 pj = s_pos[buf][c];
                        const float4 sj = s_sh[buf][c];
                        const float ix = pj.x - sj.x, iy = pj.y - sj.y, iz = pj.z - sj.z;
                        const float gx = fmaxf(0.0f, fmaxf(wlx - ix, ix - whx));
                        const float gy = fmaxf(0.0f, fmaxf(wly - iy, iy - why));
                        const float gz = fmaxf(0.0f, fmaxf(wlz - iz, iz - whz));
                        near = gx * gx + gy * gy + gz * gz < rlp2;
                    }
                    const unsigned long long m = __ballot(near);
                    if (near) s_widx[wave][wn + __popcll(m & ((1ull << lane) - 1ull))] = (unsigned short)c;
                    wn += __popcll(m);
                }
                __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
                __builtin_amdgcn_wave_barrier();
                __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
                for (int u = 0; u < wn; u += 2 * LPA) {
                    uint32_t e0 = 0, e1 = 0;
                    const bool ok0 = test(buf, u + sub, wn, e0);
                    const bool ok1 = test(buf, u + LPA + sub, wn, e1);
                    append(ok0, e0);
                    append(ok1, e1);
                }
                __syncthreads();
                if (more) stage(buf ^ 1, keep_next, rp, rs);
                __syncthreads();
            }
        }
    }
    if (sub == 0) a.ncount[slot] = active ? cnt : 0;
    int m = sub == 0 ? cnt : 0;
    for (int o = 32; o >= 1; o >>= 1) m = max(m, __shfl_xor(m, o, 64));
    if (lane == 0) s_wmax[wave] = m;
    // copy the tile back from scratch into the primary arrays (new order)
    if (tid < 64) {
        const int s = tile * 64 + tid;
        const float4 x = a.pos_s[s];
        a.pos[cb][s] = x;
        a.pos_build[s] = x;
        a.vel[s] = a.vel_s[s];
        a.force[s] = a.force_s[s];
        a.ids[s] = a.ids_s[s];
        a.typ[s] = a.typ_s[s];
    }
    __syncthreads();
    if (tid == 0) {   // longest row of the tile; k_nl_commit opens the new list epoch
        int m = s_wmax[0];
        for (int w = 1; w < NW; w++) m = max(m, s_wmax[w]);
        a.tilemax[tile] = max(a.tilemax[tile], m);   // running maximum since the host cleared it
    }
}

// Per-wave list build: one wavefront per APW = 64/LPA consecutive slots.  Each
// wave lists the image cells within r_list of its own atoms' bounding box (a
// wave-private table in LDS), streams their members in chunks of 64 (one
// candidate per lane; the next chunk's loads are in flight while the current
// one is tested), drops candidates whose image is r_list or farther from the
// box, tests the rest with the canonical predicate and appends hits in
// candidate order.  Hits are gathered per atom in LDS and written to the row
// in bursts.  Waves never wait for each other until the tile's epilogue.
template <bool EXCL>
__global__ __launch_bounds__(512) void k_nl_wbuild(BuildArgs a) {
    constexpr int LPA = 8, APW = 64 / LPA, NW = 8;   // 8 waves x 8 atoms = one 64-slot tile
    constexpr int WT = 128;                         // table entries per wave and pass
    constexpr int HB = 64;                          // buffered hits per atom
    __shared__ int s_tstart[NW][WT];
    __shared__ int s_tpref[NW][WT + 1];
    __shared__ int s_tshift[NW][WT];
    __shared__ float4 s_cpos[NW][64];               // staged candidates (x, y, z, bits of j | tj << 29)
    __shared__ float4 s_csh[NW][64];                // (-kx L, -ky L, -kz L, bits of jid)
    __shared__ uint32_t s_hits[NW][APW][HB];
    __shared__ int s_wmax[NW];
    if (a.gate->n_stale == 0) return;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int tile = blockIdx.x;
    const int r = __builtin_amdgcn_readfirstlane(a.tile_rep[tile]);
    const RepState st = a.st[r];
    if (!stale(st)) return;
    const RepParam& p = a.rep[r];
    const int cb = cur_buf(st);
    const int grp = lane / LPA, sub = lane % LPA;
    const int slot = tile * 64 + wave * APW + grp;
    const int id = a.ids_s[slot];
    const bool active = id >= 0;
    const float4 xi = a.pos_s[slot];
    int4 ex = make_int4(-1, -1, -1, -1);
    if (EXCL && active) ex = a.excl[id];
    const unsigned long long below = (1ull << lane) - 1ull;
    auto wave_sync = [&]() {
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    };

    // the wave's box over its active atoms
    float lx = active ? xi.x : 3.0e38f, ly = active ? xi.y : 3.0e38f, lz = active ? xi.z : 3.0e38f;
    float hx = active ? xi.x : -3.0e38f, hy = active ? xi.y : -3.0e38f, hz = active ? xi.z : -3.0e38f;
    for (int o = 32; o >= LPA; o >>= 1) {
        lx = fminf(lx, __shfl_xor(lx, o, 64)); ly = fminf(ly, __shfl_xor(ly, o, 64)); lz = fminf(lz, __shfl_xor(lz, o, 64));
        hx = fmaxf(hx, __shfl_xor(hx, o, 64)); hy = fmaxf(hy, __shfl_xor(hy, o, 64)); hz = fmaxf(hz, __shfl_xor(hz, o, 64));
    }
    int cnt = 0;        // hits of this lane's atom (all lanes of the group agree)
    int nbuf = 0;       // buffered, not yet written
    uint32_t* rowp = a.nlist + (size_t)slot * a.cap;
    const int gs = __builtin_ctz(a.row_g);   // row layout: lanes per atom of the step kernel, a power of 2
    // write the buffered hits of every atom of the wave to the rows (lanes of a group
    // write consecutive logical entries)
    auto flush = [&]() {
        wave_sync();
        const int base = cnt - nbuf;
        for (int k = sub; k < nbuf; k += LPA) {
            const int e = base + k;
            if (e < a.cap) rowp[lpos_sh(e, gs)] = s_hits[wave][grp][k];
        }
        nbuf = 0;
        wave_sync();
    };
    if (lx <= hx) {
        const float rlp = sqrtf(a.rl2) * 1.0001f + 1e-4f;
        const float rlp2 = rlp * rlp;
        const Axis ax = make_axis(lx, hx, rlp, p.halfL[0], p.cinv[0], p.cedge[0], p.nc[0]);
        const Axis ay = make_axis(ly, hy, rlp, p.halfL[1], p.cinv[1], p.cedge[1], p.nc[1]);
        const Axis az = make_axis(lz, hz, rlp, p.halfL[2], p.cinv[2], p.cedge[2], p.nc[2]);
        const int nreg = ax.nr * ay.nr * az.nr;
        const int cell_off = p.cell_off, cs_off = p.cs_off, ncx = p.nc[0], ncy = p.nc[1];
        auto probe = [=](int c, int& cs, int& shift) -> int {
            int kx, ky, kz, wx, wy, wz;
            const int cyz = floor_div(c, ax.nr, ax.inv_nr), cz = floor_div(cyz, ay.nr, ay.inv_nr);
            const float ddx = ax.gap(ax.lo + (c - cyz * ax.nr), kx, wx);
            const float ddy = ay.gap(ay.lo + (cyz - cz * ay.nr), ky, wy);
            const float ddz = az.gap(az.lo + cz, kz, wz);
            if (ddx * ddx + ddy * ddy + ddz * ddz >= rlp2) return 0;
            const int rank = a.cell_rank[cell_off + wx + ncx * (wy + ncy * wz)];
            cs = a.cell_start[cs_off + rank];
            shift = (kx + 8) | ((ky + 8) << 4) | ((kz + 8) << 8);
            return a.cell_start[cs_off + rank + 1] - cs;
        };
        int c0 = 0;
        while (c0 < nreg) {
            // ---- wave table: up to WT non-empty cells, in region order
            int ntab = 0, T = 0;
            while (c0 < nreg && ntab <= WT - 64) {
                const int c = c0 + lane;
                int cs = 0, sh = 0, n = 0;
                if (c < nreg) n = probe(c, cs, sh);
                const unsigned long long m = __ballot(n > 0);
                int inc = n;   // inclusive scan of counts over the wave
                for (int o = 1; o < 64; o <<= 1) {
                    const int t = __shfl_up(inc, o, 64);
                    if (lane >= o) inc += t;
                }
                if (n > 0) {
                    const int k = ntab + __popcll(m & below);
                    s_tstart[wave][k] = cs;
                    s_tpref[wave][k] = T + inc - n;
                    s_tshift[wave][k] = sh;
                }
                ntab += __popcll(m);
                T += __shfl(inc, 63, 64);
                c0 += 64;
            }
            if (lane == 0) s_tpref[wave][ntab] = T;
            wave_sync();
            // ---- stream the table's candidates, 64 per chunk
            auto fetch = [&](int q, float4& rp, float4& rs) -> bool {
                if (q >= T) return false;
                int lo_ = 0, hi_ = ntab - 1;   // last entry with pref <= q
                while (lo_ < hi_) {
                    const int mid = (lo_ + hi_ + 1) >> 1;
                    if (s_tpref[wave][mid] <= q) lo_ = mid; else hi_ = mid - 1;
                }
                const int j = s_tstart[wave][lo_] + (q - s_tpref[wave][lo_]);
                const int sh = s_tshift[wave][lo_];
                rp = a.pos_s[j];
                // -k L is exact for the |k| <= 1 images that can pass the predicate
                rs = make_float4(-(float)((sh & 15) - 8) * p.L[0], -(float)(((sh >> 4) & 15) - 8) * p.L[1],
                                 -(float)(((sh >> 8) & 15) - 8) * p.L[2], EXCL ? __int_as_float(a.ids_s[j]) : 0.0f);
                const float ix = rp.x - rs.x, iy = rp.y - rs.y, iz = rp.z - rs.z;   // image position
                const float gx = fmaxf(0.0f, fmaxf(lx - ix, ix - hx));
                const float gy = fmaxf(0.0f, fmaxf(ly - iy, iy - hy));
                const float gz = fmaxf(0.0f, fmaxf(lz - iz, iz - hz));
                rp.w = __int_as_float((int)((uint32_t)j | ((uint32_t)a.typ_s[j] << MD_JBITS)));
                return gx * gx + gy * gy + gz * gz < rlp2;
            };
            float4 rp = make_float4(0, 0, 0, 0), rs = rp;
            bool keep = fetch(lane, rp, rs);
            for (int q0 = 0; q0 < T; q0 += 64) {
                // stage this chunk's survivors (stable compaction), then fetch the next chunk
                const unsigned long long m = __ballot(keep);
                if (keep) {
                    const int k = __popcll(m & below);
                    s_cpos[wave][k] = rp;
                    s_csh[wave][k] = rs;
                }
                const int wn = __popcll(m);
                keep = fetch(q0 + 64 + lane, rp, rs);
                wave_sync();
                for (int u = 0; u < wn; u += LPA) {
                    const int q = u + sub;
                    bool ok = false;
                    uint32_t ent = 0;
                    if (q < wn) {
                        const float4 pj = s_cpos[wave][q];
                        const float4 sj = s_csh[wave][q];
                        // (xi - xj) - k L: the canonical minimum-image displacement for this image
                        const float dx = (xi.x - pj.x) + sj.x;
                        const float dy = (xi.y - pj.y) + sj.y;
                        const float dz = (xi.z - pj.z) + sj.z;
                        const float rsq = fmaf(dz, dz, fmaf(dy, dy, dx * dx));
                        ent = (uint32_t)__float_as_int(pj.w);
                        ok = active && rsq < a.rl2 && (int)(ent & MD_JMASK) != slot;
                        if (EXCL) {
                            const int jid = __float_as_int(sj.w);
                            ok = ok && jid != ex.x && jid != ex.y && jid != ex.z && jid != ex.w;
                        }
                    }
                    // rank among this atom's lanes: hits stay in candidate order
                    const unsigned long long hm = __ballot(ok);
                    const uint32_t gb = (uint32_t)(hm >> (lane & ~(LPA - 1))) & 0xFFu;
                    if (ok) s_hits[wave][grp][nbuf + __popc(gb & ((1u << sub) - 1u))] = ent;
                    const int add = __popc(gb);
                    cnt += add;
                    nbuf += add;
                    // some atom's buffer may not take another LPA hits: write all out
                    const unsigned long long full = __ballot(nbuf > HB - LPA);
                    if (full) flush();
                }
                wave_sync();
            }
        }
    }
    if (__ballot(nbuf > 0)) flush();   // wave-uniform
    if (sub == 0) a.ncount[slot] = active ? cnt : 0;
    int mx = sub == 0 ? cnt : 0;
    for (int o = 32; o >= 1; o >>= 1) mx = max(mx, __shfl_xor(mx, o, 64));
    if (lane == 0) s_wmax[wave] = mx;
    __syncthreads();
    // copy the tile back from scratch into the primary arrays (new order)
    if (tid < 64) {
        const int s = tile * 64 + tid;
        const float4 x = a.pos_s[s];
        a.pos[cb][s] = x;
        a.pos_build[s] = x;
        a.vel[s] = a.vel_s[s];
        a.force[s] = a.force_s[s];
        a.ids[s] = a.ids_s[s];
        a.typ[s] = a.typ_s[s];
    }
    __syncthreads();
    if (tid == 0) {   // longest row of the tile; k_nl_commit opens the new list epoch
        int m = s_wmax[0];
        for (int w = 1; w < NW; w++) m = max(m, s_wmax[w]);
        a.tilemax[tile] = max(a.tilemax[tile], m);   // running maximum since the host cleared it
    }
}

// Candidate-parallel build: one workgroup (64 / AW wavefronts) per 64-slot tile,
// AW (8 or 16) atoms per wavefront.  The wavefront lists the image cells within
// r_list of its atoms' bounding box, streams their members, keeps those whose
// image lies within r_list of the box (stable compaction into a 128-entry LDS
// ring) and, per chunk of 64 kept candidates, lets each lane test one
// candidate against the AW atoms in turn.  The hits of one atom in a chunk are
// one ballot: they go to count + (hits in lower lanes), so rows are filled in
// candidate order with coalesced stores and no per-hit LDS traffic.
template <bool EXCL, int AW>
__global__ __launch_bounds__(64 * (64 / AW)) void k_nl_jbuild(BuildArgs a) {
    constexpr int NW = 64 / AW, WT = 256, RQ = 128;
    __shared__ float4 s_tcell[NW][WT];  // table row: (-kx L, -ky L, -kz L, bits of (cell start - pref))
    __shared__ int s_tpref[NW][WT + 1]; // first candidate index of each row; [ntab] = total
    __shared__ int s_mk[NW][64];        // row-boundary marks of a 64-candidate chunk (tagged)
    __shared__ float4 s_qp[NW][RQ];     // ring of kept candidates (x, y, z, bits of j | tj << 29)
    __shared__ float4 s_qs[NW][RQ];     // (-kx L, -ky L, -kz L, bits of jid)
    __shared__ int4 s_ex[NW][EXCL ? AW : 1];
    __shared__ int s_wmax[NW];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int tile = blockIdx.x;
    const int slot0 = tile * 64 + wave * AW;
    // the gate, the replica index and the wavefront's atoms in one round trip
    const int n_stale = a.gate->n_stale;
    const int r = __builtin_amdgcn_readfirstlane(a.tile_rep[tile]);
    const int id_pre = a.ids_s[slot0 + (lane & (AW - 1))];
    const float4 x_pre = a.pos_s[slot0 + (lane & (AW - 1))];
    loads_issued();
    if (n_stale == 0) return;
    const RepState st = a.st[r];
    loads_issued();
    if (!stale(st)) return;
    const RepParam& p = a.rep[r];
    const int cb = cur_buf(st);
    const unsigned long long below = (1ull << lane) - 1ull;
    s_mk[wave][lane] = 0;   // tags start at 1
    auto wave_sync = [&]() {
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    };
    bool active = false;
    float4 xi = make_float4(0.f, 0.f, 0.f, 0.f);
    if (lane < AW) {
        const int id = id_pre;
        active = id >= 0;
        xi = x_pre;
        if (EXCL) s_ex[wave][lane] = active ? a.excl[id] : make_int4(-1, -1, -1, -1);
    }
    // the wavefront's box over its active atoms
    float lx = active ? xi.x : 3.0e38f, ly = active ? xi.y : 3.0e38f, lz = active ? xi.z : 3.0e38f;
    float hx = active ? xi.x : -3.0e38f, hy = active ? xi.y : -3.0e38f, hz = active ? xi.z : -3.0e38f;
    for (int o = 32; o >= 1; o >>= 1) {
        lx = fminf(lx, __shfl_xor(lx, o, 64)); ly = fminf(ly, __shfl_xor(ly, o, 64)); lz = fminf(lz, __shfl_xor(lz, o, 64));
        hx = fmaxf(hx, __shfl_xor(hx, o, 64)); hy = fmaxf(hy, __shfl_xor(hy, o, 64)); hz = fmaxf(hz, __shfl_xor(hz, o, 64));
    }
    wave_sync();
    const int gs = __builtin_ctz(a.row_g);   // row layout: lanes per atom of the step kernel, a power of 2
    const int cap = a.cap;
    const float rl2 = a.rl2;
    uint32_t* const wrow = a.nlist + (size_t)slot0 * cap;   // row of the wavefront's first atom
    // the atoms as wavefront-uniform values (scalar registers); an empty slot sits
    // far outside the box, so no candidate passes the distance test for it
    float axr[AW], ayr[AW], azr[AW];
    int cnt[AW];
    {
        const float fx = active ? xi.x : 1.0e30f, fy = active ? xi.y : 1.0e30f, fz = active ? xi.z : 1.0e30f;
#pragma unroll
        for (int i = 0; i < AW; i++) {
            axr[i] = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(fx), i));
            ayr[i] = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(fy), i));
            azr[i] = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(fz), i));
            cnt[i] = 0;
        }
    }

    // test ring entries [head, head + n) against the wavefront's atoms
    auto process = [&](int head, int n) {
        const int q = (head + lane) & (RQ - 1);
        const bool valid = lane < n;
        const float4 pj = s_qp[wave][q];
        const float4 sj = s_qs[wave][q];
        const uint32_t ent = __float_as_uint(pj.w);
        const int js = (int)(ent & MD_JMASK);
        const int jid = __float_as_int(sj.w);
        const int jrel = js - slot0;   // the candidate's offset from the wavefront's first atom
#pragma unroll
        for (int i = 0; i < AW; i++) {
            // (xi - xj) - k L: the canonical minimum-image displacement for this image
            const float dx = (axr[i] - pj.x) + sj.x;
            const float dy = (ayr[i] - pj.y) + sj.y;
            const float dz = (azr[i] - pj.z) + sj.z;
            const float rsq = fmaf(dz, dz, fmaf(dy, dy, dx * dx));
            bool ok = valid & (rsq < rl2) & (jrel != i);
            if (EXCL) {
                const int4 ex = s_ex[wave][i];
                ok = ok & (jid != ex.x) & (jid != ex.y) & (jid != ex.z) & (jid != ex.w);
            }
            const unsigned long long m = __builtin_amdgcn_ballot_w64(ok);
            if (m) {
                const uint32_t e = __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32),
                                                             __builtin_amdgcn_mbcnt_lo((uint32_t)m, (uint32_t)cnt[i]));
                if (ok && e < (uint32_t)cap) wrow[(uint32_t)(i * cap) + lpos_u(e, gs)] = ent;
                cnt[i] += __popcll(m);
            }
        }
    };

    if (lx <= hx) {
        const float rlp = sqrtf(rl2) * 1.0001f + 1e-4f;
        const float rlp2 = rlp * rlp;
        const Axis ax = make_axis(lx, hx, rlp, p.halfL[0], p.cinv[0], p.cedge[0], p.nc[0]);
        const Axis ay = make_axis(ly, hy, rlp, p.halfL[1], p.cinv[1], p.cedge[1], p.nc[1]);
        const Axis az = make_axis(lz, hz, rlp, p.halfL[2], p.cinv[2], p.cedge[2], p.nc[2]);
        const int nreg = ax.nr * ay.nr * az.nr;
        const int cell_off = p.cell_off, cs_off = p.cs_off, ncx = p.nc[0], ncy = p.nc[1];
        // Cell probe without branches around the loads, so the loads of a group of
        // probes are in flight together: a cell beyond r_list of the box reads
        // entry 0 and reports no members.
        auto probe = [=](int c, int& cs, int& shift) -> int {
            int kx, ky, kz, wx, wy, wz;
            const int cyz = floor_div(c, ax.nr, ax.inv_nr), cz = floor_div(cyz, ay.nr, ay.inv_nr);
            const float ddx = ax.gap(ax.lo + (c - cyz * ax.nr), kx, wx);
            const float ddy = ay.gap(ay.lo + (cyz - cz * ay.nr), ky, wy);
            const float ddz = az.gap(az.lo + cz, kz, wz);
            const bool in = c < nreg && ddx * ddx + ddy * ddy + ddz * ddz < rlp2;
            const int rank = a.cell_rank[in ? cell_off + wx + ncx * (wy + ncy * wz) : cell_off];
            cs = a.cell_start[cs_off + rank];
            const int ce = a.cell_start[cs_off + rank + 1];
            shift = (kx + 8) | ((ky + 8) << 4) | ((kz + 8) << 8);
            return in ? ce - cs : 0;
        };
        // Row of candidate q0 + lane, given row k_lo holding q0: the rows starting
        // inside the chunk mark their start position, and a lane's row is k_lo plus
        // the marks at or below it.  Also advances k_lo to the row holding q0 + 64.
        int ktag = 0;
        auto locate = [&](int q0, int& k_lo, int ntab) -> int {
            ++ktag;
            const int kk = k_lo + 1 + lane;
            if (kk < ntab) {
                const int rel = s_tpref[wave][kk] - q0;   // >= 1: row k_lo holds q0
                if (rel < 64) s_mk[wave][rel] = ktag;
            }
            wave_sync();
            const unsigned long long M = __ballot(s_mk[wave][lane] == ktag);
            const unsigned long long upto = lane == 63 ? ~0ull : ((2ull << lane) - 1ull);
            const int k = k_lo + __popcll(M & upto);
            const int k63 = k_lo + __popcll(M);
            k_lo = (k63 + 1 < ntab && s_tpref[wave][k63 + 1] <= q0 + 64) ? k63 + 1 : k63;
            return k;
        };
        // candidate q0 + lane: its row, then the loads (consumed one chunk later, so
        // they overlap the chunk before)
        struct Cand { float4 p, s; int typ, id, j; bool valid; };
        auto fetch = [&](int q0, int& k_lo, int ntab, int T) -> Cand {
            Cand c;
            const int q = q0 + lane;
            c.valid = q < T;
            const int k = locate(q0, k_lo, ntab);
            c.s = s_tcell[wave][c.valid ? k : 0];
            c.j = c.valid ? __float_as_int(c.s.w) + q : slot0;
            c.p = a.pos_s[c.j];
            c.typ = a.typ_s[c.j];
            c.id = EXCL ? a.ids_s[c.j] : 0;
            return c;
        };
        int head = 0, tail = 0;   // ring positions (wavefront-uniform, taken mod RQ)
        int c0 = 0;
        while (c0 < nreg) {
            // ---- wavefront table: up to WT non-empty cells, in region order, probed
            //      in groups of four 64-cell batches
            int ntab = 0, T = 0;
            while (c0 < nreg && ntab <= WT - 256) {
                int cs[4], sh[4], n[4];
#pragma unroll
                for (int b = 0; b < 4; b++) n[b] = probe(c0 + 64 * b + lane, cs[b], sh[b]);
#pragma unroll
                for (int b = 0; b < 4; b++) {
                    const unsigned long long m = __ballot(n[b] > 0);
                    int inc = n[b];   // inclusive scan of counts over the wavefront
                    for (int o = 1; o < 64; o <<= 1) {
                        const int t = __shfl_up(inc, o, 64);
                        if (lane >= o) inc += t;
                    }
                    if (n[b] > 0) {
                        const int k = ntab + __popcll(m & below);
                        const int pref = T + inc - n[b];
                        // -k L is exact for the image offsets of the region
                        s_tcell[wave][k] = make_float4(-(float)((sh[b] & 15) - 8) * p.L[0],
                                                       -(float)(((sh[b] >> 4) & 15) - 8) * p.L[1],
                                                       -(float)(((sh[b] >> 8) & 15) - 8) * p.L[2],
                                                       __int_as_float(cs[b] - pref));
                        s_tpref[wave][k] = pref;
                    }
                    ntab += __popcll(m);
                    T += __shfl(inc, 63, 64);
                }
                c0 += 256;
            }
            if (lane == 0) s_tpref[wave][ntab] = T;
            wave_sync();
            if (T == 0) continue;
            // ---- stream the table's candidates, 64 at a time, through the ring
            int k_lo = 0;
            Cand nx = fetch(0, k_lo, ntab, T);
            for (int q0 = 0; q0 < T; q0 += 64) {
                const Cand cu = nx;
                if (q0 + 64 < T) nx = fetch(q0 + 64, k_lo, ntab, T);   // wavefront-uniform test
                bool keep = false;
                float4 rp = cu.p, rs = make_float4(0.f, 0.f, 0.f, 0.f);
                if (cu.valid) {
                    rs = make_float4(cu.s.x, cu.s.y, cu.s.z, __int_as_float(cu.id));
                    const float ix = rp.x - rs.x, iy = rp.y - rs.y, iz = rp.z - rs.z;   // image position
                    const float gx = fmaxf(0.0f, fmaxf(lx - ix, ix - hx));
                    const float gy = fmaxf(0.0f, fmaxf(ly - iy, iy - hy));
                    const float gz = fmaxf(0.0f, fmaxf(lz - iz, iz - hz));
                    rp.w = __int_as_float((int)((uint32_t)cu.j | ((uint32_t)cu.typ << MD_JBITS)));
                    keep = gx * gx + gy * gy + gz * gz < rlp2;
                }
                const unsigned long long m = __ballot(keep);
                if (keep) {
                    const int k = (tail + __popcll(m & below)) & (RQ - 1);
                    s_qp[wave][k] = rp;
                    s_qs[wave][k] = rs;
                }
                tail += __popcll(m);
                wave_sync();
                if (tail - head >= 64) {   // at most 127 entries are pending
                    process(head, 64);
                    head += 64;
                    wave_sync();
                }
            }
        }
        if (tail > head) process(head, tail - head);
    }
    int mine = 0, mx = 0;
#pragma unroll
    for (int i = 0; i < AW; i++) {
        if (lane == i) mine = cnt[i];
        mx = max(mx, cnt[i]);
    }
    if (lane < AW) a.ncount[slot0 + lane] = active ? mine : 0;
    if (lane == 0) s_wmax[wave] = mx;
    __syncthreads();
    // copy the tile back from scratch into the primary arrays (new order)
    if (tid < 64) {
        const int s = tile * 64 + tid;
        const float4 x = a.pos_s[s];
        a.pos[cb][s] = x;
        a.pos_build[s] = x;
        a.vel[s] = a.vel_s[s];
        a.force[s] = a.force_s[s];
        a.ids[s] = a.ids_s[s];
        a.typ[s] = a.typ_s[s];
    }
    if (tid == 0) {   // longest row of the tile; k_nl_commit opens the new list epoch
        int m = s_wmax[0];
        for (int w = 1; w < NW; w++) m = max(m, s_wmax[w]);
        a.tilemax[tile] = max(a.tilemax[tile], m);   // running maximum since the host cleared it
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

// ---------------------------------------------------------------- launchers

// commit = 0: a deferred-finalize block follows and its first step kernel opens the new lists
extern "C" hipError_t md_launch_rebuild(const BuildArgs& a, int commit, hipStream_t st) {
    const int nb = (a.nslot + 255) / 256;
    hipLaunchKernelGGL(k_nl_bin, dim3(nb), dim3(256), 0, st, a);
    hipLaunchKernelGGL(k_nl_scan, dim3(a.nrep), dim3(1024), 0, st, a);
    hipLaunchKernelGGL(k_nl_scatter, dim3(nb), dim3(256), 0, st, a);
    hipLaunchKernelGGL(k_nl_permute, dim3(nb), dim3(256), 0, st, a);
    const dim3 ntiles(a.nslot / 64);
    if (a.build_kernel == 2) {
        if (a.has_excl) hipLaunchKernelGGL((k_nl_jbuild<true, 16>), ntiles, dim3(256), 0, st, a);
        else hipLaunchKernelGGL((k_nl_jbuild<false, 16>), ntiles, dim3(256), 0, st, a);
    } else if (a.build_kernel == 3) {
        if (a.has_excl) hipLaunchKernelGGL((k_nl_jbuild<true, 8>), ntiles, dim3(512), 0, st, a);
        else hipLaunchKernelGGL((k_nl_jbuild<false, 8>), ntiles, dim3(512), 0, st, a);
    } else if (a.build_kernel == 1) {
        if (a.has_excl) hipLaunchKernelGGL((k_nl_wbuild<true>), ntiles, dim3(512), 0, st, a);
        else hipLaunchKernelGGL((k_nl_wbuild<false>), ntiles, dim3(512), 0, st, a);
    } else {
        switch ((a.has_excl ? 100 : 0) + a.build_lanes) {
        case 4:   hipLaunchKernelGGL((k_nl_build<false, 4>), ntiles, dim3(256), 0, st, a); break;
        case 8:   hipLaunchKernelGGL((k_nl_build<false, 8>), ntiles, dim3(512), 0, st, a); break;
        case 104: hipLaunchKernelGGL((k_nl_build<true, 4>), ntiles, dim3(256), 0, st, a); break;
        case 108: hipLaunchKernelGGL((k_nl_build<true, 8>), ntiles, dim3(512), 0, st, a); break;
        default:  return hipErrorInvalidValue;
        }
    }
    if (commit) hipLaunchKernelGGL(k_nl_commit, dim3((a.nrep + 3) / 4), dim3(256), 0, st, a);
    return hipGetLastError();
}

extern "C" hipError_t md_launch_force_trip(RepState* s, Gate* g, int nrep, hipStream_t st) {
    hipLaunchKernelGGL(k_force_trip, dim3(1), dim3(256), 0, st, s, g, nrep);
    return hipGetLastError();
}
