// nlist.hip -- binned cell-list neighbor build with space-filling-curve
// ordering of the atoms (the reference's intended neighborlist_binned,
// src/neighborlist_binned.cu:20-119 / src/neighborlist_kernel_o.cu:20-261,
// which never compiled; the list semantics are those of the O(N^2) build it
// replaces, src/neighborlist.cu:231-297).
//
// A rebuild is five kernels, each gated per replica by the rebuild flag that
// the previous position update raised (skin criterion, src/neighborlist.cu:157-176):
//   k_nl_bin      cell of every atom (cells are ranked along a Hilbert curve),
//                 arrival rank in the cell by atomic count
//   k_nl_scan     exclusive scan of cell counts -> cell_start (one block per replica)
//   k_nl_scatter  atom -> provisional slot (cell order, arrival order)
//   k_nl_permute  deterministic order inside a cell (by original index);
//                 gathers pos/vel/force/id/type into scratch in the new order
//   k_nl_build    one wavefront per 64-slot tile: tile bounding box, walk the
//                 image cells within r_list of it, test every member with the
//                 canonical FP32 predicate (wave-uniform j -> scalar loads),
//                 append (j | type_j << 29); copy the tile's scratch back.
// Bytes per atom per rebuild (algorithmic): read pos/vel/force/id/type 56 B,
// write them twice (scratch + copy back) 112 B, plus the list 4*K B.
#include "md_device.h"

namespace {

__device__ __forceinline__ int cell_coord(float x, float halfL, float cinv, int nc) {
    int k = (int)floorf((x + halfL) * cinv);
    return min(max(k, 0), nc - 1);
}

__device__ __forceinline__ int floordiv(int a, int b) {  // b > 0
    int q = a / b;
    return (a % b != 0 && ((a < 0) != (b < 0))) ? q - 1 : q;
}

}  // namespace

__global__ __launch_bounds__(256) void k_nl_bin(BuildArgs a) {
    int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= a.nslot) return;
    int r = a.tile_rep[s >> 6];
    if (!a.flag[r]) return;
    int id = a.ids[s];
    if (id < 0) return;
    const RepParam& p = a.rep[r];
    float4 x = a.pos[s];
    int cx = cell_coord(x.x, p.halfL[0], p.cinv[0], p.nc[0]);
    int cy = cell_coord(x.y, p.halfL[1], p.cinv[1], p.nc[1]);
    int cz = cell_coord(x.z, p.halfL[2], p.cinv[2], p.nc[2]);
    int lin = cx + p.nc[0] * (cy + p.nc[1] * cz);
    int gc = p.cell_off + a.cell_rank[p.cell_off + lin];
    a.key[s] = gc;
    a.rank[s] = atomicAdd(&a.cell_count[gc], 1);
}

__global__ __launch_bounds__(1024) void k_nl_scan(BuildArgs a) {
    __shared__ int wsum[16];
    const int r = blockIdx.x;
    if (!a.flag[r]) return;
    const RepParam& p = a.rep[r];
    const int nc = p.nc[0] * p.nc[1] * p.nc[2];
    const int per = (nc + 1023) / 1024;
    const int b = min(nc, (int)threadIdx.x * per), e = min(nc, b + per);
    int sum = 0;
    for (int k = b; k < e; k++) sum += a.cell_count[p.cell_off + k];
    // block exclusive scan of per-thread sums
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
    int run = p.slot_off + wsum[w] + inc - sum;
    for (int k = b; k < e; k++) {
        int c = a.cell_count[p.cell_off + k];
        a.cell_start[p.cs_off + k] = run;
        run += c;
        a.cell_count[p.cell_off + k] = 0;   // ready for the next rebuild
    }
    if (threadIdx.x == 0) {
        a.cell_start[p.cs_off + nc] = p.slot_off + p.natom;
        atomicAdd(a.rebuild_count, 1ULL);
    }
}

__global__ __launch_bounds__(256) void k_nl_scatter(BuildArgs a, int* tmpid) {
    int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= a.nslot) return;
    int r = a.tile_rep[s >> 6];
    if (!a.flag[r]) return;
    int id = a.ids[s];
    if (id < 0) return;
    int gc = a.key[s];
    int t = a.cell_start[gc + r] + a.rank[s];
    a.tmp[t] = s;
    tmpid[t] = id;
}

__global__ __launch_bounds__(256) void k_nl_permute(BuildArgs a, const int* tmpid) {
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= a.nslot) return;
    int r = a.tile_rep[t >> 6];
    if (!a.flag[r]) return;
    const RepParam& p = a.rep[r];
    if (t >= p.slot_off + p.natom) return;
    int o = a.tmp[t];
    int myid = tmpid[t];
    int gc = a.key[o];
    int cs = a.cell_start[gc + r], ce = a.cell_start[gc + r + 1];
    int cnt = 0;
    for (int u = cs; u < ce; u++) cnt += (tmpid[u] < myid);
    int f = cs + cnt;
    a.pos_s[f] = a.pos[o];
    a.vel_s[f] = a.vel[o];
    a.force_s[f] = a.force[o];
    a.ids_s[f] = myid;
    a.typ_s[f] = a.typ[o];
    a.slot_of[myid] = f;
}

template <bool EXCL>
__global__ __launch_bounds__(256) void k_nl_build(BuildArgs a) {
    const int lane = threadIdx.x & 63;
    const int tile = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (tile * 64 >= a.nslot) return;
    const int r = __builtin_amdgcn_readfirstlane(a.tile_rep[tile]);
    if (!a.flag[r]) return;
    const RepParam& p = a.rep[r];
    const int slot = tile * 64 + lane;
    const int id = a.ids_s[slot];
    const bool active = id >= 0;
    const float4 xi = a.pos_s[slot];

    // tile bounding box over active lanes
    float bmin[3], bmax[3];
    float xv[3] = {xi.x, xi.y, xi.z};
    for (int d = 0; d < 3; d++) {
        bmin[d] = active ? xv[d] : 3.0e38f;
        bmax[d] = active ? xv[d] : -3.0e38f;
    }
    for (int o = 32; o >= 1; o >>= 1)
        for (int d = 0; d < 3; d++) {
            bmin[d] = fminf(bmin[d], __shfl_xor(bmin[d], o, 64));
            bmax[d] = fmaxf(bmax[d], __shfl_xor(bmax[d], o, 64));
        }
    int count = 0;
    const bool any = bmin[0] <= bmax[0];
    int4 ex = make_int4(-1, -1, -1, -1);
    if (EXCL && active) ex = a.excl[id];
    if (any) {
        const float rlp = sqrtf(a.rl2) * 1.0001f + 1e-4f;
        const float rlp2 = rlp * rlp;
        int lo[3], hi[3];
        for (int d = 0; d < 3; d++) {
            lo[d] = (int)floorf((bmin[d] + p.halfL[d] - rlp) * p.cinv[d]);
            hi[d] = (int)floorf((bmax[d] + p.halfL[d] + rlp) * p.cinv[d]);
            // coordinates are wrapped into the box, so the span never needs more than
            // two box lengths of cells; the bound keeps the walk finite for any input
            lo[d] = max(lo[d], -2 * p.nc[d]);
            hi[d] = min(hi[d], 3 * p.nc[d]);
        }
        const size_t row = (size_t)slot * a.cap;
        for (int cz = lo[2]; cz <= hi[2]; cz++) {
            const int kz = floordiv(cz, p.nc[2]), wz = cz - kz * p.nc[2];
            const float zl = cz * p.cedge[2] - p.halfL[2], zh = zl + p.cedge[2];
            const float dz0 = fmaxf(0.0f, fmaxf(zl - bmax[2], bmin[2] - zh));
            if (dz0 * dz0 >= rlp2) continue;
            for (int cy = lo[1]; cy <= hi[1]; cy++) {
                const int ky = floordiv(cy, p.nc[1]), wy = cy - ky * p.nc[1];
                const float yl = cy * p.cedge[1] - p.halfL[1], yh = yl + p.cedge[1];
                const float dy0 = fmaxf(0.0f, fmaxf(yl - bmax[1], bmin[1] - yh));
                if (dz0 * dz0 + dy0 * dy0 >= rlp2) continue;
                for (int cx = lo[0]; cx <= hi[0]; cx++) {
                    const int kx = floordiv(cx, p.nc[0]), wx = cx - kx * p.nc[0];
                    const float xl = cx * p.cedge[0] - p.halfL[0], xh = xl + p.cedge[0];
                    const float dx0 = fmaxf(0.0f, fmaxf(xl - bmax[0], bmin[0] - xh));
                    if (dz0 * dz0 + dy0 * dy0 + dx0 * dx0 >= rlp2) continue;
                    const int lin = wx + p.nc[0] * (wy + p.nc[1] * wz);
                    const int rank = a.cell_rank[p.cell_off + lin];
                    const int cs = __builtin_amdgcn_readfirstlane(a.cell_start[p.cs_off + rank]);
                    const int ce = __builtin_amdgcn_readfirstlane(a.cell_start[p.cs_off + rank + 1]);
                    // image shift of this cell copy: displacement = (xi - xj) - k L (canonical form)
                    const float skx = (float)kx, sky = (float)ky, skz = (float)kz;
                    for (int j = cs; j < ce; j++) {
                        const float4 pj = a.pos_s[j];
                        const int jid = a.ids_s[j];
                        const int tj = a.typ_s[j];
                        const float dx = fmaf(-p.L[0], skx, xi.x - pj.x);
                        const float dy = fmaf(-p.L[1], sky, xi.y - pj.y);
                        const float dz = fmaf(-p.L[2], skz, xi.z - pj.z);
                        const float rsq = fmaf(dz, dz, fmaf(dy, dy, dx * dx));
                        bool ok = active && rsq < a.rl2 && j != slot;
                        if (EXCL) ok = ok && jid != ex.x && jid != ex.y && jid != ex.z && jid != ex.w;
                        if (ok) {
                            if (count < a.cap) a.nlist[row + count] = (uint32_t)j | ((uint32_t)tj << MD_JBITS);
                            count++;
                        }
                    }
                }
            }
        }
    }
    a.ncount[slot] = active ? count : 0;
    int m = count;
    for (int o = 32; o >= 1; o >>= 1) m = max(m, __shfl_xor(m, o, 64));
    if (lane == 0) atomicMax(&a.maxcount[r], m);
    // copy the tile back from scratch into the primary arrays (new order)
    a.pos_dst[slot] = xi;
    a.pos_build[slot] = xi;
    a.vel_dst[slot] = a.vel_s[slot];
    a.force_dst[slot] = a.force_s[slot];
    a.ids_dst[slot] = id;
    a.typ_dst[slot] = a.typ_s[slot];
}

// ---------------------------------------------------------------- launchers

extern "C" hipError_t md_launch_rebuild(const BuildArgs& a, int* tmpid, hipStream_t st) {
    const int nb = (a.nslot + 255) / 256;
    hipLaunchKernelGGL(k_nl_bin, dim3(nb), dim3(256), 0, st, a);
    hipLaunchKernelGGL(k_nl_scan, dim3(a.nrep), dim3(1024), 0, st, a);
    hipLaunchKernelGGL(k_nl_scatter, dim3(nb), dim3(256), 0, st, a, tmpid);
    hipLaunchKernelGGL(k_nl_permute, dim3(nb), dim3(256), 0, st, a, (const int*)tmpid);
    const int ntiles = a.nslot / 64;
    if (a.has_excl)
        hipLaunchKernelGGL(k_nl_build<true>, dim3((ntiles + 3) / 4), dim3(256), 0, st, a);
    else
        hipLaunchKernelGGL(k_nl_build<false>, dim3((ntiles + 3) / 4), dim3(256), 0, st, a);
    return hipGetLastError();
}
