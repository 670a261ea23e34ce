// persist.hip -- small single systems: every step of a run in one launch.
//
// For one replica of at most MD_PK_SLOTS slots (Lennard-Jones only, no bonded
// terms) a step is microseconds of pair work spread over a handful of tiles,
// and a launch per step (plus its finalize) costs more than the step.  Here a
// few workgroups of 1024 threads, one per CU and all resident at once, run
// every step of a run and meet at a grid barrier once per step:
//   - workgroup w owns the 1024/G slots from w 1024/G (G lanes per atom) and
//     keeps their position, velocity, force and build position in registers;
//   - after each barrier it stages the positions of the whole system into LDS
//     (at most 64 KB) and walks its atoms' rows there, minimum image per pair
//     (src/neighborlist.cu:271-273), LJ force as src/pair_forces.cu:190-195;
//   - the rows are private to the lanes: lane l of an atom lists the
//     neighbors j = l (mod G) within r_list, 16-bit slot indices in LDS.  They
//     are built by a scan of the staged system at the run's first step and
//     whenever the previous step's skin check tripped anywhere (each workgroup
//     reads every workgroup's flag at the barrier, so all rebuild together);
//     the build wraps the atoms into the box, as the tile path's build does;
//   - at the barrier each workgroup also publishes its FP64 partial sums (2 KE,
//     LJ energy, virial), so the Berendsen factor (src/integrator.cu:146-147,
//     from the previous step's temperature) and the statistics record of each
//     step come from the same sums, added in the same order, everywhere.
// Integrators: leapfrog (src/nve_leapfrog.cu:37-55, src/nvt_leapfrog_ber.cu:35-55)
// and NVE velocity Verlet in the leapfrog form of the step kernel (each step
// ends with the second half kick and begins the next with the first).
//
// The barrier: plain stores of the published data, then an agent-scope release
// fence and a relaxed store of the workgroup's flag; one wavefront polls every
// workgroup's flag with relaxed loads, then an agent-scope acquire fence.  Flags
// alternate between two banks by barrier parity: a workgroup one barrier ahead
// never overwrites a flag another is still reading.  Every wait is bounded
// (MD_PK_TIMEOUT ticks of the 100 MHz wall clock): a workgroup that times out
// sets MD_PK_ERR_WAIT and leaves, the others leave at their next wait, and the
// host reports the error.
#include <hip/hip_ext.h>

#include <type_traits>

#include "md_device.h"
#include "md_reduce.h"

namespace {

constexpr int NTH = MD_PK_THREADS;
constexpr unsigned FL_TRIP = 1u, FL_ROWS = 2u, FL_STOP = 4u;   // flag bits below the barrier count

// Grid barrier number b with the workgroup's flag bits; returns the OR of every
// workgroup's bits (FL_STOP also when this workgroup's wait timed out).  Called
// by every thread of the workgroup.
__device__ unsigned pk_barrier(const PersistArgs& a, int b, unsigned bits, unsigned* s_bits) {
    const int tid = threadIdx.x, lane = tid & 63;
    unsigned* fl = a.flags + (b & 1) * MD_PK_WG;
    __syncthreads();   // the workgroup's published stores precede thread 0's release fence
    if (tid == 0) {
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
        __hip_atomic_store(fl + blockIdx.x, ((unsigned)(b + 1) << 3) | bits, __ATOMIC_RELAXED,
                           __HIP_MEMORY_SCOPE_AGENT);
    }
    if (tid < 64) {
        const unsigned want = (unsigned)(b + 1);
        const unsigned long long t0 = wall_clock64();
        unsigned mine = 0;
        bool late = false;
        for (;;) {
            bool ok = true;
            if (lane < a.nwg) {
                const unsigned v = __hip_atomic_load(fl + lane, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                ok = (v >> 3) >= want;
                mine = v & 7u;
            }
            if (__all(ok)) break;
            if (wall_clock64() - t0 > MD_PK_TIMEOUT) {
                late = true;
                break;
            }
            __builtin_amdgcn_s_sleep(1);
        }
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
        unsigned all = 0;
        if (__ballot((mine & FL_TRIP) != 0)) all |= FL_TRIP;
        if (__ballot((mine & FL_ROWS) != 0)) all |= FL_ROWS;
        if (__ballot((mine & FL_STOP) != 0)) all |= FL_STOP;
        if (late) {
            all |= FL_STOP;
            if (lane == 0) atomicOr(a.err, MD_PK_ERR_WAIT);
        }
        if (lane == 0) *s_bits = all;
    }
    __syncthreads();
    return *s_bits;
}

}  // namespace

// G lanes per atom (1024 / G atoms per workgroup); NT1: one atom type (the LJ
// parameters sit in registers).
template <int G, bool NT1>
__global__ __launch_bounds__(MD_PK_THREADS) void k_persist(const PersistArgs a) {
    constexpr int APW = NTH / G, NW = NTH / 64;
    extern __shared__ float4 s_dyn[];
    float4* s_pos = s_dyn;                                                 // [nslot] x, y, z, type (-1: padding)
    float* s_tyf = reinterpret_cast<float*>(s_dyn + a.nslot);              // [nslot] type as a float
    // lane rows: entry p of thread t in the 16-bit half p & 1 of word (p >> 1) NTH + t
    uint16_t* s_rows = reinterpret_cast<uint16_t*>(s_tyf + a.nslot);       // [cap_l / 2][NTH] words
    __shared__ float4 s_lj[64];
    __shared__ double s_red[NW][3];
    __shared__ double s_lam;         // Berendsen factor of the current step
    __shared__ unsigned s_bits;      // flag bits of all workgroups at the last barrier
    __shared__ unsigned s_wgb;       // this workgroup's flag bits of the current step
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6, sub = tid & (G - 1);
    const int slot = blockIdx.x * APW + tid / G;
    const bool inr = slot < a.nslot;
    const int id = inr ? a.ids[slot] : -1;
    const bool active = id >= 0;
    const int ty = active ? a.typ[slot] : 0;
    const RepParam& p = a.rep[0];
    const float L0 = p.L[0], L1 = p.L[1], L2 = p.L[2];
    const float Li0 = p.Linv[0], Li1 = p.Linv[1], Li2 = p.Linv[2];
    const RepState s0 = a.st[0];
    const long long step0 = s0.step;
    const int cb0 = (int)(step0 & 1);   // between runs phase = 0: the positions of step0
    float4 x = make_float4(0.f, 0.f, 0.f, 0.f), v = x, f = x;
    if (inr) {
        x = a.pos[cb0][slot];
        v = a.vel[slot];
        f = a.force[slot];
    }
    for (int k = tid; k < a.nslot; k += NTH) s_tyf[k] = a.ids[k] >= 0 ? (float)a.typ[k] : -1.0f;
    if (tid < 64) s_lj[tid] = a.ljtab[tid];
    const double dtd = a.dt, taud = a.tau;
    double lamd = a.leapfrog ? a.th[0].S : 1.0;   // leapfrog: the factor of the run's first step
    const bool need_lam = a.leapfrog && a.thermostat == 1;
    const bool need_ke = need_lam || a.stats != nullptr;   // the published sums have a reader
    int b = 0;                                      // barriers passed
    unsigned bits = 0;
    if (!a.leapfrog) {
        // velocity Verlet: the first half kick and drift of the run's first step
        // (k_vv_kick's arithmetic with S = 1)
        if (active) {
            const float h = 0.5f * a.dt * v.w;
            v.x = fmaf(f.x, h, v.x);
            v.y = fmaf(f.y, h, v.y);
            v.z = fmaf(f.z, h, v.z);
            x = make_float4(fmaf(v.x, a.dt, x.x), fmaf(v.y, a.dt, x.y), fmaf(v.z, a.dt, x.z), x.w);
            if (sub == 0) a.pos[cb0 ^ 1][slot] = x;
        }
        bits = pk_barrier(a, b++, 0u, &s_bits);
    }
    float4 xb = x;                  // build position (skin check)
    int cnt = 0;                    // this lane's row length
    int nbuild = 0;
    float4 lj1 = make_float4(0.f, 0.f, 0.f, 0.f);
    __syncthreads();                // s_lj, s_tyf
    if (NT1) lj1 = s_lj[0];
    const int nsteps = (int)a.nsteps;
    int done = 0;
    for (int k = 0; k < nsteps && !(bits & FL_STOP); k++) {
        // positions of the step whose forces this iteration computes
        const long long sx = a.leapfrog ? step0 + k : step0 + k + 1;
        const float4* __restrict__ pin = a.pos[(int)(sx & 1)];
        if (tid == 0) s_wgb = 0;   // every thread read the last value before the previous barrier
        for (int q = tid; q < a.nslot; q += NTH) {
            const float4 y = pin[q];
            s_pos[q] = make_float4(y.x, y.y, y.z, s_tyf[q]);
        }
        // The previous step's sums, wanted by the Berendsen factor (every workgroup)
        // and the statistics record (workgroup 0): loaded now by wavefront 0, added
        // after the pair loop (in the same order everywhere).  The bank stays intact
        // until this workgroup has passed its next barrier.
        const bool want_sums = k > 0 && wave == 0 && (need_lam || (blockIdx.x == 0 && a.stats));
        double t0 = 0.0, t1 = 0.0, t2 = 0.0;
        if (want_sums && lane < a.nwg) {
            const double* pb = a.pub + (size_t)((b - 1) & 1) * MD_PK_WG * 4;
            t0 = pb[lane * 4 + 0];
            t1 = pb[lane * 4 + 1];
            t2 = pb[lane * 4 + 2];
        }
        __syncthreads();
        // rows: at the first step and after any skin trip.  The atom's G lanes test
        // the slots j = c G + lane in chunks c and compact the hits in slot order,
        // entry e of the row going to lane e mod G: the lanes' rows differ in length
        // by at most one.
        unsigned mybits = 0;
        if (k == 0 || (bits & FL_TRIP)) {
            if (active) {
                x.x = wrap_coord(x.x, L0, Li0);
                x.y = wrap_coord(x.y, L1, Li1);
                x.z = wrap_coord(x.z, L2, Li2);
            }
            xb = x;
            const int g0 = lane & ~(G - 1);   // the group's first lane in the wavefront
            int n = 0;                        // the atom's entries so far (the same in its G lanes)
            for (int j0 = 0; j0 < a.nslot; j0 += G) {
                const int j = j0 + sub;
                bool hit = false;
                if (active && j < a.nslot) {
                    const float4 pj = s_pos[j];
                    const float dx = minimg(x.x - pj.x, L0, Li0);
                    const float dy = minimg(x.y - pj.y, L1, Li1);
                    const float dz = minimg(x.z - pj.z, L2, Li2);
                    const float rsq = fmaf(dz, dz, fmaf(dy, dy, dx * dx));
                    hit = rsq < a.rl2 && j != slot && pj.w >= 0.0f;
                }
                const unsigned gm = (unsigned)(__ballot(hit) >> g0) & ((1u << G) - 1u);
                if (hit) {
                    const int e = n + __builtin_popcount(gm & ((1u << sub) - 1u));
                    const int q = e / G;
                    if (q < a.cap_l) s_rows[((q >> 1) * NTH + (tid - sub) + (e % G)) * 2 + (q & 1)] = (uint16_t)j;
                }
                n += __builtin_popcount(gm);
            }
            cnt = n > sub ? (n - sub + G - 1) / G : 0;
            // an odd row ends in half a word: its other half names the atom itself (a
            // valid slot whose pair the walk masks)
            if ((cnt & 1) && cnt < a.cap_l) s_rows[((cnt >> 1) * NTH + tid) * 2 + 1] = (uint16_t)slot;
            if ((n + G - 1) / G > a.cap_l) mybits |= FL_ROWS;
            nbuild++;
        }
        const bool en = a.energy_every > 0 && ((step0 + k) % a.energy_every) == 0;
        // pair forces of the own atom (this lane's share of its row), two entries per
        // LDS word so that both neighbors' loads are in flight together
        float fx = 0.f, fy = 0.f, fz = 0.f;
        double elj = 0.0, evi = 0.0;
        const int nr = min(cnt, a.cap_l);
        const float4* s_lji = s_lj + ty * 8;
        const uint32_t* s_w = reinterpret_cast<const uint32_t*>(s_rows) + tid;
        auto walk = [&](auto en_tag) {
            constexpr bool EN = decltype(en_tag)::value;
            auto pair = [&](const float4& pj, bool ok) {
                const float dx = minimg(x.x - pj.x, L0, Li0);
                const float dy = minimg(x.y - pj.y, L1, Li1);
                const float dz = minimg(x.z - pj.z, L2, Li2);
                const float rsq = fmaf(dz, dz, fmaf(dy, dy, dx * dx));
                const float4 lj = NT1 ? lj1 : s_lji[(int)pj.w];
                float ff = pair_ff_lj(rsq, lj);
                const bool use = ok && !(rsq > a.rc2);   // src/pair_forces.cu:195
                ff = use ? ff : 0.0f;
                fx = fmaf(ff, dx, fx);
                fy = fmaf(ff, dy, fy);
                fz = fmaf(ff, dz, fz);
                if (EN && use) {
                    const float ri2 = __builtin_amdgcn_rcpf(rsq);
                    const float ri6 = ri2 * ri2 * ri2;
                    elj += (double)(fmaf(lj.x * (1.0f / 12.0f) * ri6, ri6, -lj.y * (1.0f / 6.0f) * ri6) - lj.z);
                    evi += (double)(ff * rsq);
                }
            };
            for (int e = 0; e < nr; e += 2) {
                const uint32_t w = s_w[(e >> 1) * NTH];
                const float4 p0 = s_pos[w & 0xffffu];
                const float4 p1 = s_pos[w >> 16];
                pair(p0, true);
                pair(p1, e + 1 < nr);
            }
        };
        if (en) walk(std::true_type{});
        else walk(std::false_type{});
        fx = group_sum<G>(fx);
        fy = group_sum<G>(fy);
        fz = group_sum<G>(fz);
        if (want_sums) {   // the previous step's totals: Berendsen factor and record
            t0 = wave_sum_d(t0);
            t1 = wave_sum_d(t1);
            t2 = wave_sum_d(t2);
            const double T = t0 / ((double)p.nf * MD_KB_REF);
            if (need_lam && lane == 0) s_lam = sqrt(1.0 + dtd / taud * ((double)p.T0 / T - 1.0));
            const long long ks = k - 1;
            if (blockIdx.x == 0 && lane == 0 && a.stats && ks < a.stat_cap) {
                const bool en_p = a.energy_every > 0 && ((step0 + ks) % a.energy_every) == 0;
                StatRec rec;
                rec.ke = 0.5 * t0;
                rec.temp = T;
                rec.xi = a.th[0].xi;
                rec.vxi = a.th[0].vxi;
                rec.volume = (double)L0 * (double)L1 * (double)L2;
                rec.has_energy = en_p ? 1 : 0;
                rec.pad = 0;
                rec.pe_lj = en_p ? 0.5 * t1 : 0.0;
                rec.pe_coul = 0.0;
                rec.pe_bond = 0.0;
                rec.pe_angle = 0.0;
                rec.virial = en_p ? 0.5 * t2 : 0.0;
                rec.conserved = rec.ke + rec.pe_lj;
                a.stats[ks] = rec;
            }
        }
        if (need_lam && k > 0) {
            __syncthreads();   // s_lam
            lamd = s_lam;
        }
        // integrate (every lane of the group holds the same bits; lane 0 stores)
        double dke = 0.0;
        const bool last = k + 1 == nsteps;
        if (active) {
            f = make_float4(fx, fy, fz, 0.0f);
            const float lam = (float)lamd;
            bool drift = true;
            if (a.leapfrog) {
                const float massif = a.dt * v.w;   // dt / m (src/nve_leapfrog.cu:38)
                v.x = fmaf(fx, massif, v.x) * lam;
                v.y = fmaf(fy, massif, v.y) * lam;
                v.z = fmaf(fz, massif, v.z) * lam;
                dke = ((double)v.x * v.x + (double)v.y * v.y + (double)v.z * v.z) / (double)v.w;
            } else {
                // v(t) = v(t - dt/2) + F dt/2m ends the step; v(t + dt/2) = v(t) + F dt/2m begins the next
                const float h = 0.5f * a.dt * v.w;
                v.x = fmaf(fx, h, v.x);
                v.y = fmaf(fy, h, v.y);
                v.z = fmaf(fz, h, v.z);
                dke = ((double)v.x * v.x + (double)v.y * v.y + (double)v.z * v.z) / (double)v.w;
                drift = !last;
                if (drift) {
                    v.x = fmaf(fx, h, v.x);
                    v.y = fmaf(fy, h, v.y);
                    v.z = fmaf(fz, h, v.z);
                }
            }
            if (drift) {
                x = make_float4(fmaf(v.x, a.dt, x.x), fmaf(v.y, a.dt, x.y), fmaf(v.z, a.dt, x.z), x.w);
                if (sub == 0) a.pos[(int)((sx + 1) & 1)][slot] = x;
                const float ddx = minimg(x.x - xb.x, L0, Li0), ddy = minimg(x.y - xb.y, L1, Li1),
                            ddz = minimg(x.z - xb.z, L2, Li2);
                if (fmaf(ddz, ddz, fmaf(ddy, ddy, ddx * ddx)) >= a.skin2q) mybits |= FL_TRIP;   // src/neighborlist.cu:174
                if (!(fabsf(x.x) <= L0 && fabsf(x.y) <= L1 && fabsf(x.z) <= L2)) mybits |= FL_STOP;
            }
            if (sub != 0) dke = 0.0;
        }
        // the workgroup's partial sums and flag bits
        const double q0 = need_ke ? wave_sum_d(dke) : 0.0;
        const double q1 = en ? wave_sum_d(elj) : 0.0;
        const double q2 = en ? wave_sum_d(evi) : 0.0;
        const unsigned wb = (__ballot((mybits & FL_TRIP) != 0) ? FL_TRIP : 0u) |
                            (__ballot((mybits & FL_ROWS) != 0) ? FL_ROWS : 0u) |
                            (__ballot((mybits & FL_STOP) != 0) ? FL_STOP : 0u);
        if (lane == 0) {
            s_red[wave][0] = q0;
            s_red[wave][1] = q1;
            s_red[wave][2] = q2;
            if (wb) atomicOr(&s_wgb, wb);
        }
        __syncthreads();
        unsigned wgbits = s_wgb;
        if (tid < 3 && need_ke) {
            double t = s_red[0][tid];
            for (int w = 1; w < NW; w++) t += s_red[w][tid];
            a.pub[((size_t)(b & 1) * MD_PK_WG + blockIdx.x) * 4 + tid] = t;
        }
        if ((wgbits & FL_STOP) && tid == 0) atomicOr(a.err, 1);   // an atom left the box: unstable
        if ((wgbits & FL_ROWS) && tid == 0) atomicOr(a.err, MD_PK_ERR_ROWS);
        if (wgbits & FL_ROWS) wgbits |= FL_STOP;   // the run is replayed with longer rows
        bits = pk_barrier(a, b++, wgbits, &s_bits);
        done = k + 1;
    }
    // the last step's record and factor
    if (done == nsteps && nsteps > 0) {
        if (wave == 0) {
            const double* pb = a.pub + (size_t)((b - 1) & 1) * MD_PK_WG * 4;
            double t0 = 0.0, t1 = 0.0, t2 = 0.0;
            if (lane < a.nwg) {
                t0 = pb[lane * 4 + 0];
                t1 = pb[lane * 4 + 1];
                t2 = pb[lane * 4 + 2];
            }
            t0 = wave_sum_d(t0);
            t1 = wave_sum_d(t1);
            t2 = wave_sum_d(t2);
            if (blockIdx.x == 0 && lane == 0) {
                const long long ks = nsteps - 1;
                const double T = t0 / ((double)p.nf * MD_KB_REF);
                if (a.stats && ks < a.stat_cap) {
                    const bool en = a.energy_every > 0 && ((step0 + ks) % a.energy_every) == 0;
                    StatRec rec;
                    rec.ke = 0.5 * t0;
                    rec.temp = T;
                    rec.xi = a.th[0].xi;
                    rec.vxi = a.th[0].vxi;
                    rec.volume = (double)L0 * (double)L1 * (double)L2;
                    rec.has_energy = en ? 1 : 0;
                    rec.pad = 0;
                    rec.pe_lj = en ? 0.5 * t1 : 0.0;
                    rec.pe_coul = 0.0;
                    rec.pe_bond = 0.0;
                    rec.pe_angle = 0.0;
                    rec.virial = en ? 0.5 * t2 : 0.0;
                    rec.conserved = rec.ke + rec.pe_lj;
                    a.stats[ks] = rec;
                }
                double lam = 1.0;
                if (a.leapfrog && a.thermostat == 1) lam = sqrt(1.0 + dtd / taud * ((double)p.T0 / T - 1.0));
                if (a.leapfrog) a.th[0].S = lam;
                RepState& st = a.st[0];
                st.step = step0 + nsteps;
                st.nb = s0.nb + nbuild;
            }
        }
        if (active && sub == 0) {
            a.vel[slot] = v;
            a.force[slot] = f;
        }
    }
}

extern "C" hipError_t md_launch_persist(const PersistArgs& a, int lanes, size_t shmem, hipStream_t st, hipEvent_t ev0,
                                        hipEvent_t ev1) {
    const dim3 grid(a.nwg), block(MD_PK_THREADS);
    const bool nt1 = a.ntype == 1;
    hipError_t rc = hipSuccess;
    auto go = [&](auto kern) {
        // more than 64 KB of dynamic LDS is requested (one workgroup per CU)
        rc = hipFuncSetAttribute(reinterpret_cast<const void*>(kern), hipFuncAttributeMaxDynamicSharedMemorySize,
                                 (int)shmem);
        if (rc != hipSuccess) return;
        if (ev0) hipExtLaunchKernelGGL(kern, grid, block, shmem, st, ev0, ev1, 0, a);
        else hipLaunchKernelGGL(kern, grid, block, shmem, st, a);
    };
    switch (lanes) {
    case 4: nt1 ? go(k_persist<4, true>) : go(k_persist<4, false>); break;
    case 8: nt1 ? go(k_persist<8, true>) : go(k_persist<8, false>); break;
    case 16: nt1 ? go(k_persist<16, true>) : go(k_persist<16, false>); break;
    default: return hipErrorInvalidValue;
    }
    if (rc != hipSuccess) return rc;
    return hipGetLastError();
}

// zero the barrier flags before a launch (a fill kernel, so a caller may capture it)
__global__ void k_persist_clear(unsigned* flags) {
    if (threadIdx.x < 2 * MD_PK_WG) flags[threadIdx.x] = 0u;
}

extern "C" hipError_t md_launch_persist_clear(unsigned* flags, hipStream_t st) {
    hipLaunchKernelGGL(k_persist_clear, dim3(1), dim3(2 * MD_PK_WG), 0, st, flags);
    return hipGetLastError();
}
