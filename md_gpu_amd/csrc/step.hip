// step.hip -- the fused MD step: pair forces (LJ + Wolf / reaction-field /
// damped Coulomb), harmonic bonds and angles, and the integrator epilogue,
// in one launch per step.
//
// Reference arithmetic restated (FP32, same operation order):
//   pair loop          src/pair_forces.cu:151-200 (Wolf), :60-126 (RF)
//   bonds / angles     src/bonded_forces.cu:50-91, :119-188
//   leapfrog NVE/NVT   src/nve_leapfrog.cu:37-55, src/nvt_leapfrog_ber.cu:35-55
//   Berendsen lambda   src/integrator.cu:146-155 (computed here on the device,
//                      from the previous step's temperature, as the reference does)
//   velocity Verlet    textbook half kicks (the reference's vverlet_nve_pre
//                      applies the half kick twice, src/nve_vverlet.cu:49-54)
//   Nose-Hoover        MTK chain-of-one Trotter half steps (Martyna, Tuckerman,
//                      Klein, Mol. Phys. 87 (1996) 1117), no reference code exists
//
// Work decomposition: one workgroup per 64-slot tile, G lanes per atom (64 G
// threads).  The workgroup stages the tile's neighbor set into LDS (the
// positions of the recorded images, md_rows.h), then each atom's G lanes walk
// its row of 16-bit local indices (the LDS-staged tiles of
// src/pair_forces_tiles.cu:98-155, restricted to the atoms that are
// neighbors).  Per-atom sums are reduced across the G lanes with DPP row
// moves; KE / energies / virial and the skin-check flag are reduced per
// workgroup into FP64 partial slots, one per tile (plain stores).
// k_step_finalize, launched next, reduces a replica's slots and computes its
// temperature, thermostat update and statistics record, folds the skin-check
// flag into the list state and advances the step counter.  No device-scope
// atomics on shared addresses in the hot kernel: with thousands of workgroups
// they serialise behind each other.
// A replica is skipped when its list is stale or its run target is reached,
// so a captured segment can be replayed with no host decision.
//
// Positions run unwrapped between list builds (the build wraps them), so the
// drift x += v dt has no wrap here; an atom is at most skin/2 outside the box
// before the skin check forces a build.
//
// Algorithmic work per ordered pair (FMA = 2, add/mul/rsqrt/rcp/exp2/rint = 1):
// LJ + Wolf 65 flop, LJ + RF 45, LJ 36, Wolf only 60 (src/pair_forces.cu:151-200).
#include <hip/hip_ext.h>

#include <algorithm>
#include <type_traits>

#include "md_reduce.h"
#include "md_rows.h"

namespace {

// MTK chain-of-one half step: returns the velocity scale, updates xi, vxi, K.
__device__ double nh_half(double& xi, double& vxi, double& K, double nf, double kT0, double Q, double dt) {
    double G = (2.0 * K - nf * kT0) / Q;
    vxi += 0.25 * dt * G;
    double s = exp(-0.5 * dt * vxi);
    K *= s * s;
    xi += 0.5 * dt * vxi;
    G = (2.0 * K - nf * kT0) / Q;
    vxi += 0.25 * dt * G;
    return s;
}

// Leapfrog end of step for one replica: temperature, the Berendsen factor for
// the next step (src/integrator.cu:146-147), list age, skin/age trip and step
// count.  The finalize kernel and the deferred finalize inside the step kernel
// both call it, so the two schedules give bit-identical trajectories.
__device__ __forceinline__ void lf_done(const StepArgs& a, float nf, float T0, RepState& s, double ke2, bool tn,
                                        double& T, double& lam) {
    const double dt = a.dt, tau = a.tau;
    T = ke2 / ((double)nf * MD_KB_REF);
    lam = 1.0;
    if (a.thermostat == 1) lam = sqrt(1.0 + dt / tau * ((double)T0 / T - 1.0));
    s.age = s.age + 1;
    s.step = s.step + 1;
    // list refresh by age (ahead of the skin trip), so queued steps do not stall;
    // the skin check of the positions just written makes the list stale for them
    const bool aged = a.max_age > 0 && s.age >= a.max_age;
    if (tn || aged) {
        if (s.nb & 1) s.trip[1] = 1;
        else s.trip[0] = 1;
    }
    // velocity Verlet in leapfrog form: the run's last step began no next step
    if (a.vvfused && s.step >= s.target) s.phase = 0;
}

// Statistics record of a finished step (index kstat of the run).
__device__ __forceinline__ void put_record(const StepArgs& a, int r, long long kstat, bool en, const double* v,
                                           double ke, double temp, double xi, double vxi, double nh_terms) {
    if (!(a.stats && kstat >= 0 && kstat < a.stat_cap)) return;
    StatRec rec;
    rec.ke = ke;
    rec.temp = temp;
    rec.xi = xi;
    rec.vxi = vxi;
    const RepParam& pb = a.rep[r];   // the box of the step's forces (a coupled box is scaled after the record)
    rec.volume = (double)pb.L[0] * (double)pb.L[1] * (double)pb.L[2];
    rec.has_energy = en ? 1 : 0;
    rec.pad = 0;
    rec.pe_lj = 0.5 * v[ACC_LJ];
    rec.pe_coul = 0.5 * v[ACC_COUL];
    rec.pe_bond = v[ACC_BOND];
    rec.pe_angle = v[ACC_ANGLE];
    rec.virial = 0.5 * v[ACC_VIR];
    rec.conserved = rec.ke + rec.pe_lj + rec.pe_coul + rec.pe_bond + rec.pe_angle + nh_terms;
    a.stats[(size_t)kstat * a.nrep + r] = rec;
}

// Gate counters after a replica went from state o to state n within one leapfrog
// launch or block (target unchanged).
__device__ __forceinline__ void lf_gates(const StepArgs& a, const RepState& o, const RepState& n) {
    const int ds = (int)stale(n) - (int)stale(o);
    const int dr = (int)(!stale(n) && n.step < n.target) - (int)(!stale(o) && o.step < o.target);
    if (ds) atomicAdd(&a.gate->n_stale, ds);
    if (dr) atomicAdd(&a.gate->n_run, dr);
}

// Runs once per replica and step-kernel launch, by thread 0 of the replica's
// workgroup in k_step_finalize; v = the replica's reduced partial sums, tn = the
// skin check tripped during the position update.
__device__ void finalize_replica(const StepArgs& a, int r, const RepParam& p, const RepState& st0, ThState th,
                                 long long kstat, bool en, const double* v, bool tn) {
    const double dt = a.dt, tau = a.tau;
    if (a.mode == MODE_LF) {
        RepState s = st0;
        double T, lam;
        lf_done(a, p.nf, p.T0, s, v[ACC_KE2], tn, T, lam);
        put_record(a, r, kstat, en, v, 0.5 * v[ACC_KE2], T, th.xi, th.vxi, 0.0);
        a.th[r].S = lam;
        RepState& st = a.st[r];
        st.age = s.age;
        st.step = s.step;
        st.phase = s.phase;
        st.trip[0] = s.trip[0];
        st.trip[1] = s.trip[1];
        lf_gates(a, st0, s);
        return;
    }
    double nh_terms = 0.0, ke = 0.0, temp = 0.0, rxi = th.xi, rvxi = th.vxi;
    if (a.mode == MODE_VV) {
        double K = 0.5 * v[ACC_KE2];
        double xi = th.xi, vxi = th.vxi, s1 = 1.0;
        if (a.thermostat == 2) {
            s1 = nh_half(xi, vxi, K, p.nf, p.kT0, p.Q, dt);
        } else if (a.thermostat == 1) {
            double T = 2.0 * K / ((double)p.nf * MD_KB_REF);
            s1 = sqrt(1.0 + dt / tau * ((double)p.T0 / T - 1.0));
            K *= s1 * s1;
        }
        th.S_end = s1;
        th.xi_end = xi;
        th.vxi_end = vxi;
        th.ke2_cur = 2.0 * K;
        ke = K;
        temp = 2.0 * K / ((double)p.nf * MD_KB_REF);
        rxi = xi;
        rvxi = vxi;
        if (a.thermostat == 2) nh_terms = 0.5 * p.Q * vxi * vxi + (double)p.nf * p.kT0 * xi;
        double s2 = 1.0;
        if (a.thermostat == 2) s2 = nh_half(xi, vxi, K, p.nf, p.kT0, p.Q, dt);
        th.xi = xi;
        th.vxi = vxi;
        th.S = s1 * s2;
    }
    put_record(a, r, kstat, en, v, ke, temp, rxi, rvxi, nh_terms);
    if (a.mode == MODE_FORCE) return;
    a.th[r] = th;
    RepState& st = a.st[r];
    const int age = st0.age + 1;
    const long long step = st0.step + 1;
    st.age = age;
    st.step = step;
    const bool aged = a.max_age > 0 && age >= a.max_age;
    st.phase = 0;
    atomicSub(&a.gate->n_force, 1);
    if (aged) {
        st.trip[st0.nb & 1] = 1;
        atomicAdd(&a.gate->n_stale, 1);
    } else if (step < st0.target) {
        atomicAdd(&a.gate->n_run, 1);
    }
}

// Totals over the NW wavefronts of the workgroup of per-thread sums v; every
// thread gets the totals.  The order of the additions is fixed, so every caller
// with the same workgroup size gets the same bits.  ALL = false: only KE and
// the trip flag.
template <int NW, bool ALL>
__device__ __forceinline__ void reduce_slots(double (*s_red)[MD_NSTAT], double* v) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
    for (int k = 0; k < MD_NSTAT; k++) {
        if (!ALL && k != ACC_KE2 && k != MD_NSTAT - 1) continue;
        v[k] = wave_sum_d(v[k]);
    }
    __syncthreads();   // s_red is free
    if (lane == 0) {
#pragma unroll
        for (int k = 0; k < MD_NSTAT; k++) s_red[wave][k] = v[k];
    }
    __syncthreads();
    // thread k adds quantity k over the wavefronts in order; the totals go to row 0
    double t = 0.0;
    if (threadIdx.x < MD_NSTAT) {
        t = s_red[0][threadIdx.x];
        for (int w = 1; w < NW; w++) t += s_red[w][threadIdx.x];
    }
    __syncthreads();
    if (threadIdx.x < MD_NSTAT) s_red[0][threadIdx.x] = t;
    __syncthreads();
#pragma unroll
    for (int k = 0; k < MD_NSTAT; k++) v[k] = s_red[0][k];
}

// Sums of a replica's partial slots (its tiles t0 .. t1-1) in one buffer of part
// ([MD_NSTAT][ntiles]: one plane per quantity, so ALL = false reads 16 bytes a
// tile).  Thread k adds tiles t0 + k, t0 + k + blockDim, ... in that order.
template <int NW, bool ALL>
__device__ __forceinline__ void sum_tiles(const double* part, int t0, int t1, int ntiles,
                                          double (*s_red)[MD_NSTAT], double* v) {
#pragma unroll
    for (int k = 0; k < MD_NSTAT; k++) v[k] = 0.0;
    for (int t = t0 + (int)threadIdx.x; t < t1; t += (int)blockDim.x) {
#pragma unroll
        for (int j = 0; j < MD_NSTAT; j++)
            if (ALL || j == ACC_KE2 || j == MD_NSTAT - 1) v[j] += part[(size_t)j * ntiles + t];
    }
    reduce_slots<NW, ALL>(s_red, v);
}

// One pair: force factor (F_i += ff * d) and, with EN, the pair energies.
// Wolf / damped forces come from the cubic table of h(r) in LDS (s_ct); the
// energies keep the closed form.
template <int COUL, bool EN>
__device__ __forceinline__ float pair_ff(const StepArgs& a, const float4* s_ct, float qi, float qj, float rsq,
                                         const float4 lj, float& ulj, float& ucoul) {
    const float ri = __builtin_amdgcn_rsqf(rsq);
    const float ri2 = ri * ri, ri3 = ri2 * ri;
    const float ri6 = ri3 * ri3, ri8 = ri6 * ri2;
    const float flj = fmaf(lj.x, ri6, -lj.y) * ri8;   // lj1 r^-14 - lj2 r^-8, src/pair_forces.cu:190-193
    float ff = flj;
    if (COUL == 0) {           // reaction field, src/pair_forces.cu:88-91
        const float qq = qi * qj;
        ff = fmaf(qq * (ri - a.reacf * rsq), ri2, flj);
        if (EN) ucoul = qq * (ri + 0.5f * a.reacf * rsq - a.ecut);
    } else if (COUL == 1 || COUL == 2) {   // Wolf / damped, src/pair_forces.cu:174-186
        const float qq = qi * qj;
        const float r = rsq * ri;
        const float hv = ctab_eval(s_ct, r * a.ct_scale);   // A(r) - r A(rc)/rc
        ff = fmaf(hv * qq, ri3, flj);                      // (A(r)/r - A(rc)/rc) qq / r^2
        if (EN) {
            const float tt = __builtin_amdgcn_rcpf(fmaf(a.palfa, r, 1.0f));
            const float poly = ((((1.061405429f * tt - 1.453152027f) * tt + 1.421413741f) * tt - 0.284496736f) * tt +
                                0.254829592f) * tt;
            const float ex = __builtin_amdgcn_exp2f(a.alfasq * rsq);
            float u = poly * ex * ri - a.ecut;
            if (COUL == 1) u += a.wolf * (__builtin_amdgcn_logf(r) - a.log2rc) * 0.69314718056f;
            ucoul = qq * u;
        }
    }
    if (EN) ulj = fmaf(lj.x * (1.0f / 12.0f) * ri6, ri6, -lj.y * (1.0f / 6.0f) * ri6) - lj.z;
    return ff;
}

}  // namespace

// ENMODE 0: no energies (compiled out); 1: energies where requested (runtime, per replica)
// NT2: at most two atom types, so the LJ parameters of a pair are one of the
// two entries of the atom's table row, held in registers and picked by a select
// (no LDS read per pair); otherwise the row is read from LDS per pair.
template <int G, int COUL, int ENMODE, bool NT2>
__global__ __launch_bounds__(64 * G, COUL == 3 ? 5 : 4) void k_step(const StepArgs a) {
    constexpr int NW = G;                 // wavefronts: 64 atoms x G lanes
    extern __shared__ float4 s_j[];       // [jcap + 1] the tile's staged neighbor set
    __shared__ float4 s_lj[64];
    __shared__ float4 s_ct[(COUL == 1 || COUL == 2) ? MD_CTAB + 1 : 1];
    __shared__ float4 s_sh[27];
    __shared__ double s_red[NW][MD_NSTAT];
    static_assert(MD_CTAB == 256, "the table is loaded by the first 256 threads");
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int sub = tid & (G - 1);
    const int tile = xcd_block(blockIdx.x, gridDim.x);   // tiles contiguous per XCD
    if (a.prefilter) {
        // Many replicas, no deferred finalize: a tile whose replica does not step in
        // this launch (its list waits for the next rebuild chain, or it reached its
        // target) leaves after two scalar loads, before the speculative loads below
        // (about 16 KB per tile) are issued.
        const int r0 = __builtin_amdgcn_readfirstlane(a.tile_rep[tile]);
        const RepState s0 = a.st[r0];
        bool go;
        if (a.mode == MODE_FORCE) go = !stale(s0);
        else if (a.mode == MODE_LF) go = !stale(s0) && s0.step < s0.target;
        else go = s0.phase == 1 && !stale(s0);
        if (!go) return;   // workgroup-uniform
    }
    // lane group tid / G takes the atom of that rank in the tile's row-length order
    const int slot = tile * 64 + a.rperm[tile * 64 + tid / G];
    // Loads that do not depend on the replica's state, issued before any is used:
    // the gate, the tables, the replica's constants and state, the tile's set
    // entries (JPRE per thread; entries past the set are loaded and not used:
    // every entry of the buffer is some slot), the atom's data in both position
    // buffers, its row's first two blocks, and its velocity and build position
    // (epilogue).
    // (LJ-only systems have sparser sets and need the registers for occupancy)
    constexpr int NTH = 64 * G, JPRE = (COUL == 3 ? 16 : 32) / G;
    const Gate gate = *a.gate;
    const int r = __builtin_amdgcn_readfirstlane(a.tile_rep[tile]);
    const float4 lj_v = a.ljtab[tid & 63];
    float4 ct_v = make_float4(0.f, 0.f, 0.f, 0.f), ct_e = ct_v;
    if (COUL == 1 || COUL == 2) {
        if (tid < 256) ct_v = a.ctab[tid];
        ct_e = a.ctab[MD_CTAB];
    }
    const int jn = min(a.jcount[tile], a.jcap);
    const int id = a.ids[slot], tyi = a.typ[slot], nci = a.ncount[slot];
    RepState s = a.lag ? a.lst[(size_t)(a.par ^ 1) * a.nrep + r] : a.st[r];
    const RepParam& pr = a.rep[r];
    const int so = pr.slot_off, ns = pr.nslot;
    const float nf = pr.nf, T0 = pr.T0;
    const float L0 = pr.L[0], L1 = pr.L[1], L2 = pr.L[2];
    const float Li0 = pr.Linv[0], Li1 = pr.Linv[1], Li2 = pr.Linv[2];
    const double c_S = a.th[r].S;
    const uint32_t* __restrict__ jt = a.jset + (size_t)tile * a.jcap;
    uint32_t je[JPRE];
#pragma unroll
    for (int u = 0; u < JPRE; u++) je[u] = jt[min(tid + u * NTH, a.jcap - 1)];
    const float4 xa = a.pos[0][slot], xb = a.pos[1][slot];
    const uint4* __restrict__ row4 = reinterpret_cast<const uint4*>(a.nl16 + (size_t)slot * a.cap);
    const uint4 q0 = row4[sub];
    const uint4 q1 = row4[(a.cap >= 16 * G ? G : 0) + sub];   // block 1 when the capacity holds one
    float4 v_e = make_float4(0.f, 0.f, 0.f, 0.f), pb_e = v_e, fr_e = v_e;
    if (sub == 0) {
        v_e = a.vel[slot];
        pb_e = a.pos_build[slot];
        if (COUL == 2 && a.frec) fr_e = a.frec[slot];   // PME: long-range force of the atom
    }
    loads_issued();
    if (tid < 64) s_lj[tid] = lj_v;
    if (COUL == 1 || COUL == 2) {
        if (tid < 256) s_ct[tid] = ct_v;
        if (tid == 0) s_ct[MD_CTAB] = ct_e;
    }
    stage_shift_table(s_sh, L0, L1, L2);
    // nothing can step (and no list was just rebuilt for this block to open): idle launch
    if (a.mode == MODE_LF && gate.n_run == 0 && !(a.open && gate.n_stale != 0)) return;
    if (a.mode == MODE_VV && gate.n_force == 0) return;
    const int t0 = so / 64, t1 = (so + ns) / 64;
    const bool leader = tile == t0;   // the workgroup of the replica's first tile
    // The position buffer this step reads: the deferred finalize below advances the
    // step counter (and ends a fused velocity-Verlet run's phase) first.
    const bool lagsum = a.lag && s.ran;
    RepState sp = s;
    if (lagsum) {
        sp.step += 1;
        if (a.vvfused && sp.step >= sp.target) sp.phase = 0;
    }
    const float4* __restrict__ pin = a.pos[cur_buf(sp)];
    if (!a.lagmode) {
        // No launch of this schedule hands state on: a tile whose replica does not
        // step now (its list is stale and waits for the next rebuild chain, or it
        // reached its target) leaves before gathering and staging its set.
        bool go;
        if (a.mode == MODE_FORCE) go = !stale(s);
        else if (a.mode == MODE_LF) go = !stale(s) && s.step < s.target;
        else go = s.phase == 1 && !stale(s);
        if (!go) return;   // workgroup-uniform: one replica per tile
    }
    // Round trip 2: the neighbors' positions and, for the deferred finalize, the
    // previous launch's slots of the replica (added in the order of sum_tiles).
    float4 xg[JPRE];
#pragma unroll
    for (int u = 0; u < JPRE; u++) xg[u] = pin[(tid + u * NTH < jn) ? (je[u] & MD_SLOTMASK) : 0u];
    const bool en_prev = lagsum && a.energy_every > 0 && (s.step % a.energy_every) == 0;
    const bool sum_all = leader && en_prev && a.stats;
    double v[MD_NSTAT];
#pragma unroll
    for (int k = 0; k < MD_NSTAT; k++) v[k] = 0.0;
    if (lagsum) {
        const double* prev = a.part + (size_t)(a.par ^ 1) * MD_NSTAT * a.ntiles;
        for (int t = t0 + tid; t < t1; t += NTH) {
#pragma unroll
            for (int j = 0; j < MD_NSTAT; j++)
                if (sum_all || j == ACC_KE2 || j == MD_NSTAT - 1) v[j] += prev[(size_t)j * a.ntiles + t];
        }
    }
    // the staged set: the image of each neighbor (x + k L per shift code, the
    // table s_sh), in w the type of j (packed into the charge's low mantissa bits,
    // or as a float without charges); local index 0 = the dummy
    constexpr bool CHG = COUL != 3;
    __syncthreads();   // s_sh visible
#pragma unroll
    for (int u = 0; u < JPRE; u++) {
        const int k = tid + u * NTH;
        if (k < jn) {
            const uint32_t e = je[u];
            const float4 sv = s_sh[(e >> MD_SLOTBITS) & 31u];
            s_j[k + 1] = make_float4(xg[u].x + sv.x, xg[u].y + sv.y, xg[u].z + sv.z, staged_w<CHG>(xg[u].w, e));
        }
    }
    for (int k = tid + JPRE * NTH; k < jn; k += NTH) {   // sets beyond JPRE entries per thread
        const uint32_t e = jt[k];
        const float4 x = pin[e & MD_SLOTMASK];
        const float4 sv = s_sh[(e >> MD_SLOTBITS) & 31u];
        s_j[k + 1] = make_float4(x.x + sv.x, x.y + sv.y, x.z + sv.z, staged_w<CHG>(x.w, e));
    }
    if (tid == 0) s_j[0] = make_float4(MD_FAR, MD_FAR, MD_FAR, 0.0f);
    double lamd = 1.0;
    if (a.lag) {
        // Deferred finalize of the previous launch (same block): every workgroup sums
        // the replica's KE and trip flag from the previous partial slots and advances
        // the hand-over state; the leader also writes the statistics record.  Same
        // arithmetic as k_step_finalize.
        lamd = s.lam;
        if (lagsum) {
            if (sum_all) reduce_slots<NW, true>(s_red, v);
            else reduce_slots<NW, false>(s_red, v);
            const long long kst = s.step - s.start;
            double T;
            lf_done(a, nf, T0, s, v[ACC_KE2], v[MD_NSTAT - 1] != 0.0, T, lamd);
            if (leader && tid == 0 && a.stats)
                put_record(a, r, kst, en_prev, v, 0.5 * v[ACC_KE2], T, a.th[r].xi, a.th[r].vxi, 0.0);
        }
    } else {
        if (a.open && stale(s)) {   // rebuilt by the block's chain: the new list epoch opens
            s.nb += 1;
            s.age = 0;
        }
        if (a.mode == MODE_LF) lamd = c_S;
    }
    bool run_r;
    if (a.mode == MODE_FORCE) run_r = !stale(s);
    else if (a.mode == MODE_LF) run_r = !stale(s) && s.step < s.target;
    else run_r = s.phase == 1 && !stale(s);
    if (a.lagmode && leader && tid == 0) {   // hand-over to the next launch
        RepState o = s;
        o.ran = run_r ? 1 : 0;
        o.lam = lamd;
        a.lst[(size_t)a.par * a.nrep + r] = o;
    }
    if (!run_r) return;   // workgroup-uniform: one replica per tile
    const bool en = ENMODE != 0 && ((a.mode == MODE_FORCE) ? (a.force_energy != 0)
                                                           : (a.energy_every > 0 && (s.step % a.energy_every) == 0));
    const int cb = cur_buf(s);
    float4* pout = a.pos[cb ^ 1];
    const float lam = (float)lamd;
    const bool drift = !a.vvfused || s.step + 1 < s.target;
    const float4 xi = cb ? xb : xa;
    const bool active = id >= 0;
    const int nn = active ? min(nci, a.cap) : 0;
    __syncthreads();   // tables and the staged set visible

    const float qi = xi.w * a.qfac * a.qq_scale;
    const float4* s_lji = s_lj + (active ? tyi : 0) * 8;
    const float4 lj_a = s_lji[0], lj_b = s_lji[1];
    const float4 lj_d = make_float4(lj_b.x - lj_a.x, lj_b.y - lj_a.y, lj_b.z - lj_a.z, 0.0f);
    float fx = 0.f, fy = 0.f, fz = 0.f;
    double elj = 0.0, eco = 0.0, evi = 0.0;
    auto pair = [&](const float4& pj, auto en_tag) {
        constexpr bool EN = decltype(en_tag)::value;
        float4 lj;
        if (NT2 && !CHG) {   // w = type (0 or 1) as a float: the row entry by interpolation
            lj = make_float4(fmaf(pj.w, lj_d.x, lj_a.x), fmaf(pj.w, lj_d.y, lj_a.y), fmaf(pj.w, lj_d.z, lj_a.z), 0.0f);
        } else if (NT2) {
            const bool tb = (__float_as_uint(pj.w) & 1u) != 0;   // type 1 (types are 0 or 1)
            lj = make_float4(tb ? lj_b.x : lj_a.x, tb ? lj_b.y : lj_a.y, tb ? lj_b.z : lj_a.z, 0.0f);
        } else {
            lj = s_lji[CHG ? staged_type(pj) : (int)pj.w];
        }
        const float dx = xi.x - pj.x, dy = xi.y - pj.y, dz = xi.z - pj.z;
        const float rsq = fmaf(dz, dz, fmaf(dy, dy, dx * dx));
        float ulj = 0.0f, ucoul = 0.0f;
        float ff;
        if (COUL == 3 && !EN) ff = pair_ff_lj(rsq, lj);
        else ff = pair_ff<COUL, EN>(a, s_ct, qi, pj.w, rsq, lj, ulj, ucoul);
        const bool use = !(rsq > a.rc2);   // cutoff as src/pair_forces.cu:195 (zero for rsq > rc2)
        ff = use ? ff : 0.0f;
        fx = fmaf(ff, dx, fx);
        fy = fmaf(ff, dy, fy);
        fz = fmaf(ff, dz, fz);
        if (EN && use) {
            elj += (double)ulj;
            eco += (double)ucoul;
            evi += (double)(ff * rsq);
        }
    };
    // lane groups hold the tile's atoms by decreasing row length (rperm)
    if (ENMODE != 0 && en)
        walk_row_sorted<G>(s_j, row4, nn, sub, q0, q1, [&](const float4& pj) { pair(pj, std::true_type{}); });
    else
        walk_row_sorted<G>(s_j, row4, nn, sub, q0, q1, [&](const float4& pj) { pair(pj, std::false_type{}); });

    // bonded terms (lanes of the atom's group share them)
    double ebo = 0.0, ean = 0.0, bvi = 0.0;
    if (a.bonded && active) {
        const int nb = a.nbond[id];
        for (int b = sub; b < nb; b += G) {          // src/bonded_forces.cu:50-84
            const uint32_t e = a.bond_list[(size_t)id * 4 + b];
            const int js = a.slot_of[e & MD_JMASK];
            const int bt = (int)(e >> MD_JBITS);
            const float4 pj = pin[js];
            const float dx = minimg(xi.x - pj.x, L0, Li0);
            const float dy = minimg(xi.y - pj.y, L1, Li1);
            const float dz = minimg(xi.z - pj.z, L2, Li2);
            const float rsq = fmaf(dz, dz, fmaf(dy, dy, dx * dx));
            const float kk = a.bk2[bt], r0 = a.br0[bt];
            const float fa = kk * (r0 * __builtin_amdgcn_rsqf(rsq) - 1.0f);
            fx = fmaf(fa, dx, fx);
            fy = fmaf(fa, dy, fy);
            fz = fmaf(fa, dz, fz);
            if (ENMODE != 0 && en) {
                const float dr = sqrtf(rsq) - r0;
                ebo += 0.25 * (double)kk * (double)dr * (double)dr;   // each bond is seen from both atoms
                bvi += 0.5 * (double)(fa * rsq);
            }
        }
        const int na = a.nang[id];
        for (int q = sub; q < na; q += G) {          // src/bonded_forces.cu:119-188
            const int4 e = a.ang_list[(size_t)id * 4 + q];
            const int role = e.x & 0xff, at = e.x >> 8;
            const float4 pA = pin[a.slot_of[e.y]];
            const float4 pk = pin[a.slot_of[e.z]];
            float4 pi = xi, pc = pA;
            if (role == 1) { pi = pA; pc = xi; }
            float ax = minimg(pi.x - pc.x, L0, Li0), ay = minimg(pi.y - pc.y, L1, Li1), az = minimg(pi.z - pc.z, L2, Li2);
            float ai_ = __builtin_amdgcn_rsqf(fmaf(az, az, fmaf(ay, ay, ax * ax)));
            float bx = minimg(pk.x - pc.x, L0, Li0), by = minimg(pk.y - pc.y, L1, Li1), bz = minimg(pk.z - pc.z, L2, Li2);
            float bi = __builtin_amdgcn_rsqf(fmaf(bz, bz, fmaf(by, by, bx * bx)));
            const float ab = fmaf(az, bz, fmaf(ay, by, ax * bx));
            float cang = ab * ai_ * bi;
            cang = fminf(fmaxf(cang, -0.9999999f), 0.9999999f);
            const float th = acosf(cang);
            const float fk = a.ak2[at] * (th - a.at0[at]) * __builtin_amdgcn_rsqf(1.0f - cang * cang) * ai_ * bi;
            const float ca = ab * ai_ * ai_;
            float gx = bx - ca * ax, gy = by - ca * ay, gz = bz - ca * az;
            if (role == 1) {
                const float cbb = ab * bi * bi;
                gx = cbb * bx - ax - gx;
                gy = cbb * by - ay - gy;
                gz = cbb * bz - az - gz;
            }
            fx = fmaf(fk, gx, fx);
            fy = fmaf(fk, gy, fy);
            fz = fmaf(fk, gz, fz);
            if (ENMODE != 0 && en) {
                const double d = (double)th - (double)a.at0[at];
                ean += (0.5 / 3.0) * (double)a.ak2[at] * d * d;   // each angle is seen from 3 atoms
            }
        }
    }
    fx = group_sum<G>(fx);
    fy = group_sum<G>(fy);
    fz = group_sum<G>(fz);
    double dke = 0.0;
    int trip = 0;
    if (active && sub == 0) {
        if (COUL == 2 && a.frec) {
            fx += fr_e.x;
            fy += fr_e.y;
            fz += fr_e.z;
        }
        a.force[slot] = make_float4(fx, fy, fz, 0.0f);
        if (a.mode == MODE_LF) {
            float4 v = v_e;                 // loaded in the first round trip
            const float4 pb = pb_e;         // for the skin check
            if (a.vvfused) {
                // velocity Verlet, NVE: v(t) = v(t - dt/2) + F dt/2m ends the step (its kinetic
                // energy is the step's); v(t + dt/2) = v(t) + F dt/2m begins the next.  The same
                // operations as the kick kernel and the force-half epilogue below.
                const float h = 0.5f * a.dt * v.w;
                v.x = fmaf(fx, h, v.x);
                v.y = fmaf(fy, h, v.y);
                v.z = fmaf(fz, h, v.z);
                dke = ((double)v.x * v.x + (double)v.y * v.y + (double)v.z * v.z) / (double)v.w;
                if (drift) {
                    v.x = fmaf(fx, h, v.x * lam);
                    v.y = fmaf(fy, h, v.y * lam);
                    v.z = fmaf(fz, h, v.z * lam);
                }
            } else {
                const float massif = a.dt * v.w;          // dt / m_int (src/nve_leapfrog.cu:38)
                v.x = fmaf(fx, massif, v.x) * lam;
                v.y = fmaf(fy, massif, v.y) * lam;
                v.z = fmaf(fz, massif, v.z) * lam;
                dke = ((double)v.x * v.x + (double)v.y * v.y + (double)v.z * v.z) / (double)v.w;
            }
            a.vel[slot] = v;
            if (drift) {
                const float x = fmaf(v.x, a.dt, xi.x), y = fmaf(v.y, a.dt, xi.y), z = fmaf(v.z, a.dt, xi.z);
                pout[slot] = make_float4(x, y, z, xi.w);
                if (!a.baro) {
                    const float ddx = minimg(x - pb.x, L0, Li0), ddy = minimg(y - pb.y, L1, Li1), ddz = minimg(z - pb.z, L2, Li2);
                    if (fmaf(ddz, ddz, fmaf(ddy, ddy, ddx * ddx)) >= a.skin2q) trip = 1;   // src/neighborlist.cu:174
                } else {
                    // Coupled box: since the build every pair vector was scaled by m = L / Lb
                    // besides the atoms' own moves d = x - m x_build, so a pair's distance is
                    // at least m r_list - |d_i| - |d_j|: the list holds while |d| stays below
                    // (min(m) r_list - r_c) / 2 (skin / 2 when m = 1).
                    const float m0 = L0 / pr.Lb[0], m1 = L1 / pr.Lb[1], m2 = L2 / pr.Lb[2];
                    const float ddx = minimg(x - m0 * pb.x, L0, Li0), ddy = minimg(y - m1 * pb.y, L1, Li1),
                                ddz = minimg(z - m2 * pb.z, L2, Li2);
                    const float h = 0.5f * (fminf(m0, fminf(m1, m2)) * a.rl - a.rc);
                    if (!(h > 0.0f) || fmaf(ddz, ddz, fmaf(ddy, ddy, ddx * ddx)) >= h * h) trip = 1;
                }
                if (!(fabsf(x) <= L0 && fabsf(y) <= L1 && fabsf(z) <= L2)) *a.err = 1;
            }
        } else if (a.mode == MODE_VV) {
            float4 v = v_e;
            const float h = 0.5f * a.dt * v.w;
            v.x = fmaf(fx, h, v.x);
            v.y = fmaf(fy, h, v.y);
            v.z = fmaf(fz, h, v.z);
            a.vel[slot] = v;
            dke = ((double)v.x * v.x + (double)v.y * v.y + (double)v.z * v.z) / (double)v.w;
        }
    }
    // the workgroup's partial sums go to the tile's slot (plain stores; the next
    // kernel reduces the replica's slots)
    double vals[MD_NSTAT] = {dke, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    if (ENMODE != 0 && en) {
        vals[ACC_LJ] = elj;
        vals[ACC_COUL] = eco;
        vals[ACC_BOND] = ebo;
        vals[ACC_ANGLE] = ean;
        vals[ACC_VIR] = evi + 2.0 * bvi;
    }
#pragma unroll
    for (int k = 0; k < MD_NSTAT - 2; k++) {
        if (ENMODE == 0 && k != ACC_KE2) continue;
        vals[k] = wave_sum_d(vals[k]);
    }
    const unsigned long long tb = __ballot(trip != 0);
    if (lane == 0) {
#pragma unroll
        for (int k = 0; k < MD_NSTAT; k++) s_red[wave][k] = vals[k];
        s_red[wave][MD_NSTAT - 1] = tb ? 1.0 : 0.0;
    }
    __syncthreads();
    if (tid < MD_NSTAT) {
        double t = s_red[0][tid];
#pragma unroll
        for (int w = 1; w < NW; w++) t += s_red[w][tid];
        if (COUL == 2 && a.pme_e && ENMODE != 0 && en && leader) {
            // PME: the replica's long-range terms (the slots hold sums over ordered
            // pairs, halved by the finalize)
            if (tid == ACC_COUL) t += 2.0 * a.pme_e[0];
            if (tid == ACC_VIR) t += 2.0 * a.pme_e[1];
        }
        a.part[((size_t)a.par * MD_NSTAT + tid) * a.ntiles + tile] = t;
    }
}

// After k_step: one workgroup per replica (64 G threads, the step kernel's
// workgroup size, so the slot sums have the deferred finalize's bits) sums the
// partial slots of the replica's tiles and finalizes it.  After a leapfrog block
// with deferred finalize (a.lagmode) the replica's state comes from the last
// launch's hand-over copy lst[par]: this kernel finalizes that launch's step,
// writes the state back and settles the gate counters for the whole block.
template <int G>
__global__ __launch_bounds__(64 * G) void k_step_finalize(const StepArgs a) {
    constexpr int NW = G;
    __shared__ double s_red[NW][MD_NSTAT];
    const int r = blockIdx.x;
    const double* part = a.part + (size_t)a.par * MD_NSTAT * a.ntiles;
    const RepState s = a.lagmode ? a.lst[(size_t)a.par * a.nrep + r] : a.st[r];
    const RepParam p = a.rep[r];
    ThState th;
    RepState s_old;
    if (threadIdx.x == 0) {
        th = a.th[r];
        if (a.lagmode) s_old = a.st[r];
    }
    const int t0 = p.slot_off / 64, t1 = (p.slot_off + p.nslot) / 64;
    double v[MD_NSTAT];
    sum_tiles<NW, true>(part, t0, t1, a.ntiles, s_red, v);   // used only if the replica ran
    if (a.lagmode) {
        // No launch of the block passed the gate when none could step, and lst[par]
        // was not written: nothing to do.  (A replica that could step keeps
        // n_run >= 1 until its own workgroup here; one that could not is unchanged.)
        if (threadIdx.x == 0 && (a.gate->n_run != 0 || (a.open && a.gate->n_stale != 0))) {
            RepState n = s;
            double lam = s.lam;
            if (s.ran) {
                double T;
                lf_done(a, p.nf, p.T0, n, v[ACC_KE2], v[MD_NSTAT - 1] != 0.0, T, lam);
                const bool en = a.energy_every > 0 && (s.step % a.energy_every) == 0;
                put_record(a, r, s.step - s.start, en, v, 0.5 * v[ACC_KE2], T, th.xi, th.vxi, 0.0);
            }
            a.th[r].S = lam;
            RepState& st = a.st[r];
            st.age = n.age;
            st.step = n.step;
            st.phase = n.phase;
            st.nb = n.nb;
            st.trip[0] = n.trip[0];
            st.trip[1] = n.trip[1];
            lf_gates(a, s_old, n);
        }
        return;
    }
    // No gate test: when the step kernel left early no replica passes the test below.
    bool run;
    if (a.mode == MODE_FORCE) run = !stale(s);
    else if (a.mode == MODE_LF) run = !stale(s) && s.step < s.target;
    else run = s.phase == 1 && !stale(s);
    if (run && threadIdx.x == 0) {
        const long long kstat = (a.mode == MODE_FORCE) ? 0 : s.step - s.start;
        const bool en = (a.mode == MODE_FORCE) ? (a.force_energy != 0)
                                               : (a.energy_every > 0 && (s.step % a.energy_every) == 0);
        finalize_replica(a, r, p, s, th, kstat, en, v, v[MD_NSTAT - 1] != 0.0);
    }
    if (a.baro && a.mode == MODE_LF) {
        // Berendsen pressure coupling (Berendsen et al., J. Chem. Phys. 81, 3684
        // (1984)), isotropic: P = (sum m v^2 + W) / (3 V) from the step's kinetic
        // energy and virial (every step computes energies under coupling), then the
        // positions just written and the box are scaled by mu = cbrt(1 - k (P0 - P)).
        __shared__ float s_mu;
        if (threadIdx.x == 0) {
            float mu = 1.0f;
            if (run) {
                const double V = (double)p.L[0] * (double)p.L[1] * (double)p.L[2];
                const double P = (v[ACC_KE2] + 0.5 * v[ACC_VIR]) / (3.0 * V) * MD_BAR_PER_KJMOL_A3;
                const double m3 = 1.0 - (double)a.baro_k * ((double)a.p0 - P);
                mu = (float)cbrt(fmax(m3, 0.125));   // at most a halving of the box edge per step
                RepParam& w = a.rep[r];
                for (int d = 0; d < 3; d++) {
                    const float L = (float)((double)p.L[d] * (double)mu);
                    w.L[d] = L;
                    w.Linv[d] = (float)(1.0 / (double)L);   // as the host sets it
                    w.halfL[d] = 0.5f * L;
                    w.cinv[d] = (float)p.nc[d] / L;
                    w.cedge[d] = L / (float)p.nc[d];
                    if (!(a.rl < 0.5f * L)) *a.err = 2;   // the list radius must stay below half the box
                }
            }
            s_mu = mu;
        }
        __syncthreads();
        const float mu = s_mu;
        if (mu != 1.0f) {
            float4* x = a.pos[cur_buf(s) ^ 1] + p.slot_off;   // the buffer this launch wrote
            for (int k = threadIdx.x; k < p.nslot; k += blockDim.x) {
                float4 q = x[k];
                q.x *= mu;
                q.y *= mu;
                q.z *= mu;
                x[k] = q;
            }
        }
    }
}

// Velocity-Verlet begin-of-step kernel: v = S v + F dt/2m; x += v dt; skin
// check (no wrap: positions run free between list builds).  With
// apply_drift = 0 it only applies the end-of-step thermostat scale (run end).
// Each wavefront stores its 64-slot tile's skin check flag; k_kick_finalize
// folds them into the list state and moves the replica to the force phase.
// Two round trips: the gate, the replica index and every per-atom value (both
// position buffers) first, the replica's state and constants second.
__global__ __launch_bounds__(256) void k_vv_kick(const KickArgs a) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= a.nslot) return;   // whole wavefronts (nslot is a multiple of 64)
    const int lane = threadIdx.x & 63;
    const int n_run = a.gate->n_run;
    const int r = __builtin_amdgcn_readfirstlane(a.tile_rep[s >> 6]);
    const int id = a.ids[s];
    float4 v = a.vel[s];
    const float4 f = a.force[s];
    const float4 x0 = a.pos[0][s], x1 = a.pos[1][s];
    const float4 pb = a.pos_build[s];
    loads_issued();
    if (a.apply_drift && n_run == 0) return;   // nothing to kick: idle launch
    const RepState st = a.st[r];
    const RepParam& p = a.rep[r];
    const float L0 = p.L[0], L1 = p.L[1], L2 = p.L[2], Li0 = p.Linv[0], Li1 = p.Linv[1], Li2 = p.Linv[2];
    const float S = (float)(a.apply_drift ? a.th[r].S : a.th[r].S_end);
    loads_issued();
    if (!a.apply_drift) {
        if (id < 0) return;
        v.x *= S; v.y *= S; v.z *= S;
        a.vel[s] = v;
        return;
    }
    if (!(st.phase == 0 && !stale(st) && st.step < st.target)) return;   // wave-uniform
    const int cb = cur_buf(st);
    bool trip = false;
    if (id >= 0) {
        const float h = 0.5f * a.dt * v.w;
        v.x = fmaf(f.x, h, v.x * S);
        v.y = fmaf(f.y, h, v.y * S);
        v.z = fmaf(f.z, h, v.z * S);
        a.vel[s] = v;
        const float4 xi = cb ? x1 : x0;
        const float x = fmaf(v.x, a.dt, xi.x), y = fmaf(v.y, a.dt, xi.y), z = fmaf(v.z, a.dt, xi.z);
        a.pos[cb ^ 1][s] = make_float4(x, y, z, xi.w);
        const float ddx = minimg(x - pb.x, L0, Li0);
        const float ddy = minimg(y - pb.y, L1, Li1);
        const float ddz = minimg(z - pb.z, L2, Li2);
        trip = fmaf(ddz, ddz, fmaf(ddy, ddy, ddx * ddx)) >= a.skin2q;
        if (!(fabsf(x) <= L0 && fabsf(y) <= L1 && fabsf(z) <= L2)) *a.err = 1;
    }
    const unsigned long long tb = __ballot(trip);
    if (lane == 0) a.tflag[s >> 6] = tb ? 1 : 0;
}

// Thermostat bookkeeping at run boundaries (one workgroup).
//   mode 0 (velocity-Verlet run start): begin-of-step half from ke2_cur -> S
//   mode 1 (velocity-Verlet run end):   restore the state after the end-of-step half
__global__ void k_thermo_fix(ThState* th, const RepParam* rep, int nrep, int thermostat, float dt, int mode) {
    for (int r = threadIdx.x; r < nrep; r += blockDim.x) {
        ThState& t = th[r];
        if (mode == 0) {
            if (thermostat == 2) {
                double K = 0.5 * t.ke2_cur, xi = t.xi, vxi = t.vxi;
                double s = nh_half(xi, vxi, K, rep[r].nf, rep[r].kT0, rep[r].Q, dt);
                t.xi = xi;
                t.vxi = vxi;
                t.S = s;
            } else {
                t.S = 1.0;
            }
        } else {
            t.xi = t.xi_end;
            t.vxi = t.vxi_end;
            t.S = 1.0;
            t.S_end = 1.0;
        }
    }
}

// After k_vv_kick: per replica (one wavefront each, or with wide = 1 a whole
// workgroup for the few-replica case) OR the tile flags of the skin check, mark
// the list stale when one tripped and move the replica to the force phase.  Gate
// counters change once per workgroup; no early exit before the barriers.
__global__ __launch_bounds__(256) void k_kick_finalize(const KickArgs a, int wide) {
    __shared__ int s_d[4][3];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int r = wide ? (int)blockIdx.x : (int)blockIdx.x * 4 + wave;
    int moved = 0, tripped = 0, nbpar = 0;
    if (r < a.nrep) {
        const RepState st = a.st[r];
        if (st.phase == 0 && !stale(st) && st.step < st.target) {   // the replica kicked
            const RepParam& p = a.rep[r];
            const int t0 = p.slot_off / 64, nt = p.nslot / 64;
            const int k0 = wide ? (int)threadIdx.x : lane, kstep = wide ? 256 : 64;
            int t = 0;
#pragma unroll 4
            for (int k = k0; k < nt; k += kstep) t |= a.tflag[t0 + k];
            tripped = __ballot(t != 0) ? 1 : 0;
            moved = 1;
            nbpar = st.nb & 1;
        }
    }
    if (wide) tripped = __syncthreads_or(tripped);
    const bool owner = wide ? threadIdx.x == 0 : lane == 0;
    if (owner && moved) {
        if (tripped) a.st[r].trip[nbpar] = 1;
        a.st[r].phase = 1;
    }
    if (lane == 0) {
        const bool count = !wide || wave == 0;
        s_d[wave][0] = count ? moved : 0;
        s_d[wave][1] = count ? tripped : 0;
        s_d[wave][2] = count ? moved - tripped : 0;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        int n = 0, nt = 0, nf = 0;
        for (int w = 0; w < 4; w++) { n += s_d[w][0]; nt += s_d[w][1]; nf += s_d[w][2]; }
        if (a.lfgate) n = nt, nf = 0;   // leapfrog-form steps follow: only a tripped replica stops
        if (n) atomicSub(&a.gate->n_run, n);
        if (nt) atomicAdd(&a.gate->n_stale, nt);
        if (nf) atomicAdd(&a.gate->n_force, nf);
    }
}

// Start of a run: target = step + nsteps for every replica; recount the gates.
__global__ void k_set_target(RepState* st, Gate* g, int nrep, long long nsteps) {
    __shared__ int cnt[3];
    if (threadIdx.x < 3) cnt[threadIdx.x] = 0;
    __syncthreads();
    int ns = 0, nr = 0, nf = 0;
    for (int r = threadIdx.x; r < nrep; r += blockDim.x) {
        RepState& s = st[r];
        s.start = s.step;
        s.target = s.step + nsteps;
        const bool st_ = stale(s);
        ns += st_;
        nr += !st_ && s.phase == 0 && s.step < s.target;
        nf += !st_ && s.phase == 1;
    }
    atomicAdd(&cnt[0], ns);
    atomicAdd(&cnt[1], nr);
    atomicAdd(&cnt[2], nf);
    __syncthreads();
    if (threadIdx.x == 0) {
        g->n_stale = cnt[0];
        g->n_run = cnt[1];
        g->n_force = cnt[2];
    }
}

// Run report: one workgroup reduces the replicas' step counters and the
// builds' maxima (see RunReport).
__global__ __launch_bounds__(1024) void k_report(const RepState* st, int nrep, const int* tmax, int ntiles,
                                                 const int* err, RunReport* out) {
    constexpr int NT = 1024, NW = NT / 64, U = 4;
    __shared__ long long s_step[NW];
    __shared__ int s_row[NW], s_set[NW];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    long long mn = 0x7fffffffffffffffLL;
    int row = 0, set = 0;
    for (int r = tid; r < nrep; r += NT) mn = min(mn, st[r].step);
    // U independent loads of each array in flight per thread
    for (int t0 = tid; t0 < ntiles; t0 += U * NT) {
        int rv[U], sv[U];
#pragma unroll
        for (int u = 0; u < U; u++) {
            const int t = t0 + u * NT;
            rv[u] = t < ntiles ? tmax[t] : 0;
            sv[u] = t < ntiles ? tmax[ntiles + t] : 0;
        }
#pragma unroll
        for (int u = 0; u < U; u++) {
            row = max(row, rv[u]);
            set = max(set, sv[u]);
        }
    }
    for (int o = 32; o >= 1; o >>= 1) {
        mn = min(mn, (long long)__shfl_xor(mn, o, 64));
        row = max(row, __shfl_xor(row, o, 64));
        set = max(set, __shfl_xor(set, o, 64));
    }
    if (lane == 0) {
        s_step[wave] = mn;
        s_row[wave] = row;
        s_set[wave] = set;
    }
    __syncthreads();
    if (tid == 0) {
        RunReport rep;
        rep.min_step = s_step[0];
        rep.rowmax = s_row[0];
        rep.setmax = s_set[0];
        for (int w = 1; w < NW; w++) {
            rep.min_step = min(rep.min_step, s_step[w]);
            rep.rowmax = max(rep.rowmax, s_row[w]);
            rep.setmax = max(rep.setmax, s_set[w]);
        }
        rep.err = *err;
        rep.pad = 0;
        *out = rep;
    }
}

// Copy of the engine state a run may have to be rolled back to (one launch;
// the caller passes matching buffer pairs).  nbytes per pair, 16-byte units.
struct CopyPlan {
    const float4* src[12];
    float4* dst[12];
    long long n16[12];   // 16-byte units per pair
    int npair;
};
__global__ __launch_bounds__(256) void k_copy_plan(const CopyPlan c) {
    const long long stride = (long long)gridDim.x * blockDim.x;
    for (int k = 0; k < c.npair; k++) {
        const float4* __restrict__ s = c.src[k];
        float4* __restrict__ d = c.dst[k];
        for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < c.n16[k]; i += stride) d[i] = s[i];
    }
}

// ---------------------------------------------------------------- launchers

// With ev0/ev1 the launch records the kernel's own start and end in the two events
// (hipExtLaunchKernelGGL), so their elapsed time is the kernel's execution time.
template <typename K>
static void launch_k(K kern, const StepArgs& a, int grid, int threads, size_t shmem, hipStream_t st, hipEvent_t ev0,
                     hipEvent_t ev1) {
    if (ev0) hipExtLaunchKernelGGL(kern, dim3(grid), dim3(threads), shmem, st, ev0, ev1, 0, a);
    else hipLaunchKernelGGL(kern, dim3(grid), dim3(threads), shmem, st, a);
}

template <int G, int EM, bool NT2>
static void launch_ge(const StepArgs& a, int coul, size_t shmem, hipStream_t st, hipEvent_t ev0, hipEvent_t ev1) {
    const int grid = a.ntiles;
    switch (coul) {
    case 0: launch_k(k_step<G, 0, EM, NT2>, a, grid, 64 * G, shmem, st, ev0, ev1); break;
    case 1: launch_k(k_step<G, 1, EM, NT2>, a, grid, 64 * G, shmem, st, ev0, ev1); break;
    case 2: launch_k(k_step<G, 2, EM, NT2>, a, grid, 64 * G, shmem, st, ev0, ev1); break;
    default: launch_k(k_step<G, 3, EM, NT2>, a, grid, 64 * G, shmem, st, ev0, ev1); break;
    }
}

template <int G>
static void launch_g(const StepArgs& a, int coul, size_t shmem, hipStream_t st, hipEvent_t ev0, hipEvent_t ev1) {
    const bool em = (a.mode == MODE_FORCE) ? a.force_energy != 0 : a.energy_every > 0;
    const bool nt2 = a.ntype <= 2;
    if (em) {
        if (nt2) launch_ge<G, 1, true>(a, coul, shmem, st, ev0, ev1);
        else launch_ge<G, 1, false>(a, coul, shmem, st, ev0, ev1);
    } else {
        if (nt2) launch_ge<G, 0, true>(a, coul, shmem, st, ev0, ev1);
        else launch_ge<G, 0, false>(a, coul, shmem, st, ev0, ev1);
    }
}

// dynamic LDS of the step kernel: the staged set and its dummy
static size_t step_shmem(const StepArgs& a) { return (size_t)(a.jcap + 1) * sizeof(float4); }

extern "C" hipError_t md_launch_step(const StepArgs& a, int lanes, int coul, hipStream_t st, hipEvent_t ev0,
                                     hipEvent_t ev1) {
    const size_t sh = step_shmem(a);
    switch (lanes) {
    case 4: launch_g<4>(a, coul, sh, st, ev0, ev1); break;
    case 8: launch_g<8>(a, coul, sh, st, ev0, ev1); break;
    case 16: launch_g<16>(a, coul, sh, st, ev0, ev1); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

extern "C" hipError_t md_launch_step_finalize(const StepArgs& a, int lanes, hipStream_t st) {
    switch (lanes) {
    case 4: hipLaunchKernelGGL(k_step_finalize<4>, dim3(a.nrep), dim3(256), 0, st, a); break;
    case 8: hipLaunchKernelGGL(k_step_finalize<8>, dim3(a.nrep), dim3(512), 0, st, a); break;
    case 16: hipLaunchKernelGGL(k_step_finalize<16>, dim3(a.nrep), dim3(1024), 0, st, a); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

extern "C" hipError_t md_launch_kick(const KickArgs& a, hipStream_t st) {
    hipLaunchKernelGGL(k_vv_kick, dim3((a.nslot + 255) / 256), dim3(256), 0, st, a);
    if (a.apply_drift) {
        const int wide = a.nrep <= 32 ? 1 : 0;   // few replicas: a workgroup each
        hipLaunchKernelGGL(k_kick_finalize, dim3(wide ? a.nrep : (a.nrep + 3) / 4), dim3(256), 0, st, a, wide);
    }
    return hipGetLastError();
}

extern "C" hipError_t md_launch_thermo_fix(ThState* th, const RepParam* rep, int nrep, int thermostat, float dt,
                                           int mode, hipStream_t st) {
    hipLaunchKernelGGL(k_thermo_fix, dim3(1), dim3(256), 0, st, th, rep, nrep, thermostat, dt, mode);
    return hipGetLastError();
}

extern "C" hipError_t md_launch_set_target(RepState* s, Gate* g, int nrep, long long nsteps, hipStream_t st) {
    hipLaunchKernelGGL(k_set_target, dim3(1), dim3(256), 0, st, s, g, nrep, nsteps);
    return hipGetLastError();
}

// Sums over every atom of the force and of the momentum m_int v (vel.w holds
// 1 / m_int): one workgroup, each thread a strided share of the slots, then a
// tree in LDS (fixed order: the same bits every call).
__global__ __launch_bounds__(1024) void k_state_sums(const float4* __restrict__ vel, const float4* __restrict__ force,
                                                     const int* __restrict__ ids, int nslot, double* out) {
    __shared__ double s[6][1024];
    double a[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    for (int k = threadIdx.x; k < nslot; k += blockDim.x) {
        if (ids[k] < 0) continue;
        const float4 f = force[k], v = vel[k];
        const double m = 1.0 / (double)v.w;
        a[0] += f.x; a[1] += f.y; a[2] += f.z;
        a[3] += m * v.x; a[4] += m * v.y; a[5] += m * v.z;
    }
    for (int c = 0; c < 6; c++) s[c][threadIdx.x] = a[c];
    __syncthreads();
    for (int h = blockDim.x / 2; h > 0; h >>= 1) {
        if ((int)threadIdx.x < h)
            for (int c = 0; c < 6; c++) s[c][threadIdx.x] += s[c][threadIdx.x + h];
        __syncthreads();
    }
    if (threadIdx.x < 6) out[threadIdx.x] = s[threadIdx.x][0];
}

extern "C" hipError_t md_launch_state_sums(const float4* vel, const float4* force, const int* ids, int nslot, double* out,
                                           hipStream_t st) {
    hipLaunchKernelGGL(k_state_sums, dim3(1), dim3(1024), 0, st, vel, force, ids, nslot, out);
    return hipGetLastError();
}

extern "C" hipError_t md_launch_report(const RepState* st, int nrep, const int* tmax, int ntiles, const int* err,
                                       RunReport* out, hipStream_t stream) {
    hipLaunchKernelGGL(k_report, dim3(1), dim3(1024), 0, stream, st, nrep, tmax, ntiles, err, out);
    return hipGetLastError();
}

// copies: npair (src, dst, bytes) with 16-byte aligned sizes
extern "C" hipError_t md_launch_copies(int npair, const void* const* src, void* const* dst, const long long* bytes,
                                       hipStream_t st) {
    if (npair < 0 || npair > 12) return hipErrorInvalidValue;
    CopyPlan c{};
    long long most = 0;
    for (int k = 0; k < npair; k++) {
        if (bytes[k] % 16) return hipErrorInvalidValue;
        c.src[k] = static_cast<const float4*>(src[k]);
        c.dst[k] = static_cast<float4*>(dst[k]);
        c.n16[k] = bytes[k] / 16;
        most = std::max(most, c.n16[k]);
    }
    c.npair = npair;
    const int grid = (int)std::max(1LL, std::min(2048LL, (most + 255) / 256));
    hipLaunchKernelGGL(k_copy_plan, dim3(grid), dim3(256), 0, st, c);
    return hipGetLastError();
}
