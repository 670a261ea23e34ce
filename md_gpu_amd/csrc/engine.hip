// engine.hip -- host side of the MD step engine and its C ABI (md_api.h).
//
// One engine = one HIP stream on one device, holding R >= 1 replicas that
// share a force field.  A single system is R = 1.  The reference's per-step
// host round trips (skin check flag, KE partials: src/neighborlist.cu:127-129,
// src/integrator.cu:152) are replaced by device-side flags and reductions, so
// a run is queued as replays of one captured hipGraph segment and the host
// synchronises only to confirm that every replica reached its step target.
//
// A graph segment is `blocks` x ([rebuild chain] + check_every x [step]), with
//   step (leapfrog)         = k_step(LF)
//   step (velocity Verlet)  = k_vv_kick + k_step(VV)
// Every kernel decides per replica on the device whether it runs (RepState:
// list valid, velocity-Verlet phase, step target), so one captured segment is
// replayed until every replica reaches its target; a replica whose update
// tripped the skin criterion (src/neighborlist.cu:36,157-176) idles until the
// next rebuild chain in the segment.
#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include <hipfft/hipfft.h>

#include "md_device.h"
#include "mdgpu/md_api.h"

extern "C" hipError_t md_launch_rebuild(const BuildArgs& a, int commit, hipStream_t st);
extern "C" hipError_t md_launch_force_trip(RepState* s, Gate* g, int nrep, hipStream_t st);
extern "C" hipError_t md_launch_set_target(RepState* s, Gate* g, int nrep, long long nsteps, hipStream_t st);
extern "C" hipError_t md_launch_step(const StepArgs& a, int lanes, int coul, hipStream_t st, hipEvent_t ev0,
                                     hipEvent_t ev1);
extern "C" hipError_t md_launch_step_finalize(const StepArgs& a, int lanes, hipStream_t st);
extern "C" hipError_t md_launch_kick(const KickArgs& a, hipStream_t st);
extern "C" hipError_t md_launch_thermo_fix(ThState* th, const RepParam* rep, int nrep, int thermostat, float dt,
                                           int mode, hipStream_t st);
extern "C" hipError_t md_launch_fm_adjoint(const StepArgs& a, double* part, double* out, int lanes, int coul, int nt,
                                           hipStream_t st, hipEvent_t ev0, hipEvent_t ev1);
extern "C" hipError_t md_launch_fm_fused(const StepArgs& a, double* part, double* out, int lanes, int coul, int nt,
                                         hipStream_t st, hipEvent_t ev0, hipEvent_t ev1);
extern "C" hipError_t md_launch_copies(int npair, const void* const* src, void* const* dst, const long long* bytes,
                                       hipStream_t st);
extern "C" hipError_t md_launch_check_lists(const BuildArgs& a, int row_g, int* bad, hipStream_t st);
extern "C" hipError_t md_launch_state_sums(const float4* vel, const float4* force, const int* ids, int nslot, double* out,
                                           hipStream_t st);
extern "C" hipError_t md_launch_persist(const PersistArgs& a, int lanes, size_t shmem, hipStream_t st, hipEvent_t ev0,
                                        hipEvent_t ev1);
extern "C" hipError_t md_launch_persist_clear(unsigned* flags, hipStream_t st);
extern "C" hipError_t md_launch_settle(const SettleArgs& a, hipStream_t st);
extern "C" int md_launch_pme(const PmeArgs& a, hipfftHandle r2c, hipfftHandle c2r, hipStream_t st);
extern "C" hipError_t md_launch_report(const RepState* st, int nrep, const int* tmax, int ntiles, const int* err,
                                       RunReport* out, hipStream_t stream);

namespace {

thread_local std::string g_err;

int fail(int code, const char* fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_err = buf;
    return code;
}

#define HIPCHECK(x)                                                                              \
    do {                                                                                         \
        hipError_t e_ = (x);                                                                     \
        if (e_ != hipSuccess) return fail(MD_E_HIP, "%s:%d %s: %s", __FILE__, __LINE__, #x,     \
                                          hipGetErrorString(e_));                                \
    } while (0)

#define CHECK(x)                 \
    do {                         \
        int rc_ = (x);           \
        if (rc_ != MD_OK) return rc_; \
    } while (0)

const int kSegment = 96;         // steps queued per full graph segment (blocks x check_every)
const double kPi = 3.14159265358979323846;

// Device buffer of n elements; the allocation is a multiple of 16 bytes (so a
// buffer can be copied in 16-byte units) and zero when alloc returns.  The
// zero fill runs on the null stream and is waited for here: the engine's own
// stream is non-blocking and does not order itself after null-stream work, so
// a fill still in flight could land after the engine's first kernels.
template <class T>
struct DBuf {
    T* p = nullptr;
    size_t n = 0;
    int alloc(size_t count) {
        release();
        if (count == 0) count = 1;
        const size_t nb = (count * sizeof(T) + 15) / 16 * 16;
        hipError_t e = hipMalloc((void**)&p, nb);
        if (e != hipSuccess) {
            p = nullptr;
            return fail(MD_E_HIP, "hipMalloc(%zu bytes): %s", nb, hipGetErrorString(e));
        }
        n = count;
        e = hipMemset(p, 0, nb);
        if (e == hipSuccess) e = hipStreamSynchronize(nullptr);
        if (e != hipSuccess) return fail(MD_E_HIP, "hipMemset: %s", hipGetErrorString(e));
        return MD_OK;
    }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
    }
    size_t bytes() const { return (n * sizeof(T) + 15) / 16 * 16; }
};

// Skilling's transpose form of the Hilbert index ("Programming the Hilbert
// curve", AIP Conf. Proc. 707, 381 (2004)), 3 axes of `bits` bits each.
uint64_t hilbert3(unsigned x, unsigned y, unsigned z, int bits) {
    unsigned X[3] = {x, y, z};
    unsigned M = 1u << (bits - 1);
    for (unsigned Q = M; Q > 1; Q >>= 1) {
        unsigned P = Q - 1;
        for (int i = 0; i < 3; i++) {
            if (X[i] & Q) {
                X[0] ^= P;
            } else {
                unsigned t = (X[0] ^ X[i]) & P;
                X[0] ^= t;
                X[i] ^= t;
            }
        }
    }
    for (int i = 1; i < 3; i++) X[i] ^= X[i - 1];
    unsigned t = 0;
    for (unsigned Q = M; Q > 1; Q >>= 1)
        if (X[2] & Q) t ^= Q - 1;
    for (int i = 0; i < 3; i++) X[i] ^= t;
    uint64_t h = 0;
    for (int b = bits - 1; b >= 0; b--)
        for (int i = 0; i < 3; i++) h = (h << 1) | ((X[i] >> b) & 1u);
    return h;
}

double as_poly(double t) {   // A&S 7.1.26 polynomial, src/simulation.h:39-43
    return ((((1.061405429 * t - 1.453152027) * t + 1.421413741) * t - 0.284496736) * t + 0.254829592) * t;
}

// A(r) = (2 alpha / sqrt(pi)) r exp(-alpha^2 r^2) + erfc_AS(alpha r) and dA/dr, in FP64,
// with the reference's erfc (src/pair_forces.cu:174-183: E = poly(t) exp(-alpha^2 r^2),
// t = 1 / (1 + p alpha r))
void wolf_A(double r, double a, double* A, double* dA) {
    const double p = 0.3275911, sp = 2.0 / std::sqrt(kPi) * a;
    const double t = 1.0 / (1.0 + p * a * r);
    const double poly = as_poly(t);
    const double dpoly = ((((5 * 1.061405429 * t - 4 * 1.453152027) * t + 3 * 1.421413741) * t -
                           2 * 0.284496736) * t + 0.254829592);   // d poly / dt
    const double ex = std::exp(-a * a * r * r);
    *A = (sp * r + poly) * ex;
    *dA = (sp + dpoly * (-p * a * t * t)) * ex + (sp * r + poly) * ex * (-2.0 * a * a * r);
}

// Cubic Hermite table of h(r) = A(r) - r A(rc)/rc (Wolf) or A(r) (damped) on MD_CTAB
// equal intervals of [0, rc].  The pair force factor is then qq h(r) / r^3, the
// same law as src/pair_forces.cu:174-186 evaluated in FP64 at the nodes; the cubic
// reproduces it to ~2e-7 relative in FP32.
std::vector<float4> coulomb_table(double a, double rc, bool shifted) {
    double Arc, dArc;
    wolf_A(rc, a, &Arc, &dArc);
    const double w = shifted ? Arc / rc : 0.0, h = rc / MD_CTAB;
    std::vector<double> v(MD_CTAB + 1), d(MD_CTAB + 1);
    for (int k = 0; k <= MD_CTAB; k++) {
        double A, dA;
        wolf_A(k * h, a, &A, &dA);
        v[k] = A - w * k * h;
        d[k] = (dA - w) * h;   // derivative per unit of the local coordinate
    }
    std::vector<float4> t(MD_CTAB + 1);
    for (int k = 0; k < MD_CTAB; k++) {
        const double c2 = 3.0 * (v[k + 1] - v[k]) - 2.0 * d[k] - d[k + 1];
        const double c3 = 2.0 * (v[k] - v[k + 1]) + d[k] + d[k + 1];
        t[k] = make_float4((float)v[k], (float)d[k], (float)c2, (float)c3);
    }
    t[MD_CTAB] = make_float4((float)v[MD_CTAB], 0.0f, 0.0f, 0.0f);   // h(rc), for r = rc exactly
    return t;
}

}  // namespace

struct md_engine {
    int dev = 0;
    hipStream_t st = nullptr;
    int ncu = 256;
    std::string arch;

    bool has_ff = false, has_integ = false, has_sys = false, has_state = false;
    md_forcefield ff{};
    md_integrator in{};
    float rc2 = 0, rl2 = 0, skin2q = 0;
    float palfa = 0, spiialfa = 0, alfasq = 0, wolf = 0, reacf = 0, ecut = 0, log2rc = 0;
    int coul_kernel = MD_COUL_NONE;

    int R = 0, N = 0, S = 0, ncell_total = 0;
    std::vector<int> natom, atom_off, slot_off;
    std::vector<float> box, T0;
    std::vector<int> type_h;
    std::vector<float> charge_h, mass_h;
    std::vector<RepParam> rep_h;
    std::vector<int> tile_rep_h;
    int nb = 0, na = 0;
    bool bonded = false, has_excl = false;

    DBuf<float4> pos[2], vel, force, pos_build, pos_s, vel_s, force_s, target;
    DBuf<int> ids, typ, ids_s, typ_s, slot_of, key, rank, tmp, tmpid;
    DBuf<int> cell_count, cell_start, cell_rank;
    DBuf<uint16_t> nl16;           // [S][cap] rows of local indices into the tile sets
    DBuf<uint32_t> jset;           // [S/64][jcap] tile neighbor sets
    DBuf<int> ncount, jcount;
    DBuf<uint8_t> rperm;           // [S] per tile: atoms by decreasing row length
    DBuf<int> tmax;                // [2][S/64] builds' longest row / largest set per tile (running maxima)
    DBuf<int> tflag;               // [S/64] skin-check flag per tile (kick)
    DBuf<RunReport> report;        // run progress for the host
    DBuf<int> badcnt;              // list invariant counters (md_check_lists)
    DBuf<int4> excl, ang_list;
    std::vector<int4> excl_bond;   // host copy of the 1-2 exclusions (md_set_topology)
    int nrigid = 0;                // rigid 3-site molecules held by SETTLE (settle.hip)
    float rig_oh = 0, rig_hh = 0;  // their O-H and H-H distances
    std::vector<int> rig_h;        // host copy: nrigid x (O, H1, H2)
    DBuf<int4> rig_mol;            // [nrigid] (O, H1, H2, 0)
    DBuf<int> rig_of;              // [N] molecule of an O atom, -2 for an H atom, -1 unconstrained
    DBuf<uint32_t> bond_list;
    DBuf<int> nbond, nang;
    DBuf<RepParam> rep, snap_rep;
    DBuf<double> sums;             // [6] force and momentum sums (md_get_sums)
    DBuf<int> tile_rep;
    DBuf<RepState> rst;
    DBuf<Gate> gate;
    DBuf<ThState> th;
    DBuf<double> part, fm_acc;     // part: [2][MD_NSTAT][S/64] partial sums of the step kernel, one per tile
    DBuf<double> fm_part;          // [S/64][2 + 2 nt^2] per-tile slots of the force-matching kernels
    DBuf<float4> fm_ljtab;         // LJ table of the trial parameters (force matching)
    float4 fm_ljtab_h[64];         // its host staging copy (alive until the next evaluation)
    DBuf<RepState> lst;            // [2][R] hand-over state of leapfrog blocks with deferred finalize
    DBuf<int> err;
    long long reb_base = 0;        // list builds of states uploaded before the current one
    DBuf<StatRec> stats;
    int stat_cap = 0;
    DBuf<float4> ljtab, ctab;
    // snapshot for overflow recovery
    DBuf<float4> snap_pos0, snap_pos1, snap_vel, snap_force;
    DBuf<int> snap_ids, snap_typ, snap_slot_of;
    DBuf<ThState> snap_th;
    DBuf<RepState> snap_rst;
    DBuf<Gate> snap_gate;
    int cap = 0;                   // row capacity, entries per atom (a multiple of 128)
    int jcap = 0;                  // tile set capacity, entries per tile
    int jcap_max = 0;              // largest set the step kernel's LDS holds on this device

    long long step = 0;          // steps completed (all replicas advance together between calls)
    int lanes = 8, grid = 0, cell_div = 2, check_every = 8, build_lanes = 4, submit = 0, refresh = 0, build_kernel = 0;
    int defer = 1;                 // leapfrog: finalize each step inside the next step kernel of the block
    int vv_fuse = 1;               // NVE velocity Verlet in leapfrog form (one step kernel per step)
    int fm_fused = 1;              // force matching without bonded terms: one pass for loss and gradient
    int persist = 0;               // small single systems: all steps of a run in one launch (persist.hip)
    bool pk_off = false;           // the persistent kernel's rows do not fit its LDS for this system
    int pk_cap = 0;                // its row entries per lane (0: from the density and pk_req)
    int pk_req = 0;                // list capacity the caller set (entries per atom): bounds the lane rows
    long long pk_runs = 0, pk_replays = 0;   // runs on the persistent kernel, replays after a row overflow
    size_t lds_max = 65536;        // LDS per workgroup on this device
    DBuf<unsigned> pk_flags;       // [2][MD_PK_WG] barrier flags
    DBuf<double> pk_pub;           // [2][MD_PK_WG][4] partial sums published at the barriers
    static constexpr double kQueueRatio0 = 1.15;
    double queue_ratio = kQueueRatio0;   // launched / completed steps, learned per launch configuration
    int fm_nframes = 0;
    // smooth PME (pme.hip): requested grid and order (order 0: none); the buffers,
    // tables and FFT plans are made for the current slots and box by pme_prepare
    int pme_order = 0, pme_K[3] = {0, 0, 0};
    bool pme_on = false;           // prepared: the step chain includes the PME kernels
    bool vv_force_stale = false;   // the force array (velocity Verlet's first half kick) predates a PME change
    int pme_S = 0;                 // slot count the buffers were made for
    double pme_eself = 0, pme_qsum2 = 0;   // self term; (sum q)^2 for the background term
    int pme_nconv = 0, pme_ngather = 0;
    DBuf<unsigned long long> pme_qacc;
    DBuf<float> pme_qgrid;
    DBuf<float2> pme_sgrid, pme_tab;
    DBuf<float4> pme_frec;
    DBuf<double> pme_part, pme_out;
    DBuf<char> pme_work;
    hipfftHandle pme_r2c = nullptr, pme_c2r = nullptr;
    // segment graphs by key (check interval, energy interval, blocks); two
    // executable copies each, launched alternately, so the next replay can be
    // submitted while the previous one still runs
    struct SegGraph { hipGraphExec_t ex[2]; int next; };
    std::map<std::string, SegGraph> graphs;

    bool timing = false;
    double t_step_ms = 0, t_reb_ms = 0;
    long long n_step = 0, n_reb = 0;
    std::vector<hipEvent_t> ev_pool;

    ~md_engine() {
        drop_graphs();
        drop_pme_plans();
        for (auto e : ev_pool) (void)hipEventDestroy(e);
        if (st) (void)hipStreamDestroy(st);
    }
    void drop_graphs() {
        for (auto& kv : graphs)
            for (auto x : kv.second.ex) (void)hipGraphExecDestroy(x);
        graphs.clear();
    }
    int ntiles() const { return S / MD_TILE; }
    void drop_pme_plans() {
        if (pme_r2c) (void)hipfftDestroy(pme_r2c);
        if (pme_c2r) (void)hipfftDestroy(pme_c2r);
        pme_r2c = pme_c2r = nullptr;
    }
};

namespace {

int check_ready(md_engine* e, bool need_state) {
    if (!e) return fail(MD_E_INVALID, "null engine");
    if (!e->has_ff) return fail(MD_E_STATE, "force field not set");
    if (!e->has_integ) return fail(MD_E_STATE, "integrator not set");
    if (!e->has_sys) return fail(MD_E_STATE, "system not set");
    if (need_state && !e->has_state) return fail(MD_E_STATE, "state not uploaded");
    return MD_OK;
}

// ---- derived constants -------------------------------------------------
void derive_ff(md_engine* e) {
    const md_forcefield& f = e->ff;
    float rc = f.rc;
    e->rc2 = rc * rc;                                                  // src/simulation.cu:96
    double rl = (double)f.skin + std::sqrt((double)e->rc2);
    e->rl2 = (float)(rl * rl);                                         // src/neighborlist.cu:35
    e->skin2q = f.skin * f.skin / 4.0f;                                // src/neighborlist.cu:36
    double a = f.alpha, rcd = f.rc;
    e->palfa = (float)(a * 0.3275911);                                 // src/forces.cu:98
    e->spiialfa = (float)(2.0 / std::sqrt(kPi) * a);                   // :100
    e->alfasq = (float)(-a * a / std::log(2.0));                       // :102
    e->wolf = 0.f;
    e->ecut = 0.f;
    e->reacf = 0.f;
    e->log2rc = (float)std::log2(rcd);
    if (f.coulomb == MD_COUL_WOLF) {                                   // :104-109 (in FP64 here)
        double t = 1.0 / (1.0 + a * 0.3275911 * rcd);
        double ex = std::exp(-a * a * rcd * rcd);
        e->wolf = (float)((2.0 / std::sqrt(kPi) * a * rcd + as_poly(t)) * ex / rcd);
        e->ecut = (float)(as_poly(t) * ex / rcd);
    } else if (f.coulomb == MD_COUL_RF) {                              // src/simulation.cu:86
        double kr = 2.0 * ((double)f.eps_rf - 1.0) / (2.0 * (double)f.eps_rf + 1.0) / (rcd * rcd * rcd);
        e->reacf = (float)kr;
        e->ecut = (float)(1.0 / rcd + 0.5 * kr * rcd * rcd);
    }
}

// LJ table of the step kernel: (lj1, lj2, energy shift at rc, 0) per type pair
void make_ljtab(const md_engine* e, const float* lj1, const float* lj2, float4* t) {
    const int nt = e->ff.ntype;
    for (int k = 0; k < 64; k++) t[k] = make_float4(0, 0, 0, 0);
    double rc2 = (double)e->ff.rc * e->ff.rc;
    for (int i = 0; i < nt; i++)
        for (int j = 0; j < nt; j++) {
            double l1 = lj1[i * nt + j], l2 = lj2[i * nt + j];
            double ri6 = 1.0 / (rc2 * rc2 * rc2);
            double shift = l1 / 12.0 * ri6 * ri6 - l2 / 6.0 * ri6;
            t[i * 8 + j] = make_float4((float)l1, (float)l2, (float)shift, 0.f);
        }
}

int upload_ljtab(md_engine* e, const float* lj1, const float* lj2) {
    std::vector<float4> t(64);
    make_ljtab(e, lj1, lj2, t.data());
    HIPCHECK(hipMemcpyAsync(e->ljtab.p, t.data(), 64 * sizeof(float4), hipMemcpyHostToDevice, e->st));
    HIPCHECK(hipStreamSynchronize(e->st));
    return MD_OK;
}

// ---- exclusions and rigid molecules ------------------------------------
// The device exclusion list: the 1-2 bonded partners of md_set_topology and the
// other atoms of each rigid molecule (at most 4 per atom).
int upload_exclusions(md_engine* e) {
    const int N = e->N;
    std::vector<int4> ex = e->excl_bond;
    ex.resize(N, make_int4(-1, -1, -1, -1));
    for (int m = 0; m < e->nrigid; m++) {
        const int* a = &e->rig_h[3 * (size_t)m];
        for (int u = 0; u < 3; u++)
            for (int v = 0; v < 3; v++) {
                if (u == v) continue;
                int* sl = &ex[a[u]].x;
                bool have = false;
                int k = 0;
                for (; k < 4 && sl[k] >= 0; k++) have |= sl[k] == a[v];
                if (have) continue;
                if (k == 4) return fail(MD_E_CAPACITY, "atom %d has more than 4 excluded partners", a[u]);
                sl[k] = a[v];
            }
    }
    bool any = false;
    for (const int4& q : ex) any |= q.x >= 0;
    HIPCHECK(hipMemcpy(e->excl.p, ex.data(), N * sizeof(int4), hipMemcpyHostToDevice));
    e->has_excl = any;
    return MD_OK;
}

// Temperature degrees of freedom: 3N - 3 per replica (src/integrator.cu:155),
// less 3 per rigid molecule of the replica.
int upload_dof(md_engine* e) {
    std::vector<int> nmol(e->R, 0);
    for (int m = 0; m < e->nrigid; m++) {
        const int o = e->rig_h[3 * (size_t)m];
        const int r = (int)(std::upper_bound(e->atom_off.begin(), e->atom_off.end(), o) - e->atom_off.begin()) - 1;
        nmol[r]++;
    }
    for (int r = 0; r < e->R; r++) {
        RepParam& p = e->rep_h[r];
        p.nf = (float)(3.0 * e->natom[r] - 3.0 - 3.0 * nmol[r]);
        p.Q = (double)p.nf * p.kT0 * (double)e->in.tau * e->in.tau;
    }
    HIPCHECK(hipMemcpy(e->rep.p, e->rep_h.data(), e->R * sizeof(RepParam), hipMemcpyHostToDevice));
    return MD_OK;
}

// ---- cells --------------------------------------------------------------
int setup_cells(md_engine* e) {
    double rl = std::sqrt((double)e->rl2);
    std::vector<int> ranks;
    int off = 0;
    for (int r = 0; r < e->R; r++) {
        RepParam& p = e->rep_h[r];
        int nmax = 1;
        for (int d = 0; d < 3; d++) {
            double L = p.L[d];
            if (rl >= 0.5 * L) return fail(MD_E_INVALID, "replica %d: list radius %.4f must be below half the box (%.4f)", r, rl, L);
            int nc = (int)std::floor(L * e->cell_div / rl);
            nc = std::max(1, std::min(nc, 1024));
            p.nc[d] = nc;
            p.cedge[d] = (float)(L / nc);
            p.cinv[d] = (float)(nc / L);
            nmax = std::max(nmax, nc);
        }
        int bits = 1;
        while ((1 << bits) < nmax) bits++;
        int ncell = p.nc[0] * p.nc[1] * p.nc[2];
        std::vector<std::pair<uint64_t, int>> h(ncell);
        for (int cz = 0; cz < p.nc[2]; cz++)
            for (int cy = 0; cy < p.nc[1]; cy++)
                for (int cx = 0; cx < p.nc[0]; cx++) {
                    int lin = cx + p.nc[0] * (cy + p.nc[1] * cz);
                    h[lin] = {hilbert3(cx, cy, cz, bits), lin};
                }
        std::sort(h.begin(), h.end());
        ranks.resize(off + ncell);
        for (int k = 0; k < ncell; k++) ranks[off + h[k].second] = k;
        p.cell_off = off;
        p.cs_off = off + r;
        off += ncell;
    }
    e->ncell_total = off;
    CHECK(e->cell_rank.alloc(off));
    CHECK(e->cell_count.alloc(off));
    CHECK(e->cell_start.alloc(off + e->R));
    HIPCHECK(hipMemcpy(e->cell_rank.p, ranks.data(), off * sizeof(int), hipMemcpyHostToDevice));
    HIPCHECK(hipMemcpy(e->rep.p, e->rep_h.data(), e->R * sizeof(RepParam), hipMemcpyHostToDevice));
    e->drop_graphs();
    return MD_OK;
}

double max_density(const md_engine* e) {
    double dens = 0;
    for (int r = 0; r < e->R; r++) {
        double V = (double)e->box[3 * r] * e->box[3 * r + 1] * e->box[3 * r + 2];
        dens = std::max(dens, e->natom[r] / V);
    }
    return dens;
}

// row capacity: the atoms within r_list at the densest replica, with margin
int estimate_cap(md_engine* e) {
    double rl = std::sqrt((double)e->rl2);
    double k = max_density(e) * 4.0 / 3.0 * kPi * rl * rl * rl;
    int c = (int)(k * 1.4 + 64);
    c = std::min(c, e->N);
    return std::max(32, (c + 31) / 32 * 32);
}

// tile set capacity: the atoms within r_list of a 64-atom tile (a cube of the
// tile's volume grown by r_list: Steiner's formula for the parallel body)
int estimate_jcap(md_engine* e) {
    double rl = std::sqrt((double)e->rl2), dens = max_density(e);
    double a = std::cbrt(64.0 / std::max(dens, 1e-12));
    double V = a * a * a + 6 * a * a * rl + 3 * kPi * a * rl * rl + 4.0 / 3.0 * kPi * rl * rl * rl;
    int c = (int)(V * dens * 1.3 + 128);
    return (c + 63) / 64 * 64;
}

int alloc_list(md_engine* e, int cap, int jcap) {
    // a multiple of 128 entries = 8G for every G: rows hold whole blocks (lpos16) and stay 256-byte aligned
    cap = (std::max(cap, 128) + 127) / 128 * 128;
    jcap = (std::max(jcap, 64) + 63) / 64 * 64;
    if (jcap > e->jcap_max)
        return fail(MD_E_CAPACITY, "a 64-atom tile has %d neighbors; the step kernel's LDS holds %d on this device",
                    jcap, e->jcap_max);
    if (cap > 65535) return fail(MD_E_CAPACITY, "rows of %d entries exceed the 16-bit set index", cap);
    e->cap = cap;
    e->jcap = jcap;
    CHECK(e->nl16.alloc((size_t)e->S * cap));
    CHECK(e->jset.alloc((size_t)e->ntiles() * jcap));
    e->drop_graphs();
    return MD_OK;
}

// grow the capacities after a build reported rows of `row` and sets of `set` entries
int grow_list(md_engine* e, int row, int set) {
    const int nc = row > e->cap ? (int)(row * 1.25) / 32 * 32 + 64 : e->cap;
    const int nj = set > e->jcap ? std::min((int)(set * 1.2) / 64 * 64 + 128, std::max(set, e->jcap_max)) : e->jcap;
    return alloc_list(e, nc, nj);
}

int ensure_stats(md_engine* e, int nrec) {
    nrec = std::max(nrec, 1);
    if (e->stat_cap < nrec) {
        CHECK(e->stats.alloc((size_t)nrec * e->R));
        e->stat_cap = nrec;
        e->drop_graphs();
    }
    return MD_OK;
}

// NVE velocity Verlet runs in leapfrog form (one step kernel per step)
bool vv_fused(const md_engine* e) {
    return e->in.integrator == MD_VVERLET && e->in.thermostat == MD_TH_NONE && e->vv_fuse;
}
// the step kernel runs in leapfrog form: leapfrog, or fused velocity Verlet
bool lf_form(const md_engine* e) { return e->in.integrator == MD_LEAPFROG || vv_fused(e); }

BuildArgs build_args(md_engine* e) {
    BuildArgs b{};
    b.pos[0] = e->pos[0].p; b.pos[1] = e->pos[1].p; b.vel = e->vel.p; b.force = e->force.p;
    b.ids = e->ids.p; b.typ = e->typ.p;
    b.pos_s = e->pos_s.p; b.vel_s = e->vel_s.p; b.force_s = e->force_s.p; b.ids_s = e->ids_s.p; b.typ_s = e->typ_s.p;
    b.pos_build = e->pos_build.p;
    b.key = e->key.p; b.rank = e->rank.p; b.tmp = e->tmp.p; b.tmpid = e->tmpid.p;
    b.cell_count = e->cell_count.p; b.cell_start = e->cell_start.p; b.cell_rank = e->cell_rank.p;
    b.slot_of = e->slot_of.p; b.nl16 = e->nl16.p; b.ncount = e->ncount.p; b.rperm = e->rperm.p;
    b.jset = e->jset.p; b.jcount = e->jcount.p;
    b.tilemax = e->tmax.p; b.jmax = e->tmax.p + e->ntiles();
    b.excl = e->excl.p; b.st = e->rst.p; b.gate = e->gate.p;
    b.tile_rep = e->tile_rep.p; b.rep = e->rep.p; b.repw = e->rep.p;
    b.nslot = e->S; b.nrep = e->R; b.cap = e->cap; b.jcap = e->jcap; b.rl2 = e->rl2; b.has_excl = e->has_excl ? 1 : 0;
    b.row_g = e->lanes;
    b.build_nw = e->build_lanes;
    b.lfgate = lf_form(e) ? 1 : 0;
    return b;
}

StepArgs step_args(md_engine* e, int mode, int energy_every) {
    StepArgs a{};
    a.pos[0] = e->pos[0].p; a.pos[1] = e->pos[1].p; a.vel = e->vel.p; a.force = e->force.p;
    a.ids = e->ids.p; a.typ = e->typ.p; a.nl16 = e->nl16.p; a.ncount = e->ncount.p; a.pos_build = e->pos_build.p;
    a.rperm = e->rperm.p;
    a.jset = e->jset.p; a.jcount = e->jcount.p;
    a.tile_rep = e->tile_rep.p; a.rep = e->rep.p; a.st = e->rst.p; a.gate = e->gate.p; a.th = e->th.p; a.part = e->part.p;
    a.lst = e->lst.p; a.lagmode = 0; a.lag = 0; a.par = 0; a.open = 0;
    a.prefilter = e->R > 1 ? 1 : 0;   // cleared for deferred-finalize blocks (every tile takes part)
    a.vvfused = vv_fused(e) ? 1 : 0;
    a.err = e->err.p; a.stats = e->stats.p; a.stat_cap = e->stat_cap;
    a.ljtab = e->ljtab.p;
    a.ctab = e->ctab.p;
    a.ct_scale = (float)(MD_CTAB / (double)e->ff.rc);
    a.slot_of = e->slot_of.p; a.bond_list = e->bond_list.p; a.nbond = e->nbond.p; a.ang_list = e->ang_list.p; a.nang = e->nang.p;
    a.target = e->target.p;
    a.cap = e->cap; a.jcap = e->jcap; a.ntype = e->ff.ntype; a.nrep = e->R; a.ntiles = e->ntiles(); a.mode = mode;
    a.thermostat = e->in.thermostat; a.bonded = e->bonded ? 1 : 0;
    a.energy_every = energy_every; a.force_energy = 0;
    a.rc2 = e->rc2; a.qfac = e->ff.qfac; a.palfa = e->palfa; a.spiialfa = e->spiialfa; a.alfasq = e->alfasq;
    a.wolf = e->wolf; a.reacf = e->reacf; a.ecut = e->ecut; a.log2rc = e->log2rc; a.qq_scale = 1.0f;
    a.dt = e->in.dt; a.tau = e->in.tau; a.skin2q = e->skin2q;
    a.max_age = e->refresh ? e->check_every : 0;
    a.baro = (mode == MODE_LF && e->in.barostat == MD_BARO_BERENDSEN) ? 1 : 0;
    if (a.baro) a.energy_every = 1;   // the pressure needs the virial of every step
    a.p0 = e->in.press0;
    a.baro_k = (float)((double)e->in.compress * e->in.dt / (double)(e->in.tau_p > 0 ? e->in.tau_p : 1.0f));
    a.rl = std::sqrt(e->rl2);
    a.rc = e->ff.rc;
    a.frec = e->pme_on ? e->pme_frec.p : nullptr;
    a.pme_e = e->pme_on ? e->pme_out.p : nullptr;
    for (int k = 0; k < 8; k++) {
        a.bk2[k] = e->ff.bond_k2[k]; a.br0[k] = e->ff.bond_r0[k];
        a.ak2[k] = e->ff.angle_k2[k]; a.at0[k] = e->ff.angle_t0[k];
    }
    return a;
}

KickArgs kick_args(md_engine* e, int apply_drift) {
    KickArgs k{};
    k.pos[0] = e->pos[0].p; k.pos[1] = e->pos[1].p; k.vel = e->vel.p; k.force = e->force.p; k.ids = e->ids.p;
    k.pos_build = e->pos_build.p; k.tile_rep = e->tile_rep.p; k.rep = e->rep.p; k.st = e->rst.p; k.gate = e->gate.p; k.th = e->th.p;
    k.err = e->err.p; k.tflag = e->tflag.p; k.nslot = e->S; k.nrep = e->R; k.dt = e->in.dt; k.skin2q = e->skin2q; k.apply_drift = apply_drift;
    k.lfgate = lf_form(e) ? 1 : 0;
    return k;
}

hipEvent_t ev_get(md_engine* e, size_t i) {
    while (e->ev_pool.size() <= i) {
        hipEvent_t ev;
        if (hipEventCreate(&ev) != hipSuccess) return nullptr;
        e->ev_pool.push_back(ev);
    }
    return e->ev_pool[i];
}

// HIP-event timing of direct (non-graph) launches
struct TimedSpan {
    md_engine* e;
    std::vector<std::pair<hipEvent_t, hipEvent_t>> steps, rebs;
    size_t used = 2;   // events 0 and 1 are reserved for whole-run timing
    std::pair<hipEvent_t, hipEvent_t> take() {
        hipEvent_t a = ev_get(e, used++), b = ev_get(e, used++);
        return {a, b};
    }
};

// commit = false: the caller queues a deferred-finalize block whose first step
// kernel opens the rebuilt lists
int enqueue_rebuild(md_engine* e, TimedSpan* ts, bool commit = true) {
    BuildArgs b = build_args(e);
    std::pair<hipEvent_t, hipEvent_t> ev{};
    if (ts) { ev = ts->take(); HIPCHECK(hipEventRecord(ev.first, e->st)); }
    HIPCHECK(md_launch_rebuild(b, commit ? 1 : 0, e->st));
    if (ts) { HIPCHECK(hipEventRecord(ev.second, e->st)); ts->rebs.push_back(ev); }
    return MD_OK;
}

SettleArgs settle_args(md_engine* e, int par) {
    SettleArgs g{};
    g.pos[0] = e->pos[0].p; g.pos[1] = e->pos[1].p; g.vel = e->vel.p; g.pos_build = e->pos_build.p;
    g.ids = e->ids.p; g.tile_rep = e->tile_rep.p; g.rep = e->rep.p; g.st = e->rst.p; g.slot_of = e->slot_of.p;
    g.rig_of = e->rig_of.p; g.rig_mol = e->rig_mol.p; g.part = e->part.p; g.err = e->err.p;
    g.par = par; g.ntiles = e->ntiles(); g.dt = e->in.dt; g.skin2q = e->skin2q;
    g.rc_hh = 0.5f * e->rig_hh;
    g.h = (float)std::sqrt((double)e->rig_oh * e->rig_oh - 0.25 * (double)e->rig_hh * e->rig_hh);
    return g;
}

// M_p(w + j), j = 0..p-1 (the recursion of pme.hip, in FP64)
void bspline_host(int p, double w, double* th) {
    double a[MD_PME_MAX_ORDER + 1] = {0};
    a[0] = w;
    a[1] = 1.0 - w;
    for (int k = 2; k < p; k++) {
        double b[MD_PME_MAX_ORDER + 1];
        for (int j = 0; j <= k; j++) b[j] = ((w + j) * (j < k ? a[j] : 0.0) + (k + 1 - w - j) * (j > 0 ? a[j - 1] : 0.0)) / k;
        for (int j = 0; j <= k; j++) a[j] = b[j];
    }
    for (int j = 0; j < p; j++) th[j] = a[j];
}

// Buffers, tables and FFT plans of the requested PME for the current slots and
// box (once; again after md_set_system or md_set_pme).
int pme_prepare(md_engine* e) {
    if (e->pme_order == 0) {
        e->pme_on = false;
        return MD_OK;
    }
    if (e->pme_on) return MD_OK;
    if (e->R != 1) return fail(MD_E_INVALID, "PME runs one replica per engine (this one has %d)", e->R);
    if (e->coul_kernel != MD_COUL_DAMPED)
        return fail(MD_E_INVALID, "PME needs the damped Coulomb term (lkspace 2) and charged atoms");
    if (!(e->ff.alpha > 0.0f)) return fail(MD_E_INVALID, "PME needs a positive splitting parameter alpha");
    const int p = e->pme_order;
    const int* K = e->pme_K;
    for (int d = 0; d < 3; d++)
        if (K[d] < p) return fail(MD_E_INVALID, "PME grid edge %d has %d points, fewer than the order %d", d, K[d], p);
    e->drop_graphs();   // captured segments hold the previous buffers
    e->drop_pme_plans();
    const size_t n = (size_t)K[0] * K[1] * K[2], nh = (size_t)K[0] * K[1] * (K[2] / 2 + 1);
    e->pme_nconv = (int)std::min<size_t>((nh + 255) / 256, 2048);
    e->pme_ngather = (e->S + 63) / 64;   // 4 lanes per slot, 256 lanes per workgroup
    CHECK(e->pme_qacc.alloc(n));
    CHECK(e->pme_qgrid.alloc(n));
    CHECK(e->pme_sgrid.alloc(nh));
    CHECK(e->pme_tab.alloc((size_t)K[0] + K[1] + K[2]));
    CHECK(e->pme_frec.alloc(e->S));
    CHECK(e->pme_part.alloc(2 * (size_t)(e->pme_nconv + e->pme_ngather)));
    CHECK(e->pme_out.alloc(2));
    // per axis: the signed frequency m and |b(m)|^2 (Essmann et al. eq. 4.4); the
    // kernels divide m by the box edge of the step
    const double pi = 3.14159265358979323846;
    double mint[MD_PME_MAX_ORDER];
    bspline_host(p, 0.0, mint);   // M_p(j)
    std::vector<float2> tab;
    for (int d = 0; d < 3; d++)
        for (int m = 0; m < K[d]; m++) {
            double re = 0.0, im = 0.0;
            for (int k = 0; k <= p - 2; k++) {
                re += mint[k + 1] * std::cos(2.0 * pi * m * k / K[d]);
                im += mint[k + 1] * std::sin(2.0 * pi * m * k / K[d]);
            }
            const int ms = m <= K[d] / 2 ? m : m - K[d];
            tab.push_back(make_float2((float)ms, (float)(1.0 / (re * re + im * im))));
        }
    HIPCHECK(hipMemcpy(e->pme_tab.p, tab.data(), tab.size() * sizeof(float2), hipMemcpyHostToDevice));
    // self and neutralizing-background terms
    double s1 = 0.0, s2 = 0.0;
    for (int i = 0; i < e->N; i++) {
        s1 += e->charge_h[i];
        s2 += (double)e->charge_h[i] * e->charge_h[i];
    }
    const double beta = e->ff.alpha;
    e->pme_eself = -(double)e->ff.qfac * beta / std::sqrt(pi) * s2;
    e->pme_qsum2 = s1 * s1;
    // FFT plans on the engine's stream with one work area allocated here, so a
    // captured segment makes no allocation
    size_t ws1 = 0, ws2 = 0;
    auto fft_ok = [](hipfftResult r) { return r == HIPFFT_SUCCESS; };
    if (!fft_ok(hipfftCreate(&e->pme_r2c)) || !fft_ok(hipfftCreate(&e->pme_c2r)))
        return fail(MD_E_HIP, "hipfftCreate failed");
    if (!fft_ok(hipfftSetAutoAllocation(e->pme_r2c, 0)) || !fft_ok(hipfftSetAutoAllocation(e->pme_c2r, 0)))
        return fail(MD_E_HIP, "hipfftSetAutoAllocation failed");
    if (!fft_ok(hipfftMakePlan3d(e->pme_r2c, K[0], K[1], K[2], HIPFFT_R2C, &ws1)) ||
        !fft_ok(hipfftMakePlan3d(e->pme_c2r, K[0], K[1], K[2], HIPFFT_C2R, &ws2)))
        return fail(MD_E_HIP, "hipFFT plan for a %d x %d x %d grid failed", K[0], K[1], K[2]);
    CHECK(e->pme_work.alloc(std::max<size_t>(std::max(ws1, ws2), 16)));
    if (!fft_ok(hipfftSetWorkArea(e->pme_r2c, e->pme_work.p)) || !fft_ok(hipfftSetWorkArea(e->pme_c2r, e->pme_work.p)) ||
        !fft_ok(hipfftSetStream(e->pme_r2c, e->st)) || !fft_ok(hipfftSetStream(e->pme_c2r, e->st)))
        return fail(MD_E_HIP, "hipFFT work area or stream set-up failed");
    e->pme_S = e->S;
    e->pme_on = true;
    return MD_OK;
}

PmeArgs pme_args(md_engine* e, int mode) {
    PmeArgs g{};
    g.pos[0] = e->pos[0].p; g.pos[1] = e->pos[1].p; g.ids = e->ids.p; g.slot_of = e->slot_of.p; g.excl = e->excl.p;
    g.st = e->rst.p; g.qacc = e->pme_qacc.p; g.qgrid = e->pme_qgrid.p; g.sgrid = e->pme_sgrid.p; g.tab = e->pme_tab.p;
    g.frec = e->pme_frec.p; g.part = e->pme_part.p; g.out = e->pme_out.p;
    g.K0 = e->pme_K[0]; g.K1 = e->pme_K[1]; g.K2 = e->pme_K[2]; g.order = e->pme_order;
    g.nslot = e->S; g.mode = mode; g.has_excl = e->has_excl ? 1 : 0;
    g.nconv = e->pme_nconv; g.ngather = e->pme_ngather;
    g.rep = e->rep.p;
    g.ik0 = 1.0 / e->pme_K[0]; g.ik1 = 1.0 / e->pme_K[1]; g.ik2 = 1.0 / e->pme_K[2];
    const double beta = e->ff.alpha, pi = 3.14159265358979323846;
    g.beta = (float)beta;
    g.qfac = e->ff.qfac;
    g.pi2_b2 = (float)(pi * pi / (beta * beta));
    g.e_self = e->pme_eself; g.qsum2 = e->pme_qsum2;
    return g;
}

int enqueue_pme(md_engine* e, int mode) {
    const int rc = md_launch_pme(pme_args(e, mode), e->pme_r2c, e->pme_c2r, e->st);
    if (rc == 2) return fail(MD_E_HIP, "hipFFT execution failed");
    if (rc != 0) return fail(MD_E_HIP, "PME kernel launch failed");
    return MD_OK;
}

int enqueue_stepkernel(md_engine* e, const StepArgs& a, TimedSpan* ts, bool finalize = true) {
    // long-range part first: the step kernel adds its forces and energy
    if (e->pme_on) CHECK(enqueue_pme(e, a.mode));
    std::pair<hipEvent_t, hipEvent_t> ev{};
    if (ts) ev = ts->take();   // the launch records the kernel's own start and end
    HIPCHECK(md_launch_step(a, e->lanes, e->coul_kernel, e->st, ev.first, ev.second));
    if (ts) ts->steps.push_back(ev);
    // rigid molecules: SETTLE on the step's positions, before the finalize reads the slots
    if (e->nrigid > 0 && a.mode == MODE_LF) HIPCHECK(md_launch_settle(settle_args(e, a.par), e->st));
    if (finalize) HIPCHECK(md_launch_step_finalize(a, e->lanes, e->st));
    return MD_OK;
}

// A leapfrog block of n steps with deferred finalize: each step kernel finalizes
// the previous one of the block on its way in; one finalize kernel closes the
// block (the rebuild chain and the host see only closed blocks).
int enqueue_lf_block(md_engine* e, int energy_every, int n) {
    StepArgs a = step_args(e, MODE_LF, energy_every);
    a.lagmode = 1;
    a.open = 1;
    a.prefilter = 0;
    for (int k = 0; k < n; k++) {
        a.lag = k > 0 ? 1 : 0;
        a.par = k & 1;
        CHECK(enqueue_stepkernel(e, a, nullptr, false));
    }
    HIPCHECK(md_launch_step_finalize(a, e->lanes, e->st));
    return MD_OK;
}

// One step launch (the kernels skip replicas that cannot advance).  With
// chain_mid a gated rebuild chain sits between the velocity-Verlet kick and
// the force kernel, so the step always completes (timed launches).
int enqueue_step(md_engine* e, int energy_every, TimedSpan* ts, bool chain_mid = false) {
    if (lf_form(e)) {
        CHECK(enqueue_stepkernel(e, step_args(e, MODE_LF, energy_every), ts));
    } else {
        HIPCHECK(md_launch_kick(kick_args(e, 1), e->st));
        if (chain_mid) CHECK(enqueue_rebuild(e, ts));
        CHECK(enqueue_stepkernel(e, step_args(e, MODE_VV, energy_every), ts));
    }
    return MD_OK;
}

// Copy of the state a run may be rolled back to (restore = false), or back
// (restore = true): one copy kernel for all arrays.
// Pressure coupling changes the boxes on the device: copy them to the host
// parameters (wrapping at download, cell set-up and later uploads use them).
int sync_box(md_engine* e) {
    HIPCHECK(hipMemcpy(e->rep_h.data(), e->rep.p, e->R * sizeof(RepParam), hipMemcpyDeviceToHost));
    for (int r = 0; r < e->R; r++)
        for (int d = 0; d < 3; d++) e->box[3 * r + d] = e->rep_h[r].L[d];
    return MD_OK;
}

// Between runs every replica is at step e->step with phase 0, so its positions
// are in pos[e->step & 1] (the other buffer is overwritten by the next step) and
// only velocity Verlet reads the force array at a run's start (first half
// kick): the snapshot holds what a replay from the run's start reads.
int snapshot(md_engine* e, bool restore) {
    struct Pair { void* a; void* b; size_t n; };
    const int cb = (int)(e->step & 1);
    DBuf<float4>& sp = cb ? e->snap_pos1 : e->snap_pos0;
    std::vector<Pair> list = {
        {sp.p, e->pos[cb].p, e->pos[cb].bytes()},
        {e->snap_vel.p, e->vel.p, e->vel.bytes()},
        {e->snap_ids.p, e->ids.p, e->ids.bytes()}, {e->snap_typ.p, e->typ.p, e->typ.bytes()},
        {e->snap_slot_of.p, e->slot_of.p, e->slot_of.bytes()}, {e->snap_th.p, e->th.p, e->th.bytes()},
        {e->snap_rst.p, e->rst.p, e->rst.bytes()}, {e->snap_gate.p, e->gate.p, e->gate.bytes()},
        {e->snap_rep.p, e->rep.p, e->rep.bytes()}};   // the box, under pressure coupling
    if (e->in.integrator == MD_VVERLET) list.push_back({e->snap_force.p, e->force.p, e->force.bytes()});
    const int np = (int)list.size();
    const void* src[12];
    void* dst[12];
    long long nb[12];
    for (int k = 0; k < np; k++) {
        src[k] = restore ? list[k].a : list[k].b;
        dst[k] = restore ? list[k].b : list[k].a;
        nb[k] = (long long)list[k].n;
    }
    HIPCHECK(md_launch_copies(np, src, dst, nb, e->st));
    return MD_OK;
}

// The builds' longest rows and largest sets per tile accumulate in tmax until cleared.
int clear_rowmax(md_engine* e) {
    HIPCHECK(hipMemsetAsync(e->tmax.p, 0, e->tmax.bytes(), e->st));
    return MD_OK;
}

// Queue the run report and its copy; the caller synchronises the stream.
int queue_report(md_engine* e, RunReport* host) {
    HIPCHECK(md_launch_report(e->rst.p, e->R, e->tmax.p, e->ntiles(), e->err.p, e->report.p, e->st));
    HIPCHECK(hipMemcpyAsync(host, e->report.p, sizeof(RunReport), hipMemcpyDeviceToHost, e->st));
    return MD_OK;
}

bool overflowed(const md_engine* e, const RunReport& r) { return r.rowmax > e->cap || r.setmax > e->jcap; }

// Rebuild all replicas now at the current positions (host synchronous),
// growing the list capacities until every row and tile set fits.
int full_rebuild(md_engine* e) {
    for (int attempt = 0; attempt < 6; attempt++) {
        HIPCHECK(md_launch_force_trip(e->rst.p, e->gate.p, e->R, e->st));
        CHECK(clear_rowmax(e));
        CHECK(enqueue_rebuild(e, nullptr));
        RunReport rr;
        CHECK(queue_report(e, &rr));
        HIPCHECK(hipStreamSynchronize(e->st));
        if (!overflowed(e, rr)) return MD_OK;
        CHECK(grow_list(e, rr.rowmax, rr.setmax));
    }
    return fail(MD_E_CAPACITY, "neighbor list capacity could not be sized");
}

// Queue a rebuild chain (it runs only for replicas whose list is stale).
int refresh_lists(md_engine* e) {
    CHECK(clear_rowmax(e));
    CHECK(enqueue_rebuild(e, nullptr));
    RunReport rr;
    CHECK(queue_report(e, &rr));
    HIPCHECK(hipStreamSynchronize(e->st));
    if (overflowed(e, rr)) CHECK(full_rebuild(e));
    return MD_OK;
}

int read_states(md_engine* e, std::vector<RepState>& s) {
    s.resize(e->R);
    HIPCHECK(hipMemcpyAsync(s.data(), e->rst.p, e->R * sizeof(RepState), hipMemcpyDeviceToHost, e->st));
    HIPCHECK(hipStreamSynchronize(e->st));
    return MD_OK;
}

}  // namespace

// ======================================================================= C ABI
// ---- fixed scratch of md_fm_loss_grad_noalloc ----------------------------
// Byte layout: [0, 512) the sums (2 + 2 k^2 <= 34 doubles), [512, 1536) the
// trial LJ table (64 float4), [2048, MD_FM_SCRATCH_BYTES) the per-tile slots.
// Host code never passes `scratch` itself to a kernel or a copy: its device
// address is resolved per device with hipGetSymbolAddress (kept in
// g_scratch_ptr) at md_create, with that device current.
constexpr size_t kScratchAcc = 0, kScratchLj = 512, kScratchPart = 2048;
constexpr int kMaxDevices = 64;
static __device__ __align__(256) unsigned char scratch[MD_FM_SCRATCH_BYTES];
static std::mutex g_scratch_mu;
static void* g_scratch_ptr[kMaxDevices] = {};

// the current device must be `dev`
static int scratch_setup(int dev) {
    if (dev >= kMaxDevices) return MD_OK;   // md_fm_loss_grad_noalloc then reports the device as unsupported
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (g_scratch_ptr[dev]) return MD_OK;
    void* p = nullptr;
    hipError_t r = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (r != hipSuccess || !p)
        return fail(MD_E_HIP, "hipGetSymbolAddress(scratch) on device %d: %s", dev, hipGetErrorString(r));
    g_scratch_ptr[dev] = p;
    return MD_OK;
}

extern "C" {

const char* md_last_error(void) { return g_err.c_str(); }

const char* md_version(void) {
    return "mdgpu 0.2 (gfx950, hipcc " __clang_version__ ")";
}

int md_create(md_engine** out, int device) {
    if (!out) return fail(MD_E_INVALID, "null output pointer");
    *out = nullptr;
    int ndev = 0;
    HIPCHECK(hipGetDeviceCount(&ndev));
    if (device < 0 || device >= ndev) return fail(MD_E_INVALID, "device %d not present (%d devices)", device, ndev);
    HIPCHECK(hipSetDevice(device));
    md_engine* e = new md_engine();
    e->dev = device;
    hipDeviceProp_t prop;
    size_t lds = 65536;
    if (hipGetDeviceProperties(&prop, device) == hipSuccess) {
        e->ncu = prop.multiProcessorCount;
        e->arch = prop.gcnArchName;
        lds = prop.sharedMemPerBlock;
    }
    e->lds_max = lds;
    // the step and force-matching kernels hold the staged set and the dummy
    // (16 bytes each) next to at most 8 KB of static LDS
    e->jcap_max = (int)std::min<size_t>(MD_JMAX, (lds - 8192) / 16 - 1) / 64 * 64;
    if (hipStreamCreateWithFlags(&e->st, hipStreamNonBlocking) != hipSuccess) {
        delete e;
        return fail(MD_E_HIP, "hipStreamCreate failed");
    }
    if (e->ljtab.alloc(64) || e->ctab.alloc(MD_CTAB + 1) || e->err.alloc(1) || e->fm_acc.alloc(2 + 2 * 64) ||
        e->report.alloc(1) || e->badcnt.alloc(8) || e->sums.alloc(6)) {
        md_destroy(e);
        return MD_E_HIP;
    }
    if (!ev_get(e, 1)) {   // events 0 and 1 exist from here on: timing allocates no events later
        md_destroy(e);
        return fail(MD_E_HIP, "hipEventCreate failed");
    }
    if (int rc = scratch_setup(device)) {
        md_destroy(e);
        return rc;
    }
    *out = e;
    return MD_OK;
}

int md_destroy(md_engine* e) {
    if (!e) return MD_OK;
    (void)hipSetDevice(e->dev);
    (void)hipStreamSynchronize(e->st);
    delete e;
    return MD_OK;
}

int md_set_forcefield(md_engine* e, const md_forcefield* f) {
    if (!e || !f) return fail(MD_E_INVALID, "null argument");
    if (f->ntype < 1 || f->ntype > MD_MAX_TYPES) return fail(MD_E_INVALID, "ntype %d outside 1..%d", f->ntype, MD_MAX_TYPES);
    if (!(f->rc > 0) || !(f->skin >= 0)) return fail(MD_E_INVALID, "rc must be > 0 and skin >= 0");
    if (f->coulomb < 0 || f->coulomb > 3) return fail(MD_E_INVALID, "unknown coulomb kind %d", f->coulomb);
    if (f->nbond_types < 0 || f->nbond_types > 8 || f->nangle_types < 0 || f->nangle_types > 8)
        return fail(MD_E_INVALID, "bond/angle types must be 0..8");
    HIPCHECK(hipSetDevice(e->dev));
    e->ff = *f;
    if (e->ff.qfac == 0.0f) e->ff.qfac = 1389.35427611722f;
    derive_ff(e);
    CHECK(upload_ljtab(e, e->ff.lj1, e->ff.lj2));
    if (e->ff.coulomb == MD_COUL_WOLF || e->ff.coulomb == MD_COUL_DAMPED) {
        const std::vector<float4> t = coulomb_table(e->ff.alpha, e->ff.rc, e->ff.coulomb == MD_COUL_WOLF);
        HIPCHECK(hipMemcpy(e->ctab.p, t.data(), (MD_CTAB + 1) * sizeof(float4), hipMemcpyHostToDevice));
    }
    e->has_ff = true;
    e->has_sys = false;   // cells depend on r_list: set the system again
    e->has_state = false;
    e->drop_graphs();
    return MD_OK;
}

int md_set_integrator(md_engine* e, const md_integrator* in) {
    if (!e || !in) return fail(MD_E_INVALID, "null argument");
    if (in->integrator != MD_LEAPFROG && in->integrator != MD_VVERLET) return fail(MD_E_INVALID, "unknown integrator");
    if (in->thermostat < 0 || in->thermostat > 2) return fail(MD_E_INVALID, "unknown thermostat");
    if (in->integrator == MD_LEAPFROG && in->thermostat == MD_TH_NOSEHOOVER)
        return fail(MD_E_INVALID, "Nose-Hoover needs the velocity-Verlet integrator");
    if (!(in->dt > 0)) return fail(MD_E_INVALID, "dt must be > 0");
    if (in->thermostat != MD_TH_NONE && !(in->tau > 0)) return fail(MD_E_INVALID, "tau must be > 0");
    if (in->barostat != MD_BARO_NONE && in->barostat != MD_BARO_BERENDSEN) return fail(MD_E_INVALID, "unknown barostat");
    if (in->barostat == MD_BARO_BERENDSEN) {
        if (in->integrator != MD_LEAPFROG) return fail(MD_E_INVALID, "pressure coupling needs the leapfrog integrator");
        if (!(in->tau_p > 0) || !(in->compress >= 0)) return fail(MD_E_INVALID, "tau_p must be > 0 and compress >= 0");
    }
    if (e->has_state && e->in.integrator != in->integrator)
        return fail(MD_E_STATE, "switching the integrator needs a new upload (velocity convention differs)");
    // A new velocity-Verlet thermostat starts from the current velocities (leapfrog-form NVE
    // runs do not keep the thermostat's kinetic-energy record): a state round trip resets it.
    const bool rethermo = e->has_state && e->has_integ && in->integrator == MD_VVERLET &&
                          in->thermostat != e->in.thermostat;
    std::vector<float> P, V;
    if (rethermo) {
        P.resize(3 * (size_t)e->N);
        V.resize(3 * (size_t)e->N);
        CHECK(md_download_state(e, P.data(), V.data(), nullptr));
    }
    e->in = *in;
    e->has_integ = true;
    e->drop_graphs();
    if (e->has_sys) {
        for (int r = 0; r < e->R; r++) {
            RepParam& p = e->rep_h[r];
            p.Q = (double)p.nf * p.kT0 * (double)in->tau * in->tau;
        }
        HIPCHECK(hipMemcpy(e->rep.p, e->rep_h.data(), e->R * sizeof(RepParam), hipMemcpyHostToDevice));
    }
    if (rethermo) CHECK(md_upload_state(e, P.data(), V.data()));
    return MD_OK;
}

int md_set_system(md_engine* e, int nrep, const int* natom, const float* box, const float* T0,
                  const int* type, const float* charge, const float* mass) {
    if (!e || !natom || !box || !type || !charge || !mass) return fail(MD_E_INVALID, "null argument");
    if (!e->has_ff) return fail(MD_E_STATE, "set the force field first");
    if (nrep < 1) return fail(MD_E_INVALID, "nrep must be >= 1");
    HIPCHECK(hipSetDevice(e->dev));
    HIPCHECK(hipStreamSynchronize(e->st));
    e->pme_on = false;   // PME buffers, tables and plans are made again for the new system
    long long N = 0, S = 0;
    for (int r = 0; r < nrep; r++) {
        if (natom[r] < 2) return fail(MD_E_INVALID, "replica %d has %d atoms", r, natom[r]);
        for (int d = 0; d < 3; d++)
            if (!(box[3 * r + d] > 0)) return fail(MD_E_INVALID, "replica %d: box must be positive", r);
        N += natom[r];
        S += (natom[r] + MD_TILE - 1) / MD_TILE * MD_TILE;
    }
    if (S > (1LL << MD_SLOTBITS)) return fail(MD_E_CAPACITY, "too many atoms (%lld slots; limit %d)", S, 1 << MD_SLOTBITS);
    for (long long i = 0; i < N; i++) {
        if (type[i] < 0 || type[i] >= e->ff.ntype) return fail(MD_E_INVALID, "atom %lld: type %d outside 0..%d", i, type[i], e->ff.ntype - 1);
        if (!(mass[i] > 0)) return fail(MD_E_INVALID, "atom %lld: mass must be > 0", i);
    }
    e->R = nrep;
    e->natom.assign(natom, natom + nrep);
    e->nrigid = 0;   // a new system: no rigid molecules, no exclusions until given
    e->rig_h.clear();
    e->excl_bond.clear();
    e->box.assign(box, box + 3 * nrep);
    e->T0.assign(nrep, 0.f);
    if (T0) e->T0.assign(T0, T0 + nrep);
    e->atom_off.resize(nrep);
    e->slot_off.resize(nrep);
    for (int r = 0, n = 0, s = 0; r < nrep; r++) {
        e->atom_off[r] = n;
        e->slot_off[r] = s;
        n += natom[r];
        s += (natom[r] + MD_TILE - 1) / MD_TILE * MD_TILE;
    }
    e->N = (int)N;
    e->S = (int)S;
    e->type_h.assign(type, type + N);
    e->charge_h.assign(charge, charge + N);
    e->mass_h.assign(mass, mass + N);
    bool charged = false;
    for (long long i = 0; i < N; i++)
        if (charge[i] != 0.0f) charged = true;
    e->coul_kernel = charged ? e->ff.coulomb : MD_COUL_NONE;
    e->rep_h.assign(nrep, RepParam{});
    e->tile_rep_h.assign(S / MD_TILE, 0);
    for (int r = 0; r < nrep; r++) {
        RepParam& p = e->rep_h[r];
        for (int d = 0; d < 3; d++) {
            p.L[d] = box[3 * r + d];
            p.Linv[d] = (float)(1.0 / (double)box[3 * r + d]);         // src/simulation.cu:290-292
            p.halfL[d] = 0.5f * box[3 * r + d];
            p.Lb[d] = box[3 * r + d];
        }
        p.slot_off = e->slot_off[r];
        p.nslot = (natom[r] + MD_TILE - 1) / MD_TILE * MD_TILE;
        p.natom = natom[r];
        p.atom_off = e->atom_off[r];
        p.T0 = e->T0[r];
        p.nf = (float)(3.0 * natom[r] - 3.0);                           // src/integrator.cu:155
        p.kT0 = MD_KB_REF * (double)p.T0;
        p.Q = (double)p.nf * p.kT0 * (double)e->in.tau * e->in.tau;
        for (int t = p.slot_off / MD_TILE; t < (p.slot_off + p.nslot) / MD_TILE; t++) e->tile_rep_h[t] = r;
    }
    e->drop_graphs();
    const size_t Sz = S;
    for (auto* b : {&e->pos[0], &e->pos[1], &e->vel, &e->force, &e->pos_build, &e->pos_s, &e->vel_s, &e->force_s,
                    &e->target, &e->snap_pos0, &e->snap_pos1, &e->snap_vel, &e->snap_force})
        CHECK(b->alloc(Sz));
    for (auto* b : {&e->ids, &e->typ, &e->ids_s, &e->typ_s, &e->key, &e->rank, &e->tmp, &e->tmpid, &e->ncount,
                    &e->snap_ids, &e->snap_typ})
        CHECK(b->alloc(Sz));
    CHECK(e->jcount.alloc(S / MD_TILE));
    CHECK(e->rperm.alloc(Sz));
    CHECK(e->slot_of.alloc(N));
    CHECK(e->snap_slot_of.alloc(N));
    CHECK(e->rep.alloc(nrep));
    CHECK(e->snap_rep.alloc(nrep));
    CHECK(e->tile_rep.alloc(S / MD_TILE));
    CHECK(e->rst.alloc(nrep));
    CHECK(e->snap_rst.alloc(nrep));
    CHECK(e->gate.alloc(1));
    CHECK(e->snap_gate.alloc(1));
    CHECK(e->th.alloc(nrep));
    CHECK(e->snap_th.alloc(nrep));
    CHECK(e->part.alloc(2 * (size_t)(S / MD_TILE) * MD_NSTAT));
    CHECK(e->lst.alloc(2 * (size_t)nrep));
    CHECK(e->tmax.alloc(2 * (size_t)(S / MD_TILE)));
    CHECK(e->tflag.alloc(S / MD_TILE));
    e->stat_cap = 0;
    CHECK(ensure_stats(e, 1));
    std::vector<int> neg(S, -1);
    HIPCHECK(hipMemcpy(e->ids.p, neg.data(), S * sizeof(int), hipMemcpyHostToDevice));
    HIPCHECK(hipMemcpy(e->ids_s.p, neg.data(), S * sizeof(int), hipMemcpyHostToDevice));
    HIPCHECK(hipMemcpy(e->tile_rep.p, e->tile_rep_h.data(), (S / MD_TILE) * sizeof(int), hipMemcpyHostToDevice));
    CHECK(setup_cells(e));
    CHECK(alloc_list(e, estimate_cap(e), std::min(estimate_jcap(e), e->jcap_max)));
    e->nb = e->na = 0;
    e->bonded = e->has_excl = false;
    CHECK(e->excl.alloc(N));
    CHECK(e->bond_list.alloc(4 * (size_t)N));
    CHECK(e->ang_list.alloc(4 * (size_t)N));
    CHECK(e->nbond.alloc(N));
    CHECK(e->nang.alloc(N));
    e->has_sys = true;
    e->has_state = false;
    e->fm_nframes = 0;
    return MD_OK;
}

int md_set_topology(md_engine* e, int nb, const int* bonds, int na, const int* angles) {
    if (!e || !e->has_sys) return fail(MD_E_STATE, "set the system first");
    if (nb < 0 || na < 0 || (nb && !bonds) || (na && !angles)) return fail(MD_E_INVALID, "bad topology arrays");
    HIPCHECK(hipSetDevice(e->dev));
    const int N = e->N;
    std::vector<int4> ex(N, make_int4(-1, -1, -1, -1));
    std::vector<uint32_t> bl(4 * (size_t)N, 0);
    std::vector<int> nbo(N, 0), nan_(N, 0), nex(N, 0);
    std::vector<int4> al(4 * (size_t)N, make_int4(0, 0, 0, 0));
    for (int b = 0; b < nb; b++) {
        int i = bonds[3 * b], j = bonds[3 * b + 1], t = bonds[3 * b + 2];
        if (i < 0 || i >= N || j < 0 || j >= N || i == j) return fail(MD_E_INVALID, "bond %d: bad atoms %d %d", b, i, j);
        if (t < 0 || t >= e->ff.nbond_types) return fail(MD_E_INVALID, "bond %d: type %d not defined", b, t);
        for (int s = 0; s < 2; s++) {
            int a = s ? j : i, c = s ? i : j;
            if (nbo[a] >= 4) return fail(MD_E_CAPACITY, "atom %d has more than 4 bonds (src/forces.cu:118)", a);
            bl[(size_t)a * 4 + nbo[a]++] = (uint32_t)c | ((uint32_t)t << MD_JBITS);
            int* sl = &ex[a].x;
            sl[nex[a]++] = c;   // 1-2 exclusion (src/neighborlist.cu:74-97, slot bug not copied)
        }
    }
    for (int q = 0; q < na; q++) {
        int i = angles[4 * q], j = angles[4 * q + 1], k = angles[4 * q + 2], t = angles[4 * q + 3];
        if (i < 0 || i >= N || j < 0 || j >= N || k < 0 || k >= N) return fail(MD_E_INVALID, "angle %d: bad atoms", q);
        if (t < 0 || t >= e->ff.nangle_types) return fail(MD_E_INVALID, "angle %d: type %d not defined", q, t);
        int who[3] = {i, j, k};
        int4 ent[3] = {make_int4(0 | (t << 8), j, k, 0), make_int4(1 | (t << 8), i, k, 0), make_int4(0 | (t << 8), j, i, 0)};
        for (int s = 0; s < 3; s++) {   // per-atom angle list, src/forces.cu:154-173
            int a = who[s];
            if (nan_[a] >= 4) return fail(MD_E_CAPACITY, "atom %d is in more than 4 angles (src/forces.cu:147)", a);
            al[(size_t)a * 4 + nan_[a]++] = ent[s];
        }
    }
    e->excl_bond = ex;
    CHECK(upload_exclusions(e));
    HIPCHECK(hipMemcpy(e->bond_list.p, bl.data(), bl.size() * sizeof(uint32_t), hipMemcpyHostToDevice));
    HIPCHECK(hipMemcpy(e->nbond.p, nbo.data(), N * sizeof(int), hipMemcpyHostToDevice));
    HIPCHECK(hipMemcpy(e->ang_list.p, al.data(), al.size() * sizeof(int4), hipMemcpyHostToDevice));
    HIPCHECK(hipMemcpy(e->nang.p, nan_.data(), N * sizeof(int), hipMemcpyHostToDevice));
    e->nb = nb;
    e->na = na;
    e->bonded = nb > 0 || na > 0;
    e->drop_graphs();
    if (e->has_state) CHECK(full_rebuild(e));
    return MD_OK;
}

int md_set_rigid(md_engine* e, int nmol, const int* mol, float d_oh, float d_hh) {
    if (!e || !e->has_sys) return fail(MD_E_STATE, "set the system first");
    if (nmol < 0 || (nmol > 0 && !mol)) return fail(MD_E_INVALID, "bad molecule array");
    if (nmol > 0 && !(d_oh > 0.0f && d_hh > 0.0f && d_hh < 2.0f * d_oh))
        return fail(MD_E_INVALID, "rigid geometry needs 0 < d_hh < 2 d_oh (got %g, %g)", d_oh, d_hh);
    HIPCHECK(hipSetDevice(e->dev));
    HIPCHECK(hipStreamSynchronize(e->st));
    const int N = e->N;
    std::vector<int> of(N, -1);
    for (int m = 0; m < nmol; m++) {
        const int* a = mol + 3 * (size_t)m;
        int rep_of[3];
        for (int u = 0; u < 3; u++) {
            if (a[u] < 0 || a[u] >= N) return fail(MD_E_INVALID, "molecule %d: atom %d out of range", m, a[u]);
            if (of[a[u]] != -1) return fail(MD_E_INVALID, "atom %d is in two rigid molecules", a[u]);
            of[a[u]] = u == 0 ? m : -2;
            rep_of[u] = (int)(std::upper_bound(e->atom_off.begin(), e->atom_off.end(), a[u]) - e->atom_off.begin()) - 1;
        }
        if (rep_of[1] != rep_of[0] || rep_of[2] != rep_of[0])
            return fail(MD_E_INVALID, "molecule %d spans two replicas", m);
        if (e->mass_h[a[1]] != e->mass_h[a[2]]) return fail(MD_E_INVALID, "molecule %d: the two H masses differ", m);
    }
    e->nrigid = nmol;
    e->rig_oh = d_oh;
    e->rig_hh = d_hh;
    e->rig_h.assign(mol, mol + 3 * (size_t)nmol);
    if (nmol > 0) {
        std::vector<int4> mh(nmol);
        for (int m = 0; m < nmol; m++) mh[m] = make_int4(mol[3 * m], mol[3 * m + 1], mol[3 * m + 2], 0);
        CHECK(e->rig_mol.alloc(nmol));
        HIPCHECK(hipMemcpy(e->rig_mol.p, mh.data(), nmol * sizeof(int4), hipMemcpyHostToDevice));
    }
    CHECK(e->rig_of.alloc(N));
    HIPCHECK(hipMemcpy(e->rig_of.p, of.data(), N * sizeof(int), hipMemcpyHostToDevice));
    CHECK(upload_exclusions(e));
    CHECK(upload_dof(e));
    e->drop_graphs();
    if (e->has_state) CHECK(full_rebuild(e));
    return MD_OK;
}

int md_set_pme(md_engine* e, const int* grid, int order) {
    if (!e) return fail(MD_E_INVALID, "null engine");
    if (order != 0) {
        if (order != 4 && order != 6 && order != 8) return fail(MD_E_INVALID, "PME order must be 4, 6 or 8 (got %d)", order);
        if (!grid) return fail(MD_E_INVALID, "null grid");
        for (int d = 0; d < 3; d++)
            if (grid[d] < order || grid[d] > 1024)
                return fail(MD_E_INVALID, "PME grid edge %d must be in [order, 1024] (got %d)", d, grid[d]);
    }
    HIPCHECK(hipSetDevice(e->dev));
    HIPCHECK(hipStreamSynchronize(e->st));
    e->pme_order = order;
    for (int d = 0; d < 3; d++) e->pme_K[d] = order ? grid[d] : 0;
    e->pme_on = false;
    e->vv_force_stale = e->has_state;
    e->drop_graphs();
    return MD_OK;
}

int md_upload_state(md_engine* e, const float* pos, const float* vel) {
    CHECK(check_ready(e, false));
    if (!pos) return fail(MD_E_INVALID, "null positions");
    HIPCHECK(hipSetDevice(e->dev));
    HIPCHECK(hipStreamSynchronize(e->st));
    if (e->has_state) {   // the builds of the state being replaced stay counted
        std::vector<RepState> old;
        CHECK(read_states(e, old));
        for (const RepState& x : old) e->reb_base += x.nb;
    }
    const int S = e->S;
    std::vector<float4> P(S, make_float4(0, 0, 0, 0)), V(S, make_float4(0, 0, 0, 0));
    std::vector<int> ids(S, -1), typ(S, 0), slot_of(e->N);
    std::vector<ThState> th(e->R);
    std::vector<RepState> rs(e->R);
    for (int r = 0; r < e->R; r++) {
        double ke2 = 0;
        for (int k = 0; k < e->natom[r]; k++) {
            int i = e->atom_off[r] + k, s = e->slot_off[r] + k;
            float x[3];
            for (int d = 0; d < 3; d++) {
                float L = e->box[3 * r + d];
                float v = pos[3 * (size_t)i + d];
                if (!std::isfinite(v)) return fail(MD_E_INVALID, "atom %d: non-finite position", i);
                v -= L * std::floor(v / L + 0.5f);   // into the centred box
                if (v >= 0.5f * L) v -= L;
                x[d] = v;
            }
            P[s] = make_float4(x[0], x[1], x[2], e->charge_h[i]);
            float invm = (float)(1.0 / (1e4 * (double)e->mass_h[i]));   // src/simulation.cu:177
            float v3[3] = {0, 0, 0};
            if (vel)
                for (int d = 0; d < 3; d++) v3[d] = vel[3 * (size_t)i + d];
            V[s] = make_float4(v3[0], v3[1], v3[2], invm);
            ke2 += (double)(v3[0] * (double)v3[0] + v3[1] * (double)v3[1] + v3[2] * (double)v3[2]) / (double)invm;
            ids[s] = i;
            typ[s] = e->type_h[i];
            slot_of[i] = s;
        }
        th[r] = ThState{0, 0, 1.0, 1.0, 0, 0, ke2, 0};
        RepState z;
        memset(&z, 0, sizeof z);
        z.step = z.target = z.start = e->step;
        z.trip[0] = 1;   // no list yet
        rs[r] = z;
    }
    for (int b = 0; b < 2; b++) HIPCHECK(hipMemcpy(e->pos[b].p, P.data(), S * sizeof(float4), hipMemcpyHostToDevice));
    HIPCHECK(hipMemcpy(e->vel.p, V.data(), S * sizeof(float4), hipMemcpyHostToDevice));
    HIPCHECK(hipMemset(e->force.p, 0, S * sizeof(float4)));
    HIPCHECK(hipMemcpy(e->ids.p, ids.data(), S * sizeof(int), hipMemcpyHostToDevice));
    HIPCHECK(hipMemcpy(e->typ.p, typ.data(), S * sizeof(int), hipMemcpyHostToDevice));
    HIPCHECK(hipMemcpy(e->slot_of.p, slot_of.data(), e->N * sizeof(int), hipMemcpyHostToDevice));
    HIPCHECK(hipMemcpy(e->th.p, th.data(), e->R * sizeof(ThState), hipMemcpyHostToDevice));
    HIPCHECK(hipMemcpy(e->rst.p, rs.data(), e->R * sizeof(RepState), hipMemcpyHostToDevice));
    HIPCHECK(hipMemset(e->err.p, 0, sizeof(int)));
    e->has_state = true;
    CHECK(full_rebuild(e));
    if (e->in.integrator == MD_VVERLET) {
        // velocity Verlet needs F(x0) before the first half kick
        StepArgs a = step_args(e, MODE_FORCE, 0);
        a.stats = nullptr;
        CHECK(enqueue_stepkernel(e, a, nullptr));
        HIPCHECK(hipStreamSynchronize(e->st));
        e->vv_force_stale = e->pme_order > 0 && !e->pme_on;   // md_run adds the PME part first
    }
    return MD_OK;
}

int md_download_state(md_engine* e, float* pos, float* vel, float* force) {
    CHECK(check_ready(e, true));
    HIPCHECK(hipSetDevice(e->dev));
    std::vector<RepState> rs;
    CHECK(read_states(e, rs));
    const int S = e->S;
    std::vector<int> ids(S);
    HIPCHECK(hipMemcpy(ids.data(), e->ids.p, S * sizeof(int), hipMemcpyDeviceToHost));
    std::vector<float4> buf(S), buf2;
    float* outs[3] = {pos, vel, force};
    for (int k = 0; k < 3; k++) {
        if (!outs[k]) continue;
        if (k == 0) {   // positions: each replica's current buffer
            buf2.resize(S);
            HIPCHECK(hipMemcpy(buf.data(), e->pos[0].p, S * sizeof(float4), hipMemcpyDeviceToHost));
            HIPCHECK(hipMemcpy(buf2.data(), e->pos[1].p, S * sizeof(float4), hipMemcpyDeviceToHost));
        } else {
            HIPCHECK(hipMemcpy(buf.data(), k == 1 ? e->vel.p : e->force.p, S * sizeof(float4), hipMemcpyDeviceToHost));
        }
        for (int s = 0; s < S; s++) {
            int i = ids[s];
            if (i < 0) continue;
            const int r = e->tile_rep_h[s / MD_TILE];
            float4 v = (k == 0 && cur_buf(rs[r]) == 1) ? buf2[s] : buf[s];
            if (k == 0) {   // positions run unwrapped between list builds: wrap as the build does
                const RepParam& p = e->rep_h[r];
                v.x = wrap_coord(v.x, p.L[0], p.Linv[0]);
                v.y = wrap_coord(v.y, p.L[1], p.Linv[1]);
                v.z = wrap_coord(v.z, p.L[2], p.Linv[2]);
            }
            outs[k][3 * (size_t)i] = v.x;
            outs[k][3 * (size_t)i + 1] = v.y;
            outs[k][3 * (size_t)i + 2] = v.z;
        }
    }
    return MD_OK;
}

int md_compute_forces(md_engine* e, float* force, double* energy, double* virial) {
    CHECK(check_ready(e, true));
    HIPCHECK(hipSetDevice(e->dev));
    CHECK(pme_prepare(e));
    CHECK(refresh_lists(e));
    StepArgs a = step_args(e, MODE_FORCE, 0);
    a.force_energy = 1;
    CHECK(enqueue_stepkernel(e, a, nullptr));
    e->vv_force_stale = false;   // the force array now holds F(x) with the PME part as set
    std::vector<StatRec> rec(e->R);
    HIPCHECK(hipMemcpyAsync(rec.data(), e->stats.p, e->R * sizeof(StatRec), hipMemcpyDeviceToHost, e->st));
    HIPCHECK(hipStreamSynchronize(e->st));
    for (int r = 0; r < e->R; r++) {
        if (energy) {
            energy[4 * r] = rec[r].pe_lj;
            energy[4 * r + 1] = rec[r].pe_coul;
            energy[4 * r + 2] = rec[r].pe_bond;
            energy[4 * r + 3] = rec[r].pe_angle;
        }
        if (virial) virial[r] = rec[r].virial;
    }
    if (force) CHECK(md_download_state(e, nullptr, nullptr, force));
    return MD_OK;
}

}  // extern "C"

namespace {

// One segment: `blocks` x ([gated rebuild chain] + check_every steps).
int enqueue_segment(md_engine* e, int energy_every, int blocks) {
    const int ce = std::max(1, e->check_every);
    for (int b = 0; b < blocks; b++) {
        if (e->defer && lf_form(e) && !e->in.barostat && e->nrigid == 0 && !e->pme_on) {
            // the block's first step kernel opens the rebuilt lists (pressure coupling scales the
            // box in the finalize kernel after every step; SETTLE runs between a step kernel and
            // its finalize)
            CHECK(enqueue_rebuild(e, nullptr, false));
            CHECK(enqueue_lf_block(e, energy_every, ce));
        } else {
            CHECK(enqueue_rebuild(e, nullptr));
            for (int k = 0; k < ce; k++) CHECK(enqueue_step(e, energy_every, nullptr));
        }
    }
    return MD_OK;
}

// blocks of a full segment
int segment_blocks(const md_engine* e) { return std::max(1, kSegment / std::max(1, e->check_every)); }

// Queue one segment of `blocks` blocks: a replay of its captured graph, or direct launches.
int launch_one(md_engine* e, int energy_every, int blocks) {
    if (e->submit == 1) return enqueue_segment(e, energy_every, blocks);
    const int ce = std::max(1, e->check_every);
    char keyb[96];
    snprintf(keyb, sizeof keyb, "%d:%d:%d", ce, energy_every, blocks);
    std::string key(keyb);
    auto it = e->graphs.find(key);
    if (it == e->graphs.end()) {
        hipGraph_t g;
        HIPCHECK(hipStreamBeginCapture(e->st, hipStreamCaptureModeThreadLocal));
        const int rc = enqueue_segment(e, energy_every, blocks);
        hipError_t cerr = hipStreamEndCapture(e->st, &g);
        if (rc != MD_OK) return rc;
        if (cerr != hipSuccess) return fail(MD_E_HIP, "graph capture: %s", hipGetErrorString(cerr));
        md_engine::SegGraph sg{{nullptr, nullptr}, 0};
        hipError_t ie = hipGraphInstantiate(&sg.ex[0], g, nullptr, nullptr, 0);
        if (ie == hipSuccess) ie = hipGraphInstantiate(&sg.ex[1], g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        if (ie != hipSuccess) {
            for (auto x : sg.ex)
                if (x) (void)hipGraphExecDestroy(x);
            return fail(MD_E_HIP, "graph instantiate: %s", hipGetErrorString(ie));
        }
        it = e->graphs.emplace(key, sg).first;
    }
    md_engine::SegGraph& sg = it->second;
    HIPCHECK(hipGraphLaunch(sg.ex[sg.next], e->st));
    sg.next ^= 1;
    return MD_OK;
}

// Queue `nblocks` blocks: full segments, then the rest as segments of
// power-of-two block counts (a short run does not replay a whole segment of
// idle launches, and at most log2 of a segment's blocks graphs per interval
// are captured).
int launch_blocks(md_engine* e, int energy_every, long long nblocks) {
    const int full = segment_blocks(e);
    for (; nblocks >= full; nblocks -= full) CHECK(launch_one(e, energy_every, full));
    for (int b = 1 << 30; b >= 1; b >>= 1)
        if (nblocks & b) CHECK(launch_one(e, energy_every, b));
    return MD_OK;
}

// ---- the persistent kernel for small single systems (persist.hip) ----------
// One replica of at most MD_PK_SLOTS slots, Lennard-Jones only (no charges,
// bonded terms or exclusions), leapfrog (NVE or Berendsen) or NVE velocity
// Verlet, no pressure coupling, and r_list below half the box (the kernel's
// pairs take the minimum image).
bool persist_eligible(const md_engine* e) {
    if (e->pk_off || e->R != 1 || e->S > MD_PK_SLOTS) return false;
    if (e->bonded || e->has_excl || e->coul_kernel != MD_COUL_NONE || e->ff.ntype > 8) return false;
    if (e->in.barostat != MD_BARO_NONE) return false;
    const bool lf = e->in.integrator == MD_LEAPFROG && e->in.thermostat != MD_TH_NOSEHOOVER;
    const bool vv = e->in.integrator == MD_VVERLET && e->in.thermostat == MD_TH_NONE;
    if (!lf && !vv) return false;
    const float rl = std::sqrt(e->rl2);
    for (int d = 0; d < 3; d++)
        if (!(2.0f * rl < e->box[d])) return false;
    return true;
}
bool persist_ok(const md_engine* e) { return e->persist && persist_eligible(e); }

// Dynamic LDS of a launch: staged positions and types (20 B per slot) and the
// lane rows (2 B per entry and thread); at least half the CU's 160 KB, so that
// no two of the workgroups share a CU.
size_t persist_lds(const md_engine* e, int cap_l) {
    const size_t need = (size_t)e->S * 20 + (size_t)cap_l * MD_PK_THREADS * 2;
    return std::max(need, (size_t)81920);
}

// Row entries per lane: 1.5 x the mean count of neighbors within r_list over
// the atom's lanes, plus 6, or less when the caller set a smaller list
// capacity (md_set_list_capacity: entries per atom over the lanes, plus 2).  A
// longer row is caught and the run replayed with rows twice as long.
int persist_cap(const md_engine* e) {
    if (e->pk_cap > 0) return e->pk_cap;
    const double V = (double)e->box[0] * e->box[1] * e->box[2];
    const double rl = std::sqrt((double)e->rl2);
    const double n_rl = e->N / V * 4.0 / 3.0 * kPi * rl * rl * rl;
    int c = (int)std::ceil(1.5 * n_rl / e->lanes) + 6;
    if (e->pk_req > 0) c = std::min(c, (e->pk_req + e->lanes - 1) / e->lanes + 2);
    return (c + 1) & ~1;
}

// One run on the persistent kernel.  *use_tiles = 1 (state unchanged) when the
// rows do not fit the kernel's LDS: the caller runs the tile path instead.
int run_persist(md_engine* e, int nsteps, int energy_every, bool want_stats, int* use_tiles) {
    *use_tiles = 0;
    const int G = e->lanes, apw = MD_PK_THREADS / G;
    const int nwg = (e->S + apw - 1) / apw;
    if (nwg > MD_PK_WG) {   // (S <= MD_PK_SLOTS and apw >= 64 keep nwg <= 64)
        *use_tiles = 1;
        return MD_OK;
    }
    if (!e->pk_flags.p) CHECK(e->pk_flags.alloc(2 * MD_PK_WG));
    if (!e->pk_pub.p) CHECK(e->pk_pub.alloc(2 * MD_PK_WG * 4));
    const size_t static_lds = 2048;   // the kernel's own tables and sums (1.4 KB)
    for (;;) {
        const int cap_l = persist_cap(e);
        if (persist_lds(e, cap_l) + static_lds > e->lds_max) {
            e->pk_off = true;   // rows this long do not fit: the tile path from now on
            *use_tiles = 1;
            return MD_OK;
        }
        CHECK(snapshot(e, false));
        HIPCHECK(md_launch_persist_clear(e->pk_flags.p, e->st));
        HIPCHECK(hipMemsetAsync(e->err.p, 0, sizeof(int), e->st));
        PersistArgs a{};
        a.pos[0] = e->pos[0].p; a.pos[1] = e->pos[1].p; a.vel = e->vel.p; a.force = e->force.p;
        a.ids = e->ids.p; a.typ = e->typ.p; a.ljtab = e->ljtab.p;
        a.st = e->rst.p; a.th = e->th.p; a.rep = e->rep.p;
        a.flags = e->pk_flags.p; a.pub = e->pk_pub.p; a.err = e->err.p;
        a.stats = want_stats ? e->stats.p : nullptr; a.stat_cap = e->stat_cap;
        a.nslot = e->S; a.nwg = nwg; a.nsteps = nsteps; a.energy_every = energy_every;
        a.leapfrog = e->in.integrator == MD_LEAPFROG ? 1 : 0;
        a.thermostat = e->in.thermostat; a.ntype = e->ff.ntype; a.cap_l = cap_l;
        a.dt = e->in.dt; a.tau = e->in.tau; a.rc2 = e->rc2; a.rl2 = e->rl2; a.skin2q = e->skin2q;
        hipEvent_t ev0 = nullptr, ev1 = nullptr;
        if (e->timing) {
            ev0 = ev_get(e, 2);
            ev1 = ev_get(e, 3);
            if (!ev0 || !ev1) return fail(MD_E_HIP, "hipEventCreate failed");
        }
        HIPCHECK(md_launch_persist(a, G, persist_lds(e, cap_l), e->st, ev0, ev1));
        int err = 0;
        HIPCHECK(hipMemcpyAsync(&err, e->err.p, sizeof(int), hipMemcpyDeviceToHost, e->st));
        HIPCHECK(hipStreamSynchronize(e->st));
        if (err & MD_PK_ERR_WAIT) return fail(MD_E_HIP, "persistent kernel: a grid barrier wait timed out");
        if (err & MD_PK_ERR_ROWS) {   // a row outgrew its length: replay the run with longer rows
            CHECK(snapshot(e, true));
            HIPCHECK(hipStreamSynchronize(e->st));
            e->pk_cap = 2 * cap_l;
            e->pk_replays++;
            continue;
        }
        if (err) return fail(MD_E_UNSTABLE, "non-finite coordinates or an atom beyond the box (steps %lld..%lld)",
                             e->step, e->step + nsteps);
        if (e->timing) {   // counted per step: one launch runs them all
            float ms = 0;
            HIPCHECK(hipEventElapsedTime(&ms, ev0, ev1));
            e->t_step_ms += ms;
            e->n_step += nsteps;
        }
        e->pk_runs++;
        break;
    }
    // the tile path's lists were built for earlier positions
    HIPCHECK(md_launch_force_trip(e->rst.p, e->gate.p, e->R, e->st));
    return MD_OK;
}

}  // namespace

extern "C" {

int md_run(md_engine* e, int nsteps, int energy_every, md_step_stats* stats) {
    CHECK(check_ready(e, true));
    CHECK(pme_prepare(e));
    if (e->vv_force_stale && e->in.integrator == MD_VVERLET) {
        // velocity Verlet's first half kick reads F(x0): with the PME part as now set
        HIPCHECK(hipSetDevice(e->dev));
        CHECK(refresh_lists(e));
        StepArgs a = step_args(e, MODE_FORCE, 0);
        a.stats = nullptr;
        CHECK(enqueue_stepkernel(e, a, nullptr));
        HIPCHECK(hipStreamSynchronize(e->st));
    }
    e->vv_force_stale = false;
    if (nsteps < 0) return fail(MD_E_INVALID, "nsteps < 0");
    if (nsteps == 0) return MD_OK;
    if (energy_every < 0) return fail(MD_E_INVALID, "energy_every < 0");
    if (e->nrigid > 0 && (e->in.integrator != MD_LEAPFROG || e->in.barostat != MD_BARO_NONE))
        return fail(MD_E_INVALID, "rigid molecules need the leapfrog integrator and no pressure coupling");
    HIPCHECK(hipSetDevice(e->dev));
    const bool vv = e->in.integrator == MD_VVERLET, fused = vv_fused(e);
    if (stats) CHECK(ensure_stats(e, nsteps));
    const long long target = e->step + nsteps;
    bool tiles = true;   // the tile path (else the persistent kernel ran the whole run)
    if (persist_ok(e)) {
        int use_tiles = 0;
        CHECK(run_persist(e, nsteps, energy_every, stats != nullptr, &use_tiles));
        tiles = use_tiles != 0;
    }
    int retries = 0;
    while (tiles) {   // one attempt; repeated only after a list-capacity overflow
        CHECK(snapshot(e, false));
        CHECK(clear_rowmax(e));
        HIPCHECK(hipMemsetAsync(e->err.p, 0, sizeof(int), e->st));
        HIPCHECK(md_launch_set_target(e->rst.p, e->gate.p, e->R, nsteps, e->st));
        if (vv) HIPCHECK(md_launch_thermo_fix(e->th.p, e->rep.p, e->R, e->in.thermostat, e->in.dt, 0, e->st));
        if (fused) {
            // leapfrog-form velocity Verlet: a valid list, then the first half-kick and drift of
            // the run's first step; every step kernel then ends one step and begins the next
            CHECK(enqueue_rebuild(e, nullptr));
            HIPCHECK(md_launch_kick(kick_args(e, 1), e->st));
        }
        long long launched = 0;
        RunReport rr{};
        if (e->timing) {
            // direct launches, one rebuild chain per step, HIP events around each kernel
            TimedSpan span{e, {}, {}, 2};
            for (int k = 0; k < nsteps; k++) {
                CHECK(enqueue_rebuild(e, &span));
                CHECK(enqueue_step(e, energy_every, &span, true));
            }
            launched = nsteps;
            CHECK(queue_report(e, &rr));
            HIPCHECK(hipStreamSynchronize(e->st));
            for (auto& pr : span.steps) {
                float ms = 0;
                HIPCHECK(hipEventElapsedTime(&ms, pr.first, pr.second));
                e->t_step_ms += ms;
                e->n_step++;
            }
            for (auto& pr : span.rebs) {
                float ms = 0;
                HIPCHECK(hipEventElapsedTime(&ms, pr.first, pr.second));
                e->t_reb_ms += ms;
                e->n_reb++;
            }
        } else {
            const int ce = std::max(1, e->check_every);
            long long want = (long long)std::ceil(nsteps * e->queue_ratio);
            for (int round = 0; round < 64; round++) {
                const long long nbl = std::max(1LL, (want + ce - 1) / ce);
                CHECK(launch_blocks(e, energy_every, nbl));
                launched += nbl * ce;
                CHECK(queue_report(e, &rr));
                HIPCHECK(hipStreamSynchronize(e->st));
                const long long lag = target - rr.min_step;
                if (lag <= 0 || overflowed(e, rr) || rr.err) break;
                want = (long long)std::ceil(lag * std::max(e->queue_ratio, 1.0) * 1.1) + 1;
            }
        }
        if (overflowed(e, rr)) {   // a row or tile set overflowed: roll the run back with larger capacities
            // (checked first: an attempt with truncated rows may also have gone unstable)
            if (++retries > 4)
                return fail(MD_E_CAPACITY, "neighbor list capacity keeps overflowing (rows %d > %d or sets %d > %d)",
                            rr.rowmax, e->cap, rr.setmax, e->jcap);
            CHECK(snapshot(e, true));
            CHECK(grow_list(e, rr.rowmax, rr.setmax));
            CHECK(full_rebuild(e));
            continue;
        }
        if (rr.err == 2) return fail(MD_E_UNSTABLE, "the coupled box shrank below twice the list radius (steps %lld..%lld)", e->step, target);
        if (rr.err) return fail(MD_E_UNSTABLE, "non-finite coordinates or an atom beyond the box (steps %lld..%lld)", e->step, target);
        if (rr.min_step != target) {
            std::vector<RepState> rs;
            CHECK(read_states(e, rs));
            for (auto& s : rs)
                if (s.step != target) return fail(MD_E_STATE, "replica stopped at step %lld of %lld", s.step, target);
        }
        if (!e->timing) {   // learn how many launches a step costs (stalls until the next rebuild chain)
            // Over-queuing costs an idle launch per surplus step; under-queuing one more host
            // round trip.  Follow a lower ratio at once, a higher one gradually.
            const double ratio = std::max(1.0, (double)launched / nsteps * 0.98);
            e->queue_ratio = ratio < e->queue_ratio ? ratio : 0.7 * e->queue_ratio + 0.3 * ratio;
        }
        break;
    }
    if (tiles && vv && !fused) {   // (leapfrog form: the last step kernel left full-step velocities)
        HIPCHECK(md_launch_kick(kick_args(e, 0), e->st));
        HIPCHECK(md_launch_thermo_fix(e->th.p, e->rep.p, e->R, e->in.thermostat, e->in.dt, 1, e->st));
    }
    if (stats) {
        std::vector<StatRec> rec((size_t)nsteps * e->R);
        HIPCHECK(hipMemcpyAsync(rec.data(), e->stats.p, rec.size() * sizeof(StatRec), hipMemcpyDeviceToHost, e->st));
        HIPCHECK(hipStreamSynchronize(e->st));
        for (size_t k = 0; k < rec.size(); k++) {
            const StatRec& s = rec[k];
            md_step_stats& o = stats[k];
            o.ke = s.ke; o.temp = s.temp; o.pe_lj = s.pe_lj; o.pe_coul = s.pe_coul; o.pe_bond = s.pe_bond;
            o.pe_angle = s.pe_angle; o.virial = s.virial; o.conserved = s.conserved; o.xi = s.xi;
            o.vxi = s.vxi; o.volume = s.volume; o.has_energy = s.has_energy; o.pad = 0;
        }
    }
    HIPCHECK(hipStreamSynchronize(e->st));
    e->step = target;
    if (e->in.barostat) CHECK(sync_box(e));
    return MD_OK;
}

int md_get_neighbors(md_engine* e, int cap, int* counts, int* nbrs) {
    CHECK(check_ready(e, true));
    if (!counts || cap < 0 || (cap > 0 && !nbrs)) return fail(MD_E_INVALID, "bad output arrays");
    HIPCHECK(hipSetDevice(e->dev));
    CHECK(refresh_lists(e));
    const int S = e->S, nt = e->ntiles();
    std::vector<int> ids(S), nc(S), jc(nt);
    std::vector<uint16_t> L((size_t)S * e->cap);
    std::vector<uint32_t> J((size_t)nt * e->jcap);
    HIPCHECK(hipMemcpy(ids.data(), e->ids.p, S * sizeof(int), hipMemcpyDeviceToHost));
    HIPCHECK(hipMemcpy(nc.data(), e->ncount.p, S * sizeof(int), hipMemcpyDeviceToHost));
    HIPCHECK(hipMemcpy(jc.data(), e->jcount.p, nt * sizeof(int), hipMemcpyDeviceToHost));
    HIPCHECK(hipMemcpy(L.data(), e->nl16.p, L.size() * sizeof(uint16_t), hipMemcpyDeviceToHost));
    HIPCHECK(hipMemcpy(J.data(), e->jset.p, J.size() * sizeof(uint32_t), hipMemcpyDeviceToHost));
    int worst = 0;
    for (int s = 0; s < S; s++) {
        int i = ids[s];
        if (i < 0) continue;
        counts[i] = nc[s];
        worst = std::max(worst, nc[s]);
        const int t = s / MD_TILE;
        int c = std::min(std::min(nc[s], cap), e->cap);
        for (int k = 0; k < c; k++) {
            const int li = L[(size_t)s * e->cap + lpos16(k, e->lanes)] >> 4;   // entries are byte offsets
            if (li < 1 || li > jc[t]) return fail(MD_E_STATE, "row of slot %d holds set index %d outside 1..%d", s, li, jc[t]);
            nbrs[(size_t)i * cap + k] = ids[J[(size_t)t * e->jcap + li - 1] & MD_SLOTMASK];
        }
    }
    if (worst > cap) return fail(MD_E_CAPACITY, "rows need %d entries, cap is %d", worst, cap);
    if (worst > e->cap) return fail(MD_E_CAPACITY, "device list overflowed (%d > %d)", worst, e->cap);
    return MD_OK;
}

int md_get_sums(md_engine* e, double* force_sum, double* momentum_sum) {
    CHECK(check_ready(e, true));
    HIPCHECK(hipSetDevice(e->dev));
    HIPCHECK(md_launch_state_sums(e->vel.p, e->force.p, e->ids.p, e->S, e->sums.p, e->st));
    double h[6];
    HIPCHECK(hipMemcpyAsync(h, e->sums.p, sizeof h, hipMemcpyDeviceToHost, e->st));
    HIPCHECK(hipStreamSynchronize(e->st));
    for (int d = 0; d < 3; d++) {
        if (force_sum) force_sum[d] = h[d];
        if (momentum_sum) momentum_sum[d] = h[3 + d];
    }
    return MD_OK;
}

int md_get_box(md_engine* e, float* box) {
    if (!e || !box) return fail(MD_E_INVALID, "null argument");
    if (!e->has_sys) return fail(MD_E_STATE, "system not set");
    HIPCHECK(hipSetDevice(e->dev));
    HIPCHECK(hipStreamSynchronize(e->st));
    CHECK(sync_box(e));
    for (int k = 0; k < 3 * e->R; k++) box[k] = e->box[k];
    return MD_OK;
}

int md_get_thresholds(md_engine* e, float* rc2, float* rl2, float* skin2q) {
    if (!e || !e->has_ff) return fail(MD_E_STATE, "force field not set");
    if (rc2) *rc2 = e->rc2;
    if (rl2) *rl2 = e->rl2;
    if (skin2q) *skin2q = e->skin2q;
    return MD_OK;
}

int md_get_info(md_engine* e, long long* rebuilds, long long* steps, int* cap, int* lanes, int* grid) {
    if (!e) return fail(MD_E_INVALID, "null engine");
    if (rebuilds) {   // every list build opens an epoch: the replicas' epoch counters add up the builds
        long long c = e->reb_base;
        if (e->has_state) {
            std::vector<RepState> rs;
            HIPCHECK(hipSetDevice(e->dev));
            CHECK(read_states(e, rs));
            for (const RepState& x : rs) c += x.nb;
        }
        *rebuilds = c;
    }
    if (steps) *steps = e->step;
    if (cap) *cap = e->cap;
    if (lanes) *lanes = e->lanes;
    if (grid) *grid = e->has_sys ? e->ntiles() : 0;   // one step-kernel workgroup per 64-slot tile
    return MD_OK;
}

int md_set_list_capacity(md_engine* e, int cap) {
    CHECK(check_ready(e, true));
    if (cap < 1) return fail(MD_E_INVALID, "capacity must be >= 1");
    HIPCHECK(hipSetDevice(e->dev));
    HIPCHECK(hipStreamSynchronize(e->st));
    CHECK(alloc_list(e, cap, e->jcap));
    e->pk_req = cap;
    e->pk_cap = 0;
    e->pk_off = false;
    HIPCHECK(md_launch_force_trip(e->rst.p, e->gate.p, e->R, e->st));   // every list is rebuilt when next needed
    HIPCHECK(hipStreamSynchronize(e->st));
    return MD_OK;
}

int md_set_launch(md_engine* e, const md_launch_config* c) {
    if (!e || !c) return fail(MD_E_INVALID, "null argument");
    if (c->lanes != 4 && c->lanes != 8 && c->lanes != 16)
        return fail(MD_E_INVALID, "lanes must be 4, 8 or 16");
    if (c->grid < 0) return fail(MD_E_INVALID, "grid must be >= 0");
    if (c->cell_div < 1 || c->cell_div > 3) return fail(MD_E_INVALID, "cell_div must be 1..3");
    if (c->check_every < 1 || c->check_every > kSegment)
        return fail(MD_E_INVALID, "check_every must be 1..%d", kSegment);
    if (c->build_lanes != 4 && c->build_lanes != 8)
        return fail(MD_E_INVALID, "build_lanes must be 4 or 8");
    if (c->submit != 0 && c->submit != 1) return fail(MD_E_INVALID, "submit must be 0 (graphs) or 1 (direct)");
    if (c->refresh != 0 && c->refresh != 1) return fail(MD_E_INVALID, "refresh must be 0 or 1");
    if (c->build_kernel != 0) return fail(MD_E_INVALID, "build_kernel must be 0 (the tile neighbor-set build)");
    if (c->defer != 0 && c->defer != 1) return fail(MD_E_INVALID, "defer must be 0 or 1");
    if (c->vv_fuse != 0 && c->vv_fuse != 1) return fail(MD_E_INVALID, "vv_fuse must be 0 or 1");
    if (c->fm_fused != 0 && c->fm_fused != 1) return fail(MD_E_INVALID, "fm_fused must be 0 or 1");
    if (c->persist != 0 && c->persist != 1) return fail(MD_E_INVALID, "persist must be 0 or 1");
    HIPCHECK(hipSetDevice(e->dev));
    HIPCHECK(hipStreamSynchronize(e->st));
    // the row layout (lpos16) follows the step kernel's lanes
    const bool relayout = c->lanes != e->lanes;
    e->lanes = c->lanes;
    e->grid = c->grid;
    e->check_every = c->check_every;
    e->build_lanes = c->build_lanes;
    e->submit = c->submit;
    e->refresh = c->refresh;
    e->build_kernel = c->build_kernel;
    e->defer = c->defer;
    e->vv_fuse = c->vv_fuse;
    e->fm_fused = c->fm_fused;
    if (c->persist != e->persist || relayout) {   // lane rows are sized per lanes per atom
        e->pk_off = false;
        e->pk_cap = 0;
    }
    e->persist = c->persist;
    e->queue_ratio = md_engine::kQueueRatio0;   // what a step costs in launches depends on the configuration
    e->drop_graphs();
    if (c->cell_div != e->cell_div) {
        e->cell_div = c->cell_div;
        if (e->has_sys) {
            CHECK(setup_cells(e));
            if (e->has_state) CHECK(full_rebuild(e));
        }
    } else if (relayout && e->has_sys && e->has_state) {
        CHECK(full_rebuild(e));
    }
    return MD_OK;
}

int md_get_launch(md_engine* e, md_launch_config* c) {
    if (!e || !c) return fail(MD_E_INVALID, "null argument");
    c->lanes = e->lanes;
    c->grid = e->grid;
    c->cell_div = e->cell_div;
    c->check_every = e->check_every;
    c->build_lanes = e->build_lanes;
    c->submit = e->submit;
    c->refresh = e->refresh;
    c->build_kernel = e->build_kernel;
    c->defer = e->defer;
    c->vv_fuse = e->vv_fuse;
    c->fm_fused = e->fm_fused;
    c->persist = e->persist;
    return MD_OK;
}

int md_persist_info(md_engine* e, long long* runs, long long* replays, int* row_len) {
    if (!e) return fail(MD_E_INVALID, "null engine");
    if (runs) *runs = e->pk_runs;
    if (replays) *replays = e->pk_replays;
    if (row_len) *row_len = (e->has_sys && e->has_ff) ? persist_cap(e) : 0;
    return MD_OK;
}

int md_timing_enable(md_engine* e, int enable) {
    if (!e) return fail(MD_E_INVALID, "null engine");
    e->timing = enable != 0;
    e->t_step_ms = e->t_reb_ms = 0;
    e->n_step = e->n_reb = 0;
    return MD_OK;
}

int md_timing_get(md_engine* e, double* step_ms, long long* step_n, double* reb_ms, long long* reb_n) {
    if (!e) return fail(MD_E_INVALID, "null engine");
    if (step_ms) *step_ms = e->t_step_ms;
    if (step_n) *step_n = e->n_step;
    if (reb_ms) *reb_ms = e->t_reb_ms;
    if (reb_n) *reb_n = e->n_reb;
    return MD_OK;
}

int md_autotune(md_engine* e, int nsteps, double* best_ms) {
    CHECK(check_ready(e, true));
    CHECK(pme_prepare(e));
    if (nsteps < 1) return fail(MD_E_INVALID, "nsteps must be >= 1");
    HIPCHECK(hipSetDevice(e->dev));
    const bool saved_timing = e->timing;
    const double saved_t[2] = {e->t_step_ms, e->t_reb_ms};
    const long long saved_n[2] = {e->n_step, e->n_reb};
    hipEvent_t t0 = ev_get(e, 0), t1 = ev_get(e, 1);
    // the physical state is kept: every trial starts from it
    const int N = e->N;
    std::vector<float> P(3 * (size_t)N), V(3 * (size_t)N);
    std::vector<ThState> th(e->R);
    const long long step0 = e->step;
    CHECK(md_download_state(e, P.data(), V.data(), nullptr));
    HIPCHECK(hipMemcpy(th.data(), e->th.p, e->R * sizeof(ThState), hipMemcpyDeviceToHost));
    auto restore = [&]() -> int {
        e->step = step0;
        CHECK(md_upload_state(e, P.data(), V.data()));
        HIPCHECK(hipMemcpy(e->th.p, th.data(), e->R * sizeof(ThState), hipMemcpyHostToDevice));
        return MD_OK;
    };
    md_launch_config best;
    CHECK(md_get_launch(e, &best));
    // stage 1, cell grid x lanes per atom: the cell edge sets the build's candidate
    // count, the lanes the step kernel's shape; build time (three forced rebuilds)
    // plus step-kernel time (HIP events) of each pair
    struct Shape { double ms; int cell_div, lanes; };
    std::vector<Shape> shapes;
    const int ks = std::max(8, std::min(nsteps, 64));
    for (int d = 1; d <= 3; d++)
        for (int l : {4, 8, 16}) {
            md_launch_config c = best;
            c.cell_div = d;
            c.lanes = l;
            c.persist = 0;
            if (md_set_launch(e, &c) != MD_OK) continue;   // e.g. a tile set beyond the LDS
            CHECK(restore());
            HIPCHECK(hipEventRecord(t0, e->st));
            for (int k = 0; k < 3; k++) {
                HIPCHECK(md_launch_force_trip(e->rst.p, e->gate.p, e->R, e->st));
                CHECK(enqueue_rebuild(e, nullptr));
            }
            HIPCHECK(hipEventRecord(t1, e->st));
            HIPCHECK(hipEventSynchronize(t1));
            float bms = 0;
            HIPCHECK(hipEventElapsedTime(&bms, t0, t1));
            CHECK(restore());
            CHECK(md_timing_enable(e, 1));
            CHECK(md_run(e, ks, 0, nullptr));
            const double sms = e->t_step_ms / std::max(1LL, e->n_step);
            CHECK(md_timing_enable(e, 0));
            // a build every ~8 steps
            shapes.push_back({sms + bms / 3.0 / 8.0, d, l});
        }
    if (shapes.empty()) return fail(MD_E_CAPACITY, "no launch shape fits this system");
    std::sort(shapes.begin(), shapes.end(), [](const Shape& x, const Shape& y) { return x.ms < y.ms; });
    shapes.resize(std::min<size_t>(shapes.size(), 3));
    // the tile build's wavefronts per tile (the same lists: build time only), at the best shape
    {
        double bbest = 1e30;
        for (int nw : {4, 8}) {
            md_launch_config c = best;
            c.cell_div = shapes[0].cell_div;
            c.lanes = shapes[0].lanes;
            c.build_lanes = nw;
            c.persist = 0;
            if (md_set_launch(e, &c) != MD_OK) continue;
            CHECK(restore());
            HIPCHECK(hipEventRecord(t0, e->st));
            for (int k = 0; k < 3; k++) {
                HIPCHECK(md_launch_force_trip(e->rst.p, e->gate.p, e->R, e->st));
                CHECK(enqueue_rebuild(e, nullptr));
            }
            HIPCHECK(hipEventRecord(t1, e->st));
            HIPCHECK(hipEventSynchronize(t1));
            float bms = 0;
            HIPCHECK(hipEventElapsedTime(&bms, t0, t1));
            if (bms < bbest) {
                bbest = bms;
                best.build_lanes = nw;
            }
        }
    }
    // stage 2, whole runs (best of two): launch mode (age refresh, submission, deferred
    // finalize), then shape x rebuild-check interval.  A trial times runs of nsteps steps each (the
    // length of the caller's md_run calls: a short run pays its first rebuild chain and
    // waits for the slowest replica, which long runs amortise), at least 4 segments' worth.
    double best_run = 1e30;
    const long long total = std::max<long long>(nsteps, 4LL * kSegment);
    md_launch_config base = best;
    base.persist = 0;
    const bool lf = lf_form(e);
    // per-step time of configuration c: the best of `reps` trials, each a sequence of
    // md_run(nsteps) calls covering `steps` steps after a warm-up run
    auto time_run = [&](const md_launch_config& c, long long steps, int reps, double& per) -> int {
        CHECK(md_set_launch(e, &c));
        per = 1e30;
        for (int rep = 0; rep < reps; rep++) {
            CHECK(restore());
            CHECK(md_run(e, std::min(nsteps, kSegment), 0, nullptr));   // warm up (graph capture, queue ratio)
            long long done = 0;
            HIPCHECK(hipEventRecord(t0, e->st));
            while (done < steps) {
                CHECK(md_run(e, nsteps, 0, nullptr));
                done += nsteps;
            }
            HIPCHECK(hipEventRecord(t1, e->st));
            HIPCHECK(hipEventSynchronize(t1));
            float ms = 0;
            HIPCHECK(hipEventElapsedTime(&ms, t0, t1));
            per = std::min(per, (double)ms / (double)done);
        }
        return MD_OK;
    };
    struct Trial { double ms; md_launch_config c; };
    std::vector<Trial> trials;
    // With PME a step launched after a skin trip (and skipped until the next
    // rebuild chain) still runs its FFTs: short check intervals pay off there.
    const std::vector<int> ces = e->pme_on ? std::vector<int>{1, 2, 3, 4, 6, 8}
                                           : std::vector<int>{3, 4, 6, 8, 12, 16, 24, 32};
    auto trial = [&](const Shape& sh, int sub, int rf, int df, int ce, std::vector<Trial>& into) -> int {
        md_launch_config c = base;
        c.cell_div = sh.cell_div;
        c.lanes = sh.lanes;
        c.check_every = ce;
        c.submit = sub;
        c.refresh = rf;
        c.defer = df;
        double per = 1e30;
        CHECK(time_run(c, total, 2, per));
        into.push_back({per, c});
        return MD_OK;
    };
    // 2a: submission mode x age refresh x deferred finalize at the best shape and a
    // middle check interval; the two best modes go on
    std::vector<Trial> modes;
    for (int sub : {0, 1})
        for (int rf : {0, 1})
            for (int df : {1, 0}) {
                if (df == 0 && !lf) continue;              // deferral applies to leapfrog-form steps only
                if (df == 1 && e->pme_on && lf) continue;  // PME steps finalize in their own kernel
                CHECK(trial(shapes[0], sub, rf, df, ces[ces.size() / 2], modes));
            }
    std::sort(modes.begin(), modes.end(), [](const Trial& x, const Trial& y) { return x.ms < y.ms; });
    trials = modes;
    modes.resize(std::min<size_t>(modes.size(), 2));
    // 2b: the check interval for the two best shapes and modes
    for (size_t k = 0; k < std::min<size_t>(shapes.size(), 2); k++)
        for (const Trial& m : modes)
            for (int ce : ces) {
                if (k == 0 && ce == ces[ces.size() / 2]) continue;   // timed in 2a
                CHECK(trial(shapes[k], m.c.submit, m.c.refresh, m.c.defer, ce, trials));
            }
    // the persistent kernel (small single systems), one trial per lanes per atom
    if (persist_eligible(e))
        for (int l : {4, 8, 16}) {
            md_launch_config c = base;
            c.cell_div = shapes[0].cell_div;
            c.lanes = l;
            c.persist = 1;
            double per = 1e30;
            CHECK(time_run(c, total, 2, per));
            trials.push_back({per, c});
        }
    // stage 3: the four fastest again, longer and three times each (the minimum
    // over many short noisy trials favours a lucky one)
    std::sort(trials.begin(), trials.end(), [](const Trial& x, const Trial& y) { return x.ms < y.ms; });
    trials.resize(std::min<size_t>(trials.size(), 4));
    for (const Trial& t : trials) {
        double per = 1e30;
        CHECK(time_run(t.c, 2 * total, 3, per));
        if (per < best_run) {
            best_run = per;
            best = t.c;
        }
    }
    CHECK(md_set_launch(e, &best));
    CHECK(restore());
    e->timing = saved_timing;
    e->t_step_ms = saved_t[0];
    e->t_reb_ms = saved_t[1];
    e->n_step = saved_n[0];
    e->n_reb = saved_n[1];
    if (best_ms) *best_ms = best_run;
    return MD_OK;
}

int md_rebuild_bench(md_engine* e, int reps, double* ms) {
    CHECK(check_ready(e, true));
    if (reps < 1) return fail(MD_E_INVALID, "reps must be >= 1");
    HIPCHECK(hipSetDevice(e->dev));
    CHECK(full_rebuild(e));   // settles the capacity
    hipEvent_t t0 = ev_get(e, 0), t1 = ev_get(e, 1);
    HIPCHECK(hipEventRecord(t0, e->st));
    for (int k = 0; k < reps; k++) {
        HIPCHECK(md_launch_force_trip(e->rst.p, e->gate.p, e->R, e->st));
        CHECK(enqueue_rebuild(e, nullptr));
    }
    HIPCHECK(hipEventRecord(t1, e->st));
    HIPCHECK(hipEventSynchronize(t1));
    float t = 0;
    HIPCHECK(hipEventElapsedTime(&t, t0, t1));
    if (ms) *ms = t / reps;
    return MD_OK;
}

int md_copy_bench(md_engine* e, long long nbytes, int reps, double* ms) {
    if (!e || nbytes <= 0 || reps < 1) return fail(MD_E_INVALID, "bad arguments");
    HIPCHECK(hipSetDevice(e->dev));
    void *a = nullptr, *b = nullptr;
    HIPCHECK(hipMalloc(&a, nbytes));
    HIPCHECK(hipMalloc(&b, nbytes));
    HIPCHECK(hipMemsetAsync(a, 1, nbytes, e->st));
    hipEvent_t t0 = ev_get(e, 0), t1 = ev_get(e, 1);
    HIPCHECK(hipMemcpyAsync(b, a, nbytes, hipMemcpyDeviceToDevice, e->st));
    HIPCHECK(hipEventRecord(t0, e->st));
    for (int k = 0; k < reps; k++) HIPCHECK(hipMemcpyAsync(b, a, nbytes, hipMemcpyDeviceToDevice, e->st));
    HIPCHECK(hipEventRecord(t1, e->st));
    HIPCHECK(hipEventSynchronize(t1));
    float t = 0;
    HIPCHECK(hipEventElapsedTime(&t, t0, t1));
    (void)hipFree(a);
    (void)hipFree(b);
    if (ms) *ms = t / reps;
    return MD_OK;
}

// ---------------------------------------------------------------- list invariants
int md_check_lists(md_engine* e, int* counts) {
    CHECK(check_ready(e, true));
    if (!counts) return fail(MD_E_INVALID, "null counts");
    HIPCHECK(hipSetDevice(e->dev));
    HIPCHECK(hipMemsetAsync(e->badcnt.p, 0, e->badcnt.bytes(), e->st));
    HIPCHECK(md_launch_check_lists(build_args(e), e->lanes, e->badcnt.p, e->st));
    int bad[8];
    HIPCHECK(hipMemcpyAsync(bad, e->badcnt.p, sizeof bad, hipMemcpyDeviceToHost, e->st));
    HIPCHECK(hipStreamSynchronize(e->st));
    for (int k = 0; k < MD_NCHECK; k++) counts[k] = bad[k];
    return MD_OK;
}

// ---------------------------------------------------------------- force matching
int md_fm_set_frames(md_engine* e, int nframes, const float* pos, const float* target_force) {
    CHECK(check_ready(e, false));
    if (nframes < 1 || !pos || !target_force) return fail(MD_E_INVALID, "bad frames");
    if (e->fm_nframes == 0 && e->R != 1) return fail(MD_E_STATE, "force matching needs a single-replica system");
    const int n = e->natom[0];
    // frames become replicas of the base system (replica 0)
    std::vector<int> na(nframes, n), ty((size_t)nframes * n);
    std::vector<float> bx(3 * (size_t)nframes), t0(nframes, e->T0[0]), q((size_t)nframes * n), m((size_t)nframes * n);
    std::vector<int> base_t(e->type_h.begin(), e->type_h.begin() + n);
    std::vector<float> base_q(e->charge_h.begin(), e->charge_h.begin() + n), base_m(e->mass_h.begin(), e->mass_h.begin() + n);
    float b0[3] = {e->box[0], e->box[1], e->box[2]};
    for (int f = 0; f < nframes; f++) {
        for (int d = 0; d < 3; d++) bx[3 * f + d] = b0[d];
        for (int i = 0; i < n; i++) {
            ty[(size_t)f * n + i] = base_t[i];
            q[(size_t)f * n + i] = base_q[i];
            m[(size_t)f * n + i] = base_m[i];
        }
    }
    CHECK(md_set_system(e, nframes, na.data(), bx.data(), t0.data(), ty.data(), q.data(), m.data()));
    e->step = 0;
    CHECK(md_upload_state(e, pos, nullptr));
    // targets in slot order
    std::vector<int> so(e->N);
    HIPCHECK(hipMemcpy(so.data(), e->slot_of.p, e->N * sizeof(int), hipMemcpyDeviceToHost));
    std::vector<float4> T(e->S, make_float4(0, 0, 0, 0));
    for (int i = 0; i < e->N; i++)
        T[so[i]] = make_float4(target_force[3 * (size_t)i], target_force[3 * (size_t)i + 1], target_force[3 * (size_t)i + 2], 0.f);
    HIPCHECK(hipMemcpy(e->target.p, T.data(), e->S * sizeof(float4), hipMemcpyHostToDevice));
    e->fm_nframes = nframes;
    return MD_OK;
}

// One loss-and-gradient evaluation with the given device buffers: ljtab_d (64
// float4, the trial table), part_d (ntiles x nacc doubles, per-tile slots) and
// acc_d (nacc doubles, the sums); nacc = 2 + 2 ntk^2.  Every slot of part_d and
// acc_d is written before it is read, so nothing needs clearing.
static int fm_evaluate(md_engine* e, const float* lj1, const float* lj2, float qscale, float4* ljtab_d, double* part_d,
                       double* acc_d, double* loss, double* g_lj1, double* g_lj2, double* g_qscale) {
    const int nt = e->ff.ntype;
    const int ntk = nt == 1 ? 1 : (nt == 2 ? 2 : 4);
    // trial LJ table in its own buffer: the engine's table is left alone, and the
    // upload is ordered on the stream (one synchronisation per evaluation, at the end)
    HIPCHECK(hipStreamSynchronize(e->st));   // the staging copy of the previous evaluation is free
    make_ljtab(e, lj1, lj2, e->fm_ljtab_h);
    HIPCHECK(hipMemcpyAsync(ljtab_d, e->fm_ljtab_h, 64 * sizeof(float4), hipMemcpyHostToDevice, e->st));
    StepArgs a = step_args(e, MODE_FORCE, 0);
    a.ljtab = ljtab_d;
    a.stats = nullptr;
    const int coul = e->coul_kernel;
    hipEvent_t ev0 = e->timing ? ev_get(e, 0) : nullptr, ev1 = e->timing ? ev_get(e, 1) : nullptr;
    if (e->timing && (!ev0 || !ev1)) return fail(MD_E_HIP, "hipEventCreate failed");
    if (e->fm_fused && !e->bonded) {   // one pass: force and gradient sums together
        a.qq_scale = qscale;
        HIPCHECK(md_launch_fm_fused(a, part_d, acc_d, e->lanes, coul, ntk, e->st, ev0, ev1));
    } else {                            // force pass (bonded terms included), then the adjoint walk
        a.qq_scale = qscale * qscale;
        HIPCHECK(md_launch_step(a, e->lanes, coul, e->st, nullptr, nullptr));
        a.qq_scale = qscale;
        HIPCHECK(md_launch_fm_adjoint(a, part_d, acc_d, e->lanes, coul, ntk, e->st, ev0, ev1));
    }
    double g[2 + 2 * 16];
    const int nacc = 2 + 2 * ntk * ntk;
    HIPCHECK(hipMemcpyAsync(g, acc_d, nacc * sizeof(double), hipMemcpyDeviceToHost, e->st));
    HIPCHECK(hipStreamSynchronize(e->st));
    if (e->timing) {   // the per-tile kernel's duration counts as this evaluation's "step"
        float ms = 0;
        HIPCHECK(hipEventElapsedTime(&ms, ev0, ev1));
        e->t_step_ms += ms;
        e->n_step++;
    }
    const double M = 3.0 * (double)e->N;
    if (loss) *loss = g[0] / M;
    if (g_qscale) *g_qscale = 2.0 / M * g[1];
    for (int i = 0; i < nt; i++)
        for (int j = 0; j < nt; j++) {
            if (g_lj1) g_lj1[i * nt + j] = 2.0 / M * g[2 + i * ntk + j];
            if (g_lj2) g_lj2[i * nt + j] = 2.0 / M * g[2 + ntk * ntk + i * ntk + j];
        }
    return MD_OK;
}

static int fm_check(md_engine* e, const float* lj1, const float* lj2) {
    CHECK(check_ready(e, true));
    if (e->pme_order) return fail(MD_E_INVALID, "force matching has no PME term: set the order to 0");
    if (!e->fm_nframes) return fail(MD_E_STATE, "call md_fm_set_frames first");
    if (!lj1 || !lj2) return fail(MD_E_INVALID, "null LJ tables");
    if (e->ff.ntype > 4) return fail(MD_E_CAPACITY, "force matching supports at most 4 atom types");
    HIPCHECK(hipSetDevice(e->dev));
    return MD_OK;
}

static int fm_nacc(const md_engine* e) {
    const int nt = e->ff.ntype;
    const int ntk = nt == 1 ? 1 : (nt == 2 ? 2 : 4);
    return 2 + 2 * ntk * ntk;
}

int md_fm_loss_grad(md_engine* e, const float* lj1, const float* lj2, float qscale, double* loss, double* g_lj1,
                    double* g_lj2, double* g_qscale) {
    CHECK(fm_check(e, lj1, lj2));
    if (!e->fm_ljtab.p) CHECK(e->fm_ljtab.alloc(64));
    const size_t nslot = (size_t)e->ntiles() * fm_nacc(e);
    if (e->fm_part.n < nslot) CHECK(e->fm_part.alloc(nslot));
    return fm_evaluate(e, lj1, lj2, qscale, e->fm_ljtab.p, e->fm_part.p, e->fm_acc.p, loss, g_lj1, g_lj2, g_qscale);
}

int md_fm_loss_grad_noalloc(md_engine* e, const float* lj1, const float* lj2, float qscale, double* loss,
                            double* g_lj1, double* g_lj2, double* g_qscale) {
    CHECK(fm_check(e, lj1, lj2));
    void* base = e->dev < kMaxDevices ? g_scratch_ptr[e->dev] : nullptr;
    if (!base) return fail(MD_E_STATE, "no scratch address for device %d", e->dev);
    const size_t need = (size_t)e->ntiles() * fm_nacc(e) * sizeof(double);
    if (need > MD_FM_SCRATCH_BYTES - kScratchPart)
        return fail(MD_E_CAPACITY, "force matching over %d tiles needs %zu bytes of per-tile sums; the fixed scratch holds %zu",
                    e->ntiles(), need, (size_t)(MD_FM_SCRATCH_BYTES - kScratchPart));
    char* b = static_cast<char*>(base);
    return fm_evaluate(e, lj1, lj2, qscale, reinterpret_cast<float4*>(b + kScratchLj),
                       reinterpret_cast<double*>(b + kScratchPart), reinterpret_cast<double*>(b + kScratchAcc), loss,
                       g_lj1, g_lj2, g_qscale);
}

}  // extern "C"
