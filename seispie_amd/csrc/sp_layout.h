// seispie_hip -- device data layout: state slots, shot descriptors, step
// arguments, cross-lane rotations.
// Part of seispie_hip.hip: included there, inside its anonymous namespace,
// after <hip/hip_runtime.h> and the headers before this one.
#pragma once

// ------------------------------------------------------------------ layout
// Stress, velocity and displacement fields per physics.
//   SH:   s = {sxy, szy}                         v = {vy}          u = {uy}
//   P-SV: s = {sxx, szz, sxz}                    v = {vx, vz}      u = {ux, uz}
//   spin: s = {sxx, szz, sxz, syx_c, syy_c, syz_c} v = {vx, vz, vy_c} u = {ux, uz, uy_c}
// State slots of a shot: two stress sets, two velocity sets, one u set.
__host__ __device__ constexpr int n_s(int p) { return p == SP_SH ? 2 : (p == SP_PSV ? 3 : 6); }
__host__ __device__ constexpr int n_v(int p) { return p == SP_SH ? 1 : (p == SP_PSV ? 2 : 3); }
__host__ __device__ constexpr int n_slots(int p) { return 2 * n_s(p) + 3 * n_v(p); }
__host__ __device__ constexpr int slot_s(int p, int set, int f) { return set * n_s(p) + f; }
__host__ __device__ constexpr int slot_v(int p, int set, int f) { return 2 * n_s(p) + set * n_v(p) + f; }
__host__ __device__ constexpr int slot_u(int p, int f) { return 2 * n_s(p) + 2 * n_v(p) + f; }
// snapshot fields: u (before the step) then v (after it), solver.py:193-242
__host__ __device__ constexpr int n_snap(int p) { return 2 * n_v(p); }
// components of the records (SH y; P-SV x, z; spin x, z, y_c), of the
// injected sources (SH y; P-SV/spin x, z) and of the misfit.
__host__ __device__ constexpr int n_rec(int p) { return n_v(p); }
__host__ __device__ constexpr int n_inj(int p) { return p == SP_SH ? 1 : 2; }
__host__ __device__ constexpr int n_mis(int p) { return p == SP_SH ? 1 : 2; }
// material slots of a grid
constexpr int M_RHO = SP_F_RHO, M_LAM = SP_F_LAM, M_MU = SP_F_MU, M_NU = SP_F_NU, M_J = SP_F_J;
constexpr int M_GAM = 8;  // float32(mu_c + nu_c), as the reference adds them (solver.py:140)
constexpr int N_MAT = 9;

struct ShotDesc {
    const float *mat;   // [N_MAT][Npad]
    const float *bx;    // [nx] x-side sponge factors (already float32-rounded)
    const double *bzb;  // [nz] bottom factor or 1
    const double *bzt;  // [nz] top factor or 1
    float *state;       // [n_slots][Npad]
    float *rec;         // [n_rec][nrec][nt]
    float *snap;        // [nsa][n_snap][Npad] or null
    const float *stf;   // [n_inj][nsrc][nt]
    float *adstf;       // [n_mis][nrec][nt]
    const int *src_k, *src_row, *rec_k, *rec_row;  // sorted by k
    int nsrc, nrec, nx, nz;
    long long Npad;
    double c1x, c2x, c1z, c2z;  // 9/(8h), 1/(24h)
    // exact mode: 9/(8h) and 1/(24h) as unevaluated sums hi + lo of two doubles
    double k9hx, k9lx, k1hx, k1lx, k9hz, k9lz, k1hz, k1lz;
    const int *src_i, *rec_i;  // x row (k / nz) of each sorted source / receiver
};

// read-only, wave-uniform data through the constant address space: scalar
// loads, counted by lgkmcnt and never by the vector memory counter
typedef const int __attribute__((address_space(4))) *cint_p;
typedef const float __attribute__((address_space(4))) *cflt_p;
// a shot descriptor read the same way: its fields stay scalar loads even
// where the kernel has stored before (in a branch of a row loop)
typedef const __attribute__((address_space(4))) struct ShotDesc cdesc_t;

// The injection lists of a launch: sources (forward) or the receivers with
// their adjoint sources (adjoint), sorted by grid index
struct InjLists {
    cint_p k, row, xi;  // grid index, row of the time function, x row
    cflt_p val;         // time functions [n_inj][nrows][nt]
    int n;
    long long nrows;
};
__device__ __forceinline__ InjLists inj_lists(cdesc_t *dp, int adjoint)
{
    InjLists L;
    L.k = (cint_p)(adjoint ? dp->rec_k : dp->src_k);
    L.row = (cint_p)(adjoint ? dp->rec_row : dp->src_row);
    L.xi = (cint_p)(adjoint ? dp->rec_i : dp->src_i);
    L.val = (cflt_p)(adjoint ? (const float *)dp->adstf : dp->stf);
    L.n = adjoint ? dp->nrec : dp->nsrc;
    L.nrows = L.n;
    return L;
}
// A descriptor pointer the optimiser cannot follow from one row to the
// next: what a row loop reads through it in a rarely taken branch (list
// bases, record buffer) is loaded there, from the scalar cache, instead of
// being held in (and spilled from) SGPRs for the whole march.
__device__ __forceinline__ cdesc_t *fresh(cdesc_t *dp)
{
    asm volatile("" : "+s"(dp));
    return dp;
}

// Upcoming injection entries in a slice of LDS per wavefront, filled
// before the march: word l of each of its four 64-word rows holds list
// entry base + l (grid index, x row, time-function values 0 and 1 at the
// queue's step), so a row that injects reads them with broadcast LDS
// reads.  Read from the lists in that row instead, they are a chain of
// dependent scalar loads (entry, then its time-function row, then the
// value) that the march waits for in every row holding an entry -- with
// adjoint sources at receivers every second row, 14% of a spin launch.
// The host launches the queue instances only when no work item reaches
// more than IQ_CAP entries, so the march holds no refill: code that
// refilled the slice, or read the lists past it, inside the march cost
// 3-5% of every launch, injections or not.  The fill loads only the rounds
// of 64 entries the item reaches (the count comes with the work item): a
// fill is latency at the start of the march, 2-4% of a 26-row one.  row is
// the next entry's x row (0x7fffffff: none left).
constexpr int IQ_CAP = 128;               // entries a queue holds
// modes of a two-step launch (template parameter M of the kernels)
constexpr int STEP_SNAP = 0, STEP_PLAIN = 1, STEP_QUEUE = 2;
constexpr int IQ_WORDS = 4 * IQ_CAP;      // [k | x row | value 0 | value 1][IQ_CAP]
struct InjQ {
    int *s;  // the queue's LDS slice
    int l, row;  // the next entry's word, its x row
};
// entries from, from + 1, ... at time step `it` into slice s, in rounds of
// 64: the `count` entries the work item reaches (at most IQ_CAP) and the
// one after them, whose row ends the item's injections
template <int P>
__device__ __forceinline__ void injq_fill(InjQ &q, int *s, cdesc_t *dp, int adjoint, int nt, int it, int from,
                                          int count)
{
    const InjLists L = inj_lists(dp, adjoint);
    const int lane = (int)(threadIdx.x & 63);
    int xi0 = 0x7fffffff;
    const int fill = min(count + 1, IQ_CAP);
    for (int h = 0; h < fill; h += 64) {
        const int l = h + lane, e = from + l;
        int k = 0x7fffffff, xi = 0x7fffffff;
        float v0 = 0.f, v1 = 0.f;
        if (e < L.n) {
            k = L.k[e];
            xi = L.xi[e];
            const long long rw = L.row[e];
            v0 = L.val[rw * nt + it];
            if constexpr (P != SP_SH) v1 = L.val[(L.nrows + rw) * nt + it];
        }
        s[l] = k;
        s[IQ_CAP + l] = xi;
        s[2 * IQ_CAP + l] = __float_as_int(v0);
        if constexpr (P != SP_SH) s[3 * IQ_CAP + l] = __float_as_int(v1);
        if (h == 0) xi0 = xi;
    }
    q.s = s;
    q.l = 0;
    q.row = __builtin_amdgcn_readlane(xi0, 0);
}
// the entries of x row r, in list order (sources: stf_dsy / stf_dsxz,
// solver.py:78-93): apply(k, value0, value1) for each.  Every lane reads
// the same LDS word, which a lane of the same wavefront wrote.
template <int P, class F>
__device__ __forceinline__ void injq_row(InjQ &q, int r, F &&apply)
{
    if (r != q.row) return;
    do {
        __builtin_amdgcn_wave_barrier();
        const int kk = q.s[q.l];
        const float s0 = __int_as_float(q.s[2 * IQ_CAP + q.l]);
        float s1 = 0.f;
        if constexpr (P != SP_SH) s1 = __int_as_float(q.s[3 * IQ_CAP + q.l]);
        apply(kk, s0, s1);
        // the host guarantees no item reaches entry IQ_CAP; the fill covers
        // the entry after the item's last one
        q.row = ++q.l < IQ_CAP ? __builtin_amdgcn_readfirstlane(q.s[IQ_CAP + q.l]) : 0x7fffffff;
    } while (q.row == r);
}

struct StepArgs {
    const ShotDesc *shots;
    const int4 *items;  // per wavefront: (shot, j0, i0, i1) and the first source with k >= (i0-4) nz,
                        // the first receiver with k >= (i0-4) nz, the first with k >= i0 nz and
                        // the receivers in rows i0-4 .. i1+4 (at most IQ_CAP)
    int nitems;
    int it, nt;         // step of this launch (plus *itp when itp is set)
    const int *itp;     // device step counter of a replayed graph, or null
    double dt;
    int rd;        // stress/velocity set read this step
    int snap_idx;  // -1: no snapshot this step (one-step kernel k_step)
    // two-step kernel: the snapshot of the pair's first step (snap_v) and
    // of its second step (snap_u), -1: none; the kernel writes both halves
    // of each (u before the step, v after it)
    int snap_v, snap_u;
    // replayed graphs: per step of the run, the snapshot slot taken after
    // it or -1 ([nt + 1]; null: snap_v / snap_u as given)
    const int *snap_tab;
    int adjoint;   // inject adjoint sources at the receivers, record nothing
    int shift;     // cells of a strip before its first output cell (>= 3)
    int wout;      // output cells per strip (shift + wout <= 61; two cells per lane: 6 + 116)
};

// z neighbours by whole-wavefront DPP rotations: lane L receives the value
// of lane L-1 (lane_m1) or L+1 (lane_p1), lane 0 / 63 wrapping to 63 / 0 as
// a 64-wide shuffle does (checked on gfx950).  One VALU move each, where a
// shuffle through the LDS crossbar waits on the LDS queue.  Every caller
// runs them with all 64 lanes active.
__device__ __forceinline__ float lane_m1(float v)
{
    return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), 0x13C /* wave_ror:1 */, 0xf, 0xf, false));
}
__device__ __forceinline__ float lane_p1(float v)
{
    return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), 0x134 /* wave_rol:1 */, 0xf, 0xf, false));
}
