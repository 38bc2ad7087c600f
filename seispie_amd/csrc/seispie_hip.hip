// seispie_hip -- MI355X (gfx950) solver hot path of seispie's fd2d plugin.
//
// What the reference does (seispie/solver/fd2d/solver.py) is a Python time
// loop that launches 6 (SH), 7 (P-SV) or 13 (spin) Numba kernels per step,
// one thread per cell, each kernel streaming whole arrays through HBM.
// Here one launch advances every shot of a batch by one or two time steps:
//
//   * k_step2<PHYS> (two steps per launch, the default wherever the state
//     between the steps is not needed), k_step3<PHYS> (one step; the same
//     march with buffer loads and stores) and k_step<PHYS> (plain one-step
//     fallback for grids too large for 32-bit buffer offsets).  In each,
//     a wavefront owns a strip of 64 consecutive z cells
//     (z is the contiguous axis, k = i*nz + j) and marches along x through
//     a chunk of rows.  It keeps a 4-row register window of the old
//     stresses and of the new velocities, takes z neighbours with
//     wavefront (DPP) rotations, and writes new velocity, displacement and stress
//     for the 58 interior lanes.  Stress divergence, source injection,
//     sponge, velocity/displacement update, velocity gradient, stress
//     update, receiver recording and adjoint snapshots are all fused, so a
//     cell's state is read and written once per launch.  Stresses and
//     velocities ping-pong between two buffers because neighbouring strips
//     read the old values in their halo.
//   * k_interact<PHYS>: kernel accumulation every sae adjoint steps.
//   * k_misfit, k_smooth_x/z, k_vps2lm, ...: the smaller pieces.
//
// Numerics follow the reference: float32 storage, arithmetic in float64
// wherever Numba promotes (int literal * float32, Python-float dt) and the
// few float32 products it keeps, rounded to float32 at the points where
// the reference stores an array.  By default (SP_OPT_EXACT = 1) every
// float64 operation is the reference's, in its order and with one rounding
// each: divisions by 8h and 24h are products with a two-double split of the
// quotient that the host verifies to round identically for every float32
// input, divisions by rho use a correctly rounded reciprocal and one
// correction step, and nothing is contracted into fused multiply-adds.
// SP_OPT_EXACT = 0 multiplies by reciprocals and contracts, which moves
// single results by one float32 rounding step on rare cells.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <map>
#include <mutex>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <numeric>
#include <string>
#include <vector>

#include "../../include/seispie_hip.h"

namespace {

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
};

// read-only, wave-uniform data through the constant address space: scalar
// loads, counted by lgkmcnt and never by the vector memory counter
typedef const int __attribute__((address_space(4))) *cint_p;
typedef const float __attribute__((address_space(4))) *cflt_p;

struct StepArgs {
    const ShotDesc *shots;
    const int4 *items;  // per wavefront: (shot, j0, i0, i1) and the first source with k >= (i0-4) nz,
                        // the first receiver with k >= (i0-4) nz and the first with k >= i0 nz
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
    int adjoint;   // inject adjoint sources at the receivers, record nothing
    int shift;     // cells of a strip before its first output cell (>= 3)
    int wout;      // output cells per strip (shift + wout <= 61; k_step2w: 6 + 116)
};

// float32 operations the reference performs in float32 (float32 * float32
// stays float32 under Numba's typing); contraction off so they are never
// folded into a float64 multiply-add.
__device__ __forceinline__ float mul_rn(float a, float b)
{
#pragma clang fp contract(off)
    return a * b;
}
__device__ __forceinline__ float add_rn(float a, float b)
{
#pragma clang fp contract(off)
    return a + b;
}
__device__ __forceinline__ float sub_rn(float a, float b)
{
#pragma clang fp contract(off)
    return a - b;
}

// D- at row r from values at r-2, r-1, r, r+1 (solver.py:22-27, 38-43)
__device__ __forceinline__ double dminus(float m2, float m1, float c0, float p1, double c1, double c2)
{
    return c1 * (double)sub_rn(c0, m1) - c2 * (double)sub_rn(p1, m2);
}
// D+ at row q from values at q-1, q, q+1, q+2 (solver.py:30-35, 46-51)
__device__ __forceinline__ double dplus(float m1, float c0, float p1, float p2, double c1, double c2)
{
    return c1 * (double)sub_rn(p1, c0) - c2 * (double)sub_rn(p2, m1);
}

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

// ------------------------------------------------------------ step kernel
// Global-address-space pointers: loads through them become global_load with
// a scalar base and one 32-bit offset per row shared by every field.
typedef const float __attribute__((address_space(1))) *gcf;
typedef float __attribute__((address_space(1))) *gf;

// 1/x in float64 to within an ulp: hardware estimate + two Newton steps
__device__ __forceinline__ double rcp64(double x)
{
    double r = __builtin_amdgcn_rcp(x);
    r = fma(r, fma(-x, r, 1.0), r);
    return fma(r, fma(-x, r, 1.0), r);
}

// ---- exact mode: the reference's float64 operations, one rounding each, in
// its order (no contraction), so results equal the oracle bit for bit.
// a / c correctly rounded from rc = RN(1/c): q = RN(a rc), the remainder
// a - q c is exact by FMA, and RN(q + r rc) is the correctly rounded quotient
// (Markstein's correction step; also checked against IEEE division on
// 2e8 random operands of these ranges).
__device__ __forceinline__ double rn_div(double a, double c, double rc)
{
#pragma clang fp contract(off)
    const double q = a * rc;
    const double r = __builtin_fma(-q, c, a);
    return __builtin_fma(r, rc, q);
}
// RN(1/x) for float32 x: three Newton steps from the hardware estimate.  Two
// already give the correctly rounded reciprocal of every float32 mantissa
// from any start within 2^-20 (exhaustive check); the third keeps it so from
// starts as far off as 2^-14.
__device__ __forceinline__ double rcp_rn(double x)
{
    double r = __builtin_amdgcn_rcp(x);
    r = __builtin_fma(r, __builtin_fma(-x, r, 1.0), r);
    r = __builtin_fma(r, __builtin_fma(-x, r, 1.0), r);
    return __builtin_fma(r, __builtin_fma(-x, r, 1.0), r);
}
// RN(num x / den) as fma(x, hi, x lo) with num/den split into hi + lo (the
// host checks the identity for every float32 mantissa of x)
__device__ __forceinline__ double kmul(double x, double hi, double lo)
{
    return __builtin_fma(x, hi, x * lo);
}
// D- and D+ as the reference writes them: 9 (a) / (8h) - (b) / (24h)
__device__ __forceinline__ double dminus_x(float m2, float m1, float c0, float p1, double k9h, double k9l,
                                           double k1h, double k1l)
{
#pragma clang fp contract(off)
    return kmul((double)sub_rn(c0, m1), k9h, k9l) - kmul((double)sub_rn(p1, m2), k1h, k1l);
}
__device__ __forceinline__ double dplus_x(float m1, float c0, float p1, float p2, double k9h, double k9l,
                                          double k1h, double k1l)
{
#pragma clang fp contract(off)
    return kmul((double)sub_rn(p1, c0), k9h, k9l) - kmul((double)sub_rn(p2, m1), k1h, k1l);
}
// add_vy / add_vxz: v = bound (v + dt ds / rho), then u = u + v dt
__device__ __forceinline__ float vel_x(float vo, float ds, double dt, double rho, double irho, double b)
{
#pragma clang fp contract(off)
    return (float)(b * ((double)vo + rn_div(dt * (double)ds, rho, irho)));
}
__device__ __forceinline__ float disp_x(float u, float v, double dt)
{
#pragma clang fp contract(off)
    return (float)((double)u + (double)v * dt);
}
// s + f * g, one rounding per operation (add_sy and the spin couple stresses)
__device__ __forceinline__ float sadd_x(float s, double f, double g)
{
#pragma clang fp contract(off)
    return (float)((double)s + f * g);
}
// add_sxz: s + dt (l2m d1 + float32(lam d2))
__device__ __forceinline__ float snorm_x(float s, double dt, double l2m, float d1, float lam_d2)
{
#pragma clang fp contract(off)
    return (float)((double)s + dt * (l2m * (double)d1 + (double)lam_d2));
}
// add_sy_c: syy_c + 2 dt nu (-vy_c - 0.5 float32(dvzdx - dvxdz))
__device__ __forceinline__ float srot_x(float s, double dt2, float nu, float vyc, float curl)
{
#pragma clang fp contract(off)
    return (float)((double)s + dt2 * (double)nu * (-(double)vyc - 0.5 * (double)curl));
}

// the updates in either mode (fast: contraction allowed)
template <bool EX>
__device__ __forceinline__ float upd_sadd(float s, double f, double g)
{
    if constexpr (EX) return sadd_x(s, f, g);
    else return (float)((double)s + f * g);
}
template <bool EX>
__device__ __forceinline__ float upd_snorm(float s, double dt, double l2m, float d1, float lam_d2)
{
    if constexpr (EX) return snorm_x(s, dt, l2m, d1, lam_d2);
    else return (float)((double)s + dt * (l2m * (double)d1 + (double)lam_d2));
}
template <bool EX>
__device__ __forceinline__ float upd_srot(float s, double dt2, float nu, float vyc, float curl)
{
    if constexpr (EX) return srot_x(s, dt2, nu, vyc, curl);
    else return (float)((double)s + dt2 * (double)nu * (-(double)vyc - 0.5 * (double)curl));
}
template <bool EX>
__device__ __forceinline__ float upd_disp(float u, float v, double dt)
{
    if constexpr (EX) return disp_x(u, v, dt);
    else return (float)((double)u + (double)v * dt);
}
// derivative coefficients of a shot in both forms
struct DCo {
    double c1, c2, k9h, k9l, k1h, k1l;
};
template <bool EX>
__device__ __forceinline__ double dmn(float m2, float m1, float c0, float p1, const DCo &k)
{
    if constexpr (EX) return dminus_x(m2, m1, c0, p1, k.k9h, k.k9l, k.k1h, k.k1l);
    else return dminus(m2, m1, c0, p1, k.c1, k.c2);
}
template <bool EX>
__device__ __forceinline__ double dpl(float m1, float c0, float p1, float p2, const DCo &k)
{
    if constexpr (EX) return dplus_x(m1, c0, p1, p2, k.k9h, k.k9l, k.k1h, k.k1l);
    else return dplus(m1, c0, p1, p2, k.c1, k.c2);
}

// Everything iteration m of the row march reads from memory, fetched one
// iteration ahead so that a wavefront's loads overlap its own arithmetic.
// Row and lane indices are clamped into the grid so every load is valid;
// values fetched for rows or lanes outside the grid are never used.
template <int P>
struct RowLoads {
    float s[n_s(P)];   // old stresses, row m
    float v[n_v(P)];   // old velocities, row r = m-1
    float u[n_v(P)];   // displacement, row r
    float rho, jj, bx; // row r
    float lam, mu, nu, gam;  // row q = m-3
};

struct StepPtrs {
    gcf s_in[6];
    gcf v_in[3];
    gcf u_in[3];
    gcf rho, jj, lam, mu, nu, gam, bx;
};

template <int P>
__device__ __forceinline__ void load_row(RowLoads<P> &p, const StepPtrs &q, int m, int nx, int nz, int jc,
                                         int i0, int i1)
{
    constexpr int NS = n_s(P), NV = n_v(P);
    const int ms = min(max(m, 0), nx - 1);
    const int mr = min(max(m - 1, 0), nx - 1);
    const int mq = min(max(m - 3, i0), i1 - 1);
    const unsigned os = (unsigned)(ms * nz + jc), orr = (unsigned)(mr * nz + jc), oq = (unsigned)(mq * nz + jc);
#pragma unroll
    for (int f = 0; f < NS; f++) p.s[f] = q.s_in[f][os];
#pragma unroll
    for (int f = 0; f < NV; f++) {
        p.v[f] = q.v_in[f][orr];
        p.u[f] = q.u_in[f][orr];
    }
    p.rho = q.rho[orr];
    p.bx = q.bx[mr];
    p.jj = 1.f;
    if constexpr (P == SP_SPIN) p.jj = q.jj[orr];
    p.mu = q.mu[oq];
    p.lam = p.nu = p.gam = 0.f;
    if constexpr (P != SP_SH) p.lam = q.lam[oq];
    if constexpr (P == SP_SPIN) {
        p.nu = q.nu[oq];
        p.gam = q.gam[oq];
    }
}

#ifndef SP_STEP_WAVES
#define SP_STEP_WAVES 1
#endif
template <int P, bool EX>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(SP_STEP_WAVES))) void k_step(StepArgs a)
{
    constexpr int NS = n_s(P), NV = n_v(P);
    // Blocks are dispatched round-robin over the 8 XCDs; remap them so that
    // each XCD (and its L2) walks a contiguous run of the work list.
    const int nb = gridDim.x, b = blockIdx.x;
    const int q8 = nb >> 3, r8 = nb & 7, bx8 = b & 7, by8 = b >> 3;
    const int lb = bx8 < r8 ? bx8 * (q8 + 1) + by8 : r8 * (q8 + 1) + (bx8 - r8) * q8 + by8;
    const int wave = __builtin_amdgcn_readfirstlane(lb * 4 + (threadIdx.x >> 6));
    if (wave >= a.nitems) return;  // whole wavefront leaves together
    const int L = threadIdx.x & 63;
    const int4 item = a.items[2 * wave];
    if (item.x < 0) return;        // padding entry of the work list
    const int4 item2 = a.items[2 * wave + 1];
    const ShotDesc &d = a.shots[item.x];
    const int itx = a.itp ? a.it + *(cint_p)a.itp : a.it;
    const int nx = d.nx, nz = d.nz, i0 = item.z, i1 = item.w;
    const int j = item.y - 3 + L;
    const int jc = min(max(j, 0), nz - 1);
    const bool jin = j >= 0 && j < nz;
    const bool outl = jin && L >= 3 && L < 61;
    const bool zm_ok = j >= 2 && j < nz - 2, zp_ok = j >= 1 && j < nz - 2;
    const long long NP = d.Npad;
    const double dt = a.dt, dt2 = 2.0 * a.dt;
    const DCo KX{d.c1x, d.c2x, d.k9hx, d.k9lx, d.k1hx, d.k1lx}, KZ{d.c1z, d.c2z, d.k9hz, d.k9lz, d.k1hz, d.k1lz};
    const int rd = a.rd, wr = a.rd ^ 1;
    const double bzb = d.bzb[jc], bzt = d.bzt[jc];

    gf st = (gf)d.state;
    gcf mat = (gcf)d.mat;
    StepPtrs q;
#pragma unroll
    for (int f = 0; f < NS; f++) q.s_in[f] = st + slot_s(P, rd, f) * NP;
#pragma unroll
    for (int f = 0; f < NV; f++) {
        q.v_in[f] = st + slot_v(P, rd, f) * NP;
        q.u_in[f] = st + slot_u(P, f) * NP;
    }
    q.rho = mat + M_RHO * NP;
    q.jj = mat + M_J * NP;
    q.lam = mat + M_LAM * NP;
    q.mu = mat + M_MU * NP;
    q.nu = mat + M_NU * NP;
    q.gam = mat + M_GAM * NP;
    q.bx = (gcf)d.bx;
    gf s_out[NS];
    gf v_out[NV];
    gf u_out[NV];
#pragma unroll
    for (int f = 0; f < NS; f++) s_out[f] = st + slot_s(P, wr, f) * NP;
#pragma unroll
    for (int f = 0; f < NV; f++) {
        v_out[f] = st + slot_v(P, wr, f) * NP;
        u_out[f] = st + slot_u(P, f) * NP;
    }

    // injection list: sources (forward) or receivers with adjoint sources
    const int *ik = a.adjoint ? d.rec_k : d.src_k;
    const int *irow = a.adjoint ? d.rec_row : d.src_row;
    const int ni = a.adjoint ? d.nrec : d.nsrc;
    const float *ival = a.adjoint ? d.adstf : d.stf;
    const long long nrows = a.adjoint ? d.nrec : d.nsrc;
    // first list entries at or after the chunk's rows, found on the host
    int pi = a.adjoint ? item2.y : item2.x;
    int pr = item2.z;
    const bool record = !a.adjoint;

    float S0[NS], S1[NS], S2[NS], S3[NS];
    float W0[NV], W1[NV], W2[NV], W3[NV];
#pragma unroll
    for (int f = 0; f < NS; f++) S0[f] = S1[f] = S2[f] = S3[f] = 0.f;
#pragma unroll
    for (int f = 0; f < NV; f++) W0[f] = W1[f] = W2[f] = W3[f] = 0.f;

    // two rows in flight ahead of the one being computed
    RowLoads<P> n1, n2;
    load_row<P>(n1, q, i0 - 3, nx, nz, jc, i0, i1);
    load_row<P>(n2, q, i0 - 2, nx, nz, jc, i0, i1);
    for (int m = i0 - 3; m <= i1 + 2; ++m) {
        const RowLoads<P> cur = n1;
        n1 = n2;
        if (m + 2 <= i1 + 2) load_row<P>(n2, q, m + 2, nx, nz, jc, i0, i1);
#pragma unroll
        for (int f = 0; f < NS; f++) {
            S0[f] = S1[f];
            S1[f] = S2[f];
            S2[f] = S3[f];
            S3[f] = cur.s[f];
        }
        if (m < i0) continue;

        // ---- new velocities of row r = m-1 (stress window rows r-2..r+1)
        const int r = m - 1;
        float vn[NV];
#pragma unroll
        for (int f = 0; f < NV; f++) vn[f] = 0.f;
        if (r >= 0 && r < nx) {
            const bool xm_ok = r >= 2 && r < nx - 2;
            float dsv[NV];
            if constexpr (P == SP_SH) {
                // div_sy (solver.py:54-58)
                float zm1 = lane_m1(S2[1]), zm2 = lane_m1(zm1), zp1 = lane_p1(S2[1]);
                double ddx = xm_ok ? dmn<EX>(S0[0], S1[0], S2[0], S3[0], KX) : 0.0;
                double ddz = zm_ok ? dmn<EX>(zm2, zm1, S2[1], zp1, KZ) : 0.0;
                dsv[0] = (float)(ddx + ddz);
            } else {
                // div_sxz (solver.py:61-66)
                float xzm1 = lane_m1(S2[2]), xzm2 = lane_m1(xzm1), xzp1 = lane_p1(S2[2]);
                float zzm1 = lane_m1(S2[1]), zzm2 = lane_m1(zzm1), zzp1 = lane_p1(S2[1]);
                double dxx = xm_ok ? dmn<EX>(S0[0], S1[0], S2[0], S3[0], KX) : 0.0;
                double dxxz = xm_ok ? dmn<EX>(S0[2], S1[2], S2[2], S3[2], KX) : 0.0;
                double dzxz = zm_ok ? dmn<EX>(xzm2, xzm1, S2[2], xzp1, KZ) : 0.0;
                double dzzz = zm_ok ? dmn<EX>(zzm2, zzm1, S2[1], zzp1, KZ) : 0.0;
                dsv[0] = (float)(dxx + dzxz);
                dsv[1] = (float)(dxxz + dzzz);
                if constexpr (P == SP_SPIN) {
                    // div_sy on the couple stresses, then div_sxyz_c (solver.py:69-75, 217-218)
                    float yzm1 = lane_m1(S2[5]), yzm2 = lane_m1(yzm1), yzp1 = lane_p1(S2[5]);
                    float yym1 = lane_m1(S2[4]), yym2 = lane_m1(yym1), yyp1 = lane_p1(S2[4]);
                    double dxyx = xm_ok ? dmn<EX>(S0[3], S1[3], S2[3], S3[3], KX) : 0.0;
                    double dzyz = zm_ok ? dmn<EX>(yzm2, yzm1, S2[5], yzp1, KZ) : 0.0;
                    float dsy = (float)(dxyx + dzyz);
                    double dzyy = zm_ok ? dmn<EX>(yym2, yym1, S2[4], yyp1, KZ) : 0.0;
                    double dxyy = xm_ok ? dmn<EX>(S0[4], S1[4], S2[4], S3[4], KX) : 0.0;
                    dsv[0] = (float)((double)dsv[0] + dzyy);
                    dsv[1] = (float)((double)dsv[1] - dxyy);
                    dsv[2] = (float)((double)dsy + 2.0 * (double)S2[4]);
                }
            }
            // stf_dsy / stf_dsxz (solver.py:78-93): float32 adds, list order
            const int kend = (r + 1) * nz, kr = r * nz + j;
            while (pi < ni && ik[pi] < kend) {
                if (ik[pi] == kr) {
                    const int row = irow[pi];
                    dsv[0] = add_rn(dsv[0], ival[(long long)row * a.nt + itx]);
                    if constexpr (P != SP_SH) dsv[1] = add_rn(dsv[1], ival[(nrows + row) * a.nt + itx]);
                }
                ++pi;
            }
            // add_vy / add_vxz (solver.py:95-109) with the separable sponge of
            // set_bound (fd2d.py:47-68): f32(f32(bx*bottom)*top)
            const float b1 = (float)((double)cur.bx * bzb);
            const double b = (double)(float)((double)b1 * bzt);
            if constexpr (EX) {
                const double rr = (double)cur.rho, irr = rcp_rn(rr);
#pragma unroll
                for (int f = 0; f < NV; f++) {
                    double rf = rr, ir = irr;
                    if constexpr (P == SP_SPIN) {
                        if (f == 2) {
                            rf = (double)cur.jj;
                            ir = rcp_rn(rf);
                        }
                    }
                    const float v = vel_x(cur.v[f], dsv[f], dt, rf, ir, b);
                    vn[f] = jin ? v : 0.f;
                }
            } else {
                const double dtr = dt * rcp64((double)cur.rho);
#pragma unroll
                for (int f = 0; f < NV; f++) {
                    double dtm = dtr;
                    if constexpr (P == SP_SPIN) {
                        if (f == 2) dtm = dt * rcp64((double)cur.jj);
                    }
                    vn[f] = jin ? (float)(b * ((double)cur.v[f] + (double)dsv[f] * dtm)) : 0.f;
                }
            }
            if (outl && r >= i0 && r < i1) {
                const unsigned k = (unsigned)(r * nz + j);
                float un[NV];
#pragma unroll
                for (int f = 0; f < NV; f++) {
                    un[f] = upd_disp<EX>(cur.u[f], vn[f], dt);
                    u_out[f][k] = un[f];
                    v_out[f][k] = vn[f];
                }
                if (a.snap_idx >= 0) {
                    gf sp = (gf)d.snap + (long long)a.snap_idx * n_snap(P) * NP;
#pragma unroll
                    for (int f = 0; f < NV; f++) {
                        sp[f * NP + k] = cur.u[f];
                        sp[(NV + f) * NP + k] = vn[f];
                    }
                }
                // save_obs (solver.py:151-156): displacement after the step
                if (record) {
                    for (int q2 = pr; q2 < d.nrec && d.rec_k[q2] < kend; ++q2) {
                        if (d.rec_k[q2] == kr) {
                            const int row = d.rec_row[q2];
#pragma unroll
                            for (int f = 0; f < NV; f++)
                                d.rec[((long long)f * d.nrec + row) * a.nt + itx] = un[f];
                        }
                    }
                }
            }
            if (record && r >= i0 && r < i1)
                while (pr < d.nrec && d.rec_k[pr] < kend) ++pr;
        }
#pragma unroll
        for (int f = 0; f < NV; f++) {
            W0[f] = W1[f];
            W1[f] = W2[f];
            W2[f] = W3[f];
            W3[f] = vn[f];
        }

        // ---- new stresses of row qr = m-3 (velocity window rows qr-1..qr+2)
        const int qr = m - 3;
        if (qr < i0) continue;
        const bool xp_ok = qr >= 1 && qr < nx - 2;
        const unsigned kq = (unsigned)(qr * nz + j);
        if constexpr (P == SP_SH) {
            // div_vy + add_sy (solver.py:112-117, 128-133)
            float zm1 = lane_m1(W1[0]), zp1 = lane_p1(W1[0]), zp2 = lane_p1(zp1);
            if (outl) {
                const float dvdx = xp_ok ? (float)dpl<EX>(W0[0], W1[0], W2[0], W3[0], KX) : 0.f;
                const float dvdz = zp_ok ? (float)dpl<EX>(zm1, W1[0], zp1, zp2, KZ) : 0.f;
                const double dtmu = dt * (double)cur.mu;
                s_out[0][kq] = upd_sadd<EX>(S0[0], dtmu, (double)dvdx);
                s_out[1][kq] = upd_sadd<EX>(S0[1], dtmu, (double)dvdz);
            }
        } else {
            // div_vxz + add_sxz (solver.py:119-126, 143-149)
            float xm1 = lane_m1(W1[0]), xp1 = lane_p1(W1[0]), xp2 = lane_p1(xp1);
            float zm1 = lane_m1(W1[1]), zp1 = lane_p1(W1[1]), zp2 = lane_p1(zp1);
            float cm1 = 0.f, cp1 = 0.f, cp2 = 0.f;
            if constexpr (P == SP_SPIN) {
                cm1 = lane_m1(W1[2]);
                cp1 = lane_p1(W1[2]);
                cp2 = lane_p1(cp1);
            }
            if (outl) {
                const float dvxdx = xp_ok ? (float)dpl<EX>(W0[0], W1[0], W2[0], W3[0], KX) : 0.f;
                const float dvzdx = xp_ok ? (float)dpl<EX>(W0[1], W1[1], W2[1], W3[1], KX) : 0.f;
                const float dvxdz = zp_ok ? (float)dpl<EX>(xm1, W1[0], xp1, xp2, KZ) : 0.f;
                const float dvzdz = zp_ok ? (float)dpl<EX>(zm1, W1[1], zp1, zp2, KZ) : 0.f;
                const float lam = cur.lam, mu = cur.mu;
                const double l2m = (double)lam + 2.0 * (double)mu;
                s_out[0][kq] = upd_snorm<EX>(S0[0], dt, l2m, dvxdx, mul_rn(lam, dvzdz));
                s_out[1][kq] = upd_snorm<EX>(S0[1], dt, l2m, dvzdz, mul_rn(lam, dvxdx));
                s_out[2][kq] = upd_sadd<EX>(S0[2], dt, (double)mul_rn(mu, add_rn(dvxdz, dvzdx)));
                if constexpr (P == SP_SPIN) {
                    // div_vy on vy_c + add_sy_c (solver.py:135-141, 227-228)
                    const float dcdx = xp_ok ? (float)dpl<EX>(W0[2], W1[2], W2[2], W3[2], KX) : 0.f;
                    const float dcdz = zp_ok ? (float)dpl<EX>(cm1, W1[2], cp1, cp2, KZ) : 0.f;
                    const double dtg = dt * (double)cur.gam;
                    s_out[4][kq] = upd_srot<EX>(S0[4], dt2, cur.nu, W1[2], sub_rn(dvzdx, dvxdz));
                    s_out[3][kq] = upd_sadd<EX>(S0[3], dtg, (double)dcdx);
                    s_out[5][kq] = upd_sadd<EX>(S0[5], dtg, (double)dcdz);
                }
            }
        }
    }
}

// ------------------------------------------------- step kernel, version 3
// Same algorithm and rounding as k_step; the difference is memory-pipeline
// behaviour.  On gfx9 loads and stores share one in-order vmcnt counter, so
// the compiler can only wait for exactly the prefetched row (and not drain
// every outstanding load and store) when each row iteration issues the same
// sequence of vector memory operations.  Hence:
//   * buffer loads/stores with one 32-bit byte offset per row, the field's
//     slot offset in an SGPR and the base in a buffer resource;
//   * loads always issue (rows and lanes clamped into the grid); stores
//     always issue, with an offset past the end of the buffer for lanes or
//     rows that must not be written -- the buffer range check drops them;
//   * snapshots and receiver records are separate kernels (k_snap,
//     k_record); source injection reads the time function with scalar
//     loads only;
//   * the row loop is unrolled by four with windows indexed by row & 3, so
//     the register windows never move.
typedef __amdgpu_buffer_rsrc_t rsrc_t;
constexpr unsigned OOB = 0x40000000u;  // >= every per-shot buffer size (checked on the host)
__device__ __forceinline__ rsrc_t make_rsrc(const void *p, long long bytes)
{
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(p), 0, (int)bytes, 0x00020000);
}
__device__ __forceinline__ float bld(rsrc_t r, unsigned voff, int soff)
{
    return __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(r, (int)voff, soff, 0));
}
__device__ __forceinline__ void bst(rsrc_t r, unsigned voff, int soff, float v)
{
    __builtin_amdgcn_raw_buffer_store_b32(__float_as_uint(v), r, (int)voff, soff, 0);
}

template <int P>
struct Row3 {
    float s[n_s(P)];   // old stresses, row m
    float v[n_v(P)];   // old velocities, row m-1
    float u[n_v(P)];   // displacement, row m-1
    float rho, jj, bx; // row m-1
    float lam, mu, nu, gam;  // row m-3
};

template <int P>
__device__ __forceinline__ void load3(Row3<P> &p, rsrc_t RS, rsrc_t RM, cflt_p bxp, int NP4, int rd, int mp, int nx,
                                      int nz, int jc, int i0, int i1)
{
    constexpr int NS = n_s(P), NV = n_v(P);
    // each field's row is clamped to the rows the chunk needs, so the extra
    // prologue/epilogue iterations re-read cached rows instead of new ones:
    // stresses i0-3..i1+2, velocities i0-1..i1+1, u and material i0..i1-1
    const int ms = min(max(mp, max(i0 - 3, 0)), min(i1 + 2, nx - 1));
    const int mr = min(max(mp - 1, max(i0 - 1, 0)), min(i1 + 1, nx - 1));
    const int mu_ = min(max(mp - 1, i0), i1 - 1);
    const int mq = min(max(mp - 3, i0), i1 - 1);
    const unsigned os = (unsigned)(ms * nz + jc) * 4u, orr = (unsigned)(mr * nz + jc) * 4u,
                   ou = (unsigned)(mu_ * nz + jc) * 4u, oq = (unsigned)(mq * nz + jc) * 4u;
#pragma unroll
    for (int f = 0; f < NS; f++) p.s[f] = bld(RS, os, slot_s(P, rd, f) * NP4);
#pragma unroll
    for (int f = 0; f < NV; f++) {
        p.v[f] = bld(RS, orr, slot_v(P, rd, f) * NP4);
        p.u[f] = bld(RS, ou, slot_u(P, f) * NP4);
    }
    p.rho = bld(RM, orr, M_RHO * NP4);
    p.bx = bxp[mr];
    p.jj = 1.f;
    if constexpr (P == SP_SPIN) p.jj = bld(RM, orr, M_J * NP4);
    p.mu = bld(RM, oq, M_MU * NP4);
    p.lam = p.nu = p.gam = 0.f;
    if constexpr (P != SP_SH) p.lam = bld(RM, oq, M_LAM * NP4);
    if constexpr (P == SP_SPIN) {
        p.nu = bld(RM, oq, M_NU * NP4);
        p.gam = bld(RM, oq, M_GAM * NP4);
    }
}

template <int P, bool EX>
__global__ __launch_bounds__(256) void k_step3(StepArgs a)
{
    constexpr int NS = n_s(P), NV = n_v(P);
    const int nb = gridDim.x, bb = blockIdx.x;
    const int q8 = nb >> 3, r8 = nb & 7, bx8 = bb & 7, by8 = bb >> 3;
    const int lb = bx8 < r8 ? bx8 * (q8 + 1) + by8 : r8 * (q8 + 1) + (bx8 - r8) * q8 + by8;
    const int wave = __builtin_amdgcn_readfirstlane(lb * 4 + (threadIdx.x >> 6));
    if (wave >= a.nitems) return;
    const int L = threadIdx.x & 63;
    const int4 item = a.items[2 * wave];
    if (item.x < 0) return;
    const int4 item2 = a.items[2 * wave + 1];
    const ShotDesc &d = a.shots[item.x];
    const int itx = a.itp ? a.it + *(cint_p)a.itp : a.it;
    const int nx = d.nx, nz = d.nz, i0 = item.z, i1 = item.w;
    const int j = item.y - a.shift + L;
    const int jc = min(max(j, 0), nz - 1);
    const bool jin = j >= 0 && j < nz;
    const bool outl = jin && L >= a.shift && L < a.shift + a.wout;
    const bool zm_ok = j >= 2 && j < nz - 2, zp_ok = j >= 1 && j < nz - 2;
    const int NP4 = (int)(d.Npad * 4);  // bytes per field
    const double dt = a.dt, dt2 = 2.0 * a.dt;
    const DCo KX{d.c1x, d.c2x, d.k9hx, d.k9lx, d.k1hx, d.k1lx}, KZ{d.c1z, d.c2z, d.k9hz, d.k9lz, d.k1hz, d.k1lz};
    const int rd = a.rd, wr = a.rd ^ 1;
    const double bzb = d.bzb[jc], bzt = d.bzt[jc];
    const rsrc_t RS = make_rsrc(d.state, (long long)n_slots(P) * NP4);
    const rsrc_t RM = make_rsrc(d.mat, (long long)N_MAT * NP4);
    const cflt_p bxp = (cflt_p)d.bx;

    // injection list: sources (forward) or receivers with adjoint sources
    const cint_p ik = (cint_p)(a.adjoint ? d.rec_k : d.src_k);
    const cint_p irow = (cint_p)(a.adjoint ? d.rec_row : d.src_row);
    const int ni = a.adjoint ? d.nrec : d.nsrc;
    const cflt_p ival = (cflt_p)(a.adjoint ? (const float *)d.adstf : d.stf);
    const long long nrows = a.adjoint ? d.nrec : d.nsrc;
    int pi = __builtin_amdgcn_readfirstlane(a.adjoint ? item2.y : item2.x);
    while (pi < ni && ik[pi] < (i0 - 1) * nz) ++pi;  // the list pointer starts at row i0-4
    int src_row = pi < ni ? ik[pi] / nz : 0x7fffffff;

    float S[4][NS], W[4][NV];
    Row3<P> PF[4];
#pragma unroll
    for (int u = 0; u < 4; u++) {
#pragma unroll
        for (int f = 0; f < NS; f++) S[u][f] = 0.f;
#pragma unroll
        for (int f = 0; f < NV; f++) W[u][f] = 0.f;
    }

    // prologue (i0 is a multiple of 4): rows i0-3, i0-2, i0-1 fill stress
    // slots 1, 2, 3; rows up to i0+1 are prefetched.
    load3<P>(PF[1], RS, RM, bxp, NP4, rd, i0 - 3, nx, nz, jc, i0, i1);
    load3<P>(PF[2], RS, RM, bxp, NP4, rd, i0 - 2, nx, nz, jc, i0, i1);
    load3<P>(PF[3], RS, RM, bxp, NP4, rd, i0 - 1, nx, nz, jc, i0, i1);
    load3<P>(PF[0], RS, RM, bxp, NP4, rd, i0, nx, nz, jc, i0, i1);
#pragma unroll
    for (int f = 0; f < NS; f++) {
        S[1][f] = PF[1].s[f];
        S[2][f] = PF[2].s[f];
        S[3][f] = PF[3].s[f];
    }
    load3<P>(PF[1], RS, RM, bxp, NP4, rd, i0 + 1, nx, nz, jc, i0, i1);

    const int mend = i1 + 2;
    for (int mb = i0; mb <= mend; mb += 4) {
#pragma unroll
        for (int u = 0; u < 4; u++) {
            const int m = mb + u;
            if (m > mend) break;
            const int A = (u + 1) & 3, B = (u + 2) & 3, C = (u + 3) & 3, D = u;
            // ---- prefetch row m+2 into slot (u+2)&3 (always issued)
            load3<P>(PF[B], RS, RM, bxp, NP4, rd, m + 2, nx, nz, jc, i0, i1);
            const Row3<P> &cur = PF[D];
#pragma unroll
            for (int f = 0; f < NS; f++) S[D][f] = cur.s[f];
            // stress window rows m-3, m-2, m-1, m = slots A, B, C, D
            float(&Sa)[NS] = S[A];
            float(&Sb)[NS] = S[B];
            float(&Sc)[NS] = S[C];
            float(&Sd)[NS] = S[D];

            // ---- new velocities of row r = m-1
            const int r = m - 1;
            const bool rin = r >= 0 && r < nx;
            const bool xm_ok = r >= 2 && r < nx - 2;
            float dsv[NV];
            if constexpr (P == SP_SH) {
                float zm1 = lane_m1(Sc[1]), zm2 = lane_m1(zm1), zp1 = lane_p1(Sc[1]);
                double ddx = xm_ok ? dmn<EX>(Sa[0], Sb[0], Sc[0], Sd[0], KX) : 0.0;
                double ddz = zm_ok ? dmn<EX>(zm2, zm1, Sc[1], zp1, KZ) : 0.0;
                dsv[0] = (float)(ddx + ddz);
            } else {
                float xzm1 = lane_m1(Sc[2]), xzm2 = lane_m1(xzm1), xzp1 = lane_p1(Sc[2]);
                float zzm1 = lane_m1(Sc[1]), zzm2 = lane_m1(zzm1), zzp1 = lane_p1(Sc[1]);
                double dxx = xm_ok ? dmn<EX>(Sa[0], Sb[0], Sc[0], Sd[0], KX) : 0.0;
                double dxxz = xm_ok ? dmn<EX>(Sa[2], Sb[2], Sc[2], Sd[2], KX) : 0.0;
                double dzxz = zm_ok ? dmn<EX>(xzm2, xzm1, Sc[2], xzp1, KZ) : 0.0;
                double dzzz = zm_ok ? dmn<EX>(zzm2, zzm1, Sc[1], zzp1, KZ) : 0.0;
                dsv[0] = (float)(dxx + dzxz);
                dsv[1] = (float)(dxxz + dzzz);
                if constexpr (P == SP_SPIN) {
                    float yzm1 = lane_m1(Sc[5]), yzm2 = lane_m1(yzm1), yzp1 = lane_p1(Sc[5]);
                    float yym1 = lane_m1(Sc[4]), yym2 = lane_m1(yym1), yyp1 = lane_p1(Sc[4]);
                    double dxyx = xm_ok ? dmn<EX>(Sa[3], Sb[3], Sc[3], Sd[3], KX) : 0.0;
                    double dzyz = zm_ok ? dmn<EX>(yzm2, yzm1, Sc[5], yzp1, KZ) : 0.0;
                    float dsy = (float)(dxyx + dzyz);
                    double dzyy = zm_ok ? dmn<EX>(yym2, yym1, Sc[4], yyp1, KZ) : 0.0;
                    double dxyy = xm_ok ? dmn<EX>(Sa[4], Sb[4], Sc[4], Sd[4], KX) : 0.0;
                    dsv[0] = (float)((double)dsv[0] + dzyy);
                    dsv[1] = (float)((double)dsv[1] - dxyy);
                    dsv[2] = (float)((double)dsy + 2.0 * (double)Sc[4]);
                }
            }
            if (r == src_row) {
                // stf_dsy / stf_dsxz (solver.py:78-93): float32 adds, list
                // order; every read here is wave-uniform (scalar loads)
                const int kend = (r + 1) * nz, kr = r * nz + j;
                while (pi < ni && ik[pi] < kend) {
                    const int kk = ik[pi];
                    const int row = irow[pi];
                    const float s0 = ival[(long long)row * a.nt + itx];
                    float s1 = 0.f;
                    if constexpr (P != SP_SH) s1 = ival[(nrows + row) * a.nt + itx];
                    if (kk == kr) {
                        dsv[0] = add_rn(dsv[0], s0);
                        if constexpr (P != SP_SH) dsv[1] = add_rn(dsv[1], s1);
                    }
                    ++pi;
                }
                src_row = pi < ni ? ik[pi] / nz : 0x7fffffff;
            }
            const float b1 = (float)((double)cur.bx * bzb);
            const double b = (double)(float)((double)b1 * bzt);
            float vn[NV], un[NV];
            if constexpr (EX) {
                const double rr = (double)cur.rho, irr = rcp_rn(rr);
#pragma unroll
                for (int f = 0; f < NV; f++) {
                    double rf = rr, ir = irr;
                    if constexpr (P == SP_SPIN) {
                        if (f == 2) {
                            rf = (double)cur.jj;
                            ir = rcp_rn(rf);
                        }
                    }
                    const float v = vel_x(cur.v[f], dsv[f], dt, rf, ir, b);
                    vn[f] = (rin && jin) ? v : 0.f;
                    un[f] = disp_x(cur.u[f], vn[f], dt);
                }
            } else {
                const double dtr = dt * rcp64((double)cur.rho);
#pragma unroll
                for (int f = 0; f < NV; f++) {
                    double dtm = dtr;
                    if constexpr (P == SP_SPIN) {
                        if (f == 2) dtm = dt * rcp64((double)cur.jj);
                    }
                    const float v = (float)(b * ((double)cur.v[f] + (double)dsv[f] * dtm));
                    vn[f] = (rin && jin) ? v : 0.f;
                    un[f] = (float)((double)cur.u[f] + (double)vn[f] * dt);
                }
            }
            {
                const unsigned ov = (outl && r >= i0 && r < i1) ? (unsigned)(r * nz + j) * 4u : OOB;
#pragma unroll
                for (int f = 0; f < NV; f++) {
                    bst(RS, ov, slot_u(P, f) * NP4, un[f]);
                    bst(RS, ov, slot_v(P, wr, f) * NP4, vn[f]);
                }
            }
#pragma unroll
            for (int f = 0; f < NV; f++) W[C][f] = vn[f];

            // ---- new stresses of row q = m-3 (velocity rows q-1..q+2 = slots D, A, B, C)
            const int qr = m - 3;
            float(&Wa)[NV] = W[D];
            float(&Wb)[NV] = W[A];
            float(&Wc)[NV] = W[B];
            float(&Wd)[NV] = W[C];
            const bool xp_ok = qr >= 1 && qr < nx - 2;
            const unsigned oq = (outl && qr >= i0 && qr < i1) ? (unsigned)(qr * nz + j) * 4u : OOB;
            if constexpr (P == SP_SH) {
                float zm1 = lane_m1(Wb[0]), zp1 = lane_p1(Wb[0]), zp2 = lane_p1(zp1);
                const float dvdx = xp_ok ? (float)dpl<EX>(Wa[0], Wb[0], Wc[0], Wd[0], KX) : 0.f;
                const float dvdz = zp_ok ? (float)dpl<EX>(zm1, Wb[0], zp1, zp2, KZ) : 0.f;
                const double dtmu = dt * (double)cur.mu;
                bst(RS, oq, slot_s(P, wr, 0) * NP4, upd_sadd<EX>(Sa[0], dtmu, (double)dvdx));
                bst(RS, oq, slot_s(P, wr, 1) * NP4, upd_sadd<EX>(Sa[1], dtmu, (double)dvdz));
            } else {
                float xm1 = lane_m1(Wb[0]), xp1 = lane_p1(Wb[0]), xp2 = lane_p1(xp1);
                float zm1 = lane_m1(Wb[1]), zp1 = lane_p1(Wb[1]), zp2 = lane_p1(zp1);
                const float dvxdx = xp_ok ? (float)dpl<EX>(Wa[0], Wb[0], Wc[0], Wd[0], KX) : 0.f;
                const float dvzdx = xp_ok ? (float)dpl<EX>(Wa[1], Wb[1], Wc[1], Wd[1], KX) : 0.f;
                const float dvxdz = zp_ok ? (float)dpl<EX>(xm1, Wb[0], xp1, xp2, KZ) : 0.f;
                const float dvzdz = zp_ok ? (float)dpl<EX>(zm1, Wb[1], zp1, zp2, KZ) : 0.f;
                const float lam = cur.lam, mu = cur.mu;
                const double l2m = (double)lam + 2.0 * (double)mu;
                bst(RS, oq, slot_s(P, wr, 0) * NP4, upd_snorm<EX>(Sa[0], dt, l2m, dvxdx, mul_rn(lam, dvzdz)));
                bst(RS, oq, slot_s(P, wr, 1) * NP4, upd_snorm<EX>(Sa[1], dt, l2m, dvzdz, mul_rn(lam, dvxdx)));
                bst(RS, oq, slot_s(P, wr, 2) * NP4, upd_sadd<EX>(Sa[2], dt, (double)mul_rn(mu, add_rn(dvxdz, dvzdx))));
                if constexpr (P == SP_SPIN) {
                    float cm1 = lane_m1(Wb[2]), cp1 = lane_p1(Wb[2]), cp2 = lane_p1(cp1);
                    const float dcdx = xp_ok ? (float)dpl<EX>(Wa[2], Wb[2], Wc[2], Wd[2], KX) : 0.f;
                    const float dcdz = zp_ok ? (float)dpl<EX>(cm1, Wb[2], cp1, cp2, KZ) : 0.f;
                    const double dtg = dt * (double)cur.gam;
                    bst(RS, oq, slot_s(P, wr, 4) * NP4, upd_srot<EX>(Sa[4], dt2, cur.nu, Wb[2], sub_rn(dvzdx, dvxdz)));
                    bst(RS, oq, slot_s(P, wr, 3) * NP4, upd_sadd<EX>(Sa[3], dtg, (double)dcdx));
                    bst(RS, oq, slot_s(P, wr, 5) * NP4, upd_sadd<EX>(Sa[5], dtg, (double)dcdz));
                }
            }
        }
    }
}

// ------------------------------------------------- two steps per launch
// k_step2 advances a strip by two time steps (it and it+1) in one row march,
// keeping the intermediate state in registers: the state is read once and
// written once per two steps.  The intermediate stresses and velocities of a
// chunk's halo rows are recomputed by each chunk that needs them; the strip
// keeps 6 lanes of halo on each side (lanes [6, 58) are output), and the
// chunk reads 6 rows before and 5 after its own.  Every value is computed
// with the same operations, in the same order, as two launches of k_step3,
// so results are identical bit for bit.  The result goes to the other
// stress/velocity set (rd ^ 1), like one step, so no wavefront overwrites
// rows another one still reads.
//
// Iteration m of the march (rows relative to m):
//   A  v(n+1)  at row m-1 from s(n) rows m-3..m   (source at step it)
//   B  s(n+1)  at row m-3 from v(n+1) rows m-4..m-1
//   C  v(n+2)  at row m-4 from s(n+1) rows m-6..m-3 (source at step it+1);
//      u(n+1), u(n+2) and the receiver records of both steps at row m-4
//   D  s(n+2)  at row m-6 from v(n+2) rows m-7..m-4
// Windows of four rows are indexed by row & 3 (loop unrolled by four).

template <int P>
struct Row2 {
    float s[n_s(P)];     // s(n), row m
    float v[n_v(P)];     // v(n), row m-1
    float u[n_v(P)];     // u(n), row m-4
    float rhoA, jA;      // row m-1 (j: micro-inertia, spin only)
    float lamB, muB;     // row m-3
    float nuB, gamB;     // row m-3 (spin only: nu, mu_c + nu_c)
    // rows m-4 and m-6 of the same materials are the values loaded three
    // iterations earlier; the march keeps them in register windows
};

// The rows are wave-uniform, so each load's row offset goes into the scalar
// offset and the per-lane offset vj = 4 jc is the same for every load (no
// vector address arithmetic per row).  Every row is clamped into the grid,
// so every load stays inside its buffer.
template <int P>
__device__ __forceinline__ void load2(Row2<P> &p, rsrc_t RS, rsrc_t RM, cflt_p bxp, int NP4, int rd, int mp, int nx,
                                      int nz, unsigned vj, int i0, int i1)
{
    constexpr int NS = n_s(P), NV = n_v(P);
    // rows clamped to what the chunk needs (and to the grid)
    const int ms = min(max(mp, max(i0 - 6, 0)), min(i1 + 5, nx - 1));
    const int ma = min(max(mp - 1, max(i0 - 4, 0)), min(i1 + 4, nx - 1));
    const int mb = min(max(mp - 3, max(i0 - 3, 0)), min(i1 + 2, nx - 1));
    const int mu_ = min(max(mp - 4, i0), i1 - 1);
    const int rb = nz * 4;
    const int os = ms * rb, oa = ma * rb, ob = mb * rb, ou = mu_ * rb;
#pragma unroll
    for (int f = 0; f < NS; f++) p.s[f] = bld(RS, vj, slot_s(P, rd, f) * NP4 + os);
#pragma unroll
    for (int f = 0; f < NV; f++) {
        p.v[f] = bld(RS, vj, slot_v(P, rd, f) * NP4 + oa);
        p.u[f] = bld(RS, vj, slot_u(P, f) * NP4 + ou);
    }
    p.rhoA = bld(RM, vj, M_RHO * NP4 + oa);
    p.muB = bld(RM, vj, M_MU * NP4 + ob);
    p.lamB = 0.f;
    if constexpr (P != SP_SH) p.lamB = bld(RM, vj, M_LAM * NP4 + ob);
    p.jA = 1.f;
    p.nuB = p.gamB = 0.f;
    if constexpr (P == SP_SPIN) {
        p.jA = bld(RM, vj, M_J * NP4 + oa);
        p.nuB = bld(RM, vj, M_NU * NP4 + ob);
        p.gamB = bld(RM, vj, M_GAM * NP4 + ob);
    }
}

// stress divergence of one velocity row (div_sy / div_sxz, solver.py:54-66)
// from the stress window rows r-2..r+1 (sa..sd)
template <int P, bool EX>
__device__ __forceinline__ void vel_row(float (&dsv)[n_v(P)], const float (&sa)[n_s(P)], const float (&sb)[n_s(P)],
                                        const float (&sc)[n_s(P)], const float (&sd)[n_s(P)], bool xm_ok, bool zm_ok,
                                        int L, const DCo &KX, const DCo &KZ)
{
    if constexpr (P == SP_SH) {
        float zm1 = lane_m1(sc[1]), zm2 = lane_m1(zm1), zp1 = lane_p1(sc[1]);
        double ddx = xm_ok ? dmn<EX>(sa[0], sb[0], sc[0], sd[0], KX) : 0.0;
        double ddz = zm_ok ? dmn<EX>(zm2, zm1, sc[1], zp1, KZ) : 0.0;
        dsv[0] = (float)(ddx + ddz);
    } else {
        float xzm1 = lane_m1(sc[2]), xzm2 = lane_m1(xzm1), xzp1 = lane_p1(sc[2]);
        float zzm1 = lane_m1(sc[1]), zzm2 = lane_m1(zzm1), zzp1 = lane_p1(sc[1]);
        double dxx = xm_ok ? dmn<EX>(sa[0], sb[0], sc[0], sd[0], KX) : 0.0;
        double dxxz = xm_ok ? dmn<EX>(sa[2], sb[2], sc[2], sd[2], KX) : 0.0;
        double dzxz = zm_ok ? dmn<EX>(xzm2, xzm1, sc[2], xzp1, KZ) : 0.0;
        double dzzz = zm_ok ? dmn<EX>(zzm2, zzm1, sc[1], zzp1, KZ) : 0.0;
        dsv[0] = (float)(dxx + dzxz);
        dsv[1] = (float)(dxxz + dzzz);
        if constexpr (P == SP_SPIN) {
            // div_sy on the couple stresses, then div_sxyz_c (solver.py:69-75, 217-218)
            float yzm1 = lane_m1(sc[5]), yzm2 = lane_m1(yzm1), yzp1 = lane_p1(sc[5]);
            float yym1 = lane_m1(sc[4]), yym2 = lane_m1(yym1), yyp1 = lane_p1(sc[4]);
            double dxyx = xm_ok ? dmn<EX>(sa[3], sb[3], sc[3], sd[3], KX) : 0.0;
            double dzyz = zm_ok ? dmn<EX>(yzm2, yzm1, sc[5], yzp1, KZ) : 0.0;
            float dsy = (float)(dxyx + dzyz);
            double dzyy = zm_ok ? dmn<EX>(yym2, yym1, sc[4], yyp1, KZ) : 0.0;
            double dxyy = xm_ok ? dmn<EX>(sa[4], sb[4], sc[4], sd[4], KX) : 0.0;
            dsv[0] = (float)((double)dsv[0] + dzyy);
            dsv[1] = (float)((double)dsv[1] - dxyy);
            dsv[2] = (float)((double)dsy + 2.0 * (double)sc[4]);
        }
    }
}

// velocity update (add_vy / add_vxz, solver.py:96-117) from the divergence
// plus sources, the old v, 1/rho and the sponge factors
// `zint`: every lane of the strip lies outside the top/bottom sponges; with
// bx = 1 the sponge factor is exactly 1 and its arithmetic is skipped
// (the same float64 result, bit for bit).
template <int P, bool EX>
__device__ __forceinline__ void vel_finish(float (&vn)[n_v(P)], const float (&dsv)[n_v(P)], const float (&vo)[n_v(P)],
                                           float rho, float jj, float bxr, double bzb, double bzt, bool zint, bool rin,
                                           bool jin, double dt)
{
    if constexpr (EX) {
        double b = 1.0;
        if (!(zint && bxr == 1.f)) {
            const float b1 = (float)((double)bxr * bzb);
            b = (double)(float)((double)b1 * bzt);
        }
        const double rr = (double)rho, irr = rcp_rn(rr);
#pragma unroll
        for (int f = 0; f < n_v(P); f++) {
            double rf = rr, ir = irr;
            if constexpr (P == SP_SPIN) {
                if (f == 2) {  // vy_c divides by the micro-inertia j (solver.py:226)
                    rf = (double)jj;
                    ir = rcp_rn(rf);
                }
            }
            const float v = vel_x(vo[f], dsv[f], dt, rf, ir, b);
            vn[f] = (rin && jin) ? v : 0.f;
        }
        return;
    }
    const double dtr = dt * rcp64((double)rho);
    double dtj = dtr;
    if constexpr (P == SP_SPIN) dtj = dt * rcp64((double)jj);
    if (zint && bxr == 1.f) {
#pragma unroll
        for (int f = 0; f < n_v(P); f++) {
            const float v = (float)((double)vo[f] + (double)dsv[f] * (f == 2 ? dtj : dtr));
            vn[f] = (rin && jin) ? v : 0.f;
        }
        return;
    }
    const float b1 = (float)((double)bxr * bzb);
    const double b = (double)(float)((double)b1 * bzt);
#pragma unroll
    for (int f = 0; f < n_v(P); f++) {
        const float v = (float)(b * ((double)vo[f] + (double)dsv[f] * (f == 2 ? dtj : dtr)));
        vn[f] = (rin && jin) ? v : 0.f;
    }
}

// stress update of one row (add_sy / add_sxz, solver.py:120-148) from the
// velocity window rows q-1..q+2 (wa..wd) and the old stresses
template <int P, bool EX>
__device__ __forceinline__ void str_row(float (&sn)[n_s(P)], const float (&wa)[n_v(P)], const float (&wb)[n_v(P)],
                                        const float (&wc)[n_v(P)], const float (&wd)[n_v(P)], const float (&so)[n_s(P)],
                                        float lam, float mu, float nu, float gam, bool xp_ok, bool zp_ok, int L,
                                        const DCo &KX, const DCo &KZ, double dt)
{
    if constexpr (P == SP_SH) {
        float zm1 = lane_m1(wb[0]), zp1 = lane_p1(wb[0]), zp2 = lane_p1(zp1);
        const float dvdx = xp_ok ? (float)dpl<EX>(wa[0], wb[0], wc[0], wd[0], KX) : 0.f;
        const float dvdz = zp_ok ? (float)dpl<EX>(zm1, wb[0], zp1, zp2, KZ) : 0.f;
        const double dtmu = dt * (double)mu;
        sn[0] = upd_sadd<EX>(so[0], dtmu, (double)dvdx);
        sn[1] = upd_sadd<EX>(so[1], dtmu, (double)dvdz);
    } else {
        float xm1 = lane_m1(wb[0]), xp1 = lane_p1(wb[0]), xp2 = lane_p1(xp1);
        float zm1 = lane_m1(wb[1]), zp1 = lane_p1(wb[1]), zp2 = lane_p1(zp1);
        const float dvxdx = xp_ok ? (float)dpl<EX>(wa[0], wb[0], wc[0], wd[0], KX) : 0.f;
        const float dvzdx = xp_ok ? (float)dpl<EX>(wa[1], wb[1], wc[1], wd[1], KX) : 0.f;
        const float dvxdz = zp_ok ? (float)dpl<EX>(xm1, wb[0], xp1, xp2, KZ) : 0.f;
        const float dvzdz = zp_ok ? (float)dpl<EX>(zm1, wb[1], zp1, zp2, KZ) : 0.f;
        const double l2m = (double)lam + 2.0 * (double)mu;
        sn[0] = upd_snorm<EX>(so[0], dt, l2m, dvxdx, mul_rn(lam, dvzdz));
        sn[1] = upd_snorm<EX>(so[1], dt, l2m, dvzdz, mul_rn(lam, dvxdx));
        sn[2] = upd_sadd<EX>(so[2], dt, (double)mul_rn(mu, add_rn(dvxdz, dvzdx)));
        if constexpr (P == SP_SPIN) {
            // div_vy on vy_c + add_sy_c (solver.py:135-141, 227-228)
            float cm1 = lane_m1(wb[2]), cp1 = lane_p1(wb[2]), cp2 = lane_p1(cp1);
            const float dcdx = xp_ok ? (float)dpl<EX>(wa[2], wb[2], wc[2], wd[2], KX) : 0.f;
            const float dcdz = zp_ok ? (float)dpl<EX>(cm1, wb[2], cp1, cp2, KZ) : 0.f;
            const double dtg = dt * (double)gam;
            sn[4] = upd_srot<EX>(so[4], 2.0 * dt, nu, wb[2], sub_rn(dvzdx, dvxdz));
            sn[3] = upd_sadd<EX>(so[3], dtg, (double)dcdx);
            sn[5] = upd_sadd<EX>(so[5], dtg, (double)dcdz);
        }
    }
}

// sources (forward) or adjoint sources (adjoint) of velocity row r at step
// `it`: float32 adds in list order (stf_dsy / stf_dsxz, solver.py:78-93)
template <int P>
__device__ __forceinline__ void inject(float (&dsv)[n_v(P)], int r, int j, int nz, int &pi, int &src_row, int ni,
                                       cint_p ik, cint_p irow, cflt_p ival, long long nrows, int nt, int it)
{
    if (r != src_row) return;
    const int kend = (r + 1) * nz, kr = r * nz + j;
    while (pi < ni && ik[pi] < kend) {
        const int kk = ik[pi];
        const int row = irow[pi];
        const float s0 = ival[(long long)row * nt + it];
        float s1 = 0.f;
        if constexpr (P != SP_SH) s1 = ival[(nrows + row) * nt + it];
        if (kk == kr) {
            dsv[0] = add_rn(dsv[0], s0);
            if constexpr (P != SP_SH) dsv[1] = add_rn(dsv[1], s1);
        }
        ++pi;
    }
    src_row = pi < ni ? ik[pi] / nz : 0x7fffffff;
}

template <int P, bool EX>
__global__ __launch_bounds__(256) void k_step2(StepArgs a)
{
    constexpr int NS = n_s(P), NV = n_v(P);
    const int nb = gridDim.x, bb = blockIdx.x;
    const int q8 = nb >> 3, r8 = nb & 7, bx8 = bb & 7, by8 = bb >> 3;
    const int lb = bx8 < r8 ? bx8 * (q8 + 1) + by8 : r8 * (q8 + 1) + (bx8 - r8) * q8 + by8;
    const int wave = __builtin_amdgcn_readfirstlane(lb * 4 + (threadIdx.x >> 6));
    if (wave >= a.nitems) return;
    const int L = threadIdx.x & 63;
    const int4 item = a.items[2 * wave];
    if (item.x < 0) return;
    const int4 item2 = a.items[2 * wave + 1];
    const ShotDesc &d = a.shots[item.x];
    const int itx = a.itp ? a.it + *(cint_p)a.itp : a.it;
    const int nx = d.nx, nz = d.nz, i0 = item.z, i1 = item.w;
    const int j = item.y - a.shift + L;
    const int jc = min(max(j, 0), nz - 1);
    const bool jin = j >= 0 && j < nz;
    const bool outl = jin && L >= a.shift && L < a.shift + a.wout;
    const bool zm_ok = j >= 2 && j < nz - 2, zp_ok = j >= 1 && j < nz - 2;
    const int NP4 = (int)(d.Npad * 4);
    const double dt = a.dt;
    const DCo KX{d.c1x, d.c2x, d.k9hx, d.k9lx, d.k1hx, d.k1lx}, KZ{d.c1z, d.c2z, d.k9hz, d.k9lz, d.k1hz, d.k1lz};
    const int rd = a.rd, wr = a.rd ^ 1;
    const double bzb = d.bzb[jc], bzt = d.bzt[jc];
    const bool zint = __all(bzb == 1.0 && bzt == 1.0);
    const rsrc_t RS = make_rsrc(d.state, (long long)n_slots(P) * NP4);
    const rsrc_t RM = make_rsrc(d.mat, (long long)N_MAT * NP4);
    const cflt_p bxp = (cflt_p)d.bx;
    const unsigned vj = (unsigned)jc * 4u;            // per-lane load offset
    const unsigned vo = outl ? (unsigned)j * 4u : OOB;  // per-lane store offset (OOB: not written)

    const cint_p ik = (cint_p)(a.adjoint ? d.rec_k : d.src_k);
    const cint_p irow = (cint_p)(a.adjoint ? d.rec_row : d.src_row);
    const int ni = a.adjoint ? d.nrec : d.nsrc;
    const cflt_p ival = (cflt_p)(a.adjoint ? (const float *)d.adstf : d.stf);
    const long long nrows = a.adjoint ? d.nrec : d.nsrc;
    // injection pointers: step it from row i0-4, step it+1 from row i0-1
    int p1 = __builtin_amdgcn_readfirstlane(a.adjoint ? item2.y : item2.x);
    int p2 = p1;
    while (p2 < ni && ik[p2] < (i0 - 1) * nz) ++p2;
    int row1 = p1 < ni ? ik[p1] / nz : 0x7fffffff;
    int row2 = p2 < ni ? ik[p2] / nz : 0x7fffffff;
    // receivers recorded by this chunk (forward only): rows i0..i1-1
    const cint_p rk = (cint_p)d.rec_k;
    const cint_p rrow = (cint_p)d.rec_row;
    int pr = __builtin_amdgcn_readfirstlane(item2.z);
    int rec_row = (!a.adjoint && pr < d.nrec) ? rk[pr] / nz : 0x7fffffff;
    const rsrc_t RR = make_rsrc(d.rec, (long long)n_rec(P) * d.nrec * a.nt * 4);

    float S0[4][NS], W1[4][NV], S1[4][NS], W2[4][NV];
    // materials of rows m-1 (rho, j) and m-3 (lam, mu, nu, gam) by
    // iteration: three iterations later they are rows m-4 and m-6
    float RW[4], JW[4], LW[4], MW[4], NW[4], GW[4];
#pragma unroll
    for (int u = 0; u < 4; u++) {
        RW[u] = JW[u] = 1.f;
        LW[u] = MW[u] = NW[u] = GW[u] = 0.f;
    }
    Row2<P> PF[4];
#pragma unroll
    for (int u = 0; u < 4; u++) {
#pragma unroll
        for (int f = 0; f < NS; f++) S0[u][f] = S1[u][f] = 0.f;
#pragma unroll
        for (int f = 0; f < NV; f++) W1[u][f] = W2[u][f] = 0.f;
    }

    // the march starts at m0 = i0 - 3: s(n) rows m0-3..m0-1 fill window slots 1..3
    const int m0 = i0 - 3, mend = i1 + 5;
    load2<P>(PF[1], RS, RM, bxp, NP4, rd, m0 - 3, nx, nz, vj, i0, i1);
    load2<P>(PF[2], RS, RM, bxp, NP4, rd, m0 - 2, nx, nz, vj, i0, i1);
    load2<P>(PF[3], RS, RM, bxp, NP4, rd, m0 - 1, nx, nz, vj, i0, i1);
    load2<P>(PF[0], RS, RM, bxp, NP4, rd, m0, nx, nz, vj, i0, i1);
#pragma unroll
    for (int f = 0; f < NS; f++) {
        S0[1][f] = PF[1].s[f];
        S0[2][f] = PF[2].s[f];
        S0[3][f] = PF[3].s[f];
    }
    load2<P>(PF[1], RS, RM, bxp, NP4, rd, m0 + 1, nx, nz, vj, i0, i1);

    for (int mb = m0; mb <= mend; mb += 4) {
#pragma unroll
        for (int u = 0; u < 4; u++) {
            const int m = mb + u;
            if (m > mend) break;
            const int A = (u + 1) & 3, B = (u + 2) & 3, C = (u + 3) & 3, D = u;
            load2<P>(PF[B], RS, RM, bxp, NP4, rd, m + 2, nx, nz, vj, i0, i1);
            const Row2<P> &cur = PF[D];
#pragma unroll
            for (int f = 0; f < NS; f++) S0[D][f] = cur.s[f];

            // ---- A: v(n+1) at row m-1; s(n) rows m-3..m = slots A, B, C, D
            {
                const int r = m - 1;
                const bool rin = r >= 0 && r < nx, xm_ok = r >= 2 && r < nx - 2;
                float dsv[NV], vn[NV];
                vel_row<P, EX>(dsv, S0[A], S0[B], S0[C], S0[D], xm_ok, zm_ok, L, KX, KZ);
                inject<P>(dsv, r, j, nz, p1, row1, ni, ik, irow, ival, nrows, a.nt, itx);
                const float bxr = bxp[min(max(r, 0), nx - 1)];
                vel_finish<P, EX>(vn, dsv, cur.v, cur.rhoA, cur.jA, bxr, bzb, bzt, zint, rin, jin, dt);
                RW[D] = cur.rhoA;
                JW[D] = cur.jA;
#pragma unroll
                for (int f = 0; f < NV; f++) W1[C][f] = vn[f];
            }
            // ---- B: s(n+1) at row m-3; v(n+1) rows m-4..m-1 = slots D, A, B, C
            {
                const int q = m - 3;
                const bool xp_ok = q >= 1 && q < nx - 2;
                float sn[NS];
                str_row<P, EX>(sn, W1[D], W1[A], W1[B], W1[C], S0[A], cur.lamB, cur.muB, cur.nuB, cur.gamB, xp_ok, zp_ok, L, KX, KZ, dt);
                LW[D] = cur.lamB;
                MW[D] = cur.muB;
                NW[D] = cur.nuB;
                GW[D] = cur.gamB;
#pragma unroll
                for (int f = 0; f < NS; f++) S1[A][f] = sn[f];
            }
            // ---- C: v(n+2) at row m-4; s(n+1) rows m-6..m-3 = slots B, C, D, A
            {
                const int r = m - 4;
                const bool rin = r >= 0 && r < nx, xm_ok = r >= 2 && r < nx - 2;
                float dsv[NV], vn[NV];
                vel_row<P, EX>(dsv, S1[B], S1[C], S1[D], S1[A], xm_ok, zm_ok, L, KX, KZ);
                inject<P>(dsv, r, j, nz, p2, row2, ni, ik, irow, ival, nrows, a.nt, itx + 1);
                const float bxr = bxp[min(max(r, 0), nx - 1)];
                vel_finish<P, EX>(vn, dsv, W1[D], RW[A], JW[A], bxr, bzb, bzt, zint, rin, jin, dt);
                float u1[NV], u2[NV];
#pragma unroll
                for (int f = 0; f < NV; f++) {
                    u1[f] = upd_disp<EX>(cur.u[f], W1[D][f], dt);
                    u2[f] = upd_disp<EX>(u1[f], vn[f], dt);
                    W2[D][f] = vn[f];
                }
                // rows outside the chunk: every lane's offset is out of range
                const bool rok = r >= i0 && r < i1;
                const unsigned ov = rok ? vo : OOB;
                const int orow = rok ? r * nz * 4 : 0;
#pragma unroll
                for (int f = 0; f < NV; f++) {
                    bst(RS, ov, slot_u(P, f) * NP4 + orow, u2[f]);
                    bst(RS, ov, slot_v(P, wr, f) * NP4 + orow, vn[f]);
                }
                // adjoint snapshots of either step (solver.py:193-203, 236-242):
                // [isa][u before the step | v after it][NV]
                if (a.snap_v >= 0) {
                    const rsrc_t RP = make_rsrc(d.snap + (long long)a.snap_v * 2 * NV * (NP4 / 4), 2LL * NV * NP4);
#pragma unroll
                    for (int f = 0; f < NV; f++) {
                        bst(RP, ov, f * NP4 + orow, cur.u[f]);
                        bst(RP, ov, (NV + f) * NP4 + orow, W1[D][f]);
                    }
                }
                if (a.snap_u >= 0) {
                    const rsrc_t RP = make_rsrc(d.snap + (long long)a.snap_u * 2 * NV * (NP4 / 4), 2LL * NV * NP4);
#pragma unroll
                    for (int f = 0; f < NV; f++) {
                        bst(RP, ov, f * NP4 + orow, u1[f]);
                        bst(RP, ov, (NV + f) * NP4 + orow, vn[f]);
                    }
                }
                // save_obs of both steps (solver.py:151-156) for receivers in
                // this row; rows past the chunk belong to the next chunk
                if (r == rec_row && r < i1) {
                    const int kend = (r + 1) * nz, kr = r * nz + j;
                    while (pr < d.nrec && rk[pr] < kend) {
                        const int row = rrow[pr];
                        const unsigned orc = (outl && rk[pr] == kr) ? (unsigned)(row * a.nt + itx) * 4u : OOB;
                        const int cs = d.nrec * a.nt * 4;
#pragma unroll
                        for (int f = 0; f < NV; f++) {
                            bst(RR, orc, f * cs, u1[f]);
                            bst(RR, orc, f * cs + 4, u2[f]);
                        }
                        ++pr;
                    }
                    rec_row = pr < d.nrec ? rk[pr] / nz : 0x7fffffff;
                }
            }
            // ---- D: s(n+2) at row m-6; v(n+2) rows m-7..m-4 = slots A, B, C, D
            {
                const int q = m - 6;
                const bool xp_ok = q >= 1 && q < nx - 2;
                float sn[NS];
                str_row<P, EX>(sn, W2[A], W2[B], W2[C], W2[D], S1[B], LW[A], MW[A], NW[A], GW[A], xp_ok, zp_ok, L, KX, KZ, dt);
                const bool qok = q >= i0 && q < i1;
                const unsigned oq = qok ? vo : OOB;
                const int orow = qok ? q * nz * 4 : 0;
#pragma unroll
                for (int f = 0; f < NS; f++) bst(RS, oq, slot_s(P, wr, f) * NP4 + orow, sn[f]);
            }
        }
    }
}

// ------------------------------------- two steps per launch, two cells per lane
// k_step2w is k_step2 with every lane holding two consecutive z cells
// (j0 + 2L and j0 + 2L + 1): a wavefront covers 128 cells, 116 of them
// output, and every load and store moves both cells at once (8 bytes per
// lane), so a row issues half the memory instructions per cell.  The
// arithmetic per cell is the same as in k_step2, operation for operation.
// Needs even nz (a lane's two cells are then both inside or both outside
// the grid and every access is 8-byte aligned); SH and P-SV.
template <int P>
struct Row2W {
    float s[n_s(P)][2];  // s(n), row m
    float v[n_v(P)][2];  // v(n), row m-1
    float u[n_v(P)][2];  // u(n), row m-4
    float rhoA[2];       // row m-1
    float lamB[2], muB[2];  // row m-3
};

typedef unsigned v2u_t __attribute__((ext_vector_type(2)));
__device__ __forceinline__ void bld2(float (&x)[2], rsrc_t r, unsigned voff, int soff)
{
    const v2u_t w = __builtin_amdgcn_raw_buffer_load_b64(r, (int)voff, soff, 0);
    x[0] = __uint_as_float(w.x);
    x[1] = __uint_as_float(w.y);
}
__device__ __forceinline__ void bst2(rsrc_t r, unsigned voff, int soff, const float (&x)[2])
{
    v2u_t w;
    w.x = __float_as_uint(x[0]);
    w.y = __float_as_uint(x[1]);
    __builtin_amdgcn_raw_buffer_store_b64(w, r, (int)voff, soff, 0);
}

template <int P>
__device__ __forceinline__ void load2w(Row2W<P> &p, rsrc_t RS, rsrc_t RM, int NP4, int rd, int mp, int nx, int nz,
                                       unsigned vj, int i0, int i1)
{
    constexpr int NS = n_s(P), NV = n_v(P);
    // the same row clamps as load2
    const int ms = min(max(mp, max(i0 - 6, 0)), min(i1 + 5, nx - 1));
    const int ma = min(max(mp - 1, max(i0 - 4, 0)), min(i1 + 4, nx - 1));
    const int mb = min(max(mp - 3, max(i0 - 3, 0)), min(i1 + 2, nx - 1));
    const int mu_ = min(max(mp - 4, i0), i1 - 1);
    const int rb = nz * 4;
    const int os = ms * rb, oa = ma * rb, ob = mb * rb, ou = mu_ * rb;
#pragma unroll
    for (int f = 0; f < NS; f++) bld2(p.s[f], RS, vj, slot_s(P, rd, f) * NP4 + os);
#pragma unroll
    for (int f = 0; f < NV; f++) {
        bld2(p.v[f], RS, vj, slot_v(P, rd, f) * NP4 + oa);
        bld2(p.u[f], RS, vj, slot_u(P, f) * NP4 + ou);
    }
    bld2(p.rhoA, RM, vj, M_RHO * NP4 + oa);
    bld2(p.muB, RM, vj, M_MU * NP4 + ob);
    p.lamB[0] = p.lamB[1] = 0.f;
    if constexpr (P != SP_SH) bld2(p.lamB, RM, vj, M_LAM * NP4 + ob);
}

// z neighbours of a two-cell field x: cell 0 sits at z = j0 + 2L, cell 1 at
// z + 1; the previous lane holds z - 2, z - 1 and the next z + 2, z + 3
__device__ __forceinline__ void zminus_w(float (&m2)[2], float (&m1)[2], float (&p1)[2], const float (&x)[2])
{
    const float a_m = lane_m1(x[0]), b_m = lane_m1(x[1]), a_p = lane_p1(x[0]);
    m2[0] = a_m;  m1[0] = b_m;  p1[0] = x[1];
    m2[1] = b_m;  m1[1] = x[0]; p1[1] = a_p;
}
__device__ __forceinline__ void zplus_w(float (&m1)[2], float (&p1)[2], float (&p2)[2], const float (&x)[2])
{
    const float b_m = lane_m1(x[1]), a_p = lane_p1(x[0]), b_p = lane_p1(x[1]);
    m1[0] = b_m;  p1[0] = x[1]; p2[0] = a_p;
    m1[1] = x[0]; p1[1] = a_p;  p2[1] = b_p;
}

// vel_row for two cells per lane (div_sy / div_sxz, solver.py:54-66)
template <int P, bool EX>
__device__ __forceinline__ void vel_row_w(float (&dsv)[n_v(P)][2], const float (&sa)[n_s(P)][2],
                                          const float (&sb)[n_s(P)][2], const float (&sc)[n_s(P)][2],
                                          const float (&sd)[n_s(P)][2], bool xm_ok, const bool (&zm_ok)[2],
                                          const DCo &KX, const DCo &KZ)
{
    if constexpr (P == SP_SH) {
        float zm2[2], zm1[2], zp1[2];
        zminus_w(zm2, zm1, zp1, sc[1]);
#pragma unroll
        for (int c = 0; c < 2; c++) {
            double ddx = xm_ok ? dmn<EX>(sa[0][c], sb[0][c], sc[0][c], sd[0][c], KX) : 0.0;
            double ddz = zm_ok[c] ? dmn<EX>(zm2[c], zm1[c], sc[1][c], zp1[c], KZ) : 0.0;
            dsv[0][c] = (float)(ddx + ddz);
        }
    } else {
        float xzm2[2], xzm1[2], xzp1[2], zzm2[2], zzm1[2], zzp1[2];
        zminus_w(xzm2, xzm1, xzp1, sc[2]);
        zminus_w(zzm2, zzm1, zzp1, sc[1]);
#pragma unroll
        for (int c = 0; c < 2; c++) {
            double dxx = xm_ok ? dmn<EX>(sa[0][c], sb[0][c], sc[0][c], sd[0][c], KX) : 0.0;
            double dxxz = xm_ok ? dmn<EX>(sa[2][c], sb[2][c], sc[2][c], sd[2][c], KX) : 0.0;
            double dzxz = zm_ok[c] ? dmn<EX>(xzm2[c], xzm1[c], sc[2][c], xzp1[c], KZ) : 0.0;
            double dzzz = zm_ok[c] ? dmn<EX>(zzm2[c], zzm1[c], sc[1][c], zzp1[c], KZ) : 0.0;
            dsv[0][c] = (float)(dxx + dzxz);
            dsv[1][c] = (float)(dxxz + dzzz);
        }
    }
}

// str_row for two cells per lane (add_sy / add_sxz, solver.py:120-148)
template <int P, bool EX>
__device__ __forceinline__ void str_row_w(float (&sn)[n_s(P)][2], const float (&wa)[n_v(P)][2],
                                          const float (&wb)[n_v(P)][2], const float (&wc)[n_v(P)][2],
                                          const float (&wd)[n_v(P)][2], const float (&so)[n_s(P)][2],
                                          const float (&lam)[2], const float (&mu)[2], bool xp_ok,
                                          const bool (&zp_ok)[2], const DCo &KX, const DCo &KZ, double dt)
{
    if constexpr (P == SP_SH) {
        float zm1[2], zp1[2], zp2[2];
        zplus_w(zm1, zp1, zp2, wb[0]);
#pragma unroll
        for (int c = 0; c < 2; c++) {
            const float dvdx = xp_ok ? (float)dpl<EX>(wa[0][c], wb[0][c], wc[0][c], wd[0][c], KX) : 0.f;
            const float dvdz = zp_ok[c] ? (float)dpl<EX>(zm1[c], wb[0][c], zp1[c], zp2[c], KZ) : 0.f;
            const double dtmu = dt * (double)mu[c];
            sn[0][c] = upd_sadd<EX>(so[0][c], dtmu, (double)dvdx);
            sn[1][c] = upd_sadd<EX>(so[1][c], dtmu, (double)dvdz);
        }
    } else {
        float xm1[2], xp1[2], xp2[2], zm1[2], zp1[2], zp2[2];
        zplus_w(xm1, xp1, xp2, wb[0]);
        zplus_w(zm1, zp1, zp2, wb[1]);
#pragma unroll
        for (int c = 0; c < 2; c++) {
            const float dvxdx = xp_ok ? (float)dpl<EX>(wa[0][c], wb[0][c], wc[0][c], wd[0][c], KX) : 0.f;
            const float dvzdx = xp_ok ? (float)dpl<EX>(wa[1][c], wb[1][c], wc[1][c], wd[1][c], KX) : 0.f;
            const float dvxdz = zp_ok[c] ? (float)dpl<EX>(xm1[c], wb[0][c], xp1[c], xp2[c], KZ) : 0.f;
            const float dvzdz = zp_ok[c] ? (float)dpl<EX>(zm1[c], wb[1][c], zp1[c], zp2[c], KZ) : 0.f;
            const double l2m = (double)lam[c] + 2.0 * (double)mu[c];
            sn[0][c] = upd_snorm<EX>(so[0][c], dt, l2m, dvxdx, mul_rn(lam[c], dvzdz));
            sn[1][c] = upd_snorm<EX>(so[1][c], dt, l2m, dvzdz, mul_rn(lam[c], dvxdx));
            sn[2][c] = upd_sadd<EX>(so[2][c], dt, (double)mul_rn(mu[c], add_rn(dvxdz, dvzdx)));
        }
    }
}

// sources (forward) or adjoint sources (adjoint) of velocity row r, two cells
// per lane: float32 adds in list order (stf_dsy / stf_dsxz, solver.py:78-93)
template <int P>
__device__ __forceinline__ void inject_w(float (&dsv)[n_v(P)][2], int r, int jl, int nz, int &pi, int &src_row,
                                         int ni, cint_p ik, cint_p irow, cflt_p ival, long long nrows, int nt, int it)
{
    if (r != src_row) return;
    const int kend = (r + 1) * nz, kl = r * nz + jl;  // kl: this lane's cell 0
    while (pi < ni && ik[pi] < kend) {
        const int kk = ik[pi];
        const int row = irow[pi];
        const float s0 = ival[(long long)row * nt + it];
        float s1 = 0.f;
        if constexpr (P != SP_SH) s1 = ival[(nrows + row) * nt + it];
#pragma unroll
        for (int c = 0; c < 2; c++) {
            if (kk == kl + c) {
                dsv[0][c] = add_rn(dsv[0][c], s0);
                if constexpr (P != SP_SH) dsv[1][c] = add_rn(dsv[1][c], s1);
            }
        }
        ++pi;
    }
    src_row = pi < ni ? ik[pi] / nz : 0x7fffffff;
}

template <int P, bool EX>
__global__ __launch_bounds__(256) void k_step2w(StepArgs a)
{
    static_assert(P == SP_SH || P == SP_PSV, "two cells per lane: SH and P-SV");
    constexpr int NS = n_s(P), NV = n_v(P);
    const int nb = gridDim.x, bb = blockIdx.x;
    const int q8 = nb >> 3, r8 = nb & 7, bx8 = bb & 7, by8 = bb >> 3;
    const int lb = bx8 < r8 ? bx8 * (q8 + 1) + by8 : r8 * (q8 + 1) + (bx8 - r8) * q8 + by8;
    const int wave = __builtin_amdgcn_readfirstlane(lb * 4 + (threadIdx.x >> 6));
    if (wave >= a.nitems) return;
    const int L = threadIdx.x & 63;
    const int4 item = a.items[2 * wave];
    if (item.x < 0) return;
    const int4 item2 = a.items[2 * wave + 1];
    const ShotDesc &d = a.shots[item.x];
    const int itx = a.itp ? a.it + *(cint_p)a.itp : a.it;
    const int nx = d.nx, nz = d.nz, i0 = item.z, i1 = item.w;
    const int jl = item.y - a.shift + 2 * L;          // cell 0 of this lane (even)
    const bool jin = jl >= 0 && jl < nz;               // both cells (nz even)
    const bool outl = jin && 2 * L >= a.shift && 2 * L < a.shift + a.wout;
    bool zm_ok[2], zp_ok[2];
    double bzb[2], bzt[2];
#pragma unroll
    for (int c = 0; c < 2; c++) {
        const int j = jl + c;
        zm_ok[c] = j >= 2 && j < nz - 2;
        zp_ok[c] = j >= 1 && j < nz - 2;
        const int jc = min(max(j, 0), nz - 1);
        bzb[c] = d.bzb[jc];
        bzt[c] = d.bzt[jc];
    }
    const int jp = min(max(jl, 0), nz - 2);            // clamped pair start
    const int NP4 = (int)(d.Npad * 4);
    const double dt = a.dt;
    const DCo KX{d.c1x, d.c2x, d.k9hx, d.k9lx, d.k1hx, d.k1lx}, KZ{d.c1z, d.c2z, d.k9hz, d.k9lz, d.k1hz, d.k1lz};
    const int rd = a.rd, wr = a.rd ^ 1;
    const bool zint = __all(bzb[0] == 1.0 && bzt[0] == 1.0 && bzb[1] == 1.0 && bzt[1] == 1.0);
    const rsrc_t RS = make_rsrc(d.state, (long long)n_slots(P) * NP4);
    const rsrc_t RM = make_rsrc(d.mat, (long long)N_MAT * NP4);
    const cflt_p bxp = (cflt_p)d.bx;
    const unsigned vj = (unsigned)jp * 4u;
    const unsigned vo = outl ? (unsigned)jl * 4u : OOB;

    const cint_p ik = (cint_p)(a.adjoint ? d.rec_k : d.src_k);
    const cint_p irow = (cint_p)(a.adjoint ? d.rec_row : d.src_row);
    const int ni = a.adjoint ? d.nrec : d.nsrc;
    const cflt_p ival = (cflt_p)(a.adjoint ? (const float *)d.adstf : d.stf);
    const long long nrows = a.adjoint ? d.nrec : d.nsrc;
    int p1 = __builtin_amdgcn_readfirstlane(a.adjoint ? item2.y : item2.x);
    int p2 = p1;
    while (p2 < ni && ik[p2] < (i0 - 1) * nz) ++p2;
    int row1 = p1 < ni ? ik[p1] / nz : 0x7fffffff;
    int row2 = p2 < ni ? ik[p2] / nz : 0x7fffffff;
    const cint_p rk = (cint_p)d.rec_k;
    const cint_p rrow = (cint_p)d.rec_row;
    int pr = __builtin_amdgcn_readfirstlane(item2.z);
    int rec_row = (!a.adjoint && pr < d.nrec) ? rk[pr] / nz : 0x7fffffff;
    const rsrc_t RR = make_rsrc(d.rec, (long long)n_rec(P) * d.nrec * a.nt * 4);

    float S0[4][NS][2], W1[4][NV][2], S1[4][NS][2], W2[4][NV][2];
    float RW[4][2], LW[4][2], MW[4][2];
#pragma unroll
    for (int u = 0; u < 4; u++) {
#pragma unroll
        for (int c = 0; c < 2; c++) {
            RW[u][c] = 1.f;
            LW[u][c] = MW[u][c] = 0.f;
#pragma unroll
            for (int f = 0; f < NS; f++) S0[u][f][c] = S1[u][f][c] = 0.f;
#pragma unroll
            for (int f = 0; f < NV; f++) W1[u][f][c] = W2[u][f][c] = 0.f;
        }
    }
    Row2W<P> PF[4];

    const int m0 = i0 - 3, mend = i1 + 5;
    load2w<P>(PF[1], RS, RM, NP4, rd, m0 - 3, nx, nz, vj, i0, i1);
    load2w<P>(PF[2], RS, RM, NP4, rd, m0 - 2, nx, nz, vj, i0, i1);
    load2w<P>(PF[3], RS, RM, NP4, rd, m0 - 1, nx, nz, vj, i0, i1);
    load2w<P>(PF[0], RS, RM, NP4, rd, m0, nx, nz, vj, i0, i1);
#pragma unroll
    for (int f = 0; f < NS; f++)
#pragma unroll
        for (int c = 0; c < 2; c++) {
            S0[1][f][c] = PF[1].s[f][c];
            S0[2][f][c] = PF[2].s[f][c];
            S0[3][f][c] = PF[3].s[f][c];
        }
    load2w<P>(PF[1], RS, RM, NP4, rd, m0 + 1, nx, nz, vj, i0, i1);

    for (int mb = m0; mb <= mend; mb += 4) {
#pragma unroll
        for (int u = 0; u < 4; u++) {
            const int m = mb + u;
            if (m > mend) break;
            const int A = (u + 1) & 3, B = (u + 2) & 3, C = (u + 3) & 3, D = u;
            load2w<P>(PF[B], RS, RM, NP4, rd, m + 2, nx, nz, vj, i0, i1);
            const Row2W<P> &cur = PF[D];
#pragma unroll
            for (int f = 0; f < NS; f++)
#pragma unroll
                for (int c = 0; c < 2; c++) S0[D][f][c] = cur.s[f][c];

            // ---- A: v(n+1) at row m-1
            {
                const int r = m - 1;
                const bool rin = r >= 0 && r < nx, xm_ok = r >= 2 && r < nx - 2;
                float dsv[NV][2], vn[NV][2];
                vel_row_w<P, EX>(dsv, S0[A], S0[B], S0[C], S0[D], xm_ok, zm_ok, KX, KZ);
                inject_w<P>(dsv, r, jl, nz, p1, row1, ni, ik, irow, ival, nrows, a.nt, itx);
                const float bxr = bxp[min(max(r, 0), nx - 1)];
#pragma unroll
                for (int c = 0; c < 2; c++) {
                    float dc[NV], vc[NV], vo_[NV];
#pragma unroll
                    for (int f = 0; f < NV; f++) {
                        dc[f] = dsv[f][c];
                        vo_[f] = cur.v[f][c];
                    }
                    vel_finish<P, EX>(vc, dc, vo_, cur.rhoA[c], 1.f, bxr, bzb[c], bzt[c], zint, rin, jin, dt);
#pragma unroll
                    for (int f = 0; f < NV; f++) W1[C][f][c] = vc[f];
                    RW[D][c] = cur.rhoA[c];
                }
            }
            // ---- B: s(n+1) at row m-3
            {
                const int q = m - 3;
                const bool xp_ok = q >= 1 && q < nx - 2;
                str_row_w<P, EX>(S1[A], W1[D], W1[A], W1[B], W1[C], S0[A], cur.lamB, cur.muB, xp_ok, zp_ok, KX, KZ, dt);
#pragma unroll
                for (int c = 0; c < 2; c++) {
                    LW[D][c] = cur.lamB[c];
                    MW[D][c] = cur.muB[c];
                }
            }
            // ---- C: v(n+2) at row m-4
            {
                const int r = m - 4;
                const bool rin = r >= 0 && r < nx, xm_ok = r >= 2 && r < nx - 2;
                float dsv[NV][2], u1[NV][2], u2[NV][2];
                vel_row_w<P, EX>(dsv, S1[B], S1[C], S1[D], S1[A], xm_ok, zm_ok, KX, KZ);
                inject_w<P>(dsv, r, jl, nz, p2, row2, ni, ik, irow, ival, nrows, a.nt, itx + 1);
                const float bxr = bxp[min(max(r, 0), nx - 1)];
#pragma unroll
                for (int c = 0; c < 2; c++) {
                    float dc[NV], vc[NV], vo_[NV];
#pragma unroll
                    for (int f = 0; f < NV; f++) {
                        dc[f] = dsv[f][c];
                        vo_[f] = W1[D][f][c];
                    }
                    vel_finish<P, EX>(vc, dc, vo_, RW[A][c], 1.f, bxr, bzb[c], bzt[c], zint, rin, jin, dt);
#pragma unroll
                    for (int f = 0; f < NV; f++) {
                        u1[f][c] = upd_disp<EX>(cur.u[f][c], W1[D][f][c], dt);
                        u2[f][c] = upd_disp<EX>(u1[f][c], vc[f], dt);
                        W2[D][f][c] = vc[f];
                    }
                }
                const bool rok = r >= i0 && r < i1;
                const unsigned ov = rok ? vo : OOB;
                const int orow = rok ? r * nz * 4 : 0;
#pragma unroll
                for (int f = 0; f < NV; f++) {
                    bst2(RS, ov, slot_u(P, f) * NP4 + orow, u2[f]);
                    bst2(RS, ov, slot_v(P, wr, f) * NP4 + orow, W2[D][f]);
                }
                if (a.snap_v >= 0) {
                    const rsrc_t RP = make_rsrc(d.snap + (long long)a.snap_v * 2 * NV * (NP4 / 4), 2LL * NV * NP4);
#pragma unroll
                    for (int f = 0; f < NV; f++) {
                        bst2(RP, ov, f * NP4 + orow, cur.u[f]);
                        bst2(RP, ov, (NV + f) * NP4 + orow, W1[D][f]);
                    }
                }
                if (a.snap_u >= 0) {
                    const rsrc_t RP = make_rsrc(d.snap + (long long)a.snap_u * 2 * NV * (NP4 / 4), 2LL * NV * NP4);
#pragma unroll
                    for (int f = 0; f < NV; f++) {
                        bst2(RP, ov, f * NP4 + orow, u1[f]);
                        bst2(RP, ov, (NV + f) * NP4 + orow, W2[D][f]);
                    }
                }
                // save_obs of both steps for receivers in this row
                if (r == rec_row && r < i1) {
                    const int kend = (r + 1) * nz, kl = r * nz + jl;
                    while (pr < d.nrec && rk[pr] < kend) {
                        const int row = rrow[pr];
                        const int c = rk[pr] - kl;  // 0 or 1 in the lane holding the receiver
                        const unsigned orc = (outl && (c == 0 || c == 1)) ? (unsigned)(row * a.nt + itx) * 4u : OOB;
                        const int cs = d.nrec * a.nt * 4;
#pragma unroll
                        for (int f = 0; f < NV; f++) {
                            bst(RR, orc, f * cs, c == 1 ? u1[f][1] : u1[f][0]);
                            bst(RR, orc, f * cs + 4, c == 1 ? u2[f][1] : u2[f][0]);
                        }
                        ++pr;
                    }
                    rec_row = pr < d.nrec ? rk[pr] / nz : 0x7fffffff;
                }
            }
            // ---- D: s(n+2) at row m-6
            {
                const int q = m - 6;
                const bool xp_ok = q >= 1 && q < nx - 2;
                float sn[NS][2];
                str_row_w<P, EX>(sn, W2[A], W2[B], W2[C], W2[D], S1[B], LW[A], MW[A], xp_ok, zp_ok, KX, KZ, dt);
                const bool qok = q >= i0 && q < i1;
                const unsigned oq = qok ? vo : OOB;
                const int orow = qok ? q * nz * 4 : 0;
#pragma unroll
                for (int f = 0; f < NS; f++) bst2(RS, oq, slot_s(P, wr, f) * NP4 + orow, sn[f]);
            }
        }
    }
}

// save_obs (solver.py:151-156) for every receiver of a batch, after a step:
// one thread per (shot, receiver) of the batch.
__global__ __launch_bounds__(256) void k_record(const ShotDesc *shots, int nshots, int maxrec, int it, int nt, int P,
                                                const int *itp)
{
    if (itp) it += *itp;
    const int t = blockIdx.x * 256 + threadIdx.x;
    const int s = t / maxrec, q = t - s * maxrec;
    if (s >= nshots) return;
    const ShotDesc &d = shots[s];
    if (q >= d.nrec) return;
    const int k = d.rec_k[q], row = d.rec_row[q];
    const int nv = n_v(P);
    for (int f = 0; f < nv; f++)
        d.rec[((long long)f * d.nrec + row) * nt + it] = d.state[(long long)(2 * n_s(P) + 2 * nv + f) * d.Npad + k];
}

// advances the step counter of a replayed graph (one thread, vector store)
__global__ void k_advance(int *counter, int steps)
{
    if (threadIdx.x == 0 && blockIdx.x == 0) counter[0] += steps;
}

// Adjoint snapshots (solver.py:193-203, 236-242): u before a step (part 0)
// or the new v after it (part 1) into snapshot isa, for every shot.
__global__ __launch_bounds__(256) void k_snap(const ShotDesc *shots, int nshots, long long maxN, int isa, int part,
                                              int vset, int P)
{
    const long long k = (long long)blockIdx.x * 256 + threadIdx.x;
    const int s = blockIdx.y;
    if (s >= nshots) return;
    const ShotDesc &d = shots[s];
    if (k >= (long long)d.nx * d.nz) return;
    const int nv = n_v(P), ns = n_s(P);
    const long long NP = d.Npad;
    float *dst = d.snap + ((long long)isa * 2 * nv + part * nv) * NP + k;
    const float *src = d.state + (part == 0 ? (long long)(2 * ns + 2 * nv) : (long long)(2 * ns + vset * nv)) * NP + k;
    for (int f = 0; f < nv; f++) dst[f * NP] = src[f * NP];
}

// ------------------------------------------------------ interaction kernel
struct InteractArgs {
    const ShotDesc *shots;
    const int *list;  // shots of this grid in the batch
    int nlist;
    int nx, nz;
    long long Npad;
    float *kern;      // [SP_NKERN][Npad]
    int cur;          // velocity set holding the adjoint velocity
    int isae;
    double ndt;
    DCo KX, KZ;  // derivative constants (the exact split form is used)
};

// D+ in x and z in the reference's operation order (runs every sae steps,
// so always in the exact form)
__device__ __forceinline__ float dpx(const float *u, int i, long long k, int nx, int nz, const DCo &K)
{
    if (i >= 1 && i < nx - 2)
        return (float)dpl<true>(u[k - nz], u[k], u[k + nz], u[k + 2 * (long long)nz], K);
    return 0.f;
}
__device__ __forceinline__ float dpz(const float *u, int j, long long k, int nz, const DCo &K)
{
    if (j >= 1 && j < nz - 2) return (float)dpl<true>(u[k - 1], u[k], u[k + 1], u[k + 2], K);
    return 0.f;
}

// SH: interaction_muy on D+ of the adjoint and the forward uy
// (solver.py:159-163, 315-321).  P-SV and spin: this library's kernels,
// accumulated in float64 from the float32 derivative fields:
//   k_lam += div_a div ndt;  k_mu += (2 exx_a exx + 2 ezz_a ezz + g_a g) ndt
//   k_rho += (vx_a vx + vz_a vz) ndt;  spin: k_nu += 4 E_a E ndt,
//   k_j += vyc_a vyc ndt, k_muc += grad phi_a . grad phi ndt,
//   with g = dz ux + dx uz and E = -uy_c - (dx uz - dz ux)/2.
template <int P>
__global__ __launch_bounds__(256) void k_interact(InteractArgs a)
{
#pragma clang fp contract(off)
    const long long N = (long long)a.nx * a.nz;
    const long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= N) return;
    const int i = (int)(k / a.nz), j = (int)(k - (long long)i * a.nz);
    const int nx = a.nx, nz = a.nz;
    const long long NP = a.Npad;
    float *K = a.kern;
    if constexpr (P == SP_SH) {
        float km = K[SP_K_MU * NP + k];
        double kh = K[SP_K_HESS * NP + k], kr = K[SP_K_RHO * NP + k];
        for (int s = 0; s < a.nlist; s++) {
            const ShotDesc &d = a.shots[a.list[s]];
            const float *ua = d.state + slot_u(P, 0) * NP;
            const float *uf = d.snap + (long long)a.isae * n_snap(P) * NP;
            const float ax = dpx(ua, i, k, nx, nz, a.KX), fx = dpx(uf, i, k, nx, nz, a.KX);
            const float az = dpz(ua, j, k, nz, a.KZ), fz = dpz(uf, j, k, nz, a.KZ);
            km = (float)((double)km + (double)add_rn(mul_rn(ax, fx), mul_rn(az, fz)) * a.ndt);
            const double vf = uf[NP + k], va = d.state[slot_v(P, a.cur, 0) * NP + k];
            kr = (double)(float)(kr + va * vf * a.ndt);
            kh = (double)(float)(kh + vf * vf * a.ndt);
        }
        K[SP_K_MU * NP + k] = km;
        K[SP_K_RHO * NP + k] = (float)kr;
        K[SP_K_HESS * NP + k] = (float)kh;
        return;
    }
    double kl = K[SP_K_LAM * NP + k], kmu = K[SP_K_MU * NP + k], kr = K[SP_K_RHO * NP + k];
    double kh = K[SP_K_HESS * NP + k];
    double knu = 0, kj = 0, kc = 0;
    if constexpr (P == SP_SPIN) {
        knu = K[SP_K_NU * NP + k];
        kj = K[SP_K_J * NP + k];
        kc = K[SP_K_MUC * NP + k];
    }
    for (int s = 0; s < a.nlist; s++) {
        const ShotDesc &d = a.shots[a.list[s]];
        const float *uxa = d.state + slot_u(P, 0) * NP, *uza = d.state + slot_u(P, 1) * NP;
        const float *sn = d.snap + (long long)a.isae * n_snap(P) * NP;
        const float *uxf = sn, *uzf = sn + NP;
        const double axx = dpx(uxa, i, k, nx, nz, a.KX), axz = dpz(uxa, j, k, nz, a.KZ);
        const double azx = dpx(uza, i, k, nx, nz, a.KX), azz = dpz(uza, j, k, nz, a.KZ);
        const double fxx = dpx(uxf, i, k, nx, nz, a.KX), fxz = dpz(uxf, j, k, nz, a.KZ);
        const double fzx = dpx(uzf, i, k, nx, nz, a.KX), fzz = dpz(uzf, j, k, nz, a.KZ);
        const double vxa = d.state[slot_v(P, a.cur, 0) * NP + k], vza = d.state[slot_v(P, a.cur, 1) * NP + k];
        const double vxf = sn[(n_v(P) + 0) * NP + k], vzf = sn[(n_v(P) + 1) * NP + k];
        // the same expressions, in the same order, as the oracle
        const double diva = axx + azz, divf = fxx + fzz, sha = axz + azx, shf = fxz + fzx;
        kl = (double)(float)(kl + diva * divf * a.ndt);
        kmu = (double)(float)(kmu + (2.0 * axx * fxx + 2.0 * azz * fzz + sha * shf) * a.ndt);
        kr = (double)(float)(kr + (vxa * vxf + vza * vzf) * a.ndt);
        kh = (double)(float)(kh + (vxf * vxf + vzf * vzf) * a.ndt);
        if constexpr (P == SP_SPIN) {
            const float *uca = d.state + slot_u(P, 2) * NP, *ucf = sn + 2 * NP;
            const double ea = -(double)uca[k] - 0.5 * (azx - axz);
            const double ef = -(double)ucf[k] - 0.5 * (fzx - fxz);
            knu = (double)(float)(knu + 4.0 * ea * ef * a.ndt);
            const double vca = d.state[slot_v(P, a.cur, 2) * NP + k], vcf = sn[5 * NP + k];
            kj = (double)(float)(kj + vca * vcf * a.ndt);
            const double cax = dpx(uca, i, k, nx, nz, a.KX), caz = dpz(uca, j, k, nz, a.KZ);
            const double cfx = dpx(ucf, i, k, nx, nz, a.KX), cfz = dpz(ucf, j, k, nz, a.KZ);
            kc = (double)(float)(kc + (cax * cfx + caz * cfz) * a.ndt);
        }
    }
    K[SP_K_LAM * NP + k] = (float)kl;
    K[SP_K_MU * NP + k] = (float)kmu;
    K[SP_K_RHO * NP + k] = (float)kr;
    K[SP_K_HESS * NP + k] = (float)kh;
    if constexpr (P == SP_SPIN) {
        K[SP_K_NU * NP + k] = (float)knu;
        K[SP_K_J * NP + k] = (float)kj;
        K[SP_K_MUC * NP + k] = (float)kc;
    }
}

// ---------------------------------------------------------- misfit kernel
// diff kernel of misfit/waveform.py:5-25 plus the norm's partial sums.
// One block per (shot, component, receiver) triple of the batch.
struct MisfitArgs {
    const ShotDesc *shots;
    const int4 *triples;  // (shot, comp, rec, -)
    const float *const *data;  // per shot, [n_rec][nrec][nt]
    double *partial;
    const double *taper;  // [nt], computed on the host with the C library's cos
    int n, nt;
};

__global__ __launch_bounds__(256) void k_misfit(MisfitArgs a)
{
    const int b = blockIdx.x;
    if (b >= a.n) return;
    const int4 t = a.triples[b];
    const ShotDesc &d = a.shots[t.x];
    const int nt = a.nt;
    const long long base = ((long long)t.y * d.nrec + t.z) * nt;
    const float *syn = a.data[t.x] + base;  // observed data ("syn" in fd2d.py:305)
    const float *obs = d.rec + base;         // current synthetics
    float *ad = d.adstf + base;
    double acc = 0.0;
    for (int it = threadIdx.x; it < nt; it += blockDim.x) {
        const double taper = a.taper[it];
        const float res = (float)((double)sub_rn(syn[it], obs[it]) * taper);
        ad[nt - it - 1] = (float)((double)res * taper * 2.0);
        acc += (double)res * (double)res;
    }
    __shared__ double red[256];
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) a.partial[b] = red[0];
}

// ------------------------------------------------------- smoothing kernels
// apply_gauxxian_x / _z with init_gausian's normaliser (fd2d.py:71-108).
// g[d] = exp(-d^2/(2 s^2)) / (sqrt(2 pi) s) for |i-n| = d, in float64.
__global__ __launch_bounds__(256) void k_smooth_x(const float *data, float *gtmp, const double *g, int nx, int nz)
{
#pragma clang fp contract(off)
    const long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= (long long)nx * nz) return;
    const int i = (int)(k / nz), j = (int)(k - (long long)i * nz);
    double s = 0.0;
    for (int n = 0; n < nx; n++) s += g[i > n ? i - n : n - i] * (double)data[(long long)n * nz + j];
    gtmp[k] = (float)s;
}

__global__ __launch_bounds__(256) void k_smooth_z(float *data, const float *gtmp, const double *g,
                                                  const double *sumx, const double *sumz, int nx, int nz)
{
#pragma clang fp contract(off)
    const long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= (long long)nx * nz) return;
    const int i = (int)(k / nz), j = (int)(k - (long long)i * nz);
    const float *row = gtmp + (long long)i * nz;
    double s = 0.0;
    for (int n = 0; n < nz; n++) s += g[j > n ? j - n : n - j] * (double)row[n];
    const float gsum = (float)(sumx[i] * sumz[j]);
    data[k] = (float)(s / (double)gsum);
}

// ---------------------------------------------------------- small kernels
__global__ void k_fill(float *p, long long n, float v)
{
    const long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n) p[k] = v;
}

// vps2lm: fd2d.py:21-33 (lam holds vp and mu holds vs on entry)
__global__ void k_vps2lm(float *lam, float *mu, const float *rho, long long n)
{
#pragma clang fp contract(off)
    const long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const float vp = lam[k], vs = mu[k];
    if (vp > vs) lam[k] = (float)((double)rho[k] * ((double)(vp * vp) - 2.0 * (double)vs * (double)vs));
    else lam[k] = 0.f;
    mu[k] = rho[k] * vs * vs;
}

__global__ void k_gamma(float *gam, const float *muc, const float *nuc, long long n)
{
    const long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n) gam[k] = add_rn(muc[k], nuc[k]);
}

// vp/vs/rho kernels from lam/mu/rho kernels: lam = rho (vp^2 - 2 vs^2),
// mu = rho vs^2 (the parameterisation of vps2lm).
__global__ void k_kern_vps(float *K, const float *mat, long long n, long long NP)
{
#pragma clang fp contract(off)
    const long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const double rho = mat[M_RHO * NP + k], lam = mat[M_LAM * NP + k], mu = mat[M_MU * NP + k];
    const double vp = sqrt((lam + 2.0 * mu) / rho), vs = sqrt(mu / rho);
    const double kl = K[SP_K_LAM * NP + k], km = K[SP_K_MU * NP + k], kr = K[SP_K_RHO * NP + k];
    K[SP_K_VP * NP + k] = (float)(2.0 * rho * vp * kl);
    K[SP_K_VS * NP + k] = (float)(2.0 * rho * vs * (km - 2.0 * kl));
    K[SP_K_RHOP * NP + k] = (float)((vp * vp - 2.0 * vs * vs) * kl + vs * vs * km + kr);
}

// ===================================================================== host
struct Grid {
    int nx = 0, nz = 0;
    float dx = 0, dz = 0;
    long long N = 0, Npad = 0;
    float *mat = nullptr;   // [N_MAT][Npad]
    float *bx = nullptr;    // [nx]
    double *bz = nullptr;   // [2][nz]: bottom, top
    float *kern = nullptr;  // [SP_NKERN][Npad]
};

struct Shot {
    int grid = 0, nsrc = 0, nrec = 0;
    std::vector<int> src_sorted, rec_sorted;  // grid indices in list order
    int *d_idx = nullptr;  // src_k, src_row, rec_k, rec_row
    float *stf = nullptr, *state = nullptr, *rec = nullptr, *data = nullptr, *adstf = nullptr;
    float *snap = nullptr;  // into the context's snapshot pool
    float *state2 = nullptr;  // second state block: forward recomputation from checkpoints
    float *ckpt = nullptr;    // [nckpt][n_s + 2 n_v][Npad] full forward states
    size_t ckpt_bytes = 0;
};

}  // namespace

struct sp_ctx {
    int device = 0;
    hipStream_t own = nullptr, stream = nullptr;
    int physics = 0, nt = 0;
    double dt = 0;
    std::vector<Grid> grids;
    std::vector<Shot> shots;
    std::string err;
    // tuning
    int chunk = 0;         // 0: automatic
    int order = 0;
    int kernel = 1;
    int cur_kernel = 1;        // kernel of the prepared batch
    int shift = 3, wout = 58;  // strip geometry of the step kernel
    int fuse = 1;              // two steps per launch where nothing observes the state between them
    int exact = 1;             // step kernels follow the reference's float64 operations exactly
    int *d_counter = nullptr;  // step counter read by the kernels of a replayed graph
    const int *itp = nullptr;  // set while a graph is captured
    int cur_fuse = 0;          // the prepared batch uses the two-step strip geometry
    int graph = 0;
    long long snap_mb = 0;
    // snapshot pool
    float *snap_pool = nullptr;
    size_t snap_pool_bytes = 0;
    int snaps_valid = 0, snaps_sae = 0;  // snapshots of every shot held from the last sp_forward
    int cur_set = 0;      // stress/velocity set holding the state after the last forward run
    int fwd_next = -1;    // next step of a forward run advanced by sp_forward_steps (-1: none)
    // batch descriptors
    ShotDesc *d_desc = nullptr;
    size_t d_desc_cap = 0;
    ShotDesc *d_desc2 = nullptr;  // the same shots on their second state block
    size_t d_desc2_cap = 0;
    float *seg_pool = nullptr;    // snapshots of one recomputed segment
    size_t seg_pool_bytes = 0;
    int ckpt_C = 0, ckpt_valid = 0;
    int4 *d_items = nullptr;
    size_t d_items_cap = 0;
    int4 *d_items_w = nullptr;  // two cells per lane: the two-step work list
    size_t d_items_w_cap = 0;
    int n_items_w = 0;
    int *d_list = nullptr;
    size_t d_list_cap = 0;
    // timing of the last run
    hipEvent_t ev0 = nullptr, ev1 = nullptr, ev2 = nullptr;
    long long t_launches = 0;
    double t_step_ms = 0, t_total_ms = 0, t_bytes = 0;
};

namespace {

#define HIPCHECK(ctx, call)                                                                  \
    do {                                                                                     \
        hipError_t e_ = (call);                                                              \
        if (e_ != hipSuccess) {                                                              \
            (ctx)->err = std::string(#call) + ": " + hipGetErrorString(e_);                   \
            return 2;                                                                        \
        }                                                                                    \
    } while (0)

// Runs `undo` when the scope ends unless dismissed: partial allocations of
// a failed call are released and the context is left as it was.
template <typename F>
struct ScopeUndo {
    F undo;
    bool armed = true;
    ~ScopeUndo() { if (armed) undo(); }
};
template <typename F>
ScopeUndo<F> on_failure(F f) { return ScopeUndo<F>{f}; }

// device buffer owned for the length of one call
template <typename T>
struct DevTmp {
    T *p = nullptr;
    DevTmp() = default;
    DevTmp(const DevTmp &) = delete;
    DevTmp &operator=(const DevTmp &) = delete;
    ~DevTmp() { if (p) (void)hipFree(p); }
};

int fail(sp_ctx *ctx, const std::string &msg)
{
    ctx->err = msg;
    return 1;
}

long long pad64(long long n) { return (n + 63) / 64 * 64; }

int dcopy(sp_ctx *ctx, void *dst, const void *src, size_t bytes, int dst_dev, int src_dev)
{
    hipMemcpyKind kind = dst_dev ? (src_dev ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice)
                                 : (src_dev ? hipMemcpyDeviceToHost : hipMemcpyHostToHost);
    HIPCHECK(ctx, hipMemcpyAsync(dst, src, bytes, kind, ctx->stream));
    if (!dst_dev || !src_dev) HIPCHECK(ctx, hipStreamSynchronize(ctx->stream));
    return 0;
}

int fill(sp_ctx *ctx, float *p, long long n, float v)
{
    if (n <= 0) return 0;
    hipLaunchKernelGGL(k_fill, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, ctx->stream, p, n, v);
    HIPCHECK(ctx, hipGetLastError());
    return 0;
}

template <typename T>
int ensure(sp_ctx *ctx, T *&p, size_t &cap, size_t n)
{
    if (n <= cap) return 0;
    if (p) HIPCHECK(ctx, hipFree(p));
    p = nullptr;
    cap = 0;  // a failed allocation leaves an empty buffer, never a stale size
    HIPCHECK(ctx, hipMalloc(&p, n * sizeof(T)));
    cap = n;
    return 0;
}

int check_ctx(sp_ctx *ctx)
{
    if (!ctx) return 3;
    if (!ctx->physics) return fail(ctx, "sp_set_physics has not been called");
    HIPCHECK(ctx, hipSetDevice(ctx->device));
    return 0;
}

// sizes of one shot's buffers
long long rec_floats(const sp_ctx *c, const Shot &s) { return (long long)n_rec(c->physics) * s.nrec * c->nt; }
long long adj_floats(const sp_ctx *c, const Shot &s) { return (long long)n_mis(c->physics) * s.nrec * c->nt; }
size_t snap_bytes(const sp_ctx *c, const Shot &s, int sae)
{
    return (size_t)(c->nt / sae) * n_snap(c->physics) * (size_t)c->grids[s.grid].Npad * sizeof(float);
}

// num/den as hi + lo: hi = RN(num/den), lo = RN((num - hi den)/den), the
// remainder num - hi den being exact by FMA
void split_quotient(double num, double den, double &hi, double &lo)
{
    hi = num / den;
    lo = std::fma(-hi, den, num) / den;
}

// Exact mode multiplies by the split constant, fma(x, hi, x lo), where the
// reference divides, RN(num x / den).  For float32 x and a divisor 8h or 24h
// of float32 h, num x / den is an integer of at most 28 bits over an odd
// integer below 2^26: never a float64 rounding midpoint and at least 2^-80
// (relative) from each, while the split product is within ~2^-105 of it, so
// both round alike.  The identity is still checked here once per constant
// over all 2^23 float32 mantissas (the exponent of x only scales both sides).
bool split_quotient_exact(double num, double den)
{
    static std::mutex mu;
    static std::map<std::pair<double, double>, bool> known;
    std::lock_guard<std::mutex> lock(mu);
    auto f = known.find({num, den});
    if (f != known.end()) return f->second;
    double hi, lo;
    split_quotient(num, den, hi, lo);
    bool ok = std::isfinite(hi) && std::isfinite(lo);
    for (long m = 1L << 23; ok && m < (1L << 24); m++) {
        const double x = (double)m;
        ok = std::fma(x, hi, x * lo) == (num * x) / den;
    }
    known[{num, den}] = ok;
    return ok;
}

int exact_mode_ok(sp_ctx *c, const std::vector<int> &batch)
{
    if (!c->exact) return 0;
    for (int s : batch) {
        const Grid &g = c->grids[c->shots[s].grid];
        for (double h : {(double)g.dx, (double)g.dz})
            if (!split_quotient_exact(9.0, 8.0 * h) || !split_quotient_exact(1.0, 24.0 * h))
                return fail(c, "exact mode: the derivative constants of grid spacing " + std::to_string(h) +
                                   " do not reproduce the reference's divisions; use the default mode");
    }
    return 0;
}

ShotDesc make_desc(const sp_ctx *c, const Shot &s)
{
    const Grid &g = c->grids[s.grid];
    ShotDesc d;
    d.mat = g.mat;
    d.bx = g.bx;
    d.bzb = g.bz;
    d.bzt = g.bz + g.nz;
    d.state = s.state;
    d.rec = s.rec;
    d.snap = s.snap;
    d.stf = s.stf;
    d.adstf = s.adstf;
    d.src_k = s.d_idx;
    d.src_row = s.d_idx + s.nsrc;
    d.rec_k = s.d_idx + 2 * s.nsrc;
    d.rec_row = s.d_idx + 2 * s.nsrc + s.nrec;
    d.nsrc = s.nsrc;
    d.nrec = s.nrec;
    d.nx = g.nx;
    d.nz = g.nz;
    d.Npad = g.Npad;
    d.c1x = 9.0 / (8.0 * (double)g.dx);
    d.c2x = 1.0 / (24.0 * (double)g.dx);
    d.c1z = 9.0 / (8.0 * (double)g.dz);
    d.c2z = 1.0 / (24.0 * (double)g.dz);
    split_quotient(9.0, 8.0 * (double)g.dx, d.k9hx, d.k9lx);
    split_quotient(1.0, 24.0 * (double)g.dx, d.k1hx, d.k1lx);
    split_quotient(9.0, 8.0 * (double)g.dz, d.k9hz, d.k9lz);
    split_quotient(1.0, 24.0 * (double)g.dz, d.k1hz, d.k1lz);
    return d;
}

// The buffer-addressed kernel needs every per-shot buffer below the
// out-of-range store offset; larger grids use the plain kernel.
int batch_kernel(const sp_ctx *c, const std::vector<int> &batch)
{
    if (c->kernel != 1) return 0;
    for (int s : batch) {
        const Grid &g = c->grids[c->shots[s].grid];
        const long long bytes = (long long)std::max(n_slots(c->physics), N_MAT) * g.Npad * 4;
        if (bytes >= (long long)OOB) return 0;
    }
    return 1;
}

// strip geometry of the prepared batch: the two-step kernel needs 6 halo
// lanes (and the single-step kernel runs on the same strips in that mode)
int strip_shift(const sp_ctx *c) { return c->cur_fuse ? 6 : (c->cur_kernel == 1 ? c->shift : 3); }
int strip_width(const sp_ctx *c) { return c->cur_fuse ? 52 : (c->cur_kernel == 1 ? c->wout : 58); }

// 0: one step per launch; 1: two steps per launch (k_step2); 2: two steps
// per launch with two cells per lane (k_step2w: SH and P-SV, every grid of
// the batch with even nz; single steps keep the k_step2 strips)
int fuse_ok(const sp_ctx *c, const std::vector<int> &batch)
{
    if (!c->fuse || batch_kernel(c, batch) != 1) return 0;
    if (c->fuse == 2 && c->physics != SP_SPIN) {
        bool even = true;
        for (int s : batch) even = even && c->grids[c->shots[s].grid].nz % 2 == 0;
        if (even) return 2;
    }
    return 1;
}

// Steps the launch at step `it` may take: two when fusion is on and the
// state after step it is observed by nothing on the host side -- no
// interaction after step it, no checkpoint before step it+1 -- and both
// steps are before `end`.  Adjoint snapshots do not split a pair: the
// two-step kernel writes the halves that fall between its steps.
int steps_next(const sp_ctx *c, int it, int end, int inter_sae, int ckpt_C)
{
    if (!c->cur_fuse || it + 2 > end) return 1;
    if (inter_sae > 0 && it % inter_sae == 0) return 1;
    if (ckpt_C > 0 && (it + 1) % ckpt_C == 0) return 1;
    return 2;
}

int auto_chunk(const sp_ctx *c, const std::vector<int> &batch)
{
    if (c->chunk > 0) return batch_kernel(c, batch) == 1 ? (c->chunk + 3) / 4 * 4 : c->chunk;
    // enough wavefronts to fill 256 CUs several times over, long chunks otherwise
    long long waves64 = 0;
    for (int s : batch) {
        const Grid &g = c->grids[c->shots[s].grid];
        waves64 += (long long)((g.nz + 57) / 58) * ((g.nx + 63) / 64);
    }
    if (waves64 >= 2048) return 16;
    if (waves64 >= 256) return 8;
    return 4;
}

// Upload descriptors and the wavefront work list of a batch.
int prepare_batch(sp_ctx *c, const std::vector<int> &batch, int *nitems_out)
{
    if (exact_mode_ok(c, batch)) return 2;
    std::vector<ShotDesc> desc;
    for (int s : batch) desc.push_back(make_desc(c, c->shots[s]));
    const int chunk = auto_chunk(c, batch);
    c->cur_kernel = batch_kernel(c, batch);
    c->cur_fuse = fuse_ok(c, batch);
    std::vector<int4> items;
    int W = 0;
    auto push = [&](int b, int j0, int i0) {
        const Shot &sh = c->shots[batch[b]];
        const Grid &g = c->grids[sh.grid];
        const int i1 = std::min(i0 + chunk, g.nx);
        auto first = [](const std::vector<int> &v, long long key) {
            return (int)(std::lower_bound(v.begin(), v.end(), (int)std::max(key, -1LL)) - v.begin());
        };
        items.push_back(make_int4(b, j0, i0, i1));
        // list pointers from row i0-4 (the two-step kernel's first velocity
        // row; the single-step kernels skip forward to row i0-1)
        items.push_back(make_int4(first(sh.src_sorted, (long long)(i0 - 4) * g.nz),
                                  first(sh.rec_sorted, (long long)(i0 - 4) * g.nz),
                                  first(sh.rec_sorted, (long long)i0 * g.nz), 0));
    };
    auto build = [&](int width) {
        W = width;
        items.clear();
        if (c->order == 3) {
            // strip-group-major: a block's four wavefronts are four z-adjacent
            // strips of one shot (one ~1 KB contiguous segment per field row);
            // the next blocks take the same strips of the following shots of
            // the grid, which reuse the material rows from L2.
            std::vector<std::vector<int>> per_grid(c->grids.size());
            for (int b = 0; b < (int)batch.size(); b++) per_grid[c->shots[batch[b]].grid].push_back(b);
            for (size_t g = 0; g < per_grid.size(); g++) {
                const Grid &G = c->grids[g];
                const int nstrip = (G.nz + W - 1) / W;
                for (int i0 = 0; i0 < G.nx; i0 += chunk)
                    for (int sg = 0; sg < nstrip; sg += 4)
                        for (int b : per_grid[g]) {
                            for (int sp = sg; sp < sg + 4; sp++) {
                                if (sp < nstrip) push(b, sp * W, i0);
                                else items.push_back(make_int4(-1, 0, 0, 0)), items.push_back(make_int4(0, 0, 0, 0));
                            }
                        }
            }
        } else if (c->order == 2) {
            // tile-major: groups of four shots of one grid walk the tiles
            // (chunk-major, strips inner); a block's four wavefronts are the
            // four shots of one tile, so they share its material rows, and
            // consecutive blocks are neighbouring strips, so they share halos.
            std::vector<std::vector<int>> per_grid(c->grids.size());
            for (int b = 0; b < (int)batch.size(); b++) per_grid[c->shots[batch[b]].grid].push_back(b);
            for (size_t g = 0; g < per_grid.size(); g++) {
                const auto &list = per_grid[g];
                const Grid &G = c->grids[g];
                for (size_t s0 = 0; s0 < list.size(); s0 += 4) {
                    const size_t s1 = std::min(list.size(), s0 + 4);
                    for (int i0 = 0; i0 < G.nx; i0 += chunk)
                        for (int j0 = 0; j0 < G.nz; j0 += W)
                            for (size_t q = s0; q < s1; q++) push(list[q], j0, i0);
                }
            }
        } else if (c->order == 1) {
            // shot-inner: consecutive wavefronts of a block share a tile of the
            // material and differ in shot (same grid) or are neighbours in z
            int maxs = 0, maxc = 0;
            for (int s : batch) {
                const Grid &g = c->grids[c->shots[s].grid];
                maxs = std::max(maxs, (g.nz + W - 1) / W);
                maxc = std::max(maxc, (g.nx + chunk - 1) / chunk);
            }
            for (int ch = 0; ch < maxc; ch++)
                for (int sp = 0; sp < maxs; sp++)
                    for (int b = 0; b < (int)batch.size(); b++) {
                        const Grid &g = c->grids[c->shots[batch[b]].grid];
                        const int i0 = ch * chunk, j0 = sp * W;
                        if (i0 >= g.nx || j0 >= g.nz) continue;
                        push(b, j0, i0);
                    }
        } else {
            for (int b = 0; b < (int)batch.size(); b++) {
                const Grid &g = c->grids[c->shots[batch[b]].grid];
                for (int i0 = 0; i0 < g.nx; i0 += chunk)
                    for (int j0 = 0; j0 < g.nz; j0 += W) push(b, j0, i0);
            }
        }
    };
    if (c->cur_fuse == 2) {
        build(116);
        if (ensure(c, c->d_items_w, c->d_items_w_cap, items.size())) return 2;
        HIPCHECK(c, hipMemcpyAsync(c->d_items_w, items.data(), items.size() * sizeof(int4), hipMemcpyHostToDevice,
                                   c->stream));
        HIPCHECK(c, hipStreamSynchronize(c->stream));
        c->n_items_w = (int)items.size() / 2;
    }
    build(strip_width(c));
    if (ensure(c, c->d_desc, c->d_desc_cap, desc.size())) return 2;
    if (ensure(c, c->d_items, c->d_items_cap, items.size())) return 2;
    HIPCHECK(c, hipMemcpyAsync(c->d_desc, desc.data(), desc.size() * sizeof(ShotDesc), hipMemcpyHostToDevice, c->stream));
    HIPCHECK(c, hipMemcpyAsync(c->d_items, items.data(), items.size() * sizeof(int4), hipMemcpyHostToDevice, c->stream));
    HIPCHECK(c, hipStreamSynchronize(c->stream));
    *nitems_out = (int)items.size() / 2;
    return 0;
}

// One launch from step `it`: `nsteps` (1 or 2) steps of every shot of the
// prepared batch, reading stress/velocity set `rd` and writing set rd ^ 1.
int launch_step(sp_ctx *c, int nitems, int it, int snap_idx, int adjoint, int rd, int nsteps = 1,
                const ShotDesc *desc = nullptr, int snap_v = -1, int snap_u = -1)
{
    StepArgs a;
    a.snap_v = snap_v;
    a.snap_u = snap_u;
    a.shots = desc ? desc : c->d_desc;
    a.items = c->d_items;
    a.nitems = nitems;
    a.it = it;
    a.itp = c->itp;
    a.nt = c->nt;
    a.dt = c->dt;
    a.rd = rd;
    a.snap_idx = snap_idx;
    a.adjoint = adjoint;
    a.shift = strip_shift(c);
    a.wout = strip_width(c);
    dim3 grid((unsigned)((nitems + 3) / 4)), block(256);
    const bool ex = c->exact;
    if (nsteps == 2 && c->cur_fuse == 2) {
        a.items = c->d_items_w;
        a.nitems = c->n_items_w;
        a.wout = 116;
        grid = dim3((unsigned)((c->n_items_w + 3) / 4));
        if (c->physics == SP_SH) {
            if (ex) hipLaunchKernelGGL((k_step2w<SP_SH, true>), grid, block, 0, c->stream, a);
            else hipLaunchKernelGGL((k_step2w<SP_SH, false>), grid, block, 0, c->stream, a);
        } else {
            if (ex) hipLaunchKernelGGL((k_step2w<SP_PSV, true>), grid, block, 0, c->stream, a);
            else hipLaunchKernelGGL((k_step2w<SP_PSV, false>), grid, block, 0, c->stream, a);
        }
    } else if (nsteps == 2) {
        if (!c->cur_fuse) return fail(c, "internal: two-step launch without the two-step geometry");
        switch (c->physics) {
        case SP_SH:
            if (ex) hipLaunchKernelGGL((k_step2<SP_SH, true>), grid, block, 0, c->stream, a);
            else hipLaunchKernelGGL((k_step2<SP_SH, false>), grid, block, 0, c->stream, a);
            break;
        case SP_PSV:
            if (ex) hipLaunchKernelGGL((k_step2<SP_PSV, true>), grid, block, 0, c->stream, a);
            else hipLaunchKernelGGL((k_step2<SP_PSV, false>), grid, block, 0, c->stream, a);
            break;
        default:
            if (ex) hipLaunchKernelGGL((k_step2<SP_SPIN, true>), grid, block, 0, c->stream, a);
            else hipLaunchKernelGGL((k_step2<SP_SPIN, false>), grid, block, 0, c->stream, a);
            break;
        }
    } else if (c->cur_kernel == 1) {
        switch (c->physics) {
        case SP_SH:
            if (ex) hipLaunchKernelGGL((k_step3<SP_SH, true>), grid, block, 0, c->stream, a);
            else hipLaunchKernelGGL((k_step3<SP_SH, false>), grid, block, 0, c->stream, a);
            break;
        case SP_PSV:
            if (ex) hipLaunchKernelGGL((k_step3<SP_PSV, true>), grid, block, 0, c->stream, a);
            else hipLaunchKernelGGL((k_step3<SP_PSV, false>), grid, block, 0, c->stream, a);
            break;
        default:
            if (ex) hipLaunchKernelGGL((k_step3<SP_SPIN, true>), grid, block, 0, c->stream, a);
            else hipLaunchKernelGGL((k_step3<SP_SPIN, false>), grid, block, 0, c->stream, a);
            break;
        }
    } else {
        switch (c->physics) {
        case SP_SH:
            if (ex) hipLaunchKernelGGL((k_step<SP_SH, true>), grid, block, 0, c->stream, a);
            else hipLaunchKernelGGL((k_step<SP_SH, false>), grid, block, 0, c->stream, a);
            break;
        case SP_PSV:
            if (ex) hipLaunchKernelGGL((k_step<SP_PSV, true>), grid, block, 0, c->stream, a);
            else hipLaunchKernelGGL((k_step<SP_PSV, false>), grid, block, 0, c->stream, a);
            break;
        default:
            if (ex) hipLaunchKernelGGL((k_step<SP_SPIN, true>), grid, block, 0, c->stream, a);
            else hipLaunchKernelGGL((k_step<SP_SPIN, false>), grid, block, 0, c->stream, a);
            break;
        }
    }
    HIPCHECK(c, hipGetLastError());
    c->t_launches++;
    return 0;
}

double step_alg_bytes(const sp_ctx *c, const std::vector<int> &batch)
{
    // 8 B per state field per shot cell, 4 B per material field per grid cell
    const int nst = n_s(c->physics) + 2 * n_v(c->physics);
    const int nmat = c->physics == SP_SH ? 2 : (c->physics == SP_PSV ? 3 : 7);
    std::vector<int> seen(c->grids.size(), 0);
    double b = 0;
    for (int s : batch) {
        const Grid &g = c->grids[c->shots[s].grid];
        b += 8.0 * nst * (double)g.N;
        if (!seen[c->shots[s].grid]) {
            seen[c->shots[s].grid] = 1;
            b += 4.0 * nmat * (double)g.N;
        }
    }
    return b;
}

int clear_state(sp_ctx *c, const std::vector<int> &batch)
{
    for (int s : batch)
        if (fill(c, c->shots[s].state, (long long)n_slots(c->physics) * c->grids[c->shots[s].grid].Npad, 0.f)) return 2;
    return 0;
}

// Forward steps [0, nt) of a prepared batch without snapshots, as replays of
// one captured graph: a block of `block` steps (their step launches and
// receiver records, step numbers relative to a device counter) and a
// one-thread kernel that advances the counter; the steps past the last
// whole block run eagerly.  The block returns the stress/velocity set to
// where it started (an even number of launches), so every replay starts
// from the same set.  Needs a stream other than the null stream.
int run_forward_graph(sp_ctx *c, int nitems, int nb, int maxrec, int &set)
{
    const int nt = c->nt;
    int block = 64;
    while (block > 4 && block > nt) block /= 2;
    auto launches_in = [&](int steps) {
        int n = 0;
        for (int it = 0; it < steps;) {
            it += steps_next(c, it, steps, 0, 0);
            n++;
        }
        return n;
    };
    if (launches_in(block) % 2) block -= 2;  // keep the set parity of a block even
    const int nblk = nt / block;
    if (nblk < 2) return 1;  // not worth a graph: the caller runs eagerly
    if (!c->d_counter) HIPCHECK(c, hipMalloc(&c->d_counter, sizeof(int)));
    hipLaunchKernelGGL(k_fill, dim3(1), dim3(256), 0, c->stream, (float *)c->d_counter, 1LL, 0.f);  // 0.f = int 0
    HIPCHECK(c, hipGetLastError());
    const dim3 rgrid((unsigned)(((long long)nb * maxrec + 255) / 256));
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    HIPCHECK(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
    c->itp = c->d_counter;
    int s = set, rc = 0, nl = 0;
    for (int it = 0; it < block && !rc;) {
        const int ns = steps_next(c, it, block, 0, 0);
        rc = launch_step(c, nitems, it, -1, 0, s, ns);
        s ^= 1;
        nl++;
        if (!rc && ns == 1)
            hipLaunchKernelGGL(k_record, rgrid, dim3(256), 0, c->stream, c->d_desc, nb, maxrec, it, c->nt, c->physics,
                               c->itp);
        it += ns;
    }
    if (!rc) hipLaunchKernelGGL(k_advance, dim3(1), dim3(64), 0, c->stream, c->d_counter, block);
    c->itp = nullptr;
    const hipError_t ec = hipStreamEndCapture(c->stream, &graph);
    if (rc) {
        if (graph) (void)hipGraphDestroy(graph);
        return rc;
    }
    HIPCHECK(c, ec);
    const hipError_t ei = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    HIPCHECK(c, ei);
    for (int b = 0; b < nblk; b++) {
        const hipError_t el = hipGraphLaunch(exec, c->stream);
        if (el != hipSuccess) {
            (void)hipGraphExecDestroy(exec);
            HIPCHECK(c, el);
        }
    }
    HIPCHECK(c, hipGraphExecDestroy(exec));
    c->t_launches += (long long)nl * (nblk - 1);  // the capture counted one block
    // the remaining steps, eagerly
    for (int it = nblk * block; it < nt;) {
        const int ns = steps_next(c, it, nt, 0, 0);
        if (launch_step(c, nitems, it, -1, 0, s, ns)) return 2;
        s ^= 1;
        if (ns == 1)
            hipLaunchKernelGGL(k_record, rgrid, dim3(256), 0, c->stream, c->d_desc, nb, maxrec, it, c->nt, c->physics,
                               nullptr);
        HIPCHECK(c, hipGetLastError());
        it += ns;
    }
    set = s;
    return 0;
}

// forward run of a batch (solver.py:166-242); snapshots when sae > 0
int run_forward_batch(sp_ctx *c, const std::vector<int> &batch, int sae)
{
    int nitems = 0;
    if (prepare_batch(c, batch, &nitems)) return 2;
    if (clear_state(c, batch)) return 2;
    const int nsa = sae > 0 ? c->nt / sae : 0;
    const int nb = (int)batch.size();
    int maxrec = 1;
    long long maxN = 1;
    for (int s : batch) {
        maxrec = std::max(maxrec, c->shots[s].nrec);
        maxN = std::max(maxN, c->grids[c->shots[s].grid].N);
    }
    const dim3 rgrid((unsigned)(((long long)nb * maxrec + 255) / 256)), sgrid((unsigned)((maxN + 255) / 256), (unsigned)nb);
    const bool split = c->cur_kernel == 1;  // snapshots and records as kernels of their own
    int set = 0;
    if (c->graph && sae == 0 && split && c->stream) {
        const int rc = run_forward_graph(c, nitems, nb, maxrec, set);
        if (rc == 0) {
            c->cur_set = set;
            c->t_bytes += step_alg_bytes(c, batch) * c->nt;
            return 0;
        }
        if (rc != 1) return rc;
    }
    for (int it = 0; it < c->nt;) {
        const int ns = steps_next(c, it, c->nt, 0, 0);
        // snapshots after step it and (two-step launches) after step it+1,
        // stored in reverse order (solver.py:193-203)
        const int isa = (sae > 0 && (it + 1) % sae == 0) ? nsa - (it + 1) / sae : -1;
        const int isb = (ns == 2 && sae > 0 && (it + 2) % sae == 0) ? nsa - (it + 2) / sae : -1;
        // single steps of the buffer-addressed kernel: snapshot copies before
        // and after; two-step launches write their snapshots themselves
        const bool copy = split && ns == 1 && isa >= 0;
        if (copy)
            hipLaunchKernelGGL(k_snap, sgrid, dim3(256), 0, c->stream, c->d_desc, nb, maxN, isa, 0, 0, c->physics);
        if (launch_step(c, nitems, it, split ? -1 : isa, 0, set, ns, nullptr, ns == 2 ? isa : -1, isb)) return 2;
        set ^= 1;
        if (split) {
            if (copy)
                hipLaunchKernelGGL(k_snap, sgrid, dim3(256), 0, c->stream, c->d_desc, nb, maxN, isa, 1, set, c->physics);
            // the two-step kernel records both of its steps itself
            if (ns == 1)
                hipLaunchKernelGGL(k_record, rgrid, dim3(256), 0, c->stream, c->d_desc, nb, maxrec, it, c->nt,
                                   c->physics, c->itp);
            HIPCHECK(c, hipGetLastError());
        }
        it += ns;
    }
    c->cur_set = set;
    c->t_bytes += step_alg_bytes(c, batch) * c->nt;
    return 0;
}

int run_misfit_batch(sp_ctx *c, const std::vector<int> &batch, double *norms_out)
{
    int nitems = 0;
    if (prepare_batch(c, batch, &nitems)) return 2;
    const int ncm = n_mis(c->physics);
    std::vector<int4> trip;
    std::vector<const float *> data;
    for (int b = 0; b < (int)batch.size(); b++) {
        const Shot &s = c->shots[batch[b]];
        if (!s.data) return fail(c, "sp_set_data has not been called for shot " + std::to_string(batch[b]));
        data.push_back(s.data);
        for (int cc = 0; cc < ncm; cc++)
            for (int r = 0; r < s.nrec; r++) trip.push_back(make_int4(b, cc, r, 0));
    }
    if (trip.empty()) {
        for (int b = 0; b < (int)batch.size(); b++)
            for (int cc = 0; cc < ncm; cc++) norms_out[(size_t)batch[b] * ncm + cc] = 0.0;
        return 0;
    }
    // cosine taper of the first and last tenth of the trace (waveform.py:5-25)
    std::vector<double> taper(c->nt);
    {
        const double t_end = (double)(c->nt - 1) * c->dt;
        const double tw = t_end / 10.0, t_min = tw, t_max = t_end - tw;
        for (int it = 0; it < c->nt; it++) {
            const double tt = (double)it * c->dt;
            if (tt <= t_min) taper[it] = 0.5 + 0.5 * std::cos(M_PI * (t_min - tt) / tw);
            else if (tt >= t_max) taper[it] = 0.5 + 0.5 * std::cos(M_PI * (t_max - tt) / tw);
            else taper[it] = 1.0;
        }
    }
    DevTmp<int4> d_trip;
    DevTmp<const float *> d_data;
    DevTmp<double> d_part, d_taper;
    HIPCHECK(c, hipMalloc(&d_taper.p, taper.size() * sizeof(double)));
    HIPCHECK(c, hipMemcpyAsync(d_taper.p, taper.data(), taper.size() * sizeof(double), hipMemcpyHostToDevice, c->stream));
    HIPCHECK(c, hipMalloc(&d_trip.p, trip.size() * sizeof(int4)));
    HIPCHECK(c, hipMalloc(&d_data.p, data.size() * sizeof(float *)));
    HIPCHECK(c, hipMalloc(&d_part.p, trip.size() * sizeof(double)));
    HIPCHECK(c, hipMemcpyAsync(d_trip.p, trip.data(), trip.size() * sizeof(int4), hipMemcpyHostToDevice, c->stream));
    HIPCHECK(c, hipMemcpyAsync(d_data.p, data.data(), data.size() * sizeof(float *), hipMemcpyHostToDevice, c->stream));
    MisfitArgs a;
    a.shots = c->d_desc;
    a.triples = d_trip.p;
    a.data = d_data.p;
    a.partial = d_part.p;
    a.taper = d_taper.p;
    a.n = (int)trip.size();
    a.nt = c->nt;
    hipLaunchKernelGGL(k_misfit, dim3((unsigned)trip.size()), dim3(256), 0, c->stream, a);
    HIPCHECK(c, hipGetLastError());
    std::vector<double> part(trip.size());
    HIPCHECK(c, hipMemcpyAsync(part.data(), d_part.p, part.size() * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPCHECK(c, hipStreamSynchronize(c->stream));
    size_t p = 0;
    for (int b = 0; b < (int)batch.size(); b++) {
        const Shot &s = c->shots[batch[b]];
        for (int cc = 0; cc < ncm; cc++) {
            double ss = 0.0;
            for (int r = 0; r < s.nrec; r++) ss += part[p++];
            norms_out[(size_t)batch[b] * ncm + cc] = std::sqrt(ss);
        }
    }
    return 0;
}

// interaction launches of one adjoint step for the shots of a batch,
// grouped by grid (lists prepared by interact_lists)
struct InteractPlan {
    std::vector<std::vector<int>> per_grid;
    std::vector<int> off;
};

int interact_lists(sp_ctx *c, const std::vector<int> &batch, InteractPlan &pl)
{
    pl.per_grid.assign(c->grids.size(), {});
    for (int b = 0; b < (int)batch.size(); b++) pl.per_grid[c->shots[batch[b]].grid].push_back(b);
    std::vector<int> flat;
    pl.off.assign(c->grids.size() + 1, 0);
    for (size_t g = 0; g < pl.per_grid.size(); g++) {
        pl.off[g] = (int)flat.size();
        flat.insert(flat.end(), pl.per_grid[g].begin(), pl.per_grid[g].end());
    }
    pl.off[pl.per_grid.size()] = (int)flat.size();
    if (ensure(c, c->d_list, c->d_list_cap, std::max<size_t>(flat.size(), 1))) return 2;
    HIPCHECK(c, hipMemcpyAsync(c->d_list, flat.data(), flat.size() * sizeof(int), hipMemcpyHostToDevice, c->stream));
    return 0;
}

int launch_interact(sp_ctx *c, const InteractPlan &pl, int cur, int isae, double ndt)
{
    for (size_t g = 0; g < pl.per_grid.size(); g++) {
        if (pl.per_grid[g].empty()) continue;
        const Grid &G = c->grids[g];
        InteractArgs a;
        a.shots = c->d_desc;
        a.list = c->d_list + pl.off[g];
        a.nlist = (int)pl.per_grid[g].size();
        a.nx = G.nx;
        a.nz = G.nz;
        a.Npad = G.Npad;
        a.kern = G.kern;
        a.cur = cur;
        a.isae = isae;
        a.ndt = ndt;
        a.KX.c1 = 9.0 / (8.0 * (double)G.dx);
        a.KX.c2 = 1.0 / (24.0 * (double)G.dx);
        a.KZ.c1 = 9.0 / (8.0 * (double)G.dz);
        a.KZ.c2 = 1.0 / (24.0 * (double)G.dz);
        split_quotient(9.0, 8.0 * (double)G.dx, a.KX.k9h, a.KX.k9l);
        split_quotient(1.0, 24.0 * (double)G.dx, a.KX.k1h, a.KX.k1l);
        split_quotient(9.0, 8.0 * (double)G.dz, a.KZ.k9h, a.KZ.k9l);
        split_quotient(1.0, 24.0 * (double)G.dz, a.KZ.k1h, a.KZ.k1l);
        dim3 gr((unsigned)((G.N + 255) / 256)), bl(256);
        switch (c->physics) {
        case SP_SH: hipLaunchKernelGGL(k_interact<SP_SH>, gr, bl, 0, c->stream, a); break;
        case SP_PSV: hipLaunchKernelGGL(k_interact<SP_PSV>, gr, bl, 0, c->stream, a); break;
        default: hipLaunchKernelGGL(k_interact<SP_SPIN>, gr, bl, 0, c->stream, a); break;
        }
        HIPCHECK(c, hipGetLastError());
    }
    return 0;
}

int run_adjoint_batch(sp_ctx *c, const std::vector<int> &batch, int sae)
{
    int nitems = 0;
    if (prepare_batch(c, batch, &nitems)) return 2;
    if (clear_state(c, batch)) return 2;
    InteractPlan pl;
    if (interact_lists(c, batch, pl)) return 2;
    const double ndt = (double)sae * c->dt;
    int set = 0;
    for (int it = 0; it < c->nt;) {
        const int ns = steps_next(c, it, c->nt, sae, 0);
        if (launch_step(c, nitems, it, -1, 1, set, ns)) return 2;
        set ^= 1;
        const int last = it + ns - 1;
        if (last % sae == 0)
            if (launch_interact(c, pl, set, last / sae, ndt)) return 2;
        it += ns;
    }
    c->t_bytes += step_alg_bytes(c, batch) * c->nt;
    return 0;
}

// ---- checkpointed forward and adjoint ----------------------------------
// A checkpoint is the full state before step k*C: the stresses and
// velocities of the set holding it, and the displacement.
int ckpt_fields(int P) { return n_s(P) + 2 * n_v(P); }

// save: set `set` of the state -> checkpoint k; restore: checkpoint k -> set
// `set` of the second state block
int ckpt_copy(sp_ctx *c, Shot &S, int k, bool save, int set)
{
    const int P = c->physics;
    const Grid &G = c->grids[S.grid];
    const size_t fb = (size_t)G.Npad * sizeof(float);
    float *ck = S.ckpt + (size_t)k * ckpt_fields(P) * G.Npad;
    float *st = save ? S.state : S.state2;
    const size_t parts[3][3] = {{0, (size_t)slot_s(P, set, 0), (size_t)n_s(P)},
                                {(size_t)n_s(P), (size_t)slot_v(P, set, 0), (size_t)n_v(P)},
                                {(size_t)(n_s(P) + n_v(P)), (size_t)slot_u(P, 0), (size_t)n_v(P)}};
    for (auto &q : parts) {
        float *a = ck + q[0] * G.Npad, *b = st + q[1] * G.Npad;
        HIPCHECK(c, hipMemcpyAsync(save ? a : b, save ? b : a, q[2] * fb, hipMemcpyDeviceToDevice, c->stream));
    }
    return 0;
}

// forward run of a batch that keeps checkpoints every C steps (records too)
int run_forward_ckpt_batch(sp_ctx *c, const std::vector<int> &batch, int C)
{
    const int P = c->physics;
    const int K = (c->nt + C - 1) / C;
    for (int s : batch) {
        Shot &S = c->shots[s];
        const size_t need = (size_t)K * ckpt_fields(P) * c->grids[S.grid].Npad * sizeof(float);
        if (S.ckpt_bytes != need) {
            if (S.ckpt) HIPCHECK(c, hipFree(S.ckpt));
            S.ckpt = nullptr;
            S.ckpt_bytes = 0;
            HIPCHECK(c, hipMalloc(&S.ckpt, need));
            S.ckpt_bytes = need;
        }
    }
    int nitems = 0;
    if (prepare_batch(c, batch, &nitems)) return 2;
    if (clear_state(c, batch)) return 2;
    const int nb = (int)batch.size();
    int maxrec = 1;
    for (int s : batch) maxrec = std::max(maxrec, c->shots[s].nrec);
    const dim3 rgrid((unsigned)(((long long)nb * maxrec + 255) / 256));
    int set = 0;
    for (int it = 0; it < c->nt;) {
        if (it % C == 0)
            for (int s : batch)
                if (ckpt_copy(c, c->shots[s], it / C, true, set)) return 2;
        const int ns = steps_next(c, it, c->nt, 0, C);
        if (launch_step(c, nitems, it, -1, 0, set, ns)) return 2;
        set ^= 1;
        if (c->cur_kernel == 1 && ns == 1) {
            hipLaunchKernelGGL(k_record, rgrid, dim3(256), 0, c->stream, c->d_desc, nb, maxrec, it, c->nt, c->physics,
                               nullptr);
            HIPCHECK(c, hipGetLastError());
        }
        it += ns;
    }
    c->t_bytes += step_alg_bytes(c, batch) * c->nt;
    return 0;
}

// adjoint run of a batch; the forward snapshots it needs are recomputed
// segment by segment (last segment first) from the checkpoints, on each
// shot's second state block.
int run_adjoint_ckpt_batch(sp_ctx *c, const std::vector<int> &batch, int sae, int C)
{
    const int P = c->physics, nt = c->nt, nsa = nt / sae;
    const int cap = C / sae + 2;  // snapshots per segment, at most
    size_t need = 0;
    for (int s : batch) need += (size_t)cap * n_snap(P) * c->grids[c->shots[s].grid].Npad * sizeof(float);
    if (need > c->seg_pool_bytes) {
        if (c->seg_pool) HIPCHECK(c, hipFree(c->seg_pool));
        c->seg_pool = nullptr;
        c->seg_pool_bytes = 0;
        HIPCHECK(c, hipMalloc(&c->seg_pool, need));
        c->seg_pool_bytes = need;
    }
    size_t off = 0;
    for (auto &sh : c->shots) sh.snap = nullptr;
    for (int s : batch) {
        Shot &S = c->shots[s];
        S.snap = c->seg_pool + off / sizeof(float);
        off += (size_t)cap * n_snap(P) * c->grids[S.grid].Npad * sizeof(float);
        if (!S.state2) {
            const long long nf = (long long)n_slots(P) * c->grids[S.grid].Npad;
            HIPCHECK(c, hipMalloc(&S.state2, (size_t)nf * sizeof(float)));
            if (fill(c, S.state2, nf, 0.f)) return 2;
        }
    }
    int nitems = 0;
    if (prepare_batch(c, batch, &nitems)) return 2;
    // descriptors of the same shots on their second state block
    std::vector<ShotDesc> d2;
    for (int s : batch) {
        ShotDesc d = make_desc(c, c->shots[s]);
        d.state = c->shots[s].state2;
        d2.push_back(d);
    }
    if (ensure(c, c->d_desc2, c->d_desc2_cap, d2.size())) return 2;
    HIPCHECK(c, hipMemcpyAsync(c->d_desc2, d2.data(), d2.size() * sizeof(ShotDesc), hipMemcpyHostToDevice, c->stream));
    if (clear_state(c, batch)) return 2;
    InteractPlan pl;
    if (interact_lists(c, batch, pl)) return 2;
    const int nb = (int)batch.size();
    long long maxN = 1;
    for (int s : batch) maxN = std::max(maxN, c->grids[c->shots[s].grid].N);
    const dim3 sgrid((unsigned)((maxN + 255) / 256), (unsigned)nb);
    const double ndt = (double)sae * c->dt;
    int seg_loaded = -1, isa_base = 0, set = 0;
    for (int a = 0; a < nt;) {
        const int ns = steps_next(c, a, nt, sae, 0);
        if (launch_step(c, nitems, a, -1, 1, set, ns)) return 2;
        set ^= 1;
        const int last = a + ns - 1;
        a += ns;
        if (last % sae) continue;
        const int isae = last / sae, seg = (nt - last - 1) / C;
        if (seg != seg_loaded) {
            const int lo = seg * C, hi = std::min(nt, lo + C);
            isa_base = nsa - hi / sae;
            for (int s : batch)
                if (ckpt_copy(c, c->shots[s], seg, false, 0)) return 2;
            int set2 = 0;
            for (int it = lo; it < hi;) {
                const int ns2 = steps_next(c, it, hi, 0, 0);
                const int isa = (it + 1) % sae == 0 ? nsa - (it + 1) / sae - isa_base : -1;
                const int isb = (ns2 == 2 && (it + 2) % sae == 0) ? nsa - (it + 2) / sae - isa_base : -1;
                const bool copy = ns2 == 1 && isa >= 0;
                if (copy)
                    hipLaunchKernelGGL(k_snap, sgrid, dim3(256), 0, c->stream, c->d_desc2, nb, maxN, isa, 0, 0,
                                       c->physics);
                if (launch_step(c, nitems, it, -1, 0, set2, ns2, c->d_desc2, ns2 == 2 ? isa : -1, isb)) return 2;
                set2 ^= 1;
                if (copy)
                    hipLaunchKernelGGL(k_snap, sgrid, dim3(256), 0, c->stream, c->d_desc2, nb, maxN, isa, 1, set2,
                                       c->physics);
                it += ns2;
            }
            HIPCHECK(c, hipGetLastError());
            seg_loaded = seg;
        }
        if (launch_interact(c, pl, set, isae - isa_base, ndt)) return 2;
    }
    c->t_bytes += step_alg_bytes(c, batch) * nt;
    return 0;
}

// split all shots into batches whose snapshots fit the pool
int plan_batches(sp_ctx *c, int sae, std::vector<std::vector<int>> &out)
{
    out.clear();
    const int ns = (int)c->shots.size();
    if (sae <= 0) {
        std::vector<int> all(ns);
        std::iota(all.begin(), all.end(), 0);
        out.push_back(all);
        return 0;
    }
    size_t budget;
    if (c->snap_mb > 0) {
        budget = (size_t)c->snap_mb << 20;
    } else {
        size_t fr = 0, tot = 0;
        HIPCHECK(c, hipMemGetInfo(&fr, &tot));
        budget = (size_t)((double)(fr + c->snap_pool_bytes) * 0.6);
    }
    std::vector<int> cur;
    size_t used = 0;
    for (int s = 0; s < ns; s++) {
        const size_t b = snap_bytes(c, c->shots[s], sae);
        if (b > budget) return fail(c, "snapshots of one shot exceed the snapshot memory budget");
        if (used + b > budget) {
            out.push_back(cur);
            cur.clear();
            used = 0;
        }
        cur.push_back(s);
        used += b;
    }
    if (!cur.empty()) out.push_back(cur);
    return 0;
}

int assign_snapshots(sp_ctx *c, const std::vector<int> &batch, int sae)
{
    size_t need = 0;
    for (int s : batch) need += snap_bytes(c, c->shots[s], sae);
    if (need > c->snap_pool_bytes) {
        if (c->snap_pool) HIPCHECK(c, hipFree(c->snap_pool));
        c->snap_pool = nullptr;
        c->snap_pool_bytes = 0;
        HIPCHECK(c, hipMalloc(&c->snap_pool, need));
        c->snap_pool_bytes = need;
    }
    size_t off = 0;
    for (auto &sh : c->shots) sh.snap = nullptr;
    for (int s : batch) {
        c->shots[s].snap = c->snap_pool + off / sizeof(float);
        off += snap_bytes(c, c->shots[s], sae);
    }
    return 0;
}

int timing_begin(sp_ctx *c)
{
    c->fwd_next = -1;  // every run replaces the wavefields a stepwise forward run left
    c->t_launches = 0;
    c->t_bytes = 0;
    if (!c->ev0) {
        HIPCHECK(c, hipEventCreate(&c->ev0));
        HIPCHECK(c, hipEventCreate(&c->ev1));
        HIPCHECK(c, hipEventCreate(&c->ev2));
    }
    HIPCHECK(c, hipEventRecord(c->ev0, c->stream));
    return 0;
}

int timing_end(sp_ctx *c)
{
    HIPCHECK(c, hipEventRecord(c->ev1, c->stream));
    HIPCHECK(c, hipEventSynchronize(c->ev1));
    float ms = 0;
    HIPCHECK(c, hipEventElapsedTime(&ms, c->ev0, c->ev1));
    c->t_total_ms = ms;
    c->t_step_ms = ms;
    return 0;
}

// separable sponge factors of set_bound (fd2d.py:47-68)
void sponge_factors(int nx, int nz, int width, double alpha, int left, int right, int bottom, int top,
                    std::vector<float> &bx, std::vector<double> &bz)
{
    bx.assign(nx, 1.0f);
    bz.assign(2 * (size_t)nz, 1.0);
    for (int i = 0; i < nx; i++) {
        float b = 1.0f;
        if (left && i + 1 < width) {
            const double aw = alpha * (double)(width - i - 1);
            b = (float)((double)b * std::exp(-aw * aw));
        }
        if (right && i > nx - width) {
            const double aw = alpha * (double)(width + i - nx);
            b = (float)((double)b * std::exp(-aw * aw));
        }
        bx[i] = b;
    }
    for (int j = 0; j < nz; j++) {
        if (bottom && j > nz - width) {
            const double aw = alpha * (double)(width + j - nz);
            bz[j] = std::exp(-aw * aw);
        }
        if (top && j + 1 < width) {
            const double aw = alpha * (double)(width - j - 1);
            bz[nz + j] = std::exp(-aw * aw);
        }
    }
}

}  // namespace

// ====================================================================== ABI
extern "C" {

int sp_create(int device, sp_ctx **out)
{
    if (!out) return 3;
    *out = nullptr;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) return 4;
    if (device < 0 || device >= n) return 5;
    if (hipSetDevice(device) != hipSuccess) return 6;
    sp_ctx *c = new sp_ctx();
    c->device = device;
    if (hipStreamCreateWithFlags(&c->own, hipStreamNonBlocking) != hipSuccess) {
        delete c;
        return 7;
    }
    c->stream = c->own;
    *out = c;
    return 0;
}

static void free_grid(Grid &g)
{
    if (g.mat) (void)hipFree(g.mat);
    if (g.bx) (void)hipFree(g.bx);
    if (g.bz) (void)hipFree(g.bz);
    if (g.kern) (void)hipFree(g.kern);
    g = Grid();
}

static void free_shot(Shot &s)
{
    if (s.state2) (void)hipFree(s.state2);
    if (s.ckpt) (void)hipFree(s.ckpt);
    if (s.d_idx) (void)hipFree(s.d_idx);
    if (s.stf) (void)hipFree(s.stf);
    if (s.state) (void)hipFree(s.state);
    if (s.rec) (void)hipFree(s.rec);
    if (s.data) (void)hipFree(s.data);
    if (s.adstf) (void)hipFree(s.adstf);
    s = Shot();
}

int sp_destroy(sp_ctx *c)
{
    if (!c) return 0;
    (void)hipSetDevice(c->device);
    (void)hipStreamSynchronize(c->stream);
    for (auto &s : c->shots) free_shot(s);
    for (auto &g : c->grids) {
        free_grid(g);
    }
    if (c->snap_pool) (void)hipFree(c->snap_pool);
    if (c->d_desc) (void)hipFree(c->d_desc);
    if (c->d_desc2) (void)hipFree(c->d_desc2);
    if (c->d_counter) (void)hipFree(c->d_counter);
    if (c->seg_pool) (void)hipFree(c->seg_pool);
    if (c->d_items) (void)hipFree(c->d_items);
    if (c->d_items_w) (void)hipFree(c->d_items_w);
    if (c->d_list) (void)hipFree(c->d_list);
    if (c->ev0) (void)hipEventDestroy(c->ev0);
    if (c->ev1) (void)hipEventDestroy(c->ev1);
    if (c->ev2) (void)hipEventDestroy(c->ev2);
    if (c->own) (void)hipStreamDestroy(c->own);
    delete c;
    return 0;
}

const char *sp_last_error(const sp_ctx *c) { return c ? c->err.c_str() : "null context"; }

int sp_set_stream(sp_ctx *c, void *stream)
{
    if (!c) return 3;
    c->stream = (hipStream_t)stream;  // NULL is the device's null stream, as in HIP
    return 0;
}

int sp_sync(sp_ctx *c)
{
    if (!c) return 3;
    HIPCHECK(c, hipSetDevice(c->device));
    HIPCHECK(c, hipStreamSynchronize(c->stream));
    return 0;
}

int sp_set_option(sp_ctx *c, int key, long long v)
{
    if (!c) return 3;
    switch (key) {
    case SP_OPT_CHUNK:
        if (v < 0 || v > 4096) return fail(c, "chunk must be in [0, 4096]");
        c->chunk = (int)v;
        return 0;
    case SP_OPT_ORDER:
        if (v < 0 || v > 3) return fail(c, "order must be 0, 1, 2 or 3");
        c->order = (int)v;
        return 0;
    case SP_OPT_GRAPH: c->graph = v ? 1 : 0; return 0;  // forward runs without snapshots, own or set stream
    case SP_OPT_KERNEL: c->kernel = v ? 1 : 0; return 0;
    case SP_OPT_FUSE: c->fuse = v == 2 ? 2 : (v ? 1 : 0); return 0;
    case SP_OPT_EXACT: c->exact = v ? 1 : 0; return 0;
    case SP_OPT_STRIP: {
        const int sh = (int)(v >> 16), wo = (int)(v & 0xffff);
        if (sh < 3 || wo < 1 || sh + wo > 61) return fail(c, "strip geometry needs shift >= 3 and shift + width <= 61");
        c->shift = sh;
        c->wout = wo;
        return 0;
    }
    case SP_OPT_SNAP_MB:
        if (v < 0) return fail(c, "snapshot budget must be >= 0");
        c->snap_mb = v;
        return 0;
    default: return fail(c, "unknown option " + std::to_string(key));
    }
}

int sp_set_physics(sp_ctx *c, int physics, int nt, double dt)
{
    if (!c) return 3;
    if (physics != SP_SH && physics != SP_PSV && physics != SP_SPIN) return fail(c, "physics must be SP_SH, SP_PSV or SP_SPIN");
    if (nt <= 0) return fail(c, "nt must be positive");
    if (!(dt > 0)) return fail(c, "dt must be positive");
    if (!c->shots.empty() && (physics != c->physics || nt != c->nt))
        return fail(c, "clear the shots before changing physics or nt");
    c->physics = physics;
    c->nt = nt;
    c->dt = dt;
    return 0;
}

int sp_add_grid(sp_ctx *c, int nx, int nz, float dx, float dz, int param, const float *const *fields, int *grid_out)
{
    if (int e = check_ctx(c)) return e;
    if (nx < 1 || nz < 1) return fail(c, "grid must have nx >= 1 and nz >= 1");
    if (!(dx > 0) || !(dz > 0)) return fail(c, "dx and dz must be positive");
    if (!fields) return fail(c, "fields is null");
    if (param < 0 || param > 2) return fail(c, "param must be 0 (vp, vs, rho), 1 (spin list) or 2 (lambda, mu, rho)");
    for (int f = 0; f < (param == 1 ? 8 : 3); f++)
        if (!fields[f]) return fail(c, "missing model field");
    Grid g;
    auto undo = on_failure([&] { free_grid(g); });
    g.nx = nx;
    g.nz = nz;
    g.dx = dx;
    g.dz = dz;
    g.N = (long long)nx * nz;
    g.Npad = pad64(g.N);
    const size_t nb = (size_t)g.N * sizeof(float);
    HIPCHECK(c, hipMalloc(&g.mat, (size_t)N_MAT * g.Npad * sizeof(float)));
    HIPCHECK(c, hipMalloc(&g.kern, (size_t)SP_NKERN * g.Npad * sizeof(float)));
    HIPCHECK(c, hipMalloc(&g.bx, (size_t)nx * sizeof(float)));
    HIPCHECK(c, hipMalloc(&g.bz, 2 * (size_t)nz * sizeof(double)));
    Grid &G = g;
    if (fill(c, G.mat, (long long)N_MAT * G.Npad, 0.f)) return 2;
    if (fill(c, G.kern, (long long)SP_NKERN * G.Npad, 0.f)) return 2;
    float *m = G.mat;
    auto up = [&](int slot, const float *src) -> int {
        HIPCHECK(c, hipMemcpyAsync(m + (size_t)slot * G.Npad, src, nb, hipMemcpyHostToDevice, c->stream));
        return 0;
    };
    if (param == 0) {  // vp, vs, rho
        if (up(M_LAM, fields[0]) || up(M_MU, fields[1]) || up(M_RHO, fields[2])) return 2;
        hipLaunchKernelGGL(k_vps2lm, dim3((unsigned)((G.N + 255) / 256)), dim3(256), 0, c->stream,
                           m + (size_t)M_LAM * G.Npad, m + (size_t)M_MU * G.Npad, m + (size_t)M_RHO * G.Npad, G.N);
        HIPCHECK(c, hipGetLastError());
    } else if (param == 1) {  // lambda, mu, nu, j, lambda_c, mu_c, nu_c, rho
        const int slots[8] = {M_LAM, M_MU, M_NU, M_J, SP_F_LAMC, SP_F_MUC, SP_F_NUC, M_RHO};
        for (int f = 0; f < 8; f++)
            if (up(slots[f], fields[f])) return 2;
        hipLaunchKernelGGL(k_gamma, dim3((unsigned)((G.N + 255) / 256)), dim3(256), 0, c->stream,
                           m + (size_t)M_GAM * G.Npad, m + (size_t)SP_F_MUC * G.Npad, m + (size_t)SP_F_NUC * G.Npad, G.N);
        HIPCHECK(c, hipGetLastError());
    } else {  // 2: lambda, mu, rho
        if (up(M_LAM, fields[0]) || up(M_MU, fields[1]) || up(M_RHO, fields[2])) return 2;
    }
    // no sponge until sp_set_sponge: all factors 1
    std::vector<float> bx(nx, 1.0f);
    std::vector<double> bz(2 * (size_t)nz, 1.0);
    HIPCHECK(c, hipMemcpyAsync(G.bx, bx.data(), bx.size() * sizeof(float), hipMemcpyHostToDevice, c->stream));
    HIPCHECK(c, hipMemcpyAsync(G.bz, bz.data(), bz.size() * sizeof(double), hipMemcpyHostToDevice, c->stream));
    HIPCHECK(c, hipStreamSynchronize(c->stream));
    undo.armed = false;
    c->grids.push_back(g);
    if (grid_out) *grid_out = (int)c->grids.size() - 1;
    return 0;
}

int sp_set_sponge(sp_ctx *c, int grid, int width, double alpha, int left, int right, int bottom, int top)
{
    if (int e = check_ctx(c)) return e;
    if (grid < 0 || grid >= (int)c->grids.size()) return fail(c, "bad grid id");
    Grid &G = c->grids[grid];
    std::vector<float> bx;
    std::vector<double> bz;
    sponge_factors(G.nx, G.nz, width, alpha, left, right, bottom, top, bx, bz);
    HIPCHECK(c, hipMemcpyAsync(G.bx, bx.data(), bx.size() * sizeof(float), hipMemcpyHostToDevice, c->stream));
    HIPCHECK(c, hipMemcpyAsync(G.bz, bz.data(), bz.size() * sizeof(double), hipMemcpyHostToDevice, c->stream));
    HIPCHECK(c, hipStreamSynchronize(c->stream));
    return 0;
}

int sp_set_field(sp_ctx *c, int grid, int field, const float *src, int on_dev)
{
    if (int e = check_ctx(c)) return e;
    if (grid < 0 || grid >= (int)c->grids.size()) return fail(c, "bad grid id");
    if (field < 0 || field > SP_F_NUC) return fail(c, "bad field id");
    Grid &G = c->grids[grid];
    float *m = G.mat;
    if (dcopy(c, m + (size_t)field * G.Npad, src, (size_t)G.N * sizeof(float), 1, on_dev)) return 2;
    if (field == SP_F_MUC || field == SP_F_NUC) {
        hipLaunchKernelGGL(k_gamma, dim3((unsigned)((G.N + 255) / 256)), dim3(256), 0, c->stream,
                           m + (size_t)M_GAM * G.Npad, m + (size_t)SP_F_MUC * G.Npad, m + (size_t)SP_F_NUC * G.Npad, G.N);
        HIPCHECK(c, hipGetLastError());
    }
    c->snaps_valid = 0;
    c->ckpt_valid = 0;
    c->fwd_next = -1;
    return 0;
}

int sp_get_field(sp_ctx *c, int grid, int field, float *dst, int on_dev)
{
    if (int e = check_ctx(c)) return e;
    if (grid < 0 || grid >= (int)c->grids.size()) return fail(c, "bad grid id");
    if (field < 0 || field > SP_F_NUC) return fail(c, "bad field id");
    Grid &G = c->grids[grid];
    return dcopy(c, dst, G.mat + (size_t)field * G.Npad, (size_t)G.N * sizeof(float), on_dev, 1);
}

int sp_add_shot(sp_ctx *c, int grid, int nsrc, const int *src_k, const float *stf, int nrec, const int *rec_k, int *shot_out)
{
    if (int e = check_ctx(c)) return e;
    c->fwd_next = -1;
    if (grid < 0 || grid >= (int)c->grids.size()) return fail(c, "bad grid id");
    if (nsrc < 0 || nrec < 0) return fail(c, "negative source or receiver count");
    if ((nsrc && (!src_k || !stf)) || (nrec && !rec_k)) return fail(c, "null source/receiver arrays");
    const Grid &G = c->grids[grid];
    for (int s = 0; s < nsrc; s++)
        if (src_k[s] < 0 || src_k[s] >= G.N) return fail(c, "source index outside the grid");
    for (int r = 0; r < nrec; r++)
        if (rec_k[r] < 0 || rec_k[r] >= G.N) return fail(c, "receiver index outside the grid");
    Shot s;
    auto undo = on_failure([&] { free_shot(s); });
    s.grid = grid;
    s.nsrc = nsrc;
    s.nrec = nrec;
    // index lists sorted by grid index, ties in input order
    std::vector<int> so(nsrc), ro(nrec);
    std::iota(so.begin(), so.end(), 0);
    std::iota(ro.begin(), ro.end(), 0);
    std::stable_sort(so.begin(), so.end(), [&](int a, int b) { return src_k[a] < src_k[b]; });
    std::stable_sort(ro.begin(), ro.end(), [&](int a, int b) { return rec_k[a] < rec_k[b]; });
    std::vector<int> idx(2 * (size_t)nsrc + 2 * (size_t)nrec + 1, 0);
    for (int q = 0; q < nsrc; q++) {
        idx[q] = src_k[so[q]];
        idx[nsrc + q] = so[q];
    }
    for (int q = 0; q < nrec; q++) {
        idx[2 * nsrc + q] = rec_k[ro[q]];
        idx[2 * nsrc + nrec + q] = ro[q];
    }
    s.src_sorted.assign(idx.begin(), idx.begin() + nsrc);
    s.rec_sorted.assign(idx.begin() + 2 * nsrc, idx.begin() + 2 * nsrc + nrec);
    const int P = c->physics;
    HIPCHECK(c, hipMalloc(&s.d_idx, idx.size() * sizeof(int)));
    HIPCHECK(c, hipMemcpyAsync(s.d_idx, idx.data(), idx.size() * sizeof(int), hipMemcpyHostToDevice, c->stream));
    const size_t nstf = (size_t)n_inj(P) * nsrc * c->nt;
    HIPCHECK(c, hipMalloc(&s.stf, std::max<size_t>(nstf, 1) * sizeof(float)));
    if (nstf) HIPCHECK(c, hipMemcpyAsync(s.stf, stf, nstf * sizeof(float), hipMemcpyHostToDevice, c->stream));
    HIPCHECK(c, hipMalloc(&s.state, (size_t)n_slots(P) * G.Npad * sizeof(float)));
    const long long nr = std::max(rec_floats(c, s), 1LL), na = std::max(adj_floats(c, s), 1LL);
    HIPCHECK(c, hipMalloc(&s.rec, (size_t)nr * sizeof(float)));
    HIPCHECK(c, hipMalloc(&s.adstf, (size_t)na * sizeof(float)));
    if (fill(c, s.rec, nr, 0.f) || fill(c, s.adstf, na, 0.f)) return 2;
    HIPCHECK(c, hipStreamSynchronize(c->stream));
    undo.armed = false;
    c->shots.push_back(s);
    c->snaps_valid = 0;
    if (shot_out) *shot_out = (int)c->shots.size() - 1;
    return 0;
}

int sp_clear_shots(sp_ctx *c)
{
    if (!c) return 3;
    HIPCHECK(c, hipSetDevice(c->device));
    HIPCHECK(c, hipStreamSynchronize(c->stream));
    for (auto &s : c->shots) free_shot(s);
    c->shots.clear();
    c->snaps_valid = 0;
    c->ckpt_valid = 0;
    c->fwd_next = -1;
    return 0;
}

int sp_num_components(const sp_ctx *c) { return c ? n_rec(c->physics) : 0; }

int sp_forward(sp_ctx *c, int sae)
{
    if (int e = check_ctx(c)) return e;
    if (sae < 0) return fail(c, "adjoint_interval must be >= 0");
    if (sae > 0 && c->nt % sae) return fail(c, "nt must be a multiple of adjoint_interval");
    if (c->shots.empty()) return fail(c, "no shots");
    std::vector<std::vector<int>> batches;
    if (plan_batches(c, sae, batches)) return 2;
    if (batches.size() > 1) return fail(c, "snapshots of all shots exceed the snapshot memory budget; use sp_gradient");
    c->snaps_valid = 0;
    if (timing_begin(c)) return 2;
    if (sae > 0 && assign_snapshots(c, batches[0], sae)) return 2;
    if (run_forward_batch(c, batches[0], sae)) return 2;
    if (timing_end(c)) return 2;
    if (sae > 0) {
        c->snaps_valid = 1;
        c->snaps_sae = sae;
    }
    return 0;
}

int sp_forward_steps(sp_ctx *c, int it0, int it1)
{
    if (int e = check_ctx(c)) return e;
    if (c->shots.empty()) return fail(c, "no shots");
    if (it0 < 0 || it1 <= it0 || it1 > c->nt) return fail(c, "steps must satisfy 0 <= it0 < it1 <= nt");
    if (it0 > 0 && it0 != c->fwd_next) return fail(c, "sp_forward_steps must continue from step " + std::to_string(c->fwd_next));
    std::vector<int> all(c->shots.size());
    std::iota(all.begin(), all.end(), 0);
    const int nb = (int)all.size();
    int maxrec = 1;
    for (int s : all) maxrec = std::max(maxrec, c->shots[s].nrec);
    const dim3 rgrid((unsigned)(((long long)nb * maxrec + 255) / 256));
    int nitems = 0;
    if (prepare_batch(c, all, &nitems)) return 2;
    if (it0 == 0) {
        c->snaps_valid = 0;
        c->ckpt_valid = 0;
    c->fwd_next = -1;
        if (clear_state(c, all)) return 2;
        c->cur_set = 0;
    }
    if (timing_begin(c)) return 2;
    int set = c->cur_set;
    for (int it = it0; it < it1;) {
        const int ns = steps_next(c, it, it1, 0, 0);
        if (launch_step(c, nitems, it, -1, 0, set, ns)) return 2;
        set ^= 1;
        if (c->cur_kernel == 1 && ns == 1) {
            hipLaunchKernelGGL(k_record, rgrid, dim3(256), 0, c->stream, c->d_desc, nb, maxrec, it, c->nt, c->physics,
                               nullptr);
            HIPCHECK(c, hipGetLastError());
        }
        it += ns;
    }
    c->cur_set = set;
    c->fwd_next = it1 < c->nt ? it1 : -1;
    c->t_bytes += step_alg_bytes(c, all) * (it1 - it0);
    return timing_end(c);
}

int sp_get_wavefield(sp_ctx *c, int shot, int kind, float *dst, int on_dev)
{
    if (int e = check_ctx(c)) return e;
    if (shot < 0 || shot >= (int)c->shots.size()) return fail(c, "bad shot id");
    const int P = c->physics;
    const Shot &S = c->shots[shot];
    const Grid &G = c->grids[S.grid];
    int first, n;
    switch (kind) {
    case SP_WF_V: first = slot_v(P, c->cur_set, 0), n = n_v(P); break;
    case SP_WF_U: first = slot_u(P, 0), n = n_v(P); break;
    case SP_WF_S: first = slot_s(P, c->cur_set, 0), n = n_s(P); break;
    default: return fail(c, "bad wavefield kind");
    }
    for (int f = 0; f < n; f++)
        if (dcopy(c, dst + (size_t)f * G.N, S.state + (size_t)(first + f) * G.Npad, (size_t)G.N * sizeof(float), on_dev, 1))
            return 2;
    return 0;
}

int sp_get_records(sp_ctx *c, int shot, float *dst, int on_dev)
{
    if (int e = check_ctx(c)) return e;
    if (shot < 0 || shot >= (int)c->shots.size()) return fail(c, "bad shot id");
    const Shot &s = c->shots[shot];
    return dcopy(c, dst, s.rec, (size_t)rec_floats(c, s) * sizeof(float), on_dev, 1);
}

int sp_set_data(sp_ctx *c, int shot, const float *src, int on_dev)
{
    if (int e = check_ctx(c)) return e;
    if (shot < 0 || shot >= (int)c->shots.size()) return fail(c, "bad shot id");
    Shot &s = c->shots[shot];
    const size_t n = (size_t)std::max(rec_floats(c, s), 1LL);
    if (!s.data) HIPCHECK(c, hipMalloc(&s.data, n * sizeof(float)));
    return dcopy(c, s.data, src, (size_t)rec_floats(c, s) * sizeof(float), 1, on_dev);
}

int sp_misfit(sp_ctx *c, double *norms_out)
{
    if (int e = check_ctx(c)) return e;
    if (!norms_out) return fail(c, "norms_out is null");
    std::vector<int> all(c->shots.size());
    std::iota(all.begin(), all.end(), 0);
    return run_misfit_batch(c, all, norms_out);
}

int sp_adjoint(sp_ctx *c, int sae)
{
    if (int e = check_ctx(c)) return e;
    if (sae <= 0 || c->nt % sae) return fail(c, "nt must be a positive multiple of adjoint_interval");
    if (!c->snaps_valid || c->snaps_sae != sae) return fail(c, "no snapshots for this adjoint_interval: run sp_forward(ctx, sae) first");
    for (auto &g : c->grids)
        if (fill(c, g.kern, (long long)SP_NKERN * g.Npad, 0.f)) return 2;
    std::vector<int> all(c->shots.size());
    std::iota(all.begin(), all.end(), 0);
    if (timing_begin(c)) return 2;
    if (run_adjoint_batch(c, all, sae)) return 2;
    return timing_end(c);
}

int sp_gradient(sp_ctx *c, int sae, double *norms_out)
{
    if (int e = check_ctx(c)) return e;
    if (sae <= 0 || c->nt % sae) return fail(c, "nt must be a positive multiple of adjoint_interval");
    if (!norms_out) return fail(c, "norms_out is null");
    if (c->shots.empty()) return fail(c, "no shots");
    std::vector<std::vector<int>> batches;
    if (plan_batches(c, sae, batches)) return 2;
    for (auto &g : c->grids)
        if (fill(c, g.kern, (long long)SP_NKERN * g.Npad, 0.f)) return 2;
    c->snaps_valid = 0;
    if (timing_begin(c)) return 2;
    double bytes = 0;
    long long launches = 0;
    for (auto &b : batches) {
        if (assign_snapshots(c, b, sae)) return 2;
        if (run_forward_batch(c, b, sae)) return 2;
        if (run_misfit_batch(c, b, norms_out)) return 2;
        if (run_adjoint_batch(c, b, sae)) return 2;
        bytes += c->t_bytes;
        launches += c->t_launches;
        c->t_bytes = 0;
        c->t_launches = 0;
    }
    if (timing_end(c)) return 2;
    c->t_bytes = bytes;
    c->t_launches = launches;
    if (batches.size() == 1) {
        c->snaps_valid = 1;
        c->snaps_sae = sae;
    }
    return 0;
}

int sp_forward_ckpt(sp_ctx *c, int C)
{
    if (int e = check_ctx(c)) return e;
    if (c->shots.empty()) return fail(c, "no shots");
    if (C < 2 || C % 2) return fail(c, "checkpoint interval must be even and >= 2");
    std::vector<int> all(c->shots.size());
    std::iota(all.begin(), all.end(), 0);
    c->ckpt_valid = 0;
    c->fwd_next = -1;
    c->snaps_valid = 0;
    if (timing_begin(c)) return 2;
    if (run_forward_ckpt_batch(c, all, C)) return 2;
    if (timing_end(c)) return 2;
    c->ckpt_C = C;
    c->ckpt_valid = 1;
    return 0;
}

int sp_set_adjoint_source(sp_ctx *c, int shot, const float *src, int on_dev)
{
    if (int e = check_ctx(c)) return e;
    if (shot < 0 || shot >= (int)c->shots.size()) return fail(c, "bad shot id");
    Shot &s = c->shots[shot];
    return dcopy(c, s.adstf, src, (size_t)adj_floats(c, s) * sizeof(float), 1, on_dev);
}

int sp_adjoint_ckpt(sp_ctx *c, int sae)
{
    if (int e = check_ctx(c)) return e;
    if (!c->ckpt_valid) return fail(c, "no checkpoints: run sp_forward_ckpt first");
    if (sae <= 0 || c->nt % sae) return fail(c, "nt must be a positive multiple of adjoint_interval");
    for (auto &g : c->grids)
        if (fill(c, g.kern, (long long)SP_NKERN * g.Npad, 0.f)) return 2;
    std::vector<int> all(c->shots.size());
    std::iota(all.begin(), all.end(), 0);
    if (timing_begin(c)) return 2;
    if (run_adjoint_ckpt_batch(c, all, sae, c->ckpt_C)) return 2;
    c->snaps_valid = 0;
    return timing_end(c);
}

int sp_set_model_vps(sp_ctx *c, int grid, const float *vp, const float *vs, const float *rho, int on_dev)
{
    if (int e = check_ctx(c)) return e;
    if (grid < 0 || grid >= (int)c->grids.size()) return fail(c, "bad grid id");
    Grid &G = c->grids[grid];
    float *m = G.mat;
    const size_t nb = (size_t)G.N * sizeof(float);
    if (dcopy(c, m + (size_t)M_LAM * G.Npad, vp, nb, 1, on_dev) || dcopy(c, m + (size_t)M_MU * G.Npad, vs, nb, 1, on_dev) ||
        dcopy(c, m + (size_t)M_RHO * G.Npad, rho, nb, 1, on_dev))
        return 2;
    hipLaunchKernelGGL(k_vps2lm, dim3((unsigned)((G.N + 255) / 256)), dim3(256), 0, c->stream,
                       m + (size_t)M_LAM * G.Npad, m + (size_t)M_MU * G.Npad, m + (size_t)M_RHO * G.Npad, G.N);
    HIPCHECK(c, hipGetLastError());
    c->snaps_valid = 0;
    c->ckpt_valid = 0;
    c->fwd_next = -1;
    return 0;
}

int sp_get_kernel(sp_ctx *c, int grid, int which, float *dst, int on_dev)
{
    if (int e = check_ctx(c)) return e;
    if (grid < 0 || grid >= (int)c->grids.size()) return fail(c, "bad grid id");
    if (which < 0 || which >= SP_NKERN) return fail(c, "bad kernel id");
    Grid &G = c->grids[grid];
    return dcopy(c, dst, G.kern + (size_t)which * G.Npad, (size_t)G.N * sizeof(float), on_dev, 1);
}

int sp_kernels_vps(sp_ctx *c, int grid)
{
    if (int e = check_ctx(c)) return e;
    if (grid < 0 || grid >= (int)c->grids.size()) return fail(c, "bad grid id");
    Grid &G = c->grids[grid];
    hipLaunchKernelGGL(k_kern_vps, dim3((unsigned)((G.N + 255) / 256)), dim3(256), 0, c->stream, G.kern, G.mat, G.N, G.Npad);
    HIPCHECK(c, hipGetLastError());
    return 0;
}

static int smooth_dev(sp_ctx *c, float *field, int nx, int nz, int sigma)
{
    if (sigma <= 0) return fail(c, "smooth sigma must be positive");
    const int nm = std::max(nx, nz);
    const double s = (double)sigma;
    const double cc = 1.0 / (std::sqrt(2.0 * M_PI) * s);
    std::vector<double> g(nm), sx(nx), sz(nz);
    for (int d = 0; d < nm; d++) g[d] = cc * std::exp(-(double)((long long)d * d) / (2.0 * s * s));
    for (int i = 0; i < nx; i++) {
        double a = 0.0;
        for (int n = 0; n < nx; n++) a += g[std::abs(i - n)];
        sx[i] = a;
    }
    for (int j = 0; j < nz; j++) {
        double a = 0.0;
        for (int n = 0; n < nz; n++) a += g[std::abs(j - n)];
        sz[j] = a;
    }
    DevTmp<double> dg;
    DevTmp<float> tmp;
    const long long N = (long long)nx * nz;
    HIPCHECK(c, hipMalloc(&dg.p, (size_t)(nm + nx + nz) * sizeof(double)));
    HIPCHECK(c, hipMalloc(&tmp.p, (size_t)N * sizeof(float)));
    HIPCHECK(c, hipMemcpyAsync(dg.p, g.data(), nm * sizeof(double), hipMemcpyHostToDevice, c->stream));
    HIPCHECK(c, hipMemcpyAsync(dg.p + nm, sx.data(), nx * sizeof(double), hipMemcpyHostToDevice, c->stream));
    HIPCHECK(c, hipMemcpyAsync(dg.p + nm + nx, sz.data(), nz * sizeof(double), hipMemcpyHostToDevice, c->stream));
    const dim3 gr((unsigned)((N + 255) / 256)), bl(256);
    hipLaunchKernelGGL(k_smooth_x, gr, bl, 0, c->stream, field, tmp.p, dg.p, nx, nz);
    hipLaunchKernelGGL(k_smooth_z, gr, bl, 0, c->stream, field, tmp.p, dg.p, dg.p + nm, dg.p + nm + nx, nx, nz);
    HIPCHECK(c, hipGetLastError());
    HIPCHECK(c, hipStreamSynchronize(c->stream));
    return 0;
}

int sp_smooth_kernel(sp_ctx *c, int grid, int which, int sigma)
{
    if (int e = check_ctx(c)) return e;
    if (grid < 0 || grid >= (int)c->grids.size()) return fail(c, "bad grid id");
    if (which < 0 || which >= SP_NKERN) return fail(c, "bad kernel id");
    Grid &G = c->grids[grid];
    return smooth_dev(c, G.kern + (size_t)which * G.Npad, G.nx, G.nz, sigma);
}

int sp_smooth(sp_ctx *c, float *field, int nx, int nz, int sigma, int on_dev)
{
    if (!c) return 3;
    HIPCHECK(c, hipSetDevice(c->device));
    if (nx < 1 || nz < 1 || !field) return fail(c, "bad field");
    if (on_dev) return smooth_dev(c, field, nx, nz, sigma);
    float *d = nullptr;
    const size_t nb = (size_t)nx * nz * sizeof(float);
    HIPCHECK(c, hipMalloc(&d, nb));
    if (dcopy(c, d, field, nb, 1, 0) || smooth_dev(c, d, nx, nz, sigma) || dcopy(c, field, d, nb, 0, 1)) {
        (void)hipFree(d);
        return 2;
    }
    HIPCHECK(c, hipFree(d));
    return 0;
}

int sp_autotune(sp_ctx *c, int steps, int *chunk_out, int *order_out)
{
    if (int e = check_ctx(c)) return e;
    if (c->shots.empty()) return fail(c, "no shots");
    if (steps < 1) return fail(c, "steps must be >= 1");
    steps = std::min(steps, c->nt);
    std::vector<int> all(c->shots.size());
    std::iota(all.begin(), all.end(), 0);
    // results are kept per process for the batch's shape: physics, kernel
    // choice, arithmetic mode and every shot's grid
    std::string key = std::to_string(c->physics) + "/" + std::to_string(c->kernel) + "/" + std::to_string(c->fuse) +
                      "/" + std::to_string(c->exact) + "/" + std::to_string(steps);
    for (const Shot &sh : c->shots)
        key += ":" + std::to_string(c->grids[sh.grid].nx) + "x" + std::to_string(c->grids[sh.grid].nz);
    static std::mutex mu;
    static std::map<std::string, std::array<int, 3>> cache;
    {
        std::lock_guard<std::mutex> lock(mu);
        auto f = cache.find(key);
        if (f != cache.end()) {
            c->chunk = f->second[0];
            c->order = f->second[1];
            c->fuse = f->second[2];
            if (chunk_out) *chunk_out = c->chunk;
            if (order_out) *order_out = c->order;
            return 0;
        }
    }
    const int chunks[] = {4, 8, 16, 32, 64, 128};
    // one or two steps per launch, and two steps with one or two cells per
    // lane where the batch allows it (all three give the same results)
    std::vector<int> fuses = {c->fuse};
    if (c->fuse && fuse_ok(c, all)) {
        fuses.push_back(0);
        const int user = c->fuse;
        c->fuse = 2;
        const bool wide = fuse_ok(c, all) == 2;
        c->fuse = user;
        if (wide) fuses.push_back(user == 2 ? 1 : 2);
    }
    float best = 1e30f;
    int bc = 16, bo = 0, bf = c->fuse;
    if (timing_begin(c)) return 2;
    for (int fu : fuses) {
        c->fuse = fu;
        for (int order = 0; order < 4; order++) {
            for (int ch : chunks) {
                c->chunk = ch;
                c->order = order;
                int nitems = 0;
                if (prepare_batch(c, all, &nitems) || clear_state(c, all)) return 2;
                if (launch_step(c, nitems, 0, -1, 0, 0)) return 2;
                HIPCHECK(c, hipEventRecord(c->ev0, c->stream));
                int set = 1;
                for (int it = 0; it < steps;) {
                    const int ns = steps_next(c, it, steps, 0, 0);
                    if (launch_step(c, nitems, it, -1, 0, set, ns)) return 2;
                    set ^= 1;
                    it += ns;
                }
                HIPCHECK(c, hipEventRecord(c->ev1, c->stream));
                HIPCHECK(c, hipEventSynchronize(c->ev1));
                float ms = 0;
                HIPCHECK(c, hipEventElapsedTime(&ms, c->ev0, c->ev1));
                if (ms < best) {
                    best = ms;
                    bc = ch;
                    bo = order;
                    bf = fu;
                }
            }
        }
    }
    c->chunk = bc;
    c->order = bo;
    c->fuse = bf;
    c->snaps_valid = 0;
    {
        std::lock_guard<std::mutex> lock(mu);
        cache[key] = {bc, bo, bf};
    }
    if (chunk_out) *chunk_out = bc;
    if (order_out) *order_out = bo;
    return 0;
}

int sp_get_option(const sp_ctx *c, int key, long long *value)
{
    if (!c || !value) return 3;
    switch (key) {
    case SP_OPT_CHUNK: *value = c->chunk; return 0;
    case SP_OPT_ORDER: *value = c->order; return 0;
    case SP_OPT_GRAPH: *value = c->graph; return 0;
    case SP_OPT_SNAP_MB: *value = c->snap_mb; return 0;
    case SP_OPT_KERNEL: *value = c->kernel; return 0;
    case SP_OPT_STRIP: *value = ((long long)c->shift << 16) | c->wout; return 0;
    case SP_OPT_FUSE: *value = c->fuse; return 0;
    case SP_OPT_EXACT: *value = c->exact; return 0;
    default: return 1;
    }
}

int sp_last_timing(const sp_ctx *c, long long *launches, double *step_ms, double *total_ms, double *bytes)
{
    if (!c) return 3;
    if (launches) *launches = c->t_launches;
    if (step_ms) *step_ms = c->t_step_ms;
    if (total_ms) *total_ms = c->t_total_ms;
    if (bytes) *bytes = c->t_bytes;
    return 0;
}

}  // extern "C"
