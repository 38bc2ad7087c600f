// seispie_hip -- k_step2<P, EX, W, SQ, M> and k_step2g, two time steps per
// launch with W = 1 or 2 z cells per lane.
// Part of seispie_hip.hip: included there, inside its anonymous namespace,
// after <hip/hip_runtime.h> and the headers before this one.
#pragma once

// ------------------------------------------------- two steps per launch
// k_step2 advances a strip by two time steps (it and it+1) in one row march,
// keeping the intermediate state in registers: the state is read once and
// written once per two steps.  The intermediate stresses and velocities of a
// chunk's halo rows are recomputed by each chunk that needs them; the strip
// keeps 6 cells of halo on each side, and the chunk reads 6 rows before and
// 5 after its own.  Every value is computed with the same operations, in the
// same order, as two launches of k_step3, so results are identical bit for
// bit.  The result goes to the other stress/velocity set (rd ^ 1), like one
// step, so no wavefront overwrites rows another one still reads.
//
// W cells per lane: lane L holds z cells j0 + W L .. j0 + W L + W - 1 of
// every field.
//   W = 1: a wavefront covers 64 cells, lanes [6, 58) are output.
//   W = 2: 128 cells, 116 of them output, and every load and
//     store moves both cells at once (8 bytes per lane), so a row issues
//     half the memory instructions per cell.  Needs even nz (a lane's two
//     cells are then both inside or both outside the grid and every access
//     is 8-byte aligned); SH and P-SV (spin would need more than 256 VGPRs).
//
// Iteration m of the march (rows relative to m):
//   A  v(n+1)  at row m-1 from s(n) rows m-3..m   (source at step it)
//   B  s(n+1)  at row m-3 from v(n+1) rows m-4..m-1
//   C  v(n+2)  at row m-4 from s(n+1) rows m-6..m-3 (source at step it+1);
//      u(n+1), u(n+2) and the receiver records of both steps at row m-4
//   D  s(n+2)  at row m-6 from v(n+2) rows m-7..m-4
// Windows of four rows are indexed by row & 3 (loop unrolled by four).

// W consecutive floats of a lane at its offset: one 4- or 8-byte access
typedef unsigned v2u_t __attribute__((ext_vector_type(2)));
template <int W>
__device__ __forceinline__ void bldw(float (&x)[W], rsrc_t r, unsigned voff, int soff)
{
    if constexpr (W == 1) {
        x[0] = bld(r, voff, soff);
    } else {
        const v2u_t w = __builtin_amdgcn_raw_buffer_load_b64(r, (int)voff, soff, 0);
        x[0] = __uint_as_float(w.x);
        x[1] = __uint_as_float(w.y);
    }
}
template <int W>
__device__ __forceinline__ void bstw(rsrc_t r, unsigned voff, int soff, const float (&x)[W])
{
    if constexpr (W == 1) {
        bst(r, voff, soff, x[0]);
    } else {
        v2u_t w;
        w.x = __float_as_uint(x[0]);
        w.y = __float_as_uint(x[1]);
        __builtin_amdgcn_raw_buffer_store_b64(w, r, (int)voff, soff, 0);
    }
}

template <int P, int W>
struct Row2 {
    float s[n_s(P)][W];     // s(n), row m
    float v[n_v(P)][W];     // v(n), row m-1
    float u[n_v(P)][W];     // u(n), row m-4
    float rhoA[W], jA[W];   // row m-1 (j: micro-inertia, spin only)
    float lamB[W], muB[W];  // row m-3
    float nuB[W], gamB[W];  // row m-3 (spin only: nu, mu_c + nu_c)
    // rows m-4 and m-6 of the same materials are the values loaded three
    // iterations earlier; the march keeps them in register windows
};

// The rows are wave-uniform, so each load's row offset goes into the scalar
// offset and the per-lane offset vj = 4 jc is the same for every load (no
// vector address arithmetic per row).  Every row is clamped into the grid,
// so every load stays inside its buffer.
template <int P, int W>
__device__ __forceinline__ void load2(Row2<P, W> &p, rsrc_t RS, rsrc_t RM, int NP4, int rd, int mp, int nx, int nz,
                                      unsigned vj, int i0, int i1)
{
    constexpr int NS = n_s(P), NV = n_v(P);
    // rows clamped to the rows the chunk reads, i0-6 .. i1+5 (and to the
    // grid): one pair of bounds for every field, so the march keeps two
    // scalars live for them instead of eight; the few prologue and epilogue
    // rows this lets in are loaded but feed no output
    const int lo = max(i0 - 6, 0), hi = min(i1 + 5, nx - 1);
    const int ms = min(max(mp, lo), hi), ma = min(max(mp - 1, lo), hi);
    const int mb = min(max(mp - 3, lo), hi), mu_ = min(max(mp - 4, lo), hi);
    const int rb = nz * 4;
    const int os = ms * rb, oa = ma * rb, ob = mb * rb, ou = mu_ * rb;
#pragma unroll
    for (int f = 0; f < NS; f++) bldw<W>(p.s[f], RS, vj, slot_s(P, rd, f) * NP4 + os);
#pragma unroll
    for (int f = 0; f < NV; f++) {
        bldw<W>(p.v[f], RS, vj, slot_v(P, rd, f) * NP4 + oa);
        bldw<W>(p.u[f], RS, vj, slot_u(P, f) * NP4 + ou);
    }
    bldw<W>(p.rhoA, RM, vj, M_RHO * NP4 + oa);
    bldw<W>(p.muB, RM, vj, M_MU * NP4 + ob);
#pragma unroll
    for (int c = 0; c < W; c++) {
        p.lamB[c] = 0.f;
        p.jA[c] = 1.f;
        p.nuB[c] = p.gamB[c] = 0.f;
    }
    if constexpr (P != SP_SH) bldw<W>(p.lamB, RM, vj, M_LAM * NP4 + ob);
    if constexpr (P == SP_SPIN) {
        bldw<W>(p.jA, RM, vj, M_J * NP4 + oa);
        bldw<W>(p.nuB, RM, vj, M_NU * NP4 + ob);
        bldw<W>(p.gamB, RM, vj, M_GAM * NP4 + ob);
    }
}

// z neighbours of a field held W cells per lane.  For D- (at z: z-2, z-1,
// z+1) and D+ (at z: z-1, z+1, z+2); with two cells, the lane's own pair
// plus one rotated cell from each side covers both.
template <int W>
__device__ __forceinline__ void zminus(float (&m2)[W], float (&m1)[W], float (&p1)[W], const float (&x)[W])
{
    if constexpr (W == 1) {
        m1[0] = lane_m1(x[0]);
        m2[0] = lane_m1(m1[0]);
        p1[0] = lane_p1(x[0]);
    } else {
        const float a_m = lane_m1(x[0]), b_m = lane_m1(x[1]), a_p = lane_p1(x[0]);
        m2[0] = a_m;  m1[0] = b_m;  p1[0] = x[1];
        m2[1] = b_m;  m1[1] = x[0]; p1[1] = a_p;
    }
}
template <int W>
__device__ __forceinline__ void zplus(float (&m1)[W], float (&p1)[W], float (&p2)[W], const float (&x)[W])
{
    if constexpr (W == 1) {
        m1[0] = lane_m1(x[0]);
        p1[0] = lane_p1(x[0]);
        p2[0] = lane_p1(p1[0]);
    } else {
        const float b_m = lane_m1(x[1]), a_p = lane_p1(x[0]), b_p = lane_p1(x[1]);
        m1[0] = b_m;  p1[0] = x[1]; p2[0] = a_p;
        m1[1] = x[0]; p1[1] = a_p;  p2[1] = b_p;
    }
}

// ------------------------------------------------- column groups
// In a column group the wavefronts of one block march the same rows of one
// shot and together cover a span of its z column, wavefront g holding cells
// j0 + 64 W g .. j0 + 64 W (g + 1) - 1.  A z neighbour across a wavefront
// boundary is the neighbouring wavefront's edge cell, passed through LDS
// instead of recomputed in halo lanes, so only the span's ends inside the
// column keep halo cells.
//
// Cells next to a wavefront's edges: l0, l1 are the two cells before its
// first cell (the previous wavefront's last two), r0, r1 the two after its
// last cell (the next wavefront's first two).
struct Nbr {
    float l0, l1, r0, r1;
};

// the fields whose z neighbours the stencils take: stresses in div_sy /
// div_sxz / div_sxyz_c (szy; szz, sxz; and syy_c, syz_c for spin) and
// velocities in div_vy / div_vxz (vy; vx, vz; and vy_c) (solver.py:54-75,
// 112-126)
constexpr int n_zs(int p) { return p == SP_SH ? 1 : (p == SP_PSV ? 2 : 4); }
constexpr int n_zv(int p) { return p == SP_SH ? 1 : (p == SP_PSV ? 2 : 3); }
// exchanged rows per iteration: s(n), v(n+1), s(n+1), v(n+2)
constexpr int n_xch(int p) { return 2 * n_zs(p) + 2 * n_zv(p); }
// wavefronts per block: at most 8, so a kernel may use 256 VGPRs (with 16,
// 128 VGPRs make the one-cell P-SV march spill)
constexpr int grp_max(int, int) { return 8; }

// D- neighbours (z-2, z-1, z+1) as zminus, the lanes at the wavefront's
// edges (lf: first lane, ll: last lane) taking them from nb.  Every
// rotation runs first, with all 64 lanes active, and the edge values are
// selected afterwards: a rotation inside a conditional would run with the
// edge lane switched off, and its neighbour would read a disabled lane.
template <int W>
__device__ __forceinline__ void zminus_g(float (&m2)[W], float (&m1)[W], float (&p1)[W], const float (&x)[W],
                                         const Nbr &nb, bool lf, bool ll)
{
    if constexpr (W == 1) {
        const float rm = lane_m1(x[0]), rp = lane_p1(x[0]);
        m1[0] = lf ? nb.l1 : rm;
        const float rmm = lane_m1(m1[0]);
        m2[0] = lf ? nb.l0 : rmm;
        p1[0] = ll ? nb.r0 : rp;
    } else {
        const float ra = lane_m1(x[0]), rb = lane_m1(x[1]), rp = lane_p1(x[0]);
        const float a_m = lf ? nb.l0 : ra, b_m = lf ? nb.l1 : rb;
        const float a_p = ll ? nb.r0 : rp;
        m2[0] = a_m;  m1[0] = b_m;  p1[0] = x[1];
        m2[1] = b_m;  m1[1] = x[0]; p1[1] = a_p;
    }
}
// D+ neighbours (z-1, z+1, z+2) as zplus
template <int W>
__device__ __forceinline__ void zplus_g(float (&m1)[W], float (&p1)[W], float (&p2)[W], const float (&x)[W],
                                        const Nbr &nb, bool lf, bool ll)
{
    if constexpr (W == 1) {
        const float rm = lane_m1(x[0]), rp = lane_p1(x[0]);
        m1[0] = lf ? nb.l1 : rm;
        p1[0] = ll ? nb.r0 : rp;
        const float rpp = lane_p1(p1[0]);
        p2[0] = ll ? nb.r1 : rpp;
    } else {
        const float rb = lane_m1(x[1]), ra = lane_p1(x[0]), rc = lane_p1(x[1]);
        const float b_m = lf ? nb.l1 : rb;
        const float a_p = ll ? nb.r0 : ra, b_p = ll ? nb.r1 : rc;
        m1[0] = b_m;  p1[0] = x[1]; p2[0] = a_p;
        m1[1] = x[0]; p1[1] = a_p;  p2[1] = b_p;
    }
}
// strips (G false) or column groups (G true; nb[k] for the k-th exchanged field)
template <int W, bool G>
__device__ __forceinline__ void zm(float (&m2)[W], float (&m1)[W], float (&p1)[W], const float (&x)[W],
                                   const Nbr *nb, int k, bool lf, bool ll)
{
    if constexpr (G) zminus_g<W>(m2, m1, p1, x, nb[k], lf, ll);
    else zminus<W>(m2, m1, p1, x);
}
template <int W, bool G>
__device__ __forceinline__ void zp(float (&m1)[W], float (&p1)[W], float (&p2)[W], const float (&x)[W],
                                   const Nbr *nb, int k, bool lf, bool ll)
{
    if constexpr (G) zplus_g<W>(m1, p1, p2, x, nb[k], lf, ll);
    else zplus<W>(m1, p1, p2, x);
}

// stress divergence of one velocity row (div_sy / div_sxz, solver.py:54-66)
// from the stress window rows r-2..r+1 (sa..sd); column groups take the
// z neighbours of sc's fields 1, 2, 4, 5 across wavefront edges from nb
template <int P, bool EX, int W, bool G = false>
__device__ __forceinline__ void vel_row(float (&dsv)[n_v(P)][W], const float (&sa)[n_s(P)][W],
                                        const float (&sb)[n_s(P)][W], const float (&sc)[n_s(P)][W],
                                        const float (&sd)[n_s(P)][W], bool xm_ok, const bool (&zm_ok)[W],
                                        const DCo &KX, const DCo &KZ, const Nbr *nb = nullptr, bool lf = false,
                                        bool ll = false)
{
    if constexpr (P == SP_SH) {
        float zm2[W], zm1[W], zp1[W];
        zm<W, G>(zm2, zm1, zp1, sc[1], nb, 0, lf, ll);
#pragma unroll
        for (int c = 0; c < W; c++) {
            const double ddx = xm_ok ? dmn<EX>(sa[0][c], sb[0][c], sc[0][c], sd[0][c], KX) : 0.0;
            const double ddz = lsel(zm_ok[c], dmn<EX>(zm2[c], zm1[c], sc[1][c], zp1[c], KZ));
            dsv[0][c] = (float)(ddx + ddz);
        }
    } else {
        float xzm2[W], xzm1[W], xzp1[W], zzm2[W], zzm1[W], zzp1[W];
        zm<W, G>(xzm2, xzm1, xzp1, sc[2], nb, 1, lf, ll);
        zm<W, G>(zzm2, zzm1, zzp1, sc[1], nb, 0, lf, ll);
#pragma unroll
        for (int c = 0; c < W; c++) {
            const double dxx = xm_ok ? dmn<EX>(sa[0][c], sb[0][c], sc[0][c], sd[0][c], KX) : 0.0;
            const double dxxz = xm_ok ? dmn<EX>(sa[2][c], sb[2][c], sc[2][c], sd[2][c], KX) : 0.0;
            const double dzxz = lsel(zm_ok[c], dmn<EX>(xzm2[c], xzm1[c], sc[2][c], xzp1[c], KZ));
            const double dzzz = lsel(zm_ok[c], dmn<EX>(zzm2[c], zzm1[c], sc[1][c], zzp1[c], KZ));
            dsv[0][c] = (float)(dxx + dzxz);
            dsv[1][c] = (float)(dxxz + dzzz);
        }
        if constexpr (P == SP_SPIN) {
            // div_sy on the couple stresses, then div_sxyz_c (solver.py:69-75, 217-218)
            float yzm2[W], yzm1[W], yzp1[W], yym2[W], yym1[W], yyp1[W];
            zm<W, G>(yzm2, yzm1, yzp1, sc[5], nb, 3, lf, ll);
            zm<W, G>(yym2, yym1, yyp1, sc[4], nb, 2, lf, ll);
#pragma unroll
            for (int c = 0; c < W; c++) {
                const double dxyx = xm_ok ? dmn<EX>(sa[3][c], sb[3][c], sc[3][c], sd[3][c], KX) : 0.0;
                const double dzyz = lsel(zm_ok[c], dmn<EX>(yzm2[c], yzm1[c], sc[5][c], yzp1[c], KZ));
                const float dsy = (float)(dxyx + dzyz);
                const double dzyy = lsel(zm_ok[c], dmn<EX>(yym2[c], yym1[c], sc[4][c], yyp1[c], KZ));
                const double dxyy = xm_ok ? dmn<EX>(sa[4][c], sb[4][c], sc[4][c], sd[4][c], KX) : 0.0;
                dsv[0][c] = (float)((double)dsv[0][c] + dzyy);
                dsv[1][c] = (float)((double)dsv[1][c] - dxyy);
                dsv[2][c] = (float)((double)dsy + 2.0 * (double)sc[4][c]);
            }
        }
    }
}

// velocity update (add_vy / add_vxz, solver.py:96-117) of one cell from the
// divergence plus sources, the old v, 1/rho and the sponge factors
// `zint`: every lane of the strip lies outside the top/bottom sponges; with
// bx = 1 the sponge factor is exactly 1 and its arithmetic is skipped
// (the same float64 result, bit for bit).
// irr, irj: RN(1/rho), RN(1/j) for the exact arithmetic (the caller
// computes them, or reuses them from the pair's first step)
template <int P, bool EX>
__device__ __forceinline__ void vel_finish(float (&vn)[n_v(P)], const float (&dsv)[n_v(P)], const float (&vo)[n_v(P)],
                                           float rho, float jj, double irr, double irj, float bxr, double bzb,
                                           double bzt, bool zint, bool rin, bool jin, double dt)
{
    if constexpr (EX) {
        double b = 1.0;
        if (!(zint && bxr == 1.f)) {
            const float b1 = (float)((double)bxr * bzb);
            b = (double)(float)((double)b1 * bzt);
        }
        const double rr = (double)rho;
#pragma unroll
        for (int f = 0; f < n_v(P); f++) {
            double rf = rr, ir = irr;
            if constexpr (P == SP_SPIN) {
                if (f == 2) {  // vy_c divides by the micro-inertia j (solver.py:226)
                    rf = (double)jj;
                    ir = irj;
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
// velocity window rows q-1..q+2 (wa..wd) and the old stresses; column
// groups take the z neighbours of wb's fields 0, 1, 2 from nb
template <int P, bool EX, int W, bool G = false>
__device__ __forceinline__ void str_row(float (&sn)[n_s(P)][W], const float (&wa)[n_v(P)][W],
                                        const float (&wb)[n_v(P)][W], const float (&wc)[n_v(P)][W],
                                        const float (&wd)[n_v(P)][W], const float (&so)[n_s(P)][W],
                                        const float (&lam)[W], const float (&mu)[W], const float (&nu)[W],
                                        const float (&gam)[W], bool xp_ok, const bool (&zp_ok)[W], const DCo &KX,
                                        const DCo &KZ, double dt, const Nbr *nb = nullptr, bool lf = false,
                                        bool ll = false)
{
    if constexpr (P == SP_SH) {
        float zm1[W], zp1[W], zp2[W];
        zp<W, G>(zm1, zp1, zp2, wb[0], nb, 0, lf, ll);
#pragma unroll
        for (int c = 0; c < W; c++) {
            const float dvdx = xp_ok ? (float)dpl<EX>(wa[0][c], wb[0][c], wc[0][c], wd[0][c], KX) : 0.f;
            const float dvdz = lsel(zp_ok[c], (float)dpl<EX>(zm1[c], wb[0][c], zp1[c], zp2[c], KZ));
            const double dtmu = dt * (double)mu[c];
            sn[0][c] = upd_sadd<EX>(so[0][c], dtmu, (double)dvdx);
            sn[1][c] = upd_sadd<EX>(so[1][c], dtmu, (double)dvdz);
        }
    } else {
        float xm1[W], xp1[W], xp2[W], zm1[W], zp1[W], zp2[W];
        zp<W, G>(xm1, xp1, xp2, wb[0], nb, 0, lf, ll);
        zp<W, G>(zm1, zp1, zp2, wb[1], nb, 1, lf, ll);
        float cm1[W], cp1[W], cp2[W];
        if constexpr (P == SP_SPIN) zp<W, G>(cm1, cp1, cp2, wb[2], nb, 2, lf, ll);
#pragma unroll
        for (int c = 0; c < W; c++) {
            const float dvxdx = xp_ok ? (float)dpl<EX>(wa[0][c], wb[0][c], wc[0][c], wd[0][c], KX) : 0.f;
            const float dvzdx = xp_ok ? (float)dpl<EX>(wa[1][c], wb[1][c], wc[1][c], wd[1][c], KX) : 0.f;
            const float dvxdz = lsel(zp_ok[c], (float)dpl<EX>(xm1[c], wb[0][c], xp1[c], xp2[c], KZ));
            const float dvzdz = lsel(zp_ok[c], (float)dpl<EX>(zm1[c], wb[1][c], zp1[c], zp2[c], KZ));
            const double l2m = (double)lam[c] + 2.0 * (double)mu[c];
            sn[0][c] = upd_snorm<EX>(so[0][c], dt, l2m, dvxdx, mul_rn(lam[c], dvzdz));
            sn[1][c] = upd_snorm<EX>(so[1][c], dt, l2m, dvzdz, mul_rn(lam[c], dvxdx));
            sn[2][c] = upd_sadd<EX>(so[2][c], dt, (double)mul_rn(mu[c], add_rn(dvxdz, dvzdx)));
            if constexpr (P == SP_SPIN) {
                // div_vy on vy_c + add_sy_c (solver.py:135-141, 227-228)
                const float dcdx = xp_ok ? (float)dpl<EX>(wa[2][c], wb[2][c], wc[2][c], wd[2][c], KX) : 0.f;
                const float dcdz = lsel(zp_ok[c], (float)dpl<EX>(cm1[c], wb[2][c], cp1[c], cp2[c], KZ));
                const double dtg = dt * (double)gam[c];
                sn[4][c] = upd_srot<EX>(so[4][c], 2.0 * dt, nu[c], wb[2][c], sub_rn(dvzdx, dvxdz));
                sn[3][c] = upd_sadd<EX>(so[3][c], dtg, (double)dcdx);
                sn[5][c] = upd_sadd<EX>(so[5][c], dtg, (double)dcdz);
            }
        }
    }
}

// sources (forward) or adjoint sources (adjoint) of velocity row r at step
// `it`: float32 adds in list order (stf_dsy / stf_dsxz, solver.py:78-93);
// j is the lane's first cell; pi and src_row: the next list entry and its
// x row
template <int P, int W>
__device__ __forceinline__ void inject(float (&dsv)[n_v(P)][W], int r, int j, int nz, int &pi, int &src_row,
                                       cdesc_t *dp, int adjoint, int nt, int it)
{
    if (r != src_row) return;
    const InjLists L = inj_lists(fresh(dp), adjoint);
    const int kend = (r + 1) * nz, kr = r * nz + j;
    while (pi < L.n && L.k[pi] < kend) {
        const int kk = L.k[pi];
        const int row = L.row[pi];
        const float s0 = L.val[(long long)row * nt + it];
        float s1 = 0.f;
        if constexpr (P != SP_SH) s1 = L.val[(L.nrows + row) * nt + it];
#pragma unroll
        for (int c = 0; c < W; c++) {
            if (kk == kr + c) {
                dsv[0][c] = add_rn(dsv[0][c], s0);
                if constexpr (P != SP_SH) dsv[1][c] = add_rn(dsv[1][c], s1);
            }
        }
        ++pi;
    }
    src_row = pi < L.n ? L.xi[pi] : 0x7fffffff;
}
// the same adds from an injection queue in LDS (adjoint launches, with an
// adjoint source at every receiver)
template <int P, int W>
__device__ __forceinline__ void inject_q(float (&dsv)[n_v(P)][W], int r, int j, int nz, InjQ &q)
{
    const int kr = r * nz + j;
    injq_row<P>(q, r, [&](int kk, float s0, float s1) {
#pragma unroll
        for (int c = 0; c < W; c++) {
            if (kk == kr + c) {
                dsv[0][c] = add_rn(dsv[0][c], s0);
                if constexpr (P != SP_SH) dsv[1][c] = add_rn(dsv[1][c], s1);
            }
        }
    });
}

// The march of one work item.  Strips (G false): four independent
// wavefronts per block, each a strip with its own halo lanes.  Column
// groups (G true): a block is one work item, its wavefronts cover 64 W
// cells each of the shot's z column from cell item.y on and exchange edge
// cells through xl, the LDS ring [4 iterations][grp_max][n_xch][4 cells]:
// the rows a wavefront computes in iteration m go to slot m & 3, and the
// stencils of iteration m read the neighbours' rows of iterations m-1 and
// m-2, after one barrier per iteration.  Groups that do not reach the
// column's ends keep 6 halo cells there, as strips do; wavefronts past the
// column's last cell take part in the barriers only.
//
// M, the launch's mode: STEP_SNAP forward launches that may store adjoint
// snapshots; STEP_PLAIN launches that store none (most forward launches,
// and adjoint launches that read the injection lists); STEP_QUEUE adjoint
// launches with the adjoint sources queued in LDS.  Code for snapshots the
// launch does not take costs column groups and spin 1-2.4% per launch (two
// branches per row); SH and P-SV strips run their STEP_SNAP instances for
// launches without snapshots too (0.5% faster there).
template <int P, bool EX, int W, bool G, bool SQ, int M>
__device__ __forceinline__ void step2_body(StepArgs a, float *xl, int *iq)
{
    static_assert(W == 1 || (W == 2 && P != SP_SPIN), "two cells per lane: SH and P-SV");
    constexpr int NS = n_s(P), NV = n_v(P);
    constexpr bool IQ = M == STEP_QUEUE, SNAP = M == STEP_SNAP;
    constexpr int GM = grp_max(P, W), NX = n_xch(P), NZS = n_zs(P), NZV = n_zv(P);
    // strips skip the arithmetic of prologue rows that feed no output row of
    // the chunk (3 of the first 9 iterations' s(n+1) rows, 6 v(n+2) rows, 9
    // s(n+2) rows); in column groups the branches cost more than they save
    constexpr bool SKIP = !G;
    // x-edge tests (a row's D- or D+ in x is zero within two rows of the
    // grid's edges) as selects of unconditionally computed differences in
    // column groups, two-cell strips and spin, where the scalar branches they
    // otherwise become end scheduling regions (0.2-0.9% faster); one-cell
    // P-SV strips measured 1.3% faster with the branches
    auto xsel = [](bool b) { return (G || W == 2 || P == SP_SPIN) ? lane_bool(b) : b; };
    const int nb = gridDim.x, bb = blockIdx.x;
    const int q8 = nb >> 3, r8 = nb & 7, bx8 = bb & 7, by8 = bb >> 3;
    const int lb = bx8 < r8 ? bx8 * (q8 + 1) + by8 : r8 * (q8 + 1) + (bx8 - r8) * q8 + by8;
    const int gw = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);  // wavefront of the block
    const int wave = __builtin_amdgcn_readfirstlane(G ? lb : lb * 4 + gw);
    if (wave >= a.nitems) return;
    const int L = threadIdx.x & 63;
    const int4 item = a.items[2 * wave];
    if (item.x < 0) return;
    const int4 item2 = a.items[2 * wave + 1];
    cdesc_t &d = ((cdesc_t *)a.shots)[item.x];
    const int itx = a.itp ? a.it + *(cint_p)a.itp : a.it;
    // snapshot slots of the pair's two steps (solver.py:193-203): given, or
    // in a replayed graph looked up for the step
    int snap_v = a.snap_v, snap_u = a.snap_u;
    if (a.snap_tab) {
        snap_v = ((cint_p)a.snap_tab)[itx];
        snap_u = ((cint_p)a.snap_tab)[itx + 1];
    }
    const int nx = d.nx, nz = d.nz, i0 = item.z, i1 = item.w;
    // the march starts at m0 = i0 - 3 and ends at mend (one barrier per
    // iteration in column groups)
    const int m0 = i0 - 3, mend = i1 + 5;
    if constexpr (G) {
        if (item.y + 64 * W * gw >= nz) {
            for (int m = m0; m <= mend; m++) __syncthreads();
            return;
        }
    }
    // the lane's first cell (column groups: the group starts at cell item.y)
    const int j = G ? item.y + W * (64 * gw + L) : item.y - a.shift + W * L;
    const int jc = min(max(j, 0), nz - W);   // the first cell it loads
    const bool jin = j >= 0 && j < nz;       // every cell of the lane (nz is a multiple of W)
    // column groups: a group's output cells are [glo, ghi), 6 halo cells
    // kept at each of its edges inside the column; gact wavefronts hold
    // cells of the column
    int glo = 0, ghi = nz, gact = 1;
    if constexpr (G) {
        const int span = 64 * W * (int)(blockDim.x >> 6);
        glo = item.y == 0 ? 0 : item.y + 6;
        ghi = item.y + span >= nz ? nz : item.y + span - 6;
        gact = min((int)(blockDim.x >> 6), (nz - item.y + 64 * W - 1) / (64 * W));
    }
    const bool outl = G ? jin && j >= glo && j < ghi : jin && W * L >= a.shift && W * L < a.shift + a.wout;
    // column groups: edge lanes, their cells' positions in a ring entry, and
    // the neighbouring wavefronts; the group's first and last wavefronts
    // read their own entries there, values that reach only halo cells or
    // masked stencils
    const bool lf = L == 0, ll = L == 63;
    const bool xw = W == 1 ? (L < 2 || L >= 62) : (lf || ll);
    const int xpos = W == 1 ? (L < 2 ? L : L - 60) : (lf ? 0 : 2);
    const int gp = max(gw - 1, 0), gn = min(gw + 1, gact - 1);
    bool zm_ok[W], zp_ok[W];
    double bzb[W], bzt[W];
    bool bz_one = true;
#pragma unroll
    for (int c = 0; c < W; c++) {
        const int jj = j + c, jb = min(max(jj, 0), nz - 1);
        zm_ok[c] = jj >= 2 && jj < nz - 2;
        zp_ok[c] = jj >= 1 && jj < nz - 2;
        bzb[c] = d.bzb[jb];
        bzt[c] = d.bzt[jb];
        bz_one = bz_one && bzb[c] == 1.0 && bzt[c] == 1.0;
    }
    const int NP4 = (int)(d.Npad * 4);
    const double dt = a.dt;
    // square cells (SQ: dx = dz in every grid of the batch): one set of
    // difference coefficients serves both axes, eight scalar registers fewer
    // for the march (1-2% faster, the same values)
    const DCo KX{d.c1x, d.c2x, d.k9hx, d.k9lx, d.k1hx, d.k1lx};
    const DCo KZ = SQ ? KX : DCo{d.c1z, d.c2z, d.k9hz, d.k9lz, d.k1hz, d.k1lz};
    const int rd = a.rd, wr = a.rd ^ 1;
    const bool zint = __all(bz_one);
    const rsrc_t RS = make_rsrc(d.state, (long long)n_slots(P) * NP4);
    const rsrc_t RM = make_rsrc(d.mat, (long long)N_MAT * NP4);
    const cflt_p bxp = (cflt_p)d.bx;
    const unsigned vj = (unsigned)jc * 4u;            // per-lane load offset
    const unsigned vo = outl ? (unsigned)j * 4u : OOB;  // per-lane store offset (OOB: not written)

    // injection pointers: step it from row i0-4, step it+1 from row i0-1;
    // IQ (launches in which no item reaches more than IQ_CAP entries):
    // queues of the entries in LDS
    int p1, p2, row1, row2;
    InjQ q1, q2;
    {
        const InjLists L = inj_lists(&d, a.adjoint);
        p1 = __builtin_amdgcn_readfirstlane(a.adjoint ? item2.y : item2.x);
        p2 = p1;
        while (p2 < L.n && L.k[p2] < (i0 - 1) * nz) ++p2;
        if constexpr (IQ) {
            injq_fill<P>(q1, iq + 2 * IQ_WORDS * gw, &d, a.adjoint, a.nt, itx, p1, item2.w);
            injq_fill<P>(q2, iq + (2 * gw + 1) * IQ_WORDS, &d, a.adjoint, a.nt, itx + 1, p2, item2.w);
        } else {
            row1 = p1 < L.n ? L.xi[p1] : 0x7fffffff;
            row2 = p2 < L.n ? L.xi[p2] : 0x7fffffff;
        }
    }
    // receivers recorded by this chunk (forward only): rows i0..i1-1; the
    // next two receivers' cells and record rows wait in scalar registers
    // (a row's test uses values loaded rows earlier)
    int pr = __builtin_amdgcn_readfirstlane(item2.z), nk1, nr1, nk2, nr2;
    {
        const cint_p rk = (cint_p)d.rec_k, rrow = (cint_p)d.rec_row;
        const int nrec = a.adjoint ? 0 : d.nrec;
        nk1 = pr < nrec ? rk[pr] : 0x7fffffff;
        nr1 = pr < nrec ? rrow[pr] : 0;
        nk2 = pr + 1 < nrec ? rk[pr + 1] : 0x7fffffff;
        nr2 = pr + 1 < nrec ? rrow[pr + 1] : 0;
    }

    float S0[4][NS][W], W1[4][NV][W], S1[4][NS][W], W2[4][NV][W];
    // materials of rows m-1 (rho, j) and m-3 (lam, mu, nu, gam) by
    // iteration: three iterations later they are rows m-4 and m-6
    float RW[4][W], JW[4][W], LW[4][W], MW[4][W], NW[4][W], GW[4][W];
    // spin, exact arithmetic: RN(1/rho), RN(1/j) of the same rows
    double IRW[4][W], IJW[4][W];
#pragma unroll
    for (int u = 0; u < 4; u++) {
#pragma unroll
        for (int c = 0; c < W; c++) {
            RW[u][c] = JW[u][c] = 1.f;
            IRW[u][c] = IJW[u][c] = 1.0;
            LW[u][c] = MW[u][c] = NW[u][c] = GW[u][c] = 0.f;
#pragma unroll
            for (int f = 0; f < NS; f++) S0[u][f][c] = S1[u][f][c] = 0.f;
#pragma unroll
            for (int f = 0; f < NV; f++) W1[u][f][c] = W2[u][f][c] = 0.f;
        }
    }
    Row2<P, W> PF[4];

    // column groups: the edge cells of field f of a row computed in
    // iteration m go to ring entry x (x0 + k for the k-th exchanged field)
    // of slot m & 3; nbr reads what the neighbouring wavefronts put there
    auto put = [&](int sl, int x, const float (&v)[W]) {
        if (xw) {
            float *e = xl + ((sl * GM + gw) * NX + x) * 4 + xpos;
            e[0] = v[0];
            if constexpr (W == 2) e[1] = v[1];
        }
    };
    auto nbr = [&](int sl, int x) {
        const float *pv = xl + ((sl * GM + gp) * NX + x) * 4 + 2;
        const float *nv = xl + ((sl * GM + gn) * NX + x) * 4;
        return Nbr{pv[0], pv[1], nv[0], nv[1]};
    };
    // exchanged fields (z neighbours taken by the stencils): stresses 1, 2
    // (SH: 1; spin: 1, 2, 4, 5) and velocities 0, 1 (SH: 0; spin: 0, 1, 2)
    auto zsf = [](int k) { return P == SP_SH ? 1 : (k < 2 ? k + 1 : k + 2); };
    auto put_s = [&](int sl, int x0, const float (&s)[NS][W]) {
#pragma unroll
        for (int k = 0; k < NZS; k++) put(sl, x0 + k, s[zsf(k)]);
    };
    auto put_v = [&](int sl, int x0, const float (&v)[NV][W]) {
#pragma unroll
        for (int k = 0; k < NZV; k++) put(sl, x0 + k, v[k]);
    };
    // ring entries: s(n) [0, NZS), v(n+1) [NZS, NZS+NZV), s(n+1), v(n+2)
    constexpr int X0 = 0, X1 = NZS, X2 = NZS + NZV, X3 = 2 * NZS + NZV;
    const int sl0 = m0 & 3;  // slot of iteration m0 (m0 may be negative)

    // s(n) rows m0-3..m0-1 fill window slots 1..3
    load2<P, W>(PF[1], RS, RM, NP4, rd, m0 - 3, nx, nz, vj, i0, i1);
    load2<P, W>(PF[2], RS, RM, NP4, rd, m0 - 2, nx, nz, vj, i0, i1);
    load2<P, W>(PF[3], RS, RM, NP4, rd, m0 - 1, nx, nz, vj, i0, i1);
    load2<P, W>(PF[0], RS, RM, NP4, rd, m0, nx, nz, vj, i0, i1);
#pragma unroll
    for (int f = 0; f < NS; f++)
#pragma unroll
        for (int c = 0; c < W; c++) {
            S0[1][f][c] = PF[1].s[f][c];
            S0[2][f][c] = PF[2].s[f][c];
            S0[3][f][c] = PF[3].s[f][c];
        }
    // the s(n) row the first iteration's phase A centres on
    if constexpr (G) put_s((sl0 + 3) & 3, X0, S0[3]);

    for (int mb = m0; mb <= mend; mb += 4) {
#pragma unroll
        for (int u = 0; u < 4; u++) {
            const int m = mb + u;
            if (m > mend) break;
            const int A = (u + 1) & 3, B = (u + 2) & 3, C = (u + 3) & 3, D = u;
            // ring slots of this iteration and the two before it
            const int slm = (sl0 + u) & 3, sl1 = (sl0 + u + 3) & 3, sl2 = (sl0 + u + 2) & 3;
            Nbr nb[NZS > NZV ? NZS : NZV];
            if constexpr (G) __syncthreads();
            // The field offsets (slot x field bytes) are loop-invariant, but
            // kept as such they exhaust the SGPRs and are spilled to VGPR
            // lanes, one v_readlane per use.  Hiding the field size and the
            // stress/velocity set from the optimiser for each row makes the
            // compiler recompute them on the scalar unit instead (1-3%
            // faster).
            int np4 = NP4, rdv = rd;
            asm volatile("" : "+s"(np4), "+s"(rdv));
            const int NP4 = np4, rd = rdv, wr = rdv ^ 1;
            // prefetch one row ahead: the row of the next iteration (a
            // second row in flight measured no faster and costs about 70
            // VGPRs with two cells per lane)
            load2<P, W>(PF[A], RS, RM, NP4, rd, m + 1, nx, nz, vj, i0, i1);
            const Row2<P, W> &cur = PF[D];
#pragma unroll
            for (int f = 0; f < NS; f++)
#pragma unroll
                for (int c = 0; c < W; c++) S0[D][f][c] = cur.s[f][c];
            if constexpr (G) put_s(slm, X0, S0[D]);

            // ---- A: v(n+1) at row m-1; s(n) rows m-3..m = slots A, B, C, D
            {
                const int r = m - 1;
                const bool rin = r >= 0 && r < nx, xm_ok = xsel(r >= 2 && r < nx - 2);
                float dsv[NV][W];
                if constexpr (G) {
#pragma unroll
                    for (int k = 0; k < NZS; k++) nb[k] = nbr(sl1, X0 + k);
                }
                vel_row<P, EX, W, G>(dsv, S0[A], S0[B], S0[C], S0[D], xm_ok, zm_ok, KX, KZ, nb, lf, ll);
                if constexpr (IQ) inject_q<P, W>(dsv, r, j, nz, q1);
                else inject<P, W>(dsv, r, j, nz, p1, row1, &d, a.adjoint, a.nt, itx);
                const float bxr = bxp[min(max(r, 0), nx - 1)];
#pragma unroll
                for (int c = 0; c < W; c++) {
                    float dc[NV], vc[NV], vn[NV];
#pragma unroll
                    for (int f = 0; f < NV; f++) {
                        dc[f] = dsv[f][c];
                        vc[f] = cur.v[f][c];
                    }
                    // exact reciprocals of rho and j; spin keeps them for the
                    // pair's second step at this row (three iterations on)
                    double irr = 0.0, irj = 0.0;
                    if constexpr (EX) {
                        irr = rcp_rn((double)cur.rhoA[c]);
                        if constexpr (P == SP_SPIN) {
                            irj = rcp_rn((double)cur.jA[c]);
                            IRW[D][c] = irr;
                            IJW[D][c] = irj;
                        }
                    }
                    vel_finish<P, EX>(vn, dc, vc, cur.rhoA[c], cur.jA[c], irr, irj, bxr, bzb[c], bzt[c], zint, rin,
                                      jin, dt);
                    RW[D][c] = cur.rhoA[c];
                    JW[D][c] = cur.jA[c];
#pragma unroll
                    for (int f = 0; f < NV; f++) W1[C][f][c] = vn[f];
                }
                if constexpr (G) put_v(slm, X1, W1[C]);
            }
            // ---- B: s(n+1) at row m-3; v(n+1) rows m-4..m-1 = slots D, A, B, C
            {
                const int q = m - 3;
                const bool xp_ok = xsel(q >= 1 && q < nx - 2);
                float sn[NS][W];
                if constexpr (G) {
#pragma unroll
                    for (int k = 0; k < NZV; k++) nb[k] = nbr(sl2, X1 + k);
                }
                if (!SKIP || m >= i0) {  // rows before i0-3 feed no output row of the chunk
                    str_row<P, EX, W, G>(sn, W1[D], W1[A], W1[B], W1[C], S0[A], cur.lamB, cur.muB, cur.nuB,
                                         cur.gamB, xp_ok, zp_ok, KX, KZ, dt, nb, lf, ll);
#pragma unroll
                    for (int f = 0; f < NS; f++)
#pragma unroll
                        for (int c = 0; c < W; c++) S1[A][f][c] = sn[f][c];
                    if constexpr (G) put_s(slm, X2, S1[A]);
                }
#pragma unroll
                for (int c = 0; c < W; c++) {
                    LW[D][c] = cur.lamB[c];
                    MW[D][c] = cur.muB[c];
                    NW[D][c] = cur.nuB[c];
                    GW[D][c] = cur.gamB[c];
                }
            }
            // ---- C: v(n+2) at row m-4; s(n+1) rows m-6..m-3 = slots B, C, D, A
            {
                const int r = m - 4;
                const bool rin = r >= 0 && r < nx, xm_ok = xsel(r >= 2 && r < nx - 2);
                float dsv[NV][W], u1[NV][W], u2[NV][W];
                // rows before i0-1 feed no output row of the chunk: their
                // arithmetic is skipped (stores below are masked there)
                if (!SKIP || m >= i0 + 3) {
                    if constexpr (G) {
#pragma unroll
                        for (int k = 0; k < NZS; k++) nb[k] = nbr(sl1, X2 + k);
                    }
                    vel_row<P, EX, W, G>(dsv, S1[B], S1[C], S1[D], S1[A], xm_ok, zm_ok, KX, KZ, nb, lf, ll);
                    if constexpr (IQ) inject_q<P, W>(dsv, r, j, nz, q2);
                    else inject<P, W>(dsv, r, j, nz, p2, row2, &d, a.adjoint, a.nt, itx + 1);
                    const float bxr = bxp[min(max(r, 0), nx - 1)];
#pragma unroll
                    for (int c = 0; c < W; c++) {
                        float dc[NV], vc[NV], vn[NV];
#pragma unroll
                        for (int f = 0; f < NV; f++) {
                            dc[f] = dsv[f][c];
                            vc[f] = W1[D][f][c];
                        }
                        double irr = 0.0, irj = 0.0;
                        if constexpr (EX) {
                            if constexpr (P == SP_SPIN) {
                                irr = IRW[A][c];
                                irj = IJW[A][c];
                            } else {
                                irr = rcp_rn((double)RW[A][c]);
                            }
                        }
                        vel_finish<P, EX>(vn, dc, vc, RW[A][c], JW[A][c], irr, irj, bxr, bzb[c], bzt[c], zint, rin,
                                          jin, dt);
#pragma unroll
                        for (int f = 0; f < NV; f++) {
                            u1[f][c] = upd_disp<EX>(cur.u[f][c], W1[D][f][c], dt);
                            u2[f][c] = upd_disp<EX>(u1[f][c], vn[f], dt);
                            W2[D][f][c] = vn[f];
                        }
                    }
                    if constexpr (G) put_v(slm, X3, W2[D]);
                } else {
#pragma unroll
                    for (int f = 0; f < NV; f++)
#pragma unroll
                        for (int c = 0; c < W; c++) u1[f][c] = u2[f][c] = 0.f;
                }
                // rows outside the chunk: every lane's offset is out of range
                const bool rok = r >= i0 && r < i1;
                const unsigned ov = rok ? vo : OOB;
                const int orow = rok ? r * nz * 4 : 0;
#pragma unroll
                for (int f = 0; f < NV; f++) {
                    bstw<W>(RS, ov, slot_u(P, f) * NP4 + orow, u2[f]);
                    bstw<W>(RS, ov, slot_v(P, wr, f) * NP4 + orow, W2[D][f]);
                }
                // adjoint snapshots of either step (solver.py:193-203, 236-242):
                // [isa][u before the step | v after it][NV]
                if (SNAP && snap_v >= 0) {
                    const rsrc_t RP = make_rsrc(d.snap + (long long)snap_v * 2 * NV * (NP4 / 4), 2LL * NV * NP4);
#pragma unroll
                    for (int f = 0; f < NV; f++) {
                        bstw<W>(RP, ov, f * NP4 + orow, cur.u[f]);
                        bstw<W>(RP, ov, (NV + f) * NP4 + orow, W1[D][f]);
                    }
                }
                if (SNAP && snap_u >= 0) {
                    const rsrc_t RP = make_rsrc(d.snap + (long long)snap_u * 2 * NV * (NP4 / 4), 2LL * NV * NP4);
#pragma unroll
                    for (int f = 0; f < NV; f++) {
                        bstw<W>(RP, ov, f * NP4 + orow, u1[f]);
                        bstw<W>(RP, ov, (NV + f) * NP4 + orow, W2[D][f]);
                    }
                }
                // save_obs of both steps (solver.py:151-156) for receivers in
                // this row; rows past the chunk belong to the next chunk
                if (r < i1 && nk1 < (r + 1) * nz) {
                    cdesc_t *dq = fresh(&d);
                    const cint_p rk = (cint_p)dq->rec_k, rrow = (cint_p)dq->rec_row;
                    const int nrec = dq->nrec;
                    const rsrc_t RR = make_rsrc(dq->rec, (long long)n_rec(P) * nrec * a.nt * 4);
                    const int kr = r * nz + j, cs = nrec * a.nt * 4;
                    do {
                        const int c = nk1 - kr;  // 0 .. W-1 in the lane holding the receiver
                        const bool mine = outl && c >= 0 && c < W;
                        // column groups: only the wavefront holding the
                        // receiver stores (the group's wavefronts march in
                        // step, so stores in any of them hold up all; 2% at
                        // 500^2 x 64); strips measured no faster that way
                        if (!G || __any(mine)) {
                            const unsigned orc = mine ? (unsigned)(nr1 * a.nt + itx) * 4u : OOB;
#pragma unroll
                            for (int f = 0; f < NV; f++) {
                                bst(RR, orc, f * cs, (W == 2 && c == 1) ? u1[f][W - 1] : u1[f][0]);
                                bst(RR, orc, f * cs + 4, (W == 2 && c == 1) ? u2[f][W - 1] : u2[f][0]);
                            }
                        }
                        ++pr;
                        nk1 = nk2;
                        nr1 = nr2;
                        nk2 = pr + 1 < nrec ? rk[pr + 1] : 0x7fffffff;
                        nr2 = pr + 1 < nrec ? rrow[pr + 1] : 0;
                    } while (nk1 < (r + 1) * nz);
                }
            }
            // ---- D: s(n+2) at row m-6; v(n+2) rows m-7..m-4 = slots A, B, C, D
            {
                const int q = m - 6;
                const bool xp_ok = xsel(q >= 1 && q < nx - 2);
                float sn[NS][W];
                if (!SKIP || m >= i0 + 6) {  // rows before i0: no output (stores masked)
                    if constexpr (G) {
#pragma unroll
                        for (int k = 0; k < NZV; k++) nb[k] = nbr(sl2, X3 + k);
                    }
                    str_row<P, EX, W, G>(sn, W2[A], W2[B], W2[C], W2[D], S1[B], LW[A], MW[A], NW[A], GW[A], xp_ok,
                                         zp_ok, KX, KZ, dt, nb, lf, ll);
                } else {
#pragma unroll
                    for (int f = 0; f < NS; f++)
#pragma unroll
                        for (int c = 0; c < W; c++) sn[f][c] = 0.f;
                }
                const bool qok = q >= i0 && q < i1;
                const unsigned oq = qok ? vo : OOB;
                const int orow = qok ? q * nz * 4 : 0;
#pragma unroll
                for (int f = 0; f < NS; f++) bstw<W>(RS, oq, slot_s(P, wr, f) * NP4 + orow, sn[f]);
            }
        }
    }
}

template <int P, bool EX, int W, bool SQ, int M>
__global__ __launch_bounds__(256) void k_step2(StepArgs a)
{
    __shared__ int iq[M == STEP_QUEUE ? 4 * 2 * IQ_WORDS : 1];
    step2_body<P, EX, W, false, SQ, M>(a, nullptr, iq);
}

// column groups: blocks of up to grp_max wavefronts
template <int P, bool EX, int W, bool SQ, int M>
__global__ __launch_bounds__(64 * grp_max(P, W)) void k_step2g(StepArgs a)
{
    __shared__ float xl[4 * grp_max(P, W) * n_xch(P) * 4];
    __shared__ int iq[M == STEP_QUEUE ? grp_max(P, W) * 2 * IQ_WORDS : 1];
    step2_body<P, EX, W, true, SQ, M>(a, xl, iq);
}
