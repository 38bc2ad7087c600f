// seispie_hip -- k_step2w, two time steps per launch with two z cells per
// lane.
// Part of seispie_hip.hip: included there, inside its anonymous namespace,
// after <hip/hip_runtime.h> and the headers before this one.
#pragma once

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
