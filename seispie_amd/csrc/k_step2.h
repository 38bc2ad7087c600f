// seispie_hip -- k_step2, two time steps per launch.
// Part of seispie_hip.hip: included there, inside its anonymous namespace,
// after <hip/hip_runtime.h> and the headers before this one.
#pragma once

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
