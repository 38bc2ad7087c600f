// seispie_hip -- k_step3, one time step per launch with buffer loads and
// stores.
// Part of seispie_hip.hip: included there, inside its anonymous namespace,
// after <hip/hip_runtime.h> and the headers before this one.
#pragma once

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
//   * snapshots are a separate kernel (k_snap); receiver records of a
//     forward step are stored in a wave-uniform branch of the rows that hold
//     receivers, as the two-step kernel does; source injection reads the
//     time function with scalar loads only;
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

template <int P, bool EX, bool SQ, bool IQ>
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
    cdesc_t &d = ((cdesc_t *)a.shots)[item.x];
    const int itx = a.itp ? a.it + *(cint_p)a.itp : a.it;
    const int nx = d.nx, nz = d.nz, i0 = item.z, i1 = item.w;
    const int j = item.y - a.shift + L;
    const int jc = min(max(j, 0), nz - 1);
    const bool jin = j >= 0 && j < nz;
    const bool outl = jin && L >= a.shift && L < a.shift + a.wout;
    const bool zm_ok = j >= 2 && j < nz - 2, zp_ok = j >= 1 && j < nz - 2;
    const int NP4 = (int)(d.Npad * 4);  // bytes per field
    const double dt = a.dt, dt2 = 2.0 * a.dt;
    // square cells (SQ): one set of difference coefficients for both axes
    const DCo KX{d.c1x, d.c2x, d.k9hx, d.k9lx, d.k1hx, d.k1lx};
    const DCo KZ = SQ ? KX : DCo{d.c1z, d.c2z, d.k9hz, d.k9lz, d.k1hz, d.k1lz};
    const int rd = a.rd, wr = a.rd ^ 1;
    const double bzb = d.bzb[jc], bzt = d.bzt[jc];
    const rsrc_t RS = make_rsrc(d.state, (long long)n_slots(P) * NP4);
    const rsrc_t RM = make_rsrc(d.mat, (long long)N_MAT * NP4);
    const cflt_p bxp = (cflt_p)d.bx;

    // injection list: sources (forward) or receivers with adjoint sources;
    // IQ (launches in which no item reaches more than IQ_CAP entries): a
    // queue of the entries in LDS
    const cint_p ik = (cint_p)(a.adjoint ? d.rec_k : d.src_k);
    const cint_p irow = (cint_p)(a.adjoint ? d.rec_row : d.src_row);
    const int ni = a.adjoint ? d.nrec : d.nsrc;
    const cflt_p ival = (cflt_p)(a.adjoint ? (const float *)d.adstf : d.stf);
    const long long nrows = a.adjoint ? d.nrec : d.nsrc;
    int pi = __builtin_amdgcn_readfirstlane(a.adjoint ? item2.y : item2.x);
    while (pi < ni && ik[pi] < (i0 - 1) * nz) ++pi;  // the list pointer starts at row i0-4
    int src_row = 0x7fffffff;
    __shared__ int iq[IQ ? 4 * IQ_WORDS : 1];
    InjQ q;
    if constexpr (IQ) {
        injq_fill<P>(q, iq + IQ_WORDS * (threadIdx.x >> 6), &d, a.adjoint, a.nt, itx, pi, item2.w);
    } else {
        src_row = pi < ni ? ik[pi] / nz : 0x7fffffff;
    }
    // receivers recorded by this chunk (forward only): rows i0..i1-1.  The
    // next two receivers' cells and record rows wait in scalar registers, so
    // a row's test uses values loaded rows earlier and the loads of later
    // receivers overlap the march
    const cint_p rk = (cint_p)d.rec_k;
    const cint_p rrow = (cint_p)d.rec_row;
    const int nrec = a.adjoint ? 0 : d.nrec;
    int pr = __builtin_amdgcn_readfirstlane(item2.z);
    int nk1 = pr < nrec ? rk[pr] : 0x7fffffff, nr1 = pr < nrec ? rrow[pr] : 0;
    int nk2 = pr + 1 < nrec ? rk[pr + 1] : 0x7fffffff, nr2 = pr + 1 < nrec ? rrow[pr + 1] : 0;
    const rsrc_t RR = make_rsrc(d.rec, (long long)n_rec(P) * d.nrec * a.nt * 4);

    float S[4][NS], W[4][NV];
    Row3<P> PF[4];
#pragma unroll
    for (int u = 0; u < 4; u++) {
#pragma unroll
        for (int f = 0; f < NS; f++) S[u][f] = 0.f;
#pragma unroll
        for (int f = 0; f < NV; f++) W[u][f] = 0.f;
    }

    // prologue: rows i0-3, i0-2, i0-1 fill stress slots 1, 2, 3 (slots are
    // counted from row i0, so any i0 works); rows up to i0+1 are prefetched.
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
                double ddz = lsel(zm_ok, dmn<EX>(zm2, zm1, Sc[1], zp1, KZ));
                dsv[0] = (float)(ddx + ddz);
            } else {
                float xzm1 = lane_m1(Sc[2]), xzm2 = lane_m1(xzm1), xzp1 = lane_p1(Sc[2]);
                float zzm1 = lane_m1(Sc[1]), zzm2 = lane_m1(zzm1), zzp1 = lane_p1(Sc[1]);
                double dxx = xm_ok ? dmn<EX>(Sa[0], Sb[0], Sc[0], Sd[0], KX) : 0.0;
                double dxxz = xm_ok ? dmn<EX>(Sa[2], Sb[2], Sc[2], Sd[2], KX) : 0.0;
                double dzxz = lsel(zm_ok, dmn<EX>(xzm2, xzm1, Sc[2], xzp1, KZ));
                double dzzz = lsel(zm_ok, dmn<EX>(zzm2, zzm1, Sc[1], zzp1, KZ));
                dsv[0] = (float)(dxx + dzxz);
                dsv[1] = (float)(dxxz + dzzz);
                if constexpr (P == SP_SPIN) {
                    float yzm1 = lane_m1(Sc[5]), yzm2 = lane_m1(yzm1), yzp1 = lane_p1(Sc[5]);
                    float yym1 = lane_m1(Sc[4]), yym2 = lane_m1(yym1), yyp1 = lane_p1(Sc[4]);
                    double dxyx = xm_ok ? dmn<EX>(Sa[3], Sb[3], Sc[3], Sd[3], KX) : 0.0;
                    double dzyz = lsel(zm_ok, dmn<EX>(yzm2, yzm1, Sc[5], yzp1, KZ));
                    float dsy = (float)(dxyx + dzyz);
                    double dzyy = lsel(zm_ok, dmn<EX>(yym2, yym1, Sc[4], yyp1, KZ));
                    double dxyy = xm_ok ? dmn<EX>(Sa[4], Sb[4], Sc[4], Sd[4], KX) : 0.0;
                    dsv[0] = (float)((double)dsv[0] + dzyy);
                    dsv[1] = (float)((double)dsv[1] - dxyy);
                    dsv[2] = (float)((double)dsy + 2.0 * (double)Sc[4]);
                }
            }
            // stf_dsy / stf_dsxz (solver.py:78-93): float32 adds, list order
            if constexpr (IQ) {
                const int kr = r * nz + j;
                injq_row<P>(q, r, [&](int kk, float s0, float s1) {
                    if (kk == kr) {
                        dsv[0] = add_rn(dsv[0], s0);
                        if constexpr (P != SP_SH) dsv[1] = add_rn(dsv[1], s1);
                    }
                });
            } else if (r == src_row) {
                // every read here is wave-uniform (scalar loads)
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
            // save_obs (solver.py:151-156): displacement after the step, for
            // receivers in this row; rows past the chunk belong to the next
            if (r < i1 && nk1 < (r + 1) * nz) {
                const int kr = r * nz + j, cs = d.nrec * a.nt * 4;
                do {
                    const unsigned orc = (outl && nk1 == kr) ? (unsigned)(nr1 * a.nt + itx) * 4u : OOB;
#pragma unroll
                    for (int f = 0; f < NV; f++) bst(RR, orc, f * cs, un[f]);
                    ++pr;
                    nk1 = nk2;
                    nr1 = nr2;
                    nk2 = pr + 1 < nrec ? rk[pr + 1] : 0x7fffffff;
                    nr2 = pr + 1 < nrec ? rrow[pr + 1] : 0;
                } while (nk1 < (r + 1) * nz);
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
            // rows before i0 are no output of the chunk: their arithmetic is
            // skipped (the stores, masked there, always issue)
            float sn[NS];
            if (qr < i0) {
#pragma unroll
                for (int f = 0; f < NS; f++) sn[f] = 0.f;
            } else {
                if constexpr (P == SP_SH) {
                    float zm1 = lane_m1(Wb[0]), zp1 = lane_p1(Wb[0]), zp2 = lane_p1(zp1);
                    const float dvdx = xp_ok ? (float)dpl<EX>(Wa[0], Wb[0], Wc[0], Wd[0], KX) : 0.f;
                    const float dvdz = lsel(zp_ok, (float)dpl<EX>(zm1, Wb[0], zp1, zp2, KZ));
                    const double dtmu = dt * (double)cur.mu;
                    sn[0] = upd_sadd<EX>(Sa[0], dtmu, (double)dvdx);
                    sn[1] = upd_sadd<EX>(Sa[1], dtmu, (double)dvdz);
                } else {
                    float xm1 = lane_m1(Wb[0]), xp1 = lane_p1(Wb[0]), xp2 = lane_p1(xp1);
                    float zm1 = lane_m1(Wb[1]), zp1 = lane_p1(Wb[1]), zp2 = lane_p1(zp1);
                    const float dvxdx = xp_ok ? (float)dpl<EX>(Wa[0], Wb[0], Wc[0], Wd[0], KX) : 0.f;
                    const float dvzdx = xp_ok ? (float)dpl<EX>(Wa[1], Wb[1], Wc[1], Wd[1], KX) : 0.f;
                    const float dvxdz = lsel(zp_ok, (float)dpl<EX>(xm1, Wb[0], xp1, xp2, KZ));
                    const float dvzdz = lsel(zp_ok, (float)dpl<EX>(zm1, Wb[1], zp1, zp2, KZ));
                    const float lam = cur.lam, mu = cur.mu;
                    const double l2m = (double)lam + 2.0 * (double)mu;
                    sn[0] = upd_snorm<EX>(Sa[0], dt, l2m, dvxdx, mul_rn(lam, dvzdz));
                    sn[1] = upd_snorm<EX>(Sa[1], dt, l2m, dvzdz, mul_rn(lam, dvxdx));
                    sn[2] = upd_sadd<EX>(Sa[2], dt, (double)mul_rn(mu, add_rn(dvxdz, dvzdx)));
                    if constexpr (P == SP_SPIN) {
                        float cm1 = lane_m1(Wb[2]), cp1 = lane_p1(Wb[2]), cp2 = lane_p1(cp1);
                        const float dcdx = xp_ok ? (float)dpl<EX>(Wa[2], Wb[2], Wc[2], Wd[2], KX) : 0.f;
                        const float dcdz = lsel(zp_ok, (float)dpl<EX>(cm1, Wb[2], cp1, cp2, KZ));
                        const double dtg = dt * (double)cur.gam;
                        sn[4] = upd_srot<EX>(Sa[4], dt2, cur.nu, Wb[2], sub_rn(dvzdx, dvxdz));
                        sn[3] = upd_sadd<EX>(Sa[3], dtg, (double)dcdx);
                        sn[5] = upd_sadd<EX>(Sa[5], dtg, (double)dcdz);
                    }
                }
            }
#pragma unroll
            for (int f = 0; f < NS; f++) bst(RS, oq, slot_s(P, wr, f) * NP4, sn[f]);
        }
    }
}
