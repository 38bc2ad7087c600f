// seispie_hip -- k_step, the plain one-step kernel (grids too large for
// 32-bit buffer offsets).
// Part of seispie_hip.hip: included there, inside its anonymous namespace,
// after <hip/hip_runtime.h> and the headers before this one.
#pragma once

// ------------------------------------------------------------ step kernel
// Global-address-space pointers: loads through them become global_load with
// a scalar base and one 32-bit offset per row shared by every field.
typedef const float __attribute__((address_space(1))) *gcf;
typedef float __attribute__((address_space(1))) *gf;


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
