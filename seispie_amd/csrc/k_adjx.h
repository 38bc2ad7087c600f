// seispie_hip -- the discrete adjoint (SH, P-SV and spin): the exact
// transpose of the forward time stepping, edge bands and sponge included.
// Part of seispie_hip.hip: included there, inside its anonymous namespace,
// after <hip/hip_runtime.h> and the headers before this one.
#pragma once

// Forward step n (solver.py:205-234):  w = b (v + dt (A s + f) / M),
// u' = u + dt w, s' = s + dt C B w, records R u'  (A: D- divergence, B: D+
// strains, C: the pointwise stiffness, M: rho (and j for the spin velocity),
// b: the sponge factor).  Spin adds syx_c, syy_c, syz_c with B rows
// Dx+ vy_c, -vy_c - (Dx+ vz - Dz+ vx)/2, Dz+ vy_c and stiffness mu_c + nu_c,
// 2 nu, mu_c + nu_c; A adds Dz- syy_c to the x row, -Dx- syy_c to the z row
// and Dx- syx_c + Dz- syz_c + 2 syy_c to the vy_c row (solver.py:69-75,
// 135-141).  The adjoint
// state (s^, v^, u^) runs backwards from step nt-1:
//     u^ += R^T g(n)                                   k_xrec
//     q = dt u^ + B^T e,  e = dt C s^                  k_xw (e from k_xs)
//     v^ = b (v^ + q)                                  k_xw
//     s^ += A^T y, y = dt v^ / M, then e = dt C s^     k_xs
// and, with s^ before this step's update and w from the forward snapshot,
//     dL/dm += dt s^ . (dC/dm) B w  (lam, mu; spin nu, mu_c + nu_c),
//     dL/drho -= (w . q)_xz / rho,  spin dL/dj -= w_y q_y / j              k_xg
// Kernels are kept as dt dL/dm, the convention of the reference-style ones.
// The transposes take each neighbour only where its row of D-/D+ is valid:
//     (D-^T y)_j = c1 (y_j - y_{j+1}) - c2 (y_{j-1} - y_{j+2}),  rows 2 <= i < n-2
//     (D+^T y)_j = c1 (y_{j-1} - y_j) - c2 (y_{j-2} - y_{j+1}),  rows 1 <= i < n-2
// Storage is float32 in the shot's state block, arithmetic float64:
//     s^: stress set 0   v^: velocity set 0   u^: displacement
//     e : stress set 1   q : velocity set 1

__device__ __forceinline__ double tminus(const float *y, long long k, long long st, int i, int n, double c1, double c2)
{
    const double a = (i >= 2 && i < n - 2) ? (double)y[k] : 0.0;
    const double b = (i + 1 >= 2 && i + 1 < n - 2) ? (double)y[k + st] : 0.0;
    const double cm = (i - 1 >= 2 && i - 1 < n - 2) ? (double)y[k - st] : 0.0;
    const double d = (i + 2 >= 2 && i + 2 < n - 2) ? (double)y[k + 2 * st] : 0.0;
    return c1 * (a - b) - c2 * (cm - d);
}
__device__ __forceinline__ double tplus(const float *y, long long k, long long st, int i, int n, double c1, double c2)
{
    const double a = (i - 1 >= 1 && i - 1 < n - 2) ? (double)y[k - st] : 0.0;
    const double b = (i >= 1 && i < n - 2) ? (double)y[k] : 0.0;
    const double cm = (i - 2 >= 1 && i - 2 < n - 2) ? (double)y[k - 2 * st] : 0.0;
    const double d = (i + 1 >= 1 && i + 1 < n - 2) ? (double)y[k + st] : 0.0;
    return c1 * (a - b) - c2 * (cm - d);
}
// the forward's D+ of a float32 field (the same rounding of the differences)
__device__ __forceinline__ double fplus(const float *u, long long k, long long st, int i, int n, double c1, double c2)
{
    if (i >= 1 && i < n - 2)
        return c1 * (double)sub_rn(u[k + st], u[k]) - c2 * (double)sub_rn(u[k + 2 * st], u[k - st]);
    return 0.0;
}

// u^ += R^T g(n): records of step n, g = -(time-reversed adjoint source)
// as sp_misfit / sp_set_adjoint_source store it ([n_mis][nrec][nt]).  One
// thread per receiver; receivers sharing a cell are summed by the first of
// them (receivers are sorted by cell), so no two threads write one cell.
template <int P>
__global__ __launch_bounds__(256) void k_xrec(const ShotDesc *shots, int nshots, int it, int nt)
{
    const int s = blockIdx.y;
    const int r = blockIdx.x * 256 + threadIdx.x;
    if (s >= nshots) return;
    const ShotDesc &d = shots[s];
    if (r >= d.nrec) return;
    const int kr = d.rec_k[r];
    if (r > 0 && d.rec_k[r - 1] == kr) return;
    const long long NP = d.Npad;
    const long long tr = nt - 1 - it;
    for (int c = 0; c < n_mis(P); c++) {
        double acc = 0.0;
        for (int q = r; q < d.nrec && d.rec_k[q] == kr; q++)
            acc -= (double)d.adstf[((long long)c * d.nrec + d.rec_row[q]) * nt + tr];
        float *uh = d.state + slot_u(P, c) * NP + kr;
        *uh = (float)((double)*uh + acc);
    }
}

struct XAdjArgs {
    const ShotDesc *shots;
    const int *list;  // shots of this grid in the batch
    int nlist;
    int nx, nz;
    long long Npad;
    float *kern;      // [SP_NKERN][Npad]
    int slot;         // snapshot slot holding v after step n
    double dt, c1x, c2x, c1z, c2z;
};

// q = dt u^ + B^T e and v^ = b (v^ + q); q is kept (velocity set 1) for the
// density kernel.  One thread per cell and shot.
template <int P>
__global__ __launch_bounds__(256) void k_xw(const ShotDesc *shots, int nshots, long long maxN, double dt)
{
#pragma clang fp contract(off)
    constexpr int NV = n_v(P);
    const long long k = (long long)blockIdx.x * 256 + threadIdx.x;
    const int s = blockIdx.y;
    if (s >= nshots) return;
    const ShotDesc &d = shots[s];
    const int nx = d.nx, nz = d.nz;
    if (k >= (long long)nx * nz) return;
    const int i = (int)(k / nz), j = (int)(k - (long long)i * nz);
    const long long NP = d.Npad;
    float *st = d.state;
    const float b1 = (float)((double)d.bx[i] * d.bzb[j]);
    const double b = (double)(float)((double)b1 * d.bzt[j]);
    const float *e0 = st + slot_s(P, 1, 0) * NP, *e1 = st + slot_s(P, 1, 1) * NP;
    double bt[NV];
    if constexpr (P == SP_SH) {
        bt[0] = tplus(e0, k, nz, i, nx, d.c1x, d.c2x) + tplus(e1, k, 1, j, nz, d.c1z, d.c2z);
    } else {
        const float *e2 = st + slot_s(P, 1, 2) * NP;
        bt[0] = tplus(e0, k, nz, i, nx, d.c1x, d.c2x) + tplus(e2, k, 1, j, nz, d.c1z, d.c2z);
        bt[1] = tplus(e1, k, 1, j, nz, d.c1z, d.c2z) + tplus(e2, k, nz, i, nx, d.c1x, d.c2x);
        if constexpr (P == SP_SPIN) {
            const float *e3 = st + slot_s(P, 1, 3) * NP, *e4 = st + slot_s(P, 1, 4) * NP;
            const float *e5 = st + slot_s(P, 1, 5) * NP;
            bt[0] += 0.5 * tplus(e4, k, 1, j, nz, d.c1z, d.c2z);
            bt[1] -= 0.5 * tplus(e4, k, nz, i, nx, d.c1x, d.c2x);
            bt[2] = -(double)e4[k] + tplus(e3, k, nz, i, nx, d.c1x, d.c2x) + tplus(e5, k, 1, j, nz, d.c1z, d.c2z);
        }
    }
#pragma unroll
    for (int c = 0; c < NV; c++) {
        const double q = dt * (double)st[slot_u(P, c) * NP + k] + bt[c];
        float *vh = st + slot_v(P, 0, c) * NP + k;
        *vh = (float)(b * ((double)*vh + q));
        st[slot_v(P, 1, c) * NP + k] = (float)q;
    }
}

// The kernels' step-n terms: one thread per grid cell, the grid's shots in
// turn (summed in float64, then added to the float32 kernels once).  Only
// loads inside the shot loop, so the loads of successive shots overlap.
template <int P>
__global__ __launch_bounds__(256) void k_xg(XAdjArgs a)
{
#pragma clang fp contract(off)
    constexpr int NV = n_v(P);
    const long long N = (long long)a.nx * a.nz;
    const long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= N) return;
    const int nx = a.nx, nz = a.nz;
    const int i = (int)(k / nz), j = (int)(k - (long long)i * nz);
    const long long NP = a.Npad;
    const double dt = a.dt;
    double glam = 0.0, gmu = 0.0, grho = 0.0, ghess = 0.0, gnu = 0.0, gj = 0.0, gmuc = 0.0;
    for (int t = 0; t < a.nlist; t++) {
        const ShotDesc &d = a.shots[a.list[t]];
        const float *st = d.state;
        const float *w = d.snap + ((long long)a.slot * 2 * NV + NV) * NP;  // v after step n
        const double rho = d.mat[M_RHO * NP + k];
        double wq = 0.0;
#pragma unroll
        for (int c = 0; c < (P == SP_SPIN ? 2 : NV); c++) {
            const double wc = w[c * NP + k];
            wq += wc * (double)st[slot_v(P, 1, c) * NP + k];
            ghess += wc * wc;
        }
        grho -= wq / rho;
        if constexpr (P == SP_SPIN)
            gj -= (double)w[2 * NP + k] * (double)st[slot_v(P, 1, 2) * NP + k] / (double)d.mat[M_J * NP + k];
        const double s0 = st[slot_s(P, 0, 0) * NP + k], s1 = st[slot_s(P, 0, 1) * NP + k];
        if constexpr (P == SP_SH) {
            const double ex = fplus(w, k, nz, i, nx, a.c1x, a.c2x), ez = fplus(w, k, 1, j, nz, a.c1z, a.c2z);
            gmu += dt * (s0 * ex + s1 * ez);
        } else {
            const double s2 = st[slot_s(P, 0, 2) * NP + k];
            const float *wz = w + NP;
            const double exx = fplus(w, k, nz, i, nx, a.c1x, a.c2x), ezz = fplus(wz, k, 1, j, nz, a.c1z, a.c2z);
            const double dxz = fplus(w, k, 1, j, nz, a.c1z, a.c2z), dzx = fplus(wz, k, nz, i, nx, a.c1x, a.c2x);
            glam += dt * (s0 + s1) * (exx + ezz);
            gmu += dt * (2.0 * s0 * exx + 2.0 * s1 * ezz + s2 * (dxz + dzx));
            if constexpr (P == SP_SPIN) {
                const float *wy = w + 2 * NP;
                const double s3 = st[slot_s(P, 0, 3) * NP + k], s4 = st[slot_s(P, 0, 4) * NP + k];
                const double s5 = st[slot_s(P, 0, 5) * NP + k];
                const double ee = -(double)wy[k] - 0.5 * (dzx - dxz);
                gnu += dt * 2.0 * s4 * ee;
                gmuc += dt * (s3 * fplus(wy, k, nz, i, nx, a.c1x, a.c2x) + s5 * fplus(wy, k, 1, j, nz, a.c1z, a.c2z));
            }
        }
    }
    float *K = a.kern;
    if constexpr (P != SP_SH) K[SP_K_LAM * NP + k] = (float)((double)K[SP_K_LAM * NP + k] + dt * glam);
    K[SP_K_MU * NP + k] = (float)((double)K[SP_K_MU * NP + k] + dt * gmu);
    K[SP_K_RHO * NP + k] = (float)((double)K[SP_K_RHO * NP + k] + dt * grho);
    K[SP_K_HESS * NP + k] = (float)((double)K[SP_K_HESS * NP + k] + dt * ghess);
    if constexpr (P == SP_SPIN) {
        K[SP_K_NU * NP + k] = (float)((double)K[SP_K_NU * NP + k] + dt * gnu);
        K[SP_K_J * NP + k] = (float)((double)K[SP_K_J * NP + k] + dt * gj);
        K[SP_K_MUC * NP + k] = (float)((double)K[SP_K_MUC * NP + k] + dt * gmuc);
    }
}

// D-^T of y = dt v^ / rho, taken on the fly from v^ and rho at the neighbours
__device__ __forceinline__ double tminus_y(const float *vh, const float *rho, long long k, long long st, int i, int n,
                                           double c1, double c2, double dt)
{
#define YV(q, kk) (((q) >= 2 && (q) < n - 2) ? dt * (double)vh[kk] / (double)rho[kk] : 0.0)
    const double a = YV(i, k), b = YV(i + 1, k + st), cm = YV(i - 1, k - st), d = YV(i + 2, k + 2 * st);
#undef YV
    return c1 * (a - b) - c2 * (cm - d);
}

// s^ += A^T (dt v^ / rho), then e = dt C s^ for the next (earlier) step; one
// thread per cell and shot.
template <int P>
__global__ __launch_bounds__(256) void k_xs(const ShotDesc *shots, int nshots, long long maxN, double dt)
{
#pragma clang fp contract(off)
    const long long k = (long long)blockIdx.x * 256 + threadIdx.x;
    const int s = blockIdx.y;
    if (s >= nshots) return;
    const ShotDesc &d = shots[s];
    const int nx = d.nx, nz = d.nz;
    if (k >= (long long)nx * nz) return;
    const int i = (int)(k / nz), j = (int)(k - (long long)i * nz);
    const long long NP = d.Npad;
    float *st = d.state;
    const float *rho = d.mat + M_RHO * NP;
    const float *vx = st + slot_v(P, 0, 0) * NP;
    const double mu = d.mat[M_MU * NP + k];
    if constexpr (P == SP_SH) {
        const double s0 = (double)st[slot_s(P, 0, 0) * NP + k] + tminus_y(vx, rho, k, nz, i, nx, d.c1x, d.c2x, dt);
        const double s1 = (double)st[slot_s(P, 0, 1) * NP + k] + tminus_y(vx, rho, k, 1, j, nz, d.c1z, d.c2z, dt);
        const float f0 = (float)s0, f1 = (float)s1;
        st[slot_s(P, 0, 0) * NP + k] = f0;
        st[slot_s(P, 0, 1) * NP + k] = f1;
        st[slot_s(P, 1, 0) * NP + k] = (float)(dt * mu * (double)f0);
        st[slot_s(P, 1, 1) * NP + k] = (float)(dt * mu * (double)f1);
    } else {
        const float *vz = st + slot_v(P, 0, 1) * NP;
        const double s0 = (double)st[slot_s(P, 0, 0) * NP + k] + tminus_y(vx, rho, k, nz, i, nx, d.c1x, d.c2x, dt);
        const double s1 = (double)st[slot_s(P, 0, 1) * NP + k] + tminus_y(vz, rho, k, 1, j, nz, d.c1z, d.c2z, dt);
        const double s2 = (double)st[slot_s(P, 0, 2) * NP + k] + tminus_y(vx, rho, k, 1, j, nz, d.c1z, d.c2z, dt) +
                          tminus_y(vz, rho, k, nz, i, nx, d.c1x, d.c2x, dt);
        const float f0 = (float)s0, f1 = (float)s1, f2 = (float)s2;
        st[slot_s(P, 0, 0) * NP + k] = f0;
        st[slot_s(P, 0, 1) * NP + k] = f1;
        st[slot_s(P, 0, 2) * NP + k] = f2;
        const double lam = d.mat[M_LAM * NP + k], l2m = lam + 2.0 * mu;
        st[slot_s(P, 1, 0) * NP + k] = (float)(dt * (l2m * (double)f0 + lam * (double)f1));
        st[slot_s(P, 1, 1) * NP + k] = (float)(dt * (lam * (double)f0 + l2m * (double)f1));
        st[slot_s(P, 1, 2) * NP + k] = (float)(dt * mu * (double)f2);
        if constexpr (P == SP_SPIN) {
            const float *vy = st + slot_v(P, 0, 2) * NP, *jj = d.mat + M_J * NP;
            const double s3 = (double)st[slot_s(P, 0, 3) * NP + k] + tminus_y(vy, jj, k, nz, i, nx, d.c1x, d.c2x, dt);
            const double s4 = (double)st[slot_s(P, 0, 4) * NP + k] + tminus_y(vx, rho, k, 1, j, nz, d.c1z, d.c2z, dt) -
                              tminus_y(vz, rho, k, nz, i, nx, d.c1x, d.c2x, dt) +
                              2.0 * (dt * (double)vy[k] / (double)jj[k]);
            const double s5 = (double)st[slot_s(P, 0, 5) * NP + k] + tminus_y(vy, jj, k, 1, j, nz, d.c1z, d.c2z, dt);
            const float f3 = (float)s3, f4 = (float)s4, f5 = (float)s5;
            st[slot_s(P, 0, 3) * NP + k] = f3;
            st[slot_s(P, 0, 4) * NP + k] = f4;
            st[slot_s(P, 0, 5) * NP + k] = f5;
            const double gam = d.mat[M_GAM * NP + k], nu = d.mat[M_NU * NP + k];
            st[slot_s(P, 1, 3) * NP + k] = (float)(dt * gam * (double)f3);
            st[slot_s(P, 1, 4) * NP + k] = (float)(dt * 2.0 * nu * (double)f4);
            st[slot_s(P, 1, 5) * NP + k] = (float)(dt * gam * (double)f5);
        }
    }
}
