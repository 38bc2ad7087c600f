// seispie_hip -- the kernels beside the step: records, snapshots, kernel
// interaction, misfit, smoothing and small elementwise kernels.
// Part of seispie_hip.hip: included there, inside its anonymous namespace,
// after <hip/hip_runtime.h> and the headers before this one.
#pragma once

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
    int isae;         // snapshot slot
    // replayed graphs: the slot is isae + (device step counter) / sae; a slot
    // outside [0, nsa) writes nothing
    const int *itp;
    int sae, nsa;
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
    const int isae = a.itp ? a.isae + *(cint_p)a.itp / a.sae : a.isae;
    if (isae < 0 || isae >= a.nsa) return;
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
            const float *uf = d.snap + (long long)isae * n_snap(P) * NP;
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
        const float *sn = d.snap + (long long)isae * n_snap(P) * NP;
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
