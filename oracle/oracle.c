/*
 * seispie fd2d oracle -- CPU restatement of the reference solver hot path.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP product path and the
 * CPU baseline timed by bench.py.  The product (seispie_amd/) never links,
 * loads or calls it.
 *
 * Follows, kernel by kernel:
 *   seispie/solver/fd2d/solver.py   stencils :22-163, run_forward :166-242,
 *                                   run_adjoint :290-321
 *   seispie/solver/fd2d/fd2d.py     vps2lm :21-33, set_bound :47-68,
 *                                   gaussian smoothing :71-108
 *   seispie/solver/misfit/waveform.py   diff :5-25
 * See oracle_body.h for the numerics contract (float32 storage, float64
 * arithmetic where Numba promotes) and the FP64 replay.
 *
 * Build: make -C oracle   (the checker: gcc -O2 -ffp-contract=off -fopenmp -shared)
 *        make -C oracle baseline BASELINE_OUT=<dir>   (the CPU baseline:
 *        gcc -O3 -march=native -fopenmp, compiled on the host it is timed on)
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <omp.h>
#include <time.h>

#define ORC_SH 1
#define ORC_PSV 2
#define ORC_SPIN 3

typedef struct {
    int physics; /* ORC_SH, ORC_PSV, ORC_SPIN */
    int nx, nz, nt;
    float dx, dz; /* float32 scalars in the reference (fd2d.py:195-196) */
    double dt;    /* Python float (fd2d.py:115) */
    const float *rho, *lam, *mu, *nu, *jj, *mu_c, *nu_c; /* after vps2lm */
    const float *bound;
} orc_model;

typedef struct {
    int nsrc;
    const int *src_id;
    int isrc; /* -1: all sources (combine_sources = yes) */
    int nrec;
    const int *rec_id;
} orc_geom;

int orc_nsnap(int phys) { return phys == ORC_SH ? 2 : (phys == ORC_PSV ? 4 : 6); }

#define T float
#define SFX f32
#include "oracle_body.h"
#undef T
#undef SFX

#define T double
#define SFX f64
#include "oracle_body.h"
#undef T
#undef SFX

/* set_bound: fd2d.py:47-68 (each factor multiplies the stored float32) */
void orc_set_bound(float *bound, int width, double alpha, int left, int right, int bottom, int top,
                   int nx, int nz)
{
    long n = (long)nx * nz, k;
    for (k = 0; k < n; k++) {
        int j = (int)(k % nz), i = (int)(k / nz);
        double aw;
        bound[k] = 1.0f;
        if (left && i + 1 < width) {
            aw = alpha * (double)(width - i - 1);
            bound[k] = (float)((double)bound[k] * exp(-aw * aw));
        }
        if (right && i > nx - width) {
            aw = alpha * (double)(width + i - nx);
            bound[k] = (float)((double)bound[k] * exp(-aw * aw));
        }
        if (bottom && j > nz - width) {
            aw = alpha * (double)(width + j - nz);
            bound[k] = (float)((double)bound[k] * exp(-aw * aw));
        }
        if (top && j + 1 < width) {
            aw = alpha * (double)(width - j - 1);
            bound[k] = (float)((double)bound[k] * exp(-aw * aw));
        }
    }
}

/* vps2lm: fd2d.py:21-33.  In: lam = vp, mu = vs; out: lambda, mu. */
void orc_vps2lm(float *lam, float *mu, const float *rho, long n)
{
    long k;
    for (k = 0; k < n; k++) {
        float vp = lam[k], vs = mu[k];
        if (vp > vs)
            lam[k] = (float)((double)rho[k] * ((double)(vp * vp) - 2.0 * (double)vs * (double)vs));
        else
            lam[k] = 0.0f;
        mu[k] = rho[k] * vs * vs;
    }
}

/* CPU baseline: run warm + nsteps forward steps of one shot with the
 * unfused reference kernel sequence on nthreads OpenMP threads and return
 * the wall time of the last nsteps.  Sources fire every step. */
double orc_bench_forward_f32(const orc_model *m, const orc_geom *g, const float *stf_x,
                             const float *stf_y, const float *stf_z, int warm, int nsteps,
                             int nthreads)
{
    state_f32 s;
    int it;
    struct timespec t0, t1;
    omp_set_num_threads(nthreads);
    state_init_f32(&s, m);
    clear_f32(&s);
    for (it = 0; it < warm + nsteps; it++) {
        if (it == warm) clock_gettime(CLOCK_MONOTONIC, &t0);
        step_f32(&s, m, g->nsrc, g->src_id, g->isrc, stf_x, stf_y, stf_z, it % m->nt);
    }
    clock_gettime(CLOCK_MONOTONIC, &t1);
    state_free_f32(&s);
    return (double)(t1.tv_sec - t0.tv_sec) + 1e-9 * (double)(t1.tv_nsec - t0.tv_nsec);
}

/* CPU baseline of the adjoint run (solver.py:290-321): warm + nsteps steps
 * with the adjoint sources injected at the receivers and the interaction
 * every sae steps against one stored snapshot (nsnap fields of the
 * physics, values irrelevant to the timing); returns the wall time of the
 * last nsteps. */
double orc_bench_adjoint_f32(const orc_model *m, const orc_geom *g, const float *ad_x,
                             const float *ad_y, const float *ad_z, int sae, const float *snap,
                             int warm, int nsteps, int nthreads)
{
    state_f32 s;
    int it;
    long n = (long)m->nx * m->nz;
    struct timespec t0, t1;
    double ndt = (double)sae * m->dt;
    float *k = (float *)calloc((size_t)(6 * n), sizeof(float));
    omp_set_num_threads(nthreads);
    state_init_f32(&s, m);
    clear_f32(&s);
    for (it = 0; it < warm + nsteps; it++) {
        if (it == warm) clock_gettime(CLOCK_MONOTONIC, &t0);
        step_f32(&s, m, g->nrec, g->rec_id, -1, ad_x, ad_y, ad_z, it % m->nt);
        if (it % sae == 0)
            interact_f32(&s, m, snap, ndt, k, k + n, k + 2 * n, k + 3 * n, k + 4 * n, k + 5 * n);
    }
    clock_gettime(CLOCK_MONOTONIC, &t1);
    state_free_f32(&s);
    free(k);
    return (double)(t1.tv_sec - t0.tv_sec) + 1e-9 * (double)(t1.tv_nsec - t0.tv_nsec);
}

int orc_max_threads(void) { return omp_get_max_threads(); }

/* The four staggered differences on a whole grid (which: 0 = D-x,
 * 1 = D+x, 2 = D-z, 3 = D+z), float64, for operator-level tests. */
void orc_diff_f64(const double *v, double *out, int nx, int nz, double dx, double dz, int which)
{
    long n = (long)nx * nz, k;
    for (k = 0; k < n; k++) {
        int j = (int)(k % nz), i = (int)(k / nz);
        switch (which) {
        case 0: out[k] = dxm_f64(v, i, k, dx, nx, nz); break;
        case 1: out[k] = dxp_f64(v, i, k, dx, nx, nz); break;
        case 2: out[k] = dzm_f64(v, j, k, dz, nz); break;
        default: out[k] = dzp_f64(v, j, k, dz, nz); break;
        }
    }
}
