/*
 * seispie_hip -- C ABI of the MI355X (gfx950) solver library.
 *
 * The reference (icui/seispie) has no native code: its solver is the
 * Python class seispie/solver/fd2d/fd2d.py driving Numba-CUDA kernels in
 * seispie/solver/fd2d/solver.py.  Each entry point below names the
 * reference method or kernel sequence it replaces; the Python plugin
 * seispie_amd.solver.fd2d binds them through ctypes.
 *
 * Conventions
 *   - Every function returns 0 on success and a nonzero code on failure;
 *     sp_last_error(ctx) then holds a message.  Nothing throws across the ABI.
 *   - Grid arrays are float32 with k = i*nz + j (z fastest), as the
 *     reference (solver.py:14-19).
 *   - Host buffers are borrowed for the duration of the call.  Arguments
 *     named *_on_device == 1 take device pointers on the context's device
 *     (ordered on the context's stream) instead.
 *   - A context is not thread safe; use one per host thread.
 */
#ifndef SEISPIE_HIP_H
#define SEISPIE_HIP_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct sp_ctx sp_ctx;

/* physics: solver config keys sh / psv / spin (fd2d.py:116-118) */
#define SP_SH 1
#define SP_PSV 2
#define SP_SPIN 3

/* model fields of a grid (fd2d.py:172, 203-216) */
#define SP_F_RHO 0
#define SP_F_LAM 1
#define SP_F_MU 2
#define SP_F_NU 3
#define SP_F_J 4
#define SP_F_LAMC 5
#define SP_F_MUC 6
#define SP_F_NUC 7

/* sensitivity kernels of a grid.  SP_K_MU is the reference's k_mu
 * (solver.py:159-163); the others are this library's P-SV / spin
 * definitions.  SP_K_VP/VS/RHOP are the vp/vs/rho kernels derived from
 * lam/mu/rho by sp_kernels_vps. */
#define SP_K_LAM 0
#define SP_K_MU 1
#define SP_K_RHO 2
#define SP_K_NU 3
#define SP_K_J 4
#define SP_K_MUC 5
#define SP_K_VP 6
#define SP_K_VS 7
#define SP_K_RHOP 8
#define SP_K_HESS 9   /* pseudo-Hessian diagonal: sum over shots of the forward velocity energy */
#define SP_NKERN 10

/* tuning options (sp_set_option) */
#define SP_OPT_CHUNK 1      /* rows per wavefront item in the step kernel */
#define SP_OPT_ORDER 2      /* work order: 0 shot-outer (default), 1 shot-inner, 2 tile-major, 3 strip-group-major */
#define SP_OPT_GRAPH 3      /* 1: forward and adjoint runs replay blocks of time steps from captured
                               hipGraphs (forward runs with adjoint snapshots need two steps per
                               launch); checkpointed runs launch eagerly */
#define SP_OPT_SNAP_MB 4    /* snapshot memory budget in MiB (0 = auto) */
#define SP_OPT_KERNEL 5     /* step kernel: 1 buffer-addressed unrolled march (default), 0 plain */
#define SP_OPT_STRIP 6      /* strip geometry (shift << 16) | width: output lanes [shift, shift + width) */
#define SP_OPT_FUSE 7       /* 1 (default): shots advance two steps per launch wherever no snapshot,
                               kernel interaction or checkpoint needs the state between them; 2: the same
                               with two z cells per lane (SH and P-SV batches whose grids all have even nz;
                               others use 1); 3: two steps per launch in column groups, blocks whose
                               wavefronts march one span of a shot's z column together and pass the cells
                               at their edges through LDS, so halo cells are kept only at a span's ends
                               inside the column (SP_OPT_GROUP wavefronts per block); 4: column groups with
                               two z cells per lane (SH and P-SV with even nz, else as 2); 0: one step.  All
                               give the same results bit for bit; the autotuner picks among them */
#define SP_OPT_EXACT 8      /* 1 (default): the step kernels perform the reference's float64 operations one
                               by one in its order (correctly rounded quotients, no fused multiply-add), so
                               records equal the reference arithmetic bit for bit at any length; 0:
                               multiplications by reciprocals and fused multiply-adds (about 3-5% faster;
                               rounding-level differences that grow with the run length: 1.1e-5 relative
                               L2 on seismograms after 4000 steps at 1000^2, OUTSIDE the 1e-5 seismogram
                               tolerance, so runs that must meet it keep the default) */
#define SP_OPT_ADJOINT 9    /* sp_adjoint_ckpt runs 0 (default): the reference-style adjoint (the forward
                               operator reused with the adjoint sources injected at the receivers,
                               solver.py:290-321, kernels correlated every adjoint_interval steps); 1: the
                               discrete adjoint, the exact transpose of the forward time stepping with its
                               edge bands and sponge (SH, P-SV and spin, adjoint_interval 1, K = dt dL/dm exactly:
                               the gradient also beside a free surface and inside sponges, where the
                               reference-style kernels are not).  sp_gradient and sp_adjoint refuse 1. */
#define SP_OPT_GROUP 10     /* column groups (SP_OPT_FUSE 3, 4): wavefronts per block, 1..8 (0, the default:
                               4); a span covers 64 x that many x (1 or 2) cells */

/* Context and device.  Replaces fd2d.setup (fd2d.py:111-120). */
int sp_create(int device, sp_ctx **out);
int sp_destroy(sp_ctx *ctx);
const char *sp_last_error(const sp_ctx *ctx);
/* Work of the context goes to its own non-blocking stream until
 * sp_set_stream names another hipStream_t (NULL: the null stream).  The
 * new stream is ordered after the work already queued on the old one (an
 * event recorded there and waited for on the new stream), so a switch
 * needs no synchronisation by the caller. */
int sp_set_stream(sp_ctx *ctx, void *hip_stream);
int sp_sync(sp_ctx *ctx);
int sp_set_option(sp_ctx *ctx, int key, long long value);

/* Time axis and physics: solver keys nt, dt, sh/psv/spin (fd2d.py:114-118). */
int sp_set_physics(sp_ctx *ctx, int physics, int nt, double dt);

/* Model grid.  Replaces import_model's device part (fd2d.py:184-245):
 * param 0 takes fields {vp, vs, rho} and applies vps2lm (fd2d.py:21-33);
 * param 1 takes the reference's spin list {lambda, mu, nu, j, lambda_c,
 * mu_c, nu_c, rho} (fd2d.py:172); param 2 takes {lambda, mu, rho}. */
int sp_add_grid(sp_ctx *ctx, int nx, int nz, float dx, float dz, int param,
                const float *const *fields, int *grid_out);  /* fields: 3 pointers (param 0, 2) or 8 (param 1);
                                                                 on failure no grid is added */
/* set_bound: fd2d.py:47-68, 218-228 */
int sp_set_sponge(sp_ctx *ctx, int grid, int width, double alpha, int left, int right,
                  int bottom, int top);
/* update_model (fd2d.py:442-443) for any field, and read-back */
int sp_set_field(sp_ctx *ctx, int grid, int field, const float *src, int src_on_device);
int sp_get_field(sp_ctx *ctx, int grid, int field, float *dst, int dst_on_device);

/* One simulation ("shot") on a grid.  Replaces import_sources /
 * import_stations (fd2d.py:122-165) and the isrc selection of run_forward
 * (solver.py:180-183): the caller lists the sources that fire in this
 * shot.  src_k/rec_k are grid indices; stf is [ncomp_src][nsrc][nt]
 * (SH: y; P-SV and spin: x then z). */
int sp_add_shot(sp_ctx *ctx, int grid, int nsrc, const int *src_k, const float *stf, int nrec,
                const int *rec_k, int *shot_out);
int sp_clear_shots(sp_ctx *ctx);
int sp_num_components(const sp_ctx *ctx); /* records per receiver: SH 1, P-SV 2, spin 3 */

/* run_forward (solver.py:166-287) for every shot, batched.  Records the
 * displacement at the receivers.  sae > 0 also keeps the adjoint snapshots
 * (solver.py:193-203, 236-242) on the device for sp_adjoint. */
int sp_forward(sp_ctx *ctx, int sae);
/* The forward run of every shot in pieces, for field exports during the run
 * (save_snapshot, solver.py:244-261): steps [it0, it1), records included;
 * it0 = 0 starts a run (clears the wavefields), a later call continues
 * from where the previous one stopped. */
int sp_forward_steps(sp_ctx *ctx, int it0, int it1);
/* the current wavefield of a shot after sp_forward / sp_forward_steps:
 * SP_WF_V velocities [n_v][N] (SH vy; P-SV vx, vz; spin vx, vz, vy_c),
 * SP_WF_U displacements [n_v][N], SP_WF_S stresses [n_s][N] */
#define SP_WF_V 0
#define SP_WF_U 1
#define SP_WF_S 2
int sp_get_wavefield(sp_ctx *ctx, int shot, int kind, float *dst, int dst_on_device);
/* records of a shot: [ncomp][nrec][nt] (SH y; P-SV x, z; spin x, z, y) */
int sp_get_records(sp_ctx *ctx, int shot, float *dst, int dst_on_device);
/* observed data ("syn" in fd2d.py:303-309), same layout as the records */
int sp_set_data(sp_ctx *ctx, int shot, const float *src, int src_on_device);
/* misfit/waveform.py:5-34 for every shot and misfit component (SH y,
 * P-SV and spin x and z): writes the adjoint sources on the device and
 * the L2 norm of each tapered residual to out[shot*ncomp_misfit + c]. */
int sp_misfit(sp_ctx *ctx, double *norms_out);
/* run_adjoint + interaction (solver.py:290-321) for every shot, using the
 * snapshots of the last sp_forward(sae); kernels are cleared first. */
int sp_adjoint(sp_ctx *ctx, int sae);
/* run_kernel(adj=1) minus smoothing (solver.py:332-394): forward, misfit
 * and adjoint for every shot in snapshot-memory-sized sub-batches.
 * norms_out as sp_misfit.  compute_misfit (adj=0) is sp_forward(ctx, 0)
 * followed by sp_misfit. */
int sp_gradient(sp_ctx *ctx, int sae, double *norms_out);
/* Checkpointed forward and adjoint (no reference counterpart: the
 * reference keeps every adjoint snapshot in host memory, solver.py:193-203).
 * sp_forward_ckpt runs every shot, records, and keeps the full state every
 * C steps (C even).  sp_adjoint_ckpt runs the adjoint with the adjoint
 * sources currently set (sp_misfit or sp_set_adjoint_source) and
 * recomputes each C-step segment of the forward run from its checkpoint,
 * last segment first, for the snapshots of that segment: memory is
 * nt/C states plus C/sae snapshots per shot instead of nt/sae snapshots. */
int sp_forward_ckpt(sp_ctx *ctx, int C);
int sp_adjoint_ckpt(sp_ctx *ctx, int sae);
/* adjoint sources of a shot, [ncomp_misfit][nrec][nt], already time-reversed
 * (the layout sp_misfit writes, waveform.py:10,25) */
int sp_set_adjoint_source(sp_ctx *ctx, int shot, const float *src, int src_on_device);
/* a new model in vp/vs/rho, converted by vps2lm on the device (fd2d.py:21-33) */
int sp_set_model_vps(sp_ctx *ctx, int grid, const float *vp, const float *vs, const float *rho, int on_device);

/* kernels (SP_K_*) of a grid, [N] */
int sp_get_kernel(sp_ctx *ctx, int grid, int which, float *dst, int dst_on_device);
/* vp/vs/rho kernels from lam/mu/rho kernels (chain rule of vps2lm) */
int sp_kernels_vps(sp_ctx *ctx, int grid);
/* smooth (fd2d.py:71-108, 349-352): in place on a kernel of a grid */
int sp_smooth_kernel(sp_ctx *ctx, int grid, int which, int sigma);
/* the same smoothing on a caller buffer of nx*nz floats */
int sp_smooth(sp_ctx *ctx, float *field, int nx, int nz, int sigma, int on_device);

/* Allocation-free sp_smooth on a device buffer of nx*nz floats, in place.
 * No counterpart in the reference: it is for callers that cannot allocate
 * device memory (tight loops, harnesses that forbid hipMalloc).  It calls no
 * hipMalloc or hipFree; its weights and x-pass intermediate live in a
 * fixed device array of about 17 MB per GPU, held until the process ends,
 * which limits it to max(nx, nz) <= SP_NOALLOC_AXIS_MAX and nx*nz <=
 * SP_NOALLOC_MAX_CELLS (larger fields are refused; use sp_smooth).  Every
 * call on one GPU shares that array: calls that overlap on a GPU (other
 * threads or contexts) race; a call returns when its stream has finished.
 * Same results as sp_smooth(..., on_device = 1), bit for bit. */
#define SP_NOALLOC_AXIS_MAX 8192
#define SP_NOALLOC_MAX_CELLS 4194304LL
int sp_smooth_noalloc(sp_ctx *ctx, float *field_dev, int nx, int nz, int sigma);

/* Autotuner: times `steps` forward steps of the current shots for each
 * rows-per-wavefront chunk, work order and one or two steps per launch, in
 * strips or in column groups of 2, 4 or 8 wavefronts (SP_OPT_FUSE,
 * SP_OPT_GROUP); the chunks are a fixed list plus, for 2 to 6 wavefronts per
 * SIMD, the shortest chunk whose launch fits in that many.  The six fastest
 * are timed again over more steps (at least 3 x steps, about 25 ms), twice,
 * interleaved, and the one with the best time is kept for later runs and
 * reported; the choice is
 * remembered for the process and reused for the same physics and grids.
 * Overwrites wavefields and records (every run clears them first).  No
 * reference counterpart (the reference launches one thread per cell with
 * threads_per_block from the config, fd2d.py:185-187). */
int sp_autotune(sp_ctx *ctx, int steps, int *chunk_out, int *order_out);
/* current value of an option (SP_OPT_*), e.g. the autotuner's SP_OPT_FUSE choice */
int sp_get_option(const sp_ctx *ctx, int key, long long *value);

/* Timing of the last sp_forward / sp_adjoint / sp_gradient on the
 * context's stream: step-kernel launches and their total milliseconds
 * (HIP events), and the algorithmic bytes those launches move. */
int sp_last_timing(const sp_ctx *ctx, long long *step_launches, double *step_ms,
                   double *total_ms, double *alg_bytes);

#ifdef __cplusplus
}
#endif
#endif
