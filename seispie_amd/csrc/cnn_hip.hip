// cnn_hip.hip -- a small fully convolutional network (3x3 convolutions, ReLU
// between layers) with hand-written training and inference kernels for
// gfx950, used as the learned step conditioner of the FWI optimizers.
//
// No reference counterpart: icui/seispie's optimizers use the raw kernel
// as the gradient (seispie/optimize/lbfgs.py:23-59).
//
// Layout.  Activations are float32 [C][H][W] (W contiguous); for a model
// grid H = nx and W = nz, so a field of the solver (k = i*nz + j) is a
// one-channel activation as it is.  Parameters are one flat float32 array,
// per layer the weights [cin][3][3][cout] (cout contiguous) then the biases
// [cout]:
//     z_l[co][p] = b_l[co] + sum_ci sum_t W_l[ci][t][co] * a_{l-1}[ci][p + d_t]
//     a_l = relu(z_l) for every layer but the last, whose z is the output;
// d_t = (t / 3 - 1, t % 3 - 1), zero padding outside the grid.
//
// Kernels.
//   k_conv<COUT>  one output pixel per thread, all COUT outputs in registers;
//                 input tiles of 8 channels x (8+2) x (32+2) in LDS; the
//                 weights of one (ci, tap) are wave-uniform and come through
//                 scalar loads, so each LDS read feeds COUT FMAs.  Also the
//                 data gradient: a convolution of dz with W transposed and
//                 flipped (k_wflip), gated by the ReLU mask of the layer below.
//   k_wgrad       per-block partial sums of dW and db over an 8 x 32 tile,
//                 one (ci, co) pair per thread; k_reduce_blocks sums the
//                 partials in a fixed order (deterministic gradients).
//   k_mse, k_adam  loss gradient and partial sums; the Adam update.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <string>
#include <vector>

#include "../../include/seispie_cnn.h"

namespace {

constexpr int TY = 8, TX = 32;  // pixel tile of a block (256 threads)
constexpr int CCH = 8;           // input channels per LDS stage
typedef const float __attribute__((address_space(4))) *cflt_p;

template <int COUT>
__global__ __launch_bounds__(256) void k_conv(const float *__restrict__ in, int cin, const float *w,
                                              const float *b, float *__restrict__ out, int H, int W, int relu,
                                              const float *__restrict__ gate)
{
    __shared__ float tile[CCH][TY + 2][TX + 2];
    const int tx = threadIdx.x & (TX - 1), ty = threadIdx.x / TX;
    const int x0 = blockIdx.x * TX, y0 = blockIdx.y * TY;
    const int x = x0 + tx, y = y0 + ty;
    const long long HW = (long long)H * W;
    const cflt_p wc = (cflt_p)w;
    float acc[COUT];
#pragma unroll
    for (int co = 0; co < COUT; co++) acc[co] = b ? ((cflt_p)b)[co] : 0.f;
    for (int c0 = 0; c0 < cin; c0 += CCH) {
        for (int idx = threadIdx.x; idx < CCH * (TY + 2) * (TX + 2); idx += 256) {
            const int c = idx / ((TY + 2) * (TX + 2)), r = (idx / (TX + 2)) % (TY + 2), cc = idx % (TX + 2);
            const int gy = y0 - 1 + r, gx = x0 - 1 + cc;
            float v = 0.f;
            if (c0 + c < cin && gy >= 0 && gy < H && gx >= 0 && gx < W) v = in[(c0 + c) * HW + (long long)gy * W + gx];
            tile[c][r][cc] = v;
        }
        __syncthreads();
        const int cn = min(CCH, cin - c0);
        for (int c = 0; c < cn; c++) {
#pragma unroll
            for (int t = 0; t < 9; t++) {
                const float v = tile[c][ty + t / 3][tx + t % 3];
                const cflt_p wp = wc + ((long long)(c0 + c) * 9 + t) * COUT;
#pragma unroll
                for (int co = 0; co < COUT; co++) acc[co] = fmaf(wp[co], v, acc[co]);
            }
        }
        __syncthreads();
    }
    if (y >= H || x >= W) return;
    const long long p = (long long)y * W + x;
#pragma unroll
    for (int co = 0; co < COUT; co++) {
        float v = acc[co];
        if (relu) v = fmaxf(v, 0.f);
        if (gate) v = gate[co * HW + p] > 0.f ? v : 0.f;
        out[co * HW + p] = v;
    }
}

// W'[co][t][ci] = W[ci][8 - t][co]: the data gradient as a forward convolution
__global__ void k_wflip(const float *w, float *wt, int cin, int cout)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= cin * 9 * cout) return;
    const int co = i % cout, t = (i / cout) % 9, ci = i / (9 * cout);
    wt[((long long)co * 9 + (8 - t)) * cin + ci] = w[i];
}

// Partial weight and bias gradients of one 8 x 32 pixel tile:
//   part[block][ci][t][co] = sum_p dz[co][p] * a[ci][p + d_t],  part[block][cin*9*cout + co] = sum_p dz[co][p].
// Thread (cl, co) with cl = threadIdx.x / 32, co = threadIdx.x % 32 covers
// input channel c0 + cl of the current stage; cout <= 32.
__global__ __launch_bounds__(256) void k_wgrad(const float *__restrict__ a, int cin, const float *__restrict__ dz,
                                               int cout, int H, int W, float *__restrict__ part)
{
    __shared__ float ta[CCH][TY + 2][TX + 2];
    __shared__ float tz[32][TY * TX + 1];  // +1: the 32 output channels fall in different banks
    const int x0 = blockIdx.x * TX, y0 = blockIdx.y * TY;
    const long long HW = (long long)H * W;
    const long long blk = (long long)blockIdx.y * gridDim.x + blockIdx.x;
    const long long np = (long long)cin * 9 * cout + cout;
    float *pp = part + blk * np;
    for (int idx = threadIdx.x; idx < 32 * TY * TX; idx += 256) {
        const int co = idx / (TY * TX), q = idx % (TY * TX);
        const int gy = y0 + q / TX, gx = x0 + q % TX;
        float v = 0.f;
        if (co < cout && gy < H && gx < W) v = dz[co * HW + (long long)gy * W + gx];
        tz[co][q] = v;
    }
    const int cl = threadIdx.x / 32, co = threadIdx.x % 32;
    __syncthreads();
    if (cl == 0 && co < cout) {
        float s = 0.f;
        for (int q = 0; q < TY * TX; q++) s += tz[co][q];
        pp[(long long)cin * 9 * cout + co] = s;
    }
    for (int c0 = 0; c0 < cin; c0 += CCH) {
        for (int idx = threadIdx.x; idx < CCH * (TY + 2) * (TX + 2); idx += 256) {
            const int c = idx / ((TY + 2) * (TX + 2)), r = (idx / (TX + 2)) % (TY + 2), cc = idx % (TX + 2);
            const int gy = y0 - 1 + r, gx = x0 - 1 + cc;
            float v = 0.f;
            if (c0 + c < cin && gy >= 0 && gy < H && gx >= 0 && gx < W) v = a[(c0 + c) * HW + (long long)gy * W + gx];
            ta[c][r][cc] = v;
        }
        __syncthreads();
        if (c0 + cl < cin && co < cout) {
            float acc[9];
#pragma unroll
            for (int t = 0; t < 9; t++) acc[t] = 0.f;
            for (int q = 0; q < TY * TX; q++) {
                const float g = tz[co][q];
                const int qy = q / TX, qx = q % TX;
#pragma unroll
                for (int t = 0; t < 9; t++) acc[t] = fmaf(g, ta[cl][qy + t / 3][qx + t % 3], acc[t]);
            }
#pragma unroll
            for (int t = 0; t < 9; t++) pp[((long long)(c0 + cl) * 9 + t) * cout + co] = acc[t];
        }
        __syncthreads();
    }
}

// out[i] = sum over blocks of part[b][i], blocks in order
__global__ void k_reduce_blocks(const float *part, long long nblk, long long n, float *out)
{
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float s = 0.f;
    for (long long b = 0; b < nblk; b++) s += part[b * n + i];
    out[i] = s;
}

// loss = sum m (y - t)^2 / sum m;  dy = 2 m (y - t) / sum m;  per-block partial sums of m (y-t)^2
__global__ __launch_bounds__(256) void k_mse(const float *y, const float *t, const float *m, long long n, double inv,
                                             float *dy, double *part)
{
    __shared__ double red[256];
    const long long i = (long long)blockIdx.x * 256 + threadIdx.x;
    double e = 0.0;
    if (i < n) {
        const float w = m ? m[i] : 1.f;
        const float d = y[i] - t[i];
        dy[i] = (float)(2.0 * inv * (double)w * (double)d);
        e = (double)w * (double)d * (double)d;
    }
    red[threadIdx.x] = e;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

__global__ void k_adam(float *p, const float *g, float *m, float *v, long long n, float lr, float b1, float b2, float eps,
                       float c1, float c2)
{
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.f - b1) * gi;
    const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    p[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

}  // namespace

struct sp_cnn {
    int device = 0;
    hipStream_t stream = nullptr;
    std::vector<int> ch;          // channels: ch[0] inputs ... ch[L] outputs
    std::vector<long long> off;   // parameter offset of each layer's weights
    long long nparam = 0;
    float *params = nullptr, *grad = nullptr, *am = nullptr, *av = nullptr;
    long long step = 0;
    // activations of the last training forward pass: act[0] = input, act[l] = layer l output
    std::vector<float *> act;
    long long act_hw = 0;
    float *dz0 = nullptr, *dz1 = nullptr, *wt = nullptr, *part = nullptr, *tgt = nullptr;
    long long part_cap = 0;
    double *lpart = nullptr;
    long long lpart_cap = 0;
    std::string err;
};

namespace {

#define CK(net, call)                                                                            \
    do {                                                                                         \
        hipError_t e_ = (call);                                                                  \
        if (e_ != hipSuccess) {                                                                  \
            (net)->err = std::string(#call) + ": " + hipGetErrorString(e_);                       \
            return 2;                                                                            \
        }                                                                                        \
    } while (0)

int nfail(sp_cnn *n, const std::string &m)
{
    n->err = m;
    return 1;
}

int launch_conv(sp_cnn *n, const float *in, int cin, const float *w, const float *b, float *out, int cout, int H,
                int W, int relu, const float *gate)
{
    dim3 g((unsigned)((W + TX - 1) / TX), (unsigned)((H + TY - 1) / TY)), bl(256);
    switch (cout) {
    case 1: hipLaunchKernelGGL(k_conv<1>, g, bl, 0, n->stream, in, cin, w, b, out, H, W, relu, gate); break;
    case 2: hipLaunchKernelGGL(k_conv<2>, g, bl, 0, n->stream, in, cin, w, b, out, H, W, relu, gate); break;
    case 4: hipLaunchKernelGGL(k_conv<4>, g, bl, 0, n->stream, in, cin, w, b, out, H, W, relu, gate); break;
    case 8: hipLaunchKernelGGL(k_conv<8>, g, bl, 0, n->stream, in, cin, w, b, out, H, W, relu, gate); break;
    case 16: hipLaunchKernelGGL(k_conv<16>, g, bl, 0, n->stream, in, cin, w, b, out, H, W, relu, gate); break;
    case 32: hipLaunchKernelGGL(k_conv<32>, g, bl, 0, n->stream, in, cin, w, b, out, H, W, relu, gate); break;
    default: return nfail(n, "channel counts must be 1, 2, 4, 8, 16 or 32");
    }
    CK(n, hipGetLastError());
    return 0;
}

int ensure_acts(sp_cnn *n, long long hw)
{
    if (hw <= n->act_hw) return 0;
    for (float *p : n->act)
        if (p) CK(n, hipFree(p));
    for (float **p : {&n->dz0, &n->dz1, &n->tgt})
        if (*p) CK(n, hipFree(*p));
    n->act.assign(n->ch.size(), nullptr);
    n->dz0 = n->dz1 = n->tgt = nullptr;
    n->act_hw = 0;
    int cmax = 1;
    for (size_t l = 0; l < n->ch.size(); l++) {
        CK(n, hipMalloc(&n->act[l], (size_t)n->ch[l] * hw * sizeof(float)));
        cmax = std::max(cmax, n->ch[l]);
    }
    CK(n, hipMalloc(&n->dz0, (size_t)cmax * hw * sizeof(float)));
    CK(n, hipMalloc(&n->dz1, (size_t)cmax * hw * sizeof(float)));
    CK(n, hipMalloc(&n->tgt, (size_t)2 * hw * sizeof(float)));  // target and mask
    n->act_hw = hw;
    return 0;
}

int copy_in(sp_cnn *n, void *dst, const void *src, size_t bytes, int on_dev)
{
    CK(n, hipMemcpyAsync(dst, src, bytes, on_dev ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice, n->stream));
    if (!on_dev) CK(n, hipStreamSynchronize(n->stream));
    return 0;
}

int copy_out(sp_cnn *n, void *dst, const void *src, size_t bytes, int on_dev)
{
    CK(n, hipMemcpyAsync(dst, src, bytes, on_dev ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost, n->stream));
    if (!on_dev) CK(n, hipStreamSynchronize(n->stream));
    return 0;
}

// forward pass from act[0], keeping every activation
int forward_all(sp_cnn *n, int H, int W)
{
    const int L = (int)n->ch.size() - 1;
    for (int l = 1; l <= L; l++) {
        const float *w = n->params + n->off[l - 1];
        const float *b = w + (long long)n->ch[l - 1] * 9 * n->ch[l];
        if (launch_conv(n, n->act[l - 1], n->ch[l - 1], w, b, n->act[l], n->ch[l], H, W, l < L, nullptr)) return 2;
    }
    return 0;
}

}  // namespace

extern "C" {

int sp_cnn_create(int device, int nlayers, const int *channels, sp_cnn **out)
{
    if (!out || nlayers < 1 || !channels) return 1;
    *out = nullptr;
    for (int l = 0; l <= nlayers; l++)
        if (channels[l] < 1 || channels[l] > 32 || (l > 0 && (channels[l] & (channels[l] - 1)))) return 1;
    sp_cnn *n = new sp_cnn();
    n->device = device;
    if (hipSetDevice(device) != hipSuccess) {
        delete n;
        return 2;
    }
    n->ch.assign(channels, channels + nlayers + 1);
    for (int l = 1; l <= nlayers; l++) {
        n->off.push_back(n->nparam);
        n->nparam += (long long)n->ch[l - 1] * 9 * n->ch[l] + n->ch[l];
    }
    const size_t pb = (size_t)n->nparam * sizeof(float);
    if (hipMalloc(&n->params, pb) != hipSuccess || hipMalloc(&n->grad, pb) != hipSuccess ||
        hipMalloc(&n->am, pb) != hipSuccess || hipMalloc(&n->av, pb) != hipSuccess || hipMalloc(&n->wt, pb) != hipSuccess ||
        hipMemset(n->params, 0, pb) != hipSuccess || hipMemset(n->am, 0, pb) != hipSuccess ||
        hipMemset(n->av, 0, pb) != hipSuccess || hipMemset(n->grad, 0, pb) != hipSuccess) {
        delete n;
        return 2;
    }
    *out = n;
    return 0;
}

int sp_cnn_destroy(sp_cnn *n)
{
    if (!n) return 3;
    (void)hipSetDevice(n->device);
    (void)hipStreamSynchronize(n->stream);
    for (float *p : {n->params, n->grad, n->am, n->av, n->wt, n->dz0, n->dz1, n->tgt, n->part})
        if (p) (void)hipFree(p);
    for (float *p : n->act)
        if (p) (void)hipFree(p);
    if (n->lpart) (void)hipFree(n->lpart);
    delete n;
    return 0;
}

const char *sp_cnn_last_error(const sp_cnn *n) { return n ? n->err.c_str() : "null network"; }

int sp_cnn_set_stream(sp_cnn *n, void *stream)
{
    if (!n) return 3;
    n->stream = (hipStream_t)stream;
    return 0;
}

long long sp_cnn_num_params(const sp_cnn *n) { return n ? n->nparam : -1; }

int sp_cnn_set_params(sp_cnn *n, const float *params, int on_dev)
{
    if (!n) return 3;
    CK(n, hipSetDevice(n->device));
    if (copy_in(n, n->params, params, (size_t)n->nparam * sizeof(float), on_dev)) return 2;
    // a new parameter vector starts a new Adam run
    CK(n, hipMemsetAsync(n->am, 0, (size_t)n->nparam * sizeof(float), n->stream));
    CK(n, hipMemsetAsync(n->av, 0, (size_t)n->nparam * sizeof(float), n->stream));
    n->step = 0;
    return 0;
}

int sp_cnn_get_params(sp_cnn *n, float *params, int on_dev)
{
    if (!n) return 3;
    CK(n, hipSetDevice(n->device));
    return copy_out(n, params, n->params, (size_t)n->nparam * sizeof(float), on_dev);
}

int sp_cnn_get_grad(sp_cnn *n, float *grad, int on_dev)
{
    if (!n) return 3;
    CK(n, hipSetDevice(n->device));
    return copy_out(n, grad, n->grad, (size_t)n->nparam * sizeof(float), on_dev);
}

int sp_cnn_forward(sp_cnn *n, const float *x, float *y, int H, int W, int on_dev)
{
    if (!n) return 3;
    if (H < 1 || W < 1) return nfail(n, "empty input");
    CK(n, hipSetDevice(n->device));
    const long long hw = (long long)H * W;
    if (ensure_acts(n, hw)) return 2;
    if (copy_in(n, n->act[0], x, (size_t)n->ch[0] * hw * sizeof(float), on_dev)) return 2;
    if (forward_all(n, H, W)) return 2;
    const int L = (int)n->ch.size() - 1;
    return copy_out(n, y, n->act[L], (size_t)n->ch[L] * hw * sizeof(float), on_dev);
}

int sp_cnn_loss_grad(sp_cnn *n, const float *x, const float *target, const float *mask, int H, int W, int on_dev,
                     double *loss)
{
    if (!n) return 3;
    if (H < 1 || W < 1) return nfail(n, "empty input");
    CK(n, hipSetDevice(n->device));
    const int L = (int)n->ch.size() - 1;
    if (n->ch[L] != 1) return nfail(n, "the loss needs a one-channel output");
    const long long hw = (long long)H * W;
    if (ensure_acts(n, hw)) return 2;
    if (copy_in(n, n->act[0], x, (size_t)n->ch[0] * hw * sizeof(float), on_dev)) return 2;
    float *t = n->tgt, *m = mask ? n->tgt + hw : nullptr;
    if (copy_in(n, t, target, (size_t)hw * sizeof(float), on_dev)) return 2;
    double msum = (double)hw;
    if (mask) {
        if (copy_in(n, m, mask, (size_t)hw * sizeof(float), on_dev)) return 2;
        std::vector<float> hm;
        const float *mh = mask;
        if (on_dev) {
            hm.resize(hw);
            if (copy_out(n, hm.data(), m, (size_t)hw * sizeof(float), 0)) return 2;
            mh = hm.data();
        }
        msum = 0;
        for (long long i = 0; i < hw; i++) msum += mh[i];
    }
    if (forward_all(n, H, W)) return 2;
    const long long nb = (hw + 255) / 256;
    if (nb > n->lpart_cap) {
        if (n->lpart) CK(n, hipFree(n->lpart));
        n->lpart = nullptr;
        n->lpart_cap = 0;
        CK(n, hipMalloc(&n->lpart, (size_t)nb * sizeof(double)));
        n->lpart_cap = nb;
    }
    hipLaunchKernelGGL(k_mse, dim3((unsigned)nb), dim3(256), 0, n->stream, n->act[L], t, m, hw,
                       msum > 0 ? 1.0 / msum : 0.0, n->dz0, n->lpart);
    CK(n, hipGetLastError());
    if (loss) {
        std::vector<double> hp(nb);
        if (copy_out(n, hp.data(), n->lpart, (size_t)nb * sizeof(double), 0)) return 2;
        double s = 0;
        for (double v : hp) s += v;
        *loss = msum > 0 ? s / msum : 0.0;
    }
    // backward: dz0 holds dL/dz_L
    const dim3 tiles((unsigned)((W + TX - 1) / TX), (unsigned)((H + TY - 1) / TY));
    const long long nblk = (long long)tiles.x * tiles.y;
    long long need = 0;
    for (int l = 1; l <= L; l++) need = std::max(need, nblk * ((long long)n->ch[l - 1] * 9 * n->ch[l] + n->ch[l]));
    if (need > n->part_cap) {
        if (n->part) CK(n, hipFree(n->part));
        n->part = nullptr;
        n->part_cap = 0;
        CK(n, hipMalloc(&n->part, (size_t)need * sizeof(float)));
        n->part_cap = need;
    }
    float *dz = n->dz0, *dzn = n->dz1;
    for (int l = L; l >= 1; l--) {
        const int ci = n->ch[l - 1], co = n->ch[l];
        const long long np = (long long)ci * 9 * co + co;
        hipLaunchKernelGGL(k_wgrad, tiles, dim3(256), 0, n->stream, n->act[l - 1], ci, dz, co, H, W, n->part);
        CK(n, hipGetLastError());
        hipLaunchKernelGGL(k_reduce_blocks, dim3((unsigned)((np + 255) / 256)), dim3(256), 0, n->stream, n->part, nblk,
                           np, n->grad + n->off[l - 1]);
        CK(n, hipGetLastError());
        if (l > 1) {
            const long long nw = (long long)ci * 9 * co;
            hipLaunchKernelGGL(k_wflip, dim3((unsigned)((nw + 255) / 256)), dim3(256), 0, n->stream,
                               n->params + n->off[l - 1], n->wt, ci, co);
            CK(n, hipGetLastError());
            // dL/dz_{l-1} = (W_l^T * dz_l) gated by relu'(z_{l-1}) = [a_{l-1} > 0]
            if (launch_conv(n, dz, co, n->wt, nullptr, dzn, ci, H, W, 0, n->act[l - 1])) return 2;
            std::swap(dz, dzn);
        }
    }
    return 0;
}

int sp_cnn_adam_step(sp_cnn *n, float lr, float beta1, float beta2, float eps)
{
    if (!n) return 3;
    CK(n, hipSetDevice(n->device));
    n->step++;
    const float c1 = 1.f - std::pow(beta1, (float)n->step), c2 = 1.f - std::pow(beta2, (float)n->step);
    hipLaunchKernelGGL(k_adam, dim3((unsigned)((n->nparam + 255) / 256)), dim3(256), 0, n->stream, n->params, n->grad,
                       n->am, n->av, n->nparam, lr, beta1, beta2, eps, c1, c2);
    CK(n, hipGetLastError());
    return 0;
}

int sp_cnn_sync(sp_cnn *n)
{
    if (!n) return 3;
    CK(n, hipSetDevice(n->device));
    CK(n, hipStreamSynchronize(n->stream));
    return 0;
}

}  // extern "C"
