// xc_unet.hip -- the learned coarse-correction U-Net: BF16 MFMA forward and
// backward kernels with FP32 accumulation, Adam in FP32 (gfx950).
//
// No counterpart in the reference: this is the learned Stokes preconditioner.
// Network (4 levels, channels 16/32/48/64, ~226k parameters): two bias-free
// 3x3 convolutions + ReLU per level, 2x2 average pooling down, nearest
// upsampling and channel concatenation [up, skip] up, a 1x1 head to 3 outputs.
// Boundaries: x periodic, z zero-padded.  Inputs: 4 channels (log-viscosity,
// r_W, r_U, r_P, cell centred); outputs: 3 channels (e_W, e_U, e_P).
//
// Layouts.  Activations are pixel-major with channels fastest, pixel
// p = (b * W + x) * H + z (z fastest, as the fields), BF16 storage.  3x3
// weights are [cout][9][cin] with tap t = 3 (dx + 1) + (dz + 1); the flat FP32
// parameter vector keeps the layers in order e0a e0b e1a e1b e2a e2b e3a e3b
// d2a d2b d1a d1b d0a d0b head, with e0a at cin = 4 (the device pads it to 16).
//
// Convolution = implicit GEMM on gfx950's v_mfma_f32_16x16x32_bf16: one
// wavefront owns 16 output pixels and every output channel (NB blocks of 16), K
// runs over the 9 taps x cin in steps of 32 (a last step of 16 channels, when
// cin is an odd multiple of 16, on v_mfma_f32_16x16x16_bf16).  A operand: lane l
// holds 8 input channels (k = 8 (l / 16) .. +7) of output pixel (l % 16)'s
// neighbour at the tap, one 16-byte load; B operand: lane l holds
// w[co = l % 16][t][k .. k+7], one 16-byte load (K is a sum: any assignment of
// the k slots that A and B share gives the same product); D: lane l holds
// channel l % 16 of pixels 4 (l / 16) .. +3.  The backward-data pass is the same
// kernel on transposed, tap-flipped weights; the weight gradient is a GEMM over
// pixels (32 per MFMA) accumulated with FP32 atomics.
#include <math.h>
#include <string.h>

#include <vector>

#include "xc_internal.h"

namespace {

typedef unsigned short bf16_t;
typedef short s4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf8 __attribute__((ext_vector_type(8)));
typedef unsigned short u8v __attribute__((ext_vector_type(8)));
typedef float f4 __attribute__((ext_vector_type(4)));

constexpr int NT = 256;
constexpr int NLAYER = 15;
constexpr int CH0 = 16, CH1 = 32, CH2 = 48, CH3 = 64;
constexpr int CIN = 4, COUT = 3;

__device__ __forceinline__ float bf2f(bf16_t h) { return __uint_as_float((unsigned)h << 16); }
__device__ __forceinline__ bf16_t f2bf(float f) {  // round to nearest even
    unsigned u = __float_as_uint(f);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (bf16_t)(u >> 16);
}

// ---------------------------------------------------------------- 3x3 convolution
struct ConvArgs {
    const bf16_t* in;
    int in_stride, cin;        // cin: multiple of 16
    const bf16_t* w;           // [cout][9][cin]
    bf16_t* out;
    int out_stride, out_off;
    const bf16_t* mask;        // mode 1: out = y * (mask > 0)
    int mask_stride, mask_off;
    int mode;                  // 0: ReLU, 1: masked, 2: plain
    int B, H, W;
    int64_t P;                 // B * H * W
};

template <int NB>
__global__ __launch_bounds__(NT) void k_conv3(ConvArgs a) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int64_t p0 = ((int64_t)blockIdx.x * 4 + wave) * 16;
    if (p0 >= a.P) return;  // whole wavefront (uniform)
    // A-operand pixel of this lane and its coordinates
    const int64_t pa = p0 + (lane & 15);
    const bool va = pa < a.P;
    const int z = (int)(pa % a.H);
    const int64_t col = pa / a.H;                 // b * W + x
    const int x = (int)(col % a.W);
    const int64_t bW = col - x;                   // b * W
    const int kq = 4 * (lane >> 4), kq8 = 8 * (lane >> 4);
    const int c32 = a.cin & ~31;  // channels taken 32 at a time; a remaining 16 take the 16-deep MFMA
    f4 acc[NB];
#pragma unroll
    for (int n = 0; n < NB; ++n) acc[n] = f4{0.f, 0.f, 0.f, 0.f};
    for (int t = 0; t < 9; ++t) {
        const int dx = t / 3 - 1, dz = t % 3 - 1;
        const int zz = z + dz;
        int xx = x + dx;
        xx = xx < 0 ? xx + a.W : (xx >= a.W ? xx - a.W : xx);
        const bool ok = va && zz >= 0 && zz < a.H;
        const int64_t q = ((bW + xx) * a.H + zz) * a.in_stride;
        for (int c0 = 0; c0 < c32; c0 += 32) {
            u8v ar = u8v{0, 0, 0, 0, 0, 0, 0, 0};
            if (ok) ar = *reinterpret_cast<const u8v*>(a.in + q + c0 + kq8);
            const bf8 av = __builtin_bit_cast(bf8, ar);
#pragma unroll
            for (int n = 0; n < NB; ++n) {
                const int co = n * 16 + (lane & 15);
                const bf8 bv = __builtin_bit_cast(bf8, *reinterpret_cast<const u8v*>(a.w + ((int64_t)co * 9 + t) * a.cin + c0 + kq8));
                acc[n] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av, bv, acc[n], 0, 0, 0);
            }
        }
        if (c32 < a.cin) {
            s4 av = s4{0, 0, 0, 0};
            if (ok) av = *reinterpret_cast<const s4*>(a.in + q + c32 + kq);
#pragma unroll
            for (int n = 0; n < NB; ++n) {
                const int co = n * 16 + (lane & 15);
                const s4 bv = *reinterpret_cast<const s4*>(a.w + ((int64_t)co * 9 + t) * a.cin + c32 + kq);
                acc[n] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(av, bv, acc[n], 0, 0, 0);
            }
        }
    }
#pragma unroll
    for (int n = 0; n < NB; ++n) {
        const int co = n * 16 + (lane & 15);
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int64_t p = p0 + kq + r;
            if (p >= a.P) continue;
            float v = acc[n][r];
            if (a.mode == 0) v = v > 0.f ? v : 0.f;
            else if (a.mode == 1) v = bf2f(a.mask[p * a.mask_stride + a.mask_off + co]) > 0.f ? v : 0.f;
            a.out[p * a.out_stride + a.out_off + co] = f2bf(v);
        }
    }
}

// weight gradient: gw[co][t][ci] += sum_p gy[p][co] x[q_t(p)][ci]   (FP32 atomics)
struct WGradArgs {
    const bf16_t* x;
    int x_stride, cin;
    const bf16_t* gy;
    int gy_stride, cout;
    float* gw;                 // [cout][9][cin_w] with cin_w >= the channels used here
    int cin_w;
    int B, H, W;
    int64_t P;
    int chunk;                 // pixels per wavefront (multiple of 32)
};
__global__ __launch_bounds__(NT) void k_conv3_wgrad(WGradArgs a) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    // tile: (t, co block, ci block) from blockIdx.y; pixel chunk from blockIdx.x * 4 + wave
    const int nco = a.cout / 16, nci = a.cin / 16;
    const int tile = blockIdx.y;
    const int t = tile / (nco * nci);
    const int rem = tile - t * nco * nci;
    const int cb = rem / nci, ib = rem - cb * nci;
    const int dx = t / 3 - 1, dz = t % 3 - 1;
    const int64_t pbeg = ((int64_t)blockIdx.x * 4 + wave) * a.chunk;
    if (pbeg >= a.P) return;
    const int64_t pend = pbeg + a.chunk < a.P ? pbeg + a.chunk : a.P;
    const int co = cb * 16 + (lane & 15);   // A row
    const int ci = ib * 16 + (lane & 15);   // B column
    const int kq8 = 8 * (lane >> 4);
    f4 acc = f4{0.f, 0.f, 0.f, 0.f};
    for (int64_t pk = pbeg; pk < pend; pk += 32) {  // K = 32 pixels a step, 8 a lane
        u8v ar, br;
#pragma unroll
        for (int i = 0; i < 8; ++i) {
            const int64_t p = pk + kq8 + i;
            unsigned short ga = 0, xb = 0;
            if (p < pend) {
                ga = a.gy[p * a.gy_stride + co];
                const int z = (int)(p % a.H);
                const int64_t col = p / a.H;
                const int x = (int)(col % a.W);
                const int zz = z + dz;
                int xx = x + dx;
                xx = xx < 0 ? xx + a.W : (xx >= a.W ? xx - a.W : xx);
                if (zz >= 0 && zz < a.H) xb = a.x[((col - x + xx) * a.H + zz) * a.x_stride + ci];
            }
            ar[i] = ga;
            br[i] = xb;
        }
        acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf8, ar), __builtin_bit_cast(bf8, br), acc, 0, 0, 0);
    }
    // D: lane -> column ci = ib*16 + lane%16, rows co = cb*16 + 4 (lane/16) + r
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        const int cor = cb * 16 + 4 * (lane >> 4) + r;
        const int cic = ib * 16 + (lane & 15);
        if (cic < a.cin_w) atomicAdd(&a.gw[((int64_t)cor * 9 + t) * a.cin_w + cic], acc[r]);
    }
}

// ---------------------------------------------------------------- elementwise
// input FP32 [p][4] -> BF16 [p][16] (channels 4..15 zero)
__global__ __launch_bounds__(NT) void k_input(const float* __restrict__ in, bf16_t* __restrict__ out, int64_t P) {
    const int64_t k = (int64_t)blockIdx.x * NT + threadIdx.x;
    if (k >= P * 16) return;
    const int64_t p = k >> 4;
    const int c = (int)(k & 15);
    out[k] = c < CIN ? f2bf(in[p * CIN + c]) : (bf16_t)0;
}
// 2x2 average pooling (fine H x W -> H/2 x W/2), C channels
__global__ __launch_bounds__(NT) void k_pool(const bf16_t* __restrict__ in, int in_stride, int in_off,
                                             bf16_t* __restrict__ out, int out_stride, int C, int B, int Hc, int Wc) {
    const int64_t k = (int64_t)blockIdx.x * NT + threadIdx.x;
    const int64_t n = (int64_t)B * Hc * Wc * C;
    if (k >= n) return;
    const int c = (int)(k % C);
    const int64_t p = k / C;
    const int Z = (int)(p % Hc);
    const int64_t col = p / Hc;
    const int X = (int)(col % Wc);
    const int64_t b = col / Wc;
    const int Hf = 2 * Hc, Wf = 2 * Wc;
    auto at = [&](int z, int x) { return bf2f(in[((b * Wf + x) * Hf + z) * in_stride + in_off + c]); };
    const float v = 0.25f * (at(2 * Z, 2 * X) + at(2 * Z + 1, 2 * X) + at(2 * Z, 2 * X + 1) + at(2 * Z + 1, 2 * X + 1));
    out[p * out_stride + c] = f2bf(v);
}
// nearest upsampling (coarse -> fine) into channels [out_off, out_off + C) of out
__global__ __launch_bounds__(NT) void k_up(const bf16_t* __restrict__ in, int in_stride, bf16_t* __restrict__ out,
                                           int out_stride, int out_off, int C, int B, int Hf, int Wf) {
    const int64_t k = (int64_t)blockIdx.x * NT + threadIdx.x;
    const int64_t n = (int64_t)B * Hf * Wf * C;
    if (k >= n) return;
    const int c = (int)(k % C);
    const int64_t p = k / C;
    const int z = (int)(p % Hf);
    const int64_t col = p / Hf;
    const int x = (int)(col % Wf);
    const int64_t b = col / Wf;
    const int Hc = Hf / 2, Wc = Wf / 2;
    out[p * out_stride + out_off + c] = in[((b * Wc + x / 2) * Hc + z / 2) * in_stride + c];
}
// gradient of a level output e (= ReLU output), masked: g = [sum of the 4 fine cells of gup (coarse
// result of the upsampling backward) if gup] + [0.25 * gpool(coarse) if gpool] + [gskip if gskip], * (e > 0)
struct GradMergeArgs {
    const bf16_t* gup;   int gup_stride, gup_off;    // fine-level gradient of an upsampled tensor (for coarse targets)
    const bf16_t* gpool; int gpool_stride;           // coarse gradient of pool(e)
    const bf16_t* gskip; int gskip_stride, gskip_off; // same-level gradient through the skip concat
    const bf16_t* e;     int e_stride, e_off;        // level output (mask)
    bf16_t* out;                                     // [p][C]
    int C, B, H, W;
};
__global__ __launch_bounds__(NT) void k_grad_merge(GradMergeArgs a) {
    const int64_t k = (int64_t)blockIdx.x * NT + threadIdx.x;
    const int64_t n = (int64_t)a.B * a.H * a.W * a.C;
    if (k >= n) return;
    const int c = (int)(k % a.C);
    const int64_t p = k / a.C;
    const int z = (int)(p % a.H);
    const int64_t col = p / a.H;
    const int x = (int)(col % a.W);
    const int64_t b = col / a.W;
    float g = 0.f;
    if (a.gup) {  // this level is the coarse side of an upsampling: sum its 4 fine children
        const int Hf = 2 * a.H, Wf = 2 * a.W;
        for (int i = 0; i < 2; ++i)
            for (int j = 0; j < 2; ++j)
                g += bf2f(a.gup[((b * Wf + 2 * x + j) * Hf + 2 * z + i) * a.gup_stride + a.gup_off + c]);
    }
    if (a.gpool) {  // this level is the fine side of a pooling
        const int Hc = a.H / 2, Wc = a.W / 2;
        g += 0.25f * bf2f(a.gpool[((b * Wc + x / 2) * Hc + z / 2) * a.gpool_stride + c]);
    }
    if (a.gskip) g += bf2f(a.gskip[p * a.gskip_stride + a.gskip_off + c]);
    if (!(bf2f(a.e[p * a.e_stride + a.e_off + c]) > 0.f)) g = 0.f;
    a.out[p * a.C + c] = f2bf(g);
}

// ---------------------------------------------------------------- head and loss
// out[p][o] = sum_c d0[p][c] wh[o][c] (FP32); training: loss partials, gout, g_d0 (masked), gwh (atomics)
struct HeadArgs {
    const bf16_t* d0;
    const float* wh;       // [3][16] FP32
    float* out;            // [p][3]
    const float* tgt;      // [p][3] or null (inference)
    float inv_n;           // 1 / (B H W)
    bf16_t* g_d0;          // [p][16]
    float* gwh;            // [3][16]
    double* loss_part;     // per block
    int64_t P;
};
__global__ __launch_bounds__(NT) void k_head(HeadArgs a) {
    __shared__ float sg[COUT * CH0];
    __shared__ double sl[NT];
    const int64_t p = (int64_t)blockIdx.x * NT + threadIdx.x;
    if (a.tgt)
        for (int i = threadIdx.x; i < COUT * CH0; i += NT) sg[i] = 0.f;
    __syncthreads();
    double l = 0.0;
    if (p < a.P) {
        float d[CH0];
#pragma unroll
        for (int c = 0; c < CH0; ++c) d[c] = bf2f(a.d0[p * CH0 + c]);
        float go[COUT];
#pragma unroll
        for (int o = 0; o < COUT; ++o) {
            float s = 0.f;
#pragma unroll
            for (int c = 0; c < CH0; ++c) s += d[c] * a.wh[o * CH0 + c];
            a.out[p * COUT + o] = s;
            if (a.tgt) {
                const float r = s - a.tgt[p * COUT + o];
                l += 0.5 * (double)r * (double)r;
                go[o] = r * a.inv_n;
            }
        }
        if (a.tgt) {
#pragma unroll
            for (int c = 0; c < CH0; ++c) {
                float g = 0.f;
#pragma unroll
                for (int o = 0; o < COUT; ++o) {
                    g += go[o] * a.wh[o * CH0 + c];
                    atomicAdd(&sg[o * CH0 + c], go[o] * d[c]);
                }
                a.g_d0[p * CH0 + c] = f2bf(d[c] > 0.f ? g : 0.f);
            }
        }
    }
    if (!a.tgt) return;
    sl[threadIdx.x] = l;
    __syncthreads();
    for (int st = NT / 2; st > 0; st >>= 1) {
        if ((int)threadIdx.x < st) sl[threadIdx.x] += sl[threadIdx.x + st];
        __syncthreads();
    }
    if (threadIdx.x == 0) a.loss_part[blockIdx.x] = sl[0] * (double)a.inv_n;
    for (int i = threadIdx.x; i < COUT * CH0; i += NT) atomicAdd(&a.gwh[i], sg[i]);
}
__global__ void k_sum_loss(const double* part, int nb, double* out) {
    __shared__ double s[NT];
    double acc = 0.0;
    for (int i = threadIdx.x; i < nb; i += NT) acc += part[i];
    s[threadIdx.x] = acc;
    __syncthreads();
    for (int st = NT / 2; st > 0; st >>= 1) {
        if ((int)threadIdx.x < st) s[threadIdx.x] += s[threadIdx.x + st];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[0] = s[0];
}

// ---------------------------------------------------------------- parameters
struct LayerDesc {
    int cout, cin, cin_dev;    // cin_dev: channels on the device (e0a padded to 16)
    int64_t off;               // offset in the flat FP32 vector
    int64_t woff;              // offset of the BF16 [cout][9][cin_dev] copy
    int64_t toff;              // offset of the BF16 transposed copy [cin_dev][9][cout]
};
struct Layers { LayerDesc L[NLAYER]; int64_t n, nw; };

Layers make_layers() {
    const int spec[NLAYER][2] = {{CH0, CIN}, {CH0, CH0}, {CH1, CH0}, {CH1, CH1}, {CH2, CH1}, {CH2, CH2},
                                 {CH3, CH2}, {CH3, CH3}, {CH2, CH3 + CH2}, {CH2, CH2}, {CH1, CH2 + CH1},
                                 {CH1, CH1}, {CH0, CH1 + CH0}, {CH0, CH0}, {COUT, CH0}};
    Layers Ls{};
    int64_t o = 0, w = 0;
    for (int i = 0; i < NLAYER; ++i) {
        LayerDesc& d = Ls.L[i];
        d.cout = spec[i][0];
        d.cin = spec[i][1];
        d.cin_dev = d.cin < 16 ? 16 : d.cin;
        d.off = o;
        const bool head = i == NLAYER - 1;
        o += head ? (int64_t)d.cout * d.cin : (int64_t)d.cout * 9 * d.cin;
        if (!head) {
            d.woff = w;
            w += (int64_t)d.cout * 9 * d.cin_dev;
            d.toff = w;
            w += (int64_t)d.cout * 9 * d.cin_dev;
        }
    }
    Ls.n = o;
    Ls.nw = w;
    return Ls;
}

// FP32 master -> BF16 [cout][9][cin_dev] and transposed/flipped [cin_dev][9][cout]
__global__ __launch_bounds__(NT) void k_cast_w(const float* __restrict__ th, bf16_t* __restrict__ wb, LayerDesc d) {
    const int64_t k = (int64_t)blockIdx.x * NT + threadIdx.x;
    const int64_t n = (int64_t)d.cout * 9 * d.cin_dev;
    if (k >= n) return;
    const int ci = (int)(k % d.cin_dev);
    const int t = (int)((k / d.cin_dev) % 9);
    const int co = (int)(k / (9 * (int64_t)d.cin_dev));
    const float v = ci < d.cin ? th[d.off + ((int64_t)co * 9 + t) * d.cin + ci] : 0.f;
    const bf16_t b = f2bf(v);
    wb[d.woff + k] = b;
    wb[d.toff + ((int64_t)ci * 9 + (8 - t)) * d.cout + co] = b;
}

// Adam (Kingma & Ba 2015, algorithm 1) over the flat FP32 vectors; bias corrections passed in
__global__ __launch_bounds__(NT) void k_adam(float* __restrict__ th, float* __restrict__ m, float* __restrict__ v,
                                             const float* __restrict__ g, int64_t n, float lr, float b1, float b2,
                                             float c1, float c2, float eps) {
    const int64_t i = (int64_t)blockIdx.x * NT + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.f - b1) * gi;
    const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    th[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}
__global__ __launch_bounds__(NT) void k_fillf(float* a, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * NT + threadIdx.x;
    if (i < n) a[i] = 0.f;
}

// ---------------------------------------------------------------- host side
struct Net {
    Layers Ls;
    float *th = nullptr, *m = nullptr, *v = nullptr, *g = nullptr;  // FP32 [n] each, one allocation
    bf16_t* wb = nullptr;                                          // BF16 weights + transposes
    bool wb_fresh = false;
    int64_t adam_t = 0;
    int64_t cap = 0;   // level-0 pixel capacity of the activation buffers
    void* act = nullptr;
    bf16_t *in16, *ra[4], *e3, *p[4], *cat[3], *rad[3], *d[3];
    bf16_t *gA, *gB, *gC, *gcat[3];
    float* out = nullptr;
    double* lpart = nullptr;
    double* lsum = nullptr;
    bool has_params = false;
    float in_scale[4] = {1.f, 1.f, 1.f, 1.f};
    float out_scale[3] = {1.f, 1.f, 1.f};
};

void net_free(xc_ctx*, void* q) {
    Net* n = (Net*)q;
    (void)hipFree(n->th);
    (void)hipFree(n->wb);
    (void)hipFree(n->act);
    delete n;
}

inline unsigned nblk(int64_t n) { return (unsigned)((n + NT - 1) / NT); }

int get_net(xc_ctx* ctx, Net** out) {
    Net* n = (Net*)ctx->ext[XC_EXT_UNET];
    if (n) { *out = n; return XC_OK; }
    n = new Net();
    n->Ls = make_layers();
    const int64_t N = n->Ls.n;
    if (hipMalloc(&n->th, 4 * N * sizeof(float)) != hipSuccess) { delete n; return xc_fail(ctx, XC_ENOMEM, "unet: parameters"); }
    n->m = n->th + N;
    n->v = n->m + N;
    n->g = n->v + N;
    if (hipMalloc(&n->wb, n->Ls.nw * sizeof(bf16_t)) != hipSuccess) {
        (void)hipFree(n->th);
        delete n;
        return xc_fail(ctx, XC_ENOMEM, "unet: weights");
    }
    hipLaunchKernelGGL(k_fillf, dim3(nblk(4 * N)), dim3(NT), 0, ctx->stream, n->th, 4 * N);
    ctx->ext[XC_EXT_UNET] = n;
    ctx->ext_free[XC_EXT_UNET] = net_free;
    *out = n;
    return XC_OK;
}

int ensure_acts(xc_ctx* ctx, Net* n, int B, int H, int W) {
    const int64_t P0 = (int64_t)B * H * W;
    if (P0 <= n->cap) return XC_OK;
    if (n->act) {
        XC_TRY_HIP(ctx, hipStreamSynchronize(ctx->stream));
        (void)hipFree(n->act);
        n->act = nullptr;
        n->cap = 0;
    }
    const int64_t P1 = P0 / 4, P2 = P0 / 16, P3 = P0 / 64;
    auto r = [](int64_t k) { return (k + 63) & ~(int64_t)63; };  // keep every tensor 128-byte aligned
    const int64_t sizes[] = {
        r(P0 * 16),                                                    // in16
        r(P0 * CH0), r(P1 * CH1), r(P2 * CH2), r(P3 * CH3),            // ra[0..3]
        r(P3 * CH3),                                                   // e3
        r(P1 * CH0), r(P2 * CH1), r(P3 * CH2),                         // p[1..3]
        r(P0 * (CH1 + CH0)), r(P1 * (CH2 + CH1)), r(P2 * (CH3 + CH2)), // cat[0..2]
        r(P0 * CH0), r(P1 * CH1), r(P2 * CH2),                         // rad[0..2]
        r(P0 * CH0), r(P1 * CH1), r(P2 * CH2),                         // d[0..2]
        r(P0 * 16), r(P0 * 16), r(P0 * 16),                            // gA, gB, gC
        r(P0 * (CH1 + CH0)), r(P1 * (CH2 + CH1)), r(P2 * (CH3 + CH2))  // gcat[0..2]
    };
    int64_t tot = 0;
    for (int64_t s : sizes) tot += s;
    const int64_t outb = r(P0 * COUT * 2);             // FP32 output, in bf16 units
    const int64_t lpb = r(4 * ((int64_t)nblk(P0) + 2)); // FP64 loss partials + sum
    void* mem;
    const hipError_t e = hipMalloc(&mem, (tot + outb + lpb) * sizeof(bf16_t));
    if (e != hipSuccess) return xc_fail(ctx, XC_ENOMEM, "unet: activations for %lld pixels", (long long)P0);
    n->act = mem;
    bf16_t* q = (bf16_t*)mem;
    int i = 0;
    auto take = [&]() { bf16_t* t = q; q += sizes[i++]; return t; };
    n->in16 = take();
    for (int l = 0; l < 4; ++l) n->ra[l] = take();
    n->e3 = take();
    n->p[0] = nullptr;
    for (int l = 1; l < 4; ++l) n->p[l] = take();
    for (int l = 0; l < 3; ++l) n->cat[l] = take();
    for (int l = 0; l < 3; ++l) n->rad[l] = take();
    for (int l = 0; l < 3; ++l) n->d[l] = take();
    n->gA = take(); n->gB = take(); n->gC = take();
    for (int l = 0; l < 3; ++l) n->gcat[l] = take();
    n->out = (float*)q;
    q += outb;
    n->lpart = (double*)q;
    n->lsum = n->lpart + nblk(P0);
    n->cap = P0;
    return XC_OK;
}

int conv(xc_ctx* ctx, const bf16_t* in, int in_stride, int cin, const bf16_t* w, int cout, bf16_t* out,
         int out_stride, int out_off, int mode, const bf16_t* mask, int mask_stride, int mask_off, int B, int H,
         int W) {
    ConvArgs a{in, in_stride, cin, w, out, out_stride, out_off, mask, mask_stride, mask_off, mode, B, H, W,
               (int64_t)B * H * W};
    const unsigned g = (unsigned)((a.P + 63) / 64);
    switch (cout / 16) {
        case 1: hipLaunchKernelGGL(k_conv3<1>, dim3(g), dim3(NT), 0, ctx->stream, a); break;
        case 2: hipLaunchKernelGGL(k_conv3<2>, dim3(g), dim3(NT), 0, ctx->stream, a); break;
        case 3: hipLaunchKernelGGL(k_conv3<3>, dim3(g), dim3(NT), 0, ctx->stream, a); break;
        case 4: hipLaunchKernelGGL(k_conv3<4>, dim3(g), dim3(NT), 0, ctx->stream, a); break;
        case 5: hipLaunchKernelGGL(k_conv3<5>, dim3(g), dim3(NT), 0, ctx->stream, a); break;
        case 7: hipLaunchKernelGGL(k_conv3<7>, dim3(g), dim3(NT), 0, ctx->stream, a); break;
        default: return xc_fail(ctx, XC_EINVAL_ARG, "unet: unsupported channel count %d", cout);
    }
    return XC_OK;
}

void wgrad(xc_ctx* ctx, const bf16_t* x, int x_stride, int cin, const bf16_t* gy, int cout, float* gw, int cin_w,
           int B, int H, int W) {
    WGradArgs a{x, x_stride, cin, gy, cout, cout, gw, cin_w, B, H, W, (int64_t)B * H * W, 256};
    const int64_t waves = (a.P + a.chunk - 1) / a.chunk;
    const unsigned gx = (unsigned)((waves + 3) / 4);
    const unsigned gy_ = (unsigned)(9 * (cout / 16) * (cin / 16));
    hipLaunchKernelGGL(k_conv3_wgrad, dim3(gx, gy_), dim3(NT), 0, ctx->stream, a);
}

void refresh_weights(xc_ctx* ctx, Net* n) {
    if (n->wb_fresh) return;
    for (int i = 0; i < NLAYER - 1; ++i) {
        const LayerDesc& d = n->Ls.L[i];
        hipLaunchKernelGGL(k_cast_w, dim3(nblk((int64_t)d.cout * 9 * d.cin_dev)), dim3(NT), 0, ctx->stream, n->th,
                           n->wb, d);
    }
    n->wb_fresh = true;
}

int check_geom(xc_ctx* ctx, int B, int H, int W) {
    if (B < 1 || H < 8 || W < 8 || H % 8 || W % 8)
        return xc_fail(ctx, XC_EINVAL_SHAPE, "unet: batch %d, grid %d x %d (need >= 8, multiples of 8)", B, H, W);
    return XC_OK;
}

constexpr int CHS[4] = {CH0, CH1, CH2, CH3};

// forward pass; keeps every tensor the backward pass needs.  With tgt: loss and the head's gradients.
int forward(xc_ctx* ctx, Net* n, const float* in, int B, int H, int W, const float* tgt) {
    int st;
    if ((st = ensure_acts(ctx, n, B, H, W))) return st;
    refresh_weights(ctx, n);
    const Layers& Ls = n->Ls;
    const int64_t P0 = (int64_t)B * H * W;
    hipLaunchKernelGGL(k_input, dim3(nblk(P0 * 16)), dim3(NT), 0, ctx->stream, in, n->in16, P0);
    const int Hs[4] = {H, H / 2, H / 4, H / 8}, Ws[4] = {W, W / 2, W / 4, W / 8};
    // encoder; e_l (l < 3) is written straight into the skip slot of cat[l] = [up | e_l]
    const bf16_t* x = n->in16;
    int xs = 16;
    for (int l = 0; l < 4; ++l) {
        const LayerDesc& A = Ls.L[2 * l];
        const LayerDesc& Bd = Ls.L[2 * l + 1];
        if (l > 0) {
            const int cp = CHS[l - 1];
            hipLaunchKernelGGL(k_pool, dim3(nblk((int64_t)B * Hs[l] * Ws[l] * cp)), dim3(NT), 0, ctx->stream,
                               n->cat[l - 1], CHS[l] + cp, CHS[l], n->p[l], cp, cp, B, Hs[l], Ws[l]);
            x = n->p[l];
            xs = cp;
        }
        if ((st = conv(ctx, x, xs, A.cin_dev, n->wb + A.woff, A.cout, n->ra[l], A.cout, 0, 0, nullptr, 0, 0, B, Hs[l],
                       Ws[l])))
            return st;
        bf16_t* eo = l < 3 ? n->cat[l] : n->e3;
        const int eos = l < 3 ? CHS[l + 1] + CHS[l] : CH3;
        const int eoo = l < 3 ? CHS[l + 1] : 0;
        if ((st = conv(ctx, n->ra[l], A.cout, A.cout, n->wb + Bd.woff, Bd.cout, eo, eos, eoo, 0, nullptr, 0, 0, B,
                       Hs[l], Ws[l])))
            return st;
    }
    // decoder: level l = 2, 1, 0 reads cat[l] = [up(coarser output) | e_l]
    const bf16_t* coarse = n->e3;
    int cc = CH3;
    for (int l = 2; l >= 0; --l) {
        const int cs = cc + CHS[l];
        hipLaunchKernelGGL(k_up, dim3(nblk((int64_t)B * Hs[l] * Ws[l] * cc)), dim3(NT), 0, ctx->stream, coarse, cc,
                           n->cat[l], cs, 0, cc, B, Hs[l], Ws[l]);
        const LayerDesc& A = Ls.L[8 + 2 * (2 - l)];
        const LayerDesc& Bd = Ls.L[9 + 2 * (2 - l)];
        if ((st = conv(ctx, n->cat[l], cs, cs, n->wb + A.woff, A.cout, n->rad[l], A.cout, 0, 0, nullptr, 0, 0, B,
                       Hs[l], Ws[l])))
            return st;
        if ((st = conv(ctx, n->rad[l], A.cout, A.cout, n->wb + Bd.woff, Bd.cout, n->d[l], Bd.cout, 0, 0, nullptr, 0, 0,
                       B, Hs[l], Ws[l])))
            return st;
        coarse = n->d[l];
        cc = Bd.cout;
    }
    HeadArgs h{};
    h.d0 = n->d[0];
    h.wh = n->th + Ls.L[NLAYER - 1].off;
    h.out = n->out;
    h.tgt = tgt;
    h.inv_n = 1.0f / (float)P0;
    h.g_d0 = n->gA;
    h.gwh = n->g + Ls.L[NLAYER - 1].off;
    h.loss_part = n->lpart;
    h.P = P0;
    hipLaunchKernelGGL(k_head, dim3(nblk(P0)), dim3(NT), 0, ctx->stream, h);
    if (tgt) hipLaunchKernelGGL(k_sum_loss, dim3(1), dim3(NT), 0, ctx->stream, n->lpart, (int)nblk(P0), n->lsum);
    return XC_OK;
}

// one conv block backward: g = dL/d(pre-activation of conv b) (masked) -> weight grads of a and b;
// returns dL/d(block input) in gx (plain) unless gx is null
int block_back(xc_ctx* ctx, Net* n, const LayerDesc& A, const LayerDesc& Bd, const bf16_t* x, int xs,
               const bf16_t* ra, const bf16_t* g, bf16_t* gya, bf16_t* gx, int gx_stride, int B, int H, int W) {
    int st;
    wgrad(ctx, ra, A.cout, A.cout, g, Bd.cout, n->g + Bd.off, Bd.cin, B, H, W);
    if ((st = conv(ctx, g, Bd.cout, Bd.cout, n->wb + Bd.toff, Bd.cin_dev, gya, Bd.cin_dev, 0, 1, ra, A.cout, 0, B, H,
                   W)))
        return st;
    wgrad(ctx, x, xs, A.cin_dev, gya, A.cout, n->g + A.off, A.cin, B, H, W);
    if (gx && (st = conv(ctx, gya, A.cout, A.cout, n->wb + A.toff, A.cin_dev, gx, gx_stride, 0, 2, nullptr, 0, 0, B,
                         H, W)))
        return st;
    return XC_OK;
}

// backward pass after forward(..., tgt); n->g must have been zeroed before that forward pass
int backward(xc_ctx* ctx, Net* n, int B, int H, int W) {
    int st;
    const Layers& Ls = n->Ls;
    const int Hs[4] = {H, H / 2, H / 4, H / 8}, Ws[4] = {W, W / 2, W / 4, W / 8};
    bf16_t* g = n->gA;   // dL/d(pre-activation) of the current block's second conv (masked)
    bf16_t* s1 = n->gB;
    bf16_t* s2 = n->gC;
    // decoder blocks d0, d1, d2: input gradients into gcat[l]; then the upsampling's coarse gradient
    for (int l = 0; l <= 2; ++l) {
        const LayerDesc& A = Ls.L[8 + 2 * (2 - l)];
        const LayerDesc& Bd = Ls.L[9 + 2 * (2 - l)];
        const int cs = A.cin;
        if ((st = block_back(ctx, n, A, Bd, n->cat[l], cs, n->rad[l], g, s1, n->gcat[l], cs, B, Hs[l], Ws[l])))
            return st;
        // coarser output: d[l+1] for l < 2, e3 for l = 2 (mask); its gradient = sum of the 4 fine cells
        const int cc = cs - CHS[l];
        GradMergeArgs m{};
        m.gup = n->gcat[l]; m.gup_stride = cs; m.gup_off = 0;
        m.e = l < 2 ? n->d[l + 1] : n->e3; m.e_stride = cc; m.e_off = 0;
        m.out = s2; m.C = cc; m.B = B; m.H = Hs[l + 1]; m.W = Ws[l + 1];
        hipLaunchKernelGGL(k_grad_merge, dim3(nblk((int64_t)B * Hs[l + 1] * Ws[l + 1] * cc)), dim3(NT), 0,
                           ctx->stream, m);
        bf16_t* t = g; g = s2; s2 = t;
    }
    // encoder blocks e3, e2, e1, e0
    for (int l = 3; l >= 0; --l) {
        const LayerDesc& A = Ls.L[2 * l];
        const LayerDesc& Bd = Ls.L[2 * l + 1];
        const bf16_t* x = l > 0 ? n->p[l] : n->in16;
        const int xs = l > 0 ? CHS[l - 1] : 16;
        if ((st = block_back(ctx, n, A, Bd, x, xs, n->ra[l], g, s1, l > 0 ? s2 : nullptr, A.cin_dev, B, Hs[l], Ws[l])))
            return st;
        if (l == 0) break;
        // e_{l-1}: pooled into this block (s2 = dL/dp_l) and sent through the skip of cat[l-1]
        const int c = CHS[l - 1];
        GradMergeArgs m{};
        m.gpool = s2; m.gpool_stride = c;
        m.gskip = n->gcat[l - 1]; m.gskip_stride = CHS[l] + c; m.gskip_off = CHS[l];
        m.e = n->cat[l - 1]; m.e_stride = CHS[l] + c; m.e_off = CHS[l];
        m.out = g; m.C = c; m.B = B; m.H = Hs[l - 1]; m.W = Ws[l - 1];
        hipLaunchKernelGGL(k_grad_merge, dim3(nblk((int64_t)B * Hs[l - 1] * Ws[l - 1] * c)), dim3(NT), 0,
                           ctx->stream, m);
    }
    return XC_OK;
}

}  // namespace

extern "C" {

int xc_unet_num_params(int64_t* n) {
    if (!n) return XC_EINVAL_ARG;
    *n = make_layers().n;
    return XC_OK;
}

int xc_unet_set_params(xc_ctx* ctx, const float* host, int64_t n) {
    XC_CHECK_CTX(ctx);
    Net* net;
    int st;
    if ((st = get_net(ctx, &net))) return st;
    if (!host || n != net->Ls.n) return xc_fail(ctx, XC_EINVAL_ARG, "unet: expected %lld parameters", (long long)net->Ls.n);
    XC_TRY_HIP(ctx, hipMemcpyAsync(net->th, host, n * sizeof(float), hipMemcpyHostToDevice, ctx->stream));
    hipLaunchKernelGGL(k_fillf, dim3(nblk(3 * n)), dim3(NT), 0, ctx->stream, net->m, 3 * n);  // Adam moments, grads
    XC_TRY_HIP(ctx, hipStreamSynchronize(ctx->stream));
    net->adam_t = 0;
    net->wb_fresh = false;
    net->has_params = true;
    return xc_finish(ctx, "unet_set_params");
}

int xc_unet_get_params(xc_ctx* ctx, float* host, int64_t n) {
    XC_CHECK_CTX(ctx);
    Net* net;
    int st;
    if ((st = get_net(ctx, &net))) return st;
    if (!host || n != net->Ls.n) return xc_fail(ctx, XC_EINVAL_ARG, "unet: expected %lld parameters", (long long)net->Ls.n);
    XC_TRY_HIP(ctx, hipMemcpyAsync(host, net->th, n * sizeof(float), hipMemcpyDeviceToHost, ctx->stream));
    XC_TRY_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return XC_OK;
}

int xc_unet_set_io_scales(xc_ctx* ctx, const float in_scale[4], const float out_scale[3]) {
    XC_CHECK_CTX(ctx);
    Net* net;
    int st;
    if ((st = get_net(ctx, &net))) return st;
    if (!in_scale || !out_scale) return xc_fail(ctx, XC_EINVAL_ARG, "unet_set_io_scales: NULL");
    for (int k = 0; k < 4; ++k) net->in_scale[k] = in_scale[k];
    for (int k = 0; k < 3; ++k) net->out_scale[k] = out_scale[k];
    return XC_OK;
}

int xc_unet_forward(xc_ctx* ctx, const float* in, int B, int H, int W, float* out) {
    XC_CHECK_CTX(ctx);
    Net* net;
    int st;
    if ((st = check_geom(ctx, B, H, W))) return st;
    if (!in || !out) return xc_fail(ctx, XC_EINVAL_ARG, "unet_forward: NULL buffer");
    if ((st = get_net(ctx, &net))) return st;
    if ((st = forward(ctx, net, in, B, H, W, nullptr))) return st;
    XC_TRY_HIP(ctx, hipMemcpyAsync(out, net->out, (size_t)B * H * W * COUT * sizeof(float), hipMemcpyDeviceToDevice,
                                   ctx->stream));
    return xc_finish(ctx, "unet_forward");
}

int xc_unet_loss_grad(xc_ctx* ctx, const float* in, const float* tgt, int B, int H, int W, double* loss,
                      float* grad) {
    XC_CHECK_CTX(ctx);
    Net* net;
    int st;
    if ((st = check_geom(ctx, B, H, W))) return st;
    if (!in || !tgt) return xc_fail(ctx, XC_EINVAL_ARG, "unet_loss_grad: NULL buffer");
    if ((st = get_net(ctx, &net))) return st;
    const int64_t N = net->Ls.n;
    hipLaunchKernelGGL(k_fillf, dim3(nblk(N)), dim3(NT), 0, ctx->stream, net->g, N);
    if ((st = forward(ctx, net, in, B, H, W, tgt))) return st;
    if ((st = backward(ctx, net, B, H, W))) return st;
    if (grad) XC_TRY_HIP(ctx, hipMemcpyAsync(grad, net->g, N * sizeof(float), hipMemcpyDeviceToDevice, ctx->stream));
    if (loss) {
        XC_TRY_HIP(ctx, hipMemcpyAsync(loss, net->lsum, sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
        XC_TRY_HIP(ctx, hipStreamSynchronize(ctx->stream));
    }
    return xc_finish(ctx, "unet_loss_grad");
}

int xc_unet_train_step(xc_ctx* ctx, const float* in, const float* tgt, int B, int H, int W, float lr, double* loss) {
    int st;
    if ((st = xc_unet_loss_grad(ctx, in, tgt, B, H, W, loss, nullptr))) return st;
    Net* net = (Net*)ctx->ext[XC_EXT_UNET];
    const int64_t N = net->Ls.n;
    net->adam_t += 1;
    const float b1 = 0.9f, b2 = 0.999f;
    const float c1 = 1.0f - powf(b1, (float)net->adam_t), c2 = 1.0f - powf(b2, (float)net->adam_t);
    hipLaunchKernelGGL(k_adam, dim3(nblk(N)), dim3(NT), 0, ctx->stream, net->th, net->m, net->v, net->g, N, lr, b1, b2,
                       c1, c2, 1e-8f);
    net->wb_fresh = false;
    return xc_finish(ctx, "unet_train_step");
}

}  // extern "C"

int xc_unet_infer(xc_ctx* ctx, const float* in, int B, int H, int W, const float** out) {
    Net* net;
    int st;
    if ((st = check_geom(ctx, B, H, W))) return st;
    if ((st = get_net(ctx, &net))) return st;
    if ((st = forward(ctx, net, in, B, H, W, nullptr))) return st;
    *out = net->out;
    return XC_OK;
}

bool xc_unet_has_params(xc_ctx* ctx) {
    const Net* n = (const Net*)ctx->ext[XC_EXT_UNET];
    return n && n->has_params;
}

void xc_unet_io_scales(xc_ctx* ctx, float in[4], float out[3]) {
    const Net* n = (const Net*)ctx->ext[XC_EXT_UNET];
    if (in)
        for (int k = 0; k < 4; ++k) in[k] = n ? n->in_scale[k] : 1.f;
    if (out)
        for (int k = 0; k < 3; ++k) out[k] = n ? n->out_scale[k] : 1.f;
}
