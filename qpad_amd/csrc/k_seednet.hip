// k_seednet.hip -- the learned B_perp seed of the predictor-corrector
// iteration: a 1D convolution network over the radial nodes of a slice, its
// forward, backward and Adam kernels, the hooks into the slice step and the
// C ABI of the network.  It has no counterpart in QPAD; it only replaces the
// seed b_spe(xi-1) that QPAD's iteration starts from (simulation_class.f03:
// 326, 370-375), and the iteration still runs to the same tolerance.  The
// seed is the first-order extrapolation 2 B(xi-1) - B(xi-2) plus the
// network's correction (B_perp components 0 and 1, every mode).
//
// Matrix-core mapping (v_mfma_f32_16x16x32_bf16, one wave, lane l = 16 q + j):
//   A (16 x 32): lane l holds A[j][8 q + e], e = 0..7
//   B (32 x 16): lane l holds B[8 q + e][j]
//   D (16 x 16): lane l, register r holds D[4 q + r][j]
// A conv layer  z[p][co] = b[co] + sum_tap sum_ci W[tap][co][ci] h[p+tap-1][ci]
// is a GEMM with rows co, columns p and the reduction over (tap, ci).
// Activations are stored position-major, [row = p + 1][channel], with zero
// guard rows at p = -1 and p >= npos, so the taps are row offsets and every
// operand is one aligned 16-byte load.  Channels are padded to multiples of
// 32 with zero weights.  Weight gradients are reduced over fixed chunks of
// positions in a fixed order, so training is deterministic.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "kernels.h"

namespace qpad {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int kNetMaxLayers = 8;
constexpr int kNetChunk = 128;      // positions per weight-gradient workgroup
constexpr int kNetLossLog = 1 << 16;
constexpr int kNetQty = 8;          // input quantities per mode array
constexpr int kNetMaxCinP = 128;
constexpr int kNetMaxCoutP = 64;
constexpr int kNetLossThreads = 256;

struct SeedNet {
    qpad_seednet_cfg cfg;
    int layers, cin, cout, npos, npos_pad, nchunk;
    int cin_p[kNetMaxLayers], cout_p[kNetMaxLayers];
    long long w_off[kNetMaxLayers], b_off[kNetMaxLayers], wb_off[kNetMaxLayers];
    long long nparams, nwb;
    long long steps;
    int seed_on, train_on;
    int packed_slice;       // slice whose inputs are packed (0: none)
    bool fwd_done;          // forward pass of that slice done
    float *theta, *am, *av, *part, *grad;
    __bf16 *wb, *wtb;       // BF16 weights [3][cout_p][cin_p] and [3][cin_p][cout_p]
    __bf16 *h[kNetMaxLayers];   // input of layer l: [npos_pad + 2][cin_p]
    float *z[kNetMaxLayers];    // pre-activation of layer l: [npos_pad][cout_p]
    __bf16 *dz[kNetMaxLayers];  // d loss / d z of layer l: [npos_pad + 2][cout_p]
    double *scales;         // B_perp, J_perp, q_spe, q_beam maxima of the slice
    double *scale_part;     // per-workgroup maxima [nscale_blk][4]
    int nscale_blk;
    int infer_stride;       // BF16 per LDS row of the fused inference
    double *seed_buf;       // the seed of the next slice [2][nm][ld]
    hipStream_t stream;     // the network's stream (default priority)
    hipEvent_t ev_main, ev_net;
    double *bprev;          // b_spe components 0, 1 at the last two slice starts
    float *loss_part, *loss_log;
    float *xin, *tin;       // host inputs of qpad_seednet_eval
    int nloss_blk;
};

// ---------------------------------------------------------------------------
// device helpers
// ---------------------------------------------------------------------------
// logistic function for SiLU: fast exp and reciprocal (relative error near
// 1e-7, far below the BF16 rounding of the activations it feeds)
__device__ __forceinline__ float net_sigmoid(float z) { return __builtin_amdgcn_rcpf(1.0f + __expf(-z)); }

__device__ __forceinline__ f32x4 mfma_bf16(bf16x8 a, bf16x8 b, f32x4 c)
{
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}

#define NX(p, c, k, i) (p)[((size_t)(c) * nm + (size_t)(k)) * ld + (size_t)(i)]

// ---------------------------------------------------------------------------
// slice inputs: per-quantity maxima, then the packed BF16 input rows
// ---------------------------------------------------------------------------
struct NetScaleArgs {
    const double *b_spe, *cu, *q_spe, *q_beam;
    double *bsave;         // copy of b_spe components 0, 1 (the seed is
                           // written into b_spe while neighbours still read)
    double *part;          // per-workgroup maxima [gridDim.x][4]
    int nr, nm, ld;
};

// maxima over nodes 1..nr and every mode array of |B_r|,|B_phi|; |J_r|,
// |J_phi|; |q_spe|; |q_beam| (per workgroup) and the saved copy of B_perp
__global__ __launch_bounds__(256) void k_net_scale(NetScaleArgs a)
{
    __shared__ double red[4][4];
    const int nr = a.nr, nm = a.nm, ld = a.ld;
    double s[4] = {0.0, 0.0, 0.0, 0.0};
    for (int idx = blockIdx.x * blockDim.x + threadIdx.x; idx < nm * nr; idx += gridDim.x * blockDim.x) {
        const int k = idx / nr, i = idx % nr + 1;
        const double b0 = NX(a.b_spe, 0, k, i), b1 = NX(a.b_spe, 1, k, i);
        NX(a.bsave, 0, k, i) = b0;
        NX(a.bsave, 1, k, i) = b1;
        s[0] = fmax(s[0], fmax(fabs(b0), fabs(b1)));
        s[1] = fmax(s[1], fmax(fabs(NX(a.cu, 0, k, i)), fabs(NX(a.cu, 1, k, i))));
        s[2] = fmax(s[2], fabs(NX(a.q_spe, 0, k, i)));
        s[3] = fmax(s[3], fabs(NX(a.q_beam, 0, k, i)));
    }
#pragma unroll
    for (int q = 0; q < 4; q++)
        for (int o = 32; o > 0; o >>= 1) s[q] = fmax(s[q], __shfl_xor(s[q], o, 64));
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0)
        for (int q = 0; q < 4; q++) red[q][w] = s[q];
    __syncthreads();
    if (threadIdx.x < 4) {
        double m = 0.0;
        for (int v = 0; v < 4; v++) m = fmax(m, red[threadIdx.x][v]);
        a.part[blockIdx.x * 4 + threadIdx.x] = m;
    }
}

// the slice's scales from the per-workgroup maxima, into LDS (threads 0..3;
// the caller synchronises); workgroup 0 also stores them for later kernels.
// An all-zero quantity (ahead of the beam) keeps its inputs at zero.
__device__ __forceinline__ void net_final_scales(const double *part, int nsp, double *out, double *ssc)
{
    // wave 0: one workgroup's maxima per lane (nsp <= 64), then a wave max
    if (threadIdx.x < 64) {
        const int l = threadIdx.x;
        const int g = min(l, nsp - 1);
        double v[4];
#pragma unroll
        for (int q = 0; q < 4; q++) v[q] = l < nsp ? part[g * 4 + q] : 0.0;
#pragma unroll
        for (int q = 0; q < 4; q++)
            for (int o = 32; o > 0; o >>= 1) v[q] = fmax(v[q], __shfl_xor(v[q], o, 64));
        if (l < 4) {
            const double m = l == 0 ? v[0] : (l == 1 ? v[1] : (l == 2 ? v[2] : v[3]));
            const double x = (m > 0.0 && m < INFINITY) ? m : 1.0;
            ssc[l] = x;
            if (blockIdx.x == 0) out[l] = x;
        }
    }
}

// input channel ch of node i: qty * nm + k over the 8 quantities, then
// r / r_max.  Branch-free (selects, one or two unconditional loads) so that
// a thread's channels issue their loads together.
__device__ __forceinline__ float net_input(int ch, int qc, int kc, int i, int nm, int ld, const double *b1,
                                           const double *b2, const double *cu, const double *q_spe,
                                           const double *q_beam, const double *inv, float rinv)
{
    // qc, kc: ch / nm and ch % nm
    const bool field = ch < kNetQty * nm;
    const int qty = field ? qc : 0, k = field ? kc : 0;
    // qty: 0, 1 B_r, B_phi; 2, 3 their change; 4, 5 J_r, J_phi; 6 q_spe; 7 q_beam
    const int comp = qty & 1 & (qty < 6);
    const double *arr = qty < 4 ? b1 : (qty < 6 ? cu : (qty == 6 ? q_spe : q_beam));
    const int sidx = qty < 4 ? 0 : (qty < 6 ? 1 : qty - 4);
    const size_t off = ((size_t)comp * nm + (size_t)k) * ld + (size_t)i;
    const double xa = arr[off];
    const double xb = b2[off];
    const double x = (xa - ((qty == 2 || qty == 3) ? xb : 0.0)) * inv[sidx];
    return field ? (float)x : (ch == kNetQty * nm ? (float)(i - 1) * rinv : 0.0f);
}

struct NetPackArgs {
    const double *bsave;   // B_perp of slice xi-1 (saved by k_net_scale)
    const double *bold;    // B_perp of slice xi-2 (saved at the previous slice start)
    const double *cu, *q_spe, *q_beam;
    const double *part;
    int nsp;
    double *scales;
    __bf16 *h0;
    int nr, nm, ld, cin, cin_p;
    float rinv;            // 1 / nr: the radius feature (i - 1) / nr = r / r_max
};

// one thread per (group of 8 channels, node)
__global__ __launch_bounds__(256) void k_net_pack(NetPackArgs a)
{
    __shared__ double ssc[4];
    net_final_scales(a.part, a.nsp, a.scales, ssc);
    __syncthreads();
    const int nr = a.nr, nm = a.nm, ld = a.ld;
    const int ng = a.cin_p >> 3;
    const int idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= ng * nr) return;
    const int g = idx / nr, pos = idx % nr, i = pos + 1;
    const double inv[4] = {1.0 / ssc[0], 1.0 / ssc[1], 1.0 / ssc[2], 1.0 / ssc[3]};
    bf16x8 out;
    int qc = (8 * g) / nm, kc = (8 * g) % nm;
#pragma unroll
    for (int e = 0; e < 8; e++) {
        out[e] = (__bf16)net_input(8 * g + e, qc, kc, i, nm, ld, a.bsave, a.bold, a.cu, a.q_spe, a.q_beam, inv,
                                   a.rinv);
        if (++kc == nm) {
            kc = 0;
            qc++;
        }
    }
    *(bf16x8 *)(a.h0 + (size_t)i * a.cin_p + 8 * g) = out;
}

// ---------------------------------------------------------------------------
// fused inference of one slice (seed mode): inputs, every layer and the seed
// in one launch.  A workgroup owns kInferT nodes; it recomputes a halo of
// one node per remaining layer on each side, so the layers run from LDS
// without a grid-wide step between them.  LDS row r holds node position
// p0 - L + r; rows outside 0 <= p < nr stay zero (the convolution's
// padding) at every layer.  The same products in the same order as the
// layer-by-layer kernels: the outputs are identical.  Everything the
// workgroup reads from memory (all layers' weights, the input fields, the
// FP64 B_perp values of its own nodes for the seed) is requested up front,
// so the launch costs about two memory round trips.  With `store`, the
// owned rows' layer inputs and pre-activations go to global memory for the
// training step.
//
// Dynamic LDS: [2 (B(xi-1), B(xi-2))][2 comp][nm][kInferT] doubles, then two
// activation buffers [kInferRows][stride] and every layer's weights
// (BF16, the global layout).
// ---------------------------------------------------------------------------
constexpr int kInferT = 32;
constexpr int kInferThreads = 256;
constexpr int kInferRows = 64;          // >= kInferT + 2 L + 16 (tile overhang), L <= 8
constexpr int kInferMaxW = 49152;       // BF16 weights staged in LDS (96 KB)
constexpr int kInferWVec = kInferMaxW / 8 / kInferThreads;   // 16-byte loads per thread
constexpr int kInferItems = 2;          // input (row, channel group) items per thread
constexpr int kInferMaxLds = 160 * 1024;

struct NetInferArgs {
    const double *bsave, *bold, *cu, *q_spe, *q_beam;
    const double *part;
    int nsp;
    double *scales;
    double *seed;          // seed: components 0, 1 of nodes 1..nr ([2][nm][ld])
    double oscale;
    int nr, nm, ld, cin, layers, cout, store, stride, nwb;
    float rinv;
    const __bf16 *wb;      // every layer's BF16 weights
    long long wb_off[kNetMaxLayers];
    const float *bias[kNetMaxLayers];
    int cin_p[kNetMaxLayers], cout_p[kNetMaxLayers];
    __bf16 *h[kNetMaxLayers];
    float *z[kNetMaxLayers];
};

inline size_t infer_lds_bytes(int nm, int stride, int nwb)
{
    return sizeof(double) * 4 * nm * kInferT + sizeof(__bf16) * (2 * kInferRows * stride + (size_t)nwb);
}

__global__ __launch_bounds__(kInferThreads) void k_net_infer(NetInferArgs a)
{
    extern __shared__ double lds_d[];
    __shared__ double ssc[4];
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63, j = lane & 15, q = lane >> 4;
    const int L = a.layers, T = kInferT, nr = a.nr, nm = a.nm, ld = a.ld, S = a.stride;
    const int p0 = blockIdx.x * T, pbase = p0 - L;
    double *bcache = lds_d;                                   // [2][2][nm][T]
    __bf16 *buf0 = (__bf16 *)(lds_d + 4 * nm * T);
    __bf16 *bufs[2] = {buf0, buf0 + kInferRows * S};
    __bf16 *wl = buf0 + 2 * kInferRows * S;
    // 1. requests: weights, input fields of every (row, channel group) item,
    //    the FP64 B_perp of the owned rows, the scales' partial maxima
    bf16x8 wv[kInferWVec];
    const int nwv = a.nwb >> 3;
#pragma unroll
    for (int u = 0; u < kInferWVec; u++) {
        const int v = tid + u * kInferThreads;
        if (v < nwv) wv[u] = ((const bf16x8 *)a.wb)[v];
    }
    const int ng = a.cin_p[0] >> 3, nrow = T + 2 * L;
    double xa[kInferItems][8], xb[kInferItems][8];
    int sidx[kInferItems][8];
    unsigned kind[kInferItems];   // per channel 2 bits: 0 zero, 1 field, 2 radius
#pragma unroll
    for (int it = 0; it < kInferItems; it++) {
        const int idx = tid + it * kInferThreads;
        const int g = idx / nrow, r = idx % nrow, p = pbase + r;
        const bool use = idx < nrow * ng && p >= 0 && p < nr;
        const int i = use ? p + 1 : 1;
        kind[it] = 0;
        // (quantity, mode array) of channel 8 g, then stepped per channel
        int qc = (8 * g) / nm, kc = (8 * g) % nm;
#pragma unroll
        for (int e = 0; e < 8; e++) {
            const int ch = 8 * g + e;
            const bool field = use && ch < kNetQty * nm;
            const int qty = field ? qc : 0, k = field ? kc : 0;
            if (++kc == nm) {
                kc = 0;
                qc++;
            }
            const int comp = qty & 1 & (qty < 6);
            const double *arr = qty < 4 ? a.bsave : (qty < 6 ? a.cu : (qty == 6 ? a.q_spe : a.q_beam));
            const int off = (comp * nm + k) * ld + i;
            xa[it][e] = arr[off];
            xb[it][e] = (qty == 2 || qty == 3) ? a.bold[off] : 0.0;
            sidx[it][e] = qty < 4 ? 0 : (qty < 6 ? 1 : qty - 4);
            kind[it] |= (field ? 1u : (use && ch == kNetQty * nm ? 2u : 0u)) << (2 * e);
        }
    }
    static_assert((kInferT & (kInferT - 1)) == 0, "kInferT must be a power of two");
    for (int v = tid; v < 4 * nm * T; v += kInferThreads) {
        // v = (which 2 nm + c nm + k) T + r; c nm + k is the row of component c, mode array k
        const int r = v & (T - 1), row = v / T, which = row >= 2 * nm, crow = which ? row - 2 * nm : row;
        const int p = p0 + r;
        const double *src = which ? a.bold : a.bsave;
        bcache[v] = p < nr ? src[crow * ld + p + 1] : 0.0;
    }
    net_final_scales(a.part, a.nsp, a.scales, ssc);
    // 2. LDS: weights, zeroed activations
#pragma unroll
    for (int u = 0; u < kInferWVec; u++) {
        const int v = tid + u * kInferThreads;
        if (v < nwv) ((bf16x8 *)wl)[v] = wv[u];
    }
    {
        bf16x8 zero;
#pragma unroll
        for (int e = 0; e < 8; e++) zero[e] = (__bf16)0.0f;
        for (int v = tid; v < 2 * kInferRows * S / 8; v += kInferThreads) ((bf16x8 *)buf0)[v] = zero;
    }
    __syncthreads();
    // 3. the inputs, scaled
    {
        const double inv[4] = {1.0 / ssc[0], 1.0 / ssc[1], 1.0 / ssc[2], 1.0 / ssc[3]};
#pragma unroll
        for (int it = 0; it < kInferItems; it++) {
            const int idx = tid + it * kInferThreads;
            const int g = idx / nrow, r = idx % nrow, p = pbase + r;
            if (idx >= nrow * ng || p < 0 || p >= nr) continue;
            bf16x8 v;
#pragma unroll
            for (int e = 0; e < 8; e++) {
                const unsigned kd = (kind[it] >> (2 * e)) & 3u;
                const float x = kd == 1u ? (float)((xa[it][e] - xb[it][e]) * inv[sidx[it][e]])
                                         : (kd == 2u ? (float)p * a.rinv : 0.0f);
                v[e] = (__bf16)x;
            }
            *(bf16x8 *)(bufs[0] + r * S + 8 * g) = v;
            if (a.store && r >= L && r < L + T) *(bf16x8 *)(a.h[0] + (size_t)(p + 1) * a.cin_p[0] + 8 * g) = v;
        }
    }
    __syncthreads();
    // 4. the layers
    for (int l = 0; l < L; l++) {
        const __bf16 *in = bufs[l & 1];
        __bf16 *out = bufs[(l + 1) & 1];
        const int cin_p = a.cin_p[l], cout_p = a.cout_p[l], nt = cout_p >> 4, nkb = cin_p >> 5;
        const bool last = l == L - 1;
        // rows l + 1 .. T + 2L - l - 2 are exact at this layer
        const int r_lo = l + 1, nrt = (T + 2 * L - 2 * l - 2 + 15) >> 4;
        // each wave owns output tile t (nt = 2 or 4 tiles: 4 or 2 waves per
        // tile) with its weights, every tap and k-block, in registers
        const int t = wave % nt, wstep = (kInferThreads / 64) / nt;
        const __bf16 *W = wl + a.wb_off[l];
        bf16x8 av[3][kNetMaxCinP / 32];
#pragma unroll
        for (int tap = 0; tap < 3; tap++) {
            const __bf16 *wrow = W + (tap * cout_p + 16 * t + j) * cin_p + 8 * q;
#pragma unroll
            for (int kb = 0; kb < kNetMaxCinP / 32; kb++)
                if (kb < nkb) av[tap][kb] = *(const bf16x8 *)(wrow + 32 * kb);
        }
        f32x4 bv4;
#pragma unroll
        for (int rr = 0; rr < 4; rr++) bv4[rr] = a.bias[l][16 * t + 4 * q + rr];
        for (int rt = wave / nt; rt < nrt; rt += wstep) {
            const int r0 = r_lo + 16 * rt;
            f32x4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
            for (int tap = 0; tap < 3; tap++) {
                const __bf16 *hrow = in + (r0 + j + tap - 1) * S + 8 * q;
#pragma unroll
                for (int kb = 0; kb < kNetMaxCinP / 32; kb++)
                    if (kb < nkb) acc = mfma_bf16(av[tap][kb], *(const bf16x8 *)(hrow + 32 * kb), acc);
            }
            const int r = r0 + j, p = pbase + r, co0 = 16 * t + 4 * q;
            const bool inside = p >= 0 && p < nr;
            const bool own = inside && r >= L && r < L + T;
            f32x4 zz;
#pragma unroll
            for (int rr = 0; rr < 4; rr++) zz[rr] = acc[rr] + bv4[rr];
            if (last && own) {
                if (a.store) *(f32x4 *)(a.z[l] + (size_t)p * cout_p + co0) = zz;
                const double sc = ssc[0] * a.oscale;
                const int rl = r - L;
#pragma unroll
                for (int rr = 0; rr < 4; rr++) {
                    // output channel co = cc nm + k: B_perp component cc, mode array k
                    const int co = co0 + rr;
                    if (co < a.cout) {
                        const double b = bcache[co * T + rl];
                        const double bo = bcache[(2 * nm + co) * T + rl];
                        a.seed[(size_t)co * ld + p + 1] = b + (b - bo) + (double)zz[rr] * sc;
                    }
                }
            } else if (!last) {
                bf16x4 hv;
#pragma unroll
                for (int rr = 0; rr < 4; rr++) hv[rr] = inside ? (__bf16)(zz[rr] * net_sigmoid(zz[rr])) : (__bf16)0.0f;
                *(bf16x4 *)(out + r * S + co0) = hv;
                if (a.store && own) {
                    *(f32x4 *)(a.z[l] + (size_t)p * cout_p + co0) = zz;
                    *(bf16x4 *)(a.h[l + 1] + (size_t)(p + 1) * cout_p + co0) = hv;
                }
            }
        }
        __syncthreads();
    }
}

// host inputs x [npos][cin] -> BF16 rows (qpad_seednet_eval)
__global__ __launch_bounds__(256) void k_net_pack_host(const float *x, __bf16 *h0, int npos, int cin, int cin_p)
{
    const int idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= npos * cin_p) return;
    const int pos = idx / cin_p, c = idx % cin_p;
    h0[(size_t)(pos + 1) * cin_p + c] = (__bf16)(c < cin ? x[(size_t)pos * cin + c] : 0.0f);
}

// ---------------------------------------------------------------------------
// forward: one wave per 16 positions, all output channels
// ---------------------------------------------------------------------------
struct NetFwdArgs {
    const __bf16 *W;      // [3][cout_p][cin_p]
    const float *bias;    // [cout_p]
    const __bf16 *h_in;   // [npos_pad + 2][cin_p]
    float *z;             // [npos_pad][cout_p]
    __bf16 *h_out;        // [npos_pad + 2][cout_p] (hidden layers)
    int cin_p, cout_p, npos;
};

template <int NT, bool LAST>
__global__ __launch_bounds__(64) void k_net_fwd(NetFwdArgs a)
{
    const int lane = threadIdx.x, j = lane & 15, q = lane >> 4;
    const int pos = blockIdx.x * 16 + j;
    const int cin_p = a.cin_p, cout_p = a.cout_p;
    f32x4 acc[NT];
#pragma unroll
    for (int t = 0; t < NT; t++) acc[t] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
    const int nkb = cin_p >> 5;
    for (int tap = 0; tap < 3; tap++) {
        const __bf16 *hrow = a.h_in + (size_t)(pos + tap) * cin_p + 8 * q;
        const __bf16 *wrow = a.W + ((size_t)tap * cout_p + j) * cin_p + 8 * q;
        for (int kb = 0; kb < nkb; kb++) {
            const bf16x8 bv = *(const bf16x8 *)(hrow + 32 * kb);
#pragma unroll
            for (int t = 0; t < NT; t++) {
                const bf16x8 av = *(const bf16x8 *)(wrow + (size_t)16 * t * cin_p + 32 * kb);
                acc[t] = mfma_bf16(av, bv, acc[t]);
            }
        }
    }
    if (pos >= a.npos) return;
#pragma unroll
    for (int t = 0; t < NT; t++) {
        const int co0 = 16 * t + 4 * q;
        f32x4 zz;
#pragma unroll
        for (int r = 0; r < 4; r++) zz[r] = acc[t][r] + a.bias[co0 + r];
        *(f32x4 *)(a.z + (size_t)pos * cout_p + co0) = zz;
        if (!LAST) {
            bf16x4 hv;
#pragma unroll
            for (int r = 0; r < 4; r++) hv[r] = (__bf16)(zz[r] * net_sigmoid(zz[r]));
            *(bf16x4 *)(a.h_out + (size_t)(pos + 1) * cout_p + co0) = hv;
        }
    }
}

// ---------------------------------------------------------------------------
// loss: normalised target, d loss / d y, per-block sums of squares
// ---------------------------------------------------------------------------
struct NetLossArgs {
    const float *y;        // [npos_pad][cout_p]
    __bf16 *dz;            // [npos_pad + 2][cout_p]
    float *part;           // per-block sums of squared errors
    const float *tin;      // [npos][cout] host target, or null: from the slice
    const double *b_spe, *bsave, *bold, *scales;
    double oscale;
    int npos, cout, cout_p, nm, ld;
};

__global__ __launch_bounds__(kNetLossThreads) void k_net_loss(NetLossArgs a)
{
    __shared__ float red[kNetLossThreads / 64];
    const int idx = blockIdx.x * blockDim.x + threadIdx.x;
    const int pos = idx / a.cout_p, co = idx % a.cout_p;
    float d = 0.0f;
    if (pos < a.npos && co < a.cout) {
        float t;
        if (a.tin) {
            t = a.tin[(size_t)pos * a.cout + co];
        } else {
            // converged B_perp minus its first-order extrapolation
            // 2 B(xi-1) - B(xi-2), in units of oscale times the B_perp scale,
            // clipped: the correction stays within that bound, and slices
            // where B_perp first appears (scale near round-off) cannot
            // dominate the gradient
            const int nm = a.nm, ld = a.ld, c = co / nm, k = co % nm, i = pos + 1;
            const double dt = (NX(a.b_spe, c, k, i) - 2.0 * NX(a.bsave, c, k, i) + NX(a.bold, c, k, i)) /
                              (a.scales[0] * a.oscale);
            t = (float)fmin(fmax(dt, -1.0), 1.0);
        }
        d = a.y[(size_t)pos * a.cout_p + co] - t;
    }
    if (pos < a.npos)
        a.dz[(size_t)(pos + 1) * a.cout_p + co] = (__bf16)(2.0f * d / (float)((long long)a.npos * a.cout));
    float s = d * d;
    for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        float t = 0.0f;
        for (int w = 0; w < kNetLossThreads / 64; w++) t += red[w];
        a.part[blockIdx.x] = t;
    }
}

// ---------------------------------------------------------------------------
// backward through layer l: d loss / d z of layer l - 1
//   dh[p][ci] = sum_tap sum_co W[tap][co][ci] dz[p - tap + 1][co]
//   dz_prev   = dh * silu'(z_prev)
// ---------------------------------------------------------------------------
struct NetBwdArgs {
    const __bf16 *WT;      // layer l, [3][cin_p][cout_p]
    const __bf16 *dz;      // layer l, [npos_pad + 2][cout_p]
    const float *z_prev;   // layer l - 1, [npos_pad][cin_p]
    __bf16 *dz_prev;       // layer l - 1, [npos_pad + 2][cin_p]
    int cin_p, cout_p, npos;
};

template <int NT>
__global__ __launch_bounds__(64) void k_net_bwd(NetBwdArgs a)
{
    const int lane = threadIdx.x, j = lane & 15, q = lane >> 4;
    const int pos = blockIdx.x * 16 + j;
    const int cin_p = a.cin_p, cout_p = a.cout_p;
    f32x4 acc[NT];
#pragma unroll
    for (int t = 0; t < NT; t++) acc[t] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
    const int nkb = cout_p >> 5;
    for (int tap = 0; tap < 3; tap++) {
        // row of position pos - tap + 1
        const __bf16 *drow = a.dz + (size_t)(pos + 2 - tap) * cout_p + 8 * q;
        const __bf16 *wrow = a.WT + ((size_t)tap * cin_p + j) * cout_p + 8 * q;
        for (int kb = 0; kb < nkb; kb++) {
            const bf16x8 bv = *(const bf16x8 *)(drow + 32 * kb);
#pragma unroll
            for (int t = 0; t < NT; t++) {
                const bf16x8 av = *(const bf16x8 *)(wrow + (size_t)16 * t * cout_p + 32 * kb);
                acc[t] = mfma_bf16(av, bv, acc[t]);
            }
        }
    }
    if (pos >= a.npos) return;
#pragma unroll
    for (int t = 0; t < NT; t++) {
        const int ci0 = 16 * t + 4 * q;
        const f32x4 zp = *(const f32x4 *)(a.z_prev + (size_t)pos * cin_p + ci0);
        bf16x4 dv;
#pragma unroll
        for (int r = 0; r < 4; r++) {
            const float s = net_sigmoid(zp[r]);
            dv[r] = (__bf16)(acc[t][r] * (s * (1.0f + zp[r] * (1.0f - s))));
        }
        *(bf16x4 *)(a.dz_prev + (size_t)(pos + 1) * cin_p + ci0) = dv;
    }
}

// ---------------------------------------------------------------------------
// weight gradients of every layer over one chunk of positions
//   dW[tap][co][ci] = sum_p dz[p][co] h[p + tap - 1][ci],  db[co] = sum_p dz[p][co]
// ---------------------------------------------------------------------------
struct NetWgradLayer {
    const __bf16 *h;       // [npos_pad + 2][cin_p]
    const __bf16 *dz;      // [npos_pad + 2][cout_p]
    int cin_p, cout_p;
    long long w_off, b_off;
};

struct NetWgradArgs {
    NetWgradLayer L[kNetMaxLayers];
    float *part;           // [nchunk][nparams]
    long long nparams;
};

__global__ __launch_bounds__(256) void k_net_wgrad(NetWgradArgs a)
{
    __shared__ __bf16 sH[(kNetChunk + 2) * kNetMaxCinP];
    __shared__ __bf16 sD[kNetChunk * kNetMaxCoutP];
    const NetWgradLayer L = a.L[blockIdx.y];
    const int cin_p = L.cin_p, cout_p = L.cout_p;
    const int c0 = blockIdx.x * kNetChunk;
    const int tid = threadIdx.x;
    // rows c0 .. c0 + chunk + 1 of h (positions c0 - 1 .. c0 + chunk) and
    // rows c0 + 1 .. c0 + chunk of dz
    const int hv = (kNetChunk + 2) * cin_p / 8, dv = kNetChunk * cout_p / 8;
    const bf16x8 *gh = (const bf16x8 *)(L.h + (size_t)c0 * cin_p);
    const bf16x8 *gd = (const bf16x8 *)(L.dz + (size_t)(c0 + 1) * cout_p);
    for (int v = tid; v < hv; v += 256) ((bf16x8 *)sH)[v] = gh[v];
    for (int v = tid; v < dv; v += 256) ((bf16x8 *)sD)[v] = gd[v];
    __syncthreads();
    float *P = a.part + (size_t)blockIdx.x * a.nparams;
    if (tid < cout_p) {
        float s = 0.0f;
        for (int p = 0; p < kNetChunk; p++) s += (float)sD[p * cout_p + tid];
        P[L.b_off + tid] = s;
    }
    const int wave = tid >> 6, lane = tid & 63, j = lane & 15, q = lane >> 4;
    const int nco = cout_p >> 4, nci = cin_p >> 4, ntile = 3 * nco * nci;
    for (int tile = wave; tile < ntile; tile += 4) {
        const int tap = tile / (nco * nci), rem = tile % (nco * nci), t = rem / nci, u = rem % nci;
        f32x4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
        for (int ks = 0; ks < kNetChunk / 32; ks++) {
            bf16x8 av, bv;
#pragma unroll
            for (int e = 0; e < 8; e++) {
                const int pl = 32 * ks + 8 * q + e;
                av[e] = sD[pl * cout_p + 16 * t + j];
                bv[e] = sH[(pl + tap) * cin_p + 16 * u + j];
            }
            acc = mfma_bf16(av, bv, acc);
        }
#pragma unroll
        for (int r = 0; r < 4; r++) {
            const int co = 16 * t + 4 * q + r, ci = 16 * u + j;
            P[L.w_off + ((size_t)tap * cout_p + co) * cin_p + ci] = acc[r];
        }
    }
}

// ---------------------------------------------------------------------------
// gradient reduction over chunks (fixed order), Adam, BF16 weight copies
// ---------------------------------------------------------------------------
constexpr int kAdamGrad = 1, kAdamUpdate = 2, kAdamRefresh = 4;

struct NetAdamArgs {
    float *theta, *m, *v;
    const float *part;
    float *grad_out;
    __bf16 *wb, *wtb;
    const float *loss_part;
    float *loss_slot;
    long long nparams;
    int nchunk, flags, layers, nloss;
    float lr, b1, b2, eps, ic1, ic2, loss_norm;
    long long w_off[kNetMaxLayers], b_off[kNetMaxLayers], wb_off[kNetMaxLayers];
    int cin_p[kNetMaxLayers], cout_p[kNetMaxLayers];
};

__global__ __launch_bounds__(256) void k_net_adam(NetAdamArgs a)
{
    __shared__ float red[256];
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (blockIdx.x == 0 && a.loss_slot) {
        // sum of the loss partials: per-thread strided sums, then a tree
        float s = 0.0f;
        for (int b = threadIdx.x; b < a.nloss; b += 256) s += a.loss_part[b];
        red[threadIdx.x] = s;
        __syncthreads();
        for (int w = 128; w > 0; w >>= 1) {
            if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
            __syncthreads();
        }
        if (threadIdx.x == 0) *a.loss_slot = red[0] * a.loss_norm;
    }
    if (i >= a.nparams) return;
    float th = a.theta[i];
    if (a.flags & (kAdamGrad | kAdamUpdate)) {
        // chunks in order, loads issued eight at a time
        float g = 0.0f;
        for (int c0 = 0; c0 < a.nchunk; c0 += 8) {
            float v[8];
#pragma unroll
            for (int u = 0; u < 8; u++) v[u] = c0 + u < a.nchunk ? a.part[(size_t)(c0 + u) * a.nparams + i] : 0.0f;
#pragma unroll
            for (int u = 0; u < 8; u++) g += v[u];
        }
        if (a.flags & kAdamGrad) a.grad_out[i] = g;
        if (a.flags & kAdamUpdate) {
            const float mm = a.b1 * a.m[i] + (1.0f - a.b1) * g;
            const float vv = a.b2 * a.v[i] + (1.0f - a.b2) * g * g;
            a.m[i] = mm;
            a.v[i] = vv;
            th -= a.lr * (mm * a.ic1) / (sqrtf(vv * a.ic2) + a.eps);
            a.theta[i] = th;
        }
    }
    if (a.flags & (kAdamUpdate | kAdamRefresh)) {
        for (int l = 0; l < a.layers; l++) {
            const long long o = i - a.w_off[l];
            if (o < 0 || i >= a.b_off[l]) continue;
            const int cin_p = a.cin_p[l], cout_p = a.cout_p[l];
            const int tap = (int)(o / ((long long)cout_p * cin_p));
            const int rem = (int)(o % ((long long)cout_p * cin_p));
            const int co = rem / cin_p, ci = rem % cin_p;
            a.wb[a.wb_off[l] + o] = (__bf16)th;
            a.wtb[a.wb_off[l] + ((size_t)tap * cin_p + ci) * cout_p + co] = (__bf16)th;
        }
    }
}

#undef NX

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
static int net_fail(int code, const std::string &msg)
{
    set_error(msg);
    return code;
}

template <class T>
static int net_alloc(T **p, size_t n)
{
    *p = nullptr;
    if (n == 0) n = 1;
    hipError_t e = hipMalloc((void **)p, n * sizeof(T));
    if (e != hipSuccess) {
        *p = nullptr;
        return net_fail(QPAD_E_NOMEM, std::string("hipMalloc failed: ") + hipGetErrorString(e));
    }
    return 0;
}

template <class T>
static void net_free(T *&p)
{
    if (p) (void)hipFree((void *)p);
    p = nullptr;
}

void seednet_destroy(qpad_ctx *c)
{
    SeedNet *n = c->net;
    if (!n) return;
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    if (n->stream) (void)hipStreamSynchronize(n->stream);
    net_free(n->seed_buf);
    if (n->ev_main) (void)hipEventDestroy(n->ev_main);
    if (n->ev_net) (void)hipEventDestroy(n->ev_net);
    if (n->stream) (void)hipStreamDestroy(n->stream);
    c->net_pre = 0;
    net_free(n->theta); net_free(n->am); net_free(n->av); net_free(n->part); net_free(n->grad);
    net_free(n->wb); net_free(n->wtb);
    for (int l = 0; l < kNetMaxLayers; l++) { net_free(n->h[l]); net_free(n->z[l]); net_free(n->dz[l]); }
    net_free(n->scales); net_free(n->scale_part); net_free(n->bprev); net_free(n->loss_part); net_free(n->loss_log);
    net_free(n->xin); net_free(n->tin);
    delete n;
    c->net = nullptr;
}

static inline int round_up(int x, int m) { return (x + m - 1) / m * m; }

static uint64_t splitmix64(uint64_t &s)
{
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

// He-uniform weights over the used channels (fan-in 3 cin); the output layer
// starts at zero, so an untrained network seeds with the first-order
// extrapolation in xi; zero biases
static void net_init_params(const SeedNet &n, std::vector<float> &th)
{
    th.assign((size_t)n.nparams, 0.0f);
    uint64_t s = n.cfg.seed;
    for (int l = 0; l < n.layers; l++) {
        const int cin = l == 0 ? n.cin : n.cfg.width;
        const int cout = l == n.layers - 1 ? n.cout : n.cfg.width;
        const double a = l == n.layers - 1 ? 0.0 : std::sqrt(6.0 / (3.0 * cin));
        for (int tap = 0; tap < 3; tap++)
            for (int co = 0; co < cout; co++)
                for (int ci = 0; ci < cin; ci++) {
                    const double u = (double)(splitmix64(s) >> 11) * (1.0 / 9007199254740992.0);
                    th[(size_t)(n.w_off[l] + ((long long)tap * n.cout_p[l] + co) * n.cin_p[l] + ci)] =
                        (float)((2.0 * u - 1.0) * a);
                }
    }
}

static int net_adam_launch(qpad_ctx *c, hipStream_t st, int flags, float *grad_out, float *loss_slot, int nloss,
                           float loss_norm)
{
    SeedNet &n = *c->net;
    NetAdamArgs a;
    memset(&a, 0, sizeof(a));
    a.theta = n.theta;
    a.m = n.am;
    a.v = n.av;
    a.part = n.part;
    a.grad_out = grad_out;
    a.wb = n.wb;
    a.wtb = n.wtb;
    a.loss_part = n.loss_part;
    a.loss_slot = loss_slot;
    a.nparams = n.nparams;
    a.nchunk = n.nchunk;
    a.flags = flags;
    a.layers = n.layers;
    a.nloss = nloss;
    a.loss_norm = loss_norm;
    a.lr = n.cfg.lr;
    a.b1 = n.cfg.beta1;
    a.b2 = n.cfg.beta2;
    a.eps = n.cfg.eps;
    if (flags & kAdamUpdate) {
        const double t = (double)n.steps;
        a.ic1 = (float)(1.0 / (1.0 - std::pow((double)n.cfg.beta1, t)));
        a.ic2 = (float)(1.0 / (1.0 - std::pow((double)n.cfg.beta2, t)));
    }
    for (int l = 0; l < n.layers; l++) {
        a.w_off[l] = n.w_off[l];
        a.b_off[l] = n.b_off[l];
        a.wb_off[l] = n.wb_off[l];
        a.cin_p[l] = n.cin_p[l];
        a.cout_p[l] = n.cout_p[l];
    }
    const unsigned nb = (unsigned)((n.nparams + 255) / 256);
    hipLaunchKernelGGL(k_net_adam, dim3(nb), dim3(256), 0, st, a);
    QPAD_HIP_CHECK(hipGetLastError());
    return 0;
}

#define NET_FWD(NT, LAST) hipLaunchKernelGGL((k_net_fwd<NT, LAST>), grid, dim3(64), 0, st, a)

// layer-by-layer forward pass over npos positions (training without the
// seed, qpad_seednet_eval)
static int net_forward(qpad_ctx *c, hipStream_t st, int npos)
{
    SeedNet &n = *c->net;
    const dim3 grid((unsigned)(round_up(npos, 16) / 16));
    for (int l = 0; l < n.layers; l++) {
        const bool last = l == n.layers - 1;
        NetFwdArgs a;
        memset(&a, 0, sizeof(a));
        a.W = n.wb + n.wb_off[l];
        a.bias = n.theta + n.b_off[l];
        a.h_in = n.h[l];
        a.z = n.z[l];
        a.h_out = last ? nullptr : n.h[l + 1];
        a.cin_p = n.cin_p[l];
        a.cout_p = n.cout_p[l];
        a.npos = npos;
        const int nt = n.cout_p[l] / 16;
        if (last) {
            if (nt == 2) NET_FWD(2, true);
            else NET_FWD(4, true);
        } else {
            if (nt == 2) NET_FWD(2, false);
            else NET_FWD(4, false);
        }
        QPAD_HIP_CHECK(hipGetLastError());
    }
    return 0;
}

#undef NET_FWD

// loss (host target tin or the converged slice), backward, weight gradients
// and the Adam step (flags: kAdamGrad / kAdamUpdate)
static int net_backward(qpad_ctx *c, hipStream_t st, int npos, const float *tin, int flags, float *grad_out)
{
    SeedNet &n = *c->net;
    const int L = n.layers;
    {
        NetLossArgs a;
        memset(&a, 0, sizeof(a));
        a.y = n.z[L - 1];
        a.dz = n.dz[L - 1];
        a.part = n.loss_part;
        a.tin = tin;
        a.b_spe = c->f[QPAD_F_B_SPE];
        a.bsave = n.bprev + (size_t)(n.packed_slice & 1) * 2 * c->g.nm * c->g.ld;
        a.bold = n.bprev + (size_t)((n.packed_slice - 1) & 1) * 2 * c->g.nm * c->g.ld;
        a.scales = n.scales;
        a.oscale = n.cfg.out_scale;
        a.npos = npos;
        a.cout = n.cout;
        a.cout_p = n.cout_p[L - 1];
        a.nm = c->g.nm;
        a.ld = c->g.ld;
        const int nb = (npos * a.cout_p + kNetLossThreads - 1) / kNetLossThreads;
        hipLaunchKernelGGL(k_net_loss, dim3(nb), dim3(kNetLossThreads), 0, st, a);
        QPAD_HIP_CHECK(hipGetLastError());
        n.nloss_blk = nb;
    }
    const dim3 grid((unsigned)(round_up(npos, 16) / 16));
    for (int l = L - 1; l >= 1; l--) {
        NetBwdArgs a;
        a.WT = n.wtb + n.wb_off[l];
        a.dz = n.dz[l];
        a.z_prev = n.z[l - 1];
        a.dz_prev = n.dz[l - 1];
        a.cin_p = n.cin_p[l];
        a.cout_p = n.cout_p[l];
        a.npos = npos;
        if (n.cin_p[l] / 16 == 2) hipLaunchKernelGGL(k_net_bwd<2>, grid, dim3(64), 0, st, a);
        else hipLaunchKernelGGL(k_net_bwd<4>, grid, dim3(64), 0, st, a);
        QPAD_HIP_CHECK(hipGetLastError());
    }
    {
        NetWgradArgs a;
        memset(&a, 0, sizeof(a));
        for (int l = 0; l < L; l++) {
            a.L[l].h = n.h[l];
            a.L[l].dz = n.dz[l];
            a.L[l].cin_p = n.cin_p[l];
            a.L[l].cout_p = n.cout_p[l];
            a.L[l].w_off = n.w_off[l];
            a.L[l].b_off = n.b_off[l];
        }
        a.part = n.part;
        a.nparams = n.nparams;
        hipLaunchKernelGGL(k_net_wgrad, dim3((unsigned)n.nchunk, (unsigned)L), dim3(256), 0, st, a);
        QPAD_HIP_CHECK(hipGetLastError());
    }
    float *slot = nullptr;
    if (flags & kAdamUpdate) {
        n.steps++;
        slot = n.loss_log + (size_t)((n.steps - 1) % kNetLossLog);
    }
    return net_adam_launch(c, st, flags, grad_out, slot, n.nloss_blk, 1.0f / (float)((long long)npos * n.cout));
}

// k_net_apply: the seed into b_spe (fine-grained form)
__global__ __launch_bounds__(256) void k_net_apply(const double *seed, double *b_spe, int nr, int nm, int ld)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x + 1, k = blockIdx.y;
    if (i > nr) return;
    b_spe[(size_t)k * ld + i] = seed[(size_t)k * ld + i];
    b_spe[((size_t)nm + k) * ld + i] = seed[((size_t)nm + k) * ld + i];
}

// the inputs of slice j from the fields as slice j - 1 left them (b_spe =
// B_perp(xi-1), cu = the J predictor of slice j, q_spe of slice j - 1,
// q_beam of slice j): scales and saved B_perp, then in seed mode the fused
// inference into seed_buf (keeping the activations when training),
// otherwise the packed inputs for the training step's forward pass.
// q_prev: q_spe of slice j - 1; the fused slice passes its history copy,
// because the next slice's start overwrites the live q_spe while these
// kernels may still run
static int queue_slice_inputs(qpad_ctx *c, hipStream_t st, int j, const double *q_prev)
{
    SeedNet *n = c->net;
    const int nr = c->g.nr, nm = c->g.nm, ld = c->g.ld;
    const size_t half = (size_t)2 * nm * ld;
    const double *q_beam = c->h_q_beam + (size_t)j * nm * ld;
    double *bsave = n->bprev + (size_t)(j & 1) * half;
    const double *bold = n->bprev + (size_t)((j - 1) & 1) * half;
    {
        NetScaleArgs a;
        a.b_spe = c->f[QPAD_F_B_SPE];
        a.cu = c->f[QPAD_F_CU];
        a.q_spe = q_prev;
        a.q_beam = q_beam;
        a.bsave = bsave;
        a.part = n->scale_part;
        a.nr = nr;
        a.nm = nm;
        a.ld = ld;
        hipLaunchKernelGGL(k_net_scale, dim3((unsigned)n->nscale_blk), dim3(256), 0, st, a);
        QPAD_HIP_CHECK(hipGetLastError());
    }
    n->packed_slice = j;
    n->fwd_done = false;
    if (n->seed_on) {
        NetInferArgs a;
        memset(&a, 0, sizeof(a));
        a.bsave = bsave;
        a.bold = bold;
        a.cu = c->f[QPAD_F_CU];
        a.q_spe = q_prev;
        a.q_beam = q_beam;
        a.part = n->scale_part;
        a.nsp = n->nscale_blk;
        a.scales = n->scales;
        a.seed = n->seed_buf;
        a.oscale = n->cfg.out_scale;
        a.nr = nr;
        a.nm = nm;
        a.ld = ld;
        a.cin = n->cin;
        a.layers = n->layers;
        a.cout = n->cout;
        a.store = n->train_on;
        a.stride = n->infer_stride;
        a.nwb = (int)n->nwb;
        a.rinv = 1.0f / (float)nr;
        a.wb = n->wb;
        for (int l = 0; l < n->layers; l++) {
            a.wb_off[l] = n->wb_off[l];
            a.bias[l] = n->theta + n->b_off[l];
            a.cin_p[l] = n->cin_p[l];
            a.cout_p[l] = n->cout_p[l];
            a.h[l] = n->h[l];
            a.z[l] = n->z[l];
        }
        const size_t lds = infer_lds_bytes(nm, n->infer_stride, (int)n->nwb);
        hipLaunchKernelGGL(k_net_infer, dim3((unsigned)((nr + kInferT - 1) / kInferT)), dim3(kInferThreads), lds,
                           st, a);
        QPAD_HIP_CHECK(hipGetLastError());
        n->fwd_done = true;
    } else {
        NetPackArgs a;
        a.bsave = bsave;
        a.bold = bold;
        a.cu = c->f[QPAD_F_CU];
        a.q_spe = q_prev;
        a.q_beam = q_beam;
        a.part = n->scale_part;
        a.nsp = n->nscale_blk;
        a.scales = n->scales;
        a.h0 = n->h[0];
        a.nr = nr;
        a.nm = nm;
        a.ld = ld;
        a.cin = n->cin;
        a.cin_p = n->cin_p[0];
        a.rinv = 1.0f / (float)nr;
        const int tot = (n->cin_p[0] / 8) * nr;
        hipLaunchKernelGGL(k_net_pack, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, st, a);
        QPAD_HIP_CHECK(hipGetLastError());
    }
    return 0;
}

// one Adam step on the converged slice j (its inputs were queued)
static int queue_slice_training(qpad_ctx *c, hipStream_t st, int j)
{
    SeedNet *n = c->net;
    if (n->packed_slice != j) return net_fail(QPAD_E_STATE, "seed network: inputs of this slice were not packed");
    if (!n->fwd_done) QPAD_TRY(net_forward(c, st, c->g.nr));
    QPAD_TRY(net_backward(c, st, c->g.nr, nullptr, kAdamUpdate, nullptr));
    n->packed_slice = 0;
    return 0;
}

int seednet_active(const qpad_ctx *c) { return c->net && (c->net->seed_on || c->net->train_on); }

int seednet_mark(qpad_ctx *c)
{
    QPAD_HIP_CHECK(hipEventRecord(c->net->ev_main, c->stream));
    return 0;
}

int seednet_after_epilogue(qpad_ctx *c, int j, int next, bool marked)
{
    SeedNet *n = c->net;
    if (!marked) QPAD_HIP_CHECK(hipEventRecord(n->ev_main, c->stream));
    QPAD_HIP_CHECK(hipStreamWaitEvent(n->stream, n->ev_main, 0));
    if (j > 0 && n->train_on) QPAD_TRY(queue_slice_training(c, n->stream, j));
    c->net_pre = 0;
    if (next > 0) {
        // q_spe of slice next - 1 from its history copy (written by that
        // slice's epilogue; zeroed with every history at the step start)
        const size_t q_off = (size_t)(next - 1) * c->g.nm * c->g.ld;
        QPAD_TRY(queue_slice_inputs(c, n->stream, next, c->h_q_spe + q_off));
        c->net_pre = next;
    }
    QPAD_HIP_CHECK(hipEventRecord(n->ev_net, n->stream));
    return 0;
}

int seednet_before_sources(qpad_ctx *c, int j, const double **seed)
{
    SeedNet *n = c->net;
    if (c->net_pre != j) return net_fail(QPAD_E_STATE, "seed network: slice inputs not queued");
    QPAD_HIP_CHECK(hipStreamWaitEvent(c->stream, n->ev_net, 0));
    *seed = n->seed_on ? n->seed_buf : nullptr;
    c->net_pre = 0;
    return 0;
}

// the host waits for the network's stream (its last work may still read
// b_spe or write its own buffers)
int seednet_quiesce(qpad_ctx *c)
{
    if (c->net && c->net->stream) QPAD_HIP_CHECK(hipStreamSynchronize(c->net->stream));
    return 0;
}

// b_spe is zeroed at the 3D-step start: so are the saved copies
int seednet_step_begin(qpad_ctx *c)
{
    SeedNet *n = c->net;
    if (!n) return 0;
    QPAD_TRY(seednet_quiesce(c));
    QPAD_HIP_CHECK(hipMemsetAsync(n->bprev, 0, sizeof(double) * 4 * c->g.nm * c->g.ld, c->stream));
    n->packed_slice = 0;
    c->net_pre = 0;
    return 0;
}

}  // namespace qpad

using namespace qpad;

extern "C" {

int qpad_seednet_create(qpad_ctx *c, const qpad_seednet_cfg *cfg)
{
    if (!c || !cfg) return net_fail(QPAD_E_INVALID, "null argument");
    if (c->subcyc) return net_fail(QPAD_E_STATE, "the seed network does not step sub-cycled slices");
    if (cfg->width != 32 && cfg->width != 64) return net_fail(QPAD_E_INVALID, "seed network width must be 32 or 64");
    if (cfg->layers < 2 || cfg->layers > kNetMaxLayers)
        return net_fail(QPAD_E_INVALID, "seed network layers must be 2..8");
    if (!(cfg->lr > 0.0f) || !(cfg->beta1 >= 0.0f && cfg->beta1 < 1.0f) || !(cfg->beta2 >= 0.0f && cfg->beta2 < 1.0f) ||
        !(cfg->eps > 0.0f))
        return net_fail(QPAD_E_INVALID, "bad Adam settings");
    if (!(cfg->out_scale > 0.0f)) return net_fail(QPAD_E_INVALID, "seed network out_scale must be > 0");
    seednet_destroy(c);
    SeedNet *n = new SeedNet();
    memset(n, 0, sizeof(*n));
    c->net = n;
    n->cfg = *cfg;
    n->layers = cfg->layers;
    n->cin = kNetQty * c->g.nm + 1;
    n->cout = 2 * c->g.nm;
    n->npos = c->g.nr;
    n->npos_pad = round_up(c->g.nr, kNetChunk);
    n->nchunk = n->npos_pad / kNetChunk;
    long long off = 0, woff = 0;
    for (int l = 0; l < n->layers; l++) {
        n->cin_p[l] = l == 0 ? round_up(n->cin, 32) : cfg->width;
        n->cout_p[l] = l == n->layers - 1 ? round_up(n->cout, 32) : cfg->width;
        if (n->cin_p[l] > kNetMaxCinP || n->cout_p[l] > kNetMaxCoutP) {
            seednet_destroy(c);
            return net_fail(QPAD_E_INVALID, "seed network channels exceed the kernel limits");
        }
        n->w_off[l] = off;
        off += 3LL * n->cout_p[l] * n->cin_p[l];
        n->b_off[l] = off;
        off += n->cout_p[l];
        n->wb_off[l] = woff;
        woff += 3LL * n->cout_p[l] * n->cin_p[l];
    }
    n->nparams = off;
    n->nwb = woff;
    {
        int cmax = cfg->width;
        for (int l = 0; l < n->layers; l++) cmax = std::max(cmax, n->cin_p[l]);
        n->infer_stride = cmax + 8;
        const size_t lds = infer_lds_bytes(c->g.nm, n->infer_stride, (int)n->nwb);
        if (n->nwb > kInferMaxW || lds > (size_t)kInferMaxLds ||
            kInferT + 2 * n->layers + 16 > kInferRows ||
            (size_t)(kInferT + 2 * n->layers) * (n->cin_p[0] / 8) > (size_t)kInferItems * kInferThreads) {
            seednet_destroy(c);
            return net_fail(QPAD_E_INVALID, "seed network too large for the fused inference kernel");
        }
        hipError_t e = hipFuncSetAttribute((const void *)k_net_infer, hipFuncAttributeMaxDynamicSharedMemorySize,
                                           (int)lds);
        if (e != hipSuccess) {
            seednet_destroy(c);
            return net_fail(QPAD_E_HIP, std::string("hipFuncSetAttribute: ") + hipGetErrorString(e));
        }
    }
    int rc = 0;
    const size_t rows = (size_t)n->npos_pad + 2;
    if (!rc) rc = net_alloc(&n->theta, (size_t)n->nparams);
    if (!rc) rc = net_alloc(&n->am, (size_t)n->nparams);
    if (!rc) rc = net_alloc(&n->av, (size_t)n->nparams);
    if (!rc) rc = net_alloc(&n->grad, (size_t)n->nparams);
    if (!rc) rc = net_alloc(&n->part, (size_t)n->nchunk * n->nparams);
    if (!rc) rc = net_alloc(&n->wb, (size_t)n->nwb);
    if (!rc) rc = net_alloc(&n->wtb, (size_t)n->nwb);
    for (int l = 0; l < n->layers && !rc; l++) {
        rc = net_alloc(&n->h[l], rows * n->cin_p[l]);
        if (!rc) rc = net_alloc(&n->z[l], (size_t)n->npos_pad * n->cout_p[l]);
        if (!rc) rc = net_alloc(&n->dz[l], rows * n->cout_p[l]);
    }
    if (!rc) rc = net_alloc(&n->scales, 4);
    n->nscale_blk = std::max(1, std::min(64, (c->g.nm * c->g.nr + 255) / 256));
    if (!rc) rc = net_alloc(&n->scale_part, (size_t)4 * n->nscale_blk);
    if (!rc) rc = net_alloc(&n->bprev, (size_t)4 * c->g.nm * c->g.ld);
    if (!rc) rc = net_alloc(&n->loss_part, (size_t)(rows * 32 / kNetLossThreads + 1));
    if (!rc) rc = net_alloc(&n->loss_log, (size_t)kNetLossLog);
    if (!rc) rc = net_alloc(&n->seed_buf, (size_t)2 * c->g.nm * c->g.ld);
    if (!rc && hipStreamCreateWithFlags(&n->stream, hipStreamNonBlocking) != hipSuccess)
        rc = net_fail(QPAD_E_HIP, "hipStreamCreate failed");
    if (!rc && (hipEventCreateWithFlags(&n->ev_main, hipEventDisableTiming) != hipSuccess ||
                hipEventCreateWithFlags(&n->ev_net, hipEventDisableTiming) != hipSuccess))
        rc = net_fail(QPAD_E_HIP, "hipEventCreate failed");
    if (!rc) rc = net_alloc(&n->xin, (size_t)n->npos * n->cin);
    if (!rc) rc = net_alloc(&n->tin, (size_t)n->npos * n->cout);
    if (rc) {
        seednet_destroy(c);
        return rc;
    }
    // guard rows of the activations stay zero (conv padding)
    for (int l = 0; l < n->layers; l++) {
        QPAD_HIP_CHECK(hipMemsetAsync(n->h[l], 0, rows * n->cin_p[l] * sizeof(__bf16), c->stream));
        QPAD_HIP_CHECK(hipMemsetAsync(n->dz[l], 0, rows * n->cout_p[l] * sizeof(__bf16), c->stream));
    }
    QPAD_HIP_CHECK(hipMemsetAsync(n->bprev, 0, sizeof(double) * 4 * c->g.nm * c->g.ld, c->stream));
    QPAD_HIP_CHECK(hipMemsetAsync(n->seed_buf, 0, sizeof(double) * 2 * c->g.nm * c->g.ld, c->stream));
    QPAD_HIP_CHECK(hipMemsetAsync(n->loss_log, 0, sizeof(float) * kNetLossLog, c->stream));
    std::vector<float> th;
    net_init_params(*n, th);
    return qpad_seednet_set_params(c, th.data(), n->nparams);
}

int qpad_seednet_get_info(qpad_ctx *c, qpad_seednet_info *out)
{
    if (!c || !out) return net_fail(QPAD_E_INVALID, "null argument");
    if (!c->net) return net_fail(QPAD_E_STATE, "no seed network (qpad_seednet_create)");
    QPAD_TRY(seednet_quiesce(c));
    const SeedNet &n = *c->net;
    memset(out, 0, sizeof(*out));
    out->layers = n.layers;
    out->cin = n.cin;
    out->cout = n.cout;
    out->npos = n.npos;
    for (int l = 0; l < n.layers; l++) {
        out->cin_p[l] = n.cin_p[l];
        out->cout_p[l] = n.cout_p[l];
    }
    out->nparams = n.nparams;
    out->steps = n.steps;
    return 0;
}

int qpad_seednet_mode(qpad_ctx *c, int32_t seed, int32_t train)
{
    if (!c) return net_fail(QPAD_E_INVALID, "null context");
    if (!c->net) return net_fail(QPAD_E_STATE, "no seed network (qpad_seednet_create)");
    QPAD_TRY(seednet_quiesce(c));
    c->net->seed_on = seed != 0;
    c->net->train_on = train != 0;
    c->net->packed_slice = 0;
    c->net_pre = 0;
    return 0;
}

int qpad_seednet_get_params(qpad_ctx *c, float *theta, int64_t n)
{
    if (!c || !theta) return net_fail(QPAD_E_INVALID, "null argument");
    if (!c->net) return net_fail(QPAD_E_STATE, "no seed network (qpad_seednet_create)");
    QPAD_TRY(seednet_quiesce(c));
    if (n != c->net->nparams) return net_fail(QPAD_E_INVALID, "parameter count mismatch");
    QPAD_HIP_CHECK(hipMemcpyAsync(theta, c->net->theta, sizeof(float) * (size_t)n, hipMemcpyDeviceToHost, c->stream));
    QPAD_HIP_CHECK(hipStreamSynchronize(c->stream));
    return 0;
}

// new parameters restart Adam (moments and step count)
int qpad_seednet_set_params(qpad_ctx *c, const float *theta, int64_t n)
{
    if (!c || !theta) return net_fail(QPAD_E_INVALID, "null argument");
    if (!c->net) return net_fail(QPAD_E_STATE, "no seed network (qpad_seednet_create)");
    QPAD_TRY(seednet_quiesce(c));
    SeedNet &s = *c->net;
    if (n != s.nparams) return net_fail(QPAD_E_INVALID, "parameter count mismatch");
    QPAD_HIP_CHECK(hipMemcpyAsync(s.theta, theta, sizeof(float) * (size_t)n, hipMemcpyHostToDevice, c->stream));
    QPAD_HIP_CHECK(hipMemsetAsync(s.am, 0, sizeof(float) * (size_t)n, c->stream));
    QPAD_HIP_CHECK(hipMemsetAsync(s.av, 0, sizeof(float) * (size_t)n, c->stream));
    s.steps = 0;
    QPAD_TRY(net_adam_launch(c, c->stream, kAdamRefresh, nullptr, nullptr, 0, 0.0f));
    QPAD_HIP_CHECK(hipStreamSynchronize(c->stream));
    return 0;
}

int qpad_seednet_losses(qpad_ctx *c, int64_t first, int64_t n, float *out)
{
    if (!c || !out) return net_fail(QPAD_E_INVALID, "null argument");
    if (!c->net) return net_fail(QPAD_E_STATE, "no seed network (qpad_seednet_create)");
    QPAD_TRY(seednet_quiesce(c));
    const SeedNet &s = *c->net;
    if (first < 0 || n < 0 || first + n > s.steps || s.steps - first > kNetLossLog)
        return net_fail(QPAD_E_INVALID, "loss range not held");
    std::vector<float> all(kNetLossLog);
    QPAD_HIP_CHECK(hipMemcpyAsync(all.data(), s.loss_log, sizeof(float) * kNetLossLog, hipMemcpyDeviceToHost,
                                  c->stream));
    QPAD_HIP_CHECK(hipStreamSynchronize(c->stream));
    for (int64_t k = 0; k < n; k++) out[k] = all[(size_t)((first + k) % kNetLossLog)];
    return 0;
}

int qpad_seednet_seed(qpad_ctx *c, int32_t j)
{
    if (!c) return net_fail(QPAD_E_INVALID, "null context");
    if (!c->net) return net_fail(QPAD_E_STATE, "no seed network (qpad_seednet_create)");
    if (j < 1 || j > c->g.nxi) return net_fail(QPAD_E_INVALID, "slice index out of range");
    SeedNet &n = *c->net;
    QPAD_TRY(seednet_quiesce(c));
    QPAD_TRY(queue_slice_inputs(c, c->stream, j, c->f[QPAD_F_Q_SPE]));
    if (n.seed_on) {
        const dim3 grid((unsigned)((c->g.nr + 255) / 256), (unsigned)c->g.nm);
        hipLaunchKernelGGL(k_net_apply, grid, dim3(256), 0, c->stream, n.seed_buf, c->f[QPAD_F_B_SPE], c->g.nr,
                           c->g.nm, c->g.ld);
        QPAD_HIP_CHECK(hipGetLastError());
    }
    QPAD_HIP_CHECK(hipStreamSynchronize(c->stream));
    return 0;
}

int qpad_seednet_learn(qpad_ctx *c, int32_t j)
{
    if (!c) return net_fail(QPAD_E_INVALID, "null context");
    if (!c->net) return net_fail(QPAD_E_STATE, "no seed network (qpad_seednet_create)");
    if (!c->net->train_on) return net_fail(QPAD_E_STATE, "seed network is not in train mode");
    QPAD_TRY(seednet_quiesce(c));
    QPAD_TRY(queue_slice_training(c, c->stream, j));
    QPAD_HIP_CHECK(hipStreamSynchronize(c->stream));
    return 0;
}

int qpad_seednet_eval(qpad_ctx *c, const float *x, const float *t, int32_t npos, float *y, float *loss, float *grad,
                      int32_t update)
{
    if (!c || !x) return net_fail(QPAD_E_INVALID, "null argument");
    if (!c->net) return net_fail(QPAD_E_STATE, "no seed network (qpad_seednet_create)");
    SeedNet &n = *c->net;
    if (npos < 1 || npos > n.npos) return net_fail(QPAD_E_INVALID, "npos must be 1..nr");
    if ((loss || grad || update) && !t) return net_fail(QPAD_E_INVALID, "loss, gradient and update need a target");
    QPAD_TRY(seednet_quiesce(c));
    const int L = n.layers;
    const size_t rows = (size_t)n.npos_pad + 2;
    // rows past npos must read as zero padding
    for (int l = 0; l < L; l++) {
        QPAD_HIP_CHECK(hipMemsetAsync(n.h[l], 0, rows * n.cin_p[l] * sizeof(__bf16), c->stream));
        QPAD_HIP_CHECK(hipMemsetAsync(n.dz[l], 0, rows * n.cout_p[l] * sizeof(__bf16), c->stream));
    }
    QPAD_HIP_CHECK(hipMemcpyAsync(n.xin, x, sizeof(float) * (size_t)npos * n.cin, hipMemcpyHostToDevice, c->stream));
    const int tot = npos * n.cin_p[0];
    hipLaunchKernelGGL(k_net_pack_host, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, c->stream, n.xin, n.h[0],
                       npos, n.cin, n.cin_p[0]);
    QPAD_HIP_CHECK(hipGetLastError());
    QPAD_TRY(net_forward(c, c->stream, npos));
    if (t) {
        QPAD_HIP_CHECK(
            hipMemcpyAsync(n.tin, t, sizeof(float) * (size_t)npos * n.cout, hipMemcpyHostToDevice, c->stream));
        const int flags = (grad ? kAdamGrad : 0) | (update ? kAdamUpdate : 0);
        QPAD_TRY(net_backward(c, c->stream, npos, n.tin, flags, grad ? n.grad : nullptr));
    }
    if (y) {
        const int cp = n.cout_p[L - 1];
        std::vector<float> yy((size_t)npos * cp);
        QPAD_HIP_CHECK(hipMemcpyAsync(yy.data(), n.z[L - 1], sizeof(float) * yy.size(), hipMemcpyDeviceToHost,
                                      c->stream));
        QPAD_HIP_CHECK(hipStreamSynchronize(c->stream));
        for (int p = 0; p < npos; p++)
            for (int k = 0; k < n.cout; k++) y[(size_t)p * n.cout + k] = yy[(size_t)p * cp + k];
    }
    if (grad)
        QPAD_HIP_CHECK(hipMemcpyAsync(grad, n.grad, sizeof(float) * (size_t)n.nparams, hipMemcpyDeviceToHost,
                                      c->stream));
    if (loss && t) {
        std::vector<float> lp((size_t)n.nloss_blk);
        QPAD_HIP_CHECK(hipMemcpyAsync(lp.data(), n.loss_part, sizeof(float) * lp.size(), hipMemcpyDeviceToHost,
                                      c->stream));
        QPAD_HIP_CHECK(hipStreamSynchronize(c->stream));
        double s = 0.0;
        for (float v : lp) s += v;
        *loss = (float)(s / ((double)npos * n.cout));
    }
    QPAD_HIP_CHECK(hipStreamSynchronize(c->stream));
    return 0;
}

}  // extern "C"
