// k_denoise.hip -- the residual denoiser: a three-level U-Net on Gaussian-subtracted charge planes,
// trained and run on the device with bf16 MFMA products and fp32 accumulation.
//
// The network has no counterpart in the reference (it is this library's learned extension of the
// plain PIC path). Layout: planes are NHWC (plane, row, column, channel), activations bf16, all
// gradients fp32 (rounded to bf16 only where they enter an MFMA product), parameters fp32 with
// bf16 copies for the products, Adam in fp32.
//
// Layers (3x3, zero padding 1, ReLU):  e1a 2->16, e1b 16->16 | pool | e2a 16->32, e2b 32->32 | pool |
// ba 32->64, bb 64->64 | up, concat skip | d2a 96->32, d2b 32->32 | up, concat skip | d1a 48->16,
// d1b 16->16 | head 1x1 16->1 (linear). 118129 parameters. The network input's two channels are
// stored padded to eight (zero channels 2..7) so that every product reads 16-byte channel groups.
//
// Every 3x3 product is one implicit GEMM kernel (k_conv): rows = pixels, columns = output channels,
// reduction = 9 x input channels, each wave a 16-pixel x (16 NT)-channel tile of
// v_mfma_f32_16x16x16_bf16 (A: lane l holds row l%16, reduction indices 4(l/16)..4(l/16)+3;
// B: column l%16, the same reduction indices; D: rows 4(l/16)+i, column l%16). The input
// gradient is the same kernel over the flipped, channel-transposed weights; the weight gradient
// (k_wgrad) reduces over pixels in fixed-order split partials, so training is deterministic.
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <string.h>

#include <cmath>
#include <string>
#include <vector>

#include "denoise.h"
#include "octopic.h"

namespace octopic {

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ uint16_t f2bf(float f) {   // round to nearest even
    uint32_t u = __float_as_uint(f);
    if ((u & 0x7f800000u) == 0x7f800000u && (u & 0x7fffffu)) return (uint16_t)((u >> 16) | 0x40);   // NaN stays NaN
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}
__device__ __forceinline__ float bf2f(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }
__device__ __forceinline__ bool failed(const ErrorInfo *err) { return *((volatile const int32_t *)&err->code) != 0; }

constexpr int kConvThreads = 256;   // 4 waves, 64 pixels per block

// ------------------------------------------------------------------------------------- products
struct ConvArgs {
    const uint16_t *x;      // [P][cin] bf16, P = B*H*W pixels, cin % 4 == 0
    const uint16_t *w;      // [cout][kpad] bf16, reduction index k = tap * cin + ci, zero for k >= 9 cin
    const float *bias;      // [cout] or null
    int B, H, W, cin, cout, kpad, relu;
    uint16_t *y;            // bf16 activation [P][cout], or null
    float *yf;              // fp32 output [P][cout], or null
    // two-source input (the decoder's upsampling and skip concatenation, never materialised): with x2
    // set, input channels [0, cd) are x at half resolution ([B][H/2][W/2][cd], nearest upsampling) and
    // channels [cd, cin) are x2 ([P][cin - cd])
    const uint16_t *x2;
    int cd;
    uint16_t *yq;           // 2x2 average pool of the bf16 activation [B][H/2][W/2][cout], or null (tile kernels)
};

// 8 input channels c8*8 .. c8*8+7 of pixel (b, yy, xx) as one 16-byte load (cd and cin - cd are multiples of 8)
__device__ __forceinline__ uint4 load_in8(const ConvArgs &a, int b, int yy, int xx, int c) {
    if (a.x2) {
        if (c < a.cd)
            return *reinterpret_cast<const uint4 *>(a.x + (((int64_t)b * (a.H >> 1) + (yy >> 1)) * (a.W >> 1) + (xx >> 1)) * a.cd + c);
        return *reinterpret_cast<const uint4 *>(a.x2 + (((int64_t)b * a.H + yy) * a.W + xx) * (a.cin - a.cd) + (c - a.cd));
    }
    return *reinterpret_cast<const uint4 *>(a.x + (((int64_t)b * a.H + yy) * a.W + xx) * a.cin + c);
}

template <int NT>
__global__ __launch_bounds__(kConvThreads) void k_conv(ConvArgs a) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int64_t P = (int64_t)a.B * a.H * a.W;
    const int64_t p0 = (int64_t)blockIdx.x * 64 + wave * 16;
    if (p0 >= P) return;
    const int64_t p = p0 + (lane & 15);
    const int HW = a.H * a.W;
    const int b = (int)(p / HW), rem = (int)(p - (int64_t)b * HW), py = rem / a.W, px = rem - py * a.W;
    const int kq = (lane >> 4) * 4;
    const int kmax = 9 * a.cin;
    const bool prow = p < P;
    f32x4 acc[NT];
#pragma unroll
    for (int t = 0; t < NT; ++t) acc[t] = f32x4{0.f, 0.f, 0.f, 0.f};
    for (int k0 = 0; k0 < a.kpad; k0 += 16) {
        const int k = k0 + kq;
        bf16x4 av = bf16x4{0, 0, 0, 0};
        if (prow && k < kmax) {
            const int tap = k / a.cin, ci = k - tap * a.cin;
            const int yy = py + tap / 3 - 1, xx = px + tap % 3 - 1;
            if (yy >= 0 && yy < a.H && xx >= 0 && xx < a.W)
                av = *reinterpret_cast<const bf16x4 *>(a.x + (((int64_t)b * a.H + yy) * a.W + xx) * a.cin + ci);
        }
#pragma unroll
        for (int t = 0; t < NT; ++t) {
            const bf16x4 bv = *reinterpret_cast<const bf16x4 *>(a.w + (int64_t)(t * 16 + (lane & 15)) * a.kpad + k);
            acc[t] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(av, bv, acc[t], 0, 0, 0);
        }
    }
#pragma unroll
    for (int t = 0; t < NT; ++t) {
        const int co = t * 16 + (lane & 15);
        const float bb = a.bias ? a.bias[co] : 0.f;
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const int64_t q = p0 + (lane >> 4) * 4 + i;
            if (q >= P) continue;
            float v = acc[t][i] + bb;
            if (a.relu) v = v > 0.f ? v : 0.f;
            if (a.y) a.y[q * a.cout + co] = f2bf(v);
            if (a.yf) a.yf[q * a.cout + co] = v;
        }
    }
}

// The same product for inputs of 8k channels, with the block's weights staged in LDS (rows padded
// by 8 to spread the banks) and v_mfma_f32_16x16x32_bf16: lane l loads 8 consecutive reduction
// indices 8(l/16)..8(l/16)+7 of its row (A) or column (B) in one 16-byte access; A and B use the same
// index assignment, so the sum is the full reduction whatever order the instruction takes it in.
// Each wave takes MT 16-pixel tiles, the block 64 MT pixels.
template <int NT, int MT>
__global__ __launch_bounds__(kConvThreads) void k_conv8(ConvArgs a) {
    extern __shared__ __align__(16) uint16_t wl[];
    const int ldw = a.kpad + 8;
    {
        const int per_row = a.kpad / 8;
        for (int e = threadIdx.x; e < a.cout * per_row; e += blockDim.x) {
            const int r = e / per_row, c8 = e - r * per_row;
            *reinterpret_cast<uint4 *>(wl + r * ldw + c8 * 8) = *reinterpret_cast<const uint4 *>(a.w + (int64_t)r * a.kpad + c8 * 8);
        }
    }
    __syncthreads();
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int64_t P = (int64_t)a.B * a.H * a.W;
    const int64_t base = (int64_t)blockIdx.x * (64 * MT) + wave * (16 * MT);
    if (base >= P) return;
    const int HW = a.H * a.W;
    int pb[MT], py[MT], px[MT];
    bool pv[MT];
#pragma unroll
    for (int mt = 0; mt < MT; ++mt) {
        const int64_t p = base + mt * 16 + (lane & 15);
        pv[mt] = p < P;
        const int64_t pp = pv[mt] ? p : 0;
        pb[mt] = (int)(pp / HW);
        const int rem = (int)(pp - (int64_t)pb[mt] * HW);
        py[mt] = rem / a.W;
        px[mt] = rem - py[mt] * a.W;
    }
    const int kq = (lane >> 4) * 8, kmax = 9 * a.cin;
    int tap = kq / a.cin, c = kq - tap * a.cin;
    f32x4 acc[MT][NT];
#pragma unroll
    for (int mt = 0; mt < MT; ++mt)
#pragma unroll
        for (int t = 0; t < NT; ++t) acc[mt][t] = f32x4{0.f, 0.f, 0.f, 0.f};
    const uint16_t *wrow = wl + (lane & 15) * ldw;
    // A fragments of reduction step k (this lane's 8 indices at tap, channel c), zero off the mesh
    auto load_a = [&](bf16x8 (&av)[MT], int k, int tp, int cc) {
        const bool kin = k < kmax;
        const int dy = tp / 3 - 1, dx = tp - (tp / 3) * 3 - 1;
#pragma unroll
        for (int mt = 0; mt < MT; ++mt) {
            const int yy = py[mt] + dy, xx = px[mt] + dx;
            if (kin && pv[mt] && yy >= 0 && yy < a.H && xx >= 0 && xx < a.W) {
                const uint4 u = load_in8(a, pb[mt], yy, xx, cc);
                av[mt] = *reinterpret_cast<const bf16x8 *>(&u);
            } else
                av[mt] = bf16x8{};
        }
    };
    bf16x8 cur[MT], nxt[MT];
    load_a(cur, kq, tap, c);
    for (int k0 = 0; k0 < a.kpad; k0 += 32) {
        c += 32;
        while (c >= a.cin) { c -= a.cin; ++tap; }
        if (k0 + 32 < a.kpad) load_a(nxt, k0 + 32 + kq, tap, c);   // next step's loads in flight during this step
        const int k = k0 + kq;
#pragma unroll
        for (int t = 0; t < NT; ++t) {
            const bf16x8 bv = *reinterpret_cast<const bf16x8 *>(wrow + t * 16 * ldw + k);
#pragma unroll
            for (int mt = 0; mt < MT; ++mt) acc[mt][t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(cur[mt], bv, acc[mt][t], 0, 0, 0);
        }
#pragma unroll
        for (int mt = 0; mt < MT; ++mt) cur[mt] = nxt[mt];
    }
#pragma unroll
    for (int mt = 0; mt < MT; ++mt)
#pragma unroll
        for (int t = 0; t < NT; ++t) {
            const int co = t * 16 + (lane & 15);
            const float bb = a.bias ? a.bias[co] : 0.f;
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const int64_t q = base + mt * 16 + (lane >> 4) * 4 + i;
                if (q >= P) continue;
                float v = acc[mt][t][i] + bb;
                if (a.relu) v = v > 0.f ? v : 0.f;
                if (a.y) a.y[q * a.cout + co] = f2bf(v);
                if (a.yf) a.yf[q * a.cout + co] = v;
            }
        }
}

// The same product on 16 x 16-pixel tiles of one plane: the tile and its one-pixel halo (all input
// channels, pixel stride cin + 8 to spread the banks) and the weights are staged in LDS once, so each
// input element is read from memory once instead of once per tap. Wave w takes tile rows 4w..4w+3
// (one 16-pixel MFMA row tile each). Grid: planes x ceil(H/16) x ceil(W/16).
constexpr int kTile = 16, kHalo = kTile + 2;

__host__ __device__ inline size_t conv_tile_lds(int cin, int cout, int kpad) {
    return (size_t)kHalo * kHalo * (cin + 8) * 2 + (size_t)cout * (kpad + 8) * 2;
}

// Staging of n 16-byte chunks into LDS, eight independent loads in flight per thread before their
// stores (a load-store pair per iteration would serialise on the memory latency).
template <class Src>
__device__ __forceinline__ void stage16(uint16_t *lds_base, int n, const Src &src) {
    for (int e0 = threadIdx.x; e0 < n; e0 += 8 * kConvThreads) {
        uint4 v[8];
        int dst[8];
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            const int e = e0 + j * kConvThreads;
            dst[j] = -1;
            if (e < n) v[j] = src(e, dst[j]);
        }
#pragma unroll
        for (int j = 0; j < 8; ++j)
            if (dst[j] >= 0) *reinterpret_cast<uint4 *>(lds_base + dst[j]) = v[j];
    }
}

template <int NT>
__global__ __launch_bounds__(kConvThreads) void k_conv_tile(ConvArgs a) {
    extern __shared__ __align__(16) uint16_t lds[];
    const int cs = a.cin + 8, ldw = a.kpad + 8;
    uint16_t *xt = lds, *wl = lds + kHalo * kHalo * cs;
    const int tiles_x = (a.W + kTile - 1) / kTile, tiles_y = (a.H + kTile - 1) / kTile;
    const int b = blockIdx.x / (tiles_x * tiles_y), tr = blockIdx.x - b * tiles_x * tiles_y;
    const int ty0 = (tr / tiles_x) * kTile, tx0 = (tr % tiles_x) * kTile;
    {
        const int per_row = a.kpad / 8;
        stage16(wl, a.cout * per_row, [&](int e, int &dst) {
            const int r = e / per_row, c8 = e - r * per_row;
            dst = r * ldw + c8 * 8;
            return *reinterpret_cast<const uint4 *>(a.w + (int64_t)r * a.kpad + c8 * 8);
        });
        const int g8 = a.cin / 8;
        stage16(xt, kHalo * kHalo * g8, [&](int e, int &dst) {
            const int pix = e / g8, c8 = e - pix * g8, hy = pix / kHalo, hx = pix - hy * kHalo;
            const int yy = ty0 + hy - 1, xx = tx0 + hx - 1;
            dst = pix * cs + c8 * 8;
            uint4 v = make_uint4(0u, 0u, 0u, 0u);
            if (yy >= 0 && yy < a.H && xx >= 0 && xx < a.W) v = load_in8(a, b, yy, xx, c8 * 8);
            return v;
        });
    }
    __syncthreads();
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, col = lane & 15;
    const int kq = (lane >> 4) * 8, kmax = 9 * a.cin;
    int tap = kq / a.cin, c = kq - tap * a.cin;
    f32x4 acc[4][NT];
#pragma unroll
    for (int mt = 0; mt < 4; ++mt)
#pragma unroll
        for (int t = 0; t < NT; ++t) acc[mt][t] = f32x4{0.f, 0.f, 0.f, 0.f};
    const uint16_t *wrow = wl + col * ldw;
    for (int k0 = 0; k0 < a.kpad; k0 += 32) {
        const int k = k0 + kq;
        bf16x8 av[4];
        if (k < kmax) {
            const int dy = tap / 3, dx = tap - (tap / 3) * 3;   // halo coordinates: +1 offset folded in
#pragma unroll
            for (int mt = 0; mt < 4; ++mt)
                av[mt] = *reinterpret_cast<const bf16x8 *>(xt + ((wave * 4 + mt + dy) * kHalo + col + dx) * cs + c);
        } else {
#pragma unroll
            for (int mt = 0; mt < 4; ++mt) av[mt] = bf16x8{};
        }
#pragma unroll
        for (int t = 0; t < NT; ++t) {
            const bf16x8 bv = *reinterpret_cast<const bf16x8 *>(wrow + t * 16 * ldw + k);
#pragma unroll
            for (int mt = 0; mt < 4; ++mt) acc[mt][t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av[mt], bv, acc[mt][t], 0, 0, 0);
        }
        c += 32;
        while (c >= a.cin) { c -= a.cin; ++tap; }
    }
    // activations: lane holds rows wave*4 + mt, columns 4(lane/16) + i of the tile, channel col (+16 t)
    uint16_t hv[4][NT][4];
#pragma unroll
    for (int mt = 0; mt < 4; ++mt) {
        const int yy = ty0 + wave * 4 + mt;
#pragma unroll
        for (int t = 0; t < NT; ++t) {
            const int co = t * 16 + col;
            const float bb = a.bias ? a.bias[co] : 0.f;
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const int xx = tx0 + (lane >> 4) * 4 + i;
                float v = acc[mt][t][i] + bb;
                if (a.relu) v = v > 0.f ? v : 0.f;
                hv[mt][t][i] = f2bf(v);
                if (yy >= a.H || xx >= a.W) continue;
                const int64_t q = ((int64_t)b * a.H + yy) * a.W + xx;
                if (a.y) a.y[q * a.cout + co] = hv[mt][t][i];
                if (a.yf) a.yf[q * a.cout + co] = v;
            }
        }
    }
    if (a.yq) {   // 2x2 average of the bf16 activations, summed as (top pair) + (bottom pair)
        const int h2 = a.H >> 1, w2 = a.W >> 1;
#pragma unroll
        for (int mp = 0; mp < 4; mp += 2) {
            const int y2 = (ty0 + wave * 4 + mp) >> 1;
            if (y2 >= h2) continue;
#pragma unroll
            for (int t = 0; t < NT; ++t)
#pragma unroll
                for (int i = 0; i < 4; i += 2) {
                    const int x2 = (tx0 + (lane >> 4) * 4 + i) >> 1;
                    if (x2 >= w2) continue;
                    const float sum = (bf2f(hv[mp][t][i]) + bf2f(hv[mp][t][i + 1])) + (bf2f(hv[mp + 1][t][i]) + bf2f(hv[mp + 1][t][i + 1]));
                    a.yq[(((int64_t)b * h2 + y2) * w2 + x2) * a.cout + t * 16 + col] = f2bf(0.25f * sum);
                }
        }
    }
}

// The same on 8 x 8-pixel tiles (for the coarse levels, where 16 x 16 tiles leave CUs idle): each of
// the 4 waves takes one 16-pixel MFMA row tile made of two 8-pixel rows (rows 2w, 2w + 1).
constexpr int kTile8 = 8, kHalo8 = kTile8 + 2;
// 16 x 16 tiles from this many tiles up, 8 x 8 below (measured on 90 planes of 64 x 64: the coarse
// levels run 11-13 us with 8 x 8 tiles against 16-22 us with 16 x 16; the full level prefers 16 x 16,
// 16.9 against 20 us)
constexpr int64_t kConvTile16From = 1024;

__host__ __device__ inline size_t conv_tile8_lds(int cin, int cout, int kpad) {
    return (size_t)kHalo8 * kHalo8 * (cin + 8) * 2 + (size_t)cout * (kpad + 8) * 2;
}

template <int NT>
__global__ __launch_bounds__(kConvThreads) void k_conv_tile8(ConvArgs a) {
    extern __shared__ __align__(16) uint16_t lds[];
    const int cs = a.cin + 8, ldw = a.kpad + 8;
    uint16_t *xt = lds, *wl = lds + kHalo8 * kHalo8 * cs;
    const int tiles_x = (a.W + kTile8 - 1) / kTile8, tiles_y = (a.H + kTile8 - 1) / kTile8;
    const int b = blockIdx.x / (tiles_x * tiles_y), tr = blockIdx.x - b * tiles_x * tiles_y;
    const int ty0 = (tr / tiles_x) * kTile8, tx0 = (tr % tiles_x) * kTile8;
    {
        const int per_row = a.kpad / 8;
        stage16(wl, a.cout * per_row, [&](int e, int &dst) {
            const int r = e / per_row, c8 = e - r * per_row;
            dst = r * ldw + c8 * 8;
            return *reinterpret_cast<const uint4 *>(a.w + (int64_t)r * a.kpad + c8 * 8);
        });
        const int g8 = a.cin / 8;
        stage16(xt, kHalo8 * kHalo8 * g8, [&](int e, int &dst) {
            const int pix = e / g8, c8 = e - pix * g8, hy = pix / kHalo8, hx = pix - hy * kHalo8;
            const int yy = ty0 + hy - 1, xx = tx0 + hx - 1;
            dst = pix * cs + c8 * 8;
            uint4 v = make_uint4(0u, 0u, 0u, 0u);
            if (yy >= 0 && yy < a.H && xx >= 0 && xx < a.W) v = load_in8(a, b, yy, xx, c8 * 8);
            return v;
        });
    }
    __syncthreads();
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int ar = wave * 2 + ((lane & 15) >> 3), ac = lane & 7;   // this lane's A row pixel in the tile
    const int kq = (lane >> 4) * 8, kmax = 9 * a.cin;
    int tap = kq / a.cin, c = kq - tap * a.cin;
    f32x4 acc[NT];
#pragma unroll
    for (int t = 0; t < NT; ++t) acc[t] = f32x4{0.f, 0.f, 0.f, 0.f};
    const uint16_t *wrow = wl + (lane & 15) * ldw;
    for (int k0 = 0; k0 < a.kpad; k0 += 32) {
        const int k = k0 + kq;
        bf16x8 av = bf16x8{};
        if (k < kmax) {
            const int dy = tap / 3, dx = tap - (tap / 3) * 3;
            av = *reinterpret_cast<const bf16x8 *>(xt + ((ar + dy) * kHalo8 + ac + dx) * cs + c);
        }
#pragma unroll
        for (int t = 0; t < NT; ++t) {
            const bf16x8 bv = *reinterpret_cast<const bf16x8 *>(wrow + t * 16 * ldw + k);
            acc[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av, bv, acc[t], 0, 0, 0);
        }
        c += 32;
        while (c >= a.cin) { c -= a.cin; ++tap; }
    }
#pragma unroll
    for (int t = 0; t < NT; ++t) {
        const int co = t * 16 + (lane & 15);
        const float bb = a.bias ? a.bias[co] : 0.f;
        uint16_t hv[4];
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const int pix = (lane >> 4) * 4 + i;   // output row of the 16-pixel tile
            const int yy = ty0 + wave * 2 + (pix >> 3), xx = tx0 + (pix & 7);
            float v = acc[t][i] + bb;
            if (a.relu) v = v > 0.f ? v : 0.f;
            hv[i] = f2bf(v);
            if (yy >= a.H || xx >= a.W) continue;
            const int64_t q = ((int64_t)b * a.H + yy) * a.W + xx;
            if (a.y) a.y[q * a.cout + co] = hv[i];
            if (a.yf) a.yf[q * a.cout + co] = v;
        }
        if (a.yq) {   // the row below sits in lane + 32 (pixels 8..15 of the tile); pairs i, i + 1 in this lane
            int below[4];
#pragma unroll
            for (int i = 0; i < 4; ++i) below[i] = __shfl_down((int)hv[i], 32, 64);
            if ((lane >> 4) < 2) {
                const int h2 = a.H >> 1, w2 = a.W >> 1;
                const int y2 = (ty0 + wave * 2) >> 1;
#pragma unroll
                for (int i = 0; i < 4; i += 2) {
                    const int x2 = (tx0 + (lane >> 4) * 4 + i) >> 1;
                    if (y2 >= h2 || x2 >= w2) continue;
                    const float sum = (bf2f(hv[i]) + bf2f(hv[i + 1])) + (bf2f((uint16_t)below[i]) + bf2f((uint16_t)below[i + 1]));
                    a.yq[(((int64_t)b * h2 + y2) * w2 + x2) * a.cout + co] = f2bf(0.25f * sum);
                }
            }
        }
    }
}

// Weight gradient partials: part[s][co][k] = sum over the pixels of split s of dy[p][co] x_k(p),
// with x_k(p) the im2col entry k = tap * cin + ci of pixel p. Grid (kpad/16, cout/16, splits);
// each of the 4 waves takes a quarter of the split's pixels, the block sums the waves in order.
struct WgradArgs {
    const uint16_t *x;      // [P][cin] bf16 (or the half-resolution first source, see ConvArgs)
    const uint16_t *x2;     // second source of a two-source input, or null
    int cd;
    const uint16_t *dy;     // [P][cout] bf16
    int B, H, W, cin, cout, kpad;
    int64_t split_pixels;   // pixels per split (multiple of 64)
    float *part;            // [splits][cout][kpad]
};

__global__ __launch_bounds__(kConvThreads) void k_wgrad(WgradArgs a) {
    __shared__ float red[4][16][17];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int kt = blockIdx.x, ct = blockIdx.y, s = blockIdx.z;
    const int64_t P = (int64_t)a.B * a.H * a.W;
    const int HW = a.H * a.W;
    const int co = ct * 16 + (lane & 15);
    const int k = kt * 16 + (lane & 15);
    const int kmax = 9 * a.cin;
    int tap = 0, ci = 0;
    const bool kvalid = k < kmax;
    if (kvalid) { tap = k / a.cin; ci = k - tap * a.cin; }
    const int dyo = tap / 3 - 1, dxo = tap % 3 - 1;
    int64_t pbeg = (int64_t)s * a.split_pixels, pend = pbeg + a.split_pixels;
    if (pend > P) pend = P;
    const int64_t per_wave = (pend - pbeg + 3) / 4;
    int64_t wb = pbeg + wave * per_wave, we = wb + per_wave;
    if (we > pend) we = pend;
    f32x4 acc = f32x4{0.f, 0.f, 0.f, 0.f};
    for (int64_t q0 = wb; q0 < we; q0 += 16) {
        bf16x4 av, bv;
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const int64_t p = q0 + (lane >> 4) * 4 + i;
            short va = 0, vb = 0;
            if (p < we) {
                va = (short)a.dy[p * a.cout + co];
                if (kvalid) {
                    const int b = (int)(p / HW), rem = (int)(p - (int64_t)b * HW), py = rem / a.W, px = rem - py * a.W;
                    const int yy = py + dyo, xx = px + dxo;
                    if (yy >= 0 && yy < a.H && xx >= 0 && xx < a.W) {
                        if (!a.x2)
                            vb = (short)a.x[(((int64_t)b * a.H + yy) * a.W + xx) * a.cin + ci];
                        else if (ci < a.cd)
                            vb = (short)a.x[(((int64_t)b * (a.H >> 1) + (yy >> 1)) * (a.W >> 1) + (xx >> 1)) * a.cd + ci];
                        else
                            vb = (short)a.x2[(((int64_t)b * a.H + yy) * a.W + xx) * (a.cin - a.cd) + (ci - a.cd)];
                    }
                }
            }
            av[i] = va;
            bv[i] = vb;
        }
        acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(av, bv, acc, 0, 0, 0);
    }
#pragma unroll
    for (int i = 0; i < 4; ++i) red[wave][(lane >> 4) * 4 + i][lane & 15] = acc[i];
    __syncthreads();
    if (threadIdx.x < 256) {
        const int r = threadIdx.x >> 4, c = threadIdx.x & 15;
        const float v = ((red[0][r][c] + red[1][r][c]) + red[2][r][c]) + red[3][r][c];
        a.part[((int64_t)s * a.cout + ct * 16 + r) * a.kpad + kt * 16 + c] = v;
    }
}

// dW (fp32, logical layout W[co][ky][kx][ci_logical]) = sum over splits in order.
__global__ void k_wgrad_reduce(const float *part, int splits, int cout, int kpad, int cin, int cin_logical, float *dw) {
    const int n = cout * 9 * cin_logical;
    for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < n; e += gridDim.x * blockDim.x) {
        const int co = e / (9 * cin_logical), r = e - co * 9 * cin_logical, tap = r / cin_logical, ci = r - tap * cin_logical;
        const int k = tap * cin + ci;
        float s = 0.f;
        for (int j = 0; j < splits; ++j) s += part[((int64_t)j * cout + co) * kpad + k];
        dw[e] = s;
    }
}

// ------------------------------------------------------------------------------- elementwise
// Network input [P][2] fp32 -> [P][8] bf16 (channels 2..7 zero), so the first product reads whole
// 16-byte groups like the others.
__global__ void k_pack_input(const float *in, int64_t P, uint16_t *x) {
    for (int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; p < P; p += (int64_t)gridDim.x * blockDim.x) {
        uint4 v;
        v.x = (uint32_t)f2bf(in[p * 2 + 0]) | ((uint32_t)f2bf(in[p * 2 + 1]) << 16);
        v.y = v.z = v.w = 0u;
        *reinterpret_cast<uint4 *>(x + p * 8) = v;
    }
}

// 2x2 average pooling, bf16 -> bf16 (sum in fp32, in a fixed order).
__global__ void k_pool(const uint16_t *x, int B, int H, int W, int C, uint16_t *y) {
    const int h = H / 2, w = W / 2;
    const int64_t n = (int64_t)B * h * w * C;
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < n; e += (int64_t)gridDim.x * blockDim.x) {
        const int c = (int)(e % C);
        const int64_t q = e / C;
        const int xx = (int)(q % w), yy = (int)((q / w) % h), b = (int)(q / ((int64_t)w * h));
        const int64_t base = (((int64_t)b * H + 2 * yy) * W + 2 * xx) * C + c;
        const float s = (bf2f(x[base]) + bf2f(x[base + C])) + (bf2f(x[base + (int64_t)W * C]) + bf2f(x[base + (int64_t)W * C + C]));
        y[e] = f2bf(0.25f * s);
    }
}

// Gradient through up+concat: dup [B][h][w][Cd] = 2x2 sums of d[..][c < Cd]; dskip [P][Cs] = d[..][c >= Cd].
__global__ void k_up_concat_back(const float *d, int B, int H, int W, int Cd, int Cs, float *dup, float *dskip) {
    const int C = Cd + Cs, h = H / 2, w = W / 2;
    const int64_t nup = (int64_t)B * h * w * Cd, nsk = (int64_t)B * H * W * Cs;
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < nup + nsk; e += (int64_t)gridDim.x * blockDim.x) {
        if (e < nup) {
            const int c = (int)(e % Cd);
            const int64_t q = e / Cd;
            const int xx = (int)(q % w), yy = (int)((q / w) % h), b = (int)(q / ((int64_t)w * h));
            const int64_t base = (((int64_t)b * H + 2 * yy) * W + 2 * xx) * C + c;
            dup[e] = (d[base] + d[base + C]) + (d[base + (int64_t)W * C] + d[base + (int64_t)W * C + C]);
        } else {
            const int64_t f = e - nup;
            dskip[f] = d[(f / Cs) * C + Cd + f % Cs];
        }
    }
}

// Gradient through the pooling, added to the skip gradient: out[p][c] = dskip[p][c] + dq[pool(p)][c] / 4.
__global__ void k_pool_back_add(const float *dq, const float *dskip, int B, int H, int W, int C, float *out) {
    const int64_t n = (int64_t)B * H * W * C;
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < n; e += (int64_t)gridDim.x * blockDim.x) {
        const int c = (int)(e % C);
        const int64_t q = e / C;
        const int xx = (int)(q % W), yy = (int)((q / W) % H), b = (int)(q / ((int64_t)W * H));
        out[e] = dskip[e] + 0.25f * dq[(((int64_t)b * (H / 2) + yy / 2) * (W / 2) + xx / 2) * C + c];
    }
}

// ReLU gradient: dpre = da * (act > 0), as bf16 for the products; per-block channel sums of the
// fp32 dpre for the bias gradient (block j sums pixels [64 j, 64 j + 64)).
__global__ __launch_bounds__(256) void k_relu_back(const float *da, const uint16_t *act, int64_t P, int C, uint16_t *dq,
                                                   float *bias_part) {
    __shared__ float red[256];
    const int64_t p0 = (int64_t)blockIdx.x * 64;
    const int c = threadIdx.x % C, lanes = 256 / C * C;
    float s = 0.f;
    if ((int)threadIdx.x < lanes) {
        for (int64_t p = p0 + threadIdx.x / C; p < p0 + 64 && p < P; p += 256 / C) {
            const int64_t e = p * C + c;
            const float v = bf2f(act[e]) > 0.f ? da[e] : 0.f;
            dq[e] = f2bf(v);
            s += v;
        }
    }
    red[threadIdx.x] = s;
    __syncthreads();
    if ((int)threadIdx.x < C) {
        float t = 0.f;
        for (int j = threadIdx.x; j < lanes; j += C) t += red[j];
        bias_part[(int64_t)blockIdx.x * C + threadIdx.x] = t;
    }
}

// Sum of per-block partials [nb][C] in block order.
__global__ void k_sum_parts(const float *part, int nb, int C, float *out) {
    for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < C; c += gridDim.x * blockDim.x) {
        float s = 0.f;
        for (int j = 0; j < nb; ++j) s += part[(int64_t)j * C + c];
        out[c] = s;
    }
}

// Head: out[p] = sum_c act[p][c] w[c] + b (fp32).
__global__ void k_head(const uint16_t *act, const float *w, const float *b, int64_t P, float *out) {
    for (int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; p < P; p += (int64_t)gridDim.x * blockDim.x) {
        float s = 0.f;
#pragma unroll
        for (int c = 0; c < 16; ++c) s += bf2f(act[p * 16 + c]) * w[c];
        out[p] = s + b[0];
    }
}

// Loss and head gradient. dout = 2 (out - t) / P; da[p][c] = dout w[c]; per-block partials of
// sum (out - t)^2, dout act[p][c] (c < 16) and dout (slot 16), 18 floats per block of 64 pixels.
__global__ __launch_bounds__(64) void k_head_back(const float *out, const float *target, const uint16_t *act, const float *w,
                                                  int64_t P, float *da, float *part) {
    __shared__ float red[18][64];
    const int64_t p = (int64_t)blockIdx.x * 64 + threadIdx.x;
    float loss = 0.f, dout = 0.f;
    float g[16];
    if (p < P) {
        const float diff = out[p] - target[p];
        loss = diff * diff;
        dout = 2.f * diff / (float)P;
    }
#pragma unroll
    for (int c = 0; c < 16; ++c) {
        const float av = p < P ? bf2f(act[p * 16 + c]) : 0.f;
        g[c] = dout * av;
        if (p < P) da[p * 16 + c] = dout * w[c];
    }
#pragma unroll
    for (int c = 0; c < 16; ++c) red[c][threadIdx.x] = g[c];
    red[16][threadIdx.x] = dout;
    red[17][threadIdx.x] = loss;
    __syncthreads();
    if (threadIdx.x < 18) {
        float s = 0.f;
        for (int j = 0; j < 64; ++j) s += red[threadIdx.x][j];
        part[(int64_t)blockIdx.x * 18 + threadIdx.x] = s;
    }
}

__global__ void k_loss_finish(const float *tot, int64_t P, float *loss) {
    if (threadIdx.x == 0 && blockIdx.x == 0) *loss = tot[17] / (float)P;
}

// Adam over the flat fp32 parameters; omb = 1 - beta and bc = 1 - beta^t, each rounded once from
// double on the host.
__global__ void k_adam(float *p, const float *g, float *m, float *v, int64_t n, float lr, float b1, float b2, float omb1,
                       float omb2, float bc1, float bc2, float eps) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const float gi = g[i];
        const float mi = b1 * m[i] + omb1 * gi;
        const float vi = b2 * v[i] + omb2 * gi * gi;
        m[i] = mi;
        v[i] = vi;
        const float mh = mi / bc1, vh = vi / bc2;
        p[i] = p[i] - lr * mh / (sqrtf(vh) + eps);
    }
}

// bf16 product weights of one layer from its fp32 parameters W[co][ky][kx][ci_logical]:
// fwd[co][k = tap cin + ci] (cin = stored input channels, zero beyond ci_logical and k >= 9 cin);
// bwd[ci][k' = tap' cout + co] = W[co][8 - tap'][ci] for the input gradient (cin_logical rows).
__global__ void k_pack_weights(const float *W, int cout, int cin, int cin_logical, int kpad, uint16_t *fwd, uint16_t *bwd) {
    const int kpadb = (9 * cout + 31) / 32 * 32;   // backward rows padded to a multiple of 32 with zeros
    const int nf = cout * kpad, nb = bwd ? cin_logical * kpadb : 0;
    for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < nf + nb; e += gridDim.x * blockDim.x) {
        if (e < nf) {
            const int co = e / kpad, k = e - co * kpad;
            float v = 0.f;
            if (k < 9 * cin) {
                const int tap = k / cin, ci = k - tap * cin;
                if (ci < cin_logical) v = W[((int64_t)co * 9 + tap) * cin_logical + ci];
            }
            fwd[e] = f2bf(v);
        } else {
            const int f = e - nf, ci = f / kpadb, r = f - ci * kpadb;
            if (r >= 9 * cout) {
                bwd[f] = 0;
            } else {
                const int tp = r / cout, co = r - tp * cout;
                bwd[f] = f2bf(W[((int64_t)co * 9 + (8 - tp)) * cin_logical + ci]);
            }
        }
    }
}

// ------------------------------------------------------------------------ residual deposits
// Node integral of the unit Gaussian (mu, sigma) against the CIC tent / TSC spline centred on
// node xi, in units where the cell size h and xi are given (gaussian_pic.jl:206-244).
__device__ double gauss_node_weight(double xi, double h, double mu, double s, int method) {
    const double invroot = 1.0 / (s * 1.4142135623730951), norm = 1.0 / (s * 2.5066282746310002);
    auto pdf = [&](double x) { return norm * exp(-((x - mu) * (x - mu)) / (2.0 * s * s)); };
    auto m0 = [&](double A, double B) { return 0.5 * (erf((B - mu) * invroot) - erf((A - mu) * invroot)); };
    auto m1 = [&](double A, double B, double c) { return (mu - c) * m0(A, B) - s * s * (pdf(B) - pdf(A)); };
    auto m2 = [&](double A, double B, double c) {
        return (s * s + (mu - c) * (mu - c)) * m0(A, B) - s * s * ((B - mu) * pdf(B) - (A - mu) * pdf(A)) -
               2.0 * (mu - c) * s * s * (pdf(B) - pdf(A));
    };
    if (method == 0) return m0(xi - h, xi + h) + (m1(xi - h, xi, xi) - m1(xi, xi + h, xi)) / h;
    const double half = 0.5 * h;
    const double cL = xi - half, cR = xi + half, lL = xi - 3.0 * half, lR = xi - half, rL = xi + half, rR = xi + 3.0 * half;
    return 0.75 * m0(cL, cR) - m2(cL, cR, xi) / (h * h) + 1.125 * (m0(lL, lR) + m0(rL, rR)) +
           1.5 * (m1(lL, lR, xi) - m1(rL, rR, xi)) / h + 0.5 * (m2(lL, lR, xi) + m2(rL, rR, xi)) / (h * h);
}

// The same node integrals on unit cells (node i at i) from tables of the erf and the pdf at the cell
// boundaries: E[k], G[k] at p_k = k - off (off = 1.5 for TSC, whose stencil ends at half cells; 1 for
// CIC), k = 0 .. n + 2, so that neighbouring nodes share their evaluations: the formulas of
// gauss_node_weight with h = 1, with one erf and one exp per boundary point instead of per use.
__device__ __forceinline__ double gauss_node_weight_tab(int i, double mu, double s, int method, const double *E, const double *G) {
    const double xi = (double)i, h = 1.0;
    const double off = method == 0 ? 1.0 : 1.5;
    auto m0 = [&](int ka, int kb) { return 0.5 * (E[kb] - E[ka]); };
    auto m1 = [&](int ka, int kb) { return (mu - xi) * m0(ka, kb) - s * s * (G[kb] - G[ka]); };
    auto m2 = [&](int ka, int kb) {
        const double A = (double)ka - off, B = (double)kb - off;
        return (s * s + (mu - xi) * (mu - xi)) * m0(ka, kb) - s * s * ((B - mu) * G[kb] - (A - mu) * G[ka]) -
               2.0 * (mu - xi) * s * s * (G[kb] - G[ka]);
    };
    if (method == 0)   // points i - 1, i, i + 1 at k = i, i + 1, i + 2
        return m0(i, i + 2) + (m1(i, i + 1) - m1(i + 1, i + 2)) / h;
    // points i - 1.5, i - 0.5, i + 0.5, i + 1.5 at k = i .. i + 3
    return 0.75 * m0(i + 1, i + 2) - m2(i + 1, i + 2) / (h * h) + 1.125 * (m0(i, i + 1) + m0(i + 2, i + 3)) +
           1.5 * (m1(i, i + 1) - m1(i + 2, i + 3)) / h + 0.5 * (m2(i, i + 1) + m2(i + 2, i + 3)) / (h * h);
}

// Counter-based stream of the pair generator: Philox4x32-10 of (particle, sample) under a key from
// (seed, stream), four 32-bit words per particle.
__device__ __forceinline__ void philox4(uint32_t c[4], uint32_t k0, uint32_t k1) {
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        const uint64_t p0 = (uint64_t)0xD2511F53u * c[0], p1 = (uint64_t)0xCD9E8D57u * c[2];
        const uint32_t n0 = (uint32_t)(p1 >> 32) ^ c[1] ^ k0, n1 = (uint32_t)p1, n2 = (uint32_t)(p0 >> 32) ^ c[3] ^ k1,
                       n3 = (uint32_t)p0;
        c[0] = n0; c[1] = n1; c[2] = n2; c[3] = n3;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
}
__host__ __device__ inline uint64_t gen_mix(uint64_t z) {
    z += 0x9e3779b97f4a7c15ull;
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}
__host__ __device__ inline double gen_uniform(uint64_t seed, int64_t sample, int j) {
    return ((double)(gen_mix(gen_mix(seed ^ 0x5851f42d4c957f2dull) + (uint64_t)sample * 64 + (uint64_t)j) >> 11) + 0.5) *
           (1.0 / 9007199254740992.0);
}

// The distribution of training sample `sample` (a two-component Gaussian mixture in cell units):
// core widths sx, sy in [1.5, 8] cells, shifted by a dipole mode of amplitude a_d in [0, 0.3] core
// sigmas and squeezed by a quadrupole mode a_q in [0, 0.3] (sx (1 + a_q), sy (1 - a_q), either sign);
// a centred halo of fraction f_h in [0, 0.2] and width 1.3-2 core sigmas; one sample in three is a
// plain Gaussian (no modes, no halo), the form most collision slices take; the mesh centre jittered
// by up to four cells; the noisy sample draws n_noisy in [16384, 262144] (log-uniform).
struct PairDist {
    double cx, cy, sx, sy, ox, oy, hx, hy, fh;
    int64_t n_noisy;
};
__host__ __device__ inline PairDist pair_dist(uint64_t seed, int64_t sample, int H, int W) {
    PairDist d;
    const bool plain = gen_uniform(seed, sample, 10) < 1.0 / 3.0;
    const double sx0 = 1.5 + 6.5 * gen_uniform(seed, sample, 0), sy0 = 1.5 + 6.5 * gen_uniform(seed, sample, 1);
    const double ad = plain ? 0.0 : 0.3 * gen_uniform(seed, sample, 2), th = 6.283185307179586 * gen_uniform(seed, sample, 3);
    const double aq = plain ? 0.0 : 0.3 * (2.0 * gen_uniform(seed, sample, 4) - 1.0);
    d.sx = sx0 * (1.0 + aq);
    d.sy = sy0 * (1.0 - aq);
    d.ox = ad * sx0 * cos(th);
    d.oy = ad * sy0 * sin(th);
    const double hk = 1.3 + 0.7 * gen_uniform(seed, sample, 5);
    d.hx = hk * sx0;
    d.hy = hk * sy0;
    d.fh = plain ? 0.0 : 0.2 * gen_uniform(seed, sample, 6);
    d.cx = 0.5 * (W - 1) + 4.0 * (2.0 * gen_uniform(seed, sample, 7) - 1.0);
    d.cy = 0.5 * (H - 1) + 4.0 * (2.0 * gen_uniform(seed, sample, 8) - 1.0);
    d.n_noisy = (int64_t)llround(exp(log(16384.0) + (log(262144.0) - log(16384.0)) * gen_uniform(seed, sample, 9)));
    return d;
}

constexpr double kPairWeightScale = 4294967296.0;   // deposit weights in units of 2^-32
constexpr double kPairMomentScale = 16777216.0;     // moment sums in units of 2^-24

struct PairArgs {
    uint64_t seed;
    int64_t first;          // first sample id
    int H, W, method;
    int64_t n_target;
    unsigned long long *planes;   // [batch][2][H*W] (noisy, converged), zero on entry
    unsigned long long *mom;      // [batch][2][5]: accepted count, sum u, sum v, sum u^2, sum v^2, zero on entry
    int chunks;                   // blocks per (sample, stream)
};

// Grid (chunks, batch, 2). u = column (x), v = row (y), cell units; particles whose stencil leaves
// the mesh are not deposited and not counted.
__global__ __launch_bounds__(256) void k_pair_deposit(PairArgs a) {
    extern __shared__ unsigned long long plane[];
    const int k = blockIdx.y, st = blockIdx.z, HW = a.H * a.W;
    const PairDist d = pair_dist(a.seed, a.first + k, a.H, a.W);
    const int64_t n = st == 0 ? d.n_noisy : a.n_target;
    for (int c = threadIdx.x; c < HW; c += blockDim.x) plane[c] = 0ull;
    __syncthreads();
    const uint64_t key = gen_mix(a.seed * 2 + (uint64_t)st + 0x632be59bd9b4e019ull);
    long long cnt = 0, su = 0, sv = 0, suu = 0, svv = 0;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)a.chunks * blockDim.x) {
        uint32_t c[4] = {(uint32_t)i, (uint32_t)(i >> 32), (uint32_t)(a.first + k), (uint32_t)((a.first + k) >> 32)};
        philox4(c, (uint32_t)key, (uint32_t)(key >> 32));
        const float u1 = ((float)(c[0] >> 8) + 0.5f) * (1.f / 16777216.f), u2 = ((float)(c[1] >> 8) + 0.5f) * (1.f / 16777216.f);
        const float u3 = ((float)(c[2] >> 8) + 0.5f) * (1.f / 16777216.f);
        const float r = sqrtf(-2.f * logf(u1));
        float sn, cs;
        sincosf(6.2831853f * u2, &sn, &cs);
        const float n1 = fminf(fmaxf(r * cs, -5.f), 5.f), n2 = fminf(fmaxf(r * sn, -5.f), 5.f);
        const bool halo = u3 < (float)d.fh;
        const float u = halo ? (float)d.cx + (float)d.hx * n1 : (float)(d.cx + d.ox) + (float)d.sx * n1;
        const float v = halo ? (float)d.cy + (float)d.hy * n2 : (float)(d.cy + d.oy) + (float)d.sy * n2;
        if (!(u >= 1.f && u < (float)(a.W - 2) && v >= 1.f && v < (float)(a.H - 2))) continue;
        // stencil weights (CIC: 2 nodes from floor(u); TSC: 3 nodes around the nearest node)
        float wu[3], wv[3];
        int bu, bv, np;
        if (a.method == 0) {
            bu = (int)floorf(u); bv = (int)floorf(v);
            const float fu = u - bu, fv = v - bv;
            wu[0] = 1.f - fu; wu[1] = fu; wu[2] = 0.f;
            wv[0] = 1.f - fv; wv[1] = fv; wv[2] = 0.f;
            np = 2;
        } else {
            const int iu = (int)floorf(u + 0.5f), iv = (int)floorf(v + 0.5f);
            const float du = u - iu, dv = v - iv;
            wu[0] = 0.5f * (0.5f - du) * (0.5f - du); wu[1] = 0.75f - du * du; wu[2] = 0.5f * (0.5f + du) * (0.5f + du);
            wv[0] = 0.5f * (0.5f - dv) * (0.5f - dv); wv[1] = 0.75f - dv * dv; wv[2] = 0.5f * (0.5f + dv) * (0.5f + dv);
            bu = iu - 1; bv = iv - 1;
            np = 3;
        }
        for (int jv = 0; jv < np; ++jv)
            for (int ju = 0; ju < np; ++ju) {
                const unsigned long long q = (unsigned long long)llrint((double)(wu[ju] * wv[jv]) * kPairWeightScale);
                if (q) atomicAdd(&plane[(bv + jv) * a.W + bu + ju], q);
            }
        const long long qu = llrint((double)u * kPairMomentScale), qv = llrint((double)v * kPairMomentScale);
        cnt += 1;
        su += qu;
        sv += qv;
        suu += llrint((double)u * (double)u * kPairMomentScale);
        svv += llrint((double)v * (double)v * kPairMomentScale);
    }
    // moments: wave sums, then one atomic per wave (integer sums: order-free)
    long long vals[5] = {cnt, su, sv, suu, svv};
#pragma unroll
    for (int j = 0; j < 5; ++j) {
        long long t = vals[j];
        for (int o = 32; o > 0; o >>= 1) t += __shfl_down(t, o);
        if ((threadIdx.x & 63) == 0 && t) atomicAdd(&a.mom[((int64_t)k * 2 + st) * 5 + j], (unsigned long long)t);
    }
    __syncthreads();
    unsigned long long *g = a.planes + ((int64_t)k * 2 + st) * HW;
    for (int c = threadIdx.x; c < HW; c += blockDim.x)
        if (plane[c]) atomicAdd(&g[c], plane[c]);
}

// One block per sample: the neutralised Gaussian-subtracted residuals of both deposits (the
// Gaussian of each deposit's own particle moments, gaussian_pic.jl:543-576), the network input
// (residual / peak, sqrt(expected count) / peak) and the target (the converged residual rescaled to
// the noisy sample's particle count, / the same peak); monitor = rms of the input residual channel.
__global__ __launch_bounds__(256) void k_pair_finish(PairArgs a, float *input, float *target, float *monitor) {
    extern __shared__ double sh[];
    const int k = blockIdx.x, H = a.H, W = a.W, HW = H * W;
    double *gx = sh, *gy = sh + W, *gxt = sh + W + H, *gyt = sh + 2 * W + H;   // 2 (W + H) doubles
    __shared__ double red[256];
    __shared__ double par[16];
    const unsigned long long *m = a.mom + (int64_t)k * 10;
    if (threadIdx.x < 2) {
        const int st = threadIdx.x;
        const double n = (double)m[st * 5], mu = (double)(long long)m[st * 5 + 1] / kPairMomentScale / n,
                     mv = (double)(long long)m[st * 5 + 2] / kPairMomentScale / n;
        const double vu = (double)(long long)m[st * 5 + 3] / kPairMomentScale / n - mu * mu,
                     vv = (double)(long long)m[st * 5 + 4] / kPairMomentScale / n - mv * mv;
        par[st * 5 + 0] = n;
        par[st * 5 + 1] = mu;
        par[st * 5 + 2] = mv;
        par[st * 5 + 3] = sqrt(vu > 0.0 ? vu : 0.0);
        par[st * 5 + 4] = sqrt(vv > 0.0 ? vv : 0.0);
    }
    __syncthreads();
    for (int i = threadIdx.x; i < W; i += blockDim.x) {
        gx[i] = gauss_node_weight((double)i, 1.0, par[1], par[3], a.method);
        gxt[i] = gauss_node_weight((double)i, 1.0, par[6], par[8], a.method);
    }
    for (int j = threadIdx.x; j < H; j += blockDim.x) {
        gy[j] = gauss_node_weight((double)j, 1.0, par[2], par[4], a.method);
        gyt[j] = gauss_node_weight((double)j, 1.0, par[7], par[9], a.method);
    }
    __syncthreads();
    // sums: charge (noisy, converged), profile sums, profile maxima, in fixed thread order
    const unsigned long long *qn = a.planes + (int64_t)k * 2 * HW, *qt = qn + HW;
    double acc[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    for (int c = threadIdx.x; c < HW; c += blockDim.x) {
        acc[0] += (double)qn[c];
        acc[1] += (double)qt[c];
    }
    for (int i = threadIdx.x; i < W; i += blockDim.x) { acc[2] += gx[i]; acc[4] += gxt[i]; acc[6] = fmax(acc[6], gx[i]); }
    for (int j = threadIdx.x; j < H; j += blockDim.x) { acc[3] += gy[j]; acc[5] += gyt[j]; acc[7] = fmax(acc[7], gy[j]); }
    __shared__ double tot[8];
    for (int q = 0; q < 8; ++q) {
        red[threadIdx.x] = acc[q];
        __syncthreads();
        if (threadIdx.x == 0) {
            double s = 0.0;
            for (int t = 0; t < (int)blockDim.x; ++t) s = q >= 6 ? fmax(s, red[t]) : s + red[t];
            tot[q] = s;
        }
        __syncthreads();
    }
    const double qsum = tot[0] / kPairWeightScale, qsum_t = tot[1] / kPairWeightScale;
    const double amp = tot[2] * tot[3] > 0.0 ? qsum / (tot[2] * tot[3]) : 0.0;
    const double amp_t = tot[4] * tot[5] > 0.0 ? qsum_t / (tot[4] * tot[5]) : 0.0;
    const double peak = amp * tot[6] * tot[7];
    const double inv_peak = peak > 0.0 ? 1.0 / peak : 0.0;
    const double rescale = par[5] > 0.0 ? par[0] / par[5] : 0.0;   // noisy count / converged count
    double ss = 0.0;
    for (int c = threadIdx.x; c < HW; c += blockDim.x) {
        const int j = c / W, i = c - j * W;
        const double g = amp * gx[i] * gy[j];
        const double r = (double)qn[c] / kPairWeightScale - g;
        const double rt = ((double)qt[c] / kPairWeightScale - amp_t * gxt[i] * gyt[j]) * rescale;
        const float x0 = (float)(r * inv_peak);
        input[((int64_t)k * HW + c) * 2 + 0] = x0;
        input[((int64_t)k * HW + c) * 2 + 1] = (float)(sqrt(g > 0.0 ? g : 0.0) * inv_peak);
        target[(int64_t)k * HW + c] = (float)(rt * inv_peak);
        ss += (double)x0 * (double)x0;
    }
    red[threadIdx.x] = ss;
    __syncthreads();
    if (threadIdx.x == 0) {
        double s = 0.0;
        for (int t = 0; t < (int)blockDim.x; ++t) s += red[t];
        if (monitor) monitor[k] = (float)sqrt(s / HW);
    }
}

// ------------------------------------------------------------------------- in the collision
// Source moments of each interaction (gaussian_pic.jl:374-435, uncoupled): sums of the source
// slice's x, px, y, py shifted by its first member, per chunk of members, in fixed order.
__global__ __launch_bounds__(256) void k_dn_moments(DnCollideArgs a) {
    __shared__ double red[10][256];
    const int k = blockIdx.y, c = blockIdx.x;
    const Inter &it = a.inters[k];
    double acc[10] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
    if (it.valid && it.src_n > 0 && !failed(a.err)) {
        const double *X = a.st.c[it.src_beam][0] + it.src_off, *PX = a.st.c[it.src_beam][1] + it.src_off,
                     *Y = a.st.c[it.src_beam][2] + it.src_off, *PY = a.st.c[it.src_beam][3] + it.src_off;
        const double x0 = X[0], px0 = PX[0], y0 = Y[0], py0 = PY[0];
        const int64_t b = it.src_n * c / a.chunks, e = it.src_n * (c + 1) / a.chunks;
        for (int64_t i = b + threadIdx.x; i < e; i += blockDim.x) {
            const double dx = X[i] - x0, dpx = PX[i] - px0, dy = Y[i] - y0, dpy = PY[i] - py0;
            acc[0] += dx; acc[1] += dpx; acc[2] += dy; acc[3] += dpy;
            acc[4] += dx * dx; acc[5] += dpx * dpx; acc[6] += dy * dy; acc[7] += dpy * dpy;
            acc[8] += dx * dpx; acc[9] += dy * dpy;
        }
    }
#pragma unroll
    for (int j = 0; j < 10; ++j) red[j][threadIdx.x] = acc[j];
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w)
#pragma unroll
            for (int j = 0; j < 10; ++j) red[j][threadIdx.x] += red[j][threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x < 10) a.mom_part[((int64_t)k * a.chunks + c) * 10 + threadIdx.x] = red[threadIdx.x][0];
}

// One block per field plane: the Gaussian of the source drifted by the plane's s
// (gaussian_pic.jl:482-491) on the plane's mesh, its neutralised subtraction (gaussian_pic.jl:543-576)
// and the network input (residual / peak, sqrt(expected count) / peak); the monitor is the rms of the
// input residual. A plane is passed to the network only when the source has at least two members and
// finite positive drifted widths (the reference's plain-PIC fallback, gaussian_pic.jl:525-530) and
// its monitor is within the limit.
__global__ __launch_bounds__(256) void k_dn_prepare(DnCollideArgs a) {
    extern __shared__ double g[];   // gx [nx], gy [ny]
    __shared__ double red[256];
    __shared__ double par[8];
    __shared__ double tot[5];
    const int p = blockIdx.x, nx = a.nx, ny = a.ny, HW = nx * ny;
    const Plane &pl = a.planes[p];
    const Inter &it = a.inters[pl.inter];
    __shared__ double smf[10];
    if (threadIdx.x < 10) {
        double t = 0.0;
        for (int c = 0; c < a.chunks; ++c) t += a.mom_part[((int64_t)pl.inter * a.chunks + c) * 10 + threadIdx.x];
        smf[threadIdx.x] = t;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double sm[10];
        for (int j = 0; j < 10; ++j) sm[j] = smf[j];
        const int64_t n = it.src_n;
        int ok = it.valid && n >= 2 && !failed(a.err);
        double mux = 0, muy = 0, sigx = 0, sigy = 0;
        if (ok) {
            const int b = it.src_beam;
            const double x0 = a.st.c[b][0][it.src_off], px0 = a.st.c[b][1][it.src_off], y0 = a.st.c[b][2][it.src_off],
                         py0 = a.st.c[b][3][it.src_off];
            const double invn = 1.0 / (double)n;
            const double dmx = sm[0] * invn, dmpx = sm[1] * invn, dmy = sm[2] * invn, dmpy = sm[3] * invn;
            const double mx = x0 + dmx, mpx = px0 + dmpx, my = y0 + dmy, mpy = py0 + dmpy;
            double varx = sm[4] * invn - dmx * dmx, varpx = sm[5] * invn - dmpx * dmpx;
            double vary = sm[6] * invn - dmy * dmy, varpy = sm[7] * invn - dmpy * dmpy;
            varx = varx > 0.0 ? varx : 0.0; varpx = varpx > 0.0 ? varpx : 0.0;
            vary = vary > 0.0 ? vary : 0.0; varpy = varpy > 0.0 ? varpy : 0.0;
            const double cxpx = sm[8] * invn - dmx * dmpx, cypy = sm[9] * invn - dmy * dmpy;
            const double s = pl.s;
            mux = mx + mpx * s;
            muy = my + mpy * s;
            const double vx = varx + 2.0 * s * cxpx + s * s * varpx, vy = vary + 2.0 * s * cypy + s * s * varpy;
            sigx = sqrt(vx > 0.0 ? vx : 0.0);
            sigy = sqrt(vy > 0.0 ? vy : 0.0);
            ok = sigx > 0.0 && sigy > 0.0 && isfinite(sigx) && isfinite(sigy) && isfinite(mux) && isfinite(muy);
        }
        par[0] = ok;
        par[1] = (mux - pl.x0) / pl.hx;
        par[2] = (muy - pl.y0) / pl.hy;
        par[3] = sigx / pl.hx;
        par[4] = sigy / pl.hy;
    }
    __syncthreads();
    const bool ok = par[0] != 0.0;
    {
        // erf and pdf at the cell boundaries of both axes (nx + 3 and ny + 3 points), then the nodes
        double *Ex = g + nx + ny, *Gx = Ex + (nx + 3), *Ey = Gx + (nx + 3), *Gy = Ey + (ny + 3);
        const double off = a.method == 0 ? 1.0 : 1.5;
        if (ok) {
            for (int k = threadIdx.x; k < nx + 3 + ny + 3; k += blockDim.x) {
                const bool xa = k < nx + 3;
                const int kk = xa ? k : k - (nx + 3);
                const double mu = xa ? par[1] : par[2], sg = xa ? par[3] : par[4];
                const double invroot = 1.0 / (sg * 1.4142135623730951), norm = 1.0 / (sg * 2.5066282746310002);
                const double p = (double)kk - off;
                (xa ? Ex : Ey)[kk] = erf((p - mu) * invroot);
                (xa ? Gx : Gy)[kk] = norm * exp(-((p - mu) * (p - mu)) / (2.0 * sg * sg));
            }
        }
        __syncthreads();
        for (int i = threadIdx.x; i < nx; i += blockDim.x) g[i] = ok ? gauss_node_weight_tab(i, par[1], par[3], a.method, Ex, Gx) : 0.0;
        for (int j = threadIdx.x; j < ny; j += blockDim.x) g[nx + j] = ok ? gauss_node_weight_tab(j, par[2], par[4], a.method, Ey, Gy) : 0.0;
    }
    __syncthreads();
    const unsigned long long *q = a.charge + (int64_t)p * a.charge_stride;
    double acc[5] = {0, 0, 0, 0, 0};   // charge, sum gx, sum gy, max gx, max gy
    for (int c = threadIdx.x; c < HW; c += blockDim.x) acc[0] += (double)(long long)q[c] * a.inv_scale;
    for (int i = threadIdx.x; i < nx; i += blockDim.x) { acc[1] += g[i]; acc[3] = fmax(acc[3], g[i]); }
    for (int j = threadIdx.x; j < ny; j += blockDim.x) { acc[2] += g[nx + j]; acc[4] = fmax(acc[4], g[nx + j]); }
    for (int v = 0; v < 5; ++v) {   // fixed-order tree reductions
        red[threadIdx.x] = acc[v];
        __syncthreads();
        for (int w = 128; w > 0; w >>= 1) {
            if ((int)threadIdx.x < w) red[threadIdx.x] = v >= 3 ? fmax(red[threadIdx.x], red[threadIdx.x + w]) : red[threadIdx.x] + red[threadIdx.x + w];
            __syncthreads();
        }
        if (threadIdx.x == 0) tot[v] = red[0];
        __syncthreads();
    }
    const double amp = tot[1] * tot[2] > 0.0 ? tot[0] / (tot[1] * tot[2]) : 0.0;
    const double peak = amp * tot[3] * tot[4];
    const double inv_peak = peak > 0.0 ? 1.0 / peak : 0.0;
    double ss = 0.0;
    float *in = a.input + (int64_t)p * HW * 2;
    for (int c = threadIdx.x; c < HW; c += blockDim.x) {
        const int j = c / nx, i = c - j * nx;
        const double gg = amp * g[i] * g[nx + j];
        const float r = (float)(((double)(long long)q[c] * a.inv_scale - gg) * inv_peak);
        in[2 * c] = r;
        in[2 * c + 1] = (float)(sqrt(gg > 0.0 ? gg : 0.0) * inv_peak);
        ss += (double)r * (double)r;
    }
    red[threadIdx.x] = ss;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        const double mon = sqrt(red[0] / HW);
        double *m = a.meta + (int64_t)p * 4;
        m[0] = amp;
        m[1] = peak;
        m[2] = mon;
        m[3] = (ok && peak > 0.0 && isfinite(mon) && mon <= a.limit) ? 1.0 : 0.0;
    }
    double *pr = a.prof + (int64_t)p * (nx + ny);
    for (int c = threadIdx.x; c < nx + ny; c += blockDim.x) pr[c] = g[c];
}

// One block per field plane: where the plane is used, charge = amp gx gy + peak (out - mean(out)),
// so the plane keeps its deposited total; elsewhere the plain deposit stays.
__global__ __launch_bounds__(256) void k_dn_apply(DnCollideArgs a, const float *out) {
    __shared__ double red[256];
    const int p = blockIdx.x, nx = a.nx, ny = a.ny, HW = nx * ny;
    const double *m = a.meta + (int64_t)p * 4;
    const bool use = m[3] != 0.0;
    if (threadIdx.x == 0) atomicAdd(&a.counters[use ? 0 : 1], 1ull);
    if (!use) return;
    const float *o = out + (int64_t)p * HW;
    double t = 0.0;
    for (int c = threadIdx.x; c < HW; c += blockDim.x) t += (double)o[c];
    red[threadIdx.x] = t;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    const double mean = red[0] / HW, amp = m[0], peak = m[1];
    const double *gx = a.prof + (int64_t)p * (nx + ny), *gy = gx + nx;
    unsigned long long *q = a.charge + (int64_t)p * a.charge_stride;
    for (int c = threadIdx.x; c < HW; c += blockDim.x) {
        const int j = c / nx, i = c - j * nx;
        const double v = amp * gx[i] * gy[j] + peak * ((double)o[c] - mean);
        q[c] = (unsigned long long)llrint(v * a.scale);
    }
}

// ------------------------------------------------------------------ Gaussian-subtracted solver
// Source moments of interaction d = 2k + dir - 1 of a batch (gaussian_pic.jl:374-435, uncoupled), before
// the geometry stage: the source slice of pair k is beam 1's slice i (dir 1) or beam 2's slice j (dir 2).
// Sums of x, px, y, py shifted by the slice's first member, per chunk of members, in fixed order.
__global__ __launch_bounds__(256) void k_gpic_moments(GpicMomentArgs a) {
    __shared__ double red[10][256];
    const int d = blockIdx.y, c = blockIdx.x, k = d >> 1, side = d & 1;
    const BatchPair pr = a.pairs[k];
    const SliceTable *tab = side == 0 ? a.tab1 : a.tab2;
    const int sl = side == 0 ? pr.i : pr.j;
    const int64_t off = tab->offset[sl], n = tab->count[sl];
    double acc[10] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
    if (n > 0 && !failed(a.err)) {
        const double *X = a.st.c[side][0] + off, *PX = a.st.c[side][1] + off, *Y = a.st.c[side][2] + off,
                     *PY = a.st.c[side][3] + off;
        const double x0 = X[0], px0 = PX[0], y0 = Y[0], py0 = PY[0];
        if (a.shift && c == 0 && threadIdx.x == 0) {
            double *sh = a.shift + (int64_t)d * 4;
            sh[0] = x0; sh[1] = px0; sh[2] = y0; sh[3] = py0;
        }
        const int64_t b = n * c / a.chunks, e = n * (c + 1) / a.chunks;
        for (int64_t i = b + threadIdx.x; i < e; i += blockDim.x) {
            const double dx = X[i] - x0, dpx = PX[i] - px0, dy = Y[i] - y0, dpy = PY[i] - py0;
            acc[0] += dx; acc[1] += dpx; acc[2] += dy; acc[3] += dpy;
            acc[4] += dx * dx; acc[5] += dpx * dpx; acc[6] += dy * dy; acc[7] += dpy * dpy;
            acc[8] += dx * dpx; acc[9] += dy * dpy;
        }
    }
#pragma unroll
    for (int j = 0; j < 10; ++j) red[j][threadIdx.x] = acc[j];
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w)
#pragma unroll
            for (int j = 0; j < 10; ++j) red[j][threadIdx.x] += red[j][threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x < 10) a.mom_part[((int64_t)d * a.chunks + c) * 10 + threadIdx.x] = red[threadIdx.x][0];
}

// One block per field plane of an interaction that subtracts (gaussian_pic.jl:543-595): the erf profile of
// the source's Gaussian drifted to the plane's boundary, on the plane's mesh (node integrals against the
// CIC tent / TSC spline, in cell units), scaled to the deposited total (neutralize) or to the slice
// population, subtracted from the fixed-point deposit.  The deposit's total is an exact integer sum.
__global__ __launch_bounds__(256) void k_gpic_residual(GpicResidualArgs a) {
    extern __shared__ double g[];   // gx [nx], gy [ny], then the erf / pdf tables
    __shared__ double red[256];
    __shared__ long long ired[256];
    __shared__ double tot[2];
    __shared__ long long qtot;
    const int p = blockIdx.x, nx = a.nx, ny = a.ny, HW = nx * ny;
    const Plane &pl = a.planes[p];
    const Inter &it = a.inters[pl.inter];
    if (!it.valid || !it.gauss || failed(a.err)) return;
    const GaussBnd &gb = it.g[p - it.plane0];
    const double mu_x = (gb.mux - pl.x0) / pl.hx, mu_y = (gb.muy - pl.y0) / pl.hy;
    const double sg_x = gb.sigx / pl.hx, sg_y = gb.sigy / pl.hy;
    {
        double *Ex = g + nx + ny, *Gx = Ex + (nx + 3), *Ey = Gx + (nx + 3), *Gy = Ey + (ny + 3);
        const double off = a.method == 0 ? 1.0 : 1.5;
        for (int k = threadIdx.x; k < nx + 3 + ny + 3; k += blockDim.x) {
            const bool xa = k < nx + 3;
            const int kk = xa ? k : k - (nx + 3);
            const double mu = xa ? mu_x : mu_y, sg = xa ? sg_x : sg_y;
            const double invroot = 1.0 / (sg * 1.4142135623730951), norm = 1.0 / (sg * 2.5066282746310002);
            const double q = (double)kk - off;
            (xa ? Ex : Ey)[kk] = erf((q - mu) * invroot);
            (xa ? Gx : Gy)[kk] = norm * exp(-((q - mu) * (q - mu)) / (2.0 * sg * sg));
        }
        __syncthreads();
        for (int i = threadIdx.x; i < nx; i += blockDim.x) g[i] = gauss_node_weight_tab(i, mu_x, sg_x, a.method, Ex, Gx);
        for (int j = threadIdx.x; j < ny; j += blockDim.x) g[nx + j] = gauss_node_weight_tab(j, mu_y, sg_y, a.method, Ey, Gy);
    }
    __syncthreads();
    unsigned long long *q = a.charge + (int64_t)p * a.charge_stride;
    long long qs = 0;
    for (int c = threadIdx.x; c < HW; c += blockDim.x) qs += (long long)q[c];
    double sx = 0.0, sy = 0.0;
    for (int i = threadIdx.x; i < nx; i += blockDim.x) sx += g[i];
    for (int j = threadIdx.x; j < ny; j += blockDim.x) sy += g[nx + j];
    ired[threadIdx.x] = qs;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) ired[threadIdx.x] += ired[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) qtot = ired[0];
    for (int v = 0; v < 2; ++v) {   // fixed-order tree reductions of the profile sums
        red[threadIdx.x] = v == 0 ? sx : sy;
        __syncthreads();
        for (int w = 128; w > 0; w >>= 1) {
            if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
            __syncthreads();
        }
        if (threadIdx.x == 0) tot[v] = red[0];
        __syncthreads();
    }
    const double sgx = tot[0], sgy = tot[1];
    double amp = (double)it.src_n;
    if (a.neutralize && sgx * sgy > 0.0) amp = ((double)qtot * a.inv_scale) / (sgx * sgy);
    for (int c = threadIdx.x; c < HW; c += blockDim.x) {
        const int j = c / nx, i = c - j * nx;
        const double v = g[i] * (amp * g[nx + j]);
        q[c] = (unsigned long long)((long long)q[c] - llrint(v * a.scale));
    }
}

hipError_t launch_gpic_moments(const GpicMomentArgs &a, hipStream_t s) {
    if (a.npairs <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_gpic_moments, dim3(a.chunks, 2 * a.npairs), dim3(256), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_gpic_residual(const GpicResidualArgs &a, hipStream_t s) {
    if (a.nplanes <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_gpic_residual, dim3(a.nplanes), dim3(256), (size_t)(3 * (a.nx + a.ny) + 12) * sizeof(double), s, a);
    return hipGetLastError();
}

hipError_t launch_dn_prepare(const DnCollideArgs &a, hipStream_t s) {
    if (a.nplanes <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_dn_moments, dim3(a.chunks, a.ninters), dim3(256), 0, s, a);
    // LDS: the profiles (nx + ny) and the erf / pdf tables (2 (nx + 3) + 2 (ny + 3)) in doubles
    hipLaunchKernelGGL(k_dn_prepare, dim3(a.nplanes), dim3(256), (size_t)(3 * (a.nx + a.ny) + 12) * sizeof(double), s, a);
    return hipGetLastError();
}

hipError_t launch_dn_apply(const DnCollideArgs &a, const float *out, hipStream_t s) {
    if (a.nplanes <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_dn_apply, dim3(a.nplanes), dim3(256), 0, s, a, out);
    return hipGetLastError();
}

// --------------------------------------------------------------------------------- the network
namespace {

struct LayerSpec { const char *name; int cin, cout; };
constexpr LayerSpec kLayers[10] = {{"e1a", 2, 16}, {"e1b", 16, 16}, {"e2a", 16, 32}, {"e2b", 32, 32}, {"ba", 32, 64},
                                   {"bb", 64, 64},  {"d2a", 96, 32}, {"d2b", 32, 32}, {"d1a", 48, 16}, {"d1b", 16, 16}};
enum { E1A, E1B, E2A, E2B, BA, BB, D2A, D2B, D1A, D1B };

inline int ceil16(int v) { return (v + 15) / 16 * 16; }
inline int ceil32(int v) { return (v + 31) / 32 * 32; }
inline int grid_for(int64_t n, int threads = 256) {
    int64_t g = (n + threads - 1) / threads;
    if (g < 1) g = 1;
    if (g > 16384) g = 16384;
    return (int)g;
}

}  // namespace

struct DenoiserImpl {
    int device = 0, H = 0, W = 0, maxb = 0;
    hipStream_t stream = nullptr;
    std::string err;
    int64_t nparams = 0;
    int64_t woff[10], boff[10], hoff = 0;   // offsets into the flat parameter vector
    int cin_s[10], kpad[10];                 // stored input channels, padded reduction length
    float *params = nullptr, *grad = nullptr, *m = nullptr, *v = nullptr;
    uint16_t *wf[10] = {}, *wb[10] = {};
    int64_t step = 0;
    // activations (bf16) at full (P1), half (P2), quarter (P3) resolution pixel counts
    uint16_t *x0 = nullptr, *a[10] = {}, *q1 = nullptr, *q2 = nullptr;
    float *out = nullptr;
    // gradients (fp32) and their bf16 copies
    float *gA = nullptr, *gB = nullptr, *gS1 = nullptr, *gS2 = nullptr, *gpart = nullptr, *hpart = nullptr, *bpart = nullptr;
    uint16_t *dq = nullptr;
    int64_t gpart_floats = 0;
    unsigned long long *pair_planes = nullptr, *pair_mom = nullptr;   // pair generator: [maxb][2][H*W], [maxb][10]
    std::vector<void *> allocs;

    template <class T>
    bool alloc(T *&p, int64_t n) {
        void *q = nullptr;
        if (hipMalloc(&q, (size_t)(n > 0 ? n : 1) * sizeof(T)) != hipSuccess) { err = "device allocation failed"; return false; }
        allocs.push_back(q);
        p = reinterpret_cast<T *>(q);
        return true;
    }
    ~DenoiserImpl() {
        for (void *q : allocs) hipFree(q);
    }
    int64_t P(int level, int B) const { return (int64_t)B * (H >> level) * (W >> level); }
    int level_of(int l) const { return (l == E1A || l == E1B || l == D1A || l == D1B) ? 0 : (l == BA || l == BB) ? 2 : 1; }

    hipError_t pack_all() {
        for (int l = 0; l < 10; ++l) {
            hipLaunchKernelGGL(k_pack_weights, dim3(64), dim3(256), 0, stream, params + woff[l], kLayers[l].cout, cin_s[l],
                               kLayers[l].cin, kpad[l], wf[l], l == E1A ? nullptr : wb[l]);
        }
        return hipGetLastError();
    }

    // Returns whether the pooled activation (yq) was written by the product kernel.
    bool conv(int l, const uint16_t *in, int B, int level, bool relu, uint16_t *y, float *yf, bool backward,
              const uint16_t *in2 = nullptr, int cd = 0, uint16_t *yq = nullptr) {
        ConvArgs c{};
        c.x = in;
        c.x2 = in2;
        c.cd = cd;
        c.yq = yq;
        c.B = B;
        c.H = H >> level;
        c.W = W >> level;
        if (!backward) {
            c.w = wf[l]; c.bias = params + boff[l]; c.cin = cin_s[l]; c.cout = kLayers[l].cout; c.kpad = kpad[l];
        } else {
            c.w = wb[l]; c.bias = nullptr; c.cin = kLayers[l].cout; c.cout = kLayers[l].cin; c.kpad = ceil32(9 * kLayers[l].cout);
        }
        c.relu = relu ? 1 : 0;
        c.y = y;
        c.yf = yf;
        const int64_t Pn = (int64_t)B * c.H * c.W;
        const int64_t tiles16 = (int64_t)B * ((c.H + kTile - 1) / kTile) * ((c.W + kTile - 1) / kTile);
        if (c.cin % 8 == 0 && tiles16 < kConvTile16From && conv_tile8_lds(c.cin, c.cout, c.kpad) <= 160 * 1024 - 1024) {
            const size_t lds = conv_tile8_lds(c.cin, c.cout, c.kpad);
            const dim3 gt((unsigned)(B * ((c.H + kTile8 - 1) / kTile8) * ((c.W + kTile8 - 1) / kTile8)));
#define OCTOPIC_CONVT8(NT_)                                                                                       \
    {                                                                                                            \
        hipFuncSetAttribute((const void *)k_conv_tile8<NT_>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds); \
        hipLaunchKernelGGL((k_conv_tile8<NT_>), gt, dim3(kConvThreads), lds, stream, c);                         \
    }
            switch (c.cout / 16) {
                case 1: OCTOPIC_CONVT8(1) break;
                case 2: OCTOPIC_CONVT8(2) break;
                case 3: OCTOPIC_CONVT8(3) break;
                case 4: OCTOPIC_CONVT8(4) break;
                case 6: OCTOPIC_CONVT8(6) break;
                default: break;
            }
#undef OCTOPIC_CONVT8
            return c.yq != nullptr;
        }
        if (c.cin % 8 == 0 && conv_tile_lds(c.cin, c.cout, c.kpad) <= 160 * 1024 - 1024) {
            const size_t lds = conv_tile_lds(c.cin, c.cout, c.kpad);
            const dim3 gt((unsigned)(B * ((c.H + kTile - 1) / kTile) * ((c.W + kTile - 1) / kTile)));
#define OCTOPIC_CONVT(NT_)                                                                                       \
    {                                                                                                           \
        hipFuncSetAttribute((const void *)k_conv_tile<NT_>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds); \
        hipLaunchKernelGGL((k_conv_tile<NT_>), gt, dim3(kConvThreads), lds, stream, c);                         \
    }
            switch (c.cout / 16) {
                case 1: OCTOPIC_CONVT(1) break;
                case 2: OCTOPIC_CONVT(2) break;
                case 3: OCTOPIC_CONVT(3) break;
                case 4: OCTOPIC_CONVT(4) break;
                case 6: OCTOPIC_CONVT(6) break;
                default: break;
            }
#undef OCTOPIC_CONVT
            return c.yq != nullptr;
        }
        if (c.cin % 8 == 0) {
            const size_t lds = (size_t)c.cout * (c.kpad + 8) * sizeof(uint16_t);
            const int nt = c.cout / 16, mt = nt <= 2 ? 4 : 2;
            const dim3 g8((unsigned)((Pn + 64 * mt - 1) / (64 * mt)));
#define OCTOPIC_CONV8(NT_, MT_)                                                                                  \
    {                                                                                                           \
        hipFuncSetAttribute((const void *)k_conv8<NT_, MT_>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds); \
        hipLaunchKernelGGL((k_conv8<NT_, MT_>), g8, dim3(kConvThreads), lds, stream, c);                        \
    }
            switch (nt) {
                case 1: OCTOPIC_CONV8(1, 4) break;
                case 2: OCTOPIC_CONV8(2, 4) break;
                case 3: OCTOPIC_CONV8(3, 2) break;
                case 4: OCTOPIC_CONV8(4, 2) break;
                case 6: OCTOPIC_CONV8(6, 2) break;
                default: break;
            }
#undef OCTOPIC_CONV8
            return false;
        }
        const dim3 g((unsigned)((Pn + 63) / 64));
        switch (c.cout / 16) {
            case 1: hipLaunchKernelGGL(k_conv<1>, g, dim3(kConvThreads), 0, stream, c); break;
            case 2: hipLaunchKernelGGL(k_conv<2>, g, dim3(kConvThreads), 0, stream, c); break;
            case 3: hipLaunchKernelGGL(k_conv<3>, g, dim3(kConvThreads), 0, stream, c); break;
            case 4: hipLaunchKernelGGL(k_conv<4>, g, dim3(kConvThreads), 0, stream, c); break;
            case 6: hipLaunchKernelGGL(k_conv<6>, g, dim3(kConvThreads), 0, stream, c); break;
            default: break;
        }
        return false;
    }

    void forward(const float *input, int B) {
        const int64_t P0 = P(0, B);
        hipLaunchKernelGGL(k_pack_input, dim3(grid_for(P0)), dim3(256), 0, stream, input, P0, x0);
        conv(E1A, x0, B, 0, true, a[E1A], nullptr, false);
        // the pools are written by the products that make their inputs (k_pool only when a product cannot)
        if (!conv(E1B, a[E1A], B, 0, true, a[E1B], nullptr, false, nullptr, 0, q1))
            hipLaunchKernelGGL(k_pool, dim3(grid_for(P(1, B) * 16)), dim3(256), 0, stream, a[E1B], B, H, W, 16, q1);
        conv(E2A, q1, B, 1, true, a[E2A], nullptr, false);
        if (!conv(E2B, a[E2A], B, 1, true, a[E2B], nullptr, false, nullptr, 0, q2))
            hipLaunchKernelGGL(k_pool, dim3(grid_for(P(2, B) * 32)), dim3(256), 0, stream, a[E2B], B, H >> 1, W >> 1, 32, q2);
        conv(BA, q2, B, 2, true, a[BA], nullptr, false);
        conv(BB, a[BA], B, 2, true, a[BB], nullptr, false);
        // decoder inputs: upsampled deep features and the skip, read from their own tensors
        conv(D2A, a[BB], B, 1, true, a[D2A], nullptr, false, a[E2B], 64);
        conv(D2B, a[D2A], B, 1, true, a[D2B], nullptr, false);
        conv(D1A, a[D2B], B, 0, true, a[D1A], nullptr, false, a[E1B], 32);
        conv(D1B, a[D1A], B, 0, true, a[D1B], nullptr, false);
        hipLaunchKernelGGL(k_head, dim3(grid_for(P0)), dim3(256), 0, stream, a[D1B], params + hoff, params + hoff + 16, P0, out);
    }

    // Gradient of layer l given dL/d(activation) in da (fp32 [P][cout]); input-gradient written to
    // dx (fp32 [P][cin]) unless l is the first layer.
    void layer_back(int l, const uint16_t *in, const float *da, int B, float *dx, const uint16_t *in2 = nullptr, int cd = 0) {
        const int level = level_of(l), C = kLayers[l].cout;
        const int64_t Pn = P(level, B);
        const int nb = (int)((Pn + 63) / 64);
        hipLaunchKernelGGL(k_relu_back, dim3(nb), dim3(256), 0, stream, da, a[l], Pn, C, dq, bpart);
        hipLaunchKernelGGL(k_sum_parts, dim3(1), dim3(64), 0, stream, bpart, nb, C, grad + boff[l]);
        WgradArgs w{};
        w.x = in; w.x2 = in2; w.cd = cd; w.dy = dq; w.B = B; w.H = H >> level; w.W = W >> level; w.cin = cin_s[l]; w.cout = C; w.kpad = kpad[l];
        w.split_pixels = 1024;
        const int splits = (int)((Pn + w.split_pixels - 1) / w.split_pixels);
        w.part = gpart;
        hipLaunchKernelGGL(k_wgrad, dim3(kpad[l] / 16, C / 16, splits), dim3(kConvThreads), 0, stream, w);
        hipLaunchKernelGGL(k_wgrad_reduce, dim3(grid_for((int64_t)C * 9 * kLayers[l].cin)), dim3(256), 0, stream, gpart, splits, C,
                           kpad[l], cin_s[l], kLayers[l].cin, grad + woff[l]);
        if (dx) conv(l, dq, B, level, false, nullptr, dx, true);
    }

    void backward(const float *target, int B) {
        const int64_t P0 = P(0, B), P1 = P(1, B), P2 = P(2, B);
        const int nb0 = (int)((P0 + 63) / 64);
        // head
        hipLaunchKernelGGL(k_head_back, dim3(nb0), dim3(64), 0, stream, out, target, a[D1B], params + hoff, P0, gA, hpart);
        hipLaunchKernelGGL(k_sum_parts, dim3(1), dim3(64), 0, stream, hpart, nb0, 18, hpart + (int64_t)nb0 * 18);
        // d1b, d1a (its input gradient is [P0][48])
        layer_back(D1B, a[D1A], gA, B, gB);
        layer_back(D1A, a[D2B], gB, B, gA, a[E1B], 32);
        hipLaunchKernelGGL(k_up_concat_back, dim3(grid_for(P0 * 48)), dim3(256), 0, stream, gA, B, H, W, 32, 16, gB, gS1);
        // d2b (da = gB [P1][32]), d2a (input gradient [P1][96])
        layer_back(D2B, a[D2A], gB, B, gA);
        layer_back(D2A, a[BB], gA, B, gB, a[E2B], 64);
        hipLaunchKernelGGL(k_up_concat_back, dim3(grid_for(P1 * 96)), dim3(256), 0, stream, gB, B, H >> 1, W >> 1, 64, 32, gA, gS2);
        // bottleneck (da = gA [P2][64])
        layer_back(BB, a[BA], gA, B, gB);
        layer_back(BA, q2, gB, B, gA);
        // e2b: da = skip gradient + pooled gradient
        hipLaunchKernelGGL(k_pool_back_add, dim3(grid_for(P1 * 32)), dim3(256), 0, stream, gA, gS2, B, H >> 1, W >> 1, 32, gB);
        layer_back(E2B, a[E2A], gB, B, gA);
        layer_back(E2A, q1, gA, B, gB);
        hipLaunchKernelGGL(k_pool_back_add, dim3(grid_for(P0 * 16)), dim3(256), 0, stream, gB, gS1, B, H, W, 16, gA);
        layer_back(E1B, a[E1A], gA, B, gB);
        layer_back(E1A, x0, gB, B, nullptr);
        (void)P2;
    }
};

namespace {
// Head partial totals live after the per-block partials: [nb0][18] then [18].
void head_grads(DenoiserImpl &d, int B, float *loss_dev) {
    const int nb0 = (int)((d.P(0, B) + 63) / 64);
    const float *tot = d.hpart + (int64_t)nb0 * 18;
    hipMemcpyAsync(d.grad + d.hoff, tot, 16 * sizeof(float), hipMemcpyDeviceToDevice, d.stream);
    hipMemcpyAsync(d.grad + d.hoff + 16, tot + 16, sizeof(float), hipMemcpyDeviceToDevice, d.stream);
    if (loss_dev) hipLaunchKernelGGL(k_loss_finish, dim3(1), dim3(64), 0, d.stream, tot, d.P(0, B), loss_dev);
}
}  // namespace

}  // namespace octopic

using octopic::DenoiserImpl;

struct octopic_denoiser {
    DenoiserImpl impl;
};

namespace octopic {
const float *denoiser_run(octopic_denoiser *d, const float *input, int batch, hipStream_t s, hipError_t *e) {
    DenoiserImpl &m = d->impl;
    hipStream_t keep = m.stream;
    m.stream = s;
    m.forward(input, batch);
    m.stream = keep;
    *e = hipGetLastError();
    return m.out;
}
void denoiser_shape(const octopic_denoiser *d, int *height, int *width, int *max_batch) {
    *height = d->impl.H;
    *width = d->impl.W;
    *max_batch = d->impl.maxb;
}
}  // namespace octopic

namespace {
int fail(octopic_denoiser *d, int code, const std::string &msg) {
    if (d) d->impl.err = msg;
    return code;
}
int hip_check(octopic_denoiser *d, hipError_t e, const char *what) {
    if (e == hipSuccess) return OCTOPIC_OK;
    return fail(d, OCTOPIC_ERR_DEVICE, std::string(what) + ": " + hipGetErrorString(e));
}

// Deterministic He-normal initial weights (biases zero) from a 64-bit seed (splitmix64 + Box-Muller).
uint64_t mix64(uint64_t &s) {
    uint64_t z = (s += 0x9e3779b97f4a7c15ull);
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}
double normal01(uint64_t &s) {
    const double u1 = ((double)(mix64(s) >> 11) + 0.5) * (1.0 / 9007199254740992.0);
    const double u2 = ((double)(mix64(s) >> 11) + 0.5) * (1.0 / 9007199254740992.0);
    return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
}
}  // namespace

extern "C" {

int64_t octopic_denoiser_num_params(void) {
    int64_t n = 0;
    for (const auto &l : octopic::kLayers) n += (int64_t)l.cout * 9 * l.cin + l.cout;
    return n + 17;
}

int octopic_denoiser_create(int32_t device, int32_t height, int32_t width, int32_t max_batch, uint64_t seed,
                            octopic_denoiser **out) {
    using namespace octopic;
    if (!out) return OCTOPIC_ERR_INVALID_ARGUMENT;
    *out = nullptr;
    if (height < 16 || width < 16 || height % 16 || width % 16 || max_batch < 1 || (int64_t)height * width * max_batch > (1ll << 26))
        return OCTOPIC_ERR_INVALID_ARGUMENT;
    if (hipSetDevice(device) != hipSuccess) return OCTOPIC_ERR_DEVICE;
    auto *d = new octopic_denoiser();
    DenoiserImpl &m = d->impl;
    m.device = device;
    m.H = height;
    m.W = width;
    m.maxb = max_batch;
    int64_t off = 0;
    for (int l = 0; l < 10; ++l) {
        m.woff[l] = off;
        off += (int64_t)kLayers[l].cout * 9 * kLayers[l].cin;
        m.boff[l] = off;
        off += kLayers[l].cout;
        m.cin_s[l] = l == E1A ? 8 : kLayers[l].cin;
        m.kpad[l] = ceil32(9 * m.cin_s[l]);
    }
    m.hoff = off;
    m.nparams = off + 17;
    const int64_t P0 = (int64_t)max_batch * height * width, P1 = P0 / 4, P2 = P0 / 16;
    bool ok = m.alloc(m.params, m.nparams) && m.alloc(m.grad, m.nparams) && m.alloc(m.m, m.nparams) && m.alloc(m.v, m.nparams);
    for (int l = 0; ok && l < 10; ++l) {
        ok = m.alloc(m.wf[l], (int64_t)kLayers[l].cout * m.kpad[l]) &&
             (l == E1A || m.alloc(m.wb[l], (int64_t)kLayers[l].cin * ceil32(9 * kLayers[l].cout)));
        const int64_t Pl = m.level_of(l) == 0 ? P0 : m.level_of(l) == 1 ? P1 : P2;
        ok = ok && m.alloc(m.a[l], Pl * kLayers[l].cout);
    }
    // fp32 gradient buffers sized for the widest level: P0 x 48, P1 x 96 (both 12 P0)
    const int64_t gmax = P0 * 48 > P1 * 96 ? P0 * 48 : P1 * 96;
    // weight-gradient partials: splits (1024 pixels each) x cout x kpad, largest over layers
    int64_t gp = 0;
    for (int l = 0; l < 10; ++l) {
        const int64_t Pl = m.level_of(l) == 0 ? P0 : m.level_of(l) == 1 ? P1 : P2;
        const int64_t v = (Pl + 1023) / 1024 * kLayers[l].cout * m.kpad[l];
        if (v > gp) gp = v;
    }
    m.gpart_floats = gp;
    ok = ok && m.alloc(m.x0, P0 * 8) && m.alloc(m.q1, P1 * 16) && m.alloc(m.q2, P2 * 32) && m.alloc(m.out, P0) && m.alloc(m.gA, gmax) && m.alloc(m.gB, gmax) && m.alloc(m.gS1, P0 * 16) &&
         m.alloc(m.gS2, P1 * 32) && m.alloc(m.gpart, gp) && m.alloc(m.hpart, ((P0 + 63) / 64 + 1) * 18) &&
         m.alloc(m.bpart, (P0 + 63) / 64 * 64) && m.alloc(m.dq, gmax) && m.alloc(m.pair_planes, P0 * 2) &&
         m.alloc(m.pair_mom, (int64_t)max_batch * 10);
    if (!ok) {
        delete d;
        return OCTOPIC_ERR_OOM;
    }
    std::vector<float> h((size_t)m.nparams, 0.f);
    uint64_t s = seed;
    for (int l = 0; l < 10; ++l) {
        const double sd = std::sqrt(2.0 / (9.0 * kLayers[l].cin));
        for (int64_t i = 0; i < (int64_t)kLayers[l].cout * 9 * kLayers[l].cin; ++i) h[(size_t)(m.woff[l] + i)] = (float)(sd * normal01(s));
    }
    for (int c = 0; c < 16; ++c) h[(size_t)(m.hoff + c)] = (float)(0.25 * normal01(s));
    if (octopic_denoiser_set_params(d, h.data(), m.nparams) != OCTOPIC_OK) {
        delete d;
        return OCTOPIC_ERR_DEVICE;
    }
    *out = d;
    return OCTOPIC_OK;
}

void octopic_denoiser_destroy(octopic_denoiser *d) {
    if (!d) return;
    hipSetDevice(d->impl.device);
    if (d->impl.stream) hipStreamSynchronize(d->impl.stream);
    else hipDeviceSynchronize();
    delete d;
}

const char *octopic_denoiser_last_error(const octopic_denoiser *d) { return d ? d->impl.err.c_str() : "null denoiser"; }

int octopic_denoiser_set_stream(octopic_denoiser *d, void *stream) {
    if (!d) return OCTOPIC_ERR_INVALID_ARGUMENT;
    d->impl.stream = (hipStream_t)stream;
    return OCTOPIC_OK;
}

int octopic_denoiser_get_params(const octopic_denoiser *d, float *host, int64_t n) {
    if (!d || !host) return OCTOPIC_ERR_INVALID_ARGUMENT;
    auto *m = const_cast<octopic_denoiser *>(d);
    if (n != d->impl.nparams) return fail(m, OCTOPIC_ERR_INVALID_ARGUMENT, "parameter count mismatch");
    hipSetDevice(d->impl.device);
    if (int r = hip_check(m, hipStreamSynchronize(d->impl.stream), "synchronize")) return r;
    return hip_check(m, hipMemcpy(host, d->impl.params, (size_t)n * sizeof(float), hipMemcpyDeviceToHost), "copy parameters");
}

int octopic_denoiser_set_params(octopic_denoiser *d, const float *host, int64_t n) {
    if (!d || !host) return OCTOPIC_ERR_INVALID_ARGUMENT;
    DenoiserImpl &m = d->impl;
    if (n != m.nparams) return fail(d, OCTOPIC_ERR_INVALID_ARGUMENT, "parameter count mismatch");
    hipSetDevice(m.device);
    if (int r = hip_check(d, hipStreamSynchronize(m.stream), "synchronize")) return r;
    if (int r = hip_check(d, hipMemcpy(m.params, host, (size_t)n * sizeof(float), hipMemcpyHostToDevice), "copy parameters")) return r;
    if (int r = hip_check(d, hipMemsetAsync(m.m, 0, (size_t)n * sizeof(float), m.stream), "reset optimiser")) return r;
    if (int r = hip_check(d, hipMemsetAsync(m.v, 0, (size_t)n * sizeof(float), m.stream), "reset optimiser")) return r;
    m.step = 0;
    if (int r = hip_check(d, m.pack_all(), "pack weights")) return r;
    return hip_check(d, hipStreamSynchronize(m.stream), "synchronize");
}

int octopic_denoiser_forward(octopic_denoiser *d, const float *input, int32_t batch, float *output) {
    if (!d || !input || !output) return OCTOPIC_ERR_INVALID_ARGUMENT;
    DenoiserImpl &m = d->impl;
    if (batch < 1 || batch > m.maxb) return fail(d, OCTOPIC_ERR_INVALID_ARGUMENT, "batch outside [1, max_batch]");
    hipSetDevice(m.device);
    m.forward(input, batch);
    if (int r = hip_check(d, hipGetLastError(), "denoiser forward")) return r;
    return hip_check(d, hipMemcpyAsync(output, m.out, (size_t)m.P(0, batch) * sizeof(float), hipMemcpyDeviceToDevice, m.stream),
                     "copy output");
}

int octopic_denoiser_gradient(octopic_denoiser *d, const float *input, const float *target, int32_t batch, float *loss,
                              float *grad_out) {
    if (!d || !input || !target) return OCTOPIC_ERR_INVALID_ARGUMENT;
    DenoiserImpl &m = d->impl;
    if (batch < 1 || batch > m.maxb) return fail(d, OCTOPIC_ERR_INVALID_ARGUMENT, "batch outside [1, max_batch]");
    hipSetDevice(m.device);
    m.forward(input, batch);
    m.backward(target, batch);
    octopic::head_grads(m, batch, loss);
    if (int r = hip_check(d, hipGetLastError(), "denoiser backward")) return r;
    if (grad_out)
        return hip_check(d, hipMemcpyAsync(grad_out, m.grad, (size_t)m.nparams * sizeof(float), hipMemcpyDeviceToDevice, m.stream),
                         "copy gradient");
    return OCTOPIC_OK;
}

int octopic_denoiser_train_step(octopic_denoiser *d, const float *input, const float *target, int32_t batch, float lr,
                                float *loss, float *grad_out) {
    if (int r = octopic_denoiser_gradient(d, input, target, batch, loss, grad_out)) return r;
    DenoiserImpl &m = d->impl;
    if (!(lr > 0.f) || !std::isfinite(lr)) return fail(d, OCTOPIC_ERR_INVALID_ARGUMENT, "learning rate must be finite and positive");
    m.step += 1;
    const float b1 = 0.9f, b2 = 0.999f;
    const float bc1 = (float)(1.0 - std::pow(0.9, (double)m.step)), bc2 = (float)(1.0 - std::pow(0.999, (double)m.step));
    hipLaunchKernelGGL(octopic::k_adam, dim3(octopic::grid_for(m.nparams)), dim3(256), 0, m.stream, m.params, m.grad, m.m, m.v,
                       m.nparams, lr, b1, b2, (float)(1.0 - 0.9), (float)(1.0 - 0.999), bc1, bc2, 1e-8f);
    if (int r = hip_check(d, hipGetLastError(), "adam")) return r;
    return hip_check(d, m.pack_all(), "pack weights");
}

int octopic_denoiser_make_pairs(octopic_denoiser *d, uint64_t seed, int64_t first_sample, int32_t batch, int64_t n_target,
                                int32_t method, float *input, float *target, float *monitor, uint64_t *raw_planes,
                                uint64_t *raw_moments) {
    using namespace octopic;
    if (!d || !input || !target) return OCTOPIC_ERR_INVALID_ARGUMENT;
    DenoiserImpl &m = d->impl;
    if (batch < 1 || batch > m.maxb) return fail(d, OCTOPIC_ERR_INVALID_ARGUMENT, "batch outside [1, max_batch]");
    if (n_target < 1 || n_target > (1ll << 31)) return fail(d, OCTOPIC_ERR_INVALID_ARGUMENT, "n_target outside [1, 2^31]");
    if (method != 0 && method != 1) return fail(d, OCTOPIC_ERR_INVALID_ARGUMENT, "method must be 0 (CIC) or 1 (TSC)");
    if ((int64_t)m.H * m.W * 8 > 160 * 1024) return fail(d, OCTOPIC_ERR_UNSUPPORTED, "pair deposits need H*W <= 20480");
    hipSetDevice(m.device);
    const int64_t HW = (int64_t)m.H * m.W;
    if (int r = hip_check(d, hipMemsetAsync(m.pair_planes, 0, (size_t)batch * 2 * HW * 8, m.stream), "clear")) return r;
    if (int r = hip_check(d, hipMemsetAsync(m.pair_mom, 0, (size_t)batch * 10 * 8, m.stream), "clear")) return r;
    PairArgs a{};
    a.seed = seed;
    a.first = first_sample;
    a.H = m.H;
    a.W = m.W;
    a.method = method;
    a.n_target = n_target;
    a.planes = m.pair_planes;
    a.mom = m.pair_mom;
    int64_t ch = n_target / (256 * 256);
    a.chunks = (int)(ch < 4 ? 4 : ch > 256 ? 256 : ch);
    const size_t lds = (size_t)HW * 8;
    hipFuncSetAttribute((const void *)k_pair_deposit, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    hipLaunchKernelGGL(k_pair_deposit, dim3(a.chunks, batch, 2), dim3(256), lds, m.stream, a);
    hipLaunchKernelGGL(k_pair_finish, dim3(batch), dim3(256), (size_t)2 * (m.H + m.W) * sizeof(double), m.stream, a, input, target,
                       monitor);
    if (int r = hip_check(d, hipGetLastError(), "pair generator")) return r;
    if (raw_planes &&
        hipMemcpyAsync(raw_planes, m.pair_planes, (size_t)batch * 2 * HW * 8, hipMemcpyDeviceToDevice, m.stream) != hipSuccess)
        return fail(d, OCTOPIC_ERR_DEVICE, "copy planes");
    if (raw_moments && hipMemcpyAsync(raw_moments, m.pair_mom, (size_t)batch * 10 * 8, hipMemcpyDeviceToDevice, m.stream) != hipSuccess)
        return fail(d, OCTOPIC_ERR_DEVICE, "copy moments");
    return OCTOPIC_OK;
}

}  // extern "C"
