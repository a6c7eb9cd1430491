// nr_unet.hip — learned U-Net denoiser for gfx950 (include/nectar_denoise.h).
//
// Tensors are NHWC bf16 ("views": base, channel stride, channel offset).  Every
// 3x3 convolution is an implicit GEMM on the matrix cores:
//   forward  D[co][pix] = sum_k W[co][k] X[k][pix], k = (tap, ci)
//   input gradient: the same kernel with the weights flipped and transposed
//   weight gradient dW[co][tap][ci] = sum_pix G[co][pix] X[ci][pix+tap]
// using v_mfma_f32_16x16x32_bf16 (16 x 16 output tile, K = 32, fp32
// accumulation).  Forward/input-gradient convolutions stage a halo tile of 32
// input channels in LDS and run all 9 taps from it; the encoder's 2x2 average
// pool is fused into its second convolution's epilogue and the decoder's skip
// concatenation + nearest upsample into its first convolution's tile loader, so
// neither the pooled-away nor the upsampled tensor makes an extra HBM round
// trip.  The weight gradient transposes 32-pixel tiles through LDS so the
// reduction runs over pixels.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <random>
#include <string>
#include <type_traits>
#include <vector>

#include "../../include/nectar_denoise.h"
#include "nr_error.h"

namespace {

typedef __bf16 bf16x8 __attribute__((__vector_size__(8 * sizeof(__bf16))));
typedef float f32x4 __attribute__((__vector_size__(4 * sizeof(float))));
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));   // 16 bytes as one register quad

int fail(int code, const std::string& m) { nr_error_slot() = m; return code; }
#define UCK(expr)                                                                         \
    do {                                                                                  \
        hipError_t e_ = (expr);                                                           \
        if (e_ != hipSuccess) return fail(NR_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

__host__ __device__ inline uint16_t f2bf(float f) {          // round to nearest even
    uint32_t u;
    std::memcpy(&u, &f, 4);
    if ((u & 0x7F800000u) == 0x7F800000u) return (uint16_t)((u >> 16) | ((u & 0xFFFFu) ? 0x40u : 0u));
    u += 0x7FFFu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}
typedef __bf16 bf16x2 __attribute__((__vector_size__(2 * sizeof(__bf16))));
__device__ inline uint32_t pk_bf16(float lo, float hi) {   // v_cvt_pk_bf16_f32: round to nearest even
    bf16x2 v = {(__bf16)lo, (__bf16)hi};
    return __builtin_bit_cast(uint32_t, v);
}
__host__ __device__ inline float bf2f(uint16_t h) {
    uint32_t u = (uint32_t)h << 16;
    float f;
    std::memcpy(&f, &u, 4);
    return f;
}

struct View {             // element (pixel p, channel c) at p[pix * cs + off + c]
    uint16_t* p;
    int cs, off;
};

// ---------------------------------------------------------------------------
// network inputs
// ---------------------------------------------------------------------------
// Planar fp32 render layers of N images of H0 x W0: beauty, albedo, normal
// (N x 3 x H0 x W0) and depth (N x 1 x H0 x W0).
struct Inputs {
    const float *b, *al, *nm, *dp;
    int H0, W0;
};

// The 16 network input channels of pixel (n, y, x) (zero outside the image):
// log1p(beauty), log1p(albedo), the clamped normal, log1p(depth), 6 zeros.
// NaN samples become 0 (fmaxf returns the other operand).  Split into the ten
// loads (from pixel 0 where the pixel lies outside, so loads never branch) and
// the transform, so a loader can have every load in flight before it transforms.
__device__ inline bool input_load(const Inputs& in, int n, int y, int x, float raw[10]) {
    const bool inside = y >= 0 && x >= 0 && y < in.H0 && x < in.W0;
    const size_t hw = (size_t)in.H0 * in.W0, pix = inside ? (size_t)y * in.W0 + x : 0;
    const size_t base = (size_t)(inside ? n : 0) * 3 * hw + pix;
#pragma unroll
    for (int c = 0; c < 3; c++) {
        raw[c] = in.b[base + c * hw];
        raw[3 + c] = in.al[base + c * hw];
        raw[6 + c] = in.nm[base + c * hw];
    }
    raw[9] = in.dp[(size_t)(inside ? n : 0) * hw + pix];
    return inside;
}
__device__ inline void input_transform(const float raw[10], bool inside, float f[16]) {
#pragma unroll
    for (int k = 0; k < 16; k++) f[k] = 0.0f;
    if (!inside) return;
#pragma unroll
    for (int c = 0; c < 3; c++) {
        f[c] = log1pf(fmaxf(raw[c], 0.0f));
        f[3 + c] = log1pf(fmaxf(raw[3 + c], 0.0f));
        const float nv = raw[6 + c];
        f[6 + c] = (nv == nv) ? fminf(fmaxf(nv, -1.0f), 1.0f) : 0.0f;
    }
    f[9] = log1pf(fmaxf(raw[9], 0.0f));
}
__device__ inline void input_pixel(const Inputs& in, int n, int y, int x, float f[16]) {
    float raw[10];
    const bool inside = input_load(in, n, y, x, raw);
    input_transform(raw, inside, f);
}

// ---------------------------------------------------------------------------
// convolution (forward and input gradient)
// ---------------------------------------------------------------------------
// LDS-tiled: a workgroup (ROWS waves) computes ROWS rows x 64 pixels x (16 TM)
// output channels.  For each chunk of 32 input channels the (ROWS + 2) x 66-pixel
// halo tile and the chunk's weights of the workgroup's output channels are
// staged in LDS once and feed all 9 taps.  With several chunks (MC) the next
// chunk's tile and weights are loaded into registers while this chunk's MFMAs
// run, so a workgroup waits on memory once per chunk, not once per tap; a
// single-chunk layer holds no such registers (more waves per SIMD).  The tile is
// stored [row][8-channel group][column]: the 16 lanes that share a channel group
// read 16 consecutive 16-byte columns, one contiguous 256-byte run with no bank
// conflict.  The weights are stored [tap][8-channel group][output channel], so
// the 16 lanes that share a channel group read 16 consecutive rows, again one
// contiguous 256-byte run.  Global weights are [rows][9][c32].
enum { LD_PLAIN = 0, LD_UP = 1, LD_INPUT = 2 };      // tile loader
enum { EP_PLAIN = 0, EP_POOL = 1, EP_OUT = 2 };      // epilogue

struct ConvLdsArgs {
    const uint16_t* in; int in_cs, in_off, cin_avail;   // channels present in the input tensor
    const uint16_t* up; int up_cs, up_off, up_c0;      // LD_UP: channels >= up_c0 from the nearest 2x upsample of an (H/2 x W/2) tensor
    Inputs src;                                        // LD_INPUT: the planar render layers; EP_OUT: beauty for the residual
    int c32;                                           // padded channels per tap of the weights
    int N, H, W;
    const uint16_t* w;                                 // [rows][9][c32]
    const float* bias;
    int cout;
    int relu;
    uint16_t* out; int out_cs, out_off;
    uint16_t* pool; int pool_cs, pool_off;             // EP_POOL: also the 2x2 average of the stored values (H, W even)
    const uint16_t* mask; int mask_cs, mask_off;
    float* out_f32;                                    // EP_OUT: log-space result (N x H x W x 4), else
    float* out_final;                                  // EP_OUT: expm1(max(y, 0)), planar N x 3 x H0 x W0
};

// value of the lane next to this one (lane ^ 1), through DPP quad_perm [1,0,3,2]
__device__ inline float lane_xor1(float v) {
    return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), 0xB1, 0xF, 0xF, false));
}

constexpr int CL_TILE_W = 64;

// LD: which tile loader (LD_UP: the decoder's skip + upsampled channels; LD_INPUT:
// the network input computed from the render layers, one chunk).  EP: the
// epilogue (EP_POOL: the encoder's fused 2x2 average pool; EP_OUT: the output
// convolution, residual added).  ROWS: output rows (waves) per workgroup.  MC:
// more than one chunk (prefetch the next one).  All compile-time, so each variant
// carries only its own registers.
// Waves per SIMD the LDS allows (two 8-wave or three 4-wave workgroups per CU); the
// register budget is held to it, except that 48 output channels per workgroup get two
// waves' worth of registers (they would spill at three).
template <int TM, int LD, int EP, int ROWS, bool MC>
__global__ void __launch_bounds__(64 * ROWS) __attribute__((amdgpu_waves_per_eu(ROWS == 8 ? 4 : TM == 3 ? 2 : 3)))
k_conv_lds(ConvLdsArgs a) {
    constexpr int NT = 64 * ROWS;                                // threads
    __shared__ u32x4 tile[ROWS + 2][4][CL_TILE_W + 2];           // [row][8-channel group][column]
    constexpr int NW = 9 * 16 * TM * 4;                          // 16-byte weight entries per chunk
    constexpr int NWL = (NW + NT - 1) / NT;
    __shared__ u32x4 wl[NW];                                     // [tap][8-channel group][output channel]
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int r = lane & 15, kq = lane >> 4;
    const int x0 = blockIdx.x * CL_TILE_W;
    const int rowblocks = (a.H + ROWS - 1) / ROWS;
    const int n = blockIdx.y / rowblocks, y0 = (blockIdx.y - n * rowblocks) * ROWS;
    const int y = y0 + wave;
    const int cobase = blockIdx.z * 16 * TM;
    f32x4 acc[TM][4];
#pragma unroll
    for (int m = 0; m < TM; m++)
#pragma unroll
        for (int t = 0; t < 4; t++) acc[m][t] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
    // this lane's output-channel biases, loaded with the first chunk
    float bias[TM][4];
#pragma unroll
    for (int m = 0; m < TM; m++)
#pragma unroll
        for (int j = 0; j < 4; j++) {
            const int co = cobase + m * 16 + kq * 4 + j;
            bias[m][j] = (a.bias && co < a.cout) ? a.bias[co] : 0.0f;
        }
    const int nchunks = MC ? (a.c32 >> 5) : 1;
    constexpr int NV = (ROWS + 2) * (CL_TILE_W + 2) * 4;        // 16-byte tile entries
    constexpr int NLD = (NV + NT - 1) / NT;
    // LD_UP chunks made only of upsampled channels: the coarse tile, in the fine tile's memory
    constexpr int CROWS = ROWS / 2 + 2, CCOLS = CL_TILE_W / 2 + 2;
    constexpr int NVC = CROWS * CCOLS * 4;
    constexpr int NLDC = (NVC + NT - 1) / NT;
    static_assert(CROWS * CCOLS <= (ROWS + 2) * (CL_TILE_W + 2), "coarse tile fits the fine one");
    auto ctile = reinterpret_cast<u32x4 (*)[4][CCOLS]>(&tile[0][0][0]);
    constexpr int NP = (ROWS + 2) * (CL_TILE_W + 2);             // tile pixels
    constexpr int NPI = LD == LD_INPUT ? (NP + NT - 1) / NT : 1;  // LD_INPUT: tile pixels per thread
    u32x4 val[LD == LD_INPUT ? 1 : NLD], wv[NWL];
    u32x4 cval[LD == LD_UP ? NLDC : 1];                          // LD_UP: a coarse chunk's loads
    float raw[NPI][10];                                          // LD_INPUT: the render-layer values
    uint32_t okm = 0u;                                           // val[k] / raw[k] lies inside the image
    // EP_OUT: the residual's beauty values of this lane's pixels (used by lanes kq = 0)
    float resid[EP == EP_OUT ? 4 : 1][3];
    if constexpr (EP == EP_OUT) {
        const Inputs& sr = a.src;
        const size_t hw0 = (size_t)sr.H0 * sr.W0;
#pragma unroll
        for (int t = 0; t < 4; t++) {
            const int x = x0 + t * 16 + r;
            const bool inside = y < sr.H0 && x < sr.W0;
            const size_t q = inside ? (size_t)n * 3 * hw0 + (size_t)y * sr.W0 + x : 0;
#pragma unroll
            for (int j = 0; j < 3; j++) resid[t][j] = sr.b[q + j * hw0];
        }
    }
    // A chunk's weights: global -> registers (wv) -> LDS (wl).  Every load and store is
    // unconditional: a thread past entry NW loads and stores entry NW - 1 again, the same
    // value to the same place as its owner.
    auto load_w = [&](int c) {
#pragma unroll
        for (int k = 0; k < NWL; k++) {
            const int v = min((int)threadIdx.x + k * NT, NW - 1);
            const int q = v & 3, rt = v >> 2;
            const int row = rt / 9, tap = rt - row * 9;
            wv[k] = *reinterpret_cast<const u32x4*>(a.w + ((size_t)(cobase + row) * 9 + tap) * a.c32 + c * 32 + q * 8);
        }
    };
    auto store_w = [&]() {
#pragma unroll
        for (int k = 0; k < NWL; k++) {
            const int v = min((int)threadIdx.x + k * NT, NW - 1);
            const int q = v & 3, rt = v >> 2;
            const int row = rt / 9, tap = rt - row * 9;
            wl[(tap * 4 + q) * 16 * TM + row] = wv[k];
        }
    };
    // The fine tile of chunk c (input-tensor channels; LD_UP: channels from up_c0 on read
    // the coarse tensor at (y/2, x/2)).  Branch-free: entries outside the image (or past
    // the tile) load the tensor's first 16 bytes and are stored as zeros (okm), so no load
    // waits under divergent control flow.
    auto load_fine = [&](int c) {
#pragma unroll
        for (int k = 0; k < NLD; k++) {
            const int v = threadIdx.x + k * NT;
            const int g = v & 3, px = v >> 2;
            const int row = px / (CL_TILE_W + 2), col = px - row * (CL_TILE_W + 2);
            const int yy = y0 - 1 + row, xx = x0 - 1 + col, ch = c * 32 + g * 8;
            const bool ok = v < NV && yy >= 0 && yy < a.H && xx >= 0 && xx < a.W && ch < a.cin_avail;
            const uint16_t* src =
                !ok ? a.in
                    : (LD == LD_UP && ch >= a.up_c0)
                          ? a.up + ((size_t)(n * (a.H >> 1) + (yy >> 1)) * (a.W >> 1) + (xx >> 1)) * a.up_cs + a.up_off + (ch - a.up_c0)
                          : a.in + ((size_t)(n * a.H + yy) * a.W + xx) * a.in_cs + a.in_off + ch;
            val[k] = *reinterpret_cast<const u32x4*>(src);
            okm = ok ? (okm | (1u << k)) : (okm & ~(1u << k));
        }
    };
    auto store_fine = [&]() {
#pragma unroll
        for (int k = 0; k < NLD; k++) {
            const int v = threadIdx.x + k * NT;
            const int g = v & 3, px = v >> 2;
            const int row = px / (CL_TILE_W + 2), col = px - row * (CL_TILE_W + 2);
            if (v < NV) tile[row][g][col] = ((okm >> k) & 1u) ? val[k] : u32x4{0u, 0u, 0u, 0u};
        }
    };
    // LD_UP, a chunk made only of upsampled channels: its (ROWS/2 + 2) x 34 coarse tile
    // (a quarter of the loads), in the fine tile's memory
    auto load_coarse = [&](int c) {
#pragma unroll
        for (int k = 0; k < NLDC; k++) {
            const int v = threadIdx.x + k * NT;
            const int g = v & 3, cp = v >> 2;
            const int crow = cp / CCOLS, ccol = cp - crow * CCOLS;
            const int cy = (y0 >> 1) - 1 + crow, cx = (x0 >> 1) - 1 + ccol, ch = c * 32 + g * 8;
            const bool ok = v < NVC && cy >= 0 && cy < (a.H >> 1) && cx >= 0 && cx < (a.W >> 1) && ch < a.cin_avail;
            const uint16_t* src =
                !ok ? a.in : a.up + ((size_t)(n * (a.H >> 1) + cy) * (a.W >> 1) + cx) * a.up_cs + a.up_off + (ch - a.up_c0);
            cval[k] = *reinterpret_cast<const u32x4*>(src);
            okm = ok ? (okm | (1u << k)) : (okm & ~(1u << k));
        }
    };
    auto store_coarse = [&]() {
#pragma unroll
        for (int k = 0; k < NLDC; k++) {
            const int v = threadIdx.x + k * NT;
            const int g = v & 3, cp = v >> 2;
            const int crow = cp / CCOLS, ccol = cp - crow * CCOLS;
            if (v < NVC) ctile[crow][g][ccol] = ((okm >> k) & 1u) ? cval[k] : u32x4{0u, 0u, 0u, 0u};
        }
    };
    // LD_INPUT (one chunk): the render-layer values of this thread's tile pixels, then
    // the network input channels (groups 0-1; groups 2-3 are zero)
    auto load_input = [&]() {
#pragma unroll
        for (int k = 0; k < NPI; k++) {
            const int px = min((int)threadIdx.x + k * NT, NP - 1);
            const int row = px / (CL_TILE_W + 2), col = px - row * (CL_TILE_W + 2);
            const int yy = y0 - 1 + row, xx = x0 - 1 + col;
            const bool ok = yy < a.H && xx < a.W && input_load(a.src, n, yy, xx, raw[k]);
            okm = ok ? (okm | (1u << k)) : (okm & ~(1u << k));
        }
    };
    auto store_input = [&]() {
#pragma unroll
        for (int k = 0; k < NPI; k++) {
            const int px = threadIdx.x + k * NT;
            if (px >= NP) continue;
            const int row = px / (CL_TILE_W + 2), col = px - row * (CL_TILE_W + 2);
            float f[16];
            input_transform(raw[k], (okm >> k) & 1u, f);
            tile[row][0][col] = u32x4{pk_bf16(f[0], f[1]), pk_bf16(f[2], f[3]), pk_bf16(f[4], f[5]), pk_bf16(f[6], f[7])};
            tile[row][1][col] = u32x4{pk_bf16(f[8], f[9]), pk_bf16(f[10], f[11]), pk_bf16(f[12], f[13]),
                                      pk_bf16(f[14], f[15])};
            tile[row][2][col] = u32x4{0u, 0u, 0u, 0u};
            tile[row][3][col] = u32x4{0u, 0u, 0u, 0u};
        }
    };
    // The 9 taps of one staged chunk.  A fine tile position (row, col) of a coarse
    // chunk is the coarse tile's ((row + 1) / 2, (col + 1) / 2).
    auto mfma_chunk = [&](auto coarse_tag) {
        constexpr bool CO = decltype(coarse_tag)::value;
#pragma unroll
        for (int tap = 0; tap < 9; tap++) {
            const int dy = tap / 3, dx = tap - 3 * (tap / 3);
            bf16x8 b[4];
#pragma unroll
            for (int t = 0; t < 4; t++) {
                if constexpr (CO) b[t] = __builtin_bit_cast(bf16x8, ctile[(wave + dy + 1) >> 1][kq][(t * 16 + r + dx + 1) >> 1]);
                else b[t] = __builtin_bit_cast(bf16x8, tile[wave + dy][kq][t * 16 + r + dx]);
            }
#pragma unroll
            for (int m = 0; m < TM; m++) {
                const bf16x8 af = __builtin_bit_cast(bf16x8, wl[(tap * 4 + kq) * 16 * TM + m * 16 + r]);
#pragma unroll
                for (int t = 0; t < 4; t++)
                    acc[m][t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(af, b[t], acc[m][t], 0, 0, 0);
            }
        }
    };
    auto run = [&](auto coarse_tag, int cb, int ce) {
        constexpr bool CO = decltype(coarse_tag)::value;
        for (int c = cb; c < ce; c++) {
            load_w(c);
            if constexpr (LD == LD_INPUT) load_input();
            else if constexpr (CO) load_coarse(c);
            else load_fine(c);
            if (c > 0) __syncthreads();        // every wave is done with the previous chunk
            store_w();
            if constexpr (LD == LD_INPUT) store_input();
            else if constexpr (CO) store_coarse();
            else store_fine();
            __syncthreads();
            mfma_chunk(coarse_tag);
        }
    };
    if constexpr (LD == LD_UP) {
        // the chunks holding any skip channel, then those made only of upsampled ones
        const int nfine = min(nchunks, (a.up_c0 + 31) >> 5);
        run(std::false_type{}, 0, nfine);
        run(std::true_type{}, nfine, nchunks);
    } else {
        run(std::false_type{}, 0, nchunks);
    }
    if constexpr (EP == EP_POOL) {
        // Forward encoder output: bias + ReLU, bf16 store, and the 2x2 average pool of
        // the stored values.  The horizontal pair is lanes r, r^1; the vertical pair is
        // waves 2i, 2i+1 (tile rows are 4-aligned), exchanged through the (no longer
        // needed) input tile.  H and W are even, so a pair is entirely in or out.
        const bool rowv = y < a.H;
        float* xch = reinterpret_cast<float*>(&tile[0][0][0]);   // [ROWS/2][16][32] floats per m
#pragma unroll
        for (int m = 0; m < TM; m++) {
            const int co0 = cobase + m * 16 + kq * 4;
            float hs[4][4];
#pragma unroll
            for (int t = 0; t < 4; t++) {
                const int x = x0 + t * 16 + r;
                float v[4];
#pragma unroll
                for (int j = 0; j < 4; j++) {
                    float xv = acc[m][t][j] + bias[m][j];
                    v[j] = a.relu ? fmaxf(xv, 0.0f) : xv;
                }
                uint2 pk;
                pk.x = pk_bf16(v[0], v[1]);
                pk.y = pk_bf16(v[2], v[3]);
                if (rowv && x < a.W)
                    *reinterpret_cast<uint2*>(a.out + (((size_t)n * a.H + y) * a.W + x) * a.out_cs + a.out_off + co0) = pk;
                const float s[4] = {bf2f((uint16_t)(pk.x & 0xFFFFu)), bf2f((uint16_t)(pk.x >> 16)),
                                    bf2f((uint16_t)(pk.y & 0xFFFFu)), bf2f((uint16_t)(pk.y >> 16))};
#pragma unroll
                for (int j = 0; j < 4; j++) hs[t][j] = s[j] + lane_xor1(s[j]);
            }
            __syncthreads();                   // tile reads (or the previous m's exchange) are done
            if ((wave & 1) && !(r & 1)) {
#pragma unroll
                for (int t = 0; t < 4; t++)
#pragma unroll
                    for (int j = 0; j < 4; j++) xch[((wave >> 1) * 16 + t * 4 + j) * 32 + (lane >> 1)] = hs[t][j];
            }
            __syncthreads();
            if (!(wave & 1) && !(r & 1) && rowv) {
#pragma unroll
                for (int t = 0; t < 4; t++) {
                    const int x = x0 + t * 16 + r;
                    if (x >= a.W) continue;
                    float o[4];
#pragma unroll
                    for (int j = 0; j < 4; j++) o[j] = 0.25f * (hs[t][j] + xch[((wave >> 1) * 16 + t * 4 + j) * 32 + (lane >> 1)]);
                    uint2 pk;
                    pk.x = pk_bf16(o[0], o[1]);
                    pk.y = pk_bf16(o[2], o[3]);
                    const size_t q = ((size_t)n * (a.H >> 1) + (y >> 1)) * (a.W >> 1) + (x >> 1);
                    *reinterpret_cast<uint2*>(a.pool + q * a.pool_cs + a.pool_off + co0) = pk;
                }
            }
        }
        return;
    }
    if (y >= a.H) return;
    if constexpr (EP == EP_OUT) {
        // output convolution (channels 0-2 of the first 16-row tile, lanes kq = 0): the
        // residual log1p(beauty) added in fp32; either the log-space value for the loss
        // (out_f32, padded dims, 4-stride) or the final radiance expm1(max(y, 0)) into the
        // planar image of the unpadded size
        if (kq != 0) return;
        const Inputs& s = a.src;
        const size_t hw0 = (size_t)s.H0 * s.W0;
#pragma unroll
        for (int t = 0; t < 4; t++) {
            const int x = x0 + t * 16 + r;
            if (x >= a.W) continue;
            const bool inside = y < s.H0 && x < s.W0;
            const size_t p = ((size_t)n * a.H + y) * a.W + x;
            const size_t q = (size_t)n * 3 * hw0 + (size_t)y * s.W0 + x;
#pragma unroll
            for (int j = 0; j < 3; j++) {
                const float res = inside ? log1pf(fmaxf(resid[t][j], 0.0f)) : 0.0f;
                const float v = acc[0][t][j] + bias[0][j] + res;
                if (a.out_f32) a.out_f32[p * 4 + j] = v;
                else if (inside) a.out_final[q + j * hw0] = expm1f(fmaxf(v, 0.0f));
            }
            if (a.out_f32) a.out_f32[p * 4 + 3] = 0.0f;
        }
        return;
    }
#pragma unroll
    for (int m = 0; m < TM; m++)
#pragma unroll
        for (int t = 0; t < 4; t++) {
            const int x = x0 + t * 16 + r;
            if (x >= a.W) continue;
            const size_t p = ((size_t)n * a.H + y) * a.W + x;
            const int co0 = cobase + m * 16 + kq * 4;
            float v[4];
#pragma unroll
            for (int j = 0; j < 4; j++) {
                const int co = co0 + j;
                float xv = acc[m][t][j] + bias[m][j];
                if (a.relu) xv = fmaxf(xv, 0.0f);
                if (a.mask && co < a.cout && !(bf2f(a.mask[p * a.mask_cs + a.mask_off + co]) > 0.0f)) xv = 0.0f;
                v[j] = xv;
            }
            uint16_t* o = a.out + p * a.out_cs + a.out_off + co0;
            if (co0 + 3 < a.cout) {
                uint2 pk;
                pk.x = pk_bf16(v[0], v[1]);
                pk.y = pk_bf16(v[2], v[3]);
                *reinterpret_cast<uint2*>(o) = pk;
            } else {
                for (int j = 0; j < 4; j++)
                    if (co0 + j < a.cout) o[j] = f2bf(v[j]);
            }
        }
}

// ---------------------------------------------------------------------------
// weight gradient: one wave per (tap, 16 co, 16 ci) block over a pixel range
// ---------------------------------------------------------------------------
struct WgradArgs {
    const uint16_t* x; int x_cs, x_off, cin;
    const uint16_t* up; int up_cs, up_off, up_c0;   // channels >= up_c0: nearest 2x upsample of an H/2 x W/2 tensor
    const uint16_t* g; int g_cs, g_off, cout;
    int N, H, W;
    int nci;              // ci blocks of 16
    int chunks_per_wg;    // 32-pixel chunks per workgroup
    float* dw;            // [cout][9][cin]
};

__global__ void __launch_bounds__(64) k_wgrad(WgradArgs a) {
    __shared__ __attribute__((aligned(16))) uint16_t gT[16 * 40];   // [co][pix], rows padded to 40
    __shared__ __attribute__((aligned(16))) uint16_t xT[16 * 40];   // [ci][pix]
    const int lane = threadIdx.x;
    const int tap = blockIdx.y;
    const int co0 = (blockIdx.z / a.nci) * 16, ci0 = (blockIdx.z % a.nci) * 16;
    const int dy = tap / 3 - 1, dx = tap % 3 - 1;
    const int P = a.N * a.H * a.W;
    const int nchunks = (P + 31) / 32;
    const int c_beg = blockIdx.x * a.chunks_per_wg;
    const int c_end = min(nchunks, c_beg + a.chunks_per_wg);
    const int lp = lane >> 1, half = lane & 1;
    f32x4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
    for (int c = c_beg; c < c_end; c++) {
        const int p = c * 32 + lp;
        uint4 gv = make_uint4(0u, 0u, 0u, 0u), xv = make_uint4(0u, 0u, 0u, 0u);
        if (p < P) {
            gv = *reinterpret_cast<const uint4*>(a.g + (size_t)p * a.g_cs + a.g_off + co0 + half * 8);
            const int px = p % a.W, py = (p / a.W) % a.H;
            const int yy = py + dy, xx = px + dx;
            if (yy >= 0 && yy < a.H && xx >= 0 && xx < a.W) {
                const int ch = ci0 + half * 8;
                const int n = p / (a.W * a.H);
                const uint16_t* src =
                    (a.up && ch >= a.up_c0)
                        ? a.up + ((size_t)(n * (a.H >> 1) + (yy >> 1)) * (a.W >> 1) + (xx >> 1)) * a.up_cs + a.up_off + (ch - a.up_c0)
                        : a.x + (size_t)(p + dy * a.W + dx) * a.x_cs + a.x_off + ch;
                xv = *reinterpret_cast<const uint4*>(src);
            }
        }
        const uint16_t* gs = reinterpret_cast<const uint16_t*>(&gv);
        const uint16_t* xs = reinterpret_cast<const uint16_t*>(&xv);
        for (int j = 0; j < 8; j++) {
            gT[(half * 8 + j) * 40 + lp] = gs[j];
            xT[(half * 8 + j) * 40 + lp] = xs[j];
        }
        __syncthreads();
        const int r = lane & 15, kq = lane >> 4;
        uint4 av = *reinterpret_cast<const uint4*>(&gT[r * 40 + kq * 8]);
        uint4 bv = *reinterpret_cast<const uint4*>(&xT[r * 40 + kq * 8]);
        acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8, av), __builtin_bit_cast(bf16x8, bv),
                                                      acc, 0, 0, 0);
        __syncthreads();
    }
    const int ci = ci0 + (lane & 15);
    for (int j = 0; j < 4; j++) {
        const int co = co0 + (lane >> 4) * 4 + j;
        if (co < a.cout && ci < a.cin) atomicAdd(&a.dw[((size_t)co * 9 + tap) * a.cin + ci], acc[j]);
    }
}

__global__ void __launch_bounds__(256) k_bias_grad(const uint16_t* g, int g_cs, int g_off, int cout, int cpad, int P,
                                                   int pix_per_block, float* db) {
    const int lanes_per_pix = cpad;
    const int phases = 256 / lanes_per_pix;
    const int t = threadIdx.x;
    const int c = t % lanes_per_pix, ph = t / lanes_per_pix;
    if (ph >= phases) return;
    const int p0 = blockIdx.x * pix_per_block;
    const int p1 = min(P, p0 + pix_per_block);
    float s = 0.0f;
    for (int p = p0 + ph; p < p1; p += phases) s += bf2f(g[(size_t)p * g_cs + g_off + c]);
    if (c < cout) atomicAdd(&db[c], s);
}

// ---------------------------------------------------------------------------
// pooling, upsampling, input/output transforms, loss, optimizer
// ---------------------------------------------------------------------------
// The forward pass pools in the encoder convolution's epilogue and upsamples in
// the decoder convolution's loader; their adjoints below move 8 channels (16 B)
// per thread (every tensor's channel count, stride and offset are multiples of 8).
__device__ inline void ld8(const uint16_t* p, float* f) {
    uint4 v = *reinterpret_cast<const uint4*>(p);
    const uint32_t w[4] = {v.x, v.y, v.z, v.w};
    for (int k = 0; k < 4; k++) { f[2 * k] = bf2f((uint16_t)(w[k] & 0xFFFFu)); f[2 * k + 1] = bf2f((uint16_t)(w[k] >> 16)); }
}
__device__ inline void st8(uint16_t* p, const float* f) {
    uint32_t w[4];
    for (int k = 0; k < 4; k++) w[k] = pk_bf16(f[2 * k], f[2 * k + 1]);
    *reinterpret_cast<uint4*>(p) = make_uint4(w[0], w[1], w[2], w[3]);
}

// d[p] = (skip[p] + 0.25 * gdown[p/2]) * (mask[p] > 0); dims of d: H x W
__global__ void k_pool_bwd(View gdown, View skip, View mask, View d, int C, int N, int H, int W) {
    const int C8 = C >> 3;
    const size_t total = (size_t)N * H * W * C8;
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= total) return;
    const int c = (int)(i % C8) * 8;
    const size_t p = i / C8;
    const int x = (int)(p % W), y = (int)((p / W) % H), n = (int)(p / ((size_t)W * H));
    const size_t q = ((size_t)n * (H >> 1) + (y >> 1)) * (W >> 1) + (x >> 1);
    float g[8], sk[8], m[8], o[8];
    ld8(gdown.p + q * gdown.cs + gdown.off + c, g);
    ld8(mask.p + p * mask.cs + mask.off + c, m);
    if (skip.p) ld8(skip.p + p * skip.cs + skip.off + c, sk);
    for (int k = 0; k < 8; k++) {
        float v = 0.25f * g[k];
        if (skip.p) v += sk[k];
        o[k] = m[k] > 0.0f ? v : 0.0f;
    }
    st8(d.p + p * d.cs + d.off + c, o);
}

// d[q] = sum of g over the 2x2 children of q, times (mask[q] > 0); dst dims Ho x Wo
__global__ void k_up_bwd(View g, View mask, View d, int C, int N, int Ho, int Wo) {
    const int C8 = C >> 3;
    const size_t total = (size_t)N * Ho * Wo * C8;
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= total) return;
    const int c = (int)(i % C8) * 8;
    const size_t q = i / C8;
    const int xo = (int)(q % Wo), yo = (int)((q / Wo) % Ho), n = (int)(q / ((size_t)Wo * Ho));
    const int W = Wo * 2;
    const size_t p00 = ((size_t)n * (Ho * 2) + 2 * yo) * W + 2 * xo;
    float a[8], b[8], e[8], f[8], m[8], o[8];
    ld8(g.p + p00 * g.cs + g.off + c, a);
    ld8(g.p + (p00 + 1) * g.cs + g.off + c, b);
    ld8(g.p + (p00 + W) * g.cs + g.off + c, e);
    ld8(g.p + (p00 + W + 1) * g.cs + g.off + c, f);
    ld8(mask.p + q * mask.cs + mask.off + c, m);
    for (int k = 0; k < 8; k++) o[k] = m[k] > 0.0f ? (a[k] + b[k] + e[k] + f[k]) : 0.0f;
    st8(d.p + q * d.cs + d.off + c, o);
}

// the network input of N images (padded dims Hp x Wp) as the 16-channel bf16 X0; the
// first convolution of the inference forward computes the same values in its loader
__global__ void k_prep(Inputs in, int N, int Hp, int Wp, uint16_t* x0) {
    const size_t total = (size_t)N * Hp * Wp;
    size_t p = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= total) return;
    const int x = (int)(p % Wp), y = (int)((p / Wp) % Hp), n = (int)(p / ((size_t)Wp * Hp));
    float f[16];
    input_pixel(in, n, y, x, f);
    uint32_t pk[8];
    for (int k = 0; k < 8; k++) pk[k] = pk_bf16(f[2 * k], f[2 * k + 1]);
    uint4* o = reinterpret_cast<uint4*>(x0 + p * 16);
    o[0] = make_uint4(pk[0], pk[1], pk[2], pk[3]);
    o[1] = make_uint4(pk[4], pk[5], pk[6], pk[7]);
}

__global__ void k_target(const float* t, int N, int H, int W, float* tlog) {
    const size_t total = (size_t)N * H * W;
    size_t p = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= total) return;
    const int x = (int)(p % W), y = (int)((p / W) % H), n = (int)(p / ((size_t)W * H));
    const size_t hw = (size_t)H * W, base = (size_t)n * 3 * hw + (size_t)y * W + x;
    for (int c = 0; c < 4; c++) tlog[p * 4 + c] = c < 3 ? log1pf(fmaxf(t[base + c * hw], 0.0f)) : 0.0f;
}

// L1 loss on log radiance; gradient into a 16-channel bf16 tensor; per-block partial sums
__global__ void __launch_bounds__(256) k_loss(const float* y, const float* t, int P, float inv_count, uint16_t* gy,
                                              float* partial) {
    __shared__ float red[256];
    size_t p = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    float s = 0.0f;
    if (p < (size_t)P) {
        uint16_t g[16];
        for (int c = 0; c < 16; c++) g[c] = 0;
        for (int c = 0; c < 3; c++) {
            float d = y[p * 4 + c] - t[p * 4 + c];
            s += fabsf(d);
            g[c] = f2bf(d > 0.0f ? inv_count : (d < 0.0f ? -inv_count : 0.0f));
        }
        uint32_t pk[8];
        for (int k = 0; k < 8; k++) pk[k] = (uint32_t)g[2 * k] | ((uint32_t)g[2 * k + 1] << 16);
        uint4* o = reinterpret_cast<uint4*>(gy + p * 16);
        o[0] = make_uint4(pk[0], pk[1], pk[2], pk[3]);
        o[1] = make_uint4(pk[4], pk[5], pk[6], pk[7]);
    }
    red[threadIdx.x] = s;
    __syncthreads();
    for (int k = 128; k > 0; k >>= 1) {
        if ((int)threadIdx.x < k) red[threadIdx.x] += red[threadIdx.x + k];
        __syncthreads();
    }
    if (threadIdx.x == 0) partial[blockIdx.x] = red[0];
}

__global__ void k_adam(float* prm, const float* g, float* m, float* v, size_t n, float lr, float b1, float b2,
                       float c1, float c2, float eps) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float gi = g[i];
    float mi = b1 * m[i] + (1.0f - b1) * gi;
    float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    prm[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

// fp32 [cout][9][cin] -> bf16 [rows][9][c32] (forward: rows = output channels)
__global__ void k_pack_fwd_tap(const float* w, int cout, int cin, int c32, int rows, uint16_t* o) {
    const size_t total = (size_t)rows * 9 * c32;
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= total) return;
    const int ci = (int)(i % c32);
    const int tap = (int)((i / c32) % 9);
    const int co = (int)(i / ((size_t)9 * c32));
    o[i] = f2bf(co < cout && ci < cin ? w[((size_t)co * 9 + tap) * cin + ci] : 0.0f);
}

// input-gradient weights, rows = input channels: element (ci, tap, co) = W[co][8 - tap][ci]
__global__ void k_pack_bwd_tap(const float* w, int cout, int cin, int c32, int rows, uint16_t* o) {
    const size_t total = (size_t)rows * 9 * c32;
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= total) return;
    const int co = (int)(i % c32);
    const int tap = (int)((i / c32) % 9);
    const int ci = (int)(i / ((size_t)9 * c32));
    o[i] = f2bf(ci < cin && co < cout ? w[((size_t)co * 9 + (8 - tap)) * cin + ci] : 0.0f);
}

inline int up32(int x) { return (x + 31) & ~31; }
inline int up16(int x) { return (x + 15) & ~15; }
inline unsigned blocks(size_t n, unsigned b) { return (unsigned)((n + b - 1) / b); }

// ---------------------------------------------------------------------------
// network description
// ---------------------------------------------------------------------------
constexpr int LV = 5;
constexpr int CH[LV] = {32, 48, 64, 96, 128};
constexpr int IN_CH = 10, IN_PAD = 16;

struct Layer {
    int cin, cout;            // actual channels of the parameters
    int cin_t;                // channels of the input tensor the forward kernel reads (multiple of 8)
    size_t w_off, b_off;      // offsets into the flat parameter vector
    int cout_pad, cin_pad;    // rows of the forward / input-gradient packings (multiples of 16)
    int fc32, bc32;           // channels per tap of those packings (multiples of 32)
    uint16_t* wf2 = nullptr;
    uint16_t* wb2 = nullptr;
};

struct Net {
    std::vector<Layer> L;     // enc0a..enc4b (10), dec3a..dec0b (8), out (1)
    size_t n_params = 0;
    Net() {
        auto add = [&](int cin, int cout, int cin_t) {
            Layer l{};
            l.cin = cin; l.cout = cout; l.cin_t = cin_t;
            l.w_off = n_params; n_params += (size_t)cout * 9 * cin;
            l.b_off = n_params; n_params += (size_t)cout;
            l.cout_pad = up16(cout);
            l.cin_pad = up16(cin);
            l.fc32 = up32(cin_t);
            l.bc32 = up32(l.cout_pad);
            L.push_back(l);
        };
        int prev = IN_CH, prev_t = IN_PAD;
        for (int l = 0; l < LV; l++) {
            add(prev, CH[l], prev_t);
            add(CH[l], CH[l], CH[l]);
            prev = prev_t = CH[l];
        }
        for (int l = LV - 2; l >= 0; l--) {
            add(CH[l] + CH[l + 1], CH[l], CH[l] + CH[l + 1]);
            add(CH[l], CH[l], CH[l]);
        }
        add(CH[0], 3, CH[0]);
    }
    Layer& enc(int l, int b) { return L[2 * l + b]; }
    Layer& dec(int l, int b) { return L[10 + 2 * (LV - 2 - l) + b]; }
    Layer& out() { return L[18]; }
};

}  // namespace

struct nr_unet {
    int device = 0;
    hipStream_t stream = nullptr;
    Net net;
    float* prm = nullptr;      // fp32 master parameters
    float* grad = nullptr;
    float* m = nullptr;
    float* v = nullptr;
    uint16_t* wmem = nullptr;  // packed bf16 weights of all layers
    int adam_t = 0;
    // activations for the current shape
    int N = 0, H = 0, W = 0;
    void* amem = nullptr;
    size_t amem_bytes = 0;
    // X: level input, A: first conv, E: encoder output (skip), DA/DO: decoder convs
    uint16_t *X[LV] = {}, *A[LV] = {}, *E[LV] = {}, *DA[LV] = {}, *DO[LV] = {};
    uint16_t *gX[LV] = {}, *gA[LV] = {}, *gE[LV] = {}, *gD[LV] = {}, *gDA[LV] = {}, *gDO[LV] = {}, *gY = nullptr;
    float *y = nullptr, *tlog = nullptr, *partial = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    float last_ms = 0.0f;
    double last_flops = 0.0;
};

namespace {

int pack_all(nr_unet* u) {
    for (Layer& l : u->net.L) {
        size_t nf2 = (size_t)l.cout_pad * 9 * l.fc32;
        hipLaunchKernelGGL(k_pack_fwd_tap, dim3(blocks(nf2, 256)), dim3(256), 0, u->stream, u->prm + l.w_off, l.cout,
                           l.cin, l.fc32, l.cout_pad, l.wf2);
        size_t nb2 = (size_t)l.cin_pad * 9 * l.bc32;
        hipLaunchKernelGGL(k_pack_bwd_tap, dim3(blocks(nb2, 256)), dim3(256), 0, u->stream, u->prm + l.w_off, l.cout,
                           l.cin, l.bc32, l.cin_pad, l.wb2);
    }
    UCK(hipGetLastError());
    return NR_OK;
}

// (Re)allocate activation and gradient tensors for N images of H x W (multiples of 16).
int ensure_shape(nr_unet* u, int N, int H, int W) {
    if (u->N == N && u->H == H && u->W == W) return NR_OK;
    if (u->amem) { UCK(hipFree(u->amem)); u->amem = nullptr; }
    size_t off = 0;
    std::vector<std::pair<void**, size_t>> plan;
    auto want = [&](void** p, size_t bytes) { plan.push_back({p, bytes}); };
    for (int l = 0; l < LV; l++) {
        const size_t P = (size_t)N * (H >> l) * (W >> l);
        const int cin = l == 0 ? IN_PAD : CH[l - 1];
        want((void**)&u->X[l], P * cin * 2);
        want((void**)&u->A[l], P * CH[l] * 2);
        want((void**)&u->gX[l], P * cin * 2);
        want((void**)&u->gA[l], P * CH[l] * 2);
        want((void**)&u->gE[l], P * CH[l] * 2);
        want((void**)&u->E[l], P * CH[l] * 2);
        if (l < LV - 1) {
            const int cd = CH[l] + CH[l + 1];
            want((void**)&u->gD[l], P * cd * 2);
            want((void**)&u->DA[l], P * CH[l] * 2);
            want((void**)&u->DO[l], P * CH[l] * 2);
            want((void**)&u->gDA[l], P * CH[l] * 2);
            want((void**)&u->gDO[l], P * CH[l] * 2);
        }
    }
    const size_t P0 = (size_t)N * H * W;
    want((void**)&u->gY, P0 * 16 * 2);
    want((void**)&u->y, P0 * 4 * 4);
    want((void**)&u->tlog, P0 * 4 * 4);
    want((void**)&u->partial, (P0 / 256 + 1) * 4);
    for (auto& pr : plan) off += (pr.second + 255) & ~(size_t)255;
    UCK(hipMalloc(&u->amem, off));
    u->amem_bytes = off;
    size_t o = 0;
    for (auto& pr : plan) {
        *pr.first = (char*)u->amem + o;
        o += (pr.second + 255) & ~(size_t)255;
    }
    u->N = N; u->H = H; u->W = W;
    return NR_OK;
}

double conv_flops(const Layer& l, size_t P) { return 2.0 * P * 9.0 * l.cin * l.cout; }

// What a convolution reads besides its input tensor and writes besides its output.
struct ConvOpt {
    View up{nullptr, 0, 0};        // input channels from cin_groups*8 on: the nearest 2x upsample of this H/2 x W/2 tensor
    int up_groups = 0;
    View pool{nullptr, 0, 0};      // also write the 2x2 average of the (ReLU, bf16) output here
    View mask{nullptr, 0, 0};      // input gradient: zero where the forward activation is not positive
    const Inputs* input = nullptr; // first layer: compute the tile from these render layers
    const Inputs* resid = nullptr; // output layer: add log1p(beauty) of these layers; write out_f32 or out_final
    float* out_f32 = nullptr;
    float* out_final = nullptr;
};

template <int TM, int LD, int EP>
void launch_conv_shape(int rows, bool mc, dim3 grid, hipStream_t st, const ConvLdsArgs& a) {
    constexpr bool single = LD == LD_INPUT || EP == EP_OUT;     // one chunk by construction
    if (rows == 8) {
        if (mc && !single) hipLaunchKernelGGL((k_conv_lds<TM, LD, EP, 8, !single>), grid, dim3(512), 0, st, a);
        else hipLaunchKernelGGL((k_conv_lds<TM, LD, EP, 8, false>), grid, dim3(512), 0, st, a);
    } else {
        if (mc && !single) hipLaunchKernelGGL((k_conv_lds<TM, LD, EP, 4, !single>), grid, dim3(256), 0, st, a);
        else hipLaunchKernelGGL((k_conv_lds<TM, LD, EP, 4, false>), grid, dim3(256), 0, st, a);
    }
}

void launch_conv(nr_unet* u, const Layer& l, bool backward, View in, int cin_groups, int N, int H, int W, View out,
                 int cout, bool relu, const float* bias, const ConvOpt& o = ConvOpt{}) {
    ConvLdsArgs a{};
    a.in = in.p; a.in_cs = in.cs; a.in_off = in.off; a.cin_avail = (cin_groups + o.up_groups) * 8;
    a.up = o.up.p; a.up_cs = o.up.cs; a.up_off = o.up.off; a.up_c0 = cin_groups * 8;
    a.pool = o.pool.p; a.pool_cs = o.pool.cs; a.pool_off = o.pool.off;
    if (o.input) a.src = *o.input;
    if (o.resid) a.src = *o.resid;
    a.c32 = backward ? l.bc32 : l.fc32;
    a.N = N; a.H = H; a.W = W;
    a.w = backward ? l.wb2 : l.wf2;
    a.bias = bias;
    a.cout = cout;
    a.relu = relu;
    a.out = out.p; a.out_cs = out.cs; a.out_off = out.off;
    a.mask = o.mask.p; a.mask_cs = o.mask.cs; a.mask_off = o.mask.off;
    a.out_f32 = o.out_f32; a.out_final = o.out_final;
    const int cpad = up16(cout);
    int TM = (cpad % 48 == 0) ? 3 : (cpad % 32 == 0) ? 2 : 1;
    if (o.resid) TM = 1;                       // three output channels: one 16-row tile
    const int64_t cols = (W + CL_TILE_W - 1) / CL_TILE_W;
    // 8 rows (waves) per workgroup where the layer still has at least 2048 workgroups
    // (full resolution: fewer halo rows per output row, more waves per CU); else 4
    const int rows = cols * N * ((H + 7) / 8) * (cpad / (16 * TM)) >= 2048 ? 8 : 4;
    const int64_t tiles = cols * N * ((H + rows - 1) / rows);
    // coarse levels have few pixel tiles: fewer output channels per workgroup
    // (same arithmetic per channel) until the launch has about 4 workgroups per CU
    while (TM > 1 && tiles * (cpad / (16 * TM)) < 1024) {
        int t = TM - 1;
        while (t > 1 && (cpad / 16) % t != 0) t--;
        TM = t;
    }
    const bool mc = a.c32 > 32;
    dim3 grid((unsigned)cols, (unsigned)(N * ((H + rows - 1) / rows)), cpad / (16 * TM));
    hipStream_t st = u->stream;
    if (o.resid) {
        launch_conv_shape<1, LD_PLAIN, EP_OUT>(rows, false, grid, st, a);
    } else if (o.input) {
        if (TM == 2) launch_conv_shape<2, LD_INPUT, EP_PLAIN>(rows, false, grid, st, a);
        else launch_conv_shape<1, LD_INPUT, EP_PLAIN>(rows, false, grid, st, a);
    } else {
        const int mode = a.pool ? 2 : a.up ? 1 : 0;
#define NR_CONV_CASE(T)                                                                \
    case T:                                                                            \
        if (mode == 2) launch_conv_shape<T, LD_PLAIN, EP_POOL>(rows, mc, grid, st, a);  \
        else if (mode == 1) launch_conv_shape<T, LD_UP, EP_PLAIN>(rows, mc, grid, st, a); \
        else launch_conv_shape<T, LD_PLAIN, EP_PLAIN>(rows, mc, grid, st, a);          \
        break;
        switch (TM) {
            NR_CONV_CASE(3)
            NR_CONV_CASE(2)
            default: NR_CONV_CASE(1)
        }
#undef NR_CONV_CASE
    }
}

void launch_wgrad(nr_unet* u, const Layer& l, View x, int cin, View g, int cout, int N, int H, int W,
                  View up = View{nullptr, 0, 0}, int up_c0 = 0) {
    WgradArgs a{};
    a.x = x.p; a.x_cs = x.cs; a.x_off = x.off; a.cin = cin;
    a.up = up.p; a.up_cs = up.cs; a.up_off = up.off; a.up_c0 = up_c0;
    a.g = g.p; a.g_cs = g.cs; a.g_off = g.off; a.cout = cout;
    a.N = N; a.H = H; a.W = W;
    a.nci = up16(cin) / 16;
    const int nco = up16(cout) / 16;
    const int P = N * H * W;
    const int nchunks = (P + 31) / 32;
    // about 2048 workgroups per layer
    const int per_tap_blocks = 9 * nco * a.nci;
    int want_x = std::max(1, 2048 / per_tap_blocks);
    a.chunks_per_wg = std::max(4, (nchunks + want_x - 1) / want_x);
    a.dw = u->grad + l.w_off;
    dim3 grid((nchunks + a.chunks_per_wg - 1) / a.chunks_per_wg, 9, nco * a.nci);
    hipLaunchKernelGGL(k_wgrad, grid, dim3(64), 0, u->stream, a);
    const int cpad = up16(cout);
    const int ppb = 4096;
    hipLaunchKernelGGL(k_bias_grad, dim3(blocks(P, ppb)), dim3(256), 0, u->stream, g.p, g.cs, g.off, cout, cpad, P, ppb,
                       u->grad + l.b_off);
}

// The forward pass.  Training (`in` null): the first layer reads X0, which k_prep
// wrote for the weight gradient, and the output layer writes the log-space result
// to y for the loss.  Inference (`in` given): the first layer computes its input
// tile from the render layers and the output layer writes the final radiance to
// `final`.  Both read the same bf16 inputs and compute the same values.
void forward(nr_unet* u, int N, int H, int W, const Inputs& src, bool fused_input, float* final_out) {
    Net& n = u->net;
    for (int l = 0; l < LV; l++) {
        const int h = H >> l, w = W >> l;
        const int cin_t = l == 0 ? IN_PAD : CH[l - 1];
        View x{u->X[l], cin_t, 0}, a{u->A[l], CH[l], 0}, e{u->E[l], CH[l], 0};
        Layer& la = n.enc(l, 0);
        Layer& lb = n.enc(l, 1);
        ConvOpt oa;
        if (l == 0 && fused_input) oa.input = &src;
        launch_conv(u, la, false, x, cin_t / 8, N, h, w, a, CH[l], true, u->prm + la.b_off, oa);
        ConvOpt ob;
        if (l < LV - 1) ob.pool = View{u->X[l + 1], CH[l], 0};
        launch_conv(u, lb, false, a, CH[l] / 8, N, h, w, e, CH[l], true, u->prm + lb.b_off, ob);
    }
    for (int l = LV - 2; l >= 0; l--) {
        const int h = H >> l, w = W >> l;
        Layer& la = n.dec(l, 0);
        Layer& lb = n.dec(l, 1);
        ConvOpt oa;
        oa.up = l == LV - 2 ? View{u->E[LV - 1], CH[l + 1], 0} : View{u->DO[l + 1], CH[l + 1], 0};
        oa.up_groups = CH[l + 1] / 8;
        launch_conv(u, la, false, View{u->E[l], CH[l], 0}, CH[l] / 8, N, h, w, View{u->DA[l], CH[l], 0}, CH[l], true,
                    u->prm + la.b_off, oa);
        launch_conv(u, lb, false, View{u->DA[l], CH[l], 0}, CH[l] / 8, N, h, w, View{u->DO[l], CH[l], 0}, CH[l], true,
                    u->prm + lb.b_off);
    }
    Layer& lo = n.out();
    ConvOpt oo;
    oo.resid = &src;
    if (final_out) oo.out_final = final_out;
    else oo.out_f32 = u->y;
    launch_conv(u, lo, false, View{u->DO[0], CH[0], 0}, CH[0] / 8, N, H, W, View{nullptr, 0, 0}, 3, false,
                u->prm + lo.b_off, oo);
}

void backward(nr_unet* u, int N, int H, int W) {
    Net& n = u->net;
    auto masked = [](View m) { ConvOpt o; o.mask = m; return o; };
    Layer& lo = n.out();
    // output layer
    launch_wgrad(u, lo, View{u->DO[0], CH[0], 0}, CH[0], View{u->gY, 16, 0}, 3, N, H, W);
    launch_conv(u, lo, true, View{u->gY, 16, 0}, 2, N, H, W, View{u->gDO[0], CH[0], 0}, CH[0], false,
                    nullptr, masked(View{u->DO[0], CH[0], 0}));
    // decoder, finest level first
    for (int l = 0; l < LV - 1; l++) {
        const int h = H >> l, w = W >> l;
        const int cd = CH[l] + CH[l + 1];
        Layer& la = n.dec(l, 0);
        Layer& lb = n.dec(l, 1);
        launch_wgrad(u, lb, View{u->DA[l], CH[l], 0}, CH[l], View{u->gDO[l], CH[l], 0}, CH[l], N, h, w);
        launch_conv(u, lb, true, View{u->gDO[l], CH[l], 0}, CH[l] / 8, N, h, w, View{u->gDA[l], CH[l], 0}, CH[l], false,
                    nullptr, masked(View{u->DA[l], CH[l], 0}));
        View coarse = l == LV - 2 ? View{u->E[LV - 1], CH[l + 1], 0} : View{u->DO[l + 1], CH[l + 1], 0};
        launch_wgrad(u, la, View{u->E[l], CH[l], 0}, cd, View{u->gDA[l], CH[l], 0}, CH[l], N, h, w, coarse, CH[l]);
        launch_conv(u, la, true, View{u->gDA[l], CH[l], 0}, CH[l] / 8, N, h, w, View{u->gD[l], cd, 0}, cd, false,
                    nullptr);
        // upsampled half of the concatenation -> the coarser decoder output (or the bottleneck)
        View dst = l == LV - 2 ? View{u->gE[LV - 1], CH[l + 1], 0} : View{u->gDO[l + 1], CH[l + 1], 0};
        size_t tot = (size_t)N * (h >> 1) * (w >> 1) * (CH[l + 1] / 8);
        hipLaunchKernelGGL(k_up_bwd, dim3(blocks(tot, 256)), dim3(256), 0, u->stream, View{u->gD[l], cd, CH[l]},
                           coarse, dst, CH[l + 1], N, h >> 1, w >> 1);
    }
    // encoder, coarsest level first
    for (int l = LV - 1; l >= 0; l--) {
        const int h = H >> l, w = W >> l;
        const int cin_t = l == 0 ? IN_PAD : CH[l - 1];
        const int cin = l == 0 ? IN_CH : CH[l - 1];
        Layer& la = n.enc(l, 0);
        Layer& lb = n.enc(l, 1);
        if (l < LV - 1) {
            const int cd = CH[l] + CH[l + 1];
            size_t tot = (size_t)N * h * w * (CH[l] / 8);
            hipLaunchKernelGGL(k_pool_bwd, dim3(blocks(tot, 256)), dim3(256), 0, u->stream,
                               View{u->gX[l + 1], CH[l], 0}, View{u->gD[l], cd, 0}, View{u->E[l], CH[l], 0},
                               View{u->gE[l], CH[l], 0}, CH[l], N, h, w);
        }
        launch_wgrad(u, lb, View{u->A[l], CH[l], 0}, CH[l], View{u->gE[l], CH[l], 0}, CH[l], N, h, w);
        launch_conv(u, lb, true, View{u->gE[l], CH[l], 0}, CH[l] / 8, N, h, w, View{u->gA[l], CH[l], 0}, CH[l], false,
                    nullptr, masked(View{u->A[l], CH[l], 0}));
        launch_wgrad(u, la, View{u->X[l], cin_t, 0}, cin, View{u->gA[l], CH[l], 0}, CH[l], N, h, w);
        if (l > 0)
            launch_conv(u, la, true, View{u->gA[l], CH[l], 0}, CH[l] / 8, N, h, w, View{u->gX[l], cin_t, 0}, cin, false,
                    nullptr);
    }
}

double net_flops(nr_unet* u, size_t P0) {
    double f = 0.0;
    Net& n = u->net;
    for (int l = 0; l < LV; l++) {
        size_t P = P0 >> (2 * l);
        f += conv_flops(n.enc(l, 0), P) + conv_flops(n.enc(l, 1), P);
        if (l < LV - 1) f += conv_flops(n.dec(l, 0), P) + conv_flops(n.dec(l, 1), P);
    }
    return f + conv_flops(n.out(), P0);
}

int check_dims(int N, int H, int W) {
    if (N <= 0 || H <= 0 || W <= 0) return fail(NR_ERR_INVALID, "bad image dimensions");
    if ((size_t)N * H * W * 224 * 2 > ((size_t)1 << 37)) return fail(NR_ERR_LIMIT, "image batch too large");
    return NR_OK;
}

int train_common(nr_unet* u, const Inputs& in, const float* t, int N, int H, int W, float* loss) {
    if (!u || !in.b || !in.al || !in.nm || !in.dp || !t) return fail(NR_ERR_INVALID, "null argument");
    if (int e = check_dims(N, H, W)) return e;
    if (H % 16 || W % 16) return fail(NR_ERR_INVALID, "training crops must be multiples of 16");
    UCK(hipSetDevice(u->device));
    if (int e = ensure_shape(u, N, H, W)) return e;
    const size_t P = (size_t)N * H * W;
    hipLaunchKernelGGL(k_prep, dim3(blocks(P, 256)), dim3(256), 0, u->stream, in, N, H, W, u->X[0]);
    hipLaunchKernelGGL(k_target, dim3(blocks(P, 256)), dim3(256), 0, u->stream, t, N, H, W, u->tlog);
    forward(u, N, H, W, in, false, nullptr);
    const unsigned nb = blocks(P, 256);
    hipLaunchKernelGGL(k_loss, dim3(nb), dim3(256), 0, u->stream, u->y, u->tlog, (int)P, 1.0f / (3.0f * (float)P), u->gY,
                       u->partial);
    UCK(hipMemsetAsync(u->grad, 0, u->net.n_params * sizeof(float), u->stream));
    backward(u, N, H, W);
    UCK(hipGetLastError());
    std::vector<float> part(nb);
    UCK(hipMemcpyAsync(part.data(), u->partial, nb * sizeof(float), hipMemcpyDeviceToHost, u->stream));
    UCK(hipStreamSynchronize(u->stream));
    double s = 0.0;
    for (float x : part) s += x;
    if (loss) *loss = (float)(s / (3.0 * (double)P));
    return NR_OK;
}

}  // namespace

extern "C" {

size_t nr_unet_param_count(void) { return Net().n_params; }

nr_unet* nr_unet_create(int device, uint32_t seed) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) { fail(NR_ERR_HIP, "no HIP device visible"); return nullptr; }
    if (device < 0 || device >= n) { fail(NR_ERR_INVALID, "bad device"); return nullptr; }
    auto* u = new nr_unet();
    u->device = device;
    auto bad = [&](const char* m) { nr_error_slot() = m; nr_unet_destroy(u); return (nr_unet*)nullptr; };
    if (hipSetDevice(device) != hipSuccess) return bad("hipSetDevice");
    if (hipStreamCreateWithFlags(&u->stream, hipStreamNonBlocking) != hipSuccess) return bad("stream");
    if (hipEventCreate(&u->e0) != hipSuccess || hipEventCreate(&u->e1) != hipSuccess) return bad("events");
    const size_t np = u->net.n_params;
    if (hipMalloc((void**)&u->prm, np * 4) != hipSuccess || hipMalloc((void**)&u->grad, np * 4) != hipSuccess ||
        hipMalloc((void**)&u->m, np * 4) != hipSuccess || hipMalloc((void**)&u->v, np * 4) != hipSuccess)
        return bad("hipMalloc parameters");
    size_t wbytes = 0;
    for (Layer& l : u->net.L)
        wbytes += ((size_t)l.cout_pad * 9 * l.fc32 + (size_t)l.cin_pad * 9 * l.bc32) * 2 + 512;
    if (hipMalloc((void**)&u->wmem, wbytes) != hipSuccess) return bad("hipMalloc weights");
    size_t o = 0;
    for (Layer& l : u->net.L) {
        l.wf2 = u->wmem + o; o += ((size_t)l.cout_pad * 9 * l.fc32 + 127) & ~(size_t)127;
        l.wb2 = u->wmem + o; o += ((size_t)l.cin_pad * 9 * l.bc32 + 127) & ~(size_t)127;
    }
    // He-normal weights, zero biases; the output layer starts small (near-identity residual)
    std::vector<float> h(np, 0.0f);
    std::mt19937 rng(seed);
    for (size_t li = 0; li < u->net.L.size(); li++) {
        const Layer& l = u->net.L[li];
        const float sd = std::sqrt(2.0f / (9.0f * (float)l.cin)) * (li + 1 == u->net.L.size() ? 0.1f : 1.0f);
        std::normal_distribution<float> nd(0.0f, sd);
        for (size_t i = 0; i < (size_t)l.cout * 9 * l.cin; i++) h[l.w_off + i] = nd(rng);
    }
    if (nr_unet_set_params(u, h.data(), np) != NR_OK) { nr_unet_destroy(u); return nullptr; }
    return u;
}

void nr_unet_destroy(nr_unet* u) {
    if (!u) return;
    (void)hipSetDevice(u->device);
    if (u->stream) (void)hipStreamSynchronize(u->stream);
    for (void* p : {(void*)u->prm, (void*)u->grad, (void*)u->m, (void*)u->v, (void*)u->wmem, u->amem})
        if (p) (void)hipFree(p);
    if (u->e0) (void)hipEventDestroy(u->e0);
    if (u->e1) (void)hipEventDestroy(u->e1);
    if (u->stream) (void)hipStreamDestroy(u->stream);
    delete u;
}

int nr_unet_set_params(nr_unet* u, const float* p, size_t n) {
    if (!u || !p) return fail(NR_ERR_INVALID, "null argument");
    if (n != u->net.n_params) return fail(NR_ERR_INVALID, "parameter count mismatch");
    UCK(hipSetDevice(u->device));
    UCK(hipMemcpyAsync(u->prm, p, n * 4, hipMemcpyHostToDevice, u->stream));
    UCK(hipMemsetAsync(u->m, 0, n * 4, u->stream));
    UCK(hipMemsetAsync(u->v, 0, n * 4, u->stream));
    u->adam_t = 0;
    if (int e = pack_all(u)) return e;
    UCK(hipStreamSynchronize(u->stream));
    return NR_OK;
}

int nr_unet_get_params(nr_unet* u, float* p, size_t n) {
    if (!u || !p) return fail(NR_ERR_INVALID, "null argument");
    if (n != u->net.n_params) return fail(NR_ERR_INVALID, "parameter count mismatch");
    UCK(hipSetDevice(u->device));
    UCK(hipMemcpyAsync(p, u->prm, n * 4, hipMemcpyDeviceToHost, u->stream));
    UCK(hipStreamSynchronize(u->stream));
    return NR_OK;
}

int nr_unet_forward(nr_unet* u, const float* b, const float* al, const float* nm, const float* dp, int32_t N, int32_t H,
                    int32_t W, float* out) {
    if (!u || !b || !al || !nm || !dp || !out) return fail(NR_ERR_INVALID, "null argument");
    if (int e = check_dims(N, H, W)) return e;
    UCK(hipSetDevice(u->device));
    const int Hp = up16(H), Wp = up16(W);
    if (int e = ensure_shape(u, N, Hp, Wp)) return e;
    const size_t P = (size_t)N * Hp * Wp;
    UCK(hipEventRecord(u->e0, u->stream));
    forward(u, N, Hp, Wp, Inputs{b, al, nm, dp, H, W}, true, out);
    UCK(hipGetLastError());
    UCK(hipEventRecord(u->e1, u->stream));
    UCK(hipEventSynchronize(u->e1));
    UCK(hipEventElapsedTime(&u->last_ms, u->e0, u->e1));
    u->last_flops = net_flops(u, P);
    return NR_OK;
}

int nr_unet_gradients(nr_unet* u, const float* b, const float* al, const float* nm, const float* dp, const float* t,
                      int32_t N, int32_t H, int32_t W, float* grad, float* loss) {
    if (!grad) return fail(NR_ERR_INVALID, "null argument");
    if (int e = train_common(u, Inputs{b, al, nm, dp, H, W}, t, N, H, W, loss)) return e;
    UCK(hipMemcpy(grad, u->grad, u->net.n_params * 4, hipMemcpyDeviceToHost));
    return NR_OK;
}

int nr_unet_train_step(nr_unet* u, const float* b, const float* al, const float* nm, const float* dp, const float* t,
                       int32_t N, int32_t H, int32_t W, float lr, float* loss) {
    if (int e = train_common(u, Inputs{b, al, nm, dp, H, W}, t, N, H, W, loss)) return e;
    u->adam_t++;
    const float b1 = 0.9f, b2 = 0.999f;
    const float c1 = 1.0f - std::pow(b1, (float)u->adam_t), c2 = 1.0f - std::pow(b2, (float)u->adam_t);
    const size_t np = u->net.n_params;
    hipLaunchKernelGGL(k_adam, dim3(blocks(np, 256)), dim3(256), 0, u->stream, u->prm, u->grad, u->m, u->v, np, lr, b1,
                       b2, c1, c2, 1e-8f);
    if (int e = pack_all(u)) return e;
    UCK(hipStreamSynchronize(u->stream));
    return NR_OK;
}

int nr_denoise(nr_ctx* ctx, int view, nr_unet* u, float* out) {
    if (!ctx || !u || !out) return fail(NR_ERR_INVALID, "null argument");
    const float* b = nr_aov_device_ptr(ctx, view, NR_LAYER_BEAUTY);
    const float* al = nr_aov_device_ptr(ctx, view, NR_LAYER_DIFFUSE);
    const float* nm = nr_aov_device_ptr(ctx, view, NR_LAYER_WORLD_NORMAL);
    const float* dp = nr_aov_device_ptr(ctx, view, NR_LAYER_DEPTH);
    if (!b || !al || !nm || !dp) return fail(NR_ERR_STATE, "view has no beauty/diffuse/world-normal/depth layers");
    int32_t H = 0, W = 0;
    if (nr_view_size(ctx, view, &H, &W) != NR_OK) return NR_ERR_STATE;
    if (nr_sync(ctx) != NR_OK) return NR_ERR_HIP;
    return nr_unet_forward(u, b, al, nm, dp, 1, H, W, out);
}

int nr_unet_last_stats(nr_unet* u, float* ms, double* flops) {
    if (!u) return fail(NR_ERR_INVALID, "null argument");
    if (ms) *ms = u->last_ms;
    if (flops) *flops = u->last_flops;
    return NR_OK;
}

}  // extern "C"
