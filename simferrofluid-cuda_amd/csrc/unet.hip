// Kernels of the 3-D U-Net preconditioner on gfx950.
//
// Tensors are channels-last [batch][x][y][z][c], activations BF16 (stored as
// uint16), activation gradients FP32, master weights FP32 with BF16 copies
// [tap][cout][cin] for the convolutions.
//
// k_conv3 is an implicit GEMM: M = output voxels of a (TX x 4 x 16) tile,
// N = output channels, K = 27 taps x input channels, on
// v_mfma_f32_16x16x32_bf16 with FP32 accumulation. The input tile with its
// one-voxel halo is staged in LDS; each of the 4 wavefronts owns TX row blocks
// of 16 voxels and all N. Lane layouts (checked on the device): A[m][k] from
// lane m%16 + 16*(k/8), B[k][n] from lane n + 16*(k/8), D[m][n] in lane
// n + 16*(m/4), element m%4.
#include "device_util.hpp"
#include "unet.hpp"

namespace sfp {
namespace {

typedef __bf16 v8bf __attribute__((ext_vector_type(8)));
typedef float v4f __attribute__((ext_vector_type(4)));

__device__ __forceinline__ float bf2f(uint16_t h) { return __uint_as_float(((unsigned)h) << 16); }
__device__ __forceinline__ uint16_t f2bf(float f) {
    // round to nearest even; NaN stays NaN
    const unsigned u = __float_as_uint(f);
    if ((u & 0x7fffffffu) > 0x7f800000u) return (uint16_t)((u >> 16) | 0x40);
    return (uint16_t)((u + 0x7fffu + ((u >> 16) & 1u)) >> 16);
}

struct ConvArgs {
    const uint16_t *in;   // [B][X][Y][Z][CIN]
    const uint16_t *w;    // [tap][COUT][CIN]
    const float *bias;    // [COUT] or null
    const uint16_t *skip; // [B][X][Y][Z][COUT] or null: added after the ReLU
    uint16_t *out;        // BF16 output or null
    float *out32;         // FP32 output or null
    uint16_t *relu_copy;  // relu(acc + bias) before the skip, or null
    int B, X, Y, Z;
    int relu;
    int up;  // 1: `in` is [B][X/2][Y/2][Z/2][CIN], read nearest-up-sampled
    const double *gate;  // nonzero *gate: return at once (a stopped batch of FCG iterations)
};

template <int CIN>
__device__ __forceinline__ void chunk_slice(int chunk, int g, int &tap, int &c0) {
    if constexpr (CIN == 8) {
        tap = chunk * 4 + g;
        c0 = 0;
    } else if constexpr (CIN == 16) {
        tap = chunk * 2 + (g >> 1);
        c0 = 8 * (g & 1);
    } else if constexpr (CIN == 32) {
        tap = chunk;
        c0 = 8 * g;
    } else {
        tap = chunk >> 1;
        c0 = 32 * (chunk & 1) + 8 * g;
    }
}

template <int CIN>
constexpr int num_chunks() {
    return CIN == 8 ? 7 : CIN == 16 ? 14 : CIN == 32 ? 27 : 54;
}

template <int CIN, int COUT, int TX>
__global__ __launch_bounds__(256) void k_conv3(ConvArgs a) {
    constexpr int HX = TX + 2, HY = 6, HZ = 18;
    constexpr int NB = COUT / 16;
    constexpr int RB = TX;  // 16-row blocks per wavefront
    __shared__ __attribute__((aligned(16))) uint16_t tile[HX * HY * HZ * CIN];
    if (a.gate && *a.gate != 0.0) return;
    const int tid = threadIdx.x;
    const int tx_n = (a.X + TX - 1) / TX, ty_n = (a.Y + 3) / 4, tz_n = (a.Z + 15) / 16;
    unsigned t = blockIdx.x;
    const int bz = t % tz_n;
    t /= tz_n;
    const int by = t % ty_n;
    t /= ty_n;
    const int bx = t % tx_n;
    const int bb = t / tx_n;
    const int x0 = bx * TX, y0 = by * 4, z0 = bz * 16;
    // stage the halo tile
    constexpr int PARTS = CIN / 8;
    constexpr int TOTAL = HX * HY * HZ * PARTS;
    for (int idx = tid; idx < TOTAL; idx += 256) {
        const int vox = idx / PARTS, part = idx - vox * PARTS;
        const int hz = vox % HZ, hy = (vox / HZ) % HY, hx = vox / (HZ * HY);
        const int gx = x0 + hx - 1, gy = y0 + hy - 1, gz = z0 + hz - 1;
        uint4 v = make_uint4(0, 0, 0, 0);
        if (gx >= 0 && gy >= 0 && gz >= 0 && gx < a.X && gy < a.Y && gz < a.Z) {
            const size_t gi =
                a.up ? ((((size_t)bb * (a.X >> 1) + (gx >> 1)) * (a.Y >> 1) + (gy >> 1)) * (a.Z >> 1) + (gz >> 1)) *
                               CIN + part * 8
                     : ((((size_t)bb * a.X + gx) * a.Y + gy) * a.Z + gz) * CIN + part * 8;
            v = *reinterpret_cast<const uint4 *>(a.in + gi);
        }
        *reinterpret_cast<uint4 *>(tile + vox * CIN + part * 8) = v;
    }
    __syncthreads();
    const int lane = tid & 63, wave = tid >> 6;
    const int g = lane >> 4, r16 = lane & 15;
    v4f acc[RB][NB];
#pragma unroll
    for (int rb = 0; rb < RB; rb++)
#pragma unroll
        for (int nb = 0; nb < NB; nb++) acc[rb][nb] = v4f{0.f, 0.f, 0.f, 0.f};
    for (int chunk = 0; chunk < num_chunks<CIN>(); chunk++) {
        int tap, c0;
        chunk_slice<CIN>(chunk, g, tap, c0);
        const bool valid = tap < 27;
        const int dx = tap / 9, dy = (tap / 3) % 3, dz = tap % 3;
        v8bf bf[NB];
#pragma unroll
        for (int nb = 0; nb < NB; nb++) {
            uint4 u = make_uint4(0, 0, 0, 0);
            if (valid)
                u = *reinterpret_cast<const uint4 *>(a.w + ((size_t)(tap * COUT + nb * 16 + r16) * CIN + c0));
            bf[nb] = __builtin_bit_cast(v8bf, u);
        }
#pragma unroll
        for (int rb = 0; rb < RB; rb++) {
            const int row = wave * (16 * RB) + rb * 16 + r16;
            const int tz = row & 15, ty = (row >> 4) & 3, tx = row >> 6;
            uint4 u = make_uint4(0, 0, 0, 0);
            if (valid)
                u = *reinterpret_cast<const uint4 *>(tile + (((tx + dx) * HY + (ty + dy)) * HZ + (tz + dz)) * CIN + c0);
            const v8bf af = __builtin_bit_cast(v8bf, u);
#pragma unroll
            for (int nb = 0; nb < NB; nb++)
                acc[rb][nb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(af, bf[nb], acc[rb][nb], 0, 0, 0);
        }
    }
    // epilogue
#pragma unroll
    for (int rb = 0; rb < RB; rb++)
#pragma unroll
        for (int j = 0; j < 4; j++) {
            const int row = wave * (16 * RB) + rb * 16 + 4 * g + j;
            const int tz = row & 15, ty = (row >> 4) & 3, tx = row >> 6;
            const int gx = x0 + tx, gy = y0 + ty, gz = z0 + tz;
            if (gx >= a.X || gy >= a.Y || gz >= a.Z) continue;
            const size_t vox = (((size_t)bb * a.X + gx) * a.Y + gy) * a.Z + gz;
#pragma unroll
            for (int nb = 0; nb < NB; nb++) {
                const int ch = nb * 16 + r16;
                float v = acc[rb][nb][j];
                if (a.bias) v += a.bias[ch];
                if (a.relu) v = v > 0.f ? v : 0.f;
                if (a.relu_copy) a.relu_copy[vox * COUT + ch] = f2bf(v);
                if (a.skip) v += bf2f(a.skip[vox * COUT + ch]);
                if (a.out32)
                    a.out32[vox * COUT + ch] = v;
                else
                    a.out[vox * COUT + ch] = f2bf(v);
            }
        }
}

// Weight gradient partials as an implicit GEMM on MFMA: for tap blockIdx.y
// and voxel chunk blockIdx.x, part[chunk][tap][cout][cin] =
// sum_v g[v][cout] x[v + tap][cin] (M = output channels, N = input channels,
// K = the chunk's voxels; BF16 products, exact in FP32, FP32 accumulation);
// the centre-tap blocks also store bias partials bpart[chunk][cout] =
// sum_v g[v][cout]. Every K step stages 128 voxels of g and of the shifted
// x in LDS, transposed (channel-major, 8 consecutive voxels per 16-byte
// fragment); wavefront w takes the step's voxels 32 w .. 32 w + 31 for every
// output tile, and the four partial sets are added in a fixed order at the
// end, so the result does not depend on timing. Input channels below 16 are
// padded with zeros.
template <int CIN, int COUT>
__global__ __launch_bounds__(256) void k_conv_dw(const uint16_t *__restrict__ x,
                                                 const uint16_t *__restrict__ g, int B, int X,
                                                 int Y, int Z, int chunk_len,
                                                 float *__restrict__ part,
                                                 float *__restrict__ bpart) {
    constexpr int KV = 128;                       // voxels per K step (4 x 32)
    constexpr int CINP = CIN < 16 ? 16 : CIN;     // padded to the MFMA's 16 columns
    constexpr int MB = COUT / 16, NB = CINP / 16, T = MB * NB;
    constexpr int PAIRS = CIN * COUT;
    __shared__ __attribute__((aligned(16))) uint16_t sg[COUT][KV];
    __shared__ __attribute__((aligned(16))) uint16_t sx[CINP][KV];
    __shared__ float red[4][T][4][64];            // each wavefront's partial tiles
    const int tap = blockIdx.y;
    const int dx = tap / 9 - 1, dy = (tap / 3) % 3 - 1, dz = tap % 3 - 1;
    const size_t nv = (size_t)B * X * Y * Z;
    const size_t v0 = (size_t)blockIdx.x * chunk_len;
    const size_t v1 = v0 + chunk_len < nv ? v0 + chunk_len : nv;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int gq = lane >> 4, r16 = lane & 15;
    v4f acc[T];
#pragma unroll
    for (int t = 0; t < T; t++) acc[t] = v4f{0.f, 0.f, 0.f, 0.f};
    float bacc = 0.f;
    if (CINP > CIN) {  // the padding columns stay zero
        for (int e = tid; e < (CINP - CIN) * KV; e += 256) sx[CIN + e / KV][e % KV] = 0;
    }
    for (size_t vb = v0; vb < v1; vb += KV) {
        __syncthreads();  // the previous step's fragments are read
        // g: 8 channels (16 bytes) of one voxel per thread and round
        for (int e = tid; e < KV * (COUT / 8); e += 256) {
            const int vi = e / (COUT / 8), q = e - vi * (COUT / 8);
            const size_t v = vb + vi;
            uint4 u = make_uint4(0, 0, 0, 0);
            if (v < v1) u = *reinterpret_cast<const uint4 *>(g + v * COUT + 8 * q);
            const uint16_t *h = reinterpret_cast<const uint16_t *>(&u);
#pragma unroll
            for (int c = 0; c < 8; c++) sg[8 * q + c][vi] = h[c];
        }
        // x at the tapped voxel (zero outside the volume)
        for (int e = tid; e < KV * (CIN / 8); e += 256) {
            const int vi = e / (CIN / 8), q = e - vi * (CIN / 8);
            const size_t v = vb + vi;
            uint4 u = make_uint4(0, 0, 0, 0);
            if (v < v1) {
                size_t r = v;
                const int z = (int)(r % Z);
                r /= Z;
                const int y = (int)(r % Y);
                r /= Y;
                const int xx = (int)(r % X);
                const size_t bb = r / X;
                const int jx = xx + dx, jy = y + dy, jz = z + dz;
                if (jx >= 0 && jy >= 0 && jz >= 0 && jx < X && jy < Y && jz < Z)
                    u = *reinterpret_cast<const uint4 *>(x + (((bb * X + jx) * Y + jy) * Z + jz) * CIN + 8 * q);
            }
            const uint16_t *h = reinterpret_cast<const uint16_t *>(&u);
#pragma unroll
            for (int c = 0; c < 8; c++) sx[8 * q + c][vi] = h[c];
        }
        __syncthreads();
        // this wavefront's 32 voxels of the step: k = 32 wave + 8 gq .. + 7
        const int k0 = 32 * wave + 8 * gq;
        v8bf af[MB], bf[NB];
#pragma unroll
        for (int mb = 0; mb < MB; mb++)
            af[mb] = __builtin_bit_cast(v8bf, *reinterpret_cast<const uint4 *>(&sg[16 * mb + r16][k0]));
#pragma unroll
        for (int nb = 0; nb < NB; nb++)
            bf[nb] = __builtin_bit_cast(v8bf, *reinterpret_cast<const uint4 *>(&sx[16 * nb + r16][k0]));
#pragma unroll
        for (int mb = 0; mb < MB; mb++)
#pragma unroll
            for (int nb = 0; nb < NB; nb++)
                acc[mb * NB + nb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(af[mb], bf[nb], acc[mb * NB + nb], 0, 0, 0);
        if (tap == 13 && tid < COUT) {
            float sb = 0.f;
            for (int i = 0; i < KV; i++) sb += bf2f(sg[tid][i]);
            bacc += sb;
        }
    }
    // add the four wavefronts' tiles in a fixed order
#pragma unroll
    for (int t = 0; t < T; t++)
#pragma unroll
        for (int j = 0; j < 4; j++) red[wave][t][j][lane] = acc[t][j];
    __syncthreads();
    for (int e = tid; e < T * 4 * 64; e += 256) {
        const int t = e / 256, j = (e / 64) & 3, l = e & 63;
        const float sum = ((red[0][t][j][l] + red[1][t][j][l]) + red[2][t][j][l]) + red[3][t][j][l];
        // C[row][col] of tile t: row = 4 (l >> 4) + j (output channel), col = l & 15 (input channel)
        const int co = 16 * (t / NB) + 4 * (l >> 4) + j, ci = 16 * (t % NB) + (l & 15);
        if (ci < CIN) part[((size_t)blockIdx.x * 27 + tap) * PAIRS + co * CIN + ci] = sum;
    }
    if (tap == 13 && tid < COUT) bpart[(size_t)blockIdx.x * COUT + tid] = bacc;
}

// out[i] += sum over chunks of part[chunk][i] (fixed order).
__global__ __launch_bounds__(256) void k_reduce_chunks(const float *__restrict__ part, int chunks,
                                                       size_t n, float *__restrict__ out) {
    const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    float s = 0.f;
    for (int c = 0; c < chunks; c++) s += part[(size_t)c * n + i];
    out[i] += s;
}

__global__ __launch_bounds__(256) void k_pool2(const uint16_t *__restrict__ in, int B, int X, int Y,
                                               int Z, int C, uint16_t *__restrict__ out,
                                               const double *gate) {
    if (gate && *gate != 0.0) return;
    const int Xo = X / 2, Yo = Y / 2, Zo = Z / 2;
    const size_t n = (size_t)B * Xo * Yo * Zo * C;
    const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    size_t r = i;
    const int c = (int)(r % C);
    r /= C;
    const int z = (int)(r % Zo);
    r /= Zo;
    const int y = (int)(r % Yo);
    r /= Yo;
    const int x = (int)(r % Xo);
    const size_t b = r / Xo;
    float s = 0.f;
    for (int d = 0; d < 8; d++) {
        const size_t v = ((b * X + 2 * x + (d >> 2)) * Y + 2 * y + ((d >> 1) & 1)) * Z + 2 * z + (d & 1);
        s += bf2f(in[v * C + c]);
    }
    out[i] = f2bf(s * 0.125f);
}

// gin[fine] = g[coarse] / 8, optionally plus add[fine]
__global__ __launch_bounds__(256) void k_pool2_bwd(const float *__restrict__ g, const float *__restrict__ add,
                                                   int B, int X, int Y, int Z, int C,
                                                   float *__restrict__ gin) {
    const size_t n = (size_t)B * X * Y * Z * C;
    const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    size_t r = i;
    const int c = (int)(r % C);
    r /= C;
    const int z = (int)(r % Z);
    r /= Z;
    const int y = (int)(r % Y);
    r /= Y;
    const int x = (int)(r % X);
    const size_t b = r / X;
    const size_t vc = ((b * (X / 2) + x / 2) * (Y / 2) + y / 2) * (Z / 2) + z / 2;
    float v = 0.125f * g[vc * C + c];
    if (add) v += add[i];
    gin[i] = v;
}

__global__ __launch_bounds__(256) void k_up2(const uint16_t *__restrict__ in, int B, int Xc, int Yc,
                                             int Zc, int C, uint16_t *__restrict__ out) {
    const int X = 2 * Xc, Y = 2 * Yc, Z = 2 * Zc;
    const size_t n = (size_t)B * X * Y * Z * C;
    const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    size_t r = i;
    const int c = (int)(r % C);
    r /= C;
    const int z = (int)(r % Z);
    r /= Z;
    const int y = (int)(r % Y);
    r /= Y;
    const int x = (int)(r % X);
    const size_t b = r / X;
    out[i] = in[(((b * Xc + x / 2) * Yc + y / 2) * Zc + z / 2) * C + c];
}

// gcoarse = sum of the 8 children of g (fine)
__global__ __launch_bounds__(256) void k_up2_bwd(const float *__restrict__ g, int B, int Xc, int Yc,
                                                 int Zc, int C, float *__restrict__ gc) {
    const size_t n = (size_t)B * Xc * Yc * Zc * C;
    const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    size_t r = i;
    const int c = (int)(r % C);
    r /= C;
    const int z = (int)(r % Zc);
    r /= Zc;
    const int y = (int)(r % Yc);
    r /= Yc;
    const int x = (int)(r % Xc);
    const size_t b = r / Xc;
    const int X = 2 * Xc, Y = 2 * Yc, Z = 2 * Zc;
    float s = 0.f;
    for (int d = 0; d < 8; d++) {
        const size_t v = ((b * X + 2 * x + (d >> 2)) * Y + 2 * y + ((d >> 1) & 1)) * Z + 2 * z + (d & 1);
        s += g[v * C + c];
    }
    gc[i] = s;
}

// g16 = bf16(g * (act > 0))
__global__ __launch_bounds__(256) void k_relu_mask(const float *__restrict__ g,
                                                   const uint16_t *__restrict__ act, size_t n,
                                                   uint16_t *__restrict__ g16) {
    const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    g16[i] = bf2f(act[i]) > 0.f ? f2bf(g[i]) : (uint16_t)0;
}

// z[v] = bo + sum_c wo[c] d0[v][c]
__global__ __launch_bounds__(256) void k_head_fwd(const uint16_t *__restrict__ d0, const float *__restrict__ wb,
                                                  size_t nv, float *__restrict__ z, const double *gate) {
    if (gate && *gate != 0.0) return;
    const size_t v = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (v >= nv) return;
    float s = wb[16];
    const uint16_t *p = d0 + v * 16;
#pragma unroll
    for (int c = 0; c < 16; c++) s += wb[c] * bf2f(p[c]);
    z[v] = s;
}

// gd0[v][c] = dz[v] wo[c]; per-block partials of sum dz d0[v][c] (16) and sum dz
__global__ __launch_bounds__(256) void k_head_bwd(const uint16_t *__restrict__ d0, const float *__restrict__ wb,
                                                  const float *__restrict__ dz, size_t nv,
                                                  float *__restrict__ gd0, double *__restrict__ part) {
    const size_t v = (size_t)blockIdx.x * 256 + threadIdx.x;
    double acc[17];
#pragma unroll
    for (int c = 0; c < 17; c++) acc[c] = 0.0;
    if (v < nv) {
        const float d = dz[v];
#pragma unroll
        for (int c = 0; c < 16; c++) {
            gd0[v * 16 + c] = d * wb[c];
            acc[c] = (double)d * bf2f(d0[v * 16 + c]);
        }
        acc[16] = d;
    }
    block_partials<17>(acc, part);
}

// per-block partials of (mask (z - t)^2, mask)
__global__ __launch_bounds__(256) void k_mse(const float *__restrict__ z, const float *__restrict__ t,
                                             const float *__restrict__ m, size_t nv,
                                             double *__restrict__ part) {
    const size_t v = (size_t)blockIdx.x * 256 + threadIdx.x;
    double acc[2] = {0.0, 0.0};
    if (v < nv) {
        const double mm = m ? m[v] : 1.0;
        const double d = (double)z[v] - t[v];
        acc[0] = mm * d * d;
        acc[1] = mm;
    }
    block_partials<2>(acc, part);
}

__global__ __launch_bounds__(256) void k_sum_partials(const double *__restrict__ part, unsigned nb,
                                                      int n, double *__restrict__ out) {
    // n <= 17 columns, each of nb partials
    for (int q = 0; q < n; q++) {
        double v[1] = {0.0};
        for (unsigned b = threadIdx.x; b < nb; b += 256) v[0] += part[(size_t)q * nb + b];
        block_sum<1>(v);
        if (threadIdx.x == 0) out[q] = v[0];
    }
}

__global__ __launch_bounds__(256) void k_cast_sums(const double *__restrict__ sums, int n,
                                                   float *__restrict__ out) {
    if ((int)threadIdx.x < n) out[threadIdx.x] = (float)sums[threadIdx.x];
}

__global__ __launch_bounds__(256) void k_mse_grad(const float *__restrict__ z, const float *__restrict__ t,
                                                  const float *__restrict__ m, size_t nv,
                                                  const double *__restrict__ sums, float *__restrict__ dz) {
    const size_t v = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (v >= nv) return;
    const double cnt = sums[1];
    const float mm = m ? m[v] : 1.f;
    dz[v] = cnt > 0 ? (float)(mm * 2.0 * ((double)z[v] - t[v]) / cnt) : 0.f;
}

__global__ __launch_bounds__(256) void k_adam(float *__restrict__ w, const float *__restrict__ g,
                                              float *__restrict__ m, float *__restrict__ v, size_t n,
                                              float lr, float b1, float b2, float eps, float c1, float c2) {
    const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.f - b1) * gi;
    const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    w[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

// BF16 forward weights [tap][cout][cin] (taps padded with zeros) and the
// flipped, transposed copy for the input gradient: wt[tap'][ci][co] = w[26-tap'][co][ci].
__global__ __launch_bounds__(256) void k_cast_weights(const float *__restrict__ w, int cin, int cout,
                                                      int taps_fwd, int taps_bwd,
                                                      uint16_t *__restrict__ wf,
                                                      uint16_t *__restrict__ wt) {
    const size_t n = 27 * (size_t)cin * cout;
    const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (i < (size_t)taps_fwd * cin * cout) wf[i] = i < n ? f2bf(w[i]) : (uint16_t)0;
    if (wt && i < (size_t)taps_bwd * cin * cout) {
        uint16_t val = 0;
        if (i < n) {
            const int tp = (int)(i / ((size_t)cin * cout));
            const int rem = (int)(i - (size_t)tp * cin * cout);
            const int ci = rem / cout, co = rem - ci * cout;  // wt index [tp][ci][co]
            val = f2bf(w[((size_t)(26 - tp) * cout + co) * cin + ci]);
        }
        wt[i] = val;
    }
}

// FP32 [nv][4] -> BF16 [nv][8] (channels 4..7 zero)
__global__ __launch_bounds__(256) void k_pack_input(const float *__restrict__ in, size_t nv,
                                                    uint16_t *__restrict__ out) {
    const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= nv * 8) return;
    const size_t v = i >> 3;
    const int c = (int)(i & 7);
    out[i] = c < 4 ? f2bf(in[v * 4 + c]) : (uint16_t)0;
}

unsigned blocks_for(size_t n) { return (unsigned)((n + 255) / 256); }

template <int CIN, int COUT>
void conv_dispatch(hipStream_t st, const ConvArgs &a) {
    constexpr int TX = CIN >= 64 ? 2 : 4;
    const unsigned tiles = (unsigned)((size_t)a.B * ((a.X + TX - 1) / TX) * ((a.Y + 3) / 4) * ((a.Z + 15) / 16));
    hipLaunchKernelGGL((k_conv3<CIN, COUT, TX>), dim3(tiles), dim3(256), 0, st, a);
    SFP_LAUNCH_CHECK();
}

template <int CIN, int COUT>
void dw_dispatch(hipStream_t st, const uint16_t *x, const uint16_t *g, int B, int X, int Y, int Z,
                 float *part, float *bpart, int chunks, int chunk_len) {
    hipLaunchKernelGGL((k_conv_dw<CIN, COUT>), dim3(chunks, 27), dim3(256), 0, st, x, g, B, X, Y, Z,
                       chunk_len, part, bpart);
    SFP_LAUNCH_CHECK();
}

}  // namespace

void unet_conv(hipStream_t st, int cin, int cout, const uint16_t *in, const uint16_t *w,
               const float *bias, const uint16_t *skip, uint16_t *out, float *out32,
               uint16_t *relu_copy, int B, int X, int Y, int Z, bool relu, bool up_input,
               const double *gate) {
    if (up_input && (X % 2 || Y % 2 || Z % 2)) throw ArgError("up-sampled input needs even extents");
    ConvArgs a{in, w, bias, skip, out, out32, relu_copy, B, X, Y, Z, relu ? 1 : 0, up_input ? 1 : 0, gate};
#define SFP_CONV(CI, CO) \
    if (cin == CI && cout == CO) return conv_dispatch<CI, CO>(st, a);
    SFP_CONV(8, 16)
    SFP_CONV(16, 16)
    SFP_CONV(16, 32)
    SFP_CONV(32, 16)
    SFP_CONV(32, 32)
    SFP_CONV(32, 64)
    SFP_CONV(64, 32)
    SFP_CONV(64, 64)
#undef SFP_CONV
    throw ArgError("unsupported convolution shape");
}

void unet_conv_dw(hipStream_t st, int cin, int cout, const uint16_t *x, const uint16_t *g, int B,
                  int X, int Y, int Z, float *part, float *dw, float *db) {
    const size_t nv = (size_t)B * X * Y * Z;
    const int chunks = (int)std::min<size_t>(kDwChunks, nv);
    const int chunk_len = (int)((nv + chunks - 1) / chunks);
    const size_t n = 27 * (size_t)cin * cout;
    float *bpart = part + (size_t)chunks * n;
#define SFP_DW(CI, CO) \
    if (cin == CI && cout == CO) dw_dispatch<CI, CO>(st, x, g, B, X, Y, Z, part, bpart, chunks, chunk_len); else
    SFP_DW(8, 16)
    SFP_DW(16, 16)
    SFP_DW(16, 32)
    SFP_DW(32, 16)
    SFP_DW(32, 32)
    SFP_DW(32, 64)
    SFP_DW(64, 32)
    SFP_DW(64, 64)
    throw ArgError("unsupported convolution shape");
#undef SFP_DW
    hipLaunchKernelGGL(k_reduce_chunks, dim3(blocks_for(n)), dim3(256), 0, st, part, chunks, n, dw);
    SFP_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_reduce_chunks, dim3(1), dim3(256), 0, st, bpart, chunks, (size_t)cout, db);
    SFP_LAUNCH_CHECK();
}

void unet_pool2(hipStream_t st, const uint16_t *in, int B, int X, int Y, int Z, int C, uint16_t *out,
                const double *gate) {
    const size_t n = (size_t)B * (X / 2) * (Y / 2) * (Z / 2) * C;
    hipLaunchKernelGGL(k_pool2, dim3(blocks_for(n)), dim3(256), 0, st, in, B, X, Y, Z, C, out, gate);
    SFP_LAUNCH_CHECK();
}

void unet_pool2_bwd(hipStream_t st, const float *g, const float *add, int B, int X, int Y, int Z,
                    int C, float *gin) {
    const size_t n = (size_t)B * X * Y * Z * C;
    hipLaunchKernelGGL(k_pool2_bwd, dim3(blocks_for(n)), dim3(256), 0, st, g, add, B, X, Y, Z, C, gin);
    SFP_LAUNCH_CHECK();
}

void unet_up2(hipStream_t st, const uint16_t *in, int B, int Xc, int Yc, int Zc, int C, uint16_t *out) {
    const size_t n = (size_t)B * 8 * Xc * Yc * Zc * C;
    hipLaunchKernelGGL(k_up2, dim3(blocks_for(n)), dim3(256), 0, st, in, B, Xc, Yc, Zc, C, out);
    SFP_LAUNCH_CHECK();
}

void unet_up2_bwd(hipStream_t st, const float *g, int B, int Xc, int Yc, int Zc, int C, float *gc) {
    const size_t n = (size_t)B * Xc * Yc * Zc * C;
    hipLaunchKernelGGL(k_up2_bwd, dim3(blocks_for(n)), dim3(256), 0, st, g, B, Xc, Yc, Zc, C, gc);
    SFP_LAUNCH_CHECK();
}

void unet_relu_mask(hipStream_t st, const float *g, const uint16_t *act, size_t n, uint16_t *g16) {
    hipLaunchKernelGGL(k_relu_mask, dim3(blocks_for(n)), dim3(256), 0, st, g, act, n, g16);
    SFP_LAUNCH_CHECK();
}

void unet_head_fwd(hipStream_t st, const uint16_t *d0, const float *wb, size_t nv, float *z,
                   const double *gate) {
    hipLaunchKernelGGL(k_head_fwd, dim3(blocks_for(nv)), dim3(256), 0, st, d0, wb, nv, z, gate);
    SFP_LAUNCH_CHECK();
}

void unet_head_bwd(hipStream_t st, const uint16_t *d0, const float *wb, const float *dz, size_t nv,
                   float *gd0, double *part, double *sums, float *gwb) {
    const unsigned nb = blocks_for(nv);
    hipLaunchKernelGGL(k_head_bwd, dim3(nb), dim3(256), 0, st, d0, wb, dz, nv, gd0, part);
    SFP_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_sum_partials, dim3(1), dim3(256), 0, st, part, nb, 17, sums);
    SFP_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_cast_sums, dim3(1), dim3(256), 0, st, sums, 17, gwb);
    SFP_LAUNCH_CHECK();
}

void unet_mse(hipStream_t st, const float *z, const float *t, const float *m, size_t nv,
              double *part, double *sums, float *dz) {
    const unsigned nb = blocks_for(nv);
    hipLaunchKernelGGL(k_mse, dim3(nb), dim3(256), 0, st, z, t, m, nv, part);
    SFP_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_sum_partials, dim3(1), dim3(256), 0, st, part, nb, 2, sums);
    SFP_LAUNCH_CHECK();
    if (dz) {
        hipLaunchKernelGGL(k_mse_grad, dim3(nb), dim3(256), 0, st, z, t, m, nv, sums, dz);
        SFP_LAUNCH_CHECK();
    }
}

void unet_adam(hipStream_t st, float *w, const float *g, float *m, float *v, size_t n, float lr,
               int step) {
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    const float c1 = 1.f - powf(b1, (float)step), c2 = 1.f - powf(b2, (float)step);
    hipLaunchKernelGGL(k_adam, dim3(blocks_for(n)), dim3(256), 0, st, w, g, m, v, n, lr, b1, b2, eps, c1, c2);
    SFP_LAUNCH_CHECK();
}

void unet_cast_weights(hipStream_t st, const float *w, int cin, int cout, uint16_t *wf, uint16_t *wt) {
    const int tf = conv_taps(cin), tb = conv_taps(cout);
    const size_t n = (size_t)std::max(tf, tb) * cin * cout;
    hipLaunchKernelGGL(k_cast_weights, dim3(blocks_for(n)), dim3(256), 0, st, w, cin, cout, tf, tb, wf, wt);
    SFP_LAUNCH_CHECK();
}

void unet_pack_input(hipStream_t st, const float *in, size_t nv, uint16_t *out) {
    hipLaunchKernelGGL(k_pack_input, dim3(blocks_for(nv * 8)), dim3(256), 0, st, in, nv, out);
    SFP_LAUNCH_CHECK();
}

}  // namespace sfp
