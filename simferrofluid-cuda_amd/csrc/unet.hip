// Kernels of the 3-D U-Net preconditioner on gfx950.
//
// Tensors are channels-last [batch][x][y][z][c], activations BF16 (stored as
// uint16), activation gradients FP32, master weights FP32 with BF16 copies
// [tap][cout][cin] for the convolutions.
//
// k_conv3r is an implicit GEMM: M = output voxels (z rows of 16), N = output
// channels (blocks of 16), K = 27 taps x input channels, on
// v_mfma_f32_16x16x32_bf16 with FP32 accumulation. Lane layouts (checked on
// the device): A[m][k] from lane m%16 + 16*(k/8), B[k][n] from lane
// n + 16*(k/8), D[m][n] in lane n + 16*(m/4), element m%4.
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

// The 3x3x3 convolution (zero padding), optionally reading its input
// nearest-up-sampled, with bias, ReLU, a skip tensor added after the ReLU and
// BF16 or FP32 output. Each input row's fragments are reused in registers by
// every output row it is a tap of. A wavefront owns one 16-channel block of
// outputs (nb) on PY z rows (y0 .. y0 + PY - 1, 16 voxels of z each) and
// marches along x over a run of output planes [xa, xe). Its weight fragments
// (9 x KC, all 27 taps of its output channels) stay in VGPRs for the march.
// At each step it loads the fragments of input plane gx (PY + 2 rows, each
// row shifted by dz = -1, 0, 1 in z, times the channel chunks: KC per row)
// straight from global memory (neighbouring wavefronts' rows meet in L1/L2),
// one plane ahead of their use, and each fragment feeds the MFMAs of up to 9
// output rows: planes gx + 1, gx, gx - 1 (dx = 0, 1, 2) times the rows it
// neighbours in y. Accumulators of three output planes rotate; plane gx - 1
// is complete after the step and goes through the epilogue.
template <int CIN>
constexpr int kchunks() {
    return CIN == 8 ? 1 : CIN == 16 ? 2 : CIN == 32 ? 3 : 6;
}
// K chunk kc of an input row, lane group g (8 of the MFMA's 32 K values): the
// z shift dz (0..2), the first channel c, and whether the slice exists.
template <int CIN>
__device__ __forceinline__ void kslice(int kc, int g, int &dz, int &c, bool &ok) {
    if constexpr (CIN == 8) {  // dz = g, all 8 channels; group 3 empty
        dz = g;
        c = 0;
        ok = g < 3;
    } else if constexpr (CIN == 16) {  // two z shifts per chunk, 8 channels per group
        dz = 2 * kc + (g >> 1);
        c = 8 * (g & 1);
        ok = dz < 3;
    } else if constexpr (CIN == 32) {
        dz = kc;
        c = 8 * g;
        ok = true;
    } else {
        dz = kc % 3;
        c = 32 * (kc / 3) + 8 * g;
        ok = true;
    }
}

// 16 zero bytes: lanes whose operand lies outside the volume (or outside
// the K chunk) load from here, so no loaded register is overwritten later.
__device__ __attribute__((aligned(16))) uint16_t g_zero8[8] = {0, 0, 0, 0, 0, 0, 0, 0};

template <int CIN, int COUT, int PY, int WPS>
__global__ __launch_bounds__(256, WPS) void k_conv3r(ConvArgs a, int run, int xruns) {
    constexpr int KC = kchunks<CIN>();
    constexpr int NB = COUT / 16;
    if (a.gate && *a.gate != 0.0) return;
    const uint16_t *__restrict__ in = a.in;
    const uint16_t *__restrict__ wts = a.w;
    const int X = a.X, Y = a.Y, Z = a.Z;
    const int lane = threadIdx.x & 63, g = lane >> 4, r16 = lane & 15;
    const int ypn = (Y + PY - 1) / PY, zbn = (Z + 15) / 16;
    unsigned w = blockIdx.x * 4 + (threadIdx.x >> 6);
    const int nb = (int)(w % NB);
    w /= NB;
    const int yp = (int)(w % (unsigned)ypn);
    w /= (unsigned)ypn;
    const int zb = (int)(w % (unsigned)zbn);
    w /= (unsigned)zbn;
    const int xr = (int)(w % (unsigned)xruns);
    const int b = (int)(w / (unsigned)xruns);
    if (b >= a.B) return;  // uniform over the wavefront
    const int y0 = yp * PY, z0 = zb * 16, xa = xr * run;
    if (xa >= X) return;
    const int xe = min(xa + run, X);
    // weights [dx][dy][kc]: B[k][n] of lane n + 16 (k / 8)
    v8bf wf[3][3][KC];
#pragma unroll
    for (int dx = 0; dx < 3; dx++)
#pragma unroll
        for (int dy = 0; dy < 3; dy++)
#pragma unroll
            for (int kc = 0; kc < KC; kc++) {
                int dz, c;
                bool ok;
                kslice<CIN>(kc, g, dz, c, ok);
                const uint16_t *p =
                    ok ? wts + ((size_t)((dx * 9 + dy * 3 + dz) * COUT + nb * 16 + r16) * CIN + c) : g_zero8;
                wf[dx][dy][kc] = __builtin_bit_cast(v8bf, *reinterpret_cast<const uint4 *>(p));
            }
    // this lane's input voxel (z) and channel offset per chunk: A[m][k] of
    // lane m + 16 (k / 8), m = z - z0 of the output row; the element offset
    // within an (x, y) row of the (possibly half-resolution) input
    const int up = a.up;
    const int Xs = up ? X >> 1 : X, Ys = up ? Y >> 1 : Y, Zs = up ? Z >> 1 : Z;
    int zoff[KC];
    bool zok[KC];
#pragma unroll
    for (int kc = 0; kc < KC; kc++) {
        int dz, c;
        bool ok;
        kslice<CIN>(kc, g, dz, c, ok);
        const int zi = z0 + r16 + dz - 1;
        zok[kc] = ok && zi >= 0 && zi < Z;
        const int cz = min(max(zi, 0), Z - 1);  // a valid address either way
        zoff[kc] = (up ? cz >> 1 : cz) * CIN + c;
    }
    v4f acc[3][PY];
#pragma unroll
    for (int s = 0; s < 3; s++)
#pragma unroll
        for (int oy = 0; oy < PY; oy++) acc[s][oy] = v4f{0.f, 0.f, 0.f, 0.f};
    const int steps = xe - xa + 2;  // input planes xa - 1 .. xe
    // The fragments of input plane gx: rows y0 - 1 .. y0 + PY, zero outside
    // the volume. Branch-free (the address is selected, never the loaded
    // value), so the next plane's loads stay in flight during this plane's
    // MFMAs.
    auto load_plane = [&](v8bf(&f)[PY + 2][KC], int gx) {
        const bool xok = gx >= 0 && gx < X;
        const int cx = min(max(gx, 0), X - 1);
        const size_t xrow = ((size_t)b * Xs + (up ? cx >> 1 : cx)) * Ys;
#pragma unroll
        for (int iy = 0; iy < PY + 2; iy++) {
            const int gy = y0 + iy - 1;
            const bool ok = xok && gy >= 0 && gy < Y;
            const int cy = min(max(gy, 0), Y - 1);
            const uint16_t *row = in + (xrow + (up ? cy >> 1 : cy)) * Zs * CIN;
#pragma unroll
            for (int kc = 0; kc < KC; kc++) {
                const uint16_t *p = ok && zok[kc] ? row + zoff[kc] : g_zero8;
                f[iy][kc] = __builtin_bit_cast(v8bf, *reinterpret_cast<const uint4 *>(p));
            }
        }
    };
    v8bf cur[PY + 2][KC], nxt[PY + 2][KC];
    load_plane(cur, xa - 1);
    const int ch = nb * 16 + r16;
    const float bias = a.bias ? a.bias[ch] : 0.f;
    const uint16_t *__restrict__ skip = a.skip;
    // step t reads input plane gx = xa - 1 + t; with U = t mod 3, output
    // plane gx + 1 - dx accumulates in slot (U - dx) mod 3
    for (int t0 = 0; t0 < steps; t0 += 3) {
#pragma unroll
        for (int U = 0; U < 3; U++) {
            const int t = t0 + U;
            if (t >= steps) break;  // uniform
            const int gx = xa - 1 + t;
            // the skip values of the plane this step completes, loaded before
            // the next plane's fragments so that waiting for them leaves those
            // in flight
            uint16_t skv[PY][4];
            if (skip && t >= 2) {
#pragma unroll
                for (int oy = 0; oy < PY; oy++)
#pragma unroll
                    for (int j = 0; j < 4; j++) {
                        const int gy = y0 + oy, gz = z0 + 4 * g + j;
                        const size_t vox = (((size_t)b * X + (gx - 1)) * Y + gy) * Z + gz;
                        skv[oy][j] = *(gy < Y && gz < Z ? skip + vox * COUT + ch : g_zero8);
                    }
            }
            load_plane(nxt, gx + 1);  // (one plane past the run at the end: unused)
#pragma unroll
            for (int iy = 0; iy < PY + 2; iy++)
#pragma unroll
                for (int dx = 0; dx < 3; dx++)
#pragma unroll
                    for (int dy = 0; dy < 3; dy++) {
                        const int oy = iy - dy;
                        if (oy < 0 || oy >= PY) continue;  // known after unrolling
#pragma unroll
                        for (int kc = 0; kc < KC; kc++)
                            acc[(U - dx + 3) % 3][oy] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(
                                cur[iy][kc], wf[dx][dy][kc], acc[(U - dx + 3) % 3][oy], 0, 0, 0);
                    }
            // planes before the run (gx + 1 - dx < xa) took contributions in
            // the first two steps: clear their slots
            if (t == 0) {
#pragma unroll
                for (int oy = 0; oy < PY; oy++) acc[1][oy] = acc[2][oy] = v4f{0.f, 0.f, 0.f, 0.f};
            } else if (t == 1) {
#pragma unroll
                for (int oy = 0; oy < PY; oy++) acc[2][oy] = v4f{0.f, 0.f, 0.f, 0.f};
            }
            // output plane gx - 1 has had its three input planes
            if (t >= 2) {
                const int S = (U + 1) % 3;
                const int ox = gx - 1;
#pragma unroll
                for (int oy = 0; oy < PY; oy++) {
                    const int gy = y0 + oy;
                    if (gy < Y) {
#pragma unroll
                        for (int j = 0; j < 4; j++) {
                            const int gz = z0 + 4 * g + j;
                            if (gz >= Z) continue;
                            const size_t vox = (((size_t)b * X + ox) * Y + gy) * Z + gz;
                            float v = acc[S][oy][j] + bias;
                            if (a.relu) v = v > 0.f ? v : 0.f;
                            if (a.relu_copy) a.relu_copy[vox * COUT + ch] = f2bf(v);
                            if (skip) v += bf2f(skv[oy][j]);
                            if (a.out32)
                                a.out32[vox * COUT + ch] = v;
                            else
                                a.out[vox * COUT + ch] = f2bf(v);
                        }
                    }
                    acc[S][oy] = v4f{0.f, 0.f, 0.f, 0.f};
                }
            }
#pragma unroll
            for (int iy = 0; iy < PY + 2; iy++)
#pragma unroll
                for (int kc = 0; kc < KC; kc++) cur[iy][kc] = nxt[iy][kc];
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
    // rows per wavefront and the wavefronts per SIMD the launch bound asks
    // for, fastest of PY = 1..4 measured on MI355X for each shape at its
    // U-Net level (128^3, 64^3, 32^3); more rows reuse each loaded fragment
    // more often but hold more VGPRs (the weight fragments alone take 36 to
    // 216)
    constexpr int PY = CIN == 8 ? 4 : CIN == 16 ? 2 : 1;
    constexpr int WPS = CIN <= 16 ? 2 : 1;
    // x runs: about two wavefronts per SIMD of the 256 CUs
    const long long base = (long long)a.B * ((a.Z + 15) / 16) * ((a.Y + PY - 1) / PY) * (COUT / 16);
    const long long target = 2048;
    int xruns = (int)std::min<long long>(a.X, std::max<long long>(1, (target + base - 1) / base));
    const int run = (a.X + xruns - 1) / xruns;
    xruns = (a.X + run - 1) / run;
    const long long waves = base * xruns;
    hipLaunchKernelGGL((k_conv3r<CIN, COUT, PY, WPS>), dim3((unsigned)((waves + 3) / 4)), dim3(256), 0, st, a,
                       run, xruns);
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
