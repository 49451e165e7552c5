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

#include <type_traits>

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
    unsigned in_bytes;   // size of `in` (at most kInLimit)
    unsigned out_elems;  // B X Y Z COUT: the elements of out, out32, relu_copy and skip
};

// Raw buffer resources over the tensors: a load at an offset past the
// buffer's size reads 0 and a store there is dropped, so the padding and the
// partial tiles need no branches or address selects. An input offset is the
// sum of a row part and a (z, channel) part, either of which is kInOutside
// when outside the volume: with the input at most kInLimit bytes, the sum is
// then past it too. Output offsets outside the volume are kOutOutside (the
// outputs are at most 2 GiB).
constexpr unsigned kInLimit = 0x40000000u, kInOutside = 0x40000000u, kOutOutside = 0x80000000u;
typedef unsigned u4v __attribute__((ext_vector_type(4)));
__device__ __forceinline__ __amdgpu_buffer_rsrc_t buf_rsrc(const void *p, unsigned bytes) {
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(p), 0, (int)bytes, 0x00020000);
}
__device__ __forceinline__ int ch_of(int nb, int r16) { return nb * 16 + r16; }
__device__ __forceinline__ uint16_t bf16_bits(float f) {  // round to nearest even (v_cvt_pk_bf16_f32)
    return __builtin_bit_cast(uint16_t, (__bf16)f);
}

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

// An input row is loaded once per plane: fragment C holds voxel z0 + m of
// lane m (its channel slice), and H the two halo voxels, z0 - 1 in lane m = 0
// and z0 + 16 in lane m = 15 (of each 16-lane row). The z-shifted fragments
// of the K chunks come from C by DPP shifts within the 16-lane rows (lane m
// takes lane m - 1, or m + 1), with H at the row's end, so they equal the
// fragments a load of the shifted voxels would give, bit for bit.
template <int CIN>
constexpr int kcentres() {  // C fragments per row: the channel halves of CIN = 64
    return CIN == 64 ? 2 : 1;
}
// the channel offset of lane group g's slice in centre fragment h
template <int CIN>
__device__ __forceinline__ int centre_channel(int h, int g) {
    if constexpr (CIN == 8) return 0;
    else if constexpr (CIN == 16) return 8 * (g & 1);
    else if constexpr (CIN == 32) return 8 * g;
    else return 32 * h + 8 * g;
}
// whether lane group g of centre fragment h needs the halo below (z0 - 1) or above (z0 + 16)
template <int CIN>
__device__ __forceinline__ bool halo_lo(int g) {
    return CIN == 8 ? g == 0 : CIN == 16 ? g < 2 : true;
}
template <int CIN>
__device__ __forceinline__ bool halo_hi(int g) {
    return CIN == 8 ? g == 2 : CIN == 16 ? g < 2 : true;
}
// x of lane m - 1 (from_below) or m + 1 within each 16-lane row; the lane at
// the row's end, which has no such neighbour, keeps `end` (bound_ctrl off:
// the DPP move leaves it unwritten), so the halo needs no select.
__device__ __forceinline__ v8bf dpp_shift(v8bf x, v8bf end, bool from_below) {
    const u4v u = __builtin_bit_cast(u4v, x), e = __builtin_bit_cast(u4v, end);
    u4v r;
#pragma unroll
    for (int q = 0; q < 4; q++)
        r[q] = from_below ? (unsigned)__builtin_amdgcn_update_dpp((int)e[q], (int)u[q], 0x111, 0xf, 0xf, false)
                          : (unsigned)__builtin_amdgcn_update_dpp((int)e[q], (int)u[q], 0x101, 0xf, 0xf, false);
    return __builtin_bit_cast(v8bf, r);
}
// The KC chunks of one row (kslice's layout) from its centre and halo fragments.
template <int CIN>
__device__ __forceinline__ void row_chunks(const v8bf (&c)[kcentres<CIN>()], const v8bf (&h)[kcentres<CIN>()],
                                           int g, v8bf (&out)[kchunks<CIN>()]) {
    const v8bf zero = __builtin_bit_cast(v8bf, u4v{0u, 0u, 0u, 0u});
#pragma unroll
    for (int q = 0; q < kcentres<CIN>(); q++) {
        const v8bf lo = dpp_shift(c[q], h[q], true);   // voxel z0 + m - 1 (row_shr:1; m = 0: the halo)
        const v8bf hi = dpp_shift(c[q], h[q], false);  // voxel z0 + m + 1 (row_shl:1; m = 15: the halo)
        if constexpr (CIN == 8) {
            out[0] = g == 0 ? lo : g == 1 ? c[0] : g == 2 ? hi : zero;
        } else if constexpr (CIN == 16) {
            out[0] = g < 2 ? lo : c[0];
            out[1] = g < 2 ? hi : zero;
        } else {  // 32, 64: chunk dz + 3 q
            out[3 * q] = lo;
            out[3 * q + 1] = c[q];
            out[3 * q + 2] = hi;
        }
    }
}

// UP: the input is read nearest-up-sampled (a.up set) and PY is even, so y0
// is even and rows y0 + 2k - 1 and y0 + 2k (k >= 1) read the same source row:
// the second takes the first's fragments instead of a load.
template <int PY, bool UP>
constexpr int src_row(int iy) {
    return UP && PY % 2 == 0 && iy >= 2 && iy <= PY && iy % 2 == 0 ? iy - 1 : iy;
}

template <int CIN, int COUT, int PY, int WPS, bool UP = false>
__global__ __launch_bounds__(256, WPS) void k_conv3r(ConvArgs a, int run, int xruns) {
    static_assert(!UP || PY % 2 == 0, "the shared source rows need an even row count");
    constexpr int KC = kchunks<CIN>();
    constexpr int NB = COUT / 16;
    if (a.gate && *a.gate != 0.0) return;
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
    // this lane's input voxel (z) and channel byte offset per chunk within an
    // (x, y) row of the (possibly half-resolution) input: A[m][k] of lane
    // m + 16 (k / 8), m = z - z0 of the output row
    const int up = UP ? 1 : a.up;
    const int Xs = up ? X >> 1 : X, Ys = up ? Y >> 1 : Y, Zs = up ? Z >> 1 : Z;
    const auto rin = buf_rsrc(a.in, a.in_bytes);
    // 16 input channels and more: one centre and one halo fragment per row
    // and channel half (the chunks derived by DPP); 8 channels: the single
    // chunk, which already holds the three z shifts, loaded directly
    constexpr bool kDpp = CIN >= 16;
    constexpr int NC = kDpp ? kcentres<CIN>() : KC;
    // byte offsets within an (x, y) row: of the centre and halo fragments, or
    // of the chunks
    unsigned zc[NC], zh[NC];
#pragma unroll
    for (int q = 0; q < NC; q++) {
        if constexpr (kDpp) {
            const int c = centre_channel<CIN>(q, g);
            auto zoff = [&](int zi) { return (unsigned)(((up ? zi >> 1 : zi) * CIN + c) * 2); };
            const int zi = z0 + r16;
            zc[q] = zi < Z ? zoff(zi) : kInOutside;
            zh[q] = kInOutside;
            if (r16 == 0 && halo_lo<CIN>(g) && z0 > 0) zh[q] = zoff(z0 - 1);
            if (r16 == 15 && halo_hi<CIN>(g) && z0 + 16 < Z) zh[q] = zoff(z0 + 16);
        } else {
            int dz, c;
            bool ok;
            kslice<CIN>(q, g, dz, c, ok);
            const int zi = z0 + r16 + dz - 1;
            zc[q] = ok && zi >= 0 && zi < Z ? (unsigned)(((up ? zi >> 1 : zi) * CIN + c) * 2) : kInOutside;
            zh[q] = kInOutside;
        }
    }
    const unsigned row_bytes = (unsigned)(Zs * CIN * 2);
    // Offsets that do not change along the march, per lane: the input rows
    // y0 - 1 .. y0 + PY (their y part plus the z/channel part; kInOutside
    // parts where outside the volume), and the output values' (y, z, channel)
    // part (kOutOutside outside). Each step adds its plane's part, a value
    // common to the wavefront.
    unsigned rzc[PY + 2][NC], rzh[PY + 2][NC];
#pragma unroll
    for (int iy = 0; iy < PY + 2; iy++) {
        const int gy = y0 + iy - 1;
        const unsigned ry = gy >= 0 && gy < Y ? (unsigned)(up ? gy >> 1 : gy) * row_bytes : kInOutside;
#pragma unroll
        for (int q = 0; q < NC; q++) {
            rzc[iy][q] = ry + zc[q];
            rzh[iy][q] = ry + zh[q];
        }
    }
    unsigned ob[PY][4];  // byte offsets of the 16-bit outputs within a plane
#pragma unroll
    for (int oy = 0; oy < PY; oy++)
#pragma unroll
        for (int j = 0; j < 4; j++) {
            const int gy = y0 + oy, gz = z0 + 4 * g + j;
            ob[oy][j] = gy < Y && gz < Z ? (unsigned)((gy * Z + gz) * COUT + ch_of(nb, r16)) * 2 : kOutOutside;
        }
    v4f acc[3][PY];
#pragma unroll
    for (int s = 0; s < 3; s++)
#pragma unroll
        for (int oy = 0; oy < PY; oy++) acc[s][oy] = v4f{0.f, 0.f, 0.f, 0.f};
    const int steps = xe - xa + 2;  // input planes xa - 1 .. xe
    // The centre and halo fragments of input plane gx: rows y0 - 1 .. y0 +
    // PY, zero outside the volume (offsets past the buffer), so the next
    // plane's loads stay in flight during this plane's MFMAs.
    auto load_plane = [&](v8bf(&fc)[PY + 2][NC], v8bf(&fh)[PY + 2][NC], int gx) {
        // the plane's part (every term below 1 GiB or kInOutside: the sum of
        // three stays below 4 GiB, and past the buffer if any term is outside)
        const unsigned po = gx >= 0 && gx < X ? (unsigned)((b * Xs + (up ? gx >> 1 : gx)) * Ys) * row_bytes
                                              : kInOutside;
#pragma unroll
        for (int iy = 0; iy < PY + 2; iy++)
#pragma unroll
            for (int q = 0; q < NC; q++) {
                if (src_row<PY, UP>(iy) != iy) continue;  // (known after unrolling)
                fc[iy][q] = __builtin_bit_cast(v8bf, __builtin_amdgcn_raw_buffer_load_b128(rin, po + rzc[iy][q], 0, 0));
                if constexpr (kDpp)
                    fh[iy][q] =
                        __builtin_bit_cast(v8bf, __builtin_amdgcn_raw_buffer_load_b128(rin, po + rzh[iy][q], 0, 0));
            }
    };
    const int ch = nb * 16 + r16;
    const float bias = a.bias ? a.bias[ch] : 0.f;
    const auto rskip = buf_rsrc(a.skip, a.skip ? 2 * a.out_elems : 0);
    const auto rout = buf_rsrc(a.out, a.out ? 2 * a.out_elems : 0);
    const auto rout32 = buf_rsrc(a.out32, a.out32 ? 4 * a.out_elems : 0);
    const auto rrelu = buf_rsrc(a.relu_copy, a.relu_copy ? 2 * a.out_elems : 0);
    const bool has_skip = a.skip != nullptr, has_relu_copy = a.relu_copy != nullptr;
    const bool relu = a.relu != 0, f32 = a.out32 != nullptr;
    // Two fragment buffers alternate with the step's parity and three
    // accumulator slots rotate with the step mod 3: the march is unrolled by
    // six, so every index is a constant and no fragment is copied. Step t
    // reads input plane gx = xa - 1 + t; output plane gx + 1 - dx
    // accumulates in slot (t - dx) mod 3.
    v8bf fc[2][PY + 2][NC], fh[2][PY + 2][NC];
    load_plane(fc[0], fh[0], xa - 1);
    // Lc: whether this step loads the next plane (with UP, every other step:
    // planes 2k - 1 and 2k read the same source plane, see below)
    auto step = [&](auto Pc, auto Uc, auto Lc, int t) {
        constexpr int P = decltype(Pc)::value, U = decltype(Uc)::value;
        constexpr bool L = decltype(Lc)::value;
        const int gx = xa - 1 + t;
        // byte offset of output plane gx - 1 (the plane this step completes)
        // in the 16-bit outputs
        const unsigned pb = (unsigned)((b * X + (gx - 1)) * Y * Z * COUT) * 2;
        // the skip values of that plane, loaded before the next plane's
        // fragments so that waiting for them leaves those in flight
        uint16_t skv[PY][4];
        if (has_skip && t >= 2) {
#pragma unroll
            for (int oy = 0; oy < PY; oy++)
#pragma unroll
                for (int j = 0; j < 4; j++)
                    skv[oy][j] = __builtin_amdgcn_raw_buffer_load_b16(rskip, pb + ob[oy][j], 0, 0);
        }
        if constexpr (L) load_plane(fc[1 - P], fh[1 - P], gx + 1);  // (one past the run at the end: unused)
#pragma unroll
        for (int iy = 0; iy < PY + 2; iy++) {
            v8bf ck[KC];
            if constexpr (kDpp)
                row_chunks<CIN>(fc[P][src_row<PY, UP>(iy)], fh[P][src_row<PY, UP>(iy)], g, ck);
            else
#pragma unroll
                for (int kc = 0; kc < KC; kc++) ck[kc] = fc[P][src_row<PY, UP>(iy)][kc];
#pragma unroll
            for (int dx = 0; dx < 3; dx++)
#pragma unroll
                for (int dy = 0; dy < 3; dy++) {
                    const int oy = iy - dy;
                    if (oy < 0 || oy >= PY) continue;  // known after unrolling
#pragma unroll
                    for (int kc = 0; kc < KC; kc++)
                        acc[(U - dx + 3) % 3][oy] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(
                            ck[kc], wf[dx][dy][kc], acc[(U - dx + 3) % 3][oy], 0, 0, 0);
                }
        }
        // planes before the run (gx + 1 - dx < xa) took contributions in
        // the first two steps: clear their slots
        if (t == 0) {
#pragma unroll
            for (int oy = 0; oy < PY; oy++) acc[(U + 1) % 3][oy] = acc[(U + 2) % 3][oy] = v4f{0.f, 0.f, 0.f, 0.f};
        } else if (t == 1) {
#pragma unroll
            for (int oy = 0; oy < PY; oy++) acc[(U + 1) % 3][oy] = v4f{0.f, 0.f, 0.f, 0.f};
        }
        // output plane gx - 1 has had its three input planes
        if (t >= 2) {
            constexpr int S = (U + 1) % 3;
#pragma unroll
            for (int oy = 0; oy < PY; oy++) {
#pragma unroll
                for (int j = 0; j < 4; j++) {
                    const unsigned o = pb + ob[oy][j];  // past the buffers where outside the volume
                    float v = acc[S][oy][j] + bias;
                    if (relu) v = v > 0.f ? v : 0.f;
                    if (has_relu_copy) __builtin_amdgcn_raw_buffer_store_b16(bf16_bits(v), rrelu, o, 0, 0);
                    if (has_skip) v += bf2f(skv[oy][j]);
                    if (f32)  // (FP32 output: twice the 16-bit offset)
                        __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned, v), rout32,
                                                              ob[oy][j] == kOutOutside ? kOutOutside : 2 * o, 0, 0);
                    else
                        __builtin_amdgcn_raw_buffer_store_b16(bf16_bits(v), rout, o, 0, 0);
                }
                acc[S][oy] = v4f{0.f, 0.f, 0.f, 0.f};
            }
        }
    };
    using I0 = std::integral_constant<int, 0>;
    using I1 = std::integral_constant<int, 1>;
    using I2 = std::integral_constant<int, 2>;
    using LD = std::true_type;
    using NL = std::false_type;
    if constexpr (UP) {
        // The run starts on an even plane (the launcher makes runs even), so
        // step t's plane xa - 1 + t reads source plane (xa - 1 + t) >> 1: the
        // same as step t - 1's for even t >= 2. Steps with even t load the
        // next source plane, into the other buffer; odd steps load nothing,
        // and the next step reads their buffer again. Buffer of step t:
        // ((t + 1) >> 1) mod 2, period 4; with the accumulator slots' period
        // 3 the march is unrolled by twelve.
        for (int t0 = 0; t0 < steps; t0 += 12) {  // uniform over the wavefront
            step(I0{}, I0{}, LD{}, t0);
            if (t0 + 1 >= steps) break;
            step(I1{}, I1{}, NL{}, t0 + 1);
            if (t0 + 2 >= steps) break;
            step(I1{}, I2{}, LD{}, t0 + 2);
            if (t0 + 3 >= steps) break;
            step(I0{}, I0{}, NL{}, t0 + 3);
            if (t0 + 4 >= steps) break;
            step(I0{}, I1{}, LD{}, t0 + 4);
            if (t0 + 5 >= steps) break;
            step(I1{}, I2{}, NL{}, t0 + 5);
            if (t0 + 6 >= steps) break;
            step(I1{}, I0{}, LD{}, t0 + 6);
            if (t0 + 7 >= steps) break;
            step(I0{}, I1{}, NL{}, t0 + 7);
            if (t0 + 8 >= steps) break;
            step(I0{}, I2{}, LD{}, t0 + 8);
            if (t0 + 9 >= steps) break;
            step(I1{}, I0{}, NL{}, t0 + 9);
            if (t0 + 10 >= steps) break;
            step(I1{}, I1{}, LD{}, t0 + 10);
            if (t0 + 11 >= steps) break;
            step(I0{}, I2{}, NL{}, t0 + 11);
        }
    } else {
        for (int t0 = 0; t0 < steps; t0 += 6) {  // uniform over the wavefront
            step(I0{}, I0{}, LD{}, t0);
            if (t0 + 1 >= steps) break;
            step(I1{}, I1{}, LD{}, t0 + 1);
            if (t0 + 2 >= steps) break;
            step(I0{}, I2{}, LD{}, t0 + 2);
            if (t0 + 3 >= steps) break;
            step(I1{}, I0{}, LD{}, t0 + 3);
            if (t0 + 4 >= steps) break;
            step(I0{}, I1{}, LD{}, t0 + 4);
            if (t0 + 5 >= steps) break;
            step(I1{}, I2{}, LD{}, t0 + 5);
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
    constexpr int PY = CIN == 8 ? 4 : COUT == 64 ? 1 : 2;
    constexpr int WPS = CIN <= 32 ? 2 : 1;
    // x runs: about two wavefronts per SIMD of the 256 CUs
    const long long base = (long long)a.B * ((a.Z + 15) / 16) * ((a.Y + PY - 1) / PY) * (COUT / 16);
    const long long target = 2048;
    int xruns = (int)std::min<long long>(a.X, std::max<long long>(1, (target + base - 1) / base));
    const int run = (a.X + xruns - 1) / xruns;
    xruns = (a.X + run - 1) / run;
    const long long waves = base * xruns;
    // the decoder's up-sampled inputs (32 -> 16 and 64 -> 32 channels) take
    // the kernel that reads each shared source row once
    // (its runs start on even planes: each loaded source plane serves two)
    if constexpr (PY % 2 == 0 && ((CIN == 32 && COUT == 16) || (CIN == 64 && COUT == 32))) {
        if (a.up) {
            const int run2 = (run + 1) & ~1;
            const int xruns2 = (a.X + run2 - 1) / run2;
            const long long waves2 = base * xruns2;
            hipLaunchKernelGGL((k_conv3r<CIN, COUT, PY, WPS, true>), dim3((unsigned)((waves2 + 3) / 4)), dim3(256), 0,
                               st, a, run2, xruns2);
            SFP_LAUNCH_CHECK();
            return;
        }
    }
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
    const size_t in_bytes = (size_t)B * (up_input ? (size_t)(X / 2) * (Y / 2) * (Z / 2) : (size_t)X * Y * Z) * cin * 2;
    const size_t out_elems = (size_t)B * X * Y * Z * cout;
    if (in_bytes > kInLimit || out_elems * 4 > kOutOutside)
        throw ArgError("convolution tensors too large for 32-bit buffer offsets");
    ConvArgs a{in, w, bias, skip, out, out32, relu_copy, B, X, Y, Z, relu ? 1 : 0, up_input ? 1 : 0, gate,
               (unsigned)in_bytes, (unsigned)out_elems};
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
