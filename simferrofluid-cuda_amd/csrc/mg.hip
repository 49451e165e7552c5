// Geometric-multigrid kernels for gfx950: FP32 Chebyshev smoothing with the
// coarse-grid correction fused into the first post-smoothing step, one-pass
// pairs of steps (and of the last step with the restriction), residual +
// restriction, FP64 Galerkin coarsening over 2x2x2 aggregates, the dense
// coarsest solve, and the small levels of a cycle in one single-block launch.
//
// Threads map to cells with z (the unit-stride axis of the reference layout)
// across the 64 lanes of a wavefront and y across the 4 wavefronts of a
// block; each block covers one x plane of its tile (see cell_tile).
#include "device_util.hpp"
#include "kernels.hpp"

#include <algorithm>
#include <initializer_list>
#include <type_traits>

namespace sfp {

Launch cell_launch(Dims d, int planes, int max_planes) {
    Launch L;
    unsigned tz = 64;
    while (tz > 1 && (int)tz / 2 >= d.nz) tz /= 2;
    const unsigned ty = kBlock / tz;
    L.cg.gx = (d.nz + tz - 1) / tz;
    L.cg.gy = (d.ny + ty - 1) / ty;
    if (planes <= 0) {
        // march over several x planes per block (x-neighbour reuse in
        // registers) while keeping >= ~2048 blocks to fill 256 CUs
        const long long tiles = (long long)L.cg.gx * L.cg.gy * d.nx;
        planes = (int)std::min<long long>(max_planes, std::max<long long>(1, tiles / 2048));
    }
    L.cg.planes = planes;
    const unsigned gz = (d.nx + L.cg.planes - 1) / L.cg.planes;
    L.block = dim3(tz, ty, 1);
    L.grid = dim3(L.cg.gx * L.cg.gy * gz, 1, 1);
    return L;
}

namespace {

// true when a batched solve has already converged (see KrylovScalars::stop)
__device__ __forceinline__ bool gated(const double *gate) { return gate && *gate != 0.0; }


template <bool FROM64>
__global__ __launch_bounds__(kBlock) void k_smooth_first(CellGrid cg, Dims d, Op32 A, const double *r64,
                                                         float *b, float *x, float c0,
                                                         const double *gate) {
    if (gated(gate)) return;
    int k, j, i0;
    cell_tile(cg, k, j, i0);
    if (k >= d.nz || j >= d.ny) return;
    for (int i = i0, ie = min(i0 + cg.planes, d.nx); i < ie; i++) {
        const size_t c = d.idx(i, j, k);
        float bv;
        if constexpr (FROM64) {
            bv = (float)r64[c];
            b[c] = bv;
        } else {
            bv = b[c];
        }
        x[c] = c0 * rcp_or_zero(A.diag[c]) * bv;
    }
}

// One Chebyshev step xout = x + d', d' = cd d + cr D^{-1} (b - A x), marching
// over the block's x planes with the x-direction neighbours kept in
// registers and the z-direction neighbours (iterate and z coupling) taken
// from the adjacent lanes of the tile row (only the row's end lanes load
// them). The input iterate x is, by mode:
//   SM_FIRST   : the zero-start first step c0 D^{-1} b, recomputed at every
//                stencil point (b read from r64 and rounded when r64 != null,
//                which also stores b); the previous direction equals it
//   SM_PROLONG : xin + P e (piecewise-constant coarse correction)
//   otherwise  : xin
// Every lane of a row takes part in the shuffles, so lanes past the grid
// edge stay in the loop with their loads and stores masked.
template <int MODE>
__global__ __launch_bounds__(kBlock) void k_smooth(CellGrid cg, Dims d, Op32 A,
                                                   const double *__restrict__ r64,
                                                   float *__restrict__ b,
                                                   const float *__restrict__ xin,
                                                   float *__restrict__ xout, float *dvec,
                                                   float c0, float cd, float cr,
                                                   const float *__restrict__ e, Dims dc,
                                                   const double *gate) {
    if (gated(gate)) return;
    constexpr bool kFirst = (MODE & SM_FIRST) != 0;
    constexpr bool kProlong = (MODE & SM_PROLONG) != 0;
    constexpr bool kXFirst = (MODE & SM_XFIRST) != 0;
    int k, j, i0;
    cell_tile(cg, k, j, i0);
    const int tz = blockDim.x;
    const int lane = threadIdx.x;  // position in the tile row
    const bool on = k < d.nz && j < d.ny;
    if (i0 >= d.nx) return;  // uniform over the block
    const Strides s(d);
    auto bval = [&](size_t q) -> float { return r64 ? (float)r64[q] : b[q]; };
    // coarse-grid offsets (prolongation): plane stride, and this thread's
    // column and its y neighbours' columns within a coarse plane
    const size_t dcsx = (size_t)dc.ny * dc.nz;
    size_t eo = 0, eym = 0, eyp = 0;
    if constexpr (kProlong) {
        if (on) {
            eo = dc.idx(0, j >> 1, k >> 1);
            if (j > 0) eym = dc.idx(0, (j - 1) >> 1, k >> 1);
            if (j + 1 < d.ny) eyp = dc.idx(0, (j + 1) >> 1, k >> 1);
        }
    }
    // iterate value at a neighbour point q (cell coordinates qi, qj, qk; for
    // the prolongation, ecol = the point's column offset on the coarse grid)
    auto xval = [&](size_t q, int qi, size_t ecol) -> float {
        if constexpr (kFirst) {
            return c0 * rcp_or_zero(A.diag[q]) * bval(q);
        } else if constexpr (kProlong) {
            return xin[q] + e[(size_t)(qi >> 1) * dcsx + ecol];
        } else {
            return xin[q];
        }
    };
    const int ie = min(i0 + cg.planes, d.nx);
    const bool zlo = k > 0, zhi = k + 1 < d.nz;
    size_t c = on ? d.idx(i0, j, k) : 0;
    float xm = 0.f, xc = 0.f, cxc = 0.f;
    bool primed = false;  // xm, xc, cxc hold this plane's values
    const bool ylo = on && j > 0, yhi = on && j + 1 < d.ny;
    for (int i = i0; i < ie; i++, c += s.sx) {
        const float dg = on ? A.diag[c] : 0.f;
        // a wavefront whose cells are all inactive has nothing to update (their
        // vectors hold zeros and couple with weight 0): skip it whole
        if (!__any(on && dg > 0.f)) {
            primed = false;
            continue;
        }
        if (!primed) {
            xm = (on && i > 0) ? xval(c - s.sx, i - 1, eo) : 0.f;
            xc = on ? xval(c, i, eo) : 0.f;
            cxc = on ? A.cpl[0][c] : 0.f;
            primed = true;
        }
        // every load of the plane is issued before the lane exchange
        const bool up = i + 1 < d.nx;
        const float xp = (on && up) ? xval(c + s.sx, i + 1, eo) : 0.f;
        const float cxp = (on && up) ? A.cpl[0][c + s.sx] : 0.f;
        const float czm = on ? A.cpl[2][c] : 0.f;
        const float cym = ylo ? A.cpl[1][c] : 0.f;
        const float xym = ylo ? xval(c - s.sy, i, eym) : 0.f;
        const float cyp = yhi ? A.cpl[1][c + s.sy] : 0.f;
        const float xyp = yhi ? xval(c + s.sy, i, eyp) : 0.f;
        float bc = 0.f;
        if constexpr (kFirst) {
            bc = on ? bval(c) : 0.f;
        } else {
            bc = on ? b[c] : 0.f;
        }
        float dprev = 0.f;
        if constexpr ((MODE & SM_READ_D) != 0) dprev = on ? dvec[c] : 0.f;
        float xin_c = 0.f;
        if constexpr (kProlong) xin_c = on ? xin[c] : 0.f;
        float xzm = __shfl_up(xc, 1, tz);
        float xzp = __shfl_down(xc, 1, tz);
        float czp = __shfl_down(czm, 1, tz);
        if (lane == 0) xzm = (on && zlo) ? xval(c - 1, i, kProlong ? dc.idx(0, j >> 1, (k - 1) >> 1) : 0) : 0.f;
        if (lane == tz - 1) {
            xzp = (on && zhi) ? xval(c + 1, i, kProlong ? dc.idx(0, j >> 1, (k + 1) >> 1) : 0) : 0.f;
            czp = (on && zhi) ? A.cpl[2][c + 1] : 0.f;
        }
        if (on) {
            float off = fmaf(cxc, xm, cxp * xp);
            if (j > 0) off = fmaf(cym, xym, off);
            if (j + 1 < d.ny) off = fmaf(cyp, xyp, off);
            if (zlo) off = fmaf(czm, xzm, off);
            if (zhi) off = fmaf(czp, xzp, off);
            if constexpr (kFirst) {
                if (r64) b[c] = bc;
            }
            // centre iterate: the coarse correction only lands on active cells
            float x0 = xc;
            if constexpr (kProlong) {
                if (!(dg > 0.f)) x0 = xin_c;
            }
            const float r = bc - fmaf(dg, x0, -off);
            if constexpr (kFirst || kXFirst) dprev = x0;
            const float dn = fmaf(cd, dprev, cr * rcp_or_zero(dg) * r);
            xout[c] = x0 + dn;
            if constexpr ((MODE & SM_WRITE_D) != 0) dvec[c] = dn;
        }
        xm = xc;
        xc = xp;
        cxc = cxp;
    }
}

// The same step (modes without SM_FIRST) over the flat cell index, four
// consecutive cells of one z row per lane (grids with nz a multiple of 4):
// 16-byte loads of every array at the cell, its x and y neighbours and the
// upper x and y couplings, one more iterate value and z coupling at each end
// of the group. With SM_PROLONG the coarse correction (two coarse cells per
// group and row) is added to every iterate value read and, on active cells,
// to the centre, as k_prolong_add followed by a plain step would. Every cell's
// arithmetic is k_smooth's (same operands, same order), so the two give the
// same bits. A wavefront whose 256 cells are all inactive writes nothing.
constexpr unsigned kSmoothFlatCells = 4 * kBlock;

// Row-sum FP32 diagonal (Op32::pinned_rowsum) of one cell from its six
// couplings in the order -x, +x, -y, +y, -z, +z (0 where there is no
// neighbour): summed in FP64 (exact for couplings within 2^27 of each other)
// and rounded once, the assembly's arithmetic, so the bits are the stored
// diagonal's.
__device__ __forceinline__ float rowsum_diag(float xl, float xh, float yl, float yh, float zl, float zh) {
    return (float)(((((((double)xl + (double)xh) + (double)yl) + (double)yh) + (double)zl) + (double)zh));
}
// The same for the four cells of a group: the lower x and y couplings (cxc,
// cym), the upper ones (cxp, cyp: the next plane's and row's lower), the
// lower z ones (cz) and the one above the group (czhi).
typedef float f4v __attribute__((ext_vector_type(4)));
__device__ __forceinline__ f4v rowsum_diag4(f4v cxc, f4v cxp, f4v cym, f4v cyp, f4v cz, float czhi) {
    f4v dg;
#pragma unroll
    for (int q = 0; q < 4; q++) dg[q] = rowsum_diag(cxc[q], cxp[q], cym[q], cyp[q], cz[q], q < 3 ? cz[q + 1] : czhi);
    return dg;
}
// The group starting at cell c (c a multiple of 4, nz a multiple of 4) holds
// none of the cells whose diagonal is not the row sum (1, nz, nz*ny).
__device__ __forceinline__ bool rowsum_group(const Op32 &A, size_t c, Dims d) {
    return A.pinned_rowsum && c != 0 && c != (size_t)d.nz && c != (size_t)d.ny * d.nz;
}

template <int MODE, bool kRowsum>
__global__ __launch_bounds__(kBlock) void k_smooth_flat(Dims d, Op32 A, const float *__restrict__ b,
                                                        const float *__restrict__ xin,
                                                        float *__restrict__ xout, float *dvec,
                                                        float cd, float cr,
                                                        const float *__restrict__ e, Dims dc,
                                                        const double *gate) {
    static_assert((MODE & SM_FIRST) == 0, "flat step: no fused first step");
    constexpr bool kProlong = (MODE & SM_PROLONG) != 0;
    typedef float f4 __attribute__((ext_vector_type(4)));
    typedef float f2 __attribute__((ext_vector_type(2)));
    if (gated(gate)) return;
    const Strides s(d);
    const size_t c = (size_t)blockIdx.x * kSmoothFlatCells + 4 * (size_t)flat_tid();
    const bool on = c < d.cells();  // cells() is a multiple of 4
    const f4 z4 = {0.f, 0.f, 0.f, 0.f};
    const bool rs = kRowsum && on && rowsum_group(A, c, d);  // every group there has an active cell
    f4 dg = (on && !rs) ? *(const f4 *)(A.diag + c) : z4;
    const bool act = rs || dg.x > 0.f || dg.y > 0.f || dg.z > 0.f || dg.w > 0.f;
    if (!__any(act) || !on) return;
    // (i, j, k) of the first cell (the launcher keeps cells() < 2^32)
    const unsigned cu = (unsigned)c;
    const unsigned t = cu / (unsigned)d.nz;
    const int k = (int)(cu - t * (unsigned)d.nz);
    const unsigned iu = t / (unsigned)d.ny;
    const int i = (int)iu, j = (int)(t - iu * (unsigned)d.ny);
    const bool up = i + 1 < d.nx, ylo = j > 0, yhi = j + 1 < d.ny;
    const bool zl = k > 0, zh = k + 4 < d.nz;  // neighbours below the first, above the last
    f4 xc = *(const f4 *)(xin + c);
    f4 xm = i > 0 ? *(const f4 *)(xin + c - s.sx) : z4;
    f4 xp = up ? *(const f4 *)(xin + c + s.sx) : z4;
    const f4 cxc = *(const f4 *)(A.cpl[0] + c);
    const f4 cxp = up ? *(const f4 *)(A.cpl[0] + c + s.sx) : z4;
    const f4 cym = ylo ? *(const f4 *)(A.cpl[1] + c) : z4;
    f4 xym = ylo ? *(const f4 *)(xin + c - s.sy) : z4;
    const f4 cyp = yhi ? *(const f4 *)(A.cpl[1] + c + s.sy) : z4;
    f4 xyp = yhi ? *(const f4 *)(xin + c + s.sy) : z4;
    const f4 cz = *(const f4 *)(A.cpl[2] + c);
    float xzlo = zl ? xin[c - 1] : 0.f;
    float xzhi = zh ? xin[c + 4] : 0.f;
    const float czhi = zh ? A.cpl[2][c + 4] : 0.f;
    if (rs) dg = rowsum_diag4(cxc, cxp, cym, cyp, cz, czhi);
    const f4 bc = *(const f4 *)(b + c);
    f4 dprev = z4;
    if constexpr ((MODE & SM_READ_D) != 0) dprev = *(const f4 *)(dvec + c);
    const f4 xin_c = xc;
    if constexpr (kProlong) {
        // coarse values of the group's two coarse cells in this row and in the
        // neighbouring rows and planes (16-byte aligned groups: k >> 1 even)
        const int kc = k >> 1;
        auto e2 = [&](int ii, int jj) { return *(const f2 *)(e + dc.idx(ii >> 1, jj >> 1, kc)); };
        const f2 ec = e2(i, j);
        const f2 e_xm = i > 0 ? e2(i - 1, j) : f2{0.f, 0.f};
        const f2 e_xp = up ? e2(i + 1, j) : f2{0.f, 0.f};
        const f2 e_ym = ylo ? e2(i, j - 1) : f2{0.f, 0.f};
        const f2 e_yp = yhi ? e2(i, j + 1) : f2{0.f, 0.f};
        const float ezlo = zl ? e[dc.idx(i >> 1, j >> 1, kc - 1)] : 0.f;
        const float ezhi = zh ? e[dc.idx(i >> 1, j >> 1, kc + 2)] : 0.f;
#pragma unroll
        for (int q = 0; q < 4; q++) {
            xc[q] = xc[q] + ec[q >> 1];
            xm[q] = xm[q] + e_xm[q >> 1];
            xp[q] = xp[q] + e_xp[q >> 1];
            xym[q] = xym[q] + e_ym[q >> 1];
            xyp[q] = xyp[q] + e_yp[q >> 1];
        }
        xzlo = xzlo + ezlo;
        xzhi = xzhi + ezhi;
    }
    f4 xo, dno;
#pragma unroll
    for (int q = 0; q < 4; q++) {
        const float xzm = q > 0 ? xc[q - 1] : xzlo;
        const float xzp = q < 3 ? xc[q + 1] : xzhi;
        const float czp = q < 3 ? cz[q + 1] : czhi;
        float off = fmaf(cxc[q], xm[q], cxp[q] * xp[q]);
        if (ylo) off = fmaf(cym[q], xym[q], off);
        if (yhi) off = fmaf(cyp[q], xyp[q], off);
        if (q > 0 || zl) off = fmaf(cz[q], xzm, off);
        if (q < 3 || zh) off = fmaf(czp, xzp, off);
        // centre iterate: the coarse correction only lands on active cells
        const float x0 = (kProlong && !(dg[q] > 0.f)) ? xin_c[q] : xc[q];
        const float r = bc[q] - fmaf(dg[q], x0, -off);
        const float dp = (MODE & SM_XFIRST) != 0 ? x0 : dprev[q];
        const float dn = fmaf(cd, dp, cr * rcp_or_zero(dg[q]) * r);
        xo[q] = x0 + dn;
        dno[q] = dn;
    }
    *(f4 *)(xout + c) = xo;
    if constexpr ((MODE & SM_WRITE_D) != 0) *(f4 *)(dvec + c) = dno;
}

// Two passes over the grid in one (temporal blocking). Step A is a
// Chebyshev step from iterate x1 and direction d1:
//   d2 = cdA d1 + crA D^{-1} (b - A x1),   x2 = x1 + d2.
// Step B is, with kRestrict false, a second Chebyshev step
//   x3 = x2 + cdB d2 + crB D^{-1} (b - A x2)   (written to xo; d2 not stored),
// and with kRestrict true the restriction of the residual of x2,
//   bc = R (b - A x2)   (x2 written to xo),
// with the per-cell arithmetic of the separate launches (k_smooth with
// SM_READ_D | SM_WRITE_D, then SM_READ_D; or k_smooth with SM_READ_D, then
// k_restrict), so results agree bit for bit.
//
// A block of 64 lanes (z) x kPairRows wavefronts (y) marches over x planes,
// computing x2 on its whole 64 x kPairRows tile one plane ahead of step B;
// the rows exchange x2 through LDS, and step B covers the tile's interior
// (lanes 1..62, rows 1..kPairRows-2), so adjacent tiles overlap by one cell
// on each side. The interior starts on even z, y and x and has even extents,
// so each coarse cell's 8 children lie in one block: z pairs through the lane
// pair, y pairs through LDS, x pairs in a register. A wavefront whose row is
// inactive on a plane skips its work there (x2 there is 0, as the buffers
// already hold). xo must not alias x1.
//
// LDS holds x2 (and cpl_y) of three consecutive planes: the one barrier per
// plane then also separates a plane's reads from the store that reuses its
// buffer (issued two planes later, by a wavefront at most one barrier ahead).
constexpr int kPairRows = 16;
constexpr int kPairZ = 62, kPairY = kPairRows - 2;  // interior extents (tile strides)

template <bool kRestrict, bool kRowsum>
__global__ __launch_bounds__(64 * kPairRows, 8) void k_smooth_pair(
    CellGrid cg, Dims d, Op32 A, const float *__restrict__ b, const float *__restrict__ x1,
    float *__restrict__ xo, const float *__restrict__ d1, float cdA, float crA, float cdB,
    float crB, Dims dc, float *__restrict__ bc, const double *gate) {
    if (gated(gate)) return;
    __shared__ float sx2[3][kPairRows][64];
    __shared__ float scy[3][kPairRows][64];  // cpl_y of the same three planes
    __shared__ float sres[kRestrict ? kPairRows : 1][64];
    // Blocks go to the 8 XCDs in turn (block b to XCD b % 8); renumbered so
    // that each XCD takes a contiguous run of tiles, the rows, planes and
    // cache lines that neighbouring tiles share are read into one L2 (the
    // tiles overlap, and their z extents are not line-aligned). A tail of
    // fewer than 8 blocks keeps its numbers.
    unsigned L = blockIdx.x;
    {
        const unsigned q = gridDim.x / 8;
        if (L < 8 * q) L = (L % 8) * q + L / 8;
    }
    const unsigned per = cg.gx * cg.gy;
    const unsigned bx = L / per, rem = L - bx * per;
    const unsigned by = rem / cg.gx, bz = rem - by * cg.gx;
    const int lane = threadIdx.x, row = threadIdx.y;
    const int k = (int)bz * kPairZ - 1 + lane;
    const int j = (int)by * kPairY - 1 + row;
    const int i0 = (int)bx * cg.planes;
    if (i0 >= d.nx) return;  // uniform over the block
    const int ie = min(i0 + cg.planes, d.nx);
    const bool on = k >= 0 && k < d.nz && j >= 0 && j < d.ny;
    const bool interior = on && lane >= 1 && lane <= kPairZ && row >= 1 && row <= kPairY;
    const bool zlo = k > 0, zhi = k + 1 < d.nz, ylo = on && j > 0, yhi = on && j + 1 < d.ny;
    const Strides s(d);
    const size_t c0 = on ? d.idx(0, j, k) : 0;  // cell (i, j, k) = c0 + i sx when on

    // one plane of step A, kept for step B. Step B takes the couplings of the
    // upper faces from where they are stored: cpl_x(i+1) from the next
    // plane, cpl_y(j+1) from the next row (LDS), cpl_z(k+1) from the next
    // lane; where that cell is inactive the coupling is 0 either way.
    struct Plane {
        float x2, d2, dg, bc, cxc, cym, czm;
        bool live;
    };
    // step-A pipeline for the next plane q: x1 at q-1 and q, cpl_x at q
    float xm = 0.f, xc = 0.f, cxc = 0.f;
    bool primed = false;
    auto step_a = [&](int q) -> Plane {
        Plane P{};
        if (q >= d.nx) return P;
        const size_t c = c0 + q * s.sx;
        // a row-sum diagonal is rebuilt below (cells 1, nz and nz*ny read it)
        const bool rs = kRowsum && on && A.pinned_rowsum && c != 1 && c != (size_t)d.nz && c != s.sx;
        float dg = (on && !rs) ? A.diag[c] : 0.f;
        P.live = __any(on && (rs || dg > 0.f));
        if (!P.live) {
            primed = false;
            return P;
        }
        if (!primed) {
            xm = (on && q > 0) ? x1[c - s.sx] : 0.f;
            xc = on ? x1[c] : 0.f;
            cxc = on ? A.cpl[0][c] : 0.f;
            primed = true;
        }
        const bool up = q + 1 < d.nx;
        const float xp = (on && up) ? x1[c + s.sx] : 0.f;
        const float cxp = (on && up) ? A.cpl[0][c + s.sx] : 0.f;
        const float czm = on ? A.cpl[2][c] : 0.f;
        const float cym = ylo ? A.cpl[1][c] : 0.f;
        const float xym = ylo ? x1[c - s.sy] : 0.f;
        const float cyp = yhi ? A.cpl[1][c + s.sy] : 0.f;
        const float xyp = yhi ? x1[c + s.sy] : 0.f;
        const float bc = on ? b[c] : 0.f;
        const float dprev = on ? d1[c] : 0.f;
        float xzm = __shfl_up(xc, 1, 64);
        float xzp = __shfl_down(xc, 1, 64);
        float czp = __shfl_down(czm, 1, 64);
        if (lane == 0) xzm = (on && zlo) ? x1[c - 1] : 0.f;
        if (lane == 63) {
            xzp = (on && zhi) ? x1[c + 1] : 0.f;
            czp = (on && zhi) ? A.cpl[2][c + 1] : 0.f;
        }
        if (rs) dg = rowsum_diag(cxc, cxp, cym, cyp, czm, czp);
        if (on) {
            float off = fmaf(cxc, xm, cxp * xp);
            if (j > 0) off = fmaf(cym, xym, off);
            if (j + 1 < d.ny) off = fmaf(cyp, xyp, off);
            if (zlo) off = fmaf(czm, xzm, off);
            if (zhi) off = fmaf(czp, xzp, off);
            const float r = bc - fmaf(dg, xc, -off);
            const float dn = fmaf(cdA, dprev, crA * rcp_or_zero(dg) * r);
            P.x2 = xc + dn;
            if constexpr (kRestrict) {
                if (interior) xo[c] = P.x2;
            } else {
                P.d2 = dn;
            }
        }
        P.dg = dg;
        P.bc = bc;
        P.cxc = cxc;
        P.cym = cym;
        P.czm = czm;
        xm = xc;
        xc = xp;
        cxc = cxp;
        return P;
    };

    Plane pm{}, pc{};  // x2 at planes i - 1 and i
    if (i0 > 0) pm = step_a(i0 - 1);
    pc = step_a(i0);
    int bcur = i0 % 3;  // LDS buffer of plane i
    sx2[bcur][row][lane] = pc.x2;
    scy[bcur][row][lane] = pc.cym;
    float acc = 0.f;  // restriction: sum over the x pair
    // the lane of each coarse cell's first child stores the sum
    const bool writer = kRestrict && interior && (k & 1) == 0 && (j & 1) == 0;
    const size_t bco = writer ? dc.idx(0, j >> 1, k >> 1) : 0;
    const size_t dcsx = (size_t)dc.ny * dc.nz;
    for (int i = i0; i < ie; i++) {
        const int bnext = bcur == 2 ? 0 : bcur + 1;  // buffer of plane i + 1
        const Plane pn = step_a(i + 1);
        sx2[bnext][row][lane] = pn.x2;
        scy[bnext][row][lane] = pn.cym;
        __syncthreads();
        const float cxp = pn.cxc;
        const float czp = __shfl_down(pc.czm, 1, 64);  // lane 63 is never interior
        if constexpr (!kRestrict) {
            if (pc.live) {
                const float xzm = __shfl_up(pc.x2, 1, 64);
                const float xzp = __shfl_down(pc.x2, 1, 64);
                if (interior) {
                    const float xym = sx2[bcur][row - 1][lane];
                    const float xyp = sx2[bcur][row + 1][lane];
                    const float cyp = scy[bcur][row + 1][lane];
                    float off = fmaf(pc.cxc, pm.x2, cxp * pn.x2);
                    if (j > 0) off = fmaf(pc.cym, xym, off);
                    if (j + 1 < d.ny) off = fmaf(cyp, xyp, off);
                    if (zlo) off = fmaf(pc.czm, xzm, off);
                    if (zhi) off = fmaf(czp, xzp, off);
                    const float r = pc.bc - fmaf(pc.dg, pc.x2, -off);
                    const float dn = fmaf(cdB, pc.d2, crB * rcp_or_zero(pc.dg) * r);
                    xo[c0 + i * s.sx] = pc.x2 + dn;
                }
            }
        } else {
            // residual of x2 on the interior's active cells (the operand
            // order of k_restrict), then the 2 x 2 x 2 sums
            float r = 0.f;
            if (pc.live) {
                const float xzm = __shfl_up(pc.x2, 1, 64);
                const float xzp = __shfl_down(pc.x2, 1, 64);
                if (interior && pc.dg > 0.f) {
                    const float xym = sx2[bcur][row - 1][lane];
                    const float xyp = sx2[bcur][row + 1][lane];
                    const float cyp = scy[bcur][row + 1][lane];
                    float off = 0.f;
                    if (i > 0) off = fmaf(pc.cxc, pm.x2, off);
                    if (i + 1 < d.nx) off = fmaf(cxp, pn.x2, off);
                    if (j > 0) off = fmaf(pc.cym, xym, off);
                    if (j + 1 < d.ny) off = fmaf(cyp, xyp, off);
                    if (zlo) off = fmaf(pc.czm, xzm, off);
                    if (zhi) off = fmaf(czp, xzp, off);
                    r = pc.bc - fmaf(pc.dg, pc.x2, -off);
                }
            }
            r += __shfl_down(r, 1, 64);  // z pair: lanes of even k take k + 1
            sres[row][lane] = r;
            __syncthreads();
            if (row < kPairRows - 1) r += sres[row + 1][lane];  // y pair (rows of even j)
            acc += r;                                             // x pair
            if ((i & 1) == 1 || i + 1 == d.nx) {
                if (writer) bc[bco + (size_t)(i >> 1) * dcsx] = acc;
                acc = 0.f;
            }
        }
        pm = pc;
        pc = pn;
        bcur = bnext;
    }
}

// The same pair of steps (step A, then step B with kRestrict false) on grids
// whose z rows hold at most 256 cells, in groups of four: a lane holds four
// consecutive cells of a z row (16-byte accesses), a wavefront a whole row,
// a block kRowPairRows rows of y, and the block marches over its run of x
// planes. Tiles overlap by one row on each side in y (rows 0 and
// kRowPairRows - 1 are step A's halo) and not at all in z. The per-cell
// arithmetic is k_smooth_flat's (two launches: SM_READ_D | SM_WRITE_D, or
// SM_XFIRST | SM_WRITE_D when d1 is x1 (kDx), then SM_READ_D), bit for bit.
//
// Each iteration of the march, for plane i:
//   1. puts x1 and cpl_y of plane i + 2 (loaded one iteration earlier) in
//      LDS, with the out-of-tile rows next to the tile (rows 0 and
//      kRowPairRows - 1 load them);
//   2. issues the loads of x1, cpl_x, cpl_y of plane i + 3 and of cpl_z, b,
//      d1 and the diagonal of plane i + 2, so one iteration hides their
//      latency;
//   3. computes step A on plane i + 1 (x1 of the rows above and below and
//      their cpl_y from LDS) and puts x2 in LDS; one barrier;
//   4. computes step B on plane i (x2 of the three planes and of the rows
//      above and below from LDS) and stores it on the tile's interior rows.
// LDS rings: x1 two planes deep (written an iteration before it is read, and
// rewritten only after the next barrier), x2 three (written before the
// barrier that precedes its readers, which a wavefront one barrier ahead must
// not disturb) and cpl_y four (read by step A in one iteration and by step B
// in the next). `rowact` (null: every row active) marks the (plane, row) z
// rows holding an active cell: the others load nothing and stay 0, as the
// vectors hold there.
constexpr int kRowPairRows = 12;
constexpr int kRowPairY = kRowPairRows - 2;  // interior rows: the tiles' stride in y
// the out-of-tile rows take one value per thread: 3 parts of 256 z values
static_assert(64 * kRowPairRows >= 3 * 256, "k_smooth_pair_rows: too few threads for the halo rows");

template <bool kRowsum, bool kDx>
__global__ __launch_bounds__(64 * kRowPairRows, 1) void k_smooth_pair_rows(
    Dims d, Op32 A, const float *__restrict__ b, const float *__restrict__ x1,
    float *__restrict__ xo, const float *__restrict__ d1, float cdA, float crA, float cdB,
    float crB, const unsigned char *__restrict__ rowact, int planes, unsigned ytiles,
    const double *gate) {
    typedef float f4 __attribute__((ext_vector_type(4)));
    if (gated(gate)) return;
    __shared__ f4 sx1[2][kRowPairRows + 2][64];  // x1 of rows -1 .. kRowPairRows (index row + 1)
    __shared__ f4 scy[4][kRowPairRows + 1][64];  // cpl_y of rows 0 .. kRowPairRows
    __shared__ f4 sx2[3][kRowPairRows][64];      // x2 of rows 0 .. kRowPairRows - 1
    // logical tiles run y fastest; each XCD (block b goes to XCD b % 8) takes
    // a contiguous run of them, so the rows two tiles share meet in one L2
    unsigned L = blockIdx.x;
    {
        const unsigned q = gridDim.x / 8;
        if (L < 8 * q) L = (L % 8) * q + L / 8;
    }
    const unsigned by = L % ytiles, bx = L / ytiles;
    const int i0 = (int)bx * planes;
    if (i0 >= d.nx) return;  // uniform over the block
    const int ie = min(i0 + planes, d.nx);
    const int lane = threadIdx.x, row = threadIdx.y;
    const int k = 4 * lane;
    const int j = (int)by * kRowPairY - 1 + row;
    const bool zon = k < d.nz;
    const bool yin = j >= 0 && j < d.ny;
    const bool interior = row >= 1 && row <= kRowPairY;
    const bool ylo = j > 0, yhi = j + 1 < d.ny;
    const bool zl = k > 0, zh = k + 4 < d.nz;
    const Strides s(d);
    // element offsets fit 32 bits (the launcher keeps cells() < 2^32); lanes
    // past the row's end read its last group (a cache hit) and drop it
    const unsigned c0 = yin ? (unsigned)d.idx(0, j, min(k, d.nz - 4)) : 0u;
    const unsigned sx = (unsigned)s.sx;
    // the out-of-tile rows next to the tile (x1 of rows j0 - 1 and
    // j0 + kRowPairRows - 2, cpl_y of the latter), one value per thread: thread
    // t of the block takes z = t mod 256 of part t / 256
    const int t = row * 64 + lane;
    const int hpart = t >> 8, hz = t & 255;  // part 3: nothing
    const int jt = (int)by * kRowPairY - 1;  // the tile's row 0
    const int jh = hpart == 0 ? jt - 1 : jt + kRowPairRows;
    const bool hon_any = hpart < 3 && jh >= 0 && jh < d.ny;  // uniform over the wavefront
    const unsigned ch = hon_any ? (unsigned)d.idx(0, jh, min(hz, d.nz - 1)) : 0u;
    const f4 z4 = {0.f, 0.f, 0.f, 0.f};
    // activity of this row and of the out-of-tile row for planes i0 - 2 ..
    // i0 + 61 (bit q - i0 + 2), one lane per plane, read once
    auto row_mask = [&](int jj) -> unsigned long long {
        const int q = i0 - 2 + lane;
        const bool a = q >= 0 && q < d.nx && jj >= 0 && jj < d.ny &&
                       (!rowact || rowact[(size_t)q * d.ny + jj] != 0);
        return __ballot(a);
    };
    const unsigned long long amask = row_mask(j);
    const unsigned long long hmask_lo = row_mask(jt - 1), hmask_hi = row_mask(jt + kRowPairRows);
    const unsigned long long hmask = hpart == 0 ? hmask_lo : hmask_hi;
    auto active = [&](int q, unsigned long long m) { return ((m >> (q - i0 + 2)) & 1ull) != 0; };
    auto ld = [&](const float *p) { return zon ? *(const f4 *)p : z4; };

    struct Early {
        f4 x, cx, cy;  // x1, cpl_x, cpl_y
        float h;       // this thread's value of the out-of-tile rows
    };
    struct Rest {
        f4 cz, bv, dv, dg;
    };
    // loads branch on the wavefront's row activity only (uniform)
    auto load_early = [&](int q) {
        Early e{z4, z4, z4, 0.f};
        if (active(q, amask)) {
            const unsigned c = c0 + (unsigned)q * sx;
            e.x = ld(x1 + c);
            e.cx = ld(A.cpl[0] + c);
            e.cy = ld(A.cpl[1] + c);
        }
        if (hon_any && active(q, hmask)) {
            const unsigned h = ch + (unsigned)q * sx;
            const float v = hpart == 2 ? A.cpl[1][h] : x1[h];
            e.h = hz < d.nz ? v : 0.f;
        }
        return e;
    };
    auto load_rest = [&](int q) {
        Rest r{z4, z4, z4, z4};
        if (active(q, amask)) {
            const unsigned c = c0 + (unsigned)q * sx;
            r.cz = ld(A.cpl[2] + c);
            r.bv = ld(b + c);
            if (!kDx) r.dv = ld(d1 + c);
            // a row-sum diagonal is rebuilt; the groups of cells 1, nz and
            // nz*ny (on planes 0 and 1, rows 0 and 1) read it
            if (!kRowsum || !A.pinned_rowsum || (q <= 1 && j <= 1)) r.dg = ld(A.diag + c);
        }
        return r;
    };
    auto put_early = [&](int q, const Early &e) {
        sx1[q & 1][row + 1][lane] = e.x;
        scy[q & 3][row][lane] = e.cy;
        if (hpart == 0) ((float *)&sx1[q & 1][0][0])[hz] = e.h;
        else if (hpart == 1) ((float *)&sx1[q & 1][kRowPairRows + 1][0])[hz] = e.h;
        else if (hpart == 2) ((float *)&scy[q & 3][kRowPairRows][0])[hz] = e.h;
    };
    // one Chebyshev step of the four cells at plane q from iterate xc (x
    // neighbours xm, xp; y neighbours and couplings given), direction dp
    auto cheb = [&](f4 xc, f4 xm, f4 xp, f4 xym, f4 xyp, f4 cxc, f4 cxp, f4 cym, f4 cyp, f4 cz,
                    f4 bv, f4 dg, f4 rdg, f4 dp, float cd, float cr, f4 &dn_out) {
        // every lane takes part in the shuffles: a lane outside the exec mask
        // does not supply its value to ds_bpermute (the reader gets 0)
        const float xzlo = __shfl_up(xc.w, 1, 64);
        const float xzhi = __shfl_down(xc.x, 1, 64);
        const float czdn = __shfl_down(cz.x, 1, 64);
        const float czhi = zh ? czdn : 0.f;
        f4 xo4;
#pragma unroll
        for (int q = 0; q < 4; q++) {
            const float xzm = q > 0 ? xc[q - 1] : xzlo;
            const float xzp = q < 3 ? xc[q + 1] : xzhi;
            const float czp = q < 3 ? cz[q + 1] : czhi;
            float off = fmaf(cxc[q], xm[q], cxp[q] * xp[q]);
            if (ylo) off = fmaf(cym[q], xym[q], off);
            if (yhi) off = fmaf(cyp[q], xyp[q], off);
            if (q > 0 || zl) off = fmaf(cz[q], xzm, off);
            if (q < 3 || zh) off = fmaf(czp, xzp, off);
            const float r = bv[q] - fmaf(dg[q], xc[q], -off);
            const float dn = fmaf(cd, dp[q], cr * rdg[q] * r);
            xo4[q] = xc[q] + dn;
            dn_out[q] = dn;
        }
        return xo4;
    };
    // the diagonal at plane q: loaded, or rebuilt from the couplings
    auto diag_at = [&](int q, f4 cxc, f4 cxp, f4 cym, f4 cyp, const Rest &r) {
        f4 dg = r.dg;
        if (kRowsum && A.pinned_rowsum) {
            const unsigned c = c0 + (unsigned)q * sx;
            const float czdn = __shfl_down(r.cz.x, 1, 64);  // all lanes (see cheb)
            const float czhi = zh ? czdn : 0.f;
            if (rowsum_group(A, c, d)) dg = rowsum_diag4(cxc, cxp, cym, cyp, r.cz, czhi);
        }
        return dg;
    };
    // step A at plane q: x2 (returned) and d2; x1 of this row at planes q and
    // q + 1 is read back from LDS (this wavefront wrote it). Step B forms the
    // plane's diagonal and its reciprocal again from the same operands (bit
    // for bit the same values): carried in registers for an iteration they
    // made the kernel spill.
    auto step_a = [&](int q, f4 xm, const Early &e, f4 cxp, const Rest &r, f4 &d2) {
        const f4 xc = sx1[q & 1][row + 1][lane], xp = sx1[(q + 1) & 1][row + 1][lane];
        const f4 xym = sx1[q & 1][row][lane], xyp = sx1[q & 1][row + 2][lane];
        const f4 cym = scy[q & 3][row][lane], cyp = scy[q & 3][row + 1][lane];
        const f4 dg = diag_at(q, e.cx, cxp, cym, cyp, r);
        f4 rdg;
#pragma unroll
        for (int k = 0; k < 4; k++) rdg[k] = rcp_or_zero(dg[k]);
        d2 = z4;
        if (!active(q, amask)) return z4;  // uniform over the wavefront
        f4 x2 = cheb(xc, xm, xp, xym, xyp, e.cx, cxp, cym, cyp, r.cz, r.bv, dg, rdg, kDx ? xc : r.dv,
                     cdA, crA, d2);
        if (!zon) x2 = z4;
        return x2;
    };

    // ---- prologue: planes i0 - 1 and i0 through step A, plane i0 + 1 in LDS
    const f4 xpre = load_early(i0 - 2).x;
    Early em = load_early(i0 - 1), e0 = load_early(i0), e1 = load_early(i0 + 1);
    Early e2 = load_early(i0 + 2);
    Rest rm = load_rest(i0 - 1), r0 = load_rest(i0), r1 = load_rest(i0 + 1);
    // step A at i0 - 1 reads planes i0 - 1 and i0 from the x1 ring, step A at
    // i0 planes i0 and i0 + 1 (which takes the buffer of i0 - 1)
    put_early(i0 - 1, em);
    put_early(i0, e0);
    __syncthreads();
    f4 dtmp, d2c;
    const f4 x2m = step_a(i0 - 1, xpre, em, e0.cx, rm, dtmp);
    __syncthreads();
    put_early(i0 + 1, e1);
    __syncthreads();
    const f4 x2c = step_a(i0, em.x, e0, e1.cx, r0, d2c);
    sx2[(i0 + 2) % 3][row][lane] = x2m;  // (i0 - 1) mod 3
    sx2[i0 % 3][row][lane] = x2c;
    __syncthreads();
    // Rings indexed by plane - i0 mod 3 (the march is unrolled by three, so
    // every index is a constant and nothing is copied between iterations):
    // E holds x1 / cpl_x / cpl_y of planes i + 1 and i + 2 and receives plane
    // i + 3 in the slot of plane i; R the rest of planes i, i + 1 (and i + 2,
    // loaded into the slot of plane i - 1); D2 the direction of planes i and
    // i + 1 from step A.
    Early E[3] = {e0, e1, e2};
    Rest R[3] = {r0, r1, Rest{z4, z4, z4, z4}};
    f4 D2[3];
    D2[0] = d2c;
    f4 xa = e0.x;     // x1 at plane i
    f4 cxa = e0.cx;   // cpl_x at plane i
    f4 x2prev = x2m;  // x2 at plane i - 1 (this row)
    (void)x2c;
    auto iter = [&](auto Uc, int i) {
        constexpr int U = decltype(Uc)::value, U1 = (U + 1) % 3, U2 = (U + 2) % 3;
        // 1. loads for the next iteration, in flight during this one's two
        // steps: their slots hold planes i and i - 1, which this iteration
        // no longer reads (plane i's x1 and cpl_x are in xa, cxa)
        E[U] = load_early(i + 3);
        R[U2] = load_rest(i + 2);
        // 2. x1 and cpl_y of plane i + 2 into LDS
        put_early(i + 2, E[U2]);
        // 3. step A on plane i + 1
        const f4 x2n = step_a(i + 1, xa, E[U1], E[U2].cx, R[U1], D2[U1]);
        sx2[(i + 1) % 3][row][lane] = x2n;
        __syncthreads();
        // 4. step B on plane i (the tile's interior rows)
        if (interior) {
            const int bi = i % 3;
            const f4 xc = sx2[bi][row][lane];
            if (active(i, amask)) {
                const f4 xym = sx2[bi][row - 1][lane], xyp = sx2[bi][row + 1][lane];
                const f4 cym = scy[i & 3][row][lane], cyp = scy[i & 3][row + 1][lane];
                const f4 dgB = diag_at(i, cxa, E[U1].cx, cym, cyp, R[U]);
                f4 rdgB;
#pragma unroll
                for (int k = 0; k < 4; k++) rdgB[k] = rcp_or_zero(dgB[k]);
                f4 dn;
                const f4 x3 = cheb(xc, x2prev, x2n, xym, xyp, cxa, E[U1].cx, cym, cyp, R[U].cz, R[U].bv, dgB,
                                   rdgB, D2[U], cdB, crB, dn);
                if (zon) *(f4 *)(xo + c0 + (unsigned)i * sx) = x3;
            }
            x2prev = xc;
        }
        xa = E[U1].x;
        cxa = E[U1].cx;
    };
    for (int i = i0; i < ie; i += 3) {  // uniform over the block
        iter(std::integral_constant<int, 0>{}, i);
        if (i + 1 >= ie) break;
        iter(std::integral_constant<int, 1>{}, i + 1);
        if (i + 2 >= ie) break;
        iter(std::integral_constant<int, 2>{}, i + 2);
    }
}

// x += P e on active cells (the piecewise-constant coarse correction), in
// place, before a plain first post-smoothing step: on large grids this pass
// and the step cost less than the step that adds the correction at every
// neighbour it reads (the coarse values' gathers cost a third of that step).
__global__ __launch_bounds__(kBlock) void k_prolong_add(CellGrid cg, Dims d, const float *__restrict__ diag32,
                                                        const float *__restrict__ e, Dims dc,
                                                        float *__restrict__ x, const double *gate) {
    if (gated(gate)) return;
    int k, j, i0;
    cell_tile(cg, k, j, i0);
    if (k >= d.nz || j >= d.ny) return;
    const Strides s(d);
    const size_t dcsx = (size_t)dc.ny * dc.nz;
    const size_t eo = dc.idx(0, j >> 1, k >> 1);
    size_t c = d.idx(i0, j, k);
    for (int i = i0, ie = min(i0 + cg.planes, d.nx); i < ie; i++, c += s.sx) {
        if (!(diag32[c] > 0.f)) continue;
        x[c] = x[c] + e[(size_t)(i >> 1) * dcsx + eo];
    }
}

// Residual b - A x on the fine grid, summed over each coarse cell's (up to 8)
// children. Threads map to fine cells like the smoother (coalesced rows,
// x-neighbours in registers, z-neighbours from adjacent lanes); the children
// are added pairwise: z through the lane pair, y through LDS (rows 2m and
// 2m+1 of a block), x over the block's plane pairs. The block's plane count
// is even and starts on an even plane, and tz, ty are even (or tz = 1).
__global__ __launch_bounds__(kBlock) void k_restrict(CellGrid cg, Dims df, Op32 A, const float *__restrict__ b,
                                                     const float *__restrict__ x, Dims dc,
                                                     float *__restrict__ bc, float *__restrict__ x0c,
                                                     const float *__restrict__ diag32c, float c0,
                                                     const double *gate) {
    if (gated(gate)) return;
    __shared__ float sh[kBlock];
    int k, j, i0;
    cell_tile(cg, k, j, i0);
    const int tz = blockDim.x, lane = threadIdx.x, row = threadIdx.y;
    if (i0 >= df.nx) return;  // uniform over the block
    const bool on = k < df.nz && j < df.ny;
    const bool writer = on && (k & 1) == 0 && (j & 1) == 0;
    const Strides s(df);
    const int ie = min(i0 + cg.planes, df.nx);
    const bool zlo = k > 0, zhi = k + 1 < df.nz;
    size_t c = on ? df.idx(i0, j, k) : 0;
    float xm = 0.f, xc = 0.f, cxc = 0.f;
    bool primed = false;  // xm, xc, cxc hold this plane's values
    float acc = 0.f;
    for (int i = i0; i < ie; i++, c += s.sx) {
        const bool up = i + 1 < df.nx;
        const float dg = on ? A.diag[c] : 0.f;
        float r = 0.f;
        // a wavefront whose cells are all inactive contributes 0: skip its loads
        if (__any(on && dg > 0.f)) {
            if (!primed) {
                xm = (on && i > 0) ? x[c - s.sx] : 0.f;
                xc = on ? x[c] : 0.f;
                cxc = on ? A.cpl[0][c] : 0.f;
                primed = true;
            }
            const float xp = (on && up) ? x[c + s.sx] : 0.f;
            const float cxp = (on && up) ? A.cpl[0][c + s.sx] : 0.f;
            const float czm = on ? A.cpl[2][c] : 0.f;
            float xzm = __shfl_up(xc, 1, tz);
            float xzp = __shfl_down(xc, 1, tz);
            float czp = __shfl_down(czm, 1, tz);
            if (lane == 0) xzm = (on && zlo) ? x[c - 1] : 0.f;
            if (lane == tz - 1) {
                xzp = (on && zhi) ? x[c + 1] : 0.f;
                czp = (on && zhi) ? A.cpl[2][c + 1] : 0.f;
            }
            if (on && dg > 0.f) {
                float off = 0.f;
                if (i > 0) off = fmaf(cxc, xm, off);
                if (up) off = fmaf(cxp, xp, off);
                if (j > 0) off = fmaf(A.cpl[1][c], x[c - s.sy], off);
                if (j + 1 < df.ny) off = fmaf(A.cpl[1][c + s.sy], x[c + s.sy], off);
                if (zlo) off = fmaf(czm, xzm, off);
                if (zhi) off = fmaf(czp, xzp, off);
                r = b[c] - fmaf(dg, xc, -off);
            }
            xm = xc;
            xc = xp;
            cxc = cxp;
        } else {
            primed = false;
        }
        if (tz > 1) r += __shfl_xor(r, 1, tz);  // z pair
        sh[row * tz + lane] = r;
        __syncthreads();
        if ((row & 1) == 0) r += sh[(row + 1) * tz + lane];  // y pair
        __syncthreads();
        acc += r;  // x pair
        if ((i & 1) == 1 || !up) {
            if (writer) {
                const size_t q = dc.idx(i >> 1, j >> 1, k >> 1);
                bc[q] = acc;
                if (x0c) x0c[q] = c0 * rcp_or_zero(diag32c[q]) * acc;
            }
            acc = 0.f;
        }
    }
}

// The same restriction over the flat cell index (grids with nz a multiple of
// 4): a wavefront covers 64 cells of a z row in two rows and two planes, four
// cells per lane (lanes 0-15: row j, plane i; 16-31: row j+1; 32-63: plane
// i+1), so a coarse cell's children are added in k_restrict's order -- the z
// pair in the lane, the y pair across lanes 16 apart, the x pair across lanes
// 32 apart, onto a zero -- and the first 16 lanes store two coarse values
// each. Each active cell's residual is k_restrict's arithmetic, so the two
// give the same values. A wavefront whose 256 cells are all inactive only
// stores zeros.
constexpr int kRestrictFlatZ = 64;  // z cells per wavefront

template <bool kRowsum>
__global__ __launch_bounds__(kBlock) void k_restrict_flat(Dims df, Op32 A, const float *__restrict__ b,
                                                          const float *__restrict__ x, Dims dc,
                                                          float *__restrict__ bc, float *__restrict__ x0c,
                                                          const float *__restrict__ diag32c, float c0,
                                                          const double *gate) {
    typedef float f4 __attribute__((ext_vector_type(4)));
    typedef float f2 __attribute__((ext_vector_type(2)));
    if (gated(gate)) return;
    const int lane = threadIdx.x & 63;
    const unsigned wave = (unsigned)(blockIdx.x * kBlock + threadIdx.x) >> 6;
    const unsigned zch = (unsigned)(df.nz + kRestrictFlatZ - 1) / kRestrictFlatZ;
    const unsigned yp = (unsigned)(df.ny + 1) / 2;
    const unsigned kz = wave % zch, t = wave / zch;
    const unsigned jp = t % yp, ip = t / yp;
    if (ip >= (unsigned)(df.nx + 1) / 2) return;  // uniform over the wavefront (no barriers follow)
    const int i = 2 * (int)ip + (lane >> 5), j = 2 * (int)jp + ((lane >> 4) & 1);
    const int k = kRestrictFlatZ * (int)kz + 4 * (lane & 15);
    const bool on = i < df.nx && j < df.ny && k < df.nz;
    const Strides s(df);
    const size_t c = on ? df.idx(i, j, k) : 0;
    const f4 z4 = {0.f, 0.f, 0.f, 0.f};
    const bool rs = kRowsum && on && rowsum_group(A, c, df);  // every group there has an active cell
    f4 dg = (on && !rs) ? *(const f4 *)(A.diag + c) : z4;
    f4 r = z4;
    if (__any(rs || dg.x > 0.f || dg.y > 0.f || dg.z > 0.f || dg.w > 0.f) && on) {
        const bool up = i + 1 < df.nx, ylo = j > 0, yhi = j + 1 < df.ny;
        const bool zl = k > 0, zh = k + 4 < df.nz;
        const f4 xc = *(const f4 *)(x + c);
        const f4 xm = i > 0 ? *(const f4 *)(x + c - s.sx) : z4;
        const f4 xp = up ? *(const f4 *)(x + c + s.sx) : z4;
        const f4 cxc = *(const f4 *)(A.cpl[0] + c);
        const f4 cxp = up ? *(const f4 *)(A.cpl[0] + c + s.sx) : z4;
        const f4 cym = ylo ? *(const f4 *)(A.cpl[1] + c) : z4;
        const f4 xym = ylo ? *(const f4 *)(x + c - s.sy) : z4;
        const f4 cyp = yhi ? *(const f4 *)(A.cpl[1] + c + s.sy) : z4;
        const f4 xyp = yhi ? *(const f4 *)(x + c + s.sy) : z4;
        const f4 cz = *(const f4 *)(A.cpl[2] + c);
        const float xzlo = zl ? x[c - 1] : 0.f;
        const float xzhi = zh ? x[c + 4] : 0.f;
        const float czhi = zh ? A.cpl[2][c + 4] : 0.f;
        if (rs) dg = rowsum_diag4(cxc, cxp, cym, cyp, cz, czhi);
        const f4 bv = *(const f4 *)(b + c);
#pragma unroll
        for (int q = 0; q < 4; q++) {
            if (!(dg[q] > 0.f)) continue;
            const float xzm = q > 0 ? xc[q - 1] : xzlo;
            const float xzp = q < 3 ? xc[q + 1] : xzhi;
            const float czp = q < 3 ? cz[q + 1] : czhi;
            float off = 0.f;
            if (i > 0) off = fmaf(cxc[q], xm[q], off);
            if (up) off = fmaf(cxp[q], xp[q], off);
            if (ylo) off = fmaf(cym[q], xym[q], off);
            if (yhi) off = fmaf(cyp[q], xyp[q], off);
            if (q > 0 || zl) off = fmaf(cz[q], xzm, off);
            if (q < 3 || zh) off = fmaf(czp, xzp, off);
            r[q] = bv[q] - fmaf(dg[q], xc[q], -off);
        }
    }
    float s0 = r.x + r.y, s1 = r.z + r.w;  // z pairs
    s0 += __shfl_xor(s0, 16);              // y pairs
    s1 += __shfl_xor(s1, 16);
    const float o0 = __shfl_xor(s0, 32), o1 = __shfl_xor(s1, 32);  // the other plane's
    if (lane < 16 && on) {
        f2 v;
        v.x = 0.f;
        v.x += s0;
        v.x += o0;
        v.y = 0.f;
        v.y += s1;
        v.y += o1;
        const size_t q = dc.idx((int)ip, (int)jp, k >> 1);
        *(f2 *)(bc + q) = v;
        if (x0c) {
            const f2 dgc = *(const f2 *)(diag32c + q);
            f2 x0;
            x0.x = c0 * rcp_or_zero(dgc.x) * v.x;
            x0.y = c0 * rcp_or_zero(dgc.y) * v.y;
            *(f2 *)(x0c + q) = x0;
        }
    }
}

// Zeros the level vectors (and, if given, one FP64 vector of the first
// level) at the inactive cells (FP32 diagonal 0): a new operator may make
// inactive cells that were unknowns before, and the values left there would
// survive in the groups of inactive cells the stencil kernels skip. Active
// cells are written before they are read in every cycle, so they need no
// clearing. Every level in one launch (the small levels' launches would each
// cost more than their bytes), one cell per lane (consecutive lanes,
// consecutive cells: the sparse stores stay coalesced).
__global__ __launch_bounds__(kBlock) void k_clear_inactive(ClearArgs a) {
    const unsigned gid = blockIdx.x * kBlock + threadIdx.x;
    if (gid >= a.start[a.nlev]) return;
    int l = 0;
    while (gid >= a.start[l + 1]) l++;
    const unsigned c = gid - a.start[l];
    if (a.diag32[l][c] > 0.f) return;
#pragma unroll
    for (int k = 0; k < 4; k++) a.v[l][k][c] = 0.f;
    if (l == 0 && a.w) a.w[c] = 0.0;
}

// Galerkin coarsening over the fine grid (the mapping of k_restrict): each
// fine cell forms its contributions to its aggregate -- the active cell's
// row sum diag - (its 6 couplings), its lower-face couplings where it lies on
// the aggregate's lower layer of an axis, its upper-face couplings where it
// lies on the upper layer -- and the 8 children are added pairwise (z through
// the lane pair, y through LDS, x over the block's plane pairs). Same operator
// as a per-aggregate loop over the children, up to the order of the sums.
// The coarse couplings are the summed ones times `scale`, the coarse diagonal
// the summed row sum times `rs_scale` plus the coarse couplings.
constexpr int kCoarsenSums = 8;  // row sum, 3 lower, 3 upper, active count

__global__ __launch_bounds__(kBlock) void k_coarsen_fine(CellGrid cg, Dims df, Op64 Af,
                                                         const float *__restrict__ diag32f, Dims dc,
                                                         double scale, double rs_scale, double *diag64c,
                                                         MutFaces cpl64c, float *diag32c,
                                                         float *cpl32c0, float *cpl32c1,
                                                         float *cpl32c2) {
    __shared__ double sh[kCoarsenSums][kBlock];
    int k, j, i0;
    cell_tile(cg, k, j, i0);
    const int tz = blockDim.x, lane = threadIdx.x, row = threadIdx.y;
    if (i0 >= df.nx) return;  // uniform over the block
    const bool on = k < df.nz && j < df.ny;
    const bool writer = on && (k & 1) == 0 && (j & 1) == 0;
    const Strides s(df);
    const int ie = min(i0 + cg.planes, df.nx);
    const int t = row * tz + lane;
    float *cpl32c[3] = {cpl32c0, cpl32c1, cpl32c2};
    double acc[kCoarsenSums];
#pragma unroll
    for (int q = 0; q < kCoarsenSums; q++) acc[q] = 0.0;
    for (int i = i0; i < ie; i++) {
        double v[kCoarsenSums];
#pragma unroll
        for (int q = 0; q < kCoarsenSums; q++) v[q] = 0.0;
        if (on) {
            const size_t c = df.idx(i, j, k);
            const bool up[3] = {i + 1 < df.nx, j + 1 < df.ny, k + 1 < df.nz};
            const double cl[3] = {Af.cpl[0][c], Af.cpl[1][c], Af.cpl[2][c]};
            const double cu[3] = {up[0] ? Af.cpl[0][c + s.sx] : 0.0, up[1] ? Af.cpl[1][c + s.sy] : 0.0,
                                  up[2] ? Af.cpl[2][c + 1] : 0.0};
            const int idx[3] = {i, j, k};
            if (diag32f[c] > 0.f) {
                double off = cl[0] + cl[1] + cl[2];
                if (up[0]) off += cu[0];
                if (up[1]) off += cu[1];
                if (up[2]) off += cu[2];
                v[0] = Af.diag[c] - off;
                v[7] = 1.0;
            }
#pragma unroll
            for (int a = 0; a < 3; a++) {
                if ((idx[a] & 1) == 0) v[1 + a] = cl[a];  // lower layer: the face below
                else v[4 + a] = cu[a];                     // upper layer: the face above
            }
        }
        if (tz > 1) {
#pragma unroll
            for (int q = 0; q < kCoarsenSums; q++) v[q] += __shfl_xor(v[q], 1, tz);  // z pair
        }
#pragma unroll
        for (int q = 0; q < kCoarsenSums; q++) sh[q][t] = v[q];
        __syncthreads();
        if ((row & 1) == 0) {
#pragma unroll
            for (int q = 0; q < kCoarsenSums; q++) v[q] += sh[q][t + tz];  // y pair
        }
        __syncthreads();
#pragma unroll
        for (int q = 0; q < kCoarsenSums; q++) acc[q] += v[q];  // x pair
        if ((i & 1) == 1 || i + 1 == df.nx) {
            if (writer) {
                const int I = i >> 1, J = j >> 1, K = k >> 1;
                const int IC[3] = {I, J, K};
                double lower[3], upper[3];
                for (int a = 0; a < 3; a++) {
                    lower[a] = IC[a] > 0 ? scale * acc[1 + a] : 0.0;
                    upper[a] = IC[a] + 1 < dc.n(a) ? scale * acc[4 + a] : 0.0;
                }
                const bool active = acc[7] > 0.0;
                const size_t cc = dc.idx(I, J, K);
                if (active) {
                    const double dg = rs_scale * acc[0] + lower[0] + upper[0] + lower[1] + upper[1] +
                                      lower[2] + upper[2];
                    diag64c[cc] = dg;
                    diag32c[cc] = dg > 0.0 ? (float)dg : 0.f;
                } else {
                    diag64c[cc] = 1.0;
                    diag32c[cc] = 0.f;
                }
                for (int a = 0; a < 3; a++) {
                    const double cv = active ? lower[a] : 0.0;
                    cpl64c.a[a][cc] = cv;
                    cpl32c[a][cc] = (float)cv;
                }
            }
#pragma unroll
            for (int q = 0; q < kCoarsenSums; q++) acc[q] = 0.0;
        }
    }
}

// The same coarsening over the flat cell index (grids with nz a multiple of
// 4), in k_restrict_flat's mapping: a wavefront covers 64 cells of a z row in
// two rows and two planes, four cells per lane (lanes 0-15: row j, plane i;
// 16-31: row j+1; 32-63: plane i+1). Each cell's contributions are
// k_coarsen_fine's, and the children are added in its order -- the z pair in
// the lane, the y pair across lanes 16 apart, the x pair across lanes 32
// apart onto a zero -- so the coarse operator is the same to the bit; the
// first 16 lanes store two coarse cells each.
__global__ __launch_bounds__(kBlock) void k_coarsen_flat(Dims df, Op64 Af, const float *__restrict__ diag32f,
                                                         Dims dc, double scale, double rs_scale,
                                                         double *diag64c, MutFaces cpl64c,
                                                         float *diag32c, float *cpl32c0, float *cpl32c1,
                                                         float *cpl32c2) {
    typedef double d2 __attribute__((ext_vector_type(2)));
    typedef float f4 __attribute__((ext_vector_type(4)));
    const int lane = threadIdx.x & 63;
    const unsigned wave = (unsigned)(blockIdx.x * kBlock + threadIdx.x) >> 6;
    const unsigned zch = (unsigned)(df.nz + 63) / 64;
    const unsigned yp = (unsigned)(df.ny + 1) / 2;
    const unsigned kz = wave % zch, t = wave / zch;
    const unsigned jp = t % yp, ip = t / yp;
    if (ip >= (unsigned)(df.nx + 1) / 2) return;  // uniform over the wavefront (no barriers follow)
    const int i = 2 * (int)ip + (lane >> 5), j = 2 * (int)jp + ((lane >> 4) & 1);
    const int k = 64 * (int)kz + 4 * (lane & 15);
    const bool on = i < df.nx && j < df.ny && k < df.nz;
    const Strides s(df);
    // v[h][q]: the sums of coarse z cell k / 2 + h (row sum, 3 lower, 3 upper, active count)
    double v[2][kCoarsenSums];
#pragma unroll
    for (int h = 0; h < 2; h++)
#pragma unroll
        for (int q = 0; q < kCoarsenSums; q++) v[h][q] = 0.0;
    if (on) {
        const size_t c = df.idx(i, j, k);
        const bool upx = i + 1 < df.nx, upy = j + 1 < df.ny;
        auto ld4 = [&](const double *p, double (&o)[4]) {
            const d2 a = *(const d2 *)p, b = *(const d2 *)(p + 2);
            o[0] = a.x;
            o[1] = a.y;
            o[2] = b.x;
            o[3] = b.y;
        };
        double cl0[4], cl1[4], cl2[4], cu0[4] = {0.0, 0.0, 0.0, 0.0}, cu1[4] = {0.0, 0.0, 0.0, 0.0}, dg[4];
        ld4(Af.cpl[0] + c, cl0);
        ld4(Af.cpl[1] + c, cl1);
        ld4(Af.cpl[2] + c, cl2);
        if (upx) ld4(Af.cpl[0] + c + s.sx, cu0);
        if (upy) ld4(Af.cpl[1] + c + s.sy, cu1);
        const double cz4 = k + 4 < df.nz ? Af.cpl[2][c + 4] : 0.0;  // the face above the group
        ld4(Af.diag + c, dg);
        const f4 d32 = *(const f4 *)(diag32f + c);
#pragma unroll
        for (int q = 0; q < 4; q++) {
            const int kq = k + q;
            const bool upz = kq + 1 < df.nz;
            const double cl[3] = {cl0[q], cl1[q], cl2[q]};
            const double cu[3] = {upx ? cu0[q] : 0.0, upy ? cu1[q] : 0.0,
                                  upz ? (q < 3 ? cl2[q + 1] : cz4) : 0.0};
            const bool upv[3] = {upx, upy, upz};
            const int idx[3] = {i, j, kq};
            double w[kCoarsenSums];
#pragma unroll
            for (int e = 0; e < kCoarsenSums; e++) w[e] = 0.0;
            if (d32[q] > 0.f) {
                double off = cl[0] + cl[1] + cl[2];
                if (upv[0]) off += cu[0];
                if (upv[1]) off += cu[1];
                if (upv[2]) off += cu[2];
                w[0] = dg[q] - off;
                w[7] = 1.0;
            }
#pragma unroll
            for (int a = 0; a < 3; a++) {
                if ((idx[a] & 1) == 0) w[1 + a] = cl[a];  // lower layer: the face below
                else w[4 + a] = cu[a];                     // upper layer: the face above
            }
            // the z pair: the even cell's sums first, then the odd cell's added
#pragma unroll
            for (int e = 0; e < kCoarsenSums; e++) v[q >> 1][e] = (q & 1) ? v[q >> 1][e] + w[e] : w[e];
        }
    }
    // y pair (the row-j lanes keep row j's sums first), then the x pair onto 0
    double acc[2][kCoarsenSums];
#pragma unroll
    for (int h = 0; h < 2; h++)
#pragma unroll
        for (int e = 0; e < kCoarsenSums; e++) {
            const double y = v[h][e] + __shfl_xor(v[h][e], 16, 64);
            const double o = __shfl_xor(y, 32, 64);
            double a = 0.0;
            a += y;
            if (2 * (int)ip + 1 < df.nx) a += o;
            acc[h][e] = a;
        }
    if (lane < 16 && on) {
        float *cpl32c[3] = {cpl32c0, cpl32c1, cpl32c2};
        const int I = (int)ip, J = (int)jp;
#pragma unroll
        for (int h = 0; h < 2; h++) {
            const int K = (k >> 1) + h;
            const int IC[3] = {I, J, K};
            double lower[3], upper[3];
            for (int a = 0; a < 3; a++) {
                lower[a] = IC[a] > 0 ? scale * acc[h][1 + a] : 0.0;
                upper[a] = IC[a] + 1 < dc.n(a) ? scale * acc[h][4 + a] : 0.0;
            }
            const bool active = acc[h][7] > 0.0;
            const size_t cc = dc.idx(I, J, K);
            if (active) {
                const double dgc =
                    rs_scale * acc[h][0] + lower[0] + upper[0] + lower[1] + upper[1] + lower[2] + upper[2];
                diag64c[cc] = dgc;
                diag32c[cc] = dgc > 0.0 ? (float)dgc : 0.f;
            } else {
                diag64c[cc] = 1.0;
                diag32c[cc] = 0.f;
            }
            for (int a = 0; a < 3; a++) {
                const double cv = active ? lower[a] : 0.0;
                cpl64c.a[a][cc] = cv;
                cpl32c[a][cc] = (float)cv;
            }
        }
    }
}

// Dense inverse of the coarsest operator (n <= 64 cells; SPD up to
// pure-Neumann blocks) by in-place Gauss-Jordan in LDS, rows stored with a
// fixed stride of 64 so the element loops split an index with a shift and a
// mask. Pivots below 1e-12 of the largest diagonal (the pure-Neumann
// direction) are dropped: their row and column stay 0. One block of
// kInvertThreads: every thread owns S * S / kInvertThreads matrix entries, so
// a pivot step is a few FP64 operations per thread between two barriers.
constexpr int kInvertThreads = 1024;

__global__ __launch_bounds__(kInvertThreads) void k_coarse_invert(Dims d, Op64 A, float *ainv) {
    constexpr int S = kMaxCoarse;  // row stride
    static_assert(S == 64, "the element split below assumes a stride of 64");
    __shared__ double M[S * S];
    __shared__ double rowk[S], colk[S];
    __shared__ double scale_sh;
    const int n = (int)d.cells();
    const int tid = threadIdx.x;
    const Strides s(d);
    for (int e = tid; e < S * S; e += kInvertThreads) M[e] = 0.0;
    __syncthreads();
    for (int c = tid; c < n; c += kInvertThreads) {
        const int i = c / (d.ny * d.nz), j = (c / d.nz) % d.ny, k = c % d.nz;
        M[c * S + c] = A.diag[c];
        if (i > 0) M[c * S + (c - (int)s.sx)] = -A.cpl[0][c];
        if (i + 1 < d.nx) M[c * S + (c + (int)s.sx)] = -A.cpl[0][c + s.sx];
        if (j > 0) M[c * S + (c - (int)s.sy)] = -A.cpl[1][c];
        if (j + 1 < d.ny) M[c * S + (c + (int)s.sy)] = -A.cpl[1][c + s.sy];
        if (k > 0) M[c * S + (c - 1)] = -A.cpl[2][c];
        if (k + 1 < d.nz) M[c * S + (c + 1)] = -A.cpl[2][c + 1];
    }
    __syncthreads();
    if (tid == 0) {
        double m = 0.0;
        for (int c = 0; c < n; c++) m = fmax(m, fabs(M[c * S + c]));
        scale_sh = m;
    }
    __syncthreads();
    const double tiny = 1e-12 * scale_sh;
    for (int kk = 0; kk < n; kk++) {
        const double p = M[kk * S + kk];
        if (tid < n) {
            rowk[tid] = M[kk * S + tid];
            colk[tid] = M[tid * S + kk];
        }
        __syncthreads();
        if (fabs(p) <= tiny) {
            if (tid < n) {
                M[kk * S + tid] = 0.0;
                M[tid * S + kk] = 0.0;
            }
        } else {
            const double pinv = 1.0 / p;  // the same value in every thread
            for (int e = tid; e < S * S; e += kInvertThreads) {
                const int r = e >> 6, c = e & (S - 1);
                if (r >= n || c >= n) continue;
                double v;
                if (r == kk && c == kk)
                    v = pinv;
                else if (r == kk)
                    v = rowk[c] * pinv;
                else if (c == kk)
                    v = -colk[r] * pinv;
                else
                    v = M[e] - colk[r] * rowk[c] * pinv;
                M[e] = v;
            }
        }
        __syncthreads();
    }
    for (int e = tid; e < n * n; e += kInvertThreads) ainv[e] = (float)M[(e / n) * S + (e % n)];
}

__global__ __launch_bounds__(kBlock) void k_coarse_solve(int n, const float *ainv, const float *b,
                                                         float *x, const double *gate) {
    if (gated(gate)) return;
    __shared__ float bs[kMaxCoarse];
    const int tid = threadIdx.x;
    for (int t = tid; t < n; t += kBlock) bs[t] = b[t];
    __syncthreads();
    for (int r = tid; r < n; r += kBlock) {
        float acc = 0.f;
        for (int c = 0; c < n; c++) acc = fmaf(ainv[r * n + c], bs[c], acc);
        x[r] = acc;
    }
}

// ---------------------------------------------------------------- V-cycle tail
// One block runs the coarse end of the V-cycle (levels of at most
// kTailMaxCells cells) with the levels' arrays in LDS: each sweep covers a
// level's cells with the block's threads and ends at a barrier. The per-cell
// arithmetic is that of k_smooth, k_restrict and k_coarse_solve (neighbour
// values equal to what those take from adjacent lanes), so the result is the
// same bit for bit.
//
// The block is one CU, so its sweeps are bound by instructions and LDS
// bandwidth: a step takes a group of four consecutive cells of a z row per
// thread with 16-byte LDS reads (the index arithmetic once per group), the
// zero-start first iterate c0 D^-1 b is formed once per cell in a sweep of
// its own, the restriction gives each of a coarse cell's eight children to a
// lane of an aligned group of eight (the children's residuals added across
// those lanes in k_restrict's order), and the operators come in with all of
// their loads in flight at once.
constexpr int kTailThreads = 1024;
constexpr int kTailLdsFloats = (int)(kTailLdsBytes / sizeof(float));

// x = c0 D^-1 b on every cell of the level (k_smooth's SM_FIRST iterate).
__device__ void tail_first(const TailLevel &L, float *__restrict__ x, float c0) {
    typedef float f4 __attribute__((ext_vector_type(4)));
    const int n = (int)L.d.cells();
    if ((n & 3) == 0) {
        for (int g = threadIdx.x; 4 * g < n; g += kTailThreads) {
            const f4 dg = *(const f4 *)(L.A.diag + 4 * g), bv = *(const f4 *)(L.b + 4 * g);
            f4 xv;
#pragma unroll
            for (int q = 0; q < 4; q++) xv[q] = c0 * rcp_or_zero(dg[q]) * bv[q];
            *(f4 *)(x + 4 * g) = xv;
        }
    } else {
        for (int q = threadIdx.x; q < n; q += kTailThreads) x[q] = c0 * rcp_or_zero(L.A.diag[q]) * L.b[q];
    }
    __syncthreads();
}

// One Chebyshev step (modes without SM_FIRST: the first iterate is stored by
// tail_first and read with SM_XFIRST). Levels whose z rows hold whole groups
// of four cells take k_smooth_flat's mapping: a thread a group, 16-byte LDS
// reads of the group and of its x and y neighbour groups (a wavefront's
// reads then touch consecutive banks), one value at each z end. Other levels
// take one cell per thread.
template <int MODE>
__device__ void tail_smooth(const TailLevel &L, const float *__restrict__ xin, float *__restrict__ xout,
                            float cd, float cr, const float *__restrict__ e, Dims dc) {
    typedef float f4 __attribute__((ext_vector_type(4)));
    typedef float f2 __attribute__((ext_vector_type(2)));
    static_assert((MODE & SM_FIRST) == 0, "tail step: the first iterate comes from tail_first");
    constexpr bool kProlong = (MODE & SM_PROLONG) != 0;
    constexpr bool kXFirst = (MODE & SM_XFIRST) != 0;
    const Dims d = L.d;
    const Op32 A = L.A;
    const float *__restrict__ b = L.b;
    const int sx = d.ny * d.nz, sy = d.nz;
    const int n = (int)d.cells();
    if ((d.nz & 3) == 0) {
        const f4 z4 = {0.f, 0.f, 0.f, 0.f};
        for (int g = threadIdx.x; 4 * g < n; g += kTailThreads) {
            const int c = 4 * g;
            const int t = c / d.nz, k = c - t * d.nz;
            const int i = t / d.ny, j = t - i * d.ny;
            const bool up = i + 1 < d.nx, ylo = j > 0, yhi = j + 1 < d.ny;
            const bool zl = k > 0, zh = k + 4 < d.nz;  // neighbours below the first, above the last
            const f4 dg = *(const f4 *)(A.diag + c);
            f4 xc = *(const f4 *)(xin + c);
            f4 xm = i > 0 ? *(const f4 *)(xin + c - sx) : z4;
            f4 xp = up ? *(const f4 *)(xin + c + sx) : z4;
            const f4 cxc = *(const f4 *)(A.cpl[0] + c);
            const f4 cxp = up ? *(const f4 *)(A.cpl[0] + c + sx) : z4;
            const f4 cym = ylo ? *(const f4 *)(A.cpl[1] + c) : z4;
            f4 xym = ylo ? *(const f4 *)(xin + c - sy) : z4;
            const f4 cyp = yhi ? *(const f4 *)(A.cpl[1] + c + sy) : z4;
            f4 xyp = yhi ? *(const f4 *)(xin + c + sy) : z4;
            const f4 cz = *(const f4 *)(A.cpl[2] + c);
            float xzlo = zl ? xin[c - 1] : 0.f;
            float xzhi = zh ? xin[c + 4] : 0.f;
            const float czhi = zh ? A.cpl[2][c + 4] : 0.f;
            const f4 bc = *(const f4 *)(b + c);
            f4 dprev = z4;
            if constexpr ((MODE & SM_READ_D) != 0) dprev = *(const f4 *)(L.dv + c);
            const f4 xin_c = xc;
            if constexpr (kProlong) {
                // the group's two coarse cells in this row and in the
                // neighbouring rows and planes (k >> 1 even: 8-byte aligned)
                const int kc = k >> 1;
                auto e2 = [&](int ii, int jj) {
                    return *(const f2 *)(e + ((ii >> 1) * dc.ny + (jj >> 1)) * dc.nz + kc);
                };
                const f2 ec = e2(i, j);
                const f2 e_xm = i > 0 ? e2(i - 1, j) : f2{0.f, 0.f};
                const f2 e_xp = up ? e2(i + 1, j) : f2{0.f, 0.f};
                const f2 e_ym = ylo ? e2(i, j - 1) : f2{0.f, 0.f};
                const f2 e_yp = yhi ? e2(i, j + 1) : f2{0.f, 0.f};
                const int er = ((i >> 1) * dc.ny + (j >> 1)) * dc.nz;
                const float ezlo = zl ? e[er + kc - 1] : 0.f;
                const float ezhi = zh ? e[er + kc + 2] : 0.f;
#pragma unroll
                for (int q = 0; q < 4; q++) {
                    xc[q] = xc[q] + ec[q >> 1];
                    xm[q] = xm[q] + e_xm[q >> 1];
                    xp[q] = xp[q] + e_xp[q >> 1];
                    xym[q] = xym[q] + e_ym[q >> 1];
                    xyp[q] = xyp[q] + e_yp[q >> 1];
                }
                xzlo = xzlo + ezlo;
                xzhi = xzhi + ezhi;
            }
            f4 xo, dno;
#pragma unroll
            for (int q = 0; q < 4; q++) {
                const float xzm = q > 0 ? xc[q - 1] : xzlo;
                const float xzp = q < 3 ? xc[q + 1] : xzhi;
                const float czp = q < 3 ? cz[q + 1] : czhi;
                float off = fmaf(cxc[q], xm[q], cxp[q] * xp[q]);
                if (ylo) off = fmaf(cym[q], xym[q], off);
                if (yhi) off = fmaf(cyp[q], xyp[q], off);
                if (q > 0 || zl) off = fmaf(cz[q], xzm, off);
                if (q < 3 || zh) off = fmaf(czp, xzp, off);
                // centre iterate: the coarse correction only lands on active cells
                const float x0 = (kProlong && !(dg[q] > 0.f)) ? xin_c[q] : xc[q];
                const float r = bc[q] - fmaf(dg[q], x0, -off);
                const float dp = kXFirst ? x0 : dprev[q];
                const float dn = fmaf(cd, dp, cr * rcp_or_zero(dg[q]) * r);
                xo[q] = x0 + dn;
                dno[q] = dn;
            }
            *(f4 *)(xout + c) = xo;
            if constexpr ((MODE & SM_WRITE_D) != 0) *(f4 *)(L.dv + c) = dno;
        }
        __syncthreads();
        return;
    }
    const int nyz = d.ny * d.nz;
    auto xval = [&](int q, int qi, int qj, int qk) -> float {
        if constexpr (kProlong) return xin[q] + e[dc.idx(qi >> 1, qj >> 1, qk >> 1)];
        else return xin[q];
    };
    for (int c = threadIdx.x; c < n; c += kTailThreads) {
        const int i = c / nyz, rem = c - i * nyz, j = rem / d.nz, k = rem - j * d.nz;
        const float dg = A.diag[c];
        const bool up = i + 1 < d.nx, zlo = k > 0, zhi = k + 1 < d.nz;
        const float xm = i > 0 ? xval(c - sx, i - 1, j, k) : 0.f;
        const float xc = xval(c, i, j, k);
        const float cxc = A.cpl[0][c];
        const float xp = up ? xval(c + sx, i + 1, j, k) : 0.f;
        const float cxp = up ? A.cpl[0][c + sx] : 0.f;
        const float czm = A.cpl[2][c];
        const float cym = j > 0 ? A.cpl[1][c] : 0.f;
        const float xym = j > 0 ? xval(c - sy, i, j - 1, k) : 0.f;
        const float cyp = j + 1 < d.ny ? A.cpl[1][c + sy] : 0.f;
        const float xyp = j + 1 < d.ny ? xval(c + sy, i, j + 1, k) : 0.f;
        const float bc = b[c];
        float dprev = 0.f;
        if constexpr ((MODE & SM_READ_D) != 0) dprev = L.dv[c];
        const float xzm = zlo ? xval(c - 1, i, j, k - 1) : 0.f;
        const float xzp = zhi ? xval(c + 1, i, j, k + 1) : 0.f;
        const float czp = zhi ? A.cpl[2][c + 1] : 0.f;
        float x0 = xc;
        if constexpr (kProlong) {
            if (!(dg > 0.f)) x0 = xin[c];
        }
        float off = fmaf(cxc, xm, cxp * xp);
        if (j > 0) off = fmaf(cym, xym, off);
        if (j + 1 < d.ny) off = fmaf(cyp, xyp, off);
        if (zlo) off = fmaf(czm, xzm, off);
        if (zhi) off = fmaf(czp, xzp, off);
        const float r = bc - fmaf(dg, x0, -off);
        if constexpr (kXFirst) dprev = x0;
        const float dn = fmaf(cd, dprev, cr * rcp_or_zero(dg) * r);
        xout[c] = x0 + dn;
        if constexpr ((MODE & SM_WRITE_D) != 0) L.dv[c] = dn;
    }
    __syncthreads();
}

// bc = R (b - A x): k_restrict's residual per child and its pairwise sums (z
// pair, then y pair, then x pair onto 0). Lane 8m + (4 di + 2 dj + dk) of a
// wavefront takes child (di, dj, dk) of one coarse cell; the sums run across
// those lanes (every lane of the group takes part in each exchange).
__device__ void tail_restrict(const TailLevel &F, const float *__restrict__ x, float *__restrict__ bcv,
                              Dims dc) {
    const Dims d = F.d;
    const Op32 A = F.A;
    const Strides s(d);
    const int nyz = dc.ny * dc.nz;
    const int nw = 8 * (int)dc.cells();  // a multiple of 8: whole groups, one wavefront each
    // residual of a child (0 outside the grid or inactive): all loads from
    // clamped addresses, then a select
    auto res = [&](int iu, int ju, int ku) -> float {
        const bool in = iu < d.nx && ju < d.ny && ku < d.nz;
        const int i = min(iu, d.nx - 1), j = min(ju, d.ny - 1), k = min(ku, d.nz - 1);
        const size_t c = d.idx(i, j, k);
        const float dg = A.diag[c];
        const bool up = i + 1 < d.nx, zlo = k > 0, zhi = k + 1 < d.nz;
        const float xc = x[c], bcv = F.b[c];
        const float cxm = A.cpl[0][c], cym = A.cpl[1][c], czm = A.cpl[2][c];
        const float xm = i > 0 ? x[c - s.sx] : 0.f, ym = j > 0 ? x[c - s.sy] : 0.f;
        const float zm = zlo ? x[c - 1] : 0.f;
        const float cxp = up ? A.cpl[0][c + s.sx] : 0.f, xp = up ? x[c + s.sx] : 0.f;
        const bool yup = j + 1 < d.ny;
        const float cyp = yup ? A.cpl[1][c + s.sy] : 0.f, yp = yup ? x[c + s.sy] : 0.f;
        const float czp = zhi ? A.cpl[2][c + 1] : 0.f, zp = zhi ? x[c + 1] : 0.f;
        float off = 0.f;
        if (i > 0) off = fmaf(cxm, xm, off);
        if (up) off = fmaf(cxp, xp, off);
        if (j > 0) off = fmaf(cym, ym, off);
        if (yup) off = fmaf(cyp, yp, off);
        if (zlo) off = fmaf(czm, zm, off);
        if (zhi) off = fmaf(czp, zp, off);
        const float r = bcv - fmaf(dg, xc, -off);
        return (in && dg > 0.f) ? r : 0.f;
    };
    for (int w = threadIdx.x; w < nw; w += kTailThreads) {
        const int t = w >> 3, ch = w & 7;
        const int I = t / nyz, rem = t - I * nyz, J = rem / dc.nz, K = rem - J * dc.nz;
        const int di = ch >> 2, dj = (ch >> 1) & 1, dk = ch & 1;
        const float r = res(2 * I + di, 2 * J + dj, 2 * K + dk);
        const float rz = __shfl_xor(r, 1, 64);                  // the z pair's other child
        const float zp = d.nz > 1 ? (dk ? rz + r : r + rz) : r;  // (r0 + r1 on both lanes)
        const float zy = __shfl_xor(zp, 2, 64);
        const float ay = dj ? zy + zp : zp + zy;  // zp[0] + zp[1]
        const float ax = __shfl_xor(ay, 4, 64);   // the other x layer's sum
        if (ch == 0) {
            float acc = 0.f;
            acc += ay;
            if (2 * I + 1 < d.nx) acc += ax;
            bcv[t] = acc;
        }
    }
    __syncthreads();
}

__global__ __launch_bounds__(kTailThreads) void k_vcycle_tail(TailArgs t, const double *gate) {
    typedef float f4 __attribute__((ext_vector_type(4)));
    if (gated(gate)) return;
    // every array of the tail's levels lives in LDS for the whole cycle (8
    // per level, levels one after another): the operators and the first
    // level's b are copied in, the first level's result is copied out
    __shared__ __attribute__((aligned(16))) float lds[kTailLdsFloats];
    auto level = [&](int l) -> TailLevel {
        size_t off = 0;
        for (int q = 0; q < l; q++) off += 8 * t.lv[q].d.cells();
        const size_t n = t.lv[l].d.cells();
        float *p = lds + off;
        return TailLevel{t.lv[l].d, Op32{p, {p + n, p + 2 * n, p + 3 * n}}, p + 4 * n, p + 5 * n,
                         p + 6 * n, p + 7 * n};
    };
    const int tid = threadIdx.x;
    const int last = t.nlev - 1;
    const int nC = (int)t.lv[last].d.cells();
    // The coarsest solve reads the dense inverse from the first level's
    // direction array, which no sweep uses between that level's last
    // pre-smoothing step and its first post-smoothing step (which writes it
    // before reading): loaded here with the operators, stored there.
    const bool ainv_lds = (int)t.lv[0].d.cells() >= nC * nC && (nC * nC) % 4 == 0 &&
                          ((uintptr_t)t.ainv & 15) == 0;
    f4 ainv4 = {0.f, 0.f, 0.f, 0.f};
    if (ainv_lds && 4 * tid < nC * nC) ainv4 = *(const f4 *)(t.ainv + 4 * tid);
    // Copy-in: where every level has a multiple of 4 cells (16-byte aligned
    // LDS arrays and source arrays), each thread loads at most one 16-byte
    // vector of each array (kTailMaxCells / 4 <= kTailThreads) of the first
    // three levels, all in flight together, then stores them.
    bool vec = true;
    for (int l = 0; l < t.nlev; l++) {
        const TailLevel &G = t.lv[l];
        vec = vec && G.d.cells() % 4 == 0 &&
              (((uintptr_t)G.A.diag | (uintptr_t)G.A.cpl[0] | (uintptr_t)G.A.cpl[1] | (uintptr_t)G.A.cpl[2]) & 15) == 0;
    }
    vec = vec && ((uintptr_t)t.lv[0].b & 15) == 0;
    int lstart = 0;  // levels still to copy by the scalar loop
    if (vec) {
        constexpr int kV = 3;
        f4 v[kV][5];
#pragma unroll
        for (int l = 0; l < kV; l++) {
            const TailLevel &G = t.lv[l < t.nlev ? l : 0];
            const bool on = l < t.nlev && 4 * tid < (int)G.d.cells();
            const f4 z = {0.f, 0.f, 0.f, 0.f};
            v[l][0] = on ? *(const f4 *)(G.A.diag + 4 * tid) : z;
            v[l][1] = on ? *(const f4 *)(G.A.cpl[0] + 4 * tid) : z;
            v[l][2] = on ? *(const f4 *)(G.A.cpl[1] + 4 * tid) : z;
            v[l][3] = on ? *(const f4 *)(G.A.cpl[2] + 4 * tid) : z;
            v[l][4] = (on && l == 0) ? *(const f4 *)(G.b + 4 * tid) : z;
        }
#pragma unroll
        for (int l = 0; l < kV; l++) {
            if (l < t.nlev && 4 * tid < (int)t.lv[l].d.cells()) {
                const TailLevel L = level(l);
                *(f4 *)(const_cast<float *>(L.A.diag) + 4 * tid) = v[l][0];
                *(f4 *)(const_cast<float *>(L.A.cpl[0]) + 4 * tid) = v[l][1];
                *(f4 *)(const_cast<float *>(L.A.cpl[1]) + 4 * tid) = v[l][2];
                *(f4 *)(const_cast<float *>(L.A.cpl[2]) + 4 * tid) = v[l][3];
                if (l == 0) *(f4 *)(L.b + 4 * tid) = v[l][4];
            }
        }
        lstart = kV;
    }
    for (int l = lstart; l < t.nlev; l++) {
        const TailLevel L = level(l);
        const TailLevel &G = t.lv[l];
        for (int q = tid; q < (int)G.d.cells(); q += kTailThreads) {
            const_cast<float *>(L.A.diag)[q] = G.A.diag[q];
            const_cast<float *>(L.A.cpl[0])[q] = G.A.cpl[0][q];
            const_cast<float *>(L.A.cpl[1])[q] = G.A.cpl[1][q];
            const_cast<float *>(L.A.cpl[2])[q] = G.A.cpl[2][q];
            if (l == 0) L.b[q] = G.b[q];
        }
    }
    __syncthreads();
    unsigned in_x1 = 0;  // bit l: level l's pre-smoothed iterate is in x1
    // descend: pre-smoothing and restriction (as the separate launches);
    // coefficient set 0 on the tail's first level, set 1 below it
    for (int l = 0; l < last; l++) {
        const TailLevel L = level(l);
        const int k = l == 0 ? 0 : 1;
        const int deg = t.deg[k];
        const float *cd = t.cd[k], *cr = t.cr[k];
        float *c = L.x0, *o = L.x1;
        const Dims dc = t.lv[l + 1].d;
        tail_first(L, c, cr[0]);
        if (deg > 1) {
            if (deg > 2)
                tail_smooth<SM_XFIRST | SM_WRITE_D>(L, c, o, cd[1], cr[1], nullptr, dc);
            else
                tail_smooth<SM_XFIRST>(L, c, o, cd[1], cr[1], nullptr, dc);
            float *tmp = c;
            c = o;
            o = tmp;
            for (int s = 2; s < deg; s++) {
                if (s < deg - 1)
                    tail_smooth<SM_READ_D | SM_WRITE_D>(L, c, o, cd[s], cr[s], nullptr, dc);
                else
                    tail_smooth<SM_READ_D>(L, c, o, cd[s], cr[s], nullptr, dc);
                tmp = c;
                c = o;
                o = tmp;
            }
        }
        tail_restrict(L, c, level(l + 1).b, dc);
        if (c == L.x1) in_x1 |= 1u << l;
    }
    // coarsest: dense solve (k_coarse_solve's order)
    const TailLevel C = level(last);
    {
        if (ainv_lds) {
            // stored transposed (column c at a + c nC), so the row threads
            // read consecutive words
            float *a = level(0).dv;
            if (4 * tid < nC * nC) {
                const int r = (4 * tid) / nC, c0 = 4 * tid - r * nC;
#pragma unroll
                for (int q = 0; q < 4; q++) a[(c0 + q) * nC + r] = ainv4[q];
            }
            __syncthreads();
            for (int r = tid; r < nC; r += kTailThreads) {
                float acc = 0.f;
                for (int c = 0; c < nC; c++) acc = fmaf(a[c * nC + r], C.b[c], acc);
                C.x0[r] = acc;
            }
        } else {
            for (int r = tid; r < nC; r += kTailThreads) {
                float acc = 0.f;
                for (int c = 0; c < nC; c++) acc = fmaf(t.ainv[r * nC + c], C.b[c], acc);
                C.x0[r] = acc;
            }
        }
        __syncthreads();
    }
    // ascend: prolongation fused into the first post-smoothing step
    const float *e = C.x0;
    for (int l = last - 1; l >= 0; l--) {
        const TailLevel L = level(l);
        const int k = l == 0 ? 0 : 1;
        const int deg = t.deg[k];
        const float *cd = t.cd[k], *cr = t.cr[k];
        const Dims dc = t.lv[l + 1].d;
        float *c = (in_x1 >> l) & 1u ? L.x1 : L.x0;
        float *o = c == L.x0 ? L.x1 : L.x0;
        for (int s = 0; s < deg; s++) {
            const bool wd = s < deg - 1;
            if (s == 0) {
                if (wd)
                    tail_smooth<SM_PROLONG | SM_WRITE_D>(L, c, o, 0.f, cr[0], e, dc);
                else
                    tail_smooth<SM_PROLONG>(L, c, o, 0.f, cr[0], e, dc);
            } else {
                if (wd)
                    tail_smooth<SM_READ_D | SM_WRITE_D>(L, c, o, cd[s], cr[s], nullptr, dc);
                else
                    tail_smooth<SM_READ_D>(L, c, o, cd[s], cr[s], nullptr, dc);
            }
            float *tmp = c;
            c = o;
            o = tmp;
        }
        e = c;
    }
    // the first level's result (in its x1, where the level-by-level cycle leaves it)
    for (int q = tid; q < (int)t.lv[0].d.cells(); q += kTailThreads) t.lv[0].x1[q] = e[q];
}

}  // namespace

void launch_smooth_first(hipStream_t st, Dims d, Op32 A, const double *r64, float *b, float *x,
                         float c0, const double *gate) {
    Launch L = cell_launch(d);
    if (r64)
        hipLaunchKernelGGL(k_smooth_first<true>, L.grid, L.block, 0, st, L.cg, d, A, r64, b, x, c0, gate);
    else
        hipLaunchKernelGGL(k_smooth_first<false>, L.grid, L.block, 0, st, L.cg, d, A, r64, b, x, c0, gate);
    SFP_LAUNCH_CHECK();
}

bool flat_step(Dims d) { return d.nz % 4 == 0 && d.cells() < (size_t(1) << 32); }

// The flat step where it applies: no fused first step, z rows of whole
// 4-cell groups, 16-byte aligned arrays (8-byte aligned coarse values). On
// MI355X 12 % faster than x-plane tiles at 256^3, 17 % at 128^3.
static bool smooth_flat(Dims d, int mode, const double *r64, const float *ecoarse,
                        std::initializer_list<const float *> arrays) {
    if ((mode & SM_FIRST) || r64 || !flat_step(d)) return false;
    if ((mode & SM_PROLONG) && (uintptr_t)ecoarse % 8) return false;
    for (const float *a : arrays)
        if ((uintptr_t)a % 16) return false;
    return true;
}

void launch_smooth(hipStream_t st, int mode, Dims d, Op32 A, const double *r64, float *b,
                   const float *xin, float *xout, float *dvec, float c0, float cd, float cr,
                   const float *ecoarse, const double *gate) {
    if (smooth_flat(d, mode, r64, ecoarse, {A.diag, A.cpl[0], A.cpl[1], A.cpl[2], b, xin, xout, dvec})) {
        const unsigned nb = (unsigned)((d.cells() + kSmoothFlatCells - 1) / kSmoothFlatCells);
        const Dims dc = d.coarse();
#define SFP_FLAT_CASE(M)                                                                           \
    case M:                                                                                        \
        if (A.pinned_rowsum)                                                                       \
            hipLaunchKernelGGL((k_smooth_flat<M, true>), dim3(nb), dim3(kBlock), 0, st, d, A, b,   \
                               xin, xout, dvec, cd, cr, ecoarse, dc, gate);                        \
        else                                                                                       \
            hipLaunchKernelGGL((k_smooth_flat<M, false>), dim3(nb), dim3(kBlock), 0, st, d, A, b,  \
                               xin, xout, dvec, cd, cr, ecoarse, dc, gate);                        \
        break;
        switch (mode) {
            SFP_FLAT_CASE(SM_PROLONG)
            SFP_FLAT_CASE(SM_PROLONG | SM_WRITE_D)
            SFP_FLAT_CASE(SM_XFIRST)
            SFP_FLAT_CASE(SM_XFIRST | SM_WRITE_D)
            SFP_FLAT_CASE(SM_READ_D)
            SFP_FLAT_CASE(SM_READ_D | SM_WRITE_D)
            SFP_FLAT_CASE(0)
            SFP_FLAT_CASE(SM_WRITE_D)
            default:
                throw ArgError("unsupported smoother mode");
        }
#undef SFP_FLAT_CASE
        SFP_LAUNCH_CHECK();
        return;
    }
    // 8 planes per block measured best for the shuffle-based step (256^3:
    // 122 us vs 124 us at 16 and 123 us at 4)
    Launch L = cell_launch(d, 0, 8);
    Dims dc = d.coarse();
#define SFP_SMOOTH_CASE(M)                                                                    \
    case M:                                                                                   \
        hipLaunchKernelGGL(k_smooth<M>, L.grid, L.block, 0, st, L.cg, d, A, r64, b, xin, xout, \
                           dvec, c0, cd, cr, ecoarse, dc, gate);                              \
        break;
    switch (mode) {
        SFP_SMOOTH_CASE(SM_FIRST)
        SFP_SMOOTH_CASE(SM_FIRST | SM_WRITE_D)
        SFP_SMOOTH_CASE(SM_XFIRST)
        SFP_SMOOTH_CASE(SM_XFIRST | SM_WRITE_D)
        SFP_SMOOTH_CASE(SM_READ_D)
        SFP_SMOOTH_CASE(SM_READ_D | SM_WRITE_D)
        SFP_SMOOTH_CASE(SM_PROLONG)
        SFP_SMOOTH_CASE(SM_PROLONG | SM_WRITE_D)
        SFP_SMOOTH_CASE(0)
        SFP_SMOOTH_CASE(SM_WRITE_D)
        default:
            throw ArgError("unsupported smoother mode");
    }
#undef SFP_SMOOTH_CASE
    SFP_LAUNCH_CHECK();
}

Launch pair_launch(Dims d, bool even_planes) {
    Launch L;
    L.cg.gx = (d.nz + kPairZ - 1) / kPairZ;
    L.cg.gy = (d.ny + kPairY - 1) / kPairY;
    const long long tiles = (long long)L.cg.gx * L.cg.gy * d.nx;
    L.cg.planes = (int)std::min<long long>(16, std::max<long long>(1, tiles / 1024));
    if (even_planes && (L.cg.planes & 1)) L.cg.planes++;  // x pairs stay in one block
    L.block = dim3(64, kPairRows, 1);
    L.grid = dim3(L.cg.gx * L.cg.gy * ((d.nx + L.cg.planes - 1) / L.cg.planes), 1, 1);
    return L;
}

bool row_pair(Dims d) { return d.nz % 4 == 0 && d.nz > 192 && d.nz <= 256 && d.cells() < (size_t(1) << 32); }

// Compute units of the current device (cached per device).
static int device_cus() {
    static int cus[64] = {0};
    int dev = 0;
    SFP_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= 64) return 256;
    if (!cus[dev]) {
        int v = 0;
        SFP_HIP(hipDeviceGetAttribute(&v, hipDeviceAttributeMultiprocessorCount, dev));
        cus[dev] = v > 0 ? v : 256;
    }
    return cus[dev];
}

// rowact[i ny + j] = 1 when the z row (i, j) holds an active cell: one
// wavefront per row.
__global__ __launch_bounds__(kBlock) void k_row_activity(Dims d, const float *__restrict__ diag32,
                                                         unsigned char *__restrict__ act) {
    const size_t row = ((size_t)blockIdx.x * kBlock + threadIdx.x) / 64;
    const int lane = threadIdx.x & 63;
    if (row >= (size_t)d.nx * d.ny) return;  // uniform over the wavefront
    const size_t c = row * (size_t)d.nz;
    bool a = false;
    for (int k = lane; k < d.nz; k += 64) a = a || diag32[c + k] > 0.f;
    const unsigned long long m = __ballot(a);
    if (lane == 0) act[row] = m != 0ull ? 1 : 0;
}

void launch_row_activity(hipStream_t st, Dims d, const float *diag32, unsigned char *act) {
    const size_t threads = (size_t)d.nx * d.ny * 64;
    hipLaunchKernelGGL(k_row_activity, dim3((unsigned)((threads + kBlock - 1) / kBlock)), dim3(kBlock), 0, st,
                       d, diag32, act);
    SFP_LAUNCH_CHECK();
}

void launch_smooth_pair(hipStream_t st, Dims d, Op32 A, const float *b, const float *x1, float *x3,
                        const float *d1, float cdA, float crA, float cdB, float crB,
                        const double *gate, const unsigned char *rowact) {
    if (x1 == x3) throw ArgError("launch_smooth_pair: x3 aliases x1");
    bool rows = row_pair(d);
    for (const float *a : {A.diag, A.cpl[0], A.cpl[1], A.cpl[2], b, x1, (const float *)x3, d1})
        rows = rows && (uintptr_t)a % 16 == 0;
    if (rows) {
        // one block per compute unit, in one round: as many x runs per tile
        // column as the columns leave CUs for (at most 56 planes a run, the
        // activity masks' reach)
        const unsigned ytiles = (unsigned)((d.ny + kRowPairY - 1) / kRowPairY);
        const int runs_max = std::max(1, device_cus() / (int)ytiles);
        int planes = std::min(56, std::max(1, (d.nx + runs_max - 1) / runs_max));
        const unsigned runs = (unsigned)((d.nx + planes - 1) / planes);
        const dim3 grid(ytiles * runs), block(64, kRowPairRows);
        const bool dx = d1 == x1;
        const unsigned char *ra = A.pinned_rowsum ? nullptr : rowact;
#define SFP_ROWPAIR(RS, DX)                                                                      \
    hipLaunchKernelGGL((k_smooth_pair_rows<RS, DX>), grid, block, 0, st, d, A, b, x1, x3, d1, cdA, \
                       crA, cdB, crB, ra, planes, ytiles, gate)
        if (A.pinned_rowsum) {
            if (dx) SFP_ROWPAIR(true, true); else SFP_ROWPAIR(true, false);
        } else {
            if (dx) SFP_ROWPAIR(false, true); else SFP_ROWPAIR(false, false);
        }
#undef SFP_ROWPAIR
        SFP_LAUNCH_CHECK();
        return;
    }
    const Launch L = pair_launch(d, false);
    if (A.pinned_rowsum)
        hipLaunchKernelGGL((k_smooth_pair<false, true>), L.grid, L.block, 0, st, L.cg, d, A, b, x1, x3,
                           d1, cdA, crA, cdB, crB, Dims{}, nullptr, gate);
    else
        hipLaunchKernelGGL((k_smooth_pair<false, false>), L.grid, L.block, 0, st, L.cg, d, A, b, x1, x3,
                           d1, cdA, crA, cdB, crB, Dims{}, nullptr, gate);
    SFP_LAUNCH_CHECK();
}

void launch_smooth_restrict(hipStream_t st, Dims d, Op32 A, const float *b, const float *x1,
                            float *x2, const float *d1, float cd, float cr, float *bc,
                            const double *gate) {
    if (x1 == x2) throw ArgError("launch_smooth_restrict: x2 aliases x1");
    const Launch L = pair_launch(d, true);
    if (A.pinned_rowsum)
        hipLaunchKernelGGL((k_smooth_pair<true, true>), L.grid, L.block, 0, st, L.cg, d, A, b, x1, x2,
                           d1, cd, cr, 0.f, 0.f, d.coarse(), bc, gate);
    else
        hipLaunchKernelGGL((k_smooth_pair<true, false>), L.grid, L.block, 0, st, L.cg, d, A, b, x1, x2,
                           d1, cd, cr, 0.f, 0.f, d.coarse(), bc, gate);
    SFP_LAUNCH_CHECK();
}

void launch_prolong_add(hipStream_t st, Dims d, const float *diag32, const float *ecoarse, float *x,
                        const double *gate) {
    Launch L = cell_launch(d, 0, 8);
    hipLaunchKernelGGL(k_prolong_add, L.grid, L.block, 0, st, L.cg, d, diag32, ecoarse, d.coarse(), x, gate);
    SFP_LAUNCH_CHECK();
}

void launch_clear_inactive(hipStream_t st, const ClearArgs &a) {
    if (a.nlev < 1 || a.nlev > kClearMaxLevels) throw ArgError("launch_clear_inactive: bad level count");
    const unsigned groups = a.start[a.nlev];
    const unsigned nb = (groups + kBlock - 1) / kBlock;
    hipLaunchKernelGGL(k_clear_inactive, dim3(nb), dim3(kBlock), 0, st, a);
    SFP_LAUNCH_CHECK();
}

void launch_restrict(hipStream_t st, Dims df, Op32 A, const float *b, const float *x, float *bc,
                     const double *gate, float *x0c, const float *diag32c, float c0) {
    if (x0c && !diag32c) throw ArgError("launch_restrict: x0c needs the coarse diagonal");
    Dims dc = df.coarse();
    bool flat = flat_step(df) && (uintptr_t)bc % 8 == 0 && (uintptr_t)x0c % 8 == 0 &&
                (uintptr_t)diag32c % 8 == 0;
    for (const float *a : {A.diag, A.cpl[0], A.cpl[1], A.cpl[2], b, x}) flat = flat && (uintptr_t)a % 16 == 0;
    if (flat) {
        const size_t waves = (size_t)((df.nz + kRestrictFlatZ - 1) / kRestrictFlatZ) *
                             ((df.ny + 1) / 2) * ((df.nx + 1) / 2);
        const unsigned nb = (unsigned)((waves * 64 + kBlock - 1) / kBlock);
        if (A.pinned_rowsum)
            hipLaunchKernelGGL(k_restrict_flat<true>, dim3(nb), dim3(kBlock), 0, st, df, A, b, x, dc,
                               bc, x0c, diag32c, c0, gate);
        else
            hipLaunchKernelGGL(k_restrict_flat<false>, dim3(nb), dim3(kBlock), 0, st, df, A, b, x, dc,
                               bc, x0c, diag32c, c0, gate);
        SFP_LAUNCH_CHECK();
        return;
    }
    Launch L = cell_launch(df, 0, 8);
    if (L.cg.planes & 1) L = cell_launch(df, L.cg.planes + 1);  // plane pairs stay in one block
    hipLaunchKernelGGL(k_restrict, L.grid, L.block, 0, st, L.cg, df, A, b, x, dc, bc, x0c, diag32c, c0,
                       gate);
    SFP_LAUNCH_CHECK();
}

void launch_coarsen(hipStream_t st, Dims df, Op64 Af, const float *diag32f, Dims dc,
                    double scale, double rs_scale, double *diag64c, double *const cpl64c[3],
                    float *diag32c, float *const cpl32c[3]) {
    MutFaces m{{cpl64c[0], cpl64c[1], cpl64c[2]}};
    bool flat = flat_step(df) && (uintptr_t)diag32f % 16 == 0;
    for (const double *a : {Af.diag, Af.cpl[0], Af.cpl[1], Af.cpl[2]}) flat = flat && (uintptr_t)a % 16 == 0;
    if (flat) {
        const size_t waves = (size_t)((df.nz + 63) / 64) * ((df.ny + 1) / 2) * ((df.nx + 1) / 2);
        const unsigned nb = (unsigned)((waves * 64 + kBlock - 1) / kBlock);
        hipLaunchKernelGGL(k_coarsen_flat, dim3(nb), dim3(kBlock), 0, st, df, Af, diag32f, dc, scale, rs_scale,
                           diag64c, m, diag32c, cpl32c[0], cpl32c[1], cpl32c[2]);
        SFP_LAUNCH_CHECK();
        return;
    }
    // over the fine grid, with even plane pairs per block (as the restriction)
    Launch L = cell_launch(df, 0, 8);
    if (L.cg.planes & 1) L = cell_launch(df, L.cg.planes + 1);
    hipLaunchKernelGGL(k_coarsen_fine, L.grid, L.block, 0, st, L.cg, df, Af, diag32f, dc, scale, rs_scale,
                       diag64c, m, diag32c, cpl32c[0], cpl32c[1], cpl32c[2]);
    SFP_LAUNCH_CHECK();
}

void launch_coarse_invert(hipStream_t st, Dims d, Op64 A, float *ainv) {
    if ((int)d.cells() > kMaxCoarse) throw ArgError("coarsest level too large for the dense solve");
    hipLaunchKernelGGL(k_coarse_invert, dim3(1), dim3(kInvertThreads), 0, st, d, A, ainv);
    SFP_LAUNCH_CHECK();
}

void launch_coarse_solve(hipStream_t st, int n, const float *ainv, const float *b, float *x,
                         const double *gate) {
    if (n > kMaxCoarse) throw ArgError("coarsest level too large for the dense solve");
    hipLaunchKernelGGL(k_coarse_solve, dim3(1), dim3(kBlock), 0, st, n, ainv, b, x, gate);
    SFP_LAUNCH_CHECK();
}

void launch_vcycle_tail(hipStream_t st, const TailArgs &t, const double *gate) {
    if (t.nlev < 2 || t.nlev > kTailMaxLevels) throw ArgError("vcycle tail: bad level count");
    for (int k = 0; k < 2; k++)
        if (t.deg[k] < 1 || t.deg[k] > 8) throw ArgError("vcycle tail: bad degree");
    if (t.lv[0].d.cells() > kTailMaxCells) throw ArgError("vcycle tail: level too large");
    if ((int)t.lv[t.nlev - 1].d.cells() > kMaxCoarse) throw ArgError("vcycle tail: coarsest too large");
    if (tail_lds_bytes(t) > kTailLdsBytes) throw ArgError("vcycle tail: levels exceed the LDS");
    hipLaunchKernelGGL(k_vcycle_tail, dim3(1), dim3(kTailThreads), 0, st, t, gate);
    SFP_LAUNCH_CHECK();
}

}  // namespace sfp
