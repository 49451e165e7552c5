// FP64 kernels of the outer flexible CG (Notay's FCG with one stored
// direction): residual, matvec fused with its inner products, and the fused
// vector update. Each reduction writes per-block partial sums; a one-block
// finalize kernel adds them in block order (deterministic) and runs the
// scalar recurrences on the device, so the host only reads ||r||^2.
#include "device_util.hpp"
#include "kernels.hpp"
#include "solver.hpp"

#include <algorithm>

namespace sfp {
namespace {

enum FinalizeOp : int { FIN_RESIDUAL = 0, FIN_MATVEC = 1, FIN_UPDATE = 2, FIN_COUNT = 3 };

// true when a batched solve has already converged (see KrylovScalars::stop)
__device__ __forceinline__ bool gated(const double *gate) { return gate && *gate != 0.0; }

// Streams that no other thread of the launch reads again (FCG vectors) go
// through non-temporal loads and stores: measured 5-10 % faster on MI355X
// for the matvec and the update than cached accesses.
template <class T>
__device__ __forceinline__ T ld_stream(const T *a) {
    return __builtin_nontemporal_load(a);
}
template <class T>
__device__ __forceinline__ void st_stream(T *a, T v) {
    __builtin_nontemporal_store(v, a);
}
typedef double double2v __attribute__((ext_vector_type(2)));
typedef float float2v __attribute__((ext_vector_type(2)));

// The stencil kernels below march over the block's x planes with the
// x-neighbours in registers and take the z-neighbours (value and z coupling)
// from the adjacent lanes of the tile row; only the row's end lanes load
// them. The off-diagonal sum keeps the order x-, x+, y-, y+, z-, z+.
__global__ __launch_bounds__(kBlock) void k_residual64(CellGrid cg, Dims d, Op64 A,
                                                       const double *__restrict__ b,
                                                       const double *__restrict__ x,
                                                       double *__restrict__ r,
                                                       double *__restrict__ partials,
                                                       float *__restrict__ b32,
                                                       float *__restrict__ x0,
                                                       const float *__restrict__ diag32,
                                                       float c0) {
    int k, j, i0;
    cell_tile(cg, k, j, i0);
    const Strides s(d);
    const int tz = blockDim.x, lane = threadIdx.x;
    const bool on = k < d.nz && j < d.ny && i0 < d.nx;
    const bool zlo = k > 0, zhi = k + 1 < d.nz;
    const int ie = min(i0 + cg.planes, d.nx);
    double v[2] = {0.0, 0.0};  // r.r, b.b
    size_t c = on ? d.idx(i0, j, k) : 0;
    // the FP32 copy of r and the cycle's first iterate (the FCG update's arithmetic)
    auto store32 = [&](size_t q, double rq) {
        if (!b32) return;
        const float bf = (float)rq;
        b32[q] = bf;
        if (x0) x0[q] = c0 * rcp_or_zero(diag32[q]) * bf;
    };
    if (!x) {
        if (on)
            for (int i = i0; i < ie; i++, c += s.sx) {
                const double bc = b[c];
                if (r) r[c] = bc;
                store32(c, bc);
                v[0] += bc * bc;
                v[1] += bc * bc;
            }
        block_partials<2>(v, partials);
        return;
    }
    double xm = (on && i0 > 0) ? x[c - s.sx] : 0.0;
    double xc = on ? x[c] : 0.0;
    double cxc = on ? A.cpl[0][c] : 0.0;
    for (int i = i0; i < ie; i++, c += s.sx) {
        const bool up = i + 1 < d.nx;
        const double xp = (on && up) ? x[c + s.sx] : 0.0;
        const double cxp = (on && up) ? A.cpl[0][c + s.sx] : 0.0;
        const double czm = on ? A.cpl[2][c] : 0.0;
        double xzm = __shfl_up(xc, 1, tz);
        double xzp = __shfl_down(xc, 1, tz);
        double czp = __shfl_down(czm, 1, tz);
        if (lane == 0) xzm = (on && zlo) ? x[c - 1] : 0.0;
        if (lane == tz - 1) {
            xzp = (on && zhi) ? x[c + 1] : 0.0;
            czp = (on && zhi) ? A.cpl[2][c + 1] : 0.0;
        }
        if (on) {
            double off = 0.0;
            if (i > 0) off += cxc * xm;
            if (up) off += cxp * xp;
            if (j > 0) off += A.cpl[1][c] * x[c - s.sy];
            if (j + 1 < d.ny) off += A.cpl[1][c + s.sy] * x[c + s.sy];
            if (zlo) off += czm * xzm;
            if (zhi) off += czp * xzp;
            const double bc = b[c];
            const double rc = bc - (A.diag[c] * xc - off);
            if (r) r[c] = rc;
            store32(c, rc);
            v[0] += rc * rc;
            v[1] += bc * bc;
        }
        xm = xc;
        xc = xp;
        cxc = cxp;
    }
    block_partials<2>(v, partials);
}

// The same residual over the flat cell index, two consecutive cells of one z
// row per lane (grids with nz even), the mapping of k_fcg_matvec_flat: 16-byte
// loads of the FP64 operator and vectors at the pair and at its x and y
// neighbour pairs, one extra x value and z coupling at each end. Each cell's
// off-diagonal sum keeps k_residual64's order, so r is the same to the bit; a
// row-sum diagonal is rebuilt from the couplings as in the matvec. r, and b32
// with x0, may be null (then not stored: a convergence check needs only the
// sums).
constexpr unsigned kResidualCells = 2 * kBlock;

__global__ __launch_bounds__(kBlock) void k_residual64_flat(Dims d, Op64 A,
                                                            const double *__restrict__ b,
                                                            const double *__restrict__ x,
                                                            double *__restrict__ r,
                                                            double *__restrict__ partials,
                                                            float *__restrict__ b32,
                                                            float *__restrict__ x0,
                                                            const float *__restrict__ diag32,
                                                            float c0) {
    const Strides s(d);
    const size_t c = (size_t)blockIdx.x * kResidualCells + 2 * (size_t)flat_tid();
    double v[2] = {0.0, 0.0};  // r.r, b.b
    if (c < d.cells()) {       // cells() is even: then c + 1 is a cell too
        const double2v d0 = {0.0, 0.0};
        const double2v bc = ld_stream((const double2v *)(b + c));
        double2v rc = bc;
        if (x) {
            const unsigned cu = (unsigned)c;
            const unsigned t = cu / (unsigned)d.nz;
            const int k = (int)(cu - t * (unsigned)d.nz);
            const unsigned iu = t / (unsigned)d.ny;
            const int i = (int)iu, j = (int)(t - iu * (unsigned)d.ny);
            const bool xl = i > 0, xh = i + 1 < d.nx, yl = j > 0, yh = j + 1 < d.ny;
            const bool zl = k > 0, zh = k + 2 < d.nz;
            const double2v xc = *(const double2v *)(x + c);
            const double2v xxm = xl ? *(const double2v *)(x + c - s.sx) : d0;
            const double2v xxp = xh ? *(const double2v *)(x + c + s.sx) : d0;
            const double2v xym = yl ? *(const double2v *)(x + c - s.sy) : d0;
            const double2v xyp = yh ? *(const double2v *)(x + c + s.sy) : d0;
            const double xzm = zl ? x[c - 1] : 0.0;
            const double xzp = zh ? x[c + 2] : 0.0;
            const double2v cxm = xl ? *(const double2v *)(A.cpl[0] + c) : d0;
            const double2v cxp = xh ? *(const double2v *)(A.cpl[0] + c + s.sx) : d0;
            const double2v cym = yl ? *(const double2v *)(A.cpl[1] + c) : d0;
            const double2v cyp = yh ? *(const double2v *)(A.cpl[1] + c + s.sy) : d0;
            const double2v cz = *(const double2v *)(A.cpl[2] + c);
            const double czp = zh ? A.cpl[2][c + 2] : 0.0;
            double2v dg;
            if (A.pinned_rowsum && cu != 0 && cu != (unsigned)d.nz && cu != (unsigned)s.sx) {
                dg.x = ((((cxm.x + cxp.x) + cym.x) + cyp.x) + cz.x) + cz.y;
                dg.y = ((((cxm.y + cxp.y) + cym.y) + cyp.y) + cz.y) + czp;
            } else {
                dg = *(const double2v *)(A.diag + c);
            }
            double o0 = 0.0, o1 = 0.0;
            if (xl) {
                o0 += cxm.x * xxm.x;
                o1 += cxm.y * xxm.y;
            }
            if (xh) {
                o0 += cxp.x * xxp.x;
                o1 += cxp.y * xxp.y;
            }
            if (yl) {
                o0 += cym.x * xym.x;
                o1 += cym.y * xym.y;
            }
            if (yh) {
                o0 += cyp.x * xyp.x;
                o1 += cyp.y * xyp.y;
            }
            if (zl) o0 += cz.x * xzm;
            o0 += cz.y * xc.y;
            o1 += cz.y * xc.x;
            if (zh) o1 += czp * xzp;
            rc.x = bc.x - (dg.x * xc.x - o0);
            rc.y = bc.y - (dg.y * xc.y - o1);
        }
        if (r) st_stream((double2v *)(r + c), rc);
        if (b32) {
            const float2v bf = {(float)rc.x, (float)rc.y};
            *(float2v *)(b32 + c) = bf;
            if (x0) {
                const float2v dd = *(const float2v *)(diag32 + c);
                const float2v xf = {c0 * rcp_or_zero(dd.x) * bf.x, c0 * rcp_or_zero(dd.y) * bf.y};
                *(float2v *)(x0 + c) = xf;
            }
        }
        v[0] = rc.x * rc.x + rc.y * rc.y;
        v[1] = bc.x * bc.x + bc.y * bc.y;
    }
    block_partials<2>(v, partials);
}

// `active` (the FP32 diagonal, 0 = inactive cell) or null: with it, a
// wavefront whose cells are all inactive is skipped (z is 0 there and the
// vectors it would write are never read by active cells). The flag of the
// next plane is loaded one plane ahead, and every load of a plane is issued
// before the lane exchange.
__global__ __launch_bounds__(kBlock) void k_fcg_matvec(CellGrid cg, Dims d, Op64 A,
                                                       const float *__restrict__ active,
                                                       const float *__restrict__ z,
                                                       const double *__restrict__ r,
                                                       const double *__restrict__ qprev,
                                                       double *__restrict__ w, int first,
                                                       double *__restrict__ partials,
                                                       const double *gate) {
    if (gated(gate)) return;
    int k, j, i0;
    cell_tile(cg, k, j, i0);
    const Strides s(d);
    const int tz = blockDim.x, lane = threadIdx.x;
    const bool on = k < d.nz && j < d.ny && i0 < d.nx;
    const bool zlo = k > 0, zhi = k + 1 < d.nz, ylo = on && j > 0, yhi = on && j + 1 < d.ny;
    const int ie = min(i0 + cg.planes, d.nx);
    double v[3] = {0.0, 0.0, 0.0};  // z.w, z.q_prev, z.r
    size_t c = on ? d.idx(i0, j, k) : 0;
    float zm = 0.f, zc = 0.f;
    double cxc = 0.0;
    bool primed = false;  // zm, zc, cxc hold this plane's values
    float act_next = (active && on) ? active[c] : 0.f;
    for (int i = i0; i < ie; i++, c += s.sx) {
        if (active) {
            const float act = act_next;
            act_next = (on && i + 1 < ie) ? active[c + s.sx] : 0.f;
            if (!__any(on && act > 0.f)) {
                primed = false;
                continue;
            }
        }
        if (!primed) {
            zm = (on && i > 0) ? z[c - s.sx] : 0.f;
            zc = on ? z[c] : 0.f;
            cxc = on ? A.cpl[0][c] : 0.0;
            primed = true;
        }
        const bool up = i + 1 < d.nx;
        const float zp = (on && up) ? z[c + s.sx] : 0.f;
        const double cxp = (on && up) ? A.cpl[0][c + s.sx] : 0.0;
        const double czm = on ? A.cpl[2][c] : 0.0;
        const double cym = ylo ? A.cpl[1][c] : 0.0;
        const double cyp = yhi ? A.cpl[1][c + s.sy] : 0.0;
        const float zym = ylo ? z[c - s.sy] : 0.f;
        const float zyp = yhi ? z[c + s.sy] : 0.f;
        const double dg = on ? A.diag[c] : 0.0;
        const double qp = (on && !first) ? ld_stream(qprev + c) : 0.0;
        const double rc = on ? ld_stream(r + c) : 0.0;
        float zzm = __shfl_up(zc, 1, tz);
        float zzp = __shfl_down(zc, 1, tz);
        double czp = __shfl_down(czm, 1, tz);
        if (lane == 0) zzm = (on && zlo) ? z[c - 1] : 0.f;
        if (lane == tz - 1) {
            zzp = (on && zhi) ? z[c + 1] : 0.f;
            czp = (on && zhi) ? A.cpl[2][c + 1] : 0.0;
        }
        if (on) {
            double off = 0.0;
            if (i > 0) off += cxc * (double)zm;
            if (up) off += cxp * (double)zp;
            if (j > 0) off += cym * (double)zym;
            if (j + 1 < d.ny) off += cyp * (double)zyp;
            if (zlo) off += czm * (double)zzm;
            if (zhi) off += czp * (double)zzp;
            const double zd = (double)zc;
            const double wc = dg * zd - off;
            st_stream(w + c, wc);
            v[0] += zd * wc;
            if (!first) v[1] += zd * qp;
            v[2] += zd * rc;
        }
        zm = zc;
        zc = zp;
        cxc = cxp;
    }
    block_partials<3>(v, partials);
}

// The same matvec over the flat cell index, two consecutive cells of one z row
// per lane (grids with nz even): 16-byte loads of the FP64 operator and
// vectors, 8-byte loads of z and its x and y neighbours, one extra z value and
// z coupling at each end of the pair. Every cell's w and dot terms are formed
// in the tiled kernel's order, so w is the same to the bit. A wavefront whose
// 128 cells are all inactive is skipped.
constexpr unsigned kMatvecCells = 2 * kBlock;

__global__ __launch_bounds__(kBlock) void k_fcg_matvec_flat(Dims d, Op64 A,
                                                            const float *__restrict__ active,
                                                            const float *__restrict__ z,
                                                            const double *__restrict__ r,
                                                            const double *__restrict__ qprev,
                                                            double *__restrict__ w, int first,
                                                            double *__restrict__ partials,
                                                            const double *gate) {
    if (gated(gate)) return;
    const Strides s(d);
    const size_t c = (size_t)blockIdx.x * kMatvecCells + 2 * (size_t)flat_tid();
    const bool on = c < d.cells();  // cells() is even: then c + 1 is a cell too
    double v[3] = {0.0, 0.0, 0.0};  // z.w, z.q_prev, z.r
    bool act = on;
    if (active && on) {
        const float2v a = *(const float2v *)(active + c);
        act = a.x > 0.f || a.y > 0.f;
    }
    if (__any(act) && on) {
        // (i, j, k) of the first cell (the launcher keeps cells() < 2^32)
        const unsigned cu = (unsigned)c;
        const unsigned t = cu / (unsigned)d.nz;
        const int k = (int)(cu - t * (unsigned)d.nz);
        const unsigned iu = t / (unsigned)d.ny;
        const int i = (int)iu, j = (int)(t - iu * (unsigned)d.ny);
        const bool xl = i > 0, xh = i + 1 < d.nx, yl = j > 0, yh = j + 1 < d.ny;
        const bool zl = k > 0, zh = k + 2 < d.nz;  // neighbours below the first, above the second
        const float2v f0 = {0.f, 0.f};
        const double2v d0 = {0.0, 0.0};
        const float2v zc = *(const float2v *)(z + c);
        const float2v zxm = xl ? *(const float2v *)(z + c - s.sx) : f0;
        const float2v zxp = xh ? *(const float2v *)(z + c + s.sx) : f0;
        const float2v zym = yl ? *(const float2v *)(z + c - s.sy) : f0;
        const float2v zyp = yh ? *(const float2v *)(z + c + s.sy) : f0;
        const float zzm = zl ? z[c - 1] : 0.f;
        const float zzp = zh ? z[c + 2] : 0.f;
        const double2v cxm = xl ? *(const double2v *)(A.cpl[0] + c) : d0;
        const double2v cxp = xh ? *(const double2v *)(A.cpl[0] + c + s.sx) : d0;
        const double2v cym = yl ? *(const double2v *)(A.cpl[1] + c) : d0;
        const double2v cyp = yh ? *(const double2v *)(A.cpl[1] + c + s.sy) : d0;
        const double2v cz = *(const double2v *)(A.cpl[2] + c);  // below each cell of the pair
        const double czp = zh ? A.cpl[2][c + 2] : 0.0;          // above the second
        // A row-sum diagonal is rebuilt from the couplings (the same sums in
        // the same order as the assembly's, so the same bits); the pairs
        // holding cell 0 and its upper neighbours read the stored one.
        double2v dg;
        if (A.pinned_rowsum && cu != 0 && cu != (unsigned)d.nz && cu != (unsigned)s.sx) {
            dg.x = ((((cxm.x + cxp.x) + cym.x) + cyp.x) + cz.x) + cz.y;
            dg.y = ((((cxm.y + cxp.y) + cym.y) + cyp.y) + cz.y) + czp;
        } else {
            dg = *(const double2v *)(A.diag + c);
        }
        const double2v qp = first ? d0 : ld_stream((const double2v *)(qprev + c));
        const double2v rc = ld_stream((const double2v *)(r + c));
        double o0 = 0.0, o1 = 0.0;
        if (xl) {
            o0 += cxm.x * (double)zxm.x;
            o1 += cxm.y * (double)zxm.y;
        }
        if (xh) {
            o0 += cxp.x * (double)zxp.x;
            o1 += cxp.y * (double)zxp.y;
        }
        if (yl) {
            o0 += cym.x * (double)zym.x;
            o1 += cym.y * (double)zym.y;
        }
        if (yh) {
            o0 += cyp.x * (double)zyp.x;
            o1 += cyp.y * (double)zyp.y;
        }
        if (zl) o0 += cz.x * (double)zzm;
        o0 += cz.y * (double)zc.y;
        o1 += cz.y * (double)zc.x;
        if (zh) o1 += czp * (double)zzp;
        const double z0 = (double)zc.x, z1 = (double)zc.y;
        double2v wc;
        wc.x = dg.x * z0 - o0;
        wc.y = dg.y * z1 - o1;
        st_stream((double2v *)(w + c), wc);
        v[0] = z0 * wc.x + z1 * wc.y;
        if (!first) v[1] = z0 * qp.x + z1 * qp.y;
        v[2] = z0 * rc.x + z1 * rc.y;
    }
    block_partials<3>(v, partials);
}

// The update is pointwise: it runs over the flat cell index with two
// consecutive cells per lane (16-byte loads and stores of the FP64 vectors),
// one pair per thread (kUpdateCells per block). A wavefront whose 128 cells
// are all inactive is skipped (p, q, x, r stay 0 there). With b32 it also
// stores the new residual rounded to FP32: the right-hand side the next
// V-cycle starts from, which then need not read the FP64 residual.
constexpr unsigned kUpdateCells = 2 * kBlock;

__global__ __launch_bounds__(kBlock) void k_fcg_update(size_t n, const float *__restrict__ active,
                                                       const float *z,  // may be x0: read before x0 is written
                                                       const double *__restrict__ w,
                                                       double *__restrict__ p,
                                                       double *__restrict__ q,
                                                       double *__restrict__ x,
                                                       double *__restrict__ r,
                                                       float *__restrict__ b32,
                                                       float *x0, float c0, int first, int xmode,
                                                       double *__restrict__ partials,
                                                       const KrylovScalars *__restrict__ sc,
                                                       const double *gate) {
    if (gated(gate)) return;
    const double alpha = sc->alpha, beta = sc->beta, alpha_prev = sc->alpha_prev;
    double v[1] = {0.0};
    const size_t c = (size_t)blockIdx.x * kUpdateCells + 2 * (size_t)flat_tid();
    const bool pair = c + 1 < n;  // else at most the grid's last cell
    bool act = c < n;
    float2v dg = {0.f, 0.f};  // FP32 diagonal (the activity flag)
    if (active) {
        if (pair) dg = *(const float2v *)(active + c);
        else if (c < n) dg.x = active[c];
        act = dg.x > 0.f || dg.y > 0.f;
    }
    if (__any(act)) {
        if (pair) {
            const float2v zf = *(const float2v *)(z + c);
            const double2v zc = {(double)zf.x, (double)zf.y};
            const double2v wc = ld_stream((const double2v *)(w + c));
            const double2v r0 = ld_stream((const double2v *)(r + c));
            double2v pc, qc, pp = {0.0, 0.0};
            if (first) {
                pc = zc;
                qc = wc;
            } else {
                pp = ld_stream((const double2v *)(p + c));
                const double2v qq = ld_stream((const double2v *)(q + c));
                pc.x = zc.x - beta * pp.x;
                pc.y = zc.y - beta * pp.y;
                qc.x = wc.x - beta * qq.x;
                qc.y = wc.y - beta * qq.y;
            }
            double2v rc;
            rc.x = r0.x - alpha * qc.x;
            rc.y = r0.y - alpha * qc.y;
            st_stream((double2v *)(p + c), pc);
            st_stream((double2v *)(q + c), qc);
            st_stream((double2v *)(r + c), rc);
            if (xmode != X_LATER) {
                double2v xn = ld_stream((const double2v *)(x + c));
                if (xmode == X_BOTH) {
                    xn.x += alpha_prev * pp.x;
                    xn.y += alpha_prev * pp.y;
                }
                xn.x += alpha * pc.x;
                xn.y += alpha * pc.y;
                st_stream((double2v *)(x + c), xn);
            }
            if (b32) {
                const float2v bf = {(float)rc.x, (float)rc.y};
                *(float2v *)(b32 + c) = bf;
                if (x0) {
                    const float2v xf = {c0 * rcp_or_zero(dg.x) * bf.x, c0 * rcp_or_zero(dg.y) * bf.y};
                    *(float2v *)(x0 + c) = xf;
                }
            }
            v[0] += rc.x * rc.x;
            v[0] += rc.y * rc.y;
        } else if (c < n) {
            const double zc = (double)z[c];
            const double pp = first ? 0.0 : p[c];
            const double pc = first ? zc : zc - beta * pp;
            const double qc = first ? w[c] : w[c] - beta * q[c];
            p[c] = pc;
            q[c] = qc;
            if (xmode != X_LATER) {
                double xn = x[c];
                if (xmode == X_BOTH) xn += alpha_prev * pp;
                xn += alpha * pc;
                x[c] = xn;
            }
            const double rc = r[c] - alpha * qc;
            r[c] = rc;
            if (b32) {
                b32[c] = (float)rc;
                if (x0) x0[c] = c0 * rcp_or_zero(dg.x) * b32[c];
            }
            v[0] += rc * rc;
        }
    }
    block_partials<1>(v, partials);
}

// x += alpha p (the update an X_LATER iteration left pending), over the
// update's groups of cells.
__global__ __launch_bounds__(kBlock) void k_fcg_flush(size_t n, const float *__restrict__ active,
                                                      const double *__restrict__ p,
                                                      double *__restrict__ x,
                                                      const KrylovScalars *__restrict__ sc) {
    const double alpha = sc->alpha;
    const size_t c = (size_t)blockIdx.x * kUpdateCells + 2 * (size_t)flat_tid();
    const bool pair = c + 1 < n;
    bool act = c < n;
    if (active) {
        float2v dg = {0.f, 0.f};
        if (pair) dg = *(const float2v *)(active + c);
        else if (c < n) dg.x = active[c];
        act = dg.x > 0.f || dg.y > 0.f;
    }
    if (!__any(act)) return;
    if (pair) {
        const double2v pc = ld_stream((const double2v *)(p + c));
        double2v xn = ld_stream((const double2v *)(x + c));
        xn.x += alpha * pc.x;
        xn.y += alpha * pc.y;
        st_stream((double2v *)(x + c), xn);
    } else if (c < n) {
        x[c] += alpha * p[c];
    }
}

__global__ __launch_bounds__(kBlock) void k_jacobi(CellGrid cg, Dims d, Op32 A,
                                                   const double *__restrict__ r,
                                                   float *__restrict__ z, const double *gate) {
    if (gated(gate)) return;
    int k, j, i0;
    cell_tile(cg, k, j, i0);
    if (k >= d.nz || j >= d.ny) return;
    for (int i = i0, ie = min(i0 + cg.planes, d.nx); i < ie; i++) {
        const size_t c = d.idx(i, j, k);
        const float dg = A.diag[c];
        z[c] = dg > 0.f ? (float)(r[c] / (double)dg) : 0.f;
    }
}

__global__ __launch_bounds__(kBlock) void k_count_active(CellGrid cg, Dims d,
                                                         const float *__restrict__ diag32,
                                                         double *__restrict__ partials) {
    int k, j, i0;
    cell_tile(cg, k, j, i0);
    double v[1] = {0.0};
    if (k < d.nz && j < d.ny) {
        for (int i = i0, ie = min(i0 + cg.planes, d.nx); i < ie; i++)
            if (diag32[d.idx(i, j, k)] > 0.f) v[0] += 1.0;
    }
    block_partials<1>(v, partials);
}

// One block: add the partials and advance the scalar recurrences.
// The finalize kernel is one block of kFinalizeThreads (wide, so the partial
// sums of the largest grids take few rounds of loads).
constexpr int kFinalizeThreads = 1024;

// Start of a batch of gated iterations: clear the stop flag and the iteration
// count, set the threshold on ||r||^2 (one thread; no host copy needed).
__global__ __launch_bounds__(64) void k_batch_begin(KrylovScalars *sc, double stop_rr, double stall_rr) {
    if (threadIdx.x == 0) {
        sc->stop = 0.0;
        sc->iters = 0.0;
        sc->stop_rr = stop_rr;
        sc->stall_rr = stall_rr;
        sc->stalled = 0.0;
    }
}

__global__ __launch_bounds__(kFinalizeThreads) void k_finalize(int op, unsigned nb, int first,
                                                     const double *__restrict__ partials,
                                                     KrylovScalars *sc, const double *gate) {
    if (gated(gate)) return;
    double v[3] = {0.0, 0.0, 0.0};
    if (op == FIN_MATVEC) {
        sum_partials<3, kFinalizeThreads, 16>(partials, nb, v);
    } else if (op == FIN_RESIDUAL) {
        double t[2];
        sum_partials<2, kFinalizeThreads, 16>(partials, nb, t);
        v[0] = t[0];
        v[1] = t[1];
    } else {
        double t[1];
        sum_partials<1, kFinalizeThreads, 16>(partials, nb, t);
        v[0] = t[0];
    }
    if (flat_tid() != 0) return;
    if (op == FIN_RESIDUAL) {
        sc->rr = v[0];
        sc->bb = v[1];
    } else if (op == FIN_MATVEC) {
        const double zw = v[0], zq = v[1], zr = v[2];
        double beta = 0.0, pq = zw;
        if (!first && sc->pq_prev != 0.0) {
            // p = z - beta p_prev with beta = z.q_prev / p_prev.q_prev makes p
            // A-orthogonal to p_prev; then p.q = z.w - beta z.q_prev, and
            // p.r = z.r because p_prev.r = 0 after the previous step.
            beta = zq / sc->pq_prev;
            pq = zw - beta * zq;
        }
        sc->zw = zw;
        sc->zq = zq;
        sc->zr = zr;
        sc->beta = beta;
        sc->pq = pq;
        sc->alpha_prev = sc->alpha;
        sc->alpha = pq != 0.0 ? zr / pq : 0.0;
    } else if (op == FIN_UPDATE) {
        const double rr_prev = sc->rr;
        sc->rr = v[0];
        sc->pq_prev = sc->pq;
        sc->iters += 1.0;
        if (gate && v[0] <= sc->stop_rr) sc->stop = 1.0;  // later gated kernels return at once
        if (gate && sc->stall_rr > 0.0 && v[0] > sc->stall_rr * rr_prev) {
            sc->stalled = 1.0;  // the host switches to the plain V-cycle after the batch
            sc->stop = 1.0;
        }
    } else {
        sc->pad[0] = v[0];
    }
}

// Above kStageMin partials the sums take two launches: kStageThreads-wide
// blocks add fixed chunks of kStageChunk partials each (every block one
// round of loads instead of the single block's many), the finalize adds the
// chunks' sums. The order is fixed by nb, so results stay deterministic.
constexpr unsigned kStageMin = 4096, kStageChunk = 1024, kStageMaxBlocks = 64;
constexpr int kStageThreads = 1024;

template <int N>
__global__ __launch_bounds__(kStageThreads) void k_partials_stage(unsigned nb, unsigned chunk,
                                                                  const double *__restrict__ partials,
                                                                  double *__restrict__ out,
                                                                  const double *gate) {
    if (gated(gate)) return;
    const unsigned b0 = blockIdx.x * chunk;  // < nb: the launcher sizes the grid
    const unsigned n = min(chunk, nb - b0);
    double v[N];
    sum_partials<N, kStageThreads, 8>(partials + b0, n, v, nb);
    if (flat_tid() == 0)
#pragma unroll
        for (int q = 0; q < N; q++) out[(size_t)q * gridDim.x + blockIdx.x] = v[q];
}

// `partials` holds the N sums of op for nb blocks ([q * nb + b]); the stage
// writes its sums after them (reduction_partials leaves room).
void finalize(hipStream_t st, int op, unsigned nb, int first, double *partials, KrylovScalars *sc,
              const double *gate = nullptr) {
    const int N = op == FIN_MATVEC ? 3 : op == FIN_RESIDUAL ? 2 : 1;
    if (nb > kStageMin) {
        const unsigned chunk = std::max(kStageChunk, (nb + kStageMaxBlocks - 1) / kStageMaxBlocks);
        const unsigned G = (nb + chunk - 1) / chunk;
        double *out = partials + (size_t)N * nb;
        if (N == 3)
            hipLaunchKernelGGL(k_partials_stage<3>, dim3(G), dim3(kStageThreads), 0, st, nb, chunk, partials, out,
                               gate);
        else if (N == 2)
            hipLaunchKernelGGL(k_partials_stage<2>, dim3(G), dim3(kStageThreads), 0, st, nb, chunk, partials, out,
                               gate);
        else
            hipLaunchKernelGGL(k_partials_stage<1>, dim3(G), dim3(kStageThreads), 0, st, nb, chunk, partials, out,
                               gate);
        SFP_LAUNCH_CHECK();
        partials = out;
        nb = G;
    }
    hipLaunchKernelGGL(k_finalize, dim3(1), dim3(kFinalizeThreads), 0, st, op, nb, first, partials,
                       sc, gate);
    SFP_LAUNCH_CHECK();
}

}  // namespace

// FCG matvec: the flat kernel where z rows hold whole cell pairs (nz even;
// 11 % faster than x-plane tiles at 256^3 and 9 % at 384^3 on MI355X),
// else x-plane tiles of 4 planes per block (more blocks in flight than the 16
// of the other reductions: 6 % faster at 256^3).
static bool matvec_flat(Dims d) { return d.nz % 2 == 0 && d.cells() < (size_t(1) << 32); }
static Launch matvec_launch(Dims d) { return cell_launch(d, 0, 4); }
static unsigned matvec_blocks(Dims d) {
    return matvec_flat(d) ? (unsigned)((d.cells() + kMatvecCells - 1) / kMatvecCells)
                          : matvec_launch(d).blocks();
}
static unsigned update_blocks(Dims d) { return (unsigned)((d.cells() + kUpdateCells - 1) / kUpdateCells); }

size_t reduction_partials(Dims d) {
    // The largest set of partials, then room for the two-stage finalize's chunk sums.
    return std::max({3 * (size_t)cell_launch(d, kReducePlanes).blocks(),
                     3 * (size_t)matvec_blocks(d), (size_t)update_blocks(d)}) +
           3 * (size_t)kStageMaxBlocks;
}

void launch_count_active(hipStream_t st, Dims d, const float *diag32, double *partials,
                         KrylovScalars *sc) {
    Launch L = cell_launch(d, kReducePlanes);
    hipLaunchKernelGGL(k_count_active, L.grid, L.block, 0, st, L.cg, d, diag32, partials);
    SFP_LAUNCH_CHECK();
    finalize(st, FIN_COUNT, L.blocks(), 0, partials, sc);
}

void launch_batch_begin(hipStream_t st, KrylovScalars *sc, double stop_rr, double stall_rr) {
    hipLaunchKernelGGL(k_batch_begin, dim3(1), dim3(64), 0, st, sc, stop_rr, stall_rr);
    SFP_LAUNCH_CHECK();
}

void launch_residual64(hipStream_t st, Dims d, Op64 A, const double *b, const double *x,
                       double *r, double *partials, KrylovScalars *sc, float *b32, float *x0,
                       const float *diag32, float c0) {
    if (x0 && (!b32 || !diag32)) throw ArgError("launch_residual64: x0 needs b32 and the diagonal");
    if (matvec_flat(d)) {
        for (const void *a : {(const void *)A.diag, (const void *)A.cpl[0], (const void *)A.cpl[1],
                              (const void *)A.cpl[2], (const void *)b, (const void *)x, (const void *)r})
            if ((uintptr_t)a % 16) throw ArgError("launch_residual64: FP64 array not 16-byte aligned");
        if ((uintptr_t)b32 % 8 || (uintptr_t)x0 % 8 || (uintptr_t)diag32 % 8)
            throw ArgError("launch_residual64: b32, x0 or the diagonal not 8-byte aligned");
        const unsigned nb = (unsigned)((d.cells() + kResidualCells - 1) / kResidualCells);
        hipLaunchKernelGGL(k_residual64_flat, dim3(nb), dim3(kBlock), 0, st, d, A, b, x, r, partials, b32,
                           x0, diag32, c0);
        SFP_LAUNCH_CHECK();
        finalize(st, FIN_RESIDUAL, nb, 0, partials, sc);
        return;
    }
    Launch L = cell_launch(d, kReducePlanes);
    hipLaunchKernelGGL(k_residual64, L.grid, L.block, 0, st, L.cg, d, A, b, x, r, partials, b32, x0,
                       diag32, c0);
    SFP_LAUNCH_CHECK();
    finalize(st, FIN_RESIDUAL, L.blocks(), 0, partials, sc);
}

void launch_fcg_matvec(hipStream_t st, Dims d, Op64 A, const float *active, const float *z,
                       const double *r, const double *qprev, double *w, int first,
                       double *partials, KrylovScalars *sc, const double *gate, hipEvent_t kernel_done) {
    if (matvec_flat(d)) {
        // pairs of cells: 16-byte (8-byte for z and the flags) aligned vectors
        for (const void *a : {(const void *)A.diag, (const void *)A.cpl[0], (const void *)A.cpl[1],
                              (const void *)A.cpl[2], (const void *)r, (const void *)qprev, (const void *)w})
            if ((uintptr_t)a % 16) throw ArgError("launch_fcg_matvec: FP64 array not 16-byte aligned");
        if ((uintptr_t)z % 8 || (uintptr_t)active % 8)
            throw ArgError("launch_fcg_matvec: z or the diagonal not 8-byte aligned");
        const unsigned nb = matvec_blocks(d);
        hipLaunchKernelGGL(k_fcg_matvec_flat, dim3(nb), dim3(kBlock), 0, st, d, A, active, z, r,
                           qprev, w, first, partials, gate);
        SFP_LAUNCH_CHECK();
        if (kernel_done) SFP_HIP(hipEventRecord(kernel_done, st));
        finalize(st, FIN_MATVEC, nb, first, partials, sc, gate);
        return;
    }
    Launch L = matvec_launch(d);
    hipLaunchKernelGGL(k_fcg_matvec, L.grid, L.block, 0, st, L.cg, d, A, active, z, r, qprev, w,
                       first, partials, gate);
    SFP_LAUNCH_CHECK();
    if (kernel_done) SFP_HIP(hipEventRecord(kernel_done, st));
    finalize(st, FIN_MATVEC, L.blocks(), first, partials, sc, gate);
}

void launch_fcg_update(hipStream_t st, Dims d, const float *active, const float *z,
                       const double *w, double *p, double *q, double *x, double *r, float *b32,
                       float *x0, float c0, int first, int xmode, double *partials,
                       KrylovScalars *sc, const double *gate, hipEvent_t kernel_done) {
    if (xmode == X_BOTH && first) throw ArgError("launch_fcg_update: X_BOTH needs a previous direction");
    // two cells per lane: 16-byte (8-byte for z) aligned vectors
    for (const void *a : {(const void *)w, (const void *)p, (const void *)q, (const void *)x, (const void *)r})
        if ((uintptr_t)a % 16) throw ArgError("launch_fcg_update: FP64 vector not 16-byte aligned");
    if ((uintptr_t)z % 8 || (uintptr_t)b32 % 8 || (uintptr_t)x0 % 8)
        throw ArgError("launch_fcg_update: z, b32 or x0 not 8-byte aligned");
    if (x0 && (!b32 || !active)) throw ArgError("launch_fcg_update: x0 needs b32 and the diagonal");
    const unsigned nb = update_blocks(d);
    hipLaunchKernelGGL(k_fcg_update, dim3(nb), dim3(kBlock), 0, st, d.cells(), active, z, w, p, q,
                       x, r, b32, x0, c0, first, xmode, partials, sc, gate);
    SFP_LAUNCH_CHECK();
    if (kernel_done) SFP_HIP(hipEventRecord(kernel_done, st));
    finalize(st, FIN_UPDATE, nb, first, partials, sc, gate);
}

void launch_fcg_flush(hipStream_t st, Dims d, const float *active, const double *p, double *x,
                      const KrylovScalars *sc) {
    const unsigned nb = update_blocks(d);
    hipLaunchKernelGGL(k_fcg_flush, dim3(nb), dim3(kBlock), 0, st, d.cells(), active, p, x, sc);
    SFP_LAUNCH_CHECK();
}

void launch_jacobi(hipStream_t st, Dims d, Op32 A, const double *r, float *z, const double *gate) {
    Launch L = cell_launch(d);
    hipLaunchKernelGGL(k_jacobi, L.grid, L.block, 0, st, L.cg, d, A, r, z, gate);
    SFP_LAUNCH_CHECK();
}

}  // namespace sfp
