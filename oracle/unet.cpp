// FP64 CPU reference of the 3-level 3-D U-Net used as the learned
// preconditioner: forward pass and backward pass (MSE loss) over a batch of
// channels-last volumes. Test infrastructure only (see oracle.h); pinned by
// finite-difference gradient checks in tests/test_unet_oracle.py.
//
// Network (c = channels, every conv 3x3x3 stride 1 zero padding, ReLU after
// every conv except the last 1x1):
//   L0: x(8) -c0-> a0(16) -c1-> e0(16)            pool2 -> p0
//   L1: p0(16) -c2-> a1(32) -c3-> e1(32)          pool2 -> p1
//   L2: p1(32) -c4-> a2(64) -c5-> e2(64)          up2   -> u1
//   L1: relu(c6(u1)) + e1 = s1(32) -c7-> d1(32)   up2   -> u0
//   L0: relu(c8(u0)) + e0 = s0(16) -c9-> d0(16)   1x1   -> z(1)
// pool2 = 2x2x2 average, up2 = nearest. Parameter vector: for each conv k,
// weights [27][cout][cin] then bias [cout]; then the 1x1 weights [16], bias.
#include "oracle.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

namespace {

struct Conv {
    int cin, cout;
};
const Conv kConvs[10] = {{8, 16}, {16, 16}, {16, 32}, {32, 32}, {32, 64},
                         {64, 64}, {64, 32}, {32, 32}, {32, 16}, {16, 16}};

long long param_count() {
    long long n = 0;
    for (auto c : kConvs) n += 27LL * c.cin * c.cout + c.cout;
    return n + 16 + 1;
}

struct Vol {  // [b][x][y][z][c]
    int b = 0, x = 0, y = 0, z = 0, c = 0;
    std::vector<double> v;
    Vol() = default;
    Vol(int B, int X, int Y, int Z, int C) : b(B), x(X), y(Y), z(Z), c(C), v((size_t)B * X * Y * Z * C, 0.0) {}
    size_t at(int ib, int ix, int iy, int iz, int ic) const {
        return ((((size_t)ib * x + ix) * y + iy) * z + iz) * c + ic;
    }
};

void conv_fwd(const Vol &in, const double *w, const double *bias, int cout, Vol &out) {
    out = Vol(in.b, in.x, in.y, in.z, cout);
    const int cin = in.c;
#pragma omp parallel for collapse(2) schedule(static)
    for (int b = 0; b < in.b; b++)
        for (int ix = 0; ix < in.x; ix++)
            for (int iy = 0; iy < in.y; iy++)
                for (int iz = 0; iz < in.z; iz++)
                    for (int co = 0; co < cout; co++) {
                        double s = bias[co];
                        for (int t = 0; t < 27; t++) {
                            const int jx = ix + t / 9 - 1, jy = iy + (t / 3) % 3 - 1, jz = iz + t % 3 - 1;
                            if (jx < 0 || jy < 0 || jz < 0 || jx >= in.x || jy >= in.y || jz >= in.z) continue;
                            const double *xi = &in.v[in.at(b, jx, jy, jz, 0)];
                            const double *wt = w + ((size_t)t * cout + co) * cin;
                            for (int ci = 0; ci < cin; ci++) s += wt[ci] * xi[ci];
                        }
                        out.v[out.at(b, ix, iy, iz, co)] = s;
                    }
}

// gin = conv^T(gout); gw, gb accumulate
void conv_bwd(const Vol &in, const double *w, const Vol &gout, Vol *gin, double *gw, double *gb) {
    const int cin = in.c, cout = gout.c;
    if (gin) *gin = Vol(in.b, in.x, in.y, in.z, cin);
    for (int b = 0; b < in.b; b++)
        for (int ix = 0; ix < in.x; ix++)
            for (int iy = 0; iy < in.y; iy++)
                for (int iz = 0; iz < in.z; iz++)
                    for (int co = 0; co < cout; co++) {
                        const double g = gout.v[gout.at(b, ix, iy, iz, co)];
                        if (g == 0.0) continue;
                        gb[co] += g;
                        for (int t = 0; t < 27; t++) {
                            const int jx = ix + t / 9 - 1, jy = iy + (t / 3) % 3 - 1, jz = iz + t % 3 - 1;
                            if (jx < 0 || jy < 0 || jz < 0 || jx >= in.x || jy >= in.y || jz >= in.z) continue;
                            const size_t xi = in.at(b, jx, jy, jz, 0);
                            const size_t wi = ((size_t)t * cout + co) * cin;
                            for (int ci = 0; ci < cin; ci++) {
                                gw[wi + ci] += g * in.v[xi + ci];
                                if (gin) gin->v[xi + ci] += g * w[wi + ci];
                            }
                        }
                    }
}

void relu(Vol &a) {
    for (auto &x : a.v) x = x > 0 ? x : 0.0;
}

void relu_mask(const Vol &act, Vol &g) {
    for (size_t i = 0; i < g.v.size(); i++)
        if (!(act.v[i] > 0)) g.v[i] = 0.0;
}

Vol pool2(const Vol &a) {
    Vol o(a.b, a.x / 2, a.y / 2, a.z / 2, a.c);
    for (int b = 0; b < o.b; b++)
        for (int x = 0; x < o.x; x++)
            for (int y = 0; y < o.y; y++)
                for (int z = 0; z < o.z; z++)
                    for (int c = 0; c < o.c; c++) {
                        double s = 0;
                        for (int d = 0; d < 8; d++)
                            s += a.v[a.at(b, 2 * x + (d >> 2), 2 * y + ((d >> 1) & 1), 2 * z + (d & 1), c)];
                        o.v[o.at(b, x, y, z, c)] = s * 0.125;
                    }
    return o;
}

Vol pool2_bwd(const Vol &g, const Vol &shape) {
    Vol o(shape.b, shape.x, shape.y, shape.z, shape.c);
    for (int b = 0; b < o.b; b++)
        for (int x = 0; x < o.x; x++)
            for (int y = 0; y < o.y; y++)
                for (int z = 0; z < o.z; z++)
                    for (int c = 0; c < o.c; c++) o.v[o.at(b, x, y, z, c)] = 0.125 * g.v[g.at(b, x / 2, y / 2, z / 2, c)];
    return o;
}

Vol up2(const Vol &a) {
    Vol o(a.b, a.x * 2, a.y * 2, a.z * 2, a.c);
    for (int b = 0; b < o.b; b++)
        for (int x = 0; x < o.x; x++)
            for (int y = 0; y < o.y; y++)
                for (int z = 0; z < o.z; z++)
                    for (int c = 0; c < o.c; c++) o.v[o.at(b, x, y, z, c)] = a.v[a.at(b, x / 2, y / 2, z / 2, c)];
    return o;
}

Vol up2_bwd(const Vol &g) {
    Vol o(g.b, g.x / 2, g.y / 2, g.z / 2, g.c);
    for (int b = 0; b < g.b; b++)
        for (int x = 0; x < g.x; x++)
            for (int y = 0; y < g.y; y++)
                for (int z = 0; z < g.z; z++)
                    for (int c = 0; c < g.c; c++) o.v[o.at(b, x / 2, y / 2, z / 2, c)] += g.v[g.at(b, x, y, z, c)];
    return o;
}

void add(Vol &a, const Vol &b) {
    for (size_t i = 0; i < a.v.size(); i++) a.v[i] += b.v[i];
}

struct Params {
    const double *w[10], *b[10], *wo, *bo;
    explicit Params(const double *p) {
        for (int k = 0; k < 10; k++) {
            w[k] = p;
            p += 27LL * kConvs[k].cin * kConvs[k].cout;
            b[k] = p;
            p += kConvs[k].cout;
        }
        wo = p;
        bo = p + 16;
    }
};

struct Acts {
    Vol x, a0, e0, p0, a1, e1, p1, a2, e2, u1, r6, s1, d1, u0, r8, s0, d0;
    std::vector<double> z;
};

void forward(const Params &P, Acts &A) {
    conv_fwd(A.x, P.w[0], P.b[0], 16, A.a0), relu(A.a0);
    conv_fwd(A.a0, P.w[1], P.b[1], 16, A.e0), relu(A.e0);
    A.p0 = pool2(A.e0);
    conv_fwd(A.p0, P.w[2], P.b[2], 32, A.a1), relu(A.a1);
    conv_fwd(A.a1, P.w[3], P.b[3], 32, A.e1), relu(A.e1);
    A.p1 = pool2(A.e1);
    conv_fwd(A.p1, P.w[4], P.b[4], 64, A.a2), relu(A.a2);
    conv_fwd(A.a2, P.w[5], P.b[5], 64, A.e2), relu(A.e2);
    A.u1 = up2(A.e2);
    conv_fwd(A.u1, P.w[6], P.b[6], 32, A.r6), relu(A.r6);
    A.s1 = A.r6, add(A.s1, A.e1);
    conv_fwd(A.s1, P.w[7], P.b[7], 32, A.d1), relu(A.d1);
    A.u0 = up2(A.d1);
    conv_fwd(A.u0, P.w[8], P.b[8], 16, A.r8), relu(A.r8);
    A.s0 = A.r8, add(A.s0, A.e0);
    conv_fwd(A.s0, P.w[9], P.b[9], 16, A.d0), relu(A.d0);
    const size_t nv = A.d0.v.size() / 16;
    A.z.assign(nv, 0.0);
    for (size_t i = 0; i < nv; i++) {
        double s = P.bo[0];
        for (int c = 0; c < 16; c++) s += P.wo[c] * A.d0.v[i * 16 + c];
        A.z[i] = s;
    }
}

}  // namespace

extern "C" {

long long orc_unet_param_count(void) { return param_count(); }

double orc_unet_forward(int batch, int nx, int ny, int nz, const double *params, const double *input,
                        const double *target, const double *mask, double *output) {
    Params P(params);
    Acts A;
    A.x = Vol(batch, nx, ny, nz, 8);
    const size_t nv = (size_t)batch * nx * ny * nz;
    for (size_t i = 0; i < nv; i++)
        for (int c = 0; c < 4; c++) A.x.v[i * 8 + c] = input[i * 4 + c];
    forward(P, A);
    if (output) std::memcpy(output, A.z.data(), nv * sizeof(double));
    if (!target) return 0.0;
    double loss = 0, cnt = 0;
    for (size_t i = 0; i < nv; i++) {
        const double m = mask ? mask[i] : 1.0;
        loss += m * (A.z[i] - target[i]) * (A.z[i] - target[i]);
        cnt += m;
    }
    return cnt > 0 ? loss / cnt : 0.0;
}

double orc_unet_gradients(int batch, int nx, int ny, int nz, const double *params, const double *input,
                          const double *target, const double *mask, double *grad) {
    Params P(params);
    Acts A;
    A.x = Vol(batch, nx, ny, nz, 8);
    const size_t nv = (size_t)batch * nx * ny * nz;
    for (size_t i = 0; i < nv; i++)
        for (int c = 0; c < 4; c++) A.x.v[i * 8 + c] = input[i * 4 + c];
    forward(P, A);
    double loss = 0, cnt = 0;
    for (size_t i = 0; i < nv; i++) {
        const double m = mask ? mask[i] : 1.0;
        loss += m * (A.z[i] - target[i]) * (A.z[i] - target[i]);
        cnt += m;
    }
    const long long np = param_count();
    std::fill(grad, grad + np, 0.0);
    if (cnt == 0) return 0.0;
    // gradient pointers inside grad, same layout as params
    double *gw[10], *gb[10];
    double *g = grad;
    for (int k = 0; k < 10; k++) {
        gw[k] = g;
        g += 27LL * kConvs[k].cin * kConvs[k].cout;
        gb[k] = g;
        g += kConvs[k].cout;
    }
    double *gwo = g, *gbo = g + 16;
    Vol gd0(batch, nx, ny, nz, 16);
    for (size_t i = 0; i < nv; i++) {
        const double m = mask ? mask[i] : 1.0;
        const double dz = m * 2.0 * (A.z[i] - target[i]) / cnt;
        gbo[0] += dz;
        for (int c = 0; c < 16; c++) {
            gwo[c] += dz * A.d0.v[i * 16 + c];
            gd0.v[i * 16 + c] = dz * P.wo[c];
        }
    }
    relu_mask(A.d0, gd0);
    Vol gs0, gr8, gu0, gd1, gs1, gr6, gu1, ge2, ga2, gp1, ge1, ga1, gp0, ge0, ga0;
    conv_bwd(A.s0, P.w[9], gd0, &gs0, gw[9], gb[9]);
    gr8 = gs0, relu_mask(A.r8, gr8);
    conv_bwd(A.u0, P.w[8], gr8, &gu0, gw[8], gb[8]);
    gd1 = up2_bwd(gu0), relu_mask(A.d1, gd1);
    conv_bwd(A.s1, P.w[7], gd1, &gs1, gw[7], gb[7]);
    gr6 = gs1, relu_mask(A.r6, gr6);
    conv_bwd(A.u1, P.w[6], gr6, &gu1, gw[6], gb[6]);
    ge2 = up2_bwd(gu1), relu_mask(A.e2, ge2);
    conv_bwd(A.a2, P.w[5], ge2, &ga2, gw[5], gb[5]);
    relu_mask(A.a2, ga2);
    conv_bwd(A.p1, P.w[4], ga2, &gp1, gw[4], gb[4]);
    ge1 = pool2_bwd(gp1, A.e1);
    add(ge1, gs1);  // skip into s1
    relu_mask(A.e1, ge1);
    conv_bwd(A.a1, P.w[3], ge1, &ga1, gw[3], gb[3]);
    relu_mask(A.a1, ga1);
    conv_bwd(A.p0, P.w[2], ga1, &gp0, gw[2], gb[2]);
    ge0 = pool2_bwd(gp0, A.e0);
    add(ge0, gs0);  // skip into s0
    relu_mask(A.e0, ge0);
    conv_bwd(A.a0, P.w[1], ge0, &ga0, gw[1], gb[1]);
    relu_mask(A.a0, ga0);
    conv_bwd(A.x, P.w[0], ga0, nullptr, gw[0], gb[0]);
    return loss / cnt;
}

}  // extern "C"
