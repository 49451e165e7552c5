// 3-D U-Net of the learned preconditioner: kernel launchers and the host-side
// network (weights, activations, forward, backward, Adam).
//
// Network (every conv 3x3x3, zero padding, ReLU; pool = 2x2x2 average,
// up = nearest):
//   L0: x(8) -c0-> a0(16) -c1-> e0(16)            pool -> p0
//   L1: p0(16) -c2-> a1(32) -c3-> e1(32)          pool -> p1
//   L2: p1(32) -c4-> a2(64) -c5-> e2(64)          up   -> u1
//   L1: relu(c6(u1)) + e1 = s1(32) -c7-> d1(32)   up   -> u0
//   L0: relu(c8(u0)) + e0 = s0(16) -c9-> d0(16)   1x1  -> z
// Parameters (FP32): per conv k, weights [27][cout][cin] then bias [cout];
// then the 1x1 head weights [16] and bias.
#pragma once
#include "sfp_internal.hpp"

#include <algorithm>
#include <cstdint>

namespace sfp {

constexpr int kUNetConvs = 10;
constexpr int kConvCin[kUNetConvs] = {8, 16, 16, 32, 32, 64, 64, 32, 32, 16};
constexpr int kConvCout[kUNetConvs] = {16, 16, 32, 32, 64, 64, 32, 32, 16, 16};
constexpr int kDwChunks = 64;  // voxel chunks of the weight-gradient reduction

// Taps stored for a convolution with `cin` input channels: padded so that the
// 32-wide K chunks of the MFMA cover whole taps.
constexpr int conv_taps(int cin) { return cin <= 16 ? 28 : 27; }

long long unet_param_count();

// launchers (unet.hip)
// up_input: `in` holds the half-resolution tensor, read nearest-up-sampled
// (the up-sampling fused into the staging). gate (the forward's kernels): the
// launch returns at once while *gate is nonzero (a stopped batch of FCG
// iterations).
void unet_conv(hipStream_t st, int cin, int cout, const uint16_t *in, const uint16_t *w,
               const float *bias, const uint16_t *skip, uint16_t *out, float *out32,
               uint16_t *relu_copy, int B, int X, int Y, int Z, bool relu, bool up_input = false,
               const double *gate = nullptr);
void unet_conv_dw(hipStream_t st, int cin, int cout, const uint16_t *x, const uint16_t *g, int B,
                  int X, int Y, int Z, float *part, float *dw, float *db);
void unet_pool2(hipStream_t st, const uint16_t *in, int B, int X, int Y, int Z, int C, uint16_t *out,
                const double *gate = nullptr);
void unet_pool2_bwd(hipStream_t st, const float *g, const float *add, int B, int X, int Y, int Z,
                    int C, float *gin);
void unet_up2(hipStream_t st, const uint16_t *in, int B, int Xc, int Yc, int Zc, int C, uint16_t *out);
void unet_up2_bwd(hipStream_t st, const float *g, int B, int Xc, int Yc, int Zc, int C, float *gc);
void unet_relu_mask(hipStream_t st, const float *g, const uint16_t *act, size_t n, uint16_t *g16);
void unet_head_fwd(hipStream_t st, const uint16_t *d0, const float *wb, size_t nv, float *z,
                   const double *gate = nullptr);
void unet_head_bwd(hipStream_t st, const uint16_t *d0, const float *wb, const float *dz, size_t nv,
                   float *gd0, double *part, double *sums, float *gwb);
void unet_mse(hipStream_t st, const float *z, const float *t, const float *m, size_t nv,
              double *part, double *sums, float *dz);
void unet_adam(hipStream_t st, float *w, const float *g, float *m, float *v, size_t n, float lr,
               int step);
void unet_cast_weights(hipStream_t st, const float *w, int cin, int cout, uint16_t *wf, uint16_t *wt);
void unet_pack_input(hipStream_t st, const float *in, size_t nv, uint16_t *out);

struct UNetShape {
    int B = 0, X = 0, Y = 0, Z = 0;
    size_t voxels(int level) const {
        return (size_t)B * (X >> level) * (Y >> level) * (Z >> level);
    }
};

class UNet {
  public:
    UNet();
    // parameters (host FP32, unet_param_count() values)
    void set_params(hipStream_t st, const float *host);
    void get_params(hipStream_t st, float *host) const;
    void init_random(hipStream_t st, unsigned long long seed);
    // activations for volumes of this shape (each axis a multiple of 4)
    void ensure(const UNetShape &s, bool training);
    const UNetShape &shape() const { return shape_; }
    // x8: BF16 [B][X][Y][Z][8]; z: FP32 [B][X][Y][Z]; every launch returns at
    // once while *gate is nonzero (gate null: always runs)
    void forward(hipStream_t st, const uint16_t *x8, float *z, const double *gate = nullptr);
    // profiling: with events (2 per convolution, kUNetConvs pairs), every
    // convolution of the forward is bracketed by its pair; null: none
    void set_conv_events(hipEvent_t *events) { conv_events_ = events; }
    // masked MSE against t (FP32 [voxels]); returns the loss on the host
    double loss_and_backward(hipStream_t st, const uint16_t *x8, const float *t, const float *m,
                             double *host_sums);
    void adam_step(hipStream_t st, float lr);
    float *grads() { return grad_.p; }
    uint16_t *input_buffer() { return x8_.p; }
    float *output_buffer() { return z_.p; }

  private:
    void cast_weights(hipStream_t st);
    void conv(hipStream_t st, int k, int level, const uint16_t *in, const uint16_t *skip,
              uint16_t *out, uint16_t *relu_copy, bool up_input = false, const double *gate = nullptr);
    int dims(int level, int axis) const;
    hipEvent_t *conv_events_ = nullptr;
    long long off_w_[kUNetConvs], off_b_[kUNetConvs], off_head_;
    DevBuf<float> w_, grad_, m_, v_;
    DevBuf<uint16_t> wf_[kUNetConvs], wt_[kUNetConvs];
    UNetShape shape_;
    bool training_ = false;
    int step_ = 0;
    // activations (BF16)
    DevBuf<uint16_t> x8_, a0_, e0_, p0_, a1_, e1_, p1_, a2_, e2_, u1_, r6_, s1_, d1_, u0_, r8_, s0_, d0_;
    DevBuf<float> z_, dz_;
    // backward scratch
    DevBuf<float> gA_, gB_, gC_, gs0_, gs1_, part_;
    DevBuf<uint16_t> g16_;
    DevBuf<double> dpart_, dsums_;
};

}  // namespace sfp
