// Host side of the U-Net: parameter layout, forward, backward and Adam.
#include "unet.hpp"

#include <cmath>
#include <cstring>
#include <vector>

namespace sfp {

long long unet_param_count() {
    long long n = 0;
    for (int k = 0; k < kUNetConvs; k++) n += 27LL * kConvCin[k] * kConvCout[k] + kConvCout[k];
    return n + 16 + 1;
}

UNet::UNet() {
    long long o = 0;
    for (int k = 0; k < kUNetConvs; k++) {
        off_w_[k] = o;
        o += 27LL * kConvCin[k] * kConvCout[k];
        off_b_[k] = o;
        o += kConvCout[k];
    }
    off_head_ = o;
    const size_t n = (size_t)unet_param_count();
    w_.ensure(n);
    grad_.ensure(n);
    m_.ensure(n);
    v_.ensure(n);
    SFP_HIP(hipMemset(w_.p, 0, n * sizeof(float)));
    SFP_HIP(hipMemset(m_.p, 0, n * sizeof(float)));
    SFP_HIP(hipMemset(v_.p, 0, n * sizeof(float)));
    SFP_HIP(hipMemset(grad_.p, 0, n * sizeof(float)));
    for (int k = 0; k < kUNetConvs; k++) {
        wf_[k].ensure((size_t)conv_taps(kConvCin[k]) * kConvCin[k] * kConvCout[k]);
        wt_[k].ensure((size_t)conv_taps(kConvCout[k]) * kConvCin[k] * kConvCout[k]);
    }
}

void UNet::cast_weights(hipStream_t st) {
    for (int k = 0; k < kUNetConvs; k++)
        unet_cast_weights(st, w_.p + off_w_[k], kConvCin[k], kConvCout[k], wf_[k].p, k > 0 ? wt_[k].p : nullptr);
}

void UNet::set_params(hipStream_t st, const float *host) {
    SFP_HIP(hipMemcpyAsync(w_.p, host, w_.n * sizeof(float), hipMemcpyHostToDevice, st));
    SFP_HIP(hipMemsetAsync(m_.p, 0, m_.n * sizeof(float), st));
    SFP_HIP(hipMemsetAsync(v_.p, 0, v_.n * sizeof(float), st));
    step_ = 0;
    cast_weights(st);
    SFP_HIP(hipStreamSynchronize(st));
}

void UNet::get_params(hipStream_t st, float *host) const {
    SFP_HIP(hipMemcpyAsync(host, w_.p, w_.n * sizeof(float), hipMemcpyDeviceToHost, st));
    SFP_HIP(hipStreamSynchronize(st));
}

void UNet::init_random(hipStream_t st, unsigned long long seed) {
    // He-normal weights from SplitMix64 + Box-Muller; zero biases
    std::vector<float> p((size_t)unet_param_count(), 0.f);
    unsigned long long s = seed;
    auto next = [&]() {
        unsigned long long z = (s += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        z ^= z >> 31;
        return ((z >> 11) + 0.5) * (1.0 / 9007199254740992.0);
    };
    auto normal = [&]() { return std::sqrt(-2.0 * std::log(next())) * std::cos(6.283185307179586 * next()); };
    for (int k = 0; k < kUNetConvs; k++) {
        const double stdv = std::sqrt(2.0 / (27.0 * (k == 0 ? 4 : kConvCin[k])));
        for (long long i = 0; i < 27LL * kConvCin[k] * kConvCout[k]; i++) {
            const int ci = (int)(i % kConvCin[k]);
            p[off_w_[k] + i] = (k == 0 && ci >= 4) ? 0.f : (float)(stdv * normal());
        }
    }
    for (int c = 0; c < 16; c++) p[off_head_ + c] = (float)(0.25 * normal());
    set_params(st, p.data());
}

int UNet::dims(int level, int axis) const {
    const int n = axis == 0 ? shape_.X : axis == 1 ? shape_.Y : shape_.Z;
    return n >> level;
}

void UNet::ensure(const UNetShape &s, bool training) {
    if (s.X % 4 || s.Y % 4 || s.Z % 4 || s.B < 1 || s.X < 4 || s.Y < 4 || s.Z < 4)
        throw ArgError("U-Net volumes need every axis a positive multiple of 4");
    shape_ = s;
    training_ = training;
    const size_t n0 = s.voxels(0), n1 = s.voxels(1), n2 = s.voxels(2);
    x8_.ensure(n0 * 8);
    for (auto *b : {&a0_, &e0_, &r8_, &s0_, &d0_}) b->ensure(n0 * 16);
    if (training) u0_.ensure(n0 * 32);
    p0_.ensure(n1 * 16);
    for (auto *b : {&a1_, &e1_, &r6_, &s1_, &d1_}) b->ensure(n1 * 32);
    if (training) u1_.ensure(n1 * 64);
    p1_.ensure(n2 * 32);
    a2_.ensure(n2 * 64);
    e2_.ensure(n2 * 64);
    z_.ensure(n0);
    if (training) {
        dz_.ensure(n0);
        gA_.ensure(n0 * 16);
        gB_.ensure(n0 * 32);
        gs0_.ensure(n0 * 16);
        gs1_.ensure(n1 * 32);
        g16_.ensure(n0 * 16);
        part_.ensure((size_t)kDwChunks * (27 * 64 * 64 + 64));
    }
    dpart_.ensure(17 * ((n0 + 255) / 256) + 17);
    dsums_.ensure(17);
}

void UNet::conv(hipStream_t st, int k, int level, const uint16_t *in, const uint16_t *skip,
                uint16_t *out, uint16_t *relu_copy, bool up_input, const double *gate) {
    if (conv_events_) SFP_HIP(hipEventRecord(conv_events_[2 * k], st));
    unet_conv(st, kConvCin[k], kConvCout[k], in, wf_[k].p, w_.p + off_b_[k], skip, out, nullptr, relu_copy,
              shape_.B, dims(level, 0), dims(level, 1), dims(level, 2), true, up_input, gate);
    if (conv_events_) SFP_HIP(hipEventRecord(conv_events_[2 * k + 1], st));
}

void UNet::forward(hipStream_t st, const uint16_t *x8, float *z, const double *gate) {
    const int B = shape_.B;
    uint16_t *keep8 = training_ ? r8_.p : nullptr, *keep6 = training_ ? r6_.p : nullptr;
    conv(st, 0, 0, x8, nullptr, a0_.p, nullptr, false, gate);
    conv(st, 1, 0, a0_.p, nullptr, e0_.p, nullptr, false, gate);
    unet_pool2(st, e0_.p, B, dims(0, 0), dims(0, 1), dims(0, 2), 16, p0_.p, gate);
    conv(st, 2, 1, p0_.p, nullptr, a1_.p, nullptr, false, gate);
    conv(st, 3, 1, a1_.p, nullptr, e1_.p, nullptr, false, gate);
    unet_pool2(st, e1_.p, B, dims(1, 0), dims(1, 1), dims(1, 2), 32, p1_.p, gate);
    conv(st, 4, 2, p1_.p, nullptr, a2_.p, nullptr, false, gate);
    conv(st, 5, 2, a2_.p, nullptr, e2_.p, nullptr, false, gate);
    // training keeps the up-sampled tensors (the weight gradients read them);
    // inference reads the half-resolution ones through the convolution
    if (training_) {
        unet_up2(st, e2_.p, B, dims(2, 0), dims(2, 1), dims(2, 2), 64, u1_.p);
        conv(st, 6, 1, u1_.p, e1_.p, s1_.p, keep6, false, gate);
    } else {
        conv(st, 6, 1, e2_.p, e1_.p, s1_.p, keep6, true, gate);
    }
    conv(st, 7, 1, s1_.p, nullptr, d1_.p, nullptr, false, gate);
    if (training_) {
        unet_up2(st, d1_.p, B, dims(1, 0), dims(1, 1), dims(1, 2), 32, u0_.p);
        conv(st, 8, 0, u0_.p, e0_.p, s0_.p, keep8, false, gate);
    } else {
        conv(st, 8, 0, d1_.p, e0_.p, s0_.p, keep8, true, gate);
    }
    conv(st, 9, 0, s0_.p, nullptr, d0_.p, nullptr, false, gate);
    unet_head_fwd(st, d0_.p, w_.p + off_head_, shape_.voxels(0), z, gate);
}

double UNet::loss_and_backward(hipStream_t st, const uint16_t *x8, const float *t, const float *m,
                               double *host_sums) {
    if (!training_) throw StateError("network buffers were sized for inference");
    const int B = shape_.B;
    const size_t n0 = shape_.voxels(0), n1 = shape_.voxels(1), n2 = shape_.voxels(2);
    forward(st, x8, z_.p);
    SFP_HIP(hipMemsetAsync(grad_.p, 0, grad_.n * sizeof(float), st));
    unet_mse(st, z_.p, t, m, n0, dpart_.p, dsums_.p, dz_.p);
    double mse[2];
    SFP_HIP(hipMemcpyAsync(mse, dsums_.p, 2 * sizeof(double), hipMemcpyDeviceToHost, st));
    SFP_HIP(hipStreamSynchronize(st));
    unet_head_bwd(st, d0_.p, w_.p + off_head_, dz_.p, n0, gA_.p, dpart_.p, dsums_.p, grad_.p + off_head_);
    auto L = [&](int level, int a) { return dims(level, a); };
    auto dw = [&](int k, int level, const uint16_t *x, const uint16_t *g) {
        unet_conv_dw(st, kConvCin[k], kConvCout[k], x, g, B, L(level, 0), L(level, 1), L(level, 2), part_.p,
                     grad_.p + off_w_[k], grad_.p + off_b_[k]);
    };
    auto dx = [&](int k, int level, const uint16_t *g, float *out) {
        unet_conv(st, kConvCout[k], kConvCin[k], g, wt_[k].p, nullptr, nullptr, nullptr, out, nullptr, B,
                  L(level, 0), L(level, 1), L(level, 2), false);
    };
    // level 0 decoder
    unet_relu_mask(st, gA_.p, d0_.p, n0 * 16, g16_.p);
    dw(9, 0, s0_.p, g16_.p);
    dx(9, 0, g16_.p, gs0_.p);
    unet_relu_mask(st, gs0_.p, r8_.p, n0 * 16, g16_.p);
    dw(8, 0, u0_.p, g16_.p);
    dx(8, 0, g16_.p, gB_.p);
    // level 1 decoder
    unet_up2_bwd(st, gB_.p, B, L(1, 0), L(1, 1), L(1, 2), 32, gA_.p);
    unet_relu_mask(st, gA_.p, d1_.p, n1 * 32, g16_.p);
    dw(7, 1, s1_.p, g16_.p);
    dx(7, 1, g16_.p, gs1_.p);
    unet_relu_mask(st, gs1_.p, r6_.p, n1 * 32, g16_.p);
    dw(6, 1, u1_.p, g16_.p);
    dx(6, 1, g16_.p, gB_.p);
    // level 2
    unet_up2_bwd(st, gB_.p, B, L(2, 0), L(2, 1), L(2, 2), 64, gA_.p);
    unet_relu_mask(st, gA_.p, e2_.p, n2 * 64, g16_.p);
    dw(5, 2, a2_.p, g16_.p);
    dx(5, 2, g16_.p, gB_.p);
    unet_relu_mask(st, gB_.p, a2_.p, n2 * 64, g16_.p);
    dw(4, 2, p1_.p, g16_.p);
    dx(4, 2, g16_.p, gA_.p);
    // level 1 encoder (the skip into s1 adds gs1)
    unet_pool2_bwd(st, gA_.p, gs1_.p, B, L(1, 0), L(1, 1), L(1, 2), 32, gB_.p);
    unet_relu_mask(st, gB_.p, e1_.p, n1 * 32, g16_.p);
    dw(3, 1, a1_.p, g16_.p);
    dx(3, 1, g16_.p, gA_.p);
    unet_relu_mask(st, gA_.p, a1_.p, n1 * 32, g16_.p);
    dw(2, 1, p0_.p, g16_.p);
    dx(2, 1, g16_.p, gB_.p);
    // level 0 encoder (the skip into s0 adds gs0)
    unet_pool2_bwd(st, gB_.p, gs0_.p, B, L(0, 0), L(0, 1), L(0, 2), 16, gA_.p);
    unet_relu_mask(st, gA_.p, e0_.p, n0 * 16, g16_.p);
    dw(1, 0, a0_.p, g16_.p);
    dx(1, 0, g16_.p, gB_.p);
    unet_relu_mask(st, gB_.p, a0_.p, n0 * 16, g16_.p);
    dw(0, 0, x8, g16_.p);
    if (host_sums) {
        host_sums[0] = mse[0];
        host_sums[1] = mse[1];
    }
    return mse[1] > 0 ? mse[0] / mse[1] : 0.0;
}

void UNet::adam_step(hipStream_t st, float lr) {
    step_++;
    unet_adam(st, w_.p, grad_.p, m_.p, v_.p, w_.n, lr, step_);
    cast_weights(st);
}

}  // namespace sfp
