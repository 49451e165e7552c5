// C ABI of the learned preconditioner (include/sfp_learned.h).
#include "../../include/sfp_learned.h"
#include "context.hpp"
#include "learned.hpp"

#include <string>

using namespace sfp;

struct sfp_unet_s {
    int device = 0;
    hipStream_t stream = nullptr;
    UNet *net = nullptr;
    UNetShape shape;
    bool training = false;
    DevBuf<float> in, tgt, msk;
    std::string last_error;
};

namespace {

thread_local std::string g_unet_error;

template <class F>
sfp_status run_net(sfp_unet n, F &&f) {
    if (!n) return SFP_ERR_ARGUMENT;
    try {
        SFP_HIP(hipSetDevice(n->device));
        f();
        n->last_error.clear();
        return SFP_OK;
    } catch (const ArgError &e) {
        n->last_error = e.what();
        return SFP_ERR_ARGUMENT;
    } catch (const StateError &e) {
        n->last_error = e.what();
        return SFP_ERR_STATE;
    } catch (const HipError &e) {
        n->last_error = e.what();
        return SFP_ERR_HIP;
    } catch (const std::exception &e) {
        n->last_error = e.what();
        return SFP_ERR_ARGUMENT;
    }
}

void check_count(long long count) {
    if (count != unet_param_count()) throw ArgError("parameter count does not match the network");
}

}  // namespace

extern "C" {

long long sfp_unet_param_count(void) { return unet_param_count(); }

sfp_status sfp_set_network(sfp_context ctx, const float *params, long long count) {
    if (!ctx) return SFP_ERR_ARGUMENT;
    try {
        SFP_HIP(hipSetDevice(ctx->device));
        if (!params) throw ArgError("null parameters");
        check_count(count);
        if (!ctx->net) ctx->net = std::make_unique<UNet>();
        ctx->net->set_params(ctx->ws.stream, params);
        ctx->ws.net = ctx->net.get();
        ctx->last_error.clear();
        return SFP_OK;
    } catch (const ArgError &e) {
        ctx->last_error = e.what();
        return SFP_ERR_ARGUMENT;
    } catch (const std::exception &e) {
        ctx->last_error = e.what();
        return SFP_ERR_HIP;
    }
}

sfp_status sfp_learned_features(sfp_context ctx, int which, const double *phi, float *out) {
    if (!ctx || !out) return SFP_ERR_ARGUMENT;
    try {
        SFP_HIP(hipSetDevice(ctx->device));
        System *s = nullptr;
        if (which == 0 && ctx->pressure_ready && ctx->pressure_unknowns > 0) s = &ctx->pressure;
        if (which == 1 && ctx->field_ready) s = &ctx->magnetic;
        if (!s) throw StateError("no such system assembled");
        if (s->num_levels() < 2) throw StateError("the hierarchy has a single level");
        const Dims d = s->dims();
        const size_t nc = s->level(1).d.cells();
        DevBuf<double> dphi;
        DevBuf<float> f3;
        f3.ensure(3 * nc);
        if (phi) {
            dphi.ensure(d.cells());
            SFP_HIP(hipMemcpyAsync(dphi.p, phi, d.cells() * sizeof(double), hipMemcpyHostToDevice,
                                   ctx->ws.stream));
        }
        Workspace &ws = ctx->ws;
        const double *keep_phi = ws.phi;
        ws.phi = dphi.p;
        ws.spacing = ctx->spacing;
        try {
            s->learned_features(ws, f3.p);
        } catch (...) {
            ws.phi = keep_phi;
            throw;
        }
        ws.phi = keep_phi;
        SFP_HIP(hipMemcpyAsync(out, f3.p, 3 * nc * sizeof(float), hipMemcpyDeviceToHost, ws.stream));
        SFP_HIP(hipStreamSynchronize(ws.stream));
        return SFP_OK;
    } catch (const StateError &e) {
        ctx->last_error = e.what();
        return SFP_ERR_STATE;
    } catch (const std::exception &e) {
        ctx->last_error = e.what();
        return SFP_ERR_HIP;
    }
}

sfp_status sfp_unet_create(sfp_unet *out, int device, int batch, int nx, int ny, int nz, int training) {
    if (!out) return SFP_ERR_ARGUMENT;
    *out = nullptr;
    if (batch < 1 || nx < 4 || ny < 4 || nz < 4 || nx % 4 || ny % 4 || nz % 4) {
        g_unet_error = "every axis must be a positive multiple of 4";
        return SFP_ERR_ARGUMENT;
    }
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || device < 0 || device >= count) {
        g_unet_error = "no such HIP device";
        return SFP_ERR_NO_DEVICE;
    }
    sfp_unet n = new sfp_unet_s;
    n->device = device;
    n->shape = UNetShape{batch, nx, ny, nz};
    n->training = training != 0;
    try {
        SFP_HIP(hipSetDevice(device));
        SFP_HIP(hipStreamCreateWithFlags(&n->stream, hipStreamNonBlocking));
        n->net = new UNet();
        n->net->ensure(n->shape, n->training);
        const size_t nv = n->shape.voxels(0);
        n->in.ensure(nv * 4);
        n->tgt.ensure(nv);
        n->msk.ensure(nv);
    } catch (const std::exception &e) {
        g_unet_error = e.what();
        sfp_unet_destroy(n);
        return SFP_ERR_HIP;
    }
    *out = n;
    return SFP_OK;
}

void sfp_unet_destroy(sfp_unet n) {
    if (!n) return;
    (void)hipSetDevice(n->device);
    if (n->stream) (void)hipStreamSynchronize(n->stream);
    delete n->net;
    hipStream_t s = n->stream;
    delete n;
    if (s) (void)hipStreamDestroy(s);
}

sfp_status sfp_unet_set_params(sfp_unet n, const float *params, long long count) {
    return run_net(n, [&] {
        if (!params) throw ArgError("null parameters");
        check_count(count);
        n->net->set_params(n->stream, params);
    });
}

sfp_status sfp_unet_get_params(sfp_unet n, float *params, long long count) {
    return run_net(n, [&] {
        if (!params) throw ArgError("null parameters");
        check_count(count);
        n->net->get_params(n->stream, params);
    });
}

sfp_status sfp_unet_init_params(sfp_unet n, unsigned long long seed) {
    return run_net(n, [&] { n->net->init_random(n->stream, seed); });
}

sfp_status sfp_unet_forward(sfp_unet n, const float *input, float *output) {
    return run_net(n, [&] {
        if (!input || !output) throw ArgError("null array");
        const size_t nv = n->shape.voxels(0);
        SFP_HIP(hipMemcpyAsync(n->in.p, input, nv * 4 * sizeof(float), hipMemcpyHostToDevice, n->stream));
        unet_pack_input(n->stream, n->in.p, nv, n->net->input_buffer());
        n->net->forward(n->stream, n->net->input_buffer(), n->net->output_buffer());
        SFP_HIP(hipMemcpyAsync(output, n->net->output_buffer(), nv * sizeof(float), hipMemcpyDeviceToHost,
                               n->stream));
        SFP_HIP(hipStreamSynchronize(n->stream));
    });
}

sfp_status sfp_unet_profile_forward(sfp_unet n, int reps, double *forward_ms, double *conv_ms) {
    return run_net(n, [&] {
        if (reps < 1 || !forward_ms || !conv_ms) throw ArgError("reps >= 1 and non-null outputs");
        hipEvent_t ev[2 * kUNetConvs + 2] = {};
        for (auto &e : ev) SFP_HIP(hipEventCreate(&e));
        struct Cleanup {
            hipEvent_t *e;
            UNet *net;
            ~Cleanup() {
                net->set_conv_events(nullptr);
                for (int q = 0; q < 2 * kUNetConvs + 2; q++) (void)hipEventDestroy(e[q]);
            }
        } cleanup{ev, n->net};
        double conv[kUNetConvs] = {0}, total = 0;
        n->net->set_conv_events(ev);
        for (int r = 0; r < reps; r++) {
            SFP_HIP(hipEventRecord(ev[2 * kUNetConvs], n->stream));
            n->net->forward(n->stream, n->net->input_buffer(), n->net->output_buffer());
            SFP_HIP(hipEventRecord(ev[2 * kUNetConvs + 1], n->stream));
            SFP_HIP(hipEventSynchronize(ev[2 * kUNetConvs + 1]));
            float ms = 0.f;
            SFP_HIP(hipEventElapsedTime(&ms, ev[2 * kUNetConvs], ev[2 * kUNetConvs + 1]));
            total += ms;
            for (int k = 0; k < kUNetConvs; k++) {
                SFP_HIP(hipEventElapsedTime(&ms, ev[2 * k], ev[2 * k + 1]));
                conv[k] += ms;
            }
        }
        *forward_ms = total / reps;
        for (int k = 0; k < kUNetConvs; k++) conv_ms[k] = conv[k] / reps;
    });
}

sfp_status sfp_unet_gradients(sfp_unet n, const float *input, const float *target, const float *mask,
                              double *loss, float *grads) {
    return run_net(n, [&] {
        if (!input || !target || !grads) throw ArgError("null array");
        if (!n->training) throw StateError("network created for inference");
        const size_t nv = n->shape.voxels(0);
        SFP_HIP(hipMemcpyAsync(n->in.p, input, nv * 4 * sizeof(float), hipMemcpyHostToDevice, n->stream));
        SFP_HIP(hipMemcpyAsync(n->tgt.p, target, nv * sizeof(float), hipMemcpyHostToDevice, n->stream));
        if (mask)
            SFP_HIP(hipMemcpyAsync(n->msk.p, mask, nv * sizeof(float), hipMemcpyHostToDevice, n->stream));
        unet_pack_input(n->stream, n->in.p, nv, n->net->input_buffer());
        const double l = n->net->loss_and_backward(n->stream, n->net->input_buffer(), n->tgt.p,
                                                   mask ? n->msk.p : nullptr, nullptr);
        SFP_HIP(hipMemcpyAsync(grads, n->net->grads(), unet_param_count() * sizeof(float),
                               hipMemcpyDeviceToHost, n->stream));
        SFP_HIP(hipStreamSynchronize(n->stream));
        if (loss) *loss = l;
    });
}

sfp_status sfp_unet_train_step_device(sfp_unet n, const float *input, const float *target,
                                      const float *mask, float lr, double *loss) {
    return run_net(n, [&] {
        if (!input || !target) throw ArgError("null array");
        if (!n->training) throw StateError("network created for inference");
        const size_t nv = n->shape.voxels(0);
        unet_pack_input(n->stream, input, nv, n->net->input_buffer());
        const double l = n->net->loss_and_backward(n->stream, n->net->input_buffer(), target, mask, nullptr);
        n->net->adam_step(n->stream, lr);
        SFP_HIP(hipStreamSynchronize(n->stream));
        if (loss) *loss = l;
    });
}

const char *sfp_unet_last_error(sfp_unet n) { return n ? n->last_error.c_str() : g_unet_error.c_str(); }

}  // extern "C"
