/*
 * sfp_profile.h -- per-kernel timing of the solver's hot kernels, measured
 * with HIP events recorded around each launch on the context's stream.
 * No counterpart in the reference (it has no profiler hooks; its StopWatch,
 * core/StopWatch.h:6-47, times whole phases on the host).
 */
#ifndef SFP_PROFILE_H
#define SFP_PROFILE_H
#include "sfp.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef enum sfp_kernel_id {
    SFP_KERNEL_SMOOTH = 0,   /* fine-level single Chebyshev steps (all but the fused first), FP32 */
    SFP_KERNEL_MATVEC64 = 1, /* flexible-CG matvec fused with its inner products, FP64 */
    SFP_KERNEL_UPDATE64 = 2, /* flexible-CG vector update, FP64 */
    SFP_KERNEL_RESTRICT = 3, /* fine-level residual + restriction, FP32 */
    SFP_KERNEL_SMOOTH_PAIR = 4, /* fine-level pairs of Chebyshev steps in one pass, FP32 */
    SFP_KERNEL_SMOOTH_RESTRICT = 5 /* fine-level last pre-smoothing step + restriction, FP32 */
} sfp_kernel_id;

typedef struct sfp_kernel_profile {
    long long launches; /* launches timed since sfp_profile_enable */
    double total_ms;    /* summed event-to-event time */
    double bytes;       /* summed algorithmic bytes (each distinct array once per launch) */
} sfp_kernel_profile;

/* Start (enable != 0) or stop timing; either way the counters are reset. */
sfp_status sfp_profile_enable(sfp_context ctx, int enable);
sfp_status sfp_profile_query(sfp_context ctx, int kernel, sfp_kernel_profile *out);

#ifdef __cplusplus
}
#endif
#endif /* SFP_PROFILE_H */
