// Mesh module (placeholder until the binning kernels land).
#include "ncg_internal.hpp"

extern "C" ncg_status ncg_mesh_pass(ncg_ctx *ctx, const ncg_body *, int32_t, const int32_t *,
                                    double, double, double, int32_t, int32_t,
                                    const ncg_replay_ops *, ncg_mesh_stats *)
{
    if (!ctx) return NCG_ERR_INVALID;
    return ncg_fail(ctx, NCG_ERR_UNSUPPORTED, "mesh pass not built yet");
}

extern "C" ncg_status ncg_mesh_gravity(ncg_ctx *ctx, const ncg_body *, int32_t, const int32_t *,
                                       const ncg_central *, double, double, ncg_precision,
                                       double *, double *)
{
    if (!ctx) return NCG_ERR_INVALID;
    return ncg_fail(ctx, NCG_ERR_UNSUPPORTED, "mesh gravity not built yet");
}

extern "C" ncg_status ncg_mesh_pairs(ncg_ctx *ctx, ncg_pair *, int64_t, int64_t *)
{
    if (!ctx) return NCG_ERR_INVALID;
    return ncg_fail(ctx, NCG_ERR_UNSUPPORTED, "mesh pairs not built yet");
}
