/*
 * gp_ops.hip -- one translation unit per supported (phases, mu fields) pair: the build compiles this file once per
 * pair of GP_COMBOS (gp_internal.h) with -DGP_N=<phases> -DGP_NC=<mu fields>, and gp_hip.hip reaches the kernels of
 * a handle through gp_ops_<N>_<NC>.
 */
#include "gp_step_impl.h"

#if !defined(GP_N) || !defined(GP_NC)
#error "gp_ops.hip is compiled with -DGP_N=<phases> -DGP_NC=<mu fields>"
#endif

#define GP_OPS_NAME2(a, b) gp_ops_##a##_##b
#define GP_OPS_NAME(a, b) GP_OPS_NAME2(a, b)

const GpOps *GP_OPS_NAME(GP_N, GP_NC)(int precision) { return gpk::make_ops<GP_N, GP_NC>(precision); }
