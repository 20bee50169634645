// frsdg_ctx.hpp — the library context shared by the translation units of
// libfrsdg.so, and small host helpers.
#pragma once
#include "../../include/frsdg.h"
#include "frsdg_kernels.hpp"
#include "frsdg_limiter.hpp"
#include "frsdg_transform.hpp"
#include "frsdg_tables.hpp"

#include <hip/hip_runtime.h>

#include <cstdint>

struct frsdg_ctx
{
    using Tables = frsdg::Tables;
    using Res2DP = frsdg::Res2DP;
    using Res1DP = frsdg::Res1DP;
    // problem and setup tables
    frsdg_desc d;
    int        n   = 0;
    long long  E   = 0;
    long long  nd  = 0; // number of DOFs
    Tables     tb;
    double*    d_tab  = nullptr; // D, dL, dR, filt
    double*    d_vmet = nullptr;
    double*    d_fxm  = nullptr;
    double*    d_fym  = nullptr;
    double*    d_csi  = nullptr;
    double*    d_rinv = nullptr; // 1/radius per node (radial coordinate system)
    Res2DP     p2{};
    Res1DP     p1{};
    // lazily allocated scratch
    double* w1  = nullptr; // RK work states
    double* w2  = nullptr;
    double* red = nullptr; // reduction partials
    double* wb  = nullptr; // adjoint lifted face cotangents [E][4][n][4]
    double* wbeta = nullptr; // WENO smoothness of every element line [E][2][n][4]
    int*      health          = nullptr; // stage-output health flags of explicit stepping (device)
    int*      health_chk      = nullptr; // frsdg_check_state flags (device)
    int       stop_mask       = 1;       // health flags that stop frsdg_ssprk_run (default: non-finite)
    double*   adj_pool        = nullptr; // unsteady adjoint: snapshots + work states
    long long adj_pool_states = 0;
    // WENO limiter constants (apply_weno without the spectral filter)
    frsdg::LimP lim{};
    // node addressing of the log-variable change of variables (cond_sys)
    frsdg::NodeMap nm{};
    // pre_jv_eval snapshot
    const double* jv_u     = nullptr;
    double        jv_sigma = 0.0;
    double        jv_t     = 0.0;
    // graph cache for ssprk_run
    hipGraphExec_t gexec = nullptr;
    hipGraph_t     graph = nullptr;
    double*        g_u   = nullptr;
    double         g_dt  = 0.0;
    int            g_st  = 0;
    int            g_n   = 0;
    // Newton-Krylov workspace (frsdg_newton.hip)
    struct SolverWs* ws = nullptr;
};

namespace frsdg
{
    inline int hip_rc(hipError_t e) { return e == hipSuccess ? FRSDG_OK : FRSDG_ERR_HIP; }
#define FR_TRY(x)                                                                                  \
    do                                                                                             \
    {                                                                                              \
        hipError_t _e = (x);                                                                       \
        if (_e != hipSuccess) return FRSDG_ERR_HIP;                                                \
    } while (0)

    inline bool aligned(const void* p) { return (reinterpret_cast<std::uintptr_t>(p) & 15u) == 0; }

    // defined in frsdg_newton.hip
    void free_solver(frsdg_ctx* c);
    // defined in frsdg_capi.hip: F = sigma (x - un) + R(x) in one residual
    // launch (2D, conservative variables); FRSDG_ERR_UNSUPPORTED elsewhere
    int ifunction_be(frsdg_ctx* c, const double* x, const double* un, double sigma, double* F, hipStream_t s);
} // namespace frsdg
