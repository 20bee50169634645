// frsdg_batch.hip — many independent FR problems (uneven mesh size, order and
// gas model) advanced together by persistent kernels that walk the
// concatenated (instance, element-tile) work list with the same tile function
// the single-problem kernels use; the last stage carries the post-step
// spectral filter.  Two launch modes: one kernel per stage over all orders
// (each tile dispatches on its instance's n = p+1; LDS and registers sized for
// the largest case), or one kernel per stage and order (specialised).  A step
// is 3 (or 3 x #orders) launches regardless of the number of instances, and
// frsdg_batch_ssprk_run replays captured steps from a hipGraph.
#include "frsdg_ctx.hpp"

#include <vector>

using namespace frsdg;

namespace
{
    struct BatchInst
    {
        Res2DP    P;
        long long off;    // first DOF of the instance in the concatenated state
        long long tile0;  // first tile index of the instance
        long long ntiles;
        int       n;
        int       tx, ty; // element tile shape
        int       filt;   // post-step filter active
        int       curved;
        int       pad;
    };

    union BatchSmem
    {
        TileSmem<2, double, false> a2;
        TileSmem<3, double, false> a3;
        TileSmem<4, double, false> a4;
        TileSmem<5, double, false> a5;
        TileSmem<6, double, false> a6;
        TileSmem<7, double, false> a7;
        TileSmem<8, double, false> a8;
        TileSmem<9, double, false> a9;
        TileSmem<2, double, true>  c2;
        TileSmem<3, double, true>  c3;
        TileSmem<4, double, true>  c4;
        TileSmem<5, double, true>  c5;
        TileSmem<6, double, true>  c6;
        TileSmem<7, double, true>  c7;
        TileSmem<8, double, true>  c8;
        TileSmem<9, double, true>  c9;
    };

    struct StageArgs
    {
        const double* us;
        const double* u0;
        double*       out;
        const double* dt; // per-instance time step (device)
        double        a0, b, cfac;
        int           last;
    };

    __global__ __launch_bounds__(256) void k_batch_stage(const BatchInst* __restrict__ tab, int ninst,
                                                         long long total, StageArgs A)
    {
        __shared__ BatchSmem sm;
        const long long slot = xcd_tile(blockIdx.x, gridDim.x);
        for (long long t = slot; t < total; t += gridDim.x)
        {
            int lo = 0, hi = ninst - 1;
            while (lo < hi)
            {
                const int mid = (lo + hi + 1) >> 1;
                if (tab[mid].tile0 <= t)
                    lo = mid;
                else
                    hi = mid - 1;
            }
            const BatchInst& I   = tab[lo];
            const IOStage    io  = {A.us + I.off, A.u0 + I.off, A.out + I.off, A.a0, A.b, A.cfac * A.dt[lo]};
            const long long  blk = t - I.tile0;
            const bool       flt = A.last && I.filt;
#define FR_BT(NN)                                                                                            \
    case NN:                                                                                                 \
        if (I.curved)                                                                                        \
            res2d_tile<NN, IOStage, true>(I.P, io, blk, I.tx, I.ty, flt, sm.c##NN);                          \
        else                                                                                                 \
            res2d_tile<NN, IOStage, false>(I.P, io, blk, I.tx, I.ty, flt, sm.a##NN);                         \
        break;
            switch (I.n)
            {
                FR_BT(2) FR_BT(3) FR_BT(4) FR_BT(5) FR_BT(6) FR_BT(7) FR_BT(8) FR_BT(9)
            }
#undef FR_BT
            __syncthreads(); // LDS is reused by the next tile
        }
    }

    // one launch per (order, curvature): the tiles of all instances of the
    // group, fully specialised (registers and LDS sized for this order only)
    template <int N, bool CURVED>
    __global__ __launch_bounds__(256) void k_batch_stage_n(const BatchInst* __restrict__ tab,
                                                           const int* __restrict__ gidx,
                                                           const long long* __restrict__ gtile0, int ng,
                                                           long long total, StageArgs A)
    {
        __shared__ TileSmem<N, double, CURVED> sm;
        const long long slot = xcd_tile(blockIdx.x, gridDim.x);
        for (long long t = slot; t < total; t += gridDim.x)
        {
            int lo = 0, hi = ng - 1;
            while (lo < hi)
            {
                const int mid = (lo + hi + 1) >> 1;
                if (gtile0[mid] <= t)
                    lo = mid;
                else
                    hi = mid - 1;
            }
            const int        k  = gidx[lo];
            const BatchInst& I  = tab[k];
            const IOStage    io = {A.us + I.off, A.u0 + I.off, A.out + I.off, A.a0, A.b, A.cfac * A.dt[k]};
            res2d_tile<N, IOStage, CURVED>(I.P, io, t - gtile0[lo], I.tx, I.ty, A.last && I.filt, sm);
            __syncthreads();
        }
    }
} // namespace

struct frsdg_batch
{
    std::vector<frsdg_ctx*> ctx;
    std::vector<long long>  off;   // count + 1
    std::vector<BatchInst>  inst;
    BatchInst*              d_tab = nullptr;
    double*                 d_dt  = nullptr;
    double*                 w1    = nullptr;
    double*                 w2    = nullptr;
    long long               total_tiles = 0;
    long long               nd          = 0;
    int                     grid        = 0; // persistent workgroups (0: one workgroup per tile)
    int                     epb_div     = 1;
    hipGraph_t              graph = nullptr;
    hipGraphExec_t          gexec = nullptr;
    double*                 g_u   = nullptr;
    int                     g_ok  = 0;
    int                     mode  = 1; // 0 one launch for all orders, 1 one launch per order
    // per-order groups
    struct Group
    {
        int        n;
        int        curved;
        int        ng;
        long long  total;
        int*       d_idx = nullptr;
        long long* d_t0  = nullptr;
    };
    std::vector<Group> groups;
};

namespace
{
    int rebuild_tiles(frsdg_batch* b)
    {
        long long t = 0;
        for (auto& I : b->inst)
        {
            tile_shape(I.n, b->epb_div, I.tx, I.ty);
            I.tile0  = t;
            I.ntiles = tile_count(I.P.nx, I.P.ny, I.tx, I.ty);
            t += I.ntiles;
        }
        b->total_tiles = t;
        FR_TRY(hipMemcpy(b->d_tab, b->inst.data(), sizeof(BatchInst) * b->inst.size(), hipMemcpyHostToDevice));
        for (auto& g : b->groups)
        {
            if (g.d_idx) (void) hipFree(g.d_idx);
            if (g.d_t0) (void) hipFree(g.d_t0);
        }
        b->groups.clear();
        for (int key = 0; key < 16; ++key)
        {
            const int              n = 2 + (key >> 1), cv = key & 1;
            std::vector<int>       idx;
            std::vector<long long> t0;
            long long              tot = 0;
            for (size_t k = 0; k < b->inst.size(); ++k)
                if (b->inst[k].n == n && b->inst[k].curved == cv)
                {
                    idx.push_back(static_cast<int>(k));
                    t0.push_back(tot);
                    tot += b->inst[k].ntiles;
                }
            if (idx.empty()) continue;
            frsdg_batch::Group g;
            g.n      = n;
            g.curved = cv;
            g.ng     = static_cast<int>(idx.size());
            g.total = tot;
            FR_TRY(hipMalloc(&g.d_idx, sizeof(int) * idx.size()));
            FR_TRY(hipMalloc(&g.d_t0, sizeof(long long) * t0.size()));
            FR_TRY(hipMemcpy(g.d_idx, idx.data(), sizeof(int) * idx.size(), hipMemcpyHostToDevice));
            FR_TRY(hipMemcpy(g.d_t0, t0.data(), sizeof(long long) * t0.size(), hipMemcpyHostToDevice));
            b->groups.push_back(g);
        }
        if (b->gexec) (void) hipGraphExecDestroy(b->gexec);
        if (b->graph) (void) hipGraphDestroy(b->graph);
        b->gexec = nullptr;
        b->graph = nullptr;
        return FRSDG_OK;
    }

    int batch_step(frsdg_batch* b, double* u, hipStream_t s)
    {
        const int ninst = static_cast<int>(b->inst.size());
        auto launch     = [&](const double* us, double a0, double bb, double* out, int last)
        {
            StageArgs A{us, u, out, b->d_dt, a0, bb, 1.0, last};
            if (b->mode == 0)
            {
                const unsigned gr = static_cast<unsigned>(b->grid > 0 ? b->grid : b->total_tiles);
                hipLaunchKernelGGL(k_batch_stage, dim3(gr), dim3(256), 0, s, b->d_tab, ninst, b->total_tiles, A);
                return hip_rc(hipGetLastError());
            }
            for (const auto& g : b->groups)
            {
                const unsigned gr = static_cast<unsigned>(b->grid > 0 ? std::min<long long>(b->grid, g.total) : g.total);
#define FR_BN(NN)                                                                                              \
    case NN:                                                                                                   \
        if (g.curved)                                                                                          \
            hipLaunchKernelGGL((k_batch_stage_n<NN, true>), dim3(gr), dim3(256), 0, s, b->d_tab, g.d_idx, g.d_t0, \
                               g.ng, g.total, A);                                                             \
        else                                                                                                   \
            hipLaunchKernelGGL((k_batch_stage_n<NN, false>), dim3(gr), dim3(256), 0, s, b->d_tab, g.d_idx,       \
                               g.d_t0, g.ng, g.total, A);                                                     \
        break;
                switch (g.n)
                {
                    FR_BN(2) FR_BN(3) FR_BN(4) FR_BN(5) FR_BN(6) FR_BN(7) FR_BN(8) FR_BN(9)
                }
#undef FR_BN
            }
            return hip_rc(hipGetLastError());
        };
        int rc;
        if ((rc = launch(u, 0.0, 1.0, b->w1, 0))) return rc;
        if ((rc = launch(b->w1, 0.75, 0.25, b->w2, 0))) return rc;
        return launch(b->w2, 1.0 / 3.0, 2.0 / 3.0, u, 1);
    }
} // namespace

extern "C" {

int frsdg_batch_create(const frsdg_desc* descs, int count, frsdg_batch** out)
{
    if (!descs || count < 1 || !out) return FRSDG_ERR_INVALID;
    *out   = nullptr;
    auto* b = new frsdg_batch{};
    auto fail = [&](int rc)
    {
        frsdg_batch_destroy(b);
        return rc;
    };
    b->off.push_back(0);
    for (int k = 0; k < count; ++k)
    {
        if (descs[k].dims != 2) return fail(FRSDG_ERR_UNSUPPORTED);
        frsdg_ctx* c = nullptr;
        int rc = frsdg_create(&descs[k], &c);
        if (rc) return fail(rc);
        b->ctx.push_back(c);
        BatchInst I{};
        I.P    = c->p2;
        I.off  = b->off.back();
        I.n      = c->n;
        I.filt   = descs[k].apply_filter;
        I.curved = c->p2.curved;
        b->inst.push_back(I);
        b->off.push_back(b->off.back() + c->nd);
    }
    b->nd = b->off.back();
    if (hipMalloc(&b->d_tab, sizeof(BatchInst) * count) != hipSuccess) return fail(FRSDG_ERR_HIP);
    if (hipMalloc(&b->d_dt, sizeof(double) * count) != hipSuccess) return fail(FRSDG_ERR_HIP);
    if (hipMalloc(&b->w1, sizeof(double) * b->nd) != hipSuccess) return fail(FRSDG_ERR_HIP);
    if (hipMalloc(&b->w2, sizeof(double) * b->nd) != hipSuccess) return fail(FRSDG_ERR_HIP);
    int dev = 0, ncu = 256;
    (void) hipGetDevice(&dev);
    (void) hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev);
    b->grid = 4 * ncu;
    int rc = rebuild_tiles(b);
    if (rc) return fail(rc);
    *out = b;
    return FRSDG_OK;
}

int frsdg_batch_destroy(frsdg_batch* b)
{
    if (!b) return FRSDG_OK;
    if (b->gexec) (void) hipGraphExecDestroy(b->gexec);
    if (b->graph) (void) hipGraphDestroy(b->graph);
    for (auto& g : b->groups)
    {
        if (g.d_idx) (void) hipFree(g.d_idx);
        if (g.d_t0) (void) hipFree(g.d_t0);
    }
    for (auto* c : b->ctx) frsdg_destroy(c);
    for (void* p : {static_cast<void*>(b->d_tab), static_cast<void*>(b->d_dt), static_cast<void*>(b->w1),
                    static_cast<void*>(b->w2)})
        if (p) (void) hipFree(p);
    delete b;
    return FRSDG_OK;
}

long long frsdg_batch_ndofs(const frsdg_batch* b) { return b ? b->nd : -1; }

int frsdg_batch_offsets(const frsdg_batch* b, long long* off)
{
    if (!b || !off) return FRSDG_ERR_INVALID;
    for (size_t k = 0; k < b->off.size(); ++k) off[k] = b->off[k];
    return FRSDG_OK;
}

int frsdg_batch_instance(frsdg_batch* b, int k, frsdg_ctx** out)
{
    if (!b || !out || k < 0 || k >= static_cast<int>(b->ctx.size())) return FRSDG_ERR_INVALID;
    *out = b->ctx[k];
    return FRSDG_OK;
}

int frsdg_batch_set_tuning(frsdg_batch* b, int grid, int epb_div, int mode)
{
    if (!b || grid < 0 || epb_div < 1 || epb_div > 64 || mode < 0 || mode > 1) return FRSDG_ERR_INVALID;
    b->grid    = grid;
    b->epb_div = epb_div;
    b->mode    = mode;
    return rebuild_tiles(b);
}

int frsdg_batch_set_dt(frsdg_batch* b, const double* dt)
{
    if (!b || !dt) return FRSDG_ERR_INVALID;
    FR_TRY(hipMemcpy(b->d_dt, dt, sizeof(double) * b->inst.size(), hipMemcpyHostToDevice));
    return FRSDG_OK;
}

int frsdg_batch_ssprk_step(frsdg_batch* b, double* u, void* stream)
{
    if (!b || !u) return FRSDG_ERR_INVALID;
    if (!aligned(u)) return FRSDG_ERR_ALIGN;
    return batch_step(b, u, static_cast<hipStream_t>(stream));
}

int frsdg_batch_ssprk_run(frsdg_batch* b, double* u, int nsteps, void* stream)
{
    if (!b || !u || nsteps < 0) return FRSDG_ERR_INVALID;
    if (!aligned(u)) return FRSDG_ERR_ALIGN;
    hipStream_t s = static_cast<hipStream_t>(stream);
    constexpr int CH = 8; // steps per captured graph
    if (!b->gexec || b->g_u != u)
    {
        if (b->gexec) (void) hipGraphExecDestroy(b->gexec);
        if (b->graph) (void) hipGraphDestroy(b->graph);
        b->gexec = nullptr;
        b->graph = nullptr;
        hipStream_t cs;
        FR_TRY(hipStreamCreateWithFlags(&cs, hipStreamNonBlocking));
        if (hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal) != hipSuccess)
        {
            (void) hipStreamDestroy(cs);
            return FRSDG_ERR_HIP;
        }
        int rr = FRSDG_OK;
        for (int k = 0; k < CH && !rr; ++k) rr = batch_step(b, u, cs);
        hipGraph_t g = nullptr;
        const hipError_t ec = hipStreamEndCapture(cs, &g);
        (void) hipStreamDestroy(cs);
        if (rr || ec != hipSuccess) return rr ? rr : FRSDG_ERR_HIP;
        b->graph = g;
        FR_TRY(hipGraphInstantiate(&b->gexec, g, nullptr, nullptr, 0));
        b->g_u = u;
    }
    for (int k = 0; k < nsteps / CH; ++k) FR_TRY(hipGraphLaunch(b->gexec, s));
    for (int k = 0; k < nsteps % CH; ++k)
    {
        int rc = batch_step(b, u, s);
        if (rc) return rc;
    }
    return FRSDG_OK;
}

// Reference point for the batched sweep: the same step with one set of
// launches per instance (3 stage launches each), using each instance's own
// context and work buffers.
int frsdg_batch_ssprk_step_each(frsdg_batch* b, const double* dt, double* u, void* stream)
{
    if (!b || !u || !dt) return FRSDG_ERR_INVALID;
    for (size_t k = 0; k < b->ctx.size(); ++k)
    {
        int rc = frsdg_ssprk_step(b->ctx[k], 0.0, dt[k], u + b->off[k], 3, stream);
        if (rc) return rc;
    }
    return FRSDG_OK;
}

} // extern "C"
