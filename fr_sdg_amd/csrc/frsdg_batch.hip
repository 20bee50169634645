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
        int*          health; // per-instance health flags of the stored stage outputs (device)
    };

    // Work list of one launch: tile ranges of its instances.  Up to KMAX
    // entries travel by value in the kernel arguments (the per-tile instance
    // search then reads the constant cache instead of issuing dependent HBM
    // loads); longer lists are searched in device memory.
    constexpr int KMAX = 64;
    // mode 2 chunk size: three stage states of the chunk's instances (about
    // half of the 256 MB last-level cache)
    constexpr long long CHUNK_BYTES = 128LL << 20;
    struct WorkList
    {
        int                    ng;
        long long              total;
        const int*       __restrict__ d_idx; // ng entries (device), used when ng > KMAX
        const long long* __restrict__ d_t0;
        long long              t0[KMAX];
        int                    idx[KMAX];
    };
    // instance slot k and local tile of global tile t (uniform across the
    // workgroup; the results are pinned to scalar registers so that the
    // instance record is read with scalar loads)
    __device__ __forceinline__ int work_find_any(const WorkList& W, long long t, long long& local);
    __device__ __forceinline__ int work_find(const WorkList& W, long long t, long long& local)
    {
        long long  l;
        const int  k  = work_find_any(W, t, l);
        const auto lo = static_cast<unsigned>(static_cast<unsigned long long>(l));
        const auto hi = static_cast<unsigned>(static_cast<unsigned long long>(l) >> 32);
        local         = static_cast<long long>((static_cast<unsigned long long>(__builtin_amdgcn_readfirstlane(hi)) << 32) |
                                               __builtin_amdgcn_readfirstlane(lo));
        return __builtin_amdgcn_readfirstlane(k);
    }
    __device__ __forceinline__ int work_find_any(const WorkList& W, long long t, long long& local)
    {
        int lo = 0, hi = W.ng - 1;
        if (W.ng <= KMAX)
        {
            while (lo < hi)
            {
                const int mid = (lo + hi + 1) >> 1;
                if (W.t0[mid] <= t)
                    lo = mid;
                else
                    hi = mid - 1;
            }
            local = t - W.t0[lo];
            return W.idx[lo];
        }
        while (lo < hi)
        {
            const int mid = (lo + hi + 1) >> 1;
            if (W.d_t0[mid] <= t)
                lo = mid;
            else
                hi = mid - 1;
        }
        local = t - W.d_t0[lo];
        return W.d_idx[lo];
    }

    // Instance record k through the constant address space, field by field:
    // every field becomes a scalar load of a uniform address.  A plain copy of
    // tab[k] compiles to ~44 vector loads per wavefront (the record is ~300
    // bytes) ahead of the tile's own loads.  The host pass of the
    // single-source compile has no address spaces and takes the plain copy.
    __device__ __forceinline__ BatchInst load_inst(const BatchInst* tab, int k)
    {
#if defined(__HIP_DEVICE_COMPILE__)
        typedef const __attribute__((address_space(4))) BatchInst CInst;
        CInst*    c = (CInst*) (tab + k);
        BatchInst r;
        r.P.nx      = c->P.nx;
        r.P.ny      = c->P.ny;
        r.P.E       = c->P.E;
        r.P.per_x   = c->P.per_x;
        r.P.per_y   = c->P.per_y;
        r.P.curved  = c->P.curved;
        r.P.riemann = c->P.riemann;
        r.P.xi_x    = c->P.xi_x;
        r.P.eta_y   = c->P.eta_y;
        r.P.sc_x    = c->P.sc_x;
        r.P.sc_y    = c->P.sc_y;
        r.P.vmet    = c->P.vmet;
        r.P.fxm     = c->P.fxm;
        r.P.fym     = c->P.fym;
        r.P.tab     = c->P.tab;
        r.P.gas.kind  = c->P.gas.kind;
        r.P.gas.gamma = c->P.gas.gamma;
        r.P.gas.a     = c->P.gas.a;
        r.P.gas.b     = c->P.gas.b;
#pragma unroll
        for (int f = 0; f < 4; ++f)
        {
            r.P.bc.type[f] = c->P.bc.type[f];
#pragma unroll
            for (int q = 0; q < 4; ++q) r.P.bc.state[f][q] = c->P.bc.state[f][q];
        }
        r.P.bc.freeze = c->P.bc.freeze;
        r.P.rinv      = c->P.rinv;
        r.off         = c->off;
        r.tile0       = c->tile0;
        r.ntiles      = c->ntiles;
        r.n           = c->n;
        r.tx          = c->tx;
        r.ty          = c->ty;
        r.filt        = c->filt;
        r.curved      = c->curved;
        r.pad         = 0;
        return r;
#else
        return tab[k];
#endif
    }

    template <bool FILT>
    __global__ __launch_bounds__(256) void k_batch_stage(const BatchInst* __restrict__ tab, const WorkList W,
                                                         StageArgs A)
    {
        __shared__ BatchSmem sm;
        const long long slot = xcd_tile(blockIdx.x, gridDim.x);
        for (long long t = slot; t < W.total; t += gridDim.x)
        {
            long long       blk;
            const int       k   = work_find(W, t, blk);
            const BatchInst I   = load_inst(tab, k);
            const IOStage    io  = {A.us + I.off, A.u0 + I.off, A.out + I.off, A.a0, A.b, A.cfac * A.dt[k], A.health + k};
            const bool       flt = A.last && I.filt;
#define FR_BT(NN)                                                                                            \
    case NN:                                                                                                 \
        if (I.curved)                                                                                        \
            res2d_tile<NN, IOStage, true, FILT>(I.P, io, blk, I.tx, I.ty, flt, sm.c##NN);                          \
        else                                                                                                 \
            res2d_tile<NN, IOStage, false, FILT>(I.P, io, blk, I.tx, I.ty, flt, sm.a##NN);                         \
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
    // the same with one workgroup per tile (grid = tiles of the group) and the
    // tile shape fixed at compile time (SHAPE: batch_tile_dim's choice): no
    // loop and no run-time shape arithmetic, so the register budget is the
    // single-problem kernel's (≤ 128 VGPRs, 4 waves per SIMD; a cap of 96 for
    // 5 waves spills 60-88 bytes per thread at n = 5, 7, 8)
    template <int N, bool CURVED, bool FILT, int SHAPE>
    __global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4))) void k_batch_tile_n(const BatchInst* __restrict__ tab, const WorkList W,
                                                          StageArgs A)
    {
        using T          = Tile2D<N, double>;
        constexpr int TX = batch_tile_dim(N, SHAPE, true), TY = batch_tile_dim(N, SHAPE, false);
        static_assert(TX * TY <= T::EPB, "tile larger than the LDS of the order");
        __shared__ TileSmem<N, double, CURVED> sm;
        long long       blk;
        const int       k  = work_find(W, xcd_tile(blockIdx.x, gridDim.x), blk);
        const BatchInst I  = load_inst(tab, k);
        const IOStage   io = {A.us + I.off, A.u0 + I.off, A.out + I.off, A.a0, A.b, A.cfac * A.dt[k], A.health + k};
        res2d_tile<N, IOStage, CURVED, FILT>(I.P, io, blk, TX, TY, I.filt, sm);
    }

    template <int N, bool CURVED, bool FILT>
    __global__ __launch_bounds__(256) void k_batch_stage_n(const BatchInst* __restrict__ tab, const WorkList W,
                                                           StageArgs A)
    {
        __shared__ TileSmem<N, double, CURVED> sm;
        const long long slot = xcd_tile(blockIdx.x, gridDim.x);
        for (long long t = slot; t < W.total; t += gridDim.x)
        {
            long long       blk;
            const int       k  = work_find(W, t, blk);
            const BatchInst I  = load_inst(tab, k);
            const IOStage    io = {A.us + I.off, A.u0 + I.off, A.out + I.off, A.a0, A.b, A.cfac * A.dt[k], A.health + k};
            res2d_tile<N, IOStage, CURVED, FILT>(I.P, io, blk, I.tx, I.ty, I.filt, sm);
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
    int*                    d_health = nullptr; // per instance: health flags of its stage outputs (frsdg_batch_health)
    long long               total_tiles = 0;
    long long               nd          = 0;
    int                     grid        = 0; // persistent workgroups (0: one workgroup per tile)
    int                     shape[10]   = {}; // element tile shape choice per order n (batch_tile_dim)
    hipGraph_t              graph = nullptr;
    hipGraphExec_t          gexec = nullptr;
    double*                 g_u   = nullptr;
    int                     g_ok  = 0;
    int                     mode  = 1; // 0 one launch for all orders, 1 one launch per order
    // launch groups: all instances (mode 0), or one per (order, mesh kind) (mode 1)
    struct Group
    {
        int        n      = 0;
        int        curved = 0;
        WorkList   W{};
        int*       d_idx = nullptr;
        long long* d_t0  = nullptr;
    };
    Group              all;
    std::vector<Group> groups;
    std::vector<Group> chunks; // mode 2: groups cut into cache-sized runs of instances
};

namespace
{
    void free_groups(frsdg_batch* b)
    {
        auto rel = [](frsdg_batch::Group& g)
        {
            if (g.d_idx) (void) hipFree(g.d_idx);
            if (g.d_t0) (void) hipFree(g.d_t0);
            g.d_idx = nullptr;
            g.d_t0  = nullptr;
        };
        rel(b->all);
        for (auto& g : b->groups) rel(g);
        for (auto& g : b->chunks) rel(g);
        b->groups.clear();
        b->chunks.clear();
    }

    int rebuild_tiles(frsdg_batch* b)
    {
        long long t = 0;
        for (auto& I : b->inst)
        {
            I.tx = batch_tile_dim(I.n, b->shape[I.n], true);
            I.ty = batch_tile_dim(I.n, b->shape[I.n], false);
            I.tile0  = t;
            I.ntiles = tile_count(I.P.nx, I.P.ny, I.tx, I.ty);
            t += I.ntiles;
        }
        b->total_tiles = t;
        FR_TRY(hipMemcpy(b->d_tab, b->inst.data(), sizeof(BatchInst) * b->inst.size(), hipMemcpyHostToDevice));
        free_groups(b);
        // fills a group's work list from the instances selected by `pick`
        auto make = [&](frsdg_batch::Group& g, auto pick) -> int
        {
            std::vector<int>       idx;
            std::vector<long long> t0;
            long long              tot = 0;
            for (size_t k = 0; k < b->inst.size(); ++k)
                if (pick(b->inst[k]))
                {
                    idx.push_back(static_cast<int>(k));
                    t0.push_back(tot);
                    tot += b->inst[k].ntiles;
                }
            g.W.ng    = static_cast<int>(idx.size());
            g.W.total = tot;
            for (size_t q = 0; q < idx.size() && q < static_cast<size_t>(KMAX); ++q)
            {
                g.W.t0[q]  = t0[q];
                g.W.idx[q] = idx[q];
            }
            if (idx.empty()) return FRSDG_OK;
            FR_TRY(hipMalloc(&g.d_idx, sizeof(int) * idx.size()));
            FR_TRY(hipMalloc(&g.d_t0, sizeof(long long) * t0.size()));
            FR_TRY(hipMemcpy(g.d_idx, idx.data(), sizeof(int) * idx.size(), hipMemcpyHostToDevice));
            FR_TRY(hipMemcpy(g.d_t0, t0.data(), sizeof(long long) * t0.size(), hipMemcpyHostToDevice));
            g.W.d_idx = g.d_idx;
            g.W.d_t0  = g.d_t0;
            return FRSDG_OK;
        };
        {
            const int rc = make(b->all, [](const BatchInst&) { return true; });
            if (rc) return rc;
        }
        for (int key = 0; key < 16; ++key)
        {
            const int          n = 2 + (key >> 1), cv = key & 1;
            frsdg_batch::Group g;
            g.n      = n;
            g.curved = cv;
            const int rc = make(g, [&](const BatchInst& I) { return I.n == n && I.curved == cv; });
            if (g.W.ng > 0) b->groups.push_back(g);
            if (rc) return rc;
            // cache-sized chunks of this group: consecutive instances whose three
            // stage states (u, w1, w2) together stay within CHUNK_BYTES
            size_t k0 = 0;
            while (true)
            {
                while (k0 < b->inst.size() && !(b->inst[k0].n == n && b->inst[k0].curved == cv)) ++k0;
                if (k0 >= b->inst.size()) break;
                size_t    k1    = k0;
                long long bytes = 0;
                while (k1 < b->inst.size())
                {
                    const BatchInst& I = b->inst[k1];
                    if (!(I.n == n && I.curved == cv))
                    {
                        ++k1;
                        continue;
                    }
                    const long long ib = 24LL * (b->off[k1 + 1] - b->off[k1]);
                    if (bytes > 0 && bytes + ib > CHUNK_BYTES) break;
                    bytes += ib;
                    ++k1;
                }
                frsdg_batch::Group c;
                c.n      = n;
                c.curved = cv;
                const size_t lo = k0, hi = k1;
                const int    rc2 = make(c, [&](const BatchInst& I)
                                        {
                                            const size_t k = static_cast<size_t>(&I - b->inst.data());
                                            return k >= lo && k < hi && I.n == n && I.curved == cv;
                                        });
                if (c.W.ng > 0) b->chunks.push_back(c);
                if (rc2) return rc2;
                k0 = k1;
            }
        }
        if (b->gexec) (void) hipGraphExecDestroy(b->gexec);
        if (b->graph) (void) hipGraphDestroy(b->graph);
        b->gexec = nullptr;
        b->graph = nullptr;
        return FRSDG_OK;
    }

    int batch_step(frsdg_batch* b, double* u, hipStream_t s)
    {
        // one stage over the instances of group g
        auto launch_group = [&](const frsdg_batch::Group& g, const StageArgs& A, int last)
        {
            const WorkList& W  = g.W;
            const unsigned  gr = static_cast<unsigned>(b->grid > 0 ? std::min<long long>(b->grid, W.total) : W.total);
#define FR_BK(K, ...) hipLaunchKernelGGL((K<__VA_ARGS__>), dim3(gr), dim3(256), 0, s, b->d_tab, W, A)
#define FR_BT(NN, C)                                                                                          \
    {                                                                                                         \
        if (b->grid > 0)                                                                                      \
        {                                                                                                     \
            if (last) FR_BK(k_batch_stage_n, NN, C, true); else FR_BK(k_batch_stage_n, NN, C, false);         \
        }                                                                                                     \
        else if (b->shape[NN] == 0)                                                                           \
        {                                                                                                     \
            if (last) FR_BK(k_batch_tile_n, NN, C, true, 0); else FR_BK(k_batch_tile_n, NN, C, false, 0);     \
        }                                                                                                     \
        else if (b->shape[NN] == 1)                                                                           \
        {                                                                                                     \
            if (last) FR_BK(k_batch_tile_n, NN, C, true, 1); else FR_BK(k_batch_tile_n, NN, C, false, 1);     \
        }                                                                                                     \
        else                                                                                                  \
        {                                                                                                     \
            if (last) FR_BK(k_batch_tile_n, NN, C, true, 2); else FR_BK(k_batch_tile_n, NN, C, false, 2);     \
        }                                                                                                     \
    }
#define FR_BN(NN)                                                                                             \
    case NN:                                                                                                  \
        if (g.curved)                                                                                         \
            FR_BT(NN, true)                                                                                   \
        else                                                                                                  \
            FR_BT(NN, false)                                                                                  \
        break;
            switch (g.n)
            {
                FR_BN(2) FR_BN(3) FR_BN(4) FR_BN(5) FR_BN(6) FR_BN(7) FR_BN(8) FR_BN(9)
            }
#undef FR_BN
#undef FR_BT
#undef FR_BK
            return hip_rc(hipGetLastError());
        };
        const StageArgs S1{u, u, b->w1, b->d_dt, 0.0, 1.0, 1.0, 0, b->d_health};
        const StageArgs S2{b->w1, u, b->w2, b->d_dt, 0.75, 0.25, 1.0, 0, b->d_health};
        const StageArgs S3{b->w2, u, u, b->d_dt, 1.0 / 3.0, 2.0 / 3.0, 1.0, 1, b->d_health};
        int             rc;
        if (b->mode == 2)
        {
            // each cache-sized chunk takes its three stages back to back
            for (const auto& c : b->chunks)
            {
                if ((rc = launch_group(c, S1, 0))) return rc;
                if ((rc = launch_group(c, S2, 0))) return rc;
                if ((rc = launch_group(c, S3, 1))) return rc;
            }
            return FRSDG_OK;
        }
        auto launch = [&](const StageArgs& A, int last)
        {
            if (b->mode == 0)
            {
                const WorkList& W  = b->all.W;
                const unsigned  gr = static_cast<unsigned>(b->grid > 0 ? std::min<long long>(b->grid, W.total) : W.total);
                if (last)
                    hipLaunchKernelGGL(k_batch_stage<true>, dim3(gr), dim3(256), 0, s, b->d_tab, W, A);
                else
                    hipLaunchKernelGGL(k_batch_stage<false>, dim3(gr), dim3(256), 0, s, b->d_tab, W, A);
                return hip_rc(hipGetLastError());
            }
            for (const auto& g : b->groups)
            {
                const int r = launch_group(g, A, last);
                if (r) return r;
            }
            return FRSDG_OK;
        };
        if ((rc = launch(S1, 0))) return rc;
        if ((rc = launch(S2, 0))) return rc;
        return launch(S3, 1);
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
        // 2D explicit stages only: no WENO limiter sweep, no implicit change of variables
        if (descs[k].dims != 2 || (descs[k].apply_weno && !descs[k].apply_filter) || descs[k].cond_sys ||
            descs[k].coord_sys)
            return fail(FRSDG_ERR_UNSUPPORTED);
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
    if (hipMalloc(&b->d_health, sizeof(int) * count) != hipSuccess) return fail(FRSDG_ERR_HIP);
    if (hipMemset(b->d_health, 0, sizeof(int) * count) != hipSuccess) return fail(FRSDG_ERR_HIP);
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
    free_groups(b);
    for (auto* c : b->ctx) frsdg_destroy(c);
    for (void* p : {static_cast<void*>(b->d_tab), static_cast<void*>(b->d_dt), static_cast<void*>(b->w1),
                    static_cast<void*>(b->w2), static_cast<void*>(b->d_health)})
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
    if (!b || grid < 0 || (epb_div != 1 && epb_div != 2) || mode < 0 || mode > 2) return FRSDG_ERR_INVALID;
    b->grid = grid;
    for (int& sh : b->shape) sh = epb_div - 1;
    b->mode = mode;
    return rebuild_tiles(b);
}

int frsdg_batch_set_tile_shape(frsdg_batch* b, int n, int shape)
{
    if (!b || n < 2 || n > 9 || shape < 0 || shape > 2) return FRSDG_ERR_INVALID;
    b->shape[n] = shape;
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
    const int     K  = static_cast<int>(b->ctx.size());
    // the health flags of this run (outside any capture, so a memset is fine);
    // a run stops after the first graph of steps in which an instance raised a
    // flag of its own stop mask (frsdg_set_stop_mask on frsdg_batch_instance)
    FR_TRY(hipMemsetAsync(b->d_health, 0, sizeof(int) * K, s));
    std::vector<int> flags(K);
    auto healthy = [&]() -> int
    {
        FR_TRY(hipMemcpyAsync(flags.data(), b->d_health, sizeof(int) * K, hipMemcpyDeviceToHost, s));
        FR_TRY(hipStreamSynchronize(s));
        for (int k = 0; k < K; ++k)
            if (flags[k] & b->ctx[k]->stop_mask) return FRSDG_ERR_NUMERIC;
        return FRSDG_OK;
    };
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
    int rc;
    for (int k = 0; k < nsteps / CH; ++k)
    {
        FR_TRY(hipGraphLaunch(b->gexec, s));
        if ((rc = healthy())) return rc;
    }
    for (int k = 0; k < nsteps % CH; ++k)
        if ((rc = batch_step(b, u, s))) return rc;
    return healthy();
}

int frsdg_batch_health(frsdg_batch* b, int* flags, int clear, void* stream)
{
    if (!b || !flags) return FRSDG_ERR_INVALID;
    hipStream_t s = static_cast<hipStream_t>(stream);
    const size_t bytes = sizeof(int) * b->ctx.size();
    FR_TRY(hipMemcpyAsync(flags, b->d_health, bytes, hipMemcpyDeviceToHost, s));
    if (clear) FR_TRY(hipMemsetAsync(b->d_health, 0, bytes, s));
    FR_TRY(hipStreamSynchronize(s));
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
