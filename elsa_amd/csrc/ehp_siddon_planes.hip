// ehp_siddon_planes.hip -- Siddon forward projection over plane-major volume views (FP32, FP64).
//
// The forward of Siddon's method (elsa/projectors/SiddonsMethod.cpp:76-124, forward :115-117;
// traversal of TraverseAABB.cpp:6-184) with one lane per ray, reading the zero-padded
// plane-major views of the Joseph forward (that axis slowest, the fan-out axis fastest). A
// wavefront holds 8 x 8 neighbouring pixels; its rays are grouped by driving axis A (their
// largest |d|) and step direction along it, and each group walks the planes of A together:
// on plane p every ray takes its own traversal steps whose voxel lies on p (sid_walk_plane,
// at most 3), so all lanes read the same plane at neighbouring positions. One loop per lane
// would put rays that enter through a side face, each at its own plane, on a cache line of
// their own at every step. The steps are siddon_traverse's, in its order, so the sum is the
// one-thread-per-ray kernel's bit for bit, in FP32 and in FP64 (SiddonsMethodCUDA<double>: FP64
// views and sums, the FP32 lengths of the FP32 traversal, as that kernel forms them). The
// traversal state of every ray comes from the per-operator table of the tiled back projection.
#include "ehp_waves.hpp"

namespace ehp
{
    namespace
    {
        // The group's rays walk the planes of axis A from traversal coordinate t0 (the plane, or
        // its negative when NEG), until none is left inside the volume. FB2: the view's fastest
        // axis is b2 (else b1). T: the view's element type, in which the sum runs.
        template <int D, int A, bool NEG, bool FB2, class T>
        __device__ inline T sid_planes(const SidRec& r, const Dims& dims, const PadView& V, int t0)
        {
            constexpr uint32_t lg = sizeof(T) == 8 ? 3u : 2u; // log2 of the element size
            constexpr int b1 = TileAxes<D, A>::b1, b2 = TileAxes<D, A>::b2;
            constexpr int f = (D == 3 && FB2) ? b2 : b1; // fastest axis of the view
            constexpr int m = (D == 3) ? (FB2 ? b1 : b2) : 0; // rows (3D)
            SidWalk w;
            sid_walk_init<D>(r, w);
            T acc = T(0);
            constexpr int dir = NEG ? -1 : 1;
            for (int p = NEG ? -t0 : t0; p >= 0 && p < dims.vol[A]; p += dir) {
                if (!__any(w.alive))
                    break;
                const char* pb = (const char*) V.base + (uint32_t) p * V.PS4; // the plane, for the whole group
                sid_walk_plane<D>(r, dims.vol, A, p, w, [&](int c0, int c1, int c2, float len) {
                    const int c[3] = {c0, c1, c2};
                    const uint32_t off =
                        (D == 3 ? mul24((uint32_t) (c[m] + 2), V.RS4) : 0u) + ((uint32_t) (c[f] + 2) << lg);
                    T v;
                    __builtin_memcpy(&v, pb + off, sizeof(T));
                    acc += v * (T) len;
                });
            }
            return acc;
        }

        template <int D, int K, class T>
        __device__ __forceinline__ void sid_group(int keys, int key, const SidRec& r, const Dims& dims,
                                                  const ViewSet& vs, int t, T& val)
        {
            if (!((keys >> K) & 1))
                return;
            constexpr int A = K / 2;
            constexpr bool NEG = (K & 1) != 0;
            const bool mine = key == K;
            const int t0 = wave_min(mine ? t : 0x7fffffff);
            if (mine)
                val = (D == 3 && ((vs.fb2 >> A) & 1)) ? sid_planes<D, A, NEG, true, T>(r, dims, vs.v[A], t0)
                                                      : sid_planes<D, A, NEG, false, T>(r, dims, vs.v[A], t0);
        }

        // Block id of the forward's tiles (fwd_pixel's order): the rays of its wavefronts, grouped
        // by driving axis and direction, each group walking its planes together.
        template <int D, class T>
        __device__ __forceinline__ void sid_fwd_block(const SidRec* __restrict__ table, const Dims& dims,
                                                      const ViewSet& views, const int tiles_u, const int64_t id,
                                                      T* __restrict__ sino)
        {
            int u, v, pose;
            fwd_pixel<D>(dims, tiles_u, id, u, v, pose);
            const bool inside = u < dims.det[0] && v < dims.det[1];
            const int64_t ray = (int64_t) pose * dims.pix_per_pose + (int64_t) v * dims.det[0] + u;
            SidRec r{};
            int key = -1, t = 0;
            if (inside) {
                r = table[ray];
                if (sid_hit(r)) {
                    const int a = sid_axis(r);
                    const bool neg = sid_step(r, a) < 0;
                    key = 2 * a + (neg ? 1 : 0);
                    t = neg ? -(int) r.cur0[a] : (int) r.cur0[a];
                }
            }
            const int keys = wave_or(key >= 0 ? 1 << key : 0);
            T val = T(0);
            sid_group<D, 0>(keys, key, r, dims, views, t, val);
            sid_group<D, 1>(keys, key, r, dims, views, t, val);
            sid_group<D, 2>(keys, key, r, dims, views, t, val);
            sid_group<D, 3>(keys, key, r, dims, views, t, val);
            if (D == 3) {
                sid_group<D, 2 * D - 2>(keys, key, r, dims, views, t, val);
                sid_group<D, 2 * D - 1>(keys, key, r, dims, views, t, val);
            }
            if (inside)
                sino[ray] = val;
        }

        template <int D, class T>
        __global__ void __launch_bounds__(kFwdThreads)
            siddon_fwd_planes_kernel(const SidRec* __restrict__ table, const Dims dims, const ViewSet views,
                                     const int tiles_u, const int per_xcd, const int64_t total,
                                     T* __restrict__ sino)
        {
            const int64_t id = (int64_t) (blockIdx.x % 8u) * per_xcd + (blockIdx.x / 8u);
            if (id >= total)
                return;
            sid_fwd_block<D, T>(table, dims, views, tiles_u, id, sino);
        }

        // the forwards of many operators in one launch (block b: block b - job.block0 of job
        // block_job[b], as that operator's own launch would run it)
        template <int D>
        __global__ void __launch_bounds__(kFwdThreads)
            siddon_fwd_planes_batch_kernel(const SidFwdJob* __restrict__ jobs, const int* __restrict__ block_job,
                                           const int per_xcd, const int64_t total, float* __restrict__ sinos)
        {
            const int64_t id = (int64_t) (blockIdx.x % 8u) * per_xcd + (blockIdx.x / 8u);
            if (id >= total)
                return;
            const SidFwdJob& J = jobs[block_job[id]];
            sid_fwd_block<D, float>(J.table, J.dims, J.views, J.tiles_u, id - J.block0, sinos + J.sino_off);
        }
    } // namespace

    template <class T>
    hipError_t siddon_forward_planes(const SidRec* table, const Dims& dims, const PadLayout* layouts,
                                     void* const* views, int axes_present, const T* vol, T* sino, hipStream_t s)
    {
        constexpr uint32_t es = sizeof(T);
        ViewSet vs{};
        for (int a = 0; a < dims.dim; ++a) {
            if (!(axes_present & (1 << a)))
                continue;
            const PadLayout& L = layouts[a];
            if (L.pairs)
                return hipErrorInvalidValue; // the Siddon walk reads single voxels of plain views
            hipError_t e = joseph_view_copy<T>(dims, L, vol, (T*) views[a], s);
            if (e != hipSuccess)
                return e;
            vs.v[a] = PadView{views[a], L.PS * es, L.RS * es, L.size};
            const int b2 = dims.dim == 3 ? (a == 2 ? 1 : 2) : -1;
            if (L.f == b2)
                vs.fb2 |= 1 << a;
        }
        const FwdGrid fg = fwd_grid(dims);
        if (fg.per_xcd * 8 > (int64_t) 0x7fffffff)
            return hipErrorInvalidValue;
        const dim3 grid((unsigned) (fg.per_xcd * 8));
        if (dims.dim == 2)
            hipLaunchKernelGGL((siddon_fwd_planes_kernel<2, T>), grid, dim3(kFwdThreads), 0, s, table, dims, vs,
                               fg.tiles_u, (int) fg.per_xcd, fg.total, sino);
        else
            hipLaunchKernelGGL((siddon_fwd_planes_kernel<3, T>), grid, dim3(kFwdThreads), 0, s, table, dims, vs,
                               fg.tiles_u, (int) fg.per_xcd, fg.total, sino);
        return hipGetLastError();
    }

    template hipError_t siddon_forward_planes<float>(const SidRec*, const Dims&, const PadLayout*, void* const*, int,
                                                     const float*, float*, hipStream_t);
    template hipError_t siddon_forward_planes<double>(const SidRec*, const Dims&, const PadLayout*, void* const*, int,
                                                      const double*, double*, hipStream_t);

    // Jobs of a batched Siddon forward: the view copies of every operator (one job per view)
    // and the forward blocks (one job per operator).
    void siddon_forward_batch_jobs(const JosephBatchItem* items, int n, std::vector<CopyJob>& cjobs,
                                   std::vector<int>& cblocks, std::vector<SidFwdJob>& fjobs, std::vector<int>& fblocks)
    {
        cjobs.clear();
        cblocks.clear();
        fjobs.clear();
        fblocks.clear();
        for (int j = 0; j < n; ++j) {
            const JosephBatchItem& it = items[j];
            SidFwdJob fj{};
            fj.table = it.sidtable;
            fj.dims = it.dims;
            fj.sino_off = it.sino_off;
            for (int a = 0; a < it.dims.dim; ++a) {
                if (!(it.axes_present & (1 << a)))
                    continue;
                const PadLayout& L = it.layouts[a];
                CopyJob cj{};
                cj.view = it.views[a];
                cj.dims = it.dims;
                cj.L = L;
                cj.vol_off = it.vol_off;
                cj.block0 = (int64_t) cblocks.size();
                const int64_t nb = view_copy_blocks(it.dims, L, cj.gx, cj.gy);
                cblocks.insert(cblocks.end(), (size_t) nb, (int) cjobs.size());
                cjobs.push_back(cj);
                fj.views.v[a] = PadView{it.views[a], L.PS * 4u, L.RS * 4u, L.size};
                const int b2 = it.dims.dim == 3 ? (a == 2 ? 1 : 2) : -1;
                if (L.f == b2)
                    fj.views.fb2 |= 1 << a;
            }
            const FwdGrid fg = fwd_grid(it.dims);
            fj.tiles_u = fg.tiles_u;
            fj.block0 = (int64_t) fblocks.size();
            fblocks.insert(fblocks.end(), (size_t) fg.total, (int) fjobs.size());
            fjobs.push_back(fj);
        }
    }

    hipError_t siddon_forward_batch(int dim, const CopyJob* cjobs, const int* cblocks, int64_t ncblocks,
                                    const SidFwdJob* fjobs, const int* fblocks, int64_t nfblocks, const float* vols,
                                    float* sinos, hipStream_t s)
    {
        hipError_t e = view_copy_batch(cjobs, cblocks, ncblocks, vols, s);
        if (e != hipSuccess)
            return e;
        const int64_t per_xcd = (nfblocks + 7) / 8;
        if (per_xcd * 8 > (int64_t) 0x7fffffff)
            return hipErrorInvalidValue;
        if (nfblocks > 0) {
            const dim3 grid((unsigned) (per_xcd * 8));
            if (dim == 2)
                hipLaunchKernelGGL((siddon_fwd_planes_batch_kernel<2>), grid, dim3(kFwdThreads), 0, s, fjobs, fblocks,
                                   (int) per_xcd, nfblocks, sinos);
            else
                hipLaunchKernelGGL((siddon_fwd_planes_batch_kernel<3>), grid, dim3(kFwdThreads), 0, s, fjobs, fblocks,
                                   (int) per_xcd, nfblocks, sinos);
        }
        return hipGetLastError();
    }
} // namespace ehp
