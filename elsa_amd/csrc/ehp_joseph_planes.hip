// ehp_joseph_planes.hip -- Joseph forward projection over plane-major volume views.
//
// The forward of elsa's Joseph method (elsa/projectors/JosephsMethod.cpp:140-198, forward
// branch :187-194) with one lane per ray. FP32 and bfloat16: each sample's bilinear
// interpolation in lerp form (plane_taps_add) and the sum in FP64, elsa's samples, floors and
// weights, elsa's sums to FP32 rounding. FP64: elsa's weights and sums over FP64 views
// (plane_taps_add64). A wavefront holds a patch of neighbouring detector
// pixels (8 x 8, or 16 x 4 over row-pair views), and its rays step through the volume's
// planes together, so at each step its lanes
// read the same plane at neighbouring in-plane positions. Those reads coalesce only if the axis
// along which the rays fan out along u is the fastest one in memory: rays driving along x fan
// out along z (circular trajectories about y), which in the input's x-fastest layout puts
// every lane on a different cache line.
// So each driving axis reads its own view of the volume, made at the start of the call (one
// read and one write of the volume): that axis slowest, the fan-out axis fastest, and two
// zero voxels around each plane so that taps outside the volume need no bounds test. 3D
// operators keep row-pair views (each element holds a voxel and the one in the next row), so
// that a sample's four taps are one load (16 bytes in FP32, 8 in bfloat16).
//
// Blocks of 256 lanes cover 32 x 8 pixels (64 x 4 with 16 x 4 wavefronts; 2D: 256 x 1) of one
// pose. Tiles are ordered detector-row band first, then pose, then column, and each XCD
// (consecutive workgroups go to different XCDs, so the block index is remapped) takes a
// contiguous run: the poses of a band then run together on one XCD. For a rotation about y a band of rows sees the same
// slab of the volume from every pose, so that slab stays in the XCD's L2.
#include "ehp_waves.hpp"

namespace ehp
{
    namespace
    {
        constexpr int kThreads = kFwdThreads;
        constexpr int kTile = 32; // transpose tile

        // View with x fastest: one lane per voxel, rows copied as they are.
        template <class E>
        __device__ __forceinline__ void view_copy_rows(const E* __restrict__ in, E* __restrict__ out,
                                                       const Dims& dims, const PadLayout& L, int64_t block)
        {
            const int64_t i = block * 256 + threadIdx.x;
            if (i >= dims.nvox)
                return;
            const int x = (int) (i % dims.vol[0]);
            const int64_t r = i / dims.vol[0];
            const int y = (int) (r % dims.vol[1]);
            const int z = (int) (r / dims.vol[1]);
            pad_store<E>(L, out, x, y, z, in[i]);
        }

        // View with f != x fastest: a 32 x 32 LDS tile over the (x, f) plane, reads coalesced
        // along x, writes along f. Block (bx, by, bz) of the grid (ceil(nx / 32), ceil(n_f / 32),
        // n_t), t the third axis.
        template <class E>
        __device__ __forceinline__ void view_copy_transpose(const E* __restrict__ in, E* __restrict__ out,
                                                            const Dims& dims, const PadLayout& L, int bx, int by,
                                                            int bz)
        {
            __shared__ E tile[kTile][kTile + 1];
            const int f = L.f;
            const int t = 3 - f; // the axis that is neither x nor f
            const int x0 = bx * kTile, f0 = by * kTile, ct = bz;
            const int tx = threadIdx.x % kTile, ty = threadIdx.x / kTile; // 32 x 8
            const int nx = dims.vol[0], nf = dims.vol[f];
            const int64_t in_f = f == 1 ? dims.strides[1] : dims.strides[2];
            const int64_t in_t = t == 1 ? dims.strides[1] : dims.strides[2];
#pragma unroll
            for (int j = 0; j < kTile; j += 8) {
                const int cx = x0 + tx, cf = f0 + ty + j;
                if (cx < nx && cf < nf)
                    tile[ty + j][tx] = in[cx + cf * in_f + (int64_t) ct * in_t];
            }
            __syncthreads();
#pragma unroll
            for (int j = 0; j < kTile; j += 8) {
                const int cx = x0 + ty + j, cf = f0 + tx;
                if (cx < nx && cf < nf) {
                    const int c1 = f == 1 ? cf : ct, c2 = f == 1 ? ct : cf;
                    pad_store<E>(L, out, cx, c1, c2, tile[tx][ty + j]);
                }
            }
        }

        template <int D, int A, bool NEG, bool V16, bool F64>
        __device__ inline double fwd_span(const RayRec& r, const Dims& dims, const ViewSet& vs, int t0, int niter,
                                          int s_first, int full0, int full1)
        {
            if constexpr (F64) {
                // FP64 views hold one element per voxel
                if (D == 3 && (vs.fb2 >> A) & 1)
                    return plane_span_forward<D, A, NEG, true, false, false, true>(r, dims, vs.v[A], t0, niter,
                                                                                   s_first, full0, full1);
                return plane_span_forward<D, A, NEG, false, false, false, true>(r, dims, vs.v[A], t0, niter, s_first,
                                                                                full0, full1);
            }
            if (D == 3 && (vs.pairs >> A) & 1) {
                if ((vs.fb2 >> A) & 1)
                    return plane_span_forward<D, A, NEG, true, V16, true>(r, dims, vs.v[A], t0, niter, s_first,
                                                                          full0, full1);
                return plane_span_forward<D, A, NEG, false, V16, true>(r, dims, vs.v[A], t0, niter, s_first, full0,
                                                                       full1);
            }
            if (D == 3 && (vs.fb2 >> A) & 1)
                return plane_span_forward<D, A, NEG, true, V16>(r, dims, vs.v[A], t0, niter, s_first, full0, full1);
            return plane_span_forward<D, A, NEG, false, V16>(r, dims, vs.v[A], t0, niter, s_first, full0, full1);
        }

        // The rays of group K = 2 * axis + direction of a wavefront share one plane loop over the
        // group's span. Called with every lane active; `keys` holds the groups present.
        template <int D, int K, bool V16, bool F64>
        __device__ __forceinline__ void fwd_group(int keys, int key, const RayRec& r, const Dims& dims,
                                                  const ViewSet& vs, int tfirst, int tlast, int s0, double& val)
        {
            if (!((keys >> K) & 1))
                return;
            const bool mine = key == K;
            const int t0 = wave_min(mine ? tfirst : 0x7fffffff);
            const int t1 = wave_max(mine ? tlast : -0x7fffffff);
            const int ta = wave_max(mine ? tfirst : -0x7fffffff); // every ray of the group has a sample
            const int tb = wave_min(mine ? tlast : 0x7fffffff);   // in [ta, tb]
            const int full0 = ta - t0, full1 = tb - t0 + 1 > full0 ? tb - t0 + 1 : full0;
            if (mine)
                val = fwd_span<D, K / 2, (K & 1) != 0, V16, F64>(r, dims, vs, t0, t1 - t0 + 1, s0 + t0 - tfirst,
                                                                 full0, full1);
        }

        template <class E>
        __global__ void __launch_bounds__(256)
            view_copy_rows_kernel(const E* __restrict__ in, E* __restrict__ out, const Dims dims, const PadLayout L)
        {
            view_copy_rows<E>(in, out, dims, L, blockIdx.x);
        }

        template <class E>
        __global__ void __launch_bounds__(256)
            view_copy_transpose_kernel(const E* __restrict__ in, E* __restrict__ out, const Dims dims,
                                       const PadLayout L)
        {
            view_copy_transpose<E>(in, out, dims, L, blockIdx.x, blockIdx.y, blockIdx.z);
        }

        // Block id of one pose-band tile of the forward: id = (band * nposes + pose) * tiles_u + tu.
        // Each wavefront holds WU x (64 / WU) neighbouring pixels (2D: 64 of the row). Its rays fall into
        // groups by driving axis and direction (in practice one or two), and each group runs one
        // plane loop (plane_span_forward), so that at every step all its lanes read one plane. A
        // per-lane loop would put rays that enter through a side face of the volume, each at its
        // own plane, on a cache line of their own at every step. Every lane of the block takes
        // part in the wavefront reductions; lanes past the detector edge join no group.
        template <int D, class OT = float, int WU = kFwdWU>
        __device__ __forceinline__ void fwd_planes_block(const PoseGeom* __restrict__ geo, const Dims& dims,
                                                         const ViewSet& views, const int tiles_u, const int64_t id,
                                                         OT* __restrict__ sino)
        {
            int u, v, pose;
            fwd_pixel<D, WU>(dims, tiles_u, id, u, v, pose);
            const bool inside = u < dims.det[0] && v < dims.det[1];
            RayRec r{};
            int key = -1, tfirst = 0, tlast = 0, s0 = 0;
            if (inside) {
                float o[3], d[3];
                make_ray_px<D>(geo[pose], dims, u, v, o, d);
                JRay jr;
                joseph_setup<D>(o, d, dims.vol, jr);
                r = make_ray_rec<D>(jr);
                if (plane_ray_span(r, dims.vol[jr.axis], tfirst, tlast, s0))
                    key = 2 * jr.axis + ((r.meta & 4) ? 1 : 0);
            }
            constexpr bool V16 = sizeof(OT) == 2; // bfloat16 operators read bfloat16 views
            constexpr bool F64 = sizeof(OT) == 8; // FP64 operators read FP64 views
            const int keys = wave_or(key >= 0 ? 1 << key : 0);
            double val = 0.0;
            fwd_group<D, 0, V16, F64>(keys, key, r, dims, views, tfirst, tlast, s0, val);
            fwd_group<D, 1, V16, F64>(keys, key, r, dims, views, tfirst, tlast, s0, val);
            fwd_group<D, 2, V16, F64>(keys, key, r, dims, views, tfirst, tlast, s0, val);
            fwd_group<D, 3, V16, F64>(keys, key, r, dims, views, tfirst, tlast, s0, val);
            if (D == 3) {
                fwd_group<D, 2 * D - 2, V16, F64>(keys, key, r, dims, views, tfirst, tlast, s0, val);
                fwd_group<D, 2 * D - 1, V16, F64>(keys, key, r, dims, views, tfirst, tlast, s0, val);
            }
            if (inside) {
                OT* out = sino + (int64_t) pose * dims.pix_per_pose + (int64_t) v * dims.det[0] + u;
                if constexpr (F64)
                    *out = val;
                else
                    *out = down<OT>((float) val);
            }
        }

        template <int D, class OT, int WU = kFwdWU>
        __global__ void __launch_bounds__(kThreads)
            joseph_fwd_planes_kernel(const PoseGeom* __restrict__ geo, const Dims dims, const ViewSet views,
                                     const int tiles_u, const int per_xcd, const int64_t total,
                                     OT* __restrict__ sino)
        {
            const int64_t id = (int64_t) (blockIdx.x % 8u) * per_xcd + (blockIdx.x / 8u);
            if (id >= total)
                return;
            fwd_planes_block<D, OT, WU>(geo, dims, views, tiles_u, id, sino);
        }

        // ---- batches: one launch over the jobs of many operators; block b runs block
        // b - job.block0 of job block_job[b], exactly as the single-operator launch would

        __global__ void __launch_bounds__(256)
            view_copy_batch_kernel(const CopyJob* __restrict__ jobs, const int* __restrict__ block_job,
                                   const float* __restrict__ vols)
        {
            const CopyJob& J = jobs[block_job[blockIdx.x]];
            const int64_t b = (int64_t) blockIdx.x - J.block0;
            const float* in = vols + J.vol_off;
            if (J.L.f == 0) {
                view_copy_rows<float>(in, J.view, J.dims, J.L, b);
            } else {
                const int64_t gxy = (int64_t) J.gx * J.gy;
                const int bz = (int) (b / gxy);
                const int r = (int) (b - bz * gxy);
                view_copy_transpose<float>(in, J.view, J.dims, J.L, r % J.gx, r / J.gx, bz);
            }
        }

        template <int D>
        __global__ void __launch_bounds__(kThreads)
            joseph_fwd_planes_batch_kernel(const FwdJob* __restrict__ jobs, const int* __restrict__ block_job,
                                           const int per_xcd, const int64_t total, float* __restrict__ sinos)
        {
            const int64_t id = (int64_t) (blockIdx.x % 8u) * per_xcd + (blockIdx.x / 8u);
            if (id >= total)
                return;
            const FwdJob& J = jobs[block_job[id]];
            fwd_planes_block<D, float>(J.geo, J.dims, J.views, J.tiles_u, id - J.block0, sinos + J.sino_off);
        }
    } // namespace

    template <class E>
    hipError_t joseph_view_copy(const Dims& dims, const PadLayout& L, const E* vol, E* view, hipStream_t s)
    {
        if (L.f < 0 || L.f > 2 || L.f == L.A)
            return hipErrorInvalidValue;
        if (L.f == 0) {
            const int64_t blocks = (dims.nvox + 255) / 256;
            hipLaunchKernelGGL((view_copy_rows_kernel<E>), dim3((unsigned) blocks), dim3(256), 0, s, vol, view, dims,
                               L);
        } else {
            const int t = 3 - L.f;
            const dim3 grid((unsigned) ((dims.vol[0] + kTile - 1) / kTile),
                            (unsigned) ((dims.vol[L.f] + kTile - 1) / kTile), (unsigned) dims.vol[t]);
            hipLaunchKernelGGL((view_copy_transpose_kernel<E>), grid, dim3(256), 0, s, vol, view, dims, L);
        }
        return hipGetLastError();
    }

    // T = float (FP32 operators), bf16 (bfloat16 storage: bfloat16 volume, views and sinogram,
    // FP32 arithmetic) or double (FP64 operators: FP64 views, elsa's FP64 sums)
    template <class T>
    hipError_t joseph_forward_planes(const PoseGeom* geo, const Dims& dims, const PadLayout* layouts,
                                     void* const* views, int axes_present, const T* vol, T* sino, hipStream_t s)
    {
        constexpr uint32_t es = sizeof(T);
        ViewSet vs{};
        for (int a = 0; a < dims.dim; ++a) {
            if (!(axes_present & (1 << a)))
                continue;
            const PadLayout& L = layouts[a];
            hipError_t e = joseph_view_copy<T>(dims, L, vol, (T*) views[a], s);
            if (e != hipSuccess)
                return e;
            vs.v[a] = PadView{views[a], L.PS * es, L.RS * es, L.size};
            const int b2 = dims.dim == 3 ? (a == 2 ? 1 : 2) : -1;
            if (L.f == b2)
                vs.fb2 |= 1 << a;
            if (L.pairs)
                vs.pairs |= 1 << a;
        }
        // row-pair views (3D) run with wider wavefronts
        const bool pairs = dims.dim == 3 && vs.pairs != 0;
        const FwdGrid fg = fwd_grid(dims, pairs ? kFwdWUPairs : kFwdWU);
        const int tiles_u = fg.tiles_u;
        const int64_t total = fg.total, per_xcd = fg.per_xcd;
        if (per_xcd * 8 > (int64_t) 0x7fffffff)
            return hipErrorInvalidValue;
        const dim3 grid((unsigned) (per_xcd * 8));
        if (dims.dim == 2)
            hipLaunchKernelGGL((joseph_fwd_planes_kernel<2, T>), grid, dim3(kThreads), 0, s, geo, dims, vs, tiles_u,
                               (int) per_xcd, total, sino);
        else if (pairs)
            hipLaunchKernelGGL((joseph_fwd_planes_kernel<3, T, kFwdWUPairs>), grid, dim3(kThreads), 0, s, geo, dims,
                               vs, tiles_u, (int) per_xcd, total, sino);
        else
            hipLaunchKernelGGL((joseph_fwd_planes_kernel<3, T>), grid, dim3(kThreads), 0, s, geo, dims, vs, tiles_u,
                               (int) per_xcd, total, sino);
        return hipGetLastError();
    }

    template hipError_t joseph_view_copy<float>(const Dims&, const PadLayout&, const float*, float*, hipStream_t);
    template hipError_t joseph_view_copy<bf16>(const Dims&, const PadLayout&, const bf16*, bf16*, hipStream_t);
    template hipError_t joseph_view_copy<double>(const Dims&, const PadLayout&, const double*, double*, hipStream_t);
    template hipError_t joseph_forward_planes<float>(const PoseGeom*, const Dims&, const PadLayout*, void* const*, int,
                                                     const float*, float*, hipStream_t);
    template hipError_t joseph_forward_planes<bf16>(const PoseGeom*, const Dims&, const PadLayout*, void* const*, int,
                                                    const bf16*, bf16*, hipStream_t);
    template hipError_t joseph_forward_planes<double>(const PoseGeom*, const Dims&, const PadLayout*, void* const*,
                                                      int, const double*, double*, hipStream_t);

    // Jobs of a batched forward: the view copies of every operator (one job per view) and the
    // forward blocks (one job per operator); block_job maps each block to its job.
    void joseph_forward_batch_jobs(const JosephBatchItem* items, int n, std::vector<CopyJob>& cjobs,
                                   std::vector<int>& cblocks, std::vector<FwdJob>& fjobs, std::vector<int>& fblocks)
    {
        cjobs.clear();
        cblocks.clear();
        fjobs.clear();
        fblocks.clear();
        for (int j = 0; j < n; ++j) {
            const JosephBatchItem& it = items[j];
            FwdJob fj{};
            fj.geo = it.geo;
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
                if (L.pairs)
                    fj.views.pairs |= 1 << a;
            }
            const FwdGrid fg = fwd_grid(it.dims);
            fj.tiles_u = fg.tiles_u;
            const int64_t nb = fg.total;
            fj.block0 = (int64_t) fblocks.size();
            fblocks.insert(fblocks.end(), (size_t) nb, (int) fjobs.size());
            fjobs.push_back(fj);
        }
    }

    // Blocks of the view copy of one volume into its view (rows: a flat range; transpose:
    // the grid (gx, gy, n_t) flattened with x fastest).
    int64_t view_copy_blocks(const Dims& dims, const PadLayout& L, int& gx, int& gy)
    {
        if (L.f == 0) {
            gx = gy = 0;
            return (dims.nvox + 255) / 256;
        }
        gx = (dims.vol[0] + kTile - 1) / kTile;
        gy = (dims.vol[L.f] + kTile - 1) / kTile;
        return (int64_t) gx * gy * dims.vol[3 - L.f];
    }

    hipError_t view_copy_batch(const CopyJob* cjobs, const int* cblocks, int64_t ncblocks, const float* vols,
                               hipStream_t s)
    {
        if (ncblocks > 0x7fffffff)
            return hipErrorInvalidValue;
        if (ncblocks > 0)
            hipLaunchKernelGGL(view_copy_batch_kernel, dim3((unsigned) ncblocks), dim3(256), 0, s, cjobs, cblocks, vols);
        return hipGetLastError();
    }

    hipError_t joseph_forward_batch(int dim, const CopyJob* cjobs, const int* cblocks, int64_t ncblocks,
                                    const FwdJob* fjobs, const int* fblocks, int64_t nfblocks, const float* vols,
                                    float* sinos, hipStream_t s)
    {
        if (nfblocks > 0x7fffffff)
            return hipErrorInvalidValue;
        hipError_t e = view_copy_batch(cjobs, cblocks, ncblocks, vols, s);
        if (e != hipSuccess)
            return e;
        const int64_t per_xcd = (nfblocks + 7) / 8;
        if (per_xcd * 8 > (int64_t) 0x7fffffff)
            return hipErrorInvalidValue;
        if (nfblocks > 0) {
            const dim3 grid((unsigned) (per_xcd * 8));
            if (dim == 2)
                hipLaunchKernelGGL((joseph_fwd_planes_batch_kernel<2>), grid, dim3(kThreads), 0, s, fjobs, fblocks,
                                   (int) per_xcd, nfblocks, sinos);
            else
                hipLaunchKernelGGL((joseph_fwd_planes_batch_kernel<3>), grid, dim3(kThreads), 0, s, fjobs, fblocks,
                                   (int) per_xcd, nfblocks, sinos);
        }
        return hipGetLastError();
    }
} // namespace ehp
