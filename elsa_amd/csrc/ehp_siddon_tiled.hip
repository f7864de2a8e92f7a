// ehp_siddon_tiled.hip -- Siddon back projection with LDS-privatised volume tiles (FP32, FP64).
//
// The adjoint of Siddon's method (elsa/projectors/SiddonsMethod.cpp:76-124, adjoint
// :119-120) without global atomics, on the scheme of the tiled Joseph back projection: the
// volume is cut into tiles long along one driving axis, one workgroup owns a tile for all
// poses and sums in LDS every crossing of every ray (driving along that axis) that passes
// through it, then adds the tile to the volume once; one launch per driving axis. A ray
// enters a tile through siddon_box, which restarts its traversal exactly at the tile's
// boundary, so every crossing -- voxel and length -- is the one the forward visits.
//
// Sums are 64-bit fixed point as in the Joseph kernel: one exponent per call from max|y|
// and the slot bound, each contribution y * length rounded once to the unit (fixed_round; in
// FP32 the product is exact in FP64).
#include "ehp_kernels.hpp"

namespace ehp
{
    namespace
    {
        constexpr int kThreads = 256;
    } // namespace

    template <int D>
    __global__ void __launch_bounds__(256)
        siddon_table_kernel(const PoseGeom* __restrict__ geo, const Dims dims, SidRec* __restrict__ table)
    {
        const int pose = blockIdx.y;
        const int64_t pix = (int64_t) blockIdx.x * blockDim.x + threadIdx.x;
        if (pix >= dims.pix_per_pose)
            return;
        const int u = (int) (pix % dims.det[0]);
        const int v = (int) (pix / dims.det[0]);
        float o[3], d[3];
        make_ray_px<D>(geo[pose], dims, u, v, o, d);
        table[(int64_t) pose * dims.pix_per_pose + pix] = siddon_setup<D>(o, d, dims.vol);
    }

    __global__ void __launch_bounds__(256)
        siddon_pose_axes_kernel(const SidRec* __restrict__ table, const Dims dims, int* __restrict__ pose_axes)
    {
        const int pose = blockIdx.y;
        const int64_t pix = (int64_t) blockIdx.x * blockDim.x + threadIdx.x;
        if (pix >= dims.pix_per_pose)
            return;
        const SidRec r = table[(int64_t) pose * dims.pix_per_pose + pix];
        if (sid_hit(r))
            atomicOr(pose_axes + pose, 1 << sid_axis(r));
    }

    hipError_t siddon_ray_table(const PoseGeom* geo, const Dims& dims, SidRec* table, int* pose_axes, hipStream_t s)
    {
        const dim3 grid((unsigned) ((dims.pix_per_pose + 255) / 256), (unsigned) dims.nposes);
        if (dims.dim == 2)
            hipLaunchKernelGGL((siddon_table_kernel<2>), grid, dim3(256), 0, s, geo, dims, table);
        else
            hipLaunchKernelGGL((siddon_table_kernel<3>), grid, dim3(256), 0, s, geo, dims, table);
        hipError_t e = launch_fill<int>(pose_axes, dims.nposes, 0, s);
        if (e != hipSuccess)
            return e;
        hipLaunchKernelGGL(siddon_pose_axes_kernel, grid, dim3(256), 0, s, table, dims, pose_axes);
        return hipGetLastError();
    }

    size_t siddon_tile_lds_bytes(const Dims& dims, const TileCfg& tc, int batch)
    {
        return (size_t) tc.LA * tc.TB * (dims.dim == 3 ? tc.TC : 1) * sizeof(long long) + batch_lds_bytes(batch);
    }

    // One tile of the back projection along driving axis A (tile < ntA * ntB * ntC); T = float or
    // double (sinogram and volume).
    template <int D, int A, class T>
    __device__ __forceinline__ void sid_adj_tiled_block(const PoseGeom* __restrict__ geo,
                                                        const SidRec* __restrict__ table, const Dims& dims,
                                                        const int* __restrict__ pose_axes, const TileCfg& tc,
                                                        const int ntA, const int ntB, const int tile, const int batch,
                                                        const int* __restrict__ fixexp, const T* __restrict__ sino,
                                                        T* __restrict__ vol)
    {
        constexpr int b1 = TileAxes<D, A>::b1, b2 = TileAxes<D, A>::b2;
        extern __shared__ long long lds[];

        const int tid = threadIdx.x;
        const TileBox t = make_tile<D, A>(dims, tc, tile % ntA, (tile / ntA) % ntB, tile / (ntA * ntB));
        const int nslots = tc.LA * tc.TB * (D == 3 ? tc.TC : 1);
        long long* iacc = lds;
        const BatchLds bl = batch_lds(lds, nslots, batch);
        CandRect* s_rect = bl.rect;
        int *s_cnt = bl.cnt, *s_pref = bl.pref;
        for (int i = tid; i < nslots; i += kThreads)
            iacc[i] = 0;
        // the tile's voxel box
        int lo[3] = {0, 0, 0}, hi[3] = {0, 0, 0};
        lo[A] = t.A0;
        hi[A] = t.A1 - 1;
        lo[b1] = t.B0;
        hi[b1] = t.B1 - 1;
        if (D == 3) {
            lo[b2] = t.C0;
            hi[b2] = t.C1 - 1;
        }
        float flo[3] = {0.0f, 0.0f, 0.0f}, fhi[3] = {0.0f, 0.0f, 0.0f};
#pragma unroll
        for (int i = 0; i < D; ++i) {
            flo[i] = (float) lo[i];
            fhi[i] = (float) (hi[i] + 1);
        }
        const int eg = fixexp[0];
        __syncthreads();

        for (int pbase = 0; eg != kNonFinite && pbase < dims.nposes; pbase += batch) {
            if (tid < batch) {
                const int p = pbase + tid;
                int u0 = 0, u1 = -1, v0 = 0, v1 = -1;
                int cnt = 0;
                if (p < dims.nposes && (pose_axes[p] & (1 << A))) {
                    if (box_bbox<D>(geo[p], flo, fhi, dims, u0, u1, v0, v1))
                        cnt = (u1 - u0 + 1) * (v1 - v0 + 1);
                }
                CandRect cr;
                cr.nu = u1 - u0 + 1;
                cr.rnu = 1.0f / (float) (u1 - u0 + 1);
                cr.base = (long long) p * dims.pix_per_pose + (long long) v0 * dims.det[0] + u0;
                s_rect[tid] = cr;
                s_cnt[tid] = cnt;
            }
            __syncthreads();
            if (tid == 0) {
                s_pref[0] = 0;
                for (int j = 0; j < batch; ++j)
                    s_pref[j + 1] = s_pref[j] + s_cnt[j];
            }
            __syncthreads();
            const int total = s_pref[batch];
            // ray index of candidate k of this batch: pixel (col, row) of its pose's rectangle
            auto cand_ray = [&](int k) -> int64_t {
                const int j = batch_pose_slot(s_pref, batch, k);
                const CandRect cr = s_rect[j];
                int row, col;
                rect_split(k - s_pref[j], cr.nu, cr.rnu, row, col);
                return cr.base + (int64_t) (mul24((uint32_t) row, (uint32_t) dims.det[0]) + (uint32_t) col);
            };
            T yv_next = T(0);
            SidRec r_next{};
            if (tid < total) {
                const int64_t ray = cand_ray(tid);
                yv_next = sino[ray];
                r_next = table[ray];
            }
            for (int k = tid; k < total; k += kThreads) {
                const T yv = yv_next;
                const SidRec r = r_next;
                if (k + kThreads < total) {
                    const int64_t ray = cand_ray(k + kThreads);
                    yv_next = sino[ray];
                    r_next = table[ray];
                }
                if (yv == T(0) || sid_axis(r) != A)
                    continue;
                // y in units of q (a power-of-two scaling, exact); each crossing adds
                // y * length rounded once to an integer (|y * length| < 2^51 units: the slot
                // bound is at least 4096 and lengths are below 2)
                const double yq = ldexp((double) yv, 62 - eg);
                siddon_box<D>(r, lo, hi, [&](int c0, int c1, int c2, float w) {
                    if (w != 0.0f) {
                        const int slot = sid_slot<D, A>(t, tc, c0, c1, c2);
                        atomicAdd(reinterpret_cast<unsigned long long*>(iacc + slot), fixed_round(yq, (double) w));
                    }
                });
            }
            __syncthreads();
        }

        // add the tile to the volume
        const double q = eg == kNonFinite ? __longlong_as_double(0x7ff8000000000000ll) : ldexp(1.0, eg - 62);
        const int TBe = tc.TB, TCe = D == 3 ? tc.TC : 1;
        for (int i = tid; i < nslots; i += kThreads) {
            const int xb = i % TBe;
            const int rest = i / TBe;
            const int xc = rest % TCe;
            const int xa = rest / TCe;
            const int ca = t.A0 + xa, cb = t.B0 + xb, cc = t.C0 + xc;
            if (ca >= t.A1 || cb >= t.B1 || (D == 3 && cc >= t.C1))
                continue;
            int c[3] = {0, 0, 0};
            c[A] = ca;
            c[b1] = cb;
            if (D == 3)
                c[b2] = cc;
            const int64_t lin = c[0] + c[1] * dims.strides[1] + (int64_t) c[2] * dims.strides[2];
            vol[lin] += (T) ((double) iacc[i] * q);
        }
    }

    template <int D, int A, class T>
    __global__ void __launch_bounds__(kThreads)
        siddon_adj_tiled_kernel(const PoseGeom* __restrict__ geo, const SidRec* __restrict__ table, const Dims dims,
                                const int* __restrict__ pose_axes, const TileCfg tc, const int ntA, const int ntB,
                                const int ntC, const int tiles_per_xcd, const int batch,
                                const int* __restrict__ fixexp, const T* __restrict__ sino, T* __restrict__ vol)
    {
        // consecutive workgroups land on different XCDs: give each XCD a contiguous block of tiles
        const int tile = (int) (blockIdx.x % 8u) * tiles_per_xcd + (int) (blockIdx.x / 8u);
        if (tile >= ntA * ntB * ntC)
            return;
        sid_adj_tiled_block<D, A, T>(geo, table, dims, pose_axes, tc, ntA, ntB, tile, batch, fixexp, sino, vol);
    }

    // the tiles of axis A of many operators in one launch (block b: tile b - job.block0 of job
    // block_job[b])
    template <int D, int A>
    __global__ void __launch_bounds__(kThreads)
        siddon_adj_tiled_batch_kernel(const SidAdjJob* __restrict__ jobs, const int* __restrict__ block_job,
                                      const int per_xcd, const int64_t total, const int* __restrict__ fixexp,
                                      const float* __restrict__ sinos, float* __restrict__ vols)
    {
        const int64_t id = (int64_t) (blockIdx.x % 8u) * per_xcd + (blockIdx.x / 8u);
        if (id >= total)
            return;
        const SidAdjJob& J = jobs[block_job[id]];
        sid_adj_tiled_block<D, A, float>(J.geo, J.table, J.dims, J.pose_axes, J.tc, J.ntA, J.ntB,
                                         (int) (id - J.block0), J.batch, fixexp + J.item, sinos + J.sino_off,
                                         vols + J.vol_off);
    }

    template <class T>
    hipError_t siddon_adjoint_tiled(const PoseGeom* geo, const SidRec* table, const Dims& dims, const int* pose_axes,
                                    int axes_present, const TileCfg& tc, int batch, double slot_bound,
                                    unsigned int* d_maxbits, int* d_fixexp, const T* sino, T* vol, hipStream_t s)
    {
        if (batch < 1 || batch > kMaxPoseBatch)
            return hipErrorInvalidValue;
        hipError_t e = launch_fill<T>(vol, dims.nvox, T(0), s);
        if (e == hipSuccess)
            e = fixed_point_exponent_launch(sino, dims.nrays, slot_bound, d_maxbits, d_fixexp, s);
        if (e != hipSuccess)
            return e;
        const size_t lds = siddon_tile_lds_bytes(dims, tc, batch);
        for (int a = 0; a < dims.dim; ++a) {
            if (!(axes_present & (1 << a)))
                continue;
            int ntA, ntB, ntC;
            tile_counts(dims, tc, a, ntA, ntB, ntC);
            const int per_xcd = (ntA * ntB * ntC + 7) / 8;
            const dim3 grid((unsigned) (per_xcd * 8));
#define EHP_SID_TILED(Dv, Av)                                                                                      \
    hipLaunchKernelGGL((siddon_adj_tiled_kernel<Dv, Av, T>), grid, dim3(kThreads), lds, s, geo, table, dims,         \
                       pose_axes, tc, ntA, ntB, ntC, per_xcd, batch, d_fixexp, sino, vol)
            if (dims.dim == 2) {
                if (a == 0)
                    EHP_SID_TILED(2, 0);
                else
                    EHP_SID_TILED(2, 1);
            } else {
                if (a == 0)
                    EHP_SID_TILED(3, 0);
                else if (a == 1)
                    EHP_SID_TILED(3, 1);
                else
                    EHP_SID_TILED(3, 2);
            }
#undef EHP_SID_TILED
            e = hipGetLastError();
            if (e != hipSuccess)
                return e;
        }
        return hipSuccess;
    }

    template hipError_t siddon_adjoint_tiled<float>(const PoseGeom*, const SidRec*, const Dims&, const int*, int,
                                                    const TileCfg&, int, double, unsigned int*, int*, const float*,
                                                    float*, hipStream_t);
    template hipError_t siddon_adjoint_tiled<double>(const PoseGeom*, const SidRec*, const Dims&, const int*, int,
                                                     const TileCfg&, int, double, unsigned int*, int*,
                                                     const double*, double*, hipStream_t);

    // The tiles of every operator per driving axis (one job per operator that has rays of that
    // axis).
    void siddon_adjoint_batch_jobs(const JosephBatchItem* items, int n, std::vector<SidAdjJob> ajobs[3],
                                   std::vector<int> ablocks[3], size_t lds[3])
    {
        for (int a = 0; a < 3; ++a) {
            ajobs[a].clear();
            ablocks[a].clear();
            lds[a] = 0;
        }
        for (int j = 0; j < n; ++j) {
            const JosephBatchItem& it = items[j];
            for (int a = 0; a < it.dims.dim; ++a) {
                if (!(it.axes_present & (1 << a)))
                    continue;
                SidAdjJob aj{};
                aj.geo = it.geo;
                aj.table = it.sidtable;
                aj.pose_axes = it.pose_axes;
                aj.dims = it.dims;
                aj.tc = it.tile;
                aj.batch = it.batch;
                aj.item = j;
                aj.sino_off = it.sino_off;
                aj.vol_off = it.vol_off;
                int ntC;
                tile_counts(it.dims, it.tile, a, aj.ntA, aj.ntB, ntC);
                aj.block0 = (int64_t) ablocks[a].size();
                ablocks[a].insert(ablocks[a].end(), (size_t) aj.ntA * aj.ntB * ntC, (int) ajobs[a].size());
                ajobs[a].push_back(aj);
                const size_t need = siddon_tile_lds_bytes(it.dims, it.tile, it.batch);
                lds[a] = need > lds[a] ? need : lds[a];
            }
        }
    }

    // The tile launches of a batched back projection (the volumes already cleared and the
    // exponents d_fixexp[item] computed by fixed_point_exponent_batch).
    hipError_t siddon_adjoint_batch(int dim, const SidAdjJob* const ajobs[3], const int* const ablocks[3],
                                    const int64_t nablocks[3], const size_t lds[3], const int* d_fixexp,
                                    const float* sinos, float* vols, hipStream_t s)
    {
        for (int a = 0; a < dim; ++a) {
            const int64_t total = nablocks[a];
            if (total == 0)
                continue;
            const int64_t per_xcd = (total + 7) / 8;
            if (per_xcd * 8 > (int64_t) 0x7fffffff)
                return hipErrorInvalidValue;
            const dim3 grid((unsigned) (per_xcd * 8));
#define EHP_SID_BATCH(Dv, Av)                                                                                      \
    hipLaunchKernelGGL((siddon_adj_tiled_batch_kernel<Dv, Av>), grid, dim3(kThreads), lds[a], s, ajobs[a], ablocks[a], \
                       (int) per_xcd, total, d_fixexp, sinos, vols)
            if (dim == 2) {
                if (a == 0)
                    EHP_SID_BATCH(2, 0);
                else
                    EHP_SID_BATCH(2, 1);
            } else {
                if (a == 0)
                    EHP_SID_BATCH(3, 0);
                else if (a == 1)
                    EHP_SID_BATCH(3, 1);
                else
                    EHP_SID_BATCH(3, 2);
            }
#undef EHP_SID_BATCH
            const hipError_t e = hipGetLastError();
            if (e != hipSuccess)
                return e;
        }
        return hipSuccess;
    }
} // namespace ehp
