// ehp_joseph_tiled.hip -- Joseph back projection with LDS-privatised volume tiles (FP32).
//
// The matched adjoint of the Joseph forward (elsa/projectors/JosephsMethod.cpp:140-198,
// adjoint branch :180-186) without global atomics. The volume is cut into tiles that are
// long along one driving axis a (LA planes) and TB x TC voxels across. One workgroup owns a
// tile for all poses: it accumulates in LDS every sample of every ray whose driving axis is
// a and whose bilinear footprint touches the tile, then adds the tile to the volume once.
// One launch per driving axis that occurs in the trajectory ("pass"); each voxel is owned by
// exactly one tile per pass, so passes are plain read-modify-write.
//
// Accumulation: LDS float atomics are slow on this part (a ds_add_f32 wave operation costs
// about 15 times a ds_add_u64), so the tile is summed with 64-bit integer atomics in fixed
// point: units of q = 2^(eg - 62), eg chosen per call from max|y| and a bound on what one slot
// can receive, so that no sum can overflow. Each ray rounds its contributions to int32 in its
// own power-of-two scale (error at most 2^-30 of y * len) and shifts them into units of q: a
// conversion, a sign extension and a 64-bit shift per tap instead of a float-to-int64
// conversion. Integer sums do not depend on the order of the additions, so the result is
// bitwise reproducible, whatever the tile shape, and more accurate than a sequential FP32
// accumulation. A non-finite y makes the whole output NaN (the fixed-point sums cannot carry
// it to just the voxels it touches).
//
// The per-ray state comes from a table built once per operator from the same functions as
// the one-thread-per-ray kernels (make_ray, joseph_setup), and the per-sample weights use the
// same arithmetic as joseph_sample, so they are bit-identical to the forward's.
#include "ehp_kernels.hpp"

namespace ehp
{
    namespace
    {
        constexpr int kThreads = 256;
        // LDS (address space 3) pointers: 32-bit byte addresses
        typedef __attribute__((address_space(3))) unsigned long long lds_u64;
        typedef __attribute__((address_space(3))) char lds_char;
    } // namespace

    // ---------------------------------------------------------------------------------
    // per-ray table

    template <int D>
    __global__ void __launch_bounds__(256)
        joseph_ray_table_kernel(const PoseGeom* __restrict__ geo, const Dims dims, RayRec* __restrict__ table)
    {
        const int pose = blockIdx.y;
        const int64_t pix = (int64_t) blockIdx.x * blockDim.x + threadIdx.x;
        if (pix >= dims.pix_per_pose)
            return;
        const int u = (int) (pix % dims.det[0]);
        const int v = (int) (pix / dims.det[0]);
        float o[3], d[3];
        make_ray_px<D>(geo[pose], dims, u, v, o, d);
        JRay jr;
        joseph_setup<D>(o, d, dims.vol, jr);
        table[(int64_t) pose * dims.pix_per_pose + pix] = make_ray_rec<D>(jr);
    }

    hipError_t joseph_ray_table(const PoseGeom* geo, const Dims& dims, RayRec* table, hipStream_t s)
    {
        const dim3 grid((unsigned) ((dims.pix_per_pose + 255) / 256), (unsigned) dims.nposes);
        if (dims.dim == 2)
            hipLaunchKernelGGL((joseph_ray_table_kernel<2>), grid, dim3(256), 0, s, geo, dims, table);
        else
            hipLaunchKernelGGL((joseph_ray_table_kernel<3>), grid, dim3(256), 0, s, geo, dims, table);
        return hipGetLastError();
    }

    // ---------------------------------------------------------------------------------
    // fixed-point scale of one back projection

    // Largest |y| in two stages without atomics: each block writes the maximum of its share
    // as the bit pattern of |y| (for non-negative floats, and NaN above infinity, the integer
    // order of the bits is the order of the values), one block reduces those.
    template <class U>
    __device__ inline U block_max_bits(U m)
    {
        for (int off = 32; off > 0; off >>= 1) {
            const U o = __shfl_down(m, off, 64);
            m = o > m ? o : m;
        }
        __shared__ U red[4];
        if ((threadIdx.x & 63) == 0)
            red[threadIdx.x >> 6] = m;
        __syncthreads();
        if (threadIdx.x == 0)
            for (int w = 1; w < (int) (blockDim.x >> 6); ++w)
                m = red[w] > m ? red[w] : m;
        return m; // valid in thread 0
    }

    template <class T>
    __global__ void __launch_bounds__(256)
        absmax_kernel(const T* __restrict__ x, int64_t n, unsigned int* __restrict__ partial)
    {
        unsigned int m = 0u;
        const int64_t stride = (int64_t) gridDim.x * blockDim.x;
        for (int64_t i = (int64_t) blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
            const unsigned int u = __float_as_uint(fabsf(up(x[i])));
            m = u > m ? u : m;
        }
        m = block_max_bits(m);
        if (threadIdx.x == 0)
            partial[blockIdx.x] = m;
    }

    // the fixed-point exponent eg for the slot bound of the operator, from the partial maxima
    __global__ void __launch_bounds__(256)
        scale_kernel(const unsigned int* __restrict__ partial, int nblocks, double bound, int* __restrict__ fixexp)
    {
        unsigned int m = 0u;
        for (int i = threadIdx.x; i < nblocks; i += blockDim.x)
            m = partial[i] > m ? partial[i] : m;
        m = block_max_bits(m);
        if (threadIdx.x == 0)
            fixexp[0] = fixed_point_exponent(__uint_as_float(m), bound);
    }

    // The same in FP64: the bit patterns of non-negative doubles order as the values do.
    __global__ void __launch_bounds__(256)
        absmax64_kernel(const double* __restrict__ x, int64_t n, unsigned long long* __restrict__ partial)
    {
        unsigned long long m = 0ull;
        const int64_t stride = (int64_t) gridDim.x * blockDim.x;
        for (int64_t i = (int64_t) blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
            const unsigned long long u = (unsigned long long) __double_as_longlong(fabs(x[i]));
            m = u > m ? u : m;
        }
        m = block_max_bits(m);
        if (threadIdx.x == 0)
            partial[blockIdx.x] = m;
    }

    __global__ void __launch_bounds__(256)
        scale64_kernel(const unsigned long long* __restrict__ partial, int nblocks, double bound,
                       int* __restrict__ fixexp)
    {
        unsigned long long m = 0ull;
        for (int i = threadIdx.x; i < nblocks; i += blockDim.x)
            m = partial[i] > m ? partial[i] : m;
        m = block_max_bits(m);
        if (threadIdx.x == 0)
            fixexp[0] = fixed_point_exponent(__longlong_as_double((long long) m), bound);
    }

    // d_maxbits holds kAbsmaxBlocks 64-bit partial maxima (FP32 and bfloat16 use its first half
    // as 32-bit ones)
    template <class T>
    hipError_t fixed_point_exponent_launch(const T* y, int64_t n, double slot_bound, unsigned int* d_maxbits,
                                           int* d_fixexp, hipStream_t s)
    {
        // 16 elements per thread at least
        int64_t blocks = (n + 256 * 16 - 1) / (256 * 16);
        blocks = blocks > kAbsmaxBlocks ? kAbsmaxBlocks : (blocks < 1 ? 1 : blocks);
        if constexpr (sizeof(T) == 8) {
            unsigned long long* part = reinterpret_cast<unsigned long long*>(d_maxbits);
            hipLaunchKernelGGL(absmax64_kernel, dim3((unsigned) blocks), dim3(256), 0, s, y, n, part);
            hipLaunchKernelGGL(scale64_kernel, dim3(1), dim3(256), 0, s, part, (int) blocks, slot_bound, d_fixexp);
        } else {
            hipLaunchKernelGGL((absmax_kernel<T>), dim3((unsigned) blocks), dim3(256), 0, s, y, n, d_maxbits);
            hipLaunchKernelGGL(scale_kernel, dim3(1), dim3(256), 0, s, d_maxbits, (int) blocks, slot_bound, d_fixexp);
        }
        return hipGetLastError();
    }

    template hipError_t fixed_point_exponent_launch<float>(const float*, int64_t, double, unsigned int*, int*,
                                                           hipStream_t);
    template hipError_t fixed_point_exponent_launch<bf16>(const bf16*, int64_t, double, unsigned int*, int*,
                                                          hipStream_t);
    template hipError_t fixed_point_exponent_launch<double>(const double*, int64_t, double, unsigned int*, int*,
                                                            hipStream_t);

    // ---------------------------------------------------------------------------------
    // the tiled back projection

    // The candidate rays of one pose batch of one tile, all added into the tile's slots.
    // Candidate k (k < total) is pixel k - pref[j] of pose j's rectangle, j the last pose with
    // pref[j] <= k. Each lane walks its candidates k = tid, tid + 256, ... in order, so it keeps
    // its pose j (and j's rectangle) in registers and moves on when k passes pref[j + 1]. The
    // ray state and y of the next candidate are loaded while the current one is added (two
    // register sets taken in turn).
    template <int D, int A, bool LOW, class ST>
    __device__ __forceinline__ void adj_tile_batch(const RayRec* __restrict__ table, const ST* __restrict__ sino,
                                                   const int det0, const CandRect* s_rect, const int* s_pref,
                                                   const int total, const TileBox& t, const TileOff& to,
                                                   const int eg)
    {
        using YT = typename Acc<ST>::type; // FP64 for FP64 sinograms, else FP32
        const int tid = threadIdx.x;
        int j = 0, pj = 0, pn = s_pref[1];
        CandRect cr = s_rect[0];
        auto fetch = [&](int k, RayRec& r, YT& yv) {
            while (k >= pn) {
                ++j;
                pj = pn;
                pn = s_pref[j + 1];
                cr = s_rect[j];
            }
            int row, col;
            rect_split(k - pj, cr.nu, cr.rnu, row, col);
            const int64_t ray = cr.base + (int64_t) (mul24((uint32_t) row, (uint32_t) det0) + (uint32_t) col);
            yv = up(sino[ray]);
            r = table[ray];
        };
        auto process = [&](const RayRec& r, YT yv) {
            if constexpr (sizeof(YT) == 8) {
                // FP64: each contribution is elsa's (y * len) * weight in FP64 (the weight in FP32,
                // as the one-thread-per-ray FP64 kernel forms it), rounded to the nearest multiple of
                // q; yq is y * len in units of q (a power-of-two scaling, exact)
                const double yq = ldexp(yv * (double) r.len, 62 - eg);
                tile_candidate_fx<D, A, LOW, TAP_WEIGHT>(r, t, to, yq, [&](uint32_t off, float w) {
                    __hip_atomic_fetch_add((lds_u64*) (uintptr_t) off, fixed_round(yq, (double) w), __ATOMIC_RELAXED,
                                           __HIP_MEMORY_SCOPE_WORKGROUP);
                });
            } else {
                // FP32 (and bfloat16, widened): elsa's FP32 product y * len in units of q (exact),
                // times the exact weights, rounded once to the nearest multiple of q
                const double yq = ldexp((double) (yv * r.len), 62 - eg);
                // off: the slot's LDS byte address (the tile's base included)
                tile_candidate_fx<D, A, LOW, TAP_FIXED>(r, t, to, yq, [&](uint32_t off, unsigned long long v) {
                    __hip_atomic_fetch_add((lds_u64*) (uintptr_t) off, v, __ATOMIC_RELAXED,
                                           __HIP_MEMORY_SCOPE_WORKGROUP);
                });
            }
        };
        RayRec ra{}, rb{};
        YT ya = 0, yb = 0;
        if (tid < total)
            fetch(tid, ra, ya);
        for (int k = tid; k < total; k += 2 * kThreads) {
            const int k1 = k + kThreads, k2 = k + 2 * kThreads;
            if (k1 < total)
                fetch(k1, rb, yb);
            process(ra, ya);
            if (k2 < total)
                fetch(k2, ra, ya);
            if (k1 < total)
                process(rb, yb);
        }
    }

    // One tile of the back projection along driving axis A (tile < ntA * ntB * ntC).
    template <int D, int A, class ST = float>
    __device__ __forceinline__ void adj_tiled_block(const PoseGeom* __restrict__ geo,
                                                    const RayRec* __restrict__ table, const Dims& dims,
                                                    const int* __restrict__ pose_axes, const TileCfg& tc,
                                                    const int ntA, const int ntB, const int tile, const int batch,
                                                    const int* __restrict__ fixexp, const ST* __restrict__ sino,
                                                    typename Acc<ST>::type* __restrict__ vol)
    {
        extern __shared__ long long lds[];

        const int tid = threadIdx.x;
        const TileBox t = make_tile<D, A>(dims, tc, tile % ntA, (tile / ntA) % ntB, tile / (ntA * ntB));
        const TileOff to = tile_offsets(t, (uint32_t) (uintptr_t) (lds_char*) lds);
        const bool low = tile_low_face<D>(t);
        long long* iacc = lds;
        const BatchLds bl = batch_lds(lds, t.nslots, batch);
        CandRect* s_rect = bl.rect;
        int *s_cnt = bl.cnt, *s_pref = bl.pref;
        for (int i = tid; i < t.nslots; i += kThreads)
            iacc[i] = 0;
        const int eg = fixexp[0];
        __syncthreads();

        for (int pbase = 0; eg != kNonFinite && pbase < dims.nposes; pbase += batch) {
            if (tid < batch) {
                const int p = pbase + tid;
                int u0 = 0, u1 = -1, v0 = 0, v1 = -1;
                int cnt = 0;
                if (p < dims.nposes && (pose_axes[p] & (1 << A))) {
                    if (tile_bbox<D, A>(geo[p], t, dims, u0, u1, v0, v1))
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
            if (low)
                adj_tile_batch<D, A, true>(table, sino, dims.det[0], s_rect, s_pref, total, t, to, eg);
            else
                adj_tile_batch<D, A, false>(table, sino, dims.det[0], s_rect, s_pref, total, t, to, eg);
            __syncthreads();
        }

        // add the owned voxels of the tile to the volume
        const double q = eg == kNonFinite ? __longlong_as_double(0x7ff8000000000000ll) : ldexp(1.0, eg - 62);
        for (int i = tid; i < t.nslots; i += kThreads) {
            const int64_t lin = tile_slot_voxel<D, A>(t, dims, i);
            if (lin >= 0)
                vol[lin] += (typename Acc<ST>::type) ((double) iacc[i] * q);
        }
    }

    template <int D, int A, class ST>
    __global__ void __launch_bounds__(kThreads)
        joseph_adj_tiled_kernel(const PoseGeom* __restrict__ geo, const RayRec* __restrict__ table, const Dims dims,
                                const int* __restrict__ pose_axes, const TileCfg tc, const int ntA, const int ntB,
                                const int ntC, const int tiles_per_xcd, const int batch,
                                const int* __restrict__ fixexp, const ST* __restrict__ sino,
                                typename Acc<ST>::type* __restrict__ vol)
    {
        // consecutive workgroups land on different XCDs: give each XCD a contiguous block of tiles
        const int tile = (int) (blockIdx.x % 8u) * tiles_per_xcd + (int) (blockIdx.x / 8u);
        if (tile >= ntA * ntB * ntC)
            return;
        adj_tiled_block<D, A, ST>(geo, table, dims, pose_axes, tc, ntA, ntB, tile, batch, fixexp, sino, vol);
    }

    // the tiles of axis A of many operators in one launch (block b: tile b - job.block0 of job
    // block_job[b])
    template <int D, int A>
    __global__ void __launch_bounds__(kThreads)
        joseph_adj_tiled_batch_kernel(const AdjJob* __restrict__ jobs, const int* __restrict__ block_job,
                                      const int per_xcd, const int64_t total, const int* __restrict__ fixexp,
                                      const float* __restrict__ sinos, float* __restrict__ vols)
    {
        const int64_t id = (int64_t) (blockIdx.x % 8u) * per_xcd + (blockIdx.x / 8u);
        if (id >= total)
            return;
        const int j = block_job[id];
        const AdjJob& J = jobs[j];
        adj_tiled_block<D, A, float>(J.geo, J.table, J.dims, J.pose_axes, J.tc, J.ntA, J.ntB, (int) (id - J.block0),
                                     J.batch, fixexp + J.item, sinos + J.sino_off, vols + J.vol_off);
    }

    // max|y| per operator of a batch: block b writes the maximum of its share of segment
    // block_seg[b] to partial[b]; then one block per segment reduces its partials
    __global__ void __launch_bounds__(256)
        absmax_batch_kernel(const float* __restrict__ y, const SegJob* __restrict__ segs,
                            const int* __restrict__ block_seg, unsigned int* __restrict__ partial)
    {
        const SegJob& S = segs[block_seg[blockIdx.x]];
        const int64_t b = (int64_t) blockIdx.x - S.block0;
        unsigned int m = 0u;
        const int64_t stride = (int64_t) S.nblocks * 256;
        for (int64_t i = b * 256 + threadIdx.x; i < S.n; i += stride) {
            const unsigned int u = __float_as_uint(fabsf(y[S.off + i]));
            m = u > m ? u : m;
        }
        m = block_max_bits(m);
        if (threadIdx.x == 0)
            partial[blockIdx.x] = m;
    }

    __global__ void __launch_bounds__(256)
        scale_batch_kernel(const unsigned int* __restrict__ partial, const SegJob* __restrict__ segs,
                           const double* __restrict__ bounds, int* __restrict__ fixexp)
    {
        const int j = blockIdx.x;
        const SegJob& S = segs[j];
        unsigned int m = 0u;
        for (int i = threadIdx.x; i < S.nblocks; i += blockDim.x) {
            const unsigned int u = partial[S.block0 + i];
            m = u > m ? u : m;
        }
        m = block_max_bits(m);
        if (threadIdx.x == 0)
            fixexp[j] = fixed_point_exponent(__uint_as_float(m), bounds[j]);
    }

    // Which driving axes occur among the rays of each pose (bit a set when some ray that hits
    // the volume drives along a).
    __global__ void __launch_bounds__(256)
        pose_axes_kernel(const RayRec* __restrict__ table, const Dims dims, int* __restrict__ pose_axes)
    {
        const int pose = blockIdx.y;
        const int64_t pix = (int64_t) blockIdx.x * blockDim.x + threadIdx.x;
        if (pix >= dims.pix_per_pose)
            return;
        const RayRec r = table[(int64_t) pose * dims.pix_per_pose + pix];
        if (r.nsteps > 0)
            atomicOr(pose_axes + pose, 1 << (r.meta & 3));
    }

    hipError_t joseph_pose_axes(const RayRec* table, const Dims& dims, int* pose_axes, hipStream_t s)
    {
        hipError_t e = launch_fill<int>(pose_axes, dims.nposes, 0, s);
        if (e != hipSuccess)
            return e;
        const dim3 grid((unsigned) ((dims.pix_per_pose + 255) / 256), (unsigned) dims.nposes);
        hipLaunchKernelGGL(pose_axes_kernel, grid, dim3(256), 0, s, table, dims, pose_axes);
        return hipGetLastError();
    }

    size_t joseph_tile_lds_bytes(const Dims& dims, const TileCfg& tc, int batch)
    {
        const int PC = dims.dim == 3 ? tc.TC + 2 : 1;
        return (size_t) tc.LA * PC * (tc.TB + 2) * sizeof(long long) + batch_lds_bytes(batch);
    }

    // ST: the sinogram's storage type (float, or bf16 for bfloat16 operators); the volume is
    // FP32 either way.
    template <class ST>
    hipError_t joseph_adjoint_tiled(const PoseGeom* geo, const RayRec* table, const Dims& dims, const int* pose_axes,
                                    int axes_present, const TileCfg& tc, int batch, double slot_bound,
                                    unsigned int* d_maxbits, int* d_fixexp, const ST* sino,
                                    typename Acc<ST>::type* vol, hipStream_t s)
    {
        using VT = typename Acc<ST>::type;
        if (batch < 1 || batch > kMaxPoseBatch)
            return hipErrorInvalidValue;
        hipError_t e = launch_fill<VT>(vol, dims.nvox, VT(0), s);
        if (e == hipSuccess)
            e = fixed_point_exponent_launch(sino, dims.nrays, slot_bound, d_maxbits, d_fixexp, s);
        if (e != hipSuccess)
            return e;
        const size_t lds = joseph_tile_lds_bytes(dims, tc, batch);
        for (int a = 0; a < dims.dim; ++a) {
            if (!(axes_present & (1 << a)))
                continue;
            int ntA, ntB, ntC;
            tile_counts(dims, tc, a, ntA, ntB, ntC);
            const int per_xcd = (ntA * ntB * ntC + 7) / 8;
            const dim3 grid((unsigned) (per_xcd * 8));
#define EHP_ADJ_TILED(Dv, Av)                                                                                      \
    hipLaunchKernelGGL((joseph_adj_tiled_kernel<Dv, Av, ST>), grid, dim3(kThreads), lds, s, geo, table, dims,         \
                       pose_axes, tc, ntA, ntB, ntC, per_xcd, batch, d_fixexp, sino, vol)
            if (dims.dim == 2) {
                if (a == 0)
                    EHP_ADJ_TILED(2, 0);
                else
                    EHP_ADJ_TILED(2, 1);
            } else {
                if (a == 0)
                    EHP_ADJ_TILED(3, 0);
                else if (a == 1)
                    EHP_ADJ_TILED(3, 1);
                else
                    EHP_ADJ_TILED(3, 2);
            }
#undef EHP_ADJ_TILED
            e = hipGetLastError();
            if (e != hipSuccess)
                return e;
        }
        return hipSuccess;
    }

    template hipError_t joseph_adjoint_tiled<float>(const PoseGeom*, const RayRec*, const Dims&, const int*, int,
                                                    const TileCfg&, int, double, unsigned int*, int*, const float*,
                                                    float*, hipStream_t);
    template hipError_t joseph_adjoint_tiled<bf16>(const PoseGeom*, const RayRec*, const Dims&, const int*, int,
                                                   const TileCfg&, int, double, unsigned int*, int*, const bf16*,
                                                   float*, hipStream_t);
    template hipError_t joseph_adjoint_tiled<double>(const PoseGeom*, const RayRec*, const Dims&, const int*, int,
                                                     const TileCfg&, int, double, unsigned int*, int*, const double*,
                                                     double*, hipStream_t);

    // Segments of a batch's sinograms (absmax) and the tiles of every operator per driving
    // axis (one job per operator that has rays of that axis).
    // One segment per operator (offset and size in a packed vector), each reduced by up to 256
    // blocks of the batched absmax.
    void batch_segments(const int64_t* offs, const int64_t* sizes, int n, std::vector<SegJob>& segs,
                        std::vector<int>& sblocks)
    {
        segs.clear();
        sblocks.clear();
        for (int j = 0; j < n; ++j) {
            SegJob sg{};
            sg.off = offs[j];
            sg.n = sizes[j];
            int64_t nb = (sg.n + 256 * 16 - 1) / (256 * 16);
            nb = nb > 256 ? 256 : (nb < 1 ? 1 : nb);
            sg.nblocks = (int) nb;
            sg.block0 = (int64_t) sblocks.size();
            sblocks.insert(sblocks.end(), (size_t) nb, j);
            segs.push_back(sg);
        }
    }

    hipError_t fixed_point_exponent_batch(int nitems, const SegJob* segs, const int* sblocks, int64_t nsblocks,
                                          const double* bounds, unsigned int* d_maxbits, int* d_fixexp,
                                          const float* y, hipStream_t s)
    {
        if (nsblocks > 0x7fffffff)
            return hipErrorInvalidValue;
        hipLaunchKernelGGL(absmax_batch_kernel, dim3((unsigned) nsblocks), dim3(256), 0, s, y, segs, sblocks,
                           d_maxbits);
        hipLaunchKernelGGL(scale_batch_kernel, dim3((unsigned) nitems), dim3(256), 0, s, d_maxbits, segs, bounds,
                           d_fixexp);
        return hipGetLastError();
    }

    void joseph_adjoint_batch_jobs(const JosephBatchItem* items, int n, std::vector<SegJob>& segs,
                                   std::vector<int>& sblocks, std::vector<AdjJob> ajobs[3],
                                   std::vector<int> ablocks[3], size_t lds[3])
    {
        std::vector<int64_t> offs(n), sizes(n);
        for (int j = 0; j < n; ++j) {
            offs[j] = items[j].sino_off;
            sizes[j] = items[j].dims.nrays;
        }
        batch_segments(offs.data(), sizes.data(), n, segs, sblocks);
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
                AdjJob aj{};
                aj.geo = it.geo;
                aj.table = it.table;
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
                const size_t need = joseph_tile_lds_bytes(it.dims, it.tile);
                lds[a] = need > lds[a] ? need : lds[a];
            }
        }
    }

    hipError_t joseph_adjoint_batch(int dim, int nitems, const SegJob* segs, const int* sblocks, int64_t nsblocks,
                                    const double* bounds, const AdjJob* const ajobs[3], const int* const ablocks[3],
                                    const int64_t nablocks[3], const size_t lds[3], int64_t nvol_total,
                                    unsigned int* d_maxbits, int* d_fixexp, const float* sinos, float* vols,
                                    hipStream_t s)
    {
        hipError_t e = launch_fill<float>(vols, nvol_total, 0.0f, s);
        if (e == hipSuccess)
            e = fixed_point_exponent_batch(nitems, segs, sblocks, nsblocks, bounds, d_maxbits, d_fixexp, sinos, s);
        if (e != hipSuccess)
            return e;
        for (int a = 0; a < dim; ++a) {
            const int64_t total = nablocks[a];
            if (total == 0)
                continue;
            const int64_t per_xcd = (total + 7) / 8;
            if (per_xcd * 8 > (int64_t) 0x7fffffff)
                return hipErrorInvalidValue;
            const dim3 grid((unsigned) (per_xcd * 8));
#define EHP_ADJ_BATCH(Dv, Av)                                                                                      \
    hipLaunchKernelGGL((joseph_adj_tiled_batch_kernel<Dv, Av>), grid, dim3(kThreads), lds[a], s, ajobs[a], ablocks[a], \
                       (int) per_xcd, total, d_fixexp, sinos, vols)
            if (dim == 2) {
                if (a == 0)
                    EHP_ADJ_BATCH(2, 0);
                else
                    EHP_ADJ_BATCH(2, 1);
            } else {
                if (a == 0)
                    EHP_ADJ_BATCH(3, 0);
                else if (a == 1)
                    EHP_ADJ_BATCH(3, 1);
                else
                    EHP_ADJ_BATCH(3, 2);
            }
#undef EHP_ADJ_BATCH
            e = hipGetLastError();
            if (e != hipSuccess)
                return e;
        }
        return hipGetLastError();
    }
} // namespace ehp
