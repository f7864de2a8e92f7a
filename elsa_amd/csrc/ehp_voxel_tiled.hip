// ehp_voxel_tiled.hip -- voxel-driven forward projection with LDS-privatised detector tiles
// (FP32 arithmetic with the volume and sinogram stored in FP32 or bfloat16; FP64).
//
// The forward of elsa's voxel projectors (VoxelComputation.h:110-152: every voxel adds
// x * basis(distance) to the detector pixels of its footprint) without global atomics. One
// workgroup owns one detector tile of one pose and accumulates it in LDS: it walks the planes
// of the volume along the pose's driving axis, takes on each plane the voxels whose footprint
// can reach the tile (frustum_plane_range), computes each one's footprint exactly as the
// one-thread-per-voxel kernels do and adds the taps that fall in the tile. Every pixel then
// belongs to one tile and is written once, so the output needs no clearing. Tiles are ordered
// tile first, pose second: the poses of a tile, which read nearly the same column of the
// volume, run together on one XCD.
//
// Sums are 64-bit fixed point as in the tiled back projections: one exponent per call from
// max|x| and a bound on what one pixel can receive, each tap's x * weight rounded once to the
// unit (fixed_round; exact product in FP64 for FP32 data). With bfloat16 storage the voxels
// are widened exactly and each pixel's sum is rounded to bfloat16 once, so the result is the
// FP32 operator's on the same (bfloat16) volume, rounded once. FP64 operators take an FP64
// table and form each weight as the one-thread-per-voxel FP64 kernel does (tap_weight, the
// divisions of VoxelComputation.h), so their terms are that kernel's.
#include "ehp_kernels.hpp"

#include <type_traits>

namespace ehp
{
    namespace
    {
        constexpr int kThreads = 256;
        constexpr int kMaxLut = kVoxMaxLut;

        // The frustum rectangle of one plane: first voxel along b1 and b2, width along b1 and its
        // reciprocal (rect_split), and the linear index of its first voxel (volumes below 2^31
        // voxels with strides below 2^24, whose indices are formed with 24-bit multiplications)
        struct PlaneRect {
            int blo, clo, nb;
            float rnb;
            uint32_t base;
        };
        // LDS of the plane tables of a chunk of kThreads planes: rectangles, counts and their prefix
        constexpr size_t kPlaneLds = kThreads * (sizeof(PlaneRect) + sizeof(int)) + (kThreads + 1) * sizeof(int);
    } // namespace

    // Tile-pose id (tile * nposes + pose) of one operator's forward, whose pose has driving axis
    // A; LT = float (FP32 and bfloat16 storage) or double (FP64) is the table's type.
    // The static LDS of the forward, declared once per kernel (each instantiation of a function
    // declaring its own would get its own copy): FP64 the table, FP32 its pair table
    // (lerp_table_entry, tap_weight_fast); the plane tables of a chunk.
    template <class LT>
    struct VoxFwdLds {
        LT* lut;
        PlaneRect* pr;
        int* cnt;
        int* pref;
    };

    template <int D, int A, class ST, class LT>
    __device__ __forceinline__ void vox_fwd_tiled_axis(const PoseGeom* __restrict__ geo, const Dims& dims,
                                                       const LT* __restrict__ lut, const int n_lut,
                                                       const float support, const int diff,
                                                       const float* __restrict__ pose_margin, const int TU,
                                                       const int TV, const int ntu, const int64_t id,
                                                       const int* __restrict__ fixexp,
                                                       const ST* __restrict__ vol, ST* __restrict__ sino,
                                                       const VoxFwdLds<LT>& L)
    {
        extern __shared__ long long lds[];
        LT* s_lut = L.lut;
        PlaneRect* s_pr = L.pr;
        int* s_cnt = L.cnt;
        int* s_pref = L.pref;
        const int tid = threadIdx.x;
        // id = tile * nposes + pose
        const int pose = (int) (id % dims.nposes);
        const int tile = (int) (id / dims.nposes);
        const int tu = tile % ntu, tv = tile / ntu;
        const int u0 = tu * TU, v0 = tv * TV;
        const int u1 = (u0 + TU < dims.det[0] ? u0 + TU : dims.det[0]) - 1;
        const int v1 = (v0 + TV < dims.det[1] ? v0 + TV : dims.det[1]) - 1;
        long long* acc = lds;
        const int RS = vox_acc_stride(TU, TV); // slot (v - v0, u - u0) at (v - v0) * RS + u - u0
        for (int i = tid; i < RS * TV; i += kThreads)
            acc[i] = 0;
        for (int i = tid; i < (sizeof(LT) == 8 ? n_lut : lerp_table_size(n_lut)); i += kThreads) {
            if constexpr (sizeof(LT) == 8)
                s_lut[i] = lut[i];
            else
                lerp_table_entry(lut, n_lut, i, s_lut);
        }
        const int eg = fixexp[0];
        const float lut_index_scale = (float) (n_lut - 1) / support;
        const bool lut_tail_zero = lut[n_lut - 1] == LT(0); // tap_weight_fast's TZ (FP32)
        const PoseGeom& g = geo[pose];
        float cdir[12];
        const int ncorner = frustum_corners<D>(g, u0, u1, v0, v1, pose_margin[pose], cdir);
        __syncthreads();

        constexpr int b1 = (D == 2) ? 1 - A : (A == 0 ? 1 : 0);
        constexpr int b2 = (D == 2) ? 2 : (A == 2 ? 1 : 2);
        const bool small = dims.nvox < (1ll << 31) && dims.strides[1] < (1 << 24) && dims.strides[2] < (1 << 24);
        const uint32_t sA = (uint32_t) dims.strides[A], s1 = (uint32_t) dims.strides[b1];
        const uint32_t s2 = D == 3 ? (uint32_t) dims.strides[b2] : 0u;
        // The planes in chunks of up to kThreads: one thread per plane computes the plane's frustum
        // rectangle (frustum_plane_range) into LDS, then every thread walks its candidates
        // k = tid, tid + kThreads, ... over all planes of the chunk in order, keeping its plane jp in
        // registers (candidate k is entry k - pref[jp] of plane jp's rectangle, jp the last plane with
        // pref[jp] <= k). A chunk holds fewer planes when its candidates could pass 2^31 - 1.
        const int64_t psz = (int64_t) dims.vol[b1] * (D == 3 ? dims.vol[b2] : 1);
        int pch = kThreads;
        while (pch > 1 && (int64_t) pch * psz > 0x7fffffffll)
            pch >>= 1;
        for (int pc0 = 0; eg != kNonFinite && pc0 < dims.vol[A]; pc0 += pch) {
            const int npl = dims.vol[A] - pc0 < pch ? dims.vol[A] - pc0 : pch;
            if (tid < npl) {
                int blo, bhi, clo, chi, cnt = 0;
                if (frustum_plane_range<D>(g, dims, A, pc0 + tid, cdir, ncorner, blo, bhi, clo, chi))
                    cnt = (bhi - blo + 1) * (chi - clo + 1);
                const int nb = cnt > 0 ? bhi - blo + 1 : 1;
                const uint32_t base = small && cnt > 0
                                          ? (uint32_t) (pc0 + tid) * sA + (uint32_t) blo * s1 + (uint32_t) clo * s2
                                          : 0u;
                s_pr[tid] = PlaneRect{blo, clo, nb, 1.0f / (float) nb, base};
                s_cnt[tid] = cnt;
            }
            __syncthreads();
            if (tid == 0) {
                s_pref[0] = 0;
                for (int q = 0; q < npl; ++q)
                    s_pref[q + 1] = s_pref[q] + s_cnt[q];
            }
            __syncthreads();
            const int total = s_pref[npl];
            int jp = 0, pj = 0, pn = s_pref[1];
            PlaneRect pr = s_pr[0];
            // candidate k's voxel c and its linear index (k increasing from call to call)
            auto locate = [&](int k, int* c) -> int64_t {
                while (k >= pn) {
                    ++jp;
                    pj = pn;
                    pn = s_pref[jp + 1];
                    pr = s_pr[jp];
                }
                int row, col;
                rect_split(k - pj, pr.nb, pr.rnb, row, col);
                c[A] = pc0 + jp;
                c[b1] = pr.blo + col;
                if (D == 3)
                    c[b2] = pr.clo + row;
                return small ? (int64_t) (pr.base + mul24((uint32_t) col, s1) + mul24((uint32_t) row, s2))
                             : c[0] + c[1] * dims.strides[1] + (int64_t) c[2] * dims.strides[2];
            };
            using XT = decltype(up(vol[0]));
            // the taps of voxel c with value xv
            auto visit = [&](const int* c, const XT xv) {
                if (xv == 0)
                    return;
                // FP64: footprint()'s divisions, as the one-thread-per-voxel FP64 kernel; FP32:
                // footprint_fast, as the gather back projection
                Footprint fp;
                float rs = 0.0f;
                if constexpr (sizeof(LT) == 8)
                    footprint<D>(g, dims, c[0], c[1], c[2], support, fp);
                else
                    footprint_fast<D>(g, dims, c[0], c[1], c[2], support, fp, rs);
                const int lo0 = fp.lo[0] > u0 ? fp.lo[0] : u0, hi0 = fp.hi[0] < u1 ? fp.hi[0] : u1;
                const int lo1 = fp.lo[1] > v0 ? fp.lo[1] : v0, hi1 = fp.hi[1] < v1 ? fp.hi[1] : v1;
                if (lo0 > hi0 || lo1 > hi1)
                    return;
                // x in units of q (a power-of-two scaling, exact); each tap adds x * w rounded once
                // (|x w| < 2^51 units: the pixel bound is at least 4096 times the table's bound)
                const double xq = ldexp((double) xv, 62 - eg);
                if constexpr (sizeof(LT) == 8) {
                    // FP64: the weights of visit() and tap_weight, as the one-thread-per-voxel kernel
                    for (int j = lo1; j <= hi1; ++j) {
                        const float d1 = D == 3 ? (fp.center[1] - (float) j) / fp.scaling : 0.0f;
                        for (int i = lo0; i <= hi0; ++i) {
                            const float d0 = (fp.center[0] - (float) i) / fp.scaling;
                            const float dist = D == 2 ? sqrtf(d0 * d0) : sqrtf(d0 * d0 + d1 * d1);
                            const double w = tap_weight<D, double>(s_lut, n_lut, support, dist, d0, diff != 0);
                            if (w != 0.0)
                                atomicAdd(reinterpret_cast<unsigned long long*>(acc + (j - v0) * RS + (i - u0)),
                                          fixed_round(xq, w));
                        }
                    }
                } else {
                    // the distance as visit() forms it, with the reciprocal rs of the scaling per
                    // voxel instead of a division per tap (and tap_weight_fast's square root and
                    // table index): within a few ulp of the one-thread-per-voxel kernels' weights,
                    // and the gather back projection's weights bit for bit. Two columns per step, so
                    // that their table reads are in flight together (the second is skipped past hi0).
                    // A table ending in 0 takes tap_weight_fast's TZ form.
                    auto taps = [&](auto tzc) {
                        constexpr bool TZ = decltype(tzc)::value;
                        for (int j = lo1; j <= hi1; ++j) {
                            const float d1 = D == 3 ? (fp.center[1] - (float) j) * rs : 0.0f;
                            const float d1s = d1 * d1;
                            unsigned long long* row = reinterpret_cast<unsigned long long*>(acc + (j - v0) * RS);
                            float fi = (float) lo0; // i as a float (exact), without a conversion per tap
                            for (int i = lo0; i <= hi0; i += 2, fi += 2.0f) {
                                const float d0a = (fp.center[0] - fi) * rs;
                                const float d0b = (fp.center[0] - (fi + 1.0f)) * rs;
                                const float wa = tap_weight_fast<D, TZ>(s_lut, n_lut, lut_index_scale,
                                                                        __builtin_fmaf(d0a, d0a, d1s), d0a, diff != 0);
                                const float wb = tap_weight_fast<D, TZ>(s_lut, n_lut, lut_index_scale,
                                                                        __builtin_fmaf(d0b, d0b, d1s), d0b, diff != 0);
                                if (wa != 0.0f)
                                    atomicAdd(row + (i - u0), fixed_round(xq, (double) wa));
                                if (i < hi0 && wb != 0.0f)
                                    atomicAdd(row + (i + 1 - u0), fixed_round(xq, (double) wb));
                            }
                        }
                    };
                    if (lut_tail_zero)
                        taps(std::true_type{});
                    else
                        taps(std::false_type{});
                }
            };
            // each lane loads its next candidate's value before it visits the current one, so that
            // the load's latency is spent on that visit
            int cur[3] = {0, 0, 0};
            XT xcur = 0;
            if (tid < total)
                xcur = up(vol[locate(tid, cur)]);
            for (int k = tid; k < total; k += kThreads) {
                int nxt[3] = {0, 0, 0};
                XT xnxt = 0;
                if (k + kThreads < total)
                    xnxt = up(vol[locate(k + kThreads, nxt)]);
                visit(cur, xcur);
                cur[0] = nxt[0];
                cur[1] = nxt[1];
                cur[2] = nxt[2];
                xcur = xnxt;
            }
            __syncthreads(); // the next chunk rewrites the plane tables
        }
        const double q = eg == kNonFinite ? __longlong_as_double(0x7ff8000000000000ll) : ldexp(1.0, eg - 62);
        ST* sp = sino + (int64_t) pose * dims.pix_per_pose;
        for (int i = tid; i < TU * TV; i += kThreads) {
            const int u = u0 + i % TU, v = v0 + i / TU;
            if (u <= u1 && v <= v1)
                sp[(int64_t) v * dims.det[0] + u] =
                    down<ST>((typename Acc<ST>::type) ((double) acc[(v - v0) * RS + (u - u0)] * q));
        }
    }

    // Tile-pose id of one operator's forward: the pose's driving axis as a template parameter
    template <int D, class ST, class LT = float>
    __device__ __forceinline__ void vox_fwd_tiled_block(const PoseGeom* __restrict__ geo, const Dims& dims,
                                                        const LT* __restrict__ lut, const int n_lut,
                                                        const float support, const int diff,
                                                        const int* __restrict__ pose_axis,
                                                        const float* __restrict__ pose_margin, const int TU,
                                                        const int TV, const int ntu, const int64_t id,
                                                        const int* __restrict__ fixexp,
                                                        const ST* __restrict__ vol, ST* __restrict__ sino)
    {
        __shared__ LT s_lut[sizeof(LT) == 8 ? kMaxLut : 2 * kVoxLutSpan];
        __shared__ PlaneRect s_pr[kThreads];
        __shared__ int s_cnt[kThreads];
        __shared__ int s_pref[kThreads + 1];
        const VoxFwdLds<LT> L{s_lut, s_pr, s_cnt, s_pref};
        const int A = pose_axis[(int) (id % dims.nposes)];
        if (A == 0)
            vox_fwd_tiled_axis<D, 0, ST, LT>(geo, dims, lut, n_lut, support, diff, pose_margin, TU, TV, ntu, id,
                                             fixexp, vol, sino, L);
        else if (D == 2 || A == 1)
            vox_fwd_tiled_axis<D, 1, ST, LT>(geo, dims, lut, n_lut, support, diff, pose_margin, TU, TV, ntu, id,
                                             fixexp, vol, sino, L);
        else
            vox_fwd_tiled_axis<D, D == 3 ? 2 : 1, ST, LT>(geo, dims, lut, n_lut, support, diff, pose_margin, TU, TV,
                                                          ntu, id, fixexp, vol, sino, L);
    }

    // DIFF: the phase-contrast projectors' weights (a compile-time flag, so that the taps of the
    // others carry no select)
    template <int D, class ST, class LT, bool DIFF>
    __global__ void __launch_bounds__(kThreads)
        voxel_fwd_tiled_kernel(const PoseGeom* __restrict__ geo, const Dims dims, const LT* __restrict__ lut,
                               const int n_lut, const float support, const int* __restrict__ pose_axis,
                               const float* __restrict__ pose_margin, const int TU, const int TV, const int ntu,
                               const int per_xcd, const int64_t total, const int* __restrict__ fixexp,
                               const ST* __restrict__ vol, ST* __restrict__ sino)
    {
        const int64_t id = (int64_t) (blockIdx.x % 8u) * per_xcd + (blockIdx.x / 8u);
        if (id >= total)
            return;
        vox_fwd_tiled_block<D, ST, LT>(geo, dims, lut, n_lut, support, DIFF ? 1 : 0, pose_axis, pose_margin, TU, TV,
                                       ntu, id, fixexp, vol, sino);
    }

    // the forwards of many operators in one launch (block b: tile-pose b - job.block0 of job
    // block_job[b], as that operator's own launch would run it)
    template <int D>
    __global__ void __launch_bounds__(kThreads)
        voxel_fwd_tiled_batch_kernel(const VoxFwdJob* __restrict__ jobs, const int* __restrict__ block_job,
                                     const int per_xcd, const int64_t total, const int* __restrict__ fixexp,
                                     const float* __restrict__ vols, float* __restrict__ sinos)
    {
        const int64_t id = (int64_t) (blockIdx.x % 8u) * per_xcd + (blockIdx.x / 8u);
        if (id >= total)
            return;
        const VoxFwdJob& J = jobs[block_job[id]];
        vox_fwd_tiled_block<D, float>(J.geo, J.dims, J.lut, J.n_lut, J.support, J.diff, J.pose_axis, J.pose_margin,
                                      J.TU, J.TV, J.ntu, id - J.block0, fixexp + J.item,
                                      vols + J.vol_off, sinos + J.sino_off);
    }

    // dynamic LDS (the tile's sums) plus the table: FP64 entries (lut_bytes 8) or the FP32 pair
    // table (lut_bytes 4), 8 bytes per entry either way
    size_t voxel_tile_lds_bytes(const TileCfg& tc, size_t lut_bytes)
    {
        return (size_t) vox_acc_stride(tc.LA, tc.TB) * tc.TB * sizeof(long long) +
               (lut_bytes == 8 ? (size_t) kMaxLut * 8 : (size_t) kVoxLutSpan * 2 * lut_bytes) + kPlaneLds;
    }

    int voxel_max_lut()
    {
        return kMaxLut;
    }

    // tc.LA x tc.TB: tile width x height in pixels (height 1 in 2D); LT: the table's type (FP64
    // for FP64 operators, else FP32)
    template <class ST, class LT>
    hipError_t voxel_forward_tiled(const PoseGeom* geo, const Dims& dims, const LT* lut, int n_lut, float support,
                                   int diff, const int* pose_axis, const float* pose_margin, const TileCfg& tc,
                                   double slot_bound, unsigned int* d_maxbits, int* d_fixexp, const ST* vol, ST* sino,
                                   hipStream_t s)
    {
        if (n_lut < 2 || n_lut > kMaxLut)
            return hipErrorInvalidValue;
        hipError_t e = fixed_point_exponent_launch(vol, dims.nvox, slot_bound, d_maxbits, d_fixexp, s);
        if (e != hipSuccess)
            return e;
        const int TU = tc.LA, TV = dims.dim == 3 ? tc.TB : 1;
        const int ntu = (dims.det[0] + TU - 1) / TU, ntv = (dims.det[1] + TV - 1) / TV;
        const int64_t total = (int64_t) ntu * ntv * dims.nposes;
        const int64_t per_xcd = (total + 7) / 8;
        if (per_xcd * 8 > (int64_t) 0x7fffffff)
            return hipErrorInvalidValue;
        const dim3 grid((unsigned) (per_xcd * 8));
        const size_t lds = (size_t) vox_acc_stride(TU, TV) * TV * sizeof(long long);
        auto launch = [&](auto kernel) {
            hipLaunchKernelGGL(kernel, grid, dim3(kThreads), lds, s, geo, dims, lut, n_lut, support, pose_axis,
                               pose_margin, TU, TV, ntu, (int) per_xcd, total, d_fixexp, vol, sino);
        };
        if (dims.dim == 2)
            diff ? launch(voxel_fwd_tiled_kernel<2, ST, LT, true>) : launch(voxel_fwd_tiled_kernel<2, ST, LT, false>);
        else
            diff ? launch(voxel_fwd_tiled_kernel<3, ST, LT, true>) : launch(voxel_fwd_tiled_kernel<3, ST, LT, false>);
        return hipGetLastError();
    }

    template hipError_t voxel_forward_tiled<float, float>(const PoseGeom*, const Dims&, const float*, int, float,
                                                          int, const int*, const float*, const TileCfg&, double,
                                                          unsigned int*, int*, const float*, float*, hipStream_t);
    template hipError_t voxel_forward_tiled<bf16, float>(const PoseGeom*, const Dims&, const float*, int, float,
                                                         int, const int*, const float*, const TileCfg&, double,
                                                         unsigned int*, int*, const bf16*, bf16*, hipStream_t);
    template hipError_t voxel_forward_tiled<double, double>(const PoseGeom*, const Dims&, const double*, int, float,
                                                            int, const int*, const float*, const TileCfg&, double,
                                                            unsigned int*, int*, const double*, double*,
                                                            hipStream_t);

    // Jobs of a batched voxel forward: one per operator, its tile-pose blocks back to back; the
    // launch's dynamic LDS is the largest tile's.
    void voxel_forward_batch_jobs(const VoxBatchItem* items, int n, std::vector<VoxFwdJob>& jobs,
                                  std::vector<int>& blocks, size_t& lds)
    {
        jobs.clear();
        blocks.clear();
        lds = 0;
        for (int j = 0; j < n; ++j) {
            const VoxBatchItem& it = items[j];
            VoxFwdJob J{};
            J.geo = it.geo;
            J.dims = it.dims;
            J.lut = it.lut;
            J.n_lut = it.n_lut;
            J.support = it.support;
            J.diff = it.diff;
            J.pose_axis = it.pose_axes;
            J.pose_margin = it.pose_margin;
            J.TU = it.tile.LA;
            J.TV = it.dims.dim == 3 ? it.tile.TB : 1;
            J.ntu = (it.dims.det[0] + J.TU - 1) / J.TU;
            const int ntv = (it.dims.det[1] + J.TV - 1) / J.TV;
            J.item = j;
            J.vol_off = it.vol_off;
            J.sino_off = it.sino_off;
            J.block0 = (int64_t) blocks.size();
            blocks.insert(blocks.end(), (size_t) J.ntu * ntv * it.dims.nposes, (int) jobs.size());
            jobs.push_back(J);
            const size_t need = (size_t) vox_acc_stride(J.TU, J.TV) * J.TV * sizeof(long long);
            lds = need > lds ? need : lds;
        }
    }

    // The tile launch of a batched voxel forward (the exponents d_fixexp[item] computed from
    // each operator's volume by fixed_point_exponent_batch).
    hipError_t voxel_forward_batch(int dim, const VoxFwdJob* jobs, const int* blocks, int64_t nblocks, size_t lds,
                                   const int* d_fixexp, const float* vols, float* sinos, hipStream_t s)
    {
        const int64_t per_xcd = (nblocks + 7) / 8;
        if (per_xcd * 8 > (int64_t) 0x7fffffff)
            return hipErrorInvalidValue;
        if (nblocks == 0)
            return hipSuccess;
        const dim3 grid((unsigned) (per_xcd * 8));
        if (dim == 2)
            hipLaunchKernelGGL((voxel_fwd_tiled_batch_kernel<2>), grid, dim3(kThreads), lds, s, jobs, blocks,
                               (int) per_xcd, nblocks, d_fixexp, vols, sinos);
        else
            hipLaunchKernelGGL((voxel_fwd_tiled_batch_kernel<3>), grid, dim3(kThreads), lds, s, jobs, blocks,
                               (int) per_xcd, nblocks, d_fixexp, vols, sinos);
        return hipGetLastError();
    }
} // namespace ehp
