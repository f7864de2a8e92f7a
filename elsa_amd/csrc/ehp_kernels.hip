// ehp_kernels.hip -- HIP kernels of the X-ray projectors for gfx950.
//
// Joseph (forward / matched adjoint), Siddon (forward / adjoint) and the voxel-driven
// blob / B-spline projector. The discretizations are those of elsa's CPU projectors:
//   Joseph  elsa/projectors/JosephsMethod.cpp:140-198 + DrivingDirectionTraversal.cpp:6-106
//   Siddon  elsa/projectors/SiddonsMethod.cpp:76-124 + TraverseAABB.cpp:6-184
//   voxel   elsa/projectors/VoxelComputation.h:40-185 + Luts.hpp:60-78
// The Joseph kernels place sample s of a ray at pos0 + s*step (one FMA per transverse axis)
// instead of accumulating the step, so the forward and the adjoint see identical weights.
#include "ehp_kernels.hpp"

#include <type_traits>

namespace ehp
{
    // ---------------------------------------------------------------------------------
    // fill

    template <class T>
    __global__ void __launch_bounds__(256) fill_kernel(T* __restrict__ p, int64_t n, T value)
    {
        const int64_t stride = (int64_t) gridDim.x * blockDim.x;
        for (int64_t i = (int64_t) blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
            p[i] = value;
    }

    template <class T>
    hipError_t launch_fill(T* p, int64_t n, T value, hipStream_t s)
    {
        if (n <= 0)
            return hipSuccess;
        int64_t blocks = (n + 255) / 256;
        if (blocks > 8192)
            blocks = 8192;
        hipLaunchKernelGGL(fill_kernel<T>, dim3((unsigned) blocks), dim3(256), 0, s, p, n, value);
        return hipGetLastError();
    }

    __global__ void __launch_bounds__(256) to_bf16_kernel(const float* __restrict__ src, bf16* __restrict__ dst, int64_t n)
    {
        const int64_t stride = (int64_t) gridDim.x * blockDim.x;
        for (int64_t i = (int64_t) blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
            dst[i] = float_to_bf16(src[i]);
    }

    hipError_t launch_to_bf16(const float* src, bf16* dst, int64_t n, hipStream_t s)
    {
        if (n <= 0)
            return hipSuccess;
        int64_t blocks = (n + 255) / 256;
        if (blocks > 8192)
            blocks = 8192;
        hipLaunchKernelGGL(to_bf16_kernel, dim3((unsigned) blocks), dim3(256), 0, s, src, dst, n);
        return hipGetLastError();
    }

    // ---------------------------------------------------------------------------------
    // Joseph: one thread per ray (pixel of one pose). Grid: x = pixel block, y = pose.

    // The samples of one ray with its driving axis A fixed at compile time.
    template <int D, int A, class T>
    __device__ inline T joseph_ray_forward(const JRay& jr, const Dims& dims, const T* __restrict__ vol)
    {
        // elsa's association: acc += (value * len) * weight per tap
        const T len = (T) jr.len;
        T acc = T(0);
        for (int s = 0; s < jr.nsteps; ++s) {
            JSample js;
            joseph_sample<D, A>(jr, s, js);
            joseph_taps<D>(js, dims.vol, dims.strides, [&](int64_t lin, float wt) { acc += vol[lin] * len * (T) wt; });
        }
        return acc;
    }

    template <int D, int A, class T>
    __device__ inline void joseph_ray_adjoint(const JRay& jr, const Dims& dims, T yl, T* __restrict__ vol)
    {
        for (int s = 0; s < jr.nsteps; ++s) {
            JSample js;
            joseph_sample<D, A>(jr, s, js);
            joseph_taps<D>(js, dims.vol, dims.strides, [&](int64_t lin, float wt) {
                if (wt != 0.0f)
                    atomic_add(vol + lin, yl * (T) wt);
            });
        }
    }

    template <int D, class T>
    __global__ void __launch_bounds__(256)
        joseph_fwd_kernel(const PoseGeom* __restrict__ geo, const Dims dims, const T* __restrict__ vol,
                          T* __restrict__ sino)
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
        T acc;
        if (jr.axis == 0)
            acc = joseph_ray_forward<D, 0, T>(jr, dims, vol);
        else if (D == 2 || jr.axis == 1)
            acc = joseph_ray_forward<D, 1, T>(jr, dims, vol);
        else
            acc = joseph_ray_forward<D, D - 1, T>(jr, dims, vol);
        sino[(int64_t) pose * dims.pix_per_pose + pix] = acc;
    }

    template <int D, class T>
    __global__ void __launch_bounds__(256)
        joseph_adj_kernel(const PoseGeom* __restrict__ geo, const Dims dims, const T* __restrict__ sino,
                          T* __restrict__ vol)
    {
        const int pose = blockIdx.y;
        const int64_t pix = (int64_t) blockIdx.x * blockDim.x + threadIdx.x;
        if (pix >= dims.pix_per_pose)
            return;
        const T yv = sino[(int64_t) pose * dims.pix_per_pose + pix];
        if (yv == T(0))
            return;
        const int u = (int) (pix % dims.det[0]);
        const int v = (int) (pix / dims.det[0]);
        float o[3], d[3];
        make_ray_px<D>(geo[pose], dims, u, v, o, d);
        JRay jr;
        joseph_setup<D>(o, d, dims.vol, jr);
        const T yl = yv * (T) jr.len;
        if (jr.axis == 0)
            joseph_ray_adjoint<D, 0, T>(jr, dims, yl, vol);
        else if (D == 2 || jr.axis == 1)
            joseph_ray_adjoint<D, 1, T>(jr, dims, yl, vol);
        else
            joseph_ray_adjoint<D, D - 1, T>(jr, dims, yl, vol);
    }

    // Per-ray traversal state (axis, sample count) and the total sample count.
    template <int D>
    __global__ void __launch_bounds__(256)
        joseph_info_kernel(const PoseGeom* __restrict__ geo, const Dims dims, int32_t* __restrict__ axis,
                           int64_t* __restrict__ nsteps, unsigned long long* __restrict__ total)
    {
        const int pose = blockIdx.y;
        const int64_t pix = (int64_t) blockIdx.x * blockDim.x + threadIdx.x;
        int n = 0;
        if (pix < dims.pix_per_pose) {
            const int u = (int) (pix % dims.det[0]);
            const int v = (int) (pix / dims.det[0]);
            float o[3], d[3];
            make_ray_px<D>(geo[pose], dims, u, v, o, d);
            JRay jr;
            joseph_setup<D>(o, d, dims.vol, jr);
            n = jr.nsteps;
            const int64_t r = (int64_t) pose * dims.pix_per_pose + pix;
            if (axis)
                axis[r] = jr.axis;
            if (nsteps)
                nsteps[r] = jr.nsteps;
        }
        if (total) {
            unsigned long long c = (unsigned long long) n;
            for (int off = 32; off > 0; off >>= 1)
                c += __shfl_down(c, off, 64);
            if ((threadIdx.x & 63) == 0 && c)
                atomicAdd(total, c);
        }
    }

    // ---------------------------------------------------------------------------------
    // Siddon: one thread per ray; the traversal of TraverseAABB with the direct rule for the
    // crossing parameters (siddon_traverse, ehp_siddon.hpp).

    template <int D, bool ADJ, bool COUNT, class T>
    __device__ inline unsigned long long siddon_ray(const float* o, const float* d, const Dims& dims,
                                                    const T* in, T* out, int64_t ray)
    {
        T acc = T(0);
        unsigned long long crossings = 0;
        const T yv = (ADJ && !COUNT) ? in[ray] : T(0);
        const SidRec r = siddon_setup<D>(o, d, dims.vol);
        siddon_traverse<D>(r, dims.vol, [&](int c0, int c1, int c2, float w) {
            const int64_t idx = c0 + c1 * dims.strides[1] + (int64_t) c2 * dims.strides[2];
            if (COUNT)
                ++crossings;
            else if (ADJ) {
                if (w != 0.0f)
                    atomic_add(out + idx, yv * (T) w);
            } else
                acc += in[idx] * (T) w;
        });
        if (!ADJ && !COUNT)
            out[ray] = acc;
        return crossings;
    }

    template <int D, bool ADJ, bool COUNT, class T>
    __global__ void __launch_bounds__(256)
        siddon_kernel(const PoseGeom* __restrict__ geo, const Dims dims, const T* __restrict__ in,
                      T* __restrict__ out, unsigned long long* count)
    {
        const int pose = blockIdx.y;
        const int64_t pix = (int64_t) blockIdx.x * blockDim.x + threadIdx.x;
        const bool valid = pix < dims.pix_per_pose;
        if (!COUNT && !valid)
            return;
        unsigned long long c = 0;
        if (valid) {
            const int u = (int) (pix % dims.det[0]);
            const int v = (int) (pix / dims.det[0]);
            float o[3], d[3];
            make_ray_px<D>(geo[pose], dims, u, v, o, d);
            c = siddon_ray<D, ADJ, COUNT, T>(o, d, dims, in, out, (int64_t) pose * dims.pix_per_pose + pix);
        }
        if (COUNT) {
            for (int off = 32; off > 0; off >>= 1)
                c += __shfl_down(c, off, 64);
            if ((threadIdx.x & 63) == 0 && c)
                atomicAdd(count, c);
        }
    }

    // ---------------------------------------------------------------------------------
    // Voxel-driven blob / B-spline projector.

    // Adjoint: one thread per voxel gathers over all poses (no atomics). ST is the storage type
    // (bfloat16 storage: widened exactly, summed in FP32, each voxel rounded once).
    template <int D, class T, class ST>
    __global__ void __launch_bounds__(256)
        voxel_adj_kernel(const PoseGeom* __restrict__ geo, const Dims dims, const T* __restrict__ lut, int n_lut,
                         float support, const int diff, const int per_xcd, const ST* __restrict__ sino,
                         ST* __restrict__ vol)
    {
        // the FP32 gather's order (voxel_adj_fast_kernel): in 3D y slowest and a contiguous run of
        // blocks per XCD, so that an XCD's detector rows stay in its L2; each voxel's sum is the same
        const int64_t blk = (int64_t) (blockIdx.x % 8u) * per_xcd + (blockIdx.x / 8u);
        const int64_t vox = blk * blockDim.x + threadIdx.x;
        if (vox >= dims.nvox)
            return;
        const int x = (int) (vox % dims.vol[0]);
        int y, z;
        if (D == 3) {
            z = (int) ((vox / dims.vol[0]) % dims.vol[2]);
            y = (int) (vox / ((int64_t) dims.vol[0] * dims.vol[2]));
        } else {
            y = (int) ((vox / dims.vol[0]) % dims.vol[1]);
            z = 0;
        }
        T acc = T(0);
        for (int p = 0; p < dims.nposes; ++p) {
            Footprint fp;
            footprint<D>(geo[p], dims, x, y, z, support, fp);
            const ST* sp = sino + (int64_t) p * dims.pix_per_pose;
            visit<D, T>(fp, dims, [&](int64_t pix, float dist, float d0) {
                acc += tap_weight<D, T>(lut, n_lut, support, dist, d0, diff != 0) * up(sp[pix]);
            });
        }
        vol[x + (int64_t) y * dims.strides[1] + (int64_t) z * dims.strides[2]] = down<ST>(acc);
    }

    // The default FP32 gather: the footprints and taps of the tiled forward (footprint_fast, one
    // reciprocal of the scaling per voxel and pose instead of a division per tap,
    // tap_weight_fast's square root and table index), so forward and back projection weigh every
    // tap alike. The pair table of the n_lut <=
    // kVoxMaxLut entries (lerp_table_entry) is read from LDS. ST: float or bfloat16 storage.
    // YSLOW: the thread's index vox counts the voxels with y slowest (x fastest, then z), else in
    // the volume's order.
    template <int D, class ST, bool YSLOW = false>
    __device__ __forceinline__ void vox_adj_fast(const PoseGeom* __restrict__ geo, const Dims& dims,
                                                 const float* __restrict__ lut, int n_lut, float support,
                                                 const int diff, const ST* __restrict__ sino, ST* __restrict__ vol,
                                                 const int64_t vox)
    {
        __shared__ float s_lut[2 * kVoxLutSpan];
        for (int i = threadIdx.x; i < lerp_table_size(n_lut); i += blockDim.x)
            lerp_table_entry(lut, n_lut, i, s_lut);
        __syncthreads();
        if (vox >= dims.nvox)
            return;
        const int x = (int) (vox % dims.vol[0]);
        int y, z;
        if (YSLOW) {
            z = (int) ((vox / dims.vol[0]) % dims.vol[2]);
            y = (int) (vox / ((int64_t) dims.vol[0] * dims.vol[2]));
        } else {
            y = (int) ((vox / dims.vol[0]) % dims.vol[1]);
            z = (int) (vox / ((int64_t) dims.vol[0] * dims.vol[1]));
        }
        const float kidx = (float) (n_lut - 1) / support;
        // a table ending in 0 takes tap_weight_fast's TZ form
        auto gather = [&](auto tzc) {
            constexpr bool TZ = decltype(tzc)::value;
            float acc = 0.0f;
            for (int p = 0; p < dims.nposes; ++p) {
                Footprint fp;
                float rs;
                footprint_fast<D>(geo[p], dims, x, y, z, support, fp, rs);
                const ST* sp = sino + (int64_t) p * dims.pix_per_pose;
                for (int j = fp.lo[1]; j <= fp.hi[1]; ++j) {
                    const float d1 = D == 3 ? (fp.center[1] - (float) j) * rs : 0.0f;
                    const float d1s = d1 * d1;
                    float fi = (float) fp.lo[0]; // i as a float (exact), without a conversion per tap
                    for (int i = fp.lo[0]; i <= fp.hi[0]; ++i, fi += 1.0f) {
                        const float d0 = (fp.center[0] - fi) * rs;
                        // one FMA per tap: the sum of the taps in pose and pixel order, each product
                        // unrounded
                        acc = __builtin_fmaf(tap_weight_fast<D, TZ>(s_lut, n_lut, kidx, __builtin_fmaf(d0, d0, d1s),
                                                                    d0, diff != 0),
                                             up(sp[(int64_t) j * dims.det[0] + i]), acc);
                    }
                }
            }
            return acc;
        };
        const float acc = lut[n_lut - 1] == 0.0f ? gather(std::true_type{}) : gather(std::false_type{});
        vol[x + (int64_t) y * dims.strides[1] + (int64_t) z * dims.strides[2]] = down<ST>(acc);
    }

    // DIFF: the phase-contrast projectors' weights (a compile-time flag, so that the taps of the
    // others carry no select). In 3D the voxels are taken with y slowest and each XCD gets a
    // contiguous run of blocks (consecutive blocks land on different XCDs): a slab of y, whose
    // detector rows over all poses stay in that XCD's L2 when the trajectory turns about y.
    template <int D, class ST, bool DIFF>
    __global__ void __launch_bounds__(256)
        voxel_adj_fast_kernel(const PoseGeom* __restrict__ geo, const Dims dims, const float* __restrict__ lut,
                              int n_lut, float support, const int per_xcd, const ST* __restrict__ sino,
                              ST* __restrict__ vol)
    {
        const int64_t blk = (int64_t) (blockIdx.x % 8u) * per_xcd + (blockIdx.x / 8u);
        if (blk * blockDim.x >= dims.nvox)
            return; // the whole block
        vox_adj_fast<D, ST, D == 3>(geo, dims, lut, n_lut, support, DIFF ? 1 : 0, sino, vol,
                                    blk * blockDim.x + threadIdx.x);
    }

    // the gathers of many operators in one launch (block b: block b - job.block0 of job
    // block_job[b])
    template <int D>
    __global__ void __launch_bounds__(256)
        voxel_adj_fast_batch_kernel(const VoxAdjJob* __restrict__ jobs, const int* __restrict__ block_job,
                                    const float* __restrict__ sinos, float* __restrict__ vols)
    {
        const VoxAdjJob& J = jobs[block_job[blockIdx.x]];
        vox_adj_fast<D, float>(J.geo, J.dims, J.lut, J.n_lut, J.support, J.diff, sinos + J.sino_off,
                               vols + J.vol_off, ((int64_t) blockIdx.x - J.block0) * blockDim.x + threadIdx.x);
    }

    void voxel_adjoint_batch_jobs(const VoxBatchItem* items, int n, std::vector<VoxAdjJob>& jobs,
                                  std::vector<int>& blocks)
    {
        jobs.clear();
        blocks.clear();
        for (int j = 0; j < n; ++j) {
            const VoxBatchItem& it = items[j];
            VoxAdjJob J{};
            J.geo = it.geo;
            J.dims = it.dims;
            J.lut = it.lut;
            J.n_lut = it.n_lut;
            J.support = it.support;
            J.diff = it.diff;
            J.vol_off = it.vol_off;
            J.sino_off = it.sino_off;
            J.block0 = (int64_t) blocks.size();
            blocks.insert(blocks.end(), (size_t) ((it.dims.nvox + 255) / 256), (int) jobs.size());
            jobs.push_back(J);
        }
    }

    hipError_t voxel_adjoint_batch(int dim, const VoxAdjJob* jobs, const int* blocks, int64_t nblocks,
                                   const float* sinos, float* vols, hipStream_t s)
    {
        if (nblocks > 0x7fffffff)
            return hipErrorInvalidValue;
        if (nblocks == 0)
            return hipSuccess;
        if (dim == 2)
            hipLaunchKernelGGL((voxel_adj_fast_batch_kernel<2>), dim3((unsigned) nblocks), dim3(256), 0, s, jobs,
                               blocks, sinos, vols);
        else
            hipLaunchKernelGGL((voxel_adj_fast_batch_kernel<3>), dim3((unsigned) nblocks), dim3(256), 0, s, jobs,
                               blocks, sinos, vols);
        return hipGetLastError();
    }

    template <class ST>
    hipError_t voxel_adjoint_fast(const PoseGeom* geo, const Dims& dims, const float* lut, int n_lut, float support,
                                  int diff, const ST* sino, ST* vol, hipStream_t s)
    {
        const int64_t per_xcd = ((dims.nvox + 255) / 256 + 7) / 8;
        if (per_xcd * 8 > (int64_t) 0x7fffffff)
            return hipErrorInvalidValue;
        const dim3 grid((unsigned) (per_xcd * 8));
        auto launch = [&](auto kernel) {
            hipLaunchKernelGGL(kernel, grid, dim3(256), 0, s, geo, dims, lut, n_lut, support, (int) per_xcd, sino,
                               vol);
        };
        if (dims.dim == 2)
            diff ? launch(voxel_adj_fast_kernel<2, ST, true>) : launch(voxel_adj_fast_kernel<2, ST, false>);
        else
            diff ? launch(voxel_adj_fast_kernel<3, ST, true>) : launch(voxel_adj_fast_kernel<3, ST, false>);
        return hipGetLastError();
    }

    template hipError_t voxel_adjoint_fast<float>(const PoseGeom*, const Dims&, const float*, int, float, int,
                                                  const float*, float*, hipStream_t);
    template hipError_t voxel_adjoint_fast<bf16>(const PoseGeom*, const Dims&, const float*, int, float, int,
                                                 const bf16*, bf16*, hipStream_t);

    // Forward: one thread per (voxel, pose) scatters into the detector.
    template <int D, class T>
    __global__ void __launch_bounds__(256)
        voxel_fwd_kernel(const PoseGeom* __restrict__ geo, const Dims dims, const T* __restrict__ lut, int n_lut,
                         float support, const int diff, const T* __restrict__ vol, T* __restrict__ sino)
    {
        const int p = blockIdx.y;
        const int64_t vox = (int64_t) blockIdx.x * blockDim.x + threadIdx.x;
        if (vox >= dims.nvox)
            return;
        const int x = (int) (vox % dims.vol[0]);
        const int y = (int) ((vox / dims.vol[0]) % dims.vol[1]);
        const int z = (int) (vox / ((int64_t) dims.vol[0] * dims.vol[1]));
        const T xv = vol[vox];
        if (xv == T(0))
            return;
        Footprint fp;
        footprint<D>(geo[p], dims, x, y, z, support, fp);
        T* sp = sino + (int64_t) p * dims.pix_per_pose;
        visit<D, T>(fp, dims, [&](int64_t pix, float dist, float d0) {
            const T w = tap_weight<D, T>(lut, n_lut, support, dist, d0, diff != 0);
            if (w != T(0))
                atomic_add(sp + pix, xv * w);
        });
    }

    // Footprint taps of the voxel projector (the roofline's unit): the clamped footprint
    // rectangle of every voxel and pose, counted per voxel over all poses, reduced per block,
    // one atomic per block.
    template <int D>
    __global__ void __launch_bounds__(256)
        voxel_count_kernel(const PoseGeom* __restrict__ geo, const Dims dims, float support,
                           unsigned long long* __restrict__ count)
    {
        const int64_t vox = (int64_t) blockIdx.x * blockDim.x + threadIdx.x;
        unsigned long long taps = 0;
        if (vox < dims.nvox) {
            const int x = (int) (vox % dims.vol[0]);
            const int y = (int) ((vox / dims.vol[0]) % dims.vol[1]);
            const int z = (int) (vox / ((int64_t) dims.vol[0] * dims.vol[1]));
            for (int p = 0; p < dims.nposes; ++p) {
                Footprint fp;
                footprint<D>(geo[p], dims, x, y, z, support, fp);
                const int nu = fp.hi[0] - fp.lo[0] + 1;
                const int nv = fp.hi[1] - fp.lo[1] + 1;
                taps += (nu > 0 && nv > 0) ? (unsigned long long) nu * nv : 0;
            }
        }
        for (int off = 32; off > 0; off >>= 1)
            taps += __shfl_down(taps, off, 64);
        __shared__ unsigned long long part[4];
        if ((threadIdx.x & 63) == 0)
            part[threadIdx.x >> 6] = taps;
        __syncthreads();
        if (threadIdx.x == 0) {
            const unsigned long long t = part[0] + part[1] + part[2] + part[3];
            if (t)
                atomicAdd(count, t);
        }
    }

    // ---------------------------------------------------------------------------------
    // launchers

    static dim3 ray_grid(const Dims& dims)
    {
        return dim3((unsigned) ((dims.pix_per_pose + 255) / 256), (unsigned) dims.nposes);
    }

    template <class T>
    hipError_t joseph_forward(const PoseGeom* geo, const Dims& dims, const T* vol, T* sino, hipStream_t s)
    {
        if (dims.dim == 2)
            hipLaunchKernelGGL((joseph_fwd_kernel<2, T>), ray_grid(dims), dim3(256), 0, s, geo, dims, vol, sino);
        else
            hipLaunchKernelGGL((joseph_fwd_kernel<3, T>), ray_grid(dims), dim3(256), 0, s, geo, dims, vol, sino);
        return hipGetLastError();
    }

    template <class T>
    hipError_t joseph_adjoint(const PoseGeom* geo, const Dims& dims, const T* sino, T* vol, hipStream_t s)
    {
        hipError_t e = launch_fill<T>(vol, dims.nvox, T(0), s);
        if (e != hipSuccess)
            return e;
        if (dims.dim == 2)
            hipLaunchKernelGGL((joseph_adj_kernel<2, T>), ray_grid(dims), dim3(256), 0, s, geo, dims, sino, vol);
        else
            hipLaunchKernelGGL((joseph_adj_kernel<3, T>), ray_grid(dims), dim3(256), 0, s, geo, dims, sino, vol);
        return hipGetLastError();
    }

    hipError_t joseph_info(const PoseGeom* geo, const Dims& dims, int32_t* axis, int64_t* nsteps,
                           unsigned long long* total, hipStream_t s)
    {
        if (dims.dim == 2)
            hipLaunchKernelGGL((joseph_info_kernel<2>), ray_grid(dims), dim3(256), 0, s, geo, dims, axis, nsteps, total);
        else
            hipLaunchKernelGGL((joseph_info_kernel<3>), ray_grid(dims), dim3(256), 0, s, geo, dims, axis, nsteps, total);
        return hipGetLastError();
    }

    template <class T>
    hipError_t siddon_forward(const PoseGeom* geo, const Dims& dims, const T* vol, T* sino, hipStream_t s)
    {
        if (dims.dim == 2)
            hipLaunchKernelGGL((siddon_kernel<2, false, false, T>), ray_grid(dims), dim3(256), 0, s, geo, dims, vol, sino,
                               nullptr);
        else
            hipLaunchKernelGGL((siddon_kernel<3, false, false, T>), ray_grid(dims), dim3(256), 0, s, geo, dims, vol, sino,
                               nullptr);
        return hipGetLastError();
    }

    template <class T>
    hipError_t siddon_adjoint(const PoseGeom* geo, const Dims& dims, const T* sino, T* vol, hipStream_t s)
    {
        hipError_t e = launch_fill<T>(vol, dims.nvox, T(0), s);
        if (e != hipSuccess)
            return e;
        if (dims.dim == 2)
            hipLaunchKernelGGL((siddon_kernel<2, true, false, T>), ray_grid(dims), dim3(256), 0, s, geo, dims, sino, vol,
                               nullptr);
        else
            hipLaunchKernelGGL((siddon_kernel<3, true, false, T>), ray_grid(dims), dim3(256), 0, s, geo, dims, sino, vol,
                               nullptr);
        return hipGetLastError();
    }

    hipError_t siddon_count(const PoseGeom* geo, const Dims& dims, unsigned long long* total, hipStream_t s)
    {
        if (dims.dim == 2)
            hipLaunchKernelGGL((siddon_kernel<2, false, true, float>), ray_grid(dims), dim3(256), 0, s, geo, dims,
                               (const float*) nullptr, (float*) nullptr, total);
        else
            hipLaunchKernelGGL((siddon_kernel<3, false, true, float>), ray_grid(dims), dim3(256), 0, s, geo, dims,
                               (const float*) nullptr, (float*) nullptr, total);
        return hipGetLastError();
    }

    static dim3 vox_pose_grid(const Dims& dims)
    {
        return dim3((unsigned) ((dims.nvox + 255) / 256), (unsigned) dims.nposes);
    }

    template <class T>
    hipError_t voxel_forward(const PoseGeom* geo, const Dims& dims, const T* lut, int n_lut, float support,
                             int diff, const T* vol, T* sino, hipStream_t s)
    {
        hipError_t e = launch_fill<T>(sino, dims.nrays, T(0), s);
        if (e != hipSuccess)
            return e;
        if (dims.dim == 2)
            hipLaunchKernelGGL((voxel_fwd_kernel<2, T>), vox_pose_grid(dims), dim3(256), 0, s, geo, dims, lut, n_lut,
                               support, diff, vol, sino);
        else
            hipLaunchKernelGGL((voxel_fwd_kernel<3, T>), vox_pose_grid(dims), dim3(256), 0, s, geo, dims, lut, n_lut,
                               support, diff, vol, sino);
        return hipGetLastError();
    }

    template <class T, class ST>
    hipError_t voxel_adjoint(const PoseGeom* geo, const Dims& dims, const T* lut, int n_lut, float support,
                             int diff, const ST* sino, ST* vol, hipStream_t s)
    {
        const int64_t per_xcd = ((dims.nvox + 255) / 256 + 7) / 8;
        if (per_xcd * 8 > (int64_t) 0x7fffffff)
            return hipErrorInvalidValue;
        const dim3 grid((unsigned) (per_xcd * 8));
        if (dims.dim == 2)
            hipLaunchKernelGGL((voxel_adj_kernel<2, T, ST>), grid, dim3(256), 0, s, geo, dims, lut, n_lut, support,
                               diff, (int) per_xcd, sino, vol);
        else
            hipLaunchKernelGGL((voxel_adj_kernel<3, T, ST>), grid, dim3(256), 0, s, geo, dims, lut, n_lut, support,
                               diff, (int) per_xcd, sino, vol);
        return hipGetLastError();
    }

    hipError_t voxel_count(const PoseGeom* geo, const Dims& dims, float support, unsigned long long* total,
                           hipStream_t s)
    {
        const dim3 grid((unsigned) ((dims.nvox + 255) / 256));
        if (dims.dim == 2)
            hipLaunchKernelGGL((voxel_count_kernel<2>), grid, dim3(256), 0, s, geo, dims, support, total);
        else
            hipLaunchKernelGGL((voxel_count_kernel<3>), grid, dim3(256), 0, s, geo, dims, support, total);
        return hipGetLastError();
    }

#define EHP_INSTANTIATE(T)                                                                                         \
    template hipError_t launch_fill<T>(T*, int64_t, T, hipStream_t);                                             \
    template hipError_t joseph_forward<T>(const PoseGeom*, const Dims&, const T*, T*, hipStream_t);              \
    template hipError_t joseph_adjoint<T>(const PoseGeom*, const Dims&, const T*, T*, hipStream_t);              \
    template hipError_t siddon_forward<T>(const PoseGeom*, const Dims&, const T*, T*, hipStream_t);              \
    template hipError_t siddon_adjoint<T>(const PoseGeom*, const Dims&, const T*, T*, hipStream_t);              \
    template hipError_t voxel_forward<T>(const PoseGeom*, const Dims&, const T*, int, float, int, const T*, T*,   \
                                         hipStream_t);                                                           \
    template hipError_t voxel_adjoint<T, T>(const PoseGeom*, const Dims&, const T*, int, float, int, const T*,    \
                                            T*, hipStream_t);

    EHP_INSTANTIATE(float)
    EHP_INSTANTIATE(double)
    template hipError_t launch_fill<int>(int*, int64_t, int, hipStream_t);
    template hipError_t launch_fill<bf16>(bf16*, int64_t, bf16, hipStream_t);
    template hipError_t launch_fill<unsigned int>(unsigned int*, int64_t, unsigned int, hipStream_t);
#undef EHP_INSTANTIATE
} // namespace ehp
