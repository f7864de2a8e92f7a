// ehp_capi.hip -- the extern "C" boundary of libelsahip.so (declared in include/elsa_hip.h).
//
// Owns the device state elsa's GPU projector classes build in their constructors (per-pose
// geometry, the basis-function table) and dispatches apply/applyAdjoint to the kernels.
// Errors come back as status codes with a per-thread message instead of elsa's exceptions
// (elsa/projectors_cuda/JosephsMethodCUDA.cpp:17-69, SiddonsMethodCUDA.cpp:17-30).
#include "elsa_hip.h"
#include "ehp_kernels.hpp"
#include "ehp_solvers.hpp"

#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "ehp_op.hpp"

namespace
{
    thread_local std::string g_last_error;

    int fail(int code, const char* fmt, ...)
    {
        char buf[512];
        va_list ap;
        va_start(ap, fmt);
        std::vsnprintf(buf, sizeof(buf), fmt, ap);
        va_end(ap);
        g_last_error = buf;
        return code;
    }

    int hip_fail(hipError_t e, const char* what)
    {
        return fail(EHP_ERR_HIP, "%s: %s", what, hipGetErrorString(e));
    }

    bool valid_dtype(int dtype) { return dtype == EHP_F32 || dtype == EHP_F64 || dtype == EHP_BF16; }

    size_t dtype_size(int dtype)
    {
        return dtype == EHP_F64 ? sizeof(double) : (dtype == EHP_BF16 ? sizeof(ehp::bf16) : sizeof(float));
    }

    ehp::Dims make_dims(const ehp_geometry* g)
    {
        ehp::Dims d{};
        d.dim = g->dim;
        for (int i = 0; i < 3; ++i)
            d.vol[i] = i < g->dim ? g->vol[i] : 1;
        d.det[0] = g->det[0];
        d.det[1] = g->dim == 3 ? g->det[1] : 1;
        d.nposes = (int) g->nposes;
        d.strides[0] = 1;
        for (int i = 1; i < 3; ++i)
            d.strides[i] = d.strides[i - 1] * d.vol[i - 1];
        d.nvox = (int64_t) d.vol[0] * d.vol[1] * d.vol[2];
        d.pix_per_pose = (int64_t) d.det[0] * d.det[1];
        d.nrays = d.pix_per_pose * d.nposes;
        return d;
    }

    int new_op(const ehp_geometry* g, int kind, int dtype, ehp_op** out)
    {
        if (!out)
            return fail(EHP_ERR_ARGUMENT, "output handle is NULL");
        *out = nullptr;
        if (!g)
            return fail(EHP_ERR_ARGUMENT, "geometry is NULL");
        if (!valid_dtype(dtype))
            return fail(EHP_ERR_ARGUMENT, "unsupported dtype %d", dtype);
        ehp_op* op = new ehp_op();
        op->kind = kind;
        op->dtype = dtype;
        op->dims = make_dims(g);
        hipError_t e = hipGetDevice(&op->device);
        if (e != hipSuccess) {
            delete op;
            return hip_fail(e, "hipGetDevice");
        }
        e = ehp::order_init(op->order);
        if (e != hipSuccess) {
            delete op;
            return hip_fail(e, "event of the operator");
        }
        const size_t bytes = g->host.size() * sizeof(ehp::PoseGeom);
        e = hipMalloc((void**) &op->d_geo, bytes);
        if (e != hipSuccess) {
            delete op;
            return fail(EHP_ERR_ALLOC, "hipMalloc of %zu bytes of geometry failed: %s", bytes, hipGetErrorString(e));
        }
        e = hipMemcpy(op->d_geo, g->host.data(), bytes, hipMemcpyHostToDevice);
        if (e != hipSuccess) {
            (void) hipFree(op->d_geo);
            delete op;
            return hip_fail(e, "geometry upload");
        }
        op->host_geo = g->host;
        if (!g->pix.empty()) {
            const size_t pbytes = g->pix.size() * sizeof(float);
            e = hipMalloc((void**) &op->d_pix, pbytes);
            if (e == hipSuccess)
                e = hipMemcpy(op->d_pix, g->pix.data(), pbytes, hipMemcpyHostToDevice);
            if (e != hipSuccess) {
                (void) hipFree(op->d_geo);
                (void) hipFree(op->d_pix);
                delete op;
                return fail(EHP_ERR_ALLOC, "detector coordinates: %s", hipGetErrorString(e));
            }
            op->dims.pix = op->d_pix;
            op->host_pix = g->pix;
            op->curved_tiles = ehp::curved_candidates(op->dims, g->pix.data());
        }
        *out = op;
        return EHP_OK;
    }

    // host copy of a curved detector's planar centres (nullptr for a planar detector)
    const float* curved_pix(const ehp_op* op) { return op->host_pix.empty() ? nullptr : op->host_pix.data(); }

    // Sample count of one pass, computed once on the device.
    int count_samples(ehp_op* op)
    {
        unsigned long long* d_total = nullptr;
        hipError_t e = hipMalloc((void**) &d_total, sizeof(unsigned long long));
        if (e != hipSuccess)
            return fail(EHP_ERR_ALLOC, "hipMalloc of the sample counter failed: %s", hipGetErrorString(e));
        e = hipMemset(d_total, 0, sizeof(unsigned long long));
        if (e == hipSuccess) {
            if (op->kind == KIND_JOSEPH)
                e = ehp::joseph_info(op->d_geo, op->dims, nullptr, nullptr, d_total, nullptr);
            else if (op->kind == KIND_SIDDON)
                e = ehp::siddon_count(op->d_geo, op->dims, d_total, nullptr);
            else
                e = ehp::voxel_count(op->d_geo, op->dims, op->support, d_total, nullptr);
        }
        unsigned long long h = 0;
        if (e == hipSuccess)
            e = hipMemcpy(&h, d_total, sizeof(h), hipMemcpyDeviceToHost);
        (void) hipFree(d_total);
        if (e != hipSuccess)
            return hip_fail(e, "sample count");
        op->samples = h;
        return EHP_OK;
    }
} // namespace

extern "C" {

const char* ehp_last_error(void) { return g_last_error.c_str(); }

const char* ehp_version(void) { return "elsa_amd 0.1 (gfx950)"; }

int ehp_geometry_create(int dim, int64_t nposes, const int64_t* vol_dims, const int64_t* det_dims,
                        const float* P, const float* Pinv, const float* C, const float* ext,
                        const float* sdd, ehp_geometry** out)
{
    if (!out)
        return fail(EHP_ERR_ARGUMENT, "output handle is NULL");
    *out = nullptr;
    if (dim != 2 && dim != 3)
        return fail(EHP_ERR_ARGUMENT, "only 2D and 3D geometries are supported (got %d)", dim);
    if (nposes <= 0)
        return fail(EHP_ERR_ARGUMENT, "geometry list was empty");
    if (nposes > 65535)
        return fail(EHP_ERR_UNSUPPORTED, "at most 65535 poses per operator (got %lld)", (long long) nposes);
    if (!vol_dims || !det_dims || !P || !Pinv || !C || !ext || !sdd)
        return fail(EHP_ERR_ARGUMENT, "NULL geometry array");
    ehp_geometry* g = new ehp_geometry();
    g->dim = dim;
    g->nposes = nposes;
    int64_t nvox = 1, npix = 1;
    for (int i = 0; i < 3; ++i)
        g->vol[i] = 1;
    g->det[0] = g->det[1] = 1;
    for (int i = 0; i < dim; ++i) {
        if (vol_dims[i] <= 0 || vol_dims[i] > (1 << 20)) {
            delete g;
            return fail(EHP_ERR_ARGUMENT, "invalid volume size %lld along axis %d", (long long) vol_dims[i], i);
        }
        g->vol[i] = (int) vol_dims[i];
        nvox *= vol_dims[i];
    }
    for (int i = 0; i < dim - 1; ++i) {
        if (det_dims[i] <= 0 || det_dims[i] > (1 << 20)) {
            delete g;
            return fail(EHP_ERR_ARGUMENT, "invalid detector size %lld along axis %d", (long long) det_dims[i], i);
        }
        g->det[i] = (int) det_dims[i];
        npix *= det_dims[i];
    }
    if (nvox >= (int64_t(1) << 40) || npix * nposes >= (int64_t(1) << 40)) {
        delete g;
        return fail(EHP_ERR_UNSUPPORTED, "problem too large");
    }
    g->host.resize((size_t) nposes);
    const int nm = dim * (dim + 1);
    for (int64_t p = 0; p < nposes; ++p) {
        ehp::PoseGeom& pg = g->host[(size_t) p];
        std::memset(&pg, 0, sizeof(pg));
        for (int i = 0; i < nm; ++i) {
            pg.P[i] = P[p * nm + i];
            pg.Pinv[i] = Pinv[p * nm + i];
            pg.ext[i] = ext[p * nm + i];
        }
        for (int i = 0; i < dim; ++i)
            pg.C[i] = C[p * dim + i];
        pg.sdd = sdd[p];
    }
    *out = g;
    return EHP_OK;
}

void ehp_geometry_destroy(ehp_geometry* g) { delete g; }

int ehp_geometry_set_pixel_coords(ehp_geometry* g, const float* coords)
{
    if (!g)
        return fail(EHP_ERR_ARGUMENT, "geometry is NULL");
    if (!coords) {
        g->pix.clear();
        return EHP_OK;
    }
    const size_t n = (size_t) g->det[0] * (g->dim == 3 ? g->det[1] : 1) * (size_t) (g->dim - 1);
    for (size_t i = 0; i < n; ++i)
        if (!std::isfinite(coords[i]))
            return fail(EHP_ERR_ARGUMENT, "detector coordinate %zu is not finite", i);
    g->pix.assign(coords, coords + n);
    return EHP_OK;
}

namespace
{
    // Tile shape of the LDS-privatised back projections, shrunk until it fits the device's LDS
    // per workgroup. Joseph 3D: 8 x 17 x 17 (9.15 ms against 9.3 ms for 8 x 15 x 15 and 9.7 ms for
    // 8 x 19 x 19 at 256^3 / 360 poses; odd widths, as an even slot row stride costs LDS bank
    // conflicts: 12.6 ms for 8 x 18 x 18).
    size_t tile_lds_bytes(int kind, const ehp::Dims& dims, const ehp::TileCfg& tc, int dtype = EHP_F32)
    {
        if (kind == KIND_VOXEL)
            return ehp::voxel_tile_lds_bytes(tc, dtype == EHP_F64 ? sizeof(double) : sizeof(float));
        return kind == KIND_SIDDON ? ehp::siddon_tile_lds_bytes(dims, tc) : ehp::joseph_tile_lds_bytes(dims, tc);
    }

    // cus: compute units of the device. A 3D back projection launch is one workgroup per tile; a
    // small volume gets thinner tiles (fewer planes each) until a launch has about 8 workgroups
    // per CU, which keeps the device busy (128^3: 1024 tiles of 8 planes, 2048 of 4).
    ehp::TileCfg default_tile(int kind, const ehp::Dims& dims, int max_lds, int cus)
    {
        ehp::TileCfg tc;
        if (kind == KIND_SIDDON)
            tc = dims.dim == 3 ? ehp::TileCfg{8, 16, 16} : ehp::TileCfg{4, 16, 1};
        else
            tc = dims.dim == 3 ? ehp::TileCfg{8, 17, 17} : ehp::TileCfg{4, 15, 1};
        while (tc.LA > 1 && (int) tile_lds_bytes(kind, dims, tc) > max_lds)
            tc.LA /= 2;
        while (dims.dim == 3 && tc.LA > 2) {
            int ntA, ntB, ntC;
            ehp::tile_counts(dims, tc, 0, ntA, ntB, ntC);
            if ((int64_t) ntA * ntB * ntC >= (int64_t) 8 * cus)
                break;
            tc.LA /= 2;
        }
        return tc;
    }

    int device_cus(int device)
    {
        int cus = 0;
        if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess || cus <= 0)
            cus = 256;
        return cus;
    }

    // Poses per step of the tiled back projections: 32, the most a step takes. In 2D few rays
    // meet a tile per pose, so many poses keep the workgroup's lanes busy (0.105 -> 0.078 ms at
    // 256^2 / 180 poses); in 3D fewer steps mean fewer barriers and candidate setups (256^3 /
    // 360 poses: 7.65 ms with 8 poses per step, 7.35 ms with 32), and the batch arrays add
    // under 1 KB of LDS to the tile's 23 KB.
    int default_pose_batch(const ehp::Dims&)
    {
        return ehp::kMaxPoseBatch;
    }

    // Fastest axis of the plane-major view for driving axis a: the transverse axis along
    // which the rays of neighbouring pixels in a detector row spread most, summed over the
    // poses that have rays of axis a (the first column of Pinv is that spread per pixel).
    int fan_axis(const ehp_op* op, const std::vector<int>& pose_axes, int a)
    {
        const ehp::Dims& dm = op->dims;
        if (dm.dim == 2)
            return 1 - a;
        double score[3] = {0.0, 0.0, 0.0};
        for (int p = 0; p < dm.nposes; ++p) {
            if (!(pose_axes[p] & (1 << a)))
                continue;
            for (int i = 0; i < 3; ++i)
                score[i] += std::fabs((double) op->host_geo[p].Pinv[i]);
        }
        int best = -1;
        for (int i = 0; i < 3; ++i)
            if (i != a && (best < 0 || score[i] > score[best]))
                best = i;
        return best;
    }

    // Zero-padded plane-major views of the volume for the plane-walking forwards, one per
    // driving axis that occurs (pose_axes: the axes per pose), in the operator's storage type:
    // 32-bit byte offsets inside a view, so below 4 GiB each; the padding is zeroed once here and
    // never written again. Sets op->planes when every view fits.
    int setup_views(ehp_op* op, const std::vector<int>& h)
    {
        const ehp::Dims& dm = op->dims;
        hipError_t e = hipSuccess;
        // elements per view below 2^32 bytes: 2^30 FP32 / bfloat16 (the latter's byte offsets
        // would allow more; one limit keeps the FP32 and bfloat16 views alike), 2^29 FP64
        const int64_t max_elems = (int64_t) 1 << (op->dtype == EHP_F64 ? 29 : 30);
        // 3D Joseph forwards (FP32, bfloat16) read row-pair views (one load per sample) when those fit
        bool pairs = op->kind == KIND_JOSEPH && (op->dtype == EHP_F32 || op->dtype == EHP_BF16) && dm.dim == 3;
        for (int a = 0; pairs && a < dm.dim; ++a)
            if ((op->axes_present & (1 << a)) && ehp::pad_layout(dm, a, fan_axis(op, h, a), true).size >= max_elems)
                pairs = false;
        bool planes = true;
        for (int a = 0; a < dm.dim; ++a) {
            op->layouts[a] = ehp::pad_layout(dm, a, fan_axis(op, h, a), pairs);
            if ((op->axes_present & (1 << a)) && op->layouts[a].size >= max_elems)
                planes = false;
        }
        const bool b16 = op->dtype == EHP_BF16;
        const size_t esize = b16 ? sizeof(ehp::bf16) : (op->dtype == EHP_F64 ? sizeof(double) : sizeof(float));
        for (int a = 0; planes && a < dm.dim; ++a) {
            if (!(op->axes_present & (1 << a)))
                continue;
            const size_t vbytes = (size_t) op->layouts[a].size * esize + (b16 ? ehp::kViewSlack : 0);
            e = hipMalloc((void**) &op->d_view[a], vbytes);
            if (e != hipSuccess)
                return fail(EHP_ERR_ALLOC, "hipMalloc of a %zu-byte volume view failed: %s", vbytes,
                            hipGetErrorString(e));
            if (b16)
                e = ehp::launch_fill<ehp::bf16>((ehp::bf16*) op->d_view[a], op->layouts[a].size + ehp::kViewSlack / 2,
                                                ehp::bf16{0}, nullptr);
            else if (op->dtype == EHP_F64)
                e = ehp::launch_fill<double>((double*) op->d_view[a], op->layouts[a].size, 0.0, nullptr);
            else
                e = ehp::launch_fill<float>((float*) op->d_view[a], op->layouts[a].size, 0.0f, nullptr);
            if (e != hipSuccess)
                return hip_fail(e, "volume view");
        }
        op->planes = planes;
        return EHP_OK;
    }

    // Joseph (FP32, bfloat16, FP64): per-ray table, driving axes per pose, plane-major views for
    // the forward in the operator's storage type and, when the fixed-point bound exists, the
    // LDS-tiled back projection.
    int setup_fast(ehp_op* op)
    {
        const ehp::Dims& dm = op->dims;
        const int n = dm.nposes;
        const size_t tbytes = (size_t) dm.nrays * sizeof(ehp::RayRec);
        hipError_t e = hipMalloc((void**) &op->d_table, tbytes);
        if (e != hipSuccess)
            return fail(EHP_ERR_ALLOC, "hipMalloc of the %zu-byte ray table failed: %s", tbytes, hipGetErrorString(e));
        e = hipMalloc((void**) &op->d_pose_axes, n * sizeof(int));
        if (e == hipSuccess)
            e = hipMalloc((void**) &op->d_maxbits, ehp::kAbsmaxBlocks * sizeof(unsigned long long));
        if (e == hipSuccess)
            e = hipMalloc((void**) &op->d_fixexp, 2 * sizeof(int));
        if (e != hipSuccess)
            return fail(EHP_ERR_ALLOC, "hipMalloc of the Joseph state failed: %s", hipGetErrorString(e));
        e = ehp::joseph_ray_table(op->d_geo, dm, op->d_table, nullptr);
        if (e == hipSuccess)
            e = ehp::joseph_pose_axes(op->d_table, dm, op->d_pose_axes, nullptr);
        std::vector<int> h(n);
        if (e == hipSuccess)
            e = hipMemcpy(h.data(), op->d_pose_axes, n * sizeof(int), hipMemcpyDeviceToHost);
        if (e != hipSuccess)
            return hip_fail(e, "ray table");
        op->axes_present = 0;
        for (int v : h)
            op->axes_present |= v;

        int rc = setup_views(op, h);
        if (rc != EHP_OK)
            return rc;
        const bool b16 = op->dtype == EHP_BF16;
        const bool planes = op->planes;
        if (b16 && planes) {
            const size_t abytes = (size_t) dm.nvox * sizeof(float);
            e = hipMalloc((void**) &op->d_acc32, abytes);
            if (e != hipSuccess)
                return fail(EHP_ERR_ALLOC, "hipMalloc of the %zu-byte FP32 back projection failed: %s", abytes,
                            hipGetErrorString(e));
        }
        op->planes = planes;
        // the views' padding is cleared on the null stream; finish before any caller stream runs
        e = hipDeviceSynchronize();
        if (e != hipSuccess)
            return hip_fail(e, "volume view");

        // back projection: the bound keeps the 64-bit fixed-point sums in range; without it,
        // and for curved detectors whose planar centres lack the structure the tiles' candidate
        // search needs, the one-thread-per-ray kernel runs
        const double bound = ehp::fixed_point_slot_bound(dm, op->host_geo.data(), curved_pix(op));
        if (!(bound > 0.0) || bound > 1.0e18 || (op->d_pix && !op->curved_tiles) || !ehp::tiled_detector_fits(dm)) {
            (void) hipFree(op->d_table);
            op->d_table = nullptr;
            return EHP_OK;
        }
        op->slot_bound = bound;
        int max_lds = 65536;
        if (hipDeviceGetAttribute(&max_lds, hipDeviceAttributeMaxSharedMemoryPerBlock, op->device) != hipSuccess)
            max_lds = 65536;
        op->tile = default_tile(KIND_JOSEPH, dm, max_lds, device_cus(op->device));
        op->batch = default_pose_batch(dm);
        op->tiled = true;
        return EHP_OK;
    }

    // Siddon (FP32, FP64): when the fixed-point bound exists, the per-ray table, the driving axes
    // per pose, the LDS-tiled back projection and the plane-major views of the forward.
    int setup_siddon_tiled(ehp_op* op)
    {
        const ehp::Dims& dm = op->dims;
        const double bound = ehp::fixed_point_slot_bound(dm, op->host_geo.data(), curved_pix(op));
        if (!(bound > 0.0) || bound > 1.0e18 || (op->d_pix && !op->curved_tiles) || !ehp::tiled_detector_fits(dm))
            return EHP_OK; // keep the one-thread-per-ray kernel
        for (int i = 0; i < dm.dim; ++i)
            if (dm.vol[i] > 32767)
                return EHP_OK; // the table stores voxel indices as 16-bit values
        const int n = dm.nposes;
        const size_t tbytes = (size_t) dm.nrays * sizeof(ehp::SidRec);
        hipError_t e = hipMalloc((void**) &op->d_sidtable, tbytes);
        if (e != hipSuccess)
            return fail(EHP_ERR_ALLOC, "hipMalloc of the %zu-byte ray table failed: %s", tbytes, hipGetErrorString(e));
        e = hipMalloc((void**) &op->d_pose_axes, n * sizeof(int));
        if (e == hipSuccess)
            e = hipMalloc((void**) &op->d_maxbits, ehp::kAbsmaxBlocks * sizeof(unsigned long long));
        if (e == hipSuccess)
            e = hipMalloc((void**) &op->d_fixexp, 2 * sizeof(int));
        if (e != hipSuccess)
            return fail(EHP_ERR_ALLOC, "hipMalloc of the Siddon state failed: %s", hipGetErrorString(e));
        e = ehp::siddon_ray_table(op->d_geo, dm, op->d_sidtable, op->d_pose_axes, nullptr);
        std::vector<int> h(n);
        if (e == hipSuccess)
            e = hipMemcpy(h.data(), op->d_pose_axes, n * sizeof(int), hipMemcpyDeviceToHost);
        if (e != hipSuccess)
            return hip_fail(e, "Siddon ray table");
        op->axes_present = 0;
        for (int v : h)
            op->axes_present |= v;
        // the forward walks plane-major views (one element per voxel) plane by plane
        const int rc = setup_views(op, h);
        if (rc != EHP_OK)
            return rc;
        e = hipDeviceSynchronize(); // the views' padding is cleared on the null stream
        if (e != hipSuccess)
            return hip_fail(e, "volume view");
        op->slot_bound = bound;
        int max_lds = 65536;
        if (hipDeviceGetAttribute(&max_lds, hipDeviceAttributeMaxSharedMemoryPerBlock, op->device) != hipSuccess)
            max_lds = 65536;
        op->tile = default_tile(KIND_SIDDON, dm, max_lds, device_cus(op->device));
        op->batch = default_pose_batch(dm);
        op->tiled = true;
        return EHP_OK;
    }
    // Voxel projectors (FP32, bfloat16, FP64): the detector-tiled forward, when every source lies
    // outside the volume (the footprint scaling is then bounded).
    int setup_voxel_tiled(ehp_op* op, const void* host_lut)
    {
        const ehp::Dims& dm = op->dims;
        const int n = dm.nposes;
        if (op->n_lut > ehp::voxel_max_lut())
            return EHP_OK;
        std::vector<int> axis(n);
        std::vector<float> margin(n);
        double smax_all = 0.0;
        for (int p = 0; p < n; ++p) {
            const double smax = ehp::voxel_max_scaling(op->host_geo[p], dm);
            if (!(smax > 0.0) || smax > 1.0e6)
                return EHP_OK; // keep the one-thread-per-voxel kernel
            smax_all = smax > smax_all ? smax : smax_all;
            margin[p] = (float) (op->support * smax + 1.0);
            axis[p] = dm.dim == 3 ? ehp::pose_axis<3>(op->host_geo[p], dm) : ehp::pose_axis<2>(op->host_geo[p], dm);
        }
        // the table (FP64 for FP64 operators, else FP32)
        double lutmax = 0.0;
        for (int i = 0; i < op->n_lut; ++i) {
            const double v = op->dtype == EHP_F64 ? ((const double*) host_lut)[i] : ((const float*) host_lut)[i];
            lutmax = std::fabs(v) > lutmax ? std::fabs(v) : lutmax;
        }
        // the 3D phase-contrast weight is the table times d_u, |d_u| <= |d| <= support where the
        // table is not zero (with a little room for the rounding of |d|)
        if (op->differential && dm.dim == 3)
            lutmax = lutmax * (op->support > 1.0f ? op->support : 1.0f) * 1.001;
        if (!(lutmax < 1.0e30))
            return EHP_OK;
        double ls = 1.0;
        while (ls < lutmax)
            ls *= 2.0;
        op->slot_bound = ehp::voxel_slot_bound(dm, op->support, smax_all, ls);
        hipError_t e = hipMalloc((void**) &op->d_pose_axes, n * sizeof(int));
        if (e == hipSuccess)
            e = hipMalloc((void**) &op->d_pose_margin, n * sizeof(float));
        if (e == hipSuccess)
            e = hipMalloc((void**) &op->d_maxbits, ehp::kAbsmaxBlocks * sizeof(unsigned long long));
        if (e == hipSuccess)
            e = hipMalloc((void**) &op->d_fixexp, 2 * sizeof(int));
        if (e != hipSuccess)
            return fail(EHP_ERR_ALLOC, "hipMalloc of the voxel projector state failed: %s", hipGetErrorString(e));
        e = hipMemcpy(op->d_pose_axes, axis.data(), n * sizeof(int), hipMemcpyHostToDevice);
        if (e == hipSuccess)
            e = hipMemcpy(op->d_pose_margin, margin.data(), n * sizeof(float), hipMemcpyHostToDevice);
        if (e != hipSuccess)
            return hip_fail(e, "voxel projector state");
        op->tile = dm.dim == 3 ? ehp::TileCfg{32, 32, 1} : ehp::TileCfg{256, 1, 1};
        op->tiled = true;
        return EHP_OK;
    }
} // namespace

int ehp_joseph_create(const ehp_geometry* g, int dtype, int flags, ehp_op** out)
{
    if (flags & ~EHP_SIMPLE_KERNELS)
        return fail(EHP_ERR_ARGUMENT, "unknown Joseph flags 0x%x", flags);
    if (dtype == EHP_BF16 && (flags & EHP_SIMPLE_KERNELS))
        return fail(EHP_ERR_UNSUPPORTED, "bfloat16 Joseph operators run the default kernels only");
    int rc = new_op(g, KIND_JOSEPH, dtype, out);
    if (rc != EHP_OK)
        return rc;
    rc = count_samples(*out);
    if (rc == EHP_OK && !(flags & EHP_SIMPLE_KERNELS))
        rc = setup_fast(*out);
    if (rc == EHP_OK && dtype == EHP_BF16 && !((*out)->planes && (*out)->tiled))
        rc = fail(EHP_ERR_UNSUPPORTED, "this geometry has no plane-major / tiled Joseph kernels, which bfloat16 needs");
    if (rc != EHP_OK) {
        ehp_destroy(*out);
        *out = nullptr;
    }
    return rc;
}

int ehp_siddon_create(const ehp_geometry* g, int dtype, int flags, ehp_op** out)
{
    if (flags & ~EHP_SIMPLE_KERNELS)
        return fail(EHP_ERR_ARGUMENT, "unknown Siddon flags 0x%x", flags);
    if (dtype == EHP_BF16)
        return fail(EHP_ERR_UNSUPPORTED, "bfloat16 storage is implemented for Joseph only");
    int rc = new_op(g, KIND_SIDDON, dtype, out);
    if (rc != EHP_OK)
        return rc;
    rc = count_samples(*out);
    if (rc == EHP_OK && (dtype == EHP_F32 || dtype == EHP_F64) && !(flags & EHP_SIMPLE_KERNELS))
        rc = setup_siddon_tiled(*out);
    if (rc != EHP_OK) {
        ehp_destroy(*out);
        *out = nullptr;
    }
    return rc;
}

int ehp_voxel_create(const ehp_geometry* g, int dtype, const void* lut, int n_lut, double support, int flags,
                     ehp_op** out)
{
    if (flags & ~(EHP_SIMPLE_KERNELS | EHP_DIFFERENTIAL))
        return fail(EHP_ERR_ARGUMENT, "unknown voxel projector flags 0x%x", flags);
    if (dtype == EHP_BF16 && (flags & EHP_SIMPLE_KERNELS))
        return fail(EHP_ERR_UNSUPPORTED, "bfloat16 voxel projectors run the default kernels only");
    if (g && !g->pix.empty())
        return fail(EHP_ERR_UNSUPPORTED, "the voxel projectors take planar detectors only (as elsa's)");
    if (!lut || n_lut < 2)
        return fail(EHP_ERR_ARGUMENT, "the basis-function table needs at least 2 entries");
    if (!(support > 0.0))
        return fail(EHP_ERR_ARGUMENT, "the basis-function support must be positive");
    int rc = new_op(g, KIND_VOXEL, dtype, out);
    if (rc != EHP_OK)
        return rc;
    ehp_op* op = *out;
    op->n_lut = n_lut;
    op->support = (float) support;
    op->differential = (flags & EHP_DIFFERENTIAL) ? 1 : 0;
    // bfloat16 operators compute in FP32 and take an FP32 table
    const size_t bytes = (size_t) n_lut * (dtype == EHP_BF16 ? sizeof(float) : dtype_size(dtype));
    hipError_t e = hipMalloc(&op->d_lut, bytes);
    if (e != hipSuccess) {
        ehp_destroy(op);
        *out = nullptr;
        return fail(EHP_ERR_ALLOC, "hipMalloc of the basis-function table failed: %s", hipGetErrorString(e));
    }
    e = hipMemcpy(op->d_lut, lut, bytes, hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        ehp_destroy(op);
        *out = nullptr;
        return hip_fail(e, "basis-function table upload");
    }
    rc = count_samples(op);
    if (rc == EHP_OK && !(flags & EHP_SIMPLE_KERNELS))
        rc = setup_voxel_tiled(op, lut);
    if (rc == EHP_OK && dtype == EHP_BF16 && !op->tiled)
        rc = fail(EHP_ERR_UNSUPPORTED, "bfloat16 storage needs the tiled voxel forward, which this geometry or "
                                       "table does not allow (source inside the volume, or a table over %d entries)",
                  ehp::voxel_max_lut());
    if (rc != EHP_OK) {
        ehp_destroy(op);
        *out = nullptr;
    }
    return rc;
}

void ehp_destroy(ehp_op* op)
{
    if (!op)
        return;
    int cur = -1;
    if (hipGetDevice(&cur) != hipSuccess)
        cur = -1; // restore nothing afterwards
    if (cur != op->device)
        (void) hipSetDevice(op->device);
    for (void* p : {(void*) op->d_geo, op->d_lut, (void*) op->d_pose_axes, (void*) op->d_table,
                    (void*) op->d_sidtable, (void*) op->d_pose_margin, (void*) op->d_maxbits,
                    (void*) op->d_fixexp, op->d_view[0], op->d_view[1], op->d_view[2], (void*) op->d_acc32,
                    (void*) op->d_pix, op->d_stage})
        if (p)
            (void) hipFree(p);
    if (cur >= 0 && cur != op->device)
        (void) hipSetDevice(cur);
    delete op;
}

// ---------------------------------------------------------------------------------------
// back projection without allocation

namespace
{
    // The scratch of ehp_adjoint_noalloc: the fixed-point scale of one back projection, that is
    // the per-block maxima of |y| (kAbsmaxBlocks 64-bit words) and, kScaleOffset bytes in, the
    // exponent. Every kernel that reads a word of it writes that word first in the same call
    // (absmax before scale, scale before the tiles), so nothing needs clearing. One copy per
    // GPU, held until the process ends; in host code only its device address, looked up per
    // device, is used.
    constexpr size_t kScaleOffset = ehp::kAbsmaxBlocks * sizeof(unsigned long long);
    constexpr size_t kScratchBytes = kScaleOffset + 256;
    static_assert(kScaleOffset % 256 == 0, "the exponent must be 256-byte aligned");
    static __device__ __align__(256) unsigned char scratch[kScratchBytes];

    constexpr int kMaxDevices = 64;
    std::mutex g_scratch_mu;
    void* g_scratch_ptr[kMaxDevices] = {};

    // Device address of this GPU's scratch array (the current device), looked up once per
    // device.
    int scratch_address(int device, unsigned char** out)
    {
        if (device < 0 || device >= kMaxDevices)
            return fail(EHP_ERR_UNSUPPORTED, "device %d is beyond the %d devices of the no-alloc scratch", device,
                        kMaxDevices);
        std::lock_guard<std::mutex> lock(g_scratch_mu);
        if (!g_scratch_ptr[device]) {
            void* p = nullptr;
            const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
            if (e != hipSuccess || !p)
                return hip_fail(e != hipSuccess ? e : hipErrorInvalidSymbol, "address of the no-alloc scratch");
            g_scratch_ptr[device] = p;
        }
        *out = (unsigned char*) g_scratch_ptr[device];
        return EHP_OK;
    }
} // namespace

int ehp_adjoint_noalloc(ehp_op* op, const void* d_sino, void* d_vol, void* stream)
{
    if (!op)
        return fail(EHP_ERR_ARGUMENT, "operator is NULL");
    if (!d_sino || !d_vol)
        return fail(EHP_ERR_ARGUMENT, "NULL data pointer");
    if (d_sino == d_vol)
        return fail(EHP_ERR_ARGUMENT, "input and output must not alias");
    const bool joseph = op->kind == KIND_JOSEPH && (op->dtype == EHP_F32 || op->dtype == EHP_F64);
    const bool siddon = op->kind == KIND_SIDDON && (op->dtype == EHP_F32 || op->dtype == EHP_F64);
    if (!op->tiled || !(joseph || siddon))
        return fail(EHP_ERR_UNSUPPORTED, "ehp_adjoint_noalloc takes FP32/FP64 Joseph and Siddon operators "
                                         "with a tiled back projection only");
    // the scratch this call needs: the partial maxima of at most kAbsmaxBlocks blocks and the exponent
    const size_t need = kScaleOffset + 2 * sizeof(int);
    if (need > kScratchBytes)
        return fail(EHP_ERR_ARGUMENT, "the back projection needs %zu bytes of scratch, the no-alloc maximum is %zu",
                    need, kScratchBytes);
    int cur = -1;
    hipError_t e = hipGetDevice(&cur);
    if (e != hipSuccess)
        return hip_fail(e, "hipGetDevice");
    if (cur != op->device)
        return fail(EHP_ERR_ARGUMENT, "operator lives on device %d, current device is %d", op->device, cur);
    unsigned char* sp = nullptr;
    int rc = scratch_address(cur, &sp);
    if (rc != EHP_OK)
        return rc;
    unsigned int* d_maxbits = (unsigned int*) sp;
    int* d_fixexp = (int*) (sp + kScaleOffset);
    hipStream_t s = (hipStream_t) stream;
    std::lock_guard<std::mutex> lock(op->order.mu);
    const bool capturing = ehp::stream_capturing(s);
    e = ehp::order_begin(op->order, s, capturing);
    if (e != hipSuccess)
        return hip_fail(e, "ordering after the previous call on this operator");
    const ehp::Dims& dm = op->dims;
    if (siddon && op->dtype == EHP_F32)
        e = ehp::siddon_adjoint_tiled<float>(op->d_geo, op->d_sidtable, dm, op->d_pose_axes, op->axes_present,
                                             op->tile, op->batch, op->slot_bound, d_maxbits, d_fixexp,
                                             (const float*) d_sino, (float*) d_vol, s);
    else if (siddon)
        e = ehp::siddon_adjoint_tiled<double>(op->d_geo, op->d_sidtable, dm, op->d_pose_axes, op->axes_present,
                                              op->tile, op->batch, op->slot_bound, d_maxbits, d_fixexp,
                                              (const double*) d_sino, (double*) d_vol, s);
    else if (op->dtype == EHP_F32)
        e = ehp::joseph_adjoint_tiled<float>(op->d_geo, op->d_table, dm, op->d_pose_axes, op->axes_present, op->tile,
                                             op->batch, op->slot_bound, d_maxbits, d_fixexp, (const float*) d_sino,
                                             (float*) d_vol, s);
    else
        e = ehp::joseph_adjoint_tiled<double>(op->d_geo, op->d_table, dm, op->d_pose_axes, op->axes_present,
                                              op->tile, op->batch, op->slot_bound, d_maxbits, d_fixexp,
                                              (const double*) d_sino, (double*) d_vol, s);
    if (e != hipSuccess)
        return hip_fail(e, "adjoint launch");
    e = ehp::order_end(op->order, s, capturing);
    if (e != hipSuccess)
        return hip_fail(e, "recording the end of the call");
    return EHP_OK;
}

int ehp_op_sizes(const ehp_op* op, int64_t* n_vol, int64_t* n_sino)
{
    if (!op)
        return fail(EHP_ERR_ARGUMENT, "operator is NULL");
    if (n_vol)
        *n_vol = op->dims.nvox;
    if (n_sino)
        *n_sino = op->dims.nrays;
    return EHP_OK;
}

int ehp_sample_count(const ehp_op* op, uint64_t* samples)
{
    if (!op || !samples)
        return fail(EHP_ERR_ARGUMENT, "NULL argument");
    *samples = op->samples;
    return EHP_OK;
}

static hipError_t launch(ehp_op* op, const void* in, void* out, hipStream_t s, bool adjoint);

// Whether the kernels of this direction add into the output with global atomics (the
// one-thread-per-ray / per-voxel scatters) rather than writing each element once.
static bool scatters_into_output(const ehp_op* op, bool adjoint)
{
    if (op->dtype == EHP_BF16)
        return false;
    // FP32 and FP64 alike: the detector-tiled voxel forward and the LDS-tiled Joseph and Siddon
    // back projections write each element from one workgroup
    if (op->kind == KIND_VOXEL)
        return !adjoint && !op->tiled;
    return adjoint && !op->tiled;
}

// Memory from hipMalloc (not managed, not host memory mapped for the device).
static bool plain_device_memory(const void* p)
{
    hipPointerAttribute_t a;
    if (hipPointerGetAttributes(&a, p) != hipSuccess) {
        (void) hipGetLastError(); // clear the error of an unknown pointer
        return false;
    }
    return a.type == hipMemoryTypeDevice && !a.isManaged;
}

// Argument checks, then the launches in order after the previous call on the operator.
static int apply(ehp_op* op, const void* in, void* out, void* stream, bool adjoint)
{
    if (!op)
        return fail(EHP_ERR_ARGUMENT, "operator is NULL");
    if (!in || !out)
        return fail(EHP_ERR_ARGUMENT, "NULL data pointer");
    if (in == out)
        return fail(EHP_ERR_ARGUMENT, "input and output must not alias");
    int cur = -1;
    hipError_t e = hipGetDevice(&cur);
    if (e != hipSuccess)
        return hip_fail(e, "hipGetDevice");
    if (cur != op->device)
        return fail(EHP_ERR_ARGUMENT, "operator lives on device %d, current device is %d", op->device, cur);
    hipStream_t s = (hipStream_t) stream;
    std::lock_guard<std::mutex> lock(op->order.mu);
    const bool capturing = ehp::stream_capturing(s);
    e = ehp::order_begin(op->order, s, capturing);
    if (e != hipSuccess)
        return hip_fail(e, "ordering after the previous call on this operator");
    // Atomic scatters go to plain device memory: into the output itself when it is that,
    // else into the operator's staging buffer and then one copy (hardware float atomics are
    // not guaranteed on managed or host-coherent memory).
    void* dst = out;
    size_t out_bytes = 0;
    if (scatters_into_output(op, adjoint) && !plain_device_memory(out)) {
        out_bytes = (size_t) (adjoint ? op->dims.nvox : op->dims.nrays) * dtype_size(op->dtype);
        if (op->stage_bytes < out_bytes) {
            if (capturing)
                return fail(EHP_ERR_UNSUPPORTED, "an output outside plain device memory cannot be captured "
                                                 "(the operator would allocate its staging buffer)");
            (void) hipFree(op->d_stage);
            op->d_stage = nullptr;
            op->stage_bytes = 0;
            e = hipMalloc(&op->d_stage, out_bytes);
            if (e != hipSuccess)
                return fail(EHP_ERR_ALLOC, "hipMalloc of a %zu-byte staging buffer failed: %s", out_bytes,
                            hipGetErrorString(e));
            op->stage_bytes = out_bytes;
        }
        dst = op->d_stage;
    }
    e = launch(op, in, dst, s, adjoint);
    if (e != hipSuccess)
        return hip_fail(e, adjoint ? "adjoint launch" : "forward launch");
    if (dst != out) {
        e = hipMemcpyAsync(out, dst, out_bytes, hipMemcpyDeviceToDevice, s);
        if (e != hipSuccess)
            return hip_fail(e, "copy of the staged output");
    }
    e = ehp::order_end(op->order, s, capturing);
    if (e != hipSuccess)
        return hip_fail(e, "recording the end of the call");
    return EHP_OK;
}

// The kernels of one forward or back projection, enqueued on s.
static hipError_t launch(ehp_op* op, const void* in, void* out, hipStream_t s, bool adjoint)
{
    hipError_t e = hipSuccess;
    const ehp::Dims& dm = op->dims;
    if (op->dtype == EHP_BF16 && op->kind == KIND_VOXEL) {
        // bfloat16 storage, FP32 arithmetic: the tiled forward rounds each pixel's fixed-point
        // sum once, the gather each voxel's FP32 sum
        if (!adjoint)
            e = ehp::voxel_forward_tiled<ehp::bf16, float>(op->d_geo, dm, (const float*) op->d_lut, op->n_lut,
                                                           op->support, op->differential, op->d_pose_axes,
                                                           op->d_pose_margin, op->tile, op->slot_bound, op->d_maxbits,
                                                           op->d_fixexp, (const ehp::bf16*) in, (ehp::bf16*) out, s);
        else
            e = ehp::voxel_adjoint_fast<ehp::bf16>(op->d_geo, dm, (const float*) op->d_lut, op->n_lut, op->support,
                                                   op->differential, (const ehp::bf16*) in, (ehp::bf16*) out, s);
    } else if (op->dtype == EHP_BF16) {
        // bfloat16 storage, FP32 arithmetic: the forward reads bfloat16 views and rounds each ray
        // sum once; the back projection sums in fixed point into FP32 and rounds each voxel once
        if (!adjoint) {
            e = ehp::joseph_forward_planes<ehp::bf16>(op->d_geo, dm, op->layouts, op->d_view, op->axes_present,
                                                      (const ehp::bf16*) in, (ehp::bf16*) out, s);
        } else {
            e = ehp::joseph_adjoint_tiled<ehp::bf16>(op->d_geo, op->d_table, dm, op->d_pose_axes, op->axes_present,
                                                     op->tile, op->batch, op->slot_bound, op->d_maxbits, op->d_fixexp,
                                                     (const ehp::bf16*) in, op->d_acc32, s);
            if (e == hipSuccess)
                e = ehp::launch_to_bf16(op->d_acc32, (ehp::bf16*) out, dm.nvox, s);
        }
    } else if (op->dtype == EHP_F32) {
        const float* x = (const float*) in;
        float* y = (float*) out;
        if (op->kind == KIND_JOSEPH && adjoint && op->tiled)
            e = ehp::joseph_adjoint_tiled(op->d_geo, op->d_table, dm, op->d_pose_axes, op->axes_present, op->tile,
                                          op->batch, op->slot_bound, op->d_maxbits, op->d_fixexp, x, y, s);
        else if (op->kind == KIND_JOSEPH && !adjoint && op->planes)
            e = ehp::joseph_forward_planes<float>(op->d_geo, dm, op->layouts, op->d_view, op->axes_present, x, y, s);
        else if (op->kind == KIND_JOSEPH)
            e = adjoint ? ehp::joseph_adjoint<float>(op->d_geo, dm, x, y, s) : ehp::joseph_forward<float>(op->d_geo, dm, x, y, s);
        else if (op->kind == KIND_SIDDON && !adjoint && op->planes)
            e = ehp::siddon_forward_planes<float>(op->d_sidtable, dm, op->layouts, op->d_view, op->axes_present, x, y,
                                                  s);
        else if (op->kind == KIND_SIDDON && adjoint && op->tiled)
            e = ehp::siddon_adjoint_tiled(op->d_geo, op->d_sidtable, dm, op->d_pose_axes, op->axes_present, op->tile,
                                          op->batch, op->slot_bound, op->d_maxbits, op->d_fixexp, x, y, s);
        else if (op->kind == KIND_SIDDON)
            e = adjoint ? ehp::siddon_adjoint<float>(op->d_geo, dm, x, y, s) : ehp::siddon_forward<float>(op->d_geo, dm, x, y, s);
        else if (!adjoint && op->tiled)
            e = ehp::voxel_forward_tiled<float, float>(op->d_geo, dm, (const float*) op->d_lut, op->n_lut,
                                                       op->support, op->differential, op->d_pose_axes,
                                                       op->d_pose_margin, op->tile, op->slot_bound, op->d_maxbits,
                                                       op->d_fixexp, x, y, s);
        else if (adjoint && op->tiled)
            e = ehp::voxel_adjoint_fast<float>(op->d_geo, dm, (const float*) op->d_lut, op->n_lut, op->support,
                                               op->differential, x, y, s);
        else
            e = adjoint ? ehp::voxel_adjoint<float>(op->d_geo, dm, (const float*) op->d_lut, op->n_lut, op->support,
                                                    op->differential, x, y, s)
                        : ehp::voxel_forward<float>(op->d_geo, dm, (const float*) op->d_lut, op->n_lut, op->support,
                                                    op->differential, x, y, s);
    } else {
        const double* x = (const double*) in;
        double* y = (double*) out;
        if (op->kind == KIND_JOSEPH && adjoint && op->tiled)
            e = ehp::joseph_adjoint_tiled<double>(op->d_geo, op->d_table, dm, op->d_pose_axes, op->axes_present,
                                                  op->tile, op->batch, op->slot_bound, op->d_maxbits, op->d_fixexp, x,
                                                  y, s);
        else if (op->kind == KIND_JOSEPH && !adjoint && op->planes)
            e = ehp::joseph_forward_planes<double>(op->d_geo, dm, op->layouts, op->d_view, op->axes_present, x, y, s);
        else if (op->kind == KIND_JOSEPH)
            e = adjoint ? ehp::joseph_adjoint<double>(op->d_geo, dm, x, y, s) : ehp::joseph_forward<double>(op->d_geo, dm, x, y, s);
        else if (op->kind == KIND_SIDDON && !adjoint && op->planes)
            e = ehp::siddon_forward_planes<double>(op->d_sidtable, dm, op->layouts, op->d_view, op->axes_present, x,
                                                   y, s);
        else if (op->kind == KIND_SIDDON && adjoint && op->tiled)
            e = ehp::siddon_adjoint_tiled<double>(op->d_geo, op->d_sidtable, dm, op->d_pose_axes, op->axes_present,
                                                  op->tile, op->batch, op->slot_bound, op->d_maxbits, op->d_fixexp, x,
                                                  y, s);
        else if (op->kind == KIND_SIDDON)
            e = adjoint ? ehp::siddon_adjoint<double>(op->d_geo, dm, x, y, s) : ehp::siddon_forward<double>(op->d_geo, dm, x, y, s);
        else if (!adjoint && op->tiled)
            e = ehp::voxel_forward_tiled<double, double>(op->d_geo, dm, (const double*) op->d_lut, op->n_lut,
                                                         op->support, op->differential, op->d_pose_axes,
                                                         op->d_pose_margin, op->tile, op->slot_bound, op->d_maxbits,
                                                         op->d_fixexp, x, y, s);
        else
            e = adjoint ? ehp::voxel_adjoint<double>(op->d_geo, dm, (const double*) op->d_lut, op->n_lut, op->support,
                                                     op->differential, x, y, s)
                        : ehp::voxel_forward<double>(op->d_geo, dm, (const double*) op->d_lut, op->n_lut, op->support,
                                                     op->differential, x, y, s);
    }
    return e;
}

int ehp_forward(ehp_op* op, const void* d_vol, void* d_sino, void* stream)
{
    return apply(op, d_vol, d_sino, stream, false);
}

int ehp_adjoint(ehp_op* op, const void* d_sino, void* d_vol, void* stream)
{
    return apply(op, d_sino, d_vol, stream, true);
}

int ehp_tile(ehp_op* op, int la, int tb, int tc, int* la_out, int* tb_out, int* tc_out)
{
    if (!op)
        return fail(EHP_ERR_ARGUMENT, "operator is NULL");
    if (!op->tiled)
        return fail(EHP_ERR_ARGUMENT, "operator does not use a tiled kernel");
    std::lock_guard<std::mutex> lock(op->order.mu);
    if (la > 0) {
        const bool vox = op->kind == KIND_VOXEL;
        if (tb <= 0 || (op->dims.dim == 3 && !vox && tc <= 0))
            return fail(EHP_ERR_ARGUMENT, "tile extents must be positive");
        ehp::TileCfg t{la, tb, op->dims.dim == 3 ? tc : 1};
        if (vox)
            t = ehp::TileCfg{la, op->dims.dim == 3 ? tb : 1, 1};
        int max_lds = 65536;
        if (hipDeviceGetAttribute(&max_lds, hipDeviceAttributeMaxSharedMemoryPerBlock, op->device) != hipSuccess)
            max_lds = 65536;
        const size_t need = tile_lds_bytes(op->kind, op->dims, t, op->dtype);
        if ((int64_t) need > max_lds)
            return fail(EHP_ERR_ARGUMENT, "tile needs %zu bytes of LDS, the device allows %d", need, max_lds);
        op->tile = t;
    }
    if (la_out)
        *la_out = op->tile.LA;
    if (tb_out)
        *tb_out = op->tile.TB;
    if (tc_out)
        *tc_out = op->tile.TC;
    return EHP_OK;
}

int ehp_pose_batch(ehp_op* op, int batch, int* batch_out)
{
    if (!op)
        return fail(EHP_ERR_ARGUMENT, "operator is NULL");
    if (!op->tiled || op->kind == KIND_VOXEL)
        return fail(EHP_ERR_ARGUMENT, "operator has no tiled back projection");
    std::lock_guard<std::mutex> lock(op->order.mu);
    if (batch > 0) {
        if (batch > ehp::kMaxPoseBatch)
            return fail(EHP_ERR_ARGUMENT, "pose batch %d exceeds %d", batch, ehp::kMaxPoseBatch);
        op->batch = batch;
    }
    if (batch_out)
        *batch_out = op->batch;
    return EHP_OK;
}

int ehp_joseph_ray_info(const ehp_op* op, int32_t* axis, int64_t* nsteps)
{
    if (!op || !axis || !nsteps)
        return fail(EHP_ERR_ARGUMENT, "NULL argument");
    if (op->kind != KIND_JOSEPH)
        return fail(EHP_ERR_ARGUMENT, "not a Joseph operator");
    const int64_t n = op->dims.nrays;
    int32_t* d_axis = nullptr;
    int64_t* d_steps = nullptr;
    hipError_t e = hipMalloc((void**) &d_axis, n * sizeof(int32_t));
    if (e == hipSuccess)
        e = hipMalloc((void**) &d_steps, n * sizeof(int64_t));
    if (e != hipSuccess) {
        (void) hipFree(d_axis);
        return fail(EHP_ERR_ALLOC, "hipMalloc for ray info failed: %s", hipGetErrorString(e));
    }
    e = ehp::joseph_info(op->d_geo, op->dims, d_axis, d_steps, nullptr, nullptr);
    if (e == hipSuccess)
        e = hipMemcpy(axis, d_axis, n * sizeof(int32_t), hipMemcpyDeviceToHost);
    if (e == hipSuccess)
        e = hipMemcpy(nsteps, d_steps, n * sizeof(int64_t), hipMemcpyDeviceToHost);
    (void) hipFree(d_axis);
    (void) hipFree(d_steps);
    if (e != hipSuccess)
        return hip_fail(e, "ray info");
    return EHP_OK;
}

} // extern "C"

// ---- solvers -----------------------------------------------------------------------------

int ehp_solver_fail(int code, const char* msg)
{
    return fail(code, "%s", msg);
}

namespace
{
    // one operator as a batch of one for the solvers
    struct SingleOps {
        int64_t doff[2], roff[2];
        ehp::SolveOps ops;
        explicit SingleOps(ehp_op* op)
        {
            doff[0] = roff[0] = 0;
            doff[1] = op->dims.nvox;
            roff[1] = op->dims.nrays;
            ops.nseg = 1;
            ops.doff = doff;
            ops.roff = roff;
            ops.shapes = {op->dims.dim, op->dims.vol[0], op->dims.vol[1], op->dims.vol[2]};
            ops.fwd = [op](const void* x, void* y, hipStream_t s) { return ehp_forward(op, x, y, s); };
            ops.adj = [op](const void* y, void* x, hipStream_t s) { return ehp_adjoint(op, y, x, s); };
        }
    };
} // namespace

namespace
{
    int solver_args(const ehp_op* op, const void* b, const void* x, int iterations)
    {
        if (!op)
            return fail(EHP_ERR_ARGUMENT, "operator is NULL");
        if (!b || !x)
            return fail(EHP_ERR_ARGUMENT, "NULL data pointer");
        if (b == x)
            return fail(EHP_ERR_ARGUMENT, "b and x must not alias");
        if (iterations < 0)
            return fail(EHP_ERR_ARGUMENT, "negative iteration count %d", iterations);
        int cur = -1;
        hipError_t e = hipGetDevice(&cur);
        if (e != hipSuccess)
            return hip_fail(e, "hipGetDevice");
        if (cur != op->device)
            return fail(EHP_ERR_ARGUMENT, "operator lives on device %d, current device is %d", op->device, cur);
        return EHP_OK;
    }
} // namespace

int ehp_cgls(ehp_op* op, const void* d_b, void* d_x, int use_x0, int iterations, double eps, double tol, int flags,
             double* residuals, int* iterations_done, void* stream)
{
    int rc = solver_args(op, d_b, d_x, iterations);
    if (rc != EHP_OK)
        return rc;
    if (flags & ~EHP_SOLVE_GRAPH)
        return fail(EHP_ERR_ARGUMENT, "unknown solver flags 0x%x", flags);
    SingleOps A(op);
    return ehp::solve_cgls(A.ops, op->dtype, d_b, d_x, use_x0, iterations, eps, tol, (flags & EHP_SOLVE_GRAPH) != 0,
                           residuals, iterations_done, (hipStream_t) stream);
}

int ehp_apgd(ehp_op* op, const void* d_b, void* d_x, int use_x0, int iterations, double mu, double l1_sigma,
             double epsilon, int flags, double* grad_sq_norms, int* iterations_done, double* mu_used, void* stream)
{
    int rc = solver_args(op, d_b, d_x, iterations);
    if (rc != EHP_OK)
        return rc;
    if (flags & ~EHP_SOLVE_GRAPH)
        return fail(EHP_ERR_ARGUMENT, "unknown solver flags 0x%x", flags);
    SingleOps A(op);
    return ehp::solve_apgd(A.ops, op->dtype, d_b, d_x, use_x0, iterations, &mu, l1_sigma, epsilon,
                           (flags & EHP_SOLVE_GRAPH) != 0, grad_sq_norms, iterations_done, mu_used,
                           (hipStream_t) stream);
}

int ehp_admml2(ehp_op* op, const void* d_b, void* d_x, int use_x0, int iterations, double tau, double l1_sigma,
               int ninner, int flags, double* residuals, double* tau_used, void* stream)
{
    int rc = solver_args(op, d_b, d_x, iterations);
    if (rc != EHP_OK)
        return rc;
    if (flags & ~(EHP_SOLVE_GRAPH | EHP_ADMM_FINITE_DIFFERENCES))
        return fail(EHP_ERR_ARGUMENT, "unknown solver flags 0x%x", flags);
    if (ninner < 0)
        return fail(EHP_ERR_ARGUMENT, "negative inner iteration count %d", ninner);
    SingleOps A(op);
    return ehp::solve_admml2(A.ops, op->dtype, d_b, d_x, use_x0, iterations, tau, l1_sigma, ninner,
                             (flags & EHP_ADMM_FINITE_DIFFERENCES) != 0, (flags & EHP_SOLVE_GRAPH) != 0, residuals,
                             tau_used, (hipStream_t) stream);
}

int ehp_cgls_refined(ehp_op* op_hi, ehp_op* op_lo, const void* d_b, void* d_x, int use_x0, int outer, int inner,
                     int flags, double* residuals, void* stream)
{
    int rc = solver_args(op_hi, d_b, d_x, outer);
    if (rc != EHP_OK)
        return rc;
    if (!op_lo)
        return fail(EHP_ERR_ARGUMENT, "operator is NULL");
    if (op_hi->dtype != EHP_F32 || op_lo->dtype != EHP_BF16)
        return fail(EHP_ERR_ARGUMENT, "refinement needs an FP32 operator and a bfloat16 operator");
    if (op_hi->kind != op_lo->kind || op_hi->dims.nvox != op_lo->dims.nvox || op_hi->dims.nrays != op_lo->dims.nrays ||
        op_hi->host_geo.size() != op_lo->host_geo.size() ||
        std::memcmp(op_hi->host_geo.data(), op_lo->host_geo.data(), op_hi->host_geo.size() * sizeof(ehp::PoseGeom)))
        return fail(EHP_ERR_ARGUMENT, "the two operators must have the same kind and geometry");
    if (inner < 1)
        return fail(EHP_ERR_ARGUMENT, "inner iterations must be at least 1 (got %d)", inner);
    if (flags & ~EHP_SOLVE_GRAPH)
        return fail(EHP_ERR_ARGUMENT, "unknown solver flags 0x%x", flags);
    SingleOps hi(op_hi), lo(op_lo);
    return ehp::solve_cgls_refined(hi.ops, lo.ops, (const float*) d_b, (float*) d_x, use_x0, outer, inner,
                                   (flags & EHP_SOLVE_GRAPH) != 0, residuals, (hipStream_t) stream);
}

int ehp_power_iterations(ehp_op* op, int niters, double* L, void* stream)
{
    if (!op || !L)
        return fail(EHP_ERR_ARGUMENT, "NULL argument");
    if (niters < 1)
        return fail(EHP_ERR_ARGUMENT, "power iterations need niters >= 1 (got %d)", niters);
    int cur = -1;
    hipError_t e = hipGetDevice(&cur);
    if (e != hipSuccess)
        return hip_fail(e, "hipGetDevice");
    if (cur != op->device)
        return fail(EHP_ERR_ARGUMENT, "operator lives on device %d, current device is %d", op->device, cur);
    SingleOps A(op);
    return ehp::solve_power_iterations(A.ops, op->dtype, niters, L, (hipStream_t) stream);
}
