// ehp_batch.hip -- batches of independent problems (declared in include/elsa_hip.h).
//
// A batch applies many operators -- each with its own volume size, trajectory and detector --
// as one block-diagonal operator over packed volumes and sinograms. For FP32 Joseph or FP32
// Siddon operators with the fast kernels (all of one kind) a projector pass is one launch over
// all problems per kernel: the launch's blocks are the concatenation of the blocks every
// operator's own launch would run (job tables built here once), so each problem's result is
// bit for bit its operator's, while the GPU is filled by the batch as a whole instead of by one
// small problem at a time. Otherwise the batch calls the operators one after another.
#include "elsa_hip.h"
#include "ehp_op.hpp"
#include "ehp_solvers.hpp"

#include <algorithm>
#include <cstdarg>
#include <cstdio>
#include <mutex>
#include <vector>

int ehp_solver_fail(int code, const char* msg);

struct ehp_batch {
    int device = 0;
    int dtype = EHP_F32;
    int dim = 0;
    std::vector<ehp_op*> ops;
    std::vector<int64_t> vol_off, sino_off; // n + 1
    bool fused = false;
    int kind = KIND_JOSEPH; // of the fused passes
    // fused Siddon passes (device tables; the copy, segment and block tables are shared)
    ehp::SidFwdJob* d_sfjobs = nullptr;
    ehp::SidAdjJob* d_sajobs[3] = {nullptr, nullptr, nullptr};
    // fused voxel passes (the forward's blocks in d_fblocks, its segments over the volumes)
    ehp::VoxFwdJob* d_vfjobs = nullptr;
    ehp::VoxAdjJob* d_vajobs = nullptr;
    int* d_vablocks = nullptr;
    int64_t nvablocks = 0;
    // fused Joseph passes (device tables)
    ehp::CopyJob* d_cjobs = nullptr;
    int* d_cblocks = nullptr;
    int64_t ncblocks = 0;
    ehp::FwdJob* d_fjobs = nullptr;
    int* d_fblocks = nullptr;
    int64_t nfblocks = 0;
    ehp::SegJob* d_segs = nullptr;
    int* d_sblocks = nullptr;
    int64_t nsblocks = 0;
    double* d_bounds = nullptr;
    ehp::AdjJob* d_ajobs[3] = {nullptr, nullptr, nullptr};
    int* d_ablocks[3] = {nullptr, nullptr, nullptr};
    int64_t nablocks[3] = {0, 0, 0};
    size_t lds[3] = {0, 0, 0};
    unsigned int* d_maxbits = nullptr;
    int* d_fixexp = nullptr;
    std::vector<void*> allocs;
    ehp::OrderGuard order; // fused passes: ordered with each other and with the operators' own calls

    ~ehp_batch()
    {
        for (void* p : allocs)
            (void) hipFree(p);
    }
};

namespace
{
    int bfail(int code, const char* fmt, ...)
    {
        char buf[512];
        va_list ap;
        va_start(ap, fmt);
        std::vsnprintf(buf, sizeof(buf), fmt, ap);
        va_end(ap);
        return ehp_solver_fail(code, buf);
    }

    template <class T>
    hipError_t upload(ehp_batch* b, const std::vector<T>& h, T** d)
    {
        *d = nullptr;
        if (h.empty())
            return hipSuccess;
        hipError_t e = hipMalloc((void**) d, h.size() * sizeof(T));
        if (e != hipSuccess)
            return e;
        b->allocs.push_back((void*) *d);
        return hipMemcpy(*d, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice);
    }

    bool fusable(const ehp_op* op, int kind, int dim)
    {
        if (op->kind != kind || op->dtype != EHP_F32 || !op->tiled || op->dims.dim != dim)
            return false;
        return op->kind == KIND_VOXEL || op->planes;
    }

    int build_fused(ehp_batch* b)
    {
        const int n = (int) b->ops.size();
        std::vector<ehp::JosephBatchItem> items(n);
        std::vector<double> bounds(n);
        for (int j = 0; j < n; ++j) {
            const ehp_op* op = b->ops[j];
            ehp::JosephBatchItem& it = items[j];
            it.geo = op->d_geo;
            it.table = op->d_table;
            it.sidtable = op->d_sidtable;
            it.pose_axes = op->d_pose_axes;
            for (int a = 0; a < 3; ++a) {
                it.views[a] = (float*) op->d_view[a];
                it.layouts[a] = op->layouts[a];
            }
            it.dims = op->dims;
            it.tile = op->tile;
            it.axes_present = op->axes_present;
            it.batch = op->batch;
            it.vol_off = b->vol_off[j];
            it.sino_off = b->sino_off[j];
            bounds[j] = op->slot_bound;
        }
        std::vector<ehp::CopyJob> cj;
        std::vector<ehp::FwdJob> fj;
        std::vector<ehp::SidFwdJob> sfj;
        std::vector<int> cb, fb, sb;
        std::vector<ehp::SegJob> sg;
        std::vector<ehp::AdjJob> aj[3];
        std::vector<ehp::SidAdjJob> saj[3];
        std::vector<int> ab[3];
        const bool sid = b->kind == KIND_SIDDON, vox = b->kind == KIND_VOXEL;
        std::vector<ehp::VoxFwdJob> vfj;
        std::vector<ehp::VoxAdjJob> vaj;
        std::vector<int> vab;
        if (vox) {
            // the forward's fixed-point scale comes from each problem's volume
            std::vector<ehp::VoxBatchItem> vitems(n);
            std::vector<int64_t> offs(n), sizes(n);
            for (int j = 0; j < n; ++j) {
                const ehp_op* op = b->ops[j];
                ehp::VoxBatchItem& v = vitems[j];
                v.geo = op->d_geo;
                v.dims = op->dims;
                v.lut = (const float*) op->d_lut;
                v.n_lut = op->n_lut;
                v.support = op->support;
                v.diff = op->differential;
                v.pose_axes = op->d_pose_axes;
                v.pose_margin = op->d_pose_margin;
                v.tile = op->tile;
                v.vol_off = b->vol_off[j];
                v.sino_off = b->sino_off[j];
                offs[j] = b->vol_off[j];
                sizes[j] = op->dims.nvox;
            }
            ehp::voxel_forward_batch_jobs(vitems.data(), n, vfj, fb, b->lds[0]);
            ehp::batch_segments(offs.data(), sizes.data(), n, sg, sb);
            ehp::voxel_adjoint_batch_jobs(vitems.data(), n, vaj, vab);
        } else if (sid) {
            std::vector<int64_t> offs(n), sizes(n);
            for (int j = 0; j < n; ++j) {
                offs[j] = b->sino_off[j];
                sizes[j] = b->ops[j]->dims.nrays;
            }
            ehp::siddon_forward_batch_jobs(items.data(), n, cj, cb, sfj, fb);
            ehp::batch_segments(offs.data(), sizes.data(), n, sg, sb);
            ehp::siddon_adjoint_batch_jobs(items.data(), n, saj, ab, b->lds);
        } else {
            ehp::joseph_forward_batch_jobs(items.data(), n, cj, cb, fj, fb);
            ehp::joseph_adjoint_batch_jobs(items.data(), n, sg, sb, aj, ab, b->lds);
        }
        int max_lds = 65536;
        if (hipDeviceGetAttribute(&max_lds, hipDeviceAttributeMaxSharedMemoryPerBlock, b->device) != hipSuccess)
            max_lds = 65536;
        for (int a = 0; a < 3; ++a)
            if ((int64_t) b->lds[a] > max_lds)
                return bfail(EHP_ERR_ARGUMENT, "batch tiles need %zu bytes of LDS, the device allows %d", b->lds[a],
                             max_lds);
        hipError_t e = upload(b, cj, &b->d_cjobs);
        if (e == hipSuccess)
            e = upload(b, cb, &b->d_cblocks);
        if (e == hipSuccess)
            e = vox ? upload(b, vfj, &b->d_vfjobs) : (sid ? upload(b, sfj, &b->d_sfjobs) : upload(b, fj, &b->d_fjobs));
        if (e == hipSuccess && vox) {
            e = upload(b, vaj, &b->d_vajobs);
            if (e == hipSuccess)
                e = upload(b, vab, &b->d_vablocks);
            b->nvablocks = (int64_t) vab.size();
        }
        if (e == hipSuccess)
            e = upload(b, fb, &b->d_fblocks);
        if (e == hipSuccess)
            e = upload(b, sg, &b->d_segs);
        if (e == hipSuccess)
            e = upload(b, sb, &b->d_sblocks);
        if (e == hipSuccess)
            e = upload(b, bounds, &b->d_bounds);
        for (int a = 0; a < 3 && e == hipSuccess; ++a) {
            e = sid ? upload(b, saj[a], &b->d_sajobs[a]) : upload(b, aj[a], &b->d_ajobs[a]);
            if (e == hipSuccess)
                e = upload(b, ab[a], &b->d_ablocks[a]);
            b->nablocks[a] = (int64_t) ab[a].size();
        }
        if (e == hipSuccess) {
            e = hipMalloc((void**) &b->d_maxbits, (sb.size() + 1) * sizeof(unsigned int)); // one per max-|y| block
            if (e == hipSuccess)
                b->allocs.push_back(b->d_maxbits);
        }
        if (e == hipSuccess) {
            e = hipMalloc((void**) &b->d_fixexp, n * sizeof(int));
            if (e == hipSuccess)
                b->allocs.push_back(b->d_fixexp);
        }
        if (e != hipSuccess)
            return bfail(EHP_ERR_ALLOC, "batch tables: %s", hipGetErrorString(e));
        b->ncblocks = (int64_t) cb.size();
        b->nfblocks = (int64_t) fb.size();
        b->nsblocks = (int64_t) sb.size();
        b->fused = true;
        return EHP_OK;
    }

    int check_device(const ehp_batch* b)
    {
        int cur = -1;
        hipError_t e = hipGetDevice(&cur);
        if (e != hipSuccess)
            return bfail(EHP_ERR_HIP, "hipGetDevice: %s", hipGetErrorString(e));
        if (cur != b->device)
            return bfail(EHP_ERR_ARGUMENT, "batch lives on device %d, current device is %d", b->device, cur);
        return EHP_OK;
    }

    int batch_apply(ehp_batch* b, const void* in, void* out, hipStream_t s, bool adjoint)
    {
        if (!b->fused) {
            const size_t es = b->dtype == EHP_F64 ? sizeof(double) : (b->dtype == EHP_BF16 ? 2 : sizeof(float));
            for (size_t j = 0; j < b->ops.size(); ++j) {
                const int64_t ai = adjoint ? b->sino_off[j] : b->vol_off[j];
                const int64_t ao = adjoint ? b->vol_off[j] : b->sino_off[j];
                const void* pi = (const char*) in + ai * es;
                void* po = (char*) out + ao * es;
                const int rc = adjoint ? ehp_adjoint(b->ops[j], pi, po, s) : ehp_forward(b->ops[j], pi, po, s);
                if (rc != EHP_OK)
                    return rc;
            }
            return EHP_OK;
        }
        // A fused pass rewrites the operators' views and the batch's scale: it runs after the
        // previous call on the batch and on every operator of it, and the operators' next calls
        // run after it. The mutexes are taken in address order, so two callers cannot deadlock.
        std::vector<std::mutex*> mus{&b->order.mu};
        for (ehp_op* op : b->ops)
            mus.push_back(&op->order.mu);
        std::sort(mus.begin(), mus.end());
        std::vector<std::unique_lock<std::mutex>> locks;
        locks.reserve(mus.size());
        for (std::mutex* m : mus)
            locks.emplace_back(*m);
        const bool capturing = ehp::stream_capturing(s);
        hipError_t e = ehp::order_begin(b->order, s, capturing);
        std::vector<const ehp::SharedEvent*> waited;
        for (size_t j = 0; e == hipSuccess && j < b->ops.size(); ++j) {
            const ehp::OrderGuard& g = b->ops[j]->order;
            if (!g.last || std::find(waited.begin(), waited.end(), g.last.get()) != waited.end())
                continue;
            waited.push_back(g.last.get());
            e = ehp::order_begin(g, s, capturing);
        }
        if (e != hipSuccess)
            return bfail(EHP_ERR_HIP, "ordering after the previous calls: %s", hipGetErrorString(e));
        if (!adjoint && b->kind == KIND_VOXEL) {
            e = ehp::fixed_point_exponent_batch((int) b->ops.size(), b->d_segs, b->d_sblocks, b->nsblocks,
                                                b->d_bounds, b->d_maxbits, b->d_fixexp, (const float*) in, s);
            if (e == hipSuccess)
                e = ehp::voxel_forward_batch(b->dim, b->d_vfjobs, b->d_fblocks, b->nfblocks, b->lds[0], b->d_fixexp,
                                             (const float*) in, (float*) out, s);
        } else if (adjoint && b->kind == KIND_VOXEL) {
            e = ehp::voxel_adjoint_batch(b->dim, b->d_vajobs, b->d_vablocks, b->nvablocks, (const float*) in,
                                         (float*) out, s);
        } else if (!adjoint && b->kind == KIND_SIDDON) {
            e = ehp::siddon_forward_batch(b->dim, b->d_cjobs, b->d_cblocks, b->ncblocks, b->d_sfjobs, b->d_fblocks,
                                          b->nfblocks, (const float*) in, (float*) out, s);
        } else if (!adjoint) {
            e = ehp::joseph_forward_batch(b->dim, b->d_cjobs, b->d_cblocks, b->ncblocks, b->d_fjobs, b->d_fblocks,
                                          b->nfblocks, (const float*) in, (float*) out, s);
        } else if (b->kind == KIND_SIDDON) {
            e = ehp::launch_fill<float>((float*) out, b->vol_off.back(), 0.0f, s);
            if (e == hipSuccess)
                e = ehp::fixed_point_exponent_batch((int) b->ops.size(), b->d_segs, b->d_sblocks, b->nsblocks,
                                                    b->d_bounds, b->d_maxbits, b->d_fixexp, (const float*) in, s);
            const ehp::SidAdjJob* sj[3] = {b->d_sajobs[0], b->d_sajobs[1], b->d_sajobs[2]};
            const int* ab[3] = {b->d_ablocks[0], b->d_ablocks[1], b->d_ablocks[2]};
            if (e == hipSuccess)
                e = ehp::siddon_adjoint_batch(b->dim, sj, ab, b->nablocks, b->lds, b->d_fixexp, (const float*) in,
                                              (float*) out, s);
        } else {
            const ehp::AdjJob* aj[3] = {b->d_ajobs[0], b->d_ajobs[1], b->d_ajobs[2]};
            const int* ab[3] = {b->d_ablocks[0], b->d_ablocks[1], b->d_ablocks[2]};
            e = ehp::joseph_adjoint_batch(b->dim, (int) b->ops.size(), b->d_segs, b->d_sblocks, b->nsblocks,
                                          b->d_bounds, aj, ab, b->nablocks, b->lds, b->vol_off.back(), b->d_maxbits,
                                          b->d_fixexp, (const float*) in, (float*) out, s);
        }
        if (e != hipSuccess)
            return bfail(EHP_ERR_HIP, "%s launch: %s", adjoint ? "batch adjoint" : "batch forward",
                         hipGetErrorString(e));
        e = ehp::order_end(b->order, s, capturing);
        if (e != hipSuccess)
            return bfail(EHP_ERR_HIP, "recording the end of the pass: %s", hipGetErrorString(e));
        if (!capturing)
            for (ehp_op* op : b->ops) {
                op->order.last = b->order.last;
                op->order.last_stream = s;
            }
        return EHP_OK;
    }

    int solver_args(const ehp_batch* b, const void* db, const void* dx, int iterations)
    {
        if (!b)
            return bfail(EHP_ERR_ARGUMENT, "batch is NULL");
        if (!db || !dx)
            return bfail(EHP_ERR_ARGUMENT, "NULL data pointer");
        if (db == dx)
            return bfail(EHP_ERR_ARGUMENT, "b and x must not alias");
        if (iterations < 0)
            return bfail(EHP_ERR_ARGUMENT, "negative iteration count %d", iterations);
        return check_device(b);
    }

    ehp::SolveOps solve_ops(ehp_batch* b)
    {
        ehp::SolveOps A;
        A.nseg = (int) b->ops.size();
        A.doff = b->vol_off.data();
        A.roff = b->sino_off.data();
        for (const ehp_op* op : b->ops) {
            A.shapes.push_back(op->dims.dim);
            for (int i = 0; i < 3; ++i)
                A.shapes.push_back(op->dims.vol[i]);
        }
        A.fwd = [b](const void* x, void* y, hipStream_t s) { return batch_apply(b, x, y, s, false); };
        A.adj = [b](const void* y, void* x, hipStream_t s) { return batch_apply(b, y, x, s, true); };
        return A;
    }
} // namespace

extern "C" {

int ehp_batch_create(ehp_op* const* ops, int n, ehp_batch** out)
{
    if (!out)
        return bfail(EHP_ERR_ARGUMENT, "output handle is NULL");
    *out = nullptr;
    if (!ops || n <= 0)
        return bfail(EHP_ERR_ARGUMENT, "a batch needs at least one operator");
    for (int j = 0; j < n; ++j) {
        if (!ops[j])
            return bfail(EHP_ERR_ARGUMENT, "operator %d is NULL", j);
        if (ops[j]->dtype != ops[0]->dtype)
            return bfail(EHP_ERR_ARGUMENT, "operator %d has another data type than operator 0", j);
        if (ops[j]->device != ops[0]->device)
            return bfail(EHP_ERR_ARGUMENT, "operator %d lives on another device than operator 0", j);
        for (int k = 0; k < j; ++k)
            if (ops[k] == ops[j])
                return bfail(EHP_ERR_ARGUMENT, "operator %d appears twice", j);
    }
    ehp_batch* b = new ehp_batch();
    b->device = ops[0]->device;
    b->dtype = ops[0]->dtype;
    b->dim = ops[0]->dims.dim;
    b->ops.assign(ops, ops + n);
    b->vol_off.assign(n + 1, 0);
    b->sino_off.assign(n + 1, 0);
    bool fuse = true;
    b->kind = ops[0]->kind;
    for (int j = 0; j < n; ++j) {
        b->vol_off[j + 1] = b->vol_off[j] + ops[j]->dims.nvox;
        b->sino_off[j + 1] = b->sino_off[j] + ops[j]->dims.nrays;
        fuse = fuse && fusable(ops[j], b->kind, b->dim);
    }
    int cur = -1;
    if (hipGetDevice(&cur) != hipSuccess || cur != b->device) {
        delete b;
        return bfail(EHP_ERR_ARGUMENT, "the operators live on device %d, current device is %d", ops[0]->device, cur);
    }
    if (fuse) {
        const int rc = build_fused(b);
        if (rc != EHP_OK) {
            delete b;
            return rc;
        }
    }
    *out = b;
    return EHP_OK;
}

void ehp_batch_destroy(ehp_batch* b)
{
    if (!b)
        return;
    const int dev = b->device;
    int cur = -1;
    (void) hipGetDevice(&cur);
    if (cur != dev)
        (void) hipSetDevice(dev);
    delete b;
    if (cur >= 0 && cur != dev)
        (void) hipSetDevice(cur);
}

int ehp_batch_layout(const ehp_batch* b, int64_t* vol_offsets, int64_t* sino_offsets, int* fused)
{
    if (!b)
        return bfail(EHP_ERR_ARGUMENT, "batch is NULL");
    for (size_t j = 0; j < b->vol_off.size(); ++j) {
        if (vol_offsets)
            vol_offsets[j] = b->vol_off[j];
        if (sino_offsets)
            sino_offsets[j] = b->sino_off[j];
    }
    if (fused)
        *fused = b->fused ? 1 : 0;
    return EHP_OK;
}

static int batch_entry(ehp_batch* b, const void* in, void* out, void* stream, bool adjoint)
{
    if (!b)
        return bfail(EHP_ERR_ARGUMENT, "batch is NULL");
    if (!in || !out)
        return bfail(EHP_ERR_ARGUMENT, "NULL data pointer");
    if (in == out)
        return bfail(EHP_ERR_ARGUMENT, "input and output must not alias");
    const int rc = check_device(b);
    if (rc != EHP_OK)
        return rc;
    return batch_apply(b, in, out, (hipStream_t) stream, adjoint);
}

int ehp_batch_forward(ehp_batch* b, const void* d_vols, void* d_sinos, void* stream)
{
    return batch_entry(b, d_vols, d_sinos, stream, false);
}

int ehp_batch_adjoint(ehp_batch* b, const void* d_sinos, void* d_vols, void* stream)
{
    return batch_entry(b, d_sinos, d_vols, stream, true);
}

int ehp_batch_cgls(ehp_batch* b, const void* d_b, void* d_x, int use_x0, int iterations, double eps, double tol,
                   int flags, double* residuals, int* iterations_done, void* stream)
{
    int rc = solver_args(b, d_b, d_x, iterations);
    if (rc != EHP_OK)
        return rc;
    if (flags & ~EHP_SOLVE_GRAPH)
        return bfail(EHP_ERR_ARGUMENT, "unknown solver flags 0x%x", flags);
    const ehp::SolveOps A = solve_ops(b);
    return ehp::solve_cgls(A, b->dtype, d_b, d_x, use_x0, iterations, eps, tol, (flags & EHP_SOLVE_GRAPH) != 0,
                           residuals, iterations_done, (hipStream_t) stream);
}

int ehp_batch_apgd(ehp_batch* b, const void* d_b, void* d_x, int use_x0, int iterations, const double* mu,
                   double l1_sigma, double epsilon, int flags, double* grad_sq_norms, int* iterations_done,
                   double* mu_used, void* stream)
{
    int rc = solver_args(b, d_b, d_x, iterations);
    if (rc != EHP_OK)
        return rc;
    if (flags & ~EHP_SOLVE_GRAPH)
        return bfail(EHP_ERR_ARGUMENT, "unknown solver flags 0x%x", flags);
    const ehp::SolveOps A = solve_ops(b);
    return ehp::solve_apgd(A, b->dtype, d_b, d_x, use_x0, iterations, mu, l1_sigma, epsilon,
                           (flags & EHP_SOLVE_GRAPH) != 0, grad_sq_norms, iterations_done, mu_used,
                           (hipStream_t) stream);
}

int ehp_batch_admml2(ehp_batch* b, const void* d_b, void* d_x, int use_x0, int iterations, double tau,
                     double l1_sigma, int ninner, int flags, double* residuals, double* tau_used, void* stream)
{
    int rc = solver_args(b, d_b, d_x, iterations);
    if (rc != EHP_OK)
        return rc;
    if (flags & ~(EHP_SOLVE_GRAPH | EHP_ADMM_FINITE_DIFFERENCES))
        return bfail(EHP_ERR_ARGUMENT, "unknown solver flags 0x%x", flags);
    if (ninner < 0)
        return bfail(EHP_ERR_ARGUMENT, "negative inner iteration count %d", ninner);
    const ehp::SolveOps A = solve_ops(b);
    return ehp::solve_admml2(A, b->dtype, d_b, d_x, use_x0, iterations, tau, l1_sigma, ninner,
                             (flags & EHP_ADMM_FINITE_DIFFERENCES) != 0, (flags & EHP_SOLVE_GRAPH) != 0, residuals,
                             tau_used, (hipStream_t) stream);
}

int ehp_batch_power_iterations(ehp_batch* b, int niters, double* L, void* stream)
{
    if (!b || !L)
        return bfail(EHP_ERR_ARGUMENT, "NULL argument");
    if (niters < 1)
        return bfail(EHP_ERR_ARGUMENT, "power iterations need niters >= 1 (got %d)", niters);
    const int rc = check_device(b);
    if (rc != EHP_OK)
        return rc;
    const ehp::SolveOps A = solve_ops(b);
    return ehp::solve_power_iterations(A, b->dtype, niters, L, (hipStream_t) stream);
}

} // extern "C"
