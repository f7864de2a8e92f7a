/*
 * elsa_hip.h -- C-ABI of libelsahip.so, the MI355X (gfx950) X-ray projectors.
 *
 * These entry points replace the static C++ launcher members that elsa's GPU projector
 * classes call (not extern "C" in elsa):
 *   TraverseJosephsCUDA<T,dim>::traverseForward / traverseAdjoint
 *       elsa/projectors_cuda/projector_kernels/TraverseJosephsCUDA.cuh:61-65, :89-93
 *   TraverseSiddonsCUDA<T,dim>::traverseForward / traverseAdjoint
 *       elsa/projectors_cuda/projector_kernels/TraverseSiddonsCUDA.cuh:60-89
 *   ProjectVoxelsCUDA<T,dim,adjoint,type>::project
 *       elsa/projectors_cuda/projector_kernels/ProjectVoxelsCUDA.cuh:49-52
 * together with the per-operator device state that the host classes set up in their
 * constructors (JosephsMethodCUDA.cpp:9-92, SiddonsMethodCUDA.cpp:13-59,
 * VoxelProjectorCUDA.cpp:16-54 with VoxelCUDAHelper.cuh:18-48).
 *
 * Conventions
 *  - Every call returns EHP_OK (0) on success or an EHP_ERR_* code; the message of the
 *    last failure on the calling thread is ehp_last_error().
 *  - Geometry is elsa's real_t = float, matrices column-major as Eigen stores them:
 *      P    dim x (dim+1)   projection matrix          Geometry::getProjectionMatrix()
 *      Pinv (dim+1) x dim   inverse projection matrix  Geometry::getInverseProjectionMatrix()
 *      C    dim             camera centre              Geometry::getCameraCenter()
 *      ext  dim x (dim+1)   extrinsic matrix           Geometry::getExtrinsicMatrix()
 *      sdd  1               source-detector distance   Geometry::getSourceDetectorDistance()
 *    one block per pose, poses contiguous.
 *  - Data layout is elsa's (DataDescriptor.cpp:25-28): volume x fastest, then y (then z);
 *    sinogram u fastest, then v (3D), then the pose index.
 *  - Data pointers are device pointers to the first element of the container (a view's
 *    begin pointer, never the parent storage). Outputs are overwritten, never accumulated.
 *  - Calls are asynchronous on the given HIP stream (hipStream_t passed as void*; NULL is
 *    the default stream). Inputs must stay valid until the stream reaches the call.
 *  - Calls on one operator are ordered by the library, whatever stream or host thread they
 *    come from: an operator owns device scratch that each call rewrites (volume views, the
 *    fixed-point scale of a back projection), so a call enqueues under the operator's mutex
 *    and its stream first waits for the event recorded after the operator's previous call
 *    when that call went to another stream. A fused batch pass is ordered the same way with
 *    the batch and every operator in it. Calls on different operators run concurrently.
 *    Inside a stream capture (hipGraph) the capture's own order applies; replays of a
 *    captured graph are ordered by the caller.
 *  - Host-only helpers (ehp_geometry_pose_*, ehp_*_lut) never touch the GPU.
 */
#ifndef ELSA_HIP_H
#define ELSA_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum ehp_status {
    EHP_OK = 0,
    EHP_ERR_ARGUMENT = 1,    /* elsa: InvalidArgumentError / std::invalid_argument */
    EHP_ERR_HIP = 2,         /* a HIP runtime call failed (elsa: std::logic_error) */
    EHP_ERR_ALLOC = 3,       /* device allocation failed (elsa: std::bad_alloc) */
    EHP_ERR_UNSUPPORTED = 4  /* valid request this build does not implement */
};

/* flags of ehp_joseph_create */
enum ehp_joseph_flags {
    EHP_SIMPLE_KERNELS = 1, /* one thread per ray (Joseph, Siddon) or per voxel and pose (voxel
                               projectors) in both directions, atomic scatters; the reference
                               point for the tiled kernels */
    EHP_DIFFERENTIAL = 2    /* ehp_voxel_create only: the phase-contrast (differential) weight
                               of elsa's PhaseContrast*VoxelProjectorCUDA, lut(|d|) * sgn(d_u)
                               in 2D and lut(|d|) * d_u in 3D (ProjectVoxelsCUDA.cu:73-76,
                               :160-165; VoxelComputation.h:26-38), with d_u the detector-u
                               component of the scaled distance; pass the derivative table in
                               2D and the normalized-gradient table in 3D */
};

enum ehp_dtype {
    EHP_F32 = 0,  /* data_t = float  */
    EHP_F64 = 1,  /* data_t = double */
    EHP_BF16 = 2  /* bfloat16 storage (raw bits, the upper half of an IEEE single) of volume and
                     sinogram, FP32 arithmetic: Joseph and the voxel projectors with the default
                     (fast) kernels only. No counterpart in elsa; the storage form of
                     mixed-precision reconstruction */
};

typedef struct ehp_geometry ehp_geometry; /* device-resident per-pose matrices      */
typedef struct ehp_op ehp_op;             /* one projector (Joseph / Siddon / voxel) */
typedef struct ehp_batch ehp_batch;       /* many independent projectors as one         */

/* Message of the last failing call on this thread ("" if none). */
const char* ehp_last_error(void);

/* Library version string. */
const char* ehp_version(void);

/* Upload the geometry of a trajectory (elsa: the ctor-time upload of JosephsMethodCUDA.cpp:
 * 33-91 and transferGeometries, VoxelCUDAHelper.cuh:18-48). vol_dims has dim entries,
 * det_dims dim-1 entries. All pointer arguments are host memory. */
int ehp_geometry_create(int dim, int64_t nposes, const int64_t* vol_dims, const int64_t* det_dims,
                        const float* P, const float* Pinv, const float* C, const float* ext,
                        const float* sdd, ehp_geometry** out);
void ehp_geometry_destroy(ehp_geometry* g);

/* A curved detector (elsa: CurvedDetectorDescriptor, CurvedDetectorDescriptor.cpp:97-119):
 * the planar detector coordinate of each pixel centre, det_dims[0] x det_dims[1] pixels (u
 * fastest), dim - 1 floats each (host memory; NULL returns to a planar detector). Operators
 * created from the geometry afterwards trace the ray of each pixel through that coordinate
 * instead of u + 0.5 (v + 0.5). Joseph and Siddon take it: their tiled back projections find
 * each tile's candidate pixels by binary search over the planar coordinates when these have a
 * curved detector's structure (planar u the same on every row and increasing along it, planar
 * v increasing down every column; otherwise one thread per ray). The voxel projectors, like
 * elsa's, take planar detectors only. elsa's GPU projectors ignore curved detectors; its CPU
 * projectors honour them, as these do. */
int ehp_geometry_set_pixel_coords(ehp_geometry* g, const float* coords);

/* Joseph's method (elsa: JosephsMethodCUDA(volume, detector, fast), JosephsMethodCUDA.h:78-79,
 * with the discretization of the CPU JosephsMethod, JosephsMethod.cpp:140-198). The adjoint
 * is always the matched one (elsa's fast=false); elsa's unmatched fast=true back projection
 * has no counterpart. flags: 0 or EHP_SIMPLE_KERNELS. */
int ehp_joseph_create(const ehp_geometry* g, int dtype, int flags, ehp_op** out);

/* Siddon's method (elsa: SiddonsMethodCUDA(volume, detector), SiddonsMethodCUDA.h:71-72,
 * traversal of TraverseAABB.cpp:6-184, crossing parameters T0 + k * tDelta with one rounding
 * each). FP32 and FP64 forward projections walk plane-major views of the volume and back
 * projections run LDS-tiled; flags: 0 or EHP_SIMPLE_KERNELS. */
int ehp_siddon_create(const ehp_geometry* g, int dtype, int flags, ehp_op** out);

/* Voxel-driven projector with a radially symmetric basis function sampled in a table of
 * n_lut values over [0, support] (elsa: BlobVoxelProjectorCUDA / BSplineVoxelProjectorCUDA,
 * VoxelProjectorCUDA.h:39-43, :82-84; the table is ProjectedBlob::get_lut() or
 * ProjectedBSpline::get_lut(), Blobs.h:249-317, BSplines.h:132-208). lut is host memory of
 * the op's dtype (FP32 for EHP_BF16 operators, which compute in FP32). Forward projections
 * (FP32, bfloat16, FP64) run on LDS-tiled detector tiles; flags: 0,
 * EHP_SIMPLE_KERNELS (one thread per voxel and pose, atomic splat) and EHP_DIFFERENTIAL (the
 * phase-contrast projectors PhaseContrastBlobVoxelProjectorCUDA /
 * PhaseContrastBSplineVoxelProjectorCUDA, VoxelProjectorCUDA.h:109-191). */
int ehp_voxel_create(const ehp_geometry* g, int dtype, const void* lut, int n_lut, double support, int flags,
                     ehp_op** out);

void ehp_destroy(ehp_op* op);

/* Forward projection Ax of the volume d_vol into d_sino (elsa: LinearOperator::apply). */
int ehp_forward(ehp_op* op, const void* d_vol, void* d_sino, void* stream);

/* Back projection A^T y of d_sino into d_vol (elsa: LinearOperator::applyAdjoint). */
int ehp_adjoint(ehp_op* op, const void* d_sino, void* d_vol, void* stream);

/* ehp_adjoint for callers that cannot allocate device memory (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc). No counterpart in elsa. The same
 * kernels and result, bit for bit, but the scratch of the back projection (the fixed-point
 * scale: per-block maxima of |y| and the exponent, 16.6 KB) lives in one static device array
 * of the library, 16.6 KB per GPU held until the process ends, instead of the operator's; the
 * call allocates nothing. Takes FP32 and FP64 Joseph and Siddon operators with a tiled back
 * projection (EHP_ERR_UNSUPPORTED otherwise: bfloat16, voxel or EHP_SIMPLE_KERNELS operators,
 * and geometries without a tiled back projection). Every call on a GPU shares that GPU's array: calls that overlap
 * on one GPU (different operators on different streams, or threads) race and must be ordered
 * by the caller; calls on one operator are ordered as for ehp_adjoint. */
int ehp_adjoint_noalloc(ehp_op* op, const void* d_sino, void* d_vol, void* stream);

/* Number of elements of the domain (volume) and range (sinogram). */
int ehp_op_sizes(const ehp_op* op, int64_t* n_vol, int64_t* n_sino);

/* Joseph: exact number of samples (ray x driving-plane crossings inside the volume) of one
 * pass, the unit of the roofline. Siddon: number of voxel crossings. Voxel projectors:
 * number of footprint taps. Computed on the device at create time. */
int ehp_sample_count(const ehp_op* op, uint64_t* samples);

/* Tile shape of the operator's LDS-tiled kernel (FP32). Joseph and Siddon back projections:
 * la planes along the driving axis, tb x tc voxels across (tc ignored in 2D). Voxel forward
 * projection: la x tb detector pixels (tb ignored in 2D, tc ignored). No counterpart in elsa;
 * it is the knob the tile autotuner turns. Fails with EHP_ERR_ARGUMENT if the tile does not
 * fit the device's LDS or the operator has no tiled kernel. Queries only when out pointers
 * are given and la <= 0. */
int ehp_tile(ehp_op* op, int la, int tb, int tc, int* la_out, int* tb_out, int* tc_out);

/* Poses per step of the tiled back projections (Joseph, Siddon; 1-32, default 8): each step
 * gathers the candidate rays of that many poses for one tile, so larger steps keep more lanes
 * busy where few rays meet a tile (2D). No counterpart in elsa; a knob of the tile autotuner.
 * Queries only when batch <= 0. */
int ehp_pose_batch(ehp_op* op, int batch, int* batch_out);

/* Per-ray Joseph traversal state as the kernels compute it (for parity checks): driving
 * axis and number of samples of every ray, host memory of n_sino entries each. */
int ehp_joseph_ray_info(const ehp_op* op, int32_t* axis, int64_t* nsteps);

/* ---- solvers on the device ---------------------------------------------------------------
 * The iterations of elsa's solvers with every scalar (step length, momentum, norms,
 * convergence) kept in device memory: norms and dot products are FP64 reductions in a fixed
 * order, vector updates run in the operator's data type. elsa's early return on convergence
 * becomes a device flag after which x no longer changes. With EHP_SOLVE_GRAPH one iteration is
 * captured into a hipGraph and replayed. Each call allocates its workspace, is ordered after
 * `stream` and returns when the solve has finished. x is overwritten with the result
 * (it is the start value when use_x0 != 0). */
enum ehp_solve_flags {
    EHP_SOLVE_GRAPH = 1,
    /* ehp_admml2 / ehp_batch_admml2 only: the regularisation operator A is elsa's
     * FiniteDifferences(domain) (forward differences on every axis) instead of the identity */
    EHP_ADMM_FINITE_DIFFERENCES = 2
};

/* CGLS on min ||Ax - b||^2 + eps^2 ||x||^2 (elsa: CGLS(A, b, eps, tol).solve(iterations, x0),
 * CGLS.cpp:19-91): stops before an iteration when ||A^T s - eps^2 x||^2 < tol. residuals
 * (host, iterations + 1 entries or NULL) receives ||s|| = ||b - Ax|| at the start and after
 * each iteration taken (-1 for iterations not taken); iterations_done the number taken. */
int ehp_cgls(ehp_op* op, const void* d_b, void* d_x, int use_x0, int iterations, double eps, double tol, int flags,
             double* residuals, int* iterations_done, void* stream);

/* APGD / FISTA on 0.5 ||Ax - b||^2 + g(x) with g = l1_sigma ||x||_1 (l1_sigma >= 0, elsa's
 * ProximalL1) or g = 0 (l1_sigma < 0, ProximalIdentity) (elsa: APGD(A, b, prox, mu, epsilon)
 * .solve(iterations, x0), APGD.cpp:14-102). mu <= 0 takes elsa's 0.45 / L with L from 5
 * power iterations on A^T A; mu_used receives the step. Stops after the iteration in which
 * ||grad||^2 <= epsilon; grad_sq_norms (host, iterations + 1 entries or NULL) receives
 * ||grad||^2 per iteration taken (-1 otherwise). */
int ehp_apgd(ehp_op* op, const void* d_b, void* d_x, int use_x0, int iterations, double mu, double l1_sigma,
             double epsilon, int flags, double* grad_sq_norms, int* iterations_done, double* mu_used, void* stream);

/* ADMM with an L2 data term (elsa: ADMML2(op, b, A, proxg, tau, ninneriters).solve(iterations,
 * x0), ADMML2.cpp:20-97) for the regularisation operator A = identity, or with
 * EHP_ADMM_FINITE_DIFFERENCES A = forward finite differences on every axis (elsa:
 * FiniteDifferences(domain), FiniteDifferences.cpp:106-206; the anisotropic TV problem with
 * ProximalL1). Per iteration the x-update runs ninner CGNE iterations (CGNE.cpp, via
 * RegularizedInversion.cpp) on [op; lambda A] with right-hand side [b; z - u] and
 * lambda = 1 / sqrt(tau), starting from the current x; then z = prox(A x + u) with threshold
 * tau * l1_sigma (ProximalL1; l1_sigma < 0: identity) and u += A x - z. z and u live in A's
 * range (dim * nvox values per problem for finite differences). tau <= 0 takes elsa's default
 * 0.9 / powerIterations(A^T A): 0.9 for the identity, 5 power iterations on the device for
 * finite differences (one problem only: a batch with finite differences needs tau > 0);
 * tau_used receives it. residuals (host, iterations + 1 or NULL) receives ||b - op x|| (FP64)
 * at the start and after each iteration; asking for it runs the iterations without graph
 * capture. */
int ehp_admml2(ehp_op* op, const void* d_b, void* d_x, int use_x0, int iterations, double tau, double l1_sigma,
               int ninner, int flags, double* residuals, double* tau_used, void* stream);

/* CGLS with iterative refinement in mixed precision (no counterpart in elsa): x (FP32) takes
 * `outer` steps x += d. Each d approximately minimises ||A d - r|| by `inner` CGLS iterations
 * (the ehp_cgls iteration, eps = 0) on op_lo, a bfloat16 operator; r = b - A x is the
 * residual of op_hi, the FP32 operator of the same geometry, formed in FP64 and rounded to
 * bfloat16 for the inner solve. d_b and d_x are FP32. residuals (host, outer + 1 entries or
 * NULL) receives ||b - A x|| (FP64) before the first step and after each. */
int ehp_cgls_refined(ehp_op* op_hi, ehp_op* op_lo, const void* d_b, void* d_x, int use_x0, int outer, int inner,
                     int flags, double* residuals, void* stream);

/* Largest eigenvalue estimate of A^T A by power iterations from the all-ones vector
 * (elsa: powerIterations(adjoint(A) * A, niters), PowerIterations.cpp:10-27). */
int ehp_power_iterations(ehp_op* op, int niters, double* L, void* stream);

/* ---- batches of independent problems -----------------------------------------------------
 * A batch is the block-diagonal operator diag(A_0, ..., A_{n-1}) of n operators of one data
 * type on one device, each with its own volume, trajectory and detector: its volume is the
 * operators' volumes packed back to back (problem j at element vol_offsets[j]), its sinogram
 * likewise. No counterpart in elsa, whose BlockLinearOperator (BlockLinearOperator.h:31-38)
 * stacks operators that share a domain (ROW) or a range (COL); applied to problem j's block, a
 * batch gives exactly what ehp_forward / ehp_adjoint of A_j give (bit for bit), and its
 * solvers run elsa's iteration on every problem independently, each with its own scalars and
 * its own convergence test. When all operators are FP32 Joseph operators, or all are FP32
 * Siddon operators, or all are FP32 voxel projectors (blob, B-spline, phase contrast), of one
 * dimension with the default (fast) kernels, every projector pass is one launch over all
 * problems per kernel (plus the view copies, the zero fill and the fixed-point scales);
 * otherwise the
 * batch calls the operators one after another. The operators must outlive the batch; tile
 * shapes are taken at creation; a batch and its operators share device state, and the library
 * orders calls on either (see Conventions). */
int ehp_batch_create(ehp_op* const* ops, int n, ehp_batch** out);
void ehp_batch_destroy(ehp_batch* b);

/* Element offsets of each problem in the packed volume and sinogram (n + 1 entries each, the
 * last one the total; host memory; either may be NULL). *fused is 1 when passes are single
 * launches (NULL allowed). */
int ehp_batch_layout(const ehp_batch* b, int64_t* vol_offsets, int64_t* sino_offsets, int* fused);

int ehp_batch_forward(ehp_batch* b, const void* d_vols, void* d_sinos, void* stream);
int ehp_batch_adjoint(ehp_batch* b, const void* d_sinos, void* d_vols, void* stream);

/* ehp_cgls / ehp_apgd / ehp_power_iterations / ehp_admml2 on every problem of the batch at once. Per-problem
 * outputs are host arrays of n rows: residuals and grad_sq_norms n x (iterations + 1),
 * iterations_done, mu_used and L n entries. mu: n step lengths (entries <= 0, or mu NULL:
 * 0.45 / L_j from power iterations on that problem). */
int ehp_batch_cgls(ehp_batch* b, const void* d_b, void* d_x, int use_x0, int iterations, double eps, double tol,
                   int flags, double* residuals, int* iterations_done, void* stream);
int ehp_batch_apgd(ehp_batch* b, const void* d_b, void* d_x, int use_x0, int iterations, const double* mu,
                   double l1_sigma, double epsilon, int flags, double* grad_sq_norms, int* iterations_done,
                   double* mu_used, void* stream);
int ehp_batch_power_iterations(ehp_batch* b, int niters, double* L, void* stream);
int ehp_batch_admml2(ehp_batch* b, const void* d_b, void* d_x, int use_x0, int iterations, double tau,
                     double l1_sigma, int ninner, int flags, double* residuals, double* tau_used, void* stream);

/* ---- host-only helpers (no GPU) -------------------------------------------------------- */

/* One pose of elsa's Geometry (elsa/core/Geometry.cpp:12-51 for dim 2, :53-109 for dim 3):
 * source-to-centre s2c, centre-to-detector c2d, volume spacing and origin (dim), detector
 * spacing and origin (dim-1), rotation angles in radians (dim 2: gamma only), principal point
 * offset (dim-1) and centre-of-rotation offset (dim). Writes P, Pinv, C, ext, sdd as above. */
int ehp_geometry_pose(int dim, float s2c, float c2d, const float* vol_spacing,
                      const float* vol_origin, const float* det_spacing, const float* det_origin,
                      float gamma, float beta, float alpha, const float* principal_offset,
                      const float* rotation_offset, float* P, float* Pinv, float* C, float* ext,
                      float* sdd);

/* Circular trajectory of elsa's CircleTrajectoryGenerator::createTrajectory
 * (CircleTrajectoryGenerator.cpp:9-24): nposes angles LinSpaced over [0, arc_degrees]
 * (both ends included), detector size det_dims (or floor(N*sqrt 2) per axis when det_dims is
 * NULL; the chosen size is written back to det_out), volume spacing 1. Writes nposes blocks. */
int ehp_circle_trajectory(int dim, int64_t nposes, const int64_t* vol_dims, float arc_degrees,
                          float s2c, float c2d, const int64_t* det_dims, int64_t* det_out,
                          float* P, float* Pinv, float* C, float* ext, float* sdd);

/* Planar coordinates of the pixel centres of a curved detector of det_dims pixels spanning
 * `angle` radians at source-detector distance s2d (CurvedDetectorDescriptor::setup,
 * CurvedDetectorDescriptor.cpp:36-95; getPlanarCoords()). Writes det_dims[0] x det_dims[1] x
 * (dim - 1) floats. */
int ehp_curved_detector_coords(int dim, const int64_t* det_dims, float angle, float s2d, float* out);

/* LUT of elsa's ProjectedBlob<data_t, n> (Blobs.h:249-317) and ProjectedBSpline<data_t, n>
 * (BSplines.h:132-208), in the given dtype; *support receives the table's support. */
int ehp_blob_lut(int dtype, double radius, double alpha, int64_t order, int n, void* out,
                 double* support);
int ehp_bspline_lut(int dtype, int64_t dim, int64_t order, int n, void* out, double* support);

/* Derivative tables of the phase-contrast projectors: ProjectedBlob::get_derivative_lut()
 * (normalized = 0) or get_normalized_gradient_lut() (normalized = 1) (Blobs.h:249-317 with
 * blob_derivative_projected :129-153 and blob_normalized_derivative_projected :183-207), and
 * ProjectedBSpline's (BSplines.h:132-208 with :40-45, :83-92). Order 0 gives zeros. */
int ehp_blob_derivative_lut(int dtype, double radius, double alpha, int64_t order, int n, int normalized, void* out,
                            double* support);
int ehp_bspline_derivative_lut(int dtype, int64_t dim, int64_t order, int n, int normalized, void* out,
                               double* support);

#ifdef __cplusplus
}
#endif

#endif /* ELSA_HIP_H */
