/*
 * dare_hip.h — C-ABI of the MI355X (gfx950) finite-volume hot path of DaRe.
 *
 * Plain pointers and sizes only; every device pointer is HBM memory of the context's
 * device, every call is asynchronous on the context's stream unless it returns host data.
 * All functions return 0 (DARE_OK) or a negative DARE_ERR_* code; dare_last_error() gives
 * a thread-local message. No C++ exception crosses this boundary.
 *
 * Each entry point names the reference interface it replaces (hpc-dave/dare v0.3.1,
 * file:line). The reference is a C++ template library, not an FFI; the C++ host mirror
 * in dare_amd/include/dare/ keeps its class names and call surface and calls these
 * functions.
 */
#ifndef DARE_HIP_H_
#define DARE_HIP_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define DARE_OK 0
#define DARE_ERR_INVALID (-1)      /* bad argument or shape */
#define DARE_ERR_OOM (-2)          /* device allocation failed */
#define DARE_ERR_HIP (-3)          /* HIP runtime error */
#define DARE_ERR_NOT_CONVERGED (-4)
#define DARE_ERR_BREAKDOWN (-5)    /* Krylov breakdown (rho, gamma, mu or tau == 0) */
#define DARE_ERR_UNSUPPORTED (-6)

/* Limiter ids. Equations/FluxLimiter.h:41,73,105,145 (+ van Leer, closed form). */
#define DARE_LIMITER_CDS 0
#define DARE_LIMITER_UPWIND 1
#define DARE_LIMITER_VANALBADA 2
#define DARE_LIMITER_MINMOD 3
#define DARE_LIMITER_VANLEER 4

/* Ghost fill modes per boundary face (W,E,S,N,B,T). */
#define DARE_GHOST_LEAVE 0
#define DARE_GHOST_PERIODIC 1   /* HaloBuffer self-exchange of a periodic single-rank grid */
#define DARE_GHOST_DIRICHLET 2  /* 2 v - phi(mirror) */
#define DARE_GHOST_NEUMANN 3    /* phi(mirror) */
#define DARE_GHOST_CONSTANT 4   /* v */

/* Boundary edits applied during assembly (counter_diffusion2D.cpp:125-161 pattern). */
#define DARE_BC_NONE 0
#define DARE_BC_DIRICHLET 1
#define DARE_BC_NEUMANN 2
/* the boundary cell itself carries the value: identity row (centre 1, rhs = x0 = value, no
 * neighbours) -- the Dirichlet condition of a staggered axis, whose first and last cells sit
 * on the boundary faces (the functor of test_integration_Cartesian_Trilinos.cpp:155-220) */
#define DARE_BC_FIXED 3

/* Right preconditioners of the solver (dare_solver_params.precond). */
#define DARE_PRECOND_NONE 0
#define DARE_PRECOND_JACOBI 1   /* M = diag(A): Ifpack2 "RELAXATION", Jacobi, one sweep, damping 1 */
/* aggregation multigrid V-cycle on the diagonal view (one-component 3-D operators): the
 * stand-in for MueLu "AMG" (MatrixSystem/TrilinosSolver.inl:180-246) */
#define DARE_PRECOND_MG 2
/* Jacobi in BiCGStab2's left slot (BiCGStab2.inl:276-280 and the M^-1 A applications of
 * steps 6, 9, 13): the method runs on diag(A)^-1 A x = diag(A)^-1 b, its residual and stopping
 * test are the preconditioned ones; FP64 solve only */
#define DARE_PRECOND_JACOBI_LEFT 3

typedef struct dare_ctx dare_ctx;
typedef struct dare_csr dare_csr;

/* Grid of one single-rank subdomain. Replaces Grid::Cartesian<Dim> + CartesianRepresentation
 * (Grid/Cartesian/CartesianMesh.inl:35-112, CartesianRepresentation.inl:32-112). Storage of a
 * field: data[local_ordinal * ncomp + c], local ordinal over the ghost-padded box, last axis
 * fastest (Data/GridVector.inl:37-41,69-76). */
typedef struct {
    int32_t dim;          /* 1, 2 or 3 */
    int32_t ncomp;        /* components per cell (N), 1..8 */
    int32_t ghost;        /* ghost layers per side */
    int32_t n[3];         /* internal resolution; n[d] = 1 for d >= dim */
    int32_t periodic[3];  /* non-zero: periodic along axis d */
    double size[3];       /* domain lengths */
} dare_grid;

/* Terms of the per-cell assembly functor, evaluated on the device. Replaces the user lambda
 * handed to Trilinos<SC>::Build (MatrixSystem/Trilinos.inl:50-60,308-328) for the operator
 *   mb = ddt(phi); mb -= div(D_f * grad(mb)); mb += div(tvd * phi); rhs += V s; reaction; BCs.
 * Pointers are device pointers (or NULL). */
typedef struct {
    double dt;               /* > 0: implicit transient term (Equations/DDT.inl:41-81) */
    double diff_const;       /* constant diffusivity, used when diff_cell == NULL */
    const double* diff_cell; /* per-cell diffusivity, padded layout, face value = harmonic mean */
    double diff_comp[8];     /* per-component constant diffusivity when use_diff_comp != 0 */
    int32_t use_diff_comp;
    int32_t conv_limiter;    /* < 0: none, else implicit TVD convection (TVD_Cartesian.inl:185-261) */
    const double* vel[3];    /* face velocities (face arrays, 1 component) */
    const double* source;    /* per-cell source (padded layout), rhs += V * s */
    const double* react;     /* ncomp == 2: per-cell 2x2 block [a b; c d], internal order */
    int32_t bc_type[6];      /* W,E,S,N,B,T: DARE_BC_* */
    double bc_value[6];
    /* Time scheme of the diffusion and convection divergences, Divergence<Grid, TimeScheme>
     * (Equations/TimeDiscretizationSchemes.h:29-76): div_levels = 0 (or 1) is EULER_BACKWARD
     * (weights {1}); 2 or 3 levels take div_weights[0..levels-1] (EULER_FORWARD {0, 1},
     * CRANK_NICHOLSON {0.5, 0.5}, ADAMS_BASHFORT {0, 1.5, -0.5}, ADAMS_MOULTON {5/12, 8/12,
     * -1/12}). As in the reference's matrix form (Divergence_Cartesian.inl:58-66,141-207) the
     * explicit levels reuse the implicit level's face stencil: the coefficients are scaled by
     * w0 and the face right-hand sides (the TVD deferred correction) enter once unweighted and
     * once per explicit level t with weight w_t. */
    double div_weights[3];
    int32_t div_levels;
} dare_terms;

/* Solver parameters. Replaces the Teuchos::ParameterList keys "Convergence Tolerance",
 * "Maximum Iterations" of BiCGStab2 (MatrixSystem/BiCGStab2.inl:95-111). */
typedef struct {
    double tol;              /* convergence tolerance */
    int32_t tol_kind;        /* 0: ||r|| <= tol (reference RNORM), 1: ||r|| <= tol * ||b|| */
    int32_t max_it;          /* maximum BiCGStab2 passes */
    double inner_tol;        /* mixed precision: inner FP32 relative reduction per refinement;
                              * an inner solve stops at max(inner_tol * ||r||, tol_abs / 4), so the
                              * last one reduces only as far as the outer threshold needs */
    int32_t max_refine;      /* mixed precision: maximum refinement steps */
    int32_t inner_max_it;    /* mixed precision: passes per inner solve */
    double* history;         /* optional host array, residual norm after every pass */
    int32_t history_cap;
    int32_t precond;         /* 0: none (the reference default, M == null); 1: Jacobi, the right
                              * preconditioner M = diag(A) (BiCGStab2.inl:353-365 applies M^-1 on
                              * the right); the diagonal-view operator is column-scaled, a
                              * generic CSR matrix gets a scaled copy of its values; 2:
                              * multigrid V-cycle M^-1 applied before every operator
                              * application (DARE_PRECOND_MG); 3: Jacobi on the left
                              * (DARE_PRECOND_JACOBI_LEFT, dare_bicgstab2_solve only) */
} dare_solver_params;

typedef struct {
    int32_t converged;       /* 1 if the stopping test held */
    int32_t iterations;      /* BiCGStab2 passes (BiCGStab2.inl:578) — total over refinements */
    int32_t status;          /* 0 or DARE_ERR_BREAKDOWN */
    int32_t refinements;     /* mixed precision: outer steps */
    double residual;         /* last recursive residual norm (FP64 true residual for mixed) */
    double bnorm;            /* ||b|| */
    int32_t history_len;
} dare_solver_result;

/* ---- context, memory (replaces ScopeGuard/ExecutionManager set-up, ScopeGuard.cpp:30-151) */
int dare_ctx_create(int32_t device, dare_ctx** out);
int dare_ctx_destroy(dare_ctx* ctx);
int dare_ctx_set_stream(dare_ctx* ctx, void* hip_stream);   /* NULL: the null stream */
void* dare_ctx_get_stream(dare_ctx* ctx);
int dare_ctx_synchronize(dare_ctx* ctx);
const char* dare_last_error(void);
int32_t dare_abi_version(void);

int dare_malloc(dare_ctx* ctx, size_t bytes, void** out);
int dare_free(dare_ctx* ctx, void* p);
int dare_memcpy_h2d(dare_ctx* ctx, void* dst, const void* src, size_t bytes);
int dare_memcpy_d2h(dare_ctx* ctx, void* dst, const void* src, size_t bytes);
int dare_memcpy_d2d(dare_ctx* ctx, void* dst, const void* src, size_t bytes);
int dare_fill_f64(dare_ctx* ctx, double* p, int64_t n, double v);

/* Bandwidth probe: dst[j] = sum of `ratio` doubles of src, src holding n_dst * ratio doubles,
 * every access a coalesced non-temporal 16-byte one (ratio 1, 2, 4 or 8; n_dst a multiple of 512;
 * 16-byte aligned). ratio 1 is a copy; ratio 8 has the read-heavy mix of the solver's SpMV. The
 * attainable HBM bandwidth the measurements quote beside each roofline. No reference
 * counterpart. grid <= 0 picks eight work-groups per CU; layout 0 deals each XCD one contiguous
 * slab (as the solver's kernels do), 1 strides the whole grid over the vectors, 2 is layout 0
 * with src read as `ratio` separate arrays of n_dst doubles (the slot-major diagonal view), 3 is
 * layout 2 read-only: each lane sums all it reads and writes the total once, as two doubles at
 * dst[2 (256 b + t)] for lane t of work-group b (dst then needs 512 doubles per work-group). */
int dare_stream_probe(dare_ctx* ctx, int64_t n_dst, const double* src, double* dst, int32_t ratio, int32_t grid,
                      int32_t layout);

/* ---- stencils (device pointers, FP64) ------------------------------------------------- */
/* Ghost fill. Replaces Field::ExchangeHaloCells -> HaloBuffer::Exchange on one rank
 * (Data/Field.inl:101-103, MPI/HaloBuffer.inl:196-208, SingleHaloBuffer.inl:95-127). */
int dare_ghost_fill(dare_ctx* ctx, const dare_grid* g, const int32_t* mode6,
                    const double* value6, double* phi);
/* Face gradient of every face. Replaces Gradient::operator()(GridVector)
 * (Grid/Cartesian/Gradient_Cartesian.inl:62-96) evaluated over the field. */
int dare_grad_faces(dare_ctx* ctx, const dare_grid* g, const double* phi,
                    double* fx, double* fy, double* fz);
/* Value divergence of face arrays. Replaces Divergence::ApplyDivergence(FaceValueStencil)
 * (Grid/Cartesian/Divergence_Cartesian.inl:211-235). Writes internal cells of out. */
int dare_div_faces(dare_ctx* ctx, const dare_grid* g, const double* fx, const double* fy,
                   const double* fz, double* out);
/* Face velocities of the TVD operator from cell-centred velocity fields (padded layout, one
 * component each, ghosts filled by the caller): face m along axis d takes (u[m-1] + u[m]) * 0.5.
 * Replaces the cell-centred source of TVD(grid, ordinal, Vector<Dim, const GridVector*>)
 * (TVD_Cartesian.inl:27-38) -> math::InterpolateToFace (Grid/Cartesian/Interpolation_Cartesian.h:
 * 203-219,298-354). The face arrays feed dare_tvd_* and dare_terms.vel; a constant velocity
 * (TVD(grid, ordinal, Vector<Dim, SC>), TVD_Cartesian.inl:41-53) is face arrays filled with its
 * components (dare_fill_f64). g->ncomp is ignored. */
int dare_interpolate_to_faces(dare_ctx* ctx, const dare_grid* g, const double* ux, const double* uy,
                              const double* uz, double* fx, double* fy, double* fz);
/* TVD face flux phi_f u_f. Replaces TVD::Interpolate(GridVector) times the face velocity
 * (TVD_Cartesian.inl:27-38,122-183; Stencils_Cartesian.inl:652-681). */
int dare_tvd_face_flux(dare_ctx* ctx, const dare_grid* g, int32_t limiter, const double* phi,
                       const double* ux, const double* uy, const double* uz,
                       double* fx, double* fy, double* fz);
/* Fused explicit TVD step phi_out = phi - (dt/V) div(phi_f u_f) on internal cells. */
int dare_tvd_explicit_step(dare_ctx* ctx, const dare_grid* g, int32_t limiter,
                           const double* phi, const double* ux, const double* uy,
                           const double* uz, double dt, double* phi_out);

/* ---- matrix assembly ------------------------------------------------------------------- */
int dare_csr_create(dare_ctx* ctx, dare_csr** out);
int dare_csr_destroy(dare_csr* csr);
/* Row pointers and sorted columns of the per-cell operator. Replaces the structure part of
 * Trilinos<SC>::BuildNew + MatrixBlock::Finalize + fillComplete (Trilinos.inl:287-385,
 * MatrixBlock_Cartesian.inl:128-240). */
int dare_csr_build_structure(dare_ctx* ctx, dare_csr* csr, const dare_grid* g,
                             const dare_terms* t);
/* Values, right-hand side and initial guess (x0 = phi_old). Replaces the functor loop of
 * Trilinos.inl:308-328 and the value half of :372-383 (BuildReplace semantics, :389-457). */
int dare_csr_fill_values(dare_ctx* ctx, dare_csr* csr, const dare_grid* g,
                         const dare_terms* t, const double* phi_old, double* rhs, double* x0);
/* Host-assembled CSR (generic functor path). Replaces insertGlobalValues/fillComplete. */
int dare_csr_upload_host(dare_ctx* ctx, dare_csr* csr, int64_t nrows, int64_t nnz,
                         const int32_t* row_ptr, const int32_t* col, const double* val);
int dare_csr_info(const dare_csr* csr, int64_t* nrows, int64_t* nnz, const int32_t** row_ptr,
                  const int32_t** col, const double** val);
int dare_csr_download(dare_ctx* ctx, const dare_csr* csr, int32_t* row_ptr, int32_t* col,
                      double* val);
/* Solution back to the field. Replaces Trilinos::CopyTo (Trilinos.inl:168-177). */
int dare_copy_to_field(dare_ctx* ctx, const dare_grid* g, const double* x, double* phi);

/* ---- Krylov primitives (the Tpetra calls of BiCGStab2.inl: apply, dot, norm2, update) -- */
int dare_spmv(dare_ctx* ctx, const dare_csr* csr, const double* x, double* y);
int dare_dot(dare_ctx* ctx, int64_t n, const double* a, const double* b, double* out_host);
/* Fused y = A x and <y, w> (the SpMV+dot launch of every BiCGStab2 step 6/9/13). out_host may
 * be NULL (no readback, fully asynchronous). */
int dare_spmv_dot(dare_ctx* ctx, const dare_csr* csr, const double* x, double* y, const double* w,
                  double* out_host);
int dare_axpby(dare_ctx* ctx, int64_t n, double a, const double* x, double b, double* y);
/* r = b - A x (the operator the solver uses) and, if norm_host != NULL, ||r|| (BiCGStab2.inl:265-266). */
int dare_residual(dare_ctx* ctx, const dare_csr* csr, const double* b, const double* x, double* r, double* norm_host);
/* *out = 1 when the FP64 solves, dare_residual, dare_spmv_dot (and dare_spmv without a
 * periodic axis) apply A through the symmetric traversal of its diagonal view (E, N, T of each
 * row read as W, S, B of its neighbours, with results bit-identical to the seven-slot one: the
 * view is exactly symmetric, by construction of the last fill or by the device's slot-by-slot
 * check), 0 otherwise (no diagonal view, two components, an odd k extent, convection, a failed
 * check). No counterpart in the reference; waits for the context stream. */
int dare_csr_symmetric(dare_ctx* ctx, const dare_csr* csr, int32_t* out);
/* enable = 0: every FP64 and FP32 traversal of this matrix reads its seven slots (the results
 * are the same bits either way); 1 (the default): symmetric views take the four-slot
 * traversal. No counterpart in the reference. */
int dare_csr_set_symmetric_traversal(dare_csr* csr, int32_t enable);

/* ---- solvers ------------------------------------------------------------------------- */
/* FP64 BiCGStab(2). Replaces TrilinosSolver::Solve(BumbleBee, "BICGSTAB2", ...) ->
 * BiCGStab2::Solve (TrilinosSolver.inl:71-93,150-152; BiCGStab2.inl:167-582). x holds x0. */
int dare_bicgstab2_solve(dare_ctx* ctx, const dare_csr* csr, const double* b, double* x,
                         const dare_solver_params* p, dare_solver_result* res);
/* Allocation-free FP64 BiCGStab(2): the same method, kernels and results as
 * dare_bicgstab2_solve, with its work vectors, reduction partials and scalars in a static
 * device array (15.2 MB per device, held until the process ends) instead of context memory.
 * No counterpart in the reference: for callers that cannot allocate device memory (tight
 * serving loops, harnesses that forbid hipMalloc). It calls no hipMalloc / hipFree.
 * Limits: at most 262144 rows (a 64^3 one-component grid; larger systems return
 * DARE_ERR_INVALID), no preconditioner (precond != 0 returns DARE_ERR_UNSUPPORTED), and the
 * matrix must already hold its values (dare_csr_fill_values or dare_csr_upload_host). Every
 * call on one device shares that device's array: calls that overlap on one GPU race. */
int dare_bicgstab2_solve_noalloc(dare_ctx* ctx, const dare_csr* csr, const double* b, double* x,
                                 const dare_solver_params* p, dare_solver_result* res);
/* FP32 inner BiCGStab(2) inside FP64 iterative refinement; r = b - A x, the update of x and
 * the stopping test in FP64. No reference counterpart (north-star precision form). */
int dare_refine_solve_mixed(dare_ctx* ctx, const dare_csr* csr, const double* b, double* x,
                            const dare_solver_params* p, dare_solver_result* res);

/* ---- batched solves ---------------------------------------------------------------------
 * Many independent systems in one set of persistent launches. The reference solves one
 * system per TrilinosSolver::Solve call (TrilinosSolver.inl:71-93); a batch is the same
 * BiCGStab2 (BiCGStab2.inl:167-582) run for every system with a per-system convergence mask.
 * Systems are dare_csr handles built by dare_csr_build_structure on 3-D grids (all with the
 * same number of components, 1 or 2); their values come from dare_csr_fill_values. The batch
 * owns the concatenated right-hand sides and solutions: system s occupies rows
 * [row_offsets[s], row_offsets[s+1]) of b and x (fill them, e.g. by passing b + offset and
 * x + offset to dare_csr_fill_values). */
typedef struct dare_batch dare_batch;
typedef struct {
    int32_t tile_k;   /* lanes along k per work-group: 16, 32, 64, or 0 = by each system's nz */
    int32_t tile_i;   /* planes per tile */
    int32_t grid;     /* work-groups of the persistent launches (multiple of 8) */
    int32_t order;    /* work list: 0 system-major, 1 XCD slabs, 2 XCD slabs largest first */
} dare_batch_config;

int dare_batch_create(dare_ctx* ctx, int32_t nsys, dare_csr* const* systems, dare_batch** out);
int dare_batch_destroy(dare_batch* batch);
int dare_batch_vectors(dare_batch* batch, double** b, double** x, int64_t* row_offsets /* nsys + 1 */);
int dare_batch_set_config(dare_ctx* ctx, dare_batch* batch, const dare_batch_config* cfg);
int dare_batch_get_config(const dare_batch* batch, dare_batch_config* cfg);
/* Times every candidate configuration on this batch (FP32 operator sweep, `reps` launches
 * each), keeps the fastest and returns it with its time per sweep in microseconds. */
int dare_batch_autotune(dare_ctx* ctx, dare_batch* batch, int32_t reps, dare_batch_config* best, double* best_us);
/* Solves every system. mixed = 0: FP64 BiCGStab2 (tol, tol_kind, max_it as in
 * dare_bicgstab2_solve); mixed = 1: FP32 inner solves inside FP64 refinement (as in
 * dare_refine_solve_mixed). graph_passes > 0 captures that many passes in a HIP graph that is
 * replayed until every system has stopped; 0 launches pass by pass. results[nsys]. */
int dare_batch_solve(dare_ctx* ctx, dare_batch* batch, const dare_solver_params* p, int32_t mixed,
                     int32_t graph_passes, dare_solver_result* results);

/* ---- learned warm start ------------------------------------------------------------------
 * A 3-level 3-D convolutional encoder-decoder (16/32/64 channels, 3x3x3 kernels, BF16 weights
 * and activations, FP32 accumulation on the matrix cores) that predicts the solution of the
 * next implicit step from phi^n, phi^{n-1} and the right-hand side. No reference counterpart:
 * the reference starts BiCGStab2 from phi^n (MatrixSystem/Trilinos.inl:316-317); the guess
 * only replaces x0, the solver still iterates to its FP64 tolerance.
 * Sequence: pack -> fwd -> unpack (inference); pack -> fwd -> loss -> bwd -> step (training).
 * Grids: 3-D, one component, every size a multiple of 4 and at most max_dims. */
typedef struct dare_convnet dare_convnet;
int dare_convnet_create(dare_ctx* ctx, int32_t in_channels, const int32_t* max_dims, dare_convnet** out);
int dare_convnet_destroy(dare_convnet* net);
/* number of FP32 master parameters; layer_offsets[12] (may be NULL). Layer l holds
 * [27][cout][cin] (tap = (dz+1)*9 + (dy+1)*3 + (dx+1), dx along the fastest grid index). */
int dare_convnet_num_params(const dare_convnet* net, int64_t* n, int64_t* layer_offsets);
int dare_convnet_set_params(dare_ctx* ctx, dare_convnet* net, const float* host);
int dare_convnet_get_params(dare_ctx* ctx, const dare_convnet* net, float* host_w, float* host_grad);
/* crop: NULL for the whole grid, else {origin[3], size[3]} of a sub-box of internal cells
 * (training on crops; the network's zero padding then applies at the crop's faces). */
/* Inputs: phi^n, phi^{n-1} (padded fields) and resid = b - A x_base in solver row order, the
 * residual of a classical guess x_base (phi^n, or 2 phi^n - phi^{n-1}), which resid_scale
 * (dt / V for the implicit Euler term) maps to field units. Channels: phi^n / max|phi^n|,
 * (phi^n - phi^{n-1}) / max|.|, resid * resid_scale / S with S = max|resid * resid_scale|,
 * then the extras. The guess (unpack) is x_base + S * out; the loss compares out with
 * (x_true - x_base) / S. The input tensor has 8 BF16 channels (at most 5 extras); layer 0 maps
 * them to 16. */
int dare_convnet_pack(dare_ctx* ctx, dare_convnet* net, const dare_grid* g, const double* phi, const double* phi_prev,
                      const double* resid, double resid_scale, int32_t nextra, const double* const* extras,
                      const double* extra_scale, const double* extra_offset, const int32_t* crop);
int dare_conv_net_fwd(dare_ctx* ctx, dare_convnet* net);
/* Inference mode (off after create): the forward pass does not store the decoder's ReLU
 * outputs a7 and a9, which only the backward pass reads (at 256^3 that is 0.67 GB less HBM
 * traffic per forward pass); every other tensor and the output are unchanged. While the last
 * forward pass ran in this mode dare_conv_net_bwd refuses with DARE_ERR_INVALID. */
int dare_convnet_set_inference(dare_convnet* net, int32_t on);
/* Coarsening (1 after create): with factor 2 the network runs at half resolution. pack feeds
 * each network voxel the mean of the channel values of its 2x2x2 cells; unpack adds S times
 * the trilinear interpolation of the output at every cell centre (weights 3/4, 1/4 of the two
 * nearest voxels per axis, the nearest repeated at the box faces); loss compares the output
 * with the 2x2x2 mean of (x_true - x_base) / S. Boxes are then multiples of 8 whose halves fit
 * max_dims, which counts network voxels. No reference counterpart. */
int dare_convnet_set_coarsening(dare_convnet* net, int32_t factor);
/* pack -> fwd (inference mode) -> unpack of the whole grid in one call: the three steps'
 * kernels are captured as a HIP graph at the first call (and again whenever any argument
 * differs from the captured call's) and replayed on the context's stream; the results equal
 * those of the three separate calls. Not for use while the context's stream is itself being
 * captured. No reference counterpart. */
int dare_convnet_predict(dare_ctx* ctx, dare_convnet* net, const dare_grid* g, const double* phi,
                         const double* phi_prev, const double* resid, double resid_scale, int32_t nextra,
                         const double* const* extras, const double* extra_scale, const double* extra_offset,
                         const double* x_base, double* x_guess);
int dare_convnet_unpack(dare_ctx* ctx, dare_convnet* net, const dare_grid* g, const double* x_base, double* x_guess,
                        const int32_t* crop);
int dare_convnet_loss(dare_ctx* ctx, dare_convnet* net, const dare_grid* g, const double* x_base, const double* x_true,
                      double* loss_host, const int32_t* crop);
int dare_conv_net_bwd(dare_ctx* ctx, dare_convnet* net);
int dare_conv_net_step(dare_ctx* ctx, dare_convnet* net, float lr, float beta1, float beta2, float eps);
/* Device pointers of x0, out, dout and the activations (a1 e1 p1 a3 e2 p2 a5 e3 u2 a7 s2 d2 u1
 * a9 s1 d1) and the three input scales. The upsampled u2, u1 are only read inside the decoder
 * convs by the forward pass; the backward pass writes them. a7 and a9 keep the values of the
 * last forward pass outside inference mode. */
int dare_convnet_tensors(const dare_convnet* net, void** ptrs, double* scales_host);

#ifdef __cplusplus
}
#endif

#endif  // DARE_HIP_H_
