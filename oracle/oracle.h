/*
 * CPU oracle for the DaRe finite-volume hot path — TEST INFRASTRUCTURE ONLY.
 *
 * This is a plain C++ restatement of the reference algorithm (hpc-dave/dare v0.3.1),
 * used to check the HIP path and to time a CPU baseline. Nothing in the shipped
 * library links, loads or calls it; only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py do.
 *
 * Every function cites the reference file:line it restates. Parity status per function
 * (what pins it) is listed in oracle.cpp next to each restatement.
 */
#ifndef DARE_ORACLE_H_
#define DARE_ORACLE_H_

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Grid descriptor, same memory layout as dare_grid in include/dare_hip.h. */
typedef struct {
    int32_t dim;          /* 1, 2 or 3 */
    int32_t ncomp;        /* components per cell (N) */
    int32_t ghost;        /* ghost layers per side (g) */
    int32_t n[3];         /* internal resolution; n[d] = 1 for d >= dim */
    int32_t periodic[3];  /* non-zero: periodic along axis */
    double size[3];       /* domain lengths */
} orc_grid;

/* Assembly terms, same layout as dare_terms in include/dare_hip.h. */
typedef struct {
    double dt;               /* > 0: implicit transient term (DDT, IMPLICIT scheme) */
    double diff_const;       /* constant diffusivity used when diff_cell is NULL */
    const double* diff_cell; /* per-cell diffusivity (padded layout, N comps): harmonic face mean */
    double diff_comp[8];     /* per-component constant diffusivity when use_diff_comp != 0 */
    int32_t use_diff_comp;
    int32_t conv_limiter;    /* < 0: no convection; else limiter id for implicit TVD */
    const double* vel[3];    /* face velocities (face arrays, 1 comp) for convection */
    const double* source;    /* per-cell source (padded layout, N comps), adds V*s to rhs */
    const double* react;     /* per-cell 2x2 reaction block (internal order, 4 doubles) or NULL */
    int32_t bc_type[6];      /* W,E,S,N,B,T: 0 none/periodic, 1 Dirichlet mirror, 2 Neumann */
    double bc_value[6];
    double div_weights[3];   /* Divergence<Grid, TimeScheme> weights of the diffusion and */
    int32_t div_levels;      /* convection terms; 0: EULER_BACKWARD {1} */
} orc_terms;

enum { ORC_CDS = 0, ORC_UPWIND = 1, ORC_VANALBADA = 2, ORC_MINMOD = 3, ORC_VANLEER = 4 };

/* Limiter value Psi(r). FluxLimiter.h:41-173 (+ van Leer closed form). */
double orc_limiter(int32_t id, double r);
/* Diffusion_N0D1 series. AnalyticalSolutions/Diffusion.h:46-58. */
double orc_diffusion_n0d1(double tau, double zeta, int32_t n_modes);

/* Index maps. CartesianRepresentation.inl:315-333, 373-450. */
int64_t orc_num_cells_padded(const orc_grid* g);
int64_t orc_num_cells_internal(const orc_grid* g);
int64_t orc_map_internal_to_local(const orc_grid* g, int64_t n_internal);
int64_t orc_num_faces(const orc_grid* g, int32_t axis);
/* Cell centre coordinates of a padded index. CartesianRepresentation.inl:230-238. */
void orc_cell_center(const orc_grid* g, const int32_t* ind, double* pos);

/* Field-wide face gradient: face arrays fx/fy/fz (N comps). Gradient_Cartesian.inl:62-96. */
int orc_grad_faces(const orc_grid* g, const double* phi, double* fx, double* fy, double* fz);
/* Value divergence of face arrays into padded cell layout. Divergence_Cartesian.inl:211-235. */
int orc_div_faces(const orc_grid* g, const double* fx, const double* fy, const double* fz, double* out);
/* Cell-centred velocity fields (padded, 1 component each) -> face arrays, the cell-centred
 * source of the TVD constructor (TVD_Cartesian.inl:27-38 -> InterpolateToFace,
 * Interpolation_Cartesian.h:298-354): face value (u[m-1] + u[m]) * 0.5. */
int orc_interpolate_to_faces(const orc_grid* g, const double* ux, const double* uy, const double* uz,
                             double* fx, double* fy, double* fz);
/* TVD face flux phi_f * u_f per face. TVD_Cartesian.inl:27-38,122-183 + Stencils_Cartesian.inl:652-681. */
int orc_tvd_face_flux(const orc_grid* g, int32_t limiter, const double* phi,
                      const double* ux, const double* uy, const double* uz,
                      double* fx, double* fy, double* fz);
/* Per-cell TVD matrix stencil (nb[6], c[6], rhs[6] for component comp) of internal cell
 * cell_int. TVD::operator*(GridVector), TVD_Cartesian.inl:185-261. */
int orc_tvd_matrix_cell(const orc_grid* g, int32_t limiter, const double* phi,
                        const double* ux, const double* uy, const double* uz,
                        int64_t cell_int, int32_t comp, double* nb6, double* c6, double* rhs6);
/* Explicit TVD advection step phi_out = phi - (dt/V) div(phi_f u_f) on internal cells. */
int orc_tvd_explicit_step(const orc_grid* g, int32_t limiter, const double* phi,
                          const double* ux, const double* uy, const double* uz,
                          double dt, double* phi_out);
/* Ghost fill. mode[6] per boundary face: 0 leave, 1 periodic (HaloBuffer.inl:32-208),
 * 2 Dirichlet mirror 2v-phi, 3 zero-gradient mirror, 4 constant v. */
int orc_ghost_fill(const orc_grid* g, const int32_t* mode, const double* value, double* phi);

/* Per-cell MatrixBlock assembly -> sorted CSR. Trilinos.inl:287-385,
 * MatrixBlock_Cartesian.inl:128-240, MatrixBlockBase.inl:229-295.
 * row_ptr[nrows+1], col[cap], val[cap], rhs[nrows], x0[nrows]. Returns nnz or < 0. */
int64_t orc_assemble(const orc_grid* g, const orc_terms* t, const double* phi_old,
                     int32_t* row_ptr, int32_t* col, double* val, int64_t cap,
                     double* rhs, double* x0);

/* y = A x for CSR. */
void orc_spmv(int64_t nrows, const int32_t* row_ptr, const int32_t* col, const double* val,
              const double* x, double* y);
double orc_dot(int64_t n, const double* a, const double* b);

/* BiCGStab(2). BiCGStab2.inl:167-582 (RNORM absolute test on the recursive residual).
 * hist (optional, length hist_cap) receives the residual norm after every pass.
 * Returns 1 converged, 0 not converged; iters = passes, resid = last residual norm. */
int orc_bicgstab2(int64_t nrows, const int32_t* row_ptr, const int32_t* col, const double* val,
                  const double* b, double* x, double tol, int32_t max_it,
                  int32_t* iters, double* resid, double* hist, int32_t hist_cap);

/* Copy CSR-ordered solution back into the padded field. Trilinos.inl:168-177. */
void orc_copy_to_field(const orc_grid* g, const double* x, double* phi);

/* Threads used by the OpenMP loops of the oracle (1 = scalar). */
void orc_set_threads(int32_t n);
int32_t orc_get_threads(void);

#ifdef __cplusplus
}
#endif

#endif  // DARE_ORACLE_H_
