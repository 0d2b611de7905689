/*
 * shearwave.h - C ABI of the MI355X (gfx950) shear_fdtd hot path.
 *
 * Every entry point is extern "C", takes plain device pointers (caller-owned,
 * SoA: one contiguous C-order (nx, ny, nz) array per component, z fastest),
 * sizes and an explicit hipStream_t passed as void*, and returns an int status
 * (SW_OK = 0). On failure, sw_last_error() returns a thread-local message.
 * No C++ exception crosses the boundary. The library allocates no device
 * memory: scratch space is passed in as a workspace.
 *
 * The reference (pinton-lab/shearwave) is pure Python; each declaration names
 * the Python function it replaces as file:line in src/shearwave/.
 */
#ifndef SHEARWAVE_H
#define SHEARWAVE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SW_ABI_VERSION 5

/* status codes */
#define SW_OK 0
#define SW_ERR_ARG 1          /* invalid argument (shape, pointer, option) */
#define SW_ERR_HIP 2          /* HIP runtime error */
#define SW_ERR_WORKSPACE 3    /* workspace too small */
#define SW_ERR_UNSUPPORTED 4  /* option combination not supported */

/* storage types */
#define SW_F32 0
#define SW_F64 1

/* time-loop state forms */
#define SW_FORM_INCREMENT 0 /* stores u_n and d_n = u_n - u_{n-1} */
#define SW_FORM_LEAPFROG 1  /* stores u_n and u_{n-1}, as solver.py:148-150, 234 */

/* time-step kernel choice */
#define SW_KERNEL_AUTO 0
#define SW_KERNEL_DIRECT 1  /* one voxel per thread, neighbours through the caches */
#define SW_KERNEL_MARCH 2   /* x-marching LDS tiles with register queues, one step per pass */
#define SW_KERNEL_MARCH2 3  /* as MARCH, but two increment-form steps per pass wherever
                               no output falls between them (AUTO picks it) */
#define SW_KERNEL_MARCH3 4  /* as MARCH2, but three force-free float32 steps per pass
                               wherever no output falls between them */

/* CG stop reasons */
#define SW_CG_RUNNING 0
#define SW_CG_CONVERGED 1   /* sqrt(rs/rs0) < tol            (solver.py:1044) */
#define SW_CG_DENOM 2       /* p.Ap non-finite or <= 0         (solver.py:1036) */
#define SW_CG_NONFINITE 3   /* r.r non-finite                  (solver.py:1042) */
#define SW_CG_MAX_ITERS 4   /* loop exhausted                  (solver.py:1033) */
#define SW_CG_EARLY_EXIT 5  /* max|div f| < 1e-12, no solve    (solver.py:672)  */

/* Last error message of the calling thread ("" when none). */
const char* sw_last_error(void);

/* SW_ABI_VERSION of the loaded library. */
int sw_abi_version(void);

/* Number of compute units and the gfx architecture name of the current device. */
int sw_device_info(int* cu_count, char* arch, int arch_len);

/* ---------------------------------------------------------------- stencils */

/* laplacian_center (solver.py:686-695): compact 7-point Laplacian, 0 on the
 * boundary layer. dtype SW_F32 or SW_F64; f and out have n = nx*ny*nz elements. */
int sw_laplacian_center(int dtype, const void* f, void* out, int64_t nx, int64_t ny, int64_t nz,
                        double dx, double dy, double dz, void* stream);

/* divergence_center (solver.py:708-717): 2h central divergence, 0 on the boundary layer. */
int sw_divergence_center(int dtype, const void* fx, const void* fy, const void* fz, void* out,
                         int64_t nx, int64_t ny, int64_t nz, double dx, double dy, double dz, void* stream);

/* gradient_center (solver.py:720-729): 2h central gradient, 0 on the boundary layer. */
int sw_gradient_center(int dtype, const void* phi, void* gx, void* gy, void* gz, int64_t nx, int64_t ny,
                       int64_t nz, double dx, double dy, double dz, void* stream);

/* ------------------------------------------------------------- projection */

typedef struct sw_cg_info {
    int64_t iters;      /* number of phi updates performed */
    double relres;      /* last sqrt(r.r / rs0) of the residual the CG iterates on */
    double max_div;     /* max |div_c(mask f)| of the input */
    int32_t reason;     /* SW_CG_* */
    int32_t pad;
    double true_relres; /* sqrt(|b - A phi|^2 / rs0), b = -div_c(mask f) in float64, from phi as
                           returned (-1 where not computed) */
} sw_cg_info;

/* Bytes of device workspace sw_project_vector_field needs for an nx*ny*nz grid. */
int64_t sw_project_workspace_bytes(int64_t nx, int64_t ny, int64_t nz);

/* project_vector_field (solver.py:644-683) with poisson_cg (solver.py:1010-1050).
 * Overwrites fx, fy, fz (dtype storage) with mask*(mask*f - grad_c phi), phi
 * solving -div_c grad_c phi = -div_c(mask*f) by CG on the interior. phi, the
 * dot products and the residual norm are float64; the CG vectors r and p are
 * stored in vec_dtype: SW_F64 as poisson_cg keeps them, or SW_F32 (p rounded
 * as it is formed, and every use of p reads that stored value, so phi and r
 * stay consistent). If scale_out is non-NULL (three arrays of dtype), it also
 * receives scale * f_S * inv_rho (inv_rho: dtype array, or NULL for 1).
 * `info` is a host pointer, filled after the call synchronises `stream`. */
int sw_project_vector_field(int dtype, int vec_dtype, void* fx, void* fy, void* fz, int64_t nx, int64_t ny,
                            int64_t nz, double dx, double dy, double dz, double tol, int64_t max_iters,
                            void* workspace, int64_t workspace_bytes, void* const* scale_out, const void* inv_rho,
                            double scale, sw_cg_info* info, void* stream);

/* Whether sw_project_vector_field, given a workspace of
 * sw_project_workspace_bytes(nx, ny, nz), runs the spectral-preconditioned CG
 * (1) or the plain one (0). The preconditioner (an exact solve of the
 * interior operator -div_c grad_c on each of its 8 parity sub-lattices, by
 * closed-form sine/cosine bases along each axis) is taken for float32 CG
 * vectors on grids whose parity chains have at most 256 points, when tol <=
 * 1e-7 and max_iters is at least the plain CG's iteration estimate, i.e. where
 * the plain CG of poisson_cg (solver.py:1010-1050) meets tol as well: both
 * stop on sqrt(r.r / r0.r0) < tol, and the preconditioned one checks its true
 * residual. Environment SW_CG_PRECOND=0 / 1 turns it off / on wherever
 * possible. No reference counterpart. */
int sw_project_preconditioned(int vec_dtype, int64_t nx, int64_t ny, int64_t nz, double tol, int64_t max_iters);

/* No reference counterpart: for callers that cannot allocate device memory
 * (tight serving loops, harnesses that forbid hipMalloc). Like
 * sw_project_vector_field it synchronizes `stream` to poll the CG's stop flag,
 * so it cannot itself be recorded into a graph; it allocates nothing, so a
 * caller's captured work around it needs no allocation either.
 * sw_project_vector_field with the CG workspace taken from a
 * static device array of the library instead of a caller's buffer, for grids
 * of at most sw_project_noalloc_max_voxels() (2^20) voxels; larger grids are
 * rejected with SW_ERR_ARG. It allocates nothing. Each GPU holds one such
 * array (33.3 MiB: the CG vectors and the preconditioner's tables) until the
 * process ends, shared by every call on that
 * GPU: calls that overlap on one GPU race. */
int64_t sw_project_noalloc_max_voxels(void);
int sw_project_vector_field_noalloc(int dtype, int vec_dtype, void* fx, void* fy, void* fz, int64_t nx, int64_t ny,
                                    int64_t nz, double dx, double dy, double dz, double tol, int64_t max_iters,
                                    void* const* scale_out, const void* inv_rho, double scale, sw_cg_info* info,
                                    void* stream);

/* ---------------------------------------------------------------- time loop */

/* Material coefficients (solver.py:124-145, 1072-1079): from float64 rho, mu
 * and eta arrays (eta may be NULL) writes mu/rho, eta/rho and 1/rho in dtype. */
int sw_prepare_coefficients(int dtype, const double* rho, const double* mu, const double* eta, void* mu_over_rho,
                            void* eta_over_rho, void* inv_rho, int64_t n, void* stream);

typedef struct sw_fdtd_state {
    int32_t dtype;  /* SW_F32 / SW_F64 */
    int32_t form;   /* SW_FORM_* */
    int64_t nx, ny, nz;
    double dx, dy, dz, dt;
    double vel_damp;
    const void* mu_over_rho;  /* (nx,ny,nz) */
    const void* eta_over_rho; /* (nx,ny,nz) or NULL when inviscid */
    const void* force[3];     /* dt^2 * f_S / rho per component, or NULL */
    void* u[3];               /* u_n                                 (in/out, rotated) */
    void* b[3];               /* d_n (increment) or u_{n-1} (leapfrog)  (in/out, rotated) */
    void* u_next[3];          /* scratch for u_{n+1}                 (rotated) */
    void* b_next[3];          /* scratch for d_{n+1} (increment form only; rotated) */
    int32_t kernel;           /* SW_KERNEL_* */
    int32_t tile_y;           /* march: y rows per tile (0 = auto); MARCH2: rows of the two-step tile
                                 (float32 8 or 12, float64 4), its single steps use the default */
    int32_t x_chunk;          /* march kernel: x planes per tile (0 = auto) */
    int32_t tile_z;           /* march kernel: z width of a tile, 64 or 256 (0 = auto) */
} sw_fdtd_state;

/* Optional per-step outputs of sw_fdtd_run. Step indices are global:
 * step0 is the index of the first step of this call, so a solve run in
 * several calls (host callback, reprojection) numbers its outputs as one. */
typedef struct sw_fdtd_outputs {
    int64_t step0;
    /* every snap_stride global steps (0 = never): u as AoS (nx,ny,nz,3) into
     * snap_out + k*3n, k = (step+1)/snap_stride - 1, as shear_fdtd_jax's
     * snapshot_stride (solver.py:469-507) */
    int64_t snap_stride;
    void* snap_out;
    /* every inv_stride global steps (0 = never): the four strain invariants of
     * u (strain.py:16-135) as SoA (4,nx,ny,nz) into inv_out + k*4n */
    int64_t inv_stride;
    void* inv_out;
    /* after every step: u_z at n_traces flat voxel indices (device int64
     * array) into trace_out[k*trace_ld + step], as shear_fdtd_jax's
     * trace_indices / return_traces (solver.py:396-413, 444-447) */
    const int64_t* trace_idx;
    int64_t n_traces;
    void* trace_out;
    int64_t trace_ld;
    /* non-finite detection (no reference counterpart: the reference lets an
     * unstable run fill u with inf/NaN silently). Every finite_stride global
     * steps (0 = never) u is checked, after the first pass that ends at or
     * past such a step, so the check never splits a pass. If u holds an inf or
     * NaN and *nonfinite_step (device int64; the caller sets it to -1) is still
     * negative, it becomes the global index of that pass's last step. */
    int64_t finite_stride;
    int64_t* nonfinite_step;
    /* how the in-loop invariants (inv_stride) are computed: 0 = by the next
     * pass when that is a force-free two-step pass (it reads the state
     * anyway and stores the invariants of its input state), else, and after
     * the last pass of the call, by their own kernel right away; 1 = always
     * by their own kernel right after the step. The values are the same. */
    int64_t inv_mode;
} sw_fdtd_outputs;

/* The time loop of shear_fdtd (solver.py:194-239) for n_steps steps with
 * separable forcing f(x,t) = f_S(x) env[it] (solver.py:189-197). env is a host
 * array of n_steps doubles for these steps (NULL = all zero). Rotates the
 * state's buffers in place: after the call u/b hold the state after the last
 * step. `out` may be NULL. */
int sw_fdtd_run(sw_fdtd_state* st, int64_t n_steps, const double* env, const sw_fdtd_outputs* out, void* stream);

/* Initial state (solver.py:147-171): u = mask*u0 (u0 AoS (nx,ny,nz,3) or NULL
 * for zero); u_prev = mask*(u - dt*v0) when v0 (AoS) is given, else 0. Writes
 * u and, per form, b = u - u_prev (increment) or b = u_prev (leapfrog). */
int sw_import_state(int dtype, int form, const void* u0, const void* v0, double dt, void* const* u, void* const* b,
                    int64_t nx, int64_t ny, int64_t nz, void* stream);

/* Returned fields (solver.py:241-243): u as AoS and v = (u - u_prev)/dt as
 * AoS, i.e. b*inv_dt (increment) or (u - b)*inv_dt (leapfrog). Either output may be NULL. */
int sw_export_state(int dtype, int form, const void* const* u, const void* const* b, double inv_dt, void* u_aos,
                    void* v_aos, int64_t n, void* stream);

/* out[c] = alpha*a[c] + beta*b[c] for the three components of n elements
 * (state conversions around reprojection, solver.py:236-239). */
int sw_combine3(int dtype, const void* const* a, const void* const* b, double alpha, double beta, void* const* out,
                int64_t n, void* stream);

/* out[k*out_stride] = f[idx[k]] for k < count (device int64 flat indices into
 * the n elements of f): the per-step sensor traces of shear_fdtd_jax
 * (solver.py:396-413, 444-447). An index outside [0, n) reads nothing and
 * yields NaN (the kernel checks every index, so a bad one cannot fault). */
int sw_gather_traces(int dtype, const void* f, int64_t n, const int64_t* idx, int64_t count, void* out,
                     int64_t out_stride, void* stream);

/* ---------------------------------------------------------------- batching */

/* One solve of a batch (float32 storage, increment form), as device memory.
 * A batch advances all its solves with ONE persistent launch per pass: work
 * items are (solve, z tile, y tile, x chunk) quadruples of int32 over
 * tile_y-row, 64-column tiles (tile_y 4, 8 or 12, one per launch), and a device-side global pass counter gives each solve its
 * pass k = gpass - start (active while 0 <= k < n_passes; its state is read
 * from parity k&1 and written to the other parity). A pass is one step, or
 * two steps when plan says so: plan[k] = (first local step << 1)
 * | (1 if the pass is two steps). plan == NULL means one step per pass
 * (n_passes = n_steps, local step = k). Each solve's y-z plane (ny*nz) holds
 * fewer than 2^28 voxels: the kernels address a plane with 32-bit offsets.
 * No reference counterpart: the reference runs one solve per call
 * (solver.py:27). */
typedef struct sw_batch_solve {
    int32_t nx, ny, nz;
    int32_t xchunk;                  /* x planes per work item */
    float cx, cy, cz;                /* dt^2 / h^2 */
    float vx, vy, vz;                /* dt / h^2 */
    float damp;                      /* 1 - vel_damp */
    float pad0;
    int64_t start;                   /* global pass of this solve's first pass */
    int64_t n_steps;
    const float* mu_over_rho;
    const float* eta_over_rho;       /* NULL when inviscid */
    const float* force[3];           /* dt^2 f_S / rho, or NULL */
    const double* env;               /* device, n_steps values, or NULL */
    float* u[2][3];                  /* u by parity */
    float* d[2][3];                  /* d = u - u_prev by parity */
    const int32_t* plan;             /* device, n_passes entries, or NULL */
    int64_t n_passes;
} sw_batch_solve;

/* Advance the batch by n_passes global passes: per pass one persistent launch
 * over all `items` (tiles of tile_y = 4, 8 or 12 rows, as the items were
 * cut; blocks = 0 picks one resident wave) and one
 * increment of gpass[0]. gpass is device memory of SW_BATCH_PASS_STATE int64:
 * gpass[0] the global pass, gpass[1..9] the launch's work-claim counters,
 * which the caller zeroes once and every launch leaves at zero. The item list
 * is cut into 8 equal ranges, one per XCD; blocks claim items of their XCD's
 * range in list order, then of the others': list neighbouring tiles next to
 * each other. */
#define SW_BATCH_PASS_STATE 16
int sw_batch_step(const sw_batch_solve* solves, const int32_t* items, int64_t n_items, int64_t* gpass, int32_t tile_y,
                  int32_t blocks, int64_t n_passes, void* stream);

/* The same n_passes captured once into a hipGraph (no memset nodes), to be
 * replayed with sw_graph_launch; the graph keeps the pointers it was built with. */
typedef struct sw_graph sw_graph;
int sw_batch_graph_create(const sw_batch_solve* solves, const int32_t* items, int64_t n_items, int64_t* gpass,
                          int32_t tile_y, int32_t blocks, int64_t n_passes, sw_graph** graph);
int sw_graph_launch(sw_graph* graph, void* stream);
int sw_graph_destroy(sw_graph* graph);

/* One solve's projection in a batched projection (project_vector_field,
 * solver.py:644-683, for many solves at once), as device memory; ny*nz
 * below 2^28 as for sw_batch_solve. */
typedef struct sw_batch_cg {
    int32_t nx, ny, nz;
    int32_t xchunk;                  /* x planes per CG work item: a multiple of 48, or >= nx */
    double dx, dy, dz, tol;
    int64_t max_iters;
    float* f[3];                     /* force in, f_S out (in place) */
    float* out[3];                   /* scale * f_S * inv_rho, or NULL */
    const float* inv_rho;            /* or NULL for 1 */
    double scale;
    double* phi;                     /* float64 potential, nx*ny*nz */
    float *r, *p0, *p1;              /* float32 CG vectors, nx*ny*nz each (dot products stay float64) */
    double* part;                    /* 2 * n_slots float64 partial slots, n_slots =
                                        ceil(nx/48) ceil(ny/16) ceil(nz/64) */
    int32_t item0, n_items;          /* the caller's bookkeeping of this solve's items (the
                                        kernels do not read it) */
} sw_batch_cg;

/* Bytes of device CG state per solve. */
int64_t sw_cg_state_bytes(void);

/* Project every solve of the batch; CG items are (solve, z tile, y tile,
 * x chunk) quadruples over 16-row tiles (64 z columns wide). Dot products are
 * summed per 48-plane x sub-chunk of a tile, then over a solve's slots in a
 * fixed order: a solve's result depends neither on xchunk nor on the other
 * solves of the batch. The descriptors and items are checked first (one
 * small device-to-host copy): SW_ERR_ARG unless every tile of every solve is
 * listed exactly once and each xchunk is valid. `states` is device
 * memory of (n_solves + 1) * sw_cg_state_bytes() bytes: a CG state per solve,
 * then the kernels' work-claim counters (blocks claim items in list order, so
 * list the longest first). Runs until every solve has stopped; info (host,
 * n_solves entries) receives each solve's report. */
int sw_batch_project(const sw_batch_cg* solves, int32_t n_solves, const int32_t* items, int64_t n_items,
                     void* states, sw_cg_info* info, void* stream);

/* -------------------------------------------------------- differentiation */

/* One step of the adjoint of sw_fdtd_run's update (increment form), for
 * gradients with respect to mu/rho and eta/rho. No reference counterpart: the
 * reference has no derivative code. Given the adjoint (ub_next, db_next) of
 * the state after the step and the forward state (u, d) before it, writes the
 * adjoint (ub, db) of the state before the step and accumulates
 * grad_m += dJ/d(mu/rho), grad_e += dJ/d(eta/rho) (float64 arrays; grad_e and
 * eta_over_rho NULL when inviscid). */
int sw_adjoint_step(int dtype, const void* const* u, const void* const* d, const void* const* ub_next,
                    const void* const* db_next, void* const* ub, void* const* db, const void* mu_over_rho,
                    const void* eta_over_rho, double* grad_m, double* grad_e, int64_t nx, int64_t ny, int64_t nz,
                    double dx, double dy, double dz, double dt, double vel_damp, void* stream);

/* f[idx[k]] += grad[k*stride] for k < count (unique device int64 flat indices
 * into the n elements of f): the misfit gradient of sensor traces entering the
 * adjoint of u_z. An index outside [0, n) is skipped. No reference
 * counterpart (the reference has no adjoint; its JAX backend differentiates
 * shear_fdtd_jax, solver.py:246). */
int sw_inject_traces(int dtype, void* f, int64_t n, const int64_t* idx, int64_t count, const void* grad,
                     int64_t stride, void* stream);

/* One two-step pass of sw_fdtd_run's update (increment form; steps with
 * env2[0] and env2[1], a host array of two doubles or NULL for zero) from
 * (u, b) into (u_next, b_next), that also writes the state after the first
 * step into (u_mid, d_mid): the values a single step writes, bit for bit. The
 * state's buffers are not rotated. For the adjoint's recomputation of the
 * forward states of a segment (the forward loop, solver.py:194-234, runs such
 * passes). No reference counterpart. SW_ERR_UNSUPPORTED on grids the
 * x-marching pass does not take (nz < 32, ny < 8, nx < 8). */
int sw_fdtd_step2_mid(const sw_fdtd_state* st, const double* env2, void* const* u_mid, void* const* d_mid,
                      void* stream);

/* n_passes force-free two-step passes from (u, b) into (u_next, b_next), with
 * the state's tile_y and x_chunk, each reading the same state: the buffers
 * are not rotated, so the state is unchanged and u_next / b_next receive one
 * pass's result. For the launch-shape autotuner, which times candidate tile
 * shapes on a solve's own buffers before its loop. No reference counterpart.
 * SW_ERR_UNSUPPORTED where sw_fdtd_step2_mid is. */
int sw_fdtd_trial_passes(const sw_fdtd_state* st, int64_t n_passes, void* stream);

/* ------------------------------------------------------------------ strain */

/* compute_strain_tensor (strain.py:16-70): Voigt [exx,eyy,ezz,gyz,gxz,gxy],
 * engineering shear, written AoS (nx,ny,nz,6). */
int sw_strain_tensor(int dtype, const void* ux, const void* uy, const void* uz, void* eps, int64_t nx, int64_t ny,
                     int64_t nz, double dx, double dy, double dz, void* stream);

/* compute_strain_invariants (strain.py:73-135) on `count` Voigt 6-vectors (AoS):
 * volumetric, deviatoric, max_shear (dtype) and von_mises, which is float64 for
 * both dtypes because the reference multiplies by np.sqrt(2/3), a float64
 * scalar, and NumPy 2 promotes the float32 result (strain.py:123). */
int sw_strain_invariants(int dtype, const void* eps, int64_t count, void* vol, void* dev, void* max_shear,
                         void* von_mises, void* stream);

/* Fused compute_strain_invariants(compute_strain_tensor(u)): the four
 * invariants straight from the displacement, without the tensor in memory. */
int sw_strain_invariants_from_u(int dtype, const void* ux, const void* uy, const void* uz, void* vol, void* dev,
                                void* max_shear, void* von_mises, int64_t nx, int64_t ny, int64_t nz, double dx,
                                double dy, double dz, void* stream);

/* --------------------------------------------------------- radiation force */

/* Pressure (and particle velocity) fields are (n_vox, nt) with time fastest,
 * i.e. the reference's (nx, ny, nz, nt) arrays; maps are (n_vox). A gate is
 * either `gate`, n_gate device int32 time indices in [0, nt) (repeats
 * allowed, as NumPy fancy indexing), or, with gate == NULL, the contiguous
 * range [gate_start, gate_start + n_gate). n_gate = 0 gives NaN, the mean of
 * an empty gate. */

/* env[t] = sum over voxels of p^2 (float64, deterministic): the input of the
 * auto-gate of compute_radiation_force (radiation_force.py:66-72), whose
 * mean is env / n_vox. workspace: sw_pressure_energy_workspace bytes. */
int64_t sw_pressure_energy_workspace(int64_t n_vox, int64_t nt);
int sw_pressure_energy(int dtype, const void* p, int64_t n_vox, int64_t nt, double* env, void* workspace,
                       int64_t workspace_bytes, void* stream);

/* compute_radiation_force (radiation_force.py:27-83): per voxel
 * b = 2 alpha_Np mean_gate(p^2) / (rho c) / c, alpha_Np = alpha f0[MHz] / 8.685889638 * 100. */
int sw_radiation_force(int dtype, const void* p, int64_t n_vox, int64_t nt, const int32_t* gate, int64_t gate_start,
                       int64_t n_gate, const void* c, const void* rho, const void* alpha_coeff, double f0, void* out,
                       void* stream);

/* compute_radiation_force_velocity (radiation_force.py:86-157): per voxel and
 * component b_i = 2 alpha_Np mean_gate(p v_i) / c. */
int sw_radiation_force_velocity(int dtype, const void* p, const void* vx, const void* vy, const void* vz,
                                int64_t n_vox, int64_t nt, const int32_t* gate, int64_t gate_start, int64_t n_gate,
                                const void* c, const void* alpha_coeff, double f0, void* bx, void* by, void* bz,
                                void* stream);

/* --------------------------------------------------------- acoustic strain */

/* Records, maps and gates as for the radiation force; K = rho c^2.
 * compute_acoustic_strain_volumetric (acoustic_strain.py:75-104):
 * out (n_vox, n_gate) = -p[:, gate] / K. */
int sw_acoustic_strain_volumetric(int dtype, const void* p, int64_t n_vox, int64_t nt, const int32_t* gate,
                                  int64_t gate_start, int64_t n_gate, const void* c, const void* rho, void* out,
                                  void* stream);

/* compute_acoustic_strain_peak (acoustic_strain.py:107-151): per voxel
 * peak = max|p| / K, rms = sqrt(mean p^2) / K, mean = -mean(p) / K. */
int sw_acoustic_strain_peak(int dtype, const void* p, int64_t n_vox, int64_t nt, const int32_t* gate,
                            int64_t gate_start, int64_t n_gate, const void* c, const void* rho, void* peak, void* rms,
                            void* mean, void* stream);

/* compute_acoustic_strain_gradient (acoustic_strain.py:154-203): gradient_center
 * of -p_t / K for every gated sample; gx, gy, gz are (nx, ny, nz, n_gate). */
int sw_acoustic_strain_gradient(int dtype, const void* p, int64_t nx, int64_t ny, int64_t nz, int64_t nt,
                                const int32_t* gate, int64_t gate_start, int64_t n_gate, const void* c,
                                const void* rho, double dx, double dy, double dz, void* gx, void* gy, void* gz,
                                void* stream);

/* compute_acoustic_strain_gradient_peak (acoustic_strain.py:206-277): float64
 * running max of |grad|, |d/dx|, |d/dy|, |d/dz| and sqrt(mean |grad|^2). */
int sw_acoustic_strain_gradient_peak(int dtype, const void* p, int64_t nx, int64_t ny, int64_t nz, int64_t nt,
                                     const int32_t* gate, int64_t gate_start, int64_t n_gate, const void* c,
                                     const void* rho, double dx, double dy, double dz, double* peak_mag,
                                     double* peak_gx, double* peak_gy, double* peak_gz, double* rms_mag,
                                     void* stream);

/* ------------------------------------------------------- static elasticity */

typedef struct sw_static_info {
    int64_t iters;   /* completed CG iterations (u updates) */
    float relres;    /* last sqrt(r.r / rs0), NaN before the first iteration */
    int32_t reason;  /* SW_CG_CONVERGED, SW_CG_DENOM (p.Ap <= 0) or SW_CG_MAX_ITERS */
    int32_t pad;
} sw_static_info;

/* Bytes of device workspace sw_elastic_solve_static needs for an nx*ny*nz grid. */
int64_t sw_static_workspace_bytes(int64_t nx, int64_t ny, int64_t nz);

/* elastic_solve_static (solver.py:569-625): CG on (-mu lap) u = bz with a zero
 * Dirichlet boundary, float32 throughout as the reference. bz, mu and u are
 * float32 device arrays of nx*ny*nz voxels; mu == NULL uses mu_scalar
 * (rounded to float32) everywhere. The dot products are summed in NumPy's
 * float32 order (8192-element buffers, pairwise within), so u, every scalar
 * and the iteration count equal the reference's. tol_is_f32 != 0 compares
 * rel < tol in float32 (a Python float tol), 0 in float64 (a float64 NumPy
 * scalar). history, if non-NULL, is a float32 device array of 2*max_iters
 * receiving (r.r, rel) per iteration. `info` is a host pointer, filled after
 * the call synchronises `stream`. */
int sw_elastic_solve_static(const float* bz, const float* mu, double mu_scalar, int64_t nx, int64_t ny, int64_t nz,
                            double dx, double dy, double dz, double tol, int tol_is_f32, int64_t max_iters,
                            float* u, void* workspace, int64_t workspace_bytes, float* history,
                            sw_static_info* info, void* stream);

/* ------------------------------------------------------------------ layout */

/* (count, 3) AoS <-> three SoA arrays of `count` elements (solver.py:148 keeps
 * vectors as (nx,ny,nz,3); the kernels use SoA). */
int sw_aos3_to_soa(int dtype, const void* aos, void* x, void* y, void* z, int64_t count, void* stream);
int sw_soa_to_aos3(int dtype, const void* x, const void* y, const void* z, void* aos, int64_t count, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* SHEARWAVE_H */
