/*
 * pll.h -- C-ABI of the MI355X-native PowerLLEL channel-flow hot path.
 *
 * One opaque context owns every device buffer of one channel-flow problem on
 * one GPU: the seven halo-padded fields (u, v, w, u1, v1, w1, p), the mesh
 * coefficients, the FFT twiddles, the pre-factored tridiagonal pivots and the
 * reduction scratch.  Each entry point replaces one routine of the reference's
 * module interface; the citation on each line is the routine it replaces
 * (paths relative to the PowerLLEL_Channel_SYCL source tree).
 *
 * Layout contract (src/mod_type.hh:107-117,172-179): a field is
 * (nx+2)*(ny+2)*(nz+2) values, x fastest, one halo cell per side, logical
 * indices 0..n+1.  Host buffers passed to the upload/download calls use this
 * layout in FP64, whatever the context's precision.  The Poisson work array
 * (var_xpen, src/mod_poisson_solver.cc:196-214) is nx*ny*nz values ordered
 * [k][j][i] with the reference's "perm" packing of each real DFT
 * (src/mod_fft.cc:79-86).
 *
 * Semantics: calls are ordered on the context's HIP stream and return
 * before the GPU finishes, except those that hand a value back to the host
 * (means, checks, downloads), which synchronise.  pll_synchronize() waits for
 * everything queued.  Every call returns PLL_OK or a pll_status code;
 * pll_last_error() describes the last failure on the calling thread.  Nothing
 * aborts the process.  One context must not be driven by two host threads at
 * the same time.
 */
#ifndef PLL_H
#define PLL_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PLL_ABI_VERSION 1

typedef struct pll_context pll_context;

typedef enum pll_status {
  PLL_OK = 0,
  PLL_ERR_ARG = 1,         /* invalid argument or unsupported size          */
  PLL_ERR_HIP = 2,         /* a HIP runtime call failed                     */
  PLL_ERR_STATE = 3,       /* call not valid in the context's current state */
  PLL_ERR_UNSUPPORTED = 4, /* feature not available in this build           */
  PLL_ERR_NOMEM = 5,       /* device allocation failed                      */
  PLL_ERR_IO = 6           /* a file could not be read or written           */
} pll_status;

typedef enum pll_precision {
  PLL_FP64 = 0,      /* every field, transform and solve in FP64           */
  PLL_FP32_MIXED = 1 /* FP32 fields/stencils/FFTs; FP64 for the kx=ky=0 and
                        low-wavenumber tridiagonal solves and for the
                        bulk-velocity reduction                            */
} pll_precision;

typedef enum pll_field {
  PLL_U = 0, PLL_V = 1, PLL_W = 2,
  PLL_U1 = 3, PLL_V1 = 4, PLL_W1 = 5,
  PLL_P = 6,
  PLL_NFIELD = 7
} pll_field;

typedef enum pll_profile { PLL_PROFILE_UNI = 0, PLL_PROFILE_POI = 1, PLL_PROFILE_LOG = 2 } pll_profile;

/* Problem definition: the PARA/MESH/TIME/PHYSICAL/FORCE keys of the
 * reference's param.json that the hot path reads (src/mod_parameters.hh:12-88). */
typedef struct pll_config {
  int nx, ny, nz;          /* global grid; even (src/mod_parameters.cc:170-176)   */
  double lx, ly, lz;
  double re, u_ref, l_ref; /* re_inv = u_ref*l_ref/re (src/mod_parameters.cc:259) */
  double dt;
  double u_crf;            /* convecting reference frame velocity                */
  int is_forced[3];
  double vel_force[3];
  int precision;           /* pll_precision                                      */
  int pin_singular;        /* 1: the kx=ky=0 Neumann pivot is exactly zero, so the
                              solve takes the reference's b==0 branch
                              (src/mod_poisson_solver.cc:341-346); PLL_FP32_MIXED
                              always pins (p cannot carry the round-off constant) */
  int device;              /* HIP device ordinal                                 */
} pll_config;

/* Max-with-location record of the runtime checks (src/mod_utils.hh:8-12). */
typedef struct pll_value_index {
  double value;
  int rank;
  int ig, jg, kg;
} pll_value_index;

/* ---- library --------------------------------------------------------------- */
int pll_abi_version(void);
const char *pll_last_error(void);
/* Fills cfg with the defaults of examples/template/param.json. */
void pll_config_default(pll_config *cfg);

/* ---- mod_mesh ---------------------------------------------------------------- */
/* initMeshByFunc (src/mod_mesh.cc:95-125): dzf_global[0..nz+1]. */
int pll_mesh_by_func(int mesh_type, int stretch_func, double stretch_ratio, int nz, double lz,
                     double *dzf_global);

/* ---- context (allocVariables + initMesh + init_poisson_solver) -------------- */
/* allocVariables (src/mod_variables.cc:4-12), initMesh coefficients
 * (src/mod_mesh.cc:54-65) and init_poisson_solver
 * (src/mod_poisson_solver.cc:180-315) in one call.  dzf_global is the
 * halo-inclusive spacing array dzf_global[0..nz+1] (src/main.cc:55-57). */
int pll_create(const pll_config *cfg, const double *dzf_global, pll_context **out);
/* free_poisson_solver + freeVariables + freeMesh (src/mod_poisson_solver.cc:512-537). */
int pll_destroy(pll_context *ctx);
int pll_get_config(const pll_context *ctx, pll_config *cfg);
/* Bytes of device memory the context holds: every buffer it allocated (fields,
 * spare fields, work array, pivots, twiddles, mesh, reduction scratch, and the
 * statistics sums and closure buffers when enabled). */
int pll_device_bytes(const pll_context *ctx, size_t *bytes);

/* ---- field access (Array3D::data(), src/mod_type.hh:169-170) ----------------- */
int pll_field_upload(pll_context *ctx, int field, const double *host);
int pll_field_download(pll_context *ctx, int field, double *host);
/* Device address of a field and its element size (8 for FP64, 4 for FP32). */
int pll_field_device_ptr(pll_context *ctx, int field, void **dev, int *elem_bytes);
/* Poisson work array var_xpen (nx*ny*nz, [k][j][i]). */
int pll_spec_upload(pll_context *ctx, const double *host);
int pll_spec_download(pll_context *ctx, double *host);

/* ---- mod_initFlow ------------------------------------------------------------ */
/* initFlow (src/mod_initFlow.cc:47-102) with this library's documented noise
 * stream: for interior point L = (k-1)*nx*ny + (j-1)*nx + (i-1) of component
 * d, a std::minstd_rand engine seeded with seeds[d] is advanced L steps and
 * drawn once (x); the point gets 2*((x-1)/2147483646 - 0.5)*noise_intensity.
 * Keeps the reference's Mean(nz, ny, ...) normalisation (src/mod_initFlow.cc:92). */
int pll_init_flow(pll_context *ctx, int profile, double u0, int with_noise,
                  double noise_intensity, const uint32_t seeds[3]);

/* ---- mod_calcVel -------------------------------------------------------------- */
/* transform2CRF (src/mod_calcVel.cc:13-25) on u; vel_force[0] -= vel_crf. */
int pll_transform2crf(pll_context *ctx, double vel_crf);
/* timeIntVelRK1 (src/mod_calcVel.cc:131-138): u1,v1,w1 = u + dt*F(u). */
int pll_time_int_vel_rk1(pll_context *ctx);
/* timeIntVelRK2 (src/mod_calcVel.cc:244-251): u = (u1 + dt*F(u1) + u)/2. */
int pll_time_int_vel_rk2(pll_context *ctx);
/* correctVel (src/mod_calcVel.cc:253-278): u -= dt*grad(p). */
int pll_correct_vel(pll_context *ctx);
/* forceVel (src/mod_calcVel.cc:280-320).  mean_out (may be NULL) receives the
 * bulk means of the forced components; passing NULL keeps the call asynchronous. */
int pll_force_vel(pll_context *ctx, double mean_out[3]);

/* ---- mod_calcRHS -------------------------------------------------------------- */
/* calcRHS (src/mod_calcRHS.cc:7-33): p = div(u)/dt. */
int pll_calc_rhs(pll_context *ctx);

/* ---- poisson_solver / mod_fft ------------------------------------------------- */
/* execute_poisson_solver (src/mod_poisson_solver.cc:454-510) on p. */
int pll_poisson_execute(pll_context *ctx);
/* execute_fft_fwd_xpen (src/mod_fft.cc:60-88): p interior -> work array. */
int pll_fft_fwd_xpen(pll_context *ctx);
/* execute_fft_fwd_ypen (src/mod_fft.cc:122-153): work array, in place. */
int pll_fft_fwd_ypen(pll_context *ctx);
/* ssolve_trid (src/mod_poisson_solver.cc:317-357): work array, in place. */
int pll_solve_trid(pll_context *ctx);
/* execute_fft_bwd_ypen (src/mod_fft.cc:155-187): work array, in place. */
int pll_fft_bwd_ypen(pll_context *ctx);
/* execute_fft_bwd_xpen (src/mod_fft.cc:90-120): work array -> p interior, scaled 1/(nx*ny). */
int pll_fft_bwd_xpen(pll_context *ctx);
/* get_eigen_values (src/mod_fft.cc:196-222): perm-ordered eigenvalues, unscaled. */
int pll_eigen_values(int n, double *lambda);

/* ---- mod_updateBound ---------------------------------------------------------- */
/* updateBoundVel (src/mod_updateBound.cc:229-251): which = 0 -> (u,v,w), 1 -> (u1,v1,w1). */
int pll_update_bound_vel(pll_context *ctx, int which);
/* imposeBCVel (src/mod_updateBound.cc:183-227) at one rank: only the wall
 * conditions (x and y neighbours are the rank itself, so no periodic copy):
 * u(0) = 2(0-u_crf) - u(1), v(0) = -v(1), w(0) = 0 and the same at the top
 * with w(nz) = w(nz+1) = 0, over the whole x/y extent including halo cells. */
int pll_impose_bc_vel(pll_context *ctx, int which);
/* updateBoundP (src/mod_updateBound.cc:253-296). */
int pll_update_bound_p(pll_context *ctx);

/* ---- mod_utils ------------------------------------------------------------------ */
/* Mean (src/mod_utils.cc:30-51): sum(var*dzflzi)/nx_global/ny_global, FP64. */
int pll_mean(pll_context *ctx, int field, int nx_global, int ny_global, double *out);
/* calcMaxDiv (src/mod_utils.cc:164-205) and calcMaxCFL (src/mod_utils.cc:108-150). */
int pll_calc_max_div(pll_context *ctx, pll_value_index *out);
int pll_calc_max_cfl(pll_context *ctx, pll_value_index *out);
/* checkNaN (src/mod_utils.cc:77-106): *passed = 1 when no NaN in the whole array. */
int pll_check_nan(pll_context *ctx, int field, int *passed);

/* ---- mod_statistics ------------------------------------------------------------ */
/* The per-point sums of calcStat (src/mod_statistics.cc:30-106), kept on the
 * device in FP64.  Quantity index: 0 u, 1 v, 2 w, 3 u2, 4 v2, 5 w2, 6 uv,
 * 7 uw, 8 vw, 9 p, 10 p2 (src/mod_dataIO.cc:39-63), from cell-centred
 * uc = (u(i-1)+u(i))/2 + u_crf, vc = (v(j-1)+v(j))/2, wc = (w(k-1)+w(k))/2.
 * Arrays are interior-only nz*ny*nx, x fastest (the checkpoint layout). */
/* Select quantities (which[t] != 0), allocate their sums and zero them. */
int pll_stat_enable(pll_context *ctx, const int which[11]);
/* Add one sample of the current u, v, w, p to every selected sum. */
int pll_stat_accumulate(pll_context *ctx);
int pll_stat_download(pll_context *ctx, int var, double *host);
int pll_stat_upload(pll_context *ctx, int var, const double *host);
/* host[k-1] = sum over plane k of the sum array of var (k = 1..nz). */
int pll_stat_plane_sums(pll_context *ctx, int var, double *host);

/* ---- learned closure ------------------------------------------------------------ */
/* No counterpart in the reference.  A 10 -> 64 -> 64 -> 1 perceptron (ReLU
 * hidden layers, softplus output) with BF16 weights and FP32 accumulation on
 * MFMA, predicting a normalised eddy viscosity from 10 normalised features
 * (9 filtered velocity gradients, wall distance).  Parameters (FP32 master):
 * W1[16][64] (rows 10..15 zero), b1[64], W2[64][64], b2[64], W3[64], b3.
 * Sample arrays are device pointers: x[n][10], y[n], out[n] (FP32). */
typedef struct pll_closure pll_closure;
int pll_closure_create(uint64_t seed, pll_closure **out);
int pll_closure_destroy(pll_closure *c);
int pll_closure_num_params(void);
int pll_closure_get_params(pll_closure *c, float *host);
int pll_closure_set_params(pll_closure *c, const float *host);
/* Gradient of the mean squared error of softplus(o) against y (host copy), and the loss. */
int pll_closure_gradient(pll_closure *c, const float *x, const float *y, long n, float *grad_host,
                         double *loss);
/* pll_closure_gradient without allocating: no counterpart in the reference;
 * for callers that cannot allocate device memory (code under HIP graph
 * capture, tight training loops).  The kernels' partial gradients, chunk sums
 * and result live in one static device array of 6.1 MB per GPU (every
 * training grid fits: at most 1024 waves); the call allocates nothing on the
 * device or host.  All calls on one GPU share that array, so calls that
 * overlap on one GPU (other streams, other threads, other closures) race.
 * The closure's device (current at pll_closure_create) must be current. */
int pll_closure_gradient_noalloc(pll_closure *c, const float *x, const float *y, long n,
                                 float *grad_host, double *loss);
/* One Adam step (beta 0.9, 0.999, eps 1e-8) on the full batch. */
int pll_closure_train_step(pll_closure *c, const float *x, const float *y, long n, double lr);
int pll_closure_predict(pll_closure *c, const float *x, long n, float *out);
int pll_closure_synchronize(pll_closure *c);
/* Scales of the network's inputs and output: features are the 9 velocity
 * gradients at a cell centre times gscale (du/dz, dv/dz, dw/dx, dw/dy negated
 * in the upper half channel, so both walls look alike) and the distance to the
 * nearer wall times yscale; the eddy viscosity is nuscale * softplus(o). */
int pll_closure_set_scales(pll_closure *c, double gscale, double yscale, double nuscale);
int pll_closure_get_scales(pll_closure *c, double *gscale, double *yscale, double *nuscale);
/* Attach a closure to a context (NULL detaches; detach before destroying the
 * closure).  With a closure attached, timeIntVelRK1/RK2 and pll_step add the
 * divergence of the eddy-viscosity stress 2 nu_t S_ij to the momentum
 * tendency (src/mod_calcVel.cc:27-242), with nu_t evaluated by the network
 * inside the stage kernel; pll_step then runs the unfused sequence
 * src/main.cc:117-153.  Training the closure while attached is ordered
 * before the context's next stage. */
int pll_closure_attach(pll_context *ctx, pll_closure *c);
/* Save a closure (parameters and scales) to a file / create one from such a
 * file: "PLLCLOS1", int32 parameter count, the FP32 parameters, the three FP64
 * scales, in host byte order. */
int pll_closure_save(pll_closure *c, const char *path);
int pll_closure_load(const char *path, pll_closure **out);
/* nu_t of the context's last closure stage, interior nz*ny*nx, x fastest. */
int pll_closure_nut(pll_context *ctx, double *host);
/* Box filter of the fine context's u, v, w onto the coarse context's grid
 * (factors nx_f/nx_c, ny_f/ny_c, nz_f/nz_c; every coarse z cell the union of
 * fine ones): each coarse face value is the area average of the fine face
 * values on that face, so divergence-free stays divergence-free; the coarse
 * halos are refreshed.  Blocks until done. */
int pll_closure_filter(pll_context *fine, pll_context *coarse);
/* Training samples of the fine context's current state: filters it into the
 * coarse context, then for each of the n = nx_c*ny_c*nz_c coarse cells writes
 * the features (x[n][10], device) and the label (y[n], device): the
 * least-squares eddy viscosity of the subgrid stress of the box
 * max(0, -tau^d_ij S_ij) / (2 S_ij S_ij + 1e-4/gscale^2) / nuscale, at most 50.
 * Blocks until done. */
int pll_closure_samples(pll_context *fine, pll_context *coarse, pll_closure *c, float *x,
                        float *y);

/* ---- decomp_2d ------------------------------------------------------------------ */
/* The 2D pencil decomposition for p_row x p_col ranks whose pencils all live
 * on this GPU (decomp_2d_init + decomp_info_init, src/decomp_2d.cc:41-207).
 * Rank r has coordinates (r / p_col, r % p_col); every pencil keeps x fastest;
 * device buffers hold all ranks' pencils of one orientation in rank order. */
typedef struct pll_decomp pll_decomp;
int pll_decomp_create(int nx, int ny, int nz, int p_row, int p_col, pll_decomp **out);
int pll_decomp_destroy(pll_decomp *d);
/* pencil 0 = x, 1 = y, 2 = z: 1-based start, size and element offset of rank's block
 * (decomp_info xst/xsz, yst/ysz, zst/zsz, src/decomp_2d.hh:35-37). */
int pll_decomp_pencil(const pll_decomp *d, int pencil, int rank, int st[3], int sz[3], long *offset);
int pll_decomp_total(const pll_decomp *d, int pencil, long *total);
/* transpose_{x_to_y,y_to_x,y_to_z,z_to_y}_real (src/decomp_2d.cc:522-589) on
 * device buffers of 8-byte (or 4-byte) values: a bit-exact permutation.
 * Synchronous on return. */
int pll_transpose_x_to_y_real(pll_decomp *d, const void *src, void *dst, int elem_bytes);
int pll_transpose_y_to_x_real(pll_decomp *d, const void *src, void *dst, int elem_bytes);
int pll_transpose_y_to_z_real(pll_decomp *d, const void *src, void *dst, int elem_bytes);
int pll_transpose_z_to_y_real(pll_decomp *d, const void *src, void *dst, int elem_bytes);

/* ---- main loop ------------------------------------------------------------------ */
/* nsteps repetitions of src/main.cc:116-154: RK1, BCvel(u1), RK2, BCvel(u),
 * calcRHS, Poisson, BCp, correctVel, forceVel, BCvel(u).  Asynchronous. */
int pll_step(pll_context *ctx, int nsteps);
int pll_synchronize(pll_context *ctx);

/* ---- per-kernel timing (HIP events on the context's stream) ------------------- */
/* When enabled, each kernel launch is bracketed by HIP events; totals are
 * accumulated at pll_timing_read().  Names are stable identifiers. */
int pll_timing_enable(pll_context *ctx, int on);
int pll_timing_reset(pll_context *ctx);
int pll_timing_count(pll_context *ctx, int *nkernels);
int pll_timing_read(pll_context *ctx, int index, const char **name, double *total_ms,
                    long *launches);

#ifdef __cplusplus
}
#endif

#endif /* PLL_H */
