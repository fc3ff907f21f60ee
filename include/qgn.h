/* C-ABI of the MI355X (gfx950) QGN spectral right-hand side and time stepper.
 *
 * Drop-in replacement for the parts of iangrooms/QGN_Fortran that do the
 * arithmetic: the MKL DFTI plans and transforms, and the QGN_Module
 * procedures Initialize / TimeStep / GetRHS / WriteQP.  A Fortran host binds
 * these with iso_c_binding (qgn_fortran_amd/host/qgn_c.f90).
 *
 * Conventions
 *   - Arrays use the reference's Fortran memory layout: physical fields
 *     real(nx,ny,nz), spectra complex(nx/2+1,ny,nz), x fastest, complex
 *     interleaved (re,im) as C double[2] / Fortran complex(c_double_complex).
 *   - Every function returns 0 on success or a negative QGN_E* code; the
 *     reference ignored every DFTI status (QGN_Module.f90:180,452-455).
 *     qgn_last_error() gives a message for the calling thread.
 *   - The caller owns host arrays; a context owns its device buffers, plans
 *     and HIP stream.  A context is not thread-safe; use one host thread per
 *     context.  The prognostic state stays in device memory between steps.
 */
#ifndef QGN_H
#define QGN_H

#ifdef __cplusplus
extern "C" {
#endif

enum {
  QGN_OK = 0,
  QGN_EINVAL = -1,    /* bad argument or unsupported size */
  QGN_EHIP = -2,      /* HIP runtime error */
  QGN_ENOMEM = -3,    /* device allocation failed */
  QGN_ENONFINITE = -4,/* NaN/Inf in the error norm of an adaptive step */
  QGN_EREJECT = -5    /* adaptive step rejected more than max_rejects times */
};

/* TimeStep include fragments (QGN_Module.f90:306) */
enum { QGN_AB3 = 0, QGN_RK4 = 1, QGN_BSRK32 = 2, QGN_ARK43 = 3, QGN_ARK43_QGLEITH = 4 };
/* arithmetic of the RHS: FP64 throughout, or FP32 FFT/Jacobian/N_j with
 * FP64 state, stage sums and error norm */
enum { QGN_FP64 = 0, QGN_MIXED = 1 };

/* Mirrors parameters.f90:1-31 (sizes become run-time values). */
typedef struct qgn_params {
  int nx, ny, nz;      /* nx even; nx/2 and ny products of 2, 3, 5 */
  double Lx, Ly, Htot; /* m */
  double beta;         /* 1/(m s) */
  double A2, A8;       /* m^2/s, m^8/s */
  double r0;           /* Ekman: drag on layer nz is r0*Htot/H(nz) */
  double C_d;          /* quadratic drag */
  double TOL;          /* adaptive-step error tolerance (max-norm, PV units) */
  double dt_max;       /* s; parameters.f90:31 uses pi/(beta*Lx) */
  int is_explicit;     /* parameters.f90:28 */
  int scheme;          /* QGN_AB3 .. QGN_ARK43_QGLEITH */
  int precision;       /* QGN_FP64 or QGN_MIXED */
  int max_rejects;     /* bound on the reject loop; <= 0 means 1000 */
} qgn_params;

typedef struct qgn_ctx qgn_ctx;

/* Module set-up.  Replaces the vertical-structure and plan parts of
 * Initialize (QGN_Module.f90:53-175).  Sizes: nx even, nx/2 and ny
 * products of 2, 3 and 5, nx <= 2048, ny <= 4096.  nx/2 and ny in {32, 48,
 * 64, 96, 128, 192, 256, 384, 512, 768, 1024, 1536, 2048} with nx, ny <=
 * 2048 run the compiled column and row passes; every other size runs a
 * generic path of run-time-plan transforms (the same results, more HBM
 * traffic, no tuned tiles and no closure network).  H (nz, already rescaled to Htot),
 * Modes (nz x nz, column-major, Modes(:,k) = mode k), EVals (nz), uBar (nz),
 * qyBar (nz) are the module variables of QGN_Module.f90:22-26. */
int qgn_create(qgn_ctx** ctx, const qgn_params* p, const double* H, const double* Modes,
               const double* EVals, const double* uBar, const double* qyBar);
/* Cleanup (QGN_Module.f90:448-458).  Returns QGN_EINVAL and destroys
 * nothing while the context belongs to an ensemble. */
int qgn_destroy(qgn_ctx* ctx);

/* ---- Batches of members (ensembles) ------------------------------------
 * No counterpart in the reference, which runs one model per process.
 * A batch stacks `members` models of one grid size along z: each member has
 * p->nz layers and its own vertical structure (H, EVals, uBar, qyBar: nz
 * values per member; Modes: nz*nz per member, member after member) and its
 * own Ekman coefficient r0[m] (NULL: p->r0 for all); the other parameters
 * are shared.  State arrays of the other entry points then hold
 * members*nz layers, member after member.  qgn_create is the batch of one. */
int qgn_create_batch(qgn_ctx** ctx, const qgn_params* p, int members, const double* H, const double* Modes,
                     const double* EVals, const double* uBar, const double* qyBar, const double* r0);

/* Per-member time and step size (arrays of `members`); resets the
 * controller memory to FirstCall (err0 = TOL) and the counters. */
int qgn_batch_set_time(qgn_ctx* ctx, const double* t, const double* dt);

/* Enqueue `attempts` ARK43 / ARK43_QGLeith attempts of every member with
 * the accept/reject and PI.3.4 control of ARK43*.f90 on the device, each
 * member with its own dt: an accepted attempt advances that member, a
 * rejected one shrinks its dt.  The attempt is captured once into a HIP
 * graph (and a graph of 8 chained attempts, launched once per 8); this call
 * only launches them and returns without waiting. */
int qgn_batch_step(qgn_ctx* ctx, int attempts);

/* Wait for the context's work and read per-member t, dt, accepted steps,
 * rejected attempts and status (0 ok, 1 non-finite error norm, 2 more than
 * max_rejects consecutive rejections; such a member stops advancing).  Any
 * argument may be NULL.  Returns QGN_ENONFINITE / QGN_EREJECT if a member
 * has that status. */
int qgn_batch_get_time(qgn_ctx* ctx, double* t, double* dt, long long* steps, long long* rejects, int* status);

/* Several batches (e.g. one per grid size) advanced together by graph
 * launches (one per 8 attempts), the batches' attempt graphs as independent
 * children that may run concurrently (so no two batches may share a
 * closure network).  Each
 * context needs qgn_batch_set_time first; qgn_ensemble_step is ordered
 * after the work already queued on every context and before what is
 * queued on them afterwards, and returns without waiting. */
typedef struct qgn_ens qgn_ens;
/* The ensemble's graph holds its contexts (qgn_destroy refuses them until
 * qgn_ensemble_destroy) and the closure networks they had attached when it
 * was created (which keep their buffers until qgn_ensemble_destroy, even
 * after qgn_set_closure(ctx, NULL, ...)). */
int qgn_ensemble_create(qgn_ens** ens, qgn_ctx* const* ctxs, int n);
/* `attempts` attempts of every member: one launch of a graph of 8 chained
 * attempts per 8, one single-attempt launch per remaining attempt. */
int qgn_ensemble_step(qgn_ens* ens, int attempts);
int qgn_ensemble_destroy(qgn_ens* ens);

/* Initialize's transform of the initial condition (QGN_Module.f90:177-183):
 * forward r2c of qp real(nx,ny,nz), DeAlias, store as the device state.
 * q_hat (may be NULL) receives the dealiased spectrum. */
int qgn_initialize(qgn_ctx* ctx, const double* qp, double* q_hat);

/* Set / read the prognostic spectrum q_hat complex(nx/2+1,ny,nz).
 * qgn_set_state applies DeAlias (the state is always dealiased). */
int qgn_set_state(qgn_ctx* ctx, const double* q_hat);
int qgn_get_state(qgn_ctx* ctx, double* q_hat);

/* TimeStep (QGN_Module.f90:302-309) with the scheme chosen at create:
 * advances the device state by one accepted step, updates t and dt. */
int qgn_time_step(qgn_ctx* ctx, double* t, double* dt);

/* GetRHS (QGN_Module.f90:258-296): rhs = RHS(q_hat), both
 * complex(nx/2+1,ny,nz) on the host.  The input is dealiased first, the
 * precondition the reference states for Jacobian (QGN_Module.f90:213). */
int qgn_rhs(qgn_ctx* ctx, const double* q_hat, double* rhs);

/* GetPsi (QGN_Module.f90:315-341) on the host: psi_hat from q_hat. */
int qgn_get_psi(qgn_ctx* ctx, const double* q_hat, double* psi_hat);

/* WriteQP's fields (QGN_Module.f90:348-361) of the current state:
 * q_phys and psi_phys real(nx,ny,nz) = c2r(.)/(nx*ny).  Either may be NULL. */
int qgn_fields(qgn_ctx* ctx, double* q_phys, double* psi_phys);

/* WriteQPmodal's fields (QGN_Module.f90:381-416) of the current state:
 * modal q, psi and Jacobian (the full, not dealiased, flux-form Jacobian),
 * each real(nx,ny,nz) = c2r(.)/(nx*ny) with mode k in slot k.  Any may be
 * NULL.  Allocates its work arrays on first use. */
int qgn_modal_fields(qgn_ctx* ctx, double* q_mode, double* psi_mode, double* j_mode);

/* Isotropic spectra of the current state, per layer, each (nbins, nz) with
 * the bin fastest: ke = 1/2 sum k^2 |psi_hat|^2 and ens = 1/2 sum |q_hat|^2
 * over the wavenumbers of the full plane with floor(|k|/dk + 1/2) = bin,
 * dk = min(2 pi/Lx, 2 pi/Ly), normalised by (nx ny)^2 so that the sums over
 * bins are the domain means of (u^2+v^2)/2 and q^2/2 of the physical fields.
 * Wavenumbers beyond the last bin are dropped; 1 <= nbins <= 4096.  No
 * counterpart in the reference: QGN_Driver.f90:52-55 leaves this
 * diagnostics hook to the user. */
int qgn_spectra(qgn_ctx* ctx, int nbins, double* ke, double* ens);

/* QG-Leith coefficients A8L(nz) of the current state
 * (ARK43_QGLeith.f90:62-73). */
int qgn_leith(qgn_ctx* ctx, double* a8l);

typedef struct qgn_stats {
  long long steps;      /* accepted steps */
  long long attempts;   /* accepted + rejected */
  long long rejects;
  long long rhs_evals;
  double last_err;      /* err1 of the last attempt */
  double err0;          /* controller memory */
} qgn_stats;
int qgn_get_stats(qgn_ctx* ctx, qgn_stats* s);

/* The HIP stream (hipStream_t) every kernel of the context runs on. */
int qgn_get_stream(qgn_ctx* ctx, void** stream);

/* Launch choices of the mixed path's FFT passes.  At creation a mixed
 * context picks the tile width of each column pass (2, 4, 8, 16 columns;
 * inverse and forward separately) and the row group of the fused row pass
 * (2, 4, 8 rows) by timing whole GetRHS evaluations on its own buffers, one
 * pass's candidates at a time with the others at their current picks, two
 * sweeps, a candidate displacing the current pick only when 2% faster
 * (QGN_AUTOTUNE=0 in the environment keeps the fixed rules: 8 columns for
 * the inverse and 4 for the forward column pass up to ny = 1024, 4 rows).  qgn_get_tuning returns the choices (0 = fixed rule)
 * and, in times_us[3][4], the GetRHS time of each candidate in the last
 * sweep (column inverse and forward by tile width index, row pass by group
 * index); qgn_set_tuning forces choices.  No counterpart in the reference. */
int qgn_get_tuning(qgn_ctx* ctx, int* col_inv, int* col_rhs, int* row, float* times_us);
int qgn_set_tuning(qgn_ctx* ctx, int col_inv, int col_rhs, int row);

/* Kernel timing with HIP events on the context stream: after
 * qgn_prof_enable(ctx, id), every launch of kernel `id` (QGN_K_ALL: of every
 * kernel below; -1: none) outside a captured graph is bracketed by events;
 * qgn_prof_read returns the launch count and summed milliseconds since
 * enable, qgn_prof_read_all the same per kernel id (arrays of QGN_K_COUNT);
 * both synchronise the stream.  No counterpart in the reference. */
/* Kernel ids: the three FFT passes of GetRHS; the projection (GetPsi: on
 * the mixed path with nz <= 16 the MFMA projection of every GetRHS, else the
 * projection of q); the stage kernels (the ARK stage sums, on the FP64 path
 * fused with their projection; on the mixed path with nz <= 16 also the
 * FP32 copy of q before the first GetRHS); the candidate update; the error
 * norm's transform. */
enum { QGN_K_ALL = -2, QGN_K_ROW_JAC = 0, QGN_K_COL_INV = 1, QGN_K_COL_RHS = 2, QGN_K_PROJECT = 3,
       QGN_K_STAGE = 4, QGN_K_UPDATE = 5, QGN_K_ERRNORM = 6, QGN_K_COUNT = 7 };
int qgn_prof_enable(qgn_ctx* ctx, int kernel_id);
int qgn_prof_read(qgn_ctx* ctx, int* count, double* total_ms);
int qgn_prof_read_all(qgn_ctx* ctx, int* counts, double* ms);

/* ---- MKL DFTI replacement (MKL_DFTI_Setup.f90:1-52) --------------------
 * Batched 2-D real transforms with the reference's descriptor:
 * DFTI_REAL, lengths [nx,ny], NUMBER_OF_TRANSFORMS = howmany,
 * CCE half spectrum along x, out of place, unnormalised both ways,
 * input distance nx*ny / output distance (nx/2+1)*ny.
 * Backward transforms use only the real part of the kx=0 and kx=nx/2 bins
 * after the y inverse, as DFTI does for non-Hermitian input. */
typedef struct qgn_dft qgn_dft;
int qgn_dft_create(qgn_dft** h, int precision, int nx, int ny, int howmany);
/* DftiComputeForward(g2s_q, gridE, specE), host arrays */
int qgn_dft_forward(qgn_dft* h, const double* grid, double* spec);
/* DftiComputeBackward(s2g_q, specE, gridE), host arrays */
int qgn_dft_backward(qgn_dft* h, const double* spec, double* grid);
/* device-pointer variants on a caller stream (NULL = the handle's stream);
 * precision QGN_MIXED handles take float data */
int qgn_dft_forward_device(qgn_dft* h, const void* grid, void* spec, void* stream);
int qgn_dft_backward_device(qgn_dft* h, const void* spec, void* grid, void* stream);
int qgn_dft_free(qgn_dft* h);

/* The same transforms behind DFTI's descriptor calls, in the order
 * MKL_DFTI_Setup.f90:11-50 makes them (create, set values, commit) and as
 * QGN_Module.f90 computes and frees them (DftiComputeForward at :180,
 * :238-241, :287-290; DftiComputeBackward at :228-234; DftiFreeDescriptor
 * at :452-455).  Lengths are in Fortran order {nx, ny}; strides are
 * {offset, x stride, y stride} as the reference sets them.  Commit accepts
 * the descriptions of real(nx,ny,howmany) <-> complex(nx/2+1,ny,howmany)
 * arrays, x fastest, out of place, unnormalised, and returns QGN_EINVAL
 * with a message for anything else; a compute call checks that the strides
 * fit its direction.  Host arrays, DFTI_DOUBLE data. */
typedef struct qgn_dfti qgn_dfti;
enum { QGN_DFTI_DOUBLE = 1, QGN_DFTI_REAL = 2, QGN_DFTI_COMPLEX_COMPLEX = 3, QGN_DFTI_CCE_FORMAT = 4,
       QGN_DFTI_NOT_INPLACE = 5, QGN_DFTI_INPLACE = 6 };
enum { QGN_DFTI_NUMBER_OF_TRANSFORMS = 1, QGN_DFTI_INPUT_DISTANCE = 2, QGN_DFTI_OUTPUT_DISTANCE = 3,
       QGN_DFTI_CONJUGATE_EVEN_STORAGE = 4, QGN_DFTI_PACKED_FORMAT = 5, QGN_DFTI_PLACEMENT = 6,
       QGN_DFTI_INPUT_STRIDES = 7, QGN_DFTI_OUTPUT_STRIDES = 8, QGN_DFTI_FORWARD_SCALE = 9,
       QGN_DFTI_BACKWARD_SCALE = 10 };
int qgn_dfti_create_descriptor(qgn_dfti** h, int precision, int domain, int dimension, const long long* lengths);
int qgn_dfti_set_value_int(qgn_dfti* h, int param, long long value);
int qgn_dfti_set_value_ints(qgn_dfti* h, int param, const long long* values, int n);
int qgn_dfti_set_value_double(qgn_dfti* h, int param, double value);
int qgn_dfti_commit_descriptor(qgn_dfti* h);
int qgn_dfti_compute_forward(qgn_dfti* h, const double* grid, double* spec);
int qgn_dfti_compute_backward(qgn_dfti* h, const double* spec, double* grid);
int qgn_dfti_free_descriptor(qgn_dfti** h);

/* The transforms of qgn_dft_forward_device / qgn_dft_backward_device
 * (forward != 0: r2c, DftiComputeForward; else c2r, DftiComputeBackward)
 * without a handle, allocating nothing: for callers that cannot allocate
 * device memory (work recorded under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc).  No counterpart in the
 * reference.  in/out are device arrays in the DFTI layout of the handle
 * variants (float data for QGN_MIXED); stream NULL is the null stream; the
 * call returns without waiting.  The intermediate spectrum and twiddle
 * tables live in a fixed 64 MiB device array of the library on each GPU
 * (held until the process ends): howmany*ny*(nx/2+1) complex values plus
 * the tables must fit it, else QGN_EINVAL (1024^2 FP64: up to 7 transforms).
 * Calls on one GPU share that array, so calls that overlap on one GPU
 * race. */
int qgn_dft_compute_noalloc(int precision, int forward, int nx, int ny, int howmany, const void* in, void* out,
                            void* stream);

/* ---- LAPACK DGEEV replacement for the vertical structure ---------------
 * QGN_Module.f90:84-148: H <- Htot*H/sum(H); tridiagonal stretching matrix
 * L from S(0..nz) (only S(1..nz-1) used); eigenpairs of L (real, since L is
 * similar to a symmetric matrix), each mode normalised so that
 * sum(H*Modes(:,k)**2)/Htot = 1 with Modes(1,k) > 0; qyBar = -L uBar.
 * Modes are returned in descending eigenvalue order (barotropic first);
 * DGEEV's order is unspecified and only changes WriteQPmodal's mode order. */
int qgn_vertical_modes(int nz, double Htot, const double* H_in, const double* S, const double* uBar,
                       double* H_out, double* L, double* Modes, double* EVals, double* qyBar);

/* ---- Learned subgrid closure: a 4-layer 3x3 CNN ------------------------
 * No counterpart in the reference (whose hook for a closure is the QG-Leith
 * term, ARK43_QGLeith.f90:62-79).  Layers cin -> 32 -> 32 -> 32 -> cout with
 * ReLU between them, periodic padding; BF16 weights and activations with
 * FP32 accumulation on the matrix cores, FP32 master weights and Adam.
 * cin in {16, 32}, cout in {8, 16}.  Tensors are (n, H, W, C), channel
 * fastest, with W % 16 == 0 and n*H*W % 64 == 0.
 * params: FP32, for each layer W[co][tap][ci] then b[co], tap = 3 (dy+1) +
 * (dx+1) for the input offset (dy, dx); qgn_cnn_nparams gives the length. */
typedef struct qgn_cnn qgn_cnn;
int qgn_cnn_create(qgn_cnn** h, int cin, int cout, const float* params);
/* Returns QGN_EINVAL and destroys nothing while a model has the network
 * attached (qgn_set_closure) or an ensemble holds it. */
int qgn_cnn_destroy(qgn_cnn* h);
int qgn_cnn_nparams(qgn_cnn* h, int* n);
/* host copies of the parameters and of the last gradients (either may be NULL) */
int qgn_cnn_get_params(qgn_cnn* h, float* params, float* grads);
int qgn_cnn_set_params(qgn_cnn* h, const float* params);
/* device pointers: x bf16 bit patterns (n, H, W, cin), y fp32 (n, H, W, cout).
 * The network's kernels run on its own stream, ordered with no other
 * stream: device inputs (x, target) must be complete when a call is made
 * (synchronise the stream that wrote them), and each call has finished
 * its device work when it returns. */
/* While the network is attached to a model or held by an ensemble, its
 * activation buffers are fixed at the attached model's pixel count: a pass
 * on more pixels returns QGN_EINVAL instead of reallocating them. */
int qgn_cnn_forward(qgn_cnn* h, const void* x, int n, int H, int W, float* y);
/* gradients of the mean-squared error of the last forward's y against the
 * device array target (n, H, W, cout); *loss receives the error.  The
 * backward pass reads the x and y of that forward: they must stay allocated
 * and unchanged until it returns.  After a closure evaluation (an RHS of a
 * model that has the network attached) the last forward is the closure's,
 * and backward returns QGN_EINVAL until qgn_cnn_forward runs again. */
int qgn_cnn_backward(qgn_cnn* h, const float* target, double* loss);
/* one Adam step (bias-corrected) with the last gradients */
int qgn_cnn_adam(qgn_cnn* h, float lr, float beta1, float beta2, float eps);
/* forward + backward + Adam(lr, 0.9, 0.999, 1e-8) */
int qgn_cnn_train_step(qgn_cnn* h, const void* x, const float* target, int n, int H, int W, float* y, float lr,
                       double* loss);
const char* qgn_cnn_last_error(void);

/* Attach the network as a learned subgrid PV forcing (NULL detaches): every
 * RHS evaluation of the context (qgn_rhs and every stage of every stepper)
 * adds r2c(F) on the kept modes, where F = out_scale * net(in_scale * (q,
 * psi)) is evaluated on the physical fields of each member (input channels
 * q_1..q_nz, psi_1..psi_nz; output one forcing per layer).  The network
 * needs cin = 2 nz and cout = nz; nx % 16 == 0 (and nx <= 1024 on the
 * FP64 path).  The context keeps the pointer and holds the network until it
 * is detached or the context destroyed (qgn_cnn_destroy refuses a held
 * network).  An ensemble built before this call keeps the attempt it
 * captured, and with it its hold on the network it had. */
int qgn_set_closure(qgn_ctx* ctx, qgn_cnn* net, const double* in_scale, const double* out_scale);

const char* qgn_last_error(void);
const char* qgn_version(void);

#ifdef __cplusplus
}
#endif
#endif /* QGN_H */
