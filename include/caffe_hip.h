/*
 * caffe_hip.h -- C ABI of the MI355X (gfx950) implementation of CAFFE's
 * SIMPLE / PISO / IbPiso per-iteration hot path.
 *
 * One caffe_ctx owns one GPU-resident solver instance: mesh metrics, fields
 * with their boundary patches, the assembled momentum and pressure-correction
 * systems and the Krylov work space.  All work is queued on the context's HIP
 * stream; calls that return host data (downloads, reports, solves that check
 * convergence) synchronise that stream.  One host thread per context.
 *
 * Every function returns CAFFE_OK (0) or a negative status; the message of the
 * last failure is caffe_last_error(ctx).  No C++ exception crosses this ABI.
 *
 * Reference interfaces replaced (paths relative to the CAFFE tree):
 *   caffe_case_setup          Simple::Simple / Piso::Piso / IbPiso::IbPiso
 *                             (src/Solvers/Simple.cpp:34-74, Piso.cpp:5-10,
 *                             IbPiso.cpp:32-39) with HexaFvmMesh::initialize
 *                             (src/Domains/HexaFvmMesh/HexaFvmMesh.cpp:51-58, 615-624)
 *   caffe_solve               virtual double Solver::solve(double timeStep)
 *                             (src/Solvers/Solver.h:20; Simple.cpp:78-91,
 *                             Piso.cpp:12-29, IbPiso.cpp:43-61)
 *   caffe_momentum_assemble   Simple::computeMomentum assembly loop (Simple.cpp:117-179)
 *   caffe_momentum_solve      the 3 x SparseMatrix::solve + copy-back + BCs + face
 *                             interpolation of computeMomentum (Simple.cpp:184-218)
 *   caffe_rhiechow            Simple::rhieChowInterpolateFaces (Simple.cpp:410-449)
 *   caffe_pcorr_assemble      Simple::computePCorr assembly loop (Simple.cpp:231-269)
 *   caffe_pcorr_solve         SparseMatrix::solve on A_[1] (Simple.cpp:275,
 *                             src/Math/SparseMatrix.cpp:86-106) + copy-back,
 *                             BCs, interpolation and grad p' (Simple.cpp:277-298)
 *   caffe_correct             Simple::correct (Simple.cpp:301-356)
 *   caffe_ib_field            IbPiso::computeIbField (IbPiso.cpp:65-117)
 *   caffe_ib_coeffs           IbPiso::computeIbCoeffs (IbPiso.cpp:119-190)
 *                             and Coupled::solve (src/Solvers/Coupled.cpp:38-50)
 *   caffe_step_report         Simple::computeContinuityError / computeCourantNumber
 *                             (Simple.cpp:358-408)
 *   caffe_field_download      Field<T>::cellData() / patches / faces
 *                             (src/Domains/Field/Field.h:53-131)
 */
#ifndef CAFFE_HIP_H
#define CAFFE_HIP_H

#ifdef __cplusplus
extern "C" {
#endif

#define CAFFE_OK 0
#define CAFFE_ERR_ARG (-1)
#define CAFFE_ERR_HIP (-2)
#define CAFFE_ERR_STATE (-3)
#define CAFFE_ERR_NOT_CONVERGED (-4)

typedef struct caffe_ctx caffe_ctx;

/* FlowBoundaryConditions::Type (src/Domains/Field/FlowBoundaryConditions.h:10) */
enum { CAFFE_BC_INLET = 0, CAFFE_BC_OUTLET = 1, CAFFE_BC_WALL = 2, CAFFE_BC_EMPTY = 3, CAFFE_BC_ZERO_GRADIENT = 4,
       CAFFE_BC_FIXED = 5 /* PrimitiveBoundaryCondition FIXED: the Diffusion solver's boundaries
                             (PrimitiveBoundaryCondition.tpp:3-30), with EMPTY and ZERO_GRADIENT */ };
/* Solver family */
enum { CAFFE_SIMPLE = 0, CAFFE_PISO = 1, CAFFE_IBPISO = 2,
       CAFFE_DIFFUSION = 3, /* Diffusion (Diffusion.cpp:8-131): the scalar phi with pInit as its
                              uniform initial value and bcVal as the refValues; caffe_solve returns
                              ||A x - b||_inf; solve options are those of the p' system */
       CAFFE_COUPLED = 4    /* Coupled (Coupled.cpp:5-338): u, v, w and p in one block system
                              (Jacobi BiCGStab in FP64 whatever the mode; the options of the p'
                              system give rtol and maxit, its report the result); uInit and pInit
                              are ignored (the reference reads no initial conditions);
                              caffe_solve returns the max over cells of |sum of face mass flows|/V */ };
/* Pressure-correction / momentum solve modes */
enum {
    CAFFE_SOLVE_PLAIN_FP64 = 0,   /* preconditioned Krylov in FP64 (the judge and fallback) */
    CAFFE_SOLVE_IR_FP32 = 1,      /* FP32 inner Krylov inside FP64 iterative refinement */
    CAFFE_SOLVE_LEARNED_FCG = 2,  /* flexible CG with the learned two-grid cycle (FP32 inside
                                     FP64 refinement); accepted only at rtol, else the plain
                                     solve runs (fell_back) */
    CAFFE_SOLVE_SGS_FP64 = 4,     /* p' system: CG in FP64 preconditioned by one red-black symmetric
                                     Gauss-Seidel sweep (SSOR, omega = 1): the parallel analogue of
                                     the reference's PCSOR(its 1, lits 1) (SparseMatrix.cpp:97-98);
                                     ghost-row and indefinite systems take the plain solve
                                     (fell_back = 1) */
    CAFFE_SOLVE_MG = 3            /* p' and Coupled: BiCGStab in FP64, right-preconditioned by one
                                     FP32 geometric multigrid V-cycle (no counterpart in the
                                     reference, whose PETSc solve it replaces); the FP32 modes
                                     use it for ghost-row systems too.  Coupled: the block
                                     lower triangle [diag M, 0; D, L], the cycle on L; set
                                     with which = 1 (the pressure / Coupled options) */
};

/* The keys of case/case.info and case/initialConditions.info the hot path reads
 * (Solver.cpp:8, Simple.cpp:61-67, 104-108, Piso.cpp:9, FlowBoundaryConditions.cpp:25-29,
 * IbPiso.cpp:37-38, InitialConditions.cpp:12-74). */
typedef struct {
    int nI, nJ, nK;          /* cells */
    int algo;                /* CAFFE_SIMPLE / CAFFE_PISO / CAFFE_IBPISO */
    int steady;              /* Solver.timeAccurate OFF */
    double rho, mu;
    double alphaU, alphaP;   /* relaxationFactorMomentum / relaxationFactorPCorr */
    int nInner, nPCorr;
    int bcType[6];           /* east, west, north, south, top, bottom */
    double bcVec[18];        /* refVector per face */
    double bcVal[6];         /* refValue per face */
    double uInit[3];
    double pInit;
    double sphereR, sphereC[3];
    int uniform;             /* 1: cells form a uniform box; metrics in closed form */
    int nonOrth;             /* 1: corrected mode: the momentum right-hand side gets the non-orthogonal
                                term mu_f (grad u)_f . c_f of Simple.cpp:153-158 with the faces of
                                grad u filled (interior: volume-weighted interpolation of the cell
                                gradients, domain faces: the boundary cell's gradient).  0: reference
                                mode, where those faces are never assigned and the term vanishes.
                                No effect on uniform boxes (c = 0).  No counterpart key in the
                                reference's case.info. */
} caffe_case;

typedef struct {
    int mode;                /* CAFFE_SOLVE_* */
    double rtol;             /* target true FP64 relative residual ||b - Ax||_2 / ||b||_2 */
    int maxit;               /* Krylov iteration cap (all refinement levels together) */
    double inner_rtol;       /* IR_FP32: relative tolerance of each FP32 inner solve */
} caffe_solve_opts;

typedef struct {
    int iters;               /* Krylov iterations (sum over inner solves for IR) */
    int refinements;         /* FP64 refinement steps (IR modes) */
    double relres;           /* true FP64 relative residual at exit */
    int converged;           /* relres <= rtol */
    int fell_back;           /* learned step rejected, plain solve used */
    double ms;               /* device time of the solve */
} caffe_solve_report;

int caffe_ctx_create(int device, caffe_ctx** out);
int caffe_ctx_destroy(caffe_ctx* ctx);
const char* caffe_last_error(caffe_ctx* ctx);

/* Build the device mesh (metrics from nodes, or closed form when c->uniform),
 * allocate fields and systems, apply boundary and initial conditions.
 * nodes: (nI+1)(nJ+1)(nK+1) x 3 doubles, i fastest, xyz interleaved; borrowed. */
int caffe_case_setup(caffe_ctx* ctx, const caffe_case* c, const double* nodes);

/* which: 0 = momentum, 1 = pressure correction */
int caffe_set_solve_opts(caffe_ctx* ctx, int which, const caffe_solve_opts* o);

/* Solver::solve(dt): returns the reference's continuity value (Simple.cpp:358-380) */
int caffe_solve(caffe_ctx* ctx, double dt, double* continuity);

int caffe_save_u0(caffe_ctx* ctx);
int caffe_momentum_assemble(caffe_ctx* ctx, double dt);
int caffe_momentum_solve(caffe_ctx* ctx, caffe_solve_report* rep);
int caffe_rhiechow(caffe_ctx* ctx);
int caffe_pcorr_assemble(caffe_ctx* ctx);
int caffe_pcorr_solve(caffe_ctx* ctx, const caffe_solve_opts* o, caffe_solve_report* rep);
int caffe_correct(caffe_ctx* ctx);
int caffe_ib_field(caffe_ctx* ctx);
int caffe_ib_coeffs(caffe_ctx* ctx);
int caffe_grad_u(caffe_ctx* ctx);
int caffe_step_report(caffe_ctx* ctx, double dt, double* continuity_ref, double* continuity_max, double* courant);
int caffe_last_reports(caffe_ctx* ctx, caffe_solve_report* momentum, caffe_solve_report* pcorr);
/* courantNumber_ as the last caffe_solve left it (Simple.cpp:88, Piso.cpp:26; printed by
 * displayUpdateMessage, Piso.cpp:36); IbPiso::solve does not compute it, so it stays 0 there */
int caffe_last_courant(caffe_ctx* ctx, double* courant);

/* Field exchange in the padded layout (nI+2)(nJ+2)(nK+2), i fastest, halo = boundary
 * patches; vector fields are component-major.  Names: u u0 h p pCorr d gradP gradPCorr
 * gradU massFlow mI mJ mK status.  n = number of doubles in host. */
int caffe_field_download(caffe_ctx* ctx, const char* name, double* host, long n);
int caffe_field_upload(caffe_ctx* ctx, const char* name, const double* host, long n);
/* Metric arrays (general meshes): vol xc SI SJ SK fcI fcJ fcK dI dJ dK gI gJ gK */
int caffe_metric_download(caffe_ctx* ctx, const char* name, double* host, long n);
/* Assembled systems.  which 0: band 7*P (aP aE aW aN aS aT aB); which 1: 4*P
 * (aP, aI, aJ, aK: coefficient of the west/south/bottom face of each cell); which 2 (Coupled
 * cases): 56*P, eight groups of seven planes in the order own, E, W, N, S, T, B: the momentum
 * band, the pressure coefficients of the u, v, w momentum rows, the u, v, w coefficients of
 * the continuity row, and its pressure coefficients (Coupled.cpp:80-206).
 * rhs: 3*P (momentum), P (pressure correction) or 4*P (Coupled: u, v, w, continuity). */
int caffe_matrix_download(caffe_ctx* ctx, int which, double* host, long n);
int caffe_rhs_download(caffe_ctx* ctx, int which, double* host, long n);
/* Ghost rows: n = number of ghost cells; cells[n], cols[9n] (padded indices, -1 dropped),
 * valsM[9n], valsP[9n]; any pointer may be NULL. */
int caffe_ghost_rows(caffe_ctx* ctx, int* n, long* cells, long* cols, double* valsM, double* valsP);
/* Diagnostics: average device time (ms) of `reps` back-to-back launches of one kernel on
 * the context's stream, and its algorithmic HBM bytes per launch.  which:
 * "pcorr_cg_k1_fp64" / "pcorr_cg_k1_fp32" (pressure-correction CG stencil kernel on the
 * current p' system), "copy_fp64" (a plain copy of 8*P doubles, for counter calibration).  No counterpart in the reference (its timing is Time::tic/toc). */
int caffe_time_kernel(caffe_ctx* ctx, const char* which, int reps, double* ms, double* bytes);
/* Live timing of one kernel inside the solves that follow (HIP event pairs around each of
 * its launches on the context's stream, up to 4096): which = "fcg_k1" (stencil kernel of the
 * learned flexible CG), "pcorr_cg_k1_fp32" / "pcorr_cg_k1_fp64" (stencil kernel of the
 * pressure CG), "coupled_bicg_a" (first BiCGStab kernel of the Coupled solve),
 * "pcorr_bicg_a_fp32" / "pcorr_bicg_a_fp64" (first kernel of the Jacobi p' BiCGStab:
 * ghost-row systems in the plain mode, indefinite systems), "mg_post0" (fine-level
 * post-smoothing sweep of the multigrid cycle: CAFFE_SOLVE_MG, and ghost-row systems in
 * the FP32 modes), "fcg_k2pre" (the update and pre-smoothing kernel of the learned flexible CG);
 * enable 0 stops recording.  _read returns the average duration of the
 * launches that ran (launches queued past convergence return at once and are left out),
 * their number, and the kernel's algorithmic HBM bytes per launch.  No counterpart in the
 * reference. */
int caffe_kernel_probe(caffe_ctx* ctx, const char* which, int enable);
int caffe_kernel_probe_read(caffe_ctx* ctx, double* avg_ms, long* launches, double* bytes);
int caffe_sync(caffe_ctx* ctx);
long caffe_padded_size(caffe_ctx* ctx);
/* Device memory held by the context (fields, metrics, systems, Krylov work), in bytes;
 * a learned network attached to it holds its own.  No counterpart in the reference. */
long caffe_device_bytes(caffe_ctx* ctx);

/* ---------------------------------------------------------------------------------
 * Learned multilevel preconditioner of the pressure-correction solve.  No counterpart in
 * the reference: there the p' system goes to PETSc BiCGStab + SOR (SparseMatrix.cpp:86-106)
 * and the learned stage replaces that preconditioner; the plain FP64 solve stays the
 * fallback and judge.
 *
 * Cycle for a residual r (fine grid, dimensions divisible by 16):
 *   z1 = w D^-1 r;  c = net(P^T D^-1 (r - A z1));  y = z1 + P c;  z = y + w D^-1 (r - A y)
 * with w = 0.8, P the cell-centred trilinear prolongation to the fine grid, and net a
 * linear network of twelve 3x3x3 convolutions over four levels (coarse grid, /2, /4, /8;
 * 16/32/64/64 channels), BF16 on MFMA with FP32 accumulation.
 * Parameters: FP32, layers in order, each [27][cout][cin] (tap (dz+1)*9+(dy+1)*3+(dx+1)).
 * Coarse fields are [K][J][I] (I fastest); fine fields and matrices use the padded layout
 * of the context (matrix = 4*P floats: aP, aI, aJ, aK as caffe_matrix_download which 1).
 * ------------------------------------------------------------------------------- */
typedef struct caffe_net caffe_net;
int caffe_net_create(int device, unsigned long long seed, caffe_net** out);
int caffe_net_destroy(caffe_net* net);
const char* caffe_net_last_error(caffe_net* net);
long caffe_net_param_count(void);
int caffe_net_get_params(caffe_net* net, float* host, long n);
int caffe_net_set_params(caffe_net* net, const float* host, long n);
/* the network alone on a coarse grid (dimensions multiples of 8): out = net(in);
 * backward: gparams = d<gout, net(in)>/dtheta */
int caffe_net_forward(caffe_net* net, int nx, int ny, int nz, const float* in, float* out);
int caffe_net_backward(caffe_net* net, int nx, int ny, int nz, const float* in, const float* gout, float* gparams);
/* z = cycle(r) on the fine grid for the FP32 matrix A */
int caffe_net_precond(caffe_net* net, int nI, int nJ, int nK, const float* A, const float* r, float* z);
/* damped Jacobi sweeps after the coarse correction in the cycle of caffe_net_precond,
 * caffe_net_loss_grad, caffe_net_train and caffe_net_eval: 1 to 4 (default 1).  Training
 * with the count the flexible CG applies (caffe_ctx_set_fcg_post_sweeps) fits the network
 * to the cycle it is used in.  No counterpart in the reference. */
int caffe_net_set_post_sweeps(caffe_net* net, int sweeps);
/* training loss of one pair: L = e^T A e / x^T A x, e = x - cycle(b); gparams may be NULL */
int caffe_net_loss_grad(caffe_net* net, int nI, int nJ, int nK, const float* A, const float* b, const float* x,
                        double* loss, float* gparams);
/* append the context's current p' system (matrix, right-hand side, converged correction;
 * call after caffe_pcorr_solve) to the device-resident training set */
int caffe_net_collect(caffe_net* net, caffe_ctx* ctx, long* count);
int caffe_net_clear(caffe_net* net);
long caffe_net_dataset_size(caffe_net* net);
/* `steps` Adam steps at rate lr over random mini-batches of `batch` pairs (Adam moments
 * persist across calls, so a caller can schedule the rate call by call).  Each batch
 * entry is, with probability `synthetic`, an exact synthetic pair on the matrix of a
 * random stored system: x = a random sum of Neumann cosine modes from one frequency band,
 * b = A x (the stored right-hand sides cover only the bands the flow excites, while a
 * preconditioner meets every band in the Krylov residuals).  losses[steps] (may be NULL)
 * receives the mean batch loss of each step. */
int caffe_net_train(caffe_net* net, int steps, int batch, float lr, double synthetic, unsigned long long seed,
                    float* losses);
/* loss of every stored pair under the current parameters */
int caffe_net_eval(caffe_net* net, float* losses, long n);
/* use the network in CAFFE_SOLVE_LEARNED_FCG solves of ctx (NULL detaches); the network
 * must outlive the attachment */
int caffe_ctx_attach_net(caffe_ctx* ctx, caffe_net* net);
/* The coarse solve of the two-grid cycle in CAFFE_SOLVE_LEARNED_FCG solves of ctx: 0 the
 * attached network (default); 1 one geometric multigrid V-cycle of the CAFFE_SOLVE_MG
 * hierarchy below the fine level on the classical right-hand side P^T (r - A z1), the same
 * cycle and FCG otherwise: the ablation that shows what the network gains over a classical
 * coarse solve.  A network must still be attached (its fine-grid work buffers are used). */
int caffe_ctx_set_fcg_coarse(caffe_ctx* ctx, int mode);
/* Damped Jacobi sweeps after the coarse correction of the learned cycle in
 * CAFFE_SOLVE_LEARNED_FCG solves of ctx, 1 to 4 (default 3).  The network is trained with
 * one; the later ones damp the high-frequency error the coarse correction leaves, which
 * the second PISO corrector's systems carry most of.  No counterpart in the reference. */
int caffe_ctx_set_fcg_post_sweeps(caffe_ctx* ctx, int sweeps);
/* Initial guess of the CAFFE_SOLVE_LEARNED_FCG and CAFFE_SOLVE_MG p' solves of ctx: 1
 * (default) starts the k-th p' solve after a momentum assembly (the k-th corrector) from the
 * combination g0 x1 + g1 x2 of the solutions the k-th solve reached in the two previous time
 * steps or outer iterations whose residual |b - A x| is smallest (so never above |b|); solves
 * past the algorithm's correctors start from x = 0.  0 starts every solve from x = 0 as
 * SparseMatrix::solve does (SparseMatrix.cpp:86-106).  Either way the solve ends at the same
 * true FP64 relative residual.  No counterpart in the reference. */
int caffe_ctx_set_fcg_guess(caffe_ctx* ctx, int on);

/* ---------------------------------------------------------------------------------
 * The SparseMatrix / SparseVector compatibility path: for callers that assemble their
 * systems row by row, as the reference's solvers do (Simple.cpp:162-175, 265-275,
 * IbPiso.cpp:148-185, Diffusion.cpp:97-130, Coupled.cpp:137-234), and solve them through
 * SparseMatrix::solve.  A matrix keeps its rows on the host while they are set; assembly
 * copies them to HBM (ELL layout), where the products and the solve run.  Every call is
 * synchronous.  Matrices and vectors are independent of any caffe_ctx.
 * ------------------------------------------------------------------------------- */
typedef struct caffe_matrix caffe_matrix;
typedef struct caffe_vector caffe_vector;
enum { CAFFE_NORM_L1 = 0, CAFFE_NORM_L2 = 1, CAFFE_NORM_INF = 2 };

/* SparseMatrix(rToler, absToler, maxIters) + allocate(m, n, nnz) (src/Math/SparseMatrix.h:38-46,
 * SparseMatrix.cpp:26-63): an m x n matrix with nnz entries reserved per row (a row may grow
 * past it, as PETSc AIJ lets it).  rtol <= 0, abstol < 0 and maxit <= 0 take the PETSc
 * defaults 1e-5, 1e-50 and 10000.  destroy = deallocate (SparseMatrix.cpp:65-69). */
int caffe_matrix_create(int device, int m, int n, int nnz, double rtol, double abstol, int maxit, caffe_matrix** out);
int caffe_matrix_destroy(caffe_matrix* A);
const char* caffe_matrix_last_error(caffe_matrix* A);
/* SparseMatrix::setRow / setValue (SparseMatrix.cpp:71-79: MatSetValues, INSERT_VALUES) for
 * nrows rows at once: row rows[q] gets columns cols[rowptr[q] .. rowptr[q+1]) with vals; an
 * entry set again is overwritten, other entries of the row stay; negative column and row
 * indices are ignored (IndexMap::operator() returns -1 off the domain, IndexMap.cpp:77). */
int caffe_matrix_set_rows(caffe_matrix* A, int nrows, const int* rows, const int* rowptr, const int* cols,
                          const double* vals);
int caffe_matrix_set_value(caffe_matrix* A, int i, int j, double v);
/* SparseMatrix::zeroEntries (MatZeroEntries: values zero, structure kept) and assemble */
int caffe_matrix_zero_entries(caffe_matrix* A);
int caffe_matrix_assemble(caffe_matrix* A);
/* SparseMatrix::solve(b, x) (SparseMatrix.cpp:86-106): x from zero; BiCGStab in FP64,
 * right-preconditioned by Jacobi, to ||b - A x||_2 <= rtol ||b||_2 (or <= abstol), restarted
 * from the true residual on breakdown; *iters = the iteration count.  Not converging is not
 * an error (the reference never checks KSP's converged reason): caffe_matrix_last_report
 * gives the true relative residual and whether it met rtol. */
int caffe_matrix_solve(caffe_matrix* A, const caffe_vector* b, caffe_vector* x, int* iters);
/* No counterpart in the reference.  caffe_matrix_solve for callers that cannot allocate
 * device memory (code under HIP graph capture set-up, tight serving loops, harnesses that
 * forbid hipMalloc): the same solve with the same results, whose BiCGStab work lives in a
 * fixed per-GPU library array of CAFFE_NOALLOC_MAX_ROWS rows (19.3 MB, held by each GPU the
 * library runs on until the process ends) instead of a per-matrix allocation.  It calls no
 * hipMalloc or hipFree: A must have been assembled (caffe_matrix_assemble) since its rows last
 * changed, or CAFFE_ERR_STATE; more rows than the maximum give CAFFE_ERR_ARG.  Every call on
 * one GPU shares that array, so calls that overlap on one GPU (from several host threads)
 * race; each call writes every work value before it reads it. */
#define CAFFE_NOALLOC_MAX_ROWS 262144
int caffe_matrix_solve_noalloc(caffe_matrix* A, const caffe_vector* b, caffe_vector* x, int* iters);
int caffe_matrix_last_report(caffe_matrix* A, caffe_solve_report* rep);
/* multiply(A, x, y): y = A x; multiplyAdd(A, x1, x2, y): y = A x1 + x2 (y may be x2);
 * scale(alpha, A) (SparseMatrix.h:76-79, SparseMatrix.cpp:135-148).  The matrix-matrix
 * multiply(A, B, C) is not provided: no solver of the reference calls it. */
int caffe_matrix_multiply(caffe_matrix* A, const caffe_vector* x, caffe_vector* y);
int caffe_matrix_multiply_add(caffe_matrix* A, const caffe_vector* x1, const caffe_vector* x2, caffe_vector* y);
int caffe_matrix_scale(double alpha, caffe_matrix* A);
/* SparseVector (src/Math/SparseVector.h:8-43): allocate(m), setValue (VecSetValues,
 * INSERT_VALUES; negative indices ignored), operator() (VecGetValues), zeroEntries,
 * l1Norm / l2Norm / infNorm, and scale(alpha, v) (Diffusion.cpp:127) */
int caffe_vector_create(int device, int m, caffe_vector** out);
int caffe_vector_destroy(caffe_vector* v);
const char* caffe_vector_last_error(caffe_vector* v);
int caffe_vector_set_values(caffe_vector* v, int n, const int* idx, const double* vals);
int caffe_vector_get_values(caffe_vector* v, int n, const int* idx, double* out);
int caffe_vector_zero_entries(caffe_vector* v);
int caffe_vector_scale(double alpha, caffe_vector* v);
int caffe_vector_norm(caffe_vector* v, int type, double* out);

#ifdef __cplusplus
}
#endif
#endif
