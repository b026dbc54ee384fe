/*
 * caffe_oracle.h -- CPU FP64 restatement of CAFFE's SIMPLE/PISO/IbPiso hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the correctness oracle and the
 * CPU baseline.  Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load it, and only as the checker: the product path
 * (caffe_old_amd/, libcaffe_hip.so) never links, loads or calls it.
 *
 * Every function restates the reference algorithm with the same IEEE
 * operation order (built with -ffp-contract=off) and cites the reference
 * file:line it follows.  Device-layout export: every cell array is exported
 * in the padded layout (nI+2)*(nJ+2)*(nK+2), i fastest, where the one-cell
 * halo holds the six boundary patches; face arrays use the same shape, the
 * I-face array entry at padded cell (i,j,k) being the WEST face of that cell
 * (so the east face of the last cell lands in the halo slot).
 */
#ifndef CAFFE_ORACLE_H
#define CAFFE_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* FlowBoundaryConditions::Type (FlowBoundaryConditions.h:10) */
enum { OC_INLET = 0, OC_OUTLET = 1, OC_WALL = 2, OC_EMPTY = 3, OC_ZERO_GRADIENT = 4,
       OC_FIXED = 5 /* PrimitiveBoundaryCondition FIXED, the Diffusion solver's boundaries */ };
/* Solver family (Simple.h, Piso.h, IbPiso.h) */
enum { OC_SIMPLE = 0, OC_PISO = 1, OC_IBPISO = 2, OC_DIFFUSION = 3 /* Diffusion.h */,
       OC_COUPLED = 4 /* Coupled.h: the 4-variable (u, v, w, p) block system */ };
/* Linear solver used in place of PETSc KSPBCGS + PCSOR (SparseMatrix.cpp:86-106) */
enum { OC_KSP_PETSC_LIKE = 0, OC_KSP_TIGHT = 1 };

typedef struct {
    int nI, nJ, nK;              /* cells */
    int algo;                    /* OC_SIMPLE / OC_PISO / OC_IBPISO */
    int steady;                  /* Solver.timeAccurate OFF -> 1 (Solver.cpp:8-11) */
    double rho, mu;              /* Solver.rho, Solver.mu (Simple.cpp:104-108) */
    double alphaU, alphaP;       /* relaxationFactorMomentum / PCorr (Simple.cpp:66-67) */
    int nInner, nPCorr;          /* numberOfInnerIterations / numberOfPressureCorrections */
    int bcType[6];               /* east, west, north, south, top, bottom */
    double bcVec[18];            /* refVector per face */
    double bcVal[6];             /* refValue per face */
    double uInit[3];             /* uniform initial u (InitialConditions.tpp:32-36) */
    double pInit;                /* uniform initial p */
    double sphereR, sphereC[3];  /* ImmersedBoundaries.Sphere (IbPiso.cpp:37-38) */
    int kspKind;                 /* OC_KSP_PETSC_LIKE / OC_KSP_TIGHT */
    double rtol;                 /* PETSc-like: rtol on preconditioned norm; tight: true relres */
    int maxit;
    int nBlocks;                 /* row blocks of the block-local SSOR ("MPI ranks") */
    int nThreads;                /* OpenMP threads for assembly / Krylov (1 = serial, deterministic) */
    int nonOrth;                 /* 1: corrected mode, the non-orthogonal term mu_f (grad u)_f . c_f of
                                    Simple.cpp:153-158 with the faces of grad u filled (interior:
                                    volume-weighted interpolation; domain faces: the cell's own
                                    gradient); 0: reference mode, the faces stay zero (the term
                                    vanishes, as in the reference) */
} oc_case;

void*  oc_create(const oc_case* c, const double* nodes /* (nI+1)(nJ+1)(nK+1) x 3, i fastest */);
void   oc_destroy(void* h);

/* Solver::solve(dt) for the configured algo (Simple.cpp:78-91, Piso.cpp:12-29, IbPiso.cpp:43-61) */
double oc_solve(void* h, double dt);

/* Sub-steps, in the reference's own order */
void   oc_save_u0(void* h);                        /* uField0_ = uField_ */
void   oc_compute_momentum(void* h, double dt);    /* Simple.cpp:110-221 */
void   oc_assemble_momentum(void* h, double dt);   /* Simple.cpp:117-179 only */
void   oc_compute_pcorr(void* h);                  /* Simple.cpp:223-299 */
void   oc_assemble_pcorr(void* h);                 /* Simple.cpp:231-269 only */
void   oc_correct(void* h);                        /* Simple.cpp:301-356 */
void   oc_rhie_chow(void* h);                      /* Simple.cpp:410-449 */
void   oc_compute_ib_field(void* h);               /* IbPiso.cpp:65-117 */
void   oc_compute_ib_coeffs(void* h);              /* IbPiso.cpp:119-190 */
double oc_continuity(void* h);                     /* Simple.cpp:358-380 (order-dependent scan) */
double oc_continuity_max(void* h);                 /* max |m_c|/V_c over active cells */
double oc_courant(void* h, double dt);             /* Simple.cpp:382-408 */
void   oc_grad_u(void* h);                         /* FvVectorScheme.cpp:56-74 on uField_ */

/* State exchange in the padded device layout; returns number of doubles, -1 on bad name */
long   oc_get(void* h, const char* name, double* out);
long   oc_set(void* h, const char* name, const double* in);
long   oc_get_metric(void* h, const char* name, double* out);

/* Assembled systems.  which: 0 = momentum (A_[0]), 1 = pressure correction (A_[1]),
 * 2 = the Coupled system (4 * nrows rows: u, v, w, p blocks; its right-hand side is vec 3) */
int    oc_nrows(void* h);
long   oc_matrix_nnz(void* h, int which);
void   oc_matrix_csr(void* h, int which, int* rowptr, int* cols, double* vals);
void   oc_rhs(void* h, int vec, double* out);       /* b_[vec], row order */
void   oc_imap(void* h, int* rowid, int* status);   /* unpadded, i fastest */
void   oc_stats(void* h, int* itersMomentum, int* itersPCorr, double* lastRelres);

/* Geometry/IB helpers restated from Geometry.cpp / Sphere.cpp / Interpolation.cpp */
double oc_hex_volume(const double* pts24);
void   oc_hex_centroid(const double* pts24, double* out3);
double oc_quad_area(const double* pts12);
void   oc_quad_centroid(const double* pts12, double* out3);
int    oc_inside_hex(const double* pt3, const double* pts24);
void   oc_trilinear(const double* pts24, const double* x3, double* beta8);
void   oc_transfinite_nodes(const double* verts24, int nNodesI, int nNodesJ, int nNodesK, double* out);

#ifdef __cplusplus
}
#endif
#endif
