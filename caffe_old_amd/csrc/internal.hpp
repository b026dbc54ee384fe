// internal.hpp -- the device-resident solver context behind the C ABI, and the
// host-side launchers each kernel translation unit exports.
#pragma once
#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "../../include/caffe_hip.h"
#include "common.hpp"

namespace caffe {

// Device scalar slots (double) used by the Krylov kernels.
enum {
    SC_BI = 0,     // BiCGStab state, 16 slots per system (up to 3 systems)
    SC_CG = 48,    // CG state
    SC_MISC = 64,  // reductions read back by the host
    SC_FCGR = 128, // the learned FCG's per-round records (learned.hip)
    N_SCAL = 192
};
enum { FL_MDONE = 0, FL_PDONE = 1, FL_MITER = 2, FL_PITER = 3, N_FLAGS = 8 };

enum { PROBE_NONE = 0, PROBE_FCG_K1 = 1, PROBE_CG_K1_FP32 = 2, PROBE_CG_K1_FP64 = 3, PROBE_CP_A = 4,
       PROBE_PBI_A_FP32 = 5, PROBE_PBI_A_FP64 = 6, PROBE_MG_POST0 = 7, PROBE_FCG_K2PRE = 8 };

// Coupled system (Coupled.cpp:54-228): coefficient planes of P doubles per cell, each group
// of seven in the neighbour order own, E, W, N, S, T, B
enum {
    CP_M = 0,        // momentum a_P / omega and a_nb, shared by u, v, w
    CP_G = 7,        // + 7c: pressure coefficients of the momentum row of component c
    CP_D = 28,       // + 7c: velocity-c coefficients of the continuity row
    CP_L = 49,       // pressure coefficients of the continuity row
    CP_PLANES = 56
};
constexpr int PROBE_MAX = 4096;   // launches recorded per probe window

struct Overlay {          // ghost-cell rows (IbPiso.cpp:148-185), 9 entries each
    int n = 0;
    int* cell = nullptr;      // padded index of the ghost cell
    int* cols = nullptr;      // 9 * n padded column cells (-1 = dropped)
    double* valM = nullptr;   // 9 * n momentum-row values
    double* valP = nullptr;   // 9 * n pressure-row values
    float* valP32 = nullptr;
    int* rowOf = nullptr;     // P: overlay row of a padded cell or -1
};

// Multigrid hierarchy of the p' system (mg.hip): level 0 is the context's own grid and
// storage; each coarser level holds its operator (aP and the three low-face coefficients,
// FP32), right-hand side and two iterates.  f is the aggregation factor per direction
// from the finer level.
constexpr int MG_MAX_LEVELS = 16;
struct MgLevel {
    Dims D{};
    int f[3] = {1, 1, 1};
    float* A = nullptr;
    float* b = nullptr;
    float* z = nullptr;
    float* t = nullptr;
    float* w = nullptr;   // damped Jacobi weights 0.8 / aP
};
struct MgHier {
    int nl = 0;                      // levels, 0 until planned for the grid
    MgLevel L[MG_MAX_LEVELS];
    double* cinv = nullptr;          // dense inverse of the coarsest operator
    float* z0[3] = {nullptr, nullptr, nullptr};   // level 0: the two preconditioned vectors, a ping-pong
    float* w0 = nullptr;             // level 0 Jacobi weights
    float* g1 = nullptr;             // first-sweep values of the ghost cells, one per ghost row
    int g1n = 0;
    int lastIts = 0;                 // iterations of the last solve's first round
    int lastItsGauge = 0;            // the same for the Coupled null-vector solve
    hipGraphExec_t iter2 = nullptr;  // two BiCGStab iterations, captured once per context
    bool graphOff = false;           // capture failed: iterations are launched one by one
};

struct Ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    std::string err;
    bool ready = false;

    caffe_case cs{};
    Dims D{};
    Geo geo{};
    bool uniform = true;
    bool hasIB = false;
    bool singularP = false;   // pure Neumann p' system (null space = constants)
    bool indefiniteP = false; // the Sym4 system is symmetric but indefinite (unsteady Diffusion): BiCGStab, not CG
    int uT[6], pT[6], flowT[6];
    bool massCorr[6];

    // fields, padded layout
    double *u[3], *u0[3], *h[3], *p, *pc, *d, *gp[3], *gpc[3], *gu[9], *mflow;
    double *mF[3], *mF2[3];   // mass flux face arrays (ping-pong for correct())
    int8_t* st = nullptr;     // IndexMap statuses, halo = INACTIVE
    int* rowId = nullptr;     // host-visible row numbering is rebuilt on demand

    // assembled systems
    double* Am = nullptr;     // 7 * P  momentum band: aP, aE, aW, aN, aS, aT, aB
    double* bm = nullptr;     // 3 * P
    double* xm = nullptr;     // 3 * P
    double* Ap = nullptr;     // 4 * P  p' system: aP, aI (west face), aJ, aK
    double* bp = nullptr;     // P
    double* xp = nullptr;     // P
    float* Ap32 = nullptr;    // 4 * P
    Overlay ov;

    // Krylov work
    double* wm = nullptr;     // momentum BiCGStab: 12 vectors * 3 * P
    double* wp = nullptr;     // pressure: 8 * P
    float* wp32 = nullptr;    // FP32 inner: 8 * P
    // FP32 BiCGStab work of the ghost-row p' system (11 * P).  Every work buffer is viewed
    // with one element type only: kernels write interior cells, and the stencils read the
    // zero halo through boundary-face coefficients (nonzero at a fixed-pressure outlet),
    // so a buffer shared between FP32 and FP64 views would carry stale bytes there.
    float* wb32 = nullptr;
    // FP32 momentum solves inside FP64 refinement (allocated at first use): the band in
    // FP32 (7 * P) and the BiCGStab work block plus the correction (33 * P)
    float* Am32 = nullptr;
    int momChunk[4] = {4, 4, 4, 4};   // BiCGStab iterations per momentum FP32 round in the last solve
    float* wm32 = nullptr;
    struct { double* part = nullptr; } red;  // 3 partial-sum buffers of MAX_RED_BLOCKS rows
    double* scal = nullptr;   // N_SCAL device scalars
    int* flags = nullptr;     // N_FLAGS device ints
    double* hscal = nullptr;  // pinned host mirror
    int* hflags = nullptr;

    // metric arrays of general meshes
    double* metric = nullptr;
    double* dnodes = nullptr; // node array of general meshes (IB cell location)
    std::vector<void*> allocs;
    size_t deviceBytes = 0;   // bytes held in allocs

    caffe_solve_opts momOpts{}, pcOpts{};
    caffe_solve_report lastMom{}, lastPc{};
    int pcSolvesThisStep = 0;  // p' solves since the last momentum assembly (the corrector)
    bool ibReady = false;     // ghost rows computed for the current case
    double courant = 0.;      // last Courant number (Simple.cpp:88)
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    hipEvent_t evR = nullptr;                   // marks a queued read-back into hscal
    hipEvent_t evP0 = nullptr, evP1 = nullptr;  // the last p' solve, start and end
    bool pcMsPending = false;                   // lastPc.ms still to be read from evP0/evP1

    // live timing of one kernel inside the solves (event pairs around its launches)
    struct Probe {
        int which = 0;                    // PROBE_*
        std::vector<hipEvent_t> ev;       // 2 per launch
        int used = 0;
        // the event pair for the next launch of `kind`, or false when not recording
        bool slot(int kind, hipEvent_t& start, hipEvent_t& stop) {
            if (which != kind || used >= PROBE_MAX) return false;
            start = ev[2 * used];
            stop = ev[2 * used + 1];
            ++used;
            return true;
        }
    } probe;

    // Coupled solver: coefficient planes (CP_PLANES * P), right-hand side and solution
    // (u, v, w, p blocks of P), BiCGStab work and the null-vector solve (52 * P)
    double* cA = nullptr;
    double* cb = nullptr;
    double* cx = nullptr;
    double* cw = nullptr;

    MgHier mg;

    // learned preconditioner (optional)
    void* net = nullptr;
    int fcgCoarse = 0;        // coarse solve of its two-grid cycle: 0 the network, 1 a multigrid V-cycle
    int fcgPost = 3;          // damped Jacobi sweeps after its coarse correction (1 to 4)
    // the learned FCG's rounds queued without a read-back: count and length of each, in two
    // slots (PISO alternates two correctors whose systems differ), each with the norm of the
    // right-hand side it was made on
    int fcgSolves = 0;
    int fcgPlanN[2] = {0, 0};
    int fcgPlan[2][8] = {};
    double fcgPlanB[2] = {0., 0.};
    // the initial guess of a learned or multigrid p' solve: the best combination of the
    // solutions of the same corrector's solves in the two previous time steps (or outer
    // iterations); per corrector a ring of two, stepGen counts momentum assemblies
    int fcgGuess = 1;
    long stepGen = 0;
    double* fcgPrev[2][2] = {{nullptr, nullptr}, {nullptr, nullptr}};
    long fcgPrevGen[2][2] = {{-1, -1}, {-1, -1}};
};

// ---- launchers (fv_kernels.hip)
void launch_halo(Ctx& c, double* const* f, int nf, const int* types);
void launch_mom_assemble(Ctx& c, double dt);
void launch_mom_copyback(Ctx& c);
void launch_rhie_chow(Ctx& c);
void launch_pc_assemble(Ctx& c);
void launch_pc_copyback(Ctx& c);
void launch_grad_scalar(Ctx& c, const double* phi, double* const* out);
void launch_grad_vector(Ctx& c);
void launch_correct(Ctx& c);
double continuity_quirk(Ctx& c);
double continuity_max(Ctx& c);
// Diffusion (Diffusion.cpp:33-131): assembly into the p' storage, the extrapolated-face
// gradient of phi, and ||A phi - b||_inf
void launch_diff_assemble(Ctx& c, double dt);
void launch_copy_cells(Ctx& c, const double* src, double* dst);   // non-INACTIVE cells
void launch_grad_extrap(Ctx& c, const double* phi, double* const* g1, double* const* out);
double diff_resid_inf(Ctx& c, const double* x);
double courant(Ctx& c, double dt);
// the Courant number's reduction and its copy to the host mirror, queued; courant_queued
// reads it once the stream has been synchronised
void courant_queue(Ctx& c, double dt);
double courant_queued(const Ctx& c);
// Coupled (Coupled.cpp:54-338): assembly of the 4-variable system, the field update after its
// solve (u, p, grad p, h, Rhie-Chow faces), the continuity error
void launch_cp_assemble(Ctx& c, double dt);
void launch_cp_update(Ctx& c, const double* x);
double cp_continuity(Ctx& c);
void launch_fill(Ctx& c, double* a, double v, long n);
double time_copy(Ctx& c, double* a, double* b, int n, int reps);
void launch_set_inactive_fields(Ctx& c);

// ---- mesh (mesh.hip)
void compute_metrics(Ctx& c, const double* dNodes);

// ---- Krylov (krylov.hip)
int momentum_solve(Ctx& c, const caffe_solve_opts& o, caffe_solve_report* rep);
// lazyMs: the report's ms is left to caffe_last_reports (no wait for the solve's end)
int pcorr_solve(Ctx& c, const caffe_solve_opts& o, caffe_solve_report* rep, bool lazyMs = false);
void pcorr_remove_mean(Ctx& c);
// x = g0 x0 + g1 x1 (x1 = null: g1 = 0), the combination whose residual b - shift - A x is
// smallest in the 2-norm (never above |b - shift|)
void pcorr_scaled_guess(Ctx& c, const double* x0, const double* x1, double shift);
// the guess of this p' solve (the solutions of its corrector in the two previous steps), if
// any; and, after a converged solve, its solution kept for the next steps
void pcorr_apply_guess(Ctx& c, double shift);
void pcorr_save_guess(Ctx& c);
void pcorr_resolve_ms(Ctx& c);
// the Coupled system: Jacobi BiCGStab over the joint (u, v, w, p) vector in FP64, to a true
// relative residual rtol; p is returned with zero mean when no face fixes it
int coupled_solve(Ctx& c, const caffe_solve_opts& o, caffe_solve_report* rep);
void pcorr_make_fp32(Ctx& c);
// ---- multigrid (mg.hip): BiCGStab on the p' system with the FP32 V-cycle as right
// preconditioner, FP64 recurrences restarted from the true residual; continues from c.xp.
// Returns the iteration count, or -1 when the system does not qualify (8 cells or fewer,
// or an indefinite system)
int mg_bicgstab(Ctx& c, double shift, double rtol, int maxit, int* refs, double* relOut);
// The classical coarse solve of the learned two-grid cycle (its ablation): prepare builds the
// hierarchy's operators for the current p' system (false unless level 1 is the half grid);
// vcycle runs one V-cycle of levels 1 .. on b1 (level-1 right-hand side, unpadded, i fastest)
// into z1, returning at once when *done is set
bool mg_coarse_prepare(Ctx& c);
void mg_coarse_vcycle(Ctx& c, const float* b1, float* z1, const double* done);
void add_vec(Ctx& c, double* x, const double* d, int n);
double time_cg_k1(Ctx& c, bool fp32, int reps);

// ---- immersed boundary (ib.hip)
int ib_field(Ctx& c);
int ib_coeffs(Ctx& c);

// zero-filled device allocation owned by the context (freed by freeAll)
float* falloc(Ctx& c, size_t n);
double* dalloc64(Ctx& c, size_t n);

// FP64 true residual of the p' system (r = b - shift - A x into wp + 6P): ||r||, ||b - shift||
void pcorr_true_resid(Ctx& c, double shift, double* rr, double* bb);
// the same residual left on the device: (r,r), (b,b), sum r at c.scal + SC_MISC + 8
void pcorr_true_resid_dev(Ctx& c, double shift);

// ---- learned preconditioner (learned.hip): flexible CG with the learned two-grid
// cycle, FP32 iterations inside FP64 refinement; returns the iteration count, or -1
// when the system does not qualify (ghost rows, dimensions not divisible by 16)
int learned_fcg(Ctx& c, double shift, double bNorm, double rtol, int maxit, double delta, int* refs, double* relOut,
                bool* meanQueued);

#define HIPCHECK(x)                                                        \
    do {                                                                   \
        hipError_t e_ = (x);                                               \
        if (e_ != hipSuccess) throw std::string(#x) + ": " + hipGetErrorString(e_); \
    } while (0)

inline int blocksFor(long n, int bs, int cap = 1 << 30) {
    long b = (n + bs - 1) / bs;
    if (b > cap) b = cap;
    if (b < 1) b = 1;
    return (int)b;
}

}  // namespace caffe

struct caffe_ctx {
    caffe::Ctx c;
};
