// krylov.hpp -- launch shape, cell loop and the BiCGStab state shared by the Krylov
// solvers (krylov.hip) and the multigrid-preconditioned BiCGStab (mg.hip).
#pragma once
#include "internal.hpp"

namespace caffe {

constexpr int KBS = 1024;
constexpr int KGRID = 256;   // one block of 16 waves per CU: the prologue that adds the partial rows
                             // is paid once per CU, and the rows are few

// interior cells in one contiguous chunk per block (XCD-aware chunk order): q = interior
// id, p = padded index
__device__ inline int cellIdx(const Dims& D, int q) { int i, j, k; return D.cell(q, i, j, k); }
#define FOR_CELLS(D, q, p)                                                                 \
    for (int q##c_ = ((D).N + (int)gridDim.x - 1) / (int)gridDim.x, q = xcdChunk() * q##c_ + (int)threadIdx.x, \
             q##e_ = min((D).N, q - (int)threadIdx.x + q##c_), p = 0;                      \
         q < q##e_ && ((p = cellIdx((D), q)), true); q += blockDim.x)

#define BLOCK0 (blockIdx.x == 0 && threadIdx.x == 0)

struct LoopCtl { double tol2; int maxit; double nLive; };

// BiCGStab state slots (doubles), B_STRIDE per system
enum { B_RHOCUR = 0, B_RHOPREV, B_ALPHA, B_OMEGA, B_BB, B_RR, B_IT, B_DONE, B_BRK,
       B_TOL2, B_MAXIT,   // loop limits kept on the device (the multigrid BiCGStab, whose
                          // iterations replay as a captured graph)
       B_STRIDE = 16 };

// The state of the NV systems, loaded by every block in one batch at kernel start, ahead
// of the partial sums of the previous kernel
template <int NV>
struct BiState {
    double bb[NV], done[NV], it[NV], brk[NV], om[NV], rhoPrev[NV], al[NV], rho[NV];
    __device__ __forceinline__ void load(const double* st) {
        for (int s = 0; s < NV; ++s) {
            const double* c = st + B_STRIDE * s;
            bb[s] = c[B_BB]; done[s] = c[B_DONE]; it[s] = c[B_IT]; brk[s] = c[B_BRK];
            om[s] = c[B_OMEGA]; rhoPrev[s] = c[B_RHOPREV]; al[s] = c[B_ALPHA]; rho[s] = c[B_RHOCUR];
        }
    }
    __device__ __forceinline__ bool any() const {
        bool a = false;
        for (int s = 0; s < NV; ++s) a |= done[s] == 0.;
        return a;
    }
};

// KA head, every block: the convergence test on the partials (rh, r), (r, r) of KC,
// beta and the omega of the previous step; block 0 records them
template <int NV>
__device__ __forceinline__ void biaHead(const double* partIn, int nb, double* st, const LoopCtl& lc,
                                        const BiState<NV>& S, double* beta, double* om, bool* live) {
    double t[2 * NV];
    readPartials<KBS, 2 * NV>(partIn, nb, t);
    for (int s = 0; s < NV; ++s) {
        const double rho = t[2 * s], rr = t[2 * s + 1];
        const double bb = S.bb[s] < 0. ? rr : S.bb[s];
        double done = S.done[s];
        if (done == 0.) {
            if (rr <= lc.tol2 * bb || S.it[s] >= lc.maxit || rr == 0.) done = 1.;
            else if (S.brk[s] != 0. || rho == 0.) done = 2.;
        }
        live[s] = done == 0.;
        om[s] = S.om[s];
        beta[s] = S.it[s] == 0. ? 0. : (rho / S.rhoPrev[s]) * (S.al[s] / S.om[s]);
        if (BLOCK0) {
            double* w = st + B_STRIDE * s;
            w[B_BB] = bb; w[B_RR] = rr; w[B_RHOCUR] = rho; w[B_DONE] = done;
        }
    }
}
// KB head: alpha = rho / (rh, v)
template <int NV>
__device__ __forceinline__ void bibHead(const double* partIn, int nb, double* st, const BiState<NV>& S, double* al,
                                        bool* live) {
    double t[NV];
    readPartials<KBS, NV>(partIn, nb, t);
    for (int s = 0; s < NV; ++s) {
        live[s] = S.done[s] == 0. && t[s] != 0.;
        al[s] = live[s] ? S.rho[s] / t[s] : 0.;
        if (BLOCK0) {
            double* w = st + B_STRIDE * s;
            w[B_ALPHA] = al[s];
            if (S.done[s] == 0. && t[s] == 0.) w[B_BRK] = 1.;
        }
    }
}

// The operator of the Coupled system (Coupled.cpp:54-234): planes CP_M (momentum), CP_G
// (pressure gradient, per component), CP_D (continuity, per velocity component) and CP_L
// (continuity, pressure) over the offsets own, E, W, N, S, T, B.
struct CpOp {
    const double* A; int P, sJ, sK;
    // y = the four rows of cell p applied to f(v, c): variable v at padded cell c
    template <class F>
    __device__ __forceinline__ void apply(F f, int p, double* y) const {
        const int o[7] = {0, 1, -1, sJ, -sJ, sK, -sK};
        double xp[7];
#pragma unroll
        for (int m = 0; m < 7; ++m) xp[m] = f(3, p + o[m]);
        double cont = 0.;
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            double mom = 0.;
#pragma unroll
            for (int m = 0; m < 7; ++m) {
                const double xv = f(c, p + o[m]);
                mom += A[(CP_M + m) * P + p] * xv + A[(CP_G + 7 * c + m) * P + p] * xp[m];
                cont += A[(CP_D + 7 * c + m) * P + p] * xv;
            }
            y[c] = mom;
        }
#pragma unroll
        for (int m = 0; m < 7; ++m) cont += A[(CP_L + m) * P + p] * xp[m];
        y[3] = cont;
    }
};

// The Coupled system with the block-triangular multigrid preconditioner (mg.hip).
// cp_mg_setup: the multigrid hierarchy of the assembled system's continuity-pressure block
// (false when the grid is too small for a coarse level).  cp_mg_run: BiCGStab on
// A x = b - shift (shift on the continuity rows) from x = 0 to the recurrence tolerance
// rtol; returns the iterations, *done the state's done flag.  The iterations are queued in
// chunks between read-backs of the state: firstChunk, then 2, 4, ... 16 (2 * firstChunk, ...
// when firstChunk is 4 or less).
bool cp_mg_setup(Ctx& c);
int cp_mg_run(Ctx& c, const CpOp& A, const double* b, double shift, double* x, double rtol, int maxit, int firstChunk,
              int* doneOut);

}  // namespace caffe
