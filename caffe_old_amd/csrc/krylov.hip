// krylov.hip -- preconditioned Krylov solvers for the momentum and pressure-
// correction systems (the replacement of SparseMatrix::solve, SparseMatrix.cpp:86-106).
//
// Every solve starts from x = 0, as KSPSolve does with PETSc's default initial guess.
// Convergence is judged on the true FP64 residual ||b - A x||_2 <= rtol ||b||_2.
//
// Kernel design.  All Krylov scalars stay in device memory.  Each kernel
//   1. adds up the per-block partial sums the previous kernel wrote (every block does
//      this redundantly, in block order, so all blocks hold identical scalars and the
//      result does not depend on scheduling),
//   2. derives its step length (alpha, beta, omega) and the convergence flag,
//   3. applies the 7-point operator to an operand formed on the fly at the stencil
//      points from the previous iterates (so there is no separate vector-update pass),
//   4. writes its own partial sums.
// No kernel uses a device-scope fence or atomic: on CDNA the L2 of each XCD is not
// coherent with the others, so such fences write back L2; kernel boundaries give the
// ordering instead.  Block 0 records the scalars the next kernel or the host needs.
//
// Systems:
//   momentum  7-point band (aP aE aW aN aS aT aB), three right-hand sides sharing the
//             coefficients, 9-entry ghost rows; BiCGStab with Jacobi, FP64.
//   pressure  symmetric storage (aP + one coefficient per face); CG with Jacobi.  FP64,
//             or FP32 iterations with FP64 reliable updates (the true residual is
//             recomputed in FP64 and the FP32 recurrence restarted from it, keeping the
//             search direction).  With ghost rows (nonsymmetric) BiCGStab is used.
#include <hip/hip_ext.h>

#include "internal.hpp"
#include "krylov.hpp"
#include "sym4.hpp"

#include <cmath>
#include <type_traits>

namespace caffe {

// ------------------------------------------------------------------ operators
template <class T>
struct BandT {             // momentum matrix; T = float only without ghost rows
    const T* A; int P, sJ, sK; OvView ov; const int8_t* st;
    template <class F>
    __device__ double apply(F x, int p) const {
        if (ov.rowOf) {
            int r = ov.rowOf[p];
            if (r >= 0) {
                const double* v = (const double*)ov.vals + 9 * r;
                double s = v[0] * x(p);
                for (int m = 1; m < 9; ++m) { int c = ov.cols[9 * r + m]; if (c >= 0) s += v[m] * x(c); }
                return s;
            }
        }
        return (double)A[p] * x(p) + (double)A[P + p] * x(p + 1) + (double)A[2 * P + p] * x(p - 1)
             + (double)A[3 * P + p] * x(p + sJ) + (double)A[4 * P + p] * x(p - sJ) + (double)A[5 * P + p] * x(p + sK)
             + (double)A[6 * P + p] * x(p - sK);
    }
    __device__ double diag(int p) const { return (double)A[p]; }
};
using Band7 = BandT<double>;

// ================================================================== CG
// state slots (doubles)
enum { C_RZCUR = 0, C_RZPREV, C_BB, C_RR, C_IT, C_DONE, C_RATIO, C_PQ, C_PBUF };

// r = s*(b - shift) on live cells, z = r / aP, p = 0 (zeroP), x = 0 (zeroX);
// partials (r,z), (r,r)
template <class T, class TB>
__global__ void k_cg_init(Dims D, Sym4<T> A, const TB* b, double shift, double s, T* r, T* z, T* p0, T* p1, T* x,
                          int zeroP, int zeroX, double* part, double* st, double ratio) {
    double v[4] = {0., 0., 0., 0.};
    FOR_CELLS(D, q, p) {
        const bool live = !A.st || A.st[p] != ST_INACTIVE;
        const T rr = (T)(live ? ((double)b[p] - shift) * s : 0.);
        const double zz = (double)rr / A.diag(p);
        r[p] = rr; z[p] = (T)zz;
        if (zeroP) { p0[p] = 0; p1[p] = 0; }
        if (zeroX) x[p] = 0;
        v[0] += (double)rr * zz;
        v[1] += (double)rr * (double)rr;
        v[2] += (double)(T)zz;
        v[3] += (double)rr;
    }
    writePartials<KBS, 4>(v, part);
    if (BLOCK0) {
        if (zeroP) st[C_IT] = 0.;
        st[C_DONE] = 0.; st[C_BB] = -1.; st[C_RATIO] = ratio;
    }
}

// K1: [rz, rr from the previous partials] convergence; beta; p_new = z + beta p_old formed
// at the stencil points; q = A p_new; partial (p_new, q)
// For a singular (pure Neumann) system lc.nLive > 0: the preconditioned residual is
// projected onto zero mean, z - mean(z), so that neither the search directions nor the
// iterate drift along the constant null vector (rounding puts a small component there).
// The two direction buffers alternate with the parity of the completed-iteration count
// kept on the device, so a chunk of queued iterations that stops early (convergence)
// leaves the latest direction where the next K1 looks for it.
template <class T>
__global__ void k_cg_k1(Dims D, Sym4<T> A, const T* z, T* pb0, T* pb1, T* qv, const double* partIn, int nb,
                        double* partOut, double* st, LoopCtl lc) {
    // the scalars written by earlier kernels are loaded together, ahead of the partial sums
    const double sDone = st[C_DONE], sIt = st[C_IT], sBb = st[C_BB], sRzPrev = st[C_RZPREV], sRatio = st[C_RATIO];
    if (sDone != 0.) return;
    const int par = ((int)sIt) & 1;
    const T* pold = par ? pb1 : pb0;
    T* pnew = par ? pb0 : pb1;
    double t[4];
    readPartials<KBS, 4>(partIn, nb, t);
    const double zbar = lc.nLive > 0. ? t[2] / lc.nLive : 0.;
    const double rz = t[0] - zbar * t[3];
    // residual norm without its (unreachable) mean component
    const double rr = lc.nLive > 0. ? fmax(t[1] - t[3] * t[3] / lc.nLive, 0.) : t[1];
    const double bb = sBb < 0. ? rr : sBb;
    const bool first = sIt == 0.;
    const bool done = rr <= lc.tol2 * bb || sIt >= lc.maxit || rr == 0.;
    if (BLOCK0) {
        st[C_BB] = bb; st[C_RR] = rr; st[C_RZCUR] = rz; st[C_PBUF] = par ? 0. : 1.;
        if (done) st[C_DONE] = 1.;
    }
    if (done) return;
    const double beta = first ? 0. : (rz / sRzPrev) * sRatio;
    double v[1] = {0.};
    // (singular systems have no ghost rows: halo cells meet zero coefficients, so the
    // shifted value there is never used)
    auto pf = [&](int c) { return ((double)z[c] - zbar) + beta * (double)pold[c]; };
    FOR_CELLS(D, q, p) {
        const double pn = pf(p);
        const double y = A.apply(pf, p);
        pnew[p] = (T)pn; qv[p] = (T)y;
        v[0] += pn * y;
    }
    writePartials<KBS, 1>(v, partOut);
}

// K2: [alpha = rz / (p, q)] x += alpha p; r -= alpha q; z = r / aP; partials (r,z), (r,r)
template <class T>
__global__ void k_cg_k2(Dims D, Sym4<T> A, T* x, T* r, T* z, const T* pb0, const T* pb1, const T* qv, const double* partIn,
                        int nb, double* partOut, double* st) {
    const double sDone = st[C_DONE], sPbuf = st[C_PBUF], sRz = st[C_RZCUR], sIt = st[C_IT];
    if (sDone != 0.) return;
    const T* pv = sPbuf != 0. ? pb1 : pb0;  // the direction K1 of this iteration wrote
    double t[1];
    readPartials<KBS, 1>(partIn, nb, t);
    const double alpha = t[0] != 0. ? sRz / t[0] : 0.;
    if (BLOCK0) {
        st[C_RZPREV] = sRz; st[C_IT] = sIt + 1.; st[C_RATIO] = 1.; st[C_PQ] = t[0];
    }
    double v[4] = {0., 0., 0., 0.};
    FOR_CELLS(D, q, p) {
        const double xn = (double)x[p] + alpha * (double)pv[p];
        const T rt = (T)((double)r[p] - alpha * (double)qv[p]);
        const T zn = (T)((double)rt / A.diag(p));
        x[p] = (T)xn; r[p] = rt; z[p] = zn;
        v[0] += (double)rt * (double)zn;
        v[1] += (double)rt * (double)rt;
        v[2] += (double)zn;
        v[3] += (double)rt;
    }
    writePartials<KBS, 4>(v, partOut);
}

// ================================================================== BiCGStab, NV systems
template <int NV, class T, class Op>
__global__ void k_bi_init(Dims D, Op A, const double* b, double shift, T* r, T* rh, T* p0, T* v0, T* x, T* dinv,
                          double* part, double* st) {
    double acc[2 * NV];
    for (int s = 0; s < 2 * NV; ++s) acc[s] = 0.;
    const int P = D.P;
    FOR_CELLS(D, q, p) {
        const bool live = !A.st || A.st[p] != ST_INACTIVE;
        dinv[p] = (T)(1. / A.diag(p));
        for (int s = 0; s < NV; ++s) {
            const int o = s * P + p;
            const T bv = (T)(live ? b[o] - shift : 0.);
            r[o] = bv; rh[o] = bv; p0[o] = 0; v0[o] = 0; x[o] = 0;
            acc[2 * s] += (double)bv * (double)bv;
            acc[2 * s + 1] += (double)bv * (double)bv;
        }
    }
    writePartials<KBS, 2 * NV>(acc, part);
    if (BLOCK0) {
        for (int s = 0; s < NV; ++s) {
            double* c = st + B_STRIDE * s;
            c[B_IT] = 0.; c[B_DONE] = 0.; c[B_BB] = -1.; c[B_RHOPREV] = 1.; c[B_ALPHA] = 1.; c[B_OMEGA] = 1.; c[B_BRK] = 0.;
        }
    }
}

// KA: [rho, rr] convergence; beta; p_new = r + beta (p - omega v) at the stencil points;
// v_new = A D^-1 p_new; partial (rh, v_new)
template <int NV, class T, class Op>
__global__ void k_bi_a(Dims D, Op A, const T* r, const T* rh, const T* pold, T* pnew, const T* vold, T* vnew,
                       const T* dinv, const double* partIn, int nb, double* partOut, double* st, LoopCtl lc) {
    BiState<NV> S;
    S.load(st);
    if (!S.any()) return;
    double beta[NV], om[NV];
    bool live[NV];
    biaHead<NV>(partIn, nb, st, lc, S, beta, om, live);
    double acc[NV];
    for (int s = 0; s < NV; ++s) acc[s] = 0.;
    const int P = D.P;
    FOR_CELLS(D, q, p) {
        for (int s = 0; s < NV; ++s) {
            if (!live[s]) continue;
            const int o = s * P;
            auto pf = [&](int c) { return (double)r[o + c] + beta[s] * ((double)pold[o + c] - om[s] * (double)vold[o + c]); };
            auto ph = [&](int c) { return pf(c) * (double)dinv[c]; };
            const double pn = pf(p);
            const double vn = A.apply(ph, p);
            pnew[o + p] = (T)pn; vnew[o + p] = (T)vn;
            acc[s] += (double)rh[o + p] * vn;
        }
    }
    writePartials<KBS, NV>(acc, partOut);
}

// KB: [alpha = rho / (rh, v)] s = r - alpha v at the stencil points; t = A D^-1 s;
// partials (t,s), (t,t)
template <int NV, class T, class Op>
__global__ void k_bi_b(Dims D, Op A, const T* r, const T* v, T* sv, T* tv, const T* dinv, const double* partIn, int nb,
                       double* partOut, double* st) {
    BiState<NV> S;
    S.load(st);
    if (!S.any()) return;
    double al[NV];
    bool live[NV];
    bibHead<NV>(partIn, nb, st, S, al, live);
    double acc[2 * NV];
    for (int s = 0; s < 2 * NV; ++s) acc[s] = 0.;
    const int P = D.P;
    FOR_CELLS(D, q, p) {
        for (int s = 0; s < NV; ++s) {
            if (!live[s]) continue;
            const int o = s * P;
            auto sf = [&](int c) { return (double)r[o + c] - al[s] * (double)v[o + c]; };
            auto sh = [&](int c) { return sf(c) * (double)dinv[c]; };
            const double ss = sf(p);
            const double tt = A.apply(sh, p);
            sv[o + p] = (T)ss; tv[o + p] = (T)tt;
            acc[2 * s] += tt * ss;
            acc[2 * s + 1] += tt * tt;
        }
    }
    writePartials<KBS, 2 * NV>(acc, partOut);
}

// KA and KB for the three momentum components on the band: the seven coefficients and
// D^-1 at the seven stencil points are loaded once per cell for the three components (the
// generic kernels reload them per component, as their stores may alias them).  Same
// arithmetic as Band7::apply.  G: the system has ghost rows; a ghost cell (about 0.1% of
// the cells) evaluates its 9-entry row (FP64 values, as BandT::apply reads them) on the
// operand formed at the row's cells.
template <class T, bool G>
__device__ __forceinline__ bool bandGhost(const BandT<T>& A, int p, int& r) {
    r = G ? A.ov.rowOf[p] : -1;
    return G && r >= 0;
}
template <class T, class F>
__device__ __forceinline__ double ghostApply(const BandT<T>& A, int r, int p, F x) {
    const double* v = (const double*)A.ov.vals + 9 * r;
    double s = v[0] * x(p);
    for (int m = 1; m < 9; ++m) {
        const int c = A.ov.cols[9 * r + m];
        if (c >= 0) s += v[m] * x(c);
    }
    return s;
}
template <class T, bool G>
__global__ void k_bi_a_band(Dims D, BandT<T> A, const T* __restrict__ r, const T* __restrict__ rh,
                            const T* __restrict__ pold, T* __restrict__ pnew, const T* __restrict__ vold,
                            T* __restrict__ vnew, const T* __restrict__ dinv, const double* partIn, int nb,
                            double* partOut, double* st, LoopCtl lc) {
    BiState<3> S;
    S.load(st);
    if (!S.any()) return;
    double beta[3], om[3];
    bool live[3];
    biaHead<3>(partIn, nb, st, lc, S, beta, om, live);
    double acc[3] = {0., 0., 0.};
    const int P = D.P;
    const int o[7] = {0, 1, -1, D.sJ, -D.sJ, D.sK, -D.sK};
    FOR_CELLS(D, q, p) {
        double a[7], di[7];
#pragma unroll
        for (int m = 0; m < 7; ++m) { a[m] = (double)A.A[m * P + p]; di[m] = (double)dinv[p + o[m]]; }
        int gr;
        const bool ghost = bandGhost<T, G>(A, p, gr);
#pragma unroll
        for (int s = 0; s < 3; ++s) {
            if (!live[s]) continue;
            const int b = s * P + p;
            double ph[7], pn = 0.;
#pragma unroll
            for (int m = 0; m < 7; ++m) {
                const double pf = (double)r[b + o[m]] + beta[s] * ((double)pold[b + o[m]] - om[s] * (double)vold[b + o[m]]);
                if (m == 0) pn = pf;
                ph[m] = pf * di[m];
            }
            double vn = a[0] * ph[0] + a[1] * ph[1] + a[2] * ph[2] + a[3] * ph[3] + a[4] * ph[4] + a[5] * ph[5]
                      + a[6] * ph[6];
            if (ghost) {
                const int so = s * P;
                vn = ghostApply(A, gr, p, [&](int c) {
                    return ((double)r[so + c] + beta[s] * ((double)pold[so + c] - om[s] * (double)vold[so + c])) * (double)dinv[c];
                });
            }
            pnew[b] = (T)pn; vnew[b] = (T)vn;
            acc[s] += (double)rh[b] * vn;
        }
    }
    writePartials<KBS, 3>(acc, partOut);
}
template <class T, bool G>
__global__ void k_bi_b_band(Dims D, BandT<T> A, const T* __restrict__ r, const T* __restrict__ v, T* __restrict__ sv,
                            T* __restrict__ tv, const T* __restrict__ dinv, const double* partIn, int nb,
                            double* partOut, double* st) {
    BiState<3> S;
    S.load(st);
    if (!S.any()) return;
    double al[3];
    bool live[3];
    bibHead<3>(partIn, nb, st, S, al, live);
    double acc[6] = {0., 0., 0., 0., 0., 0.};
    const int P = D.P;
    const int o[7] = {0, 1, -1, D.sJ, -D.sJ, D.sK, -D.sK};
    FOR_CELLS(D, q, p) {
        double a[7], di[7];
#pragma unroll
        for (int m = 0; m < 7; ++m) { a[m] = (double)A.A[m * P + p]; di[m] = (double)dinv[p + o[m]]; }
        int gr;
        const bool ghost = bandGhost<T, G>(A, p, gr);
#pragma unroll
        for (int s = 0; s < 3; ++s) {
            if (!live[s]) continue;
            const int b = s * P + p;
            double sh[7], ss = 0.;
#pragma unroll
            for (int m = 0; m < 7; ++m) {
                const double sf = (double)r[b + o[m]] - al[s] * (double)v[b + o[m]];
                if (m == 0) ss = sf;
                sh[m] = sf * di[m];
            }
            double tt = a[0] * sh[0] + a[1] * sh[1] + a[2] * sh[2] + a[3] * sh[3] + a[4] * sh[4] + a[5] * sh[5]
                      + a[6] * sh[6];
            if (ghost) {
                const int so = s * P;
                tt = ghostApply(A, gr, p, [&](int c) {
                    return ((double)r[so + c] - al[s] * (double)v[so + c]) * (double)dinv[c];
                });
            }
            sv[b] = (T)ss; tv[b] = (T)tt;
            acc[2 * s] += tt * ss;
            acc[2 * s + 1] += tt * tt;
        }
    }
    writePartials<KBS, 6>(acc, partOut);
}

// KA and KB on the band without ghost rows, on column tiles (ColMap): each thread walks
// its column up k and keeps the operand it formed at k - 1 and k (p = r + beta (p_old -
// omega v_old) in KA, s = r - alpha v in KB) and D^-1 there in registers, so a cell loads
// the operand's arrays at its four side neighbours and above only (KA: 60 loads per cell
// for three components instead of 80).  At 128^3 that takes KA from 70-73 to 62-66 us and KB
// from 52 to 45 us with three live components, although the counted reads from below L2
// grow (KA 1.43x -> 1.59x its algorithmic reads, KB 1.21x -> 1.41x: the tile-edge
// neighbours): the kernels are bound by their loads' latency, not by those bytes.  The
// arithmetic of each cell is k_bi_a_band's / k_bi_b_band's.
static_assert(CTX * CTY == KBS, "one thread per column of a tile");
template <class T>
__global__ __launch_bounds__(KBS) void k_bi_a_col(Dims D, BandT<T> A, ColMap cm, const T* __restrict__ r,
                                                  const T* __restrict__ rh, const T* __restrict__ pold,
                                                  T* __restrict__ pnew, const T* __restrict__ vold,
                                                  T* __restrict__ vnew, const T* __restrict__ dinv,
                                                  const double* partIn, int nb, double* partOut, double* st, LoopCtl lc) {
    BiState<3> S;
    S.load(st);
    if (!S.any()) return;
    double beta[3], om[3];
    bool live[3];
    biaHead<3>(partIn, nb, st, lc, S, beta, om, live);
    double acc[3] = {0., 0., 0.};
    const int P = D.P, sJ = D.sJ, sK = D.sK;
    auto opnd = [&](int s, int q) {
        const int b = s * P + q;
        return (double)r[b] + beta[s] * ((double)pold[b] - om[s] * (double)vold[b]);
    };
    for (int v = xcdChunk(), vn = cm.ntx * cm.nty * cm.nseg; v < vn; v += (int)gridDim.x) {
        int p, k, ke;
        if (!colItem(D, cm, v, p, k, ke)) continue;
        double pd[3] = {0., 0., 0.}, pc[3] = {0., 0., 0.};   // the operand at k - 1 and k
#pragma unroll
        for (int s = 0; s < 3; ++s)
            if (live[s]) { pd[s] = opnd(s, p - sK); pc[s] = opnd(s, p); }
        double dd = (double)dinv[p - sK], dc = (double)dinv[p];
        for (; k < ke; ++k, p += sK) {
            T a[7];   // converted where used, which keeps the registers of the FP32 values
#pragma unroll
            for (int m = 0; m < 7; ++m) a[m] = A.A[m * P + p];
            const T dE = dinv[p + 1], dW = dinv[p - 1], dN = dinv[p + sJ], dS = dinv[p - sJ], dU = dinv[p + sK];
#pragma unroll 1
            for (int s = 0; s < 3; ++s) {   // one component at a time: the three at once need more than 128 registers
                if (!live[s]) continue;
                const int b = s * P + p;
                const double pu = opnd(s, p + sK);
                const double ph[7] = {pc[s] * dc, opnd(s, p + 1) * (double)dE, opnd(s, p - 1) * (double)dW,
                                      opnd(s, p + sJ) * (double)dN, opnd(s, p - sJ) * (double)dS, pu * (double)dU, pd[s] * dd};
                const double vv = (double)a[0] * ph[0] + (double)a[1] * ph[1] + (double)a[2] * ph[2] + (double)a[3] * ph[3]
                                + (double)a[4] * ph[4] + (double)a[5] * ph[5] + (double)a[6] * ph[6];
                pnew[b] = (T)pc[s]; vnew[b] = (T)vv;
                acc[s] += (double)rh[b] * vv;
                pd[s] = pc[s]; pc[s] = pu;
            }
            dd = dc; dc = (double)dU;
        }
    }
    writePartials<KBS, 3>(acc, partOut);
}
template <class T>
__global__ __launch_bounds__(KBS) void k_bi_b_col(Dims D, BandT<T> A, ColMap cm, const T* __restrict__ r,
                                                  const T* __restrict__ v, T* __restrict__ sv, T* __restrict__ tv,
                                                  const T* __restrict__ dinv, const double* partIn, int nb, double* partOut,
                                                  double* st) {
    BiState<3> S;
    S.load(st);
    if (!S.any()) return;
    double al[3];
    bool live[3];
    bibHead<3>(partIn, nb, st, S, al, live);
    double acc[6] = {0., 0., 0., 0., 0., 0.};
    const int P = D.P, sJ = D.sJ, sK = D.sK;
    auto opnd = [&](int s, int q) {
        const int b = s * P + q;
        return (double)r[b] - al[s] * (double)v[b];
    };
    for (int w = xcdChunk(), wn = cm.ntx * cm.nty * cm.nseg; w < wn; w += (int)gridDim.x) {
        int p, k, ke;
        if (!colItem(D, cm, w, p, k, ke)) continue;
        double sd[3] = {0., 0., 0.}, sc[3] = {0., 0., 0.};
#pragma unroll
        for (int s = 0; s < 3; ++s)
            if (live[s]) { sd[s] = opnd(s, p - sK); sc[s] = opnd(s, p); }
        double dd = (double)dinv[p - sK], dc = (double)dinv[p];
        for (; k < ke; ++k, p += sK) {
            T a[7];   // converted where used, which keeps the registers of the FP32 values
#pragma unroll
            for (int m = 0; m < 7; ++m) a[m] = A.A[m * P + p];
            const T dE = dinv[p + 1], dW = dinv[p - 1], dN = dinv[p + sJ], dS = dinv[p - sJ], dU = dinv[p + sK];
#pragma unroll 1
            for (int s = 0; s < 3; ++s) {
                if (!live[s]) continue;
                const int b = s * P + p;
                const double su = opnd(s, p + sK);
                const double sh[7] = {sc[s] * dc, opnd(s, p + 1) * (double)dE, opnd(s, p - 1) * (double)dW,
                                      opnd(s, p + sJ) * (double)dN, opnd(s, p - sJ) * (double)dS, su * (double)dU, sd[s] * dd};
                const double tt = (double)a[0] * sh[0] + (double)a[1] * sh[1] + (double)a[2] * sh[2] + (double)a[3] * sh[3]
                                + (double)a[4] * sh[4] + (double)a[5] * sh[5] + (double)a[6] * sh[6];
                sv[b] = (T)sc[s]; tv[b] = (T)tt;
                acc[2 * s] += tt * sc[s];
                acc[2 * s + 1] += tt * tt;
                sd[s] = sc[s]; sc[s] = su;
            }
            dd = dc; dc = (double)dU;
        }
    }
    writePartials<KBS, 6>(acc, partOut);
}

// KC: [omega = (t,s)/(t,t)] x += alpha D^-1 p + omega D^-1 s; r = s - omega t;
// partials (rh, r), (r, r)
template <int NV, class T, class Op>
__global__ void k_bi_c(Dims D, Op A, T* x, T* r, const T* rh, const T* pv, const T* sv, const T* tv, const T* dinv,
                       const double* partIn, int nb, double* partOut, double* st) {
    // B_BRK is 1 when KB found (rh, v) = 0 (it then wrote no s, t); this kernel marks a zero
    // omega with 2, which other blocks may already see: they still update, as block 0 does
    double sDone[NV], sBrk[NV], sAl[NV], sRho[NV], sIt[NV];
    bool any = false;
    for (int s = 0; s < NV; ++s) {
        const double* c = st + B_STRIDE * s;
        sDone[s] = c[B_DONE]; sBrk[s] = c[B_BRK]; sAl[s] = c[B_ALPHA]; sRho[s] = c[B_RHOCUR]; sIt[s] = c[B_IT];
        any |= sDone[s] == 0.;
    }
    if (!any) return;
    double t[2 * NV];
    readPartials<KBS, 2 * NV>(partIn, nb, t);
    double al[NV], om[NV];
    bool live[NV];
    for (int s = 0; s < NV; ++s) {
        live[s] = sDone[s] == 0. && sBrk[s] != 1.;
        al[s] = sAl[s];
        om[s] = t[2 * s + 1] > 0. ? t[2 * s] / t[2 * s + 1] : 0.;
        if (BLOCK0 && live[s]) {
            double* w = st + B_STRIDE * s;
            w[B_OMEGA] = om[s]; w[B_RHOPREV] = sRho[s]; w[B_IT] = sIt[s] + 1.;
            if (om[s] == 0.) w[B_BRK] = 2.;
        }
    }
    double acc[2 * NV];
    for (int s = 0; s < 2 * NV; ++s) acc[s] = 0.;
    const int P = D.P;
    FOR_CELLS(D, q, p) {
        // all loads of the cell first (one wait, not one per system); a finished system's
        // x, p, s, t are read and not used (at 128^3: 47.8 -> 42.7 us for the first
        // launch of a momentum solve)
        const double di = (double)dinv[p];
        T xv[NV], pb[NV], sb[NV], tb[NV], rb[NV], hb[NV];
#pragma unroll
        for (int s = 0; s < NV; ++s) {
            const int o = s * P + p;
            xv[s] = x[o]; pb[s] = pv[o]; sb[s] = sv[o]; tb[s] = tv[o]; rb[s] = r[o]; hb[s] = rh[o];
        }
#pragma unroll
        for (int s = 0; s < NV; ++s) {
            const int o = s * P + p;
            double rn;
            if (live[s]) {
                const double xn = (double)xv[s] + al[s] * ((double)pb[s] * di) + om[s] * ((double)sb[s] * di);
                rn = (double)sb[s] - om[s] * (double)tb[s];
                x[o] = (T)xn; r[o] = (T)rn;
            } else {
                rn = (double)rb[s];
            }
            acc[2 * s] += (double)hb[s] * rn;
            acc[2 * s + 1] += rn * rn;
        }
    }
    writePartials<KBS, 2 * NV>(acc, partOut);
}

// ================================================================== residual and utilities
// r = (b - shift) - A x in FP64 over NV systems; partials (r,r), (b,b) per system
template <int NV, class Op>
__global__ void k_resid(Dims D, Op A, const double* b, double shift, const double* x, double* r, double* part) {
    double acc[3 * NV];
    for (int s = 0; s < 3 * NV; ++s) acc[s] = 0.;
    const int P = D.P;
    FOR_CELLS(D, q, p) {
        const bool live = !A.st || A.st[p] != ST_INACTIVE;
        for (int s = 0; s < NV; ++s) {
            const int o = s * P;
            auto xf = [&](int c) { return x[o + c]; };
            const double bv = live ? b[o + p] - shift : 0.;
            const double rv = live ? bv - A.apply(xf, p) : 0.;
            if (r) r[o + p] = rv;
            acc[3 * s] += rv * rv;
            acc[3 * s + 1] += bv * bv;
            acc[3 * s + 2] += rv;
        }
    }
    writePartials<KBS, 3 * NV>(acc, part);
}

// sum and count of a field over non-inactive cells
__global__ void k_sum(Dims D, const int8_t* st, const double* b, double* part) {
    double acc[2] = {0., 0.};
    FOR_CELLS(D, q, p) {
        if (st[p] == ST_INACTIVE) continue;
        acc[0] += b[p];
        acc[1] += 1.;
    }
    writePartials<KBS, 2>(acc, part);
}

template <int K>
__global__ void k_finalize(const double* part, int nb, double* out) {
    double t[K];
    readPartials<KBS, K>(part, nb, t);
    if (threadIdx.x == 0)
        for (int q = 0; q < K; ++q) out[q] = t[q];
}

__global__ void k_add_const(Dims D, const int8_t* st, double* x, double c) {
    FOR_CELLS(D, q, p) { if (st[p] != ST_INACTIVE) x[p] += c; }
}
// k_sum with the sum of squares as a third partial
__global__ void k_sum_sq(Dims D, const int8_t* st, const double* b, double* part) {
    double acc[3] = {0., 0., 0.};
    FOR_CELLS(D, q, p) {
        if (st[p] == ST_INACTIVE) continue;
        acc[0] += b[p];
        acc[1] += 1.;
        acc[2] += b[p] * b[p];
    }
    writePartials<KBS, 3>(acc, part);
}
// x -= the mean that k_sum + k_finalize<2> left in m (sum, count), on the live cells
__global__ void k_sub_mean(Dims D, const int8_t* st, double* x, const double* m) {
    const double mean = m[1] > 0. ? m[0] / m[1] : 0.;
    FOR_CELLS(D, q, p) { if (st[p] != ST_INACTIVE) x[p] -= mean; }
}
__global__ void k_axpy_mixed(Dims D, double* x, const float* e, double s) {
    FOR_CELLS(D, q, p) { x[p] += s * (double)e[p]; }
}
__global__ void k_cast_f32(const double* a, float* o, int n) {
    int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q < n) o[q] = (float)a[q];
}
__global__ void k_zero_d(double* a, int n) {
    int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q < n) a[q] = 0.;
}
__global__ void k_add(double* x, const double* d, int n) {
    int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q < n) x[q] += d[q];
}

// ================================================================== host drivers
static int gridFor(const Dims& D) { return blocksFor(D.N, KBS, KGRID); }

static void readBack(Ctx& c, const double* dev, int slot, int n) {
    HIPCHECK(hipMemcpyAsync(c.hscal + slot, dev, n * sizeof(double), hipMemcpyDeviceToHost, c.stream));
    HIPCHECK(hipStreamSynchronize(c.stream));
}

void add_vec(Ctx& c, double* x, const double* d, int n) {
    hipLaunchKernelGGL(k_add, dim3(blocksFor(n, KBS)), dim3(KBS), 0, c.stream, x, d, n);
}

// mean over non-inactive cells
// the mean of v over the live cells and the norm of v minus that mean, in one read-back
// queued in two halves: meanNormQueue launches the reduction and its copy to the host,
// meanNormWait waits for that copy alone, so work queued in between keeps the GPU busy
// while the host turns round
static void meanNormQueue(Ctx& c, const double* v) {
    const int g = gridFor(c.D);
    hipLaunchKernelGGL(k_sum_sq, dim3(g), dim3(KBS), 0, c.stream, c.D, c.st, v, c.red.part);
    hipLaunchKernelGGL(k_finalize<3>, dim3(1), dim3(KBS), 0, c.stream, c.red.part, g, c.scal + SC_MISC + 4);
    HIPCHECK(hipMemcpyAsync(c.hscal + SC_MISC + 4, c.scal + SC_MISC + 4, 3 * sizeof(double), hipMemcpyDeviceToHost, c.stream));
    HIPCHECK(hipEventRecord(c.evR, c.stream));
}
static double meanNormWait(Ctx& c, double* norm) {
    HIPCHECK(hipEventSynchronize(c.evR));
    const double s = c.hscal[SC_MISC + 4], n = c.hscal[SC_MISC + 5], sq = c.hscal[SC_MISC + 6];
    const double m = n > 0. ? s / n : 0.;
    *norm = std::sqrt(std::max(sq - n * m * m, 0.));
    return m;
}

static double meanOf(Ctx& c, const double* v) {
    const int g = gridFor(c.D);
    hipLaunchKernelGGL(k_sum, dim3(g), dim3(KBS), 0, c.stream, c.D, c.st, v, c.red.part);
    hipLaunchKernelGGL(k_finalize<2>, dim3(1), dim3(KBS), 0, c.stream, c.red.part, g, c.scal + SC_MISC + 4);
    readBack(c, c.scal + SC_MISC + 4, SC_MISC + 4, 2);
    return c.hscal[SC_MISC + 5] > 0. ? c.hscal[SC_MISC + 4] / c.hscal[SC_MISC + 5] : 0.;
}

static OvView ovM(Ctx& c) { return OvView{c.hasIB ? c.ov.rowOf : nullptr, c.ov.cols, c.ov.valM}; }
static OvView ovP(Ctx& c) { return OvView{c.hasIB ? c.ov.rowOf : nullptr, c.ov.cols, c.ov.valP}; }
static OvView ovP32(Ctx& c) { return OvView{c.hasIB ? c.ov.rowOf : nullptr, c.ov.cols, c.ov.valP32}; }
static Band7 bandM(Ctx& c) { return Band7{c.Am, c.D.P, c.D.sJ, c.D.sK, ovM(c), c.st}; }
template <class T> static Sym4<T> symP(Ctx& c);
template <> Sym4<double> symP<double>(Ctx& c) { return Sym4<double>{c.Ap, c.D.P, c.D.sJ, c.D.sK, ovP(c), c.st}; }
template <> Sym4<float> symP<float>(Ctx& c) { return Sym4<float>{c.Ap32, c.D.P, c.D.sJ, c.D.sK, ovP32(c), c.st}; }

// true FP64 residual: r = (b - shift) - A x; returns ||r||, ||b - shift|| per system.
// For the singular (pure Neumann) pressure system the mean of r (rounding of the zero
// row sums) is not part of the residual of the problem and is left out of ||r||.
template <int NV, class Op>
static void trueResid(Ctx& c, Op A, const double* b, double shift, const double* x, double* r, double* rr, double* bb,
                      bool project = false) {
    const int g = gridFor(c.D);
    hipLaunchKernelGGL((k_resid<NV, Op>), dim3(g), dim3(KBS), 0, c.stream, c.D, A, b, shift, x, r, c.red.part);
    hipLaunchKernelGGL(k_finalize<3 * NV>, dim3(1), dim3(KBS), 0, c.stream, c.red.part, g, c.scal + SC_MISC + 8);
    readBack(c, c.scal + SC_MISC + 8, SC_MISC + 8, 3 * NV);
    for (int s = 0; s < NV; ++s) {
        const double r2 = c.hscal[SC_MISC + 8 + 3 * s], b2 = c.hscal[SC_MISC + 9 + 3 * s], rs = c.hscal[SC_MISC + 10 + 3 * s];
        rr[s] = std::sqrt(project ? std::max(r2 - rs * rs / (double)c.D.N, 0.) : r2);
        bb[s] = std::sqrt(b2);
    }
}

// BiCGStab over NV systems on the work block w (10 * NV * P entries of T).
// Returns the iteration count (max over systems); on exit done[] holds 1 (converged
// on the recurrence), 2 (breakdown) or 0 (iteration cap).
template <int NV, class T, class Op>
static int runBicgstab(Ctx& c, Op A, const double* b, double shift, T* x, T* w, double rtol, int maxit, int chunk,
                       int* doneOut) {
    const Dims& D = c.D;
    const int P = D.P, NVP = NV * P;
    T* r = w; T* rh = r + NVP; T* p0 = rh + NVP; T* p1 = p0 + NVP; T* v0 = p1 + NVP; T* v1 = v0 + NVP;
    T* s = v1 + NVP; T* t = s + NVP; T* dinv = t + NVP;
    double* st = c.scal + SC_BI;
    double* partC = c.red.part;
    double* partA = c.red.part + MAX_RED_BLOCKS * RED_STRIDE;
    double* partB = c.red.part + 2 * MAX_RED_BLOCKS * RED_STRIDE;
    const int g = gridFor(D);
    const ColMap cm = colMap(D, g);
    LoopCtl lc{rtol * rtol, maxit, 0.};
    hipLaunchKernelGGL((k_bi_init<NV, T, Op>), dim3(g), dim3(KBS), 0, c.stream, D, A, b, shift, r, rh, p0, v0, x, dinv, partC, st);
    T *pc = p0, *pn = p1, *vc = v0, *vn = v1;
    int it = 0;
    while (true) {
        for (int q = 0; q < chunk; ++q) {
            if constexpr (NV == 3 && std::is_same<Op, BandT<T>>::value) {
                if (!A.ov.rowOf) {
                    hipLaunchKernelGGL(k_bi_a_col<T>, dim3(g), dim3(KBS), 0, c.stream, D, A, cm, (const T*)r, (const T*)rh,
                                       (const T*)pc, pn, (const T*)vc, vn, (const T*)dinv, (const double*)partC, g, partA, st, lc);
                    hipLaunchKernelGGL(k_bi_b_col<T>, dim3(g), dim3(KBS), 0, c.stream, D, A, cm, (const T*)r, (const T*)vn, s, t,
                                       (const T*)dinv, (const double*)partA, g, partB, st);
                } else {
                    hipLaunchKernelGGL((k_bi_a_band<T, true>), dim3(g), dim3(KBS), 0, c.stream, D, A, (const T*)r, (const T*)rh,
                                       (const T*)pc, pn, (const T*)vc, vn, (const T*)dinv, (const double*)partC, g, partA, st, lc);
                    hipLaunchKernelGGL((k_bi_b_band<T, true>), dim3(g), dim3(KBS), 0, c.stream, D, A, (const T*)r, (const T*)vn, s, t,
                                       (const T*)dinv, (const double*)partA, g, partB, st);
                }
            } else {
                hipEvent_t e0, e1;
                // the p' BiCGStab (one system: ghost-row or indefinite p' systems) can be probed
                if (NV == 1 && c.probe.slot(std::is_same<T, float>::value ? PROBE_PBI_A_FP32 : PROBE_PBI_A_FP64, e0, e1))
                    hipExtLaunchKernelGGL((k_bi_a<NV, T, Op>), dim3(g), dim3(KBS), 0, c.stream, e0, e1, 0, D, A, (const T*)r,
                                          (const T*)rh, (const T*)pc, pn, (const T*)vc, vn, (const T*)dinv,
                                          (const double*)partC, g, partA, st, lc);
                else
                    hipLaunchKernelGGL((k_bi_a<NV, T, Op>), dim3(g), dim3(KBS), 0, c.stream, D, A, r, rh, pc, pn, vc, vn, dinv, partC, g, partA, st, lc);
                hipLaunchKernelGGL((k_bi_b<NV, T, Op>), dim3(g), dim3(KBS), 0, c.stream, D, A, r, vn, s, t, dinv, partA, g, partB, st);
            }
            hipLaunchKernelGGL((k_bi_c<NV, T, Op>), dim3(g), dim3(KBS), 0, c.stream, D, A, x, r, rh, pn, s, t, dinv, partB, g, partC, st);
            std::swap(pc, pn); std::swap(vc, vn);
        }
        readBack(c, st, SC_BI, B_STRIDE * NV);
        bool all = true;
        it = 0;
        // the next chunk: the iterations the observed rate of decrease still needs, worst
        // system, plus one (a system is seen converged by the head of the iteration after
        // its last); 2 without an estimate.  (Doubling the chunk queued up to 7 iterations
        // past convergence: their launches return at once but cost the host's launch time.)
        int need = 0;
        for (int q = 0; q < NV; ++q) {
            const double* h = c.hscal + SC_BI + B_STRIDE * q;
            all &= h[B_DONE] != 0.;
            it = std::max(it, (int)h[B_IT]);
            if (h[B_DONE] != 0.) continue;
            const double rr = h[B_RR], bb = h[B_BB], itq = h[B_IT];
            int est = 2;
            if (itq > 0. && rr > 0. && bb > 0. && rr < bb) {
                const double rate = std::log(rr / bb) / itq;
                est = (int)std::ceil(std::log(lc.tol2 * bb / rr) / rate) + 1;
            }
            need = std::max(need, est);
        }
        if (all || it >= maxit) break;
        chunk = std::max(1, std::min(16, need));
    }
    for (int q = 0; q < NV; ++q) doneOut[q] = (int)c.hscal[SC_BI + B_STRIDE * q + B_DONE];
    return it;
}

// ---------------------------------------------------------------- momentum
__global__ void k_add_f32_3(Dims D, double* x, const float* e) {
    FOR_CELLS(D, q, p) {
        x[p] += (double)e[p];
        x[D.P + p] += (double)e[D.P + p];
        x[2 * D.P + p] += (double)e[2 * D.P + p];
    }
}

// FP32 BiCGStab rounds inside FP64 refinement (ghost rows in FP64 values).  x starts at
// zero; each round solves A dx = r64 for the three components in FP32 from dx = 0 to a
// relative tolerance of at most 1e-5 (or only what is still needed), adds dx in FP64
// and recomputes the true FP64 residual, until the worst component meets rtol.
static int momentumIR(Ctx& c, const caffe_solve_opts& o, int* rounds, double* relOut) {
    const Dims& D = c.D;
    const int P = D.P;
    if (!c.Am32) {
        c.Am32 = falloc(c, 7 * (size_t)P);
        c.wm32 = falloc(c, 33 * (size_t)P);
    }
    hipLaunchKernelGGL(k_cast_f32, dim3(blocksFor(7 * P, KBS)), dim3(KBS), 0, c.stream, c.Am, c.Am32, 7 * P);
    Band7 A = bandM(c);
    // ghost rows keep their FP64 values (BandT reads the overlay in FP64 whatever T is)
    BandT<float> A32{c.Am32, P, D.sJ, D.sK, ovM(c), c.st};
    float* w = c.wm32;
    float* dx = c.wm32 + 30 * (size_t)P;
    double* rt = c.wm + 33 * (size_t)P;
    const int g = gridFor(D);
    hipLaunchKernelGGL(k_zero_d, dim3(blocksFor(3 * P, KBS)), dim3(KBS), 0, c.stream, c.xm, 3 * P);
    double rr[3], bb[3];
    auto worst = [&]() { double m = 0.; for (int s = 0; s < 3; ++s) m = std::max(m, bb[s] > 0. ? rr[s] / bb[s] : 0.); return m; };
    // x starts at zero, so the first round's right-hand side is b itself and its
    // relative residual is 1: no residual pass before it
    double rel = 1.;
    const double* rhs = c.bm;
    int iters = 0, done[3];
    *rounds = 0;
    while (rel > o.rtol && iters < o.maxit && *rounds < 20) {
        const double inner = std::min(1e-5 > 0.5 * o.rtol / rel ? 1e-5 : 0.5 * o.rtol / rel, 0.5);
        // the round queues as many iterations between read-backs as the same round of the
        // previous solve needed (4 at first)
        const int slot = std::min(*rounds, 3);
        const int it = runBicgstab<3, float>(c, A32, rhs, 0., dx, w, inner, o.maxit - iters, c.momChunk[slot], done);
        c.momChunk[slot] = std::max(1, std::min(32, it + 1));   // + 1: the launch that sees convergence
        iters += it;
        hipLaunchKernelGGL(k_add_f32_3, dim3(g), dim3(KBS), 0, c.stream, D, c.xm, (const float*)dx);
        trueResid<3>(c, A, c.bm, 0., c.xm, rt, rr, bb);
        rhs = rt;
        const double prev = rel;
        rel = worst();
        ++*rounds;
        if (!(rel < 0.5 * prev)) break;   // no progress in FP32: finish in FP64
    }
    *relOut = rel;
    return iters;
}

int momentum_solve(Ctx& c, const caffe_solve_opts& o, caffe_solve_report* rep) {
    HIPCHECK(hipEventRecord(c.ev0, c.stream));
    const int P = c.D.P;
    Band7 A = bandM(c);
    if (o.mode == CAFFE_SOLVE_IR_FP32) {
        int rounds = 0, fell = 0;
        double rel = 0.;
        int iters = momentumIR(c, o, &rounds, &rel);
        if (!(rel <= o.rtol)) {   // finish from the current x in FP64
            int done[3];
            double* dx = c.wm + 30 * P;
            double* rt = c.wm + 33 * P;
            double rr[3], bb[3];
            for (int r = 0; r < 10 && !(rel <= o.rtol) && iters < o.maxit; ++r) {
                const double sub = std::min(0.5, 0.5 * o.rtol / rel);
                iters += runBicgstab<3, double>(c, A, rt, 0., dx, c.wm, sub, o.maxit - iters, 4, done);
                add_vec(c, c.xm, dx, 3 * P);
                trueResid<3>(c, A, c.bm, 0., c.xm, rt, rr, bb);
                rel = 0.;
                for (int s = 0; s < 3; ++s) rel = std::max(rel, bb[s] > 0. ? rr[s] / bb[s] : 0.);
                ++rounds;
            }
            fell = 1;
        }
        HIPCHECK(hipEventRecord(c.ev1, c.stream));
        HIPCHECK(hipEventSynchronize(c.ev1));
        float ms = 0.f;
        HIPCHECK(hipEventElapsedTime(&ms, c.ev0, c.ev1));
        if (rep) {
            rep->iters = iters; rep->refinements = rounds; rep->relres = rel; rep->converged = rel <= o.rtol;
            rep->fell_back = fell; rep->ms = ms;
        }
        return rel <= o.rtol ? CAFFE_OK : CAFFE_ERR_NOT_CONVERGED;
    }
    int done[3];
    double* w = c.wm;              // 10 * 3 * P
    double* dx = c.wm + 30 * P;    // 3 * P
    double* rt = c.wm + 33 * P;    // 3 * P
    int iters = runBicgstab<3, double>(c, A, c.bm, 0., c.xm, w, o.rtol, o.maxit, 4, done);
    double rr[3], bb[3];
    trueResid<3>(c, A, c.bm, 0., c.xm, rt, rr, bb);
    auto worst = [&]() { double m = 0.; for (int s = 0; s < 3; ++s) m = std::max(m, bb[s] > 0. ? rr[s] / bb[s] : 0.); return m; };
    double rel = worst();
    int restarts = 0;
    while (rel > o.rtol && restarts < 10 && iters < o.maxit) {
        // continue from the true residual: solve A dx = r_true, x += dx
        const double sub = std::min(0.5, 0.5 * o.rtol / rel);
        iters += runBicgstab<3, double>(c, A, rt, 0., dx, w, sub, o.maxit - iters, 4, done);
        add_vec(c, c.xm, dx, 3 * P);
        trueResid<3>(c, A, c.bm, 0., c.xm, rt, rr, bb);
        rel = worst();
        ++restarts;
    }
    HIPCHECK(hipEventRecord(c.ev1, c.stream));
    HIPCHECK(hipEventSynchronize(c.ev1));
    float ms = 0.f;
    HIPCHECK(hipEventElapsedTime(&ms, c.ev0, c.ev1));
    if (rep) {
        rep->iters = iters; rep->refinements = restarts; rep->relres = rel; rep->converged = rel <= o.rtol;
        rep->fell_back = 0; rep->ms = ms;
    }
    return rel <= o.rtol ? CAFFE_OK : CAFFE_ERR_NOT_CONVERGED;
}

// ---------------------------------------------------------------- Coupled
// The 4-variable system of Coupled.cpp:54-234 (replacing A_[0].solve on it).  One BiCGStab
// over the joint vector (u, v, w, p blocks of P): its dot products add up all four blocks.
// Jacobi preconditioning (right): the diagonal is a_P / omega in the momentum rows and the
// d-Laplacian diagonal in the continuity rows.  In the CAFFE_SOLVE_MG mode the block-
// triangular multigrid preconditioner (cp_mg_run) replaces it.

// r = b (minus shift on the continuity rows), rh = r, p = v = x = 0, D^-1; partials (rh, r), (r, r)
__global__ void k_cp_init(Dims D, CpOp A, const double* b, double shift, double* r, double* rh, double* p0, double* v0,
                          double* x, double* dinv, double* part, double* st) {
    double acc[2] = {0., 0.};
    const int P = D.P;
    FOR_CELLS(D, q, p) {
        const double dm = 1. / A.A[CP_M * P + p], dp = 1. / A.A[CP_L * P + p];
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const int o = v * P + p;
            const double bv = v == 3 ? b[o] - shift : b[o];
            dinv[o] = v == 3 ? dp : dm;
            r[o] = bv; rh[o] = bv; p0[o] = 0.; v0[o] = 0.; x[o] = 0.;
            acc[0] += bv * bv;
            acc[1] += bv * bv;
        }
    }
    writePartials<KBS, 2>(acc, part);
    if (BLOCK0) {
        st[B_IT] = 0.; st[B_DONE] = 0.; st[B_BB] = -1.; st[B_RHOPREV] = 1.; st[B_ALPHA] = 1.; st[B_OMEGA] = 1.; st[B_BRK] = 0.;
    }
}

// KA: p_new = r + beta (p - omega v) at the stencil points; v = A D^-1 p_new; partial (rh, v)
__global__ void k_cp_a(Dims D, CpOp A, const double* __restrict__ r, const double* __restrict__ rh,
                       const double* __restrict__ pold, double* __restrict__ pnew, const double* __restrict__ vold,
                       double* __restrict__ vnew, const double* __restrict__ dinv, const double* partIn, int nb,
                       double* partOut, double* st, LoopCtl lc) {
    BiState<1> S;
    S.load(st);
    if (!S.any()) return;
    double beta, om;
    bool live;
    biaHead<1>(partIn, nb, st, lc, S, &beta, &om, &live);
    if (!live) return;
    double acc[1] = {0.};
    const int P = D.P;
    FOR_CELLS(D, q, p) {
        auto pf = [&](int v, int c) { const int o = v * P + c; return r[o] + beta * (pold[o] - om * vold[o]); };
        auto ph = [&](int v, int c) { return pf(v, c) * dinv[v * P + c]; };
        double y[4];
        A.apply(ph, p, y);
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const int o = v * P + p;
            pnew[o] = pf(v, p); vnew[o] = y[v];
            acc[0] += rh[o] * y[v];
        }
    }
    writePartials<KBS, 1>(acc, partOut);
}

// KB: s = r - alpha v at the stencil points; t = A D^-1 s; partials (t, s), (t, t)
__global__ void k_cp_b(Dims D, CpOp A, const double* __restrict__ r, const double* __restrict__ v,
                       double* __restrict__ sv, double* __restrict__ tv, const double* __restrict__ dinv,
                       const double* partIn, int nb, double* partOut, double* st) {
    BiState<1> S;
    S.load(st);
    if (!S.any()) return;
    double al;
    bool live;
    bibHead<1>(partIn, nb, st, S, &al, &live);
    if (!live) return;
    double acc[2] = {0., 0.};
    const int P = D.P;
    FOR_CELLS(D, q, p) {
        auto sf = [&](int w, int c) { const int o = w * P + c; return r[o] - al * v[o]; };
        auto sh = [&](int w, int c) { return sf(w, c) * dinv[w * P + c]; };
        double y[4];
        A.apply(sh, p, y);
#pragma unroll
        for (int w = 0; w < 4; ++w) {
            const int o = w * P + p;
            const double ss = sf(w, p);
            sv[o] = ss; tv[o] = y[w];
            acc[0] += y[w] * ss;
            acc[1] += y[w] * y[w];
        }
    }
    writePartials<KBS, 2>(acc, partOut);
}

// KC: omega = (t, s) / (t, t); x += alpha D^-1 p + omega D^-1 s; r = s - omega t; partials (rh, r), (r, r)
__global__ void k_cp_c(Dims D, double* x, double* r, const double* rh, const double* pv, const double* sv,
                       const double* tv, const double* dinv, const double* partIn, int nb, double* partOut, double* st) {
    const double sDone = st[B_DONE], sBrk = st[B_BRK], sAl = st[B_ALPHA], sRho = st[B_RHOCUR], sIt = st[B_IT];
    if (sDone != 0.) return;
    double t[2];
    readPartials<KBS, 2>(partIn, nb, t);
    // B_BRK is 1 when KB found (rh, v) = 0 (it then wrote no s, t); this kernel marks a zero
    // omega with 2, which the other blocks may already see: they still update, as block 0 does
    const bool live = sBrk != 1.;
    const double al = sAl;
    const double om = t[1] > 0. ? t[0] / t[1] : 0.;
    if (BLOCK0 && live) {
        st[B_OMEGA] = om; st[B_RHOPREV] = sRho; st[B_IT] = sIt + 1.;
        if (om == 0.) st[B_BRK] = 2.;
    }
    double acc[2] = {0., 0.};
    const int P = D.P;
    FOR_CELLS(D, q, p) {
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const int o = v * P + p;
            double rn = r[o];
            if (live) {
                const double di = dinv[o];
                x[o] = x[o] + al * (pv[o] * di) + om * (sv[o] * di);
                rn = sv[o] - om * tv[o];
                r[o] = rn;
            }
            acc[0] += rh[o] * rn;
            acc[1] += rn * rn;
        }
    }
    writePartials<KBS, 2>(acc, partOut);
}

// r = b (minus shift on the continuity rows) - A x in FP64; partials (r, r), (b, b), sum of
// the continuity block of r, sum of the p block of x
__global__ void k_cp_resid(Dims D, CpOp A, const double* b, double shift, const double* x, double* r, double* part) {
    double acc[4] = {0., 0., 0., 0.};
    const int P = D.P;
    FOR_CELLS(D, q, p) {
        auto xf = [&](int v, int c) { return x[v * P + c]; };
        double y[4];
        A.apply(xf, p, y);
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const int o = v * P + p;
            const double bv = v == 3 ? b[o] - shift : b[o];
            const double rv = bv - y[v];
            r[o] = rv;
            acc[0] += rv * rv;
            acc[1] += bv * bv;
        }
        acc[2] += r[3 * P + p];
        acc[3] += x[3 * P + p];
    }
    writePartials<KBS, 4>(acc, part);
}

__global__ void k_cp_sum_p(Dims D, const double* b, double* part) {
    double acc[1] = {0.};
    FOR_CELLS(D, q, p) { acc[0] += b[3 * D.P + p]; }
    writePartials<KBS, 1>(acc, part);
}
__global__ void k_cp_add(Dims D, double* x, const double* d) {
    FOR_CELLS(D, q, p) {
#pragma unroll
        for (int v = 0; v < 4; ++v) x[v * D.P + p] += d[v * D.P + p];
    }
}

// Jacobi BiCGStab on A x = b from x = 0, to the recurrence tolerance rtol
static int runCoupled(Ctx& c, CpOp A, const double* b, double shift, double* x, double rtol, int maxit, int chunk,
                      int* doneOut) {
    const Dims& D = c.D;
    const size_t V = 4 * (size_t)D.P;
    double* r = c.cw; double* rh = r + V; double* p0 = rh + V; double* p1 = p0 + V; double* v0 = p1 + V;
    double* v1 = v0 + V; double* s = v1 + V; double* t = s + V; double* dinv = t + V;
    double* st = c.scal + SC_BI;
    double* partC = c.red.part;
    double* partA = c.red.part + MAX_RED_BLOCKS * RED_STRIDE;
    double* partB = c.red.part + 2 * MAX_RED_BLOCKS * RED_STRIDE;
    const int g = gridFor(D);
    LoopCtl lc{rtol * rtol, maxit, 0.};
    hipLaunchKernelGGL(k_cp_init, dim3(g), dim3(KBS), 0, c.stream, D, A, b, shift, r, rh, p0, v0, x, dinv, partC, st);
    double *pc = p0, *pn = p1, *vc = v0, *vn = v1;
    int it = 0;
    while (true) {
        for (int q = 0; q < chunk; ++q) {
            hipEvent_t e0, e1;
            if (c.probe.slot(PROBE_CP_A, e0, e1))
                hipExtLaunchKernelGGL(k_cp_a, dim3(g), dim3(KBS), 0, c.stream, e0, e1, 0, D, A, (const double*)r,
                                      (const double*)rh, (const double*)pc, pn, (const double*)vc, vn, (const double*)dinv,
                                      (const double*)partC, g, partA, st, lc);
            else
                hipLaunchKernelGGL(k_cp_a, dim3(g), dim3(KBS), 0, c.stream, D, A, (const double*)r, (const double*)rh,
                                   (const double*)pc, pn, (const double*)vc, vn, (const double*)dinv, (const double*)partC, g,
                                   partA, st, lc);
            hipLaunchKernelGGL(k_cp_b, dim3(g), dim3(KBS), 0, c.stream, D, A, (const double*)r, (const double*)vn, s, t,
                               (const double*)dinv, (const double*)partA, g, partB, st);
            hipLaunchKernelGGL(k_cp_c, dim3(g), dim3(KBS), 0, c.stream, D, x, r, (const double*)rh, (const double*)pn,
                               (const double*)s, (const double*)t, (const double*)dinv, (const double*)partB, g, partC, st);
            std::swap(pc, pn); std::swap(vc, vn);
        }
        readBack(c, st, SC_BI, B_STRIDE);
        it = (int)c.hscal[SC_BI + B_IT];
        if (c.hscal[SC_BI + B_DONE] != 0. || it >= maxit) break;
        if (chunk < 32) chunk *= 2;
    }
    *doneOut = (int)c.hscal[SC_BI + B_DONE];
    return it;
}

// g = -A e_p (e_p: p = 1, u = 0): minus the row sums of the pressure coefficients;
// partial sum of g . g
__global__ void k_cp_nullrhs(Dims D, CpOp A, double* g, double* part) {
    double acc[1] = {0.};
    const int P = D.P;
    FOR_CELLS(D, q, p) {
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const int base = v == 3 ? CP_L : CP_G + 7 * v;
            double sum = 0.;
#pragma unroll
            for (int m = 0; m < 7; ++m) sum += A.A[(base + m) * P + p];
            g[v * P + p] = -sum;
            acc[0] += sum * sum;
        }
    }
    writePartials<KBS, 1>(acc, part);
}
// x -= beta (e_p + z)
__global__ void k_cp_gauge(Dims D, double* x, const double* z, double beta) {
    FOR_CELLS(D, q, p) {
#pragma unroll
        for (int v = 0; v < 4; ++v) x[v * D.P + p] -= beta * ((v == 3 ? 1. : 0.) + z[v * D.P + p]);
    }
}

namespace {
struct CpSolver {
    Ctx& c;
    CpOp A;
    int g;
    double* rt;
    double* dx;
    bool mg;
    // the multigrid runs queue first one iteration fewer than the last solve of the same kind
    // took in its first run (the first solve 4, a restart 2): a queued iteration past
    // convergence costs more than one more read-back.  lastIts is the main solve's count or
    // the null-vector solve's, so neither sizes the other's chunks.
    int run(const double* b, double shift, double* x, double rtol, int maxit, int* done, bool restart, int& lastIts) {
        if (!mg) return runCoupled(c, A, b, shift, x, rtol, maxit, 8, done);
        const int first = restart ? 2 : (lastIts > 1 ? lastIts - 1 : 4);
        const int it = cp_mg_run(c, A, b, shift, x, rtol, maxit, first, done);
        if (!restart) lastIts = it;
        return it;
    }
    // r = b - shift - A x: returns ||r|| (without the component along the continuity rows'
    // zero sum when singular), ||b - shift|| and the sum of the p block of x
    void resid(const double* b, double shift, const double* x, double* rr, double* bb, double* xsum) {
        const Dims& D = c.D;
        hipLaunchKernelGGL(k_cp_resid, dim3(g), dim3(KBS), 0, c.stream, D, A, b, shift, x, rt, c.red.part);
        hipLaunchKernelGGL(k_finalize<4>, dim3(1), dim3(KBS), 0, c.stream, c.red.part, g, c.scal + SC_MISC + 8);
        readBack(c, c.scal + SC_MISC + 8, SC_MISC + 8, 4);
        const double r2 = c.hscal[SC_MISC + 8], rs = c.hscal[SC_MISC + 10];
        *rr = std::sqrt(c.singularP ? std::max(r2 - rs * rs / (double)D.N, 0.) : r2);
        *bb = std::sqrt(c.hscal[SC_MISC + 9]);
        *xsum = c.hscal[SC_MISC + 11];
    }
    double meanP(const double* v) {
        const Dims& D = c.D;
        hipLaunchKernelGGL(k_cp_sum_p, dim3(g), dim3(KBS), 0, c.stream, D, v, c.red.part);
        hipLaunchKernelGGL(k_finalize<1>, dim3(1), dim3(KBS), 0, c.stream, c.red.part, g, c.scal + SC_MISC + 4);
        readBack(c, c.scal + SC_MISC + 4, SC_MISC + 4, 1);
        return c.hscal[SC_MISC + 4] / (double)D.N;
    }
    // A x = b - shift from x = 0 to a true relative residual rtol: BiCGStab, restarted from the
    // true residual while that gains; returns the relative residual
    double solve(const double* b, double shift, double* x, double rtol, int maxit, int* iters, int* restarts,
                 double* xsum, int& lastIts) {
        int done = 0;
        *iters += run(b, shift, x, rtol, maxit - *iters, &done, false, lastIts);
        double rr, bb;
        resid(b, shift, x, &rr, &bb, xsum);
        double rel = bb > 0. ? rr / bb : 0.;
        for (int k = 0; rel > rtol && k < 20 && *iters < maxit; ++k) {
            const double sub = std::min(0.5, 0.5 * rtol / rel);
            *iters += run(rt, 0., dx, sub, maxit - *iters, &done, true, lastIts);
            hipLaunchKernelGGL(k_cp_add, dim3(g), dim3(KBS), 0, c.stream, c.D, x, (const double*)dx);
            const double prev = rel;
            resid(b, shift, x, &rr, &bb, xsum);
            rel = bb > 0. ? rr / bb : 0.;
            ++*restarts;
            if (!(rel < prev)) break;
        }
        return rel;
    }
};
}  // namespace

int coupled_solve(Ctx& c, const caffe_solve_opts& o, caffe_solve_report* rep) {
    HIPCHECK(hipEventRecord(c.ev0, c.stream));
    const Dims& D = c.D;
    const size_t V = 4 * (size_t)D.P;
    const bool wantMg = o.mode == CAFFE_SOLVE_MG;
    const bool mg = wantMg && cp_mg_setup(c);
    CpSolver S{c, CpOp{c.cA, D.P, D.sJ, D.sK}, gridFor(D), c.cw + 36 * (size_t)D.P, c.cw + 40 * (size_t)D.P, mg};
    double* gv = c.cw + 44 * (size_t)D.P;   // -A e_p
    double* z = gv + V;                     // A z = -A e_p
    // Without a fixed-pressure face the continuity rows add up to zero: the system is
    // singular, its right-hand side loses its (rounding) mean on those rows, and the solution
    // is moved along the null vector n = e_p + z (A z = -A e_p) to zero-mean p.  With planar
    // faces A e_p = 0 and n is the constant pressure; on cells with non-planar faces (normals
    // from one triangle of each face) n carries velocity components, and z costs one more
    // solve, made only when a plain shift of p would move the residual by more than
    // 1e-3 rtol.
    const double shift = c.singularP ? S.meanP(c.cb) : 0.;
    int iters = 0, restarts = 0;
    double xs = 0.;
    double rel = S.solve(c.cb, shift, c.cx, o.rtol, o.maxit, &iters, &restarts, &xs, c.mg.lastIts);
    if (c.singularP) {
        const double xbar = xs / (double)D.N;
        const int g = S.g;
        hipLaunchKernelGGL(k_cp_nullrhs, dim3(g), dim3(KBS), 0, c.stream, D, S.A, gv, c.red.part);
        hipLaunchKernelGGL(k_finalize<1>, dim3(1), dim3(KBS), 0, c.stream, c.red.part, g, c.scal + SC_MISC + 4);
        readBack(c, c.scal + SC_MISC + 4, SC_MISC + 4, 1);
        const double gnorm = std::sqrt(c.hscal[SC_MISC + 4]);
        double rr, bb, tmp;
        S.resid(c.cb, shift, c.cx, &rr, &bb, &tmp);
        double zbar = 0.;
        bool plain = true;
        if (std::fabs(xbar) * gnorm > 1e-3 * o.rtol * bb) {
            double zs = 0.;
            const double zsh = S.meanP(gv);
            const double zrel = S.solve(gv, zsh, z, o.rtol, o.maxit, &iters, &restarts, &zs, c.mg.lastItsGauge);
            zbar = zs / (double)D.N;
            plain = !(zrel <= o.rtol) || !std::isfinite(zbar);   // no null vector: shift p only
        }
        if (plain) {
            zbar = 0.;
            hipLaunchKernelGGL(k_zero_d, dim3(blocksFor((int)V, KBS)), dim3(KBS), 0, c.stream, z, (int)V);
        }
        hipLaunchKernelGGL(k_cp_gauge, dim3(g), dim3(KBS), 0, c.stream, D, c.cx, (const double*)z, xbar / (1. + zbar));
        S.resid(c.cb, shift, c.cx, &rr, &bb, &tmp);
        rel = bb > 0. ? rr / bb : 0.;
    }
    HIPCHECK(hipEventRecord(c.ev1, c.stream));
    HIPCHECK(hipEventSynchronize(c.ev1));
    float ms = 0.f;
    HIPCHECK(hipEventElapsedTime(&ms, c.ev0, c.ev1));
    if (rep) {
        rep->iters = iters; rep->refinements = restarts; rep->relres = rel; rep->converged = rel <= o.rtol;
        rep->fell_back = wantMg && !mg; rep->ms = ms;
    }
    return rel <= o.rtol ? CAFFE_OK : CAFFE_ERR_NOT_CONVERGED;
}

void pcorr_make_fp32(Ctx& c) {
    const int n = 4 * c.D.P;
    hipLaunchKernelGGL(k_cast_f32, dim3(blocksFor(n, KBS)), dim3(KBS), 0, c.stream, c.Ap, c.Ap32, n);
    if (c.hasIB && c.ov.n > 0) {
        const int m = 9 * c.ov.n;
        hipLaunchKernelGGL(k_cast_f32, dim3(blocksFor(m, KBS)), dim3(KBS), 0, c.stream, c.ov.valP, c.ov.valP32, m);
    }
}

// ---------------------------------------------------------------- pressure correction
// CG on the symmetric p' system, continuing from the current x in c.xp.  T = double:
// the iterate is c.xp itself.  T = float: the FP32 iterate e holds a correction in
// units of the scale s (r32 = s r64); at every reliable update x64 += e / s, the true
// residual r64 = b - A x64 is recomputed in FP64, r32 restarts from it with a new
// scale and the search direction is rescaled.  Updates happen when the recurrence
// residual has dropped by `delta` since the last one.
template <class T>
static int cgSolve(Ctx& c, double shift, double rtol, int maxit, double delta, int* updates, double* relOut) {
    const Dims& D = c.D;
    const int P = D.P;
    const int g = gridFor(D);
    constexpr bool mixed = std::is_same<T, float>::value;
    Sym4<T> A = symP<T>(c);
    Sym4<double> A64 = symP<double>(c);
    T* w = mixed ? (T*)c.wp32 : (T*)c.wp;
    T* r = w; T* z = r + P; T* p0 = z + P; T* p1 = p0 + P; T* qv = p1 + P; T* e = qv + P;
    double* r64 = c.wp + 6 * P;
    double* st = c.scal + SC_CG;
    double* partB = c.red.part;                                   // (r,z), (r,r)
    double* partA = c.red.part + MAX_RED_BLOCKS * RED_STRIDE;     // (p,q)
    double rr, bb;
    trueResid<1>(c, A64, c.bp, shift, c.xp, r64, &rr, &bb, c.singularP);
    double rel = bb > 0. ? rr / bb : 0.;
    *updates = 0;
    if (rel <= rtol || bb == 0.) { *relOut = rel; return 0; }
    double scale = mixed ? 1.0 / rr : 1.0;
    T* x = mixed ? e : (T*)c.xp;
    // the double iterate continues from xp (x is not zeroed); the FP32 correction starts at 0
    hipLaunchKernelGGL((k_cg_init<T, double>), dim3(g), dim3(KBS), 0, c.stream, D, A, r64, 0., scale, r, z, p0, p1, x, 1,
                       mixed ? 1 : 0, partB, st, 1.0);
    const double tolIn = mixed ? delta : std::min(0.5, rtol / rel);
    LoopCtl lc{tolIn * tolIn, maxit, c.singularP ? (double)D.N : 0.};
    int it = 0;
    int chunk = 8;
    while (true) {
        for (int q = 0; q < chunk; ++q) {
            hipEvent_t e0, e1;
            if (c.probe.slot(mixed ? PROBE_CG_K1_FP32 : PROBE_CG_K1_FP64, e0, e1))
                hipExtLaunchKernelGGL((k_cg_k1<T>), dim3(g), dim3(KBS), 0, c.stream, e0, e1, 0, D, A, (const T*)z, p0, p1, qv,
                                      (const double*)partB, g, partA, st, lc);
            else
                hipLaunchKernelGGL((k_cg_k1<T>), dim3(g), dim3(KBS), 0, c.stream, D, A, z, p0, p1, qv, partB, g, partA, st, lc);
            hipLaunchKernelGGL((k_cg_k2<T>), dim3(g), dim3(KBS), 0, c.stream, D, A, x, r, z, p0, p1, qv, partA, g, partB, st);
        }
        readBack(c, st, SC_CG, 8);
        it = (int)c.hscal[SC_CG + C_IT];
        const bool recDone = c.hscal[SC_CG + C_DONE] != 0.;
        if (chunk < 32) chunk *= 2;
        if (!recDone && it < maxit) continue;
        // reliable update / convergence check on the true FP64 residual
        if (mixed) hipLaunchKernelGGL(k_axpy_mixed, dim3(g), dim3(KBS), 0, c.stream, D, c.xp, (const float*)e, 1.0 / scale);
        trueResid<1>(c, A64, c.bp, shift, c.xp, r64, &rr, &bb, c.singularP);
        rel = rr / bb;
        ++*updates;
        if (rel <= rtol || it >= maxit || *updates > 200) break;
        // restart the recurrence from the true residual, keeping the search direction
        const double ns = mixed ? 1.0 / rr : 1.0;
        const double ratio = mixed ? scale / ns : 1.0;
        hipLaunchKernelGGL((k_cg_init<T, double>), dim3(g), dim3(KBS), 0, c.stream, D, A, r64, 0., ns, r, z, p0, p1, x, 0,
                           mixed ? 1 : 0, partB, st, ratio);
        scale = ns;
        const double tolNext = mixed ? delta : std::min(0.5, rtol / rel);
        lc.tol2 = tolNext * tolNext;
    }
    *relOut = rel;
    return it;
}

// ---------------------------------------------------------------- red-black SGS-PCG
// CG preconditioned by one symmetric Gauss-Seidel sweep (SSOR with omega = 1) in red-black
// order, the parallel analogue of PETSc's PCSOR with its = lits = 1 and the symmetric sweep
// (SparseMatrix.cpp:97-98).  Cell (i, j, k) is red when i + j + k is even.  z = M^-1 r is the
// forward sweep (red: z = r / aP; black: z = (r - sum a z_red) / aP) followed by the backward
// sweep, whose black half repeats the forward one (black cells couple only to red cells) and
// is skipped, and whose red half is z = (r - sum a z_black) / aP.  M is symmetric, so the
// iteration is CG.  An iteration: k_cg_k1 (direction, A p), k_sgs_upd (x, r and the forward
// red half), k_sgs_black, k_sgs_red (backward red half, partial sums for the next k_cg_k1).
__device__ __forceinline__ int cellColour(const Dims& D, int q) {
    int i, j, k;
    D.cell(q, i, j, k);
    return (i + j + k) & 1;
}
// the off-diagonal part of row p applied to z
__device__ __forceinline__ double offDiag(const Sym4<double>& A, const double* z, int p) {
    return A.A[A.P + p] * z[p - 1] + A.A[A.P + p + 1] * z[p + 1] + A.A[2 * A.P + p] * z[p - A.sJ] +
           A.A[2 * A.P + p + A.sJ] * z[p + A.sJ] + A.A[3 * A.P + p] * z[p - A.sK] + A.A[3 * A.P + p + A.sK] * z[p + A.sK];
}
// [alpha = rz / (p, q)] x += alpha p; r -= alpha q; red cells: z = r / aP
__global__ void k_sgs_upd(Dims D, Sym4<double> A, double* x, double* r, double* z, const double* pb0, const double* pb1,
                          const double* qv, const double* partIn, int nb, double* st) {
    const double sDone = st[C_DONE], sPbuf = st[C_PBUF], sRz = st[C_RZCUR], sIt = st[C_IT];
    if (sDone != 0.) return;
    const double* pv = sPbuf != 0. ? pb1 : pb0;
    double t[1];
    readPartials<KBS, 1>(partIn, nb, t);
    const double alpha = t[0] != 0. ? sRz / t[0] : 0.;
    if (BLOCK0) { st[C_RZPREV] = sRz; st[C_IT] = sIt + 1.; st[C_RATIO] = 1.; st[C_PQ] = t[0]; }
    FOR_CELLS(D, q, p) {
        x[p] += alpha * pv[p];
        const double rn = r[p] - alpha * qv[p];
        r[p] = rn;
        if (cellColour(D, q) == 0) z[p] = rn / A.diag(p);
    }
}
// black cells: z = (r - sum_nb a z_red) / aP
__global__ void k_sgs_black(Dims D, Sym4<double> A, const double* r, double* z, const double* st) {
    if (st[C_DONE] != 0.) return;
    FOR_CELLS(D, q, p) {
        if (cellColour(D, q) == 1) z[p] = (r[p] - offDiag(A, z, p)) / A.diag(p);
    }
}
// red cells: z = (r - sum_nb a z_black) / aP; partials (r, z), (r, r), sum z, sum r over
// all cells (as k_cg_k2 writes them)
__global__ void k_sgs_red(Dims D, Sym4<double> A, const double* r, double* z, double* partOut, const double* st) {
    if (st[C_DONE] != 0.) return;
    double v[4] = {0., 0., 0., 0.};
    FOR_CELLS(D, q, p) {
        const double rv = r[p];
        double zv;
        if (cellColour(D, q) == 0) { zv = (rv - offDiag(A, z, p)) / A.diag(p); z[p] = zv; }
        else zv = z[p];
        v[0] += rv * zv; v[1] += rv * rv; v[2] += zv; v[3] += rv;
    }
    writePartials<KBS, 4>(v, partOut);
}

// SGS-PCG in FP64 on the symmetric p' system from the current c.xp, to the true FP64 relative
// residual rtol (the recurrence aims at what is still needed; restarts from the true
// residual keep the direction, as cgSolve does)
static int sgsCgSolve(Ctx& c, double shift, double rtol, int maxit, int* updates, double* relOut) {
    const Dims& D = c.D;
    const int P = D.P;
    const int g = gridFor(D);
    Sym4<double> A = symP<double>(c);
    double* w = c.wp;
    double* r = w; double* z = r + P; double* p0 = z + P; double* p1 = p0 + P; double* qv = p1 + P;
    double* r64 = c.wp + 6 * P;
    double* st = c.scal + SC_CG;
    double* partB = c.red.part;
    double* partA = c.red.part + MAX_RED_BLOCKS * RED_STRIDE;
    double rr, bb;
    trueResid<1>(c, A, c.bp, shift, c.xp, r64, &rr, &bb, c.singularP);
    double rel = bb > 0. ? rr / bb : 0.;
    *updates = 0;
    if (rel <= rtol || bb == 0.) { *relOut = rel; return 0; }
    // k_cg_init: r = b - A x (from r64), z = r / aP on every cell (the forward red half for the
    // red cells), x kept; then the black half and the backward red half complete z = M^-1 r
    auto start = [&](int zeroP) {
        hipLaunchKernelGGL((k_cg_init<double, double>), dim3(g), dim3(KBS), 0, c.stream, D, A, r64, 0., 1.0, r, z, p0, p1,
                           (double*)c.xp, zeroP, 0, partB, st, 1.0);
        hipLaunchKernelGGL(k_sgs_black, dim3(g), dim3(KBS), 0, c.stream, D, A, (const double*)r, z, (const double*)st);
        hipLaunchKernelGGL(k_sgs_red, dim3(g), dim3(KBS), 0, c.stream, D, A, (const double*)r, z, partB, (const double*)st);
    };
    start(1);
    double tolIn = std::min(0.5, rtol / rel);
    LoopCtl lc{tolIn * tolIn, maxit, c.singularP ? (double)D.N : 0.};
    int it = 0, chunk = 8;
    while (true) {
        for (int q = 0; q < chunk; ++q) {
            hipLaunchKernelGGL((k_cg_k1<double>), dim3(g), dim3(KBS), 0, c.stream, D, A, (const double*)z, p0, p1, qv,
                               (const double*)partB, g, partA, st, lc);
            hipLaunchKernelGGL(k_sgs_upd, dim3(g), dim3(KBS), 0, c.stream, D, A, (double*)c.xp, r, z, (const double*)p0,
                               (const double*)p1, (const double*)qv, (const double*)partA, g, st);
            hipLaunchKernelGGL(k_sgs_black, dim3(g), dim3(KBS), 0, c.stream, D, A, (const double*)r, z, (const double*)st);
            hipLaunchKernelGGL(k_sgs_red, dim3(g), dim3(KBS), 0, c.stream, D, A, (const double*)r, z, partB, (const double*)st);
        }
        readBack(c, st, SC_CG, 8);
        it = (int)c.hscal[SC_CG + C_IT];
        const bool recDone = c.hscal[SC_CG + C_DONE] != 0.;
        if (chunk < 32) chunk *= 2;
        if (!recDone && it < maxit) continue;
        trueResid<1>(c, A, c.bp, shift, c.xp, r64, &rr, &bb, c.singularP);
        rel = rr / bb;
        ++*updates;
        if (rel <= rtol || it >= maxit || *updates > 200) break;
        start(0);
        tolIn = std::min(0.5, rtol / rel);
        lc.tol2 = tolIn * tolIn;
    }
    *relOut = rel;
    return it;
}

// BiCGStab on the nonsymmetric (ghost-row) p' system, restarted from the true FP64
// residual; T = float runs each restart in FP32 to relative tolerance `inner`.
template <class T>
static int bicgPressure(Ctx& c, double shift, double rtol, int maxit, double inner, int* refs, double* relOut) {
    const Dims& D = c.D;
    const int P = D.P;
    Sym4<T> A = symP<T>(c);
    Sym4<double> A64 = symP<double>(c);
    // FP64: the momentum work block (double views only); FP32: its own block
    constexpr bool f32 = std::is_same<T, float>::value;
    T* w = f32 ? (T*)c.wb32 : (T*)c.wm;
    T* e = f32 ? (T*)(c.wb32 + 10 * (size_t)P) : (T*)(c.wm + 12 * (size_t)P);
    double* r64 = c.wp + 6 * P;
    double rr, bb;
    trueResid<1>(c, A64, c.bp, shift, c.xp, r64, &rr, &bb);
    double rel = bb > 0. ? rr / bb : 0.;
    int it = 0, done[1];
    *refs = 0;
    while (rel > rtol && it < maxit && *refs < 60) {
        const double tol = std::is_same<T, float>::value ? inner : std::min(0.5, 0.5 * rtol / rel);
        it += runBicgstab<1, T>(c, A, r64, 0., e, w, tol, maxit - it, 8, done);
        if (std::is_same<T, float>::value) hipLaunchKernelGGL(k_axpy_mixed, dim3(gridFor(D)), dim3(KBS), 0, c.stream, D, c.xp, (const float*)e, 1.0);
        else add_vec(c, c.xp, (const double*)e, P);
        const double prev = rel;
        trueResid<1>(c, A64, c.bp, shift, c.xp, r64, &rr, &bb);
        rel = rr / bb;
        ++*refs;
        if (!(rel < prev)) break;
    }
    *relOut = rel;
    return it;
}

void pcorr_true_resid(Ctx& c, double shift, double* rr, double* bb) {
    trueResid<1>(c, symP<double>(c), c.bp, shift, c.xp, c.wp + 6 * c.D.P, rr, bb, c.singularP);
}

void pcorr_true_resid_dev(Ctx& c, double shift) {
    const int g = gridFor(c.D);
    hipLaunchKernelGGL((k_resid<1, Sym4<double>>), dim3(g), dim3(KBS), 0, c.stream, c.D, symP<double>(c), (const double*)c.bp,
                       shift, (const double*)c.xp, c.wp + 6 * c.D.P, c.red.part);
    hipLaunchKernelGGL(k_finalize<3>, dim3(1), dim3(KBS), 0, c.stream, c.red.part, g, c.scal + SC_MISC + 8);
}

// The initial guess x = g0 x0 + g1 x1 that minimises |b - shift - A x|: q_i = A x_i with the
// partial sums of the normal equations (q0,q0) (q0,q1) (q1,q1) (q0,b') (q1,b'), then the
// 2 x 2 solve and the combination (g1 = 0 without x1, or when q0 and q1 are too close to
// parallel for the solve; g = 0 when q0 = 0; inactive cells keep x = 0)
template <bool TWO>
__global__ void k_guess_dots(Dims D, Sym4<double> A, const double* b, double shift, const double* x0, const double* x1,
                             double* part) {
    double acc[5] = {0., 0., 0., 0., 0.};
    const int P = D.P, sJ = D.sJ, sK = D.sK;
    FOR_CELLS(D, q, p) {
        if (A.st && A.st[p] == ST_INACTIVE) continue;
        const double bv = b[p] - shift;
        double q0, q1 = 0.;
        if (A.ov.rowOf) {   // ghost rows (the learned solve takes none): the generic rows
            q0 = A.apply([&](int c) { return x0[c]; }, p);
            if (TWO) q1 = A.apply([&](int c) { return x1[c]; }, p);
        } else {            // the seven coefficients loaded once for both products, in apply's order
            const double a[7] = {A.A[p], A.A[P + p], A.A[P + p + 1], A.A[2 * P + p], A.A[2 * P + p + sJ], A.A[3 * P + p],
                                 A.A[3 * P + p + sK]};
            const int o[7] = {0, -1, 1, -sJ, sJ, -sK, sK};
            auto prod = [&](const double* x) {
                return a[0] * x[p + o[0]] + a[1] * x[p + o[1]] + a[2] * x[p + o[2]] + a[3] * x[p + o[3]] + a[4] * x[p + o[4]]
                     + a[5] * x[p + o[5]] + a[6] * x[p + o[6]];
            };
            q0 = prod(x0);
            if (TWO) q1 = prod(x1);
        }
        acc[0] += q0 * q0;
        acc[3] += q0 * bv;
        if (TWO) {
            acc[1] += q0 * q1;
            acc[2] += q1 * q1;
            acc[4] += q1 * bv;
        }
    }
    writePartials<KBS, 5>(acc, part);
}
template <bool TWO>
__global__ void k_guess_scale(Dims D, const int8_t* st, const double* x0, const double* x1, double* x, const double* part,
                              int nb) {
    double t[5];
    readPartials<KBS, 5>(part, nb, t);
    double g0 = t[0] > 0. ? t[3] / t[0] : 0., g1 = 0.;
    if (TWO) {
        const double det = t[0] * t[2] - t[1] * t[1];
        if (det > 1e-10 * t[0] * t[2]) {
            g0 = (t[3] * t[2] - t[1] * t[4]) / det;
            g1 = (t[0] * t[4] - t[1] * t[3]) / det;
        }
    }
    FOR_CELLS(D, q, p) {
        const double v = TWO ? g0 * x0[p] + g1 * x1[p] : g0 * x0[p];
        x[p] = (st && st[p] == ST_INACTIVE) ? 0. : v;
    }
}
// The p' solves' initial guesses.  The k-th p' solve after a momentum assembly (the k-th
// corrector) takes slot min(k, 2) - 1 of a ring per slot holding the solutions of the two
// previous steps; solves past the algorithm's correctors in a step (a caller solving the same
// system again) neither take nor leave a guess, so they start from x = 0.
static int guessSlot(const Ctx& c) {
    const int nPC = c.cs.algo == CAFFE_SIMPLE ? 1 : std::max(c.cs.nPCorr, 1);
    const int k = c.pcSolvesThisStep;   // 1 for the first solve after the assembly
    return c.fcgGuess && k >= 1 && k <= nPC ? std::min(k, 2) - 1 : -1;
}
void pcorr_apply_guess(Ctx& c, double shift) {
    const int slot = guessSlot(c);
    if (slot < 0) return;
    const int e1 = (int)((c.stepGen - 1) & 1), e2 = (int)((c.stepGen - 2) & 1);
    const double* x1 = c.fcgPrevGen[slot][e1] == c.stepGen - 1 ? c.fcgPrev[slot][e1] : nullptr;
    const double* x2 = c.fcgPrevGen[slot][e2] == c.stepGen - 2 ? c.fcgPrev[slot][e2] : nullptr;
    if (x1) pcorr_scaled_guess(c, x1, x2, shift);
}
void pcorr_save_guess(Ctx& c) {
    const int slot = guessSlot(c);
    if (slot < 0) return;
    const int en = (int)(c.stepGen & 1);
    if (!c.fcgPrev[slot][en]) c.fcgPrev[slot][en] = dalloc64(c, c.D.P);
    HIPCHECK(hipMemcpyAsync(c.fcgPrev[slot][en], c.xp, (size_t)c.D.P * sizeof(double), hipMemcpyDeviceToDevice, c.stream));
    c.fcgPrevGen[slot][en] = c.stepGen;
}

void pcorr_scaled_guess(Ctx& c, const double* x0, const double* x1, double shift) {
    const int g = gridFor(c.D);
    if (x1) {
        hipLaunchKernelGGL(k_guess_dots<true>, dim3(g), dim3(KBS), 0, c.stream, c.D, symP<double>(c), (const double*)c.bp, shift,
                           x0, x1, c.red.part);
        hipLaunchKernelGGL(k_guess_scale<true>, dim3(g), dim3(KBS), 0, c.stream, c.D, (const int8_t*)c.st, x0, x1, c.xp,
                           (const double*)c.red.part, g);
    } else {
        hipLaunchKernelGGL(k_guess_dots<false>, dim3(g), dim3(KBS), 0, c.stream, c.D, symP<double>(c), (const double*)c.bp,
                           shift, x0, x0, c.red.part);
        hipLaunchKernelGGL(k_guess_scale<false>, dim3(g), dim3(KBS), 0, c.stream, c.D, (const int8_t*)c.st, x0, x0, c.xp,
                           (const double*)c.red.part, g);
    }
}

// the solution's mean over the cells that carry a row, removed on the device
void pcorr_remove_mean(Ctx& c) {
    const int g = gridFor(c.D);
    hipLaunchKernelGGL(k_sum, dim3(g), dim3(KBS), 0, c.stream, c.D, c.st, (const double*)c.xp, c.red.part);
    hipLaunchKernelGGL(k_finalize<2>, dim3(1), dim3(KBS), 0, c.stream, c.red.part, g, c.scal + SC_MISC + 4);
    hipLaunchKernelGGL(k_sub_mean, dim3(g), dim3(KBS), 0, c.stream, c.D, c.st, c.xp, (const double*)(c.scal + SC_MISC + 4));
}

// the time of a solve whose report left it pending
void pcorr_resolve_ms(Ctx& c) {
    if (!c.pcMsPending) return;
    HIPCHECK(hipEventSynchronize(c.evP1));
    float ms = 0.f;
    HIPCHECK(hipEventElapsedTime(&ms, c.evP0, c.evP1));
    c.lastPc.ms = ms;
    c.pcMsPending = false;
}

int pcorr_solve(Ctx& c, const caffe_solve_opts& o, caffe_solve_report* rep, bool lazyMs) {
    c.pcMsPending = false;   // a pending time belongs to the report this solve replaces
    ++c.pcSolvesThisStep;
    HIPCHECK(hipEventRecord(c.evP0, c.stream));
    const Dims& D = c.D;
    const int P = D.P;
    const bool fp32 = o.mode != CAFFE_SOLVE_PLAIN_FP64 && o.mode != CAFFE_SOLVE_SGS_FP64;
    // Neumann null space: the right-hand side loses its (rounding) mean, the solution is
    // returned with zero mean over the cells that carry a row.  The mean's read-back is
    // waited for after the launches that do not need it.
    double bNorm = 0.;   // |b - shift| where known without another read-back (0: unknown)
    if (c.singularP) meanNormQueue(c, c.bp);
    hipLaunchKernelGGL(k_zero_d, dim3(blocksFor(P, KBS)), dim3(KBS), 0, c.stream, c.xp, P);
    if (fp32) pcorr_make_fp32(c);
    const double shift = c.singularP ? meanNormWait(c, &bNorm) : 0.;
    int iters = 0, refs = 0, fell = 0;
    double rel = 1.;
    bool meanDone = false;   // the solution's mean already removed (queued by the learned solve)
    const bool sym = !c.hasIB && !c.indefiniteP;
    if (o.mode == CAFFE_SOLVE_PLAIN_FP64) {
        if (sym) iters += cgSolve<double>(c, shift, o.rtol, o.maxit, 0., &refs, &rel);
        else iters += bicgPressure<double>(c, shift, o.rtol, o.maxit, 0., &refs, &rel);
    } else if (o.mode == CAFFE_SOLVE_SGS_FP64) {
        // ghost-row systems (not a 7-point red-black coupling) and indefinite ones take the
        // plain solve
        if (sym && !c.hasIB) iters += sgsCgSolve(c, shift, o.rtol, o.maxit, &refs, &rel);
        else { fell = 1; iters += bicgPressure<double>(c, shift, o.rtol, o.maxit, 0., &refs, &rel); }
    } else {
        bool learnedDone = false;
        if (o.mode == CAFFE_SOLVE_LEARNED_FCG) {
            // learned two-grid cycle inside flexible CG; the inner FP32 solve restarts from
            // the true FP64 residual each time it has gained `inner_rtol` (default 1e-4)
            bool meanQ = false;
            const int it = c.net ? learned_fcg(c, shift, bNorm, o.rtol, o.maxit, o.inner_rtol > 0. ? o.inner_rtol : 1e-4, &refs, &rel,
                                               &meanQ) : -1;
            if (it >= 0) { iters += it; learnedDone = true; meanDone = meanQ; }
            else fell = 1;   // no network, or a system it does not apply to: IR below
        }
        if (!learnedDone) {
            const double delta = o.inner_rtol > 0. ? o.inner_rtol : 1e-2;
            if (sym && o.mode != CAFFE_SOLVE_MG) {
                iters += cgSolve<float>(c, shift, o.rtol, o.maxit, delta, &refs, &rel);
            } else {
                // the multigrid mode, and ghost-row systems in the FP32 modes: FP32 Jacobi
                // BiCGStab does not reach 1e-8 on those (on the 256x64x64 channel every inner
                // tolerance from 1e-2 to 1e-5 fell back), the FP64 recurrence with FP32
                // V-cycles does.  Systems the cycle does not take go to the FP64 solve.
                const int it = mg_bicgstab(c, shift, o.rtol, o.maxit, &refs, &rel);
                if (it >= 0) {
                    iters += it;
                } else {
                    fell = 1;
                    iters += bicgPressure<double>(c, shift, o.rtol, o.maxit, 0., &refs, &rel);
                }
            }
        }
        if (!(rel <= o.rtol)) {  // the plain FP64 solve stays as fallback and judge, from x = 0 (also for NaN)
            fell = 1;
            meanDone = false;
            int r2 = 0;
            hipLaunchKernelGGL(k_zero_d, dim3(blocksFor(P, KBS)), dim3(KBS), 0, c.stream, c.xp, P);
            if (sym) iters += cgSolve<double>(c, shift, o.rtol, o.maxit, 0., &r2, &rel);
            else iters += bicgPressure<double>(c, shift, o.rtol, o.maxit, 0., &r2, &rel);
            refs += r2;
        }
    }
    if (c.singularP && !meanDone) pcorr_remove_mean(c);   // the mean stays on the device: no read-back
    HIPCHECK(hipEventRecord(c.evP1, c.stream));
    if (rep) {
        rep->iters = iters; rep->refinements = refs; rep->relres = rel; rep->converged = rel <= o.rtol;
        rep->fell_back = fell; rep->ms = 0.;
    }
    if (lazyMs && rep == &c.lastPc) {
        c.pcMsPending = true;
    } else {
        HIPCHECK(hipEventSynchronize(c.evP1));
        float ms = 0.f;
        HIPCHECK(hipEventElapsedTime(&ms, c.evP0, c.evP1));
        if (rep) rep->ms = ms;
    }
    return rel <= o.rtol ? CAFFE_OK : CAFFE_ERR_NOT_CONVERGED;
}

// Timing of the dominant kernel: K1 of the pressure-correction CG (stencil SpMV fused
// with the search-direction update and the (p, Ap) partial sum) on the current p'
// system, launched `reps` times between two events on the context's stream.
template <class T>
static void timeK1(Ctx& c, Sym4<T> A, T* w, int reps) {
    const Dims& D = c.D;
    const int P = D.P;
    const int g = gridFor(D);
    double* st = c.scal + SC_CG;
    double* partB = c.red.part;
    double* partA = c.red.part + MAX_RED_BLOCKS * RED_STRIDE;
    LoopCtl lc{0., 1 << 30, 0.};  // never converges: every launch runs
    T* r = w; T* z = r + P; T* p0 = z + P; T* p1 = p0 + P; T* q = p1 + P; T* x = q + P;
    hipLaunchKernelGGL((k_cg_init<T, double>), dim3(g), dim3(KBS), 0, c.stream, D, A, c.bp, 0., 1.0, r, z, p0, p1, x, 1, 1, partB, st, 1.0);
    hipLaunchKernelGGL((k_cg_k1<T>), dim3(g), dim3(KBS), 0, c.stream, D, A, z, p0, p1, q, partB, g, partA, st, lc);
    HIPCHECK(hipEventRecord(c.ev0, c.stream));
    for (int i = 0; i < reps; ++i)
        hipLaunchKernelGGL((k_cg_k1<T>), dim3(g), dim3(KBS), 0, c.stream, D, A, z, p0, p1, q, partB, g, partA, st, lc);
    HIPCHECK(hipEventRecord(c.ev1, c.stream));
}

double time_cg_k1(Ctx& c, bool fp32, int reps) {
    if (fp32) { pcorr_make_fp32(c); timeK1<float>(c, symP<float>(c), c.wp32, reps); }
    else timeK1<double>(c, symP<double>(c), c.wp, reps);
    float ms = 0.f;
    HIPCHECK(hipEventSynchronize(c.ev1));
    HIPCHECK(hipEventElapsedTime(&ms, c.ev0, c.ev1));
    return (double)ms / reps;
}

}  // namespace caffe
