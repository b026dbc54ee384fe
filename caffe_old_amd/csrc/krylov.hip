// krylov.hip -- preconditioned Krylov solvers for the momentum and pressure-
// correction systems (the replacement of SparseMatrix::solve, SparseMatrix.cpp:86-106).
//
// Every solve starts from x = 0, as KSPSolve does with PETSc's default initial guess.
// Convergence is on the true residual norm ||b - A x||_2 <= rtol ||b||_2; the
// recurrence residual is checked every iteration and the true residual at exit
// (a restart from it if they disagree).
//
// Design: all Krylov scalars live in device memory.  Each kernel both applies the
// operator (7-point stencil, operands formed on the fly at the 7 stencil points from
// the previous iterates, so no separate vector-update pass) and produces the partial
// dot products; the last block to finish combines the partials in block order (a
// fixed order) and computes the next scalar (alpha, omega, beta) on the device.  The
// host queues a chunk of iterations and reads one flag per chunk.
//
// Systems:
//   momentum     7-point band (aP aE aW aN aS aT aB), 3 right-hand sides sharing the
//                coefficients, plus 9-entry ghost rows; BiCGStab, Jacobi.
//   pressure     symmetric storage (aP + one coefficient per face), pure Neumann or
//                Dirichlet outlet; PCG with Jacobi in FP64, or FP32 PCG inside FP64
//                iterative refinement.  With ghost rows the matrix is nonsymmetric and
//                BiCGStab is used instead.
#include "internal.hpp"

#include <cmath>

namespace caffe {

constexpr int KBS = 256;
constexpr int KGRID_CAP = 2048;

// ------------------------------------------------------------------ operators
struct OvView {
    const int* rowOf;      // nullptr when there are no ghost rows
    const long* cols;
    const void* vals;      // double or float, 9 per row
};

template <class T>
struct Band7 {             // momentum matrix
    const double* A; long P; long sJ, sK; OvView ov; const int8_t* st;
    template <class F>
    __device__ double apply(F x, long p) const {
        if (ov.rowOf) {
            int r = ov.rowOf[p];
            if (r >= 0) {
                const double* v = (const double*)ov.vals + 9L * r;
                double s = v[0] * x(p);
                for (int m = 1; m < 9; ++m) { long c = ov.cols[9L * r + m]; if (c >= 0) s += v[m] * x(c); }
                return s;
            }
        }
        return A[p] * x(p) + A[P + p] * x(p + 1) + A[2 * P + p] * x(p - 1) + A[3 * P + p] * x(p + sJ)
             + A[4 * P + p] * x(p - sJ) + A[5 * P + p] * x(p + sK) + A[6 * P + p] * x(p - sK);
    }
};

template <class T>
struct Sym4 {              // pressure-correction matrix
    const T* A; long P; long sJ, sK; OvView ov; const int8_t* st;
    template <class F>
    __device__ double apply(F x, long p) const {
        if (ov.rowOf) {
            if (st[p] == ST_INACTIVE) return x(p);
            int r = ov.rowOf[p];
            if (r >= 0) {
                const T* v = (const T*)ov.vals + 9L * r;
                double s = (double)v[0] * x(p);
                for (int m = 1; m < 9; ++m) { long c = ov.cols[9L * r + m]; if (c >= 0) s += (double)v[m] * x(c); }
                return s;
            }
        }
        return (double)A[p] * x(p) + (double)A[P + p] * x(p - 1) + (double)A[P + p + 1] * x(p + 1)
             + (double)A[2 * P + p] * x(p - sJ) + (double)A[2 * P + p + sJ] * x(p + sJ)
             + (double)A[3 * P + p] * x(p - sK) + (double)A[3 * P + p + sK] * x(p + sK);
    }
};

// ------------------------------------------------------------------ slot layouts
enum { B_RHO = 0, B_ALPHA, B_OMEGA, B_BETA, B_BB, B_RR, B_DONE, B_IT };  // BiCGStab
enum { C_RZ = 0, C_ALPHA, C_BETA, C_BB, C_RR, C_DONE, C_IT, C_PQ };     // CG

struct LoopCtl { double tol2; int maxit; };

// grid-stride over interior cells: q = interior id, p = padded index
__device__ inline long cellIdx(const Dims& D, long q) { int i, j, k; return D.cell(q, i, j, k); }
#define FOR_CELLS(D, q, p)                                                                    \
    for (long q = (long)blockIdx.x * blockDim.x + threadIdx.x, p = 0;                        \
         q < (D).N && ((p = cellIdx((D), q)), true); q += (long)gridDim.x * blockDim.x)

// ================================================================== PCG (pressure)
// init: r = b - mean (if singular), z = r / aP, p = 0, bb, rz
template <class T>
__global__ void k_cg_init(Dims D, Sym4<T> A, const T* b, double mean, T* r, T* z, T* p0, T* p1, T* x,
                          double* sc, RedBuf rb) {
    double v[2] = {0., 0.};
    FOR_CELLS(D, q, p) {
        const bool live = !A.st || A.st[p] != ST_INACTIVE;
        double bv = live ? (double)b[p] - mean : 0.;
        T rr = (T)bv;
        double zz = (double)rr / (double)A.A[p];
        r[p] = rr; z[p] = (T)zz; p0[p] = 0; p1[p] = 0; x[p] = 0;
        v[0] += (double)rr * (double)rr;
        v[1] += (double)rr * zz;
    }
    gridReduce<KBS, 2>(v, rb, [&](const double* t) {
        sc[C_BB] = t[0]; sc[C_RR] = t[0]; sc[C_RZ] = t[1]; sc[C_BETA] = 0.; sc[C_IT] = 0.;
        sc[C_DONE] = (t[0] == 0.) ? 1. : 0.;
    });
}

// K1: p_new = z + beta p_old (on the fly at the stencil), q = A p_new, partial (p_new, q)
template <class T>
__global__ void k_cg_k1(Dims D, Sym4<T> A, const T* z, const T* pold, T* pnew, T* qv, double* sc, RedBuf rb) {
    if (sc[C_DONE] != 0.) return;
    const double beta = sc[C_BETA];
    double v[1] = {0.};
    auto pf = [&](long c) { return (double)z[c] + beta * (double)pold[c]; };
    FOR_CELLS(D, q, p) {
        double pn = pf(p);
        double y = A.apply(pf, p);
        pnew[p] = (T)pn; qv[p] = (T)y;
        v[0] += pn * y;
    }
    gridReduce<KBS, 1>(v, rb, [&](const double* t) {
        sc[C_PQ] = t[0];
        sc[C_ALPHA] = t[0] != 0. ? sc[C_RZ] / t[0] : 0.;
    });
}

// K2: x += alpha p; r -= alpha q; z = r / aP; partials (r,z), (r,r); beta, convergence
template <class T>
__global__ void k_cg_k2(Dims D, Sym4<T> A, T* x, T* r, T* z, const T* pv, const T* qv, double* sc, RedBuf rb, LoopCtl lc) {
    if (sc[C_DONE] != 0.) return;
    const double alpha = sc[C_ALPHA];
    double v[2] = {0., 0.};
    FOR_CELLS(D, q, p) {
        double xn = (double)x[p] + alpha * (double)pv[p];
        double rn = (double)r[p] - alpha * (double)qv[p];
        T rt = (T)rn;
        double zn = (double)rt / (double)A.A[p];
        x[p] = (T)xn; r[p] = rt; z[p] = (T)zn;
        v[0] += (double)rt * zn;
        v[1] += (double)rt * (double)rt;
    }
    gridReduce<KBS, 2>(v, rb, [&](const double* t) {
        double rzOld = sc[C_RZ];
        sc[C_RZ] = t[0];
        sc[C_BETA] = rzOld != 0. ? t[0] / rzOld : 0.;
        sc[C_RR] = t[1];
        sc[C_IT] += 1.;
        if (t[1] <= lc.tol2 * sc[C_BB] || sc[C_IT] >= lc.maxit || sc[C_PQ] == 0.) sc[C_DONE] = 1.;
    });
}

// ================================================================== BiCGStab, NV systems
// Vectors of system v are at offset v*P.  Slots of system v at sc + 8v.
template <int NV, class T, class Op>
__global__ void k_bi_init(Dims D, Op A, const T* b, double mean, T* r, T* rh, T* p0, T* p1, T* v0, T* v1, T* x,
                          T* dinv, double* sc, RedBuf rb) {
    double acc[NV];
    for (int s = 0; s < NV; ++s) acc[s] = 0.;
    const long P = D.P;
    FOR_CELLS(D, q, p) {
        const bool live = !A.st || A.st[p] != ST_INACTIVE;
        dinv[p] = (T)(1. / (double)A.A[p]);
        for (int s = 0; s < NV; ++s) {
            const long o = s * P + p;
            double bv = live ? (double)b[o] - mean : 0.;
            r[o] = (T)bv; rh[o] = (T)bv; p0[o] = 0; p1[o] = 0; v0[o] = 0; v1[o] = 0; x[o] = 0;
            acc[s] += bv * bv;
        }
    }
    gridReduce<KBS, NV>(acc, rb, [&](const double* t) {
        for (int s = 0; s < NV; ++s) {
            double* c = sc + 8 * s;
            c[B_BB] = t[s]; c[B_RR] = t[s]; c[B_RHO] = t[s];  // rho = (rh, r) = (r, r)
            c[B_ALPHA] = 1.; c[B_OMEGA] = 1.; c[B_BETA] = 0.; c[B_IT] = 0.;
            c[B_DONE] = (t[s] == 0.) ? 1. : 0.;
        }
    });
}

// KA: p_new = r + beta (p_old - omega v_old); v_new = A D^-1 p_new; partial (rh, v_new)
template <int NV, class T, class Op>
__global__ void k_bi_a(Dims D, Op A, const T* r, const T* rh, const T* pold, T* pnew, const T* vold, T* vnew,
                       const T* dinv, double* sc, RedBuf rb) {
    bool any = false;
    for (int s = 0; s < NV; ++s) any |= sc[8 * s + B_DONE] == 0.;
    if (!any) return;
    double acc[NV];
    double beta[NV], om[NV];
    bool live[NV];
    for (int s = 0; s < NV; ++s) {
        acc[s] = 0.; beta[s] = sc[8 * s + B_BETA]; om[s] = sc[8 * s + B_OMEGA]; live[s] = sc[8 * s + B_DONE] == 0.;
    }
    const long P = D.P;
    FOR_CELLS(D, q, p) {
        for (int s = 0; s < NV; ++s) {
            if (!live[s]) continue;
            const long o = s * P;
            auto pf = [&](long c) { return (double)r[o + c] + beta[s] * ((double)pold[o + c] - om[s] * (double)vold[o + c]); };
            auto ph = [&](long c) { return pf(c) * (double)dinv[c]; };
            double pn = pf(p);
            double vn = A.apply(ph, p);
            pnew[o + p] = (T)pn; vnew[o + p] = (T)vn;
            acc[s] += (double)rh[o + p] * vn;
        }
    }
    gridReduce<KBS, NV>(acc, rb, [&](const double* t) {
        for (int s = 0; s < NV; ++s) {
            double* c = sc + 8 * s;
            if (c[B_DONE] != 0.) continue;
            if (t[s] == 0.) { c[B_DONE] = 2.; c[B_ALPHA] = 0.; }  // breakdown: host restarts
            else c[B_ALPHA] = c[B_RHO] / t[s];
        }
    });
}

// KB: s = r - alpha v (on the fly); t = A D^-1 s; partials (t,s), (t,t)
template <int NV, class T, class Op>
__global__ void k_bi_b(Dims D, Op A, const T* r, const T* v, T* sv, T* tv, const T* dinv, double* sc, RedBuf rb) {
    bool any = false;
    for (int s = 0; s < NV; ++s) any |= sc[8 * s + B_DONE] == 0.;
    if (!any) return;
    double acc[2 * NV];
    double al[NV];
    bool live[NV];
    for (int s = 0; s < NV; ++s) { acc[2 * s] = acc[2 * s + 1] = 0.; al[s] = sc[8 * s + B_ALPHA]; live[s] = sc[8 * s + B_DONE] == 0.; }
    const long P = D.P;
    FOR_CELLS(D, q, p) {
        for (int s = 0; s < NV; ++s) {
            if (!live[s]) continue;
            const long o = s * P;
            auto sf = [&](long c) { return (double)r[o + c] - al[s] * (double)v[o + c]; };
            auto sh = [&](long c) { return sf(c) * (double)dinv[c]; };
            double ss = sf(p);
            double tt = A.apply(sh, p);
            sv[o + p] = (T)ss; tv[o + p] = (T)tt;
            acc[2 * s] += tt * ss;
            acc[2 * s + 1] += tt * tt;
        }
    }
    gridReduce<KBS, 2 * NV>(acc, rb, [&](const double* t) {
        for (int s = 0; s < NV; ++s) {
            double* c = sc + 8 * s;
            if (c[B_DONE] != 0.) continue;
            c[B_OMEGA] = t[2 * s + 1] > 0. ? t[2 * s] / t[2 * s + 1] : 0.;
        }
    });
}

// KC: x += alpha D^-1 p + omega D^-1 s; r = s - omega t; partials (rh, r), (r, r)
template <int NV, class T, class Op>
__global__ void k_bi_c(Dims D, Op A, T* x, T* r, const T* rh, const T* pv, const T* sv, const T* tv, const T* dinv,
                       double* sc, RedBuf rb, LoopCtl lc) {
    bool any = false;
    for (int s = 0; s < NV; ++s) any |= sc[8 * s + B_DONE] == 0.;
    if (!any) return;
    double acc[2 * NV];
    double al[NV], om[NV];
    bool live[NV];
    for (int s = 0; s < NV; ++s) {
        acc[2 * s] = acc[2 * s + 1] = 0.; al[s] = sc[8 * s + B_ALPHA]; om[s] = sc[8 * s + B_OMEGA]; live[s] = sc[8 * s + B_DONE] == 0.;
    }
    const long P = D.P;
    FOR_CELLS(D, q, p) {
        const double di = (double)dinv[p];
        for (int s = 0; s < NV; ++s) {
            if (!live[s]) continue;
            const long o = s * P + p;
            double xn = (double)x[o] + al[s] * ((double)pv[o] * di) + om[s] * ((double)sv[o] * di);
            double rn = (double)sv[o] - om[s] * (double)tv[o];
            x[o] = (T)xn; r[o] = (T)rn;
            acc[2 * s] += (double)rh[o] * rn;
            acc[2 * s + 1] += rn * rn;
        }
    }
    gridReduce<KBS, 2 * NV>(acc, rb, [&](const double* t) {
        for (int s = 0; s < NV; ++s) {
            double* c = sc + 8 * s;
            if (c[B_DONE] != 0.) continue;
            double rhoN = t[2 * s];
            c[B_RR] = t[2 * s + 1];
            c[B_BETA] = (c[B_RHO] != 0. && c[B_OMEGA] != 0.) ? (rhoN / c[B_RHO]) * (c[B_ALPHA] / c[B_OMEGA]) : 0.;
            c[B_RHO] = rhoN;
            c[B_IT] += 1.;
            if (t[2 * s + 1] <= lc.tol2 * c[B_BB] || c[B_IT] >= lc.maxit) c[B_DONE] = 1.;
            else if (c[B_OMEGA] == 0. || rhoN == 0.) c[B_DONE] = 2.;
        }
    });
}

// ================================================================== residual and utilities
// r = b - mean - A x in FP64 over NV systems; partials (r,r) per system, (b,b) per system
template <int NV, class Op>
__global__ void k_resid(Dims D, Op A, const double* b, double mean, const double* x, double* r, double* sc, RedBuf rb) {
    double acc[2 * NV];
    for (int s = 0; s < 2 * NV; ++s) acc[s] = 0.;
    const long P = D.P;
    FOR_CELLS(D, q, p) {
        const bool live = !A.st || A.st[p] != ST_INACTIVE;
        for (int s = 0; s < NV; ++s) {
            const long o = s * P;
            auto xf = [&](long c) { return x[o + c]; };
            double bv = live ? b[o + p] - mean : 0.;
            double rv = live ? bv - A.apply(xf, p) : 0.;
            if (r) r[o + p] = rv;
            acc[2 * s] += rv * rv;
            acc[2 * s + 1] += bv * bv;
        }
    }
    gridReduce<KBS, 2 * NV>(acc, rb, [&](const double* t) {
        for (int s = 0; s < 2 * NV; ++s) sc[s] = t[s];
    });
}

// mean of b over non-inactive cells (Neumann null-space projection)
__global__ void k_mean(Dims D, const int8_t* st, const double* b, double* sc, RedBuf rb) {
    double acc[2] = {0., 0.};
    FOR_CELLS(D, q, p) {
        if (st[p] == ST_INACTIVE) continue;
        acc[0] += b[p];
        acc[1] += 1.;
    }
    gridReduce<KBS, 2>(acc, rb, [&](const double* t) { sc[0] = t[1] > 0. ? t[0] / t[1] : 0.; });
}

// x += s * e (FP64 accumulate of the FP32 correction)
__global__ void k_axpy_mixed(Dims D, double* x, const float* e, double s) {
    FOR_CELLS(D, q, p) { x[p] += s * (double)e[p]; }
}
// r32 = r64 * s
__global__ void k_to_f32(Dims D, const double* r, float* o, double s) {
    FOR_CELLS(D, q, p) { o[p] = (float)(r[p] * s); }
}
__global__ void k_cast_f32(const double* a, float* o, long n) {
    long q = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (q < n) o[q] = (float)a[q];
}
__global__ void k_zero_d(double* a, long n) {
    long q = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (q < n) a[q] = 0.;
}

// ================================================================== host drivers
static int gridFor(const Dims& D) { return blocksFor(D.N, KBS, KGRID_CAP); }

static void readBack(Ctx& c, int slot, int n) {
    HIPCHECK(hipMemcpyAsync(c.hscal + slot, c.scal + slot, n * sizeof(double), hipMemcpyDeviceToHost, c.stream));
    HIPCHECK(hipStreamSynchronize(c.stream));
}

static double computeMean(Ctx& c, const double* b) {
    hipLaunchKernelGGL(k_mean, dim3(gridFor(c.D)), dim3(KBS), 0, c.stream, c.D, c.st, b, c.scal + SC_MISC + 4, c.red);
    readBack(c, SC_MISC + 4, 1);
    return c.hscal[SC_MISC + 4];
}

static OvView ovM(Ctx& c) { return OvView{c.hasIB ? c.ov.rowOf : nullptr, c.ov.cols, c.ov.valM}; }
static OvView ovP(Ctx& c) { return OvView{c.hasIB ? c.ov.rowOf : nullptr, c.ov.cols, c.ov.valP}; }
static OvView ovP32(Ctx& c) { return OvView{c.hasIB ? c.ov.rowOf : nullptr, c.ov.cols, c.ov.valP32}; }

// Generic BiCGStab over NV systems.  Returns iterations (max over systems).
template <int NV, class T, class Op>
static int runBicgstab(Ctx& c, Op A, const T* b, double mean, T* x, T* work, int slot, double rtol, int maxit,
                       int chunk) {
    const Dims& D = c.D;
    const long P = D.P, NVP = (long)NV * P;
    T* r = work; T* rh = r + NVP; T* p0 = rh + NVP; T* p1 = p0 + NVP; T* v0 = p1 + NVP; T* v1 = v0 + NVP;
    T* s = v1 + NVP; T* t = s + NVP; T* dinv = t + NVP;
    double* sc = c.scal + slot;
    const int g = gridFor(D);
    LoopCtl lc{rtol * rtol, maxit};
    hipLaunchKernelGGL((k_bi_init<NV, T, Op>), dim3(g), dim3(KBS), 0, c.stream, D, A, b, mean, r, rh, p0, p1, v0, v1, x, dinv, sc, c.red);
    int it = 0;
    T *pc = p0, *pn = p1, *vc = v0, *vn = v1;
    while (true) {
        for (int q = 0; q < chunk; ++q) {
            hipLaunchKernelGGL((k_bi_a<NV, T, Op>), dim3(g), dim3(KBS), 0, c.stream, D, A, r, rh, pc, pn, vc, vn, dinv, sc, c.red);
            hipLaunchKernelGGL((k_bi_b<NV, T, Op>), dim3(g), dim3(KBS), 0, c.stream, D, A, r, vn, s, t, dinv, sc, c.red);
            hipLaunchKernelGGL((k_bi_c<NV, T, Op>), dim3(g), dim3(KBS), 0, c.stream, D, A, x, r, rh, pn, s, t, dinv, sc, c.red, lc);
            std::swap(pc, pn); std::swap(vc, vn);
        }
        readBack(c, slot, 8 * NV);
        bool all = true;
        it = 0;
        for (int q = 0; q < NV; ++q) {
            all &= c.hscal[slot + 8 * q + B_DONE] != 0.;
            it = std::max(it, (int)c.hscal[slot + 8 * q + B_IT]);
        }
        if (all || it >= maxit) break;
    }
    return it;
}

template <class T>
static int runCg(Ctx& c, Sym4<T> A, const T* b, double mean, T* x, T* work, int slot, double rtol, int maxit, int chunk) {
    const Dims& D = c.D;
    const long P = D.P;
    T* r = work; T* z = r + P; T* p0 = z + P; T* p1 = p0 + P; T* qv = p1 + P;
    double* sc = c.scal + slot;
    const int g = gridFor(D);
    LoopCtl lc{rtol * rtol, maxit};
    hipLaunchKernelGGL((k_cg_init<T>), dim3(g), dim3(KBS), 0, c.stream, D, A, b, mean, r, z, p0, p1, x, sc, c.red);
    T *pc = p0, *pn = p1;
    int it = 0;
    while (true) {
        for (int q = 0; q < chunk; ++q) {
            hipLaunchKernelGGL((k_cg_k1<T>), dim3(g), dim3(KBS), 0, c.stream, D, A, z, pc, pn, qv, sc, c.red);
            hipLaunchKernelGGL((k_cg_k2<T>), dim3(g), dim3(KBS), 0, c.stream, D, A, x, r, z, pn, qv, sc, c.red, lc);
            std::swap(pc, pn);
        }
        readBack(c, slot, 8);
        it = (int)c.hscal[slot + C_IT];
        if (c.hscal[slot + C_DONE] != 0. || it >= maxit) break;
    }
    return it;
}

// true FP64 residual of the momentum systems; returns max relres over components
static double momResid(Ctx& c, double* rbuf) {
    Band7<double> A{c.Am, c.D.P, c.D.sJ, c.D.sK, ovM(c), c.st};
    hipLaunchKernelGGL((k_resid<3, Band7<double>>), dim3(gridFor(c.D)), dim3(KBS), 0, c.stream, c.D, A, c.bm, 0., c.xm, rbuf,
                       c.scal + SC_MISC + 8, c.red);
    readBack(c, SC_MISC + 8, 6);
    double mx = 0.;
    for (int s = 0; s < 3; ++s) {
        double rr = c.hscal[SC_MISC + 8 + 2 * s], bb = c.hscal[SC_MISC + 8 + 2 * s + 1];
        double rel = bb > 0. ? std::sqrt(rr / bb) : 0.;
        mx = std::max(mx, rel);
    }
    return mx;
}

__global__ void k_add(double* x, const double* d, long n) {
    long q = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (q < n) x[q] += d[q];
}
void add_vec(Ctx& c, double* x, const double* d, long n) {
    hipLaunchKernelGGL(k_add, dim3(blocksFor(n, KBS)), dim3(KBS), 0, c.stream, x, d, n);
}

int momentum_solve(Ctx& c, const caffe_solve_opts& o, caffe_solve_report* rep) {
    HIPCHECK(hipEventRecord(c.ev0, c.stream));
    Band7<double> A{c.Am, c.D.P, c.D.sJ, c.D.sK, ovM(c), c.st};
    const long P = c.D.P;
    int iters = runBicgstab<3, double>(c, A, c.bm, 0., c.xm, c.wm, SC_M, o.rtol, o.maxit, 4);
    double rel = momResid(c, c.wm);
    // restart from the true residual if the recurrence drifted (or broke down)
    int restarts = 0;
    while (rel > o.rtol && restarts < 3 && iters < o.maxit) {
        double* dx = c.wm + 10 * 3 * P;
        double* rr = c.wm + 11 * 3 * P;
        HIPCHECK(hipMemcpyAsync(rr, c.wm, 3 * P * sizeof(double), hipMemcpyDeviceToDevice, c.stream));
        const double sub = std::min(0.5, 0.5 * o.rtol / rel);
        iters += runBicgstab<3, double>(c, A, rr, 0., dx, c.wm, SC_M, sub, o.maxit - iters, 4);
        add_vec(c, c.xm, dx, 3 * P);
        rel = momResid(c, c.wm);
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


void pcorr_make_fp32(Ctx& c) {
    long n = 4 * c.D.P;
    hipLaunchKernelGGL(k_cast_f32, dim3(blocksFor(n, KBS)), dim3(KBS), 0, c.stream, c.Ap, c.Ap32, n);
    if (c.hasIB && c.ov.n > 0) {
        long m = 9L * c.ov.n;
        hipLaunchKernelGGL(k_cast_f32, dim3(blocksFor(m, KBS)), dim3(KBS), 0, c.stream, c.ov.valP, c.ov.valP32, m);
    }
}

// true FP64 residual of the pressure system: r = (b - mean) - A x; returns relres, stores rr, bb
static double pcResid(Ctx& c, double mean, double* r, double* rrOut, double* bbOut) {
    Sym4<double> A{c.Ap, c.D.P, c.D.sJ, c.D.sK, ovP(c), c.st};
    hipLaunchKernelGGL((k_resid<1, Sym4<double>>), dim3(gridFor(c.D)), dim3(KBS), 0, c.stream, c.D, A, c.bp, mean, c.xp, r,
                       c.scal + SC_IR, c.red);
    readBack(c, SC_IR, 2);
    double rr = c.hscal[SC_IR], bb = c.hscal[SC_IR + 1];
    if (rrOut) *rrOut = rr;
    if (bbOut) *bbOut = bb;
    return bb > 0. ? std::sqrt(rr / bb) : 0.;
}

int pcorr_solve(Ctx& c, const caffe_solve_opts& o, caffe_solve_report* rep) {
    HIPCHECK(hipEventRecord(c.ev0, c.stream));
    const Dims& D = c.D;
    const long P = D.P;
    const double mean = c.singularP ? computeMean(c, c.bp) : 0.;
    int iters = 0, refs = 0, fell = 0;
    double rel = 1.;
    const bool sym = !c.hasIB;
    double* r64 = c.wp + 5 * P;
    auto plain = [&](double tol, int cap) {
        Sym4<double> A{c.Ap, P, D.sJ, D.sK, ovP(c), c.st};
        if (sym) iters += runCg<double>(c, A, c.bp, mean, c.xp, c.wp, SC_P, tol, cap, 16);
        else iters += runBicgstab<1, double>(c, A, c.bp, mean, c.xp, c.wm, SC_P, tol, cap, 8);
        rel = pcResid(c, mean, r64, nullptr, nullptr);
    };
    if (o.mode == CAFFE_SOLVE_PLAIN_FP64) {
        plain(o.rtol, o.maxit);
        // restart from the true residual when the recurrence drifted
        int restarts = 0;
        while (rel > o.rtol && restarts < 3 && iters < o.maxit) {
            double* dx = c.wp + 6 * P;
            Sym4<double> A{c.Ap, P, D.sJ, D.sK, ovP(c), c.st};
            double sub = std::min(0.5, 0.5 * o.rtol / rel);
            if (sym) iters += runCg<double>(c, A, r64, 0., dx, c.wp, SC_P, sub, o.maxit - iters, 16);
            else iters += runBicgstab<1, double>(c, A, r64, 0., dx, c.wm, SC_P, sub, o.maxit - iters, 8);
            add_vec(c, c.xp, dx, P);
            rel = pcResid(c, mean, r64, nullptr, nullptr);
            ++restarts;
        }
        refs = restarts;
    } else {
        // FP32 inner solves inside FP64 iterative refinement.  x starts at 0 (or at the
        // learned guess); each step solves A e = r / ||r|| in FP32 and adds ||r|| e.
        hipLaunchKernelGGL(k_zero_d, dim3(blocksFor(P, KBS)), dim3(KBS), 0, c.stream, c.xp, P);
        if (o.mode == CAFFE_SOLVE_LEARNED_FCG && c.net) {
            double rr0, bb0;
            pcResid(c, mean, r64, &rr0, &bb0);
            if (learned_apply_guess(c, r64, c.xp, 1.0) != CAFFE_OK) fell = 1;
        }
        pcorr_make_fp32(c);
        Sym4<float> A32{c.Ap32, P, D.sJ, D.sK, ovP32(c), c.st};
        float* e32 = c.wp32 + 5 * P;
        const double inner = o.inner_rtol > 0. ? o.inner_rtol : 1e-4;
        double rr, bb;
        rel = pcResid(c, mean, r64, &rr, &bb);
        while (rel > o.rtol && iters < o.maxit && refs < 50) {
            const double nr = std::sqrt(rr);
            hipLaunchKernelGGL(k_to_f32, dim3(gridFor(D)), dim3(KBS), 0, c.stream, D, r64, c.wp32 + 6 * P, 1.0 / nr);
            const float* rhs = c.wp32 + 6 * P;
            if (sym) iters += runCg<float>(c, A32, rhs, 0., e32, c.wp32, SC_P, inner, o.maxit - iters, 16);
            else iters += runBicgstab<1, float>(c, A32, rhs, 0., e32, (float*)c.wm, SC_P, inner, o.maxit - iters, 8);
            hipLaunchKernelGGL(k_axpy_mixed, dim3(gridFor(D)), dim3(KBS), 0, c.stream, D, c.xp, e32, nr);
            ++refs;
            double relNew = pcResid(c, mean, r64, &rr, &bb);
            if (!(relNew < rel)) {  // stagnation: fall back to the plain FP64 solve
                rel = relNew;
                break;
            }
            rel = relNew;
        }
        if (rel > o.rtol) {
            fell = 1;
            plain(o.rtol, o.maxit);
        }
    }
    HIPCHECK(hipEventRecord(c.ev1, c.stream));
    HIPCHECK(hipEventSynchronize(c.ev1));
    float ms = 0.f;
    HIPCHECK(hipEventElapsedTime(&ms, c.ev0, c.ev1));
    if (rep) {
        rep->iters = iters; rep->refinements = refs; rep->relres = rel; rep->converged = rel <= o.rtol;
        rep->fell_back = fell; rep->ms = ms;
    }
    return rel <= o.rtol ? CAFFE_OK : CAFFE_ERR_NOT_CONVERGED;
}

}  // namespace caffe
