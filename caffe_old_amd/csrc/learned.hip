// learned.hip -- the learned two-grid preconditioner of the pressure-correction solve,
// its training loss, the device-resident training set, and flexible CG (FCG(1),
// FP32 iterations inside FP64 iterative refinement) that uses it.  No counterpart in the
// reference (SparseMatrix::solve, SparseMatrix.cpp:86-106, is PETSc BiCGStab + SOR);
// the plain FP64 solve of krylov.hip stays the fallback and the judge.
//
// The cycle for a residual r on the fine grid (w = damped-Jacobi weight):
//   z1 = w D^-1 r
//   v  = P^T D^-1 (r - A z1)                 restriction (adjoint of P) to the coarse grid
//   c  = net(v)                              the multilevel convolution network
//   y  = z1 + P c                            P = cell-centred trilinear prolongation
//   z  = y + w D^-1 (r - A y)
// The network is linear, so z = M r with a fixed linear M.  Flexible CG applies the cycle
// with the last line repeated (3 sweeps by default, caffe_ctx_set_fcg_post_sweeps): the
// high-frequency error the coarse correction leaves is damped there rather than by Krylov
// iterations (the second PISO corrector's systems: 24 -> 18 iterations at 128^3).
// Training minimises the relative energy error of one cycle (one post-smoothing sweep)
// applied to a right-hand side b whose converged solution x is known:
// L = (x - M b)^T A (x - M b) / x^T A x.
#include <algorithm>
#include <cmath>
#include <cstring>

#include <hip/hip_ext.h>

#include "net.hpp"

namespace caffe {

// One 1024-thread block per CU: a kernel that first adds the previous kernel's partial
// rows pays that latency once per CU instead of once per round of smaller blocks, and
// the rows to add are few (measured on the FCG stencil kernel: 21.1 -> 16.8 us at 128^3).
constexpr int FB = 1024;    // threads per block of the fine-grid kernels
constexpr int FGRID = 256;  // blocks (partial-sum rows), at most MAX_RED_BLOCKS

struct SymF {  // symmetric p' matrix: aP, and the coefficient of each cell's low face
    const float* A;
    int P, sJ, sK;
    __device__ float aP(int p) const { return A[p]; }
    __device__ float inv(int p) const { const float a = A[p]; return a != 0.f ? 1.f / a : 0.f; }
    template <class F>
    __device__ float offdiag(F x, int p) const {
        return A[P + p] * x(p - 1) + A[P + p + 1] * x(p + 1) + A[2 * P + p] * x(p - sJ) + A[2 * P + p + sJ] * x(p + sJ)
             + A[3 * P + p] * x(p - sK) + A[3 * P + p + sK] * x(p + sK);
    }
    template <class F>
    __device__ float apply(F x, int p) const { return A[p] * x(p) + offdiag(x, p); }
};
static SymF symF(const float* A, const Dims& D) { return SymF{A, D.P, D.sJ, D.sK}; }

// interior cells in one contiguous chunk per block (XCD-aware chunk order)
#define FOR_FINE(D, q, p)                                                                  \
    for (int q##c_ = ((D).N + (int)gridDim.x - 1) / (int)gridDim.x, q = xcdChunk() * q##c_ + (int)threadIdx.x, \
             q##e_ = min((D).N, q - (int)threadIdx.x + q##c_), p = 0;                      \
         q < q##e_ && ((p = fineIdx((D), q)), true); q += blockDim.x)
__device__ inline int fineIdx(const Dims& D, int q) { int i, j, k; return D.cell(q, i, j, k); }

// column tiles (ColMap, common.hpp): one thread per column of a tile
static_assert(CTX * CTY == FB, "one thread per column of a tile");
#define FOR_COLS(D, M, p)                                                                           \
    for (int v_ = xcdChunk(), vn_ = (M).ntx * (M).nty * (M).nseg; v_ < vn_; v_ += (int)gridDim.x)   \
        for (int t_ = v_ / (M).nseg, i_ = (t_ % (M).ntx) * CTX + (int)(threadIdx.x % CTX),          \
                 j_ = (t_ / (M).ntx) * CTY + (int)(threadIdx.x / CTX), k_ = (v_ % (M).nseg) * (M).kper, \
                 ke_ = min((D).nK, k_ + (M).kper), p = 0;                                          \
             i_ < (D).nI && j_ < (D).nJ && k_ < ke_ && ((p = (D).pidx(i_, j_, k_)), true); ++k_)
__device__ __forceinline__ bool fdone(const double* done) { return done && *done != 0.; }
static int fgrid(const Dims& D) { return blocksFor(D.N, FB, FGRID); }


// 1-D cell-centred linear prolongation: fine f takes 3/4 of coarse f>>1 and 1/4 of its
// other neighbour, clamped to [0, nc)
__device__ __forceinline__ void pw(int f, int nc, int& c0, int& c1) {
    c0 = f >> 1;
    c1 = (f & 1) ? c0 + 1 : c0 - 1;
    c1 = c1 < 0 ? 0 : (c1 >= nc ? nc - 1 : c1);
}

// ---------------------------------------------------------------- cycle kernels
// w = D^-1 (r - A (w D^-1 r))
__global__ void k_tg_pre(Dims D, SymF S, const float* r, float* w, const double* done) {
    if (fdone(done)) return;
    auto z1 = [&](int c) { return NET_OMEGA * S.inv(c) * r[c]; };
    FOR_FINE(D, q, p) { w[p] = S.inv(p) * (r[p] - S.apply(z1, p)); }
}

// v = P^T w on the coarse grid (nI/2 x nJ/2 x nK/2), fine w padded.  Coarse cell C
// gathers fine cells 2C-1 .. 2C+2 in each direction; the weight of fine f is 3/4 when
// f>>1 == C plus 1/4 when its other neighbour is C.  The halo of w is zero, so fine
// cells just outside the grid (weight 0) are read through the padded layout unmasked.
// SC: the fine values are w scaled cell by cell by sc (the classical coarse right-hand side
// P^T (D w) of the ablation, sc = aP)
// the 1-D weights of fine cells 2C-1 .. 2C+2 in coarse cell C of P^T (nc coarse, nf fine)
__device__ __forceinline__ void restrictWeights(int C, int nc, int nf, float (&wt)[4]) {
#pragma unroll
    for (int m = 0; m < 4; ++m) {
        const int ff = 2 * C - 1 + m;
        float s = 0.f;
        if (ff >= 0 && ff < nf) {
            int c0, c1;
            pw(ff, nc, c0, c1);
            if (c0 == C) s += 0.75f;
            if (c1 == C) s += 0.25f;
        }
        wt[m] = s;
    }
}
template <bool SC = false>
__device__ __forceinline__ void restrictAt(const Dims& D, const float* w, float* v, int e, const float* sc = nullptr) {
    const int cx = D.nI >> 1, cy = D.nJ >> 1, cz = D.nK >> 1;
    if (e >= cx * cy * cz) return;
    const int I = e % cx, J = (e / cx) % cy, K = e / (cx * cy);
    float wx[4], wy[4], wz[4];
    restrictWeights(I, cx, D.nI, wx);
    restrictWeights(J, cy, D.nJ, wy);
    restrictWeights(K, cz, D.nK, wz);
    // padded index of fine cell (2I-1, 2J-1, 2K-1): lies in the halo at the low faces
    const int p0 = (2 * I) + (2 * J) * D.sJ + (2 * K) * D.sK;
    float s = 0.f;
#pragma unroll
    for (int c = 0; c < 4; ++c) {
        float sy = 0.f;
#pragma unroll
        for (int b = 0; b < 4; ++b) {
            const int o = p0 + c * D.sK + b * D.sJ;
            const float* row = w + o;
            if (SC) {
                const float* sr = sc + o;
                sy += wy[b] * (wx[0] * (row[0] * sr[0]) + wx[1] * (row[1] * sr[1]) + wx[2] * (row[2] * sr[2])
                               + wx[3] * (row[3] * sr[3]));
            } else {
                sy += wy[b] * (wx[0] * row[0] + wx[1] * row[1] + wx[2] * row[2] + wx[3] * row[3]);
            }
        }
        s += wz[c] * sy;
    }
    v[e] = s;
}
__global__ void k_tg_restrict(Dims D, const float* w, float* v, const double* done) {
    if (fdone(done)) return;
    restrictAt(D, w, v, blockIdx.x * blockDim.x + threadIdx.x);
}

// (P c) at fine cell (i,j,k)
__device__ __forceinline__ float prolongAt(const float* c, int cx, int cy, int cz, int i, int j, int k) {
    int x0, x1, y0, y1, z0, z1;
    pw(i, cx, x0, x1);
    pw(j, cy, y0, y1);
    pw(k, cz, z0, z1);
    auto C = [&](int a, int b, int d) { return c[(d * cy + b) * cx + a]; };
    const float ax = 0.75f, bx = 0.25f;
    const float lo = ax * (ax * C(x0, y0, z0) + bx * C(x1, y0, z0)) + bx * (ax * C(x0, y1, z0) + bx * C(x1, y1, z0));
    const float hi = ax * (ax * C(x0, y0, z1) + bx * C(x1, y0, z1)) + bx * (ax * C(x0, y1, z1) + bx * C(x1, y1, z1));
    return ax * lo + bx * hi;
}

// y = w D^-1 r + P c
__global__ void k_tg_post1(Dims D, SymF S, const float* r, const float* c, float* y, const double* done) {
    if (fdone(done)) return;
    const int cx = D.nI >> 1, cy = D.nJ >> 1, cz = D.nK >> 1;
    for (int q = blockIdx.x * blockDim.x + threadIdx.x; q < D.N; q += gridDim.x * blockDim.x) {
        int i, j, k;
        const int p = D.cell(q, i, j, k);
        y[p] = NET_OMEGA * S.inv(p) * r[p] + prolongAt(c, cx, cy, cz, i, j, k);
    }
}

// z = y + w D^-1 (r - A y); with xs: write e = xs - z instead
__global__ void k_tg_post2(Dims D, SymF S, const float* r, const float* y, float* z, const float* xs) {
    auto yf = [&](int c) { return y[c]; };
    FOR_FINE(D, q, p) {
        const float zz = y[p] + NET_OMEGA * S.inv(p) * (r[p] - S.apply(yf, p));
        z[p] = xs ? xs[p] - zz : zz;
    }
}

// one transposed post-smoothing sweep of the gradient: h = g - w A (D^-1 g)
__global__ void k_l_sweepT(Dims D, SymF S, const float* g, float* h) {
    auto gi = [&](int c) { return S.inv(c) * g[c]; };
    FOR_FINE(D, q, p) { h[p] = g[p] - NET_OMEGA * S.apply(gi, p); }
}

// ---------------------------------------------------------------- loss and its gradient
// Ae = A e; partial (e, Ae)
__global__ void k_l_spmv(Dims D, SymF S, const float* e, float* Ae, double* part) {
    double v[1] = {0.};
    auto ef = [&](int c) { return e[c]; };
    FOR_FINE(D, q, p) {
        const float a = S.apply(ef, p);
        Ae[p] = a;
        v[0] += (double)e[p] * (double)a;
    }
    writePartials<FB, 1>(v, part);
}
__global__ void k_l_loss(const double* part, int nb, const double* nrm, float* loss) {
    double t[1];
    readPartials<FB, 1>(part, nb, t);
    if (threadIdx.x == 0) *loss += (float)(t[0] / *nrm);
}
__global__ void k_l_final(const double* part, int nb, double* out) {
    double t[1];
    readPartials<FB, 1>(part, nb, t);
    if (threadIdx.x == 0) *out = t[0];
}
// g = s Ae (s = -2 / x^T A x);  h = g - w A (D^-1 g)   (dL/dy = (I - w A D^-1) dL/dz)
__global__ void k_l_h(Dims D, SymF S, const float* Ae, const double* nrm, float* h) {
    const float s = (float)(-2.0 / *nrm);
    auto gi = [&](int c) { return S.inv(c) * s * Ae[c]; };
    FOR_FINE(D, q, p) { h[p] = s * Ae[p] - NET_OMEGA * S.apply(gi, p); }
}

// Synthetic exact pair: x = sum of Neumann cosine modes (eigenvectors of the constant-
// coefficient Laplacian on the box), b = A x computed by the caller.
constexpr int N_MODES = 8;
struct Modes {
    int k[N_MODES][3];
    float a[N_MODES];
};
__global__ void k_modes(Dims D, Modes m, float* x) {
    const float pi = 3.14159265358979f;
    for (int q = blockIdx.x * blockDim.x + threadIdx.x; q < D.N; q += gridDim.x * blockDim.x) {
        int i, j, k;
        const int p = D.cell(q, i, j, k);
        float s = 0.f;
        for (int t = 0; t < N_MODES; ++t)
            s += m.a[t] * cosf(pi * m.k[t][0] * (i + 0.5f) / D.nI) * cosf(pi * m.k[t][1] * (j + 0.5f) / D.nJ)
                 * cosf(pi * m.k[t][2] * (k + 0.5f) / D.nK);
        x[p] = s;
    }
}
__global__ void k_spmv_f(Dims D, SymF S, const float* x, float* y) {
    auto xf = [&](int c) { return x[c]; };
    FOR_FINE(D, q, p) { y[p] = S.apply(xf, p); }
}

// ---------------------------------------------------------------- host: cycle and loss
enum { FV_R = 0, FV_E, FV_W, FV_Y, FV_Z, FV_P0, FV_P1, FV_Q0, FV_Q1, FV_H, FV_AE, FV_XS, FV_BS, FV_R1, FV_DI, N_FV };

void tg_ensure_fine(Net& n, const Dims& D) {
    if (n.fineP == D.P && n.fw) return;
    HIPCHECK(hipDeviceSynchronize());
    for (void* p : n.fineAllocs) (void)hipFree(p);
    n.fineAllocs.clear();
    n.fw = (float*)net_alloc(n, (size_t)N_FV * D.P * sizeof(float), n.fineAllocs);
    n.fineP = D.P;
}

static bool netDimsOk(const Dims& D) { return D.nI % 16 == 0 && D.nJ % 16 == 0 && D.nK % 16 == 0 && D.nI >= 16; }

static void restrictLaunch(hipStream_t s, const Dims& D, const float* w, float* v, const double* done) {
    const int nc = (D.nI / 2) * (D.nJ / 2) * (D.nK / 2);
    hipLaunchKernelGGL(k_tg_restrict, dim3((nc + 255) / 256), dim3(256), 0, s, D, w, v, done);
}

// the n.post post-smoothing sweeps from y: the last one writes z (or e = xs - z with xs);
// the earlier ones alternate through the FV_Z and FV_P0 vectors (free outside an FCG solve)
static void postSweeps(Net& n, hipStream_t s, const Dims& D, const SymF& S, const float* r, const float* y, float* z,
                       const float* xs) {
    const int g = fgrid(D);
    float* t[2] = {n.fw + (size_t)FV_Z * D.P, n.fw + (size_t)FV_P0 * D.P};
    const float* src = y;
    for (int q = 0; q < n.post - 1; ++q) {
        hipLaunchKernelGGL(k_tg_post2, dim3(g), dim3(FB), 0, s, D, S, r, src, t[q & 1], (const float*)nullptr);
        src = t[q & 1];
    }
    hipLaunchKernelGGL(k_tg_post2, dim3(g), dim3(FB), 0, s, D, S, r, src, z, xs);
}

void tg_apply(Net& n, hipStream_t s, const Dims& D, const float* A, const float* r, float* z, const double* done) {
    float* w = n.fw + (size_t)FV_W * D.P;
    float* y = n.fw + (size_t)FV_Y * D.P;
    const SymF S = symF(A, D);
    const int g = fgrid(D);
    hipLaunchKernelGGL(k_tg_pre, dim3(g), dim3(FB), 0, s, D, S, r, w, done);
    restrictLaunch(s, D, w, n.vin, done);
    net_forward(n, s, n.vin, n.vout, done);
    hipLaunchKernelGGL(k_tg_post1, dim3(g), dim3(FB), 0, s, D, S, r, n.vout, y, done);
    postSweeps(n, s, D, S, r, y, z, nullptr);
}

// loss of the pair (A, b, x): x^T A x first (device scalar), then the cycle on b, the
// error e = x - cycle(b), e^T A e; with the gradient, dL/dc = P^T (I - w A D^-1)^m (-2 A e / x^T A x)
// for m = n.post post-smoothing sweeps
static void lossCore(Net& n, hipStream_t s, const Dims& D, const float* A, const float* b, const float* x, bool withGrad,
                     float* lossDev) {
    float* w = n.fw + (size_t)FV_W * D.P;
    float* y = n.fw + (size_t)FV_Y * D.P;
    float* e = n.fw + (size_t)FV_E * D.P;
    float* Ae = n.fw + (size_t)FV_AE * D.P;
    float* h = n.fw + (size_t)FV_H * D.P;
    double* nrm = n.scal + 16;
    const SymF S = symF(A, D);
    const int g = fgrid(D);
    hipLaunchKernelGGL(k_l_spmv, dim3(g), dim3(FB), 0, s, D, S, x, Ae, n.part);
    hipLaunchKernelGGL(k_l_final, dim3(1), dim3(FB), 0, s, n.part, g, nrm);
    hipLaunchKernelGGL(k_tg_pre, dim3(g), dim3(FB), 0, s, D, S, b, w, (const double*)nullptr);
    restrictLaunch(s, D, w, n.vin, nullptr);
    net_forward(n, s, n.vin, n.vout, nullptr);
    hipLaunchKernelGGL(k_tg_post1, dim3(g), dim3(FB), 0, s, D, S, b, n.vout, y, (const double*)nullptr);
    postSweeps(n, s, D, S, b, y, e, x);
    hipLaunchKernelGGL(k_l_spmv, dim3(g), dim3(FB), 0, s, D, S, e, Ae, n.part);
    hipLaunchKernelGGL(k_l_loss, dim3(1), dim3(FB), 0, s, n.part, g, (const double*)nrm, lossDev);
    if (!withGrad) return;
    hipLaunchKernelGGL(k_l_h, dim3(g), dim3(FB), 0, s, D, S, Ae, (const double*)nrm, h);
    float* t[2] = {n.fw + (size_t)FV_Q0 * D.P, h};   // the further transposed sweeps: h -> Q0 -> h ...
    float* hs = h;
    for (int q = 0; q < n.post - 1; ++q) {
        hipLaunchKernelGGL(k_l_sweepT, dim3(g), dim3(FB), 0, s, D, S, (const float*)hs, t[q & 1]);
        hs = t[q & 1];
    }
    restrictLaunch(s, D, hs, n.gout, nullptr);
    net_backward(n, s, n.vin, n.gout);
}

void tg_loss(Net& n, hipStream_t s, const Sample& smp, bool withGrad, float* lossDev) {
    lossCore(n, s, smp.D, smp.A, smp.b, smp.x, withGrad, lossDev);
}

// loss of a synthetic exact pair on the matrix of `smp`: x from `m`, b = A x
void tg_loss_synthetic(Net& n, hipStream_t s, const Sample& smp, const Modes& m, bool withGrad, float* lossDev) {
    const Dims& D = smp.D;
    float* x = n.fw + (size_t)FV_XS * D.P;
    float* b = n.fw + (size_t)FV_BS * D.P;
    const int g = fgrid(D);
    hipLaunchKernelGGL(k_modes, dim3(g), dim3(FB), 0, s, D, m, x);
    hipLaunchKernelGGL(k_spmv_f, dim3(g), dim3(FB), 0, s, D, symF(smp.A, D), (const float*)x, b);
    lossCore(n, s, D, smp.A, b, x, withGrad, lossDev);
}

// x^T A x of a sample
static double sampleNorm(Net& n, hipStream_t s, const Sample& smp) {
    float* Ae = n.fw + (size_t)FV_AE * smp.D.P;
    const int g = fgrid(smp.D);
    hipLaunchKernelGGL(k_l_spmv, dim3(g), dim3(FB), 0, s, smp.D, symF(smp.A, smp.D), smp.x, Ae, n.part);
    hipLaunchKernelGGL(k_l_final, dim3(1), dim3(FB), 0, s, n.part, g, n.scal);
    HIPCHECK(hipMemcpyAsync(n.hscal, n.scal, sizeof(double), hipMemcpyDeviceToHost, s));
    HIPCHECK(hipStreamSynchronize(s));
    return n.hscal[0];
}

// ---------------------------------------------------------------- flexible CG
// state slots; the round's loop limits F_TOL2 and F_MAXIT live on the device, so the
// iteration's launches are the same in every round
enum { F_RR = 0, F_BB, F_IT, F_DONE, F_PQ, F_SQ, F_PAR, F_FIRSTRR, F_SR, F_PR, F_TOL2, F_MAXIT };
struct FcgCtl { double nLive; };   // live cells of a singular system (0: nonsingular)

// r = s r64 (live cells), e = 0; partials (r,r), sum r; the round's loop limits
__global__ void k_f_init(Dims D, const double* r64, double s, float* r, float* e, double* part, double* st, double tol2,
                         int maxit) {
    double v[2] = {0., 0.};
    FOR_FINE(D, q, p) {
        const float rr = (float)(s * r64[p]);
        r[p] = rr;
        e[p] = 0.f;
        v[0] += (double)rr * (double)rr;
        v[1] += (double)rr;
    }
    writePartials<FB, 2>(v, part);
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        st[F_IT] = 0.; st[F_DONE] = 0.; st[F_BB] = -1.; st[F_TOL2] = tol2; st[F_MAXIT] = (double)maxit;
    }
}

// Refinement rounds queued without a host read-back.  Each round's record at
// c.scal + SC_FCGR: the relative residual it started from, whether it was skipped, its
// scale (1 / |r|) and iteration count, and whether its queued iterations ran out first.
enum { FR_REL = 0, FR_STOP, FR_SCALE, FR_ITS, FR_CAPPED, FR_STRIDE };
constexpr int FCG_MAX_PLAN = 8;
struct RoundCtl {
    double rtol, delta, nCells;   // nCells > 0: the residual norm drops the mean (pure Neumann)
    int round, maxit;
};

// The start of round rc.round from the true residual (r,r), (b,b), sum r that the previous
// kernels left in fin, with the decisions of the host loop made on the device: the round is
// skipped when the residual meets rtol, the iteration budget is spent, or the last round
// gained less than 2x or made no iteration.  Otherwise r = r64 / |r| (live cells), e = 0,
// partials (r,r), sum r, and the round's loop limits (its tolerance gains `delta`, or only
// what rtol still needs).  Every block decides from values earlier kernels wrote; block 0
// writes this round's record and the FCG state, which no block of this kernel reads.
__global__ void k_f_init_round(Dims D, const double* r64, const double* fin, double* rec, float* r, float* e,
                               double* part, double* st, RoundCtl rc) {
    const double r2 = fin[0], b2 = fin[1], rs = fin[2];
    double itsSoFar = 0., prevRel = 0., prevStop = 0., prevIts = 1.;
    for (int q = 0; q < rc.round; ++q) itsSoFar += rec[q * FR_STRIDE + FR_ITS];
    if (rc.round > 0) {
        const double* pr = rec + (rc.round - 1) * FR_STRIDE;
        prevRel = pr[FR_REL]; prevStop = pr[FR_STOP]; prevIts = pr[FR_ITS];
    }
    const double rr = sqrt(rc.nCells > 0. ? fmax(r2 - rs * rs / rc.nCells, 0.) : r2), bb = sqrt(b2);
    const double rel = bb > 0. ? rr / bb : 0.;
    const bool stop = prevStop != 0. || !(rel > rc.rtol) || itsSoFar >= (double)rc.maxit ||
                      (rc.round > 0 && (!(rel < 0.5 * prevRel) || prevIts == 0.));
    const double need = fmax(rc.delta, 0.5 * rc.rtol / rel);
    const double scale = 1. / rr;
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        double* my = rec + rc.round * FR_STRIDE;
        my[FR_REL] = rel; my[FR_STOP] = stop ? 1. : 0.; my[FR_SCALE] = scale; my[FR_ITS] = 0.; my[FR_CAPPED] = 0.;
        st[F_IT] = 0.; st[F_DONE] = stop ? 1. : 0.; st[F_BB] = -1.; st[F_TOL2] = need * need;
        st[F_MAXIT] = fmin(300., (double)rc.maxit - itsSoFar);
    }
    if (stop) return;
    double v[2] = {0., 0.};
    FOR_FINE(D, q, p) {
        const float rv = (float)(scale * r64[p]);
        r[p] = rv;
        e[p] = 0.f;
        v[0] += (double)rv * (double)rv;
        v[1] += (double)rv;
    }
    writePartials<FB, 2>(v, part);
}

// x += e / scale at the end of a queued round (nothing for a skipped one), and e = 0, so
// that the round can go on from its FCG state if its queued iterations ran out; block 0
// records the round's iteration count and whether they ran out before its test
__global__ void k_f_axpy_round(Dims D, double* x, float* e, double* rec, const double* st, int round) {
    const double* my = rec + round * FR_STRIDE;
    if (my[FR_STOP] != 0.) return;
    const double sc = 1. / my[FR_SCALE];
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        rec[round * FR_STRIDE + FR_ITS] = st[F_IT];
        rec[round * FR_STRIDE + FR_CAPPED] = st[F_DONE] == 0. ? 1. : 0.;
    }
    FOR_FINE(D, q, p) {
        x[p] += sc * (double)e[p];
        e[p] = 0.f;
    }
}

// The residual alternates between two buffers by iteration parity (r[F_IT & 1] is the
// current one), so that k_f_k2pre can write the next residual while its neighbours
// still read the current one.

// w = D^-1 (r - A w D^-1 r) for the first iteration of a round (r = r0)
__global__ void k_f_pre0(Dims D, SymF S, const float* dinv, const float* r, float* w, const double* st) {
    if (st[F_DONE] != 0.) return;
    auto z1 = [&](int c) { return NET_OMEGA * dinv[c] * r[c]; };
    FOR_FINE(D, q, p) { w[p] = dinv[p] * (r[p] - S.apply(z1, p)); }
}

// Convergence check on the partials of (r,r), sum r (from k_f_init or k_f_k2pre), then
// v = P^T w.  Block 0 raises F_DONE in this kernel, so no block returns on it before the
// block-wide reduction: a block whose waves read the flag on either side of that write
// would otherwise split at the barrier.  When the flag is read as set, the partials
// already meet the criterion, so every wave reaches the same decision.
__global__ void k_f_restrict(Dims D, const float* w, float* v, const double* partIn, int nb, double* st, FcgCtl lc) {
    const double sBb = st[F_BB], sDone = st[F_DONE], sIt = st[F_IT], tol2 = st[F_TOL2], maxit = st[F_MAXIT];
    double t[2];
    readPartials<256, 2>(partIn, nb, t);
    const double rr = lc.nLive > 0. ? fmax(t[0] - t[1] * t[1] / lc.nLive, 0.) : t[0];
    const double bb = sBb < 0. ? rr : sBb;
    const bool done = sDone != 0. || rr <= tol2 * bb || sIt >= maxit || rr == 0.;
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        st[F_BB] = bb; st[F_RR] = rr; st[F_SR] = t[1];
        if (sIt == 1.) st[F_FIRSTRR] = rr;
        if (done) st[F_DONE] = 1.;
    }
    if (done) return;
    restrictAt(D, w, v, blockIdx.x * blockDim.x + threadIdx.x);
}

// the ablation's classical coarse right-hand side v = P^T (D w) = P^T (r - A z1)
__global__ void k_f_restrict_dw(Dims D, const float* w, const float* aP, float* v, const double* st) {
    if (st[F_DONE] != 0.) return;
    restrictAt<true>(D, w, v, blockIdx.x * blockDim.x + threadIdx.x, aP);
}

// The post-smoothing of the cycle in one pass, without y in memory.  With s = P c and
// w = D^-1 (r - A z1) from the pre-smoothing (k_f_k2pre / k_f_pre0),
//   z = y + w D^-1 (r - A y) = z1 + s + w (w_pre - D^-1 A s),   z1 = w D^-1 r,
// since r - A y = (r - A z1) - A s.  s is formed at the seven stencil points from the
// coarse correction c (a 1/64-size array that stays in cache): each thread walks its
// column up k (column tiles as in k_f_k1) keeping, for its own column and the four side
// columns, the bilinear (x, y) interpolant of c on three coarse planes (k >> 1 - 1 .. + 1);
// a step to an even k shifts them and loads one coarse plane (a 3x3 patch).  Each s value
// has prolongAt's products in prolongAt's order, so s is bit-identical to P c.  Cells
// outside the grid take s = 0 (the halo of y).  Partials (z, q_old), sum z, (z, r).
template <bool PART>   // PART: this pass ends the post-smoothing and writes the partial sums
__global__ void k_f_post(Dims D, SymF S, ColMap cm, const float* __restrict__ dinv, const float* r0, const float* r1,
                         const float* __restrict__ wpre, const float* __restrict__ c, float* z, const float* q0,
                         const float* q1, double* part, const double* st) {
    const double sDone = st[F_DONE], sIt = st[F_IT];
    if (sDone != 0.) return;
    const int it = (int)sIt;
    const float* __restrict__ qold = (it & 1) ? q0 : q1;
    const float* __restrict__ r = (it & 1) ? r1 : r0;
    const int P = S.P, sK = S.sK;
    const float* __restrict__ A = S.A;
    const int cx = D.nI >> 1, cy = D.nJ >> 1, cz = D.nK >> 1;
    const float ax = 0.75f, bx = 0.25f;
    struct Plane { float r, d, w, q, a[7]; };
    auto loadPlane = [&](int p, Plane& L, bool a5) {   // a5 false: a[5] is the plane below's a[6]
        L.r = r[p]; L.d = dinv[p]; L.w = wpre[p]; L.q = PART ? qold[p] : 0.f;
        L.a[0] = A[p]; L.a[1] = A[P + p]; L.a[2] = A[P + p + 1]; L.a[3] = A[2 * P + p];
        L.a[4] = A[2 * P + p + S.sJ]; if (a5) L.a[5] = A[3 * P + p]; L.a[6] = A[3 * P + p + sK];
    };
    double v[3] = {0., 0., 0.};
    for (int v_ = xcdChunk(), vn = cm.ntx * cm.nty * cm.nseg; v_ < vn; v_ += (int)gridDim.x) {
        const int t_ = v_ / cm.nseg;
        const int i = (t_ % cm.ntx) * CTX + (int)(threadIdx.x % CTX), j = (t_ / cm.ntx) * CTY + (int)(threadIdx.x / CTX);
        int k = (v_ % cm.nseg) * cm.kper;
        const int ke = min(D.nK, k + cm.kper);
        if (!(i < D.nI && j < D.nJ) || k >= ke) continue;
        int p = D.pidx(i, j, k);
        Plane cur, nxt;
        loadPlane(p, cur, true);
        // the 3x3 coarse patch around (i >> 1, j >> 1), indices clamped as pw clamps them
        const int X = i >> 1, Y = j >> 1;
        const int xa = max(X - 1, 0), xb = X, xc = min(X + 1, cx - 1);
        const int ya = max(Y - 1, 0), yb = Y, yc = min(Y + 1, cy - 1);
        const bool pi = i & 1, pj = j & 1;
        const bool inW = i > 0, inE = i + 1 < D.nI, inS = j > 0, inN = j + 1 < D.nJ;
        // B[col][slot]: bilinear value of columns C, W, E, S, N on coarse planes m-1, m, m+1
        float B[5][3];
        struct Row { float l, c, h; };   // x = X - 1, X, X + 1 (clamped), scalars: no array to index
        auto pick = [](bool b, float t, float f) { return b ? t : f; };
        auto plane = [&](int m, int slot) {
            const int mc = min(max(m, 0), cz - 1);
            const float* c0 = c + (size_t)mc * cy * cx;
            auto row = [&](int y) { const float* q = c0 + y * cx; return Row{q[xa], q[xb], q[xc]}; };
            const Row R0 = row(ya), R1 = row(yb), R2 = row(yc);
            // 1-D pairs (x0, x1) of pw for the column's fine index
            auto rowC = [&](const Row& R) { return ax * R.c + bx * pick(pi, R.h, R.l); };
            auto rowW = [&](const Row& R) { return pi ? ax * R.c + bx * R.l : ax * R.l + bx * R.c; };
            auto rowE = [&](const Row& R) { return pi ? ax * R.h + bx * R.c : ax * R.c + bx * R.h; };
            // the centre column's second y row: R2 for odd j, else R0
            const Row& Ry = pj ? R2 : R0;
            B[0][slot] = ax * rowC(R1) + bx * rowC(Ry);
            B[1][slot] = ax * rowW(R1) + bx * rowW(Ry);
            B[2][slot] = ax * rowE(R1) + bx * rowE(Ry);
            B[3][slot] = pj ? ax * rowC(R1) + bx * rowC(R0) : ax * rowC(R0) + bx * rowC(R1);   // j - 1
            B[4][slot] = pj ? ax * rowC(R2) + bx * rowC(R1) : ax * rowC(R1) + bx * rowC(R2);   // j + 1
        };
        int m = k >> 1;
        plane(m - 1, 0);
        plane(m, 1);
        plane(m + 1, 2);
        // s of column col at fine plane f, whose coarse pair is (f >> 1, its neighbour):
        // slots (1, f odd ? 2 : 0) when f >> 1 == m, (2, 1) when f >> 1 == m + 1 (f even)
        auto sAt = [&](int col, int f) {
            if ((f >> 1) == m) return (f & 1) ? ax * B[col][1] + bx * B[col][2] : ax * B[col][1] + bx * B[col][0];
            if ((f >> 1) > m) return ax * B[col][2] + bx * B[col][1];
            return ax * B[col][0] + bx * B[col][1];   // f = k - 1 odd below an even k
        };
        float sd = k > 0 ? sAt(0, k - 1) : 0.f;
        float sc = sAt(0, k);
        for (; k < ke; ++k, p += sK) {
            if (k + 1 < ke) { loadPlane(p + sK, nxt, false); nxt.a[5] = cur.a[6]; }
            const float su = k + 1 < D.nK ? sAt(0, k + 1) : 0.f;
            const float sw = inW ? sAt(1, k) : 0.f, se = inE ? sAt(2, k) : 0.f;
            const float ss = inS ? sAt(3, k) : 0.f, sn = inN ? sAt(4, k) : 0.f;
            const float As = cur.a[0] * sc + (cur.a[1] * sw + cur.a[2] * se + cur.a[3] * ss + cur.a[4] * sn
                                              + cur.a[5] * sd + cur.a[6] * su);
            const float zz = (NET_OMEGA * cur.d * cur.r + sc) + NET_OMEGA * (cur.w - cur.d * As);
            z[p] = zz;
            if (PART) {
                v[0] += (double)zz * (double)cur.q;
                v[1] += (double)zz;
                v[2] += (double)zz * (double)cur.r;
            }
            sd = sc; sc = su;
            cur = nxt;
            if (k & 1) {   // the next plane is even: shift the window up one coarse plane
#pragma unroll
                for (int col = 0; col < 5; ++col) { B[col][0] = B[col][1]; B[col][1] = B[col][2]; }
                ++m;
                if (k + 1 < ke) plane(m + 1, 2);
            }
        }
    }
    if (PART) writePartials<FB, 3>(v, part);
}

// One more damped Jacobi sweep after the cycle's post-smoothing: z = y + w D^-1 (r - A y)
// for y from k_f_post; partials (z, q_old), sum z, (z, r) (they replace k_f_post's).
// Column tiles and loads one plane ahead as in k_f_k1.
template <bool PART>
__global__ void k_f_sweep(Dims D, SymF S, ColMap cm, const float* __restrict__ dinv, const float* r0, const float* r1,
                          const float* __restrict__ y, float* z, const float* q0, const float* q1, double* part,
                          const double* st) {
    const double sDone = st[F_DONE], sIt = st[F_IT];
    if (sDone != 0.) return;
    const int it = (int)sIt;
    const float* __restrict__ qold = (it & 1) ? q0 : q1;
    const float* __restrict__ r = (it & 1) ? r1 : r0;
    const int P = S.P, sJ = S.sJ, sK = S.sK;
    const float* __restrict__ A = S.A;
    struct Plane { float ys[4], yu, a[7], r, d, q; };
    auto loadPlane = [&](int p, Plane& L, bool a5) {   // a5 false: a[5] is the plane below's a[6]
        const int o[4] = {-1, 1, -sJ, sJ};
#pragma unroll
        for (int m = 0; m < 4; ++m) L.ys[m] = y[p + o[m]];
        L.yu = y[p + sK];
        L.a[0] = A[p]; L.a[1] = A[P + p]; L.a[2] = A[P + p + 1]; L.a[3] = A[2 * P + p];
        L.a[4] = A[2 * P + p + sJ]; if (a5) L.a[5] = A[3 * P + p]; L.a[6] = A[3 * P + p + sK];
        L.r = r[p]; L.d = dinv[p]; L.q = PART ? qold[p] : 0.f;
    };
    double v[3] = {0., 0., 0.};
    for (int v_ = xcdChunk(), vn = cm.ntx * cm.nty * cm.nseg; v_ < vn; v_ += (int)gridDim.x) {
        const int t_ = v_ / cm.nseg;
        const int i = (t_ % cm.ntx) * CTX + (int)(threadIdx.x % CTX), j = (t_ / cm.ntx) * CTY + (int)(threadIdx.x / CTX);
        int k = (v_ % cm.nseg) * cm.kper;
        const int ke = min(D.nK, k + cm.kper);
        if (!(i < D.nI && j < D.nJ) || k >= ke) continue;
        int p = D.pidx(i, j, k);
        float yd = y[p - sK], yc = y[p];
        Plane cur, nxt;
        loadPlane(p, cur, true);
        for (; k < ke; ++k, p += sK) {
            if (k + 1 < ke) { loadPlane(p + sK, nxt, false); nxt.a[5] = cur.a[6]; }
            const float Ay = cur.a[0] * yc + (cur.a[1] * cur.ys[0] + cur.a[2] * cur.ys[1] + cur.a[3] * cur.ys[2]
                                              + cur.a[4] * cur.ys[3] + cur.a[5] * yd + cur.a[6] * cur.yu);
            const float zz = yc + NET_OMEGA * cur.d * (cur.r - Ay);
            z[p] = zz;
            if (PART) {
                v[0] += (double)zz * (double)cur.q;
                v[1] += (double)zz;
                v[2] += (double)zz * (double)cur.r;
            }
            yd = yc; yc = cur.yu;
            cur = nxt;
        }
    }
    if (PART) writePartials<FB, 3>(v, part);
}

// beta = -(z - zbar, q_old) / (p_old, q_old); p = (z - zbar) + beta p_old formed at the
// stencil points; q = A p; partials (p,q), sum q.  (p, r) needs no pass of its own:
// alpha of the previous step makes (p_old, r) = 0, so (p, r) = (z - zbar, r), from the
// partials of k_f_post; block 0 records it for k_f_k2pre.
__global__ void k_f_k1(Dims D, SymF S, ColMap cm, const float* z, float* p0, float* p1, float* q0, float* q1,
                       const double* partIn, int nb, double* partOut, double* st, FcgCtl lc) {
    // the scalars written by earlier kernels are loaded together, ahead of the partial sums
    const double sDone = st[F_DONE], sIt = st[F_IT], sSq = st[F_SQ], sPq = st[F_PQ], sSr = st[F_SR];
    if (sDone != 0.) return;
    const int it = (int)sIt;
    const float* pold = (it & 1) ? p0 : p1;
    float* pnew = (it & 1) ? p1 : p0;
    float* qnew = (it & 1) ? q1 : q0;
    // Each thread walks its column up k with a rolling window: the k - 1, k and k + 1 values
    // of z and p_old stay in registers, and the loads of the next plane (its four side
    // neighbours, the value above it and the seven matrix entries) are issued before the
    // current cell's stores, which the compiler may not move past them.  (Two planes ahead
    // measured no faster.)  The first column's first loads need no scalar, so they are
    // issued before the partial sums are read, and the two latencies overlap (at 128^3
    // every block takes one column tile).  The arithmetic is SymF::apply's, in its order.
    const int P = S.P, sJ = S.sJ, sK = S.sK;
    const float* A = S.A;
    struct Plane { float zs[4], ps[4], zu, pu, a[7]; };   // sides W E S N, k + 1 (u), matrix
    auto loadPlane = [&](int p, Plane& L, bool a5) {   // a5 false: a[5] is the plane below's a[6]
        const int o[4] = {-1, 1, -sJ, sJ};
#pragma unroll
        for (int m = 0; m < 4; ++m) { L.zs[m] = z[p + o[m]]; L.ps[m] = pold[p + o[m]]; }
        L.zu = z[p + sK]; L.pu = pold[p + sK];
        L.a[0] = A[p]; L.a[1] = A[P + p]; L.a[2] = A[P + p + 1]; L.a[3] = A[2 * P + p];
        L.a[4] = A[2 * P + p + sJ]; if (a5) L.a[5] = A[3 * P + p]; L.a[6] = A[3 * P + p + sK];
    };
    // work item v: this thread's column (i, j) and its k range [k, ke); false if empty
    auto item = [&](int v_, int& p, int& k, int& ke) {
        const int t_ = v_ / cm.nseg;
        const int i = (t_ % cm.ntx) * CTX + (int)(threadIdx.x % CTX), j = (t_ / cm.ntx) * CTY + (int)(threadIdx.x / CTX);
        k = (v_ % cm.nseg) * cm.kper;
        ke = min(D.nK, k + cm.kper);
        if (!(i < D.nI && j < D.nJ) || k >= ke) return false;
        p = D.pidx(i, j, k);
        return true;
    };
    const int v0 = xcdChunk(), vn = cm.ntx * cm.nty * cm.nseg;
    int p = 0, k = 0, ke = 0;
    bool live = v0 < vn && item(v0, p, k, ke);
    float zd = 0.f, pd = 0.f, zc = 0.f, pc = 0.f;   // k - 1, k
    Plane cur, nxt;
    if (live) {
        zd = z[p - sK]; pd = pold[p - sK]; zc = z[p]; pc = pold[p];
        loadPlane(p, cur, true);
    }
    double t[3];
    readPartials<FB, 3>(partIn, nb, t);
    const double zbar = lc.nLive > 0. ? t[1] / lc.nLive : 0.;
    const double zq = t[0] - zbar * sSq;
    const double beta = it == 0 || sPq == 0. ? 0. : -zq / sPq;
    if (blockIdx.x == 0 && threadIdx.x == 0) { st[F_PAR] = (double)(it & 1); st[F_PR] = t[2] - zbar * sSr; }
    const float zb = (float)zbar, bt = (float)beta;
    double v[2] = {0., 0.};
    auto f = [&](float zz, float pp) { return (zz - zb) + bt * pp; };
    for (int v_ = v0; v_ < vn; v_ += (int)gridDim.x) {
        if (v_ != v0) {
            live = item(v_, p, k, ke);
            if (live) {
                zd = z[p - sK]; pd = pold[p - sK]; zc = z[p]; pc = pold[p];
                loadPlane(p, cur, true);
            }
        }
        if (!live) continue;
        for (; k < ke; ++k, p += sK) {
            if (k + 1 < ke) { loadPlane(p + sK, nxt, false); nxt.a[5] = cur.a[6]; }
            const float fc = f(zc, pc);
            const float qq = cur.a[0] * fc + (cur.a[1] * f(cur.zs[0], cur.ps[0]) + cur.a[2] * f(cur.zs[1], cur.ps[1])
                                              + cur.a[3] * f(cur.zs[2], cur.ps[2]) + cur.a[4] * f(cur.zs[3], cur.ps[3])
                                              + cur.a[5] * f(zd, pd) + cur.a[6] * f(cur.zu, cur.pu));
            pnew[p] = fc;
            qnew[p] = qq;
            v[0] += (double)fc * (double)qq;
            v[1] += (double)qq;
            zd = zc; pd = pc; zc = cur.zu; pc = cur.pu;
            cur = nxt;
        }
    }
    writePartials<FB, 2>(v, partOut);
}

// alpha = (p,r)/(p,q); e += alpha p; r' = r - alpha q into the other residual buffer;
// partials (r',r'), sum r'; and the next iteration's pre-smoothing from r' formed at the
// stencil points: w = D^-1 (r' - A w D^-1 r').  The convergence test on these partials
// is made by the next k_f_restrict, so a w computed after the last iteration is unused.
// Column tiles as in k_f_k1: each thread walks its column up k with z1 = w D^-1 r' of the
// k - 1, k and k + 1 cells in registers; the loads of the next plane (r, q and D^-1 at the
// four side neighbours and above, the seven matrix entries, e and p) are issued before the
// current cell's stores.  D^-1 comes from the array k_f_dinv fills once per solve (the
// same values SymF::inv gives); the first plane's loads go ahead of the partial sums.
// The arithmetic is SymF::apply's, in its order.
__global__ void k_f_k2pre(Dims D, SymF S, ColMap cm, const float* __restrict__ dinv, float* e, const float* r0,
                          const float* r1, float* w, const float* p0, const float* p1, const float* q0, const float* q1,
                          const double* partIn, int nb, double* partOut, double* st) {
    const double sDone = st[F_DONE], sPar = st[F_PAR], sPr = st[F_PR], sIt = st[F_IT];
    if (sDone != 0.) return;
    const int par = (int)sPar;               // this iteration's parity, set by k_f_k1
    const float* __restrict__ pv = par ? p1 : p0;
    const float* __restrict__ qv = par ? q1 : q0;
    const float* __restrict__ rc = par ? r1 : r0;
    float* __restrict__ rn = (float*)(par ? r0 : r1);
    const int P = S.P, sJ = S.sJ, sK = S.sK;
    const float* __restrict__ A = S.A;
    struct Plane { float rs[4], qs[4], ds[4], ru, qu, du, a[7], ee, pp; };   // sides W E S N, k + 1 (u)
    auto loadPlane = [&](int p, Plane& L, bool a5) {   // a5 false: a[5] is the plane below's a[6]
        const int o[4] = {-1, 1, -sJ, sJ};
#pragma unroll
        for (int m = 0; m < 4; ++m) { L.rs[m] = rc[p + o[m]]; L.qs[m] = qv[p + o[m]]; L.ds[m] = dinv[p + o[m]]; }
        L.ru = rc[p + sK]; L.qu = qv[p + sK]; L.du = dinv[p + sK];
        L.a[0] = A[p]; L.a[1] = A[P + p]; L.a[2] = A[P + p + 1]; L.a[3] = A[2 * P + p];
        L.a[4] = A[2 * P + p + sJ]; if (a5) L.a[5] = A[3 * P + p]; L.a[6] = A[3 * P + p + sK];
        L.ee = e[p]; L.pp = pv[p];
    };
    auto item = [&](int v_, int& p, int& k, int& ke) {
        const int t_ = v_ / cm.nseg;
        const int i = (t_ % cm.ntx) * CTX + (int)(threadIdx.x % CTX), j = (t_ / cm.ntx) * CTY + (int)(threadIdx.x / CTX);
        k = (v_ % cm.nseg) * cm.kper;
        ke = min(D.nK, k + cm.kper);
        if (!(i < D.nI && j < D.nJ) || k >= ke) return false;
        p = D.pidx(i, j, k);
        return true;
    };
    const int v0 = xcdChunk(), vn = cm.ntx * cm.nty * cm.nseg;
    int p = 0, k = 0, ke = 0;
    bool live = v0 < vn && item(v0, p, k, ke);
    // raw values below and at the cell (r, q, D^-1), loaded before alpha is known
    float rd = 0.f, qd = 0.f, dd = 0.f, rcc = 0.f, qcc = 0.f, dcc = 0.f;
    Plane cur, nxt;
    if (live) {
        rd = rc[p - sK]; qd = qv[p - sK]; dd = dinv[p - sK];
        rcc = rc[p]; qcc = qv[p]; dcc = dinv[p];
        loadPlane(p, cur, true);
    }
    double t[2];
    readPartials<FB, 2>(partIn, nb, t);
    const double alpha = t[0] != 0. ? sPr / t[0] : 0.;
    if (blockIdx.x == 0 && threadIdx.x == 0) { st[F_PQ] = t[0]; st[F_SQ] = t[1]; st[F_IT] = sIt + 1.; }
    const float al = (float)alpha;
    auto rnew = [&](float rr_, float qq_) { return rr_ - al * qq_; };
    auto z1 = [&](float d_, float rr_, float qq_) { return NET_OMEGA * d_ * rnew(rr_, qq_); };
    double v[2] = {0., 0.};
    for (int v_ = v0; v_ < vn; v_ += (int)gridDim.x) {
        if (v_ != v0) {
            live = item(v_, p, k, ke);
            if (live) {
                rd = rc[p - sK]; qd = qv[p - sK]; dd = dinv[p - sK];
                rcc = rc[p]; qcc = qv[p]; dcc = dinv[p];
                loadPlane(p, cur, true);
            }
        }
        if (!live) continue;
        float zd = z1(dd, rd, qd);
        float zc = z1(dcc, rcc, qcc);
        for (; k < ke; ++k, p += sK) {
            if (k + 1 < ke) { loadPlane(p + sK, nxt, false); nxt.a[5] = cur.a[6]; }
            const float rp = rnew(rcc, qcc);
            const float zu = z1(cur.du, cur.ru, cur.qu);
            const float Az = cur.a[0] * zc + (cur.a[1] * z1(cur.ds[0], cur.rs[0], cur.qs[0])
                                              + cur.a[2] * z1(cur.ds[1], cur.rs[1], cur.qs[1])
                                              + cur.a[3] * z1(cur.ds[2], cur.rs[2], cur.qs[2])
                                              + cur.a[4] * z1(cur.ds[3], cur.rs[3], cur.qs[3]) + cur.a[5] * zd
                                              + cur.a[6] * zu);
            e[p] = cur.ee + al * cur.pp;
            rn[p] = rp;
            w[p] = dcc * (rp - Az);
            v[0] += (double)rp * (double)rp;
            v[1] += (double)rp;
            zd = zc; zc = zu;
            rcc = cur.ru; qcc = cur.qu; dcc = cur.du;
            cur = nxt;
        }
    }
    writePartials<FB, 2>(v, partOut);
}

// D^-1 of the FP32 p' matrix over the whole padded array, as SymF::inv computes it, once per
// solve (the halo holds the identity rows' 1)
__global__ void k_f_dinv(SymF S, float* dinv) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p < S.P) dinv[p] = S.inv(p);
}

__global__ void k_f_axpy(Dims D, double* x, const float* e, double s) {
    FOR_FINE(D, q, p) { x[p] += s * (double)e[p]; }
}

// The launches of one FCG iteration: restriction (with the convergence test), the network,
// the post-smoothing (k_f_post, then k_f_sweep for each further sweep), the direction and
// the update with the next pre-smoothing.  The same
// launches serve every iteration of every round: the parities and the round's loop limits
// live in the device state.  (Replaying them from a captured HIP graph measured no faster
// at 128^3, 3.79 against 3.76 ms per solve: the gaps between the dependent kernels are not
// launch overhead.)
struct FcgIter {
    Ctx& c;
    Net& n;
    SymF S;
    ColMap cm;
    int g;
    float *r, *r1, *e, *w, *z, *p0, *p1, *q0, *q1, *dinv, *y, *h;
    double *st, *partA, *partB, *partC;
    FcgCtl lc;
    bool classical;   // the coarse solve is one multigrid V-cycle (the ablation), not the network
    // probe: record the kernel probe's events around its launches
    void launch(hipStream_t s, bool probe) const {
        const Dims& D = c.D;
        const int nc = (D.nI / 2) * (D.nJ / 2) * (D.nK / 2);
        const double* done = st + F_DONE;
        // (Fusing the restriction into the network's input layer, with the restriction of
        // the layer's halo tile formed in LDS from the fine cells under it, measured slower:
        // 215 against 199 us per iteration at 128^3, results bit-identical.)
        hipLaunchKernelGGL(k_f_restrict, dim3((nc + 255) / 256), dim3(256), 0, s, D, (const float*)w, n.vin,
                           (const double*)partC, g, st, lc);
        if (classical) {
            hipLaunchKernelGGL(k_f_restrict_dw, dim3((nc + 255) / 256), dim3(256), 0, s, D, (const float*)w, S.A, n.vin,
                               (const double*)st);
            mg_coarse_vcycle(c, n.vin, n.vout, done);
        } else {
            net_forward(n, s, n.vin, n.vout, done);
        }
        // the post-smoothing sweeps end in z; the extra ones alternate through y and h
        const int extra = c.fcgPost - 1;
        float* bufs[2] = {y, h};
        float* dst = extra == 0 ? z : bufs[(extra - 1) & 1];
        hipLaunchKernelGGL(extra == 0 ? k_f_post<true> : k_f_post<false>, dim3(g), dim3(FB), 0, s, D, S, cm, (const float*)dinv, (const float*)r,
                           (const float*)r1, (const float*)w, (const float*)n.vout, dst, (const float*)q0,
                           (const float*)q1, partA, (const double*)st);
        for (int q = extra - 1; q >= 0; --q) {
            float* nd = q == 0 ? z : bufs[(q - 1) & 1];
            hipLaunchKernelGGL(q == 0 ? k_f_sweep<true> : k_f_sweep<false>, dim3(g), dim3(FB), 0, s, D, S, cm, (const float*)dinv, (const float*)r,
                               (const float*)r1, (const float*)dst, nd, (const float*)q0, (const float*)q1, partA,
                               (const double*)st);
            dst = nd;
        }
        hipEvent_t e0, e1;
        if (probe && c.probe.slot(PROBE_FCG_K1, e0, e1))   // dispatch-bracketing events: the kernel's own time
            hipExtLaunchKernelGGL(k_f_k1, dim3(g), dim3(FB), 0, s, e0, e1, 0, D, S, cm, (const float*)z, p0, p1, q0, q1,
                                  (const double*)partA, g, partB, st, lc);
        else
            hipLaunchKernelGGL(k_f_k1, dim3(g), dim3(FB), 0, s, D, S, cm, (const float*)z, p0, p1, q0, q1,
                               (const double*)partA, g, partB, st, lc);
        if (probe && c.probe.slot(PROBE_FCG_K2PRE, e0, e1))
            hipExtLaunchKernelGGL(k_f_k2pre, dim3(g), dim3(FB), 0, s, e0, e1, 0, D, S, cm, (const float*)dinv, e,
                                  (const float*)r, (const float*)r1, w, (const float*)p0, (const float*)p1, (const float*)q0,
                                  (const float*)q1, (const double*)partB, g, partC, st);
        else
            hipLaunchKernelGGL(k_f_k2pre, dim3(g), dim3(FB), 0, s, D, S, cm, (const float*)dinv, e, (const float*)r,
                               (const float*)r1, w, (const float*)p0, (const float*)p1, (const float*)q0, (const float*)q1,
                               (const double*)partB, g, partC, st);
    }
};

int learned_fcg(Ctx& c, double shift, double bNorm, double rtol, int maxit, double delta, int* refs, double* relOut,
                bool* meanQueued) {
    Net& n = *(Net*)c.net;
    const Dims& D = c.D;
    *refs = 0;
    *meanQueued = false;
    if (!netDimsOk(D) || c.hasIB) return -1;
    HIPCHECK(hipSetDevice(n.device));
    tg_ensure_fine(n, D);
    if (!(n.nx == D.nI / 2 && n.ny == D.nJ / 2 && n.nz == D.nK / 2)) net_set_dims(n, D.nI / 2, D.nJ / 2, D.nK / 2, false);
    hipStream_t s = c.stream;
    if (!n.packed) net_pack(n, s);
    const int P = D.P;
    const int g = fgrid(D);
    auto fv = [&](int slot) { return n.fw + (size_t)slot * P; };
    const FcgIter I{c, n, symF(c.Ap32, D), colMap(D, g), g,
                    fv(FV_R), fv(FV_R1), fv(FV_E), fv(FV_W), fv(FV_Z), fv(FV_P0), fv(FV_P1), fv(FV_Q0), fv(FV_Q1),
                    fv(FV_DI), fv(FV_Y), fv(FV_H), c.scal + SC_CG, c.red.part, c.red.part + MAX_RED_BLOCKS * RED_STRIDE,
                    c.red.part + 2 * MAX_RED_BLOCKS * RED_STRIDE, FcgCtl{c.singularP ? (double)D.N : 0.},
                    c.fcgCoarse == 1};
    if (I.classical && !mg_coarse_prepare(c)) return -1;
    double* st = I.st;
    hipLaunchKernelGGL(k_f_dinv, dim3(blocksFor(P, 256)), dim3(256), 0, s, I.S, I.dinv);
    double rr = 0., bb = 0., rel = 0.;
    int its = 0;
    int lastRound = 0;
    bool stalled = false;
    // this solve's rounds: iterations and whether the queued ones ran out before the test
    int nRec = 0, recIts[FCG_MAX_PLAN] = {}, recCapped[FCG_MAX_PLAN] = {};
    auto record = [&](int it_, bool capped) {
        if (nRec < FCG_MAX_PLAN) { recIts[nRec] = it_; recCapped[nRec] = capped ? 1 : 0; }
        ++nRec;
    };
    // the plan slot: with |b| known, the slot made on the nearest |b| (a slot made on a |b|
    // more than 3x away is not used; an empty slot then takes this solve's rounds); without
    // it, the slots alternate
    int slot = c.fcgSolves & 1;
    if (bNorm > 0.) {
        auto dist = [&](int q) { return c.fcgPlanN[q] > 0 && c.fcgPlanB[q] > 0. ? std::fabs(std::log(bNorm / c.fcgPlanB[q])) : 1e300; };
        const double d0 = dist(0), d1 = dist(1), far = std::log(3.);
        if (d0 < 1e300 || d1 < 1e300) {
            slot = d0 <= d1 ? 0 : 1;
            if (std::min(d0, d1) > far) {   // too far: the other slot's plan is not used
                if (c.fcgPlanN[1 - slot] == 0) slot = 1 - slot;
                c.fcgPlanN[slot] = 0;
            }
        }
    }
    const double nLive = c.singularP ? (double)D.N : 0.;
    double* rec = c.scal + SC_FCGR;
    const double* fin = c.scal + SC_MISC + 8;
    auto relFrom = [&](const double* f) {   // the host formula of pcorr_true_resid
        rr = std::sqrt(nLive > 0. ? std::max(f[0] - f[2] * f[2] / nLive, 0.) : f[0]);
        bb = std::sqrt(f[1]);
        rel = bb > 0. ? rr / bb : 0.;
    };
    // the initial guess: the combination of the solutions this corrector's solves reached in
    // the two previous time steps (or outer iterations) with the smallest residual.  At
    // 128^3, once the flow has developed, the first corrector's systems then start at a
    // relative residual of about 0.03 (0.1 from the last solution alone) and take 10-11
    // iterations, not 14.
    pcorr_apply_guess(c, shift);
    if (c.fcgPlanN[slot] > 0) {
        // the rounds of the slot's last solve, queued with their decisions on the device and
        // one read-back at the end
        const int nq = c.fcgPlanN[slot];
        pcorr_true_resid_dev(c, shift);
        for (int q = 0; q < nq; ++q) {
            const RoundCtl rc{rtol, delta, nLive, q, maxit};
            hipLaunchKernelGGL(k_f_init_round, dim3(g), dim3(FB), 0, s, D, (const double*)(c.wp + 6 * P), fin, rec, I.r,
                               I.e, I.partC, st, rc);
            hipLaunchKernelGGL(k_f_pre0, dim3(g), dim3(FB), 0, s, D, I.S, (const float*)I.dinv, (const float*)I.r, I.w,
                               (const double*)st);
            for (int k = 0; k < c.fcgPlan[slot][q]; ++k) I.launch(s, true);
            hipLaunchKernelGGL(k_f_axpy_round, dim3(g), dim3(FB), 0, s, D, c.xp, I.e, rec, (const double*)st, q);
            pcorr_true_resid_dev(c, shift);
        }
        HIPCHECK(hipMemcpyAsync(c.hscal + SC_MISC + 8, fin, 3 * sizeof(double), hipMemcpyDeviceToHost, s));
        HIPCHECK(hipMemcpyAsync(c.hscal + SC_FCGR, rec, nq * FR_STRIDE * sizeof(double), hipMemcpyDeviceToHost, s));
        HIPCHECK(hipEventRecord(c.evR, s));
        // the solution's mean removal goes behind the read-back, so that the GPU runs it while
        // the host turns round.  It leaves the residual as it is (the residual of the singular
        // system leaves out its mean); the continuation below changes x and clears the flag.
        if (c.singularP) {
            pcorr_remove_mean(c);
            *meanQueued = true;
        }
        HIPCHECK(hipEventSynchronize(c.evR));
        relFrom(c.hscal + SC_MISC + 8);
        bool anyStop = false;
        int qLast = -1;
        for (int q = 0; q < nq; ++q) {
            const double* h = c.hscal + SC_FCGR + q * FR_STRIDE;
            if (h[FR_STOP] != 0.) { anyStop = true; break; }
            const int it = (int)h[FR_ITS];
            its += it;
            ++*refs;
            lastRound = it;
            record(it, h[FR_CAPPED] != 0.);
            qLast = q;
        }
        // a round the device skipped with the residual above rtol and iterations left is the
        // host loop's "no progress" exit
        stalled = anyStop && rel > rtol && its < maxit;
        const double* hl = qLast >= 0 ? c.hscal + SC_FCGR + qLast * FR_STRIDE : nullptr;
        if (!stalled && rel > rtol && hl && hl[FR_CAPPED] != 0.) {
            // the last round's queued iterations ran out before its test (this system needs more
            // than the plan): the round goes on from its FCG state, adding to x what it gains
            *meanQueued = false;
            const int it0 = (int)hl[FR_ITS];
            const double scale = hl[FR_SCALE], prev = hl[FR_REL];
            int it = it0, chunk = 1;
            while (true) {
                for (int k = 0; k < chunk; ++k) I.launch(s, true);
                HIPCHECK(hipMemcpyAsync(c.hscal + SC_CG, st, 12 * sizeof(double), hipMemcpyDeviceToHost, s));
                HIPCHECK(hipStreamSynchronize(s));
                it = (int)c.hscal[SC_CG + F_IT];
                if (c.hscal[SC_CG + F_DONE] != 0. || it - it0 >= 300) break;
                chunk = 2;
            }
            its += it - it0;
            lastRound = it;
            if (nRec - 1 < FCG_MAX_PLAN) { recIts[nRec - 1] = it; recCapped[nRec - 1] = 0; }
            hipLaunchKernelGGL(k_f_axpy, dim3(g), dim3(FB), 0, s, D, c.xp, (const float*)I.e, 1.0 / scale);
            pcorr_true_resid(c, shift, &rr, &bb);
            rel = rr / bb;
            stalled = !(rel < 0.5 * prev);
        }
    } else {
        pcorr_true_resid(c, shift, &rr, &bb);
        rel = bb > 0. ? rr / bb : 0.;
    }
    while (!stalled && rel > rtol && its < maxit && *refs < 40) {
        // the inner solve gains `delta`, or only what the last round still needs
        const double need = std::max(delta, 0.5 * rtol / rel);
        const double tol2 = need * need;
        // a refinement round that needs more than 300 iterations means the cycle does not
        // fit this system: the round ends there and the progress test below decides
        const int roundMax = std::min(maxit - its, 300);
        const double scale = 1.0 / rr;
        hipLaunchKernelGGL(k_f_init, dim3(g), dim3(FB), 0, s, D, (const double*)(c.wp + 6 * P), scale, I.r, I.e, I.partC, st,
                           tol2, roundMax);
        hipLaunchKernelGGL(k_f_pre0, dim3(g), dim3(FB), 0, s, D, I.S, (const float*)I.dinv, (const float*)I.r, I.w,
                           (const double*)st);
        // iterations are queued in chunks between convergence checks: the first chunk of a
        // round is the previous round's count, later ones the count the observed rate of
        // decrease still needs (a queued iteration past convergence costs 16 launches
        // that return at once)
        int chunk = lastRound > 0 ? lastRound : 6, it = 0;
        while (true) {
            for (int k = 0; k < chunk; ++k) I.launch(s, true);
            HIPCHECK(hipMemcpyAsync(c.hscal + SC_CG, st, 8 * sizeof(double), hipMemcpyDeviceToHost, s));
            HIPCHECK(hipStreamSynchronize(s));
            it = (int)c.hscal[SC_CG + F_IT];
            if (c.hscal[SC_CG + F_DONE] != 0. || it >= roundMax) break;
            // F_RR is the residual checked before the last queued iteration (it - 1 done)
            const double rr0 = c.hscal[SC_CG + F_RR], bb0 = c.hscal[SC_CG + F_BB];
            chunk = 4;
            if (it > 1 && rr0 > 0. && bb0 > 0. && rr0 < bb0) {
                const double rate = std::log(rr0 / bb0) / (it - 1);          // log decrease per iteration
                const double still = std::log(tol2 * bb0 / rr0) / rate;    // iterations still needed from it - 1
                chunk = std::max(1, std::min(32, (int)std::ceil(still)));
            }
        }
        lastRound = it;
        its += it;
        record(it, false);
        hipLaunchKernelGGL(k_f_axpy, dim3(g), dim3(FB), 0, s, D, c.xp, (const float*)I.e, 1.0 / scale);
        const double prev = rel;
        pcorr_true_resid(c, shift, &rr, &bb);
        rel = rr / bb;
        ++*refs;
        if (!(rel < 0.5 * prev) || it == 0) break;
    }
    // the plan for the next solve of this slot: this solve's rounds if it converged, as long
    // as they ran (one iteration longer where the queued ones ran out before the round's
    // test).  A queued iteration past convergence costs ~60 us of launches that return at
    // once; a round that runs out is continued after the read-back, at one more read-back.
    c.fcgPlanN[slot] = (rel <= rtol && nRec > 0 && nRec <= FCG_MAX_PLAN) ? nRec : 0;
    c.fcgPlanB[slot] = bNorm;
    for (int q = 0; q < c.fcgPlanN[slot]; ++q) c.fcgPlan[slot][q] = recIts[q] + (recCapped[q] ? 1 : 0);
    if (rel <= rtol) pcorr_save_guess(c);   // the next steps' guess for this corrector
    ++c.fcgSolves;
    *relOut = rel;
    return its;
}

}  // namespace caffe

// ================================================================== C ABI
using namespace caffe;

namespace {

template <class F>
int nguard(caffe_net* x, F f) {
    if (!x) return CAFFE_ERR_ARG;
    try {
        HIPCHECK(hipSetDevice(x->n.device));
        return f(x->n);
    } catch (const std::string& e) {
        x->n.err = e;
        return CAFFE_ERR_HIP;
    } catch (const std::exception& e) {
        x->n.err = e.what();
        return CAFFE_ERR_HIP;
    }
}

// splitmix64 + Box-Muller normal deviates (deterministic initialisation)
struct Rng {
    uint64_t s;
    uint64_t next() {
        uint64_t z = (s += 0x9e3779b97f4a7c15ULL);
        z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
        z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
        return z ^ (z >> 31);
    }
    double uni() { return ((next() >> 11) + 0.5) * (1.0 / 9007199254740992.0); }
    double normal() { return std::sqrt(-2.0 * std::log(uni())) * std::cos(6.283185307179586 * uni()); }
};

// A synthetic pair's modes: one band per pair, kmax = 2^B with B uniform in
// [0, log2(n/2)] (up to the coarse grid's highest mode), wave numbers uniform in
// [0, kmax] per direction, N(0,1) amplitudes.  Each band weighs the same in the loss.
Modes drawModes(Rng& r, const Dims& D) {
    const int nmin = std::min(D.nI, std::min(D.nJ, D.nK));
    int lmax = 0;
    while ((2 << lmax) <= nmin / 2) ++lmax;
    const int B = (int)(r.next() % (uint64_t)(lmax + 1));
    const int kmax = 1 << B;
    Modes m;
    for (int t = 0; t < N_MODES; ++t) {
        for (int d = 0; d < 3; ++d) m.k[t][d] = (int)(r.next() % (uint64_t)(kmax + 1));
        if (m.k[t][0] == 0 && m.k[t][1] == 0 && m.k[t][2] == 0) m.k[t][r.next() % 3] = 1;
        m.a[t] = (float)r.normal();
    }
    return m;
}

void freeNet(Net& n) {
    (void)hipDeviceSynchronize();
    for (void* p : n.dimAllocs) (void)hipFree(p);
    for (void* p : n.fineAllocs) (void)hipFree(p);
    for (auto& smp : n.data) { (void)hipFree(smp.A); (void)hipFree(smp.b); (void)hipFree(smp.x); }
    for (void* p : n.allocs) (void)hipFree(p);
    n.dimAllocs.clear(); n.fineAllocs.clear(); n.data.clear(); n.allocs.clear();
}

Dims fineDims(int nI, int nJ, int nK) {
    Dims D;
    D.nI = nI; D.nJ = nJ; D.nK = nK;
    D.sJ = nI + 2; D.sK = (nI + 2) * (nJ + 2);
    D.P = D.sK * (nK + 2);
    D.N = nI * nJ * nK;
    return D;
}

__global__ void k_collect(Dims D, const double* Ap, const double* bp, const double* xp, const int8_t* st, double shift,
                          double xshift, double s, float* A, float* b, float* x) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= D.P) return;
    for (int m = 0; m < 4; ++m) A[m * D.P + p] = (float)Ap[m * D.P + p];
    int i = p % D.sJ, j = (p / D.sJ) % (D.nJ + 2), k = p / D.sK;
    const bool in = i >= 1 && i <= D.nI && j >= 1 && j <= D.nJ && k >= 1 && k <= D.nK;
    const bool live = in && st[p] != ST_INACTIVE;
    b[p] = live ? (float)((bp[p] - shift) * s) : 0.f;
    x[p] = live ? (float)((xp[p] - xshift) * s) : 0.f;
}
__global__ void k_negate(float* a, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) a[i] = -a[i];
}

// sum and count of b over the cells that carry a row
__global__ void k_sum_live(Dims D, const double* b, const int8_t* st, double* part) {
    double v[2] = {0., 0.};
    FOR_FINE(D, q, p) {
        if (st[p] == ST_INACTIVE) continue;
        v[0] += b[p];
        v[1] += 1.;
    }
    writePartials<FB, 2>(v, part);
}
__global__ void k_sumsq(Dims D, const double* b, const int8_t* st, double shift, double* part) {
    double v[1] = {0.};
    FOR_FINE(D, q, p) {
        if (st[p] == ST_INACTIVE) continue;
        const double d = b[p] - shift;
        v[0] += d * d;
    }
    writePartials<FB, 1>(v, part);
}
__global__ void k_final2(const double* part, int nb, double* out) {
    double t[2];
    readPartials<FB, 2>(part, nb, t);
    if (threadIdx.x == 0) { out[0] = t[0]; out[1] = t[1]; }
}

Sample uploadSample(Net& n, const Dims& D, const float* A, const float* b, const float* x) {
    Sample smp;
    smp.D = D;
    HIPCHECK(hipMalloc(&smp.A, 4 * (size_t)D.P * sizeof(float)));
    HIPCHECK(hipMalloc(&smp.b, (size_t)D.P * sizeof(float)));
    HIPCHECK(hipMalloc(&smp.x, (size_t)D.P * sizeof(float)));
    HIPCHECK(hipMemcpy(smp.A, A, 4 * (size_t)D.P * sizeof(float), hipMemcpyHostToDevice));
    HIPCHECK(hipMemcpy(smp.b, b, (size_t)D.P * sizeof(float), hipMemcpyHostToDevice));
    if (x) HIPCHECK(hipMemcpy(smp.x, x, (size_t)D.P * sizeof(float), hipMemcpyHostToDevice));
    else HIPCHECK(hipMemset(smp.x, 0, (size_t)D.P * sizeof(float)));
    (void)n;
    return smp;
}
void freeSample(Sample& s) {
    (void)hipFree(s.A); (void)hipFree(s.b); (void)hipFree(s.x);
    s.A = s.b = s.x = nullptr;
}

int checkFine(Net& n, int nI, int nJ, int nK) {
    if (nI < 16 || nJ < 16 || nK < 16 || nI % 16 || nJ % 16 || nK % 16) {
        n.err = "the learned preconditioner needs fine-grid dimensions divisible by 16";
        return CAFFE_ERR_ARG;
    }
    return CAFFE_OK;
}

}  // namespace

extern "C" {

int caffe_net_create(int device, unsigned long long seed, caffe_net** out) {
    if (!out) return CAFFE_ERR_ARG;
    *out = nullptr;
    caffe_net* x = new caffe_net();
    x->n.device = device;
    int rc = nguard(x, [&](Net& n) {
        HIPCHECK(hipStreamCreateWithFlags(&n.stream, hipStreamNonBlocking));
        HIPCHECK(hipEventCreate(&n.ev0));
        HIPCHECK(hipEventCreate(&n.ev1));
        long o = 0, po = 0;
        for (int l = 0; l < NET_LAYERS; ++l) {
            n.off[l] = o;
            n.poff[l] = po;
            const long L = 27L * kNet[l].cin * kNet[l].cout;
            o += L;
            if (l != NET_IN && l != NET_OUT) po += L;
        }
        n.poff[NET_OUT] = po;   // the forward packing also holds the out layer's BF16 hi/lo columns
        n.nParams = o;
        n.theta = (float*)net_alloc(n, o * 4, n.allocs);
        n.grad = (float*)net_alloc(n, o * 4, n.allocs);
        n.m1 = (float*)net_alloc(n, o * 4, n.allocs);
        n.m2 = (float*)net_alloc(n, o * 4, n.allocs);
        n.wf = (uint16_t*)net_alloc(n, (po + 27 * 256) * 2, n.allocs);
        n.wb = (uint16_t*)net_alloc(n, po * 2, n.allocs);
        n.part = (double*)net_alloc(n, 2 * MAX_RED_BLOCKS * RED_STRIDE * sizeof(double), n.allocs);
        n.scal = (double*)net_alloc(n, 64 * sizeof(double), n.allocs);
        HIPCHECK(hipHostMalloc((void**)&n.hscal, 64 * sizeof(double), hipHostMallocDefault));
        // variance-preserving initialisation of a linear network; the output layer small
        std::vector<float> h(o);
        Rng rng{seed * 0x2545F4914F6CDD1DULL + 1};
        for (int l = 0; l < NET_LAYERS; ++l) {
            const long L = 27L * kNet[l].cin * kNet[l].cout;
            const double sd = (l == NET_OUT ? 0.1 : 1.0) / std::sqrt(27.0 * kNet[l].cin);
            for (long i = 0; i < L; ++i) h[n.off[l] + i] = (float)(sd * rng.normal());
        }
        HIPCHECK(hipMemcpy(n.theta, h.data(), o * 4, hipMemcpyHostToDevice));
        net_pack(n, n.stream);
        HIPCHECK(hipStreamSynchronize(n.stream));
        return CAFFE_OK;
    });
    if (rc != CAFFE_OK) {
        freeNet(x->n);
        delete x;
        return rc;
    }
    *out = x;
    return CAFFE_OK;
}

int caffe_net_destroy(caffe_net* x) {
    if (!x) return CAFFE_ERR_ARG;
    Net& n = x->n;
    (void)hipSetDevice(n.device);
    freeNet(n);
    if (n.hscal) (void)hipHostFree(n.hscal);
    if (n.ev0) (void)hipEventDestroy(n.ev0);
    if (n.ev1) (void)hipEventDestroy(n.ev1);
    if (n.stream) (void)hipStreamDestroy(n.stream);
    delete x;
    return CAFFE_OK;
}

const char* caffe_net_last_error(caffe_net* x) { return x ? x->n.err.c_str() : "null network"; }

long caffe_net_param_count(void) {
    long o = 0;
    for (int l = 0; l < NET_LAYERS; ++l) o += 27L * kNet[l].cin * kNet[l].cout;
    return o;
}

int caffe_net_get_params(caffe_net* x, float* host, long n_) {
    return nguard(x, [&](Net& n) {
        if (!host || n_ != n.nParams) { n.err = "parameter count mismatch"; return CAFFE_ERR_ARG; }
        HIPCHECK(hipStreamSynchronize(n.stream));
        HIPCHECK(hipMemcpy(host, n.theta, n.nParams * 4, hipMemcpyDeviceToHost));
        return CAFFE_OK;
    });
}

int caffe_net_set_params(caffe_net* x, const float* host, long n_) {
    return nguard(x, [&](Net& n) {
        if (!host || n_ != n.nParams) { n.err = "parameter count mismatch"; return CAFFE_ERR_ARG; }
        for (long i = 0; i < n_; ++i)
            if (!std::isfinite(host[i])) { n.err = "non-finite parameter"; return CAFFE_ERR_ARG; }
        HIPCHECK(hipStreamSynchronize(n.stream));
        HIPCHECK(hipMemcpy(n.theta, host, n.nParams * 4, hipMemcpyHostToDevice));
        HIPCHECK(hipMemset(n.m1, 0, n.nParams * 4));
        HIPCHECK(hipMemset(n.m2, 0, n.nParams * 4));
        n.adamT = 0;
        net_pack(n, n.stream);
        HIPCHECK(hipStreamSynchronize(n.stream));
        return CAFFE_OK;
    });
}

int caffe_net_forward(caffe_net* x, int nx, int ny, int nz, const float* in, float* out) {
    return nguard(x, [&](Net& n) {
        if (!in || !out || nx < 8 || ny < 8 || nz < 8 || nx % 8 || ny % 8 || nz % 8) {
            n.err = "coarse dimensions must be multiples of 8";
            return CAFFE_ERR_ARG;
        }
        net_set_dims(n, nx, ny, nz, n.trainBufs);
        const size_t nv = (size_t)nx * ny * nz;
        HIPCHECK(hipMemcpyAsync(n.vin, in, nv * 4, hipMemcpyHostToDevice, n.stream));
        net_forward(n, n.stream, n.vin, n.vout, nullptr);
        HIPCHECK(hipMemcpyAsync(out, n.vout, nv * 4, hipMemcpyDeviceToHost, n.stream));
        HIPCHECK(hipStreamSynchronize(n.stream));
        return CAFFE_OK;
    });
}

int caffe_net_backward(caffe_net* x, int nx, int ny, int nz, const float* in, const float* gout, float* gparams) {
    return nguard(x, [&](Net& n) {
        if (!in || !gout || !gparams || nx < 8 || ny < 8 || nz < 8 || nx % 8 || ny % 8 || nz % 8) {
            n.err = "coarse dimensions must be multiples of 8";
            return CAFFE_ERR_ARG;
        }
        net_set_dims(n, nx, ny, nz, true);
        const size_t nv = (size_t)nx * ny * nz;
        HIPCHECK(hipMemcpyAsync(n.vin, in, nv * 4, hipMemcpyHostToDevice, n.stream));
        HIPCHECK(hipMemcpyAsync(n.gout, gout, nv * 4, hipMemcpyHostToDevice, n.stream));
        HIPCHECK(hipMemsetAsync(n.grad, 0, n.nParams * 4, n.stream));
        net_forward(n, n.stream, n.vin, n.vout, nullptr);
        net_backward(n, n.stream, n.vin, n.gout);
        HIPCHECK(hipMemcpyAsync(gparams, n.grad, n.nParams * 4, hipMemcpyDeviceToHost, n.stream));
        HIPCHECK(hipStreamSynchronize(n.stream));
        return CAFFE_OK;
    });
}

int caffe_net_set_post_sweeps(caffe_net* x, int sweeps) {
    return nguard(x, [&](Net& n) {
        if (sweeps < 1 || sweeps > 4) { n.err = "post-smoothing sweeps of the cycle: 1 to 4"; return CAFFE_ERR_ARG; }
        n.post = sweeps;
        return CAFFE_OK;
    });
}

int caffe_net_precond(caffe_net* x, int nI, int nJ, int nK, const float* A, const float* r, float* z) {
    return nguard(x, [&](Net& n) {
        if (!A || !r || !z) { n.err = "null argument"; return CAFFE_ERR_ARG; }
        if (int rc = checkFine(n, nI, nJ, nK)) return rc;
        const Dims D = fineDims(nI, nJ, nK);
        HIPCHECK(hipStreamSynchronize(n.stream));
        Sample smp = uploadSample(n, D, A, r, nullptr);
        tg_ensure_fine(n, D);
        net_set_dims(n, nI / 2, nJ / 2, nK / 2, n.trainBufs);
        float* zd = n.fw + (size_t)FV_E * D.P;   // not one of the vectors the cycle works in
        HIPCHECK(hipMemsetAsync(zd, 0, (size_t)D.P * 4, n.stream));
        tg_apply(n, n.stream, D, smp.A, smp.b, zd, nullptr);
        HIPCHECK(hipMemcpyAsync(z, zd, (size_t)D.P * 4, hipMemcpyDeviceToHost, n.stream));
        HIPCHECK(hipStreamSynchronize(n.stream));
        freeSample(smp);
        return CAFFE_OK;
    });
}

int caffe_net_loss_grad(caffe_net* x, int nI, int nJ, int nK, const float* A, const float* b, const float* xs,
                        double* loss, float* gparams) {
    return nguard(x, [&](Net& n) {
        if (!A || !b || !xs || !loss) { n.err = "null argument"; return CAFFE_ERR_ARG; }
        if (int rc = checkFine(n, nI, nJ, nK)) return rc;
        const Dims D = fineDims(nI, nJ, nK);
        HIPCHECK(hipStreamSynchronize(n.stream));
        Sample smp = uploadSample(n, D, A, b, xs);
        tg_ensure_fine(n, D);
        net_set_dims(n, nI / 2, nJ / 2, nK / 2, gparams != nullptr || n.trainBufs);
        smp.nrm = sampleNorm(n, n.stream, smp);
        if (!(smp.nrm > 0.)) { freeSample(smp); n.err = "x^T A x is not positive"; return CAFFE_ERR_ARG; }
        float* ld = (float*)(n.scal + 8);
        HIPCHECK(hipMemsetAsync(ld, 0, sizeof(float), n.stream));
        if (gparams) HIPCHECK(hipMemsetAsync(n.grad, 0, n.nParams * 4, n.stream));
        tg_loss(n, n.stream, smp, gparams != nullptr, ld);
        float lh = 0.f;
        HIPCHECK(hipMemcpyAsync(&lh, ld, sizeof(float), hipMemcpyDeviceToHost, n.stream));
        if (gparams) HIPCHECK(hipMemcpyAsync(gparams, n.grad, n.nParams * 4, hipMemcpyDeviceToHost, n.stream));
        HIPCHECK(hipStreamSynchronize(n.stream));
        freeSample(smp);
        *loss = lh;
        return CAFFE_OK;
    });
}

int caffe_net_collect(caffe_net* x, caffe_ctx* cx, long* count) {
    return nguard(x, [&](Net& n) {
        if (!cx) { n.err = "null context"; return CAFFE_ERR_ARG; }
        Ctx& c = cx->c;
        if (!c.ready) { n.err = "context has no case"; return CAFFE_ERR_STATE; }
        if (c.hasIB) { n.err = "ghost-cell (nonsymmetric) systems are not collected"; return CAFFE_ERR_STATE; }
        if (c.device != n.device) { n.err = "context and network are on different devices"; return CAFFE_ERR_ARG; }
        if (int rc = checkFine(n, c.D.nI, c.D.nJ, c.D.nK)) return rc;
        const Dims& D = c.D;
        hipStream_t s = c.stream;
        const int g = fgrid(D);
        // A singular (pure-Neumann) system: b is made compatible by removing its mean, and
        // the mean of x, which A does not see, is removed too.  Near a steady state the
        // correction is small against its free constant, and x^T A x of the undemeaned FP32
        // pair is then lost to cancellation.
        double shift = 0., xshift = 0.;
        if (c.singularP) {
            auto liveMean = [&](const void* v) {
                hipLaunchKernelGGL(k_sum_live, dim3(g), dim3(FB), 0, s, D, (const double*)v, (const int8_t*)c.st, n.part);
                hipLaunchKernelGGL(k_final2, dim3(1), dim3(FB), 0, s, n.part, g, n.scal);
                HIPCHECK(hipMemcpyAsync(n.hscal, n.scal, 2 * sizeof(double), hipMemcpyDeviceToHost, s));
                HIPCHECK(hipStreamSynchronize(s));
                return n.hscal[1] > 0. ? n.hscal[0] / n.hscal[1] : 0.;
            };
            shift = liveMean(c.bp);
            xshift = liveMean(c.xp);
        }
        hipLaunchKernelGGL(k_sumsq, dim3(g), dim3(FB), 0, s, D, (const double*)c.bp, (const int8_t*)c.st, shift, n.part);
        hipLaunchKernelGGL(k_l_final, dim3(1), dim3(FB), 0, s, n.part, g, n.scal);
        HIPCHECK(hipMemcpyAsync(n.hscal, n.scal, sizeof(double), hipMemcpyDeviceToHost, s));
        HIPCHECK(hipStreamSynchronize(s));
        const double bn = std::sqrt(n.hscal[0]);
        if (!(bn > 0.) || !std::isfinite(bn)) { n.err = "zero or non-finite right-hand side"; return CAFFE_ERR_STATE; }
        Sample smp;
        smp.D = D;
        HIPCHECK(hipMalloc(&smp.A, 4 * (size_t)D.P * sizeof(float)));
        HIPCHECK(hipMalloc(&smp.b, (size_t)D.P * sizeof(float)));
        HIPCHECK(hipMalloc(&smp.x, (size_t)D.P * sizeof(float)));
        hipLaunchKernelGGL(k_collect, dim3((D.P + 255) / 256), dim3(256), 0, s, D, (const double*)c.Ap, (const double*)c.bp,
                           (const double*)c.xp, (const int8_t*)c.st, shift, xshift, 1.0 / bn, smp.A, smp.b, smp.x);
        HIPCHECK(hipStreamSynchronize(s));
        tg_ensure_fine(n, D);
        smp.nrm = sampleNorm(n, n.stream, smp);
        // the assembled p' matrix has aP = -(sum of the face coefficients) (negative
        // semi-definite); the pair is stored as (-A, -b), which has the same solution and
        // a positive energy norm.  The cycle is unchanged by negating A and r together,
        // so the solver applies it to the assembled system as it stands.
        if (smp.nrm < 0.) {
            hipLaunchKernelGGL(k_negate, dim3((4 * D.P + 255) / 256), dim3(256), 0, n.stream, smp.A, 4 * D.P);
            hipLaunchKernelGGL(k_negate, dim3((D.P + 255) / 256), dim3(256), 0, n.stream, smp.b, D.P);
            smp.nrm = -smp.nrm;
        }
        if (!(smp.nrm > 0.) || !std::isfinite(smp.nrm)) { freeSample(smp); n.err = "x^T A x vanishes"; return CAFFE_ERR_STATE; }
        n.data.push_back(smp);
        if (count) *count = (long)n.data.size();
        return CAFFE_OK;
    });
}

int caffe_net_clear(caffe_net* x) {
    return nguard(x, [&](Net& n) {
        HIPCHECK(hipStreamSynchronize(n.stream));
        for (auto& smp : n.data) freeSample(smp);
        n.data.clear();
        return CAFFE_OK;
    });
}

long caffe_net_dataset_size(caffe_net* x) { return x ? (long)x->n.data.size() : -1; }

int caffe_net_train(caffe_net* x, int steps, int batch, float lr, double synthetic, unsigned long long seed,
                    float* losses) {
    return nguard(x, [&](Net& n) {
        if (steps < 1 || batch < 1 || !(lr > 0.f) || !(synthetic >= 0. && synthetic <= 1.)) {
            n.err = "bad training arguments";
            return CAFFE_ERR_ARG;
        }
        if (n.data.empty()) { n.err = "empty training set"; return CAFFE_ERR_STATE; }
        const Dims D = n.data[0].D;
        for (auto& smp : n.data)
            if (smp.D.nI != D.nI || smp.D.nJ != D.nJ || smp.D.nK != D.nK) { n.err = "training samples differ in size"; return CAFFE_ERR_STATE; }
        hipStream_t s = n.stream;
        tg_ensure_fine(n, D);
        net_set_dims(n, D.nI / 2, D.nJ / 2, D.nK / 2, true);
        if ((long)steps > n.lossCap) {
            n.losses = (float*)net_alloc(n, (size_t)steps * sizeof(float), n.allocs);
            n.lossCap = steps;
        }
        HIPCHECK(hipMemsetAsync(n.losses, 0, (size_t)steps * sizeof(float), s));
        Rng rng{seed + 0x51ed2701ULL};
        const size_t ns = n.data.size();
        for (int t = 0; t < steps; ++t) {
            HIPCHECK(hipMemsetAsync(n.grad, 0, n.nParams * 4, s));
            for (int b = 0; b < batch; ++b) {
                const Sample& smp = n.data[rng.next() % ns];
                if (rng.uni() < synthetic) tg_loss_synthetic(n, s, smp, drawModes(rng, D), true, n.losses + t);
                else tg_loss(n, s, smp, true, n.losses + t);
            }
            net_adam(n, s, lr, 1.0f / batch);
        }
        std::vector<float> h(steps);
        HIPCHECK(hipMemcpyAsync(h.data(), n.losses, (size_t)steps * 4, hipMemcpyDeviceToHost, s));
        HIPCHECK(hipStreamSynchronize(s));
        for (int t = 0; t < steps; ++t) {
            h[t] /= (float)batch;
            if (losses) losses[t] = h[t];
        }
        if (!std::isfinite(h[steps - 1])) { n.err = "training diverged (non-finite loss)"; return CAFFE_ERR_NOT_CONVERGED; }
        return CAFFE_OK;
    });
}

int caffe_net_eval(caffe_net* x, float* losses, long count) {
    return nguard(x, [&](Net& n) {
        if (!losses || count != (long)n.data.size()) { n.err = "loss buffer size mismatch"; return CAFFE_ERR_ARG; }
        if (n.data.empty()) return CAFFE_OK;
        hipStream_t s = n.stream;
        const Dims D = n.data[0].D;
        tg_ensure_fine(n, D);
        net_set_dims(n, D.nI / 2, D.nJ / 2, D.nK / 2, n.trainBufs);
        float* ld = (float*)(n.scal + 8);
        for (long i = 0; i < count; ++i) {
            HIPCHECK(hipMemsetAsync(ld, 0, sizeof(float), s));
            tg_loss(n, s, n.data[i], false, ld);
            HIPCHECK(hipMemcpyAsync(losses + i, ld, sizeof(float), hipMemcpyDeviceToHost, s));
        }
        HIPCHECK(hipStreamSynchronize(s));
        return CAFFE_OK;
    });
}

int caffe_ctx_attach_net(caffe_ctx* cx, caffe_net* x) {
    if (!cx) return CAFFE_ERR_ARG;
    Ctx& c = cx->c;
    if (x && x->n.device != c.device) { c.err = "network and context are on different devices"; return CAFFE_ERR_ARG; }
    c.net = x ? (void*)&x->n : nullptr;
    return CAFFE_OK;
}

}  // extern "C"
