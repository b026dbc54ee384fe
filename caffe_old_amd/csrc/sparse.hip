// sparse.hip -- the SparseMatrix / SparseVector compatibility path of the C ABI
// (include/caffe_hip.h, caffe_matrix_* and caffe_vector_*).
//
// The reference's solvers build their systems row by row through SparseMatrix::setRow and
// SparseVector::setValue and solve them with SparseMatrix::solve (src/Math/SparseMatrix.h:34-81,
// SparseMatrix.cpp:47-106, SparseVector.h:8-43).  Here a matrix keeps its rows on the host
// while they are set, with PETSc's MatSetValues(INSERT_VALUES) semantics: an entry set again
// is overwritten, negative column (and row) indices are ignored, and nothing is zeroed
// between assemblies.  Assembly packs the rows into an ELL copy in HBM, column-major (entry q
// of row r at q * m + r, so a wave's 64 rows read 64 consecutive values), padded with zero
// entries on the row's own column.  solve() runs BiCGStab in FP64, right-preconditioned by
// Jacobi, from x = 0 to a true relative residual rtol (restarted from the true residual on
// breakdown), and returns the iteration count; non-convergence is not an error, as in the
// reference (SparseMatrix.cpp:86-106 never checks the converged reason).
//
// Deterministic reductions: each kernel writes one row of block partial sums, the next kernel
// (or a one-block finalize) adds them in block order (readPartials in common.hpp).
#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>
#include <string>
#include <utility>
#include <vector>

#include "internal.hpp"

using namespace caffe;

namespace {

constexpr int SBS = 256;          // rows per block step
constexpr int SGRID = 1024;       // at most MAX_RED_BLOCKS blocks
static_assert(SGRID <= MAX_RED_BLOCKS, "partial rows");

// scalar slots of one solve
// (S_DONE is written by block 0 of k_sp_a and k_sp_b, whose other blocks compute the same
// value from the same partials, so a block that already sees it skips nothing it would have
// done; k_sp_c flags a zero omega in S_BRKC instead, since every block must still update)
enum { S_RHO = 0, S_RHOPREV, S_ALPHA, S_OMEGA, S_BB, S_RR, S_CIT, S_ITS, S_DONE, S_TOL2, S_ATOL2, S_MAXIT, S_BRKC,
       S_RES0, S_RES1, S_RES2, N_SSLOTS = 32 };
constexpr int N_PART_ROWS = 3;    // three partial buffers (ping-pong between the kernels)

// work of one solve: 8 vectors of m doubles, the Jacobi weights, three partial buffers and the
// scalar slots (in doubles)
inline size_t solveWorkDoubles(long m) { return 9 * (size_t)m + (size_t)N_PART_ROWS * MAX_RED_BLOCKS * RED_STRIDE + N_SSLOTS; }

struct Ell {
    int m, W;
    const int* __restrict__ cols;
    const double* __restrict__ vals;
    template <class F>
    __device__ __forceinline__ double apply(int r, F f) const {
        double s = 0.;
        for (int q = 0; q < W; ++q) {
            const long e = (long)q * m + r;
            s += vals[e] * f(cols[e]);
        }
        return s;
    }
};

// rows of block b: one contiguous chunk per block
#define FOR_ROWS(m, r)                                                                            \
    for (int r##c_ = ((m) + (int)gridDim.x - 1) / (int)gridDim.x, r = (int)blockIdx.x * r##c_ + (int)threadIdx.x, \
             r##e_ = min((m), (int)blockIdx.x * r##c_ + r##c_);                                   \
         r < r##e_; r += blockDim.x)

// ---------------------------------------------------------------- BiCGStab kernels
// (re)start from the iterate x: r = b - A x, rh = r, p = v = 0, Jacobi weights; partials
// (rh, r), (r, r), (b, b)
__global__ void k_sp_init(Ell A, const double* b, const double* x, double* r, double* rh, double* p, double* v,
                          double* dinv, double* part, double* st) {
    double acc[3] = {0., 0., 0.};
    FOR_ROWS(A.m, i) {
        double d = 0.;
        for (int q = 0; q < A.W; ++q) {
            const long e = (long)q * A.m + i;
            if (A.cols[e] == i) d += A.vals[e];
        }
        dinv[i] = d != 0. ? 1. / d : 1.;
        const double bv = b[i];
        const double rv = bv - A.apply(i, [&](int c) { return x[c]; });
        r[i] = rv; rh[i] = rv; p[i] = 0.; v[i] = 0.;
        acc[0] += rv * rv; acc[1] += rv * rv; acc[2] += bv * bv;
    }
    writePartials<SBS, 3>(acc, part);
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        st[S_CIT] = 0.; st[S_DONE] = 0.; st[S_BRKC] = 0.; st[S_RHOPREV] = 1.; st[S_ALPHA] = 1.; st[S_OMEGA] = 1.;
    }
}

// KA: the convergence test on (rh, r), (r, r); beta; p_new = r + beta (p - omega v) and
// v_new = A D^-1 p_new with p_new formed at each column; partial (rh, v_new)
__global__ void k_sp_a(Ell A, const double* __restrict__ r, const double* __restrict__ rh, const double* __restrict__ pold,
                       double* __restrict__ pnew, const double* __restrict__ vold, double* __restrict__ vnew,
                       const double* __restrict__ dinv, const double* partIn, int nb, double* partOut, double* st) {
    if (st[S_DONE] != 0.) return;
    const double cit = st[S_CIT], its = st[S_ITS], om = st[S_OMEGA], rhoPrev = st[S_RHOPREV], al = st[S_ALPHA];
    const double tol2 = st[S_TOL2], atol2 = st[S_ATOL2], maxit = st[S_MAXIT], bbSt = st[S_BB], brkc = st[S_BRKC];
    double t[3];
    readPartials<SBS, 3>(partIn, nb, t);
    const double rho = t[0], rr = t[1];
    const double bb = cit == 0. ? t[2] : bbSt;
    double done = 0.;
    if (rr <= tol2 * bb || rr <= atol2) done = 1.;
    else if (its >= maxit) done = 3.;
    else if (brkc != 0. || rho == 0. || !(rho == rho)) done = 2.;
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        st[S_RHO] = rho; st[S_RR] = rr; st[S_DONE] = done;
        if (cit == 0.) st[S_BB] = bb;
    }
    if (done != 0.) return;
    const double beta = cit == 0. ? 0. : (rho / rhoPrev) * (al / om);
    double acc[1] = {0.};
    FOR_ROWS(A.m, i) {
        auto pf = [&](int c) { return r[c] + beta * (pold[c] - om * vold[c]); };
        const double pn = pf(i);
        const double vn = A.apply(i, [&](int c) { return dinv[c] * pf(c); });
        pnew[i] = pn; vnew[i] = vn;
        acc[0] += rh[i] * vn;
    }
    writePartials<SBS, 1>(acc, partOut);
}

// KB: alpha = rho / (rh, v); s = r - alpha v and t = A D^-1 s with s formed at each column;
// partials (t, s), (t, t)
__global__ void k_sp_b(Ell A, const double* __restrict__ r, const double* __restrict__ v, double* __restrict__ s,
                       double* __restrict__ tv, const double* __restrict__ dinv, const double* partIn, int nb, double* partOut,
                       double* st) {
    if (st[S_DONE] != 0.) return;
    const double rho = st[S_RHO];
    double t[1];
    readPartials<SBS, 1>(partIn, nb, t);
    const bool brk = t[0] == 0. || !(t[0] == t[0]);
    const double al = brk ? 0. : rho / t[0];
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        st[S_ALPHA] = al;
        if (brk) st[S_DONE] = 2.;
    }
    if (brk) return;
    double acc[2] = {0., 0.};
    FOR_ROWS(A.m, i) {
        auto sf = [&](int c) { return r[c] - al * v[c]; };
        const double sv = sf(i);
        const double tt = A.apply(i, [&](int c) { return dinv[c] * sf(c); });
        s[i] = sv; tv[i] = tt;
        acc[0] += tt * sv; acc[1] += tt * tt;
    }
    writePartials<SBS, 2>(acc, partOut);
}

// KC: omega = (t, s) / (t, t); x += alpha D^-1 p + omega D^-1 s; r = s - omega t;
// partials (rh, r), (r, r), 0
__global__ void k_sp_c(int m, double* __restrict__ x, double* __restrict__ r, const double* __restrict__ rh,
                       const double* __restrict__ p, const double* __restrict__ s, const double* __restrict__ tv,
                       const double* __restrict__ dinv, const double* partIn, int nb, double* partOut, double* st) {
    if (st[S_DONE] != 0.) return;
    const double al = st[S_ALPHA], rho = st[S_RHO], cit = st[S_CIT], its = st[S_ITS];
    double t[2];
    readPartials<SBS, 2>(partIn, nb, t);
    const double om = t[1] > 0. ? t[0] / t[1] : 0.;
    const bool ok = om == om;
    const double omu = ok ? om : 0.;
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        st[S_OMEGA] = omu; st[S_RHOPREV] = rho; st[S_CIT] = cit + 1.; st[S_ITS] = its + 1.;
        if (omu == 0.) st[S_BRKC] = 1.;   // x and r are still updated (omega = 0 is a valid step)
    }
    double acc[3] = {0., 0., 0.};
    FOR_ROWS(m, i) {
        const double di = dinv[i];
        x[i] += al * (di * p[i]) + omu * (di * s[i]);
        const double rn = s[i] - omu * tv[i];
        r[i] = rn;
        acc[0] += rh[i] * rn; acc[1] += rn * rn;
    }
    writePartials<SBS, 3>(acc, partOut);
}

// true residual: partials (r, r), (b, b)
__global__ void k_sp_resid(Ell A, const double* b, const double* x, double* part) {
    double acc[2] = {0., 0.};
    FOR_ROWS(A.m, i) {
        const double rv = b[i] - A.apply(i, [&](int c) { return x[c]; });
        acc[0] += rv * rv; acc[1] += b[i] * b[i];
    }
    writePartials<SBS, 2>(acc, part);
}

template <int K>
__global__ void k_sp_finalize(const double* part, int nb, double* out) {
    double t[K];
    readPartials<SBS, K>(part, nb, t);
    if (threadIdx.x == 0)
        for (int q = 0; q < K; ++q) out[q] = t[q];
}

__global__ void k_sp_zero(double* a, long n) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) a[i] = 0.;
}
__global__ void k_sp_setscal(double* st, double tol2, double atol2, double maxit, double its) {
    if (threadIdx.x == 0) { st[S_TOL2] = tol2; st[S_ATOL2] = atol2; st[S_MAXIT] = maxit; st[S_ITS] = its; }
}

// ---------------------------------------------------------------- products and vectors
// y = A x (+ x2): MatMult / MatMultAdd (SparseMatrix.cpp:135-143)
__global__ void k_sp_mult(Ell A, const double* x, const double* x2, double* y) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= A.m) return;
    const double s = A.apply(i, [&](int c) { return x[c]; });
    y[i] = x2 ? s + x2[i] : s;
}
__global__ void k_sp_scale(double* a, double s, long n) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) a[i] *= s;
}
// partials sum |v|, sum v^2 per block; max |v| in row slot 2 (folded by k_sp_norm_final)
__global__ void k_sp_norm(const double* v, int n, double* part) {
    __shared__ double lds[SBS / 64];
    double a1 = 0., a2 = 0., mx = 0.;
    FOR_ROWS(n, i) {
        const double x = v[i];
        a1 += fabs(x); a2 += x * x; mx = fmax(mx, fabs(x));
    }
    const double s1 = blockSum<SBS>(a1, lds);
    const double s2 = blockSum<SBS>(a2, lds);
    const double m = blockMax<SBS>(mx, lds);
    if (threadIdx.x == 0) {
        part[blockIdx.x * RED_STRIDE + 0] = s1;
        part[blockIdx.x * RED_STRIDE + 1] = s2;
        part[blockIdx.x * RED_STRIDE + 2] = m;
    }
}
__global__ void k_sp_norm_final(const double* part, int nb, double* out) {
    if (threadIdx.x != 0) return;
    double s1 = 0., s2 = 0., m = 0.;
    for (int b = 0; b < nb; ++b) {
        s1 += part[b * RED_STRIDE]; s2 += part[b * RED_STRIDE + 1]; m = fmax(m, part[b * RED_STRIDE + 2]);
    }
    out[0] = s1; out[1] = std::sqrt(s2); out[2] = m;
}

int gridRows(long m) { return blocksFor(m, SBS, SGRID); }

constexpr long NOALLOC_MAX_ROWS = CAFFE_NOALLOC_MAX_ROWS;
constexpr size_t NOALLOC_DOUBLES = 9 * (size_t)NOALLOC_MAX_ROWS + (size_t)N_PART_ROWS * MAX_RED_BLOCKS * RED_STRIDE + N_SSLOTS;

}  // namespace

// The scratch of caffe_matrix_solve_noalloc: the BiCGStab work (vectors, partial sums and
// scalar slots) of systems up to CAFFE_NOALLOC_MAX_ROWS rows, 19.3 MB, held by every GPU the
// library runs on until the process ends.  Host code never uses this name as an address: it
// uses the address hipGetSymbolAddress returns for the device (g_scratch below).
static __device__ __align__(256) double scratch[NOALLOC_DOUBLES];

// ---------------------------------------------------------------- host objects
struct caffe_vector {
    int device = 0;
    hipStream_t stream = nullptr;
    std::string err;
    int m = 0;
    double* dev = nullptr;
    double* part = nullptr;       // norm partials and results
    double* hout = nullptr;       // pinned read-back (3 doubles)
    std::vector<double> host;
    int newer = 0;                // 0: copies agree, 1: host newer, 2: device newer
};

struct caffe_matrix {
    int device = 0;
    hipStream_t stream = nullptr;
    std::string err;
    int m = 0, n = 0, nnzPre = 0;
    double rtol = 1e-5, abstol = 1e-50;
    int maxit = 10000;
    std::vector<std::vector<std::pair<int, double>>> rows;   // sorted by column
    bool dirty = true;
    int W = 0, Wcap = 0;
    int* dcols = nullptr;
    double* dvals = nullptr;
    double* work = nullptr;       // solveWorkDoubles(m), allocated at the first caffe_matrix_solve
    double* hst = nullptr;        // pinned mirror of the scalar slots
    hipEvent_t ev0 = nullptr, ev1 = nullptr;   // solve timing
    caffe_solve_report last{};
};

namespace {

template <class O, class F>
int sguard(O* o, F f) {
    if (!o) return CAFFE_ERR_ARG;
    try {
        HIPCHECK(hipSetDevice(o->device));
        return f(*o);
    } catch (const std::string& e) {
        o->err = e;
        return CAFFE_ERR_HIP;
    } catch (const std::exception& e) {
        o->err = e.what();
        return CAFFE_ERR_HIP;
    }
}

void vecToDevice(caffe_vector& v) {
    if (v.newer != 1) return;
    HIPCHECK(hipMemcpyAsync(v.dev, v.host.data(), sizeof(double) * v.m, hipMemcpyHostToDevice, v.stream));
    HIPCHECK(hipStreamSynchronize(v.stream));
    v.newer = 0;
}
void vecToHost(caffe_vector& v) {
    if (v.newer != 2) return;
    HIPCHECK(hipMemcpyAsync(v.host.data(), v.dev, sizeof(double) * v.m, hipMemcpyDeviceToHost, v.stream));
    HIPCHECK(hipStreamSynchronize(v.stream));
    v.newer = 0;
}

Ell ellOf(const caffe_matrix& A) { return Ell{A.m, A.W, A.dcols, A.dvals}; }

// MatAssemblyBegin/End (SparseMatrix.cpp:90-93, 108-112): pack the rows into the ELL copy
void assemble(caffe_matrix& A) {
    if (!A.dirty) return;
    int W = 1;
    for (const auto& r : A.rows) W = std::max(W, (int)r.size());
    if (W > A.Wcap) {
        if (A.dcols) { (void)hipFree(A.dcols); A.dcols = nullptr; }
        if (A.dvals) { (void)hipFree(A.dvals); A.dvals = nullptr; }
        HIPCHECK(hipMalloc((void**)&A.dcols, sizeof(int) * (size_t)W * A.m));
        HIPCHECK(hipMalloc((void**)&A.dvals, sizeof(double) * (size_t)W * A.m));
        A.Wcap = W;
    }
    A.W = W;
    std::vector<int> c((size_t)W * A.m);
    std::vector<double> v((size_t)W * A.m);
    for (int r = 0; r < A.m; ++r) {
        const auto& row = A.rows[r];
        for (int q = 0; q < W; ++q) {
            const size_t e = (size_t)q * A.m + r;
            if (q < (int)row.size()) { c[e] = row[q].first; v[e] = row[q].second; }
            else { c[e] = r < A.n ? r : 0; v[e] = 0.; }
        }
    }
    HIPCHECK(hipMemcpyAsync(A.dcols, c.data(), sizeof(int) * c.size(), hipMemcpyHostToDevice, A.stream));
    HIPCHECK(hipMemcpyAsync(A.dvals, v.data(), sizeof(double) * v.size(), hipMemcpyHostToDevice, A.stream));
    HIPCHECK(hipStreamSynchronize(A.stream));
    A.dirty = false;
}

// MatSetValues(INSERT_VALUES) on one row; negative columns ignored
int setRow(caffe_matrix& A, int row, int nc, const int* cols, const double* vals) {
    if (row < 0) return CAFFE_OK;   // MatSetValues ignores negative row indices
    if (row >= A.m) { A.err = "row index out of range"; return CAFFE_ERR_ARG; }
    auto& R = A.rows[row];
    for (int q = 0; q < nc; ++q) {
        const int c = cols[q];
        if (c < 0) continue;
        if (c >= A.n) { A.err = "column index out of range"; return CAFFE_ERR_ARG; }
        auto it = std::lower_bound(R.begin(), R.end(), c, [](const std::pair<int, double>& e, int k) { return e.first < k; });
        if (it != R.end() && it->first == c) it->second = vals[q];
        else R.insert(it, std::make_pair(c, vals[q]));
    }
    A.dirty = true;
    return CAFFE_OK;
}

constexpr int MAX_DEVICES = 64;
std::mutex g_scratchMu;
double* g_scratch[MAX_DEVICES] = {nullptr};

// the device address of scratch, once per device, with that device current
void initScratch(int device) {
    if (device < 0 || device >= MAX_DEVICES) throw std::string("device index out of range");
    std::lock_guard<std::mutex> lk(g_scratchMu);
    if (g_scratch[device]) return;
    void* p = nullptr;
    HIPCHECK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    g_scratch[device] = (double*)p;
}

// BiCGStab on the assembled A from x = 0 into x, work in w (solveWorkDoubles(m)).  Every
// value the kernels read is written first in this call (k_sp_init and the fill below), so
// the work buffer may hold anything on entry.
int runSolve(caffe_matrix& A, const double* b, double* x, double* w, caffe_solve_report* rep) {
    const int m = A.m;
    Ell E = ellOf(A);
    double* r = w; double* rh = r + m; double* p0 = rh + m; double* p1 = p0 + m; double* v0 = p1 + m;
    double* v1 = v0 + m; double* s = v1 + m; double* t = s + m; double* dinv = t + m;
    double* partC = dinv + m;
    double* partA = partC + MAX_RED_BLOCKS * RED_STRIDE;
    double* partB = partA + MAX_RED_BLOCKS * RED_STRIDE;
    double* st = partB + MAX_RED_BLOCKS * RED_STRIDE;
    const int g = gridRows(m);
    hipLaunchKernelGGL(k_sp_zero, dim3(blocksFor(m, 256)), dim3(256), 0, A.stream, x, (long)m);
    HIPCHECK(hipEventRecord(A.ev0, A.stream));
    // the recurrence aims at half the tolerance; the true residual decides
    const double rt = A.rtol, tol2 = 0.25 * rt * rt, atol2 = A.abstol * A.abstol;
    int its = 0, cycles = 0;
    double relres = 1., best = 2.;
    bool conv = false;
    while (true) {
        hipLaunchKernelGGL(k_sp_setscal, dim3(1), dim3(64), 0, A.stream, st, tol2, atol2, (double)A.maxit, (double)its);
        hipLaunchKernelGGL(k_sp_init, dim3(g), dim3(SBS), 0, A.stream, E, b, (const double*)x, r, rh, p0, v0, dinv, partC, st);
        double *pc = p0, *pn = p1, *vc = v0, *vn = v1;
        int chunk = 4;
        double done = 0.;
        while (true) {
            for (int q = 0; q < chunk; ++q) {
                hipLaunchKernelGGL(k_sp_a, dim3(g), dim3(SBS), 0, A.stream, E, (const double*)r, (const double*)rh,
                                   (const double*)pc, pn, (const double*)vc, vn, (const double*)dinv, (const double*)partC, g,
                                   partA, st);
                hipLaunchKernelGGL(k_sp_b, dim3(g), dim3(SBS), 0, A.stream, E, (const double*)r, (const double*)vn, s, t,
                                   (const double*)dinv, (const double*)partA, g, partB, st);
                hipLaunchKernelGGL(k_sp_c, dim3(g), dim3(SBS), 0, A.stream, m, x, r, (const double*)rh, (const double*)pn,
                                   (const double*)s, (const double*)t, (const double*)dinv, (const double*)partB, g, partC, st);
                std::swap(pc, pn); std::swap(vc, vn);
            }
            HIPCHECK(hipMemcpyAsync(A.hst, st, sizeof(double) * N_SSLOTS, hipMemcpyDeviceToHost, A.stream));
            HIPCHECK(hipStreamSynchronize(A.stream));
            done = A.hst[S_DONE];
            its = (int)A.hst[S_ITS];
            if (done != 0.) break;
            if (chunk < 32) chunk *= 2;
        }
        ++cycles;
        hipLaunchKernelGGL(k_sp_resid, dim3(g), dim3(SBS), 0, A.stream, E, b, (const double*)x, partA);
        hipLaunchKernelGGL(k_sp_finalize<2>, dim3(1), dim3(SBS), 0, A.stream, (const double*)partA, g, st + S_RES0);
        HIPCHECK(hipMemcpyAsync(A.hst, st, sizeof(double) * N_SSLOTS, hipMemcpyDeviceToHost, A.stream));
        HIPCHECK(hipStreamSynchronize(A.stream));
        const double rr = A.hst[S_RES0], bb = A.hst[S_RES1];
        relres = bb > 0. ? std::sqrt(rr / bb) : std::sqrt(rr);
        conv = relres <= rt || std::sqrt(rr) <= A.abstol;
        // restart from the true residual while it improves and iterations remain
        if (conv || its >= A.maxit || !(relres < best) || cycles >= 100) break;
        best = relres;
    }
    HIPCHECK(hipEventRecord(A.ev1, A.stream));
    HIPCHECK(hipEventSynchronize(A.ev1));
    float ms = 0.f;
    HIPCHECK(hipEventElapsedTime(&ms, A.ev0, A.ev1));
    A.last = caffe_solve_report{its, cycles, relres, conv ? 1 : 0, 0, (double)ms};
    if (rep) *rep = A.last;
    return its;
}

int checkSolveArgs(caffe_matrix& A, const caffe_vector* b, const caffe_vector* x) {
    if (!b || !x) { A.err = "null vector"; return CAFFE_ERR_ARG; }
    if (A.m != A.n) { A.err = "solve needs a square matrix"; return CAFFE_ERR_ARG; }
    if (b->m != A.m || x->m != A.m) { A.err = "vector length does not match the matrix"; return CAFFE_ERR_ARG; }
    if (b->device != A.device || x->device != A.device) { A.err = "matrix and vectors are on different devices"; return CAFFE_ERR_ARG; }
    if (b == x) { A.err = "b and x must be different vectors"; return CAFFE_ERR_ARG; }
    return CAFFE_OK;
}

}  // namespace

extern "C" {

int caffe_matrix_create(int device, int m, int n, int nnz, double rtol, double abstol, int maxit, caffe_matrix** out) {
    if (!out) return CAFFE_ERR_ARG;
    *out = nullptr;
    if (m < 1 || n < 1 || nnz < 0) return CAFFE_ERR_ARG;
    caffe_matrix* A = new caffe_matrix();
    A->device = device;
    A->m = m; A->n = n; A->nnzPre = nnz;
    if (rtol > 0.) A->rtol = rtol;      // PETSC_DEFAULT (a negative value) keeps the default
    if (abstol >= 0.) A->abstol = abstol;
    if (maxit > 0) A->maxit = maxit;
    int rc = sguard(A, [&](caffe_matrix& a) {
        HIPCHECK(hipStreamCreateWithFlags(&a.stream, hipStreamNonBlocking));
        HIPCHECK(hipHostMalloc((void**)&a.hst, sizeof(double) * N_SSLOTS, hipHostMallocDefault));
        HIPCHECK(hipEventCreate(&a.ev0));
        HIPCHECK(hipEventCreate(&a.ev1));
        a.rows.assign(m, {});
        for (auto& r : a.rows) r.reserve(nnz);
        initScratch(device);
        return CAFFE_OK;
    });
    if (rc != CAFFE_OK) {
        caffe_matrix_destroy(A);
        return rc;
    }
    *out = A;
    return CAFFE_OK;
}

int caffe_matrix_destroy(caffe_matrix* A) {
    if (!A) return CAFFE_ERR_ARG;
    (void)hipSetDevice(A->device);
    if (A->stream) (void)hipStreamSynchronize(A->stream);
    if (A->dcols) (void)hipFree(A->dcols);
    if (A->dvals) (void)hipFree(A->dvals);
    if (A->work) (void)hipFree(A->work);
    if (A->hst) (void)hipHostFree(A->hst);
    if (A->ev0) (void)hipEventDestroy(A->ev0);
    if (A->ev1) (void)hipEventDestroy(A->ev1);
    if (A->stream) (void)hipStreamDestroy(A->stream);
    delete A;
    return CAFFE_OK;
}

const char* caffe_matrix_last_error(caffe_matrix* A) { return A ? A->err.c_str() : "null matrix"; }

int caffe_matrix_set_rows(caffe_matrix* A, int nrows, const int* rows, const int* rowptr, const int* cols,
                          const double* vals) {
    return sguard(A, [&](caffe_matrix& a) {
        if (nrows < 0 || (nrows > 0 && (!rows || !rowptr || !cols || !vals))) { a.err = "null row arguments"; return CAFFE_ERR_ARG; }
        for (int q = 0; q < nrows; ++q) {
            const int nc = rowptr[q + 1] - rowptr[q];
            if (nc < 0) { a.err = "row pointers must not decrease"; return CAFFE_ERR_ARG; }
            const int rc = setRow(a, rows[q], nc, cols + rowptr[q], vals + rowptr[q]);
            if (rc != CAFFE_OK) return rc;
        }
        return CAFFE_OK;
    });
}

int caffe_matrix_set_value(caffe_matrix* A, int i, int j, double v) {
    return sguard(A, [&](caffe_matrix& a) { return setRow(a, i, 1, &j, &v); });
}

int caffe_matrix_zero_entries(caffe_matrix* A) {
    // MatZeroEntries: the values become zero, the nonzero structure stays
    return sguard(A, [&](caffe_matrix& a) {
        for (auto& r : a.rows)
            for (auto& e : r) e.second = 0.;
        a.dirty = true;
        return CAFFE_OK;
    });
}

int caffe_matrix_assemble(caffe_matrix* A) {
    return sguard(A, [&](caffe_matrix& a) { assemble(a); return CAFFE_OK; });
}

int caffe_matrix_solve(caffe_matrix* A, const caffe_vector* b, caffe_vector* x, int* iters) {
    return sguard(A, [&](caffe_matrix& a) {
        int rc = checkSolveArgs(a, b, x);
        if (rc != CAFFE_OK) return rc;
        assemble(a);
        vecToDevice(*const_cast<caffe_vector*>(b));
        if (!a.work) HIPCHECK(hipMalloc((void**)&a.work, sizeof(double) * solveWorkDoubles(a.m)));
        const int its = runSolve(a, b->dev, x->dev, a.work, nullptr);
        x->newer = 2;
        if (iters) *iters = its;
        return CAFFE_OK;
    });
}

int caffe_matrix_solve_noalloc(caffe_matrix* A, const caffe_vector* b, caffe_vector* x, int* iters) {
    // no hipMalloc, hipFree or allocating routine: the work is scratch (its address taken at
    // caffe_matrix_create), the matrix must be assembled, and b reaches HBM by a copy into its
    // existing buffer
    return sguard(A, [&](caffe_matrix& a) {
        int rc = checkSolveArgs(a, b, x);
        if (rc != CAFFE_OK) return rc;
        if (a.m > NOALLOC_MAX_ROWS) {
            a.err = "caffe_matrix_solve_noalloc: " + std::to_string(a.m) + " rows, more than CAFFE_NOALLOC_MAX_ROWS (" +
                    std::to_string(NOALLOC_MAX_ROWS) + ")";
            return CAFFE_ERR_ARG;
        }
        if (a.dirty) {
            a.err = "caffe_matrix_solve_noalloc: rows changed since the last caffe_matrix_assemble (assembly may allocate)";
            return CAFFE_ERR_STATE;
        }
        double* w = g_scratch[a.device];
        if (!w) { a.err = "caffe_matrix_solve_noalloc: scratch address not initialised for this device"; return CAFFE_ERR_STATE; }
        vecToDevice(*const_cast<caffe_vector*>(b));
        const int its = runSolve(a, b->dev, x->dev, w, nullptr);
        x->newer = 2;
        if (iters) *iters = its;
        return CAFFE_OK;
    });
}

int caffe_matrix_last_report(caffe_matrix* A, caffe_solve_report* rep) {
    return sguard(A, [&](caffe_matrix& a) { if (!rep) return CAFFE_ERR_ARG; *rep = a.last; return CAFFE_OK; });
}

int caffe_matrix_multiply(caffe_matrix* A, const caffe_vector* x, caffe_vector* y) {
    return caffe_matrix_multiply_add(A, x, nullptr, y);
}

int caffe_matrix_multiply_add(caffe_matrix* A, const caffe_vector* x1, const caffe_vector* x2, caffe_vector* y) {
    return sguard(A, [&](caffe_matrix& a) {
        if (!x1 || !y) { a.err = "null vector"; return CAFFE_ERR_ARG; }
        if (x1->m != a.n || y->m != a.m || (x2 && x2->m != a.m)) { a.err = "vector length does not match the matrix"; return CAFFE_ERR_ARG; }
        if (y == x1) { a.err = "y must differ from x1"; return CAFFE_ERR_ARG; }
        assemble(a);
        vecToDevice(*const_cast<caffe_vector*>(x1));
        if (x2) vecToDevice(*const_cast<caffe_vector*>(x2));
        vecToDevice(*y);   // y may alias x2 (MatMultAdd allows it); its host edits land first
        hipLaunchKernelGGL(k_sp_mult, dim3(blocksFor(a.m, 256)), dim3(256), 0, a.stream, ellOf(a), (const double*)x1->dev,
                           (const double*)(x2 ? x2->dev : nullptr), y->dev);
        HIPCHECK(hipStreamSynchronize(a.stream));
        y->newer = 2;
        return CAFFE_OK;
    });
}

int caffe_matrix_scale(double alpha, caffe_matrix* A) {
    return sguard(A, [&](caffe_matrix& a) {
        for (auto& r : a.rows)
            for (auto& e : r) e.second *= alpha;
        a.dirty = true;
        return CAFFE_OK;
    });
}

int caffe_vector_create(int device, int m, caffe_vector** out) {
    if (!out) return CAFFE_ERR_ARG;
    *out = nullptr;
    if (m < 1) return CAFFE_ERR_ARG;
    caffe_vector* v = new caffe_vector();
    v->device = device;
    v->m = m;
    int rc = sguard(v, [&](caffe_vector& a) {
        HIPCHECK(hipStreamCreateWithFlags(&a.stream, hipStreamNonBlocking));
        HIPCHECK(hipMalloc((void**)&a.dev, sizeof(double) * (size_t)m));
        HIPCHECK(hipMalloc((void**)&a.part, sizeof(double) * ((size_t)MAX_RED_BLOCKS * RED_STRIDE + 8)));
        HIPCHECK(hipHostMalloc((void**)&a.hout, sizeof(double) * 4, hipHostMallocDefault));
        HIPCHECK(hipMemsetAsync(a.dev, 0, sizeof(double) * (size_t)m, a.stream));
        HIPCHECK(hipStreamSynchronize(a.stream));
        a.host.assign(m, 0.);
        a.newer = 0;
        return CAFFE_OK;
    });
    if (rc != CAFFE_OK) {
        caffe_vector_destroy(v);
        return rc;
    }
    *out = v;
    return CAFFE_OK;
}

int caffe_vector_destroy(caffe_vector* v) {
    if (!v) return CAFFE_ERR_ARG;
    (void)hipSetDevice(v->device);
    if (v->stream) (void)hipStreamSynchronize(v->stream);
    if (v->dev) (void)hipFree(v->dev);
    if (v->part) (void)hipFree(v->part);
    if (v->hout) (void)hipHostFree(v->hout);
    if (v->stream) (void)hipStreamDestroy(v->stream);
    delete v;
    return CAFFE_OK;
}

const char* caffe_vector_last_error(caffe_vector* v) { return v ? v->err.c_str() : "null vector"; }

int caffe_vector_set_values(caffe_vector* v, int n, const int* idx, const double* vals) {
    // VecSetValues(INSERT_VALUES): negative indices ignored
    return sguard(v, [&](caffe_vector& a) {
        if (n < 0 || (n > 0 && (!idx || !vals))) { a.err = "null arguments"; return CAFFE_ERR_ARG; }
        for (int q = 0; q < n; ++q)
            if (idx[q] >= a.m) { a.err = "index out of range"; return CAFFE_ERR_ARG; }
        vecToHost(a);
        for (int q = 0; q < n; ++q)
            if (idx[q] >= 0) a.host[idx[q]] = vals[q];
        if (n > 0) a.newer = 1;
        return CAFFE_OK;
    });
}

int caffe_vector_get_values(caffe_vector* v, int n, const int* idx, double* out) {
    // SparseVector::operator() (SparseVector.cpp:52-59, VecGetValues)
    return sguard(v, [&](caffe_vector& a) {
        if (n < 0 || (n > 0 && (!idx || !out))) { a.err = "null arguments"; return CAFFE_ERR_ARG; }
        for (int q = 0; q < n; ++q)
            if (idx[q] < 0 || idx[q] >= a.m) { a.err = "index out of range"; return CAFFE_ERR_ARG; }
        vecToHost(a);
        for (int q = 0; q < n; ++q) out[q] = a.host[idx[q]];
        return CAFFE_OK;
    });
}

int caffe_vector_zero_entries(caffe_vector* v) {
    return sguard(v, [&](caffe_vector& a) {
        std::fill(a.host.begin(), a.host.end(), 0.);
        hipLaunchKernelGGL(k_sp_zero, dim3(blocksFor(a.m, 256)), dim3(256), 0, a.stream, a.dev, (long)a.m);
        HIPCHECK(hipStreamSynchronize(a.stream));
        a.newer = 0;
        return CAFFE_OK;
    });
}

int caffe_vector_scale(double alpha, caffe_vector* v) {
    return sguard(v, [&](caffe_vector& a) {
        vecToDevice(a);
        hipLaunchKernelGGL(k_sp_scale, dim3(blocksFor(a.m, 256)), dim3(256), 0, a.stream, a.dev, alpha, (long)a.m);
        HIPCHECK(hipStreamSynchronize(a.stream));
        a.newer = 2;
        return CAFFE_OK;
    });
}

int caffe_vector_norm(caffe_vector* v, int type, double* out) {
    return sguard(v, [&](caffe_vector& a) {
        if (!out || type < CAFFE_NORM_L1 || type > CAFFE_NORM_INF) { a.err = "bad norm arguments"; return CAFFE_ERR_ARG; }
        vecToDevice(a);
        const int g = gridRows(a.m);
        hipLaunchKernelGGL(k_sp_norm, dim3(g), dim3(SBS), 0, a.stream, (const double*)a.dev, a.m, a.part);
        hipLaunchKernelGGL(k_sp_norm_final, dim3(1), dim3(64), 0, a.stream, (const double*)a.part, g,
                           a.part + (size_t)MAX_RED_BLOCKS * RED_STRIDE);
        HIPCHECK(hipMemcpyAsync(a.hout, a.part + (size_t)MAX_RED_BLOCKS * RED_STRIDE, sizeof(double) * 3,
                                hipMemcpyDeviceToHost, a.stream));
        HIPCHECK(hipStreamSynchronize(a.stream));
        *out = a.hout[type];
        return CAFFE_OK;
    });
}

}  // extern "C"
