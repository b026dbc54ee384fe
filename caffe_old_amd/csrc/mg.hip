// mg.hip -- geometric multigrid preconditioner for the pressure-correction systems, and
// the BiCGStab that uses it.
//
// The reference hands each p' system (assembled in Simple.cpp:231-269; with the ghost-cell
// rows of IbPiso.cpp:148-185 it is nonsymmetric) to a preconditioned PETSc Krylov solve.
// With a pointwise preconditioner the iteration count grows with the square of the
// longest cell path between fixed-pressure faces: a 256-cell channel needs thousands.
// Here BiCGStab keeps its recurrences in FP64 and is preconditioned on the right by one
// V-cycle in FP32:
//   - level 0 is the system itself, ghost rows included; damped Jacobi (weight 0.8),
//     two sweeps before and two after the coarse correction, each sweep ending with the
//     ghost rows solved exactly for their cells;
//   - coarser levels aggregate 2 x 2 x 2 cells (1 along a direction of 2 cells or fewer).
//     A coarse face coefficient is the sum of the fine face coefficients across it divided
//     by the aggregation factor along its normal (4x the area at 2x the distance), so the
//     coarse operator is the finite-volume Laplacian of the coarse cells; the fixed-
//     pressure part of a diagonal carries over the same way.  Ghost cells join the coarse
//     graph through their faces to fluid cells; inactive cells drop out, and a coarse cell
//     made of inactive cells only is an identity row;
//   - restriction sums the fine residuals of an aggregate, prolongation is cell-centred
//     trilinear (3/4 and 1/4 per direction);
//   - the coarsest level (8 cells or fewer) is solved with its dense inverse, formed once
//     per solve by one workgroup.
// Fixed sweep counts and the exact coarse solve make the cycle a fixed linear map, which
// BiCGStab requires of its preconditioner.
#include <hip/hip_ext.h>

#include "internal.hpp"
#include "krylov.hpp"
#include "sym4.hpp"

#include <algorithm>
#include <cmath>

namespace caffe {

constexpr int MBS = 256;          // block of the per-level kernels (one thread per cell)
constexpr double MG_W = 0.8;      // Jacobi damping
constexpr int MG_COARSEST = 8;    // cells of the level solved densely (a 64-cell level took
                                  // 262 us to invert in one workgroup; 8 cells take a few)

__host__ __device__ inline Sym4<float> levOp(const MgLevel& L) {
    return Sym4<float>{L.A, L.D.P, L.D.sJ, L.D.sK, OvView{nullptr, nullptr, nullptr}, nullptr};
}

// Damped Jacobi weights w = 0.8 / aP of a level's storage, formed once per solve (zero on
// the halo, whose diagonal is zero).  The sweeps then run in FP32 without divisions.
__global__ void k_mg_weights(const float* A, int P, float* w) {
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q < P) {
        const float a = A[q];
        w[q] = a != 0.f ? (float)(MG_W / (double)a) : 0.f;
    }
}

// the plain stencil of the p' storage at p on the values x at p, W, E, S, N, B, T
__device__ __forceinline__ float stencil7(const float* A, int P, int sJ, int sK, int p, const float (&x)[7]) {
    return A[p] * x[0] + A[P + p] * x[1] + A[P + p + 1] * x[2] + A[2 * P + p] * x[3] + A[2 * P + p + sJ] * x[4]
         + A[3 * P + p] * x[5] + A[3 * P + p + sK] * x[6];
}
__device__ __forceinline__ void nbrs(int p, int sJ, int sK, int (&nb)[7]) {
    nb[0] = p; nb[1] = p - 1; nb[2] = p + 1; nb[3] = p - sJ; nb[4] = p + sJ; nb[5] = p - sK; nb[6] = p + sK;
}

// two damped Jacobi sweeps from zero, at cell p: z1 = w b, z = z1 + w (b - A z1)
template <class B>
__device__ __forceinline__ float preSmooth(const Sym4<float>& A, const float* wt, B b, int p) {
    int nb[7];
    nbrs(p, A.sJ, A.sK, nb);
    float z1[7];
#pragma unroll
    for (int m = 0; m < 7; ++m) z1[m] = wt[nb[m]] * (float)b(nb[m]);
    return z1[0] + wt[p] * ((float)b(p) - stencil7(A.A, A.P, A.sJ, A.sK, p, z1));
}

// Ghost rows on the fine level.  A ghost row ties its cell to the interpolated value at the
// image point (IbPiso.cpp:148-185); damped Jacobi smooths such rows badly (on the
// 64 x 16 x 16 channel the cycle then needs 40 BiCGStab iterations, 7 when they are solved
// exactly).  So every fine-level sweep ends with each ghost row solved exactly for its own
// cell, from the values this sweep gave the other cells of the row (their input values
// where those are ghost cells too).  The sweep kernels run the plain stencil on every
// cell, and one thread per ghost row then writes the ghost cells (k_mg_first_*,
// k_mg_ghost_*): it reads other cells than those it writes, so the result does not depend
// on thread order.
// preSmooth on the fine level.  A fluid cell's ghost neighbours take their first-sweep
// value g1 (the ghost row solved from w b, formed by k_mg_first_*); a ghost cell gets its
// g1 here and its second sweep from k_mg_ghost_pre.  The seven neighbours' values, weights
// and ghost-row numbers (G: the system has ghost rows) are loaded together; a cell next to
// a ghost cell then loads those neighbours' g1.  An inactive cell's stored row is the
// identity with zero faces, so the plain stencil is its row.
template <bool G, class B>
__device__ __forceinline__ float preSmooth0(const Sym4<float>& A, const float* __restrict__ wt, B b,
                                            const float* __restrict__ g1, int p) {
    int nb[7];
    nbrs(p, A.sJ, A.sK, nb);
    float z1[7];
    int gr[7];
#pragma unroll
    for (int m = 0; m < 7; ++m) {
        z1[m] = wt[nb[m]] * (float)b(nb[m]);
        gr[m] = G ? A.ov.rowOf[nb[m]] : -1;
    }
    if (G) {
        bool anyG = false;
#pragma unroll
        for (int m = 0; m < 7; ++m) anyG |= gr[m] >= 0;
        if (anyG) {
#pragma unroll
            for (int m = 0; m < 7; ++m)
                if (gr[m] >= 0) z1[m] = g1[gr[m]];
            if (gr[0] >= 0) return z1[0];
        }
    }
    return z1[0] + wt[p] * ((float)b(p) - stencil7(A.A, A.P, A.sJ, A.sK, p, z1));
}

// g1 = the ghost rows solved from w b (the first sweep from zero), one thread per row
template <class B>
__device__ __forceinline__ void firstGhost(const Sym4<float>& A, const float* wt, B b, const int* cell, float* g1, int r) {
    const int p = cell[r];
    const float* v = (const float*)A.ov.vals + 9 * r;
    int cols[8];
#pragma unroll
    for (int m = 0; m < 8; ++m) cols[m] = A.ov.cols[9 * r + 1 + m];
    double s = b(p);
#pragma unroll
    for (int m = 0; m < 8; ++m) {
        const int c = cols[m] >= 0 ? cols[m] : p;
        const double zJ = (double)(wt[c] * (float)b(c));
        s -= cols[m] >= 0 ? (double)v[1 + m] * zJ : 0.;
    }
    g1[r] = (float)(s / (double)v[0]);
}

// ------------------------------------------------------------------ hierarchy set-up
// Fine coefficients seen by the aggregation: face(dir, i, j, k) is the stored low-face
// coefficient of cell (i, j, k) along dir; surplus(i, j, k) the part of -aP that no
// stored face accounts for (fixed-pressure boundary faces), zero where the row is not a
// plain fluid row.
struct FineP {            // level 0: the FP64 p' storage and the cell statuses
    const double* A; const int8_t* st; Dims D;
    __device__ double face(int dir, int i, int j, int k) const { return A[(1 + dir) * D.P + D.pidx(i, j, k)]; }
    __device__ double surplus(int i, int j, int k) const {
        const int p = D.pidx(i, j, k), P = D.P;
        if (st[p] != ST_ACTIVE) return 0.;
        const double s = -(A[p] + A[P + p] + A[P + p + 1] + A[2 * P + p] + A[2 * P + p + D.sJ] + A[3 * P + p]
                           + A[3 * P + p + D.sK]);
        return s > 1e-12 * -A[p] ? s : 0.;
    }
};
struct FineL {            // a coarse level as the fine side of the next aggregation
    const float* A; Dims D;
    __device__ double face(int dir, int i, int j, int k) const { return A[(1 + dir) * D.P + D.pidx(i, j, k)]; }
    __device__ double surplus(int i, int j, int k) const {
        const int p = D.pidx(i, j, k), P = D.P;
        const double a = A[p];
        if (a > 0.) return 0.;   // identity row of an inactive aggregate
        const double s = -(a + (double)A[P + p] + A[P + p + 1] + A[2 * P + p] + A[2 * P + p + D.sJ] + A[3 * P + p]
                           + A[3 * P + p + D.sK]);
        return s > 1e-6 * -a ? s : 0.;
    }
};

template <class Fine>
__global__ void k_mg_coarsen(Fine F, MgLevel c) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= c.D.N) return;
    const Dims& Df = F.D;
    int I, J, K;
    const int pc = c.D.cell(e, I, J, K);
    const int i0 = c.f[0] * I, i1 = min(i0 + c.f[0], Df.nI);
    const int j0 = c.f[1] * J, j1 = min(j0 + c.f[1], Df.nJ);
    const int k0 = c.f[2] * K, k1 = min(k0 + c.f[2], Df.nK);
    double lo[3] = {0., 0., 0.}, hi[3] = {0., 0., 0.}, sur = 0.;
    for (int k = k0; k < k1; ++k)
        for (int j = j0; j < j1; ++j) {
            if (I > 0) lo[0] += F.face(0, i0, j, k);
            if (I + 1 < c.D.nI) hi[0] += F.face(0, i0 + c.f[0], j, k);
        }
    for (int k = k0; k < k1; ++k)
        for (int i = i0; i < i1; ++i) {
            if (J > 0) lo[1] += F.face(1, i, j0, k);
            if (J + 1 < c.D.nJ) hi[1] += F.face(1, i, j0 + c.f[1], k);
        }
    for (int j = j0; j < j1; ++j)
        for (int i = i0; i < i1; ++i) {
            if (K > 0) lo[2] += F.face(2, i, j, k0);
            if (K + 1 < c.D.nK) hi[2] += F.face(2, i, j, k0 + c.f[2]);
        }
    for (int k = k0; k < k1; ++k)
        for (int j = j0; j < j1; ++j)
            for (int i = i0; i < i1; ++i) sur += F.surplus(i, j, k);
    double aP = 0.;
    for (int d = 0; d < 3; ++d) {
        lo[d] /= c.f[d];
        hi[d] /= c.f[d];
        aP -= lo[d] + hi[d];
    }
    aP -= sur / (double)max(c.f[0], max(c.f[1], c.f[2]));
    if (aP == 0.) aP = 1.;   // no face and no fixed value: an aggregate of inactive cells
    const int P = c.D.P;
    c.A[pc] = (float)aP;
    c.A[P + pc] = (float)lo[0];
    c.A[2 * P + pc] = (float)lo[1];
    c.A[3 * P + pc] = (float)lo[2];
}

// Dense inverse of the coarsest operator (n <= MG_COARSEST cells) by in-place Gauss-Jordan without
// pivoting (the operator is symmetric negative definite, or made so by grounding cell 0
// when no face fixes the pressure); one workgroup, inverse stored row-major.
__global__ void k_mg_dense(MgLevel L, int n, int ground, double* inv) {
    __shared__ double M[MG_COARSEST][MG_COARSEST + 1];
    __shared__ double col[MG_COARSEST];
    const Dims& D = L.D;
    const int P = D.P;
    for (int e = threadIdx.x; e < n * n; e += blockDim.x) M[e / n][e % n] = 0.;
    __syncthreads();
    if ((int)threadIdx.x < n) {
        const int q = threadIdx.x;
        int i, j, k;
        const int p = D.cell(q, i, j, k);
        M[q][q] = L.A[p];
        if (i > 0) M[q][q - 1] = L.A[P + p];
        if (i + 1 < D.nI) M[q][q + 1] = L.A[P + p + 1];
        if (j > 0) M[q][q - D.nI] = L.A[2 * P + p];
        if (j + 1 < D.nJ) M[q][q + D.nI] = L.A[2 * P + p + D.sJ];
        if (k > 0) M[q][q - D.nI * D.nJ] = L.A[3 * P + p];
        if (k + 1 < D.nK) M[q][q + D.nI * D.nJ] = L.A[3 * P + p + D.sK];
    }
    __syncthreads();
    if (ground && threadIdx.x == 0) {
        const double d = M[0][0];
        for (int m = 0; m < n; ++m) { M[0][m] = 0.; M[m][0] = 0.; }
        M[0][0] = d != 0. ? d : -1.;
    }
    __syncthreads();
    for (int k = 0; k < n; ++k) {
        if (threadIdx.x == 0) {
            const double piv = 1. / M[k][k];
            M[k][k] = 1.;
            for (int m = 0; m < n; ++m) M[k][m] *= piv;
        }
        __syncthreads();
        if ((int)threadIdx.x < n) col[threadIdx.x] = M[threadIdx.x][k];
        __syncthreads();
        for (int e = threadIdx.x; e < n * n; e += blockDim.x) {
            const int r = e / n, m = e % n;
            if (r == k) continue;
            M[r][m] = (m == k ? 0. : M[r][m]) - col[r] * M[k][m];
        }
        __syncthreads();
    }
    for (int e = threadIdx.x; e < n * n; e += blockDim.x) inv[e] = M[e / n][e % n];
}

// ------------------------------------------------------------------ cycle kernels
__device__ __forceinline__ bool mgDone(const double* st) { return st[B_DONE] != 0.; }

// coarse level: z = two Jacobi sweeps from zero on b
__global__ void k_mg_pre(MgLevel L, const double* st) {
    if (mgDone(st)) return;
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= L.D.N) return;
    int i, j, k;
    const int p = L.D.cell(e, i, j, k);
    L.z[p] = preSmooth(levOp(L), L.w, [&](int c) { return L.b[c]; }, p);
}

// zout = zin + w (b - A zin), coarse level
__global__ void k_mg_post(MgLevel L, const float* zin, float* zout, const double* st) {
    if (mgDone(st)) return;
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= L.D.N) return;
    int i, j, k;
    const int p = L.D.cell(e, i, j, k);
    int nb[7];
    nbrs(p, L.D.sJ, L.D.sK, nb);
    float x[7];
#pragma unroll
    for (int m = 0; m < 7; ++m) x[m] = zin[nb[m]];
    zout[p] = x[0] + L.w[p] * (L.b[p] - stencil7(L.A, L.D.P, L.D.sJ, L.D.sK, p, x));
}

// c.b = the sum over each aggregate of the fine residuals res(p).  Eight consecutive lanes
// take the (up to) eight fine cells of one aggregate, lane s at offset (s & 1, s >> 1 & 1,
// s >> 2); the butterfly over those lanes adds in a fixed order.
template <class R>
__device__ __forceinline__ void restrictAgg(const Dims& Df, R res, const MgLevel& c, int t) {
    const int e = t >> 3, sub = t & 7;
    int I = 0, J = 0, K = 0, pc = 0;
    double s = 0.;
    if (e < c.D.N) {
        pc = c.D.cell(e, I, J, K);
        const int di = sub & 1, dj = (sub >> 1) & 1, dk = sub >> 2;
        const int i = c.f[0] * I + di, j = c.f[1] * J + dj, k = c.f[2] * K + dk;
        if (di < c.f[0] && dj < c.f[1] && dk < c.f[2] && i < Df.nI && j < Df.nJ && k < Df.nK) s = res(Df.pidx(i, j, k));
    }
    s += __shfl_xor(s, 1, 64);
    s += __shfl_xor(s, 2, 64);
    s += __shfl_xor(s, 4, 64);
    if (e < c.D.N && sub == 0) c.b[pc] = (float)s;
}
// residual b - A z of a level's plain stencil at p
__device__ __forceinline__ float levResid(const float* A, const float* b, const float* z, const Dims& D, int p) {
    int nb[7];
    nbrs(p, D.sJ, D.sK, nb);
    float x[7];
#pragma unroll
    for (int m = 0; m < 7; ++m) x[m] = z[nb[m]];
    return b[p] - stencil7(A, D.P, D.sJ, D.sK, p, x);
}
// Level 0: the plain stencil's residual on fluid and inactive cells (an inactive row is the
// identity, zero at b = z = 0); ghost rows restrict nothing.  Their residuals are values
// of the interpolation constraint, not flux balances, and after the sweep's exact solve
// they hold only the coupling among ghost cells: restricted, they spoil the coarse right-
// hand side (on the 128 x 32 x 32 channel with a sphere of radius 6.4 cells the cycle then
// needs 12 BiCGStab iterations instead of 9).  G: the system has ghost rows.
template <bool G>
__global__ void k_mg_restrict0(Dims Df, const float* __restrict__ A, const int* __restrict__ rowOf,
                               const double* __restrict__ b, const float* __restrict__ z, MgLevel c, const double* st) {
    if (mgDone(st)) return;
    restrictAgg(Df, [&](int p) -> double {
        if (G && rowOf[p] >= 0) return 0.;
        int nb[7];
        nbrs(p, Df.sJ, Df.sK, nb);
        float x[7];
#pragma unroll
        for (int m = 0; m < 7; ++m) x[m] = z[nb[m]];
        return b[p] - (double)stencil7(A, Df.P, Df.sJ, Df.sK, p, x);
    }, c, blockIdx.x * blockDim.x + threadIdx.x);
}
__global__ void k_mg_restrict(MgLevel f, MgLevel c, const double* st) {
    if (mgDone(st)) return;
    restrictAgg(f.D, [&](int p) { return (double)levResid(f.A, f.b, f.z, f.D, p); }, c, blockIdx.x * blockDim.x + threadIdx.x);
}

// cell-centred linear interpolation along one direction: fine index f, aggregation factor
// fa, nc coarse cells -> coarse cells c0, c1 with weights w0, w1
__device__ __forceinline__ void lin(int f, int fa, int nc, int& c0, int& c1, float& w0, float& w1) {
    if (fa == 1) { c0 = c1 = f; w0 = 1.f; w1 = 0.f; return; }
    c0 = f >> 1;
    c1 = (f & 1) ? c0 + 1 : c0 - 1;
    c1 = c1 < 0 ? 0 : (c1 >= nc ? nc - 1 : c1);
    w0 = 0.75f; w1 = 0.25f;
}

// zf += P zc on the fine level (cells of status INACTIVE keep their value)
__global__ void k_mg_prolong(Dims Df, MgLevel c, float* zf, const int8_t* stf, const double* st) {
    if (mgDone(st)) return;
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= Df.N) return;
    int i, j, k;
    const int p = Df.cell(e, i, j, k);
    if (stf && stf[p] == ST_INACTIVE) return;
    int cx[2], cy[2], cz[2];
    float wx[2], wy[2], wz[2];
    lin(i, c.f[0], c.D.nI, cx[0], cx[1], wx[0], wx[1]);
    lin(j, c.f[1], c.D.nJ, cy[0], cy[1], wy[0], wy[1]);
    lin(k, c.f[2], c.D.nK, cz[0], cz[1], wz[0], wz[1]);
    float s = 0.f;
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b)
#pragma unroll
            for (int d = 0; d < 2; ++d) s += wz[a] * wy[b] * wx[d] * c.z[c.D.pidx(cx[d], cy[b], cz[a])];
    zf[p] += s;
}

// The bottom of the cycle in one workgroup: from level L[0] (its b restricted) down to the
// coarsest level L[n - 1] and back up to L[0].z, the per-level steps of the kernels above
// separated by workgroup barriers (global memory written before a barrier is visible to
// the whole workgroup after it).  It replaces five launches per level on levels small
// enough for one workgroup (at most MG_TAIL_CELLS cells).
constexpr int MG_TAIL_LEVELS = 8;
constexpr int MG_TAIL_CELLS = 4096;
struct MgTail { int n; MgLevel L[MG_TAIL_LEVELS]; };
__global__ void __launch_bounds__(1024) k_mg_tail(MgTail T, const double* inv, const double* st) {
    if (mgDone(st)) return;
    const int tid = threadIdx.x, bd = blockDim.x;
    for (int l = 0; l < T.n - 1; ++l) {
        const MgLevel& L = T.L[l];
        for (int e = tid; e < L.D.N; e += bd) {
            int i, j, k;
            const int p = L.D.cell(e, i, j, k);
            L.z[p] = preSmooth(levOp(L), L.w, [&](int c) { return L.b[c]; }, p);
        }
        __syncthreads();
        const MgLevel& C = T.L[l + 1];
        for (int base = 0; base < 8 * C.D.N; base += bd)   // every lane of a wave takes part in the butterfly
            restrictAgg(L.D, [&](int p) { return (double)levResid(L.A, L.b, L.z, L.D, p); }, C, base + tid);
        __syncthreads();
    }
    {   // coarsest level: z = inverse * b
        __shared__ double bs[MG_COARSEST];
        const MgLevel& L = T.L[T.n - 1];
        const int n = L.D.N;
        int i, j, k;
        if (tid < n) bs[tid] = L.b[L.D.cell(tid, i, j, k)];
        __syncthreads();
        if (tid < n) {
            double acc = 0.;
            for (int m = 0; m < n; ++m) acc += inv[tid * n + m] * bs[m];
            L.z[L.D.cell(tid, i, j, k)] = (float)acc;
        }
        __syncthreads();
    }
    for (int l = T.n - 2; l >= 0; --l) {
        const MgLevel& L = T.L[l];
        const MgLevel& C = T.L[l + 1];
        for (int e = tid; e < L.D.N; e += bd) {
            int i, j, k;
            const int p = L.D.cell(e, i, j, k);
            int cx[2], cy[2], cz[2];
            float wx[2], wy[2], wz[2];
            lin(i, C.f[0], C.D.nI, cx[0], cx[1], wx[0], wx[1]);
            lin(j, C.f[1], C.D.nJ, cy[0], cy[1], wy[0], wy[1]);
            lin(k, C.f[2], C.D.nK, cz[0], cz[1], wz[0], wz[1]);
            float s = 0.f;
#pragma unroll
            for (int a = 0; a < 2; ++a)
#pragma unroll
                for (int b = 0; b < 2; ++b)
#pragma unroll
                    for (int d = 0; d < 2; ++d) s += wz[a] * wy[b] * wx[d] * C.z[C.D.pidx(cx[d], cy[b], cz[a])];
            L.z[p] += s;
        }
        __syncthreads();
        for (int sweep = 0; sweep < 2; ++sweep) {
            const float* zin = sweep == 0 ? L.z : L.t;
            float* zout = sweep == 0 ? L.t : L.z;
            for (int e = tid; e < L.D.N; e += bd) {
                int i, j, k;
                const int p = L.D.cell(e, i, j, k);
                int nb[7];
                nbrs(p, L.D.sJ, L.D.sK, nb);
                float x[7];
#pragma unroll
                for (int m = 0; m < 7; ++m) x[m] = zin[nb[m]];
                zout[p] = x[0] + L.w[p] * (L.b[p] - stencil7(L.A, L.D.P, L.D.sJ, L.D.sK, p, x));
            }
            __syncthreads();
        }
    }
}

// level 0: zout = zin + w (b - A zin) on every cell with the plain stencil (the ghost
// cells are rewritten by k_mg_ghost_post)
__global__ void k_mg_post0(Dims D, const float* __restrict__ A, const float* __restrict__ wt, const double* __restrict__ b,
                           const float* __restrict__ zin, float* __restrict__ zout, const double* st) {
    if (mgDone(st)) return;
    FOR_CELLS(D, q, p) {
        int nb[7];
        nbrs(p, D.sJ, D.sK, nb);
        float x[7];
#pragma unroll
        for (int m = 0; m < 7; ++m) x[m] = zin[nb[m]];
        zout[p] = x[0] + wt[p] * ((float)b[p] - stencil7(A, D.P, D.sJ, D.sK, p, x));
    }
}

// the ghost rows after the pre-smoothing: z holds the second sweep on fluid cells, g1 the
// first sweep of the ghost cells.  The row's eight cells, their ghost-row numbers and
// values are loaded together.
__global__ void k_mg_ghost_pre(int n, const int* cell, Sym4<float> A, const double* b, const float* g1, float* z,
                               const double* st) {
    if (mgDone(st)) return;
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n) return;
    const int p = cell[r];
    const float* v = (const float*)A.ov.vals + 9 * r;
    int cols[8], gr[8];
    float zc[8], gc[8];
#pragma unroll
    for (int m = 0; m < 8; ++m) cols[m] = A.ov.cols[9 * r + 1 + m];
#pragma unroll
    for (int m = 0; m < 8; ++m) {
        const int c = cols[m] >= 0 ? cols[m] : p;
        gr[m] = A.ov.rowOf[c];
        zc[m] = z[c];
    }
#pragma unroll
    for (int m = 0; m < 8; ++m) gc[m] = g1[gr[m] >= 0 ? gr[m] : r];
    double s = b[p];
#pragma unroll
    for (int m = 0; m < 8; ++m)
        if (cols[m] >= 0) s -= (double)v[1 + m] * (double)(gr[m] >= 0 ? gc[m] : zc[m]);
    z[p] = (float)(s / (double)v[0]);
}
// the ghost rows after a post-smoothing sweep zin -> zout: fluid cells of the row at their
// swept value, ghost cells at their input value
__global__ void k_mg_ghost_post(int n, const int* cell, Sym4<float> A, const double* b, const float* zin, float* zout,
                                const double* st) {
    if (mgDone(st)) return;
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n) return;
    const int p = cell[r];
    const float* v = (const float*)A.ov.vals + 9 * r;
    int cols[8];
    float zi[8], zo[8];
    int gr[8];
#pragma unroll
    for (int m = 0; m < 8; ++m) cols[m] = A.ov.cols[9 * r + 1 + m];
#pragma unroll
    for (int m = 0; m < 8; ++m) {
        const int c = cols[m] >= 0 ? cols[m] : p;
        gr[m] = A.ov.rowOf[c];
        zi[m] = zin[c];
        zo[m] = zout[c];
    }
    double s = b[p];
#pragma unroll
    for (int m = 0; m < 8; ++m)
        if (cols[m] >= 0) s -= (double)v[1 + m] * (double)(gr[m] >= 0 ? zi[m] : zo[m]);
    zout[p] = (float)(s / (double)v[0]);
}

// ------------------------------------------------------------------ BiCGStab kernels
// r = b on live cells, rh = r, p = v = x = 0; partials (rh, r), (r, r)
__global__ void k_mg_init(Dims D, const int8_t* stc, const double* b, double* r, double* rh, double* p0, double* p1,
                          double* v, double* x, double* part, double* st, double tol2, int maxit) {
    double acc[2] = {0., 0.};
    FOR_CELLS(D, q, p) {
        const double bv = stc[p] != ST_INACTIVE ? b[p] : 0.;
        r[p] = bv; rh[p] = bv; p0[p] = 0.; p1[p] = 0.; v[p] = 0.; x[p] = 0.;
        acc[0] += bv * bv;
        acc[1] += bv * bv;
    }
    writePartials<KBS, 2>(acc, part);
    if (BLOCK0) {
        st[B_IT] = 0.; st[B_DONE] = 0.; st[B_BB] = -1.; st[B_RHOPREV] = 1.; st[B_ALPHA] = 1.; st[B_OMEGA] = 1.;
        st[B_BRK] = 0.; st[B_TOL2] = tol2; st[B_MAXIT] = (double)maxit;
    }
}
__device__ __forceinline__ LoopCtl mgLoop(const double* st) { return LoopCtl{st[B_TOL2], (int)st[B_MAXIT], 0.}; }

// The ghost rows' first sweep ahead of the two level-0 heads: each computes the head's
// beta (alpha) from the same partial sums and state, without recording anything.
__global__ void k_mg_first_p(int n, const int* cell, Sym4<float> A, const float* wt, float* g1, const double* r,
                             const double* pold, const double* v, const double* partIn, int nb, const double* st) {
    BiState<1> S;
    S.load(st);
    if (!S.any()) return;
    const LoopCtl lc = mgLoop(st);
    double t[2];
    readPartials<MBS, 2>(partIn, nb, t);
    const double rho = t[0], rr = t[1], bb = S.bb[0] < 0. ? rr : S.bb[0];
    if (rr <= lc.tol2 * bb || S.it[0] >= lc.maxit || rr == 0. || S.brk[0] != 0. || rho == 0.) return;
    const double bt = S.it[0] == 0. ? 0. : (rho / S.rhoPrev[0]) * (S.al[0] / S.om[0]), w = S.om[0];
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q < n) firstGhost(A, wt, [&](int c) { return r[c] + bt * (pold[c] - w * v[c]); }, cell, g1, q);
}
__global__ void k_mg_first_s(int n, const int* cell, Sym4<float> A, const float* wt, float* g1, const double* r,
                             const double* v, const double* partIn, int nb, const double* st) {
    BiState<1> S;
    S.load(st);
    if (!S.any()) return;
    double t[1];
    readPartials<MBS, 1>(partIn, nb, t);
    if (t[0] == 0.) return;
    const double a = S.rho[0] / t[0];
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q < n) firstGhost(A, wt, [&](int c) { return r[c] - a * v[c]; }, cell, g1, q);
}

// [convergence on (rh, r), (r, r); beta] p_new = r + beta (p - omega v); the cycle's
// level-0 pre-smoothing of p_new into z (G: the system has ghost rows)
template <bool G>
__global__ void k_mg_pre_p(Dims D, Sym4<float> A, const float* __restrict__ wt, const float* __restrict__ g1,
                           const double* __restrict__ r, const double* __restrict__ pold, double* __restrict__ pnew,
                           const double* __restrict__ v, float* __restrict__ z, const double* partIn, int nb, double* st) {
    BiState<1> S;
    S.load(st);
    if (!S.any()) return;
    const LoopCtl lc = mgLoop(st);
    double beta[1], om[1];
    bool live[1];
    biaHead<1>(partIn, nb, st, lc, S, beta, om, live);
    if (!live[0]) return;
    const double bt = beta[0], w = om[0];
    auto pf = [&](int c) { return r[c] + bt * (pold[c] - w * v[c]); };
    FOR_CELLS(D, q, p) {
        const double pp = pf(p);
        const float zz = preSmooth0<G>(A, wt, pf, g1, p);
        pnew[p] = pp;
        z[p] = zz;
    }
}

// [alpha = rho / (rh, v)] s = r - alpha v; the level-0 pre-smoothing of s into z
template <bool G>
__global__ void k_mg_pre_s(Dims D, Sym4<float> A, const float* __restrict__ wt, const float* __restrict__ g1,
                           const double* __restrict__ r, const double* __restrict__ v, double* __restrict__ s,
                           float* __restrict__ z, const double* partIn, int nb, double* st) {
    BiState<1> S;
    S.load(st);
    if (!S.any()) return;
    double al[1];
    bool live[1];
    bibHead<1>(partIn, nb, st, S, al, live);
    if (!live[0]) return;
    const double a = al[0];
    auto sf = [&](int c) { return r[c] - a * v[c]; };
    FOR_CELLS(D, q, p) {
        const double ss = sf(p);
        const float zz = preSmooth0<G>(A, wt, sf, g1, p);
        s[p] = ss;
        z[p] = zz;
    }
}

// v = A ph; partial (rh, v)
__global__ void k_mg_v(Dims D, Sym4<double> A, const float* ph, const double* rh, double* v, double* partOut,
                       const double* st) {
    if (mgDone(st)) return;
    double acc[1] = {0.};
    auto xf = [&](int c) { return (double)ph[c]; };
    FOR_CELLS(D, q, p) {
        const double vv = A.apply(xf, p);
        v[p] = vv;
        acc[0] += rh[p] * vv;
    }
    writePartials<KBS, 1>(acc, partOut);
}

// t = A sh; partials (t, s), (t, t)
__global__ void k_mg_t(Dims D, Sym4<double> A, const float* sh, const double* s, double* t, double* partOut,
                       const double* st) {
    if (mgDone(st) || st[B_BRK] == 1.) return;
    double acc[2] = {0., 0.};
    auto xf = [&](int c) { return (double)sh[c]; };
    FOR_CELLS(D, q, p) {
        const double tt = A.apply(xf, p);
        t[p] = tt;
        acc[0] += tt * s[p];
        acc[1] += tt * tt;
    }
    writePartials<KBS, 2>(acc, partOut);
}

// [omega = (t, s) / (t, t)] x += alpha ph + omega sh; r = s - omega t; partials (rh, r), (r, r)
__global__ void k_mg_x(Dims D, double* x, double* r, const double* rh, const float* ph, const float* sh, const double* s,
                       const double* t, const double* partIn, int nb, double* partOut, double* st) {
    const double sDone = st[B_DONE], sBrk = st[B_BRK], sAl = st[B_ALPHA], sRho = st[B_RHOCUR], sIt = st[B_IT];
    if (sDone != 0.) return;
    double tt[2];
    readPartials<KBS, 2>(partIn, nb, tt);
    const bool live = sBrk != 1.;   // KB found (rh, v) = 0: no s, t this step
    const double om = tt[1] > 0. ? tt[0] / tt[1] : 0.;
    if (BLOCK0 && live) {
        st[B_OMEGA] = om; st[B_RHOPREV] = sRho; st[B_IT] = sIt + 1.;
        if (om == 0.) st[B_BRK] = 2.;
    }
    double acc[2] = {0., 0.};
    FOR_CELLS(D, q, p) {
        double rn;
        if (live) {
            x[p] += sAl * (double)ph[p] + om * (double)sh[p];
            rn = s[p] - om * t[p];
            r[p] = rn;
        } else {
            rn = r[p];
        }
        acc[0] += rh[p] * rn;
        acc[1] += rn * rn;
    }
    writePartials<KBS, 2>(acc, partOut);
}

// ------------------------------------------------------------------ host side
static int kgrid(const Dims& D) { return blocksFor(D.N, KBS, KGRID); }
static int lgrid(const Dims& D) { return blocksFor(D.N, MBS); }
static int rgrid(const Dims& Dc) { return blocksFor(8L * Dc.N, MBS); }   // 8 lanes per coarse cell

static Dims mkDims(int nI, int nJ, int nK) {
    Dims D;
    D.nI = nI; D.nJ = nJ; D.nK = nK;
    D.sJ = nI + 2; D.sK = (nI + 2) * (nJ + 2);
    D.P = D.sK * (nK + 2);
    D.N = nI * nJ * nK;
    return D;
}

// Levels for the context's grid, allocated once per grid: each level halves every
// dimension of more than 2 cells (rounding up) until MG_COARSEST cells or fewer remain.
// Returns false when the grid itself is that small.
static bool mgPlan(Ctx& c) {
    MgHier& H = c.mg;
    if (H.nl > 0) return H.nl > 1;
    const Dims& D = c.D;
    H.L[0].D = D;
    H.nl = 1;
    Dims cur = D;
    while (cur.N > MG_COARSEST && H.nl < MG_MAX_LEVELS) {
        int f[3], n[3] = {cur.nI, cur.nJ, cur.nK};
        for (int d = 0; d < 3; ++d) f[d] = n[d] > 2 ? 2 : 1;
        if (f[0] == 1 && f[1] == 1 && f[2] == 1) break;
        MgLevel& L = H.L[H.nl];
        L.D = mkDims((n[0] + f[0] - 1) / f[0], (n[1] + f[1] - 1) / f[1], (n[2] + f[2] - 1) / f[2]);
        for (int d = 0; d < 3; ++d) L.f[d] = f[d];
        const size_t Pl = (size_t)L.D.P;
        L.A = falloc(c, 4 * Pl);
        L.b = falloc(c, Pl);
        L.z = falloc(c, Pl);
        L.t = falloc(c, Pl);
        L.w = falloc(c, Pl);
        cur = L.D;
        ++H.nl;
    }
    if (H.nl == 1 || cur.N > MG_COARSEST) { H.nl = 1; return false; }
    for (int q = 0; q < 3; ++q) H.z0[q] = falloc(c, (size_t)D.P);
    H.w0 = falloc(c, (size_t)D.P);
    double* inv = nullptr;
    HIPCHECK(hipMalloc(&inv, sizeof(double) * MG_COARSEST * MG_COARSEST));
    c.allocs.push_back(inv);
    c.deviceBytes += sizeof(double) * MG_COARSEST * MG_COARSEST;
    H.cinv = inv;
    return true;
}

// coarse operators of the current p' system (its FP32 copy made), the Jacobi weights of
// every level and the dense inverse of the coarsest
static void mgBuild(Ctx& c) {
    MgHier& H = c.mg;
    if (c.hasIB && c.ov.n > H.g1n) {
        H.g1 = falloc(c, (size_t)c.ov.n);
        H.g1n = c.ov.n;
        if (H.iter2) {   // the captured iterations hold the old pointer
            HIPCHECK(hipStreamSynchronize(c.stream));
            HIPCHECK(hipGraphExecDestroy(H.iter2));
            H.iter2 = nullptr;
        }
    }
    hipLaunchKernelGGL(k_mg_weights, dim3(blocksFor(c.D.P, MBS)), dim3(MBS), 0, c.stream, (const float*)c.Ap32, c.D.P, H.w0);
    hipLaunchKernelGGL(k_mg_coarsen<FineP>, dim3(lgrid(H.L[1].D)), dim3(MBS), 0, c.stream, FineP{c.Ap, c.st, c.D}, H.L[1]);
    for (int l = 2; l < H.nl; ++l)
        hipLaunchKernelGGL(k_mg_coarsen<FineL>, dim3(lgrid(H.L[l].D)), dim3(MBS), 0, c.stream,
                           FineL{H.L[l - 1].A, H.L[l - 1].D}, H.L[l]);
    for (int l = 1; l < H.nl; ++l)
        hipLaunchKernelGGL(k_mg_weights, dim3(blocksFor(H.L[l].D.P, MBS)), dim3(MBS), 0, c.stream, (const float*)H.L[l].A,
                           H.L[l].D.P, H.L[l].w);
    // The coarse operators carry no fixed value when no face fixes the pressure: they are
    // singular and the coarsest is grounded.  That holds with ghost rows too, although the
    // fine system is then not singular (a ghost row whose image stencil reaches into the
    // solid drops those zero-valued cells, which anchors it): the coarse levels have no
    // solid cells.  Inverting the ungrounded coarsest there ruined the cycle (54 BiCGStab
    // iterations to 1e-2 on a 16^3 closed cavity with a sphere, 7 grounded).
    bool fixedFace = false;
    for (int f = 0; f < 6; ++f) fixedFace |= c.pT[f] == BC_FIXED;
    const MgLevel& Lc = H.L[H.nl - 1];
    hipLaunchKernelGGL(k_mg_dense, dim3(1), dim3(MBS), 0, c.stream, Lc, Lc.D.N, fixedFace ? 0 : 1, H.cinv);
}

// k_mg_tail with the levels' iterates, sweep buffers and right-hand sides in LDS: the same
// steps in the same order (results bit-identical), but a step between two barriers then
// waits on LDS instead of on global memory that the previous step wrote, and the tail is
// about twenty such steps.  The operators and Jacobi weights stay in global memory (read
// only).  On entry the first level's b is copied in, with every other buffer zeroed (the
// halos the stencils read are zero, as in global memory); on exit the first level's z is
// copied out for the prolongation above it.  Used when the levels' 3 P floats fit in
// MG_TAIL_LDS (padded cells: 16^3 with its coarser levels takes 84 KB).
constexpr int MG_TAIL_LDS = 24576;   // floats (96 KB)
// level l of the tail with z, t and b in LDS (levels packed in order, 3 P floats each)
__device__ __forceinline__ MgLevel ldsLevel(const MgTail& T, float* lds, int l) {
    MgLevel L = T.L[l];
    int off = 0;
    for (int m = 0; m < l; ++m) off += 3 * T.L[m].D.P;
    L.z = lds + off;
    L.t = lds + off + L.D.P;
    L.b = lds + off + 2 * L.D.P;
    return L;
}
__global__ void __launch_bounds__(1024) k_mg_tail_lds(MgTail T, const double* inv, const double* st) {
    if (mgDone(st)) return;
    __shared__ float lds[MG_TAIL_LDS];
    const int tid = threadIdx.x, bd = blockDim.x;
    int used = 0;
    for (int l = 0; l < T.n; ++l) used += 3 * T.L[l].D.P;
    for (int q = tid; q < used; q += bd) lds[q] = 0.f;
    __syncthreads();
    {
        const MgLevel L0 = ldsLevel(T, lds, 0);
        for (int q = tid; q < L0.D.P; q += bd) L0.b[q] = T.L[0].b[q];
    }
    __syncthreads();
    for (int l = 0; l < T.n - 1; ++l) {
        const MgLevel L = ldsLevel(T, lds, l);
        for (int e = tid; e < L.D.N; e += bd) {
            int i, j, k;
            const int p = L.D.cell(e, i, j, k);
            L.z[p] = preSmooth(levOp(L), L.w, [&](int c) { return L.b[c]; }, p);
        }
        __syncthreads();
        const MgLevel C = ldsLevel(T, lds, l + 1);
        for (int base = 0; base < 8 * C.D.N; base += bd)
            restrictAgg(L.D, [&](int p) { return (double)levResid(L.A, L.b, L.z, L.D, p); }, C, base + tid);
        __syncthreads();
    }
    {
        __shared__ double bs[MG_COARSEST];
        const MgLevel L = ldsLevel(T, lds, T.n - 1);
        const int n = L.D.N;
        int i, j, k;
        if (tid < n) bs[tid] = L.b[L.D.cell(tid, i, j, k)];
        __syncthreads();
        if (tid < n) {
            double acc = 0.;
            for (int m = 0; m < n; ++m) acc += inv[tid * n + m] * bs[m];
            L.z[L.D.cell(tid, i, j, k)] = (float)acc;
        }
        __syncthreads();
    }
    for (int l = T.n - 2; l >= 0; --l) {
        const MgLevel L = ldsLevel(T, lds, l);
        const MgLevel C = ldsLevel(T, lds, l + 1);
        for (int e = tid; e < L.D.N; e += bd) {
            int i, j, k;
            const int p = L.D.cell(e, i, j, k);
            int cx[2], cy[2], cz[2];
            float wx[2], wy[2], wz[2];
            lin(i, C.f[0], C.D.nI, cx[0], cx[1], wx[0], wx[1]);
            lin(j, C.f[1], C.D.nJ, cy[0], cy[1], wy[0], wy[1]);
            lin(k, C.f[2], C.D.nK, cz[0], cz[1], wz[0], wz[1]);
            float s = 0.f;
#pragma unroll
            for (int a = 0; a < 2; ++a)
#pragma unroll
                for (int b = 0; b < 2; ++b)
#pragma unroll
                    for (int d = 0; d < 2; ++d) s += wz[a] * wy[b] * wx[d] * C.z[C.D.pidx(cx[d], cy[b], cz[a])];
            L.z[p] += s;
        }
        __syncthreads();
        for (int sweep = 0; sweep < 2; ++sweep) {
            const float* zin = sweep == 0 ? L.z : L.t;
            float* zout = sweep == 0 ? L.t : L.z;
            for (int e = tid; e < L.D.N; e += bd) {
                int i, j, k;
                const int p = L.D.cell(e, i, j, k);
                int nb[7];
                nbrs(p, L.D.sJ, L.D.sK, nb);
                float x[7];
#pragma unroll
                for (int m = 0; m < 7; ++m) x[m] = zin[nb[m]];
                zout[p] = x[0] + L.w[p] * (L.b[p] - stencil7(L.A, L.D.P, L.D.sJ, L.D.sK, p, x));
            }
            __syncthreads();
        }
    }
    const MgLevel L0 = ldsLevel(T, lds, 0);
    for (int e = tid; e < L0.D.N; e += bd) {
        int i, j, k;
        const int p = L0.D.cell(e, i, j, k);
        T.L[0].z[p] = L0.z[p];
    }
}

// levels 1 .. nl - 1 of the cycle: from level 1's right-hand side L[1].b to its iterate
// L[1].z (two sweeps from zero, restriction, the coarser levels, prolongation, two sweeps)
static void mgCoarseLevels(Ctx& c, const double* st) {
    MgHier& H = c.mg;
    hipStream_t s = c.stream;
    const int nl = H.nl;
    // levels lt .. nl - 1 run in one workgroup
    int lt = nl - 1;
    while (lt > 1 && H.L[lt - 1].D.N <= MG_TAIL_CELLS && nl - (lt - 1) <= MG_TAIL_LEVELS) --lt;
    for (int l = 1; l < lt; ++l) {
        hipLaunchKernelGGL(k_mg_pre, dim3(lgrid(H.L[l].D)), dim3(MBS), 0, s, H.L[l], st);
        hipLaunchKernelGGL(k_mg_restrict, dim3(rgrid(H.L[l + 1].D)), dim3(MBS), 0, s, H.L[l], H.L[l + 1], st);
    }
    MgTail T;
    T.n = nl - lt;
    long need = 0;
    for (int l = lt; l < nl; ++l) {
        T.L[l - lt] = H.L[l];
        need += 3L * H.L[l].D.P;
    }
    if (need <= MG_TAIL_LDS) hipLaunchKernelGGL(k_mg_tail_lds, dim3(1), dim3(1024), 0, s, T, (const double*)H.cinv, st);
    else hipLaunchKernelGGL(k_mg_tail, dim3(1), dim3(1024), 0, s, T, (const double*)H.cinv, st);
    for (int l = lt - 1; l >= 1; --l) {
        const MgLevel& L = H.L[l];
        hipLaunchKernelGGL(k_mg_prolong, dim3(lgrid(L.D)), dim3(MBS), 0, s, L.D, H.L[l + 1], L.z, (const int8_t*)nullptr, st);
        hipLaunchKernelGGL(k_mg_post, dim3(lgrid(L.D)), dim3(MBS), 0, s, L, (const float*)L.z, L.t, st);
        hipLaunchKernelGGL(k_mg_post, dim3(lgrid(L.D)), dim3(MBS), 0, s, L, (const float*)L.t, L.z, st);
    }
}

// the cycle after the level-0 pre-smoothing: z (level 0) holds that smoothing of b
static void mgCycleRest(Ctx& c, const Sym4<float>& A0, const double* b, float* z, const double* st) {
    MgHier& H = c.mg;
    const Dims& D = c.D;
    hipStream_t s = c.stream;
    const int g0 = kgrid(D);
    const bool ghosts = A0.ov.rowOf && c.ov.n > 0;
    const int gg = blocksFor(c.ov.n, MBS);
    if (ghosts)
        hipLaunchKernelGGL(k_mg_ghost_pre, dim3(gg), dim3(MBS), 0, s, c.ov.n, (const int*)c.ov.cell, A0, b,
                           (const float*)H.g1, z, st);
    if (ghosts)
        hipLaunchKernelGGL(k_mg_restrict0<true>, dim3(rgrid(H.L[1].D)), dim3(MBS), 0, s, D, (const float*)A0.A,
                           (const int*)A0.ov.rowOf, b, (const float*)z, H.L[1], st);
    else
        hipLaunchKernelGGL(k_mg_restrict0<false>, dim3(rgrid(H.L[1].D)), dim3(MBS), 0, s, D, (const float*)A0.A,
                           (const int*)nullptr, b, (const float*)z, H.L[1], st);
    mgCoarseLevels(c, st);
    hipLaunchKernelGGL(k_mg_prolong, dim3(lgrid(D)), dim3(MBS), 0, s, D, H.L[1], z, (const int8_t*)c.st, st);
    for (int sweep = 0; sweep < 2; ++sweep) {
        const float* zin = sweep == 0 ? z : H.z0[2];
        float* zout = sweep == 0 ? H.z0[2] : z;
        hipEvent_t e0, e1;
        if (c.probe.slot(PROBE_MG_POST0, e0, e1))   // dispatch-bracketing events: the kernel's own time
            hipExtLaunchKernelGGL(k_mg_post0, dim3(g0), dim3(KBS), 0, s, e0, e1, 0, D, A0.A, (const float*)H.w0, b, zin, zout,
                                  st);
        else
            hipLaunchKernelGGL(k_mg_post0, dim3(g0), dim3(KBS), 0, s, D, A0.A, (const float*)H.w0, b, zin, zout, st);
        if (ghosts)
            hipLaunchKernelGGL(k_mg_ghost_post, dim3(gg), dim3(MBS), 0, s, c.ov.n, (const int*)c.ov.cell, A0, b, zin, zout, st);
    }
}

// The buffers of the multigrid BiCGStab: FP64 vectors in the momentum work block (double
// views only, halos stay zero), FP32 preconditioned vectors on the fine grid
struct MgVecs {
    double *r, *rh, *p0, *p1, *v, *sv, *tv;
    float *ph, *sh;
};
static MgVecs mgVecs(Ctx& c) {
    const size_t P = (size_t)c.D.P;
    double* W = c.wm;
    return MgVecs{W, W + P, W + 2 * P, W + 3 * P, W + 4 * P, W + 5 * P, W + 6 * P, c.mg.z0[0], c.mg.z0[1]};
}

// one BiCGStab iteration: direction p from pc into pn, both V-cycles, the update of e
static void mgIteration(Ctx& c, const MgVecs& V, const double* pc, double* pn, double* e) {
    MgHier& H = c.mg;
    const Dims& D = c.D;
    const int P = D.P;
    hipStream_t s = c.stream;
    const OvView ov64{c.hasIB ? c.ov.rowOf : nullptr, c.ov.cols, c.ov.valP};
    const OvView ov32{c.hasIB ? c.ov.rowOf : nullptr, c.ov.cols, c.ov.valP32};
    const Sym4<double> A64{c.Ap, P, D.sJ, D.sK, ov64, c.st};
    const Sym4<float> A32{c.Ap32, P, D.sJ, D.sK, ov32, c.st};
    double* st = c.scal + SC_BI;
    double* partC = c.red.part;
    double* partA = c.red.part + MAX_RED_BLOCKS * RED_STRIDE;
    double* partB = c.red.part + 2 * MAX_RED_BLOCKS * RED_STRIDE;
    const int g = kgrid(D);
    const bool ghosts = A32.ov.rowOf && c.ov.n > 0;
    const int gg = blocksFor(c.ov.n, MBS);
    if (ghosts) {
        hipLaunchKernelGGL(k_mg_first_p, dim3(gg), dim3(MBS), 0, s, c.ov.n, (const int*)c.ov.cell, A32, (const float*)H.w0,
                           H.g1, (const double*)V.r, pc, (const double*)V.v, (const double*)partC, g, (const double*)st);
        hipLaunchKernelGGL(k_mg_pre_p<true>, dim3(g), dim3(KBS), 0, s, D, A32, (const float*)H.w0, (const float*)H.g1,
                           (const double*)V.r, pc, pn, (const double*)V.v, V.ph, (const double*)partC, g, st);
    } else {
        hipLaunchKernelGGL(k_mg_pre_p<false>, dim3(g), dim3(KBS), 0, s, D, A32, (const float*)H.w0, (const float*)H.g1,
                           (const double*)V.r, pc, pn, (const double*)V.v, V.ph, (const double*)partC, g, st);
    }
    mgCycleRest(c, A32, pn, V.ph, st);
    hipLaunchKernelGGL(k_mg_v, dim3(g), dim3(KBS), 0, s, D, A64, (const float*)V.ph, (const double*)V.rh, V.v, partA,
                       (const double*)st);
    if (ghosts) {
        hipLaunchKernelGGL(k_mg_first_s, dim3(gg), dim3(MBS), 0, s, c.ov.n, (const int*)c.ov.cell, A32, (const float*)H.w0,
                           H.g1, (const double*)V.r, (const double*)V.v, (const double*)partA, g, (const double*)st);
        hipLaunchKernelGGL(k_mg_pre_s<true>, dim3(g), dim3(KBS), 0, s, D, A32, (const float*)H.w0, (const float*)H.g1,
                           (const double*)V.r, (const double*)V.v, V.sv, V.sh, (const double*)partA, g, st);
    } else {
        hipLaunchKernelGGL(k_mg_pre_s<false>, dim3(g), dim3(KBS), 0, s, D, A32, (const float*)H.w0, (const float*)H.g1,
                           (const double*)V.r, (const double*)V.v, V.sv, V.sh, (const double*)partA, g, st);
    }
    mgCycleRest(c, A32, V.sv, V.sh, st);
    hipLaunchKernelGGL(k_mg_t, dim3(g), dim3(KBS), 0, s, D, A64, (const float*)V.sh, (const double*)V.sv, V.tv, partB,
                       (const double*)st);
    hipLaunchKernelGGL(k_mg_x, dim3(g), dim3(KBS), 0, s, D, e, V.r, (const double*)V.rh, (const float*)V.ph,
                       (const float*)V.sh, (const double*)V.sv, (const double*)V.tv, (const double*)partB, g, partC, st);
}

// Two iterations (the direction buffers back at their starting parity) captured once per
// context as a HIP graph: every pointer and launch shape in them is fixed per context, and
// the loop limits live in the device state.  The graph holds kernel launches only.  A
// capture that fails leaves the iterations launched one by one.
static bool mgGraph(Ctx& c, const MgVecs& V, double* e) {
    MgHier& H = c.mg;
    if (H.iter2) return true;
    if (H.graphOff) return false;
    hipStream_t s = c.stream;
    hipGraph_t graph = nullptr;
    bool ok = hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal) == hipSuccess;
    if (ok) {
        mgIteration(c, V, V.p0, V.p1, e);
        mgIteration(c, V, V.p1, V.p0, e);
        ok = hipStreamEndCapture(s, &graph) == hipSuccess && graph;
    }
    if (ok) ok = hipGraphInstantiate(&H.iter2, graph, nullptr, nullptr, 0) == hipSuccess;
    if (graph) (void)hipGraphDestroy(graph);
    if (!ok) {
        (void)hipGetLastError();
        H.iter2 = nullptr;
        H.graphOff = true;
    }
    return ok;
}

// BiCGStab with the cycle on A e = b from e = 0, to ||r|| <= rtol ||b|| on its recurrence.
// Returns the iteration count; done holds 1 (converged), 2 (breakdown) or 0 (cap).  The
// iterations replay the captured graph, two at a time, except while a kernel probe records
// (its events cannot be part of a graph).
static int mgRun(Ctx& c, const double* b, double* e, double rtol, int maxit, int firstChunk, int* doneOut) {
    const Dims& D = c.D;
    hipStream_t s = c.stream;
    const MgVecs V = mgVecs(c);
    double* st = c.scal + SC_BI;
    const int g = kgrid(D);
    const double tol2 = rtol * rtol;
    hipLaunchKernelGGL(k_mg_init, dim3(g), dim3(KBS), 0, s, D, (const int8_t*)c.st, b, V.r, V.rh, V.p0, V.p1, V.v, e,
                       c.red.part, st, tol2, maxit);
    const bool graph = c.probe.which == PROBE_NONE && mgGraph(c, V, e);
    double *pc = V.p0, *pn = V.p1;
    int it = 0, chunk = firstChunk;
    while (true) {
        if (graph) {
            for (int q = 0; q < (chunk + 1) / 2; ++q) HIPCHECK(hipGraphLaunch(c.mg.iter2, s));
        } else {
            for (int q = 0; q < chunk; ++q) {
                mgIteration(c, V, pc, pn, e);
                std::swap(pc, pn);
            }
        }
        HIPCHECK(hipMemcpyAsync(c.hscal + SC_BI, st, B_STRIDE * sizeof(double), hipMemcpyDeviceToHost, s));
        HIPCHECK(hipStreamSynchronize(s));
        it = (int)c.hscal[SC_BI + B_IT];
        if (c.hscal[SC_BI + B_DONE] != 0. || it >= maxit) break;
        // B_RR is the residual tested before the last queued iteration (it - 1 done): queue
        // what the observed rate of decrease still needs (an iteration queued past
        // convergence costs 55 launches that return at once)
        const double rr0 = c.hscal[SC_BI + B_RR], bb0 = c.hscal[SC_BI + B_BB];
        chunk = 4;
        if (it > 1 && rr0 > 0. && bb0 > 0. && rr0 < bb0) {
            const double rate = std::log(rr0 / bb0) / (it - 1);
            const double need = std::log(tol2 * bb0 / rr0) / rate - 1.;
            chunk = std::max(1, std::min(16, (int)std::ceil(need)));
        }
    }
    *doneOut = (int)c.hscal[SC_BI + B_DONE];
    return it;
}

// ---- the classical coarse solve of the learned preconditioner's two-grid cycle (its
// ablation): one V-cycle of this hierarchy below the fine level
__global__ void k_mg_lvl_in(Dims D, const float* src, float* dst) {   // unpadded -> padded interior
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= D.N) return;
    int i, j, k;
    dst[D.cell(e, i, j, k)] = src[e];
}
__global__ void k_mg_lvl_out(Dims D, const float* src, float* dst, const double* done) {
    if (done && *done != 0.) return;
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= D.N) return;
    int i, j, k;
    dst[e] = src[D.cell(e, i, j, k)];
}

bool mg_coarse_prepare(Ctx& c) {
    if (c.indefiniteP || c.hasIB || !mgPlan(c)) return false;
    const Dims& L1 = c.mg.L[1].D;
    if (L1.nI * 2 != c.D.nI || L1.nJ * 2 != c.D.nJ || L1.nK * 2 != c.D.nK) return false;
    mgBuild(c);
    return true;
}

void mg_coarse_vcycle(Ctx& c, const float* b1, float* z1, const double* done) {
    MgHier& H = c.mg;
    const Dims& D1 = H.L[1].D;
    // the level kernels read only the done slot B_DONE of their state pointer
    const double* st = done - B_DONE;
    hipLaunchKernelGGL(k_mg_lvl_in, dim3(lgrid(D1)), dim3(MBS), 0, c.stream, D1, b1, H.L[1].b);
    mgCoarseLevels(c, st);
    hipLaunchKernelGGL(k_mg_lvl_out, dim3(lgrid(D1)), dim3(MBS), 0, c.stream, D1, (const float*)H.L[1].z, z1, done);
}

int mg_bicgstab(Ctx& c, double shift, double rtol, int maxit, int* refs, double* relOut) {
    *refs = 0;
    if (c.indefiniteP) return -1;   // the coarse Laplacians do not model a reaction term
    if (!mgPlan(c)) return -1;
    mgBuild(c);   // on the FP32 copy of the matrix pcorr_solve made
    const int P = c.D.P;
    double* e = c.wm + 7 * (size_t)P;
    const double* r64 = c.wp + 6 * (size_t)P;
    // the initial guess, as in the learned solve: the least-residual combination of the
    // solutions this corrector reached in the two previous steps
    pcorr_apply_guess(c, shift);
    double rr, bb;
    pcorr_true_resid(c, shift, &rr, &bb);
    double rel = bb > 0. ? rr / bb : 0.;
    int its = 0, done = 0;
    // each round solves A e = r_true from e = 0 and adds e: the recurrence residual drifts
    // from the true one, and a breakdown restarts from the true residual
    while (rel > rtol && its < maxit && *refs < 30) {
        const double tol = std::min(0.5, 0.5 * rtol / rel);
        // the first round queues one iteration fewer than the last solve's first round needed
        // (one more read-back costs less than an iteration that returns at once: ~55 launches)
        const int first = *refs == 0 ? (c.mg.lastIts > 1 ? c.mg.lastIts - 1 : 4) : 2;
        const int it = mgRun(c, r64, e, tol, maxit - its, first, &done);
        if (*refs == 0) c.mg.lastIts = it;
        its += it;
        add_vec(c, c.xp, e, P);
        const double prev = rel;
        pcorr_true_resid(c, shift, &rr, &bb);
        rel = rr / bb;
        ++*refs;
        if (!(rel < prev)) break;
    }
    if (rel <= rtol) pcorr_save_guess(c);   // the next steps' guess for this corrector
    *relOut = rel;
    return its;
}

// ------------------------------------------------------------------ Coupled
// The Coupled system [M G; D L] [u; p] = b (Coupled.cpp:54-234) solved by BiCGStab right-
// preconditioned with the block lower triangle [diag M, 0; D, L]:
//   z_u = v_u / aP_M (each velocity component),   z_p = L^-1 (v_p - D z_u),
// where L^-1 is one V-cycle of the multigrid above on -L: the continuity rows' d-Laplacian
// with the sign of the p' storage (negative diagonal, positive faces).  Jacobi-preconditioned
// BiCGStab needs 73 iterations to 1e-8 on a 16^3 cavity and 134 on 32^3, growing with the
// grid; the triangle with an exact L^-1 needs 12 on both.  The velocity part of a
// preconditioned vector is recomputed from its FP64 vector where a kernel needs it; the
// pressure part is the cycle's FP32 output.

// the p' storage <- -L: diagonal and west, south, bottom faces (the assembly stores the
// boundary faces as zero); halo rows are written zero
__global__ void k_cpm_fill(Dims D, const double* __restrict__ A, double* __restrict__ Ap) {
    const int P = D.P;
    for (int q = blockIdx.x * blockDim.x + threadIdx.x; q < P; q += gridDim.x * blockDim.x) {
        const int k = q / D.sK, j = (q - k * D.sK) / D.sJ, i = q - k * D.sK - j * D.sJ;
        const bool in = i >= 1 && i <= D.nI && j >= 1 && j <= D.nJ && k >= 1 && k <= D.nK;
        Ap[q] = in ? -A[CP_L * P + q] : 0.;
        Ap[P + q] = in ? -A[(CP_L + 2) * P + q] : 0.;
        Ap[2 * P + q] = in ? -A[(CP_L + 4) * P + q] : 0.;
        Ap[3 * P + q] = in ? -A[(CP_L + 6) * P + q] : 0.;
    }
}

// r = b (minus shift on the continuity rows), rh = r, p = v = x = 0, dinv = 1 / aP_M on the
// velocity blocks; partials (rh, r), (r, r)
__global__ void k_cpm_init(Dims D, CpOp A, const double* b, double shift, double* r, double* rh, double* p0, double* p1,
                           double* v, double* x, double* dinv, double* part, double* st) {
    double acc[2] = {0., 0.};
    const int P = D.P;
    FOR_CELLS(D, q, p) {
        const double dm = 1. / A.A[CP_M * P + p];
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            const int o = c * P + p;
            const double bv = c == 3 ? b[o] - shift : b[o];
            if (c < 3) dinv[o] = dm;
            r[o] = bv; rh[o] = bv; p0[o] = 0.; p1[o] = 0.; v[o] = 0.; x[o] = 0.;
            acc[0] += bv * bv;
            acc[1] += bv * bv;
        }
    }
    writePartials<KBS, 2>(acc, part);
    if (BLOCK0) {
        st[B_IT] = 0.; st[B_DONE] = 0.; st[B_BB] = -1.; st[B_RHOPREV] = 1.; st[B_ALPHA] = 1.; st[B_OMEGA] = 1.; st[B_BRK] = 0.;
    }
}

// the right-hand side of the pressure block at cell p: D (y_u / aP_M) - y_p, y given at any
// cell by yf(component, cell)
template <class F>
__device__ __forceinline__ double cpmRhs(const CpOp& A, F yf, const double* __restrict__ dinv, int p) {
    const int P = A.P;
    const int o[7] = {0, 1, -1, A.sJ, -A.sJ, A.sK, -A.sK};
    double s = -yf(3, p);
#pragma unroll
    for (int c = 0; c < 3; ++c) {
#pragma unroll
        for (int m = 0; m < 7; ++m) {
            const int q = p + o[m];
            s += A.A[(CP_D + 7 * c + m) * P + p] * (yf(c, q) * dinv[c * P + q]);
        }
    }
    return s;
}

// [convergence on (rh, r), (r, r); beta] pn = r + beta (p - omega v); tb = D (pn_u / aP_M) - pn_p
__global__ void k_cpm_head_p(Dims D, CpOp A, const double* __restrict__ r, const double* __restrict__ pold,
                             double* __restrict__ pnew, const double* __restrict__ v, const double* __restrict__ dinv,
                             double* __restrict__ tb, const double* partIn, int nb, double* st, LoopCtl lc) {
    BiState<1> S;
    S.load(st);
    if (!S.any()) return;
    double beta[1], om[1];
    bool live[1];
    biaHead<1>(partIn, nb, st, lc, S, beta, om, live);
    if (!live[0]) return;
    const double bt = beta[0], w = om[0];
    const int P = D.P;
    auto pf = [&](int c, int q) { const int o = c * P + q; return r[o] + bt * (pold[o] - w * v[o]); };
    FOR_CELLS(D, q, p) {
#pragma unroll
        for (int c = 0; c < 4; ++c) pnew[c * P + p] = pf(c, p);
        tb[p] = cpmRhs(A, pf, dinv, p);
    }
}

// [alpha = rho / (rh, v)] s = r - alpha v; tb = D (s_u / aP_M) - s_p
__global__ void k_cpm_head_s(Dims D, CpOp A, const double* __restrict__ r, const double* __restrict__ v,
                             double* __restrict__ s, const double* __restrict__ dinv, double* __restrict__ tb,
                             const double* partIn, int nb, double* st) {
    BiState<1> S;
    S.load(st);
    if (!S.any()) return;
    double al[1];
    bool live[1];
    bibHead<1>(partIn, nb, st, S, al, live);
    if (!live[0]) return;
    const double a = al[0];
    const int P = D.P;
    auto sf = [&](int c, int q) { const int o = c * P + q; return r[o] - a * v[o]; };
    FOR_CELLS(D, q, p) {
#pragma unroll
        for (int c = 0; c < 4; ++c) s[c * P + p] = sf(c, p);
        tb[p] = cpmRhs(A, sf, dinv, p);
    }
}

// level 0 of the cycle on -L: z = two damped Jacobi sweeps from zero on b
__global__ void k_cpm_pre0(Dims D, Sym4<float> A, const float* __restrict__ wt, const double* __restrict__ b,
                           float* __restrict__ z, const double* st) {
    if (mgDone(st)) return;
    FOR_CELLS(D, q, p) { z[p] = preSmooth(A, wt, [&](int c) { return b[c]; }, p); }
}

// v = A ph (ph_u = pn_u / aP_M, ph_p the cycle's output); partial (rh, v)
__global__ void k_cpm_v(Dims D, CpOp A, const double* __restrict__ pn, const double* __restrict__ dinv,
                        const float* __restrict__ zp, const double* __restrict__ rh, double* __restrict__ v, double* partOut,
                        const double* st) {
    if (mgDone(st)) return;
    double acc[1] = {0.};
    const int P = D.P;
    auto ph = [&](int c, int q) { return c < 3 ? pn[c * P + q] * dinv[c * P + q] : (double)zp[q]; };
    FOR_CELLS(D, q, p) {
        double y[4];
        A.apply(ph, p, y);
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            v[c * P + p] = y[c];
            acc[0] += rh[c * P + p] * y[c];
        }
    }
    writePartials<KBS, 1>(acc, partOut);
}

// t = A sh; partials (t, s), (t, t)
__global__ void k_cpm_t(Dims D, CpOp A, const double* __restrict__ s, const double* __restrict__ dinv,
                        const float* __restrict__ zs, double* __restrict__ t, double* partOut, const double* st) {
    if (mgDone(st) || st[B_BRK] == 1.) return;
    double acc[2] = {0., 0.};
    const int P = D.P;
    auto sh = [&](int c, int q) { return c < 3 ? s[c * P + q] * dinv[c * P + q] : (double)zs[q]; };
    FOR_CELLS(D, q, p) {
        double y[4];
        A.apply(sh, p, y);
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            t[c * P + p] = y[c];
            acc[0] += y[c] * s[c * P + p];
            acc[1] += y[c] * y[c];
        }
    }
    writePartials<KBS, 2>(acc, partOut);
}

// [omega = (t, s) / (t, t)] x += alpha ph + omega sh; r = s - omega t; partials (rh, r), (r, r)
__global__ void k_cpm_x(Dims D, double* x, double* r, const double* rh, const double* pn, const double* s,
                        const double* t, const double* dinv, const float* zp, const float* zs, const double* partIn,
                        int nb, double* partOut, double* st) {
    const double sDone = st[B_DONE], sBrk = st[B_BRK], sAl = st[B_ALPHA], sRho = st[B_RHOCUR], sIt = st[B_IT];
    if (sDone != 0.) return;
    double tt[2];
    readPartials<KBS, 2>(partIn, nb, tt);
    const bool live = sBrk != 1.;   // the s head found (rh, v) = 0: no s, t this step
    const double om = tt[1] > 0. ? tt[0] / tt[1] : 0.;
    if (BLOCK0 && live) {
        st[B_OMEGA] = om; st[B_RHOPREV] = sRho; st[B_IT] = sIt + 1.;
        if (om == 0.) st[B_BRK] = 2.;
    }
    double acc[2] = {0., 0.};
    const int P = D.P;
    FOR_CELLS(D, q, p) {
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            const int o = c * P + p;
            double rn = r[o];
            if (live) {
                const double ph = c < 3 ? pn[o] * dinv[o] : (double)zp[p];
                const double sh = c < 3 ? s[o] * dinv[o] : (double)zs[p];
                x[o] += sAl * ph + om * sh;
                rn = s[o] - om * t[o];
                r[o] = rn;
            }
            acc[0] += rh[o] * rn;
            acc[1] += rn * rn;
        }
    }
    writePartials<KBS, 2>(acc, partOut);
}

bool cp_mg_setup(Ctx& c) {
    if (!mgPlan(c)) return false;
    hipLaunchKernelGGL(k_cpm_fill, dim3(blocksFor(c.D.P, MBS)), dim3(MBS), 0, c.stream, c.D, (const double*)c.cA, c.Ap);
    pcorr_make_fp32(c);
    mgBuild(c);
    return true;
}

int cp_mg_run(Ctx& c, const CpOp& A, const double* b, double shift, double* x, double rtol, int maxit, int firstChunk,
              int* doneOut) {
    const Dims& D = c.D;
    const size_t V = 4 * (size_t)D.P;
    double* r = c.cw; double* rh = r + V; double* p0 = rh + V; double* p1 = p0 + V; double* v = p1 + V;
    double* tb = v + V; double* s = tb + V; double* t = s + V; double* dinv = t + V;
    double* st = c.scal + SC_BI;
    double* partC = c.red.part;
    double* partA = c.red.part + MAX_RED_BLOCKS * RED_STRIDE;
    double* partB = c.red.part + 2 * MAX_RED_BLOCKS * RED_STRIDE;
    const Sym4<float> A32{c.Ap32, D.P, D.sJ, D.sK, OvView{nullptr, nullptr, nullptr}, c.st};
    float* zp = c.mg.z0[0];
    float* zs = c.mg.z0[1];
    const float* w0 = c.mg.w0;
    // up to 4 blocks of 16 waves per CU: the operator kernels hold 59-64 VGPRs, so a SIMD
    // takes 8 of their waves, and one block per CU left it at 4
    const int g = blocksFor(D.N, KBS, 4 * KGRID);
    hipStream_t q = c.stream;
    LoopCtl lc{rtol * rtol, maxit, 0.};
    hipLaunchKernelGGL(k_cpm_init, dim3(g), dim3(KBS), 0, q, D, A, b, shift, r, rh, p0, p1, v, x, dinv, partC, st);
    double *pc = p0, *pn = p1;
    int it = 0, chunk = std::max(firstChunk, 1), next = firstChunk > 4 ? 2 : 2 * chunk;
    while (true) {
        for (int k = 0; k < chunk; ++k) {
            hipLaunchKernelGGL(k_cpm_head_p, dim3(g), dim3(KBS), 0, q, D, A, (const double*)r, (const double*)pc, pn,
                               (const double*)v, (const double*)dinv, tb, (const double*)partC, g, st, lc);
            hipLaunchKernelGGL(k_cpm_pre0, dim3(g), dim3(KBS), 0, q, D, A32, w0, (const double*)tb, zp, (const double*)st);
            mgCycleRest(c, A32, tb, zp, st);
            hipLaunchKernelGGL(k_cpm_v, dim3(g), dim3(KBS), 0, q, D, A, (const double*)pn, (const double*)dinv,
                               (const float*)zp, (const double*)rh, v, partA, (const double*)st);
            hipLaunchKernelGGL(k_cpm_head_s, dim3(g), dim3(KBS), 0, q, D, A, (const double*)r, (const double*)v, s,
                               (const double*)dinv, tb, (const double*)partA, g, st);
            hipLaunchKernelGGL(k_cpm_pre0, dim3(g), dim3(KBS), 0, q, D, A32, w0, (const double*)tb, zs, (const double*)st);
            mgCycleRest(c, A32, tb, zs, st);
            hipLaunchKernelGGL(k_cpm_t, dim3(g), dim3(KBS), 0, q, D, A, (const double*)s, (const double*)dinv,
                               (const float*)zs, t, partB, (const double*)st);
            hipLaunchKernelGGL(k_cpm_x, dim3(g), dim3(KBS), 0, q, D, x, r, (const double*)rh, (const double*)pn,
                               (const double*)s, (const double*)t, (const double*)dinv, (const float*)zp, (const float*)zs,
                               (const double*)partB, g, partC, st);
            std::swap(pc, pn);
        }
        HIPCHECK(hipMemcpyAsync(c.hscal + SC_BI, st, B_STRIDE * sizeof(double), hipMemcpyDeviceToHost, q));
        HIPCHECK(hipStreamSynchronize(q));
        it = (int)c.hscal[SC_BI + B_IT];
        if (c.hscal[SC_BI + B_DONE] != 0. || it >= maxit) break;
        chunk = next;
        next = std::min(2 * next, 16);
    }
    *doneOut = (int)c.hscal[SC_BI + B_DONE];
    return it;
}

}  // namespace caffe
