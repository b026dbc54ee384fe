// fv_kernels.hip -- finite-volume assembly, Rhie-Chow, correction, gradients,
// boundary patches, continuity and Courant kernels of the SIMPLE/PISO step.
//
// Each kernel restates one loop of the reference with the same IEEE operation
// order (the library is built with -ffp-contract=off), so assembled
// coefficients match the reference arithmetic.  Face values of d, h, p, p' and u
// are not stored: an interior face value is recomputed from the two cell values
// with the volume weight g (FvScheme.tpp:12-40) exactly when the reference would
// read the stored one, and a boundary face value is the halo (patch) value,
// which Field::set*FacesFromPatch keeps equal to the face (Field.tpp:280-404).
#include "internal.hpp"

namespace caffe {

__device__ inline double dmin(double a, double b) { return (b < a) ? b : a; }  // std::min
__device__ inline double dmax(double a, double b) { return (a < b) ? b : a; }  // std::max

constexpr int BS = 256;

// ------------------------------------------------------------------ boundary patches
// PrimitiveBoundaryCondition::setBoundaries (PrimitiveBoundaryCondition.tpp:42-142):
// ZERO_GRADIENT and EMPTY faces copy the boundary cell plane into the patch
// (Field.tpp:364-404); FIXED faces keep their patch.
struct HaloArgs { double* f[12]; int nf; int types[6]; };

__global__ void k_halo(Dims D, HaloArgs a) {
    const int face = blockIdx.y;
    if (a.types[face] == BC_FIXED) return;
    long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    int n1, n2;
    if (face < 2) { n1 = D.nJ; n2 = D.nK; } else if (face < 4) { n1 = D.nI; n2 = D.nK; } else { n1 = D.nI; n2 = D.nJ; }
    if (t >= (long)n1 * n2) return;
    int x = (int)(t % n1), y = (int)(t / n1);
    long pin, pout;
    switch (face) {
    case 0: pin = D.pidx(D.nI - 1, x, y); pout = pin + 1; break;
    case 1: pin = D.pidx(0, x, y); pout = pin - 1; break;
    case 2: pin = D.pidx(x, D.nJ - 1, y); pout = pin + D.sJ; break;
    case 3: pin = D.pidx(x, 0, y); pout = pin - D.sJ; break;
    case 4: pin = D.pidx(x, y, D.nK - 1); pout = pin + D.sK; break;
    default: pin = D.pidx(x, y, 0); pout = pin - D.sK; break;
    }
    for (int q = 0; q < a.nf; ++q) a.f[q][pout] = a.f[q][pin];
}

void launch_halo(Ctx& c, double* const* f, int nf, const int* types) {
    const Dims& D = c.D;
    long mx = (long)D.nJ * D.nK;
    mx = std::max(mx, (long)D.nI * D.nK);
    mx = std::max(mx, (long)D.nI * D.nJ);
    for (int base = 0; base < nf; base += 12) {
        HaloArgs a;
        a.nf = std::min(12, nf - base);
        for (int q = 0; q < a.nf; ++q) a.f[q] = f[base + q];
        for (int q = 0; q < 6; ++q) a.types[q] = types[q];
        dim3 grid(blocksFor(mx, BS), 6);
        hipLaunchKernelGGL(k_halo, grid, dim3(BS), 0, c.stream, D, a);
    }
}

// ------------------------------------------------------------------ momentum assembly
// Simple::computeMomentum assembly loop (Simple.cpp:117-179) with the BC fold of
// PrimitiveBoundaryCondition::setImplicitBoundaryCoefficients (PrimitiveBoundaryCondition.tpp:144-255).
// The non-orthogonal term mu_f (grad u)_f . c_f is identically zero in the reference:
// gradUField_ faces are never assigned (Simple.cpp:153-158, 355), so in reference mode it
// is not evaluated.  In the corrected mode (caffe_case.nonOrth, general meshes) the face
// value of grad u is the volume-weighted interpolation of the two cell gradients on an
// interior face and the cell's own gradient on a domain face, and c is the face metric
// stored as seen from the low-side cell (the high-side cell's is its negation).
struct MomArgs {
    const int8_t* st;
    const double *mI, *mJ, *mK;
    const double *u[3], *u0[3], *gp[3];
    const double* gu[9];      // grad u cells, row-major (d u_r / d x_c at 3 r + c); corrected mode
    double* A; double* b; double* d;
    double rho, mu, om, dt;
    int steady;
    int nonOrth;
    int uT[6];
};

// (grad u)_f . c over one face in the corrected mode: the tensor times the vector by rows
// (Tensor3D.cpp:223-228), the tensor interpolated at an interior face (FvScheme.tpp:12-40)
__device__ inline void nonOrthFace(const Geo& g, const MomArgs& a, long p, long s, int dir, bool hi, bool bnd,
                                   double* out) {
    double T[9];
    if (bnd) {
        for (int q = 0; q < 9; ++q) T[q] = a.gu[q][p];
    } else if (hi) {
        const double w = G<false>::gw(g, dir, p + s);
        for (int q = 0; q < 9; ++q) T[q] = w * a.gu[q][p] + (1. - w) * a.gu[q][p + s];
    } else {
        const double w = G<false>::gw(g, dir, p);
        for (int q = 0; q < 9; ++q) T[q] = w * a.gu[q][p - s] + (1. - w) * a.gu[q][p];
    }
    const long f = hi ? p + s : p;
    V3 c = mk(g.cF[dir][0][f], g.cF[dir][1][f], g.cF[dir][2][f]);
    if (!hi) c = -c;
    for (int r = 0; r < 3; ++r) out[r] = T[3 * r] * c.x + T[3 * r + 1] * c.y + T[3 * r + 2] * c.z;
}

template <bool U>
__global__ void k_mom_assemble(Geo g, MomArgs a) {
    const Dims& D = g.D;
    long q = xcdCell();
    if (q >= D.N) return;
    int i, j, k;
    const long p = D.cell(q, i, j, k);
    const long P = D.P, sJ = D.sJ, sK = D.sK;
    const int8_t s = a.st[p];
    if (s != ST_ACTIVE) {
        if (s == ST_INACTIVE) {  // no row in the reference; inert identity row here
            a.A[p] = 1.;
            for (int c = 1; c < 7; ++c) a.A[c * P + p] = 0.;
            a.b[p] = 0.; a.b[P + p] = 0.; a.b[2 * P + p] = 0.;
        }
        return;
    }
    const double vol = G<U>::vol(g, p);
    const double a0P = a.steady ? 0. : a.rho * vol / a.dt;
    const double mE = a.mI[p + 1], mW = a.mI[p], mN = a.mJ[p + sJ], mS = a.mJ[p], mT = a.mK[p + sK], mB = a.mK[p];
    const double dE = G<U>::d(g, 0, p + 1, i == D.nI - 1), dW = G<U>::d(g, 0, p, i == 0);
    const double dN = G<U>::d(g, 1, p + sJ, j == D.nJ - 1), dS = G<U>::d(g, 1, p, j == 0);
    const double dT = G<U>::d(g, 2, p + sK, k == D.nK - 1), dB = G<U>::d(g, 2, p, k == 0);
    const double mu = a.mu;
    double A[7];
    A[1] = dmin(mE, 0.) - mu * dE;
    A[2] = -dmax(mW, 0.) - mu * dW;
    A[3] = dmin(mN, 0.) - mu * dN;
    A[4] = -dmax(mS, 0.) - mu * dS;
    A[5] = dmin(mT, 0.) - mu * dT;
    A[6] = -dmax(mB, 0.) - mu * dB;
    A[0] = (dmax(mE, 0.) + mu * dE - dmin(mW, 0.) + mu * dW + dmax(mN, 0.) + mu * dN - dmin(mS, 0.) + mu * dS
            + dmax(mT, 0.) + mu * dT - dmin(mB, 0.) + mu * dB + a0P) / a.om;
    double bx = a.u0[0][p] * a0P, by = a.u0[1][p] * a0P, bz = a.u0[2][p] * a0P;
    const double mv = -vol;
    bx += a.gp[0][p] * mv; by += a.gp[1][p] * mv; bz += a.gp[2][p] * mv;
    const double r = (1. - a.om) * A[0];
    bx += a.u[0][p] * r; by += a.u[1][p] * r; bz += a.u[2][p] * r;
    if (!U && a.nonOrth) {   // Simple.cpp:153-158, faces E, W, N, S, T, B
        const long str[3] = {1, sJ, sK};
        const int ijk[3] = {i, j, k}, nn[3] = {D.nI, D.nJ, D.nK};
#pragma unroll
        for (int f = 0; f < 6; ++f) {
            const int dir = f >> 1;
            const bool hi = (f & 1) == 0;
            const bool bnd = hi ? ijk[dir] == nn[dir] - 1 : ijk[dir] == 0;
            double t[3];
            nonOrthFace(g, a, p, str[dir], dir, hi, bnd, t);
            bx += mu * t[0]; by += mu * t[1]; bz += mu * t[2];
        }
    }
    // boundary fold, in the reference's face order E, W, N, S, T, B
    const long nb[7] = {0, p + 1, p - 1, p + sJ, p - sJ, p + sK, p - sK};
    const bool atB[7] = {false, i == D.nI - 1, i == 0, j == D.nJ - 1, j == 0, k == D.nK - 1, k == 0};
#pragma unroll
    for (int f = 1; f < 7; ++f) {
        if (!atB[f]) continue;
        if (a.uT[f - 1] == BC_FIXED) {
            const long h = nb[f];
            bx -= a.u[0][h] * A[f]; by -= a.u[1][h] * A[f]; bz -= a.u[2][h] * A[f];
        } else {
            A[0] += A[f]; A[f] = 0.;
        }
    }
    a.A[p] = A[0];
#pragma unroll
    for (int f = 1; f < 7; ++f) {
        // PETSc drops columns -1: off-domain and INACTIVE neighbours (IndexMap.cpp:47-78)
        const bool keep = !atB[f] && a.st[nb[f]] != ST_INACTIVE;
        a.A[f * P + p] = keep ? A[f] : 0.;
    }
    a.b[p] = bx; a.b[P + p] = by; a.b[2 * P + p] = bz;
    a.d[p] = vol / A[0];
}

void launch_mom_assemble(Ctx& c, double dt) {
    ++c.stepGen;   // a new time step (or outer iteration): the p' solves' guesses age by one
    c.pcSolvesThisStep = 0;
    MomArgs a;
    a.st = c.st; a.mI = c.mF[0]; a.mJ = c.mF[1]; a.mK = c.mF[2];
    for (int q = 0; q < 3; ++q) { a.u[q] = c.u[q]; a.u0[q] = c.u0[q]; a.gp[q] = c.gp[q]; }
    for (int q = 0; q < 9; ++q) a.gu[q] = c.gu[q];
    a.A = c.Am; a.b = c.bm; a.d = c.d;
    a.rho = c.cs.rho; a.mu = c.cs.mu; a.om = c.cs.alphaU; a.dt = dt; a.steady = c.cs.steady;
    a.nonOrth = (c.cs.nonOrth && !c.uniform && c.gu[0]) ? 1 : 0;
    for (int q = 0; q < 6; ++q) a.uT[q] = c.uT[q];
    int nb = blocksFor(c.D.N, BS);
    if (c.uniform) hipLaunchKernelGGL(k_mom_assemble<true>, dim3(nb), dim3(BS), 0, c.stream, c.geo, a);
    else hipLaunchKernelGGL(k_mom_assemble<false>, dim3(nb), dim3(BS), 0, c.stream, c.geo, a);
}

// Simple.cpp:189-206: u = x for non-inactive cells; h = u + d grad p (ACTIVE) or u (GHOST)
__global__ void k_mom_copyback(Dims D, const int8_t* st, const double* x, double* u0, double* u1, double* u2,
                               double* h0, double* h1, double* h2, const double* d,
                               const double* g0, const double* g1, const double* g2) {
    long q = xcdCell();
    if (q >= D.N) return;
    int i, j, k;
    const long p = D.cell(q, i, j, k);
    const int8_t s = st[p];
    if (s == ST_INACTIVE) return;
    const long P = D.P;
    double ux = x[p], uy = x[P + p], uz = x[2 * P + p];
    u0[p] = ux; u1[p] = uy; u2[p] = uz;
    if (s == ST_ACTIVE) {
        const double dd = d[p];
        h0[p] = ux + dd * g0[p]; h1[p] = uy + dd * g1[p]; h2[p] = uz + dd * g2[p];
    } else {
        h0[p] = ux; h1[p] = uy; h2[p] = uz;
    }
}

void launch_mom_copyback(Ctx& c) {
    hipLaunchKernelGGL(k_mom_copyback, dim3(blocksFor(c.D.N, BS)), dim3(BS), 0, c.stream, c.D, c.st, c.xm,
                       c.u[0], c.u[1], c.u[2], c.h[0], c.h[1], c.h[2], c.d, c.gp[0], c.gp[1], c.gp[2]);
}

// face value of a cell field on the high (E/N/T) face of cell p in direction dir
template <bool U>
__device__ inline double faceHi(const Geo& g, const double* phi, long p, long s, int dir, bool bnd) {
    if (bnd) return phi[p + s];
    const double w = G<U>::gw(g, dir, p + s);
    return w * phi[p] + (1. - w) * phi[p + s];
}
// face value on the low (W/S/B) face of cell p: the high face of the low neighbour
template <bool U>
__device__ inline double faceLo(const Geo& g, const double* phi, long p, long s, int dir, bool bnd) {
    if (bnd) return phi[p - s];
    const double w = G<U>::gw(g, dir, p);
    return w * phi[p - s] + (1. - w) * phi[p];
}

// ------------------------------------------------------------------ Rhie-Chow
// Simple::rhieChowInterpolateFaces (Simple.cpp:410-449)
struct RcArgs {
    const int8_t* st;
    const double *h[3], *d, *p;
    double *mF[3];
    double rho;
};

template <bool U>
__global__ void k_rhie_chow(Geo g, RcArgs a) {
    const Dims& D = g.D;
    long q = xcdCell();
    if (q >= D.N) return;
    int ijk[3];
    const long p = D.cell(q, ijk[0], ijk[1], ijk[2]);
    if (a.st[p] == ST_INACTIVE) return;
    const long str[3] = {1, D.sJ, D.sK};
    const int nn[3] = {D.nI, D.nJ, D.nK};
    const double pP = a.p[p];
#pragma unroll
    for (int dir = 0; dir < 3; ++dir) {
        const long s = str[dir];
        if (ijk[dir] == 0) {  // W/S/B domain face, area normal -S
            V3 hf = mk(a.h[0][p - s], a.h[1][p - s], a.h[2][p - s]);
            V3 Sw = -G<U>::S(g, dir, p);
            double m = -dot(hf, Sw) + a.d[p - s] * (a.p[p - s] - pP) * G<U>::d(g, dir, p, true);
            m *= a.rho;
            a.mF[dir][p] = m;
        }
        const bool bnd = ijk[dir] == nn[dir] - 1;
        V3 hf = mk(faceHi<U>(g, a.h[0], p, s, dir, bnd), faceHi<U>(g, a.h[1], p, s, dir, bnd), faceHi<U>(g, a.h[2], p, s, dir, bnd));
        const double df = faceHi<U>(g, a.d, p, s, dir, bnd);
        double m = dot(hf, G<U>::S(g, dir, p + s)) - df * (a.p[p + s] - pP) * G<U>::d(g, dir, p + s, bnd);
        m *= a.rho;
        a.mF[dir][p + s] = m;
    }
}

void launch_rhie_chow(Ctx& c) {
    RcArgs a;
    a.st = c.st;
    for (int q = 0; q < 3; ++q) { a.h[q] = c.h[q]; a.mF[q] = c.mF[q]; }
    a.d = c.d; a.p = c.p; a.rho = c.cs.rho;
    int nb = blocksFor(c.D.N, BS);
    if (c.uniform) hipLaunchKernelGGL(k_rhie_chow<true>, dim3(nb), dim3(BS), 0, c.stream, c.geo, a);
    else hipLaunchKernelGGL(k_rhie_chow<false>, dim3(nb), dim3(BS), 0, c.stream, c.geo, a);
}

// ------------------------------------------------------------------ pressure-correction assembly
// Simple::computePCorr assembly (Simple.cpp:231-269).  Stored symmetric: aP per cell
// and one coefficient per face (aI at entry p couples cells p-1 and p).  An interior
// face shared by two ACTIVE cells gets the same bits from both sides (d_f is one face
// value, dE(i) == dW(i+1)), so it is written once, by the high-side cell; a face
// toward a non-ACTIVE neighbour is written by the ACTIVE side.
struct PcArgs {
    const int8_t* st;
    const double *d, *mI, *mJ, *mK, *pc;
    double *A, *b;
    double rho;
    int pT[6];
};

template <bool U>
__global__ void k_pc_assemble(Geo g, PcArgs a) {
    const Dims& D = g.D;
    long q = xcdCell();
    if (q >= D.N) return;
    int i, j, k;
    const long p = D.cell(q, i, j, k);
    const long P = D.P, sJ = D.sJ, sK = D.sK;
    const int8_t s = a.st[p];
    if (s != ST_ACTIVE) {
        if (s == ST_INACTIVE) { a.A[p] = 1.; a.b[p] = 0.; }
        return;
    }
    const bool atB[7] = {false, i == D.nI - 1, i == 0, j == D.nJ - 1, j == 0, k == D.nK - 1, k == 0};
    const double rho = a.rho;
    double A[7];
    A[1] = rho * faceHi<U>(g, a.d, p, 1, 0, atB[1]) * G<U>::d(g, 0, p + 1, atB[1]);
    A[2] = rho * faceLo<U>(g, a.d, p, 1, 0, atB[2]) * G<U>::d(g, 0, p, atB[2]);
    A[3] = rho * faceHi<U>(g, a.d, p, sJ, 1, atB[3]) * G<U>::d(g, 1, p + sJ, atB[3]);
    A[4] = rho * faceLo<U>(g, a.d, p, sJ, 1, atB[4]) * G<U>::d(g, 1, p, atB[4]);
    A[5] = rho * faceHi<U>(g, a.d, p, sK, 2, atB[5]) * G<U>::d(g, 2, p + sK, atB[5]);
    A[6] = rho * faceLo<U>(g, a.d, p, sK, 2, atB[6]) * G<U>::d(g, 2, p, atB[6]);
    A[0] = -(A[1] + A[2] + A[3] + A[4] + A[5] + A[6]);
    double b = a.mI[p + 1] - a.mI[p] + a.mJ[p + sJ] - a.mJ[p] + a.mK[p + sK] - a.mK[p];
    const long nb[7] = {0, p + 1, p - 1, p + sJ, p - sJ, p + sK, p - sK};
#pragma unroll
    for (int f = 1; f < 7; ++f) {
        if (!atB[f]) continue;
        if (a.pT[f - 1] == BC_FIXED) b -= A[f] * a.pc[nb[f]];
        else { A[0] += A[f]; A[f] = 0.; }
    }
    a.A[p] = A[0];
    a.b[p] = b;
    // west/south/bottom faces: always written by this cell
    a.A[P + p] = (!atB[2] && a.st[p - 1] != ST_INACTIVE) ? A[2] : 0.;
    a.A[2 * P + p] = (!atB[4] && a.st[p - sJ] != ST_INACTIVE) ? A[4] : 0.;
    a.A[3 * P + p] = (!atB[6] && a.st[p - sK] != ST_INACTIVE) ? A[6] : 0.;
    // east/north/top faces: written here unless an ACTIVE neighbour writes them
    if (atB[1] || a.st[p + 1] != ST_ACTIVE) a.A[P + p + 1] = (!atB[1] && a.st[p + 1] != ST_INACTIVE) ? A[1] : 0.;
    if (atB[3] || a.st[p + sJ] != ST_ACTIVE) a.A[2 * P + p + sJ] = (!atB[3] && a.st[p + sJ] != ST_INACTIVE) ? A[3] : 0.;
    if (atB[5] || a.st[p + sK] != ST_ACTIVE) a.A[3 * P + p + sK] = (!atB[5] && a.st[p + sK] != ST_INACTIVE) ? A[5] : 0.;
}

void launch_pc_assemble(Ctx& c) {
    PcArgs a;
    a.st = c.st; a.d = c.d; a.mI = c.mF[0]; a.mJ = c.mF[1]; a.mK = c.mF[2]; a.pc = c.pc;
    a.A = c.Ap; a.b = c.bp; a.rho = c.cs.rho;
    for (int q = 0; q < 6; ++q) a.pT[q] = c.pT[q];
    int nb = blocksFor(c.D.N, BS);
    if (c.uniform) hipLaunchKernelGGL(k_pc_assemble<true>, dim3(nb), dim3(BS), 0, c.stream, c.geo, a);
    else hipLaunchKernelGGL(k_pc_assemble<false>, dim3(nb), dim3(BS), 0, c.stream, c.geo, a);
}

// Simple.cpp:277-289
__global__ void k_copy_noninactive(Dims D, const int8_t* st, const double* src, double* dst) {
    long q = xcdCell();
    if (q >= D.N) return;
    int i, j, k;
    const long p = D.cell(q, i, j, k);
    if (st[p] != ST_INACTIVE) dst[p] = src[p];
}
void launch_copy_cells(Ctx& c, const double* src, double* dst) {
    hipLaunchKernelGGL(k_copy_noninactive, dim3(blocksFor(c.D.N, BS)), dim3(BS), 0, c.stream, c.D, c.st, src, dst);
}
void launch_pc_copyback(Ctx& c) {
    hipLaunchKernelGGL(k_copy_noninactive, dim3(blocksFor(c.D.N, BS)), dim3(BS), 0, c.stream, c.D, c.st, c.xp, c.pc);
}

// ------------------------------------------------------------------ gradients
// FvScalarScheme::computeCellCenteredGradients, DIVERGENCE_THEOREM (FvScalarScheme.cpp:48-66),
// every cell: grad = (phi_E S_E + phi_W S_W + ... + phi_B S_B) / V
template <bool U>
__global__ void k_grad_scalar(Geo g, const double* phi, double* o0, double* o1, double* o2) {
    const Dims& D = g.D;
    long q = xcdCell();
    if (q >= D.N) return;
    int ijk[3];
    const long p = D.cell(q, ijk[0], ijk[1], ijk[2]);
    const long str[3] = {1, D.sJ, D.sK};
    const int nn[3] = {D.nI, D.nJ, D.nK};
    V3 acc = mk(0., 0., 0.);
    bool first = true;
#pragma unroll
    for (int dir = 0; dir < 3; ++dir) {
        const long s = str[dir];
        const double fe = faceHi<U>(g, phi, p, s, dir, ijk[dir] == nn[dir] - 1);
        const double fw = faceLo<U>(g, phi, p, s, dir, ijk[dir] == 0);
        V3 te = G<U>::S(g, dir, p + s) * fe;
        V3 tw = (-G<U>::S(g, dir, p)) * fw;
        if (first) { acc = te + tw; first = false; }
        else { acc = acc + te; acc = acc + tw; }
    }
    acc = acc / G<U>::vol(g, p);
    o0[p] = acc.x; o1[p] = acc.y; o2[p] = acc.z;
}

void launch_grad_scalar(Ctx& c, const double* phi, double* const* out) {
    int nb = blocksFor(c.D.N, BS);
    if (c.uniform) hipLaunchKernelGGL(k_grad_scalar<true>, dim3(nb), dim3(BS), 0, c.stream, c.geo, phi, out[0], out[1], out[2]);
    else hipLaunchKernelGGL(k_grad_scalar<false>, dim3(nb), dim3(BS), 0, c.stream, c.geo, phi, out[0], out[1], out[2]);
}

// FvVectorScheme::computeCellCenteredGradients, DIVERGENCE_THEOREM (FvVectorScheme.cpp:56-74):
// grad u = sum_f tensor(u_f, S_f) / V, row-major (tensor: Vector3D.cpp:176-181)
struct GuArgs { const double* u[3]; double* o[9]; };
template <bool U>
__global__ void k_grad_vector(Geo g, GuArgs a) {
    const Dims& D = g.D;
    long q = xcdCell();
    if (q >= D.N) return;
    int ijk[3];
    const long p = D.cell(q, ijk[0], ijk[1], ijk[2]);
    const long str[3] = {1, D.sJ, D.sK};
    const int nn[3] = {D.nI, D.nJ, D.nK};
    double t[9];
    bool first = true;
#pragma unroll
    for (int dir = 0; dir < 3; ++dir) {
        const long s = str[dir];
        const bool bh = ijk[dir] == nn[dir] - 1, bl = ijk[dir] == 0;
        double ue[3], uw[3];
        for (int c = 0; c < 3; ++c) { ue[c] = faceHi<U>(g, a.u[c], p, s, dir, bh); uw[c] = faceLo<U>(g, a.u[c], p, s, dir, bl); }
        V3 Se = G<U>::S(g, dir, p + s), Sw = -G<U>::S(g, dir, p);
        double te[9] = {ue[0] * Se.x, ue[0] * Se.y, ue[0] * Se.z, ue[1] * Se.x, ue[1] * Se.y, ue[1] * Se.z, ue[2] * Se.x, ue[2] * Se.y, ue[2] * Se.z};
        double tw[9] = {uw[0] * Sw.x, uw[0] * Sw.y, uw[0] * Sw.z, uw[1] * Sw.x, uw[1] * Sw.y, uw[1] * Sw.z, uw[2] * Sw.x, uw[2] * Sw.y, uw[2] * Sw.z};
        for (int c = 0; c < 9; ++c) {
            if (first) t[c] = te[c] + tw[c];
            else { t[c] = t[c] + te[c]; t[c] = t[c] + tw[c]; }
        }
        first = false;
    }
    const double v = G<U>::vol(g, p);
    for (int c = 0; c < 9; ++c) a.o[c][p] = t[c] / v;
}

void launch_grad_vector(Ctx& c) {
    GuArgs a;
    for (int q = 0; q < 3; ++q) a.u[q] = c.u[q];
    for (int q = 0; q < 9; ++q) a.o[q] = c.gu[q];
    int nb = blocksFor(c.D.N, BS);
    if (c.uniform) hipLaunchKernelGGL(k_grad_vector<true>, dim3(nb), dim3(BS), 0, c.stream, c.geo, a);
    else hipLaunchKernelGGL(k_grad_vector<false>, dim3(nb), dim3(BS), 0, c.stream, c.geo, a);
}

// ------------------------------------------------------------------ correct
// Simple::correct (Simple.cpp:303-346).  Each face is corrected once, by its low-side
// cell when that cell is not INACTIVE (W/S/B domain faces by the boundary cell, only
// for OUTLET faces).  The new fluxes go to a second buffer so every cell can form its
// own continuity sum from the corrected values of all six faces without a race.
struct CorrArgs {
    const int8_t* st;
    const double *mO[3];
    double *mN[3];
    double *mflow, *p, *u[3];
    const double *pc, *d, *gpc[3];
    double rho, ap;
    int mc[6];
};

template <bool U>
__global__ void k_correct(Geo g, CorrArgs a) {
    const Dims& D = g.D;
    long q = xcdCell();
    if (q >= D.N) return;
    int ijk[3];
    const long p = D.cell(q, ijk[0], ijk[1], ijk[2]);
    const long str[3] = {1, D.sJ, D.sK};
    const int nn[3] = {D.nI, D.nJ, D.nK};
    const int8_t s = a.st[p];
    if (s == ST_INACTIVE) {
#pragma unroll
        for (int dir = 0; dir < 3; ++dir) {
            a.mN[dir][p + str[dir]] = a.mO[dir][p + str[dir]];
            if (ijk[dir] == 0) a.mN[dir][p] = a.mO[dir][p];
        }
        return;
    }
    const double rho = a.rho, pcP = a.pc[p];
    double mHi[3], mLo[3];
#pragma unroll
    for (int dir = 0; dir < 3; ++dir) {
        const long s2 = str[dir];
        // high face (E/N/T), owned by this cell
        const bool bh = ijk[dir] == nn[dir] - 1;
        double mh = a.mO[dir][p + s2];
        if (!bh || a.mc[2 * dir]) {
            const double df = faceHi<U>(g, a.d, p, s2, dir, bh);
            mh += -rho * df * (a.pc[p + s2] - pcP) * G<U>::d(g, dir, p + s2, bh);
        }
        a.mN[dir][p + s2] = mh;
        mHi[dir] = mh;
        // low face (W/S/B)
        const bool bl = ijk[dir] == 0;
        double ml = a.mO[dir][p];
        if (bl) {
            if (a.mc[2 * dir + 1]) {
                const double df = a.d[p - s2];
                ml += rho * df * (a.pc[p - s2] - pcP) * G<U>::d(g, dir, p, true);
            }
            a.mN[dir][p] = ml;
        } else if (a.st[p - s2] != ST_INACTIVE) {
            // the low neighbour's high-face correction, evaluated as it evaluates it
            const double df = faceLo<U>(g, a.d, p, s2, dir, false);
            ml += -rho * df * (pcP - a.pc[p - s2]) * G<U>::d(g, dir, p, false);
        }
        mLo[dir] = ml;
    }
    a.mflow[p] = mHi[0] - mLo[0] + mHi[1] - mLo[1] + mHi[2] - mLo[2];
    a.p[p] += a.ap * pcP;
    const double md = -a.d[p];
    a.u[0][p] += a.gpc[0][p] * md;
    a.u[1][p] += a.gpc[1][p] * md;
    a.u[2][p] += a.gpc[2][p] * md;
}

void launch_correct(Ctx& c) {
    CorrArgs a;
    a.st = c.st;
    for (int q = 0; q < 3; ++q) { a.mO[q] = c.mF[q]; a.mN[q] = c.mF2[q]; a.u[q] = c.u[q]; a.gpc[q] = c.gpc[q]; }
    a.mflow = c.mflow; a.p = c.p; a.pc = c.pc; a.d = c.d; a.rho = c.cs.rho; a.ap = c.cs.alphaP;
    for (int q = 0; q < 6; ++q) a.mc[q] = c.massCorr[q] ? 1 : 0;
    int nb = blocksFor(c.D.N, BS);
    if (c.uniform) hipLaunchKernelGGL(k_correct<true>, dim3(nb), dim3(BS), 0, c.stream, c.geo, a);
    else hipLaunchKernelGGL(k_correct<false>, dim3(nb), dim3(BS), 0, c.stream, c.geo, a);
    for (int q = 0; q < 3; ++q) std::swap(c.mF[q], c.mF2[q]);
}

// ------------------------------------------------------------------ continuity and Courant
// Simple::computeContinuityError (Simple.cpp:358-380) in the reference's order: scan the
// ACTIVE cells (i fastest, then j, then k) with e = 0; whenever |m| > e, e = |m| / V.
// The scan is a sequence of rounds, each finding the first ACTIVE cell at or after
// `start` with |m| > e (a grid-wide first-index search) and then moving e and start on;
// the number of rounds is the number of updates the sequential scan makes, which is
// small (after the first hit e = |m| / V exceeds |m| by the inverse cell volume).
// State cs: [0] e, [1] start, [2] done.
template <bool U>
__global__ void k_cq_scan(Geo g, const int8_t* st, const double* mflow, const double* cs, double* part) {
    const Dims& D = g.D;
    if (cs[2] != 0.) return;   // not written by this kernel: every thread of a block agrees
    const double e = cs[0];
    const long start = (long)cs[1];
    const long chunk = ((long)D.N + gridDim.x - 1) / gridDim.x;
    const long b0 = (long)blockIdx.x * chunk, b1 = min((long)D.N, b0 + chunk);
    long best = D.N;
    for (long q = max(b0, start) + threadIdx.x; q < b1; q += blockDim.x) {
        int i, j, k;
        const long p = D.cell((int)q, i, j, k);
        if (st[p] == ST_ACTIVE && fabs(mflow[p]) > e) { best = q; break; }
    }
    for (int o = 32; o > 0; o >>= 1) { const long t = __shfl_xor(best, o, 64); best = t < best ? t : best; }
    __shared__ long sm[BS / 64];
    if ((threadIdx.x & 63) == 0) sm[threadIdx.x >> 6] = best;
    __syncthreads();
    if (threadIdx.x == 0) {
        long m = D.N;
        for (int w = 0; w < BS / 64; ++w) m = sm[w] < m ? sm[w] : m;
        part[blockIdx.x * RED_STRIDE] = (double)m;
    }
}
template <bool U>
__global__ void k_cq_step(Geo g, const double* mflow, const double* part, int nb, double* cs) {
    const Dims& D = g.D;
    if (cs[2] != 0.) return;
    long best = D.N;
    for (int b = threadIdx.x; b < nb; b += blockDim.x) {
        const long v = (long)part[b * RED_STRIDE];
        best = v < best ? v : best;
    }
    for (int o = 32; o > 0; o >>= 1) { const long t = __shfl_xor(best, o, 64); best = t < best ? t : best; }
    __shared__ long sm[BS / 64];
    if ((threadIdx.x & 63) == 0) sm[threadIdx.x >> 6] = best;
    __syncthreads();
    if (threadIdx.x == 0) {
        long m = D.N;
        for (int w = 0; w < BS / 64; ++w) m = sm[w] < m ? sm[w] : m;
        if (m >= D.N) {
            cs[2] = 1.;
        } else {
            int i, j, k;
            const long p = D.cell((int)m, i, j, k);
            cs[0] = fabs(mflow[p]) / G<U>::vol(g, p);
            cs[1] = (double)(m + 1);
        }
    }
}
__global__ void k_cq_init(double* cs) {
    if (threadIdx.x == 0) { cs[0] = 0.; cs[1] = 0.; cs[2] = 0.; }
}

template <bool U, int MODE>
__global__ void k_maxreduce(Geo g, const int8_t* st, const double* mflow, const double* u0, const double* u1,
                            const double* u2, double dt, double* part) {
    const Dims& D = g.D;
    __shared__ double lds[BS / 64];
    double mx = 0.;
    for (long q = (long)blockIdx.x * blockDim.x + threadIdx.x; q < D.N; q += (long)gridDim.x * blockDim.x) {
        int i, j, k;
        const long p = D.cell(q, i, j, k);
        double v;
        if (MODE == 0) {  // continuity max over ACTIVE cells
            if (st[p] != ST_ACTIVE) continue;
            v = fabs(mflow[p]) / G<U>::vol(g, p);
        } else {  // Simple::computeCourantNumber (Simple.cpp:382-408), every cell
            V3 xc = G<U>::xc(g, i, j, k, p);
            V3 rI = (G<U>::fc(g, 0, i + 1, j, k, p + 1) - xc) - (G<U>::fc(g, 0, i, j, k, p) - xc);
            V3 rJ = (G<U>::fc(g, 1, i, j + 1, k, p + D.sJ) - xc) - (G<U>::fc(g, 1, i, j, k, p) - xc);
            V3 rK = (G<U>::fc(g, 2, i, j, k + 1, p + D.sK) - xc) - (G<U>::fc(g, 2, i, j, k, p) - xc);
            V3 uu = mk(u0[p], u1[p], u2[p]);
            double uI = fabs(dot(uu, unitv(rI))), uJ = fabs(dot(uu, unitv(rJ))), uK = fabs(dot(uu, unitv(rK)));
            v = dt * (uI / mag(rI) + uJ / mag(rJ) + uK / mag(rK));
        }
        mx = v > mx ? v : mx;
    }
    double r = blockMax<BS>(mx, lds);
    if (threadIdx.x == 0) part[blockIdx.x] = r;
}

__global__ void k_max_final(const double* part, int n, double* out) {
    __shared__ double lds[BS / 64];
    double mx = 0.;
    for (int q = threadIdx.x; q < n; q += BS) mx = part[q] > mx ? part[q] : mx;
    double r = blockMax<BS>(mx, lds);
    if (threadIdx.x == 0) *out = r;
}

static double readScalar(Ctx& c, int slot) {
    HIPCHECK(hipMemcpyAsync(c.hscal + slot, c.scal + slot, sizeof(double), hipMemcpyDeviceToHost, c.stream));
    HIPCHECK(hipStreamSynchronize(c.stream));
    return c.hscal[slot];
}

double continuity_quirk(Ctx& c) {
    double* cs = c.scal + SC_MISC;
    const int nb = blocksFor(c.D.N, BS, 1024);
    hipLaunchKernelGGL(k_cq_init, dim3(1), dim3(64), 0, c.stream, cs);
    for (long round = 0; round <= (long)c.D.N; round += 4) {
        for (int r = 0; r < 4; ++r) {
            if (c.uniform) {
                hipLaunchKernelGGL(k_cq_scan<true>, dim3(nb), dim3(BS), 0, c.stream, c.geo, c.st, c.mflow, (const double*)cs, c.red.part);
                hipLaunchKernelGGL(k_cq_step<true>, dim3(1), dim3(BS), 0, c.stream, c.geo, c.mflow, (const double*)c.red.part, nb, cs);
            } else {
                hipLaunchKernelGGL(k_cq_scan<false>, dim3(nb), dim3(BS), 0, c.stream, c.geo, c.st, c.mflow, (const double*)cs, c.red.part);
                hipLaunchKernelGGL(k_cq_step<false>, dim3(1), dim3(BS), 0, c.stream, c.geo, c.mflow, (const double*)c.red.part, nb, cs);
            }
        }
        HIPCHECK(hipMemcpyAsync(c.hscal + SC_MISC, cs, 3 * sizeof(double), hipMemcpyDeviceToHost, c.stream));
        HIPCHECK(hipStreamSynchronize(c.stream));
        if (c.hscal[SC_MISC + 2] != 0.) break;
    }
    return c.hscal[SC_MISC];
}

// the max reduction of mode 0 (continuity) or 1 (Courant) into the scalar slot
static void maxReduceLaunch(Ctx& c, int mode, double dt, int slot) {
    const int nb = blocksFor(c.D.N, BS, 2048);
    double* part = c.red.part;
    if (mode == 0) {
        if (c.uniform) hipLaunchKernelGGL((k_maxreduce<true, 0>), dim3(nb), dim3(BS), 0, c.stream, c.geo, c.st, c.mflow, c.u[0], c.u[1], c.u[2], dt, part);
        else hipLaunchKernelGGL((k_maxreduce<false, 0>), dim3(nb), dim3(BS), 0, c.stream, c.geo, c.st, c.mflow, c.u[0], c.u[1], c.u[2], dt, part);
    } else {
        if (c.uniform) hipLaunchKernelGGL((k_maxreduce<true, 1>), dim3(nb), dim3(BS), 0, c.stream, c.geo, c.st, c.mflow, c.u[0], c.u[1], c.u[2], dt, part);
        else hipLaunchKernelGGL((k_maxreduce<false, 1>), dim3(nb), dim3(BS), 0, c.stream, c.geo, c.st, c.mflow, c.u[0], c.u[1], c.u[2], dt, part);
    }
    hipLaunchKernelGGL(k_max_final, dim3(1), dim3(BS), 0, c.stream, part, nb, c.scal + slot);
}
static double maxReduce(Ctx& c, int mode, double dt) {
    maxReduceLaunch(c, mode, dt, SC_MISC + 1);
    return readScalar(c, SC_MISC + 1);
}
double continuity_max(Ctx& c) { return maxReduce(c, 0, 0.); }
double courant(Ctx& c, double dt) { return maxReduce(c, 1, dt); }
// a slot no other reduction of the step uses (the continuity scan holds SC_MISC + 0..2)
constexpr int SC_COURANT = SC_MISC + 16;
void courant_queue(Ctx& c, double dt) {
    maxReduceLaunch(c, 1, dt, SC_COURANT);
    HIPCHECK(hipMemcpyAsync(c.hscal + SC_COURANT, c.scal + SC_COURANT, sizeof(double), hipMemcpyDeviceToHost, c.stream));
}
double courant_queued(const Ctx& c) { return c.hscal[SC_COURANT]; }

// ------------------------------------------------------------------ Coupled
// Coupled::computeMomentumAndPressure (Coupled.cpp:54-280): the four rows of a cell (u, v,
// w momentum and continuity) in planes of P doubles, neighbour order own, E, W, N, S, T, B
// (the column order of Coupled.cpp:118-132 and 179-214).  A coefficient whose column is an
// off-domain neighbour (IndexMap gives -1, and PETSc drops it) is stored as zero.
// Interpolation weights are the reference's: gE = V_E / (V_E + V_P), gW = V_W / (V_W + V_P)
// in the cell's own interior, 0 on a domain face (HexaFvmMesh.cpp:884-907).
template <bool U>
__device__ __forceinline__ void cpWeights(const Geo& g, long p, const long* nb, const bool* atB, double* w) {
    const double v = G<U>::vol(g, p);
#pragma unroll
    for (int f = 0; f < 6; ++f) {
        if (atB[f + 1]) { w[f] = 0.; continue; }
        const double vn = G<U>::vol(g, nb[f + 1]);
        w[f] = vn / (vn + v);
    }
}
// outward area normals of the six faces, E W N S T B
template <bool U>
__device__ __forceinline__ void cpNormals(const Geo& g, long p, const Dims& D, V3* S) {
    S[0] = G<U>::S(g, 0, p + 1); S[1] = -G<U>::S(g, 0, p);
    S[2] = G<U>::S(g, 1, p + D.sJ); S[3] = -G<U>::S(g, 1, p);
    S[4] = G<U>::S(g, 2, p + D.sK); S[5] = -G<U>::S(g, 2, p);
}
__device__ __forceinline__ double comp(const V3& v, int c) { return c == 0 ? v.x : (c == 1 ? v.y : v.z); }

struct CpArgs {
    const double *mI, *mJ, *mK;
    const double *u[3], *u0[3], *p, *d, *gp[3];
    double *A, *b, *dOut;
    double rho, mu, om, dt;
    int steady;
    int uT[6], pT[6];
};

// Coupled.cpp:66-144: momentum rows (14 coefficients: a_P / omega and a_nb shared by the
// components, then the pressure coefficients of component c) and d = V / a_P
template <bool U>
__global__ void k_cp_momentum(Geo g, CpArgs a) {
    const Dims& D = g.D;
    long q = xcdCell();
    if (q >= D.N) return;
    int i, j, k;
    const long p = D.cell(q, i, j, k);
    const long P = D.P, sJ = D.sJ, sK = D.sK;
    const double vol = G<U>::vol(g, p);
    const double a0P = a.steady ? 0. : a.rho * vol / a.dt;
    const double mE = a.mI[p + 1], mW = a.mI[p], mN = a.mJ[p + sJ], mS = a.mJ[p], mT = a.mK[p + sK], mB = a.mK[p];
    const double dE = G<U>::d(g, 0, p + 1, i == D.nI - 1), dW = G<U>::d(g, 0, p, i == 0);
    const double dN = G<U>::d(g, 1, p + sJ, j == D.nJ - 1), dS = G<U>::d(g, 1, p, j == 0);
    const double dT = G<U>::d(g, 2, p + sK, k == D.nK - 1), dB = G<U>::d(g, 2, p, k == 0);
    const double mu = a.mu;
    double A[7];
    A[0] = (dmax(mE, 0.) + mu * dE - dmin(mW, 0.) + mu * dW + dmax(mN, 0.) + mu * dN - dmin(mS, 0.) + mu * dS
            + dmax(mT, 0.) + mu * dT - dmin(mB, 0.) + mu * dB + a0P) / a.om;
    A[1] = dmin(mE, 0.) - mu * dE;
    A[2] = -dmax(mW, 0.) - mu * dW;
    A[3] = dmin(mN, 0.) - mu * dN;
    A[4] = -dmax(mS, 0.) - mu * dS;
    A[5] = dmin(mT, 0.) - mu * dT;
    A[6] = -dmax(mB, 0.) - mu * dB;
    double bu[3];
    const double r = (1. - a.om) * A[0];
#pragma unroll
    for (int c = 0; c < 3; ++c) { bu[c] = a0P * a.u0[c][p]; bu[c] += r * a.u[c][p]; }
    const long nb[7] = {p, p + 1, p - 1, p + sJ, p - sJ, p + sK, p - sK};
    const bool atB[7] = {false, i == D.nI - 1, i == 0, j == D.nJ - 1, j == 0, k == D.nK - 1, k == 0};
#pragma unroll
    for (int f = 1; f < 7; ++f) {
        if (!atB[f]) continue;
        if (a.uT[f - 1] == BC_FIXED) {
#pragma unroll
            for (int c = 0; c < 3; ++c) bu[c] -= A[f] * a.u[c][nb[f]];
        } else {
            A[0] += A[f]; A[f] = 0.;
        }
    }
#pragma unroll
    for (int m = 0; m < 7; ++m) a.A[(CP_M + m) * P + p] = atB[m] ? 0. : A[m];
    V3 S[6];
    cpNormals<U>(g, p, D, S);
    double w[6];
    cpWeights<U>(g, p, nb, atB, w);
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        double G7[7];
        G7[0] = w[0] * comp(S[0], c) + w[1] * comp(S[1], c) + w[2] * comp(S[2], c) + w[3] * comp(S[3], c)
              + w[4] * comp(S[4], c) + w[5] * comp(S[5], c);
#pragma unroll
        for (int f = 0; f < 6; ++f) G7[f + 1] = (1. - w[f]) * comp(S[f], c);
#pragma unroll
        for (int f = 1; f < 7; ++f) {
            if (!atB[f]) continue;
            if (a.pT[f - 1] == BC_FIXED) bu[c] -= G7[f] * a.p[nb[f]];
            else { G7[0] += G7[f]; G7[f] = 0.; }
        }
#pragma unroll
        for (int m = 0; m < 7; ++m) a.A[(CP_G + 7 * c + m) * P + p] = atB[m] ? 0. : G7[m];
        a.b[c * P + p] = bu[c];
    }
    a.dOut[p] = vol / A[0];
}

// Coupled.cpp:150-226: the continuity row.  Velocity coefficients rho g S(c) with the
// velocity fold (its right-hand side keeps component c of the vector fold), the d-Laplacian
// of p, and the right-hand side -sum rho d (grad p)_f . S_f with the pressure fold.  d faces:
// interpolated inside, the patch on a domain face; (grad p)_f: interpolated inside, zero on a
// domain face (never assigned there).
template <bool U>
__global__ void k_cp_continuity(Geo g, CpArgs a) {
    const Dims& D = g.D;
    long q = xcdCell();
    if (q >= D.N) return;
    int i, j, k;
    const long p = D.cell(q, i, j, k);
    const long P = D.P, sJ = D.sJ, sK = D.sK;
    const long nb[7] = {p, p + 1, p - 1, p + sJ, p - sJ, p + sK, p - sK};
    const bool atB[7] = {false, i == D.nI - 1, i == 0, j == D.nJ - 1, j == 0, k == D.nK - 1, k == 0};
    const double rho = a.rho;
    V3 S[6];
    cpNormals<U>(g, p, D, S);
    double w[6];
    cpWeights<U>(g, p, nb, atB, w);
    double bp = 0.;
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        double Dc[7];
        Dc[0] = rho * w[0] * comp(S[0], c) + rho * w[1] * comp(S[1], c) + rho * w[2] * comp(S[2], c)
              + rho * w[3] * comp(S[3], c) + rho * w[4] * comp(S[4], c) + rho * w[5] * comp(S[5], c);
#pragma unroll
        for (int f = 0; f < 6; ++f) Dc[f + 1] = rho * (1. - w[f]) * comp(S[f], c);
        double bu = 0.;
#pragma unroll
        for (int f = 1; f < 7; ++f) {
            if (!atB[f]) continue;
            if (a.uT[f - 1] == BC_FIXED) bu -= Dc[f] * a.u[c][nb[f]];
            else { Dc[0] += Dc[f]; Dc[f] = 0.; }
        }
        bp += bu;
#pragma unroll
        for (int m = 0; m < 7; ++m) a.A[(CP_D + 7 * c + m) * P + p] = atB[m] ? 0. : Dc[m];
    }
    const long str[3] = {1, sJ, sK};
    double df[6], dd[6];
    V3 gf[6];
#pragma unroll
    for (int dir = 0; dir < 3; ++dir) {
        const long s = str[dir];
        const bool bh = atB[2 * dir + 1], bl = atB[2 * dir + 2];
        df[2 * dir] = faceHi<U>(g, a.d, p, s, dir, bh);
        df[2 * dir + 1] = faceLo<U>(g, a.d, p, s, dir, bl);
        dd[2 * dir] = G<U>::d(g, dir, p + s, bh);
        dd[2 * dir + 1] = G<U>::d(g, dir, p, bl);
        gf[2 * dir] = bh ? mk(0., 0., 0.)
                         : mk(faceHi<U>(g, a.gp[0], p, s, dir, false), faceHi<U>(g, a.gp[1], p, s, dir, false),
                              faceHi<U>(g, a.gp[2], p, s, dir, false));
        gf[2 * dir + 1] = bl ? mk(0., 0., 0.)
                             : mk(faceLo<U>(g, a.gp[0], p, s, dir, false), faceLo<U>(g, a.gp[1], p, s, dir, false),
                                  faceLo<U>(g, a.gp[2], p, s, dir, false));
    }
    double L[7];
    L[0] = rho * df[0] * dd[0] + rho * df[1] * dd[1] + rho * df[2] * dd[2] + rho * df[3] * dd[3] + rho * df[4] * dd[4]
         + rho * df[5] * dd[5];
#pragma unroll
    for (int f = 0; f < 6; ++f) L[f + 1] = -rho * df[f] * dd[f];
    bp -= rho * df[0] * dot(gf[0], S[0]) + rho * df[1] * dot(gf[1], S[1]) + rho * df[2] * dot(gf[2], S[2])
        + rho * df[3] * dot(gf[3], S[3]) + rho * df[4] * dot(gf[4], S[4]) + rho * df[5] * dot(gf[5], S[5]);
#pragma unroll
    for (int f = 1; f < 7; ++f) {
        if (!atB[f]) continue;
        if (a.pT[f - 1] == BC_FIXED) bp -= L[f] * a.p[nb[f]];
        else { L[0] += L[f]; L[f] = 0.; }
    }
#pragma unroll
    for (int m = 0; m < 7; ++m) a.A[(CP_L + m) * P + p] = atB[m] ? 0. : L[m];
    a.b[3 * P + p] = bp;
}

static CpArgs cpArgs(Ctx& c, double dt) {
    CpArgs a;
    a.mI = c.mF[0]; a.mJ = c.mF[1]; a.mK = c.mF[2];
    for (int q = 0; q < 3; ++q) { a.u[q] = c.u[q]; a.u0[q] = c.u0[q]; a.gp[q] = c.gp[q]; }
    a.p = c.p; a.d = c.d; a.A = c.cA; a.b = c.cb; a.dOut = c.d;
    a.rho = c.cs.rho; a.mu = c.cs.mu; a.om = c.cs.alphaU; a.dt = dt; a.steady = c.cs.steady;
    for (int q = 0; q < 6; ++q) { a.uT[q] = c.uT[q]; a.pT[q] = c.pT[q]; }
    return a;
}

// assembly of the whole system: momentum rows and d, the d patches (zero gradient), the
// continuity rows
void launch_cp_assemble(Ctx& c, double dt) {
    const CpArgs a = cpArgs(c, dt);
    const int nb = blocksFor(c.D.N, BS);
    if (c.uniform) hipLaunchKernelGGL(k_cp_momentum<true>, dim3(nb), dim3(BS), 0, c.stream, c.geo, a);
    else hipLaunchKernelGGL(k_cp_momentum<false>, dim3(nb), dim3(BS), 0, c.stream, c.geo, a);
    int zg[6] = {BC_ZERO_GRADIENT, BC_ZERO_GRADIENT, BC_ZERO_GRADIENT, BC_ZERO_GRADIENT, BC_ZERO_GRADIENT, BC_ZERO_GRADIENT};
    double* dd[1] = {c.d};
    launch_halo(c, dd, 1, zg);
    if (c.uniform) hipLaunchKernelGGL(k_cp_continuity<true>, dim3(nb), dim3(BS), 0, c.stream, c.geo, a);
    else hipLaunchKernelGGL(k_cp_continuity<false>, dim3(nb), dim3(BS), 0, c.stream, c.geo, a);
}

// Coupled.cpp:239-255: u and p from the solution
__global__ void k_cp_copyback(Dims D, const double* x, double* u0, double* u1, double* u2, double* p) {
    long q = xcdCell();
    if (q >= D.N) return;
    int i, j, k;
    const long c = D.cell(q, i, j, k);
    const long P = D.P;
    u0[c] = x[c]; u1[c] = x[P + c]; u2[c] = x[2 * P + c]; p[c] = x[3 * P + c];
}

// Coupled.cpp:258: the divergence-theorem gradient of p from its face values, which this
// solver never interpolates: zero on interior faces, the patch on domain faces
template <bool U>
__global__ void k_cp_gradp(Geo g, const double* phi, double* o0, double* o1, double* o2) {
    const Dims& D = g.D;
    long q = xcdCell();
    if (q >= D.N) return;
    int ijk[3];
    const long p = D.cell(q, ijk[0], ijk[1], ijk[2]);
    const long str[3] = {1, D.sJ, D.sK};
    const int nn[3] = {D.nI, D.nJ, D.nK};
    V3 acc = mk(0., 0., 0.);
    bool first = true;
#pragma unroll
    for (int dir = 0; dir < 3; ++dir) {
        const long s = str[dir];
        const double fe = ijk[dir] == nn[dir] - 1 ? phi[p + s] : 0.;
        const double fw = ijk[dir] == 0 ? phi[p - s] : 0.;
        V3 te = G<U>::S(g, dir, p + s) * fe;
        V3 tw = (-G<U>::S(g, dir, p)) * fw;
        if (first) { acc = te + tw; first = false; }
        else { acc = acc + te; acc = acc + tw; }
    }
    acc = acc / G<U>::vol(g, p);
    o0[p] = acc.x; o1[p] = acc.y; o2[p] = acc.z;
}

// Coupled.cpp:261-273: h = u + d grad p
__global__ void k_cp_h(Dims D, const double* u0, const double* u1, const double* u2, const double* d,
                       const double* g0, const double* g1, const double* g2, double* h0, double* h1, double* h2) {
    long q = xcdCell();
    if (q >= D.N) return;
    int i, j, k;
    const long p = D.cell(q, i, j, k);
    const double dd = d[p];
    h0[p] = u0[p] + dd * g0[p]; h1[p] = u1[p] + dd * g1[p]; h2[p] = u2[p] + dd * g2[p];
}

// Coupled::rhieChowInterpolateFaces (Coupled.cpp:282-314): as Simple's without the density
// factor, and on a W/S/B domain face the pressure term has the sign of the E/N/T faces
template <bool U>
__global__ void k_cp_rhie_chow(Geo g, RcArgs a) {
    const Dims& D = g.D;
    long q = xcdCell();
    if (q >= D.N) return;
    int ijk[3];
    const long p = D.cell(q, ijk[0], ijk[1], ijk[2]);
    const long str[3] = {1, D.sJ, D.sK};
    const int nn[3] = {D.nI, D.nJ, D.nK};
    const double pP = a.p[p];
#pragma unroll
    for (int dir = 0; dir < 3; ++dir) {
        const long s = str[dir];
        if (ijk[dir] == 0) {
            V3 hf = mk(a.h[0][p - s], a.h[1][p - s], a.h[2][p - s]);
            V3 Sw = -G<U>::S(g, dir, p);
            a.mF[dir][p] = -dot(hf, Sw) - a.d[p - s] * (a.p[p - s] - pP) * G<U>::d(g, dir, p, true);
        }
        const bool bnd = ijk[dir] == nn[dir] - 1;
        V3 hf = mk(faceHi<U>(g, a.h[0], p, s, dir, bnd), faceHi<U>(g, a.h[1], p, s, dir, bnd), faceHi<U>(g, a.h[2], p, s, dir, bnd));
        const double df = faceHi<U>(g, a.d, p, s, dir, bnd);
        a.mF[dir][p + s] = dot(hf, G<U>::S(g, dir, p + s)) - df * (a.p[p + s] - pP) * G<U>::d(g, dir, p + s, bnd);
    }
}

// Coupled.cpp:239-279: u and p from the solution, the p patches, grad p and its faces, h and
// its patches, the u patches, and the Rhie-Chow faces
void launch_cp_update(Ctx& c, const double* x) {
    const Dims& D = c.D;
    const int nb = blocksFor(D.N, BS);
    hipLaunchKernelGGL(k_cp_copyback, dim3(nb), dim3(BS), 0, c.stream, D, x, c.u[0], c.u[1], c.u[2], c.p);
    double* pf[1] = {c.p};
    launch_halo(c, pf, 1, c.pT);
    if (c.uniform) hipLaunchKernelGGL(k_cp_gradp<true>, dim3(nb), dim3(BS), 0, c.stream, c.geo, c.p, c.gp[0], c.gp[1], c.gp[2]);
    else hipLaunchKernelGGL(k_cp_gradp<false>, dim3(nb), dim3(BS), 0, c.stream, c.geo, c.p, c.gp[0], c.gp[1], c.gp[2]);
    hipLaunchKernelGGL(k_cp_h, dim3(nb), dim3(BS), 0, c.stream, D, c.u[0], c.u[1], c.u[2], c.d, c.gp[0], c.gp[1], c.gp[2],
                       c.h[0], c.h[1], c.h[2]);
    double* uh[6] = {c.h[0], c.h[1], c.h[2], c.u[0], c.u[1], c.u[2]};
    launch_halo(c, uh, 6, c.uT);
    RcArgs a;
    a.st = c.st;
    for (int q = 0; q < 3; ++q) { a.h[q] = c.h[q]; a.mF[q] = c.mF[q]; }
    a.d = c.d; a.p = c.p; a.rho = 1.;
    if (c.uniform) hipLaunchKernelGGL(k_cp_rhie_chow<true>, dim3(nb), dim3(BS), 0, c.stream, c.geo, a);
    else hipLaunchKernelGGL(k_cp_rhie_chow<false>, dim3(nb), dim3(BS), 0, c.stream, c.geo, a);
}

// Coupled::computeContinuityError (Coupled.cpp:316-338): max over every cell of
// |mE - mW + mN - mS + mT - mB| / V
template <bool U>
__global__ void k_cp_continuity_max(Geo g, const double* mI, const double* mJ, const double* mK, double* part) {
    const Dims& D = g.D;
    __shared__ double lds[BS / 64];
    double mx = 0.;
    for (long q = (long)blockIdx.x * blockDim.x + threadIdx.x; q < D.N; q += (long)gridDim.x * blockDim.x) {
        int i, j, k;
        const long p = D.cell(q, i, j, k);
        const double v = fabs(mI[p + 1] - mI[p] + mJ[p + D.sJ] - mJ[p] + mK[p + D.sK] - mK[p]) / G<U>::vol(g, p);
        mx = v > mx ? v : mx;
    }
    double r = blockMax<BS>(mx, lds);
    if (threadIdx.x == 0) part[blockIdx.x] = r;
}

double cp_continuity(Ctx& c) {
    const int nb = blocksFor(c.D.N, BS, 2048);
    if (c.uniform) hipLaunchKernelGGL(k_cp_continuity_max<true>, dim3(nb), dim3(BS), 0, c.stream, c.geo, c.mF[0], c.mF[1], c.mF[2], c.red.part);
    else hipLaunchKernelGGL(k_cp_continuity_max<false>, dim3(nb), dim3(BS), 0, c.stream, c.geo, c.mF[0], c.mF[1], c.mF[2], c.red.part);
    hipLaunchKernelGGL(k_max_final, dim3(1), dim3(BS), 0, c.stream, c.red.part, nb, c.scal + SC_MISC + 1);
    return readScalar(c, SC_MISC + 1);
}

// ------------------------------------------------------------------ Diffusion
// Diffusion::solve assembly loop (Diffusion.cpp:40-75): a_nb = d_f (the face metric),
// a0 = -(sum a_nb) + a0P with a0P = V/dt when unsteady (the sign as in the reference),
// b = a0P phi; the fold of PrimitiveBoundaryCondition::setImplicitBoundaryCoefficients
// (PrimitiveBoundaryCondition.tpp:144-255) in east..bottom order.  Stored like the p'
// system: aP and each cell's low-face coefficient, zero on domain faces.
struct DiffArgs { const double* phi; double* A; double* b; double dt; int unsteady; int pT[6]; };

template <bool U>
__global__ void k_diff_assemble(Geo g, DiffArgs a) {
    const Dims& D = g.D;
    long q = xcdCell();
    if (q >= D.N) return;
    int i, j, k;
    const long p = D.cell(q, i, j, k);
    const long P = D.P, sJ = D.sJ, sK = D.sK;
    const bool atB[7] = {false, i == D.nI - 1, i == 0, j == D.nJ - 1, j == 0, k == D.nK - 1, k == 0};
    double A[7];
    A[1] = G<U>::d(g, 0, p + 1, atB[1]);
    A[2] = G<U>::d(g, 0, p, atB[2]);
    A[3] = G<U>::d(g, 1, p + sJ, atB[3]);
    A[4] = G<U>::d(g, 1, p, atB[4]);
    A[5] = G<U>::d(g, 2, p + sK, atB[5]);
    A[6] = G<U>::d(g, 2, p, atB[6]);
    const double a0P = a.unsteady ? G<U>::vol(g, p) / a.dt : 0.;
    A[0] = -(A[1] + A[2] + A[3] + A[4] + A[5] + A[6]) + a0P;
    double b = a0P * a.phi[p];
    const long nb[7] = {0, p + 1, p - 1, p + sJ, p - sJ, p + sK, p - sK};
#pragma unroll
    for (int f = 1; f < 7; ++f) {
        if (!atB[f]) continue;
        if (a.pT[f - 1] == BC_FIXED) b -= A[f] * a.phi[nb[f]];
        else { A[0] += A[f]; A[f] = 0.; }
    }
    a.A[p] = A[0];
    a.b[p] = b;
    a.A[P + p] = atB[2] ? 0. : A[2];
    a.A[2 * P + p] = atB[4] ? 0. : A[4];
    a.A[3 * P + p] = atB[6] ? 0. : A[6];
    if (atB[1]) a.A[P + p + 1] = 0.;
    if (atB[3]) a.A[2 * P + p + sJ] = 0.;
    if (atB[5]) a.A[3 * P + p + sK] = 0.;
}

void launch_diff_assemble(Ctx& c, double dt) {
    DiffArgs a;
    a.phi = c.p; a.A = c.Ap; a.b = c.bp; a.dt = dt; a.unsteady = c.cs.steady ? 0 : 1;
    for (int f = 0; f < 6; ++f) a.pT[f] = c.pT[f];
    const int nb = blocksFor(c.D.N, BS);
    if (c.uniform) hipLaunchKernelGGL(k_diff_assemble<true>, dim3(nb), dim3(BS), 0, c.stream, c.geo, a);
    else hipLaunchKernelGGL(k_diff_assemble<false>, dim3(nb), dim3(BS), 0, c.stream, c.geo, a);
}

// FvScalarScheme::extrapolateInteriorFaces (FvScalarScheme.cpp:70-160) followed by the
// divergence-theorem gradient (FvScalarScheme.cpp:48-66), one domain.  An interior face
// takes 0.5 (phi_L + grad_L . r_L + phi_H + grad_H . r_H) from the gradient of the
// interpolated faces (g1), clamped to [min, max] of its low-side cell and that cell's six
// neighbours (the reference sets it while visiting the low-side cell); domain faces keep
// the patch value.
template <bool U>
__device__ __forceinline__ double cellMax(const double* f, long p, const Dims& D, bool mx) {
    const long o[6] = {1, -1, D.sJ, -D.sJ, D.sK, -D.sK};
    double v = f[p];
#pragma unroll
    for (int q = 0; q < 6; ++q) v = mx ? fmax(v, f[p + o[q]]) : fmin(v, f[p + o[q]]);
    return v;
}
template <bool U>
__device__ __forceinline__ double extrapFace(const Geo& g, const double* f, const double* const* g1, int dir, long lo,
                                             int il, int jl, int kl) {
    // the face between cell lo (indices il, jl, kl) and its high-side neighbour
    const Dims& D = g.D;
    const long s = dir == 0 ? 1 : (dir == 1 ? D.sJ : D.sK);
    const long hi = lo + s;
    const int ih = il + (dir == 0), jh = jl + (dir == 1), kh = kl + (dir == 2);
    const V3 fc = G<U>::fc(g, dir, ih, jh, kh, hi);
    const V3 rl = fc - G<U>::xc(g, il, jl, kl, lo), rh = fc - G<U>::xc(g, ih, jh, kh, hi);
    const double dl = g1[0][lo] * rl.x + g1[1][lo] * rl.y + g1[2][lo] * rl.z;
    const double dh = g1[0][hi] * rh.x + g1[1][hi] * rh.y + g1[2][hi] * rh.z;
    const double v = 0.5 * (f[lo] + dl + f[hi] + dh);
    const double up = cellMax<U>(f, lo, D, true), dn = cellMax<U>(f, lo, D, false);
    return fmax(fmin(up, v), dn);
}
struct ExtrapArgs { const double* phi; const double* g1[3]; double* o[3]; };
template <bool U>
__global__ void k_grad_extrap(Geo g, ExtrapArgs a) {
    const Dims& D = g.D;
    long q = xcdCell();
    if (q >= D.N) return;
    int ijk[3];
    const long p = D.cell(q, ijk[0], ijk[1], ijk[2]);
    const long str[3] = {1, D.sJ, D.sK};
    const int nn[3] = {D.nI, D.nJ, D.nK};
    V3 acc = mk(0., 0., 0.);
    bool first = true;
#pragma unroll
    for (int dir = 0; dir < 3; ++dir) {
        const long s = str[dir];
        const double fe = ijk[dir] < nn[dir] - 1 ? extrapFace<U>(g, a.phi, a.g1, dir, p, ijk[0], ijk[1], ijk[2]) : a.phi[p + s];
        const double fw = ijk[dir] > 0
            ? extrapFace<U>(g, a.phi, a.g1, dir, p - s, ijk[0] - (dir == 0), ijk[1] - (dir == 1), ijk[2] - (dir == 2))
            : a.phi[p - s];
        V3 te = G<U>::S(g, dir, p + s) * fe;
        V3 tw = (-G<U>::S(g, dir, p)) * fw;
        if (first) { acc = te + tw; first = false; }
        else { acc = acc + te; acc = acc + tw; }
    }
    acc = acc / G<U>::vol(g, p);
    a.o[0][p] = acc.x; a.o[1][p] = acc.y; a.o[2][p] = acc.z;
}
void launch_grad_extrap(Ctx& c, const double* phi, double* const* g1, double* const* out) {
    ExtrapArgs a;
    a.phi = phi;
    for (int q = 0; q < 3; ++q) { a.g1[q] = g1[q]; a.o[q] = out[q]; }
    const int nb = blocksFor(c.D.N, BS);
    if (c.uniform) hipLaunchKernelGGL(k_grad_extrap<true>, dim3(nb), dim3(BS), 0, c.stream, c.geo, a);
    else hipLaunchKernelGGL(k_grad_extrap<false>, dim3(nb), dim3(BS), 0, c.stream, c.geo, a);
}

// ||A x - b||_inf over the rows (Diffusion.cpp:127-130: scale(-1, b); res = A x + b)
__global__ void k_diff_resid(Dims D, const double* A, const double* b, const double* x, double* part) {
    __shared__ double lds[BS / 64];
    const long P = D.P, sJ = D.sJ, sK = D.sK;
    double mx = 0.;
    for (long q = (long)blockIdx.x * blockDim.x + threadIdx.x; q < D.N; q += (long)gridDim.x * blockDim.x) {
        int i, j, k;
        const long p = D.cell(q, i, j, k);
        const double ax = A[p] * x[p] + A[P + p] * x[p - 1] + A[P + p + 1] * x[p + 1] + A[2 * P + p] * x[p - sJ]
                        + A[2 * P + p + sJ] * x[p + sJ] + A[3 * P + p] * x[p - sK] + A[3 * P + p + sK] * x[p + sK];
        const double r = fabs(ax + (-b[p]));
        mx = r > mx ? r : mx;
    }
    const double r = blockMax<BS>(mx, lds);
    if (threadIdx.x == 0) part[blockIdx.x] = r;
}
double diff_resid_inf(Ctx& c, const double* x) {
    const int nb = std::min(blocksFor(c.D.N, BS), 1024);
    hipLaunchKernelGGL(k_diff_resid, dim3(nb), dim3(BS), 0, c.stream, c.D, (const double*)c.Ap, (const double*)c.bp, x, c.red.part);
    hipLaunchKernelGGL(k_max_final, dim3(1), dim3(BS), 0, c.stream, c.red.part, nb, c.scal + SC_MISC + 1);
    return readScalar(c, SC_MISC + 1);
}

// plain copy of n doubles: the reference kernel that calibrates HBM byte counters
__global__ void k_copy(const double* __restrict__ a, double* __restrict__ b, int n) {
    int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q < n) b[q] = a[q];
}
double time_copy(Ctx& c, double* a, double* b, int n, int reps) {
    hipLaunchKernelGGL(k_copy, dim3(blocksFor(n, BS)), dim3(BS), 0, c.stream, a, b, n);
    HIPCHECK(hipEventRecord(c.ev0, c.stream));
    for (int r = 0; r < reps; ++r) hipLaunchKernelGGL(k_copy, dim3(blocksFor(n, BS)), dim3(BS), 0, c.stream, a, b, n);
    HIPCHECK(hipEventRecord(c.ev1, c.stream));
    HIPCHECK(hipEventSynchronize(c.ev1));
    float ms = 0.f;
    HIPCHECK(hipEventElapsedTime(&ms, c.ev0, c.ev1));
    return (double)ms / reps;
}

__global__ void k_fill(double* a, double v, long n) {
    long q = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (q < n) a[q] = v;
}
void launch_fill(Ctx& c, double* a, double v, long n) {
    hipLaunchKernelGGL(k_fill, dim3(blocksFor(n, BS)), dim3(BS), 0, c.stream, a, v, n);
}

// IbPiso.cpp:74-80: every SOLID cell (INACTIVE, and the IB/GHOST cells picked from
// them afterwards) starts the time step with u = 0, p = 0
__global__ void k_zero_inactive(Dims D, const int8_t* st, double* u0, double* u1, double* u2, double* p) {
    long q = xcdCell();
    if (q >= D.N) return;
    int i, j, k;
    const long pp = D.cell(q, i, j, k);
    if (st[pp] != ST_ACTIVE) { u0[pp] = 0.; u1[pp] = 0.; u2[pp] = 0.; p[pp] = 0.; }
}
void launch_set_inactive_fields(Ctx& c) {
    hipLaunchKernelGGL(k_zero_inactive, dim3(blocksFor(c.D.N, BS)), dim3(BS), 0, c.stream, c.D, c.st, c.u[0], c.u[1], c.u[2], c.p);
}

}  // namespace caffe
