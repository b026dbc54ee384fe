// ib.hip -- ghost-cell immersed boundary for a sphere (IbPiso.cpp:65-190).
//
// computeIbField: SOLID if |x_c - c| <= r (Sphere.cpp:49-55), FLUID otherwise; a SOLID
// cell with a FLUID 6-neighbour (off-domain neighbours read the FLUID patch) becomes
// IB/GHOST and gets a matrix row, other SOLID cells are INACTIVE (no row).
// computeIbCoeffs: image point = 2 * nearestIntersect(x_c) - x_c; the 8 cell centres
// enclosing it (HexaFvmMesh.cpp:378-418) give trilinear weights beta = b M^-1 over the
// monomials [1, xyz, xy, yz, xz, x, y, z] (Interpolation.cpp:135-181); the momentum row
// is (1 + beta_self) u_G + sum beta_m u_m = 0 and the pressure row is -p_G + sum beta_m p_m
// = 0, where beta_self survives in the pressure row only when the ghost is the 8th
// enclosing cell (the reset at IbPiso.cpp:170).  Right-hand sides of ghost rows stay 0.
//
// For a fixed sphere on a fixed mesh both results are the same every time step and
// inner iteration, so they are computed once and kept until the case changes.
#include "internal.hpp"

#include <vector>

namespace caffe {

__device__ inline V3 qCentroid(const V3& a, const V3& b, const V3& c, const V3& d) {  // Geometry.cpp:14-20
    V3 d1 = a + 0.5 * (c - a);
    V3 d2 = b + 0.5 * (d - b);
    return d1 + 0.5 * (d2 - d1);
}
__device__ inline V3 qNormal(const V3& a, const V3& b, const V3& c) {  // Geometry.cpp:22-25
    return 0.5 * (cross(b - a, c - a) + cross(b - a, c - a));
}
// Geometry.cpp:178-256
__device__ inline bool insideHexD(const V3& pt, const V3* p) {
    const int F[6][4] = {{1, 2, 6, 5}, {0, 4, 7, 3}, {2, 3, 7, 6}, {0, 1, 5, 4}, {4, 5, 6, 7}, {0, 3, 2, 1}};
    for (int f = 0; f < 6; ++f) {
        V3 xc = qCentroid(p[F[f][0]], p[F[f][1]], p[F[f][2]], p[F[f][3]]);
        V3 n = qNormal(p[F[f][0]], p[F[f][1]], p[F[f][2]]);
        if (dot(xc - pt, n) < 0.) return false;
    }
    return true;
}

struct IbArgs {
    double r, cx, cy, cz;
    int8_t* st;
    int* kind;  // 0 FLUID, 1 IB, 2 SOLID (IbPiso.h:44)
};

template <bool U>
__global__ void k_ib_solid(Geo g, IbArgs a) {
    const Dims& D = g.D;
    long q = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= D.N) return;
    int i, j, k;
    long p = D.cell(q, i, j, k);
    V3 x = G<U>::xc(g, i, j, k, p);
    V3 cen = mk(a.cx, a.cy, a.cz);
    bool solid = mag(x - cen) <= a.r;
    a.kind[p] = solid ? 2 : 0;
    a.st[p] = solid ? ST_INACTIVE : ST_ACTIVE;
}

__global__ void k_ib_ghost(Dims D, IbArgs a) {
    long q = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= D.N) return;
    int i, j, k;
    long p = D.cell(q, i, j, k);
    if (a.kind[p] != 2) return;
    auto fluid = [&](int ii, int jj, int kk, long pp) {
        if (ii < 0 || jj < 0 || kk < 0 || ii >= D.nI || jj >= D.nJ || kk >= D.nK) return true;  // FLUID patch
        return a.kind[pp] == 0;
    };
    if (fluid(i + 1, j, k, p + 1) || fluid(i - 1, j, k, p - 1) || fluid(i, j + 1, k, p + D.sJ) || fluid(i, j - 1, k, p - D.sJ)
        || fluid(i, j, k + 1, p + D.sK) || fluid(i, j, k - 1, p - D.sK)) {
        a.st[p] = ST_GHOST;  // kind stays SOLID for the neighbour test of other cells
    }
}

template <bool U>
__device__ inline void hexOfCell(const Geo& g, const double* nodes, int i, int j, int k, V3* t) {
    if (U) {
        const double x0 = i * g.hx, x1 = (i + 1) * g.hx, y0 = j * g.hy, y1 = (j + 1) * g.hy, z0 = k * g.hz, z1 = (k + 1) * g.hz;
        t[0] = mk(x0, y0, z0); t[1] = mk(x1, y0, z0); t[2] = mk(x1, y1, z0); t[3] = mk(x0, y1, z0);
        t[4] = mk(x0, y0, z1); t[5] = mk(x1, y0, z1); t[6] = mk(x1, y1, z1); t[7] = mk(x0, y1, z1);
    } else {
        const Dims& D = g.D;
        auto nd = [&](int a, int b, int c) { long q = ((long)c * (D.nJ + 1) + b) * (D.nI + 1) + a; return mk(nodes[3 * q], nodes[3 * q + 1], nodes[3 * q + 2]); };
        t[0] = nd(i, j, k); t[1] = nd(i + 1, j, k); t[2] = nd(i + 1, j + 1, k); t[3] = nd(i, j + 1, k);
        t[4] = nd(i, j, k + 1); t[5] = nd(i + 1, j, k + 1); t[6] = nd(i + 1, j + 1, k + 1); t[7] = nd(i, j + 1, k + 1);
    }
}

// Gauss-Jordan inverse with partial pivoting, then beta = b M^-1
__device__ inline void trilinear(const V3* pts, const V3& ip, double* beta) {
    double M[8][8], Inv[8][8];
    for (int i = 0; i < 8; ++i) {
        double x = pts[i].x, y = pts[i].y, z = pts[i].z;
        M[i][0] = 1.; M[i][1] = x * y * z; M[i][2] = x * y; M[i][3] = y * z; M[i][4] = x * z; M[i][5] = x; M[i][6] = y; M[i][7] = z;
        for (int j = 0; j < 8; ++j) Inv[i][j] = (i == j) ? 1. : 0.;
    }
    for (int col = 0; col < 8; ++col) {
        int piv = col;
        for (int r = col + 1; r < 8; ++r) if (fabs(M[r][col]) > fabs(M[piv][col])) piv = r;
        if (piv != col) for (int j = 0; j < 8; ++j) { double t = M[piv][j]; M[piv][j] = M[col][j]; M[col][j] = t; t = Inv[piv][j]; Inv[piv][j] = Inv[col][j]; Inv[col][j] = t; }
        double dd = M[col][col];
        for (int j = 0; j < 8; ++j) { M[col][j] /= dd; Inv[col][j] /= dd; }
        for (int r = 0; r < 8; ++r) {
            if (r == col) continue;
            double f = M[r][col];
            if (f == 0.) continue;
            for (int j = 0; j < 8; ++j) { M[r][j] -= f * M[col][j]; Inv[r][j] -= f * Inv[col][j]; }
        }
    }
    double x = ip.x, y = ip.y, z = ip.z;
    double b[8] = {1., x * y * z, x * y, y * z, x * z, x, y, z};
    for (int j = 0; j < 8; ++j) {
        double s = 0.;
        for (int i = 0; i < 8; ++i) s += b[i] * Inv[i][j];
        beta[j] = s;
    }
}

struct CoefArgs {
    int n;
    const int* cell;
    int* cols;
    double *valM, *valP;
    double *Am, *Ap;     // diagonal of ghost rows for the Jacobi preconditioner
    const int8_t* st;
    const double* nodes;
    double r, cx, cy, cz;
    int* fail;
};

template <bool U>
__global__ void k_ib_coeffs(Geo g, CoefArgs a) {
    const Dims& D = g.D;
    int gi = blockIdx.x * blockDim.x + threadIdx.x;
    if (gi >= a.n) return;
    const int p = a.cell[gi];
    int rem = p;
    const int k = (int)(rem / D.sK) - 1; rem %= D.sK;
    const int j = (int)(rem / D.sJ) - 1;
    const int i = (int)(rem % D.sJ) - 1;
    V3 xc = G<U>::xc(g, i, j, k, p);
    V3 cen = mk(a.cx, a.cy, a.cz);
    V3 bp = cen + a.r * unitv(xc - cen);   // Sphere::nearestIntersect
    V3 ip = 2. * bp - xc;
    // HexaFvmMesh::locateCell: first cell in k,j,i order whose hex holds the point
    int i0 = -1, j0 = -1, k0 = -1;
    V3 t[8];
    if (U) {
        int ci = (int)floor(ip.x / g.hx), cj = (int)floor(ip.y / g.hy), ck = (int)floor(ip.z / g.hz);
        for (int kk = ck - 1; kk <= ck + 1 && i0 < 0; ++kk)
            for (int jj = cj - 1; jj <= cj + 1 && i0 < 0; ++jj)
                for (int ii = ci - 1; ii <= ci + 1 && i0 < 0; ++ii) {
                    if (ii < 0 || jj < 0 || kk < 0 || ii >= D.nI || jj >= D.nJ || kk >= D.nK) continue;
                    hexOfCell<U>(g, a.nodes, ii, jj, kk, t);
                    if (insideHexD(ip, t)) { i0 = ii; j0 = jj; k0 = kk; }
                }
    } else {
        for (int kk = 0; kk < D.nK && i0 < 0; ++kk)
            for (int jj = 0; jj < D.nJ && i0 < 0; ++jj)
                for (int ii = 0; ii < D.nI && i0 < 0; ++ii) {
                    hexOfCell<U>(g, a.nodes, ii, jj, kk, t);
                    if (insideHexD(ip, t)) { i0 = ii; j0 = jj; k0 = kk; }
                }
    }
    if (i0 < 0) { atomicAdd(a.fail, 1); return; }
    // locateEnclosingCells: first of the 8 cell-centre hexes around it that holds the point
    int bi = -1, bj = -1, bk = -1;
    for (int kk = k0 - 1; kk <= k0 && bi < 0; ++kk)
        for (int jj = j0 - 1; jj <= j0 && bi < 0; ++jj)
            for (int ii = i0 - 1; ii <= i0 && bi < 0; ++ii) {
                if (ii < 0 || jj < 0 || kk < 0 || ii + 1 >= D.nI || jj + 1 >= D.nJ || kk + 1 >= D.nK) { bi = -2; break; }
                t[0] = G<U>::xc(g, ii, jj, kk, D.pidx(ii, jj, kk)); t[1] = G<U>::xc(g, ii + 1, jj, kk, D.pidx(ii + 1, jj, kk));
                t[2] = G<U>::xc(g, ii + 1, jj + 1, kk, D.pidx(ii + 1, jj + 1, kk)); t[3] = G<U>::xc(g, ii, jj + 1, kk, D.pidx(ii, jj + 1, kk));
                t[4] = G<U>::xc(g, ii, jj, kk + 1, D.pidx(ii, jj, kk + 1)); t[5] = G<U>::xc(g, ii + 1, jj, kk + 1, D.pidx(ii + 1, jj, kk + 1));
                t[6] = G<U>::xc(g, ii + 1, jj + 1, kk + 1, D.pidx(ii + 1, jj + 1, kk + 1)); t[7] = G<U>::xc(g, ii, jj + 1, kk + 1, D.pidx(ii, jj + 1, kk + 1));
                if (insideHexD(ip, t)) { bi = ii; bj = jj; bk = kk; }
            }
    if (bi < 0) { atomicAdd(a.fail, 1); return; }
    const int di[8] = {0, 1, 1, 0, 0, 1, 1, 0}, dj[8] = {0, 0, 1, 1, 0, 0, 1, 1}, dk[8] = {0, 0, 0, 0, 1, 1, 1, 1};
    int ii[8], jj[8], kk[8];
    V3 pts[8];
    for (int m = 0; m < 8; ++m) {
        ii[m] = bi + di[m]; jj[m] = bj + dj[m]; kk[m] = bk + dk[m];
        pts[m] = G<U>::xc(g, ii[m], jj[m], kk[m], D.pidx(ii[m], jj[m], kk[m]));
    }
    double beta[8];
    trilinear(pts, ip, beta);
    // IbPiso.cpp:148-185
    double vm[9], vp[9];
    int cl[9];
    for (int m = 0; m < 9; ++m) { vm[m] = 0.; vp[m] = 0.; cl[m] = -1; }
    vm[0] = 1.;
    cl[0] = p;
    for (int m = 0; m < 8; ++m) {
        const int pm = D.pidx(ii[m], jj[m], kk[m]);
        if (ii[m] == i && jj[m] == j && kk[m] == k) { vm[0] += beta[m]; cl[m + 1] = -1; }
        else { vm[m + 1] = beta[m]; cl[m + 1] = a.st[pm] == ST_INACTIVE ? -1 : pm; }
    }
    for (int m = 0; m < 8; ++m) {
        vp[0] = -1.;
        if (ii[m] == i && jj[m] == j && kk[m] == k) vp[0] += beta[m];
        else vp[m + 1] = beta[m];
    }
    for (int m = 0; m < 9; ++m) {
        a.cols[9L * gi + m] = cl[m];
        a.valM[9L * gi + m] = vm[m];
        a.valP[9L * gi + m] = vp[m];
    }
    a.Am[p] = vm[0];
    a.Ap[p] = vp[0];
}

static void ensureOverlay(Ctx& c, int n) {
    auto al = [&](size_t bytes) { void* ptr = nullptr; HIPCHECK(hipMalloc(&ptr, bytes)); HIPCHECK(hipMemset(ptr, 0, bytes)); c.allocs.push_back(ptr); return ptr; };
    if (!c.ov.rowOf) c.ov.rowOf = (int*)al(sizeof(int) * c.D.P);
    if (n > 0) {
        c.ov.cell = (int*)al(sizeof(int) * n);
        c.ov.cols = (int*)al(sizeof(int) * 9 * n);
        c.ov.valM = (double*)al(sizeof(double) * 9 * n);
        c.ov.valP = (double*)al(sizeof(double) * 9 * n);
        c.ov.valP32 = (float*)al(sizeof(float) * 9 * n);
    }
    c.ov.n = n;
}

int ib_field(Ctx& c) {
    const Dims& D = c.D;
    static_assert(sizeof(int) == 4, "");
    int* kind = nullptr;
    HIPCHECK(hipMallocAsync((void**)&kind, sizeof(int) * D.P, c.stream));
    HIPCHECK(hipMemsetAsync(kind, 0, sizeof(int) * D.P, c.stream));
    IbArgs a{c.cs.sphereR, c.cs.sphereC[0], c.cs.sphereC[1], c.cs.sphereC[2], c.st, kind};
    const int nb = blocksFor(D.N, 256);
    if (c.uniform) hipLaunchKernelGGL(k_ib_solid<true>, dim3(nb), dim3(256), 0, c.stream, c.geo, a);
    else hipLaunchKernelGGL(k_ib_solid<false>, dim3(nb), dim3(256), 0, c.stream, c.geo, a);
    hipLaunchKernelGGL(k_ib_ghost, dim3(nb), dim3(256), 0, c.stream, D, a);
    HIPCHECK(hipFreeAsync(kind, c.stream));
    launch_set_inactive_fields(c);
    // ghost list, in k,j,i order (IndexMap::generateLocalIndices order)
    std::vector<int8_t> st(D.P);
    HIPCHECK(hipMemcpyAsync(st.data(), c.st, D.P, hipMemcpyDeviceToHost, c.stream));
    HIPCHECK(hipStreamSynchronize(c.stream));
    std::vector<int> cells;
    std::vector<int> rowOf(D.P, -1);
    for (int k = 0; k < D.nK; ++k)
        for (int j = 0; j < D.nJ; ++j)
            for (int i = 0; i < D.nI; ++i) {
                int p = D.pidx(i, j, k);
                if (st[p] == ST_GHOST) { rowOf[p] = (int)cells.size(); cells.push_back(p); }
            }
    ensureOverlay(c, (int)cells.size());
    HIPCHECK(hipMemcpyAsync(c.ov.rowOf, rowOf.data(), sizeof(int) * D.P, hipMemcpyHostToDevice, c.stream));
    if (!cells.empty()) HIPCHECK(hipMemcpyAsync(c.ov.cell, cells.data(), sizeof(int) * cells.size(), hipMemcpyHostToDevice, c.stream));
    // ghost and inactive rows carry a zero right-hand side (never assembled)
    HIPCHECK(hipStreamSynchronize(c.stream));
    return CAFFE_OK;
}

int ib_coeffs(Ctx& c) {
    if (c.ov.n == 0) return CAFFE_OK;
    int* fail = c.flags + 6;
    HIPCHECK(hipMemsetAsync(fail, 0, sizeof(int), c.stream));
    CoefArgs a{c.ov.n, c.ov.cell, c.ov.cols, c.ov.valM, c.ov.valP, c.Am, c.Ap, c.st, nullptr,
               c.cs.sphereR, c.cs.sphereC[0], c.cs.sphereC[1], c.cs.sphereC[2], fail};
    a.nodes = c.dnodes;
    const int nb = blocksFor(c.ov.n, 64);
    if (c.uniform) hipLaunchKernelGGL(k_ib_coeffs<true>, dim3(nb), dim3(64), 0, c.stream, c.geo, a);
    else hipLaunchKernelGGL(k_ib_coeffs<false>, dim3(nb), dim3(64), 0, c.stream, c.geo, a);
    int hf = 0;
    HIPCHECK(hipMemcpyAsync(&hf, fail, sizeof(int), hipMemcpyDeviceToHost, c.stream));
    HIPCHECK(hipStreamSynchronize(c.stream));
    if (hf) {
        c.err = "computeIbCoeffs: an image point has no enclosing cell-centre hexahedron";
        return CAFFE_ERR_STATE;
    }
    return CAFFE_OK;
}

}  // namespace caffe
