/*
 * caffe_oracle.cpp -- CPU FP64 restatement of CAFFE's SIMPLE/PISO/IbPiso hot path.
 *
 * TEST INFRASTRUCTURE ONLY: the correctness oracle and the CPU baseline.  Only
 * tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load it, and
 * only as the checker.  The product path never links or calls it.
 *
 * Build: g++ -O2 -ffp-contract=off -fopenmp (oracle/Makefile).  With
 * -ffp-contract=off every multiply and add is rounded separately, in the
 * order the reference writes them, so the arithmetic is the reference's.
 *
 * Reference files restated (paths relative to the reference root):
 *   src/Math/Vector3D.cpp, Tensor3D.cpp, Geometry.cpp, Sphere.cpp,
 *   Interpolation.cpp:135-181 (LU inverse in place of LAPACKE dgetrf/dgetri),
 *   src/DataStructures/Array3D.tpp, src/Domains/HexaFvmMesh/HexaFvmMesh.cpp,
 *   IndexMap.cpp, HexaMeshGen.cpp:32-79, src/Domains/Field/Field.tpp,
 *   PrimitiveBoundaryCondition.{tpp,cpp}, FlowBoundaryConditions.cpp,
 *   SimpleBoundaryCondition.cpp, src/Schemes/FvScheme.tpp,
 *   FvScalarScheme.cpp, FvVectorScheme.cpp, src/Solvers/Solver.cpp,
 *   Simple.cpp, Piso.cpp, IbPiso.cpp, src/Math/SparseMatrix.cpp:86-106.
 *
 * Definitions where the reference is undefined or delegates to PETSc:
 *   - refValues_ of PrimitiveBoundaryCondition(Field&) are uninitialised in the
 *     reference (PrimitiveBoundaryCondition.tpp:33-40); here they are T() = 0.
 *   - KSPBCGS + PCSOR(1,1) (SparseMatrix.cpp:95-102) is restated as left-
 *     preconditioned BiCGStab with a block-local symmetric Gauss-Seidel sweep
 *     (PETSc's documented defaults: rtol on the preconditioned norm, x0 = 0).
 *     OC_KSP_TIGHT instead runs right-preconditioned BiCGStab to a true
 *     relative residual, so fields do not depend on the Krylov path.
 */
#include "caffe_oracle.h"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

// ---------------------------------------------------------------- Vector3D
// Vector3D.cpp:8-193.  operator/ multiplies by the reciprocal (Vector3D.cpp:98-107).
struct V3 {
    double x, y, z;
    V3() : x(0.), y(0.), z(0.) {}
    V3(double a, double b, double c) : x(a), y(b), z(c) {}
    V3& operator+=(const V3& r) { x += r.x; y += r.y; z += r.z; return *this; }
    V3& operator-=(const V3& r) { x -= r.x; y -= r.y; z -= r.z; return *this; }
    V3& operator*=(double s) { x *= s; y *= s; z *= s; return *this; }
    double& operator()(int c) { return c == 0 ? x : (c == 1 ? y : z); }
    double operator()(int c) const { return c == 0 ? x : (c == 1 ? y : z); }
};
inline V3 operator+(V3 l, const V3& r) { return l += r; }
inline V3 operator-(V3 l, const V3& r) { return l -= r; }
inline V3 operator-(const V3& r) { return V3(-r.x, -r.y, -r.z); }
inline V3 operator*(V3 l, double s) { return l *= s; }
inline V3 operator*(double s, V3 r) { return r *= s; }
inline V3 operator/(V3 l, double s) { s = 1. / s; return l *= s; }
inline double dot(const V3& u, const V3& v) { return u.x * v.x + u.y * v.y + u.z * v.z; }
inline V3 cross(const V3& u, const V3& v) {
    return V3(u.y * v.z - u.z * v.y, u.z * v.x - u.x * v.z, u.x * v.y - u.y * v.x);
}
inline double mag(const V3& u) { return std::sqrt(u.x * u.x + u.y * u.y + u.z * u.z); }
inline V3 unitVector(const V3& u) { double inv = 1. / mag(u); return V3(inv * u.x, inv * u.y, inv * u.z); }

// ---------------------------------------------------------------- Tensor3D
// Tensor3D.cpp: row-major xx..zz; operator/ divides each entry (Tensor3D.cpp:155-167).
struct T3 {
    double a[9];
    T3() { for (int i = 0; i < 9; ++i) a[i] = 0.; }
    T3& operator+=(const T3& r) { for (int i = 0; i < 9; ++i) a[i] += r.a[i]; return *this; }
};
inline T3 operator+(T3 l, const T3& r) { return l += r; }
inline T3 operator/(T3 l, double s) { for (int i = 0; i < 9; ++i) l.a[i] /= s; return l; }
// scalar times tensor, each entry (Tensor3D.cpp operator*=)
inline T3 operator*(double s, T3 t) { for (int i = 0; i < 9; ++i) t.a[i] *= s; return t; }
// tensor(u, v) = u v^T (Vector3D.cpp:176-181)
inline T3 tensor(const V3& u, const V3& v) {
    T3 t;
    t.a[0] = u.x * v.x; t.a[1] = u.x * v.y; t.a[2] = u.x * v.z;
    t.a[3] = u.y * v.x; t.a[4] = u.y * v.y; t.a[5] = u.y * v.z;
    t.a[6] = u.z * v.x; t.a[7] = u.z * v.y; t.a[8] = u.z * v.z;
    return t;
}
// dot(T, v) = T v by rows (Tensor3D.cpp:223-228)
inline V3 dot(const T3& t, const V3& v) {
    return V3(t.a[0] * v.x + t.a[1] * v.y + t.a[2] * v.z,
              t.a[3] * v.x + t.a[4] * v.y + t.a[5] * v.z,
              t.a[6] * v.x + t.a[7] * v.y + t.a[8] * v.z);
}

// ---------------------------------------------------------------- Geometry
// Geometry.cpp:5-12
double quadArea(const V3* p) {
    V3 r01 = p[1] - p[0], r02 = p[2] - p[0], r03 = p[3] - p[0];
    return 0.5 * (mag(cross(r02, r01)) + mag(cross(r03, r02)));
}
// Geometry.cpp:14-20
V3 quadCentroid(const V3* p) {
    V3 d1 = p[0] + 0.5 * (p[2] - p[0]);
    V3 d2 = p[1] + 0.5 * (p[3] - p[1]);
    return d1 + 0.5 * (d2 - d1);
}
// Geometry.cpp:22-25 (the same cross product added twice, then halved)
V3 quadNormal(const V3* p) {
    return 0.5 * (cross(p[1] - p[0], p[2] - p[0]) + cross(p[1] - p[0], p[2] - p[0]));
}
// Geometry.cpp:40-59 (Pletcher, Tannehill & Anderson)
double hexVolume(const V3* p) {
    V3 r16 = p[6] - p[1], r25 = p[5] - p[2];
    V3 s1265 = 0.5 * cross(r16, r25);
    V3 r57 = p[7] - p[5], r64 = p[4] - p[6];
    V3 s5674 = 0.5 * cross(r57, r64);
    V3 r27 = p[7] - p[2], r36 = p[6] - p[3];
    V3 s2376 = 0.5 * cross(r27, r36);
    return std::fabs(dot(s1265 + s5674 + s2376, p[6] - p[0]) / 3.);
}
// Geometry.cpp:258-277
double tetVolume(const V3* p) {
    V3 a = p[1] - p[0], b = p[2] - p[0], c = p[3] - p[0];
    return std::fabs(dot(a, cross(b, c))) / 6.;
}
V3 tetCentroid(const V3* p) {
    V3 s(0., 0., 0.);
    for (int i = 0; i < 4; ++i) s += p[i];
    return 0.25 * s;
}
// Geometry.cpp:61-109 (5-tet split, volume-weighted)
V3 hexCentroid(const V3* p) {
    static const int T[5][4] = {{4, 3, 6, 7}, {4, 1, 6, 5}, {4, 1, 6, 3}, {4, 0, 1, 3}, {3, 1, 2, 6}};
    V3 tet[4], cent[5], c(0., 0., 0.);
    double vol[5], sum = 0.;
    for (int i = 0; i < 5; ++i) {
        for (int q = 0; q < 4; ++q) tet[q] = p[T[i][q]];
        cent[i] = tetCentroid(tet);
        vol[i] = tetVolume(tet);
        sum += vol[i];
    }
    for (int i = 0; i < 5; ++i) c += vol[i] * cent[i];
    return c / sum;
}
// Geometry.cpp:178-256: inside if dot(faceXc - point, faceNorm) >= 0 for all six faces
bool insideHex(const V3& pt, const V3* p) {
    static const int F[6][4] = {{1, 2, 6, 5}, {0, 4, 7, 3}, {2, 3, 7, 6}, {0, 1, 5, 4}, {4, 5, 6, 7}, {0, 3, 2, 1}};
    V3 q[4];
    for (int f = 0; f < 6; ++f) {
        for (int n = 0; n < 4; ++n) q[n] = p[F[f][n]];
        V3 xc = quadCentroid(q), nrm = quadNormal(q);
        if (dot(xc - pt, nrm) < 0.) return false;
    }
    return true;
}

// Interpolation.cpp:135-181: beta = b M^-1 over the monomials [1, xyz, xy, yz, xz, x, y, z].
// The inverse is Gauss-Jordan with partial pivoting (the reference calls LAPACKE dgetrf/dgetri).
void trilinearCoeffs(const V3* pts, const V3& ip, double beta[8]) {
    double M[8][8], Inv[8][8];
    for (int i = 0; i < 8; ++i) {
        double x = pts[i].x, y = pts[i].y, z = pts[i].z;
        M[i][0] = 1.; M[i][1] = x * y * z; M[i][2] = x * y; M[i][3] = y * z;
        M[i][4] = x * z; M[i][5] = x; M[i][6] = y; M[i][7] = z;
        for (int j = 0; j < 8; ++j) Inv[i][j] = (i == j) ? 1. : 0.;
    }
    for (int col = 0; col < 8; ++col) {
        int piv = col;
        for (int r = col + 1; r < 8; ++r)
            if (std::fabs(M[r][col]) > std::fabs(M[piv][col])) piv = r;
        if (piv != col)
            for (int j = 0; j < 8; ++j) { std::swap(M[piv][j], M[col][j]); std::swap(Inv[piv][j], Inv[col][j]); }
        double d = M[col][col];
        for (int j = 0; j < 8; ++j) { M[col][j] /= d; Inv[col][j] /= d; }
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

// ---------------------------------------------------------------- Array3D
// Array3D.tpp:40-58: index k*sizeI*sizeJ + j*sizeI + i, value-initialised storage.
template <class T>
struct A3 {
    int nI = 0, nJ = 0, nK = 0;
    std::vector<T> v;
    void resize(int a, int b, int c) { nI = a; nJ = b; nK = c; v.assign((size_t)a * b * c, T()); }
    T& operator()(int i, int j, int k) { return v[((size_t)k * nJ + j) * nI + i]; }
    const T& operator()(int i, int j, int k) const { return v[((size_t)k * nJ + j) * nI + i]; }
    void assign(const T& t) { std::fill(v.begin(), v.end(), t); }
};

// ---------------------------------------------------------------- Mesh
// HexaFvmMesh.cpp:626-918 (no boundary meshes: one domain).
struct Mesh {
    int nI, nJ, nK;
    A3<V3> nodes;
    A3<V3> xc; A3<double> vol;
    A3<V3> fcI, fcJ, fcK, nI_, nJ_, nK_;  // face centres / area normals
    A3<V3> c2cI, c2cJ, c2cK;               // cell-to-cell vectors
    A3<V3> rfE, rfW, rfN, rfS, rfT, rfB;   // cell-to-face vectors
    A3<double> dE, dW, dN, dS, dT, dB, gE, gW, gN, gS, gT, gB;
    A3<V3> cE, cW, cN, cS, cT, cB;

    V3 SE(int i, int j, int k) const { return nI_(i + 1, j, k); }
    V3 SW(int i, int j, int k) const { return -nI_(i, j, k); }
    V3 SN(int i, int j, int k) const { return nJ_(i, j + 1, k); }
    V3 SS(int i, int j, int k) const { return -nJ_(i, j, k); }
    V3 ST(int i, int j, int k) const { return nK_(i, j, k + 1); }
    V3 SB(int i, int j, int k) const { return -nK_(i, j, k); }
    // HexaFvmMesh.cpp:218-294: at a physical boundary the cell-to-face vector is used
    V3 rCellE(int i, int j, int k) const { return i == nI - 1 ? rfE(i, j, k) : c2cI(i, j, k); }
    V3 rCellW(int i, int j, int k) const { return i == 0 ? rfW(i, j, k) : -c2cI(i - 1, j, k); }
    V3 rCellN(int i, int j, int k) const { return j == nJ - 1 ? rfN(i, j, k) : c2cJ(i, j, k); }
    V3 rCellS(int i, int j, int k) const { return j == 0 ? rfS(i, j, k) : -c2cJ(i, j - 1, k); }
    V3 rCellT(int i, int j, int k) const { return k == nK - 1 ? rfT(i, j, k) : c2cK(i, j, k); }
    V3 rCellB(int i, int j, int k) const { return k == 0 ? rfB(i, j, k) : -c2cK(i, j, k - 1); }

    void hexNodes(int i, int j, int k, V3* t) const {
        t[0] = nodes(i, j, k); t[1] = nodes(i + 1, j, k); t[2] = nodes(i + 1, j + 1, k); t[3] = nodes(i, j + 1, k);
        t[4] = nodes(i, j, k + 1); t[5] = nodes(i + 1, j, k + 1); t[6] = nodes(i + 1, j + 1, k + 1); t[7] = nodes(i, j + 1, k + 1);
    }

    void init(int a, int b, int c, const double* nd) {
        nI = a; nJ = b; nK = c;
        nodes.resize(nI + 1, nJ + 1, nK + 1);
        for (size_t q = 0; q < nodes.v.size(); ++q) nodes.v[q] = V3(nd[3 * q], nd[3 * q + 1], nd[3 * q + 2]);
        // initializeCells (HexaFvmMesh.cpp:626-655)
        xc.resize(nI, nJ, nK); vol.resize(nI, nJ, nK);
        V3 t[8];
        for (int k = 0; k < nK; ++k) for (int j = 0; j < nJ; ++j) for (int i = 0; i < nI; ++i) {
            hexNodes(i, j, k, t);
            xc(i, j, k) = hexCentroid(t);
            vol(i, j, k) = hexVolume(t);
        }
        // initializeCellToCellParameters (:657-697)
        c2cI.resize(nI - 1 > 0 ? nI - 1 : 0, nJ, nK); c2cJ.resize(nI, nJ - 1 > 0 ? nJ - 1 : 0, nK); c2cK.resize(nI, nJ, nK - 1 > 0 ? nK - 1 : 0);
        for (int k = 0; k < nK; ++k) for (int j = 0; j < nJ; ++j) for (int i = 0; i < nI; ++i) {
            if (i < nI - 1) c2cI(i, j, k) = xc(i + 1, j, k) - xc(i, j, k);
            if (j < nJ - 1) c2cJ(i, j, k) = xc(i, j + 1, k) - xc(i, j, k);
            if (k < nK - 1) c2cK(i, j, k) = xc(i, j, k + 1) - xc(i, j, k);
        }
        // initializeFaces (:699-789): normal = unit(cross(p1-p0, p2-p0)) * quadArea
        V3 q[4];
        fcI.resize(nI + 1, nJ, nK); nI_.resize(nI + 1, nJ, nK);
        for (int k = 0; k < nK; ++k) for (int j = 0; j < nJ; ++j) for (int i = 0; i < nI + 1; ++i) {
            q[0] = nodes(i, j, k); q[1] = nodes(i, j + 1, k); q[2] = nodes(i, j + 1, k + 1); q[3] = nodes(i, j, k + 1);
            V3 tv = unitVector(cross(q[1] - q[0], q[2] - q[0]));
            double ar = quadArea(q);
            fcI(i, j, k) = quadCentroid(q); nI_(i, j, k) = tv * ar;
        }
        fcJ.resize(nI, nJ + 1, nK); nJ_.resize(nI, nJ + 1, nK);
        for (int k = 0; k < nK; ++k) for (int j = 0; j < nJ + 1; ++j) for (int i = 0; i < nI; ++i) {
            q[0] = nodes(i, j, k); q[1] = nodes(i, j, k + 1); q[2] = nodes(i + 1, j, k + 1); q[3] = nodes(i + 1, j, k);
            V3 tv = unitVector(cross(q[1] - q[0], q[2] - q[0]));
            double ar = quadArea(q);
            fcJ(i, j, k) = quadCentroid(q); nJ_(i, j, k) = tv * ar;
        }
        fcK.resize(nI, nJ, nK + 1); nK_.resize(nI, nJ, nK + 1);
        for (int k = 0; k < nK + 1; ++k) for (int j = 0; j < nJ; ++j) for (int i = 0; i < nI; ++i) {
            q[0] = nodes(i, j, k); q[1] = nodes(i + 1, j, k); q[2] = nodes(i + 1, j + 1, k); q[3] = nodes(i, j + 1, k);
            V3 tv = unitVector(cross(q[1] - q[0], q[2] - q[0]));
            double ar = quadArea(q);
            fcK(i, j, k) = quadCentroid(q); nK_(i, j, k) = tv * ar;
        }
        // initializeCellToFaceParameters (:791-839)
        rfE.resize(nI, nJ, nK); rfW.resize(nI, nJ, nK); rfN.resize(nI, nJ, nK);
        rfS.resize(nI, nJ, nK); rfT.resize(nI, nJ, nK); rfB.resize(nI, nJ, nK);
        for (int k = 0; k < nK; ++k) for (int j = 0; j < nJ; ++j) for (int i = 0; i < nI; ++i) {
            rfE(i, j, k) = fcI(i + 1, j, k) - xc(i, j, k);
            rfW(i, j, k) = fcI(i, j, k) - xc(i, j, k);
            rfN(i, j, k) = fcJ(i, j + 1, k) - xc(i, j, k);
            rfS(i, j, k) = fcJ(i, j, k) - xc(i, j, k);
            rfT(i, j, k) = fcK(i, j, k + 1) - xc(i, j, k);
            rfB(i, j, k) = fcK(i, j, k) - xc(i, j, k);
        }
        // computeMeshMetrics (:841-918)
        A3<double>* ds[6] = {&dE, &dW, &dN, &dS, &dT, &dB};
        A3<double>* gs[6] = {&gE, &gW, &gN, &gS, &gT, &gB};
        A3<V3>* cs[6] = {&cE, &cW, &cN, &cS, &cT, &cB};
        for (int f = 0; f < 6; ++f) { ds[f]->resize(nI, nJ, nK); gs[f]->resize(nI, nJ, nK); cs[f]->resize(nI, nJ, nK); }
        for (int k = 0; k < nK; ++k) for (int j = 0; j < nJ; ++j) for (int i = 0; i < nI; ++i) {
            V3 S[6] = {SE(i, j, k), SW(i, j, k), SN(i, j, k), SS(i, j, k), ST(i, j, k), SB(i, j, k)};
            V3 R[6] = {rCellE(i, j, k), rCellW(i, j, k), rCellN(i, j, k), rCellS(i, j, k), rCellT(i, j, k), rCellB(i, j, k)};
            for (int f = 0; f < 6; ++f) {
                (*ds[f])(i, j, k) = dot(S[f], S[f]) / dot(S[f], R[f]);
                (*cs[f])(i, j, k) = S[f] - R[f] * dot(S[f], S[f]) / dot(S[f], R[f]);
            }
            gE(i, j, k) = i < nI - 1 ? vol(i + 1, j, k) / (vol(i + 1, j, k) + vol(i, j, k)) : 0.;
            gW(i, j, k) = i > 0 ? vol(i - 1, j, k) / (vol(i - 1, j, k) + vol(i, j, k)) : 0.;
            gN(i, j, k) = j < nJ - 1 ? vol(i, j + 1, k) / (vol(i, j + 1, k) + vol(i, j, k)) : 0.;
            gS(i, j, k) = j > 0 ? vol(i, j - 1, k) / (vol(i, j - 1, k) + vol(i, j, k)) : 0.;
            gT(i, j, k) = k < nK - 1 ? vol(i, j, k + 1) / (vol(i, j, k + 1) + vol(i, j, k)) : 0.;
            gB(i, j, k) = k > 0 ? vol(i, j, k - 1) / (vol(i, j, k - 1) + vol(i, j, k)) : 0.;
        }
    }

    // HexaFvmMesh.cpp:345-376: first cell in k,j,i order whose hex contains the point.
    // The k,j,i scan is kept; slabs (k) and rows (j,k) whose node bounding box, widened by
    // 1e-6 of its extent, misses the point are skipped whole.  On meshes with planar faces
    // the six half-space test accepts only points of the closed hex, which lies in that box,
    // so the first hit is the reference's.
    std::vector<double> slabBox, rowBox;  // 6 doubles per slab / per row: min xyz, max xyz
    static void grow(double* b, const V3& p) {
        b[0] = std::min(b[0], p.x); b[1] = std::min(b[1], p.y); b[2] = std::min(b[2], p.z);
        b[3] = std::max(b[3], p.x); b[4] = std::max(b[4], p.y); b[5] = std::max(b[5], p.z);
    }
    static bool inBox(const double* b, const V3& p) {
        double tx = 1e-6 * (b[3] - b[0]) + 1e-300, ty = 1e-6 * (b[4] - b[1]) + 1e-300, tz = 1e-6 * (b[5] - b[2]) + 1e-300;
        return p.x >= b[0] - tx && p.x <= b[3] + tx && p.y >= b[1] - ty && p.y <= b[4] + ty && p.z >= b[2] - tz && p.z <= b[5] + tz;
    }
    void buildBoxes() {
        const double big = 1e300;
        slabBox.assign((size_t)6 * nK, 0.); rowBox.assign((size_t)6 * nK * nJ, 0.);
        for (int k = 0; k < nK; ++k) {
            double* sb = &slabBox[6 * (size_t)k];
            sb[0] = sb[1] = sb[2] = big; sb[3] = sb[4] = sb[5] = -big;
            for (int j = 0; j < nJ; ++j) {
                double* rb = &rowBox[6 * ((size_t)k * nJ + j)];
                rb[0] = rb[1] = rb[2] = big; rb[3] = rb[4] = rb[5] = -big;
                for (int dk = 0; dk < 2; ++dk) for (int dj = 0; dj < 2; ++dj) for (int i = 0; i <= nI; ++i)
                    grow(rb, nodes(i, j + dj, k + dk));
                for (int q = 0; q < 3; ++q) { sb[q] = std::min(sb[q], rb[q]); sb[q + 3] = std::max(sb[q + 3], rb[q + 3]); }
            }
        }
    }
    bool locateCell(const V3& pt, int& ii, int& jj, int& kk) {
        if (slabBox.empty()) buildBoxes();
        V3 t[8];
        for (int k = 0; k < nK; ++k) {
            if (!inBox(&slabBox[6 * (size_t)k], pt)) continue;
            for (int j = 0; j < nJ; ++j) {
                if (!inBox(&rowBox[6 * ((size_t)k * nJ + j)], pt)) continue;
                for (int i = 0; i < nI; ++i) {
                    hexNodes(i, j, k, t);
                    if (insideHex(pt, t)) { ii = i; jj = j; kk = k; return true; }
                }
            }
        }
        return false;
    }
    // HexaFvmMesh.cpp:378-418
    bool locateEnclosingCells(const V3& pt, int ii[8], int jj[8], int kk[8]) {
        int i0, j0, k0;
        if (!locateCell(pt, i0, j0, k0)) return false;
        V3 t[8];
        for (int k = k0 - 1; k <= k0; ++k) for (int j = j0 - 1; j <= j0; ++j) for (int i = i0 - 1; i <= i0; ++i) {
            if (i < 0 || j < 0 || k < 0 || i + 1 >= nI || j + 1 >= nJ || k + 1 >= nK) return false;
            t[0] = xc(i, j, k); t[1] = xc(i + 1, j, k); t[2] = xc(i + 1, j + 1, k); t[3] = xc(i, j + 1, k);
            t[4] = xc(i, j, k + 1); t[5] = xc(i + 1, j, k + 1); t[6] = xc(i + 1, j + 1, k + 1); t[7] = xc(i, j + 1, k + 1);
            if (insideHex(pt, t)) {
                const int di[8] = {0, 1, 1, 0, 0, 1, 1, 0}, dj[8] = {0, 0, 1, 1, 0, 0, 1, 1}, dk[8] = {0, 0, 0, 0, 1, 1, 1, 1};
                for (int m = 0; m < 8; ++m) { ii[m] = i + di[m]; jj[m] = j + dj[m]; kk[m] = k + dk[m]; }
                return true;
            }
        }
        return false;
    }
};

// ---------------------------------------------------------------- IndexMap
// IndexMap.cpp:34-148 (serial: off-domain lookups return -1, :77)
enum CellStatus { ACTIVE = 0, GHOST = 1, INACTIVE = 2 };
struct IndexMap {
    int nI, nJ, nK, nActive;
    A3<int> status, gidx;
    void init(int a, int b, int c) { nI = a; nJ = b; nK = c; status.resize(a, b, c); gidx.resize(a, b, c); status.assign(ACTIVE); generate(); }
    void generate() {
        nActive = 0;
        for (size_t q = 0; q < status.v.size(); ++q) {
            if (status.v[q] == INACTIVE) gidx.v[q] = -1;
            else gidx.v[q] = nActive++;
        }
    }
    int operator()(int i, int j, int k) const {
        if (i >= 0 && i < nI && j >= 0 && j < nJ && k >= 0 && k < nK) return gidx(i, j, k);
        return -1;
    }
    bool isActive(int i, int j, int k) const { return status(i, j, k) == ACTIVE; }
    bool isInactive(int i, int j, int k) const { return status(i, j, k) == INACTIVE; }
};

// ---------------------------------------------------------------- Field
// Field.tpp:31-128, 270-404.  Patches are one-cell ghost layers.
template <class T>
struct Field {
    const Mesh* m = nullptr;
    bool conserved = true;
    A3<T> cells, fI, fJ, fK;
    A3<T> pE, pW, pN, pS, pT, pB;  // east..bottom patches
    void init(const Mesh& mesh, bool cons) {
        m = &mesh; conserved = cons;
        int a = mesh.nI, b = mesh.nJ, c = mesh.nK;
        cells.resize(a, b, c);
        if (cons) { fI.resize(a + 1, b, c); fJ.resize(a, b + 1, c); fK.resize(a, b, c + 1); }
        pE.resize(1, b, c); pW.resize(1, b, c); pN.resize(a, 1, c); pS.resize(a, 1, c); pT.resize(a, b, 1); pB.resize(a, b, 1);
    }
    T& operator()(int i, int j, int k) {
        if (i >= 0 && j >= 0 && k >= 0 && i < cells.nI && j < cells.nJ && k < cells.nK) return cells(i, j, k);
        if (i < 0) return pW(-i - 1, j, k);
        if (i >= cells.nI) return pE(i - cells.nI, j, k);
        if (j < 0) return pS(i, -j - 1, k);
        if (j >= cells.nJ) return pN(i, j - cells.nJ, k);
        if (k < 0) return pB(i, j, -k - 1);
        return pT(i, j, k - cells.nK);
    }
    T& faceE(int i, int j, int k) { return fI(i + 1, j, k); }
    T& faceW(int i, int j, int k) { return fI(i, j, k); }
    T& faceN(int i, int j, int k) { return fJ(i, j + 1, k); }
    T& faceS(int i, int j, int k) { return fJ(i, j, k); }
    T& faceT(int i, int j, int k) { return fK(i, j, k + 1); }
    T& faceB(int i, int j, int k) { return fK(i, j, k); }
    // Field.tpp:59-64: faces copy only when both are CONSERVED
    void copyFrom(const Field& o) {
        cells = o.cells;
        if (conserved && o.conserved) { fI = o.fI; fJ = o.fJ; fK = o.fK; }
    }
    // Field.tpp:250-259, 280-362
    void setFixedPatches(const T* r) {
        pE.assign(r[0]); pW.assign(r[1]); pN.assign(r[2]); pS.assign(r[3]); pT.assign(r[4]); pB.assign(r[5]);
    }
    void facesFromPatch(int f) {
        int a = cells.nI, b = cells.nJ, c = cells.nK;
        if (!conserved) return;
        switch (f) {
        case 0: for (int k = 0; k < c; ++k) for (int j = 0; j < b; ++j) fI(a, j, k) = pE(0, j, k); break;
        case 1: for (int k = 0; k < c; ++k) for (int j = 0; j < b; ++j) fI(0, j, k) = pW(0, j, k); break;
        case 2: for (int k = 0; k < c; ++k) for (int i = 0; i < a; ++i) fJ(i, b, k) = pN(i, 0, k); break;
        case 3: for (int k = 0; k < c; ++k) for (int i = 0; i < a; ++i) fJ(i, 0, k) = pS(i, 0, k); break;
        case 4: for (int j = 0; j < b; ++j) for (int i = 0; i < a; ++i) fK(i, j, c) = pT(i, j, 0); break;
        case 5: for (int j = 0; j < b; ++j) for (int i = 0; i < a; ++i) fK(i, j, 0) = pB(i, j, 0); break;
        }
    }
    // Field.tpp:364-404 (getSubfield of the boundary cell plane, then faces from the patch)
    void zeroGradient(int f) {
        int a = cells.nI, b = cells.nJ, c = cells.nK;
        switch (f) {
        case 0: for (int k = 0; k < c; ++k) for (int j = 0; j < b; ++j) pE(0, j, k) = cells(a - 1, j, k); break;
        case 1: for (int k = 0; k < c; ++k) for (int j = 0; j < b; ++j) pW(0, j, k) = cells(0, j, k); break;
        case 2: for (int k = 0; k < c; ++k) for (int i = 0; i < a; ++i) pN(i, 0, k) = cells(i, b - 1, k); break;
        case 3: for (int k = 0; k < c; ++k) for (int i = 0; i < a; ++i) pS(i, 0, k) = cells(i, 0, k); break;
        case 4: for (int j = 0; j < b; ++j) for (int i = 0; i < a; ++i) pT(i, j, 0) = cells(i, j, c - 1); break;
        case 5: for (int j = 0; j < b; ++j) for (int i = 0; i < a; ++i) pB(i, j, 0) = cells(i, j, 0); break;
        }
        facesFromPatch(f);
    }
    void setAll(const T& t) {
        cells.assign(t); fI.assign(t); fJ.assign(t); fK.assign(t);
        T r[6] = {t, t, t, t, t, t};
        setFixedPatches(r);
    }
};

// ---------------------------------------------------------------- boundary conditions
// PrimitiveBoundaryCondition.tpp:33-294
enum PType { P_FIXED = 0, P_ZERO_GRADIENT = 1, P_EMPTY = 2 };
template <class T>
struct PrimBC {
    Field<T>* f = nullptr;
    int types[6];
    T ref[6];
    void init(Field<T>& fld) { f = &fld; for (int i = 0; i < 6; ++i) { types[i] = P_FIXED; ref[i] = T(); } }
    void setFixed() { f->setFixedPatches(ref); for (int q = 0; q < 6; ++q) f->facesFromPatch(q); }
    void changeType(int i, int t, const T& r) { types[i] = t; ref[i] = r; if (t == P_FIXED) setFixed(); }
    void setBoundaries() { for (int q = 0; q < 6; ++q) if (types[q] != P_FIXED) f->zeroGradient(q); }
    // PrimitiveBoundaryCondition.tpp:144-255
    void fold(int i, int j, int k, double* a, T& b) {
        const Field<T>& F = *f;
        int uI = F.cells.nI - 1, uJ = F.cells.nJ - 1, uK = F.cells.nK - 1;
        if (i == uI) { if (types[0] == P_FIXED) b -= a[1] * F.pE(0, j, k); else { a[0] += a[1]; a[1] = 0.; } }
        if (i == 0) { if (types[1] == P_FIXED) b -= a[2] * F.pW(0, j, k); else { a[0] += a[2]; a[2] = 0.; } }
        if (j == uJ) { if (types[2] == P_FIXED) b -= a[3] * F.pN(i, 0, k); else { a[0] += a[3]; a[3] = 0.; } }
        if (j == 0) { if (types[3] == P_FIXED) b -= a[4] * F.pS(i, 0, k); else { a[0] += a[4]; a[4] = 0.; } }
        if (k == uK) { if (types[4] == P_FIXED) b -= a[5] * F.pT(i, j, 0); else { a[0] += a[5]; a[5] = 0.; } }
        if (k == 0) { if (types[5] == P_FIXED) b -= a[6] * F.pB(i, j, 0); else { a[0] += a[6]; a[6] = 0.; } }
    }
};

// ---------------------------------------------------------------- sparse matrix
// SparseMatrix.cpp:47-106: AIJ rows, INSERT_VALUES, negative columns dropped.
template <int W>
struct RowT { int n = 0; int col[W]; double val[W]; };
template <int W>
struct SpMatT {
    std::vector<RowT<W>> rows;
    void allocate(int n) { rows.assign(n, RowT<W>()); }
    void setRow(int r, int nc, const int* cols, const double* vals) {
        RowT<W>& R = rows[r];
        for (int q = 0; q < nc; ++q) {
            int c = cols[q];
            if (c < 0) continue;
            int p = 0;
            while (p < R.n && R.col[p] != c) ++p;
            if (p < R.n) { R.val[p] = vals[q]; continue; }
            // keep columns sorted, as AIJ does
            int at = R.n;
            while (at > 0 && R.col[at - 1] > c) { R.col[at] = R.col[at - 1]; R.val[at] = R.val[at - 1]; --at; }
            R.col[at] = c; R.val[at] = vals[q]; ++R.n;
        }
    }
};
using SpMat = SpMatT<9>;     // 7-point rows and 9-entry ghost rows
using SpMat28 = SpMatT<28>;  // Coupled: 14-entry momentum rows, 28-entry continuity rows

struct KspOpts { int kind; double rtol; int maxit; int nBlocks; int nThreads; };

// Deterministic dot product: fixed chunks summed in order (independent of thread count).
double dotv(const std::vector<double>& a, const std::vector<double>& b, int nt) {
    const int NCH = 64;
    size_t n = a.size();
    double part[NCH];
    #pragma omp parallel for num_threads(nt) schedule(static)
    for (int c = 0; c < NCH; ++c) {
        size_t lo = n * c / NCH, hi = n * (c + 1) / NCH;
        double s = 0.;
        for (size_t q = lo; q < hi; ++q) s += a[q] * b[q];
        part[c] = s;
    }
    double s = 0.;
    for (int c = 0; c < NCH; ++c) s += part[c];
    return s;
}

template <class M>
void spmv(const M& A, const std::vector<double>& x, std::vector<double>& y, int nt) {
    int n = (int)A.rows.size();
    #pragma omp parallel for num_threads(nt) schedule(static)
    for (int r = 0; r < n; ++r) {
        const auto& R = A.rows[r];
        double s = 0.;
        for (int q = 0; q < R.n; ++q) s += R.val[q] * x[R.col[q]];
        y[r] = s;
    }
}

// PCSOR(its=1, lits=1), SOR_LOCAL_SYMMETRIC_SWEEP, omega = 1, zero initial guess:
// one forward and one backward Gauss-Seidel sweep inside each row block; couplings
// to other blocks are ignored (processor-local SOR across ranks).
template <class M>
void ssorApply(const M& A, const std::vector<double>& r, std::vector<double>& z, int nBlocks, int nt) {
    int n = (int)A.rows.size();
    #pragma omp parallel for num_threads(nt) schedule(static)
    for (int bl = 0; bl < nBlocks; ++bl) {
        int lo = (int)((long)n * bl / nBlocks), hi = (int)((long)n * (bl + 1) / nBlocks);
        for (int i = lo; i < hi; ++i) z[i] = 0.;
        for (int i = lo; i < hi; ++i) {
            const auto& R = A.rows[i];
            double s = r[i], d = 1.;
            for (int q = 0; q < R.n; ++q) {
                int c = R.col[q];
                if (c == i) d = R.val[q];
                else if (c >= lo && c < hi) s -= R.val[q] * z[c];
            }
            z[i] = s / d;
        }
        for (int i = hi - 1; i >= lo; --i) {
            const auto& R = A.rows[i];
            double s = r[i], d = 1.;
            for (int q = 0; q < R.n; ++q) {
                int c = R.col[q];
                if (c == i) d = R.val[q];
                else if (c >= lo && c < hi) s -= R.val[q] * z[c];
            }
            z[i] = s / d;
        }
    }
}

// Left-preconditioned BiCGStab with the preconditioned residual norm (PETSc KSPBCGS defaults).
template <class M>
int bicgstabLeft(const M& A, const std::vector<double>& b, std::vector<double>& x, const KspOpts& o, double* relres) {
    int n = (int)b.size(), nt = o.nThreads;
    std::vector<double> r(n), rp(n), p(n, 0.), v(n, 0.), s(n), t(n), tmp(n);
    std::fill(x.begin(), x.end(), 0.);
    ssorApply(A, b, r, o.nBlocks, nt);
    rp = r;
    double rnorm0 = std::sqrt(dotv(r, r, nt));
    if (rnorm0 == 0.) { if (relres) *relres = 0.; return 0; }
    double rhoOld = 1., alpha = 1., omegaOld = 1.;
    int it = 0;
    for (it = 0; it < o.maxit; ++it) {
        double rho = dotv(r, rp, nt);
        double beta = (rho / rhoOld) * (alpha / omegaOld);
        for (int q = 0; q < n; ++q) p[q] = r[q] + beta * (p[q] - omegaOld * v[q]);
        spmv(A, p, tmp, nt); ssorApply(A, tmp, v, o.nBlocks, nt);
        double d = dotv(v, rp, nt);
        if (d == 0.) break;
        alpha = rho / d;
        for (int q = 0; q < n; ++q) s[q] = r[q] - alpha * v[q];
        spmv(A, s, tmp, nt); ssorApply(A, tmp, t, o.nBlocks, nt);
        double tt = dotv(t, t, nt);
        double omega = tt > 0. ? dotv(t, s, nt) / tt : 0.;
        for (int q = 0; q < n; ++q) { x[q] += alpha * p[q] + omega * s[q]; r[q] = s[q] - omega * t[q]; }
        double rn = std::sqrt(dotv(r, r, nt));
        rhoOld = rho; omegaOld = omega;
        if (rn <= o.rtol * rnorm0 || omega == 0.) { ++it; if (relres) *relres = rn / rnorm0; return it; }
    }
    if (relres) *relres = -1.;
    return it;
}

// Right-preconditioned BiCGStab to a true relative residual, restarted from the true residual.
// project: the rows and unknowns from projFrom on carry a Neumann null space (constants):
// the right-hand side loses its (rounding) mean there and the solution is returned with
// zero mean there.  The residual's mean over those rows comes only from the rounding of
// the zero row sums; no Krylov step can reduce it, so it is removed from every residual
// and left out of the residual norm (otherwise a tight tolerance stalls on it); the
// preconditioned directions are projected to zero mean there too, so the iterate gathers
// no constant whose rounding in A x would floor the true residual.
// Safeguards near the attainable accuracy: a cycle restarts on breakdown (rho or (r^, v)
// vanishing, a non-finite step length), when its recurrence residual grows 1e3-fold over
// its best, or after 100 iterations without a new best; the best iterate (by the true
// residual at the restarts) is kept, and the solve ends, returning it, once a restart no
// longer improves on it (a restart from the same iterate would replay the same cycle).
template <class M>
int bicgstabTight(const M& A, const std::vector<double>& bIn, std::vector<double>& x, const KspOpts& o,
                  bool project, double* relres, int projFrom = 0, bool gauge = true) {
    int n = (int)bIn.size(), nt = o.nThreads;
    std::vector<double> b = bIn;
    const int np = n - projFrom;
    const bool proj = project && np > 0;
    auto demean = [&](std::vector<double>& w) {
        if (!proj) return;
        double m = 0.;
        for (int q = projFrom; q < n; ++q) m += w[q];
        m /= np;
        for (int q = projFrom; q < n; ++q) w[q] -= m;
    };
    demean(b);
    std::fill(x.begin(), x.end(), 0.);
    double bn = std::sqrt(dotv(b, b, nt));
    if (bn == 0.) { if (relres) *relres = 0.; return 0; }
    std::vector<double> r(n), rh(n), p(n), v(n), s(n), t(n), ph(n), sh(n), xbest = x;
    int total = 0;
    double rrBest = 2.;
    const double tiny = 1e-30;
    for (int restart = 0; restart < 200 && total < o.maxit; ++restart) {
        spmv(A, x, r, nt);
        for (int q = 0; q < n; ++q) r[q] = b[q] - r[q];
        demean(r);
        const double rr = std::sqrt(dotv(r, r, nt)) / bn;
        if (!(rr < rrBest)) break;  // no gain since the last restart: the best iterate stands
        rrBest = rr;
        xbest = x;
        if (rr <= o.rtol) break;
        rh = r; p = r;
        const double rhn = std::sqrt(dotv(rh, rh, nt));
        double rho = dotv(rh, r, nt);
        double rnMin = rr;
        int itMin = total;
        for (; total < o.maxit; ++total) {
            ssorApply(A, p, ph, o.nBlocks, nt);
            demean(ph);  // the preconditioned direction projected off the null space
            spmv(A, ph, v, nt);
            demean(v);
            double d = dotv(rh, v, nt);
            double alpha = rho / d;
            if (!(std::fabs(d) > tiny * rhn * std::sqrt(dotv(v, v, nt))) || !std::isfinite(alpha)) break;
            for (int q = 0; q < n; ++q) s[q] = r[q] - alpha * v[q];
            ssorApply(A, s, sh, o.nBlocks, nt);
            demean(sh);
            spmv(A, sh, t, nt);
            demean(t);
            double tt = dotv(t, t, nt);
            double omega = tt > 0. ? dotv(t, s, nt) / tt : 0.;
            if (!std::isfinite(omega)) break;
            for (int q = 0; q < n; ++q) { x[q] += alpha * ph[q] + omega * sh[q]; r[q] = s[q] - omega * t[q]; }
            double rn = std::sqrt(dotv(r, r, nt)) / bn;
            if (rn <= 0.5 * o.rtol || omega == 0.) { ++total; break; }
            if (rn < rnMin) { rnMin = rn; itMin = total; }
            if (!(rn < 1e3 * rnMin) || total - itMin > 100) { ++total; break; }
            double rhoN = dotv(rh, r, nt);
            if (!(std::fabs(rhoN) > tiny * rhn * rn * bn)) { ++total; break; }
            double beta = (rhoN / rho) * (alpha / omega);
            rho = rhoN;
            for (int q = 0; q < n; ++q) p[q] = r[q] + beta * (p[q] - omega * v[q]);
        }
    }
    // the cycle that ran last may have ended worse than the best restart point
    spmv(A, x, r, nt);
    for (int q = 0; q < n; ++q) r[q] = b[q] - r[q];
    demean(r);
    if (!(std::sqrt(dotv(r, r, nt)) / bn < rrBest)) x = xbest;
    if (proj && gauge) {  // gauge of the Neumann problem: zero-mean solution
        double m = 0.;
        for (int q = projFrom; q < n; ++q) m += x[q];
        m /= np;
        for (int q = projFrom; q < n; ++q) x[q] -= m;
    }
    spmv(A, x, r, nt);
    for (int q = 0; q < n; ++q) r[q] = b[q] - r[q];
    demean(r);
    if (relres) *relres = std::sqrt(dotv(r, r, nt)) / bn;
    return total;
}

// ---------------------------------------------------------------- the solver
struct Oracle {
    oc_case cs;
    Mesh mesh;
    IndexMap iMap;
    Field<V3> u, u0, uStar, h, gradP, gradPC;
    Field<double> p, rho, mu, mflow, pC, d;
    Field<T3> gradU;
    Field<double> phi;      // Diffusion: phiField_ (CONSERVED)
    Field<V3> gradPhi;      // Diffusion: gradPhiField_ (AUXILLARY)
    PrimBC<double> phiBC;
    PrimBC<V3> uBC, hBC;
    PrimBC<double> pBC, rhoBC, muBC, dBC, pCBC;
    int flowType[6];
    SpMat A[2];
    std::vector<double> bv[3], xv[3];
    SpMat28 Ac;                    // Coupled: the 4-variable system (Coupled.cpp:20)
    std::vector<double> bc, xc;
    A3<int> ib;  // IbPiso cell types: 0 FLUID, 1 IB, 2 SOLID (IbPiso.h:44); patches are FLUID
    int itersMom = 0, itersPC = 0;
    double lastRelres = 0.;
    bool pcSingular = false;

    bool massCorr(int f) const { return flowType[f] == OC_OUTLET; }  // SimpleBoundaryCondition.cpp:68-96

    void createMatrices(int nnz) {
        (void)nnz;
        for (int q = 0; q < 2; ++q) A[q].allocate(iMap.nActive);
        for (int q = 0; q < 3; ++q) { bv[q].assign(iMap.nActive, 0.); xv[q].assign(iMap.nActive, 0.); }
    }

    void init(const oc_case* c, const double* nodes) {
        cs = *c;
        mesh.init(c->nI, c->nJ, c->nK, nodes);
        iMap.init(c->nI, c->nJ, c->nK);
        if (c->algo == OC_DIFFUSION) { initDiffusion(c); return; }
        u.init(mesh, true); p.init(mesh, true); rho.init(mesh, true); mu.init(mesh, true);
        mflow.init(mesh, true); pC.init(mesh, true); gradU.init(mesh, true); gradP.init(mesh, true);
        gradPC.init(mesh, true); u0.init(mesh, false); uStar.init(mesh, false); d.init(mesh, true); h.init(mesh, true);
        // FlowBoundaryConditions.cpp:3-73
        uBC.init(u); pBC.init(p); rhoBC.init(rho); muBC.init(mu); hBC.init(h); dBC.init(d);
        for (int i = 0; i < 6; ++i) {
            int t = c->bcType[i];
            flowType[i] = t;
            V3 rv(c->bcVec[3 * i], c->bcVec[3 * i + 1], c->bcVec[3 * i + 2]);
            double rp = c->bcVal[i];
            if (t == OC_INLET || t == OC_WALL) {
                uBC.changeType(i, P_FIXED, rv); pBC.changeType(i, P_ZERO_GRADIENT, rp); hBC.changeType(i, P_FIXED, rv);
            } else if (t == OC_OUTLET) {
                uBC.changeType(i, P_ZERO_GRADIENT, rv); pBC.changeType(i, P_FIXED, rp); hBC.changeType(i, P_ZERO_GRADIENT, rv);
            } else if (t == OC_EMPTY) {
                uBC.changeType(i, P_EMPTY, rv); pBC.changeType(i, P_EMPTY, rp); hBC.changeType(i, P_EMPTY, rv);
            } else {
                uBC.changeType(i, P_ZERO_GRADIENT, rv); pBC.changeType(i, P_ZERO_GRADIENT, rp); hBC.changeType(i, P_ZERO_GRADIENT, rv);
            }
            rhoBC.changeType(i, P_ZERO_GRADIENT, 0.); muBC.changeType(i, P_ZERO_GRADIENT, 0.); dBC.changeType(i, P_ZERO_GRADIENT, 0.);
        }
        // SimpleBoundaryCondition.cpp:37-57: p' mirrors the p types with reference 0
        pCBC.init(pC);
        pcSingular = true;
        for (int i = 0; i < 6; ++i) { pCBC.changeType(i, pBC.types[i], 0.); if (pBC.types[i] == P_FIXED) pcSingular = false; }
        rho.setAll(c->rho); mu.setAll(c->mu);
        ib.resize(c->nI, c->nJ, c->nK);
        if (c->algo == OC_COUPLED) {
            // Coupled.cpp:5-34: createMatrices(4, 1, 1, 28); no initial conditions are read,
            // so u and p start at zero
            const int n4 = 4 * iMap.nActive;
            Ac.allocate(n4);
            bc.assign(n4, 0.); xc.assign(n4, 0.);
            return;
        }
        createMatrices(7);  // Simple.cpp:56
        // Simple.cpp:104-108, 71-73 (uniform initial conditions set cells only)
        u.cells.assign(V3(c->uInit[0], c->uInit[1], c->uInit[2]));
        p.cells.assign(c->pInit);
    }

    int nt() const { return cs.nThreads > 0 ? cs.nThreads : 1; }
    KspOpts kopts() const { return KspOpts{cs.kspKind, cs.rtol, cs.maxit, cs.nBlocks > 0 ? cs.nBlocks : 1, nt()}; }
    int kspSolve(int which, int vec, bool project) {
        KspOpts o = kopts();
        if (o.kind == OC_KSP_TIGHT) return bicgstabTight(A[which], bv[vec], xv[vec], o, project, &lastRelres);
        return bicgstabLeft(A[which], bv[vec], xv[vec], o, &lastRelres);
    }

    // FvScheme.tpp:12-40, VOLUME_WEIGHTED, interior faces only (no boundary meshes)
    template <class T>
    void interpolate(Field<T>& F) {
        const Mesh& M = mesh;
        #pragma omp parallel for num_threads(nt()) schedule(static)
        for (int k = 0; k < M.nK; ++k) for (int j = 0; j < M.nJ; ++j) for (int i = 0; i < M.nI; ++i) {
            if (i < M.nI - 1) F.faceE(i, j, k) = M.gE(i, j, k) * F(i, j, k) + (1. - M.gE(i, j, k)) * F(i + 1, j, k);
            if (j < M.nJ - 1) F.faceN(i, j, k) = M.gN(i, j, k) * F(i, j, k) + (1. - M.gN(i, j, k)) * F(i, j + 1, k);
            if (k < M.nK - 1) F.faceT(i, j, k) = M.gT(i, j, k) * F(i, j, k) + (1. - M.gT(i, j, k)) * F(i, j, k + 1);
        }
    }
    // FvScalarScheme.cpp:48-66
    void gradScalar(Field<double>& F, Field<V3>& G) {
        const Mesh& M = mesh;
        #pragma omp parallel for num_threads(nt()) schedule(static)
        for (int k = 0; k < M.nK; ++k) for (int j = 0; j < M.nJ; ++j) for (int i = 0; i < M.nI; ++i) {
            G(i, j, k) = (F.faceE(i, j, k) * M.SE(i, j, k) + F.faceW(i, j, k) * M.SW(i, j, k)
                          + F.faceN(i, j, k) * M.SN(i, j, k) + F.faceS(i, j, k) * M.SS(i, j, k)
                          + F.faceT(i, j, k) * M.ST(i, j, k) + F.faceB(i, j, k) * M.SB(i, j, k)) / M.vol(i, j, k);
        }
    }
    // FvVectorScheme.cpp:56-74
    void gradVector(Field<V3>& F, Field<T3>& G) {
        const Mesh& M = mesh;
        #pragma omp parallel for num_threads(nt()) schedule(static)
        for (int k = 0; k < M.nK; ++k) for (int j = 0; j < M.nJ; ++j) for (int i = 0; i < M.nI; ++i) {
            G(i, j, k) = (tensor(F.faceE(i, j, k), M.SE(i, j, k)) + tensor(F.faceW(i, j, k), M.SW(i, j, k))
                          + tensor(F.faceN(i, j, k), M.SN(i, j, k)) + tensor(F.faceS(i, j, k), M.SS(i, j, k))
                          + tensor(F.faceT(i, j, k), M.ST(i, j, k)) + tensor(F.faceB(i, j, k), M.SB(i, j, k))) / M.vol(i, j, k);
        }
    }

    // Simple.cpp:117-179
    void assembleMomentum(double dt) {
        const Mesh& M = mesh;
        uStar.copyFrom(u);
        const double om = cs.alphaU;
        #pragma omp parallel for num_threads(nt()) schedule(static)
        for (int k = 0; k < M.nK; ++k) for (int j = 0; j < M.nJ; ++j) for (int i = 0; i < M.nI; ++i) {
            if (!iMap.isActive(i, j, k)) continue;
            double a0P = cs.steady ? 0. : rho(i, j, k) * M.vol(i, j, k) / dt;
            double a[7];
            double mE = mflow.faceE(i, j, k), mW = mflow.faceW(i, j, k), mN = mflow.faceN(i, j, k);
            double mS = mflow.faceS(i, j, k), mT = mflow.faceT(i, j, k), mB = mflow.faceB(i, j, k);
            double muE = mu.faceE(i, j, k), muW = mu.faceW(i, j, k), muN = mu.faceN(i, j, k);
            double muS = mu.faceS(i, j, k), muT = mu.faceT(i, j, k), muB = mu.faceB(i, j, k);
            a[1] = std::min(mE, 0.) - muE * M.dE(i, j, k);
            a[2] = -std::max(mW, 0.) - muW * M.dW(i, j, k);
            a[3] = std::min(mN, 0.) - muN * M.dN(i, j, k);
            a[4] = -std::max(mS, 0.) - muS * M.dS(i, j, k);
            a[5] = std::min(mT, 0.) - muT * M.dT(i, j, k);
            a[6] = -std::max(mB, 0.) - muB * M.dB(i, j, k);
            a[0] = (std::max(mE, 0.) + muE * M.dE(i, j, k) - std::min(mW, 0.) + muW * M.dW(i, j, k)
                    + std::max(mN, 0.) + muN * M.dN(i, j, k) - std::min(mS, 0.) + muS * M.dS(i, j, k)
                    + std::max(mT, 0.) + muT * M.dT(i, j, k) - std::min(mB, 0.) + muB * M.dB(i, j, k) + a0P) / om;
            V3 b = a0P * u0(i, j, k);
            b += -M.vol(i, j, k) * gradP(i, j, k);
            b += (1. - om) * a[0] * uStar(i, j, k);
            b += muE * dot(gradU.faceE(i, j, k), M.cE(i, j, k));
            b += muW * dot(gradU.faceW(i, j, k), M.cW(i, j, k));
            b += muN * dot(gradU.faceN(i, j, k), M.cN(i, j, k));
            b += muS * dot(gradU.faceS(i, j, k), M.cS(i, j, k));
            b += muT * dot(gradU.faceT(i, j, k), M.cT(i, j, k));
            b += muB * dot(gradU.faceB(i, j, k), M.cB(i, j, k));
            uBC.fold(i, j, k, a, b);
            int row = iMap(i, j, k);
            int cols[7] = {row, iMap(i + 1, j, k), iMap(i - 1, j, k), iMap(i, j + 1, k), iMap(i, j - 1, k), iMap(i, j, k + 1), iMap(i, j, k - 1)};
            A[0].setRow(row, 7, cols, a);
            bv[0][row] = b.x; bv[1][row] = b.y; bv[2][row] = b.z;
            d(i, j, k) = M.vol(i, j, k) / a[0];
        }
    }

    // Simple.cpp:110-221
    void computeMomentum(double dt) {
        const Mesh& M = mesh;
        assembleMomentum(dt);
        itersMom = 0;
        for (int c = 0; c < 3; ++c) {
            itersMom += kspSolve(0, c, false);
            for (int k = 0; k < M.nK; ++k) for (int j = 0; j < M.nJ; ++j) for (int i = 0; i < M.nI; ++i) {
                if (iMap.isInactive(i, j, k)) continue;
                u(i, j, k)(c) = xv[c][iMap(i, j, k)];
                if (iMap.isActive(i, j, k)) h(i, j, k)(c) = u(i, j, k)(c) + d(i, j, k) * gradP(i, j, k)(c);
                else h(i, j, k)(c) = u(i, j, k)(c);
            }
        }
        uBC.setBoundaries(); dBC.setBoundaries(); hBC.setBoundaries();
        interpolate(d); interpolate(h);
        rhieChow();
    }

    // Simple.cpp:231-269
    void assemblePCorr() {
        const Mesh& M = mesh;
        #pragma omp parallel for num_threads(nt()) schedule(static)
        for (int k = 0; k < M.nK; ++k) for (int j = 0; j < M.nJ; ++j) for (int i = 0; i < M.nI; ++i) {
            if (!iMap.isActive(i, j, k)) continue;
            double a[7];
            a[1] = rho.faceE(i, j, k) * d.faceE(i, j, k) * M.dE(i, j, k);
            a[2] = rho.faceW(i, j, k) * d.faceW(i, j, k) * M.dW(i, j, k);
            a[3] = rho.faceN(i, j, k) * d.faceN(i, j, k) * M.dN(i, j, k);
            a[4] = rho.faceS(i, j, k) * d.faceS(i, j, k) * M.dS(i, j, k);
            a[5] = rho.faceT(i, j, k) * d.faceT(i, j, k) * M.dT(i, j, k);
            a[6] = rho.faceB(i, j, k) * d.faceB(i, j, k) * M.dB(i, j, k);
            a[0] = -(a[1] + a[2] + a[3] + a[4] + a[5] + a[6]);
            double b = mflow.faceE(i, j, k) - mflow.faceW(i, j, k) + mflow.faceN(i, j, k) - mflow.faceS(i, j, k)
                       + mflow.faceT(i, j, k) - mflow.faceB(i, j, k);
            pCBC.fold(i, j, k, a, b);
            int row = iMap(i, j, k);
            int cols[7] = {row, iMap(i + 1, j, k), iMap(i - 1, j, k), iMap(i, j + 1, k), iMap(i, j - 1, k), iMap(i, j, k + 1), iMap(i, j, k - 1)};
            A[1].setRow(row, 7, cols, a);
            bv[0][row] = b;
        }
    }

    // Simple.cpp:223-299
    void computePCorr() {
        const Mesh& M = mesh;
        assemblePCorr();
        bool proj = pcSingular && cs.algo != OC_IBPISO;
        itersPC = kspSolve(1, 0, proj);
        for (int k = 0; k < M.nK; ++k) for (int j = 0; j < M.nJ; ++j) for (int i = 0; i < M.nI; ++i) {
            if (iMap.isInactive(i, j, k)) continue;
            pC(i, j, k) = xv[0][iMap(i, j, k)];
        }
        pCBC.setBoundaries();
        interpolate(pC);
        gradScalar(pC, gradPC);
    }

    // Simple.cpp:301-356
    void correct() {
        const Mesh& M = mesh;
        const double aP = cs.alphaP;
        for (int k = 0; k < M.nK; ++k) for (int j = 0; j < M.nJ; ++j) for (int i = 0; i < M.nI; ++i) {
            if (iMap.isInactive(i, j, k)) continue;
            if (i == 0 && massCorr(1))
                mflow.faceW(i, j, k) += rho.faceW(i, j, k) * d.faceW(i, j, k) * (pC(i - 1, j, k) - pC(i, j, k)) * M.dW(i, j, k);
            if (j == 0 && massCorr(3))
                mflow.faceS(i, j, k) += rho.faceS(i, j, k) * d.faceS(i, j, k) * (pC(i, j - 1, k) - pC(i, j, k)) * M.dS(i, j, k);
            if (k == 0 && massCorr(5))
                mflow.faceB(i, j, k) += rho.faceB(i, j, k) * d.faceB(i, j, k) * (pC(i, j, k - 1) - pC(i, j, k)) * M.dB(i, j, k);
            if (i < M.nI - 1 || massCorr(0))
                mflow.faceE(i, j, k) += -rho.faceE(i, j, k) * d.faceE(i, j, k) * (pC(i + 1, j, k) - pC(i, j, k)) * M.dE(i, j, k);
            if (j < M.nJ - 1 || massCorr(2))
                mflow.faceN(i, j, k) += -rho.faceN(i, j, k) * d.faceN(i, j, k) * (pC(i, j + 1, k) - pC(i, j, k)) * M.dN(i, j, k);
            if (k < M.nK - 1 || massCorr(4))
                mflow.faceT(i, j, k) += -rho.faceT(i, j, k) * d.faceT(i, j, k) * (pC(i, j, k + 1) - pC(i, j, k)) * M.dT(i, j, k);
            mflow(i, j, k) = mflow.faceE(i, j, k) - mflow.faceW(i, j, k) + mflow.faceN(i, j, k) - mflow.faceS(i, j, k)
                             + mflow.faceT(i, j, k) - mflow.faceB(i, j, k);
            p(i, j, k) += aP * pC(i, j, k);
            u(i, j, k) += -d(i, j, k) * gradPC(i, j, k);
        }
        pBC.setBoundaries(); uBC.setBoundaries();
        interpolate(p); interpolate(u);
        gradScalar(p, gradP);
        gradVector(u, gradU);
        if (cs.nonOrth) gradUFaces();
    }

    // Corrected mode only (the reference never assigns gradUField_'s faces, Simple.cpp:355):
    // interior faces of grad u by the volume-weighted interpolation of FvScheme.tpp:12-40,
    // domain faces from the boundary cell's own gradient.
    void gradUFaces() {
        interpolate(gradU);
        const Mesh& M = mesh;
        for (int k = 0; k < M.nK; ++k) for (int j = 0; j < M.nJ; ++j) for (int i = 0; i < M.nI; ++i) {
            const T3 g = gradU(i, j, k);
            if (i == 0) gradU.faceW(i, j, k) = g;
            if (i == M.nI - 1) gradU.faceE(i, j, k) = g;
            if (j == 0) gradU.faceS(i, j, k) = g;
            if (j == M.nJ - 1) gradU.faceN(i, j, k) = g;
            if (k == 0) gradU.faceB(i, j, k) = g;
            if (k == M.nK - 1) gradU.faceT(i, j, k) = g;
        }
    }

    // Simple.cpp:410-449
    void rhieChow() {
        const Mesh& M = mesh;
        #pragma omp parallel for num_threads(nt()) schedule(static)
        for (int k = 0; k < M.nK; ++k) for (int j = 0; j < M.nJ; ++j) for (int i = 0; i < M.nI; ++i) {
            if (iMap.isInactive(i, j, k)) continue;
            if (i == 0) {
                mflow.faceW(i, j, k) = -dot(h.faceW(i, j, k), M.SW(i, j, k)) + d.faceW(i, j, k) * (p(i - 1, j, k) - p(i, j, k)) * M.dW(i, j, k);
                mflow.faceW(i, j, k) *= rho.faceW(i, j, k);
            }
            if (j == 0) {
                mflow.faceS(i, j, k) = -dot(h.faceS(i, j, k), M.SS(i, j, k)) + d.faceS(i, j, k) * (p(i, j - 1, k) - p(i, j, k)) * M.dS(i, j, k);
                mflow.faceS(i, j, k) *= rho.faceS(i, j, k);
            }
            if (k == 0) {
                mflow.faceB(i, j, k) = -dot(h.faceB(i, j, k), M.SB(i, j, k)) + d.faceB(i, j, k) * (p(i, j, k - 1) - p(i, j, k)) * M.dB(i, j, k);
                mflow.faceB(i, j, k) *= rho.faceB(i, j, k);
            }
            mflow.faceE(i, j, k) = dot(h.faceE(i, j, k), M.SE(i, j, k)) - d.faceE(i, j, k) * (p(i + 1, j, k) - p(i, j, k)) * M.dE(i, j, k);
            mflow.faceN(i, j, k) = dot(h.faceN(i, j, k), M.SN(i, j, k)) - d.faceN(i, j, k) * (p(i, j + 1, k) - p(i, j, k)) * M.dN(i, j, k);
            mflow.faceT(i, j, k) = dot(h.faceT(i, j, k), M.ST(i, j, k)) - d.faceT(i, j, k) * (p(i, j, k + 1) - p(i, j, k)) * M.dT(i, j, k);
            mflow.faceE(i, j, k) *= rho.faceE(i, j, k);
            mflow.faceN(i, j, k) *= rho.faceN(i, j, k);
            mflow.faceT(i, j, k) *= rho.faceT(i, j, k);
        }
    }

    // Simple.cpp:358-380: order-dependent scan, compared unnormalised, stored normalised
    double continuity() {
        const Mesh& M = mesh;
        double e = 0.;
        for (int k = 0; k < M.nK; ++k) for (int j = 0; j < M.nJ; ++j) for (int i = 0; i < M.nI; ++i) {
            if (!iMap.isActive(i, j, k)) continue;
            if (std::fabs(mflow(i, j, k)) > e) e = std::fabs(mflow(i, j, k)) / M.vol(i, j, k);
        }
        return e;
    }
    double continuityMax() {
        const Mesh& M = mesh;
        double e = 0.;
        for (int k = 0; k < M.nK; ++k) for (int j = 0; j < M.nJ; ++j) for (int i = 0; i < M.nI; ++i) {
            if (!iMap.isActive(i, j, k)) continue;
            double v = std::fabs(mflow(i, j, k)) / M.vol(i, j, k);
            if (v > e) e = v;
        }
        return e;
    }
    // Simple.cpp:382-408 (all cells)
    double courant(double dt) {
        const Mesh& M = mesh;
        double mx = 0.;
        for (int k = 0; k < M.nK; ++k) for (int j = 0; j < M.nJ; ++j) for (int i = 0; i < M.nI; ++i) {
            V3 rI = M.rfE(i, j, k) - M.rfW(i, j, k);
            V3 rJ = M.rfN(i, j, k) - M.rfS(i, j, k);
            V3 rK = M.rfT(i, j, k) - M.rfB(i, j, k);
            double uI = std::fabs(dot(u(i, j, k), unitVector(rI)));
            double uJ = std::fabs(dot(u(i, j, k), unitVector(rJ)));
            double uK = std::fabs(dot(u(i, j, k), unitVector(rK)));
            double co = dt * (uI / mag(rI) + uJ / mag(rJ) + uK / mag(rK));
            if (co > mx) mx = co;
        }
        return mx;
    }

    // IbPiso.cpp:65-117
    void computeIbField() {
        const Mesh& M = mesh;
        V3 cen(cs.sphereC[0], cs.sphereC[1], cs.sphereC[2]);
        for (int k = 0; k < M.nK; ++k) for (int j = 0; j < M.nJ; ++j) for (int i = 0; i < M.nI; ++i) {
            if (mag(M.xc(i, j, k) - cen) <= cs.sphereR) {  // Sphere.cpp:49-55
                ib(i, j, k) = 2; iMap.status(i, j, k) = INACTIVE;
                u(i, j, k) = V3(0., 0., 0.); p(i, j, k) = 0.;
            } else {
                ib(i, j, k) = 0; iMap.status(i, j, k) = ACTIVE;
            }
        }
        auto ibAt = [&](int i, int j, int k) -> int {  // off-domain reads the (FLUID) patch
            if (i < 0 || j < 0 || k < 0 || i >= M.nI || j >= M.nJ || k >= M.nK) return 0;
            return ib(i, j, k);
        };
        for (int k = 0; k < M.nK; ++k) for (int j = 0; j < M.nJ; ++j) for (int i = 0; i < M.nI; ++i) {
            if (ib(i, j, k) != 2) continue;
            if (ibAt(i + 1, j, k) == 0 || ibAt(i - 1, j, k) == 0 || ibAt(i, j + 1, k) == 0 || ibAt(i, j - 1, k) == 0
                || ibAt(i, j, k + 1) == 0 || ibAt(i, j, k - 1) == 0) {
                ib(i, j, k) = 1; iMap.status(i, j, k) = GHOST;
            }
        }
        iMap.generate();
        createMatrices(9);  // destroyMatrices(); createMatrices(2, 3, 9)
    }

    // IbPiso.cpp:119-190
    void computeIbCoeffs() {
        Mesh& M = mesh;
        V3 cen(cs.sphereC[0], cs.sphereC[1], cs.sphereC[2]);
        for (int k = 0; k < M.nK; ++k) for (int j = 0; j < M.nJ; ++j) for (int i = 0; i < M.nI; ++i) {
            if (ib(i, j, k) != 1) continue;
            V3 xc = M.xc(i, j, k);
            V3 bp = cen + cs.sphereR * unitVector(xc - cen);  // Sphere.cpp:23-30
            V3 ip = 2. * bp - xc;
            int ii[8], jj[8], kk[8], cols[9];
            double vals[9], beta[8];
            if (!M.locateEnclosingCells(ip, ii, jj, kk)) continue;
            V3 pts[8];
            for (int m = 0; m < 8; ++m) pts[m] = M.xc(ii[m], jj[m], kk[m]);
            trilinearCoeffs(pts, ip, beta);
            for (int q = 0; q < 9; ++q) vals[q] = 0.;
            vals[0] = 1.;
            cols[0] = iMap(i, j, k);
            for (int m = 0; m < 8; ++m) {
                if (ii[m] == i && jj[m] == j && kk[m] == k) { vals[0] += beta[m]; cols[m + 1] = -1; }
                else { vals[m + 1] = beta[m]; cols[m + 1] = iMap(ii[m], jj[m], kk[m]); }
            }
            A[0].setRow(iMap(i, j, k), 9, cols, vals);
            for (int m = 0; m < 8; ++m) {
                vals[0] = -1;  // reset inside the loop (IbPiso.cpp:170)
                cols[0] = iMap(i, j, k);
                if (ii[m] == i && jj[m] == j && kk[m] == k) { vals[0] += beta[m]; cols[m + 1] = -1; }
                else { vals[m + 1] = beta[m]; cols[m + 1] = iMap(ii[m], jj[m], kk[m]); }
            }
            A[1].setRow(iMap(i, j, k), 9, cols, vals);
        }
    }

    // ---- Diffusion (Diffusion.cpp:8-131)
    // Diffusion::Diffusion: phi CONSERVED, gradPhi AUXILLARY, PrimitiveBoundaryCondition<double>
    // (PrimitiveBoundaryCondition.tpp:3-30): every face's type (fixed / zeroGradient /
    // empty) and refValue, then setFixedBoundaries once, so that all six patches and
    // boundary faces start at their refValue; one 7-point matrix; uniform initial phi on
    // the cells.
    void initDiffusion(const oc_case* c) {
        phi.init(mesh, true);
        gradPhi.init(mesh, false);
        phiBC.init(phi);
        pcSingular = c->steady != 0;
        for (int i = 0; i < 6; ++i) {
            const int t = c->bcType[i];
            phiBC.types[i] = t == OC_FIXED ? P_FIXED : (t == OC_EMPTY ? P_EMPTY : P_ZERO_GRADIENT);
            phiBC.ref[i] = c->bcVal[i];
            if (phiBC.types[i] == P_FIXED) pcSingular = false;
        }
        phiBC.setFixed();
        A[0].allocate(iMap.nActive);
        bv[0].assign(iMap.nActive, 0.);
        xv[0].assign(iMap.nActive, 0.);
        phi.cells.assign(c->pInit);
    }
    // FvScalarScheme::extrapolateInteriorFaces (FvScalarScheme.cpp:70-160), one domain:
    // interpolate, gradient, then each interior E/N/T face from both cells' extrapolations,
    // clamped to the range of the low-side cell and its six neighbours
    void extrapolateFaces(Field<double>& F, Field<V3>& G) {
        const Mesh& M = mesh;
        interpolate(F);
        gradScalar(F, G);
        for (int k = 0; k < M.nK; ++k) for (int j = 0; j < M.nJ; ++j) for (int i = 0; i < M.nI; ++i) {
            double up = F(i, j, k), lo = F(i, j, k);
            const double nb[6] = {F(i + 1, j, k), F(i - 1, j, k), F(i, j + 1, k), F(i, j - 1, k), F(i, j, k + 1), F(i, j, k - 1)};
            for (int q = 0; q < 6; ++q) { up = std::max(up, nb[q]); lo = std::min(lo, nb[q]); }
            if (i < M.nI - 1) {
                double v = 0.5 * (F(i, j, k) + dot(G(i, j, k), M.rfE(i, j, k)) + F(i + 1, j, k) + dot(G(i + 1, j, k), M.rfW(i + 1, j, k)));
                F.faceE(i, j, k) = std::max(std::min(up, v), lo);
            }
            if (j < M.nJ - 1) {
                double v = 0.5 * (F(i, j, k) + dot(G(i, j, k), M.rfN(i, j, k)) + F(i, j + 1, k) + dot(G(i, j + 1, k), M.rfS(i, j + 1, k)));
                F.faceN(i, j, k) = std::max(std::min(up, v), lo);
            }
            if (k < M.nK - 1) {
                double v = 0.5 * (F(i, j, k) + dot(G(i, j, k), M.rfT(i, j, k)) + F(i, j, k + 1) + dot(G(i, j, k + 1), M.rfB(i, j, k + 1)));
                F.faceT(i, j, k) = std::max(std::min(up, v), lo);
            }
        }
    }
    // Diffusion::solve (Diffusion.cpp:33-131); returns ||A x - b||_inf
    double solveDiffusion(double dt) {
        const Mesh& M = mesh;
        const bool unsteady = cs.steady == 0;
        for (int k = 0; k < M.nK; ++k) for (int j = 0; j < M.nJ; ++j) for (int i = 0; i < M.nI; ++i) {
            const double a0P = unsteady ? M.vol(i, j, k) / dt : 0.;
            double a[7];
            a[1] = M.dE(i, j, k); a[2] = M.dW(i, j, k); a[3] = M.dN(i, j, k);
            a[4] = M.dS(i, j, k); a[5] = M.dT(i, j, k); a[6] = M.dB(i, j, k);
            a[0] = -(a[1] + a[2] + a[3] + a[4] + a[5] + a[6]) + a0P;
            double b = a0P * phi(i, j, k);
            phiBC.fold(i, j, k, a, b);
            const int cols[7] = {iMap(i, j, k), iMap(i + 1, j, k), iMap(i - 1, j, k), iMap(i, j + 1, k),
                                 iMap(i, j - 1, k), iMap(i, j, k + 1), iMap(i, j, k - 1)};
            A[0].setRow(iMap(i, j, k), 7, cols, a);
            bv[0][iMap(i, j, k)] = b;
        }
        std::fill(xv[0].begin(), xv[0].end(), 0.);
        itersPC = kspSolve(0, 0, pcSingular);
        for (int k = 0; k < M.nK; ++k) for (int j = 0; j < M.nJ; ++j) for (int i = 0; i < M.nI; ++i)
            phi(i, j, k) = xv[0][iMap(i, j, k)];
        phiBC.setBoundaries();
        extrapolateFaces(phi, gradPhi);
        gradScalar(phi, gradPhi);
        // scale(-1, b); multiplyAdd(A, x, b, res): res = A x - b
        std::vector<double> ax(bv[0].size());
        spmv(A[0], xv[0], ax, nt());
        double r = 0.;
        for (size_t q = 0; q < ax.size(); ++q) r = std::max(r, std::fabs(ax[q] + (-bv[0][q])));
        return r;
    }

    // ---- Coupled (Coupled.cpp:38-338)
    // Row numbering of IndexMap::operator()(i, j, k, varSetNo) (IndexMap.cpp:47-52): the
    // cell's row plus varSetNo * nActive; u, v, w are variable sets 0-2 and p is set 3.
    int row4(int i, int j, int k, int v) const { const int r = iMap(i, j, k); return r < 0 ? -1 : r + v * iMap.nActive; }

    // Coupled.cpp:54-228: the momentum rows (14 entries), d = V / a_P, the d boundaries and
    // faces, then the continuity rows (28 entries)
    void assembleCoupled(double dt) {
        const Mesh& M = mesh;
        uStar.copyFrom(u);
        const double om = cs.alphaU;
        for (int k = 0; k < M.nK; ++k) for (int j = 0; j < M.nJ; ++j) for (int i = 0; i < M.nI; ++i) {
            if (!iMap.isActive(i, j, k)) continue;
            const double a0P = cs.steady ? 0. : rho(i, j, k) * M.vol(i, j, k) / dt;
            const double mE = mflow.faceE(i, j, k), mW = mflow.faceW(i, j, k), mN = mflow.faceN(i, j, k);
            const double mS = mflow.faceS(i, j, k), mT = mflow.faceT(i, j, k), mB = mflow.faceB(i, j, k);
            const double muE = mu.faceE(i, j, k), muW = mu.faceW(i, j, k), muN = mu.faceN(i, j, k);
            const double muS = mu.faceS(i, j, k), muT = mu.faceT(i, j, k), muB = mu.faceB(i, j, k);
            double a[14];
            a[0] = (std::max(mE, 0.) + muE * M.dE(i, j, k) - std::min(mW, 0.) + muW * M.dW(i, j, k)
                    + std::max(mN, 0.) + muN * M.dN(i, j, k) - std::min(mS, 0.) + muS * M.dS(i, j, k)
                    + std::max(mT, 0.) + muT * M.dT(i, j, k) - std::min(mB, 0.) + muB * M.dB(i, j, k) + a0P) / om;
            a[1] = std::min(mE, 0.) - muE * M.dE(i, j, k);
            a[2] = -std::max(mW, 0.) - muW * M.dW(i, j, k);
            a[3] = std::min(mN, 0.) - muN * M.dN(i, j, k);
            a[4] = -std::max(mS, 0.) - muS * M.dS(i, j, k);
            a[5] = std::min(mT, 0.) - muT * M.dT(i, j, k);
            a[6] = -std::max(mB, 0.) - muB * M.dB(i, j, k);
            V3 bu = a0P * u0(i, j, k);
            bu += (1. - om) * a[0] * uStar(i, j, k);
            uBC.fold(i, j, k, a, bu);
            const V3 S[6] = {M.SE(i, j, k), M.SW(i, j, k), M.SN(i, j, k), M.SS(i, j, k), M.ST(i, j, k), M.SB(i, j, k)};
            const double g[6] = {M.gE(i, j, k), M.gW(i, j, k), M.gN(i, j, k), M.gS(i, j, k), M.gT(i, j, k), M.gB(i, j, k)};
            for (int c = 0; c < 3; ++c) {
                a[7] = g[0] * S[0](c) + g[1] * S[1](c) + g[2] * S[2](c) + g[3] * S[3](c) + g[4] * S[4](c) + g[5] * S[5](c);
                for (int f = 0; f < 6; ++f) a[8 + f] = (1. - g[f]) * S[f](c);
                const int row = row4(i, j, k, c);
                const int cols[14] = {row, row4(i + 1, j, k, c), row4(i - 1, j, k, c), row4(i, j + 1, k, c),
                                      row4(i, j - 1, k, c), row4(i, j, k + 1, c), row4(i, j, k - 1, c),
                                      row4(i, j, k, 3), row4(i + 1, j, k, 3), row4(i - 1, j, k, 3), row4(i, j + 1, k, 3),
                                      row4(i, j - 1, k, 3), row4(i, j, k + 1, 3), row4(i, j, k - 1, 3)};
                pBC.fold(i, j, k, &a[7], bu(c));
                Ac.setRow(row, 14, cols, a);
                bc[row] = bu(c);
            }
            d(i, j, k) = M.vol(i, j, k) / a[0];
        }
        dBC.setBoundaries();
        interpolate(d);
        for (int k = 0; k < M.nK; ++k) for (int j = 0; j < M.nJ; ++j) for (int i = 0; i < M.nI; ++i) {
            if (!iMap.isActive(i, j, k)) continue;
            const int row = row4(i, j, k, 3);
            double bp = 0., a[7];
            const V3 S[6] = {M.SE(i, j, k), M.SW(i, j, k), M.SN(i, j, k), M.SS(i, j, k), M.ST(i, j, k), M.SB(i, j, k)};
            const double g[6] = {M.gE(i, j, k), M.gW(i, j, k), M.gN(i, j, k), M.gS(i, j, k), M.gT(i, j, k), M.gB(i, j, k)};
            const double rf[6] = {rho.faceE(i, j, k), rho.faceW(i, j, k), rho.faceN(i, j, k), rho.faceS(i, j, k),
                                  rho.faceT(i, j, k), rho.faceB(i, j, k)};
            for (int c = 0; c < 3; ++c) {
                a[0] = rf[0] * g[0] * S[0](c) + rf[1] * g[1] * S[1](c) + rf[2] * g[2] * S[2](c) + rf[3] * g[3] * S[3](c)
                       + rf[4] * g[4] * S[4](c) + rf[5] * g[5] * S[5](c);
                for (int f = 0; f < 6; ++f) a[1 + f] = rf[f] * (1. - g[f]) * S[f](c);
                const int cols[7] = {row4(i, j, k, c), row4(i + 1, j, k, c), row4(i - 1, j, k, c), row4(i, j + 1, k, c),
                                     row4(i, j - 1, k, c), row4(i, j, k + 1, c), row4(i, j, k - 1, c)};
                V3 bu(0., 0., 0.);
                uBC.fold(i, j, k, a, bu);
                bp += bu(c);
                Ac.setRow(row, 7, cols, a);
            }
            const double df[6] = {d.faceE(i, j, k), d.faceW(i, j, k), d.faceN(i, j, k), d.faceS(i, j, k),
                                  d.faceT(i, j, k), d.faceB(i, j, k)};
            const double dd[6] = {M.dE(i, j, k), M.dW(i, j, k), M.dN(i, j, k), M.dS(i, j, k), M.dT(i, j, k), M.dB(i, j, k)};
            a[0] = rf[0] * df[0] * dd[0] + rf[1] * df[1] * dd[1] + rf[2] * df[2] * dd[2] + rf[3] * df[3] * dd[3]
                   + rf[4] * df[4] * dd[4] + rf[5] * df[5] * dd[5];
            for (int f = 0; f < 6; ++f) a[1 + f] = -rf[f] * df[f] * dd[f];
            const int cols[7] = {row, row4(i + 1, j, k, 3), row4(i - 1, j, k, 3), row4(i, j + 1, k, 3),
                                 row4(i, j - 1, k, 3), row4(i, j, k + 1, 3), row4(i, j, k - 1, 3)};
            const V3 gf[6] = {gradP.faceE(i, j, k), gradP.faceW(i, j, k), gradP.faceN(i, j, k), gradP.faceS(i, j, k),
                              gradP.faceT(i, j, k), gradP.faceB(i, j, k)};
            bp -= rf[0] * df[0] * dot(gf[0], S[0]) + rf[1] * df[1] * dot(gf[1], S[1]) + rf[2] * df[2] * dot(gf[2], S[2])
                  + rf[3] * df[3] * dot(gf[3], S[3]) + rf[4] * df[4] * dot(gf[4], S[4]) + rf[5] * df[5] * dot(gf[5], S[5]);
            pBC.fold(i, j, k, a, bp);
            Ac.setRow(row, 7, cols, a);
            bc[row] = bp;
        }
    }

    // Gauge of a Coupled system without a fixed-pressure face (tight mode; the reference
    // returns whatever its iteration reaches).  The continuity rows add up to zero, so the
    // system is singular; its null vector is n = e_p + z with A z = -A e_p (e_p: p = 1,
    // u = 0).  With planar faces A e_p = 0 (the face normals of a cell add up to zero) and n
    // is the constant pressure; on cells with non-planar faces (normals from one triangle of
    // each face, HexaFvmMesh.cpp:699-789) n has velocity components.  The solution is moved
    // along n to zero-mean p.
    void gaugeCoupled(const KspOpts& o) {
        const int N = iMap.nActive, n4 = 4 * N;
        std::vector<double> ep(n4, 0.), ae(n4), z(n4, 0.), r(n4);
        for (int q = 3 * N; q < n4; ++q) ep[q] = 1.;
        spmv(Ac, ep, ae, o.nThreads);
        double g2 = 0.;
        for (int q = 0; q < n4; ++q) { ae[q] = -ae[q]; g2 += ae[q] * ae[q]; }
        double xbar = 0., bbar = 0.;
        for (int q = 3 * N; q < n4; ++q) { xbar += xc[q]; bbar += bc[q]; }
        xbar /= N; bbar /= N;
        double b2 = 0.;
        for (int q = 0; q < n4; ++q) { const double v = q >= 3 * N ? bc[q] - bbar : bc[q]; b2 += v * v; }
        // a plain shift of p changes the residual by xbar * A e_p: below 1e-3 rtol it is kept
        double zbar = 0.;
        if (std::fabs(xbar) * std::sqrt(g2) > 1e-3 * o.rtol * std::sqrt(b2)) {
            double rr = 0.;
            bicgstabTight(Ac, ae, z, o, true, &rr, 3 * N, false);
            for (int q = 3 * N; q < n4; ++q) zbar += z[q];
            zbar /= N;
        }
        const double beta = xbar / (1. + zbar);
        for (int q = 0; q < n4; ++q) xc[q] -= beta * (ep[q] + z[q]);
    }

    // Coupled.cpp:282-314: no density factor, and the W/S/B domain faces take the pressure
    // difference with the sign of the E/N/T faces
    void rhieChowCoupled() {
        const Mesh& M = mesh;
        for (int k = 0; k < M.nK; ++k) for (int j = 0; j < M.nJ; ++j) for (int i = 0; i < M.nI; ++i) {
            if (iMap.isInactive(i, j, k)) continue;
            if (i == 0)
                mflow.faceW(i, j, k) = -dot(h.faceW(i, j, k), M.SW(i, j, k)) - d.faceW(i, j, k) * (p(i - 1, j, k) - p(i, j, k)) * M.dW(i, j, k);
            if (j == 0)
                mflow.faceS(i, j, k) = -dot(h.faceS(i, j, k), M.SS(i, j, k)) - d.faceS(i, j, k) * (p(i, j - 1, k) - p(i, j, k)) * M.dS(i, j, k);
            if (k == 0)
                mflow.faceB(i, j, k) = -dot(h.faceB(i, j, k), M.SB(i, j, k)) - d.faceB(i, j, k) * (p(i, j, k - 1) - p(i, j, k)) * M.dB(i, j, k);
            mflow.faceE(i, j, k) = dot(h.faceE(i, j, k), M.SE(i, j, k)) - d.faceE(i, j, k) * (p(i + 1, j, k) - p(i, j, k)) * M.dE(i, j, k);
            mflow.faceN(i, j, k) = dot(h.faceN(i, j, k), M.SN(i, j, k)) - d.faceN(i, j, k) * (p(i, j + 1, k) - p(i, j, k)) * M.dN(i, j, k);
            mflow.faceT(i, j, k) = dot(h.faceT(i, j, k), M.ST(i, j, k)) - d.faceT(i, j, k) * (p(i, j, k + 1) - p(i, j, k)) * M.dT(i, j, k);
        }
    }

    // Coupled.cpp:54-280.  The solve starts from x = 0 (KSPSolve's default guess).  The p
    // faces are never interpolated by this solver, so the divergence-theorem gradient sees
    // zero on interior faces and the patch on domain faces (Coupled.cpp:257-259).
    void computeMomentumAndPressure(double dt) {
        const Mesh& M = mesh;
        assembleCoupled(dt);
        std::fill(xc.begin(), xc.end(), 0.);
        const int N = iMap.nActive;
        KspOpts o = kopts();
        if (o.kind == OC_KSP_TIGHT) {
            itersPC = bicgstabTight(Ac, bc, xc, o, pcSingular, &lastRelres, 3 * N, false);
            if (pcSingular) gaugeCoupled(o);
        } else {
            itersPC = bicgstabLeft(Ac, bc, xc, o, &lastRelres);
        }
        for (int k = 0; k < M.nK; ++k) for (int j = 0; j < M.nJ; ++j) for (int i = 0; i < M.nI; ++i) {
            if (iMap.isInactive(i, j, k)) continue;
            for (int c = 0; c < 3; ++c) u(i, j, k)(c) = xc[row4(i, j, k, c)];
            p(i, j, k) = xc[row4(i, j, k, 3)];
        }
        pBC.setBoundaries();
        gradScalar(p, gradP);
        interpolate(gradP);
        for (int k = 0; k < M.nK; ++k) for (int j = 0; j < M.nJ; ++j) for (int i = 0; i < M.nI; ++i) {
            if (iMap.isActive(i, j, k)) h(i, j, k) = u(i, j, k) + d(i, j, k) * gradP(i, j, k);
            else if (iMap.status(i, j, k) == GHOST) h(i, j, k) = u(i, j, k);
        }
        hBC.setBoundaries();
        interpolate(h);
        uBC.setBoundaries();
        rhieChowCoupled();
    }

    // Coupled.cpp:316-338: every cell, |sum of outward mass flows| / V
    double continuityCoupled() {
        const Mesh& M = mesh;
        double e = 0.;
        for (int k = 0; k < M.nK; ++k) for (int j = 0; j < M.nJ; ++j) for (int i = 0; i < M.nI; ++i) {
            const double v = std::fabs(mflow.faceE(i, j, k) - mflow.faceW(i, j, k) + mflow.faceN(i, j, k) - mflow.faceS(i, j, k)
                                       + mflow.faceT(i, j, k) - mflow.faceB(i, j, k)) / M.vol(i, j, k);
            if (v > e) e = v;
        }
        return e;
    }

    double solveCoupled(double dt) {
        u0.copyFrom(u);
        const int nInner = cs.steady ? 1 : cs.nInner;  // Coupled.cpp:25-28
        for (int it = 0; it < nInner; ++it) computeMomentumAndPressure(dt);
        return continuityCoupled();
    }

    double solve(double dt) {
        if (cs.algo == OC_DIFFUSION) return solveDiffusion(dt);
        if (cs.algo == OC_COUPLED) return solveCoupled(dt);
        u0.copyFrom(u);
        if (cs.algo == OC_IBPISO) computeIbField();
        int nInner = cs.steady ? 1 : cs.nInner;  // Simple.cpp:61-64
        int nPC = cs.algo == OC_SIMPLE ? 1 : cs.nPCorr;
        for (int it = 0; it < nInner; ++it) {
            if (cs.algo == OC_IBPISO) computeIbCoeffs();
            computeMomentum(dt);
            for (int q = 0; q < nPC; ++q) { computePCorr(); correct(); }
        }
        return continuity();
    }

    // ---- padded-layout exchange
    size_t P() const { return (size_t)(mesh.nI + 2) * (mesh.nJ + 2) * (mesh.nK + 2); }
    size_t pidx(int i, int j, int k) const { return ((size_t)(k + 1) * (mesh.nJ + 2) + (j + 1)) * (mesh.nI + 2) + (i + 1); }
    // cells + patches of a field component into padded storage
    template <class T, class G>
    void exportCells(Field<T>& F, double* out, G get) {
        const int a = mesh.nI, b = mesh.nJ, c = mesh.nK;
        std::fill(out, out + P(), 0.);
        for (int k = -1; k <= c; ++k) for (int j = -1; j <= b; ++j) for (int i = -1; i <= a; ++i) {
            int outs = (i < 0 || i >= a) + (j < 0 || j >= b) + (k < 0 || k >= c);
            if (outs > 1) continue;
            out[pidx(i, j, k)] = get(F(i, j, k));
        }
    }
    template <class T, class S>
    void importCells(Field<T>& F, const double* in, S set) {
        const int a = mesh.nI, b = mesh.nJ, c = mesh.nK;
        for (int k = -1; k <= c; ++k) for (int j = -1; j <= b; ++j) for (int i = -1; i <= a; ++i) {
            int outs = (i < 0 || i >= a) + (j < 0 || j >= b) + (k < 0 || k >= c);
            if (outs > 1) continue;
            set(F(i, j, k), in[pidx(i, j, k)]);
        }
    }
    template <class T, class G>
    void exportFaces(Field<T>& F, int dir, double* out, G get) {
        std::fill(out, out + P(), 0.);
        A3<T>& A = dir == 0 ? F.fI : (dir == 1 ? F.fJ : F.fK);
        for (int k = 0; k < A.nK; ++k) for (int j = 0; j < A.nJ; ++j) for (int i = 0; i < A.nI; ++i)
            out[pidx(i, j, k)] = get(A(i, j, k));
    }
    template <class T, class S>
    void importFaces(Field<T>& F, int dir, const double* in, S set) {
        A3<T>& A = dir == 0 ? F.fI : (dir == 1 ? F.fJ : F.fK);
        for (int k = 0; k < A.nK; ++k) for (int j = 0; j < A.nJ; ++j) for (int i = 0; i < A.nI; ++i)
            set(A(i, j, k), in[pidx(i, j, k)]);
    }
    template <class T, class G>
    void exportMetricCells(const A3<T>& A, double* out, G get) {
        std::fill(out, out + P(), 0.);
        for (int k = 0; k < A.nK; ++k) for (int j = 0; j < A.nJ; ++j) for (int i = 0; i < A.nI; ++i)
            out[pidx(i, j, k)] = get(A(i, j, k));
    }
};

Oracle* O(void* h) { return static_cast<Oracle*>(h); }

template <class M>
void toCsr(const M& A, int* rowptr, int* cols, double* vals) {
    long q = 0;
    rowptr[0] = 0;
    for (size_t r = 0; r < A.rows.size(); ++r) {
        for (int e = 0; e < A.rows[r].n; ++e) { cols[q] = A.rows[r].col[e]; vals[q] = A.rows[r].val[e]; ++q; }
        rowptr[r + 1] = (int)q;
    }
}

}  // namespace

// ==================================================================== C ABI
extern "C" {

void* oc_create(const oc_case* c, const double* nodes) {
    Oracle* o = new Oracle();
    o->init(c, nodes);
    return o;
}
void oc_destroy(void* h) { delete O(h); }
double oc_solve(void* h, double dt) { return O(h)->solve(dt); }
void oc_save_u0(void* h) { O(h)->u0.copyFrom(O(h)->u); }
void oc_compute_momentum(void* h, double dt) { O(h)->computeMomentum(dt); }
void oc_assemble_momentum(void* h, double dt) { O(h)->assembleMomentum(dt); }
void oc_compute_pcorr(void* h) { O(h)->computePCorr(); }
void oc_assemble_pcorr(void* h) { O(h)->assemblePCorr(); }
void oc_correct(void* h) { O(h)->correct(); }
void oc_rhie_chow(void* h) { O(h)->rhieChow(); }
void oc_compute_ib_field(void* h) { O(h)->computeIbField(); }
void oc_compute_ib_coeffs(void* h) { O(h)->computeIbCoeffs(); }
double oc_continuity(void* h) { return O(h)->continuity(); }
double oc_continuity_max(void* h) { return O(h)->continuityMax(); }
double oc_courant(void* h, double dt) { return O(h)->courant(dt); }
void oc_grad_u(void* h) { O(h)->gradVector(O(h)->u, O(h)->gradU); }

long oc_get(void* h, const char* name, double* out) {
    Oracle& o = *O(h);
    std::string n(name);
    size_t P = o.P();
    auto comp = [](int c) { return [c](const V3& v) { return v(c); }; };
    auto sc = [](const double& v) { return v; };
    if (n == "u" || n == "u0" || n == "h" || n == "gradP" || n == "gradPCorr") {
        Field<V3>& F = n == "u" ? o.u : n == "u0" ? o.u0 : n == "h" ? o.h : n == "gradP" ? o.gradP : o.gradPC;
        for (int c = 0; c < 3; ++c) o.exportCells(F, out + c * P, comp(c));
        return (long)(3 * P);
    }
    if (n == "p" || n == "pCorr" || n == "d" || n == "massFlow") {
        Field<double>& F = n == "p" ? o.p : n == "pCorr" ? o.pC : n == "d" ? o.d : o.mflow;
        o.exportCells(F, out, sc);
        return (long)P;
    }
    if (n == "phi") { o.exportCells(o.phi, out, sc); return (long)P; }
    if (n == "gradPhi") {
        for (int c = 0; c < 3; ++c) o.exportCells(o.gradPhi, out + c * P, comp(c));
        return (long)(3 * P);
    }
    if (n == "gradU") {
        for (int c = 0; c < 9; ++c) o.exportCells(o.gradU, out + c * P, [c](const T3& t) { return t.a[c]; });
        return (long)(9 * P);
    }
    if (n == "mI" || n == "mJ" || n == "mK") {
        o.exportFaces(o.mflow, n == "mI" ? 0 : (n == "mJ" ? 1 : 2), out, sc);
        return (long)P;
    }
    if (n == "ib") {
        o.exportMetricCells(o.ib, out, [](const int& v) { return (double)v; });
        return (long)P;
    }
    return -1;
}

long oc_set(void* h, const char* name, const double* in) {
    Oracle& o = *O(h);
    std::string n(name);
    size_t P = o.P();
    // Fields whose face values the hot path reads (d, h, p, p', u) get faces consistent
    // with the new cells: interior faces interpolated (FvScheme.tpp:12-40), domain faces
    // equal to the patch (Field.tpp:280-362), as they are at every point of use.
    if (n == "u" || n == "u0" || n == "h" || n == "gradP" || n == "gradPCorr") {
        Field<V3>& F = n == "u" ? o.u : n == "u0" ? o.u0 : n == "h" ? o.h : n == "gradP" ? o.gradP : o.gradPC;
        for (int c = 0; c < 3; ++c) o.importCells(F, in + c * P, [c](V3& v, double x) { v(c) = x; });
        if (n == "u" || n == "h") { o.interpolate(F); for (int f = 0; f < 6; ++f) F.facesFromPatch(f); }
        return (long)(3 * P);
    }
    if (n == "p" || n == "pCorr" || n == "d" || n == "massFlow") {
        Field<double>& F = n == "p" ? o.p : n == "pCorr" ? o.pC : n == "d" ? o.d : o.mflow;
        o.importCells(F, in, [](double& v, double x) { v = x; });
        if (n != "massFlow") { o.interpolate(F); for (int f = 0; f < 6; ++f) F.facesFromPatch(f); }
        return (long)P;
    }
    if (n == "mI" || n == "mJ" || n == "mK") {
        o.importFaces(o.mflow, n == "mI" ? 0 : (n == "mJ" ? 1 : 2), in, [](double& v, double x) { v = x; });
        return (long)P;
    }
    if (n == "gradU") {  // cells; in the corrected mode the faces follow as correct() forms them
        for (int c = 0; c < 9; ++c) o.importCells(o.gradU, in + c * P, [c](T3& t, double x) { t.a[c] = x; });
        if (o.cs.nonOrth) o.gradUFaces();
        return (long)(9 * P);
    }
    return -1;
}

long oc_get_metric(void* h, const char* name, double* out) {
    Oracle& o = *O(h);
    Mesh& M = o.mesh;
    std::string n(name);
    size_t P = o.P();
    auto sc = [](const double& v) { return v; };
    auto comp = [](int c) { return [c](const V3& v) { return v(c); }; };
    if (n == "vol") { o.exportMetricCells(M.vol, out, sc); return (long)P; }
    if (n == "xc") { for (int c = 0; c < 3; ++c) o.exportMetricCells(M.xc, out + c * P, comp(c)); return (long)(3 * P); }
    if (n == "SI" || n == "SJ" || n == "SK") {
        A3<V3>& A = n == "SI" ? M.nI_ : (n == "SJ" ? M.nJ_ : M.nK_);
        for (int c = 0; c < 3; ++c) o.exportMetricCells(A, out + c * P, comp(c));
        return (long)(3 * P);
    }
    if (n == "fcI" || n == "fcJ" || n == "fcK") {
        A3<V3>& A = n == "fcI" ? M.fcI : (n == "fcJ" ? M.fcJ : M.fcK);
        for (int c = 0; c < 3; ++c) o.exportMetricCells(A, out + c * P, comp(c));
        return (long)(3 * P);
    }
    const char* dn[6] = {"dE", "dW", "dN", "dS", "dT", "dB"};
    const char* gn[6] = {"gE", "gW", "gN", "gS", "gT", "gB"};
    const char* cn[6] = {"cE", "cW", "cN", "cS", "cT", "cB"};
    A3<double>* ds[6] = {&M.dE, &M.dW, &M.dN, &M.dS, &M.dT, &M.dB};
    A3<double>* gs[6] = {&M.gE, &M.gW, &M.gN, &M.gS, &M.gT, &M.gB};
    A3<V3>* cvs[6] = {&M.cE, &M.cW, &M.cN, &M.cS, &M.cT, &M.cB};
    for (int f = 0; f < 6; ++f) {
        if (n == dn[f]) { o.exportMetricCells(*ds[f], out, sc); return (long)P; }
        if (n == gn[f]) { o.exportMetricCells(*gs[f], out, sc); return (long)P; }
        if (n == cn[f]) { for (int c = 0; c < 3; ++c) o.exportMetricCells(*cvs[f], out + c * P, comp(c)); return (long)(3 * P); }
    }
    return -1;
}

int oc_nrows(void* h) { return O(h)->iMap.nActive; }
long oc_matrix_nnz(void* h, int which) {
    long s = 0;
    if (which == 2) { for (const auto& R : O(h)->Ac.rows) s += R.n; return s; }
    for (const auto& R : O(h)->A[which].rows) s += R.n;
    return s;
}
void oc_matrix_csr(void* h, int which, int* rowptr, int* cols, double* vals) {
    if (which == 2) toCsr(O(h)->Ac, rowptr, cols, vals);
    else toCsr(O(h)->A[which], rowptr, cols, vals);
}
void oc_rhs(void* h, int vec, double* out) {
    const std::vector<double>& b = vec == 3 ? O(h)->bc : O(h)->bv[vec];
    std::copy(b.begin(), b.end(), out);
}
void oc_imap(void* h, int* rowid, int* status) {
    IndexMap& m = O(h)->iMap;
    for (size_t q = 0; q < m.gidx.v.size(); ++q) { rowid[q] = m.gidx.v[q]; status[q] = m.status.v[q]; }
}
void oc_stats(void* h, int* im, int* ip, double* rr) {
    if (im) *im = O(h)->itersMom;
    if (ip) *ip = O(h)->itersPC;
    if (rr) *rr = O(h)->lastRelres;
}

static void load8(const double* p, V3* t) { for (int q = 0; q < 8; ++q) t[q] = V3(p[3 * q], p[3 * q + 1], p[3 * q + 2]); }
double oc_hex_volume(const double* p) { V3 t[8]; load8(p, t); return hexVolume(t); }
void oc_hex_centroid(const double* p, double* o) { V3 t[8]; load8(p, t); V3 c = hexCentroid(t); o[0] = c.x; o[1] = c.y; o[2] = c.z; }
double oc_quad_area(const double* p) { V3 t[4]; for (int q = 0; q < 4; ++q) t[q] = V3(p[3 * q], p[3 * q + 1], p[3 * q + 2]); return quadArea(t); }
void oc_quad_centroid(const double* p, double* o) {
    V3 t[4]; for (int q = 0; q < 4; ++q) t[q] = V3(p[3 * q], p[3 * q + 1], p[3 * q + 2]);
    V3 c = quadCentroid(t); o[0] = c.x; o[1] = c.y; o[2] = c.z;
}
int oc_inside_hex(const double* pt, const double* p) { V3 t[8]; load8(p, t); return insideHex(V3(pt[0], pt[1], pt[2]), t) ? 1 : 0; }
void oc_trilinear(const double* p, const double* x, double* beta) { V3 t[8]; load8(p, t); trilinearCoeffs(t, V3(x[0], x[1], x[2]), beta); }

// HexaMeshGen.cpp:32-79 / StructuredMesh.cpp:77-129: transfinite fill from 8 vertices
void oc_transfinite_nodes(const double* vv, int nNI, int nNJ, int nNK, double* out) {
    V3 v[8];
    load8(vv, v);
    A3<V3> nd;
    nd.resize(nNI, nNJ, nNK);
    int uI = nNI - 1, uJ = nNJ - 1, uK = nNK - 1;
    for (int k = 0; k <= uK; ++k) {
        double s = double(k) / double(uK);
        nd(0, 0, k) = (v[4] - v[0]) * s + v[0];
        nd(0, uJ, k) = (v[7] - v[3]) * s + v[3];
        nd(uI, 0, k) = (v[5] - v[1]) * s + v[1];
        nd(uI, uJ, k) = (v[6] - v[2]) * s + v[2];
        V3 t1 = nd(0, uJ, k) - nd(0, 0, k), t2 = nd(uI, uJ, k) - nd(uI, 0, k);
        for (int j = 0; j <= uJ; ++j) {
            s = double(j) / double(uJ);
            nd(0, j, k) = t1 * s + nd(0, 0, k);
            nd(uI, j, k) = t2 * s + nd(uI, 0, k);
        }
    }
    for (int k = 0; k <= uK; ++k) for (int j = 0; j <= uJ; ++j) {
        V3 t1 = nd(uI, j, k) - nd(0, j, k);
        for (int i = 0; i <= uI; ++i) {
            double s = double(i) / double(uI);
            nd(i, j, k) = t1 * s + nd(0, j, k);
        }
    }
    for (size_t q = 0; q < nd.v.size(); ++q) { out[3 * q] = nd.v[q].x; out[3 * q + 1] = nd.v[q].y; out[3 * q + 2] = nd.v[q].z; }
}

}  // extern "C"
