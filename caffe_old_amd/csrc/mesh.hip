// mesh.hip -- finite-volume metrics of a general structured hex mesh, computed on
// the device from the node array with the reference's formulas and operation order
// (HexaFvmMesh.cpp:626-918, Geometry.cpp:5-109, 258-277), so they are bit-identical
// to a CPU evaluation built without contraction.  Uniform boxes skip this and use
// closed forms (common.hpp, G<true>).
#include "internal.hpp"

namespace caffe {

__device__ inline double quadAreaD(const V3* p) {  // Geometry.cpp:5-12
    V3 r01 = p[1] - p[0], r02 = p[2] - p[0], r03 = p[3] - p[0];
    return 0.5 * (mag(cross(r02, r01)) + mag(cross(r03, r02)));
}
__device__ inline V3 quadCentroidD(const V3* p) {  // Geometry.cpp:14-20
    V3 d1 = p[0] + 0.5 * (p[2] - p[0]);
    V3 d2 = p[1] + 0.5 * (p[3] - p[1]);
    return d1 + 0.5 * (d2 - d1);
}
__device__ inline double hexVolumeD(const V3* p) {  // Geometry.cpp:40-59
    V3 s1265 = 0.5 * cross(p[6] - p[1], p[5] - p[2]);
    V3 s5674 = 0.5 * cross(p[7] - p[5], p[4] - p[6]);
    V3 s2376 = 0.5 * cross(p[7] - p[2], p[6] - p[3]);
    return fabs(dot(s1265 + s5674 + s2376, p[6] - p[0]) / 3.);
}
__device__ inline double tetVolumeD(const V3& a0, const V3& a1, const V3& a2, const V3& a3) {
    V3 a = a1 - a0, b = a2 - a0, c = a3 - a0;
    return fabs(dot(a, cross(b, c))) / 6.;
}
__device__ inline V3 tetCentroidD(const V3& a0, const V3& a1, const V3& a2, const V3& a3) {
    V3 s = mk(0., 0., 0.);
    s = s + a0; s = s + a1; s = s + a2; s = s + a3;
    return 0.25 * s;
}
__device__ inline V3 hexCentroidD(const V3* p) {  // Geometry.cpp:61-109
    const int T[5][4] = {{4, 3, 6, 7}, {4, 1, 6, 5}, {4, 1, 6, 3}, {4, 0, 1, 3}, {3, 1, 2, 6}};
    V3 cent[5];
    double vol[5], sum = 0.;
    for (int i = 0; i < 5; ++i) {
        cent[i] = tetCentroidD(p[T[i][0]], p[T[i][1]], p[T[i][2]], p[T[i][3]]);
        vol[i] = tetVolumeD(p[T[i][0]], p[T[i][1]], p[T[i][2]], p[T[i][3]]);
        sum += vol[i];
    }
    V3 c = mk(0., 0., 0.);
    for (int i = 0; i < 5; ++i) c = c + vol[i] * cent[i];
    return c / sum;
}

struct MeshOut {
    const double* nodes;  // (nI+1)(nJ+1)(nK+1) x 3, i fastest
    double* vol;
    double* xc[3];
    double* S[3][3];
    double* fc[3][3];
    double* d[3];
    double* g[3];
    double* c[3][3];      // c = S - r d seen from the low-side cell (null: not computed)
};

__device__ inline V3 node(const Dims& D, const double* n, int i, int j, int k) {
    long q = ((long)k * (D.nJ + 1) + j) * (D.nI + 1) + i;
    return mk(n[3 * q], n[3 * q + 1], n[3 * q + 2]);
}

__global__ void k_cells(Dims D, MeshOut m) {
    long q = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= D.N) return;
    int i, j, k;
    long p = D.cell(q, i, j, k);
    V3 t[8] = {node(D, m.nodes, i, j, k), node(D, m.nodes, i + 1, j, k), node(D, m.nodes, i + 1, j + 1, k), node(D, m.nodes, i, j + 1, k),
               node(D, m.nodes, i, j, k + 1), node(D, m.nodes, i + 1, j, k + 1), node(D, m.nodes, i + 1, j + 1, k + 1), node(D, m.nodes, i, j + 1, k + 1)};
    V3 c = hexCentroidD(t);
    m.xc[0][p] = c.x; m.xc[1][p] = c.y; m.xc[2][p] = c.z;
    m.vol[p] = hexVolumeD(t);
}

// HexaFvmMesh::initializeFaces (HexaFvmMesh.cpp:699-789): normal = unit(cross(p1-p0, p2-p0)) * area
__global__ void k_faces(Dims D, MeshOut m, int dir) {
    const int ni = D.nI + (dir == 0), nj = D.nJ + (dir == 1), nk = D.nK + (dir == 2);
    long q = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= (long)ni * nj * nk) return;
    int i = (int)(q % ni); long r = q / ni; int j = (int)(r % nj); int k = (int)(r / nj);
    V3 t[4];
    if (dir == 0) { t[0] = node(D, m.nodes, i, j, k); t[1] = node(D, m.nodes, i, j + 1, k); t[2] = node(D, m.nodes, i, j + 1, k + 1); t[3] = node(D, m.nodes, i, j, k + 1); }
    else if (dir == 1) { t[0] = node(D, m.nodes, i, j, k); t[1] = node(D, m.nodes, i, j, k + 1); t[2] = node(D, m.nodes, i + 1, j, k + 1); t[3] = node(D, m.nodes, i + 1, j, k); }
    else { t[0] = node(D, m.nodes, i, j, k); t[1] = node(D, m.nodes, i + 1, j, k); t[2] = node(D, m.nodes, i + 1, j + 1, k); t[3] = node(D, m.nodes, i, j + 1, k); }
    V3 tv = unitv(cross(t[1] - t[0], t[2] - t[0]));
    double ar = quadAreaD(t);
    V3 c = quadCentroidD(t);
    V3 s = tv * ar;
    long f = D.pidx(i, j, k);
    m.fc[dir][0][f] = c.x; m.fc[dir][1][f] = c.y; m.fc[dir][2][f] = c.z;
    m.S[dir][0][f] = s.x; m.S[dir][1][f] = s.y; m.S[dir][2][f] = s.z;
}

// HexaFvmMesh::computeMeshMetrics (HexaFvmMesh.cpp:841-918), one value per face.
// Interior: dE(low) = |S|^2 / (S . (xc_hi - xc_lo)), which equals dW(high) bit for bit;
// gE(low) = V_hi / (V_hi + V_lo).  Domain faces use the cell-to-face vector
// (HexaFvmMesh.cpp:218-294) and g = 0.
__global__ void k_face_metrics(Dims D, MeshOut m, int dir) {
    const int ni = D.nI + (dir == 0), nj = D.nJ + (dir == 1), nk = D.nK + (dir == 2);
    long q = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= (long)ni * nj * nk) return;
    int i = (int)(q % ni); long r = q / ni; int j = (int)(r % nj); int k = (int)(r / nj);
    const long s = dir == 0 ? 1 : (dir == 1 ? D.sJ : D.sK);
    const int idx = dir == 0 ? i : (dir == 1 ? j : k);
    const int nmax = dir == 0 ? D.nI : (dir == 1 ? D.nJ : D.nK);
    long f = D.pidx(i, j, k);
    V3 S = mk(m.S[dir][0][f], m.S[dir][1][f], m.S[dir][2][f]);
    V3 fcv = mk(m.fc[dir][0][f], m.fc[dir][1][f], m.fc[dir][2][f]);
    double dval, gval;
    // c = S - r (S.S) / (S.r) (HexaFvmMesh.cpp:879-884; Vector3D's operator/ multiplies by
    // the reciprocal), stored as seen from the low-side cell: the high-side cell's c is its
    // negation bit for bit, and on a west/south/bottom domain face (seen only from the cell
    // above it) the negation of that cell's c is stored
    V3 cv;
    if (idx == 0) {  // west/south/bottom domain face, seen from cell f with normal -S
        V3 Sw = -S;
        V3 xcv = mk(m.xc[0][f], m.xc[1][f], m.xc[2][f]);
        V3 rf = fcv - xcv;
        dval = dot(Sw, Sw) / dot(Sw, rf);
        gval = 0.;
        cv = -(Sw - rf * dot(Sw, Sw) / dot(Sw, rf));
    } else if (idx == nmax) {  // east/north/top domain face, seen from cell f - s
        V3 xcv = mk(m.xc[0][f - s], m.xc[1][f - s], m.xc[2][f - s]);
        V3 rf = fcv - xcv;
        dval = dot(S, S) / dot(S, rf);
        gval = 0.;
        cv = S - rf * dot(S, S) / dot(S, rf);
    } else {
        V3 lo = mk(m.xc[0][f - s], m.xc[1][f - s], m.xc[2][f - s]);
        V3 hi = mk(m.xc[0][f], m.xc[1][f], m.xc[2][f]);
        V3 rc = hi - lo;
        dval = dot(S, S) / dot(S, rc);
        gval = m.vol[f] / (m.vol[f] + m.vol[f - s]);
        cv = S - rc * dot(S, S) / dot(S, rc);
    }
    m.d[dir][f] = dval;
    m.g[dir][f] = gval;
    if (m.c[dir][0]) { m.c[dir][0][f] = cv.x; m.c[dir][1][f] = cv.y; m.c[dir][2][f] = cv.z; }
}

void compute_metrics(Ctx& c, const double* dNodes) {
    MeshOut m;
    m.nodes = dNodes;
    m.vol = const_cast<double*>(c.geo.vol);
    for (int a = 0; a < 3; ++a) {
        m.xc[a] = const_cast<double*>(c.geo.xc[a]);
        m.d[a] = const_cast<double*>(c.geo.dF[a]);
        m.g[a] = const_cast<double*>(c.geo.gF[a]);
        for (int b = 0; b < 3; ++b) {
            m.S[a][b] = const_cast<double*>(c.geo.S[a][b]);
            m.fc[a][b] = const_cast<double*>(c.geo.fc[a][b]);
            m.c[a][b] = const_cast<double*>(c.geo.cF[a][b]);
        }
    }
    const Dims& D = c.D;
    hipLaunchKernelGGL(k_cells, dim3(blocksFor(D.N, 256)), dim3(256), 0, c.stream, D, m);
    for (int dir = 0; dir < 3; ++dir) {
        long nf = (long)(D.nI + (dir == 0)) * (D.nJ + (dir == 1)) * (D.nK + (dir == 2));
        hipLaunchKernelGGL(k_faces, dim3(blocksFor(nf, 256)), dim3(256), 0, c.stream, D, m, dir);
    }
    for (int dir = 0; dir < 3; ++dir) {
        long nf = (long)(D.nI + (dir == 0)) * (D.nJ + (dir == 1)) * (D.nK + (dir == 2));
        hipLaunchKernelGGL(k_face_metrics, dim3(blocksFor(nf, 256)), dim3(256), 0, c.stream, D, m, dir);
    }
}

}  // namespace caffe
