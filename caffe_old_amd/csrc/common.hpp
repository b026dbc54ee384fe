// common.hpp -- device-side layout, geometry access and deterministic reductions
// shared by every CAFFE hot-path kernel.
//
// Layout: every cell array is padded to (nI+2)*(nJ+2)*(nK+2), i fastest.  The
// one-cell halo is the Field's six boundary patches (Field.tpp:89-128: (-1,j,k)
// is the west patch), so stencils never branch on the domain edge.  Face arrays
// use the same shape: entry p of an I-face array is the WEST face of padded cell
// p, so the east face of cell p is entry p+1 (the last one lands in the halo
// slot).  This keeps faceE(i) == faceW(i+1) a single storage location, as
// Field::facesI_ does (Field.h:96-101).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace caffe {

enum : int8_t { ST_ACTIVE = 0, ST_GHOST = 1, ST_INACTIVE = 2 };  // IndexMap.h:39
enum : int { BC_FIXED = 0, BC_ZERO_GRADIENT = 1, BC_EMPTY = 2 };    // PrimitiveBoundaryCondition.h:16
enum : int { F_INLET = 0, F_OUTLET = 1, F_WALL = 2, F_EMPTY = 3, F_ZEROGRAD = 4 };  // FlowBoundaryConditions.h:10

struct V3 { double x, y, z; };
__host__ __device__ inline V3 mk(double x, double y, double z) { V3 v; v.x = x; v.y = y; v.z = z; return v; }
__host__ __device__ inline V3 operator+(V3 a, V3 b) { return mk(a.x + b.x, a.y + b.y, a.z + b.z); }
__host__ __device__ inline V3 operator-(V3 a, V3 b) { return mk(a.x - b.x, a.y - b.y, a.z - b.z); }
__host__ __device__ inline V3 operator-(V3 a) { return mk(-a.x, -a.y, -a.z); }
__host__ __device__ inline V3 operator*(V3 a, double s) { return mk(a.x * s, a.y * s, a.z * s); }
__host__ __device__ inline V3 operator*(double s, V3 a) { return mk(a.x * s, a.y * s, a.z * s); }
// Vector3D::operator/= multiplies by the reciprocal (Vector3D.cpp:98-107)
__host__ __device__ inline V3 operator/(V3 a, double s) { s = 1. / s; return mk(a.x * s, a.y * s, a.z * s); }
__host__ __device__ inline double dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__host__ __device__ inline V3 cross(V3 u, V3 v) {
    return mk(u.y * v.z - u.z * v.y, u.z * v.x - u.x * v.z, u.x * v.y - u.y * v.x);
}
__host__ __device__ inline double mag(V3 u) { return sqrt(u.x * u.x + u.y * u.y + u.z * u.z); }
__host__ __device__ inline V3 unitv(V3 u) { double inv = 1. / mag(u); return mk(inv * u.x, inv * u.y, inv * u.z); }

// Padded grid dimensions (32-bit indices: the padded grid holds < 2^31 cells).
struct Dims {
    int nI, nJ, nK;
    int sJ;      // nI + 2
    int sK;      // (nI + 2) * (nJ + 2)
    int P;       // padded cell count
    int N;       // interior cell count
    __host__ __device__ int pidx(int i, int j, int k) const { return (i + 1) + (j + 1) * sJ + (k + 1) * sK; }
    // interior linear id (i fastest) -> padded index and (i,j,k)
    __host__ __device__ int cell(int q, int& i, int& j, int& k) const {
        const unsigned uq = (unsigned)q, ni = (unsigned)nI, nj = (unsigned)nJ;
        const unsigned r = uq / ni;
        i = (int)(uq - r * ni);
        const unsigned kk = r / nj;
        j = (int)(r - kk * nj);
        k = (int)kk;
        return pidx(i, j, k);
    }
};

// Mesh metrics.  UNIFORM meshes (a box of equal cells) evaluate every metric in
// closed form from (hx, hy, hz); general meshes read per-face/per-cell arrays
// computed on device from the node array with the reference's formulas
// (HexaFvmMesh.cpp:626-918).
struct Geo {
    Dims D;
    double hx, hy, hz;
    const double* vol;
    const double* S[3][3];   // S[dir][xyz]: area normal of the low (W/S/B) face of each cell
    const double* dF[3];     // d = |S|^2 / (S . r) of that face
    const double* gF[3];     // gE/gN/gT of the low-side cell (weight of the low cell's value)
    const double* cF[3][3];  // c = S - r d of that face, seen from the low-side cell
    const double* xc[3];     // cell centroid
    const double* fc[3][3];  // face centroid of the low face
};

template <bool U> struct G;

template <> struct G<true> {
    __device__ static double vol(const Geo& g, long) { return g.hx * g.hy * g.hz; }
    __device__ static V3 S(const Geo& g, int dir, long) {
        return dir == 0 ? mk(g.hy * g.hz, 0., 0.) : (dir == 1 ? mk(0., g.hx * g.hz, 0.) : mk(0., 0., g.hx * g.hy));
    }
    // d at a face: |S| / |r|, r = cell-to-cell distance, or cell-to-face (h/2) on the domain boundary
    __device__ static double d(const Geo& g, int dir, long, bool bnd) {
        double v = dir == 0 ? g.hy * g.hz / g.hx : (dir == 1 ? g.hx * g.hz / g.hy : g.hx * g.hy / g.hz);
        return bnd ? 2. * v : v;
    }
    __device__ static double gw(const Geo&, int, long) { return 0.5; }
    __device__ static V3 xc(const Geo& g, int i, int j, int k, long) {
        return mk((i + 0.5) * g.hx, (j + 0.5) * g.hy, (k + 0.5) * g.hz);
    }
    __device__ static V3 fc(const Geo& g, int dir, int i, int j, int k, long) {
        return dir == 0 ? mk(i * g.hx, (j + 0.5) * g.hy, (k + 0.5) * g.hz)
             : (dir == 1 ? mk((i + 0.5) * g.hx, j * g.hy, (k + 0.5) * g.hz) : mk((i + 0.5) * g.hx, (j + 0.5) * g.hy, k * g.hz));
    }
};

template <> struct G<false> {
    __device__ static double vol(const Geo& g, long p) { return g.vol[p]; }
    __device__ static V3 S(const Geo& g, int dir, long f) { return mk(g.S[dir][0][f], g.S[dir][1][f], g.S[dir][2][f]); }
    __device__ static double d(const Geo& g, int dir, long f, bool) { return g.dF[dir][f]; }
    __device__ static double gw(const Geo& g, int dir, long f) { return g.gF[dir][f]; }
    __device__ static V3 xc(const Geo& g, int, int, int, long p) { return mk(g.xc[0][p], g.xc[1][p], g.xc[2][p]); }
    __device__ static V3 fc(const Geo& g, int dir, int, int, int, long f) { return mk(g.fc[dir][0][f], g.fc[dir][1][f], g.fc[dir][2][f]); }
};

// ------------------------------------------------------------------ reductions
// Deterministic block reduction: wave64 butterfly then a fixed LDS tree.
template <int BS>
__device__ inline double blockSum(double v, double* lds) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    if (lane == 0) lds[wid] = v;
    __syncthreads();
    double r = 0.;
    if (threadIdx.x == 0) {
        for (int w = 0; w < BS / 64; ++w) r += lds[w];
    }
    __syncthreads();
    return r;  // valid in thread 0
}
template <int BS>
__device__ inline double blockMax(double v, double* lds) {
    for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    if (lane == 0) lds[wid] = v;
    __syncthreads();
    double r = 0.;
    if (threadIdx.x == 0) {
        r = lds[0];
        for (int w = 1; w < BS / 64; ++w) r = fmax(r, lds[w]);
    }
    __syncthreads();
    return r;
}

// Deterministic grid reductions without device-scope fences.  A producing kernel
// writes one row of up to RED_STRIDE partial sums per block; a consuming kernel (or a
// one-block finalize kernel) adds the rows in block order.  Kernel boundaries give
// the visibility, so no block ever waits for, or fences against, another.
constexpr int MAX_RED_BLOCKS = 1024;
constexpr int RED_STRIDE = 16;   // partial sums per row (at most)

template <int BS, int K>
__device__ inline void writePartials(const double (&v)[K], double* buf) {
    static_assert(K <= RED_STRIDE, "more partial sums than a row holds");
    __shared__ double lds[BS / 64][K];
    double w[K];
#pragma unroll
    for (int q = 0; q < K; ++q) w[q] = v[q];
    // butterfly rounds outside, sums inside: the K exchanges of a round are independent
#pragma unroll
    for (int o = 32; o > 0; o >>= 1)
#pragma unroll
        for (int q = 0; q < K; ++q) w[q] += __shfl_xor(w[q], o, 64);
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    if (lane == 0)
#pragma unroll
        for (int q = 0; q < K; ++q) lds[wid][q] = w[q];
    __syncthreads();
    if ((int)threadIdx.x < K) {
        double s = 0.;
        for (int wv = 0; wv < BS / 64; ++wv) s += lds[wv][threadIdx.x];
        buf[blockIdx.x * RED_STRIDE + threadIdx.x] = s;
    }
}

// every thread of the block receives the K totals of nb rows: the first wave adds
// rows lane, lane+64, ... in order and folds them by a butterfly; one barrier.  The rows
// of a lane are loaded in groups (four at K <= 4, two above) before any is added, so a
// grid of up to 256 blocks pays one or two memory latencies, not four; the addition order
// is the row order either way.
template <int BS, int K>
__device__ inline void readPartials(const double* buf, int nb, double (&out)[K]) {
    static_assert(K <= RED_STRIDE, "more partial sums than a row holds");
    constexpr int GR = K <= 4 ? 4 : 2;
    __shared__ double res[K];
    if (threadIdx.x < 64) {
        double a[K];
#pragma unroll
        for (int q = 0; q < K; ++q) a[q] = 0.;
        for (int b0 = threadIdx.x; b0 < nb; b0 += 64 * GR) {
            double v[GR][K];
#pragma unroll
            for (int r = 0; r < GR; ++r) {
                const int b = b0 + 64 * r;
#pragma unroll
                for (int q = 0; q < K; ++q) v[r][q] = b < nb ? buf[b * RED_STRIDE + q] : 0.;
            }
#pragma unroll
            for (int r = 0; r < GR; ++r)
#pragma unroll
                for (int q = 0; q < K; ++q)
                    if (b0 + 64 * r < nb) a[q] += v[r][q];
        }
#pragma unroll
        for (int o = 32; o > 0; o >>= 1)
#pragma unroll
            for (int q = 0; q < K; ++q) a[q] += __shfl_xor(a[q], o, 64);
        if (threadIdx.x == 0)
#pragma unroll
            for (int q = 0; q < K; ++q) res[q] = a[q];
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < K; ++q) out[q] = res[q];
}

// Block b of a launch whose grid is a multiple of 8 takes chunk (b % 8) * (nb / 8) + b / 8:
// blocks are dealt to the 8 XCDs round-robin, so each XCD then works on one contiguous
// eighth of the cells and the stencil neighbours it reads stay in its own L2.
__device__ inline int xcdChunk() {
    const int nb = gridDim.x, b = blockIdx.x;
    return (nb & 7) ? b : (b & 7) * (nb >> 3) + (b >> 3);
}
// interior cell of a one-thread-per-cell kernel (grid = ceil(N / blockDim)) in the
// XCD-aware chunk order: an XCD's blocks take one contiguous eighth of the cells, so the
// +-sJ and +-sK neighbours a cell reads are mostly in that XCD's L2
__device__ inline long xcdCell() { return (long)xcdChunk() * blockDim.x + threadIdx.x; }

// Column tiles: a block of CTX x CTY threads takes that many columns of (i, j) through a
// range of k planes, one thread per column, k ascending, so that the k - 1 and k + 1
// neighbours a cell reads were, or will next be, loaded by the same thread or block (the
// plane-chunk order fetches them again from the cache levels below L2, or from HBM on large
// grids).  Work item v (tile v / nseg, k segment v % nseg) goes to blocks in the
// XCD-aware order, so an XCD holds whole tiles.
constexpr int CTX = 64, CTY = 16;
struct ColMap { int ntx, nty, nseg, kper; };
inline ColMap colMap(const Dims& D, int g) {
    ColMap m;
    m.ntx = (D.nI + CTX - 1) / CTX;
    m.nty = (D.nJ + CTY - 1) / CTY;
    const int t = m.ntx * m.nty;
    m.nseg = g / t > 1 ? g / t : 1;
    m.kper = (D.nK + m.nseg - 1) / m.nseg;
    return m;
}
// work item v of the map for this thread: the padded index p of its first cell and its k
// range [k, ke); false when the thread has no cell in the item
__device__ inline bool colItem(const Dims& D, const ColMap& cm, int v, int& p, int& k, int& ke) {
    const int t = v / cm.nseg;
    const int i = (t % cm.ntx) * CTX + (int)(threadIdx.x % CTX), j = (t / cm.ntx) * CTY + (int)(threadIdx.x / CTX);
    k = (v % cm.nseg) * cm.kper;
    ke = min(D.nK, k + cm.kper);
    if (!(i < D.nI && j < D.nJ) || k >= ke) return false;
    p = D.pidx(i, j, k);
    return true;
}

}  // namespace caffe
