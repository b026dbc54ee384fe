// net.hip -- the multilevel convolution network of the learned pressure-correction
// preconditioner (no counterpart in the reference, whose p' solve is PETSc BiCGStab +
// SOR, SparseMatrix.cpp:86-106).
//
// Kernels
//   k_conv         3x3x3 convolution on MFMA (v_mfma_f32_16x16x16_bf16): BF16 operands,
//                  FP32 accumulation.  One workgroup computes a TXxTYxTZ output tile for
//                  16 output channels.  A single staging pass puts every input channel of
//                  the tile and its one-voxel halo in LDS together with the 27xCINx16
//                  weights (up to 147 KB of the 160 KB a gfx950 workgroup may hold), so
//                  the MFMA loop over taps and 16-channel input blocks runs without a
//                  barrier.  A wave owns TX*TY*TZ/64 groups of 16 voxels; per tap and input
//                  block it reads one B fragment (weights) and one A fragment per group.
//                  Smaller tiles on the coarse levels give those small grids more workgroups.
//                  The staging applies the layer's input operator on the fly: plain,
//                  2x2x2 average pooling of the finer level, or nearest up-sampling of the
//                  coarser level plus the skip features.  The same kernel computes the
//                  input gradient with flipped taps and swapped channels (FP32 in/out).
//   k_conv_wgrad   weight gradient on MFMA: D[co][ci] += G^T X over the voxels of a
//                  chunk of tiles for all 27 taps, reduced across waves in LDS in a fixed
//                  order, one partial row per workgroup, then k_reduce_parts.
//   k_inf16/k_outf16  the 1->16 input and 16->1 output layers on MFMA, their FP32 operands
//                  (input x, weights) split into a BF16 high part and remainder;
//                  k_out_ginput / k_edge_wgrad their training gradients in FP32 scalar code.
// Layouts: activations [z][y][x][C] channels last (BF16), gradients FP32 in the same
// layout, FP32 parameters [27][cout][cin] with tap t = (dz+1)*9 + (dy+1)*3 + (dx+1).
#include <cmath>
#include <cstring>

#include "net.hpp"
#include "netdev.hpp"

namespace caffe {

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ uint2 pack4(const float (&v)[4]) {
    uint2 r;
    r.x = f2bf(v[0]) | (f2bf(v[1]) << 16);
    r.y = f2bf(v[2]) | (f2bf(v[3]) << 16);
    return r;
}
__device__ __forceinline__ void ld4(const uint16_t* p, float (&v)[4]) {
    const uint2 u = *(const uint2*)p;
    v[0] = bf2f(u.x & 0xffffu); v[1] = bf2f(u.x >> 16); v[2] = bf2f(u.y & 0xffffu); v[3] = bf2f(u.y >> 16);
}
__device__ __forceinline__ void ld4(const float* p, float (&v)[4]) {
    const float4 f = *(const float4*)p;
    v[0] = f.x; v[1] = f.y; v[2] = f.z; v[3] = f.w;
}
__device__ __forceinline__ bool isDone(const double* done) { return done && *done != 0.; }
// The forward kernels always get a flag (net_forward passes a slot that stays zero when the
// caller has none), load it first and test it once their staging loads are in flight, so
// the flag's latency overlaps theirs.
__device__ __forceinline__ double flagOf(const double* done) { return *done; }

constexpr int NB = 256;   // threads per block of the network kernels
constexpr int LS = 20;    // LDS row of the weight-gradient kernel: 16 BF16 channels + 4 padding

// Channels [c, c+4) of the staged input at voxel (gx,gy,gz) of a level with dims
// (nx,ny,nz), C channels, as 4 packed BF16.
template <int MODE, class TIN>
__device__ __forceinline__ uint2 stage4(const TIN* in, const TIN* skip, int C, int nx, int ny, int nz, int gx, int gy,
                                        int gz, int c) {
    if (MODE == MODE_PLAIN) {
        const size_t o = ((size_t)(gz * ny + gy) * nx + gx) * C + c;
        if constexpr (sizeof(TIN) == 2) {
            return *(const uint2*)(in + o);
        } else {
            float v[4];
            ld4(in + o, v);
            return pack4(v);
        }
    } else if (MODE == MODE_POOL) {  // the finer level has dims (2nx, 2ny, 2nz)
        float s[4] = {0.f, 0.f, 0.f, 0.f};
        const int fx = 2 * nx, fy = 2 * ny;
        for (int dz = 0; dz < 2; ++dz)
            for (int dy = 0; dy < 2; ++dy)
                for (int dx = 0; dx < 2; ++dx) {
                    float v[4];
                    ld4(in + ((size_t)((2 * gz + dz) * fy + 2 * gy + dy) * fx + 2 * gx + dx) * C + c, v);
                    for (int q = 0; q < 4; ++q) s[q] += v[q];
                }
        for (int q = 0; q < 4; ++q) s[q] *= 0.125f;
        return pack4(s);
    } else {  // UPADD: coarser level has dims (nx/2, ny/2, nz/2)
        float a[4], b[4];
        ld4(in + ((size_t)((gz >> 1) * (ny >> 1) + (gy >> 1)) * (nx >> 1) + (gx >> 1)) * C + c, a);
        ld4(skip + ((size_t)(gz * ny + gy) * nx + gx) * C + c, b);
        for (int q = 0; q < 4; ++q) a[q] += b[q];
        return pack4(a);
    }
}

// Channels [c, c+8) of the staged input, as 8 packed BF16 (16-byte loads).
template <int MODE, class TIN>
__device__ __forceinline__ uint4 stage8(const TIN* in, const TIN* skip, int C, int nx, int ny, int nz, int gx, int gy,
                                        int gz, int c) {
    const uint2 lo = stage4<MODE, TIN>(in, skip, C, nx, ny, nz, gx, gy, gz, c);
    const uint2 hi = stage4<MODE, TIN>(in, skip, C, nx, ny, nz, gx, gy, gz, c + 4);
    return make_uint4(lo.x, lo.y, hi.x, hi.y);
}

struct ConvArgs {
    const void* in;
    const void* skip;
    const uint16_t* w;     // packed [cout/16][cin/16][27][16 co][16 ci]
    void* out;
    int nx, ny, nz, cin, cout;
    int accum;             // FP32 output: add to it
    const double* done;
};

template <int TX, int TY, int TZ, int CIN, int MODE, class TIN, class TOUT>
__global__ __launch_bounds__(NB) void k_conv(ConvArgs a) {
    if (isDone(a.done)) return;
    constexpr int HX = TX + 2, HY = TY + 2, HZ = TZ + 2, HV = HX * HY * HZ;
    constexpr int TV = TX * TY * TZ, G = TV / 64;
    constexpr int NCB = CIN / 16, LSI = CIN + 4, C4 = CIN / 4;  // LDS row: CIN channels + 4 padding
    static_assert(G >= 1 && TV % 64 == 0, "tile must hold a multiple of 64 voxels");
    __shared__ __align__(16) uint16_t sIn[HV * LSI];
    __shared__ __align__(16) uint16_t sW[27 * 16 * LSI];   // [t][co][ci] of this block of 16 outputs
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int n = lane & 15, kq = lane >> 4;
    const int ntx = (a.nx + TX - 1) / TX, nty = (a.ny + TY - 1) / TY;
    int b = blockIdx.x;
    const int bx = b % ntx;
    b /= ntx;
    const int by = b % nty, bz = b / nty;
    const int ox = bx * TX, oy = by * TY, oz = bz * TZ;
    const int cob = blockIdx.y;
    const TIN* in = (const TIN*)a.in;
    const TIN* sk = (const TIN*)a.skip;

    // one staging pass: every input channel of the halo tile, and the weights.  All
    // global loads of a thread are issued before its LDS stores, so their latencies
    // overlap instead of adding up.
    constexpr int C8 = CIN / 8, NI = HV * C8, NIT = (NI + NB - 1) / NB;
    constexpr int NWI = NCB * 27 * 16 * 2, NWT = (NWI + NB - 1) / NB;
    const uint16_t* ws = a.w + (size_t)cob * NCB * (27 * 256);
    uint4 vin[NIT], vw[NWT];
#pragma unroll
    for (int u = 0; u < NIT; ++u) {
        const int it = tid + u * NB;
        vin[u] = make_uint4(0u, 0u, 0u, 0u);
        if (it < NI) {
            const int hv = it / C8, c8 = it - hv * C8;
            const int hx = hv % HX, hy = (hv / HX) % HY, hz = hv / (HX * HY);
            const int gx = ox + hx - 1, gy = oy + hy - 1, gz = oz + hz - 1;
            if (gx >= 0 && gx < a.nx && gy >= 0 && gy < a.ny && gz >= 0 && gz < a.nz)
                vin[u] = stage8<MODE, TIN>(in, sk, CIN, a.nx, a.ny, a.nz, gx, gy, gz, c8 * 8);
        }
    }
#pragma unroll
    for (int u = 0; u < NWT; ++u) {
        const int it = tid + u * NB;
        if (it < NWI) vw[u] = *(const uint4*)(ws + (size_t)it * 8);   // packed [cb][t][co][ci]
    }
#pragma unroll
    for (int u = 0; u < NIT; ++u) {
        const int it = tid + u * NB;
        if (it < NI) {
            const int hv = it / C8, c8 = it - hv * C8;
            uint16_t* d = &sIn[hv * LSI + c8 * 8];
            *(uint2*)d = make_uint2(vin[u].x, vin[u].y);
            *(uint2*)(d + 4) = make_uint2(vin[u].z, vin[u].w);
        }
    }
#pragma unroll
    for (int u = 0; u < NWT; ++u) {
        const int it = tid + u * NB;
        if (it < NWI) {
            const int h = it & 1;
            int r = it >> 1;
            const int co = r % 16;
            r /= 16;
            const int t = r % 27, cb = r / 27;
            uint16_t* d = &sW[(t * 16 + co) * LSI + cb * 16 + h * 8];
            *(uint2*)d = make_uint2(vw[u].x, vw[u].y);
            *(uint2*)(d + 4) = make_uint2(vw[u].z, vw[u].w);
        }
    }
    __syncthreads();

    int base[G];
#pragma unroll
    for (int g = 0; g < G; ++g) {
        const int vi = (wave * G + g) * 16 + n;
        const int tx = vi % TX, ty = (vi / TX) % TY, tz = vi / (TX * TY);
        base[g] = ((tz + 1) * HY + (ty + 1)) * HX + (tx + 1);
    }
    f32x4 acc[G];
#pragma unroll
    for (int g = 0; g < G; ++g) acc[g] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll 3
    for (int t = 0; t < 27; ++t) {
        const int dz = t / 9 - 1, dy = (t / 3) % 3 - 1, dx = t % 3 - 1;
        const int toff = (dz * HY + dy) * HX + dx;
#pragma unroll
        for (int cb = 0; cb < NCB; ++cb) {
            const bf16x4 bw = *(const bf16x4*)&sW[(t * 16 + n) * LSI + cb * 16 + kq * 4];
#pragma unroll
            for (int g = 0; g < G; ++g) {
                const bf16x4 av = *(const bf16x4*)&sIn[(base[g] + toff) * LSI + cb * 16 + kq * 4];
                acc[g] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(av, bw, acc[g], 0, 0, 0);
            }
        }
    }
#pragma unroll
    for (int g = 0; g < G; ++g) {
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int vi = (wave * G + g) * 16 + kq * 4 + q;
            const int tx = vi % TX, ty = (vi / TX) % TY, tz = vi / (TX * TY);
            const int gx = ox + tx, gy = oy + ty, gz = oz + tz;
            if (gx < a.nx && gy < a.ny && gz < a.nz) {
                const size_t o = ((size_t)(gz * a.ny + gy) * a.nx + gx) * a.cout + cob * 16 + n;
                if constexpr (sizeof(TOUT) == 2) {
                    ((uint16_t*)a.out)[o] = (uint16_t)f2bf(acc[g][q]);
                } else {
                    float* o32 = (float*)a.out;
                    o32[o] = a.accum ? o32[o] + acc[g][q] : acc[g][q];
                }
            }
        }
    }
}

// ---------------------------------------------------------------- forward convolution
// k_convf: the forward pass of the ten MFMA layers (BF16 in, BF16 out), on
// v_mfma_f32_16x16x32_bf16.  The reduction index runs over (tap, input channel) in steps
// of 32: two taps per step at 16 input channels (a zero 28th tap completes the last
// pair), one tap per step at 32, half a tap at 64.  LDS rows hold a voxel's channels (or
// a weight row's input channels) without padding; their 16-byte chunks are XOR-swizzled
// by row, so the eight lanes that read one chunk index of eight consecutive rows hit
// eight distinct bank groups.  Without padding four level-0 workgroups fit in a CU and a
// 64^3 layer runs in one round.  The output tile goes through LDS for 16-byte stores;
// with POOLOUT the same pass writes the 2x2x2 average of the tile's BF16 outputs (summed
// in dz, dy, dx order and scaled by 1/8, the order the POOL staging of the backward
// pass uses), which the next level's first layer reads as its plain input.
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));

struct ConvFArgs {
    const uint16_t* in;     // PLAIN: input at this level; UPADD: the coarser level's output
    const uint16_t* skip;   // UPADD: skip features at this level
    const uint16_t* w;      // packed [cout/16][cin/16][27][16 co][16 ci]
    uint16_t* out;          // [v][cout]
    uint16_t* pooled;       // POOLOUT: [v/8][cout] at the next level
    int nx, ny, nz, cout;
    const double* done;
};

template <int C8>
__device__ __forceinline__ int swz(int r, int c) {
    // 64 banks of 4 bytes: sixteen consecutive rows read at one chunk index must cover
    // 256 distinct bytes
    constexpr int S = C8 == 2 ? 3 : (C8 == 4 ? 2 : 1);
    return c ^ ((r >> S) & (C8 - 1));
}
__device__ __forceinline__ uint4 addbf8(uint4 x, uint4 y) {
    const uint32_t xs[4] = {x.x, x.y, x.z, x.w}, ys[4] = {y.x, y.y, y.z, y.w};
    uint32_t r[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const float lo = bf2f(xs[i] & 0xffffu) + bf2f(ys[i] & 0xffffu);
        const float hi = bf2f(xs[i] >> 16) + bf2f(ys[i] >> 16);
        r[i] = f2bf(lo) | (f2bf(hi) << 16);
    }
    return make_uint4(r[0], r[1], r[2], r[3]);
}
__device__ __forceinline__ int tapOff(int t, int HX, int HY) {
    if (t >= 27) return 0;   // the zero tap: any row in bounds
    const int dz = t / 9 - 1, dy = (t / 3) % 3 - 1, dx = t % 3 - 1;
    return (dz * HY + dy) * HX + dx;
}

template <int TX, int TY, int TZ, int CIN, int MODE, bool POOLOUT>
__global__ __launch_bounds__(NB) void k_convf(ConvFArgs a) {
    const double flag = flagOf(a.done);
    constexpr int HX = TX + 2, HY = TY + 2, HZ = TZ + 2, HV = HX * HY * HZ;
    constexpr int TV = TX * TY * TZ, G = TV / 64;
    constexpr int C8 = CIN / 8;                               // 16-byte chunks per row
    constexpr int NT = CIN == 16 ? 28 : 27;                   // taps held in LDS
    constexpr int NK = CIN == 16 ? 14 : 27 * (CIN / 32);      // K = 32 steps
    static_assert(G >= 1 && TV % 64 == 0, "tile must hold a multiple of 64 voxels");
    static_assert(TV * 16 <= HV * CIN, "output tile must fit in the input tile's LDS");
    __shared__ __align__(16) uint16_t sIn[HV * CIN];
    __shared__ __align__(16) uint16_t sW[NT * 16 * CIN];    // [t][co][ci] of this block's 16 outputs
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int n = lane & 15, kq = lane >> 4;
    const int ntx = (a.nx + TX - 1) / TX, nty = (a.ny + TY - 1) / TY;
    int b = blockIdx.x;
    const int bx = b % ntx;
    b /= ntx;
    const int by = b % nty, bz = b / nty;
    const int ox = bx * TX, oy = by * TY, oz = bz * TZ;
    const int cob = blockIdx.y;

    // staging: all global loads of a thread are issued before its LDS stores
    constexpr int NI = HV * C8, NIT = (NI + NB - 1) / NB;
    constexpr int NWI = (CIN / 16) * 27 * 16 * 2, NWT = (NWI + NB - 1) / NB;
    const uint16_t* ws = a.w + (size_t)cob * (CIN / 16) * (27 * 256);
    uint4 vin[NIT], vsk[MODE == MODE_UPADD ? NIT : 1], vw[NWT];
#pragma unroll
    for (int u = 0; u < NIT; ++u) {
        const int it = tid + u * NB;
        vin[u] = make_uint4(0u, 0u, 0u, 0u);
        if (MODE == MODE_UPADD) vsk[u] = make_uint4(0u, 0u, 0u, 0u);
        if (it < NI) {
            const int hv = it / C8, c = it - hv * C8;
            const int hx = hv % HX, hy = (hv / HX) % HY, hz = hv / (HX * HY);
            const int gx = ox + hx - 1, gy = oy + hy - 1, gz = oz + hz - 1;
            if (gx >= 0 && gx < a.nx && gy >= 0 && gy < a.ny && gz >= 0 && gz < a.nz) {
                if (MODE == MODE_PLAIN) {
                    vin[u] = *(const uint4*)(a.in + ((size_t)(gz * a.ny + gy) * a.nx + gx) * CIN + 8 * c);
                } else {
                    const int cx = a.nx >> 1, cy = a.ny >> 1;
                    vin[u] = *(const uint4*)(a.in + ((size_t)((gz >> 1) * cy + (gy >> 1)) * cx + (gx >> 1)) * CIN + 8 * c);
                    vsk[u] = *(const uint4*)(a.skip + ((size_t)(gz * a.ny + gy) * a.nx + gx) * CIN + 8 * c);
                }
            }
        }
    }
#pragma unroll
    for (int u = 0; u < NWT; ++u) {
        const int it = tid + u * NB;
        vw[u] = make_uint4(0u, 0u, 0u, 0u);
        if (it < NWI) vw[u] = *(const uint4*)(ws + (size_t)it * 8);   // packed [cb][t][co][ci]
    }
    if (flag != 0.) return;
#pragma unroll
    for (int u = 0; u < NIT; ++u) {
        const int it = tid + u * NB;
        if (it < NI) {
            const int hv = it / C8, c = it - hv * C8;
            const uint4 v = MODE == MODE_UPADD ? addbf8(vin[u], vsk[u]) : vin[u];
            *(uint4*)&sIn[hv * CIN + 8 * swz<C8>(hv, c)] = v;
        }
    }
#pragma unroll
    for (int u = 0; u < NWT; ++u) {
        const int it = tid + u * NB;
        if (it < NWI) {
            const int h = it & 1;
            int r = it >> 1;
            const int co = r % 16;
            r /= 16;
            const int t = r % 27, cb = r / 27;
            const int row = t * 16 + co;
            *(uint4*)&sW[row * CIN + 8 * swz<C8>(row, cb * 2 + h)] = vw[u];
        }
    }
    if (CIN == 16 && tid < 32) {   // the zero 28th tap
        const int row = 27 * 16 + (tid >> 1);
        *(uint4*)&sW[row * CIN + 8 * (tid & 1)] = make_uint4(0u, 0u, 0u, 0u);
    }
    __syncthreads();

    int base[G];
#pragma unroll
    for (int g = 0; g < G; ++g) {
        const int vi = (wave * G + g) * 16 + n;
        const int tx = vi % TX, ty = (vi / TX) % TY, tz = vi / (TX * TY);
        base[g] = ((tz + 1) * HY + (ty + 1)) * HX + (tx + 1);
    }
    f32x4 acc[G];
#pragma unroll
    for (int g = 0; g < G; ++g) acc[g] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll 2
    for (int kc = 0; kc < NK; ++kc) {
        int t, c;
        if (CIN == 16) { t = 2 * kc + (kq >> 1); c = kq & 1; }
        else if (CIN == 32) { t = kc; c = kq; }
        else { t = kc >> 1; c = (kc & 1) * 4 + kq; }
        const int wrow = t * 16 + n;
        const bf16x8 bw = *(const bf16x8*)&sW[wrow * CIN + 8 * swz<C8>(wrow, c)];
        const int toff = tapOff(t, HX, HY);
#pragma unroll
        for (int g = 0; g < G; ++g) {
            const int row = base[g] + toff;
            const bf16x8 av = *(const bf16x8*)&sIn[row * CIN + 8 * swz<C8>(row, c)];
            acc[g] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av, bw, acc[g], 0, 0, 0);
        }
    }

    // epilogue: the BF16 tile in LDS ([TV][16], two swizzled chunks per voxel)
    __syncthreads();
    uint16_t* sOut = sIn;
#pragma unroll
    for (int g = 0; g < G; ++g)
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int vi = (wave * G + g) * 16 + kq * 4 + q;
            sOut[vi * 16 + 8 * swz<2>(vi, n >> 3) + (n & 7)] = (uint16_t)f2bf(acc[g][q]);
        }
    __syncthreads();
    for (int it = tid; it < TV * 2; it += NB) {
        const int vi = it >> 1, h = it & 1;
        const int tx = vi % TX, ty = (vi / TX) % TY, tz = vi / (TX * TY);
        const int gx = ox + tx, gy = oy + ty, gz = oz + tz;
        if (gx < a.nx && gy < a.ny && gz < a.nz)
            *(uint4*)(a.out + ((size_t)(gz * a.ny + gy) * a.nx + gx) * a.cout + cob * 16 + 8 * h) =
                *(const uint4*)&sOut[vi * 16 + 8 * swz<2>(vi, h)];
    }
    if constexpr (POOLOUT) {
        constexpr int PX = TX / 2, PY = TY / 2, PZ = TZ / 2, PV = PX * PY * PZ;
        const int px0 = ox >> 1, py0 = oy >> 1, pz0 = oz >> 1;
        const int pnx = a.nx >> 1, pny = a.ny >> 1, pnz = a.nz >> 1;
        for (int it = tid; it < PV * 16; it += NB) {
            const int pv = it >> 4, ch = it & 15;
            const int px = pv % PX, py = (pv / PX) % PY, pz = pv / (PX * PY);
            float s = 0.f;
#pragma unroll
            for (int dz = 0; dz < 2; ++dz)
#pragma unroll
                for (int dy = 0; dy < 2; ++dy)
#pragma unroll
                    for (int dx = 0; dx < 2; ++dx) {
                        const int vi = ((2 * pz + dz) * TY + 2 * py + dy) * TX + 2 * px + dx;
                        s += bf2f(sOut[vi * 16 + 8 * swz<2>(vi, ch >> 3) + (ch & 7)]);
                    }
            s *= 0.125f;
            const int gx = px0 + px, gy = py0 + py, gz = pz0 + pz;
            if (gx < pnx && gy < pny && gz < pnz)
                a.pooled[((size_t)(gz * pny + gy) * pnx + gx) * a.cout + cob * 16 + ch] = (uint16_t)f2bf(s);
        }
    }
}

// k_convf16: the 16-input-channel layers on 16-voxel x rows (tile 16x4x4), on
// v_mfma_f32_16x16x16_bf16 with one tap per MFMA.  The 27 weight fragments of the
// block's 16 outputs stay in registers for the whole tile (54 VGPRs), so the loop reads
// only activations from LDS: per (dz, dy) and 16-voxel row one 8-byte fragment (lane n
// holds voxel n), and the dx = -1 and dx = +1 fragments are that row shifted by one lane
// (DPP row_shr:1 / row_shl:1 within the 16-lane rows), with the two halo voxels from a
// second read of the row.  This reads two thirds of the LDS bytes of one read per tap.
__device__ __forceinline__ uint32_t shr1(uint32_t old, uint32_t v) {   // lane n <- lane n-1; lane 0 keeps old
    return (uint32_t)__builtin_amdgcn_update_dpp((int)old, (int)v, 0x111, 0xf, 0xf, false);
}
__device__ __forceinline__ uint32_t shl1(uint32_t old, uint32_t v) {   // lane n <- lane n+1; lane 15 keeps old
    return (uint32_t)__builtin_amdgcn_update_dpp((int)old, (int)v, 0x101, 0xf, 0xf, false);
}
__device__ __forceinline__ bf16x4 asbf4(uint32_t x, uint32_t y) {
    const uint2 u = make_uint2(x, y);
    return *(const bf16x4*)&u;
}

template <int MODE, bool POOLOUT>
__global__ __launch_bounds__(NB) void k_convf16(ConvFArgs a) {
    const double flag = flagOf(a.done);
    constexpr int TX = 16, TY = 4, TZ = 4, CIN = 16, C8 = 2;
    constexpr int HX = TX + 2, HY = TY + 2, HZ = TZ + 2, HV = HX * HY * HZ;
    constexpr int TV = TX * TY * TZ, G = TV / 64;
    static_assert(TV * 16 <= HV * CIN, "output tile must fit in the input tile's LDS");
    __shared__ __align__(16) uint16_t sIn[HV * CIN];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int n = lane & 15, kq = lane >> 4;
    const int ntx = (a.nx + TX - 1) / TX, nty = (a.ny + TY - 1) / TY;
    int b = blockIdx.x;
    const int bx = b % ntx;
    b /= ntx;
    const int by = b % nty, bz = b / nty;
    const int ox = bx * TX, oy = by * TY, oz = bz * TZ;
    const int cob = blockIdx.y;

    constexpr int NI = HV * C8, NIT = (NI + NB - 1) / NB;
    uint4 vin[NIT], vsk[MODE == MODE_UPADD ? NIT : 1];
#pragma unroll
    for (int u = 0; u < NIT; ++u) {
        const int it = tid + u * NB;
        vin[u] = make_uint4(0u, 0u, 0u, 0u);
        if (MODE == MODE_UPADD) vsk[u] = make_uint4(0u, 0u, 0u, 0u);
        if (it < NI) {
            const int hv = it / C8, c = it - hv * C8;
            const int hx = hv % HX, hy = (hv / HX) % HY, hz = hv / (HX * HY);
            const int gx = ox + hx - 1, gy = oy + hy - 1, gz = oz + hz - 1;
            if (gx >= 0 && gx < a.nx && gy >= 0 && gy < a.ny && gz >= 0 && gz < a.nz) {
                if (MODE == MODE_PLAIN) {
                    vin[u] = *(const uint4*)(a.in + ((size_t)(gz * a.ny + gy) * a.nx + gx) * CIN + 8 * c);
                } else {
                    const int cx = a.nx >> 1, cy = a.ny >> 1;
                    vin[u] = *(const uint4*)(a.in + ((size_t)((gz >> 1) * cy + (gy >> 1)) * cx + (gx >> 1)) * CIN + 8 * c);
                    vsk[u] = *(const uint4*)(a.skip + ((size_t)(gz * a.ny + gy) * a.nx + gx) * CIN + 8 * c);
                }
            }
        }
    }
    if (flag != 0.) return;
#pragma unroll
    for (int u = 0; u < NIT; ++u) {
        const int it = tid + u * NB;
        if (it < NI) {
            const int hv = it / C8, c = it - hv * C8;
            const uint4 v = MODE == MODE_UPADD ? addbf8(vin[u], vsk[u]) : vin[u];
            *(uint4*)&sIn[hv * CIN + 8 * swz<C8>(hv, c)] = v;
        }
    }
    // weights (after the staging, so that their registers and the staging's do not
    // overlap): lane (n, kq) holds w[t][ci = 4kq..4kq+3][co = n] for all 27 taps
    const uint16_t* ws = a.w + (size_t)cob * (27 * 256) + n * 16 + 4 * kq;
    uint2 wr[27];
#pragma unroll
    for (int t = 0; t < 27; ++t) wr[t] = *(const uint2*)(ws + t * 256);
    __syncthreads();

    // lane (n, kq): channels 4kq..4kq+3 (8 bytes) of an LDS row
    auto frag = [&](int row) { return *(const uint2*)&sIn[row * CIN + 8 * swz<C8>(row, kq >> 1) + 4 * (kq & 1)]; };
    // the second read of a row: lanes 0-7 read voxel n-1, lanes 8-15 voxel n+1 (always
    // inside the halo row), so that lane 0 holds voxel -1 and lane 15 voxel 16; reading
    // it in every lane keeps the loop free of branches
    const int hoff = n < 8 ? -1 : 1;
    int base[G];
#pragma unroll
    for (int g = 0; g < G; ++g) {
        const int vi = (wave * G + g) * 16 + n;   // a group is one x row: tx = n
        const int ty = (vi / TX) % TY, tz = vi / (TX * TY);
        base[g] = ((tz + 1) * HY + (ty + 1)) * HX + (n + 1);
    }
    f32x4 acc[G];
#pragma unroll
    for (int g = 0; g < G; ++g) acc[g] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int p = 0; p < 9; ++p) {
        const int dz = p / 3 - 1, dy = p % 3 - 1, t0 = 3 * p;   // taps t0 (dx -1), t0+1, t0+2
        const int poff = (dz * HY + dy) * HX;
#pragma unroll
        for (int g = 0; g < G; ++g) {
            const int row = base[g] + poff;
            const uint2 c0 = frag(row);
            const uint2 h = frag(row + hoff);
            const bf16x4 am = asbf4(shr1(h.x, c0.x), shr1(h.y, c0.y));
            const bf16x4 ap = asbf4(shl1(h.x, c0.x), shl1(h.y, c0.y));
            acc[g] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(am, asbf4(wr[t0].x, wr[t0].y), acc[g], 0, 0, 0);
            acc[g] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(asbf4(c0.x, c0.y), asbf4(wr[t0 + 1].x, wr[t0 + 1].y), acc[g], 0, 0, 0);
            acc[g] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(ap, asbf4(wr[t0 + 2].x, wr[t0 + 2].y), acc[g], 0, 0, 0);
        }
    }

    __syncthreads();
    uint16_t* sOut = sIn;
#pragma unroll
    for (int g = 0; g < G; ++g)
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int vi = (wave * G + g) * 16 + kq * 4 + q;
            sOut[vi * 16 + 8 * swz<2>(vi, n >> 3) + (n & 7)] = (uint16_t)f2bf(acc[g][q]);
        }
    __syncthreads();
    for (int it = tid; it < TV * 2; it += NB) {
        const int vi = it >> 1, hh = it & 1;
        const int tx = vi % TX, ty = (vi / TX) % TY, tz = vi / (TX * TY);
        const int gx = ox + tx, gy = oy + ty, gz = oz + tz;
        if (gx < a.nx && gy < a.ny && gz < a.nz)
            *(uint4*)(a.out + ((size_t)(gz * a.ny + gy) * a.nx + gx) * a.cout + cob * 16 + 8 * hh) =
                *(const uint4*)&sOut[vi * 16 + 8 * swz<2>(vi, hh)];
    }
    if constexpr (POOLOUT) {
        constexpr int PX = TX / 2, PY = TY / 2, PZ = TZ / 2, PV = PX * PY * PZ;
        const int px0 = ox >> 1, py0 = oy >> 1, pz0 = oz >> 1;
        const int pnx = a.nx >> 1, pny = a.ny >> 1, pnz = a.nz >> 1;
        for (int it = tid; it < PV * 16; it += NB) {
            const int pv = it >> 4, ch = it & 15;
            const int px = pv % PX, py = (pv / PX) % PY, pz = pv / (PX * PY);
            float sum = 0.f;
#pragma unroll
            for (int dz = 0; dz < 2; ++dz)
#pragma unroll
                for (int dy = 0; dy < 2; ++dy)
#pragma unroll
                    for (int dx = 0; dx < 2; ++dx) {
                        const int vi = ((2 * pz + dz) * TY + 2 * py + dy) * TX + 2 * px + dx;
                        sum += bf2f(sOut[vi * 16 + 8 * swz<2>(vi, ch >> 3) + (ch & 7)]);
                    }
            sum *= 0.125f;
            const int gx = px0 + px, gy = py0 + py, gz = pz0 + pz;
            if (gx < pnx && gy < pny && gz < pnz)
                a.pooled[((size_t)(gz * pny + gy) * pnx + gx) * a.cout + cob * 16 + ch] = (uint16_t)f2bf(sum);
        }
    }
}

struct WgArgs {
    const uint16_t* x;     // source activation (BF16) of the staged input
    const uint16_t* skip;
    const float* g;        // gradient of the layer output, FP32 [v][cout]
    float* part;           // [chunk][27][cout][cin]
    int nx, ny, nz, cin, cout;
    int tilesPerChunk, ntiles;
};

template <int TX, int TY, int TZ, int MODE>
__global__ __launch_bounds__(NB) void k_conv_wgrad(WgArgs a) {
    constexpr int HX = TX + 2, HY = TY + 2, HZ = TZ + 2, HV = HX * HY * HZ;
    constexpr int TV = TX * TY * TZ, G = TV / 64, GS = TV + 4;
    __shared__ __align__(16) uint16_t sX[HV * LS];
    __shared__ __align__(16) uint16_t sG[16 * GS];
    __shared__ __align__(16) float sR[27 * 256];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int n = lane & 15, kq = lane >> 4;
    const int ntx = (a.nx + TX - 1) / TX, nty = (a.ny + TY - 1) / TY;
    const int ncib = a.cin >> 4;
    const int cob = blockIdx.y / ncib, cib = blockIdx.y % ncib;

    int hb[G][4];
#pragma unroll
    for (int g = 0; g < G; ++g)
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int vi = (wave * G + g) * 16 + kq * 4 + q;
            const int tx = vi % TX, ty = (vi / TX) % TY, tz = vi / (TX * TY);
            hb[g][q] = ((tz + 1) * HY + (ty + 1)) * HX + (tx + 1);
        }
    f32x4 acc[27];
#pragma unroll
    for (int t = 0; t < 27; ++t) acc[t] = f32x4{0.f, 0.f, 0.f, 0.f};

    const int t0 = blockIdx.x * a.tilesPerChunk;
    const int t1 = min(a.ntiles, t0 + a.tilesPerChunk);
    for (int tile = t0; tile < t1; ++tile) {
        int b = tile;
        const int bx = b % ntx;
        b /= ntx;
        const int by = b % nty, bz = b / nty;
        const int ox = bx * TX, oy = by * TY, oz = bz * TZ;
        __syncthreads();
        for (int it = tid; it < HV * 4; it += NB) {
            const int hv = it >> 2, c4 = it & 3;
            const int hx = hv % HX, hy = (hv / HX) % HY, hz = hv / (HX * HY);
            const int gx = ox + hx - 1, gy = oy + hy - 1, gz = oz + hz - 1;
            uint2 v = make_uint2(0u, 0u);
            if (gx >= 0 && gx < a.nx && gy >= 0 && gy < a.ny && gz >= 0 && gz < a.nz)
                v = stage4<MODE, uint16_t>(a.x, a.skip, a.cin, a.nx, a.ny, a.nz, gx, gy, gz, cib * 16 + c4 * 4);
            *(uint2*)&sX[hv * LS + c4 * 4] = v;
        }
        for (int it = tid; it < TV * 4; it += NB) {
            const int vi = it >> 2, c4 = it & 3;
            const int tx = vi % TX, ty = (vi / TX) % TY, tz = vi / (TX * TY);
            const int gx = ox + tx, gy = oy + ty, gz = oz + tz;
            float v[4] = {0.f, 0.f, 0.f, 0.f};
            if (gx < a.nx && gy < a.ny && gz < a.nz)
                ld4(a.g + ((size_t)(gz * a.ny + gy) * a.nx + gx) * a.cout + cob * 16 + c4 * 4, v);
#pragma unroll
            for (int q = 0; q < 4; ++q) sG[(c4 * 4 + q) * GS + vi] = (uint16_t)f2bf(v[q]);
        }
        __syncthreads();
#pragma unroll
        for (int g = 0; g < G; ++g) {
            const bf16x4 av = *(const bf16x4*)&sG[n * GS + (wave * G + g) * 16 + kq * 4];
#pragma unroll
            for (int t = 0; t < 27; ++t) {
                const int dz = t / 9 - 1, dy = (t / 3) % 3 - 1, dx = t % 3 - 1;
                const int toff = (dz * HY + dy) * HX + dx;
                bf16x4 bx4;
#pragma unroll
                for (int q = 0; q < 4; ++q) bx4[q] = (short)sX[(hb[g][q] + toff) * LS + n];
                acc[t] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(av, bx4, acc[t], 0, 0, 0);
            }
        }
    }
    // fixed-order reduction over the four waves
    for (int w = 1; w < NB / 64; ++w) {
        __syncthreads();
        if (wave == w)
            for (int t = 0; t < 27; ++t)
                for (int q = 0; q < 4; ++q) sR[(t * 4 + q) * 64 + lane] = acc[t][q];
        __syncthreads();
        if (wave == 0)
            for (int t = 0; t < 27; ++t)
                for (int q = 0; q < 4; ++q) acc[t][q] += sR[(t * 4 + q) * 64 + lane];
    }
    if (wave == 0) {
        float* dst = a.part + (size_t)blockIdx.x * 27 * a.cout * a.cin;
        for (int t = 0; t < 27; ++t)
            for (int q = 0; q < 4; ++q) {
                const int co = cob * 16 + kq * 4 + q, ci = cib * 16 + n;
                dst[((size_t)t * a.cout + co) * a.cin + ci] = acc[t][q];
            }
    }
}

// dst[i] (+)= sum over chunks of part[chunk][i], chunks in order
__global__ void k_reduce_parts(const float* part, int nChunk, int L, float* dst, int accum) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= L) return;
    float s = 0.f;
    for (int c = 0; c < nChunk; ++c) s += part[(size_t)c * L + i];
    dst[i] = accum ? dst[i] + s : s;
}

// ---------------------------------------------------------------- edge layers (FP32)
__device__ __forceinline__ void vox(int v, int nx, int ny, int& x, int& y, int& z) {
    x = v % nx;
    const int r = v / nx;
    y = r % ny;
    z = r / ny;
}

// Neighbour (i+dx, j+dy, k+dz) of a voxel, clamped into the grid, and whether it lies
// inside: loads through a clamped index and a mask have no branch, so the compiler issues
// all 27 of them before the first use.
__device__ __forceinline__ int nbr(int i, int j, int k, int t, int nx, int ny, int nz, bool& in) {
    const int a = i + t % 3 - 1, b = j + (t / 3) % 3 - 1, c = k + t / 9 - 1;
    in = a >= 0 && a < nx && b >= 0 && b < ny && c >= 0 && c < nz;
    const int ac = min(max(a, 0), nx - 1), bc = min(max(b, 0), ny - 1), cc = min(max(c, 0), nz - 1);
    return (cc * ny + bc) * nx + ac;
}

// in: 1 -> 16 channels on MFMA (v_mfma_f32_16x16x32_bf16), tile 16x4x4 as k_outf16.  The
// reduction index is the tap: lane (n, kq) holds the FP32 input at taps 8kq .. 8kq+7 of
// voxel n of a 16-voxel x row (taps 27-31 zero) split into a BF16 high part and remainder,
// x = xh + xl, and the weights W[t][co = n][0] likewise, w = wh + wl; three products
// xh wh + xh wl + xl wh (each exact in FP32, accumulated in FP32) carry x w to 2^-16
// relative, well inside the BF16 rounding of the layer's output.  The output tile goes
// through LDS and leaves in 16-byte stores.
__global__ __launch_bounds__(NB) void k_inf16(const float* x, const float* W, uint16_t* out, int nx, int ny, int nz,
                                              const double* done) {
    const double flag = flagOf(done);
    constexpr int TX = 16, TY = 4, TZ = 4;
    constexpr int HX = TX + 2, HY = TY + 2, HZ = TZ + 2, HV = HX * HY * HZ;
    constexpr int TV = TX * TY * TZ, G = TV / 64;
    __shared__ float sx[HV];
    __shared__ __align__(16) uint16_t sOut[TV * 16];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int n = lane & 15, kq = lane >> 4;
    const int ntx = (nx + TX - 1) / TX, nty = (ny + TY - 1) / TY;
    int b = blockIdx.x;
    const int bx = b % ntx;
    b /= ntx;
    const int by = b % nty, bz = b / nty;
    const int ox = bx * TX, oy = by * TY, oz = bz * TZ;
    constexpr int NIT = (HV + NB - 1) / NB;
    float tv[NIT];
#pragma unroll
    for (int u = 0; u < NIT; ++u) {
        const int hv = tid + u * NB;
        tv[u] = 0.f;
        if (hv < HV) {
            const int hx = hv % HX, hy = (hv / HX) % HY, hz = hv / (HX * HY);
            const int gx = ox + hx - 1, gy = oy + hy - 1, gz = oz + hz - 1;
            if (gx >= 0 && gx < nx && gy >= 0 && gy < ny && gz >= 0 && gz < nz) tv[u] = x[(gz * ny + gy) * nx + gx];
        }
    }
    // weights of taps 8kq .. 8kq+7 for output channel n, split high / remainder
    bf16x8 wh, wl;
#pragma unroll
    for (int j = 0; j < 8; ++j) {   // v_cvt_pk_bf16_f32: round to nearest even
        const int t = 8 * kq + j;
        const float w = t < 27 ? W[t * 16 + n] : 0.f;
        wh[j] = (__bf16)w;
        wl[j] = (__bf16)(w - (float)wh[j]);
    }
    if (flag != 0.) return;
#pragma unroll
    for (int u = 0; u < NIT; ++u)
        if (tid + u * NB < HV) sx[tid + u * NB] = tv[u];
    __syncthreads();
    int toff[8];
#pragma unroll
    for (int j = 0; j < 8; ++j) {
        const int t = min(8 * kq + j, 26);   // taps past 26 read any in-bounds row: their weights are zero
        toff[j] = ((t / 9 - 1) * HY + ((t / 3) % 3 - 1)) * HX + (t % 3 - 1);
    }
#pragma unroll
    for (int g = 0; g < G; ++g) {
        const int vi = (wave * G + g) * 16 + n;   // a group is one x row: tx = n
        const int ty = (vi / TX) % TY, tz = vi / (TX * TY);
        const int hb = ((tz + 1) * HY + (ty + 1)) * HX + (n + 1);
        bf16x8 xh, xl;
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            const float v = sx[hb + toff[j]];
            xh[j] = (__bf16)v;
            xl[j] = (__bf16)(v - (float)xh[j]);
        }
        f32x4 acc = f32x4{0.f, 0.f, 0.f, 0.f};
        acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(xh, wh, acc, 0, 0, 0);
        acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(xh, wl, acc, 0, 0, 0);
        acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(xl, wh, acc, 0, 0, 0);
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int vo = (wave * G + g) * 16 + kq * 4 + q;
            sOut[vo * 16 + 8 * swz<2>(vo, n >> 3) + (n & 7)] = (uint16_t)f2bf(acc[q]);
        }
    }
    __syncthreads();
    for (int it = tid; it < TV * 2; it += NB) {
        const int vi = it >> 1, h = it & 1;
        const int tx = vi % TX, ty = (vi / TX) % TY, tz = vi / (TX * TY);
        const int gx = ox + tx, gy = oy + ty, gz = oz + tz;
        if (gx < nx && gy < ny && gz < nz)
            *(uint4*)(out + ((size_t)(gz * ny + gy) * nx + gx) * 16 + 8 * h) = *(const uint4*)&sOut[vi * 16 + 8 * swz<2>(vi, h)];
    }
}

// out: 16 -> 1 channel on MFMA, in k_convf16's row scheme (tile 16x4x4, 16-voxel x rows,
// one v_mfma_f32_16x16x16_bf16 per tap, the dx = +-1 operands by DPP row shifts).  The
// FP32 weights W[t][0][ci] enter as two BF16 output columns, their high part hi = bf16(w)
// and remainder lo = bf16(w - hi) (w = hi + lo to 2^-17 relative; k_pack_out), so column 0
// of the product holds sum a hi and column 1 sum a lo, each product exact in FP32 and the
// sums accumulated in FP32; y = column 0 + column 1.  Columns 2-15 are zero weights.
__global__ __launch_bounds__(NB) void k_outf16(const uint16_t* in, const uint16_t* wpk, float* y, int nx, int ny, int nz,
                                               const double* done) {
    const double flag = flagOf(done);
    constexpr int TX = 16, TY = 4, TZ = 4, CIN = 16, C8 = 2;
    constexpr int HX = TX + 2, HY = TY + 2, HZ = TZ + 2, HV = HX * HY * HZ;
    constexpr int TV = TX * TY * TZ, G = TV / 64;
    __shared__ __align__(16) uint16_t sIn[HV * CIN];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int n = lane & 15, kq = lane >> 4;
    const int ntx = (nx + TX - 1) / TX, nty = (ny + TY - 1) / TY;
    int b = blockIdx.x;
    const int bx = b % ntx;
    b /= ntx;
    const int by = b % nty, bz = b / nty;
    const int ox = bx * TX, oy = by * TY, oz = bz * TZ;

    constexpr int NI = HV * C8, NIT = (NI + NB - 1) / NB;
    uint4 vin[NIT];
#pragma unroll
    for (int u = 0; u < NIT; ++u) {
        const int it = tid + u * NB;
        vin[u] = make_uint4(0u, 0u, 0u, 0u);
        if (it < NI) {
            const int hv = it / C8, c = it - hv * C8;
            const int hx = hv % HX, hy = (hv / HX) % HY, hz = hv / (HX * HY);
            const int gx = ox + hx - 1, gy = oy + hy - 1, gz = oz + hz - 1;
            if (gx >= 0 && gx < nx && gy >= 0 && gy < ny && gz >= 0 && gz < nz)
                vin[u] = *(const uint4*)(in + ((size_t)(gz * ny + gy) * nx + gx) * CIN + 8 * c);
        }
    }
    if (flag != 0.) return;
#pragma unroll
    for (int u = 0; u < NIT; ++u) {
        const int it = tid + u * NB;
        if (it < NI) {
            const int hv = it / C8, c = it - hv * C8;
            *(uint4*)&sIn[hv * CIN + 8 * swz<C8>(hv, c)] = vin[u];
        }
    }
    // lane (n, kq) holds wpk[t][co = n][ci = 4kq..4kq+3] for all 27 taps
    const uint16_t* ws = wpk + n * 16 + 4 * kq;
    uint2 wr[27];
#pragma unroll
    for (int t = 0; t < 27; ++t) wr[t] = *(const uint2*)(ws + t * 256);
    __syncthreads();

    auto frag = [&](int row) { return *(const uint2*)&sIn[row * CIN + 8 * swz<C8>(row, kq >> 1) + 4 * (kq & 1)]; };
    const int hoff = n < 8 ? -1 : 1;
    int base[G];
#pragma unroll
    for (int g = 0; g < G; ++g) {
        const int vi = (wave * G + g) * 16 + n;
        const int ty = (vi / TX) % TY, tz = vi / (TX * TY);
        base[g] = ((tz + 1) * HY + (ty + 1)) * HX + (n + 1);
    }
    f32x4 acc[G];
#pragma unroll
    for (int g = 0; g < G; ++g) acc[g] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int p = 0; p < 9; ++p) {
        const int dz = p / 3 - 1, dy = p % 3 - 1, t0 = 3 * p;
        const int poff = (dz * HY + dy) * HX;
#pragma unroll
        for (int g = 0; g < G; ++g) {
            const int row = base[g] + poff;
            const uint2 c0 = frag(row);
            const uint2 h = frag(row + hoff);
            const bf16x4 am = asbf4(shr1(h.x, c0.x), shr1(h.y, c0.y));
            const bf16x4 ap = asbf4(shl1(h.x, c0.x), shl1(h.y, c0.y));
            acc[g] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(am, asbf4(wr[t0].x, wr[t0].y), acc[g], 0, 0, 0);
            acc[g] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(asbf4(c0.x, c0.y), asbf4(wr[t0 + 1].x, wr[t0 + 1].y), acc[g], 0, 0, 0);
            acc[g] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(ap, asbf4(wr[t0 + 2].x, wr[t0 + 2].y), acc[g], 0, 0, 0);
        }
    }
    // lane (0, kq) of group g: y at the row's voxels tx = 4kq .. 4kq+3 = column 0 + column 1
#pragma unroll
    for (int g = 0; g < G; ++g) {
        float v[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) v[q] = acc[g][q] + __shfl_down(acc[g][q], 1, 64);
        const int vi = (wave * G + g) * 16;
        const int ty = (vi / TX) % TY, tz = vi / (TX * TY);
        const int gx = ox + 4 * kq, gy = oy + ty, gz = oz + tz;
        if (n == 0 && gx < nx && gy < ny && gz < nz)   // nx is a multiple of 4: the 4 voxels are in
            *(float4*)(y + ((size_t)(gz * ny + gy) * nx + gx)) = make_float4(v[0], v[1], v[2], v[3]);
    }
}

// gradient of the out layer's input: ga[v][ci] += sum_t W[t][0][ci] g[v - off(t)]
__global__ void k_out_ginput(const float* g, const float* W, float* ga, int nx, int ny, int nz) {
    __shared__ float sW[27 * 16];
    for (int i = threadIdx.x; i < 27 * 16; i += blockDim.x) sW[i] = W[i];
    __syncthreads();
    const int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= nx * ny * nz) return;
    int i, j, k;
    vox(v, nx, ny, i, j, k);
    float gv[27];
#pragma unroll
    for (int t = 0; t < 27; ++t) {
        bool in;
        const int u = nbr(i, j, k, 26 - t, nx, ny, nz, in);   // v - off(t) = v + off(26 - t)
        const float val = g[u];
        gv[t] = in ? val : 0.f;
    }
    float acc[16];
#pragma unroll
    for (int c = 0; c < 16; ++c) acc[c] = 0.f;
#pragma unroll
    for (int t = 0; t < 27; ++t)
#pragma unroll
        for (int c = 0; c < 16; ++c) acc[c] += sW[t * 16 + c] * gv[t];
    float* o = ga + (size_t)v * 16;
    for (int c = 0; c < 16; ++c) o[c] += acc[c];
}

// Weight gradients of the edge layers, one (tap, channel) pair per thread, a chunk of
// voxels per block:
//   IN  (x: [v] FP32 input, g: [v][16] FP32):  dW[t][c] = sum_v g[v][c] x[v+off(t)]
//   OUT (x: [v][16] BF16 input, g: [v] FP32):  dW[t][c] = sum_v g[v] x[v+off(t)][c]
template <bool IN>
__global__ void k_edge_wgrad(const void* xv, const float* g, float* part, int nx, int ny, int nz, int vpc) {
    const int pr = threadIdx.x;
    if (pr >= 27 * 16) return;
    const int t = pr / 16, c = pr % 16;
    const int dx = t % 3 - 1, dy = (t / 3) % 3 - 1, dz = t / 9 - 1;
    const int nv = nx * ny * nz;
    const int v0 = blockIdx.x * vpc, v1 = min(nv, v0 + vpc);
    float s = 0.f;
    for (int v = v0; v < v1; ++v) {
        int i, j, k;
        vox(v, nx, ny, i, j, k);
        const int a = i + dx, b = j + dy, cz = k + dz;
        if (a < 0 || a >= nx || b < 0 || b >= ny || cz < 0 || cz >= nz) continue;
        const int u = (cz * ny + b) * nx + a;
        if (IN) s += g[(size_t)v * 16 + c] * ((const float*)xv)[u];
        else s += g[v] * bf2f(((const uint16_t*)xv)[(size_t)u * 16 + c]);
    }
    part[(size_t)blockIdx.x * 432 + pr] = s;
}

// ---------------------------------------------------------------- adjoints of staging (FP32)
// g_fine[v][c] += g[v/2][c] / 8   (fine dims 2nx x 2ny x 2nz)
__global__ void k_pool_adj(const float* g, float* gf, int nx, int ny, int nz, int C) {
    const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const long nf = 8L * nx * ny * nz * C;
    if (e >= nf) return;
    const int c = (int)(e % C);
    const long v = e / C;
    const int fx = 2 * nx, fy = 2 * ny;
    const int i = (int)(v % fx), j = (int)((v / fx) % fy), k = (int)(v / ((long)fx * fy));
    gf[e] += 0.125f * g[((size_t)((k >> 1) * ny + (j >> 1)) * nx + (i >> 1)) * C + c];
}
// g_coarse[V][c] += sum of the 8 children of V   (coarse dims nx x ny x nz)
__global__ void k_up_adj(const float* g, float* gc, int nx, int ny, int nz, int C) {
    const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (long)nx * ny * nz * C) return;
    const int c = (int)(e % C);
    const long v = e / C;
    const int i = (int)(v % nx), j = (int)((v / nx) % ny), k = (int)(v / ((long)nx * ny));
    const int fx = 2 * nx, fy = 2 * ny;
    float s = 0.f;
    for (int dz = 0; dz < 2; ++dz)
        for (int dy = 0; dy < 2; ++dy)
            for (int dx = 0; dx < 2; ++dx)
                s += g[((size_t)((2 * k + dz) * fy + 2 * j + dy) * fx + 2 * i + dx) * C + c];
    gc[e] += s;
}
__global__ void k_addf(const float* a, float* b, long n) {
    const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (e < n) b[e] += a[e];
}
__global__ void k_zerof(float* a, long n) {
    const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (e < n) a[e] = 0.f;
}

// ---------------------------------------------------------------- parameters
// forward packing: dst[cob][cib][t][co16][ci16] = W[t][co][ci]
// input-gradient packing (flip): layer with cin' = cout, cout' = cin and
//   W'[t][co' = ci][ci' = co] = W[26 - t][co][ci]
__global__ void k_pack(const float* W, uint16_t* dst, int cin, int cout, int flip) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    const int L = 27 * cin * cout;
    if (e >= L) return;
    const int ci_ = cin, co_ = cout;
    // e enumerates the packed layout of the (possibly flipped) layer
    const int pin = flip ? co_ : ci_, pout = flip ? ci_ : co_;
    int r = e;
    const int cil = r % 16; r /= 16;
    const int col = r % 16; r /= 16;
    const int t = r % 27; r /= 27;
    const int ncb = pin / 16;
    const int cib = r % ncb, cob = r / ncb;
    const int pco = cob * 16 + col, pci = cib * 16 + cil;
    float v;
    if (!flip) v = W[((size_t)t * co_ + pco) * ci_ + pci];
    else v = W[((size_t)(26 - t) * co_ + pci) * ci_ + pco];
    (void)pout;
    dst[e] = (uint16_t)f2bf(v);
}

// the out layer for k_outf16: dst[t][co][ci] with co 0 = bf16(W[t][0][ci]), co 1 = its BF16
// remainder, co 2-15 zero
__global__ void k_pack_out(const float* W, uint16_t* dst) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= 27 * 256) return;
    const int ci = e % 16, co = (e / 16) % 16, t = e / 256;
    const float w = W[t * 16 + ci];
    const uint32_t hi = f2bf(w);
    dst[e] = co == 0 ? (uint16_t)hi : (co == 1 ? (uint16_t)f2bf(w - bf2f(hi)) : (uint16_t)0);
}

// Gradient clipping: the global norm of the (batch-averaged) gradient is limited to
// 5x a running average of the previous steps' norms.  A deep linear network trained by
// Adam otherwise meets rare gradient spikes while the second moments are still small,
// and one such step can throw the loss up by orders of magnitude.
// clip[0] = running norm (0 before the first step), clip[1] = factor for this step.
__global__ void k_gnorm_part(const float* g, long n, float gscale, double* part) {
    double v[1] = {0.};
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
        const double x = (double)g[i] * gscale;
        v[0] += x * x;
    }
    writePartials<256, 1>(v, part);
}
__global__ void k_gnorm_final(const double* part, int nb, double* clip) {
    double t[1];
    readPartials<256, 1>(part, nb, t);
    if (threadIdx.x == 0) {
        const double nrm = sqrt(t[0]);
        const double lim = clip[0] > 0. ? 5. * clip[0] : nrm;
        const double used = nrm > lim ? lim : nrm;
        clip[1] = (nrm > 0. && std::isfinite(nrm)) ? used / nrm : 0.;   // a non-finite gradient is dropped
        if (std::isfinite(used)) clip[0] = clip[0] > 0. ? 0.9 * clip[0] + 0.1 * used : used;
    }
}

__global__ void k_adam(float* th, const float* g, float* m1, float* m2, long n, float lr, float gscale, const double* clip,
                       float b1, float b2, float c1, float c2, float eps) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i] * gscale * (float)clip[1];
    const float a = b1 * m1[i] + (1.f - b1) * gi;
    const float b = b2 * m2[i] + (1.f - b2) * gi * gi;
    m1[i] = a;
    m2[i] = b;
    th[i] -= lr * (a / c1) / (sqrtf(b / c2) + eps);
}

// ---------------------------------------------------------------- host side
void* net_alloc(Net& n, size_t bytes, std::vector<void*>& list) {
    void* p = nullptr;
    HIPCHECK(hipMalloc(&p, bytes));
    HIPCHECK(hipMemset(p, 0, bytes));
    list.push_back(p);
    return p;
}

static inline int lvlDim(int n0, int L) { return n0 >> L; }

void net_set_dims(Net& n, int nx, int ny, int nz, bool train) {
    if (n.nx == nx && n.ny == ny && n.nz == nz && (n.trainBufs || !train)) return;
    HIPCHECK(hipDeviceSynchronize());
    for (void* p : n.dimAllocs) (void)hipFree(p);
    n.dimAllocs.clear();
    for (int l = 0; l < NET_LAYERS; ++l) { n.act[l] = nullptr; n.gact[l] = nullptr; n.pooled[l] = nullptr; }
    n.nx = nx; n.ny = ny; n.nz = nz;
    n.trainBufs = train;
    const size_t v0 = (size_t)nx * ny * nz;
    size_t gst = 0;
    for (int l = 0; l < NET_LAYERS; ++l) {
        const LayerDef& d = kNet[l];
        const size_t vl = (size_t)lvlDim(nx, d.level) * lvlDim(ny, d.level) * lvlDim(nz, d.level);
        if (l != NET_OUT) {
            n.act[l] = (uint16_t*)net_alloc(n, vl * d.cout * 2, n.dimAllocs);
            if (train) n.gact[l] = (float*)net_alloc(n, vl * d.cout * 4, n.dimAllocs);
        }
        gst = std::max(gst, vl * d.cin);
        if (d.mode == MODE_POOL)   // the producer's epilogue writes the pooled input of this layer
            n.pooled[d.src] = (uint16_t*)net_alloc(n, vl * d.cin * 2, n.dimAllocs);
    }
    n.vin = (float*)net_alloc(n, v0 * 4, n.dimAllocs);
    n.vout = (float*)net_alloc(n, v0 * 4, n.dimAllocs);
    n.gout = (float*)net_alloc(n, v0 * 4, n.dimAllocs);
    if (train) n.gstage = (float*)net_alloc(n, gst * 4, n.dimAllocs);
}

void net_pack(Net& n, hipStream_t s) {
    for (int l = 1; l < NET_OUT; ++l) {
        const LayerDef& d = kNet[l];
        const int L = 27 * d.cin * d.cout;
        hipLaunchKernelGGL(k_pack, dim3((L + 255) / 256), dim3(256), 0, s, n.theta + n.off[l], n.wf + n.poff[l], d.cin, d.cout, 0);
        hipLaunchKernelGGL(k_pack, dim3((L + 255) / 256), dim3(256), 0, s, n.theta + n.off[l], n.wb + n.poff[l], d.cin, d.cout, 1);
    }
    hipLaunchKernelGGL(k_pack_out, dim3(27), dim3(256), 0, s, n.theta + n.off[NET_OUT], n.wf + n.poff[NET_OUT]);
    n.packed = true;
}

// Tile by level size: 16x4x4 on grids of 32 and more, 8x4x4 on 16, 4x4x4 below (more
// workgroups where the grid is small).
template <int CIN, int MODE, class TIN, class TOUT>
static void convLaunchC(hipStream_t s, const ConvArgs& a) {
    if (a.nx >= 32) {
        const int nt = ((a.nx + 15) / 16) * ((a.ny + 3) / 4) * ((a.nz + 3) / 4);
        hipLaunchKernelGGL((k_conv<16, 4, 4, CIN, MODE, TIN, TOUT>), dim3(nt, a.cout / 16), dim3(NB), 0, s, a);
    } else if (a.nx >= 16) {
        const int nt = ((a.nx + 7) / 8) * ((a.ny + 3) / 4) * ((a.nz + 3) / 4);
        hipLaunchKernelGGL((k_conv<8, 4, 4, CIN, MODE, TIN, TOUT>), dim3(nt, a.cout / 16), dim3(NB), 0, s, a);
    } else {
        const int nt = ((a.nx + 3) / 4) * ((a.ny + 3) / 4) * ((a.nz + 3) / 4);
        hipLaunchKernelGGL((k_conv<4, 4, 4, CIN, MODE, TIN, TOUT>), dim3(nt, a.cout / 16), dim3(NB), 0, s, a);
    }
}
template <int MODE, class TIN, class TOUT>
static void convLaunch(hipStream_t s, const ConvArgs& a) {
    if (a.cin == 16) convLaunchC<16, MODE, TIN, TOUT>(s, a);
    else if (a.cin == 32) convLaunchC<32, MODE, TIN, TOUT>(s, a);
    else if (a.cin == 64) convLaunchC<64, MODE, TIN, TOUT>(s, a);
    else throw std::string("convolution input channels must be 16, 32 or 64");
}

// Forward tiles by level size: 16x4x4 on grids of 64 and more, 8x4x4 on 32, 4x4x4 below
// (more workgroups where the grid is small); the 16-input-channel layers take k_convf16
// from 32 up.  Measured on the learned solve at 128^3 (coarse grid 64^3): 211 us per FCG
// iteration with 16x4x4 on 32^3 and 8x4x4 on 16^3, 199 us with this choice, results
// identical (a tile's outputs do not depend on its shape).
template <int CIN, int MODE, bool POOLOUT>
static void convFLaunchC(hipStream_t s, const ConvFArgs& a) {
    if (CIN == 16 && a.nx >= 32) {
        const int nt = ((a.nx + 15) / 16) * ((a.ny + 3) / 4) * ((a.nz + 3) / 4);
        hipLaunchKernelGGL((k_convf16<MODE, POOLOUT>), dim3(nt, a.cout / 16), dim3(NB), 0, s, a);
    } else if (a.nx >= 64) {
        const int nt = ((a.nx + 15) / 16) * ((a.ny + 3) / 4) * ((a.nz + 3) / 4);
        hipLaunchKernelGGL((k_convf<16, 4, 4, CIN, MODE, POOLOUT>), dim3(nt, a.cout / 16), dim3(NB), 0, s, a);
    } else if (a.nx >= 32) {
        const int nt = ((a.nx + 7) / 8) * ((a.ny + 3) / 4) * ((a.nz + 3) / 4);
        hipLaunchKernelGGL((k_convf<8, 4, 4, CIN, MODE, POOLOUT>), dim3(nt, a.cout / 16), dim3(NB), 0, s, a);
    } else {
        const int nt = ((a.nx + 3) / 4) * ((a.ny + 3) / 4) * ((a.nz + 3) / 4);
        hipLaunchKernelGGL((k_convf<4, 4, 4, CIN, MODE, POOLOUT>), dim3(nt, a.cout / 16), dim3(NB), 0, s, a);
    }
}
template <int MODE, bool POOLOUT>
static void convFLaunch(hipStream_t s, const ConvFArgs& a, int cin) {
    if (cin == 16) convFLaunchC<16, MODE, POOLOUT>(s, a);
    else if (cin == 32) convFLaunchC<32, MODE, POOLOUT>(s, a);
    else if (cin == 64) convFLaunchC<64, MODE, POOLOUT>(s, a);
    else throw std::string("convolution input channels must be 16, 32 or 64");
}

void net_forward(Net& n, hipStream_t s, const float* vin, float* vout, const double* done) {
    if (!done) done = n.scal + NET_ZERO_SLOT;   // a flag that stays zero
    const int nt16 = ((n.nx + 15) / 16) * ((n.ny + 3) / 4) * ((n.nz + 3) / 4);
    hipLaunchKernelGGL(k_inf16, dim3(nt16), dim3(NB), 0, s, vin, (const float*)(n.theta + n.off[0]), n.act[0], n.nx, n.ny,
                       n.nz, done);
    for (int l = 1; l < NET_OUT; ++l) {
        const LayerDef& d = kNet[l];
        ConvFArgs a;
        // a POOL layer reads the average its source layer's epilogue wrote
        a.in = d.mode == MODE_POOL ? n.pooled[d.src] : n.act[d.src];
        a.skip = d.skip >= 0 ? n.act[d.skip] : nullptr;
        a.w = n.wf + n.poff[l];
        a.out = n.act[l];
        a.pooled = n.pooled[l];
        a.nx = lvlDim(n.nx, d.level); a.ny = lvlDim(n.ny, d.level); a.nz = lvlDim(n.nz, d.level);
        a.cout = d.cout; a.done = done;
        if (d.mode == MODE_UPADD) {
            if (a.pooled) convFLaunch<MODE_UPADD, true>(s, a, d.cin);
            else convFLaunch<MODE_UPADD, false>(s, a, d.cin);
        } else {
            if (a.pooled) convFLaunch<MODE_PLAIN, true>(s, a, d.cin);
            else convFLaunch<MODE_PLAIN, false>(s, a, d.cin);
        }
    }
    hipLaunchKernelGGL(k_outf16, dim3(nt16), dim3(NB), 0, s, (const uint16_t*)n.act[NET_OUT - 1],
                       (const uint16_t*)(n.wf + n.poff[NET_OUT]), vout, n.nx, n.ny, n.nz, done);
}

static void ensurePart(Net& n, long need) {
    if (need <= n.wpartCap) return;
    HIPCHECK(hipDeviceSynchronize());
    if (n.wpart) {
        for (auto it = n.allocs.begin(); it != n.allocs.end(); ++it)
            if (*it == n.wpart) { n.allocs.erase(it); break; }
        (void)hipFree(n.wpart);
    }
    n.wpart = (float*)net_alloc(n, need * sizeof(float), n.allocs);
    n.wpartCap = need;
}

template <int MODE>
static void wgradLaunch(Net& n, hipStream_t s, WgArgs a, float* dst) {
    int TX = a.nx >= 16 ? 16 : (a.nx >= 8 ? 8 : 4);
    const int nt = ((a.nx + TX - 1) / TX) * ((a.ny + 3) / 4) * ((a.nz + 3) / 4);
    const int pairs = (a.cout / 16) * (a.cin / 16);
    int nChunk = std::max(1, std::min(nt, 512 / pairs));
    a.tilesPerChunk = (nt + nChunk - 1) / nChunk;
    nChunk = (nt + a.tilesPerChunk - 1) / a.tilesPerChunk;
    a.ntiles = nt;
    const long L = 27L * a.cin * a.cout;
    ensurePart(n, L * nChunk);
    a.part = n.wpart;
    if (TX == 16) hipLaunchKernelGGL((k_conv_wgrad<16, 4, 4, MODE>), dim3(nChunk, pairs), dim3(NB), 0, s, a);
    else if (TX == 8) hipLaunchKernelGGL((k_conv_wgrad<8, 4, 4, MODE>), dim3(nChunk, pairs), dim3(NB), 0, s, a);
    else hipLaunchKernelGGL((k_conv_wgrad<4, 4, 4, MODE>), dim3(nChunk, pairs), dim3(NB), 0, s, a);
    hipLaunchKernelGGL(k_reduce_parts, dim3((L + 255) / 256), dim3(256), 0, s, n.wpart, nChunk, (int)L, dst, 1);
}

static void edgeWgrad(Net& n, hipStream_t s, bool in, const void* x, const float* g, float* dst) {
    const int nv = n.nx * n.ny * n.nz;
    const int vpc = 256;
    const int nChunk = (nv + vpc - 1) / vpc;
    ensurePart(n, 432L * nChunk);
    if (in) hipLaunchKernelGGL(k_edge_wgrad<true>, dim3(nChunk), dim3(448), 0, s, x, g, n.wpart, n.nx, n.ny, n.nz, vpc);
    else hipLaunchKernelGGL(k_edge_wgrad<false>, dim3(nChunk), dim3(448), 0, s, x, g, n.wpart, n.nx, n.ny, n.nz, vpc);
    hipLaunchKernelGGL(k_reduce_parts, dim3((432 + 255) / 256), dim3(256), 0, s, n.wpart, nChunk, 432, dst, 1);
}

void net_backward(Net& n, hipStream_t s, const float* vin, const float* gout) {
    // zero the activation gradients
    for (int l = 0; l < NET_OUT; ++l) {
        const LayerDef& d = kNet[l];
        const long e = (long)lvlDim(n.nx, d.level) * lvlDim(n.ny, d.level) * lvlDim(n.nz, d.level) * d.cout;
        hipLaunchKernelGGL(k_zerof, dim3((e + 255) / 256), dim3(256), 0, s, n.gact[l], e);
    }
    const int nv0 = n.nx * n.ny * n.nz;
    // out layer
    edgeWgrad(n, s, false, n.act[NET_OUT - 1], gout, n.grad + n.off[NET_OUT]);
    hipLaunchKernelGGL(k_out_ginput, dim3((nv0 + 255) / 256), dim3(256), 0, s, gout, n.theta + n.off[NET_OUT], n.gact[NET_OUT - 1], n.nx, n.ny, n.nz);
    for (int l = NET_OUT - 1; l >= 1; --l) {
        const LayerDef& d = kNet[l];
        const int lx = lvlDim(n.nx, d.level), ly = lvlDim(n.ny, d.level), lz = lvlDim(n.nz, d.level);
        WgArgs w;
        w.x = n.act[d.src];
        w.skip = d.skip >= 0 ? n.act[d.skip] : nullptr;
        w.g = n.gact[l];
        w.nx = lx; w.ny = ly; w.nz = lz; w.cin = d.cin; w.cout = d.cout;
        float* dst = n.grad + n.off[l];
        if (d.mode == MODE_PLAIN) wgradLaunch<MODE_PLAIN>(n, s, w, dst);
        else if (d.mode == MODE_POOL) wgradLaunch<MODE_POOL>(n, s, w, dst);
        else wgradLaunch<MODE_UPADD>(n, s, w, dst);
        // gradient of the staged input: the conv with flipped taps and swapped channels
        ConvArgs a;
        a.in = n.gact[l];
        a.skip = nullptr;
        a.w = n.wb + n.poff[l];
        a.nx = lx; a.ny = ly; a.nz = lz; a.cin = d.cout; a.cout = d.cin; a.done = nullptr;
        if (d.mode == MODE_PLAIN) {
            a.out = n.gact[d.src];
            a.accum = 1;
            convLaunch<MODE_PLAIN, float, float>(s, a);
        } else {
            a.out = n.gstage;
            a.accum = 0;
            convLaunch<MODE_PLAIN, float, float>(s, a);
            const long e = (long)lx * ly * lz * d.cin;
            if (d.mode == MODE_POOL) {
                hipLaunchKernelGGL(k_pool_adj, dim3((8 * e + 255) / 256), dim3(256), 0, s, n.gstage, n.gact[d.src], lx, ly, lz, d.cin);
            } else {
                hipLaunchKernelGGL(k_up_adj, dim3((e / 8 + 255) / 256), dim3(256), 0, s, n.gstage, n.gact[d.src], lx / 2, ly / 2, lz / 2, d.cin);
                hipLaunchKernelGGL(k_addf, dim3((e + 255) / 256), dim3(256), 0, s, n.gstage, n.gact[d.skip], e);
            }
        }
    }
    edgeWgrad(n, s, true, vin, n.gact[NET_IN], n.grad + n.off[NET_IN]);
}

void net_adam(Net& n, hipStream_t s, float lr, float gscale) {
    const float b1 = 0.9f, b2 = 0.999f;
    ++n.adamT;
    const float c1 = 1.f - std::pow(b1, (float)n.adamT), c2 = 1.f - std::pow(b2, (float)n.adamT);
    double* clip = n.scal + 24;
    if (n.adamT == 1) HIPCHECK(hipMemsetAsync(clip, 0, 2 * sizeof(double), s));
    const int nb = 512;
    hipLaunchKernelGGL(k_gnorm_part, dim3(nb), dim3(256), 0, s, (const float*)n.grad, n.nParams, gscale, n.part);
    hipLaunchKernelGGL(k_gnorm_final, dim3(1), dim3(256), 0, s, (const double*)n.part, nb, clip);
    hipLaunchKernelGGL(k_adam, dim3((n.nParams + 255) / 256), dim3(256), 0, s, n.theta, n.grad, n.m1, n.m2, n.nParams, lr,
                       gscale, (const double*)clip, b1, b2, c1, c2, 1e-8f);
    net_pack(n, s);
}

}  // namespace caffe
