// netdev.hpp -- device helpers shared by the network kernels (net.hip) and the flexible-CG
// kernels that feed the network (learned.hip): BF16 conversion, the edge layers' tile
// (16x4x4 output voxels and their one-voxel halo, one thread per output voxel) and the
// input layer (1 -> 16 channels) evaluated from a staged halo tile.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace caffe {

__device__ __forceinline__ float bf2f(uint32_t h) { return __uint_as_float(h << 16); }
// round to nearest even (NaN stays NaN)
__device__ __forceinline__ uint32_t f2bf(float f) {
    uint32_t u = __float_as_uint(f);
    if ((u & 0x7fffffffu) > 0x7f800000u) return (u >> 16) | 0x40u;
    u += 0x7fffu + ((u >> 16) & 1u);
    return u >> 16;
}

// The edge layers stage a 16x4x4 output tile with its one-voxel halo in LDS (clamped
// index and mask: no branch around the loads, which therefore issue together), then one
// thread computes one output voxel from LDS.
constexpr int ETX = 16, ETY = 4, ETZ = 4, EHX = ETX + 2, EHY = ETY + 2, EHZ = ETZ + 2, EHV = EHX * EHY * EHZ;
__device__ __forceinline__ void edgeTile(int nx, int ny, int& ox, int& oy, int& oz) {
    const int ntx = (nx + ETX - 1) / ETX, nty = (ny + ETY - 1) / ETY;
    int b = blockIdx.x;
    ox = (b % ntx) * ETX;
    b /= ntx;
    oy = (b % nty) * ETY;
    oz = (b / nty) * ETZ;
}
// halo voxel hv of the tile at (ox,oy,oz): clamped global index and whether it is inside
__device__ __forceinline__ int haloVox(int hv, int ox, int oy, int oz, int nx, int ny, int nz, bool& in) {
    const int hx = hv % EHX, hy = (hv / EHX) % EHY, hz = hv / (EHX * EHY);
    const int gx = ox + hx - 1, gy = oy + hy - 1, gz = oz + hz - 1;
    in = gx >= 0 && gx < nx && gy >= 0 && gy < ny && gz >= 0 && gz < nz;
    return (min(max(gz, 0), nz - 1) * ny + min(max(gy, 0), ny - 1)) * nx + min(max(gx, 0), nx - 1);
}
inline int edgeBlocks(int nx, int ny, int nz) {
    return ((nx + ETX - 1) / ETX) * ((ny + ETY - 1) / ETY) * ((nz + ETZ - 1) / ETZ);
}

// The input layer at the tile voxel of thread tid: 16 channels from the staged halo tile
// sx ([hz][hy][hx], zero outside the grid) and the weights sW ([27][16] FP32), rounded to
// BF16 and stored channels last.
__device__ __forceinline__ void inLayerTile(const float* sx, const float* sW, uint16_t* out, int ox, int oy, int oz,
                                            int nx, int ny, int nz, int tid) {
    const int tx = tid % ETX, ty = (tid / ETX) % ETY, tz = tid / (ETX * ETY);
    const int gx = ox + tx, gy = oy + ty, gz = oz + tz;
    const int hb = ((tz + 1) * EHY + (ty + 1)) * EHX + (tx + 1);
    float acc[16];
#pragma unroll
    for (int c = 0; c < 16; ++c) acc[c] = 0.f;
#pragma unroll 3
    for (int t = 0; t < 27; ++t) {
        const float xv = sx[hb + ((t / 9 - 1) * EHY + ((t / 3) % 3 - 1)) * EHX + (t % 3 - 1)];
#pragma unroll
        for (int c4 = 0; c4 < 4; ++c4) {
            const float4 w4 = *(const float4*)&sW[t * 16 + 4 * c4];
            acc[4 * c4] += w4.x * xv;
            acc[4 * c4 + 1] += w4.y * xv;
            acc[4 * c4 + 2] += w4.z * xv;
            acc[4 * c4 + 3] += w4.w * xv;
        }
    }
    if (gx < nx && gy < ny && gz < nz) {
        uint32_t w[8];
#pragma unroll
        for (int c = 0; c < 8; ++c) w[c] = f2bf(acc[2 * c]) | (f2bf(acc[2 * c + 1]) << 16);
        uint4* o = (uint4*)(out + ((size_t)(gz * ny + gy) * nx + gx) * 16);
        o[0] = make_uint4(w[0], w[1], w[2], w[3]);
        o[1] = make_uint4(w[4], w[5], w[6], w[7]);
    }
}

}  // namespace caffe
