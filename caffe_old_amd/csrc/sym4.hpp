// sym4.hpp -- the pressure-correction operator: symmetric 7-point storage (aP and one
// coefficient per low face) with optional 9-entry ghost rows, shared by the Krylov
// solvers and the multigrid preconditioner.
#pragma once
#include "internal.hpp"

namespace caffe {

struct OvView {
    const int* rowOf;      // nullptr when there are no ghost rows
    const int* cols;
    const void* vals;      // double or float, 9 per row
};

template <class T>
struct Sym4 {              // pressure-correction matrix
    const T* A; int P, sJ, sK; OvView ov; const int8_t* st;
    template <class F>
    __device__ double apply(F x, int p) const {
        if (ov.rowOf) {
            if (st[p] == ST_INACTIVE) return x(p);
            int r = ov.rowOf[p];
            if (r >= 0) {
                const T* v = (const T*)ov.vals + 9 * r;
                int cols[8];
#pragma unroll
                for (int m = 0; m < 8; ++m) cols[m] = ov.cols[9 * r + 1 + m];
                double s = (double)v[0] * x(p);
#pragma unroll
                for (int m = 0; m < 8; ++m)
                    if (cols[m] >= 0) s += (double)v[1 + m] * x(cols[m]);
                return s;
            }
        }
        return (double)A[p] * x(p) + (double)A[P + p] * x(p - 1) + (double)A[P + p + 1] * x(p + 1)
             + (double)A[2 * P + p] * x(p - sJ) + (double)A[2 * P + p + sJ] * x(p + sJ)
             + (double)A[3 * P + p] * x(p - sK) + (double)A[3 * P + p + sK] * x(p + sK);
    }
    __device__ double diag(int p) const { return (double)A[p]; }
};

}  // namespace caffe
