// drop_in.cpp -- the reference's calling pattern, compiled against include/caffe_hip.hpp and
// linked to libcaffe_hip.so, as a maintainer would bind the GPU path into the reference:
//
//  1. a SparseMatrix / SparseVector system assembled row by row the way Diffusion::solve
//     does it (Diffusion.cpp:61-130): 7-point rows whose off-domain neighbours come back as
//     column -1 from IndexMap (IndexMap.cpp:77) and are dropped by setRow; solve(b, x) from
//     x = 0; the residual res = A x - b by scale(-1, b) and multiplyAdd, and its infNorm;
//  2. three PISO steps of a 16^3 lid-driven cavity through FlowSolver (the Solver interface,
//     Modules/caffePiso.cpp:30-40), writing the cell values of u, v, w and p to argv[1]
//     (doubles, Array3D order) for the comparison with the oracle.
// Prints one "ok" line per part; exits nonzero on failure.
#include <cmath>
#include <cstdio>
#include <vector>

#include "caffe_hip.hpp"

using namespace caffe_gpu;

static int sparsePart() {
    const int n = 20, N = n * n * n;
    auto id = [&](int i, int j, int k) { return (i < 0 || j < 0 || k < 0 || i >= n || j >= n || k >= n) ? -1 : (k * n + j) * n + i; };
    SparseMatrix A(1e-10, -2., 20000);
    A.allocate(N, N, 7);
    SparseVector b(N), x(N), xe(N), res(N);
    // a diffusion operator with a cell-dependent coefficient and fixed values outside
    std::vector<double> exact(N);
    for (int q = 0; q < N; ++q) exact[q] = std::sin(0.1 * q) + 0.01 * (q % 17);
    for (int k = 0; k < n; ++k)
        for (int j = 0; j < n; ++j)
            for (int i = 0; i < n; ++i) {
                int cols[7] = {id(i, j, k), id(i + 1, j, k), id(i - 1, j, k), id(i, j + 1, k), id(i, j - 1, k),
                               id(i, j, k + 1), id(i, j, k - 1)};
                double kap = 1.0 + 0.5 * std::cos(0.3 * (i + 2 * j + 3 * k));
                double vals[7];
                vals[0] = 0.;
                for (int m = 1; m < 7; ++m) { vals[m] = kap; vals[0] -= kap; }   // boundary neighbours stay in a0
                vals[0] -= 0.1;                                                   // an unsteady a0P term
                A.setRow(cols[0], 7, cols, vals);
            }
    // b = A x_exact, formed by the GPU product
    for (int q = 0; q < N; ++q) xe.setValue(q, exact[q]);
    multiply(A, xe, b);
    const int its = A.solve(b, x);
    double err = 0., nrm = 0.;
    for (int q = 0; q < N; q += 7) { err = std::fmax(err, std::fabs(x(q) - exact[q])); nrm = std::fmax(nrm, std::fabs(exact[q])); }
    scale(-1., b);
    multiplyAdd(A, x, b, res);
    const double r = res.infNorm();
    const caffe_solve_report rep = A.lastReport();
    std::printf("drop_in sparse: %d iterations, relres %.3e, max error %.3e, ||A x - b||_inf %.3e\n", its, rep.relres,
                err / nrm, r);
    return (its > 0 && rep.converged && err / nrm < 1e-7 && r < 1e-7) ? 0 : 1;
}

static int pisoPart(const char* out) {
    const int n = 16;
    caffe_case c{};
    c.nI = c.nJ = c.nK = n;
    c.algo = CAFFE_PISO;
    c.steady = 0;
    c.rho = 1.;
    c.mu = 1e-3;
    c.alphaU = c.alphaP = 1.;
    c.nInner = 1;
    c.nPCorr = 2;
    for (int f = 0; f < 6; ++f) c.bcType[f] = CAFFE_BC_WALL;
    c.bcVec[3 * 2] = 1.;   // north lid (1, 0, 0)
    c.uniform = 1;
    std::vector<double> nodes;
    for (int k = 0; k <= n; ++k)
        for (int j = 0; j <= n; ++j)
            for (int i = 0; i <= n; ++i) { nodes.push_back(i * (1.0 / n)); nodes.push_back(j * (1.0 / n)); nodes.push_back(k * (1.0 / n)); }
    FlowSolver piso(c, nodes);
    caffe_solve_opts o{CAFFE_SOLVE_PLAIN_FP64, 1e-12, 200000, 1e-4};
    piso.setSolveOptions(0, o);
    piso.setSolveOptions(1, o);
    double e = 0.;
    for (int s = 0; s < 3; ++s) { e = piso.solve(0.5 / n); piso.displayUpdateMessage(); }
    FILE* f = std::fopen(out, "wb");
    if (!f) return 1;
    for (int q = 0; q < 3; ++q) {
        std::vector<double> u = piso.cellData("u", q, 3);
        std::fwrite(u.data(), sizeof(double), u.size(), f);
    }
    std::vector<double> p = piso.cellData("p");
    std::fwrite(p.data(), sizeof(double), p.size(), f);
    std::fclose(f);
    std::printf("drop_in piso: 3 steps, continuity %.3e\n", e);
    return 0;
}

int main(int argc, char** argv) {
    if (argc < 2) { std::fprintf(stderr, "usage: drop_in <output file>\n"); return 2; }
    try {
        if (sparsePart() != 0) { std::printf("drop_in sparse FAILED\n"); return 1; }
        std::printf("drop_in sparse ok\n");
        if (pisoPart(argv[1]) != 0) { std::printf("drop_in piso FAILED\n"); return 1; }
        std::printf("drop_in piso ok\n");
    } catch (const char* msg) {
        std::printf("drop_in error: %s\n", msg);
        return 1;
    }
    return 0;
}
