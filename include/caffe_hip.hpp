/*
 * caffe_hip.hpp -- C++ classes over the C ABI of caffe_hip.h, with the public shape of the
 * reference's classes, for a maintainer who swaps the GPU path into the reference's own
 * drivers and solvers:
 *
 *   caffe_gpu::SparseVector  src/Math/SparseVector.h:8-43   (allocate, setValue, operator(),
 *                                                            zeroEntries, l1/l2/infNorm)
 *   caffe_gpu::SparseMatrix  src/Math/SparseMatrix.h:34-81  (allocate, setValue, setRow,
 *                                                            zeroEntries, solve, assemble,
 *                                                            iLower/iUpper; multiply,
 *                                                            multiplyAdd, scale)
 *   caffe_gpu::FlowSolver    src/Solvers/Solver.h:12-39     (solve(dt), displayUpdateMessage)
 *                            for Simple / Piso / IbPiso / Diffusion / Coupled
 *
 * Header only; it adds no symbol of its own to libcaffe_hip.so.  Failures throw const char*,
 * as Output::raiseException does (src/RunControl/Output.cpp:61-65) and as the reference's
 * modules catch (Modules/caffePiso.cpp:46).  The message stays valid until the next call on
 * the same object.
 */
#ifndef CAFFE_HIP_HPP
#define CAFFE_HIP_HPP

#include <cstdio>
#include <vector>

#include "caffe_hip.h"

namespace caffe_gpu {

class SparseMatrix;

class SparseVector {
public:
    SparseVector() {}
    explicit SparseVector(int m, int device = 0) { allocate(m, device); }
    ~SparseVector() { deallocate(); }
    SparseVector(const SparseVector&) = delete;
    SparseVector& operator=(const SparseVector&) = delete;

    void allocate(int m, int device = 0) {
        deallocate();
        if (caffe_vector_create(device, m, &v_) != CAFFE_OK) throw "in \"SparseVector::allocate\", caffe_vector_create failed";
    }
    void deallocate() { if (v_) { caffe_vector_destroy(v_); v_ = nullptr; } }
    void setValue(int i, double value) { check(caffe_vector_set_values(v_, 1, &i, &value)); }
    void zeroEntries() { check(caffe_vector_zero_entries(v_)); }
    double operator()(int i) { double x = 0.; check(caffe_vector_get_values(v_, 1, &i, &x)); return x; }
    double l1Norm() { return norm(CAFFE_NORM_L1); }
    double l2Norm() { return norm(CAFFE_NORM_L2); }
    double infNorm() { return norm(CAFFE_NORM_INF); }
    caffe_vector* handle() const { return v_; }

private:
    caffe_vector* v_ = nullptr;
    void check(int rc) { if (rc != CAFFE_OK) throw caffe_vector_last_error(v_); }
    double norm(int t) { double x = 0.; check(caffe_vector_norm(v_, t, &x)); return x; }
    friend void scale(double alpha, SparseVector& v);
};

class SparseMatrix {
public:
    /* PETSC_DEFAULT (-2) for any argument keeps the PETSc default (rtol 1e-5, abstol 1e-50,
       10000 iterations), as SparseMatrix(double, double, int) does (SparseMatrix.cpp:26-33) */
    explicit SparseMatrix(double rToler = -2., double absToler = -2., int maxIters = -2, int device = 0)
        : rtol_(rToler), atol_(absToler), maxit_(maxIters), device_(device) {}
    SparseMatrix(int m, int n, int nnz) : SparseMatrix() { allocate(m, n, nnz); }
    ~SparseMatrix() { deallocate(); }
    SparseMatrix(const SparseMatrix&) = delete;
    SparseMatrix& operator=(const SparseMatrix&) = delete;

    void allocate(int m, int n, int nnz) {
        deallocate();
        if (caffe_matrix_create(device_, m, n, nnz, rtol_, atol_, maxit_, &A_) != CAFFE_OK)
            throw "in \"SparseMatrix::allocate\", caffe_matrix_create failed";
        m_ = m;
    }
    void deallocate() { if (A_) { caffe_matrix_destroy(A_); A_ = nullptr; } }
    void setValue(int i, int j, double value) { check(caffe_matrix_set_value(A_, i, j, value)); }
    void setRow(int rowNo, int nCols, int colNos[], double values[]) {
        const int rowptr[2] = {0, nCols};
        check(caffe_matrix_set_rows(A_, 1, &rowNo, rowptr, colNos, values));
    }
    void zeroEntries() { check(caffe_matrix_zero_entries(A_)); }
    int solve(const SparseVector& b, SparseVector& x) {
        int its = 0;
        check(caffe_matrix_solve(A_, b.handle(), x.handle(), &its));
        return its;
    }
    void assemble() { check(caffe_matrix_assemble(A_)); }
    int iLower() const { return 0; }       // one GPU owns every row
    int iUpper() const { return m_; }
    caffe_solve_report lastReport() { caffe_solve_report r{}; check(caffe_matrix_last_report(A_, &r)); return r; }
    caffe_matrix* handle() const { return A_; }

private:
    caffe_matrix* A_ = nullptr;
    double rtol_, atol_;
    int maxit_, device_, m_ = 0;
    void check(int rc) const { if (rc != CAFFE_OK) throw caffe_matrix_last_error(A_); }
    friend void multiply(const SparseMatrix&, const SparseVector&, SparseVector&);
    friend void multiplyAdd(const SparseMatrix&, const SparseVector&, const SparseVector&, SparseVector&);
    friend void scale(double alpha, SparseMatrix& A);
};

inline void multiply(const SparseMatrix& A, const SparseVector& x, SparseVector& b) {
    A.check(caffe_matrix_multiply(A.A_, x.handle(), b.handle()));
}
inline void multiplyAdd(const SparseMatrix& A, const SparseVector& x1, const SparseVector& x2, SparseVector& b) {
    A.check(caffe_matrix_multiply_add(A.A_, x1.handle(), x2.handle(), b.handle()));
}
inline void scale(double alpha, SparseMatrix& A) { A.check(caffe_matrix_scale(alpha, A.A_)); }
inline void scale(double alpha, SparseVector& v) { v.check(caffe_vector_scale(alpha, v.v_)); }

/* One GPU-resident Simple / Piso / IbPiso / Diffusion / Coupled instance (c.algo).  The
   reference constructs these from Input and HexaFvmMesh (Simple.cpp:34-74); here the caller
   fills caffe_case from the same case.info keys and passes the StructuredMesh nodes
   ((nI+1)(nJ+1)(nK+1) x 3, i fastest, xyz interleaved). */
class FlowSolver {
public:
    FlowSolver(const caffe_case& c, const std::vector<double>& nodes, int device = 0) {
        if (caffe_ctx_create(device, &ctx_) != CAFFE_OK) throw "in \"FlowSolver\", caffe_ctx_create failed";
        check(caffe_case_setup(ctx_, &c, nodes.data()));
        nI_ = c.nI; nJ_ = c.nJ; nK_ = c.nK;
    }
    ~FlowSolver() { caffe_ctx_destroy(ctx_); }
    FlowSolver(const FlowSolver&) = delete;
    FlowSolver& operator=(const FlowSolver&) = delete;

    void setSolveOptions(int which, const caffe_solve_opts& o) { check(caffe_set_solve_opts(ctx_, which, &o)); }
    /* Solver::solve(timeStep) (Solver.h:20): the continuity value */
    double solve(double timeStep) {
        double e = 0.;
        check(caffe_solve(ctx_, timeStep, &e));
        continuity_ = e;
        return e;
    }
    void displayUpdateMessage() {
        caffe_solve_report m{}, p{};
        check(caffe_last_reports(ctx_, &m, &p));
        std::printf("Solver: completed iteration.\nSolver: Krylov iterations (momentum, p'): %d, %d\n"
                    "Solver: Max continuity error : %f\n", m.iters, p.iters, continuity_);
    }
    /* cell values of a scalar field (or one component of a vector field) in Array3D order
       (i fastest), the halo of boundary patches stripped */
    std::vector<double> cellData(const char* name, int component = 0, int ncomp = 1) {
        const long P = caffe_padded_size(ctx_);
        std::vector<double> all(ncomp * P), out((size_t)nI_ * nJ_ * nK_);
        check(caffe_field_download(ctx_, name, all.data(), (long)all.size()));
        const long sJ = nI_ + 2, sK = (long)(nI_ + 2) * (nJ_ + 2);
        size_t q = 0;
        for (int k = 0; k < nK_; ++k)
            for (int j = 0; j < nJ_; ++j)
                for (int i = 0; i < nI_; ++i) out[q++] = all[component * P + (i + 1) + (j + 1) * sJ + (k + 1) * sK];
        return out;
    }
    caffe_ctx* handle() const { return ctx_; }

private:
    caffe_ctx* ctx_ = nullptr;
    int nI_ = 0, nJ_ = 0, nK_ = 0;
    double continuity_ = 0.;
    void check(int rc) { if (rc != CAFFE_OK) throw caffe_last_error(ctx_); }
};

}  // namespace caffe_gpu

#endif
