// capi.hip -- the C ABI (include/caffe_hip.h): context, case setup, the
// Solver::solve orchestration of SIMPLE / PISO / IbPiso, and field exchange.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "internal.hpp"

using namespace caffe;

namespace {

template <class F>
int guard(caffe_ctx* x, F f) {
    if (!x) return CAFFE_ERR_ARG;
    try {
        return f(x->c);
    } catch (const std::string& e) {
        x->c.err = e;
        return CAFFE_ERR_HIP;
    } catch (const std::exception& e) {
        x->c.err = e.what();
        return CAFFE_ERR_HIP;
    }
}

double* dalloc(Ctx& c, size_t n) {
    void* p = nullptr;
    HIPCHECK(hipMalloc(&p, n * sizeof(double)));
    HIPCHECK(hipMemsetAsync(p, 0, n * sizeof(double), c.stream));
    c.allocs.push_back(p);
    c.deviceBytes += n * sizeof(double);
    return (double*)p;
}
template <class T>
T* talloc(Ctx& c, size_t n) {
    void* p = nullptr;
    HIPCHECK(hipMalloc(&p, n * sizeof(T)));
    HIPCHECK(hipMemsetAsync(p, 0, n * sizeof(T), c.stream));
    c.allocs.push_back(p);
    c.deviceBytes += n * sizeof(T);
    return (T*)p;
}

void freeAll(Ctx& c) {
    if (c.stream) (void)hipStreamSynchronize(c.stream);
    for (void* p : c.allocs) (void)hipFree(p);
    c.allocs.clear();
    c.deviceBytes = 0;
    c.wb32 = nullptr; c.Am32 = nullptr; c.wm32 = nullptr;
    c.ov = Overlay();
    if (c.mg.iter2) (void)hipGraphExecDestroy(c.mg.iter2);
    c.mg = MgHier();
    c.ready = false;
}

// FlowBoundaryConditions.cpp:44-71 mapping of flow BCs onto u/h and p types
void mapTypes(Ctx& c) {
    for (int f = 0; f < 6; ++f) {
        int t = c.cs.bcType[f];
        c.flowT[f] = t;
        if (t == F_INLET || t == F_WALL) { c.uT[f] = BC_FIXED; c.pT[f] = BC_ZERO_GRADIENT; }
        else if (t == F_OUTLET) { c.uT[f] = BC_ZERO_GRADIENT; c.pT[f] = BC_FIXED; }
        else if (t == F_EMPTY) { c.uT[f] = BC_EMPTY; c.pT[f] = BC_EMPTY; }
        else { c.uT[f] = BC_ZERO_GRADIENT; c.pT[f] = BC_ZERO_GRADIENT; }
        c.massCorr[f] = (t == F_OUTLET);  // SimpleBoundaryCondition.cpp:68-96
    }
}

// Initial patch values.  PrimitiveBoundaryCondition::changeType(FIXED) re-applies all six
// refValues_ each time (PrimitiveBoundaryCondition.tpp:257-294), and refValues_ not yet
// set hold T() (uninitialised in the reference; zero here).  After the east..bottom loop
// of FlowBoundaryConditions a patch therefore holds its refValue if some FIXED face was
// processed at or after it, and zero otherwise.
void initialPatch(const int* types, const double* refs, int ncomp, double* out /*6*ncomp*/) {
    int last = -1;
    for (int f = 0; f < 6; ++f) if (types[f] == BC_FIXED) last = f;
    for (int f = 0; f < 6; ++f)
        for (int q = 0; q < ncomp; ++q) out[f * ncomp + q] = (f <= last) ? refs[f * ncomp + q] : 0.;
}

void setHostPatch(const Dims& D, std::vector<double>& a, int face, double v) {
    int n1, n2;
    if (face < 2) { n1 = D.nJ; n2 = D.nK; } else if (face < 4) { n1 = D.nI; n2 = D.nK; } else { n1 = D.nI; n2 = D.nJ; }
    for (int y = 0; y < n2; ++y)
        for (int x = 0; x < n1; ++x) {
            long p;
            switch (face) {
            case 0: p = D.pidx(D.nI, x, y); break;
            case 1: p = D.pidx(-1, x, y); break;
            case 2: p = D.pidx(x, D.nJ, y); break;
            case 3: p = D.pidx(x, -1, y); break;
            case 4: p = D.pidx(x, y, D.nK); break;
            default: p = D.pidx(x, y, -1); break;
            }
            a[p] = v;
        }
}

void upload(Ctx& c, double* dst, const std::vector<double>& src) {
    HIPCHECK(hipMemcpyAsync(dst, src.data(), src.size() * sizeof(double), hipMemcpyHostToDevice, c.stream));
    HIPCHECK(hipStreamSynchronize(c.stream));
}

int setup(Ctx& c, const caffe_case* cs, const double* nodes) {
    if (!cs || !nodes) { c.err = "null case or nodes"; return CAFFE_ERR_ARG; }
    if (cs->nI < 1 || cs->nJ < 1 || cs->nK < 1) { c.err = "cell counts must be positive"; return CAFFE_ERR_ARG; }
    if (cs->algo < 0 || cs->algo > 4) { c.err = "unknown algorithm"; return CAFFE_ERR_ARG; }
    for (int f = 0; f < 6; ++f) {
        const int t = cs->bcType[f];
        const bool ok = cs->algo == CAFFE_DIFFUSION
            ? (t == CAFFE_BC_FIXED || t == CAFFE_BC_ZERO_GRADIENT || t == CAFFE_BC_EMPTY)   // PrimitiveBoundaryCondition.tpp:17-24
            : (t >= 0 && t <= 4);                                                          // FlowBoundaryConditions.cpp:31-42
        if (!ok) { c.err = "invalid boundary type"; return CAFFE_ERR_ARG; }
    }
    if (cs->alphaU == 0. || (!cs->steady && false)) { c.err = "relaxationFactorMomentum must be nonzero"; return CAFFE_ERR_ARG; }
    freeAll(c);
    c.cs = *cs;
    if (cs->algo == CAFFE_COUPLED) {
        // the Coupled constructor reads no initial conditions (Coupled.cpp:5-34)
        c.cs.uInit[0] = c.cs.uInit[1] = c.cs.uInit[2] = 0.;
        c.cs.pInit = 0.;
    }
    Dims& D = c.D;
    D.nI = cs->nI; D.nJ = cs->nJ; D.nK = cs->nK;
    D.sJ = D.nI + 2; D.sK = (long)(D.nI + 2) * (D.nJ + 2);
    D.P = D.sK * (D.nK + 2);
    D.N = (long)D.nI * D.nJ * D.nK;
    if (D.P > (long)INT32_MAX) { c.err = "mesh too large for 32-bit row ids"; return CAFFE_ERR_ARG; }
    mapTypes(c);
    c.hasIB = (cs->algo == CAFFE_IBPISO);
    c.singularP = !c.hasIB;
    c.indefiniteP = false;
    if (cs->algo == CAFFE_DIFFUSION) {
        // phi uses the p' storage and solver: singular when steady without a fixed face,
        // indefinite when unsteady (a0 = -sum a_nb + V/dt, Diffusion.cpp:60)
        for (int f = 0; f < 6; ++f) {
            const int t = cs->bcType[f];
            c.pT[f] = t == CAFFE_BC_FIXED ? BC_FIXED : (t == CAFFE_BC_EMPTY ? BC_EMPTY : BC_ZERO_GRADIENT);
            c.uT[f] = BC_FIXED;
            c.massCorr[f] = false;
        }
        c.singularP = cs->steady != 0;
        c.indefiniteP = cs->steady == 0;
    }
    for (int f = 0; f < 6; ++f) if (c.pT[f] == BC_FIXED) c.singularP = false;

    // ---- mesh
    const long nn = (long)(D.nI + 1) * (D.nJ + 1) * (D.nK + 1);
    c.uniform = cs->uniform != 0;
    Geo& g = c.geo;
    std::memset(&g, 0, sizeof(g));
    g.D = D;
    if (c.uniform) {
        // nodes must be the Cartesian box (StructuredMesh.cpp:131-148) with origin 0
        const double lx = nodes[3 * (long)D.nI], ly = nodes[3 * (long)D.nJ * (D.nI + 1) + 1];
        const double lz = nodes[3 * (long)D.nK * (D.nJ + 1) * (D.nI + 1) + 2];
        g.hx = lx / D.nI; g.hy = ly / D.nJ; g.hz = lz / D.nK;
        double worst = 0.;
        for (int k = 0; k <= D.nK; ++k)
            for (int j = 0; j <= D.nJ; ++j)
                for (int i = 0; i <= D.nI; ++i) {
                    long q = ((long)k * (D.nJ + 1) + j) * (D.nI + 1) + i;
                    worst = std::max(worst, std::fabs(nodes[3 * q] - i * g.hx) / g.hx);
                    worst = std::max(worst, std::fabs(nodes[3 * q + 1] - j * g.hy) / g.hy);
                    worst = std::max(worst, std::fabs(nodes[3 * q + 2] - k * g.hz) / g.hz);
                }
        if (!(worst <= 1e-9)) { c.err = "uniform=1 but the nodes are not a uniform box at the origin"; return CAFFE_ERR_ARG; }
    } else {
        c.dnodes = dalloc(c, 3 * nn);
        HIPCHECK(hipMemcpyAsync(c.dnodes, nodes, 3 * nn * sizeof(double), hipMemcpyHostToDevice, c.stream));
        double* m = dalloc(c, 28 * D.P);
        c.metric = m;
        long o = 0;
        auto nx = [&]() { double* r = m + o * D.P; ++o; return r; };
        g.vol = nx();
        for (int a = 0; a < 3; ++a) g.xc[a] = nx();
        for (int a = 0; a < 3; ++a) for (int b = 0; b < 3; ++b) g.S[a][b] = nx();
        for (int a = 0; a < 3; ++a) for (int b = 0; b < 3; ++b) g.fc[a][b] = nx();
        for (int a = 0; a < 3; ++a) g.dF[a] = nx();
        for (int a = 0; a < 3; ++a) g.gF[a] = nx();
        if (cs->nonOrth) {   // the corrected mode's c metric: 9 more arrays
            double* mc = dalloc(c, 9 * D.P);
            for (int a = 0; a < 3; ++a) for (int b = 0; b < 3; ++b) g.cF[a][b] = mc + (3 * a + b) * (long)D.P;
        }
        compute_metrics(c, c.dnodes);
    }

    // ---- fields
    for (int q = 0; q < 3; ++q) {
        c.u[q] = dalloc(c, D.P); c.u0[q] = dalloc(c, D.P); c.h[q] = dalloc(c, D.P);
        c.gp[q] = dalloc(c, D.P); c.gpc[q] = dalloc(c, D.P);
        c.mF[q] = dalloc(c, D.P); c.mF2[q] = dalloc(c, D.P);
    }
    for (int q = 0; q < 9; ++q) c.gu[q] = nullptr;
    // the corrected mode keeps grad u (Simple.cpp:355) for the next momentum assembly
    if (cs->nonOrth && !c.uniform) for (int q = 0; q < 9; ++q) c.gu[q] = dalloc(c, D.P);
    c.p = dalloc(c, D.P); c.pc = dalloc(c, D.P); c.d = dalloc(c, D.P); c.mflow = dalloc(c, D.P);
    c.st = talloc<int8_t>(c, D.P);
    {
        std::vector<int8_t> st(D.P, ST_INACTIVE);
        for (int k = 0; k < D.nK; ++k) for (int j = 0; j < D.nJ; ++j) for (int i = 0; i < D.nI; ++i) st[D.pidx(i, j, k)] = ST_ACTIVE;
        HIPCHECK(hipMemcpyAsync(c.st, st.data(), D.P, hipMemcpyHostToDevice, c.stream));
        HIPCHECK(hipStreamSynchronize(c.stream));
    }
    // systems and work
    c.Am = dalloc(c, 7 * D.P); c.bm = dalloc(c, 3 * D.P); c.xm = dalloc(c, 3 * D.P);
    c.Ap = dalloc(c, 4 * D.P); c.bp = dalloc(c, D.P); c.xp = dalloc(c, D.P);
    c.Ap32 = talloc<float>(c, 4 * D.P);
    c.wm = dalloc(c, 12 * 3 * D.P);
    c.wp = dalloc(c, 8 * D.P);
    c.wp32 = talloc<float>(c, 8 * D.P);
    if (c.hasIB) c.wb32 = talloc<float>(c, 11 * D.P);
    if (cs->algo == CAFFE_COUPLED) {
        c.cA = dalloc(c, (size_t)CP_PLANES * D.P);
        c.cb = dalloc(c, 4 * (size_t)D.P);
        c.cx = dalloc(c, 4 * (size_t)D.P);
        c.cw = dalloc(c, 52 * (size_t)D.P);
        // identity diagonals until assembled
        launch_fill(c, c.cA + (size_t)CP_M * D.P, 1.0, D.P);
        launch_fill(c, c.cA + (size_t)CP_L * D.P, 1.0, D.P);
    }
    // identity rows everywhere until assembled (keeps the Jacobi diagonal finite)
    launch_fill(c, c.Am, 1.0, D.P);
    launch_fill(c, c.Ap, 1.0, D.P);

    // ---- boundary and initial conditions (Simple.cpp:52-73)
    double uref[18], pref[6], zero6[18] = {0};
    for (int f = 0; f < 6; ++f) { for (int q = 0; q < 3; ++q) uref[3 * f + q] = cs->bcVec[3 * f + q]; pref[f] = cs->bcVal[f]; }
    double up[18], pp[6];
    initialPatch(c.uT, uref, 3, up);
    initialPatch(c.pT, pref, 1, pp);
    // the Diffusion solver's boundary condition sets all six patches to their refValue once
    // (PrimitiveBoundaryCondition.tpp:3-30)
    if (cs->algo == CAFFE_DIFFUSION) for (int f = 0; f < 6; ++f) pp[f] = pref[f];
    std::vector<double> host(D.P);
    for (int q = 0; q < 3; ++q) {
        std::fill(host.begin(), host.end(), 0.);
        for (int k = 0; k < D.nK; ++k) for (int j = 0; j < D.nJ; ++j) for (int i = 0; i < D.nI; ++i) host[D.pidx(i, j, k)] = c.cs.uInit[q];
        for (int f = 0; f < 6; ++f) setHostPatch(D, host, f, up[3 * f + q]);
        upload(c, c.u[q], host);
        // h: same types and references as u, interior zero
        for (int k = 0; k < D.nK; ++k) for (int j = 0; j < D.nJ; ++j) for (int i = 0; i < D.nI; ++i) host[D.pidx(i, j, k)] = 0.;
        upload(c, c.h[q], host);
    }
    std::fill(host.begin(), host.end(), 0.);
    for (int k = 0; k < D.nK; ++k) for (int j = 0; j < D.nJ; ++j) for (int i = 0; i < D.nI; ++i) host[D.pidx(i, j, k)] = c.cs.pInit;
    for (int f = 0; f < 6; ++f) setHostPatch(D, host, f, pp[f]);
    upload(c, c.p, host);
    (void)zero6;

    // ---- reductions, scalars, events
    c.red.part = dalloc(c, (size_t)3 * MAX_RED_BLOCKS * RED_STRIDE + 4096);
    c.scal = dalloc(c, N_SCAL);
    c.flags = talloc<int>(c, N_FLAGS);
    if (!c.hscal) {
        HIPCHECK(hipHostMalloc((void**)&c.hscal, N_SCAL * sizeof(double), hipHostMallocDefault));
        HIPCHECK(hipHostMalloc((void**)&c.hflags, N_FLAGS * sizeof(int), hipHostMallocDefault));
    }
    c.momOpts = caffe_solve_opts{CAFFE_SOLVE_PLAIN_FP64, 1e-8, 2000, 0.};
    c.pcOpts = caffe_solve_opts{CAFFE_SOLVE_PLAIN_FP64, 1e-8, 50000, 1e-4};
    HIPCHECK(hipStreamSynchronize(c.stream));
    c.ready = true;
    c.ibReady = false;
    return CAFFE_OK;
}

// ---- the steps of Simple / Piso / IbPiso
int saveU0(Ctx& c) {
    for (int q = 0; q < 3; ++q)
        HIPCHECK(hipMemcpyAsync(c.u0[q], c.u[q], c.D.P * sizeof(double), hipMemcpyDeviceToDevice, c.stream));
    return CAFFE_OK;
}

int momentumSolve(Ctx& c, caffe_solve_report* rep) {
    int rc = momentum_solve(c, c.momOpts, &c.lastMom);
    if (rep) *rep = c.lastMom;
    launch_mom_copyback(c);
    // Simple.cpp:213-218: u, d, h boundaries, then d and h faces (recomputed on use)
    double* uh[6] = {c.u[0], c.u[1], c.u[2], c.h[0], c.h[1], c.h[2]};
    launch_halo(c, uh, 6, c.uT);
    int zg[6] = {BC_ZERO_GRADIENT, BC_ZERO_GRADIENT, BC_ZERO_GRADIENT, BC_ZERO_GRADIENT, BC_ZERO_GRADIENT, BC_ZERO_GRADIENT};
    double* dd[1] = {c.d};
    launch_halo(c, dd, 1, zg);
    return rc == CAFFE_ERR_NOT_CONVERGED ? CAFFE_OK : rc;
}

int pcorrSolve(Ctx& c, const caffe_solve_opts& o, caffe_solve_report* rep) {
    // inside a step (no report asked for) the solve's time is read when the report is
    int rc = pcorr_solve(c, o, &c.lastPc, rep == nullptr);
    if (rep) *rep = c.lastPc;
    launch_pc_copyback(c);
    double* f[1] = {c.pc};
    launch_halo(c, f, 1, c.pT);
    double* g[3] = {c.gpc[0], c.gpc[1], c.gpc[2]};
    launch_grad_scalar(c, c.pc, g);
    return rc == CAFFE_ERR_NOT_CONVERGED ? CAFFE_OK : rc;
}

int correctStep(Ctx& c) {
    launch_correct(c);
    double* pf[1] = {c.p};
    launch_halo(c, pf, 1, c.pT);
    double* uf[3] = {c.u[0], c.u[1], c.u[2]};
    launch_halo(c, uf, 3, c.uT);
    double* g[3] = {c.gp[0], c.gp[1], c.gp[2]};
    launch_grad_scalar(c, c.p, g);
    if (c.gu[0]) launch_grad_vector(c);  // gradUField_ (Simple.cpp:355) when a caller keeps it
    return CAFFE_OK;
}

int ibPrepare(Ctx& c) {
    if (c.ibReady) {
        launch_set_inactive_fields(c);  // IbPiso.cpp:74-80 runs every step
        return CAFFE_OK;
    }
    int rc = ib_field(c);
    if (rc != CAFFE_OK) return rc;
    rc = ib_coeffs(c);
    if (rc != CAFFE_OK) return rc;
    c.ibReady = true;
    return CAFFE_OK;
}

// Diffusion::solve (Diffusion.cpp:33-131): assemble, solve (from x = 0), phi = x, the
// boundary patches, the extrapolated-face gradient; returns ||A x - b||_inf
int diffusionStep(Ctx& c, double dt, double* resInf) {
    launch_diff_assemble(c, dt);
    int rc = pcorr_solve(c, c.pcOpts, &c.lastPc);
    if (rc != CAFFE_OK && rc != CAFFE_ERR_NOT_CONVERGED) return rc;
    launch_copy_cells(c, c.xp, c.p);
    double* f[1] = {c.p};
    launch_halo(c, f, 1, c.pT);
    double* g1[3] = {c.gpc[0], c.gpc[1], c.gpc[2]};
    double* g[3] = {c.gp[0], c.gp[1], c.gp[2]};
    launch_grad_scalar(c, c.p, g1);
    launch_grad_extrap(c, c.p, g1, g);
    const double r = diff_resid_inf(c, c.p);
    if (resInf) *resInf = r;
    return CAFFE_OK;
}

// Coupled::solve (Coupled.cpp:38-50): u0 = u, then numberOfInnerIterations (1 when steady)
// passes of computeMomentumAndPressure (Coupled.cpp:54-280: assemble the 4-variable system,
// solve it from x = 0, update u, p, grad p, h and the Rhie-Chow faces); returns the
// continuity error of Coupled.cpp:316-338
int coupledStep(Ctx& c, double dt, double* continuity) {
    saveU0(c);
    const int nInner = c.cs.steady ? 1 : c.cs.nInner;
    for (int it = 0; it < nInner; ++it) {
        launch_cp_assemble(c, dt);
        const int rc = coupled_solve(c, c.pcOpts, &c.lastPc);
        if (rc != CAFFE_OK && rc != CAFFE_ERR_NOT_CONVERGED) return rc;
        launch_cp_update(c, c.cx);
    }
    const double e = cp_continuity(c);
    if (continuity) *continuity = e;
    return CAFFE_OK;
}

int solveStep(Ctx& c, double dt, double* continuity) {
    if (!c.ready) { c.err = "caffe_case_setup has not been called"; return CAFFE_ERR_STATE; }
    if (c.cs.algo == CAFFE_DIFFUSION) return diffusionStep(c, dt, continuity);
    if (c.cs.algo == CAFFE_COUPLED) return coupledStep(c, dt, continuity);
    saveU0(c);
    int rc;
    if (c.hasIB && (rc = ibPrepare(c)) != CAFFE_OK) return rc;
    const int nInner = c.cs.steady ? 1 : c.cs.nInner;
    const int nPC = c.cs.algo == CAFFE_SIMPLE ? 1 : c.cs.nPCorr;
    for (int it = 0; it < nInner; ++it) {
        launch_mom_assemble(c, dt);
        if ((rc = momentumSolve(c, nullptr)) != CAFFE_OK) return rc;
        launch_rhie_chow(c);
        for (int q = 0; q < nPC; ++q) {
            launch_pc_assemble(c);
            if ((rc = pcorrSolve(c, c.pcOpts, nullptr)) != CAFFE_OK) return rc;
            correctStep(c);
        }
    }
    // Piso.cpp:26 / Simple.cpp:88: the Courant number's reduction is queued ahead of the
    // continuity scan, whose read-back brings both values
    const bool co = c.cs.algo != CAFFE_IBPISO;
    if (co) courant_queue(c, dt);
    double e = continuity_quirk(c);
    if (co) c.courant = courant_queued(c);
    if (continuity) *continuity = e;
    return CAFFE_OK;
}

struct FieldRef { double* ptr[9]; int nc; };

bool fieldByName(Ctx& c, const std::string& n, FieldRef& f) {
    f.nc = 0;
    auto set3 = [&](double* const* a) { for (int q = 0; q < 3; ++q) f.ptr[q] = a[q]; f.nc = 3; };
    if (n == "u") set3(c.u);
    else if (n == "u0") set3(c.u0);
    else if (n == "h") set3(c.h);
    else if (n == "gradP") set3(c.gp);
    else if (n == "gradPCorr") set3(c.gpc);
    else if (n == "mI" || n == "mJ" || n == "mK") { f.ptr[0] = c.mF[n == "mI" ? 0 : (n == "mJ" ? 1 : 2)]; f.nc = 1; }
    else if (n == "p" || n == "phi") { f.ptr[0] = c.p; f.nc = 1; }
    else if (n == "gradPhi") set3(c.gp);
    else if (n == "pCorr") { f.ptr[0] = c.pc; f.nc = 1; }
    else if (n == "d") { f.ptr[0] = c.d; f.nc = 1; }
    else if (n == "massFlow") { f.ptr[0] = c.mflow; f.nc = 1; }
    else if (n == "gradU") {
        if (!c.gu[0]) return false;
        for (int q = 0; q < 9; ++q) f.ptr[q] = c.gu[q];
        f.nc = 9;
    }
    return f.nc > 0;
}

}  // namespace

namespace caffe {
float* falloc(Ctx& c, size_t n) { return talloc<float>(c, n); }
double* dalloc64(Ctx& c, size_t n) { return talloc<double>(c, n); }
}  // namespace caffe

extern "C" {

int caffe_ctx_create(int device, caffe_ctx** out) {
    if (!out) return CAFFE_ERR_ARG;
    *out = nullptr;
    caffe_ctx* x = new caffe_ctx();
    int rc = guard(x, [&](Ctx& c) {
        c.device = device;
        HIPCHECK(hipSetDevice(device));
        HIPCHECK(hipStreamCreateWithFlags(&c.stream, hipStreamNonBlocking));
        HIPCHECK(hipEventCreate(&c.ev0));
        HIPCHECK(hipEventCreate(&c.ev1));
        HIPCHECK(hipEventCreateWithFlags(&c.evR, hipEventDisableTiming));
        HIPCHECK(hipEventCreate(&c.evP0));
        HIPCHECK(hipEventCreate(&c.evP1));
        return CAFFE_OK;
    });
    if (rc != CAFFE_OK) {
        delete x;
        return rc;
    }
    *out = x;
    return CAFFE_OK;
}

int caffe_ctx_destroy(caffe_ctx* x) {
    if (!x) return CAFFE_ERR_ARG;
    Ctx& c = x->c;
    (void)hipSetDevice(c.device);
    freeAll(c);
    if (c.hscal) (void)hipHostFree(c.hscal);
    if (c.hflags) (void)hipHostFree(c.hflags);
    for (auto e : c.probe.ev) (void)hipEventDestroy(e);
    if (c.ev0) (void)hipEventDestroy(c.ev0);
    if (c.ev1) (void)hipEventDestroy(c.ev1);
    for (hipEvent_t e : {c.evR, c.evP0, c.evP1}) if (e) (void)hipEventDestroy(e);
    if (c.stream) (void)hipStreamDestroy(c.stream);
    delete x;
    return CAFFE_OK;
}

const char* caffe_last_error(caffe_ctx* x) { return x ? x->c.err.c_str() : "null context"; }

int caffe_case_setup(caffe_ctx* x, const caffe_case* cs, const double* nodes) {
    return guard(x, [&](Ctx& c) { HIPCHECK(hipSetDevice(c.device)); return setup(c, cs, nodes); });
}

int caffe_set_solve_opts(caffe_ctx* x, int which, const caffe_solve_opts* o) {
    return guard(x, [&](Ctx& c) {
        if (!o || which < 0 || which > 1 || !(o->rtol > 0.) || o->maxit < 1) { c.err = "invalid solve options"; return CAFFE_ERR_ARG; }
        if (o->mode < CAFFE_SOLVE_PLAIN_FP64 || o->mode > CAFFE_SOLVE_SGS_FP64) { c.err = "unknown solve mode"; return CAFFE_ERR_ARG; }
        if (which == 0 && o->mode != CAFFE_SOLVE_PLAIN_FP64 && o->mode != CAFFE_SOLVE_IR_FP32) {
            c.err = "momentum solves are FP64 or FP32 inside FP64 refinement";
            return CAFFE_ERR_ARG;
        }
        (which == 0 ? c.momOpts : c.pcOpts) = *o;
        return CAFFE_OK;
    });
}

int caffe_solve(caffe_ctx* x, double dt, double* continuity) {
    return guard(x, [&](Ctx& c) { return solveStep(c, dt, continuity); });
}

#define READY(c) if (!c.ready) { c.err = "caffe_case_setup has not been called"; return CAFFE_ERR_STATE; }

int caffe_save_u0(caffe_ctx* x) { return guard(x, [&](Ctx& c) { READY(c); return saveU0(c); }); }
int caffe_momentum_assemble(caffe_ctx* x, double dt) {
    return guard(x, [&](Ctx& c) { READY(c); launch_mom_assemble(c, dt); return CAFFE_OK; });
}
int caffe_momentum_solve(caffe_ctx* x, caffe_solve_report* rep) {
    return guard(x, [&](Ctx& c) { READY(c); return momentumSolve(c, rep); });
}
int caffe_rhiechow(caffe_ctx* x) { return guard(x, [&](Ctx& c) { READY(c); launch_rhie_chow(c); return CAFFE_OK; }); }
int caffe_pcorr_assemble(caffe_ctx* x) { return guard(x, [&](Ctx& c) { READY(c); launch_pc_assemble(c); return CAFFE_OK; }); }
int caffe_pcorr_solve(caffe_ctx* x, const caffe_solve_opts* o, caffe_solve_report* rep) {
    return guard(x, [&](Ctx& c) {
        READY(c);
        if (o && (o->mode < CAFFE_SOLVE_PLAIN_FP64 || o->mode > CAFFE_SOLVE_SGS_FP64)) { c.err = "unknown solve mode"; return CAFFE_ERR_ARG; }
        return pcorrSolve(c, o ? *o : c.pcOpts, rep);
    });
}
int caffe_correct(caffe_ctx* x) { return guard(x, [&](Ctx& c) { READY(c); return correctStep(c); }); }
int caffe_ib_field(caffe_ctx* x) {
    return guard(x, [&](Ctx& c) { READY(c); if (!c.hasIB) { c.err = "not an IbPiso case"; return CAFFE_ERR_STATE; } c.ibReady = false; return ib_field(c); });
}
int caffe_ib_coeffs(caffe_ctx* x) {
    return guard(x, [&](Ctx& c) { READY(c); if (!c.hasIB) { c.err = "not an IbPiso case"; return CAFFE_ERR_STATE; } int rc = ib_coeffs(c); if (rc == CAFFE_OK) c.ibReady = true; return rc; });
}
int caffe_grad_u(caffe_ctx* x) {
    return guard(x, [&](Ctx& c) {
        READY(c);
        if (!c.gu[0]) for (int q = 0; q < 9; ++q) c.gu[q] = dalloc(c, c.D.P);
        launch_grad_vector(c);
        return CAFFE_OK;
    });
}
int caffe_step_report(caffe_ctx* x, double dt, double* cref, double* cmax, double* co) {
    return guard(x, [&](Ctx& c) {
        READY(c);
        if (cref) *cref = continuity_quirk(c);
        if (cmax) *cmax = continuity_max(c);
        if (co) *co = courant(c, dt);
        return CAFFE_OK;
    });
}
int caffe_last_courant(caffe_ctx* x, double* co) {
    return guard(x, [&](Ctx& c) {
        if (!co) { c.err = "null argument"; return CAFFE_ERR_ARG; }
        *co = c.courant;
        return CAFFE_OK;
    });
}
int caffe_last_reports(caffe_ctx* x, caffe_solve_report* m, caffe_solve_report* p) {
    return guard(x, [&](Ctx& c) {
        pcorr_resolve_ms(c);
        if (m) *m = c.lastMom;
        if (p) *p = c.lastPc;
        return CAFFE_OK;
    });
}

int caffe_field_download(caffe_ctx* x, const char* name, double* host, long n) {
    return guard(x, [&](Ctx& c) {
        READY(c);
        if (!name || !host) { c.err = "null argument"; return CAFFE_ERR_ARG; }
        const long P = c.D.P;
        std::string nm(name);
        if (nm == "status") {
            if (n != P) { c.err = "size mismatch"; return CAFFE_ERR_ARG; }
            std::vector<int8_t> st(P);
            HIPCHECK(hipMemcpyAsync(st.data(), c.st, P, hipMemcpyDeviceToHost, c.stream));
            HIPCHECK(hipStreamSynchronize(c.stream));
            for (long q = 0; q < P; ++q) host[q] = st[q];
            return CAFFE_OK;
        }
        FieldRef f;
        if (!fieldByName(c, nm, f)) { c.err = "unknown field " + nm; return CAFFE_ERR_ARG; }
        if (n != f.nc * P) { c.err = "size mismatch for " + nm; return CAFFE_ERR_ARG; }
        for (int q = 0; q < f.nc; ++q)
            HIPCHECK(hipMemcpyAsync(host + q * P, f.ptr[q], P * sizeof(double), hipMemcpyDeviceToHost, c.stream));
        HIPCHECK(hipStreamSynchronize(c.stream));
        return CAFFE_OK;
    });
}

int caffe_field_upload(caffe_ctx* x, const char* name, const double* host, long n) {
    return guard(x, [&](Ctx& c) {
        READY(c);
        if (!name || !host) { c.err = "null argument"; return CAFFE_ERR_ARG; }
        FieldRef f;
        std::string nm(name);
        if (!fieldByName(c, nm, f)) { c.err = "unknown field " + nm; return CAFFE_ERR_ARG; }
        const long P = c.D.P;
        if (n != f.nc * P) { c.err = "size mismatch for " + nm; return CAFFE_ERR_ARG; }
        for (int q = 0; q < f.nc; ++q)
            HIPCHECK(hipMemcpyAsync(f.ptr[q], host + q * P, P * sizeof(double), hipMemcpyHostToDevice, c.stream));
        HIPCHECK(hipStreamSynchronize(c.stream));
        return CAFFE_OK;
    });
}

int caffe_metric_download(caffe_ctx* x, const char* name, double* host, long n) {
    return guard(x, [&](Ctx& c) {
        READY(c);
        if (c.uniform) { c.err = "uniform meshes keep no metric arrays"; return CAFFE_ERR_STATE; }
        std::string nm(name ? name : "");
        const Geo& g = c.geo;
        std::vector<const double*> src;
        if (nm == "vol") src = {g.vol};
        else if (nm == "xc") src = {g.xc[0], g.xc[1], g.xc[2]};
        else if (nm == "SI" || nm == "SJ" || nm == "SK") { int a = nm == "SI" ? 0 : (nm == "SJ" ? 1 : 2); src = {g.S[a][0], g.S[a][1], g.S[a][2]}; }
        else if (nm == "fcI" || nm == "fcJ" || nm == "fcK") { int a = nm == "fcI" ? 0 : (nm == "fcJ" ? 1 : 2); src = {g.fc[a][0], g.fc[a][1], g.fc[a][2]}; }
        else if (nm == "dI" || nm == "dJ" || nm == "dK") src = {g.dF[nm == "dI" ? 0 : (nm == "dJ" ? 1 : 2)]};
        else if (nm == "gI" || nm == "gJ" || nm == "gK") src = {g.gF[nm == "gI" ? 0 : (nm == "gJ" ? 1 : 2)]};
        else { c.err = "unknown metric " + nm; return CAFFE_ERR_ARG; }
        const long P = c.D.P;
        if (n != (long)src.size() * P) { c.err = "size mismatch"; return CAFFE_ERR_ARG; }
        for (size_t q = 0; q < src.size(); ++q)
            HIPCHECK(hipMemcpyAsync(host + q * P, src[q], P * sizeof(double), hipMemcpyDeviceToHost, c.stream));
        HIPCHECK(hipStreamSynchronize(c.stream));
        return CAFFE_OK;
    });
}

int caffe_matrix_download(caffe_ctx* x, int which, double* host, long n) {
    return guard(x, [&](Ctx& c) {
        READY(c);
        const long P = c.D.P;
        const double* src = which == 0 ? c.Am : (which == 1 ? c.Ap : c.cA);
        long need = which == 0 ? 7 * P : (which == 1 ? 4 * P : (long)CP_PLANES * P);
        if (which < 0 || which > 2 || (which == 2 && !c.cA) || !host || n != need) {
            c.err = "bad matrix download arguments";
            return CAFFE_ERR_ARG;
        }
        HIPCHECK(hipMemcpyAsync(host, src, need * sizeof(double), hipMemcpyDeviceToHost, c.stream));
        HIPCHECK(hipStreamSynchronize(c.stream));
        return CAFFE_OK;
    });
}

int caffe_rhs_download(caffe_ctx* x, int which, double* host, long n) {
    return guard(x, [&](Ctx& c) {
        READY(c);
        const long P = c.D.P;
        const double* src = which == 0 ? c.bm : (which == 1 ? c.bp : c.cb);
        long need = which == 0 ? 3 * P : (which == 1 ? P : 4 * P);
        if (which < 0 || which > 2 || (which == 2 && !c.cb) || !host || n != need) {
            c.err = "bad rhs download arguments";
            return CAFFE_ERR_ARG;
        }
        HIPCHECK(hipMemcpyAsync(host, src, need * sizeof(double), hipMemcpyDeviceToHost, c.stream));
        HIPCHECK(hipStreamSynchronize(c.stream));
        return CAFFE_OK;
    });
}

int caffe_ghost_rows(caffe_ctx* x, int* n, long* cells, long* cols, double* vm, double* vp) {
    return guard(x, [&](Ctx& c) {
        READY(c);
        if (!n) { c.err = "null count"; return CAFFE_ERR_ARG; }
        const int m = c.ov.n;
        *n = m;
        if (m > 0) {
            std::vector<int> ce(m), co(9 * (size_t)m);
            HIPCHECK(hipMemcpyAsync(ce.data(), c.ov.cell, sizeof(int) * m, hipMemcpyDeviceToHost, c.stream));
            HIPCHECK(hipMemcpyAsync(co.data(), c.ov.cols, sizeof(int) * 9 * m, hipMemcpyDeviceToHost, c.stream));
            HIPCHECK(hipStreamSynchronize(c.stream));
            if (cells) for (int q = 0; q < m; ++q) cells[q] = ce[q];
            if (cols) for (int q = 0; q < 9 * m; ++q) cols[q] = co[q];
            if (vm) HIPCHECK(hipMemcpyAsync(vm, c.ov.valM, sizeof(double) * 9 * m, hipMemcpyDeviceToHost, c.stream));
            if (vp) HIPCHECK(hipMemcpyAsync(vp, c.ov.valP, sizeof(double) * 9 * m, hipMemcpyDeviceToHost, c.stream));
        }
        HIPCHECK(hipStreamSynchronize(c.stream));
        return CAFFE_OK;
    });
}

int caffe_time_kernel(caffe_ctx* x, const char* which, int reps, double* ms, double* bytes) {
    return guard(x, [&](Ctx& c) {
        READY(c);
        std::string w(which ? which : "");
        if (reps < 1 || !ms) { c.err = "bad timing arguments"; return CAFFE_ERR_ARG; }
        if (w == "copy_fp64") {  // calibration: copy 8 * P doubles (read + write)
            const int n = 8 * c.D.P;
            *ms = time_copy(c, c.wm, c.wm + 12 * c.D.P, n, reps);
            if (bytes) *bytes = 2.0 * 8.0 * n;
            return CAFFE_OK;
        }
        const bool f32 = (w == "pcorr_cg_k1_fp32");
        if (!f32 && w != "pcorr_cg_k1_fp64") { c.err = "unknown kernel " + w; return CAFFE_ERR_ARG; }
        *ms = time_cg_k1(c, f32, reps);
        // algorithmic bytes per interior cell: aP, aI, aJ, aK + z + p_old read, p_new + Ap written
        if (bytes) *bytes = (double)c.D.N * (f32 ? 4.0 : 8.0) * 8.0;
        return CAFFE_OK;
    });
}

int caffe_kernel_probe(caffe_ctx* x, const char* which, int enable) {
    return guard(x, [&](Ctx& c) {
        READY(c);
        std::string w(which ? which : "");
        int kind = PROBE_NONE;
        if (w == "fcg_k1") kind = PROBE_FCG_K1;
        else if (w == "pcorr_cg_k1_fp32") kind = PROBE_CG_K1_FP32;
        else if (w == "pcorr_cg_k1_fp64") kind = PROBE_CG_K1_FP64;
        else if (w == "coupled_bicg_a") kind = PROBE_CP_A;
        else if (w == "pcorr_bicg_a_fp32") kind = PROBE_PBI_A_FP32;
        else if (w == "pcorr_bicg_a_fp64") kind = PROBE_PBI_A_FP64;
        else if (w == "mg_post0") kind = PROBE_MG_POST0;
        else if (w == "fcg_k2pre") kind = PROBE_FCG_K2PRE;
        else { c.err = "unknown probe kernel " + w; return CAFFE_ERR_ARG; }
        if (kind == PROBE_CP_A && !c.cA) { c.err = "not a Coupled case"; return CAFFE_ERR_STATE; }
        HIPCHECK(hipStreamSynchronize(c.stream));
        if (enable && c.probe.ev.empty()) {
            c.probe.ev.resize(2 * PROBE_MAX);
            // timing-only events: no system-scope fence, so the probe neither adds an L2
            // writeback to the probed kernel's time nor slows the work that follows
            for (auto& e : c.probe.ev) HIPCHECK(hipEventCreateWithFlags(&e, hipEventDisableSystemFence));
        }
        c.probe.which = enable ? kind : PROBE_NONE;
        c.probe.used = 0;
        return CAFFE_OK;
    });
}

int caffe_kernel_probe_read(caffe_ctx* x, double* avg_ms, long* launches, double* bytes) {
    return guard(x, [&](Ctx& c) {
        READY(c);
        if (!avg_ms || !launches) { c.err = "null argument"; return CAFFE_ERR_ARG; }
        HIPCHECK(hipStreamSynchronize(c.stream));
        std::vector<double> d;
        for (int i = 0; i < c.probe.used; ++i) {
            float ms = 0.f;
            HIPCHECK(hipEventElapsedTime(&ms, c.probe.ev[2 * i], c.probe.ev[2 * i + 1]));
            d.push_back(ms);
        }
        // launches queued past convergence return at once: count those that ran
        std::vector<double> srt = d;
        std::sort(srt.begin(), srt.end());
        const double med = srt.empty() ? 0. : srt[srt.size() / 2];
        double sum = 0.;
        long n = 0;
        for (double v : d)
            if (v >= 0.5 * med) { sum += v; ++n; }
        *avg_ms = n ? sum / n : 0.;
        *launches = n;
        // algorithmic bytes per launch per interior cell: the stencil kernels read the
        // search direction's two inputs (z, p_old) and the four matrix arrays and write p
        // and q: 8 x 4 B in FP32, 8 x 8 B in FP64
        // The Coupled BiCGStab kernel KA reads the 56 coefficient planes, r, p_old, v_old, rh
        // and D^-1 of the four variables and writes p and v: 84 doubles per cell.  The p'
        // BiCGStab KA reads the four matrix arrays, r, rh, p_old, v_old and D^-1 and writes
        // p and v: 11 values per cell (ghost rows, about 0.1% of the cells, not counted).
        // The multigrid post-smoothing sweep on the fine level reads the four FP32 matrix
        // arrays, the FP64 right-hand side and the FP32 iterate and writes the new iterate:
        // 32 B per cell.  The learned FCG's update and pre-smoothing kernel (k_f_k2pre) reads
        // e, p, r, q and the four FP32 matrix arrays and writes e, the next r and the
        // pre-smoothed w: 44 B per cell (r and q at the stencil points come from cache).
        double per = 32.;
        switch (c.probe.which) {
        case PROBE_CP_A: per = 672.; break;
        case PROBE_CG_K1_FP64: per = 64.; break;
        case PROBE_PBI_A_FP32: per = 44.; break;
        case PROBE_PBI_A_FP64: per = 88.; break;
        case PROBE_MG_POST0: per = 32.; break;
        case PROBE_FCG_K2PRE: per = 44.; break;
        default: break;
        }
        if (bytes) *bytes = per * (double)c.D.N;
        return CAFFE_OK;
    });
}

int caffe_ctx_set_fcg_coarse(caffe_ctx* x, int mode) {
    return guard(x, [&](Ctx& c) {
        if (mode != 0 && mode != 1) { c.err = "coarse solve: 0 (the network) or 1 (a multigrid V-cycle)"; return CAFFE_ERR_ARG; }
        c.fcgCoarse = mode;
        return CAFFE_OK;
    });
}

int caffe_ctx_set_fcg_guess(caffe_ctx* x, int on) {
    return guard(x, [&](Ctx& c) {
        if (on != 0 && on != 1) { c.err = "initial guess of the learned solve: 0 (x = 0) or 1"; return CAFFE_ERR_ARG; }
        c.fcgGuess = on;
        if (!on) for (auto& s : c.fcgPrevGen) s[0] = s[1] = -1;
        return CAFFE_OK;
    });
}

int caffe_ctx_set_fcg_post_sweeps(caffe_ctx* x, int sweeps) {
    return guard(x, [&](Ctx& c) {
        if (sweeps < 1 || sweeps > 4) { c.err = "post-smoothing sweeps of the learned cycle: 1 to 4"; return CAFFE_ERR_ARG; }
        c.fcgPost = sweeps;
        return CAFFE_OK;
    });
}

int caffe_sync(caffe_ctx* x) {
    return guard(x, [&](Ctx& c) { HIPCHECK(hipStreamSynchronize(c.stream)); return CAFFE_OK; });
}

long caffe_padded_size(caffe_ctx* x) { return x ? x->c.D.P : -1; }

long caffe_device_bytes(caffe_ctx* x) { return x ? (long)x->c.deviceBytes : -1; }

}  // extern "C"
