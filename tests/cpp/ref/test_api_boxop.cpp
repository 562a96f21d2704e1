// Proto's own operator headers compiled unchanged against this library's C++
// API: examples/_common/BoxOp_Laplace.H and tests/BoxOp_Test{Laplace,Flux,
// Source,Diag}.H from the reference tree (include path set by the Makefile;
// nothing of them is copied here).  Every result is compared bit for bit with
// what Proto's CPU path computed on the same inputs (tests/golden/boxop.bin,
// written by tests/golden/make_golden.py):
//   - LevelMultigridSolver<BoxOp_Laplace>: residual history and solution, on
//     the resident path and on the generic path (Proto's MGLevel recursion
//     over per-box BoxOp launches);
//   - LevelOp<BoxOp_TestLaplace>: the operator on a periodic level;
//   - BoxOp_TestFlux / _TestSource / _TestDiag: the operator and its fluxes.
#include <cstdint>
#include <fstream>
#include <map>
#include <string>

#include "Proto.H"
#include "check.H"

#include "BoxOp_Laplace.H"
#include "BoxOp_TestLaplace.H"
#include "BoxOp_TestFlux.H"
#include "BoxOp_TestSource.H"
#include "BoxOp_TestDiag.H"

using namespace Proto;

static std::map<std::string, std::vector<double>> g_gold;

static void loadGolden(const char* path)
{
    std::ifstream f(path, std::ios::binary);
    if (!f) { throw std::runtime_error(std::string("cannot open ") + path); }
    int32_t n = 0;
    f.read((char*)&n, 4);
    for (int i = 0; i < n; i++) {
        int32_t len = 0;
        f.read((char*)&len, 4);
        std::string name(len, '\0');
        f.read(&name[0], len);
        int64_t c = 0;
        f.read((char*)&c, 8);
        std::vector<double> v(c);
        f.read((char*)v.data(), 8 * c);
        g_gold[name] = v;
    }
    if (!f) { throw std::runtime_error("truncated golden file"); }
}

static const std::vector<double>& gold(const std::string& k)
{
    auto it = g_gold.find(k);
    if (it == g_gold.end()) { throw std::runtime_error("golden record missing: " + k); }
    return it->second;
}

static bool sameBits(const double* a, const double* b, size_t n)
{
    return std::memcmp(a, b, sizeof(double) * n) == 0;
}

template<unsigned int C>
static std::vector<double> toHost(const BoxData<double, C, DEVICE>& d)
{
    std::vector<double> h(d.size());
    proto_memcpy<DEVICE, HOST>(d.data(), h.data(), sizeof(double) * h.size());
    return h;
}

static DisjointBoxLayout periodicLayout(int n, int bs)
{
    ProblemDomain pd(Box::Cube(n), true);
    return DisjointBoxLayout(pd, Point::Ones(bs));
}

// global n^3 array (x fastest) <-> the valid cells of a level
static void fromGlobal(LevelBoxData<double>& L, const std::vector<double>& g, int n)
{
    for (auto it : L.layout()) {
        BoxData<double, 1, HOST> h(L[it].box());
        proto_memcpy<DEVICE, HOST>(L[it].data(), h.data(), sizeof(double) * h.size());
        const Box v = L.layout()[it];
        for (auto p = v.begin(); p != v.end(); ++p) {
            const Point q = *p;
            h(q) = g[q[0] + (size_t)n * (q[1] + (size_t)n * q[2])];
        }
        proto_memcpy<HOST, DEVICE>(h.data(), L[it].data(), sizeof(double) * h.size());
    }
}
static std::vector<double> toGlobal(const LevelBoxData<double>& L, int n)
{
    std::vector<double> g((size_t)n * n * n);
    for (auto it : L.layout()) {
        BoxData<double, 1, HOST> h(L[it].box());
        proto_memcpy<DEVICE, HOST>(L[it].data(), h.data(), sizeof(double) * h.size());
        const Box v = L.layout()[it];
        for (auto p = v.begin(); p != v.end(); ++p) {
            const Point q = *p;
            g[q[0] + (size_t)n * (q[1] + (size_t)n * q[2])] = h(q);
        }
    }
    return g;
}

// examples/LevelPoisson/exec/main.cpp:91-127 as a Proto user writes it
static void fasSolve(bool resident)
{
    const int n = 32;
    const double dx = 1.0 / n;
    DisjointBoxLayout layout = periodicLayout(n, 16);
    LevelBoxData<double> state(layout, BoxOp_Laplace<double>::ghost()), force(layout, Point::Zeros()),
        res(layout, Point::Zeros());
    state.setToZero();
    force.setToZero();
    fromGlobal(force, gold("fas:rhs"), n);
    LevelMultigridSolver<BoxOp_Laplace, double> mg(layout, Point::Ones(2), 4, dx);
    CHECK(mg.resident());  // the example's operator is recognised
    if (!resident) { mg.allowResident(false); }
    CHECK_EQ(mg.resident(), resident);
    const auto& hist = gold("fas:hist");
    const double r = mg.solve(state, force, 20, 1e-10);
    const auto& h = mg.history();
    std::printf("  %s path: %zu norms (golden %zu), last %.6e\n", resident ? "resident" : "generic", h.size(),
                hist.size(), r);
    CHECK_EQ(h.size(), hist.size());
    CHECK(h.size() == hist.size() && sameBits(h.data(), hist.data(), h.size()));
    CHECK_EQ(r, hist.back());
    const auto phi = toGlobal(state, n);
    CHECK(sameBits(phi.data(), gold("fas:phi").data(), phi.size()));
    // resnorm after the solve: Proto's last norm again, and m_residual filled
    const double r2 = mg.resnorm(res, state, force);
    CHECK_EQ(r2, hist.back());
    CHECK_EQ(res.absMax(), r2);
    CHECK_EQ(mg.res().absMax(), r2);
}

TEST(LevelMultigridSolverBoxOpLaplaceResident) { fasSolve(true); }
TEST(LevelMultigridSolverBoxOpLaplaceGeneric) { fasSolve(false); }

TEST(LevelMultigridSolverOtherOperatorIsGeneric)
{
    // BoxOp_TestLaplace has the same flux but no spectral radius of its own
    // (BoxOp's default 1), so Jacobi weights differ: not the resident operator
    const int n = 16;
    DisjointBoxLayout layout = periodicLayout(n, 8);
    LevelMultigridSolver<BoxOp_TestLaplace, double> mg(layout, Point::Ones(2), 2, 1.0 / n);
    CHECK(!mg.resident());
}

TEST(LevelOpTestLaplaceApply)
{
    const int n = 16;
    DisjointBoxLayout layout = periodicLayout(n, 8);
    Array<double, DIM> dx;
    dx.fill(1.0 / n);
    LevelOp<BoxOp_TestLaplace, double> op(layout, dx);
    CHECK(op.ghost() == Point::Ones(2));
    LevelBoxData<double> src(layout, BoxOp_TestLaplace<double>::ghost()), dst(layout, Point::Zeros());
    src.setToZero();
    fromGlobal(src, gold("levelop:phi"), n);
    op.applyBC(src);  // PeriodicBC: the exchange
    dst.setToZero();
    op(dst, src);
    const auto out = toGlobal(dst, n);
    CHECK(sameBits(out.data(), gold("levelop:out").data(), out.size()));
}

// tests/BoxOpTests.cpp:9-210 set-ups on seeded inputs
template<template<typename, MemType> class OP>
static void boxOpCase(int which)
{
    const int n = 16;
    const double dx = 1.0 / n;
    Box B0 = Box::Cube(n);
    std::array<bool, DIM> per;
    per.fill(false);
    ProblemDomain domain(B0, per);
    DisjointBoxLayout layout(domain, B0.sizes());
    OP<double, DEVICE> op(layout, *layout.begin(), dx);
    if (which == 2) { op.setDiagScale(31.0); } else { op.setFluxScale(31.0); }
    BoxData<double, 1> src(B0.grow(OP<double, DEVICE>::ghost())), aux(B0.grow(OP<double, DEVICE>::auxGhost())),
        dst(B0);
    proto_memcpy<HOST, DEVICE>(gold("boxop:src").data(), src.data(), sizeof(double) * src.size());
    proto_memcpy<HOST, DEVICE>(gold("boxop:aux").data(), aux.data(), sizeof(double) * aux.size());
    Array<BoxData<double, 1>, DIM> flx;
    for (int d = 0; d < DIM; d++) { flx[d].define(B0.grow((Centering)d)); }
    if (which == 0) {
        op(dst, flx, src, aux, 17.0);
    } else {
        op(dst, src, aux, 17.0);
    }
    const auto out = toHost(dst);
    CHECK(sameBits(out.data(), gold("boxop" + std::to_string(which) + ":dst").data(), out.size()));
    if (which == 0) {
        const auto& gf = gold("boxop0:flux");
        size_t off = 0;
        for (int d = 0; d < DIM; d++) {
            const auto f = toHost(flx[d]);
            CHECK(off + f.size() <= gf.size() && sameBits(f.data(), gf.data() + off, f.size()));
            off += f.size();
        }
    }
}

TEST(BoxOpTestFlux) { boxOpCase<BoxOp_TestFlux>(0); }
TEST(BoxOpTestSource) { boxOpCase<BoxOp_TestSource>(1); }
TEST(BoxOpTestDiag) { boxOpCase<BoxOp_TestDiag>(2); }

TEST(BoxOpDefaults)
{
    // BoxOp::define: d = 0, f = 1; spectral radius of BoxOp_Laplace
    DisjointBoxLayout layout = periodicLayout(16, 8);
    BoxOp_Laplace<double> op(layout, 0, 1.0 / 16);
    CHECK_EQ(op.diagScale(), 0.0);
    CHECK_EQ(op.fluxScale(), 1.0);
    CHECK_EQ(op.spectralRadius(), (2.0 * DIM) / std::pow(1.0 / 16, 2));
    CHECK(op.box() == layout[0]);
    CHECK_EQ(BoxOp_Laplace<double>::numState(), 1u);
    CHECK_EQ(BoxOp_Laplace<double>::numAux(), 0u);
    CHECK_EQ(BoxOp_Laplace<double>::order(), 4);
}

int main(int argc, char** argv)
{
    if (argc < 2) {
        std::printf("usage: %s tests/golden/boxop.bin\n", argv[0]);
        return 2;
    }
    loadGolden(argv[1]);
    return run_all();
}
