// Stencil / BoxData / Reduction / InterpStencil through Proto's C++ API on the
// device: the checks of tests/StencilTests.cpp, ReductionTests.cpp and
// InterpTests.cpp rewritten for this API.
#include "Proto.H"
#include "check.H"

using namespace Proto;

template<typename T, unsigned int C>
static std::vector<T> toHost(const BoxData<T, C, DEVICE>& d)
{
    BoxData<T, C, HOST> h(d.box());
    d.copyTo(h);
    return std::vector<T>(h.data(), h.data() + h.size());
}

template<typename T>
static void fillHost(BoxData<T, 1, DEVICE>& d, const std::function<T(const Point&)>& f)
{
    BoxData<T, 1, HOST> h(d.box());
    for (auto it = d.box().begin(); it != d.box().end(); ++it) { h(*it) = f(*it); }
    h.copyTo(d);
}

TEST(LaplacianOfParaboloid)
{
    // examples/_archived/Multigrid/exec/Multigrid.cpp:233-240: L(sum x^2) = 2 DIM
    Box B = Box::Cube(12);
    BoxData<double> src(B.grow(1)), dst(B);
    fillHost<double>(src, [](const Point& p) { double s = 0; for (int d = 0; d < DIM; d++) s += p[d] * p[d]; return s; });
    auto L = Stencil<double>::Laplacian();
    CHECK_EQ(L.kernelKind(), 7);
    dst |= L(src);
    for (double v : toHost(dst)) { CHECK_EQ(v, 2.0 * DIM); }
    auto L27 = Stencil<double>::Laplacian_27();
    CHECK_EQ(L27.kernelKind(), 27);
    dst |= L27(src);
    for (double v : toHost(dst)) { CHECK_NEAR(v, 2.0 * DIM, 1e-12); }
}

TEST(ApplyForms)
{
    // tests/StencilTests.cpp:292-335: L(src), L(src, box), |= and += agree bit for bit
    Box B = Box::Cube(16).shift(Point::Ones(-3));
    BoxData<double> src(B.grow(1));
    fillHost<double>(src, [](const Point& p) { return std::sin(0.3 * p[0]) + std::cos(0.7 * p[1] + 0.1 * p[2]); });
    auto L = Stencil<double>::Laplacian();
    BoxData<double> a = L(src, 17.0);
    BoxData<double> b(B), c(B);
    b |= L(src, 17.0);
    c.setVal(0);
    c += L(src, 17.0);
    CHECK(a.box() == B);
    CHECK(toHost(a) == toHost(b));
    CHECK(toHost(b) == toHost(c));
    // += accumulates: twice the stencil is (c) + (c), up to the order of additions
    c += L(src, 17.0);
    auto hc = toHost(c), hb = toHost(b);
    for (size_t i = 0; i < hc.size(); i++) { CHECK_NEAR(hc[i], 2 * hb[i], 1e-9 * std::abs(hb[i]) + 1e-12); }
}

TEST(BoxInference)
{
    auto L = Stencil<double>::Laplacian();
    Box B = Box::Cube(8);
    CHECK(L.indexRange(B.grow(1)) == B);
    CHECK(L.range(B.grow(1)) == B);
    CHECK(L.domain(B) == B.grow(1));
    auto A = Stencil<double>::AvgDown(2);
    CHECK(A.indexRange(Box::Cube(16)) == Box::Cube(8));
    CHECK(A.range(Box::Cube(16)) == Box::Cube(8));
}

TEST(AverageDownAndSum)
{
    Box F = Box::Cube(8);
    BoxData<double> fine(F, 3.25);
    auto A = Stencil<double>::AvgDown(2);
    BoxData<double> crse = A(fine);
    CHECK(crse.box() == Box::Cube(4));
    for (double v : toHost(crse)) { CHECK_EQ(v, 3.25); }
    auto S = Stencil<double>::Sum(2);
    BoxData<double> s = S(fine);
    for (double v : toHost(s)) { CHECK_EQ(v, 3.25 * (1 << DIM)); }
}

TEST(StencilAlgebra)
{
    // addCoef merges repeated offsets in first-insertion order
    Stencil<double> S = 1.0 * Shift::Zeros() + 2.0 * Shift::Basis(0, 1) + 0.5 * Shift::Zeros();
    CHECK_EQ(S.size(), 2u);
    CHECK_EQ(S.coefs()[0], 1.5);
    CHECK(S.offsets()[1] == Point::Basis(0, 1));
    // composition: (1 + shift) * (1 - shift) = 1 - shift^2 (shift cancels)
    Stencil<double> P = 1.0 * Shift::Zeros() + 1.0 * Shift::Basis(0);
    Stencil<double> M = 1.0 * Shift::Zeros() - 1.0 * Shift::Basis(0);
    Stencil<double> Q = P * M;
    CHECK_EQ(Q.size(), 3u);
    CHECK_EQ(Q.coefs()[1], 0.0);
    CHECK_EQ(Q.coefs()[2], -1.0);
    // Laplacian_27's FP64 weights do not sum to 0
    auto L27 = Stencil<double>::Laplacian_27();
    double sum = 0;
    for (double c : L27.coefs()) { sum += c; }
    CHECK(sum != 0.0 && std::abs(sum) < 1e-15);
    BoxData<double> one(Box::Cube(6).grow(1), 1.0), out(Box::Cube(6));
    out |= L27(one);
    for (double v : toHost(out)) { CHECK_EQ(v, sum); }
}

TEST(FloatStencil)
{
    Box B = Box::Cube(10);
    BoxData<float> src(B.grow(1)), dst(B);
    fillHost<float>(src, [](const Point& p) { return (float)(p[0] * p[0] + p[1]); });
    auto L = Stencil<float>::Laplacian();
    dst |= L(src, 2.0f);
    for (float v : toHost(dst)) { CHECK_EQ(v, 4.0f); }
}

TEST(Reductions)
{
    // tests/ReductionTests.cpp:17-31: data +-ii, closed-form answers
    const int n = 64;
    Box B(Point::Ones(4));
    BoxData<double, 1, HOST> h(B);
    int ii = 0;
    for (auto it = B.begin(); it != B.end(); ++it, ++ii) { h(*it) = (ii % 2 == 0) ? ii : -ii; }
    BoxData<double> d(B);
    h.copyTo(d);
    long sumAbs = 0, sum = 0, maxv = 0, minv = 0;
    for (int k = 0; k < (int)B.size(); k++) {
        long v = (k % 2 == 0) ? k : -k;
        sumAbs += std::labs(v);
        sum += v;
        maxv = std::max(maxv, v);
        minv = std::min(minv, v);
    }
    CHECK_EQ(d.sum(), (double)sum);
    CHECK_EQ(d.absMax(), (double)std::max(std::labs(maxv), std::labs(minv)));
    CHECK_EQ(d.max(), (double)maxv);
    CHECK_EQ(d.min(), (double)minv);
    Reduction<double, SumAbs> r;
    d.reduce(r);
    d.reduce(r);  // folds into the running value
    CHECK_EQ(r.fetch(), 2.0 * sumAbs);
    (void)n;
    // Max of negative data is numeric_limits<double>::min() (Proto_ReductionImplem.H:68)
    BoxData<double> neg(B, -5.0);
    CHECK_EQ(neg.max(), std::numeric_limits<double>::min());
}

TEST(InterpConstantAndLinear)
{
    // tests/InterpTests.cpp:42-113 (Constant is exact injection); Linear of a
    // linear function is exact with Proto's vertex-style weights
    Box C = Box::Cube(5).grow(1);
    BoxData<double> crse(C);
    fillHost<double>(crse, [](const Point& p) { return 1.0 + p[0] + 2.0 * p[1] - 3.0 * (DIM > 2 ? p[DIM - 1] : 0); });
    BoxData<double> fine(Box::Cube(10));
    auto I = InterpStencil<double>::Constant(Point::Ones(2));
    fine |= I(crse);
    BoxData<double, 1, HOST> hf(fine.box()), hc(crse.box());
    fine.copyTo(hf);
    crse.copyTo(hc);
    for (auto it = fine.box().begin(); it != fine.box().end(); ++it) {
        Point p = *it;
        Point q = Point::Zeros();
        for (int d = 0; d < DIM; d++) { q[d] = Box::floordiv(p[d], 2); }
        CHECK_EQ(hf(p), hc(q));
    }
    auto Lin = InterpStencil<double>::Linear(Point::Ones(2));
    fine.setVal(0);
    fine += Lin(crse);
    fine.copyTo(hf);
    for (auto it = fine.box().begin(); it != fine.box().end(); ++it) {
        Point p = *it;
        double x[DIM];
        for (int d = 0; d < DIM; d++) { x[d] = p[d] / 2.0; }
        double want = 1.0 + x[0] + 2.0 * x[1] - 3.0 * (DIM > 2 ? x[DIM - 1] : 0);
        CHECK_NEAR(hf(p), want, 1e-12);
    }
}

TEST(CopyToWithShiftAndIncrement)
{
    Box B = Box::Cube(6);
    BoxData<double> a(B, 2.0), b(B.shift(Point::Ones(3)), 0.0);
    a.copyTo(b, B, Point::Ones(3));
    for (double v : toHost(b)) { CHECK_EQ(v, 2.0); }
    BoxData<double> c(B, 1.0);
    c.increment(a, 0.25);
    for (double v : toHost(c)) { CHECK_EQ(v, 1.5); }
    c *= 4.0;
    c -= 1.0;
    for (double v : toHost(c)) { CHECK_EQ(v, 5.0); }
}

int main() { return run_all(); }
