// forall, LevelBoxData, averageDown and LevelMultigrid through Proto's C++
// API: the checks of tests/ForallTests.cpp and tests/LevelBoxDataTests.cpp,
// and the archived LevelMultigrid driver written exactly as Proto users write
// it, compared with the solver of this library.
#include "Proto.H"
#include "check.H"

using namespace Proto;

PROTO_KERNEL_START
void sineF(const Point& p, Var<double>& v, double h)
{
    v(0) = 1.0;
    for (int d = 0; d < DIM; d++) { v(0) = v(0) * sin(2.0 * M_PI * (p[d] * h + 0.5 * h)); }
}
PROTO_KERNEL_END(sineF, sine)

PROTO_KERNEL_START
void pointIdF(const Point& p, Var<double>& v, int n)
{
    double id = 0, f = 1;
    for (int d = 0; d < DIM; d++) { id += (((p[d] % n) + n) % n) * f; f *= n; }
    v(0) = id;
}
PROTO_KERNEL_END(pointIdF, pointId)

template<typename T>
static std::vector<T> toHost(const BoxData<T, 1, DEVICE>& d)
{
    BoxData<T, 1, HOST> h(d.box());
    d.copyTo(h);
    return std::vector<T>(h.data(), h.data() + h.size());
}

TEST(ForallResidualMatchesStencil)
{
    // r = f - L(u) written with forall equals the same program's Stencil apply
    Box B = Box::Cube(20);
    BoxData<double> u(B.grow(1)), f(B);
    forallInPlace_p(sine, u, 1.0 / 20);
    forallInPlace_p(sine, f, 1.0 / 7);
    auto L = Stencil<double>::Laplacian_27();
    BoxData<double> Lu = L(u, 400.0);
    BoxData<double> r = forall<double>([] PROTO_LAMBDA(Var<double>& rv, Var<double>& fv, Var<double>& lv) {
        rv(0) = fv(0) - lv(0);
    }, f, Lu);
    CHECK(r.box() == B);
    auto hr = toHost(r), hf = toHost(f), hl = toHost(Lu);
    for (size_t i = 0; i < hr.size(); i++) { CHECK_EQ(hr[i], hf[i] - hl[i]); }
}

TEST(ForallPointAndScalar)
{
    // tests/ForallTests.cpp forallInPlace_p: the functor sees the point and a scalar
    Box B = Box(Point::Ones(-2), Point::Ones(5));
    BoxData<double> d(B);
    forallInPlace_p([] PROTO_LAMBDA(const Point& p, Var<double>& v, double s) { v(0) = s * p.sum(); }, d, 0.5);
    BoxData<double, 1, HOST> h(B);
    d.copyTo(h);
    for (auto it = B.begin(); it != B.end(); ++it) { CHECK_EQ(h(*it), 0.5 * (*it).sum()); }
}

TEST(ForallOpAndBoxArgument)
{
    // forallOp_p counts flops per point; forall with a Box argument runs on that box only
    Box B = Box(Point::Zeros(), Point::Ones(7));
    Box S = Box(Point::Ones(2), Point::Ones(4));
    BoxData<double> d(B);
    d.setVal(-1.0);
    const unsigned long long f0 = forallFlops();
    BoxData<double> a = forallOp_p<double>(3, "idx", [] PROTO_LAMBDA(const Point& p, Var<double>& v, Var<double>& src) {
        v(0) = src(0) + p[0];
    }, d);
    CHECK_EQ(forallFlops() - f0, 3ULL * B.size());
    CHECK(a.box() == B);
    BoxData<double> b = forall<double>([] PROTO_LAMBDA(Var<double>& v, Var<double>& src) { v(0) = 2.0 * src(0); }, S, a);
    CHECK(b.box() == S);
    forallInPlaceOp(1, "scale", [] PROTO_LAMBDA(Var<double>& v) { v(0) = v(0) + 1.0; }, S, d);
    BoxData<double, 1, HOST> ha(B), hb(S), hd(B);
    a.copyTo(ha);
    b.copyTo(hb);
    d.copyTo(hd);
    for (auto it = B.begin(); it != B.end(); ++it) {
        CHECK_EQ(ha(*it), -1.0 + (*it)[0]);
        CHECK_EQ(hd(*it), S.contains(*it) ? 0.0 : -1.0);
    }
    for (auto it = S.begin(); it != S.end(); ++it) { CHECK_EQ(hb(*it), 2.0 * (-1.0 + (*it)[0])); }
}

TEST(ExchangePeriodicPointId)
{
    // tests/LevelBoxDataTests.cpp:345-429: after exchange every ghost cell holds
    // the value of the cell it is a periodic image of
    const int n = 16;
    ProblemDomain pd(Box::Cube(n), true);
    DisjointBoxLayout layout(pd, Point::Ones(8));
    CHECK_EQ(layout.size(), 1 << DIM);
    LevelBoxData<double> data(layout, Point::Ones(2));
    data.setVal(-1.0);
    for (auto it : layout) { forallInPlace_p(pointId, layout[it], data[it], n); }
    data.exchange();
    for (auto it : layout) {
        BoxData<double, 1, HOST> h(data[it].box());
        data[it].copyTo(h);
        for (auto p = h.box().begin(); p != h.box().end(); ++p) {
            double id = 0, f = 1;
            for (int d = 0; d < DIM; d++) { id += ((((*p)[d] % n) + n) % n) * f; f *= n; }
            CHECK_EQ(h(*p), id);
        }
    }
}

TEST(LayoutMortonOrder)
{
    ProblemDomain pd(Box::Cube(32), true);
    DisjointBoxLayout layout(pd, Point::Ones(8));
    for (int i = 1; i < layout.size(); i++) { CHECK(mortonIndex(layout.point(i - 1)) < mortonIndex(layout.point(i))); }
    CHECK(layout.point(0) == Point::Zeros());
    CHECK(layout.point(1) == Point::Basis(0));
}

TEST(AverageDownLevel)
{
    ProblemDomain pd(Box::Cube(16), true);
    DisjointBoxLayout fineL(pd, Point::Ones(8));
    DisjointBoxLayout crseL(pd.coarsen(Point::Ones(2)), Point::Ones(8));
    LevelBoxData<double> fine(fineL, Point::Ones(1)), crse(crseL, Point::Zeros());
    fine.setVal(2.5);
    averageDown(crse, fine, 2);
    CHECK_EQ(crse.absMax(), 2.5);
    CHECK_EQ(crse.sum(), 2.5 * (1 << (3 * DIM)));
}

// The archived correction-scheme driver's resnorm (LevelMultigrid.cpp:75-97),
// written with this API as a Proto user writes it.
static double userResnorm(LevelBoxData<double>& phi, LevelBoxData<double>& rhs, double dx)
{
    phi.exchange();
    double hsqinv = 1. / (dx * dx);
    Reduction<double, Abs> rxn;
    auto L = Stencil<double>::Laplacian();
    for (auto dit : phi) {
        BoxData<double> res(phi.layout()[dit]);
        res.setVal(0.);
        res -= rhs[dit];
        res += L(phi[dit], hsqinv);
        res.reduce(rxn);
    }
    return rxn.fetch();
}

TEST(LevelMultigridSolve)
{
    const int n = 32;
    double dx = 1.0 / n;
    ProblemDomain pd(Box::Cube(n), true);
    DisjointBoxLayout layout(pd, Point::Ones(16));
    LevelBoxData<double> rho(layout, Point::Zeros()), phi(layout, Point::Ones());
    rho.setToZero();
    phi.setToZero();
    for (auto dit : layout) { forallInPlace_p(sine, rho[dit], dx); }
    for (int form = 0; form < 3; form++) {
        phi.setToZero();
        MultigridOptions opt = form == 1 ? MultigridOptions::Jacobi() : MultigridOptions::GSRB();
        if (form == 2) { opt.mixedPrecision(); }
        LevelMultigrid mg(layout, dx, 4, opt);
        double r0 = mg.resnorm(phi, rho);
        CHECK_EQ(r0, userResnorm(phi, rho, dx));
        std::vector<double> hist;
        int iters = mg.solve(phi, rho, 20, 1e-10, &hist);
        std::printf("  form %d: %d cycles, %.3e -> %.3e (hierarchy %zu bytes)\n", form, iters, hist.front(),
                    hist.back(), mg.bytes());
        CHECK(hist.back() < 1e-10 * hist.front());
        CHECK_EQ(hist.back(), userResnorm(phi, rho, dx));
    }
}

int main() { return run_all(); }
