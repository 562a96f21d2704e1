// InterpStencil::Quadratic and InterpStencil::FiniteVolume through the C++ API
// on the device, against the outputs Proto wrote for the same inputs
// (tests/golden/interp_hi.bin, path in argv[1]): per case int32 [kind, order,
// ratio, replace, clo[3], chi[3], flo[3], fhi[3]], then float64 crse, fine0
// and the expected fine.  And interpBoundaries on a partially refined level
// against Proto's result (tests/golden/interp_bnd.bin, path in argv[2]).
#include <cstdio>
#include <cstring>

#include "Proto.H"
#include "check.H"

using namespace Proto;

static const char* g_path = nullptr;
static const char* g_bnd = nullptr;

static Box mk(const int* lo, const int* hi) { return Box(Point(lo[0], lo[1], lo[2]), Point(hi[0], hi[1], hi[2])); }

TEST(QuadraticAndFiniteVolumeMatchProto)
{
    FILE* f = std::fopen(g_path, "rb");
    CHECK(f != nullptr);
    if (!f) { return; }
    int ncase = 0;
    int head[16];
    while (std::fread(head, sizeof(int), 16, f) == 16) {
        const int kind = head[0], order = head[1], ratio = head[2], replace = head[3];
        const Box cb = mk(head + 4, head + 7), fb = mk(head + 10, head + 13);
        std::vector<double> crse(cb.size()), fine0(fb.size()), want(fb.size());
        CHECK(std::fread(crse.data(), sizeof(double), crse.size(), f) == crse.size());
        CHECK(std::fread(fine0.data(), sizeof(double), fine0.size(), f) == fine0.size());
        CHECK(std::fread(want.data(), sizeof(double), want.size(), f) == want.size());
        BoxData<double, 1, HOST> hc(crse.data(), cb), hf(fine0.data(), fb);
        BoxData<double> dc(cb), df(fb);
        hc.copyTo(dc);
        hf.copyTo(df);
        InterpStencil<double> I = kind == 2 ? InterpStencil<double>::Quadratic(Point::Ones(ratio))
                                            : InterpStencil<double>::FiniteVolume(Point::Ones(ratio), order);
        if (replace) { df |= I(dc); } else { df += I(dc); }
        BoxData<double, 1, HOST> got(fb);
        df.copyTo(got);
        const bool same = std::memcmp(got.data(), want.data(), sizeof(double) * want.size()) == 0;
        if (!same) { std::printf("  case kind %d order %d ratio %d replace %d differs\n", kind, order, ratio, replace); }
        CHECK(same);
        ncase++;
    }
    std::fclose(f);
    CHECK_EQ(ncase, 14);
}

// coarse periodic 16^3 in 8^3 boxes (ghost 1), ratio-2 fine level over patches
// {1,2,3}^3 of the 32^3 domain (ghost 2), every ghost starting at -7; valid
// cells from the fixture's global arrays; every fine box compared whole
TEST(InterpBoundariesMatchesProto)
{
    FILE* f = std::fopen(g_bnd, "rb");
    CHECK(f != nullptr);
    if (!f) { return; }
    int ncase = 0;
    CHECK(std::fread(&ncase, sizeof(int), 1, f) == 1);
    std::vector<double> cg(16 * 16 * 16), fg(32 * 32 * 32);
    CHECK(std::fread(cg.data(), sizeof(double), cg.size(), f) == cg.size());
    CHECK(std::fread(fg.data(), sizeof(double), fg.size(), f) == fg.size());
    ProblemDomain cpd(Box::Cube(16), true), fpd(Box::Cube(32), true);
    DisjointBoxLayout cl(cpd, Point::Ones(8));
    std::vector<Point> patches;
    const Box pb(Point::Ones(1), Point::Ones(3));
    for (auto it = pb.begin(); it != pb.end(); ++it) { patches.push_back(*it); }
    DisjointBoxLayout fl(fpd, patches, Point::Ones(8));
    CHECK(fl.onLevelBoundary(Point::Ones(1)) && !fl.onLevelBoundary(Point::Ones(2)));
    // whole ghosted boxes: valid cells from g, ghosts -7
    auto fill = [](LevelBoxData<double>& L, const std::vector<double>& g, int n) {
        for (auto it : L.layout()) {
            const Box v = L.layout()[it], gb = L[it].box();
            BoxData<double, 1, HOST> h(gb);
            for (auto p = gb.begin(); p != gb.end(); ++p) {
                const Point q = *p;
                h(q) = v.contains(q) ? g[q[0] + (long)n * (q[1] + (long)n * q[2])] : -7.0;
            }
            h.copyTo(L[it]);
        }
    };
    int done = 0;
    for (int k = 0; k < ncase; k++) {
        int head[2];
        CHECK(std::fread(head, sizeof(int), 2, f) == 2);
        const int kind = head[0], nb = head[1];
        std::vector<int> lows(3 * nb);
        std::vector<double> want((size_t)nb * 12 * 12 * 12);
        CHECK(std::fread(lows.data(), sizeof(int), lows.size(), f) == lows.size());
        CHECK(std::fread(want.data(), sizeof(double), want.size(), f) == want.size());
        LevelBoxData<double> crse(cl, Point::Ones(1)), fine(fl, Point::Ones(2));
        fill(crse, cg, 16);
        fill(fine, fg, 32);
        InterpStencil<double> I = kind == 0 ? InterpStencil<double>::Constant(Point::Ones(2))
                                : (kind == 1 ? InterpStencil<double>::Linear(Point::Ones(2))
                                             : InterpStencil<double>::Quadratic(Point::Ones(2)));
        interpBoundaries(crse, fine, I);
        CHECK_EQ(fl.size(), nb);
        bool same = true;
        for (auto it : fl) {
            for (int d = 0; d < 3; d++) { same = same && fl[it].low()[d] == lows[3 * it + d]; }
            BoxData<double, 1, HOST> got(fine[it].box());
            fine[it].copyTo(got);
            same = same && std::memcmp(got.data(), want.data() + (size_t)it * 1728, 1728 * sizeof(double)) == 0;
        }
        if (!same) { std::printf("  interpBoundaries kind %d differs\n", kind); }
        CHECK(same);
        done++;
    }
    std::fclose(f);
    CHECK_EQ(done, 3);
}

int main(int argc, char** argv)
{
    g_path = argc > 1 ? argv[1] : "tests/golden/interp_hi.bin";
    g_bnd = argc > 2 ? argv[2] : "tests/golden/interp_bnd.bin";
    return run_all();
}
