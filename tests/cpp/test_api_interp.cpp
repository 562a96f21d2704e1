// InterpStencil::Quadratic and InterpStencil::FiniteVolume through the C++ API
// on the device, against the outputs Proto wrote for the same inputs
// (tests/golden/interp_hi.bin, path in argv[1]): per case int32 [kind, order,
// ratio, replace, clo[3], chi[3], flo[3], fhi[3]], then float64 crse, fine0
// and the expected fine.
#include <cstdio>
#include <cstring>

#include "Proto.H"
#include "check.H"

using namespace Proto;

static const char* g_path = nullptr;

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

int main(int argc, char** argv)
{
    g_path = argc > 1 ? argv[1] : "tests/golden/interp_hi.bin";
    return run_all();
}
