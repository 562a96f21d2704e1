// Prints the coefficient lists of this library's Stencil<T> library (the C++
// mirror of Proto_StencilDefs.H) in the form of the golden tables that Proto
// wrote (tests/golden/stencil_lib*.json): one line per stencil,
//   key|c0,c1,...|o0,o1,...|sr|dr|ds
// with coefficients as C99 hex floats and offsets as 3 ints each.  Host only:
// building a stencil touches no device.
#include <cstdio>
#include <string>

#include "Proto.H"

using namespace Proto;

template<typename T>
static void dump(const std::string& key, const Stencil<T>& S)
{
    std::printf("%s|", key.c_str());
    for (size_t j = 0; j < S.coefs().size(); j++) { std::printf("%s%a", j ? "," : "", (double)S.coefs()[j]); }
    std::printf("|");
    for (size_t j = 0; j < S.offsets().size(); j++) {
        for (int d = 0; d < 3; d++) { std::printf("%s%d", (j || d) ? "," : "", d < DIM ? S.offsets()[j][d] : 0); }
    }
    auto pt = [](const Point& p, int fill) {
        std::printf("|%d,%d,%d", p[0], DIM > 1 ? p[1] : fill, DIM > 2 ? p[2] : fill);
    };
    pt(S.srcRatio(), 1);
    pt(S.destRatio(), 1);
    pt(S.destShift(), 0);
    std::printf("\n");
}

int main()
{
    const std::string D = std::to_string(DIM) + "D/";
    auto k3 = [&](const char* n, int a, int b, int c) {
        return D + n + "/" + std::to_string(a) + "_" + std::to_string(b) + "_" + std::to_string(c);
    };
    auto k1 = [&](const char* n, int a, const char* t) { return D + n + "/" + std::to_string(a) + "/" + t; };
    // the table of tests/golden/stencil_lib.json
    dump(k1("Laplacian", 0, "f64"), Stencil<double>::Laplacian());
    dump(k1("Laplacian", 0, "f32"), Stencil<float>::Laplacian());
#if DIM == 3
    dump(k1("Laplacian_27", 0, "f64"), Stencil<double>::Laplacian_27());
    dump(k1("Laplacian_27", 0, "f32"), Stencil<float>::Laplacian_27());
    dump(k1("Laplacian_19", 0, "f64"), Stencil<double>::Laplacian_19());
    dump(k1("Laplacian_19", 0, "f32"), Stencil<float>::Laplacian_19());
#else
    dump(k1("Laplacian_9", 0, "f64"), Stencil<double>::Laplacian_9());
    dump(k1("Laplacian_9", 0, "f32"), Stencil<float>::Laplacian_9());
#endif
    dump(k1("AvgDown", 2, "f64"), Stencil<double>::AvgDown(2));
    dump(k1("AvgDown", 2, "f32"), Stencil<float>::AvgDown(2));
    dump(k1("Sum", 2, "f64"), Stencil<double>::Sum(2));
    dump(k1("Sum", 2, "f32"), Stencil<float>::Sum(2));
    for (int d = 0; d < DIM; d++) {
        dump(k1("FluxDivergence", d, "f64"), Stencil<double>::FluxDivergence(d));
        dump(k1("FluxDivergence", d, "f32"), Stencil<float>::FluxDivergence(d));
        dump(k1("DiffCellToFaceLo", d, "f64"), Stencil<double>::DiffCellToFace(d, Side::Lo));
        dump(k1("DiffCellToFaceHi", d, "f64"), Stencil<double>::DiffCellToFace(d, Side::Hi));
    }
    // the table of tests/golden/stencil_lib_ext.json
    dump(k3("Identity", 0, 0, 0), Stencil<double>::Identity());
    for (int d = 0; d < DIM; d++) {
        dump(k3("LaplacianFace", d, 2, 0), Stencil<double>::LaplacianFace(d, 2));
        for (int s = 0; s < 2; s++) {
            const Side::LoHiSide side = s ? Side::Hi : Side::Lo;
            for (int o : {2, 4, 5}) { dump(k3("CellToFace", d, s, o), Stencil<double>::CellToFace(d, side, o)); }
            for (int o : {2, 4}) { dump(k3("DiffCellToFace", d, s, o), Stencil<double>::DiffCellToFace(d, side, o)); }
            for (int r : {2, 4}) { dump(k3("AvgDownFace", d, s, r), Stencil<double>::AvgDownFace(d, side, r)); }
        }
        for (int o : {2, 4, 5}) {
            dump(k3("CellToFaceL", d, o, 0), Stencil<double>::CellToFaceL(d, o));
            dump(k3("CellToFaceH", d, o, 0), Stencil<double>::CellToFaceH(d, o));
        }
        for (int o : {2, 4}) {
            dump(k3("faceToCell", d, o, 0), Stencil<double>::faceToCell(d, o));
            dump(k3("cornersToFaces", d, o, 0), Stencil<double>::cornersToFaces(d, o));
        }
        const int no[][2] = {{1, 1}, {1, 2}, {1, 4}, {1, 6}, {1, 12}, {2, 2}, {2, 4}, {2, 6}, {2, 12}, {3, 2},
                             {3, 4}, {4, 2}, {4, 4}, {5, 2}, {6, 4}, {7, 2}, {12, 2}};
        for (auto& q : no) { dump(k3("Derivative", q[0], d, q[1]), Stencil<double>::Derivative(q[0], d, q[1])); }
    }
    for (int o : {2, 4}) { dump(k3("CornersToCells", o, 0, 0), Stencil<double>::CornersToCells(o)); }
    return 0;
}
