// TEST INFRASTRUCTURE ONLY: the compiled-reference checker.
//
// This translation unit is built by oracle/Makefile against Proto's own headers
// where they lie in the reference tree (nothing of the reference is copied
// here).  It exposes Proto's CPU path through a small extern "C" surface so that
// tests/golden/make_golden.py can write golden vectors and so that the CPU tests
// can cross-check the C restatement in oracle/cpu_ref.c against the real thing.
// Only tests/, the smoke() check and bench.py's cpu_baseline leg may load it.
//
// Every entry point below calls Proto code; the only logic of our own is the
// copying of plain arrays in and out of Proto's BoxData / LevelBoxData, and the
// GSRB multigrid class, which composes Proto Stencils exactly as
// examples/_archived/AMRFAS_old/src/LaplaceOp.H:44-123 does (that file itself
// needs Chombo and does not compile) inside the correction-scheme V-cycle of
// examples/_archived/LevelMultigrid/src/LevelMultigrid.cpp:51-160.
#include "Proto.H"
// Proto is header-only with non-inline globals, so the archived solver's source
// file is compiled inside this translation unit, from where it lies.
#include "LevelMultigrid.cpp"
#include "BoxOp_Laplace.H"
// the test operators of Proto's own BoxOp / LevelOp tests (tests/BoxOp_Test*.H)
#include "BoxOp_TestLaplace.H"
#include "BoxOp_TestFlux.H"
#include "BoxOp_TestSource.H"
#include "BoxOp_TestDiag.H"

#include <cstring>
#include <vector>
#include <memory>

using namespace Proto;

namespace {

Point mkpt(const int* v)
{
    int t[DIM];
    for (int d = 0; d < DIM; d++) { t[d] = v[d]; }
    return Point(t);
}

Box mkbox(const int* lo, const int* hi) { return Box(mkpt(lo), mkpt(hi)); }

template<typename T>
Stencil<T> mkstencil(int n, const T* coefs, const int* offs,
                     const int* sr, const int* dr, const int* ds)
{
    Stencil<T> S;
    for (int j = 0; j < n; j++)
    {
        S += Stencil<T>(Shift(mkpt(offs + 3 * j)), coefs[j]);
    }
    S.srcRatio() = mkpt(sr);
    S.destRatio() = mkpt(dr);
    S.destShift() = mkpt(ds);
    return S;
}

template<typename T>
int dump_stencil(const Stencil<T>& S, T* coefs, int* offs, int* sr, int* dr, int* ds, int maxn)
{
    int n = S.size();
    if (n > maxn) { return -n; }
    for (int j = 0; j < n; j++)
    {
        coefs[j] = S.coefs()[j];
        for (int d = 0; d < 3; d++) { offs[3 * j + d] = (d < DIM) ? S.offsets()[j][d] : 0; }
    }
    for (int d = 0; d < 3; d++)
    {
        sr[d] = (d < DIM) ? S.srcRatio()[d] : 1;
        dr[d] = (d < DIM) ? S.destRatio()[d] : 1;
        ds[d] = (d < DIM) ? S.destShift()[d] : 0;
    }
    return n;
}

template<typename T>
int lib_stencil(const char* name, int arg, T* coefs, int* offs, int* sr, int* dr, int* ds, int maxn)
{
    Stencil<T> S;
    std::string nm(name);
    if (nm == "Laplacian") { S = Stencil<T>::Laplacian(); }
#if DIM == 3
    else if (nm == "Laplacian_27") { S = Stencil<T>::Laplacian_27(); }
    else if (nm == "Laplacian_19") { S = Stencil<T>::Laplacian_19(); }
#endif
#if DIM == 2
    else if (nm == "Laplacian_9") { S = Stencil<T>::Laplacian_9(); }
#endif
    else if (nm == "AvgDown") { S = Stencil<T>::AvgDown(arg); }
    else if (nm == "Sum") { S = Stencil<T>::Sum(arg); }
    else if (nm == "FluxDivergence") { S = Stencil<T>::FluxDivergence(arg); }
    else { return 0; }
    return dump_stencil(S, coefs, offs, sr, dr, ds, maxn);
}

// DiffCellToFace only instantiates for double (its order-2 branch builds a
// Stencil<double> whatever T is, Proto_StencilDefs.H:252).
int lib_stencil_face(const char* name, int arg, double* coefs, int* offs, int* sr, int* dr, int* ds, int maxn)
{
    std::string nm(name);
    Stencil<double> S;
    if (nm == "DiffCellToFaceLo") { S = Stencil<double>::DiffCellToFace(arg, Side::Lo); }
    else if (nm == "DiffCellToFaceHi") { S = Stencil<double>::DiffCellToFace(arg, Side::Hi); }
    else { return lib_stencil<double>(name, arg, coefs, offs, sr, dr, ds, maxn); }
    return dump_stencil(S, coefs, offs, sr, dr, ds, maxn);
}

// The rest of the stencil library (Proto_StencilDefs.H:152-229, 273-300,
// 356-382, 392-527), FP64, with up to three integer arguments.
int lib_stencil_ext(const char* name, int a, int b, int c, double* coefs, int* offs, int* sr, int* dr, int* ds,
                    int maxn)
{
    std::string nm(name);
    Stencil<double> S;
    const Side::LoHiSide side = b ? Side::Hi : Side::Lo;
    if (nm == "Identity") { S = Stencil<double>::Identity(); }
    else if (nm == "LaplacianFace") { S = Stencil<double>::LaplacianFace(a, b); }
    else if (nm == "CellToFace") { S = Stencil<double>::CellToFace(a, side, c); }
    else if (nm == "CellToFaceL") { S = Stencil<double>::CellToFaceL(a, b); }
    else if (nm == "CellToFaceH") { S = Stencil<double>::CellToFaceH(a, b); }
    else if (nm == "DiffCellToFace") { S = Stencil<double>::DiffCellToFace(a, side, c); }
    else if (nm == "Derivative") { S = Stencil<double>::Derivative(a, b, c); }
    else if (nm == "AvgDownFace") { S = Stencil<double>::AvgDownFace(a, side, c); }
    else if (nm == "faceToCell") { S = Stencil<double>::faceToCell(a, b); }
    else if (nm == "cornersToFaces") { S = Stencil<double>::cornersToFaces(a, b); }
    else if (nm == "CornersToCells") { S = Stencil<double>::CornersToCells(a); }
    else { return 0; }
    return dump_stencil(S, coefs, offs, sr, dr, ds, maxn);
}

template<typename T>
void stencil_apply(int n, const T* coefs, const int* offs, const int* sr, const int* dr, const int* ds,
                   const T* src, const int* slo, const int* shi,
                   T* dst, const int* dlo, const int* dhi,
                   const int* blo, const int* bhi, int infer, int replace, T scale, int* outbox)
{
    Stencil<T> S = mkstencil<T>(n, coefs, offs, sr, dr, ds);
    BoxData<T, 1, HOST> s(src, mkbox(slo, shi));
    BoxData<T, 1, HOST> d(dst, mkbox(dlo, dhi));
    Box b = infer ? Box() : mkbox(blo, bhi);
    if (infer) { b = S.indexDomain(d.box()) & S.indexRange(s.box()); }
    for (int k = 0; k < 3; k++)
    {
        outbox[k] = (k < DIM) ? b.low()[k] : 0;
        outbox[3 + k] = (k < DIM) ? b.high()[k] : 0;
    }
    if (b.empty()) { return; }
    S.apply(s, d, b, replace != 0, scale);
}

template<typename T, Operation OP>
T reduce_host(const T* data, long n)
{
    Reduction<T, OP, HOST> rxn;
    rxn.reduce(data, (size_t)n);
    return rxn.fetchLocal();
}

template<typename T>
T reduce_dispatch(int op, const T* data, long n)
{
    switch (op)
    {
        case 0: return reduce_host<T, Max>(data, n);
        case 1: return reduce_host<T, Min>(data, n);
        case 2: return reduce_host<T, Abs>(data, n);
        case 3: return reduce_host<T, Sum>(data, n);
        case 4: return reduce_host<T, SumAbs>(data, n);
        case 5: return reduce_host<T, SumSquare>(data, n);
        default: return 0;
    }
}

// global x-fastest array <-> LevelBoxData (valid cells only)
void level_from_global(LevelBoxData<double>& L, const double* g, int n)
{
    for (auto dit : L.layout())
    {
        BoxData<double>& b = L[dit];
        Box v = L.layout()[dit];
        for (auto it = v.begin(); it != v.end(); ++it)
        {
            Point p = *it;
            long gi = 0, f = 1;
            for (int d = 0; d < DIM; d++) { gi += p[d] * f; f *= n; }
            b(p) = g[gi];
        }
    }
}

void level_to_global(const LevelBoxData<double>& L, double* g, int n)
{
    for (auto dit : L.layout())
    {
        const BoxData<double>& b = L[dit];
        Box v = L.layout()[dit];
        for (auto it = v.begin(); it != v.end(); ++it)
        {
            Point p = *it;
            long gi = 0, f = 1;
            for (int d = 0; d < DIM; d++) { gi += p[d] * f; f *= n; }
            g[gi] = b(p);
        }
    }
}

PROTO_KERNEL_START
void axpyResidualF(Var<double>& a_r, const Var<double>& a_f, const Var<double>& a_lu)
{
    a_r(0) = a_f(0) - a_lu(0);
}
PROTO_KERNEL_END(axpyResidualF, axpyResidual)

DisjointBoxLayout periodic_layout(int n, int boxSize)
{
    Box domain = Box::Cube(n);
    Array<bool, DIM> per;
    per.fill(true);
    ProblemDomain pd(domain, per);
    return DisjointBoxLayout(pd, Point::Ones(boxSize));
}

// Correction-scheme V-cycle with red-black Gauss-Seidel relaxation.  The
// structure (coarseResidual / resnorm / fineInterp / vCycle) is that of
// LevelMultigrid.cpp:51-160; relax() is the GSRB of LaplaceOp.H:56-73,95-123
// written against LevelBoxData instead of Chombo's LevelData.
class GSRBMultigrid
{
public:
    GSRBMultigrid(const DisjointBoxLayout& a_layout, double a_dx, int a_level,
                  int a_pre, int a_post, int a_bottom, bool a_linear)
    {
        m_layout = a_layout;
        m_level = a_level;
        m_dx = a_dx;
        m_pre = a_pre; m_post = a_post; m_bottom = a_bottom;
        m_linear = a_linear;
        m_fineInterp = a_linear ? InterpStencil<double>::Linear(Point::Ones(2))
                                : InterpStencil<double>::Constant(Point::Ones(2));
        // LaplaceOp::init (LaplaceOp.H:56-73)
        m_lambda = -(m_dx * m_dx) / (4.0 * DIM);
        m_lambda *= 2.0;
        Stencil<double> Lap = Stencil<double>::Laplacian();
        Box K = Box::Cube(2);
        for (auto iter = K.begin(); iter != K.end(); ++iter)
        {
            int index = K.index(*iter);
            m_GSRB[index] = Lap;
            m_GSRB[index] *= (-1.0 / (m_dx * m_dx)) * Shift(*iter);
            m_GSRB[index].srcRatio() = Point::Ones(2);
            m_GSRB[index].destRatio() = Point::Ones(2);
            m_GSRB[index].destShift() = *iter;
            m_GSRB_I[index] = 1.0 * Shift(*iter);
            m_GSRB_I[index].srcRatio() = Point::Ones(2);
            m_GSRB_I[index].destRatio() = Point::Ones(2);
            m_GSRB_I[index].destShift() = *iter;
        }
        Point boxsize = m_layout.boxSize();
        ProblemDomain pd = m_layout.domain();
        if (m_level > 0)
        {
            Point refRatio = Point::Ones(2);
            DisjointBoxLayout dblCoarseLocal = m_layout.coarsen(refRatio);
            DisjointBoxLayout dblCoarse;
            ProblemDomain pdCoarse = pd.coarsen(refRatio);
            if (pdCoarse.sizes() % boxsize == Point::Zeros())
            {
                dblCoarse = DisjointBoxLayout(pdCoarse, boxsize);
            } else {
                dblCoarse = DisjointBoxLayout(pdCoarse, pdCoarse.sizes());
            }
            m_resc.define(dblCoarse, Point::Zeros());
            m_delta.define(dblCoarse, Point::Ones());
            m_localCoarse.define(dblCoarseLocal, m_linear ? Point::Ones() : Point::Zeros());
            m_coarse = std::make_shared<GSRBMultigrid>(dblCoarse, 2 * m_dx, m_level - 1,
                                                       m_pre, m_post, m_bottom, m_linear);
        }
    }

    void coarseResidual(LevelBoxData<double>& a_resc, LevelBoxData<double>& a_phi,
                        LevelBoxData<double>& a_rhs)
    {
        a_phi.exchange();
        double hsqinv = 1. / (m_dx * m_dx);
        auto L = Stencil<double>::Laplacian();
        auto A = Stencil<double>::AvgDown(2);
        for (auto dit : a_phi)
        {
            BoxData<double>& phi = a_phi[dit];
            BoxData<double>& rhs = a_rhs[dit];
            BoxData<double>& rescLocal = m_localCoarse[dit];
            BoxData<double> res(a_phi.layout()[dit]);
            res.setVal(0.);
            res += rhs;
            res += L(phi, -hsqinv);
            rescLocal |= A(res);
        }
        m_localCoarse.copyTo(a_resc);
    }

    double resnorm(LevelBoxData<double>& a_phi, LevelBoxData<double>& a_rhs)
    {
        a_phi.exchange();
        double hsqinv = 1. / (m_dx * m_dx);
        Reduction<double, Abs, HOST> rxn;
        auto L = Stencil<double>::Laplacian();
        for (auto dit : a_phi)
        {
            BoxData<double>& phi = a_phi[dit];
            BoxData<double>& rhs = a_rhs[dit];
            BoxData<double> res(a_phi.layout()[dit]);
            res.setVal(0.);
            res -= rhs;
            res += L(phi, hsqinv);
            res.reduce(rxn);
        }
        return rxn.fetch();
    }

    void relax(LevelBoxData<double>& a_phi, LevelBoxData<double>& a_rhs, int a_n)
    {
        Box K = Box::Cube(2);
        for (int ii = 0; ii < a_n; ii++)
        {
            for (int icolor = 0; icolor < 2; icolor++)
            {
                a_phi.exchange();
                for (auto dit : a_phi)
                {
                    BoxData<double>& phi = a_phi[dit];
                    BoxData<double>& rhs = a_rhs[dit];
                    BoxData<double> temp(a_phi.layout()[dit]);
                    for (auto kiter = K.begin(); kiter != K.end(); ++kiter)
                    {
                        int color = 0;
                        for (int d = 0; d < DIM; d++) { color += (*kiter)[d]; }
                        color = color % 2;
                        if (color == icolor)
                        {
                            int index = K.index(*kiter);
                            temp |= m_GSRB[index](phi);
                            temp += m_GSRB_I[index](rhs);
                            phi += m_GSRB_I[index](temp, m_lambda);
                        }
                    }
                }
            }
        }
    }

    void fineInterp(LevelBoxData<double>& a_phi, LevelBoxData<double>& a_delta)
    {
        a_delta.copyTo(m_localCoarse);
        if (m_linear) { m_localCoarse.exchange(); }
        for (auto dit : a_phi)
        {
            BoxData<double>& phi = a_phi[dit];
            BoxData<double>& delta = m_localCoarse[dit];
            phi += m_fineInterp(delta);
        }
    }

    void vCycle(LevelBoxData<double>& a_phi, LevelBoxData<double>& a_rhs)
    {
        if (m_level > 0)
        {
            relax(a_phi, a_rhs, m_pre);
            coarseResidual(m_resc, a_phi, a_rhs);
            m_delta.setToZero();
            m_coarse->vCycle(m_delta, m_resc);
            fineInterp(a_phi, m_delta);
            relax(a_phi, a_rhs, m_post);
        } else {
            relax(a_phi, a_rhs, m_bottom);
        }
    }

private:
    DisjointBoxLayout m_layout;
    int m_level, m_pre, m_post, m_bottom;
    bool m_linear;
    double m_dx, m_lambda;
    Stencil<double> m_GSRB[1 << DIM];
    Stencil<double> m_GSRB_I[1 << DIM];
    InterpStencil<double> m_fineInterp;
    LevelBoxData<double> m_localCoarse, m_resc, m_delta;
    std::shared_ptr<GSRBMultigrid> m_coarse;
};

} // namespace

extern "C" {

int ref_dim() { return DIM; }

int ref_stencil_lib(const char* name, int arg, double* coefs, int* offs, int* sr, int* dr, int* ds, int maxn)
{ return lib_stencil_face(name, arg, coefs, offs, sr, dr, ds, maxn); }

int ref_stencil_lib_f(const char* name, int arg, float* coefs, int* offs, int* sr, int* dr, int* ds, int maxn)
{ return lib_stencil<float>(name, arg, coefs, offs, sr, dr, ds, maxn); }

// Stencil<double>::apply (Proto_StencilImplem.H:412-431 -> hostApply 725-828).
// infer != 0 lets Proto infer the index box; the box used is written to outbox[6].
void ref_stencil_apply(int n, const double* coefs, const int* offs, const int* sr, const int* dr, const int* ds,
                       const double* src, const int* slo, const int* shi,
                       double* dst, const int* dlo, const int* dhi,
                       const int* blo, const int* bhi, int infer, int replace, double scale, int* outbox)
{ stencil_apply<double>(n, coefs, offs, sr, dr, ds, src, slo, shi, dst, dlo, dhi, blo, bhi, infer, replace, scale, outbox); }

void ref_stencil_apply_f(int n, const float* coefs, const int* offs, const int* sr, const int* dr, const int* ds,
                         const float* src, const int* slo, const int* shi,
                         float* dst, const int* dlo, const int* dhi,
                         const int* blo, const int* bhi, int infer, int replace, float scale, int* outbox)
{ stencil_apply<float>(n, coefs, offs, sr, dr, ds, src, slo, shi, dst, dlo, dhi, blo, bhi, infer, replace, scale, outbox); }

// Reduction<T,OP,HOST> (Proto_ReductionImplem.H:64-102,157-192); op = Max,Min,Abs,Sum,SumAbs,SumSquare.
double ref_reduce(int op, const double* data, long n) { return reduce_dispatch<double>(op, data, n); }
int ref_stencil_lib_ext(const char* name, int a, int b, int c, double* coefs, int* offs, int* sr, int* dr,
                        int* ds, int maxn)
{
    return lib_stencil_ext(name, a, b, c, coefs, offs, sr, dr, ds, maxn);
}
float ref_reduce_f(int op, const float* data, long n) { return reduce_dispatch<float>(op, data, n); }

// InterpStencil<double>::{Constant,Linear}(ratio) applied as fine (+= or |=) I(crse)
// (Proto_InterpStencilImplem.H:83-112, Proto_InterpStencilDefs.H:1-36).
void ref_interp_apply(int linear, int ratio, const double* crse, const int* clo, const int* chi,
                      double* fine, const int* flo, const int* fhi, int replace)
{
    InterpStencil<double> I = linear ? InterpStencil<double>::Linear(Point::Ones(ratio))
                                     : InterpStencil<double>::Constant(Point::Ones(ratio));
    BoxData<double, 1, HOST> c(crse, mkbox(clo, chi));
    BoxData<double, 1, HOST> f(fine, mkbox(flo, fhi));
    if (replace) { f |= I(c); } else { f += I(c); }
}

// The same with Proto's other InterpStencils (Proto_InterpStencilDefs.H):
// kind 2 Quadratic(ratio), kind 3 FiniteVolume(ratio, order).
void ref_interp_apply_kind(int kind, int order, int ratio, const double* crse, const int* clo, const int* chi,
                           double* fine, const int* flo, const int* fhi, int replace)
{
    InterpStencil<double> I = kind == 2 ? InterpStencil<double>::Quadratic(Point::Ones(ratio))
                                        : InterpStencil<double>::FiniteVolume(Point::Ones(ratio), order);
    BoxData<double, 1, HOST> c(crse, mkbox(clo, chi));
    BoxData<double, 1, HOST> f(fine, mkbox(flo, fhi));
    if (replace) { f |= I(c); } else { f += I(c); }
}

// interpBoundaries (Proto_LevelBoxDataImplem.H:953-990): a periodic coarse
// level of 16^3 cells in 8^3 boxes (1 ghost) and a ratio-2 fine level on the
// 32^3 domain covering the 27 patches {1,2,3}^3 of its 8^3-box grid (2
// ghosts).  Coarse valid cells come from crse_g (16^3, x fastest), fine valid
// cells from fine_g (32^3), every ghost cell starts at `fill`.  kind 0
// Constant, 1 Linear, 2 Quadratic (ratio 2).  Output: for each fine box in
// layout order its low corner (3 ints) and its ghosted data, x fastest.
int ref_interp_boundaries(int kind, const double* crse_g, const double* fine_g, double fill, int* lows, double* out)
{
    Array<bool, DIM> per;
    per.fill(true);
    ProblemDomain cpd(Box::Cube(16), per), fpd(Box::Cube(32), per);
    DisjointBoxLayout clayout(cpd, Point::Ones(8));
    std::vector<Point> patches;
    for (auto it = Box(Point::Ones(1), Point::Ones(3)).begin(); it != Box(Point::Ones(1), Point::Ones(3)).end(); ++it)
    {
        patches.push_back(*it);
    }
    DisjointBoxLayout flayout(fpd, patches, Point::Ones(8));
    LevelBoxData<double> crse(clayout, Point::Ones(1)), fine(flayout, Point::Ones(2));
    for (auto dit : clayout) { crse[dit].setVal(fill); }
    for (auto dit : flayout) { fine[dit].setVal(fill); }
    level_from_global(crse, crse_g, 16);
    level_from_global(fine, fine_g, 32);
    InterpStencil<double> I = kind == 0 ? InterpStencil<double>::Constant(Point::Ones(2))
                            : (kind == 1 ? InterpStencil<double>::Linear(Point::Ones(2))
                                         : InterpStencil<double>::Quadratic(Point::Ones(2)));
    interpBoundaries(crse, fine, I);
    int ib = 0;
    long off = 0;
    for (auto dit : flayout)
    {
        Box v = flayout[dit];
        for (int d = 0; d < 3; d++) { lows[3 * ib + d] = (d < DIM) ? v.low()[d] : 0; }
        const BoxData<double>& b = fine[dit];
        std::memcpy(out + off, b.data(), b.box().size() * sizeof(double));
        off += b.box().size();
        ib++;
    }
    return ib;
}

// LevelBoxData::exchange on a periodic n^DIM domain tiled by boxSize^DIM boxes
// with `ghost` ghost layers (Proto_LevelBoxDataImplem.H:313-341).  Valid cells
// come from the global array g; ghost cells start at `fill`.  Output: for each
// box in layout order its low corner (3 ints) and its ghosted data, x fastest.
int ref_level_exchange(int n, int boxSize, int ghost, const double* g, double fill, int* lows, double* out)
{
    DisjointBoxLayout layout = periodic_layout(n, boxSize);
    LevelBoxData<double> L(layout, Point::Ones(ghost));
    for (auto dit : layout) { L[dit].setVal(fill); }
    level_from_global(L, g, n);
    L.exchange();
    int ib = 0;
    long off = 0;
    for (auto dit : layout)
    {
        Box v = layout[dit];
        for (int d = 0; d < 3; d++) { lows[3 * ib + d] = (d < DIM) ? v.low()[d] : 0; }
        const BoxData<double>& b = L[dit];
        std::memcpy(out + off, b.data(), b.box().size() * sizeof(double));
        off += b.box().size();
        ib++;
    }
    return ib;
}

// LevelBoxData<double, 2>::linearOut (Proto_LevelBoxDataImplem.H:412-507 over
// BoxData::linearOut, Proto_BoxDataImplem.H:1619-1648) of a level whose every
// cell -- ghosts included -- holds value(p, c) = p.x + 1000 p.y + 1e6 p.z +
// 0.5 c (p the global, unwrapped point).  Returns the byte count written.
long ref_level_linear_out(int n, int boxSize, int ghost, double* buffer, int* lows)
{
    DisjointBoxLayout layout = periodic_layout(n, boxSize);
    LevelBoxData<double, 2> L(layout, Point::Ones(ghost));
    int ib = 0;
    for (auto dit : layout)
    {
        Box v = layout[dit];
        for (int d = 0; d < 3; d++) { lows[3 * ib + d] = (d < DIM) ? v.low()[d] : 0; }
        ib++;
        BoxData<double, 2>& b = L[dit];
        for (auto it = b.box().begin(); it != b.box().end(); ++it)
        {
            Point p = *it;
            for (int c = 0; c < 2; c++)
            {
                b(p, c) = p[0] + 1000.0 * p[1] + (DIM > 2 ? 1e6 * p[2] : 0.0) + 0.5 * c;
            }
        }
    }
    L.linearOut(buffer);
    return (long)L.linearSize();
}

// Archived correction-scheme multigrid with damped point Jacobi
// (examples/_archived/LevelMultigrid/src/LevelMultigrid.cpp, compiled from its
// own source).  phi (in/out) and rhs are global x-fastest arrays; hist gets
// resnorm before the first V-cycle and after each one.  Returns #V-cycles done.
int ref_mg_jacobi_solve(int n, int boxSize, int numLevels, const double* rhs, double* phi,
                        int maxIter, double tol, double* hist)
{
    DisjointBoxLayout layout = periodic_layout(n, boxSize);
    double dx = 1.0 / n;
    LevelBoxData<double> R(layout, Point::Zeros());
    LevelBoxData<double> P(layout, Point::Ones());
    R.setToZero();
    P.setToZero();
    level_from_global(R, rhs, n);
    level_from_global(P, phi, n);
    LevelMultigrid mg(layout, dx, numLevels);
    double res0 = mg.resnorm(P, R);
    hist[0] = res0;
    int it = 0;
    for (; it < maxIter; it++)
    {
        mg.vCycle(P, R);
        double r = mg.resnorm(P, R);
        hist[it + 1] = r;
        if (r < tol * res0) { it++; break; }
    }
    level_to_global(P, phi, n);
    return it;
}

// Correction-scheme multigrid with GSRB relaxation (class GSRBMultigrid above).
int ref_mg_gsrb_solve(int n, int boxSize, int numLevels, int pre, int post, int bottom, int linear,
                      const double* rhs, double* phi, int maxIter, double tol, double* hist)
{
    DisjointBoxLayout layout = periodic_layout(n, boxSize);
    double dx = 1.0 / n;
    LevelBoxData<double> R(layout, Point::Zeros());
    LevelBoxData<double> P(layout, Point::Ones());
    R.setToZero();
    P.setToZero();
    level_from_global(R, rhs, n);
    level_from_global(P, phi, n);
    GSRBMultigrid mg(layout, dx, numLevels, pre, post, bottom, linear != 0);
    double res0 = mg.resnorm(P, R);
    hist[0] = res0;
    int it = 0;
    for (; it < maxIter; it++)
    {
        mg.vCycle(P, R);
        double r = mg.resnorm(P, R);
        hist[it + 1] = r;
        if (r < tol * res0) { it++; break; }
    }
    level_to_global(P, phi, n);
    return it;
}

// The config-2 program as a Proto user writes it (one operator):
//   BoxData<double> Lu(box); Lu |= L(u, scale);          (Stencil apply)
//   forallInPlace(axpyResidual, r, f, Lu);               (r = f - Lu)
//   amax = r.absMax(); sum = r.sum();                    (Reduction)
// which = 0: Laplacian(), 1: Laplacian_27().  Used as the CPU baseline.
void ref_config2_step(int which, const double* u, const int* ulo, const int* uhi, const double* f,
                      double* r, const int* lo, const int* hi, double scale, double* amax, double* sum)
{
    Stencil<double> L;
#if DIM == 3
    L = which ? Stencil<double>::Laplacian_27() : Stencil<double>::Laplacian();
#else
    L = Stencil<double>::Laplacian();
#endif
    Box b = mkbox(lo, hi);
    BoxData<double, 1, HOST> U(u, mkbox(ulo, uhi));
    BoxData<double, 1, HOST> F(f, b);
    BoxData<double, 1, HOST> R(r, b);
    BoxData<double, 1, HOST> Lu(b);
    Lu |= L(U, scale);
    forallInPlace(axpyResidual, R, F, Lu);
    *amax = R.absMax();
    *sum = R.sum();
}

// One GSRB sweep on a single patch (one BoxData with `g` ghost layers held
// fixed, no exchange), with the sub-lattice stencils of LaplaceOp.H:56-73
// applied as its relax() does (95-123) -- colour by (x+y+z) parity of the
// box-local index, the box's low corner being 0.
void ref_patch_gsrb(int nx, int ny, int nz, int g, double dx, const double* rhs, double* phi)
{
    Box valid(Point::Zeros(), Point(nx - 1, ny - 1, nz - 1));
    BoxData<double, 1, HOST> P(phi, valid.grow(g));
    BoxData<double, 1, HOST> Rh(rhs, valid);
    double lambda = -(dx * dx) / (4.0 * DIM);
    lambda *= 2.0;
    Stencil<double> Lap = Stencil<double>::Laplacian();
    Stencil<double> GS[1 << DIM], GI[1 << DIM];
    Box K = Box::Cube(2);
    for (auto iter = K.begin(); iter != K.end(); ++iter)
    {
        int index = K.index(*iter);
        GS[index] = Lap;
        GS[index] *= (-1.0 / (dx * dx)) * Shift(*iter);
        GS[index].srcRatio() = Point::Ones(2);
        GS[index].destRatio() = Point::Ones(2);
        GS[index].destShift() = *iter;
        GI[index] = 1.0 * Shift(*iter);
        GI[index].srcRatio() = Point::Ones(2);
        GI[index].destRatio() = Point::Ones(2);
        GI[index].destShift() = *iter;
    }
    for (int icolor = 0; icolor < 2; icolor++)
    {
        BoxData<double> temp(valid);
        for (auto kiter = K.begin(); kiter != K.end(); ++kiter)
        {
            int color = 0;
            for (int d = 0; d < DIM; d++) { color += (*kiter)[d]; }
            color = color % 2;
            if (color == icolor)
            {
                int index = K.index(*kiter);
                temp |= GS[index](P);
                temp += GI[index](Rh);
                P += GI[index](temp, lambda);
            }
        }
    }
}

// One GSRB relaxation pass (niter iterations) on a single level, no coarse grids.
void ref_gsrb_relax(int n, int boxSize, int niter, const double* rhs, double* phi)
{
    DisjointBoxLayout layout = periodic_layout(n, boxSize);
    LevelBoxData<double> R(layout, Point::Zeros());
    LevelBoxData<double> P(layout, Point::Ones());
    R.setToZero();
    P.setToZero();
    level_from_global(R, rhs, n);
    level_from_global(P, phi, n);
    GSRBMultigrid mg(layout, 1.0 / n, 0, 0, 0, 0, false);
    mg.relax(P, R, niter);
    level_to_global(P, phi, n);
}

// Library FAS multigrid: LevelMultigridSolver<BoxOp_Laplace,double,HOST>
// (include/utils/implem/Proto_LevelMultigridSolverImplem.H:82-433) with the
// 4th-order flux-form operator of examples/_common/BoxOp_Laplace.H.
int ref_mg_fas_solve(int n, int boxSize, int numLevels, const double* rhs, double* phi,
                     int maxIter, double tol, double* hist)
{
    DisjointBoxLayout layout = periodic_layout(n, boxSize);
    double dx = 1.0 / n;
    LevelBoxData<double, 1, HOST> state(layout, Point::Ones(2));
    LevelBoxData<double, 1, HOST> force(layout, Point::Zeros());
    LevelBoxData<double, 1, HOST> res(layout, Point::Zeros());
    state.setToZero();
    force.setToZero();
    level_from_global(force, rhs, n);
    level_from_global(state, phi, n);
    LevelMultigridSolver<BoxOp_Laplace, double, HOST> mg(layout, Point::Ones(2), numLevels, dx);
    double res0 = mg.resnorm(res, state, force);
    hist[0] = res0;
    int it = 0;
    for (; it < maxIter; it++)
    {
        double r = mg.solve(state, force, 1, 0.0);
        hist[it + 1] = r;
        if (r < tol * res0) { it++; break; }
    }
    level_to_global(state, phi, n);
    return it;
}

// LevelOp<BoxOp_TestLaplace> applied to a periodic level
// (tests/LevelOpTests.cpp:22-60 with a seeded state): state from the global
// n^3 array (2 ghost layers, exchanged), out = L(state) with Proto's default
// scalings (flux scale 1, no diagonal), written back as a global array.
void ref_levelop_apply(int n, int boxSize, const double* phi, double* out)
{
    DisjointBoxLayout layout = periodic_layout(n, boxSize);
    Array<double, DIM> dx;
    dx.fill(1.0 / n);
    LevelOp<BoxOp_TestLaplace, double, PeriodicBC, HOST> op(layout, dx);
    LevelBoxData<double, 1, HOST> state(layout, Point::Ones(2));
    LevelBoxData<double, 1, HOST> dst(layout, Point::Zeros());
    state.setToZero();
    level_from_global(state, phi, n);
    state.exchange();
    dst.setToZero();
    op(dst, state);
    level_to_global(dst, out, n);
}

// BoxOp_TestFlux / _TestSource / _TestDiag (which 0 / 1 / 2) on the single box
// [0, n)^3 of a non-periodic domain, as tests/BoxOpTests.cpp:9-210 sets them
// up (flux scale 31 for flux and source, diagonal scale 31 for diag, operator
// scale 17), with seeded src and aux on the box grown by one.  dst gets L on
// the box; flux (which 0 only) the DIM face fluxes one after another.
void ref_boxop_apply(int which, int n, const double* src, const double* aux, double* dst, double* flux)
{
    Box B0 = Box::Cube(n);
    Array<bool, DIM> per;
    per.fill(false);
    ProblemDomain domain(B0, per);
    DisjointBoxLayout layout(domain, B0.sizes());
    const double dx = 1.0 / n;
    BoxData<double, 1, HOST> hsrc(B0.grow(1)), haux(B0.grow(1)), hdst(B0);
    std::memcpy(hsrc.data(), src, sizeof(double) * hsrc.box().size());
    std::memcpy(haux.data(), aux, sizeof(double) * haux.box().size());
    Array<BoxData<double, 1, HOST>, DIM> fl;
    for (int d = 0; d < DIM; d++) { fl[d].define(B0.grow((Centering)d)); }
    if (which == 0) {
        BoxOp_TestFlux<double, HOST> op(layout, *layout.begin(), dx);
        op.setFluxScale(31.0);
        op(hdst, fl, hsrc, haux, 17.0);
    } else if (which == 1) {
        BoxOp_TestSource<double, HOST> op(layout, *layout.begin(), dx);
        op.setFluxScale(31.0);
        op(hdst, hsrc, haux, 17.0);
    } else {
        BoxOp_TestDiag<double, HOST> op(layout, *layout.begin(), dx);
        op.setDiagScale(31.0);
        op(hdst, hsrc, haux, 17.0);
    }
    std::memcpy(dst, hdst.data(), sizeof(double) * B0.size());
    if (which == 0 && flux) {
        long off = 0;
        for (int d = 0; d < DIM; d++) {
            std::memcpy(flux + off, fl[d].data(), sizeof(double) * fl[d].box().size());
            off += fl[d].box().size();
        }
    }
}

} // extern "C"
