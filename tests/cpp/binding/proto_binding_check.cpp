// Compile-and-link check of the Proto-side binding (proto_amd/binding/
// Proto_protomi.H) against Proto's own headers, where they lie in the
// reference tree, and against libprotomi.so: every binding template is
// instantiated for the types Proto's DEVICE paths use, and the shared object
// is linked with --no-undefined, so each C-ABI symbol it calls must resolve.
// Built by g++ on the host (no device code); nothing is run.
#include "Proto.H"
#include "Proto_protomi.H"

using namespace Proto;

// what Stencil<T>::protoApply<DEVICE> becomes
template int protomi::apply<double, 1, 1, 1>(const protomi::StencilHandle<double>&,
                                             const BoxData<double, 1, DEVICE, 1, 1>&,
                                             BoxData<double, 1, DEVICE, 1, 1>&, const Box&, bool, double, pm_stream_t);
template int protomi::apply<float, 3, 1, 1>(const protomi::StencilHandle<float>&, const BoxData<float, 3, DEVICE, 1, 1>&,
                                            BoxData<float, 3, DEVICE, 1, 1>&, const Box&, bool, float, pm_stream_t);
template class protomi::StencilHandle<double>;
template class protomi::StencilHandle<float>;
// what Reduction<T,OP,DEVICE>::reduce / reset become
template int protomi::reduce<double, Abs>(const double*, size_t, double*, pm_stream_t);
template int protomi::reduce<double, Sum>(const double*, size_t, double*, pm_stream_t);
template int protomi::reduce<float, Max>(const float*, size_t, float*, pm_stream_t);
template int protomi::reset<double, Min>(double*, pm_stream_t);
// what LevelBoxData::exchange becomes
template int protomi::exchange<double, 1>(const protomi::LayoutHandle&, LevelBoxData<double, 1, DEVICE>&, pm_stream_t);
template int protomi::exchange<float, 2>(const protomi::LayoutHandle&, LevelBoxData<float, 2, DEVICE>&, pm_stream_t);

// A Proto-side entry point that uses them together, as Stencil::protoApply,
// Reduction and the FAS solver would after the change.
extern "C" int proto_binding_demo(const DisjointBoxLayout* layout, const BoxData<double, 1, DEVICE>* src,
                                  BoxData<double, 1, DEVICE>* dst, double* devTotal,
                                  LevelBoxData<double, 1, DEVICE>* state, LevelBoxData<double, 1, DEVICE>* force)
{
    static_assert(Abs == 2 && SumSquare == 5, "Proto's Operation order is the C ABI's");
    protomi::StencilHandle<double> L(Stencil<double>::Laplacian());
    int rc = protomi::apply(L, *src, *dst, dst->box(), true, 1.0, nullptr);
    if (rc == PM_OK) { rc = protomi::reset<double, Abs>(devTotal, nullptr); }
    if (rc == PM_OK) { rc = protomi::reduce<double, Abs>(dst->data(), dst->size(), devTotal, nullptr); }
    protomi::LayoutHandle lh(*layout);
    if (rc == PM_OK) { rc = protomi::exchange(lh, *state, nullptr); }
    protomi::FasSolver fas;
    if (rc == PM_OK) { rc = fas.define(lh, 3, 1.0 / 32); }
    std::vector<double> hist(11);
    int iters = 0;
    if (rc == PM_OK) { rc = fas.solve(*state, *force, 10, 1e-10, hist.data(), &iters, nullptr); }
    return rc;
}
