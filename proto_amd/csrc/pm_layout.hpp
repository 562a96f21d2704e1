// Device view of a pm_layout (every box the same size, on the box grid of
// the domain) and the cell addressing shared by the level kernels.
#pragma once

#include <vector>

#include "pm_internal.hpp"

struct pm_layout {
    int dim;
    int domain[3], bs[3], periodic[3], nb[3];
    int nboxes;
    std::vector<int> lows;
    int* d_lows;
    int* d_table;
};

namespace pm {

struct LayoutDev {
    int dim, nboxes;
    int domain[3], bs[3], periodic[3], nb[3];
    const int* lows;
    const int* table;
};

inline LayoutDev ldev(const pm_layout* L)
{
    LayoutDev d;
    d.dim = L->dim;
    d.nboxes = L->nboxes;
    for (int k = 0; k < 3; k++) {
        d.domain[k] = L->domain[k];
        d.bs[k] = L->bs[k];
        d.periodic[k] = L->periodic[k];
        d.nb[k] = L->nb[k];
    }
    d.lows = L->d_lows;
    d.table = L->d_table;
    return d;
}

struct G3 {
    int g[3];
};

inline G3 g3(const int* g)
{
    G3 r;
    for (int k = 0; k < 3; k++) { r.g[k] = g ? g[k] : 0; }
    return r;
}

// elements of one ghosted box
__host__ __device__ inline long gvol(const LayoutDev& L, const G3& g)
{
    return (long)(L.bs[0] + 2 * g.g[0]) * (L.bs[1] + 2 * g.g[1]) * (L.bs[2] + 2 * g.g[2]);
}

// index of the cell at box-relative (x, y, z) (0 = box low) inside the ghosted storage
__device__ inline long rel_index(const LayoutDev& L, const G3& g, int x, int y, int z)
{
    const long X = L.bs[0] + 2 * g.g[0], Y = L.bs[1] + 2 * g.g[1];
    return (long)(x + g.g[0]) + X * ((long)(y + g.g[1]) + Y * (long)(z + g.g[2]));
}

// Box owning global cell p (periodic images folded in); -1 if none.  On
// return p holds the image inside the domain.
__device__ inline int owner(const LayoutDev& L, int p[3])
{
    int b[3] = {0, 0, 0};
    for (int k = 0; k < 3; k++) {
        if (p[k] < 0 || p[k] >= L.domain[k]) {
            if (!L.periodic[k]) { return -1; }
            p[k] = ((p[k] % L.domain[k]) + L.domain[k]) % L.domain[k];
        }
        b[k] = p[k] / L.bs[k];
    }
    return L.table[b[0] + L.nb[0] * (b[1] + L.nb[1] * b[2])];
}

// global cell value of level data (valid cells of the owner; periodic images)
template<typename T>
__device__ inline T level_value(const LayoutDev& L, const G3& g, const T* data, int x, int y, int z)
{
    int p[3] = {x, y, z};
    const int o = owner(L, p);
    const int* lo = L.lows + 3 * o;
    return data[(long)o * gvol(L, g) + rel_index(L, g, p[0] - lo[0], p[1] - lo[1], p[2] - lo[2])];
}

}  // namespace pm
