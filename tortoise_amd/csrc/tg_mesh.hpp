// tg_mesh.hpp -- structured simplex mesh arithmetic usable on host and device.
//
// Sections are the cells of an nx x ny (x nz) box; each owns dim! simplices.
// Element id: 2*(sx + sy*nx) + t (2D), 6*(sx + (sy + sz*ny)*nx) + t (3D)
// (src/Geometry/Structured/DirectSpace/Mesh.cpp:412-420).  The vertices of
// simplex t walk from the section's origin corner to the opposite corner,
// switching on one axis per step; the axis order is the t-th permutation in
// lexicographic order (the depth-first construction of Mesh.cpp:55-87).
// Nodes are the vertices of the lattice: id = ix + (nx+1)*(iy + (ny+1)*iz).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace tg {

struct MeshDims {
    int dim;
    int nsec[3];
};

__host__ __device__ inline int factorial(int d) {
    int r = 1;
    for (int i = 2; i <= d; ++i) r *= i;
    return r;
}

// Corner offsets of vertex a of simplex t inside its section, as a bit mask
// (bit d set: +1 along axis d).
__host__ __device__ inline int simplex_vertex_mask(int dim, int t, int a) {
    int avail[3] = {0, 1, 2};
    int navail = dim, rem = t, mask = 0;
    for (int k = 0; k < a; ++k) {
        int f = factorial(dim - 1 - k);
        int idx = rem / f;
        rem %= f;
        mask |= 1 << avail[idx];
        for (int q = idx; q < navail - 1; ++q) avail[q] = avail[q + 1];
        --navail;
    }
    return mask;
}

__host__ __device__ inline int64_t lattice_node(const MeshDims& m, const int* c) {
    int64_t id = 0, stride = 1;
    for (int d = 0; d < m.dim; ++d) {
        id += stride * c[d];
        stride *= (m.nsec[d] + 1);
    }
    return id;
}

// Node ids of simplex e (reference vertex order) -> out[0..dim].
__host__ __device__ inline void element_nodes(const MeshDims& m, int64_t e, int64_t* out) {
    int nes = factorial(m.dim);
    int64_t s = e / nes;
    int t = (int)(e % nes);
    int sc[3] = {0, 0, 0};
    for (int d = 0; d < m.dim; ++d) {
        sc[d] = (int)(s % m.nsec[d]);
        s /= m.nsec[d];
    }
    for (int a = 0; a <= m.dim; ++a) {
        int mask = simplex_vertex_mask(m.dim, t, a);
        int c[3];
        for (int d = 0; d < m.dim; ++d) c[d] = sc[d] + ((mask >> d) & 1);
        out[a] = lattice_node(m, c);
    }
}

// Incident (simplex, local vertex) pairs of a lattice node; returns the
// count (<= 6 in 2D, <= 24 in 3D).  Sections are visited in increasing
// corner-mask order and simplices in increasing t.
__host__ __device__ inline int node_star(const MeshDims& m, int64_t node, int64_t* elems,
                                         int* locs) {
    int nc[3] = {0, 0, 0};
    int64_t r = node;
    for (int d = 0; d < m.dim; ++d) {
        nc[d] = (int)(r % (m.nsec[d] + 1));
        r /= (m.nsec[d] + 1);
    }
    int nes = factorial(m.dim);
    int cnt = 0;
    for (int msk = 0; msk < (1 << m.dim); ++msk) {
        int sc[3];
        bool ok = true;
        int64_t sid = 0, stride = 1;
        for (int d = 0; d < m.dim; ++d) {
            sc[d] = nc[d] - ((msk >> d) & 1);
            if (sc[d] < 0 || sc[d] >= m.nsec[d]) ok = false;
            sid += stride * sc[d];
            stride *= m.nsec[d];
        }
        if (!ok) continue;
        for (int t = 0; t < nes; ++t) {
            for (int a = 0; a <= m.dim; ++a) {
                int vm = simplex_vertex_mask(m.dim, t, a);
                if (vm == msk) {
                    elems[cnt] = sid * nes + t;
                    locs[cnt] = a;
                    ++cnt;
                }
            }
        }
    }
    return cnt;
}

}  // namespace tg
