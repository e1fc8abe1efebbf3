// sfx_mesh_host.h -- host-side mesh containers and builders (tessellation
// stays on the host; these produce the arrays the flux-pass kernels read).
#pragma once

#include <cstdint>
#include <vector>

namespace sfx {

// Raw mesh as a tessellator hands it over (reference semantics of VorTess
// faces and cells).  SoA, component-major ([ndim][C] etc.).
struct HostMesh {
    int ndim = 2;
    int64_t C = 0, F = 0;
    std::vector<double> pos, volume, centroid;          // cells
    std::vector<int32_t> left, right;                   // faces; right -1: mirror
    std::vector<double> area, midpoint, shift;          // faces
    std::vector<uint8_t> ghost;                         // right is a periodic copy
    std::vector<int32_t> twin;                          // other copy of a periodic face
    std::vector<double> verts;                          // 2D: v0x v0y v1x v1y per face
    std::vector<int64_t> fv_off;                        // [F+1] face vertex lists
    std::vector<double> fverts;                         // [ndim][fv_off[F]]
    std::vector<int64_t> cf_off;                        // per-cell faces (walk order)
    std::vector<int32_t> cf_face;
};

int voronoi2d_periodic(int64_t C, const double* pos, double Lx, double Ly, HostMesh& m,
                       int nthreads);
int voronoi3d_periodic(int64_t C, const double* pos, double Lx, double Ly, double Lz,
                       HostMesh& m, int nthreads);

}  // namespace sfx
