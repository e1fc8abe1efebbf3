// mesh_voronoi.cpp -- host construction of periodic Voronoi meshes and of the
// per-face geometry the flux-pass kernels read.
//
// The reference builds its faces in VorTess::construct (Shadowfax
// src/VorTess.cpp:285-396) from a Delaunay tessellation that needs
// Boost.Multiprecision.  Here each cell is built independently by clipping a
// box with the bisector half-planes of its neighbours (found through a grid
// hash over the periodic box); the face list then follows the reference's
// contract:
//   * cells in generator order; each cell's faces in counter-clockwise order;
//   * a face is created by the first of its two cells to be visited (its
//     "left"); a neighbour that is a periodic copy (ghost, reference id 1,
//     src/DelTess_mpi.cpp:151-170) never owns a face, so every periodic
//     boundary face exists twice, once from each real cell;
//   * face geometry as VorFace::calculate_area / calculate_midpoint
//     (src/VorFace.cpp:335-336,398-401; area set to 0 when the two vertices
//     are closer than 1e-13, src/VorTess.cpp:340-342), cell volume and
//     centroid from the faces' stored vertices as VorCell::calculate_volume /
//     calculate_centroid (src/VorCell.cpp:297-315,375-385).
#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <thread>
#include <unordered_map>
#include <vector>

#include "sfx_mesh_host.h"

namespace sfx {
namespace {

struct Cand {
    double x, y;      // position of the (possibly shifted) generator
    int32_t j;        // generator index
    double sx, sy;    // periodic shift applied
};

struct Edge {  // polygon edge from vertex k to k+1, created by candidate c
    int32_t c;
};

struct Cell2 {
    std::vector<double> vx, vy;  // CCW vertices
    std::vector<int32_t> lab;    // lab[k]: candidate creating edge k -> k+1
    std::vector<Cand> cand;
};

// clip convex polygon (CCW) by {x : (x - m).u <= 0}
void clip(Cell2& P, const Cand& q, double xi, double yi, int32_t label) {
    const double ux = q.x - xi, uy = q.y - yi;
    const double mx = 0.5 * (q.x + xi), my = 0.5 * (q.y + yi);
    const size_t n = P.vx.size();
    std::vector<double> s(n);
    bool any_out = false;
    for(size_t k = 0; k < n; k++) {
        s[k] = (P.vx[k] - mx) * ux + (P.vy[k] - my) * uy;
        any_out |= s[k] > 0.0;
    }
    if(!any_out) return;
    std::vector<double> nx, ny;
    std::vector<int32_t> nl;
    nx.reserve(n + 2);
    ny.reserve(n + 2);
    nl.reserve(n + 2);
    for(size_t k = 0; k < n; k++) {
        const size_t k1 = (k + 1) % n;
        const bool in0 = s[k] <= 0.0, in1 = s[k1] <= 0.0;
        if(in0) {
            nx.push_back(P.vx[k]);
            ny.push_back(P.vy[k]);
            nl.push_back(P.lab[k]);
        }
        if(in0 != in1) {
            const double t = s[k] / (s[k] - s[k1]);
            nx.push_back(P.vx[k] + t * (P.vx[k1] - P.vx[k]));
            ny.push_back(P.vy[k] + t * (P.vy[k1] - P.vy[k]));
            // leaving the half-plane: the new edge runs along the bisector
            nl.push_back(in0 ? label : P.lab[k]);
        }
    }
    P.vx.swap(nx);
    P.vy.swap(ny);
    P.lab.swap(nl);
}

struct Grid2 {
    double Lx, Ly, hx, hy;
    int nx, ny;
    std::vector<int64_t> off;
    std::vector<int32_t> idx;
};

Grid2 make_grid(int64_t C, const double* x, const double* y, double Lx, double Ly) {
    Grid2 g;
    g.Lx = Lx;
    g.Ly = Ly;
    const double h = std::sqrt(Lx * Ly / (double)std::max<int64_t>(C, 1));
    g.nx = std::max(1, (int)std::floor(Lx / h));
    g.ny = std::max(1, (int)std::floor(Ly / h));
    g.hx = Lx / g.nx;
    g.hy = Ly / g.ny;
    const int64_t nb = (int64_t)g.nx * g.ny;
    g.off.assign(nb + 1, 0);
    std::vector<int32_t> b(C);
    for(int64_t i = 0; i < C; i++) {
        int bx = std::min(g.nx - 1, std::max(0, (int)std::floor(x[i] / g.hx)));
        int by = std::min(g.ny - 1, std::max(0, (int)std::floor(y[i] / g.hy)));
        b[i] = by * g.nx + bx;
        g.off[b[i] + 1]++;
    }
    for(int64_t k = 0; k < nb; k++) g.off[k + 1] += g.off[k];
    g.idx.resize(C);
    std::vector<int64_t> fill(g.off.begin(), g.off.end() - 1);
    for(int64_t i = 0; i < C; i++) g.idx[fill[b[i]]++] = (int32_t)i;
    return g;
}

// Build cell i: candidates ring by ring until the security radius holds.
bool build_cell(const Grid2& g, int64_t C, const double* x, const double* y, int64_t i,
                Cell2& P) {
    const double xi = x[i], yi = y[i];
    const int bx = std::min(g.nx - 1, std::max(0, (int)std::floor(xi / g.hx)));
    const int by = std::min(g.ny - 1, std::max(0, (int)std::floor(yi / g.hy)));
    const double half = 0.5 * std::max(g.Lx, g.Ly) * 2.0;
    P.vx = {xi - half, xi + half, xi + half, xi - half};
    P.vy = {yi - half, yi - half, yi + half, yi + half};
    P.lab = {-1, -1, -1, -1};
    P.cand.clear();
    const int rmax = std::max(g.nx, g.ny) + 2;
    auto visit = [&](int dx, int dy) {
        int cx = bx + dx, cy = by + dy;
        int wx = ((cx % g.nx) + g.nx) % g.nx, wy = ((cy % g.ny) + g.ny) % g.ny;
        double sx = (double)((cx - wx) / g.nx) * g.Lx;
        double sy = (double)((cy - wy) / g.ny) * g.Ly;
        const int64_t b = (int64_t)wy * g.nx + wx;
        for(int64_t k = g.off[b]; k < g.off[b + 1]; k++) {
            const int32_t j = g.idx[k];
            if(j == i && sx == 0.0 && sy == 0.0) continue;
            Cand q{x[j] + sx, y[j] + sy, j, sx, sy};
            P.cand.push_back(q);
            clip(P, q, xi, yi, (int32_t)P.cand.size() - 1);
        }
    };
    for(int r = 0; r <= rmax; r++) {
        if(r == 0) {
            visit(0, 0);
        } else {
            for(int d = -r; d <= r; d++) {
                visit(d, -r);
                visit(d, r);
            }
            for(int d = -r + 1; d <= r - 1; d++) {
                visit(-r, d);
                visit(r, d);
            }
        }
        // security radius: generators outside ring r are at least r*h away
        double R2 = 0.0;
        for(size_t k = 0; k < P.vx.size(); k++) {
            double dx = P.vx[k] - xi, dy = P.vy[k] - yi;
            R2 = std::max(R2, dx * dx + dy * dy);
        }
        const double reach = r * std::min(g.hx, g.hy);
        bool boxed = false;
        for(int32_t l : P.lab) boxed |= l < 0;
        if(!boxed && 4.0 * R2 <= reach * reach) return true;
    }
    for(int32_t l : P.lab)
        if(l < 0) return false;
    return true;
}

}  // namespace

int voronoi2d_periodic(int64_t C, const double* pos, double Lx, double Ly, HostMesh& m,
                       int nthreads) {
    const double* x = pos;
    const double* y = pos + C;
    m = HostMesh();
    m.ndim = 2;
    m.C = C;
    Grid2 g = make_grid(C, x, y, Lx, Ly);
    std::vector<Cell2> cells(C);
    if(nthreads <= 0) nthreads = (int)std::max(1u, std::thread::hardware_concurrency());
    std::atomic<int64_t> next(0);
    std::atomic<bool> ok(true);
    auto work = [&]() {
        for(;;) {
            int64_t i0 = next.fetch_add(256);
            if(i0 >= C) break;
            for(int64_t i = i0; i < std::min(C, i0 + 256); i++) {
                if(!build_cell(g, C, x, y, i, cells[i])) ok = false;
                // rotate so the first edge is the one whose neighbour direction
                // has the smallest angle in [-pi, pi): a deterministic start
                Cell2& P = cells[i];
                const size_t n = P.vx.size();
                size_t best = 0;
                double ba = 1e300;
                for(size_t k = 0; k < n; k++) {
                    const Cand& q = P.cand[P.lab[k]];
                    double a = std::atan2(q.y - y[i], q.x - x[i]);
                    if(a < ba) {
                        ba = a;
                        best = k;
                    }
                }
                std::rotate(P.vx.begin(), P.vx.begin() + best, P.vx.end());
                std::rotate(P.vy.begin(), P.vy.begin() + best, P.vy.end());
                std::rotate(P.lab.begin(), P.lab.begin() + best, P.lab.end());
            }
        }
    };
    std::vector<std::thread> th;
    for(int t = 0; t < nthreads; t++) th.emplace_back(work);
    for(auto& t : th) t.join();
    if(!ok) return -1;

    // faces, in cell order; a face belongs to the first real cell visiting it
    m.pos.assign(pos, pos + 2 * C);
    m.cf_off.assign(C + 1, 0);
    std::vector<int64_t> first_face(C + 1, 0);
    for(int64_t i = 0; i < C; i++) {
        const Cell2& P = cells[i];
        int64_t nnew = 0;
        for(int32_t l : P.lab) {
            const Cand& q = P.cand[l];
            const bool ghost = q.sx != 0.0 || q.sy != 0.0;
            if(ghost || q.j > i) nnew++;
        }
        first_face[i + 1] = first_face[i] + nnew;
        m.cf_off[i + 1] = m.cf_off[i] + (int64_t)P.lab.size();
    }
    const int64_t F = first_face[C];
    m.F = F;
    m.left.resize(F);
    m.right.resize(F);
    m.area.resize(F);
    m.midpoint.resize(2 * F);
    m.shift.resize(2 * F);
    m.ghost.resize(F);
    m.twin.assign(F, -1);
    m.verts.resize(4 * F);
    m.cf_face.resize(m.cf_off[C]);
    // face index of (creator i, neighbour j) for real-real faces
    std::vector<std::unordered_map<int32_t, int64_t>> owned(C);
    for(int64_t i = 0; i < C; i++) {
        const Cell2& P = cells[i];
        int64_t f = first_face[i];
        const size_t n = P.vx.size();
        for(size_t k = 0; k < n; k++) {
            const Cand& q = P.cand[P.lab[k]];
            const bool ghost = q.sx != 0.0 || q.sy != 0.0;
            if(!(ghost || q.j > i)) continue;
            const size_t k1 = (k + 1) % n;
            m.left[f] = (int32_t)i;
            m.right[f] = q.j;
            m.ghost[f] = ghost ? 1 : 0;
            m.shift[f] = q.sx;
            m.shift[F + f] = q.sy;
            const double v0x = P.vx[k], v0y = P.vy[k], v1x = P.vx[k1], v1y = P.vy[k1];
            m.verts[4 * f + 0] = v0x;
            m.verts[4 * f + 1] = v0y;
            m.verts[4 * f + 2] = v1x;
            m.verts[4 * f + 3] = v1y;
            double d0 = v0x - v1x, d1 = v0y - v1y;
            double area = std::sqrt(d0 * d0 + d1 * d1);
            if(std::sqrt((v0x - v1x) * (v0x - v1x) + (v0y - v1y) * (v0y - v1y)) < 1.e-13)
                area = 0.;
            m.area[f] = area;
            m.midpoint[f] = 0.5 * (v0x + v1x);
            m.midpoint[F + f] = 0.5 * (v0y + v1y);
            if(!ghost) owned[i][q.j] = f;
            f++;
        }
    }
    for(int64_t i = 0; i < C; i++) {
        const Cell2& P = cells[i];
        int64_t f = first_face[i];
        for(size_t k = 0; k < P.lab.size(); k++) {
            const Cand& q = P.cand[P.lab[k]];
            const bool ghost = q.sx != 0.0 || q.sy != 0.0;
            int64_t face;
            if(ghost || q.j > i) {
                face = f++;
            } else {
                auto it = owned[q.j].find((int32_t)i);
                if(it == owned[q.j].end()) return -2;  // inconsistent tessellation
                face = it->second;
            }
            m.cf_face[m.cf_off[i] + k] = (int32_t)face;
        }
    }
    // twins of periodic faces: (i -> j, s) pairs with (j -> i, -s)
    {
        std::unordered_map<uint64_t, std::vector<int64_t>> key;
        for(int64_t f = 0; f < F; f++) {
            if(!m.ghost[f]) continue;
            uint64_t a = (uint64_t)std::min(m.left[f], m.right[f]);
            uint64_t b = (uint64_t)std::max(m.left[f], m.right[f]);
            key[(a << 32) | b].push_back(f);
        }
        for(auto& kv : key) {
            auto& v = kv.second;
            for(size_t p = 0; p < v.size(); p++)
                for(size_t q = 0; q < v.size(); q++) {
                    int64_t f = v[p], h = v[q];
                    if(f == h || m.twin[f] >= 0 || m.twin[h] >= 0) continue;
                    if(m.left[f] == m.right[h] && m.right[f] == m.left[h] &&
                       m.shift[f] == -m.shift[h] && m.shift[F + f] == -m.shift[F + h]) {
                        m.twin[f] = h;
                        m.twin[h] = f;
                    }
                }
        }
    }
    // cell volume and centroid from the faces' stored vertices
    m.volume.resize(C);
    m.centroid.resize(2 * C);
    for(int64_t i = 0; i < C; i++) {
        const double cx = x[i], cy = y[i];
        const int64_t j0 = m.cf_off[i], j1 = m.cf_off[i + 1];
        // VorCell::calculate_volume, 2D (src/VorCell.cpp:375-385): faces in reverse
        double V = 0.;
        for(int64_t j = j1; j-- > j0;) {
            const int64_t f = m.cf_face[j];
            double ax = m.verts[4 * f + 0] - cx, ay = m.verts[4 * f + 1] - cy;
            double bxx = m.verts[4 * f + 2] - cx, byy = m.verts[4 * f + 3] - cy;
            // a face referenced from its right cell lies in the same frame
            // (only real-real faces are shared)
            V += 0.5 * std::fabs(ax * byy - ay * bxx);
        }
        m.volume[i] = V;
        // VorCell::calculate_centroid, 2D (src/VorCell.cpp:297-315)
        double Ct0 = 0, Ct1 = 0, At = 0;
        for(int64_t j = j0; j < j1; j++) {
            const int64_t f = m.cf_face[j];
            const double p0x = m.verts[4 * f + 0], p0y = m.verts[4 * f + 1];
            const double p1x = m.verts[4 * f + 2], p1y = m.verts[4 * f + 3];
            double c0 = (p0x + p1x + cx) / 3.;
            double c1 = (p0y + p1y + cy) / 3.;
            double ax = p0x - cx, ay = p0y - cy, bxx = p1x - cx, byy = p1y - cy;
            double A = 0.5 * std::fabs(ax * byy - ay * bxx);
            Ct0 += A * c0;
            Ct1 += A * c1;
            At += A;
        }
        m.centroid[i] = Ct0 / At;
        m.centroid[C + i] = Ct1 / At;
    }
    return 0;
}

}  // namespace sfx

namespace sfx {
int voronoi3d_periodic(int64_t, const double*, double, double, double, HostMesh&, int) {
    return -4;
}
}  // namespace sfx
