// mesh_voronoi.cpp -- host construction of 2D periodic Voronoi meshes.
//
// The reference builds its faces in VorTess::construct (Shadowfax
// src/VorTess.cpp:285-396) from a Delaunay tessellation that needs
// Boost.Multiprecision.  Here each cell is built independently by clipping a
// box with the bisector half-planes of its neighbours, closest first
// (candidates from a grid hash over the periodic box); the face list then
// follows the reference's contract (mesh_assemble.h):
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
#include <thread>
#include <vector>

#include "mesh_assemble.h"
#include "sfx_mesh_host.h"

namespace sfx {
namespace {

struct Cand2 {
    double x, y;  // position of the (possibly shifted) generator
    double d2;    // squared distance to the cell's generator
    int32_t j;    // generator index
    int8_t s[2];  // periodic shift in box lengths
};

// Convex polygon, counter-clockwise; lab[k]: candidate whose bisector made
// the edge from vertex k to k+1 (-1: the initial box).
struct Poly2 {
    std::vector<double> vx, vy;
    std::vector<int32_t> lab;
    std::vector<double> s, nx, ny;  // scratch
    std::vector<int32_t> nl;

    double rmax2(double xi, double yi) const {
        double r = 0.0;
        for(size_t k = 0; k < vx.size(); k++) {
            double dx = vx[k] - xi, dy = vy[k] - yi;
            r = std::max(r, dx * dx + dy * dy);
        }
        return r;
    }

    // clip by {x : (x - m).u <= 0}, m = (c + q)/2, u = q - c
    bool clip(double xi, double yi, const Cand2& q, int32_t label) {
        const double ux = q.x - xi, uy = q.y - yi;
        const double mx = 0.5 * (q.x + xi), my = 0.5 * (q.y + yi);
        const size_t n = vx.size();
        s.resize(n);
        bool any_out = false;
        for(size_t k = 0; k < n; k++) {
            s[k] = (vx[k] - mx) * ux + (vy[k] - my) * uy;
            any_out |= s[k] > 0.0;
        }
        if(!any_out) return false;
        nx.clear();
        ny.clear();
        nl.clear();
        for(size_t k = 0; k < n; k++) {
            const size_t k1 = (k + 1) % n;
            const bool in0 = s[k] <= 0.0, in1 = s[k1] <= 0.0;
            if(in0) {
                nx.push_back(vx[k]);
                ny.push_back(vy[k]);
                nl.push_back(lab[k]);
            }
            if(in0 != in1) {
                const double t = s[k] / (s[k] - s[k1]);
                nx.push_back(vx[k] + t * (vx[k1] - vx[k]));
                ny.push_back(vy[k] + t * (vy[k1] - vy[k]));
                // leaving the half-plane: the new edge runs along the bisector
                nl.push_back(in0 ? label : lab[k]);
            }
        }
        vx.swap(nx);
        vy.swap(ny);
        lab.swap(nl);
        return true;
    }
};

struct Grid2 {
    int n[2];
    double L[2], h[2];
    std::vector<int64_t> off;
    std::vector<int32_t> idx;
};

Grid2 make_grid(int64_t C, const double* x, const double* y, double Lx, double Ly) {
    Grid2 g;
    g.L[0] = Lx;
    g.L[1] = Ly;
    const double hh = std::sqrt(Lx * Ly * 2.0 / (double)std::max<int64_t>(C, 1));
    for(int d = 0; d < 2; d++) {
        g.n[d] = std::max(1, (int)std::floor(g.L[d] / hh));
        g.h[d] = g.L[d] / g.n[d];
    }
    const int64_t nb = (int64_t)g.n[0] * g.n[1];
    g.off.assign(nb + 1, 0);
    std::vector<int64_t> b(C);
    for(int64_t i = 0; i < C; i++) {
        int bx = std::min(g.n[0] - 1, std::max(0, (int)std::floor(x[i] / g.h[0])));
        int by = std::min(g.n[1] - 1, std::max(0, (int)std::floor(y[i] / g.h[1])));
        b[i] = (int64_t)by * g.n[0] + bx;
        g.off[b[i] + 1]++;
    }
    for(int64_t k = 0; k < nb; k++) g.off[k + 1] += g.off[k];
    g.idx.resize(C);
    std::vector<int64_t> fill(g.off.begin(), g.off.end() - 1);
    for(int64_t i = 0; i < C; i++) g.idx[fill[b[i]]++] = (int32_t)i;
    return g;
}

// Build cell i: candidates ring by ring (Chebyshev distance in grid cells),
// closest first, until every vertex lies within half the distance covered
// by the rings gathered so far (the security-radius test).
bool build_cell(const Grid2& g, const double* x, const double* y, int64_t i, Poly2& P,
                std::vector<Cand2>& cand, CellFaces<2>& out) {
    const double xi = x[i], yi = y[i];
    const double c[2] = {xi, yi};
    int gc[2];
    for(int d = 0; d < 2; d++)
        gc[d] = std::min(g.n[d] - 1, std::max(0, (int)std::floor(c[d] / g.h[d])));
    const double half = std::max(g.L[0], g.L[1]);
    P.vx = {xi - half, xi + half, xi + half, xi - half};
    P.vy = {yi - half, yi - half, yi + half, yi + half};
    P.lab = {-1, -1, -1, -1};
    cand.clear();
    size_t next = 0;
    double r2 = P.rmax2(xi, yi);
    for(int k = 0; k < 64; k++) {
        for(int oy = -k; oy <= k; oy++)
            for(int ox = -k; ox <= k; ox++) {
                if(std::max(std::abs(ox), std::abs(oy)) != k) continue;
                const int o[2] = {ox, oy};
                int cc[2];
                int8_t sh[2];
                for(int d = 0; d < 2; d++) {
                    int t = gc[d] + o[d];
                    int w = (int)std::floor((double)t / g.n[d]);
                    cc[d] = t - w * g.n[d];
                    sh[d] = (int8_t)w;
                }
                const int64_t cell = (int64_t)cc[1] * g.n[0] + cc[0];
                for(int64_t p = g.off[cell]; p < g.off[cell + 1]; p++) {
                    const int32_t jj = g.idx[p];
                    if(jj == i && sh[0] == 0 && sh[1] == 0) continue;
                    Cand2 q;
                    q.x = x[jj] + sh[0] * g.L[0];
                    q.y = y[jj] + sh[1] * g.L[1];
                    const double dx = q.x - xi, dy = q.y - yi;
                    q.d2 = dx * dx + dy * dy;
                    q.j = jj;
                    q.s[0] = sh[0];
                    q.s[1] = sh[1];
                    cand.push_back(q);
                }
            }
        std::sort(cand.begin() + next, cand.end(), [](const Cand2& a, const Cand2& b) {
            return a.d2 < b.d2 || (a.d2 == b.d2 && a.j < b.j);
        });
        double D = 1e300;
        for(int d = 0; d < 2; d++) {
            D = std::min(D, c[d] - (gc[d] - k) * g.h[d]);
            D = std::min(D, (gc[d] + k + 1) * g.h[d] - c[d]);
        }
        const double D2 = D * D;
        while(next < cand.size() && cand[next].d2 < D2 && cand[next].d2 < 4.0 * r2) {
            if(P.clip(xi, yi, cand[next], (int32_t)next)) r2 = P.rmax2(xi, yi);
            next++;
        }
        if(4.0 * r2 <= D2) {
            const size_t n = P.vx.size();
            for(size_t e = 0; e < n; e++)
                if(P.lab[e] < 0) return false;  // box edge survived
            // start at the edge whose neighbour direction has the smallest
            // angle in [-pi, pi): a deterministic first face
            size_t best = 0;
            double ba = 1e300;
            for(size_t e = 0; e < n; e++) {
                const Cand2& q = cand[P.lab[e]];
                double a = std::atan2(q.y - yi, q.x - xi);
                if(a < ba) {
                    ba = a;
                    best = e;
                }
            }
            out.clear();
            for(size_t t = 0; t < n; t++) {
                const size_t e = (best + t) % n, e1 = (e + 1) % n;
                const Cand2& q = cand[P.lab[e]];
                out.j.push_back(q.j);
                out.sh.push_back(q.s[0]);
                out.sh.push_back(q.s[1]);
                out.vx.push_back(P.vx[e]);
                out.vx.push_back(P.vy[e]);
                out.vx.push_back(P.vx[e1]);
                out.vx.push_back(P.vy[e1]);
                out.voff.push_back((int32_t)(out.vx.size() / 2));
            }
            return true;
        }
    }
    return false;
}

}  // namespace

int voronoi2d_periodic(int64_t C, const double* pos, double Lx, double Ly, HostMesh& m,
                       int nthreads) {
    const double* x = pos;
    const double* y = pos + C;
    const double L[2] = {Lx, Ly};
    for(int64_t i = 0; i < C; i++)
        if(!(x[i] >= 0.0 && x[i] < Lx && y[i] >= 0.0 && y[i] < Ly)) return -3;
    m = HostMesh();
    if(nthreads <= 0) nthreads = (int)std::max(1u, std::thread::hardware_concurrency());
    Grid2 g = make_grid(C, x, y, Lx, Ly);
    std::vector<CellFaces<2>> cells(C);
    std::atomic<bool> ok(true);
    {
        std::atomic<int64_t> next(0);
        std::vector<std::thread> th;
        for(int t = 0; t < nthreads; t++)
            th.emplace_back([&]() {
                Poly2 P;
                std::vector<Cand2> cand;
                for(;;) {
                    int64_t i0 = next.fetch_add(256);
                    if(i0 >= C) break;
                    for(int64_t i = i0; i < std::min(C, i0 + 256); i++)
                        if(!build_cell(g, x, y, i, P, cand, cells[i])) ok = false;
                }
            });
        for(auto& t : th) t.join();
    }
    if(!ok) return -1;
    return assemble<2>(C, pos, L, cells, m, nthreads);
}

}  // namespace sfx
