// mesh_voronoi3d.cpp -- host construction of 3D periodic Voronoi meshes.
//
// The reference builds its faces in VorTess::construct (Shadowfax
// src/VorTess.cpp:153-285, 3D branch) from a Delaunay tessellation whose
// predicates need Boost.Multiprecision.  Here each cell is built on its own:
// a box around the generator is clipped by the bisector planes of the
// neighbours, closest first (candidates from a grid hash over the periodic
// box), until every vertex lies within half the distance of the closest
// neighbour not yet considered (the security-radius test).  Cells are built
// in parallel; the face list then follows the reference's contract:
//   * cells in generator order;
//   * a face is created by the first of its two real cells to be visited
//     (its "left"); a neighbour that is a periodic copy (ghost, reference id
//     1, src/DelTess_mpi.cpp:151-170) never owns a face, so every periodic
//     boundary face exists twice, once from each real cell;
//   * the face keeps the vertex list of its creating cell, with the
//     reference's rule for near-coincident vertices (a vertex closer than
//     1e-13 to the previous or to the first one is skipped, a face with fewer
//     than 3 vertices gets area 0; src/VorTess.cpp:226-268);
//   * midpoint then area as VorFace::calculate_midpoint / calculate_area
//     (src/VorFace.cpp:293-338,369-403), volume and centroid from the faces'
//     stored vertex lists as VorCell::calculate_volume / calculate_centroid
//     (src/VorCell.cpp:257-296,355-374).
#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <thread>
#include <vector>

#include "mesh_assemble.h"
#include "sfx_mesh_host.h"

namespace sfx {
namespace {

struct Cand3 {
    double x, y, z;   // position of the (possibly shifted) generator
    double d2;        // squared distance to the cell's generator
    int32_t j;        // generator index
    int8_t s[3];      // periodic shift in box lengths
};

// Convex polyhedron: vertices, faces as index cycles (counter-clockwise seen
// from outside), each face labelled with the candidate whose bisector made it
// (-1: the initial box).
struct Poly {
    std::vector<double> v;          // xyz per vertex
    std::vector<int32_t> fidx;      // concatenated cycles
    std::vector<int32_t> foff;      // [nf+1]
    std::vector<int32_t> flab;      // [nf]
    // scratch
    std::vector<double> s, nv;
    std::vector<int32_t> remap, nidx, noff, nlab;
    std::vector<uint8_t> onp;
    std::vector<int64_t> ekey;
    std::vector<int32_t> eval, succ_from, succ_to;

    int nvert() const { return (int)(v.size() / 3); }
    int nface() const { return (int)flab.size(); }

    void box(double cx, double cy, double cz, double h) {
        v.clear();
        for(int k = 0; k < 8; k++) {
            v.push_back(cx + ((k & 1) ? h : -h));
            v.push_back(cy + ((k & 2) ? h : -h));
            v.push_back(cz + ((k & 4) ? h : -h));
        }
        // outward counter-clockwise cycles of the cube (vertex bit 0: x,
        // bit 1: y, bit 2: z)
        static const int32_t cyc[6][4] = {{0, 4, 6, 2}, {1, 3, 7, 5}, {0, 1, 5, 4},
                                          {2, 6, 7, 3}, {0, 2, 3, 1}, {4, 5, 7, 6}};
        fidx.clear();
        foff.assign(1, 0);
        flab.clear();
        for(int f = 0; f < 6; f++) {
            for(int k = 0; k < 4; k++) fidx.push_back(cyc[f][k]);
            foff.push_back((int32_t)fidx.size());
            flab.push_back(-1);
        }
    }

    double rmax2(double cx, double cy, double cz) const {
        double r = 0.0;
        for(size_t k = 0; k < v.size(); k += 3) {
            double dx = v[k] - cx, dy = v[k + 1] - cy, dz = v[k + 2] - cz;
            r = std::max(r, dx * dx + dy * dy + dz * dz);
        }
        return r;
    }

    int32_t edge_vertex(int32_t a, int32_t b) {
        int64_t key = (int64_t)std::min(a, b) << 32 | (int64_t)std::max(a, b);
        for(size_t k = 0; k < ekey.size(); k++)
            if(ekey[k] == key) return eval[k];
        const double t = s[a] / (s[a] - s[b]);
        const double* A = &v[3 * a];
        const double* B = &v[3 * b];
        int32_t id = (int32_t)(nv.size() / 3);
        nv.push_back(A[0] + t * (B[0] - A[0]));
        nv.push_back(A[1] + t * (B[1] - A[1]));
        nv.push_back(A[2] + t * (B[2] - A[2]));
        onp.push_back(1);
        ekey.push_back(key);
        eval.push_back(id);
        return id;
    }

    // Clip by {x : (x - m).u <= 0}, m = (c + q)/2, u = q - c.  Returns false
    // when the plane does not cut.
    bool clip(double cx, double cy, double cz, const Cand3& q, int32_t label) {
        const double ux = q.x - cx, uy = q.y - cy, uz = q.z - cz;
        const double mx = 0.5 * (q.x + cx), my = 0.5 * (q.y + cy), mz = 0.5 * (q.z + cz);
        const double eps = 1e-13 * (ux * ux + uy * uy + uz * uz);
        const int n = nvert();
        s.resize(n);
        bool any_out = false;
        for(int k = 0; k < n; k++) {
            s[k] = (v[3 * k] - mx) * ux + (v[3 * k + 1] - my) * uy + (v[3 * k + 2] - mz) * uz;
            any_out |= s[k] > eps;
        }
        if(!any_out) return false;
        // kept vertices: IN (s < -eps) and ON (|s| <= eps); ON ones lie on the plane
        nv.clear();
        onp.clear();
        remap.assign(n, -1);
        for(int k = 0; k < n; k++) {
            if(s[k] <= eps) {
                remap[k] = (int32_t)(nv.size() / 3);
                nv.push_back(v[3 * k]);
                nv.push_back(v[3 * k + 1]);
                nv.push_back(v[3 * k + 2]);
                onp.push_back(s[k] >= -eps ? 1 : 0);
            }
        }
        ekey.clear();
        eval.clear();
        nidx.clear();
        noff.assign(1, 0);
        nlab.clear();
        succ_from.clear();
        succ_to.clear();
        const int nf = nface();
        for(int f = 0; f < nf; f++) {
            const int32_t o0 = foff[f], o1 = foff[f + 1], len = o1 - o0;
            const size_t start = nidx.size();
            for(int32_t k = 0; k < len; k++) {
                const int32_t a = fidx[o0 + k], b = fidx[o0 + (k + 1) % len];
                const bool ka = s[a] <= eps, kb = s[b] <= eps;
                const bool ia = s[a] < -eps, ib = s[b] < -eps;
                if(ka) nidx.push_back(remap[a]);
                if((ia && !kb) || (!ka && ib)) nidx.push_back(edge_vertex(a, b));
            }
            const size_t cnt = nidx.size() - start;
            if(cnt < 3) {
                nidx.resize(start);
                continue;
            }
            // edges of this face lying in the cutting plane bound the new face,
            // traversed the other way
            for(size_t k = 0; k < cnt; k++) {
                const int32_t a = nidx[start + k], b = nidx[start + (k + 1) % cnt];
                if(onp[a] && onp[b] && a != b) {
                    succ_from.push_back(b);
                    succ_to.push_back(a);
                }
            }
            noff.push_back((int32_t)nidx.size());
            nlab.push_back(flab[f]);
        }
        // chain the new face
        const size_t ne = succ_from.size();
        const size_t start = nidx.size();
        bool chained = ne >= 3;
        if(chained) {
            int32_t cur = succ_from[0];
            for(size_t step = 0; step < ne; step++) {
                nidx.push_back(cur);
                size_t hit = ne, nhit = 0;
                for(size_t e = 0; e < ne; e++)
                    if(succ_from[e] == cur) {
                        hit = e;
                        nhit++;
                    }
                if(nhit != 1) {
                    chained = false;
                    break;
                }
                cur = succ_to[hit];
                if(cur == succ_from[0]) {
                    chained = step + 1 == ne;
                    break;
                }
            }
            if(chained && cur != succ_from[0]) chained = false;
        }
        if(!chained) {
            // fallback: every on-plane vertex, sorted by angle around the
            // normal u (counter-clockwise seen from outside)
            nidx.resize(start);
            std::vector<int32_t> pts;
            for(int32_t k = 0; k < (int32_t)onp.size(); k++)
                if(onp[k]) pts.push_back(k);
            if(pts.size() >= 3) {
                double ccx = 0, ccy = 0, ccz = 0;
                for(int32_t k : pts) {
                    ccx += nv[3 * k];
                    ccy += nv[3 * k + 1];
                    ccz += nv[3 * k + 2];
                }
                ccx /= pts.size();
                ccy /= pts.size();
                ccz /= pts.size();
                // orthonormal frame (e1, e2) in the plane with e1 x e2 = u/|u|
                double un = std::sqrt(ux * ux + uy * uy + uz * uz);
                double nx = ux / un, ny = uy / un, nz = uz / un;
                double ax = std::fabs(nx) < 0.9 ? 1.0 : 0.0, ay = ax == 1.0 ? 0.0 : 1.0, az = 0.0;
                double e1x = ay * nz - az * ny, e1y = az * nx - ax * nz, e1z = ax * ny - ay * nx;
                double e1n = std::sqrt(e1x * e1x + e1y * e1y + e1z * e1z);
                e1x /= e1n;
                e1y /= e1n;
                e1z /= e1n;
                double e2x = ny * e1z - nz * e1y, e2y = nz * e1x - nx * e1z,
                       e2z = nx * e1y - ny * e1x;
                std::vector<std::pair<double, int32_t>> ang;
                for(int32_t k : pts) {
                    double dx = nv[3 * k] - ccx, dy = nv[3 * k + 1] - ccy, dz = nv[3 * k + 2] - ccz;
                    ang.push_back({std::atan2(dx * e2x + dy * e2y + dz * e2z,
                                              dx * e1x + dy * e1y + dz * e1z),
                                   k});
                }
                std::sort(ang.begin(), ang.end());
                for(auto& p : ang) nidx.push_back(p.second);
            }
        }
        if(nidx.size() - start >= 3) {
            noff.push_back((int32_t)nidx.size());
            nlab.push_back(label);
        } else {
            nidx.resize(start);
        }
        // drop vertices no face references
        std::vector<int32_t>& used = remap;
        used.assign(nv.size() / 3, -1);
        int32_t nu = 0;
        for(int32_t& k : nidx)
            if(used[k] < 0) used[k] = nu++;
        v.resize(3 * (size_t)nu);
        for(int32_t k = 0; k < (int32_t)used.size(); k++)
            if(used[k] >= 0) {
                v[3 * used[k]] = nv[3 * k];
                v[3 * used[k] + 1] = nv[3 * k + 1];
                v[3 * used[k] + 2] = nv[3 * k + 2];
            }
        for(int32_t& k : nidx) k = used[k];
        fidx.swap(nidx);
        foff.swap(noff);
        flab.swap(nlab);
        return true;
    }
};

struct Grid3 {
    int n[3];
    double L[3], h[3];
    std::vector<int64_t> off;   // [ncell+1]
    std::vector<int32_t> idx;   // generators by grid cell
};

Grid3 make_grid(int64_t C, const double* const X[3], const double L[3]) {
    Grid3 g;
    const double vol = L[0] * L[1] * L[2];
    const double hh = std::cbrt(vol * 2.0 / (double)C);  // ~2 generators per cell
    int64_t ncell = 1;
    for(int d = 0; d < 3; d++) {
        g.L[d] = L[d];
        g.n[d] = std::max(1, (int)std::floor(L[d] / hh));
        g.h[d] = L[d] / g.n[d];
        ncell *= g.n[d];
    }
    g.off.assign(ncell + 1, 0);
    std::vector<int64_t> cell(C);
    for(int64_t i = 0; i < C; i++) {
        int64_t c = 0;
        for(int d = 2; d >= 0; d--) {
            int k = (int)std::floor(X[d][i] / g.h[d]);
            k = std::min(std::max(k, 0), g.n[d] - 1);
            c = c * g.n[d] + k;
        }
        cell[i] = c;
        g.off[c + 1]++;
    }
    for(int64_t c = 0; c < ncell; c++) g.off[c + 1] += g.off[c];
    g.idx.resize(C);
    std::vector<int64_t> fill(g.off.begin(), g.off.end() - 1);
    for(int64_t i = 0; i < C; i++) g.idx[fill[cell[i]]++] = (int32_t)i;
    return g;
}

bool build_cell(const Grid3& g, const double* const X[3], int64_t i, Poly& P,
                std::vector<Cand3>& cand, CellFaces<3>& out) {
    const double cx = X[0][i], cy = X[1][i], cz = X[2][i];
    const double c[3] = {cx, cy, cz};
    int gc[3];
    for(int d = 0; d < 3; d++) {
        gc[d] = (int)std::floor(c[d] / g.h[d]);
        gc[d] = std::min(std::max(gc[d], 0), g.n[d] - 1);
    }
    const double Lmax = std::max(g.L[0], std::max(g.L[1], g.L[2]));
    P.box(cx, cy, cz, Lmax);
    cand.clear();  // every candidate gathered so far; face labels index it
    size_t next = 0;         // candidates[next..] not yet tried, sorted
    double r2 = P.rmax2(cx, cy, cz);
    for(int k = 0; k < 64; k++) {
        // gather ring k (Chebyshev distance k in grid cells)
        for(int oz = -k; oz <= k; oz++)
            for(int oy = -k; oy <= k; oy++)
                for(int ox = -k; ox <= k; ox++) {
                    if(std::max(std::abs(ox), std::max(std::abs(oy), std::abs(oz))) != k)
                        continue;
                    const int o[3] = {ox, oy, oz};
                    int cc[3];
                    int8_t sh[3];
                    for(int d = 0; d < 3; d++) {
                        int t = gc[d] + o[d];
                        int w = (int)std::floor((double)t / g.n[d]);
                        cc[d] = t - w * g.n[d];
                        sh[d] = (int8_t)w;
                    }
                    const int64_t cell = ((int64_t)cc[2] * g.n[1] + cc[1]) * g.n[0] + cc[0];
                    for(int64_t p = g.off[cell]; p < g.off[cell + 1]; p++) {
                        const int32_t jj = g.idx[p];
                        if(jj == i && sh[0] == 0 && sh[1] == 0 && sh[2] == 0) continue;
                        Cand3 q;
                        q.x = X[0][jj] + sh[0] * g.L[0];
                        q.y = X[1][jj] + sh[1] * g.L[1];
                        q.z = X[2][jj] + sh[2] * g.L[2];
                        const double dx = q.x - cx, dy = q.y - cy, dz = q.z - cz;
                        q.d2 = dx * dx + dy * dy + dz * dz;
                        q.j = jj;
                        q.s[0] = sh[0];
                        q.s[1] = sh[1];
                        q.s[2] = sh[2];
                        cand.push_back(q);
                    }
                }
        std::sort(cand.begin() + next, cand.end(),
                  [](const Cand3& a, const Cand3& b) {
                      return a.d2 < b.d2 || (a.d2 == b.d2 && a.j < b.j);
                  });
        // radius around the generator fully covered by rings 0..k
        double D = 1e300;
        for(int d = 0; d < 3; d++) {
            D = std::min(D, c[d] - (gc[d] - k) * g.h[d]);
            D = std::min(D, (gc[d] + k + 1) * g.h[d] - c[d]);
        }
        const double D2 = D * D;
        while(next < cand.size() && cand[next].d2 < D2 && cand[next].d2 < 4.0 * r2) {
            if(P.clip(cx, cy, cz, cand[next], (int32_t)next)) r2 = P.rmax2(cx, cy, cz);
            next++;
        }
        if(4.0 * r2 <= D2) {
            // done: every generator within 2 r of the centre has been tried
            out.clear();
            for(int f = 0; f < P.nface(); f++) {
                const int32_t lab = P.flab[f];
                if(lab < 0) return false;  // box face survived: not enough neighbours
                const Cand3& q = cand[lab];
                out.j.push_back(q.j);
                out.sh.push_back(q.s[0]);
                out.sh.push_back(q.s[1]);
                out.sh.push_back(q.s[2]);
                // the reference's vertex filter (src/VorTess.cpp:226-262): skip a
                // vertex closer than 1e-13 to the previous one or to the first
                const int32_t o0 = P.foff[f], o1 = P.foff[f + 1];
                const double* first = &P.v[3 * P.fidx[o0]];
                const double* prev = nullptr;
                for(int32_t t = o0; t < o1; t++) {
                    const double* p = &P.v[3 * P.fidx[t]];
                    auto dist = [](const double* a, const double* b) {
                        double dx = a[0] - b[0], dy = a[1] - b[1], dz = a[2] - b[2];
                        return std::sqrt(dx * dx + dy * dy + dz * dz);
                    };
                    if(prev == nullptr || (dist(p, prev) > 1.e-13 && dist(p, first) > 1.e-13)) {
                        out.vx.push_back(p[0]);
                        out.vx.push_back(p[1]);
                        out.vx.push_back(p[2]);
                    }
                    prev = p;
                }
                out.voff.push_back((int32_t)(out.vx.size() / 3));
            }
            return true;
        }
    }
    return false;
}

}  // namespace

int voronoi3d_periodic(int64_t C, const double* pos, double Lx, double Ly, double Lz,
                       HostMesh& m, int nthreads) {
    const double* X[3] = {pos, pos + C, pos + 2 * C};
    const double L[3] = {Lx, Ly, Lz};
    for(int64_t i = 0; i < C; i++)
        for(int d = 0; d < 3; d++)
            if(!(X[d][i] >= 0.0 && X[d][i] < L[d])) return -3;
    m = HostMesh();
    m.ndim = 3;
    m.C = C;
    if(nthreads <= 0) nthreads = (int)std::max(1u, std::thread::hardware_concurrency());
    Grid3 g = make_grid(C, X, L);
    std::vector<CellFaces<3>> cells(C);
    std::atomic<bool> ok(true);
    {
        std::atomic<int64_t> next(0);
        std::vector<std::thread> th;
        for(int t = 0; t < nthreads; t++)
            th.emplace_back([&]() {
                Poly P;
                std::vector<Cand3> cand;
                for(;;) {
                    int64_t i0 = next.fetch_add(128);
                    if(i0 >= C) break;
                    for(int64_t i = i0; i < std::min(C, i0 + 128); i++)
                        if(!build_cell(g, X, i, P, cand, cells[i])) ok = false;
                }
            });
        for(auto& t : th) t.join();
    }
    if(!ok) return -1;
    return assemble<3>(C, pos, L, cells, m, nthreads);
}

}  // namespace sfx
