// mesh_assemble.h -- from independently built Voronoi cells to the face list
// the flux pass reads, following the reference's VorTess contract
// (Shadowfax src/VorTess.cpp:153-396):
//   * cells in generator order, each cell's faces in the order its builder
//     produced them;
//   * a face is created by the first of its two real cells to be visited (its
//     "left"); a neighbour that is a periodic copy (ghost, reference id 1,
//     src/DelTess_mpi.cpp:151-170) never owns a face, so every periodic
//     boundary face exists twice, once from each real cell;
//   * the face keeps the vertex list of its creating cell; cell volume and
//     centroid are computed from the faces' stored vertex lists.
// Each cell is built independently, so round-off can leave a sliver face on
// one side of a real-real pair only; such a face is dropped from both sides.
// Host code; all loops over cells and faces run on nthreads threads.
#pragma once

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdint>
#include <thread>
#include <vector>

#include "sfx_mesh_host.h"

namespace sfx {

// One cell as its builder leaves it: for each face the neighbour generator,
// its periodic shift in box lengths, and the face's vertex list.
template <int ND>
struct CellFaces {
    std::vector<int32_t> j;
    std::vector<int8_t> sh;     // ND per face
    std::vector<int32_t> voff;  // [nf+1] into vx
    std::vector<double> vx;     // ND per vertex
    void clear() {
        j.clear();
        sh.clear();
        voff.assign(1, 0);
        vx.clear();
    }
    bool ghost(size_t f) const {
        for(int d = 0; d < ND; d++)
            if(sh[ND * f + d]) return true;
        return false;
    }
};

template <class Fn>
void parallel_for(int64_t n, int nthreads, Fn fn) {
    std::atomic<int64_t> next(0);
    std::vector<std::thread> th;
    for(int t = 0; t < nthreads; t++)
        th.emplace_back([&]() {
            for(;;) {
                int64_t i0 = next.fetch_add(256);
                if(i0 >= n) break;
                for(int64_t i = i0; i < std::min(n, i0 + 256); i++) fn(i);
            }
        });
    for(auto& t : th) t.join();
}

// VorFace::calculate_area / calculate_midpoint (src/VorFace.cpp:293-403)
// with the construction-time rules of src/VorTess.cpp:262-268,336-343.
inline void face_geometry_2d(const double* v, int nv, double* mid, double* area) {
    const double v0x = v[0], v0y = v[1], v1x = v[2], v1y = v[3];
    (void)nv;
    double d0 = v0x - v1x, d1 = v0y - v1y;
    double a = std::sqrt(d0 * d0 + d1 * d1);
    // try to solve problems with accuracy (VorTess.cpp:340-342)
    if(std::sqrt((v0x - v1x) * (v0x - v1x) + (v0y - v1y) * (v0y - v1y)) < 1.e-13) a = 0.;
    *area = a;
    mid[0] = 0.5 * (v0x + v1x);
    mid[1] = 0.5 * (v0y + v1y);
}

inline void face_geometry_3d(const double* vx, int nvx, double* mid, double* area) {
    if(nvx < 3) {
        // a face with fewer than 3 vertices is not used (area 0,
        // src/VorTess.cpp:262-264); its midpoint is the vertices' mean
        double m[3] = {0, 0, 0};
        for(int k = 0; k < nvx; k++)
            for(int d = 0; d < 3; d++) m[d] += vx[3 * k + d];
        const double nn = nvx > 0 ? (double)nvx : 1.0;
        for(int d = 0; d < 3; d++) mid[d] = m[d] / nn;
        *area = 0.;
        return;
    }
    const double* v0 = vx;
    double a = 0., m0 = 0., m1 = 0., m2 = 0.;
    for(int i = 1; i < nvx - 1; i++) {
        const double* vi = vx + 3 * i;
        const double* vj = vx + 3 * (i + 1);
        double ab[3] = {vi[0] - vj[0], vi[1] - vj[1], vi[2] - vj[2]};
        double ac[3] = {vi[0] - v0[0], vi[1] - v0[1], vi[2] - v0[2]};
        double ab2 = ab[0] * ab[0] + ab[1] * ab[1] + ab[2] * ab[2];
        double ac2 = ac[0] * ac[0] + ac[1] * ac[1] + ac[2] * ac[2];
        double abac = ab[0] * ac[0] + ab[1] * ac[1] + ab[2] * ac[2];
        double tri = 0.5 * std::sqrt(std::fabs(ab2 * ac2 - abac * abac));
        a += tri;
        m0 += (vi[0] + vj[0] + v0[0]) * tri;
        m1 += (vi[1] + vj[1] + v0[1]) * tri;
        m2 += (vi[2] + vj[2] + v0[2]) * tri;
    }
    m0 /= 3. * a;
    m1 /= 3. * a;
    m2 /= 3. * a;
    mid[0] = m0;
    mid[1] = m1;
    mid[2] = m2;
    if(m0 != m0) {
        *area = 0.;
        return;
    }
    double A = 0.;
    for(int i = 0; i < nvx; i++) {
        const double* vi = vx + 3 * i;
        const double* vj = vx + 3 * ((i + 1) % nvx);
        double ab[3] = {vi[0] - vj[0], vi[1] - vj[1], vi[2] - vj[2]};
        double ac[3] = {vi[0] - m0, vi[1] - m1, vi[2] - m2};
        double ab2 = ab[0] * ab[0] + ab[1] * ab[1] + ab[2] * ab[2];
        double ac2 = ac[0] * ac[0] + ac[1] * ac[1] + ac[2] * ac[2];
        double abac = ab[0] * ac[0] + ab[1] * ac[1] + ab[2] * ac[2];
        A += 0.5 * std::sqrt(std::fabs(ab2 * ac2 - abac * abac));
    }
    if(A != A) A = 0.;
    *area = A;
}

// VorCell::calculate_centroid / calculate_volume (src/VorCell.cpp:257-396)
// over the cell's faces (cf list order) and their stored vertex lists.
inline void cell_geometry_2d(const HostMesh& m, int64_t i, const double* cp, double* vol,
                             double* cen) {
    const int64_t C = m.C, NV = m.fv_off[m.F];
    const int64_t j0 = m.cf_off[i], j1 = m.cf_off[i + 1];
    auto P = [&](int64_t f, int t, int d) { return m.fverts[d * NV + m.fv_off[f] + t]; };
    (void)C;
    double V = 0.;
    for(int64_t j = j1; j-- > j0;) {  // 2D volume loops the faces in reverse
        const int64_t f = m.cf_face[j];
        double ax = P(f, 0, 0) - cp[0], ay = P(f, 0, 1) - cp[1];
        double bx = P(f, 1, 0) - cp[0], by = P(f, 1, 1) - cp[1];
        V += 0.5 * std::fabs(ax * by - ay * bx);
    }
    *vol = V;
    double Ct0 = 0, Ct1 = 0, At = 0;
    for(int64_t j = j0; j < j1; j++) {
        const int64_t f = m.cf_face[j];
        const double p0x = P(f, 0, 0), p0y = P(f, 0, 1), p1x = P(f, 1, 0), p1y = P(f, 1, 1);
        double c0 = (p0x + p1x + cp[0]) / 3.;
        double c1 = (p0y + p1y + cp[1]) / 3.;
        double ax = p0x - cp[0], ay = p0y - cp[1], bx = p1x - cp[0], by = p1y - cp[1];
        double A = 0.5 * std::fabs(ax * by - ay * bx);
        Ct0 += A * c0;
        Ct1 += A * c1;
        At += A;
    }
    cen[0] = Ct0 / At;
    cen[1] = Ct1 / At;
}

inline void cell_geometry_3d(const HostMesh& m, int64_t i, const double* cp, double* vol,
                             double* cen) {
    const int64_t NV = m.fv_off[m.F];
    double Ct[3] = {0, 0, 0}, At = 0, V = 0;
    auto tet = [&](int64_t f, int64_t j, double* Cc) {
        const int64_t v0 = m.fv_off[f];
        auto P = [&](int64_t t, int d) { return m.fverts[d * NV + v0 + t]; };
        if(Cc)
            for(int d = 0; d < 3; d++) Cc[d] = 0.25 * (P(0, d) + P(j, d) + P(j + 1, d) + cp[d]);
        double r1[3], r2[3], r3[3];
        for(int d = 0; d < 3; d++) {
            r1[d] = P(0, d) - P(j, d);
            r2[d] = P(j, d) - P(j + 1, d);
            r3[d] = P(j + 1, d) - cp[d];
        }
        double A = std::fabs(r1[0] * r2[1] * r3[2] + r1[1] * r2[2] * r3[0] +
                             r1[2] * r2[0] * r3[1] - r1[2] * r2[1] * r3[0] -
                             r2[2] * r3[1] * r1[0] - r3[2] * r1[1] * r2[0]);
        return A / 6;
    };
    for(int64_t e = m.cf_off[i]; e < m.cf_off[i + 1]; e++) {
        const int64_t f = m.cf_face[e];
        const int64_t nvx = m.fv_off[f + 1] - m.fv_off[f];
        for(int64_t j = 1; j + 1 < nvx; j++) {
            double Cc[3];
            double A = tet(f, j, Cc);
            Ct[0] += Cc[0] * A;
            Ct[1] += Cc[1] * A;
            Ct[2] += Cc[2] * A;
            At += A;
        }
    }
    for(int64_t e = m.cf_off[i]; e < m.cf_off[i + 1]; e++) {
        const int64_t f = m.cf_face[e];
        const int64_t nvx = m.fv_off[f + 1] - m.fv_off[f];
        for(int64_t j = 1; j + 1 < nvx; j++) V += tet(f, j, nullptr);
    }
    *vol = V;
    for(int d = 0; d < 3; d++) cen[d] = Ct[d] / At;
}

template <int ND>
bool has_real_neighbour(const CellFaces<ND>& o, int32_t j) {
    for(size_t f = 0; f < o.j.size(); f++)
        if(o.j[f] == j && !o.ghost(f)) return true;
    return false;
}

// Returns 0, or -2 (inconsistent tessellation), -5 (too many faces for
// 32-bit indices).
template <int ND>
int assemble(int64_t C, const double* pos, const double* L, std::vector<CellFaces<ND>>& cells,
             HostMesh& m, int nthreads) {
    m.ndim = ND;
    m.C = C;
    std::vector<std::vector<uint8_t>> keep(C);
    parallel_for(C, nthreads, [&](int64_t i) {
        const CellFaces<ND>& o = cells[i];
        keep[i].assign(o.j.size(), 1);
        for(size_t f = 0; f < o.j.size(); f++)
            if(!o.ghost(f) && !has_real_neighbour(cells[o.j[f]], (int32_t)i)) keep[i][f] = 0;
    });
    auto creates = [&](int64_t i, size_t f) {
        return cells[i].ghost(f) || cells[i].j[f] > (int32_t)i;
    };
    m.pos.assign(pos, pos + ND * C);
    m.cf_off.assign(C + 1, 0);
    std::vector<int64_t> first_face(C + 1, 0), first_vert(C + 1, 0);
    parallel_for(C, nthreads, [&](int64_t i) {
        const CellFaces<ND>& o = cells[i];
        int64_t nnew = 0, nk = 0, nvx = 0;
        for(size_t f = 0; f < o.j.size(); f++) {
            if(!keep[i][f]) continue;
            nk++;
            if(creates(i, f)) {
                nnew++;
                nvx += o.voff[f + 1] - o.voff[f];
            }
        }
        first_face[i + 1] = nnew;
        first_vert[i + 1] = nvx;
        m.cf_off[i + 1] = nk;
    });
    for(int64_t i = 0; i < C; i++) {
        first_face[i + 1] += first_face[i];
        first_vert[i + 1] += first_vert[i];
        m.cf_off[i + 1] += m.cf_off[i];
    }
    const int64_t F = first_face[C];
    if(F >= INT32_MAX || m.cf_off[C] >= INT32_MAX) return -5;
    const int64_t NV = first_vert[C];
    m.F = F;
    m.left.resize(F);
    m.right.resize(F);
    m.area.resize(F);
    m.midpoint.resize(ND * F);
    m.shift.resize(ND * F);
    m.ghost.resize(F);
    m.twin.assign(F, -1);
    m.fv_off.resize(F + 1);
    m.fv_off[F] = NV;
    m.fverts.resize(ND * NV);
    m.cf_face.resize(m.cf_off[C]);
    parallel_for(C, nthreads, [&](int64_t i) {
        const CellFaces<ND>& o = cells[i];
        int64_t f = first_face[i], vo = first_vert[i];
        for(size_t k = 0; k < o.j.size(); k++) {
            if(!keep[i][k] || !creates(i, k)) continue;
            m.left[f] = (int32_t)i;
            m.right[f] = o.j[k];
            m.ghost[f] = o.ghost(k) ? 1 : 0;
            for(int d = 0; d < ND; d++) m.shift[d * F + f] = o.sh[ND * k + d] * L[d];
            const int nvx = o.voff[k + 1] - o.voff[k];
            m.fv_off[f] = vo;
            for(int t = 0; t < nvx; t++)
                for(int d = 0; d < ND; d++)
                    m.fverts[d * NV + vo + t] = o.vx[ND * (o.voff[k] + t) + d];
            double mid[ND], area;
            if(ND == 2)
                face_geometry_2d(&o.vx[2 * o.voff[k]], nvx, mid, &area);
            else
                face_geometry_3d(&o.vx[3 * o.voff[k]], nvx, mid, &area);
            m.area[f] = area;
            for(int d = 0; d < ND; d++) m.midpoint[d * F + f] = mid[d];
            vo += nvx;
            f++;
        }
    });
    // per-cell face lists; a real-real face created by j < i is found among
    // j's created faces
    std::atomic<bool> consistent(true);
    parallel_for(C, nthreads, [&](int64_t i) {
        const CellFaces<ND>& o = cells[i];
        int64_t f = first_face[i], e = m.cf_off[i];
        for(size_t k = 0; k < o.j.size(); k++) {
            if(!keep[i][k]) continue;
            int64_t face = -1;
            if(creates(i, k)) {
                face = f++;
            } else {
                const int32_t jj = o.j[k];
                for(int64_t h = first_face[jj]; h < first_face[jj + 1]; h++)
                    if(m.right[h] == (int32_t)i && !m.ghost[h]) {
                        face = h;
                        break;
                    }
                if(face < 0) {
                    consistent = false;
                    face = 0;
                }
            }
            m.cf_face[e++] = (int32_t)face;
        }
    });
    if(!consistent) return -2;
    // twins of periodic faces: (i -> j, s) with (j -> i, -s)
    parallel_for(F, nthreads, [&](int64_t f) {
        if(!m.ghost[f]) return;
        const int32_t i = m.left[f], jj = m.right[f];
        for(int64_t h = first_face[jj]; h < first_face[jj + 1]; h++) {
            if(!m.ghost[h] || m.right[h] != i || h == f) continue;
            bool opp = true;
            for(int d = 0; d < ND; d++) opp &= m.shift[d * F + h] == -m.shift[d * F + f];
            if(opp) {
                m.twin[f] = (int32_t)h;
                break;
            }
        }
    });
    cells.clear();
    cells.shrink_to_fit();
    m.volume.resize(C);
    m.centroid.resize(ND * C);
    parallel_for(C, nthreads, [&](int64_t i) {
        double cp[ND], cen[ND], vol;
        for(int d = 0; d < ND; d++) cp[d] = pos[d * C + i];
        if(ND == 2)
            cell_geometry_2d(m, i, cp, &vol, cen);
        else
            cell_geometry_3d(m, i, cp, &vol, cen);
        m.volume[i] = vol;
        for(int d = 0; d < ND; d++) m.centroid[d * C + i] = cen[d];
    });
    return 0;
}

}  // namespace sfx
