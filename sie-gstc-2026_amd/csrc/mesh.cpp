// Host-side mesh upload and per-face precomputation for the assembly kernels.
#include "sie_internal.hpp"
#include "dunavant_table.h"

#include <cmath>

namespace sie {

void quad_host(int rule, std::vector<float>& xy, std::vector<float>& w) {
    require(rule >= 1 && rule <= 20, "Dunavant quadrature rule must be in [1, 20]");
    const int n = kDunavantOrder[rule], off = kDunavantOffset[rule];
    xy.assign(kDunavantXY + 2 * off, kDunavantXY + 2 * off + 2 * n);
    w.assign(kDunavantW + off, kDunavantW + off + n);
}

void quad_upload(QuadDev& q, int rule, hipStream_t s) {
    if (q.rule == rule && q.n > 0) return;
    std::vector<float> xy, w;
    quad_host(rule, xy, w);
    q.xy.alloc(xy.size());
    q.w.alloc(w.size());
    q.xy.upload(xy.data(), xy.size(), s);
    q.w.upload(w.data(), w.size(), s);
    SIE_HIP(hipStreamSynchronize(s));
    q.rule = rule;
    q.n = (int)w.size();
}

void build_mesh(Mesh& m, const double* V, int nv, const int* F, int nf, const int* FE, int ne,
                hipStream_t s) {
    require(V && F && FE, "SIE mesh must contain vertices, faces, and edges");
    require(nv > 0 && nf > 0 && ne > 0, "SIE mesh must contain vertices, faces, and edges");
    m.nv = nv; m.nf = nf; m.ne = ne;
    m.V.assign(V, V + 3 * (size_t)nv);
    m.F.assign(F, F + 3 * (size_t)nf);
    m.FE.assign(FE, FE + 3 * (size_t)nf);
    for (int k = 0; k < 2; ++k) { m.ef[k].assign(ne, -1); m.el[k].assign(ne, -1); }
    for (int f = 0; f < nf; ++f)
        for (int j = 0; j < 3; ++j) {
            const int v = F[3 * f + j];
            require(v >= 0 && v < nv, "face references a vertex index out of range");
            const int e = FE[3 * f + j];
            require(e >= 0 && e < ne, "face-to-edge mapping references an edge out of range");
            if (m.ef[0][e] < 0) { m.ef[0][e] = f; m.el[0][e] = j; }
            else {
                require(m.ef[1][e] < 0, "RWG edge " + std::to_string(e) + " is shared by more than two faces");
                m.ef[1][e] = f; m.el[1][e] = j;
            }
        }
    for (int e = 0; e < ne; ++e)
        require(m.ef[1][e] >= 0, "RWG edge " + std::to_string(e) + " belongs to a single face (mesh not closed)");

    FaceSoA& h = m.h;
    h.cx.resize(nf); h.cy.resize(nf); h.cz.resize(nf); h.area.resize(nf);
    for (int a = 0; a < 3; ++a) {
        for (int x = 0; x < 3; ++x) h.o[a][x].resize(nf);
        h.len[a].resize(nf); h.edge[a].resize(nf); h.vid[a].resize(nf);
    }
    for (int f = 0; f < nf; ++f) {
        double v[3][3];
        for (int a = 0; a < 3; ++a)
            for (int x = 0; x < 3; ++x) v[a][x] = V[3 * (size_t)F[3 * f + a] + x];
        // (v0 + v1 + v2) / 3.f as sie_kernels_assembly.cuh:193-194 forms it
        const double c[3] = {(v[0][0] + v[1][0] + v[2][0]) / 3.0, (v[0][1] + v[1][1] + v[2][1]) / 3.0,
                             (v[0][2] + v[1][2] + v[2][2]) / 3.0};
        h.cx[f] = c[0]; h.cy[f] = c[1]; h.cz[f] = c[2];
        for (int a = 0; a < 3; ++a)
            for (int x = 0; x < 3; ++x) h.o[a][x][f] = v[a][x] - c[x];
        for (int j = 0; j < 3; ++j) {
            // getEdgeInfo (mesh_geometry.cuh:51-80)
            const int a = j, b = (j + 1) % 3;
            const double dx = v[b][0] - v[a][0], dy = v[b][1] - v[a][1], dz = v[b][2] - v[a][2];
            const double l = std::sqrt(dx * dx + dy * dy + dz * dz);
            h.len[j][f] = (F[3 * f + a] > F[3 * f + b] ? 1.0 : -1.0) * l;
            h.edge[j][f] = FE[3 * f + j];
            h.vid[j][f] = F[3 * f + j];
        }
        const double e1[3] = {v[1][0] - v[0][0], v[1][1] - v[0][1], v[1][2] - v[0][2]};
        const double e2[3] = {v[2][0] - v[0][0], v[2][1] - v[0][1], v[2][2] - v[0][2]};
        const double cr[3] = {e1[1] * e2[2] - e1[2] * e2[1], e1[2] * e2[0] - e1[0] * e2[2], e1[0] * e2[1] - e1[1] * e2[0]};
        h.area[f] = 0.5 * std::sqrt(cr[0] * cr[0] + cr[1] * cr[1] + cr[2] * cr[2]);
    }
    // device SoA
    std::vector<double> c(3 * (size_t)nf), o(9 * (size_t)nf), len(3 * (size_t)nf);
    std::vector<int> edge(3 * (size_t)nf), vid(3 * (size_t)nf);
    for (int f = 0; f < nf; ++f) {
        c[f] = h.cx[f]; c[nf + f] = h.cy[f]; c[2 * (size_t)nf + f] = h.cz[f];
        for (int a = 0; a < 3; ++a) {
            for (int x = 0; x < 3; ++x) o[(3 * a + x) * (size_t)nf + f] = h.o[a][x][f];
            len[a * (size_t)nf + f] = h.len[a][f];
            edge[a * (size_t)nf + f] = h.edge[a][f];
            vid[a * (size_t)nf + f] = h.vid[a][f];
        }
    }
    m.d_c.alloc(c.size()); m.d_c.upload(c.data(), c.size(), s);
    m.d_o.alloc(o.size()); m.d_o.upload(o.data(), o.size(), s);
    m.d_len.alloc(len.size()); m.d_len.upload(len.data(), len.size(), s);
    m.d_edge.alloc(edge.size()); m.d_edge.upload(edge.data(), edge.size(), s);
    m.d_vid.alloc(vid.size()); m.d_vid.upload(vid.data(), vid.size(), s);
    m.d_V.alloc(m.V.size()); m.d_V.upload(m.V.data(), m.V.size(), s);
    m.d_F.alloc(m.F.size()); m.d_F.upload(m.F.data(), m.F.size(), s);
    std::vector<int> ef(2 * (size_t)ne), el(2 * (size_t)ne);
    for (int e = 0; e < ne; ++e) {
        ef[e] = m.ef[0][e]; ef[ne + e] = m.ef[1][e];
        el[e] = m.el[0][e]; el[ne + e] = m.el[1][e];
    }
    m.d_ef.alloc(ef.size()); m.d_ef.upload(ef.data(), ef.size(), s);
    m.d_el.alloc(el.size()); m.d_el.upload(el.data(), el.size(), s);
    SIE_HIP(hipStreamSynchronize(s));
}

}  // namespace sie

// ------------------------------------------------------------- generators
#include <array>
#include <map>
#include <utility>

namespace sie {

// RWG edge numbering in first-encounter order over faces and local edges
// (spherical_surface.cpp:100-135): face f, local edge j = (v_j, v_{j+1}).
void number_edges(const std::vector<int>& F, std::vector<int>& E, std::vector<int>& FE) {
    const size_t nf = F.size() / 3;
    std::map<std::pair<int, int>, int> ids;
    FE.assign(3 * nf, -1);
    E.clear();
    for (size_t f = 0; f < nf; ++f)
        for (int j = 0; j < 3; ++j) {
            const int a = F[3 * f + j], b = F[3 * f + (j + 1) % 3];
            const auto key = std::minmax(a, b);
            auto it = ids.find(key);
            int id;
            if (it == ids.end()) {
                id = (int)ids.size();
                ids.emplace(key, id);
            } else {
                id = it->second;
            }
            FE[3 * f + j] = id;
        }
    E.assign(2 * ids.size(), 0);
    for (const auto& kv : ids) {
        E[2 * kv.second] = kv.first.first;
        E[2 * kv.second + 1] = kv.first.second;
    }
}

// Icosahedral sphere: SphericalSurface(radius, level) (spherical_surface.cpp:19-98):
// 12-vertex icosahedron, level-fold 4:1 midpoint subdivision with unit-sphere
// projection, scaling, outward counter-clockwise orientation.
void icosphere(float radius, int level, std::vector<double>& V, std::vector<int>& F) {
    const double phi = (1.0 + std::sqrt(5.0)) / 2.0;
    std::vector<std::array<double, 3>> P = {{-1, phi, 0}, {1, phi, 0}, {-1, -phi, 0}, {1, -phi, 0},
                                            {0, -1, phi}, {0, 1, phi}, {0, -1, -phi}, {0, 1, -phi},
                                            {phi, 0, -1}, {phi, 0, 1}, {-phi, 0, -1}, {-phi, 0, 1}};
    auto unit = [](std::array<double, 3> v) {
        const double s = 1.0 / std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
        return std::array<double, 3>{v[0] * s, v[1] * s, v[2] * s};
    };
    for (auto& p : P) p = unit(p);
    std::vector<std::array<int, 3>> T = {{0, 11, 5}, {0, 5, 1},  {0, 1, 7},   {0, 7, 10}, {0, 10, 11},
                                         {1, 5, 9},  {5, 11, 4}, {11, 10, 2}, {10, 7, 6}, {7, 1, 8},
                                         {3, 9, 4},  {3, 4, 2},  {3, 2, 6},   {3, 6, 8},  {3, 8, 9},
                                         {4, 9, 5},  {2, 4, 11}, {6, 2, 10},  {8, 6, 7},  {9, 8, 1}};
    for (int l = 0; l < level; ++l) {
        std::map<std::pair<int, int>, int> mid;
        auto midpoint = [&](int i, int j) {
            const auto key = std::minmax(i, j);
            auto it = mid.find(key);
            if (it != mid.end()) return it->second;
            const std::array<double, 3> m = {(P[i][0] + P[j][0]) * 0.5, (P[i][1] + P[j][1]) * 0.5,
                                             (P[i][2] + P[j][2]) * 0.5};
            P.push_back(unit(m));
            return mid[key] = (int)P.size() - 1;
        };
        std::vector<std::array<int, 3>> nt;
        nt.reserve(T.size() * 4);
        for (const auto& t : T) {
            const int a = t[0], b = t[1], c = t[2];
            const int ab = midpoint(a, b), bc = midpoint(b, c), ca = midpoint(c, a);
            nt.push_back({a, ab, ca});
            nt.push_back({ab, b, bc});
            nt.push_back({ca, bc, c});
            nt.push_back({ab, bc, ca});
        }
        T.swap(nt);
    }
    V.resize(3 * P.size());
    for (size_t i = 0; i < P.size(); ++i)
        for (int x = 0; x < 3; ++x) V[3 * i + x] = P[i][x] * (double)radius;
    F.resize(3 * T.size());
    for (size_t f = 0; f < T.size(); ++f) {
        auto& t = T[f];
        const double* v0 = &V[3 * t[0]];
        const double* v1 = &V[3 * t[1]];
        const double* v2 = &V[3 * t[2]];
        const double e1[3] = {v1[0] - v0[0], v1[1] - v0[1], v1[2] - v0[2]};
        const double e2[3] = {v2[0] - v0[0], v2[1] - v0[1], v2[2] - v0[2]};
        const double n[3] = {e1[1] * e2[2] - e1[2] * e2[1], e1[2] * e2[0] - e1[0] * e2[2], e1[0] * e2[1] - e1[1] * e2[0]};
        const double c[3] = {(v0[0] + v1[0] + v2[0]) / 3.0, (v0[1] + v1[1] + v2[1]) / 3.0, (v0[2] + v1[2] + v2[2]) / 3.0};
        if (n[0] * c[0] + n[1] * c[1] + n[2] * c[2] < 0) std::swap(t[1], t[2]);
        F[3 * f] = t[0]; F[3 * f + 1] = t[1]; F[3 * f + 2] = t[2];
    }
}

// Circular-arc spindle: the surface of revolution about the z axis of a
// circular arc of radius `arc_radius` whose chord (length `chord`) lies on the
// axis; profile rho(z) = sqrt(R^2 - z^2) - sqrt(R^2 - (L/2)^2), |z| <= L/2.
// Rings of vertices at n_rings-1 interior stations (uniform in arc length),
// each with a vertex count proportional to its circumference (target edge
// length h = arc length / n_rings), stitched ring to ring by angular merge;
// the two tips are single vertices. Faces are outward counter-clockwise.
void spindle(double chord, double arc_radius, int n_rings, std::vector<double>& V, std::vector<int>& F) {
    require(chord > 0 && arc_radius > chord / 2 && n_rings >= 3, "invalid spindle parameters");
    const double L2 = chord / 2, R = arc_radius, d = std::sqrt(R * R - L2 * L2);
    const double a0 = std::asin(L2 / R);  // profile angle: z = R sin(t), rho = R cos(t) - d, t in [-a0, a0]
    const double h = 2 * a0 * R / n_rings;
    const double two_pi = 2.0 * std::acos(-1.0);
    V.clear();
    F.clear();
    // tip (south) vertex
    V.insert(V.end(), {0.0, 0.0, -L2});
    std::vector<int> ring_start, ring_n;
    for (int k = 1; k < n_rings; ++k) {
        const double t = -a0 + 2 * a0 * k / n_rings;
        const double z = R * std::sin(t), rho = R * std::cos(t) - d;
        int n = (int)std::lround(two_pi * rho / h);
        if (n < 6) n = 6;
        ring_start.push_back((int)V.size() / 3);
        ring_n.push_back(n);
        const double off = (k & 1) ? 0.5 : 0.0;  // stagger alternate rings
        for (int i = 0; i < n; ++i) {
            const double ph = two_pi * (i + off) / n;
            V.insert(V.end(), {rho * std::cos(ph), rho * std::sin(ph), z});
        }
    }
    const int north = (int)V.size() / 3;
    V.insert(V.end(), {0.0, 0.0, L2});
    auto ang = [&](int r, int i) {
        const double off = ((r + 1) & 1) ? 0.5 : 0.0;
        return two_pi * (i + off) / ring_n[r];
    };
    // south fan: outward normal points to -z there; (tip, next, this) is CCW seen from outside
    {
        const int s0 = ring_start[0], n = ring_n[0];
        for (int i = 0; i < n; ++i) F.insert(F.end(), {0, s0 + (i + 1) % n, s0 + i});
    }
    for (size_t r = 0; r + 1 < ring_n.size(); ++r) {
        const int a0s = ring_start[r], na = ring_n[r], b0s = ring_start[r + 1], nb = ring_n[r + 1];
        int i = 0, j = 0;
        // start both rings at the vertex with the smallest angle
        while (i < na || j < nb) {
            const double ta = i < na ? ang((int)r, i + 1) : 1e30;
            const double tb = j < nb ? ang((int)r + 1, j + 1) : 1e30;
            const int ai = a0s + i % na, aj = a0s + (i + 1) % na;
            const int bi = b0s + j % nb, bj = b0s + (j + 1) % nb;
            if (ta <= tb) {  // advance on the lower ring
                F.insert(F.end(), {ai, aj, bi});
                ++i;
            } else {
                F.insert(F.end(), {ai, bj, bi});
                ++j;
            }
        }
    }
    {
        const int s0 = ring_start.back(), n = ring_n.back();
        for (int i = 0; i < n; ++i) F.insert(F.end(), {north, s0 + i, s0 + (i + 1) % n});
    }
}

}  // namespace sie

extern "C" {

static int emit_mesh(const std::vector<double>& V, const std::vector<int>& F, double* Vo, int* Fo, int* Eo, int* FEo,
                     int* counts) {
    std::vector<int> E, FE;
    sie::number_edges(F, E, FE);
    counts[0] = (int)(V.size() / 3);
    counts[1] = (int)(F.size() / 3);
    counts[2] = (int)(E.size() / 2);
    if (Vo) std::copy(V.begin(), V.end(), Vo);
    if (Fo) std::copy(F.begin(), F.end(), Fo);
    if (Eo) std::copy(E.begin(), E.end(), Eo);
    if (FEo) std::copy(FE.begin(), FE.end(), FEo);
    return SIE_OK;
}

int sie_icosphere(float radius, int level, double* V, int* F, int* E, int* FE, int* counts) {
    try {
        sie::require(counts != nullptr, "counts pointer is null");
        sie::require(level >= 0 && level <= 9 && radius > 0, "invalid icosphere radius or level");
        std::vector<double> v;
        std::vector<int> f;
        sie::icosphere(radius, level, v, f);
        sie::set_last_error("");
        return emit_mesh(v, f, V, F, E, FE, counts);
    } catch (const sie::Error& e) {
        sie::set_last_error(e.what());
        return e.code;
    }
}

int sie_spindle(double chord, double arc_radius, int n_rings, double* V, int* F, int* E, int* FE, int* counts) {
    try {
        sie::require(counts != nullptr, "counts pointer is null");
        std::vector<double> v;
        std::vector<int> f;
        sie::spindle(chord, arc_radius, n_rings, v, f);
        sie::set_last_error("");
        return emit_mesh(v, f, V, F, E, FE, counts);
    } catch (const sie::Error& e) {
        sie::set_last_error(e.what());
        return e.code;
    }
}

}  // extern "C"
