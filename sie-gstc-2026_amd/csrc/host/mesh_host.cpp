// MeshData front-end: generators over the C ABI, edge building and the OBJ reader.
#include <algorithm>
#include <array>
#include <cstdlib>
#include <fstream>
#include <map>
#include <sstream>
#include <utility>

#include "host_common.hpp"
#include "sie_mom/mesh/circular_ogive.hpp"
#include "sie_mom/mesh/obj_mesh_loader.hpp"

namespace sie_mom {
namespace mesh {

static_assert(sizeof(vec3_t) == 3 * sizeof(double), "vec3_t must be three packed doubles");
static_assert(sizeof(int3) == 3 * sizeof(int) && sizeof(int2) == 2 * sizeof(int), "packed int vectors");

namespace {
template <typename Gen>
MeshData generate(Gen gen) {
    int counts[3] = {0, 0, 0};
    detail::check(gen(nullptr, nullptr, nullptr, nullptr, counts));
    MeshData m;
    m.vertices.resize(counts[0]);
    m.faces.resize(counts[1]);
    m.edges.resize(counts[2]);
    m.faces_edges.resize(counts[1]);
    detail::check(gen(reinterpret_cast<double*>(m.vertices.data()), reinterpret_cast<int*>(m.faces.data()),
                      reinterpret_cast<int*>(m.edges.data()), reinterpret_cast<int*>(m.faces_edges.data()), counts));
    return m;
}
}  // namespace

SphericalSurface::SphericalSurface(float radius, int refinement_level) {
    data_ = generate([&](double* v, int* f, int* e, int* fe, int* c) {
        return sie_icosphere(radius, refinement_level, v, f, e, fe, c);
    });
}

MeshData make_spindle(double chord, double arc_radius, int n_rings) {
    return generate([&](double* v, int* f, int* e, int* fe, int* c) {
        return sie_spindle(chord, arc_radius, n_rings, v, f, e, fe, c);
    });
}

void build_edges(MeshData& m) {
    std::map<std::pair<int, int>, int> index;
    std::vector<int> uses;
    m.edges.clear();
    m.faces_edges.assign(m.faces.size(), int3{});
    for (std::size_t f = 0; f < m.faces.size(); ++f) {
        const int v[3] = {m.faces[f].x, m.faces[f].y, m.faces[f].z};
        int fe[3];
        for (int j = 0; j < 3; ++j) {
            const auto key = std::minmax(v[j], v[(j + 1) % 3]);
            auto it = index.find(key);
            if (it == index.end()) {
                it = index.emplace(key, (int)m.edges.size()).first;
                m.edges.push_back(int2{key.first, key.second});
                uses.push_back(0);
            }
            fe[j] = it->second;
            ++uses[it->second];
        }
        m.faces_edges[f] = int3{fe[0], fe[1], fe[2]};
    }
    for (std::size_t e = 0; e < uses.size(); ++e)
        if (uses[e] != 2)
            throw std::invalid_argument("mesh is not a closed 2-manifold: edge " + std::to_string(e) + " has " +
                                        std::to_string(uses[e]) + " faces");
}

namespace {
int parse_index(const std::string& tok, std::size_t seen, const std::string& where) {
    const std::string head = tok.substr(0, tok.find('/'));
    if (head.empty()) throw std::runtime_error(where + ": face has an empty vertex index");
    char* end = nullptr;
    const long v = std::strtol(head.c_str(), &end, 10);
    if (*end != '\0' || v == 0) throw std::runtime_error(where + ": invalid face vertex index '" + head + "'");
    if (v > 0) return (int)(v - 1);
    const long long k = (long long)seen + v;
    if (k < 0 || k >= (long long)seen) throw std::runtime_error(where + ": relative face vertex index is out of range");
    return (int)k;
}

// The reference loader's checks (obj_mesh_loader.cpp:139-358): every vertex
// used, no degenerate or duplicate faces, every edge shared by exactly two
// faces with opposite directions, one connected component, positive volume.
void validate_closed_surface(const MeshData& m, const std::string& path) {
    const auto& V = m.vertices;
    double lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (const auto& v : V) {
        if (!std::isfinite(v.x) || !std::isfinite(v.y) || !std::isfinite(v.z))
            throw std::runtime_error("Invalid OBJ mesh " + path + ": non-finite vertex coordinate");
        const double c[3] = {v.x, v.y, v.z};
        for (int i = 0; i < 3; ++i) { lo[i] = std::min(lo[i], c[i]); hi[i] = std::max(hi[i], c[i]); }
    }
    double scale = std::sqrt((hi[0] - lo[0]) * (hi[0] - lo[0]) + (hi[1] - lo[1]) * (hi[1] - lo[1]) +
                             (hi[2] - lo[2]) * (hi[2] - lo[2]));
    if (scale == 0.0) scale = 1.0;
    const double area_tol = 1e-10 * scale * scale, vol_tol = 1e-10 * scale * scale * scale;
    std::vector<char> used(V.size(), 0);
    std::map<std::array<int, 3>, int> seen_faces;
    std::map<std::pair<int, int>, std::vector<std::pair<int, int>>> occ;  // edge -> (face, direction)
    int degenerate = 0, duplicate = 0;
    double vol6 = 0.0;
    for (std::size_t f = 0; f < m.faces.size(); ++f) {
        const int idx[3] = {m.faces[f].x, m.faces[f].y, m.faces[f].z};
        for (int i : idx) {
            if (i < 0 || (std::size_t)i >= V.size())
                throw std::runtime_error("Invalid OBJ mesh " + path + ": face contains an out-of-range vertex index");
            used[i] = 1;
        }
        std::array<int, 3> key = {idx[0], idx[1], idx[2]};
        std::sort(key.begin(), key.end());
        if (seen_faces[key]++ > 0) ++duplicate;
        const vec3_t a = V[idx[0]], b = V[idx[1]], c = V[idx[2]];
        const double u[3] = {b.x - a.x, b.y - a.y, b.z - a.z}, w[3] = {c.x - a.x, c.y - a.y, c.z - a.z};
        const double n[3] = {u[1] * w[2] - u[2] * w[1], u[2] * w[0] - u[0] * w[2], u[0] * w[1] - u[1] * w[0]};
        if (0.5 * std::sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]) <= area_tol) ++degenerate;
        vol6 += a.x * (b.y * c.z - b.z * c.y) - a.y * (b.x * c.z - b.z * c.x) + a.z * (b.x * c.y - b.y * c.x);
        for (int j = 0; j < 3; ++j) {
            const int s0 = idx[j], s1 = idx[(j + 1) % 3];
            const auto e = std::minmax(s0, s1);
            occ[e].push_back({(int)f, s0 == e.first ? 1 : -1});
        }
    }
    int boundary = 0, nonmanifold = 0, misoriented = 0;
    std::vector<std::vector<int>> adj(m.faces.size());
    for (const auto& kv : occ) {
        const auto& o = kv.second;
        if (o.size() == 1) ++boundary;
        else if (o.size() > 2) ++nonmanifold;
        if (o.size() == 2) {
            if (o[0].second == o[1].second) ++misoriented;
            adj[o[0].first].push_back(o[1].first);
            adj[o[1].first].push_back(o[0].first);
        }
    }
    int components = 0;
    std::vector<char> vis(m.faces.size(), 0);
    for (std::size_t s = 0; s < m.faces.size(); ++s) {
        if (vis[s]) continue;
        ++components;
        std::vector<int> stack{(int)s};
        vis[s] = 1;
        while (!stack.empty()) {
            const int f = stack.back();
            stack.pop_back();
            for (int g : adj[f])
                if (!vis[g]) { vis[g] = 1; stack.push_back(g); }
        }
    }
    std::vector<std::string> problems;
    const long orphans = (long)std::count(used.begin(), used.end(), 0);
    if (orphans) problems.push_back(std::to_string(orphans) + " unreferenced vertex/vertices");
    if (degenerate) problems.push_back(std::to_string(degenerate) + " degenerate face(s)");
    if (duplicate) problems.push_back(std::to_string(duplicate) + " duplicate face(s)");
    if (boundary) problems.push_back(std::to_string(boundary) + " boundary edge(s)");
    if (nonmanifold) problems.push_back(std::to_string(nonmanifold) + " non-manifold edge(s)");
    if (components != 1) problems.push_back(std::to_string(components) + " connected component(s), expected 1");
    if (misoriented) problems.push_back(std::to_string(misoriented) + " inconsistently oriented interior edge(s)");
    if (vol6 / 6.0 <= vol_tol) problems.push_back("signed volume is not positive; face winding is not outward");
    if (!problems.empty()) {
        std::string msg = "Invalid OBJ mesh " + path + ": ";
        for (std::size_t i = 0; i < problems.size(); ++i) msg += (i ? "; " : "") + problems[i];
        throw std::runtime_error(msg);
    }
}
}  // namespace

MeshData load_obj(const std::string& path) {
    std::ifstream in(path);
    if (!in) throw std::runtime_error("Failed to open OBJ file: " + path);
    MeshData m;
    std::string line;
    std::size_t lineno = 0;
    while (std::getline(in, line)) {
        ++lineno;
        const std::string where = path + ":" + std::to_string(lineno);
        std::istringstream ss(line);
        std::string tag;
        if (!(ss >> tag) || tag[0] == '#') continue;
        if (tag == "v") {
            vec3_t v;
            if (!(ss >> v.x >> v.y >> v.z)) throw std::runtime_error(where + ": vertex requires three coordinates");
            m.vertices.push_back(v);
        } else if (tag == "f") {
            std::vector<std::string> toks;
            std::string t;
            while (ss >> t) {
                if (t[0] == '#') break;
                toks.push_back(t);
            }
            if (toks.size() != 3) throw std::runtime_error(where + ": only triangular faces are supported");
            const std::size_t seen = m.vertices.size();
            m.faces.push_back(int3{parse_index(toks[0], seen, where), parse_index(toks[1], seen, where),
                                   parse_index(toks[2], seen, where)});
        }
    }
    if (m.vertices.empty()) throw std::runtime_error("Invalid OBJ mesh " + path + ": no vertices");
    if (m.faces.empty()) throw std::runtime_error("Invalid OBJ mesh " + path + ": no faces");
    validate_closed_surface(m, path);
    build_edges(m);
    return m;
}

}  // namespace mesh

namespace host {
MeshHandle::MeshHandle(const mesh::MeshData& m) {
    if (m.vertices.empty() || m.faces.empty() || m.edges.empty())
        throw std::invalid_argument("SIE mesh must contain vertices, faces, and edges");
    if (m.faces_edges.size() != m.faces.size())
        throw std::invalid_argument("mesh face-to-edge mapping size does not match face count");
    detail::check(sie_mesh_create(reinterpret_cast<const double*>(m.vertices.data()), (int)m.vertices.size(),
                                  reinterpret_cast<const int*>(m.faces.data()), (int)m.faces.size(),
                                  reinterpret_cast<const int*>(m.faces_edges.data()), (int)m.edges.size(), &h_));
}
}  // namespace host
}  // namespace sie_mom

// C ABI of the OBJ loader (two calls, as the mesh generators: NULL arrays give
// the counts). Defined here, with the loader, in the C++ layer of the library.
namespace sie {
void set_last_error(const std::string& m);  // the C ABI's thread-local message (capi.cpp)
}

extern "C" int sie_obj_load(const char* path, double* vertices, int* faces, int* edges, int* face_edges,
                            int* counts) {
    try {
        if (!path || !counts) throw std::invalid_argument("null OBJ path or counts");
        const sie_mom::mesh::MeshData m = sie_mom::mesh::load_obj(path);
        counts[0] = (int)m.vertices.size();
        counts[1] = (int)m.faces.size();
        counts[2] = (int)m.edges.size();
        if (vertices && faces && edges && face_edges) {
            for (std::size_t i = 0; i < m.vertices.size(); ++i) {
                vertices[3 * i] = m.vertices[i].x;
                vertices[3 * i + 1] = m.vertices[i].y;
                vertices[3 * i + 2] = m.vertices[i].z;
            }
            for (std::size_t i = 0; i < m.faces.size(); ++i) {
                faces[3 * i] = m.faces[i].x;
                faces[3 * i + 1] = m.faces[i].y;
                faces[3 * i + 2] = m.faces[i].z;
                face_edges[3 * i] = m.faces_edges[i].x;
                face_edges[3 * i + 1] = m.faces_edges[i].y;
                face_edges[3 * i + 2] = m.faces_edges[i].z;
            }
            for (std::size_t i = 0; i < m.edges.size(); ++i) {
                edges[2 * i] = m.edges[i].x;
                edges[2 * i + 1] = m.edges[i].y;
            }
        }
        sie::set_last_error("");
        return SIE_OK;
    } catch (const std::invalid_argument& e) {
        sie::set_last_error(e.what());
        return SIE_ERR_INVALID;
    } catch (const std::exception& e) {
        sie::set_last_error(e.what());
        return SIE_ERR_INVALID;
    }
}
