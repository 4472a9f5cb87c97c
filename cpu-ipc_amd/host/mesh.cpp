// This Source Code Form is subject to the terms of the Mozilla Public License, v. 2.0.
// If a copy of the MPL was not distributed with this file, You can obtain one at
// https://mozilla.org/MPL/2.0/.  Parts of this file restate logic of CPU-IPC (MPL-2.0);
// each such part cites the source it follows.
//
// Scene set-up on the host (see mesh.hpp).  Compiled with -ffp-contract=off: the rest-shape
// quantities round as the reference's expressions do.
#include "mesh.hpp"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <fstream>
#include <iomanip>
#include <map>
#include <set>
#include <sstream>
#include <stdexcept>
#include <unordered_map>
#include <unordered_set>

namespace cipc {
namespace host {

namespace {

Vec3 sub(const Vec3& a, const Vec3& b) { return {{a[0] - b[0], a[1] - b[1], a[2] - b[2]}}; }
Vec3 cross(const Vec3& a, const Vec3& b) {
    return {{a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0]}};
}
double dot(const Vec3& a, const Vec3& b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }
double norm(const Vec3& a) { return std::sqrt(dot(a, a)); }
Vec3 scaled(double s, const Vec3& a) { return {{s * a[0], s * a[1], s * a[2]}}; }

// Eigen's closed-form 3x3 inverse (cofactors of the first column give the determinant)
// on a row-major M; the result is stored column-major.
std::array<double, 9> inverse3_colmajor(const double M[3][3]) {
    auto cof = [&](int i, int j) {
        const int i1 = (i + 1) % 3, i2 = (i + 2) % 3, j1 = (j + 1) % 3, j2 = (j + 2) % 3;
        return M[i1][j1] * M[i2][j2] - M[i1][j2] * M[i2][j1];
    };
    const double c00 = cof(0, 0), c10 = cof(1, 0), c20 = cof(2, 0);
    const double det = c00 * M[0][0] + c10 * M[1][0] + c20 * M[2][0];
    const double inv = 1.0 / det;
    double R[3][3];
    R[0][0] = c00 * inv; R[0][1] = c10 * inv; R[0][2] = c20 * inv;
    R[1][0] = cof(0, 1) * inv; R[1][1] = cof(1, 1) * inv; R[1][2] = cof(2, 1) * inv;
    R[2][0] = cof(0, 2) * inv; R[2][1] = cof(1, 2) * inv; R[2][2] = cof(2, 2) * inv;
    std::array<double, 9> out;
    for (int c = 0; c < 3; ++c)
        for (int r = 0; r < 3; ++r) out[r + 3 * c] = R[r][c];
    return out;
}

void fail(std::string& err, const char* what, const std::string& path) { err = std::string(what) + " in " + path; }

}  // namespace

bool load_tetrahedra(Mesh& m, const std::string& path, double scale, const Vec3& offset, std::string& err) {
    std::ifstream in(path);
    if (!in) { err = "unable to read file " + path; return false; }
    std::vector<Vec3> verts;
    std::vector<std::array<int, 4>> tets;
    std::unordered_map<int, int> id_of;
    bool have_nodes = false, have_elements = false;
    size_t node_fields = 0;
    std::string line;
    auto read_count = [&](int& n) {
        if (!std::getline(in, line)) return false;
        std::istringstream s(line);
        return static_cast<bool>(s >> n) && n >= 0;
    };
    while (std::getline(in, line)) {
        if (!line.empty() && line.back() == '\r') line.pop_back();
        if (line == "$Nodes") {
            int n = 0;
            if (have_nodes || !read_count(n)) { fail(err, "invalid $Nodes section", path); return false; }
            verts.reserve(n);
            for (int k = 0; k < n; ++k) {
                if (!std::getline(in, line)) { fail(err, "truncated $Nodes section", path); return false; }
                std::istringstream fs(line);
                std::vector<std::string> fields;
                for (std::string f; fs >> f;) fields.push_back(f);
                if (node_fields == 0) node_fields = fields.size();
                else if (fields.size() != node_fields) { fail(err, "inconsistent Gmsh node rows", path); return false; }
                int id = 0;
                double x = 0, y = 0, z = 0;
                bool ok = false;
                std::istringstream vs(line);
                if (fields.size() == 3) {               // compact rows: coordinates only, ids implied
                    id = k + 1;
                    ok = static_cast<bool>(vs >> x >> y >> z);
                } else if (fields.size() == 4) {
                    ok = static_cast<bool>(vs >> id >> x >> y >> z);
                } else {
                    fail(err, "invalid Gmsh node field count", path);
                    return false;
                }
                if (!ok || id <= 0 || !std::isfinite(x) || !std::isfinite(y) || !std::isfinite(z) || !id_of.emplace(id, k).second) {
                    fail(err, "invalid Gmsh node", path);
                    return false;
                }
                verts.push_back({{scale * x + offset[0], scale * y + offset[1], scale * z + offset[2]}});
            }
            have_nodes = true;
        } else if (line == "$Elements") {
            int n = 0;
            if (!have_nodes || have_elements || !read_count(n)) { fail(err, "invalid $Elements section", path); return false; }
            tets.reserve(n);
            for (int k = 0; k < n; ++k) {
                if (!std::getline(in, line)) { fail(err, "truncated $Elements section", path); return false; }
                std::istringstream es(line);
                std::vector<int> f;
                for (int v; es >> v;) f.push_back(v);
                if (!es.eof() || f.empty() || f[0] <= 0) { fail(err, "invalid Gmsh element header", path); return false; }
                size_t first = 0;
                if (node_fields == 3 && f.size() == 5) {
                    first = 1;                           // compact rows: id and four nodes
                } else {
                    if (f.size() < 3 || f[2] < 0) { fail(err, "invalid Gmsh element tags", path); return false; }
                    if (f[1] != 4) continue;              // not a tetrahedron
                    if (f.size() < 7) { fail(err, "invalid tetrahedron", path); return false; }
                    first = f.size() - 4;                 // the last four fields are the nodes
                }
                std::array<int, 4> t;
                for (int c = 0; c < 4; ++c) {
                    const auto it = id_of.find(f[first + c]);
                    if (it == id_of.end()) { fail(err, "tetrahedron references an unknown node", path); return false; }
                    t[c] = it->second;
                }
                tets.push_back(t);
            }
            have_elements = true;
            break;
        }
    }
    if (!have_nodes || !have_elements || verts.empty() || tets.empty()) { fail(err, "missing tetrahedral mesh data", path); return false; }
    Vec3 lo = verts.front(), hi = verts.front();
    const int base = (int)m.x.size();
    for (const Vec3& v : verts) {
        m.x.push_back(v);
        m.velocities.push_back({{0, 0, 0}});
        m.masses.push_back(0.0);
        m.boundary.push_back(0);
        for (int d = 0; d < 3; ++d) { lo[d] = std::min(lo[d], v[d]); hi[d] = std::max(hi[d], v[d]); }
    }
    for (auto t : tets) {
        for (int& v : t) v += base;
        m.tets.push_back(t);
    }
    for (int d = 0; d < 3; ++d) { m.max_corner[d] = std::max(m.max_corner[d], hi[d]); m.min_corner[d] = std::min(m.min_corner[d], lo[d]); }
    if (norm(sub(hi, lo)) > norm(sub(m.obj_max, m.obj_min))) { m.obj_max = hi; m.obj_min = lo; }
    m.V_prev = m.x;
    return true;
}

bool load_triangles(Mesh& m, const std::string& path, double scale, const Vec3& offset, int type, std::string& err) {
    const int base = (int)m.x.size();
    std::ifstream in(path);
    if (!in) { err = "unable to read file " + path; return false; }
    Vec3 lo{{1e32, 1e32, 1e32}}, hi{{-1e32, -1e32, -1e32}};
    std::string line;
    while (std::getline(in, line)) {
        if (line.size() <= 1) continue;
        const std::string key = line.substr(0, 2);
        if (key == "v ") {
            std::stringstream ss(line.substr(2));
            double x, y, z;
            ss >> x >> y >> z;
            const Vec3 v{{scale * x + offset[0], scale * y + offset[1], scale * z + offset[2]}};
            m.x.push_back(v);
            m.velocities.push_back({{0, 0, 0}});
            m.masses.push_back(0.0);
            m.boundary.push_back(type == 2 ? 2 : 0);
            for (int d = 0; d < 3; ++d) { if (lo[d] > v[d]) lo[d] = v[d]; if (hi[d] < v[d]) hi[d] = v[d]; }
        } else if (key == "f ") {
            if (line.size() >= 1024) continue;
            std::stringstream ts(line);
            std::vector<std::string> tok;
            for (std::string b; ts >> b;) tok.push_back(b);
            for (size_t i = 3; i < tok.size(); ++i) {             // fan triangulation
                const std::string corners[3] = {tok[1], tok[i - 1], tok[i]};
                std::array<int, 3> f{{0, 0, 0}};
                for (int c = 0; c < 3; ++c) f[c] = std::atoi(corners[c].c_str()) - 1 + base;   // "a" or "a/b/c"
                if (type == 3) m.surface.push_back({{f[0], f[1], f[2], 0}});
                else m.triangles.push_back(f);
            }
        }
    }
    for (int d = 0; d < 3; ++d) {
        if (m.max_corner[d] < hi[d]) m.max_corner[d] = hi[d];
        if (m.min_corner[d] > lo[d]) m.min_corner[d] = lo[d];
    }
    // interior edges (two triangles) and their opposite vertices, in edge order
    std::map<std::pair<int, int>, std::vector<int>> adj;
    for (const auto& t : m.triangles) {
        const int a = t[0], b = t[1], c = t[2];
        adj[{std::min(a, b), std::max(a, b)}].push_back(c);
        adj[{std::min(b, c), std::max(b, c)}].push_back(a);
        adj[{std::min(a, c), std::max(a, c)}].push_back(b);
    }
    m.tri_edges.clear();
    m.tri_edge_adj.clear();
    for (const auto& e : adj) {
        if (e.second.size() != 2) continue;
        m.tri_edges.push_back({{e.first.first, e.first.second}});
        m.tri_edge_adj.push_back({{e.second[0], e.second[1]}});
    }
    m.V_prev = m.x;
    return true;
}

void get_surface(Mesh& m) {
    // The face list comes from iterating an unordered_map keyed by the first-seen vertex
    // order of each face, with the reference's hash and bucket hint: the same standard
    // library walks it in the same order.
    using Tri = std::array<int, 3>;
    const uint64_t V = m.x.size();
    const int T = (int)m.tets.size();
    auto h = [V](const Tri& t) -> size_t { return V * (V * (uint64_t)t[0] + (uint64_t)t[1]) + (uint64_t)t[2]; };
    std::unordered_map<Tri, uint64_t, decltype(h)> face(4 * (size_t)T, h);
    for (int i = 0; i < T; ++i) {
        const auto& t = m.tets[i];
        for (int j = 0; j < 4; ++j) {
            const int a = t[j % 4], b = t[(j + 1) % 4], c = t[(j + 2) % 4];
            const Tri perms[6] = {{{a, b, c}}, {{a, c, b}}, {{b, a, c}}, {{b, c, a}}, {{c, a, b}}, {{c, b, a}}};
            bool seen = false;
            for (const Tri& p : perms)
                if (face.find(p) != face.end()) { face[p] = (uint64_t)T + 1; seen = true; break; }
            if (!seen) face[Tri{{a, b, c}}] = (uint64_t)i;
        }
    }
    m.surface.clear();
    for (const auto& kv : face) {
        const uint64_t tet = kv.second;
        if (tet >= (uint64_t)T) continue;
        const Tri& f = kv.first;
        const Vec3 e1 = sub(m.x[f[1]], m.x[f[0]]), e2 = sub(m.x[f[2]], m.x[f[0]]);
        int opp = 0;
        for (int k = 0; k < 4; ++k) {
            const int v = m.tets[tet][k];
            if (v != f[0] && v != f[1] && v != f[2]) { opp = v; break; }
        }
        const Vec3 e3 = sub(m.x[opp], m.x[f[0]]);
        if (dot(cross(e1, e2), e3) < 0) m.surface.push_back({{f[0], f[1], f[2], (int)tet}});
        else m.surface.push_back({{f[0], f[2], f[1], (int)tet}});
    }
    for (const auto& t : m.triangles) m.surface.push_back({{t[0], t[1], t[2], 0}});
    m.surf_verts.clear();
    std::vector<bool> flag(m.x.size(), false);
    for (const auto& f : m.surface)
        for (int k = 0; k < 3; ++k)
            if (!flag[f[k]]) { m.surf_verts.push_back((uint64_t)f[k]); flag[f[k]] = true; }
    std::set<std::pair<uint64_t, uint64_t>> edges;
    for (const auto& f : m.surface)
        for (int k = 0; k < 3; ++k) {
            const uint64_t a = (uint64_t)f[k], b = (uint64_t)f[(k + 1) % 3];
            if (edges.find({b, a}) == edges.end()) edges.insert({a, b});
        }
    m.surf_edges.assign(edges.begin(), edges.end());
}

void init_mesh(Mesh& m) {
    const int V = (int)m.x.size();
    m.masses.assign(V, 0.0);
    m.dm_inv.clear();
    m.volumes.clear();
    m.dm_tri_inv.clear();
    m.areas.clear();
    double mass_sum = 0.0;
    for (const auto& t : m.tets) {
        const Vec3 e1 = sub(m.x[t[1]], m.x[t[0]]), e2 = sub(m.x[t[2]], m.x[t[0]]), e3 = sub(m.x[t[3]], m.x[t[0]]);
        const double D[3][3] = {{e1[0], e2[0], e3[0]}, {e1[1], e2[1], e3[1]}, {e1[2], e2[2], e3[2]}};
        // calculateVolum (Elasticity.cpp:535-561): base area times the height along the unit normal
        Vec3 n = cross(e1, e2);
        const double area = norm(n);
        n = scaled(1.0 / area, n);
        double vol = area * std::fabs(dot(n, e3)) / 6;
        vol = vol > 0 ? vol : -vol;
        for (int k = 0; k < 4; ++k) m.masses[t[k]] += vol * m.density / 4;
        mass_sum += vol * m.density;
        m.dm_inv.push_back(inverse3_colmajor(D));
        m.volumes.push_back(vol);
    }
    for (const auto& t : m.triangles) {
        // calculateDms2D_double (Elasticity.cpp:610-639): the triangle rotated so that its
        // normal is +y, uv = (x, z)
        const Vec3 e1 = sub(m.x[t[1]], m.x[t[0]]), e2 = sub(m.x[t[2]], m.x[t[0]]);
        Vec3 n = cross(e1, e2);
        n = scaled(1.0 / norm(n), n);
        const Vec3 up{{0, 1, 0}};
        const Vec3 w = cross(n, up);
        const double c = dot(n, up);
        double K[3][3] = {{0, -w[2], w[1]}, {w[2], 0, -w[0]}, {-w[1], w[0], 0}};
        double R[3][3];
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) {
                const double kk = K[i][0] * K[0][j] + K[i][1] * K[1][j] + K[i][2] * K[2][j];
                R[i][j] = (i == j ? 1.0 : 0.0) + (K[i][j] + kk / (1 + c));
            }
        double uv[3][2];
        for (int q = 0; q < 3; ++q) {
            const Vec3& p = m.x[t[q]];
            uv[q][0] = R[0][0] * p[0] + R[0][1] * p[1] + R[0][2] * p[2];
            uv[q][1] = R[2][0] * p[0] + R[2][1] * p[1] + R[2][2] * p[2];
        }
        const double a = uv[1][0] - uv[0][0], b = uv[2][0] - uv[0][0];    // column 0 = uv1 - uv0, column 1 = uv2 - uv0
        const double cc = uv[1][1] - uv[0][1], d = uv[2][1] - uv[0][1];
        const double det = a * d - b * cc;
        m.dm_tri_inv.push_back({{d / det, -cc / det, -b / det, a / det}});
        const double area = norm(cross(e1, e2)) * 0.5 * m.cloth_thickness;
        m.areas.push_back(area);
        for (int k = 0; k < 3; ++k) m.masses[t[k]] += area * m.cloth_density / 3;
        mass_sum += area * m.cloth_density;
    }
    m.mean_mass = mass_sum / V;
}

void update_material(Mesh& m) {
    const double lame_u = m.youngs / (2 * (1 + m.poisson));
    const double lame_l = m.youngs * m.poisson / ((1 + m.poisson) * (1 - 2 * m.poisson));
    m.length_rate = 4 * lame_u / 3;
    m.volume_rate = lame_l + 5 * lame_u / 6;
    m.stretch = m.cloth_youngs / (1 - m.poisson * m.poisson);
    m.shear = m.shear_youngs / (2 * (1 + m.poisson)) / m.cloth_thickness;
    m.plate = m.bend_youngs * std::pow(m.cloth_thickness, 3) / (12 * (1 - m.poisson * m.poisson));
}

void load_settings(Mesh& m, const std::string& path) {
    std::ifstream in(path);
    if (!in) throw std::runtime_error("unable to read settings " + path);
    static const char* keys[] = {"volume_mesh_density", "volume_mesh_YoungsModulus", "poisson_rate", "friction_rate",
                                 "triangle(cloth)_mesh_thickness", "triangle(cloth)_mesh_YoungsModulus",
                                 "triangle(cloth)_mesh_shearYoungsModulus", "triangle(cloth)_mesh_bendingYoungsModulus",
                                 "triangle(cloth)_mesh_density", "strainRate", "enable_collision_handling", "drag_coeff",
                                 "ipc_time_step", "Newton_solver_threshold", "IPC_ralative_dHat"};
    const std::unordered_set<std::string> known(std::begin(keys), std::end(keys));
    std::unordered_map<std::string, double> val;
    std::string k;
    double v = 0.0;
    while (in >> k) {
        if (!(in >> v)) throw std::runtime_error("invalid value for setting '" + k + "'");
        if (!known.count(k)) throw std::runtime_error("unknown setting '" + k + "'");
        if (!val.emplace(k, v).second) throw std::runtime_error("duplicate setting '" + k + "'");
    }
    for (const char* r : keys)
        if (!val.count(r)) throw std::runtime_error(std::string("missing setting '") + r + "'");
    auto pos = [&](const char* s) {
        const double x = val.at(s);
        if (!std::isfinite(x) || x <= 0.0) throw std::runtime_error(std::string(s) + " must be finite and positive");
        return x;
    };
    auto nonneg = [&](const char* s) {
        const double x = val.at(s);
        if (!std::isfinite(x) || x < 0.0) throw std::runtime_error(std::string(s) + " must be finite and non-negative");
        return x;
    };
    m.density = pos("volume_mesh_density");
    m.youngs = pos("volume_mesh_YoungsModulus");
    m.poisson = val.at("poisson_rate");
    if (!std::isfinite(m.poisson) || m.poisson <= -1.0 || m.poisson >= 0.5) throw std::runtime_error("poisson_rate must be in (-1, 0.5)");
    m.friction = nonneg("friction_rate");
    m.cloth_thickness = pos("triangle(cloth)_mesh_thickness");
    m.cloth_youngs = pos("triangle(cloth)_mesh_YoungsModulus");
    m.shear_youngs = pos("triangle(cloth)_mesh_shearYoungsModulus");
    m.bend_youngs = nonneg("triangle(cloth)_mesh_bendingYoungsModulus");
    m.cloth_density = pos("triangle(cloth)_mesh_density");
    m.strain_rate = nonneg("strainRate");
    const double coll = val.at("enable_collision_handling");
    if (coll != 0.0 && coll != 1.0) throw std::runtime_error("enable_collision_handling must be 0 or 1");
    m.use_barrier = coll != 0.0;
    m.drag = nonneg("drag_coeff");
    m.dt = pos("ipc_time_step");
    m.newton_threshold = pos("Newton_solver_threshold");
    const double rel = pos("IPC_ralative_dHat");
    m.hhat = rel * rel;
    m.fhat = 1e-6;
    m.kappa = 0.0;
    m.dtol = 1e-18;
    update_material(m);
}

void prepare_quadratic_bending(Mesh& m) {
    m.bend_verts.clear();
    m.bend_q.clear();
    std::unordered_map<int64_t, int> first_outer;
    for (const auto& t : m.triangles)
        for (int i = 0; i < 3; ++i) {
            int e0 = t[i], e1 = t[(i + 1) % 3];
            const int other = t[(i + 2) % 3];
            if (e0 > e1) std::swap(e0, e1);
            const int64_t key = ((int64_t)e0 << 32) | e1;
            const auto it = first_outer.find(key);
            if (it == first_outer.end()) { first_outer[key] = other; continue; }
            const int o1 = it->second, o2 = other;
            const Vec3 v0 = m.x[o1], v1 = m.x[e0], v2 = m.x[e1], v3 = m.x[o2];
            const Vec3 E0 = sub(v2, v1), E1 = sub(v0, v1), E2 = sub(v3, v1), E3 = sub(v0, v2), E4 = sub(v3, v2);
            const double a0 = 0.5 * norm(cross(E0, E1)), a1 = 0.5 * norm(cross(E0, E2));
            auto cot = [](const Vec3& a, const Vec3& b) { return dot(a, b) / norm(cross(a, b)); };
            const Vec3 mE0 = scaled(-1.0, E0);
            const double c01 = cot(E0, E1), c02 = cot(E0, E2), c03 = cot(mE0, E3), c04 = cot(mE0, E4);
            const double K[4] = {c03 + c04, c01 + c02, -c01 - c03, -c02 - c04};
            const double s = 3.0 / (a0 + a1);
            std::array<double, 16> Q;
            for (int r = 0; r < 4; ++r)
                for (int c = 0; c < 4; ++c) Q[4 * r + c] = s * K[r] * K[c];
            m.bend_verts.push_back({{e0, e1, o1, o2}});
            m.bend_q.push_back(Q);
        }
}

Vec3 twisting_direction(const Vec3& x, const Vec3& rest, double fraction, double dt) {
    const double a = 3.14 / 5 * dt * fraction;
    const double c = std::cos(a), s = std::sin(a);
    // rotation about x by -a (rest x < 0) or +a (rest x > 0); the update moves x to R x
    double sg;
    if (rest[0] < 0.0) sg = 1.0;
    else if (rest[0] > 0.0) sg = -1.0;
    else return {{0, 0, 0}};
    const double R[3][3] = {{1, 0, 0}, {0, c, sg * s}, {0, -(sg * s), c}};
    Vec3 out;
    for (int i = 0; i < 3; ++i) out[i] = (R[i][0] * x[0] + R[i][1] * x[1] + R[i][2] * x[2]) - x[i];
    return out;
}

void build_scene(Mesh& m, const std::string& scene, const std::string& asset_dir, const std::string& settings_path) {
    m = Mesh{};
    load_settings(m, settings_path);
    std::string err;
    auto need = [&](bool ok) {
        if (!ok) throw std::runtime_error(err);
    };
    if (scene == "twisting-mat" || scene == "twisting-mat-soft") {
        // buildTwistingMatScene (Simulator.cpp:265-325)
        m.youngs = 1e6;
        update_material(m);
        m.obj_max = m.obj_min = {{0, 0, 0}};
        m.quasi_static = true;
        m.apply_gravity = false;
        need(load_tetrahedra(m, asset_dir + "tetrahedraMesh/ipcmesh/mat40x40.msh", 1, {{0, 0, 0}}, err));
        const double eps = 1e-4;
        for (int i = 0; i < (int)m.x.size(); ++i)
            if (m.x[i][0] < m.obj_min[0] + eps || m.x[i][0] > m.obj_max[0] - eps) {
                if (scene == "twisting-mat") { m.dirichlet.push_back(i); m.boundary[i] = 1; }
                else m.soft.push_back(i);
            }
        if (scene == "twisting-mat-soft") m.soft_weight = 100.0;
        m.motion = 1;
    } else if (scene == "bunny2") {
        // buildBunny2Scene (Simulator.cpp:380-396)
        m.youngs = 1e5;
        update_material(m);
        m.obj_max = m.obj_min = {{0, 0, 0}};
        need(load_tetrahedra(m, asset_dir + "tetrahedraMesh/bunny2.msh", 0.2, {{0.0, 0.65, 0.0}}, err));
        need(load_tetrahedra(m, asset_dir + "tetrahedraMesh/bunny2.msh", 0.2, {{0, 0, 0}}, err));
    } else if (scene == "cloth-bunny") {
        // buildClothOverBunnyScene (Simulator.cpp:327-378), quadratic bending
        m.youngs = 1e4;
        update_material(m);
        m.obj_max = m.obj_min = {{0, 0, 0}};
        need(load_triangles(m, asset_dir + "triangleMesh/planes/plane1024.obj", 1, {{0, 0, 0}}, 0, err));
        prepare_quadratic_bending(m);
        const double a = 3.14159265358979323846 / 2;   // pi / 2 about x
        const double c = std::cos(a), s = std::sin(a);
        Vec3 lo{{1e32, 1e32, 1e32}}, hi{{-1e32, -1e32, -1e32}};
        for (Vec3& p : m.x) {
            const Vec3 q{{p[0], c * p[1] + -s * p[2], s * p[1] + c * p[2]}};
            p = q;
            for (int d = 0; d < 3; ++d) { if (lo[d] > p[d]) lo[d] = p[d]; if (hi[d] < p[d]) hi[d] = p[d]; }
        }
        m.max_corner = hi;
        m.min_corner = lo;
        need(load_tetrahedra(m, asset_dir + "tetrahedraMesh/bunny.msh", 0.5, {{0, -0.5, 0}}, err));
    } else {
        throw std::invalid_argument("unknown scene: " + scene);
    }
    finalize_scene(m);
}

void finalize_scene(Mesh& m) {
    // buildModels after the scene (Simulator.cpp:468-506)
    init_mesh(m);
    m.v_rest = m.x;
    m.V_prev = m.x;
    // updateInertialTarget (IPCSolver.cpp:55-70) with zero velocities
    const double g = m.apply_gravity ? -9.8 : 0.0;
    m.xtilde.resize(m.x.size());
    for (size_t i = 0; i < m.x.size(); ++i)
        for (int d = 0; d < 3; ++d) m.xtilde[i][d] = (m.V_prev[i][d] + m.velocities[i][d] * m.dt) + (d == 1 ? (g * m.dt) * m.dt : 0.0);
    const Vec3 diag = sub(m.max_corner, m.min_corner);
    m.bbox_diag2 = dot(diag, diag);
    m.hhat *= m.bbox_diag2;
    m.fhat *= m.bbox_diag2;
    m.dtol *= m.bbox_diag2;
    get_surface(m);
    if (m.surf_edges.empty()) throw std::runtime_error("simulation surface has no edges");
    double len = 0;
    for (const auto& e : m.surf_edges) len += norm(sub(m.x[e.first], m.x[e.second]));
    m.avg_edge_len = len / (m.surf_edges.size() * 3);
}

bool write_msh(const Mesh& m, const std::string& path) {
    std::ofstream o(path);
    o << "$Nodes\n" << m.x.size() << "\n";
    for (size_t i = 0; i < m.x.size(); ++i) o << i + 1 << " " << m.x[i][0] << " " << m.x[i][1] << " " << m.x[i][2] << "\n";
    o << "$EndNodes\n$Elements\n" << m.tets.size() << "\n";
    for (size_t i = 0; i < m.tets.size(); ++i)
        o << i + 1 << " 4 0 " << m.tets[i][0] + 1 << " " << m.tets[i][1] + 1 << " " << m.tets[i][2] + 1 << " " << m.tets[i][3] + 1 << "\n";
    o << "$EndElements\n";
    return o.good();
}

bool save_checkpoint(const Mesh& m, const std::string& dir) {
    std::ofstream a(dir + "/vertex.txt"), b(dir + "/vertexXtile.txt");
    if (!a || !b) return false;
    a << std::setprecision(17);
    b << std::setprecision(17);
    for (const Vec3& p : m.x) a << p[0] << " " << p[1] << " " << p[2] << "\n";
    for (const Vec3& p : m.xtilde) b << p[0] << " " << p[1] << " " << p[2] << "\n";
    a.flush();
    b.flush();
    return a.good() && b.good();
}

bool load_checkpoint(Mesh& m, const std::string& dir) {
    std::ifstream a(dir + "/vertex.txt"), b(dir + "/vertexXtile.txt");
    if (!a || !b) return false;
    const size_t V = m.x.size();
    if (m.V_prev.size() != V || m.xtilde.size() != V) return false;
    std::vector<Vec3> x(V), xt(V);
    for (size_t i = 0; i < V; ++i)
        if (!(a >> x[i][0] >> x[i][1] >> x[i][2])) return false;
    for (size_t i = 0; i < V; ++i)
        if (!(b >> xt[i][0] >> xt[i][1] >> xt[i][2])) return false;
    double extra;
    if (a >> extra || b >> extra) return false;   // written for another vertex count
    m.x = x;
    m.V_prev = x;
    m.xtilde = xt;
    return true;
}

}  // namespace host
}  // namespace cipc
