// Structured beam meshes and high-order conversions.
//
// Vertex numbering, element node order and the float arithmetic used to place
// new nodes follow the reference generators exactly, because vertex ids decide
// the colouring tie-breaks and node order decides which shape function a node
// gets:
//   grid + cell corner order   src/Mesh/Generator/BeamGenerator.cpp:5-43
//   6-tetrahedra split table   src/Mesh/Generator/BeamGenerator.cpp:86-92
//   P1 -> P2                   src/Mesh/Generator/BeamGenerator.cpp:172-214
//   P1 -> P3                   src/Mesh/Generator/BeamGenerator.cpp:216-295
//   Q1 -> Q2                   src/Mesh/Generator/BeamGenerator.cpp:297-377
//   edge / face orders         include/Mesh/Converter/MeshConverter.h:76-82,129-140
// New nodes get ids in first-encounter order (element by element, edge by
// edge); shared edges/faces are looked up by their sorted vertex ids.
// This file is compiled with -ffp-contract=off so that a*b + c*d is rounded
// like the reference's float expressions.
#include "oe_host.hpp"

#include <algorithm>
#include <array>
#include <stdexcept>
#include <unordered_map>

namespace oe {

namespace {

struct KeyHash {
    size_t operator()(const std::array<int32_t, 4>& k) const {
        uint64_t h = 1469598103934665603ull;
        for (int32_t v : k) { h ^= static_cast<uint32_t>(v); h *= 1099511628211ull; }
        return static_cast<size_t>(h);
    }
};

std::array<int32_t, 4> sorted_key(std::initializer_list<int32_t> ids) {
    std::array<int32_t, 4> k{-1, -1, -1, -1};
    int i = 0;
    for (int32_t v : ids) k[i++] = v;
    std::sort(k.begin(), k.begin() + i);
    return k;
}

void push3(std::vector<float>& g, float x, float y, float z) {
    g.push_back(x); g.push_back(y); g.push_back(z);
}

}  // namespace

HostMesh beam_mesh(Element element, int nx, int ny, int nz, float sx, float sy, float sz) {
    if (nx < 1 || ny < 1 || nz < 1) throw std::invalid_argument("beam needs at least one cell per axis");
    HostMesh m;
    const bool hexa = (element == Hexa || element == Hexa27);
    const bool tet = (element == Tetra || element == Tetra10 || element == Tetra20);
    if (!hexa && !tet) throw std::invalid_argument("unsupported element for beam generator");
    m.element = hexa ? Hexa : Tetra;
    const int X = nx + 1, Y = ny + 1, Z = nz + 1;
    const float hx = sx / static_cast<float>(X - 1);
    const float hy = sy / static_cast<float>(Y - 1);
    const float hz = sz / static_cast<float>(Z - 1);
    m.xyz.reserve(static_cast<size_t>(X) * Y * Z * 3);
    for (int z = 0; z < Z; ++z)
        for (int y = 0; y < Y; ++y)
            for (int x = 0; x < X; ++x)
                push3(m.xyz, static_cast<float>(x) * hx, static_cast<float>(y) * hy, static_cast<float>(z) * hz);
    auto id = [&](int x, int y, int z) { return x + y * X + z * Y * X; };
    // tetra split of a cell, indices into the 8 cell corners
    static const int tet6[24] = {0, 4, 6, 5, 3, 6, 2, 0, 0, 4, 7, 6, 3, 6, 0, 7, 2, 0, 6, 1, 6, 0, 5, 1};
    for (int z = 0; z < Z - 1; ++z)
        for (int y = 0; y < Y - 1; ++y)
            for (int x = 0; x < X - 1; ++x) {
                const int c[8] = {id(x, y, z),         id(x + 1, y, z),     id(x + 1, y, z + 1), id(x, y, z + 1),
                                  id(x, y + 1, z),     id(x + 1, y + 1, z), id(x + 1, y + 1, z + 1),
                                  id(x, y + 1, z + 1)};
                if (hexa) {
                    for (int i = 0; i < 8; ++i) m.topo.push_back(c[i]);
                } else {
                    for (int i = 0; i < 24; ++i) m.topo.push_back(c[tet6[i]]);
                }
            }
    if (element == Tetra10) tetra4_to_tetra10(m);
    if (element == Tetra20) tetra4_to_tetra20(m);
    if (element == Hexa27) hexa_to_hexa27(m);
    return m;
}

void tetra4_to_tetra10(HostMesh& m) {
    if (m.element != Tetra) throw std::invalid_argument("tetra4_to_tetra10 needs a P1 mesh");
    static const int edges[6][2] = {{0, 1}, {1, 2}, {0, 2}, {0, 3}, {1, 3}, {2, 3}};
    std::unordered_map<std::array<int32_t, 4>, int32_t, KeyHash> emap;
    emap.reserve(m.topo.size() * 2);
    std::vector<int32_t> out;
    out.reserve(m.topo.size() / 4 * 10);
    for (size_t i = 0; i < m.topo.size(); i += 4) {
        int32_t ids[10];
        for (int j = 0; j < 4; ++j) ids[j] = m.topo[i + j];
        for (int k = 0; k < 6; ++k) {
            int32_t a = ids[edges[k][0]], b = ids[edges[k][1]];
            if (a > b) std::swap(a, b);
            const auto key = sorted_key({a, b});
            auto it = emap.find(key);
            int32_t nid;
            if (it != emap.end()) {
                nid = it->second;
            } else {
                nid = static_cast<int32_t>(m.xyz.size() / 3);
                const float px = 0.5f * (m.xyz[3 * a] + m.xyz[3 * b]);
                const float py = 0.5f * (m.xyz[3 * a + 1] + m.xyz[3 * b + 1]);
                const float pz = 0.5f * (m.xyz[3 * a + 2] + m.xyz[3 * b + 2]);
                push3(m.xyz, px, py, pz);
                emap.emplace(key, nid);
            }
            ids[4 + k] = nid;
        }
        out.insert(out.end(), ids, ids + 10);
    }
    m.topo.swap(out);
    m.element = Tetra10;
}

void tetra4_to_tetra20(HostMesh& m) {
    if (m.element != Tetra) throw std::invalid_argument("tetra4_to_tetra20 needs a P1 mesh");
    static const int edges[6][2] = {{0, 1}, {1, 2}, {2, 0}, {0, 3}, {1, 3}, {2, 3}};
    static const int faces[4][3] = {{0, 1, 3}, {1, 2, 3}, {0, 3, 2}, {0, 2, 1}};
    // edge map: sorted key -> (first vertex as created, node at 1/3, node at 2/3)
    struct EdgeNodes { int32_t first; int32_t n[2]; };
    std::unordered_map<std::array<int32_t, 4>, EdgeNodes, KeyHash> emap;
    std::unordered_map<std::array<int32_t, 4>, int32_t, KeyHash> fmap;
    emap.reserve(m.topo.size() * 2);
    fmap.reserve(m.topo.size() * 2);
    std::vector<int32_t> out;
    out.reserve(m.topo.size() / 4 * 20);
    for (size_t i = 0; i < m.topo.size(); i += 4) {
        int32_t ids[20];
        for (int j = 0; j < 4; ++j) ids[j] = m.topo[i + j];
        int j = 4;
        for (int k = 0; k < 6; ++k) {
            const int32_t a = ids[edges[k][0]], b = ids[edges[k][1]];
            const auto key = sorted_key({a, b});
            auto it = emap.find(key);
            int32_t n0, n1;
            if (it != emap.end()) {
                n0 = it->second.n[0]; n1 = it->second.n[1];
                if (it->second.first != a) std::swap(n0, n1);   // traversed the other way
            } else {
                int32_t nn[2];
                for (int w = 0; w < 2; ++w) {
                    const float wt = static_cast<float>(w + 1) / 3.f;
                    const float om = 1.f - wt;
                    nn[w] = static_cast<int32_t>(m.xyz.size() / 3);
                    const float px = m.xyz[3 * a] * om + m.xyz[3 * b] * wt;
                    const float py = m.xyz[3 * a + 1] * om + m.xyz[3 * b + 1] * wt;
                    const float pz = m.xyz[3 * a + 2] * om + m.xyz[3 * b + 2] * wt;
                    push3(m.xyz, px, py, pz);
                }
                emap.emplace(key, EdgeNodes{a, {nn[0], nn[1]}});
                n0 = nn[0]; n1 = nn[1];
            }
            ids[j++] = n0;
            ids[j++] = n1;
        }
        for (int k = 0; k < 4; ++k) {
            const int32_t a = ids[faces[k][0]], b = ids[faces[k][1]], c = ids[faces[k][2]];
            const auto key = sorted_key({a, b, c});
            auto it = fmap.find(key);
            int32_t fid;
            if (it != fmap.end()) {
                fid = it->second;
            } else {
                const float w = 1.f / 3.f;
                fid = static_cast<int32_t>(m.xyz.size() / 3);
                float p[3];
                for (int d = 0; d < 3; ++d) p[d] = m.xyz[3 * a + d] * w + m.xyz[3 * b + d] * w + m.xyz[3 * c + d] * w;
                push3(m.xyz, p[0], p[1], p[2]);
                fmap.emplace(key, fid);
            }
            ids[j++] = fid;
        }
        out.insert(out.end(), ids, ids + 20);
    }
    m.topo.swap(out);
    m.element = Tetra20;
}

void hexa_to_hexa27(HostMesh& m) {
    if (m.element != Hexa) throw std::invalid_argument("hexa_to_hexa27 needs a Q1 mesh");
    static const int edges[12][2] = {{0, 1}, {1, 2}, {2, 3}, {3, 0}, {0, 4}, {1, 5},
                                     {2, 6}, {3, 7}, {4, 5}, {5, 6}, {6, 7}, {7, 4}};
    static const int quads[6][4] = {{0, 1, 2, 3}, {4, 5, 1, 0}, {5, 6, 2, 1},
                                    {3, 2, 6, 7}, {7, 4, 0, 3}, {7, 6, 5, 4}};
    std::unordered_map<std::array<int32_t, 4>, int32_t, KeyHash> emap, fmap;
    emap.reserve(m.topo.size() * 2);
    fmap.reserve(m.topo.size());
    std::vector<int32_t> out;
    out.reserve(m.topo.size() / 8 * 27);
    for (size_t i = 0; i < m.topo.size(); i += 8) {
        int32_t ids[27];
        for (int j = 0; j < 8; ++j) ids[j] = m.topo[i + j];
        int j = 8;
        for (int k = 0; k < 12; ++k) {
            const int32_t a = ids[edges[k][0]], b = ids[edges[k][1]];
            const auto key = sorted_key({a, b});
            auto it = emap.find(key);
            int32_t nid;
            if (it != emap.end()) {
                nid = it->second;
            } else {
                nid = static_cast<int32_t>(m.xyz.size() / 3);
                float p[3];
                for (int d = 0; d < 3; ++d) p[d] = 0.5f * (m.xyz[3 * a + d] + m.xyz[3 * b + d]);
                push3(m.xyz, p[0], p[1], p[2]);
                emap.emplace(key, nid);
            }
            ids[j++] = nid;
        }
        for (int k = 0; k < 6; ++k) {
            const int32_t a = ids[quads[k][0]], b = ids[quads[k][1]], c = ids[quads[k][2]], d4 = ids[quads[k][3]];
            const auto key = sorted_key({a, b, c, d4});
            auto it = fmap.find(key);
            int32_t fid;
            if (it != fmap.end()) {
                fid = it->second;
            } else {
                const float w = 1.f / 4.f;
                fid = static_cast<int32_t>(m.xyz.size() / 3);
                float p[3];
                for (int d = 0; d < 3; ++d)
                    p[d] = m.xyz[3 * a + d] * w + m.xyz[3 * b + d] * w + m.xyz[3 * c + d] * w + m.xyz[3 * d4 + d] * w;
                push3(m.xyz, p[0], p[1], p[2]);
                fmap.emplace(key, fid);
            }
            ids[j++] = fid;
        }
        float cc[3] = {0.f, 0.f, 0.f};
        for (int k = 0; k < 8; ++k)
            for (int d = 0; d < 3; ++d) cc[d] += m.xyz[3 * ids[k] + d];
        for (int d = 0; d < 3; ++d) cc[d] /= 8.f;
        ids[j] = static_cast<int32_t>(m.xyz.size() / 3);
        push3(m.xyz, cc[0], cc[1], cc[2]);
        out.insert(out.end(), ids, ids + 27);
    }
    m.topo.swap(out);
    m.element = Hexa27;
}

}  // namespace oe
