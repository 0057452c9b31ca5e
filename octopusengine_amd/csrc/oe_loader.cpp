// Mesh input from files, for the formats the reference reads
// (src/Mesh/Converter/MeshLoader.cpp:27-237):
//   * legacy-VTK ASCII UNSTRUCTURED_GRID with POINTS, CELLS n+1 m,
//     OFFSETS / CONNECTIVITY (vtktypeint64) and CELL_TYPES, as VTK 5.1 and the
//     reference's own writer produce; VTK cell types 10, 12, 24, 29, 71 map to
//     Tetra, Hexa, Tetra10, Hexa27, Tetra20 and triquadratic hexahedra are
//     brought back from VTK's node order to the engine's;
//   * Gmsh 2.2 ASCII (.msh): $Nodes then $Elements with element types
//     4 (Tetra), 8 (Hexa), 11 (Tetra10), 29 (Tetra20); node ids start at 1, or
//     at 0 when the first node is numbered 0.
// Coordinates are read as float, as the reference's loader does.
#include <fstream>
#include <sstream>
#include <stdexcept>

#include "oe_host.hpp"

namespace oe {

namespace {

bool starts_with(const std::string& s, const char* p) { return s.rfind(p, 0) == 0; }

bool next_nonempty(std::istream& in, std::string& line) {
    while (std::getline(in, line)) {
        if (!line.empty() && line != "\r") return true;
    }
    return false;
}

bool skip_to(std::istream& in, std::string& line, const char* tag) {
    while (std::getline(in, line))
        if (starts_with(line, tag)) return true;
    return false;
}

Element from_vtk_type(int t) {
    switch (t) {
        case 10: return Tetra;
        case 12: return Hexa;
        case 24: return Tetra10;
        case 29: return Hexa27;
        case 71: return Tetra20;
        default: throw std::invalid_argument("unsupported VTK cell type " + std::to_string(t));
    }
}

Element from_gmsh_type(int t) {
    switch (t) {
        case 4: return Tetra;
        case 8: return Hexa;
        case 11: return Tetra10;
        case 29: return Tetra20;
        default: throw std::invalid_argument("unsupported Gmsh element type " + std::to_string(t));
    }
}

void check_ids(const LoadedMesh& m) {
    const int nv = static_cast<int>(m.xyz.size() / 3);
    for (const auto& kv : m.topologies)
        for (int32_t v : kv.second)
            if (v < 0 || v >= nv) throw std::invalid_argument("element refers to a missing vertex");
}

}  // namespace

LoadedMesh load_vtk(const std::string& path) {
    std::ifstream in(path);
    if (!in) throw std::invalid_argument("cannot read " + path);
    std::string line;
    for (int i = 0; i < 3; ++i)
        if (!next_nonempty(in, line)) throw std::invalid_argument("truncated VTK header");
    if (!starts_with(line, "ASCII")) throw std::invalid_argument("only ASCII VTK files are read");
    if (!next_nonempty(in, line) || !starts_with(line, "DATASET"))
        throw std::invalid_argument("missing DATASET");
    if (line.find("UNSTRUCTURED_GRID") == std::string::npos)
        throw std::invalid_argument("only UNSTRUCTURED_GRID datasets are read");
    LoadedMesh m;
    if (!next_nonempty(in, line) || !starts_with(line, "POINTS")) throw std::invalid_argument("missing POINTS");
    int np = 0;
    {
        std::istringstream s(line.substr(6));
        s >> np;
    }
    if (np < 0) throw std::invalid_argument("bad POINTS count");
    m.xyz.resize(3 * static_cast<size_t>(np));
    for (size_t i = 0; i < m.xyz.size(); ++i)
        if (!(in >> m.xyz[i])) throw std::invalid_argument("truncated POINTS");
    if (!skip_to(in, line, "CELLS")) throw std::invalid_argument("missing CELLS");
    long long ncells1 = 0, nids = 0;
    {
        std::istringstream s(line.substr(5));
        s >> ncells1 >> nids;
    }
    if (!skip_to(in, line, "CONNECTIVITY")) throw std::invalid_argument("missing CONNECTIVITY");
    std::vector<int64_t> ids(static_cast<size_t>(nids));
    for (auto& v : ids)
        if (!(in >> v)) throw std::invalid_argument("truncated CONNECTIVITY");
    if (!skip_to(in, line, "CELL_TYPES")) throw std::invalid_argument("missing CELL_TYPES");
    int nc = 0;
    {
        std::istringstream s(line.substr(10));
        s >> nc;
    }
    size_t at = 0;
    for (int c = 0; c < nc; ++c) {
        int t = 0;
        if (!(in >> t)) throw std::invalid_argument("truncated CELL_TYPES");
        const Element e = from_vtk_type(t);
        const int S = elem_nb_vertices(e);
        if (at + S > ids.size()) throw std::invalid_argument("CONNECTIVITY shorter than CELL_TYPES");
        auto& topo = m.topologies[e];
        for (int j = 0; j < S; ++j) {
            // the Hexa27 reorder is an involution: VTK position j holds engine node vtk_hexa27_node(j)
            const int src = e == Hexa27 ? vtk_hexa27_node(j) : j;
            topo.push_back(static_cast<int32_t>(ids[at + src]));
        }
        at += S;
    }
    check_ids(m);
    return m;
}

LoadedMesh load_msh(const std::string& path) {
    std::ifstream in(path);
    if (!in) throw std::invalid_argument("cannot read " + path);
    std::string line;
    if (!skip_to(in, line, "$Nodes")) throw std::invalid_argument("missing $Nodes");
    if (!std::getline(in, line)) throw std::invalid_argument("truncated $Nodes");
    LoadedMesh m;
    int off = 1;
    bool first = true;
    while (std::getline(in, line)) {
        if (starts_with(line, "$EndNodes")) break;
        std::istringstream s(line);
        long long n = 0;
        float x = 0, y = 0, z = 0;
        if (!(s >> n >> x >> y >> z)) continue;
        if (first && n == 0) off = 0;
        first = false;
        m.xyz.push_back(x);
        m.xyz.push_back(y);
        m.xyz.push_back(z);
    }
    if (!skip_to(in, line, "$Elements")) throw std::invalid_argument("missing $Elements");
    if (!std::getline(in, line)) throw std::invalid_argument("truncated $Elements");
    while (std::getline(in, line)) {
        if (starts_with(line, "$EndElements")) break;
        std::istringstream s(line);
        long long num = 0;
        int type = 0, ntags = 0;
        if (!(s >> num >> type >> ntags)) continue;
        for (int i = 0; i < ntags; ++i) {
            long long tag;
            s >> tag;
        }
        const Element e = from_gmsh_type(type);
        auto& topo = m.topologies[e];
        for (int j = 0; j < elem_nb_vertices(e); ++j) {
            long long v = 0;
            if (!(s >> v)) throw std::invalid_argument("truncated element line");
            topo.push_back(static_cast<int32_t>(v - off));
        }
    }
    check_ids(m);
    return m;
}

LoadedMesh load_mesh(const std::string& path) {
    const auto dot = path.find_last_of('.');
    const std::string ext = dot == std::string::npos ? "" : path.substr(dot + 1);
    if (ext == "vtk") return load_vtk(path);
    if (ext == "msh") return load_msh(path);
    throw std::invalid_argument("unknown mesh file extension: " + path);
}

}  // namespace oe
