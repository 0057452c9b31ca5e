// Legacy-VTK 5.1 ASCII writer for unstructured grids, with the same file
// layout as the reference's VTK_Formater (src/Mesh/Converter/VTK_Formater.cpp):
//   header; POINTS; CELLS n+1 / OFFSETS / CONNECTIVITY (vtktypeint64);
//   CELL_TYPES; then POINT_DATA / CELL_DATA blocks of SCALARS and VECTORS.
// Numbers use the C++ stream defaults (6 significant digits), so files written
// from the same data are byte-identical to the reference's.  Triquadratic
// hexahedra are written in VTK's node order, which differs from the engine's
// (the reference's) Hexa27 order in the edge and face nodes.
#include <fstream>
#include <sstream>
#include <stdexcept>

#include "oe_host.hpp"

namespace oe {

int vtk_cell_type(Element e) {
    switch (e) {
        case Tetra: return 10;
        case Hexa: return 12;
        case Tetra10: return 24;
        case Tetra20: return 71;
        case Hexa27: return 29;
        default: return 0;
    }
}

// VTK position j of a Hexa27 cell takes engine node vtk_hexa27_node(j): the
// edge nodes 12-15 and 16-19 trade places, and the face nodes 20, 21, 22, 24
// are permuted (VTK_Formater.cpp:54-90).
int vtk_hexa27_node(int j) {
    switch (j) {
        case 12: return 16;
        case 13: return 17;
        case 14: return 18;
        case 15: return 19;
        case 16: return 12;
        case 17: return 13;
        case 18: return 14;
        case 19: return 15;
        case 20: return 24;
        case 21: return 22;
        case 22: return 21;
        case 24: return 20;
        default: return j;
    }
}

VtkWriter::VtkWriter(const std::string& path) : out_(path) {
    if (!out_) throw std::invalid_argument("cannot open " + path + " for writing");
    out_ << "# vtk DataFile Version 5.1\n";
    out_ << "vtk output\n";
    out_ << "ASCII\n\n";
}

void VtkWriter::mesh(Element e, const float* xyz, int nv, const int32_t* topo, int ne) {
    const int S = element_supported(e) ? elem_nb_vertices(e) : 0;
    if (S <= 0 || nv < 0 || ne < 0) throw std::invalid_argument("bad mesh");
    nb_vertices_ = nv;
    nb_cells_ = ne;
    out_ << "DATASET UNSTRUCTURED_GRID\n";
    out_ << "POINTS " << nv << " float\n";
    for (int i = 0; i < nv; ++i) out_ << xyz[3 * i] << " " << xyz[3 * i + 1] << " " << xyz[3 * i + 2] << "\n";
    out_ << "\n";
    const long long nidx = static_cast<long long>(ne) * S;
    out_ << "CELLS " << ne + 1 << " " << nidx << "\n";
    out_ << "OFFSETS vtktypeint64\n";
    if (ne > 0) {
        for (long long off = 0; off <= nidx; off += S) out_ << off << " ";
        out_ << "\n";
    }
    out_ << "CONNECTIVITY vtktypeint64\n";
    for (int c = 0; c < ne; ++c) {
        const int32_t* t = topo + static_cast<size_t>(c) * S;
        for (int j = 0; j < S; ++j) out_ << t[e == Hexa27 ? vtk_hexa27_node(j) : j] << " ";
        out_ << "\n";
    }
    out_ << "CELL_TYPES " << ne << "\n";
    if (ne > 0) {
        const int id = vtk_cell_type(e);
        for (int c = 0; c < ne; ++c) out_ << id << " ";
        out_ << "\n";
    }
    out_ << "\n";
}

void VtkWriter::point_data() {
    out_ << "POINT_DATA " << nb_vertices_ << "\n";
    mode_ = 1;
}

void VtkWriter::cell_data() {
    out_ << "CELL_DATA " << nb_cells_ << "\n";
    mode_ = 2;
}

void VtkWriter::check_size(int n) const {
    if (mode_ == 0) throw std::logic_error("start point or cell data first");
    if (n != (mode_ == 1 ? nb_vertices_ : nb_cells_)) throw std::invalid_argument("data size does not match");
}

void VtkWriter::scalars(const std::string& name, const float* data, int n) {
    check_size(n);
    out_ << "SCALARS " << name << " float 1\n";
    out_ << "LOOKUP_TABLE default\n";
    for (int i = 0; i < n; ++i) out_ << data[i] << " ";
    out_ << "\n";
}

void VtkWriter::vectors(const std::string& name, const float* data, int n) {
    check_size(n);
    out_ << "VECTORS " << name << " float\n";
    for (int i = 0; i < n; ++i) out_ << data[3 * i] << " " << data[3 * i + 1] << " " << data[3 * i + 2] << "  ";
    out_ << "\n";
}

void VtkWriter::close() {
    out_.close();
    if (out_.fail()) throw std::runtime_error("error while writing the VTK file");
}

}  // namespace oe
