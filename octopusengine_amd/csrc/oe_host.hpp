// Host-side (CPU) part of the engine: element shapes, beam meshes and their
// high-order conversions, graph colouring, per-(element, quadrature) constants,
// masses and the colour-major owners CSR that the HIP kernels consume.
//
// Everything here runs once per scene at set-up time; nothing here is on the
// per-step path.
#pragma once

#include <cstdint>
#include <fstream>
#include <map>
#include <string>
#include <vector>

#include "oe_q2_tensor.hpp"

namespace oe {

// Element ids keep the integer values of the reference enum
// (octopus/include/Mesh/Elements.h:4-15) so that a caller can pass the
// reference's own enum values through the C ABI unchanged.
enum Element : int {
    Line = 0, Triangle = 1, Quad = 2,
    Tetra = 3, Pyramid = 4, Prism = 5, Hexa = 6,
    Tetra10 = 7, Tetra20 = 8, Hexa27 = 9
};

// Material ids keep the reference enum values
// (octopus/include/Dynamic/FEM/ContinuousMaterial.h:8-10).
enum Material : int { Hooke = 0, StVK = 1, NeoHooke = 2, Stable_NeoHooke = 3 };

// Mass distribution, reference octopus/include/Dynamic/FEM/FEM_Shape.h:693-696.
enum Mass_Distribution : int { Uniform = 0, FemShape = 1 };

bool element_supported(int e);
int elem_nb_vertices(Element e);
int elem_nb_quadrature(Element e);
std::string element_name(Element e);

// Reference element: node coordinates, quadrature rule and the shape-function
// gradients at every quadrature point.
struct ShapeTable {
    Element element;
    int S = 0;                        // nodes per element
    int Q = 0;                        // quadrature points
    std::vector<double> nodes;        // S x 3 reference coordinates
    std::vector<double> qpoints;      // Q x 3
    std::vector<double> weights;      // Q
    std::vector<double> dN;           // Q x S x 3 : dN[(q*S + k)*3 + c]
    std::vector<double> coeff;        // basis coefficients (monomial x node)
    std::vector<int> mono;            // monomial exponents, S x 3
    // Shape function values N_k at an arbitrary point (used by tests).
    void eval(const double* x, double* N, double* dNdx) const;
};

const ShapeTable& shape_table(Element e);

// A mesh with one element type.  xyz is float, as in the reference geometry
// (std::vector<glm::vec3>).
struct HostMesh {
    Element element = Tetra;
    std::vector<float> xyz;     // 3 * N
    std::vector<int32_t> topo;  // S * E
    int nb_vertices() const { return static_cast<int>(xyz.size() / 3); }
    int nb_elements() const { return static_cast<int>(topo.size() / elem_nb_vertices(element)); }
};

// Structured beam generators (reference src/Mesh/Generator/BeamGenerator.cpp:5-48,86-92)
// and high-order conversions (:172-377).  `element` is the final element type:
// Tetra/Tetra10/Tetra20 use the 6-tetrahedra split of every cell, Hexa/Hexa27
// use one hexahedron per cell.
HostMesh beam_mesh(Element element, int nx, int ny, int nz, float sx, float sy, float sz);
void tetra4_to_tetra10(HostMesh& m);
void tetra4_to_tetra20(HostMesh& m);
void hexa_to_hexa27(HostMesh& m);

// ---------------------------------------------------------------- colouring
// Adjacency with multiplicity: every co-element pair is inserted once per
// shared element, exactly like the reference's std::multiset adjacency
// (include/Tools/Graph.h:36-54).  Stored as CSR, neighbours sorted ascending,
// duplicates kept adjacent.
struct MultiGraph {
    int n = 0;
    std::vector<int64_t> off;   // n + 1
    std::vector<int32_t> adj;   // sorted per row, with repeats
    int degree(int v) const { return static_cast<int>(off[v + 1] - off[v]); }
};

MultiGraph build_primal_graph(Element e, const std::vector<int32_t>& topo, int nb_vertices);
MultiGraph build_dual_vertex_graph(Element e, const std::vector<int32_t>& topo, int nb_vertices);

struct Coloration {
    int nb_color = 0;
    std::vector<int32_t> colors;
};

Coloration color_dsat(const MultiGraph& g);                       // Graph.h:321-371
void color_reduction(const MultiGraph& g, Coloration& c, int target);  // Graph.h:941-969
Coloration color_greedy_slf(const MultiGraph& g);                 // Graph.h:237-277
int64_t count_conflicts(const MultiGraph& g, const Coloration& c);  // Graph.h:863-881 (sum)

// ------------------------------------------------------------ FEM constants
// Per (element, quadrature point): JX_inv (column-major 3x3) and V = |det J| w
// (reference src/GPU/GPU_FEM.cu:176-214).  Packed as 10 floats per record.
struct FemConstants {
    int S = 0, Q = 0, E = 0;
    std::vector<float> jxv;     // E * Q * 10
    std::vector<float> dN;      // Q * S * 3 (float)
    std::vector<double> elem_volume;  // E
};
FemConstants build_fem_constants(const HostMesh& m);
std::vector<double> element_volumes(const HostMesh& m);

// Per-element records for meshes whose every element is an affine image of
// the reference element (X_k = a + B xi_k for all nodes, to within 8 float
// ulps of the coordinates): rec = E * 10 floats (B^-1 column-major, |det B|)
// and wq = Q quadrature weights, so that V_q = |det B| wq[q].  Returns false
// (outputs untouched) when some element is not affine.
bool build_affine_records(const HostMesh& m, std::vector<float>& rec, std::vector<float>& wq);

// Reference src/Dynamic/FEM/FEM_Shape.cpp:86-121.
std::vector<float> compute_fem_mass(const HostMesh& m, const FemConstants& fc, float density,
                                    Mass_Distribution distrib);

// Lame parameters, reference src/GPU/GPU_FEM.cu:162-163 (float arithmetic).
void lame_parameters(float young, float poisson, float* lambda, float* mu);

// Colour-major owners CSR (reference src/GPU/VBD/GPU_VBD_FEM.cu:594-690):
// slots are the vertices sorted by colour (ascending id inside a colour); each
// slot lists its incident elements in ascending element id, packed as
// (eid << 5) | local_node.
struct OwnersCSR {
    std::vector<int32_t> slot_vid;      // N
    std::vector<int32_t> slot_nb;       // N  incident element count
    std::vector<int32_t> slot_off;      // N  offset into own
    std::vector<int32_t> own;           // E * S packed (eid << 5 | ref)
    std::vector<int32_t> color_begin;   // nb_color + 1 (slot ranges)
    std::vector<int32_t> color_max_nb;  // max incident elements per colour
    std::vector<int64_t> color_incidences;  // sum of incident elements per colour
};
// by_valence: within each colour, vertices with more incident elements first
// (stable, ascending id otherwise); else ascending id.
OwnersCSR build_owners(const HostMesh& m, const Coloration& col, bool by_valence = false);

// Vertex blocks of the fused element-task pass: per colour, runs of
// consecutive slots with at most max_inc incidences and max_slots slots in
// total, 4 ints per block {first slot, slot count, first owners index,
// incidence count}, colour-major; color_begin[c] = first block of colour c.
// Throws std::invalid_argument when one vertex alone has more than max_inc
// incidences.
void build_vertex_blocks(const OwnersCSR& o, int max_inc, int max_slots, std::vector<int32_t>& blocks,
                         std::vector<int32_t>& color_begin);

// Element colour schedule for XPBD (reference src/GPU/PBD/GPU_PBD_FEM.cu:528-553).
struct ElementSchedule {
    std::vector<int32_t> eids;          // E, colour-major, ascending inside a colour
    std::vector<int32_t> color_begin;   // nb_color + 1
};
ElementSchedule build_element_schedule(const Coloration& col, int nb_elements);

// 1D factors of the 27-node hexahedron's shape functions at the Gauss
// coordinates (see oe_q2_tensor.hpp); false if the element's shape table does
// not factor with the node and point orders of that header.
bool q2_basis_1d(q2::Basis1D& out);

// Mesh files (reference VTK_Loader / Msh_Loader,
// src/Mesh/Converter/MeshLoader.cpp:27-237): float coordinates and one
// topology per element type, in the engine's node order.
struct LoadedMesh {
    std::vector<float> xyz;
    std::map<Element, std::vector<int32_t>> topologies;
};
LoadedMesh load_vtk(const std::string& path);
LoadedMesh load_msh(const std::string& path);
LoadedMesh load_mesh(const std::string& path);   // by extension: .vtk or .msh

// Legacy-VTK 5.1 ASCII unstructured-grid writer with the reference's layout
// (VTK_Formater, src/Mesh/Converter/VTK_Formater.cpp): mesh(), then
// point_data() / cell_data() blocks of scalars() and vectors().
int vtk_cell_type(Element e);
int vtk_hexa27_node(int vtk_position);
class VtkWriter {
public:
    explicit VtkWriter(const std::string& path);
    void mesh(Element e, const float* xyz, int nv, const int32_t* topo, int ne);
    void point_data();
    void cell_data();
    void scalars(const std::string& name, const float* data, int n);
    void vectors(const std::string& name, const float* data, int n);   // 3n floats
    void close();

private:
    void check_size(int n) const;
    std::ofstream out_;
    int nb_vertices_ = 0, nb_cells_ = 0;
    int mode_ = 0;   // 1 point data, 2 cell data
};

// ------------------------------------------------------------ JSON record
// The reference's DataRecorder file: {"name" : value,...} in insertion order,
// each value as its recorder writes it (oe_record.cpp).
class JsonRecord {
public:
    void time(int frame, float time, float dt, float fixed_time, float fixed_dt);
    void mesh(const int32_t counts[10], int n_vertices);
    void material(bool xpbd, int material, float poisson, float young, int iterations, int sub_iterations,
                  float density);
    void string(const std::string& name, const std::string& value);
    void flexion(float error, const float p[3], const float target[3]);
    void rotation(float rot_max, float beam_length, const float* dist, const float* angles, int n);
    std::string text() const;
    void save(const std::string& path) const;

private:
    void entry(const std::string& name, const std::string& value);
    std::vector<std::pair<std::string, std::string>> entries_;
};

int flexion_error(const float* rest, const float* pos, int n, const float follow[3], const float target[3],
                  float& error, float p[3]);
void rotation_error(Element el, const int32_t* topo, int ne, const float* rest, const float* pos, int samples,
                    float beam_length, std::vector<float>& dist, std::vector<float>& angles);

}  // namespace oe
