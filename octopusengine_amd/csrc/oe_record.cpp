// JSON experiment record and the two error measures the reference records
// with it (src/Script/Record/DataRecorder.cpp, src/Script/Record/Recorder.cpp,
// include/Script/Record/Recorder.h).  Values go through a std::ostream with
// its default formatting, as the reference writes them, so that the same
// values give the same bytes.
#include <algorithm>
#include <cmath>
#include <fstream>
#include <sstream>
#include <stdexcept>

#include "oe_host.hpp"

namespace oe {

namespace {
// reference element_name (src/Mesh/Elements.cpp:70-84)
const char* element_name(int e) {
    static const char* names[] = {"Line", "Triangle", "Quad", "Tetra", "Pyramid", "Prism", "Hexa",
                                  "Tetra10", "Tetra20", "Hexa27"};
    return e >= 0 && e < 10 ? names[e] : "";
}
// reference get_material_name (src/Dynamic/FEM/ContinuousMaterial.cpp:4-12)
const char* material_name(int m) {
    static const char* names[] = {"Hooke", "SaintVenant", "Neo-Hooke", "Stable Neo-Hooke"};
    return m >= 0 && m < 4 ? names[m] : "";
}

void float_array(std::ostream& o, const float* v, int n) {
    o << "[";
    for (int i = 0; i < n; ++i) {
        o << v[i];
        if (i < n - 1) o << ",";
    }
    o << "]";
}
}  // namespace

void JsonRecord::entry(const std::string& name, const std::string& value) {
    entries_.emplace_back(name, value);
}

// TimeRecorder::add_data_json (include/Script/Record/Recorder.h:37-47)
void JsonRecord::time(int frame, float time, float dt, float fixed_time, float fixed_dt) {
    std::ostringstream o;
    o << "{" << "\"frame\" : " << frame << "," << "\"time\" : " << time << "," << "\"dt\" : " << dt << ","
      << "\"fixed_Time\" : " << fixed_time << "," << "\"fixed_dt\" : " << fixed_dt << "" << "}";
    entry("Time", o.str());
}

// MeshRecorder::add_data_json (src/Script/Record/Recorder.cpp:13-24): element
// counts in element-id order, then the vertex count
void JsonRecord::mesh(const int32_t counts[10], int n_vertices) {
    std::ostringstream o;
    o << "{";
    for (int e = 0; e < 10; ++e)
        if (counts[e] > 0) o << "\"" << element_name(e) << "\" : " << counts[e] << ",";
    o << "\"vertices\" : " << n_vertices;
    o << "}";
    entry("mesh", o.str());
}

// FEM_Dynamic_Recorder / XPBD_FEM_Dynamic_Recorder::add_data_json
// (src/Script/Record/Recorder.cpp:27-49); both are named "XPBD_FEM_Dynamic"
// there (include/Script/Record/Recorder.h:82-84,105-107), and so here.
void JsonRecord::material(bool xpbd, int material, float poisson, float young, int iterations, int sub_iterations,
                          float density) {
    std::ostringstream o;
    o << "{" << "\"material\" : \"" << material_name(material) << "\"," << "\"poisson\" : " << poisson << ","
      << "\"young\" : " << young << ",";
    if (xpbd)
        o << "\"iteration\" : " << iterations << "," << "\"sub_itetrations\" : " << sub_iterations << ",";
    else
        o << "\"sub_step\" : " << sub_iterations << ",";
    o << "\"density\" : " << density << "}";
    entry("XPBD_FEM_Dynamic", o.str());
}

// the VTK recorders' entries (src/Script/Record/Recorder.cpp:95-97,113-115,173-175)
void JsonRecord::string(const std::string& name, const std::string& value) { entry(name, "\"" + value + "\""); }

// FEM_Flexion_error_recorder::add_data_json (src/Script/Record/Recorder.cpp:210-219)
void JsonRecord::flexion(float error, const float p[3], const float target[3]) {
    std::ostringstream o;
    o << "{" << "\"error\" : " << error << "," << "\"p\" : [" << p[0] << ", " << p[1] << ", " << p[2] << "],"
      << "\"target\" : [" << target[0] << ", " << target[1] << ", " << target[2] << "]" << "}";
    entry("flexion_error", o.str());
}

// FEM_Torsion_error_recorder::add_data_json (src/Script/Record/Recorder.cpp:317-341)
void JsonRecord::rotation(float rot_max, float beam_length, const float* dist, const float* angles, int n) {
    std::ostringstream o;
    o << "{" << "\"rot_max\" : " << rot_max << "," << "\"beam_lenght\" : " << beam_length << "," << "\"distance\" : ";
    float_array(o, dist, n);
    o << ",";
    o << "\"angles\" : ";
    float_array(o, angles, n);
    o << "}";
    entry("rotation_error", o.str());
}

// DataRecorder::save (src/Script/Record/DataRecorder.cpp:16-38)
std::string JsonRecord::text() const {
    std::ostringstream o;
    o << "{";
    for (size_t i = 0; i < entries_.size(); ++i) {
        o << "\"" << entries_[i].first << "\" : " << entries_[i].second;
        if (i + 1 < entries_.size()) o << ",";
    }
    o << "}";
    return o.str();
}

void JsonRecord::save(const std::string& path) const {
    std::ofstream f(path, std::ios::binary);
    if (!f) throw std::runtime_error("cannot open " + path);
    f << text();
    if (!f) throw std::runtime_error("write failed: " + path);
}

// FEM_Flexion_error_recorder::init / print (src/Script/Record/Recorder.cpp:190-208,
// include/Script/Record/Recorder.h:256-259): the first vertex whose rest
// position is within 1e-3 of `follow` (squared distance < 1e-6) is followed;
// error = |target - p|^2.  Returns the vertex index, or -1.
int flexion_error(const float* rest, const float* pos, int n, const float follow[3], const float target[3],
                  float& error, float p[3]) {
    for (int i = 0; i < n; ++i) {
        const float dx = rest[3 * i] - follow[0], dy = rest[3 * i + 1] - follow[1], dz = rest[3 * i + 2] - follow[2];
        if (dx * dx + dy * dy + dz * dz >= 1e-6f) continue;
        for (int c = 0; c < 3; ++c) p[c] = pos[3 * i + c];
        const float ex = target[0] - p[0], ey = target[1] - p[1], ez = target[2] - p[2];
        error = ex * ex + ey * ey + ez * ez;
        return i;
    }
    return -1;
}

// FEM_Torsion_error_recorder::init + compute_errors (src/Script/Record/Recorder.cpp:221-315).
// Samples: on every element with two corner nodes on the beam edge y = z = 0
// (rest positions), samples_per_edge points evenly spaced between those two
// corners' reference coordinates.  At each sample the rest and current
// positions are interpolated with the element's shape functions; the angle is
// that between the two (y, z) directions from the beam axis (0.5, 0.5), in
// degrees, and the distance is the current x over the beam length.  The pairs
// are sorted by distance and runs closer than 1e-4 merged (mean angle).  As in
// the reference, the "current" position interpolated is the one reported as
// distance, and the atan2 arguments keep the reference's order.
void rotation_error(Element el, const int32_t* topo, int ne, const float* rest, const float* pos, int samples,
                    float beam_length, std::vector<float>& dist, std::vector<float>& angles) {
    if (samples < 2) throw std::invalid_argument("rotation error needs at least 2 samples per edge");
    const ShapeTable& st = shape_table(el);
    const int S = st.S;
    const int lin = el == Tetra10 || el == Tetra20 ? 4 : el == Hexa27 ? 8 : S;
    dist.clear();
    angles.clear();
    std::vector<double> N(S), dN(static_cast<size_t>(S) * 3);
    for (int e = 0; e < ne; ++e) {
        const int32_t* t = topo + static_cast<size_t>(e) * S;
        int edge[2], k = 0;
        for (int j = 0; j < lin && k < 2; ++j) {
            const float* p = rest + 3 * static_cast<size_t>(t[j]);
            if (p[1] < 1e-6f && p[2] < 1e-6f) edge[k++] = j;
        }
        if (k < 2) continue;
        for (int s = 0; s < samples; ++s) {
            const float tt = static_cast<float>(s) / static_cast<float>(samples - 1);
            double r[3];
            for (int c = 0; c < 3; ++c) {
                const float a = static_cast<float>(st.nodes[3 * edge[0] + c]);
                const float b = static_cast<float>(st.nodes[3 * edge[1] + c]);
                r[c] = tt * (b - a) + a;
            }
            st.eval(r, N.data(), dN.data());
            float p[3] = {0.f, 0.f, 0.f}, p0[3] = {0.f, 0.f, 0.f};
            for (int l = 0; l < S; ++l) {
                const float w = static_cast<float>(N[l]);
                for (int c = 0; c < 3; ++c) {
                    p[c] += w * rest[3 * static_cast<size_t>(t[l]) + c];
                    p0[c] += w * pos[3 * static_cast<size_t>(t[l]) + c];
                }
            }
            // v * 1/sqrt(v.v), glm::normalize's arithmetic
            auto unit = [](float y, float z, float out[2]) {
                const float inv = 1.f / std::sqrt(y * y + z * z);
                out[0] = y * inv;
                out[1] = z * inv;
            };
            float di[2], dc[2];
            unit(p0[1] - 0.5f, p0[2] - 0.5f, di);
            unit(p[1] - 0.5f, p[2] - 0.5f, dc);
            float angle = std::fabs(std::atan2(dc[0] * di[1] - dc[1] * di[0], dc[0] * di[0] + dc[1] * di[1]));
            angle = angle * 57.295779513082320876798154814105f;   // glm::degrees
            dist.push_back(p0[0] / beam_length);
            angles.push_back(angle);
        }
    }
    if (dist.empty()) return;
    // ascending distance (the reference's exchange sort keeps equal keys in order)
    std::vector<size_t> order(dist.size());
    for (size_t i = 0; i < order.size(); ++i) order[i] = i;
    std::stable_sort(order.begin(), order.end(), [&](size_t a, size_t b) { return dist[a] < dist[b]; });
    std::vector<float> d(order.size()), a(order.size());
    for (size_t i = 0; i < order.size(); ++i) {
        d[i] = dist[order[i]];
        a[i] = angles[order[i]];
    }
    dist.clear();
    angles.clear();
    size_t i = 0;
    while (i < d.size()) {
        size_t j = 1;
        float sum = a[i];
        while (i + j < d.size() && std::fabs(d[i] - d[i + j]) < 1e-4) {
            sum += a[i + j];
            ++j;
        }
        angles.push_back(sum / static_cast<float>(j));
        dist.push_back(d[i]);
        i += j;
    }
}

}  // namespace oe
