// The reference's on-disk inputs, read and written on the host:
//  - gmsh 2.2 ASCII tet meshes: $Nodes (count, then "id x y z") and
//    $Elements (count, then lines whose last four fields are the 1-based
//    vertex ids); tetrahedra_obj::load_tetrahedraMesh
//    (StiffGIPC/load_mesh.cpp:481-615), writer TetMesh::export_mesh
//    (MeshProcess/metis_partition/src/mesh.cpp:115-141);
//  - .part files: one part id per vertex, whitespace separated;
//    tetrahedra_obj::load_parts (StiffGIPC/load_mesh.cpp:447-478);
//  - parameterSetting.txt: sixteen "name value" pairs in a fixed order;
//    LoadSettings (StiffGIPC/gl_main.cu:677-726).
// The positions are returned untransformed (the reference applies the
// scene transform while loading; callers do that themselves).
#include "../../include/sgipc.h"

#include <cctype>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

namespace {

struct Msh {
    std::vector<double> x;
    std::vector<uint32_t> tets;
    long long n = -1, T = -1;
};

// section header lines start with '$' (the reference checks substr(1, ...))
bool is_section(const std::string& line, const char* name) {
    return line.size() > 1 && line[0] == '$' && line.compare(1, std::strlen(name), name) == 0 &&
           (line.size() == 1 + std::strlen(name) || std::isspace((unsigned char)line[1 + std::strlen(name)]));
}

int read_msh(const char* path, bool fill, Msh& m, std::string& err) {
    std::ifstream in(path);
    if (!in) {
        err = std::string("cannot open ") + path;
        return SGIPC_ERR_INVALID_ARG;
    }
    std::string line;
    while (std::getline(in, line)) {
        if (!line.empty() && line.back() == '\r') line.pop_back();
        if (is_section(line, "Nodes")) {
            if (!std::getline(in, line)) break;
            m.n = std::atoll(line.c_str());
            if (m.n < 0) { err = "negative node count"; return SGIPC_ERR_INVALID_ARG; }
            if (fill) m.x.resize(3 * (size_t)m.n);
            for (long long i = 0; i < m.n; ++i) {
                if (!std::getline(in, line)) { err = "truncated $Nodes"; return SGIPC_ERR_INVALID_ARG; }
                if (!fill) continue;
                std::istringstream ss(line);
                long long id;
                double a, b, c;
                if (!(ss >> id >> a >> b >> c)) { err = "bad node line: " + line; return SGIPC_ERR_INVALID_ARG; }
                m.x[3 * i] = a; m.x[3 * i + 1] = b; m.x[3 * i + 2] = c;
            }
        } else if (is_section(line, "Elements")) {
            if (!std::getline(in, line)) break;
            m.T = std::atoll(line.c_str());
            if (m.T < 0) { err = "negative element count"; return SGIPC_ERR_INVALID_ARG; }
            if (fill) m.tets.resize(4 * (size_t)m.T);
            for (long long i = 0; i < m.T; ++i) {
                if (!std::getline(in, line)) { err = "truncated $Elements"; return SGIPC_ERR_INVALID_ARG; }
                if (!fill) continue;
                std::istringstream ss(line);
                std::vector<long long> f;
                long long v;
                while (ss >> v) f.push_back(v);
                if (f.size() < 5) { err = "bad element line: " + line; return SGIPC_ERR_INVALID_ARG; }
                for (int k = 0; k < 4; ++k) {
                    long long id = f[f.size() - 4 + k] - 1;  // last four fields, 1-based
                    if (id < 0 || (m.n >= 0 && id >= m.n)) { err = "vertex id out of range: " + line; return SGIPC_ERR_INVALID_ARG; }
                    m.tets[4 * i + k] = (uint32_t)id;
                }
            }
            break;  // the reference stops after the elements
        }
    }
    if (m.n < 0 || m.T < 0) {
        err = "no $Nodes / $Elements section";
        return SGIPC_ERR_INVALID_ARG;
    }
    return SGIPC_OK;
}

thread_local std::string g_io_error;

}  // namespace

extern "C" {

const char* sgipc_io_last_error(void) { return g_io_error.c_str(); }

int sgipc_msh_counts(const char* path, int* n_vertices, int* n_tets) {
    if (!path || !n_vertices || !n_tets) return SGIPC_ERR_INVALID_ARG;
    Msh m;
    int rc = read_msh(path, false, m, g_io_error);
    if (rc) return rc;
    if (m.n > 0x7fffffff || m.T > 0x7fffffff) { g_io_error = "mesh too large"; return SGIPC_ERR_CAPACITY; }
    *n_vertices = (int)m.n;
    *n_tets = (int)m.T;
    return SGIPC_OK;
}

int sgipc_msh_read(const char* path, double* h_x, int n_vertices, uint32_t* h_tets, int n_tets) {
    if (!path || (n_vertices > 0 && !h_x) || (n_tets > 0 && !h_tets)) return SGIPC_ERR_INVALID_ARG;
    Msh m;
    int rc = read_msh(path, true, m, g_io_error);
    if (rc) return rc;
    if (m.n != n_vertices || m.T != n_tets) {
        g_io_error = "sizes differ from the file (call sgipc_msh_counts)";
        return SGIPC_ERR_CAPACITY;
    }
    if (m.n) std::memcpy(h_x, m.x.data(), sizeof(double) * 3 * (size_t)m.n);
    if (m.T) std::memcpy(h_tets, m.tets.data(), sizeof(uint32_t) * 4 * (size_t)m.T);
    return SGIPC_OK;
}

int sgipc_msh_write(const char* path, int n_vertices, const double* h_x, int n_tets, const uint32_t* h_tets) {
    if (!path || n_vertices < 0 || n_tets < 0 || (n_vertices > 0 && !h_x) || (n_tets > 0 && !h_tets))
        return SGIPC_ERR_INVALID_ARG;
    FILE* f = std::fopen(path, "w");
    if (!f) { g_io_error = std::string("cannot write ") + path; return SGIPC_ERR_INVALID_ARG; }
    std::fprintf(f, "$MeshFormat\n2.2 0 8\n$EndMeshFormat\n$Nodes\n%d\n", n_vertices);
    for (int i = 0; i < n_vertices; ++i)
        std::fprintf(f, "%d %.17g %.17g %.17g\n", i + 1, h_x[3 * i], h_x[3 * i + 1], h_x[3 * i + 2]);
    std::fprintf(f, "$EndNodes\n$Elements\n%d\n", n_tets);
    for (int i = 0; i < n_tets; ++i)
        std::fprintf(f, "%d 4 0 %u %u %u %u\n", i + 1, h_tets[4 * i] + 1, h_tets[4 * i + 1] + 1, h_tets[4 * i + 2] + 1,
                     h_tets[4 * i + 3] + 1);
    std::fprintf(f, "$EndElements\n");
    return std::fclose(f) == 0 ? SGIPC_OK : SGIPC_ERR_INVALID_ARG;
}

int sgipc_part_read(const char* path, int* h_part, int capacity, int* n_read) {
    if (!path || !n_read || capacity < 0 || (capacity > 0 && !h_part)) return SGIPC_ERR_INVALID_ARG;
    std::ifstream in(path);
    if (!in) { g_io_error = std::string("cannot open ") + path; return SGIPC_ERR_INVALID_ARG; }
    long long cnt = 0, v;
    while (in >> v) {
        if (v < 0 || v > 0x7fffffff) { g_io_error = "part id out of range"; return SGIPC_ERR_INVALID_ARG; }
        if (cnt < capacity) h_part[cnt] = (int)v;
        ++cnt;
    }
    if (!in.eof()) { g_io_error = "non-numeric entry in the part file"; return SGIPC_ERR_INVALID_ARG; }
    *n_read = (int)cnt;
    return cnt > capacity ? SGIPC_ERR_CAPACITY : SGIPC_OK;
}

int sgipc_parameters_read(const char* path, sgipc_parameters* p) {
    if (!path || !p) return SGIPC_ERR_INVALID_ARG;
    std::ifstream in(path);
    if (!in) { g_io_error = std::string("cannot open ") + path; return SGIPC_ERR_INVALID_ARG; }
    double* dst[16] = {&p->density, &p->poisson_rate, &p->friction_rate, &p->gd_friction_rate,
                       &p->cloth_thickness, &p->cloth_youngs_modulus, &p->bend_youngs_modulus, &p->cloth_density,
                       &p->strain_rate, &p->soft_motion_rate, &p->collision_detection_buff_scale, &p->motion_rate,
                       &p->dt, &p->pcg_threshold, &p->newton_threshold, &p->relative_dhat};
    for (int i = 0; i < 16; ++i) {
        std::string name;
        double v;
        if (!(in >> name >> v)) {
            g_io_error = "parameter " + std::to_string(i + 1) + " of 16 missing";
            return SGIPC_ERR_INVALID_ARG;
        }
        *dst[i] = v;
    }
    return SGIPC_OK;
}

}  // extern "C"
