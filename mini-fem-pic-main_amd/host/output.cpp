// mini-pic's VTK writers (declared in mp_output.h): the layout of src/meshes.cpp:360-449 and
// src/particles.cpp:20-45, numbers in the C++ stream's default format (six significant
// digits, shortest of fixed and scientific), as the reference's ofstream writes them.
#include "mp_output.h"

#include <fstream>
#include <string>

namespace {

thread_local std::string g_out_err;

int fail(const std::string& m) {
    g_out_err = m;
    return MP_ERR_ARG;
}

}  // namespace

extern "C" {

const char* mp_output_last_error(void) { return g_out_err.c_str(); }

int mp_write_mesh_vtu(const mp_mesh* m, const char* path, const double* phi, const double* ef,
                      const double* ion_den) {
    if (!m || !path || !phi || !ef || !ion_den) return fail("mp_write_mesh_vtu: null argument");
    std::ofstream out(path);
    if (!out.is_open()) return fail(std::string("Failed to open file ") + path);
    const int32_t nn = mp_mesh_num_nodes(m), ne = mp_mesh_num_tets(m);
    const double* pos = mp_mesh_node_pos(m);
    const int32_t* type = mp_mesh_node_type(m);
    const mp_tetra* tets = mp_mesh_tets(m);
    out << "<?xml version=\"1.0\"?>\n";
    out << "<VTKFile type=\"UnstructuredGrid\" version=\"0.1\" byte_order=\"LittleEndian\">\n";
    out << "<UnstructuredGrid>\n";
    out << "<Piece NumberOfPoints=\"" << nn << "\" NumberOfVerts=\"0\" NumberOfLines=\"0\" ";
    out << "NumberOfStrips=\"0\" NumberOfCells=\"" << ne << "\">\n";
    out << "<Points>\n";
    out << "<DataArray type=\"Float32\" NumberOfComponents=\"3\" format=\"ascii\">\n";
    for (int32_t n = 0; n < nn; n++) out << pos[3 * n] << " " << pos[3 * n + 1] << " " << pos[3 * n + 2] << "\n";
    out << "</DataArray>\n";
    out << "</Points>\n";
    out << "<Cells>\n";
    out << "<DataArray type=\"Int32\" Name=\"connectivity\" format=\"ascii\">\n";
    for (int32_t e = 0; e < ne; e++)
        out << tets[e].con[0] << " " << tets[e].con[1] << " " << tets[e].con[2] << " " << tets[e].con[3] << "\n";
    out << "</DataArray>\n";
    out << "<DataArray type=\"Int32\" Name=\"offsets\" format=\"ascii\">\n";
    for (int64_t e = 0; e < ne; e++) out << (e + 1) * 4 << " ";
    out << "\n";
    out << "</DataArray>\n";
    out << "<DataArray type=\"UInt8\" Name=\"types\" format=\"ascii\">\n";
    for (int32_t e = 0; e < ne; e++) out << "10 ";
    out << "\n";
    out << "</DataArray>\n";
    out << "</Cells>\n";
    out << "<PointData Scalars=\"phi\">\n";
    out << "<DataArray type=\"Int32\" Name=\"node_index\" format=\"ascii\">\n";
    for (int32_t n = 0; n < nn; n++) out << n << " ";
    out << "\n";
    out << "</DataArray>\n";
    out << "<DataArray type=\"Int32\" Name=\"node_type\" format=\"ascii\">\n";
    for (int32_t n = 0; n < nn; n++) out << type[n] << " ";
    out << "\n";
    out << "</DataArray>\n";
    out << "<DataArray type=\"Float32\" Name=\"phi\" format=\"ascii\">\n";
    for (int32_t n = 0; n < nn; n++) out << phi[n] << " ";
    out << "\n";
    out << "</DataArray>\n";
    out << "<DataArray type=\"Float32\" Name=\"ion_den\" format=\"ascii\">\n";
    for (int32_t n = 0; n < nn; n++) out << ion_den[n] << " ";
    out << "\n";
    out << "</DataArray>\n";
    out << "</PointData>\n";
    out << "<CellData Vectors=\"ef\">\n";
    out << "<DataArray type=\"Float32\" NumberOfComponents=\"3\" Name=\"ef\" format=\"ascii\">\n";
    for (int32_t e = 0; e < ne; e++) out << ef[3 * e] << " " << ef[3 * e + 1] << " " << ef[3 * e + 2] << " ";
    out << "\n";
    out << "</DataArray>\n";
    out << "<DataArray type=\"Float32\" Name=\"cell_volume\" format=\"ascii\">\n";
    for (int32_t e = 0; e < ne; e++) out << tets[e].volume << " ";
    out << "\n";
    out << "</DataArray>\n";
    out << "</CellData>\n";
    out << "</Piece>\n";
    out << "</UnstructuredGrid>\n";
    out << "</VTKFile>\n";
    out.close();
    return out ? MP_OK : fail(std::string("write failed: ") + path);
}

int mp_write_particles_vtp(const char* path, const double* pos, int64_t n, int64_t stride_bytes) {
    if (!path || n < 0 || (n > 0 && !pos) || stride_bytes < 24) return fail("mp_write_particles_vtp: bad argument");
    std::ofstream out(path);
    if (!out.is_open()) return fail(std::string("Failed to open output file ") + path);
    out << "<?xml version=\"1.0\"?>\n";
    out << "<VTKFile type=\"PolyData\" version=\"0.1\" byte_order=\"LittleEndian\">\n";
    out << "<PolyData>\n";
    out << "<Piece NumberOfPoints=\"" << n << "\" NumberOfVerts=\"0\" NumberOfLines=\"0\" ";
    out << "NumberOfStrips=\"0\" NumberOfCells=\"0\">\n";
    out << "<Points>\n";
    out << "<DataArray type=\"Float32\" NumberOfComponents=\"3\" format=\"ascii\">\n";
    const char* b = (const char*)pos;
    for (int64_t k = 0; k < n; k++) {
        const double* p = (const double*)(b + k * stride_bytes);
        out << p[0] << " " << p[1] << " " << p[2] << "\n";
    }
    out << "</DataArray>\n";
    out << "</Points>\n";
    out << "</Piece>\n";
    out << "</PolyData>\n";
    out << "</VTKFile>\n";
    out.close();
    return out ? MP_OK : fail(std::string("write failed: ") + path);
}

}  // extern "C"
