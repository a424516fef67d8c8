// VTK UnstructuredGrid output of nodal fields, host side, with setters that
// take the fields straight from device memory.  Mirrors VTKWriter
// (include/io/vtk_writer.hpp, src/io/vtk_writer.cpp:47-475): the same file
// names, XML layout, number formatting (15 significant digits), node
// placement (equispaced, and its order-1 placement), cell types and data
// array encodings, so the files match the reference's byte for byte.
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstring>
#include <filesystem>
#include <fstream>
#include <iomanip>
#include <map>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "drf_internal.hpp"

namespace {

struct Field {
    int ncomp = 1;
    std::vector<double> data;
};

// RFC 4648 base64 of a byte string
std::string to_base64(const unsigned char *p, size_t len)
{
    static const char tab[] = "ABCDEFGHIJKLMNOPQRSTUVWXYZabcdefghijklmnopqrstuvwxyz0123456789+/";
    std::string s;
    s.reserve((len + 2) / 3 * 4);
    for (size_t i = 0; i < len; i += 3) {
        const uint32_t b0 = p[i], b1 = i + 1 < len ? p[i + 1] : 0u, b2 = i + 2 < len ? p[i + 2] : 0u;
        const uint32_t w = (b0 << 16) | (b1 << 8) | b2;
        s.push_back(tab[(w >> 18) & 63]);
        s.push_back(tab[(w >> 12) & 63]);
        s.push_back(i + 1 < len ? tab[(w >> 6) & 63] : '=');
        s.push_back(i + 2 < len ? tab[w & 63] : '=');
    }
    return s;
}

} // namespace

struct drf_vtk {
    std::string basename;
    int format = DRF_VTK_VTU, encoding = DRF_VTK_BINARY, order = 1;
    std::vector<double> points;               // x y z per point
    std::vector<std::vector<int64_t>> cells;  // connectivity per cell
    std::vector<int> types;                   // VTK cell type per cell
    std::map<std::string, Field> point_fields, cell_fields; // written in name order
    std::vector<std::pair<double, std::string>> steps;
    int64_t time_idx = 0;

    std::string filename(int64_t idx) const
    {
        std::ostringstream ss;
        ss << basename << "_" << std::setfill('0') << std::setw(6) << idx << ".vtu";
        return ss.str();
    }

    void array(std::ostream &o, const std::string &name, int ncomp,
               const std::vector<double> &d) const
    {
        o << "<DataArray type=\"Float64\" Name=\"" << name << "\" NumberOfComponents=\"" << ncomp
          << "\" format=\"";
        if (encoding == DRF_VTK_ASCII) {
            o << "ascii\">\n";
            for (size_t i = 0; i < d.size(); ++i)
                o << std::setprecision(15) << d[i] << (((i + 1) % ncomp == 0) ? "\n" : " ");
        } else { // binary is written as base64 too: 8-byte size, then the doubles
            o << "binary\">";
            const uint64_t nb = d.size() * sizeof(double);
            std::vector<unsigned char> buf(sizeof(uint64_t) + nb);
            std::memcpy(buf.data(), &nb, sizeof nb);
            if (nb)
                std::memcpy(buf.data() + sizeof nb, d.data(), nb);
            o << to_base64(buf.data(), buf.size());
        }
        o << "</DataArray>\n";
    }

    void write_vtu(const std::string &fn, double time) const
    {
        std::ofstream o(fn);
        if (!o)
            throw std::runtime_error("Failed to open file: " + fn);
        o << "<?xml version=\"1.0\"?>\n"
          << "<VTKFile type=\"UnstructuredGrid\" version=\"1.0\" byte_order=\"LittleEndian\"";
        if (encoding == DRF_VTK_BASE64) // declared, though the data is not compressed
            o << " compressor=\"vtkZLibDataCompressor\"";
        o << ">\n<UnstructuredGrid>\n<FieldData>\n"
          << "<DataArray type=\"Float64\" Name=\"TIME\" NumberOfTuples=\"1\" format=\"ascii\">\n"
          << std::setprecision(15) << time << "\n</DataArray>\n</FieldData>\n";
        o << "<Piece NumberOfPoints=\"" << points.size() / 3 << "\" NumberOfCells=\""
          << cells.size() << "\">\n<Points>\n";
        array(o, "Points", 3, points);
        o << "</Points>\n<Cells>\n<DataArray type=\"Int64\" Name=\"connectivity\" format=\"ascii\">\n";
        for (const auto &c : cells) {
            for (int64_t i : c)
                o << i << " ";
            o << "\n";
        }
        o << "</DataArray>\n<DataArray type=\"Int64\" Name=\"offsets\" format=\"ascii\">\n";
        int64_t off = 0;
        for (const auto &c : cells) {
            off += (int64_t)c.size();
            o << off << " ";
        }
        o << "\n</DataArray>\n<DataArray type=\"UInt8\" Name=\"types\" format=\"ascii\">\n";
        for (int t : types)
            o << t << " ";
        o << "\n</DataArray>\n</Cells>\n";
        if (!point_fields.empty()) {
            o << "<PointData>\n";
            for (const auto &[name, f] : point_fields)
                array(o, name, f.ncomp, f.data);
            o << "</PointData>\n";
        }
        if (!cell_fields.empty()) {
            o << "<CellData>\n";
            for (const auto &[name, f] : cell_fields)
                array(o, name, f.ncomp, f.data);
            o << "</CellData>\n";
        }
        o << "</Piece>\n</UnstructuredGrid>\n</VTKFile>\n";
    }
};

using namespace drf;

#define DRF_REQUIRE(cond, code, msg)                                                         \
    do {                                                                                     \
        if (!(cond))                                                                         \
            return ::drf::fail((code), (msg));                                               \
    } while (0)

extern "C" {

drf_status drf_vtk_create(const char *basename, int format, int encoding, drf_vtk **out)
{
    DRF_REQUIRE(out && basename, DRF_ERR_ARG, "drf_vtk_create: null argument");
    *out = nullptr;
    DRF_REQUIRE(format == DRF_VTK_VTU || format == DRF_VTK_PVTU, DRF_ERR_ARG,
                "drf_vtk_create: unknown format");
    DRF_REQUIRE(encoding >= DRF_VTK_ASCII && encoding <= DRF_VTK_BASE64, DRF_ERR_ARG,
                "drf_vtk_create: unknown encoding");
    try {
        const std::filesystem::path dir = std::filesystem::path(basename).parent_path();
        if (!dir.empty())
            std::filesystem::create_directories(dir);
        drf_vtk *w = new drf_vtk;
        w->basename = basename;
        w->format = format;
        w->encoding = encoding;
        *out = w;
    } catch (const std::exception &e) {
        return fail(DRF_ERR_ARG, std::string("drf_vtk_create: ") + e.what());
    }
    return DRF_OK;
}

void drf_vtk_destroy(drf_vtk *w) { delete w; }

drf_status drf_vtk_set_polynomial_order(drf_vtk *w, int order)
{
    DRF_REQUIRE(w && order >= 1, DRF_ERR_ARG, "drf_vtk_set_polynomial_order: bad argument");
    w->order = order;
    return DRF_OK;
}

drf_status drf_vtk_set_mesh(drf_vtk *w, int64_t n_elem, const double *bounds)
{
    DRF_REQUIRE(w && n_elem >= 0 && (n_elem == 0 || bounds), DRF_ERR_ARG,
                "drf_vtk_set_mesh: bad argument");
    const int p = w->order, nn = p + 1, npe = nn * nn * nn;
    w->points.clear();
    w->cells.clear();
    w->types.clear();
    w->points.reserve((size_t)n_elem * npe * 3);
    int64_t next = 0;
    for (int64_t e = 0; e < n_elem; ++e) {
        const double *b = bounds + 6 * e;
        const double mn[3] = {b[0], b[2], b[4]};
        const double sz[3] = {b[1] - b[0], b[3] - b[2], b[5] - b[4]};
        std::vector<int64_t> conn;
        conn.reserve(npe);
        // equispaced nodes; at order 1 the reference coordinate is the index
        // itself (0 or 1), which places them at the centre and the max face
        auto ref = [p](int k) { return (p == 1) ? (double)k : -1.0 + 2.0 * k / p; };
        for (int k = 0; k < nn; ++k)
            for (int j = 0; j < nn; ++j)
                for (int i = 0; i < nn; ++i) {
                    const double r[3] = {ref(i), ref(j), ref(k)};
                    for (int a = 0; a < 3; ++a)
                        w->points.push_back(mn[a] + 0.5 * (r[a] + 1.0) * sz[a]);
                    conn.push_back(next++);
                }
        if (p == 1) { // linear hexahedron in VTK corner order
            static const int vtk_corner[8] = {0, 1, 3, 2, 4, 5, 7, 6};
            std::vector<int64_t> c(8);
            for (int i = 0; i < 8; ++i)
                c[i] = conn[0] + vtk_corner[i];
            w->cells.push_back(c);
            w->types.push_back(12);
        } else { // Lagrange hexahedron, nodes in the DG order
            w->cells.push_back(conn);
            w->types.push_back(72);
        }
    }
    return DRF_OK;
}

drf_status drf_vtk_add_field(drf_vtk *w, const char *name, int ncomp, int on_cells)
{
    DRF_REQUIRE(w && name && ncomp >= 1, DRF_ERR_ARG, "drf_vtk_add_field: bad argument");
    Field f;
    f.ncomp = ncomp;
    (on_cells ? w->cell_fields : w->point_fields)[name] = f;
    return DRF_OK;
}

drf_status drf_vtk_set_field(drf_vtk *w, const char *name, int on_cells, int64_t n,
                             const double *data)
{
    DRF_REQUIRE(w && name && n >= 0 && (n == 0 || data), DRF_ERR_ARG,
                "drf_vtk_set_field: bad argument");
    auto &m = on_cells ? w->cell_fields : w->point_fields;
    auto it = m.find(name);
    DRF_REQUIRE(it != m.end(), DRF_ERR_ARG,
                std::string(on_cells ? "Unknown cell field: " : "Unknown point field: ") + name);
    it->second.data.assign(data, data + n);
    return DRF_OK;
}

drf_status drf_vtk_set_point_field_dev(drf_vtk *w, const char *name, int64_t n_elem,
                                       const void *field, int64_t elem_stride, int64_t ndof,
                                       void *stream)
{
    DRF_REQUIRE(w && name && n_elem >= 0 && ndof >= 1 && elem_stride >= ndof &&
                    (n_elem == 0 || field),
                DRF_ERR_ARG, "drf_vtk_set_point_field_dev: bad argument");
    auto it = w->point_fields.find(name);
    DRF_REQUIRE(it != w->point_fields.end(), DRF_ERR_ARG,
                std::string("Unknown point field: ") + name);
    std::vector<double> host((size_t)(n_elem * ndof));
    if (n_elem) { // one strided copy: ndof values of each element
        DRF_HIP_CHECK(hipMemcpy2DAsync(host.data(), sizeof(double) * ndof, field,
                                       sizeof(double) * elem_stride, sizeof(double) * ndof,
                                       n_elem, hipMemcpyDeviceToHost, (hipStream_t)stream));
        DRF_HIP_CHECK(hipStreamSynchronize((hipStream_t)stream));
    }
    it->second.data.swap(host);
    return DRF_OK;
}

drf_status drf_vtk_filename(const drf_vtk *w, int64_t time_idx, char *buf, int64_t buf_len)
{
    DRF_REQUIRE(w && buf && buf_len > 0 && time_idx >= 0, DRF_ERR_ARG,
                "drf_vtk_filename: bad argument");
    const std::string s = w->filename(time_idx);
    DRF_REQUIRE((int64_t)s.size() < buf_len, DRF_ERR_ARG, "drf_vtk_filename: buffer too small");
    std::memcpy(buf, s.c_str(), s.size() + 1);
    return DRF_OK;
}

drf_status drf_vtk_write(drf_vtk *w, int64_t time_idx, double time)
{
    DRF_REQUIRE(w && time_idx >= 0, DRF_ERR_ARG, "drf_vtk_write: bad argument");
    const std::string fn = w->filename(time_idx);
    try {
        w->write_vtu(fn, time); // PVTU without MPI writes the one piece
    } catch (const std::exception &e) {
        return fail(DRF_ERR_ARG, e.what());
    }
    w->steps.push_back({time, fn});
    return DRF_OK;
}

drf_status drf_vtk_write_timestep(drf_vtk *w, double time)
{
    DRF_REQUIRE(w, DRF_ERR_ARG, "drf_vtk_write_timestep: null writer");
    return drf_vtk_write(w, w->time_idx++, time);
}

drf_status drf_vtk_finalize(drf_vtk *w)
{
    DRF_REQUIRE(w, DRF_ERR_ARG, "drf_vtk_finalize: null writer");
    const std::string fn = w->basename + ".pvd";
    std::ofstream o(fn);
    DRF_REQUIRE((bool)o, DRF_ERR_ARG, "Failed to open file: " + fn);
    o << "<?xml version=\"1.0\"?>\n<VTKFile type=\"Collection\" version=\"0.1\">\n<Collection>\n";
    for (const auto &[t, f] : w->steps)
        o << "<DataSet timestep=\"" << std::setprecision(15) << t << "\" file=\"" << f << "\"/>\n";
    o << "</Collection>\n</VTKFile>\n";
    return DRF_OK;
}

} // extern "C"
