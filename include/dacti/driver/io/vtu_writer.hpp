// dacti/driver/io/vtu_writer.hpp -- frames as VTK unstructured grids.
//
// vtu_writer(name, out_dir).write_data(integrator, t, frame, field_vars,
// field_var_names) writes <out_dir>/<name>_<frame, 4 digits>.vtu with the
// content of driver/io/vtu_writer.hpp:20-110 of the reference: every active
// cell (for_each_active_cell order) is a hexahedron (VTK type 11) with the
// corners center + size/2 * quadrant(j), j in the order 0,1,3,2,4,5,7,6, and
// a 2D cell 0.01 deep in z; point data p_0.. are the primitives linearly
// reconstructed at each corner (cell.get_p(x)); cell data are the
// parameters k_0.., the metadata acti_level, error, max_dt, cfl and the
// observers' field variables.  The encoding is the VTK XML "binary" one
// (base64 of a UInt64 byte count followed by the raw little-endian values),
// which ParaView and the VTK readers take.
#ifndef DACTI_VTU_WRITER_HPP
#define DACTI_VTU_WRITER_HPP

#include <array>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <vector>

#include "../../core/types.hpp"

namespace dacti::io {

namespace detail {

inline std::string base64(const unsigned char *data, size_t n) {
    static const char tbl[] = "ABCDEFGHIJKLMNOPQRSTUVWXYZabcdefghijklmnopqrstuvwxyz0123456789+/";
    std::string out;
    out.reserve((n + 2) / 3 * 4);
    size_t i = 0;
    for (; i + 2 < n; i += 3) {
        const uint32_t v = (uint32_t) data[i] << 16 | (uint32_t) data[i + 1] << 8 | data[i + 2];
        out += tbl[v >> 18 & 63];
        out += tbl[v >> 12 & 63];
        out += tbl[v >> 6 & 63];
        out += tbl[v & 63];
    }
    if (i < n) {
        uint32_t v = (uint32_t) data[i] << 16;
        if (i + 1 < n) v |= (uint32_t) data[i + 1] << 8;
        out += tbl[v >> 18 & 63];
        out += tbl[v >> 12 & 63];
        out += i + 1 < n ? tbl[v >> 6 & 63] : '=';
        out += '=';
    }
    return out;
}

// one DataArray: a UInt64 byte count, then the values, base64 as one block
template <typename T>
std::string encode(const std::vector<T> &v) {
    const uint64_t nbytes = v.size() * sizeof(T);
    std::vector<unsigned char> buf(sizeof nbytes + nbytes);
    std::memcpy(buf.data(), &nbytes, sizeof nbytes);
    if (nbytes) std::memcpy(buf.data() + sizeof nbytes, v.data(), nbytes);
    return base64(buf.data(), buf.size());
}

template <typename T>
void data_array(std::ofstream &f, const char *type, const std::string &name, int comps, const std::vector<T> &v) {
    f << "<DataArray type=\"" << type << "\" Name=\"" << name << "\" NumberOfComponents=\"" << comps
      << "\" format=\"binary\">\n"
      << encode(v) << "\n</DataArray>\n";
}

}  // namespace detail

class vtu_writer {
    std::string simulation_name, out_dir;

public:
    vtu_writer(const std::string &simulation_name, const std::string &out_dir)
        : simulation_name(simulation_name), out_dir(out_dir) {}

    // path of a frame's file
    std::string frame_path(int frame_index) const {
        char num[32];
        std::snprintf(num, sizeof num, "%04d", frame_index);
        return out_dir + "/" + simulation_name + "_" + num + ".vtu";
    }

    template <typename I>
    void write_data(const I &integrator, scalar_t t, int frame_index,
                    const std::vector<std::vector<scalar_t>> &all_FieldVars,
                    const std::vector<std::string> &all_FieldVar_names) const {
        using Model = typename I::model_type;
        constexpr size_t D = Model::NumDims, NE = Model::NumEqns, NP = Model::NumParm, NM = I::NumMeta;
        (void) t;
        std::vector<double> points;
        std::array<std::vector<double>, NE> data_p;
        std::array<std::vector<double>, NP> data_k;
        std::array<std::vector<double>, NM> data_meta;
        std::vector<int64_t> conn, offsets;
        std::vector<uint8_t> types;
        static constexpr int order[8] = {0, 1, 3, 2, 4, 5, 7, 6};
        int64_t base = 0;
        integrator.for_each_active_cell([&](int, const auto &node) {
            const auto &cell = node.cell;
            types.push_back(11);
            offsets.push_back(base + 8);
            const vec3_t c = cell.get_center().template project<3>();
            const vec3_t s = cell.get_size().template project_with_default<3>(0.01);
            for (int j = 0; j < 8; ++j) {
                const vec3_t v = c + 0.5 * vec3_t::quadrant((size_t) order[j]).cwiseMul(s);
                points.push_back(v[0]);
                points.push_back(v[1]);
                points.push_back(v[2]);
                conn.push_back(base + order[j]);
                const auto p = cell.get_p(v.template project<D>(), cell.t);
                for (size_t k = 0; k < NE; ++k) data_p[k].push_back(p[k]);
            }
            for (size_t k = 0; k < NP; ++k) data_k[k].push_back(cell.get_k()[k]);
            const auto meta = node.get_metadata();
            for (size_t k = 0; k < NM; ++k) data_meta[k].push_back(meta[k]);
            base += 8;
        });
        const std::string path = frame_path(frame_index);
        std::ofstream f(path, std::ios::binary);
        if (!f) throw std::runtime_error("vtu_writer: cannot open " + path);
        f << "<?xml version=\"1.0\"?>\n"
          << "<VTKFile type=\"UnstructuredGrid\" version=\"1.0\" byte_order=\"LittleEndian\" "
             "header_type=\"UInt64\">\n<UnstructuredGrid>\n<Piece NumberOfPoints=\""
          << points.size() / 3 << "\" NumberOfCells=\"" << types.size() << "\">\n<Points>\n";
        detail::data_array(f, "Float64", "Points", 3, points);
        f << "</Points>\n<Cells>\n";
        detail::data_array(f, "Int64", "connectivity", 1, conn);
        detail::data_array(f, "Int64", "offsets", 1, offsets);
        detail::data_array(f, "UInt8", "types", 1, types);
        f << "</Cells>\n<PointData>\n";
        for (size_t k = 0; k < NE; ++k) detail::data_array(f, "Float64", "p_" + std::to_string(k), 1, data_p[k]);
        f << "</PointData>\n<CellData>\n";
        for (size_t k = 0; k < NP; ++k) detail::data_array(f, "Float64", "k_" + std::to_string(k), 1, data_k[k]);
        for (size_t k = 0; k < NM; ++k) detail::data_array(f, "Float64", I::METADATA_NAMES[k], 1, data_meta[k]);
        for (size_t k = 0; k < all_FieldVars.size(); ++k)
            detail::data_array(f, "Float64", all_FieldVar_names[k], 1, all_FieldVars[k]);
        f << "</CellData>\n</Piece>\n</UnstructuredGrid>\n</VTKFile>\n";
        if (!f) throw std::runtime_error("vtu_writer: write failed for " + path);
        std::printf("[info] Wrote file %s\n", path.c_str());
    }
};

}  // namespace dacti::io

#endif
