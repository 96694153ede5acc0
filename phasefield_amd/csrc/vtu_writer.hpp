// VTK XML UnstructuredGrid (.vtu) output of the field states on the uniform lattice,
// the file the reference writes at every output increment (SolutionOutput,
// include/prismspf/core/solution_output.h:53-126, VTU branch; DataOut with patches
// subdivided `degree` times, so the output points are the element's nodes and the output
// cells are the linear sub-cells between neighbouring nodes).
//
// Points: deal.II's full node set of the mesh (periodic images included), x fastest.
// Cells: VTK_QUAD (2-D) or VTK_HEXAHEDRON (3-D) between lattice neighbours.  Point data:
// one Float64 array per field, named after the field.  Field data: TIME and CYCLE.
// Arrays use the XML "binary" encoding: base64 of a UInt64 byte count followed by the
// little-endian data.  Host code only.
#pragma once

#include <cstdint>
#include <cstring>
#include <fstream>
#include <stdexcept>
#include <string>
#include <vector>

namespace ppf
{
  namespace vtu
  {
    inline void
    base64_append(std::string &out, const unsigned char *p, std::size_t n)
    {
      static const char table[] =
        "ABCDEFGHIJKLMNOPQRSTUVWXYZabcdefghijklmnopqrstuvwxyz0123456789+/";
      std::size_t i = 0;
      for (; i + 2 < n; i += 3)
        {
          const unsigned v = (unsigned(p[i]) << 16) | (unsigned(p[i + 1]) << 8) | p[i + 2];
          out += table[(v >> 18) & 63];
          out += table[(v >> 12) & 63];
          out += table[(v >> 6) & 63];
          out += table[v & 63];
        }
      if (i < n)
        {
          const unsigned v = (unsigned(p[i]) << 16) | (i + 1 < n ? unsigned(p[i + 1]) << 8 : 0u);
          out += table[(v >> 18) & 63];
          out += table[(v >> 12) & 63];
          out += i + 1 < n ? table[(v >> 6) & 63] : '=';
          out += '=';
        }
    }

    /// One binary-encoded DataArray element.
    template <typename T>
    void
    data_array(std::ofstream &os, const char *type, const std::string &name, int components,
               const std::vector<T> &v)
    {
      std::vector<unsigned char> buf(8 + v.size() * sizeof(T));
      const uint64_t             bytes = v.size() * sizeof(T);
      std::memcpy(buf.data(), &bytes, 8);
      if (!v.empty())
        std::memcpy(buf.data() + 8, v.data(), bytes);
      std::string enc;
      enc.reserve(buf.size() / 3 * 4 + 8);
      base64_append(enc, buf.data(), buf.size());
      os << "        <DataArray type=\"" << type << "\" Name=\"" << name
         << "\" NumberOfComponents=\"" << components << "\" format=\"binary\">\n"
         << enc << "\n        </DataArray>\n";
    }
  } // namespace vtu

  /// Writes `path`.  n[d]: nodes per direction of the full node set (n[2] = 1 in 2-D);
  /// lower/step: coordinates of node 0 and the node spacing; values[f]: stored field f
  /// over the full node set.  groups[g] consecutive stored fields form output array g,
  /// named names[g]: a scalar field (1) is written as one component; the components of a
  /// vector field (dim) are written as one 3-component array, the z component zero in
  /// 2-D, as deal.II's DataOut writes a vector-valued field
  /// (component_is_part_of_vector, reference solution_output.h:80-84).
  inline void
  write_vtu(const std::string &path, int dim, const int64_t (&n)[3], const double (&lower)[3],
            const double (&step)[3], const std::vector<std::string> &names,
            const std::vector<std::vector<double>> &values, const std::vector<int> &groups,
            double time, unsigned int cycle)
  {
    const int64_t n_points = n[0] * n[1] * n[2];
    for (const auto &v : values)
      if (int64_t(v.size()) != n_points)
        throw std::invalid_argument("write_vtu: field size does not match the node set");
    std::size_t n_grouped = 0;
    for (const int g : groups)
      {
        if (g != 1 && g != dim)
          throw std::invalid_argument("write_vtu: an output array has 1 or dim components");
        n_grouped += std::size_t(g);
      }
    if (names.size() != groups.size() || n_grouped != values.size())
      throw std::invalid_argument("write_vtu: names, groups and fields do not match");
    const int64_t c0 = n[0] - 1, c1 = n[1] - 1, c2 = dim == 3 ? n[2] - 1 : 1;
    const int64_t n_cells = c0 * c1 * c2;

    std::vector<double> points(3 * n_points);
    for (int64_t k = 0, p = 0; k < n[2]; ++k)
      for (int64_t j = 0; j < n[1]; ++j)
        for (int64_t i = 0; i < n[0]; ++i, ++p)
          {
            points[3 * p]     = lower[0] + double(i) * step[0];
            points[3 * p + 1] = lower[1] + double(j) * step[1];
            points[3 * p + 2] = dim == 3 ? lower[2] + double(k) * step[2] : 0.0;
          }
    const int            corners = dim == 3 ? 8 : 4;
    std::vector<int64_t> conn(corners * n_cells), offsets(n_cells);
    std::vector<uint8_t> types(n_cells, uint8_t(dim == 3 ? 12 : 9));
    auto idx = [&](int64_t i, int64_t j, int64_t k) { return i + n[0] * (j + n[1] * k); };
    int64_t c = 0;
    for (int64_t k = 0; k < c2; ++k)
      for (int64_t j = 0; j < c1; ++j)
        for (int64_t i = 0; i < c0; ++i, ++c)
          {
            int64_t *q = conn.data() + corners * c;
            // VTK corner order: counter-clockwise bottom face, then the top face.
            q[0] = idx(i, j, k);
            q[1] = idx(i + 1, j, k);
            q[2] = idx(i + 1, j + 1, k);
            q[3] = idx(i, j + 1, k);
            if (dim == 3)
              {
                q[4] = idx(i, j, k + 1);
                q[5] = idx(i + 1, j, k + 1);
                q[6] = idx(i + 1, j + 1, k + 1);
                q[7] = idx(i, j + 1, k + 1);
              }
            offsets[c] = corners * (c + 1);
          }

    std::ofstream os(path, std::ios::binary);
    if (!os)
      throw std::runtime_error("cannot open output file " + path);
    os << "<?xml version=\"1.0\"?>\n"
       << "<VTKFile type=\"UnstructuredGrid\" version=\"1.0\" byte_order=\"LittleEndian\" "
          "header_type=\"UInt64\">\n"
       << "  <UnstructuredGrid>\n"
       << "    <FieldData>\n";
    vtu::data_array(os, "Float64", "TIME", 1, std::vector<double> {time});
    vtu::data_array(os, "Int32", "CYCLE", 1, std::vector<int32_t> {int32_t(cycle)});
    os << "    </FieldData>\n"
       << "    <Piece NumberOfPoints=\"" << n_points << "\" NumberOfCells=\"" << n_cells
       << "\">\n"
       << "      <Points>\n";
    vtu::data_array(os, "Float64", "Points", 3, points);
    os << "      </Points>\n      <Cells>\n";
    vtu::data_array(os, "Int64", "connectivity", 1, conn);
    vtu::data_array(os, "Int64", "offsets", 1, offsets);
    vtu::data_array(os, "UInt8", "types", 1, types);
    os << "      </Cells>\n      <PointData>\n";
    for (std::size_t g = 0, f = 0; g < groups.size(); f += std::size_t(groups[g]), ++g)
      if (groups[g] == 1)
        vtu::data_array(os, "Float64", names[g], 1, values[f]);
      else
        {
          std::vector<double> vec(3 * n_points, 0.0);
          for (int c = 0; c < groups[g]; ++c)
            for (int64_t p = 0; p < n_points; ++p)
              vec[3 * p + c] = values[f + std::size_t(c)][p];
          vtu::data_array(os, "Float64", names[g], 3, vec);
        }
    os << "      </PointData>\n    </Piece>\n  </UnstructuredGrid>\n</VTKFile>\n";
    if (!os)
      throw std::runtime_error("writing output file " + path + " failed");
  }

  /// Every stored field written as a scalar array.
  inline void
  write_vtu(const std::string &path, int dim, const int64_t (&n)[3], const double (&lower)[3],
            const double (&step)[3], const std::vector<std::string> &names,
            const std::vector<std::vector<double>> &values, double time, unsigned int cycle)
  {
    write_vtu(path, dim, n, lower, step, names, values, std::vector<int>(values.size(), 1), time,
              cycle);
  }
} // namespace ppf
