// Writes two small .vtu files with the library's writer (host code): a 2-D lattice of
// 4 x 3 nodes and a 3-D lattice of 3 x 2 x 2 nodes, with fields f(x) = x + 10 y + 100 z
// and g = node index.  The test reads them back with an independent XML/base64 parser.
#include "vtu_writer.hpp"

#include <cstdio>

int
main(int argc, char **argv)
{
  if (argc < 2)
    return 2;
  const std::string dir = argv[1];
  for (int dim = 2; dim <= 3; ++dim)
    {
      const int64_t n[3]     = {dim == 2 ? 4 : 3, dim == 2 ? 3 : 2, dim == 2 ? 1 : 2};
      const double  lower[3] = {1.0, -2.0, 0.5};
      const double  step[3]  = {0.5, 0.25, 2.0};
      std::vector<std::vector<double>> values(2, std::vector<double>(n[0] * n[1] * n[2]));
      for (int64_t k = 0, p = 0; k < n[2]; ++k)
        for (int64_t j = 0; j < n[1]; ++j)
          for (int64_t i = 0; i < n[0]; ++i, ++p)
            {
              const double x = lower[0] + i * step[0], y = lower[1] + j * step[1];
              const double z = dim == 3 ? lower[2] + k * step[2] : 0.0;
              values[0][p]   = x + 10 * y + 100 * z;
              values[1][p]   = double(p);
            }
      ppf::write_vtu(dir + "/lattice_" + std::to_string(dim) + "d.vtu", dim, n, lower, step,
                     {"f", "g"}, values, 0.125 * dim, unsigned(7 * dim));
      // A scalar field and a vector field u (components u_x = x, u_y = y, u_z = z).
      std::vector<std::vector<double>> vec(1 + dim, std::vector<double>(n[0] * n[1] * n[2]));
      for (int64_t k = 0, p = 0; k < n[2]; ++k)
        for (int64_t j = 0; j < n[1]; ++j)
          for (int64_t i = 0; i < n[0]; ++i, ++p)
            {
              vec[0][p] = double(p);
              vec[1][p] = lower[0] + i * step[0];
              vec[2][p] = lower[1] + j * step[1];
              if (dim == 3)
                vec[3][p] = lower[2] + k * step[2];
            }
      ppf::write_vtu(dir + "/vector_" + std::to_string(dim) + "d.vtu", dim, n, lower, step,
                     {"g", "u"}, vec, {1, dim}, 1.0, 3u);
    }
  std::puts("ok");
  return 0;
}
