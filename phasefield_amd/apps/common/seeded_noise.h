// Counter-based pseudo-random numbers for initial conditions.
//
// The reference's initial conditions draw from a std::mt19937 in deal.II's DoF
// traversal order (applications/cahn_hilliard/explicit/custom_pde.h:51-55), which no
// other implementation can reproduce.  These applications instead key each draw by
// (seed, stream, lattice node), so the value at a node is independent of traversal
// order and identical in every build.  Generator: splitmix64 (Steele, Lea, Flood,
// "Fast splittable pseudorandom number generators", OOPSLA 2014); uniforms take the top
// 53 bits; normals use the Box-Muller transform.
#pragma once

#include <prismspf/config.h>

#include <cmath>
#include <cstdint>

namespace ppf_noise
{
  PPF_HD inline uint64_t
  splitmix64(uint64_t x)
  {
    x += 0x9e3779b97f4a7c15ull;
    x = (x ^ (x >> 30)) * 0xbf58476d1ce4e5b9ull;
    x = (x ^ (x >> 27)) * 0x94d049bb133111ebull;
    return x ^ (x >> 31);
  }

  /// Uniform in [0, 1) for draw `counter` of stream `stream`.
  PPF_HD inline double
  uniform01(uint64_t seed, uint64_t stream, uint64_t counter)
  {
    const uint64_t h = splitmix64(splitmix64(seed ^ (stream * 0xd1b54a32d192ed03ull)) + counter);
    return double(h >> 11) * (1.0 / 9007199254740992.0);
  }

  PPF_HD inline double
  uniform(uint64_t seed, uint64_t stream, uint64_t counter, double a, double b)
  {
    return a + (b - a) * uniform01(seed, stream, counter);
  }

  /// Standard normal via Box-Muller on draws 2*counter and 2*counter+1.
  PPF_HD inline double
  normal(uint64_t seed, uint64_t stream, uint64_t counter)
  {
    const double u1 = 1.0 - uniform01(seed, stream, 2 * counter);
    const double u2 = uniform01(seed, stream, 2 * counter + 1);
    return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
  }

  /// Lattice index of a node from its coordinates on a uniform Q_p mesh; periodic
  /// directions wrap the image plane onto the master.
  template <int dim>
  PPF_HD inline uint64_t
  node_key(const double *x, const double *lower, const double *size, const unsigned *cells,
           unsigned degree, const bool *periodic)
  {
    uint64_t key = 0, stride = 1;
    for (int d = 0; d < dim; ++d)
      {
        const unsigned n_lat = degree * cells[d];
        long           i     = std::lround((x[d] - lower[d]) / (size[d] - lower[d]) * n_lat);
        if (periodic[d] && i >= long(n_lat))
          i -= long(n_lat);
        key += uint64_t(i) * stride;
        stride *= uint64_t(n_lat) + 1;
      }
    return key;
  }
} // namespace ppf_noise
