// Tiled realization of the collocation-split operator (see operator_kernels.hpp).
//
// With GLL collocation the assembled operator of a split hook is, at lattice node i,
//
//   (A u)_i = M_i * ( r(values at i) + sum_terms beta * (L u)_i ),
//   (L u)_i = sum_d sum_o W_d(i_d, o) u_{i + o e_d},
//
// with M_i = prod_d m_d(i_d) the lumped mass, r the hook's value term and W_d the 1-D
// taps (assembled 1-D stiffness row) / (assembled 1-D lumped mass) along axis d.  This
// follows from the tensor-product stiffness K = sum_d (prod_{e!=d} M_e) (x) K_d: the
// transverse masses are M_i / m_d(i_d).  The reference computes the same operator cell
// by cell (src/solvers/mf_operator.cc:17-80, FEEvaluation with GLL collocation,
// include/prismspf/core/system_wide.h:40-61); only the order of the floating-point sums
// differs.  Output scaling M^-1 (explicit update) cancels the mass: r + beta (L u).
//
// Interior taps, in units of 1/h^2 (K and w of gll_basis.h):
//   Q1: K = [[1,-1],[-1,1]], w = (1/2, 1/2): (centre, +-1) = (2, -1);
//   Q2: K = [[7,-8,1],[-8,16,-8],[1,-8,7]]/3, w = (1/6, 2/3, 1/6): odd node
//       (8, -4, 0), vertex shared by two cells (14, -8, 1) for (centre, +-1, +-2).
// At a non-periodic boundary the row has only one cell; with mirror ghosts
// u[-k] = u[k], u[n-1+k] = u[n-1-k] the interior taps give exactly that one-sided row
// (Q2: (1,-8,14,-8,1) -> (14,-16,2); Q1: (-1,2,-1) -> (2,-2)), so taps depend only on
// the parity of the coordinate and boundaries are handled by the source index.
//
// Work decomposition (3-D): a workgroup of W wavefronts (4 or 8) owns a tile of
// 64 x (W R) nodes in the x-y plane and marches over a chunk of z-planes.  Lane = x;
// wave w owns rows w, w+W, ..., w+W(R-1) (wave-uniform y).
//   * Every stencil operand value, halo value and hook value input is loaded
//     D = p+1 planes before the plane that consumes it, into a register ring whose
//     slot indices are compile-time constants (the z-loop is unrolled by the ring
//     length), so no load is waited for in the step that issues it.
//   * The centre plane of each operand plus a p-wide halo is staged in LDS
//     (double-buffered: one barrier per plane) for the x and y taps; the z taps read
//     the ring.
//   * The hook runs inlined at every node and each output is written once.
//   * Tiles whose read range touches no Dirichlet face run a step variant without
//     any face test.
//   * Consecutive tiles are mapped to one XCD (workgroups are dealt round-robin to the
//     8 XCDs), so halo rows and chunk boundaries are re-read from the same L2.
// 2-D lattices are the single-plane case (no z ring).
#pragma once

#include "operator_kernels.hpp"
#include "vector_kernels.hpp"

#include <type_traits>
#include <utility>

namespace ppf
{
  constexpr int stencil_tx      = 64; // tile width in x: one wavefront
  constexpr int stencil_max_ops = 2;
  /// Plane steps between the load of a stencil plane and its first use (odd: the ring
  /// length stays even).  Measured at 128^3 Q2: 3 steps made the CH blocks 12.4 ->
  /// 13.8 us and the FP64 Poisson apply 61 -> 70 us (more registers, and a longer
  /// prologue per z chunk), so one step it is.
  constexpr int stencil_lookahead = 1;

  /// Arguments of the stencil kernels: arithmetic in `number`, operand, value and
  /// output arrays stored as ST (ST = double with number = float: FP64 state, FP32
  /// arithmetic).
  template <typename number, typename ST = number>
  struct StencilArgs
  {
    int    n[3];        // nodes per axis (n[2] = 1 for 2-D lattices)
    int    periodic[3];
    int    tiles_x, tiles_y, chunks_z, z_chunk, n_blocks;
    /// Workgroup -> tile order (workgroups are dealt round-robin to the 8 XCDs).
    /// 0: consecutive tiles (x fastest, then y, then z chunk) on one XCD;
    /// 1: XCD k owns the tile rows [k ty_per_xcd, (k+1) ty_per_xcd) of every z chunk and
    ///    walks them chunk by chunk, so a tile's y and z neighbours run on its XCD
    ///    (n_blocks = 8 x the largest per-XCD count; surplus workgroups exit).
    int    map_mode, ty_per_xcd;
    /// Tiles cover x < tile_n[0], y < tile_n[1]; the rest of the lattice (a few columns
    /// or rows left over by the tile widths) is processed node by node by the
    /// workgroups after the n_blocks tile workgroups: box A = [tile_n[0], n0) x [0, n1)
    /// x [0, n2) with rem_a nodes, then box B = [0, tile_n[0]) x [tile_n[1], n1) x
    /// [0, n2) with rem_b nodes.
    int     tile_n[2];
    int64_t rem_a, rem_b;
    /// Taps (centre, +-1, +-2) per axis and coordinate parity (index: i & 1 for Q2,
    /// both equal for Q1), already divided by h^2.
    number tap[3][2][3];
    /// Output scale factor per axis and node class (0 odd / 1 even or Q1 interior /
    /// 2 non-periodic boundary): the 1-D lumped mass for unscaled output, 1 for M^-1
    /// scaling, its square root for M^-1/2 scaling.
    number sfac[3][3];
    double q_lower[3], q_step[3];
    double element_volume;
    int    n_ops;
    const ST     *op[stencil_max_ops];
    unsigned int  op_dmask[stencil_max_ops]; // Dirichlet faces read as zero (0: plain)
    /// Summed coefficient of operand k in output field f.
    number        beta[PPF_MAX_FIELDS][stencil_max_ops];
    const ST     *val[PPF_MAX_FIELDS][PPF_N_SLOTS]; // value inputs of the hook
    unsigned int  val_dmask[PPF_MAX_FIELDS][PPF_N_SLOTS];
    /// Operand index k if the input is operand k (read from the ring), -1 otherwise.
    /// Every `op` and `val` pointer is valid (the host points unused entries at an
    /// array of the lattice size, and an absent operand at an array of zeros), so
    /// loads need no branch.
    signed char   val_alias[PPF_MAX_FIELDS][PPF_N_SLOTS];
    /// Union of all face masks the operator reads or writes (0: no face test at all).
    unsigned int  any_dmask;
    /// Entry-0 NaN probe of an increment, fused into its last launch (probe_result
    /// null: none).  The thread that owns node 0 writes (tag << 5) | (bad + 1) to
    /// *probe_result once its own outputs are stored, bad being the first field whose
    /// entry 0 is not finite (-1: none), and on a bad field sets the words
    /// probe_halt[i] of the bits i of probe_hmask.
    int          *probe_result;
    int          *probe_halt;
    unsigned int  probe_hmask;
    int           probe_tag;
    int           probe_n;
    const double *probe_ptr[PPF_MAX_FIELDS];
    /// Output stores with the non-temporal cache policy (1) or the default one (0).
    int           nt_store;
  };

  template <typename number, typename ST>
  __device__ void
  stencil_probe(const StencilArgs<number, ST> &a)
  {
    // Orders this thread's stores of node 0 before the reads below.
    __threadfence();
    int bad = -1;
    for (int f = 0; f < a.probe_n && f < PPF_MAX_FIELDS; ++f)
      if (bad < 0 && !isfinite(*reinterpret_cast<const volatile double *>(a.probe_ptr[f])))
        bad = f;
    if (bad >= 0 && a.probe_halt)
      for (int i = 0; i < 32; ++i)
        if ((a.probe_hmask >> i) & 1u)
          reinterpret_cast<volatile int *>(a.probe_halt)[i] = 1;
    *reinterpret_cast<volatile int *>(a.probe_result) = (a.probe_tag << 5) | (bad + 1);
  }

  /// Source node along one axis for lattice coordinate i, which may lie up to a stencil
  /// radius outside [0, n): periodic axes wrap, other axes reflect at the boundary node
  /// (the mirror ghosts above).  Coordinates further out are reached only by lanes and
  /// rows without output nodes and are clamped.
  PPF_HD PPF_INLINE int
  source_coord(int i, int n, int periodic)
  {
    if (periodic)
      i = i < 0 ? i + n : (i >= n ? i - n : i);
    else
      i = i < 0 ? -i : (i >= n ? 2 * (n - 1) - i : i);
    return i < 0 ? 0 : (i >= n ? n - 1 : i);
  }

  /// Boundary-face bits of lattice coordinate i on `axis` (bit 2 axis: lower face,
  /// bit 2 axis + 1: upper face; none in periodic axes).
  PPF_HD PPF_INLINE unsigned int
  face_bits(int i, int n, int periodic, int axis)
  {
    if (periodic)
      return 0u;
    return (i == 0 ? 1u << (2 * axis) : 0u) | (i == n - 1 ? 1u << (2 * axis + 1) : 0u);
  }

  /// Node class of coordinate i for the output scale factor.
  template <int degree>
  PPF_HD PPF_INLINE int
  scale_class(int i, int n, int periodic)
  {
    if (!periodic && (i == 0 || i == n - 1))
      return 2;
    return (degree == 2 && (i & 1)) ? 0 : 1;
  }

  /// Host-side 1-D lumped mass of scale class c, in units of h.
  template <int degree>
  inline double
  class_mass_units(int c)
  {
    if (degree == 2)
      return c == 0 ? 2.0 / 3.0 : (c == 1 ? 1.0 / 3.0 : 1.0 / 6.0);
    return c == 2 ? 0.5 : 1.0;
  }

  /// True if the read range [lo - rad, hi + rad] of a tile along an axis contains a
  /// boundary node (only non-periodic axes have one).
  PPF_HD PPF_INLINE bool
  touches_face(int lo, int hi, int rad, int n, int periodic)
  {
    return !periodic && (lo - rad <= 0 || hi + rad >= n - 1);
  }

  using buffer_rsrc = __amdgpu_buffer_rsrc_t;
  typedef unsigned int u32x2 __attribute__((__vector_size__(2 * sizeof(unsigned int))));

  /// Raw buffer resource over an array: loads and stores take a 32-bit per-lane byte
  /// offset (kept in a register for the whole kernel) plus a scalar byte offset (the
  /// plane), with no 64-bit address arithmetic.  No range limit is set (the callers'
  /// arrays stay below 4 GiB); word 3 selects the 32-bit data format.
  __device__ PPF_INLINE buffer_rsrc
  make_rsrc(const void *p)
  {
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(p), short(0), -1, 0x00020000);
  }

  template <typename number>
  __device__ PPF_INLINE number
  buffer_load(buffer_rsrc r, unsigned int voff, unsigned int soff)
  {
    if constexpr (sizeof(number) == 8)
      return __builtin_bit_cast(number, __builtin_amdgcn_raw_buffer_load_b64(r, voff, soff, 0));
    else
      return __builtin_bit_cast(number, __builtin_amdgcn_raw_buffer_load_b32(r, voff, soff, 0));
  }

  /// Output store; `nt` (uniform over the launch) selects the non-temporal cache
  /// policy, bit 1 of the buffer instruction's cache-policy operand (NT on gfx950).
  /// It keeps a launch's output from displacing its operand planes and halo rows in L2
  /// and the Infinity Cache, which pays when the output will not be read from there
  /// anyway (large lattices); a small lattice's output is read from the caches by the
  /// next launch, and there the default policy is faster.
  template <typename number>
  __device__ PPF_INLINE void
  buffer_store(number v, buffer_rsrc r, unsigned int voff, unsigned int soff, bool nt)
  {
    if constexpr (sizeof(number) == 8)
      {
        if (nt)
          __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(u32x2, v), r, voff, soff, 2);
        else
          __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(u32x2, v), r, voff, soff, 0);
      }
    else
      {
        if (nt)
          __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned int, v), r, voff,
                                                soff, 2);
        else
          __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned int, v), r, voff,
                                                soff, 0);
      }
  }

  /// Operand s at lattice coordinate (x, y, z), which may lie a stencil radius outside
  /// the lattice (source node by wrap / mirror), zero on the operand's Dirichlet faces.
  template <typename number, typename ST>
  __device__ PPF_INLINE number
  operand_at(const StencilArgs<number, ST> &a, int s, int x, int y, int z)
  {
    const int xs = source_coord(x, a.n[0], a.periodic[0]);
    const int ys = source_coord(y, a.n[1], a.periodic[1]);
    const int zs = source_coord(z, a.n[2], a.periodic[2]);
    const unsigned int fb = face_bits(xs, a.n[0], a.periodic[0], 0) |
                            face_bits(ys, a.n[1], a.periodic[1], 1) |
                            face_bits(zs, a.n[2], a.periodic[2], 2);
    const number v =
      number(a.op[s][int64_t(xs) + int64_t(a.n[0]) * (int64_t(ys) + int64_t(a.n[1]) * zs)]);
    return (fb & a.op_dmask[s]) ? number(0) : v;
  }

  /// The operator at one node, evaluated directly from global memory (the remainder
  /// workgroups of stencil_operator_kernel; same arithmetic as the tiles).
  template <class PDE, int dim, int degree, typename number, bool lhs, int block_id, int NF,
            int NS, typename ST = number>
  __device__ void
  stencil_node(const PDE &pde, const StencilArgs<number, ST> &a, const OutTable<ST> &out,
               const prismspf::SimulationTimer &timer, int x, int y, int z)
  {
    constexpr int RAD = degree;
    const int     c[3] = {x, y, z};
    number        tap[3][3], sf[3];
    unsigned int  fb = 0;
#pragma unroll
    for (int d = 0; d < 3; ++d)
      {
        const bool odd = degree == 2 && (c[d] & 1);
        for (int k = 0; k < 3; ++k)
          tap[d][k] = odd ? a.tap[d][1][k] : a.tap[d][0][k];
        const int cl = scale_class<degree>(c[d], a.n[d], a.periodic[d]);
        sf[d]        = cl == 0 ? a.sfac[d][0] : (cl == 1 ? a.sfac[d][1] : a.sfac[d][2]);
        fb |= face_bits(c[d], a.n[d], a.periodic[d], d);
      }
    // Centre tap and scale factor associated as in the tiles: x + (y + z), s_x (s_y s_z).
    const number sc = sf[0] * (sf[1] * (dim == 3 ? sf[2] : number(1)));
    number lap[NS];
#pragma unroll
    for (int s = 0; s < NS; ++s)
      {
        number acc = (tap[0][0] + (tap[1][0] + (dim == 3 ? tap[2][0] : number(0)))) *
                     operand_at(a, s, x, y, z);
#pragma unroll
        for (int o = 1; o <= RAD; ++o)
          {
            acc = fma(tap[0][o], operand_at(a, s, x - o, y, z) + operand_at(a, s, x + o, y, z), acc);
            acc = fma(tap[1][o], operand_at(a, s, x, y - o, z) + operand_at(a, s, x, y + o, z), acc);
            if constexpr (dim == 3)
              acc = fma(tap[2][o], operand_at(a, s, x, y, z - o) + operand_at(a, s, x, y, z + o), acc);
          }
        lap[s] = acc;
      }
    const int64_t node = int64_t(x) + int64_t(a.n[0]) * (int64_t(y) + int64_t(a.n[1]) * z);
    prismspf::FieldContainer<dim, degree, number> fc;
    fc.q_point[0] = number(a.q_lower[0] + x * a.q_step[0]);
    fc.q_point[1] = number(a.q_lower[1] + y * a.q_step[1]);
    if constexpr (dim == 3)
      fc.q_point[2] = number(a.q_lower[2] + z * a.q_step[2]);
    fc.element_volume = number(a.element_volume);
#pragma unroll
    for (int f = 0; f < PPF_MAX_FIELDS; ++f)
#pragma unroll
      for (int s = 0; s < PPF_N_SLOTS; ++s)
        {
          number v = number(0);
          if (f < NF)
            {
              const int fi = f < NF ? f : 0;
              const int al = a.val_alias[fi][s];
              v = (fb & a.val_dmask[fi][s]) ? number(0) : number(a.val[fi][s][node]);
              if (al >= 0)
                v = operand_at(a, al < NS ? al : 0, x, y, z);
            }
          fc.value[f][s] = v;
#pragma unroll
          for (int d = 0; d < dim; ++d)
            fc.gradient[f][s][d] = number(0);
        }
    fc.clear_terms();
    call_hook<PDE, dim, degree, number, lhs, block_id>(pde, fc, timer);
#pragma unroll
    for (int f = 0; f < NF; ++f)
      {
        if (!(out.out_mask & (1u << f)))
          continue;
        number acc = (fc.value_term_mask & (1u << f)) ? fc.value_term[f] : number(0);
#pragma unroll
        for (int k = 0; k < NS; ++k)
          acc = fma(a.beta[f][k], lap[k], acc);
        const bool dir = (fb & out.dirichlet_mask[f]) != 0;
        if (!(dir && out.dirichlet_write == dirichlet_value))
          out.dst[f][node] = dir ? ST(0) : ST(sc * acc);
      }
  }

  template <class PDE, int dim, int degree, typename number, bool lhs, int block_id, int NF,
            int WV, int R, int NS, typename ST = number>
  __global__ void __launch_bounds__(stencil_tx * WV)
  stencil_operator_kernel(const PDE                       pde,
                          const StencilArgs<number, ST>   a,
                          const OutTable<ST>              out,
                          const prismspf::SimulationTimer timer)
  {
    constexpr int RAD = degree;
    constexpr int T   = 2 * RAD + 1;
    constexpr int TY  = WV * R;
    constexpr int LX  = stencil_tx + 2 * RAD;
    constexpr int LY  = TY + 2 * RAD;
    constexpr int LA  = stencil_lookahead; // steps between a plane's load and its use
    constexpr int D   = RAD + LA;          // load distance in planes
    constexpr int S   = T + LA;            // ring length
    static_assert(LA % 2 == 1, "the ring length must be even (z parity from the phase)");
    static_assert(WV >= 2 * RAD, "the y-halo rows are loaded by one wave each");
    __shared__ number lds[2][NS][LY + 1][LX]; // row LY: spare target of idle lanes

    const int tx  = threadIdx.x;
    const int w   = __builtin_amdgcn_readfirstlane(threadIdx.y);
    const int tid = w * stencil_tx + tx;
    const int n0 = a.n[0], n1 = a.n[1], n2 = a.n[2];
    const int p0 = a.periodic[0], p1 = a.periodic[1], p2 = a.periodic[2];
    if (halted(out.halt))
      return;

    if (int(blockIdx.x) >= a.n_blocks)
      {
        // Remainder workgroup: one node per thread.
        int64_t t = int64_t(int(blockIdx.x) - a.n_blocks) * (stencil_tx * WV) +
                    threadIdx.y * stencil_tx + threadIdx.x;
        int x, y, z;
        if (t < a.rem_a)
          {
            const int w = n0 - a.tile_n[0];
            x           = a.tile_n[0] + int(t % w);
            t /= w;
            y = int(t % n1);
            z = int(t / n1);
          }
        else if (t < a.rem_a + a.rem_b)
          {
            t -= a.rem_a;
            const int w = a.tile_n[0], hgt = n1 - a.tile_n[1];
            x           = int(t % w);
            t /= w;
            y = a.tile_n[1] + int(t % hgt);
            z = int(t / hgt);
          }
        else
          return;
        stencil_node<PDE, dim, degree, number, lhs, block_id, NF, NS, ST>(pde, a, out, timer, x,
                                                                          y, z);
        return;
      }

    const int b = blockIdx.x;
    int       t_x, t_y, chk;
    if (a.map_mode == 1)
      {
        const int xcd = b % 8, i = b / 8;
        const int ty0 = xcd * a.ty_per_xcd;
        const int cnt = min(a.ty_per_xcd, a.tiles_y - ty0);
        if (cnt <= 0 || i >= a.tiles_x * cnt * a.chunks_z)
          return;
        t_x         = i % a.tiles_x;
        const int r = i / a.tiles_x;
        t_y         = ty0 + r % cnt;
        chk         = r / cnt;
      }
    else
      {
        const int q8  = a.n_blocks / 8;
        const int lb  = b < 8 * q8 ? (b % 8) * q8 + b / 8 : b;
        t_x           = lb % a.tiles_x;
        const int t_r = lb / a.tiles_x;
        t_y           = t_r % a.tiles_y;
        chk           = t_r / a.tiles_y;
      }
    const int x0  = t_x * stencil_tx;
    const int y0  = t_y * TY;
    const int zb  = dim == 3 ? chk * a.z_chunk : 0; // even for Q2 (host)
    const int ze  = dim == 3 ? min(zb + a.z_chunk, n2) : 1;
    // Byte offsets fit 32 bits: the host runs this kernel on arrays below 4 GiB only.
    const unsigned int plane_bytes = unsigned(n0) * unsigned(n1) * unsigned(sizeof(ST));

    const bool  nt = a.nt_store != 0;
    buffer_rsrc r_op[NS], r_val[NF][PPF_N_SLOTS], r_dst[NF];
#pragma unroll
    for (int s = 0; s < NS; ++s)
      r_op[s] = make_rsrc(a.op[s]);
#pragma unroll
    for (int f = 0; f < NF; ++f)
      {
        r_dst[f] = make_rsrc(out.dst[f]);
#pragma unroll
        for (int s = 0; s < PPF_N_SLOTS; ++s)
          r_val[f][s] = make_rsrc(a.val[f][s]);
      }

    // Own column (lane-constant).
    const int          x    = x0 + tx;
    const bool         xout = x < a.tile_n[0]; // owns output nodes
    const int          xs   = source_coord(x, n0, p0);
    const unsigned int fbx  = face_bits(xs, n0, p0, 0);
    const int          pxi  = degree == 2 ? (xs & 1) : 0;
    const number       wx0  = pxi ? a.tap[0][1][0] : a.tap[0][0][0];
    const number       wx1  = pxi ? a.tap[0][1][1] : a.tap[0][0][1];
    const number       wx2  = pxi ? a.tap[0][1][2] : a.tap[0][0][2];
    const int          cx   = scale_class<degree>(xs, n0, p0);
    const number       sx   = cx == 0 ? a.sfac[0][0] : (cx == 1 ? a.sfac[0][1] : a.sfac[0][2]);

    // Own rows (wave-uniform): source row, face bits, byte offset, taps, scale factor.
    int          ys[R];
    bool         yout[R];
    unsigned int fby[R], own_b[R];
    number       wy0[R], wy1[R], wy2[R], sy[R];
#pragma unroll
    for (int r = 0; r < R; ++r)
      {
        const int y = y0 + w + WV * r;
        yout[r]     = y < a.tile_n[1];
        ys[r]       = source_coord(y, n1, p1);
        fby[r]      = face_bits(ys[r], n1, p1, 1);
        own_b[r]    = (unsigned(xs) + unsigned(n0) * unsigned(ys[r])) * unsigned(sizeof(ST));
        const bool py = degree == 2 && (ys[r] & 1);
        wy0[r]        = py ? a.tap[1][1][0] : a.tap[1][0][0];
        wy1[r]        = py ? a.tap[1][1][1] : a.tap[1][0][1];
        wy2[r]        = py ? a.tap[1][1][2] : a.tap[1][0][2];
        const int cy  = scale_class<degree>(ys[r], n1, p1);
        sy[r]         = cy == 0 ? a.sfac[1][0] : (cy == 1 ? a.sfac[1][1] : a.sfac[1][2]);
      }

    // Halo assignments: wave w < 2 RAD loads one y-halo row of the tile's columns;
    // threads tid < 2 RAD TY load one x-halo node (the others repeat the y-halo load).
    // Lanes without a halo node store to the spare LDS row, so staging has no branch.
    __builtin_assume(w >= 0 && w < WV);
    const int          hy_row = w < RAD ? w : TY + w; // LDS row
    const bool         hy_on  = w < 2 * RAD;
    const int          hy_y   = source_coord(y0 - RAD + hy_row, n1, p1);
    const unsigned int fb_hy  = fbx | face_bits(hy_y, n1, p1, 1);
    const int          hx_k   = tid % (2 * RAD);
    const int          hx_row = tid / (2 * RAD); // tile row
    const int          hx_col = hx_k < RAD ? hx_k : stencil_tx + hx_k; // LDS column
    const bool         hx_on  = tid < 2 * RAD * TY;
    const int          hx_x   = source_coord(x0 - RAD + hx_col, n0, p0);
    const int          hx_y   = source_coord(y0 + hx_row, n1, p1);
    const unsigned int fb_hx  = face_bits(hx_x, n0, p0, 0) | face_bits(hx_y, n1, p1, 1);
    const unsigned int hy_b =
      (unsigned(xs) + unsigned(n0) * unsigned(hy_y)) * unsigned(sizeof(ST));
    const unsigned int hx_b =
      hx_on ? (unsigned(hx_x) + unsigned(n0) * unsigned(hx_y)) * unsigned(sizeof(ST)) : hy_b;
    const int hy_ly = hy_on ? hy_row : LY;
    const int hx_ly = hx_on ? RAD + hx_row : LY;
    const int hx_lx = hx_on ? hx_col : tx;

    // Scale factors of the z classes (the taps follow the step's parity).
    const number sz_odd = a.sfac[2][0], sz_even = a.sfac[2][1], sz_bnd = a.sfac[2][2];

    // Register rings: slot of absolute plane p = (p - zb + RAD) mod S.
    // (stored as loaded; converted to the arithmetic type where they are used, so no
    // load is waited for before its plane is reached)
    ST           ring[NS][R][S];
    ST           ring_hy[NS][S], ring_hx[NS][S];
    ST           ring_v[S][R][NF][PPF_N_SLOTS];
    unsigned int ring_fb[S]; // face bits of the plane in the slot (uniform)

    // Source plane of z (wrap / mirror only outside [0, n2)).
    auto source_z = [&](int z) {
      return unsigned(z) < unsigned(n2) ? z : source_coord(z, n2, p2);
    };
    auto load_own = [&](int z, int slot) {
      const int          zs = source_z(z);
      const unsigned int po = unsigned(zs) * plane_bytes;
      ring_fb[slot]         = face_bits(zs, n2, p2, 2);
#pragma unroll
      for (int s = 0; s < NS; ++s)
#pragma unroll
        for (int r = 0; r < R; ++r)
          ring[s][r][slot] = buffer_load<ST>(r_op[s], own_b[r], po);
    };
    auto load_inputs = [&](int z, int slot) {
      const unsigned int po = unsigned(source_z(z)) * plane_bytes;
#pragma unroll
      for (int s = 0; s < NS; ++s)
        {
          ring_hy[s][slot] = buffer_load<ST>(r_op[s], hy_b, po);
          ring_hx[s][slot] = buffer_load<ST>(r_op[s], hx_b, po);
        }
#pragma unroll
      for (int r = 0; r < R; ++r)
#pragma unroll
        for (int f = 0; f < NF; ++f)
#pragma unroll
          for (int s = 0; s < PPF_N_SLOTS; ++s)
            ring_v[slot][r][f][s] = buffer_load<ST>(r_val[f][s], own_b[r], po);
    };

    // Zeroes the operands of ring slot k on their Dirichlet faces.
    auto mask_slot = [&](int k) {
#pragma unroll
      for (int s = 0; s < NS; ++s)
#pragma unroll
        for (int r = 0; r < R; ++r)
          ring[s][r][k] = ((fbx | fby[r] | ring_fb[k]) & a.op_dmask[s]) ? ST(0)
                                                                       : ring[s][r][k];
    };

    // One plane.  MASKED: face tests (constrained reads, Dirichlet outputs); without
    // it the tile's read range contains no boundary node of a masked face.  Chunks
    // start at even z and the ring length S is even, so plane z has the parity of PH.
    auto step = [&]<int PH, bool MASKED>(std::integral_constant<int, PH>,
                                          std::integral_constant<bool, MASKED>, int z) {
      constexpr int cs  = (PH + RAD) % S; // centre slot
      constexpr int ns  = (PH + S - 1) % S;
      constexpr int buf = PH & 1;
      if constexpr (dim == 3)
        {
          load_own(min(z + D, ze - 1 + RAD), ns);
          load_inputs(min(z + D, ze - 1), ns);
        }

      // In masked tiles the plane entering the z taps (z + RAD) is zeroed on the
      // operands' Dirichlet faces once, in place; every other ring plane in use has
      // been zeroed before (or by the prologue).
      if constexpr (MASKED && dim == 3)
        mask_slot((cs + RAD) % S);
      else if constexpr (MASKED)
        mask_slot(cs);

      // Stage the centre plane.
      number ctr[NS][R];
#pragma unroll
      for (int s = 0; s < NS; ++s)
        {
#pragma unroll
          for (int r = 0; r < R; ++r)
            {
              ctr[s][r] = number(ring[s][r][cs]);
              lds[buf][s][RAD + w + WV * r][RAD + tx] = ctr[s][r];
            }
          number hyv = number(ring_hy[s][cs]), hxv = number(ring_hx[s][cs]);
          if constexpr (MASKED)
            {
              const unsigned int dm = a.op_dmask[s];
              hyv = ((fb_hy | ring_fb[cs]) & dm) ? number(0) : hyv;
              hxv = ((fb_hx | ring_fb[cs]) & dm) ? number(0) : hxv;
            }
          lds[buf][s][hy_ly][RAD + tx] = hyv;
          lds[buf][s][hx_ly][hx_lx]    = hxv;
        }

      // z taps (compile-time parity) and scale factor (uniform).
      const unsigned int fbz  = ring_fb[cs];
      const unsigned int zoff = unsigned(z) * plane_bytes;
      number             wz0 = number(0), wz1 = number(0), wz2 = number(0), sz = number(1);
      if constexpr (dim == 3)
        {
          constexpr bool pz = degree == 2 && (PH & 1);
          wz0               = pz ? a.tap[2][1][0] : a.tap[2][0][0];
          wz1               = pz ? a.tap[2][1][1] : a.tap[2][0][1];
          wz2               = pz ? a.tap[2][1][2] : a.tap[2][0][2];
          sz = (!p2 && (z == 0 || z == n2 - 1)) ? sz_bnd : (pz ? sz_odd : sz_even);
        }
      __syncthreads();

#pragma unroll
      for (int r = 0; r < R; ++r)
        {
          const number c0  = wx0 + (wy0[r] + wz0);
          const number syz = sy[r] * sz;
          const int    ly  = RAD + w + WV * r;
          const int    lx  = RAD + tx;

          number lap[NS];
#pragma unroll
          for (int s = 0; s < NS; ++s)
            {
              const number(&L)[LY + 1][LX] = lds[buf][s];
              number acc = c0 * ctr[s][r];
              acc = fma(wx1, L[ly][lx - 1] + L[ly][lx + 1], acc);
              acc = fma(wy1[r], L[ly - 1][lx] + L[ly + 1][lx], acc);
              if constexpr (dim == 3)
                acc = fma(wz1, number(ring[s][r][(cs + S - 1) % S]) +
                              number(ring[s][r][(cs + 1) % S]), acc);
              if constexpr (RAD == 2)
                {
                  acc = fma(wx2, L[ly][lx - 2] + L[ly][lx + 2], acc);
                  acc = fma(wy2[r], L[ly - 2][lx] + L[ly + 2][lx], acc);
                  if constexpr (dim == 3)
                    acc = fma(wz2, number(ring[s][r][(cs + S - 2) % S]) +
                                  number(ring[s][r][(cs + 2) % S]), acc);
                }
              lap[s] = acc;
            }

          // Hook at the node; lanes and rows without an output node evaluate it on
          // clamped values and store nothing.
          const unsigned int fb = fbx | fby[r] | fbz;
          prismspf::FieldContainer<dim, degree, number> fc;
          fc.q_point[0] = number(a.q_lower[0] + xs * a.q_step[0]);
          fc.q_point[1] = number(a.q_lower[1] + ys[r] * a.q_step[1]);
          if constexpr (dim == 3)
            fc.q_point[2] = number(a.q_lower[2] + z * a.q_step[2]);
          fc.element_volume = number(a.element_volume);
#pragma unroll
          for (int f = 0; f < PPF_MAX_FIELDS; ++f)
#pragma unroll
            for (int s = 0; s < PPF_N_SLOTS; ++s)
              {
                number v = number(0);
                if (f < NF)
                  {
                    const int fi = f < NF ? f : 0;
                    const int al = a.val_alias[fi][s];
                    v            = number(ring_v[cs][r][fi][s]);
                    if constexpr (MASKED)
                      v = (fb & a.val_dmask[fi][s]) ? number(0) : v;
                    if (al == 0)
                      v = ctr[0][r];
                    if (NS > 1 && al == 1)
                      v = ctr[NS > 1 ? 1 : 0][r];
                  }
                fc.value[f][s] = v;
#pragma unroll
                for (int d = 0; d < dim; ++d)
                  fc.gradient[f][s][d] = number(0);
              }
          fc.clear_terms();
          call_hook<PDE, dim, degree, number, lhs, block_id>(pde, fc, timer);

          const number scale = sx * syz;
          const bool   valid = xout && yout[r];
#pragma unroll
          for (int f = 0; f < NF; ++f)
            {
              if (!(out.out_mask & (1u << f)))
                continue;
              number acc = (fc.value_term_mask & (1u << f)) ? fc.value_term[f] : number(0);
#pragma unroll
              for (int k = 0; k < NS; ++k)
                acc = fma(a.beta[f][k], lap[k], acc);
              if constexpr (MASKED)
                {
                  // Dirichlet nodes: zero (operator rows dropped), or left to
                  // dirichlet_faces_kernel, which writes the constraint values.
                  const bool dir = (fb & out.dirichlet_mask[f]) != 0;
                  if (valid && !(dir && out.dirichlet_write == dirichlet_value))
                    buffer_store(dir ? ST(0) : ST(scale * acc), r_dst[f], own_b[r], zoff, nt);
                }
              else if (valid)
                buffer_store(ST(scale * acc), r_dst[f], own_b[r], zoff, nt);
            }
          // Node 0 (lane 0, wave 0, row 0, plane 0 of workgroup 0, which owns tile 0 of
          // chunk 0) closes the increment's NaN probe.
          if constexpr (PH == 0)
            if (a.probe_result && r == 0 && z == 0 && blockIdx.x == 0 && w == 0 && tx == 0)
              stencil_probe(a);
        }
    };

    auto run = [&]<bool MASKED>(std::integral_constant<bool, MASKED> m) {
      if constexpr (dim == 3)
        {
          // Prologue: own planes zb-RAD .. zb+RAD (slots 0 .. 2 RAD), halo and value
          // inputs of planes zb .. zb+RAD (slots RAD .. 2 RAD).
#pragma unroll
          for (int k = 0; k < T + LA - 1; ++k)
            load_own(min(zb - RAD + k, ze - 1 + RAD), k);
#pragma unroll
          for (int k = RAD; k < T + LA - 1; ++k)
            load_inputs(min(zb - RAD + k, ze - 1), k);
          if constexpr (MASKED)
            {
              // planes zb-RAD .. zb+RAD-1 never enter as the newest z tap
#pragma unroll
              for (int k = 0; k < 2 * RAD; ++k)
                mask_slot(k);
            }
          for (int z = zb; z < ze; z += S)
            {
              [&]<int... PH>(std::integer_sequence<int, PH...>) {
                ((z + PH < ze ? step(std::integral_constant<int, PH> {}, m, z + PH) : void()),
                 ...);
              }(std::make_integer_sequence<int, S> {});
            }
        }
      else
        {
          load_own(0, RAD);
          load_inputs(0, RAD);
          step(std::integral_constant<int, 0> {}, m, 0);
        }
    };

    const bool masked = a.any_dmask != 0 &&
                        (touches_face(x0, x0 + stencil_tx - 1, RAD, n0, p0) ||
                         touches_face(y0, y0 + TY - 1, RAD, n1, p1) ||
                         (dim == 3 && touches_face(zb, ze - 1, RAD, n2, p2)));
    if (masked)
      run(std::integral_constant<bool, true> {});
    else
      run(std::integral_constant<bool, false> {});
  }

  /// Vectors and limits of a coarse-grid CG run by one workgroup.
  struct CoarseCgArgs
  {
    double       *x, *r, *p, *ap, *z;
    const double *inv_diag;
    int64_t       n;       // block vector length
    double        tol;     // stop at |r| <= tol |r0|
    int           max_its;
  };

  /// The coarse-grid CG with the diagonal preconditioner (same recurrences and stopping
  /// rule as the host-driven loop) in one workgroup: the level operator is evaluated
  /// node by node (stencil_node, operand a.op = p, output out.dst = ap), dot products are
  /// fixed-order block sums, and no host round trip or further launch is needed.
  template <class PDE, int dim, int degree, int block_id, int NF, int NS>
  __global__ void __launch_bounds__(reduce_threads)
  coarse_cg_kernel(const PDE pde, const StencilArgs<double> a, const OutTable<double> out,
                   const prismspf::SimulationTimer timer, const CoarseCgArgs c)
  {
    __shared__ double bcast;
    const int tid = threadIdx.x, nt = blockDim.x;
    auto      dot = [&](const double *u, const double *v) {
      double s = 0.0;
      for (int64_t i = tid; i < c.n; i += nt)
        s += u[i] * v[i];
      s = block_sum(s);
      if (tid == 0)
        bcast = s;
      __syncthreads();
      const double r = bcast;
      __syncthreads();
      return r;
    };
    const int64_t n_nodes = int64_t(a.n[0]) * a.n[1] * a.n[2];
    for (int64_t i = tid; i < c.n; i += nt)
      c.x[i] = 0.0;
    const double r0 = std::sqrt(dot(c.r, c.r));
    if (!(r0 > 0.0))
      return;
    for (int64_t i = tid; i < c.n; i += nt)
      {
        c.z[i] = c.inv_diag[i] * c.r[i];
        c.p[i] = c.z[i];
      }
    __syncthreads();
    double rz = dot(c.r, c.z);
    for (int it = 0; it < c.max_its; ++it)
      {
        for (int64_t node = tid; node < n_nodes; node += nt)
          {
            const int     x = int(node % a.n[0]);
            const int64_t t = node / a.n[0];
            stencil_node<PDE, dim, degree, double, true, block_id, NF, NS>(
              pde, a, out, timer, x, int(t % a.n[1]), int(t / a.n[1]));
          }
        __syncthreads();
        const double alpha = rz / dot(c.p, c.ap);
        for (int64_t i = tid; i < c.n; i += nt)
          {
            c.x[i] = alpha * c.p[i] + 1.0 * c.x[i];
            c.r[i] = (-alpha) * c.ap[i] + 1.0 * c.r[i];
          }
        __syncthreads();
        if (std::sqrt(dot(c.r, c.r)) <= c.tol * r0)
          break;
        for (int64_t i = tid; i < c.n; i += nt)
          c.z[i] = c.inv_diag[i] * c.r[i];
        __syncthreads();
        const double rz_new = dot(c.r, c.z);
        const double beta   = rz_new / rz;
        for (int64_t i = tid; i < c.n; i += nt)
          c.p[i] = 1.0 * c.z[i] + beta * c.p[i];
        rz = rz_new;
        __syncthreads();
      }
  }

  /// Writes the constraint values of the Dirichlet faces in `mask` into dst (explicit
  /// update semantics, reference explicit_solver.h:63-77 with constraints distributed).
  /// One thread per node of a face; blockIdx.y is the face (2 axis + upper).
  template <typename number>
  __global__ void __launch_bounds__(256)
  dirichlet_faces_kernel(number *__restrict__ dst, const number *__restrict__ values, int n0,
                         int n1, int n2, int p0, int p1, int p2, unsigned int mask,
                         const int *halt)
  {
    if (halted(halt))
      return;
    const int face   = blockIdx.y;
    const int axis   = face >> 1;
    const int n[3]   = {n0, n1, n2};
    const int per[3] = {p0, p1, p2};
    if (!(mask & (1u << face)) || per[axis] || n[axis] < 1)
      return;
    const int     u = axis == 0 ? 1 : 0;
    const int     v = axis == 2 ? 1 : 2;
    const int64_t t = int64_t(blockIdx.x) * 256 + threadIdx.x;
    if (t >= int64_t(n[u]) * n[v])
      return;
    int c[3];
    c[axis]           = (face & 1) ? n[axis] - 1 : 0;
    c[u]              = int(t % n[u]);
    c[v]              = int(t / n[u]);
    const int64_t idx = int64_t(c[0]) + int64_t(n0) * (int64_t(c[1]) + int64_t(n1) * c[2]);
    dst[idx]          = values[idx];
  }
} // namespace ppf
