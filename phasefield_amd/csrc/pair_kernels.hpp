// Two consecutive explicit solve blocks in one z-march (3-D, split operators).
//
// Block A updates one field, u_A = s (r_A + beta_A K op_A), from values of the state
// before the increment; block B then updates another, u_B = s (r_B + beta_B K u_A),
// reading A's output through its stencil operand (and optionally its values).  The
// reference runs them one after the other (src/core/solver/explicit_solver.h:63-77 for
// each block in solve order), which here is two stencil launches: A's launch writes
// u_A, B's launch reads it back with its halo.  This kernel does both in one launch:
// a workgroup computes u_A on its tile grown by the stencil radius p in x and y (and p
// planes ahead in z), keeps it in shared memory and registers, and evaluates B from
// there.  u_A is still stored once (by the tile that owns each node), so the results
// are those of the two launches: the same taps, hook and summation order at every node
// (u_A on a grown-tile node outside the lattice is evaluated at its wrap / mirror source
// node, whose value the second launch would have read).
//
// Work decomposition: 4 main waves own a tile of 64 x TY nodes (lane = x, TY = 4 BR
// rows); the A region is (64 + 2p) x (TY + 2p).  Main wave w computes A on region rows
// w, w + 4, ... (AR rows) of its 64 columns, and B on the BR of those rows that lie in
// the tile; a fifth (side) wave computes A on the 2p side columns of the A region, one
// node per lane.  Operand planes are staged in shared memory with a 2p halo, A's output
// planes with a p halo (p + 1 planes kept); z taps read register rings.  One barrier per
// plane.  Lattice columns beyond the tiles (a few, e.g. 129 = 2 x 64 + 1) are processed
// node by node by pair_node_kernel, which evaluates A at the node and its stencil
// neighbours from global memory.
#pragma once

#include "stencil_kernels.hpp"

namespace ppf
{
  constexpr int pair_main_waves = 4;
  constexpr int pair_waves      = pair_main_waves + 1;

  template <typename number>
  struct PairArgs
  {
    int     n[3];
    int     periodic[3];
    int     tiles_x, tiles_y, chunks_z, z_chunk, n_blocks;
    /// Tiles own the columns x < tile_n0 (rows: all); the columns beyond are the
    /// remainder, rem = (n0 - tile_n0) n1 n2 nodes (pair_node_kernel).
    int     tile_n0;
    int64_t rem;
    number  tap[3][2][3]; ///< (centre, +-1, +-2) per axis and parity, divided by h^2
    number  sfac[3][3];   ///< output scale factor per axis and node class
    double  q_lower[3], q_step[3];
    double  element_volume;
    // Block A.
    const number *op_a;
    number        beta_a;
    int           fa;
    number       *dst_a;
    const number *val_a[PPF_MAX_FIELDS][PPF_N_SLOTS];
    signed char   alias_a[PPF_MAX_FIELDS][PPF_N_SLOTS]; ///< 0: op_a at the node, -1: val_a
    // Block B (stencil operand: A's output).
    number        beta_b;
    int           fb;
    number       *dst_b;
    const number *val_b[PPF_MAX_FIELDS][PPF_N_SLOTS];
    signed char   alias_b[PPF_MAX_FIELDS][PPF_N_SLOTS]; ///< 0: A's output at the node
    /// Entry-0 NaN probe of the increment (probe_result null: none), as in StencilArgs.
    int          *probe_result;
    int          *probe_halt;
    unsigned int  probe_hmask;
    int           probe_tag;
    int           probe_n;
    const double *probe_ptr[PPF_MAX_FIELDS];
    int           nt_store;
    const int    *halt;
  };

  // (the functions below are inlined: a non-inlined callee taking the kernel's
  // argument block by reference makes the compiler copy it to scratch memory)
  template <typename number>
  __device__ PPF_INLINE void
  pair_probe(const PairArgs<number> &a)
  {
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

  /// The hook's value term of field f (0 if it set none), selected with compile-time
  /// indices (a runtime index would put the container in scratch memory).
  template <int NF, class FC>
  __device__ PPF_INLINE auto
  pair_value_term(const FC &fc, int f)
  {
    decltype(fc.value_term[0] + 0) v = 0;
#pragma unroll
    for (int g = 0; g < NF; ++g)
      if (g == f && (fc.value_term_mask & (1u << g)))
        v = fc.value_term[g];
    return v;
  }

  /// Output scale factor of a node from its source coordinates.
  template <int degree, typename number>
  __device__ PPF_INLINE number
  pair_scale(const PairArgs<number> &a, int xs, int ys, int zs)
  {
    number s[3];
    const int c[3] = {xs, ys, zs};
#pragma unroll
    for (int d = 0; d < 3; ++d)
      {
        const int cl = scale_class<degree>(c[d], a.n[d], a.periodic[d]);
        s[d]         = cl == 0 ? a.sfac[d][0] : (cl == 1 ? a.sfac[d][1] : a.sfac[d][2]);
      }
    return s[0] * (s[1] * s[2]);
  }

  /// Taps (centre, +-1, +-2) of a node per axis from the parity of its source
  /// coordinates, selected with compile-time indices.
  template <int degree, typename number>
  __device__ PPF_INLINE void
  pair_taps(const PairArgs<number> &a, int xs, int ys, int zs, number *tx, number *ty,
            number *tz)
  {
    const bool ox = degree == 2 && (xs & 1), oy = degree == 2 && (ys & 1),
               oz = degree == 2 && (zs & 1);
#pragma unroll
    for (int o = 0; o <= degree; ++o)
      {
        tx[o] = ox ? a.tap[0][1][o] : a.tap[0][0][o];
        ty[o] = oy ? a.tap[1][1][o] : a.tap[1][0][o];
        tz[o] = oz ? a.tap[2][1][o] : a.tap[2][0][o];
      }
  }

  /// Block A at position (px, py, pz) (up to 2p outside the lattice: evaluated at its
  /// source node), from global memory; the summation order of the tiles.
  template <class PDE, int degree, typename number, int id_a, int NF>
  __device__ PPF_INLINE number
  pair_a_at(const PDE &pde, const PairArgs<number> &a, const prismspf::SimulationTimer &timer,
            int px, int py, int pz)
  {
    constexpr int RAD = degree;
    const int     n0 = a.n[0], n1 = a.n[1], n2 = a.n[2];
    const int     p0 = a.periodic[0], p1 = a.periodic[1], p2 = a.periodic[2];
    const int     xs = source_coord(px, n0, p0), ys = source_coord(py, n1, p1),
              zs = source_coord(pz, n2, p2);
    auto u = [&](int x, int y, int z) __attribute__((always_inline)) {
      return a.op_a[int64_t(source_coord(x, n0, p0)) +
                    int64_t(n0) * (int64_t(source_coord(y, n1, p1)) +
                                   int64_t(n1) * source_coord(z, n2, p2))];
    };
    number tx[RAD + 1], ty[RAD + 1], tz[RAD + 1];
    pair_taps<degree>(a, xs, ys, zs, tx, ty, tz);
    const number ctr = u(px, py, pz);
    number       acc = (tx[0] + (ty[0] + tz[0])) * ctr;
#pragma unroll
    for (int o = 1; o <= RAD; ++o)
      {
        acc = fma(tx[o], u(px - o, py, pz) + u(px + o, py, pz), acc);
        acc = fma(ty[o], u(px, py - o, pz) + u(px, py + o, pz), acc);
        acc = fma(tz[o], u(px, py, pz - o) + u(px, py, pz + o), acc);
      }
    const int64_t node = int64_t(xs) + int64_t(n0) * (int64_t(ys) + int64_t(n1) * zs);
    prismspf::FieldContainer<3, degree, number> fc;
    fc.q_point[0]     = number(a.q_lower[0] + xs * a.q_step[0]);
    fc.q_point[1]     = number(a.q_lower[1] + ys * a.q_step[1]);
    fc.q_point[2]     = number(a.q_lower[2] + zs * a.q_step[2]);
    fc.element_volume = number(a.element_volume);
#pragma unroll
    for (int f = 0; f < PPF_MAX_FIELDS; ++f)
#pragma unroll
      for (int s = 0; s < PPF_N_SLOTS; ++s)
        {
          number v = number(0);
          if (f < NF)
            v = a.alias_a[f][s] == 0 ? ctr : a.val_a[f][s][node];
          fc.value[f][s] = v;
#pragma unroll
          for (int d = 0; d < 3; ++d)
            fc.gradient[f][s][d] = number(0);
        }
    fc.clear_terms();
    call_hook<PDE, 3, degree, number, false, id_a>(pde, fc, timer);
    number out = pair_value_term<NF>(fc, a.fa);
    out = fma(a.beta_a, acc, out);
    return pair_scale<degree>(a, xs, ys, zs) * out;
  }

  /// One node of the remainder columns: A at the node (stored) and at its stencil
  /// neighbours, then B.
  template <class PDE, int degree, typename number, int id_a, int id_b, int NF>
  __device__ PPF_INLINE void
  pair_node(const PDE &pde, const PairArgs<number> &a, const prismspf::SimulationTimer &timer,
            int x, int y, int z)
  {
    constexpr int RAD = degree;
    auto A = [&](int px, int py, int pz) __attribute__((always_inline)) {
      return pair_a_at<PDE, degree, number, id_a, NF>(pde, a, timer, px, py, pz);
    };
    number tx[RAD + 1], ty[RAD + 1], tz[RAD + 1];
    pair_taps<degree>(a, x, y, z, tx, ty, tz);
    const number ctr = A(x, y, z);
    number       acc = (tx[0] + (ty[0] + tz[0])) * ctr;
    // Neighbour pairs in the tiles' order (offset 1: x, y, z; offset 2: x, y, z), one
    // pair per iteration (a rolled loop keeps one pair of A evaluations live); taps
    // chosen by selects (a runtime index would put arrays in scratch memory).
#pragma nounroll
    for (int t = 0; t < 3 * RAD; ++t)
      {
        const int    o = 1 + t / 3, d = t % 3;
        const int    dx = d == 0 ? o : 0, dy = d == 1 ? o : 0, dz = d == 2 ? o : 0;
        const number w1 = d == 0 ? tx[1] : (d == 1 ? ty[1] : tz[1]);
        const number w2 = d == 0 ? tx[RAD] : (d == 1 ? ty[RAD] : tz[RAD]);
        const number w  = o == 1 ? w1 : w2;
        acc = fma(w, A(x - dx, y - dy, z - dz) + A(x + dx, y + dy, z + dz), acc);
      }
    const int64_t node = int64_t(x) + int64_t(a.n[0]) * (int64_t(y) + int64_t(a.n[1]) * z);
    prismspf::FieldContainer<3, degree, number> fc;
    fc.q_point[0]     = number(a.q_lower[0] + x * a.q_step[0]);
    fc.q_point[1]     = number(a.q_lower[1] + y * a.q_step[1]);
    fc.q_point[2]     = number(a.q_lower[2] + z * a.q_step[2]);
    fc.element_volume = number(a.element_volume);
#pragma unroll
    for (int f = 0; f < PPF_MAX_FIELDS; ++f)
#pragma unroll
      for (int s = 0; s < PPF_N_SLOTS; ++s)
        {
          number v = number(0);
          if (f < NF)
            v = a.alias_b[f][s] == 0 ? ctr : a.val_b[f][s][node];
          fc.value[f][s] = v;
#pragma unroll
          for (int d = 0; d < 3; ++d)
            fc.gradient[f][s][d] = number(0);
        }
    fc.clear_terms();
    call_hook<PDE, 3, degree, number, false, id_b>(pde, fc, timer);
    number out = pair_value_term<NF>(fc, a.fb);
    out = fma(a.beta_b, acc, out);
    a.dst_a[node] = ctr;
    a.dst_b[node] = pair_scale<degree>(a, x, y, z) * out;
  }

  template <class PDE, int degree, typename number, int id_a, int id_b, int NF, int BR>
  __global__ void __launch_bounds__(stencil_tx * pair_waves)
  pair_kernel(const PDE pde, const PairArgs<number> a, const prismspf::SimulationTimer timer)
  {
    constexpr int RAD = degree;
    constexpr int MW  = pair_main_waves;
    constexpr int TY  = MW * BR;             // B rows of a tile
    constexpr int EA  = TY + 2 * RAD;        // A rows
    constexpr int AR  = (EA + MW - 1) / MW;  // A rows per main wave
    constexpr int MX  = stencil_tx + 4 * RAD; // staged operand plane (columns)
    constexpr int MY  = EA + 2 * RAD;         //                        (rows)
    constexpr int CX  = stencil_tx + 2 * RAD; // staged A-output plane (columns)
    constexpr int S   = 2 * RAD + 2;          // ring length; steps unrolled by S
    constexpr int NC  = RAD + 1;              // A-output planes kept in shared memory
    constexpr int H   = 2;                    // halo loads per thread
    static_assert(2 * RAD * EA <= stencil_tx, "side wave: one A node per lane");
    static_assert(2 * RAD * 4 * RAD <= stencil_tx, "side wave: one y-halo corner per lane");
    static_assert(2 * RAD <= MW, "y-halo rows: one per main wave");
    static_assert(S == 2 * NC, "A-output plane slots follow the step phase");
    __shared__ number mu[2][MY + 1][MX]; // row MY: spare target of idle lanes
    __shared__ number cn[NC][EA + 1][CX]; // row EA: spare

    const int lane = threadIdx.x;
    const int w    = __builtin_amdgcn_readfirstlane(threadIdx.y);
    const int n0 = a.n[0], n1 = a.n[1], n2 = a.n[2];
    const int p0 = a.periodic[0], p1 = a.periodic[1], p2 = a.periodic[2];
    if (halted(a.halt))
      return;

    const int b   = blockIdx.x;
    const int q8  = a.n_blocks / 8;
    const int lb  = b < 8 * q8 ? (b % 8) * q8 + b / 8 : b;
    const int t_x = lb % a.tiles_x;
    const int t_r = lb / a.tiles_x;
    const int t_y = t_r % a.tiles_y;
    const int chk = t_r / a.tiles_y;
    const int x0  = t_x * stencil_tx;
    const int y0  = t_y * TY;
    const int zb  = chk * a.z_chunk; // even for Q2 (host)
    const int ze  = min(zb + a.z_chunk, n2);
    const int zs0 = zb - 2 * RAD; // first step (A only until zb)
    const unsigned int plane_bytes = unsigned(n0) * unsigned(n1) * unsigned(sizeof(number));
    const bool         nt          = a.nt_store != 0;
    const bool         main        = w < MW;

    const buffer_rsrc r_op = make_rsrc(a.op_a);
    const buffer_rsrc r_da = make_rsrc(a.dst_a), r_db = make_rsrc(a.dst_b);
    buffer_rsrc       r_va[NF][PPF_N_SLOTS], r_vb[NF][PPF_N_SLOTS];
#pragma unroll
    for (int f = 0; f < NF; ++f)
#pragma unroll
      for (int s = 0; s < PPF_N_SLOTS; ++s)
        {
          r_va[f][s] = make_rsrc(a.val_a[f][s]);
          r_vb[f][s] = make_rsrc(a.val_b[f][s]);
        }

    // A nodes of this thread (k < AR; the side wave uses k = 0 only): source
    // coordinates, byte offset, shared-memory positions, ownership, B membership.
    // (my_k: shared-memory row written; mr_k: row read, a valid one for idle slots)
    int          xs_k[AR], ys_k[AR], mx_k[AR], my_k[AR], mr_k[AR], cx_k[AR], ey_k[AR];
    unsigned int off_k[AR];
    bool         own_a[AR], in_b[AR];
#pragma unroll
    for (int k = 0; k < AR; ++k)
      {
        int  px, e, mx, cx;
        bool valid;
        if (main)
          {
            e     = w + MW * k;
            valid = e < EA;
            px    = x0 + lane;
            mx    = 2 * RAD + lane;
            cx    = RAD + lane;
          }
        else
          {
            const int sc = lane % (2 * RAD);
            e            = lane / (2 * RAD);
            valid        = k == 0 && lane < 2 * RAD * EA;
            px           = sc < RAD ? x0 - RAD + sc : x0 + stencil_tx + (sc - RAD);
            mx           = sc < RAD ? RAD + sc : stencil_tx + RAD + sc;
            cx           = sc < RAD ? sc : stencil_tx + sc;
          }
        const int py = y0 - RAD + e;
        xs_k[k]      = source_coord(px, n0, p0);
        ys_k[k]      = source_coord(valid ? py : y0, n1, p1);
        off_k[k]     = (unsigned(xs_k[k]) + unsigned(n0) * unsigned(ys_k[k])) * unsigned(sizeof(number));
        mx_k[k]      = mx;
        my_k[k]      = valid ? RAD + e : MY;
        mr_k[k]      = valid ? RAD + e : RAD;
        cx_k[k]      = cx;
        ey_k[k]      = valid ? e : EA;
        const bool tile_row = e >= RAD && e < RAD + TY && py < n1;
        in_b[k]  = main && valid && e >= RAD && e < RAD + TY;
        own_a[k] = in_b[k] && tile_row && px < a.tile_n0 && px < n0;
      }
    // Taps and scale factors (x, y) of the A nodes; a node's factor is s_x (s_y s_z).
    number wx[AR][RAD + 1], wy[AR][RAD + 1], sx_k[AR], sy_k[AR];
#pragma unroll
    for (int k = 0; k < AR; ++k)
      {
        const bool ox = degree == 2 && (xs_k[k] & 1), oy = degree == 2 && (ys_k[k] & 1);
#pragma unroll
        for (int o = 0; o <= RAD; ++o)
          {
            wx[k][o] = ox ? a.tap[0][1][o] : a.tap[0][0][o];
            wy[k][o] = oy ? a.tap[1][1][o] : a.tap[1][0][o];
          }
        const int cxl = scale_class<degree>(xs_k[k], n0, p0);
        const int cyl = scale_class<degree>(ys_k[k], n1, p1);
        sx_k[k]       = cxl == 0 ? a.sfac[0][0] : (cxl == 1 ? a.sfac[0][1] : a.sfac[0][2]);
        sy_k[k]       = cyl == 0 ? a.sfac[1][0] : (cyl == 1 ? a.sfac[1][1] : a.sfac[1][2]);
      }

    // Halo loads of the operand (centre plane only): main waves w < 2p load one y-halo
    // row of the main columns; the side wave loads the outer x-halo columns of the A
    // rows and the y-halo rows of the side columns.
    unsigned int h_off[H];
    int          h_my[H], h_mx[H];
#pragma unroll
    for (int h = 0; h < H; ++h)
      {
        int  mx = 0, my = 0;
        bool on = false;
        if (main)
          {
            on = h == 0 && w < 2 * RAD;
            my = w < RAD ? w : EA + w;
            mx = 2 * RAD + lane;
          }
        else if (h == 0)
          {
            on           = lane < 2 * RAD * EA;
            const int sc = lane % (2 * RAD);
            mx           = sc < RAD ? sc : stencil_tx + 3 * RAD + (sc - RAD);
            my           = RAD + lane / (2 * RAD);
          }
        else
          {
            on           = lane < 2 * RAD * 4 * RAD;
            const int c4 = lane % (4 * RAD), r = lane / (4 * RAD);
            mx           = c4 < 2 * RAD ? c4 : stencil_tx + c4;
            my           = r < RAD ? r : EA + r;
          }
        const int hx = source_coord(x0 - 2 * RAD + (on ? mx : 2 * RAD), n0, p0);
        const int hy = source_coord(y0 - 2 * RAD + (on ? my : 2 * RAD), n1, p1);
        h_off[h]     = (unsigned(hx) + unsigned(n0) * unsigned(hy)) * unsigned(sizeof(number));
        h_mx[h]      = on ? mx : 0;
        h_my[h]      = on ? my : MY;
      }

    // Register rings.  Operand planes: slot of plane p = (p - zs0) mod S; A outputs:
    // slot of plane p = (p - zs0 + RAD) mod S; halo and value inputs: parity slots.
    number ring[AR][S];
    number cring[AR][S];
    number hring[H][2];
    number va[2][AR][NF][PPF_N_SLOTS];
    number vb[2][AR][NF][PPF_N_SLOTS];

    auto source_z = [&](int z) __attribute__((always_inline)) {
      return unsigned(z) < unsigned(n2) ? z : source_coord(z, n2, p2);
    };
    auto load_op = [&](int z, int slot) __attribute__((always_inline)) {
      const unsigned int po = unsigned(source_z(z)) * plane_bytes;
#pragma unroll
      for (int k = 0; k < AR; ++k)
        ring[k][slot] = buffer_load<number>(r_op, off_k[k], po);
    };
    auto load_inputs = [&](int z, int slot) __attribute__((always_inline)) { // A plane z
      const unsigned int po = unsigned(source_z(z)) * plane_bytes;
#pragma unroll
      for (int h = 0; h < H; ++h)
        hring[h][slot] = buffer_load<number>(r_op, h_off[h], po);
#pragma unroll
      for (int k = 0; k < AR; ++k)
#pragma unroll
        for (int f = 0; f < NF; ++f)
#pragma unroll
          for (int s = 0; s < PPF_N_SLOTS; ++s)
            if (a.alias_a[f][s] != 0) // uniform
              va[slot][k][f][s] = buffer_load<number>(r_va[f][s], off_k[k], po);
    };
    auto load_b_inputs = [&](int z, int slot) __attribute__((always_inline)) { // B plane z
      const unsigned int po = unsigned(source_z(z)) * plane_bytes;
#pragma unroll
      for (int k = 0; k < AR; ++k)
        if (in_b[k]) // uniform
#pragma unroll
          for (int f = 0; f < NF; ++f)
#pragma unroll
            for (int s = 0; s < PPF_N_SLOTS; ++s)
              if (a.alias_b[f][s] != 0) // uniform
                vb[slot][k][f][s] = buffer_load<number>(r_vb[f][s], off_k[k], po);
    };

    auto step = [&]<int PH>(std::integral_constant<int, PH>, int z) __attribute__((always_inline)) {
      constexpr int buf  = PH & 1;
      constexpr int za_r = (PH + RAD) % S;  // operand ring slot of the A plane
      constexpr int ca_r = (PH + 2 * RAD) % S; // A-output ring slot of the A plane
      constexpr int cn_w = (PH + RAD) % NC; // A-output plane written (A plane)
      constexpr int cn_r = PH % NC;         // A-output plane read (B plane)
      const int     za   = z + RAD;
      // Loads for the next step.
      load_op(min(z + 2 * RAD + 1, ze - 1 + 2 * RAD), (PH + S - 1) % S);
      load_inputs(min(za + 1, ze - 1 + RAD), (PH + 1) & 1);
      load_b_inputs(min(z + 1, ze - 1), (PH + 1) & 1);

      // Stage the operand's A plane.
#pragma unroll
      for (int k = 0; k < AR; ++k)
        mu[buf][my_k[k]][mx_k[k]] = ring[k][za_r];
#pragma unroll
      for (int h = 0; h < H; ++h)
        mu[buf][h_my[h]][h_mx[h]] = hring[h][PH & 1];
      __syncthreads();

      // Block A at plane za.
      const int  zas  = source_z(za);
      const bool oza = degree == 2 && (zas & 1);
      number     wz[RAD + 1];
#pragma unroll
      for (int o = 0; o <= RAD; ++o)
        wz[o] = oza ? a.tap[2][1][o] : a.tap[2][0][o];
      const int    czl  = scale_class<degree>(zas, n2, p2);
      const number sza  = czl == 0 ? a.sfac[2][0] : (czl == 1 ? a.sfac[2][1] : a.sfac[2][2]);
      const bool   a_st = za >= zb && za < ze;
      const unsigned int za_off = unsigned(za < n2 ? za : 0) * plane_bytes;
#pragma unroll
      for (int k = 0; k < AR; ++k)
        {
          const number(&M)[MY + 1][MX] = mu[buf];
          const int    my  = mr_k[k], mx = mx_k[k];
          const number ctr = ring[k][za_r];
          number       acc = (wx[k][0] + (wy[k][0] + wz[0])) * ctr;
#pragma unroll
          for (int o = 1; o <= RAD; ++o)
            {
              acc = fma(wx[k][o], M[my][mx - o] + M[my][mx + o], acc);
              acc = fma(wy[k][o], M[my - o][mx] + M[my + o][mx], acc);
              acc = fma(wz[o], ring[k][(za_r + S - o) % S] + ring[k][(za_r + o) % S], acc);
            }
          prismspf::FieldContainer<3, degree, number> fc;
          fc.q_point[0]     = number(a.q_lower[0] + xs_k[k] * a.q_step[0]);
          fc.q_point[1]     = number(a.q_lower[1] + ys_k[k] * a.q_step[1]);
          fc.q_point[2]     = number(a.q_lower[2] + zas * a.q_step[2]);
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
                    v            = a.alias_a[fi][s] == 0 ? ctr : va[PH & 1][k][fi][s];
                  }
                fc.value[f][s] = v;
#pragma unroll
                for (int d = 0; d < 3; ++d)
                  fc.gradient[f][s][d] = number(0);
              }
          fc.clear_terms();
          call_hook<PDE, 3, degree, number, false, id_a>(pde, fc, timer);
          number out = pair_value_term<NF>(fc, a.fa);
          out = fma(a.beta_a, acc, out);
          const number ua = (sx_k[k] * (sy_k[k] * sza)) * out;
          cring[k][ca_r]             = ua;
          cn[cn_w][ey_k[k]][cx_k[k]] = ua;
          if (own_a[k] && a_st)
            buffer_store(ua, r_da, off_k[k], za_off, nt);
        }

      // Block B at plane z (main waves, tile rows).
      if (z < zb)
        return;
      if (!main)
        return;
      const bool ozb = degree == 2 && (z & 1);
      number     wzb[RAD + 1];
#pragma unroll
      for (int o = 0; o <= RAD; ++o)
        wzb[o] = ozb ? a.tap[2][1][o] : a.tap[2][0][o];
      const int    czb  = scale_class<degree>(z, n2, p2);
      const number szb  = czb == 0 ? a.sfac[2][0] : (czb == 1 ? a.sfac[2][1] : a.sfac[2][2]);
      const unsigned int zoff = unsigned(z) * plane_bytes;
#pragma unroll
      for (int k = 0; k < AR; ++k)
        {
          if (!in_b[k])
            continue;
          const number(&C)[EA + 1][CX] = cn[cn_r];
          constexpr int cz  = (PH + RAD) % S; // A-output ring slot of plane z
          const int     ey  = ey_k[k], cx = cx_k[k];
          const number  ctr = cring[k][cz];
          number        acc = (wx[k][0] + (wy[k][0] + wzb[0])) * ctr;
#pragma unroll
          for (int o = 1; o <= RAD; ++o)
            {
              acc = fma(wx[k][o], C[ey][cx - o] + C[ey][cx + o], acc);
              acc = fma(wy[k][o], C[ey - o][cx] + C[ey + o][cx], acc);
              acc = fma(wzb[o], cring[k][(cz + S - o) % S] + cring[k][(cz + o) % S], acc);
            }
          prismspf::FieldContainer<3, degree, number> fc;
          fc.q_point[0]     = number(a.q_lower[0] + xs_k[k] * a.q_step[0]);
          fc.q_point[1]     = number(a.q_lower[1] + ys_k[k] * a.q_step[1]);
          fc.q_point[2]     = number(a.q_lower[2] + z * a.q_step[2]);
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
                    v            = a.alias_b[fi][s] == 0 ? ctr : vb[PH & 1][k][fi][s];
                  }
                fc.value[f][s] = v;
#pragma unroll
                for (int d = 0; d < 3; ++d)
                  fc.gradient[f][s][d] = number(0);
              }
          fc.clear_terms();
          call_hook<PDE, 3, degree, number, false, id_b>(pde, fc, timer);
          number out = pair_value_term<NF>(fc, a.fb);
          out = fma(a.beta_b, acc, out);
          if (own_a[k])
            {
              buffer_store((sx_k[k] * (sy_k[k] * szb)) * out, r_db, off_k[k], zoff, nt);
              // The thread that owns node 0 closes the increment's NaN probe after
              // storing both of its outputs there.
              if (a.probe_result && z == 0 && xs_k[k] == 0 && ys_k[k] == 0 && x0 == 0 &&
                  y0 == 0 && lane == 0)
                pair_probe(a);
            }
        }
    };

    // Prologue: operand planes zs0 .. zs0 + 2p (slots 0 .. 2p), halo and values of the
    // first A plane, values of the first B plane.
#pragma unroll
    for (int k = 0; k <= 2 * RAD; ++k)
      load_op(zs0 + k, k);
    load_inputs(zs0 + RAD, 0);
    load_b_inputs(zs0, 0);
    for (int z = zs0; z < ze; z += S)
      {
        [&]<int... PH>(std::integer_sequence<int, PH...>) {
          ((z + PH < ze ? step(std::integral_constant<int, PH> {}, z + PH) : void()), ...);
        }(std::make_integer_sequence<int, S> {});
      }
  }
  /// The remainder columns x >= tile_n0 of pair_kernel, one node per thread (a kernel
  /// of its own: its global-memory evaluation of A at 13 points would otherwise set the
  /// register and scratch budget of the tiles).
  template <class PDE, int degree, typename number, int id_a, int id_b, int NF>
  __global__ void __launch_bounds__(256)
  pair_node_kernel(const PDE pde, const PairArgs<number> a, const prismspf::SimulationTimer timer)
  {
    if (halted(a.halt))
      return;
    int64_t t = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (t >= a.rem)
      return;
    const int wr = a.n[0] - a.tile_n0;
    const int x  = a.tile_n0 + int(t % wr);
    t /= wr;
    pair_node<PDE, degree, number, id_a, id_b, NF>(pde, a, timer, x, int(t % a.n[1]),
                                                   int(t / a.n[1]));
  }
} // namespace ppf
