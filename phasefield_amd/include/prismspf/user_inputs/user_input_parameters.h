// User input parameters read from a PRISMS-PF `.prm` file.
//
// The reference reads these through deal.II's ParameterHandler
// (reference: include/prismspf/user_inputs/user_input_parameters.h:39-101 and the
// declare/assign functions in src/user_inputs/*.cc).  This header implements the
// subset of the format the matrix-free path consumes, with the reference's defaults
// and key aliases:
//   * mesh: `Rectangular mesh` {x,y,z} size / lower bound / subdivisions / periodic,
//     `global refinement` (spatial_discretization.cc:212-278, 358-492);
//   * time: `time step`, `number steps` and aliases, `start time`, `end time`
//     (temporal_discretization.cc:35-107);
//   * `boundary conditions: k` {variables, conditions} (constraint_parameters.cc:11-106);
//   * `linear solver parameters: k` (solve_parameters.cc:10-340) including the
//     `Chebyshev`, `BiCGStab`, `Richardson` and `GMRES` subsections;
//   * `newton solver parameters: k` (solve_parameters.cc:352-470);
//   * `output` {condition, number, list} (io_parameters.cc:100-150);
//   * `random seed` (miscellaneous_parameters.cc:11-48);
//   * `Model constant name = value, DOUBLE|INT|BOOL|TENSOR` (user_constants.h:960-1061).
// Unknown keys are kept and ignored, so a full reference `.prm` file loads.
#pragma once

#include <prismspf/config.h>
#include <prismspf/core/field_attributes.h>
#include <prismspf/core/lane_types.h>
#include <prismspf/core/type_enums.h>
#include <prismspf/user_inputs/solve_parameters.h>

#include <algorithm>
#include <cmath>
#include <fstream>
#include <map>
#include <random>
#include <set>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

PRISMS_PF_BEGIN_NAMESPACE

namespace prm
{
  inline std::string
  trim(const std::string &s)
  {
    const auto b = s.find_first_not_of(" \t\r\n");
    if (b == std::string::npos)
      return "";
    const auto e = s.find_last_not_of(" \t\r\n");
    return s.substr(b, e - b + 1);
  }

  inline std::vector<std::string>
  split(const std::string &s, char sep)
  {
    std::vector<std::string> out;
    std::string              item;
    std::istringstream       is(s);
    while (std::getline(is, item, sep))
      {
        item = trim(item);
        if (!item.empty())
          out.push_back(item);
      }
    return out;
  }

  inline bool
  to_bool(const std::string &s)
  {
    std::string l = s;
    std::transform(l.begin(), l.end(), l.begin(), ::tolower);
    if (l == "true" || l == "1" || l == "yes" || l == "on")
      return true;
    if (l == "false" || l == "0" || l == "no" || l == "off")
      return false;
    throw std::invalid_argument("Not a boolean: " + s);
  }

  /// Flat view of a parameter file: "subsection/.../key" -> value.
  class File
  {
  public:
    File() = default;

    static File
    from_text(const std::string &text)
    {
      File                     f;
      std::vector<std::string> stack;
      std::istringstream       is(text);
      std::string              line;
      unsigned int             lineno = 0;
      while (std::getline(is, line))
        {
          ++lineno;
          const auto hash = line.find('#');
          if (hash != std::string::npos)
            line = line.substr(0, hash);
          line = trim(line);
          if (line.empty())
            continue;
          if (line.rfind("subsection", 0) == 0)
            stack.push_back(trim(line.substr(10)));
          else if (line == "end")
            {
              if (stack.empty())
                throw std::invalid_argument("Unbalanced 'end' on line " +
                                            std::to_string(lineno));
              stack.pop_back();
            }
          else if (line.rfind("set", 0) == 0)
            {
              const auto eq = line.find('=');
              if (eq == std::string::npos)
                throw std::invalid_argument("Missing '=' on line " + std::to_string(lineno));
              std::string key = trim(line.substr(3, eq - 3));
              std::string val = trim(line.substr(eq + 1));
              std::string path;
              for (const auto &s : stack)
                path += s + "/";
              f.entries[path + key] = val;
            }
          else
            throw std::invalid_argument("Cannot parse line " + std::to_string(lineno) +
                                        ": " + line);
        }
      if (!stack.empty())
        throw std::invalid_argument("Missing 'end' for subsection " + stack.back());
      return f;
    }

    /// First present key among `keys` (the name and its aliases), or `fallback`.
    [[nodiscard]] std::string
    get(const std::string &prefix, const std::vector<std::string> &keys,
        const std::string &fallback) const
    {
      for (const auto &k : keys)
        {
          const auto it = entries.find(prefix + k);
          if (it != entries.end())
            return it->second;
        }
      return fallback;
    }

    [[nodiscard]] bool
    has_prefix(const std::string &prefix) const
    {
      const auto it = entries.lower_bound(prefix);
      return it != entries.end() && it->first.rfind(prefix, 0) == 0;
    }

    std::map<std::string, std::string> entries;
  };
} // namespace prm

/// Rectangular mesh description (reference spatial_discretization.h:140-230).
template <unsigned int dim>
struct RectangularMesh
{
  enum Boundary
  {
    Left   = 0,
    Right  = 1,
    Top    = 2,
    Bottom = 3,
    Front  = 4,
    Back   = 5
  };

  dealii::Tensor<1, dim>     size;
  dealii::Tensor<1, dim>     lower_bound;
  std::vector<unsigned int>  subdivisions = std::vector<unsigned int>(dim, 1);
  std::set<unsigned int>     periodic_directions;
};

template <unsigned int dim>
struct SpatialDiscretization
{
  RectangularMesh<dim> rectangular_mesh;
  unsigned int         global_refinement = 0;
  bool                 has_adaptivity    = false;

  /// Cells per direction: subdivisions * 2^global_refinement.
  [[nodiscard]] unsigned int
  n_cells(unsigned int d) const
  {
    return rectangular_mesh.subdivisions[d] << global_refinement;
  }
};

struct TemporalDiscretization
{
  unsigned int n_increments = 0;
  double       dt           = 0.0;
  double       initial_time = 0.0;
};

struct OutputParameters
{
  std::set<unsigned int> output_list;
  /// File output (reference src/user_inputs/io_parameters.cc:10-110): `file type`
  /// vtu|vtk|pvtu|xdmf, `subdivisions` (0: the element degree), `directory` (alias
  /// `folder name`), `file name` (the prefix before the increment).
  std::string  file_type    = "vtu";
  unsigned int subdivisions = 0;
  std::string  directory    = "solutions";
  std::string  file_name    = "solution";

  [[nodiscard]] bool
  should_output(unsigned int increment) const
  {
    return output_list.count(increment) != 0;
  }
};

/// Random-number engine of the applications' initial conditions (reference
/// miscellaneous_parameters.h:22).
using RNGEngine = std::mt19937;

/// `random seed` and the engine seeded from it (reference miscellaneous_parameters.h:
/// 58-65, set_random_seed in miscellaneous_parameters.cc:44-48; one process, so the seed
/// is not offset by an MPI rank).
struct MiscellaneousParameters
{
  unsigned int      random_seed = 2025;
  mutable RNGEngine rng {2025};
};

/// `Model constant` values (reference user_constants.h).
template <unsigned int dim>
class UserConstants
{
public:
  [[nodiscard]] double
  get_double(const std::string &name) const
  {
    return std::stod(lookup(name, "DOUBLE"));
  }

  [[nodiscard]] int
  get_int(const std::string &name) const
  {
    return std::stoi(lookup(name, "INT"));
  }

  [[nodiscard]] bool
  get_bool(const std::string &name) const
  {
    return prm::to_bool(lookup(name, "BOOL"));
  }

  [[nodiscard]] dealii::Tensor<1, dim>
  get_rank_1_tensor(const std::string &name) const
  {
    std::string v = lookup(name, "TENSOR");
    v.erase(std::remove(v.begin(), v.end(), '('), v.end());
    v.erase(std::remove(v.begin(), v.end(), ')'), v.end());
    const auto             parts = prm::split(v, ',');
    dealii::Tensor<1, dim> t;
    for (unsigned int d = 0; d < dim && d < parts.size(); ++d)
      t[d] = std::stod(parts[d]);
    return t;
  }

  /// `(E, nu), isotropic elastic constants [planestress | planestrain]` as a Voigt
  /// stiffness matrix (reference user_constants.h construct_user_constant :432-480 and
  /// get_cij_matrix :581-700, isotropic case): shear modulus mu = E / (2 (1 + nu));
  /// lambda = nu E / ((1 + nu)(1 - 2 nu)) in 3-D and for 2-D plane strain (the 2-D
  /// default), nu E / (1 - nu^2) for 2-D plane stress; C = lambda on the normal
  /// off-diagonals, lambda + 2 mu on the normal diagonal, mu on the shear diagonal.  In
  /// 1-D, C = E.  Other symmetries are not supported.
  [[nodiscard]] dealii::Tensor<2, dim *(dim + 1) / 2>
  get_elasticity_tensor(const std::string &name) const
  {
    const auto it = values.find(name);
    if (it == values.end())
      throw std::invalid_argument("Model constant " + name + " is not set");
    std::string t = it->second.second;
    std::transform(t.begin(), t.end(), t.begin(), ::tolower);
    if (t.find("elastic constants") == std::string::npos)
      throw std::invalid_argument("Model constant " + name + " is not elastic constants");
    if (prm::trim(t.substr(0, t.find("elastic constants"))) != "isotropic")
      throw std::invalid_argument("Model constant " + name +
                                  ": only isotropic elastic constants are supported");
    // Optional 2-D stress state after "elastic constants" (plane strain by default).
    const std::string state =
      prm::trim(t.substr(t.find("elastic constants") + std::string("elastic constants").size()));
    bool plane_stress = false;
    if (dim == 2)
      {
        if (state == "planestress")
          plane_stress = true;
        else if (!state.empty() && state != "planestrain")
          throw std::invalid_argument(
            "Model constant " + name + ": Wrong 2D stress state. Use planestress or planestrain.");
      }
    std::string v = it->second.first;
    v.erase(std::remove(v.begin(), v.end(), '('), v.end());
    v.erase(std::remove(v.begin(), v.end(), ')'), v.end());
    const auto parts = prm::split(v, ',');
    if (parts.size() != 2)
      throw std::invalid_argument("Model constant " + name +
                                  ": isotropic elastic constants are (E, nu)");
    const double E = std::stod(parts[0]), nu = std::stod(parts[1]);
    const double mu     = E / (2.0 * (1.0 + nu));
    const double lambda = plane_stress ? nu * E / (1.0 - nu * nu)
                                       : nu * E / ((1.0 + nu) * (1.0 - 2.0 * nu));
    dealii::Tensor<2, dim *(dim + 1) / 2> C;
    if (dim == 1)
      {
        C[0][0] = E;
        return C;
      }
    for (unsigned int a = 0; a < dim * (dim + 1) / 2; ++a)
      for (unsigned int b = 0; b < dim * (dim + 1) / 2; ++b)
        C[a][b] = a < dim && b < dim ? (a == b ? lambda + 2.0 * mu : lambda)
                                     : (a == b ? mu : 0.0);
    return C;
  }

  void
  add(const std::string &name, const std::string &raw)
  {
    // "value, TYPE" or "(a,b,c), TENSOR": the type is after the last comma.
    const auto comma = raw.rfind(',');
    if (comma == std::string::npos)
      throw std::invalid_argument("Model constant " + name + " needs a type");
    values[name] = {prm::trim(raw.substr(0, comma)), prm::trim(raw.substr(comma + 1))};
  }

  [[nodiscard]] bool
  contains(const std::string &name) const
  {
    return values.count(name) != 0;
  }

private:
  [[nodiscard]] std::string
  lookup(const std::string &name, const std::string &type) const
  {
    const auto it = values.find(name);
    if (it == values.end())
      throw std::invalid_argument("Model constant " + name + " is not set");
    std::string t = it->second.second;
    std::transform(t.begin(), t.end(), t.begin(), ::toupper);
    if (t != type && !(type == "DOUBLE" && t == "INT"))
      throw std::invalid_argument("Model constant " + name + " is not of type " + type);
    return it->second.first;
  }

  std::map<std::string, std::pair<std::string, std::string>> values;
};

template <unsigned int dim>
class UserInputParameters
{
public:
  UserInputParameters() = default;

  /// Reads a parameter file from disk.
  explicit UserInputParameters(const std::string &filename)
  {
    std::ifstream in(filename);
    if (!in)
      throw std::invalid_argument("Cannot open parameter file " + filename);
    std::stringstream ss;
    ss << in.rdbuf();
    assign(prm::File::from_text(ss.str()));
  }

  static UserInputParameters
  from_text(const std::string &text)
  {
    UserInputParameters p;
    p.assign(prm::File::from_text(text));
    return p;
  }

  SpatialDiscretization<dim>                      spatial_discretization;
  TemporalDiscretization                          temporal_discretization;
  OutputParameters                                output_parameters;
  MiscellaneousParameters                         misc_parameters;
  UserConstants<dim>                              user_constants;
  std::map<std::string, BoundaryConditionSet>     boundary_parameters;
  std::map<int, LinearSolverParameters>           linear_solve_parameters;
  std::map<int, NonlinearSolverParameters>        nonlinear_solve_parameters;

private:
  void
  assign(const prm::File &f)
  {
    static const char *axes[3] = {"x", "y", "z"};
    auto              &mesh    = spatial_discretization.rectangular_mesh;
    for (unsigned int d = 0; d < dim; ++d)
      {
        const std::string a = axes[d];
        const std::string p = "Rectangular mesh/";
        mesh.size[d] = std::stod(f.get(p, {a + " size", "size " + a, a + " upper bound",
                                           "upper bound " + a},
                                       "0.0"));
        mesh.lower_bound[d] =
          std::stod(f.get(p, {a + " lower bound", "lower bound " + a}, "0.0"));
        mesh.subdivisions[d] =
          std::stoul(f.get(p, {a + " subdivisions", "subdivisions " + a}, "1"));
        if (prm::to_bool(f.get(p, {a + " periodic", "periodic " + a}, "false")))
          mesh.periodic_directions.insert(d);
      }
    spatial_discretization.global_refinement =
      std::stoul(f.get("", {"global refinement", "global_refinement"}, "0"));
    spatial_discretization.has_adaptivity =
      prm::to_bool(f.get("", {"mesh adaptivity", "mesh_adaptivity"}, "false"));

    auto &t        = temporal_discretization;
    t.n_increments = std::stoul(f.get(
      "",
      {"final increment", "number steps", "number of steps", "num steps", "n steps",
       "last increment", "end increment", "final_increment", "number_steps",
       "number_of_steps", "num_steps", "n_steps", "last_increment", "end_increment",
       "steps", "iterations", "final iteration", "max iteration", "increments"},
      "0"));
    t.dt           = std::stod(f.get("", {"time step", "timestep", "time_step", "dt"}, "0.0"));
    t.initial_time = std::stod(f.get("", {"start time", "start_time", "t_0", "t0",
                                          "begin time", "begin_time"},
                                     "0.0"));
    const double final_time = std::stod(
      f.get("", {"end time", "end_time", "t_f", "tf", "final time", "final_time"}, "0.0"));
    if (final_time > 0.0)
      t.n_increments =
        static_cast<unsigned int>(std::ceil((final_time - t.initial_time) / t.dt));

    // Output schedule (reference io_parameters.h:33-60 for EQUAL_SPACING).
    {
      const std::string cond     = f.get("output/", {"condition"}, "EQUAL_SPACING");
      const unsigned    n_out    = std::stoul(f.get("output/", {"number"}, "10"));
      const unsigned    n_inc    = t.n_increments;
      auto             &list     = output_parameters.output_list;
      if (cond == "EQUAL_SPACING")
        {
          if (n_inc == 0)
            list.insert(0);
          else if (n_out > 0)
            {
              const unsigned period = std::max(1u, n_inc / n_out);
              for (unsigned o = 0; o <= n_inc; o += period)
                list.insert(o);
            }
        }
      for (const auto &s : prm::split(f.get("output/", {"list"}, ""), ','))
        list.insert(std::stoul(s));
      output_parameters.file_type = f.get("output/", {"file type"}, "vtu");
      if (output_parameters.file_type != "vtu" && output_parameters.file_type != "vtk" &&
          output_parameters.file_type != "pvtu" && output_parameters.file_type != "xdmf")
        throw std::invalid_argument("Invalid output file type: " + output_parameters.file_type);
      output_parameters.subdivisions = std::stoul(f.get("output/", {"subdivisions"}, "0"));
      output_parameters.directory =
        f.get("output/", {"directory", "folder name"}, "solutions");
      output_parameters.file_name = f.get("output/", {"file name"}, "solution");
    }

    misc_parameters.random_seed =
      std::stoul(f.get("", {"random seed", "random_seed", "seed"}, "2025"));
    misc_parameters.rng.seed(misc_parameters.random_seed);

    for (const auto &[key, val] : f.entries)
      if (key.rfind("Model constant ", 0) == 0)
        user_constants.add(prm::trim(key.substr(15)), val);

    for (unsigned int k = 0; k < 64; ++k)
      {
        const std::string p = "boundary conditions: " + std::to_string(k) + "/";
        if (!f.has_prefix(p))
          continue;
        ComponentConditions cc;
        const auto conds = prm::split(f.get(p, {"conditions"}, ""), ',');
        for (unsigned int b = 0; b < conds.size(); ++b)
          cc[b] = condition_from_string(conds[b]);
        const bool time_dependent =
          prm::to_bool(f.get(p, {"time dependent", "time_dependent"}, "false"));
        for (std::string name : prm::split(f.get(p, {"variables"}, ""), ','))
          {
            std::set<unsigned int> comps;
            for (unsigned int c = 0; c < 3; ++c)
              if (name.size() > 2 && name.substr(name.size() - 2) == std::string(":") + axes[c])
                comps.insert(c);
            if (comps.empty())
              for (unsigned int c = 0; c < dim; ++c)
                comps.insert(c);
            else
              name = name.substr(0, name.size() - 2);
            for (unsigned int c : comps)
              if (c < dim)
                boundary_parameters[name].component_constraints[c] = cc;
            boundary_parameters[name].time_dependent = time_dependent;
          }
      }

    static const std::map<std::string, SolverToleranceType> tol_types = {
      {"AbsoluteResidual",   AbsoluteResidual  },
      {"RMSEPerField",       RMSEPerField      },
      {"IntegratedPerField", IntegratedPerField},
      {"RMSETotal",          RMSETotal         },
      {"IntegratedTotal",    IntegratedTotal   }
    };
    static const std::map<std::string, PreconditionerType> precond_types = {
      {"None", None}, {"none", None}, {"", None}, {"Chebyshev", Chebyshev},
      {"chebyshev", Chebyshev}, {"GMG", GMG}, {"gmg", GMG}, {"MG", GMG}, {"mg", GMG},
      {"multigrid", GMG}
    };
    for (unsigned int k = 0; k < 64; ++k)
      {
        const std::string p = "linear solver parameters: " + std::to_string(k) + "/";
        if (!f.has_prefix(p))
          continue;
        LinearSolverParameters lp;
        // Inside a prm subsection the reference's declared default applies.
        lp.tolerance_type = AbsoluteResidual;
        lp.solver_type    = f.get(p, {"solver type", "solver_type", "linear solver type",
                                      "linear_solver_type", "linear solver", "linear_solver",
                                      "type"},
                                  "cg");
        const std::string tt = f.get(p, {"tolerance type"}, "AbsoluteResidual");
        if (!tol_types.count(tt))
          throw std::invalid_argument("Invalid tolerance type: " + tt);
        lp.tolerance_type = tol_types.at(tt);
        lp.tolerance      = std::stod(f.get(p, {"tolerance value", "tolerance"}, "1.0e-10"));
        lp.max_iterations = std::stoul(f.get(p, {"max iterations"}, "100"));
        lp.mg_depth       = std::stoul(f.get(p, {"mg depth", "mg_depth"}, "1"));
        const std::string pt =
          f.get(p, {"preconditioner type", "preconditioner_type", "preconditioner"}, "None");
        if (!precond_types.count(pt))
          throw std::invalid_argument("Invalid preconditioner type: " + pt);
        lp.preconditioner = precond_types.at(pt);
        const std::string c = p + "Chebyshev/";
        lp.chebyshev_parameters.degree =
          std::stoul(f.get(c, {"smoother degree"}, "5"));
        lp.chebyshev_parameters.smoothing_range =
          std::stod(f.get(c, {"smoothing range"}, "15.0"));
        lp.chebyshev_parameters.eig_cg_n_iterations =
          std::stoul(f.get(c, {"eigenvalue cg iterations"}, "10"));
        const std::string bs = p + "BiCGStab/";
        lp.bicgstab_exact_residual = prm::to_bool(f.get(bs, {"exact residual"}, "true"));
        lp.bicgstab_breakdown      = std::stod(f.get(bs, {"breakdown"}, "0.000000"));
        const std::string rs = p + "Richardson/";
        lp.richardson_omega  = std::stod(f.get(rs, {"omega"}, "1.0"));
        lp.richardson_use_preconditioned_residual =
          prm::to_bool(f.get(rs, {"use preconditioned residual"}, "false"));
        const std::string gm = p + "GMRES/";
        lp.gmres_max_basis_size = std::stoul(f.get(gm, {"max basis size"}, "30"));
        if (lp.gmres_max_basis_size < 1)
          throw std::invalid_argument("GMRES max basis size must be >= 1");
        lp.gmres_right_preconditioning =
          prm::to_bool(f.get(gm, {"right preconditioning"}, "false"));
        lp.gmres_use_default_residual =
          prm::to_bool(f.get(gm, {"use default residual"}, "true"));
        lp.gmres_orthogonalization =
          f.get(gm, {"orthogonalization strategy"}, "delayed_classical_gram_schmidt");
        if (lp.gmres_orthogonalization != "classical_gram_schmidt" &&
            lp.gmres_orthogonalization != "modified_gram_schmidt" &&
            lp.gmres_orthogonalization != "delayed_classical_gram_schmidt")
          throw std::invalid_argument("Invalid orthogonalization strategy: " +
                                      lp.gmres_orthogonalization);
        static const std::set<std::string> solver_names = {
          "richardson", "cg", "bicgstab", "gmres", "fgmres", "minres"};
        if (!solver_names.count(lp.solver_type))
          throw std::invalid_argument("Invalid solver type: " + lp.solver_type);
        // MI355X-path settings (no reference counterpart).
        const std::string lvl = f.get(p, {"mg level precision"}, "double");
        lp.level_precision =
          (lvl == "single" || lvl == "float" || lvl == "fp32") ? LevelPrecision::Single
                                                                 : LevelPrecision::Double;
        const std::string coarse = f.get(p, {"mg coarse solver"}, "smoother");
        lp.coarse_solver = coarse == "cg" ? CoarseSolver::CG : CoarseSolver::Smoother;
        lp.coarse_cg_tolerance = std::stod(f.get(p, {"mg coarse tolerance"}, "1.0e-12"));
        for (const auto &id : prm::split(f.get(p, {"solver_ids", "solver ids"}, ""), ','))
          linear_solve_parameters[std::stoi(id)] = lp;
      }
    // `newton solver parameters: k` (reference solve_parameters.cc:352-470).
    for (unsigned int k = 0; k < 64; ++k)
      {
        const std::string p = "newton solver parameters: " + std::to_string(k) + "/";
        if (!f.has_prefix(p))
          continue;
        NonlinearSolverParameters np;
        np.max_iterations = std::stoul(f.get(p, {"max iterations"}, "100"));
        if (np.max_iterations < 1)
          throw std::invalid_argument("newton solver parameters: max iterations must be >= 1");
        np.tolerance_value = std::stod(f.get(p, {"tolerance value", "tolerance"}, "1.0e-10"));
        np.step_length     = std::stod(f.get(p, {"step size"}, "1.0"));
        if (np.tolerance_value < 0.0 || np.step_length < 0.0)
          throw std::invalid_argument(
            "newton solver parameters: tolerance value and step size must be >= 0");
        for (const auto &id : prm::split(f.get(p, {"solver_ids", "solve blocks", "solve_blocks",
                                                   "solve block ids", "solve_block_ids",
                                                   "solver ids"},
                                               ""),
                                         ','))
          nonlinear_solve_parameters[std::stoi(id)] = np;
      }
  }
};

PRISMS_PF_END_NAMESPACE
