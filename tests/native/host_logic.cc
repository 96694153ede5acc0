// Host-logic checks of the shared C++ headers (dependency parsing, solve-block
// validation, parameter-file parsing).  Built by tests/test_host_logic.py with g++ and
// run on the CPU; prints one "name: value" line per check.
#include <prismspf/core/dependencies.h>
#include <prismspf/core/field_attributes.h>
#include <prismspf/core/solve_block.h>
#include <prismspf/user_inputs/user_input_parameters.h>

#include <iostream>
#include <string>
#include <vector>

using namespace prismspf;

template <class F>
std::string
error_of(F &&f)
{
  try
    {
      f();
    }
  catch (const std::exception &e)
    {
      return e.what();
    }
  return "";
}

int
main()
{
  std::vector<FieldAttributes> fields = {FieldAttributes("c"), FieldAttributes("v"),
                                         FieldAttributes("eta")};

  auto deps = make_dependency_set(fields, {"c", "hess(c)", "grad(old_1(c))", "old_3(eta)",
                                           "lhs(v)", "grad(lhs(v))", "grad(change(eta))"});
  std::cout << "c.flag: " << unsigned(deps[0].flag) << "\n";
  std::cout << "c.old1: " << unsigned(deps[0].old_flags.at(0)) << "\n";
  std::cout << "eta.old_size: " << deps[2].old_flags.size() << "\n";
  std::cout << "eta.old3: " << unsigned(deps[2].old_flags.at(2)) << "\n";
  std::cout << "v.src: " << unsigned(deps[1].src_flag) << "\n";
  std::cout << "eta.src: " << unsigned(deps[2].src_flag) << "\n";
  std::cout << "unknown: " << error_of([&] { make_dependency_set(fields, {"grad(q)", "c"}); })
            << "\n";

  SolveBlock exp;
  exp.id               = 3;
  exp.solve_type       = Explicit;
  exp.field_indices    = {0};
  exp.dependencies_rhs = make_dependency_set(fields, {"c"});
  std::cout << "own_current: " << error_of([&] { exp.validate(); }) << "\n";
  exp.dependencies_rhs = make_dependency_set(fields, {"old_1(c)"});
  exp.dependencies_lhs = make_dependency_set(fields, {"lhs(c)"});
  std::cout << "explicit_lhs: " << error_of([&] { exp.validate(); }) << "\n";
  SolveBlock lin;
  lin.id               = 4;
  lin.solve_type       = Linear;
  lin.field_indices    = {1};
  lin.dependencies_lhs = make_dependency_set(fields, {"c"});
  std::cout << "linear_no_src: " << error_of([&] { lin.validate(); }) << "\n";
  lin.dependencies_lhs = make_dependency_set(fields, {"c", "lhs(v)"});
  std::cout << "linear_ok: [" << error_of([&] { lin.validate(); }) << "]\n";

  const std::string prm = R"(
# comment line
subsection Rectangular mesh
  set x size = 12.5   # trailing comment
  set y size = 3.0
  set x subdivisions = 2
  set y subdivisions = 1
  set x periodic = true
end
set global refinement = 3
set time step = 0.25
set number steps = 10
subsection output
  set condition = EQUAL_SPACING
  set number = 3
end
set Model constant D0 = 0.7, DOUBLE
set Model constant n = 4, INT
subsection boundary conditions: 0
  set variables = c, v
  set conditions = Dirichlet, Natural, Natural, Dirichlet
end
subsection linear solver parameters: 0
  set solver_ids = 4
  set tolerance type = RMSEPerField
  set tolerance value = 1e-9
  set max iterations = 77
  set preconditioner = GMG
  set mg depth = 3
end
)";
  auto p = UserInputParameters<2>::from_text(prm);
  const auto &m = p.spatial_discretization.rectangular_mesh;
  std::cout << "size: " << m.size[0] << " " << m.size[1] << "\n";
  std::cout << "periodic: " << m.periodic_directions.count(0) << " "
            << m.periodic_directions.count(1) << "\n";
  std::cout << "refinement: " << p.spatial_discretization.global_refinement << "\n";
  std::cout << "n_cells: " << p.spatial_discretization.n_cells(0) << " "
            << p.spatial_discretization.n_cells(1) << "\n";
  std::cout << "dt: " << p.temporal_discretization.dt << "\n";
  std::cout << "n_increments: " << p.temporal_discretization.n_increments << "\n";
  std::cout << "outputs:";
  for (unsigned o : p.output_parameters.output_list)
    std::cout << " " << o;
  std::cout << "\n";
  std::cout << "D0: " << p.user_constants.get_double("D0") << "\n";
  std::cout << "n: " << p.user_constants.get_int("n") << "\n";
  std::cout << "const_type: "
            << error_of([&] { (void) p.user_constants.get_int("D0"); }) << "\n";
  const auto &bc = p.boundary_parameters.at("c").component_constraints.at(0);
  std::cout << "bc_c:";
  for (const auto &[b, cond] : bc)
    std::cout << " " << b << "=" << int(cond);
  std::cout << "\n";
  const auto &lp = p.linear_solve_parameters.at(4);
  std::cout << "lin: " << int(lp.tolerance_type) << " " << lp.tolerance << " "
            << lp.max_iterations << " " << int(lp.preconditioner) << " " << lp.mg_depth << "\n";
  std::cout << "bad_precond: "
            << error_of([&] {
                 UserInputParameters<2>::from_text(
                   "subsection linear solver parameters: 0\n set solver_ids = 1\n"
                   " set preconditioner = ILU\nend\n");
               })
            << "\n";
  std::cout << "unterminated: "
            << error_of([&] { UserInputParameters<2>::from_text("subsection output\n"); })
            << "\n";

  // Vector fields: component expansion and isotropic elastic constants.
  {
    std::vector<FieldAttributes> vf = {FieldAttributes("T"), FieldAttributes("u", Vector)};
    BoundaryConditionSet         bcs;
    bcs.component_constraints[1][0] = Condition::Dirichlet;
    vf[1].boundary_conditions       = bcs;
    SolveBlock vb;
    vb.id               = 1;
    vb.solve_type       = Linear;
    vb.field_indices    = {1};
    vb.dependencies_lhs = make_dependency_set(vf, {"grad(lhs(u))"});
    std::vector<SolveBlock> vbs = {vb};
    const auto              layout = expand_vector_fields(vf, vbs, 3);
    std::cout << "vec_names:";
    for (const auto &f : vf)
      std::cout << " " << f.name;
    std::cout << "\nvec_layout:";
    for (const auto &l : layout)
      std::cout << " " << l.declared << "." << l.component << (l.vector ? "v" : "s");
    std::cout << "\nvec_block:";
    for (auto f : vbs[0].field_indices)
      std::cout << " " << f << ":" << int(vbs[0].dependencies_lhs.at(f).src_flag);
    std::cout << "\nvec_bc: " << vf[1].boundary_conditions.component_constraints.count(0) << " "
              << int(vf[2].boundary_conditions.component_constraints.at(0).at(0)) << "\n";
    std::vector<FieldAttributes> bad = {FieldAttributes("u", Vector), FieldAttributes("T")};
    std::vector<SolveBlock>      none;
    std::cout << "vec_not_last: " << error_of([&] { expand_vector_fields(bad, none, 3); })
              << "\n";
    const auto q = UserInputParameters<3>::from_text(
      "set Model constant stiffness = (2.0,0.25), isotropic elastic constants\n"
      "set Model constant ortho = (1,2,3,4,5,6,7,8,9), orthotropic elastic constants\n");
    const auto C = q.user_constants.get_elasticity_tensor("stiffness");
    std::cout << "stiffness: " << C[0][0] << " " << C[0][1] << " " << C[3][3] << " "
              << C[0][3] << "\n";
    std::cout << "ortho: "
              << error_of([&] { (void) q.user_constants.get_elasticity_tensor("ortho"); })
              << "\n";
    const auto q2 = UserInputParameters<2>::from_text(
      "set Model constant s0 = (2.0,0.25), isotropic elastic constants\n"
      "set Model constant s1 = (2.0,0.25), isotropic elastic constants planestrain\n"
      "set Model constant s2 = (2.0,0.25), isotropic elastic constants planestress\n"
      "set Model constant s3 = (2.0,0.25), isotropic elastic constants planeflat\n");
    for (const char *nm : {"s0", "s1", "s2"})
      {
        const auto C2 = q2.user_constants.get_elasticity_tensor(nm);
        std::cout << "stiffness2d_" << nm << ": " << C2[0][0] << " " << C2[0][1] << " "
                  << C2[2][2] << "\n";
      }
    std::cout << "stiffness2d_bad: "
              << error_of([&] { (void) q2.user_constants.get_elasticity_tensor("s3"); })
              << "\n";
    const auto q1 = UserInputParameters<1>::from_text(
      "set Model constant s = (2.0,0.25), isotropic elastic constants\n");
    std::cout << "stiffness1d: " << q1.user_constants.get_elasticity_tensor("s")[0][0] << "\n";
  }
  return 0;
}
