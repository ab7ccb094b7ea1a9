// driver.cpp -- one case on one GPU, writing timings.json (run_serial_case,
// src/driver.cpp:150-310 of the reference).
//
// Phases and their boundaries are the reference's: read (case file + mesh), assembly (pattern +
// values), bcs (tractions + Dirichlet elimination), solve (setup + solve + true residual),
// output (stresses, Kirsch metrics, timings.json), total.  The matrix, RHS and solution stay in
// HBM throughout.  timings.json keeps every field of the serial path (driver.cpp:235-297, same
// number format, json.hpp:11-16) and adds a "gpu" object; VTU and CSV files are not written.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <filesystem>
#include <fstream>
#include <optional>
#include <set>
#include <sstream>

#include "cstfem_hip_io.hpp"

namespace cstfem_hip {

namespace {

using Clock = std::chrono::steady_clock;

double since(Clock::time_point t0) {
  return std::chrono::duration<double>(Clock::now() - t0).count();
}

// json.hpp:11-27
std::string num(double v) {
  if (!std::isfinite(v)) return "null";
  char buf[32];
  std::snprintf(buf, sizeof buf, "%.17g", v);
  return buf;
}

std::string str(const std::string& s) {
  std::string o = "\"";
  for (char c : s) {
    if (c == '"' || c == '\\') o += '\\';
    o += c;
  }
  return o + "\"";
}

std::string trim(const std::string& s) {
  const auto a = s.find_first_not_of(" \t");
  if (a == std::string::npos) return {};
  const auto b = s.find_last_not_of(" \t");
  return s.substr(a, b - a + 1);
}

std::vector<std::string> split(const std::string& s, char sep) {
  std::vector<std::string> out;
  std::string item;
  std::istringstream ss(s);
  while (std::getline(ss, item, sep)) out.push_back(trim(item));
  return out;
}

double number(const std::string& s, const std::string& context) {
  const char* b = s.c_str();
  char* end = nullptr;
  const double v = std::strtod(b, &end);
  if (s.empty() || end != b + s.size() || !std::isfinite(v))
    throw std::runtime_error(context + ": '" + s + "' is not a number");
  return v;
}

// "group:key=value,key=value ; group:key=value" (bc.cpp:41-70)
std::vector<std::pair<std::string, std::vector<std::pair<std::string, double>>>> group_list(
    const std::string& text, const char* what) {
  std::vector<std::pair<std::string, std::vector<std::pair<std::string, double>>>> out;
  for (const auto& item : split(text, ';')) {
    if (item.empty()) continue;
    const auto colon = item.find(':');
    if (colon == std::string::npos)
      throw std::runtime_error(std::string(what) + ": expected 'group:key=value' in '" + item + "'");
    const std::string group = trim(item.substr(0, colon));
    if (group.empty()) throw std::runtime_error(std::string(what) + ": empty group in '" + item + "'");
    std::vector<std::pair<std::string, double>> kv;
    for (const auto& a : split(item.substr(colon + 1), ',')) {
      const auto eq = a.find('=');
      if (eq == std::string::npos)
        throw std::runtime_error(std::string(what) + ": expected 'key=value' in '" + item + "'");
      const std::string key = trim(a.substr(0, eq));
      kv.emplace_back(key, number(trim(a.substr(eq + 1)), std::string(what) + " '" + group + ":" + key + "'"));
    }
    if (kv.empty()) throw std::runtime_error(std::string(what) + ": no values for '" + group + "'");
    out.emplace_back(group, std::move(kv));
  }
  return out;
}

const std::vector<std::string>& known_keys() {  // driver.cpp:129-136 plus the gpu.* keys
  static const std::vector<std::string> k = {
      "mesh", "plane", "E", "nu", "thickness", "partitioner", "solver", "rtol", "max_iter",
      "output_dir", "bc.dirichlet", "bc.traction", "kirsch.sigma_inf", "kirsch.a", "path",
      "instrument", "petsc.near_nullspace", "petsc.options", "petsc.log_view", "gpu.method",
      "gpu.mpnn_params"};
  return k;
}

void check_rc(int rc) { check(rc); }

std::vector<int> edge_nodes(const Mesh& m, const std::string& group) {
  const auto it = m.boundary_edges.find(group);
  if (it == m.boundary_edges.end()) throw std::runtime_error("unknown boundary group '" + group + "'");
  std::vector<int> e;
  for (const auto& ed : it->second) {
    e.push_back(ed[0]);
    e.push_back(ed[1]);
  }
  return e;
}

}  // namespace

int run_case(const std::string& case_path, std::ostream& out, std::ostream& err, int device) {
  try {
    const auto t_all0 = Clock::now();
    // ---- read
    auto t0 = Clock::now();
    const CaseKeys c = CaseKeys::read(case_path);
    c.report_unknown(known_keys(), err);
    const std::string mesh_path = c.text("mesh");
    const std::string out_dir = c.text("output_dir", "results");
    const std::string plane_name = c.text("plane");
    int plane;
    if (plane_name == "plane_stress")
      plane = CSTFEM_PLANE_STRESS;
    else if (plane_name == "plane_strain")
      plane = CSTFEM_PLANE_STRAIN;
    else
      throw std::runtime_error(c.source() + ": plane must be plane_stress or plane_strain, got '" +
                               plane_name + "'");
    const double E = c.number("E"), nu = c.number("nu"), t = c.number("thickness");
    if (!(E > 0)) throw std::runtime_error(c.source() + ": E must be positive");
    if (!(nu > -1.0 && nu < 0.5)) throw std::runtime_error(c.source() + ": nu must be in (-1, 0.5)");
    if (!(t > 0)) throw std::runtime_error(c.source() + ": thickness must be positive");
    const std::string solver = c.text("solver", "cg_jacobi");
    if (solver != "cg_jacobi")
      throw std::runtime_error(c.source() + ": solver must be cg_jacobi on the GPU path");
    const std::string path = c.text("path", "auto");
    if (path != "auto" && path != "serial")
      throw std::runtime_error(c.source() + ": path must be auto or serial on the GPU path");
    const std::string instrument = c.text("instrument", "true");
    if (instrument != "true" && instrument != "false")
      throw std::runtime_error(c.source() + ": instrument must be true or false");
    const double rtol = c.number("rtol", 1e-10);
    const int max_iter = c.integer("max_iter", 20000);
    const std::string gmethod = c.text("gpu.method", "cg_jacobi");
    const std::map<std::string, int> methods = {
        {"cg_jacobi", CSTFEM_METHOD_CG_FP64}, {"cg_fp64_refined", CSTFEM_METHOD_CG_FP64_REFINED},
        {"ir_fp32", CSTFEM_METHOD_IR_FP32}, {"fcg_multilevel", CSTFEM_METHOD_FCG_LEARNED},
        {"fcg_learned_multilevel", CSTFEM_METHOD_FCG_LEARNED}};
    if (!methods.count(gmethod)) throw std::runtime_error(c.source() + ": unknown gpu.method '" + gmethod + "'");
    std::vector<float> theta;
    if (gmethod == "fcg_learned_multilevel") {
      const std::string pp = c.text("gpu.mpnn_params");
      std::ifstream pf(pp, std::ios::binary);
      if (!pf) throw std::runtime_error(pp + ": cannot open file");
      theta.resize(cstfem_mpnn_num_params());
      pf.read(reinterpret_cast<char*>(theta.data()), (std::streamsize)(theta.size() * sizeof(float)));
      if (!pf) throw std::runtime_error(pp + ": expected " + std::to_string(theta.size()) + " float32 values");
    }
    const Mesh mesh = read_gmsh(mesh_path);
    const double t_read = since(t0);
    out << "mesh " << mesh_path << ": " << mesh.num_nodes() << " nodes, " << mesh.num_triangles()
        << " triangles, " << 2 * mesh.num_nodes() << " DOFs\n";

    const auto tg0 = Clock::now();
    Device dev(device);
    const double t_gpu_init = since(tg0);
    cstfem_ctx* ctx = dev.get();
    char dname[256] = {0};
    int cus = 0;
    check_rc(cstfem_device_info(ctx, dname, sizeof dname, &cus));

    // ---- assembly
    t0 = Clock::now();
    const int nn = (int)mesh.num_nodes();
    const int n = 2 * nn;
    check_rc(cstfem_upload_mesh(ctx, nn, mesh.xy.data(), (int64_t)mesh.num_triangles(), mesh.triangles.data()));
    check_rc(cstfem_set_material(ctx, E, nu, plane, t, nullptr));
    int64_t nnz = 0;
    check_rc(cstfem_build_pattern(ctx, &nnz));
    // the multilevel set-up's fine aggregation needs the pattern only: start it now
    if (gmethod == "fcg_learned_multilevel" || gmethod == "fcg_multilevel") check_rc(cstfem_mg_prepare(ctx));
    check_rc(cstfem_assemble(ctx));
    const double t_assembly = since(t0);

    // ---- boundary conditions
    t0 = Clock::now();
    check_rc(cstfem_set_rhs(ctx, nullptr, n));  // f = 0, then the loads
    const std::string traction = c.text("bc.traction", "");
    std::optional<std::pair<double, double>> kirsch;  // (sigma_inf, a)
    if (traction == "kirsch_exact") {
      const double s_inf = c.number("kirsch.sigma_inf"), a = c.number("kirsch.a");
      if (!(a > 0)) throw std::runtime_error(c.source() + ": kirsch.a must be positive");
      kirsch = std::make_pair(s_inf, a);
      for (const char* g : {"right", "top"}) {
        const auto e = edge_nodes(mesh, g);
        check_rc(cstfem_add_traction(ctx, (int)e.size() / 2, e.data(), 1, s_inf, a));
      }
    } else {
      for (const auto& [group, kv] : group_list(traction, "bc.traction")) {
        double tx = 0.0, ty = 0.0;
        std::set<std::string> seen;
        for (const auto& [key, value] : kv) {
          if (key != "tx" && key != "ty")
            throw std::runtime_error("bc.traction: unknown component '" + key + "' for group '" +
                                     group + "' (use tx or ty)");
          if (!seen.insert(key).second)
            throw std::runtime_error("bc.traction: '" + key + "' given twice for group '" + group + "'");
          (key == "tx" ? tx : ty) = value;
        }
        const auto e = edge_nodes(mesh, group);
        check_rc(cstfem_add_traction(ctx, (int)e.size() / 2, e.data(), 0, tx, ty));
      }
    }
    DirichletDofs dofs;
    for (const auto& [group, kv] : group_list(c.text("bc.dirichlet", ""), "bc.dirichlet")) {
      std::set<std::string> seen;
      for (const auto& [key, value] : kv) {
        if (key != "ux" && key != "uy")
          throw std::runtime_error("bc.dirichlet: unknown component '" + key + "' for group '" +
                                   group + "' (use ux or uy)");
        if (!seen.insert(key).second)
          throw std::runtime_error("bc.dirichlet: '" + key + "' given twice for group '" + group + "'");
      }
      for (const auto& [key, value] : kv) add_dirichlet_group(mesh, group, key == "ux" ? 0 : 1, value, dofs);
    }
    std::vector<int> dd;
    std::vector<double> dv;
    for (const auto& [dof, v] : dofs.values()) {
      dd.push_back(dof);
      dv.push_back(v);
    }
    check_rc(cstfem_apply_dirichlet(ctx, (int)dd.size(), dd.data(), dv.data()));
    const double t_bcs = since(t0);

    // ---- solve
    t0 = Clock::now();
    check_rc(cstfem_setup(ctx));
    if (gmethod == "fcg_learned_multilevel") {
      check_rc(cstfem_mpnn_set_params(ctx, theta.data(), (int)theta.size()));
      check_rc(cstfem_mpnn_build(ctx));
      check_rc(cstfem_mg_setup(ctx, 12, 500, 0, 3));
    } else if (gmethod == "fcg_multilevel") {
      check_rc(cstfem_mg_setup(ctx, 12, 500, 1, 3));
    }
    cstfem_solver_options o{};
    o.rtol = rtol;
    o.max_iter = max_iter;
    o.instrument = instrument == "true" ? 1 : 0;
    o.method = methods.at(gmethod);
    cstfem_solver_stats s{};
    check_rc(cstfem_solve(ctx, &o, &s, nullptr, nullptr));
    const SolverStats st = make_stats(s, rtol, max_iter);
    const double t_solve = since(t0);
    out << "cg_jacobi: " << st.message << "; true relative residual " << st.true_rel_residual << "\n";
    if (!st.warning.empty()) err << st.warning << "\n";
    if (!st.converged) err << "error: solver did not converge: " << st.message << "\n";

    // ---- stresses, metrics, timings.json
    t0 = Clock::now();
    std::filesystem::create_directories(out_dir);
    const std::size_t ne = mesh.num_triangles();
    std::vector<double> sig(3 * ne), vm(ne);
    check_rc(cstfem_element_stresses(ctx, sig.data(), vm.data(), nullptr, nullptr, nullptr));
    cstfem_kirsch_result km{};
    if (kirsch) {
      check_rc(cstfem_kirsch_metrics(ctx, kirsch->first, kirsch->second, &km));
      out << "Kirsch: relative L2 errors: stress (centroid) " << km.l2_rel_stress_error
          << ", displacement " << km.l2_rel_displacement_error << "; energy norm "
          << km.energy_rel_error << "; SCF at centroid (" << km.scf_x << ", " << km.scf_y
          << "): numerical " << km.scf_numerical << ", analytical " << km.scf_analytical_centroid
          << ", theoretical 3\n";
    }
    const std::filesystem::path dir(out_dir);
    std::ofstream js(dir / "timings.json");
    if (!js) throw std::runtime_error((dir / "timings.json").string() + ": cannot open");
    const double t_output_so_far = since(t0);
    const double t_total = since(t_all0);
    js << "{\n"
       << "  \"case\": " << str(case_path) << ",\n"
       << "  \"mesh\": " << str(mesh_path) << ",\n"
       << "  \"ranks\": 1,\n"
       << "  \"path\": \"serial\",\n"
       << "  \"plane\": " << str(plane_name) << ",\n"
       << "  \"nodes\": " << mesh.num_nodes() << ",\n"
       << "  \"elements\": " << ne << ",\n"
       << "  \"dofs\": " << n << ",\n"
       << "  \"nnz\": " << nnz << ",\n"
       << "  \"solver\": {\n"
       << "    \"name\": \"cg_jacobi\",\n"
       << "    \"rtol\": " << num(rtol) << ",\n"
       << "    \"max_iter\": " << max_iter << ",\n"
       << "    \"iterations\": " << st.iterations << ",\n"
       << "    \"rel_residual\": " << num(st.rel_residual) << ",\n"
       << "    \"true_rel_residual\": " << num(st.true_rel_residual) << ",\n"
       << "    \"converged\": " << (st.converged ? "true" : "false") << "\n"
       << "  },\n"
       << "  \"solve_breakdown\": ";
    if (o.instrument) {
      // one rank: min = max = mean; the device has no halo or allreduce (zero)
      const double accounted = s.setup + s.spmv_local + s.dot_local + s.vector_ops;
      const double per_spmv = s.spmv_calls > 0 ? 1.0 / (double)s.spmv_calls : 0.0;
      const std::pair<const char*, double> vals[] = {
          {"setup", s.setup}, {"spmv_local", s.spmv_local}, {"halo_post", 0.0},
          {"halo_wait", 0.0}, {"spmv_ghost", 0.0}, {"dot_local", s.dot_local},
          {"allreduce", 0.0}, {"vector_ops", s.vector_ops}, {"unaccounted", s.total - accounted},
          {"total", s.total},
          {"time_per_iteration", s.iterations > 0 ? s.total / s.iterations : 0.0},
          {"accounted_fraction", s.total > 0 ? accounted / s.total : 0.0},
          {"halo_bytes_sent_per_iteration", 0.0 * per_spmv},
          {"halo_messages_sent_per_iteration", 0.0 * per_spmv}};
      js << "{\n";
      const std::size_t nv = sizeof vals / sizeof vals[0];
      for (std::size_t i = 0; i < nv; ++i) {
        const std::string v = num(vals[i].second);
        js << "    " << str(vals[i].first) << ": {\"min\": " << v << ", \"max\": " << v
           << ", \"mean\": " << v << "}" << (i + 1 < nv ? "," : "") << "\n";
      }
      js << "  },\n";
    } else {
      js << "null,\n";
    }
    js << "  \"timings\": {\n"
       << "    \"read\": " << num(t_read) << ",\n"
       << "    \"assembly\": " << num(t_assembly) << ",\n"
       << "    \"bcs\": " << num(t_bcs) << ",\n"
       << "    \"solve\": " << num(t_solve) << ",\n"
       << "    \"output\": " << num(t_output_so_far) << ",\n"
       << "    \"total\": " << num(t_total) << "\n"
       << "  },\n"
       << "  \"kirsch\": ";
    if (kirsch) {
      js << "{\n"
         << "    \"sigma_inf\": " << num(kirsch->first) << ",\n"
         << "    \"a\": " << num(kirsch->second) << ",\n"
         << "    \"l2_rel_stress_error\": " << num(km.l2_rel_stress_error) << ",\n"
         << "    \"l2_rel_displacement_error\": " << num(km.l2_rel_displacement_error) << ",\n"
         << "    \"energy_rel_error\": " << num(km.energy_rel_error) << ",\n"
         << "    \"scf\": {\n"
         << "      \"element\": " << km.scf_element << ",\n"
         << "      \"x\": " << num(km.scf_x) << ",\n"
         << "      \"y\": " << num(km.scf_y) << ",\n"
         << "      \"numerical\": " << num(km.scf_numerical) << ",\n"
         << "      \"analytical_at_centroid\": " << num(km.scf_analytical_centroid) << ",\n"
         << "      \"theoretical\": " << num(km.scf_theoretical) << "\n"
         << "    }\n"
         << "  },\n";
    } else {
      js << "null,\n";
    }
    js << "  \"gpu\": {\n"
       << "    \"device\": " << str(dname) << ",\n"
       << "    \"compute_units\": " << cus << ",\n"
       << "    \"method\": " << str(gmethod) << ",\n"
       << "    \"outer_iterations\": " << s.outer_iterations << ",\n"
       << "    \"fallback\": " << (s.fallback == 0 ? "null" : s.fallback == 1 ? "\"jacobi_cg\"" : "\"multilevel_chebyshev\"") << ",\n"
       << "    \"solve_device_s\": " << num(s.total) << ",\n"
       << "    \"context_init_s\": " << num(t_gpu_init) << ",\n"
       << "    \"max_von_mises\": " << num(*std::max_element(vm.begin(), vm.end())) << "\n"
       << "  }\n"
       << "}\n";
    if (!js) throw std::runtime_error((dir / "timings.json").string() + ": write failed");
    out << "wrote " << (dir / "timings.json").string() << "\n"
        << "time: read " << t_read << " s, assembly " << t_assembly << " s, bcs " << t_bcs
        << " s, solve " << t_solve << " s, output " << since(t0) << " s, total " << since(t_all0)
        << " s\n";
    return st.converged ? 0 : 2;
  } catch (const std::exception& e) {
    err << "error: " << e.what() << "\n";
    return 1;
  }
}

}  // namespace cstfem_hip
