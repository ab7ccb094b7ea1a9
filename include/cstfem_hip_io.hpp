// cstfem_hip_io.hpp -- host-side input and orchestration around the C ABI, mirroring the
// reference's interfaces (namespace cstfem):
//
//   reference (file:line)                         here
//   read_gmsh            mesh.hpp:33-37           cstfem_hip::read_gmsh (Gmsh 4.1 ASCII)
//   CaseFile             case.hpp:13-44           cstfem_hip::CaseKeys (same key = value format)
//   run_serial_case      driver.hpp:60-62         cstfem_hip::run_case (one GPU; writes
//                        driver.cpp:150-310       timings.json with the reference's fields)
//
// Errors are std::runtime_error with the reference's message shapes ("file:line: message").
#pragma once

#include <istream>
#include <ostream>
#include <string>
#include <vector>

#include "cstfem_hip.hpp"

namespace cstfem_hip {

// Reads a Gmsh .msh 4.1 ASCII file (replaces read_gmsh, mesh.hpp:33-37): nodes in file order
// (0-based index = position), 3-node triangles reoriented counter-clockwise, 2-node lines
// collected per physical group name; the same files are accepted and rejected as by the
// reference's reader (mesh_io.cpp:114-356), with its messages.
Mesh read_gmsh(const std::string& path);
Mesh read_gmsh(std::istream& in, const std::string& source_name);

// Writes `mesh` as Gmsh 4.1 ASCII: node tags 1..nn in index order (so a file written here and
// read by either reader numbers the nodes identically), one surface entity holding every
// triangle, one curve entity per boundary group.
void write_gmsh(const Mesh& mesh, std::ostream& out);

// The settings of a case file (the reference's cases/*.case format): "key = value" lines,
// '#' starts a comment, each key at most once.  Typed lookups name the file and line of a
// value that does not parse.
class CaseKeys {
 public:
  static CaseKeys read(const std::string& path);
  static CaseKeys read(std::istream& in, const std::string& source_name);
  bool has(const std::string& key) const { return find(key) != nullptr; }
  std::string text(const std::string& key) const;  // required key
  std::string text(const std::string& key, const std::string& fallback) const;
  double number(const std::string& key) const;  // required key
  double number(const std::string& key, double fallback) const;
  int integer(const std::string& key, int fallback) const;
  // writes "warning: file:line: unknown key 'k'" for each key not in `known`, in file order
  int report_unknown(const std::vector<std::string>& known, std::ostream& os) const;
  const std::string& source() const { return source_; }

 private:
  struct Item {
    std::string key, value;
    int line = 0;
  };
  const Item* find(const std::string& key) const;
  const Item& need(const std::string& key) const;
  double parse_number(const Item& it) const;
  std::string source_;
  std::vector<Item> items_;  // file order
};

// Runs one case on GPU `device` (run_serial_case, driver.cpp:150-310): read the mesh, assemble,
// apply tractions and Dirichlet values, solve, recover stresses (and the Kirsch metrics for
// bc.traction = kirsch_exact), write <output_dir>/timings.json.  Extra keys:
//   gpu.method       cg_jacobi (default: the reference's FP64 recurrence), cg_fp64_refined,
//                    ir_fp32, fcg_multilevel (aggregation multilevel, Chebyshev smoother),
//                    fcg_learned_multilevel (learned smoother; gpu.mpnn_params required)
//   gpu.mpnn_params  raw little-endian float32 parameter file of the learned preconditioner
// Returns 0 on success, 2 if the solver did not converge, 1 on error (message on `err`).
int run_case(const std::string& case_path, std::ostream& out, std::ostream& err, int device = 0);

}  // namespace cstfem_hip
