// cstfem_hip_mesh <in.msh> <out-prefix>: reads a Gmsh 4.1 file with the host reader and writes
// the arrays as text (<prefix>.xy, <prefix>.tri, <prefix>.groups) for comparison with other
// readers; on a read error prints the message and returns 1.
// cstfem_hip_mesh --rewrite <in.msh> <out.msh>: reads and writes the mesh back (write_gmsh).
// No GPU is used.
#include <fstream>
#include <iomanip>
#include <iostream>

#include "cstfem_hip_io.hpp"

int main(int argc, char** argv) {
  if (argc == 4 && std::string(argv[1]) == "--rewrite") {
    try {
      const cstfem_hip::Mesh m = cstfem_hip::read_gmsh(argv[2]);
      std::ofstream o(argv[3]);
      cstfem_hip::write_gmsh(m, o);
      return o ? 0 : 1;
    } catch (const std::exception& e) {
      std::cerr << e.what() << "\n";
      return 1;
    }
  }
  if (argc != 3) {
    std::cerr << "usage: cstfem_hip_mesh <in.msh> <out-prefix> | --rewrite <in.msh> <out.msh>\n";
    return 1;
  }
  try {
    const cstfem_hip::Mesh m = cstfem_hip::read_gmsh(argv[1]);
    const std::string p = argv[2];
    std::ofstream xy(p + ".xy"), tri(p + ".tri"), gr(p + ".groups");
    xy << std::setprecision(17);
    for (std::size_t i = 0; i < m.num_nodes(); ++i) xy << m.xy[2 * i] << " " << m.xy[2 * i + 1] << "\n";
    for (std::size_t e = 0; e < m.num_triangles(); ++e)
      tri << m.triangles[3 * e] << " " << m.triangles[3 * e + 1] << " " << m.triangles[3 * e + 2] << "\n";
    for (const auto& [name, edges] : m.boundary_edges)
      for (const auto& e : edges) gr << name << " " << e[0] << " " << e[1] << "\n";
    std::cout << m.num_nodes() << " " << m.num_triangles() << " " << m.flipped_triangles << "\n";
    return 0;
  } catch (const std::exception& e) {
    std::cerr << e.what() << "\n";
    return 1;
  }
}
