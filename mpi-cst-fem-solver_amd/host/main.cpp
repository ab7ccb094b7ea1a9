// cstfem_hip_run <case file> [device]: one case on one GPU (cstfem_serial of the reference).
#include <cstdlib>
#include <iostream>

#include "cstfem_hip_io.hpp"

int main(int argc, char** argv) {
  if (argc < 2 || argc > 3) {
    std::cerr << "usage: cstfem_hip_run <case file> [device]\n";
    return 1;
  }
  return cstfem_hip::run_case(argv[1], std::cout, std::cerr, argc == 3 ? std::atoi(argv[2]) : 0);
}
