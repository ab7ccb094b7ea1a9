// Case-file keys (host/io.cpp CaseKeys) on well-formed and malformed texts, for the
// AddressSanitizer/UndefinedBehaviorSanitizer build of the host code: prints what each text
// gives (values or the error message); exit status 0 unless an expectation fails.
#include <iostream>
#include <sstream>
#include <string>
#include <vector>

#include "cstfem_hip_io.hpp"

int main() {
  int fails = 0;
  auto expect = [&](bool ok, const std::string& what) {
    if (!ok) {
      std::cout << "FAILED: " << what << "\n";
      ++fails;
    }
  };
  {
    std::istringstream in("# comment\nmesh.file = a.msh\n  material.E = 2e5  \nsolver.max_iter=300\n"
                          "empty =\n");
    const auto k = cstfem_hip::CaseKeys::read(in, "good.case");
    expect(k.text("mesh.file") == "a.msh", "text");
    expect(k.number("material.E") == 2e5, "number");
    expect(k.integer("solver.max_iter", 0) == 300, "integer");
    expect(k.number("missing", 1.5) == 1.5, "fallback");
    std::ostringstream warn;
    const int unknown = k.report_unknown({"mesh.file", "material.E", "solver.max_iter"}, warn);
    std::cout << "unknown keys: " << unknown << "\n" << warn.str();
  }
  const std::vector<std::string> bad = {"no equals sign\n", "= value\n", "a = 1\na = 2\n",
                                        "odd key = 1\n",
                                        "x = 1e999999\n", "x = abc\n", std::string("x = 1\0", 6),
                                        std::string(5000, 'k') + " = 1\n"};
  for (const auto& text : bad) {
    std::istringstream in(text);
    try {
      const auto k = cstfem_hip::CaseKeys::read(in, "bad.case");
      if (k.has("x")) std::cout << "x = " << k.number("x") << "\n";
      std::cout << "read without error\n";
    } catch (const std::exception& e) {
      std::cout << "error: " << e.what() << "\n";
    }
  }
  return fails;
}
