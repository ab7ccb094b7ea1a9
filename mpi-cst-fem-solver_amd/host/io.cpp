// io.cpp -- Gmsh 4.1 ASCII reader / writer and the case-file parser (host side).
//
// Reference behaviour followed: src/mesh_io.cpp:20-356 (sections, node numbering by file
// order, element types 1/2/15, degenerate-triangle threshold 1e-14 * bbox diagonal^2,
// clockwise triangles flipped, line elements collected per physical name, "file:line:"
// errors), src/case.cpp:25-128 (key = value, '#' comments, duplicate keys rejected).
#include <algorithm>
#include <cerrno>
#include <climits>
#include <cmath>
#include <cstdlib>
#include <fstream>
#include <iomanip>
#include <limits>
#include <sstream>
#include <unordered_map>

#include "cstfem_hip_io.hpp"

namespace cstfem_hip {

namespace {

std::string trim(const std::string& s, const char* ws = " \t\r") {
  const auto a = s.find_first_not_of(ws);
  if (a == std::string::npos) return {};
  const auto b = s.find_last_not_of(ws);
  return s.substr(a, b - a + 1);
}

// Non-blank, trimmed lines with their line numbers.
class Lines {
 public:
  Lines(std::istream& in, std::string src) : in_(in), src_(std::move(src)) {}
  bool next(std::string& out) {
    std::string raw;
    while (std::getline(in_, raw)) {
      ++no_;
      out = trim(raw);
      if (!out.empty()) return true;
    }
    return false;
  }
  std::string need(const std::string& what) {
    std::string l;
    if (!next(l)) fail("unexpected end of file while reading " + what);
    return l;
  }
  [[noreturn]] void fail(const std::string& msg) const {
    throw std::runtime_error(src_ + ":" + std::to_string(no_) + ": " + msg);
  }

 private:
  std::istream& in_;
  std::string src_;
  int no_ = 0;
};

class Fields {
 public:
  Fields(Lines& l, const std::string& line) : l_(l), ss_(line) {}
  long long i(const char* what) {
    long long v = 0;
    if (!(ss_ >> v)) l_.fail(std::string("expected integer (") + what + ")");
    return v;
  }
  double d(const char* what) {
    double v = 0.0;
    if (!(ss_ >> v)) l_.fail(std::string("expected number (") + what + ")");
    return v;
  }
  std::string rest() {
    std::string s;
    std::getline(ss_, s);
    return trim(s, " \t");
  }

 private:
  Lines& l_;
  std::istringstream ss_;
};

int as_int(Lines& l, long long v, const char* what) {
  if (v < INT_MIN || v > INT_MAX) l.fail(std::string(what) + " out of range");
  return (int)v;
}

std::size_t as_count(Lines& l, long long v, const char* what) {
  if (v < 0) l.fail(std::string(what) + " must be non-negative");
  return (std::size_t)v;
}

void end_of(Lines& l, const std::string& section) {
  const std::string line = l.need("$End" + section);
  if (line != "$End" + section) l.fail("expected $End" + section + ", got '" + line + "'");
}

struct Reader {
  Lines& l;
  Mesh& m;
  bool fmt = false, names = false, ents = false, nodes = false, elems = false;
  std::map<std::pair<int, int>, std::string> phys_name;          // (dim, tag) -> name
  std::map<std::pair<int, int>, std::vector<int>> ent_phys;      // entity -> physical tags
  std::unordered_map<long long, int> index_of;                    // node tag -> index

  void format() {
    const std::string line = l.need("$MeshFormat");
    const std::string version = line.substr(0, line.find_first_of(" \t"));
    if (version != "4.1") l.fail("unsupported .msh version '" + version + "' (need 4.1)");
    Fields f(l, line);
    f.d("version");
    if (f.i("file-type") != 0) l.fail("binary .msh files are not supported (need ASCII, file-type 0)");
    f.i("data-size");
    end_of(l, "MeshFormat");
    fmt = true;
  }

  void physical_names() {
    Fields h(l, l.need("$PhysicalNames"));
    const std::size_t n = as_count(l, h.i("numPhysicalNames"), "numPhysicalNames");
    for (std::size_t k = 0; k < n; ++k) {
      Fields f(l, l.need("physical name"));
      const int dim = as_int(l, f.i("dimension"), "dimension");
      const int tag = as_int(l, f.i("physical tag"), "physical tag");
      std::string name = f.rest();
      if (name.size() < 2 || name.front() != '"' || name.back() != '"')
        l.fail("physical name must be quoted");
      name = name.substr(1, name.size() - 2);
      if (!phys_name.emplace(std::make_pair(dim, tag), name).second)
        l.fail("duplicate physical group (" + std::to_string(dim) + ", " + std::to_string(tag) + ")");
      m.physical_groups.push_back({dim, tag, name});
    }
    end_of(l, "PhysicalNames");
    names = true;
  }

  void entities() {
    Fields h(l, l.need("$Entities"));
    std::size_t count[4];
    const char* what[4] = {"numPoints", "numCurves", "numSurfaces", "numVolumes"};
    for (int d = 0; d < 4; ++d) count[d] = as_count(l, h.i(what[d]), what[d]);
    for (int dim = 0; dim < 4; ++dim)
      for (std::size_t k = 0; k < count[dim]; ++k) {
        Fields f(l, l.need("entity"));
        const int tag = as_int(l, f.i("entity tag"), "entity tag");
        for (int c = 0; c < (dim == 0 ? 3 : 6); ++c) f.d("entity coordinate");
        const std::size_t np = as_count(l, f.i("numPhysicalTags"), "numPhysicalTags");
        std::vector<int> tags(np);
        for (auto& t : tags) t = as_int(l, f.i("physical tag"), "physical tag");
        ent_phys[{dim, tag}] = std::move(tags);
      }
    end_of(l, "Entities");
    ents = true;
  }

  void node_section() {
    Fields h(l, l.need("$Nodes"));
    const std::size_t nblocks = as_count(l, h.i("numEntityBlocks"), "numEntityBlocks");
    const std::size_t total = as_count(l, h.i("numNodes"), "numNodes");
    m.xy.reserve(2 * total);
    m.node_tags.reserve(total);
    for (std::size_t b = 0; b < nblocks; ++b) {
      Fields bh(l, l.need("node block header"));
      bh.i("entityDim");
      bh.i("entityTag");
      bh.i("parametric");
      const std::size_t n = as_count(l, bh.i("numNodesInBlock"), "numNodesInBlock");
      for (std::size_t k = 0; k < n; ++k) {
        Fields f(l, l.need("node tag"));
        const long long tag = f.i("node tag");
        if (tag <= 0) l.fail("node tag must be positive");
        if (!index_of.emplace(tag, (int)m.node_tags.size()).second)
          l.fail("duplicate node tag " + std::to_string(tag));
        m.node_tags.push_back((std::size_t)tag);
      }
      for (std::size_t k = 0; k < n; ++k) {
        Fields f(l, l.need("node coordinates"));
        const double x = f.d("x"), y = f.d("y");
        f.d("z");
        m.xy.push_back(x);
        m.xy.push_back(y);
      }
    }
    if (m.node_tags.size() != total)
      l.fail("node count mismatch: header says " + std::to_string(total) + ", blocks contain " +
             std::to_string(m.node_tags.size()));
    end_of(l, "Nodes");
    nodes = true;
  }

  void element_section() {
    if (!nodes) l.fail("$Elements must come after $Nodes");
    if (!ents) l.fail("$Elements must come after $Entities");
    double x0 = std::numeric_limits<double>::max(), x1 = std::numeric_limits<double>::lowest();
    double y0 = x0, y1 = x1;
    for (std::size_t i = 0; i < m.num_nodes(); ++i) {
      x0 = std::min(x0, m.xy[2 * i]);
      x1 = std::max(x1, m.xy[2 * i]);
      y0 = std::min(y0, m.xy[2 * i + 1]);
      y1 = std::max(y1, m.xy[2 * i + 1]);
    }
    const double diag2 = m.num_nodes() ? (x1 - x0) * (x1 - x0) + (y1 - y0) * (y1 - y0) : 0.0;
    const double min_area = 1e-14 * diag2;
    Fields h(l, l.need("$Elements"));
    const std::size_t nblocks = as_count(l, h.i("numEntityBlocks"), "numEntityBlocks");
    const std::size_t total = as_count(l, h.i("numElements"), "numElements");
    std::size_t seen = 0;
    for (std::size_t b = 0; b < nblocks; ++b) {
      Fields bh(l, l.need("element block header"));
      const int edim = as_int(l, bh.i("entityDim"), "entityDim");
      const int etag = as_int(l, bh.i("entityTag"), "entityTag");
      const int type = as_int(l, bh.i("elementType"), "elementType");
      const std::size_t n = as_count(l, bh.i("numElementsInBlock"), "numElementsInBlock");
      seen += n;
      const int npe = type == 1 ? 2 : type == 2 ? 3 : type == 15 ? 1 : 0;
      if (npe == 0)
        l.fail("unsupported element type " + std::to_string(type) +
               " (supported: 1 = 2-node line, 2 = 3-node triangle, 15 = point)");
      std::vector<std::string> groups;
      if (type == 1) {
        const auto it = ent_phys.find({edim, etag});
        if (it != ent_phys.end())
          for (int p : it->second) {
            const auto nm = phys_name.find({edim, p});
            groups.push_back(nm != phys_name.end() ? nm->second : std::to_string(p));
          }
      }
      for (std::size_t k = 0; k < n; ++k) {
        Fields f(l, l.need("element"));
        f.i("element tag");
        int conn[3] = {0, 0, 0};
        for (int c = 0; c < npe; ++c) {
          const long long tag = f.i("element node tag");
          const auto it = index_of.find(tag);
          if (tag <= 0 || it == index_of.end()) l.fail("unknown node tag " + std::to_string(tag));
          conn[c] = it->second;
        }
        if (type == 2) {
          const double* p0 = &m.xy[2 * conn[0]];
          const double* p1 = &m.xy[2 * conn[1]];
          const double* p2 = &m.xy[2 * conn[2]];
          // element.cpp:17-19
          const double area =
              0.5 * ((p1[0] - p0[0]) * (p2[1] - p0[1]) - (p2[0] - p0[0]) * (p1[1] - p0[1]));
          if (area == 0.0 || std::abs(area) < min_area) {
            std::ostringstream msg;
            msg << "degenerate triangle (area " << area << ", threshold " << min_area << ")";
            l.fail(msg.str());
          }
          if (area < 0.0) {
            std::swap(conn[1], conn[2]);
            ++m.flipped_triangles;
          }
          m.triangles.insert(m.triangles.end(), conn, conn + 3);
        } else if (type == 1) {
          for (const auto& g : groups) m.boundary_edges[g].push_back({conn[0], conn[1]});
        }
      }
    }
    if (seen != total)
      l.fail("element count mismatch: header says " + std::to_string(total) +
             ", blocks contain " + std::to_string(seen));
    end_of(l, "Elements");
    elems = true;
  }

  void skip(const std::string& section) {
    std::string line;
    while (l.next(line))
      if (line == "$End" + section) return;
    l.fail("unexpected end of file inside $" + section);
  }
};

}  // namespace

Mesh read_gmsh(std::istream& in, const std::string& source_name) {
  Lines l(in, source_name);
  Mesh m;
  Reader r{l, m};
  std::string line;
  while (l.next(line)) {
    if (line.front() != '$' || line.rfind("$End", 0) == 0)
      l.fail("expected a section header, got '" + line + "'");
    const std::string sec = line.substr(1);
    if (!r.fmt && sec != "MeshFormat") l.fail("file must start with $MeshFormat");
    if (sec == "MeshFormat") {
      if (r.fmt) l.fail("duplicate $MeshFormat section");
      r.format();
    } else if (sec == "PhysicalNames") {
      r.physical_names();
    } else if (sec == "Entities") {
      r.entities();
    } else if (sec == "Nodes") {
      r.node_section();
    } else if (sec == "Elements") {
      r.element_section();
    } else {
      r.skip(sec);
    }
  }
  const std::pair<bool, const char*> need[] = {{r.fmt, "$MeshFormat"},
                                               {r.names, "$PhysicalNames"},
                                               {r.ents, "$Entities"},
                                               {r.nodes, "$Nodes"},
                                               {r.elems, "$Elements"}};
  for (const auto& [ok, name] : need)
    if (!ok) l.fail(std::string("missing ") + name + " section");
  return m;
}

Mesh read_gmsh(const std::string& path) {
  std::ifstream in(path);
  if (!in) throw std::runtime_error(path + ": cannot open file");
  return read_gmsh(in, path);
}

void write_gmsh(const Mesh& m, std::ostream& out) {
  const std::size_t nn = m.num_nodes(), ne = m.num_triangles();
  const int ng = (int)m.boundary_edges.size();
  double x0 = 0, x1 = 0, y0 = 0, y1 = 0;
  for (std::size_t i = 0; i < nn; ++i) {
    const double x = m.xy[2 * i], y = m.xy[2 * i + 1];
    if (i == 0 || x < x0) x0 = x;
    if (i == 0 || x > x1) x1 = x;
    if (i == 0 || y < y0) y0 = y;
    if (i == 0 || y > y1) y1 = y;
  }
  out << std::setprecision(17);
  out << "$MeshFormat\n4.1 0 8\n$EndMeshFormat\n";
  out << "$PhysicalNames\n" << ng + 1 << "\n";
  int tag = 1;
  for (const auto& g : m.boundary_edges) out << "1 " << tag++ << " \"" << g.first << "\"\n";
  out << "2 " << tag << " \"domain\"\n$EndPhysicalNames\n";
  // curve entities 1..ng carry physical tags 1..ng; surface 1 carries tag ng + 1
  out << "$Entities\n0 " << ng << " 1 0\n";
  for (int c = 1; c <= ng; ++c)
    out << c << " " << x0 << " " << y0 << " 0 " << x1 << " " << y1 << " 0 1 " << c << " 0\n";
  out << "1 " << x0 << " " << y0 << " 0 " << x1 << " " << y1 << " 0 1 " << ng + 1 << " 0\n";
  out << "$EndEntities\n";
  out << "$Nodes\n1 " << nn << " 1 " << nn << "\n2 1 0 " << nn << "\n";
  for (std::size_t i = 0; i < nn; ++i) out << i + 1 << "\n";
  for (std::size_t i = 0; i < nn; ++i) out << m.xy[2 * i] << " " << m.xy[2 * i + 1] << " 0\n";
  out << "$EndNodes\n";
  std::size_t nedges = 0;
  for (const auto& g : m.boundary_edges) nedges += g.second.size();
  out << "$Elements\n" << ng + 1 << " " << ne + nedges << " 1 " << ne + nedges << "\n";
  std::size_t etag = 1;
  int c = 1;
  for (const auto& g : m.boundary_edges) {
    out << "1 " << c++ << " 1 " << g.second.size() << "\n";
    for (const auto& e : g.second) out << etag++ << " " << e[0] + 1 << " " << e[1] + 1 << "\n";
  }
  out << "2 1 2 " << ne << "\n";
  for (std::size_t e = 0; e < ne; ++e)
    out << etag++ << " " << m.triangles[3 * e] + 1 << " " << m.triangles[3 * e + 1] + 1 << " "
        << m.triangles[3 * e + 2] + 1 << "\n";
  out << "$EndElements\n";
}

// ---- case file ------------------------------------------------------------------------------

CaseFile CaseFile::load(const std::string& path) {
  std::ifstream in(path);
  if (!in) throw std::runtime_error(path + ": cannot open file");
  return parse(in, path);
}

CaseFile CaseFile::parse(std::istream& in, const std::string& source_name) {
  CaseFile cf;
  cf.source_ = source_name;
  std::string raw;
  int no = 0;
  while (std::getline(in, raw)) {
    ++no;
    const std::string line = trim(raw.substr(0, raw.find('#')));
    if (line.empty()) continue;
    const auto eq = line.find('=');
    if (eq == std::string::npos)
      throw std::runtime_error(cf.where(no) + ": expected 'key = value', got '" + line + "'");
    const std::string key = trim(line.substr(0, eq)), value = trim(line.substr(eq + 1));
    if (key.empty()) throw std::runtime_error(cf.where(no) + ": empty key");
    if (key.find_first_of(" \t") != std::string::npos)
      throw std::runtime_error(cf.where(no) + ": key '" + key + "' contains whitespace");
    const auto [it, fresh] = cf.entries_.emplace(key, Entry{value, no});
    if (!fresh)
      throw std::runtime_error(cf.where(no) + ": duplicate key '" + key +
                               "' (first defined on line " + std::to_string(it->second.line) + ")");
    cf.order_.push_back(key);
  }
  return cf;
}

bool CaseFile::has(const std::string& key) const { return entries_.count(key) != 0; }

std::string CaseFile::where(int line) const { return source_ + ":" + std::to_string(line); }

const CaseFile::Entry& CaseFile::entry(const std::string& key) const {
  const auto it = entries_.find(key);
  if (it == entries_.end()) throw std::runtime_error(source_ + ": missing required key '" + key + "'");
  return it->second;
}

std::string CaseFile::get_string(const std::string& key) const { return entry(key).value; }

std::string CaseFile::get_string(const std::string& key, const std::string& fallback) const {
  return has(key) ? get_string(key) : fallback;
}

double CaseFile::get_double(const std::string& key) const {
  const Entry& e = entry(key);
  const char* b = e.value.c_str();
  char* end = nullptr;
  errno = 0;
  const double v = std::strtod(b, &end);
  if (e.value.empty() || end != b + e.value.size())
    throw std::runtime_error(where(e.line) + ": key '" + key + "': '" + e.value + "' is not a number");
  if (errno == ERANGE || !std::isfinite(v))
    throw std::runtime_error(where(e.line) + ": key '" + key + "': '" + e.value +
                             "' is out of range or not finite");
  return v;
}

double CaseFile::get_double(const std::string& key, double fallback) const {
  return has(key) ? get_double(key) : fallback;
}

int CaseFile::get_int(const std::string& key) const {
  const Entry& e = entry(key);
  const char* b = e.value.c_str();
  char* end = nullptr;
  errno = 0;
  const long v = std::strtol(b, &end, 10);
  if (e.value.empty() || end != b + e.value.size())
    throw std::runtime_error(where(e.line) + ": key '" + key + "': '" + e.value + "' is not an integer");
  if (errno == ERANGE || v < INT_MIN || v > INT_MAX)
    throw std::runtime_error(where(e.line) + ": key '" + key + "': '" + e.value + "' is out of range");
  return (int)v;
}

int CaseFile::get_int(const std::string& key, int fallback) const {
  return has(key) ? get_int(key) : fallback;
}

std::vector<std::string> CaseFile::unknown_keys(const std::vector<std::string>& known) const {
  std::vector<std::string> out;
  for (const auto& k : order_)
    if (std::find(known.begin(), known.end(), k) == known.end()) out.push_back(k);
  return out;
}

int CaseFile::warn_unknown_keys(const std::vector<std::string>& known, std::ostream& os) const {
  const auto unknown = unknown_keys(known);
  for (const auto& k : unknown)
    os << "warning: " << where(entries_.at(k).line) << ": unknown key '" << k << "'\n";
  return (int)unknown.size();
}

}  // namespace cstfem_hip
