// io.cpp -- host-side input: the Gmsh 4.1 ASCII reader and writer, and the case-file keys.
//
// Gmsh reader.  The file is read whole and cut into non-blank lines that keep their line
// numbers.  A first pass splits it into sections ($Name ... $EndName) and checks their order;
// a second pass parses each known section from its own range of lines, so a parser can
// neither run past its $End marker nor into the next section; a last check lists the
// sections a mesh needs.  Node tags are looked up in a sorted (tag, index) table.
// The format facts kept from the reference's reader (src/mesh_io.cpp:114-356), so that both
// number a file's nodes and triangles identically and reject the same files:
//   - version 4.1, ASCII (file-type 0); $MeshFormat first; $PhysicalNames, $Entities, $Nodes
//     and $Elements required, $Nodes and $Entities before $Elements; other sections skipped;
//   - node index = position in the file; element types 1 (2-node line), 2 (3-node
//     triangle), 15 (point);
//   - a triangle is degenerate when its signed area is 0 or below 1e-14 times the squared
//     diagonal of the nodes' bounding box, and is reoriented counter-clockwise (second and
//     third node swapped) when its signed area is negative;
//   - line elements are collected under every physical name of their entity (the tag itself
//     when it has no name);
//   - errors are std::runtime_error "file:line: message" with the reference's messages, the
//     line being the last one read (the file's line count at its end).
//
// Case keys.  "key = value" per line, '#' starts a comment, a key may appear once; typed
// lookups name the file and line of a bad value.
#include <algorithm>
#include <cerrno>
#include <climits>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <iomanip>
#include <iterator>
#include <limits>
#include <sstream>

#include "cstfem_hip_io.hpp"

namespace cstfem_hip {

namespace {

bool is_space(char ch) { return ch == ' ' || ch == '\t' || ch == '\r'; }

// ---- a Gmsh file as numbered lines -----------------------------------------------------

struct MshLine {
  int no;            // 1-based line number in the file
  std::string text;  // without leading and trailing blanks
};

// [begin, end) indices into the line list of one section's body
struct Section {
  std::string name;
  int header;  // index of the $Name line
  int begin, end;
};

class MshFile {
 public:
  MshFile(std::istream& in, std::string source) : source_(std::move(source)) {
    std::string raw;
    while (std::getline(in, raw)) {
      ++physical_lines_;
      size_t a = 0, b = raw.size();
      while (a < b && is_space(raw[a])) ++a;
      while (b > a && is_space(raw[b - 1])) --b;
      if (a < b) lines_.push_back({physical_lines_, raw.substr(a, b - a)});
    }
  }
  const std::vector<MshLine>& lines() const { return lines_; }
  int physical_lines() const { return physical_lines_; }
  [[noreturn]] void fail(int line_no, const std::string& msg) const {
    throw std::runtime_error(source_ + ":" + std::to_string(line_no) + ": " + msg);
  }

 private:
  std::string source_;
  std::vector<MshLine> lines_;
  int physical_lines_ = 0;
};

// The values of one line, read left to right like a C++ stream: a number is the longest
// numeric prefix after the blanks (what follows it is left for the next read), names and
// special values (nan, inf) are not numbers.
class Fields {
 public:
  Fields(const MshFile& f, const MshLine& l) : file_(f), line_(l) {}

  long long integer(const char* what) {
    const char* p = start();
    char* end = nullptr;
    errno = 0;
    const long long v = numeric(p) ? std::strtoll(p, &end, 10) : 0;
    if (!numeric(p) || end == p || errno == ERANGE)
      file_.fail(line_.no, std::string("expected integer (") + what + ")");
    pos_ = static_cast<size_t>(end - line_.text.c_str());
    return v;
  }
  double real(const char* what) {
    const char* p = start();
    char* end = nullptr;
    errno = 0;
    const double v = numeric(p) ? std::strtod(p, &end) : 0.0;
    if (!numeric(p) || end == p || (errno == ERANGE && std::isinf(v)))
      file_.fail(line_.no, std::string("expected number (") + what + ")");
    pos_ = static_cast<size_t>(end - line_.text.c_str());
    return v;
  }
  int bounded(const char* what) {
    const long long v = integer(what);
    if (v < INT_MIN || v > INT_MAX) file_.fail(line_.no, std::string(what) + " out of range");
    return static_cast<int>(v);
  }
  size_t count(const char* what) {
    const long long v = integer(what);
    if (v < 0) file_.fail(line_.no, std::string(what) + " must be non-negative");
    return static_cast<size_t>(v);
  }
  // the rest of the line after the values read so far, leading blanks removed
  std::string rest() {
    const char* p = start();
    pos_ = line_.text.size();
    return std::string(p);
  }

 private:
  const char* start() {
    while (pos_ < line_.text.size() && is_space(line_.text[pos_])) ++pos_;
    return line_.text.c_str() + pos_;
  }
  static bool numeric(const char* p) {
    if (*p == '+' || *p == '-') ++p;
    return (*p >= '0' && *p <= '9') || *p == '.';
  }
  const MshFile& file_;
  const MshLine& line_;
  size_t pos_ = 0;
};

// Reads the lines of one section body in order; running out of lines or leaving lines
// unread are errors at the reference's line numbers.
class Body {
 public:
  Body(const MshFile& f, const Section& s) : file_(f), sec_(s), at_(s.begin) {}
  const MshLine& take(const std::string& what) {
    if (at_ >= sec_.end) {
      // the reference's sequential reader would read the $End line as data here
      const auto& l = file_.lines()[sec_.end];
      file_.fail(l.no, "expected " + what + ", got '" + l.text + "'");
    }
    last_ = at_;
    return file_.lines()[at_++];
  }
  int last_no() const { return file_.lines()[last_ >= 0 ? last_ : sec_.header].no; }
  void finish() const {
    if (at_ < sec_.end) {
      const auto& l = file_.lines()[at_];
      file_.fail(l.no, "expected $End" + sec_.name + ", got '" + l.text + "'");
    }
  }

 private:
  const MshFile& file_;
  const Section& sec_;
  int at_;
  int last_ = -1;
};

// cst_signed_area (element.cpp:17-19)
double signed_area(const double* xy, const int c[3]) {
  const double x1 = xy[2 * c[0]], y1 = xy[2 * c[0] + 1];
  const double x2 = xy[2 * c[1]], y2 = xy[2 * c[1] + 1];
  const double x3 = xy[2 * c[2]], y3 = xy[2 * c[2] + 1];
  return 0.5 * ((x2 - x1) * (y3 - y1) - (x3 - x1) * (y2 - y1));
}

class MshParser {
 public:
  explicit MshParser(const MshFile& f) : f_(f) {}

  Mesh run() {
    const std::vector<Section> secs = split();
    bool format = false, names = false, entities = false, nodes = false, elements = false;
    for (const Section& s : secs) {
      Body body(f_, s);
      if (s.name == "MeshFormat") {
        mesh_format(body);
        format = true;
      } else if (s.name == "PhysicalNames") {
        physical_names(body);
        names = true;
      } else if (s.name == "Entities") {
        entities_section(body);
        entities = true;
      } else if (s.name == "Nodes") {
        nodes_section(body);
        nodes = true;
      } else if (s.name == "Elements") {
        const int at = f_.lines()[s.header].no;
        if (!nodes) f_.fail(at, "$Elements must come after $Nodes");
        if (!entities) f_.fail(at, "$Elements must come after $Entities");
        elements_section(body);
        elements = true;
      } else {
        continue;  // a section the mesh does not need
      }
      body.finish();
    }
    const int end = f_.physical_lines();
    if (!format) f_.fail(end, "missing $MeshFormat section");
    if (!names) f_.fail(end, "missing $PhysicalNames section");
    if (!entities) f_.fail(end, "missing $Entities section");
    if (!nodes) f_.fail(end, "missing $Nodes section");
    if (!elements) f_.fail(end, "missing $Elements section");
    return std::move(mesh_);
  }

 private:
  // $Name ... $EndName ranges, in file order; $MeshFormat first and only once
  std::vector<Section> split() const {
    const auto& L = f_.lines();
    std::vector<Section> out;
    const int n = static_cast<int>(L.size());
    for (int i = 0; i < n;) {
      const std::string& h = L[i].text;
      if (h.empty() || h[0] != '$' || h.compare(0, 4, "$End") == 0)
        f_.fail(L[i].no, "expected a section header, got '" + h + "'");
      const std::string name = h.substr(1);
      if (out.empty() && name != "MeshFormat") f_.fail(L[i].no, "file must start with $MeshFormat");
      if (name == "MeshFormat" && !out.empty()) f_.fail(L[i].no, "duplicate $MeshFormat section");
      const std::string close = "$End" + name;
      int j = i + 1;
      while (j < n && L[j].text != close) ++j;
      if (j == n) {
        if (name == "MeshFormat" || name == "PhysicalNames" || name == "Entities" ||
            name == "Nodes" || name == "Elements")
          f_.fail(f_.physical_lines(), "unexpected end of file while reading " + close);
        f_.fail(f_.physical_lines(), "unexpected end of file inside $" + name);
      }
      out.push_back({name, i, i + 1, j});
      i = j + 1;
    }
    return out;
  }

  void mesh_format(Body& b) {
    const MshLine& l = b.take("$MeshFormat");
    const std::string version = l.text.substr(0, l.text.find_first_of(" \t"));
    if (version != "4.1") f_.fail(l.no, "unsupported .msh version '" + version + "' (need 4.1)");
    Fields t(f_, l);
    t.real("version");
    if (t.integer("file-type") != 0)
      f_.fail(l.no, "binary .msh files are not supported (need ASCII, file-type 0)");
    t.integer("data-size");
  }

  void physical_names(Body& b) {
    Fields head(f_, b.take("$PhysicalNames"));
    const size_t n = head.count("numPhysicalNames");
    for (size_t i = 0; i < n; ++i) {
      const MshLine& l = b.take("physical name");
      Fields t(f_, l);
      const int dim = t.bounded("dimension");
      const int tag = t.bounded("physical tag");
      const std::string q = t.rest();
      if (q.size() < 2 || q.front() != '"' || q.back() != '"')
        f_.fail(l.no, "physical name must be quoted");
      PhysicalGroup g{dim, tag, q.substr(1, q.size() - 2)};
      for (const PhysicalGroup& o : mesh_.physical_groups)
        if (o.dim == dim && o.tag == tag)
          f_.fail(l.no, "duplicate physical group (" + std::to_string(dim) + ", " +
                            std::to_string(tag) + ")");
      mesh_.physical_groups.push_back(std::move(g));
    }
  }

  void entities_section(Body& b) {
    Fields head(f_, b.take("$Entities"));
    static const char* const kCount[4] = {"numPoints", "numCurves", "numSurfaces", "numVolumes"};
    size_t per_dim[4];
    for (int d = 0; d < 4; ++d) per_dim[d] = head.count(kCount[d]);
    for (int d = 0; d < 4; ++d)
      for (size_t i = 0; i < per_dim[d]; ++i) {
        Fields t(f_, b.take("entity"));
        EntityPhys e;
        e.dim = d;
        e.tag = t.bounded("entity tag");
        for (int k = 0; k < (d == 0 ? 3 : 6); ++k) t.real("entity coordinate");  // point / box
        const size_t np = t.count("numPhysicalTags");
        for (size_t k = 0; k < np; ++k) e.phys.push_back(t.bounded("physical tag"));
        // the bounding entities of curves and surfaces follow and are not needed
        entity_phys_.push_back(std::move(e));
      }
  }

  void nodes_section(Body& b) {
    Fields head(f_, b.take("$Nodes"));
    const size_t nblocks = head.count("numEntityBlocks");
    const size_t nnodes = head.count("numNodes");
    mesh_.xy.reserve(2 * nnodes);
    mesh_.node_tags.reserve(nnodes);
    std::vector<int> tag_line;
    tag_line.reserve(nnodes);
    for (size_t blk = 0; blk < nblocks; ++blk) {
      Fields h(f_, b.take("node block header"));
      h.integer("entityDim");
      h.integer("entityTag");
      h.integer("parametric");
      const size_t n = h.count("numNodesInBlock");
      for (size_t i = 0; i < n; ++i) {
        const MshLine& l = b.take("node tag");
        const long long tag = Fields(f_, l).integer("node tag");
        if (tag <= 0) f_.fail(l.no, "node tag must be positive");
        mesh_.node_tags.push_back(static_cast<size_t>(tag));
        tag_line.push_back(l.no);
      }
      for (size_t i = 0; i < n; ++i) {
        Fields t(f_, b.take("node coordinates"));
        const double x = t.real("x"), y = t.real("y");
        t.real("z");  // planar mesh: z (and any parametric coordinates) are not used
        mesh_.xy.push_back(x);
        mesh_.xy.push_back(y);
      }
    }
    // tag -> index table; a repeated tag is reported where it first repeats in the file
    by_tag_.resize(mesh_.node_tags.size());
    for (size_t i = 0; i < by_tag_.size(); ++i) by_tag_[i] = {mesh_.node_tags[i], (int)i};
    std::sort(by_tag_.begin(), by_tag_.end());
    int dup = -1;
    for (size_t i = 1; i < by_tag_.size(); ++i)
      if (by_tag_[i].first == by_tag_[i - 1].first) {
        const int second = by_tag_[i].second;  // the later of the two in file order
        if (dup < 0 || second < dup) dup = second;
      }
    if (dup >= 0)
      f_.fail(tag_line[dup], "duplicate node tag " + std::to_string(mesh_.node_tags[dup]));
    if (mesh_.node_tags.size() != nnodes)
      f_.fail(b.last_no(), "node count mismatch: header says " + std::to_string(nnodes) +
                               ", blocks contain " + std::to_string(mesh_.node_tags.size()));
  }

  int node_index(long long tag, int line_no) const {
    const auto it = std::lower_bound(by_tag_.begin(), by_tag_.end(),
                                     std::pair<size_t, int>(static_cast<size_t>(tag), INT_MIN));
    if (tag <= 0 || it == by_tag_.end() || it->first != static_cast<size_t>(tag))
      f_.fail(line_no, "unknown node tag " + std::to_string(tag));
    return it->second;
  }

  // physical names of the line entity (dim, tag): names, or the tag number when unnamed
  std::vector<std::string> line_groups(int dim, int tag) const {
    std::vector<std::string> out;
    for (const EntityPhys& e : entity_phys_) {
      if (e.dim != dim || e.tag != tag) continue;
      out.clear();  // a repeated entity: the last definition holds
      for (int p : e.phys) {
        std::string name = std::to_string(p);
        for (const PhysicalGroup& g : mesh_.physical_groups)
          if (g.dim == dim && g.tag == p) {
            name = g.name;
            break;
          }
        out.push_back(std::move(name));
      }
    }
    return out;
  }

  void elements_section(Body& b) {
    double lo[2] = {std::numeric_limits<double>::max(), std::numeric_limits<double>::max()};
    double hi[2] = {std::numeric_limits<double>::lowest(), std::numeric_limits<double>::lowest()};
    const size_t nn = mesh_.num_nodes();
    for (size_t i = 0; i < 2 * nn; ++i) {
      lo[i & 1] = std::min(lo[i & 1], mesh_.xy[i]);
      hi[i & 1] = std::max(hi[i & 1], mesh_.xy[i]);
    }
    const double dx = hi[0] - lo[0], dy = hi[1] - lo[1];
    const double threshold = 1e-14 * (nn ? dx * dx + dy * dy : 0.0);

    Fields head(f_, b.take("$Elements"));
    const size_t nblocks = head.count("numEntityBlocks");
    const size_t nelems = head.count("numElements");
    size_t total = 0;
    for (size_t blk = 0; blk < nblocks; ++blk) {
      const MshLine& hl = b.take("element block header");
      Fields h(f_, hl);
      const int edim = h.bounded("entityDim");
      const int etag = h.bounded("entityTag");
      const int type = h.bounded("elementType");
      const size_t n = h.count("numElementsInBlock");
      total += n;
      const int nodes = type == 1 ? 2 : type == 2 ? 3 : type == 15 ? 1 : 0;
      if (!nodes)
        f_.fail(hl.no, "unsupported element type " + std::to_string(type) +
                           " (supported: 1 = 2-node line, 2 = 3-node triangle, 15 = point)");
      const std::vector<std::string> groups =
          type == 1 ? line_groups(edim, etag) : std::vector<std::string>();
      for (size_t i = 0; i < n; ++i) {
        const MshLine& l = b.take("element");
        Fields t(f_, l);
        t.integer("element tag");
        int c[3] = {0, 0, 0};
        for (int k = 0; k < nodes; ++k) c[k] = node_index(t.integer("element node tag"), l.no);
        if (type == 2) {
          const double area = signed_area(mesh_.xy.data(), c);
          if (area == 0.0 || std::fabs(area) < threshold) {
            char msg[128];
            std::snprintf(msg, sizeof msg, "degenerate triangle (area %g, threshold %g)", area,
                          threshold);
            f_.fail(l.no, msg);
          }
          if (area < 0.0) {
            std::swap(c[1], c[2]);
            ++mesh_.flipped_triangles;
          }
          mesh_.triangles.insert(mesh_.triangles.end(), c, c + 3);
        } else if (type == 1) {
          for (const std::string& g : groups) mesh_.boundary_edges[g].push_back({c[0], c[1]});
        }
      }
    }
    if (total != nelems)
      f_.fail(b.last_no(), "element count mismatch: header says " + std::to_string(nelems) +
                               ", blocks contain " + std::to_string(total));
  }

  struct EntityPhys {
    int dim = 0, tag = 0;
    std::vector<int> phys;
  };

  const MshFile& f_;
  Mesh mesh_;
  std::vector<EntityPhys> entity_phys_;
  std::vector<std::pair<size_t, int>> by_tag_;  // (node tag, index), ascending
};

std::string trimmed(const std::string& s) {
  size_t a = 0, b = s.size();
  while (a < b && is_space(s[a])) ++a;
  while (b > a && is_space(s[b - 1])) --b;
  return s.substr(a, b - a);
}

}  // namespace

Mesh read_gmsh(std::istream& in, const std::string& source_name) {
  const MshFile file(in, source_name);
  return MshParser(file).run();
}

Mesh read_gmsh(const std::string& path) {
  std::ifstream in(path);
  if (!in) throw std::runtime_error(path + ": cannot open file");
  return read_gmsh(in, path);
}

// ---- writer ------------------------------------------------------------------------------

void write_gmsh(const Mesh& m, std::ostream& out) {
  const size_t nn = m.num_nodes(), ne = m.num_triangles();
  const int ng = static_cast<int>(m.boundary_edges.size());
  double x0 = 0, x1 = 0, y0 = 0, y1 = 0;
  for (size_t i = 0; i < nn; ++i) {
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
  for (size_t i = 0; i < nn; ++i) out << i + 1 << "\n";
  for (size_t i = 0; i < nn; ++i) out << m.xy[2 * i] << " " << m.xy[2 * i + 1] << " 0\n";
  out << "$EndNodes\n";
  size_t nedges = 0;
  for (const auto& g : m.boundary_edges) nedges += g.second.size();
  out << "$Elements\n" << ng + 1 << " " << ne + nedges << " 1 " << ne + nedges << "\n";
  size_t etag = 1;
  int c = 1;
  for (const auto& g : m.boundary_edges) {
    out << "1 " << c++ << " 1 " << g.second.size() << "\n";
    for (const auto& e : g.second) out << etag++ << " " << e[0] + 1 << " " << e[1] + 1 << "\n";
  }
  out << "2 1 2 " << ne << "\n";
  for (size_t e = 0; e < ne; ++e)
    out << etag++ << " " << m.triangles[3 * e] + 1 << " " << m.triangles[3 * e + 1] + 1 << " "
        << m.triangles[3 * e + 2] + 1 << "\n";
  out << "$EndElements\n";
}

// ---- case keys ---------------------------------------------------------------------------

CaseKeys CaseKeys::read(const std::string& path) {
  std::ifstream in(path);
  if (!in) throw std::runtime_error(path + ": cannot open file");
  return read(in, path);
}

CaseKeys CaseKeys::read(std::istream& in, const std::string& source_name) {
  CaseKeys ck;
  ck.source_ = source_name;
  std::string raw;
  for (int no = 1; std::getline(in, raw); ++no) {
    const std::string line = trimmed(raw.substr(0, raw.find('#')));
    if (line.empty()) continue;
    const size_t eq = line.find('=');
    const std::string at = source_name + ":" + std::to_string(no) + ": ";
    if (eq == std::string::npos)
      throw std::runtime_error(at + "expected 'key = value', got '" + line + "'");
    Item it{trimmed(line.substr(0, eq)), trimmed(line.substr(eq + 1)), no};
    if (it.key.empty() || it.key.find_first_of(" \t") != std::string::npos)
      throw std::runtime_error(at + "bad key '" + it.key + "'");
    if (const Item* prev = ck.find(it.key))
      throw std::runtime_error(at + "key '" + it.key + "' repeats line " +
                               std::to_string(prev->line));
    ck.items_.push_back(std::move(it));
  }
  return ck;
}

const CaseKeys::Item* CaseKeys::find(const std::string& key) const {
  for (const Item& it : items_)
    if (it.key == key) return &it;
  return nullptr;
}

const CaseKeys::Item& CaseKeys::need(const std::string& key) const {
  const Item* it = find(key);
  if (!it) throw std::runtime_error(source_ + ": key '" + key + "' is required");
  return *it;
}

std::string CaseKeys::text(const std::string& key) const { return need(key).value; }

std::string CaseKeys::text(const std::string& key, const std::string& fallback) const {
  const Item* it = find(key);
  return it ? it->value : fallback;
}

double CaseKeys::parse_number(const Item& it) const {
  char* end = nullptr;
  errno = 0;
  const double v = std::strtod(it.value.c_str(), &end);
  if (it.value.empty() || *end != '\0' || errno == ERANGE || !std::isfinite(v))
    throw std::runtime_error(source_ + ":" + std::to_string(it.line) + ": '" + it.key +
                             "' needs a finite number, got '" + it.value + "'");
  return v;
}

double CaseKeys::number(const std::string& key) const { return parse_number(need(key)); }

double CaseKeys::number(const std::string& key, double fallback) const {
  const Item* it = find(key);
  return it ? parse_number(*it) : fallback;
}

int CaseKeys::integer(const std::string& key, int fallback) const {
  const Item* it = find(key);
  if (!it) return fallback;
  char* end = nullptr;
  errno = 0;
  const long long v = std::strtoll(it->value.c_str(), &end, 10);
  if (it->value.empty() || *end != '\0' || errno == ERANGE || v < INT_MIN || v > INT_MAX)
    throw std::runtime_error(source_ + ":" + std::to_string(it->line) + ": '" + key +
                             "' needs an integer, got '" + it->value + "'");
  return static_cast<int>(v);
}

int CaseKeys::report_unknown(const std::vector<std::string>& known, std::ostream& os) const {
  int count = 0;
  for (const Item& it : items_)
    if (std::find(known.begin(), known.end(), it.key) == known.end()) {
      os << "warning: " << source_ << ":" << it.line << ": unknown key '" << it.key << "'\n";
      ++count;
    }
  return count;
}

}  // namespace cstfem_hip
