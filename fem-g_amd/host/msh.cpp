// Gmsh MSH 4.1 (ASCII) reader and writer for the host mesh model.
#include <algorithm>
#include <cmath>
#include <fstream>
#include <map>
#include <sstream>
#include <stdexcept>
#include <string>
#include <unordered_map>
#include <vector>

#include "../csrc/refelem.h"
#include "model.hpp"

namespace gmshfem::model {

std::vector<int> Model::physicalEntities(int d, int tag) const
{
  for(const auto &p : physicals)
    if(p.dim == d && p.tag == tag) return p.entities;
  return {};
}

const Physical *Model::physical(const std::string &n) const
{
  for(const auto &p : physicals)
    if(p.name == n) return &p;
  return nullptr;
}

namespace {

// whitespace-separated tokens with the line they came from
class Tokens {
 public:
  explicit Tokens(std::istream &in) : _in(in) {}
  bool next(std::string &t)
  {
    while(!(_ls >> t)) {
      std::string line;
      if(!std::getline(_in, line)) return false;
      ++_line;
      _ls.clear();
      _ls.str(line);
    }
    return true;
  }
  // the rest of the current line (for quoted physical names)
  std::string rest()
  {
    std::string r;
    std::getline(_ls, r);
    return r;
  }
  std::string word(const char *what)
  {
    std::string t;
    if(!next(t)) fail(std::string("unexpected end of file, expected ") + what);
    return t;
  }
  long long integer(const char *what)
  {
    const std::string t = word(what);
    try {
      size_t k = 0;
      const long long v = std::stoll(t, &k);
      if(k != t.size()) throw std::invalid_argument(t);
      return v;
    }
    catch(const std::exception &) {
      fail(std::string("expected an integer (") + what + "), got '" + t + "'");
    }
    return 0;
  }
  double real(const char *what)
  {
    const std::string t = word(what);
    try {
      size_t k = 0;
      const double v = std::stod(t, &k);
      if(k != t.size()) throw std::invalid_argument(t);
      return v;
    }
    catch(const std::exception &) {
      fail(std::string("expected a number (") + what + "), got '" + t + "'");
    }
    return 0.;
  }
  void expect(const std::string &w)
  {
    const std::string t = word(w.c_str());
    if(t != w) fail("expected " + w + ", got '" + t + "'");
  }
  [[noreturn]] void fail(const std::string &msg) const
  {
    throw std::runtime_error("msh line " + std::to_string(_line) + ": " + msg);
  }

 private:
  std::istream &_in;
  std::istringstream _ls;
  long long _line = 0;
};

void skipSection(Tokens &tk, const std::string &name)
{
  const std::string end = "$End" + name.substr(1);
  std::string t;
  while(tk.next(t))
    if(t == end) return;
  tk.fail("missing " + end);
}

}  // namespace

std::shared_ptr<Model> readMsh(const std::string &path)
{
  std::ifstream in(path);
  if(!in) throw std::runtime_error("cannot open " + path);
  Tokens tk(in);
  std::map<std::pair<int, int>, std::string> names;       // (dim, physical tag) -> name
  std::map<std::pair<int, int>, std::vector<int>> phys;   // (dim, physical tag) -> entities
  std::vector<int64_t> ntag;                               // node tags in file order
  std::vector<double> nxyz;
  struct RawBlock {
    int dim, tag, type;
    std::vector<int64_t> conn;  // file node tags
  };
  std::vector<RawBlock> raw;
  bool format = false;
  std::string t;
  while(tk.next(t)) {
    if(t == "$MeshFormat") {
      const double version = tk.real("version");
      const long long ftype = tk.integer("file type");
      tk.integer("data size");
      if(version < 4.1 - 1e-9 || version >= 5.)
        tk.fail("MSH version " + std::to_string(version) + " (4.1 is read)");
      if(ftype != 0) tk.fail("binary MSH files are not read (ASCII only)");
      tk.expect("$EndMeshFormat");
      format = true;
    }
    else if(t == "$PhysicalNames") {
      const long long n = tk.integer("number of physical names");
      for(long long k = 0; k < n; ++k) {
        const int d = (int)tk.integer("physical dimension");
        const int pt = (int)tk.integer("physical tag");
        std::string r = tk.rest();
        const size_t a = r.find('"'), b = r.rfind('"');
        names[{d, pt}] = (a != std::string::npos && b > a) ? r.substr(a + 1, b - a - 1) : "";
      }
      tk.expect("$EndPhysicalNames");
    }
    else if(t == "$Entities") {
      long long count[4];
      for(auto &c : count) c = tk.integer("number of entities");
      for(int d = 0; d < 4; ++d)
        for(long long k = 0; k < count[d]; ++k) {
          const int tag = (int)tk.integer("entity tag");
          for(int q = 0; q < (d == 0 ? 3 : 6); ++q) tk.real("entity coordinates");
          const long long np = tk.integer("number of physical tags");
          for(long long q = 0; q < np; ++q) {
            const int pt = (int)std::llabs(tk.integer("physical tag"));
            phys[{d, pt}].push_back(tag);
          }
          if(d > 0) {
            const long long nb = tk.integer("number of bounding entities");
            for(long long q = 0; q < nb; ++q) tk.integer("bounding entity");
          }
        }
      tk.expect("$EndEntities");
    }
    else if(t == "$Nodes") {
      const long long nblocks = tk.integer("node blocks");
      const long long nnodes = tk.integer("number of nodes");
      tk.integer("min node tag");
      tk.integer("max node tag");
      ntag.reserve((size_t)nnodes);
      nxyz.reserve(3 * (size_t)nnodes);
      for(long long b = 0; b < nblocks; ++b) {
        tk.integer("entity dimension");
        tk.integer("entity tag");
        const long long param = tk.integer("parametric flag");
        const long long n = tk.integer("nodes in block");
        const size_t first = ntag.size();
        for(long long k = 0; k < n; ++k) ntag.push_back(tk.integer("node tag"));
        for(long long k = 0; k < n; ++k) {
          for(int c = 0; c < 3; ++c) nxyz.push_back(tk.real("node coordinate"));
          if(param) tk.fail("parametric node coordinates are not read");
        }
        (void)first;
      }
      if((long long)ntag.size() != nnodes) tk.fail("node count does not match the header");
      tk.expect("$EndNodes");
    }
    else if(t == "$Elements") {
      const long long nblocks = tk.integer("element blocks");
      tk.integer("number of elements");
      tk.integer("min element tag");
      tk.integer("max element tag");
      for(long long b = 0; b < nblocks; ++b) {
        RawBlock rb;
        rb.dim = (int)tk.integer("entity dimension");
        rb.tag = (int)tk.integer("entity tag");
        rb.type = (int)tk.integer("element type");
        const long long n = tk.integer("elements in block");
        gfem::ElemType et;
        if(!gfem::elem_type_info(rb.type, et)) tk.fail("element type " + std::to_string(rb.type) + " is not supported");
        rb.conn.reserve((size_t)(n * et.npe));
        for(long long k = 0; k < n; ++k) {
          tk.integer("element tag");
          for(int q = 0; q < et.npe; ++q) rb.conn.push_back(tk.integer("element node"));
        }
        raw.push_back(std::move(rb));
      }
      tk.expect("$EndElements");
    }
    else if(!t.empty() && t[0] == '$') {
      skipSection(tk, t);
    }
    else {
      tk.fail("unexpected '" + t + "'");
    }
  }
  if(!format) throw std::runtime_error(path + ": no $MeshFormat section");
  // compact node tags to 1..n in increasing tag order
  std::vector<size_t> order(ntag.size());
  for(size_t k = 0; k < order.size(); ++k) order[k] = k;
  std::sort(order.begin(), order.end(), [&](size_t a, size_t b) { return ntag[a] < ntag[b]; });
  std::unordered_map<int64_t, int64_t> remap;
  remap.reserve(ntag.size() * 2);
  std::vector<double> xyz(3 * ntag.size());
  std::vector<int64_t> fileTags(ntag.size());
  bool identity = true;
  for(size_t r = 0; r < order.size(); ++r) {
    const size_t k = order[r];
    if(!remap.emplace(ntag[k], (int64_t)r + 1).second)
      throw std::runtime_error(path + ": node tag " + std::to_string(ntag[k]) + " appears twice");
    for(int c = 0; c < 3; ++c) xyz[3 * r + c] = nxyz[3 * k + c];
    fileTags[r] = ntag[k];
    identity = identity && ntag[k] == (int64_t)r + 1;
  }
  std::vector<Block> blocks;
  int dim = 0;
  for(auto &rb : raw) {
    Block b;
    b.dim = rb.dim;
    b.tag = rb.tag;
    b.type = rb.type;
    b.conn.resize(rb.conn.size());
    for(size_t k = 0; k < rb.conn.size(); ++k) {
      const auto it = remap.find(rb.conn[k]);
      if(it == remap.end())
        throw std::runtime_error(path + ": element node " + std::to_string(rb.conn[k]) + " is not in $Nodes");
      b.conn[k] = it->second;
    }
    dim = std::max(dim, rb.dim);
    // merge blocks of the same (entity, type)
    bool merged = false;
    for(auto &o : blocks)
      if(o.dim == b.dim && o.tag == b.tag && o.type == b.type) {
        o.conn.insert(o.conn.end(), b.conn.begin(), b.conn.end());
        merged = true;
      }
    if(!merged) blocks.push_back(std::move(b));
  }
  std::string name = path;
  const size_t slash = name.find_last_of('/');
  if(slash != std::string::npos) name = name.substr(slash + 1);
  auto m = fromArrays(name, dim, xyz, blocks);
  if(!identity) m->fileNodeTags = std::move(fileTags);
  for(const auto &p : phys) {
    Physical ph;
    ph.dim = p.first.first;
    ph.tag = p.first.second;
    const auto nm = names.find(p.first);
    if(nm != names.end()) ph.name = nm->second;
    ph.entities = p.second;
    std::sort(ph.entities.begin(), ph.entities.end());
    m->physicals.push_back(std::move(ph));
  }
  return m;
}

void writeMsh(const Model &m, const std::string &path)
{
  std::ofstream out(path);
  if(!out) throw std::runtime_error("cannot write " + path);
  out.precision(17);
  out << "$MeshFormat\n4.1 0 8\n$EndMeshFormat\n";
  if(!m.physicals.empty()) {
    out << "$PhysicalNames\n" << m.physicals.size() << "\n";
    for(const auto &p : m.physicals) out << p.dim << " " << p.tag << " \"" << p.name << "\"\n";
    out << "$EndPhysicalNames\n";
  }
  const auto ents = m.entities();
  const int64_t nn = m.numNodes();
  auto fileTag = [&](int64_t t) { return m.fileNodeTags.empty() ? t : m.fileNodeTags[(size_t)t - 1]; };
  // entities: bounding boxes of their nodes, physical tags
  out << "$Entities\n";
  int count[4] = {0, 0, 0, 0};
  for(const auto &e : ents) count[e.first]++;
  out << count[0] << " " << count[1] << " " << count[2] << " " << count[3] << "\n";
  for(const auto &e : ents) {
    double mn[3] = {1e300, 1e300, 1e300}, mx[3] = {-1e300, -1e300, -1e300};
    for(int type : m.elementTypes(e.first, e.second)) {
      const Block *b = m.block(e.first, e.second, type);
      for(int64_t t : b->conn)
        for(int c = 0; c < 3; ++c) {
          mn[c] = std::min(mn[c], m.xyz[3 * (t - 1) + c]);
          mx[c] = std::max(mx[c], m.xyz[3 * (t - 1) + c]);
        }
    }
    std::vector<int> pt;
    for(const auto &p : m.physicals)
      if(p.dim == e.first && std::find(p.entities.begin(), p.entities.end(), e.second) != p.entities.end())
        pt.push_back(p.tag);
    out << e.second << " " << mn[0] << " " << mn[1] << " " << mn[2];
    if(e.first > 0) out << " " << mx[0] << " " << mx[1] << " " << mx[2];
    out << " " << pt.size();
    for(int p : pt) out << " " << p;
    if(e.first > 0) out << " 0";  // bounding entities are not kept in this model
    out << "\n";
  }
  out << "$EndEntities\n";
  // nodes, all in one block on the first entity of the highest dimension
  const std::pair<int, int> home = ents.empty() ? std::make_pair(0, 1) : ents.back();
  int64_t mnTag = nn ? fileTag(1) : 0, mxTag = 0;
  for(int64_t t = 1; t <= nn; ++t) {
    mnTag = std::min(mnTag, fileTag(t));
    mxTag = std::max(mxTag, fileTag(t));
  }
  out << "$Nodes\n1 " << nn << " " << mnTag << " " << mxTag << "\n";
  out << home.first << " " << home.second << " 0 " << nn << "\n";
  for(int64_t t = 1; t <= nn; ++t) out << fileTag(t) << "\n";
  for(int64_t t = 1; t <= nn; ++t)
    out << m.xyz[3 * (t - 1)] << " " << m.xyz[3 * (t - 1) + 1] << " " << m.xyz[3 * (t - 1) + 2] << "\n";
  out << "$EndNodes\n";
  int64_t ne = 0;
  for(const auto &b : m.blocks) ne += b.size();
  out << "$Elements\n" << m.blocks.size() << " " << ne << " " << (ne ? 1 : 0) << " " << ne << "\n";
  int64_t et = 1;
  for(const auto &e : ents)
    for(int type : m.elementTypes(e.first, e.second)) {
      const Block *b = m.block(e.first, e.second, type);
      out << b->dim << " " << b->tag << " " << b->type << " " << b->size() << "\n";
      for(int64_t k = 0; k < b->size(); ++k) {
        out << et++;
        for(int q = 0; q < b->npe; ++q) out << " " << fileTag(b->conn[(size_t)(k * b->npe + q)]);
        out << "\n";
      }
    }
  out << "$EndElements\n";
  if(!out) throw std::runtime_error("write failed: " + path);
}

}  // namespace gmshfem::model
