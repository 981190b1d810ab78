// Mesh model: the part of the gmsh model API the assembly path reads
// (nodes, entities, element blocks, Lagrange keys), plus the seeded
// structured generators used by the workloads.
//
// Reference call sites served by this model:
//   gmsh::model::getEntities / mesh::getElementTypes   problem/Formulation.cpp:1374-1397, 1430-1438
//   gmsh::model::mesh::getKeys("Lagrange")             field/FunctionSpaceInterface.cpp:319-331
//   gmsh::model::getBoundingBox                        dofs/DofsManager.cpp:127
//   gmsh::model::getDimension                          problem/Formulation.cpp:992
#pragma once

#include <cstdint>
#include <memory>
#include <string>
#include <utility>
#include <vector>

namespace gmshfem::model {

// All elements of one element type classified on one entity (dim, tag).
struct Block {
  int dim = 0, tag = 0, type = 0, order = 0, npe = 0;
  std::vector<int64_t> conn;  // node tags (1-based), element-major, gmsh local node order
  int64_t size() const { return npe ? (int64_t)conn.size() / npe : 0; }
};

// A physical group: a named set of entities of one dimension.
struct Physical {
  int dim = 0, tag = 0;
  std::string name;
  std::vector<int> entities;
};

struct Model {
  std::string name;
  int dim = 0;
  std::vector<double> xyz;  // node tag t -> xyz[3(t-1) .. 3(t-1)+2]
  std::vector<Block> blocks;
  std::vector<Physical> physicals;
  std::vector<int64_t> fileNodeTags;  // tags of nodes 1..n in the file they came from (empty: the same)

  int64_t numNodes() const { return (int64_t)xyz.size() / 3; }
  // entities in gmsh order: dimension ascending, tag ascending
  std::vector<std::pair<int, int>> entities(int dim = -1) const;
  // element types present on an entity (ascending type code)
  std::vector<int> elementTypes(int dim, int tag) const;
  std::vector<int> elementTypes(int dim) const;
  const Block *block(int dim, int tag, int type) const;
  void boundingBox(double mn[3], double mx[3]) const;
  // entities of physical group (dim, tag); empty when there is none
  std::vector<int> physicalEntities(int dim, int tag) const;
  const Physical *physical(const std::string &name) const;
  uint64_t uid = 0;  // unique id, changes with every new model
};

// Unit square [0,1]^2, n x n cells, each split along the (0,0)-(1,1) cell
// diagonal into two triangles; Lagrange nodes of the given order; interior
// vertices jittered by U(-jitter*h, jitter*h) per coordinate (mt19937_64).
// Entities: surface (2,1); curves (1,1) y=0, (1,2) x=1, (1,3) y=1, (1,4) x=0;
// points (0,1..4).
std::shared_ptr<Model> unitSquare(int n, int order, double jitter, uint64_t seed);

// Unit cube [0,1]^3, n^3 cells, each split into 6 positively oriented
// tetrahedra around the (0,0,0)-(1,1,1) diagonal; boundary triangles on the
// six faces.  Entities: volume (3,1); surfaces (2,1) x=0, (2,2) x=1,
// (2,3) y=0, (2,4) y=1, (2,5) z=0, (2,6) z=1.
std::shared_ptr<Model> unitCube(int n, int order, double jitter, uint64_t seed);

// Build a model from raw arrays (node coordinates + element blocks).
std::shared_ptr<Model> fromArrays(const std::string &name, int dim, const std::vector<double> &xyz,
                                  const std::vector<Block> &blocks);

// Gmsh MSH 4.1 files, ASCII (the format gmsh::open / gmsh::write exchange with
// GmshFEM; element node order is gmsh's, as everywhere in this model).  The
// reader takes $PhysicalNames, $Entities (physical tags), $Nodes and
// $Elements and skips other sections; node tags are compacted to 1..n in
// increasing order (the file's tags are kept in fileNodeTags).  Errors throw
// std::runtime_error with the line number.
std::shared_ptr<Model> readMsh(const std::string &path);
void writeMsh(const Model &m, const std::string &path);

// the "current" model, as gmsh::model keeps one
void setCurrent(std::shared_ptr<Model> m);
std::shared_ptr<Model> current();

}  // namespace gmshfem::model
