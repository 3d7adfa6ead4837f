// Host-side topology of one rank: node list, positions, neighbour index,
// node ranks and types, bulk list, and the .vtklb point attributes.
#pragma once
#include <cstdint>
#include <map>
#include <string>
#include <utility>
#include <vector>

struct lb_topo {
  int lattice = 0;   // LB_D2Q9 / LB_D3Q19
  int nD = 0, Q = 0;
  int rank = 0;      // 0-based rank of this file / node list
  bool zero_ghost = true;
  std::vector<int32_t> global_dims;  // padded dimensions (GLOBAL_DIMENSIONS)
  int32_t n_nodes = 0;               // incl. node 0 when zero_ghost
  std::vector<int32_t> pos;          // [n_nodes*nD], node 0 = -1
  std::vector<int32_t> neighbor;     // [n_nodes*Q], program basis order
  std::vector<int32_t> nodetype_attr;  // raw 'nodetype' attribute (0 solid / 1 fluid)
  std::vector<int32_t> node_rank;    // [n_nodes]
  std::vector<int16_t> node_type;    // [n_nodes], Nodes codes -1..4
  std::vector<int32_t> bulk;         // ascending node numbers
  // PROCESSOR blocks in file order: (neighbour rank, [(local node, node on that rank)])
  std::vector<std::pair<int, std::vector<std::pair<int32_t, int32_t>>>> processors;
  // SCALARS blocks: name -> (type, raw tokens, one per point in file order)
  std::map<std::string, std::pair<std::string, std::vector<std::string>>> attributes;
  // SCALARS blocks read from the binary twin: name -> type and the values
  // read as double and as float, one per point in file order
  struct AttrValues {
    std::string type;
    std::vector<double> d;
    std::vector<float> f;
  };
  std::map<std::string, AttrValues> attr_values;
};

namespace lbamd {
// Nodes::setupNodeType (LBnodes.h:141-209) followed by findBulkNodes
// (LBgeometry.h:11-21).  Needs neighbor, nodetype_attr, node_rank, rank.
void finish_topology(lb_topo& t);
}  // namespace lbamd
