// MutableGraph.hpp -- the host graph the apps start from (src/MutableGraph.hpp:49-80)
// and its binary loader/writer (src/MutableGraph.hpp:298-360).
//
// Undirected file format: size_t n, then per node { uint64 external id; size_t
// degree; uint32 neighbours[degree] }, neighbours sorted and already degree-
// relabelled (the reference's tools do that offline).  Directed files
// (src/MutableGraph.hpp:636-714) hold per node the out-neighbours and then the
// in-neighbours, each as { size_t count; uint32 ids[count] }.  The edge-list and
// attribute loaders are not provided.
#pragma once

#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <iostream>
#include <string>
#include <vector>

struct MutableGraph {
  size_t num_nodes;
  size_t num_edges;
  size_t max_nbrhood_size;
  bool symmetric;
  std::vector<uint64_t> *id_map;
  std::vector<uint32_t> *node_attr;
  std::vector<std::vector<uint32_t> *> *out_neighborhoods;
  std::vector<std::vector<uint32_t> *> *in_neighborhoods;
  std::vector<std::vector<uint32_t> *> *out_edge_attributes;
  std::vector<std::vector<uint32_t> *> *in_edge_attributes;

  MutableGraph(size_t num_nodes_in, size_t num_edges_in, size_t max_nbrhood_size_in, bool symmetric_in,
               std::vector<uint64_t> *id_map_in, std::vector<uint32_t> *node_attr_in,
               std::vector<std::vector<uint32_t> *> *out_neighborhoods_in,
               std::vector<std::vector<uint32_t> *> *in_neighborhoods_in,
               std::vector<std::vector<uint32_t> *> *out_edge_attributes_in,
               std::vector<std::vector<uint32_t> *> *in_edge_attributes_in)
      : num_nodes(num_nodes_in), num_edges(num_edges_in), max_nbrhood_size(max_nbrhood_size_in),
        symmetric(symmetric_in), id_map(id_map_in), node_attr(node_attr_in),
        out_neighborhoods(out_neighborhoods_in), in_neighborhoods(in_neighborhoods_in),
        out_edge_attributes(out_edge_attributes_in), in_edge_attributes(in_edge_attributes_in) {}

  ~MutableGraph() {
    for (auto *row : *out_neighborhoods) delete row;
    if (in_neighborhoods != out_neighborhoods) {
      for (auto *row : *in_neighborhoods) delete row;
      delete in_neighborhoods;
    }
    delete out_neighborhoods;
    delete id_map;
  }

  // Host CSR of the out-neighbourhoods, followed for a directed graph by the
  // in-neighbourhoods as rows n..2n-1 (row_off: rows+1 entries).
  void to_csr(std::vector<uint64_t> &row_off, std::vector<uint32_t> &nbrs) const {
    const size_t rows = symmetric ? num_nodes : 2 * num_nodes;
    row_off.assign(rows + 1, 0);
    nbrs.clear();
    nbrs.reserve(num_edges);
    for (size_t i = 0; i < rows; i++) {
      const std::vector<uint32_t> &r =
          i < num_nodes ? *out_neighborhoods->at(i) : *in_neighborhoods->at(i - num_nodes);
      nbrs.insert(nbrs.end(), r.begin(), r.end());
      row_off[i + 1] = nbrs.size();
    }
  }

  void writeUndirectedToBinary(const std::string &path) const {
    std::ofstream out(path, std::ios::binary | std::ios::out);
    if (!out) {
      std::cerr << "ERROR: cannot write " << path << std::endl;
      std::exit(1);
    }
    const size_t n = out_neighborhoods->size();
    out.write(reinterpret_cast<const char *>(&n), sizeof(n));
    for (size_t i = 0; i < n; i++) {
      const std::vector<uint32_t> &r = *out_neighborhoods->at(i);
      const size_t d = r.size();
      out.write(reinterpret_cast<const char *>(&id_map->at(i)), sizeof(uint64_t));
      out.write(reinterpret_cast<const char *>(&d), sizeof(d));
      out.write(reinterpret_cast<const char *>(r.data()), (std::streamsize)(sizeof(uint32_t) * d));
    }
  }

  static MutableGraph *directedFromBinary(const std::string &path) {
    std::ifstream in(path, std::ios::binary | std::ios::in);
    if (!in) {
      std::cerr << "ERROR: cannot open " << path << std::endl;
      std::exit(1);
    }
    auto *id_map = new std::vector<uint64_t>();
    auto *outs = new std::vector<std::vector<uint32_t> *>();
    auto *ins = new std::vector<std::vector<uint32_t> *>();
    size_t n = 0, edges = 0, maxd = 0;
    in.read(reinterpret_cast<char *>(&n), sizeof(n));
    for (size_t i = 0; in && i < n; i++) {
      uint64_t ext = 0;
      in.read(reinterpret_cast<char *>(&ext), sizeof(ext));
      id_map->push_back(ext);
      for (auto *side : {outs, ins}) {
        size_t d = 0;
        in.read(reinterpret_cast<char *>(&d), sizeof(d));
        if (!in || d > n) {
          std::cerr << "ERROR: malformed graph file " << path << " at node " << i << std::endl;
          std::exit(1);
        }
        auto *row = new std::vector<uint32_t>(d);
        in.read(reinterpret_cast<char *>(row->data()), (std::streamsize)(sizeof(uint32_t) * d));
        side->push_back(row);
        edges += d;
        if (d > maxd) maxd = d;
      }
    }
    if (!in || outs->size() != n) {
      std::cerr << "ERROR: truncated graph file " << path << std::endl;
      std::exit(1);
    }
    return new MutableGraph(n, edges, maxd, false, id_map, nullptr, outs, ins, nullptr, nullptr);
  }

  static MutableGraph *undirectedFromBinary(const std::string &path) {
    std::ifstream in(path, std::ios::binary | std::ios::in);
    if (!in) {
      std::cerr << "ERROR: cannot open " << path << std::endl;
      std::exit(1);
    }
    auto *id_map = new std::vector<uint64_t>();
    auto *rows = new std::vector<std::vector<uint32_t> *>();
    size_t n = 0, edges = 0, maxd = 0;
    in.read(reinterpret_cast<char *>(&n), sizeof(n));
    for (size_t i = 0; in && i < n; i++) {
      uint64_t ext = 0;
      size_t d = 0;
      in.read(reinterpret_cast<char *>(&ext), sizeof(ext));
      in.read(reinterpret_cast<char *>(&d), sizeof(d));
      if (!in || d > n) {
        std::cerr << "ERROR: malformed graph file " << path << " at node " << i << std::endl;
        std::exit(1);
      }
      auto *row = new std::vector<uint32_t>(d);
      in.read(reinterpret_cast<char *>(row->data()), (std::streamsize)(sizeof(uint32_t) * d));
      id_map->push_back(ext);
      rows->push_back(row);
      edges += d;
      if (d > maxd) maxd = d;
    }
    if (!in || rows->size() != n) {
      std::cerr << "ERROR: truncated graph file " << path << std::endl;
      std::exit(1);
    }
    return new MutableGraph(n, edges, maxd, true, id_map, nullptr, rows, rows, nullptr, nullptr);
  }
};
