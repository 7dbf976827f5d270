// MutableGraph.hpp -- the host graph the apps start from (src/MutableGraph.hpp:49-80),
// its binary loader/writer (src/MutableGraph.hpp:298-360), the undirected edge-list
// and directed edge-list loaders (src/MutableGraph.hpp:517-598,893-981), the
// directed writer (src/MutableGraph.hpp:636-656) and the vertex orderings
// (src/MutableGraph.hpp:126-285), which run on the device (eh_graph_reorder).
//
// Undirected file format: size_t n, then per node { uint64 external id; size_t
// degree; uint32 neighbours[degree] }, neighbours sorted and already renumbered
// (tools/undirectedEdgeListToBinary does that offline).  Directed files
// (src/MutableGraph.hpp:636-714) hold per node the out-neighbours and then the
// in-neighbours, each as { size_t count; uint32 ids[count] }.  The attribute
// loaders are not provided.
#pragma once

#include <algorithm>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <iostream>
#include <iterator>
#include <mutex>
#include <string>
#include <unordered_map>
#include <vector>

#include "runtime.hpp"

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

  // Text pairs "src dst" of 64-bit external ids, separated by spaces, tabs, line
  // breaks or 'A' (src/MutableGraph.hpp:547-556).  Calls pair(src, dst) per pair;
  // a token that is not a number or a trailing half pair is an error.
  template <class F>
  static void read_edge_pairs(const std::string &path, F &&pair) {
    std::ifstream in(path, std::ios::binary | std::ios::in);
    if (!in) {
      std::cerr << "ERROR: cannot open " << path << std::endl;
      std::exit(1);
    }
    const std::string text((std::istreambuf_iterator<char>(in)), std::istreambuf_iterator<char>());
    auto is_sep = [](char c) { return c == ' ' || c == '\t' || c == '\n' || c == '\r' || c == 'A'; };
    size_t pos = 0;
    uint64_t pending[2];
    int have = 0;
    while (true) {
      while (pos < text.size() && is_sep(text[pos])) pos++;
      if (pos == text.size()) break;
      const size_t start = pos;
      uint64_t v = 0;
      while (pos < text.size() && text[pos] >= '0' && text[pos] <= '9') v = 10 * v + (uint64_t)(text[pos++] - '0');
      if (pos == start || (pos < text.size() && !is_sep(text[pos]))) {
        std::cerr << "ERROR: malformed edge list " << path << " at byte " << start << std::endl;
        std::exit(1);
      }
      pending[have++] = v;
      if (have == 2) {
        have = 0;
        pair(pending[0], pending[1]);
      }
    }
    if (have != 0) {
      std::cerr << "ERROR: edge list " << path << " ends in half a pair" << std::endl;
      std::exit(1);
    }
  }

  static void sort_rows(std::vector<std::vector<uint32_t> *> &rows, size_t &edges, size_t &maxd) {
    for (auto *row : rows) {
      std::sort(row->begin(), row->end());
      row->erase(std::unique(row->begin(), row->end()), row->end());
      edges += row->size();
      maxd = std::max(maxd, row->size());
    }
  }

  // src/MutableGraph.hpp:517-598.  Internal ids follow first appearance; every pair
  // is stored in both directions; rows are sorted and deduplicated.  Unlike the
  // reference, a self-loop is dropped (device graphs are simple) and a malformed
  // file is an error.
  static MutableGraph *undirectedFromEdgeList(const std::string &path) {
    std::cout << path << std::endl;
    auto *id_map = new std::vector<uint64_t>();
    auto *rows = new std::vector<std::vector<uint32_t> *>();
    std::unordered_map<uint64_t, uint32_t> ids;
    auto intern = [&](uint64_t ext) -> uint32_t {
      auto it = ids.find(ext);
      if (it != ids.end()) return it->second;
      const uint32_t id = (uint32_t)rows->size();
      ids.emplace(ext, id);
      id_map->push_back(ext);
      rows->push_back(new std::vector<uint32_t>());
      return id;
    };
    read_edge_pairs(path, [&](uint64_t s, uint64_t d) {
      const uint32_t a = intern(s), b = intern(d);
      if (a == b) return;
      rows->at(a)->push_back(b);
      rows->at(b)->push_back(a);
    });
    size_t edges = 0, maxd = 0;
    sort_rows(*rows, edges, maxd);
    return new MutableGraph(rows->size(), edges, maxd, true, id_map, nullptr, rows, rows, nullptr, nullptr);
  }

  // src/MutableGraph.hpp:893-981: out-rows and in-rows, ids by first appearance,
  // both sorted and deduplicated; self-loops dropped as above.  num_edges counts
  // distinct arcs (the reference counts input lines).
  static MutableGraph *directedFromEdgeList(const std::string &path) {
    auto *id_map = new std::vector<uint64_t>();
    auto *outs = new std::vector<std::vector<uint32_t> *>();
    auto *ins = new std::vector<std::vector<uint32_t> *>();
    std::unordered_map<uint64_t, uint32_t> ids;
    auto intern = [&](uint64_t ext) -> uint32_t {
      auto it = ids.find(ext);
      if (it != ids.end()) return it->second;
      const uint32_t id = (uint32_t)outs->size();
      ids.emplace(ext, id);
      id_map->push_back(ext);
      outs->push_back(new std::vector<uint32_t>());
      ins->push_back(new std::vector<uint32_t>());
      return id;
    };
    read_edge_pairs(path, [&](uint64_t s, uint64_t d) {
      const uint32_t a = intern(s), b = intern(d);
      if (a == b) return;
      outs->at(a)->push_back(b);
      ins->at(b)->push_back(a);
    });
    size_t arcs = 0, in_arcs = 0, maxd = 0;
    sort_rows(*outs, arcs, maxd);
    sort_rows(*ins, in_arcs, maxd);
    return new MutableGraph(outs->size(), arcs, maxd, false, id_map, nullptr, outs, ins, nullptr, nullptr);
  }

  // src/MutableGraph.hpp:636-656: per node { uint64 external id; size_t out count;
  // out ids; size_t in count; in ids }.
  void writeDirectedToBinary(const std::string &path) const {
    std::ofstream out(path, std::ios::binary | std::ios::out);
    if (!out) {
      std::cerr << "ERROR: cannot write " << path << std::endl;
      std::exit(1);
    }
    out.write(reinterpret_cast<const char *>(&num_nodes), sizeof(num_nodes));
    for (size_t i = 0; i < num_nodes; i++) {
      out.write(reinterpret_cast<const char *>(&id_map->at(i)), sizeof(uint64_t));
      for (const auto *side : {out_neighborhoods, in_neighborhoods}) {
        const std::vector<uint32_t> &r = *side->at(i);
        const size_t d = r.size();
        out.write(reinterpret_cast<const char *>(&d), sizeof(d));
        out.write(reinterpret_cast<const char *>(r.data()), (std::streamsize)(sizeof(uint32_t) * d));
      }
    }
  }

  // The vertex orderings (src/MutableGraph.hpp:126-285).  The renumbering runs on
  // the device; the rows and id_map here are then replaced by the renumbered ones.
  // The reference shuffles with std::random_shuffle; here `seed` makes BFS,
  // shingles, random and the_game reproducible.
  void reorder_bfs(uint64_t seed = 0) { reorder(EH_ORDER_BFS, seed); }
  void reorder_strong_run() { reorder(EH_ORDER_STRONG_RUN, 0); }
  void reorder_random(uint64_t seed = 0) { reorder(EH_ORDER_RANDOM, seed); }
  void reorder_by_shingles(uint64_t seed = 0) { reorder(EH_ORDER_SHINGLES, seed); }
  void reorder_by_degree() { reorder(EH_ORDER_DEGREE, 0); }
  void reorder_by_rev_degree() { reorder(EH_ORDER_REV_DEGREE, 0); }
  void reorder_by_the_game(uint64_t seed = 0) { reorder(EH_ORDER_THE_GAME, seed); }

  void reorder(eh_ordering ordering, uint64_t seed) {
    if (!symmetric) throw emptyheaded::Error(EH_ERR_UNSUPPORTED, "MutableGraph::reorder: undirected graphs only");
    if (num_nodes == 0) return;
    std::vector<uint64_t> off;
    std::vector<uint32_t> nbr;
    to_csr(off, nbr);
    std::vector<uint32_t> new2old(num_nodes);
    {
      emptyheaded::Runtime &rt = emptyheaded::Runtime::get();
      std::lock_guard<std::mutex> lock(rt.mutex());
      eh_graph *g = nullptr;
      emptyheaded::check(eh_graph_from_csr(rt.ctx(), (uint32_t)num_nodes, off.data(), nbr.data(), &g),
                         "MutableGraph::reorder");
      eh_status st = eh_graph_reorder(g, ordering, seed);
      if (st == EH_OK) st = eh_graph_download(g, off.data(), nbr.data());
      if (st == EH_OK) st = eh_graph_download_perm(g, new2old.data());
      eh_graph_free(g);
      emptyheaded::check(st, "MutableGraph::reorder");
    }
    std::vector<uint64_t> ext(num_nodes);
    for (size_t r = 0; r < num_nodes; r++) {
      ext[r] = id_map->at(new2old[r]);
      out_neighborhoods->at(r)->assign(nbr.begin() + (std::ptrdiff_t)off[r], nbr.begin() + (std::ptrdiff_t)off[r + 1]);
    }
    id_map->swap(ext);
  }
};
