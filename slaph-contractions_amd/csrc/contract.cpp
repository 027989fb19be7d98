// contract(): the correlator assembly of the reference (Correlators.cpp:40-127,
// Diagram.cpp:9-147, DilutedTraceFactory.cpp:13-238) as a host planner plus
// batched device launches.
//
// Planning splits every index into a time part and a quantum-number part:
//  * the quantum-number part (which operator entries, which random-vector
//    combinations, which combinations survive the used-index masks) does not
//    depend on time and is enumerated once per contract() call;
//  * the time part (which factor slab a leaf reads, which product instance a
//    slice pair shares) is a handful of integers per slice pair.
// The device kernels compose the two through small index tables, so host work
// per chunk of slice pairs is O(slice pairs + trace ids).
#include "engine.hpp"

#include <algorithm>
#include <array>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <functional>
#include <map>
#include <set>
#include <unordered_map>

namespace slaph {

namespace {

// ----- shapes of the trace factories ----------------------------------------
// A leaf reads factor type q (0,1,2) for operator entry c_look[cidx] at a time
// key built from the trace times: Q0 (t[a]); Q1 (t[a], B(t[b]));
// Q2 (B(t[a]), t[b], B(t[c])).
struct LeafSpec {
  int q, cidx, a, b, c;
};
enum NodeKind { kLeaf = 0, kProd = 1, kTrace2 = 2, kTrace1 = 3 };
struct NodeSpec {
  NodeKind kind;
  int leaf;         // for kLeaf
  int left, right;  // node indices
};
struct Shape {
  int ntimes;
  std::vector<LeafSpec> leaves;
  std::vector<NodeSpec> nodes;  // topological; last is the trace
};

Shape make_shape(std::string const &name) {
  Shape s;
  auto leaf = [&s](int q, int cidx, int a, int b = 0, int c = 0) {
    s.leaves.push_back({q, cidx, a, b, c});
    s.nodes.push_back({kLeaf, int(s.leaves.size()) - 1, -1, -1});
    return int(s.nodes.size()) - 1;
  };
  auto prod = [&s](int l, int r) {
    s.nodes.push_back({kProd, -1, l, r});
    return int(s.nodes.size()) - 1;
  };
  auto tr2 = [&s](int l, int r) { s.nodes.push_back({kTrace2, -1, l, r}); };
  if (name == "trQ1") {  // DilutedTraceFactory.cpp:13-25
    s.ntimes = 1;
    int x = leaf(1, 0, 0, 0);
    s.nodes.push_back({kTrace1, -1, x, -1});
  } else if (name == "trQ1Q1") {  // :29-43
    s.ntimes = 2;
    int x = leaf(1, 0, 0, 1), y = leaf(1, 1, 1, 0);
    tr2(x, y);
  } else if (name == "trQ0Q2") {  // :47-61
    s.ntimes = 2;
    int x = leaf(0, 1, 1), y = leaf(2, 0, 1, 0, 1);
    tr2(x, y);
  } else if (name == "trQ1Q1Q1") {  // :65-89
    s.ntimes = 3;
    int a = leaf(1, 0, 0, 1), b = leaf(1, 1, 1, 2), p = prod(a, b), c = leaf(1, 2, 2, 0);
    tr2(p, c);
  } else if (name == "trQ1Q0Q2") {  // :95-119
    s.ntimes = 3;
    int a = leaf(0, 2, 0), b = leaf(2, 0, 0, 0, 1), p = prod(a, b), c = leaf(1, 1, 1, 2);
    tr2(p, c);
  } else if (name == "trQ1Q1Q1Q1") {  // :125-155
    s.ntimes = 4;
    int a = leaf(1, 0, 0, 1), b = leaf(1, 1, 1, 2), p = prod(a, b);
    int c = leaf(1, 2, 2, 3), d = leaf(1, 3, 3, 0), q = prod(c, d);
    tr2(p, q);
  } else if (name == "trQ2Q0Q2Q0") {  // :161-190
    s.ntimes = 4;
    int a = leaf(0, 0, 0), b = leaf(2, 1, 0, 1, 2), p = prod(a, b);
    int c = leaf(0, 2, 2), d = leaf(2, 3, 2, 3, 0), q = prod(c, d);
    tr2(p, q);
  } else if (name == "trQ2Q0Q2Q0Q2Q0") {  // :197-238
    s.ntimes = 6;
    int f0 = leaf(2, 1, 5, 0, 1), f1 = leaf(0, 2, 1), p = prod(f0, f1);
    int f2 = leaf(2, 3, 1, 2, 3);
    p = prod(p, f2);
    int f3 = leaf(0, 4, 3);
    p = prod(p, f3);
    int f4 = leaf(2, 5, 3, 4, 5);
    p = prod(p, f4);
    int f5 = leaf(0, 0, 5);
    tr2(p, f5);
  } else {
    throw std::runtime_error("unknown trace " + name);
  }
  return s;
}

// ----- random-vector templates ---------------------------------------------
struct FactorTmpl {
  std::vector<std::pair<int, int>> ric;
  std::vector<uint16_t> used;
};
struct PairTmpl {
  std::vector<std::array<int, 2>> combos;
  std::vector<uint16_t> masks;  // traces only
  int out = -1;                 // products only: output factor template
};

class Templates {
 public:
  int leaf(std::vector<std::pair<int, int>> const &rnd) {
    auto it = leaf_ids_.find(rnd);
    if (it != leaf_ids_.end()) return it->second;
    FactorTmpl t;
    for (auto const &p : rnd) {
      if (p.first > 15 || p.second > 15) throw std::runtime_error("at most 16 random vectors are supported");
      t.ric.push_back(p);
      t.used.push_back(0);
    }
    f_.push_back(t);
    leaf_ids_[rnd] = int(f_.size()) - 1;
    return int(f_.size()) - 1;
  }
  FactorTmpl const &factor(int id) const { return f_.at(id); }

  // operator*(vector<DilutedFactor>, vector<DilutedFactor>) (DilutedFactor.cpp:6-51)
  PairTmpl const &product(int a, int b) {
    auto key = std::make_pair(a, b);
    auto it = prod_.find(key);
    if (it != prod_.end()) return it->second;
    PairTmpl p;
    FactorTmpl out;
    auto const &L = f_.at(a), &R = f_.at(b);
    for (size_t i = 0; i < L.ric.size(); ++i)
      for (size_t j = 0; j < R.ric.size(); ++j) {
        if (!(L.ric[i].second == R.ric[j].first && L.ric[i].first != R.ric[j].second)) continue;
        if (L.used[i] & R.used[j]) continue;
        p.combos.push_back({int(i), int(j)});
        out.ric.push_back({L.ric[i].first, R.ric[j].second});
        out.used.push_back(uint16_t(L.used[i] | R.used[j] | (1u << L.ric[i].second)));
      }
    if (p.combos.empty())
      throw std::runtime_error(
          "vector<DilutedFactor> operator*(vector<DilutedFactor>, vector<DilutedFactor>) has an empty result.");
    f_.push_back(out);
    p.out = int(f_.size()) - 1;
    return prod_.emplace(key, p).first->second;
  }

  // factor_to_trace(L, R) (DilutedTrace.cpp:84-138)
  PairTmpl const &trace2(int a, int b) {
    auto key = std::make_pair(a, b);
    auto it = tr2_.find(key);
    if (it != tr2_.end()) return it->second;
    PairTmpl p;
    auto const &L = f_.at(a), &R = f_.at(b);
    for (size_t i = 0; i < L.ric.size(); ++i)
      for (size_t j = 0; j < R.ric.size(); ++j) {
        if (!(L.ric[i].second == R.ric[j].first && L.ric[i].first == R.ric[j].second)) continue;
        if (L.used[i] & R.used[j]) continue;
        p.combos.push_back({int(i), int(j)});
        p.masks.push_back(
            uint16_t(L.used[i] | R.used[j] | (1u << L.ric[i].second) | (1u << L.ric[i].first)));
      }
    if (p.combos.empty())
      throw std::runtime_error(
          "vector<DilutedTrace> factor_to_trace(vector<DilutedFactor>, vector<DilutedFactor>) has an empty result. "
          "Not enough random vectors?");
    return tr2_.emplace(key, p).first->second;
  }

  // factor_to_trace(v) (DilutedTrace.cpp:140-165)
  PairTmpl const &trace1(int a) {
    auto it = tr1_.find(a);
    if (it != tr1_.end()) return it->second;
    PairTmpl p;
    auto const &L = f_.at(a);
    for (size_t i = 0; i < L.ric.size(); ++i) {
      if (L.ric[i].first != L.ric[i].second) continue;
      p.combos.push_back({int(i), 0});
      p.masks.push_back(uint16_t(L.used[i] | (1u << L.ric[i].first)));
    }
    if (p.combos.empty())
      throw std::runtime_error("vector<DilutedTrace> factor_to_trace(vector<DilutedFactor>) has an empty result.");
    return tr1_.emplace(a, p).first->second;
  }

 private:
  std::vector<FactorTmpl> f_;
  std::map<std::vector<std::pair<int, int>>, int> leaf_ids_;
  std::map<std::pair<int, int>, PairTmpl> prod_, tr2_;
  std::map<int, PairTmpl> tr1_;
};

// ----- plans -------------------------------------------------------------------
struct NodePlan {
  // quantum-number part (per contract)
  std::map<std::vector<int>, int> qn_index;
  std::vector<std::vector<int>> qn_key;  // leaf entries of the subtree
  std::vector<int> qn_left, qn_right;    // child qn (leaf child: operator entry)
  std::vector<int> qn_tmpl;              // output factor template (leaf/product)
  std::vector<long long> qn_off;         // output offset of each qn
  long long stride = 0;
  std::vector<ComboProd> combos;
  long long cmaxX = 0, cmaxY = 0, cmaxO = 0;  // largest qn offset + combo index per operand
  std::vector<int> trl_qn;  // trl -> qn (in the owner's qn space)
  int owner = -1;           // node whose product table this node reads
  bool transposed = false;  // product table stored as Y^T (a trace's right operand)
  DevBuf d_combos, d_Xq, d_Yq;
  std::vector<long long> h_Xq, h_Yq;
  // time part (per chunk)
  std::map<std::vector<long long>, int> time_index;
  std::vector<int> time_left, time_right;  // child time index (product children)
  std::vector<long long> time_leafbase_l, time_leafbase_r;
  std::vector<int> sp_time;  // slice pair -> time index
  long long base = 0;        // offset in the chunk's product/trace buffer
  DevBuf d_Xt, d_Yt;
  std::vector<long long> h_Xt, h_Yt;
};

struct Family {
  std::string name;
  std::vector<Loc> locs;
  Shape shape;
  std::vector<int> tr_ids;  // trl -> tr_id
  std::map<int, int> trl_of;
  std::vector<NodePlan> nodes;
  std::vector<int> mask_off;  // trl -> offset into the global mask array
  std::vector<int> ncomb;     // trl -> number of trace values
  // mask groups of each trace vector (first-seen order of the masks)
  std::vector<int> ngroup, goff, gmask_off;  // trl -> count, offset in the family's group sums, global mask offset
  std::vector<std::vector<int>> gcount;      // trl -> combinations per group
  std::vector<GroupItem> gitems;
  std::vector<int2> gpairs;  // (left, right) combination indices, grouped
  long long gstride = 0;     // group sums per slice pair
  long long gbase = 0;       // offset of this family in the chunk's group-sum buffer
  DevBuf d_gitems, d_gpairs, d_tile;
  // trl -> earlier trl of equal value (tr(X_a Y_b) = tr(X_b Y_a) when both
  // operands read the same table at the same time), or -1
  std::vector<int> mirror_of;
  bool coincide = false;  // both trace operands read the same factors at every slice pair
  // subset sums of the trace vectors that multi-trace correlators read
  std::vector<int> zslot;  // trl -> slot, -1 if none
  std::vector<ZItem> zitems;
  long long zstride = 0;  // per slice pair
  DevBuf d_zitems;
  // left-stationary trace plan (time independent)
  std::vector<LsBlock> ls_blocks;
  std::vector<LsChunk> ls_chunks;
  std::vector<LsTask> ls_tasks;
  std::vector<long long> ls_slot_dst;  // slot -> partial index
  std::vector<int> ls_left;            // staged left matrices (offset from the time base)
  std::vector<GroupItem> ls_red;       // per item: its partial sums in ls_ridx
  std::vector<int> ls_ridx;
  long long ls_npart = 0;  // partial sums per slice pair
  double ls_bytes = 0;     // algorithmic bytes per slice pair: distinct operand matrices + partial sums
  // queued form: one record per block (one chunk each), see launch_trace_lstat_q
  bool ls_queue = false;
  std::vector<int> ls_units;
  DevBuf d_lsb, d_lsc, d_lst, d_lsd, d_lsl, d_lsr, d_lsri, d_lsu, d_lsq;
  // right-stationary slice-pair tile plan (time independent), see launch_trace_grid_r
  std::vector<GridRGroup> gr_groups;
  std::vector<int> gr_runs;
  DevBuf d_grg, d_grr;
  // GEMM form of the slice-pair tiles (every item with the same runs), see launch_trace_gemm
  std::vector<GridGTile> gg_tiles;
  std::vector<int> gg_runs;
  DevBuf d_ggt, d_ggr, d_ggp;  // d_ggp: partial tiles of the K split
  // Summed-product form of the crossed trace (see plan_xform): the product
  // node's table holds the X and W matrices, the trace items are sub-items of
  // the trace ids (several per trace id, one reduction each), xroot their
  // operand offsets.
  bool xform = false;
  std::vector<STerm> xterms;
  std::vector<SpUnit> xunits;
  DevBuf d_xterms, d_xunits;
  NodePlan xroot;
  // the table entries its items read: offset -> (label pairing: 1 swaps the
  // two random vectors of a label, 0 keeps it; label -> matrix index)
  std::map<long long, std::pair<int, std::map<std::pair<int, int>, int>>> xent;
  // symmetric-product plan of its traces (see plan_sym): units, stage and
  // tile tables, per trace id its partial sums (one per item)
  std::vector<SymUnit> sym_units;
  std::vector<int> sym_stages, sym_tiles;
  std::vector<GroupItem> sym_red;
  std::vector<int> sym_ridx;
  long long sym_npart = 0;
  double sym_bytes = 0;  // bytes per slice pair: staged matrices + partial sums
  DevBuf d_symu, d_syms, d_symt, d_symr, d_symri;
  long long npart() const { return std::max(ls_npart, sym_npart); }
};

struct CorrPlan {
  std::string name;
  // correlators in the device order (ordered by their traces, so that the
  // lanes of a resolve wave share traces): position -> [(family, trl)]
  std::vector<std::vector<std::pair<int, int>>> reqs;
  std::vector<int> orig;              // position -> index in the request list (and in datasets)
  std::vector<std::string> datasets;  // in request order
  DevBuf d_orig;
  std::vector<double> n_adm;  // admitted combinations per correlator
};

// Is the chunk a rectangular grid in the operands' time bases?  True when the
// slice pairs map one-to-one onto (distinct left base) x (distinct right base)
// with at most kGridT of each; fills the tile tables then.
bool grid_tile(std::vector<long long> const &Xt, std::vector<long long> const &Yt, GridTile &tile) {
  std::vector<long long> L, R;
  for (auto x : Xt)
    if (std::find(L.begin(), L.end(), x) == L.end()) {
      if (L.size() == size_t(kGridT)) return false;
      L.push_back(x);
    }
  for (auto y : Yt)
    if (std::find(R.begin(), R.end(), y) == R.end()) {
      if (R.size() == size_t(kGridT)) return false;
      R.push_back(y);
    }
  if (L.size() * R.size() != Xt.size()) return false;
  tile = GridTile{};
  for (auto &row : tile.sidx)
    for (auto &v : row) v = -1;
  for (size_t s = 0; s < Xt.size(); ++s) {
    int a = int(std::find(L.begin(), L.end(), Xt[s]) - L.begin());
    int b = int(std::find(R.begin(), R.end(), Yt[s]) - R.begin());
    if (tile.sidx[a][b] >= 0) return false;
    tile.sidx[a][b] = int(s);
  }
  tile.nL = int(L.size());
  tile.nR = int(R.size());
  for (size_t a = 0; a < L.size(); ++a) tile.Lt[a] = L[a];
  for (size_t b = 0; b < R.size(); ++b) tile.Rt[b] = R[b];
  return true;
}

// Orders a family's group-sum items by connected component of the (left
// entry, right entry) graph of its traces, then by left entry, so that items
// launched next to each other (and, through the XCD-aware block mapping of the
// trace kernels, on the same L2) read the same operand matrices.  Item
// destinations are stored in the items, so the order is free.
void order_items_by_component(Family &f, NodePlan const &root) {
  if (f.shape.nodes.back().kind != kTrace2 || f.gitems.empty()) return;
  std::map<std::pair<int, long long>, int> node_id;
  std::vector<int> parent;
  auto node = [&](int side, long long off) {
    auto it = node_id.emplace(std::make_pair(side, off), int(parent.size())).first;
    if (it->second == int(parent.size())) parent.push_back(it->second);
    return it->second;
  };
  std::function<int(int)> find = [&](int x) { return parent[x] == x ? x : parent[x] = find(parent[x]); };
  for (auto const &it : f.gitems) {
    int a = find(node(0, root.h_Xq.at(it.qn))), b = find(node(f.coincide ? 0 : 1, root.h_Yq.at(it.qn)));
    if (a != b) parent[std::max(a, b)] = std::min(a, b);
  }
  std::vector<std::pair<std::pair<int, long long>, int>> key(f.gitems.size());
  for (size_t k = 0; k < f.gitems.size(); ++k) {
    long long xq = root.h_Xq.at(f.gitems[k].qn);
    key[k] = {{find(node(0, xq)), xq}, int(k)};
  }
  std::stable_sort(key.begin(), key.end(), [](auto const &x, auto const &y) { return x.first < y.first; });
  std::vector<GroupItem> sorted(f.gitems.size());
  for (size_t k = 0; k < key.size(); ++k) sorted[k] = f.gitems[key[k].second];
  f.gitems = std::move(sorted);
}

// Left-stationary plan of a family's two-operand trace (time independent).
// Left combinations are cut into pieces of at most trace_lstat_chunk(D); a
// unit is (piece, left operator entry, up to kLsSlots of its items) and owns
// one partial sum per item.  Units are packed into blocks of up to
// kLsBlockChunks chunks (one LDS staging each); left entries run in order of
// the connected components of the (left entry, right entry) trace graph.
// Entry-major order (default): a block holds the pieces of one entry.
// Class-major order (SLAPH_LSTAT_CLASSES=1): the (left combination, right
// combination) pairs of all traces split into classes, the connected
// components of the pair graph (one node space when both operands read the
// same table; for a four-quark trace with four distinct random vectors a
// class is one unordered pair of end vectors), and all entries' work in one
// class runs before the next class.  Inside a chunk each slot's (i, j) pairs
// are sorted by j, consecutive pairs with the same j share one task, and the
// task list is cut into kLsWaves ranges of about equal length; each wave
// flushes its per-lane sum at the end of every run and of its range into its
// own slot copy.  ls_red lists each item's partial sums in piece order, so
// the order of every sum is fixed by the plan.
constexpr int kLsBlockChunks = 6;  // chunks (LDS stagings) per block

void plan_lstat(Family &f, NodePlan const &root, int D) {
  f.ls_blocks.clear();
  f.ls_chunks.clear();
  f.ls_tasks.clear();
  f.ls_slot_dst.clear();
  f.ls_left.clear();
  f.ls_red.clear();
  f.ls_ridx.clear();
  f.ls_units.clear();
  f.ls_npart = 0;
  // The queued form (default; SLAPH_LSTAT_BLOCKS=1 selects blocks of up to
  // kLsBlockChunks chunks in entry-major order) runs one chunk per unit in
  // component -> piece -> entry order.
  f.ls_queue = std::getenv("SLAPH_LSTAT_BLOCKS") == nullptr;
  const int CL = f.ls_queue ? trace_lstat_qchunk(D) : trace_lstat_chunk(D);
  const int PF = f.ls_queue ? trace_lstat_qring(D) : trace_lstat_ring(D);
  if (f.shape.nodes.back().kind != kTrace2 || CL <= 0 || f.gitems.empty()) return;
  std::function<int(std::vector<int> &, int)> find = [&](std::vector<int> &par, int x) {
    return par[x] == x ? x : par[x] = find(par, par[x]);
  };
  // (1) classes of combinations
  std::map<std::pair<int, int>, int> cnode_id;
  std::vector<int> cpar;
  auto cnode = [&](int side, int c) {
    auto it = cnode_id.emplace(std::make_pair(side, c), int(cpar.size())).first;
    if (it->second == int(cpar.size())) cpar.push_back(it->second);
    return it->second;
  };
  for (auto const &it : f.gitems)
    for (int c = 0; c < it.n; ++c) {
      int2 p = f.gpairs[it.list_off + c];
      int a = find(cpar, cnode(0, p.x)), b = find(cpar, cnode(f.coincide ? 0 : 1, p.y));
      if (a != b) cpar[std::max(a, b)] = std::min(a, b);
    }
  std::map<int, std::vector<int>> members;  // class root -> left combinations (ascending)
  // One class of all combinations (contiguous pieces, entry-major blocks)
  // unless SLAPH_LSTAT_CLASSES=1 selects the class-major plan: it reads about
  // half the bytes from HBM but runs shorter chunks, and is slower here.
  const bool noclass = std::getenv("SLAPH_LSTAT_CLASSES") == nullptr;
  for (auto const &kv : cnode_id)
    if (kv.first.first == 0) members[noclass ? 0 : find(cpar, kv.second)].push_back(kv.first.second);
  std::vector<std::vector<int>> pieces;  // at most CL left combinations of one class
  for (auto &kv : members) std::sort(kv.second.begin(), kv.second.end());
  {
    std::vector<std::vector<int>> cls;
    for (auto const &kv : members) cls.push_back(kv.second);
    std::sort(cls.begin(), cls.end());
    for (auto const &c : cls)
      for (size_t a = 0; a < c.size(); a += CL)
        pieces.emplace_back(c.begin() + a, c.begin() + std::min(c.size(), a + CL));
  }
  std::map<int, int> piece_of;
  for (size_t pc = 0; pc < pieces.size(); ++pc)
    for (int i : pieces[pc]) piece_of[i] = int(pc);
  // (2) left entries in trace-graph component order
  std::map<long long, std::vector<int>> by_left;
  for (size_t k = 0; k < f.gitems.size(); ++k) by_left[root.h_Xq.at(f.gitems[k].qn)].push_back(int(k));
  std::map<std::pair<int, long long>, int> node_id;
  std::vector<int> parent;
  auto node = [&](int side, long long off) {
    auto it = node_id.emplace(std::make_pair(side, off), int(parent.size())).first;
    if (it->second == int(parent.size())) parent.push_back(it->second);
    return it->second;
  };
  for (auto const &it : f.gitems) {
    int a = find(parent, node(0, root.h_Xq.at(it.qn))), b = find(parent, node(f.coincide ? 0 : 1, root.h_Yq.at(it.qn)));
    if (a != b) parent[std::max(a, b)] = std::min(a, b);
  }
  std::vector<std::pair<int, long long>> order;  // (component, left entry)
  for (auto const &kv : by_left) order.push_back({find(parent, node(0, kv.first)), kv.first});
  std::stable_sort(order.begin(), order.end(),
                   [](std::pair<int, long long> const &x, std::pair<int, long long> const &y) { return x.first < y.first; });
  // (3) blocks: per piece, consecutive left entries packed into one chunk
  std::vector<std::vector<int>> item_parts(f.gitems.size());
  struct CurChunk {
    std::vector<int> left;
    std::vector<LsTask> tasks;
  };
  struct Cur {
    std::vector<CurChunk> ch;
    int nslot = 0;
  } cur;
  auto flush = [&]() {
    if (cur.ch.empty()) {
      cur = Cur();
      return;
    }
    LsBlock b;
    b.chunk_off = int(f.ls_chunks.size());
    b.nchunk = int(cur.ch.size());
    b.slot_off = int(f.ls_slot_dst.size()) - cur.nslot;
    b.nslot = cur.nslot;
    // each chunk's tasks cut into kLsWaves ranges of about equal length; a
    // wave's range ends with a flush into the slot of its last task
    std::vector<std::array<int, kLsWaves + 1>> cut(cur.ch.size());
    for (size_t c = 0; c < cur.ch.size(); ++c) {
      auto &cc = cur.ch[c];
      const int nt = int(cc.tasks.size());
      for (int w = 0; w <= kLsWaves; ++w) cut[c][w] = int((long long)nt * w / kLsWaves);
      for (int w = 0; w < kLsWaves; ++w) {
        int last = cut[c][w + 1] - 1;
        if (last < cut[c][w]) continue;
        if (cc.tasks[last].slot < 0) {
          int k = last;
          while (cc.tasks[k].slot < 0) ++k;  // the run's own slot, found at its end
          cc.tasks[last].slot = cc.tasks[k].slot;
        }
      }
    }
    // tasks wave-major: wave w's ranges of all chunks are contiguous, each
    // padded to a multiple of the ring depth with null tasks
    std::vector<LsChunk> chs(cur.ch.size());
    for (size_t c = 0; c < cur.ch.size(); ++c) {
      chs[c].loff = int(f.ls_left.size());
      chs[c].n = int(cur.ch[c].left.size());
      f.ls_left.insert(f.ls_left.end(), cur.ch[c].left.begin(), cur.ch[c].left.end());
    }
    for (int w = 0; w < kLsWaves; ++w)
      for (size_t c = 0; c < cur.ch.size(); ++c) {
        auto const &cc = cur.ch[c];
        chs[c].wbeg[w] = int(f.ls_tasks.size());
        f.ls_tasks.insert(f.ls_tasks.end(), cc.tasks.begin() + cut[c][w], cc.tasks.begin() + cut[c][w + 1]);
        while ((f.ls_tasks.size() - chs[c].wbeg[w]) % PF != 0) {
          LsTask t = f.ls_tasks.back();
          t.ia = t.ib = -1;
          t.slot = -1;
          f.ls_tasks.push_back(t);
        }
        chs[c].wcnt[w] = int(f.ls_tasks.size()) - chs[c].wbeg[w];
      }
    f.ls_chunks.insert(f.ls_chunks.end(), chs.begin(), chs.end());
    f.ls_blocks.push_back(b);
    cur = Cur();
  };
  struct Unit {
    int pc;
    long long xq;
    size_t s0;
  };
  std::vector<Unit> units;
  const bool entry_major = noclass;
  if (f.ls_queue) {
    // per component: every piece of its entries before the next piece, so
    // that the units running together read one component-piece's matrices
    for (size_t g = 0; g < order.size();) {
      size_t h = g;
      while (h < order.size() && order[h].first == order[g].first) ++h;
      for (size_t pc = 0; pc < pieces.size(); ++pc)
        for (size_t k = g; k < h; ++k)
          for (size_t s0 = 0, nk = by_left[order[k].second].size(); s0 < nk; s0 += kLsSlots)
            units.push_back({int(pc), order[k].second, s0});
      g = h;
    }
  } else if (entry_major) {
    for (auto const &ord : order)
      for (size_t pc = 0; pc < pieces.size(); ++pc)
        for (size_t s0 = 0, nk = by_left[ord.second].size(); s0 < nk; s0 += kLsSlots)
          units.push_back({int(pc), ord.second, s0});
  } else {
    for (size_t pc = 0; pc < pieces.size(); ++pc)
      for (auto const &ord : order)
        for (size_t s0 = 0, nk = by_left[ord.second].size(); s0 < nk; s0 += kLsSlots)
          units.push_back({int(pc), ord.second, s0});
  }
  for (size_t u = 0; u < units.size(); ++u) {
    const int pc = units[u].pc;
    if (u > 0 && (f.ls_queue || (entry_major ? units[u - 1].xq != units[u].xq : units[u - 1].pc != pc))) flush();
    const size_t s0 = units[u].s0, nk = by_left[units[u].xq].size();
    const long long xq = units[u].xq;
    auto const &all = by_left[xq];
    const std::vector<int> items(all.begin() + s0, all.begin() + std::min(nk, s0 + kLsSlots));
    // this entry's pairs in the piece, per item
    std::vector<std::vector<std::pair<int, int>>> prs(items.size());  // (j, i)
    std::vector<int> used;
    int nslot = 0;
    for (size_t s = 0; s < items.size(); ++s) {
      GroupItem const &it = f.gitems[items[s]];
      for (int c = 0; c < it.n; ++c) {
        int2 p = f.gpairs[it.list_off + c];
        if (piece_of.at(p.x) != int(pc)) continue;
        prs[s].push_back({p.y, p.x});
        used.push_back(p.x);
      }
      if (!prs[s].empty()) ++nslot;
    }
    if (nslot == 0) continue;
    std::sort(used.begin(), used.end());
    used.erase(std::unique(used.begin(), used.end()), used.end());
    if (cur.nslot + nslot > kLsSlots) flush();
    if (cur.ch.empty() || int(cur.ch.back().left.size() + used.size()) > CL) {
      if (int(cur.ch.size()) == kLsBlockChunks) flush();
      cur.ch.emplace_back();
    }
    CurChunk &cc = cur.ch.back();
    std::map<int, int> pos;
    for (int i : used) {
      long long off = xq + i;
      if (off > INT32_MAX) throw std::runtime_error("trace plan: left operand index overflows");
      pos[i] = int(cc.left.size());
      cc.left.push_back(int(off));
    }
    for (size_t s = 0; s < items.size(); ++s) {
      auto &pr = prs[s];
      if (pr.empty()) continue;
      std::sort(pr.begin(), pr.end());
      const int slot = cur.nslot++;
      item_parts[items[s]].push_back(int(f.ls_npart));
      f.ls_slot_dst.push_back(f.ls_npart++);
      const long long yq = root.h_Yq.at(f.gitems[items[s]].qn);
      for (size_t a = 0; a < pr.size();) {
        LsTask t;
        long long yrel = yq + pr[a].first;
        if (yrel > INT32_MAX) throw std::runtime_error("trace plan: right operand index overflows");
        t.yrel = int(yrel);
        t.ia = short(pos.at(pr[a].second));
        t.ib = -1;
        if (a + 1 < pr.size() && pr[a + 1].first == pr[a].first) {
          t.ib = short(pos.at(pr[a + 1].second));
          a += 2;
        } else {
          a += 1;
        }
        t.slot = a == pr.size() ? slot : -1;
        cc.tasks.push_back(t);
      }
    }
  }
  flush();
  if (f.ls_queue) {
    // unit records; the task array ends with a null task, so that an empty
    // range's first entry is readable
    LsTask nt{0, -1, -1, -1};
    f.ls_tasks.push_back(nt);
    const int yr = kLsUnitHdr + kLsWaves * PF, ur = yr + CL;
    f.ls_units.assign(f.ls_blocks.size() * ur, 0);
    for (size_t b = 0; b < f.ls_blocks.size(); ++b) {
      LsBlock const &bl = f.ls_blocks[b];
      if (bl.nchunk != 1) throw std::logic_error("queued trace plan: a block must hold one chunk");
      LsChunk const &c = f.ls_chunks[bl.chunk_off];
      if (c.n < 1 || c.n > CL) throw std::logic_error("queued trace plan: chunk size out of range");
      int *r = &f.ls_units[b * ur];
      r[0] = c.n;
      r[1] = bl.slot_off;
      r[2] = bl.nslot;
      for (int w = 0; w < kLsWaves; ++w) {
        r[3 + w] = c.wbeg[w];
        r[3 + kLsWaves + w] = c.wcnt[w];
        for (int u = 0; u < PF && u < c.wcnt[w]; ++u) r[kLsUnitHdr + w * PF + u] = f.ls_tasks[c.wbeg[w] + u].yrel;
      }
      for (int m = 0; m < CL; ++m) r[yr + m] = f.ls_left[c.loff + std::min(m, c.n - 1)];
      for (int k = 0; k < bl.nslot; ++k)
        if (f.ls_slot_dst[bl.slot_off + k] != bl.slot_off + k)
          throw std::logic_error("queued trace plan: partial sums must follow slot order");
    }
  }
  // (4) per item: its partial sums, in piece order (in the summed-product
  // form the items of one trace id share one reduction, in item order)
  if (f.xform) {
    std::map<long long, int> red_of;
    std::vector<std::vector<int>> parts;
    for (size_t k = 0; k < f.gitems.size(); ++k) {
      auto it = red_of.emplace(f.gitems[k].dst, int(parts.size())).first;
      if (it->second == int(parts.size())) parts.emplace_back();
      parts[it->second].insert(parts[it->second].end(), item_parts[k].begin(), item_parts[k].end());
    }
    for (auto const &kv : red_of) {
      auto const &pl = parts[kv.second];
      f.ls_red.push_back(GroupItem{0, int(f.ls_ridx.size()), int(pl.size()), kv.first});
      f.ls_ridx.insert(f.ls_ridx.end(), pl.begin(), pl.end());
    }
  } else {
    for (size_t k = 0; k < f.gitems.size(); ++k) {
      GroupItem r{0, int(f.ls_ridx.size()), int(item_parts[k].size()), f.gitems[k].dst};
      f.ls_ridx.insert(f.ls_ridx.end(), item_parts[k].begin(), item_parts[k].end());
      f.ls_red.push_back(r);
    }
  }
  if (f.ls_npart > INT32_MAX) throw std::runtime_error("trace plan: too many partial sums");
  {
    std::set<long long> lset, rset;
    for (auto const &it : f.gitems)
      for (int c = 0; c < it.n; ++c) {
        int2 p = f.gpairs[it.list_off + c];
        lset.insert(root.h_Xq.at(it.qn) + p.x);
        rset.insert(root.h_Yq.at(it.qn) + p.y);
      }
    size_t nmat = lset.size() + rset.size();
    if (f.coincide) {  // both operands read one table at one time base
      lset.insert(rset.begin(), rset.end());
      nmat = lset.size();
    }
    f.ls_bytes = double(sizeof(double2)) * (double(nmat) * D * D + double(f.ls_npart));
  }
  if (std::getenv("SLAPH_PROGRESS")) {
    size_t null_tasks = 0, paired = 0;
    for (auto const &t : f.ls_tasks) {
      null_tasks += t.ia < 0;
      paired += t.ib >= 0;
    }
    std::fprintf(stderr, "left-stationary trace plan (D = %d): %zu blocks, %zu chunks, %zu tasks (%zu null, %zu with two left matrices), %zu staged left matrices, %lld partial sums\n",
                 D, f.ls_blocks.size(), f.ls_chunks.size(), f.ls_tasks.size(), null_tasks, paired, f.ls_left.size(),
                 (long long)f.ls_npart);
  }
  upload(f.d_lsb, f.ls_blocks);
  upload(f.d_lsc, f.ls_chunks);
  upload(f.d_lst, f.ls_tasks);
  upload(f.d_lsd, f.ls_slot_dst);
  upload(f.d_lsl, f.ls_left);
  upload(f.d_lsr, f.ls_red);
  upload(f.d_lsri, f.ls_ridx);
  if (f.ls_queue) {
    upload(f.d_lsu, f.ls_units);
    f.d_lsq.ensure(8 * sizeof(int));
  }
}

// Right-stationary plan of a family's two-operand trace for the slice-pair
// tile kernel.  Items are grouped per connected component of the (left
// entry, right entry) trace graph: the component's left entries are cut into
// blocks of kGridIPW, and for each block and each right entry one group holds
// the items of that right entry with a left entry in the block (more than
// kGridIPW such items spill into further groups).  Groups run block by block,
// so the groups in flight on one XCD read the same few left entries.  An
// item's runs are its pairs with one left combination i (one or two right
// combinations, consecutive in the item's list, as the item-wise kernel takes
// them); a group's run records are keyed by the right combination(s), and each
// item fills its slot of the first record with its key that it has not used.
void plan_gridr(Family &f, NodePlan const &root) {
  f.gr_groups.clear();
  f.gr_runs.clear();
  if (f.shape.nodes.back().kind != kTrace2 || f.gitems.empty() || std::getenv("SLAPH_GRID_ITEMS")) return;
  std::function<int(std::vector<int> &, int)> find = [&](std::vector<int> &par, int x) {
    return par[x] == x ? x : par[x] = find(par, par[x]);
  };
  std::map<std::pair<int, long long>, int> node_id;
  std::vector<int> parent;
  auto node = [&](int side, long long off) {
    auto it = node_id.emplace(std::make_pair(side, off), int(parent.size())).first;
    if (it->second == int(parent.size())) parent.push_back(it->second);
    return it->second;
  };
  for (auto const &it : f.gitems) {
    int a = find(parent, node(0, root.h_Xq.at(it.qn))), b = find(parent, node(f.coincide ? 0 : 1, root.h_Yq.at(it.qn)));
    if (a != b) parent[std::max(a, b)] = std::min(a, b);
  }
  // components in order of first appearance; per component its left entries
  // (in order of first appearance) and its items
  std::vector<int> comp_order;
  std::map<int, std::vector<long long>> comp_left;
  std::map<int, std::vector<int>> comp_items;
  for (size_t k = 0; k < f.gitems.size(); ++k) {
    const long long xq = root.h_Xq.at(f.gitems[k].qn);
    const int c = find(parent, node(0, xq));
    if (comp_items.find(c) == comp_items.end()) comp_order.push_back(c);
    comp_items[c].push_back(int(k));
    auto &cl = comp_left[c];
    if (std::find(cl.begin(), cl.end(), xq) == cl.end()) cl.push_back(xq);
  }
  auto add_group = [&](long long yq, std::vector<int> const &its) {
    GridRGroup g{};
    g.yq = yq;
    g.nitem = int(its.size());
    g.run_off = int(f.gr_runs.size() / kGridRun);
    std::map<std::pair<int, int>, std::vector<int>> recs;  // key -> record indices (relative)
    int nrec = 0;
    for (int w = 0; w < g.nitem; ++w) {
      GroupItem const &it = f.gitems[its[w]];
      g.dst[w] = it.dst;
      const long long xq = root.h_Xq.at(it.qn);
      int2 const *p = &f.gpairs[it.list_off];
      for (int c = 0; c < it.n;) {
        const bool two = c + 1 < it.n && p[c + 1].x == p[c].x;
        const int j1 = p[c].y, j2 = two ? p[c + 1].y : -1;
        const long long xo = xq + p[c].x;
        if (xo > INT32_MAX) throw std::runtime_error("grid plan: left operand index overflows");
        auto &lst = recs[{j1, j2}];
        int rec = -1;
        for (int r : lst)
          if (f.gr_runs[(size_t)(g.run_off + r) * kGridRun + 2 + w] < 0) {
            rec = r;
            break;
          }
        if (rec < 0) {
          rec = nrec++;
          lst.push_back(rec);
          f.gr_runs.insert(f.gr_runs.end(), kGridRun, -1);
          f.gr_runs[(size_t)(g.run_off + rec) * kGridRun] = j1;
          f.gr_runs[(size_t)(g.run_off + rec) * kGridRun + 1] = j2;
        }
        f.gr_runs[(size_t)(g.run_off + rec) * kGridRun + 2 + w] = int(xo);
        c += two ? 2 : 1;
      }
    }
    g.nrun = nrec;
    f.gr_groups.push_back(g);
  };
  for (int c : comp_order) {
    auto const &cl = comp_left[c];
    for (size_t b0 = 0; b0 < cl.size(); b0 += kGridIPW) {
      std::set<long long> block(cl.begin() + b0, cl.begin() + std::min(cl.size(), b0 + kGridIPW));
      std::map<long long, std::vector<int>> by_right;  // right entry -> items with a left entry in the block
      std::vector<long long> right_order;
      for (int k : comp_items[c]) {
        GroupItem const &it = f.gitems[k];
        if (!block.count(root.h_Xq.at(it.qn))) continue;
        const long long yq = root.h_Yq.at(it.qn);
        if (by_right.find(yq) == by_right.end()) right_order.push_back(yq);
        by_right[yq].push_back(k);
      }
      for (long long yq : right_order) {
        auto const &its = by_right[yq];
        for (size_t a = 0; a < its.size(); a += kGridIPW)
          add_group(yq, std::vector<int>(its.begin() + a, its.begin() + std::min(its.size(), a + kGridIPW)));
      }
    }
  }
  if (f.gr_runs.size() / kGridRun > INT32_MAX) throw std::runtime_error("grid plan: too many runs");
  upload(f.d_grg, f.gr_groups);
  upload(f.d_grr, f.gr_runs);
}

// GEMM plan of a family's two-operand trace for the slice-pair tile kernel,
// when every item has the same runs (an item's pairs with one left
// combination i, one or two right combinations, consecutive in its list) and
// no two items share (left entry, right entry): per connected component of
// the trace graph, its left and right entries in blocks of kGemmT, one tile
// per pair of blocks that holds an item.  Otherwise the plan stays empty and
// the right-stationary form runs.
void plan_gridg(Family &f, NodePlan const &root) {
  f.gg_tiles.clear();
  f.gg_runs.clear();
  if (f.shape.nodes.back().kind != kTrace2 || f.gitems.empty() || std::getenv("SLAPH_GRID_ITEMS") ||
      std::getenv("SLAPH_GRID_GROUPS"))
    return;
  auto runs_of = [&](GroupItem const &it) {
    std::vector<int> r;
    int2 const *p = &f.gpairs[it.list_off];
    for (int c = 0; c < it.n;) {
      const bool two = c + 1 < it.n && p[c + 1].x == p[c].x;
      r.push_back(p[c].x);
      r.push_back(p[c].y);
      r.push_back(two ? p[c + 1].y : -1);
      c += two ? 2 : 1;
    }
    return r;
  };
  const std::vector<int> runs = runs_of(f.gitems[0]);
  std::map<std::pair<long long, long long>, long long> item_of;  // (xq, yq) -> dst
  for (auto const &it : f.gitems) {
    if (runs_of(it) != runs) return;
    auto key = std::make_pair(root.h_Xq.at(it.qn), root.h_Yq.at(it.qn));
    if (!item_of.emplace(key, it.dst).second) return;
  }
  std::function<int(std::vector<int> &, int)> find = [&](std::vector<int> &par, int x) {
    return par[x] == x ? x : par[x] = find(par, par[x]);
  };
  std::map<std::pair<int, long long>, int> node_id;
  std::vector<int> parent;
  auto node = [&](int side, long long off) {
    auto it = node_id.emplace(std::make_pair(side, off), int(parent.size())).first;
    if (it->second == int(parent.size())) parent.push_back(it->second);
    return it->second;
  };
  for (auto const &it : f.gitems) {
    int a = find(parent, node(0, root.h_Xq.at(it.qn))), b = find(parent, node(f.coincide ? 0 : 1, root.h_Yq.at(it.qn)));
    if (a != b) parent[std::max(a, b)] = std::min(a, b);
  }
  std::vector<int> comp_order;
  std::map<int, std::vector<long long>> comp_left, comp_right;
  for (auto const &it : f.gitems) {
    const long long xq = root.h_Xq.at(it.qn), yq = root.h_Yq.at(it.qn);
    const int c = find(parent, node(0, xq));
    if (comp_left.find(c) == comp_left.end()) comp_order.push_back(c);
    auto &cl = comp_left[c];
    if (std::find(cl.begin(), cl.end(), xq) == cl.end()) cl.push_back(xq);
    auto &cr = comp_right[c];
    if (std::find(cr.begin(), cr.end(), yq) == cr.end()) cr.push_back(yq);
  }
  for (int c : comp_order) {
    auto const &cl = comp_left[c], &cr = comp_right[c];
    for (size_t e0 = 0; e0 < cl.size(); e0 += kGemmT)
      for (size_t f0 = 0; f0 < cr.size(); f0 += kGemmT) {
        GridGTile t{};
        t.nE = int(std::min<size_t>(kGemmT, cl.size() - e0));
        t.nF = int(std::min<size_t>(kGemmT, cr.size() - f0));
        bool any = false;
        for (int r = 0; r < kGemmT; ++r) {
          t.xq[r] = cl[e0 + std::min(r, t.nE - 1)];
          t.yq[r] = cr[f0 + std::min(r, t.nF - 1)];
        }
        for (int r = 0; r < kGemmT; ++r)
          for (int q = 0; q < kGemmT; ++q) {
            t.dst[r][q] = -1;
            if (r < t.nE && q < t.nF) {
              auto it = item_of.find({t.xq[r], t.yq[q]});
              if (it != item_of.end()) {
                t.dst[r][q] = it->second;
                any = true;
              }
            }
          }
        if (any) f.gg_tiles.push_back(t);
      }
  }
  f.gg_runs = runs;
  upload(f.d_ggt, f.gg_tiles);
  upload(f.d_ggr, f.gg_runs);
}

// Summed-product form of the crossed trace trQ2Q0Q2Q0 (DilutedTraceFactory.cpp:
// 161-190) when its two products read the same table.  A trace id
// (a, b, c, d) sums tr(Q0_a(r1 r2) Q2_b(r2 r3) Q0_c(r3 r4) Q2_d(r4 r1)) over
// the random-vector chains that operator* and factor_to_trace admit
// (DilutedFactor.cpp:6-51, DilutedTrace.cpp:84-138): factors exist for their
// (first, second) vector pairs, r1 != r3 (both products) and r2 != r4 (the
// used masks).  Summing r2 and r4 inside the products and removing r2 = r4 by
// inclusion and exclusion,
//   T = sum_{r1 != r3} tr(X_ab(r1 r3) X_cd(r3 r1))
//       - sum_{r2} tr(Y_bc(r2) Y_da(r2)) + sum_{r2, r1} tr(V_bc(r2 r1) V_da(r2 r1))
// with X_ab(r1 r3) = sum_{r2} Q0_a(r1 r2) Q2_b(r2 r3) over r1 != r3,
// V_bc(r2 r1) = Q2_b(r2 r1) Q0_c(r1 r2) and Y_bc(r2) = sum_{r1} V_bc(r2 r1):
// the second sum is every r2 = r4 chain (r1 = r3 included), the third puts
// back those with r1 = r3.  The product node's table holds, per time key, the X
// matrices of every product entry (a, b) and, per crossed pair (b, c) the
// trace ids read, the V matrices of each r2 followed by i Y(r2), both
// transposed (tr(A^T B^T) = tr(A B); the transpose puts Q0 on the left of
// every product, see launch_sprod): i Y is stored on both sides, tr(iY iY) =
// -tr(Y Y), so that every term is a plain trace of two table matrices.  The
// X matrices are sums over r2, so the table has no per-r2 products to
// write.  Each trace id becomes two items of the left-stationary
// plan (X pairs, W pairs) that reduce into its one group sum.  The trace
// reads 20 + 25 matrices per trace id instead of 60 (five random vectors), the
// table 45 per entry instead of 60.  Returns false, with the family unchanged,
// when a Q2 entry has more random-vector pairs than the product kernel stages
// (kSpStage); the family then keeps one product per combination.
bool plan_xform(Family &f, std::vector<FactorIndex> const &q0l, std::vector<FactorIndex> const &q2l,
                std::vector<long long> const &off0, std::vector<long long> const &off2) {
  NodePlan &p = f.nodes[2];
  NodePlan &root = f.nodes.back();
  for (auto const &key : p.qn_key)
    if (q2l.at(key.at(1)).rnd.size() > size_t(kSpStage)) return false;
  using Lab = std::pair<int, int>;
  f.xterms.clear();
  f.xunits.clear();
  f.xent.clear();
  long long tot = 0;
  long long maxq0 = 0, maxq2 = 0;
  // Dirac block pattern of a Q0 entry: its gamma matrix's row of each column
  auto perm_of = [&](int e) {
    auto const &g = gamma_table().at(q0l.at(e).gamma.at(0));
    int m = 0, seen = 0;
    for (int d = 0; d < 4; ++d) {
      m |= (g.row[d] & 3) << (2 * d);
      seen |= 1 << (g.row[d] & 3);
    }
    if (seen != 15) throw std::logic_error("summed-product form: gamma rows are not a permutation");
    return m;
  };
  // Runs of terms per Q2 operator entry (a unit of launch_sprod): kind 0 sums
  // Q0[xo + x] Q2_b[y] into output o; kind 1 stores (Q2_b[x] Q0[xo + y])^T at
  // o, o + 1, .. and i times their sum after them.
  struct Run {
    int kind, perm;
    long long xo, o;
    std::vector<int2> t;
  };
  std::map<int, std::vector<Run>> runs;  // Q2 entry -> its runs
  auto add_job = [&](long long xo, int b, long long o, std::vector<int2> const &t, int kind, int perm) {
    runs[b].push_back({kind, perm, xo, o, t});
  };
  // X part: one summed output per (r1, r3) of each product entry
  std::vector<long long> xoff(p.qn_key.size());
  std::vector<std::map<Lab, int>> xlab(p.qn_key.size());
  for (size_t q = 0; q < p.qn_key.size(); ++q) {
    const int a = p.qn_key[q].at(0), b = p.qn_key[q].at(1);
    auto const &ra = q0l.at(a).rnd, &rb = q2l.at(b).rnd;
    std::vector<Lab> order;
    std::map<Lab, std::vector<int2>> terms;
    for (size_t i = 0; i < ra.size(); ++i)
      for (size_t j = 0; j < rb.size(); ++j)
        if (ra[i].second == rb[j].first && ra[i].first != rb[j].second) {
          const Lab l{ra[i].first, rb[j].second};
          auto &t = terms[l];
          if (t.empty()) order.push_back(l);
          t.push_back(make_int2(int(i), int(j)));
        }
    xoff[q] = tot;
    for (size_t k = 0; k < order.size(); ++k) {
      add_job(off0.at(a), b, tot + (long long)k, terms[order[k]], 0, perm_of(a));
      xlab[q][order[k]] = int(k);
    }
    if (!order.empty()) f.xent[tot] = {1, xlab[q]};
    tot += (long long)order.size();
  }
  // W part, per crossed pair (b, c): labels (r2, r1) for V, (r2, -1) for i Y
  std::map<Lab, int> wid;
  std::vector<long long> woff;
  std::vector<std::map<Lab, int>> wlab;
  auto wentry = [&](int b, int c) -> int {
    auto it = wid.find({b, c});
    if (it != wid.end()) return it->second;
    auto const &rb = q2l.at(b).rnd, &rc = q0l.at(c).rnd;
    std::vector<int> r2s;
    std::map<int, std::vector<std::pair<int, int2>>> byr2;  // r2 -> (r1, (Q2 pair, Q0 pair))
    for (size_t i = 0; i < rb.size(); ++i)
      for (size_t j = 0; j < rc.size(); ++j)
        if (rb[i].second == rc[j].first && rb[i].first == rc[j].second) {
          auto &v = byr2[rb[i].first];
          if (v.empty()) r2s.push_back(rb[i].first);
          v.push_back({rb[i].second, make_int2(int(i), int(j))});
        }
    std::map<Lab, int> lab;
    int k = 0;
    for (int r2 : r2s) {
      std::vector<int2> t;
      const long long o = tot + k;
      for (auto const &e : byr2[r2]) {
        lab[{r2, e.first}] = k++;
        t.push_back(e.second);
      }
      lab[{r2, -1}] = k++;
      add_job(off0.at(c), b, o, t, 1, perm_of(c));
    }
    woff.push_back(tot);
    wlab.push_back(lab);
    if (k > 0) f.xent[tot] = {0, lab};
    tot += k;
    return wid[{b, c}] = int(woff.size()) - 1;
  };
  // items: per trace id (mirrors reuse their partner's sum) an X item and a W item
  std::vector<GroupItem> items;
  std::vector<int2> pairs;
  f.xroot = NodePlan();
  auto item = [&](long long lo, long long ro, std::map<Lab, int> const &L, std::map<Lab, int> const &R, bool swap,
                  long long dst) {
    const int off = int(pairs.size());
    for (auto const &kv : L) {
      auto it = R.find(swap ? Lab{kv.first.second, kv.first.first} : kv.first);
      if (it != R.end()) pairs.push_back(make_int2(kv.second, it->second));
    }
    const int n = int(pairs.size()) - off;
    if (n == 0) return;
    items.push_back({int(f.xroot.h_Xq.size()), off, n, dst});
    f.xroot.h_Xq.push_back(lo);
    f.xroot.h_Yq.push_back(ro);
  };
  for (size_t trl = 0; trl < f.tr_ids.size(); ++trl) {
    if (f.mirror_of[trl] >= 0) continue;
    if (f.ngroup[trl] != 1) throw std::logic_error("summed-product form: a trace with several mask groups");
    const int q = root.trl_qn[trl], ql = root.qn_left[q], qr = root.qn_right[q];
    const long long dst = f.goff[trl];
    item(xoff[ql], xoff[qr], xlab[ql], xlab[qr], true, dst);
    const int a = p.qn_key[ql].at(0), b = p.qn_key[ql].at(1), c = p.qn_key[qr].at(0), d = p.qn_key[qr].at(1);
    const int wl = wentry(b, c), wr = wentry(d, a);
    item(woff[wl], woff[wr], wlab[wl], wlab[wr], false, dst);
  }
  // units: per Q2 entry its runs, dealt to the kSpWaves waves in about equal term counts
  for (auto &kv : runs) {
    const int b = kv.first;
    SpUnit un;
    un.s0 = int(off2.at(b));
    un.np = int(q2l.at(b).rnd.size());
    if (un.np > kSpStage) throw std::logic_error("summed-product form: Q2 entry with too many random-vector pairs");
    maxq2 = std::max(maxq2, off2.at(b) + un.np - 1);
    long long nterm = 0;
    for (auto const &rn : kv.second) nterm += (long long)rn.t.size();
    long long done = 0;
    int w = 0;
    un.tb[0] = int(f.xterms.size());
    for (auto const &rn : kv.second) {
      while (w < kSpWaves - 1 && done * kSpWaves >= nterm * (w + 1)) un.tb[++w] = int(f.xterms.size());
      for (size_t k = 0; k < rn.t.size(); ++k) {
        const long long q0i = rn.xo + (rn.kind ? rn.t[k].y : rn.t[k].x);
        const int q2i = rn.kind ? rn.t[k].x : rn.t[k].y;
        const long long oo = rn.kind ? rn.o + (long long)k : rn.o;
        if (q0i > INT32_MAX || oo + 1 > INT32_MAX) throw std::runtime_error("summed-product plan: index overflows");
        if (q2i < 0 || q2i >= un.np) throw std::logic_error("summed-product plan: Q2 pair out of range");
        maxq0 = std::max(maxq0, q0i);
        STerm st;
        st.l = int(q0i);
        st.r = q2i;
        st.o = int(oo);
        st.flags = (k + 1 == rn.t.size() ? 1 : 0) | (rn.kind ? 2 : 0) | (rn.perm << 8);
        f.xterms.push_back(st);
      }
      done += (long long)rn.t.size();
    }
    while (w < kSpWaves) un.tb[++w] = int(f.xterms.size());
    f.xunits.push_back(un);
  }
  if (tot == 0 || items.empty() || f.xterms.empty()) throw std::logic_error("summed-product form: empty plan");
  f.gitems = std::move(items);
  f.gpairs = std::move(pairs);
  // the product node's table is now the X and W matrices of a time key
  p.stride = tot;
  p.cmaxO = tot - 1;
  p.cmaxX = maxq0;
  p.cmaxY = maxq2;
  root.cmaxX = root.cmaxY = tot - 1;
  upload(f.d_xterms, f.xterms);
  upload(f.d_xunits, f.xunits);
  return true;
}

// Symmetric-product plan of the summed-product traces (launch_trace_sym).
// The family's items form two trace graphs, one per label pairing (entries:
// table offsets, edges: items): the X items pair label (r1, r3) of the left
// entry with (r3, r1) of the right one, the W items equal labels, so that
// within a connected component every item value is an element of the
// symmetric matrix T[e1][e2] = sum_l tr(M_e1(l) M_e2(pi l)).  A component of
// at most `maxe` entries (kSymEntries, or SLAPH_SYM_ENTRIES for tests) is one
// piece, its tiles on and above the diagonal; a larger one is cut into chunks
// of `chunk` entries, one piece per chunk pair (a pair of distinct chunks:
// rows from the first, columns from the second).  Pieces are packed into
// units of at most kSymEntries positions, each piece (and its column part) at
// a multiple of four.  A unit's stages cover every label of its entries whose
// partner label occurs too: four cross stages (block (I, J) of l with block
// (J, I) of pi l) per label pair, one cross stage (blocks (0, 1), (1, 0)) and
// one self stage ((0, 0), (1, 1)) per label with pi l = l.  Each item has a
// partial sum of its own (its index in gitems); the reduction adds the items
// of a trace id in item order.  Returns false, leaving no plan, when an
// index does not fit the kernel's tables.
bool plan_sym(Family &f) {
  auto reset = [&]() {
    f.sym_units.clear();
    f.sym_stages.clear();
    f.sym_tiles.clear();
    f.sym_red.clear();
    f.sym_ridx.clear();
    f.sym_npart = 0;
    f.sym_bytes = 0;
  };
  reset();
  using Lab = std::pair<int, int>;
  int maxe = kSymEntries;
  if (const char *e = std::getenv("SLAPH_SYM_ENTRIES")) maxe = std::max(4, std::min(kSymEntries, std::atoi(e)));
  const int chunk = std::max(4, std::min(16, (maxe / 2) / 4 * 4));
  auto ceil4 = [](size_t n) { return int((n + 3) / 4 * 4); };
  // (1) components of both graphs (entry offsets are distinct across them)
  std::map<long long, long long> par;
  std::function<long long(long long)> find = [&](long long x) -> long long {
    auto it = par.emplace(x, x).first;
    if (it->second == x) return x;
    return it->second = find(it->second);
  };
  std::vector<std::array<long long, 2>> ends(f.gitems.size());
  for (size_t k = 0; k < f.gitems.size(); ++k) {
    const long long l = f.xroot.h_Xq.at(f.gitems[k].qn), r = f.xroot.h_Yq.at(f.gitems[k].qn);
    auto il = f.xent.find(l), ir = f.xent.find(r);
    if (il == f.xent.end() || ir == f.xent.end() || il->second.first != ir->second.first) return false;
    ends[k] = {l, r};
    const long long a = find(l), b = find(r);
    if (a != b) par[std::max(a, b)] = std::min(a, b);
  }
  std::map<long long, std::set<long long>> comp;
  std::map<long long, std::vector<int>> comp_items;
  for (size_t k = 0; k < f.gitems.size(); ++k) {
    const long long root = find(ends[k][0]);
    comp[root].insert(ends[k][0]);
    comp[root].insert(ends[k][1]);
    comp_items[root].push_back(int(k));
  }
  // (2) pieces
  struct Piece {
    std::vector<long long> rows, cols;  // cols empty: the symmetric piece of rows
    std::vector<int> items;
    int width() const { return int((rows.size() + 3) / 4 * 4 + (cols.size() + 3) / 4 * 4); }
  };
  std::vector<Piece> pieces[2];
  for (auto const &kv : comp) {
    const std::vector<long long> ent(kv.second.begin(), kv.second.end());
    const int kind = f.xent.at(ent[0]).first;
    auto const &items = comp_items[kv.first];
    if (int(ent.size()) <= maxe) {
      pieces[kind].push_back({ent, {}, items});
      continue;
    }
    std::map<long long, int> chunk_of;
    for (size_t i = 0; i < ent.size(); ++i) chunk_of[ent[i]] = int(i) / chunk;
    auto members = [&](int c) {
      return std::vector<long long>(ent.begin() + size_t(c) * chunk,
                                    ent.begin() + std::min(ent.size(), size_t(c + 1) * chunk));
    };
    std::map<std::pair<int, int>, Piece> pc;
    for (int k : items) {
      int a = chunk_of.at(ends[k][0]), b = chunk_of.at(ends[k][1]);
      if (a > b) std::swap(a, b);
      Piece &P = pc[{a, b}];
      if (P.rows.empty()) {
        P.rows = members(a);
        if (a != b) P.cols = members(b);
      }
      P.items.push_back(k);
    }
    for (auto &x : pc) pieces[kind].push_back(std::move(x.second));
  }
  long long nblocks = 0;
  for (int kind = 0; kind < 2; ++kind) {
    auto &pcs = pieces[kind];
    for (auto const &P : pcs)
      if (P.width() > kSymEntries) throw std::logic_error("symmetric-product plan: piece wider than a unit");
    std::stable_sort(pcs.begin(), pcs.end(), [](Piece const &x, Piece const &y) { return x.width() > y.width(); });
    // (3) first fit into units
    std::vector<std::vector<int>> bins;
    std::vector<int> used;
    for (size_t i = 0; i < pcs.size(); ++i) {
      size_t b = 0;
      while (b < bins.size() && used[b] + pcs[i].width() > kSymEntries) ++b;
      if (b == bins.size()) {
        bins.emplace_back();
        used.push_back(0);
      }
      bins[b].push_back(int(i));
      used[b] += pcs[i].width();
    }
    // (4) per unit: positions, stages, tiles
    const bool swap = kind == 1;
    for (size_t b = 0; b < bins.size(); ++b) {
      const int ns = used[b];
      std::vector<long long> at(ns, -1);  // entry at each position
      std::set<Lab> labels;
      std::map<std::pair<int, int>, std::array<int, 16>> tiles;
      int base = 0;
      for (int pi : bins[b]) {
        Piece const &P = pcs[size_t(pi)];
        const int cbase = base + ceil4(P.rows.size());
        std::map<long long, int> rpos, cpos;
        for (size_t i = 0; i < P.rows.size(); ++i) at[size_t(base) + i] = P.rows[i], rpos[P.rows[i]] = base + int(i);
        for (size_t i = 0; i < P.cols.size(); ++i) at[size_t(cbase) + i] = P.cols[i], cpos[P.cols[i]] = cbase + int(i);
        for (long long e : P.rows)
          for (auto const &lk : f.xent.at(e).second) labels.insert(lk.first);
        for (long long e : P.cols)
          for (auto const &lk : f.xent.at(e).second) labels.insert(lk.first);
        for (int k : P.items) {
          int pl, pr;
          if (P.cols.empty()) {
            pl = rpos.at(ends[k][0]);
            pr = rpos.at(ends[k][1]);
            if (pl / 4 > pr / 4) std::swap(pl, pr);
          } else if (rpos.count(ends[k][0])) {
            pl = rpos.at(ends[k][0]);
            pr = cpos.at(ends[k][1]);
          } else {
            pl = rpos.at(ends[k][1]);
            pr = cpos.at(ends[k][0]);
          }
          auto it = tiles.find({pl / 4, pr / 4});
          if (it == tiles.end()) {
            it = tiles.emplace(std::make_pair(pl / 4, pr / 4), std::array<int, 16>{}).first;
            it->second.fill(-1);
          }
          int idx = 4 * (pl & 3) + (pr & 3);
          if (it->second[idx] >= 0 && pl / 4 == pr / 4) idx = 4 * (pr & 3) + (pl & 3);
          if (it->second[idx] >= 0) return reset(), false;  // two items on one element of T
          it->second[idx] = k;
        }
        base = cbase + ceil4(P.cols.size());
      }
      SymUnit un{};
      if (f.sym_stages.size() > size_t(INT32_MAX) / 2) return reset(), false;
      un.st_off = int(f.sym_stages.size());
      un.ns = ns;
      auto emit = [&](int flags, Lab l0, int I0, int J0, Lab l1, int I1, int J1) -> bool {
        f.sym_stages.push_back(flags);
        for (int p = 0; p < ns; ++p)
          for (int slot = 0; slot < 2; ++slot) {
            int off = -1;
            if (at[size_t(p)] >= 0) {
              auto const &lab = f.xent.at(at[size_t(p)]).second;
              auto it = lab.find(slot ? l1 : l0);
              if (it != lab.end()) {
                const long long o = (at[size_t(p)] + it->second) * 256 + 128 * (slot ? I1 : I0) + 8 * (slot ? J1 : J0);
                if (o > INT32_MAX - 256) return false;
                off = int(o);
                ++nblocks;
              }
            }
            f.sym_stages.push_back(off);
          }
        ++un.nst;
        return true;
      };
      for (Lab const &l : labels) {
        const Lab pl = swap ? Lab{l.second, l.first} : l;
        bool ok = true;
        if (pl == l) {
          ok = emit(0, l, 0, 1, l, 1, 0) && emit(1, l, 0, 0, l, 1, 1);
        } else if (l < pl && labels.count(pl)) {
          for (int I = 0; I < 2 && ok; ++I)
            for (int J = 0; J < 2 && ok; ++J) ok = emit(0, l, I, J, pl, J, I);
        }
        if (!ok) return reset(), false;
      }
      if (un.nst % 2 && !emit(0, Lab{-2, -2}, 0, 0, Lab{-2, -2}, 0, 0)) return reset(), false;  // a zero stage
      // tiles in (row, column) order, cut into kSymWaves contiguous ranges
      const int n = int(tiles.size());
      if (n > kSymWaves * kSymWaveTiles) return reset(), false;
      un.tile_off = int(f.sym_tiles.size() / 17);
      for (auto const &t : tiles) {
        f.sym_tiles.push_back(t.first.first | (t.first.second << 8));
        f.sym_tiles.insert(f.sym_tiles.end(), t.second.begin(), t.second.end());
      }
      for (int w = 0; w <= kSymWaves; ++w) un.wt[w] = int((long long)n * w / kSymWaves);
      if (un.nst > 0 && n > 0) f.sym_units.push_back(un);
    }
  }
  // larger units first (the grid runs unit-major)
  std::stable_sort(f.sym_units.begin(), f.sym_units.end(), [](SymUnit const &x, SymUnit const &y) {
    return (long long)x.nst * x.ns > (long long)y.nst * y.ns;
  });
  std::map<long long, std::vector<int>> by_dst;
  for (size_t k = 0; k < f.gitems.size(); ++k) by_dst[f.gitems[k].dst].push_back(int(k));
  for (auto const &kv : by_dst) {
    f.sym_red.push_back(GroupItem{0, int(f.sym_ridx.size()), int(kv.second.size()), kv.first});
    f.sym_ridx.insert(f.sym_ridx.end(), kv.second.begin(), kv.second.end());
  }
  f.sym_npart = (long long)f.gitems.size();
  f.sym_bytes = double(nblocks) * 64 * sizeof(double2) + double(f.sym_npart) * sizeof(double2);
  if (std::getenv("SLAPH_PROGRESS"))
    std::fprintf(stderr, "symmetric-product trace plan: %zu units, %zu stage entries, %zu tiles, %lld staged blocks\n",
                 f.sym_units.size(), f.sym_stages.size(), f.sym_tiles.size() / 17, nblocks);
  upload(f.d_symu, f.sym_units);
  upload(f.d_syms, f.sym_stages);
  upload(f.d_symt, f.sym_tiles);
  upload(f.d_symr, f.sym_red);
  upload(f.d_symri, f.sym_ridx);
  return true;
}

// Every (factor type, time key) a leaf of this family reads at slice pair
// (t_so, t_si); the slab index is returned through `base`.
std::vector<int> trace_times(Family const &f, int tso, int tsi) {
  std::vector<int> tt(f.locs.size());
  for (size_t i = 0; i < f.locs.size(); ++i) tt[i] = f.locs[i] == Loc::source ? tso : tsi;
  return tt;
}

// Time part of one chunk of slice pairs, per product/trace node that owns a
// table: where its table starts in the chunk's product buffer, its time
// count, and its operands' time bases (uploaded once).
struct NodeChunk {
  long long base = 0;
  int ntime = 0;
  std::vector<long long> h_Xt, h_Yt;
  DevBuf d_Xt, d_Yt;
  bool grid = false;  // the trace runs as one slice-pair tile
  DevBuf d_tile;
};
struct ChunkPlan {
  size_t bp = 0, c0 = 0;  // block pair, first slice pair
  int S = 0;
  long long prod_total = 0, g_total = 0;
  std::vector<long long> gbase;               // per family: offset of its group sums
  std::vector<long long> zbase;               // per family: offset of its subset sums
  long long z_total = 0;
  std::vector<long long> pbase;               // per family: offset of its partial trace sums
  long long p_total = 0;
  std::vector<std::vector<NodeChunk>> nodes;  // per family, per node
  DevBuf d_tds, d_tdg;                        // slice pairs grouped by time separation (ascending in a group)
  int ntdg = 0;
  std::vector<DevBuf> d_jobs;                 // per diagram: resolve jobs
  std::vector<char> zonly;                    // per diagram: every job in the Moebius form
};

}  // namespace

struct ContractCache {
  int single_block_pair = -2;
  // Kahan form (reference build with SLAPH_KAHAN): every trace value is its
  // own group, summed in the reference's order by the resolve kernel
  bool per_element = false;
  std::vector<Family> fams;
  std::vector<CorrPlan> corrs;
  std::vector<std::vector<SlicePair>> bp_slices;
  std::vector<uint16_t> masks, gmasks;
  std::vector<std::vector<int>> need1, need2;
  bool post = false;  // mask groups, mirrors and normalisations are planned
  DevBuf d_gmasks;
  std::vector<uint16_t> zmask;  // subset list of the Moebius form (empty: not used)
  DevBuf d_zmask;
  std::vector<ChunkPlan> chunks;
  DevBuf d_prod, d_gv, d_cval, d_zv, d_res, d_pv;
  std::vector<hipEvent_t> ev;  // event pairs around the left-stationary trace launches
  // the launch sequence of one configuration (every chunk, resolve,
  // accumulation and finalisation), captured on the second call of the plan
  // and replayed by later calls
  hipGraph_t graph = nullptr;
  hipGraphExec_t gexec = nullptr;
  ~ContractCache() {
    if (gexec) (void)hipGraphExecDestroy(gexec);
    if (graph) (void)hipGraphDestroy(graph);
    for (auto e : ev) (void)hipEventDestroy(e);
  }
  std::vector<DevBuf> d_sum;  // per diagram: Lt x n_corr accumulators
};

// A restricted contraction: the given correlator requests at one slice pair
// (diluted_trace runs every trace id of one trace factory this way and reads
// the per-element trace values back instead of correlators).
struct Engine::ContractScope {
  std::map<std::string, std::vector<CorrelatorRequest>> requests;
  SlicePair slice{0, 0};
};

void Engine::contract(int single_block_pair, int kahan) { run_contract(plan_, single_block_pair, kahan, nullptr); }

void Engine::run_contract(std::shared_ptr<ContractCache> &plan_, int single_block_pair, int kahan,
                          ContractScope const *scope) {
  require_inputs();
  if (!have_ops_) throw std::runtime_error("operators have not been created");
  SLAPH_HIP(hipSetDevice(device_));
  auto t_start = std::chrono::steady_clock::now();
  const int dilT = gd_.dilT();
  auto B = [dilT](int t) { return t / dilT; };  // DilutionScheme::time_to_block, block type

  // The plan depends on the input file only; it is built on the first call
  // (or after configure / a budget change) and reused for every later gauge
  // configuration.
  const bool per_element = kahan != 0;
  const bool fresh = !plan_ || plan_->single_block_pair != single_block_pair || plan_->per_element != per_element;
  if (fresh) {
    plan_ = std::make_shared<ContractCache>();
    plan_->single_block_pair = single_block_pair;
    plan_->per_element = per_element;
  }
  // a call that throws leaves no half-built plan behind
  struct PlanGuard {
    std::shared_ptr<ContractCache> &p;
    bool done = false;
    ~PlanGuard() {
      if (!done) p.reset();
    }
  } guard{plan_};
  ContractCache &C = *plan_;
  auto &fams = C.fams;
  auto &corrs = C.corrs;
  auto &bp_slices = C.bp_slices;
  auto &masks = C.masks;
  auto &gmasks = C.gmasks;
  compute_q_offsets();
  if (fresh) {
    // ---- families and correlators (time independent) ----
    std::map<std::pair<std::string, std::vector<Loc>>, int> fam_of;
    for (auto const &kv : scope ? scope->requests : gd_.correlator_requests_map) {
      if (kv.second.empty()) continue;
      CorrPlan cp;
      cp.name = kv.first;
      for (auto const &req : kv.second) {
        if (req.traces.empty() || req.traces.size() > 3) throw std::runtime_error("This many traces are not implemented yet.");
        std::vector<std::pair<int, int>> refs;
        for (auto const &tr : req.traces) {
          auto key = std::make_pair(tr.name, tr.locs);
          auto it = fam_of.find(key);
          if (it == fam_of.end()) {
            Family f;
            f.name = tr.name;
            f.locs = tr.locs;
            f.shape = make_shape(tr.name);
            if (int(f.locs.size()) != f.shape.ntimes) throw std::logic_error("trace time count mismatch");
            fams.push_back(std::move(f));
            it = fam_of.emplace(key, int(fams.size()) - 1).first;
          }
          Family &f = fams[it->second];
          auto jt = f.trl_of.find(tr.tr_id);
          if (jt == f.trl_of.end()) {
            f.tr_ids.push_back(tr.tr_id);
            jt = f.trl_of.emplace(tr.tr_id, int(f.tr_ids.size()) - 1).first;
          }
          refs.push_back({it->second, jt->second});
        }
        cp.reqs.push_back(refs);
        cp.datasets.push_back(req.dataset);
      }
      // device order: by traces (family, trace id), ties in request order
      cp.orig.resize(cp.reqs.size());
      for (size_t i = 0; i < cp.orig.size(); ++i) cp.orig[i] = int(i);
      std::stable_sort(cp.orig.begin(), cp.orig.end(), [&](int x, int y) { return cp.reqs[x] < cp.reqs[y]; });
      {
        std::vector<std::vector<std::pair<int, int>>> sorted(cp.reqs.size());
        for (size_t i = 0; i < cp.orig.size(); ++i) sorted[i] = cp.reqs[cp.orig[i]];
        cp.reqs = std::move(sorted);
      }
      upload(cp.d_orig, cp.orig);
      corrs.push_back(std::move(cp));
    }

    // ---- the slice pairs of every block pair (Correlators.cpp:84-127) ----
    auto bps = block_pairs(Lt_, dilT);
    for (size_t b = 0; b < bps.size() && !scope; ++b) {
      if (single_block_pair >= 0 && int(b) != single_block_pair) continue;
      bp_slices.push_back(slice_pairs(bps[b].first, bps[b].second, Lt_, dilT));
    }
    if (scope) bp_slices.push_back({scope->slice});
    // (factor type, time key) of a leaf at trace times tt
    auto leaf_key = [&](LeafSpec const &ls, std::vector<int> const &tt) {
      if (ls.q == 0) return std::array<int, 4>{0, tt[ls.a], 0, 0};
      if (ls.q == 1) return std::array<int, 4>{1, tt[ls.a], B(tt[ls.b]), 0};
      return std::array<int, 4>{2, B(tt[ls.a]), tt[ls.b], B(tt[ls.c])};
    };
    // Do the left and right operands of the family's trace read the same
    // factors at every slice pair?
    auto operands_coincide = [&](Family const &f) {
      auto const &rs = f.shape.nodes.back();
      if (rs.kind != kTrace2) return false;
      std::function<void(int, std::vector<int> &)> leaves_of = [&](int n, std::vector<int> &out) {
        auto const &ns = f.shape.nodes[n];
        if (ns.kind == kLeaf) {
          out.push_back(ns.leaf);
          return;
        }
        leaves_of(ns.left, out);
        if (ns.right >= 0) leaves_of(ns.right, out);
      };
      std::vector<int> lL, lR;
      leaves_of(rs.left, lL);
      leaves_of(rs.right, lR);
      if (lL.size() != lR.size()) return false;
      for (auto const &sl : bp_slices)
        for (auto const &sp : sl) {
          auto tt = trace_times(f, sp.source, sp.sink);
          for (size_t k = 0; k < lL.size(); ++k)
            if (leaf_key(f.shape.leaves[lL[k]], tt) != leaf_key(f.shape.leaves[lR[k]], tt)) return false;
        }
      return true;
    };

    // ---- quantum-number planning per family ----
    static const char *qname[3] = {"Q0", "Q1", "Q2"};
    Templates tm;
    for (auto &f : fams) {
      auto const &cl = gd_.trace_indices_map.at(f.name);
      f.nodes.resize(f.shape.nodes.size());
      f.coincide = operands_coincide(f);
      // Product nodes with the same subtree shape (same factor types in the same
      // places) share one product table: an alias keeps its own time/qn index
      // maps into its owner's storage, so an instance (time key, operator
      // entries) is computed once however many nodes read it.  Exception: when
      // the trace's operands do not coincide, a product that is the trace's
      // right operand keeps a table of its own, stored transposed, so that
      // tr(X Y) = sum_e X[e] Y^T[e] reads both operands in memory order.
      std::vector<std::string> sig(f.shape.nodes.size());
      const int rroot = f.shape.nodes.back().kind == kTrace2 ? f.shape.nodes.back().right : -1;
      for (size_t n = 0; n < f.shape.nodes.size(); ++n) {
        auto const &ns = f.shape.nodes[n];
        if (ns.kind == kLeaf) sig[n] = "L" + std::to_string(f.shape.leaves[ns.leaf].q);
        else sig[n] = "P(" + sig[ns.left] + "," + (ns.right >= 0 ? sig[ns.right] : "") + ")";
        f.nodes[n].owner = int(n);
        f.nodes[n].transposed = false;
        if (ns.kind != kProd) continue;
        if (!f.coincide && int(n) == rroot) {
          f.nodes[n].transposed = true;
          continue;
        }
        for (size_t m = 0; m < n; ++m)
          if (f.shape.nodes[m].kind == kProd && sig[m] == sig[n] && !f.nodes[m].transposed) {
            f.nodes[n].owner = f.nodes[m].owner;
            break;
          }
      }
      for (size_t n = 0; n < f.shape.nodes.size(); ++n) {
        auto const &ns = f.shape.nodes[n];
        NodePlan &self = f.nodes[n];
        NodePlan &np = f.nodes[self.owner];
        bool root = n + 1 == f.shape.nodes.size();
        self.trl_qn.resize(f.tr_ids.size());
        size_t nqn_before = np.qn_key.size();
        for (size_t trl = 0; trl < f.tr_ids.size(); ++trl) {
          auto const &c = cl.at(f.tr_ids[trl]);
          if (ns.kind == kLeaf) {
            auto const &ls = f.shape.leaves[ns.leaf];
            self.trl_qn[trl] = c.at(ls.cidx);  // leaf qn = operator entry
            continue;
          }
          std::vector<int> key;
          auto append = [&](int child) {
            auto const &cs = f.shape.nodes[child];
            if (cs.kind == kLeaf) key.push_back(c.at(f.shape.leaves[cs.leaf].cidx));
            else {
              NodePlan const &co = f.nodes[f.nodes[child].owner];
              auto const &k2 = co.qn_key[f.nodes[child].trl_qn[trl]];
              key.insert(key.end(), k2.begin(), k2.end());
            }
          };
          append(ns.left);
          if (ns.right >= 0) append(ns.right);
          if (root) key.push_back(int(trl));  // every trl is its own trace output
          auto it = np.qn_index.find(key);
          if (it == np.qn_index.end()) {
            it = np.qn_index.emplace(key, int(np.qn_key.size())).first;
            np.qn_key.push_back(key);
            np.qn_left.push_back(f.nodes[ns.left].trl_qn[trl]);
            np.qn_right.push_back(ns.right >= 0 ? f.nodes[ns.right].trl_qn[trl] : 0);
          }
          self.trl_qn[trl] = it->second;
        }
        if (ns.kind == kLeaf) continue;
        if (self.owner != int(n) && np.qn_key.size() == nqn_before) continue;  // alias, nothing new
        // templates and combos (of the owner, over all its qn entries)
        auto const &os = f.shape.nodes[self.owner];
        auto tmpl_of = [&](int child, int cqn) {
          auto const &cs = f.shape.nodes[child];
          if (cs.kind == kLeaf) {
            auto const &ls = f.shape.leaves[cs.leaf];
            return tm.leaf(gd_.quarkline_lookuptable.at(qname[ls.q]).at(cqn).rnd);
          }
          return f.nodes[f.nodes[child].owner].qn_tmpl[cqn];
        };
        np.qn_off.assign(np.qn_key.size() + 1, 0);
        np.qn_tmpl.assign(np.qn_key.size(), -1);
        np.combos.clear();
        np.cmaxX = np.cmaxY = np.cmaxO = 0;
        if (root) {
          f.mask_off.assign(f.tr_ids.size(), 0);
          f.ncomb.assign(f.tr_ids.size(), 0);
        }
        for (size_t q = 0; q < np.qn_key.size(); ++q) {
          int lt = tmpl_of(os.left, np.qn_left[q]);
          PairTmpl const *pt;
          if (ns.kind == kProd) {
            pt = &tm.product(lt, tmpl_of(os.right, np.qn_right[q]));
            np.qn_tmpl[q] = pt->out;
          } else if (ns.kind == kTrace2) {
            pt = &tm.trace2(lt, tmpl_of(os.right, np.qn_right[q]));
          } else {
            pt = &tm.trace1(lt);
          }
          for (size_t k = 0; k < pt->combos.size(); ++k)
            np.combos.push_back({int(q), pt->combos[k][0], pt->combos[k][1], np.qn_off[q] + (long long)k});
          np.qn_off[q + 1] = np.qn_off[q] + (long long)pt->combos.size();
          if (root) {
            int trl = np.qn_key[q].back();
            f.mask_off[trl] = int(masks.size());
            f.ncomb[trl] = int(pt->combos.size());
            masks.insert(masks.end(), pt->masks.begin(), pt->masks.end());
          }
        }
        np.stride = np.qn_off.back();
        // operand qn offsets: leaf child -> first pair of the operator entry;
        // product child -> its qn offset
        std::vector<long long> Xq(np.qn_key.size()), Yq(np.qn_key.size(), 0);
        auto off_of = [&](int child, int cqn) -> long long {
          auto const &cs = f.shape.nodes[child];
          if (cs.kind == kLeaf) return q_off_[f.shape.leaves[cs.leaf].q].at(cqn);
          return f.nodes[f.nodes[child].owner].qn_off[cqn];
        };
        for (size_t q = 0; q < np.qn_key.size(); ++q) {
          Xq[q] = off_of(os.left, np.qn_left[q]);
          if (os.right >= 0) Yq[q] = off_of(os.right, np.qn_right[q]);
        }
        for (auto const &cb : np.combos) {
          np.cmaxX = std::max(np.cmaxX, Xq[cb.qn] + cb.i);
          np.cmaxY = std::max(np.cmaxY, Yq[cb.qn] + cb.j);
          np.cmaxO = std::max(np.cmaxO, cb.o);
        }
        upload(np.d_Xq, Xq);
        upload(np.d_Yq, Yq);
        upload(np.d_combos, np.combos);
        np.h_Xq = std::move(Xq);
        np.h_Yq = std::move(Yq);
      }
    }

    // ---- factor keys the slice pairs need ----
    std::set<std::vector<int>> need1, need2;
    for (auto const &sl : bp_slices)
      for (auto const &sp : sl)
        for (auto const &f : fams) {
          auto tt = trace_times(f, sp.source, sp.sink);
          for (auto const &ls : f.shape.leaves) {
            if (ls.q == 1) need1.insert({tt[ls.a], B(tt[ls.b])});
            if (ls.q == 2) need2.insert({B(tt[ls.a]), tt[ls.b], B(tt[ls.c])});
          }
        }
    C.need1.assign(need1.begin(), need1.end());
    C.need2.assign(need2.begin(), need2.end());
  }
  // Storage precision forms (FP32, 48-bit): the factors are built in FP64 and
  // stored rounded for the products and traces (arithmetic stays FP64)
  const int f32 = prec_factors_;  // storage form of factor slabs and product tables
  const size_t elem = storage_bytes(f32);
  {
    auto t0 = std::chrono::steady_clock::now();
    build_factors(C.need1, C.need2);
    if (f32) {
      DevBuf *src[3] = {&q0_, &q1_, &q2_}, *dst[3] = {&q0f_, &q1f_, &q2f_};
      for (int q = 0; q < 3; ++q) {
        long long n = (long long)(src[q]->bytes() / sizeof(double2));
        dst[q]->ensure(std::max<size_t>(16, elem * n));
        launch_to_storage(src[q]->as<double2>(), dst[q]->as<void>(), n, f32, stream_);
      }
      SLAPH_HIP(hipStreamSynchronize(stream_));
    }
    times_.factors_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  }
  auto slab_of = [&](int q) -> const void * {
    if (f32) return q == 0 ? q0f_.as<void>() : (q == 1 ? q1f_.as<void>() : q2f_.as<void>());
    return q == 0 ? q0_.as<void>() : (q == 1 ? q1_.as<void>() : q2_.as<void>());
  };
  auto leaf_base = [&](LeafSpec const &ls, std::vector<int> const &tt) -> long long {
    if (ls.q == 0) return (long long)tt[ls.a] * q_total_[0];
    if (ls.q == 1) return ((long long)tt[ls.a] * nb_ + B(tt[ls.b])) * q_total_[1];
    int slot = q2slot_.at(((size_t)B(tt[ls.a]) * Lt_ + tt[ls.b]) * nb_ + B(tt[ls.c]));
    if (slot < 0) throw std::logic_error("Q2 key was not built");
    return (long long)slot * q_total_[2];
  };
  if (!C.post) {
    // ---- mirror traces: when the two operands of a trace read the same table
    // at the same time for every slice pair, tr(X_a Y_b) = tr(X_b Y_a) with the
    // combinations swapped and the same used-index masks, so (b, a) reuses the
    // sums of (a, b) ----
    for (auto &f : fams) {
      size_t ntrl = f.tr_ids.size();
      f.mirror_of.assign(ntrl, -1);
      auto const &rs = f.shape.nodes.back();
      if (rs.kind != kTrace2) continue;
      auto const &ln = f.shape.nodes[rs.left], &rn = f.shape.nodes[rs.right];
      bool same_table = ln.kind == kLeaf && rn.kind == kLeaf
                            ? f.shape.leaves[ln.leaf].q == f.shape.leaves[rn.leaf].q
                            : (ln.kind == kProd && rn.kind == kProd && f.nodes[rs.left].owner == f.nodes[rs.right].owner);
      if (!same_table || !f.coincide || per_element) continue;
      NodePlan const &root = f.nodes.back();
      std::map<std::pair<int, int>, int> first;
      for (size_t trl = 0; trl < ntrl; ++trl) {
        int q = root.trl_qn[trl];
        first.emplace(std::make_pair(root.qn_left[q], root.qn_right[q]), int(trl));
      }
      for (size_t trl = 0; trl < ntrl; ++trl) {
        int q = root.trl_qn[trl];
        auto it = first.find(std::make_pair(root.qn_right[q], root.qn_left[q]));
        if (it == first.end() || it->second >= int(trl) || f.mirror_of[it->second] >= 0) continue;
        int c = it->second;
        if (f.ncomb[c] != f.ncomb[trl]) continue;
        std::vector<uint16_t> ma(masks.begin() + f.mask_off[c], masks.begin() + f.mask_off[c] + f.ncomb[c]);
        std::vector<uint16_t> mb(masks.begin() + f.mask_off[trl], masks.begin() + f.mask_off[trl] + f.ncomb[trl]);
        std::sort(ma.begin(), ma.end());
        std::sort(mb.begin(), mb.end());
        if (ma == mb) f.mirror_of[trl] = c;
      }
    }
    // A trace vector that only single-trace correlators read needs one sum over
    // all its combinations, not one per mask (resolve_request1, Diagram.cpp:9-27).
    std::vector<std::vector<char>> multi(fams.size());
    for (size_t fi = 0; fi < fams.size(); ++fi) multi[fi].assign(fams[fi].tr_ids.size(), 0);
    for (auto const &cp : corrs)
      for (auto const &r : cp.reqs)
        if (r.size() > 1)
          for (auto const &ref : r) multi[ref.first][ref.second] = 1;
    for (size_t fi = 0; fi < fams.size(); ++fi)
      for (size_t trl = 0; trl < fams[fi].tr_ids.size(); ++trl)
        if (fams[fi].mirror_of[trl] >= 0) {
          char &a = multi[fi][trl], &b = multi[fi][fams[fi].mirror_of[trl]];
          a = b = char(a | b);
        }

    // ---- mask groups per trace vector and the exact normalisation of each
    // correlator (the num_summands of DilutedTrace.cpp:5-82, Diagram.cpp:9-27) ----
    gmasks.clear();
    for (size_t fi = 0; fi < fams.size(); ++fi) {
      Family &f = fams[fi];
      NodePlan const &root = f.nodes.back();
      size_t ntrl = f.tr_ids.size();
      f.ngroup.assign(ntrl, 0);
      f.goff.assign(ntrl, 0);
      f.gmask_off.assign(ntrl, 0);
      f.gcount.assign(ntrl, {});
      f.gitems.clear();
      f.gpairs.clear();
      long long off = 0;
      for (size_t trl = 0; trl < ntrl; ++trl) {
        if (f.mirror_of[trl] >= 0) {
          int c = f.mirror_of[trl];
          f.ngroup[trl] = f.ngroup[c];
          f.goff[trl] = f.goff[c];
          f.gmask_off[trl] = f.gmask_off[c];
          f.gcount[trl] = f.gcount[c];
          continue;
        }
        std::vector<uint16_t> gm;
        std::vector<std::vector<int>> members;
        for (int c = 0; c < f.ncomb[trl]; ++c) {
          uint16_t m = masks[f.mask_off[trl] + c];
          // per-element form: one group per trace value, in element order
          size_t g = per_element ? gm.size()
                                 : (multi[fi][trl] ? std::find(gm.begin(), gm.end(), m) - gm.begin() : 0);
          if (g == gm.size()) {
            gm.push_back(m);
            members.emplace_back();
          }
          members[g].push_back(c);
        }
        f.ngroup[trl] = int(gm.size());
        f.goff[trl] = int(off);
        f.gmask_off[trl] = int(gmasks.size());
        gmasks.insert(gmasks.end(), gm.begin(), gm.end());
        int q = root.trl_qn[trl];
        for (size_t g = 0; g < gm.size(); ++g) {
          f.gitems.push_back({q, int(f.gpairs.size()), int(members[g].size()), off + (long long)g});
          for (int c : members[g]) {
            auto const &cb = root.combos.at(root.qn_off[q] + c);  // combos of qn q are stored at qn_off[q]..
            if (cb.qn != q || cb.o != root.qn_off[q] + c) throw std::logic_error("trace combo table out of order");
            f.gpairs.push_back(make_int2(cb.i, cb.j));
          }
          f.gcount[trl].push_back(int(members[g].size()));
        }
        off += (long long)gm.size();
      }
      f.gstride = off;
      // the summed-product form (plan_xform): D = 16, the crossed trace with
      // one table for both products, every trace id read by single-trace
      // correlators only, the left-stationary kernel (not the per-element,
      // group-kernel or restricted forms); SLAPH_XFORM_OFF=1 (or the tuning
      // knob) keeps the product-per-combination form, a cross-check
      f.xform = f.name == "trQ2Q0Q2Q0" && f.coincide && D_ == 16 && !per_element && !scope &&
                f.shape.nodes.size() == 7 && f.nodes[5].owner == 2 && !std::getenv("SLAPH_XFORM_OFF") &&
                !tune_.xform_off && !std::getenv("SLAPH_TRACE_GROUPS") && trace_lstat_qchunk(D_) > 0 &&
                std::none_of(multi[fi].begin(), multi[fi].end(), [](char m) { return m != 0; });
      if (f.xform)
        f.xform = plan_xform(f, gd_.quarkline_lookuptable.at("Q0"), gd_.quarkline_lookuptable.at("Q2"), q_off_[0],
                             q_off_[2]);
      if (f.xform) {
        order_items_by_component(f, f.xroot);
        upload(f.d_gitems, f.gitems);
        upload(f.d_gpairs, f.gpairs);
        plan_lstat(f, f.xroot, D_);
        if (f.ls_blocks.empty()) throw std::logic_error("summed-product form without a left-stationary trace plan");
        // the symmetric-product traces unless SLAPH_SYM_OFF=1 (or the tuning
        // knob) keeps the left-stationary kernel, a cross-check
        if (!std::getenv("SLAPH_SYM_OFF") && !tune_.sym_off) plan_sym(f);
        continue;
      }
      order_items_by_component(f, root);
      upload(f.d_gitems, f.gitems);
      upload(f.d_gpairs, f.gpairs);
      if (!per_element) {  // the per-element form runs every trace through the group kernel
        plan_lstat(f, root, D_);
        plan_gridr(f, root);
        plan_gridg(f, root);
      }
    }
    for (auto &cp : corrs) {
      cp.n_adm.resize(cp.reqs.size());
      for (size_t i = 0; i < cp.reqs.size(); ++i) {
        auto const &r = cp.reqs[i];
        auto gmk = [&](int k, int g) { return gmasks[fams[r[k].first].gmask_off[r[k].second] + g]; };
        auto cnt = [&](int k, int g) { return (long long)fams[r[k].first].gcount[r[k].second][g]; };
        auto ng = [&](int k) { return fams[r[k].first].ngroup[r[k].second]; };
        long long n = 0;
        if (r.size() == 1) {
          n = fams[r[0].first].ncomb[r[0].second];
        } else if (r.size() == 2) {
          for (int g = 0; g < ng(0); ++g)
            for (int h = 0; h < ng(1); ++h)
              if (!(gmk(0, g) & gmk(1, h))) n += cnt(0, g) * cnt(1, h);
        } else {
          for (int g = 0; g < ng(0); ++g)
            for (int h = 0; h < ng(1); ++h)
              for (int k = 0; k < ng(2); ++k)
                if (!(gmk(0, g) & gmk(1, h) & gmk(2, k))) n += cnt(0, g) * cnt(1, h) * cnt(2, k);
        }
        if (n == 0)
          throw std::runtime_error(
              "compcomp_t inner_product(vector<DilutedTrace>, vector<DilutedTrace>) has an empty result. Not enough "
              "random vectors?");
        cp.n_adm[i] = double(n);
      }
    }
    // ---- subset sums for the Moebius form of the masked inner products:
    // the subset list is every subset of every group mask of a trace vector
    // that a multi-trace correlator reads (SLAPH_RESOLVE_GROUPS=1 keeps the
    // direct group loops, a cross-check) ----
    {
      std::set<uint16_t> zs;
      for (size_t fi = 0; fi < fams.size(); ++fi)
        for (size_t trl = 0; trl < fams[fi].tr_ids.size(); ++trl) {
          Family const &f = fams[fi];
          if (!multi[fi][trl] || f.mirror_of[trl] >= 0) continue;
          for (int g = 0; g < f.ngroup[trl]; ++g) {
            const uint16_t m = gmasks[f.gmask_off[trl] + g];
            for (uint16_t sub = m;; sub = uint16_t((sub - 1) & m)) {
              zs.insert(sub);
              if (sub == 0) break;
            }
          }
        }
      C.zmask.clear();
      if (zs.size() > 1 && zs.size() <= 4096 && !std::getenv("SLAPH_RESOLVE_GROUPS") && !tune_.resolve_groups &&
          !per_element)
        C.zmask.assign(zs.begin(), zs.end());
      const long long nz = (long long)C.zmask.size();
      for (size_t fi = 0; fi < fams.size(); ++fi) {
        Family &f = fams[fi];
        f.zslot.assign(f.tr_ids.size(), -1);
        f.zitems.clear();
        if (nz > 0)
          for (size_t trl = 0; trl < f.tr_ids.size(); ++trl) {
            if (!multi[fi][trl] || f.mirror_of[trl] >= 0) continue;
            f.zslot[trl] = int(f.zitems.size());
            f.zitems.push_back({(long long)f.goff[trl], f.ngroup[trl], f.gmask_off[trl]});
          }
        for (size_t trl = 0; trl < f.tr_ids.size(); ++trl)
          if (f.mirror_of[trl] >= 0) f.zslot[trl] = f.zslot[f.mirror_of[trl]];
        f.zstride = (long long)f.zitems.size() * nz;
        upload(f.d_zitems, f.zitems);
      }
      upload(C.d_zmask, C.zmask);
    }
    upload(C.d_gmasks, gmasks);
    C.post = true;
  }

  const long long DD = (long long)D_ * D_;
  // SLAPH_TRACE_GROUPS=1 selects the one-wave-per-group trace kernel for every
  // trace (a cross-check of the tiled and left-stationary kernels).
  const bool lstat_off = std::getenv("SLAPH_TRACE_GROUPS") != nullptr || per_element;
  long long gstride_all = 0, npart_all = 0;
  for (auto const &f : fams) gstride_all += f.gstride;
  if (!lstat_off)
    for (auto const &f : fams) npart_all += f.npart();
  size_t max_corr = 1;
  for (auto const &c : corrs) max_corr = std::max(max_corr, c.reqs.size());
  const bool progress = std::getenv("SLAPH_PROGRESS") != nullptr;
  auto slab_count = [&](Family const &f, int child, long long prod_total) -> long long {
    auto const &cs = f.shape.nodes[child];
    if (cs.kind != kLeaf) return prod_total;
    int q = f.shape.leaves[cs.leaf].q;
    DevBuf const &b = q == 0 ? q0_ : (q == 1 ? q1_ : q2_);
    return (long long)(b.bytes() / (sizeof(double2) * DD));
  };

  // ---- per-chunk time plans (built once) ----
  // Chunks are whole rows (one source time, all sink times of the block) of
  // one direction of the block pair, so that time-separable traces see a
  // rectangular tile of slice pairs.  A chunk starts as a whole pass (dilT
  // rows); the row count halves while its plan exceeds the memory budget
  // (product tables + group sums + correlator values; 0 = automatic: 80 % of
  // the device memory free at this point).
  if (C.chunks.empty()) {
    size_t budget = chunk_budget_;
    if (budget == 0) {
      size_t free_b = 0, total_b = 0;
      SLAPH_HIP(hipMemGetInfo(&free_b, &total_b));
      budget = free_b / 10 * 8;
    }
    const size_t pass_len = (size_t)dilT * dilT;
    size_t chunk_rows = (size_t)dilT;
    for (size_t bp = 0; bp < bp_slices.size(); ++bp) {
      auto const &sl = bp_slices[bp];
      for (size_t c0 = 0, S_ = 0; c0 < sl.size(); c0 += S_) {
        auto plan_chunk = [&](int S) -> long long {
        long long prod_total = 0;
        for (auto &f : fams) {
          std::vector<std::vector<int>> tts(S);
          for (int s = 0; s < S; ++s) tts[s] = trace_times(f, sl[c0 + s].source, sl[c0 + s].sink);
          // subtree leaf bases per node and slice pair
          std::vector<std::vector<std::vector<long long>>> keys(f.shape.nodes.size(),
                                                                std::vector<std::vector<long long>>(S));
          for (auto &np : f.nodes) {
            np.time_index.clear();
            np.time_left.clear();
            np.time_right.clear();
            np.time_leafbase_l.clear();
            np.time_leafbase_r.clear();
            np.sp_time.assign(S, 0);
          }
          // (a) time keys; aliases index into their owner's time table
          for (size_t n = 0; n < f.shape.nodes.size(); ++n) {
            auto const &ns = f.shape.nodes[n];
            NodePlan &self = f.nodes[n];
            NodePlan &np = f.nodes[self.owner];
            bool root = n + 1 == f.shape.nodes.size();
            for (int s = 0; s < S; ++s) {
              if (ns.kind == kLeaf) {
                keys[n][s] = {leaf_base(f.shape.leaves[ns.leaf], tts[s])};
                continue;
              }
              std::vector<long long> key = keys[ns.left][s];
              if (ns.right >= 0) key.insert(key.end(), keys[ns.right][s].begin(), keys[ns.right][s].end());
              keys[n][s] = key;
              int ti;
              if (root) {
                ti = s;
              } else {
                auto it = np.time_index.find(key);
                if (it != np.time_index.end()) {
                  self.sp_time[s] = it->second;
                  continue;
                }
                ti = int(np.time_index.size());
                np.time_index.emplace(key, ti);
              }
              self.sp_time[s] = ti;
              auto child_t = [&](int child, std::vector<int> &ct, std::vector<long long> &lb) {
                if (child < 0) {
                  ct.push_back(0);
                  lb.push_back(0);
                  return;
                }
                if (f.shape.nodes[child].kind == kLeaf) {
                  ct.push_back(0);
                  lb.push_back(keys[child][s][0]);
                } else {
                  ct.push_back(f.nodes[child].sp_time[s]);
                  lb.push_back(0);
                }
              };
              child_t(ns.left, np.time_left, np.time_leafbase_l);
              child_t(ns.right, np.time_right, np.time_leafbase_r);
            }
          }
          // (b) storage of the owners' product tables
          for (size_t n = 0; n + 1 < f.shape.nodes.size(); ++n) {
            NodePlan &np = f.nodes[n];
            if (f.shape.nodes[n].kind != kProd || np.owner != int(n)) continue;
            np.base = prod_total;
            prod_total += (long long)np.time_index.size() * np.stride;
          }
          for (size_t n = 0; n < f.shape.nodes.size(); ++n) {
            NodePlan &np = f.nodes[n];
            if (np.owner != int(n)) {
              np.base = f.nodes[np.owner].base;
              np.stride = f.nodes[np.owner].stride;
            }
          }
          // (c) operand time tables of the owners: leaf -> slab base; product ->
          // base + ti * stride of the child's owner
          for (size_t n = 0; n < f.shape.nodes.size(); ++n) {
            auto const &ns = f.shape.nodes[n];
            NodePlan &np = f.nodes[n];
            if (ns.kind == kLeaf || np.owner != int(n)) continue;
            bool root = n + 1 == f.shape.nodes.size();
            int ntime = root ? S : int(np.time_index.size());
            std::vector<long long> Xt(ntime), Yt(ntime, 0);
            for (int ti = 0; ti < ntime; ++ti) {
              auto const &l = f.shape.nodes[ns.left];
              NodePlan const &lo = f.nodes[f.nodes[ns.left].owner];
              Xt[ti] = l.kind == kLeaf ? np.time_leafbase_l[ti] : lo.base + np.time_left[ti] * lo.stride;
              if (ns.right >= 0) {
                auto const &r = f.shape.nodes[ns.right];
                NodePlan const &ro = f.nodes[f.nodes[ns.right].owner];
                Yt[ti] = r.kind == kLeaf ? np.time_leafbase_r[ti] : ro.base + np.time_right[ti] * ro.stride;
              }
            }
            np.h_Xt = std::move(Xt);
            np.h_Yt = std::move(Yt);
          }
        }
        return prod_total;
        };
        int S = 0;
        long long prod_total = 0;
        for (;;) {
          S_ = std::min(chunk_rows * dilT, (c0 / pass_len + 1) * pass_len - c0);
          S_ = std::min(S_, sl.size() - c0);
          S = int(S_);
          prod_total = plan_chunk(S);
          double bytes = sizeof(double2) * (double(prod_total) * DD + double(S) * (gstride_all + npart_all + max_corr));
          if (bytes <= double(budget) || chunk_rows == 1) break;
          chunk_rows = std::max<size_t>(1, chunk_rows / 2);
        }
        ChunkPlan ch;
        ch.bp = bp;
        ch.c0 = c0;
        ch.S = S;
        ch.prod_total = prod_total;
        ch.gbase.resize(fams.size());
        for (size_t fi = 0; fi < fams.size(); ++fi) {
          ch.gbase[fi] = ch.g_total;  // trace values go straight into mask-group sums
          ch.g_total += (long long)S * fams[fi].gstride;
        }
        ch.zbase.resize(fams.size());
        for (size_t fi = 0; fi < fams.size(); ++fi) {
          ch.zbase[fi] = ch.z_total;
          ch.z_total += (long long)S * fams[fi].zstride;
        }
        ch.pbase.resize(fams.size());
        for (size_t fi = 0; fi < fams.size(); ++fi) {
          ch.pbase[fi] = ch.p_total;
          if (!lstat_off) ch.p_total += (long long)S * fams[fi].npart();
        }
        ch.nodes.resize(fams.size());
        for (size_t fi = 0; fi < fams.size(); ++fi) {
          Family &f = fams[fi];
          ch.nodes[fi].resize(f.shape.nodes.size());
          for (size_t n = 0; n < f.shape.nodes.size(); ++n) {
            auto const &ns = f.shape.nodes[n];
            NodePlan &np = f.nodes[n];
            NodeChunk &nc = ch.nodes[fi][n];
            nc.base = np.base;
            if (ns.kind == kLeaf || np.owner != int(n)) continue;
            const bool root = n + 1 == f.shape.nodes.size();
            nc.ntime = root ? S : int(np.time_index.size());
            nc.h_Xt = np.h_Xt;
            nc.h_Yt = np.h_Yt;
            upload(nc.d_Xt, nc.h_Xt);
            upload(nc.d_Yt, nc.h_Yt);
            // Conservative bounds check of every operand and output index the
            // launches can form (largest time base + largest qn/combo offset).
            auto max_of = [](std::vector<long long> const &v) {
              long long m = 0;
              for (auto x : v) m = std::max(m, x);
              return m;
            };
            if (max_of(nc.h_Xt) + np.cmaxX >= slab_count(f, ns.left, prod_total) ||
                (ns.right >= 0 && max_of(nc.h_Yt) + np.cmaxY >= slab_count(f, ns.right, prod_total)) ||
                (root ? ch.gbase[fi] + (long long)(nc.ntime - 1) * f.gstride + f.gstride > ch.g_total
                      : np.base + (long long)(nc.ntime - 1) * np.stride + np.cmaxO >= prod_total))
              throw std::logic_error("contraction plan indexes outside its buffers (" + f.name + ")");
            GridTile tile;
            if (root && ns.kind == kTrace2 && !lstat_off && grid_tile(nc.h_Xt, nc.h_Yt, tile)) {
              nc.grid = true;
              upload(nc.d_tile, std::vector<GridTile>{tile});
            }
          }
        }
        {
          std::map<int, std::vector<int>> by_t;
          for (int s = 0; s < S; ++s) by_t[time_delta(sl[c0 + s].source, sl[c0 + s].sink, Lt_)].push_back(s);
          std::vector<int> tds;
          std::vector<int4> tdg;
          for (auto const &kv : by_t) {
            tdg.push_back(make_int4(kv.first, int(tds.size()), int(kv.second.size()), 0));
            tds.insert(tds.end(), kv.second.begin(), kv.second.end());
          }
          upload(ch.d_tds, tds);
          upload(ch.d_tdg, tdg);
          ch.ntdg = int(tdg.size());
        }
        ch.d_jobs.resize(corrs.size());
        ch.zonly.assign(corrs.size(), 0);
        for (size_t c = 0; c < corrs.size(); ++c) {
          auto const &cp = corrs[c];
          std::vector<CorrJob> jobs(cp.reqs.size());
          for (size_t i = 0; i < cp.reqs.size(); ++i) {
            jobs[i].ntr = int(cp.reqs[i].size());
            jobs[i].out = int(i);
            jobs[i].n = cp.n_adm[i];
            for (size_t k = 0; k < cp.reqs[i].size(); ++k) {
              auto const &f = fams[cp.reqs[i][k].first];
              int trl = cp.reqs[i][k].second;
              jobs[i].tr[k].base = ch.gbase[cp.reqs[i][k].first] + f.goff[trl];
              jobs[i].tr[k].stride = f.gstride;
              jobs[i].tr[k].ngroup = f.ngroup[trl];
              jobs[i].tr[k].gmask_off = f.gmask_off[trl];
              if (jobs[i].tr[k].base + (long long)(S - 1) * f.gstride + f.ngroup[trl] > ch.g_total ||
                  f.gmask_off[trl] + f.ngroup[trl] > (long long)gmasks.size())
                throw std::logic_error("correlator request indexes outside the group-sum buffer");
            }
            // Moebius form when every trace of the request has subset sums
            bool z_ok = cp.reqs[i].size() > 1;
            for (size_t k = 0; k < cp.reqs[i].size(); ++k)
              z_ok = z_ok && fams[cp.reqs[i][k].first].zslot[cp.reqs[i][k].second] >= 0;
            for (size_t k = 0; k < 3; ++k) {
              jobs[i].tr[k].zbase = -1;
              jobs[i].tr[k].zstride = 0;
              jobs[i].tr[k].zq = 0;
            }
            for (size_t k = 0; z_ok && k < cp.reqs[i].size(); ++k) {
              auto const &f = fams[cp.reqs[i][k].first];
              const long long nz = (long long)C.zmask.size(), nit = (long long)f.zitems.size();
              // subset-major: Z(S_q) of item i at zbase + q * nit + i
              jobs[i].tr[k].zbase = ch.zbase[cp.reqs[i][k].first] + f.zslot[cp.reqs[i][k].second];
              jobs[i].tr[k].zstride = f.zstride;
              jobs[i].tr[k].zq = int(nit);
              if (nit > INT32_MAX ||
                  jobs[i].tr[k].zbase + (long long)(S - 1) * f.zstride + (nz - 1) * nit + 1 > ch.z_total)
                throw std::logic_error("correlator request indexes outside the subset-sum buffer");
            }
          }
          ch.zonly[c] = !jobs.empty() && std::all_of(jobs.begin(), jobs.end(), [](CorrJob const &j) {
            return j.ntr > 1 && j.tr[0].zbase >= 0;
          });
          upload(ch.d_jobs[c], jobs);
        }
        C.chunks.push_back(std::move(ch));
      }
    }
  }

  // ---- accumulators and chunk buffers (kept with the plan) ----
  auto &d_sum = C.d_sum;
  d_sum.resize(corrs.size());
  for (size_t c = 0; c < corrs.size(); ++c) d_sum[c].alloc(sizeof(double2) * (size_t)Lt_ * corrs[c].reqs.size());
  {
    long long np_max = 1, ng_max = 1, nc_max = 1;
    for (auto const &ch : C.chunks) {
      np_max = std::max(np_max, ch.prod_total * DD);
      ng_max = std::max(ng_max, ch.g_total);
      nc_max = std::max(nc_max, (long long)max_corr * ch.S);
    }
    C.d_prod.ensure(elem * np_max);
    C.d_gv.ensure(sizeof(double2) * ng_max);
    C.d_cval.ensure(sizeof(double2) * nc_max);
    long long nz_max = 1, npv_max = 1;
    for (auto const &ch : C.chunks) nz_max = std::max(nz_max, ch.z_total);
    for (auto const &ch : C.chunks) npv_max = std::max(npv_max, ch.p_total);
    C.d_zv.ensure(sizeof(double2) * nz_max);
    C.d_pv.ensure(sizeof(double2) * npv_max);
  }

  // ---- launches: per chunk, products in topological order, then traces,
  // then resolve + accumulate per diagram; no host work between chunks.
  // With graphs on, the second call of a plan captures the whole sequence
  // (everything it launches is a kernel: no memset, copy or host work) and
  // later calls replay it; the statistics of the left-stationary traces are
  // then those of the last call that ran the launches directly. ----
  EventTimer ctimer(stream_);
  const bool graph = graphs_ && !fresh && !scope && !progress;
  // HIP events around the trace launches timed for the kernel statistics:
  // kernel -> (first event of each launch's pair, algorithmic bytes)
  size_t ls_ev = 0;
  std::map<std::string, std::pair<std::vector<size_t>, double>> ev_stat;
  auto ev_begin = [&]() {
    if (ls_ev + 2 > C.ev.size()) {
      C.ev.resize(ls_ev + 2, nullptr);
      SLAPH_HIP(hipEventCreate(&C.ev[ls_ev]));
      SLAPH_HIP(hipEventCreate(&C.ev[ls_ev + 1]));
    }
    if (!graph) SLAPH_HIP(hipEventRecord(C.ev[ls_ev], stream_));
  };
  auto ev_end = [&](const char *kernel, double bytes) {
    if (graph) return;
    SLAPH_HIP(hipEventRecord(C.ev[ls_ev + 1], stream_));
    ev_stat[kernel].first.push_back(ls_ev);
    ev_stat[kernel].second += bytes;
    ls_ev += 2;
  };
  if (graph && C.gexec) {
    SLAPH_HIP(hipGraphLaunch(C.gexec, stream_));
  } else {
    struct CaptureGuard {  // ends a capture that a throwing launch left open
      hipStream_t s;
      bool on;
      ~CaptureGuard() {
        if (on) {
          hipGraph_t g = nullptr;
          (void)hipStreamEndCapture(s, &g);
          if (g) (void)hipGraphDestroy(g);
        }
      }
    } capture{stream_, false};
    if (graph) {
      SLAPH_HIP(hipStreamBeginCapture(stream_, hipStreamCaptureModeRelaxed));
      capture.on = true;
    }
    for (size_t c = 0; c < corrs.size(); ++c)
      launch_zero(d_sum[c].as<double2>(), (long long)Lt_ * corrs[c].reqs.size(), stream_);
    size_t last_bp = size_t(-1);
    for (auto const &ch : C.chunks) {
      if (progress && ch.bp != last_bp) {
        // like the reference's per-block-pair message (Correlators.cpp:92-97)
        std::fprintf(stderr, "[slaph] block pair %zu of %zu (%zu slice pairs, %d per launch)\n", ch.bp,
                     bp_slices.size(), bp_slices[ch.bp].size(), ch.S);
        std::fflush(stderr);
      }
      last_bp = ch.bp;
      const int S = ch.S;
      for (size_t fi = 0; fi < fams.size(); ++fi) {
        Family &f = fams[fi];
        for (size_t n = 0; n < f.shape.nodes.size(); ++n) {
          auto const &ns = f.shape.nodes[n];
          NodePlan &np = f.nodes[n];
          NodeChunk const &nc = ch.nodes[fi][n];
          if (ns.kind == kLeaf || np.owner != int(n)) continue;
          const bool root = n + 1 == f.shape.nodes.size();
          auto src = [&](int child) -> const void * {
            if (child < 0) return nullptr;
            auto const &cs = f.shape.nodes[child];
            if (cs.kind == kLeaf) return slab_of(f.shape.leaves[cs.leaf].q);
            return C.d_prod.as<void>();
          };
          // operands stored transposed: only a trace's right operand may be
          auto is_t = [&](int child) {
            return child >= 0 && f.shape.nodes[child].kind == kProd && f.nodes[f.nodes[child].owner].transposed;
          };
          const int rt = is_t(ns.right) ? 1 : 0;
          if (is_t(ns.left) || (!root && rt)) throw std::logic_error("transposed product table read as a factor");
          const long long gb = ch.gbase[fi];
          // the summed-product form never takes the slice-pair tile kernels
          const bool grid = nc.grid && !f.xform;
          if (!root && f.xform) {
            launch_sprod(slab_of(0), nc.d_Xt.as<long long>(), slab_of(2), nc.d_Yt.as<long long>(), C.d_prod.as<void>(),
                         nc.base, np.stride, f.d_xterms.as<STerm>(), f.d_xunits.as<SpUnit>(), int(f.xunits.size()),
                         nc.ntime, f32, stream_);
          } else if (!root) {
            launch_prod(src(ns.left), nc.d_Xt.as<long long>(), np.d_Xq.as<long long>(), src(ns.right),
                        nc.d_Yt.as<long long>(), np.d_Yq.as<long long>(), C.d_prod.as<void>(), nc.base, np.stride,
                        np.d_combos.as<ComboProd>(), int(np.combos.size()), nc.ntime, D_, np.transposed ? 1 : 0, f32,
                        tune_.launch.prod_generic, stream_);
          } else if (grid && !f.gg_tiles.empty()) {
            const int ntile = int(f.gg_tiles.size()), nrun = int(f.gg_runs.size() / 3);
            const int kp = trace_gemm_parts(ntile, nrun, tune_.launch.gemm_wg_per_cu);
            if (kp > 1) f.d_ggp.ensure(sizeof(double2) * (size_t)ntile * kp * 4096);
            launch_trace_gemm(src(ns.left), src(ns.right), f.d_ggt.as<GridGTile>(), ntile, f.d_ggr.as<int>(), nrun,
                              nc.d_tile.as<GridTile>(), C.d_gv.as<double2>(), gb, f.gstride, D_, rt, f32, kp,
                              kp > 1 ? f.d_ggp.as<double2>() : nullptr, stream_);
          } else if (grid && !f.gr_groups.empty()) {
            launch_trace_grid_r(src(ns.left), src(ns.right), f.d_grg.as<GridRGroup>(), int(f.gr_groups.size()),
                                f.d_grr.as<int>(), nc.d_tile.as<GridTile>(), 1, C.d_gv.as<double2>(), gb, f.gstride,
                                D_, rt, f32, stream_);
          } else if (grid) {
            launch_trace_grid(src(ns.left), np.d_Xq.as<long long>(), src(ns.right), np.d_Yq.as<long long>(),
                              f.d_gitems.as<GroupItem>(), int(f.gitems.size()), f.d_gpairs.as<int2>(),
                              nc.d_tile.as<GridTile>(), 1, C.d_gv.as<double2>(), gb, f.gstride, D_, rt, f32, stream_);
          } else if (ns.kind == kTrace2 && !f.ls_blocks.empty() && !lstat_off) {
            // the summed-product traces as symmetric products when planned
            const bool sym = f.xform && !f.sym_units.empty();
            const long long npart = sym ? f.sym_npart : f.ls_npart;
            if (ch.pbase[fi] + (long long)nc.ntime * npart > ch.p_total)
              throw std::logic_error("trace plan indexes outside the partial-sum buffer");
            if (sym && (rt || nc.h_Xt != nc.h_Yt))
              throw std::logic_error("symmetric-product traces need both operands in one table at one time base");
            ev_begin();
            if (sym)
              launch_trace_sym(src(ns.left), nc.d_Xt.as<long long>(), f.d_symu.as<SymUnit>(), int(f.sym_units.size()),
                               f.d_syms.as<int>(), f.d_symt.as<int>(), C.d_pv.as<double2>() + ch.pbase[fi], npart,
                               nc.ntime, f32, stream_);
            else if (f.ls_queue)
              launch_trace_lstat_q(src(ns.left), nc.d_Xt.as<long long>(), src(ns.right), nc.d_Yt.as<long long>(),
                                   f.d_lsu.as<int>(), int(f.ls_blocks.size()), f.d_lst.as<LsTask>(),
                                   f.d_lsd.as<long long>(), C.d_pv.as<double2>(), ch.pbase[fi], f.ls_npart, nc.ntime, D_,
                                   rt, f32, f.d_lsq.as<int>(), tune_.launch.lstat_wg_per_cu, stream_);
            else
              launch_trace_lstat(src(ns.left), nc.d_Xt.as<long long>(), src(ns.right), nc.d_Yt.as<long long>(),
                                 f.d_lsb.as<LsBlock>(), int(f.ls_blocks.size()), f.d_lsc.as<LsChunk>(),
                                 f.d_lsl.as<int>(), f.d_lst.as<LsTask>(), f.d_lsd.as<long long>(), C.d_pv.as<double2>(),
                                 ch.pbase[fi], f.ls_npart, nc.ntime, D_, rt, f32, stream_);
            // statistics keyed by the kernel that ran (the queued
            // left-stationary form unless SLAPH_LSTAT_BLOCKS selected blocks)
            ev_end(sym ? "k_trace_sym" : f.ls_queue ? "k_trace_lstat_q" : "k_trace_lstat",
                   (sym ? f.sym_bytes : f.ls_bytes) * nc.ntime);
            launch_ls_reduce(C.d_pv.as<double2>() + ch.pbase[fi], npart,
                             sym ? f.d_symr.as<GroupItem>() : f.d_lsr.as<GroupItem>(),
                             int(sym ? f.sym_red.size() : f.ls_red.size()),
                             sym ? f.d_symri.as<int>() : f.d_lsri.as<int>(), C.d_gv.as<double2>(), gb, f.gstride,
                             nc.ntime, stream_);
          } else {
            if (f.xform) throw std::logic_error("summed-product form needs the left-stationary traces");
            launch_trace_groups(src(ns.left), nc.d_Xt.as<long long>(), np.d_Xq.as<long long>(), src(ns.right),
                                nc.d_Yt.as<long long>(), np.d_Yq.as<long long>(), f.d_gitems.as<GroupItem>(),
                                int(f.gitems.size()), f.d_gpairs.as<int2>(), C.d_gv.as<double2>(), gb, f.gstride,
                                nc.ntime, D_, rt, f32, stream_);
          }
        }
        // subset sums of this family's trace vectors for the Moebius form
        launch_zeta(C.d_gv.as<double2>(), ch.gbase[fi], f.gstride, C.d_gmasks.as<uint16_t>(), f.d_zitems.as<ZItem>(),
                    int(f.zitems.size()), C.d_zmask.as<uint16_t>(), int(C.zmask.size()), C.d_zv.as<double2>(),
                    ch.zbase[fi], f.zstride, S, stream_);
      }
      for (size_t c = 0; c < corrs.size(); ++c) {
        int ncorr = int(corrs[c].reqs.size());
        launch_resolve(C.d_gv.as<double2>(), C.d_gmasks.as<uint16_t>(), ch.d_jobs[c].as<CorrJob>(), ncorr, S,
                       C.d_zv.as<double2>(), C.d_zmask.as<uint16_t>(), int(C.zmask.size()), C.d_cval.as<double2>(),
                       per_element ? 1 : 0, ch.zonly[c], stream_);
        launch_accumulate(C.d_cval.as<double2>(), ch.d_tds.as<int>(), ch.d_tdg.as<int4>(), ch.ntdg, ncorr,
                          d_sum[c].as<double2>(), stream_);
      }
    }
    if (scope) {  // the caller reads the trace values of the plan
      SLAPH_HIP(hipStreamSynchronize(stream_));
      guard.done = true;
      return;
    }
    // ---- results: correlator_[t][i].value() / Lt (Diagram.cpp:132-147),
    // formed on the device as [correlator][t] and copied once ----
    size_t res_total = 0;
    for (auto const &cp : corrs) res_total += (size_t)Lt_ * cp.reqs.size();
    C.d_res.ensure(sizeof(double2) * std::max<size_t>(1, res_total));
    for (size_t c = 0, off = 0; c < corrs.size(); off += (size_t)Lt_ * corrs[c].reqs.size(), ++c)
      launch_finalize(d_sum[c].as<double2>(), int(corrs[c].reqs.size()), Lt_, corrs[c].d_orig.as<int>(),
                      C.d_res.as<double2>() + off, stream_);
    if (graph) {
      capture.on = false;
      SLAPH_HIP(hipStreamEndCapture(stream_, &C.graph));
      SLAPH_HIP(hipGraphInstantiate(&C.gexec, C.graph, nullptr, nullptr, 0));
      SLAPH_HIP(hipGraphLaunch(C.gexec, stream_));
    }
  }
  results_.resize(corrs.size());
  double kms = ctimer.stop();
  (void)kms;
  for (auto const &kv : ev_stat) {
    KernelStat ks;
    for (size_t e : kv.second.first) {
      float ms = 0;
      SLAPH_HIP(hipEventElapsedTime(&ms, C.ev[e], C.ev[e + 1]));
      ks.ms += ms;
    }
    ks.launches = int(kv.second.first.size());
    ks.bytes = kv.second.second;
    kstats_[kv.first] = ks;
  }
  for (size_t c = 0, off = 0; c < corrs.size(); off += (size_t)Lt_ * corrs[c].reqs.size(), ++c) {
    DiagramResult &r = results_[c];
    if (r.name != corrs[c].name || r.names.size() != corrs[c].datasets.size()) {
      r.name = corrs[c].name;
      r.names = corrs[c].datasets;
    }
    // k_finalize wrote the values in request order
    r.values.resize((size_t)Lt_ * corrs[c].reqs.size());
    SLAPH_HIP(hipMemcpy(r.values.data(), C.d_res.as<double2>() + off, sizeof(cplx) * r.values.size(),
                        hipMemcpyDeviceToHost));
  }
  times_.contract_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t_start).count();
  guard.done = true;
}

std::vector<std::vector<std::pair<cplx, uint16_t>>> Engine::diluted_trace(std::string const &name, int t_source,
                                                                         int t_sink, std::vector<Loc> const &locs) {
  // AbstractDilutedTraceFactory::get(slice_pair, locations)
  // (DilutedTraceFactory.hpp:21-29, DilutedTraceFactory.cpp:13-238): the trace
  // vectors of every trace id of this factory, one value per random-vector
  // combination in the reference's order with its used-index mask.
  require_inputs();
  if (!have_ops_) throw std::runtime_error("operators have not been created");
  auto it = gd_.trace_indices_map.find(name);
  if (it == gd_.trace_indices_map.end()) throw std::out_of_range("unknown trace factory " + name);
  if (int(locs.size()) != make_shape(name).ntimes)
    throw std::out_of_range("trace factory " + name + " takes " + std::to_string(make_shape(name).ntimes) +
                            " locations");
  if (t_source < 0 || t_source >= Lt_ || t_sink < 0 || t_sink >= Lt_) throw std::out_of_range("timeslice out of range");
  std::vector<std::vector<std::pair<cplx, uint16_t>>> out(it->second.size());
  if (it->second.empty()) return out;
  ContractScope sc;
  for (size_t id = 0; id < it->second.size(); ++id) {
    CorrelatorRequest r;
    r.traces.push_back(TraceRequest{name, int(id), locs});
    sc.requests["trace"].push_back(r);
  }
  sc.slice = SlicePair{t_source, t_sink};
  std::shared_ptr<ContractCache> plan;
  run_contract(plan, -1, 1, &sc);
  ContractCache &C = *plan;
  if (C.fams.size() != 1 || C.chunks.size() != 1) throw std::logic_error("diluted_trace: unexpected plan");
  Family const &f = C.fams[0];
  ChunkPlan const &ch = C.chunks[0];
  std::vector<cplx> gv((size_t)ch.g_total);
  SLAPH_HIP(hipMemcpy(gv.data(), C.d_gv.as<void>(), sizeof(cplx) * gv.size(), hipMemcpyDeviceToHost));
  for (size_t trl = 0; trl < f.tr_ids.size(); ++trl)
    for (int g = 0; g < f.ngroup[trl]; ++g)
      out[f.tr_ids[trl]].push_back({gv[ch.gbase[0] + f.goff[trl] + g], C.gmasks[f.gmask_off[trl] + g]});
  return out;
}

cplx Engine::inner_product(int ntr, const cplx *const *vals, const uint16_t *const *used, const int *n, int kahan) {
  if (ntr < 1 || ntr > 3) throw std::runtime_error("This many traces are not implemented yet.");
  if (!vals || !used || !n) throw std::out_of_range("inner_product: null argument");
  std::vector<cplx> hv;
  std::vector<uint16_t> hm;
  CorrJob job{};
  job.ntr = ntr;
  for (int k = 0; k < ntr; ++k) {
    if (n[k] < 1 || !vals[k] || !used[k]) throw std::out_of_range("inner_product: empty trace vector");
    job.tr[k].base = (long long)hv.size();
    job.tr[k].stride = 0;
    job.tr[k].ngroup = n[k];
    job.tr[k].gmask_off = int(hm.size());
    job.tr[k].zbase = -1;
    job.tr[k].zstride = 0;
    hv.insert(hv.end(), vals[k], vals[k] + n[k]);
    hm.insert(hm.end(), used[k], used[k] + n[k]);
  }
  // number of admitted combinations (num_summands, DilutedTrace.cpp:15-40, 56-73)
  long long adm = 0;
  if (ntr == 1) {
    adm = n[0];
  } else if (ntr == 2) {
    for (int g = 0; g < n[0]; ++g)
      for (int h = 0; h < n[1]; ++h) adm += !(used[0][g] & used[1][h]);
  } else {
    for (int g = 0; g < n[0]; ++g)
      for (int h = 0; h < n[1]; ++h)
        for (int k = 0; k < n[2]; ++k) adm += !(used[0][g] & used[1][h] & used[2][k]);
  }
  if (adm == 0)
    throw std::runtime_error(
        "compcomp_t inner_product(vector<DilutedTrace>, vector<DilutedTrace>) has an empty result. Not enough random "
        "vectors?");
  job.n = double(adm);
  SLAPH_HIP(hipSetDevice(device_));
  DevBuf dv, dm, dj, dout;
  upload(dv, hv);
  upload(dm, hm);
  upload(dj, std::vector<CorrJob>{job});
  dout.alloc(sizeof(double2));
  launch_resolve(dv.as<double2>(), dm.as<uint16_t>(), dj.as<CorrJob>(), 1, 1, nullptr, nullptr, 0, dout.as<double2>(),
                 kahan ? 1 : 0, 0, stream_);
  cplx r;
  SLAPH_HIP(hipMemcpyAsync(&r, dout.as<void>(), sizeof(cplx), hipMemcpyDeviceToHost, stream_));
  SLAPH_HIP(hipStreamSynchronize(stream_));
  return r;
}

}  // namespace slaph
