// Host-side bookkeeping; see host.hpp.
#include "host.hpp"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <sstream>
#include <unordered_map>
#include <cerrno>
#include <sys/stat.h>

namespace slaph {

namespace {

using Mat4 = std::array<std::array<cplx, 4>, 4>;

Mat4 mul4(Mat4 const &a, Mat4 const &b) {
  Mat4 c{};
  for (int i = 0; i < 4; ++i)
    for (int j = 0; j < 4; ++j) {
      cplx s = 0;
      for (int k = 0; k < 4; ++k) s += a[i][k] * b[k][j];
      c[i][j] = s;
    }
  return c;
}

// The 16 gamma structures as dense products of the tmLQCD basis
// gamma_0..gamma_3 (reference: tests/gtest/gamma.cpp:18-104 fixes this basis
// and the order of the products).
std::array<GammaEntry, 16> build_gammas() {
  const cplx I(0, 1);
  Mat4 g0{}, g1{}, g2{}, g3{}, one{};
  g0[0][2] = g0[1][3] = g0[2][0] = g0[3][1] = 1.0;
  g1[0][3] = I; g1[1][2] = I; g1[2][1] = -I; g1[3][0] = -I;
  g2[0][3] = 1.0; g2[1][2] = -1.0; g2[2][1] = -1.0; g2[3][0] = 1.0;
  g3[0][2] = I; g3[1][3] = -I; g3[2][0] = -I; g3[3][1] = I;
  for (int i = 0; i < 4; ++i) one[i][i] = 1.0;
  Mat4 g5 = mul4(mul4(g0, g1), mul4(g2, g3));
  std::array<Mat4, 16> dense = {g0, g1, g2, g3, one, g5,
                                mul4(g0, g5), mul4(g1, g5), mul4(g2, g5), mul4(g3, g5),
                                mul4(g0, g1), mul4(g0, g2), mul4(g0, g3),
                                mul4(mul4(g0, g1), g5), mul4(mul4(g0, g2), g5), mul4(mul4(g0, g3), g5)};
  std::array<GammaEntry, 16> out{};
  for (int g = 0; g < 16; ++g) {
    for (int k = 0; k < 4; ++k) {
      int hits = 0;
      for (int r = 0; r < 4; ++r) {
        if (std::abs(dense[g][r][k]) > 0.5) {
          out[g].row[k] = r;
          out[g].value[k] = dense[g][r][k];
          ++hits;
        }
      }
      if (hits != 1) throw std::logic_error("gamma matrix is not a signed permutation");
    }
  }
  return out;
}

std::vector<std::string> split(std::string const &s, char sep) {
  std::vector<std::string> out;
  std::string cur;
  for (char c : s) {
    if (c == sep) {
      out.push_back(cur);
      cur.clear();
    } else {
      cur.push_back(c);
    }
  }
  out.push_back(cur);
  return out;
}

std::string trim(std::string const &s) {
  size_t a = s.find_first_not_of(" \t\r\n");
  if (a == std::string::npos) return "";
  size_t b = s.find_last_not_of(" \t\r\n");
  return s.substr(a, b - a + 1);
}

int to_int(std::string const &s) {
  size_t pos = 0;
  int v = std::stoi(trim(s), &pos);
  if (pos != trim(s).size()) throw std::runtime_error("not an integer: " + s);
  return v;
}

Mom parse_vec3(std::string const &tok) {  // "p(1,0,-1)" / "P(0,0,1)"
  std::string in = tok.substr(2, tok.size() - 3);
  auto parts = split(in, ',');
  if (parts.size() != 3) throw std::runtime_error("bad 3-vector " + tok);
  return {to_int(parts[0]), to_int(parts[1]), to_int(parts[2])};
}

template <typename T>
int unique_push_back(std::vector<T> &vec, T const &x) {  // init_lookup_tables.cpp:337-346
  auto it = std::find(vec.begin(), vec.end(), x);
  if (it != vec.end()) return int(it - vec.begin());
  vec.push_back(x);
  return int(vec.size()) - 1;
}

template <typename T, typename Key, typename Index>
int indexed_push_back(std::vector<T> &vec, Index &index, Key const &key, T const &x) {
  auto it = index.find(key);
  if (it != index.end()) return it->second;
  vec.push_back(x);
  index.emplace(key, int(vec.size()) - 1);
  return int(vec.size()) - 1;
}

std::string mom_string(Mom const &m) {
  return std::to_string(m[0]) + std::to_string(m[1]) + std::to_string(m[2]);
}

std::string disp_string(Displacement const &d) {
  if (d.empty()) return "000";
  std::string s;
  for (auto const &x : d) {
    s.push_back(x.first);
    s.push_back(x.second);
  }
  return s;
}

}  // namespace

const std::array<GammaEntry, 16> &gamma_table() {
  static const std::array<GammaEntry, 16> g = build_gammas();
  return g;
}

void host_gamma_tables(int (*rows)[4], double (*vals)[4][2]) {
  auto const &g = gamma_table();
  for (int i = 0; i < 16; ++i)
    for (int k = 0; k < 4; ++k) {
      rows[i][k] = g[i].row[k];
      vals[i][k][0] = g[i].value[k].real();
      vals[i][k][1] = g[i].value[k].imag();
    }
}

const char *const kTraceNames[8] = {"trQ0Q2",     "trQ1",       "trQ1Q0Q2",   "trQ1Q1",
                                    "trQ1Q1Q1",   "trQ1Q1Q1Q1", "trQ2Q0Q2Q0", "trQ2Q0Q2Q0Q2Q0"};

const std::map<std::string, DiagramSpec> &diagram_specs() {
  // Vertices and quark-line cycles of each Wick diagram
  // (reference: src/DiagramSpec.cpp:3-160).
  static const std::map<std::string, DiagramSpec> m = {
      {"C1", {{0}, {}, {{{"Q1", 0, 0}}}}},
      {"C20", {{0}, {1}, {{{"Q1", 0, 1}, {"Q1", 1, 0}}}}},
      {"C2c", {{0}, {1}, {{{"Q0", 0, 1}, {"Q2", 1, 0}}}}},
      {"C20V", {{0}, {1}, {{{"Q1", 0, 0}}, {{"Q1", 1, 1}}}}},
      {"C30", {{0, 2}, {1}, {{{"Q1", 0, 1}, {"Q1", 1, 2}, {"Q1", 2, 0}}}}},
      {"C3c", {{0, 2}, {1}, {{{"Q1", 0, 1}, {"Q0", 1, 2}, {"Q2", 2, 0}}}}},
      {"C30V", {{0, 1}, {2}, {{{"Q1", 0, 1}, {"Q1", 1, 0}}, {{"Q1", 2, 2}}}}},
      {"C40B", {{0, 3}, {1, 2}, {{{"Q1", 3, 0}, {"Q1", 0, 1}, {"Q1", 1, 2}, {"Q1", 2, 3}}}}},
      {"C4cB", {{0, 3}, {1, 2}, {{{"Q2", 3, 0}, {"Q0", 0, 1}, {"Q2", 1, 2}, {"Q0", 2, 3}}}}},
      {"C40C", {{0, 2}, {1, 3}, {{{"Q1", 3, 0}, {"Q1", 0, 1}, {"Q1", 1, 2}, {"Q1", 2, 3}}}}},
      {"C4cC", {{0, 2}, {1, 3}, {{{"Q2", 3, 0}, {"Q0", 0, 1}, {"Q2", 1, 2}, {"Q0", 2, 3}}}}},
      {"C40D", {{0, 2}, {1, 3}, {{{"Q1", 0, 1}, {"Q1", 1, 0}}, {{"Q1", 2, 3}, {"Q1", 3, 2}}}}},
      {"C4cD", {{0, 2}, {1, 3}, {{{"Q0", 0, 1}, {"Q2", 1, 0}}, {{"Q0", 2, 3}, {"Q2", 3, 2}}}}},
      {"C40V", {{0, 1}, {2, 3}, {{{"Q1", 0, 1}, {"Q1", 1, 0}}, {{"Q1", 2, 3}, {"Q1", 3, 2}}}}},
      {"C4cV", {{0, 1}, {2, 3}, {{{"Q0", 0, 1}, {"Q2", 1, 0}}, {{"Q0", 2, 3}, {"Q2", 3, 2}}}}},
      {"C6cC",
       {{0, 2, 4},
        {1, 3, 5},
        {{{"Q2", 5, 0}, {"Q0", 0, 1}, {"Q2", 1, 2}, {"Q0", 2, 3}, {"Q2", 3, 4}, {"Q0", 4, 5}}}}},
      {"C6cD",
       {{0, 2, 4},
        {1, 3, 5},
        {{{"Q0", 0, 1}, {"Q2", 1, 0}}, {{"Q0", 2, 3}, {"Q2", 3, 2}}, {{"Q0", 4, 5}, {"Q2", 5, 4}}}}},
      {"C6cCD",
       {{0, 2, 4},
        {1, 3, 5},
        {{{"Q2", 3, 0}, {"Q0", 0, 1}, {"Q2", 1, 2}, {"Q0", 2, 3}}, {{"Q0", 4, 5}, {"Q2", 5, 4}}}}},
  };
  return m;
}

GlobalData::GlobalData() {
  // Defaults of GlobalData::GlobalData (reference: src/global_data.cpp:27-46).
  momentum_cutoff = {{0, 4}, {1, 5}, {2, 6}, {3, 7}, {4, 4}};
  quarkline_lookuptable["Q0"];
  quarkline_lookuptable["Q1"];
  quarkline_lookuptable["Q2"];
  for (auto n : kTraceNames) trace_indices_map[n];
}

int GlobalData::n_rnd_total() const {
  int n = 0;
  for (auto const &q : quarks) n += q.nrnd;
  return n;
}

std::vector<Mom> momentum_combinations(int p2) {
  // Enumeration order of create_all_momentum_combinations
  // (reference: global_data_input_handling_utils.cpp:54-66).
  std::vector<Mom> out;
  for (int a = -p2; a <= p2; ++a)
    for (int b = -p2; b <= p2; ++b)
      for (int c = -p2; c <= p2; ++c)
        if (a * a + b * b + c * c == p2) out.push_back({a, b, c});
  return out;
}

Quark make_quark(std::string const &s) {  // ..._utils.cpp:133-157
  auto t = split(s, ':');
  if (t.size() != 9) throw std::runtime_error("invalid quark specification: " + s);
  Quark q;
  q.type = t[0];
  q.nrnd = to_int(t[1]);
  q.dilT_type = t[2];
  q.dilT = to_int(t[3]);
  q.dilE_type = t[4];
  q.dilE = to_int(t[5]);
  q.dilD_type = t[6];
  q.dilD = to_int(t[7]);
  q.path = t[8];
  if (q.nrnd < 1) throw std::runtime_error("quarks.quark.number_of_rnd_vec must be greater than 0");
  if (q.dilD < 1 || q.dilD > 4)
    throw std::runtime_error("quarks.quark.number_of_dilution_D must be greater than 0 and smaller than 5");
  return q;
}

Operators make_operator_list(std::string const &s) {  // ..._utils.cpp:215-275
  Operators out;
  for (auto const &op_t : split(s, ':')) {
    std::vector<int> gammas;
    std::vector<Displacement> disps;
    std::vector<std::vector<Mom>> moms;
    for (auto const &tok : split(op_t, '.')) {
      if (tok.empty()) throw std::runtime_error("There is something wrong with the operators!");
      if (tok[0] == 'g') {
        gammas.push_back(to_int(tok.substr(1)));
      } else if (tok[0] == 'd') {
        if (tok.size() > 1 && tok[1] == '0') {
          disps.push_back({});
        } else if (tok.size() > 1 && (tok[1] == '<' || tok[1] == '>')) {
          for (auto const &part : split(tok.substr(1), ',')) {
            Displacement d;
            for (auto const &dir : split(part, '|')) d.push_back({dir.at(0), dir.at(1)});
            disps.push_back(d);
          }
        } else {
          throw std::runtime_error("Something wrong with the displacement in the operator definition!");
        }
      } else if (tok[0] == 'p') {
        if (tok.size() > 1 && tok[1] == '(') {
          moms = {{parse_vec3(tok)}};
        } else {
          moms.clear();
          for (auto const &x : split(tok.substr(1), ',')) moms.push_back(momentum_combinations(to_int(x)));
        }
      } else {
        throw std::runtime_error("There is something wrong with the operators!");
      }
    }
    for (auto const &grp : moms)
      for (auto const &m : grp)
        for (auto const &d : disps) out.push_back({gammas, d, m});
  }
  return out;
}

CorrelatorSpec make_correlator(std::string const &s) {  // ..._utils.cpp:302-347
  CorrelatorSpec c;
  for (auto const &tok : split(s, ':')) {
    if (tok.empty()) throw std::runtime_error("There is something wrong with the correlators in the input file!");
    if (tok[0] == 'C') {
      c.type = tok;
    } else if (tok[0] == 'Q') {
      c.quarks.push_back(to_int(tok.substr(1)));
    } else if (tok.compare(0, 2, "Op") == 0) {
      c.ops.push_back(to_int(tok.substr(2)));
    } else if (tok[0] == 'G') {
    } else if (tok[0] == 'P') {
      if (tok.size() > 1 && tok[1] == '(') {
        c.tot_mom.push_back(parse_vec3(tok));
      } else {
        for (auto const &x : split(tok.substr(1), ','))
          for (auto const &m : momentum_combinations(to_int(x))) c.tot_mom.push_back(m);
      }
    } else {
      throw std::runtime_error("There is something wrong with the correlators in the input file!");
    }
  }
  return c;
}

void parse_infile(GlobalData &gd, std::string const &text) {
  std::istringstream in(text);
  std::string line, section;
  while (std::getline(in, line)) {
    auto hash = line.find('#');
    if (hash != std::string::npos) line = line.substr(0, hash);
    line = trim(line);
    if (line.empty()) continue;
    if (line[0] == '[') {
      section = trim(line.substr(1, line.find(']') - 1));
      continue;
    }
    auto eq = line.find('=');
    if (eq == std::string::npos) throw std::runtime_error("malformed infile line: " + line);
    std::string k = trim(line.substr(0, eq)), v = trim(line.substr(eq + 1));
    std::string key = section.empty() ? k : section + "." + k;
    if (key == "quarks.quark") {
      gd.quarks.push_back(make_quark(v));
      gd.quarks.back().id = int(gd.quarks.size()) - 1;
    } else if (key == "operator_lists.operator_list") {
      gd.operator_list.push_back(make_operator_list(v));
    } else if (key == "correlator_lists.correlator_list") {
      gd.correlator_list.push_back(make_correlator(v));
    } else if (k == "Lt") {
      gd.Lt = to_int(v);
    } else if (k == "Lx") {
      gd.Lx = to_int(v);
    } else if (k == "Ly") {
      gd.Ly = to_int(v);
    } else if (k == "Lz") {
      gd.Lz = to_int(v);
    } else if (k == "number_of_eigen_vec") {
      gd.nev = to_int(v);
    } else if (k == "path_eigenvectors") {
      gd.path_eigenvectors = v;
    } else if (k == "name_eigenvectors") {
      gd.name_eigenvectors = v;
    } else if (k == "output_path") {
      gd.path_output = v;
    } else if (k == "handling_vdaggerv") {
      gd.handling_vdaggerv = v;
    } else if (k == "path_vdaggerv") {
      gd.path_vdaggerv = v;
    } else if (k == "path_config") {
      gd.path_config = v;
    } else if (k == "alpha1") {
      gd.hyp_alpha1 = std::stod(v);
    } else if (k == "alpha2") {
      gd.hyp_alpha2 = std::stod(v);
    } else if (k == "iterations") {
      gd.hyp_iterations = to_int(v);
    } else if (k == "start_config") {
      gd.start_config = to_int(v);
    } else if (k == "end_config") {
      gd.end_config = to_int(v);
    } else if (k == "delta_config") {
      gd.delta_config = to_int(v);
    } else if (k == "momentum_cutoff_0") {
      gd.momentum_cutoff[0] = to_int(v);
    }
  }
  if (gd.Lt < 1 || gd.Lx < 1 || gd.Ly < 1 || gd.Lz < 1)
    throw std::runtime_error("lattice extents Lt, Lx, Ly, Lz are mandatory and must be > 0");
  if (gd.nev < 1) throw std::runtime_error("number_of_eigen_vec is mandatory and must be > 0");
  if (gd.quarks.empty()) throw std::runtime_error("no quarks defined");
}

namespace {

// Cartesian product, index 0 fastest (reference: include/CartesianProduct.hpp:25-73).
template <typename F>
void for_each_combination(std::vector<int> const &lengths, F &&f) {
  long long total = 1;
  for (int n : lengths) total *= n;
  std::vector<int> comb(lengths.size());
  for (long long i = 0; i < total; ++i) {
    long long ii = i;
    for (size_t j = 0; j < lengths.size(); ++j) {
      comb[j] = int(ii % lengths[j]);
      ii /= lengths[j];
    }
    f(comb);
  }
}

int norm2(Mom const &m) { return m[0] * m[0] + m[1] * m[1] + m[2] * m[2]; }

// reference: build_quantum_numbers_from_correlator_list (init_lookup_tables.cpp:42-163)
std::vector<std::vector<QuantumNumbers>> build_quantum_numbers(CorrelatorSpec const &corr,
                                                               std::vector<Operators> const &oplist,
                                                               std::map<int, int> const &cutoff) {
  std::vector<Operators const *> qn_op;
  for (int n : corr.ops) {
    if (n >= int(oplist.size()))
      throw std::runtime_error("Operator with ID " + std::to_string(n) +
                               " which is used in [correlator_lists] is not defined.");
    qn_op.push_back(&oplist[n]);
  }
  auto const &spec = diagram_specs().at(corr.type);
  std::vector<int> len_src, len_snk;
  for (int v : spec.source) len_src.push_back(int(qn_op.at(v)->size()));
  // The sink product runs over the *source* vertex lengths (init_lookup_tables.cpp:78-82).
  for (int v : spec.source) len_snk.push_back(int(qn_op.at(v)->size()));
  int nvert = int(spec.source.size() + spec.sink.size());
  std::vector<std::vector<QuantumNumbers>> out;
  std::vector<QuantumNumbers> all(nvert);
  for_each_combination(len_src, [&](std::vector<int> const &isrc) {
    Mom pso{0, 0, 0};
    int sum_norm = 0;
    for (size_t j = 0; j < spec.source.size(); ++j) {
      auto const &q = qn_op.at(spec.source[j])->at(isrc[j]);
      all[spec.source[j]] = q;
      for (int a = 0; a < 3; ++a) pso[a] += q.mom[a];
      sum_norm += norm2(q.mom);
    }
    if (!corr.tot_mom.empty() && std::find(corr.tot_mom.begin(), corr.tot_mom.end(), pso) == corr.tot_mom.end())
      return;
    auto it = cutoff.find(norm2(pso));
    if (it == cutoff.end()) throw std::out_of_range("momentum_cutoff.at(" + std::to_string(norm2(pso)) + ")");
    if (sum_norm > it->second) return;
    if (spec.sink.empty()) {
      std::vector<QuantumNumbers> src;
      for (int v : spec.source) src.push_back(all[v]);
      out.push_back(src);
      return;
    }
    for_each_combination(len_snk, [&](std::vector<int> const &isnk) {
      Mom psi{0, 0, 0};
      for (size_t j = 0; j < spec.sink.size(); ++j) {
        auto const &q = qn_op.at(spec.sink[j])->at(isnk[j]);
        all[spec.sink[j]] = q;
        for (int a = 0; a < 3; ++a) psi[a] += q.mom[a];
      }
      if (pso[0] == -psi[0] && pso[1] == -psi[1] && pso[2] == -psi[2]) out.push_back(all);
    });
  });
  return out;
}

// reference: create_rnd_vec_id (init_lookup_tables.cpp:294-335)
std::vector<std::pair<int, int>> create_rnd_vec_id(std::vector<Quark> const &quarks, int q1, int q2, bool is_loop) {
  int s1 = 0, s2 = 0;
  for (int i = 0; i < q1; ++i) s1 += quarks.at(i).nrnd;
  for (int i = 0; i < q2; ++i) s2 += quarks.at(i).nrnd;
  int e1 = s1 + quarks.at(q1).nrnd, e2 = s2 + quarks.at(q2).nrnd;
  if (quarks[q1].nrnd < 1 || quarks[q2].nrnd < 1 || (q1 == q2 && quarks[q1].nrnd < 2))
    throw std::runtime_error("There are not enough random vectors for charged correlators");
  std::vector<std::pair<int, int>> out;
  if (!is_loop) {
    for (int i = s1; i < e1; ++i)
      for (int j = s2; j < e2; ++j)
        if (i != j) out.push_back({i, j});
  } else {
    for (int i = s1; i < e1; ++i) out.push_back({i, i});
  }
  return out;
}

std::string dataset_name(std::string const &type, std::vector<std::string> const &quark_types,
                         std::vector<QuantumNumbers> const &qn) {
  // reference: build_hdf5_dataset_name (init_lookup_tables.cpp:196-211)
  std::string s = type + "_";
  for (auto const &t : quark_types) s += t;
  for (auto const &op : qn) {
    s += "_p" + mom_string(op.mom);
    s += ".d" + disp_string(op.disp);
    s += ".g";
    for (int g : op.gamma) s += std::to_string(g);
  }
  return s;
}

}  // namespace

void init_lookup_tables(GlobalData &gd) {
  // unique_push_back semantics (first-seen order) with a hash index, so that
  // tens of thousands of correlator requests do not cost a quadratic search.
  std::map<std::string, std::unordered_map<std::string, int>> req_index;
  std::map<std::string, std::map<std::vector<int>, int>> tr_index;
  for (auto const &corr : gd.correlator_list) {
    if (!diagram_specs().count(corr.type)) throw std::runtime_error("unknown diagram " + corr.type);
    auto qns = build_quantum_numbers(corr, gd.operator_list, gd.momentum_cutoff);
    std::vector<std::string> quark_types;
    for (int id : corr.quarks) quark_types.push_back(gd.quarks.at(id).type);

    // build_VdaggerV_lookup (init_lookup_tables.cpp:235-264): -p reuses p with a dagger.
    std::vector<std::vector<std::pair<int, bool>>> vdv_idx;
    for (auto const &qv : qns) {
      std::vector<std::pair<int, bool>> row;
      for (auto const &q : qv) {
        Mom mp{-q.mom[0], -q.mom[1], -q.mom[2]};
        int hit = -1;
        for (size_t k = 0; k < gd.vdaggerv_lookup.size(); ++k) {
          auto const &e = gd.vdaggerv_lookup[k];
          if (e.disp == q.disp && (e.mom == q.mom || e.mom == mp)) {
            hit = int(k);
            break;
          }
        }
        if (hit < 0) {
          gd.vdaggerv_lookup.push_back({int(gd.vdaggerv_lookup.size()), q.mom, q.disp});
          row.push_back({int(gd.vdaggerv_lookup.size()) - 1, false});
        } else {
          bool zero = q.mom == Mom{0, 0, 0};
          row.push_back({hit, !zero && gd.vdaggerv_lookup[hit].mom == mp});
        }
      }
      vdv_idx.push_back(row);
    }

    auto const &spec = diagram_specs().at(corr.type);
    size_t nql = 0;
    for (auto const &tr : spec.traces) nql += tr.size();
    for (size_t d = 0; d < qns.size(); ++d) {
      std::vector<int> ql_ids(nql, -1);
      for (auto const &tr : spec.traces) {
        for (auto const &ql : tr) {
          auto ric = create_rnd_vec_id(gd.quarks, corr.quarks.at(ql.q1), corr.quarks.at(ql.q2), ql.q1 == ql.q2);
          // build_Quarkline_lookup_one_qn (init_lookup_tables.cpp:348-370) keys by the q2 vertex.
          FactorIndex cand{vdv_idx[d][ql.q2].first, vdv_idx[d][ql.q2].second, qns[d][ql.q2].gamma, ric};
          ql_ids[ql.q2] = unique_push_back(gd.quarkline_lookuptable[ql.name], cand);
        }
      }
      CorrelatorRequest req;
      req.dataset = dataset_name(corr.type, quark_types, qns[d]);
      // make_trace_request_factories + TraceRequestFactory::make (init_lookup_tables.cpp:372-462)
      for (auto const &tr : spec.traces) {
        std::string name = "tr";
        std::vector<int> ids;
        std::vector<Loc> locs;
        for (auto const &ql : tr) {
          name += ql.name;
          ids.push_back(ql_ids[ql.q1]);
          bool src = std::find(spec.source.begin(), spec.source.end(), ql.q1) != spec.source.end();
          locs.push_back(src ? Loc::source : Loc::sink);
        }
        int tr_id = indexed_push_back(gd.trace_indices_map[name], tr_index[name], ids, ids);
        req.traces.push_back({name, tr_id, locs});
      }
      // CorrelatorRequest equality compares the trace requests only (global_data.hpp:35-37).
      std::string rkey;
      for (auto const &t : req.traces) {
        rkey += t.name + ":" + std::to_string(t.tr_id) + ":";
        for (auto l : t.locs) rkey += l == Loc::source ? 'o' : 'i';
        rkey += ";";
      }
      indexed_push_back(gd.correlator_requests_map[corr.type], req_index[corr.type], rkey, req);
    }
  }
  gd.index_of_unity = -1;
  for (auto const &e : gd.vdaggerv_lookup)
    if (e.mom == Mom{0, 0, 0} && e.disp.empty()) gd.index_of_unity = e.id;
  gd.need_gaugefield = false;
  for (auto const &e : gd.vdaggerv_lookup)
    if (!e.disp.empty()) gd.need_gaugefield = true;
}

std::string lookup_dump(GlobalData const &gd) {
  std::ostringstream o;
  o << "unity " << gd.index_of_unity << "\n";
  for (auto const &v : gd.vdaggerv_lookup)
    o << "vdv " << v.id << " " << v.mom[0] << " " << v.mom[1] << " " << v.mom[2] << " " << disp_string(v.disp) << "\n";
  for (auto const &kv : gd.quarkline_lookuptable)
    for (size_t i = 0; i < kv.second.size(); ++i) {
      auto const &e = kv.second[i];
      o << kv.first << " " << i << " " << e.id_vdaggerv << " " << int(e.dagger) << " g";
      for (int g : e.gamma) o << g << ",";
      o << " r";
      for (auto const &p : e.rnd) o << p.first << ":" << p.second << ",";
      o << "\n";
    }
  for (auto const &kv : gd.trace_indices_map)
    for (size_t i = 0; i < kv.second.size(); ++i) {
      o << kv.first << " " << i;
      for (int c : kv.second[i]) o << " " << c;
      o << "\n";
    }
  for (auto const &kv : gd.correlator_requests_map)
    for (auto const &r : kv.second) {
      o << kv.first << " " << r.dataset;
      for (auto const &t : r.traces) {
        o << " " << t.name << ":" << t.tr_id << ":";
        for (auto l : t.locs) o << (l == Loc::source ? "so" : "si");
      }
      o << "\n";
    }
  return o.str();
}

std::vector<std::pair<int, int>> block_pairs(int Lt, int block_size) {
  // DilutionScheme::operator[] order (dilution-iterator.hpp:315-328).
  int nb = Lt / block_size;
  std::vector<std::pair<int, int>> out;
  for (int a = 0; a < nb; ++a)
    for (int b = a; b < nb; ++b) out.push_back({a, b});
  return out;
}

std::vector<SlicePair> slice_pairs(int bso, int bsi, int Lt, int block_size, bool interlace) {
  // BlockIterator traversal (dilution-iterator.hpp:50-124): all (source,
  // sink) of the pair, then the swapped direction unless the blocks coincide.
  std::vector<SlicePair> out;
  int passes = bso == bsi ? 1 : 2;
  for (int p = 0; p < passes; ++p) {
    int a = p == 0 ? bso : bsi, b = p == 0 ? bsi : bso;
    if (!interlace) {
      for (int s = a * block_size; s < (a + 1) * block_size; ++s)
        for (int k = b * block_size; k < (b + 1) * block_size; ++k) out.push_back({s, k});
    } else {
      for (int s = a; s < Lt; s += block_size)
        for (int k = b; k < Lt; k += block_size) out.push_back({s, k});
    }
  }
  return out;
}

std::string dilution_iterator_text(int num_slice, int num_block, bool interlace) {
  // Same text as test_dilution_scheme (src/dilution-iterator.cpp:17-36).
  char buf[128];
  std::string name = interlace ? "I" : "B";
  int bs = num_slice / num_block;
  std::snprintf(buf, sizeof buf, "T = %d, T%s%d (Morningstar), T%s%d (Other):\n\n", num_slice, name.c_str(),
                num_block, name.c_str(), bs);
  std::string out = buf;
  for (auto const &bp : block_pairs(num_slice, bs)) {
    std::snprintf(buf, sizeof buf, "%2d => %2d\n", bp.first, bp.second);
    out += buf;
    for (auto const &sp : slice_pairs(bp.first, bp.second, num_slice, bs, interlace)) {
      std::snprintf(buf, sizeof buf, "  %2d -> %2d\n", sp.source, sp.sink);
      out += buf;
    }
  }
  return out + "\n\n";
}

int time_delta(int t_source, int t_sink, int Lt) { return std::abs((t_sink - t_source - Lt) % Lt); }

std::string perambulator_filename(Quark const &q, int config, int r, int Lt, int Vs) {
  char buf[1024];
  std::snprintf(buf, sizeof buf,
                "%s/cnfg%04d/rnd_vec_%02d/perambulator.rndvecnb%02d.%s.Tso%c%04d.Vso%c%04d.Dso%c%1d."
                "TsiF%04d.SsiF%04d.DsiF4.CsiF3.smeared0.%05d",
                q.path.c_str(), config, r, r, q.type.c_str(), q.dilT_type.back(), Lt / q.dilT, q.dilE_type.back(),
                q.dilE, q.dilD_type.back(), q.dilD, Lt, Vs, config);
  return buf;
}

std::string randomvector_filename(Quark const &q, int config, int r, int nev) {
  char buf[1024];
  std::snprintf(buf, sizeof buf, "%s/cnfg%04d/rnd_vec_%02d/randomvector.rndvecnb%02d.%s.nbev%04d.%04d",
                q.path.c_str(), config, r, r, q.type.c_str(), nev, config);
  return buf;
}

std::string eigenvector_prefix(GlobalData const &gd, int config) {
  char buf[1024];
  std::snprintf(buf, sizeof buf, "%s/%s.%04d.", gd.path_eigenvectors.c_str(), gd.name_eigenvectors.c_str(), config);
  return buf;
}

std::string vdaggerv_write_name(std::string const &path, int config, Mom const &p, Displacement const &d, int t) {
  // "/%s/cnfg%04d/" + "operators.%04d.p_%s.d_%s.t_%03d" (OperatorFactory.cpp:224-226, 306-311);
  // the displacement string is "000" when there is none (global_data_typedefs.hpp:107-116)
  char buf[2048];
  std::snprintf(buf, sizeof buf, "/%s/cnfg%04d/operators.%04d.p_%s.d_%s.t_%03d", path.c_str(), config, config,
                mom_string(p).c_str(), disp_string(d).c_str(), t);
  return buf;
}

std::string vdaggerv_read_name(std::string const &path, int config, Mom const &p, int t) {
  // "/%s/cnfg%04d/operators.%04d" + ".p_" + digits + "_.t_%03d" (OperatorFactory.cpp:331-333, 352-356)
  char buf[2048];
  std::snprintf(buf, sizeof buf, "/%s/cnfg%04d/operators.%04d.p_%s_.t_%03d", path.c_str(), config, config,
                mom_string(p).c_str(), t);
  return buf;
}

std::string vdaggerv_liuming_name(std::string const &path, int config, Mom const &p) {
  // "/%s/VdaggerV." + "p" x "p" y "p" z ".conf" + "%04d" (OperatorFactory.cpp:399-427)
  char buf[2048];
  std::snprintf(buf, sizeof buf, "/%s/VdaggerV.p%dp%dp%d.conf%04d", path.c_str(), p[0], p[1], p[2], config);
  return buf;
}

void resort_perambulator(const cplx *file, cplx *out, int Lt, int nev, int dilT, int dilE, int dilD) {
  int nb = Lt / dilT;
  long long rows = 4LL * nev * Lt, ninv = (long long)nb * dilE * dilD;
  for (int t = 0; t < Lt; ++t)
    for (int v = 0; v < nev; ++v)
      for (int d = 0; d < 4; ++d)
        for (int b = 0; b < nb; ++b)
          for (int e = 0; e < dilE; ++e)
            for (int dd = 0; dd < dilD; ++dd) {
              long long frow = 4LL * nev * t + 4LL * v + d;
              long long fcol = (long long)dilD * dilE * b + (long long)dilD * e + dd;
              long long row = 4LL * nev * t + (long long)nev * d + v;
              long long col = (long long)dilE * dilD * b + (long long)dilE * dd + e;
              out[col * rows + row] = file[frow * ninv + fcol];
            }
}

void write_binary(std::string const &path, const void *src, size_t bytes) {
  FILE *fp = std::fopen(path.c_str(), "wb");
  if (!fp) throw std::runtime_error("can't open " + path);
  size_t put = std::fwrite(src, 1, bytes, fp);
  const bool ok = std::fclose(fp) == 0 && put == bytes;
  if (!ok) throw std::runtime_error("Problems while write to " + path);
}

bool file_exists(std::string const &path) {
  struct stat st;
  return stat(path.c_str(), &st) == 0 && S_ISREG(st.st_mode);
}

void make_directories(std::string const &path) {
  std::string cur;
  for (size_t at = 0; at <= path.size();) {
    size_t next = path.find('/', at);
    if (next == std::string::npos) next = path.size();
    cur = path.substr(0, next);
    struct stat st;
    if (!cur.empty() && stat(cur.c_str(), &st) != 0 && mkdir(cur.c_str(), 0777) != 0 && errno != EEXIST)
      throw std::runtime_error("cannot create directory " + cur);
    at = next + 1;
  }
}

void read_binary(std::string const &path, void *dst, size_t bytes) {
  FILE *fp = std::fopen(path.c_str(), "rb");
  if (!fp) throw std::runtime_error("failed to open file: " + path);
  size_t got = std::fread(dst, 1, bytes, fp);
  std::fclose(fp);
  if (got != bytes)
    throw std::runtime_error("short read of " + path + ": expected " + std::to_string(bytes) + " B, got " +
                             std::to_string(got));
}

}  // namespace slaph
