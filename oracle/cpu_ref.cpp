// CPU restatement of the reference's contraction path in C++17 + OpenMP.
//
// TEST INFRASTRUCTURE ONLY: the timed CPU baseline of bench.py and a checker
// in tests/.  It follows the numpy oracle (oracle/slaph_oracle.py) function
// by function, which follows the reference (pittlerf/sLapH-contractions):
//
//   build_vdaggerv        OperatorFactory.cpp:220-326 (OpenMP over t, :263)
//   Factories::build_Q0/1/2  DilutedFactorFactory.cpp:33-253
//   factor_product        DilutedFactor.cpp:6-51 (operator*)
//   factor_to_trace       DilutedTrace.cpp:84-165
//   Traces::build         DilutedTraceFactory.cpp:13-238 (the eight recipes)
//   inner products        DilutedTrace.cpp:5-82, Diagram.cpp:9-94
//   contract              Correlators.cpp:40-127: one DiagramParts (factor and
//                         trace caches) per thread, block pairs dealt with
//                         schedule(dynamic), caches cleared after each block
//                         pair, per-thread correlator sums added at the end.
//
// The lookup tables (quark-line entries, trace index tuples, correlator
// requests) come from the oracle's init_lookup_tables through the functions
// below.  Matrices are dense, complex<double>, row-major; traces of products
// are formed as sum_ij L_ij R_ji (D^2 operations, not the D^3 of a formed
// product).  Checked against the oracle and the reference's golden L4 files
// in tests/test_cpu_ref.py.
#include <omp.h>

#include <algorithm>
#include <array>
#include <chrono>
#include <complex>
#include <cstdint>
#include <cstring>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

namespace {

using C = std::complex<double>;

struct Factor {
  std::vector<C> m;  // D x D row-major
  int r1, r2;
  uint16_t used;
};
using FactorList = std::vector<Factor>;
using FactorEntries = std::vector<FactorList>;  // per quark-line entry

struct TraceVal {
  C v;
  uint16_t used;
};
using TraceVec = std::vector<TraceVal>;

// gamma tables: column d has value gv[d] at row gr[d] (Gamma.cpp:18-246)
struct Gamma {
  int gr[4];
  C gv[4];
};

struct QlEntry {
  int vdv, dagger, gamma;
  std::vector<std::pair<int, int>> pairs;
};

struct TraceRef {
  int kind;                 // index into kTraceNames
  int tr_id;
  std::vector<int> sink;    // per location: 1 = sink, 0 = source
};
struct Request {
  std::vector<TraceRef> trs;
};

const char *kTraceNames[] = {"trQ1", "trQ1Q1", "trQ0Q2", "trQ1Q1Q1", "trQ1Q0Q2", "trQ1Q1Q1Q1", "trQ2Q0Q2Q0",
                             "trQ2Q0Q2Q0Q2Q0"};
constexpr int kNumTraces = 8;

struct Problem {
  int Lt, N, dilT, dilE, D, nrnd, nb;
  // last contract: per Q type (seconds, keys), wall seconds, busy thread-seconds
  double stats[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  std::vector<std::vector<C>> peram;   // per r: rows x cols row-major
  long long prow = 0, pcol = 0;
  std::vector<std::vector<C>> rnd;     // per r: Lt * N * 4
  std::vector<std::vector<C>> vdv;     // [id * Lt + t]: N x N row-major
  int nvdv = 0;
  std::vector<Gamma> gammas;
  std::vector<QlEntry> ql[3];
  std::vector<std::vector<std::vector<int>>> trace_ids;  // [kind][tr_id] -> ql indices
  std::vector<std::string> diagrams;
  std::vector<std::vector<Request>> requests;  // per diagram
};

// ---------------------------------------------------------------------------
// factors
// ---------------------------------------------------------------------------
class Factories {
 public:
  explicit Factories(Problem const &p) : p_(p) {}
  void clear() {
    for (auto &c : cache_) c.clear();
  }
  FactorEntries const &get(int q, std::array<int, 3> key) {
    auto it = cache_[q].find(key);
    if (it != cache_[q].end()) return it->second;
    auto t0 = std::chrono::steady_clock::now();
    FactorEntries e = q == 0 ? build_q0(key[0]) : (q == 1 ? build_q1(key[0], key[1]) : build_q2(key[0], key[1], key[2]));
    build_s[q] += std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    ++built[q];
    return cache_[q].emplace(key, std::move(e)).first->second;
  }
  double build_s[3] = {0, 0, 0};  // seconds spent building factors of each type
  long long built[3] = {0, 0, 0};  // keys built

 private:
  // VdaggerV of an entry at t, adjoint when the entry says so
  C vdv(QlEntry const &e, int t, int i, int j) const {
    auto const &m = p_.vdv[(size_t)e.vdv * p_.Lt + t];
    return e.dagger ? std::conj(m[(size_t)j * p_.N + i]) : m[(size_t)i * p_.N + j];
  }
  C rho(int r, int t, int v, int d) const { return p_.rnd[r][((size_t)t * p_.N + v) * 4 + d]; }
  C per(int r, long long row, long long col) const { return p_.peram[r][row * p_.pcol + col]; }

  // DilutedFactorFactory<Q0>::build (DilutedFactorFactory.cpp:33-101)
  FactorEntries build_q0(int t) const {
    const int N = p_.N, dE = p_.dilE, D = p_.D;
    FactorEntries res;
    std::vector<C> V((size_t)N * N), M((size_t)N * D);
    for (auto const &e : p_.ql[0]) {
      for (int i = 0; i < N; ++i)
        for (int j = 0; j < N; ++j) V[(size_t)i * N + j] = vdv(e, t, i, j);
      Gamma const &g = p_.gammas[e.gamma];
      FactorList lst;
      for (auto const &rr : e.pairs) {
        // M[i, d dE + a] = sum_{v % dE == a} V[i, v] rho2[v, d]
        std::fill(M.begin(), M.end(), C(0));
        for (int i = 0; i < N; ++i)
          for (int v = 0; v < N; ++v) {
            const C x = V[(size_t)i * N + v];
            for (int d = 0; d < 4; ++d) M[(size_t)i * D + d * dE + v % dE] += x * rho(rr.second, t, v, d);
          }
        Factor f{std::vector<C>((size_t)D * D, C(0)), rr.first, rr.second, 0};
        for (int d = 0; d < 4; ++d) {
          const int gr = g.gr[d];
          for (int i = 0; i < N; ++i) {
            const C c1 = g.gv[d] * std::conj(rho(rr.first, t, i, gr));
            const int a = i % dE;
            for (int e2 = 0; e2 < dE; ++e2) f.m[(size_t)(gr * dE + a) * D + d * dE + e2] += c1 * M[(size_t)i * D + d * dE + e2];
          }
        }
        lst.push_back(std::move(f));
      }
      res.push_back(std::move(lst));
    }
    return res;
  }

  // DilutedFactorFactory<Q1>::build (DilutedFactorFactory.cpp:103-170)
  FactorEntries build_q1(int t, int b) const {
    const int N = p_.N, dE = p_.dilE, D = p_.D;
    FactorEntries res;
    std::vector<C> rv((size_t)D * N);
    for (auto const &e : p_.ql[1]) {
      Gamma const &g = p_.gammas[e.gamma];
      FactorList lst;
      for (auto const &rr : e.pairs) {
        // rv[d dE + a, j] = sum_{i % dE == a} conj(rho1[i, d]) V[i, j]
        std::fill(rv.begin(), rv.end(), C(0));
        for (int d = 0; d < 4; ++d)
          for (int i = 0; i < N; ++i) {
            const C c1 = std::conj(rho(rr.first, t, i, d));
            C *row = &rv[(size_t)(d * dE + i % dE) * N];
            for (int j = 0; j < N; ++j) row[j] += c1 * vdv(e, t, i, j);
          }
        Factor f{std::vector<C>((size_t)D * D, C(0)), rr.first, rr.second, 0};
        for (int row = 0; row < 4; ++row)
          for (int col = 0; col < 4; ++col)
            for (int a = 0; a < dE; ++a)
              for (int e2 = 0; e2 < dE; ++e2) {
                C s = 0;
                for (int j = 0; j < N; ++j)
                  s += rv[(size_t)(row * dE + a) * N + j] *
                       per(rr.second, (long long)(t * 4 + g.gr[row]) * N + j, (long long)(b * 4 + col) * dE + e2);
                f.m[(size_t)(row * dE + a) * D + col * dE + e2] = g.gv[row] * s;
              }
        lst.push_back(std::move(f));
      }
      res.push_back(std::move(lst));
    }
    return res;
  }

  // DilutedFactorFactory<Q2>::build (DilutedFactorFactory.cpp:172-253); the
  // gamma_5 sign is -1 iff exactly one of (row, col) is in {2, 3} (:215-216)
  FactorEntries build_q2(int b1, int t, int b2) const {
    const int N = p_.N, dE = p_.dilE, D = p_.D;
    FactorEntries res;
    std::vector<C> V((size_t)N * N), M((size_t)D * 4 * N);
    for (auto const &e : p_.ql[2]) {
      for (int i = 0; i < N; ++i)
        for (int j = 0; j < N; ++j) V[(size_t)i * N + j] = vdv(e, t, i, j);
      Gamma const &g = p_.gammas[e.gamma];
      FactorList lst;
      for (auto const &rr : e.pairs) {
        // M[col dE + a, row N + j] = +- sum_i conj(P1[(t 4 + row) N + i, (b1 4 + col) dE + a]) V[i, j]
        std::fill(M.begin(), M.end(), C(0));
        for (int row = 0; row < 4; ++row)
          for (int col = 0; col < 4; ++col) {
            const double sg = ((row + col) == 3 || std::abs(row - col) > 1) ? -1.0 : 1.0;
            for (int a = 0; a < dE; ++a) {
              C *out = &M[(size_t)(col * dE + a) * 4 * N + (size_t)row * N];
              for (int i = 0; i < N; ++i) {
                const C c1 = sg * std::conj(per(rr.first, (long long)(t * 4 + row) * N + i, (long long)(b1 * 4 + col) * dE + a));
                const C *vr = &V[(size_t)i * N];
                for (int j = 0; j < N; ++j) out[j] += c1 * vr[j];
              }
            }
          }
        Factor f{std::vector<C>((size_t)D * D, C(0)), rr.first, rr.second, 0};
        for (int bd = 0; bd < 4; ++bd) {
          const int gr = g.gr[bd];
          for (int row = 0; row < 4; ++row)
            for (int col = 0; col < 4; ++col)
              for (int a = 0; a < dE; ++a)
                for (int e2 = 0; e2 < dE; ++e2) {
                  C s = 0;
                  const C *mr = &M[(size_t)(row * dE + a) * 4 * N + (size_t)bd * N];
                  for (int j = 0; j < N; ++j)
                    s += mr[j] * per(rr.second, (long long)(t * 4 + gr) * N + j, (long long)(b2 * 4 + col) * dE + e2);
                  f.m[(size_t)(row * dE + a) * D + col * dE + e2] += g.gv[bd] * s;
                }
        }
        lst.push_back(std::move(f));
      }
      res.push_back(std::move(lst));
    }
    return res;
  }

  Problem const &p_;
  std::map<std::array<int, 3>, FactorEntries> cache_[3];
};

// operator* (DilutedFactor.cpp:6-51)
FactorList factor_product(FactorList const &L, FactorList const &R, int D) {
  FactorList out;
  for (auto const &l : L)
    for (auto const &r : R) {
      if (!(l.r2 == r.r1 && l.r1 != r.r2)) continue;
      if (l.used & r.used) continue;
      Factor f{std::vector<C>((size_t)D * D, C(0)), l.r1, r.r2, uint16_t(l.used | r.used | (1u << l.r2))};
      for (int i = 0; i < D; ++i)
        for (int k = 0; k < D; ++k) {
          const C x = l.m[(size_t)i * D + k];
          const C *rr = &r.m[(size_t)k * D];
          C *o = &f.m[(size_t)i * D];
          for (int j = 0; j < D; ++j) o[j] += x * rr[j];
        }
      out.push_back(std::move(f));
    }
  if (out.empty())
    throw std::runtime_error(
        "vector<DilutedFactor> operator*(vector<DilutedFactor>, vector<DilutedFactor>) has an empty result.");
  return out;
}

// factor_to_trace(L, R) (DilutedTrace.cpp:84-138)
TraceVec factor_to_trace(FactorList const &L, FactorList const &R, int D) {
  TraceVec out;
  for (auto const &l : L)
    for (auto const &r : R) {
      if (!(l.r2 == r.r1 && l.r1 == r.r2)) continue;
      if (l.used & r.used) continue;
      C s = 0;
      for (int i = 0; i < D; ++i)
        for (int j = 0; j < D; ++j) s += l.m[(size_t)i * D + j] * r.m[(size_t)j * D + i];
      out.push_back({s, uint16_t(l.used | r.used | (1u << l.r2) | (1u << l.r1))});
    }
  if (out.empty()) throw std::runtime_error("factor_to_trace(L,R) has an empty result. Not enough random vectors?");
  return out;
}

// factor_to_trace(v) (DilutedTrace.cpp:140-165)
TraceVec factor_to_trace1(FactorList const &v, int D) {
  TraceVec out;
  for (auto const &f : v)
    if (f.r1 == f.r2) {
      C s = 0;
      for (int i = 0; i < D; ++i) s += f.m[(size_t)i * D + i];
      out.push_back({s, uint16_t(f.used | (1u << f.r1))});
    }
  if (out.empty()) throw std::runtime_error("factor_to_trace(v) has an empty result.");
  return out;
}

// ---------------------------------------------------------------------------
// traces (DilutedTraceFactory.cpp:13-238)
// ---------------------------------------------------------------------------
class Traces {
 public:
  Traces(Problem const &p, Factories &f) : p_(p), f_(f) {}
  void clear() {  // DilutedTrace1Factory::clear is a no-op (DilutedTraceFactory.hpp:61)
    for (auto it = cache_.begin(); it != cache_.end();) it = it->first.first == 0 ? std::next(it) : cache_.erase(it);
  }
  TraceVec const &get(int kind, std::vector<int> const &key, int tr_id) {
    auto ck = std::make_pair(kind, key);
    auto it = cache_.find(ck);
    if (it == cache_.end()) it = cache_.emplace(ck, build(kind, key)).first;
    return it->second.at(tr_id);
  }

 private:
  int B(int t) const { return t / p_.dilT; }
  FactorList const &q(int qt, int a, int b, int c, int entry) { return f_.get(qt, {a, b, c}).at(entry); }

  std::vector<TraceVec> build(int kind, std::vector<int> const &k) {
    auto const &idx = p_.trace_ids.at(kind);
    const int D = p_.D;
    std::vector<TraceVec> out;
    switch (kind) {
      case 0:  // trQ1
        for (auto const &c : idx) out.push_back(factor_to_trace1(q(1, k[0], B(k[0]), 0, c[0]), D));
        break;
      case 1:  // trQ1Q1
        for (auto const &c : idx)
          out.push_back(factor_to_trace(q(1, k[0], B(k[1]), 0, c[0]), q(1, k[1], B(k[0]), 0, c[1]), D));
        break;
      case 2: {  // trQ0Q2
        const int b2 = B(k[1]);
        for (auto const &c : idx) out.push_back(factor_to_trace(q(0, k[1], 0, 0, c[1]), q(2, b2, k[0], b2, c[0]), D));
        break;
      }
      case 3: {  // trQ1Q1Q1
        std::map<std::pair<int, int>, FactorList> L1;
        for (auto const &c : idx) {
          auto key = std::make_pair(c[0], c[1]);
          if (!L1.count(key)) L1[key] = factor_product(q(1, k[0], B(k[1]), 0, c[0]), q(1, k[1], B(k[2]), 0, c[1]), D);
        }
        for (auto const &c : idx) out.push_back(factor_to_trace(L1[{c[0], c[1]}], q(1, k[2], B(k[0]), 0, c[2]), D));
        break;
      }
      case 4: {  // trQ1Q0Q2
        std::map<std::pair<int, int>, FactorList> L1;
        for (auto const &c : idx) {
          auto key = std::make_pair(c[2], c[0]);
          if (!L1.count(key)) L1[key] = factor_product(q(0, k[0], 0, 0, c[2]), q(2, B(k[0]), k[0], B(k[1]), c[0]), D);
        }
        for (auto const &c : idx) out.push_back(factor_to_trace(L1[{c[2], c[0]}], q(1, k[1], B(k[2]), 0, c[1]), D));
        break;
      }
      case 5: {  // trQ1Q1Q1Q1
        std::map<std::pair<int, int>, FactorList> L1, L2;
        for (auto const &c : idx) {
          auto k1 = std::make_pair(c[0], c[1]), k2 = std::make_pair(c[2], c[3]);
          if (!L1.count(k1)) L1[k1] = factor_product(q(1, k[0], B(k[1]), 0, c[0]), q(1, k[1], B(k[2]), 0, c[1]), D);
          if (!L2.count(k2)) L2[k2] = factor_product(q(1, k[2], B(k[3]), 0, c[2]), q(1, k[3], B(k[0]), 0, c[3]), D);
        }
        for (auto const &c : idx) out.push_back(factor_to_trace(L1[{c[0], c[1]}], L2[{c[2], c[3]}], D));
        break;
      }
      case 6: {  // trQ2Q0Q2Q0 (DilutedTraceFactory.cpp:161-190)
        const int b0 = B(k[0]), b2 = B(k[2]);
        std::map<std::pair<int, int>, FactorList> L1, L2;
        for (auto const &c : idx) {
          auto k1 = std::make_pair(c[0], c[1]), k2 = std::make_pair(c[2], c[3]);
          if (!L1.count(k1)) L1[k1] = factor_product(q(0, k[0], 0, 0, c[0]), q(2, b0, k[1], b2, c[1]), D);
          if (!L2.count(k2)) L2[k2] = factor_product(q(0, k[2], 0, 0, c[2]), q(2, b2, k[3], b0, c[3]), D);
        }
        for (auto const &c : idx) out.push_back(factor_to_trace(L1[{c[0], c[1]}], L2[{c[2], c[3]}], D));
        break;
      }
      case 7: {  // trQ2Q0Q2Q0Q2Q0
        const int b1 = B(k[1]), b3 = B(k[3]), b5 = B(k[5]);
        for (auto const &c : idx) {
          FactorList chain = factor_product(q(2, b5, k[0], b1, c[1]), q(0, k[1], 0, 0, c[2]), D);
          chain = factor_product(chain, q(2, b1, k[2], b3, c[3]), D);
          chain = factor_product(chain, q(0, k[3], 0, 0, c[4]), D);
          chain = factor_product(chain, q(2, b3, k[4], b5, c[5]), D);
          out.push_back(factor_to_trace(chain, q(0, k[5], 0, 0, c[0]), D));
        }
        break;
      }
      default:
        throw std::runtime_error("unknown trace");
    }
    return out;
  }

  Problem const &p_;
  Factories &f_;
  std::map<std::pair<int, std::vector<int>>, std::vector<TraceVec>> cache_;
};

// ---------------------------------------------------------------------------
// accumulators and inner products (KahanAccumulator.hpp:14-78,
// DilutedTrace.cpp:5-82, Diagram.cpp:9-94)
// ---------------------------------------------------------------------------
struct Acc {
  bool kahan;
  C s = 0, c = 0;
  explicit Acc(bool k) : kahan(k) {}
  static void add1(double &s, double &c, double x) {  // Neumaier (KahanAccumulator.hpp:23-32)
    const double t = s + x;
    if (std::abs(s) >= std::abs(x))
      c += (s - t) + x;
    else
      c += (x - t) + s;
    s = t;
  }
  void add(C x) {
    if (!kahan) {
      s += x;
      return;
    }
    double sr = s.real(), si = s.imag(), cr = c.real(), ci = c.imag();
    add1(sr, cr, x.real());
    add1(si, ci, x.imag());
    s = C(sr, si);
    c = C(cr, ci);
  }
  C value() const { return kahan ? s + c : s; }
};

C cdiv(C z, double n) { return C(z.real() / n, z.imag() / n); }

C resolve(std::vector<TraceVec const *> const &v, bool kahan) {
  if (v.size() == 1) {
    Acc a(kahan);
    for (auto const &x : *v[0]) a.add(x.v);
    return cdiv(a.value(), double(v[0]->size()));
  }
  long long n = 0;
  Acc res(kahan);
  if (v.size() == 2) {
    for (auto const &l : *v[0]) {
      Acc rs(kahan);
      for (auto const &r : *v[1]) {
        if (l.used & r.used) continue;
        rs.add(r.v);
        ++n;
      }
      res.add(l.v * rs.value());
    }
  } else {
    for (auto const &l : *v[0])
      for (auto const &m : *v[1]) {
        Acc rs(kahan);
        for (auto const &r : *v[2]) {
          if (l.used & m.used & r.used) continue;
          rs.add(r.v);
          ++n;
        }
        res.add(l.v * m.v * rs.value());
      }
  }
  if (n == 0) throw std::runtime_error("inner_product has an empty result. Not enough random vectors?");
  return cdiv(res.value(), double(n));
}

int time_delta(int t_so, int t_si, int Lt) {  // Correlators.cpp:21-23
  const int v = (t_si - t_so - Lt) % Lt;
  return v < 0 ? -v : v;
}

std::vector<std::pair<int, int>> block_pairs(int nb) {  // dilution-iterator.hpp:315-328
  std::vector<std::pair<int, int>> out;
  for (int a = 0; a < nb; ++a)
    for (int b = a; b < nb; ++b) out.push_back({a, b});
  return out;
}

std::vector<std::pair<int, int>> slice_pairs(int bso, int bsi, int dilT) {  // block type
  std::vector<std::pair<int, int>> out;
  auto pass = [&](int a, int b) {
    for (int s = a * dilT; s < (a + 1) * dilT; ++s)
      for (int k = b * dilT; k < (b + 1) * dilT; ++k) out.push_back({s, k});
  };
  pass(bso, bsi);
  if (bso != bsi) pass(bsi, bso);
  return out;
}

struct Timing {
  double factors_s = 0, traces_s = 0;
};

}  // namespace

extern "C" {

void *cref_new(int Lt, int N, int dilT, int dilE, int nrnd) {
  auto *p = new Problem;
  p->Lt = Lt;
  p->N = N;
  p->dilT = dilT;
  p->dilE = dilE;
  p->D = 4 * dilE;
  p->nrnd = nrnd;
  p->nb = Lt / dilT;
  p->peram.resize(nrnd);
  p->rnd.resize(nrnd);
  p->trace_ids.resize(kNumTraces);
  return p;
}

void cref_free(void *h) { delete static_cast<Problem *>(h); }

void cref_set_gamma(void *h, int g, const int *rows, const double *vals) {
  auto *p = static_cast<Problem *>(h);
  if (int(p->gammas.size()) <= g) p->gammas.resize(g + 1);
  for (int d = 0; d < 4; ++d) {
    p->gammas[g].gr[d] = rows[d];
    p->gammas[g].gv[d] = C(vals[2 * d], vals[2 * d + 1]);
  }
}

void cref_set_peram(void *h, int r, const double *P, long long rows, long long cols) {
  auto *p = static_cast<Problem *>(h);
  p->prow = rows;
  p->pcol = cols;
  p->peram[r].assign(reinterpret_cast<const C *>(P), reinterpret_cast<const C *>(P) + rows * cols);
}

void cref_set_rnd(void *h, int r, const double *rho) {
  auto *p = static_cast<Problem *>(h);
  const size_t n = (size_t)p->Lt * p->N * 4;
  p->rnd[r].assign(reinterpret_cast<const C *>(rho), reinterpret_cast<const C *>(rho) + n);
}

void cref_set_nvdv(void *h, int nvdv) {
  auto *p = static_cast<Problem *>(h);
  p->nvdv = nvdv;
  p->vdv.assign((size_t)nvdv * p->Lt, std::vector<C>());
}

void cref_set_vdv(void *h, int id, int t, const double *M) {
  auto *p = static_cast<Problem *>(h);
  p->vdv[(size_t)id * p->Lt + t].assign(reinterpret_cast<const C *>(M), reinterpret_cast<const C *>(M) + (size_t)p->N * p->N);
}

void cref_add_ql(void *h, int q, int vdv, int dagger, int gamma, int npairs, const int *pairs) {
  QlEntry e{vdv, dagger, gamma, {}};
  for (int i = 0; i < npairs; ++i) e.pairs.push_back({pairs[2 * i], pairs[2 * i + 1]});
  static_cast<Problem *>(h)->ql[q].push_back(std::move(e));
}

int cref_add_trace_id(void *h, const char *name, int n, const int *idx) {
  auto *p = static_cast<Problem *>(h);
  for (int k = 0; k < kNumTraces; ++k)
    if (std::strcmp(name, kTraceNames[k]) == 0) {
      p->trace_ids[k].push_back(std::vector<int>(idx, idx + n));
      return k;
    }
  return -1;
}

int cref_add_diagram(void *h, const char *name) {
  auto *p = static_cast<Problem *>(h);
  p->diagrams.push_back(name);
  p->requests.emplace_back();
  return int(p->diagrams.size()) - 1;
}

// one request: ntr traces; names[k], tr_ids[k], nloc[k] locations in sinks
// (concatenated, 1 = sink, 0 = source)
int cref_add_request(void *h, int diagram, int ntr, const char *const *names, const int *tr_ids, const int *nloc,
                     const int *sinks) {
  auto *p = static_cast<Problem *>(h);
  Request r;
  for (int k = 0, at = 0; k < ntr; ++k) {
    TraceRef t;
    t.kind = -1;
    for (int j = 0; j < kNumTraces; ++j)
      if (std::strcmp(names[k], kTraceNames[j]) == 0) t.kind = j;
    if (t.kind < 0) return -1;
    t.tr_id = tr_ids[k];
    t.sink.assign(sinks + at, sinks + at + nloc[k]);
    at += nloc[k];
    r.trs.push_back(std::move(t));
  }
  p->requests.at(diagram).push_back(std::move(r));
  return 0;
}

// contract (Correlators.cpp:40-127) with `threads` OpenMP threads; out holds,
// per diagram in order, n_req x Lt complex values (correlator_[t][i] / Lt).
// block_pairs_limit >= 0 restricts the run to that many block pairs from
// `first_block_pair`, and slice_pairs_limit >= 0 to the first slice pairs of
// each (for timing samples).  Returns the wall seconds, or -1 with the
// message in err.
double cref_contract(void *h, int threads, int kahan, int first_block_pair, int block_pairs_limit,
                     int slice_pairs_limit, double *out, char *err, int errlen) {
  auto *p = static_cast<Problem *>(h);
  const auto bps = block_pairs(p->nb);
  const int b0 = std::max(0, first_block_pair);
  const int b1 = block_pairs_limit < 0 ? int(bps.size()) : std::min<int>(int(bps.size()), b0 + block_pairs_limit);
  std::vector<std::vector<std::vector<C>>> total(p->diagrams.size());
  for (size_t d = 0; d < p->diagrams.size(); ++d)
    total[d].assign(p->requests[d].size(), std::vector<C>(p->Lt, C(0)));
  std::string error;
  double fstat[6] = {0, 0, 0, 0, 0, 0}, busy_total = 0;
  auto t0 = std::chrono::steady_clock::now();
#pragma omp parallel num_threads(threads)
  {
    Factories fac(*p);
    Traces tr(*p, fac);
    std::vector<std::vector<std::vector<C>>> mine = total;  // per-thread correlator sums
    double busy = 0;
#pragma omp for schedule(dynamic)
    for (int b = b0; b < b1; ++b) {
      auto tb = std::chrono::steady_clock::now();
      try {
        auto sp = slice_pairs(bps[b].first, bps[b].second, p->dilT);
        if (slice_pairs_limit >= 0 && int(sp.size()) > slice_pairs_limit) sp.resize(slice_pairs_limit);
        for (size_t d = 0; d < p->diagrams.size(); ++d)
          for (auto const &s : sp) {
            const int t = time_delta(s.first, s.second, p->Lt);
            for (size_t i = 0; i < p->requests[d].size(); ++i) {
              auto const &req = p->requests[d][i];
              std::vector<TraceVec const *> vecs;
              for (auto const &trf : req.trs) {
                std::vector<int> key;
                for (int snk : trf.sink) key.push_back(snk ? s.second : s.first);
                vecs.push_back(&tr.get(trf.kind, key, trf.tr_id));
              }
              mine[d][i][t] += resolve(vecs, kahan != 0);
            }
          }
        fac.clear();
        tr.clear();
      } catch (std::exception const &e) {
#pragma omp critical(cref_error)
        error = e.what();
      }
      busy += std::chrono::duration<double>(std::chrono::steady_clock::now() - tb).count();
    }
#pragma omp critical(cref_sum)
    {
      for (int q = 0; q < 3; ++q) {
        fstat[2 * q] += fac.build_s[q];
        fstat[2 * q + 1] += double(fac.built[q]);
      }
      busy_total += busy;
    }
#pragma omp critical(cref_sum)
    for (size_t d = 0; d < total.size(); ++d)
      for (size_t i = 0; i < total[d].size(); ++i)
        for (int t = 0; t < p->Lt; ++t) total[d][i][t] += mine[d][i][t];
  }
  const double secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  for (int k = 0; k < 6; ++k) p->stats[k] = fstat[k];
  p->stats[6] = secs;
  p->stats[7] = busy_total;
  if (!error.empty()) {
    if (err && errlen > 0) {
      std::strncpy(err, error.c_str(), size_t(errlen - 1));
      err[errlen - 1] = '\0';
    }
    return -1.0;
  }
  size_t at = 0;
  for (size_t d = 0; d < total.size(); ++d)
    for (size_t i = 0; i < total[d].size(); ++i)
      for (int t = 0; t < p->Lt; ++t, ++at) {
        out[2 * at] = total[d][i][t].real() / p->Lt;
        out[2 * at + 1] = total[d][i][t].imag() / p->Lt;
      }
  return secs;
}

// Statistics of the last cref_contract: out[2 q] seconds (summed over
// threads) building factors of type Qq, out[2 q + 1] keys built, out[6] wall
// seconds, out[7] thread-seconds inside the block-pair loop.
void cref_last_stats(void *h, double *out) {
  auto *p = static_cast<Problem *>(h);
  for (int k = 0; k < 8; ++k) out[k] = p->stats[k];
}

// VdaggerV of one momentum at one timeslice (OperatorFactory.cpp:287-300):
// out = V^+ diag(phase[k / 3]) V for the K x N column-major V, out N x N
// row-major; the product in a cache-blocked loop nest on one thread.  Returns
// the wall seconds.
double cref_vdaggerv(const double *Vd, const double *phased, long long K, int N, double *outd) {
  const C *V = reinterpret_cast<const C *>(Vd);
  const C *ph = reinterpret_cast<const C *>(phased);
  C *out = reinterpret_cast<C *>(outd);
  auto t0 = std::chrono::steady_clock::now();
  std::fill(out, out + (size_t)N * N, C(0));
  const long long KB = 256;
  std::vector<C> W((size_t)KB * N), A((size_t)KB * N);
  for (long long k0 = 0; k0 < K; k0 += KB) {
    const long long kb = std::min(KB, K - k0);
    for (int j = 0; j < N; ++j)
      for (long long k = 0; k < kb; ++k) {
        const C v = V[(size_t)j * K + k0 + k];
        W[(size_t)k * N + j] = ph[(k0 + k) / 3] * v;
        A[(size_t)k * N + j] = std::conj(v);
      }
    for (int i = 0; i < N; ++i) {
      C *o = out + (size_t)i * N;
      for (long long k = 0; k < kb; ++k) {
        const C a = A[(size_t)k * N + i];
        const C *w = &W[(size_t)k * N];
        for (int j = 0; j < N; ++j) o[j] += a * w[j];
      }
    }
  }
  return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

}  // extern "C"
