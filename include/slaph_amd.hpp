/*
 * slaph_amd.hpp -- the reference's C++ class interfaces for the MI355X
 * contraction path, header-only over the C ABI of slaph_amd.h.
 *
 * A C++ caller of pittlerf/sLapH-contractions binds to these classes with the
 * reference's names, key types and call shapes (all arithmetic runs on the
 * GPU behind the C ABI):
 *
 *   reference (file:line)                                  here
 *   OperatorFactory (include/OperatorsForMesons.hpp:39-69)  slaph_amd::OperatorFactory
 *   DilutedFactorFactory<Q> (DilutedFactorFactory.hpp:28-65) slaph_amd::DilutedFactorFactory<Q>
 *   DilutedFactor (DilutedFactor.hpp:19-25)                 slaph_amd::DilutedFactor
 *   AbstractDilutedTraceFactory (DilutedTraceFactory.hpp:21-29) slaph_amd::AbstractDilutedTraceFactory
 *   DilutedTrace, inner_product (DilutedTrace.hpp:16-49)    slaph_amd::DilutedTrace, inner_product
 *   contract (Correlators.hpp:41-54)                       slaph_amd::contract
 *
 * Differences a caller sees: the lookup tables, quark list and paths that the
 * reference passes to each constructor (OperatorLookup, quarkline indices,
 * GlobalData) are the configuration of a slaph_amd::Context, built from the
 * same infile text; Eigen::MatrixXcd is slaph_amd::Mat (column-major, same
 * element order); failures throw std::runtime_error with the reference's
 * messages where it has one.
 */
#ifndef SLAPH_AMD_HPP
#define SLAPH_AMD_HPP

#include "slaph_amd.h"

#include <array>
#include <complex>
#include <cstdint>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <sys/types.h>
#include <utility>
#include <vector>

namespace slaph_amd {

using Complex = std::complex<double>;
using UsedRnd = uint16_t;  // DilutedFactor.hpp:15
using RndId = int8_t;      // typedefs.hpp

enum class Location { source = 0, sink = 1 };
enum class DilutedFactorType { Q0 = 0, Q1 = 1, Q2 = 2 };

template <DilutedFactorType qlt>
struct DilutedFactorTypeTraits;  // typedefs.hpp: number of times in the key
template <>
struct DilutedFactorTypeTraits<DilutedFactorType::Q0> {
  static constexpr int num_times = 1;  // {t}
};
template <>
struct DilutedFactorTypeTraits<DilutedFactorType::Q1> {
  static constexpr int num_times = 2;  // {t, b}
};
template <>
struct DilutedFactorTypeTraits<DilutedFactorType::Q2> {
  static constexpr int num_times = 3;  // {b1, t, b2}
};

// Dense complex matrix, column-major (the element order of Eigen::MatrixXcd).
struct Mat {
  ssize_t nrows = 0, ncols = 0;
  std::vector<Complex> values;
  Mat() = default;
  Mat(ssize_t r, ssize_t c) : nrows(r), ncols(c), values(size_t(r * c)) {}
  ssize_t rows() const { return nrows; }
  ssize_t cols() const { return ncols; }
  Complex &operator()(ssize_t r, ssize_t c) { return values[size_t(c * nrows + r)]; }
  Complex const &operator()(ssize_t r, ssize_t c) const { return values[size_t(c * nrows + r)]; }
  Complex *data() { return values.data(); }
  Complex trace() const {
    Complex s = 0;
    for (ssize_t i = 0; i < nrows && i < ncols; ++i) s += (*this)(i, i);
    return s;
  }
};

// One device context (owns the slc_ctx): configuration and inputs.
class Context {
 public:
  explicit Context(int device = 0) {
    slc_ctx *h = nullptr;
    if (slc_ctx_create(device, &h) != SLC_OK) throw std::runtime_error(slc_last_error(nullptr));
    h_.reset(h);
  }
  slc_ctx *get() const { return h_.get(); }
  void check(int rc) const {
    if (rc < 0) throw std::runtime_error(slc_last_error(h_.get()));
  }
  // read_parameters + init_lookup_tables from the text of an infile
  void configure(std::string const &infile_text) { check(slc_configure(get(), infile_text.c_str())); }
  long long query(const char *what) const {
    long long v = 0;
    check(slc_query(get(), what, &v));
    return v;
  }
  // Perambulator / RandomVector / EigenVector files of one configuration
  void read_inputs(std::string const &base_dir, int config) { check(slc_read_inputs(get(), base_dir.c_str(), config)); }
  void set_eigenvectors(int t, Mat const &V) {
    check(slc_set_eigenvectors(get(), t, reinterpret_cast<const double *>(V.values.data())));
  }
  void set_perambulator(int entity, Mat const &P) {
    check(slc_set_perambulator(get(), entity, reinterpret_cast<const double *>(P.values.data())));
  }
  void set_random_vector(int entity, std::vector<Complex> const &rho) {
    check(slc_set_random_vector(get(), entity, reinterpret_cast<const double *>(rho.data())));
  }
  // GaugeField time slices: Lx Ly Lz sites x 3 directions x
  // Eigen::Matrix3cd (column-major), read from <config_path>/conf.%04d
  // (GaugeField::read_gauge_field; "" = path_config of the infile) or handed
  // over, and their smearings and plaquette
  void read_gauge_field(int config, int slice_i, int slice_f, std::string const &config_path = "",
                        std::string const &base_dir = ".") {
    check(slc_read_gauge_field(get(), base_dir.c_str(), config_path.c_str(), config, slice_i, slice_f));
  }
  double plaque_ts(int t) const {
    double p = 0.0;
    check(slc_plaquette(get(), t, &p));
    return p;
  }
  void set_gauge_links(int t, std::vector<Complex> const &U) {
    check(slc_set_gauge_links(get(), t, reinterpret_cast<const double *>(U.data())));
  }
  void smearing_stout(double rho, int iterations) { check(slc_smear_stout(get(), rho, iterations)); }
  void smearing_hyp(double alpha1, double alpha2, int iterations) {
    check(slc_smear_hyp(get(), alpha1, alpha2, iterations));
  }
  // one eigenvector timeslice on the device (call before configure)
  void set_timeslice_streaming(bool on) { check(slc_set_timeslice_streaming(get(), on ? 1 : 0)); }
  // beside the reference: precision forms, HIP graphs, autotuning
  void set_precision(int vdaggerv, int factors) { check(slc_set_precision(get(), vdaggerv, factors)); }
  void set_graphs(bool on) { check(slc_set_graphs(get(), on ? 1 : 0)); }
  std::string autotune(int repeats) {
    check(slc_autotune(get(), repeats));
    return text_of(slc_autotune_report);
  }
  std::string tuning() const { return text_of(slc_get_tuning); }
  void set_tuning(std::string const &text) { check(slc_set_tuning(get(), text.c_str())); }

 private:
  std::string text_of(int (*fn)(const slc_ctx *, char *, int)) const {
    const int need = fn(get(), nullptr, 0);
    check(need);
    std::string out(size_t(need), '\0');
    check(fn(get(), &out[0], need));
    out.resize(size_t(need) - 1);
    return out;
  }

  struct Del {
    void operator()(slc_ctx *c) const { slc_ctx_destroy(c); }
  };
  std::unique_ptr<slc_ctx, Del> h_;
};

// OperatorFactory (include/OperatorsForMesons.hpp:39-69).  The constructor
// takes the reference's sizes and checks them against the context's
// configuration (the operator lookup, handling_vdaggerv, path_vdaggerv and
// path_config come from the same infile).
class OperatorFactory {
 public:
  OperatorFactory(Context &ctx, ssize_t Lt, ssize_t Lx, ssize_t Ly, ssize_t Lz, ssize_t nb_ev, ssize_t dilE)
      : ctx_(ctx), Lt_(Lt), nev_(nb_ev) {
    if (ctx.query("Lt") != Lt || ctx.query("Lx") != Lx || ctx.query("Ly") != Ly || ctx.query("Lz") != Lz ||
        ctx.query("nev") != nb_ev || ctx.query("dilE") != dilE)
      throw std::runtime_error("OperatorFactory: sizes differ from the configured infile");
  }
  // create_operators(filename, rnd_vec, config): the eigenvectors come from
  // Context::read_inputs (or set_eigenvectors); config names the VdaggerV
  // files of handling_vdaggerv = write / read / liuming.  Displaced
  // operators take the links of Context::read_gauge_field (the file the
  // reference's build_vdaggerv reads itself) or set_gauge_links; the
  // infile's HYP smearing is applied here, as there.
  void create_operators(int config) {
    cache_.clear();
    ctx_.check(slc_create_operators_config(ctx_.get(), config));
  }
  // return_vdaggerv(index, t): valid until the next create_operators.
  Mat const &return_vdaggerv(ssize_t index, ssize_t t) const {
    auto key = std::make_pair(index, t);
    auto it = cache_.find(key);
    if (it != cache_.end()) return it->second;
    Mat m(nev_, nev_);
    ctx_.check(slc_return_vdaggerv(ctx_.get(), int(index), int(t), reinterpret_cast<double *>(m.data())));
    return cache_.emplace(key, std::move(m)).first->second;
  }
  Context &context() const { return ctx_; }

 private:
  Context &ctx_;
  ssize_t Lt_, nev_;
  mutable std::map<std::pair<ssize_t, ssize_t>, Mat> cache_;
};

struct DilutedFactor {  // DilutedFactor.hpp:19-25
  using Data = Mat;
  Data data;
  std::pair<RndId, RndId> ric;
  UsedRnd used_rnd_ids;
};

template <int n>
using DilutedFactorsMap = std::map<std::array<ssize_t, n>, std::vector<DilutedFactor>>;

// DilutedFactorFactory<Q> (DilutedFactorFactory.hpp:28-65): operator[](key)
// builds (on the device) and memoises the factors of every operator index of
// this quark-line type at the time key, keyed by {operator index}.
template <DilutedFactorType qlt>
class DilutedFactorFactory {
 public:
  using Key = std::array<int, DilutedFactorTypeTraits<qlt>::num_times>;
  using Value = DilutedFactorsMap<1>;

  explicit DilutedFactorFactory(OperatorFactory const &meson_operator) : ctx_(meson_operator.context()) {}

  Value const &operator[](Key const &key) {
    if (Ql.count(key) == 0) build(key);
    return Ql.at(key);
  }
  void clear() { Ql.clear(); }

  void build(Key const &time_key) {
    static const char *n_what[3] = {"n_Q0", "n_Q1", "n_Q2"};
    const int q = int(qlt);
    const long long nops = ctx_.query(n_what[q]);
    const ssize_t D = 4 * ctx_.query("dilE");
    Value v;
    const int cap = 256;
    std::vector<Complex> buf(size_t(cap) * D * D);
    std::vector<int> ric(2 * cap);
    for (long long op = 0; op < nops; ++op) {
      int n = 0;
      ctx_.check(slc_diluted_factor(ctx_.get(), q, time_key.data(), int(op), reinterpret_cast<double *>(buf.data()),
                                    ric.data(), cap, &n));
      std::vector<DilutedFactor> fs;
      for (int k = 0; k < n; ++k) {
        DilutedFactor f;
        f.data = Mat(D, D);
        std::copy(buf.begin() + size_t(k) * D * D, buf.begin() + size_t(k + 1) * D * D, f.data.values.begin());
        f.ric = {RndId(ric[2 * k]), RndId(ric[2 * k + 1])};
        f.used_rnd_ids = 0;  // a factor has used no random vector in between
        fs.push_back(std::move(f));
      }
      v[{ssize_t(op)}] = std::move(fs);
    }
    Ql[time_key] = std::move(v);
  }

 private:
  std::map<Key, Value> Ql;
  Context &ctx_;
};

struct DilutedTrace {  // DilutedTrace.hpp:16-19
  Complex data;
  UsedRnd used_rnd_ids;
};
struct DilutedTraces {  // DilutedTrace.hpp:21-24
  std::vector<DilutedTrace> traces;
  bool is_sum;
};
using DilutedTracesMap = std::map<ssize_t, std::vector<DilutedTrace>>;

// The slice pair of a BlockIterator (dilution-iterator.hpp:21-191): what
// get_time (DilutedTraceFactory.cpp:5-11) reads of it.
struct BlockIterator {
  int source_, sink_;
  int source() const { return source_; }
  int sink() const { return sink_; }
};

class AbstractDilutedTraceFactory {  // DilutedTraceFactory.hpp:21-29
 public:
  virtual ~AbstractDilutedTraceFactory() {}
  virtual DilutedTracesMap const &get(BlockIterator const &slice_pair, std::vector<Location> const &locations) = 0;
  virtual void clear() = 0;
};

// DilutedTrace{1,2,3,4,6}Factory by trace name ("trQ0Q2", "trQ2Q0Q2Q0", ...):
// get() computes (on the device) and memoises the trace vectors of every
// trace id at the key the slice pair and locations give.
class DilutedTraceFactory : public AbstractDilutedTraceFactory {
 public:
  DilutedTraceFactory(Context &ctx, std::string trace) : ctx_(ctx), trace_(std::move(trace)) {}

  DilutedTracesMap const &get(BlockIterator const &slice_pair, std::vector<Location> const &locations) override {
    std::vector<int> key, locs;
    for (auto l : locations) {
      locs.push_back(int(l));
      key.push_back(l == Location::source ? slice_pair.source() : slice_pair.sink());
    }
    auto it = cache_.find(key);
    if (it != cache_.end()) return it->second;
    int n_ids = 0, total = 0;
    std::vector<int> counts(64);
    std::vector<Complex> vals(4096);
    std::vector<uint16_t> used(4096);
    for (int attempt = 0; attempt < 2; ++attempt) {
      const int rc = slc_diluted_trace(ctx_.get(), trace_.c_str(), slice_pair.source(), slice_pair.sink(), locs.data(),
                                       int(locs.size()), counts.data(), int(counts.size()),
                                       reinterpret_cast<double *>(vals.data()), used.data(), int(vals.size()), &n_ids,
                                       &total);
      if (rc == SLC_OK) break;
      if (n_ids <= int(counts.size()) && total <= int(vals.size())) ctx_.check(rc);
      counts.resize(size_t(std::max<int>(n_ids, int(counts.size()))));
      vals.resize(size_t(std::max<int>(total, int(vals.size()))));
      used.resize(vals.size());
      if (attempt == 1) ctx_.check(rc);
    }
    DilutedTracesMap m;
    for (int id = 0, at = 0; id < n_ids; ++id) {
      auto &v = m[id];
      for (int k = 0; k < counts[id]; ++k, ++at) v.push_back({vals[at], used[at]});
    }
    return cache_.emplace(key, std::move(m)).first->second;
  }
  void clear() override { cache_.clear(); }

 private:
  Context &ctx_;
  std::string trace_;
  std::map<std::vector<int>, DilutedTracesMap> cache_;
};

// inner_product (DilutedTrace.cpp:5-82) of 2 or 3 trace vectors, and the
// mean of one (resolve_request1, Diagram.cpp:9-27), on the device; kahan
// selects the sums of the SLAPH_KAHAN build.
inline Complex inner_product(Context &ctx, std::vector<DilutedTraces const *> const &vecs, bool kahan = false) {
  std::vector<std::vector<Complex>> v(vecs.size());
  std::vector<std::vector<uint16_t>> u(vecs.size());
  std::vector<const double *> vp;
  std::vector<const uint16_t *> up;
  std::vector<int> n;
  for (size_t k = 0; k < vecs.size(); ++k) {
    for (auto const &t : vecs[k]->traces) {
      v[k].push_back(t.data);
      u[k].push_back(t.used_rnd_ids);
    }
    vp.push_back(reinterpret_cast<const double *>(v[k].data()));
    up.push_back(u[k].data());
    n.push_back(int(v[k].size()));
  }
  double out[2];
  ctx.check(slc_inner_product(ctx.get(), int(vecs.size()), vp.data(), up.data(), n.data(), kahan ? 1 : 0, out));
  return Complex(out[0], out[1]);
}
inline Complex inner_product(Context &ctx, DilutedTraces const &l, DilutedTraces const &r, bool kahan = false) {
  return inner_product(ctx, {&l, &r}, kahan);
}
inline Complex inner_product(Context &ctx, DilutedTraces const &l, DilutedTraces const &m, DilutedTraces const &r,
                             bool kahan = false) {
  return inner_product(ctx, {&l, &m, &r}, kahan);
}

// One diagram's results: (HDF5 dataset name, Lt values) per correlator request.
using Correlators = std::vector<std::pair<std::string, std::vector<Complex>>>;

// contract (Correlators.cpp:40-127) of every diagram of the infile on the
// device; returns per diagram what Diagram::write (Diagram.cpp:132-147)
// stores, correlator_[t][i] / Lt.
inline std::map<std::string, Correlators> contract(Context &ctx, int single_time_slice_combination = -1,
                                                   bool kahan = false) {
  ctx.check(slc_contract(ctx.get(), single_time_slice_combination, kahan ? 1 : 0));
  const long long Lt = ctx.query("Lt");
  int nd = 0;
  ctx.check(slc_diagram_count(ctx.get(), &nd));
  std::map<std::string, Correlators> out;
  for (int d = 0; d < nd; ++d) {
    char name[64];
    ctx.check(slc_diagram_name(ctx.get(), d, name, sizeof name));
    int nc = 0;
    ctx.check(slc_correlator_count(ctx.get(), name, &nc));
    std::vector<Complex> vals(size_t(nc) * Lt);
    ctx.check(slc_fetch_correlators(ctx.get(), name, reinterpret_cast<double *>(vals.data())));
    Correlators &c = out[name];
    for (int i = 0; i < nc; ++i) {
      const int len = slc_correlator_name(ctx.get(), name, i, nullptr, 0);
      ctx.check(len);
      std::string ds(size_t(len), '\0');
      ctx.check(slc_correlator_name(ctx.get(), name, i, &ds[0], len));
      ds.resize(size_t(len) - 1);
      c.emplace_back(ds, std::vector<Complex>(vals.begin() + size_t(i) * Lt, vals.begin() + size_t(i + 1) * Lt));
    }
  }
  return out;
}

}  // namespace slaph_amd

#endif  // SLAPH_AMD_HPP
