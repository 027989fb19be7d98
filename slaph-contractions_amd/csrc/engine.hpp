// The device-resident contraction engine behind the C ABI: one per context.
//
// It mirrors the reference's data flow (main/contract.cpp:28-107):
// configure (read_parameters + init_lookup_tables) -> inputs (perambulators,
// random vectors, eigenvectors) -> create_operators (VdaggerV) -> contract.
// Every floating-point stage runs on the GPU; the host only plans.
#pragma once

#include "host.hpp"
#include "kernels.hpp"

#include <algorithm>
#include <map>
#include <memory>
#include <string>
#include <vector>

namespace slaph {

// Owning device allocation.
class DevBuf {
 public:
  DevBuf() = default;
  ~DevBuf() { reset(); }
  DevBuf(DevBuf const &) = delete;
  DevBuf &operator=(DevBuf const &) = delete;
  DevBuf(DevBuf &&o) noexcept : p_(o.p_), bytes_(o.bytes_) {
    o.p_ = nullptr;
    o.bytes_ = 0;
  }
  DevBuf &operator=(DevBuf &&o) noexcept {
    if (this != &o) {
      reset();
      p_ = o.p_;
      bytes_ = o.bytes_;
      o.p_ = nullptr;
      o.bytes_ = 0;
    }
    return *this;
  }
  void reset();
  void ensure(size_t bytes);  // grow-only
  void alloc(size_t bytes);   // exact (re)allocation
  template <typename T>
  T *as() const {
    return static_cast<T *>(p_);
  }
  size_t bytes() const { return bytes_; }

 private:
  void *p_ = nullptr;
  size_t bytes_ = 0;
};

// Elapsed device time between construction and stop() on one stream.
class EventTimer {
 public:
  explicit EventTimer(hipStream_t s);
  ~EventTimer();
  double stop();

 private:
  hipStream_t s_;
  hipEvent_t a_, b_;
};

struct ContractCache;  // contract.cpp

struct StageTimes {
  double vdaggerv_ms = 0, factors_ms = 0, contract_ms = 0;
  double vdaggerv_kernel_ms = 0;  // the VdaggerV kernels only
  int vdaggerv_launches = 0;
};

// Launches of one kernel in the last create_operators / contract: summed
// device time (HIP events around each launch on the engine stream), the
// summed algorithmic bytes the launches must move (each operand matrix read
// once, every output written once) and the summed floating-point operations
// the kernel's own algorithm performs.
struct KernelStat {
  double ms = 0, bytes = 0, flops = 0;
  int launches = 0;
};

// Blocking upload of a host vector into a grow-only device buffer.
template <typename T>
void upload(DevBuf &buf, std::vector<T> const &v) {
  buf.ensure(std::max<size_t>(sizeof(T) * v.size(), 16));
  if (!v.empty()) SLAPH_HIP(hipMemcpy(buf.as<void>(), v.data(), sizeof(T) * v.size(), hipMemcpyHostToDevice));
}

class Engine {
 public:
  explicit Engine(int device);
  ~Engine();

  void configure(std::string const &infile_text);
  GlobalData const &gd() const { return gd_; }

  // Inputs (host buffers, reference in-memory layouts).
  void set_eigenvectors(int t, const cplx *V);     // K x N column-major, K = 3 Lx Ly Lz
  void set_perambulator(int r, const cplx *P);     // resorted column-major
  void set_random_vector(int r, const cplx *rho);  // Lt * N * 4, entry 4(N t + v) + d
  void read_inputs(std::string const &base_dir, int config);
  void synthetic_inputs(unsigned long long seed, bool orthonormalise);
  void get_input(int kind, int index, cplx *out);  // streaming: V(t) is produced anew

  // OperatorFactory::create_operators / return_vdaggerv; config names the
  // VdaggerV files of handling_vdaggerv = write / read / liuming.
  void create_operators(int config);
  int last_config() const { return config_; }

  // Timeslice streaming: only one timeslice of eigenvectors is resident;
  // create_operators produces V(t) timeslice by timeslice from the eigenvector
  // files (read_inputs) or the seeded generator (synthetic_inputs), as the
  // reference reads them per timeslice (OperatorFactory.cpp:263-277).  Set
  // before the inputs (before configure() too: it allocates for the mode);
  // set_eigenvectors is not available in this mode.
  void set_streaming(bool on);
  bool streaming() const { return streaming_; }

  // Gauge links for displaced operators (GaugeField, src/GaugeField.cpp):
  // one timeslice of Lx Ly Lz sites x 3 directions x 3 x 3 (column-major).
  void set_gauge_links(int t, const cplx *U);
  void get_gauge_links(int t, cplx *U) const;
  // GaugeField::read_gauge_field (GaugeField.cpp:762-779): timeslices
  // slice_i..slice_f of the lime file <config_path>/conf.%04d (config_path
  // "" = path_config of the input file, relative paths below base_dir)
  void read_gauge_field(std::string const &base_dir, std::string const &config_path, int config, int slice_i,
                        int slice_f);
  // GaugeField::plaque_ts (GaugeField.cpp:742-760) of the stored links
  double plaquette(int t);
  // seeded SU(3) links exp(i eps H), H traceless Hermitian (Gaussian)
  void synthetic_gauge(unsigned long long seed, double eps);
  // `iterations` stout steps of every timeslice (smearing_stout, GaugeField.cpp:228-274)
  void smear_stout(double rho, int iterations);
  // `iterations` HYP steps of every timeslice (smearing_hyp, GaugeField.cpp:322-431);
  // create_operators applies the input file's HYP parameters to a copy of
  // the links by itself, as the reference does
  void smear_hyp(double alpha1, double alpha2, int iterations);
  void return_vdaggerv(int id, int t, cplx *out);

  // DilutedFactorFactory<Q>::operator[](key) for one operator index: copies
  // the D x D matrices (in random-pair order) to out, pairs to ric.
  int diluted_factor(int qtype, const int *key, int op, cplx *out, int *ric, int cap);

  // contract (Correlators.cpp:40-127); results kept on the host.
  void contract(int single_block_pair, int kahan);
  // inner_product(DilutedTraces...) (DilutedTrace.cpp:5-82) of 2 or 3 trace
  // vectors, or the mean of one (resolve_request1, Diagram.cpp:9-27), by the
  // device resolve kernel; kahan != 0 sums like the SLAPH_KAHAN build.
  cplx inner_product(int ntr, const cplx *const *vals, const uint16_t *const *used, const int *n, int kahan);
  // AbstractDilutedTraceFactory::get (DilutedTraceFactory.hpp:21-29): for
  // trace factory `name` at slice pair (t_source, t_sink) with the trace's
  // locations, the trace vector of every trace id: (value, used mask) per
  // random-vector combination, computed on the device.
  std::vector<std::vector<std::pair<cplx, uint16_t>>> diluted_trace(std::string const &name, int t_source,
                                                                    int t_sink, std::vector<Loc> const &locs);
  std::vector<std::string> diagrams() const;
  int correlator_count(std::string const &diagram) const;
  std::string correlator_name(std::string const &diagram, int i) const;
  void fetch_correlators(std::string const &diagram, cplx *out) const;  // n_corr x Lt

  void set_chunk_budget(size_t bytes) {
    chunk_budget_ = bytes;
    plan_.reset();  // chunk tables depend on the budget
  }
  // Precision forms: vdaggerv 0 = FP64, 1 = single-pass FP32 MFMA, 2 =
  // three-term split BF16 (FP32 accumulation), 3 = Ozaki int8 slices (FP64
  // combination); factors 0 = FP64, 1 / 2 = factor slabs and product tables
  // stored in FP32 / in the 48-bit form (arithmetic and every accumulation in
  // FP64).
  void set_precision(int vdaggerv, int factors) {
    if (vdaggerv < 0 || vdaggerv > 3 || factors < 0 || factors > 2)
      throw std::out_of_range("precision: vdaggerv must be 0, 1, 2 or 3, factors 0, 1 or 2");
    prec_vdv_ = vdaggerv;
    if (factors != prec_factors_) plan_.reset();
    prec_factors_ = factors;
  }
  // Schedules and launch shapes (autotune() picks them by timing the
  // current workload; set_tuning / tuning() read and write them as text
  // "key=value ...").  Every choice computes the same correlators up to
  // rounding (the orders of some sums differ).
  struct Tuning {
    LaunchTuning launch;
    int vdv_direct = 0;      // VdaggerV by the direct product instead of the sitewise Fourier form
    int resolve_groups = 0;  // masked inner products by direct group loops instead of the Moebius form
    int xform_off = 0;       // crossed four-factor traces from one product per combination, not summed products
    int sym_off = 0;         // summed-product traces by the left-stationary kernel, not symmetric products
  };
  Tuning const &tuning() const { return tune_; }
  std::string tuning_text() const;
  void set_tuning(std::string const &text);
  // Times contract() (and create_operators() for the VdaggerV form) on the
  // current inputs for each candidate of each knob in turn, keeping a
  // candidate only if it is faster by more than 0.5 %; `repeats` timed calls
  // per candidate (the fastest counts).  The choice is kept in the context
  // for this input file (configure() with the same text reuses it).
  // Returns a JSON report of every measurement.
  std::string autotune(int repeats);
  // HIP graphs for contract(): on (default), the second call of a plan
  // captures the configuration's launch sequence and later calls replay it.
  void set_graphs(bool on) {
    graphs_ = on;
    plan_.reset();
  }
  bool graphs() const { return graphs_; }
  StageTimes const &times() const { return times_; }
  KernelStat kernel_stat(std::string const &name) const {
    auto it = kstats_.find(name);
    if (it == kstats_.end()) throw std::out_of_range("no statistics for kernel " + name);
    return it->second;
  }
  hipStream_t stream() const { return stream_; }

  // Raw entry: C = A^H diag(phase) B for device buffers (tests and tools).
  void ahb_raw(const double2 *A, const double2 *B, double2 *C, const double2 *phase, int M, int N, int K,
               long long lda, long long ldb, long long ldc, int batch, long long sA, long long sB, long long sC);
  // The same product with no device allocation (job table in the per-GPU
  // fixed table; at most kAhbNoallocJobs products per call).
  void ahb_noalloc(const double2 *A, const double2 *B, double2 *C, const double2 *phase, int M, int N, int K,
                   long long lda, long long ldb, long long ldc, int batch, long long sA, long long sB, long long sC);

 private:
  void require_configured() const;
  void require_inputs() const;
  void alloc_inputs();
  const double2 *vdv_ptr(int id, bool dagger, int t) const;
  void compute_q_offsets();
  void build_factors(std::vector<std::vector<int>> const &q1keys, std::vector<std::vector<int>> const &q2keys);
  // VdaggerV of timeslices t0 .. t0 + nt - 1 whose eigenvectors start at V
  // (consecutive timeslices), for the given undisplaced / displaced ids
  // (each returns its number of kernel launches)
  int vdaggerv_fourier(std::vector<int> const &ids, int t0, int nt, const double2 *V);
  int vdaggerv_direct(std::vector<int> const &ids, int t0, int nt, const double2 *V);
  int vdaggerv_displaced(std::vector<int> const &ids, int t, const double2 *V);
  // V(t) of the configured source into dst (K x N): the seeded generator or
  // the eigenvector file
  void load_eigenvectors(int t, double2 *dst);
  void synthetic_ev(int t, double2 *dst);
  void write_vdaggerv_files(int config);
  void read_vdaggerv_files(int config, bool liuming);
  struct ContractScope;  // contract.cpp
  // contract() on `plan` (the plan member, or a private plan of a scoped
  // call); with a scope, only its requests at its slice pair, and no results.
  void run_contract(std::shared_ptr<ContractCache> &plan, int single_block_pair, int kahan, ContractScope const *scope);

  int device_;
  hipStream_t stream_ = nullptr;
  AhBJob *ahb_table_ = nullptr;  // device address of this GPU's fixed A^H B job table
  GlobalData gd_;
  bool configured_ = false;
  bool have_ev_ = false, have_peram_ = false, have_rnd_ = false, have_ops_ = false;
  int config_ = 0;  // gauge configuration of the last read_inputs
  // eigenvector source: 0 = set by the caller (resident), 1 = files with
  // prefix ev_prefix_, 2 = seeded generator (ev_seed_, ev_ortho_)
  bool streaming_ = false;
  int ev_source_ = 0;
  std::string ev_prefix_;
  unsigned long long ev_seed_ = 0;
  bool ev_ortho_ = false;
  DevBuf U_;  // gauge links, Lt x sites x 3 x 9 (allocated when displaced operators exist)
  bool have_gauge_ = false;
  DevBuf Us_, hyp_dec_, hyp_tmp_;         // HYP-smeared copy of the links and its work space
  const double2 *disp_links_ = nullptr;   // links the displacements of this create_operators use
  void hyp_links(const double2 *in, double2 *out, double alpha1, double alpha2, int iterations);
  DevBuf vdisp_, vtmp_, gram_, rinv_, gjob_;  // displacement and orthonormalisation work space
  DevBuf colexp_;                             // column exponents of the Ozaki VdaggerV form
  DevBuf bigjob_, ahb_part_;                  // displaced products: job table, split-K partial tiles
  static constexpr int kDispBatch = 4;        // displacement paths (W = D V, K x N each) per launch
  long long K_ = 0;
  int N_ = 0, D_ = 0, Lt_ = 0, nb_ = 0, nrnd_ = 0;

  DevBuf V_, rho_, peram_, vdv_, vdvdag_, phase_;
  std::vector<int> dag_slot_;  // vdaggerv id -> slot in vdvdag_, or -1

  // factor slabs
  DevBuf q0_, q1_, q2_;
  std::vector<long long> q_off_[3];  // per Q type: entry -> first pair index
  long long q_total_[3] = {0, 0, 0};
  std::vector<int> q2slot_;          // (b1*Lt + t)*nb + b2 -> slot, -1 if absent
  std::vector<char> q1need_;         // t*nb + b

  struct DiagramResult {
    std::string name;
    std::vector<std::string> names;
    std::vector<cplx> values;  // n_corr x Lt
  };
  std::vector<DiagramResult> results_;

  size_t chunk_budget_ = 0;  // 0: from the free device memory
  int prec_vdv_ = 0, prec_factors_ = 0;
  bool graphs_ = true;
  Tuning tune_;
  std::map<std::string, Tuning> tuned_;  // input file text -> autotuned choice
  std::string infile_;                   // text of the last configure()
  DevBuf q0f_, q1f_, q2f_;  // FP32 copies of the factor slabs (factors precision 1)
  struct FactorJobs;        // build_factors' job tables (engine.cpp)
  std::shared_ptr<FactorJobs> fjobs_;
  DevBuf mbuf_;  // Q2 stage-1 products M' of one timeslice
  bool phases_ready_ = false;  // momentum phases uploaded for the current configuration
  struct FtPlan;  // tile pairs, passes and phase tables of the Fourier VdaggerV (engine.cpp)
  std::shared_ptr<FtPlan> ft_;
  DevBuf vr_, ftu_;  // row-major eigenvectors and per-x sums of one timeslice chunk
  StageTimes times_;
  std::map<std::string, KernelStat> kstats_;
  // contract()'s plan: everything derived from the input file alone (quantum
  // numbers, trace groups, per-chunk index tables); rebuilt after configure()
  // or a budget change, reused by every later gauge configuration.
  std::shared_ptr<ContractCache> plan_;
};

}  // namespace slaph
