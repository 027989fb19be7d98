// Device-resident contraction engine; see engine.hpp.
#include "engine.hpp"

#include "gauge.hpp"
#include "lime.hpp"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <map>
#include <numeric>
#include <set>
#include <tuple>
#include <unordered_map>

namespace slaph {

// ---------------------------------------------------------------------------
// DevBuf
// ---------------------------------------------------------------------------
void DevBuf::reset() {
  if (p_) (void)hipFree(p_);
  p_ = nullptr;
  bytes_ = 0;
}

void DevBuf::alloc(size_t bytes) {
  if (bytes == bytes_ && p_) return;  // same size: keep the allocation (and its address)
  reset();
  if (bytes == 0) return;
  SLAPH_HIP(hipMalloc(&p_, bytes));
  bytes_ = bytes;
}

void DevBuf::ensure(size_t bytes) {
  if (bytes <= bytes_) return;
  alloc(bytes);
}

EventTimer::EventTimer(hipStream_t s) : s_(s) {
  SLAPH_HIP(hipEventCreate(&a_));
  SLAPH_HIP(hipEventCreate(&b_));
  SLAPH_HIP(hipEventRecord(a_, s_));
}

EventTimer::~EventTimer() {
  (void)hipEventDestroy(a_);
  (void)hipEventDestroy(b_);
}

double EventTimer::stop() {
  SLAPH_HIP(hipEventRecord(b_, s_));
  SLAPH_HIP(hipEventSynchronize(b_));
  float ms = 0;
  SLAPH_HIP(hipEventElapsedTime(&ms, a_, b_));
  return ms;
}

// ---------------------------------------------------------------------------
// Engine: configuration and inputs
// ---------------------------------------------------------------------------
Engine::Engine(int device) : device_(device) {
  int n = 0;
  SLAPH_HIP(hipGetDeviceCount(&n));
  if (device < 0 || device >= n) throw std::runtime_error("invalid HIP device " + std::to_string(device));
  SLAPH_HIP(hipSetDevice(device));
  // device address of the allocation-free entry's job table, looked up once
  // per device with that device current
  ahb_table_ = ahb_noalloc_table(device);
  SLAPH_HIP(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking));
}

Engine::~Engine() {
  (void)hipSetDevice(device_);
  if (stream_) (void)hipStreamDestroy(stream_);
}

void Engine::configure(std::string const &text) {
  GlobalData gd;
  parse_infile(gd, text);
  init_lookup_tables(gd);
  for (auto const &q : gd.quarks)
    if (q.dilT != gd.quarks[0].dilT || q.dilE != gd.quarks[0].dilE || q.dilD != 4)
      throw std::runtime_error("all quarks must share the dilution scheme of quark 0 with full Dirac dilution");
  if (gd.Lt % gd.quarks[0].dilT != 0) throw std::runtime_error("Lt must be a multiple of the time dilution block");
  gd_ = std::move(gd);
  configured_ = true;
  infile_ = text;
  auto tuned = tuned_.find(text);
  tune_ = tuned == tuned_.end() ? Tuning{} : tuned->second;
  K_ = 3LL * gd_.Lx * gd_.Ly * gd_.Lz;
  N_ = gd_.nev;
  D_ = 4 * gd_.dilE();
  Lt_ = gd_.Lt;
  nb_ = gd_.n_blocks();
  nrnd_ = gd_.n_rnd_total();
  if (D_ > 32) throw std::runtime_error("4*dilE above 32 is not supported");
  have_ev_ = have_peram_ = have_rnd_ = have_ops_ = have_gauge_ = false;
  ev_source_ = 0;
  results_.clear();
  plan_.reset();
  fjobs_.reset();
  ft_.reset();
  phases_ready_ = false;
  SLAPH_HIP(hipSetDevice(device_));
  alloc_inputs();
}

void Engine::require_configured() const {
  if (!configured_) throw std::runtime_error("context is not configured");
}

// Perambulators and random vectors: what the factors and the contraction
// read.  Eigenvectors are needed only to build VdaggerV (create_operators
// checks for them), not when the operators are read from files.
void Engine::require_inputs() const {
  require_configured();
  if (!have_peram_ || !have_rnd_) throw std::runtime_error("inputs are not set");
}

void Engine::alloc_inputs() {
  // streaming: one timeslice of eigenvectors resident
  V_.reset();
  V_.alloc(sizeof(double2) * K_ * N_ * (streaming_ ? 1 : Lt_));
  rho_.alloc(sizeof(double2) * (size_t)nrnd_ * Lt_ * N_ * 4);
  peram_.alloc(sizeof(double2) * (size_t)nrnd_ * gd_.peram_rows() * gd_.peram_cols());
  if (gd_.need_gaugefield) U_.alloc(sizeof(double2) * 27 * (size_t)(K_ / 3) * Lt_);
}

void Engine::set_streaming(bool on) {
  if (on == streaming_) return;
  streaming_ = on;
  if (!configured_) return;  // configure() allocates for the mode set
  have_ev_ = have_ops_ = false;
  ev_source_ = 0;
  ft_.reset();
  SLAPH_HIP(hipSetDevice(device_));
  V_.reset();
  V_.alloc(sizeof(double2) * K_ * N_ * (streaming_ ? 1 : Lt_));
}

void Engine::set_gauge_links(int t, const cplx *U) {
  require_configured();
  if (!gd_.need_gaugefield) throw std::runtime_error("no displaced operator needs gauge links");
  if (t < 0 || t >= Lt_) throw std::runtime_error("timeslice out of range");
  const size_t n = 27 * (size_t)(K_ / 3);
  SLAPH_HIP(hipSetDevice(device_));
  SLAPH_HIP(hipMemcpy(U_.as<double2>() + (size_t)t * n, U, sizeof(double2) * n, hipMemcpyHostToDevice));
  have_gauge_ = true;
}

void Engine::get_gauge_links(int t, cplx *U) const {
  require_configured();
  if (!have_gauge_) throw std::runtime_error("gauge links are not set");
  if (t < 0 || t >= Lt_) throw std::runtime_error("timeslice out of range");
  const size_t n = 27 * (size_t)(K_ / 3);
  SLAPH_HIP(hipSetDevice(device_));
  SLAPH_HIP(hipMemcpy(U, U_.as<double2>() + (size_t)t * n, sizeof(double2) * n, hipMemcpyDeviceToHost));
}

void Engine::read_gauge_field(std::string const &base, std::string const &config_path, int config, int slice_i,
                              int slice_f) {
  require_configured();
  if (!gd_.need_gaugefield) throw std::runtime_error("no displaced operator needs gauge links");
  std::string dir = config_path.empty() ? gd_.path_config : config_path;
  if (!dir.empty() && dir[0] != '/' && !base.empty() && base != ".") dir = base + "/" + dir;
  char name[32];
  std::snprintf(name, sizeof name, "/conf.%04d", config);
  const size_t n = 27 * (size_t)(K_ / 3);
  SLAPH_HIP(hipSetDevice(device_));
  have_ops_ = false;
  try {
    read_ildg_gauge(dir + name, gd_.Lx, gd_.Ly, gd_.Lz, Lt_, slice_i, slice_f, [&](int t, const cplx *U) {
      SLAPH_HIP(hipMemcpy(U_.as<double2>() + (size_t)t * n, U, sizeof(double2) * n, hipMemcpyHostToDevice));
    });
  } catch (...) {
    have_gauge_ = false;  // links of a file that failed its checks are not used
    throw;
  }
  have_gauge_ = true;
}

double Engine::plaquette(int t) {
  require_configured();
  if (!have_gauge_) throw std::runtime_error("gauge links are not set");
  if (t < 0 || t >= Lt_) throw std::runtime_error("timeslice out of range");
  const int nb = plaquette_blocks(gd_.Lx, gd_.Ly, gd_.Lz);
  SLAPH_HIP(hipSetDevice(device_));
  DevBuf part;
  part.alloc(sizeof(double) * nb);
  launch_plaquette(U_.as<double2>() + (size_t)t * 27 * (K_ / 3), gd_.Lx, gd_.Ly, gd_.Lz, part.as<double>(), stream_);
  std::vector<double> h(nb);
  SLAPH_HIP(hipMemcpyAsync(h.data(), part.as<void>(), sizeof(double) * nb, hipMemcpyDeviceToHost, stream_));
  SLAPH_HIP(hipStreamSynchronize(stream_));
  double s = 0.0;
  for (double v : h) s += v;
  return s / (6.0 * (K_ / 3));
}

void Engine::synthetic_gauge(unsigned long long seed, double eps) {
  require_configured();
  if (!gd_.need_gaugefield) throw std::runtime_error("no displaced operator needs gauge links");
  SLAPH_HIP(hipSetDevice(device_));
  launch_su3_random(U_.as<double2>(), 3LL * (K_ / 3) * Lt_, seed, eps, stream_);
  SLAPH_HIP(hipStreamSynchronize(stream_));
  have_gauge_ = true;
  have_ops_ = false;
}

void Engine::smear_stout(double rho, int iterations) {
  require_configured();
  if (!have_gauge_) throw std::runtime_error("gauge links are not set");
  if (iterations < 0) throw std::out_of_range("stout smearing: negative iteration count");
  SLAPH_HIP(hipSetDevice(device_));
  const size_t n = 27 * (size_t)(K_ / 3);
  DevBuf tmp;
  tmp.alloc(sizeof(double2) * n);
  for (int t = 0; t < Lt_; ++t) {
    double2 *u = U_.as<double2>() + (size_t)t * n;
    for (int i = 0; i < iterations; ++i) {
      launch_stout_step(u, tmp.as<double2>(), gd_.Lx, gd_.Ly, gd_.Lz, rho, stream_);
      SLAPH_HIP(hipMemcpyAsync(u, tmp.as<void>(), sizeof(double2) * n, hipMemcpyDeviceToDevice, stream_));
    }
  }
  SLAPH_HIP(hipStreamSynchronize(stream_));
  have_ops_ = false;
}

// HYP steps of every timeslice of `in` into `out` (in == out allowed): per
// step the decorated links of a timeslice into hyp_dec_, then the outer level
// into hyp_tmp_ and back.
void Engine::hyp_links(const double2 *in, double2 *out, double alpha1, double alpha2, int iterations) {
  const size_t n = 27 * (size_t)(K_ / 3);
  hyp_dec_.ensure(sizeof(double2) * 2 * n);
  hyp_tmp_.ensure(sizeof(double2) * n);
  for (int t = 0; t < Lt_; ++t) {
    const double2 *src = in + (size_t)t * n;
    double2 *dst = out + (size_t)t * n;
    if (iterations == 0 && src != dst)
      SLAPH_HIP(hipMemcpyAsync(dst, src, sizeof(double2) * n, hipMemcpyDeviceToDevice, stream_));
    for (int i = 0; i < iterations; ++i) {
      launch_hyp_decorate(i == 0 ? src : dst, hyp_dec_.as<double2>(), gd_.Lx, gd_.Ly, gd_.Lz, alpha2, stream_);
      launch_hyp_outer(i == 0 ? src : dst, hyp_dec_.as<double2>(), hyp_tmp_.as<double2>(), gd_.Lx, gd_.Ly, gd_.Lz,
                       alpha1, stream_);
      SLAPH_HIP(hipMemcpyAsync(dst, hyp_tmp_.as<void>(), sizeof(double2) * n, hipMemcpyDeviceToDevice, stream_));
    }
  }
  SLAPH_HIP(hipStreamSynchronize(stream_));
}

void Engine::smear_hyp(double alpha1, double alpha2, int iterations) {
  require_configured();
  if (!have_gauge_) throw std::runtime_error("gauge links are not set");
  if (iterations < 0) throw std::out_of_range("HYP smearing: negative iteration count");
  SLAPH_HIP(hipSetDevice(device_));
  hyp_links(U_.as<double2>(), U_.as<double2>(), alpha1, alpha2, iterations);
  have_ops_ = false;
}

void Engine::set_eigenvectors(int t, const cplx *V) {
  require_configured();
  if (streaming_)
    throw std::runtime_error("timeslice streaming: eigenvectors come from the files or the seeded generator");
  if (t < 0 || t >= Lt_) throw std::runtime_error("timeslice out of range");
  SLAPH_HIP(hipSetDevice(device_));
  SLAPH_HIP(hipMemcpy(V_.as<double2>() + (size_t)t * K_ * N_, V, sizeof(double2) * K_ * N_, hipMemcpyHostToDevice));
  have_ev_ = true;
}

void Engine::set_perambulator(int r, const cplx *P) {
  require_configured();
  if (r < 0 || r >= nrnd_) throw std::runtime_error("perambulator index out of range");
  size_t n = gd_.peram_rows() * gd_.peram_cols();
  SLAPH_HIP(hipSetDevice(device_));
  SLAPH_HIP(hipMemcpy(peram_.as<double2>() + r * n, P, sizeof(double2) * n, hipMemcpyHostToDevice));
  have_peram_ = true;
}

void Engine::set_random_vector(int r, const cplx *rho) {
  require_configured();
  if (r < 0 || r >= nrnd_) throw std::runtime_error("random vector index out of range");
  size_t n = (size_t)Lt_ * N_ * 4;
  SLAPH_HIP(hipSetDevice(device_));
  SLAPH_HIP(hipMemcpy(rho_.as<double2>() + r * n, rho, sizeof(double2) * n, hipMemcpyHostToDevice));
  have_rnd_ = true;
}

void Engine::read_inputs(std::string const &base, int config) {
  // Perambulator::read_perambulators_from_separate_files (Perambulator.cpp:82-97),
  // RandomVector::read_random_vectors_from_separate_files (RandomVector.cpp:104-113),
  // EigenVector::read_eigen_vector (EigenVector.cpp:35-76).
  require_configured();
  int Vs = gd_.Lx * gd_.Ly * gd_.Lz;
  long long prow = gd_.peram_rows(), pcol = gd_.peram_cols();
  std::vector<cplx> file(prow * pcol), sorted(prow * pcol), rnd((size_t)Lt_ * N_ * 4);
  int r = 0;
  for (auto const &q0 : gd_.quarks) {
    Quark q = q0;
    if (!q.path.empty() && q.path[0] != '/') q.path = base + "/" + q.path;
    for (int i = 0; i < q.nrnd; ++i, ++r) {
      read_binary(perambulator_filename(q, config, i, Lt_, Vs), file.data(), sizeof(cplx) * file.size());
      resort_perambulator(file.data(), sorted.data(), Lt_, N_, q.dilT, q.dilE, q.dilD);
      set_perambulator(r, sorted.data());
      read_binary(randomvector_filename(q, config, i, N_), rnd.data(), sizeof(cplx) * rnd.size());
      set_random_vector(r, rnd.data());
    }
  }
  config_ = config;
  // the reference reads eigenvectors only to build VdaggerV (build_vdaggerv,
  // OperatorFactory.cpp:263-277), not when the operators come from files
  if (gd_.handling_vdaggerv == "read" || gd_.handling_vdaggerv == "liuming") return;
  GlobalData g2 = gd_;
  if (!g2.path_eigenvectors.empty() && g2.path_eigenvectors[0] != '/') g2.path_eigenvectors = base + "/" + g2.path_eigenvectors;
  ev_prefix_ = eigenvector_prefix(g2, config);
  ev_source_ = 1;
  have_ev_ = true;
  if (streaming_) return;  // read timeslice by timeslice in create_operators
  for (int t = 0; t < Lt_; ++t) load_eigenvectors(t, V_.as<double2>() + (size_t)t * K_ * N_);
  SLAPH_HIP(hipStreamSynchronize(stream_));
}

void Engine::load_eigenvectors(int t, double2 *dst) {
  if (ev_source_ == 2) {
    synthetic_ev(t, dst);
    return;
  }
  if (ev_source_ != 1) throw std::runtime_error("inputs are not set");
  // EigenVector::read_eigen_vector (EigenVector.cpp:35-76): the file holds N
  // vectors of K complex numbers, i.e. the column-major K x N matrix
  char suffix[16];
  std::snprintf(suffix, sizeof suffix, "%03d", t);
  std::vector<cplx> V((size_t)K_ * N_);
  read_binary(ev_prefix_ + suffix, V.data(), sizeof(cplx) * V.size());
  SLAPH_HIP(hipMemcpy(dst, V.data(), sizeof(double2) * V.size(), hipMemcpyHostToDevice));
}

namespace {

// Cholesky G = L L^H of a Hermitian positive-definite N x N (column-major);
// returns R^{-1} with R = L^H upper triangular, so that V R^{-1} has
// orthonormal columns when G = V^H V.
std::vector<cplx> chol_inverse_upper(std::vector<cplx> const &G, int N) {
  std::vector<cplx> L(N * N, 0.0);
  for (int j = 0; j < N; ++j) {
    double d = G[j * N + j].real();
    for (int k = 0; k < j; ++k) d -= std::norm(L[k * N + j]);
    if (!(d > 0)) throw std::runtime_error("orthonormalisation: Gram matrix is not positive definite");
    double ljj = std::sqrt(d);
    L[j * N + j] = ljj;
    for (int i = j + 1; i < N; ++i) {
      cplx s = G[j * N + i];
      for (int k = 0; k < j; ++k) s -= L[k * N + i] * std::conj(L[k * N + j]);
      L[j * N + i] = s / ljj;
    }
  }
  // R = L^H (upper). Invert R column by column: R X = I.
  std::vector<cplx> X(N * N, 0.0);
  for (int c = 0; c < N; ++c) {
    for (int i = c; i >= 0; --i) {
      cplx s = (i == c) ? 1.0 : 0.0;
      for (int k = i + 1; k <= c; ++k) s -= std::conj(L[i * N + k]) * X[c * N + k];  // R[i,k] = conj(L[k,i])
      X[c * N + i] = s / std::conj(L[i * N + i]);
    }
  }
  return X;
}

unsigned long long seed_of(unsigned long long seed, unsigned long long kind, unsigned long long idx) {
  unsigned long long z = seed * 0x9E3779B97F4A7C15ULL + kind * 0xD1B54A32D192ED03ULL + idx * 0x8CB92BA72F3D8DD7ULL;
  z = (z ^ (z >> 33)) * 0xff51afd7ed558ccdULL;
  return z ^ (z >> 29);
}

}  // namespace

void Engine::synthetic_inputs(unsigned long long seed, bool orthonormalise) {
  // Seeded complex Gaussian eigenvectors (optionally orthonormalised by two
  // passes of Cholesky-QR), complex Gaussian perambulators and Z2xZ2 random
  // vectors with entries (+-1 +-i)/sqrt(2); everything generated on the device.
  require_configured();
  SLAPH_HIP(hipSetDevice(device_));
  size_t np = gd_.peram_rows() * gd_.peram_cols();
  for (int r = 0; r < nrnd_; ++r) {
    launch_fill_gauss(peram_.as<double2>() + r * np, np, seed_of(seed, 1, r), stream_);
    launch_fill_z2z2(rho_.as<double2>() + (size_t)r * Lt_ * N_ * 4, (long long)Lt_ * N_ * 4, seed_of(seed, 2, r),
                     stream_);
  }
  ev_seed_ = seed;
  ev_ortho_ = orthonormalise;
  ev_source_ = 2;
  if (!streaming_)  // streaming: generated timeslice by timeslice in create_operators
    for (int t = 0; t < Lt_; ++t) synthetic_ev(t, V_.as<double2>() + (size_t)t * K_ * N_);
  SLAPH_HIP(hipStreamSynchronize(stream_));
  have_ev_ = have_peram_ = have_rnd_ = true;
}

// V(t) of the seeded generator: complex Gaussian (stream seed_of(seed, 3, t)),
// orthonormalised when asked by two passes of Cholesky-QR (the Gram matrix on
// the device, its Cholesky factor on the host, V <- V R^{-1} on the device).
void Engine::synthetic_ev(int t, double2 *dst) {
  launch_fill_gauss(dst, K_ * N_, seed_of(ev_seed_, 3, t), stream_);
  if (!ev_ortho_) return;
  vtmp_.ensure(sizeof(double2) * K_ * N_);
  gram_.ensure(sizeof(double2) * N_ * N_);
  rinv_.ensure(sizeof(double2) * N_ * N_);
  VdvJob job{};
  job.V = dst;
  job.np = 1;
  job.C[0] = gram_.as<double2>();
  upload(gjob_, std::vector<VdvJob>{job});
  std::vector<cplx> G((size_t)N_ * N_);
  for (int pass = 0; pass < 2; ++pass) {
    launch_vdv(gjob_.as<VdvJob>(), 1, N_, K_, 0, stream_);
    SLAPH_HIP(hipMemcpyAsync(G.data(), gram_.as<void>(), sizeof(cplx) * G.size(), hipMemcpyDeviceToHost, stream_));
    SLAPH_HIP(hipStreamSynchronize(stream_));
    auto Rinv = chol_inverse_upper(G, N_);
    SLAPH_HIP(hipMemcpy(rinv_.as<void>(), Rinv.data(), sizeof(cplx) * Rinv.size(), hipMemcpyHostToDevice));
    launch_gemm_nn_small(dst, rinv_.as<double2>(), vtmp_.as<double2>(), K_, N_, stream_);
    SLAPH_HIP(hipMemcpyAsync(dst, vtmp_.as<void>(), sizeof(double2) * K_ * N_, hipMemcpyDeviceToDevice, stream_));
    SLAPH_HIP(hipStreamSynchronize(stream_));
  }
}

void Engine::get_input(int kind, int index, cplx *out) {
  require_inputs();
  if (kind == 0 && !have_ev_) throw std::runtime_error("inputs are not set");
  const double2 *src = nullptr;
  size_t n = 0;
  DevBuf one;
  SLAPH_HIP(hipSetDevice(device_));
  if (kind == 0 && index >= 0 && index < Lt_ && streaming_) {
    n = (size_t)K_ * N_;
    one.alloc(sizeof(double2) * n);
    load_eigenvectors(index, one.as<double2>());
    SLAPH_HIP(hipStreamSynchronize(stream_));
    src = one.as<double2>();
  } else if (kind == 0 && index >= 0 && index < Lt_) {
    n = (size_t)K_ * N_;
    src = V_.as<double2>() + index * n;
  } else if (kind == 1 && index >= 0 && index < nrnd_) {
    n = gd_.peram_rows() * gd_.peram_cols();
    src = peram_.as<double2>() + index * n;
  } else if (kind == 2 && index >= 0 && index < nrnd_) {
    n = (size_t)Lt_ * N_ * 4;
    src = rho_.as<double2>() + index * n;
  } else {
    throw std::out_of_range("get_input: bad kind or index");
  }
  SLAPH_HIP(hipSetDevice(device_));
  SLAPH_HIP(hipMemcpy(out, src, sizeof(double2) * n, hipMemcpyDeviceToHost));
}

void Engine::ahb_raw(const double2 *A, const double2 *B, double2 *C, const double2 *phase, int M, int N, int K,
                     long long lda, long long ldb, long long ldc, int batch, long long sA, long long sB, long long sC) {
  std::vector<AhBJob> jobs(batch);
  for (int i = 0; i < batch; ++i) jobs[i] = {A + i * sA, B + i * sB, C + i * sC, phase, nullptr, lda, ldb, ldc};
  DevBuf dj;
  upload(dj, jobs);
  launch_ahb(dj.as<AhBJob>(), batch, M, N, K, stream_);
  SLAPH_HIP(hipStreamSynchronize(stream_));
}

void Engine::ahb_noalloc(const double2 *A, const double2 *B, double2 *C, const double2 *phase, int M, int N, int K,
                         long long lda, long long ldb, long long ldc, int batch, long long sA, long long sB,
                         long long sC) {
  // No device allocation: the job table is built on the host stack and copied
  // into the per-GPU fixed table, which the kernel reads; the call returns
  // after the kernel has finished, so the table is free for the next call.
  if (batch < 1 || batch > kAhbNoallocJobs)
    throw std::out_of_range("slc_ahb_device_noalloc: batch must be between 1 and " +
                            std::to_string(kAhbNoallocJobs) + " (size of the fixed job table)");
  AhBJob jobs[kAhbNoallocJobs];
  for (int i = 0; i < batch; ++i) jobs[i] = {A + i * sA, B + i * sB, C + i * sC, phase, nullptr, lda, ldb, ldc};
  SLAPH_HIP(hipSetDevice(device_));
  SLAPH_HIP(hipMemcpyAsync(ahb_table_, jobs, sizeof(AhBJob) * batch, hipMemcpyHostToDevice, stream_));
  launch_ahb(ahb_table_, batch, M, N, K, stream_);
  SLAPH_HIP(hipStreamSynchronize(stream_));
}

// ---------------------------------------------------------------------------
// OperatorFactory::create_operators / build_vdaggerv (OperatorFactory.cpp:220-326)
// ---------------------------------------------------------------------------
void Engine::create_operators(int config) {
  // handling_vdaggerv (OperatorFactory::create_operators, OperatorFactory.cpp:507-520):
  // "build" / "write" form VdaggerV from the eigenvectors ("write" also stores
  // every non-trivial operator), "read" / "liuming" load stored operators.
  require_configured();
  const std::string &mode = gd_.handling_vdaggerv;
  if (mode != "build" && mode != "write" && mode != "read" && mode != "liuming")
    throw std::runtime_error("The flag handling_vdaggerv in input file is wrong!");
  SLAPH_HIP(hipSetDevice(device_));
  int nvdv = int(gd_.vdaggerv_lookup.size());
  int Vs = gd_.Lx * gd_.Ly * gd_.Lz;
  vdv_.alloc(sizeof(double2) * std::max(1, nvdv) * Lt_ * N_ * N_);
  if (mode == "read" || mode == "liuming") {
    read_vdaggerv_files(config, mode == "liuming");
  } else {
  if (!have_ev_) throw std::runtime_error("inputs are not set");
  if (!phases_ready_) {
    // Momentum phases exp(-i 2 pi p.x / L), site = (x Ly + y) Lz + z
    // (create_momenta, OperatorFactory.cpp:64-95); they depend on the input
    // file only and are uploaded once per configure.
    phase_.alloc(sizeof(double2) * std::max(1, nvdv) * Vs);
    std::vector<cplx> ph((size_t)std::max(1, nvdv) * Vs);
    for (auto const &op : gd_.vdaggerv_lookup) {
      double ipx = op.mom[0] * 2.0 * M_PI / gd_.Lx, ipy = op.mom[1] * 2.0 * M_PI / gd_.Ly,
             ipz = op.mom[2] * 2.0 * M_PI / gd_.Lz;
      for (int x = 0; x < gd_.Lx; ++x)
        for (int y = 0; y < gd_.Ly; ++y)
          for (int z = 0; z < gd_.Lz; ++z)
            ph[(size_t)op.id * Vs + (x * gd_.Ly + y) * gd_.Lz + z] =
                std::exp(cplx(0, -1) * (ipx * x + ipy * y + ipz * z));
    }
    SLAPH_HIP(hipMemcpy(phase_.as<void>(), ph.data(), sizeof(cplx) * ph.size(), hipMemcpyHostToDevice));
    phases_ready_ = true;
  }
  // Undisplaced non-trivial momenta by the Fourier (or direct) form;
  // displaced operators as V^+ (D V) without momentum phase, as the reference
  // builds them (OperatorFactory.cpp:287-300).
  std::vector<int> ids, disp_ids;
  for (auto const &op : gd_.vdaggerv_lookup) {
    if (op.id == gd_.index_of_unity)
      launch_identity(vdv_.as<double2>() + (size_t)op.id * Lt_ * N_ * N_, N_, Lt_, stream_);
    else if (!op.disp.empty())
      disp_ids.push_back(op.id);
    else
      ids.push_back(op.id);
  }
  if (!disp_ids.empty() && !have_gauge_) throw std::runtime_error("gauge links are not set (displaced operators)");
  // HYP smearing of the input file, applied to a copy of the links for this
  // call, as the reference smears the freshly read configuration in
  // create_operators (OperatorFactory.cpp:249-260)
  disp_links_ = U_.as<double2>();
  if (!disp_ids.empty() && gd_.hyp_iterations > 0) {
    Us_.ensure(U_.bytes());
    hyp_links(U_.as<double2>(), Us_.as<double2>(), gd_.hyp_alpha1, gd_.hyp_alpha2, gd_.hyp_iterations);
    disp_links_ = Us_.as<double2>();
  }
  const char *direct = std::getenv("SLAPH_VDV_DIRECT");
  const bool fourier = prec_vdv_ == 0 && !ids.empty() && gd_.Lz <= kFtMaxLz && !tune_.vdv_direct &&
                       !(direct && *direct && *direct != '0');
  EventTimer tm(stream_);
  int launches = 0;
  kstats_.erase("k_vdv_ft");
  auto plain = [&](int t0, int nt, const double2 *V) {
    if (ids.empty()) return;
    launches += fourier ? vdaggerv_fourier(ids, t0, nt, V) : vdaggerv_direct(ids, t0, nt, V);
  };
  if (!streaming_) {
    plain(0, Lt_, V_.as<double2>());
    if (!disp_ids.empty())
      for (int t = 0; t < Lt_; ++t) launches += vdaggerv_displaced(disp_ids, t, V_.as<double2>() + (size_t)t * K_ * N_);
  } else {
    // one timeslice of eigenvectors resident at a time
    for (int t = 0; t < Lt_; ++t) {
      load_eigenvectors(t, V_.as<double2>());
      plain(t, 1, V_.as<double2>());
      if (!disp_ids.empty()) launches += vdaggerv_displaced(disp_ids, t, V_.as<double2>());
    }
  }
  times_.vdaggerv_kernel_ms = tm.stop();
  times_.vdaggerv_launches = launches;
  times_.vdaggerv_ms = times_.vdaggerv_kernel_ms;
  if (mode == "write") write_vdaggerv_files(config);
  }
  // Daggered copies of the operators that some factor uses with -p
  // (build_VdaggerV_lookup's dagger flag, init_lookup_tables.cpp:257-259).
  dag_slot_.assign(std::max(1, nvdv), -1);
  int ndag = 0;
  for (auto const &kv : gd_.quarkline_lookuptable)
    for (auto const &e : kv.second)
      if (e.dagger && dag_slot_[e.id_vdaggerv] < 0) dag_slot_[e.id_vdaggerv] = ndag++;
  vdvdag_.alloc(sizeof(double2) * std::max(1, ndag) * Lt_ * N_ * N_);
  for (int id = 0; id < nvdv; ++id)
    if (dag_slot_[id] >= 0)
      launch_dagger(vdv_.as<double2>() + (size_t)id * Lt_ * N_ * N_,
                    vdvdag_.as<double2>() + (size_t)dag_slot_[id] * Lt_ * N_ * N_, N_, Lt_, stream_);
  SLAPH_HIP(hipStreamSynchronize(stream_));
  have_ops_ = true;
}

// Direct form: one V(t) panel pass per group of up to kVdvGroup momenta
// (FP64 or the single-pass FP32 precision form), or one pass per momentum
// (the split-BF16 and Ozaki precision forms).
int Engine::vdaggerv_direct(std::vector<int> const &ids, int t0, int nt, const double2 *V) {
  int Vs = gd_.Lx * gd_.Ly * gd_.Lz;
  std::vector<VdvJob> jobs;
  if (prec_vdv_ >= 2) {
    if (prec_vdv_ == 3) {  // column exponents of every timeslice
      colexp_.ensure(sizeof(int) * (size_t)nt * N_);
      for (int t = 0; t < nt; ++t) launch_colexp(V + (size_t)t * K_ * N_, K_, N_, colexp_.as<int>() + (size_t)t * N_, stream_);
    }
    for (int t = t0; t < t0 + nt; ++t)
      for (int id : ids) {
        VdvJob j{};
        j.V = V + (size_t)(t - t0) * K_ * N_;
        j.np = 1;
        j.phase[0] = phase_.as<double2>() + (size_t)id * Vs;
        j.C[0] = vdv_.as<double2>() + ((size_t)id * Lt_ + t) * N_ * N_;
        j.ex[0] = j.ex[1] = colexp_.as<int>() + (size_t)(t - t0) * N_;
        jobs.push_back(j);
      }
    DevBuf dj;
    upload(dj, jobs);
    for (size_t i = 0; i < jobs.size(); i += 65535)
      launch_vdv_split(dj.as<VdvJob>() + i, int(std::min<size_t>(65535, jobs.size() - i)), N_, K_, prec_vdv_, stream_);
    SLAPH_HIP(hipStreamSynchronize(stream_));  // before the job table is freed
    return int((jobs.size() + 65534) / 65535) + (prec_vdv_ == 3 ? nt : 0);
  }
  int ngroups = int((ids.size() + kVdvGroup - 1) / kVdvGroup);
  for (int t = t0; t < t0 + nt; ++t) {
    size_t at = 0;
    for (int g = 0; g < ngroups; ++g) {
      size_t len = ids.size() / ngroups + (size_t(g) < ids.size() % ngroups ? 1 : 0);
      VdvJob j{};
      j.V = V + (size_t)(t - t0) * K_ * N_;
      j.np = int(len);
      for (size_t k = 0; k < len; ++k, ++at) {
        int id = ids[at];
        j.phase[k] = phase_.as<double2>() + (size_t)id * Vs;
        j.C[k] = vdv_.as<double2>() + ((size_t)id * Lt_ + t) * N_ * N_;
      }
      jobs.push_back(j);
    }
  }
  DevBuf dj;
  upload(dj, jobs);
  for (size_t i = 0; i < jobs.size(); i += 65535)
    launch_vdv(dj.as<VdvJob>() + i, int(std::min<size_t>(65535, jobs.size() - i)), N_, K_, prec_vdv_, stream_);
  SLAPH_HIP(hipStreamSynchronize(stream_));  // before the job table is freed
  return int((jobs.size() + 65534) / 65535);
}

// Displaced operators at timeslice t: W = D V by the one-link displacements
// of the operator in order (each '>' forward / '<' backward along x, y or
// z), then VdaggerV = V^+ W on MFMA (OperatorFactory.cpp:287-300), the
// distinct displacement paths of a batch in one launch: in FP64 the split-K
// tall product (64 x 64 tiles), in the precision forms their VdaggerV
// kernels (A panel V, B panel the path's W).  The displacement is the
// covariant one, U_mu(x) V(x + mu) / U_mu(x - mu)^+ V(x - mu), computed from
// the unchanged input of each step (the reference updates its result in
// place while reading it, displace_eigenvectors, OperatorFactory.cpp:33-55,
// so that some sites read already-displaced neighbours).  The momentum does
// not enter a displaced operator, so operators with the same displacement
// path are equal: each path is computed once and copied.
int Engine::vdaggerv_displaced(std::vector<int> const &ids, int t, const double2 *V) {
  const size_t kn = (size_t)K_ * N_;
  const double2 *U = disp_links_ + (size_t)t * 27 * (K_ / 3);
  std::vector<int> paths;  // first id of each distinct path
  std::map<Displacement, int> first;
  for (int id : ids)
    if (first.emplace(gd_.vdaggerv_lookup.at(id).disp, id).second) paths.push_back(id);
  const int batch = std::min<int>(int(paths.size()), kDispBatch);
  vdisp_.ensure(sizeof(double2) * kn * batch);
  vtmp_.ensure(sizeof(double2) * kn);
  int launches = 0;
  auto out_of = [&](int id) { return vdv_.as<double2>() + ((size_t)id * Lt_ + t) * N_ * N_; };
  for (size_t b0 = 0; b0 < paths.size(); b0 += batch) {
    std::vector<VdvJob> jobs;
    for (size_t b = b0; b < std::min(paths.size(), b0 + batch); ++b) {
      auto const &disp = gd_.vdaggerv_lookup.at(paths[b]).disp;
      double2 *W = vdisp_.as<double2>() + (b - b0) * kn;
      const double2 *cur = V;
      const size_t steps = disp.size();
      for (size_t k = 0; k < steps; ++k) {
        const char axis = disp[k].second;
        const int mu = axis == 'x' ? 0 : (axis == 'y' ? 1 : (axis == 'z' ? 2 : -1));
        if (mu < 0 || (disp[k].first != '>' && disp[k].first != '<'))
          throw std::runtime_error("Something wrong with the displacement in the operator definition!");
        // alternate so that the last step lands in W
        double2 *out = ((steps - 1 - k) % 2 == 0) ? W : vtmp_.as<double2>();
        launch_displace(U, cur, out, gd_.Lx, gd_.Ly, gd_.Lz, N_, mu, disp[k].first == '>' ? 1 : 0, stream_);
        cur = out;
        ++launches;
      }
      VdvJob j{};
      j.V = V;
      j.W = W;
      j.np = 1;
      j.C[0] = out_of(paths[b]);
      if (prec_vdv_ == 3) {  // column exponents of V (slot 0) and of each W
        colexp_.ensure(sizeof(int) * (size_t)(1 + batch) * N_);
        if (b == b0) launch_colexp(V, K_, N_, colexp_.as<int>(), stream_);
        int *e = colexp_.as<int>() + (size_t)(1 + b - b0) * N_;
        launch_colexp(W, K_, N_, e, stream_);
        j.ex[0] = colexp_.as<int>();
        j.ex[1] = e;
      }
      jobs.push_back(j);
    }
    if (prec_vdv_ == 0) {  // FP64: the split-K tall product
      std::vector<AhbBigJob> big;
      for (auto const &j : jobs) big.push_back(AhbBigJob{j.V, j.W, j.C[0]});
      upload(bigjob_, big);
      ahb_part_.ensure(sizeof(double2) * (size_t)ahb_big_part_elems(int(big.size()), N_, K_));
      launch_ahb_big(bigjob_.as<AhbBigJob>(), int(big.size()), N_, K_, ahb_part_.as<double2>(), stream_);
    } else {
      upload(gjob_, jobs);
      if (prec_vdv_ >= 2)
        launch_vdv_split(gjob_.as<VdvJob>(), int(jobs.size()), N_, K_, prec_vdv_, stream_);
      else
        launch_vdv(gjob_.as<VdvJob>(), int(jobs.size()), N_, K_, prec_vdv_, stream_);
    }
    ++launches;
    SLAPH_HIP(hipStreamSynchronize(stream_));  // the job table and work space are reused
  }
  for (int id : ids) {
    const int src = first.at(gd_.vdaggerv_lookup.at(id).disp);
    if (src != id)
      SLAPH_HIP(hipMemcpyAsync(out_of(id), out_of(src), sizeof(double2) * N_ * N_, hipMemcpyDeviceToDevice, stream_));
  }
  return launches;
}

// Sitewise Fourier form (FP64): W_site = V_site^H V_site per 16 x 16 tile
// pair by MFMA, then the z, y and x transforms of the separable phase
// exp(-i 2 pi (px x / Lx + py y / Ly + pz z / Lz)) (create_momenta,
// OperatorFactory.cpp:64-95).  Everything below the phase tables depends on
// the input file only and is kept until the next configure().
struct Engine::FtPlan {
  int NT = 0, Npad = 0, ntp = 0, TC = 0;
  DevBuf tiles;  // int2 (I, J), I <= J
  struct Target {
    int id, c, lower;
  };
  struct Pass {
    FtPass p{};
    std::vector<Target> targets;
    DevBuf ey, zcs, ex, tg;
  };
  std::vector<Pass> passes;
  int max_nc = 0;
};

int Engine::vdaggerv_fourier(std::vector<int> const &ids, int t_first, int nt_all, const double2 *V) {
  const int Lx = gd_.Lx, Ly = gd_.Ly, Lz = gd_.Lz;
  if (!ft_) {
    auto F = std::make_shared<FtPlan>();
    F->NT = (N_ + 15) / 16;
    F->Npad = 16 * F->NT;
    std::vector<int2> tl;
    for (int I = 0; I < F->NT; ++I)
      for (int J = I; J < F->NT; ++J) tl.push_back(make_int2(I, J));
    F->ntp = int(tl.size());
    upload(F->tiles, tl);
    // Outputs: C(p) on every tile pair from the (py, pz) sums of +p, and the
    // lower triangle conj(C(-p)[i,j]) from those of -p.
    struct Want {
      int id, px, py, pz, lower;
    };
    std::vector<Want> want;
    std::map<std::pair<int, int>, int> combo_ix;
    for (int id : ids) {
      auto const &m = gd_.vdaggerv_lookup.at(id).mom;
      want.push_back({id, m[0], m[1], m[2], 0});
      if (F->NT > 1) want.push_back({id, -m[0], -m[1], -m[2], 1});
    }
    std::vector<std::pair<int, int>> combos;  // (py, pz)
    for (auto const &w : want)
      if (!combo_ix.count({w.py, w.pz})) {
        combo_ix[{w.py, w.pz}] = 0;
        combos.push_back({w.py, w.pz});
      }
    std::sort(combos.begin(), combos.end(), [](std::pair<int, int> a, std::pair<int, int> b) {
      return std::make_tuple(std::abs(a.second), a.second, a.first) <
             std::make_tuple(std::abs(b.second), b.second, b.first);
    });
    // Passes: at most kFtMaxC combinations and kFtMaxK nonzero |pz| each.
    std::vector<std::vector<std::pair<int, int>>> pc;
    std::vector<std::vector<int>> pk;
    std::map<std::pair<int, int>, std::pair<int, int>> where;  // combo -> (pass, index)
    for (auto const &cb : combos) {
      int k = std::abs(cb.second);
      bool fits = !pc.empty() && int(pc.back().size()) < kFtMaxC &&
                  (k == 0 || std::count(pk.back().begin(), pk.back().end(), k) ||
                   int(pk.back().size()) < kFtMaxK);
      if (!fits) {
        pc.emplace_back();
        pk.emplace_back();
      }
      if (k != 0 && !std::count(pk.back().begin(), pk.back().end(), k)) pk.back().push_back(k);
      where[cb] = {int(pc.size()) - 1, int(pc.back().size())};
      pc.back().push_back(cb);
    }
    const double tpi = 2.0 * M_PI;
    F->passes.resize(pc.size());
    for (size_t ip = 0; ip < pc.size(); ++ip) {
      FtPlan::Pass &P = F->passes[ip];
      P.p.nc = int(pc[ip].size());
      P.p.nk = int(pk[ip].size());
      F->max_nc = std::max(F->max_nc, P.p.nc);
      std::vector<cplx> ey((size_t)P.p.nc * Ly), zcs((size_t)kFtMaxK * Lz);  // rows past nk stay zero
      for (int c = 0; c < P.p.nc; ++c) {
        int py = pc[ip][c].first, pz = pc[ip][c].second;
        int s = pz == 0 ? -1 : int(std::find(pk[ip].begin(), pk[ip].end(), std::abs(pz)) - pk[ip].begin());
        P.p.tsel[c] = pz == 0 ? 0 : (pz > 0 ? 1 + 2 * s : 2 + 2 * s);
        for (int y = 0; y < Ly; ++y) ey[(size_t)c * Ly + y] = std::exp(cplx(0, -tpi * py * y / Ly));
      }
      for (int s = 0; s < P.p.nk; ++s)
        for (int z = 0; z < Lz; ++z) {
          double a = tpi * pk[ip][s] * z / Lz;
          zcs[(size_t)s * Lz + z] = cplx(std::cos(a), std::sin(a));
        }
      // targets sorted by combination (k_vdv_xsum reads each combination's
      // per-x sums once for a run of targets), x phases in the same order
      std::vector<std::pair<FtPlan::Target, int>> tl;  // (target, px)
      for (auto const &w : want) {
        auto at = where.at({w.py, w.pz});
        if (at.first != int(ip)) continue;
        tl.push_back({FtPlan::Target{w.id, at.second, w.lower}, w.px});
      }
      std::stable_sort(tl.begin(), tl.end(), [](auto const &a, auto const &b) { return a.first.c < b.first.c; });
      std::vector<cplx> ex;
      for (auto const &e : tl) {
        P.targets.push_back(e.first);
        for (int x = 0; x < Lx; ++x) ex.push_back(std::exp(cplx(0, -tpi * e.second * x / Lx)));
      }
      upload(P.ey, ey);
      upload(P.zcs, zcs);
      upload(P.ex, ex);
    }
    // Timeslice chunk: the row-major copy and the per-x sums within ~4 GiB.
    const double per_t = 16.0 * K_ * F->Npad + 16.0 * Lx * F->max_nc * F->ntp * 256;
    F->TC = std::max(1, std::min(Lt_, int((4.0 * (1ull << 30)) / per_t)));
    ft_ = F;
  }
  FtPlan &F = *ft_;
  // Output pointers of the targets (vdv_ may have moved since the plan).
  for (auto &P : F.passes) {
    std::vector<FtTarget> tg;
    for (auto const &t : P.targets)
      tg.push_back(FtTarget{vdv_.as<double2>() + (size_t)t.id * Lt_ * N_ * N_, t.c, t.lower});
    upload(P.tg, tg);
  }
  vr_.ensure(sizeof(double2) * (size_t)F.TC * K_ * F.Npad);
  ftu_.ensure(sizeof(double2) * (size_t)F.TC * Lx * F.max_nc * F.ntp * 256);
  int launches = 0;
  for (int c0 = 0; c0 < nt_all; c0 += F.TC) {
    const int nt = std::min(F.TC, nt_all - c0);
    launch_vrow(V + (size_t)c0 * K_ * N_, vr_.as<double2>(), K_, N_, F.Npad, nt, stream_);
    ++launches;
    for (auto &P : F.passes) {
      EventTimer tm(stream_);
      launch_vdv_ft(vr_.as<double2>(), K_ * F.Npad, FtLayout{F.Npad, 1, F.Npad}, Lx, Ly, Lz, F.tiles.as<int2>(), F.ntp,
                    nt, P.p,
                    P.ey.as<double2>(), P.zcs.as<double2>(), ftu_.as<double2>(), stream_);
      // own operation count: 24 per (element pair, site) for W = V^H V over
      // three colours, 5 per |pz| value for the z sums; bytes: V read once,
      // the per-(x, py, pz) sums written once
      KernelStat &ks = kstats_["k_vdv_ft"];
      ks.ms += tm.stop();
      ks.flops += (24.0 + 5.0 * std::max(1, P.p.nk)) * F.ntp * 256.0 * (K_ / 3) * nt;
      ks.bytes += 16.0 * K_ * N_ * nt + 16.0 * Lx * P.p.nc * F.ntp * 256.0 * nt;
      ++ks.launches;
      launch_vdv_xsum(ftu_.as<double2>(), P.p.nc, F.ntp, nt, Lx, F.tiles.as<int2>(), P.tg.as<FtTarget>(),
                      int(P.targets.size()), P.ex.as<double2>(), N_, t_first + c0, (long long)N_ * N_, stream_);
      launches += 2;
    }
  }
  return launches;
}

// write_vdaggerv (OperatorFactory.cpp:144-168, 303-313): every non-trivial
// operator at every timeslice, N x N column-major complex<double>, into
// "/<path_vdaggerv>/cnfg%04d/" (created when missing).
void Engine::write_vdaggerv_files(int config) {
  SLAPH_HIP(hipStreamSynchronize(stream_));
  char dir[2048];
  std::snprintf(dir, sizeof dir, "/%s/cnfg%04d", gd_.path_vdaggerv.c_str(), config);
  make_directories(dir);
  std::vector<cplx> m((size_t)N_ * N_);
  for (auto const &op : gd_.vdaggerv_lookup) {
    if (op.id == gd_.index_of_unity) continue;
    for (int t = 0; t < Lt_; ++t) {
      SLAPH_HIP(hipMemcpy(m.data(), vdv_ptr(op.id, false, t), sizeof(cplx) * m.size(), hipMemcpyDeviceToHost));
      write_binary(vdaggerv_write_name(gd_.path_vdaggerv, config, op.mom, op.disp, t), m.data(),
                   sizeof(cplx) * m.size());
    }
  }
}

// read_vdaggerv (OperatorFactory.cpp:328-394): one file per (operator, t),
// column-major.  read_vdaggerv_liuming (:396-493): one file per operator with
// the Lt matrices one after another, row-major; the file of -p holds the
// adjoint of VdaggerV(p) (adjointInPlace), that of +p is taken as it is.  The
// identity operator is not read.  A file that cannot be read is an error here
// (the reference constructs the runtime_error without throwing it).
void Engine::read_vdaggerv_files(int config, bool liuming) {
  std::vector<cplx> m((size_t)N_ * N_), all;
  for (auto const &op : gd_.vdaggerv_lookup) {
    double2 *dst = vdv_.as<double2>() + (size_t)op.id * Lt_ * N_ * N_;
    if (op.id == gd_.index_of_unity) {
      launch_identity(dst, N_, Lt_, stream_);
      continue;
    }
    if (!liuming) {
      for (int t = 0; t < Lt_; ++t) {
        read_binary(vdaggerv_read_name(gd_.path_vdaggerv, config, op.mom, t), m.data(), sizeof(cplx) * m.size());
        SLAPH_HIP(hipMemcpy(dst + (size_t)t * N_ * N_, m.data(), sizeof(cplx) * m.size(), hipMemcpyHostToDevice));
      }
      continue;
    }
    const Mom neg{{-op.mom[0], -op.mom[1], -op.mom[2]}};
    const std::string f1 = vdaggerv_liuming_name(gd_.path_vdaggerv, config, neg),
                      f2 = vdaggerv_liuming_name(gd_.path_vdaggerv, config, op.mom);
    const bool first = file_exists(f1);
    if (!first && !file_exists(f2)) throw std::runtime_error("can't open " + f1 + " NOR " + f2);
    all.resize((size_t)Lt_ * N_ * N_);
    read_binary(first ? f1 : f2, all.data(), sizeof(cplx) * all.size());
    for (int t = 0; t < Lt_; ++t) {
      const cplx *rm = all.data() + (size_t)t * N_ * N_;  // row-major: element (r, c) at r N + c
      for (int c = 0; c < N_; ++c)
        for (int r = 0; r < N_; ++r)
          m[(size_t)c * N_ + r] = first ? std::conj(rm[(size_t)c * N_ + r]) : rm[(size_t)r * N_ + c];
      SLAPH_HIP(hipMemcpy(dst + (size_t)t * N_ * N_, m.data(), sizeof(cplx) * m.size(), hipMemcpyHostToDevice));
    }
  }
}

const double2 *Engine::vdv_ptr(int id, bool dagger, int t) const {
  if (dagger) return vdvdag_.as<double2>() + ((size_t)dag_slot_.at(id) * Lt_ + t) * N_ * N_;
  return vdv_.as<double2>() + ((size_t)id * Lt_ + t) * N_ * N_;
}

void Engine::return_vdaggerv(int id, int t, cplx *out) {
  if (!have_ops_) throw std::runtime_error("operators have not been created");
  if (id < 0 || id >= int(gd_.vdaggerv_lookup.size()) || t < 0 || t >= Lt_)
    throw std::runtime_error("return_vdaggerv: index out of range");
  SLAPH_HIP(hipSetDevice(device_));
  SLAPH_HIP(hipMemcpy(out, vdv_ptr(id, false, t), sizeof(cplx) * N_ * N_, hipMemcpyDeviceToHost));
}

// ---------------------------------------------------------------------------
// Diluted factors on the device (DilutedFactorFactory.cpp:33-253)
// ---------------------------------------------------------------------------
void Engine::compute_q_offsets() {
  static const char *qn[3] = {"Q0", "Q1", "Q2"};
  for (int q = 0; q < 3; ++q) {
    auto const &lst = gd_.quarkline_lookuptable.at(qn[q]);
    q_off_[q].assign(lst.size() + 1, 0);
    for (size_t e = 0; e < lst.size(); ++e) q_off_[q][e + 1] = q_off_[q][e] + (long long)lst[e].rnd.size();
    q_total_[q] = q_off_[q].back();
  }
}

// The job tables of build_factors (device pointers into the slabs, the
// operators, the perambulators and the random vectors).  They depend on the
// factor keys and on those buffers' addresses only, so a repeated call with
// the same keys (every gauge configuration of a run) launches them as they are.
struct Engine::FactorJobs {
  std::vector<std::vector<int>> q1keys, q2keys;
  std::vector<const void *> sig;
  DevBuf q0g, q0o, q1j, drs;
  int nq0 = 0, nq1 = 0;
  struct PerT {
    DevBuf ag, ar, bj;  // stage-1 groups and rows, stage-2 jobs
    int ng = 0, maxj = 0, nb = 0;
  };
  std::vector<PerT> q2t;
};

void Engine::build_factors(std::vector<std::vector<int>> const &q1keys, std::vector<std::vector<int>> const &q2keys) {
  compute_q_offsets();
  const long long DD = (long long)D_ * D_;
  const long long prow = gd_.peram_rows();
  const size_t rsz = (size_t)Lt_ * N_ * 4, psz = gd_.peram_rows() * gd_.peram_cols();
  const double2 *rho = rho_.as<double2>();
  const double2 *peram = peram_.as<double2>();
  auto const &lst0 = gd_.quarkline_lookuptable.at("Q0");
  auto const &lst1 = gd_.quarkline_lookuptable.at("Q1");
  auto const &lst2 = gd_.quarkline_lookuptable.at("Q2");
  // slabs: Q0 for every timeslice, Q1 for the needed (t, b), Q2 for the
  // needed (b1, t, b2) (an unchanged size keeps the allocation)
  q0_.alloc(sizeof(double2) * std::max<long long>(1, Lt_ * q_total_[0] * DD));
  q1need_.assign((size_t)Lt_ * nb_, 0);
  for (auto const &k : q1keys) q1need_[k[0] * nb_ + k[1]] = 1;
  q1_.alloc(sizeof(double2) * std::max<long long>(1, (long long)Lt_ * nb_ * q_total_[1] * DD));
  q2slot_.assign((size_t)nb_ * Lt_ * nb_, -1);
  int nslot = 0;
  for (auto const &k : q2keys) {
    int &s = q2slot_[((size_t)k[0] * Lt_ + k[1]) * nb_ + k[2]];
    if (s < 0) s = nslot++;
  }
  q2_.alloc(sizeof(double2) * std::max<long long>(1, nslot * q_total_[2] * DD));
  const long long msz = (long long)D_ * 4 * N_;
  auto signature = [&] {
    return std::vector<const void *>{vdv_.as<void>(), vdvdag_.as<void>(), rho_.as<void>(), peram_.as<void>(),
                                     q0_.as<void>(),  q1_.as<void>(),     q2_.as<void>(),  mbuf_.as<void>()};
  };
  if (!fjobs_ || fjobs_->q1keys != q1keys || fjobs_->q2keys != q2keys || fjobs_->sig != signature()) {
    auto fj = std::make_shared<FactorJobs>();
    fj->q1keys = q1keys;
    fj->q2keys = q2keys;
    {
      // one group per (t, entry, r2): its jobs share X = VdV rho2
      std::vector<Q0Group> groups;
      std::vector<Q0Out> outs;
      for (int t = 0; t < Lt_; ++t)
        for (size_t e = 0; e < lst0.size(); ++e) {
          std::map<int, std::vector<size_t>> by_r2;
          for (size_t j = 0; j < lst0[e].rnd.size(); ++j) by_r2[lst0[e].rnd[j].second].push_back(j);
          for (auto const &kv : by_r2) {
            groups.push_back({vdv_ptr(lst0[e].id_vdaggerv, lst0[e].dagger, t), rho + kv.first * rsz + (size_t)t * N_ * 4,
                              lst0[e].gamma.at(0), int(outs.size()), int(kv.second.size())});
            for (size_t j : kv.second)
              outs.push_back({rho + lst0[e].rnd[j].first * rsz + (size_t)t * N_ * 4,
                              q0_.as<double2>() + ((long long)t * q_total_[0] + q_off_[0][e] + (long long)j) * DD});
          }
        }
      upload(fj->q0g, groups);
      upload(fj->q0o, outs);
      fj->nq0 = int(groups.size());
    }
    {
      std::vector<Q1Job> jobs;
      for (int t = 0; t < Lt_; ++t)
        for (int b = 0; b < nb_; ++b) {
          if (!q1need_[t * nb_ + b]) continue;
          for (size_t e = 0; e < lst1.size(); ++e)
            for (size_t j = 0; j < lst1[e].rnd.size(); ++j) {
              auto const &p = lst1[e].rnd[j];
              jobs.push_back(
                  {vdv_ptr(lst1[e].id_vdaggerv, lst1[e].dagger, t), rho + p.first * rsz + (size_t)t * N_ * 4,
                   peram + p.second * psz,
                   q1_.as<double2>() + (((long long)t * nb_ + b) * q_total_[1] + q_off_[1][e] + (long long)j) * DD, t,
                   b, lst1[e].gamma.at(0)});
            }
        }
      upload(fj->q1j, jobs);
      fj->nq1 = int(jobs.size());
    }
    // Q2: M' = gval(bd) s(bd,c) P1_bd^H VdV per (b1, t, operator, r1), then
    // Q2 = M' * stacked P2 blocks.  rowscale[g][bd][row] = gval_g(bd) *
    // s(bd, row / dilE); s = -1 iff exactly one of bd, c lies in {2,3}
    // (DilutedFactorFactory.cpp:215-216).
    auto const &gt = gamma_table();
    std::vector<cplx> rs(16 * 4 * D_);
    for (int g = 0; g < 16; ++g)
      for (int bd = 0; bd < 4; ++bd)
        for (int row = 0; row < D_; ++row) {
          int c = row / gd_.dilE();
          double sgn = ((bd + c) == 3 || std::abs(bd - c) > 1) ? -1.0 : 1.0;
          rs[(g * 4 + bd) * D_ + row] = gt[g].value[bd] * sgn;
        }
    upload(fj->drs, rs);
    // M' jobs per timeslice keyed by (b1, entry, r1); the M' buffer is shared
    // by the timeslices (they run one after another on the stream)
    std::vector<std::map<std::array<int, 3>, int>> mjob(Lt_);
    std::vector<std::vector<Q2BJob>> bjobs(Lt_);
    std::vector<std::vector<int>> bjob_m(Lt_);
    size_t mmax = 1;
    for (int t = 0; t < Lt_; ++t) {
      for (int b1 = 0; b1 < nb_; ++b1)
        for (int b2 = 0; b2 < nb_; ++b2) {
          int slot = q2slot_[((size_t)b1 * Lt_ + t) * nb_ + b2];
          if (slot < 0) continue;
          for (size_t e = 0; e < lst2.size(); ++e)
            for (size_t j = 0; j < lst2[e].rnd.size(); ++j) {
              auto const &p = lst2[e].rnd[j];
              std::array<int, 3> key{b1, int(e), p.first};
              auto it = mjob[t].find(key);
              if (it == mjob[t].end()) it = mjob[t].emplace(key, int(mjob[t].size())).first;
              bjobs[t].push_back({nullptr, peram + p.second * psz,
                                  q2_.as<double2>() + ((long long)slot * q_total_[2] + q_off_[2][e] + (long long)j) * DD,
                                  t, b2, lst2[e].gamma.at(0)});
              bjob_m[t].push_back(it->second);
            }
        }
      mmax = std::max(mmax, mjob[t].size());
    }
    mbuf_.ensure(sizeof(double2) * msz * mmax);
    fj->q2t.resize(Lt_);
    for (int t = 0; t < Lt_; ++t) {
      if (mjob[t].empty()) continue;
      // one group per operator entry: its jobs share the right matrix VdV(e, t)
      std::map<int, std::vector<AhBRow>> by_entry;
      for (auto const &kv : mjob[t]) {
        int b1 = kv.first[0], e = kv.first[1], r1 = kv.first[2];
        for (int bd = 0; bd < 4; ++bd)
          by_entry[e].push_back({peram + r1 * psz + (long long)(b1 * D_) * prow + (long long)(t * 4 + bd) * N_,
                                 fj->drs.as<double2>() + (lst2[e].gamma.at(0) * 4 + bd) * D_,
                                 mbuf_.as<double2>() + kv.second * msz + (long long)bd * N_ * D_});
      }
      std::vector<AhBGroup> groups;
      std::vector<AhBRow> rows;
      int maxj = 0;
      for (auto const &kv : by_entry) {
        groups.push_back({vdv_ptr(lst2[kv.first].id_vdaggerv, lst2[kv.first].dagger, t), int(rows.size()),
                          int(kv.second.size())});
        rows.insert(rows.end(), kv.second.begin(), kv.second.end());
        maxj = std::max(maxj, int(kv.second.size()));
      }
      for (size_t i = 0; i < bjobs[t].size(); ++i) bjobs[t][i].Mp = mbuf_.as<double2>() + (long long)bjob_m[t][i] * msz;
      upload(fj->q2t[t].ag, groups);
      upload(fj->q2t[t].ar, rows);
      upload(fj->q2t[t].bj, bjobs[t]);
      fj->q2t[t].ng = int(groups.size());
      fj->q2t[t].maxj = maxj;
      fj->q2t[t].nb = int(bjobs[t].size());
    }
    fj->sig = signature();
    fjobs_ = fj;
  }
  FactorJobs const &fj = *fjobs_;
  launch_q0(fj.q0g.as<Q0Group>(), fj.nq0, fj.q0o.as<Q0Out>(), N_, gd_.dilE(), stream_);
  launch_q1(fj.q1j.as<Q1Job>(), fj.nq1, N_, gd_.dilE(), prow, stream_);
  for (auto const &pt : fj.q2t) {
    if (pt.ng == 0) continue;
    launch_ahb_grouped(pt.ag.as<AhBGroup>(), pt.ng, pt.maxj, pt.ar.as<AhBRow>(), D_, N_, N_, prow, N_, D_, stream_);
    launch_q2b(pt.bj.as<Q2BJob>(), pt.nb, N_, gd_.dilE(), prow, stream_);
  }
  SLAPH_HIP(hipStreamSynchronize(stream_));
}

int Engine::diluted_factor(int qtype, const int *key, int op, cplx *out, int *ric, int cap) {
  require_inputs();
  if (!have_ops_) throw std::runtime_error("operators have not been created");
  static const char *qn[3] = {"Q0", "Q1", "Q2"};
  if (qtype < 0 || qtype > 2) throw std::runtime_error("qtype must be 0, 1 or 2");
  auto const &lst = gd_.quarkline_lookuptable.at(qn[qtype]);
  if (op < 0 || op >= int(lst.size())) throw std::runtime_error("operator index out of range");
  std::vector<std::vector<int>> k1, k2;
  long long base = 0;
  if (qtype == 1) {
    if (key[0] < 0 || key[0] >= Lt_ || key[1] < 0 || key[1] >= nb_) throw std::runtime_error("bad Q1 key");
    k1.push_back({key[0], key[1]});
  } else if (qtype == 2) {
    if (key[0] < 0 || key[0] >= nb_ || key[1] < 0 || key[1] >= Lt_ || key[2] < 0 || key[2] >= nb_)
      throw std::runtime_error("bad Q2 key");
    k2.push_back({key[0], key[1], key[2]});
  } else if (key[0] < 0 || key[0] >= Lt_) {
    throw std::runtime_error("bad Q0 key");
  }
  build_factors(k1, k2);
  const long long DD = (long long)D_ * D_;
  const double2 *src;
  if (qtype == 0) {
    base = (long long)key[0] * q_total_[0];
    src = q0_.as<double2>();
  } else if (qtype == 1) {
    base = ((long long)key[0] * nb_ + key[1]) * q_total_[1];
    src = q1_.as<double2>();
  } else {
    base = (long long)q2slot_[((size_t)key[0] * Lt_ + key[1]) * nb_ + key[2]] * q_total_[2];
    src = q2_.as<double2>();
  }
  int n = int(lst[op].rnd.size());
  if (n > cap) throw std::runtime_error("output capacity too small");
  SLAPH_HIP(hipMemcpy(out, src + (base + q_off_[qtype][op]) * DD, sizeof(cplx) * DD * n, hipMemcpyDeviceToHost));
  for (int j = 0; j < n; ++j) {
    ric[2 * j] = lst[op].rnd[j].first;
    ric[2 * j + 1] = lst[op].rnd[j].second;
  }
  return n;
}

std::vector<std::string> Engine::diagrams() const {
  std::vector<std::string> out;
  for (auto const &r : results_) out.push_back(r.name);
  return out;
}

int Engine::correlator_count(std::string const &d) const {
  for (auto const &r : results_)
    if (r.name == d) return int(r.names.size());
  throw std::runtime_error("no results for diagram " + d);
}

std::string Engine::correlator_name(std::string const &d, int i) const {
  for (auto const &r : results_)
    if (r.name == d) return r.names.at(i);
  throw std::runtime_error("no results for diagram " + d);
}

void Engine::fetch_correlators(std::string const &d, cplx *out) const {
  for (auto const &r : results_)
    if (r.name == d) {
      std::copy(r.values.begin(), r.values.end(), out);
      return;
    }
  throw std::runtime_error("no results for diagram " + d);
}

// ---------------------------------------------------------------------------
// Tuning: text form and the autotuner
// ---------------------------------------------------------------------------
namespace {
struct TuneKnob {
  const char *name;
  int Engine::Tuning::*field = nullptr;
  int LaunchTuning::*lfield = nullptr;
  std::vector<int> candidates;
  bool vdaggerv = false;  // timed on create_operators instead of contract
};
std::vector<TuneKnob> tune_knobs() {
  std::vector<TuneKnob> k(7);
  k[0] = {"vdaggerv_direct", &Engine::Tuning::vdv_direct, nullptr, {0, 1}, true};
  k[1] = {"lstat_wg_per_cu", nullptr, &LaunchTuning::lstat_wg_per_cu, {2, 1, 3}, false};
  k[2] = {"gemm_wg_per_cu", nullptr, &LaunchTuning::gemm_wg_per_cu, {16, 8, 32}, false};
  k[3] = {"prod_generic", nullptr, &LaunchTuning::prod_generic, {0, 1}, false};
  k[4] = {"resolve_groups", &Engine::Tuning::resolve_groups, nullptr, {0, 1}, false};
  k[5] = {"crossed_direct", &Engine::Tuning::xform_off, nullptr, {0, 1}, false};
  k[6] = {"crossed_lstat", &Engine::Tuning::sym_off, nullptr, {0, 1}, false};
  return k;
}
int &knob_ref(Engine::Tuning &t, TuneKnob const &k) { return k.field ? t.*(k.field) : t.launch.*(k.lfield); }
}  // namespace

std::string Engine::tuning_text() const {
  std::string out;
  Tuning t = tune_;
  for (auto const &k : tune_knobs()) {
    if (!out.empty()) out += ' ';
    out += std::string(k.name) + "=" + std::to_string(knob_ref(t, k));
  }
  return out;
}

void Engine::set_tuning(std::string const &text) {
  Tuning t = tune_;
  auto knobs = tune_knobs();
  size_t pos = 0;
  while (pos < text.size()) {
    while (pos < text.size() && (text[pos] == ' ' || text[pos] == '\t' || text[pos] == '\n')) ++pos;
    if (pos >= text.size()) break;
    size_t end = text.find_first_of(" \t\n", pos);
    if (end == std::string::npos) end = text.size();
    const std::string item = text.substr(pos, end - pos);
    pos = end;
    const size_t eq = item.find('=');
    if (eq == std::string::npos) throw std::runtime_error("tuning: expected key=value, got '" + item + "'");
    const std::string key = item.substr(0, eq), val = item.substr(eq + 1);
    auto it = std::find_if(knobs.begin(), knobs.end(), [&](TuneKnob const &k) { return key == k.name; });
    if (it == knobs.end()) throw std::runtime_error("tuning: unknown key '" + key + "'");
    char *stop = nullptr;
    const long v = std::strtol(val.c_str(), &stop, 10);
    if (val.empty() || *stop != '\0' || v < 0 || v > 64) throw std::runtime_error("tuning: bad value for " + key);
    knob_ref(t, *it) = int(v);
  }
  tune_ = t;
  if (configured_) tuned_[infile_] = t;
  plan_.reset();
  ft_.reset();
}

std::string Engine::autotune(int repeats) {
  require_inputs();
  if (!have_ops_) throw std::runtime_error("autotune: create the operators first (the contraction is timed on them)");
  if (repeats < 1 || repeats > 100) throw std::out_of_range("autotune: repeats must be in 1 .. 100");
  auto timed = [&](bool vdv) {
    double best = 1e300;
    // untimed first call(s): the plan (and the graph capture) of this choice
    for (int i = 0; i < (vdv ? 1 : 2); ++i) vdv ? create_operators(config_) : contract(-1, 0);
    for (int r = 0; r < repeats; ++r) {
      auto t0 = std::chrono::steady_clock::now();
      vdv ? create_operators(config_) : contract(-1, 0);
      best = std::min(best, std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count());
    }
    return best;
  };
  std::string report = "{\"knobs\": [";
  Tuning best = tune_;
  bool first_knob = true;
  for (auto const &k : tune_knobs()) {
    // the VdaggerV form only matters when the operators are computed here
    if (k.vdaggerv && gd_.handling_vdaggerv != "build" && gd_.handling_vdaggerv != "write") continue;
    double best_ms = -1.0;
    int best_v = knob_ref(best, k);
    std::string meas;
    for (int v : k.candidates) {
      tune_ = best;
      knob_ref(tune_, k) = v;
      plan_.reset();
      ft_.reset();
      const double ms = timed(k.vdaggerv);
      if (!meas.empty()) meas += ", ";
      meas += "\"" + std::to_string(v) + "\": " + std::to_string(ms);
      if (best_ms < 0 || ms < best_ms * 0.995) {
        best_ms = ms;
        best_v = v;
      }
    }
    knob_ref(best, k) = best_v;
    report += std::string(first_knob ? "" : ", ") + "{\"knob\": \"" + k.name + "\", \"ms\": {" + meas +
              "}, \"chosen\": " + std::to_string(best_v) + (k.vdaggerv ? ", \"stage\": \"vdaggerv\"" : "") + "}";
    first_knob = false;
  }
  tune_ = best;
  tuned_[infile_] = best;
  plan_.reset();
  ft_.reset();
  create_operators(config_);  // operators of the chosen form
  report += "], \"chosen\": \"" + tuning_text() + "\"}";
  return report;
}

}  // namespace slaph
