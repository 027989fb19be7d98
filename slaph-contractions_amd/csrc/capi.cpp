// C ABI of the contraction path (declared in include/slaph_amd.h).  No C++
// exception crosses this boundary: each entry converts them to error codes
// and keeps the message for slc_last_error().
#include "../../include/slaph_amd.h"

#include "engine.hpp"
#include "lime.hpp"

#include <cstring>
#include <new>
#include <string>

struct slc_ctx {
  std::unique_ptr<slaph::Engine> eng;
  std::string err;
  std::string tune_report;  // JSON report of the last slc_autotune
};

namespace {

// Message of the last failed slc_ctx_create or context-free call on this
// thread (no context exists to hold it); returned by slc_last_error(NULL).
thread_local std::string g_create_error;

template <typename F>
int guarded(slc_ctx *ctx, F &&f) {
  if (!ctx) return SLC_ERR_ARG;
  try {
    ctx->err.clear();
    f();
    return SLC_OK;
  } catch (std::out_of_range const &e) {
    ctx->err = e.what();
    return SLC_ERR_ARG;
  } catch (std::runtime_error const &e) {
    ctx->err = e.what();
    return std::strncmp(e.what(), "HIP error", 9) == 0 ? SLC_ERR_DEVICE : SLC_ERR_RUNTIME;
  } catch (std::exception const &e) {
    ctx->err = e.what();
    return SLC_ERR_RUNTIME;
  }
}

int copy_string(std::string const &s, char *buf, int len) {
  if (buf && len > 0) {
    size_t n = std::min<size_t>(s.size(), size_t(len) - 1);
    std::memcpy(buf, s.data(), n);
    buf[n] = 0;
  }
  return int(s.size()) + 1;
}

}  // namespace

extern "C" {

int slc_ctx_create(int device, slc_ctx **out) {
  if (!out) return SLC_ERR_ARG;
  *out = nullptr;
  slc_ctx *c = new (std::nothrow) slc_ctx;
  if (!c) return SLC_ERR_RUNTIME;
  try {
    c->eng.reset(new slaph::Engine(device));
  } catch (std::exception const &e) {
    g_create_error = e.what();
    delete c;
    return SLC_ERR_DEVICE;
  }
  *out = c;
  return SLC_OK;
}

void slc_ctx_destroy(slc_ctx *ctx) { delete ctx; }

const char *slc_last_error(const slc_ctx *ctx) { return ctx ? ctx->err.c_str() : g_create_error.c_str(); }

int slc_configure(slc_ctx *ctx, const char *text) {
  return guarded(ctx, [&] {
    if (!text) throw std::out_of_range("null infile text");
    ctx->eng->configure(text);
  });
}

int slc_query(const slc_ctx *cctx, const char *what, long long *out) {
  slc_ctx *ctx = const_cast<slc_ctx *>(cctx);
  return guarded(ctx, [&] {
    if (!what || !out) throw std::out_of_range("null argument");
    auto const &gd = ctx->eng->gd();
    std::string w = what;
    if (gd.quarks.empty()) throw std::out_of_range("context is not configured");
    if (w == "Lt") *out = gd.Lt;
    else if (w == "Lx") *out = gd.Lx;
    else if (w == "Ly") *out = gd.Ly;
    else if (w == "Lz") *out = gd.Lz;
    else if (w == "nev") *out = gd.nev;
    else if (w == "dilT") *out = gd.dilT();
    else if (w == "dilE") *out = gd.dilE();
    else if (w == "n_rnd") *out = gd.n_rnd_total();
    else if (w == "peram_rows") *out = gd.peram_rows();
    else if (w == "peram_cols") *out = gd.peram_cols();
    else if (w == "n_vdaggerv") *out = (long long)gd.vdaggerv_lookup.size();
    else if (w == "index_of_unity") *out = gd.index_of_unity;
    else if (w == "n_Q0") *out = (long long)gd.quarkline_lookuptable.at("Q0").size();
    else if (w == "n_Q1") *out = (long long)gd.quarkline_lookuptable.at("Q1").size();
    else if (w == "n_Q2") *out = (long long)gd.quarkline_lookuptable.at("Q2").size();
    else if (w == "start_config") *out = gd.start_config;
    else if (w == "end_config") *out = gd.end_config;
    else if (w == "delta_config") *out = gd.delta_config;
    else if (w == "need_gaugefield") *out = gd.need_gaugefield ? 1 : 0;
    else throw std::out_of_range("unknown query " + w);
  });
}

int slc_query_string(const slc_ctx *cctx, const char *what, char *buf, int len) {
  slc_ctx *ctx = const_cast<slc_ctx *>(cctx);
  int need = 0;
  const int rc = guarded(ctx, [&] {
    if (!what) throw std::out_of_range("null argument");
    auto const &gd = ctx->eng->gd();
    if (gd.quarks.empty()) throw std::out_of_range("context is not configured");
    std::string w = what, v;
    if (w == "output_path") v = gd.path_output;
    else if (w == "handling_vdaggerv") v = gd.handling_vdaggerv;
    else if (w == "path_vdaggerv") v = gd.path_vdaggerv;
    else if (w == "path_config") v = gd.path_config;
    else if (w == "path_eigenvectors") v = gd.path_eigenvectors;
    else if (w == "name_eigenvectors") v = gd.name_eigenvectors;
    else throw std::out_of_range("unknown query " + w);
    need = copy_string(v, buf, len);
  });
  return rc < 0 ? rc : need;
}

int slc_set_eigenvectors(slc_ctx *ctx, int t, const double *V) {
  return guarded(ctx, [&] { ctx->eng->set_eigenvectors(t, reinterpret_cast<const slaph::cplx *>(V)); });
}

int slc_set_perambulator(slc_ctx *ctx, int r, const double *P) {
  return guarded(ctx, [&] { ctx->eng->set_perambulator(r, reinterpret_cast<const slaph::cplx *>(P)); });
}

int slc_set_random_vector(slc_ctx *ctx, int r, const double *rho) {
  return guarded(ctx, [&] { ctx->eng->set_random_vector(r, reinterpret_cast<const slaph::cplx *>(rho)); });
}

int slc_read_inputs(slc_ctx *ctx, const char *base, int config) {
  return guarded(ctx, [&] { ctx->eng->read_inputs(base ? base : ".", config); });
}

int slc_synthetic_inputs(slc_ctx *ctx, unsigned long long seed, int orth) {
  return guarded(ctx, [&] { ctx->eng->synthetic_inputs(seed, orth != 0); });
}

int slc_get_input(slc_ctx *ctx, int kind, int index, double *out) {
  return guarded(ctx, [&] {
    if (!out) throw std::out_of_range("null output");
    ctx->eng->get_input(kind, index, reinterpret_cast<slaph::cplx *>(out));
  });
}

int slc_set_timeslice_streaming(slc_ctx *ctx, int on) {
  return guarded(ctx, [&] { ctx->eng->set_streaming(on != 0); });
}

int slc_set_gauge_links(slc_ctx *ctx, int t, const double *U) {
  return guarded(ctx, [&] {
    if (!U) throw std::out_of_range("null gauge links");
    ctx->eng->set_gauge_links(t, reinterpret_cast<const slaph::cplx *>(U));
  });
}

int slc_get_gauge_links(slc_ctx *ctx, int t, double *U) {
  return guarded(ctx, [&] {
    if (!U) throw std::out_of_range("null output");
    ctx->eng->get_gauge_links(t, reinterpret_cast<slaph::cplx *>(U));
  });
}

int slc_read_gauge_field(slc_ctx *ctx, const char *base, const char *config_path, int config, int slice_i,
                         int slice_f) {
  return guarded(ctx, [&] {
    ctx->eng->read_gauge_field(base ? base : ".", config_path ? config_path : "", config, slice_i, slice_f);
  });
}

int slc_read_lime_gauge(const char *path, int Lx, int Ly, int Lz, int Lt, int slice_i, int slice_f, double *U) {
  try {
    if (!path || !U) throw std::out_of_range("null argument");
    const size_t n = 27 * (size_t)Lx * Ly * Lz;
    slaph::read_ildg_gauge(path, Lx, Ly, Lz, Lt, slice_i, slice_f, [&](int t, const slaph::cplx *links) {
      std::memcpy(U + 2 * n * size_t(t - slice_i), links, sizeof(slaph::cplx) * n);
    });
    return SLC_OK;
  } catch (std::out_of_range const &e) {
    g_create_error = e.what();
    return SLC_ERR_ARG;
  } catch (std::exception const &e) {
    g_create_error = e.what();
    return SLC_ERR_RUNTIME;
  }
}

int slc_plaquette(slc_ctx *ctx, int t, double *plaquette) {
  return guarded(ctx, [&] {
    if (!plaquette) throw std::out_of_range("null output");
    *plaquette = ctx->eng->plaquette(t);
  });
}

int slc_synthetic_gauge(slc_ctx *ctx, unsigned long long seed, double eps) {
  return guarded(ctx, [&] { ctx->eng->synthetic_gauge(seed, eps); });
}

int slc_smear_stout(slc_ctx *ctx, double rho, int iterations) {
  return guarded(ctx, [&] { ctx->eng->smear_stout(rho, iterations); });
}

int slc_smear_hyp(slc_ctx *ctx, double alpha1, double alpha2, int iterations) {
  return guarded(ctx, [&] { ctx->eng->smear_hyp(alpha1, alpha2, iterations); });
}

int slc_create_operators(slc_ctx *ctx) {
  return guarded(ctx, [&] { ctx->eng->create_operators(ctx->eng->last_config()); });
}

int slc_create_operators_config(slc_ctx *ctx, int config) {
  return guarded(ctx, [&] { ctx->eng->create_operators(config); });
}

int slc_vdaggerv_filename(int kind, const char *path, int config, const int *mom, const char *displacement, int t,
                          char *buf, int len) {
  try {
    if (!path || !mom) return SLC_ERR_ARG;
    slaph::Mom p{{mom[0], mom[1], mom[2]}};
    std::string name;
    if (kind == 0) {
      slaph::Displacement d;
      std::string ds = displacement ? displacement : "";
      if (ds.size() % 2) return SLC_ERR_ARG;
      for (size_t i = 0; i < ds.size(); i += 2) d.push_back({ds[i], ds[i + 1]});
      name = slaph::vdaggerv_write_name(path, config, p, d, t);
    } else if (kind == 1) {
      name = slaph::vdaggerv_read_name(path, config, p, t);
    } else if (kind == 2) {
      name = slaph::vdaggerv_liuming_name(path, config, p);
    } else {
      return SLC_ERR_ARG;
    }
    return copy_string(name, buf, len);
  } catch (std::exception const &) {
    return SLC_ERR_RUNTIME;
  }
}

int slc_return_vdaggerv(slc_ctx *ctx, int id, int t, double *out) {
  return guarded(ctx, [&] { ctx->eng->return_vdaggerv(id, t, reinterpret_cast<slaph::cplx *>(out)); });
}

int slc_diluted_factor(slc_ctx *ctx, int qtype, const int *key, int op, double *out, int *ric, int cap, int *n) {
  return guarded(ctx, [&] {
    if (!key || !out || !ric || !n) throw std::out_of_range("null argument");
    *n = ctx->eng->diluted_factor(qtype, key, op, reinterpret_cast<slaph::cplx *>(out), ric, cap);
  });
}

int slc_inner_product(slc_ctx *ctx, int ntraces, const double *const *values, const uint16_t *const *used,
                      const int *n, int kahan, double *out) {
  return guarded(ctx, [&] {
    if (!values || !used || !n || !out || ntraces < 1 || ntraces > 3)
      throw std::out_of_range("slc_inner_product: bad arguments");
    const slaph::cplx *v[3] = {nullptr, nullptr, nullptr};
    for (int k = 0; k < ntraces; ++k) v[k] = reinterpret_cast<const slaph::cplx *>(values[k]);
    slaph::cplx r = ctx->eng->inner_product(ntraces, v, used, n, kahan);
    out[0] = r.real();
    out[1] = r.imag();
  });
}

int slc_diluted_trace(slc_ctx *ctx, const char *trace, int t_source, int t_sink, const int *locations, int nloc,
                      int *counts, int count_cap, double *values, uint16_t *used, int cap, int *n_ids, int *total) {
  return guarded(ctx, [&] {
    if (!trace || (nloc > 0 && !locations) || !n_ids || !total) throw std::out_of_range("slc_diluted_trace: null argument");
    std::vector<slaph::Loc> locs;
    for (int i = 0; i < nloc; ++i) {
      if (locations[i] != 0 && locations[i] != 1) throw std::out_of_range("location must be 0 (source) or 1 (sink)");
      locs.push_back(locations[i] == 0 ? slaph::Loc::source : slaph::Loc::sink);
    }
    auto tv = ctx->eng->diluted_trace(trace, t_source, t_sink, locs);
    long long need = 0;
    for (auto const &v : tv) need += (long long)v.size();
    *n_ids = int(tv.size());
    *total = int(need);
    if (int(tv.size()) > count_cap || need > cap || (need > 0 && (!counts || !values || !used)))
      throw std::out_of_range("slc_diluted_trace: output capacity too small (" + std::to_string(tv.size()) +
                              " trace ids, " + std::to_string(need) + " values)");
    long long at = 0;
    for (size_t id = 0; id < tv.size(); ++id) {
      counts[id] = int(tv[id].size());
      for (auto const &e : tv[id]) {
        values[2 * at] = e.first.real();
        values[2 * at + 1] = e.first.imag();
        used[at] = e.second;
        ++at;
      }
    }
  });
}

int slc_contract(slc_ctx *ctx, int single, int kahan) {
  return guarded(ctx, [&] { ctx->eng->contract(single, kahan); });
}

int slc_diagram_count(const slc_ctx *cctx, int *n) {
  slc_ctx *ctx = const_cast<slc_ctx *>(cctx);
  return guarded(ctx, [&] { *n = int(ctx->eng->diagrams().size()); });
}

int slc_diagram_name(const slc_ctx *cctx, int i, char *buf, int len) {
  slc_ctx *ctx = const_cast<slc_ctx *>(cctx);
  int need = 0;
  int rc = guarded(ctx, [&] { need = copy_string(ctx->eng->diagrams().at(i), buf, len); });
  return rc == SLC_OK ? need : rc;
}

int slc_correlator_count(const slc_ctx *cctx, const char *d, int *n) {
  slc_ctx *ctx = const_cast<slc_ctx *>(cctx);
  return guarded(ctx, [&] { *n = ctx->eng->correlator_count(d); });
}

int slc_correlator_name(const slc_ctx *cctx, const char *d, int i, char *buf, int len) {
  slc_ctx *ctx = const_cast<slc_ctx *>(cctx);
  int need = 0;
  int rc = guarded(ctx, [&] { need = copy_string(ctx->eng->correlator_name(d, i), buf, len); });
  return rc == SLC_OK ? need : rc;
}

int slc_fetch_correlators(const slc_ctx *cctx, const char *d, double *out) {
  slc_ctx *ctx = const_cast<slc_ctx *>(cctx);
  return guarded(ctx, [&] { ctx->eng->fetch_correlators(d, reinterpret_cast<slaph::cplx *>(out)); });
}

int slc_stage_time(const slc_ctx *cctx, const char *what, double *ms) {
  slc_ctx *ctx = const_cast<slc_ctx *>(cctx);
  return guarded(ctx, [&] {
    std::string w = what ? what : "";
    auto const &t = ctx->eng->times();
    if (w == "vdaggerv") *ms = t.vdaggerv_ms;
    else if (w == "vdaggerv_kernel") *ms = t.vdaggerv_kernel_ms;
    else if (w == "factors") *ms = t.factors_ms;
    else if (w == "contract") *ms = t.contract_ms;
    else throw std::out_of_range("unknown stage " + w);
  });
}

int slc_kernel_stat(const slc_ctx *cctx, const char *kernel, const char *what, double *value) {
  slc_ctx *ctx = const_cast<slc_ctx *>(cctx);
  return guarded(ctx, [&] {
    std::string w = what ? what : "";
    auto const ks = ctx->eng->kernel_stat(kernel ? kernel : "");
    if (w == "ms") *value = ks.ms;
    else if (w == "launches") *value = ks.launches;
    else if (w == "bytes") *value = ks.bytes;
    else if (w == "flops") *value = ks.flops;
    else throw std::out_of_range("unknown kernel statistic " + w);
  });
}

int slc_autotune(slc_ctx *ctx, int repeats) {
  return guarded(ctx, [&] { ctx->tune_report = ctx->eng->autotune(repeats); });
}

int slc_autotune_report(const slc_ctx *cctx, char *buf, int len) {
  slc_ctx *ctx = const_cast<slc_ctx *>(cctx);
  int need = 0;
  const int rc = guarded(ctx, [&] { need = copy_string(ctx->tune_report, buf, len); });
  return rc < 0 ? rc : need;
}

int slc_get_tuning(const slc_ctx *cctx, char *buf, int len) {
  slc_ctx *ctx = const_cast<slc_ctx *>(cctx);
  int need = 0;
  const int rc = guarded(ctx, [&] { need = copy_string(ctx->eng->tuning_text(), buf, len); });
  return rc < 0 ? rc : need;
}

int slc_set_tuning(slc_ctx *ctx, const char *text) {
  return guarded(ctx, [&] {
    if (!text) throw std::out_of_range("null tuning text");
    ctx->eng->set_tuning(text);
  });
}

int slc_set_graphs(slc_ctx *ctx, int on) {
  return guarded(ctx, [&] { ctx->eng->set_graphs(on != 0); });
}

int slc_set_chunk_budget(slc_ctx *ctx, long long bytes) {
  return guarded(ctx, [&] {
    if (bytes < 0) throw std::out_of_range("budget must not be negative");
    ctx->eng->set_chunk_budget(size_t(bytes));
  });
}

int slc_set_precision(slc_ctx *ctx, int vdaggerv, int factors) {
  return guarded(ctx, [&] { ctx->eng->set_precision(vdaggerv, factors); });
}

int slc_ahb_device(slc_ctx *ctx, const double *A, const double *B, double *C, const double *phase, int M, int N,
                   int K, long long lda, long long ldb, long long ldc, int batch, long long sA, long long sB,
                   long long sC) {
  return guarded(ctx, [&] {
    if (M <= 0 || N <= 0 || K <= 0 || batch <= 0 || lda < K || ldb < K || ldc < M)
      throw std::out_of_range("invalid matrix shape");
    ctx->eng->ahb_raw(reinterpret_cast<const double2 *>(A), reinterpret_cast<const double2 *>(B),
                      reinterpret_cast<double2 *>(C), reinterpret_cast<const double2 *>(phase), M, N, K, lda, ldb, ldc,
                      batch, sA, sB, sC);
  });
}

int slc_ahb_device_noalloc(slc_ctx *ctx, const double *A, const double *B, double *C, const double *phase, int M,
                           int N, int K, long long lda, long long ldb, long long ldc, int batch, long long sA,
                           long long sB, long long sC) {
  return guarded(ctx, [&] {
    if (M <= 0 || N <= 0 || K <= 0 || batch <= 0 || lda < K || ldb < K || ldc < M)
      throw std::out_of_range("invalid matrix shape");
    ctx->eng->ahb_noalloc(reinterpret_cast<const double2 *>(A), reinterpret_cast<const double2 *>(B),
                          reinterpret_cast<double2 *>(C), reinterpret_cast<const double2 *>(phase), M, N, K, lda, ldb,
                          ldc, batch, sA, sB, sC);
  });
}

int slc_ahb_host(slc_ctx *ctx, const double *A, const double *B, double *C, const double *phase, int M, int N,
                 int K) {
  return guarded(ctx, [&] {
    if (M <= 0 || N <= 0 || K <= 0 || !A || !B || !C) throw std::out_of_range("invalid matrix shape");
    slaph::DevBuf dA, dB, dC, dP;
    dA.alloc(sizeof(double2) * (size_t)K * M);
    dB.alloc(sizeof(double2) * (size_t)K * N);
    dC.alloc(sizeof(double2) * (size_t)M * N);
    SLAPH_HIP(hipMemcpy(dA.as<void>(), A, dA.bytes(), hipMemcpyHostToDevice));
    SLAPH_HIP(hipMemcpy(dB.as<void>(), B, dB.bytes(), hipMemcpyHostToDevice));
    if (phase) {
      dP.alloc(sizeof(double2) * (size_t)((K + 2) / 3));
      SLAPH_HIP(hipMemcpy(dP.as<void>(), phase, dP.bytes(), hipMemcpyHostToDevice));
    }
    ctx->eng->ahb_raw(dA.as<double2>(), dB.as<double2>(), dC.as<double2>(), phase ? dP.as<double2>() : nullptr, M,
                      N, K, K, K, M, 1, 0, 0, 0);
    SLAPH_HIP(hipMemcpy(C, dC.as<void>(), dC.bytes(), hipMemcpyDeviceToHost));
  });
}

int slc_lookup_dump(const char *text, char *buf, int len) {
  if (!text) return SLC_ERR_ARG;
  try {
    slaph::GlobalData gd;
    slaph::parse_infile(gd, text);
    slaph::init_lookup_tables(gd);
    return copy_string(slaph::lookup_dump(gd), buf, len);
  } catch (std::exception const &e) {
    copy_string(std::string("error: ") + e.what(), buf, len);
    return SLC_ERR_RUNTIME;
  }
}

int slc_dilution_text(int num_slice, int num_block, int interlace, char *buf, int len) {
  if (num_slice <= 0 || num_block <= 0 || num_slice % num_block) return SLC_ERR_ARG;
  return copy_string(slaph::dilution_iterator_text(num_slice, num_block, interlace != 0), buf, len);
}

}  // extern "C"
