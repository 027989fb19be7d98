/*
 * slaph_amd.h -- C ABI of the MI355X contraction path of sLapH-contractions.
 *
 * The reference (pittlerf/sLapH-contractions) has no FFI: its boundary is the
 * C++ class layer driven by main/contract.cpp.  Each entry point below
 * replaces one step of that flow, cited as file:line of the reference.  Host
 * callers keep the reference semantics: inputs are read-only, correlators are
 * sum_{slice pairs} c(t_source, t_sink) / Lt at t = (t_source - t_sink) mod Lt.
 *
 * Conventions
 *  - complex numbers are interleaved (re, im) doubles, matrices column-major;
 *  - every function returns 0 on success and a negative code on failure;
 *    slc_last_error() then holds the message (the reference's
 *    std::runtime_error text where one exists);
 *  - one context per host thread; a context is bound to one HIP device.
 */
#ifndef SLAPH_AMD_H
#define SLAPH_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct slc_ctx slc_ctx;

enum {
  SLC_OK = 0,
  SLC_ERR_ARG = -1,     /* invalid argument or state */
  SLC_ERR_RUNTIME = -2, /* reference runtime_error (e.g. "Not enough random vectors?") */
  SLC_ERR_DEVICE = -3   /* HIP error */
};

/* Context with its HIP stream and device buffers.  No reference counterpart
 * (the reference keeps its state in GlobalData and the factory objects). */
int slc_ctx_create(int device, slc_ctx **out);
void slc_ctx_destroy(slc_ctx *ctx);
/* Message of the last failure on ctx; with ctx == NULL, of the last failed
 * slc_ctx_create or slc_read_lime_gauge on this thread. */
const char *slc_last_error(const slc_ctx *ctx);

/* read_parameters + init_lookup_tables from the text of an infile
 * (reference: src/global_data.cpp:61-390, src/init_lookup_tables.cpp:470-556).
 * Allocates the device input buffers for the lattice it describes. */
int slc_configure(slc_ctx *ctx, const char *infile_text);

/* Sizes after slc_configure.  `what` is one of "Lt", "Lx", "Ly", "Lz", "nev",
 * "dilT", "dilE", "n_rnd", "peram_rows", "peram_cols", "n_vdaggerv",
 * "index_of_unity", "n_Q0", "n_Q1", "n_Q2", "start_config", "end_config",
 * "delta_config", "need_gaugefield" (1 when an operator is displaced:
 * OperatorLookup::need_gaugefield, init_lookup_tables.cpp). */
int slc_query(const slc_ctx *ctx, const char *what, long long *out);

/* String settings of the infile after slc_configure (GlobalData members,
 * include/global_data.hpp:51-107): "output_path", "handling_vdaggerv",
 * "path_vdaggerv", "path_config", "path_eigenvectors", "name_eigenvectors".
 * Copies at most len bytes (NUL-terminated) and returns the bytes needed
 * including the NUL, or a negative code. */
int slc_query_string(const slc_ctx *ctx, const char *what, char *buf, int len);

/* EigenVector::read_eigen_vector for one timeslice (reference:
 * src/EigenVector.cpp:35-76): V is (3 Lx Ly Lz) x nev, column-major. */
int slc_set_eigenvectors(slc_ctx *ctx, int t, const double *V);

/* Perambulator in the reference's in-memory layout after the re-sort of
 * Perambulator::read_perambulator (src/Perambulator.cpp:48-64): rows
 * 4 nev t + nev d + v, columns dilE 4 b + dilE d' + e, column-major. */
int slc_set_perambulator(slc_ctx *ctx, int entity, const double *P);

/* RandomVector entity (include/RandomVector.hpp:38-40): entry 4(nev t + v) + d. */
int slc_set_random_vector(slc_ctx *ctx, int entity, const double *rho);

/* Read perambulators, random vectors and eigenvectors of configuration
 * `config` from the reference's file layout, paths relative to base_dir
 * (main/contract.cpp:83-93, src/global_data_build_IO_names.cpp:7-85). */
int slc_read_inputs(slc_ctx *ctx, const char *base_dir, int config);

/* Seeded synthetic inputs generated on the device: complex Gaussian
 * eigenvectors (orthonormalised by Cholesky-QR when orthonormalise != 0),
 * complex Gaussian perambulators, Z2xZ2 random vectors.  Replaces the
 * `mock` fills of the reference (EigenVector.cpp:73-75, Perambulator.cpp:65-68). */
int slc_synthetic_inputs(slc_ctx *ctx, unsigned long long seed, int orthonormalise);

/* Timeslice streaming (on != 0): only one timeslice of eigenvectors is kept
 * on the device; slc_create_operators produces V(t) timeslice by timeslice
 * from the eigenvector files of slc_read_inputs or from the seeded generator
 * of slc_synthetic_inputs, as the reference reads them per timeslice while it
 * builds VdaggerV (src/OperatorFactory.cpp:263-277).  Call before the
 * inputs (before slc_configure as well, so that the full eigenvector array
 * is never allocated); slc_set_eigenvectors is then refused.  The mode stays
 * set across later slc_configure calls.
 * No reference counterpart as a switch (the reference always streams). */
int slc_set_timeslice_streaming(slc_ctx *ctx, int on);

/* Gauge links of displaced operators (GaugeField, src/GaugeField.cpp): one
 * timeslice t of Lx Ly Lz sites (site = (x Ly + y) Lz + z) x 3 directions
 * (x, y, z) x one 3 x 3 SU(3) matrix, column-major (the element order of
 * Eigen::Matrix3cd; GaugeField::tslices[t][site][dir]).  Links are handed
 * over here, read from lime files (slc_read_gauge_field) or generated
 * (slc_synthetic_gauge). */
int slc_set_gauge_links(slc_ctx *ctx, int t, const double *U);
int slc_get_gauge_links(slc_ctx *ctx, int t, double *U);
/* GaugeField::read_gauge_field (src/GaugeField.cpp:762-779) with
 * read_lime_gauge_field_doubleprec_timeslices (:781-928) and
 * map_timeslice_to_eigen (:111-149): timeslices slice_i..slice_f of the lime
 * file <config_path>/conf.%04d (config_path NULL or "" = path_config of the
 * input file; a relative path is taken below base_dir, NULL = ".") into the
 * links above; U_mu(x, y, z, t) of the file's ILDG binary record goes to
 * site (x Ly + y) Lz + z, direction mu, as is.  For Lx = Ly = Lz (the lattices
 * the reference reads correctly: for other shapes its reader and timeslice
 * map index sites differently) this is the reference's tslices.  Errors as
 * the reference's: "Could not open file ...", "no ildg-binary-data record
 * found in file ...", "Probably wrong lattice size or precision ...",
 * "Single precision read!"; beyond it, an ildg-format record whose lattice
 * differs from the input file, and a scidac-checksum record that does not
 * match the data, are errors too (the links are then marked not set).
 * slc_create_operators applies the input file's HYP smearing to a copy. */
int slc_read_gauge_field(slc_ctx *ctx, const char *base_dir, const char *config_path, int config, int slice_i,
                         int slice_f);
/* GaugeField::plaque_ts (src/GaugeField.cpp:742-760) of the links of
 * timeslice t: the mean over sites and ordered pairs mu != nu of the spatial
 * directions of Re tr(U_mu(x) U_nu(x+mu) U_mu(x+nu)^+ U_nu(x)^+) / 3. */
int slc_plaquette(slc_ctx *ctx, int t, double *plaquette);
/* The reader of slc_read_gauge_field into host memory, without a context or
 * a device: timeslices slice_i..slice_f of the lime file `path` on an
 * Lx x Ly x Lz x Lt lattice into U, (slice_f - slice_i + 1) x 27 Lx Ly Lz
 * complex numbers in the layout of slc_set_gauge_links, timeslice after
 * timeslice (GaugeField::read_lime_gauge_field_doubleprec_timeslices with
 * map_timeslice_to_eigen, src/GaugeField.cpp:781-928, :111-149).  The
 * message of a failure is slc_last_error(NULL). */
int slc_read_lime_gauge(const char *path, int Lx, int Ly, int Lz, int Lt, int slice_i, int slice_f, double *U);
/* Seeded SU(3) links exp(i eps H), H traceless Hermitian from complex
 * Gaussian entries, on every timeslice (no reference counterpart: synthetic
 * gauge configurations in place of the lime files). */
int slc_synthetic_gauge(slc_ctx *ctx, unsigned long long seed, double eps);
/* `iterations` steps of GaugeField::smearing_stout (GaugeField.cpp:228-274)
 * on every timeslice: three-dimensional staples, Omega = (rho / 4) C U^+,
 * U' = exp(-q) U with q the traceless anti-Hermitian part of Omega^+ - Omega
 * over 2 (the reference's factor 1/4 on rho kept).  The negative staple is
 * U_nu(x-nu)^+ U_mu(x-nu) U_nu(x-nu+mu), the gauge-covariant form; the
 * reference's smearing_stout multiplies U_mu(x-nu+mu) there (lines 255-256),
 * which breaks gauge covariance (its HYP staples, lines 358-359, use U_nu). */
int slc_smear_stout(slc_ctx *ctx, double rho, int iterations);
/* `iterations` HYP steps of every timeslice's links with the reference's
 * parameters (GaugeField::smearing_hyp, src/GaugeField.cpp:322-431: alpha2 the
 * decorated level, alpha1 the outer level, each level projected back to SU(3)
 * by proj_to_su3_imp).  slc_create_operators applies the input file's
 * HypSmearing parameters to a copy of the links by itself
 * (OperatorFactory.cpp:249-260), so this entry is for callers that smear
 * links they keep. */
int slc_smear_hyp(slc_ctx *ctx, double alpha1, double alpha2, int iterations);

/* Copy one resident input back to the host in the layouts above:
 * kind 0 = eigenvectors of timeslice `index`, 1 = perambulator `index`,
 * 2 = random vector `index`.  No reference counterpart (the reference reads
 * its inputs from files only); used to check device-generated inputs. */
int slc_get_input(slc_ctx *ctx, int kind, int index, double *out);

/* OperatorFactory::create_operators(filename, rnd_vec, config)
 * (src/OperatorFactory.cpp:220-326, 507-520) for the configuration of the
 * last slc_read_inputs (0 before any): VdaggerV(p, t) =
 * V(t)^+ diag(exp(-i p.x)) V(t) for every lookup entry, on the GPU, as the
 * infile's handling_vdaggerv says: "build"; "write" (build, then store every
 * non-trivial operator, :303-313); "read" (load them, :328-394); "liuming"
 * (load Liuming's files, :396-493).  A missing file is an error (the
 * reference constructs that runtime_error without throwing it). */
int slc_create_operators(slc_ctx *ctx);
/* The same for an explicit configuration number (names the VdaggerV files). */
int slc_create_operators_config(slc_ctx *ctx, int config);

/* Host only: the VdaggerV file names of the reference, kind 0 = written
 * ("/<path>/cnfg%04d/operators.%04d.p_<pxpypz>.d_<disp|000>.t_%03d",
 * OperatorFactory.cpp:224-226, 306-311), 1 = read
 * ("/<path>/cnfg%04d/operators.%04d.p_<pxpypz>_.t_%03d", :331-333, 352-356),
 * 2 = Liuming ("/<path>/VdaggerV.p<px>p<py>p<pz>.conf%04d", :399-427; the
 * reader tries -p first).  displacement: pairs of characters ("<x>y") or "".
 * Returns the bytes needed including the NUL (writes at most len). */
int slc_vdaggerv_filename(int kind, const char *path, int config, const int *mom, const char *displacement, int t,
                          char *buf, int len);

/* OperatorFactory::return_vdaggerv (include/OperatorsForMesons.hpp:66-69):
 * copies the nev x nev matrix of lookup entry id at timeslice t to out. */
int slc_return_vdaggerv(slc_ctx *ctx, int id, int t, double *out);

/* DilutedFactorFactory<Q>::operator[](key).at({op}) (include/
 * DilutedFactorFactory.hpp:43-49, src/DilutedFactorFactory.cpp:33-253):
 * qtype 0/1/2 = Q0 {t} / Q1 {t, b} / Q2 {b1, t, b2}.  Writes the D x D
 * matrices (D = 4 dilE) of every random-vector pair to out (capacity `cap`
 * matrices), the pairs to ric (2 ints each) and returns their count in *n. */
int slc_diluted_factor(slc_ctx *ctx, int qtype, const int *key, int op, double *out, int *ric, int cap, int *n);

/* contract() (src/Correlators.cpp:40-127): every diagram of the infile over
 * all dilution block pairs (or only block pair `single_time_slice_combination`
 * when >= 0, global_data.cpp:84-87).  kahan != 0 sums like the reference's
 * SLAPH_KAHAN build (include/KahanAccumulator.hpp:14-78): Neumaier
 * compensation inside each slice pair (the mean of resolve_request1 and
 * right_sum / result of inner_product, src/Diagram.cpp:9-27,
 * src/DilutedTrace.cpp:5-82, in the reference's element order), plain sums
 * across slice pairs (the per-slice-pair accumulators merge sum and zero
 * compensation, Diagram.cpp:96-114).  That form evaluates every trace value
 * on its own and is much slower than the default. */
int slc_contract(slc_ctx *ctx, int single_time_slice_combination, int kahan);

/* inner_product(DilutedTraces, DilutedTraces[, DilutedTraces]) of
 * src/DilutedTrace.cpp:5-44 / :46-82 (ntraces = 2 / 3), or the mean of one
 * trace vector as in resolve_request1 (src/Diagram.cpp:9-27, ntraces = 1),
 * evaluated by the device resolve kernel of slc_contract.  values[k] holds
 * n[k] complex trace values, used[k] their used random-vector masks
 * (DilutedTrace::used_rnd_ids, uint16).  kahan != 0 sums like the SLAPH_KAHAN
 * build.  out: one complex value.  No admitted combination: the reference's
 * "has an empty result" runtime_error. */
int slc_inner_product(slc_ctx *ctx, int ntraces, const double *const *values, const uint16_t *const *used,
                      const int *n, int kahan, double *out);

/* AbstractDilutedTraceFactory::get(slice_pair, locations) of the trace
 * factory `trace` ("trQ1", "trQ1Q1", "trQ0Q2", "trQ1Q1Q1", "trQ1Q0Q2",
 * "trQ1Q1Q1Q1", "trQ2Q0Q2Q0", "trQ2Q0Q2Q0Q2Q0"; include/DilutedTraceFactory.hpp:21-29,
 * src/DilutedTraceFactory.cpp:13-238) at the slice pair (t_source, t_sink),
 * each of its nloc trace times at the source (0) or the sink (1).  Fills the
 * trace vector of every trace id in trace_indices_map order: counts[id]
 * values, written one after another to values (complex) and used (their
 * DilutedTrace::used_rnd_ids masks), in the reference's element order
 * (factor_to_trace, src/DilutedTrace.cpp:84-165).  *n_ids and *total receive
 * the number of trace ids and of values; when count_cap or cap is too small
 * nothing is written and SLC_ERR_ARG is returned (call again with larger
 * buffers).  Computed on the device, like slc_contract. */
int slc_diluted_trace(slc_ctx *ctx, const char *trace, int t_source, int t_sink, const int *locations, int nloc,
                      int *counts, int count_cap, double *values, uint16_t *used, int cap, int *n_ids, int *total);

/* Results of the last slc_contract, per diagram (Diagram::write,
 * src/Diagram.cpp:132-147): names, counts, and n_corr x Lt complex values in
 * the order of the reference's correlator requests. */
int slc_diagram_count(const slc_ctx *ctx, int *n);
int slc_diagram_name(const slc_ctx *ctx, int i, char *buf, int len);
int slc_correlator_count(const slc_ctx *ctx, const char *diagram, int *n);
int slc_correlator_name(const slc_ctx *ctx, const char *diagram, int i, char *buf, int len);
int slc_fetch_correlators(const slc_ctx *ctx, const char *diagram, double *out);

/* Stage timings of the last create_operators / contract (milliseconds, HIP
 * events): what = "vdaggerv", "vdaggerv_kernel", "factors", "contract". */
int slc_stage_time(const slc_ctx *ctx, const char *what, double *ms);

/* Launch statistics of one kernel in the last contract or create_operators
 * (no reference counterpart; for roofline measurements): kernel =
 * "k_trace_sym" (crossed traces of the summed-product form as symmetric
 * products), "k_trace_lstat_q" / "k_trace_lstat" (non-separable traces on
 * the left-stationary kernel) or "k_vdv_ft" (VdaggerV, sitewise Fourier
 * form); what = "ms" (device time summed over its
 * launches, HIP events on the engine stream), "launches", "bytes"
 * (algorithmic bytes summed over the launches: each operand matrix once,
 * every output written once) or "flops" (operations of the kernel's own
 * algorithm: for k_vdv_ft 24 per (element pair, site) for V^H V over three
 * colours plus 5 per |pz| value of the pass for the z sums). */
int slc_kernel_stat(const slc_ctx *ctx, const char *kernel, const char *what, double *value);

/* Upper bound on the device memory the contraction may use for per-chunk
 * products and traces (bytes); sets the number of slice pairs per launch.
 * 0 (the default) = 80 % of the device memory free when contract starts;
 * negative values are rejected. */
int slc_set_chunk_budget(slc_ctx *ctx, long long bytes);

/* Precision forms beside the plain FP64 path (which the reference's
 * complex<double> arithmetic, OperatorFactory.cpp:294-300 and
 * DilutedFactorFactory.cpp:33-253, defines; no reference counterpart):
 *   vdaggerv: 0 = FP64 MFMA (default), 1 = single-pass FP32 MFMA (V and the
 *             phases rounded to FP32, FP32 accumulation), 2 = three-term
 *             split BF16 on the BF16 MFMA (six products per real product,
 *             FP32 accumulation), 3 = Ozaki int8: columns scaled to |x| < 1
 *             and cut into 7 slices of 7 bits, slice pairs i + j <= 8
 *             multiplied exactly on the int8 MFMA, combined in FP64;
 *   factors:  0 = FP64 (default), 1 = diluted factors and product tables
 *             stored in FP32, 2 = stored in a 48-bit form (each FP64 part
 *             rounded to its upper 48 bits: 36 mantissa bits, relative
 *             rounding error <= 2^-37); built, multiplied and traced in FP64.
 * Correlator accumulation is FP64 (optionally Kahan) in every form. */
int slc_set_precision(slc_ctx *ctx, int vdaggerv, int factors);

/* Autotuner (no reference counterpart): times slc_contract (and, when the
 * operators are built here, slc_create_operators for the VdaggerV form) on
 * the context's current inputs for each candidate of each schedule / launch
 * shape in turn -- the persistent grid of the left-stationary traces
 * (workgroups per CU), the K split of the GEMM-form traces, MFMA or generic
 * D = 16 products, Moebius or direct masked inner products, summed or
 * per-combination products of the crossed traces (crossed_direct), their
 * traces as symmetric products or on the left-stationary kernel
 * (crossed_lstat), Fourier or direct VdaggerV -- `repeats` timed calls each,
 * keeping a candidate only when
 * it is faster by more than 0.5 %.  The choice is kept in the context for
 * this input file (a later slc_configure with the same text reuses it); the
 * correlators stay equal up to rounding.  Needs inputs and operators. */
int slc_autotune(slc_ctx *ctx, int repeats);
/* JSON report of the last slc_autotune (every candidate's time in ms and the
 * choice); copies at most len bytes, returns the bytes needed incl. NUL. */
int slc_autotune_report(const slc_ctx *ctx, char *buf, int len);
/* The current choice as text "key=value ..." (slc_get_tuning, same copy
 * convention), and setting it from such text (unknown keys are an error);
 * lets a caller persist a tuning across processes. */
int slc_get_tuning(const slc_ctx *ctx, char *buf, int len);
int slc_set_tuning(slc_ctx *ctx, const char *text);

/* HIP graphs for slc_contract (no reference counterpart; on by default):
 * with on != 0 the second slc_contract of a plan captures the launch
 * sequence of one configuration (every chunk's products and traces, the
 * resolves, accumulations and the finalisation; kernels only) into a HIP
 * graph and later calls replay it.  The statistics of slc_kernel_stat for the
 * trace kernels are then those of the last call that launched directly.
 * Changing the setting drops the plan. */
int slc_set_graphs(slc_ctx *ctx, int on);

/* C = A^H diag(phase[k/3]) B on the device (batch of `batch` products with
 * element strides sA, sB, sC): the kernel behind VdaggerV
 * (OperatorFactory.cpp:294-300).  A, B, C, phase are device pointers; phase
 * may be NULL. */
int slc_ahb_device(slc_ctx *ctx, const double *A, const double *B, double *C, const double *phase, int M, int N,
                   int K, long long lda, long long ldb, long long ldc, int batch, long long sA, long long sB,
                   long long sC);

/* slc_ahb_device with no allocation.  No counterpart in the reference: it is
 * for callers that cannot allocate device memory (code under HIP graph
 * capture, serving loops, harnesses that forbid hipMalloc).  Same arguments,
 * result and kernel as slc_ahb_device, whose per-call job table (one entry
 * per product of the batch) is allocated with hipMalloc; here it lives in a
 * fixed device array of 1024 entries (64 KiB) that each GPU holds until the
 * process ends.  batch above 1024 is rejected with SLC_ERR_ARG.  The array is
 * shared by every context on a GPU and is overwritten by each call: calls
 * that overlap on one GPU race.  The call returns when the product is done. */
int slc_ahb_device_noalloc(slc_ctx *ctx, const double *A, const double *B, double *C, const double *phase, int M,
                           int N, int K, long long lda, long long ldb, long long ldc, int batch, long long sA,
                           long long sB, long long sC);

/* The same product on host buffers (A: K x M, B: K x N, C: M x N, phase of
 * length ceil(K/3) or NULL); copies through the device.  For checks. */
int slc_ahb_host(slc_ctx *ctx, const double *A, const double *B, double *C, const double *phase, int M, int N,
                 int K);

/* Host-only (no device needed): parse an infile and dump the lookup tables
 * init_lookup_tables builds (src/init_lookup_tables.cpp:470-556) as text --
 * VdaggerV lookup, Q0/Q1/Q2 DilutedFactorIndex lists, trace index lists and
 * every correlator request with its dataset name.  Returns the bytes needed
 * including the NUL (writes at most len), or a negative code. */
int slc_lookup_dump(const char *infile_text, char *buf, int len);

/* Dilution scheme text of test_dilution_scheme (src/dilution-iterator.cpp:17-36),
 * block (interlace = 0) or interlace; returns the number of bytes needed
 * including the terminating NUL (writes at most len). */
int slc_dilution_text(int num_slice, int num_block, int interlace, char *buf, int len);

#ifdef __cplusplus
}
#endif

#endif /* SLAPH_AMD_H */
