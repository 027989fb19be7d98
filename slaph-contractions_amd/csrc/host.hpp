// Host-side bookkeeping of the sLapH contraction path: infile grammar, lookup
// tables, dilution scheme, gamma tables and input file formats.  Everything
// here is integer/byte work that must match the reference bit for bit; the
// floating-point work lives in the HIP kernels.
#pragma once

#include <array>
#include <complex>
#include <cstdint>
#include <map>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

namespace slaph {

using cplx = std::complex<double>;
using Mom = std::array<int, 3>;
using Displacement = std::vector<std::pair<char, char>>;

// Sparse 4x4 Dirac matrix: column k has value[k] at row[k]
// (reference: include/Gamma.hpp:8-11, src/Gamma.cpp:18-246).
struct GammaEntry {
  int row[4];
  cplx value[4];
};
const std::array<GammaEntry, 16> &gamma_table();

// Quark line / dilution description (reference: global_data_typedefs.hpp:33-72).
struct Quark {
  std::string type;
  int nrnd = 0;
  std::string dilT_type;
  int dilT = 0;
  std::string dilE_type;
  int dilE = 0;
  std::string dilD_type;
  int dilD = 0;
  int id = 0;
  std::string path;
};

// One field operator: Dirac structure, displacement and momentum
// (reference: global_data_typedefs.hpp:82-89).
struct QuantumNumbers {
  std::vector<int> gamma;
  Displacement disp;
  Mom mom{{0, 0, 0}};
};
using Operators = std::vector<QuantumNumbers>;

struct CorrelatorSpec {  // reference: Correlators_2, global_data_typedefs.hpp:136-144
  std::string type;
  std::vector<int> quarks;
  std::vector<int> ops;
  std::vector<Mom> tot_mom;
};

struct VdaggerVQN {  // reference: VdaggerVQuantumNumbers, typedefs.hpp:44-55
  int id;
  Mom mom;
  Displacement disp;
};

// reference: DilutedFactorIndex, typedefs.hpp:105-121
struct FactorIndex {
  int id_vdaggerv = 0;
  bool dagger = false;
  std::vector<int> gamma;
  std::vector<std::pair<int, int>> rnd;
  bool operator==(FactorIndex const &o) const {
    return id_vdaggerv == o.id_vdaggerv && dagger == o.dagger && gamma == o.gamma && rnd == o.rnd;
  }
};

enum class Loc : uint8_t { source = 0, sink = 1 };

struct TraceRequest {  // reference: global_data.hpp:19-29
  std::string name;
  int tr_id = 0;
  std::vector<Loc> locs;
  bool operator==(TraceRequest const &o) const {
    return name == o.name && tr_id == o.tr_id && locs == o.locs;
  }
};

struct CorrelatorRequest {  // reference: global_data.hpp:31-38 (equality ignores the name)
  std::vector<TraceRequest> traces;
  std::string dataset;
  bool operator==(CorrelatorRequest const &o) const { return traces == o.traces; }
};

struct QuarklineSpec {  // reference: include/DiagramSpec.hpp:8-14
  const char *name;
  int q1, q2;
};
struct DiagramSpec {  // reference: include/DiagramSpec.hpp:25-28
  std::vector<int> source, sink;
  std::vector<std::vector<QuarklineSpec>> traces;
};
const std::map<std::string, DiagramSpec> &diagram_specs();

extern const char *const kTraceNames[8];

// reference: struct GlobalData (include/global_data.hpp:51-107)
struct GlobalData {
  GlobalData();
  int Lx = 0, Ly = 0, Lz = 0, Lt = 0;
  int nev = 0;
  int start_config = -1, end_config = 0, delta_config = 1;
  // defaults of read_parameters (reference: src/global_data.cpp:100-160)
  std::string path_eigenvectors = ".", name_eigenvectors = "eigenvector", path_output = "../../contractions",
              handling_vdaggerv = "build", path_vdaggerv, path_config;
  double hyp_alpha1 = 0, hyp_alpha2 = 0;  // HypPars (reference: global_data_typedefs.hpp)
  int hyp_iterations = 0;
  std::vector<Quark> quarks;
  std::vector<Operators> operator_list;
  std::vector<CorrelatorSpec> correlator_list;
  std::map<int, int> momentum_cutoff;

  std::vector<VdaggerVQN> vdaggerv_lookup;
  int index_of_unity = -1;
  bool need_gaugefield = false;
  std::map<std::string, std::vector<FactorIndex>> quarkline_lookuptable;
  std::map<std::string, std::vector<std::vector<int>>> trace_indices_map;
  std::map<std::string, std::vector<CorrelatorRequest>> correlator_requests_map;

  int n_rnd_total() const;
  int dilT() const { return quarks.at(0).dilT; }
  int dilE() const { return quarks.at(0).dilE; }
  int n_blocks() const { return Lt / dilT(); }
  long long peram_rows() const { return 4LL * nev * Lt; }
  long long peram_cols() const { return (long long)n_blocks() * dilE() * 4; }
};

// Infile grammar (reference: src/global_data_input_handling_utils.cpp:54-347,
// src/global_data.cpp:61-249).  Throws std::runtime_error on bad input.
std::vector<Mom> momentum_combinations(int p2);
Quark make_quark(std::string const &s);
Operators make_operator_list(std::string const &s);
CorrelatorSpec make_correlator(std::string const &s);
void parse_infile(GlobalData &gd, std::string const &text);

// reference: init_lookup_tables (src/init_lookup_tables.cpp:470-556)
void init_lookup_tables(GlobalData &gd);
// Text dump of the lookup tables (one line per entry; format fixed so that
// tests can compare it with an independent enumeration).
std::string lookup_dump(GlobalData const &gd);

// Dilution scheme, block type (reference: include/dilution-iterator.hpp:21-356).
struct SlicePair {
  int source, sink;
};
std::vector<std::pair<int, int>> block_pairs(int Lt, int block_size);
std::vector<SlicePair> slice_pairs(int bso, int bsi, int Lt, int block_size, bool interlace = false);
std::string dilution_iterator_text(int num_slice, int num_block, bool interlace);
int time_delta(int t_source, int t_sink, int Lt);  // Correlators.cpp:21-23

// Input files (reference: Perambulator.cpp:12-71, RandomVector.cpp:87-113,
// EigenVector.cpp:35-76, global_data_build_IO_names.cpp:7-85).
std::string perambulator_filename(Quark const &q, int config, int r, int Lt, int Vs);
std::string randomvector_filename(Quark const &q, int config, int r, int nev);
std::string eigenvector_prefix(GlobalData const &gd, int config);
// VdaggerV cache files (reference: OperatorFactory.cpp:224-226, 303-313,
// 331-364, 399-427).  The reference prefixes path_vdaggerv with "/" in all
// three names; written and read names differ (".d_<disp>.t_%03d" against
// "_.t_%03d").
std::string vdaggerv_write_name(std::string const &path, int config, Mom const &p, Displacement const &d, int t);
std::string vdaggerv_read_name(std::string const &path, int config, Mom const &p, int t);
std::string vdaggerv_liuming_name(std::string const &path, int config, Mom const &p);
// File layout -> column-major (rows 4N*t+N*d+v, cols dilE*dilD*b+dilE*d'+e).
void resort_perambulator(const cplx *file, cplx *out, int Lt, int nev, int dilT, int dilE, int dilD);
void read_binary(std::string const &path, void *dst, size_t bytes);
void write_binary(std::string const &path, const void *src, size_t bytes);
bool file_exists(std::string const &path);
void make_directories(std::string const &path);  // every missing directory of the path

}  // namespace slaph
