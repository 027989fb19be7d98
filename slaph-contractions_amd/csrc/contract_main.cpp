// The `contract` program of the MI355X path: the per-configuration loop of the
// reference's main (main/contract.cpp:28-107) with read_parameters' command
// line (src/global_data.cpp:61-249), on the C ABI (include/slaph_amd.h) and
// the HDF5 correlator writer (include/slaph_h5.h).
//
//   contract -i test_all.ini --start_config 1000 --end_config 1000
//
// reads the infile, then for every configuration reads perambulators, random
// vectors and eigenvectors from the infile's paths (relative to the working
// directory), builds (or reads) VdaggerV, contracts every diagram on the GPU
// and writes <output_path>/<diagram>_cnfg%04d.h5.  Options given on the
// command line take precedence over the infile, as with boost::program_options
// in the reference.  --kahan selects the compensated sums that the reference
// selects at build time (-DSLAPH_KAHAN); --device picks the HIP device.
#include "../../include/slaph_amd.h"
#include "../../include/slaph_h5.h"

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iostream>
#include <sstream>
#include <string>
#include <vector>

namespace {

[[noreturn]] void die(std::string const &msg) {
  std::cerr << msg << std::endl;
  std::exit(1);
}

void check(slc_ctx *ctx, int rc, const char *what) {
  if (rc < 0) die(std::string(what) + ": " + slc_last_error(ctx));
}

void usage() {
  std::cout << "Allowed options:\n"
               "  -h [ --help ]                          produce help message\n"
               "  -v [ --version ]                       print version string\n"
               "  --verbose                              accepted for compatibility\n"
               "  -i [ --input ] arg (=LapHs.in)         name of input file.\n"
               "  -o [ --output ] arg (=LapHs.out)       accepted for compatibility\n"
               "  --single_time_slice_combination arg (=-1)\n"
               "                                         Only compute this time slice combination\n"
               "  --start_config arg                     First configuration\n"
               "  --end_config arg                       Last configuration\n"
               "  --delta_config arg                     Stepsize between two configurations\n"
               "  --kahan                                Kahan-compensated sums (SLAPH_KAHAN)\n"
               "  --device arg (=0)                      HIP device\n";
}

int to_int(std::string const &opt, std::string const &v) {
  char *end = nullptr;
  long x = std::strtol(v.c_str(), &end, 10);
  if (v.empty() || *end) die("invalid value '" + v + "' for option " + opt);
  return int(x);
}

}  // namespace

int main(int ac, char **av) {
  std::string input = "LapHs.in";
  int single = -1, device = 0, kahan = 0;
  bool has_start = false, has_end = false, has_delta = false;
  int start = -1, end = 0, delta = 1;
  for (int i = 1; i < ac; ++i) {
    std::string a = av[i], v;
    auto eq = a.find('=');
    if (a.rfind("--", 0) == 0 && eq != std::string::npos) {
      v = a.substr(eq + 1);
      a = a.substr(0, eq);
    }
    auto value = [&]() -> std::string {
      if (!v.empty()) return v;
      if (i + 1 >= ac) die("the required argument for option '" + a + "' is missing");
      return av[++i];
    };
    if (a == "-h" || a == "--help") {
      usage();
      return 0;
    } else if (a == "-v" || a == "--version") {
      std::cout << "Contraction code for LapHs perambulators (MI355X contraction path).\n";
      return 0;
    } else if (a == "--verbose") {
    } else if (a == "-i" || a == "--input") {
      input = value();
    } else if (a == "-o" || a == "--output") {
      (void)value();
    } else if (a == "--single_time_slice_combination") {
      single = to_int(a, value());
    } else if (a == "--start_config") {
      start = to_int(a, value());
      has_start = true;
    } else if (a == "--end_config") {
      end = to_int(a, value());
      has_end = true;
    } else if (a == "--delta_config") {
      delta = to_int(a, value());
      has_delta = true;
    } else if (a == "--kahan") {
      kahan = 1;
    } else if (a == "--device") {
      device = to_int(a, value());
    } else {
      die("unrecognised option '" + a + "'");
    }
  }

  std::cout << "This is sLapH-contractions, MI355X contraction path\n";
  std::ifstream ifs(input);
  if (!ifs) {
    std::cout << "CANNOT open input file: " << input << "\n";
    return 1;
  }
  std::stringstream buffer;
  buffer << ifs.rdbuf();
  const std::string text = buffer.str();
  std::cout << "\nConfiguration file read in is:\n\n" << text << "\n\n---- End of configuration file ----\n\n";

  slc_ctx *ctx = nullptr;
  if (slc_ctx_create(device, &ctx) != SLC_OK) die(std::string("cannot create a context: ") + slc_last_error(nullptr));
  check(ctx, slc_configure(ctx, text.c_str()), "configure");
  long long q = 0;
  if (!has_start) start = (check(ctx, slc_query(ctx, "start_config", &q), "query"), int(q));
  if (!has_end) end = (check(ctx, slc_query(ctx, "end_config", &q), "query"), int(q));
  if (!has_delta) delta = (check(ctx, slc_query(ctx, "delta_config", &q), "query"), int(q));
  if (delta <= 0) {
    std::cerr << "The 'delta_config' option has been set to a non-positive value. This makes no sense."
              << std::endl;
    std::abort();
  }
  const int need = slc_query_string(ctx, "output_path", nullptr, 0);
  check(ctx, need, "query");
  std::vector<char> opath(size_t(need) + 1, 0);
  check(ctx, slc_query_string(ctx, "output_path", opath.data(), need), "query");
  const std::string output_path(opath.data());

  long long need_gauge = 0, Lt_all = 0;
  check(ctx, slc_query(ctx, "need_gaugefield", &need_gauge), "query");
  check(ctx, slc_query(ctx, "Lt", &Lt_all), "query");
  const int hneed = slc_query_string(ctx, "handling_vdaggerv", nullptr, 0);
  check(ctx, hneed, "query");
  std::vector<char> hbuf(size_t(hneed) + 1, 0);
  check(ctx, slc_query_string(ctx, "handling_vdaggerv", hbuf.data(), hneed), "query");
  const std::string handling(hbuf.data());

  for (int cfg = start; cfg <= end; cfg += delta) {
    std::cout << "\nprocessing configuration: " << cfg << "\n\n" << std::endl;
    check(ctx, slc_read_inputs(ctx, ".", cfg), "reading inputs");
    // build_vdaggerv reads the gauge configuration when an operator is
    // displaced (OperatorFactory.cpp:241-261)
    if (need_gauge && (handling == "build" || handling == "write"))
      check(ctx, slc_read_gauge_field(ctx, ".", nullptr, cfg, 0, int(Lt_all) - 1), "reading the gauge field");
    check(ctx, slc_create_operators_config(ctx, cfg), "create_operators");
    check(ctx, slc_contract(ctx, single, kahan), "contract");
    char ending[64];
    slh5_filename_ending(cfg, ending, sizeof ending);
    int nd = 0;
    check(ctx, slc_diagram_count(ctx, &nd), "results");
    long long Lt = 0;
    check(ctx, slc_query(ctx, "Lt", &Lt), "query");
    for (int d = 0; d < nd; ++d) {  // Diagram::write (src/Diagram.cpp:132-147)
      char diagram[64];
      check(ctx, slc_diagram_name(ctx, d, diagram, sizeof diagram), "results");
      int nc = 0;
      check(ctx, slc_correlator_count(ctx, diagram, &nc), "results");
      std::vector<double> vals(size_t(nc) * Lt * 2);
      check(ctx, slc_fetch_correlators(ctx, diagram, vals.data()), "results");
      slh5_file *f = nullptr;
      if (slh5_open(output_path.c_str(), diagram, ending, &f) != 0) die(slh5_last_error());
      for (int i = 0; i < nc; ++i) {
        const int len = slc_correlator_name(ctx, diagram, i, nullptr, 0);
        check(ctx, len, "results");
        std::vector<char> name(size_t(len) + 1, 0);
        check(ctx, slc_correlator_name(ctx, diagram, i, name.data(), len), "results");
        if (slh5_write(f, name.data(), vals.data() + size_t(i) * Lt * 2, int(Lt)) < 0) die(slh5_last_error());
      }
      slh5_close(f);
    }
  }
  slc_ctx_destroy(ctx);
  return 0;
}
