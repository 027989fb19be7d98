// A caller of the reference's class interfaces (OperatorFactory,
// DilutedFactorFactory, DilutedTraceFactory, inner_product, contract,
// WriteHDF5Correlator) on the golden L4 lattice, through include/slaph_amd.hpp
// and include/slaph_h5.hpp.  Usage: class_layer_l4 <l4 dir> <output dir>.
// Writes raw complex128 dumps and the correlator files for the test to check.
#include "slaph_amd.hpp"
#include "slaph_h5.hpp"

#include <cstdio>
#include <fstream>
#include <iostream>
#include <sstream>

using namespace slaph_amd;

static void dump(std::string const &path, std::vector<Complex> const &v) {
  std::ofstream f(path, std::ios::binary);
  f.write(reinterpret_cast<const char *>(v.data()), std::streamsize(v.size() * sizeof(Complex)));
}

int main(int ac, char **av) {
  if (ac != 3) {
    std::cerr << "usage: class_layer_l4 <l4 dir> <output dir>\n";
    return 2;
  }
  const std::string l4 = av[1], out = av[2];
  std::ifstream in(l4 + "/test_all.ini");
  std::stringstream text;
  text << in.rdbuf();
  try {
    Context ctx(0);
    ctx.configure(text.str());
    ctx.read_inputs(l4, 1000);
    OperatorFactory ops(ctx, 4, 4, 4, 4, 16, 2);
    ops.create_operators(1000);
    dump(out + "/vdv_1_2.bin", ops.return_vdaggerv(1, 2).values);

    DilutedFactorFactory<DilutedFactorType::Q0> q0(ops);
    DilutedFactorFactory<DilutedFactorType::Q2> q2(ops);
    std::vector<Complex> fq0, fq2;
    std::ofstream ric(out + "/ric.txt");
    for (auto const &f : q0[{1}].at({0})) {
      fq0.insert(fq0.end(), f.data.values.begin(), f.data.values.end());
      ric << "Q0 " << int(f.ric.first) << " " << int(f.ric.second) << "\n";
    }
    for (auto const &f : q2[{1, 2, 0}].at({3})) {
      fq2.insert(fq2.end(), f.data.values.begin(), f.data.values.end());
      ric << "Q2 " << int(f.ric.first) << " " << int(f.ric.second) << "\n";
    }
    dump(out + "/q0_t1_op0.bin", fq0);
    dump(out + "/q2_1_2_0_op3.bin", fq2);

    DilutedTraceFactory tr(ctx, "trQ0Q2");
    std::unique_ptr<AbstractDilutedTraceFactory> base(new DilutedTraceFactory(ctx, "trQ0Q2"));
    auto const &m = tr.get(BlockIterator{0, 1}, {Location::source, Location::sink});
    auto const &m2 = base->get(BlockIterator{0, 1}, {Location::source, Location::sink});
    std::vector<Complex> t0;
    std::ofstream used(out + "/used.txt");
    for (auto const &t : m.at(0)) {
      t0.push_back(t.data);
      used << t.used_rnd_ids << "\n";
    }
    if (m2.at(0).size() != m.at(0).size()) return 3;
    dump(out + "/trQ0Q2_0_1_id0.bin", t0);
    DilutedTraces a{m.at(0), false}, b{m.at(1), false};
    dump(out + "/inner_0_1.bin", {inner_product(ctx, a, b), inner_product(ctx, a, b, true)});

    auto res = contract(ctx);
    const std::string ending = correlator_filename_ending(1000);
    for (auto const &d : res) {
      WriteHDF5Correlator file(out + "/correlators", d.first, ending);
      for (auto const &c : d.second) file.write(c.second, c.first);
    }
  } catch (std::exception const &e) {
    std::cerr << "error: " << e.what() << "\n";
    return 1;
  }
  return 0;
}
