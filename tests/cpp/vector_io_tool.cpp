// Host-only driver of include/ntt/vector_io.hpp and include/ntt/measurement.hpp
// for the CPU tests, which check its output against Python big integers,
// numpy percentiles, json.loads and a Python restatement of the LCG stream.
//   vector_io_tool stage <file>             residues of every limb, one per line
//   vector_io_tool percentile <f> <v>...    linear-interpolated percentile
//   vector_io_tool json <text>              the escaped JSON string
//   vector_io_tool inputs <n> <q>...        benchmark operands: "left right" per line
//   vector_io_tool samples <name> <v>...    sample and summary lines
#include <cstdio>
#include <iostream>
#include <string>
#include <vector>

#include "ntt/measurement.hpp"
#include "ntt/vector_io.hpp"

int main(int argc, char **argv) {
  try {
    const std::string cmd = argc > 1 ? argv[1] : "";
    if (cmd == "stage" && argc == 3) {
      const auto f = ntt::read_vector_file(argv[2]);
      for (uint64_t v : ntt::stage_limbs(f)) std::printf("%llu\n", (unsigned long long)v);
      return 0;
    }
    if (cmd == "percentile" && argc >= 4) {
      std::vector<double> v;
      for (int i = 3; i < argc; ++i) v.push_back(std::stod(argv[i]));
      std::printf("%.17g\n", ntt::measurement::percentile(v, std::stod(argv[2])));
      return 0;
    }
    if (cmd == "json" && argc == 3) {
      ntt::measurement::json_string(std::cout, argv[2]);
      std::cout << "\n";
      return 0;
    }
    if (cmd == "inputs" && argc >= 4) {
      std::vector<uint64_t> q;
      for (int i = 3; i < argc; ++i) q.push_back(std::stoull(argv[i]));
      const auto [left, right] = ntt::measurement::benchmark_inputs(std::stoull(argv[2]), q);
      for (size_t i = 0; i < left.size(); ++i)
        std::printf("%llu %llu\n", (unsigned long long)left[i], (unsigned long long)right[i]);
      return 0;
    }
    if (cmd == "samples" && argc >= 4) {
      std::vector<double> v;
      for (int i = 3; i < argc; ++i) v.push_back(std::stod(argv[i]));
      ntt::measurement::write_samples(std::cout, argv[2], v);
      return 0;
    }
    std::cerr << "usage: vector_io_tool stage|percentile|json|inputs|samples ...\n";
    return 2;
  } catch (const std::exception &e) {
    std::cerr << "error: " << e.what() << "\n";
    return 1;
  }
}
