// Reads a result file written by the reference (numbers only) and prints
// the same numbers through the host API's writers (emcgpuOutput.hpp), so a
// test can compare the two texts byte for byte.
//
// usage: output_format grid <file>
//        output_format current <file> <contacts> <stepTime> <startTime>
//        output_format particles <file>
//        output_format bulkparticles <file>
//        output_format timeseries <file> <dt>
#include <cstdlib>
#include <fstream>
#include <iostream>
#include <sstream>
#include <string>
#include <vector>

#include "emcgpu/emcgpuOutput.hpp"

using namespace emcgpu;

static std::vector<std::vector<std::string>> rows(const char *path) {
  std::ifstream f(path);
  std::vector<std::vector<std::string>> r;
  std::string line;
  while (std::getline(f, line)) {
    std::istringstream is(line);
    std::vector<std::string> t;
    std::string w;
    while (is >> w)
      t.push_back(w);
    if (!t.empty())
      r.push_back(t);
  }
  return r;
}

int main(int argc, char **argv) {
  if (argc < 3)
    return 2;
  const std::string mode = argv[1];
  auto r = rows(argv[2]);
  if (mode == "grid") {
    const size_t nx = std::stoul(r[0][0]), ny = std::stoul(r[0][1]);
    std::vector<double> v;
    for (size_t j = 1; j < r.size(); j++)
      for (auto &w : r[j])
        v.push_back(std::strtod(w.c_str(), nullptr));
    if (v.size() != nx * ny)
      return 3;
    output::printGrid(std::cout, nx, ny, v.data());
  } else if (mode == "current") {
    const int nc = std::atoi(argv[3]);
    const double stepTime = std::strtod(argv[4], nullptr), startTime = std::strtod(argv[5], nullptr);
    std::vector<std::vector<int>> netto;
    std::vector<std::vector<double>> cur;
    for (auto &t : r) {
      std::vector<int> n;
      std::vector<double> c;
      for (int k = 0; k < nc; k++) {
        n.push_back(std::atoi(t[1 + k].c_str()));
        c.push_back(std::strtod(t[1 + nc + k].c_str(), nullptr));
      }
      netto.push_back(n);
      cur.push_back(c);
    }
    output::printCurrent(std::cout, netto, cur, stepTime, startTime);
  } else if (mode == "particles") {
    const double maxPos[2] = {std::strtod(r[0][0].c_str(), nullptr),
                              std::strtod(r[0][1].c_str(), nullptr)};
    const size_t n = r.size() - 1;
    std::vector<double> pos(2 * n), k(3 * n), E(n), tau(n);
    std::vector<int32_t> sub(n), valley(n);
    for (size_t i = 0; i < n; i++) {
      const auto &t = r[i + 1];
      if (std::stoul(t[0]) != i)
        return 4;
      auto d = [&](int j) { return std::strtod(t[j].c_str(), nullptr); };
      pos[2 * i] = d(1);
      pos[2 * i + 1] = d(2);
      k[3 * i] = d(3);
      k[3 * i + 1] = d(4);
      k[3 * i + 2] = d(5);
      E[i] = d(6);
      sub[i] = std::atoi(t[7].c_str());
      valley[i] = std::atoi(t[8].c_str());
      tau[i] = d(9);
    }
    output::printParticles(std::cout, maxPos, n, pos.data(), k.data(), E.data(), sub.data(),
                           valley.data(), tau.data());
  } else if (mode == "bulkparticles") {
    double maxPos[3];
    for (int d = 0; d < 3; d++)
      maxPos[d] = std::strtod(r[0][d].c_str(), nullptr);
    const size_t n = r.size() - 1;
    std::vector<double> pos(3 * n), k(3 * n), E(n);
    std::vector<int32_t> sub(n), valley(n);
    for (size_t i = 0; i < n; i++) {
      const auto &t = r[i + 1];
      if (std::stoul(t[0]) != i)
        return 4;
      for (int d = 0; d < 3; d++) {
        pos[3 * i + d] = std::strtod(t[1 + d].c_str(), nullptr);
        k[3 * i + d] = std::strtod(t[4 + d].c_str(), nullptr);
      }
      E[i] = std::strtod(t[7].c_str(), nullptr);
      sub[i] = std::atoi(t[8].c_str());
      valley[i] = std::atoi(t[9].c_str());
    }
    output::printBulkParticles(std::cout, maxPos, n, pos.data(), k.data(), E.data(), sub.data(),
                               valley.data());
  } else if (mode == "timeseries") {
    const double dt = std::strtod(argv[3], nullptr);
    std::vector<std::vector<double>> rowsV;
    for (size_t s = 0; s < r.size(); s++) {
      std::vector<double> v;
      for (size_t j = 1; j < r[s].size(); j++)
        v.push_back(std::strtod(r[s][j].c_str(), nullptr));
      rowsV.push_back(v);
    }
    output::printTimeSeries(std::cout, dt, rowsV);
  } else {
    return 2;
  }
  return 0;
}
