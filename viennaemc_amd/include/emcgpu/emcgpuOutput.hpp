// emcgpuOutput.hpp -- result files of device simulations in the reference's
// text formats, so that the reference's plotting helpers
// (helper/emcPlottingFiles/emcPlottingFiles/readResultFile.py) read them
// unchanged:
//   * grids (include/emcOutput.hpp:14-50, emcGrid::print,
//     include/emcGrid.hpp:143-165): the extent on the first line, then one
//     line per row of constant y with the values of increasing x separated
//     by single spaces;
//   * contact currents (include/emcOutput.hpp:54-72): one line per
//     non-transient step, "time netto_0 .. netto_{C-1} current_0 ..";
//   * particle dumps (emcBasicParticleHandler::print,
//     include/ParticleHandler/emcBasicParticleHandler.hpp:191-215): the
//     device extent, then "index x y kx ky kz energy subvalley valley tau"
//     per particle, lines separated (not terminated) by newlines;
//   * bulk particle dumps (basicBulkParticleHandler::print,
//     examples/bulkSimulation/basicBulkParticleHandler.hpp:227-248): the
//     box extent (three numbers), then "index x y z kx ky kz energy
//     subvalley valley" per particle, lines separated by newlines;
//   * per-step time series of the bulk example
//     (examples/bulkSimulation/bulkSimulation.cpp:171-183): "time v_0 ..
//     v_{V-1}" per step, time = step * dt, every line terminated.
// Numbers use the default std::ostream formatting (6 significant digits),
// as the reference writes them.
#ifndef EMCGPU_OUTPUT_HPP
#define EMCGPU_OUTPUT_HPP

#include <cstddef>
#include <cstdint>
#include <fstream>
#include <functional>
#include <ostream>
#include <stdexcept>
#include <string>
#include <vector>

namespace emcgpu {
namespace output {

/// a grid of nx * ny values (x fastest), each passed through f
inline void printGrid(std::ostream &os, std::size_t nx, std::size_t ny, const double *values,
                      const std::function<double(double)> &f = nullptr) {
  os << nx << " " << ny << "\n";
  for (std::size_t j = 0; j < ny; j++) {
    if (j > 0)
      os << "\n";
    for (std::size_t i = 0; i < nx; i++) {
      const double v = values[j * nx + i];
      os << (f ? f(v) : v);
      if (i + 1 < nx)
        os << " ";
    }
  }
  os << "\n";
}

/// contact currents: netto[s][c] particles that left (minus injected) and
/// current[s][c] [A] of non-transient step s, time startTime + (s+1) stepTime
inline void printCurrent(std::ostream &os, const std::vector<std::vector<int>> &netto,
                         const std::vector<std::vector<double>> &current, double stepTime,
                         double startTime) {
  for (std::size_t s = 1; s <= current.size(); s++) {
    os << startTime + (double)s * stepTime;
    for (int n : netto[s - 1])
      os << " " << n;
    for (double c : current[s - 1])
      os << " " << c;
    os << "\n";
  }
}

/// particle dump: pos[n][2] [m], k[n][3] [1/m], energy [eV], tau [s]
inline void printParticles(std::ostream &os, const double maxPos[2], std::size_t n,
                           const double *pos, const double *k, const double *energy,
                           const int32_t *sub, const int32_t *valley, const double *tau) {
  os << maxPos[0] << " " << maxPos[1] << "\n";
  for (std::size_t i = 0; i < n; i++) {
    os << i << " " << pos[2 * i] << " " << pos[2 * i + 1] << " " << k[3 * i] << " "
       << k[3 * i + 1] << " " << k[3 * i + 2] << " " << energy[i] << " " << sub[i] << " "
       << valley[i] << " " << tau[i];
    if (i + 1 < n)
      os << "\n";
  }
}

/// bulk particle dump: pos[n][3] [m], k[n][3] [1/m], energy [eV]
inline void printBulkParticles(std::ostream &os, const double maxPos[3], std::size_t n,
                               const double *pos, const double *k, const double *energy,
                               const int32_t *sub, const int32_t *valley) {
  os << maxPos[0] << " " << maxPos[1] << " " << maxPos[2] << "\n";
  for (std::size_t i = 0; i < n; i++) {
    os << i << " " << pos[3 * i] << " " << pos[3 * i + 1] << " " << pos[3 * i + 2] << " "
       << k[3 * i] << " " << k[3 * i + 1] << " " << k[3 * i + 2] << " " << energy[i] << " "
       << sub[i] << " " << valley[i];
    if (i + 1 < n)
      os << "\n";
  }
}

/// one line per step s of rows: "s * dt rows[s][0] rows[s][1] .."
template <class V>
inline void printTimeSeries(std::ostream &os, double dt, const std::vector<std::vector<V>> &rows) {
  for (std::size_t s = 0; s < rows.size(); s++) {
    os << (double)s * dt << " ";
    for (std::size_t j = 0; j < rows[s].size(); j++) {
      if (j > 0)
        os << " ";
      os << rows[s][j];
    }
    os << "\n";
  }
}

inline std::ofstream openResultFile(const std::string &name) {
  std::ofstream f(name + ".txt");
  if (!f)
    throw std::runtime_error("cannot write " + name + ".txt");
  return f;
}

}  // namespace output
}  // namespace emcgpu

#endif
