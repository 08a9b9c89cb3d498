// bulkSimulationGpu.cpp -- the reference's bulk silicon example
// (examples/bulkSimulation/bulkSimulation.cpp + examples/SiliconFunctions.hpp)
// written against the GPU host API: same valley, mechanism and handler calls,
// with the particle step running on the MI355X through libemcgpu.so.
//
// Physics: Si X valleys (m_l = 0.916, m_t = 0.196, alpha = 0.5 / eV, three
// subvalley pairs), acoustic + zero-order and first-order f/g intervalley
// phonons (Vasileska parameters), 300 K, 10 kV/cm along -x, 1000 table
// levels up to 1 eV.
//
// usage: bulkSimulationGpu [particles] [steps] [fp32|fp64] [seed] [every] [prefix]
//   prints "step <s> <avgE> <avgVd> <occ>" every <every> steps, then the
//   final per-subvalley occupancy and the GPU time per step.  With a file
//   prefix the particles carry positions in the example's 0.5 um periodic
//   box and the example's files are written (bulkSimulation.cpp:127-128,
//   171-183): <prefix>ElectronsEq.txt and <prefix>ElectronsEnd.txt (print),
//   <prefix>AvgEnergy.txt, <prefix>AvgDriftVelocity.txt and
//   <prefix>valleyOccupation.txt (one line per step, step 0 the initial
//   ensemble).
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <memory>
#include <string>
#include <vector>

#include "emcgpu/emcgpuBulkParticleHandler.hpp"

using NumType = double;
using namespace emcgpu;

int main(int argc, char **argv) {
  const SizeType nParticles = argc > 1 ? std::strtoull(argv[1], nullptr, 10) : 100000;
  const int nSteps = argc > 2 ? std::atoi(argv[2]) : 1000;
  const bool fp32 = argc > 3 ? std::strcmp(argv[3], "fp64") != 0 : true;
  const unsigned long seed = argc > 4 ? std::strtoul(argv[4], nullptr, 10) : 20261202;
  const int every = argc > 5 ? std::atoi(argv[5]) : 100;
  const bool files = argc > 6;
  const std::string prefix = files ? argv[6] : "";
  if (std::getenv("EMCGPU_THROW"))
    setErrorPolicy(ErrorPolicy::Throw);

  // material of examples/SiliconFunctions.hpp:21-25, 300 K; the extent of
  // bulkSimulation.cpp:39 when positions are written
  const emcgpuBulkDevice<NumType>::Material silicon{11.8, 2329., 1.45e16, 9040};
  emcgpuBulkDevice<NumType> device =
      files ? emcgpuBulkDevice<NumType>(silicon, 300., {5e-7, 5e-7, 5e-7})
            : emcgpuBulkDevice<NumType>(silicon, 300.);

  // particle type: 1000 levels up to 1 eV (bulkSimulation.cpp:94)
  emcgpuParticleType<NumType> electrons(1000, 1.);

  // X valley (SiliconFunctions.hpp:62-73)
  auto xValley = std::make_unique<emcNonParabolicAnisotropValley<NumType>>(
      std::array<NumType, 3>{0.916, 0.196, 0.196}, electrons.getMass(), 3, 0.5);
  xValley->setSubValleyEllipseCoordSystem(0, {1, 0, 0}, {0, 1, 0}, {0, 0, 1});
  xValley->setSubValleyEllipseCoordSystem(1, {0, 1, 0}, {1, 0, 0}, {0, 0, 1});
  xValley->setSubValleyEllipseCoordSystem(2, {0, 0, 1}, {0, 1, 0}, {1, 0, 0});
  electrons.addValley(std::move(xValley));

  // mechanisms (SiliconFunctions.hpp:78-145)
  const std::map<SizeType, std::vector<SizeType>> finalG = {{0, {0}}, {1, {1}}, {2, {2}}};
  const std::map<SizeType, std::vector<SizeType>> finalF = {
      {0, {1, 1, 2, 2}}, {1, {0, 0, 2, 2}}, {2, {0, 0, 1, 1}}};
  electrons.addScatterMechanism({0}, std::make_unique<emcAcousticScatterMechanism<NumType>>(0, 9., device));
  electrons.addScatterMechanism({0}, std::make_unique<emcZeroOrderInterValleyAbsorptionScatterMechanism<NumType>>("F", 0, finalF, 5.23e10, 0.06, device));
  electrons.addScatterMechanism({0}, std::make_unique<emcZeroOrderInterValleyEmissionScatterMechanism<NumType>>("F", 0, finalF, 5.23e10, 0.06, device));
  electrons.addScatterMechanism({0}, std::make_unique<emcZeroOrderInterValleyAbsorptionScatterMechanism<NumType>>("G", 0, finalG, 5.23e10, 0.06, device));
  electrons.addScatterMechanism({0}, std::make_unique<emcZeroOrderInterValleyEmissionScatterMechanism<NumType>>("G", 0, finalG, 5.23e10, 0.06, device));
  electrons.addScatterMechanism({0}, std::make_unique<emcFirstOrderInterValleyAbsorptionScatterMechanism<NumType>>("F", 0, finalF, 2.5, 0.023, device));
  electrons.addScatterMechanism({0}, std::make_unique<emcFirstOrderInterValleyEmissionScatterMechanism<NumType>>("F", 0, finalF, 2.5, 0.023, device));
  electrons.addScatterMechanism({0}, std::make_unique<emcFirstOrderInterValleyAbsorptionScatterMechanism<NumType>>("G", 0, finalG, 4., 0.018, device));
  electrons.addScatterMechanism({0}, std::make_unique<emcFirstOrderInterValleyEmissionScatterMechanism<NumType>>("G", 0, finalG, 4., 0.018, device));

  // handler: field 10 kV/cm along -x (bulkSimulation.cpp:42-43)
  emcgpuBulkParticleHandler<NumType> handler(
      device, electrons, {-1, 0, 0}, 10e5, nParticles, seed,
      fp32 ? emcgpuBulkParticleHandler<NumType>::FP32 : emcgpuBulkParticleHandler<NumType>::FP64);
  handler.generateInitialParticles();
  std::vector<std::vector<NumType>> avgEnergy, avgDriftVel, valleyOcc;
  if (files) {
    handler.print(prefix, "Eq");
    avgEnergy.push_back(handler.getAvgEnergy(0));
    avgDriftVel.push_back(handler.getAvgDriftVelocity(0));
    valleyOcc.push_back(handler.getValleyOccupationProbability(0));
  }

  const NumType dt = 1e-15;
  auto t0 = std::chrono::steady_clock::now();
  for (int done = 0; done < nSteps; done += every) {
    const int k = std::min(every, nSteps - done);
    handler.moveParticles(dt, k);
    if (files)
      for (int s = 0; s < k; s++) {
        avgEnergy.push_back(handler.getAvgEnergyAt(s));
        avgDriftVel.push_back(handler.getAvgDriftVelocityAt(s));
        valleyOcc.push_back(handler.getValleyOccupationProbabilityAt(s));
      }
    std::printf("step %d %.17g %.17g %.17g\n", done + k, handler.getAvgEnergy(0)[0],
                handler.getAvgDriftVelocity(0)[0], handler.getValleyOccupationProbability(0)[0]);
  }
  auto t1 = std::chrono::steady_clock::now();
  if (files) {
    handler.print(prefix, "End");
    auto fe = output::openResultFile(prefix + "AvgEnergy");
    output::printTimeSeries(fe, dt, avgEnergy);
    auto fv = output::openResultFile(prefix + "AvgDriftVelocity");
    output::printTimeSeries(fv, dt, avgDriftVel);
    auto fo = output::openResultFile(prefix + "valleyOccupation");
    output::printTimeSeries(fo, dt, valleyOcc);
  }
  auto occ = handler.getSubValleyOccupationProbability(0);
  std::printf("subvalleys %.17g %.17g %.17g\n", occ[0], occ[1], occ[2]);
  const double sec = std::chrono::duration<double>(t1 - t0).count();
  std::printf("time %.6f s, %.3e particle-updates/s\n", sec, nParticles * (double)nSteps / sec);
  return 0;
}
