// resistor2DGpu.cpp -- a 2-D n+ n n+ resistor in the form of the reference's
// examples/resistor2D/resistor2D.cpp, written against the GPU host API:
// device, Poisson solver, PM scheme, simulation parameter and simulation are
// set up the same way, and execute() runs the steps on the MI355X through
// libemcgpu.so and writes the reference's result files.
//
// Physics: one parabolic valley (m* = 0.26), acoustic + zero-order
// intervalley phonons, 300 K, Si material of examples/SiliconFunctions.hpp.
// (resistor2D itself uses Si X valleys with ionized-impurity scattering,
// whose rates depend on the doping region.)  With "regions" the intervalley
// absorption is registered for the channel region only and the emission for
// the channel and the drain region, so each doping region gets its own
// (valley, region) tables.
//
// usage: resistor2DGpu <name prefix> [small|c4] [fp32|fp64] [seed] [regions]
//   small: 20 x 6 nm, 1 nm spacing, 1e24 / m^3 with 1e25 / m^3 contact
//          regions, 0.2 V; 30 fs in 1 fs steps, 10 fs transient, the last 8
//          steps averaged, Poisson every second step, channel current over
//          5..15 nm
//   c4:    200 x 50 nodes at 1 nm, 1e23 / m^3 with 1e25 / m^3 contact
//          regions of 10 columns, 0.5 V, about 2M particles; 2 ps in 0.1 fs
//          steps, 1 ps transient, the last 1000 steps averaged
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <memory>
#include <string>

#include "emcgpu/emcgpuSimulation.hpp"

using NumType = double;
using namespace emcgpu;
using DeviceType = emcgpuDevice2D<NumType>;

int main(int argc, char **argv) {
  if (argc < 2) {
    std::fprintf(stderr, "usage: %s <name prefix> [small|c4] [fp32|fp64] [seed]\n", argv[0]);
    return 2;
  }
  const std::string prefix = argv[1];
  const bool c4 = argc > 2 && std::strcmp(argv[2], "c4") == 0;
  const bool fp64 = argc > 3 && std::strcmp(argv[3], "fp64") == 0;
  const unsigned long seed = argc > 4 ? std::strtoul(argv[4], nullptr, 10) : 20261221;
  const bool regions = argc > 5 && std::strcmp(argv[5], "regions") == 0;
  if (std::getenv("EMCGPU_THROW"))
    setErrorPolicy(ErrorPolicy::Throw);

  // material of examples/SiliconFunctions.hpp:21-25
  const DeviceType::Material silicon{11.8, 2329., 1.45e16, 9040};
  const NumType h = 1e-9;
  const int nx = c4 ? 200 : 21, ny = c4 ? 50 : 7, contactCols = c4 ? 10 : 4;
  const DeviceType::ValueVec maxPos = {(nx - 1) * h, (ny - 1) * h};
  DeviceType device{silicon, maxPos, {h, h}, 300.};
  const NumType nLow = c4 ? 1e23 : 1e24, nHigh = 1e25, bias = c4 ? 0.5 : 0.2;
  const NumType xc = (contactCols - 1) * h;
  if (c4) {
    // width that gives about 2M particles: sum of doping * cell area, with
    // half weights on the boundary nodes
    double w = 0;
    for (int iy = 0; iy < ny; iy++)
      for (int ix = 0; ix < nx; ix++) {
        double d = (ix < contactCols || ix >= nx - contactCols) ? nHigh : nLow;
        if (ix == 0 || ix == nx - 1)
          d *= 0.5;
        if (iy == 0 || iy == ny - 1)
          d *= 0.5;
        w += d * h * h;
      }
    device.setDeviceWidth(2e6 / w);
  } else {
    device.setDeviceWidth(1e-6);
  }
  device.addConstantDopingRegion({0, 0}, maxPos, nLow);
  device.addConstantDopingRegion({0, 0}, {xc, maxPos[1]}, nHigh);
  device.addConstantDopingRegion({maxPos[0] - xc, 0}, maxPos, nHigh);
  device.addOhmicContact(emcBoundaryPos::XMIN, bias, {0}, {maxPos[1]});
  device.addOhmicContact(emcBoundaryPos::XMAX, 0., {0}, {maxPos[1]});

  emcgpuPoissonSolver<NumType> solver(device, 1e-10, 50);
  emcgpuCICScheme<NumType> pmScheme;

  emcgpuSimulationParameter<NumType> param;
  if (c4)
    param.setTimes(2e-12, 1e-16, 1e-12);
  else
    param.setTimes(30e-15, 1e-15, 10e-15);
  param.setSeed(seed);
  param.setNamePrefix(prefix);
  param.setNrStepsBetweenShowProgress(c4 ? 5000 : 1000);
  param.setNrStepsForFinalAvg(c4 ? 1000 : 8);
  param.setPrecision(fp64 ? emcgpuSimulationParameter<NumType>::FP64
                          : emcgpuSimulationParameter<NumType>::FP32);

  // emcElectron(1000, 1.5, false): initial carriers from the doping
  auto electrons = std::make_unique<emcgpuParticleType<NumType>>(1000, 1.5);
  electrons->setUsePotentialForInit(false);
  electrons->addValley(
      std::make_unique<emcParabolicIsotropValley<NumType>>(0.26, electrons->getMass(), 1));
  const std::map<SizeType, std::vector<SizeType>> self = {{0, {0}}};
  const std::vector<int> all = {0, 1, 2};
  electrons->addScatterMechanism(
      all, std::make_unique<emcAcousticScatterMechanism<NumType>>(0, 9., device));
  electrons->addScatterMechanism(
      regions ? std::vector<int>{0} : all,
      std::make_unique<emcZeroOrderInterValleyAbsorptionScatterMechanism<NumType>>(
          "P", 0, self, 5e10, 0.05, device));
  electrons->addScatterMechanism(
      regions ? std::vector<int>{0, 2} : all,
      std::make_unique<emcZeroOrderInterValleyEmissionScatterMechanism<NumType>>(
          "P", 0, self, 5e10, 0.05, device));
  param.addParticleType(std::move(electrons));

  emcgpuSimulation<NumType, emcgpuCICScheme<NumType>> simulation(param, device, solver, pmScheme);
  if (!c4)
    simulation.setPoissonInterval(2);
  simulation.setChannelCurrentRegion(c4 ? 50e-9 : 5e-9, c4 ? 150e-9 : 15e-9,
                                     c4 ? 100e-9 : 10e-9);

  auto start = std::chrono::steady_clock::now();
  simulation.execute();
  auto end = std::chrono::steady_clock::now();
  const auto &cur = simulation.getCurrent();
  if (!cur.empty())
    std::printf("current %.9g %.9g\n", cur.back()[0], cur.back()[1]);
  std::printf("drift current %.9g\n", simulation.getAvgDriftCurrent());
  std::printf("GPU time: %.3f s\n", std::chrono::duration<double>(end - start).count());
  return 0;
}
