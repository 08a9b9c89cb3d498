// hotPhononGpu.cpp -- the hot-phonon loop of the reference's hot-carrier
// drivers (examples/hotPhononGa2O3/hotPhononGa2O3.cpp:274-281,
// examples/hotCarrierMHP/hotCarrierMHP.cpp:662-703) written against the GPU
// host API: one Kane Gamma valley (m* 0.28, alpha 0.3), Froehlich absorption
// + emission (60 meV, eps 3.5 / 10) reading a 256-bin LO bath (dq 2e7 /m,
// tau_LO 1 ps, V 2.7e-23 m^3), 10 kV/cm along z, dt 1 fs.  With "screened"
// the mechanisms are the screened hot-phonon ones with the q-resolved rate
// and angle, and the plasma screening (eps 3.5) follows the carriers.  The
// GPU updates the bath, the screening and the tables after every step.
//
// usage: hotPhononGpu <particles> <steps> [fp32|fp64] [seed] [plain|screened]
// prints per step: step <s> <mean energy> <mean N_q> <q_s^2>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>

#include "emcgpu/emcgpuBulkParticleHandler.hpp"

using T = double;
using namespace emcgpu;

int main(int argc, char **argv) {
  if (argc < 3) {
    std::fprintf(stderr, "usage: %s <particles> <steps> [fp32|fp64] [seed] [plain|screened]\n",
                 argv[0]);
    return 2;
  }
  const SizeType n = std::strtoul(argv[1], nullptr, 10);
  const int steps = std::atoi(argv[2]);
  const bool fp64 = argc > 3 && std::strcmp(argv[3], "fp64") == 0;
  const unsigned long seed = argc > 4 ? std::strtoul(argv[4], nullptr, 10) : 1;
  const bool screened = argc > 5 && std::strcmp(argv[5], "screened") == 0;
  if (std::getenv("EMCGPU_THROW"))
    setErrorPolicy(ErrorPolicy::Throw);

  emcgpuBulkDevice<T> device({11.8, 2329., 1.45e16, 9040}, 300.);
  emcgpuParticleType<T> electrons(2000, 2.);
  electrons.addValley(
      std::make_unique<emcNonParabolicIsotropValley<T>>(0.28, electrons.getMass(), 1, 0.3));
  auto bath = std::make_shared<emcPhononBath<T>>(256, 2e7, 1e-12, 0.06, 300., 2.7e-23);
  std::shared_ptr<emcPlasmonScreening<T>> screening;
  if (screened) {
    screening = std::make_shared<emcPlasmonScreening<T>>(3.5, true);
    electrons.addScatterMechanism({0}, std::make_unique<emcScreenedHotPhononFroehlichAbsorption3D<T>>(
                                           0, 0.06, 0.28, 3.5, 10., bath, screening, true));
    electrons.addScatterMechanism({0}, std::make_unique<emcScreenedHotPhononFroehlichEmission3D<T>>(
                                           0, 0.06, 0.28, 3.5, 10., bath, screening, true));
  } else {
    electrons.addScatterMechanism(
        {0}, std::make_unique<emcHotPhononFroehlichAbsorption3D<T>>(0, 0.06, 0.28, 3.5, 10., bath));
    electrons.addScatterMechanism(
        {0}, std::make_unique<emcHotPhononFroehlichEmission3D<T>>(0, 0.06, 0.28, 3.5, 10., bath));
  }

  emcgpuBulkParticleHandler<T> handler(
      device, electrons, {0., 0., 1.}, 10e5, n, seed,
      fp64 ? emcgpuBulkParticleHandler<T>::FP64 : emcgpuBulkParticleHandler<T>::FP32);
  if (screened)
    handler.enablePlasmaScreening(2.7e-23);
  handler.generateInitialParticles();
  if (screened)
    handler.updateScreening();
  for (int s = 1; s <= steps; s++) {
    // moveParticles, then (on the GPU) bath->update, updateScreening and
    // reinitScatterTables
    handler.moveParticles(1e-15);
    std::printf("step %d %.17g %.17g %.17g\n", s, handler.getAvgEnergy(0)[0], bath->getMeanNq(),
                screened ? handler.getScreeningQ2() : 0.0);
  }
  return 0;
}
