// The screened polar model (one Kane Gamma valley, m* 0.28, alpha 0.3;
// screened Froehlich absorption + emission, 60 meV, eps 3.5 / 10, sharing
// one emcPlasmonScreening) built through the GPU host API: prints the raw
// rates on the 2000-level table grid (0..2 eV) at q_s^2 = 1e17 / m^2, the
// screening object's Debye update, and what a host-only emcgpu context
// holds after the tables and the screening are lowered.
#include <cstdio>
#include <memory>

#include "emcgpu/emcgpuBulkParticleHandler.hpp"

using T = double;
using namespace emcgpu;

int main() {
  setErrorPolicy(ErrorPolicy::Throw);
  emcgpuParticleType<T> e(2000, 2.);
  e.addValley(std::make_unique<emcNonParabolicIsotropValley<T>>(0.28, e.getMass(), 1, 0.3));
  auto scr = std::make_shared<emcPlasmonScreening<T>>(10., true);
  scr->setQs2(1e17);
  e.addScatterMechanism({0}, std::make_unique<emcScreenedFroehlichAbsorption3D<T>>(
                                 0, 0.06, 0.28, 3.5, 10., 300., scr));
  e.addScatterMechanism({0}, std::make_unique<emcScreenedFroehlichEmission3D<T>>(
                                 0, 0.06, 0.28, 3.5, 10., 300., scr));
  for (double r : e.rawRates())
    std::printf("rate %.17g\n", r);
  emcgpu_config cfg{-1, EMCGPU_FP64, 1};
  emcgpu_ctx *ctx = nullptr;
  check(emcgpu_create(&cfg, &ctx), nullptr, "create");
  auto vd = e.valleyDescriptors();
  auto md = e.mechanismDescriptors();
  check(emcgpu_set_model(ctx, 0, vd.data(), (int)vd.size(), md.data(), (int)md.size()), ctx, "model");
  emcgpu_screening_desc sd;
  if (!e.screeningDescriptor(sd))
    return 3;
  check(emcgpu_set_screening(ctx, 0, &sd), ctx, "screening");
  double qs2 = 0;
  check(emcgpu_get_screening(ctx, 0, &qs2), ctx, "get screening");
  std::printf("ctx_qs2 %.17g\n", qs2);
  std::printf("kinds %d %d\n", md[0].kind, md[1].kind);
  // the Debye update: 1e24 / m^3 at 600 K, then disabled
  scr->update(1e24, 600.);
  std::printf("debye %.17g %.17g\n", scr->getQs2(), scr->getScreeningLength());
  emcPlasmonScreening<T> off(10., false);
  off.update(1e24, 600.);
  off.setQs2(5.);
  std::printf("off %.17g\n", off.getQs2());
  emcgpu_destroy(ctx);
  return 0;
}
