// Builds the reference's bulk-silicon particle type (examples/SiliconFunctions.hpp,
// Vasileska parameters) through the GPU host API, lowers it to a host-only
// emcgpu context (no GPU) and prints tau0 and the normalised tables, which the
// test compares with the reference's own (golden) values.  With a folder
// argument it also writes the scatter-table dumps of that type and of the
// device particle type (one parabolic valley, acoustic + zero-order
// intervalley phonons in doping regions 0, 1 and 2) into <folder>/si_ and
// <folder>/device_.
#include <cstdio>
#include <map>
#include <memory>
#include <string>
#include <vector>

#include "emcgpu/emcgpuBulkParticleHandler.hpp"

using T = double;
using namespace emcgpu;

int main(int argc, char **argv) {
  setErrorPolicy(ErrorPolicy::Throw);
  emcgpuBulkDevice<T> device({11.8, 2329., 1.45e16, 9040}, 300.);
  emcgpuParticleType<T> e(1000, 1.);
  auto x = std::make_unique<emcNonParabolicAnisotropValley<T>>(std::array<T, 3>{0.916, 0.196, 0.196},
                                                               e.getMass(), 3, 0.5);
  x->setSubValleyEllipseCoordSystem(0, {1, 0, 0}, {0, 1, 0}, {0, 0, 1});
  x->setSubValleyEllipseCoordSystem(1, {0, 1, 0}, {1, 0, 0}, {0, 0, 1});
  x->setSubValleyEllipseCoordSystem(2, {0, 0, 1}, {0, 1, 0}, {1, 0, 0});
  e.addValley(std::move(x));
  const std::map<SizeType, std::vector<SizeType>> g = {{0, {0}}, {1, {1}}, {2, {2}}};
  const std::map<SizeType, std::vector<SizeType>> f = {{0, {1, 1, 2, 2}}, {1, {0, 0, 2, 2}}, {2, {0, 0, 1, 1}}};
  e.addScatterMechanism({0}, std::make_unique<emcAcousticScatterMechanism<T>>(0, 9., device));
  e.addScatterMechanism({0}, std::make_unique<emcZeroOrderInterValleyAbsorptionScatterMechanism<T>>("F", 0, f, 5.23e10, 0.06, device));
  e.addScatterMechanism({0}, std::make_unique<emcZeroOrderInterValleyEmissionScatterMechanism<T>>("F", 0, f, 5.23e10, 0.06, device));
  e.addScatterMechanism({0}, std::make_unique<emcZeroOrderInterValleyAbsorptionScatterMechanism<T>>("G", 0, g, 5.23e10, 0.06, device));
  e.addScatterMechanism({0}, std::make_unique<emcZeroOrderInterValleyEmissionScatterMechanism<T>>("G", 0, g, 5.23e10, 0.06, device));
  e.addScatterMechanism({0}, std::make_unique<emcFirstOrderInterValleyAbsorptionScatterMechanism<T>>("F", 0, f, 2.5, 0.023, device));
  e.addScatterMechanism({0}, std::make_unique<emcFirstOrderInterValleyEmissionScatterMechanism<T>>("F", 0, f, 2.5, 0.023, device));
  e.addScatterMechanism({0}, std::make_unique<emcFirstOrderInterValleyAbsorptionScatterMechanism<T>>("G", 0, g, 4., 0.018, device));
  e.addScatterMechanism({0}, std::make_unique<emcFirstOrderInterValleyEmissionScatterMechanism<T>>("G", 0, g, 4., 0.018, device));

  emcgpu_config cfg{-1, EMCGPU_FP64, 1};
  emcgpu_ctx *ctx = nullptr;
  check(emcgpu_create(&cfg, &ctx), nullptr, "create");
  auto vd = e.valleyDescriptors();
  auto md = e.mechanismDescriptors();
  check(emcgpu_set_model(ctx, 0, vd.data(), (int)vd.size(), md.data(), (int)md.size()), ctx, "model");
  auto rates = e.rawRates();
  for (double r : rates)
    std::printf("rate %.17g\n", r);
  check(emcgpu_set_rate_tables(ctx, 0, 1000, 1., rates.data()), ctx, "tables");
  int nt = 0;
  double tau0 = 0;
  std::vector<double> cum(9 * 1000);
  check(emcgpu_get_tables(ctx, 0, 0, cum.data(), &tau0, &nt), ctx, "get");
  std::printf("tau0 %.17g\nntables %d\n", tau0, nt);
  // the reference's validation: a non-orthogonal subvalley frame is rejected
  try {
    emcNonParabolicAnisotropValley<T> bad({0.916, 0.196, 0.196}, e.getMass(), 3, 0.5);
    bad.setSubValleyEllipseCoordSystem(0, {1, 1, 0}, {1, 0, 0}, {0, 0, 1});
    std::printf("orthogonality not checked\n");
  } catch (const std::exception &ex) {
    std::printf("rejected %s\n", ex.what());
  }
  emcgpu_destroy(ctx);
  if (argc > 1) {
    const std::string dir = argv[1];
    e.writeTablesToFiles(dir + "/si_");
    emcgpuParticleType<T> d(1000, 1.5);
    d.addValley(std::make_unique<emcParabolicIsotropValley<T>>(0.26, d.getMass(), 1));
    const std::map<SizeType, std::vector<SizeType>> self = {{0, {0}}};
    d.addScatterMechanism({0, 1, 2}, std::make_unique<emcAcousticScatterMechanism<T>>(0, 9., device));
    d.addScatterMechanism({0, 1, 2}, std::make_unique<emcZeroOrderInterValleyAbsorptionScatterMechanism<T>>("P", 0, self, 5e10, 0.05, device));
    d.addScatterMechanism({0, 1, 2}, std::make_unique<emcZeroOrderInterValleyEmissionScatterMechanism<T>>("P", 0, self, 5e10, 0.05, device));
    d.writeTablesToFiles(dir + "/device_");
  }
  return 0;
}
