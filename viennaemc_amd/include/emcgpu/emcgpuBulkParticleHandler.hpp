// emcgpuBulkParticleHandler.hpp -- particle type and bulk handler of the GPU
// host API.
//
// emcgpuParticleType keeps the reference particle type's registration API
// (include/ParticleType/emcParticleType.hpp:125-153: addValley,
// addScatterMechanism(regions, mechanism)) and builds the scatter tables the
// reference way (fillScatterTables / renormalizeTables,
// include/emcScatterHandler.hpp:220-288) from the mechanisms' host-side
// getScatterRate.  emcgpuBulkParticleHandler keeps the bulk handler's
// methods (examples/bulkSimulation/basicBulkParticleHandler.hpp:93-347:
// generateInitialParticles, moveParticles, getAvgEnergy,
// getAvgDriftVelocity, getValleyOccupationProbability, setSeed,
// resetAppliedFieldStrength) and runs them on the GPU through the C ABI.
//
// Errors: the reference prints its message and aborts (emcMessage.hpp:43-65);
// that is the default here too.  setErrorPolicy(ErrorPolicy::Throw) turns
// errors into std::runtime_error instead.
#ifndef EMCGPU_BULK_PARTICLE_HANDLER_HPP
#define EMCGPU_BULK_PARTICLE_HANDLER_HPP

#include <array>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "emcgpuMechanisms.hpp"
#include "emcgpuOutput.hpp"

namespace emcgpu {

enum class ErrorPolicy { Abort, Throw };

inline ErrorPolicy &errorPolicy() {
  static ErrorPolicy p = ErrorPolicy::Abort;
  return p;
}
inline void setErrorPolicy(ErrorPolicy p) { errorPolicy() = p; }

inline void reportError(const std::string &msg) {
  if (errorPolicy() == ErrorPolicy::Throw)
    throw std::runtime_error(msg);
  std::fprintf(stderr, "ERROR: %s\n", msg.c_str());
  std::abort();
}

inline void check(emcgpu_status st, const emcgpu_ctx *ctx, const char *what) {
  if (st != EMCGPU_OK)
    reportError(std::string(what) + ": " + emcgpu_last_error(ctx));
}

/// material constants and lattice temperature of a bulk run; provides the
/// getters the mechanisms use (emcMaterial.hpp:27-37, emcDevice.hpp:185-190)
template <class T>
class emcgpuBulkDevice {
public:
  struct Material {
    T epsR, rho, ni, velSound;
    T getEpsR() const { return epsR; }
    T getRho() const { return rho; }
    T getNi() const { return ni; }
    T getVelSound() const { return velSound; }
  };
  emcgpuBulkDevice(Material m, T temperature) : mat(m), temp(temperature) {}
  /// with the extent of the periodic box [0, maxPos] [m]: the handler then
  /// keeps a position per particle (emcDevice::getMaxPos of the reference's
  /// bulk device, bulkSimulation.cpp:64-70)
  emcgpuBulkDevice(Material m, T temperature, std::array<T, 3> maxPos)
      : mat(m), temp(temperature), box(maxPos) {}
  const Material &getMaterial() const { return mat; }
  T getTemperature() const { return temp; }
  T getThermalVoltage() const { return phys::kB / phys::q * temp; }
  std::array<T, 3> getMaxPos() const { return box; }
  bool hasExtent() const { return box[0] > 0 && box[1] > 0 && box[2] > 0; }

private:
  Material mat;
  T temp;
  std::array<T, 3> box{0, 0, 0};
};

template <class T>
class emcgpuParticleType {
public:
  typedef emcAbstractValley<T> AbstractValley;
  typedef emcScatterMechanism<T> AbstractScatterMechanism;

  emcgpuParticleType(SizeType nrEnergyLevels = 1000, T maxEnergy = 4., T mass = phys::me,
                     T charge = -phys::q)
      : nLevels(nrEnergyLevels), maxE(maxEnergy), particleMass(mass), particleCharge(charge) {}

  T getMass() const { return particleMass; }
  T getCharge() const { return particleCharge; }
  /// name used in result file names (emcElectron::getName,
  /// include/ParticleType/emcElectron.hpp:38)
  std::string getName() const { return typeName; }
  void setName(const std::string &n) { typeName = n; }
  /// device runs: initial carriers per node from the equilibrium potential
  /// (emcElectron's usePotentialForInit, on by default there too,
  /// include/ParticleType/emcElectron.hpp:30-61) or from the doping
  bool usesPotentialForInit() const { return potentialInit; }
  void setUsePotentialForInit(bool on) { potentialInit = on; }
  SizeType getNrValleys() const { return valleys.size(); }
  const AbstractValley *getValley(SizeType i) const { return valleys.at(i).get(); }
  SizeType getNrEnergyLevels() const { return nLevels; }
  T getMaxEnergy() const { return maxE; }

  template <class DerivedValley>
  void addValley(std::unique_ptr<DerivedValley> &&v) {
    v->check();
    valleys.push_back(std::move(v));
    for (auto &m : mechs)
      m->setPtrValley(valleys);
  }

  /// the mechanism scatters in the listed doping regions (a bulk run uses
  /// region 0; a device run builds one table set per region when the lists
  /// differ); the regions also name the table dumps of writeTablesToFiles
  template <class DerivedMechanism>
  void addScatterMechanism(const std::vector<int> &regions,
                           std::unique_ptr<DerivedMechanism> &&m) {
    if (m->getIdxValley() >= valleys.size())
      reportError("Used index for Valley is invalid.");
    m->setPtrValley(valleys);
    m->check();
    mechs.push_back(std::move(m));
    mechRegions.push_back(regions.empty() ? std::vector<int>{0} : regions);
  }

  /// the scatter-table dumps of initScatterTables (writeTablesToFiles,
  /// include/emcScatterHandler.hpp:198-217): per mechanism and region a file
  /// <prefix><name><region><valley>ScatterMechanism.txt with one line
  /// "energy rate" per level, the energy accumulated in steps of dE
  void writeTablesToFiles(const std::string &prefix = "") const {
    const T dE = maxE / (T)nLevels;
    for (SizeType j = 0; j < mechs.size(); j++)
      for (int region : mechRegions[j]) {
        const std::string name = prefix + mechs[j]->getName() + std::to_string(region) +
                                 std::to_string(mechs[j]->getIdxValley()) +
                                 "ScatterMechanism.txt";
        std::ofstream os(name);
        if (!os)
          reportError("cannot write " + name);
        T e = dE;
        for (SizeType l = 0; l < nLevels; l++) {
          os << e << " " << mechs[j]->getScatterRate((T)(l + 1) * dE, (SizeType)region) << "\n";
          e += dE;
        }
      }
  }

  /// raw rates rates[mech][level] at energies (level+1) * maxEnergy / nLevels
  /// (fillScatterTables, emcScatterHandler.hpp:220-235)
  std::vector<double> rawRates() const {
    const T dE = maxE / (T)nLevels;
    std::vector<double> r(mechs.size() * nLevels);
    for (SizeType j = 0; j < mechs.size(); j++)
      for (SizeType l = 0; l < nLevels; l++)
        r[j * nLevels + l] = mechs[j]->getScatterRate((T)(l + 1) * dE, 0);
    return r;
  }

  /// the (valley, region) table sets of emcScatterHandler
  /// (emcScatterHandler.hpp:130-145): the mechanisms registered for region r
  /// in registration order, their descriptors and their rates at r
  std::vector<SizeType> mechanismsOfRegion(int region) const {
    std::vector<SizeType> idx;
    for (SizeType j = 0; j < mechs.size(); j++)
      for (int r : mechRegions[j])
        if (r == region) {
          idx.push_back(j);
          break;
        }
    return idx;
  }
  std::vector<emcgpu_mech_desc> mechanismDescriptors(int region) const {
    const auto all = mechanismDescriptors();
    std::vector<emcgpu_mech_desc> d;
    for (SizeType j : mechanismsOfRegion(region))
      d.push_back(all[j]);
    return d;
  }
  std::vector<double> rawRates(int region) const {
    const T dE = maxE / (T)nLevels;
    const auto idx = mechanismsOfRegion(region);
    std::vector<double> r(idx.size() * nLevels);
    for (SizeType i = 0; i < idx.size(); i++)
      for (SizeType l = 0; l < nLevels; l++)
        r[i * nLevels + l] = mechs[idx[i]]->getScatterRate((T)(l + 1) * dE, (SizeType)region);
    return r;
  }
  /// true when the first n regions do not all hold every mechanism
  bool regionDependent(int n) const {
    for (int r = 0; r < n; r++)
      if (mechanismsOfRegion(r).size() != mechs.size())
        return true;
    return false;
  }

  std::vector<emcgpu_valley_desc> valleyDescriptors() const {
    std::vector<emcgpu_valley_desc> d;
    for (auto &v : valleys)
      d.push_back(v->descriptor());
    for (auto &x : d)
      x.particle_mass = particleMass;
    return d;
  }
  std::vector<emcgpu_mech_desc> mechanismDescriptors() const {
    std::vector<emcgpu_mech_desc> d;
    const auto bl = baths();
    for (auto &m : mechs) {
      d.push_back(m->finalState());
      if (auto *bm = dynamic_cast<const emcBathMechanism<T> *>(m.get()))
        for (size_t b = 0; b < bl.size(); b++)
          if (bl[b] == bm->getBath())
            d.back().bath = (int)b;
    }
    return d;
  }
  /// the phonon baths the mechanisms read, in registration order
  std::vector<std::shared_ptr<emcPhononBath<T>>> baths() const {
    std::vector<std::shared_ptr<emcPhononBath<T>>> r;
    for (auto &m : mechs)
      if (auto *bm = dynamic_cast<const emcBathMechanism<T> *>(m.get())) {
        bool seen = false;
        for (auto &b : r)
          seen |= b == bm->getBath();
        if (!seen)
          r.push_back(bm->getBath());
      }
    return r;
  }
  /// every mechanism a built-in class (their descriptors carry the rates)
  bool allBuiltIn() const {
    for (auto &m : mechs)
      if (!m->isBuiltIn())
        return false;
    return true;
  }
  /// the screening the screened mechanisms read (the first one found; they
  /// share one emcPlasmonScreening as in the reference), or null
  const emcPlasmonScreening<T> *screening() const {
    for (auto &m : mechs)
      if (auto *s = dynamic_cast<const emcScreenedMechanism<T> *>(m.get()))
        return s->getScreening();
    return nullptr;
  }
  /// its state lowered to the C ABI (a fixed q_s^2)
  bool screeningDescriptor(emcgpu_screening_desc &d) const {
    const emcPlasmonScreening<T> *s = screening();
    if (!s)
      return false;
    d = emcgpu_screening_desc{};
    d.mode = EMCGPU_SCREENING_FIXED;
    d.enabled = s->isEnabled() ? 1 : 0;
    d.eps_static = s->getEpsStatic();
    d.qs2 = s->getQs2();
    return true;
  }

private:
  SizeType nLevels;
  T maxE, particleMass, particleCharge;
  std::string typeName = "Electrons";
  bool potentialInit = true;
  std::vector<std::unique_ptr<AbstractValley>> valleys;
  std::vector<std::unique_ptr<AbstractScatterMechanism>> mechs;
  std::vector<std::vector<int>> mechRegions;
};

template <class T>
class emcgpuBulkParticleHandler {
public:
  enum Precision { FP64 = EMCGPU_FP64, FP32 = EMCGPU_FP32 };

  /// the basicBulkParticleHandler constructor (basicBulkParticleHandler.hpp:93-120)
  /// with the particle count of the ensemble and the GPU precision; tables
  /// are built here (initScatterTables) from the mechanisms' rates
  template <class DeviceType>
  emcgpuBulkParticleHandler(const DeviceType &device, emcgpuParticleType<T> &type,
                            std::array<T, 3> fieldDirection, T fieldStrength, SizeType nrParticles,
                            unsigned long inSeed = 1, Precision precision = FP32, int gpu = 0)
      : ptype(type), dir(fieldDirection), seed(inSeed), temperature(device.getTemperature()),
        nPart(nrParticles) {
    material = emcgpu_material{0.0, (double)device.getMaterial().getRho(), 0.0,
                               (double)device.getMaterial().getVelSound()};
    emcgpu_config cfg{gpu, (int)precision, 1};
    check(emcgpu_create(&cfg, &ctx), nullptr, "emcgpu_create");
    auto vd = type.valleyDescriptors();
    auto md = type.mechanismDescriptors();
    check(emcgpu_set_model(ctx, 0, vd.data(), (int)vd.size(), md.data(), (int)md.size()), ctx,
          "emcgpu_set_model");
    // phonon baths: their state moves to the GPU, which updates them after
    // every step and rebuilds the tables from them
    baths = type.baths();
    if (!baths.empty()) {
      std::vector<emcgpu_bath_desc> bd;
      for (auto &b : baths)
        bd.push_back(b->descriptor());
      check(emcgpu_set_baths(ctx, 0, bd.data(), (int)bd.size()), ctx, "emcgpu_set_baths");
      for (size_t b = 0; b < baths.size(); b++)
        baths[b]->bind(ctx, 0, (int)b);
    }
    reinitScatterTables();
    resetAppliedFieldStrength(fieldStrength);
    size_t n = nrParticles;
    check(emcgpu_resize(ctx, &n), ctx, "emcgpu_resize");
    if (device.hasExtent()) {
      // positions drift with the particles and wrap in the box
      // (basicBulkParticleHandler.hpp:600-613)
      box = device.getMaxPos();
      const double b[3] = {box[0], box[1], box[2]};
      check(emcgpu_set_box(ctx, 0, b), ctx, "emcgpu_set_box");
    }
    check(emcgpu_obs_layout(ctx, &nObsL, &maxV, &maxDeg), ctx, "emcgpu_obs_layout");
  }
  ~emcgpuBulkParticleHandler() { emcgpu_destroy(ctx); }
  emcgpuBulkParticleHandler(const emcgpuBulkParticleHandler &) = delete;
  emcgpuBulkParticleHandler &operator=(const emcgpuBulkParticleHandler &) = delete;

  /// reinitScatterTables (emcParticleType.hpp / emcScatterHandler.hpp:
  /// 100-108): the tables from the mechanisms' current rates, with the
  /// current q_s^2 of a shared emcPlasmonScreening for the GPU final state
  void reinitScatterTables() {
    emcgpu_screening_desc sd;
    if (ptype.screeningDescriptor(sd)) {
      if (plasmaVolume > 0) {
        sd.mode = EMCGPU_SCREENING_PLASMA;
        sd.volume = plasmaVolume;
      }
      check(emcgpu_set_screening(ctx, 0, &sd), ctx, "emcgpu_set_screening");
    }
    if (!baths.empty() || plasmaVolume > 0) {
      // rates that follow a bath or the carriers are rebuilt on the GPU from
      // the descriptors (material and temperature of the device)
      if (!ptype.allBuiltIn())
        reportError("custom mechanisms cannot be combined with phonon baths or plasma screening");
      check(emcgpu_build_tables(ctx, 0, (int)ptype.getNrEnergyLevels(), ptype.getMaxEnergy(),
                                temperature, &material),
            ctx, "emcgpu_build_tables");
      return;
    }
    auto rates = ptype.rawRates();
    check(emcgpu_set_rate_tables(ctx, 0, (int)ptype.getNrEnergyLevels(), ptype.getMaxEnergy(),
                                 rates.data()),
          ctx, "emcgpu_set_rate_tables");
  }
  /// the updateScreening loop of the hot-carrier driver
  /// (examples/hotCarrierMHP/hotCarrierMHP.cpp:563-581, 693-703) on the GPU:
  /// after every step q_s^2 follows the carriers of the given volume [m^3]
  /// (static permittivity of the shared emcPlasmonScreening), the baths'
  /// window weights and the tables follow
  void enablePlasmaScreening(T volume) {
    if (!ptype.screening())
      reportError("plasma screening needs a screened mechanism");
    plasmaVolume = volume;
    reinitScatterTables();
  }
  /// updateScreening + reinitScatterTables from the current particles (the
  /// seeding before the first step, hotCarrierMHP.cpp:630-637)
  void updateScreening() { check(emcgpu_update_screening(ctx, 0), ctx, "emcgpu_update_screening"); }
  /// the q_s^2 the GPU holds [1/m^2]
  T getScreeningQ2() const {
    double q2 = 0;
    check(emcgpu_get_screening(ctx, 0, &q2), ctx, "emcgpu_get_screening");
    return (T)q2;
  }
  void setSeed(SizeType inSeed) { seed = inSeed; }
  void resetAppliedFieldStrength(T strength) {
    const double d[3] = {dir[0], dir[1], dir[2]};
    check(emcgpu_set_field(ctx, 0, d, strength, ptype.getCharge()), ctx, "emcgpu_set_field");
  }
  /// thermal initial ensemble at the device temperature (generateInitialParticles)
  /// (positions uniform in the box when the device has an extent); the
  /// getAvg* calls then return the initial ensemble's values
  void generateInitialParticles() {
    check(emcgpu_init_particles(ctx, 0, seed, temperature), ctx, "emcgpu_init_particles");
    step = 0;
    obs.clear();
    nObsSteps = 0;
    lastObs.assign((size_t)nObsL, 0.0);
    check(emcgpu_observe(ctx, lastObs.data()), ctx, "emcgpu_observe");
  }
  SizeType getNrParticles(SizeType = 0) const { return nPart; }

  /// moveParticles(dt) repeated nSteps times on the GPU; the observables of
  /// the last step are kept for the getAvg* calls
  void moveParticles(T dt, int nSteps = 1) {
    emcgpu_step_args a{};
    a.dt = dt;
    a.n_steps = nSteps;
    a.mode = EMCGPU_MODE_PRODUCTION;
    a.seed = seed;
    a.step0 = step;
    check(emcgpu_step(ctx, &a), ctx, "emcgpu_step");
    step += (uint32_t)nSteps;
    int nObs = 0, nS = 0;
    check(emcgpu_get_observables(ctx, 0, nullptr, &nObs, &nS), ctx, "emcgpu_get_observables");
    obs.resize((size_t)nObs * nS);
    check(emcgpu_get_observables(ctx, 0, obs.data(), &nObs, &nS), ctx, "emcgpu_get_observables");
    nObsSteps = nS;
    lastObs.assign(obs.end() - nObs, obs.end());
  }
  /// per-step sums of the last moveParticles call: [step][value]
  const std::vector<double> &stepObservables() const { return obs; }
  int observablesPerStep() const { return nObsL; }

  /// getAvgEnergy / getAvgDriftVelocity / getValleyOccupationProbability
  /// (basicBulkParticleHandler.hpp:289-347) of the state after the last step
  std::vector<T> getAvgEnergy(SizeType = 0) const { return perValley(lastObs.data(), 1); }
  std::vector<T> getAvgDriftVelocity(SizeType = 0) const { return perValley(lastObs.data(), 2); }
  std::vector<T> getValleyOccupationProbability(SizeType = 0) const {
    return occupation(lastObs.data());
  }
  /// the same values after step s (0 <= s < nSteps) of the last
  /// moveParticles(dt, nSteps) call: the per-step records of the bulk
  /// example (bulkSimulation.cpp:148-162) without a launch per step
  std::vector<T> getAvgEnergyAt(int s) const { return perValley(row(s), 1); }
  std::vector<T> getAvgDriftVelocityAt(int s) const { return perValley(row(s), 2); }
  std::vector<T> getValleyOccupationProbabilityAt(int s) const { return occupation(row(s)); }
  int stepsOfLastMove() const { return nObsSteps; }
  /// occupation of each (valley, subvalley)
  std::vector<T> getSubValleyOccupationProbability(SizeType valley) const {
    std::vector<T> r(ptype.getValley(valley)->getDegeneracyFactor());
    for (SizeType s = 0; s < r.size(); s++)
      r[s] = lastObs[2 + 3 * maxV + valley * maxDeg + s] / (T)nPart;
    return r;
  }

  /// print (basicBulkParticleHandler.hpp:227-248): the file
  /// <namePrefix><type name><nameSuffix>.txt with the box extent and
  /// "index x y z kx ky kz energy subvalley valley" per particle; needs a
  /// device with an extent
  void print(const std::string &namePrefix, const std::string &nameSuffix) const {
    if (!(box[0] > 0))
      reportError("print needs particle positions: construct the device with its extent");
    const size_t n = nPart;
    std::vector<double> k(3 * n), E(n), tau(n), pos(3 * n);
    std::vector<int32_t> valley(n), sub(n);
    check(emcgpu_get_particles(ctx, 0, k.data(), E.data(), tau.data(), valley.data(), sub.data()),
          ctx, "emcgpu_get_particles");
    check(emcgpu_get_positions(ctx, 0, pos.data()), ctx, "emcgpu_get_positions");
    const double b[3] = {box[0], box[1], box[2]};
    auto f = output::openResultFile(namePrefix + ptype.getName() + nameSuffix);
    output::printBulkParticles(f, b, n, pos.data(), k.data(), E.data(), sub.data(),
                               valley.data());
  }

  emcgpu_ctx *context() { return ctx; }

private:
  const double *row(int s) const {
    if (s < 0 || s >= nObsSteps)
      reportError("step index beyond the last moveParticles call");
    return obs.data() + (size_t)s * nObsL;
  }
  std::vector<T> perValley(const double *o, int which) const {
    std::vector<T> r(ptype.getNrValleys());
    for (SizeType v = 0; v < r.size(); v++) {
      const double cnt = o[2 + 3 * v];
      r[v] = cnt > 0 ? o[2 + 3 * v + which] / cnt : 0;
    }
    return r;
  }
  std::vector<T> occupation(const double *o) const {
    std::vector<T> r(ptype.getNrValleys());
    for (SizeType v = 0; v < r.size(); v++)
      r[v] = o[2 + 3 * v] / (T)nPart;
    return r;
  }

  emcgpu_ctx *ctx = nullptr;
  emcgpuParticleType<T> &ptype;
  std::array<T, 3> dir;
  unsigned long seed;
  T temperature;
  SizeType nPart;
  uint32_t step = 0;
  std::vector<double> obs, lastObs;
  int nObsL = 0, maxV = 1, maxDeg = 1, nObsSteps = 0;
  std::array<T, 3> box{0, 0, 0};
  emcgpu_material material{};
  std::vector<std::shared_ptr<emcPhononBath<T>>> baths;
  T plasmaVolume = 0;
};

}  // namespace emcgpu

#endif
