// emcgpuMechanisms.hpp -- scatter mechanisms of the GPU host API.
//
// emcScatterMechanism keeps the reference's plugin contract
// (include/ScatterMechanisms/emcScatterMechanism.hpp:17-53): a subclass
// provides getScatterRate(energy, region) on the host and names its final
// state.  The final state itself is sampled on the GPU, so instead of a
// virtual scatterParticle(particle, rng) a mechanism declares one of the
// built-in final-state kinds with its parameters (finalState()).
//
// The built-in mechanisms carry the reference's class names and constructor
// arguments (templated on the device type, which only needs
// getMaterial().getRho() / getVelSound() and getTemperature(), so the
// reference's emcDevice works as well as emcgpuBulkDevice).
#ifndef EMCGPU_MECHANISMS_HPP
#define EMCGPU_MECHANISMS_HPP

#include <cmath>
#include <limits>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "emcgpuValleys.hpp"

namespace emcgpu {

template <class T>
struct emcScatterMechanism {
  typedef emcAbstractValley<T> AbstractValley;

protected:
  SizeType idxValley;
  std::vector<const AbstractValley *> ptrValley;
  emcgpu_mech_desc desc_{};

  // final-state kinds a custom mechanism can choose
  void useIsotropicElasticFinalState() { desc_.kind = EMCGPU_MECH_ACOUSTIC; }
  void useInterValleyFinalState(SizeType finalValley, T phononEnergy, bool absorption,
                                const std::map<SizeType, std::vector<SizeType>> &finalSub) {
    desc_.kind = absorption ? EMCGPU_MECH_ZERO_IV_AB : EMCGPU_MECH_ZERO_IV_EM;
    desc_.final_valley = (int)finalValley;
    desc_.phonon_energy = phononEnergy;
    setFinalSubValleys(finalSub);
  }
  void useFroehlichFinalState(T phononEnergy, bool absorption) {
    desc_.kind = absorption ? EMCGPU_MECH_FROEHLICH_AB : EMCGPU_MECH_FROEHLICH_EM;
    desc_.phonon_energy = phononEnergy;
  }
  void setFinalSubValleys(const std::map<SizeType, std::vector<SizeType>> &finalSub) {
    if (finalSub.empty() || finalSub.begin()->second.empty())
      throw std::invalid_argument("final subvalley map is empty");
    desc_.z = (int)finalSub.begin()->second.size();
    if (desc_.z > EMCGPU_MAX_Z)
      throw std::invalid_argument("at most 8 final subvalleys per initial subvalley");
    for (const auto &kv : finalSub) {
      if (kv.first >= EMCGPU_MAX_DEG || (int)kv.second.size() != desc_.z)
        throw std::invalid_argument("Nr. of final subvalleys not consistent.");
      for (int z = 0; z < desc_.z; z++)
        desc_.final_sub[kv.first][z] = (int)kv.second[z];
    }
  }

public:
  explicit emcScatterMechanism(SizeType inIdxValley) : idxValley(inIdxValley) {
    desc_.valley = (int)inIdxValley;
    desc_.final_valley = (int)inIdxValley;
    desc_.z = 1;
  }
  virtual ~emcScatterMechanism() = default;

  /// rate of the mechanism [1/s] at the given energy [eV] (FP64, host)
  virtual T getScatterRate(T energy, SizeType idxRegion) const = 0;
  /// name used for table dumps
  virtual std::string getName() const = 0;
  virtual void check() {}

  /// the GPU final state of this mechanism
  const emcgpu_mech_desc &finalState() const { return desc_; }
  SizeType getIdxValley() const { return idxValley; }
  void setPtrValley(const std::vector<std::unique_ptr<AbstractValley>> &valleys) {
    ptrValley.clear();
    for (auto &v : valleys)
      ptrValley.push_back(v.get());
  }
};

/// emcAcousticScatterMechanism.hpp:24-72
template <class T>
class emcAcousticScatterMechanism : public emcScatterMechanism<T> {
  T scatterConst;

public:
  template <class DeviceType>
  emcAcousticScatterMechanism(SizeType inIdxValley, T sigma, const DeviceType &device)
      : emcAcousticScatterMechanism(inIdxValley, sigma, device.getMaterial().getRho(),
                                    device.getMaterial().getVelSound(), device) {}
  template <class DeviceType>
  emcAcousticScatterMechanism(SizeType inIdxValley, T sigma, T materialDensity, T velSound,
                              const DeviceType &device)
      : emcScatterMechanism<T>(inIdxValley) {
    scatterConst = phys::acousticConst(sigma, materialDensity, velSound, device.getTemperature());
    this->useIsotropicElasticFinalState();
  }
  std::string getName() const override { return "Acoustic"; }
  T getScatterRate(T energy, SizeType) const override {
    return phys::acousticRate(scatterConst, this->ptrValley[this->idxValley]->mass(), energy);
  }
};

/// zero- and first-order intervalley mechanisms
/// (emcZeroOrderInterValleyScatterMechanism.hpp, emcFirstOrderInterValleyScatterMechanism.hpp)
template <class T, bool FirstOrder, bool Absorption>
class emcInterValleyMechanism : public emcScatterMechanism<T> {
  std::string nameSuffix;
  T phononEnergy, scatterConst;
  SizeType idxFinalValley;

public:
  template <class DeviceType>
  emcInterValleyMechanism(std::string suffix, SizeType inIdxValley, SizeType inIdxFinalValley,
                          std::map<SizeType, std::vector<SizeType>> finalSubValleys,
                          T defPotential, T inPhononEnergy, const DeviceType &device)
      : emcScatterMechanism<T>(inIdxValley), nameSuffix(std::move(suffix)),
        phononEnergy(inPhononEnergy), idxFinalValley(inIdxFinalValley) {
    const int Z = (int)finalSubValleys.at(0).size();
    const T rho = device.getMaterial().getRho(), temp = device.getTemperature();
    scatterConst = FirstOrder
                       ? phys::firstOrderConst(defPotential, phononEnergy, rho, temp, Z, Absorption)
                       : phys::zeroOrderConst(defPotential, phononEnergy, rho, temp, Z, Absorption);
    this->desc_.kind = FirstOrder ? (Absorption ? EMCGPU_MECH_FIRST_IV_AB : EMCGPU_MECH_FIRST_IV_EM)
                                  : (Absorption ? EMCGPU_MECH_ZERO_IV_AB : EMCGPU_MECH_ZERO_IV_EM);
    this->desc_.final_valley = (int)inIdxFinalValley;
    this->desc_.phonon_energy = inPhononEnergy;
    this->desc_.coupling = defPotential;
    this->setFinalSubValleys(finalSubValleys);
  }
  template <class DeviceType>
  emcInterValleyMechanism(std::string suffix, SizeType inIdxValley,
                          std::map<SizeType, std::vector<SizeType>> finalSubValleys,
                          T defPotential, T inPhononEnergy, const DeviceType &device)
      : emcInterValleyMechanism(std::move(suffix), inIdxValley, inIdxValley, finalSubValleys,
                                defPotential, inPhononEnergy, device) {}

  std::string getName() const override {
    return std::string(FirstOrder ? "First" : "Zero") + "InterValley" +
           (Absorption ? "Absorption" : "Emission") + nameSuffix;
  }
  T getScatterRate(T energy, SizeType) const override {
    const auto &iv = this->ptrValley[this->idxValley]->mass();
    const auto &fv = this->ptrValley[idxFinalValley]->mass();
    const T delta = fv.bottom - iv.bottom;
    if (FirstOrder)
      return phys::firstOrderRate(scatterConst, iv, fv, energy, phononEnergy, delta, Absorption);
    return phys::zeroOrderRate(scatterConst, fv, energy, phononEnergy, delta, Absorption);
  }
  void check() override {
    if (idxFinalValley >= this->ptrValley.size())
      throw std::invalid_argument(getName() + ": idxFinalValley is not valid.");
  }
};

template <class T>
using emcZeroOrderInterValleyAbsorptionScatterMechanism = emcInterValleyMechanism<T, false, true>;
template <class T>
using emcZeroOrderInterValleyEmissionScatterMechanism = emcInterValleyMechanism<T, false, false>;
template <class T>
using emcFirstOrderInterValleyAbsorptionScatterMechanism = emcInterValleyMechanism<T, true, true>;
template <class T>
using emcFirstOrderInterValleyEmissionScatterMechanism = emcInterValleyMechanism<T, true, false>;

/// emcFroehlichInteraction.hpp:60-201 (equilibrium Bose occupation)
template <class T, bool Absorption>
class emcFroehlich3D : public emcScatterMechanism<T> {
  T phononEnergy, effMass, scatterConst, nBose;
  std::string nameSuffix;

public:
  emcFroehlich3D(SizeType inValley, T inPhononEnergy, T relEffMass, T epsHi, T epsLo,
                 T temperature, std::string inNameSuffix = "")
      : emcScatterMechanism<T>(inValley), phononEnergy(inPhononEnergy),
        effMass(relEffMass * phys::me), nameSuffix(std::move(inNameSuffix)) {
    scatterConst = phys::froehlichConst(phononEnergy, effMass, epsHi, epsLo);
    nBose = phys::boseEinstein(phononEnergy, temperature);
    this->useFroehlichFinalState(inPhononEnergy, Absorption);
    this->desc_.rel_eff_mass = relEffMass;
    this->desc_.eps_hi = epsHi;
    this->desc_.eps_lo = epsLo;
  }
  std::string getName() const override {
    return std::string(Absorption ? "FroehlichAbsorption3D" : "FroehlichEmission3D") + nameSuffix;
  }
  T getScatterRate(T energy, SizeType) const override {
    return phys::froehlichRate(scatterConst, nBose, effMass,
                               this->ptrValley[this->idxValley]->mass(), energy, phononEnergy,
                               Absorption);
  }
};

template <class T>
using emcFroehlichAbsorption3D = emcFroehlich3D<T, true>;
template <class T>
using emcFroehlichEmission3D = emcFroehlich3D<T, false>;

/// emcPlasmonScreening.hpp:52-90: the q_s^2 that screened mechanisms share
/// (Debye form q_s^2 = n q^2 / (eps_s eps0 kB T_e); 0 when disabled)
template <class T>
class emcPlasmonScreening {
  T qs2 = T(0);
  T epsStatic;
  bool enabled;

public:
  emcPlasmonScreening() = delete;
  emcPlasmonScreening(T inEpsStatic, bool inEnabled = true)
      : epsStatic(inEpsStatic), enabled(inEnabled) {}
  /// refresh from the carrier density [1/m^3] and temperature [K]
  void update(T density, T carrierTemp) {
    if (!enabled || density <= T(0) || carrierTemp <= T(0)) {
      qs2 = T(0);
      return;
    }
    qs2 = density * phys::q * phys::q / (epsStatic * phys::eps0 * phys::kB * carrierTemp);
  }
  void setQs2(T inQs2) { qs2 = enabled ? inQs2 : T(0); }
  T getQs2() const { return qs2; }
  T getQs() const { return std::sqrt(qs2); }
  T getScreeningLength() const {
    return (qs2 > T(0)) ? T(1) / std::sqrt(qs2) : std::numeric_limits<T>::infinity();
  }
  bool isEnabled() const { return enabled; }
  T getEpsStatic() const { return epsStatic; }
};

/// a mechanism that reads a shared emcPlasmonScreening; the particle type
/// hands its q_s^2 to the GPU final state (emcgpu_set_screening)
template <class T>
struct emcScreenedMechanism {
  virtual ~emcScreenedMechanism() = default;
  virtual const emcPlasmonScreening<T> *getScreening() const = 0;
};

/// emcScreenedFroehlichInteraction.hpp:101-213 (equilibrium Bose
/// occupation): the Froehlich rate with the screened log factor
/// 1/2 ln((q+^2 + q_s^2) / (q-^2 + q_s^2)) and the screened angle on the GPU
template <class T, bool Absorption>
class emcScreenedFroehlich3D : public emcScatterMechanism<T>, public emcScreenedMechanism<T> {
  T phononEnergy, effMass, scatterConst, nBose;
  std::shared_ptr<emcPlasmonScreening<T>> screening;
  std::string nameSuffix;

public:
  emcScreenedFroehlich3D(SizeType inValley, T inPhononEnergy, T relEffMass, T epsHi, T epsLo,
                         T temperature, std::shared_ptr<emcPlasmonScreening<T>> inScreening,
                         std::string inNameSuffix = "")
      : emcScatterMechanism<T>(inValley), phononEnergy(inPhononEnergy),
        effMass(relEffMass * phys::me), screening(std::move(inScreening)),
        nameSuffix(std::move(inNameSuffix)) {
    if (!screening)
      throw std::invalid_argument("screened Froehlich mechanism without a screening object");
    scatterConst = phys::froehlichConst(phononEnergy, effMass, epsHi, epsLo);
    nBose = phys::boseEinstein(phononEnergy, temperature);
    this->desc_.kind = Absorption ? EMCGPU_MECH_SCREENED_FROEHLICH_AB
                                  : EMCGPU_MECH_SCREENED_FROEHLICH_EM;
    this->desc_.phonon_energy = inPhononEnergy;
    this->desc_.rel_eff_mass = relEffMass;
    this->desc_.eps_hi = epsHi;
    this->desc_.eps_lo = epsLo;
  }
  std::string getName() const override {
    return std::string(Absorption ? "ScreenedFroehlichAbsorption3D"
                                  : "ScreenedFroehlichEmission3D") +
           nameSuffix;
  }
  T getScatterRate(T energy, SizeType) const override {
    return phys::screenedFroehlichRate(scatterConst, nBose, effMass,
                                       this->ptrValley[this->idxValley]->mass(), energy,
                                       phononEnergy, screening->getQs2(), Absorption);
  }
  const emcPlasmonScreening<T> *getScreening() const override { return screening.get(); }
};

template <class T>
using emcScreenedFroehlichAbsorption3D = emcScreenedFroehlich3D<T, true>;
template <class T>
using emcScreenedFroehlichEmission3D = emcScreenedFroehlich3D<T, false>;

}  // namespace emcgpu

#endif
