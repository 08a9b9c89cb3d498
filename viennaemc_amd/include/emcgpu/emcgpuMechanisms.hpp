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
  bool builtin_ = false;  // a built-in class whose descriptor carries its rate

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
  /// true for the built-in classes: the descriptor alone rebuilds the rate
  bool isBuiltIn() const { return builtin_; }
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
    this->desc_.coupling = sigma;
    this->builtin_ = true;
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
    this->builtin_ = true;
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
    this->builtin_ = true;
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
    this->builtin_ = true;
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

/// emcPhononBath (include/emcPhononBath.hpp:56-501): the constructor's
/// parameters and, once a handler owns the bath (bind), its state on the GPU:
/// getMeanNq, getNq, the reservoirs.  Before that the state is the initial
/// one (N_q = N0 everywhere).
template <class T>
class emcPhononBath {
  emcgpu_bath_desc d_{};
  std::vector<double> profile_;
  T n0_, nAcEq_ = 0, nToEq_ = 0;
  const emcgpu_ctx *ctx_ = nullptr;
  int inst_ = 0, index_ = -1;

  static T planck(T e, T temp) { return T(1) / (std::exp(phys::q * e / (phys::kB * temp)) - T(1)); }
  // {mean, n_ac, n_to, n0, t_ac, t_to} and N_q from the GPU
  void pull(std::vector<double> *nq, double sc[6]) const {
    std::vector<double> tmp;
    if (!nq)
      nq = &tmp;
    nq->resize(d_.n_bins);
    if (emcgpu_get_bath(ctx_, inst_, index_, nq->data(), sc) != EMCGPU_OK)
      throw std::runtime_error(emcgpu_last_error(ctx_));
  }

public:
  emcPhononBath(SizeType nrBins, T dq, T tauLO, T phononEnergy, T latticeTemp, T vSim,
                bool enableAcoustic = false, T acPhononEnergy = T(0), T tauAcoustic = T(0),
                T wRidley = T(0), T toPhononEnergy = T(0), T tauTO = T(0)) {
    if (nrBins < 1 || nrBins > EMCGPU_MAX_BINS)
      throw std::invalid_argument("emcPhononBath: number of bins out of range");
    d_.n_bins = (int)nrBins;
    d_.dq = dq;
    d_.tau_lo = tauLO;
    d_.phonon_energy = phononEnergy;
    d_.lattice_temp = latticeTemp;
    d_.volume = vSim;
    d_.acoustic = enableAcoustic ? 1 : 0;
    d_.ac_phonon_energy = acPhononEnergy;
    d_.tau_acoustic = tauAcoustic;
    d_.w_ridley = wRidley;
    d_.to_phonon_energy = toPhononEnergy;
    d_.tau_to = tauTO;
    n0_ = planck(phononEnergy, latticeTemp);
    if (enableAcoustic)
      nAcEq_ = planck(acPhononEnergy, latticeTemp);
    if (enableAcoustic && wRidley > T(0) && toPhononEnergy > T(0))
      nToEq_ = planck(toPhononEnergy, latticeTemp);
  }
  /// per-bin LO lifetime [s] (setTauLOProfile, :214-216); empty = uniform
  void setTauLOProfile(std::vector<T> inProfile) {
    if (!inProfile.empty() && inProfile.size() != (size_t)d_.n_bins)
      throw std::invalid_argument("emcPhononBath: profile size must equal the number of bins");
    profile_.assign(inProfile.begin(), inProfile.end());
  }
  /// the descriptor of the C ABI (its profile pointer valid while the bath lives)
  emcgpu_bath_desc descriptor() const {
    emcgpu_bath_desc d = d_;
    d.tau_profile = profile_.empty() ? nullptr : profile_.data();
    return d;
  }
  /// the handler that owns the bath's state on the GPU
  void bind(const emcgpu_ctx *ctx, int instance, int index) {
    ctx_ = ctx;
    inst_ = instance;
    index_ = index;
  }
  bool isBound() const { return ctx_ != nullptr; }

  SizeType getNrBins() const { return (SizeType)d_.n_bins; }
  T getDq() const { return d_.dq; }
  T qCentre(SizeType i) const { return (T(i) + T(0.5)) * d_.dq; }
  SizeType binOf(T q) const {
    const SizeType idx = static_cast<SizeType>(q / d_.dq);
    return (idx >= (SizeType)d_.n_bins) ? (SizeType)d_.n_bins - 1 : idx;
  }
  T getN0() const { return n0_; }
  T getTauLO() const { return d_.tau_lo; }
  T getLOEnergy() const { return d_.phonon_energy; }
  bool ridleyEnabled() const { return nToEq_ > T(0); }
  std::vector<T> getNqAll() const {
    if (!isBound())
      return std::vector<T>(d_.n_bins, n0_);
    std::vector<double> nq;
    double sc[6];
    pull(&nq, sc);
    return std::vector<T>(nq.begin(), nq.end());
  }
  T getNq(T q) const { return getNqAll()[binOf(q)]; }
  T getMeanNq() const {
    if (!isBound())
      return n0_;
    double sc[6];
    pull(nullptr, sc);
    return sc[0];
  }
  T getMeanNac() const {
    if (!isBound())
      return nAcEq_;
    double sc[6];
    pull(nullptr, sc);
    return sc[1];
  }
  T getMeanNTO() const {
    if (!isBound())
      return nToEq_;
    double sc[6];
    pull(nullptr, sc);
    return sc[2];
  }
  T getAcousticTemp() const {
    if (!isBound())
      return d_.lattice_temp;
    double sc[6];
    pull(nullptr, sc);
    return sc[4];
  }
  T getTOTemp() const {
    if (!isBound())
      return d_.lattice_temp;
    double sc[6];
    pull(nullptr, sc);
    return sc[5];
  }
};

/// a mechanism that reads and feeds a phonon bath
template <class T>
struct emcBathMechanism {
  virtual ~emcBathMechanism() = default;
  virtual const std::shared_ptr<emcPhononBath<T>> &getBath() const = 0;
};

/// emcHotPhononFroehlichMechanism.hpp:39-200: the Froehlich rate with the
/// bath's mean occupation; on the GPU the tables follow the bath after every
/// step and |k' - k| of each event is tallied in it
template <class T, bool Absorption>
class emcHotPhononFroehlich3D : public emcScatterMechanism<T>, public emcBathMechanism<T> {
  T phononEnergy, effMass, scatterConst;
  std::shared_ptr<emcPhononBath<T>> bath;
  std::string nameSuffix;

public:
  emcHotPhononFroehlich3D(SizeType inValley, T inPhononEnergy, T relEffMass, T epsHi, T epsLo,
                          std::shared_ptr<emcPhononBath<T>> inPhononBath,
                          std::string inNameSuffix = "")
      : emcScatterMechanism<T>(inValley), phononEnergy(inPhononEnergy),
        effMass(relEffMass * phys::me), bath(std::move(inPhononBath)),
        nameSuffix(std::move(inNameSuffix)) {
    if (!bath)
      throw std::invalid_argument("hot-phonon mechanism without a phonon bath");
    scatterConst = phys::froehlichConst(phononEnergy, effMass, epsHi, epsLo);
    this->desc_.kind = Absorption ? EMCGPU_MECH_HOT_FROEHLICH_AB : EMCGPU_MECH_HOT_FROEHLICH_EM;
    this->desc_.phonon_energy = inPhononEnergy;
    this->desc_.rel_eff_mass = relEffMass;
    this->desc_.eps_hi = epsHi;
    this->desc_.eps_lo = epsLo;
    this->builtin_ = true;
  }
  std::string getName() const override {
    return std::string(Absorption ? "HotPhononFroehlichAbsorption3D"
                                  : "HotPhononFroehlichEmission3D") +
           nameSuffix;
  }
  T getScatterRate(T energy, SizeType) const override {
    return phys::froehlichRate(scatterConst, bath->getMeanNq(), effMass,
                               this->ptrValley[this->idxValley]->mass(), energy, phononEnergy,
                               Absorption);
  }
  const std::shared_ptr<emcPhononBath<T>> &getBath() const override { return bath; }
};

template <class T>
using emcHotPhononFroehlichAbsorption3D = emcHotPhononFroehlich3D<T, true>;
template <class T>
using emcHotPhononFroehlichEmission3D = emcHotPhononFroehlich3D<T, false>;

/// emcScreenedFroehlichInteraction.hpp:218-375: the screened rate with the
/// bath's mean or (qResolved) window occupation; the angle screened or
/// (qResolved and qResolvedAngle) through the bath's q distribution
template <class T, bool Absorption>
class emcScreenedHotPhononFroehlich3D : public emcScatterMechanism<T>,
                                        public emcBathMechanism<T>,
                                        public emcScreenedMechanism<T> {
  std::shared_ptr<emcPhononBath<T>> bath;
  std::shared_ptr<emcPlasmonScreening<T>> screening;
  std::string nameSuffix;

public:
  emcScreenedHotPhononFroehlich3D(SizeType inValley, T inPhononEnergy, T relEffMass, T epsHi,
                                  T epsLo, std::shared_ptr<emcPhononBath<T>> inPhononBath,
                                  std::shared_ptr<emcPlasmonScreening<T>> inScreening,
                                  bool inQResolved = false, std::string inNameSuffix = "",
                                  bool inQResolvedAngle = true)
      : emcScatterMechanism<T>(inValley), bath(std::move(inPhononBath)),
        screening(std::move(inScreening)), nameSuffix(std::move(inNameSuffix)) {
    if (!bath || !screening)
      throw std::invalid_argument("screened hot-phonon mechanism without bath or screening");
    this->desc_.kind = Absorption ? EMCGPU_MECH_SCREENED_HOT_FROEHLICH_AB
                                  : EMCGPU_MECH_SCREENED_HOT_FROEHLICH_EM;
    this->desc_.phonon_energy = inPhononEnergy;
    this->desc_.rel_eff_mass = relEffMass;
    this->desc_.eps_hi = epsHi;
    this->desc_.eps_lo = epsLo;
    this->desc_.q_resolved = inQResolved ? 1 : 0;
    this->desc_.q_resolved_angle = inQResolvedAngle ? 1 : 0;
    this->builtin_ = true;
  }
  std::string getName() const override {
    return std::string(Absorption ? "ScreenedHotPhononFroehlichAbsorption3D"
                                  : "ScreenedHotPhononFroehlichEmission3D") +
           nameSuffix;
  }
  /// the rates live on the GPU (they follow the bath's window sums); the
  /// host has no copy of those
  T getScatterRate(T, SizeType) const override {
    throw std::logic_error(getName() + ": rates are rebuilt on the GPU from the bath");
  }
  const std::shared_ptr<emcPhononBath<T>> &getBath() const override { return bath; }
  const emcPlasmonScreening<T> *getScreening() const override { return screening.get(); }
};

template <class T>
using emcScreenedHotPhononFroehlichAbsorption3D = emcScreenedHotPhononFroehlich3D<T, true>;
template <class T>
using emcScreenedHotPhononFroehlichEmission3D = emcScreenedHotPhononFroehlich3D<T, false>;

}  // namespace emcgpu

#endif
