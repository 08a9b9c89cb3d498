// emcgpuPhysics.hpp -- FP64 host formulas of the valleys and scatter rates.
//
// Header-only.  Used by the C ABI library to build scatter tables from
// mechanism descriptors and by the C++ host API mechanism classes.  Each
// formula keeps the reference's operation order so that the tables match
// the reference's FP64 tables; paths are relative to the ViennaEMC tree.
#ifndef EMCGPU_PHYSICS_HPP
#define EMCGPU_PHYSICS_HPP

#include <array>
#include <cmath>
#include <cstddef>

// the rate formulas also run on the GPU (table construction from
// descriptors) when this header is compiled as HIP
#if defined(__HIPCC__)
#define EMCGPU_HD __host__ __device__
#else
#define EMCGPU_HD
#endif

namespace emcgpu {
namespace phys {

// include/emcConstants.hpp:9-25 (the reference's non-CODATA values)
constexpr double pi = 3.14159265358979323846;
constexpr double q = 1.60219e-19;
constexpr double kB = 1.38066e-23;
constexpr double hbar = 1.05459e-34;
constexpr double eps0 = 8.85419e-12;
constexpr double me = 9.11e-31;

enum ValleyKind { PARA_ISO = 0, NONPARA_ISO = 1, PARA_ANISO = 2, NONPARA_ANISO = 3 };

EMCGPU_HD inline bool isAniso(int kind) { return kind == PARA_ANISO || kind == NONPARA_ANISO; }
EMCGPU_HD inline bool isNonPara(int kind) { return kind == NONPARA_ISO || kind == NONPARA_ANISO; }

// Derived quantities of a valley (ValleyTypes/*: initEffMasses, initVogtFactor)
struct ValleyMass {
  double effMassCond = 0, effMassDOS = 0, alpha = 0, bottom = 0;
  std::array<double, 3> vogt{{1, 1, 1}};
  int kind = 0;

  ValleyMass() = default;
  ValleyMass(int inKind, const double rel[3], double particleMass, double inAlpha,
             double inBottom)
      : alpha(isNonPara(inKind) ? inAlpha : 0.0), bottom(inBottom), kind(inKind) {
    if (isAniso(kind)) {
      // emcNonParabolicAnistropValley.hpp:185-194
      double dos = 1.;
      for (int i = 0; i < 3; i++)
        dos = dos * rel[i];
      effMassDOS = std::pow(dos, 1. / 3.) * particleMass;
      double acc = 0.;
      for (int i = 0; i < 3; i++)
        acc += 1. / rel[i];
      effMassCond = 3. * particleMass / acc;
      // :198-204
      for (int i = 0; i < 3; i++)
        vogt[i] = std::sqrt(effMassCond / (rel[i] * particleMass));
    } else {
      effMassDOS = rel[0] * particleMass;
      effMassCond = rel[0] * particleMass;
    }
  }
  // getGamma (emcNonParabolicAnistropValley.hpp:123)
  EMCGPU_HD double gamma(double E) const { return isNonPara(kind) ? E * (1 + alpha * E) : E; }
  // getEffMassDOS at the default E = 0 used by every rate (:84-86)
  EMCGPU_HD double dosMass0() const {
    return isNonPara(kind) ? effMassDOS * std::pow(1 + 2 * alpha * 0.0, 3.) : effMassDOS;
  }
};

// emcAcousticScatterMechanism.hpp:36-45
inline double acousticConst(double sigma, double rho, double vSound, double T) {
  double cL = rho * std::pow(vSound, 2);
  return std::sqrt(2.0 * q) * std::pow(sigma * q, 2) * kB * T / (pi * cL * std::pow(hbar, 4));
}
// emcAcousticScatterMechanism.hpp:49-56
EMCGPU_HD inline double acousticRate(double C, const ValleyMass &v, double E) {
  double md = v.dosMass0();
  double alpha = v.alpha;
  double gamma = v.gamma(E);
  return C * std::pow(md, 3. / 2.) * std::sqrt(gamma) * (2 * alpha * E + 1.0);
}

// getZeroOrderScatterConst (emcZeroOrderInterValleyScatterMechanism.hpp:11-21)
inline double zeroOrderConst(double defPot, double phEnergy, double rho, double T, int Z,
                             bool absorption) {
  double result = (double)(std::size_t)Z * std::sqrt(q) * std::pow(defPot / hbar, 2) * q /
                  (pi * rho * phEnergy * std::sqrt(2));
  double nrPh = 1. / (std::exp(q * phEnergy / (kB * T)) - 1.);
  return absorption ? result * nrPh : result * (nrPh + 1);
}
// getFirstOrderScatterConst (emcFirstOrderInterValleyScatterMechanism.hpp:11-20)
inline double firstOrderConst(double defPot, double phEnergy, double rho, double T, int Z,
                              bool absorption) {
  double result = (double)(std::size_t)Z * std::sqrt(2) * std::pow(q, 5. / 2.) *
                  std::pow(defPot, 2) / (pi * rho * std::pow(hbar, 4) * phEnergy);
  double nrPh = 1. / (std::exp(q * phEnergy / (kB * T)) - 1.);
  return absorption ? result * nrPh : result * (nrPh + 1);
}

// zero-order rate (:103-117 absorption, :246-260 emission)
EMCGPU_HD inline double zeroOrderRate(double C, const ValleyMass &fin, double E, double hw, double delta,
                            bool absorption) {
  double Ef = absorption ? E + hw - delta : E - hw - delta;
  if (Ef > 0) {
    double md = fin.dosMass0();
    double alpha = fin.alpha;
    double gamma = fin.gamma(Ef);
    return C * std::pow(md, 3. / 2.) * std::sqrt(gamma) * (2 * alpha * Ef + 1.0);
  }
  return 0;
}
// first-order rate (emcFirstOrderInterValleyScatterMechanism.hpp:102-119, :250-267)
EMCGPU_HD inline double firstOrderRate(double C, const ValleyMass &ini, const ValleyMass &fin, double E,
                             double hw, double delta, bool absorption) {
  double Ef = absorption ? E + hw - delta : E - hw - delta;
  if (Ef > 0) {
    double md = fin.dosMass0();
    double alpha = fin.alpha;
    double gamma = ini.gamma(E);
    double gammaF = fin.gamma(Ef);
    return C * std::pow(md, 5. / 2.) * std::sqrt(gammaF) * (2 * alpha * Ef + 1.0) *
           (gamma + gammaF);
  }
  return 0;
}

// froehlichScatterConst3D / boseEinstein (emcFroehlichInteraction.hpp:45-57)
inline double froehlichConst(double hw, double effMass, double epsHi, double epsLo) {
  double omega0 = hw * q / hbar;
  return q * q * omega0 * effMass / (4. * pi * hbar * hbar) * (1.0 / epsHi - 1.0 / epsLo) / eps0;
}
inline double boseEinstein(double hw, double T) {
  double x = q * hw / (kB * T);
  return 1.0 / (std::exp(x) - 1.0);
}
// Froehlich rates (emcFroehlichInteraction.hpp:95-107, :167-181)
EMCGPU_HD inline double froehlichRate(double C, double nq, double effMass, const ValleyMass &v, double E,
                            double hw, bool absorption) {
  if (!absorption && E <= hw)
    return 0.0;
  double gammaI = v.gamma(E);
  double Ef = absorption ? E + hw : E - hw;
  double gammaF = v.gamma(Ef);
  if (gammaI <= 0.0 || gammaF <= 0.0)
    return 0.0;
  double kI = std::sqrt(2.0 * effMass * gammaI * q) / hbar;
  double kF = std::sqrt(2.0 * effMass * gammaF * q) / hbar;
  double ln = absorption ? std::log((kI + kF) / (kF - kI)) : std::log((kI + kF) / (kI - kF));
  return absorption ? C * nq / kI * ln : C * (nq + 1.0) / kI * ln;
}

// screenedFroehlichLogFactor (emcScreenedFroehlichInteraction.hpp:52-62):
// 1/2 ln((q+^2 + qs2) / (q-^2 + qs2)), q+- = kI +- kF; qs2 <= 0 is the bare
// ln(q+ / q-), 0 where q- vanishes
EMCGPU_HD inline double screenedLogFactor(double kI, double kF, double qs2) {
  const double p2 = (kI + kF) * (kI + kF), m2 = (kI - kF) * (kI - kF);
  if (qs2 <= 0.0)
    return m2 <= 0.0 ? 0.0 : 0.5 * std::log(p2 / m2);
  return 0.5 * std::log((p2 + qs2) / (m2 + qs2));
}
// screened Froehlich rates (emcScreenedFroehlichInteraction.hpp:128-138,
// :184-196): the Froehlich rate with the screened log factor
EMCGPU_HD inline double screenedFroehlichRate(double C, double nq, double effMass, const ValleyMass &v,
                                    double E, double hw, double qs2, bool absorption) {
  if (!absorption && E <= hw)
    return 0.0;
  const double gammaI = v.gamma(E);
  const double gammaF = v.gamma(absorption ? E + hw : E - hw);
  if (gammaI <= 0.0 || gammaF <= 0.0)
    return 0.0;
  const double kI = std::sqrt(2.0 * effMass * gammaI * q) / hbar;
  const double kF = std::sqrt(2.0 * effMass * gammaF * q) / hbar;
  return (absorption ? C * nq : C * (nq + 1.0)) / kI * screenedLogFactor(kI, kF, qs2);
}

}  // namespace phys
}  // namespace emcgpu

#endif
