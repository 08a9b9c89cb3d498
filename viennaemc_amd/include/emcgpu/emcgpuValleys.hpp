// emcgpuValleys.hpp -- valley types of the GPU host API.
//
// Same class names, constructor arguments and subvalley-axis setter as the
// reference's ValleyTypes (include/ValleyTypes/emcParabolicIsotropValley.hpp,
// emcNonParabolicIsotropValley.hpp, emcParabolicAnisotropValley.hpp,
// emcNonParabolicAnistropValley.hpp), living in namespace emcgpu.  Each valley
// lowers to an emcgpu_valley_desc for the C ABI; the band quantities used by
// the rate formulas come from emcgpuPhysics.hpp.
#ifndef EMCGPU_VALLEYS_HPP
#define EMCGPU_VALLEYS_HPP

#include <array>
#include <cstddef>
#include <cstring>
#include <stdexcept>
#include <string>

#include "../../../include/emcgpu.h"
#include "emcgpuPhysics.hpp"

namespace emcgpu {

using SizeType = std::size_t;

/// base of the valley types (emcAbstractValley.hpp:20-91, the members the
/// scatter rates need)
template <class T>
class emcAbstractValley {
protected:
  emcgpu_valley_desc desc_{};
  phys::ValleyMass mass_;

  void init(int kind, const std::array<T, 3> &rel, T particleMass, SizeType deg, T alpha,
            T bottom) {
    std::memset(&desc_, 0, sizeof(desc_));
    desc_.kind = kind;
    for (int i = 0; i < 3; i++)
      desc_.rel_mass[i] = rel[i];
    desc_.particle_mass = particleMass;
    desc_.degeneracy = (int)deg;
    desc_.alpha = alpha;
    desc_.bottom_energy = bottom;
    for (SizeType s = 0; s < EMCGPU_MAX_DEG; s++) {
      desc_.axes[s][0] = desc_.axes[s][4] = desc_.axes[s][8] = 1.0;
    }
    const double r[3] = {rel[0], rel[1], rel[2]};
    mass_ = phys::ValleyMass(kind, r, particleMass, alpha, bottom);
  }

public:
  virtual ~emcAbstractValley() = default;
  const emcgpu_valley_desc &descriptor() const { return desc_; }
  const phys::ValleyMass &mass() const { return mass_; }
  T getEffMassDOS(T = 0) const { return mass_.dosMass0(); }
  T getEffMassCond(T = 0) const { return mass_.effMassCond; }
  T getNonParabolicity() const { return mass_.alpha; }
  T getBottomEnergy() const { return mass_.bottom; }
  SizeType getDegeneracyFactor() const { return (SizeType)desc_.degeneracy; }
  T getGamma(T energy) const { return mass_.gamma(energy); }
  /// emcAbstractValley::check (emcAbstractValley.hpp:79-90)
  void check() const {
    if (getDegeneracyFactor() < 1 || getDegeneracyFactor() > EMCGPU_MAX_DEG)
      throw std::invalid_argument("Degeneracy Factor of a valley has to be in [1, 8].");
    if (desc_.alpha < 0)
      throw std::invalid_argument("NonParabolicity of a valley has to be positive.");
  }
};

/// emcParabolicIsotropValley.hpp:37-44
template <class T>
class emcParabolicIsotropValley : public emcAbstractValley<T> {
public:
  emcParabolicIsotropValley(T relEffMass, T particleMass, SizeType degFactor, T bottom = 0.) {
    this->init(EMCGPU_VALLEY_PARA_ISO, {relEffMass, relEffMass, relEffMass}, particleMass,
               degFactor, 0, bottom);
  }
};

/// emcNonParabolicIsotropValley.hpp:44-50
template <class T>
class emcNonParabolicIsotropValley : public emcAbstractValley<T> {
public:
  emcNonParabolicIsotropValley(T relEffMass, T particleMass, SizeType degFactor, T alpha,
                               T bottom = 0.) {
    this->init(EMCGPU_VALLEY_NONPARA_ISO, {relEffMass, relEffMass, relEffMass}, particleMass,
               degFactor, alpha, bottom);
  }
};

template <class T>
class emcAnisotropValleyBase : public emcAbstractValley<T> {
public:
  /// setSubValleyEllipseCoordSystem (emcNonParabolicAnistropValley.hpp:68-80):
  /// the three directions must be orthogonal; they need not be normalised
  void setSubValleyEllipseCoordSystem(SizeType idxSubValley, std::array<T, 3> dir1,
                                      std::array<T, 3> dir2, std::array<T, 3> dir3) {
    if (idxSubValley >= this->getDegeneracyFactor())
      throw std::invalid_argument("Using invalid subvalley index.");
    auto dot = [](const std::array<T, 3> &a, const std::array<T, 3> &b) {
      return a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
    };
    if (dot(dir1, dir2) != 0 || dot(dir1, dir3) != 0 || dot(dir2, dir3) != 0)
      throw std::invalid_argument(
          "The given coordinate system for a subvalley is not orthogonal, adapt that.");
    double *ax = this->desc_.axes[idxSubValley];
    for (int i = 0; i < 3; i++) {
      ax[i] = dir1[i];
      ax[3 + i] = dir2[i];
      ax[6 + i] = dir3[i];
    }
  }
};

/// emcParabolicAnisotropValley.hpp:46-54
template <class T>
class emcParabolicAnisotropValley : public emcAnisotropValleyBase<T> {
public:
  emcParabolicAnisotropValley(std::array<T, 3> relEffMass, T particleMass, SizeType degFactor,
                              T bottom = 0.) {
    this->init(EMCGPU_VALLEY_PARA_ANISO, relEffMass, particleMass, degFactor, 0, bottom);
  }
};

/// emcNonParabolicAnistropValley.hpp:47-58
template <class T>
class emcNonParabolicAnisotropValley : public emcAnisotropValleyBase<T> {
public:
  emcNonParabolicAnisotropValley(std::array<T, 3> relEffMass, T particleMass, SizeType degFactor,
                                 T alpha, T bottom = 0.) {
    this->init(EMCGPU_VALLEY_NONPARA_ANISO, relEffMass, particleMass, degFactor, alpha, bottom);
  }
};

}  // namespace emcgpu

#endif
