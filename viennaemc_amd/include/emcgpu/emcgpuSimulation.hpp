// emcgpuSimulation.hpp -- 2-D device simulations of the GPU host API.
//
// The classes keep the reference's device-simulation interface:
//   * emcgpuDevice2D: emcDevice<T, 2> (include/emcDevice.hpp:86-190:
//     extent, spacing, temperature, setDeviceWidth, addConstantDopingRegion,
//     addOhmicContact);
//   * emcgpuNGPScheme / emcgpuCICScheme / emcgpuNECScheme: the PM scheme
//     classes (include/PMSchemes/*), chosen as a template argument;
//   * emcgpuPoissonSolver: takes the place of emcSORSolver
//     (include/PoissonSolver/emcSORSolver.hpp); it solves the same discrete
//     equations with Newton + FP64 conjugate gradients and an FP32 multigrid
//     preconditioner to a relative residual, not by SOR sweeps to a
//     potential change;
//   * emcgpuSimulationParameter: emcSimulationParameter
//     (include/emcSimulationParameter.hpp: times, seed, name prefix, steps
//     between progress lines and for the final averages, particle types);
//   * emcgpuSimulation: emcSimulation (include/emcSimulation.hpp:106-193):
//     execute() computes the equilibrium, generates the particles, runs the
//     steps on the GPU (Poisson, E-field, drift / scatter, ohmic contacts,
//     charge assignment), tallies the contact currents of the
//     non-transient steps (emcSimulationResults::updateCurrent,
//     include/emcSimulationResults.hpp:118-147), averages potential and
//     concentration over the last steps and writes the reference's result
//     files; setPoissonInterval and setChannelCurrentRegion /
//     getAvgDriftCurrent as in the reference.
// One particle type (electrons) and one carrier per simulated particle.
#ifndef EMCGPU_SIMULATION_HPP
#define EMCGPU_SIMULATION_HPP

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdio>
#include <iostream>
#include <memory>
#include <set>
#include <string>
#include <vector>

#include "emcgpuBulkParticleHandler.hpp"
#include "emcgpuOutput.hpp"

namespace emcgpu {

/// emcBoundaryPos (include/emcBoundaryPos.hpp)
enum class emcBoundaryPos { XMIN = EMCGPU_XMIN, XMAX = EMCGPU_XMAX, YMIN = EMCGPU_YMIN,
                            YMAX = EMCGPU_YMAX };

/// 2-D device geometry, doping and ohmic contacts (emcDevice<T, 2>)
template <class T>
class emcgpuDevice2D {
public:
  typedef std::array<T, 2> ValueVec;
  typedef typename emcgpuBulkDevice<T>::Material Material;

  emcgpuDevice2D(const Material &material, const ValueVec &maxPos, const ValueVec &spacing,
                 T temperature = 300.)
      : mat(material), maxP(maxPos), h(spacing), temp(temperature) {}

  void setDeviceWidth(T w) { width = w; }
  void addConstantDopingRegion(const ValueVec &minPos, const ValueVec &maxPos, T doping) {
    emcgpu_doping_region r{};
    r.min_pos[0] = minPos[0];
    r.min_pos[1] = minPos[1];
    r.max_pos[0] = maxPos[0];
    r.max_pos[1] = maxPos[1];
    r.doping = doping;
    regions.push_back(r);
  }
  /// contact on a boundary over [minPos, maxPos] of the coordinate along it
  void addOhmicContact(emcBoundaryPos pos, T voltage, const std::array<T, 1> &minPos,
                       const std::array<T, 1> &maxPos) {
    emcgpu_contact c{};
    c.boundary = (int)pos;
    c.voltage = voltage;
    c.min_pos = minPos[0];
    c.max_pos = maxPos[0];
    contacts.push_back(c);
  }
  /// gate over [minPos, maxPos] of a boundary (emcDevice.hpp:195-202): oxide
  /// relative permittivity, oxide thickness [m], barrier height [V]
  void addGateContact(emcBoundaryPos pos, T voltage, const std::array<T, 1> &minPos,
                      const std::array<T, 1> &maxPos, T epsRoxide, T thickness,
                      T barrierHeight) {
    emcgpu_contact c{};
    c.boundary = (int)pos;
    c.voltage = voltage;
    c.min_pos = minPos[0];
    c.max_pos = maxPos[0];
    c.type = EMCGPU_CONTACT_GATE;
    c.eps_oxide = epsRoxide;
    c.thickness = thickness;
    c.barrier_height = barrierHeight;
    contacts.push_back(c);
  }

  const Material &getMaterial() const { return mat; }
  T getTemperature() const { return temp; }
  T getThermalVoltage() const { return phys::kB / phys::q * temp; }
  ValueVec getMaxPos() const { return maxP; }
  ValueVec getSpacing() const { return h; }
  T getDeviceWidth() const { return width; }
  SizeType getNrContacts() const { return contacts.size(); }

  /// the C descriptor (points into this object)
  emcgpu_device_desc descriptor(int pm, double tol, int maxIter) const {
    emcgpu_device_desc d{};
    d.max_pos[0] = maxP[0];
    d.max_pos[1] = maxP[1];
    d.spacing[0] = h[0];
    d.spacing[1] = h[1];
    d.width = width;
    d.temperature = temp;
    d.eps_r = mat.getEpsR();
    d.ni = mat.getNi();
    d.n_doping = (int)regions.size();
    d.doping = regions.data();
    d.n_contacts = (int)contacts.size();
    d.contacts = contacts.data();
    d.pm_scheme = pm;
    d.poisson_tol = tol;
    d.poisson_max_iter = maxIter;
    return d;
  }

private:
  Material mat;
  ValueVec maxP, h;
  T temp;
  T width = 1e-6;
  std::vector<emcgpu_doping_region> regions;
  std::vector<emcgpu_contact> contacts;
};

/// particle-mesh schemes (include/PMSchemes/emcNGPScheme.hpp,
/// emcCICScheme.hpp, emcNECScheme.hpp)
template <class T>
struct emcgpuNGPScheme {
  static constexpr int id = EMCGPU_PM_NGP;
};
template <class T>
struct emcgpuCICScheme {
  static constexpr int id = EMCGPU_PM_CIC;
};
template <class T>
struct emcgpuNECScheme {
  static constexpr int id = EMCGPU_PM_NEC;
};

/// Poisson solver of the GPU path: stops at
/// max |F| <= relTolerance * max h^2 (|C| + n + p), F the discrete equation
/// of emcSORSolver, after at most maxIter Newton / refinement iterations
template <class T>
class emcgpuPoissonSolver {
public:
  template <class DeviceType>
  emcgpuPoissonSolver(const DeviceType &, T relTolerance = 1e-10, int maxIter = 50)
      : tol(relTolerance), iters(maxIter) {}
  T getTolerance() const { return tol; }
  int getMaxIter() const { return iters; }

private:
  T tol;
  int iters;
};

template <class T>
class emcgpuSimulationParameter {
public:
  enum Precision { FP64 = EMCGPU_FP64, FP32 = EMCGPU_FP32 };

  emcgpuSimulationParameter() : emcgpuSimulationParameter(10e-12, 1.5e-16, 3e-12) {}
  emcgpuSimulationParameter(T inSimTime, T inStepTime, T inTransientTime) {
    setTimes(inSimTime, inStepTime, inTransientTime);
  }

  /// emcSimulationParameter::setTimes with its checks
  /// (include/emcSimulationParameter.hpp:175-193)
  void setTimes(T inSimTime, T inStepTime, T inTransientTime) {
    if (inStepTime < 0 || inSimTime < 0 || inTransientTime < 0)
      reportError("The time parameter can't be negative!");
    if (inStepTime == 0)
      reportError("Step Time can't be zero!");
    if (inStepTime > inSimTime || inTransientTime > inSimTime)
      reportError("stepTime, avgTime and transientTime have to be smaller than simTime!");
    simTime = inSimTime;
    stepTime = inStepTime;
    transientTime = inTransientTime;
  }
  void setSeed(SizeType s) { seedRNG = s; }
  void setNamePrefix(const std::string &p) { namePrefix = p; }
  void setNrStepsBetweenShowProgress(SizeType n) { nrStepsBetweenShowProgress = n; }
  void setNrStepsForFinalAvg(SizeType n) { nrStepsForFinalAvg = n; }
  void setNrCarriersPerPart(SizeType n) {
    if (n != 1)
      reportError("the GPU device path simulates one carrier per particle");
  }
  void setAdaptPotentialForWriteFunction(T (*f)(const T &, const emcgpuDevice2D<T> &)) {
    adaptPotentialForWrite = f;
  }
  void setPrecision(Precision p) { precision = p; }
  void setGpu(int g) { gpu = g; }
  void addParticleType(std::unique_ptr<emcgpuParticleType<T>> &&type) {
    if (particleType)
      reportError("the GPU device path simulates one particle type");
    particleType = std::move(type);
  }

  SizeType getNrSteps() const { return (SizeType)std::ceil(simTime / stepTime); }
  SizeType getNrTransientSteps() const { return (SizeType)std::ceil(transientTime / stepTime); }
  SizeType getNrNonTransientSteps() const { return getNrSteps() - getNrTransientSteps(); }
  bool isTransientStep(SizeType s) const { return s < getNrTransientSteps(); }
  T getStepTime() const { return stepTime; }
  T getTransientTime() const { return transientTime; }

  /// emcSimulationParameter::check (include/emcSimulationParameter.hpp:157-174, 203-209)
  void check() const {
    if (!particleType)
      reportError("Add at least one particleType to simulation parameter!");
    if (nrStepsForFinalAvg > getNrSteps())
      reportError("nrStepsForFinalAvg has to be smaller than total step nr.!");
  }

private:
  T simTime = 0, stepTime = 0, transientTime = 0;
  SizeType nrStepsBetweenShowProgress = 100, nrStepsForFinalAvg = 100;
  std::string namePrefix;
  T (*adaptPotentialForWrite)(const T &, const emcgpuDevice2D<T> &) = nullptr;
  std::unique_ptr<emcgpuParticleType<T>> particleType;
  SizeType seedRNG = 1;
  Precision precision = FP32;
  int gpu = 0;

  template <class, class>
  friend class emcgpuSimulation;
};

template <class T, class PMScheme>
class emcgpuSimulation {
public:
  emcgpuSimulation(emcgpuSimulationParameter<T> &inParam, emcgpuDevice2D<T> &inDevice,
                   emcgpuPoissonSolver<T> &inSolver, PMScheme = PMScheme())
      : param(inParam), device(inDevice), solver(inSolver) {
    param.check();
  }
  ~emcgpuSimulation() {
    if (ctx)
      emcgpu_destroy(ctx);
  }
  emcgpuSimulation(const emcgpuSimulation &) = delete;
  emcgpuSimulation &operator=(const emcgpuSimulation &) = delete;

  /// drift-velocity (Ramo) current over x0 <= x <= x1 (emcSimulation.hpp:64-70)
  void setChannelCurrentRegion(T x0, T x1, T length) {
    chan = {x0, x1, length};
    trackChannel = true;
  }
  /// mean of the channel current over the non-transient steps [A]
  T getAvgDriftCurrent() const { return driftCount ? driftSum / (T)driftCount : 0; }
  /// Poisson and E-field every n-th step (emcSimulation.hpp:84-86)
  void setPoissonInterval(SizeType n) { poissonInterval = n < 1 ? 1 : n; }

  /// emcSimulation::execute (include/emcSimulation.hpp:106-126)
  void execute() {
    setup();
    std::cout << "Equilibrium Characteristics ..." << std::endl;
    size_t n0 = 0;
    check(emcgpu_device_init(ctx, param.seedRNG, &n0), ctx, "emcgpu_device_init");
    std::cout << "\tNr. of particles: " << n0 << std::endl;
    writeCurrentResults("Eq");
    std::cout << "Monte Carlo Procedure ..." << std::endl;
    const SizeType total = param.getNrSteps();
    const SizeType nTr = std::min(param.getNrTransientSteps(), total);
    const SizeType avgFrom = total - std::min<SizeType>(param.nrStepsForFinalAvg, total);
    // calls end at the transient end, the start of the averaging and every
    // progress line
    std::set<SizeType> cuts = {0, nTr, avgFrom, total};
    const SizeType every = std::max<SizeType>(param.nrStepsBetweenShowProgress, 1);
    for (SizeType s = 0; s < total; s += every)
      cuts.insert(s);
    std::vector<SizeType> c(cuts.begin(), cuts.end());
    netto.clear();
    current.clear();
    driftSum = 0;
    driftCount = 0;
    for (size_t i = 0; i + 1 < c.size(); i++) {
      const SizeType a = c[i], b = c[i + 1];
      if (a % every == 0)
        std::cout << "\tNr. Iteration: \t\t" << a << " / " << total << std::endl;
      runSteps((int)(b - a), a >= avgFrom, a >= nTr);
    }
    std::cout << "\tNr. Iteration: \t\t" << total << " / " << total << std::endl;
    writeFinalResults();
  }

  /// netto particles (left minus injected) and the running-mean current [A]
  /// of every non-transient step and contact
  const std::vector<std::vector<int>> &getNettoParticles() const { return netto; }
  const std::vector<std::vector<double>> &getCurrent() const { return current; }
  emcgpu_ctx *context() { return ctx; }

private:
  void setup() {
    if (ctx)
      return;
    auto &type = *param.particleType;
    emcgpu_device_desc d =
        device.descriptor(PMScheme::id, solver.getTolerance(), solver.getMaxIter());
    d.init_from_potential = type.usesPotentialForInit() ? 1 : 0;
    // mechanisms registered for some doping regions only: one instance per
    // region holds that region's tables (the (valley, region) table sets of
    // emcScatterHandler)
    const int nRegions = d.n_doping;
    const bool perRegion = nRegions > 1 && type.regionDependent(nRegions);
    const int nInst = perRegion ? nRegions : 1;
    emcgpu_config cfg{param.gpu, (int)param.precision, nInst};
    check(emcgpu_create(&cfg, &ctx), nullptr, "emcgpu_create");
    auto vd = type.valleyDescriptors();
    for (int i = 0; i < nInst; i++) {
      auto md = perRegion ? type.mechanismDescriptors(i) : type.mechanismDescriptors();
      check(emcgpu_set_model(ctx, i, vd.data(), (int)vd.size(), md.data(), (int)md.size()), ctx,
            "emcgpu_set_model");
      emcgpu_screening_desc sd;
      if (type.screeningDescriptor(sd))
        check(emcgpu_set_screening(ctx, i, &sd), ctx, "emcgpu_set_screening");
      auto rates = perRegion ? type.rawRates(i) : type.rawRates();
      check(emcgpu_set_rate_tables(ctx, i, (int)type.getNrEnergyLevels(), type.getMaxEnergy(),
                                   rates.data()),
            ctx, "emcgpu_set_rate_tables");
    }
    d.region_tables = perRegion ? 1 : 0;
    check(emcgpu_device_setup(ctx, &d), ctx, "emcgpu_device_setup");
    check(emcgpu_device_info(ctx, ext, scal), ctx, "emcgpu_device_info");
    nContacts = (int)device.getNrContacts();
  }

  // the next n steps (the GPU's step counter runs along); accumulate: add
  // them to the final averages; counted: non-transient
  void runSteps(int n, bool accumulate, bool counted) {
    if (n <= 0)
      return;
    std::vector<int32_t> ex((size_t)n * nContacts), del(ex.size()), inj(ex.size());
    std::vector<double> chanCur(n);
    emcgpu_device_step_args a{};
    a.dt = param.stepTime;
    a.n_steps = n;
    a.seed = param.seedRNG;
    a.use_graph = 1;
    a.exits = ex.data();
    a.deleted = del.data();
    a.injected = inj.data();
    a.poisson_interval = (int)poissonInterval;
    a.accumulate = accumulate ? 1 : 0;
    if (trackChannel) {
      a.channel_x0 = chan[0];
      a.channel_x1 = chan[1];
      a.channel_length = chan[2];
      a.channel_current = chanCur.data();
    }
    check(emcgpu_device_step(ctx, &a), ctx, "emcgpu_device_step");
    if (!counted)
      return;
    // updateCurrent (emcSimulationResults.hpp:118-147): netto = left -
    // (injected - deleted); current = running mean * charge / stepTime
    const double factor = param.particleType->getCharge() / param.stepTime;
    if (nettoSum.empty())
      nettoSum.assign(nContacts, 0.0);
    for (int s = 0; s < n; s++) {
      std::vector<int> nt(nContacts);
      std::vector<double> cur(nContacts);
      for (int k = 0; k < nContacts; k++) {
        const size_t o = (size_t)s * nContacts + k;
        nt[k] = ex[o] - (inj[o] - del[o]);
        nettoSum[k] += nt[k];
        cur[k] = nettoSum[k] / (double)(netto.size() + 1) * factor;
      }
      netto.push_back(nt);
      current.push_back(cur);
      if (trackChannel) {
        driftSum += chanCur[s];
        driftCount++;
      }
    }
  }

  std::string prefix() const { return param.namePrefix; }

  T potentialForWrite(double v) const {
    if (param.adaptPotentialForWrite)
      return param.adaptPotentialForWrite((T)v, device);
    return (T)(v * scal[0]);  // undoNormalizeVoltage
  }

  void writeGrid(const std::string &name, const std::vector<double> &g,
                 const std::function<double(double)> &f) const {
    auto os = output::openResultFile(name);
    output::printGrid(os, (size_t)ext[0], (size_t)ext[1], g.data(), f);
  }

  void writeParticles(const std::string &suffix) const {
    size_t n = 0;
    check(emcgpu_device_get_particles(ctx, &n, nullptr, nullptr, nullptr, nullptr, nullptr,
                                      nullptr),
          ctx, "emcgpu_device_get_particles");
    std::vector<double> pos(2 * n), k(3 * n), E(n), tau(n);
    std::vector<int32_t> v(n), sub(n);
    check(emcgpu_device_get_particles(ctx, &n, pos.data(), k.data(), E.data(), tau.data(),
                                      v.data(), sub.data()),
          ctx, "emcgpu_device_get_particles");
    const auto mp = device.getMaxPos();
    const double maxPos[2] = {mp[0], mp[1]};
    auto os = output::openResultFile(prefix() + param.particleType->getName() + suffix);
    output::printParticles(os, maxPos, n, pos.data(), k.data(), E.data(), sub.data(), v.data(),
                           tau.data());
  }

  /// writeCurrentResults (include/emcSimulationResults.hpp:152-172) and the
  /// particle dump
  void writeCurrentResults(const std::string &suffix) const {
    const size_t nn = (size_t)ext[0] * ext[1];
    std::vector<double> pot(nn), ex_(nn), ey_(nn), conc(nn);
    check(emcgpu_device_fields(ctx, pot.data(), ex_.data(), ey_.data(), conc.data(), nullptr),
          ctx, "emcgpu_device_fields");
    const double ni = device.getMaterial().getNi();
    writeGrid(prefix() + "Potential" + suffix, pot, [this](double v) {
      return (double)potentialForWrite(v);
    });
    writeGrid(prefix() + param.particleType->getName() + "Conc" + suffix, conc,
              [ni](double v) { return v * ni; });
    writeGrid(prefix() + "EFieldX" + suffix, ex_, nullptr);
    writeGrid(prefix() + "EFieldY" + suffix, ey_, nullptr);
    writeParticles(suffix);
  }

  /// writeFinalResults (include/emcSimulationResults.hpp:176-202)
  void writeFinalResults() const {
    const size_t nn = (size_t)ext[0] * ext[1];
    std::vector<double> pot(nn), conc(nn);
    long long nAvg = 0;
    check(emcgpu_device_averages(ctx, pot.data(), conc.data(), &nAvg, 0), ctx,
          "emcgpu_device_averages");
    const std::string name = param.particleType->getName();
    {
      auto os = output::openResultFile(prefix() + name + "Current");
      output::printCurrent(os, netto, current, param.stepTime, param.transientTime);
    }
    const double ni = device.getMaterial().getNi();
    writeGrid(prefix() + name + "ConcAvg", conc, [ni](double v) { return v * ni; });
    writeGrid(prefix() + "PotentialAvg", pot, [this](double v) {
      return (double)potentialForWrite(v);
    });
    writeParticles("Final");
  }

  emcgpuSimulationParameter<T> &param;
  emcgpuDevice2D<T> &device;
  emcgpuPoissonSolver<T> &solver;
  emcgpu_ctx *ctx = nullptr;
  int ext[2] = {0, 0};
  double scal[4] = {0, 0, 0, 0};
  int nContacts = 0;
  SizeType poissonInterval = 1;
  bool trackChannel = false;
  std::array<T, 3> chan{{0, 0, 1}};
  double driftSum = 0;
  SizeType driftCount = 0;
  std::vector<double> nettoSum;
  std::vector<std::vector<int>> netto;
  std::vector<std::vector<double>> current;
};

}  // namespace emcgpu

#endif
