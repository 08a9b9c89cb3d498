/*
 * emc_oracle.c -- CPU restatement of ViennaEMC's per-particle bulk EMC step.
 *
 * TEST INFRASTRUCTURE ONLY (the checker).  Loaded by tests/, by
 * __graft_entry__.smoke() and by bench.py's cpu_baseline leg; never by the
 * product library.
 *
 * Each function restates the reference algorithm with the same FP64
 * operation order so that, built with gcc on x86-64 without FMA
 * contraction, it reproduces the reference bit for bit.  File:line
 * citations are relative to the ViennaEMC source tree.
 */
#include "emc_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

/* ======================================================================
 * Random numbers
 * ====================================================================== */

/* std::mt19937_64 as fixed by [rand.predef]: w=64 n=312 m=156 r=31
 * a=0xB5026F5AA96619E9 u=29 d=0x5555555555555555 s=17 b=0x71D67FFFEDA60000
 * t=37 c=0xFFF7EEE000000000 l=43 f=6364136223846793005.
 * The engine behind emcRNG (include/emcUtil.hpp:15). */
void orc_mt64_seed(orc_mt64 *g, uint64_t seed) {
  g->mt[0] = seed;
  for (int i = 1; i < 312; i++)
    g->mt[i] = 6364136223846793005ULL * (g->mt[i - 1] ^ (g->mt[i - 1] >> 62)) +
               (uint64_t)i;
  g->mti = 312;
}

static void orc_mt64_twist(orc_mt64 *g) {
  const uint64_t upper = 0xFFFFFFFF80000000ULL, lower = 0x7FFFFFFFULL;
  const uint64_t a = 0xB5026F5AA96619E9ULL;
  for (int i = 0; i < 312; i++) {
    uint64_t y = (g->mt[i] & upper) | (g->mt[(i + 1) % 312] & lower);
    uint64_t v = g->mt[(i + 156) % 312] ^ (y >> 1);
    if (y & 1ULL)
      v ^= a;
    g->mt[i] = v;
  }
  g->mti = 0;
}

uint64_t orc_mt64_next(orc_mt64 *g) {
  if (g->mti >= 312)
    orc_mt64_twist(g);
  uint64_t x = g->mt[g->mti++];
  x ^= (x >> 29) & 0x5555555555555555ULL;
  x ^= (x << 17) & 0x71D67FFFEDA60000ULL;
  x ^= (x << 37) & 0xFFF7EEE000000000ULL;
  x ^= x >> 43;
  return x;
}

void orc_mt64_fill(uint64_t seed, uint64_t *out, size_t n) {
  orc_mt64 g;
  orc_mt64_seed(&g, seed);
  for (size_t i = 0; i < n; i++)
    out[i] = orc_mt64_next(&g);
}

uint64_t orc_splitmix64(uint64_t x) {
  uint64_t z = x + 0x9E3779B97F4A7C15ULL;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}

uint64_t orc_particle_seed(uint64_t seed, uint64_t pid) {
  return orc_splitmix64(seed ^ pid);
}

void orc_mt64_fill_particles(uint64_t seed, uint64_t p0, size_t np,
                             uint64_t *out, size_t n, size_t stride) {
  for (size_t p = 0; p < np; p++)
    orc_mt64_fill(orc_particle_seed(seed, p0 + p), out + p * stride, n);
}

/* libstdc++ generate_canonical<double, 53>(mt19937_64): one draw,
 * double(x) / 2^64, clamped to nextafter(1, 0) (bits/random.tcc:3348). */
double orc_canonical(uint64_t x) {
  double r = (double)x / 18446744073709551616.0;
  if (r >= 1.0)
    r = nextafter(1.0, 0.0);
  return r;
}

void orc_rng_init_mt(orc_rng *r, uint64_t seed) {
  memset(r, 0, sizeof(*r));
  r->kind = 0;
  orc_mt64_seed(&r->mt, seed);
}

void orc_rng_init_buf(orc_rng *r, const uint64_t *buf, size_t len) {
  memset(r, 0, sizeof(*r));
  r->kind = 1;
  r->buf = buf;
  r->len = len;
}

uint64_t orc_rng_raw(orc_rng *r) {
  r->ndraws++;
  if (r->kind == 0)
    return orc_mt64_next(&r->mt);
  if (r->pos >= r->len) {
    r->overflow = 1;
    return 0;
  }
  return r->buf[r->pos++];
}

/* uniform_real_distribution: canonical * (b - a) + a (bits/random.h) */
double orc_rng_uniform(orc_rng *r, double a, double b) {
  double c = orc_canonical(orc_rng_raw(r));
  return c * (b - a) + a;
}

/* Philox-4x32-10, Salmon, Moraes, Dror, Shaw, "Parallel random numbers: as
 * easy as 1, 2, 3", SC'11: multipliers 0xD2511F53 / 0xCD9E8D57, Weyl key
 * increments 0x9E3779B9 / 0xBB67AE85. */
void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2],
                       uint32_t out[4]) {
  uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
  uint32_t k0 = key[0], k1 = key[1];
  for (int r = 0; r < 10; r++) {
    uint64_t p0 = (uint64_t)0xD2511F53u * c0;
    uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
    uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
    uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
    uint32_t n0 = hi1 ^ c1 ^ k0;
    uint32_t n1 = lo1;
    uint32_t n2 = hi0 ^ c3 ^ k1;
    uint32_t n3 = lo0;
    c0 = n0;
    c1 = n1;
    c2 = n2;
    c3 = n3;
    k0 += 0x9E3779B9u;
    k1 += 0xBB67AE85u;
  }
  out[0] = c0;
  out[1] = c1;
  out[2] = c2;
  out[3] = c3;
}


/* ======================================================================
 * Vector helpers: include/emcUtil.hpp:25-175
 * ====================================================================== */

/* square(): res = 0; res += v_i * v_i in order (emcUtil.hpp:25-31) */
static double orc_square3(const double v[3]) {
  double res = 0;
  res += v[0] * v[0];
  res += v[1] * v[1];
  res += v[2] * v[2];
  return res;
}

static double orc_norm3(const double v[3]) { return sqrt(orc_square3(v)); }

/* initRandomDirection (emcUtil.hpp:131-139) */
void orc_init_random_direction(double norm, double rand1, double rand2,
                               double res[3]) {
  double phi = 2 * ORC_PI * rand1;
  double cosTheta = 1 - 2 * rand2;
  res[0] = norm * sqrt(1 - cosTheta * cosTheta) * cos(phi);
  res[1] = norm * sqrt(1 - cosTheta * cosTheta) * sin(phi);
  res[2] = norm * cosTheta;
}

/* initRandomDirectionWithRespectToCurrentK (emcUtil.hpp:141-175) */
static void orc_init_random_direction_wrt_k(const double k[3],
                                            double cosTheta, double rand,
                                            double res[3]) {
  double kxy = sqrt(k[0] * k[0] + k[1] * k[1]);
  double normK = sqrt(kxy * kxy + k[2] * k[2]);
  if (normK == 0.0) {
    res[0] = res[1] = res[2] = 0.0;
    return;
  }
  double ct0 = k[2] / normK;
  double st0 = kxy / normK;
  double cfi0 = (kxy > 0.0) ? k[0] / kxy : 1.0;
  double sfi0 = (kxy > 0.0) ? k[1] / kxy : 0.0;
  double st = sqrt(1.0 - cosTheta * cosTheta);
  double phi = 2.0 * ORC_PI * rand;
  double kxp = normK * st * cos(phi);
  double kyp = normK * st * sin(phi);
  double kzp = normK * cosTheta;
  res[0] = kxp * cfi0 * ct0 - kyp * sfi0 + kzp * cfi0 * st0;
  res[1] = kxp * sfi0 * ct0 + kyp * cfi0 + kzp * sfi0 * st0;
  res[2] = -kxp * st0 + kzp * ct0;
}

/* ======================================================================
 * Valleys: include/ValleyTypes/
 * ====================================================================== */

static int orc_is_aniso(const orc_valley *v) {
  return v->kind == ORC_PARA_ANISO || v->kind == ORC_NONPARA_ANISO;
}

void orc_valley_init(orc_valley *v, int kind, const double relMass[3],
                     double particleMass, int deg, double alpha,
                     double bottom) {
  memset(v, 0, sizeof(*v));
  v->kind = kind;
  v->relMass[0] = relMass[0];
  v->relMass[1] = relMass[1];
  v->relMass[2] = relMass[2];
  v->particleMass = particleMass;
  v->deg = deg;
  v->alpha = (kind == ORC_NONPARA_ISO || kind == ORC_NONPARA_ANISO) ? alpha : 0.0;
  v->bottom = bottom;
  /* identity rotation per subvalley (emcNonParabolicAnistropValley.hpp:174-183) */
  for (int s = 0; s < ORC_MAX_DEG; s++) {
    for (int i = 0; i < 9; i++)
      v->rot[s][i] = 0.0;
    v->rot[s][0] = v->rot[s][4] = v->rot[s][8] = 1.0;
  }
  if (orc_is_aniso(v)) {
    /* initEffMasses (emcNonParabolicAnistropValley.hpp:185-194) */
    double dos = 1.;
    for (int i = 0; i < 3; i++)
      dos = dos * relMass[i];
    v->effMassDOS = pow(dos, 1. / 3.) * particleMass;
    double acc = 0.;
    for (int i = 0; i < 3; i++)
      acc += 1. / relMass[i];
    v->effMassCond = 3. * particleMass / acc;
    /* initVogtFactor (:198-204): sqrt(getEffMassCond() / (m_i * m0)) */
    for (int i = 0; i < 3; i++)
      v->vogt[i] = sqrt(v->effMassCond / (relMass[i] * particleMass));
  } else {
    /* emcParabolicIsotropValley.hpp:38-44, emcNonParabolicIsotropValley.hpp:44-50 */
    v->effMassDOS = relMass[0] * particleMass;
    v->effMassCond = relMass[0] * particleMass;
    v->vogt[0] = v->vogt[1] = v->vogt[2] = 1.0;
  }
}

static void orc_normalize3(double d[3]) {
  double n = orc_norm3(d);
  if (n == 0.0)
    return;
  for (int i = 0; i < 3; i++)
    d[i] /= n;
}

/* setSubValleyEllipseCoordSystem (emcNonParabolicAnistropValley.hpp:68-80) */
void orc_valley_set_axes(orc_valley *v, int sub, const double d1[3],
                         const double d2[3], const double d3[3]) {
  double a[3], b[3], c[3];
  memcpy(a, d1, sizeof a);
  memcpy(b, d2, sizeof b);
  memcpy(c, d3, sizeof c);
  orc_normalize3(a);
  orc_normalize3(b);
  orc_normalize3(c);
  for (int i = 0; i < 3; i++) {
    v->rot[sub][i] = a[i];
    v->rot[sub][3 + i] = b[i];
    v->rot[sub][6 + i] = c[i];
  }
}

/* getGamma: E for parabolic, E(1+aE) for Kane (…Anistrop…:123, …Isotrop…:104) */
double orc_valley_gamma(const orc_valley *v, double E) {
  if (v->kind == ORC_PARA_ISO || v->kind == ORC_PARA_ANISO)
    return E;
  return E * (1 + v->alpha * E);
}

double orc_valley_np(const orc_valley *v) { return v->alpha; }

/* getEnergy (emcNonParabolicAnistropValley.hpp:109-113, parabolic :95-98) */
double orc_valley_energy(const orc_valley *v, const double k[3]) {
  if (v->kind == ORC_PARA_ISO || v->kind == ORC_PARA_ANISO)
    return ORC_HBAR * ORC_HBAR * orc_square3(k) / (2 * v->effMassCond * ORC_Q);
  double gamma = ORC_HBAR * ORC_HBAR * orc_square3(k) / (v->effMassCond * ORC_Q);
  return gamma / (1 + sqrt(1 + 2 * v->alpha * gamma));
}

/* getNormWaveVec (…Anistrop…:102-105, parabolic :90-92, iso variants) */
double orc_valley_normk(const orc_valley *v, double E) {
  switch (v->kind) {
  case ORC_PARA_ISO:
  case ORC_PARA_ANISO:
    return sqrt(2 * v->effMassCond * ORC_Q * E) / ORC_HBAR;
  case ORC_NONPARA_ISO:
    return sqrt(2 * v->effMassCond * ORC_Q * orc_valley_gamma(v, E)) / ORC_HBAR;
  default:
    return sqrt(2 * v->effMassCond * orc_valley_gamma(v, E) * ORC_Q) / ORC_HBAR;
  }
}

/* getEffMassCond (…Anistrop…:90-92) */
double orc_valley_mcond(const orc_valley *v, double E) {
  if (v->kind == ORC_PARA_ISO || v->kind == ORC_PARA_ANISO)
    return v->effMassCond;
  return v->effMassCond * (1 + 2 * E * v->alpha);
}

/* getEffMassDOS (…Anistrop…:84-86) */
double orc_valley_mdos(const orc_valley *v, double E) {
  if (v->kind == ORC_PARA_ISO || v->kind == ORC_PARA_ANISO)
    return v->effMassDOS;
  return v->effMassDOS * pow(1 + 2 * v->alpha * E, 3.);
}

/* transformToEllipseCoord (…Anistrop…:140-154) */
void orc_valley_to_ellipse(const orc_valley *v, int sub, const double in[3],
                           double out[3]) {
  if (!orc_is_aniso(v)) {
    out[0] = in[0];
    out[1] = in[1];
    out[2] = in[2];
    return;
  }
  const double *r = v->rot[sub];
  double o0 = in[0] * r[0] + in[1] * r[1] + in[2] * r[2];
  double o1 = in[0] * r[3] + in[1] * r[4] + in[2] * r[5];
  double o2 = in[0] * r[6] + in[1] * r[7] + in[2] * r[8];
  out[0] = o0;
  out[1] = o1;
  out[2] = o2;
}

/* transformToDeviceCoord (…Anistrop…:157-170) */
void orc_valley_to_device(const orc_valley *v, int sub, const double in[3],
                          double out[3]) {
  if (!orc_is_aniso(v)) {
    out[0] = in[0];
    out[1] = in[1];
    out[2] = in[2];
    return;
  }
  const double *r = v->rot[sub];
  double o0 = in[0] * r[0] + in[1] * r[3] + in[2] * r[6];
  double o1 = in[0] * r[1] + in[1] * r[4] + in[2] * r[7];
  double o2 = in[0] * r[2] + in[1] * r[5] + in[2] * r[8];
  out[0] = o0;
  out[1] = o1;
  out[2] = o2;
}

/* getVelocity (…Anistrop…:126-136; parabolic aniso :103-111; iso variants) */
void orc_valley_velocity(const orc_valley *v, const double k[3], double E,
                         int sub, double vel[3]) {
  switch (v->kind) {
  case ORC_PARA_ISO: {
    double f = ORC_HBAR / v->effMassCond;
    vel[0] = k[0] * f;
    vel[1] = k[1] * f;
    vel[2] = k[2] * f;
    return;
  }
  case ORC_NONPARA_ISO: {
    double f = ORC_HBAR /
               (v->effMassCond * sqrt(1 + 4 * v->alpha * orc_valley_gamma(v, E)));
    vel[0] = k[0] * f;
    vel[1] = k[1] * f;
    vel[2] = k[2] * f;
    return;
  }
  case ORC_PARA_ANISO: {
    double e[3];
    orc_valley_to_ellipse(v, sub, k, e);
    for (int i = 0; i < 3; i++)
      e[i] = ORC_HBAR * v->vogt[i] * e[i] / v->effMassCond;
    orc_valley_to_device(v, sub, e, vel);
    return;
  }
  default: {
    double e[3];
    orc_valley_to_ellipse(v, sub, k, e);
    double npf = sqrt(1 + 4 * v->alpha * orc_valley_gamma(v, E));
    for (int i = 0; i < 3; i++)
      e[i] = ORC_HBAR * v->vogt[i] * e[i] / (v->effMassCond * npf);
    orc_valley_to_device(v, sub, e, vel);
    return;
  }
  }
}

/* ======================================================================
 * Free flight: include/emcParticleDrift.hpp:11-35
 * ====================================================================== */
void orc_drift(double dt, orc_particle *p, const orc_valley *v, double *pos,
               int dim, const double force[3]) {
  double kOld[3], kNew[3], fE[3];
  orc_valley_to_ellipse(v, p->subValley, p->k, kOld);
  memcpy(kNew, kOld, sizeof kNew);
  orc_valley_to_ellipse(v, p->subValley, force, fE);
  for (int i = 0; i < 3; i++)
    kNew[i] += fE[i] * dt * v->vogt[i] / ORC_HBAR;
  orc_valley_to_device(v, p->subValley, kNew, p->k);
  p->energy = orc_valley_energy(v, p->k);
  if (pos) {
    double dPos[3], dDev[3];
    for (int i = 0; i < 3; i++) {
      double avgK = (kNew[i] + kOld[i]) / 2;
      dPos[i] = ORC_HBAR * v->vogt[i] * avgK * dt / orc_valley_mcond(v, p->energy);
    }
    orc_valley_to_device(v, p->subValley, dPos, dDev);
    for (int i = 0; i < dim; i++)
      pos[i] = pos[i] + dDev[i];
  }
}

/* ======================================================================
 * Scatter mechanisms
 * ====================================================================== */

/* emcAcousticScatterMechanism.hpp:36-45 */
void orc_mech_acoustic(orc_mech *m, int valley, double sigma, double rho,
                       double vSound, double T) {
  memset(m, 0, sizeof(*m));
  m->kind = ORC_MECH_ACOUSTIC;
  m->valley = valley;
  m->finalValley = valley;
  double cL = rho * pow(vSound, 2);
  m->scatterConst = sqrt(2.0 * ORC_Q) * pow(sigma * ORC_Q, 2) * ORC_KB * T /
                    (ORC_PI * cL * pow(ORC_HBAR, 4));
}

/* getZeroOrderScatterConst (emcZeroOrderInterValleyScatterMechanism.hpp:11-21),
 * getFirstOrderScatterConst (emcFirstOrderInterValleyScatterMechanism.hpp:11-20) */
void orc_mech_iv(orc_mech *m, int kind, int valley, int finalValley,
                 double defPot, double phononEnergy, double rho, double T,
                 int Z, int initDeg, const int *finalSub) {
  memset(m, 0, sizeof(*m));
  m->kind = kind;
  m->valley = valley;
  m->finalValley = finalValley;
  m->phononEnergy = phononEnergy;
  m->Z = Z;
  for (int s = 0; s < initDeg; s++)
    for (int z = 0; z < Z; z++)
      m->finalSub[s][z] = finalSub[s * Z + z];
  int isAb = (kind == ORC_MECH_ZERO_IV_AB || kind == ORC_MECH_FIRST_IV_AB);
  double result;
  if (kind == ORC_MECH_ZERO_IV_AB || kind == ORC_MECH_ZERO_IV_EM) {
    result = (double)(size_t)Z * sqrt(ORC_Q) * pow(defPot / ORC_HBAR, 2) * ORC_Q /
             (ORC_PI * rho * phononEnergy * sqrt(2));
  } else {
    result = (double)(size_t)Z * sqrt(2) * pow(ORC_Q, 5. / 2.) * pow(defPot, 2) /
             (ORC_PI * rho * pow(ORC_HBAR, 4) * phononEnergy);
  }
  double nrPh = 1. / (exp(ORC_Q * phononEnergy / (ORC_KB * T)) - 1.);
  m->scatterConst = isAb ? result * nrPh : result * (nrPh + 1);
}

/* froehlichScatterConst3D / boseEinstein (emcFroehlichInteraction.hpp:45-57) */
void orc_mech_froehlich(orc_mech *m, int kind, int valley,
                        double phononEnergy, double relEffMass, double epsHi,
                        double epsLo, double T) {
  memset(m, 0, sizeof(*m));
  m->kind = kind;
  m->valley = valley;
  m->finalValley = valley;
  m->phononEnergy = phononEnergy;
  m->effMass = relEffMass * ORC_ME;
  double omega0 = phononEnergy * ORC_Q / ORC_HBAR;
  m->scatterConst = ORC_Q * ORC_Q * omega0 * m->effMass /
                    (4. * ORC_PI * ORC_HBAR * ORC_HBAR) *
                    (1.0 / epsHi - 1.0 / epsLo) / ORC_EPS0;
  double x = ORC_Q * phononEnergy / (ORC_KB * T);
  m->nBose = 1.0 / (exp(x) - 1.0);
}

/* emcHotPhononFroehlich{Absorption,Emission}3D constructors
 * (emcHotPhononFroehlichMechanism.hpp:61-70, :144-153): the Froehlich
 * prefactor without an equilibrium occupation */
void orc_mech_hot_froehlich(orc_mech *m, int kind, int valley,
                            double phononEnergy, double relEffMass,
                            double epsHi, double epsLo, int bath) {
  orc_mech_froehlich(m, ORC_MECH_FROEHLICH_AB, valley, phononEnergy,
                     relEffMass, epsHi, epsLo, 300.0);
  m->kind = kind;
  m->nBose = 0.0;
  m->bath = bath;
}

/* ======================================================================
 * LO phonon bath: include/emcPhononBath.hpp
 * ====================================================================== */

/* planck / planckTemp (:127-137) */
static double orc_planck(double energyEV, double tempK) {
  double x = ORC_Q * energyEV / (ORC_KB * tempK);
  return 1.0 / (exp(x) - 1.0);
}
static double orc_planck_temp(double energyEV, double N) {
  if (N <= 0.0)
    return 0.0;
  return ORC_Q * energyEV / (ORC_KB * log(1.0 + 1.0 / N));
}

/* constructor (:160-192) */
void orc_bath_init(orc_bath *b, int nBins, double dq, double tauLO,
                   double phononEnergy, double latticeTemp, double Vsim,
                   int enableAcoustic, double acPhononEnergy,
                   double tauAcoustic, double wRidley, double toPhononEnergy,
                   double tauTO) {
  memset(b, 0, sizeof(*b));
  b->nBins = nBins;
  b->dq = dq;
  b->tauLO = tauLO;
  b->Vsim = Vsim;
  b->Nq = (double *)calloc((size_t)nBins, sizeof(double));
  b->nEm = (double *)calloc((size_t)nBins, sizeof(double));
  b->nAbs = (double *)calloc((size_t)nBins, sizeof(double));
  b->cumW = (double *)calloc((size_t)nBins + 1, sizeof(double));
  b->cumWN = (double *)calloc((size_t)nBins + 1, sizeof(double));
  b->loE = phononEnergy;
  b->latT = latticeTemp;
  b->N0 = orc_planck(phononEnergy, latticeTemp);
  for (int i = 0; i < nBins; i++)
    b->Nq[i] = b->N0;
  if (enableAcoustic) {
    b->hasAc = 1;
    b->acE = acPhononEnergy;
    b->tauAc = tauAcoustic;
    b->NacEq = orc_planck(acPhononEnergy, latticeTemp);
    b->Nac = b->NacEq;
  }
  if (enableAcoustic && wRidley > 0.0 && toPhononEnergy > 0.0) {
    b->hasRidley = 1;
    b->wR = (wRidley > 1.0) ? 1.0 : wRidley;
    b->toE = toPhononEnergy;
    b->tauTO = tauTO;
    b->NtoEq = orc_planck(toPhononEnergy, latticeTemp);
    b->Nto = b->NtoEq;
  }
  orc_bath_rebuild_window_sums(b);
}

/* rebuildWindowSums (:109-124): weight q/(q^2 + qs2Screen) at the bin
 * centre, prefix sums in bin order */
void orc_bath_rebuild_window_sums(orc_bath *b) {
  b->cumW[0] = 0.0;
  b->cumWN[0] = 0.0;
  for (int i = 0; i < b->nBins; i++) {
    const double q = ((double)(size_t)i + 0.5) * b->dq;
    const double denom = q * q + b->qs2Screen;
    const double w = (denom > 0.0) ? q / denom : 0.0;
    b->cumW[i + 1] = b->cumW[i] + w;
    b->cumWN[i + 1] = b->cumWN[i] + w * b->Nq[i];
  }
}

/* setScreeningQ2 (:203-208): rebuild only on a change */
void orc_bath_set_screening_q2(orc_bath *b, double qs2) {
  if (qs2 == b->qs2Screen)
    return;
  b->qs2Screen = qs2;
  orc_bath_rebuild_window_sums(b);
}

/* bin range [lo, hi) covering [qMin, qMax] (:381-389 / :426-434); returns
 * 0 when it is empty */
static int orc_bath_window(const orc_bath *b, double qMin, double qMax, long *lo, long *hi) {
  long l = (long)floor(qMin / b->dq);
  long h = (long)ceil(qMax / b->dq);
  if (l < 0)
    l = 0;
  if (h > (long)b->nBins)
    h = (long)b->nBins;
  *lo = l;
  *hi = h;
  return h - l >= 1;
}

/* getNqInWindow (:378-394) */
double orc_bath_nq_in_window(const orc_bath *b, double qMin, double qMax) {
  long lo, hi;
  if (b->nBins < 2 || qMax <= qMin)
    return orc_bath_mean_nq(b);
  if (!orc_bath_window(b, qMin, qMax, &lo, &hi))
    return orc_bath_mean_nq(b);
  const double wSum = b->cumW[hi] - b->cumW[lo];
  if (wSum <= 0.0)
    return orc_bath_mean_nq(b);
  return (b->cumWN[hi] - b->cumWN[lo]) / wSum;
}

static double orc_bath_s(const orc_bath *b, long i, int emission) {
  return emission ? (b->cumWN[i] + b->cumW[i]) : b->cumWN[i];
}

/* sampleQ (:423-458): inverse of the piecewise-linear cumulative sum of
 * f(q) q/(q^2 + qs2) over the window, f = N (absorption) or N + 1 */
double orc_bath_sample_q(const orc_bath *b, double qMin, double qMax, int emission, double r) {
  long lo, hi;
  if (b->nBins < 2 || qMax <= qMin)
    return qMin;
  if (!orc_bath_window(b, qMin, qMax, &lo, &hi))
    return qMin;
  const double sLo = orc_bath_s(b, lo, emission), sHi = orc_bath_s(b, hi, emission);
  const double span = sHi - sLo;
  if (!(span > 0.0))
    return 0.5 * (qMin + qMax);
  const double target = sLo + r * span;
  long a = lo, c = hi;
  while (c - a > 1) {
    const long mid = (a + c) / 2;
    if (orc_bath_s(b, mid, emission) <= target)
      a = mid;
    else
      c = mid;
  }
  const double sA = orc_bath_s(b, a, emission), sB = orc_bath_s(b, a + 1, emission);
  const double frac = (sB > sA) ? (target - sA) / (sB - sA) : 0.5;
  const double q = ((double)a + frac) * b->dq;
  return (qMin < ((q < qMax) ? q : qMax)) ? ((q < qMax) ? q : qMax) : qMin;
}

/* sampleQResolvedCosTheta (emcScreenedFroehlichInteraction.hpp:87-97) */
double orc_qres_cos_theta(const orc_bath *b, double kI, double kF, int emission, double r) {
  const double B = 2.0 * kI * kF;
  if (B <= 0.0)
    return 1.0 - 2.0 * r;
  const double q = orc_bath_sample_q(b, fabs(kI - kF), kI + kF, emission, r);
  const double cosTheta = (kI * kI + kF * kF - q * q) / B;
  const double c1 = (cosTheta < 1.0) ? cosTheta : 1.0;
  return (-1.0 < c1) ? c1 : -1.0;
}

void orc_bath_free(orc_bath *b) {
  free(b->cumW);
  free(b->cumWN);
  free(b->Nq);
  free(b->nEm);
  free(b->nAbs);
  free(b->tauProfile);
  memset(b, 0, sizeof(*b));
}

/* binOf (:222-225): truncation of q/dq to size_t, clamped to the last bin */
int orc_bath_bin(const orc_bath *b, double q) {
  double r = q / b->dq;
  if (!(r < (double)b->nBins))
    return b->nBins - 1;
  return (int)(size_t)r;
}

/* recordEmission / recordAbsorption (:237-253) */
void orc_bath_record(const orc_bath *b, double q, int emission) {
  int i = orc_bath_bin(b, q);
  if (emission)
    b->nEm[i] += 1.0;
  else
    b->nAbs[i] += 1.0;
}

/* update (:264-358), operation for operation */
void orc_bath_update(orc_bath *b, double dt) {
  double target = b->N0;
  if (b->hasAc) {
    double nKlemens = b->N0;
    if (b->Nac > b->NacEq) {
      double tAc = orc_planck_temp(b->acE, b->Nac);
      nKlemens = orc_planck(b->loE, tAc);
    }
    if (b->hasRidley) {
      double nRidley = b->N0;
      if (b->Nto > b->NtoEq) {
        double tTo = orc_planck_temp(b->toE, b->Nto);
        nRidley = orc_planck(b->loE, tTo);
      }
      target = (1.0 - b->wR) * nKlemens + b->wR * nRidley;
    } else {
      target = nKlemens;
    }
  }
  const int hasProfile = b->tauProfile != NULL;
  double sumW = 0.0, sumWdN = 0.0, sumWdNtau = 0.0;
  /* energy balance of the step (not in the reference), bin order: the net
   * phonons emitted by the carriers, the change of the LO population and
   * the LO decay towards the target, as phonon counts in the bath volume */
  double cEm = 0.0, cLo = 0.0, cDec = 0.0;
  for (int i = 0; i < b->nBins; i++) {
    double q = ((double)(size_t)i + 0.5) * b->dq;
    double Dph = q * q * b->dq * b->Vsim / (2.0 * ORC_PI * ORC_PI);
    double g = (Dph > 0.0) ? (b->nEm[i] - b->nAbs[i]) / (Dph * dt) : 0.0;
    const double tau_i = hasProfile ? b->tauProfile[i] : b->tauLO;
    const double nOld = b->Nq[i];
    if (b->hasAc) {
      double w = q * q;
      sumW += w;
      if (hasProfile)
        sumWdNtau += w * (b->Nq[i] - target) / tau_i;
      else
        sumWdN += w * (b->Nq[i] - target);
    }
    b->Nq[i] += g * dt - (dt / tau_i) * (b->Nq[i] - target);
    if (b->Nq[i] < 0.0)
      b->Nq[i] = 0.0;
    cEm += b->nEm[i] - b->nAbs[i];
    cLo += Dph * (b->Nq[i] - nOld);
    cDec += Dph * (dt / tau_i) * (nOld - target);
    b->nEm[i] = 0.0;
    b->nAbs[i] = 0.0;
  }
  b->energy[0] = b->loE * cEm;
  b->energy[1] = b->loE * cLo;
  b->energy[2] = b->loE * cDec;
  if (b->hasAc) {
    double fKlemens = b->hasRidley ? (1.0 - b->wR) : 1.0;
    if (!hasProfile) {
      double dNmean = (sumW > 0.0) ? sumWdN / sumW : 0.0;
      b->Nac += dt * fKlemens * dNmean / b->tauLO - dt * (b->Nac - b->NacEq) / b->tauAc;
      if (b->Nac < b->NacEq)
        b->Nac = b->NacEq;
      if (b->hasRidley) {
        b->Nto += dt * b->wR * dNmean / b->tauLO - dt * (b->Nto - b->NtoEq) / b->tauTO;
        if (b->Nto < b->NtoEq)
          b->Nto = b->NtoEq;
      }
    } else {
      double flux = (sumW > 0.0) ? sumWdNtau / sumW : 0.0;
      b->Nac += dt * fKlemens * flux - dt * (b->Nac - b->NacEq) / b->tauAc;
      if (b->Nac < b->NacEq)
        b->Nac = b->NacEq;
      if (b->hasRidley) {
        b->Nto += dt * b->wR * flux - dt * (b->Nto - b->NtoEq) / b->tauTO;
        if (b->Nto < b->NtoEq)
          b->Nto = b->NtoEq;
      }
    }
  }
  orc_bath_rebuild_window_sums(b);
}

/* getMeanNq (:461-470) */
double orc_bath_mean_nq(const orc_bath *b) {
  double sumW = 0.0, sumWN = 0.0;
  for (int i = 0; i < b->nBins; i++) {
    double q = ((double)(size_t)i + 0.5) * b->dq;
    double w = q * q;
    sumW += w;
    sumWN += w * b->Nq[i];
  }
  return (sumW > 0.0) ? sumWN / sumW : b->N0;
}

/* getAcousticTemp / getTOTemp (:480-494) */
double orc_bath_acoustic_temp(const orc_bath *b) {
  if (!b->hasAc || b->Nac <= b->NacEq)
    return b->latT;
  return orc_planck_temp(b->acE, b->Nac);
}
double orc_bath_to_temp(const orc_bath *b) {
  if (!b->hasRidley || b->Nto <= b->NtoEq)
    return b->latT;
  return orc_planck_temp(b->toE, b->Nto);
}

/* screenedFroehlichLogFactor (emcScreenedFroehlichInteraction.hpp:52-62):
 * 1/2 ln((q+^2 + qs2) / (q-^2 + qs2)), q+- = kI +- kF */
double orc_scr_log_factor(double kI, double kF, double qs2) {
  const double qPlus2 = (kI + kF) * (kI + kF);
  const double qMinus2 = (kI - kF) * (kI - kF);
  if (qs2 <= 0.0) {
    if (qMinus2 <= 0.0)
      return 0.0;
    return 0.5 * log(qPlus2 / qMinus2);
  }
  return 0.5 * log((qPlus2 + qs2) / (qMinus2 + qs2));
}

static double orc_stdmin(double a, double b);
static double orc_stdmax(double a, double b);

/* sampleScreenedFroehlichCosTheta (:65-77): the inverse of the cumulative
 * distribution of P(cos) ~ 1 / (A' - B cos), A' = kI^2 + kF^2 + qs2,
 * B = 2 kI kF; isotropic where the form degenerates */
double orc_scr_cos_theta(double kI, double kF, double qs2, double r) {
  const double B = 2.0 * kI * kF;
  if (B <= 0.0)
    return 1.0 - 2.0 * r;
  const double Ap = kI * kI + kF * kF + qs2;
  const double num = Ap - B;
  const double den = Ap + B;
  if (num <= 0.0 || den <= 0.0)
    return 1.0 - 2.0 * r;
  const double cosTheta = (Ap - den * pow(num / den, r)) / B;
  return orc_stdmax(-1.0, orc_stdmin(1.0, cosTheta));
}

/* getScatterRate of each mechanism kind */
double orc_mech_rate(const orc_model *md, orc_mech *m, double energy) {
  const orc_valley *iv = &md->valley[m->valley];
  const orc_valley *fv = &md->valley[m->finalValley];
  switch (m->kind) {
  case ORC_MECH_ACOUSTIC: { /* emcAcousticScatterMechanism.hpp:49-56 */
    double mdos = orc_valley_mdos(iv, 0);
    double alpha = orc_valley_np(iv);
    double gamma = orc_valley_gamma(iv, energy);
    return m->scatterConst * pow(mdos, 3. / 2.) * sqrt(gamma) *
           (2 * alpha * energy + 1.0);
  }
  case ORC_MECH_ZERO_IV_AB:
  case ORC_MECH_ZERO_IV_EM: { /* emcZeroOrder…:103-117 / :246-260 */
    m->deltaValley = fv->bottom - iv->bottom;
    double ef = (m->kind == ORC_MECH_ZERO_IV_AB)
                    ? energy + m->phononEnergy - m->deltaValley
                    : energy - m->phononEnergy - m->deltaValley;
    if (ef > 0) {
      double mdos = orc_valley_mdos(fv, 0);
      double alpha = orc_valley_np(fv);
      double gamma = orc_valley_gamma(fv, ef);
      return m->scatterConst * pow(mdos, 3. / 2.) * sqrt(gamma) *
             (2 * alpha * ef + 1.0);
    }
    return 0;
  }
  case ORC_MECH_FIRST_IV_AB:
  case ORC_MECH_FIRST_IV_EM: { /* emcFirstOrder…:102-119 / :250-267 */
    m->deltaValley = fv->bottom - iv->bottom;
    double ef = (m->kind == ORC_MECH_FIRST_IV_AB)
                    ? energy + m->phononEnergy - m->deltaValley
                    : energy - m->phononEnergy - m->deltaValley;
    if (ef > 0) {
      double mdos = orc_valley_mdos(fv, 0);
      double alpha = orc_valley_np(fv);
      double gamma = orc_valley_gamma(iv, energy);
      double gammaF = orc_valley_gamma(fv, ef);
      return m->scatterConst * pow(mdos, 5. / 2.) * sqrt(gammaF) *
             (2 * alpha * ef + 1.0) * (gamma + gammaF);
    }
    return 0;
  }
  case ORC_MECH_FROEHLICH_AB: { /* emcFroehlichInteraction.hpp:95-107 */
    double gammaI = orc_valley_gamma(iv, energy);
    double ef = energy + m->phononEnergy;
    double gammaF = orc_valley_gamma(iv, ef);
    if (gammaI <= 0.0 || gammaF <= 0.0)
      return 0.0;
    double kI = sqrt(2.0 * m->effMass * gammaI * ORC_Q) / ORC_HBAR;
    double kF = sqrt(2.0 * m->effMass * gammaF * ORC_Q) / ORC_HBAR;
    double ln = log((kI + kF) / (kF - kI));
    return m->scatterConst * m->nBose / kI * ln;
  }
  case ORC_MECH_FROEHLICH_EM: { /* emcFroehlichInteraction.hpp:167-181 */
    if (energy <= m->phononEnergy)
      return 0.0;
    double gammaI = orc_valley_gamma(iv, energy);
    double ef = energy - m->phononEnergy;
    double gammaF = orc_valley_gamma(iv, ef);
    if (gammaI <= 0.0 || gammaF <= 0.0)
      return 0.0;
    double kI = sqrt(2.0 * m->effMass * gammaI * ORC_Q) / ORC_HBAR;
    double kF = sqrt(2.0 * m->effMass * gammaF * ORC_Q) / ORC_HBAR;
    double ln = log((kI + kF) / (kI - kF));
    return m->scatterConst * (m->nBose + 1.0) / kI * ln;
  }
  case ORC_MECH_HOT_FROEHLICH_AB: { /* emcHotPhononFroehlichMechanism.hpp:79-91 */
    double gammaI = orc_valley_gamma(iv, energy);
    double ef = energy + m->phononEnergy;
    double gammaF = orc_valley_gamma(iv, ef);
    if (gammaI <= 0.0 || gammaF <= 0.0)
      return 0.0;
    double kI = sqrt(2.0 * m->effMass * gammaI * ORC_Q) / ORC_HBAR;
    double kF = sqrt(2.0 * m->effMass * gammaF * ORC_Q) / ORC_HBAR;
    double ln = log((kI + kF) / (kF - kI));
    double Nq = orc_bath_mean_nq(&md->bath[m->bath]);
    return m->scatterConst * Nq / kI * ln;
  }
  case ORC_MECH_SCR_HOT_FROEHLICH_AB:
  case ORC_MECH_SCR_HOT_FROEHLICH_EM: { /* emcScreenedFroehlichInteraction.hpp:257-269 / :333-348 */
    const int ab = m->kind == ORC_MECH_SCR_HOT_FROEHLICH_AB;
    if (!ab && energy <= m->phononEnergy)
      return 0.0;
    double gammaI = orc_valley_gamma(iv, energy);
    double gammaF = orc_valley_gamma(iv, ab ? energy + m->phononEnergy : energy - m->phononEnergy);
    if (gammaI <= 0.0 || gammaF <= 0.0)
      return 0.0;
    double kI = sqrt(2.0 * m->effMass * gammaI * ORC_Q) / ORC_HBAR;
    double kF = sqrt(2.0 * m->effMass * gammaF * ORC_Q) / ORC_HBAR;
    const orc_bath *bt = &md->bath[m->bath];
    const double Nq = m->qResolved ? orc_bath_nq_in_window(bt, fabs(kI - kF), kI + kF)
                                   : orc_bath_mean_nq(bt);
    return (ab ? m->scatterConst * Nq : m->scatterConst * (Nq + 1.0)) / kI *
           orc_scr_log_factor(kI, kF, md->qs2);
  }
  case ORC_MECH_SCR_FROEHLICH_AB:
  case ORC_MECH_SCR_FROEHLICH_EM: { /* emcScreenedFroehlichInteraction.hpp:128-138 / :184-196 */
    const int ab = m->kind == ORC_MECH_SCR_FROEHLICH_AB;
    if (!ab && energy <= m->phononEnergy)
      return 0.0;
    double gammaI = orc_valley_gamma(iv, energy);
    double gammaF = orc_valley_gamma(iv, ab ? energy + m->phononEnergy : energy - m->phononEnergy);
    if (gammaI <= 0.0 || gammaF <= 0.0)
      return 0.0;
    double kI = sqrt(2.0 * m->effMass * gammaI * ORC_Q) / ORC_HBAR;
    double kF = sqrt(2.0 * m->effMass * gammaF * ORC_Q) / ORC_HBAR;
    return (ab ? m->scatterConst * m->nBose : m->scatterConst * (m->nBose + 1.0)) / kI *
           orc_scr_log_factor(kI, kF, md->qs2);
  }
  case ORC_MECH_HOT_FROEHLICH_EM: { /* emcHotPhononFroehlichMechanism.hpp:159-173 */
    if (energy <= m->phononEnergy)
      return 0.0;
    double gammaI = orc_valley_gamma(iv, energy);
    double ef = energy - m->phononEnergy;
    double gammaF = orc_valley_gamma(iv, ef);
    if (gammaI <= 0.0 || gammaF <= 0.0)
      return 0.0;
    double kI = sqrt(2.0 * m->effMass * gammaI * ORC_Q) / ORC_HBAR;
    double kF = sqrt(2.0 * m->effMass * gammaF * ORC_Q) / ORC_HBAR;
    double ln = log((kI + kF) / (kI - kF));
    double Nq = orc_bath_mean_nq(&md->bath[m->bath]);
    return m->scatterConst * (Nq + 1.0) / kI * ln;
  }
  }
  return 0;
}

static int orc_is_froehlich_ab(int kind) {
  return kind == ORC_MECH_FROEHLICH_AB || kind == ORC_MECH_HOT_FROEHLICH_AB;
}

/* hot-phonon mechanisms record |q| = |k_new - k_old| after the final state
 * (emcHotPhononFroehlichMechanism.hpp:113-116, :195-198) */
static void orc_hot_record(const orc_model *md, const orc_mech *m,
                           const double kOld[3], const double kNew[3]) {
  if (m->kind != ORC_MECH_HOT_FROEHLICH_AB && m->kind != ORC_MECH_HOT_FROEHLICH_EM &&
      m->kind != ORC_MECH_SCR_HOT_FROEHLICH_AB && m->kind != ORC_MECH_SCR_HOT_FROEHLICH_EM)
    return;
  double qv[3] = {kNew[0] - kOld[0], kNew[1] - kOld[1], kNew[2] - kOld[2]};
  orc_bath_record(&md->bath[m->bath], orc_norm3(qv),
                  m->kind == ORC_MECH_HOT_FROEHLICH_EM || m->kind == ORC_MECH_SCR_HOT_FROEHLICH_EM);
}

/* std::min / std::max semantics used by the Froehlich clamp */
static double orc_stdmin(double a, double b) { return (b < a) ? b : a; }
static double orc_stdmax(double a, double b) { return (a < b) ? b : a; }

/* scatterParticle of the screened Froehlich kinds
 * (emcScreenedFroehlichInteraction.hpp:140-152 / :198-211): kI and kF from
 * the valley's |k|(E) before and after the energy change, the screened
 * angle from uniform ua, the azimuth from ub, |k| rescaled to kF */
static void orc_hot_record(const orc_model *md, const orc_mech *m,
                           const double kOld[3], const double kNew[3]);

static void orc_scr_final(const orc_model *md, const orc_mech *m, orc_particle *p,
                          double ua, double ub) {
  const orc_valley *v = &md->valley[m->valley];
  const int hot = m->kind == ORC_MECH_SCR_HOT_FROEHLICH_AB || m->kind == ORC_MECH_SCR_HOT_FROEHLICH_EM;
  const int ab = m->kind == ORC_MECH_SCR_FROEHLICH_AB || m->kind == ORC_MECH_SCR_HOT_FROEHLICH_AB;
  double kOld[3] = {p->k[0], p->k[1], p->k[2]};
  double kI = orc_valley_normk(v, p->energy);
  if (ab)
    p->energy += m->phononEnergy;
  else
    p->energy -= m->phononEnergy;
  double kF = orc_valley_normk(v, p->energy);
  /* the hot kinds (:283-303 / :358-375): q-resolved angle when both
   * switches are on, else the screened one; then the bath record */
  double cosTheta = (hot && m->qResolved && m->qResolvedAngle)
                        ? orc_qres_cos_theta(&md->bath[m->bath], kI, kF, !ab, ua)
                        : orc_scr_cos_theta(kI, kF, md->qs2, ua);
  double nk[3];
  orc_init_random_direction_wrt_k(p->k, cosTheta, ub, nk);
  memcpy(p->k, nk, sizeof nk);
  double kCurr = orc_norm3(p->k);
  if (kCurr > 0.0) {
    double s = kF / kCurr;
    for (int i = 0; i < 3; i++)
      p->k[i] = p->k[i] * s;
  }
  if (hot)
    orc_hot_record(md, m, kOld, p->k);
}

/* scatterParticle of each mechanism kind.  Two uniforms fed to
 * initRandomDirection are drawn right-to-left (g++ evaluates the call
 * arguments in that order): cos(theta) first, then phi. */
void orc_mech_scatter(const orc_model *md, const orc_mech *m,
                      orc_particle *p, orc_rng *rng) {
  switch (m->kind) {
  case ORC_MECH_SCR_FROEHLICH_AB:
  case ORC_MECH_SCR_FROEHLICH_EM:
  case ORC_MECH_SCR_HOT_FROEHLICH_AB:
  case ORC_MECH_SCR_HOT_FROEHLICH_EM: {
    /* the angle's uniform is drawn before the azimuth's (separate
     * statements) */
    double ua = orc_rng_uniform(rng, 0., 1.);
    double ub = orc_rng_uniform(rng, 0., 1.);
    orc_scr_final(md, m, p, ua, ub);
    return;
  }
  case ORC_MECH_ACOUSTIC: { /* emcAcousticScatterMechanism.hpp:70-72 */
    double n = orc_norm3(p->k);
    double r2 = orc_rng_uniform(rng, 0., 1.);
    double r1 = orc_rng_uniform(rng, 0., 1.);
    orc_init_random_direction(n, r1, r2, p->k);
    return;
  }
  case ORC_MECH_ZERO_IV_AB:
  case ORC_MECH_FIRST_IV_AB:
  case ORC_MECH_ZERO_IV_EM:
  case ORC_MECH_FIRST_IV_EM: { /* emcZeroOrder…:119-129 / :262-272 */
    p->valley = m->finalValley;
    uint64_t raw = orc_rng_raw(rng);
    p->subValley = m->finalSub[p->subValley][raw % (uint64_t)m->Z];
    if (m->kind == ORC_MECH_ZERO_IV_AB || m->kind == ORC_MECH_FIRST_IV_AB)
      p->energy += (m->phononEnergy - m->deltaValley);
    else
      p->energy -= (m->phononEnergy + m->deltaValley);
    double knew = orc_valley_normk(&md->valley[m->finalValley], p->energy);
    double r2 = orc_rng_uniform(rng, 0., 1.);
    double r1 = orc_rng_uniform(rng, 0., 1.);
    orc_init_random_direction(knew, r1, r2, p->k);
    return;
  }
  case ORC_MECH_FROEHLICH_AB:
  case ORC_MECH_FROEHLICH_EM:
  case ORC_MECH_HOT_FROEHLICH_AB:
  case ORC_MECH_HOT_FROEHLICH_EM: {
    /* emcFroehlichInteraction.hpp:109-128 / :183-200; the hot-phonon
     * kinds (emcHotPhononFroehlichMechanism.hpp:94-117 / :175-199) are the
     * same sampler followed by the bath record */
    double kOld[3] = {p->k[0], p->k[1], p->k[2]};
    double initEnergy = p->energy;
    if (orc_is_froehlich_ab(m->kind))
      p->energy += m->phononEnergy;
    else
      p->energy -= m->phononEnergy;
    double finalEnergy = p->energy;
    double f = 2.0 * sqrt(initEnergy * finalEnergy) /
               (sqrt(initEnergy) - sqrt(finalEnergy)) /
               (sqrt(initEnergy) - sqrt(finalEnergy));
    double cosTheta =
        (1.0 + f - pow(1.0 + 2.0 * f, orc_rng_uniform(rng, 0., 1.))) / f;
    cosTheta = orc_stdmax(-1.0, orc_stdmin(1.0, cosTheta));
    double kNew = orc_valley_normk(&md->valley[m->valley], p->energy);
    double nk[3];
    orc_init_random_direction_wrt_k(p->k, cosTheta,
                                    orc_rng_uniform(rng, 0., 1.), nk);
    memcpy(p->k, nk, sizeof nk);
    double kCurr = orc_norm3(p->k);
    if (kCurr > 0.0) {
      double s = kNew / kCurr;
      for (int i = 0; i < 3; i++)
        p->k[i] = p->k[i] * s;
    }
    orc_hot_record(md, m, kOld, p->k);
    return;
  }
  }
}

/* ======================================================================
 * Particle type + scatter handler (single region 0)
 * ====================================================================== */

void orc_model_init(orc_model *m, int nLevels, double maxEnergy) {
  memset(m, 0, sizeof(*m));
  m->nLevels = nLevels;
  m->maxEnergy = maxEnergy;
  /* emcScatterHandler ctor: dE(maxEnergy / nrEnergyLevels) (:73-75) */
  m->dE = maxEnergy / (double)(size_t)nLevels;
  for (int v = 0; v < ORC_MAX_VALLEYS; v++)
    m->tau0[v] = 2e-15;
  m->bathUpdate = 1;
  m->rebuildEvery = 1;
}

void orc_model_free(orc_model *m) {
  for (int v = 0; v < ORC_MAX_VALLEYS; v++) {
    free(m->cum[v]);
    m->cum[v] = NULL;
  }
  for (int b = 0; b < m->nBaths; b++)
    orc_bath_free(&m->bath[b]);
  m->nBaths = 0;
}

int orc_model_has_hot(const orc_model *m) {
  for (int j = 0; j < m->nMech; j++)
    if (m->mech[j].kind == ORC_MECH_HOT_FROEHLICH_AB ||
        m->mech[j].kind == ORC_MECH_HOT_FROEHLICH_EM ||
        m->mech[j].kind == ORC_MECH_SCR_HOT_FROEHLICH_AB ||
        m->mech[j].kind == ORC_MECH_SCR_HOT_FROEHLICH_EM)
      return 1;
  return 0;
}

void orc_model_bath_step(orc_model *m, double dt, int rebuild) {
  for (int b = 0; b < m->nBaths; b++)
    orc_bath_update(&m->bath[b], dt);
  if (rebuild && orc_model_has_hot(m))
    orc_model_build_tables(m);
}

void orc_model_end_step(orc_model *m, double dt, long long nPart, double sumE0,
                        long long cnt0) {
  if (m->scrPlasma) {
    /* hotCarrierMHP.cpp:693-703: every bath's update, updateScreening,
     * then reinitScatterTables every step */
    m->stepCount++;
    if (m->bathUpdate)
      for (int b = 0; b < m->nBaths; b++)
        orc_bath_update(&m->bath[b], dt);
    orc_model_update_screening(m, nPart, sumE0, cnt0);
    orc_model_build_tables(m);
    return;
  }
  if (m->nBaths == 0 || !m->bathUpdate)
    return;
  m->stepCount++;
  int every = m->rebuildEvery > 1 ? m->rebuildEvery : 1;
  orc_model_bath_step(m, dt, m->stepCount % every == 0);
}

void orc_model_update_screening(orc_model *m, long long nPart, double sumE0, long long cnt0) {
  double qs2 = 0.0;
  if (nPart != 0) {
    const double nS = (double)nPart / m->scrVol;
    const double avgE = (cnt0 != 0) ? sumE0 / (double)(size_t)cnt0 : 0.0;
    const double tS = 2.0 * avgE * ORC_Q / (3.0 * ORC_KB);
    if (tS > 0.0)
      qs2 += nS * ORC_Q * ORC_Q / (m->scrEps * ORC_EPS0 * ORC_KB * tS);
  }
  m->qs2 = qs2;
  for (int b = 0; b < m->nBaths; b++)
    orc_bath_set_screening_q2(&m->bath[b], qs2);
}

void orc_valley0_sums(int np, const double *E, const int32_t *valley, double *sumE0,
                      long long *cnt0) {
  double s = 0.0;
  long long c = 0;
  for (int i = 0; i < np; i++)
    if (valley[i] == 0) {
      s += E[i];
      c++;
    }
  *sumE0 = s;
  *cnt0 = c;
}

void orc_valley0_sums_aos(int np, const orc_particle *p, double *sumE0, long long *cnt0) {
  double s = 0.0;
  long long c = 0;
  for (int i = 0; i < np; i++)
    if (p[i].valley == 0) {
      s += p[i].energy;
      c++;
    }
  *sumE0 = s;
  *cnt0 = c;
}

int orc_model_add_valley(orc_model *m, const orc_valley *v) {
  if (m->nValleys >= ORC_MAX_VALLEYS)
    return -1;
  m->valley[m->nValleys] = *v;
  return m->nValleys++;
}

/* addScatterMechanism (emcScatterHandler.hpp:129-145): tables are kept per
 * (valley, region) in registration order. */
int orc_model_add_mech(orc_model *m, const orc_mech *mech) {
  if (m->nMech >= ORC_MAX_MECH)
    return -1;
  int idx = m->nMech++;
  m->mech[idx] = *mech;
  int v = mech->valley;
  m->tableMech[v][m->nTables[v]++] = idx;
  return idx;
}

/* fillScatterTables (:220-235) + renormalizeTables (:248-272) */
void orc_model_build_tables(orc_model *m) {
  for (int v = 0; v < m->nValleys; v++) {
    int nt = m->nTables[v];
    free(m->cum[v]);
    m->cum[v] = NULL;
    if (nt == 0) {
      m->tau0[v] = 2e-15;
      continue;
    }
    double *t = (double *)malloc(sizeof(double) * (size_t)nt * m->nLevels);
    for (int j = 0; j < nt; j++) {
      orc_mech *mech = &m->mech[m->tableMech[v][j]];
      for (int l = 0; l < m->nLevels; l++)
        t[(size_t)j * m->nLevels + l] =
            orc_mech_rate(m, mech, (double)(size_t)(l + 1) * m->dE);
    }
    for (int j = 1; j < nt; j++)
      for (int l = 0; l < m->nLevels; l++)
        t[(size_t)j * m->nLevels + l] =
            t[(size_t)j * m->nLevels + l] + t[(size_t)(j - 1) * m->nLevels + l];
    const double *last = t + (size_t)(nt - 1) * m->nLevels;
    double mx = last[0];
    for (int l = 1; l < m->nLevels; l++)
      if (mx < last[l])
        mx = last[l];
    for (size_t i = 0; i < (size_t)nt * m->nLevels; i++)
      t[i] /= mx;
    m->tau0[v] = 1. / mx;
    m->cum[v] = t;
  }
}

/* getEnergyLevel (:237-244): floor(E/dE) - 1 converted to size_t (x86-64
 * wraps -1 to SIZE_MAX, which the reference maps to 0; any other negative
 * value wraps to a huge index and clamps to the last level). */
int orc_energy_level(const orc_model *m, double E) {
  double v = floor(E / m->dE) - 1;
  if (v == -1.0)
    return 0;
  if (!(v >= 0.0))
    return m->nLevels - 1;
  if (v > (double)(m->nLevels - 1))
    return m->nLevels - 1;
  return (int)v;
}

/* emcScatterHandler::scatterParticle (:148-170) */
int orc_scatter(const orc_model *m, orc_particle *p, orc_rng *rng) {
  int v = p->valley;
  int nt = m->nTables[v];
  if (nt == 0)
    return -3;
  int level = orc_energy_level(m, p->energy);
  double rnd = orc_rng_uniform(rng, 0., 1.);
  const double *cum = m->cum[v];
  if (rnd > cum[(size_t)(nt - 1) * m->nLevels + level])
    return -1;
  double lower = 0, upper;
  for (int j = 0; j < nt; j++) {
    upper = cum[(size_t)j * m->nLevels + level];
    if (rnd >= lower && rnd < upper) {
      orc_mech_scatter(m, &m->mech[m->tableMech[v][j]], p, rng);
      return j;
    }
    lower = upper;
  }
  return -2;
}

/* driftParticle of the bulk handler (basicBulkParticleHandler.hpp:600-613):
 * drift with the position, then each coordinate wrapped into [0, maxPos]
 * once; pos == NULL drifts k only */
static void orc_drift_wrap(double dt, orc_particle *p, const orc_valley *v, double *pos,
                           const double box[3], const double force[3]) {
  orc_drift(dt, p, v, pos, 3, force);
  if (!pos)
    return;
  for (int d = 0; d < 3; d++) {
    if (pos[d] < 0)
      pos[d] = pos[d] + box[d];
    else if (pos[d] > box[d])
      pos[d] = pos[d] - box[d];
  }
}

/* getNewTau (ParticleType/emcParticleType.hpp:187-189) */
double orc_new_tau(const orc_model *m, int valley, orc_rng *rng) {
  return -log(orc_rng_uniform(rng, 1e-6, 1.)) * m->tau0[valley];
}

/* ======================================================================
 * Bulk loop: basicBulkParticleHandler::moveParticles (:181-225)
 * ====================================================================== */
void orc_move_particle(const orc_model *m, orc_particle *p, double tStep,
                       const double force[3], orc_rng *rng,
                       orc_step_info *info, int32_t *evCount, double *pos) {
  const orc_valley *v = &m->valley[p->valley];
  orc_drift_wrap(orc_stdmin(p->tau, tStep), p, v, pos, m->box, force);
  double tRemaining = tStep - p->tau;
  int32_t nEv = 0, nReal = 0, last = -1;
  while (tRemaining > 0) {
    int j = orc_scatter(m, p, rng);
    nEv++;
    if (j >= 0) {
      nReal++;
      last = j;
      if (evCount)
        evCount[j]++;
    }
    double newTau = orc_new_tau(m, p->valley, rng);
    p->tau += newTau;
    v = &m->valley[p->valley];
    orc_drift_wrap(orc_stdmin(tRemaining, newTau), p, v, pos, m->box, force);
    tRemaining -= newTau;
  }
  p->tau -= tStep;
  /* grain bookkeeping (:213-220); no grain mechanism: grainTau stays 1 s */
  p->grainTau -= tStep;
  if (p->grainTau <= 0)
    p->grainTau = -log(orc_rng_uniform(rng, 1e-6, 1.)) * 1.0;
  if (info) {
    info->nEvents = nEv;
    info->nReal = nReal;
    info->lastMech = last;
  }
}

int orc_run_replay(orc_model *m, int np, double *k, double *E,
                   double *tau, double *grainTau, int32_t *valley,
                   int32_t *sub, const uint64_t *draws, size_t stride,
                   size_t *cursor, int nsteps, double tStep,
                   const double force[3], int32_t *valleyOut,
                   int32_t *subOut, int32_t *nEventsOut, double *kOut,
                   double *EOut, int32_t *evCount, double *pos, double *posOut) {
  int ntMax = m->nTables[0];
  for (int v = 1; v < m->nValleys; v++)
    if (m->nTables[v] > ntMax)
      ntMax = m->nTables[v];
  /* step-major: a phonon bath couples the particles between steps */
  char *over = (char *)calloc((size_t)(np > 0 ? np : 1), 1);
  for (int s = 0; s < nsteps; s++) {
    for (int ip = 0; ip < np; ip++) {
      orc_particle p;
      memset(&p, 0, sizeof p);
      p.k[0] = k[3 * ip];
      p.k[1] = k[3 * ip + 1];
      p.k[2] = k[3 * ip + 2];
      p.energy = E[ip];
      p.tau = tau[ip];
      p.grainTau = grainTau ? grainTau[ip] : 1.0;
      p.valley = valley[ip];
      p.subValley = sub[ip];
      orc_rng rng;
      orc_rng_init_buf(&rng, draws + (size_t)ip * stride, stride);
      rng.pos = cursor[ip];
      orc_step_info info;
      orc_move_particle(m, &p, tStep, force, &rng, &info,
                        evCount ? evCount + (size_t)ip * ntMax : NULL,
                        pos ? pos + 3 * (size_t)ip : NULL);
      size_t o = (size_t)s * np + ip;
      if (pos && posOut)
        for (int d = 0; d < 3; d++)
          posOut[3 * o + d] = pos[3 * (size_t)ip + d];
      if (valleyOut)
        valleyOut[o] = p.valley;
      if (subOut)
        subOut[o] = p.subValley;
      if (nEventsOut)
        nEventsOut[o] = info.nEvents;
      if (kOut) {
        kOut[3 * o] = p.k[0];
        kOut[3 * o + 1] = p.k[1];
        kOut[3 * o + 2] = p.k[2];
      }
      if (EOut)
        EOut[o] = p.energy;
      if (rng.overflow)
        over[ip] = 1;
      cursor[ip] = rng.pos;
      k[3 * ip] = p.k[0];
      k[3 * ip + 1] = p.k[1];
      k[3 * ip + 2] = p.k[2];
      E[ip] = p.energy;
      tau[ip] = p.tau;
      if (grainTau)
        grainTau[ip] = p.grainTau;
      valley[ip] = p.valley;
      sub[ip] = p.subValley;
    }
    {
      double se0;
      long long c0;
      orc_valley0_sums(np, E, valley, &se0, &c0);
      orc_model_end_step(m, tStep, np, se0, c0);
    }
  }
  int nOverflow = 0;
  for (int ip = 0; ip < np; ip++)
    nOverflow += over[ip];
  free(over);
  return nOverflow;
}

/* emcElectron::generateInitialParticle (emcElectron.hpp:75-90) followed by
 * initParticleKSpaceMaxwellian (emcParticleInitialization.hpp:36-51) at an
 * interior node (no boundary flips). */
void orc_init_particle(const orc_model *m, orc_particle *p, double Vt,
                       orc_rng *rng) {
  memset(p, 0, sizeof(*p));
  p->region = 0;
  p->valley = (int)floor((double)(size_t)m->nValleys * orc_rng_uniform(rng, 1e-6, 1.));
  const orc_valley *v = &m->valley[p->valley];
  p->subValley = (int)floor((double)(size_t)v->deg * orc_rng_uniform(rng, 1e-6, 1.));
  p->energy = -1.5 * Vt * log(orc_rng_uniform(rng, 1e-6, 1.));
  double nk = orc_valley_normk(v, p->energy);
  double r2 = orc_rng_uniform(rng, 0., 1.);
  double r1 = orc_rng_uniform(rng, 0., 1.);
  orc_init_random_direction(nk, r1, r2, p->k);
  p->tau = orc_new_tau(m, p->valley, rng);
  p->grainTau = -log(orc_rng_uniform(rng, 1e-6, 1.)) * 1.0;
}

/* emcElectron::generateInitialParticle with initEnergyEV > 0
 * (emcElectron.hpp:84-85) -> initParticleKSpaceFixed
 * (emcParticleInitialization.hpp:57-71): no energy draw, E = E0 */
void orc_init_particle_fixed(const orc_model *m, orc_particle *p, double E0,
                             orc_rng *rng) {
  memset(p, 0, sizeof(*p));
  p->region = 0;
  p->valley = (int)floor((double)(size_t)m->nValleys * orc_rng_uniform(rng, 1e-6, 1.));
  const orc_valley *v = &m->valley[p->valley];
  p->subValley = (int)floor((double)(size_t)v->deg * orc_rng_uniform(rng, 1e-6, 1.));
  p->energy = E0;
  double nk = orc_valley_normk(v, E0);
  double r2 = orc_rng_uniform(rng, 0., 1.);
  double r1 = orc_rng_uniform(rng, 0., 1.);
  orc_init_random_direction(nk, r1, r2, p->k);
  p->tau = orc_new_tau(m, p->valley, rng);
  p->grainTau = -log(orc_rng_uniform(rng, 1e-6, 1.)) * 1.0;
}

/* getValleyOccupationProbability / getAvgEnergy / getAvgDriftVelocity
 * (basicBulkParticleHandler.hpp:289-347) */
void orc_observables(const orc_model *m, int np, const orc_particle *parts,
                     const double dir[3], double *avgE, double *avgV,
                     double *occ) {
  size_t cnt[ORC_MAX_VALLEYS] = {0};
  double sE[ORC_MAX_VALLEYS] = {0}, sV[ORC_MAX_VALLEYS] = {0};
  double cOcc[ORC_MAX_VALLEYS] = {0};
  for (int i = 0; i < np; i++) {
    const orc_particle *p = &parts[i];
    double vel[3];
    orc_valley_velocity(&m->valley[p->valley], p->k, p->energy, p->subValley, vel);
    sV[p->valley] += vel[0] * dir[0] + vel[1] * dir[1] + vel[2] * dir[2];
    sE[p->valley] += p->energy;
    cnt[p->valley]++;
    cOcc[p->valley]++;
  }
  for (int v = 0; v < m->nValleys; v++) {
    avgE[v] = cnt[v] ? sE[v] / (double)cnt[v] : 0.;
    avgV[v] = cnt[v] ? sV[v] / (double)cnt[v] : 0.;
    occ[v] = np ? cOcc[v] / (double)(size_t)np : 0.;
  }
}

/* addParticle (basicBulkParticleHandler.hpp:587-596): initParticlePos
 * (emcParticleInitialization.hpp:13-29, Dim draws), then
 * generateInitialParticle (emcElectron.hpp:75-90) with the Maxwellian
 * k-space init and its boundary flips (emcParticleInitialization.hpp:36-51) */
static void orc_shared_add(const orc_model *m, orc_rng *rng, const size_t c[3],
                           const size_t ext[3], const double spacing[3], double Vt,
                           orc_particle *parts, double *pos, int *np, int maxParticles) {
  orc_particle p;
  /* initParticlePos (emcParticleInitialization.hpp:13-29) */
  double x[3];
  for (int d = 0; d < 3; d++) {
    double u = orc_rng_uniform(rng, 0., 1.);
    if (c[d] == ext[d] - 1)
      x[d] = ((double)c[d] - u * 0.5) * spacing[d];
    else if (c[d] == 0)
      x[d] = u * 0.5 * spacing[d];
    else
      x[d] = ((double)c[d] + u - 0.5) * spacing[d];
  }
  memset(&p, 0, sizeof p);
  p.valley = (int)floor((double)(size_t)m->nValleys * orc_rng_uniform(rng, 1e-6, 1.));
  const orc_valley *v = &m->valley[p.valley];
  p.subValley = (int)floor((double)(size_t)v->deg * orc_rng_uniform(rng, 1e-6, 1.));
  /* initEnergyEV > 0: initParticleKSpaceFixed (emcParticleInitialization.hpp:57-71) */
  if (m->initEnergy > 0.0)
    p.energy = m->initEnergy;
  else
    p.energy = -1.5 * Vt * log(orc_rng_uniform(rng, 1e-6, 1.));
  double nk = orc_valley_normk(v, p.energy);
  double r2 = orc_rng_uniform(rng, 0., 1.);
  double r1 = orc_rng_uniform(rng, 0., 1.);
  orc_init_random_direction(nk, r1, r2, p.k);
  for (int d = 0; d < 3; d++)
    if ((c[d] == 0 && p.k[d] < 0) || (c[d] == ext[d] - 1 && p.k[d] > 0))
      p.k[d] *= -1;
  p.tau = orc_new_tau(m, p.valley, rng);
  p.grainTau = -log(orc_rng_uniform(rng, 1e-6, 1.)) * 1.0;
  if (*np < maxParticles) {
    for (int d = 0; d < 3; d++)
      pos[3 * (size_t)*np + d] = x[d];
    parts[(*np)++] = p;
  }
}

/* ---- full reference bulk driver with one shared stream ---------------- */
int orc_run_shared(orc_model *m, uint64_t seed, const double maxPos[3],
                   const double spacing[3], double doping, double T,
                   const double fieldDir[3], double fieldStrength,
                   double tStep, int nsteps, int maxParticles, double *obs,
                   double *kOut, double *EOut, int32_t *valleyOut,
                   int32_t *subOut, double *bathMeanOut, double *posInitOut,
                   double *posOut) {
  orc_rng rng;
  orc_rng_init_mt(&rng, seed);
  /* positions drift with the particles in the device's periodic box */
  for (int d = 0; d < 3; d++)
    m->box[d] = maxPos[d];
  /* emcDevice: extent = round(maxPos/spacing)+1 (emcUtil.hpp:115-124),
   * cellVolume = prod(spacing) (emcDevice.hpp:333-343),
   * Vt = kB/q*T (emcDevice.hpp:88) */
  size_t ext[3];
  for (int d = 0; d < 3; d++)
    ext[d] = (size_t)(round(maxPos[d] / spacing[d]) + 1);
  double cellVolume = 1.;
  for (int d = 0; d < 3; d++)
    cellVolume = cellVolume * spacing[d];
  double Vt = ORC_KB / ORC_Q * T;
  /* basicBulkParticleHandler ctor: normalise direction, scale (:93-108) */
  double dir[3] = {fieldDir[0], fieldDir[1], fieldDir[2]};
  orc_normalize3(dir);
  double field[3], force[3];
  for (int d = 0; d < 3; d++)
    field[d] = dir[d] * fieldStrength;
  for (int d = 0; d < 3; d++)
    force[d] = field[d] * -ORC_Q; /* scale(appliedField, getCharge()) */

  orc_particle *parts =
      (orc_particle *)malloc(sizeof(orc_particle) * (size_t)maxParticles);
  double *pos = (double *)malloc(sizeof(double) * 3 * (size_t)maxParticles);
  int np = 0, needed = 0;
  /* generateInitialParticles (:143-159), x-fastest coordinate order
   * (emcGrid.hpp:122-137) */
  size_t c[3] = {0, 0, 0};
  while (c[2] != ext[2]) {
    /* getInitialNrParticles, doping branch (emcElectron.hpp:48-61) */
    double dens = doping;
    for (int d = 0; d < 3; d++)
      if (c[d] == 0 || c[d] == ext[d] - 1)
        dens *= 0.5;
    double nr = dens * cellVolume;
    while (nr >= 1) {
      orc_shared_add(m, &rng, c, ext, spacing, Vt, parts, pos, &np, maxParticles);
      needed++;
      nr--;
    }
    if (orc_rng_uniform(&rng, 0., 1.) < nr) {
      orc_shared_add(m, &rng, c, ext, spacing, Vt, parts, pos, &np, maxParticles);
      needed++;
    }
    /* advanceCoord */
    c[0]++;
    for (int d = 0; d < 3; d++) {
      if (c[d] >= ext[d] && d != 2) {
        c[d + 1]++;
        for (int r = 0; r <= d; r++)
          c[r] = 0;
      }
    }
  }
  if (needed > maxParticles) {
    free(parts);
    free(pos);
    return -needed;
  }
  if (posInitOut)
    memcpy(posInitOut, pos, sizeof(double) * 3 * (size_t)np);
  int nv = m->nValleys;
  if (m->scrPlasma) {
    /* the screening seeded from the initial carriers (hotCarrierMHP.cpp:
     * 630-637) */
    double se0;
    long long c0;
    orc_valley0_sums_aos(np, parts, &se0, &c0);
    orc_model_update_screening(m, np, se0, c0);
    orc_model_build_tables(m);
  }
  orc_observables(m, np, parts, dir, obs, obs + nv, obs + 2 * nv);
  for (int s = 1; s <= nsteps; s++) {
    for (int i = 0; i < np; i++)
      orc_move_particle(m, &parts[i], tStep, force, &rng, NULL, NULL, pos + 3 * (size_t)i);
    double *o = obs + (size_t)s * 3 * nv;
    orc_observables(m, np, parts, dir, o, o + nv, o + 2 * nv);
    {
      double se0;
      long long c0;
      orc_valley0_sums_aos(np, parts, &se0, &c0);
      orc_model_end_step(m, tStep, np, se0, c0);
    }
    if (bathMeanOut)
      for (int b = 0; b < m->nBaths; b++)
        bathMeanOut[(size_t)(s - 1) * m->nBaths + b] = orc_bath_mean_nq(&m->bath[b]);
  }
  for (int i = 0; i < np; i++) {
    if (kOut) {
      kOut[3 * i] = parts[i].k[0];
      kOut[3 * i + 1] = parts[i].k[1];
      kOut[3 * i + 2] = parts[i].k[2];
    }
    if (EOut)
      EOut[i] = parts[i].energy;
    if (valleyOut)
      valleyOut[i] = parts[i].valley;
    if (subOut)
      subOut[i] = parts[i].subValley;
  }
  if (posOut)
    memcpy(posOut, pos, sizeof(double) * 3 * (size_t)np);
  free(parts);
  free(pos);
  return np;
}

/* ======================================================================
 * Production mode: Philox draws (CPU twin of the GPU kernel, FP64)
 * ====================================================================== */

/* uniform #j of the (step, pid, instance) stream: word j%4 of counter block
 * j/4, its top 24 bits forced odd, (w >> 8 | 1) * 2^-24 -- the GPU's layout
 * in both precisions (FP32 holds the value exactly) */
double orc_philox_u01_24(uint32_t w) { return (double)((w >> 8) | 1u) * 0x1p-24; }

static double orc_prod_u(uint64_t seed, uint32_t inst, uint32_t pid,
                         uint32_t step, uint32_t j) {
  uint32_t ctr[4] = {j >> 2, step, pid, inst};
  uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
  uint32_t w[4];
  orc_philox4x32_10(ctr, key, w);
  return orc_philox_u01_24(w[j & 3u]);
}

/* the reference draws free-flight times from U[1e-6, 1) */
static double orc_prod_log_u(double u) { return u * (1. - 1e-6) + 1e-6; }

void orc_init_particle_philox(const orc_model *m, orc_particle *p, double Vt,
                              uint64_t seed, uint32_t inst, uint32_t pid) {
  const uint32_t s = 0xFFFFFFFFu;
  memset(p, 0, sizeof(*p));
  double uv = orc_prod_log_u(orc_prod_u(seed, inst, pid, s, 0));
  double us = orc_prod_log_u(orc_prod_u(seed, inst, pid, s, 1));
  double ue = orc_prod_log_u(orc_prod_u(seed, inst, pid, s, 2));
  double uphi = orc_prod_u(seed, inst, pid, s, 3);
  double ucos = orc_prod_u(seed, inst, pid, s, 4);
  double ut = orc_prod_log_u(orc_prod_u(seed, inst, pid, s, 5));
  p->valley = (int)floor((double)m->nValleys * uv);
  const orc_valley *v = &m->valley[p->valley];
  p->subValley = (int)floor((double)v->deg * us);
  p->energy = -1.5 * Vt * log(ue);
  orc_init_random_direction(orc_valley_normk(v, p->energy), uphi, ucos, p->k);
  p->tau = -log(ut) * m->tau0[p->valley];
  p->grainTau = 1.0;
}

/* production initialiser at a prescribed energy: E = E0 + sigma * z with
 * z = sqrt(-2 ln u2) cos(2 pi u6) (Box-Muller on the stream's uniforms 2
 * and 6, u2 in [1e-6, 1): |z| <= 5.26), |E| taken if the draw is negative,
 * E0 if it is zero; sigma = 0 gives the reference's fixed-energy
 * initialisation (emcParticleInitialization.hpp:57-71) on Philox draws.
 * Draws 0, 1, 3, 4, 5 as in orc_init_particle_philox. */
void orc_init_particle_philox_energy(const orc_model *m, orc_particle *p,
                                     double E0, double sigma, uint64_t seed,
                                     uint32_t inst, uint32_t pid) {
  const uint32_t s = 0xFFFFFFFFu;
  memset(p, 0, sizeof(*p));
  double uv = orc_prod_log_u(orc_prod_u(seed, inst, pid, s, 0));
  double us = orc_prod_log_u(orc_prod_u(seed, inst, pid, s, 1));
  double uphi = orc_prod_u(seed, inst, pid, s, 3);
  double ucos = orc_prod_u(seed, inst, pid, s, 4);
  double ut = orc_prod_log_u(orc_prod_u(seed, inst, pid, s, 5));
  p->valley = (int)floor((double)m->nValleys * uv);
  const orc_valley *v = &m->valley[p->valley];
  p->subValley = (int)floor((double)v->deg * us);
  double E = E0;
  if (sigma > 0.0) {
    double ue = orc_prod_log_u(orc_prod_u(seed, inst, pid, s, 2));
    double ua = orc_prod_u(seed, inst, pid, s, 6);
    double z = sqrt(-2.0 * log(ue)) * cos(2.0 * ORC_PI * ua);
    E = E0 + sigma * z;
    if (E < 0.0)
      E = -E;
    if (E == 0.0)
      E = E0;
  }
  p->energy = E;
  orc_init_random_direction(orc_valley_normk(v, p->energy), uphi, ucos, p->k);
  p->tau = -log(ut) * m->tau0[p->valley];
  p->grainTau = 1.0;
}

/* final state with explicit uniforms: ua, ub feed the direction (phi, cos)
 * for the isotropic kinds and (exponent, phi) for Froehlich; the final
 * subvalley index comes from the conditional uniform uc. */
static void orc_mech_scatter_u(const orc_model *md, const orc_mech *m,
                               orc_particle *p, double ua, double ub,
                               double uc) {
  switch (m->kind) {
  case ORC_MECH_ACOUSTIC:
    orc_init_random_direction(orc_norm3(p->k), ua, ub, p->k);
    return;
  case ORC_MECH_ZERO_IV_AB:
  case ORC_MECH_FIRST_IV_AB:
  case ORC_MECH_ZERO_IV_EM:
  case ORC_MECH_FIRST_IV_EM: {
    p->valley = m->finalValley;
    int z = (int)(uc * (double)m->Z);
    if (z >= m->Z)
      z = m->Z - 1;
    if (z < 0)
      z = 0;
    p->subValley = m->finalSub[p->subValley][z];
    if (m->kind == ORC_MECH_ZERO_IV_AB || m->kind == ORC_MECH_FIRST_IV_AB)
      p->energy += (m->phononEnergy - m->deltaValley);
    else
      p->energy -= (m->phononEnergy + m->deltaValley);
    double knew = orc_valley_normk(&md->valley[m->finalValley], p->energy);
    orc_init_random_direction(knew, ua, ub, p->k);
    return;
  }
  case ORC_MECH_SCR_FROEHLICH_AB:
  case ORC_MECH_SCR_FROEHLICH_EM:
  case ORC_MECH_SCR_HOT_FROEHLICH_AB:
  case ORC_MECH_SCR_HOT_FROEHLICH_EM:
    orc_scr_final(md, m, p, ua, ub);
    return;
  case ORC_MECH_FROEHLICH_AB:
  case ORC_MECH_FROEHLICH_EM:
  case ORC_MECH_HOT_FROEHLICH_AB:
  case ORC_MECH_HOT_FROEHLICH_EM: {
    double kOld[3] = {p->k[0], p->k[1], p->k[2]};
    double initEnergy = p->energy;
    if (orc_is_froehlich_ab(m->kind))
      p->energy += m->phononEnergy;
    else
      p->energy -= m->phononEnergy;
    double finalEnergy = p->energy;
    double f = 2.0 * sqrt(initEnergy * finalEnergy) /
               (sqrt(initEnergy) - sqrt(finalEnergy)) /
               (sqrt(initEnergy) - sqrt(finalEnergy));
    double cosTheta = (1.0 + f - pow(1.0 + 2.0 * f, ua)) / f;
    cosTheta = orc_stdmax(-1.0, orc_stdmin(1.0, cosTheta));
    double kNew = orc_valley_normk(&md->valley[m->valley], p->energy);
    double nk[3];
    orc_init_random_direction_wrt_k(p->k, cosTheta, ub, nk);
    memcpy(p->k, nk, sizeof nk);
    double kCurr = orc_norm3(p->k);
    if (kCurr > 0.0) {
      double s = kNew / kCurr;
      for (int i = 0; i < 3; i++)
        p->k[i] = p->k[i] * s;
    }
    orc_hot_record(md, m, kOld, p->k);
    return;
  }
  }
}

/* one scatter event of the production mode on the (step, particle) stream:
 * selection on uniform 4e, final state on 4e+1 and 4e+2, the next free
 * flight on 4e+3 (the loop body of orc_move_particle_philox); returns tau' */
double orc_philox_event(const orc_model *m, orc_particle *p, uint64_t seed,
                        uint32_t inst, uint32_t pid, uint32_t step, uint32_t e) {
  double usel = orc_prod_u(seed, inst, pid, step, 4 * e + 0);
  double ua = orc_prod_u(seed, inst, pid, step, 4 * e + 1);
  double ub = orc_prod_u(seed, inst, pid, step, 4 * e + 2);
  double ut = orc_prod_log_u(orc_prod_u(seed, inst, pid, step, 4 * e + 3));
  int vv = p->valley;
  int nt = m->nTables[vv];
  if (nt > 0) {
    int level = orc_energy_level(m, p->energy);
    const double *cum = m->cum[vv];
    if (!(usel > cum[(size_t)(nt - 1) * m->nLevels + level])) {
      double lower = 0, upper;
      for (int j = 0; j < nt; j++) {
        upper = cum[(size_t)j * m->nLevels + level];
        if (usel >= lower && usel < upper) {
          double uc = (usel - lower) / (upper - lower);
          orc_mech_scatter_u(m, &m->mech[m->tableMech[vv][j]], p, ua, ub, uc);
          break;
        }
        lower = upper;
      }
    }
  }
  return -log(ut) * m->tau0[p->valley];
}

double orc_prod_uniform(uint64_t seed, uint32_t inst, uint32_t pid, uint32_t step,
                        uint32_t j) {
  return orc_prod_u(seed, inst, pid, step, j);
}

void orc_move_particle_philox(const orc_model *m, orc_particle *p,
                              double tStep, const double force[3],
                              uint64_t seed, uint32_t inst, uint32_t pid,
                              uint32_t step, orc_step_info *info,
                              int32_t *evCount, double *pos) {
  const orc_valley *v = &m->valley[p->valley];
  orc_drift_wrap(orc_stdmin(p->tau, tStep), p, v, pos, m->box, force);
  double tRemaining = tStep - p->tau;
  int32_t nEv = 0, nReal = 0, last = -1;
  uint32_t e = 0;
  while (tRemaining > 0) {
    double usel = orc_prod_u(seed, inst, pid, step, 4 * e + 0);
    double ua = orc_prod_u(seed, inst, pid, step, 4 * e + 1);
    double ub = orc_prod_u(seed, inst, pid, step, 4 * e + 2);
    double ut = orc_prod_log_u(orc_prod_u(seed, inst, pid, step, 4 * e + 3));
    e++;
    nEv++;
    int vv = p->valley;
    int nt = m->nTables[vv];
    if (nt > 0) {
      int level = orc_energy_level(m, p->energy);
      const double *cum = m->cum[vv];
      if (!(usel > cum[(size_t)(nt - 1) * m->nLevels + level])) {
        double lower = 0, upper;
        for (int j = 0; j < nt; j++) {
          upper = cum[(size_t)j * m->nLevels + level];
          if (usel >= lower && usel < upper) {
            double uc = (usel - lower) / (upper - lower);
            orc_mech_scatter_u(m, &m->mech[m->tableMech[vv][j]], p, ua, ub, uc);
            nReal++;
            last = j;
            if (evCount)
              evCount[j]++;
            break;
          }
          lower = upper;
        }
      }
    }
    double newTau = -log(ut) * m->tau0[p->valley];
    p->tau += newTau;
    v = &m->valley[p->valley];
    orc_drift_wrap(orc_stdmin(tRemaining, newTau), p, v, pos, m->box, force);
    tRemaining -= newTau;
  }
  p->tau -= tStep;
  if (info) {
    info->nEvents = nEv;
    info->nReal = nReal;
    info->lastMech = last;
  }
}

/* ======================================================================
 * Handle API for language bindings
 * ====================================================================== */

orc_model *orc_model_new(int nLevels, double maxEnergy) {
  orc_model *m = (orc_model *)calloc(1, sizeof(orc_model));
  orc_model_init(m, nLevels, maxEnergy);
  return m;
}

void orc_model_delete(orc_model *m) {
  if (!m)
    return;
  orc_model_free(m);
  free(m);
}

int orc_model_add_valley_params(orc_model *m, int kind, double m0, double m1,
                                double m2, double particleMass, int deg,
                                double alpha, double bottom,
                                const double *axes) {
  orc_valley v;
  double rel[3] = {m0, m1, m2};
  orc_valley_init(&v, kind, rel, particleMass, deg, alpha, bottom);
  if (axes && orc_is_aniso(&v))
    for (int s = 0; s < deg; s++)
      orc_valley_set_axes(&v, s, axes + 9 * s, axes + 9 * s + 3, axes + 9 * s + 6);
  return orc_model_add_valley(m, &v);
}

int orc_model_add_acoustic(orc_model *m, int valley, double sigma, double rho,
                           double vSound, double T) {
  orc_mech mech;
  orc_mech_acoustic(&mech, valley, sigma, rho, vSound, T);
  return orc_model_add_mech(m, &mech);
}

int orc_model_add_iv(orc_model *m, int kind, int valley, int finalValley,
                     double defPot, double phononEnergy, double rho, double T,
                     int Z, const int *finalSub) {
  orc_mech mech;
  orc_mech_iv(&mech, kind, valley, finalValley, defPot, phononEnergy, rho, T, Z,
              m->valley[valley].deg, finalSub);
  return orc_model_add_mech(m, &mech);
}

int orc_model_add_froehlich(orc_model *m, int kind, int valley,
                            double phononEnergy, double relEffMass,
                            double epsHi, double epsLo, double T) {
  orc_mech mech;
  orc_mech_froehlich(&mech, kind, valley, phononEnergy, relEffMass, epsHi, epsLo, T);
  return orc_model_add_mech(m, &mech);
}

int orc_model_add_screened_froehlich(orc_model *m, int kind, int valley,
                                     double phononEnergy, double relEffMass,
                                     double epsHi, double epsLo, double T) {
  orc_mech mech;
  orc_mech_froehlich(&mech, ORC_MECH_FROEHLICH_AB, valley, phononEnergy, relEffMass, epsHi,
                     epsLo, T);
  mech.kind = kind;
  return orc_model_add_mech(m, &mech);
}

void orc_model_set_qs2(orc_model *m, double qs2) {
  m->qs2 = qs2;
  for (int b = 0; b < m->nBaths; b++)
    orc_bath_set_screening_q2(&m->bath[b], qs2);
}

int orc_model_add_screened_hot_froehlich(orc_model *m, int kind, int valley,
                                         double phononEnergy, double relEffMass,
                                         double epsHi, double epsLo, int bath,
                                         int qResolved, int qResolvedAngle) {
  orc_mech mech;
  orc_mech_hot_froehlich(&mech, ORC_MECH_HOT_FROEHLICH_AB, valley, phononEnergy, relEffMass,
                         epsHi, epsLo, bath);
  mech.kind = kind;
  mech.qResolved = qResolved;
  mech.qResolvedAngle = qResolvedAngle;
  return orc_model_add_mech(m, &mech);
}

void orc_model_set_screening_plasma(orc_model *m, double eps, double volume) {
  m->scrPlasma = eps > 0.0;
  m->scrEps = eps;
  m->scrVol = volume;
}

void orc_model_seed_screening(orc_model *m, int np, const double *E, const int32_t *valley) {
  double se0;
  long long c0;
  orc_valley0_sums(np, E, valley, &se0, &c0);
  orc_model_update_screening(m, np, se0, c0);
  orc_model_build_tables(m);
}

void orc_model_get_bath_window(const orc_model *m, int b, double *cumW, double *cumWN) {
  const orc_bath *bt = &m->bath[b];
  memcpy(cumW, bt->cumW, sizeof(double) * (size_t)(bt->nBins + 1));
  memcpy(cumWN, bt->cumWN, sizeof(double) * (size_t)(bt->nBins + 1));
}

int orc_model_n_tables(const orc_model *m, int valley) { return m->nTables[valley]; }
int orc_model_table_mech(const orc_model *m, int valley, int j) {
  return m->tableMech[valley][j];
}
double orc_model_tau0(const orc_model *m, int valley) { return m->tau0[valley]; }
double orc_model_dE(const orc_model *m) { return m->dE; }

void orc_model_get_cum(const orc_model *m, int valley, double *out) {
  size_t n = (size_t)m->nTables[valley] * m->nLevels;
  if (m->cum[valley])
    memcpy(out, m->cum[valley], n * sizeof(double));
}

double orc_model_rate(orc_model *m, int mech, double E) {
  return orc_mech_rate(m, &m->mech[mech], E);
}

static void orc_pack(orc_particle *p, const double *k, double E, int32_t valley,
                     int32_t sub) {
  memset(p, 0, sizeof(*p));
  p->k[0] = k[0];
  p->k[1] = k[1];
  p->k[2] = k[2];
  p->energy = E;
  p->valley = valley;
  p->subValley = sub;
  p->grainTau = 1.0;
}

static void orc_unpack(const orc_particle *p, double *k, double *E,
                       int32_t *valley, int32_t *sub) {
  k[0] = p->k[0];
  k[1] = p->k[1];
  k[2] = p->k[2];
  *E = p->energy;
  *valley = p->valley;
  *sub = p->subValley;
}

int orc_scatter_one(const orc_model *m, double *k, double *E, int32_t *valley,
                    int32_t *sub, const uint64_t *draws, size_t ndraw,
                    size_t *ndrawsUsed) {
  orc_particle p;
  orc_pack(&p, k, *E, *valley, *sub);
  orc_rng r;
  orc_rng_init_buf(&r, draws, ndraw);
  int j = orc_scatter(m, &p, &r);
  orc_unpack(&p, k, E, valley, sub);
  if (ndrawsUsed)
    *ndrawsUsed = r.pos;
  return j;
}

void orc_mech_scatter_one(const orc_model *m, int mech, double *k, double *E,
                          int32_t *valley, int32_t *sub,
                          const uint64_t *draws, size_t ndraw,
                          size_t *ndrawsUsed) {
  orc_particle p;
  orc_pack(&p, k, *E, *valley, *sub);
  orc_rng r;
  orc_rng_init_buf(&r, draws, ndraw);
  orc_mech_scatter(m, &m->mech[mech], &p, &r);
  orc_unpack(&p, k, E, valley, sub);
  if (ndrawsUsed)
    *ndrawsUsed = r.pos;
}

void orc_drift_one(const orc_model *m, int valley, int sub, double dt,
                   double *k, double *E, double *pos, const double *force) {
  orc_particle p;
  orc_pack(&p, k, *E, valley, sub);
  orc_drift(dt, &p, &m->valley[valley], pos, 3, force);
  int32_t vv, ss;
  orc_unpack(&p, k, E, &vv, &ss);
}

void orc_velocity_one(const orc_model *m, int valley, int sub,
                      const double *k, double E, double *vel) {
  orc_valley_velocity(&m->valley[valley], k, E, sub, vel);
}

double orc_normk_one(const orc_model *m, int valley, double E) {
  return orc_valley_normk(&m->valley[valley], E);
}

void orc_init_particles_replay(const orc_model *m, int np, double Vt,
                               const uint64_t *draws, size_t stride,
                               size_t *cursor, double *k, double *E,
                               double *tau, double *grainTau, int32_t *valley,
                               int32_t *sub) {
  for (int i = 0; i < np; i++) {
    orc_rng r;
    orc_rng_init_buf(&r, draws + (size_t)i * stride, stride);
    r.pos = cursor[i];
    orc_particle p;
    if (m->initEnergy > 0.0)
      orc_init_particle_fixed(m, &p, m->initEnergy, &r);
    else
      orc_init_particle(m, &p, Vt, &r);
    cursor[i] = r.pos;
    orc_unpack(&p, k + 3 * i, E + i, valley + i, sub + i);
    tau[i] = p.tau;
    if (grainTau)
      grainTau[i] = p.grainTau;
  }
}

static void orc_obs_soa(const orc_model *m, int np, const double *k,
                        const double *E, const int32_t *valley,
                        const int32_t *sub, const double dir[3], double *o) {
  int nv = m->nValleys;
  size_t cnt[ORC_MAX_VALLEYS] = {0};
  double sE[ORC_MAX_VALLEYS] = {0}, sV[ORC_MAX_VALLEYS] = {0};
  for (int i = 0; i < np; i++) {
    double vel[3];
    orc_valley_velocity(&m->valley[valley[i]], k + 3 * i, E[i], sub[i], vel);
    sV[valley[i]] += vel[0] * dir[0] + vel[1] * dir[1] + vel[2] * dir[2];
    sE[valley[i]] += E[i];
    cnt[valley[i]]++;
  }
  for (int v = 0; v < nv; v++) {
    o[v] = cnt[v] ? sE[v] / (double)cnt[v] : 0.;
    o[nv + v] = cnt[v] ? sV[v] / (double)cnt[v] : 0.;
    o[2 * nv + v] = np ? (double)cnt[v] / (double)np : 0.;
  }
}

void orc_run_philox(orc_model *m, int np, double *k, double *E,
                    double *tau, int32_t *valley, int32_t *sub, int nsteps,
                    double tStep, const double force[3],
                    const double fieldDir[3], uint64_t seed,
                    uint32_t instance, uint32_t step0, double *obs,
                    int32_t *nEventsOut, double *pos) {
  int nv = m->nValleys;
  for (int s = 0; s < nsteps; s++) {
    for (int i = 0; i < np; i++) {
      orc_particle p;
      orc_pack(&p, k + 3 * i, E[i], valley[i], sub[i]);
      p.tau = tau[i];
      orc_step_info info;
      orc_move_particle_philox(m, &p, tStep, force, seed, instance, (uint32_t)i,
                               step0 + (uint32_t)s, &info, NULL,
                               pos ? pos + 3 * (size_t)i : NULL);
      orc_unpack(&p, k + 3 * i, E + i, valley + i, sub + i);
      tau[i] = p.tau;
      if (nEventsOut)
        nEventsOut[(size_t)s * np + i] = info.nEvents;
    }
    if (obs)
      orc_obs_soa(m, np, k, E, valley, sub, fieldDir, obs + (size_t)s * 3 * nv);
    {
      double se0;
      long long c0;
      orc_valley0_sums(np, E, valley, &se0, &c0);
      orc_model_end_step(m, tStep, np, se0, c0);
    }
  }
}

void orc_init_particles_philox(const orc_model *m, int np, double Vt,
                               uint64_t seed, uint32_t instance, double *k,
                               double *E, double *tau, int32_t *valley,
                               int32_t *sub, double *pos) {
  for (int i = 0; i < np; i++) {
    orc_particle p;
    orc_init_particle_philox(m, &p, Vt, seed, instance, (uint32_t)i);
    orc_unpack(&p, k + 3 * i, E + i, valley + i, sub + i);
    tau[i] = p.tau;
    /* positions uniform in the box: uniforms 8, 9, 10 of the stream */
    if (pos)
      for (int d = 0; d < 3; d++)
        pos[3 * i + d] =
            orc_prod_u(seed, instance, (uint32_t)i, 0xFFFFFFFFu, 8 + (uint32_t)d) * m->box[d];
  }
}

void orc_model_set_box(orc_model *m, const double box[3]) {
  for (int d = 0; d < 3; d++)
    m->box[d] = box[d];
}

void orc_init_particles_philox_energy(const orc_model *m, int np, double E0,
                                      double sigma, uint64_t seed,
                                      uint32_t instance, double *k, double *E,
                                      double *tau, int32_t *valley,
                                      int32_t *sub) {
  for (int i = 0; i < np; i++) {
    orc_particle p;
    orc_init_particle_philox_energy(m, &p, E0, sigma, seed, instance, (uint32_t)i);
    orc_unpack(&p, k + 3 * i, E + i, valley + i, sub + i);
    tau[i] = p.tau;
  }
}

/* ---- bath handle API ---------------------------------------------------- */
int orc_model_add_bath(orc_model *m, int nBins, double dq, double tauLO,
                       double phononEnergy, double latticeTemp, double Vsim,
                       int enableAcoustic, double acPhononEnergy,
                       double tauAcoustic, double wRidley,
                       double toPhononEnergy, double tauTO) {
  if (m->nBaths >= ORC_MAX_BATHS || nBins < 1)
    return -1;
  orc_bath_init(&m->bath[m->nBaths], nBins, dq, tauLO, phononEnergy,
                latticeTemp, Vsim, enableAcoustic, acPhononEnergy, tauAcoustic,
                wRidley, toPhononEnergy, tauTO);
  return m->nBaths++;
}

/* setTauLOProfile (emcPhononBath.hpp:214-216); NULL restores the uniform
 * lifetime */
void orc_model_set_bath_profile(orc_model *m, int b, const double *profile) {
  orc_bath *bt = &m->bath[b];
  free(bt->tauProfile);
  bt->tauProfile = NULL;
  if (profile) {
    bt->tauProfile = (double *)malloc(sizeof(double) * (size_t)bt->nBins);
    memcpy(bt->tauProfile, profile, sizeof(double) * (size_t)bt->nBins);
  }
}

int orc_model_add_hot_froehlich(orc_model *m, int kind, int valley,
                                double phononEnergy, double relEffMass,
                                double epsHi, double epsLo, int bath) {
  orc_mech mech;
  orc_mech_hot_froehlich(&mech, kind, valley, phononEnergy, relEffMass, epsHi,
                         epsLo, bath);
  return orc_model_add_mech(m, &mech);
}

/* scal: getMeanNq, getMeanNac, getMeanNTO, getN0, getAcousticTemp,
 * getTOTemp */
void orc_model_get_bath(const orc_model *m, int b, double *Nq, double *nEm,
                        double *nAbs, double *scal) {
  const orc_bath *bt = &m->bath[b];
  size_t n = (size_t)bt->nBins;
  if (Nq)
    memcpy(Nq, bt->Nq, n * sizeof(double));
  if (nEm)
    memcpy(nEm, bt->nEm, n * sizeof(double));
  if (nAbs)
    memcpy(nAbs, bt->nAbs, n * sizeof(double));
  if (scal) {
    scal[0] = orc_bath_mean_nq(bt);
    scal[1] = bt->Nac;
    scal[2] = bt->Nto;
    scal[3] = bt->N0;
    scal[4] = orc_bath_acoustic_temp(bt);
    scal[5] = orc_bath_to_temp(bt);
  }
}

void orc_model_set_bath(orc_model *m, int b, const double *Nq, double Nac,
                        double Nto) {
  orc_bath *bt = &m->bath[b];
  memcpy(bt->Nq, Nq, sizeof(double) * (size_t)bt->nBins);
  bt->Nac = Nac;
  bt->Nto = Nto;
  orc_bath_rebuild_window_sums(bt);
}

void orc_model_bath_record_one(orc_model *m, int b, double q, int emission) {
  orc_bath_record(&m->bath[b], q, emission);
}
void orc_model_get_bath_energy(const orc_model *m, int b, double out[3]) {
  for (int j = 0; j < 3; j++)
    out[j] = m->bath[b].energy[j];
}

void orc_model_bath_update_one(orc_model *m, int b, double dt) {
  orc_bath_update(&m->bath[b], dt);
}
void orc_model_set_bath_mode(orc_model *m, int update, int rebuildEvery,
                             long long stepCount) {
  m->bathUpdate = update;
  m->rebuildEvery = rebuildEvery;
  m->stepCount = stepCount;
}
void orc_model_set_init_energy(orc_model *m, double E0) { m->initEnergy = E0; }

/* ---- mt19937_64 ensemble with per-step observables (threaded) --------- */
#include <pthread.h>

typedef struct {
  const orc_model *m;
  orc_particle *parts;
  int i0, i1;
  double tStep;
  const double *force;
  orc_rng *rngs;
  int nsteps;
} orc_mt_job;

static void *orc_mt_worker(void *arg) {
  orc_mt_job *j = (orc_mt_job *)arg;
  for (int s = 0; s < j->nsteps; s++)
    for (int i = j->i0; i < j->i1; i++)
      orc_move_particle(j->m, &j->parts[i], j->tStep, j->force, &j->rngs[i], NULL,
                        NULL, NULL);
  return NULL;
}

#include <time.h>

static double orc_now(void) {
  struct timespec ts;
  clock_gettime(CLOCK_MONOTONIC, &ts);
  return (double)ts.tv_sec + 1e-9 * (double)ts.tv_nsec;
}

/* fraction of the particles in each (valley, subvalley): occ[v * ORC_MAX_DEG
 * + s] (the subValley column of basicBulkParticleHandler::print, :240) */
static void orc_sub_occupancy(const orc_model *m, int np, const orc_particle *parts,
                              double *occ) {
  size_t cnt[ORC_MAX_VALLEYS * ORC_MAX_DEG] = {0};
  for (int i = 0; i < np; i++)
    cnt[parts[i].valley * ORC_MAX_DEG + parts[i].subValley]++;
  for (int j = 0; j < m->nValleys * ORC_MAX_DEG; j++)
    occ[j] = np ? (double)cnt[j] / (double)(size_t)np : 0.;
}

void orc_run_mt(orc_model *m, int np, uint64_t seed, double Vt,
                int nsteps, double tStep, const double force[3],
                const double fieldDir[3], double *obs, int nThreads,
                double *kOut, double *EOut, int32_t *valleyOut,
                int32_t *subOut, double *stepSeconds, double *subOcc) {
  orc_particle *parts = (orc_particle *)malloc(sizeof(orc_particle) * (size_t)np);
  orc_rng *rngs = (orc_rng *)malloc(sizeof(orc_rng) * (size_t)np);
  for (int i = 0; i < np; i++) {
    orc_rng_init_mt(&rngs[i], orc_particle_seed(seed, (uint64_t)i));
    orc_init_particle(m, &parts[i], Vt, &rngs[i]);
  }
  int nv = m->nValleys;
  double dir[3] = {fieldDir[0], fieldDir[1], fieldDir[2]};
  orc_normalize3(dir);
  if (nThreads < 1)
    nThreads = 1;
  /* observables need every step: run the threads one step at a time when
   * obs is requested, else let each thread run all its steps */
  int chunkSteps = (obs || subOcc || m->nBaths || m->scrPlasma) ? 1 : nsteps;
  /* bath tallies: one private copy of the model per thread whose tally
   * pointers are thread-local, summed in thread order after each step
   * (integer-valued doubles: the sum is exact in any order) */
  int ntUsed = nThreads > 256 ? 256 : nThreads;
  orc_model *locals = NULL;
  if (m->nBaths)
    locals = (orc_model *)calloc((size_t)ntUsed, sizeof(orc_model));
  double t0 = orc_now();
  for (int s0 = 0; s0 < nsteps; s0 += chunkSteps) {
    pthread_t th[256];
    orc_mt_job jobs[256];
    int nt = ntUsed;
    for (int t = 0; t < nt; t++) {
      jobs[t].m = m;
      if (locals) {
        locals[t] = *m;
        for (int b = 0; b < m->nBaths; b++) {
          locals[t].bath[b].nEm = (double *)calloc((size_t)m->bath[b].nBins, sizeof(double));
          locals[t].bath[b].nAbs = (double *)calloc((size_t)m->bath[b].nBins, sizeof(double));
        }
        jobs[t].m = &locals[t];
      }
      jobs[t].parts = parts;
      jobs[t].i0 = (int)((long long)np * t / nt);
      jobs[t].i1 = (int)((long long)np * (t + 1) / nt);
      jobs[t].tStep = tStep;
      jobs[t].force = force;
      jobs[t].rngs = rngs;
      jobs[t].nsteps = chunkSteps;
      pthread_create(&th[t], NULL, orc_mt_worker, &jobs[t]);
    }
    for (int t = 0; t < nt; t++)
      pthread_join(th[t], NULL);
    if (obs) {
      orc_observables(m, np, parts, dir, obs + (size_t)s0 * 3 * nv,
                      obs + (size_t)s0 * 3 * nv + nv,
                      obs + (size_t)s0 * 3 * nv + 2 * nv);
    }
    if (subOcc)
      orc_sub_occupancy(m, np, parts, subOcc + (size_t)s0 * nv * ORC_MAX_DEG);
    if (locals) {
      for (int t = 0; t < nt; t++)
        for (int b = 0; b < m->nBaths; b++) {
          for (int i = 0; i < m->bath[b].nBins; i++) {
            m->bath[b].nEm[i] += locals[t].bath[b].nEm[i];
            m->bath[b].nAbs[i] += locals[t].bath[b].nAbs[i];
          }
          free(locals[t].bath[b].nEm);
          free(locals[t].bath[b].nAbs);
        }
    }
    if (locals || m->scrPlasma) {
      double se0;
      long long c0;
      orc_valley0_sums_aos(np, parts, &se0, &c0);
      orc_model_end_step(m, tStep, np, se0, c0);
    }
  }
  free(locals);
  if (stepSeconds)
    *stepSeconds = orc_now() - t0;
  for (int i = 0; i < np; i++) {
    if (kOut)
      orc_unpack(&parts[i], kOut + 3 * i, EOut + i, valleyOut + i, subOut + i);
  }
  free(parts);
  free(rngs);
}

/* single-call wrappers over a model's bath and screening state */
double orc_model_get_qs2(const orc_model *m) { return m->qs2; }
double orc_bath_nq_window_one(const orc_model *m, int b, double qMin, double qMax) {
  return orc_bath_nq_in_window(&m->bath[b], qMin, qMax);
}
double orc_bath_sample_q_one(const orc_model *m, int b, double qMin, double qMax, int emission,
                             double r) {
  return orc_bath_sample_q(&m->bath[b], qMin, qMax, emission, r);
}
double orc_qres_cos_theta_one(const orc_model *m, int b, double kI, double kF, int emission,
                              double r) {
  return orc_qres_cos_theta(&m->bath[b], kI, kF, emission, r);
}
