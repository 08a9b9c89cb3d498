/*
 * emc_oracle_device.c -- CPU restatement of ViennaEMC's 2-D device path.
 * TEST INFRASTRUCTURE ONLY (see emc_oracle_device.h).
 */
#include "emc_oracle_device.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

static double orc_min_d(double a, double b) { return (b < a) ? b : a; } /* std::min */
static double orc_max_d(double a, double b) { return (a < b) ? b : a; } /* std::max */

size_t orc_device_idx(const orc_device *d, size_t ix, size_t iy) {
  return iy * (size_t)d->ext[0] + ix;
}

/* ---- geometry ----------------------------------------------------------- */

static void orc_device_cell_volume(orc_device *d) {
  /* calcCellVolume (emcDevice.hpp:333-343) */
  d->cellVolume = 1.;
  d->cellVolume = d->cellVolume * d->h[0];
  d->cellVolume = d->cellVolume * d->h[1];
  double n0 = d->h[0] / d->debye, n1 = d->h[1] / d->debye;
  d->normCellVolume = 1.;
  d->normCellVolume = d->normCellVolume * n0;
  d->normCellVolume = d->normCellVolume * n1;
  d->cellVolume *= d->width;
  d->normCellVolume *= d->width / d->debye;
}

void orc_device_init(orc_device *d, double maxX, double maxY, double hx,
                     double hy, double T, double epsR, double ni) {
  memset(d, 0, sizeof(*d));
  d->maxPos[0] = maxX;
  d->maxPos[1] = maxY;
  d->h[0] = hx;
  d->h[1] = hy;
  for (int i = 0; i < 2; i++)
    d->ext[i] = (int)(size_t)(round(d->maxPos[i] / d->h[i]) + 1);
  d->T = T;
  d->epsR = epsR;
  d->ni = ni;
  d->Vt = ORC_KB / ORC_Q * T;
  d->debye = sqrt(ORC_EPS0 * epsR * d->Vt / ORC_Q / ni);
  size_t n = (size_t)d->ext[0] * d->ext[1];
  d->doping = (double *)malloc(sizeof(double) * n);
  d->region = (int *)malloc(sizeof(int) * n);
  for (size_t i = 0; i < n; i++) {
    d->doping[i] = ni;
    d->region[i] = -1;
  }
  for (int b = 0; b < 4; b++) {
    int len = d->ext[b < 2 ? 1 : 0];
    d->contactGrid[b] = (int *)malloc(sizeof(int) * (size_t)len);
    for (int i = 0; i < len; i++)
      d->contactGrid[b][i] = -1;
  }
  d->width = 1e-6;
  orc_device_cell_volume(d);
}

void orc_device_free(orc_device *d) {
  free(d->doping);
  free(d->region);
  for (int b = 0; b < 4; b++)
    free(d->contactGrid[b]);
  memset(d, 0, sizeof(*d));
}

void orc_device_set_width(orc_device *d, double width) {
  d->width = width;
  orc_device_cell_volume(d);
}

void orc_device_add_doping(orc_device *d, const double minPos[2],
                           const double maxPos[2], double doping) {
  size_t mn[2], mx[2];
  orc_device_pos_to_coord(d, minPos, mn);
  orc_device_pos_to_coord(d, maxPos, mx);
  /* checkAndAdaptRegionCoords (emcDopingProfile.hpp:161-207), unsigned */
  for (int i = 0; i < 2; i++) {
    size_t e = (size_t)d->ext[i];
    size_t lo = mn[i] < mx[i] ? mn[i] : mx[i];
    size_t hi = mn[i] < mx[i] ? mx[i] : mn[i];
    if (lo > e)
      return; /* completely out of bounds: the reference reports an error */
    if (hi > e) {
      mn[i] = mn[i] < e ? mn[i] : e;
      mx[i] = mx[i] < e ? mx[i] : e;
    }
    if (mx[i] < mn[i]) {
      size_t t = mn[i];
      mn[i] = mx[i];
      mx[i] = t;
    }
  }
  for (size_t iy = 0; iy < (size_t)d->ext[1]; iy++)
    for (size_t ix = 0; ix < (size_t)d->ext[0]; ix++)
      if (ix >= mn[0] && ix <= mx[0] && iy >= mn[1] && iy <= mx[1]) {
        d->region[orc_device_idx(d, ix, iy)] = d->nRegions;
        d->doping[orc_device_idx(d, ix, iy)] = doping;
      }
  if (d->nRegions < ORC_MAX_REGIONS)
    d->regionDoping[d->nRegions] = doping;
  d->nRegions++;
}

static int orc_on_boundary(const orc_device *d, const size_t c[2], int b) {
  int dim = b / 2;
  if (b % 2 == 1)
    return c[dim] == (size_t)(d->ext[dim] - 1);
  return c[dim] == 0;
}

static void orc_coord_device(const orc_device *d, size_t cs, int b, size_t c[2]) {
  int fixed = b / 2;
  c[fixed] = (b % 2 == 0) ? 0 : (size_t)(d->ext[fixed] - 1);
  c[1 - fixed] = cs;
}

static size_t orc_coord_boundary(const size_t c[2], int b) { return c[1 - b / 2]; }

int orc_device_add_ohmic(orc_device *d, int b, double voltage, double minPos,
                         double maxPos) {
  if (b < 0 || b > 3 || d->nContacts >= ORC_MAX_CONTACTS)
    return -1;
  /* posToCoord on the surface (emcDevice.hpp:301-310) */
  double hs = (b / 2 == 0) ? d->h[1] : d->h[0];
  size_t mn = (size_t)round(minPos / hs), mx = (size_t)round(maxPos / hs);
  size_t len = (size_t)d->ext[b < 2 ? 1 : 0];
  if (mn >= len || mx >= len || mx < mn)
    return -1;
  int idx = d->nContacts;
  for (size_t i = mn; i <= mx; i++)
    d->contactGrid[b][i] = idx;
  /* updateAllOccurences (emcSurface.hpp:381-408): every boundary the
   * contact's nodes lie on */
  for (size_t s = mn; s <= mx; s++) {
    size_t c[2];
    orc_coord_device(d, s, b, c);
    for (int bb = 0; bb < 4; bb++)
      if (orc_on_boundary(d, c, bb))
        d->contactGrid[bb][orc_coord_boundary(c, bb)] = idx;
  }
  d->contactV[idx] = voltage;
  d->contactType[idx] = 0;
  d->nContacts++;
  return idx;
}

int orc_device_add_gate(orc_device *d, int b, double voltage, double minPos, double maxPos,
                        double epsOxide, double thickness, double barrierHeight) {
  if (b < 0 || b > 3 || d->nContacts >= ORC_MAX_CONTACTS)
    return -1;
  double hs = (b / 2 == 0) ? d->h[1] : d->h[0];
  size_t mn = (size_t)round(minPos / hs), mx = (size_t)round(maxPos / hs);
  size_t len = (size_t)d->ext[b < 2 ? 1 : 0];
  if (mn >= len || mx >= len || mx < mn)
    return -1;
  int idx = d->nContacts;
  /* idxContactGrid[boundary].fill only: a gate is not carried to the other
   * boundaries its corner nodes lie on */
  for (size_t i = mn; i <= mx; i++)
    d->contactGrid[b][i] = idx;
  d->contactV[idx] = voltage;
  d->contactType[idx] = 2;
  d->gateEps[idx] = epsOxide;
  d->gateThick[idx] = thickness;
  d->gateBarrier[idx] = barrierHeight;
  d->nContacts++;
  return idx;
}

void orc_device_pos_to_coord(const orc_device *d, const double pos[2],
                             size_t coord[2]) {
  for (int i = 0; i < 2; i++)
    coord[i] = (size_t)round(pos[i] / d->h[i]);
}

int orc_device_is_out_of_bounds(const orc_device *d, const double pos[2]) {
  for (int i = 0; i < 2; i++)
    if (pos[i] < 0 || pos[i] > d->maxPos[i])
      return 1;
  return 0;
}

int orc_device_boundary_of(const orc_device *d, const size_t c[2]) {
  for (int b = 0; b < 4; b++)
    if (orc_on_boundary(d, c, b))
      return b;
  return ORC_BINVALID;
}

int orc_device_contact_idx(const orc_device *d, const size_t c[2]) {
  int b = orc_device_boundary_of(d, c);
  if (b == ORC_BINVALID)
    return -1;
  return d->contactGrid[b][orc_coord_boundary(c, b)];
}

int orc_device_is_ohmic(const orc_device *d, const size_t c[2]) {
  int ci = orc_device_contact_idx(d, c);
  return ci != -1 && d->contactType[ci] == 0;
}

const orc_model *orc_device_model_at(const orc_device *d, const orc_model *m, int region) {
  if (d->nRegionModels > 0 && region >= 0 && region < d->nRegionModels)
    return d->regionModel[region];
  return m;
}

int orc_device_set_region_models(orc_device *d, const orc_model *const *models, int n) {
  if (n < 0 || n > ORC_MAX_REGIONS)
    return -1;
  for (int r = 0; r < n; r++)
    d->regionModel[r] = models[r];
  d->nRegionModels = n;
  return 0;
}

/* ---- particle-mesh schemes ----------------------------------------------- */

static int orc_valid(const int ext[2], size_t ix, size_t iy) {
  return ix < (size_t)ext[0] && iy < (size_t)ext[1];
}
static size_t orc_gidx(const int ext[2], size_t ix, size_t iy) {
  return iy * (size_t)ext[0] + ix;
}
/* nodes beyond the grid (a particle exactly on the maximal boundary) are
 * where the reference aborts on an invalid coordinate; they are skipped */
static void orc_add(double *grid, const int ext[2], size_t ix, size_t iy, double v) {
  if (orc_valid(ext, ix, iy))
    grid[orc_gidx(ext, ix, iy)] += v;
}
static double orc_get(const double *grid, const int ext[2], size_t ix, size_t iy) {
  return orc_valid(ext, ix, iy) ? grid[orc_gidx(ext, ix, iy)] : 0.0;
}

void orc_pm_assign(int scheme, const double *pos, int n, double nrCarriers,
                   const double h[2], const int ext[2], double *grid) {
  for (int i = 0; i < n; i++) {
    const double *p = pos + 2 * i;
    if (scheme == ORC_PM_NGP) { /* emcNGPScheme.hpp:42-52 */
      size_t cx = (size_t)round(p[0] / h[0]), cy = (size_t)round(p[1] / h[1]);
      orc_add(grid, ext, cx, cy, nrCarriers);
    } else if (scheme == ORC_PM_CIC) { /* emcCICScheme.hpp:71-86 */
      size_t mx = (size_t)floor(p[0] / h[0]), my = (size_t)floor(p[1] / h[1]);
      double wX = p[0] / h[0] - (double)mx;
      double wY = p[1] / h[1] - (double)my;
      orc_add(grid, ext, mx, my, wX * wY * nrCarriers);
      orc_add(grid, ext, mx + 1, my, (1 - wX) * wY * nrCarriers);
      orc_add(grid, ext, mx, my + 1, wX * (1 - wY) * nrCarriers);
      orc_add(grid, ext, mx + 1, my + 1, (1 - wX) * (1 - wY) * nrCarriers);
    } else { /* emcNECScheme.hpp:62-75 */
      size_t mx = (size_t)floor(p[0] / h[0]), my = (size_t)floor(p[1] / h[1]);
      double w = 0.25 * nrCarriers;
      orc_add(grid, ext, mx, my, w);
      orc_add(grid, ext, mx + 1, my, w);
      orc_add(grid, ext, mx, my + 1, w);
      orc_add(grid, ext, mx + 1, my + 1, w);
    }
  }
}

void orc_pm_force(int scheme, const double *ex, const double *ey,
                  const double pos[2], const double h[2], const int ext[2],
                  double charge, double force[3]) {
  const double *E[2] = {ex, ey};
  force[2] = 0;
  if (scheme == ORC_PM_NGP) { /* emcNGPScheme.hpp:55-68 */
    size_t cx = (size_t)round(pos[0] / h[0]), cy = (size_t)round(pos[1] / h[1]);
    for (int dd = 0; dd < 2; dd++)
      force[dd] = charge * orc_get(E[dd], ext, cx, cy);
  } else if (scheme == ORC_PM_CIC) {
    /* emcCICScheme.hpp:122-141, including the (1 - wY) * (1 - wY) weight of
     * the last node */
    size_t mx = (size_t)floor(pos[0] / h[0]), my = (size_t)floor(pos[1] / h[1]);
    double wX = pos[0] / h[0] - (double)mx;
    double wY = pos[1] / h[1] - (double)my;
    for (int dd = 0; dd < 2; dd++)
      force[dd] = charge * (orc_get(E[dd], ext, mx, my) * wX * wY +
                            orc_get(E[dd], ext, mx + 1, my) * (1 - wX) * wY +
                            orc_get(E[dd], ext, mx, my + 1) * wX * (1 - wY) +
                            orc_get(E[dd], ext, mx + 1, my + 1) * (1 - wY) * (1 - wY));
  } else { /* emcNECScheme.hpp:99-113 */
    size_t mx = (size_t)floor(pos[0] / h[0]), my = (size_t)floor(pos[1] / h[1]);
    force[0] = charge * (orc_get(ex, ext, mx, my) + orc_get(ex, ext, mx, my + 1)) / 2;
    force[1] = charge * (orc_get(ey, ext, mx, my) + orc_get(ey, ext, mx + 1, my)) / 2;
  }
}

/* ---- 3-D particle-mesh schemes ---------------------------------------- */
static int orc_valid3(const int ext[3], size_t i, size_t j, size_t k) {
  return i < (size_t)ext[0] && j < (size_t)ext[1] && k < (size_t)ext[2];
}
static size_t orc_gidx3(const int ext[3], size_t i, size_t j, size_t k) {
  return i + (size_t)ext[0] * (j + (size_t)ext[1] * k); /* emcGrid.hpp:241-253 */
}
static void orc_add3(double *g, const int ext[3], size_t i, size_t j, size_t k, double v) {
  if (orc_valid3(ext, i, j, k))
    g[orc_gidx3(ext, i, j, k)] += v;
}
static double orc_get3(const double *g, const int ext[3], size_t i, size_t j, size_t k) {
  return orc_valid3(ext, i, j, k) ? g[orc_gidx3(ext, i, j, k)] : 0.0;
}

void orc_pm3_assign(int scheme, const int ext[3], const double h[3], const double *pos, int n,
                    double *grid) {
  for (int q = 0; q < n; q++) {
    const double *p = pos + 3 * q;
    if (scheme == ORC_PM_NGP) { /* emcNGPScheme.hpp:36-48: round(pos / h) */
      orc_add3(grid, ext, (size_t)round(p[0] / h[0]), (size_t)round(p[1] / h[1]),
               (size_t)round(p[2] / h[2]), 1.0);
      continue;
    }
    /* emcCICScheme.hpp:99-119: node (floor) takes wX wY wZ, the weights
     * multiplied left to right, times the carriers (1) */
    size_t i = (size_t)floor(p[0] / h[0]), j = (size_t)floor(p[1] / h[1]),
           k = (size_t)floor(p[2] / h[2]);
    double wX = p[0] / h[0] - (double)i, wY = p[1] / h[1] - (double)j,
           wZ = p[2] / h[2] - (double)k;
    orc_add3(grid, ext, i, j, k, wX * wY * wZ * 1.0);
    orc_add3(grid, ext, i + 1, j, k, (1 - wX) * wY * wZ * 1.0);
    orc_add3(grid, ext, i, j + 1, k, wX * (1 - wY) * wZ * 1.0);
    orc_add3(grid, ext, i + 1, j + 1, k, (1 - wX) * (1 - wY) * wZ * 1.0);
    orc_add3(grid, ext, i, j, k + 1, wX * wY * (1 - wZ) * 1.0);
    orc_add3(grid, ext, i + 1, j, k + 1, (1 - wX) * wY * (1 - wZ) * 1.0);
    orc_add3(grid, ext, i, j + 1, k + 1, wX * (1 - wY) * (1 - wZ) * 1.0);
    orc_add3(grid, ext, i + 1, j + 1, k + 1, (1 - wX) * (1 - wY) * (1 - wZ) * 1.0);
  }
}

void orc_pm3_force(int scheme, const int ext[3], const double h[3], const double *ex,
                   const double *ey, const double *ez, const double *pos, int n, double charge,
                   double *force) {
  const double *E[3] = {ex, ey, ez};
  for (int q = 0; q < n; q++) {
    const double *p = pos + 3 * q;
    double *f = force + 3 * q;
    if (scheme == ORC_PM_NGP) { /* emcNGPScheme.hpp:51-66 */
      size_t i = (size_t)round(p[0] / h[0]), j = (size_t)round(p[1] / h[1]),
             k = (size_t)round(p[2] / h[2]);
      for (int d = 0; d < 3; d++)
        f[d] = charge * orc_get3(E[d], ext, i, j, k);
      continue;
    }
    /* emcCICScheme.hpp:143-173: eight nodes, the (x+1, y+1) nodes of both
     * z layers weighted (1 - wY)(1 - wY) as in the reference, the sum times
     * the charge */
    size_t i = (size_t)floor(p[0] / h[0]), j = (size_t)floor(p[1] / h[1]),
           k = (size_t)floor(p[2] / h[2]);
    double wX = p[0] / h[0] - (double)i, wY = p[1] / h[1] - (double)j,
           wZ = p[2] / h[2] - (double)k;
    for (int d = 0; d < 3; d++) {
      const double *g = E[d];
      f[d] = (orc_get3(g, ext, i, j, k) * wX * wY * wZ +
              orc_get3(g, ext, i + 1, j, k) * (1 - wX) * wY * wZ +
              orc_get3(g, ext, i, j + 1, k) * wX * (1 - wY) * wZ +
              orc_get3(g, ext, i + 1, j + 1, k) * (1 - wY) * (1 - wY) * wZ +
              orc_get3(g, ext, i, j, k + 1) * wX * wY * (1 - wZ) +
              orc_get3(g, ext, i + 1, j, k + 1) * (1 - wX) * wY * (1 - wZ) +
              orc_get3(g, ext, i, j + 1, k + 1) * wX * (1 - wY) * (1 - wZ) +
              orc_get3(g, ext, i + 1, j + 1, k + 1) * (1 - wY) * (1 - wY) * (1 - wZ)) *
             charge;
    }
  }
}

/* calcEFieldAtGridPts (emcEFieldCalculation.hpp:13-32) in 3-D: central
 * differences (prev - next) * vt / (2 h) off the component's faces; then
 * setEFieldBoundaryValues (:57-81): on a face of the component's
 * dimension, 0 at an artificial boundary, the inner neighbour's value at a
 * contact */
void orc_pm3_efield(const int ext[3], const double h[3], double vt, const int *bc,
                    const double *pot, double *ex, double *ey, double *ez) {
  double *E[3] = {ex, ey, ez};
  const size_t nn = (size_t)ext[0] * ext[1] * ext[2];
  const size_t stride[3] = {1, (size_t)ext[0], (size_t)ext[0] * ext[1]};
  for (int d = 0; d < 3; d++) {
    for (size_t k = 0; k < (size_t)ext[2]; k++)
      for (size_t j = 0; j < (size_t)ext[1]; j++)
        for (size_t i = 0; i < (size_t)ext[0]; i++) {
          const size_t c[3] = {i, j, k};
          if (c[d] == 0 || c[d] == (size_t)(ext[d] - 1))
            continue;
          const size_t m = orc_gidx3(ext, i, j, k);
          E[d][m] = (pot[m - stride[d]] - pot[m + stride[d]]) * vt / (2 * h[d]);
        }
    for (size_t k = 0; k < (size_t)ext[2]; k++)
      for (size_t j = 0; j < (size_t)ext[1]; j++)
        for (size_t i = 0; i < (size_t)ext[0]; i++) {
          const size_t c[3] = {i, j, k};
          const size_t m = orc_gidx3(ext, i, j, k);
          const int b = bc[(size_t)d * nn + m];
          if (b == -2)
            continue;
          if (b == -1)
            E[d][m] = 0;
          else
            E[d][m] = c[d] == 0 ? E[d][m + stride[d]] : E[d][m - stride[d]];
        }
  }
}

/* setEFieldBoundaryValues (emcEFieldCalculation.hpp:57-81): the surface
 * loop runs once per dimension, as in the reference */
static void orc_efield_boundary(const orc_device *d, double *E[2]) {
  for (int rep = 0; rep < 2; rep++) {
    for (int b = 0; b < 4; b++) {
      size_t len = (size_t)d->ext[b < 2 ? 1 : 0];
      for (size_t s = 0; s < len; s++) {
        size_t c[2];
        orc_coord_device(d, s, b, c);
        int dim = b / 2;
        double *g = E[dim];
        size_t i = orc_device_idx(d, c[0], c[1]);
        if (d->contactGrid[b][s] == -1) {
          g[i] = 0;
        } else if (b % 2 == 0) {
          size_t n2[2] = {c[0], c[1]};
          n2[dim]++;
          g[i] = g[orc_device_idx(d, n2[0], n2[1])];
        } else {
          size_t n2[2] = {c[0], c[1]};
          n2[dim]--;
          g[i] = g[orc_device_idx(d, n2[0], n2[1])];
        }
      }
    }
  }
}

void orc_pm_efield(int scheme, const orc_device *d, const double *pot,
                   double *ex, double *ey) {
  double *E[2] = {ex, ey};
  for (int dim = 0; dim < 2; dim++) {
    for (size_t iy = 0; iy < (size_t)d->ext[1]; iy++)
      for (size_t ix = 0; ix < (size_t)d->ext[0]; ix++) {
        size_t c[2] = {ix, iy};
        if (c[dim] == 0 || c[dim] == (size_t)(d->ext[dim] - 1))
          continue;
        size_t nx[2] = {ix, iy}, pv[2] = {ix, iy};
        nx[dim]++;
        pv[dim]--;
        double next = pot[orc_device_idx(d, nx[0], nx[1])];
        if (scheme == ORC_PM_NEC) /* calcEFieldAtEdgeMidPts (:35-53) */
          E[dim][orc_device_idx(d, ix, iy)] =
              (pot[orc_device_idx(d, ix, iy)] - next) * d->Vt / d->h[dim];
        else /* calcEFieldAtGridPts (:13-32) */
          E[dim][orc_device_idx(d, ix, iy)] =
              (pot[orc_device_idx(d, pv[0], pv[1])] - next) * d->Vt / (2 * d->h[dim]);
      }
    orc_efield_boundary(d, E);
  }
}

/* ---- SOR Poisson ---------------------------------------------------------- */

void orc_sor_init(orc_sor *s, const orc_device *d, double accuracyV,
                  double omega) {
  memset(s, 0, sizeof(*s));
  s->accuracy = accuracyV / d->Vt;
  s->omega = omega;
  s->hNorm[0] = d->h[0] / d->debye;
  s->hNorm[1] = d->h[1] / d->debye;
  /* initDeviceCharacteristics (emcSORSolver.hpp:330-353) */
  s->hFactor[0] = s->hNorm[1] / s->hNorm[0];
  s->hFactor[1] = s->hNorm[0] / s->hNorm[1];
  s->hFactorSum = 0.;
  s->hFactorSum = s->hFactorSum + s->hFactor[0];
  s->hFactorSum = s->hFactorSum + s->hFactor[1];
  s->hProduct = 1.;
  s->hProduct = s->hProduct * s->hNorm[0];
  s->hProduct = s->hProduct * s->hNorm[1];
  s->maxCoord[0] = (size_t)(d->ext[0] - 1);
  s->maxCoord[1] = (size_t)(d->ext[1] - 1);
  size_t n = (size_t)d->ext[0] * d->ext[1];
  s->dopingNorm = (double *)malloc(sizeof(double) * n);
  for (size_t i = 0; i < n; i++)
    s->dopingNorm[i] = d->doping[i] / d->ni;
  /* gate characteristics (emcSORSolver.hpp:355-366) */
  for (int k = 0; k < d->nContacts; k++) {
    s->gateVoltage[k] = d->contactV[k] / d->Vt;
    if (d->contactType[k] == 2) {
      double gamma = d->gateEps[k] / d->epsR;
      double thicknessOx = d->gateThick[k] / d->debye;
      s->gateFactor[k] = 2 * gamma / thicknessOx;
      s->gateBarrier[k] = d->gateBarrier[k] / d->Vt;
    }
  }
}

void orc_sor_free(orc_sor *s) {
  free(s->dopingNorm);
  s->dopingNorm = NULL;
}

/* checkForGateContact (emcSORSolver.hpp:399-412): a gate on boundary b at
 * this node adds its Robin term (barrier height, plus the applied voltage
 * in the non-equilibrium solve) */
static void orc_sor_gate(const orc_sor *s, const orc_device *d, const size_t c[2], int dim,
                         int isMax, int useVApp, double *nom, double *den) {
  int b = 2 * dim + isMax;
  int ci = d->contactGrid[b][orc_coord_boundary(c, b)];
  if (ci == -1 || d->contactType[ci] != 2)
    return;
  double Vg = s->gateBarrier[ci];
  if (useVApp)
    Vg += s->gateVoltage[ci];
  *nom += s->gateFactor[ci] * Vg * s->hFactor[dim] * s->hNorm[dim];
  *den += s->gateFactor[ci] * s->hFactor[dim] * s->hNorm[dim];
}

static void orc_sor_neighbours(const orc_sor *s, const orc_device *d,
                               const double *pot, size_t ix, size_t iy,
                               double *nom, double *den, int useVApp) {
  size_t c[2] = {ix, iy};
  for (int dim = 0; dim < 2; dim++) {
    size_t pv[2] = {ix, iy}, nx[2] = {ix, iy};
    pv[dim]--;
    nx[dim]++;
    if (c[dim] != 0) {
      *nom += pot[orc_device_idx(d, pv[0], pv[1])] * s->hFactor[dim];
    } else { /* at the minimal boundary: mirrored neighbour */
      *nom += pot[orc_device_idx(d, nx[0], nx[1])] * s->hFactor[dim];
      orc_sor_gate(s, d, c, dim, 0, useVApp, nom, den);
    }
    if (c[dim] != s->maxCoord[dim]) {
      *nom += pot[orc_device_idx(d, nx[0], nx[1])] * s->hFactor[dim];
    } else {
      *nom += pot[orc_device_idx(d, pv[0], pv[1])] * s->hFactor[dim];
      orc_sor_gate(s, d, c, dim, 1, useVApp, nom, den);
    }
  }
}

static void orc_sor_set_contacts(const orc_sor *s, const orc_device *d,
                                 double *pot, int applied) {
  for (int b = 0; b < 4; b++) {
    size_t len = (size_t)d->ext[b < 2 ? 1 : 0];
    for (size_t cs = 0; cs < len; cs++) {
      int ci = d->contactGrid[b][cs];
      if (ci == -1 || d->contactType[ci] != 0)
        continue;
      size_t c[2];
      orc_coord_device(d, cs, b, c);
      size_t i = orc_device_idx(d, c[0], c[1]);
      double bi = asinh(0.5 * s->dopingNorm[i]);
      pot[i] = applied ? d->contactV[ci] / d->Vt + bi : bi;
    }
  }
}

void orc_sor_equilibrium(orc_sor *s, const orc_device *d, double *pot,
                         int resetBC) {
  if (resetBC)
    orc_sor_set_contacts(s, d, pot, 0);
  double error;
  s->sweeps = 0;
  do {
    error = 0;
    for (size_t iy = 0; iy < (size_t)d->ext[1]; iy++)
      for (size_t ix = 0; ix < (size_t)d->ext[0]; ix++) {
        size_t c[2] = {ix, iy};
        if (orc_device_is_ohmic(d, c))
          continue;
        size_t i = orc_device_idx(d, ix, iy);
        double cur = pot[i];
        double n = exp(cur);
        double p = 1. / n;
        double nom = s->hProduct * (p - n + s->dopingNorm[i] + cur * (p + n));
        double den = 2 * s->hFactorSum + s->hProduct * (n + p);
        orc_sor_neighbours(s, d, pot, ix, iy, &nom, &den, 0);
        double delta = s->omega * (nom / den - cur);
        pot[i] = cur + delta;
        if (fabs(delta) > error)
          error = fabs(delta);
      }
    s->sweeps++;
  } while (error > s->accuracy);
}

void orc_sor_nonequilibrium(orc_sor *s, const orc_device *d, double *pot,
                            const double *eConc, int resetBC) {
  if (resetBC)
    orc_sor_set_contacts(s, d, pot, 1);
  double error;
  s->sweeps = 0;
  do {
    error = 0;
    for (size_t iy = 0; iy < (size_t)d->ext[1]; iy++)
      for (size_t ix = 0; ix < (size_t)d->ext[0]; ix++) {
        size_t c[2] = {ix, iy};
        if (orc_device_is_ohmic(d, c))
          continue;
        size_t i = orc_device_idx(d, ix, iy);
        double cur = pot[i];
        double p = exp(-cur);
        double n = eConc[i];
        double nom = s->hProduct * (p - n + s->dopingNorm[i] + cur * (p + n));
        double den = 2 * s->hFactorSum + s->hProduct * (n + p);
        orc_sor_neighbours(s, d, pot, ix, iy, &nom, &den, 1);
        double delta = s->omega * (nom / den - cur);
        pot[i] = cur + delta;
        if (fabs(delta) > error)
          error = fabs(delta);
      }
    s->sweeps++;
  } while (error > s->accuracy);
}

void orc_device_init_potential(const orc_device *d, double *pot) {
  size_t n = (size_t)d->ext[0] * d->ext[1];
  for (size_t i = 0; i < n; i++)
    pot[i] = asinh(0.5 * (d->doping[i] / d->ni));
}

void orc_device_concentration(const orc_device *d, const double *nrPart,
                              double *conc) {
  double factor = 1. / d->cellVolume;
  for (size_t iy = 0; iy < (size_t)d->ext[1]; iy++)
    for (size_t ix = 0; ix < (size_t)d->ext[0]; ix++) {
      size_t i = orc_device_idx(d, ix, iy);
      double c = nrPart[i] / d->ni * factor;
      size_t cc[2] = {ix, iy};
      for (int dim = 0; dim < 2; dim++)
        if (cc[dim] == 0 || cc[dim] == (size_t)(d->ext[dim] - 1))
          c *= 2;
      conc[i] = c;
    }
}

/* ---- device particles ----------------------------------------------------- */

/* driftParticle (emcAbstractParticleHandler.hpp:238-249) +
 * handleParticleAtBoundary (emcParticleDrift.hpp:41-66) */
static int orc_device_drift(const orc_model *m, const orc_device *d, double dt,
                            orc_particle *p, double pos[2], const double force[3]) {
  orc_drift(dt, p, &m->valley[p->valley], pos, 2, force);
  int removed = 0;
  if (orc_device_is_out_of_bounds(d, pos)) {
    double pb[2];
    for (int i = 0; i < 2; i++)
      pb[i] = orc_max_d(0., orc_min_d(pos[i], d->maxPos[i]));
    size_t cb[2];
    orc_device_pos_to_coord(d, pb, cb);
    if (orc_device_is_ohmic(d, cb)) {
      removed = 1;
      pos[0] = pb[0];
      pos[1] = pb[1];
    } else { /* specular reflection (emcScatterHandler.hpp:177-189) */
      for (int i = 0; i < 2; i++) {
        if (pos[i] < 0) {
          pos[i] = -pos[i];
          p->k[i] = -p->k[i];
        } else if (pos[i] > d->maxPos[i]) {
          pos[i] = 2 * d->maxPos[i] - pos[i];
          p->k[i] = -p->k[i];
        }
      }
    }
  }
  if (!removed) {
    size_t c[2];
    orc_device_pos_to_coord(d, pos, c);
    if (c[0] < (size_t)d->ext[0] && c[1] < (size_t)d->ext[1])
      p->region = d->region[orc_device_idx(d, c[0], c[1])];
  }
  return removed;
}

int orc_device_move_particle(const orc_model *m, const orc_device *d, int pm,
                             const double *ex, const double *ey, double charge,
                             orc_particle *p, double pos[2], double tStep,
                             orc_rng *rng, int *contact, int32_t *nEvents) {
  double force[3];
  orc_pm_force(pm, ex, ey, pos, d->h, d->ext, charge, force);
  int removed = 0;
  int32_t nev = 0;
  if (p->tau >= tStep) {
    removed = orc_device_drift(m, d, tStep, p, pos, force);
  } else {
    removed = orc_device_drift(m, d, p->tau, p, pos, force);
    double tRem = tStep - p->tau;
    while (tRem > 0 && !removed) {
      /* the tables of the particle's (valley, region), the region of its
       * last drift (driftParticle, emcAbstractParticleHandler.hpp:244-247) */
      const orc_model *mr = orc_device_model_at(d, m, p->region);
      orc_scatter(mr, p, rng);
      nev++;
      double newTau = orc_new_tau(mr, p->valley, rng);
      p->tau += newTau;
      orc_pm_force(pm, ex, ey, pos, d->h, d->ext, charge, force);
      removed = orc_device_drift(m, d, orc_min_d(tRem, newTau), p, pos, force);
      tRem -= newTau;
    }
  }
  p->tau -= tStep;
  if (removed && contact) {
    size_t c[2];
    orc_device_pos_to_coord(d, pos, c);
    *contact = orc_device_contact_idx(d, c);
  }
  /* grain bookkeeping (emcBasicParticleHandler.hpp:134-138); no grain
   * mechanism: a new grainTau of -ln U[1e-6, 1) * 1 s */
  p->grainTau -= tStep;
  if (p->grainTau <= 0 && !removed)
    p->grainTau = -log(orc_rng_uniform(rng, 1e-6, 1.)) * 1.0;
  if (nEvents)
    *nEvents = nev;
  return removed;
}

/* ---- ensemble helpers ------------------------------------------------------ */

static void orc_ens_push(orc_ensemble *e, const orc_particle *p, const double pos[2]) {
  if (e->n == e->cap) {
    e->cap = e->cap ? 2 * e->cap : 1024;
    e->p = (orc_particle *)realloc(e->p, sizeof(orc_particle) * (size_t)e->cap);
    e->pos = (double *)realloc(e->pos, sizeof(double) * 2 * (size_t)e->cap);
  }
  e->p[e->n] = *p;
  e->pos[2 * e->n] = pos[0];
  e->pos[2 * e->n + 1] = pos[1];
  e->n++;
}

static void orc_ens_erase(orc_ensemble *e, int i) {
  memmove(e->p + i, e->p + i + 1, sizeof(orc_particle) * (size_t)(e->n - i - 1));
  memmove(e->pos + 2 * i, e->pos + 2 * i + 2, sizeof(double) * 2 * (size_t)(e->n - i - 1));
  e->n--;
}

/* initParticlePos (emcParticleInitialization.hpp:13-29) */
static void orc_init_pos(const orc_device *d, const size_t c[2], orc_rng *rng,
                         double pos[2]) {
  for (int i = 0; i < 2; i++) {
    if (c[i] == (size_t)(d->ext[i] - 1))
      pos[i] = ((double)c[i] - orc_rng_uniform(rng, 0., 1.) * 0.5) * d->h[i];
    else if (c[i] == 0)
      pos[i] = orc_rng_uniform(rng, 0., 1.) * 0.5 * d->h[i];
    else
      pos[i] = ((double)c[i] + orc_rng_uniform(rng, 0., 1.) - 0.5) * d->h[i];
  }
}

/* addParticle (emcBasicParticleHandler.hpp:241-254): position, then
 * emcElectron::generateInitialParticle / generateInjectedParticle
 * (emcElectron.hpp:75-104) with the Maxwellian k-space initialisation and
 * its boundary flips (emcParticleInitialization.hpp:36-51) */
static void orc_device_add_particle(const orc_model *m, const orc_device *d,
                                    const size_t c[2], orc_rng *rng, orc_ensemble *e) {
  double pos[2];
  orc_init_pos(d, c, rng, pos);
  orc_particle p;
  memset(&p, 0, sizeof p);
  p.region = d->region[orc_device_idx(d, c[0], c[1])];
  p.valley = (int)floor((double)(size_t)m->nValleys * orc_rng_uniform(rng, 1e-6, 1.));
  const orc_valley *v = &m->valley[p.valley];
  p.subValley = (int)floor((double)(size_t)v->deg * orc_rng_uniform(rng, 1e-6, 1.));
  p.energy = -1.5 * d->Vt * log(orc_rng_uniform(rng, 1e-6, 1.));
  double nk = orc_valley_normk(v, p.energy);
  double r2 = orc_rng_uniform(rng, 0., 1.);
  double r1 = orc_rng_uniform(rng, 0., 1.);
  orc_init_random_direction(nk, r1, r2, p.k);
  for (int i = 0; i < 2; i++)
    if ((c[i] == 0 && p.k[i] < 0) || (c[i] == (size_t)(d->ext[i] - 1) && p.k[i] > 0))
      p.k[i] *= -1;
  p.tau = orc_new_tau(orc_device_model_at(d, m, p.region), p.valley, rng);
  p.grainTau = -log(orc_rng_uniform(rng, 1e-6, 1.)) * 1.0;
  orc_ens_push(e, &p, pos);
}

/* getExpectedNrParticlesAtContact (emcElectron.hpp:63-73) */
static double orc_expected_at_contact(const orc_device *d, size_t ix, size_t iy) {
  double e = d->cellVolume * d->doping[orc_device_idx(d, ix, iy)];
  size_t c[2] = {ix, iy};
  for (int i = 0; i < 2; i++)
    if (c[i] == 0 || c[i] == (size_t)(d->ext[i] - 1))
      e *= 0.5;
  return e;
}

void orc_device_contact_rule(const orc_device *d, const double *pos, int n,
                             double nrCarriers, int *keepOut, int *injOut,
                             double *expOut) {
  size_t nn = (size_t)d->ext[0] * d->ext[1];
  double *nr = (double *)calloc(nn, sizeof(double));
  for (size_t iy = 0; iy < (size_t)d->ext[1]; iy++)
    for (size_t ix = 0; ix < (size_t)d->ext[0]; ix++) {
      size_t c[2] = {ix, iy};
      size_t i = orc_device_idx(d, ix, iy);
      expOut[i] = orc_device_is_ohmic(d, c) ? orc_expected_at_contact(d, ix, iy) : 0.0;
      injOut[i] = 0;
    }
  for (int k = 0; k < n; k++) {
    size_t c[2];
    orc_device_pos_to_coord(d, pos + 2 * k, c);
    keepOut[k] = 1;
    if (c[0] < (size_t)d->ext[0] && c[1] < (size_t)d->ext[1] && orc_device_is_ohmic(d, c)) {
      size_t i = orc_device_idx(d, c[0], c[1]);
      if (nr[i] < expOut[i])
        nr[i] += nrCarriers;
      else
        keepOut[k] = 0;
    }
  }
  for (size_t i = 0; i < nn; i++) {
    if (expOut[i] == 0.0)
      continue;
    double diff = expOut[i] - nr[i];
    while (diff > 0) {
      injOut[i]++;
      diff -= nrCarriers;
    }
  }
  free(nr);
}

/* ---- the device simulation ------------------------------------------------ */

/* getChannelDriftCurrent (emcBasicParticleHandler.hpp:224-238): q * carriers
 * per particle * sum of v_x over the particles with x0 <= x <= x1, in array
 * order, / length */
static double orc_channel_current(const orc_model *m, const orc_particle *p, const double *pos,
                                  int n, const double chan[3], double nrCarriers) {
  double sumVx = 0;
  for (int i = 0; i < n; i++) {
    if (pos[2 * i] < chan[0] || pos[2 * i] > chan[1])
      continue;
    double vel[3];
    orc_valley_velocity(&m->valley[p[i].valley], p[i].k, p[i].energy, p[i].subValley, vel);
    sumVx += vel[0];
  }
  return ORC_Q * nrCarriers * sumVx / chan[2];
}

int orc_device_run(const orc_device_run_cfg *cfg, double *eqPot, int *nInit,
                   int *nPartOut, double *potOut, int *exitOut, int *delOut,
                   int *injOut, double *nrPartLast, double *stateOut,
                   int maxState) {
  const orc_device *d = cfg->dev;
  orc_model *m = cfg->model;
  const int nc = d->nContacts;
  const size_t nn = (size_t)d->ext[0] * d->ext[1];
  const double charge = -ORC_Q;
  orc_rng rng;
  orc_rng_init_mt(&rng, cfg->seed);
  orc_sor sor;
  orc_sor_init(&sor, d, cfg->accuracyV, cfg->omega);
  double *pot = (double *)malloc(sizeof(double) * nn);
  double *ex = (double *)calloc(nn, sizeof(double));
  double *ey = (double *)calloc(nn, sizeof(double));
  double *nrPart = (double *)calloc(nn, sizeof(double));
  double *conc = (double *)calloc(nn, sizeof(double));
  double *expNr = (double *)calloc(nn, sizeof(double));
  orc_ensemble e;
  memset(&e, 0, sizeof e);
  for (size_t iy = 0; iy < (size_t)d->ext[1]; iy++)
    for (size_t ix = 0; ix < (size_t)d->ext[0]; ix++) {
      size_t c[2] = {ix, iy};
      if (orc_device_is_ohmic(d, c))
        expNr[orc_device_idx(d, ix, iy)] = orc_expected_at_contact(d, ix, iy);
    }

  /* calcEquilibriumCharacteristics (emcSimulation.hpp:139-146) */
  orc_device_init_potential(d, pot);
  orc_sor_equilibrium(&sor, d, pot, 1);
  if (eqPot)
    memcpy(eqPot, pot, sizeof(double) * nn);
  orc_pm_efield(cfg->pm, d, pot, ex, ey);
  for (size_t iy = 0; iy < (size_t)d->ext[1]; iy++)
    for (size_t ix = 0; ix < (size_t)d->ext[0]; ix++) {
      /* getInitialNrParticles (emcElectron.hpp:48-61): doping, or
       * exp(potential) * ni with usePotentialForInit */
      double dens = cfg->initFromPotential ? exp(pot[orc_device_idx(d, ix, iy)]) * d->ni
                                           : d->doping[orc_device_idx(d, ix, iy)];
      size_t c[2] = {ix, iy};
      for (int i = 0; i < 2; i++)
        if (c[i] == 0 || c[i] == (size_t)(d->ext[i] - 1))
          dens *= 0.5;
      double nr = dens * d->cellVolume;
      while (nr >= 1) {
        orc_device_add_particle(m, d, c, &rng, &e);
        nr -= cfg->nrCarriers;
      }
      if (orc_rng_uniform(&rng, 0., 1.) < nr)
        orc_device_add_particle(m, d, c, &rng, &e);
    }
  if (nInit)
    *nInit = e.n;
  memset(nrPart, 0, sizeof(double) * nn);
  orc_pm_assign(cfg->pm, e.pos, e.n, cfg->nrCarriers, d->h, d->ext, nrPart);
  orc_device_concentration(d, nrPart, conc);

  int resetBC = 1;
  const int every = cfg->poissonInterval > 1 ? cfg->poissonInterval : 1;
  int *rem = (int *)malloc(sizeof(int) * (size_t)(e.n + 1));
  for (int s = 0; s < cfg->nSteps; s++) {
    /* performEMCStep (emcSimulation.hpp:164-193); the field is kept between
     * solves (solvePoisson, :116) */
    if (s % every == 0) {
      orc_sor_nonequilibrium(&sor, d, pot, conc, resetBC);
      orc_pm_efield(cfg->pm, d, pot, ex, ey);
    }
    resetBC = 0; /* showProgress at step 0 clears it (:118-121) */
    if (potOut)
      memcpy(potOut + (size_t)s * nn, pot, sizeof(double) * nn);
    /* driftScatterParticles (emcBasicParticleHandler.hpp:76-145) */
    int *exits = exitOut ? exitOut + (size_t)s * nc : NULL;
    if (exits)
      memset(exits, 0, sizeof(int) * (size_t)nc);
    int nrem = 0;
    rem = (int *)realloc(rem, sizeof(int) * (size_t)(e.n + 1));
    for (int i = 0; i < e.n; i++) {
      int contact = -1;
      if (orc_device_move_particle(m, d, cfg->pm, ex, ey, charge, &e.p[i], e.pos + 2 * i,
                                   cfg->tStep, &rng, &contact, NULL)) {
        if (exits && contact >= 0)
          exits[contact]++;
        rem[nrem++] = i;
      }
    }
    for (int r = 0; r < nrem; r++)
      orc_ens_erase(&e, rem[r] - r);
    /* handleOhmicContacts (emcBasicParticleHandler.hpp:158-192) */
    int *dels = delOut ? delOut + (size_t)s * nc : NULL;
    int *injs = injOut ? injOut + (size_t)s * nc : NULL;
    if (dels)
      memset(dels, 0, sizeof(int) * (size_t)nc);
    if (injs)
      memset(injs, 0, sizeof(int) * (size_t)nc);
    memset(nrPart, 0, sizeof(double) * nn);
    for (int i = 0; i < e.n; i++) {
      size_t c[2];
      orc_device_pos_to_coord(d, e.pos + 2 * i, c);
      if (orc_device_is_ohmic(d, c)) {
        size_t gi = orc_device_idx(d, c[0], c[1]);
        if (nrPart[gi] < expNr[gi]) {
          nrPart[gi] += cfg->nrCarriers;
        } else {
          if (dels)
            dels[orc_device_contact_idx(d, c)]++;
          orc_ens_erase(&e, i);
          i--;
        }
      }
    }
    /* generateInjectedParticles (emcAbstractParticleHandler.hpp:200-216) */
    for (size_t iy = 0; iy < (size_t)d->ext[1]; iy++)
      for (size_t ix = 0; ix < (size_t)d->ext[0]; ix++) {
        size_t c[2] = {ix, iy};
        if (!orc_device_is_ohmic(d, c))
          continue;
        size_t gi = orc_device_idx(d, ix, iy);
        double diff = expNr[gi] - nrPart[gi];
        while (diff > 0) {
          orc_device_add_particle(m, d, c, &rng, &e);
          if (injs)
            injs[orc_device_contact_idx(d, c)]++;
          diff -= cfg->nrCarriers;
        }
      }
    /* assignParticlesToGrid + concentrations (emcSimulation.hpp:207-213) */
    memset(nrPart, 0, sizeof(double) * nn);
    orc_pm_assign(cfg->pm, e.pos, e.n, cfg->nrCarriers, d->h, d->ext, nrPart);
    orc_device_concentration(d, nrPart, conc);
    if (nPartOut)
      nPartOut[s] = e.n;
    /* accumulateDriftCurrent (emcSimulation.hpp:196-203) */
    if (cfg->chanOut)
      cfg->chanOut[s] = orc_channel_current(m, e.p, e.pos, e.n, cfg->chan, cfg->nrCarriers);
  }
  if (nrPartLast)
    memcpy(nrPartLast, nrPart, sizeof(double) * nn);
  if (stateOut) {
    /* per particle: x, y, kx, ky, kz, E, tau, valley, sub */
    int ns = e.n < maxState ? e.n : maxState;
    for (int i = 0; i < ns; i++) {
      double *o = stateOut + (size_t)i * 9;
      o[0] = e.pos[2 * i];
      o[1] = e.pos[2 * i + 1];
      o[2] = e.p[i].k[0];
      o[3] = e.p[i].k[1];
      o[4] = e.p[i].k[2];
      o[5] = e.p[i].energy;
      o[6] = e.p[i].tau;
      o[7] = e.p[i].valley;
      o[8] = e.p[i].subValley;
    }
  }
  int nFinal = e.n;
  free(rem);
  free(e.p);
  free(e.pos);
  free(pot);
  free(ex);
  free(ey);
  free(nrPart);
  free(conc);
  free(expNr);
  orc_sor_free(&sor);
  return nFinal;
}

/* ======================================================================
 * Production twin of the GPU device run (Philox streams)
 *
 * The GPU device path keeps the reference's algorithm and array order but
 * draws its random numbers from Philox-4x32-10 streams: particle pid in
 * step s uses the counter (event, s, pid, ORC_DOM_MOVE); a particle created
 * in step s (the initial ones: s = 0xFFFFFFFF) draws its eight initial
 * uniforms from (j, s, pid, ORC_DOM_INIT); the remainder of a node's initial
 * carrier count is one Bernoulli draw (0, 0xFFFFFFFE, node, ORC_DOM_COUNT).
 * Creation indices (pid) run in creation order.  Everything else follows
 * orc_device_run.
 * ====================================================================== */
#define ORC_DOM_MOVE 0x40000000u
#define ORC_DOM_INIT 0x40000001u
#define ORC_DOM_COUNT 0x40000002u

int orc_device_move_particle_philox(const orc_model *m, const orc_device *d, int pm,
                                    const double *ex, const double *ey, double charge,
                                    orc_particle *p, double pos[2], double tStep,
                                    uint64_t seed, uint32_t pid, uint32_t step,
                                    int *contact) {
  double force[3];
  orc_pm_force(pm, ex, ey, pos, d->h, d->ext, charge, force);
  int removed = 0;
  uint32_t e = 0;
  if (p->tau >= tStep) {
    removed = orc_device_drift(m, d, tStep, p, pos, force);
  } else {
    removed = orc_device_drift(m, d, p->tau, p, pos, force);
    double tRem = tStep - p->tau;
    while (tRem > 0 && !removed) {
      double newTau = orc_philox_event(orc_device_model_at(d, m, p->region), p, seed,
                                       ORC_DOM_MOVE, pid, step, e++);
      p->tau += newTau;
      orc_pm_force(pm, ex, ey, pos, d->h, d->ext, charge, force);
      removed = orc_device_drift(m, d, orc_min_d(tRem, newTau), p, pos, force);
      tRem -= newTau;
    }
  }
  p->tau -= tStep;
  if (removed && contact) {
    size_t c[2];
    orc_device_pos_to_coord(d, pos, c);
    *contact = orc_device_contact_idx(d, c);
  }
  return removed;
}

/* a particle created at node c: position (initParticlePos), valley,
 * subvalley, thermal energy, isotropic direction with the boundary flips,
 * free-flight time, from the eight initial uniforms of its stream */
static void orc_device_new_philox(const orc_model *m, const orc_device *d, const size_t c[2],
                                  uint64_t seed, uint32_t pid, uint32_t sw, double Vt,
                                  orc_particle *p, double pos[2]) {
  double u[8];
  for (int j = 0; j < 8; j++)
    u[j] = orc_prod_uniform(seed, ORC_DOM_INIT, pid, sw, (uint32_t)j);
  const double lo = 1e-6, sc = 1. - 1e-6;
  for (int i = 0; i < 2; i++) {
    if (c[i] == (size_t)(d->ext[i] - 1))
      pos[i] = ((double)c[i] - u[i] * 0.5) * d->h[i];
    else if (c[i] == 0)
      pos[i] = u[i] * 0.5 * d->h[i];
    else
      pos[i] = ((double)c[i] + u[i] - 0.5) * d->h[i];
  }
  memset(p, 0, sizeof(*p));
  p->region = d->region[orc_device_idx(d, c[0], c[1])];
  p->valley = (int)floor((double)m->nValleys * (u[2] * sc + lo));
  if (p->valley >= m->nValleys)
    p->valley = m->nValleys - 1;
  const orc_valley *v = &m->valley[p->valley];
  p->subValley = (int)floor((double)v->deg * (u[3] * sc + lo));
  if (p->subValley >= v->deg)
    p->subValley = v->deg - 1;
  p->energy = -1.5 * Vt * log(u[4] * sc + lo);
  orc_init_random_direction(orc_valley_normk(v, p->energy), u[5], u[6], p->k);
  for (int i = 0; i < 2; i++)
    if ((c[i] == 0 && p->k[i] < 0) || (c[i] == (size_t)(d->ext[i] - 1) && p->k[i] > 0))
      p->k[i] *= -1;
  p->tau = -log(u[7] * sc + lo) * orc_device_model_at(d, m, p->region)->tau0[p->valley];
  p->grainTau = 1.0;
}

typedef struct {
  orc_ensemble e;
  uint32_t *pid;
  int pcap;
} orc_pens;

static void orc_pens_push(orc_pens *q, const orc_particle *p, const double pos[2], uint32_t pid) {
  orc_ens_push(&q->e, p, pos);
  if (q->e.n > q->pcap) {
    q->pcap = q->e.cap;
    q->pid = (uint32_t *)realloc(q->pid, sizeof(uint32_t) * (size_t)q->pcap);
  }
  q->pid[q->e.n - 1] = pid;
}

static void orc_pens_erase(orc_pens *q, int i) {
  memmove(q->pid + i, q->pid + i + 1, sizeof(uint32_t) * (size_t)(q->e.n - i - 1));
  orc_ens_erase(&q->e, i);
}

int orc_device_run_philox(const orc_device_run_cfg *cfg, double Vt, double *eqPot,
                          int *nInit, int *nPartOut, double *potOut, int *exitOut,
                          int *delOut, int *injOut, double *nrPartLast, double *stateOut,
                          int maxState) {
  const orc_device *d = cfg->dev;
  orc_model *m = cfg->model;
  const int nc = d->nContacts;
  const size_t nn = (size_t)d->ext[0] * d->ext[1];
  const double charge = -ORC_Q;
  orc_sor sor;
  orc_sor_init(&sor, d, cfg->accuracyV, cfg->omega);
  double *pot = (double *)malloc(sizeof(double) * nn);
  double *ex = (double *)calloc(nn, sizeof(double));
  double *ey = (double *)calloc(nn, sizeof(double));
  double *nrPart = (double *)calloc(nn, sizeof(double));
  double *conc = (double *)calloc(nn, sizeof(double));
  double *expNr = (double *)calloc(nn, sizeof(double));
  orc_pens q;
  memset(&q, 0, sizeof q);
  for (size_t iy = 0; iy < (size_t)d->ext[1]; iy++)
    for (size_t ix = 0; ix < (size_t)d->ext[0]; ix++) {
      size_t c[2] = {ix, iy};
      if (orc_device_is_ohmic(d, c))
        expNr[orc_device_idx(d, ix, iy)] = orc_expected_at_contact(d, ix, iy);
    }
  orc_device_init_potential(d, pot);
  orc_sor_equilibrium(&sor, d, pot, 1);
  if (eqPot)
    memcpy(eqPot, pot, sizeof(double) * nn);
  orc_pm_efield(cfg->pm, d, pot, ex, ey);
  uint32_t nextPid = 0;
  for (size_t iy = 0; iy < (size_t)d->ext[1]; iy++)
    for (size_t ix = 0; ix < (size_t)d->ext[0]; ix++) {
      size_t c[2] = {ix, iy};
      double dens = cfg->initFromPotential ? exp(pot[orc_device_idx(d, ix, iy)]) * d->ni
                                           : d->doping[orc_device_idx(d, ix, iy)];
      for (int i = 0; i < 2; i++)
        if (c[i] == 0 || c[i] == (size_t)(d->ext[i] - 1))
          dens *= 0.5;
      double nr = dens * d->cellVolume;
      int cnt = 0;
      while (nr >= 1) {
        cnt++;
        nr -= cfg->nrCarriers;
      }
      if (orc_prod_uniform(cfg->seed, ORC_DOM_COUNT, (uint32_t)orc_device_idx(d, ix, iy),
                           0xFFFFFFFEu, 0) < nr)
        cnt++;
      for (int k = 0; k < cnt; k++) {
        orc_particle p;
        double pos[2];
        orc_device_new_philox(m, d, c, cfg->seed, nextPid, 0xFFFFFFFFu, Vt, &p, pos);
        orc_pens_push(&q, &p, pos, nextPid++);
      }
    }
  if (nInit)
    *nInit = q.e.n;
  orc_pm_assign(cfg->pm, q.e.pos, q.e.n, cfg->nrCarriers, d->h, d->ext, nrPart);
  orc_device_concentration(d, nrPart, conc);
  const int every = cfg->poissonInterval > 1 ? cfg->poissonInterval : 1;
  for (int s = 0; s < cfg->nSteps; s++) {
    if (s % every == 0) {
      orc_sor_nonequilibrium(&sor, d, pot, conc, s == 0);
      orc_pm_efield(cfg->pm, d, pot, ex, ey);
    }
    if (potOut)
      memcpy(potOut + (size_t)s * nn, pot, sizeof(double) * nn);
    int *exits = exitOut ? exitOut + (size_t)s * nc : NULL;
    if (exits)
      memset(exits, 0, sizeof(int) * (size_t)nc);
    for (int i = 0; i < q.e.n; i++) {
      int contact = -1;
      if (orc_device_move_particle_philox(m, d, cfg->pm, ex, ey, charge, &q.e.p[i],
                                          q.e.pos + 2 * i, cfg->tStep, cfg->seed, q.pid[i],
                                          (uint32_t)s, &contact)) {
        if (exits && contact >= 0)
          exits[contact]++;
        q.e.p[i].region = -9; /* marked: erased below in array order */
      }
    }
    for (int i = 0; i < q.e.n; i++)
      if (q.e.p[i].region == -9) {
        orc_pens_erase(&q, i);
        i--;
      }
    int *dels = delOut ? delOut + (size_t)s * nc : NULL;
    int *injs = injOut ? injOut + (size_t)s * nc : NULL;
    if (dels)
      memset(dels, 0, sizeof(int) * (size_t)nc);
    if (injs)
      memset(injs, 0, sizeof(int) * (size_t)nc);
    memset(nrPart, 0, sizeof(double) * nn);
    for (int i = 0; i < q.e.n; i++) {
      size_t c[2];
      orc_device_pos_to_coord(d, q.e.pos + 2 * i, c);
      if (orc_device_is_ohmic(d, c)) {
        size_t gi = orc_device_idx(d, c[0], c[1]);
        if (nrPart[gi] < expNr[gi]) {
          nrPart[gi] += cfg->nrCarriers;
        } else {
          if (dels)
            dels[orc_device_contact_idx(d, c)]++;
          orc_pens_erase(&q, i);
          i--;
        }
      }
    }
    for (size_t iy = 0; iy < (size_t)d->ext[1]; iy++)
      for (size_t ix = 0; ix < (size_t)d->ext[0]; ix++) {
        size_t c[2] = {ix, iy};
        if (!orc_device_is_ohmic(d, c))
          continue;
        size_t gi = orc_device_idx(d, ix, iy);
        double diff = expNr[gi] - nrPart[gi];
        while (diff > 0) {
          orc_particle p;
          double pos[2];
          orc_device_new_philox(m, d, c, cfg->seed, nextPid, (uint32_t)s, Vt, &p, pos);
          orc_pens_push(&q, &p, pos, nextPid++);
          if (injs)
            injs[orc_device_contact_idx(d, c)]++;
          diff -= cfg->nrCarriers;
        }
      }
    memset(nrPart, 0, sizeof(double) * nn);
    orc_pm_assign(cfg->pm, q.e.pos, q.e.n, cfg->nrCarriers, d->h, d->ext, nrPart);
    orc_device_concentration(d, nrPart, conc);
    if (nPartOut)
      nPartOut[s] = q.e.n;
    if (cfg->chanOut)
      cfg->chanOut[s] = orc_channel_current(m, q.e.p, q.e.pos, q.e.n, cfg->chan, cfg->nrCarriers);
  }
  if (nrPartLast)
    memcpy(nrPartLast, nrPart, sizeof(double) * nn);
  if (stateOut) {
    /* per particle: x, y, kx, ky, kz, E, tau, valley, sub, pid */
    int ns = q.e.n < maxState ? q.e.n : maxState;
    for (int i = 0; i < ns; i++) {
      double *o = stateOut + (size_t)i * 10;
      o[0] = q.e.pos[2 * i];
      o[1] = q.e.pos[2 * i + 1];
      o[2] = q.e.p[i].k[0];
      o[3] = q.e.p[i].k[1];
      o[4] = q.e.p[i].k[2];
      o[5] = q.e.p[i].energy;
      o[6] = q.e.p[i].tau;
      o[7] = q.e.p[i].valley;
      o[8] = q.e.p[i].subValley;
      o[9] = q.pid[i];
    }
  }
  int nFinal = q.e.n;
  free(q.e.p);
  free(q.e.pos);
  free(q.pid);
  free(pot);
  free(ex);
  free(ey);
  free(nrPart);
  free(conc);
  free(expNr);
  orc_sor_free(&sor);
  return nFinal;
}

int orc_device_run_philox_h(orc_device *d, orc_model *m, int pm, double accuracyV,
                            double omega, double tStep, int nSteps, uint64_t seed, double Vt,
                            double *eqPot, int *nInit, int *nPartOut, double *potOut,
                            int *exitOut, int *delOut, int *injOut, double *nrPartLast,
                            double *stateOut, int maxState, int poissonInterval,
                            double chanX0, double chanX1, double chanLen, double *chanOut,
                            int initFromPotential) {
  orc_device_run_cfg cfg;
  memset(&cfg, 0, sizeof cfg);
  cfg.dev = d;
  cfg.model = m;
  cfg.pm = pm;
  cfg.accuracyV = accuracyV;
  cfg.omega = omega;
  cfg.tStep = tStep;
  cfg.nSteps = nSteps;
  cfg.seed = seed;
  cfg.nrCarriers = 1.0;
  cfg.poissonInterval = poissonInterval;
  cfg.chan[0] = chanX0;
  cfg.chan[1] = chanX1;
  cfg.chan[2] = chanLen;
  cfg.chanOut = chanOut;
  cfg.initFromPotential = initFromPotential;
  return orc_device_run_philox(&cfg, Vt, eqPot, nInit, nPartOut, potOut, exitOut, delOut,
                               injOut, nrPartLast, stateOut, maxState);
}

/* ======================================================================
 * Handle API for language bindings (tests)
 * ====================================================================== */

orc_device *orc_device_new(double maxX, double maxY, double hx, double hy,
                           double T, double epsR, double ni) {
  orc_device *d = (orc_device *)calloc(1, sizeof(orc_device));
  orc_device_init(d, maxX, maxY, hx, hy, T, epsR, ni);
  return d;
}

void orc_device_delete(orc_device *d) {
  if (!d)
    return;
  orc_device_free(d);
  free(d);
}

void orc_device_info(const orc_device *d, int *ext, double *scal) {
  ext[0] = d->ext[0];
  ext[1] = d->ext[1];
  if (scal) {
    scal[0] = d->Vt;
    scal[1] = d->debye;
    scal[2] = d->cellVolume;
    scal[3] = d->normCellVolume;
  }
}

void orc_device_get_grids(const orc_device *d, double *doping, int *region,
                          int *ohmic) {
  for (size_t iy = 0; iy < (size_t)d->ext[1]; iy++)
    for (size_t ix = 0; ix < (size_t)d->ext[0]; ix++) {
      size_t i = orc_device_idx(d, ix, iy), c[2] = {ix, iy};
      if (doping)
        doping[i] = d->doping[i];
      if (region)
        region[i] = d->region[i];
      if (ohmic)
        ohmic[i] = orc_device_is_ohmic(d, c) ? orc_device_contact_idx(d, c) : -1;
    }
}

long long orc_sor_solve(const orc_device *d, double *pot, const double *conc,
                        double accuracyV, double omega, int resetBC) {
  orc_sor s;
  orc_sor_init(&s, d, accuracyV, omega);
  if (conc)
    orc_sor_nonequilibrium(&s, d, pot, conc, resetBC);
  else
    orc_sor_equilibrium(&s, d, pot, resetBC);
  long long sw = s.sweeps;
  orc_sor_free(&s);
  return sw;
}

void orc_pm_assign_h(const orc_device *d, int scheme, const double *pos, int n,
                     double nrCarriers, double *grid) {
  orc_pm_assign(scheme, pos, n, nrCarriers, d->h, d->ext, grid);
}

void orc_pm_forces_h(const orc_device *d, int scheme, const double *ex,
                     const double *ey, const double *pos, int n, double charge,
                     double *forces) {
  for (int i = 0; i < n; i++)
    orc_pm_force(scheme, ex, ey, pos + 2 * i, d->h, d->ext, charge, forces + 3 * i);
}

/* per-particle initial states from per-particle raw streams at grid nodes
 * coords[2*i] (addParticle with initParticlePos, emcBasicParticleHandler.hpp:
 * 241-254); state rows: x, y, kx, ky, kz, E, tau, grainTau, valley, sub */
void orc_device_init_replay(const orc_model *m, const orc_device *d, int np,
                            const int *coords, const uint64_t *draws,
                            size_t stride, size_t *cursor, double *state) {
  for (int i = 0; i < np; i++) {
    orc_rng r;
    orc_rng_init_buf(&r, draws + (size_t)i * stride, stride);
    r.pos = cursor[i];
    size_t c[2] = {(size_t)coords[2 * i], (size_t)coords[2 * i + 1]};
    orc_ensemble e;
    memset(&e, 0, sizeof e);
    orc_device_add_particle(m, d, c, &r, &e);
    cursor[i] = r.pos;
    double *o = state + (size_t)i * 10;
    o[0] = e.pos[0];
    o[1] = e.pos[1];
    o[2] = e.p[0].k[0];
    o[3] = e.p[0].k[1];
    o[4] = e.p[0].k[2];
    o[5] = e.p[0].energy;
    o[6] = e.p[0].tau;
    o[7] = e.p[0].grainTau;
    o[8] = e.p[0].valley;
    o[9] = e.p[0].subValley;
    free(e.p);
    free(e.pos);
  }
}

/* frozen-field device steps with per-particle raw streams; a particle that
 * leaves through a contact stops (removedStep[i] = step, contact[i]).
 * state rows as orc_device_init_replay; outputs per step (may be NULL):
 * posOut[s][i][2], kOut[s][i][3], EOut[s][i], nevOut[s][i] */
void orc_device_run_replay(const orc_model *m, const orc_device *d, int pm,
                           const double *ex, const double *ey, int np,
                           double *state, const uint64_t *draws, size_t stride,
                           size_t *cursor, int nsteps, double tStep,
                           int *removedStep, int *contact, double *posOut,
                           double *kOut, double *EOut, int32_t *nevOut) {
  for (int s = 0; s < nsteps; s++) {
    for (int i = 0; i < np; i++) {
      double *o = state + (size_t)i * 10;
      size_t so = (size_t)s * np + i;
      int32_t nev = 0;
      if (removedStep[i] < 0) {
        orc_particle p;
        memset(&p, 0, sizeof p);
        double pos[2] = {o[0], o[1]};
        p.k[0] = o[2];
        p.k[1] = o[3];
        p.k[2] = o[4];
        p.energy = o[5];
        p.tau = o[6];
        p.grainTau = o[7];
        p.valley = (int)o[8];
        p.subValley = (int)o[9];
        orc_rng r;
        orc_rng_init_buf(&r, draws + (size_t)i * stride, stride);
        r.pos = cursor[i];
        int ci = -1;
        if (orc_device_move_particle(m, d, pm, ex, ey, -ORC_Q, &p, pos, tStep, &r, &ci, &nev)) {
          removedStep[i] = s;
          contact[i] = ci;
        }
        cursor[i] = r.pos;
        o[0] = pos[0];
        o[1] = pos[1];
        o[2] = p.k[0];
        o[3] = p.k[1];
        o[4] = p.k[2];
        o[5] = p.energy;
        o[6] = p.tau;
        o[7] = p.grainTau;
        o[8] = p.valley;
        o[9] = p.subValley;
      }
      if (posOut) {
        posOut[2 * so] = o[0];
        posOut[2 * so + 1] = o[1];
      }
      if (kOut) {
        kOut[3 * so] = o[2];
        kOut[3 * so + 1] = o[3];
        kOut[3 * so + 2] = o[4];
      }
      if (EOut)
        EOut[so] = o[5];
      if (nevOut)
        nevOut[so] = nev;
    }
  }
}

int orc_device_run_h(orc_device *d, orc_model *m, int pm, double accuracyV,
                     double omega, double tStep, int nSteps, uint64_t seed,
                     double *eqPot, int *nInit, int *nPartOut, double *potOut,
                     int *exitOut, int *delOut, int *injOut, double *nrPartLast,
                     double *stateOut, int maxState, int poissonInterval, double chanX0,
                     double chanX1, double chanLen, double *chanOut, int initFromPotential) {
  orc_device_run_cfg cfg;
  memset(&cfg, 0, sizeof cfg);
  cfg.dev = d;
  cfg.model = m;
  cfg.pm = pm;
  cfg.accuracyV = accuracyV;
  cfg.omega = omega;
  cfg.tStep = tStep;
  cfg.nSteps = nSteps;
  cfg.seed = seed;
  cfg.nrCarriers = 1.0;
  cfg.poissonInterval = poissonInterval;
  cfg.chan[0] = chanX0;
  cfg.chan[1] = chanX1;
  cfg.chan[2] = chanLen;
  cfg.chanOut = chanOut;
  cfg.initFromPotential = initFromPotential;
  return orc_device_run(&cfg, eqPot, nInit, nPartOut, potOut, exitOut, delOut, injOut,
                        nrPartLast, stateOut, maxState);
}
