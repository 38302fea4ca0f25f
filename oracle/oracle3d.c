/*
 * oracle3d.c -- CPU restatement of the 3D WumingPIC step (FP64), periodic in
 * x, y and z at one rank.
 *
 * TEST INFRASTRUCTURE ONLY (see wpic_oracle.h).  Each routine follows the
 * reference routine named in its comment, statement by statement, so that at
 * one OpenMP thread the results are the reference's bit for bit under the
 * same compiler rules (IEEE double, no FMA contraction).
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#include "wpic_oracle.h"

#define NX (p->nxge - p->nxgs + 1)
#define NY (p->nye - p->nys + 1)
#define NZ (p->nze - p->nzs + 1)

static inline size_t iUP(const or3d_params *p, int idim, int ii, int j, int k, int isp) {
  return (size_t)(idim - 1) +
         (size_t)p->ndim *
             ((size_t)(ii - 1) +
              (size_t)p->np * ((size_t)(j - p->nys) + (size_t)NY * ((size_t)(k - p->nzs) + (size_t)NZ * (size_t)(isp - 1))));
}
static inline size_t iF(const or3d_params *p, int nc, int c, int i, int j, int k) {
  return (size_t)(c - 1) +
         (size_t)nc * ((size_t)(i - (p->nxgs - 2)) +
                       (size_t)(NX + 4) * ((size_t)(j - (p->nys - 2)) + (size_t)(NY + 4) * (size_t)(k - (p->nzs - 2))));
}
static inline size_t iCC(const or3d_params *p, int i, int j, int k, int isp) {
  return (size_t)(i - p->nxgs) +
         (size_t)(NX + 1) * ((size_t)(j - p->nys) + (size_t)NY * ((size_t)(k - p->nzs) + (size_t)NZ * (size_t)(isp - 1)));
}
static inline size_t iNP(const or3d_params *p, int j, int k, int isp) {
  return (size_t)(j - p->nys) + (size_t)NY * ((size_t)(k - p->nzs) + (size_t)NZ * (size_t)(isp - 1));
}
static inline size_t iGK(const or3d_params *p, int l, int i, int j, int k) {
  return (size_t)(l - 1) + 3u * ((size_t)(i - p->nxgs) + (size_t)NX * ((size_t)(j - p->nys) + (size_t)NY * (size_t)(k - p->nzs)));
}
#define UF(c, i, j, k) uf[iF(p, 6, c, i, j, k)]
#define UJ(c, i, j, k) uj[iF(p, 3, c, i, j, k)]
#define DF(c, i, j, k) df[iF(p, 6, c, i, j, k)]

/* 3d/common/field.f90:53-61 */
void or3d_field_consts(const or3d_params *p, double f[5]) {
  const double pi = M_PI;
  double c = p->c, delt = p->delt, delx = p->delx, gfac = p->gfac;
  double t = delx / (c * delt * gfac);
  f[0] = c * delt / delx;
  f[1] = gfac * f[0] * f[0];
  f[2] = 4.0 * pi * delx / c;
  f[3] = 6.0 + t * t;
  f[4] = t * t;
}

int or3d_fstate_init(const or3d_params *p, or3d_fstate *st) {
  size_t n = (size_t)(NX + 4) * (size_t)(NY + 4) * (size_t)(NZ + 4);
  st->df = (double *)calloc(6 * n, sizeof(double));
  st->gkl = (double *)calloc(3u * (size_t)NX * NY * NZ, sizeof(double));
  st->uj = (double *)calloc(3 * n, sizeof(double));
  st->cg_iters[0] = st->cg_iters[1] = st->cg_iters[2] = 0;
  return (st->df && st->gkl && st->uj) ? 0 : -1;
}

void or3d_fstate_free(or3d_fstate *st) {
  free(st->df);
  free(st->gkl);
  free(st->uj);
  st->df = st->gkl = st->uj = NULL;
}

static inline void shape3(double dh, double *m, double *z, double *pp) {
  *m = 5e-1 * (5e-1 - dh) * (5e-1 - dh);
  *z = 7.5e-1 - dh * dh;
  *pp = 5e-1 * (5e-1 + dh) * (5e-1 + dh);
}

/* particle__solv: 3d/common/particle.f90:52-233 */
static void particle_solv3(const or3d_params *p, double *gp, const double *up, const double *uf, const int *cumcnt,
                           int vay);

void or3d_particle_solv(const or3d_params *p, double *gp, const double *up, const double *uf, const int *cumcnt) {
  particle_solv3(p, gp, up, uf, cumcnt, 0);
}

/* particle__solv_vay: 3d/common/particle.f90:236-419 (the Vay kick after the
 * same staging and gather; Vay, Phys. Plasmas 15, 056701 (2008)) */
void or3d_particle_solv_vay(const or3d_params *p, double *gp, const double *up, const double *uf,
                            const int *cumcnt) {
  particle_solv3(p, gp, up, uf, cumcnt, 1);
}

static void particle_solv3(const or3d_params *p, double *gp, const double *up, const double *uf, const int *cumcnt,
                           int vay) {
  const int nxs = p->nxgs, nxe = p->nxge, nys = p->nys, nye = p->nye, nzs = p->nzs, nze = p->nze;
  const double d_delx = 1. / p->delx, delt = p->delt, c = p->c;
  const int tx = NX + 2, ty = NY + 2;
  double *tmpf = (double *)malloc(sizeof(double) * 6u * (size_t)tx * ty * (NZ + 2));
#define TF(cc, i, j, k) tmpf[(size_t)((cc)-1) + 6u * ((size_t)((i) - (nxs - 1)) + (size_t)tx * ((size_t)((j) - (nys - 1)) + (size_t)ty * (size_t)((k) - (nzs - 1))))]
#pragma omp parallel for
  for (int k = nzs - 1; k <= nze + 1; k++)
    for (int j = nys - 1; j <= nye + 1; j++)
      for (int i = nxs - 1; i <= nxe + 1; i++) {
        TF(1, i, j, k) = 2.5e-1 * (+UF(1, i, j, k) + UF(1, i, j + 1, k) + UF(1, i, j, k + 1) + UF(1, i, j + 1, k + 1));
        TF(2, i, j, k) = 2.5e-1 * (+UF(2, i, j, k) + UF(2, i + 1, j, k) + UF(2, i, j, k + 1) + UF(2, i + 1, j, k + 1));
        TF(3, i, j, k) = 2.5e-1 * (+UF(3, i, j, k) + UF(3, i + 1, j, k) + UF(3, i, j + 1, k) + UF(3, i + 1, j + 1, k));
        TF(4, i, j, k) = 5e-1 * (+UF(4, i, j, k) + UF(4, i + 1, j, k));
        TF(5, i, j, k) = 5e-1 * (+UF(5, i, j, k) + UF(5, i, j + 1, k));
        TF(6, i, j, k) = 5e-1 * (+UF(6, i, j, k) + UF(6, i, j, k + 1));
      }
#pragma omp parallel for
  for (int k = nzs; k <= nze; k++)
    for (int j = nys; j <= nye; j++)
      for (int i = nxs; i <= nxe; i++)
        for (int isp = 1; isp <= p->nsp; isp++) {
          const double q = p->q[isp - 1], r = p->r[isp - 1];
          const double fac1 = q / r * 5e-1 * delt;
          const double txxx = fac1 * fac1;
          const double fac2 = q * delt / r;
          for (int ii = cumcnt[iCC(p, i, j, k, isp)] + 1; ii <= cumcnt[iCC(p, i + 1, j, k, isp)]; ii++) {
            const double *u = &up[iUP(p, 1, ii, j, k, isp)];
            double *g = &gp[iUP(p, 1, ii, j, k, isp)];
            double shxm, shx, shxp, shym, shy, shyp, shzm, shz, shzp;
            shape3(u[0] * d_delx - 5e-1 - i, &shxm, &shx, &shxp);
            shape3(u[1] * d_delx - 5e-1 - j, &shym, &shy, &shyp);
            shape3(u[2] * d_delx - 5e-1 - k, &shzm, &shz, &shzp);
#define RX(cc, jj, kk) (+TF(cc, i - 1, jj, kk) * shxm + TF(cc, i, jj, kk) * shx + TF(cc, i + 1, jj, kk) * shxp)
#define RY(cc, kk) (+RX(cc, j - 1, kk) * shym + RX(cc, j, kk) * shy + RX(cc, j + 1, kk) * shyp)
#define G3(cc) (RY(cc, k - 1) * shzm + RY(cc, k) * shz + RY(cc, k + 1) * shzp)
            const double bpx = G3(1), bpy = G3(2), bpz = G3(3), epx = G3(4), epy = G3(5), epz = G3(6);
#undef G3
#undef RY
#undef RX
            if (vay) {
              const double w1 = u[3], w2 = u[4], w3 = u[5];
              double gam = sqrt(c * c + w1 * w1 + w2 * w2 + w3 * w3);
              const double fac1r = fac1 / gam;
              const double uvm4 = w1 + fac2 * epx + fac1r * (+w2 * bpz - w3 * bpy);
              const double uvm5 = w2 + fac2 * epy + fac1r * (+w3 * bpx - w1 * bpz);
              const double uvm6 = w3 + fac2 * epz + fac1r * (+w1 * bpy - w2 * bpx);
              const double taux = fac1 * bpx / c, tauy = fac1 * bpy / c, tauz = fac1 * bpz / c;
              const double tau2 = taux * taux + tauy * tauy + tauz * tauz;
              const double ua = (uvm4 * taux + uvm5 * tauy + uvm6 * tauz) / c;
              const double sigma = 1.0 + (uvm4 * uvm4 + uvm5 * uvm5 + uvm6 * uvm6) / (c * c) - tau2;
              const double gam2 = 0.5 * (sigma + sqrt(sigma * sigma + 4.0 * (tau2 + ua * ua)));
              gam = sqrt(gam2);
              const double s = 1.0 / (tau2 + gam2);
              g[3] = s * (gam2 * uvm4 + c * ua * taux + gam * (uvm5 * tauz - uvm6 * tauy));
              g[4] = s * (gam2 * uvm5 + c * ua * tauy + gam * (uvm6 * taux - uvm4 * tauz));
              g[5] = s * (gam2 * uvm6 + c * ua * tauz + gam * (uvm4 * tauy - uvm5 * taux));
              gam = 1.0 / gam;
              g[0] = u[0] + g[3] * delt * gam;
              g[1] = u[1] + g[4] * delt * gam;
              g[2] = u[2] + g[5] * delt * gam;
              continue;
            }
            double uvm1 = u[3] + fac1 * epx;
            double uvm2 = u[4] + fac1 * epy;
            double uvm3 = u[5] + fac1 * epz;
            double gam = sqrt(c * c + uvm1 * uvm1 + uvm2 * uvm2 + uvm3 * uvm3);
            const double igam = 1e0 / gam;
            const double fac1r = fac1 * igam;
            const double fac2r = fac2 / (gam + txxx * (bpx * bpx + bpy * bpy + bpz * bpz) * igam);
            const double uvm4 = uvm1 + fac1r * (+uvm2 * bpz - uvm3 * bpy);
            const double uvm5 = uvm2 + fac1r * (+uvm3 * bpx - uvm1 * bpz);
            const double uvm6 = uvm3 + fac1r * (+uvm1 * bpy - uvm2 * bpx);
            uvm1 = uvm1 + fac2r * (+uvm5 * bpz - uvm6 * bpy);
            uvm2 = uvm2 + fac2r * (+uvm6 * bpx - uvm4 * bpz);
            uvm3 = uvm3 + fac2r * (+uvm4 * bpy - uvm5 * bpx);
            g[3] = uvm1 + fac1 * epx;
            g[4] = uvm2 + fac1 * epy;
            g[5] = uvm3 + fac1 * epz;
            gam = 1e0 / sqrt(1e0 + (+g[3] * g[3] + g[4] * g[4] + g[5] * g[5]) / (c * c));
            g[0] = u[0] + g[3] * delt * gam;
            g[1] = u[1] + g[4] * delt * gam;
            g[2] = u[2] + g[5] * delt * gam;
          }
        }
#undef TF
  free(tmpf);
  if (p->ndim == 7) {
    size_t n = (size_t)p->np * NY * NZ * p->nsp;
#pragma omp parallel for
    for (size_t q = 0; q < n; q++) memcpy(&gp[7 * q + 6], &up[7 * q + 6], sizeof(double));
  }
}

/* ele_cur: 3d/common/field.f90:211-406.  pj arrays indexed [a][b][c] for
 * the reference's (a-2, b-2, c-2) */
static void ele_cur_cell(const or3d_params *p, double *uj, const double *up, const double *gp, const int *cumcnt,
                         int i, int j, int k) {
  const double fac = 1e0 / 3e0;
  const double d_delx = 1e0 / p->delx, d_delt = 1e0 / p->delt;
  static __thread double pjx[5][5][5], pjy[5][5][5], pjz[5][5][5];
  memset(pjx, 0, sizeof pjx);
  memset(pjy, 0, sizeof pjy);
  memset(pjz, 0, sizeof pjz);
  for (int isp = 1; isp <= p->nsp; isp++) {
    const double qdxdt = p->q[isp - 1] * p->delx * d_delt;
    for (int ii = cumcnt[iCC(p, i, j, k, isp)] + 1; ii <= cumcnt[iCC(p, i + 1, j, k, isp)]; ii++) {
      const double *u = &up[iUP(p, 1, ii, j, k, isp)];
      const double *g = &gp[iUP(p, 1, ii, j, k, isp)];
      double s0[3][5], ds[3][5];
      const int cell[3] = {i, j, k};
      for (int d = 0; d < 3; d++) {
        const double dh = u[d] * d_delx - 5e-1 - cell[d];
        s0[d][0] = 0e0;
        shape3(dh, &s0[d][1], &s0[d][2], &s0[d][3]);
        s0[d][4] = 0e0;
      }
      for (int d = 0; d < 3; d++) {
        const int i2 = (int)(g[d] * d_delx);
        const double dh = g[d] * d_delx - 5e-1 - i2;
        const int inc = i2 - cell[d];
        double s1_1, s1_2, s1_3;
        shape3(dh, &s1_1, &s1_2, &s1_3);
        /* smo_1 = -(inc-abs(inc))*5d-1+0 etc.: exactly 0 or 1, never -0 */
        const double smo_1 = (inc - abs(inc)) != 0 ? -((double)(inc - abs(inc)) * 5e-1) + 0 : 0e0;
        const double smo_2 = (double)(-abs(inc) + 1);
        const double smo_3 = (double)(inc + abs(inc)) * 5e-1 + 0;
        ds[d][0] = s1_1 * smo_1;
        ds[d][1] = s1_1 * smo_2 + s1_2 * smo_1;
        ds[d][2] = s1_2 * smo_2 + s1_3 * smo_1 + s1_1 * smo_3;
        ds[d][3] = s1_3 * smo_2 + s1_2 * smo_3;
        ds[d][4] = s1_3 * smo_3;
      }
      for (int d = 0; d < 3; d++)
        for (int m = 0; m < 5; m++) ds[d][m] = ds[d][m] - s0[d][m];
      const double *s0x = s0[0], *s0y = s0[1], *s0z = s0[2], *dsx = ds[0], *dsy = ds[1], *dsz = ds[2];
      for (int kp = 0; kp < 5; kp++)
        for (int jp = 0; jp < 5; jp++) {
          double dstmp, pjtmp;
          /* pjx */
          dstmp = ((s0y[jp] + 5e-1 * dsy[jp]) * s0z[kp] + (5e-1 * s0y[jp] + fac * dsy[jp]) * dsz[kp]) * qdxdt;
          pjtmp = -dsx[0] * dstmp;
          pjx[1][jp][kp] = pjx[1][jp][kp] + pjtmp;
          pjtmp = pjtmp - dsx[1] * dstmp;
          pjx[2][jp][kp] = pjx[2][jp][kp] + pjtmp;
          pjtmp = pjtmp - dsx[2] * dstmp;
          pjx[3][jp][kp] = pjx[3][jp][kp] + pjtmp;
          pjtmp = pjtmp - dsx[3] * dstmp;
          pjx[4][jp][kp] = pjx[4][jp][kp] + pjtmp;
          /* pjy */
          dstmp = ((s0x[jp] + 5e-1 * dsx[jp]) * s0z[kp] + (5e-1 * s0x[jp] + fac * dsx[jp]) * dsz[kp]) * qdxdt;
          pjtmp = -dsy[0] * dstmp;
          pjy[1][jp][kp] = pjy[1][jp][kp] + pjtmp;
          pjtmp = pjtmp - dsy[1] * dstmp;
          pjy[2][jp][kp] = pjy[2][jp][kp] + pjtmp;
          pjtmp = pjtmp - dsy[2] * dstmp;
          pjy[3][jp][kp] = pjy[3][jp][kp] + pjtmp;
          pjtmp = pjtmp - dsy[3] * dstmp;
          pjy[4][jp][kp] = pjy[4][jp][kp] + pjtmp;
          /* pjz */
          dstmp = ((s0x[jp] + 5e-1 * dsx[jp]) * s0y[kp] + (5e-1 * s0x[jp] + fac * dsx[jp]) * dsy[kp]) * qdxdt;
          pjtmp = -dsz[0] * dstmp;
          pjz[1][jp][kp] = pjz[1][jp][kp] + pjtmp;
          pjtmp = pjtmp - dsz[1] * dstmp;
          pjz[2][jp][kp] = pjz[2][jp][kp] + pjtmp;
          pjtmp = pjtmp - dsz[2] * dstmp;
          pjz[3][jp][kp] = pjz[3][jp][kp] + pjtmp;
          pjtmp = pjtmp - dsz[3] * dstmp;
          pjz[4][jp][kp] = pjz[4][jp][kp] + pjtmp;
        }
    }
  }
  for (int kp = -2; kp <= 2; kp++)
    for (int jp = -2; jp <= 2; jp++)
      for (int ip = -2; ip <= 2; ip++) {
        UJ(1, i + ip, j + jp, k + kp) = UJ(1, i + ip, j + jp, k + kp) + pjx[ip + 2][jp + 2][kp + 2];
        UJ(2, i + ip, j + jp, k + kp) = UJ(2, i + ip, j + jp, k + kp) + pjy[jp + 2][ip + 2][kp + 2];
        UJ(3, i + ip, j + jp, k + kp) = UJ(3, i + ip, j + jp, k + kp) + pjz[kp + 2][ip + 2][jp + 2];
      }
}

void or3d_ele_cur(const or3d_params *p, double *uj, const double *up, const double *gp, const int *cumcnt) {
  const size_t n3 = 3u * (size_t)(NX + 4) * (NY + 4) * (NZ + 4);
  memset(uj, 0, n3 * sizeof(double));
  int nth = 1;
#ifdef _OPENMP
  nth = omp_get_max_threads();
#endif
  if (nth <= 1) {
    for (int k = p->nzs; k <= p->nze; k++)
      for (int j = p->nys; j <= p->nye; j++)
        for (int i = p->nxgs; i <= p->nxge; i++) ele_cur_cell(p, uj, up, gp, cumcnt, i, j, k);
    return;
  }
#pragma omp parallel
  {
    double *priv = (double *)calloc(n3, sizeof(double));
#pragma omp for schedule(static)
    for (int k = p->nzs; k <= p->nze; k++)
      for (int j = p->nys; j <= p->nye; j++)
        for (int i = p->nxgs; i <= p->nxge; i++) ele_cur_cell(p, priv, up, gp, cumcnt, i, j, k);
#pragma omp critical
    for (size_t q = 0; q < n3; q++) uj[q] = uj[q] + priv[q];
    free(priv);
  }
}

/* boundary_periodic__curre at one rank: 3d/common/boundary_periodic.f90:676-978 */
void or3d_bc_curre(const or3d_params *p, double *uj) {
  const int nxs = p->nxgs, nxe = p->nxge, nys = p->nys, nye = p->nye, nzs = p->nzs, nze = p->nze;
  /* y fold of two ghost rows, for all k including ghosts */
  for (int k = nzs - 2; k <= nze + 2; k++)
    for (int i = nxs - 2; i <= nxe + 2; i++)
      for (int c = 1; c <= 3; c++) {
        const double a = UJ(c, i, nys - 2, k), b = UJ(c, i, nys - 1, k);
        UJ(c, i, nye - 1, k) = UJ(c, i, nye - 1, k) + a;
        UJ(c, i, nye, k) = UJ(c, i, nye, k) + b;
      }
  for (int k = nzs - 2; k <= nze + 2; k++)
    for (int i = nxs - 2; i <= nxe + 2; i++)
      for (int c = 1; c <= 3; c++) {
        const double a = UJ(c, i, nye + 1, k), b = UJ(c, i, nye + 2, k);
        UJ(c, i, nys, k) = UJ(c, i, nys, k) + a;
        UJ(c, i, nys + 1, k) = UJ(c, i, nys + 1, k) + b;
      }
  /* z fold of two ghost planes, for interior j */
  for (int j = nys; j <= nye; j++)
    for (int i = nxs - 2; i <= nxe + 2; i++)
      for (int c = 1; c <= 3; c++) {
        const double a = UJ(c, i, j, nzs - 2), b = UJ(c, i, j, nzs - 1);
        UJ(c, i, j, nze - 1) = UJ(c, i, j, nze - 1) + a;
        UJ(c, i, j, nze) = UJ(c, i, j, nze) + b;
      }
  for (int j = nys; j <= nye; j++)
    for (int i = nxs - 2; i <= nxe + 2; i++)
      for (int c = 1; c <= 3; c++) {
        const double a = UJ(c, i, j, nze + 1), b = UJ(c, i, j, nze + 2);
        UJ(c, i, j, nzs) = UJ(c, i, j, nzs) + a;
        UJ(c, i, j, nzs + 1) = UJ(c, i, j, nzs + 1) + b;
      }
  /* one ghost layer copied back: y for interior k, then z for j in nys-1..nye+1 */
  for (int k = nzs; k <= nze; k++)
    for (int i = nxs - 2; i <= nxe + 2; i++)
      for (int c = 1; c <= 3; c++) UJ(c, i, nye + 1, k) = UJ(c, i, nys, k);
  for (int k = nzs; k <= nze; k++)
    for (int i = nxs - 2; i <= nxe + 2; i++)
      for (int c = 1; c <= 3; c++) UJ(c, i, nys - 1, k) = UJ(c, i, nye, k);
  for (int j = nys - 1; j <= nye + 1; j++)
    for (int i = nxs - 2; i <= nxe + 2; i++)
      for (int c = 1; c <= 3; c++) UJ(c, i, j, nze + 1) = UJ(c, i, j, nzs);
  for (int j = nys - 1; j <= nye + 1; j++)
    for (int i = nxs - 2; i <= nxe + 2; i++)
      for (int c = 1; c <= 3; c++) UJ(c, i, j, nzs - 1) = UJ(c, i, j, nze);
  /* x fold then x copy, two layers, all j and k */
  for (int k = nzs - 2; k <= nze + 2; k++)
    for (int j = nys - 2; j <= nye + 2; j++)
      for (int c = 1; c <= 3; c++) {
        UJ(c, nxe - 1, j, k) = UJ(c, nxe - 1, j, k) + UJ(c, nxs - 2, j, k);
        UJ(c, nxe, j, k) = UJ(c, nxe, j, k) + UJ(c, nxs - 1, j, k);
        UJ(c, nxs, j, k) = UJ(c, nxs, j, k) + UJ(c, nxe + 1, j, k);
        UJ(c, nxs + 1, j, k) = UJ(c, nxs + 1, j, k) + UJ(c, nxe + 2, j, k);
      }
  for (int k = nzs - 2; k <= nze + 2; k++)
    for (int j = nys - 2; j <= nye + 2; j++)
      for (int c = 1; c <= 3; c++) {
        UJ(c, nxs - 2, j, k) = UJ(c, nxe - 1, j, k);
        UJ(c, nxs - 1, j, k) = UJ(c, nxe, j, k);
        UJ(c, nxe + 1, j, k) = UJ(c, nxs, j, k);
        UJ(c, nxe + 2, j, k) = UJ(c, nxs + 1, j, k);
      }
}

/* boundary_periodic__dfield at one rank: 3d/common/boundary_periodic.f90:458-673 */
void or3d_bc_dfield(const or3d_params *p, double *df) {
  const int nxs = p->nxgs, nxe = p->nxge, nys = p->nys, nye = p->nye, nzs = p->nzs, nze = p->nze;
  for (int k = nzs; k <= nze; k++)
    for (int i = nxs; i <= nxe; i++)
      for (int c = 1; c <= 6; c++) {
        DF(c, i, nye + 1, k) = DF(c, i, nys, k);
        DF(c, i, nye + 2, k) = DF(c, i, nys + 1, k);
      }
  for (int k = nzs; k <= nze; k++)
    for (int i = nxs; i <= nxe; i++)
      for (int c = 1; c <= 6; c++) {
        DF(c, i, nys - 2, k) = DF(c, i, nye - 1, k);
        DF(c, i, nys - 1, k) = DF(c, i, nye, k);
      }
  for (int j = nys - 2; j <= nye + 2; j++)
    for (int i = nxs; i <= nxe; i++)
      for (int c = 1; c <= 6; c++) {
        DF(c, i, j, nze + 1) = DF(c, i, j, nzs);
        DF(c, i, j, nze + 2) = DF(c, i, j, nzs + 1);
      }
  for (int j = nys - 2; j <= nye + 2; j++)
    for (int i = nxs; i <= nxe; i++)
      for (int c = 1; c <= 6; c++) {
        DF(c, i, j, nzs - 2) = DF(c, i, j, nze - 1);
        DF(c, i, j, nzs - 1) = DF(c, i, j, nze);
      }
  for (int k = nzs - 2; k <= nze + 2; k++)
    for (int j = nys - 2; j <= nye + 2; j++)
      for (int c = 1; c <= 6; c++) {
        DF(c, nxs - 2, j, k) = DF(c, nxe - 1, j, k);
        DF(c, nxs - 1, j, k) = DF(c, nxe, j, k);
        DF(c, nxe + 1, j, k) = DF(c, nxs, j, k);
        DF(c, nxe + 2, j, k) = DF(c, nxs + 1, j, k);
      }
}

/* boundary_periodic__phi at one rank: 3d/common/boundary_periodic.f90:981-1099
 * phi(nxs-1:nxe+1, nys-1:nye+1, nzs-1:nze+1) */
void or3d_bc_phi(const or3d_params *p, double *phi, int l) {
  (void)l;
  const int nxs = p->nxgs, nxe = p->nxge, nys = p->nys, nye = p->nye, nzs = p->nzs, nze = p->nze;
  const int pw = NX + 2, ph = NY + 2;
#define PH(i, j, k) phi[(size_t)((i) - (nxs - 1)) + (size_t)pw * ((size_t)((j) - (nys - 1)) + (size_t)ph * (size_t)((k) - (nzs - 1)))]
  for (int k = nzs; k <= nze; k++)
    for (int i = nxs; i <= nxe; i++) PH(i, nye + 1, k) = PH(i, nys, k);
  for (int k = nzs; k <= nze; k++)
    for (int i = nxs; i <= nxe; i++) PH(i, nys - 1, k) = PH(i, nye, k);
  for (int j = nys - 1; j <= nye + 1; j++)
    for (int i = nxs; i <= nxe; i++) PH(i, j, nze + 1) = PH(i, j, nzs);
  for (int j = nys - 1; j <= nye + 1; j++)
    for (int i = nxs; i <= nxe; i++) PH(i, j, nzs - 1) = PH(i, j, nze);
  for (int k = nzs - 1; k <= nze + 1; k++)
    for (int j = nys - 1; j <= nye + 1; j++) {
      PH(nxs - 1, j, k) = PH(nxe, j, k);
      PH(nxe + 1, j, k) = PH(nxs, j, k);
    }
#undef PH
}

/* cgm: 3d/common/field.f90:409-560 (serial dot products in (k, j, i) order) */
int or3d_cgm(const or3d_params *p, or3d_fstate *st) {
  const int nxs = p->nxgs, nxe = p->nxge, nys = p->nys, nye = p->nye, nzs = p->nzs, nze = p->nze;
  double f[5];
  or3d_field_consts(p, f);
  const double f4 = f[3], f5 = f[4];
  const int pw = NX + 2, ph = NY + 2;
  const size_t np_ = (size_t)pw * ph * (NZ + 2), nin = (size_t)NX * NY * NZ;
  double *phi = (double *)calloc(np_, sizeof(double)), *pp = (double *)calloc(np_, sizeof(double));
  double *r = (double *)calloc(nin, sizeof(double)), *b = (double *)calloc(nin, sizeof(double));
  double *ap = (double *)calloc(nin, sizeof(double));
  double *df = st->df;
#define PH(a, i, j, k) a[(size_t)((i) - (nxs - 1)) + (size_t)pw * ((size_t)((j) - (nys - 1)) + (size_t)ph * (size_t)((k) - (nzs - 1)))]
#define IN(a, i, j, k) a[(size_t)((i) - nxs) + (size_t)NX * ((size_t)((j) - nys) + (size_t)NY * (size_t)((k) - nzs))]
  int rc = 0;
  for (int l = 1; l <= 3; l++) {
    int ite = 0;
    double sum = 0e0;
    for (int k = nzs; k <= nze; k++)
      for (int j = nys; j <= nye; j++)
        for (int i = nxs; i <= nxe; i++) {
          PH(phi, i, j, k) = DF(l, i, j, k);
          IN(b, i, j, k) = f5 * st->gkl[iGK(p, l, i, j, k)];
          sum = sum + IN(b, i, j, k) * IN(b, i, j, k);
        }
    double sum_g = sum;
    const double eps = sqrt(sum_g) * or_cg_err();
    or3d_bc_phi(p, phi, l);
    double sumr = 0e0;
    for (int k = nzs; k <= nze; k++)
      for (int j = nys; j <= nye; j++)
        for (int i = nxs; i <= nxe; i++) {
          IN(r, i, j, k) = IN(b, i, j, k) + PH(phi, i, j, k - 1) + PH(phi, i, j - 1, k) + PH(phi, i - 1, j, k) -
                           f4 * PH(phi, i, j, k) + PH(phi, i + 1, j, k) + PH(phi, i, j + 1, k) + PH(phi, i, j, k + 1);
          PH(pp, i, j, k) = IN(r, i, j, k);
          sumr = sumr + IN(r, i, j, k) * IN(r, i, j, k);
        }
    double sumr_g = sumr;
    if (sqrt(sumr_g) > eps) {
      while (sum_g > eps) {
        ite = ite + 1;
        or3d_bc_phi(p, pp, l);
        double s_r = 0e0, sum2 = 0e0;
        for (int k = nzs; k <= nze; k++)
          for (int j = nys; j <= nye; j++)
            for (int i = nxs; i <= nxe; i++) {
              IN(ap, i, j, k) = -PH(pp, i, j, k - 1) - PH(pp, i, j - 1, k) - PH(pp, i - 1, j, k) +
                                f4 * PH(pp, i, j, k) - PH(pp, i + 1, j, k) - PH(pp, i, j + 1, k) - PH(pp, i, j, k + 1);
              s_r = s_r + IN(r, i, j, k) * IN(r, i, j, k);
              sum2 = sum2 + PH(pp, i, j, k) * IN(ap, i, j, k);
            }
        sumr_g = s_r;
        const double av = sumr_g / sum2;
        for (int k = nzs; k <= nze; k++)
          for (int j = nys; j <= nye; j++)
            for (int i = nxs; i <= nxe; i++) {
              PH(phi, i, j, k) = PH(phi, i, j, k) + av * PH(pp, i, j, k);
              IN(r, i, j, k) = IN(r, i, j, k) - av * IN(ap, i, j, k);
            }
        sum_g = sqrt(sumr_g);
        if (ite >= 100) {
          rc = -1;
          break;
        }
        double sum1 = 0e0;
        for (int k = nzs; k <= nze; k++)
          for (int j = nys; j <= nye; j++)
            for (int i = nxs; i <= nxe; i++) sum1 = sum1 + IN(r, i, j, k) * IN(r, i, j, k);
        const double bv = sum1 / sumr_g;
        for (int k = nzs; k <= nze; k++)
          for (int j = nys; j <= nye; j++)
            for (int i = nxs; i <= nxe; i++) PH(pp, i, j, k) = IN(r, i, j, k) + bv * PH(pp, i, j, k);
      }
    }
    st->cg_iters[l - 1] = ite;
    if (rc) break;
    for (int k = nzs; k <= nze; k++)
      for (int j = nys; j <= nye; j++)
        for (int i = nxs; i <= nxe; i++) DF(l, i, j, k) = PH(phi, i, j, k);
  }
#undef PH
#undef IN
  free(phi);
  free(pp);
  free(r);
  free(b);
  free(ap);
  return rc;
}

/* field__fdtd_i: 3d/common/field.f90:70-208 */
int or3d_field_fdtd_i(const or3d_params *p, or3d_fstate *st, double *uf, const double *up, const double *gp,
                      const int *cumcnt) {
  double *uj = st->uj;
  or3d_ele_cur(p, uj, up, gp, cumcnt);
  or3d_bc_curre(p, uj);
  return or3d_field_from_current(p, st, uf);
}

/* 3d/common/field.f90:127-206: the RHS, the three solves, E, the ghost fills
 * and the update from the current in st->uj (after its boundary fold) */
int or3d_field_from_current(const or3d_params *p, or3d_fstate *st, double *uf) {
  const int nxs = p->nxgs, nxe = p->nxge, nys = p->nys, nye = p->nye, nzs = p->nzs, nze = p->nze;
  const double pi = M_PI, delt = p->delt, gfac = p->gfac;
  double f[5];
  or3d_field_consts(p, f);
  const double f1 = f[0], f2 = f[1], f3 = f[2];
  double *uj = st->uj, *df = st->df;
#pragma omp parallel for
  for (int k = nzs; k <= nze; k++)
    for (int j = nys; j <= nye; j++)
      for (int i = nxs; i <= nxe; i++) {
        st->gkl[iGK(p, 1, i, j, k)] =
            +f2 * (+UF(1, i, j, k - 1) + UF(1, i, j - 1, k) + UF(1, i - 1, j, k) - 6e0 * UF(1, i, j, k) +
                   UF(1, i + 1, j, k) + UF(1, i, j + 1, k) + UF(1, i, j, k + 1) +
                   f3 * (-UJ(3, i, j - 1, k) + UJ(3, i, j, k) + UJ(2, i, j, k - 1) - UJ(2, i, j, k))) -
            f1 * (-UF(6, i, j - 1, k) + UF(6, i, j, k) + UF(5, i, j, k - 1) - UF(5, i, j, k));
        st->gkl[iGK(p, 2, i, j, k)] =
            +f2 * (+UF(2, i, j, k - 1) + UF(2, i, j - 1, k) + UF(2, i - 1, j, k) - 6e0 * UF(2, i, j, k) +
                   UF(2, i + 1, j, k) + UF(2, i, j + 1, k) + UF(2, i, j, k + 1) +
                   f3 * (-UJ(1, i, j, k - 1) + UJ(1, i, j, k) + UJ(3, i - 1, j, k) - UJ(3, i, j, k))) -
            f1 * (-UF(4, i, j, k - 1) + UF(4, i, j, k) + UF(6, i - 1, j, k) - UF(6, i, j, k));
        st->gkl[iGK(p, 3, i, j, k)] =
            +f2 * (+UF(3, i, j, k - 1) + UF(3, i, j - 1, k) + UF(3, i - 1, j, k) - 6e0 * UF(3, i, j, k) +
                   UF(3, i + 1, j, k) + UF(3, i, j + 1, k) + UF(3, i, j, k + 1) +
                   f3 * (-UJ(2, i - 1, j, k) + UJ(2, i, j, k) + UJ(1, i, j - 1, k) - UJ(1, i, j, k))) -
            f1 * (-UF(5, i - 1, j, k) + UF(5, i, j, k) + UF(4, i, j - 1, k) - UF(4, i, j, k));
      }
  if (or3d_cgm(p, st)) return -1;
  or3d_bc_dfield(p, df);
#pragma omp parallel for
  for (int k = nzs; k <= nze; k++)
    for (int j = nys; j <= nye; j++)
      for (int i = nxs; i <= nxe; i++) {
        DF(4, i, j, k) = +f1 * (+gfac * (-DF(3, i, j, k) + DF(3, i, j + 1, k) + DF(2, i, j, k) - DF(2, i, j, k + 1)) +
                                (-UF(3, i, j, k) + UF(3, i, j + 1, k) + UF(2, i, j, k) - UF(2, i, j, k + 1))) -
                         4e0 * pi * delt * UJ(1, i, j, k);
        DF(5, i, j, k) = +f1 * (+gfac * (-DF(1, i, j, k) + DF(1, i, j, k + 1) + DF(3, i, j, k) - DF(3, i + 1, j, k)) +
                                (-UF(1, i, j, k) + UF(1, i, j, k + 1) + UF(3, i, j, k) - UF(3, i + 1, j, k))) -
                         4e0 * pi * delt * UJ(2, i, j, k);
        DF(6, i, j, k) = +f1 * (+gfac * (-DF(2, i, j, k) + DF(2, i + 1, j, k) + DF(1, i, j, k) - DF(1, i, j + 1, k)) +
                                (-UF(2, i, j, k) + UF(2, i + 1, j, k) + UF(1, i, j, k) - UF(1, i, j + 1, k))) -
                         4e0 * pi * delt * UJ(3, i, j, k);
      }
  or3d_bc_dfield(p, df);
  const size_t n6 = 6u * (size_t)(NX + 4) * (NY + 4) * (NZ + 4);
#pragma omp parallel for
  for (size_t q = 0; q < n6; q++) uf[q] = uf[q] + df[q];
  return 0;
}

/* boundary_periodic__particle_x: 3d/common/boundary_periodic.f90:68-101 */
void or3d_bc_particle_x(const or3d_params *p, double *up, const int *np2) {
  const double d_delx = 1e0 / p->delx;
  for (int isp = 1; isp <= p->nsp; isp++)
#pragma omp parallel for
    for (int k = p->nzs; k <= p->nze; k++)
      for (int j = p->nys; j <= p->nye; j++)
        for (int ii = 1; ii <= np2[iNP(p, j, k, isp)]; ii++) {
          double *u = &up[iUP(p, 1, ii, j, k, isp)];
          const int ipos = (int)(u[0] * d_delx);
          if (ipos < p->nxgs)
            u[0] = u[0] + (p->nxge - p->nxgs + 1) * p->delx;
          else if (ipos >= p->nxge + 1)
            u[0] = u[0] - (p->nxge - p->nxgs + 1) * p->delx;
        }
}

/* boundary_periodic__particle_yz: 3d/common/boundary_periodic.f90:104-455 at
 * one rank and one thread.  Buckets (jb, kb) for jb in nys-1..nye+1 and kb in
 * nzs-1..nze+1; the y transfers move the jb = nys-1 / nye+1 buckets (all kb)
 * onto nye / nys, then the z transfers move kb = nzs-1 / nze+1 (interior jb)
 * onto nze / nzs, each appended in order. */
typedef struct {
  double *d;
  int n, cap;
} obuf;

static int ob_push(obuf *b, const double *v, int ndim) {
  if (b->n + 1 > b->cap) {
    int nc = b->cap ? 2 * b->cap : 16;
    double *nd = (double *)realloc(b->d, sizeof(double) * (size_t)ndim * nc);
    if (!nd) return -1;
    b->d = nd;
    b->cap = nc;
  }
  memcpy(&b->d[(size_t)ndim * b->n], v, sizeof(double) * ndim);
  b->n++;
  return 0;
}

static int ob_append(obuf *dst, const obuf *src, int ndim) {
  for (int q = 0; q < src->n; q++)
    if (ob_push(dst, &src->d[(size_t)ndim * q], ndim)) return -1;
  return 0;
}

int or3d_bc_particle_yz(const or3d_params *p, double *up, int *np2) {
  const int nys = p->nys, nye = p->nye, nzs = p->nzs, nze = p->nze, ndim = p->ndim;
  const double d_delx = 1e0 / p->delx;
  const int bw = NY + 2, bh = NZ + 2;
  obuf *bff = (obuf *)calloc((size_t)bw * bh, sizeof(obuf));
  int *flag = (int *)malloc(sizeof(int) * (size_t)p->np * NY * NZ);
  int *cnt2 = (int *)calloc((size_t)NY * NZ, sizeof(int));
  int rc = 0;
#define B(jb, kb) bff[(size_t)((jb) - (nys - 1)) + (size_t)bw * (size_t)((kb) - (nzs - 1))]
#define FLAG(q, j, k) flag[(size_t)((q)-1) + (size_t)p->np * ((size_t)((j)-nys) + (size_t)NY * (size_t)((k)-nzs))]
#define CNT2(j, k) cnt2[(size_t)((j)-nys) + (size_t)NY * (size_t)((k)-nzs)]
  for (int isp = 1; isp <= p->nsp && !rc; isp++) {
    for (size_t q = 0; q < (size_t)bw * bh; q++) bff[q].n = 0;
    memset(cnt2, 0, sizeof(int) * (size_t)NY * NZ);
    for (int k = nzs; k <= nze && !rc; k++)
      for (int j = nys; j <= nye && !rc; j++)
        for (int ii = 1; ii <= np2[iNP(p, j, k, isp)]; ii++) {
          double *u = &up[iUP(p, 1, ii, j, k, isp)];
          const int jpos = (int)(u[1] * d_delx), kpos = (int)(u[2] * d_delx);
          if (!(jpos == j && kpos == k)) {
            if (jpos <= p->nygs - 1)
              u[1] = u[1] + (p->nyge - p->nygs + 1) * p->delx;
            else if (jpos >= p->nyge + 1)
              u[1] = u[1] - (p->nyge - p->nygs + 1) * p->delx;
            if (kpos <= p->nzgs - 1)
              u[2] = u[2] + (p->nzge - p->nzgs + 1) * p->delx;
            else if (kpos >= p->nzge + 1)
              u[2] = u[2] - (p->nzge - p->nzgs + 1) * p->delx;
            if (jpos < nys - 1 || jpos > nye + 1 || kpos < nzs - 1 || kpos > nze + 1) {
              rc = -1;
              break;
            }
            if (ob_push(&B(jpos, kpos), u, ndim)) {
              rc = -1;
              break;
            }
            CNT2(j, k) = CNT2(j, k) + 1;
            FLAG(CNT2(j, k), j, k) = ii;
          }
        }
    if (rc) break;
    for (int kb = nzs - 1; kb <= nze + 1; kb++) {
      rc |= ob_append(&B(nye, kb), &B(nys - 1, kb), ndim);
      B(nys - 1, kb).n = 0;
    }
    for (int kb = nzs - 1; kb <= nze + 1; kb++) {
      rc |= ob_append(&B(nys, kb), &B(nye + 1, kb), ndim);
      B(nye + 1, kb).n = 0;
    }
    for (int j = nys; j <= nye; j++) {
      rc |= ob_append(&B(j, nze), &B(j, nzs - 1), ndim);
      B(j, nzs - 1).n = 0;
    }
    for (int j = nys; j <= nye; j++) {
      rc |= ob_append(&B(j, nzs), &B(j, nze + 1), ndim);
      B(j, nze + 1).n = 0;
    }
    if (rc) break;
    for (int k = nzs; k <= nze && !rc; k++)
      for (int j = nys; j <= nye; j++) {
        int *n2 = &np2[iNP(p, j, k, isp)];
        int cnt = B(j, k).n;
        int iii = 0, cnt_tmp = CNT2(j, k);
        for (int ii = 1; ii <= CNT2(j, k); ii++) {
          if (cnt == 0) {
            if (*n2 < FLAG(ii, j, k)) break;
            int out = 0;
            while (*n2 == FLAG(cnt_tmp, j, k)) {
              *n2 = *n2 - 1;
              if (*n2 < FLAG(ii, j, k)) {
                out = 1;
                break;
              }
              cnt_tmp = cnt_tmp - 1;
            }
            if (out) break;
            memcpy(&up[iUP(p, 1, FLAG(ii, j, k), j, k, isp)], &up[iUP(p, 1, *n2, j, k, isp)],
                   sizeof(double) * ndim);
            *n2 = *n2 - 1;
          } else {
            memcpy(&up[iUP(p, 1, FLAG(ii, j, k), j, k, isp)], &B(j, k).d[(size_t)ndim * iii], sizeof(double) * ndim);
            iii = iii + 1;
            cnt = cnt - 1;
          }
        }
        if (cnt > 0) {
          if (*n2 + cnt > p->np) {
            rc = -2;
            break;
          }
          for (int q = 1; q <= cnt; q++)
            memcpy(&up[iUP(p, 1, *n2 + q, j, k, isp)], &B(j, k).d[(size_t)ndim * (iii + q - 1)],
                   sizeof(double) * ndim);
        }
        *n2 = *n2 + cnt;
      }
  }
#undef B
#undef FLAG
#undef CNT2
  for (size_t q = 0; q < (size_t)bw * bh; q++) free(bff[q].d);
  free(bff);
  free(flag);
  free(cnt2);
  return rc;
}

/* sort__bucket: 3d/common/sort.f90:40-88 */
void or3d_sort_bucket(const or3d_params *p, double *gp, const double *up, int *cumcnt, const int *np2) {
  const int nxs = p->nxgs, nxe = p->nxge;
  for (int isp = 1; isp <= p->nsp; isp++)
#pragma omp parallel for
    for (int k = p->nzs; k <= p->nze; k++)
      for (int j = p->nys; j <= p->nye; j++) {
        int *cnt = (int *)calloc((size_t)(nxe - nxs + 1), sizeof(int));
        int *sum_cnt = (int *)calloc((size_t)(nxe - nxs + 2), sizeof(int));
        const int n = np2[iNP(p, j, k, isp)];
        for (int ii = 1; ii <= n; ii++) {
          const int i = (int)up[iUP(p, 1, ii, j, k, isp)];
          cnt[i - nxs] = cnt[i - nxs] + 1;
        }
        sum_cnt[0] = 0;
        cumcnt[iCC(p, nxs, j, k, isp)] = 0;
        for (int i = nxs + 1; i <= nxe + 1; i++) {
          sum_cnt[i - nxs] = sum_cnt[i - 1 - nxs] + cnt[i - 1 - nxs];
          cumcnt[iCC(p, i, j, k, isp)] = sum_cnt[i - nxs];
        }
        for (int ii = 1; ii <= n; ii++) {
          const int i = (int)up[iUP(p, 1, ii, j, k, isp)];
          memcpy(&gp[iUP(p, 1, sum_cnt[i - nxs] + 1, j, k, isp)], &up[iUP(p, 1, ii, j, k, isp)],
                 sizeof(double) * (size_t)p->ndim);
          sum_cnt[i - nxs] = sum_cnt[i - nxs] + 1;
        }
        free(cnt);
        free(sum_cnt);
      }
}

int or3d_step(const or3d_params *p, or3d_fstate *st, double *up, double *gp, double *uf, int *cumcnt, int *np2,
              int nsteps) {
  for (int it = 0; it < nsteps; it++) {
    or3d_particle_solv(p, gp, up, uf, cumcnt);
    if (or3d_field_fdtd_i(p, st, uf, up, gp, cumcnt)) return -1;
    or3d_bc_particle_x(p, gp, np2);
    if (or3d_bc_particle_yz(p, gp, np2)) return -1;
    or3d_sort_bucket(p, up, gp, cumcnt, np2);
  }
  return 0;
}

void or3d_energy(const or3d_params *p, const double *up, const double *uf, const int *np2, double out[3]) {
  const double pi = M_PI, c = p->c;
  double kin = 0e0;
  for (int isp = 1; isp <= p->nsp; isp++) {
    double vene = 0e0;
    for (int k = p->nzs; k <= p->nze; k++)
      for (int j = p->nys; j <= p->nye; j++)
        for (int ii = 1; ii <= np2[iNP(p, j, k, isp)]; ii++) {
          const double *u = &up[iUP(p, 1, ii, j, k, isp)];
          const double u2 = u[3] * u[3] + u[4] * u[4] + u[5] * u[5];
          vene = vene + p->r[isp - 1] * (sqrt(1 + u2 / (c * c)) - 1);
        }
    kin = kin + vene;
  }
  double e = 0e0, b = 0e0;
  for (int k = p->nzs; k <= p->nze; k++)
    for (int j = p->nys; j <= p->nye; j++)
      for (int i = p->nxgs; i <= p->nxge; i++) {
        b = b + UF(1, i, j, k) * UF(1, i, j, k) + UF(2, i, j, k) * UF(2, i, j, k) + UF(3, i, j, k) * UF(3, i, j, k);
        e = e + UF(4, i, j, k) * UF(4, i, j, k) + UF(5, i, j, k) * UF(5, i, j, k) + UF(6, i, j, k) * UF(6, i, j, k);
      }
  out[0] = kin;
  out[1] = e / (8 * pi);
  out[2] = b / (8 * pi);
}

void or3d_gauss(const or3d_params *p, double *G, const double *uf, const double *up, const int *np2) {
  const int nxs = p->nxgs, nys = p->nys, nzs = p->nzs;
  const int nx = NX, ny = NY, nz = NZ;
  const double pi = M_PI, d_delx = 1e0 / p->delx;
  double *rho = (double *)calloc((size_t)nx * ny * nz, sizeof(double));
  for (int isp = 1; isp <= p->nsp; isp++)
    for (int k = p->nzs; k <= p->nze; k++)
      for (int j = p->nys; j <= p->nye; j++)
        for (int ii = 1; ii <= np2[iNP(p, j, k, isp)]; ii++) {
          const double *u = &up[iUP(p, 1, ii, j, k, isp)];
          int c0[3];
          double s[3][3];
          for (int d = 0; d < 3; d++) {
            c0[d] = (int)(u[d] * d_delx);
            shape3(u[d] * d_delx - 5e-1 - c0[d], &s[d][0], &s[d][1], &s[d][2]);
          }
          for (int cz = 0; cz < 3; cz++)
            for (int cy = 0; cy < 3; cy++)
              for (int cx = 0; cx < 3; cx++) {
                const int ix = ((c0[0] - 1 + cx - nxs) % nx + nx) % nx;
                const int iy = ((c0[1] - 1 + cy - nys) % ny + ny) % ny;
                const int iz = ((c0[2] - 1 + cz - nzs) % nz + nz) % nz;
                rho[ix + (size_t)nx * (iy + (size_t)ny * iz)] += p->q[isp - 1] * s[0][cx] * s[1][cy] * s[2][cz];
              }
        }
  for (int k = 0; k < nz; k++)
    for (int j = 0; j < ny; j++)
      for (int i = 0; i < nx; i++) {
        const int I = nxs + i, J = nys + j, K = nzs + k;
        const double dive = (UF(4, I + 1, J, K) - UF(4, I, J, K)) * d_delx + (UF(5, I, J + 1, K) - UF(5, I, J, K)) * d_delx +
                            (UF(6, I, J, K + 1) - UF(6, I, J, K)) * d_delx;
        G[i + (size_t)nx * (j + (size_t)ny * k)] = dive - 4e0 * pi * rho[i + (size_t)nx * (j + (size_t)ny * k)];
      }
  free(rho);
}

/* ---------------------------------------------------------------------------
 * Moments.  mom(7, nxgs-1:nxge+1, nys-1:nye+1, nzs-1:nze+1, nsp).
 * ------------------------------------------------------------------------- */
static inline size_t iMOM(const or3d_params *p, int l, int i, int j, int k, int isp) {
  return (size_t)(l - 1) +
         7u * ((size_t)(i - (p->nxgs - 1)) +
               (size_t)(NX + 2) * ((size_t)(j - (p->nys - 1)) +
                                   (size_t)(NY + 2) * ((size_t)(k - (p->nzs - 1)) + (size_t)(NZ + 2) * (size_t)(isp - 1))));
}

/* mom_calc__accl: 3d/common/mom_calc.f90:49-216, delt = delt_in*5d-1
 * (mom_calc.f90:37); gp(1:3) = up(1:3), gp(4:6) kicked, gp(7) not written */
void or3d_mom_accl(const or3d_params *p, double *gp, const double *up, const double *uf, const int *cumcnt) {
  const int nxs = p->nxgs, nxe = p->nxge, nys = p->nys, nye = p->nye, nzs = p->nzs, nze = p->nze;
  const double d_delx = 1e0 / p->delx, delt = p->delt * 5e-1, c = p->c;
  const int tx = NX + 2, ty = NY + 2;
  double *tmpf = (double *)malloc(sizeof(double) * 6u * (size_t)tx * ty * (NZ + 2));
#define TF(cc, i, j, k) tmpf[(size_t)((cc)-1) + 6u * ((size_t)((i) - (nxs - 1)) + (size_t)tx * ((size_t)((j) - (nys - 1)) + (size_t)ty * (size_t)((k) - (nzs - 1))))]
  for (int k = nzs - 1; k <= nze + 1; k++)
    for (int j = nys - 1; j <= nye + 1; j++)
      for (int i = nxs - 1; i <= nxe + 1; i++) {
        TF(1, i, j, k) = 2.5e-1 * (+UF(1, i, j, k) + UF(1, i, j + 1, k) + UF(1, i, j, k + 1) + UF(1, i, j + 1, k + 1));
        TF(2, i, j, k) = 2.5e-1 * (+UF(2, i, j, k) + UF(2, i + 1, j, k) + UF(2, i, j, k + 1) + UF(2, i + 1, j, k + 1));
        TF(3, i, j, k) = 2.5e-1 * (+UF(3, i, j, k) + UF(3, i + 1, j, k) + UF(3, i, j + 1, k) + UF(3, i + 1, j + 1, k));
        TF(4, i, j, k) = 5e-1 * (+UF(4, i, j, k) + UF(4, i + 1, j, k));
        TF(5, i, j, k) = 5e-1 * (+UF(5, i, j, k) + UF(5, i, j + 1, k));
        TF(6, i, j, k) = 5e-1 * (+UF(6, i, j, k) + UF(6, i, j, k + 1));
      }
  for (int k = nzs; k <= nze; k++)
    for (int j = nys; j <= nye; j++)
      for (int i = nxs; i <= nxe; i++)
        for (int isp = 1; isp <= p->nsp; isp++) {
          const double q = p->q[isp - 1], r = p->r[isp - 1];
          const double fac1 = q / r * 5e-1 * delt;
          const double txxx = fac1 * fac1;
          const double fac2 = q * delt / r;
          for (int ii = cumcnt[iCC(p, i, j, k, isp)] + 1; ii <= cumcnt[iCC(p, i + 1, j, k, isp)]; ii++) {
            const double *u = &up[iUP(p, 1, ii, j, k, isp)];
            double *g = &gp[iUP(p, 1, ii, j, k, isp)];
            double shxm, shx, shxp, shym, shy, shyp, shzm, shz, shzp;
            shape3(u[0] * d_delx - 5e-1 - i, &shxm, &shx, &shxp);
            shape3(u[1] * d_delx - 5e-1 - j, &shym, &shy, &shyp);
            shape3(u[2] * d_delx - 5e-1 - k, &shzm, &shz, &shzp);
#define RX(cc, jj, kk) (+TF(cc, i - 1, jj, kk) * shxm + TF(cc, i, jj, kk) * shx + TF(cc, i + 1, jj, kk) * shxp)
#define RY(cc, kk) (+RX(cc, j - 1, kk) * shym + RX(cc, j, kk) * shy + RX(cc, j + 1, kk) * shyp)
#define G3(cc) (RY(cc, k - 1) * shzm + RY(cc, k) * shz + RY(cc, k + 1) * shzp)
            const double bpx = G3(1), bpy = G3(2), bpz = G3(3), epx = G3(4), epy = G3(5), epz = G3(6);
#undef G3
#undef RY
#undef RX
            double uvm1 = u[3] + fac1 * epx;
            double uvm2 = u[4] + fac1 * epy;
            double uvm3 = u[5] + fac1 * epz;
            const double gam = sqrt(c * c + uvm1 * uvm1 + uvm2 * uvm2 + uvm3 * uvm3);
            const double igam = 1e0 / gam;
            const double fac1r = fac1 * igam;
            const double fac2r = fac2 / (gam + txxx * (bpx * bpx + bpy * bpy + bpz * bpz) * igam);
            const double uvm4 = uvm1 + fac1r * (+uvm2 * bpz - uvm3 * bpy);
            const double uvm5 = uvm2 + fac1r * (+uvm3 * bpx - uvm1 * bpz);
            const double uvm6 = uvm3 + fac1r * (+uvm1 * bpy - uvm2 * bpx);
            uvm1 = uvm1 + fac2r * (+uvm5 * bpz - uvm6 * bpy);
            uvm2 = uvm2 + fac2r * (+uvm6 * bpx - uvm4 * bpz);
            uvm3 = uvm3 + fac2r * (+uvm4 * bpy - uvm5 * bpx);
            g[0] = u[0];
            g[1] = u[1];
            g[2] = u[2];
            g[3] = uvm1 + fac1 * epx;
            g[4] = uvm2 + fac1 * epy;
            g[5] = uvm3 + fac1 * epz;
          }
        }
#undef TF
  free(tmpf);
}

/* mom_calc__nvt: 3d/common/mom_calc.f90:219-333 (serial in the reference:
 * isp, k, j, ii).  floor() cell of the lower node, and the offset is taken
 * from the position itself, up(1)-5d-1-ih, without d_delx (:243-251). */
void or3d_mom_nvt(const or3d_params *p, double *mom, const double *up, const int *np2) {
  const double d_delx = 1e0 / p->delx, c = p->c;
  memset(mom, 0, sizeof(double) * 7u * (size_t)(NX + 2) * (NY + 2) * (NZ + 2) * p->nsp);
  for (int isp = 1; isp <= p->nsp; isp++)
    for (int k = p->nzs; k <= p->nze; k++)
      for (int j = p->nys; j <= p->nye; j++)
        for (int ii = 1; ii <= np2[iNP(p, j, k, isp)]; ii++) {
          const double *u = &up[iUP(p, 1, ii, j, k, isp)];
          const int ih = (int)floor(u[0] * d_delx - 5e-1);
          const int jh = (int)floor(u[1] * d_delx - 5e-1);
          const int kh = (int)floor(u[2] * d_delx - 5e-1);
          const double dx = u[0] - 5e-1 - ih, dxm = 1e0 - dx;
          const double dy = u[1] - 5e-1 - jh, dym = 1e0 - dy;
          const double dz = u[2] - 5e-1 - kh, dzm = 1e0 - dz;
          const double gam = 1e0 / sqrt(1e0 + (+u[3] * u[3] + u[4] * u[4] + u[5] * u[5]) / (c * c));
          for (int n = 0; n < 8; n++) {
            const int ox = n & 1, oy = (n >> 1) & 1, oz = n >> 2;
            const double wx = ox ? dx : dxm, wy = oy ? dy : dym, wz = oz ? dz : dzm;
            double *m = &mom[iMOM(p, 1, ih + ox, jh + oy, kh + oz, isp)];
            m[0] = m[0] + wx * wy * wz;
            for (int d = 0; d < 3; d++) m[1 + d] = m[1 + d] + u[3 + d] * gam * wx * wy * wz;
            for (int d = 0; d < 3; d++) m[4 + d] = m[4 + d] + u[3 + d] * u[3 + d] * gam * wx * wy * wz;
          }
        }
}

/* boundary_periodic__mom: 3d/common/boundary_periodic.f90:1102-1235 at one
 * rank.  x fold over everything, then per species the y fold (nys-1 onto
 * nye, then nye+1 onto nys) over all x nodes and z planes, then the z fold
 * (nzs-1 onto nze, then nze+1 onto nzs) over all x nodes and interior y. */
void or3d_bc_mom(const or3d_params *p, double *mom) {
  const int xs = p->nxgs, xe = p->nxge, ys = p->nys, ye = p->nye, zs = p->nzs, ze = p->nze;
#define M(l, i, j, k, s) mom[iMOM(p, l, i, j, k, s)]
  for (int s = 1; s <= p->nsp; s++)
    for (int k = zs - 1; k <= ze + 1; k++)
      for (int j = ys - 1; j <= ye + 1; j++)
        for (int l = 1; l <= 7; l++) {
          M(l, xs, j, k, s) = M(l, xs, j, k, s) + M(l, xe + 1, j, k, s);
          M(l, xe, j, k, s) = M(l, xe, j, k, s) + M(l, xs - 1, j, k, s);
        }
  for (int s = 1; s <= p->nsp; s++) {
    for (int k = zs - 1; k <= ze + 1; k++)
      for (int i = xs - 1; i <= xe + 1; i++)
        for (int l = 1; l <= 7; l++) M(l, i, ye, k, s) = M(l, i, ye, k, s) + M(l, i, ys - 1, k, s);
    for (int k = zs - 1; k <= ze + 1; k++)
      for (int i = xs - 1; i <= xe + 1; i++)
        for (int l = 1; l <= 7; l++) M(l, i, ys, k, s) = M(l, i, ys, k, s) + M(l, i, ye + 1, k, s);
    for (int j = ys; j <= ye; j++)
      for (int i = xs - 1; i <= xe + 1; i++)
        for (int l = 1; l <= 7; l++) M(l, i, j, ze, s) = M(l, i, j, ze, s) + M(l, i, j, zs - 1, s);
    for (int j = ys; j <= ye; j++)
      for (int i = xs - 1; i <= xe + 1; i++)
        for (int l = 1; l <= 7; l++) M(l, i, j, zs, s) = M(l, i, j, zs, s) + M(l, i, j, ze + 1, s);
  }
#undef M
}
