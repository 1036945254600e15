/*
 * rotwave.c -- FP64 CPU restatement of the space-group rotation of
 * eigenfunctions used by readcphi / readgeig (S/readeigen.f90:71-181): rotmto
 * for the muffin-tin coefficients (S/rotwave.f90:156-192) with the orbital
 * tables of readmnla_cphi (S/readeigen.f90:487-521), and rotipw for the plane
 * waves (S/rotwave.f90:223-245).
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP library, never part of
 * the product path.  S/ = SRC/subroutines/ of ecalj.
 */
#include <math.h>
#include <stdlib.h>

#include "oracle.h"

#define PI 3.14159265358979323846

/* rotmto: cphiout(offset_rev(miat(ibas),l,k) + m, :) =
 *   sum_m' dlmm(m, m', l) cphiin(offset(iorb) + m', :) * exp(-2 pi i (sym qin).tiat(:,ibas)).
 * Tables 1-based as read from @MNLA_CPHI (offset = first row - 1).
 * dlmm(-lmxax:lmxax, -lmxax:lmxax, 0:lmxax) of this operation; miat(nbas)
 * 1-based; tiat(3, nbas).  cphi: ndima x nband (leading dimension ndima).
 * Returns 0, or -1 when the tables do not map a block. */
int orc_rotmto(const double* qin, int ndima, int nband, int norbt, const int* ibas_tbl, const int* l_tbl,
               const int* k_tbl, const int* offset_tbl, const double* sym, const double* dlmm, int lmxax,
               const int* miat, const double* tiat, int nbas, const double complex* cin, double complex* cout) {
  int mb = 0, ml = 0, mk = 0;
  for (int i = 0; i < norbt; ++i) {
    if (ibas_tbl[i] > mb) mb = ibas_tbl[i];
    if (l_tbl[i] > ml) ml = l_tbl[i];
    if (k_tbl[i] > mk) mk = k_tbl[i];
  }
  /* offset_rev_tbl(ibas, 0:l, k) (S/readeigen.f90:515-519) */
  const int nrev = mb * (ml + 1) * mk;
  int* rev = (int*)malloc(sizeof(int) * nrev);
  for (int i = 0; i < nrev; ++i) rev[i] = -99999;
  for (int i = 0; i < norbt; ++i)
    rev[(ibas_tbl[i] - 1) + mb * (l_tbl[i] + (ml + 1) * (k_tbl[i] - 1))] = offset_tbl[i];
  double qrot[3];
  for (int i = 0; i < 3; ++i) qrot[i] = sym[i] * qin[0] + sym[i + 3] * qin[1] + sym[i + 6] * qin[2];
  double complex* ph = (double complex*)malloc(sizeof(double complex) * nbas);
  for (int b = 0; b < nbas; ++b) {
    const double t = qrot[0] * tiat[3 * b] + qrot[1] * tiat[3 * b + 1] + qrot[2] * tiat[3 * b + 2];
    ph[b] = cexp(-2.0 * PI * I * t);
  }
  const int ld = 2 * lmxax + 1;
  int rc = 0;
  for (int io = 0; io < norbt; ++io) {
    const int ibas = ibas_tbl[io], l = l_tbl[io], k = k_tbl[io];
    const int jb = miat[ibas - 1];
    if (jb < 1 || jb > mb) { rc = -1; break; }
    const int ini1 = offset_tbl[io];
    const int ini2 = rev[(jb - 1) + mb * (l + (ml + 1) * (k - 1))];
    if (ini2 < 0 || ini2 + 2 * l + 1 > ndima) { rc = -1; break; }
    const double* d = dlmm + (size_t)ld * ld * l;
    for (int b = 0; b < nband; ++b)
      for (int m = -l; m <= l; ++m) {
        double complex s = 0.0;
        for (int mp = -l; mp <= l; ++mp)
          s += d[(m + lmxax) + ld * (mp + lmxax)] * cin[(ini1 + mp + l) + (long)ndima * b];
        cout[(ini2 + m + l) + (long)ndima * b] = s * ph[ibas - 1];
      }
  }
  free(rev);
  free(ph);
  return rc;
}

/* rotipw: geigout(ngvecprev(nint(platt (sym (qin + qlat G) - qtarget))), :) =
 *   geigin(ig, :) * exp(-2 pi i (sym (qin + qlat G)).shtvg), ig = 1..ngp.
 * ngvecp(3, ngp); ngvecprev(-imx:imx)^3 1-based; geig: ngp x nband
 * (leading dimension ldg).  Returns 0 or -1 (a rotated G outside the cube or
 * without an index). */
int orc_rotipw(const double* qin, const double* qtarget, int ngp, int nband, const double* platt, const double* qlat,
               const double* sym, const int* ngvecp, const int* ngvecprev, int imx, const double* shtvg,
               const double complex* gin, double complex* gout, long ldg) {
  const int nc = 2 * imx + 1;
  for (int ig = 0; ig < ngp; ++ig) {
    double qpg[3], qpgr[3], d[3];
    for (int i = 0; i < 3; ++i)
      qpg[i] = qin[i] + qlat[i] * ngvecp[3 * ig] + qlat[i + 3] * ngvecp[3 * ig + 1] + qlat[i + 6] * ngvecp[3 * ig + 2];
    for (int i = 0; i < 3; ++i) qpgr[i] = sym[i] * qpg[0] + sym[i + 3] * qpg[1] + sym[i + 6] * qpg[2];
    for (int i = 0; i < 3; ++i) d[i] = qpgr[i] - qtarget[i];
    int n[3];
    for (int i = 0; i < 3; ++i) {
      n[i] = (int)lround(platt[i] * d[0] + platt[i + 3] * d[1] + platt[i + 6] * d[2]);
      if (n[i] < -imx || n[i] > imx) return -1;
    }
    const int ig2 = ngvecprev[(n[0] + imx) + nc * ((n[1] + imx) + nc * (n[2] + imx))];
    if (ig2 < 1 || ig2 > ngp) return -1;
    const double t = qpgr[0] * shtvg[0] + qpgr[1] * shtvg[1] + qpgr[2] * shtvg[2];
    const double complex ph = cexp(-2.0 * PI * I * t);
    for (int b = 0; b < nband; ++b) gout[(ig2 - 1) + ldg * b] = gin[ig + ldg * b] * ph;
  }
  return 0;
}
