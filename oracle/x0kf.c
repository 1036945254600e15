/*
 * x0kf.c -- FP64 CPU restatement of x0kf_v4hz_init (S/x0kf_v4h.f90:48-142).
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the host-side index lists the
 * library's callers build, never part of the product path.
 */
#include "oracle.h"
#include <math.h>

#define F3(a, i, k, j) (a)[((i) - 1) + (long)nbnbx * (((k) - 1) + (long)nqbz * ((j) - 1))]

/* job 0 (:96-104): per k the band ranges over all pairs of both signs;
 * nkmax / nkqmax only over valence bands (n <= nband).  job 1 (:105-138):
 * pairs with n1b > nkmax or n2b > nkqmax are skipped (so every core pair,
 * n1b > nband, is skipped), it = nctot + n1b (valence) or n1b - nband
 * (core), itp = ncc + n2b - nkqmin + 1 (valence) or n2b - nkqmin + 1 - nband
 * (core) with ncc = 0 for npm = 1 and nctot for npm = 2, and the weight of
 * bin iw is whw(jhw + iw - ihw) of that pair.  Returns 0, 1 when a pair's
 * bins run past nwhis (the reference stops there), 2 when npm = 2 and
 * nkqmin /= 1 (idem), 3 when the output capacity is too small.          */
int orc_x0kf_init(int nqbz, int npm, int nbnbx, const int* nbnb, const int* n1b, const int* n2b,
                  const int* ihw, const int* nhw, const int* jhw, const double* whw, int nband, int nctot,
                  int nwhis, long cap_coun, long cap_count, int* kc, int* itc, int* itpc, int* jpmc,
                  int* iwini, int* iwend, int* icouini, double* whwc, int* icounkmin, int* icounkmax,
                  int* nkmin, int* nkmax, int* nkqmin, int* nkqmax, long* ncoun_out, long* ncount_out) {
  const int ncc = (npm == 1) ? 0 : nctot;
  long icoun = 0, icount = 0;
  for (int k = 1; k <= nqbz; ++k) {
    nkmin[k - 1] = 999999;
    nkqmin[k - 1] = 999999;
    nkmax[k - 1] = -999999;
    nkqmax[k - 1] = -999999;
    for (int jpm = 1; jpm <= npm; ++jpm)
      for (int ib = 1; ib <= nbnb[(k - 1) + (long)nqbz * (jpm - 1)]; ++ib) {
        const int a = F3(n1b, ib, k, jpm), b = F3(n2b, ib, k, jpm);
        if (a < nkmin[k - 1]) nkmin[k - 1] = a;
        if (b < nkqmin[k - 1]) nkqmin[k - 1] = b;
        if (a <= nband && a > nkmax[k - 1]) nkmax[k - 1] = a;
        if (b <= nband && b > nkqmax[k - 1]) nkqmax[k - 1] = b;
      }
  }
  for (int k = 1; k <= nqbz; ++k) {
    icounkmin[k - 1] = (int)(icoun + 1);
    if (npm == 2 && nkqmin[k - 1] != 1) return 2;
    for (int jpm = 1; jpm <= npm; ++jpm)
      for (int ib = 1; ib <= nbnb[(k - 1) + (long)nqbz * (jpm - 1)]; ++ib) {
        const int a = F3(n1b, ib, k, jpm), b = F3(n2b, ib, k, jpm);
        const int i0 = F3(ihw, ib, k, jpm), nh = F3(nhw, ib, k, jpm), j0 = F3(jhw, ib, k, jpm);
        if (i0 + nh - 1 > nwhis) return 1;
        if (a > nkmax[k - 1]) continue;
        if (b > nkqmax[k - 1]) continue;
        const int it = (a <= nband) ? nctot + a : a - nband;
        const int itp = (b <= nband) ? ncc + b - nkqmin[k - 1] + 1 : b - nkqmin[k - 1] + 1 - nband;
        if (icoun >= cap_coun) return 3;
        kc[icoun] = k;
        itc[icoun] = it;
        itpc[icoun] = itp;
        jpmc[icoun] = jpm;
        iwini[icoun] = i0;
        iwend[icoun] = i0 + nh - 1;
        icouini[icoun] = (int)(icount + 1);
        ++icoun;
        for (int iw = i0; iw <= i0 + nh - 1; ++iw) {
          if (icount >= cap_count) return 3;
          whwc[icount++] = whw[j0 + iw - i0 - 1];
        }
      }
    icounkmax[k - 1] = (int)icoun;
  }
  *ncoun_out = icoun;
  *ncount_out = icount;
  return 0;
}

/* getwemax (S/m_freq.f90:197-268) and the real mesh of getfreq
 * (S/m_freq.f90:110-129).  eq: the external states' energies over every
 * q of the run (Ry); ebz: all band energies over the BZ mesh (Ry).  wemax =
 * the largest of |emaxv - ef|, |omegav - ef|, |omegac - ef|, |eminc - ef|,
 * plus ffac*esmr, where omegav / omegac start at 0 and take eq -+ 2 deltaw
 * below / above ef, emaxv / eminc start at 0 and take the band energies
 * within ffac*esmr of ef.  nw = the first iw in 3..nwhis with
 * (frhis(iw-2) + frhis(iw-1))/2 > wemax/2 (frhis in Hartree), then
 * min(nw, nwhis-1); freq_r(0) = 0, freq_r(i) = (frhis(i) + frhis(i+1))/2.
 * Returns nw, or -1 when no bin qualifies (the reference leaves nw unset). */
int orc_real_mesh(int neq, const double* eq, long nbz, const double* ebz, double ef, double deltaw, double esmr,
                  double ffac, int nwhis, const double* frhis, double* wemax_out, double* freq_r, int cap) {
  double omegav = 0.0, omegac = 0.0;
  for (int i = 0; i < neq; ++i) {
    double e = eq[i] - 2.0 * deltaw;
    if (e >= 1e20 - 1e10) continue;
    if (e < ef && e < omegav) omegav = e;
    e = eq[i] + 2.0 * deltaw;
    if (e > ef && e > omegac) omegac = e;
  }
  double emaxv = 0.0, eminc = 0.0;
  for (long i = 0; i < nbz; ++i) {
    const double e = ebz[i];
    if (e < ef + ffac * esmr && e > emaxv) emaxv = e;
    if (e > ef - ffac * esmr && e < eminc) eminc = e;
  }
  double we = fabs(emaxv - ef);
  if (fabs(omegav - ef) > we) we = fabs(omegav - ef);
  if (fabs(omegac - ef) > we) we = fabs(omegac - ef);
  if (fabs(eminc - ef) > we) we = fabs(eminc - ef);
  const double wemax = we + ffac * esmr;
  *wemax_out = wemax;
  int nw = -1;
  for (int iw = 3; iw <= nwhis; ++iw)
    if (0.5 * (frhis[iw - 3] + frhis[iw - 2]) > wemax / 2.0) {
      nw = iw;
      break;
    }
  if (nw < 0) return -1;
  if (nw > nwhis - 1) nw = nwhis - 1;
  if (nw + 1 > cap) return -1;
  freq_r[0] = 0.0;
  for (int i = 1; i <= nw; ++i) freq_r[i] = 0.5 * (frhis[i - 1] + frhis[i]);
  return nw;
}
