/*
 * c3d_oracle_int.h -- helpers shared by the 2nd- and 4th-order oracle
 * restatements (TEST INFRASTRUCTURE ONLY, not part of the product library).
 */
#ifndef C3D_ORACLE_INT_H
#define C3D_ORACLE_INT_H

#include <stddef.h>

/* unnormalised mixed-radix DFT plan with twiddles rounded from long double */
typedef struct fftplan {
  int n, nf, fac[40];
  double *twr, *twi; /* exp(-2 pi i k / n), k=0..n-1 */
} fftplan;

void orci_fft_plan_init(fftplan* p, int n);
void orci_fft_plan_free(fftplan* p);
/* FFTW R2HC / HC2R, unnormalised; wk holds 4n doubles */
void orci_r2hc(const fftplan* p, const double* in, double* out, double* wk);
void orci_hc2r(const fftplan* p, const double* in, double* out, double* wk);

/* order-independent sum: terms truncated to multiples of 2^-80, four
 * 32-bit limbs accumulated in int64 */
void orci_xsum_add(long long* acc, double t);
double orci_xsum_value(const long long* acc);

/* InverseTridiagonal (m_Tools.f90:218-245) on one line with stride fs:
 * a, b, cc hold rows 1..n at [0..n-1]; arr needs n+1 entries */
void orci_thomas_line(const double* a, const double* b, const double* cc, double* f, double* arr,
                      int n, size_t fs);

/* OpenMP thread count chosen with orc_set_threads */
int orc_get_threads(void);

#endif
