/*
 * kref_internal.h -- oracle internals (TEST INFRASTRUCTURE ONLY, see kref.h).
 */
#ifndef KREF_INTERNAL_H
#define KREF_INTERNAL_H

#include <complex.h>

#include "kref.h"

typedef double _Complex cplx;

#define KR_MAXFAC 40
#define KR_MAXRADIX 64

typedef struct {
    int   n, nf, fac[KR_MAXFAC];
    cplx *tw; /* exp(-2 pi i k/n), k < n */
} kr_plan1d;

typedef struct {
    int       nx, ny, nz;
    kr_plan1d px, py, pz;
} kr_fft3;

int  kr_plan1d_init(kr_plan1d *p, int n);
void kr_plan1d_free(kr_plan1d *p);
void kr_fft1d(const kr_plan1d *p, cplx *x, cplx *w, int sign);
int  kr_fft3_init(kr_fft3 *f, int nx, int ny, int nz);
void kr_fft3_free(kr_fft3 *f);
void kr_fft3_forward(const kr_fft3 *f, double *a);
void kr_fft3_inverse(const kr_fft3 *f, double *a);

#endif
