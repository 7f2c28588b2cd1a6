/*
 * kref_fft.c -- oracle FFT: 3D real transforms with the semantics of
 * A2a_k / A_k2a (fft_wrapper.h:641-681).
 *
 * TEST INFRASTRUCTURE ONLY (see kref.h).
 *
 * Definition being restated (the _FFT_FFTW branch, fft_wrapper.h:645 and
 * 666-670, which the Ooura branch reproduces by conjugation, 650-652 and
 * 672-679):
 *   forward  A(k) = sum_x a(x) exp(-2 pi i k.x / N)          (unnormalised)
 *   inverse  a(x) = 1/(NX NY NZ) sum_k A(k) exp(+2 pi i k.x / N)
 * on the in-place padded layout a[NX][NY][NZ_], NZ_ = 2(NZ/2+1), with the
 * spectral value of (kx, ky, kz) stored as (re, im) at [i][j][2kz].
 *
 * The algorithm here is this oracle's own: a mixed-radix Stockham autosort
 * complex FFT per line (radices 4, 2, 3, then odd primes), two real rows
 * packed into one complex row for the z axis. It does not follow Ooura's
 * split-radix; agreement with Ooura's rdft3d is checked by the tests against
 * the reference's fftsg3d.c built under oracle/_ref.
 */
#include <complex.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#include "kref_internal.h"

static void twiddle_table(cplx *tw, int n) {
    for (int k = 0; k < n; k++) {
        long double ang = 2.0L * 3.14159265358979323846264338327950288L * (long double)k / (long double)n;
        double      c   = (double)cosl(ang);
        double      s   = (double)sinl(ang);
        if ((4 * k) % n == 0) { /* exact quarter turns */
            int q = (4 * k) / n;
            c     = (q == 0) ? 1.0 : (q == 2) ? -1.0 : 0.0;
            s     = (q == 1) ? 1.0 : (q == 3) ? -1.0 : 0.0;
        }
        tw[k] = c - I * s; /* exp(-2 pi i k / n) */
    }
}

int kr_plan1d_init(kr_plan1d *p, int n) {
    memset(p, 0, sizeof(*p));
    if (n < 1) return -1;
    p->n  = n;
    int m = n;
    while (m % 4 == 0) {
        p->fac[p->nf++] = 4;
        m /= 4;
    }
    while (m % 2 == 0) {
        p->fac[p->nf++] = 2;
        m /= 2;
    }
    for (int f = 3; m > 1; f += 2) {
        while (m % f == 0) {
            if (p->nf >= KR_MAXFAC || f > KR_MAXRADIX) return -1;
            p->fac[p->nf++] = f;
            m /= f;
        }
    }
    p->tw = (cplx *)malloc(sizeof(cplx) * (size_t)n);
    if (!p->tw) return -1;
    twiddle_table(p->tw, n);
    return 0;
}

void kr_plan1d_free(kr_plan1d *p) {
    free(p->tw);
    p->tw = NULL;
}

/* x: n values, w: n scratch; sign -1 forward, +1 inverse (no scaling). */
void kr_fft1d(const kr_plan1d *p, cplx *x, cplx *w, int sign) {
    const int n  = p->n;
    cplx     *in = x, *out = w;
    int       Ns = 1;
    for (int s = 0; s < p->nf; s++) {
        const int R       = p->fac[s];
        const int m       = n / R;
        const int tstride = n / (Ns * R);
        for (int j = 0; j < m; j++) {
            const int k = j % Ns;
            cplx      v[KR_MAXRADIX];
            for (int r = 0; r < R; r++) {
                cplx t = p->tw[r * k * tstride];
                if (sign > 0) t = conj(t);
                v[r] = in[j + r * m] * t;
            }
            const int base = (j / Ns) * Ns * R + k;
            if (R == 2) {
                out[base]      = v[0] + v[1];
                out[base + Ns] = v[0] - v[1];
            } else if (R == 4) {
                cplx a = v[0] + v[2], b = v[0] - v[2];
                cplx c = v[1] + v[3], d = v[1] - v[3];
                /* multiply d by -i (forward) or +i (inverse) */
                cplx dd            = (sign < 0) ? (cimag(d) - I * creal(d)) : (-cimag(d) + I * creal(d));
                out[base]          = a + c;
                out[base + Ns]     = b + dd;
                out[base + 2 * Ns] = a - c;
                out[base + 3 * Ns] = b - dd;
            } else {
                const int rstride = n / R;
                for (int q = 0; q < R; q++) {
                    cplx acc = 0.0;
                    for (int r = 0; r < R; r++) {
                        cplx t = p->tw[((r * q) % R) * rstride];
                        if (sign > 0) t = conj(t);
                        acc += v[r] * t;
                    }
                    out[base + q * Ns] = acc;
                }
            }
        }
        cplx *tmp = in;
        in        = out;
        out       = tmp;
        Ns *= R;
    }
    if (in != x) memcpy(x, in, sizeof(cplx) * (size_t)n);
}

int kr_fft3_init(kr_fft3 *f, int nx, int ny, int nz) {
    memset(f, 0, sizeof(*f));
    f->nx = nx;
    f->ny = ny;
    f->nz = nz;
    if (nz % 2) return -1;
    if (kr_plan1d_init(&f->px, nx) || kr_plan1d_init(&f->py, ny) || kr_plan1d_init(&f->pz, nz)) return -1;
    return 0;
}

void kr_fft3_free(kr_fft3 *f) {
    kr_plan1d_free(&f->px);
    kr_plan1d_free(&f->py);
    kr_plan1d_free(&f->pz);
}

/* Forward 3D r2c, in place on the padded layout. */
void kr_fft3_forward(const kr_fft3 *f, double *a) {
    const int nx = f->nx, ny = f->ny, nz = f->nz;
    const int hnz = nz / 2 + 1, nz_ = 2 * hnz;
    const int nrow = nx * ny;
    int       nmax = nx > ny ? nx : ny;
    if (nz > nmax) nmax = nz;

#pragma omp parallel
    {
        cplx *z = (cplx *)malloc(sizeof(cplx) * (size_t)nmax);
        cplx *w = (cplx *)malloc(sizeof(cplx) * (size_t)nmax);
        /* z axis: pack rows (2p, 2p+1) into one complex row */
#pragma omp for schedule(static)
        for (int pr = 0; pr < (nrow + 1) / 2; pr++) {
            int     r0 = 2 * pr, r1 = 2 * pr + 1;
            double *A  = a + (size_t)r0 * nz_;
            double *B  = (r1 < nrow) ? a + (size_t)r1 * nz_ : NULL;
            for (int k = 0; k < nz; k++) z[k] = A[k] + I * (B ? B[k] : 0.0);
            kr_fft1d(&f->pz, z, w, -1);
            for (int k = 0; k < hnz; k++) {
                cplx zk  = z[k];
                cplx zmk = conj(z[(nz - k) % nz]);
                cplx ak  = 0.5 * (zk + zmk);
                cplx bk  = -0.5 * I * (zk - zmk);
                w[k]     = bk;
                A[2 * k]     = creal(ak);
                A[2 * k + 1] = cimag(ak);
            }
            if (B)
                for (int k = 0; k < hnz; k++) {
                    B[2 * k]     = creal(w[k]);
                    B[2 * k + 1] = cimag(w[k]);
                }
        }
        /* y axis */
#pragma omp for schedule(static)
        for (int i = 0; i < nx; i++) {
            for (int kz = 0; kz < hnz; kz++) {
                for (int j = 0; j < ny; j++) {
                    double *c = a + ((size_t)i * ny + j) * nz_ + 2 * kz;
                    z[j]      = c[0] + I * c[1];
                }
                kr_fft1d(&f->py, z, w, -1);
                for (int j = 0; j < ny; j++) {
                    double *c = a + ((size_t)i * ny + j) * nz_ + 2 * kz;
                    c[0]      = creal(z[j]);
                    c[1]      = cimag(z[j]);
                }
            }
        }
        /* x axis */
#pragma omp for schedule(static)
        for (int j = 0; j < ny; j++) {
            for (int kz = 0; kz < hnz; kz++) {
                for (int i = 0; i < nx; i++) {
                    double *c = a + ((size_t)i * ny + j) * nz_ + 2 * kz;
                    z[i]      = c[0] + I * c[1];
                }
                kr_fft1d(&f->px, z, w, -1);
                for (int i = 0; i < nx; i++) {
                    double *c = a + ((size_t)i * ny + j) * nz_ + 2 * kz;
                    c[0]      = creal(z[i]);
                    c[1]      = cimag(z[i]);
                }
            }
        }
        free(z);
        free(w);
    }
}

/* Inverse 3D c2r, in place on the padded layout, normalised. */
void kr_fft3_inverse(const kr_fft3 *f, double *a) {
    const int    nx = f->nx, ny = f->ny, nz = f->nz;
    const int    hnz = nz / 2 + 1, nz_ = 2 * hnz;
    const int    nrow  = nx * ny;
    const double scale = 1.0 / ((double)nx * (double)ny * (double)nz);
    int          nmax  = nx > ny ? nx : ny;
    if (nz > nmax) nmax = nz;

#pragma omp parallel
    {
        cplx *z = (cplx *)malloc(sizeof(cplx) * (size_t)nmax);
        cplx *w = (cplx *)malloc(sizeof(cplx) * (size_t)nmax);
#pragma omp for schedule(static)
        for (int j = 0; j < ny; j++) {
            for (int kz = 0; kz < hnz; kz++) {
                for (int i = 0; i < nx; i++) {
                    double *c = a + ((size_t)i * ny + j) * nz_ + 2 * kz;
                    z[i]      = c[0] + I * c[1];
                }
                kr_fft1d(&f->px, z, w, +1);
                for (int i = 0; i < nx; i++) {
                    double *c = a + ((size_t)i * ny + j) * nz_ + 2 * kz;
                    c[0]      = creal(z[i]);
                    c[1]      = cimag(z[i]);
                }
            }
        }
#pragma omp for schedule(static)
        for (int i = 0; i < nx; i++) {
            for (int kz = 0; kz < hnz; kz++) {
                for (int j = 0; j < ny; j++) {
                    double *c = a + ((size_t)i * ny + j) * nz_ + 2 * kz;
                    z[j]      = c[0] + I * c[1];
                }
                kr_fft1d(&f->py, z, w, +1);
                for (int j = 0; j < ny; j++) {
                    double *c = a + ((size_t)i * ny + j) * nz_ + 2 * kz;
                    c[0]      = creal(z[j]);
                    c[1]      = cimag(z[j]);
                }
            }
        }
        /* z axis: two Hermitian rows -> one complex row -> two real rows.
         * The imaginary parts of kz=0 and kz=NZ/2 are dropped (c2r). */
#pragma omp for schedule(static)
        for (int pr = 0; pr < (nrow + 1) / 2; pr++) {
            int     r0 = 2 * pr, r1 = 2 * pr + 1;
            double *A  = a + (size_t)r0 * nz_;
            double *B  = (r1 < nrow) ? a + (size_t)r1 * nz_ : NULL;
            for (int k = 0; k < hnz; k++) {
                cplx ak = A[2 * k] + I * A[2 * k + 1];
                cplx bk = B ? (B[2 * k] + I * B[2 * k + 1]) : 0.0;
                if (k == 0 || k == nz / 2) {
                    ak = creal(ak);
                    bk = creal(bk);
                }
                z[k] = ak + I * bk;
                if (k > 0 && k < nz / 2) z[nz - k] = conj(ak) + I * conj(bk);
            }
            kr_fft1d(&f->pz, z, w, +1);
            for (int k = 0; k < nz; k++) {
                A[k] = creal(z[k]) * scale;
                if (B) B[k] = cimag(z[k]) * scale;
            }
        }
        free(z);
        free(w);
    }
}

int kr_fft_forward(int nx, int ny, int nz, double *a) {
    kr_fft3 f;
    if (kr_fft3_init(&f, nx, ny, nz)) {
        kr_fft3_free(&f);
        return -1;
    }
    kr_fft3_forward(&f, a);
    kr_fft3_free(&f);
    return 0;
}

int kr_fft_inverse(int nx, int ny, int nz, double *a) {
    kr_fft3 f;
    if (kr_fft3_init(&f, nx, ny, nz)) {
        kr_fft3_free(&f);
        return -1;
    }
    kr_fft3_inverse(&f, a);
    kr_fft3_free(&f);
    return 0;
}
