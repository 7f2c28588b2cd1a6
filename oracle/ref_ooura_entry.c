/*
 * ref_ooura_entry.c -- C entry points to the reference's own Ooura FFT (fftsg3d.c,
 * fftsg.c) with exactly the call sequence of A2a_k / A_k2a's default branch
 * (fft_wrapper.h:641-681; workspace sizes from Init_fft_ooura,
 * fft_wrapper.cxx:78-87), so the oracle's FFT can be checked against it.
 *
 * TEST INFRASTRUCTURE ONLY. Built by oracle/Makefile into oracle/_ref/ from
 * the sources where they lie in the reference tree; never shipped as product.
 */
#include <math.h>
#include <stdlib.h>

void rdft3d(int n1, int n2, int n3, int isgn, double ***a, double *t, int *ip, double *w);
void rdft3dsort(int n1, int n2, int n3, int isgn, double ***a);
void init_genrand(unsigned long s);
double genrand_real3(void);
unsigned long genrand_int32(void);

typedef struct {
    int     *ip;
    double  *t, *w;
    double ***view;
} ooura_ws;

static int ws_init(ooura_ws *ws, int nx, int ny, int nz, double *a) {
    int nt = nx > ny ? nx : ny;
    int n  = nt > nz / 2 ? nt : nz / 2;
    int nz_ = 2 * (nz / 2 + 1);
    ws->ip  = (int *)calloc((size_t)(2 + (int)sqrt((double)n + 0.5)), sizeof(int));
    ws->t   = (double *)calloc((size_t)(8 * nt), sizeof(double));
    ws->w   = (double *)calloc((size_t)(n / 2 + nz / 4), sizeof(double));
    ws->view = (double ***)malloc(sizeof(double **) * (size_t)nx);
    if (!ws->ip || !ws->t || !ws->w || !ws->view) return -1;
    ws->ip[0] = 0;
    for (int i = 0; i < nx; i++) {
        ws->view[i] = (double **)malloc(sizeof(double *) * (size_t)ny);
        for (int j = 0; j < ny; j++) ws->view[i][j] = a + ((size_t)i * ny + j) * nz_;
    }
    return 0;
}

static void ws_free(ooura_ws *ws, int nx) {
    for (int i = 0; i < nx; i++) free(ws->view[i]);
    free(ws->view);
    free(ws->ip);
    free(ws->t);
    free(ws->w);
}

/* A2a_k, Ooura branch */
int ref_A2a_k(int nx, int ny, int nz, double *a) {
    ooura_ws ws;
    if (ws_init(&ws, nx, ny, nz, a)) return -1;
    rdft3d(nx, ny, nz, 1, ws.view, ws.t, ws.ip, ws.w);
    rdft3dsort(nx, ny, nz, 1, ws.view);
    size_t nc = (size_t)nx * ny * (nz / 2 + 1);
    for (size_t i = 0; i < nc; i++) a[2 * i + 1] = -a[2 * i + 1]; /* std::conj */
    ws_free(&ws, nx);
    return 0;
}

/* A_k2a, Ooura branch */
int ref_A_k2a(int nx, int ny, int nz, double *a) {
    ooura_ws ws;
    if (ws_init(&ws, nx, ny, nz, a)) return -1;
    const double scale = 2.0 / ((double)nx * ny * nz);
    size_t       nc    = (size_t)nx * ny * (nz / 2 + 1);
    for (size_t i = 0; i < nc; i++) {
        a[2 * i]     = scale * a[2 * i];
        a[2 * i + 1] = -scale * a[2 * i + 1];
    }
    rdft3dsort(nx, ny, nz, -1, ws.view);
    rdft3d(nx, ny, nz, -1, ws.view, ws.t, ws.ip, ws.w);
    ws_free(&ws, nx);
    return 0;
}

void ref_mt_seed(unsigned long s) { init_genrand(s); }
double ref_mt_real3(void) { return genrand_real3(); }
unsigned long ref_mt_int32(void) { return genrand_int32(); }
