/*
 * fft_impl.h - the oracle's mixed-radix FFT, instantiated twice by fft.c
 * (REAL = float with SFX f32, REAL = double with SFX f64).
 *
 * TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 * Recursive decimation in time: n = p*m, the p sub-transforms of length m over
 * the inputs decimated by p are combined by a twiddled length-p DFT.  Twiddles
 * come from one table of exp(sign 2 pi i j / n) computed in double and rounded
 * to REAL, the accuracy FFTW-float's precomputed twiddles also have.  The 3-D
 * transforms follow FFTW's r2c/c2r conventions (the reference plans them at
 * reionization.c:685-736 and runs them at find_HII_bubbles.c:124-163, :255-269).
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#define CAT_(a, b) a##_##b
#define CAT(a, b) CAT_(a, b)
#define FN(name) CAT(name, SFX)

typedef struct {
  int n;
  int nf;
  int fac[40];
  REAL* tw; /* 2n: exp(sign 2 pi i j / n) */
} FN(plan);

static void FN(plan_init)(FN(plan) * P, int n, int sign)
{
  P->n = n;
  P->nf = 0;
  int r = n;
  while (r % 4 == 0 && r > 4) { P->fac[P->nf++] = 4; r /= 4; }
  while (r % 2 == 0) { P->fac[P->nf++] = 2; r /= 2; }
  for (int q = 3; r > 1; q += 2)
    while (r % q == 0) { P->fac[P->nf++] = q; r /= q; }
  if (P->nf == 0) P->fac[P->nf++] = 1;
  P->tw = (REAL*)malloc(sizeof(REAL) * 2 * (size_t)n);
  for (int j = 0; j < n; j++) {
    double a = (double)sign * 2.0 * M_PI * (double)j / (double)n;
    P->tw[2 * j] = (REAL)cos(a);
    P->tw[2 * j + 1] = (REAL)sin(a);
  }
}

static void FN(plan_free)(FN(plan) * P) { free(P->tw); }

/* X[0..n) = DFT of x[0], x[xs], x[2xs], ...; tws = P->n / n */
static void FN(rec)(const REAL* x, int xs, REAL* X, int n, const int* fac, const FN(plan) * P, int tws)
{
  const int p = fac[0], m = n / p, N = P->n;
  const REAL* tw = P->tw;
  if (m == 1) {
    if (p == 1) { X[0] = x[0]; X[1] = x[1]; return; }
    for (int s = 0; s < p; s++) {
      REAL ar = 0, ai = 0;
      for (int q = 0; q < p; q++) {
        int t = (int)(((long)q * s % p) * (N / p));
        REAL xr = x[2 * (long)q * xs], xi = x[2 * (long)q * xs + 1];
        ar += xr * tw[2 * t] - xi * tw[2 * t + 1];
        ai += xr * tw[2 * t + 1] + xi * tw[2 * t];
      }
      X[2 * s] = ar;
      X[2 * s + 1] = ai;
    }
    return;
  }
  for (int q = 0; q < p; q++) FN(rec)(x + 2 * (long)q * xs, xs * p, X + 2 * q * m, m, fac + 1, P, tws * p);
  REAL tr[64], ti[64];
  for (int k = 0; k < m; k++) {
    for (int q = 0; q < p; q++) {
      REAL yr = X[2 * (q * m + k)], yi = X[2 * (q * m + k) + 1];
      if (q == 0 || k == 0) { tr[q] = yr; ti[q] = yi; continue; }
      int t = q * k * tws;
      REAL wr = tw[2 * t], wi = tw[2 * t + 1];
      tr[q] = yr * wr - yi * wi;
      ti[q] = yr * wi + yi * wr;
    }
    if (p == 2) {
      X[2 * k] = tr[0] + tr[1];
      X[2 * k + 1] = ti[0] + ti[1];
      X[2 * (k + m)] = tr[0] - tr[1];
      X[2 * (k + m) + 1] = ti[0] - ti[1];
    } else if (p == 4) {
      /* W_4 = exp(sign i pi/2) = sign * i, sign read from the table */
      REAL sg = tw[2 * (N / 4) + 1];
      REAL ar = tr[0] + tr[2], ai = ti[0] + ti[2];
      REAL br = tr[0] - tr[2], bi = ti[0] - ti[2];
      REAL cr = tr[1] + tr[3], ci = ti[1] + ti[3];
      REAL dr = tr[1] - tr[3], di = ti[1] - ti[3];
      /* (sign i) * d = (-sign di, sign dr) */
      X[2 * k] = ar + cr;
      X[2 * k + 1] = ai + ci;
      X[2 * (k + m)] = br - sg * di;
      X[2 * (k + m) + 1] = bi + sg * dr;
      X[2 * (k + 2 * m)] = ar - cr;
      X[2 * (k + 2 * m) + 1] = ai - ci;
      X[2 * (k + 3 * m)] = br + sg * di;
      X[2 * (k + 3 * m) + 1] = bi - sg * dr;
    } else {
      for (int s = 0; s < p; s++) {
        REAL ar = 0, ai = 0;
        for (int q = 0; q < p; q++) {
          int t = (int)(((long)q * s % p) * (N / p));
          ar += tr[q] * tw[2 * t] - ti[q] * tw[2 * t + 1];
          ai += tr[q] * tw[2 * t + 1] + ti[q] * tw[2 * t];
        }
        X[2 * (k + m * s)] = ar;
        X[2 * (k + m * s) + 1] = ai;
      }
    }
  }
}

static void FN(exec)(const FN(plan) * P, const REAL* in, REAL* out) { FN(rec)(in, 1, out, P->n, P->fac, P, 1); }

void FN(orc_fft1d)(const REAL* in, REAL* out, int n, int sign)
{
  FN(plan) P;
  FN(plan_init)(&P, n, sign);
  FN(exec)(&P, in, out);
  FN(plan_free)(&P);
}

#define PB 8 /* pencils gathered together so strided passes use whole cache lines */

/* complex pencils of length N along an axis of the half-spectrum array
 * c[(x*N + y)*NH1 + kz]; axis 0 = x, 1 = y.  In place. */
static void FN(pencils)(REAL* c, int N, int axis, const FN(plan) * P)
{
  const int NH1 = N / 2 + 1;
  const long sx = (long)N * NH1, sy = NH1;
  const int nblk_kz = (NH1 + PB - 1) / PB;
#pragma omp parallel
  {
    REAL* a = (REAL*)malloc(sizeof(REAL) * 2 * (size_t)N * PB);
    REAL* b = (REAL*)malloc(sizeof(REAL) * 2 * (size_t)N);
#pragma omp for schedule(static)
    for (long w = 0; w < (long)N * nblk_kz; w++) {
      int other = (int)(w / nblk_kz), kz0 = (int)(w % nblk_kz) * PB;
      int nb = (kz0 + PB <= NH1) ? PB : NH1 - kz0;
      long base = (axis == 0) ? other * sy : other * sx;
      long stride = (axis == 0) ? sx : sy;
      for (int i = 0; i < N; i++)
        for (int j = 0; j < nb; j++) {
          long o = 2 * (base + i * stride + kz0 + j);
          a[2 * ((long)j * N + i)] = c[o];
          a[2 * ((long)j * N + i) + 1] = c[o + 1];
        }
      for (int j = 0; j < nb; j++) {
        FN(exec)(P, a + 2 * (long)j * N, b);
        memcpy(a + 2 * (long)j * N, b, sizeof(REAL) * 2 * (size_t)N);
      }
      for (int i = 0; i < N; i++)
        for (int j = 0; j < nb; j++) {
          long o = 2 * (base + i * stride + kz0 + j);
          c[o] = a[2 * ((long)j * N + i)];
          c[o + 1] = a[2 * ((long)j * N + i) + 1];
        }
    }
    free(a);
    free(b);
  }
}

void FN(orc_r2c3d)(int N, const REAL* in_padded, REAL* out_herm)
{
  const int NH1 = N / 2 + 1, PR = 2 * NH1;
  FN(plan) P;
  FN(plan_init)(&P, N, -1);
#pragma omp parallel
  {
    REAL* a = (REAL*)malloc(sizeof(REAL) * 2 * (size_t)N);
    REAL* b = (REAL*)malloc(sizeof(REAL) * 2 * (size_t)N);
#pragma omp for schedule(static)
    for (long r = 0; r < (long)N * N; r++) {
      for (int z = 0; z < N; z++) {
        a[2 * z] = in_padded[r * PR + z];
        a[2 * z + 1] = 0;
      }
      FN(exec)(&P, a, b);
      memcpy(out_herm + 2 * r * NH1, b, sizeof(REAL) * 2 * (size_t)NH1);
    }
    free(a);
    free(b);
  }
  FN(pencils)(out_herm, N, 1, &P);
  FN(pencils)(out_herm, N, 0, &P);
  FN(plan_free)(&P);
}

void FN(orc_c2r3d)(int N, const REAL* in_herm, REAL* out_padded)
{
  const int NH1 = N / 2 + 1, PR = 2 * NH1;
  const size_t nc = (size_t)N * N * NH1;
  REAL* c = (REAL*)malloc(sizeof(REAL) * 2 * nc);
  memcpy(c, in_herm, sizeof(REAL) * 2 * nc);
  FN(plan) P;
  FN(plan_init)(&P, N, +1);
  FN(pencils)(c, N, 0, &P);
  FN(pencils)(c, N, 1, &P);
#pragma omp parallel
  {
    REAL* a = (REAL*)malloc(sizeof(REAL) * 2 * (size_t)N);
    REAL* b = (REAL*)malloc(sizeof(REAL) * 2 * (size_t)N);
#pragma omp for schedule(static)
    for (long r = 0; r < (long)N * N; r++) {
      const REAL* row = c + 2 * r * NH1;
      for (int k = 0; k < NH1; k++) {
        a[2 * k] = row[2 * k];
        a[2 * k + 1] = row[2 * k + 1];
      }
      for (int k = NH1; k < N; k++) { /* Hermitian completion X[N-k] = conj X[k] */
        a[2 * k] = row[2 * (N - k)];
        a[2 * k + 1] = -row[2 * (N - k) + 1];
      }
      FN(exec)(&P, a, b);
      for (int z = 0; z < N; z++) out_padded[r * PR + z] = b[2 * z];
      for (int z = N; z < PR; z++) out_padded[r * PR + z] = 0;
    }
    free(a);
    free(b);
  }
  FN(plan_free)(&P);
  free(c);
}

#undef PB
#undef FN
