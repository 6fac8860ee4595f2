/*
 * oracle.c -- TEST INFRASTRUCTURE ONLY (the oracle of the parity tests, never the product).
 *
 * CPU binary64 restatement of one MEUMAPPS-SS semi-implicit spectral step.
 * The reference (Fortran + P3DFFT + FFTW + LINPACK + MPI) cannot be built in
 * this image: P3DFFT, FFTW3 and LINPACK are absent, so this file restates the
 * algorithm from the Fortran source, single process, following the
 * reference's operation order wherever it affects rounding:
 *
 *   f_trans / inv_trans ............ ModFunc.f:15-40 (P3DFFT r2c 'fft' + 1/N^3,
 *                                    c2r 'tff', FFTW c2r semantics)
 *   invert .......................... ModFunc.f:44-72 (LU with partial pivoting
 *                                    + inverse, as LINPACK dgeco/dgedi job 11)
 *   k_space ......................... ModInitials.f:348-409
 *   gradient block .................. Multi_MEUMAPPS.f:575-615
 *   KKS / Landau chemistry .......... Multi_MEUMAPPS.f:634-845
 *   eldis / elener .................. ModElastic.f:11-691
 *   reductions ...................... Multi_MEUMAPPS.f:885-907
 *   Cahn-Hilliard update ............ Multi_MEUMAPPS.f:910-949
 *   Allen-Cahn update ............... Multi_MEUMAPPS.f:955-1001
 *   clamp + check loop .............. Multi_MEUMAPPS.f:1008-1065
 *   Langevin noise .................. Multi_MEUMAPPS.f:552-571, 963-978 with
 *                                    ran_2.f:1-43 and langevin ModFunc.f:77-100
 *   tangent-linear step ............. forward-mode derivative of the above
 *                                    (oracle_tlm_*), the reference gradient
 *                                    of the adjoint checks; no reference
 *                                    counterpart
 *
 * Deviations that only move bits in the last place: the FFT is this file's
 * own mixed-radix Cooley-Tukey (not FFTW); the LINPACK rcond estimate is
 * replaced by the exact 1-norm condition number of the computed inverse;
 * global sums are taken per z-slab and then over slabs in z order, which is
 * deterministic for any thread count.  A noise step (step <= noise_step)
 * needs oracle_set_noise first (MM_ERR_ARG otherwise).
 */
#include "oracle.h"

#include <complex.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

typedef double complex cplx;

/* ------------------------------------------------------------------ FFT -- */
typedef struct {
  int n, nf, f[48];
  cplx *w; /* w[e] = exp(-2 pi i e / n), e in [0, n) */
} plan1d;

static int plan_init(plan1d *p, int n) {
  p->n = n;
  p->nf = 0;
  int m = n;
  while (m % 4 == 0) { p->f[p->nf++] = 4; m /= 4; }
  while (m % 2 == 0) { p->f[p->nf++] = 2; m /= 2; }
  while (m % 3 == 0) { p->f[p->nf++] = 3; m /= 3; }
  while (m % 5 == 0) { p->f[p->nf++] = 5; m /= 5; }
  if (m != 1) p->f[p->nf++] = m; /* remaining prime: direct DFT */
  p->w = (cplx *)malloc(sizeof(cplx) * (size_t)n);
  if (!p->w) return -1;
  const long double twopi = 6.283185307179586476925286766559005768L;
  for (int e = 0; e < n; ++e) {
    long double a = twopi * (long double)e / (long double)n;
    p->w[e] = (double)cosl(a) - I * (double)sinl(a);
  }
  return 0;
}

static void plan_free(plan1d *p) { free(p->w); p->w = NULL; }

#define FB 16
/* Stockham autosort transform of FB lines at once, stored split and
 * interleaved: re[e*FB + j], im[e*FB + j] is element e of line j.  Stage with
 * radix r on the current length L (stride s = product of the radices done,
 * m = L/r): for p < m, q < s the inputs x[q + s(p + m t)], t < r, go through an
 * r-point DFT and the outputs, times w_L^(p t), land at y[q + s(r p + t)].
 * sign -1 forward (FFTW's e^{-i...}), +1 backward.  Work arrays t_re/t_im
 * (n*FB each) hold the ping-pong copy; the result ends in re/im. */
static void fft_batch(double *re, double *im, double *t_re, double *t_im, const plan1d *p,
                      int sign) {
  const int n = p->n;
  double *xr = re, *xi = im, *yr = t_re, *yi = t_im;
  int L = n, s = 1;
  const double sg = (double)sign;
  const double s3 = 0.86602540378443864676; /* sin(2 pi / 3) */
  for (int st = 0; st < p->nf; ++st) {
    const int r = p->f[st], m = L / r, wst = n / L;
    for (int pp = 0; pp < m; ++pp) {
      double wr[64], wi[64];
      for (int t = 1; t < r; ++t) {
        cplx w = p->w[(long)pp * t * wst];
        wr[t] = creal(w);
        wi[t] = sg < 0 ? cimag(w) : -cimag(w);
      }
      for (int q = 0; q < s; ++q) {
        const long ib = ((long)q + (long)s * pp) * FB, istep = (long)s * m * FB;
        const long ob = ((long)q + (long)s * r * pp) * FB, ostep = (long)s * FB;
        if (r == 4) {
          for (int j = 0; j < FB; ++j) {
            double a0r = xr[ib + j], a0i = xi[ib + j];
            double a1r = xr[ib + istep + j], a1i = xi[ib + istep + j];
            double a2r = xr[ib + 2 * istep + j], a2i = xi[ib + 2 * istep + j];
            double a3r = xr[ib + 3 * istep + j], a3i = xi[ib + 3 * istep + j];
            double b0r = a0r + a2r, b0i = a0i + a2i, b1r = a0r - a2r, b1i = a0i - a2i;
            double b2r = a1r + a3r, b2i = a1i + a3i, b3r = a1r - a3r, b3i = a1i - a3i;
            /* (sign i) * b3 */
            double c3r = -sg * b3i, c3i = sg * b3r;
            double u0r = b0r + b2r, u0i = b0i + b2i;
            double u1r = b1r + c3r, u1i = b1i + c3i;
            double u2r = b0r - b2r, u2i = b0i - b2i;
            double u3r = b1r - c3r, u3i = b1i - c3i;
            yr[ob + j] = u0r; yi[ob + j] = u0i;
            if (pp == 0) {
              yr[ob + ostep + j] = u1r; yi[ob + ostep + j] = u1i;
              yr[ob + 2 * ostep + j] = u2r; yi[ob + 2 * ostep + j] = u2i;
              yr[ob + 3 * ostep + j] = u3r; yi[ob + 3 * ostep + j] = u3i;
            } else {
              yr[ob + ostep + j] = u1r * wr[1] - u1i * wi[1];
              yi[ob + ostep + j] = u1r * wi[1] + u1i * wr[1];
              yr[ob + 2 * ostep + j] = u2r * wr[2] - u2i * wi[2];
              yi[ob + 2 * ostep + j] = u2r * wi[2] + u2i * wr[2];
              yr[ob + 3 * ostep + j] = u3r * wr[3] - u3i * wi[3];
              yi[ob + 3 * ostep + j] = u3r * wi[3] + u3i * wr[3];
            }
          }
        } else if (r == 2) {
          for (int j = 0; j < FB; ++j) {
            double a0r = xr[ib + j], a0i = xi[ib + j];
            double a1r = xr[ib + istep + j], a1i = xi[ib + istep + j];
            double u1r = a0r - a1r, u1i = a0i - a1i;
            yr[ob + j] = a0r + a1r; yi[ob + j] = a0i + a1i;
            if (pp == 0) {
              yr[ob + ostep + j] = u1r; yi[ob + ostep + j] = u1i;
            } else {
              yr[ob + ostep + j] = u1r * wr[1] - u1i * wi[1];
              yi[ob + ostep + j] = u1r * wi[1] + u1i * wr[1];
            }
          }
        } else if (r == 3) {
          for (int j = 0; j < FB; ++j) {
            double a0r = xr[ib + j], a0i = xi[ib + j];
            double a1r = xr[ib + istep + j], a1i = xi[ib + istep + j];
            double a2r = xr[ib + 2 * istep + j], a2i = xi[ib + 2 * istep + j];
            double tr = a1r + a2r, ti = a1i + a2i, dr = a1r - a2r, di = a1i - a2i;
            double mr = a0r - 0.5 * tr, mi = a0i - 0.5 * ti;
            /* u1 = m + sign*i*s3*d, u2 = m - sign*i*s3*d */
            double er = -sg * s3 * di, ei = sg * s3 * dr;
            double u1r = mr + er, u1i = mi + ei, u2r = mr - er, u2i = mi - ei;
            yr[ob + j] = a0r + tr; yi[ob + j] = a0i + ti;
            if (pp == 0) {
              yr[ob + ostep + j] = u1r; yi[ob + ostep + j] = u1i;
              yr[ob + 2 * ostep + j] = u2r; yi[ob + 2 * ostep + j] = u2i;
            } else {
              yr[ob + ostep + j] = u1r * wr[1] - u1i * wi[1];
              yi[ob + ostep + j] = u1r * wi[1] + u1i * wr[1];
              yr[ob + 2 * ostep + j] = u2r * wr[2] - u2i * wi[2];
              yi[ob + 2 * ostep + j] = u2r * wi[2] + u2i * wr[2];
            }
          }
        } else { /* any other radix: direct r-point DFT, then the twiddle */
          const int wr_st = n / r; /* w_r^(ts) = w_n^(ts n/r) */
          for (int j = 0; j < FB; ++j) {
            for (int t = 0; t < r; ++t) {
              double accr = 0.0, acci = 0.0;
              for (int u = 0; u < r; ++u) {
                cplx w = p->w[(long)((t * u) % r) * wr_st];
                double cr = creal(w), ci = sg < 0 ? cimag(w) : -cimag(w);
                double vr = xr[ib + u * istep + j], vi = xi[ib + u * istep + j];
                accr += vr * cr - vi * ci;
                acci += vr * ci + vi * cr;
              }
              if (pp != 0 && t != 0) {
                double zr = accr * wr[t] - acci * wi[t];
                acci = accr * wi[t] + acci * wr[t];
                accr = zr;
              }
              yr[ob + t * ostep + j] = accr; yi[ob + t * ostep + j] = acci;
            }
          }
        }
      }
    }
    double *sw;
    sw = xr; xr = yr; yr = sw;
    sw = xi; xi = yi; yi = sw;
    L = m;
    s *= r;
  }
  if (xr != re) {
    memcpy(re, xr, sizeof(double) * (size_t)n * FB);
    memcpy(im, xi, sizeof(double) * (size_t)n * FB);
  }
}

/* transform `count` lines: line l starts at base + l_off(l), element stride es.
 * Lines l, l+1, ... with consecutive offsets are gathered FB at a time so the
 * strided reads touch whole cache lines. */
static void fft_lines(cplx *data, const plan1d *p, int count, long (*off)(int, const int *),
                      const int *geo, long es, int sign, int nthreads) {
  int n = p->n;
  int nblk = (count + FB - 1) / FB;
#pragma omp parallel num_threads(nthreads)
  {
    double *w = (double *)malloc(sizeof(double) * (size_t)n * FB * 4);
    double *ar = w, *ai = w + (long)n * FB, *tr = w + 2L * n * FB, *ti = w + 3L * n * FB;
#pragma omp for schedule(static)
    for (int blk = 0; blk < nblk; ++blk) {
      int l0 = blk * FB, nb = count - l0 < FB ? count - l0 : FB;
      long o0 = off(l0, geo);
      int run = 1; /* lines l0..l0+nb-1 at consecutive offsets? */
      for (int j = 1; j < nb; ++j) run &= (off(l0 + j, geo) == o0 + j);
      for (int e = 0; e < n; ++e)
        for (int j = nb; j < FB; ++j) ar[(long)e * FB + j] = ai[(long)e * FB + j] = 0.0;
      for (int j = 0; j < nb; ++j) {
        const cplx *src = data + (run ? o0 + j : off(l0 + j, geo));
        for (int e = 0; e < n; ++e) {
          cplx v = src[(long)e * es];
          ar[(long)e * FB + j] = creal(v);
          ai[(long)e * FB + j] = cimag(v);
        }
      }
      fft_batch(ar, ai, tr, ti, p, sign);
      for (int j = 0; j < nb; ++j) {
        cplx *dst = data + (run ? o0 + j : off(l0 + j, geo));
        for (int e = 0; e < n; ++e)
          dst[(long)e * es] = CMPLX(ar[(long)e * FB + j], ai[(long)e * FB + j]);
      }
    }
    free(w);
  }
}

/* geo = {H, ny, nz} */
static long off_y(int l, const int *g) { /* l = z*H + kx ; line along y */
  int H = g[0], ny = g[1];
  long z = l / H, kx = l % H;
  return z * (long)ny * H + kx;
}
static long off_z(int l, const int *g) { /* l = y*H + kx ; line along z */
  return (long)l;
}

/* ----------------------------------------------------------- small LU inverse */
/* In-place inverse of the column-major n x n matrix a, Gaussian elimination
 * with partial pivoting (LINPACK dgefa) followed by the inverse from the
 * factors (dgedi job 1); the rcond test of ModFunc.f:55 is applied to the
 * exact 1-norm condition number.  Returns 0 or `flag`. */
int oracle_invert(double *a, int n, int flag) {
  if (n <= 0 || n > 32) return flag;
  int ipvt[32];
  double anorm = 0.0;
  for (int j = 0; j < n; ++j) {
    double s = 0.0;
    for (int i = 0; i < n; ++i) s += fabs(a[i + j * n]);
    if (s > anorm) anorm = s;
  }
  /* dgefa */
  for (int k = 0; k < n - 1; ++k) {
    int l = k;
    double amax = fabs(a[k + k * n]);
    for (int i = k + 1; i < n; ++i)
      if (fabs(a[i + k * n]) > amax) { amax = fabs(a[i + k * n]); l = i; }
    ipvt[k] = l;
    if (a[l + k * n] == 0.0) return flag;
    if (l != k) { double t = a[l + k * n]; a[l + k * n] = a[k + k * n]; a[k + k * n] = t; }
    double t = -1.0 / a[k + k * n];
    for (int i = k + 1; i < n; ++i) a[i + k * n] *= t;
    for (int j = k + 1; j < n; ++j) {
      double tj = a[l + j * n];
      if (l != k) { a[l + j * n] = a[k + j * n]; a[k + j * n] = tj; }
      for (int i = k + 1; i < n; ++i) a[i + j * n] += tj * a[i + k * n];
    }
  }
  ipvt[n - 1] = n - 1;
  if (a[(n - 1) + (n - 1) * n] == 0.0) return flag;
  /* dgedi: inverse(U) */
  for (int k = 0; k < n; ++k) {
    a[k + k * n] = 1.0 / a[k + k * n];
    double t = -a[k + k * n];
    for (int i = 0; i < k; ++i) a[i + k * n] *= t;
    for (int j = k + 1; j < n; ++j) {
      double tj = a[k + j * n];
      a[k + j * n] = 0.0;
      for (int i = 0; i <= k; ++i) a[i + j * n] += tj * a[i + k * n];
    }
  }
  /* inverse(U) * inverse(L) */
  double work[32];
  for (int kb = 0; kb < n - 1; ++kb) {
    int k = n - 2 - kb;
    for (int i = k + 1; i < n; ++i) { work[i] = a[i + k * n]; a[i + k * n] = 0.0; }
    for (int j = k + 1; j < n; ++j) {
      double tj = work[j];
      for (int i = 0; i < n; ++i) a[i + k * n] += tj * a[i + j * n];
    }
    int l = ipvt[k];
    if (l != k)
      for (int i = 0; i < n; ++i) { double t = a[i + k * n]; a[i + k * n] = a[i + l * n]; a[i + l * n] = t; }
  }
  double inorm = 0.0;
  for (int j = 0; j < n; ++j) {
    double s = 0.0;
    for (int i = 0; i < n; ++i) s += fabs(a[i + j * n]);
    if (s > inorm) inorm = s;
  }
  double rcond = (anorm > 0.0 && inorm > 0.0) ? 1.0 / (anorm * inorm) : 0.0;
  if (!(rcond > 1.0e-16)) return flag;
  return 0;
}

/* ---------------------------------------------------------------- k-space */
/* ModInitials.f:355-401: kf = i * m * fl(2 pi / N), m wrapped to (-N/2, N/2],
 * zeroed at m = N/2; kf_sq = kf*kf kept at the Nyquist index. */
void oracle_kaxis(int n, int count, double *kf_im, double *kq) {
  double pi = atan(1.0) * 4.0;
  double unit = (2.0 * pi) / (double)(float)n;
  for (int m = 0; m < count; ++m) {
    int mw = (m <= n / 2) ? m : m - n;
    double a = (double)(float)mw * unit;
    kf_im[m] = (2 * m == n) ? 0.0 : a;
    kq[m] = -(a * a);
  }
}

/* ----------------------------------------------------------------- oracle */
struct oracle {
  mm_params p;
  int nx, ny, nz, H, V, E, NPH, zdim;
  long n3, nc;
  int nthreads;
  plan1d px, py, pz;
  double *kfx, *kqx, *kfy, *kqy, *kfz, *kqz;
  double ooN;
  int phase_of[MM_MAX_VAR]; /* 1-based phase of each variant */
  /* state */
  double *phi, *cons, *u, *gu, *el_en;
  double glob_min, glob_max, phi_max, solid_frac;
  /* work */
  double *gradx, *grady, *gradz, *kdel, *fgrad, *fdiff, *elgrad, *Pphi, *Fx, *Fy, *Fz;
  double *xf, *u_old;
  cplx *sphi;  /* V spectra of phi */
  cplx *sw;    /* 12 work spectra */
  double *qinv; /* 9 per k (column-major 3x3) */
  int qinv_ready;
  int step;
  /* Langevin noise: ran_2 state (idum, then idum2, iy, iv(32), init flag) */
  int noise_on;
  double noise_sd;
  int32_t r_idum, r_state[35];
  double *phi_noise; /* V*N3, accumulated over the noise steps */
  struct tlm *t;     /* tangent-linear state (oracle_tlm_set), or NULL */
};

#define RIDX(o, x, y, z) ((long)(x) + (long)(o)->nx * ((long)(y) + (long)(o)->ny * (long)(z)))

static void *xcalloc(long n, size_t sz) { return calloc((size_t)(n > 0 ? n : 1), sz); }

oracle_t *oracle_create(const mm_params *p, int nthreads) {
  if (p->nx < 2 || p->ny < 2 || p->nz < 2 || (p->nx & 1) || (p->ny & 1) || (p->nz & 1)) return NULL;
  if (p->vartot < 1 || p->vartot > MM_MAX_VAR || p->numel < 1 || p->numel > MM_MAX_EL ||
      p->numph < 2 || p->numph > MM_MAX_PH)
    return NULL;
  int s = 0;
  for (int q = 0; q < p->numph; ++q) s += p->nvar[q];
  if (s != p->vartot + 1 || p->nvar[0] != 1) return NULL;
  oracle_t *o = (oracle_t *)calloc(1, sizeof(oracle_t));
  o->p = *p;
  o->nx = p->nx; o->ny = p->ny; o->nz = p->nz; o->H = p->nx / 2 + 1;
  o->V = p->vartot; o->E = p->numel; o->NPH = p->numph; o->zdim = p->numph * p->numel;
  o->n3 = (long)o->nx * o->ny * o->nz;
  o->nc = (long)o->H * o->ny * o->nz;
  o->nthreads = nthreads > 0 ? nthreads : 1;
  o->step = p->step;
  plan_init(&o->px, o->nx); plan_init(&o->py, o->ny); plan_init(&o->pz, o->nz);
  o->kfx = malloc(sizeof(double) * o->H); o->kqx = malloc(sizeof(double) * o->H);
  o->kfy = malloc(sizeof(double) * o->ny); o->kqy = malloc(sizeof(double) * o->ny);
  o->kfz = malloc(sizeof(double) * o->nz); o->kqz = malloc(sizeof(double) * o->nz);
  oracle_kaxis(o->nx, o->H, o->kfx, o->kqx);
  oracle_kaxis(o->ny, o->ny, o->kfy, o->kqy);
  oracle_kaxis(o->nz, o->nz, o->kfz, o->kqz);
  o->ooN = 1.0 / (double)(float)((long)o->nx * o->ny * o->nz);
  {
    int ivn = 0;
    for (int ph = 2; ph <= o->NPH; ++ph) {
      ivn += p->nvar[ph - 2];
      for (int v = ivn; v <= ivn + p->nvar[ph - 1] - 1; ++v) o->phase_of[v - 1] = ph;
    }
  }
  long n3 = o->n3, V = o->V, E = o->E;
  o->phi = xcalloc(V * n3, 8); o->cons = xcalloc(E * n3, 8);
  o->u = xcalloc(3 * n3, 8); o->gu = xcalloc(9 * n3, 8); o->el_en = xcalloc(n3, 8);
  o->gradx = xcalloc(V * n3, 8); o->grady = xcalloc(V * n3, 8); o->gradz = xcalloc(V * n3, 8);
  o->kdel = xcalloc(V * n3, 8); o->fgrad = xcalloc(V * n3, 8); o->fdiff = xcalloc(V * n3, 8);
  o->elgrad = xcalloc(V * n3, 8); o->Pphi = xcalloc(V * n3, 8);
  o->Fx = xcalloc(E * n3, 8); o->Fy = xcalloc(E * n3, 8); o->Fz = xcalloc(E * n3, 8);
  o->xf = xcalloc(6 * n3, 8); o->u_old = xcalloc(3 * n3, 8);
  o->sphi = xcalloc(V * o->nc, sizeof(cplx));
  o->sw = xcalloc(12 * o->nc, sizeof(cplx));
  o->qinv = xcalloc(9 * o->nc, 8);
  o->glob_min = 1.0; o->glob_max = 0.0; o->phi_max = 0.0; o->solid_frac = 0.0;
  if (!o->phi || !o->sw || !o->qinv) { oracle_destroy(o); return NULL; }
  return o;
}

void oracle_destroy(oracle_t *o) {
  if (!o) return;
  plan_free(&o->px); plan_free(&o->py); plan_free(&o->pz);
  double *bufs[] = {o->kfx, o->kqx, o->kfy, o->kqy, o->kfz, o->kqz, o->phi, o->cons, o->u,
                    o->gu, o->el_en, o->gradx, o->grady, o->gradz, o->kdel, o->fgrad,
                    o->fdiff, o->elgrad, o->Pphi, o->Fx, o->Fy, o->Fz, o->xf, o->u_old, o->qinv};
  for (size_t i = 0; i < sizeof(bufs) / sizeof(bufs[0]); ++i) free(bufs[i]);
  free(o->sphi); free(o->sw); free(o->phi_noise);
  oracle_tlm_off(o);
  free(o);
}

/* ran_2.f:1-43: L'Ecuyer's two multiplicative generators (Schrage's
 * decomposition keeps 32-bit products) with a 32-entry Bays-Durham shuffle.
 * st[0] = idum2, st[1] = iy, st[2..33] = iv(1..32); st[34] != 0 once the
 * DATA values (idum2 = 123456789, iv = 0, iy = 0) are in place.  am and rnmx
 * follow the Fortran literals: 1./im1 is a binary32 quotient. */
double oracle_ran2(int32_t *idum, int32_t *st) {
  const int32_t im1 = 2147483563, im2 = 2147483399, ia1 = 40014, ia2 = 40692, iq1 = 53668, iq2 = 52774,
                ir1 = 12211, ir2 = 3791, ntab = 32, ndiv = 1 + (im1 - 1) / ntab;
  if (!st[34]) {
    st[0] = 123456789; st[1] = 0;
    for (int j = 0; j < ntab; ++j) st[2 + j] = 0;
    st[34] = 1;
  }
  int32_t *iv = st + 2, k;
  if (*idum <= 0) {
    *idum = -*idum > 1 ? -*idum : 1;
    st[0] = *idum;
    for (int j = ntab + 8; j >= 1; --j) {
      k = *idum / iq1;
      *idum = ia1 * (*idum - k * iq1) - k * ir1;
      if (*idum < 0) *idum += im1;
      if (j <= ntab) iv[j - 1] = *idum;
    }
    st[1] = iv[0];
  }
  k = *idum / iq1;
  *idum = ia1 * (*idum - k * iq1) - k * ir1;
  if (*idum < 0) *idum += im1;
  k = st[0] / iq2;
  st[0] = ia2 * (st[0] - k * iq2) - k * ir2;
  if (st[0] < 0) st[0] += im2;
  int32_t j = st[1] / ndiv;
  st[1] = iv[j] - st[0];
  iv[j] = *idum;
  if (st[1] < 1) st[1] += im1 - 1;
  const double am = (double)(1.0f / (float)im1), rnmx = 1.0 - 1.2e-7;
  double r = am * (double)st[1];
  return r < rnmx ? r : rnmx;
}

/* ModFunc.f:77-100 */
double oracle_langevin(int32_t *idum, int32_t *st, double sd) {
  for (;;) {
    double v1 = oracle_ran2(idum, st), v2 = oracle_ran2(idum, st);
    v1 = 2.0 * v1 - 1.0;
    v2 = 2.0 * v2 - 1.0;
    double a = v1 * v1, b = v2 * v2;
    double rsq = a + b;
    if (rsq > 0.0 && rsq < 1.0) {
      double l = log(rsq);
      rsq = sqrt(-2.0 * l / rsq);
      double h = v2 * rsq;
      return h * sd;
    }
  }
}

/* Multi_MEUMAPPS.f:554-562 for one process, sequentially */
void oracle_noise_accumulate(int32_t *idum, int32_t *st, double sd, int nvar, long n3, double *acc) {
  for (long i = 0; i < n3; ++i)
    for (int v = 0; v < nvar; ++v) acc[v * n3 + i] += oracle_langevin(idum, st, sd);
}

int oracle_set_noise(oracle_t *o, double sd, int32_t idum, const int32_t *state) {
  if (!o->phi_noise) o->phi_noise = xcalloc(o->V * o->n3, 8);
  if (!o->phi_noise) return MM_ERR_NOMEM;
  memset(o->phi_noise, 0, sizeof(double) * o->V * o->n3);
  o->r_idum = idum;
  if (state) memcpy(o->r_state, state, sizeof(o->r_state));
  else memset(o->r_state, 0, sizeof(o->r_state));
  o->noise_sd = sd;
  o->noise_on = 1;
  return 0;
}

int oracle_get_noise(oracle_t *o, int32_t *idum, int32_t *state, double *phi_noise) {
  if (idum) *idum = o->r_idum;
  if (state) memcpy(state, o->r_state, sizeof(o->r_state));
  if (phi_noise && o->phi_noise) memcpy(phi_noise, o->phi_noise, sizeof(double) * o->V * o->n3);
  return 0;
}

int oracle_set_threads(oracle_t *o, int nthreads) {
  o->nthreads = nthreads > 0 ? nthreads : 1;
  return 0;
}

/* the check-loop values of the last step (glob_min, glob_max, phi_max and the
 * unnormalised solid sum), which the next fort.98 row reports */
void oracle_get_lagged(oracle_t *o, double *v) {
  v[0] = o->glob_min; v[1] = o->glob_max; v[2] = o->phi_max; v[3] = o->solid_frac;
}
void oracle_set_lagged(oracle_t *o, const double *v) {
  o->glob_min = v[0]; o->glob_max = v[1]; o->phi_max = v[2]; o->solid_frac = v[3];
}
int oracle_get_step(oracle_t *o) { return o->step; }
void oracle_set_step(oracle_t *o, int step) { o->step = step; }

int oracle_set_state(oracle_t *o, const double *phi, const double *cons, const double *u,
                     const double *gradu, const double *el_en) {
  long n3 = o->n3;
  if (phi) memcpy(o->phi, phi, sizeof(double) * o->V * n3);
  if (cons) memcpy(o->cons, cons, sizeof(double) * o->E * n3);
  if (u) memcpy(o->u, u, sizeof(double) * 3 * n3);
  if (gradu) memcpy(o->gu, gradu, sizeof(double) * 9 * n3);
  if (el_en) memcpy(o->el_en, el_en, sizeof(double) * n3);
  return 0;
}

int oracle_get_state(oracle_t *o, double *phi, double *cons, double *u, double *gradu,
                     double *el_en) {
  long n3 = o->n3;
  if (phi) memcpy(phi, o->phi, sizeof(double) * o->V * n3);
  if (cons) memcpy(cons, o->cons, sizeof(double) * o->E * n3);
  if (u) memcpy(u, o->u, sizeof(double) * 3 * n3);
  if (gradu) memcpy(gradu, o->gu, sizeof(double) * 9 * n3);
  if (el_en) memcpy(el_en, o->el_en, sizeof(double) * n3);
  return 0;
}

int oracle_get_work(oracle_t *o, int which, double *out) {
  long n3 = o->n3, V = o->V, E = o->E;
  switch (which) {
    case 0:
      memcpy(out, o->gradx, 8 * V * n3);
      memcpy(out + V * n3, o->grady, 8 * V * n3);
      memcpy(out + 2 * V * n3, o->gradz, 8 * V * n3);
      return 0;
    case 1: memcpy(out, o->kdel, 8 * V * n3); return 0;
    case 2: memcpy(out, o->fgrad, 8 * V * n3); return 0;
    case 3: memcpy(out, o->fdiff, 8 * V * n3); return 0;
    case 4: memcpy(out, o->elgrad, 8 * V * n3); return 0;
    case 5:
      memcpy(out, o->Fx, 8 * E * n3);
      memcpy(out + E * n3, o->Fy, 8 * E * n3);
      memcpy(out + 2 * E * n3, o->Fz, 8 * E * n3);
      return 0;
    case 6: memcpy(out, o->Pphi, 8 * V * n3); return 0;
  }
  return MM_ERR_ARG;
}

/* f_trans (ModFunc.f:15-26): r2c over x, c2c over y and z, then * ooNxyz */
static void f_trans(oracle_t *o, const double *in, cplx *out) {
  int nx = o->nx, ny = o->ny, nz = o->nz, H = o->H, nt = o->nthreads;
  int rows = ny * nz, nblk = (rows + FB - 1) / FB;
#pragma omp parallel num_threads(nt)
  {
    double *w = (double *)malloc(sizeof(double) * (size_t)nx * FB * 4);
    double *ar = w, *ai = w + (long)nx * FB, *tr = w + 2L * nx * FB, *ti = w + 3L * nx * FB;
#pragma omp for schedule(static)
    for (int blk = 0; blk < nblk; ++blk) {
      int r0 = blk * FB, nb = rows - r0 < FB ? rows - r0 : FB;
      for (int x = 0; x < nx; ++x)
        for (int j = 0; j < FB; ++j) {
          ar[(long)x * FB + j] = j < nb ? in[(long)(r0 + j) * nx + x] : 0.0;
          ai[(long)x * FB + j] = 0.0;
        }
      fft_batch(ar, ai, tr, ti, &o->px, -1);
      for (int j = 0; j < nb; ++j) {
        cplx *orow = out + (long)(r0 + j) * H;
        for (int k = 0; k < H; ++k) orow[k] = CMPLX(ar[(long)k * FB + j], ai[(long)k * FB + j]);
      }
    }
    free(w);
  }
  int geo[3] = {H, ny, nz};
  /* along y: line (z, kx), elements stride H */
  fft_lines(out, &o->py, nz * H, off_y, geo, H, -1, nt);
  /* along z: line (y, kx), elements stride H*ny */
  fft_lines(out, &o->pz, ny * H, off_z, geo, (long)H * ny, -1, nt);
  double s = o->ooN;
  long nc = o->nc;
#pragma omp parallel for num_threads(nt) schedule(static)
  for (long i = 0; i < nc; ++i) out[i] = out[i] * s;
}

/* inv_trans (ModFunc.f:30-40): c2c over z and y, then c2r over x ignoring
 * the imaginary parts of the kx = 0 and kx = Nx/2 entries.  `in` is
 * overwritten. */
static void inv_trans(oracle_t *o, cplx *in, double *out) {
  int nx = o->nx, ny = o->ny, nz = o->nz, H = o->H, nt = o->nthreads;
  int geo[3] = {H, ny, nz};
  fft_lines(in, &o->pz, ny * H, off_z, geo, (long)H * ny, +1, nt);
  fft_lines(in, &o->py, nz * H, off_y, geo, H, +1, nt);
  int rows = ny * nz, nblk = (rows + FB - 1) / FB;
#pragma omp parallel num_threads(nt)
  {
    double *w = (double *)malloc(sizeof(double) * (size_t)nx * FB * 4);
    double *ar = w, *ai = w + (long)nx * FB, *tr = w + 2L * nx * FB, *ti = w + 3L * nx * FB;
#pragma omp for schedule(static)
    for (int blk = 0; blk < nblk; ++blk) {
      int r0 = blk * FB, nb = rows - r0 < FB ? rows - r0 : FB;
      for (int j = 0; j < FB; ++j) {
        const cplx *irow = in + (long)(j < nb ? r0 + j : r0) * H;
        /* Hermitian extension; the imaginary parts of kx = 0 and Nx/2 are dropped */
        for (int k = 0; k < H; ++k) {
          ar[(long)k * FB + j] = creal(irow[k]);
          ai[(long)k * FB + j] = cimag(irow[k]);
        }
        ai[j] = 0.0;
        if ((nx & 1) == 0) ai[(long)(nx / 2) * FB + j] = 0.0;
        for (int k = H; k < nx; ++k) {
          ar[(long)k * FB + j] = ar[(long)(nx - k) * FB + j];
          ai[(long)k * FB + j] = -ai[(long)(nx - k) * FB + j];
        }
      }
      fft_batch(ar, ai, tr, ti, &o->px, +1);
      for (int j = 0; j < nb; ++j) {
        double *row = out + (long)(r0 + j) * nx;
        for (int x = 0; x < nx; ++x) row[x] = ar[(long)x * FB + j];
      }
    }
    free(w);
  }
}

int oracle_f_trans(oracle_t *o, const double *in, double *spec) {
  f_trans(o, in, (cplx *)spec);
  return 0;
}

int oracle_inv_trans(oracle_t *o, const double *spec, double *out) {
  cplx *tmp = (cplx *)malloc(sizeof(cplx) * o->nc);
  memcpy(tmp, spec, sizeof(cplx) * o->nc);
  inv_trans(o, tmp, out);
  free(tmp);
  return 0;
}

/* kf and kf_sq at spectrum index (kx, ky, kz) */
#define KX(o, i) ((o)->kfx[i])
#define GC(o, v, i, j) ((o)->p.grad_coeff[(v) * 9 + (j) * 3 + (i)]) /* 0-based */
#define TP(o, m, v) ((o)->p.trans_ppt[(v) * 6 + (m)])
#define CE(o, i, j, ph) ((o)->p.cela[((ph) - 1) * 36 + (j) * 6 + (i)]) /* ph 1-based */
#define SE(o, i, j, ph) ((o)->p.sela[((ph) - 1) * 36 + (j) * 6 + (i)])
#define GP(o, j, ph) ((o)->p.gparams[((ph) - 1) * MM_GP_STRIDE + ((j) - 1)]) /* 1-based */

static inline cplx mul_i(cplx z, double a) { /* z * (i a) */
  return (-cimag(z) * a) + I * (creal(z) * a);
}

/* the 2-form  sum G_ij k_i k_j  of Multi_MEUMAPPS.f:596-600 / :979-983 */
static inline double kgk(const oracle_t *o, int v, int kx, int ky, int kz) {
  double k1q = o->kqx[kx], k2q = o->kqy[ky], k3q = o->kqz[kz];
  double a1 = o->kfx[kx], a2 = o->kfy[ky], a3 = o->kfz[kz];
  double k1k2 = -(a1 * a2), k2k3 = -(a2 * a3), k3k1 = -(a3 * a1);
  return GC(o, v, 0, 0) * k1q + GC(o, v, 1, 1) * k2q + GC(o, v, 2, 2) * k3q +
         (GC(o, v, 0, 1) + GC(o, v, 1, 0)) * k1k2 + (GC(o, v, 1, 2) + GC(o, v, 2, 1)) * k2k3 +
         (GC(o, v, 2, 0) + GC(o, v, 0, 2)) * k3k1;
}

/* gradient block, Multi_MEUMAPPS.f:575-615 */
static void gradients(oracle_t *o) {
  long n3 = o->n3, nc = o->nc;
  int H = o->H, ny = o->ny, nt = o->nthreads;
  for (int v = 0; v < o->V; ++v) {
    cplx *S = o->sphi + v * nc;
    f_trans(o, o->phi + v * n3, S);
    cplx *gx = o->sw, *gy = o->sw + nc, *gz = o->sw + 2 * nc, *kd = o->sw + 3 * nc;
#pragma omp parallel for num_threads(nt) schedule(static)
    for (long i = 0; i < nc; ++i) {
      int kx = (int)(i % H), ky = (int)((i / H) % ny), kz = (int)(i / ((long)H * ny));
      cplx s = S[i];
      gx[i] = mul_i(s, o->kfx[kx]);
      gy[i] = mul_i(s, o->kfy[ky]);
      gz[i] = mul_i(s, o->kfz[kz]);
      kd[i] = s * kgk(o, v, kx, ky, kz);
    }
    inv_trans(o, gx, o->gradx + v * n3);
    inv_trans(o, gy, o->grady + v * n3);
    inv_trans(o, gz, o->gradz + v * n3);
    inv_trans(o, kd, o->kdel + v * n3);
  }
}

/* per-point chemistry, Multi_MEUMAPPS.f:634-845.  Returns 0 or err. */
static int chem_point(const oracle_t *o, long pt, double *fint, double *fch, double *fw) {
  const int V = o->V, E = o->E, NPH = o->NPH, zdim = o->zdim;
  long n3 = o->n3;
  double phi[MM_MAX_VAR], Ppr[MM_MAX_VAR], Pph[MM_MAX_PH + 1], Gph[MM_MAX_PH + 1];
  for (int v = 0; v < V; ++v) phi[v] = o->phi[v * n3 + pt];
  for (int q = 0; q <= NPH; ++q) { Pph[q] = 0.0; Gph[q] = 0.0; }
  int ivn = 0;
  for (int ph = 2; ph <= NPH; ++ph) {
    ivn += o->p.nvar[ph - 2];
    int fvn = ivn + o->p.nvar[ph - 1] - 1;
    for (int iv = ivn; iv <= fvn; ++iv) {
      int v = iv - 1;
      double fi = phi[v], fi_sq = fi * fi;
      double pl = fi_sq * fi * (6.0 * fi_sq - 15.0 * fi + 10.0);
      o->Pphi[v * n3 + pt] = pl;
      Ppr[v] = 30.0 * fi_sq * (fi_sq - 2.0 * fi + 1.0);
      Pph[1] += pl;
      Pph[ph] += pl;
      Gph[ph] += fi_sq * ((1.0 - fi) * (1.0 - fi));
      for (int jv = 0; jv < V; ++jv)
        if (jv != v) Gph[1] += fi_sq * (phi[jv] * phi[jv]);
    }
  }
  Pph[1] = 1.0 - Pph[1];
  /* Zintf: constant rows (Multi_MEUMAPPS.f:426-448) + mixture rows (:677-688) */
  double Z[16 * 16], Zr[16];
  memset(Z, 0, sizeof(Z));
  for (int e = 1; e <= E; ++e) {
    int i1 = e + E + 1;
    for (int ph = 1; ph <= NPH - 1; ++ph) {
      int i0 = (NPH - 1) * (e - 1) + ph, j0 = NPH * (e - 1) + ph;
      Z[(i0 - 1) + (j0 - 1) * zdim] = GP(o, e, ph);
      Z[(i0 - 1) + j0 * zdim] = -GP(o, e, ph + 1);
      Zr[i0 - 1] = GP(o, e, ph) * GP(o, i1, ph) - GP(o, e, ph + 1) * GP(o, i1, ph + 1);
    }
  }
  for (int e = 1; e <= E; ++e) {
    int i0 = zdim - E + e;
    Zr[i0 - 1] = o->cons[(e - 1) * n3 + pt];
    for (int ph = 1; ph <= NPH; ++ph) {
      int j0 = NPH * (e - 1) + ph;
      Z[(i0 - 1) + (j0 - 1) * zdim] = Pph[ph];
    }
  }
  if (oracle_invert(Z, zdim, MM_ERR_ZSOLVE)) return MM_ERR_ZSOLVE;
  double X[MM_MAX_EL + 1][MM_MAX_PH + 1];
  for (int ph = 1; ph <= NPH; ++ph)
    for (int e = 1; e <= E; ++e) {
      int i0 = NPH * (e - 1) + ph;
      double acc = 0.0;
      for (int j0 = 1; j0 <= zdim; ++j0) acc += Z[(i0 - 1) + (j0 - 1) * zdim] * Zr[j0 - 1];
      X[e][ph] = acc;
    }
  double Gf[MM_MAX_PH + 1] = {0}, Mu[MM_MAX_EL + 1] = {0};
  for (int ph = 1; ph <= NPH; ++ph) {
    double g = 0.0;
    for (int e = 1; e <= E; ++e) {
      double d = X[e][ph] - GP(o, e + E + 1, ph);
      g += GP(o, e, ph) * (d * d);
    }
    Gf[ph] = g + GP(o, E + 1, ph);
  }
  for (int e = 1; e <= E; ++e) Mu[e] = 2.0 * GP(o, e, 1) * (X[e][1] - GP(o, e + E + 1, 1));
  double fchloc = Pph[1] * Gf[1];
  double fwloc = o->p.wellph[0] * Gph[1];
  double Fx[MM_MAX_EL + 1] = {0}, Fy[MM_MAX_EL + 1] = {0}, Fz[MM_MAX_EL + 1] = {0};
  double fg[MM_MAX_VAR];
  ivn = 0;
  for (int ph = 2; ph <= NPH; ++ph) {
    ivn += o->p.nvar[ph - 2];
    int fvn = ivn + o->p.nvar[ph - 1] - 1;
    for (int iv = ivn; iv <= fvn; ++iv) {
      int v = iv - 1;
      double gx = o->gradx[v * n3 + pt], gy = o->grady[v * n3 + pt], gz = o->gradz[v * n3 + pt];
      double term = GC(o, v, 0, 0) * (gx * gx) + GC(o, v, 1, 1) * (gy * gy) +
                    GC(o, v, 2, 2) * (gz * gz) + (GC(o, v, 0, 1) + GC(o, v, 1, 0)) * gx * gy +
                    (GC(o, v, 0, 2) + GC(o, v, 2, 0)) * gx * gz +
                    (GC(o, v, 1, 2) + GC(o, v, 2, 1)) * gy * gz;
      *fint += 0.5 * term;
      double fi = phi[v];
      double termch = Gf[ph] - Gf[1];
      for (int e = 1; e <= E; ++e) {
        double xd = X[e][1] - X[e][ph];
        termch = termch + Mu[e] * xd;
        double tprc = Ppr[v] * xd;
        Fx[e] += tprc * gx;
        Fy[e] += tprc * gy;
        Fz[e] += tprc * gz;
      }
      double wg = 2.0 * o->p.wellph[ph - 1] * fi * (1.0 - fi) * (1.0 - 2.0 * fi);
      for (int jv = 0; jv < V; ++jv)
        if (jv != v) wg = wg + 4.0 * o->p.wellph[0] * fi * (phi[jv] * phi[jv]);
      fg[v] = termch * Ppr[v] + wg;
    }
    fchloc = fchloc + Pph[ph] * Gf[ph];
    fwloc = fwloc + o->p.wellph[ph - 1] * Gph[ph];
  }
  *fch += fchloc;
  *fw += fwloc;
  for (int e = 1; e <= E; ++e) {
    o->Fx[(e - 1) * n3 + pt] = Fx[e];
    o->Fy[(e - 1) * n3 + pt] = Fy[e];
    o->Fz[(e - 1) * n3 + pt] = Fz[e];
  }
  /* f_diff correction, :824-841 */
  for (int jv = 0; jv < V; ++jv) {
    double t11 = 0.0;
    int nt = 1;
    if (phi[jv] > 0.0) {
      for (int v = 0; v < V; ++v)
        if (phi[v] > 0.0 && v != jv) {
          t11 = t11 + fg[v] - o->kdel[v * n3 + pt];
          nt += 1;
        }
    }
    o->fgrad[jv * n3 + pt] = fg[jv];
    o->fdiff[jv * n3 + pt] = fg[jv] - t11 / (double)nt;
  }
  return 0;
}

static int chemistry(oracle_t *o, double *fint, double *fch, double *fw) {
  int nz = o->nz, nt = o->nthreads;
  long slab = (long)o->nx * o->ny;
  double *pf = calloc(3 * nz, sizeof(double));
  int err = 0;
#pragma omp parallel for num_threads(nt) schedule(static) reduction(| : err)
  for (int z = 0; z < nz; ++z) {
    double a = 0, b = 0, c = 0;
    for (long i = 0; i < slab; ++i) err |= chem_point(o, z * slab + i, &a, &b, &c);
    pf[3 * z] = a; pf[3 * z + 1] = b; pf[3 * z + 2] = c;
  }
  double a = 0, b = 0, c = 0;
  for (int z = 0; z < nz; ++z) { a += pf[3 * z]; b += pf[3 * z + 1]; c += pf[3 * z + 2]; }
  free(pf);
  *fint = a; *fch = b; *fw = c;
  return err ? MM_ERR_ZSOLVE : 0;
}

/* Voigt index of the tensor pair (i,j), 0-based in and out: ModElastic.f:124-126 */
static inline int voigt(int i, int j) { return (i == j) ? i : (9 - (i + 1) - (j + 1)) - 1; }

/* time-invariant acoustic-tensor inverse Qinv(k), ModElastic.f:115-191 */
static int build_qinv(oracle_t *o) {
  long nc = o->nc;
  int H = o->H, ny = o->ny;
  int err = 0;
#pragma omp parallel for num_threads(o->nthreads) schedule(static) reduction(| : err)
  for (long idx = 0; idx < nc; ++idx) {
    int kx = (int)(idx % H), ky = (int)((idx / H) % ny), kz = (int)(idx / ((long)H * ny));
    if (kx == 0 && ky == 0 && kz == 0) continue;
    double kf[3] = {o->kfx[kx], o->kfy[ky], o->kfz[kz]};
    double kq[3] = {o->kqx[kx], o->kqy[ky], o->kqz[kz]};
    double Q[9] = {0};
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) {
        int ij = voigt(i, j);
        for (int k = 0; k < 3; ++k)
          for (int l = 0; l < 3; ++l) {
            int kl = voigt(k, l);
            double kpr = (k == j) ? kq[k] : -(kf[k] * kf[j]);
            Q[i + 3 * l] += CE(o, ij, kl, 2) * kpr;
          }
      }
    if (oracle_invert(Q, 3, MM_ERR_QINV)) { err |= 1; continue; }
    memcpy(o->qinv + 9 * idx, Q, sizeof(Q));
  }
  if (err) return MM_ERR_QINV;
  o->qinv_ready = 1;
  return 0;
}

/* elener, ModElastic.f:493-691 */
static int elener(oracle_t *o) {
  const int V = o->V, NPH = o->NPH;
  long n3 = o->n3;
  double transsum[MM_MAX_VAR];
  for (int v = 0; v < V; ++v) {
    double s = 0.0;
    for (int m = 0; m < 6; ++m) s += TP(o, m, v);
    transsum[v] = s;
  }
  const double thr = (double)1.e-9f;
  int err = 0;
#pragma omp parallel for num_threads(o->nthreads) schedule(static) reduction(| : err)
  for (long pt = 0; pt < n3; ++pt) {
    double te[6] = {0}, ps = 0.0, Ppr[MM_MAX_VAR];
    for (int v = 0; v < V; ++v) {
      double fi = o->phi[v * n3 + pt], fi_sq = fi * fi;
      double pl = o->Pphi[v * n3 + pt];
      ps += pl;
      Ppr[v] = 30.0 * fi_sq * (fi_sq - 2.0 * fi + 1.0);
      for (int m = 0; m < 6; ++m) te[m] += pl * TP(o, m, v);
    }
    const double *g = o->gu; /* g[(j)*3+(i)] = d_j u_i (0-based) */
#define GU(i, j) g[((j) * 3 + (i)) * n3 + pt]
    double eps[6];
    eps[0] = o->p.e_mean[0] + GU(0, 0) - te[0];
    eps[1] = o->p.e_mean[1] + GU(1, 1) - te[1];
    eps[2] = o->p.e_mean[2] + GU(2, 2) - te[2];
    eps[3] = 2.0 * o->p.e_mean[3] + GU(2, 1) + GU(1, 2) - 2.0 * te[3];
    eps[4] = 2.0 * o->p.e_mean[4] + GU(2, 0) + GU(0, 2) - 2.0 * te[4];
    eps[5] = 2.0 * o->p.e_mean[5] + GU(0, 1) + GU(1, 0) - 2.0 * te[5];
#undef GU
    double C[36] = {0};
    int ivn = 0;
    for (int ph = 2; ph <= NPH; ++ph) {
      ivn += o->p.nvar[ph - 2];
      for (int iv = ivn; iv <= ivn + o->p.nvar[ph - 1] - 1; ++iv) {
        double pl = o->Pphi[(iv - 1) * n3 + pt];
        for (int q = 0; q < 36; ++q) C[q] = C[q] + pl * o->p.sela[(ph - 1) * 36 + q];
      }
    }
    for (int q = 0; q < 36; ++q) C[q] = C[q] + (1.0 - ps) * o->p.sela[q];
    if (oracle_invert(C, 6, MM_ERR_CINV_ENER)) { err |= 1; continue; }
    double sig[6];
    for (int a = 0; a < 6; ++a) {
      double s = 0.0;
      for (int m = 0; m < 6; ++m) s = s + C[a + 6 * m] * eps[m];
      sig[a] = s;
    }
    double el_mat = 0.0;
    for (int a = 0; a < 6; ++a) {
      double em = 0.0;
      for (int m = 0; m < 6; ++m) em = em + SE(o, a, m, 1) * sig[m];
      el_mat = el_mat + sig[a] * em;
    }
    el_mat = 0.5 * el_mat;
    double een = (1.0 - ps) * el_mat;
    ivn = 0;
    for (int ph = 2; ph <= NPH; ++ph) {
      ivn += o->p.nvar[ph - 2];
      for (int iv = ivn; iv <= ivn + o->p.nvar[ph - 1] - 1; ++iv) {
        int v = iv - 1;
        if (transsum[v] > thr) {
          double el_ppt = 0.0;
          for (int a = 0; a < 6; ++a) {
            double ep = 0.0;
            for (int m = 0; m < 6; ++m) ep = ep + SE(o, a, m, ph) * sig[m];
            el_ppt = el_ppt + sig[a] * ep;
          }
          el_ppt = 0.5 * el_ppt;
          een = een + el_ppt * o->Pphi[v * n3 + pt];
          double eg = (sig[0] * TP(o, 0, v) + sig[1] * TP(o, 1, v) + sig[2] * TP(o, 2, v) +
                       2.0 * (sig[3] * TP(o, 3, v) + sig[4] * TP(o, 4, v) + sig[5] * TP(o, 5, v)));
          o->elgrad[v * n3 + pt] = Ppr[v] * (el_mat - el_ppt) - eg * Ppr[v];
        }
      }
    }
    o->el_en[pt] = een;
  }
  return err ? MM_ERR_CINV_ENER : 0;
}

/* ------------------------------------------------- tangent-linear step --
 * Forward-mode derivative of oracle_step along a direction of the state
 * (phi, c, grad u) and of the parameter arrays grad_coeff, wellph, cmobs,
 * b_mob and trans_ppt (mm_adj_grads layout).  The tangent runs in lockstep
 * with the primal and takes the primal's decisions: the phi > 0 tests of the
 * f_diff correction (Multi_MEUMAPPS.f:824-841), the clamp (:1034-1035), the
 * elastic iteration count and its cfac sequence (ModElastic.f:433-488).  So
 * it is the derivative of the map with those decisions frozen.  Each block is
 * the differential of the primal block above it, written from the primal's
 * formulas; the compliance mixtures are differentiated through their inverses
 * (dC = -C dM C), so inhomogeneous moduli are covered too.  Regular steps
 * only (no noise, step != elastep0). */
struct tlm {
  double *dphi, *dcons, *dgu;                        /* tangent state       */
  double *dgx, *dgy, *dgz, *dkd, *dfd, *dPp, *dFx, *dFy, *dFz, *dxf;
  cplx *dsphi, *sw;                                  /* V + 12 spectra      */
  mm_adj_grads dp;                                   /* parameter direction */
};

/* sum dG_ij k_i k_j with the direction of grad_coeff (the 2-form of kgk) */
static inline double dkgk(const oracle_t *o, int v, int kx, int ky, int kz) {
  const double *G = o->t->dp.grad_coeff + v * 9; /* G[j*3+i] = (i,j) */
  double k1q = o->kqx[kx], k2q = o->kqy[ky], k3q = o->kqz[kz];
  double a1 = o->kfx[kx], a2 = o->kfy[ky], a3 = o->kfz[kz];
  return G[0] * k1q + G[4] * k2q + G[8] * k3q + (G[3] + G[1]) * (-(a1 * a2)) +
         (G[7] + G[5]) * (-(a2 * a3)) + (G[2] + G[6]) * (-(a3 * a1));
}

/* d(num / den) = dnum / den - num dden / den^2 */
static inline cplx dquot(cplx num, cplx dnum, double den, double dden) {
  return dnum / den - num * (dden / (den * den));
}

/* gradient block: dg = F^-1(i k dphi^), dkd = F^-1(kgk dphi^ + dkgk phi^) */
static void tlm_gradients(oracle_t *o) {
  struct tlm *t = o->t;
  long n3 = o->n3, nc = o->nc;
  int H = o->H, ny = o->ny, nt = o->nthreads;
  for (int v = 0; v < o->V; ++v) {
    cplx *dS = t->dsphi + v * nc;
    const cplx *S = o->sphi + v * nc;
    f_trans(o, t->dphi + v * n3, dS);
    cplx *gx = t->sw, *gy = t->sw + nc, *gz = t->sw + 2 * nc, *kd = t->sw + 3 * nc;
#pragma omp parallel for num_threads(nt) schedule(static)
    for (long i = 0; i < nc; ++i) {
      int kx = (int)(i % H), ky = (int)((i / H) % ny), kz = (int)(i / ((long)H * ny));
      gx[i] = mul_i(dS[i], o->kfx[kx]);
      gy[i] = mul_i(dS[i], o->kfy[ky]);
      gz[i] = mul_i(dS[i], o->kfz[kz]);
      kd[i] = dS[i] * kgk(o, v, kx, ky, kz) + S[i] * dkgk(o, v, kx, ky, kz);
    }
    inv_trans(o, gx, t->dgx + v * n3);
    inv_trans(o, gy, t->dgy + v * n3);
    inv_trans(o, gz, t->dgz + v * n3);
    inv_trans(o, kd, t->dkd + v * n3);
  }
}

/* chemistry at one point: dP_phi, the flux tangents and d f_diff (before the
 * elastic term).  The primal quantities are recomputed here from the
 * unchanged phi, c, grad phi and kdel. */
static void tlm_chem_point(oracle_t *o, long pt) {
  struct tlm *t = o->t;
  const int V = o->V, E = o->E, NPH = o->NPH, zdim = o->zdim;
  long n3 = o->n3;
  double phi[MM_MAX_VAR], dph[MM_MAX_VAR], Ppr[MM_MAX_VAR], dPpr[MM_MAX_VAR];
  double Pph[MM_MAX_PH + 1] = {0}, dPph[MM_MAX_PH + 1] = {0};
  for (int v = 0; v < V; ++v) {
    double fi = o->phi[v * n3 + pt], df = t->dphi[v * n3 + pt], fi_sq = fi * fi;
    phi[v] = fi;
    dph[v] = df;
    double pl = fi_sq * fi * (6.0 * fi_sq - 15.0 * fi + 10.0);
    Ppr[v] = 30.0 * fi_sq * (fi_sq - 2.0 * fi + 1.0);
    dPpr[v] = 60.0 * fi * (1.0 - fi) * (1.0 - 2.0 * fi) * df;
    double dpl = Ppr[v] * df;
    t->dPp[v * n3 + pt] = dpl;
    int ph = o->phase_of[v];
    Pph[1] += pl; Pph[ph] += pl;
    dPph[1] += dpl; dPph[ph] += dpl;
  }
  Pph[1] = 1.0 - Pph[1];
  dPph[1] = -dPph[1];
  /* Zintf and its right-hand side as in chem_point; solve Z s = r */
  double Z[16 * 16], Zr[16];
  memset(Z, 0, sizeof(Z));
  for (int e = 1; e <= E; ++e) {
    int i1 = e + E + 1;
    for (int ph = 1; ph <= NPH - 1; ++ph) {
      int i0 = (NPH - 1) * (e - 1) + ph, j0 = NPH * (e - 1) + ph;
      Z[(i0 - 1) + (j0 - 1) * zdim] = GP(o, e, ph);
      Z[(i0 - 1) + j0 * zdim] = -GP(o, e, ph + 1);
      Zr[i0 - 1] = GP(o, e, ph) * GP(o, i1, ph) - GP(o, e, ph + 1) * GP(o, i1, ph + 1);
    }
  }
  for (int e = 1; e <= E; ++e) {
    int i0 = zdim - E + e;
    Zr[i0 - 1] = o->cons[(e - 1) * n3 + pt];
    for (int ph = 1; ph <= NPH; ++ph) Z[(i0 - 1) + (NPH * (e - 1) + ph - 1) * zdim] = Pph[ph];
  }
  if (oracle_invert(Z, zdim, MM_ERR_ZSOLVE)) return; /* the primal reports it */
  double sol[16], rhs[16] = {0}, dsol[16];
  for (int i = 0; i < zdim; ++i) {
    double a = 0.0;
    for (int j = 0; j < zdim; ++j) a += Z[i + j * zdim] * Zr[j];
    sol[i] = a;
  }
  /* d s = Z^-1 (dr - dZ s): dZ lives in the mixture rows, dr in c */
  for (int e = 1; e <= E; ++e) {
    int i0 = zdim - E + e;
    double r = t->dcons[(e - 1) * n3 + pt];
    for (int ph = 1; ph <= NPH; ++ph) r -= dPph[ph] * sol[NPH * (e - 1) + ph - 1];
    rhs[i0 - 1] = r;
  }
  for (int i = 0; i < zdim; ++i) {
    double a = 0.0;
    for (int j = 0; j < zdim; ++j) a += Z[i + j * zdim] * rhs[j];
    dsol[i] = a;
  }
  double X[MM_MAX_EL + 1][MM_MAX_PH + 1], dX[MM_MAX_EL + 1][MM_MAX_PH + 1];
  for (int e = 1; e <= E; ++e)
    for (int ph = 1; ph <= NPH; ++ph) {
      X[e][ph] = sol[NPH * (e - 1) + ph - 1];
      dX[e][ph] = dsol[NPH * (e - 1) + ph - 1];
    }
  double Gf[MM_MAX_PH + 1], dGf[MM_MAX_PH + 1], Mu[MM_MAX_EL + 1], dMu[MM_MAX_EL + 1];
  for (int ph = 1; ph <= NPH; ++ph) {
    double g = 0.0, dg = 0.0;
    for (int e = 1; e <= E; ++e) {
      double d = X[e][ph] - GP(o, e + E + 1, ph);
      g += GP(o, e, ph) * d * d;
      dg += 2.0 * GP(o, e, ph) * d * dX[e][ph];
    }
    Gf[ph] = g + GP(o, E + 1, ph);
    dGf[ph] = dg;
  }
  for (int e = 1; e <= E; ++e) {
    Mu[e] = 2.0 * GP(o, e, 1) * (X[e][1] - GP(o, e + E + 1, 1));
    dMu[e] = 2.0 * GP(o, e, 1) * dX[e][1];
  }
  const double *w = o->p.wellph, *dw = t->dp.wellph;
  double dF[3][MM_MAX_EL + 1], dfg[MM_MAX_VAR];
  memset(dF, 0, sizeof(dF));
  for (int v = 0; v < V; ++v) {
    int ph = o->phase_of[v];
    double g[3] = {o->gradx[v * n3 + pt], o->grady[v * n3 + pt], o->gradz[v * n3 + pt]};
    double dg[3] = {t->dgx[v * n3 + pt], t->dgy[v * n3 + pt], t->dgz[v * n3 + pt]};
    double fi = phi[v], df = dph[v];
    double termch = Gf[ph] - Gf[1], dtermch = dGf[ph] - dGf[1];
    for (int e = 1; e <= E; ++e) {
      double xd = X[e][1] - X[e][ph], dxd = dX[e][1] - dX[e][ph];
      termch += Mu[e] * xd;
      dtermch += dMu[e] * xd + Mu[e] * dxd;
      for (int d = 0; d < 3; ++d)
        dF[d][e] += dPpr[v] * xd * g[d] + Ppr[v] * dxd * g[d] + Ppr[v] * xd * dg[d];
    }
    double h = fi * (1.0 - fi) * (1.0 - 2.0 * fi), dh = (1.0 - 6.0 * fi + 6.0 * fi * fi) * df;
    double dwg = 2.0 * dw[ph - 1] * h + 2.0 * w[ph - 1] * dh;
    for (int jv = 0; jv < V; ++jv)
      if (jv != v) {
        double pj = phi[jv], dpj = dph[jv];
        dwg += 4.0 * (dw[0] * fi * pj * pj + w[0] * df * pj * pj + w[0] * fi * 2.0 * pj * dpj);
      }
    dfg[v] = dtermch * Ppr[v] + termch * dPpr[v] + dwg;
  }
  for (int e = 1; e <= E; ++e) {
    t->dFx[(e - 1) * n3 + pt] = dF[0][e];
    t->dFy[(e - 1) * n3 + pt] = dF[1][e];
    t->dFz[(e - 1) * n3 + pt] = dF[2][e];
  }
  /* the correction with the primal's phi > 0 decisions */
  for (int jv = 0; jv < V; ++jv) {
    double dt11 = 0.0;
    int nt = 1;
    if (phi[jv] > 0.0)
      for (int v = 0; v < V; ++v)
        if (phi[v] > 0.0 && v != jv) {
          dt11 += dfg[v] - t->dkd[v * n3 + pt];
          nt += 1;
        }
    t->dfd[jv * n3 + pt] = dfg[jv] - dt11 / (double)nt;
  }
}

static void tlm_chemistry(oracle_t *o) {
#pragma omp parallel for num_threads(o->nthreads) schedule(static)
  for (long pt = 0; pt < o->n3; ++pt) tlm_chem_point(o, pt);
}

/* the compliance mixture at one point, its inverse C and dC = -C dM C.
 * guard: the (1 - sum P) matrix term only while sum P < 1 (eldis), or always
 * (elener). */
static int tlm_moduli(const oracle_t *o, long pt, int guard, double *C, double *dC) {
  const struct tlm *t = o->t;
  long n3 = o->n3;
  double M[36] = {0}, dM[36] = {0}, ps = 0.0, dps = 0.0;
  for (int v = 0; v < o->V; ++v) {
    double pl = o->Pphi[v * n3 + pt], dpl = t->dPp[v * n3 + pt];
    const double *S = o->p.sela + (o->phase_of[v] - 1) * 36;
    ps += pl;
    dps += dpl;
    for (int q = 0; q < 36; ++q) { M[q] += pl * S[q]; dM[q] += dpl * S[q]; }
  }
  if (!guard || ps < 1.0)
    for (int q = 0; q < 36; ++q) { M[q] += (1.0 - ps) * o->p.sela[q]; dM[q] -= dps * o->p.sela[q]; }
  memcpy(C, M, sizeof(M));
  if (oracle_invert(C, 6, MM_ERR_CINV_ITER)) return 1;
  double T[36];
  for (int a = 0; a < 6; ++a)
    for (int b = 0; b < 6; ++b) {
      double s = 0.0;
      for (int c = 0; c < 6; ++c) s += dM[a + 6 * c] * C[c + 6 * b];
      T[a + 6 * b] = s;
    }
  for (int a = 0; a < 6; ++a)
    for (int b = 0; b < 6; ++b) {
      double s = 0.0;
      for (int c = 0; c < 6; ++c) s += C[a + 6 * c] * T[c + 6 * b];
      dC[a + 6 * b] = -s;
    }
  return 0;
}

/* one elastic iteration: the polarization stress tangent from the
 * pre-iteration grad u, then its k-space map to the new d grad u */
static void tlm_elastic_iter(oracle_t *o, double cfac) {
  struct tlm *t = o->t;
  const int V = o->V, H = o->H, ny = o->ny, nt = o->nthreads;
  long n3 = o->n3, nc = o->nc;
  const int vmap[3][3] = {{0, 5, 4}, {5, 1, 3}, {4, 3, 2}};
#pragma omp parallel for num_threads(nt) schedule(static)
  for (long pt = 0; pt < n3; ++pt) {
    double C[36], dC[36];
    if (tlm_moduli(o, pt, 1, C, dC)) continue;
    double a[6] = {0}, da[6] = {0}; /* sum_v P_v eps0_v and its tangent */
    for (int kl = 0; kl < 6; ++kl)
      for (int v = 0; v < V; ++v) {
        a[kl] += o->Pphi[v * n3 + pt] * TP(o, kl, v);
        da[kl] += t->dPp[v * n3 + pt] * TP(o, kl, v) + o->Pphi[v * n3 + pt] * t->dp.trans_ppt[v * 6 + kl];
      }
    for (int i = 0; i < 3; ++i)
      for (int j = i; j < 3; ++j) {
        int ij = vmap[i][j];
        double s = 0.0;
        for (int k = 0; k < 3; ++k)
          for (int l = 0; l < 3; ++l) {
            int kl = vmap[k][l];
            double b = o->gu[(l * 3 + k) * n3 + pt] + o->p.e_mean[kl];
            double db = t->dgu[(l * 3 + k) * n3 + pt];
            double Cp = o->p.cela[36 + ij + 6 * kl] - C[ij + 6 * kl];
            s += cfac * (-dC[ij + 6 * kl] * b + Cp * db) + dC[ij + 6 * kl] * a[kl] + C[ij + 6 * kl] * da[kl];
          }
        t->dxf[ij * n3 + pt] = s;
      }
  }
  cplx *sxx = (cplx *)malloc(sizeof(cplx) * 6 * nc);
  for (int c = 0; c < 6; ++c) f_trans(o, t->dxf + c * n3, sxx + c * nc);
  cplx *sg = t->sw; /* 9 spectra of d grad u */
#pragma omp parallel for num_threads(nt) schedule(static)
  for (long idx = 0; idx < nc; ++idx) {
    int kx = (int)(idx % H), ky = (int)((idx / H) % ny), kz = (int)(idx / ((long)H * ny));
    double kf[3] = {o->kfx[kx], o->kfy[ky], o->kfz[kz]};
    cplx uh[3] = {0, 0, 0};
    if (kx || ky || kz) {
      cplx R[3] = {0, 0, 0};
      for (int i = 0; i < 3; ++i)
        for (int d = 0; d < 3; ++d) R[i] += mul_i(sxx[vmap[i][d] * nc + idx], kf[d]);
      const double *Q = o->qinv + 9 * idx;
      for (int i = 0; i < 3; ++i) uh[i] = Q[i] * R[0] + Q[i + 3] * R[1] + Q[i + 6] * R[2];
    }
    for (int d = 0; d < 3; ++d)
      for (int i = 0; i < 3; ++i) sg[(d * 3 + i) * nc + idx] = mul_i(uh[i], kf[d]);
  }
  free(sxx);
  for (int c = 0; c < 9; ++c) inv_trans(o, sg + c * nc, t->dgu + c * n3);
}

/* elener: d el_grad added to d f_diff (el_en is not carried) */
static void tlm_elener(oracle_t *o) {
  struct tlm *t = o->t;
  const int V = o->V;
  long n3 = o->n3;
  const double thr = (double)1.e-9f;
  const double wv[6] = {1, 1, 1, 2, 2, 2};
#pragma omp parallel for num_threads(o->nthreads) schedule(static)
  for (long pt = 0; pt < n3; ++pt) {
    double C[36], dC[36];
    if (tlm_moduli(o, pt, 0, C, dC)) continue;
    double te[6] = {0}, dte[6] = {0};
    for (int v = 0; v < V; ++v)
      for (int m = 0; m < 6; ++m) {
        te[m] += o->Pphi[v * n3 + pt] * TP(o, m, v);
        dte[m] += t->dPp[v * n3 + pt] * TP(o, m, v) + o->Pphi[v * n3 + pt] * t->dp.trans_ppt[v * 6 + m];
      }
    const double *g = o->gu, *dg = t->dgu;
#define GU(A, i, j) A[((j) * 3 + (i)) * n3 + pt]
    double eps[6] = {o->p.e_mean[0] + GU(g, 0, 0) - te[0], o->p.e_mean[1] + GU(g, 1, 1) - te[1],
                     o->p.e_mean[2] + GU(g, 2, 2) - te[2],
                     2.0 * o->p.e_mean[3] + GU(g, 2, 1) + GU(g, 1, 2) - 2.0 * te[3],
                     2.0 * o->p.e_mean[4] + GU(g, 2, 0) + GU(g, 0, 2) - 2.0 * te[4],
                     2.0 * o->p.e_mean[5] + GU(g, 0, 1) + GU(g, 1, 0) - 2.0 * te[5]};
    double deps[6] = {GU(dg, 0, 0) - dte[0], GU(dg, 1, 1) - dte[1], GU(dg, 2, 2) - dte[2],
                      GU(dg, 2, 1) + GU(dg, 1, 2) - 2.0 * dte[3], GU(dg, 2, 0) + GU(dg, 0, 2) - 2.0 * dte[4],
                      GU(dg, 0, 1) + GU(dg, 1, 0) - 2.0 * dte[5]};
#undef GU
    double sig[6], dsig[6];
    for (int a = 0; a < 6; ++a) {
      double s = 0.0, ds = 0.0;
      for (int m = 0; m < 6; ++m) {
        s += C[a + 6 * m] * eps[m];
        ds += dC[a + 6 * m] * eps[m] + C[a + 6 * m] * deps[m];
      }
      sig[a] = s;
      dsig[a] = ds;
    }
    /* 0.5 sig.S sig and its tangent for the compliance of phase ph */
#define QUAD(ph, q, dq)                                                                   \
  do {                                                                                    \
    double q_ = 0.0, dq_ = 0.0;                                                           \
    for (int a = 0; a < 6; ++a)                                                           \
      for (int m = 0; m < 6; ++m) {                                                       \
        double s_ = SE(o, a, m, ph);                                                      \
        q_ += sig[a] * s_ * sig[m];                                                       \
        dq_ += dsig[a] * s_ * sig[m] + sig[a] * s_ * dsig[m];                             \
      }                                                                                   \
    q = 0.5 * q_;                                                                         \
    dq = 0.5 * dq_;                                                                       \
  } while (0)
    double el_mat, del_mat;
    QUAD(1, el_mat, del_mat);
    for (int v = 0; v < V; ++v) {
      double ts = 0.0;
      for (int m = 0; m < 6; ++m) ts += TP(o, m, v);
      if (!(ts > thr)) continue;
      double el_ppt, del_ppt;
      QUAD(o->phase_of[v], el_ppt, del_ppt);
      double eg = 0.0, deg = 0.0;
      for (int m = 0; m < 6; ++m) {
        eg += wv[m] * sig[m] * TP(o, m, v);
        deg += wv[m] * (dsig[m] * TP(o, m, v) + sig[m] * t->dp.trans_ppt[v * 6 + m]);
      }
      double fi = o->phi[v * n3 + pt], df = t->dphi[v * n3 + pt];
      double pp = 30.0 * fi * fi * (fi * fi - 2.0 * fi + 1.0);
      double dpp = 60.0 * fi * (1.0 - fi) * (1.0 - 2.0 * fi) * df;
      t->dfd[v * n3 + pt] += dpp * (el_mat - el_ppt - eg) + pp * (del_mat - del_ppt - deg);
    }
#undef QUAD
  }
}

/* Cahn-Hilliard: c^' = (c^ + dtm i k.F^) / (1 - dtm k^2), dtm = dt M_e */
static void tlm_ch(oracle_t *o) {
  struct tlm *t = o->t;
  const int H = o->H, ny = o->ny, nt = o->nthreads;
  long n3 = o->n3, nc = o->nc;
  cplx *s = t->sw; /* c, Fx, Fy, Fz, then their tangents */
  for (int e = 0; e < o->E; ++e) {
    f_trans(o, o->cons + e * n3, s);
    f_trans(o, o->Fx + e * n3, s + nc);
    f_trans(o, o->Fy + e * n3, s + 2 * nc);
    f_trans(o, o->Fz + e * n3, s + 3 * nc);
    f_trans(o, t->dcons + e * n3, s + 4 * nc);
    f_trans(o, t->dFx + e * n3, s + 5 * nc);
    f_trans(o, t->dFy + e * n3, s + 6 * nc);
    f_trans(o, t->dFz + e * n3, s + 7 * nc);
    double dtm = o->p.t_step * o->p.cmobs[e], ddtm = o->p.t_step * t->dp.cmobs[e];
#pragma omp parallel for num_threads(nt) schedule(static)
    for (long idx = 0; idx < nc; ++idx) {
      int kx = (int)(idx % H), ky = (int)((idx / H) % ny), kz = (int)(idx / ((long)H * ny));
      double ksq = o->kqx[kx] + o->kqy[ky] + o->kqz[kz];
      cplx t1 = mul_i(s[nc + idx], o->kfx[kx]) + mul_i(s[2 * nc + idx], o->kfy[ky]) +
                mul_i(s[3 * nc + idx], o->kfz[kz]);
      cplx dt1 = mul_i(s[5 * nc + idx], o->kfx[kx]) + mul_i(s[6 * nc + idx], o->kfy[ky]) +
                 mul_i(s[7 * nc + idx], o->kfz[kz]);
      cplx num = s[idx] + dtm * t1, dnum = s[4 * nc + idx] + ddtm * t1 + dtm * dt1;
      s[4 * nc + idx] = dquot(num, dnum, 1.0 - dtm * ksq, -ddtm * ksq);
    }
    inv_trans(o, s + 4 * nc, t->dcons + e * n3);
  }
}

/* Allen-Cahn: phi^' = (phi^ - dtl fd^) / (1 - dtl kgk), dtl = dt L; the
 * tangent of the pre-clamp field */
static void tlm_ac(oracle_t *o) {
  struct tlm *t = o->t;
  const int H = o->H, ny = o->ny, nt = o->nthreads;
  long n3 = o->n3, nc = o->nc;
  cplx *sf = t->sw, *dsf = t->sw + nc, *out = t->sw + 2 * nc;
  double dtl = o->p.t_step * o->p.b_mob, ddtl = o->p.t_step * t->dp.b_mob;
  for (int v = 0; v < o->V; ++v) {
    const cplx *S = o->sphi + v * nc, *dS = t->dsphi + v * nc;
    f_trans(o, o->fdiff + v * n3, sf);
    f_trans(o, t->dfd + v * n3, dsf);
#pragma omp parallel for num_threads(nt) schedule(static)
    for (long idx = 0; idx < nc; ++idx) {
      int kx = (int)(idx % H), ky = (int)((idx / H) % ny), kz = (int)(idx / ((long)H * ny));
      double kg = kgk(o, v, kx, ky, kz), dkg = dkgk(o, v, kx, ky, kz);
      cplx num = S[idx] - dtl * sf[idx], dnum = dS[idx] - ddtl * sf[idx] - dtl * dsf[idx];
      out[idx] = dquot(num, dnum, 1.0 - dtl * kg, -(ddtl * kg + dtl * dkg));
    }
    inv_trans(o, out, t->dphi + v * n3);
  }
}

/* clamp: the tangent passes where 0 <= phi' <= 1 (primal pre-clamp field) */
static void tlm_clamp(oracle_t *o) {
  long n = (long)o->V * o->n3;
  for (long i = 0; i < n; ++i)
    if (o->phi[i] < 0.0 || o->phi[i] > 1.0) o->t->dphi[i] = 0.0;
}

void oracle_tlm_off(oracle_t *o) {
  struct tlm *t = o->t;
  if (!t) return;
  double *b[] = {t->dphi, t->dcons, t->dgu, t->dgx, t->dgy, t->dgz, t->dkd, t->dfd,
                 t->dPp, t->dFx, t->dFy, t->dFz, t->dxf};
  for (size_t i = 0; i < sizeof(b) / sizeof(b[0]); ++i) free(b[i]);
  free(t->dsphi); free(t->sw);
  free(t);
  o->t = NULL;
}

int oracle_tlm_set(oracle_t *o, const double *dphi, const double *dcons, const double *dgradu,
                   const mm_adj_grads *dpar) {
  long n3 = o->n3, V = o->V, E = o->E;
  if (!o->t) {
    struct tlm *t = (struct tlm *)calloc(1, sizeof(struct tlm));
    if (!t) return MM_ERR_NOMEM;
    o->t = t;
    t->dphi = xcalloc(V * n3, 8); t->dcons = xcalloc(E * n3, 8); t->dgu = xcalloc(9 * n3, 8);
    t->dgx = xcalloc(V * n3, 8); t->dgy = xcalloc(V * n3, 8); t->dgz = xcalloc(V * n3, 8);
    t->dkd = xcalloc(V * n3, 8); t->dfd = xcalloc(V * n3, 8); t->dPp = xcalloc(V * n3, 8);
    t->dFx = xcalloc(E * n3, 8); t->dFy = xcalloc(E * n3, 8); t->dFz = xcalloc(E * n3, 8);
    t->dxf = xcalloc(6 * n3, 8);
    t->dsphi = xcalloc(V * o->nc, sizeof(cplx));
    t->sw = xcalloc(12 * o->nc, sizeof(cplx));
    if (!t->dphi || !t->dgu || !t->dxf || !t->dsphi || !t->sw) { oracle_tlm_off(o); return MM_ERR_NOMEM; }
  }
  struct tlm *t = o->t;
  if (dphi) memcpy(t->dphi, dphi, 8 * V * n3); else memset(t->dphi, 0, 8 * V * n3);
  if (dcons) memcpy(t->dcons, dcons, 8 * E * n3); else memset(t->dcons, 0, 8 * E * n3);
  if (dgradu) memcpy(t->dgu, dgradu, 8 * 9 * n3); else memset(t->dgu, 0, 8 * 9 * n3);
  if (dpar) t->dp = *dpar; else memset(&t->dp, 0, sizeof(t->dp));
  return 0;
}

int oracle_tlm_get(oracle_t *o, double *dphi, double *dcons, double *dgradu) {
  struct tlm *t = o->t;
  long n3 = o->n3;
  if (!t) return MM_ERR_ARG;
  if (dphi) memcpy(dphi, t->dphi, 8 * o->V * n3);
  if (dcons) memcpy(dcons, t->dcons, 8 * o->E * n3);
  if (dgradu) memcpy(dgradu, t->dgu, 8 * 9 * n3);
  return 0;
}

/* slab-ordered global sums of the six convergence terms, ModElastic.f:413-430 */
static void err_sums(oracle_t *o, double err[6]) {
  int nz = o->nz;
  long slab = (long)o->nx * o->ny, n3 = o->n3;
  double *pf = calloc(6 * nz, sizeof(double));
#pragma omp parallel for num_threads(o->nthreads) schedule(static)
  for (int z = 0; z < nz; ++z) {
    double e[6] = {0};
    for (long i = z * slab; i < (z + 1) * slab; ++i)
      for (int c = 0; c < 3; ++c) {
        double un = o->u[c * n3 + i], uo = o->u_old[c * n3 + i];
        e[2 * c] += (un - uo) * (un - uo);
        e[2 * c + 1] += uo * uo;
      }
    memcpy(pf + 6 * z, e, sizeof(e));
  }
  for (int c = 0; c < 6; ++c) err[c] = 0.0;
  for (int z = 0; z < nz; ++z)
    for (int c = 0; c < 6; ++c) err[c] += pf[6 * z + c];
  free(pf);
}

/* zeroth-order displacement at step == elastep0 from o->Pphi: FFT of P_phi,
 * R with the real*8 truncation of sum_0 (ModElastic.f:29,151), u = Qinv R,
 * u(k=0) = 0, u and its nine gradients to real space (:105-236) */
static void eldis_zeroth(oracle_t *o) {
  const int V = o->V, NPH = o->NPH, H = o->H, ny = o->ny, nt = o->nthreads;
  long n3 = o->n3, nc = o->nc;
  cplx *su = o->sw;           /* 3 spectra of u            */
  cplx *sg = o->sw + 3 * nc;  /* 9 spectra of grad u       */
  cplx *sp = (cplx *)malloc(sizeof(cplx) * V * nc);
  for (int v = 0; v < V; ++v) f_trans(o, o->Pphi + v * n3, sp + v * nc);
#pragma omp parallel for num_threads(nt) schedule(static)
  for (long idx = 0; idx < nc; ++idx) {
    int kx = (int)(idx % H), ky = (int)((idx / H) % ny), kz = (int)(idx / ((long)H * ny));
    double kf[3] = {o->kfx[kx], o->kfy[ky], o->kfz[kz]};
    if (kx == 0 && ky == 0 && kz == 0) {
      for (int c = 0; c < 3; ++c) su[c * nc + idx] = 0.0;
      for (int c = 0; c < 9; ++c) sg[c * nc + idx] = 0.0;
      continue;
    }
    double Ri[3] = {0, 0, 0}; /* R(i) = i * Ri[i] */
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) {
        int ij = voigt(i, j);
        double sum0 = 0.0;
        for (int k = 0; k < 3; ++k)
          for (int l = 0; l < 3; ++l) {
            int kl = voigt(k, l);
            int ivn = 0;
            for (int ph = 2; ph <= NPH; ++ph) {
              ivn += o->p.nvar[ph - 2];
              for (int iv = ivn; iv <= ivn + o->p.nvar[ph - 1] - 1; ++iv)
                sum0 = sum0 + creal(sp[(iv - 1) * nc + idx]) * TP(o, kl, iv - 1) * CE(o, ij, kl, ph);
            }
          }
        Ri[i] = Ri[i] + sum0 * kf[j];
      }
    const double *Q = o->qinv + 9 * idx;
    for (int i = 0; i < 3; ++i) {
      /* Q(i,1)*R(1) + Q(i,2)*R(2) + Q(i,3)*R(3), R purely imaginary */
      double ui = Q[i] * Ri[0] + Q[i + 3] * Ri[1] + Q[i + 6] * Ri[2];
      cplx uh = I * ui;
      su[i * nc + idx] = uh;
      for (int d = 0; d < 3; ++d) sg[(d * 3 + i) * nc + idx] = mul_i(uh, kf[d]);
    }
  }
  free(sp);
  for (int c = 0; c < 3; ++c) inv_trans(o, su + c * nc, o->u + c * n3);
  for (int c = 0; c < 9; ++c) inv_trans(o, sg + c * nc, o->gu + c * n3);
}

/* eldis, ModElastic.f:11-489 (+ elener) */
int oracle_eldis(oracle_t *o, int *n_iter, double *cfac_out) {
  const int V = o->V, NPH = o->NPH, H = o->H, ny = o->ny, nt = o->nthreads;
  long n3 = o->n3, nc = o->nc;
  int rc;
  if (!o->qinv_ready && (rc = build_qinv(o))) return rc;
  if (o->step == o->p.elastep0) eldis_zeroth(o);
  double newfac = 0.5, cfac = 1.0, err[6], err_old[6] = {0};
  int niter = 0, converged = 0, total = 0;
  const int vmap[3][3] = {{0, 5, 4}, {5, 1, 3}, {4, 3, 2}};
  for (;;) {
    /* label 2 */
    memcpy(o->u_old, o->u, sizeof(double) * 3 * n3);
    niter += 1;
    total += 1;
    int ierr = 0;
#pragma omp parallel for num_threads(nt) schedule(static) reduction(| : ierr)
    for (long pt = 0; pt < n3; ++pt) {
      double C[36] = {0}, ps = 0.0;
      int ivn = 0;
      for (int ph = 2; ph <= NPH; ++ph) {
        ivn += o->p.nvar[ph - 2];
        for (int iv = ivn; iv <= ivn + o->p.nvar[ph - 1] - 1; ++iv) {
          double pl = o->Pphi[(iv - 1) * n3 + pt];
          ps = ps + pl;
          for (int q = 0; q < 36; ++q) C[q] = C[q] + pl * o->p.sela[(ph - 1) * 36 + q];
        }
      }
      if (ps < 1.0)
        for (int q = 0; q < 36; ++q) C[q] = C[q] + (1.0 - ps) * o->p.sela[q];
      if (oracle_invert(C, 6, MM_ERR_CINV_ITER)) { ierr |= 1; continue; }
      double Cp[36];
      for (int q = 0; q < 36; ++q) Cp[q] = o->p.cela[36 + q] - C[q];
      double gradu[3][3];
      for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) gradu[i][j] = o->gu[(j * 3 + i) * n3 + pt];
      for (int i = 0; i < 3; ++i)
        for (int j = i; j < 3; ++j) {
          int ij = vmap[i][j];
          double s = 0.0;
          for (int k = 0; k < 3; ++k)
            for (int l = 0; l < 3; ++l) {
              int kl = vmap[k][l];
              s = s + cfac * Cp[ij + 6 * kl] * (gradu[k][l] + o->p.e_mean[kl]);
              for (int v = 0; v < V; ++v)
                s = s + C[ij + 6 * kl] * o->Pphi[v * n3 + pt] * TP(o, kl, v);
            }
          o->xf[ij * n3 + pt] = s;
        }
    }
    if (ierr) return MM_ERR_CINV_ITER;
    if (o->t) tlm_elastic_iter(o, cfac);
    cplx *sx = o->sw; /* 6 spectra; then reused for u (3) and grad u (9) */
    cplx *sxx = (cplx *)malloc(sizeof(cplx) * 6 * nc);
    for (int c = 0; c < 6; ++c) f_trans(o, o->xf + c * n3, sxx + c * nc);
#pragma omp parallel for num_threads(nt) schedule(static)
    for (long idx = 0; idx < nc; ++idx) {
      int kx = (int)(idx % H), ky = (int)((idx / H) % ny), kz = (int)(idx / ((long)H * ny));
      double kf[3] = {o->kfx[kx], o->kfy[ky], o->kfz[kz]};
      cplx x1 = sxx[idx], x2 = sxx[nc + idx], x3 = sxx[2 * nc + idx], x4 = sxx[3 * nc + idx],
           x5 = sxx[4 * nc + idx], x6 = sxx[5 * nc + idx];
      cplx R[3];
      R[0] = mul_i(x1, kf[0]) + mul_i(x6, kf[1]) + mul_i(x5, kf[2]);
      R[1] = mul_i(x6, kf[0]) + mul_i(x2, kf[1]) + mul_i(x4, kf[2]);
      R[2] = mul_i(x5, kf[0]) + mul_i(x4, kf[1]) + mul_i(x3, kf[2]);
      cplx uh[3];
      if (kx == 0 && ky == 0 && kz == 0) {
        uh[0] = uh[1] = uh[2] = 0.0;
      } else {
        const double *Q = o->qinv + 9 * idx;
        for (int i = 0; i < 3; ++i) uh[i] = Q[i] * R[0] + Q[i + 3] * R[1] + Q[i + 6] * R[2];
      }
      for (int i = 0; i < 3; ++i) {
        sx[i * nc + idx] = uh[i];
        for (int d = 0; d < 3; ++d) sx[(3 + d * 3 + i) * nc + idx] = mul_i(uh[i], kf[d]);
      }
    }
    free(sxx);
    for (int c = 0; c < 3; ++c) inv_trans(o, sx + c * nc, o->u + c * n3);
    for (int c = 0; c < 9; ++c) inv_trans(o, sx + (3 + c) * nc, o->gu + c * n3);
    err_sums(o, err);
    if (err[0] / err[1] < 1.e-7 && err[2] / err[3] < 1.e-7 && err[4] / err[5] < 1.e-7) converged = 1;
    if (niter > 1 && !converged) {
      if (err[0] / err[1] > err_old[0] / err_old[1] || err[2] / err[3] > err_old[2] / err_old[3] ||
          err[4] / err[5] > err_old[4] / err_old[5]) {
        cfac = cfac * newfac;
        newfac = newfac * 0.5;
        if (newfac < 1.e-7) return MM_ERR_NEWFAC;
        niter = 0;
        converged = 0;
        continue;
      }
    }
    memcpy(err_old, err, sizeof(err));
    if (niter == 10) converged = 1;
    if (!converged || niter < 4) continue;
    if (cfac < 1.0 - 1.e-6) {
      cfac = cfac + 0.1;
      if (cfac > 1.0) cfac = 1.0;
      niter = 0;
      converged = 0;
      continue;
    }
    break;
  }
  if (n_iter) *n_iter = total;
  if (cfac_out) *cfac_out = cfac;
  return elener(o);
}

/* P_phi = p(phi) of the current phi (the interpolation of
 * Multi_MEUMAPPS.f:653), without the rest of the chemistry */
int oracle_set_pphi(oracle_t *o) {
  long n3 = o->n3;
  for (long i = 0; i < (long)o->V * n3; ++i) {
    double fi = o->phi[i], fi_sq = fi * fi;
    o->Pphi[i] = fi_sq * fi * (6.0 * fi_sq - 15.0 * fi + 10.0);
  }
  return 0;
}

/* elener alone (ModElastic.f:493-691) at the current phi, P_phi and grad u:
 * el_en and el_grad of the state as it stands */
int oracle_elener(oracle_t *o) {
  memset(o->elgrad, 0, sizeof(double) * o->V * o->n3);
  return elener(o);
}

/* the zeroth-order guess alone, from given P_phi fields (V*N3): u (3*N3)
 * and its gradients (9*N3, d_d u_i at d*3+i); either output may be NULL */
int oracle_elastic_zeroth(oracle_t *o, const double *pphi, double *u, double *gradu) {
  int rc;
  if (!o->qinv_ready && (rc = build_qinv(o))) return rc;
  memcpy(o->Pphi, pphi, sizeof(double) * o->V * o->n3);
  eldis_zeroth(o);
  if (u) memcpy(u, o->u, sizeof(double) * 3 * o->n3);
  if (gradu) memcpy(gradu, o->gu, sizeof(double) * 9 * o->n3);
  return 0;
}

int oracle_step(oracle_t *o, mm_diag *d) {
  const int V = o->V, E = o->E, H = o->H, ny = o->ny, nz = o->nz, nt = o->nthreads;
  long n3 = o->n3, nc = o->nc;
  int rc;
  const int noisy = o->step <= o->p.noise_step;
  if (noisy && !o->noise_on) return MM_ERR_ARG;
  if (o->t && (noisy || o->step == o->p.elastep0)) return MM_ERR_ARG;
  /* noise, :552-571: one langevin draw per point and variant, z/y/x/variant order */
  if (noisy) {
    for (int z = 0; z < nz; ++z)
      for (int y = 0; y < ny; ++y)
        for (int x = 0; x < o->nx; ++x)
          for (int v = 0; v < V; ++v)
            o->phi_noise[v * n3 + RIDX(o, x, y, z)] += oracle_langevin(&o->r_idum, o->r_state, o->noise_sd);
  } else if (o->phi_noise) {
    memset(o->phi_noise, 0, sizeof(double) * V * n3);
  }
  mm_diag dd;
  memset(&dd, 0, sizeof(dd));
  dd.step = o->step;
  gradients(o);
  if (o->t) tlm_gradients(o);
  double fint, fch, fw;
  if ((rc = chemistry(o, &fint, &fch, &fw))) return rc;
  if (o->t) tlm_chemistry(o);
  memset(o->elgrad, 0, sizeof(double) * V * n3);
  int niter = 0;
  double cfac = 1.0;
  if (o->step >= o->p.elastep0) {
    if ((rc = oracle_eldis(o, &niter, &cfac))) return rc;
    if (o->t) tlm_elener(o);
  } else {
    memset(o->el_en, 0, sizeof(double) * n3);
  }
  for (long i = 0; i < V * n3; ++i) o->fdiff[i] = o->fdiff[i] + o->elgrad[i];
  double el_sum = 0.0;
  {
    long slab = (long)o->nx * o->ny;
    for (int z = 0; z < nz; ++z) {
      double s = 0.0;
      for (long i = z * slab; i < (z + 1) * slab; ++i) s += o->el_en[i];
      el_sum += s;
    }
  }
  dd.fint = fint; dd.fw = fw; dd.fch = fch; dd.elen = el_sum;
  dd.etot = fint + fch + el_sum + fw;
  dd.glob_min = o->glob_min; dd.glob_max = o->glob_max; dd.phi_max = o->phi_max;
  dd.solid_tot = o->solid_frac / (double)(float)n3;
  dd.n_elastic = niter;
  dd.cfac = cfac;
  /* Cahn-Hilliard, :910-949 */
  if (o->t) tlm_ch(o);
  cplx *sc = o->sw, *sfx = o->sw + nc, *sfy = o->sw + 2 * nc, *sfz = o->sw + 3 * nc;
  for (int e = 0; e < E; ++e) {
    f_trans(o, o->cons + e * n3, sc);
    f_trans(o, o->Fx + e * n3, sfx);
    f_trans(o, o->Fy + e * n3, sfy);
    f_trans(o, o->Fz + e * n3, sfz);
    double dtm = o->p.t_step * o->p.cmobs[e];
#pragma omp parallel for num_threads(nt) schedule(static)
    for (long idx = 0; idx < nc; ++idx) {
      int kx = (int)(idx % H), ky = (int)((idx / H) % ny), kz = (int)(idx / ((long)H * ny));
      double ksq = o->kqx[kx] + o->kqy[ky] + o->kqz[kz];
      cplx t1 = mul_i(sfx[idx], o->kfx[kx]) + mul_i(sfy[idx], o->kfy[ky]) + mul_i(sfz[idx], o->kfz[kz]);
      double t2 = 1.0 - dtm * ksq;
      cplx num = sc[idx] + dtm * t1;
      sc[idx] = creal(num) / t2 + I * (cimag(num) / t2);
    }
    inv_trans(o, sc, o->cons + e * n3);
  }
  /* Allen-Cahn, :955-1001 (term3 = t_step F(phi_noise), zero past noise_step) */
  if (o->t) tlm_ac(o);
  cplx *sf = o->sw + nc, *sn = o->sw + 2 * nc;
  for (int v = 0; v < V; ++v) {
    cplx *sp = o->sphi + v * nc; /* FFT of phi from the gradient block; phi unchanged since */
    f_trans(o, o->fdiff + v * n3, sf);
    if (noisy) f_trans(o, o->phi_noise + v * n3, sn);
    double dtl = o->p.t_step * o->p.b_mob, dt = o->p.t_step;
#pragma omp parallel for num_threads(nt) schedule(static)
    for (long idx = 0; idx < nc; ++idx) {
      int kx = (int)(idx % H), ky = (int)((idx / H) % ny), kz = (int)(idx / ((long)H * ny));
      cplx t1 = dtl * sf[idx];
      double t2 = 1.0 - dtl * kgk(o, v, kx, ky, kz);
      cplx num = sp[idx] - t1;
      if (noisy) num = num + sn[idx] * dt;
      sp[idx] = creal(num) / t2 + I * (cimag(num) / t2);
    }
    inv_trans(o, sp, o->phi + v * n3);
  }
  /* check loop, :1008-1065 */
  if (o->t) tlm_clamp(o);
  {
    long slab = (long)o->nx * o->ny;
    double *pf = calloc((size_t)nz * (4 + E), sizeof(double));
#pragma omp parallel for num_threads(nt) schedule(static)
    for (int z = 0; z < nz; ++z) {
      double cmin = 1.0, cmax = 0.0, pmax = 0.0, sf = 0.0, cl[MM_MAX_EL] = {0};
      for (long i = z * slab; i < (z + 1) * slab; ++i) {
        for (int e = 0; e < E; ++e) {
          double c = o->cons[e * n3 + i];
          if (c < cmin) cmin = c;
          if (c > cmax) cmax = c;
          cl[e] = cl[e] + c;
        }
        for (int v = 0; v < V; ++v) {
          double f = o->phi[v * n3 + i];
          if (f < 0.0) f = 0.0;
          if (f > 1.0) f = 1.0;
          o->phi[v * n3 + i] = f;
          if (f > pmax) pmax = f;
          if (f > 0.0) sf = sf + f;
        }
      }
      double *q = pf + (size_t)z * (4 + E);
      q[0] = cmin; q[1] = cmax; q[2] = pmax; q[3] = sf;
      for (int e = 0; e < E; ++e) q[4 + e] = cl[e];
    }
    double cmin = 1.0, cmax = 0.0, pmax = 0.0, sfr = 0.0, cl[MM_MAX_EL] = {0};
    for (int z = 0; z < nz; ++z) {
      double *q = pf + (size_t)z * (4 + E);
      if (q[0] < cmin) cmin = q[0];
      if (q[1] > cmax) cmax = q[1];
      if (q[2] > pmax) pmax = q[2];
      sfr += q[3];
      for (int e = 0; e < E; ++e) cl[e] += q[4 + e];
    }
    free(pf);
    for (int e = 0; e < E; ++e) dd.caviter[e] = cl[e] * o->ooN;
    o->glob_min = cmin; o->glob_max = cmax; o->phi_max = pmax; o->solid_frac = sfr;
  }
  o->step += 1;
  if (d) *d = dd;
  return 0;
}
