/* CPU oracle for the RGF hot path -- TEST INFRASTRUCTURE ONLY (the checker and
 * the CPU baseline; never part of the product).  See rgf_oracle.h for what is
 * restated and how each function is pinned. */
#include "rgf_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define RE(a, i) ((a)[2 * (size_t)(i)])
#define IM(a, i) ((a)[2 * (size_t)(i) + 1])

typedef struct { double r, i; } cplx;

static inline cplx cmul(cplx a, cplx b) { cplx c = {a.r * b.r - a.i * b.i, a.r * b.i + a.i * b.r}; return c; }
/* Smith's algorithm for a/b. */
static inline cplx cdiv(cplx a, cplx b) {
  cplx c;
  if (fabs(b.r) >= fabs(b.i)) {
    double t = b.i / b.r, d = b.r + b.i * t;
    c.r = (a.r + a.i * t) / d;
    c.i = (a.i - a.r * t) / d;
  } else {
    double t = b.r / b.i, d = b.r * t + b.i;
    c.r = (a.r * t + a.i) / d;
    c.i = (a.i * t - a.r) / d;
  }
  return c;
}
static inline cplx ld(const double *a, size_t k) { cplx c = {a[2 * k], a[2 * k + 1]}; return c; }
static inline void st(double *a, size_t k, cplx c) { a[2 * k] = c.r; a[2 * k + 1] = c.i; }

void orc_params_reference(orc_params *p) {
  p->eta_dev = 0.0;      /* green_rgf.f90:3974 */
  p->eta_lead = 1.0e-3;  /* green_rgf.f90:4474 */
  p->dec_tol = 1.0e-100; /* green_rgf.f90:4457 */
  p->dec_nmax = 100;     /* green_rgf.f90:4473 */
  p->boltz = 8.61734e-05;/* green_rgf.f90:3968 */
  p->mode = ORC_MODE_COMPAT;
}

/* ferm, green_rgf.f90:4444-4447 */
double orc_ferm(double a) { return 1.0 / (1.0 + exp(a)); }

/* ---------------------------------------------------------------- zgemm */
/* C = op(A) op(B) for square n x n column-major complex matrices (alpha=1,
 * beta=0 as in every call of green_rgf.f90:4002-4155).  op(A) and op(B) are
 * first expanded into split real/imaginary column-major panels so that the
 * inner loop is a unit-stride multiply-add. */
void orc_zgemm(char ta, char tb, int n, const double *A, const double *B, double *C) {
  size_t nn = (size_t)n * n;
  double *ar = (double *)malloc(sizeof(double) * nn * 4);
  double *ai = ar + nn, *br = ar + 2 * nn, *bi = ar + 3 * nn;
  for (int k = 0; k < n; ++k)
    for (int i = 0; i < n; ++i) {
      size_t d = (size_t)i + (size_t)k * n;
      if (ta == 'n' || ta == 'N') { ar[d] = RE(A, d); ai[d] = IM(A, d); }
      else { size_t s = (size_t)k + (size_t)i * n; ar[d] = RE(A, s); ai[d] = -IM(A, s); }
      if (tb == 'n' || tb == 'N') { br[d] = RE(B, d); bi[d] = IM(B, d); }
      else { size_t s = (size_t)k + (size_t)i * n; br[d] = RE(B, s); bi[d] = -IM(B, s); }
    }
  /* br/bi now hold op(B) with element (k, j) at k + j*n (d above was built
   * with i as the row index, so op(B)(i,k) is stored at i + k*n). */
  double *cr = (double *)calloc(nn * 2, sizeof(double));
  double *ci = cr + nn;
  int j = 0;
  for (; j + 4 <= n; j += 4) {
    double *c0r = cr + (size_t)j * n, *c1r = c0r + n, *c2r = c1r + n, *c3r = c2r + n;
    double *c0i = ci + (size_t)j * n, *c1i = c0i + n, *c2i = c1i + n, *c3i = c2i + n;
    for (int k = 0; k < n; ++k) {
      const double *xr = ar + (size_t)k * n, *xi = ai + (size_t)k * n;
      double b0r = br[k + (size_t)j * n], b0i = bi[k + (size_t)j * n];
      double b1r = br[k + (size_t)(j + 1) * n], b1i = bi[k + (size_t)(j + 1) * n];
      double b2r = br[k + (size_t)(j + 2) * n], b2i = bi[k + (size_t)(j + 2) * n];
      double b3r = br[k + (size_t)(j + 3) * n], b3i = bi[k + (size_t)(j + 3) * n];
      for (int i = 0; i < n; ++i) {
        double pr = xr[i], pi = xi[i];
        c0r[i] += pr * b0r - pi * b0i; c0i[i] += pr * b0i + pi * b0r;
        c1r[i] += pr * b1r - pi * b1i; c1i[i] += pr * b1i + pi * b1r;
        c2r[i] += pr * b2r - pi * b2i; c2i[i] += pr * b2i + pi * b2r;
        c3r[i] += pr * b3r - pi * b3i; c3i[i] += pr * b3i + pi * b3r;
      }
    }
  }
  for (; j < n; ++j) {
    double *c0r = cr + (size_t)j * n, *c0i = ci + (size_t)j * n;
    for (int k = 0; k < n; ++k) {
      const double *xr = ar + (size_t)k * n, *xi = ai + (size_t)k * n;
      double b0r = br[k + (size_t)j * n], b0i = bi[k + (size_t)j * n];
      for (int i = 0; i < n; ++i) {
        c0r[i] += xr[i] * b0r - xi[i] * b0i;
        c0i[i] += xr[i] * b0i + xi[i] * b0r;
      }
    }
  }
  for (size_t d = 0; d < nn; ++d) { RE(C, d) = cr[d]; IM(C, d) = ci[d]; }
  free(cr);
  free(ar);
}

/* --------------------------------------------------------------- invert */
static inline double cabs1(cplx a) { return fabs(a.r) + fabs(a.i); }

/* LAPACK zgetf2: unblocked LU with partial pivoting (pivot = first maximum of
 * |Re|+|Im| as izamax), row interchanges applied to whole rows. */
static int zgetf2(int n, double *A, int *ipiv) {
  int info = 0;
  const double sfmin = 2.2250738585072014e-308;
  for (int j = 0; j < n; ++j) {
    int jp = j;
    double best = cabs1(ld(A, (size_t)j + (size_t)j * n));
    for (int i = j + 1; i < n; ++i) {
      double v = cabs1(ld(A, (size_t)i + (size_t)j * n));
      if (v > best) { best = v; jp = i; }
    }
    ipiv[j] = jp;
    cplx piv = ld(A, (size_t)jp + (size_t)j * n);
    if (piv.r != 0.0 || piv.i != 0.0) {
      if (jp != j)
        for (int c = 0; c < n; ++c) {
          size_t a = (size_t)j + (size_t)c * n, b = (size_t)jp + (size_t)c * n;
          cplx t = ld(A, a); st(A, a, ld(A, b)); st(A, b, t);
        }
      cplx ajj = ld(A, (size_t)j + (size_t)j * n);
      if (hypot(ajj.r, ajj.i) >= sfmin) {
        cplx one = {1.0, 0.0}, r = cdiv(one, ajj);
        for (int i = j + 1; i < n; ++i) {
          size_t a = (size_t)i + (size_t)j * n;
          st(A, a, cmul(r, ld(A, a)));
        }
      } else {
        for (int i = j + 1; i < n; ++i) {
          size_t a = (size_t)i + (size_t)j * n;
          st(A, a, cdiv(ld(A, a), ajj));
        }
      }
    } else if (info == 0) {
      info = j + 1;
    }
    /* rank-1 update of the trailing block (zgeru with alpha = -1) */
    for (int c = j + 1; c < n; ++c) {
      cplx y = ld(A, (size_t)j + (size_t)c * n);
      if (y.r == 0.0 && y.i == 0.0) continue;
      for (int i = j + 1; i < n; ++i) {
        size_t a = (size_t)i + (size_t)c * n;
        cplx x = ld(A, (size_t)i + (size_t)j * n), t = cmul(x, y), v = ld(A, a);
        v.r -= t.r; v.i -= t.i;
        st(A, a, v);
      }
    }
  }
  return info;
}

/* LAPACK ztrtri/ztrti2 (upper, non-unit) followed by the unblocked zgetri
 * solve inv(A)*L = inv(U) and the column interchanges. */
static int zgetri(int n, double *A, const int *ipiv) {
  for (int j = 0; j < n; ++j) {
    cplx d = ld(A, (size_t)j + (size_t)j * n);
    if (d.r == 0.0 && d.i == 0.0) return j + 1;
  }
  /* ztrti2 */
  for (int j = 0; j < n; ++j) {
    size_t dj = (size_t)j + (size_t)j * n;
    cplx one = {1.0, 0.0};
    cplx inv = cdiv(one, ld(A, dj));
    st(A, dj, inv);
    cplx ajj = {-inv.r, -inv.i};
    /* x := T x, T = A(0:j-1, 0:j-1) upper non-unit, x = A(0:j-1, j) */
    double *x = A + 2 * (size_t)j * n;
    for (int jj = 0; jj < j; ++jj) {
      cplx t = ld(x, jj);
      if (t.r != 0.0 || t.i != 0.0) {
        for (int i = 0; i < jj; ++i) {
          cplx a = ld(A, (size_t)i + (size_t)jj * n), p = cmul(t, a), v = ld(x, i);
          v.r += p.r; v.i += p.i;
          st(x, i, v);
        }
        st(x, jj, cmul(t, ld(A, (size_t)jj + (size_t)jj * n)));
      }
    }
    for (int i = 0; i < j; ++i) st(x, i, cmul(ajj, ld(x, i)));
  }
  /* solve inv(A) * L = inv(U) */
  double *work = (double *)malloc(sizeof(double) * 2 * (size_t)n);
  for (int j = n - 1; j >= 0; --j) {
    for (int i = j + 1; i < n; ++i) {
      size_t a = (size_t)i + (size_t)j * n;
      st(work, i, ld(A, a));
      cplx z = {0.0, 0.0};
      st(A, a, z);
    }
    if (j < n - 1) {
      /* A(:, j) -= A(:, j+1:n) * work(j+1:n)  (zgemv) */
      for (int c = j + 1; c < n; ++c) {
        cplx w = ld(work, c);
        if (w.r == 0.0 && w.i == 0.0) continue;
        for (int i = 0; i < n; ++i) {
          size_t a = (size_t)i + (size_t)j * n;
          cplx p = cmul(ld(A, (size_t)i + (size_t)c * n), w), v = ld(A, a);
          v.r -= p.r; v.i -= p.i;
          st(A, a, v);
        }
      }
    }
  }
  free(work);
  for (int j = n - 2; j >= 0; --j) {
    int jp = ipiv[j];
    if (jp != j)
      for (int i = 0; i < n; ++i) {
        size_t a = (size_t)i + (size_t)j * n, b = (size_t)i + (size_t)jp * n;
        cplx t = ld(A, a); st(A, a, ld(A, b)); st(A, b, t);
      }
  }
  return 0;
}

/* invert, green_rgf.f90:4421-4441: A <- A^-1, or A <- 0 on failure. */
int orc_zinv(int n, double *A) {
  int *ipiv = (int *)malloc(sizeof(int) * (size_t)n);
  int info = zgetf2(n, A, ipiv);
  if (info == 0) info = zgetri(n, A, ipiv);
  if (info != 0) memset(A, 0, sizeof(double) * 2 * (size_t)n * n);
  free(ipiv);
  return info;
}

/* ----------------------------------------------------------- small ops */
static double *mnew(int n) { return (double *)calloc(2 * (size_t)n * n, sizeof(double)); }
static void mcopy(int n, double *d, const double *s) { memcpy(d, s, sizeof(double) * 2 * (size_t)n * n); }
static void mzero(int n, double *d) { memset(d, 0, sizeof(double) * 2 * (size_t)n * n); }
/* d = a + b (d may alias a or b) */
static void madd(int n, double *d, const double *a, const double *b) {
  for (size_t k = 0; k < 2 * (size_t)n * n; ++k) d[k] = a[k] + b[k];
}
/* d = a - a^H, the anti-Hermitian part times two */
static void mantiherm(int n, double *d, const double *a) {
  for (int j = 0; j < n; ++j)
    for (int i = 0; i < n; ++i) {
      size_t p = (size_t)i + (size_t)j * n, q = (size_t)j + (size_t)i * n;
      RE(d, p) = RE(a, p) - RE(a, q);
      IM(d, p) = IM(a, p) + IM(a, q);
    }
}
/* d = z*I - a - b (b may be NULL) */
static void mzminus(int n, double *d, double zr, double zi, const double *a, const double *b) {
  for (size_t k = 0; k < (size_t)n * n; ++k) {
    RE(d, k) = -RE(a, k) - (b ? RE(b, k) : 0.0);
    IM(d, k) = -IM(a, k) - (b ? IM(b, k) : 0.0);
  }
  for (int i = 0; i < n; ++i) {
    size_t k = (size_t)i + (size_t)i * n;
    RE(d, k) += zr;
    IM(d, k) += zi;
  }
}
static const double *blk(const double *a, int n, int l) { return a + 2 * (size_t)n * n * (size_t)l; }
static double *blkw(double *a, int n, int l) { return a + 2 * (size_t)n * n * (size_t)l; }

/* d = z*S - a (S may be NULL = identity) */
static void mzs(int n, double *d, double zr, double zi, const double *S, const double *a) {
  if (!S) {
    mzminus(n, d, zr, zi, a, NULL);
    return;
  }
  for (size_t k = 0; k < (size_t)n * n; ++k) {
    RE(d, k) = zr * RE(S, k) - zi * IM(S, k) - RE(a, k);
    IM(d, k) = zr * IM(S, k) + zi * RE(S, k) - IM(a, k);
  }
}

/* ------------------------------------------------------------- sancho */
/* Sancho-Rubio decimation, green_rgf.f90:4451-4537, with the overlap S00
 * (NULL = identity, as green_RGF_RS passes) and the broadening, tolerance and
 * cap as parameters. */
static int sancho_impl(int n, double E, double eta_lead, double tol, int nmax, const double *S00,
                       const double *H00, const double *H10, double *G00, double *GBB,
                       int *nmax_hit, int *inv_failures);

int orc_sancho(int n, double E, double eta_lead, double tol, int nmax,
               const double *H00, const double *H10, double *G00, double *GBB,
               int *nmax_hit, int *inv_failures) {
  return sancho_impl(n, E, eta_lead, tol, nmax, NULL, H00, H10, G00, GBB, nmax_hit, inv_failures);
}

int orc_sancho_s(int n, double E, double eta_lead, double tol, int nmax, const double *S00,
                 const double *H00, const double *H10, double *G00, double *GBB,
                 int *nmax_hit, int *inv_failures) {
  return sancho_impl(n, E, eta_lead, tol, nmax, S00, H00, H10, G00, GBB, nmax_hit, inv_failures);
}

static int sancho_impl(int n, double E, double eta_lead, double tol, int nmax, const double *S00,
                       const double *H00, const double *H10, double *G00, double *GBB,
                       int *nmax_hit, int *inv_failures) {
  double *A = mnew(n), *B = mnew(n), *C = mnew(n);
  double *HBB = mnew(n), *HSS = mnew(n), *H01 = mnew(n), *Hx10 = mnew(n);
  int fails = 0, it_run = 0;
  if (nmax_hit) *nmax_hit = 0;
  mcopy(n, HBB, H00);
  mcopy(n, Hx10, H10);
  for (int j = 0; j < n; ++j)
    for (int i = 0; i < n; ++i) {
      size_t p = (size_t)i + (size_t)j * n, q = (size_t)j + (size_t)i * n;
      RE(H01, p) = RE(H10, q);
      IM(H01, p) = -IM(H10, q);
    }
  mcopy(n, HSS, H00);
  for (int it = 1; it <= nmax; ++it) {
    it_run = it;
    mzs(n, A, E, eta_lead, S00, HBB);
    if (orc_zinv(n, A)) ++fails;
    orc_zgemm('n', 'n', n, A, Hx10, B);
    orc_zgemm('n', 'n', n, H01, B, C);
    madd(n, HSS, HSS, C);
    madd(n, HBB, HBB, C);
    orc_zgemm('n', 'n', n, Hx10, B, C);
    orc_zgemm('n', 'n', n, A, H01, B);
    orc_zgemm('n', 'n', n, Hx10, B, A);
    mcopy(n, Hx10, C);
    madd(n, HBB, HBB, A);
    orc_zgemm('n', 'n', n, H01, B, C);
    mcopy(n, H01, C);
    double error = 0.0;
    for (int k = 0; k < n; ++k)
      for (int j = 0; j < n; ++j) {
        size_t p = (size_t)k + (size_t)j * n;
        error += sqrt(IM(C, p) * IM(C, p) + RE(C, p) * RE(C, p));
      }
    if (fabs(error) < tol) break;
    if (it == nmax) {
      if (nmax_hit) *nmax_hit = 1;
      mcopy(n, HSS, H00);
      mcopy(n, HBB, H00);
    }
  }
  mzs(n, G00, E, eta_lead, S00, HSS);
  if (orc_zinv(n, G00)) ++fails;
  if (GBB) {
    mzs(n, GBB, E, eta_lead, S00, HBB);
    if (orc_zinv(n, GBB)) ++fails;
  }
  if (inv_failures) *inv_failures += fails;
  free(A); free(B); free(C); free(HBB); free(HSS); free(H01); free(Hx10);
  return it_run;
}

/* ---------------------------------------------------------------- RGF */
/* green_RGF_RS, green_rgf.f90:3950-4164.  The product sequence of the
 * reference is kept (same zgemm operand order), so compat mode reproduces its
 * arithmetic, including:
 *   - the lead/device broadening, tolerance, cap and Boltzmann constant as
 *     parameters (reference values via orc_params_reference);
 *   - compat: H1i(:,:,l) is the l-1,l coupling in the forward sweep and the
 *     l,l+1 coupling in the backward sweep (green_rgf.f90:4015-4018 vs 4070);
 *     physical: H1i(:,:,l+1) is the l,l+1 coupling in the backward sweep;
 *   - compat: GGi1 reads the never-filled Glp (all zero, green_rgf.f90:3978,
 *     4098), so GGi1_l = (g_l^H H^H) G^>_{nx,nx}; physical: the true
 *     G^>_{l,l+1} = G^<_{l,l+1} + G^R_{l,l+1} - (G^R_{l+1,l})^H;
 *   - compat: the right contact uses TEMP(1) and tre uses TEMP(2)
 *     (green_rgf.f90:4051, 4058, 4152); physical: the natural assignment. */
static void rgf_impl(const orc_params *p, const double temp[2], int nm, int nx, double E,
                     const double mu[2], const double *Hii, const double *H1i,
                     const double *sigL, const double *sigG, const double *sigR,
                     double *ndens, double *pdens, double *ldos, double *tr, double *tre,
                     double *cur, double *GRi1, double *GLi1, double *GGi1, orc_diag *diag,
                     const double *leadL_H00, const double *leadR_H00) {
  (void)sigG; /* accepted but unused by the reference (green_rgf.f90:3953) */
  const int n = nm;
  const size_t bs = 2 * (size_t)n * n;
  const int phys = (p->mode == ORC_MODE_PHYSICAL);
  const double mul = mu[0], mur = mu[1], templ = temp[0], tempr = temp[1];
  const double temp_right = phys ? tempr : templ;  /* trap: TEMPl at 4051/4058 */
  const double temp_left_tre = phys ? templ : tempr; /* trap: TEMPr at 4152 */
  const double zr = E, zi = p->eta_dev;
  orc_diag dg = {0, 0, 0, 0};

  double *Gl = (double *)calloc(bs * (size_t)nx, sizeof(double));
  double *Gln = (double *)calloc(bs * (size_t)nx, sizeof(double));
  double *H00 = mnew(n), *H10 = mnew(n), *A = mnew(n), *B = mnew(n), *C = mnew(n);
  double *D = mnew(n), *G00 = mnew(n), *GBB = mnew(n), *GN0 = mnew(n), *Gn = mnew(n);
  double *Gp = mnew(n), *sig = mnew(n), *sigmal = mnew(n), *sigmar = mnew(n);
  double *H10h = mnew(n), *T1 = mnew(n);

  if (ldos) memset(ldos, 0, bs * (size_t)nx * sizeof(double));
  if (ndens) memset(ndens, 0, bs * (size_t)nx * sizeof(double));
  if (pdens) memset(pdens, 0, bs * (size_t)nx * sizeof(double));
  if (cur) memset(cur, 0, (size_t)n * n * (size_t)nx * sizeof(double));

  /* left contact, green_rgf.f90:3992-4012 */
  if (sigR) madd(n, H00, blk(Hii, n, 0), blk(sigR, n, 0)); else mcopy(n, H00, blk(Hii, n, 0));
  double *H00dev = mnew(n);
  mcopy(n, H00dev, H00);
  if (leadL_H00) mcopy(n, H00, leadL_H00);
  mcopy(n, H10, blk(H1i, n, 0));
  for (int j = 0; j < n; ++j)
    for (int i = 0; i < n; ++i) {
      size_t a = (size_t)i + (size_t)j * n, b = (size_t)j + (size_t)i * n;
      RE(H10h, a) = RE(H10, b);
      IM(H10h, a) = -IM(H10, b);
    }
  {
    int hit = 0;
    dg.iters_left = orc_sancho(n, E, p->eta_lead, p->dec_tol, p->dec_nmax, H00, H10h, G00, GBB,
                               &hit, &dg.inv_failures);
    if (hit) dg.nmax_hit |= 1;
  }
  mcopy(n, H00, H00dev);
  orc_zgemm('n', 'n', n, H10, G00, A);
  orc_zgemm('n', 'c', n, A, H10, sigmal);
  {
    double f = orc_ferm((E - mul) / (p->boltz * templ));
    mantiherm(n, sig, sigmal);
    for (size_t k = 0; k < bs; ++k) sig[k] = -sig[k] * f + (sigL ? blk(sigL, n, 0)[k] : 0.0);
  }
  mzminus(n, A, zr, zi, H00, sigmal);
  if (orc_zinv(n, A)) ++dg.inv_failures;
  mcopy(n, blkw(Gl, n, 0), A);
  orc_zgemm('n', 'n', n, A, sig, B);
  orc_zgemm('n', 'c', n, B, A, C);
  mcopy(n, blkw(Gln, n, 0), C);

  /* forward sweep, green_rgf.f90:4013-4030 */
  for (int l = 1; l < nx - 1; ++l) {
    if (sigR) madd(n, H00, blk(Hii, n, l), blk(sigR, n, l)); else mcopy(n, H00, blk(Hii, n, l));
    mcopy(n, H10, blk(H1i, n, l));
    orc_zgemm('n', 'n', n, H10, blk(Gl, n, l - 1), B);
    orc_zgemm('n', 'c', n, B, H10, C);
    mzminus(n, A, zr, zi, H00, C);
    if (orc_zinv(n, A)) ++dg.inv_failures;
    mcopy(n, blkw(Gl, n, l), A);
    mcopy(n, sig, blk(Gln, n, l - 1));
    orc_zgemm('n', 'n', n, H10, sig, B);
    orc_zgemm('n', 'c', n, B, H10, C);
    if (sigL) madd(n, C, C, blk(sigL, n, l));
    orc_zgemm('n', 'n', n, A, C, B);
    orc_zgemm('n', 'c', n, B, A, Gn);
    mcopy(n, blkw(Gln, n, l), Gn);
  }

  /* right contact, green_rgf.f90:4032-4066 */
  if (sigR) madd(n, H00, blk(Hii, n, nx - 1), blk(sigR, n, nx - 1)); else mcopy(n, H00, blk(Hii, n, nx - 1));
  mcopy(n, H10, blk(H1i, n, nx - 1));
  {
    int hit = 0;
    mcopy(n, H00dev, H00);
    dg.iters_right = orc_sancho(n, E, p->eta_lead, p->dec_tol, p->dec_nmax, leadR_H00 ? leadR_H00 : H00, H10,
                                G00, GBB, &hit, &dg.inv_failures);
    if (hit) dg.nmax_hit |= 2;
  }
  orc_zgemm('c', 'n', n, H10, G00, A);
  orc_zgemm('n', 'n', n, A, H10, sigmar);
  orc_zgemm('n', 'n', n, H10, blk(Gl, n, nx - 2), B);
  orc_zgemm('n', 'c', n, B, H10, C);
  mzminus(n, G00, zr, zi, H00, sigmar);
  for (size_t k = 0; k < bs; ++k) G00[k] -= C[k];
  if (orc_zinv(n, G00)) ++dg.inv_failures;
  if (ldos) mcopy(n, blkw(ldos, n, nx - 1), G00);
  mcopy(n, sig, blk(Gln, n, nx - 2));
  orc_zgemm('n', 'n', n, H10, sig, B);
  orc_zgemm('n', 'c', n, B, H10, C);
  double fR = orc_ferm((E - mur) / (p->boltz * temp_right));
  mantiherm(n, T1, sigmar); /* T1 = sigmar - sigmar^H */
  for (size_t k = 0; k < bs; ++k)
    sig[k] = -T1[k] * fR + C[k] + (sigL ? blk(sigL, n, nx - 1)[k] : 0.0);
  orc_zgemm('n', 'n', n, G00, sig, B);
  orc_zgemm('n', 'c', n, B, G00, Gn);
  if (ndens) mcopy(n, blkw(ndens, n, nx - 1), Gn);
  mantiherm(n, Gp, G00);
  madd(n, Gp, Gn, Gp);
  if (pdens) mcopy(n, blkw(pdens, n, nx - 1), Gp);
  for (size_t k = 0; k < bs; ++k) A[k] = -T1[k] * fR;
  orc_zgemm('n', 'n', n, A, Gp, B);
  for (size_t k = 0; k < bs; ++k) A[k] = -T1[k] * (fR - 1.0);
  orc_zgemm('n', 'n', n, A, Gn, C);
  {
    double tim = 0.0;
    for (int i = 0; i < n; ++i) {
      size_t d = (size_t)i + (size_t)i * n;
      tim = tim - (RE(B, d) - RE(C, d));
    }
    if (tr) *tr = tim;
  }

  /* backward sweep, green_rgf.f90:4069-4149 */
  for (int l = nx - 2; l >= 0; --l) {
    const int lc = phys ? l + 1 : l;  /* trap: H1i(:,:,l) is used as the l,l+1 coupling */
    mcopy(n, H10, blk(H1i, n, lc));
    const double *gl = blk(Gl, n, l), *gln = blk(Gln, n, l);
    double *curl = cur ? cur + (size_t)n * n * (size_t)l : NULL;
    /* G<_{l+1,l} part of the bond current, 4071-4079 */
    orc_zgemm('n', 'c', n, H10, gl, B);
    orc_zgemm('n', 'n', n, Gn, B, C);
    orc_zgemm('n', 'n', n, H10, gln, B);
    orc_zgemm('n', 'n', n, G00, B, A);
    madd(n, B, C, A);
    orc_zgemm('c', 'n', n, H10, B, A);
    if (curl) for (size_t k = 0; k < (size_t)n * n; ++k) curl[k] = RE(A, k);
    /* G<_{l,l+1}, 4081-4092 */
    orc_zgemm('n', 'c', n, gl, H10, B);
    orc_zgemm('n', 'n', n, B, Gn, C);
    orc_zgemm('n', 'c', n, gln, H10, B);
    orc_zgemm('n', 'c', n, B, G00, A);
    madd(n, B, C, A);
    if (GLi1) mcopy(n, blkw(GLi1, n, l), B);
    if (GGi1 && phys) mcopy(n, T1, B); /* keep G<_{l,l+1} for the physical G> */
    orc_zgemm('n', 'n', n, H10, B, A);
    if (curl) for (size_t k = 0; k < (size_t)n * n; ++k) curl[k] -= RE(A, k);
    /* G>_{l,l+1}: compat 4094-4103 (Glp == 0) */
    if (GGi1 && !phys) {
      orc_zgemm('c', 'c', n, gl, H10, B);
      orc_zgemm('n', 'n', n, B, Gp, C);
      mcopy(n, blkw(GGi1, n, l), C);
    }
    if (GGi1 && phys) {
      /* G^R_{l+1,l} = G_{l+1,l+1} H_{l+1,l} g_l with the pre-update G00 */
      orc_zgemm('n', 'n', n, G00, H10, A);
      orc_zgemm('n', 'n', n, A, gl, C);
    }
    /* G_{l,l+1} and G_{l,l}, 4105-4114 */
    mcopy(n, D, gl);
    orc_zgemm('n', 'c', n, D, H10, B);
    orc_zgemm('n', 'n', n, B, G00, GN0);
    if (GRi1) mcopy(n, blkw(GRi1, n, l), GN0);
    if (GGi1 && phys) {
      /* GGi1 = G<_{l,l+1} + G_{l,l+1} - (G_{l+1,l})^H ; C holds G_{l+1,l} */
      double *o = blkw(GGi1, n, l);
      for (int j = 0; j < n; ++j)
        for (int i = 0; i < n; ++i) {
          size_t a = (size_t)i + (size_t)j * n, b = (size_t)j + (size_t)i * n;
          RE(o, a) = RE(T1, a) + RE(GN0, a) - RE(C, b);
          IM(o, a) = IM(T1, a) + IM(GN0, a) + IM(C, b);
        }
    }
    orc_zgemm('n', 'n', n, GN0, H10, A);
    orc_zgemm('n', 'n', n, A, D, C);
    madd(n, G00, gl, C);
    if (ldos) mcopy(n, blkw(ldos, n, l), G00);
    /* G<_{l,l}, 4116-4128 */
    mcopy(n, A, Gn);
    orc_zgemm('n', 'c', n, D, H10, B);
    orc_zgemm('n', 'n', n, B, A, C);
    orc_zgemm('n', 'n', n, C, H10, A);
    orc_zgemm('n', 'c', n, A, D, C);
    madd(n, Gn, gln, C);
    mcopy(n, A, gln);
    orc_zgemm('n', 'n', n, GN0, H10, B);
    orc_zgemm('n', 'n', n, B, A, C);
    madd(n, Gn, Gn, C);
    orc_zgemm('n', 'c', n, A, H10, B);
    orc_zgemm('n', 'c', n, B, GN0, C);
    madd(n, Gn, Gn, C);
    if (ndens) mcopy(n, blkw(ndens, n, l), Gn);
    if (pdens) {
      mantiherm(n, T1, G00);
      madd(n, blkw(pdens, n, l), Gn, T1);
    }
  }

  /* tre, green_rgf.f90:4151-4160 */
  mantiherm(n, Gp, G00);
  madd(n, Gp, Gn, Gp);
  {
    double fL = orc_ferm((E - mul) / (p->boltz * temp_left_tre));
    mantiherm(n, T1, sigmal);
    for (size_t k = 0; k < bs; ++k) A[k] = -T1[k] * fL;
    orc_zgemm('n', 'n', n, A, Gp, B);
    for (size_t k = 0; k < bs; ++k) A[k] = -T1[k] * (fL - 1.0);
    orc_zgemm('n', 'n', n, A, Gn, C);
    double tim = 0.0;
    for (int i = 0; i < n; ++i) {
      size_t d = (size_t)i + (size_t)i * n;
      tim = tim + (RE(B, d) - RE(C, d));
    }
    if (tre) *tre = tim;
  }
  if (diag) *diag = dg;
  free(Gl); free(Gln);
  free(H00); free(H10); free(A); free(B); free(C); free(D); free(G00); free(GBB);
  free(GN0); free(Gn); free(Gp); free(sig); free(sigmal); free(sigmar); free(H10h); free(T1);
  free(H00dev);
  (void)mzero;
}

void orc_rgf(const orc_params *p, const double temp[2], int nm, int nx, double E,
             const double mu[2], const double *Hii, const double *H1i,
             const double *sigL, const double *sigG, const double *sigR,
             double *ndens, double *pdens, double *ldos, double *tr, double *tre,
             double *cur, double *GRi1, double *GLi1, double *GGi1, orc_diag *diag) {
  rgf_impl(p, temp, nm, nx, E, mu, Hii, H1i, sigL, sigG, sigR, ndens, pdens, ldos, tr, tre, cur,
           GRi1, GLi1, GGi1, diag, NULL, NULL);
}

void orc_rgf_fixed_leads(const orc_params *p, const double temp[2], int nm, int nx, double E,
                         const double mu[2], const double *Hii, const double *H1i,
                         const double *sigL, const double *sigR, const double *leadL_H00,
                         const double *leadR_H00, double *ndens, double *ldos, double *tr,
                         double *tre, orc_diag *diag) {
  rgf_impl(p, temp, nm, nx, E, mu, Hii, H1i, sigL, NULL, sigR, ndens, NULL, ldos, tr, tre, NULL,
           NULL, NULL, NULL, diag, leadL_H00, leadR_H00);
}

void orc_rgf_energies(const orc_params *p, const double temp[2], int nm, int nx,
                      int nE, const double *E, const double mu[2],
                      const double *Hii, const double *H1i,
                      const double *sigL, const double *sigG, const double *sigR,
                      long long sig_stride_e,
                      double *ndens, double *pdens, double *ldos, double *tr,
                      double *tre, double *cur, double *GRi1, double *GLi1,
                      double *GGi1, orc_diag *diag, int nthreads) {
  const size_t cs = 2 * (size_t)nm * nm * (size_t)nx; /* doubles per complex (nm,nm,nx) */
  const size_t rs = (size_t)nm * nm * (size_t)nx;
  (void)nthreads;
#ifdef _OPENMP
  if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel for schedule(dynamic, 1)
#endif
  for (int e = 0; e < nE; ++e) {
    size_t so = 2 * (size_t)sig_stride_e * (size_t)e; /* stride is in complex elements */
    orc_rgf(p, temp, nm, nx, E[e], mu, Hii, H1i,
            sigL ? sigL + so : NULL, sigG ? sigG + so : NULL, sigR ? sigR + so : NULL,
            ndens ? ndens + cs * e : NULL, pdens ? pdens + cs * e : NULL,
            ldos ? ldos + cs * e : NULL, tr ? tr + e : NULL, tre ? tre + e : NULL,
            cur ? cur + rs * e : NULL, GRi1 ? GRi1 + cs * e : NULL,
            GLi1 ? GLi1 + cs * e : NULL, GGi1 ? GGi1 + cs * e : NULL,
            diag ? diag + e : NULL);
  }
}

/* --------------------------------------------- block-tridiagonal Sigma */
/* d = a + s*b^H (b may be NULL) */
static void maddh(int n, double *d, const double *a, double s, const double *b) {
  for (int j = 0; j < n; ++j)
    for (int i = 0; i < n; ++i) {
      size_t p = (size_t)i + (size_t)j * n, q = (size_t)j + (size_t)i * n;
      double br = b ? RE(b, q) : 0.0, bi = b ? -IM(b, q) : 0.0;
      RE(d, p) = RE(a, p) + s * br;
      IM(d, p) = IM(a, p) + s * bi;
    }
}
/* d = a + m - m^H (adds the anti-Hermitian part of m twice over; d may alias a) */
static void maddanti(int n, double *d, const double *a, const double *m) {
  for (int j = 0; j < n; ++j)
    for (int i = 0; i < n; ++i) {
      size_t p = (size_t)i + (size_t)j * n, q = (size_t)j + (size_t)i * n;
      RE(d, p) = RE(a, p) + RE(m, p) - RE(m, q);
      IM(d, p) = IM(a, p) + IM(m, p) + IM(m, q);
    }
}

static void offdiag_impl(const orc_params *p, const double temp[2], int nm, int nx, double E,
                         const double mu[2], const double *Hii, const double *H1i,
                         const double *sigL, const double *sigR, const double *sigR_lo,
                         const double *sigR_up, const double *sigL_up, double *ndens,
                         double *pdens, double *ldos, double *tr, double *tre, double *cur,
                         double *GRi1, double *GLi1, double *GGi1, double *GRi1_lo, orc_diag *diag,
                         const double *leadL_H00, const double *leadR_H00);

void orc_rgf_offdiag(const orc_params *p, const double temp[2], int nm, int nx, double E,
                     const double mu[2], const double *Hii, const double *H1i,
                     const double *sigL, const double *sigR, const double *sigR_lo,
                     const double *sigR_up, const double *sigL_up, double *ndens,
                     double *pdens, double *ldos, double *tr, double *tre, double *cur,
                     double *GRi1, double *GLi1, double *GGi1, double *GRi1_lo, orc_diag *diag) {
  offdiag_impl(p, temp, nm, nx, E, mu, Hii, H1i, sigL, sigR, sigR_lo, sigR_up, sigL_up, ndens, pdens, ldos, tr,
               tre, cur, GRi1, GLi1, GGi1, GRi1_lo, diag, NULL, NULL);
}

void orc_rgf_offdiag_fixed_leads(const orc_params *p, const double temp[2], int nm, int nx, double E,
                                 const double mu[2], const double *Hii, const double *H1i,
                                 const double *sigL, const double *sigR, const double *sigR_lo,
                                 const double *sigR_up, const double *sigL_up, const double *leadL_H00,
                                 const double *leadR_H00, double *ndens, double *tr) {
  offdiag_impl(p, temp, nm, nx, E, mu, Hii, H1i, sigL, sigR, sigR_lo, sigR_up, sigL_up, ndens, NULL, NULL, tr,
               NULL, NULL, NULL, NULL, NULL, NULL, NULL, leadL_H00, leadR_H00);
}

static void offdiag_impl(const orc_params *p, const double temp[2], int nm, int nx, double E,
                         const double mu[2], const double *Hii, const double *H1i,
                         const double *sigL, const double *sigR, const double *sigR_lo,
                         const double *sigR_up, const double *sigL_up, double *ndens,
                         double *pdens, double *ldos, double *tr, double *tre, double *cur,
                         double *GRi1, double *GLi1, double *GGi1, double *GRi1_lo, orc_diag *diag,
                         const double *leadL_H00, const double *leadR_H00) {
  const int n = nm, N = nx;
  const size_t bs = 2 * (size_t)n * n, nn = (size_t)n * n;
  const double zr = E, zi = p->eta_dev;
  orc_diag dg = {0, 0, 0, 0};
  double *g = (double *)calloc(bs * (size_t)N, sizeof(double));
  double *gn = (double *)calloc(bs * (size_t)N, sizeof(double));
  double *H00 = mnew(n), *Th = mnew(n), *G00 = mnew(n), *A = mnew(n), *B = mnew(n), *C = mnew(n);
  double *S = mnew(n), *V = mnew(n), *d = mnew(n), *u = mnew(n), *sigmal = mnew(n), *sigmar = mnew(n);
  double *GamL = mnew(n), *GamR = mnew(n), *G = mnew(n), *Gn = mnew(n), *Q = mnew(n), *R = mnew(n);
  double *Glo = mnew(n), *M = mnew(n), *gs = mnew(n), *QGn = mnew(n), *Rd = mnew(n), *Gnn = mnew(n);
  double *Gll = mnew(n), *Gp = mnew(n), *Z0 = mnew(n); /* Z0 stays zero */
#define SB(a, l) ((a) ? blk((a), n, (l)) : NULL)
  /* leads: onsite Hii + sigR at the edge slabs, couplings H1i(:,:,1), H1i(:,:,nx) */
  if (sigR) madd(n, H00, blk(Hii, n, 0), blk(sigR, n, 0)); else mcopy(n, H00, blk(Hii, n, 0));
  if (leadL_H00) mcopy(n, H00, leadL_H00);
  maddh(n, Th, G00, 1.0, blk(H1i, n, 0)); /* G00 is zero here: Th = H1i_1^H */
  {
    int hit = 0;
    dg.iters_left = orc_sancho(n, E, p->eta_lead, p->dec_tol, p->dec_nmax, H00, Th, G00, NULL, &hit,
                               &dg.inv_failures);
    if (hit) dg.nmax_hit |= 1;
  }
  orc_zgemm('n', 'n', n, blk(H1i, n, 0), G00, A);
  orc_zgemm('n', 'c', n, A, blk(H1i, n, 0), sigmal);
  if (sigR) madd(n, H00, blk(Hii, n, N - 1), blk(sigR, n, N - 1)); else mcopy(n, H00, blk(Hii, n, N - 1));
  if (leadR_H00) mcopy(n, H00, leadR_H00);
  {
    int hit = 0;
    dg.iters_right = orc_sancho(n, E, p->eta_lead, p->dec_tol, p->dec_nmax, H00, blk(H1i, n, N - 1), G00,
                                NULL, &hit, &dg.inv_failures);
    if (hit) dg.nmax_hit |= 2;
  }
  orc_zgemm('c', 'n', n, blk(H1i, n, N - 1), G00, A);
  orc_zgemm('n', 'n', n, A, blk(H1i, n, N - 1), sigmar);
  const double fL = orc_ferm((E - mu[0]) / (p->boltz * temp[0]));
  const double fR = orc_ferm((E - mu[1]) / (p->boltz * temp[1]));
  mantiherm(n, GamL, sigmal);
  mantiherm(n, GamR, sigmar);
  for (size_t k = 0; k < bs; ++k) { GamL[k] = -GamL[k]; GamR[k] = -GamR[k]; } /* i Gamma */

  /* forward: g_l = (z - H_l - SR_l - d_l g_{l-1} u_l)^-1,
   *          g<_l = g_l (S<_l + d_l g<_{l-1} d_l^H + V - V^H) g_l^H, V = d_l g_{l-1} s_l */
  for (int l = 0; l < N; ++l) {
    double *gg = (l == N - 1) ? G : blkw(g, n, l);
    double *gng = (l == N - 1) ? Gn : blkw(gn, n, l);
    mzminus(n, A, zr, zi, blk(Hii, n, l), SB(sigR, l));
    for (size_t k = 0; k < bs; ++k) S[k] = sigL ? blk(sigL, n, l)[k] : 0.0;
    if (l == 0)
      for (size_t k = 0; k < bs; ++k) { A[k] -= sigmal[k]; S[k] += fL * GamL[k]; }
    if (l == N - 1)
      for (size_t k = 0; k < bs; ++k) { A[k] -= sigmar[k]; S[k] += fR * GamR[k]; }
    if (l > 0) {
      if (sigR_lo) madd(n, d, blk(H1i, n, l), blk(sigR_lo, n, l)); else mcopy(n, d, blk(H1i, n, l));
      maddh(n, u, sigR_up ? blk(sigR_up, n, l) : Z0, 1.0, blk(H1i, n, l));
      orc_zgemm('n', 'n', n, d, blk(g, n, l - 1), B);  /* d g */
      orc_zgemm('n', 'n', n, B, u, C);
      for (size_t k = 0; k < bs; ++k) A[k] -= C[k];
      orc_zgemm('n', 'n', n, d, blk(gn, n, l - 1), C);
      orc_zgemm('n', 'c', n, C, d, V);
      madd(n, S, S, V);
      if (sigL_up) {
        orc_zgemm('n', 'n', n, B, blk(sigL_up, n, l), V);
        maddanti(n, S, S, V);
      }
    }
    if (orc_zinv(n, A)) ++dg.inv_failures;
    mcopy(n, gg, A);
    orc_zgemm('n', 'n', n, A, S, B);
    orc_zgemm('n', 'c', n, B, A, gng);
  }
  mantiherm(n, Gp, G);
  madd(n, Gp, Gn, Gp);
  if (ldos) mcopy(n, blkw(ldos, n, N - 1), G);
  if (ndens) mcopy(n, blkw(ndens, n, N - 1), Gn);
  if (pdens) mcopy(n, blkw(pdens, n, N - 1), Gp);
  if (cur) memset(cur + nn * (size_t)(N - 1), 0, nn * sizeof(double));
  {
    /* tr = -Re Tr[iGam_R fR G>_NN - iGam_R (fR - 1) G<_NN] */
    orc_zgemm('n', 'n', n, GamR, Gp, B);
    orc_zgemm('n', 'n', n, GamR, Gn, C);
    double t = 0.0;
    for (int i = 0; i < n; ++i) {
      size_t k = (size_t)i + (size_t)i * n;
      t -= fR * RE(B, k) - (fR - 1.0) * RE(C, k);
    }
    if (tr) *tr = t;
  }
  /* backward, pair (l, l+1): u = H_{l,l+1} + SR_{l,l+1}, d = H_{l+1,l} + SR_{l+1,l}, s = S<_{l,l+1};
   * Q = g u, R = Q G = G_{l,l+1}, G_{l+1,l} = G d g,
   * G<_{l,l+1} = g< d^H G^H + g s G^H + Q G<,
   * G_ll = g + R d g, G<_ll = g< + Q G< Q^H + (R d g< - h.c.) + (g s R^H - h.c.) */
  for (int l = N - 2; l >= 0; --l) {
    const double *gl = blk(g, n, l), *gln = blk(gn, n, l), *T = blk(H1i, n, l + 1);
    if (sigR_lo) madd(n, d, T, blk(sigR_lo, n, l + 1)); else mcopy(n, d, T);
    maddh(n, u, sigR_up ? blk(sigR_up, n, l + 1) : Z0, 1.0, T);
    orc_zgemm('n', 'n', n, gl, u, Q);
    orc_zgemm('n', 'n', n, Q, G, R);
    orc_zgemm('n', 'n', n, G, d, B);
    orc_zgemm('n', 'n', n, B, gl, Glo);
    orc_zgemm('n', 'n', n, Q, Gn, QGn);
    if (sigL_up) orc_zgemm('n', 'n', n, gl, blk(sigL_up, n, l + 1), gs); else mzero(n, gs);
    orc_zgemm('n', 'c', n, gln, d, M);
    madd(n, M, M, gs);
    orc_zgemm('n', 'c', n, M, G, B);
    madd(n, Gnn, B, QGn); /* G<_{l,l+1} */
    if (GRi1) mcopy(n, blkw(GRi1, n, l), R);
    if (GRi1_lo) mcopy(n, blkw(GRi1_lo, n, l), Glo);
    if (GLi1) mcopy(n, blkw(GLi1, n, l), Gnn);
    if (GGi1) {
      double *o = blkw(GGi1, n, l);
      madd(n, o, Gnn, R);
      maddh(n, o, o, -1.0, Glo);
    }
    if (cur) {
      /* Re(H_{l,l+1} G<_{l+1,l}) - Re(H_{l+1,l} G<_{l,l+1}), G<_{l+1,l} = -(G<_{l,l+1})^H */
      orc_zgemm('c', 'c', n, T, Gnn, A);
      orc_zgemm('n', 'n', n, T, Gnn, B);
      double *cl = cur + nn * (size_t)l;
      for (size_t k = 0; k < nn; ++k) cl[k] = -RE(A, k) - RE(B, k);
    }
    orc_zgemm('n', 'n', n, R, d, Rd);
    orc_zgemm('n', 'n', n, Rd, gl, A);
    madd(n, Gll, gl, A);
    orc_zgemm('n', 'c', n, QGn, Q, A);
    madd(n, B, gln, A);
    orc_zgemm('n', 'n', n, Rd, gln, A);
    maddanti(n, B, B, A);
    orc_zgemm('n', 'c', n, gs, R, A);
    maddanti(n, B, B, A);
    mcopy(n, G, Gll);
    mcopy(n, Gn, B);
    mantiherm(n, Gp, G);
    madd(n, Gp, Gn, Gp);
    if (ldos) mcopy(n, blkw(ldos, n, l), G);
    if (ndens) mcopy(n, blkw(ndens, n, l), Gn);
    if (pdens) mcopy(n, blkw(pdens, n, l), Gp);
  }
  {
    /* tre = Re Tr[iGam_L fL G>_11 - iGam_L (fL - 1) G<_11] */
    orc_zgemm('n', 'n', n, GamL, Gp, B);
    orc_zgemm('n', 'n', n, GamL, Gn, C);
    double t = 0.0;
    for (int i = 0; i < n; ++i) {
      size_t k = (size_t)i + (size_t)i * n;
      t += fL * RE(B, k) - (fL - 1.0) * RE(C, k);
    }
    if (tre) *tre = t;
  }
#undef SB
  if (diag) *diag = dg;
  free(g); free(gn);
  free(H00); free(Th); free(G00); free(A); free(B); free(C); free(S); free(V); free(d); free(u);
  free(sigmal); free(sigmar); free(GamL); free(GamR); free(G); free(Gn); free(Q); free(R); free(Glo);
  free(M); free(gs); free(QGn); free(Rd); free(Gnn); free(Gll); free(Gp); free(Z0);
}


/* ------------------------------------------------------------ overlap */
/* green_RGF_CMS, green_rgf.f90:4167-4383: green_RGF_RS (compat arithmetic,
 * same product order) with the overlap blocks Sii: z S_l instead of z,
 * Sigma^R_l S_l and Sigma^<_l S_l instead of Sigma^R_l, Sigma^<_l (4198-4207,
 * 4222-4237), the decimation with S00 (4200, 4244), and the G^> recursion of
 * its backward sweep (4356-4374) whose left-connected G^> (Glp) is never
 * filled, so pdens_l = g_l H^H pdens_{l+1} H g_l^H for l < nx and GGi1 reads
 * that recursion's value (4306-4315).  tr uses TEMP(1) and tre TEMP(2) as in
 * the reference. */
void orc_rgf_cms(const orc_params *p, const double temp[2], int nm, int nx, double E,
                 const double mu[2], const double *Sii, const double *Hii, const double *H1i,
                 const double *sigL, const double *sigG, const double *sigR,
                 double *ndens, double *pdens, double *ldos, double *tr, double *tre,
                 double *cur, double *GRi1, double *GLi1, double *GGi1, orc_diag *diag) {
  (void)sigG;
  const int n = nm;
  const size_t bs = 2 * (size_t)n * n;
  const double mul = mu[0], mur = mu[1], templ = temp[0], tempr = temp[1];
  const double zr = E, zi = p->eta_dev;
  orc_diag dg = {0, 0, 0, 0};
  double *Gl = (double *)calloc(bs * (size_t)nx, sizeof(double));
  double *Gln = (double *)calloc(bs * (size_t)nx, sizeof(double));
  double *S00 = mnew(n), *H00 = mnew(n), *H10 = mnew(n), *H10h = mnew(n), *A = mnew(n), *B = mnew(n);
  double *C = mnew(n), *D = mnew(n), *G00 = mnew(n), *GBB = mnew(n), *GN0 = mnew(n), *Gn = mnew(n);
  double *Gp = mnew(n), *sig = mnew(n), *sigmal = mnew(n), *sigmar = mnew(n), *T1 = mnew(n);
  if (ldos) memset(ldos, 0, bs * (size_t)nx * sizeof(double));
  if (ndens) memset(ndens, 0, bs * (size_t)nx * sizeof(double));
  if (pdens) memset(pdens, 0, bs * (size_t)nx * sizeof(double));
  if (cur) memset(cur, 0, (size_t)n * n * (size_t)nx * sizeof(double));
#define ZERO_OR(a, l) ((a) ? blk((a), n, (l)) : NULL)
  /* H00 = Hii + sigR S00 ; B = sigL S00 */
#define ONSITE(l)                                                        \
  do {                                                                   \
    mcopy(n, S00, blk(Sii, n, (l)));                                     \
    if (sigR) orc_zgemm('n', 'n', n, blk(sigR, n, (l)), S00, B);         \
    else mzero(n, B);                                                    \
    madd(n, H00, blk(Hii, n, (l)), B);                                   \
  } while (0)
  /* left contact, 4196-4216 */
  ONSITE(0);
  mcopy(n, H10, blk(H1i, n, 0));
  for (int j = 0; j < n; ++j)
    for (int i = 0; i < n; ++i) {
      size_t a = (size_t)i + (size_t)j * n, b = (size_t)j + (size_t)i * n;
      RE(H10h, a) = RE(H10, b);
      IM(H10h, a) = -IM(H10, b);
    }
  {
    int hit = 0;
    dg.iters_left = orc_sancho_s(n, E, p->eta_lead, p->dec_tol, p->dec_nmax, S00, H00, H10h, G00, GBB, &hit,
                                 &dg.inv_failures);
    if (hit) dg.nmax_hit |= 1;
  }
  orc_zgemm('n', 'n', n, H10, G00, A);
  orc_zgemm('n', 'c', n, A, H10, sigmal);
  if (sigL) orc_zgemm('n', 'n', n, blk(sigL, n, 0), S00, B); else mzero(n, B);
  {
    double f = orc_ferm((E - mul) / (p->boltz * templ));
    mantiherm(n, sig, sigmal);
    for (size_t k = 0; k < bs; ++k) sig[k] = -sig[k] * f + B[k];
  }
  mzs(n, A, zr, zi, S00, H00);
  for (size_t k = 0; k < bs; ++k) A[k] -= sigmal[k];
  if (orc_zinv(n, A)) ++dg.inv_failures;
  mcopy(n, blkw(Gl, n, 0), A);
  orc_zgemm('n', 'n', n, A, sig, B);
  orc_zgemm('n', 'c', n, B, A, C);
  mcopy(n, blkw(Gln, n, 0), C);
  /* forward, 4217-4241 */
  for (int l = 1; l < nx - 1; ++l) {
    ONSITE(l);
    mcopy(n, H10, blk(H1i, n, l));
    orc_zgemm('n', 'n', n, H10, blk(Gl, n, l - 1), B);
    orc_zgemm('n', 'c', n, B, H10, C);
    mzs(n, A, zr, zi, S00, H00);
    for (size_t k = 0; k < bs; ++k) A[k] -= C[k];
    if (orc_zinv(n, A)) ++dg.inv_failures;
    mcopy(n, blkw(Gl, n, l), A);
    mcopy(n, sig, blk(Gln, n, l - 1));
    orc_zgemm('n', 'n', n, H10, sig, B);
    orc_zgemm('n', 'c', n, B, H10, C);
    if (sigL) orc_zgemm('n', 'n', n, blk(sigL, n, l), S00, B); else mzero(n, B);
    madd(n, C, C, B);
    orc_zgemm('n', 'n', n, A, C, B);
    orc_zgemm('n', 'c', n, B, A, Gn);
    mcopy(n, blkw(Gln, n, l), Gn);
  }
  /* right contact, 4242-4287 */
  ONSITE(nx - 1);
  mcopy(n, H10, blk(H1i, n, nx - 1));
  {
    int hit = 0;
    dg.iters_right = orc_sancho_s(n, E, p->eta_lead, p->dec_tol, p->dec_nmax, S00, H00, H10, G00, GBB, &hit,
                                  &dg.inv_failures);
    if (hit) dg.nmax_hit |= 2;
  }
  orc_zgemm('c', 'n', n, H10, G00, A);
  orc_zgemm('n', 'n', n, A, H10, sigmar);
  orc_zgemm('n', 'n', n, H10, blk(Gl, n, nx - 2), B);
  orc_zgemm('n', 'c', n, B, H10, C);
  mzs(n, G00, zr, zi, S00, H00);
  for (size_t k = 0; k < bs; ++k) G00[k] -= sigmar[k] + C[k];
  if (orc_zinv(n, G00)) ++dg.inv_failures;
  if (ldos) mcopy(n, blkw(ldos, n, nx - 1), G00);
  mcopy(n, sig, blk(Gln, n, nx - 2));
  orc_zgemm('n', 'n', n, H10, sig, B);
  orc_zgemm('n', 'c', n, B, H10, C);
  if (sigL) orc_zgemm('n', 'n', n, blk(sigL, n, nx - 1), S00, B); else mzero(n, B);
  const double fR = orc_ferm((E - mur) / (p->boltz * templ));
  mantiherm(n, T1, sigmar);
  for (size_t k = 0; k < bs; ++k) sig[k] = -T1[k] * fR + C[k] + B[k];
  orc_zgemm('n', 'n', n, G00, sig, B);
  orc_zgemm('n', 'c', n, B, G00, Gn);
  if (ndens) mcopy(n, blkw(ndens, n, nx - 1), Gn);
  mantiherm(n, Gp, G00);
  madd(n, Gp, Gn, Gp);
  if (pdens) mcopy(n, blkw(pdens, n, nx - 1), Gp);
  for (size_t k = 0; k < bs; ++k) A[k] = -T1[k] * fR;
  orc_zgemm('n', 'n', n, A, Gp, B);
  for (size_t k = 0; k < bs; ++k) A[k] = -T1[k] * (fR - 1.0);
  orc_zgemm('n', 'n', n, A, Gn, C);
  {
    double tim = 0.0;
    for (int i = 0; i < n; ++i) {
      size_t d = (size_t)i + (size_t)i * n;
      tim = tim - (RE(B, d) - RE(C, d));
    }
    if (tr) *tr = tim;
  }
  /* backward, 4290-4379 (H1i(:,:,l) as the l,l+1 coupling, as the reference) */
  for (int l = nx - 2; l >= 0; --l) {
    mcopy(n, H10, blk(H1i, n, l));
    const double *gl = blk(Gl, n, l), *gln = blk(Gln, n, l);
    double *curl = cur ? cur + (size_t)n * n * (size_t)l : NULL;
    orc_zgemm('n', 'c', n, H10, gl, B);
    orc_zgemm('n', 'n', n, Gn, B, C);
    orc_zgemm('n', 'n', n, H10, gln, B);
    orc_zgemm('n', 'n', n, G00, B, A);
    madd(n, B, C, A);
    orc_zgemm('c', 'n', n, H10, B, A);
    if (curl) for (size_t k = 0; k < (size_t)n * n; ++k) curl[k] = RE(A, k);
    orc_zgemm('n', 'c', n, gl, H10, B);
    orc_zgemm('n', 'n', n, B, Gn, C);
    orc_zgemm('n', 'c', n, gln, H10, B);
    orc_zgemm('n', 'c', n, B, G00, A);
    madd(n, B, C, A);
    if (GLi1) mcopy(n, blkw(GLi1, n, l), B);
    orc_zgemm('n', 'n', n, H10, B, A);
    if (curl) for (size_t k = 0; k < (size_t)n * n; ++k) curl[k] -= RE(A, k);
    if (GGi1) {
      /* g^H H^H G>_{l+1} + (Glp == 0) terms */
      orc_zgemm('c', 'c', n, gl, H10, B);
      orc_zgemm('n', 'n', n, B, Gp, C);
      mcopy(n, blkw(GGi1, n, l), C);
    }
    mcopy(n, D, gl);
    orc_zgemm('n', 'c', n, D, H10, B);
    orc_zgemm('n', 'n', n, B, G00, GN0);
    if (GRi1) mcopy(n, blkw(GRi1, n, l), GN0);
    orc_zgemm('n', 'n', n, GN0, H10, A);
    orc_zgemm('n', 'n', n, A, D, C);
    madd(n, G00, gl, C);
    if (ldos) mcopy(n, blkw(ldos, n, l), G00);
    mcopy(n, A, Gn);
    orc_zgemm('n', 'c', n, D, H10, B);
    orc_zgemm('n', 'n', n, B, A, C);
    orc_zgemm('n', 'n', n, C, H10, A);
    orc_zgemm('n', 'c', n, A, D, C);
    madd(n, Gn, gln, C);
    mcopy(n, A, gln);
    orc_zgemm('n', 'n', n, GN0, H10, B);
    orc_zgemm('n', 'n', n, B, A, C);
    madd(n, Gn, Gn, C);
    orc_zgemm('n', 'c', n, A, H10, B);
    orc_zgemm('n', 'c', n, B, GN0, C);
    madd(n, Gn, Gn, C);
    /* G>_ll recursion with Glp == 0: g H^H G> H g^H */
    orc_zgemm('n', 'c', n, D, H10, B);
    orc_zgemm('n', 'n', n, B, Gp, C);
    orc_zgemm('n', 'n', n, C, H10, A);
    orc_zgemm('n', 'c', n, A, D, Gp);
    if (ndens) mcopy(n, blkw(ndens, n, l), Gn);
    if (pdens) mcopy(n, blkw(pdens, n, l), Gp);
  }
  /* tre, 4380-4390: G^> from the identity, TEMP(2) */
  mantiherm(n, Gp, G00);
  madd(n, Gp, Gn, Gp);
  {
    double fL = orc_ferm((E - mul) / (p->boltz * tempr));
    mantiherm(n, T1, sigmal);
    for (size_t k = 0; k < bs; ++k) A[k] = -T1[k] * fL;
    orc_zgemm('n', 'n', n, A, Gp, B);
    for (size_t k = 0; k < bs; ++k) A[k] = -T1[k] * (fL - 1.0);
    orc_zgemm('n', 'n', n, A, Gn, C);
    double tim = 0.0;
    for (int i = 0; i < n; ++i) {
      size_t d = (size_t)i + (size_t)i * n;
      tim = tim + (RE(B, d) - RE(C, d));
    }
    if (tre) *tre = tim;
  }
#undef ONSITE
#undef ZERO_OR
  if (diag) *diag = dg;
  free(Gl); free(Gln);
  free(S00); free(H00); free(H10); free(H10h); free(A); free(B); free(C); free(D); free(G00); free(GBB);
  free(GN0); free(Gn); free(Gp); free(sig); free(sigmal); free(sigmar); free(T1);
}
