/*
 * sdt_oracle.c -- CPU oracle for the TCE CCSD(T) perturbative-triples path.
 *
 * TEST INFRASTRUCTURE ONLY (the checker).  See sdt_oracle.h for the list of
 * reference routines restated here and for what pins each of them.
 *
 * Built with -ffp-contract=off so that every "t3 = t3 -/+ a*b" is a rounded
 * multiply followed by a rounded add, the same two roundings the reference
 * Fortran performs; per-element summation order over the contraction index
 * follows the reference loops, so the kernels agree bit for bit with the
 * reference kernels compiled without OpenMP.
 */
#include "sdt_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define RNG(t, b) ((t)->range[(b) - 1])
#define SPN(t, b) ((t)->spin[(b) - 1])
#define SYM(t, b) ((t)->sym[(b) - 1])
#define OFS(t, b) ((t)->offset[(b) - 1])

/* ---------------------------------------------------------------- tiles -- */

/* tce_tile.F:330-355: nblocks = ceil(n/isize); range(k) = k*n/nblocks - sum
 * of the previous ranges (integer division). */
long ora_tile_ranges(long n, long tilesize, long *ranges) {
  if (n <= 0 || tilesize <= 0) return 0;
  long nb = n / tilesize;
  if (n > tilesize * nb) nb++;
  long acc = 0;
  for (long k = 1; k <= nb; k++) {
    long r = k * n / nb - acc;
    if (ranges) ranges[k - 1] = r;
    acc += r;
  }
  return nb;
}

/* ----------------------------------------------------- offsets and hash -- */

/* tce_t1_offset.F:20-56: blocks t(p h), p outer, h inner. */
long ora_t1_offset(const ora_tiles_t *t, long *hash, long *size) {
  long noab = t->noab, nvab = t->nvab;
  long length = 0;
  for (long p5b = noab + 1; p5b <= noab + nvab; p5b++)
    for (long h6b = 1; h6b <= noab; h6b++)
      if (SPN(t, p5b) == SPN(t, h6b) && (SYM(t, p5b) ^ SYM(t, h6b)) == 0 &&
          (!t->restricted || SPN(t, p5b) + SPN(t, h6b) != 4))
        length++;
  if (hash) {
    long addr = 0, sz = 0;
    hash[0] = length;
    for (long p5b = noab + 1; p5b <= noab + nvab; p5b++)
      for (long h6b = 1; h6b <= noab; h6b++)
        if (SPN(t, p5b) == SPN(t, h6b) && (SYM(t, p5b) ^ SYM(t, h6b)) == 0 &&
            (!t->restricted || SPN(t, p5b) + SPN(t, h6b) != 4)) {
          addr++;
          hash[addr] = h6b - 1 + noab * (p5b - noab - 1);
          hash[length + addr] = sz;
          sz += RNG(t, p5b) * RNG(t, h6b);
        }
    if (size) *size = sz;
  }
  return length;
}

/* tce_t2_offset.F:1-70: blocks t(p1 p2 h3 h4) with p1<=p2, h3<=h4. */
long ora_t2_offset(const ora_tiles_t *t, long *hash, long *size) {
  long noab = t->noab, nvab = t->nvab;
  long length = 0, addr = 0, sz = 0;
  for (int pass = 0; pass < 2; pass++) {
    if (pass == 1) {
      if (!hash) break;
      hash[0] = length;
    }
    for (long p1b = noab + 1; p1b <= noab + nvab; p1b++)
      for (long p2b = p1b; p2b <= noab + nvab; p2b++)
        for (long h3b = 1; h3b <= noab; h3b++)
          for (long h4b = h3b; h4b <= noab; h4b++) {
            if (SPN(t, p1b) + SPN(t, p2b) != SPN(t, h3b) + SPN(t, h4b)) continue;
            if ((SYM(t, p1b) ^ SYM(t, p2b) ^ SYM(t, h3b) ^ SYM(t, h4b)) != 0) continue;
            if (t->restricted &&
                SPN(t, p1b) + SPN(t, p2b) + SPN(t, h3b) + SPN(t, h4b) == 8)
              continue;
            if (pass == 0) {
              length++;
            } else {
              addr++;
              hash[addr] = h4b - 1 + noab * (h3b - 1 + noab * (p2b - noab - 1 +
                                                               nvab * (p1b - noab - 1)));
              hash[length + addr] = sz;
              sz += RNG(t, p1b) * RNG(t, p2b) * RNG(t, h3b) * RNG(t, h4b);
            }
          }
  }
  if (size) *size = sz;
  return length;
}

/* Integral classes the (T) path reads from v(g3 g4 g1 g2):
 *   d1  v(h7 p6 h2 h3)  -> (hole, particle, hole, hole)
 *   d2  v(p5 p6 h3 p7)  -> (particle, particle, hole, particle)
 *   s1  v(p5 p6 h2 h3)  -> (particle, particle, hole, hole) */
static int v2_triples_class(long noab, long g3b, long g4b, long g1b, long g2b) {
  int p3 = g3b > noab, p4 = g4b > noab, p1 = g1b > noab, p2 = g2b > noab;
  if (!p3 && p4 && !p1 && !p2) return 1;
  if (p3 && p4 && !p1 && p2) return 1;
  if (p3 && p4 && !p1 && !p2) return 1;
  return 0;
}

/* tce_mo2e_offset.F:1-70: blocks v(g3 g4 g1 g2), g3<=g4, g1<=g2. */
long ora_v2_offset(const ora_tiles_t *t, int subset, long *hash, long *size) {
  long noab = t->noab, nvab = t->nvab, N = noab + nvab;
  long length = 0, addr = 0, sz = 0;
  for (int pass = 0; pass < 2; pass++) {
    if (pass == 1) {
      if (!hash) break;
      hash[0] = length;
    }
    for (long g3b = 1; g3b <= N; g3b++)
      for (long g4b = g3b; g4b <= N; g4b++)
        for (long g1b = 1; g1b <= N; g1b++)
          for (long g2b = g1b; g2b <= N; g2b++) {
            if (SPN(t, g3b) + SPN(t, g4b) != SPN(t, g1b) + SPN(t, g2b)) continue;
            if ((SYM(t, g3b) ^ SYM(t, g4b) ^ SYM(t, g1b) ^ SYM(t, g2b)) != 0) continue;
            if (t->restricted &&
                SPN(t, g3b) + SPN(t, g4b) + SPN(t, g1b) + SPN(t, g2b) == 8)
              continue;
            if (subset && !v2_triples_class(noab, g3b, g4b, g1b, g2b)) continue;
            if (pass == 0) {
              length++;
            } else {
              addr++;
              hash[addr] = g2b - 1 + N * (g1b - 1 + N * (g4b - 1 + N * (g3b - 1)));
              hash[length + addr] = sz;
              sz += RNG(t, g3b) * RNG(t, g4b) * RNG(t, g1b) * RNG(t, g2b);
            }
          }
  }
  if (size) *size = sz;
  return length;
}

/* tce_hash.F:271-316: bisection on hash(2..L+1) while the window is wider
 * than 4 entries, then a linear scan of the window. */
long ora_hash(const long *hash, long key) {
  long length = hash[0];
  long less = 1, more = length, middle;
  if (length <= 0) return -1;
  for (;;) {
    if (more - less <= 4) {
      middle = -1;
      for (long i = less; i <= more; i++)
        if (hash[i] == key) middle = i;
      if (middle == -1) return -1;
      break;
    }
    middle = (less + more) / 2;
    if (hash[middle] == key) break;
    if (hash[middle] > key)
      more = middle;
    else
      less = middle;
  }
  return hash[length + middle];
}

/* get_hash_block.F / get_hash_block_ma.F / get_block.F: copy of one block. */
static int get_hash_block(const double *file, double *array, long size,
                          const long *hash, long key) {
  long off = ora_hash(hash, key);
  if (off < 0) return -1;
  memcpy(array, file + off, (size_t)size * sizeof(double));
  return 0;
}

/* ---------------------------------------------------------------- sorts -- */

/* sort/tce_sort.F:15-49.  unsorted is (a,b) with b fastest; the result has
 * index (i,j) order with the j-th source index fastest. */
void ora_sort_2(const double *unsorted, double *sorted, long a, long b, long i,
                long j, double factor) {
  long jd[3] = {0, a, b};
  long id[3];
  long ia = 0;
  for (long j1 = 1; j1 <= a; j1++) {
    id[1] = j1;
    for (long j2 = 1; j2 <= b; j2++) {
      id[2] = j2;
      ia++;
      long ib = id[j] + jd[j] * (id[i] - 1);
      sorted[ib - 1] = unsorted[ia - 1] * factor;
    }
  }
}

/* sort/tce_sort.F:50-125.  The three loop orders of the reference only differ
 * in traversal; every element is written once as unsorted*factor. */
void ora_sort_4(const double *unsorted, double *sorted, long a, long b, long c,
                long d, long i, long j, long k, long l, double factor) {
  long jd[5] = {0, a, b, c, d};
  long id[5];
  for (long j1 = 1; j1 <= a; j1++) {
    id[1] = j1;
    for (long j2 = 1; j2 <= b; j2++) {
      id[2] = j2;
      for (long j3 = 1; j3 <= c; j3++) {
        id[3] = j3;
        for (long j4 = 1; j4 <= d; j4++) {
          id[4] = j4;
          long ia = id[4] + jd[4] * (id[3] - 1 + jd[3] * (id[2] - 1 + jd[2] * (id[1] - 1)));
          long ib = id[l] + jd[l] * (id[k] - 1 + jd[k] * (id[j] - 1 + jd[j] * (id[i] - 1)));
          sorted[ib - 1] = unsorted[ia - 1] * factor;
        }
      }
    }
  }
}

/* tce_restricted.F:1-71. */
static void restricted_2(const ora_tiles_t *t, long a1, long a2, long *b1, long *b2) {
  if (t->restricted && SPN(t, a1) + SPN(t, a2) == 4) {
    *b1 = t->alpha[a1 - 1];
    *b2 = t->alpha[a2 - 1];
  } else {
    *b1 = a1;
    *b2 = a2;
  }
}
static void restricted_4(const ora_tiles_t *t, long a1, long a2, long a3, long a4,
                         long *b1, long *b2, long *b3, long *b4) {
  if (t->restricted && SPN(t, a1) + SPN(t, a2) + SPN(t, a3) + SPN(t, a4) == 8) {
    *b1 = t->alpha[a1 - 1];
    *b2 = t->alpha[a2 - 1];
    *b3 = t->alpha[a3 - 1];
    *b4 = t->alpha[a4 - 1];
  } else {
    *b1 = a1;
    *b2 = a2;
    *b3 = a3;
    *b4 = a4;
  }
}

/* -------------------------------------------------------------- kernels -- */

/* Row index codes of the kernels' loop variables. */
enum { P4 = 0, P5 = 1, P6 = 2, H1 = 3, H2 = 4, H3 = 5 };

/* Output layout of triplesx (Fortran order, fastest first) and the sign of
 * the update for each variant, as declared in ccsd_t_kernels_omp.F. */
static const int S1_L[9][6] = {
    {H3, H2, H1, P6, P5, P4}, {H3, H1, H2, P6, P5, P4}, {H1, H3, H2, P6, P5, P4},
    {H3, H2, H1, P6, P4, P5}, {H3, H1, H2, P6, P4, P5}, {H1, H3, H2, P6, P4, P5},
    {H3, H2, H1, P4, P6, P5}, {H3, H1, H2, P4, P6, P5}, {H1, H3, H2, P4, P6, P5}};
static const int S1_SIGN[9] = {+1, -1, +1, -1, +1, -1, +1, -1, +1};
static const int D1_L[9][6] = {
    {H3, H2, H1, P6, P5, P4}, {H3, H1, H2, P6, P5, P4}, {H1, H3, H2, P6, P5, P4},
    {H3, H2, H1, P5, P4, P6}, {H3, H1, H2, P5, P4, P6}, {H1, H3, H2, P5, P4, P6},
    {H3, H2, H1, P5, P6, P4}, {H3, H1, H2, P5, P6, P4}, {H1, H3, H2, P5, P6, P4}};
static const int D1_SIGN[9] = {-1, +1, -1, -1, +1, -1, +1, -1, +1};
static const int D2_L[9][6] = {
    {H3, H2, H1, P6, P5, P4}, {H2, H1, H3, P6, P5, P4}, {H2, H3, H1, P6, P5, P4},
    {H3, H2, H1, P6, P4, P5}, {H2, H1, H3, P6, P4, P5}, {H2, H3, H1, P6, P4, P5},
    {H3, H2, H1, P4, P6, P5}, {H2, H1, H3, P4, P6, P5}, {H2, H3, H1, P4, P6, P5}};
static const int D2_SIGN[9] = {-1, -1, +1, +1, +1, -1, -1, -1, +1};

/* Planted-mistake hook for tests: negate one variant of the built-in kernels
 * (family 0 = s1, 1 = d1, 2 = d2; n = 1..9; family -1 = none). */
static int g_mut_family = -1, g_mut_n = 0;
void ora_set_mutation(int family, int n) {
  g_mut_family = family;
  g_mut_n = n;
}
static int mut_sign(int family, int n, int sg) {
  return (family == g_mut_family && n == g_mut_n) ? -sg : sg;
}

/* ------------------------------------------ reduced-precision arithmetic -- */

/* Operand arithmetic of the built-in d1/d2 kernels (the s1 kernels and the
 * energy stay FP64 in every mode).  The reduced-precision modes restate the
 * library's operand conversion and represent every product EXACTLY in FP64:
 * what they leave out of the GPU arithmetic is only the FP32 rounding of
 * its accumulation, so the gap to the GPU is what an accumulation error
 * model has to cover.  ORA_ARITH_ABS adds |a|*|b| whatever the variant's
 * sign: the sum of product magnitudes such a model needs.  The two planted
 * mistakes are for tests of that model's sensitivity. */
static int g_arith = ORA_ARITH_FP64;
void ora_set_arith(int mode) { g_arith = mode; }

static unsigned f32_bits(float f) {
  unsigned u;
  memcpy(&u, &f, 4);
  return u;
}
static float bits_f32(unsigned u) {
  float f;
  memcpy(&f, &u, 4);
  return f;
}
/* bf16 of a float, round to nearest even on the upper 16 bits (finite x). */
static unsigned short bf16_rne(float f) {
  unsigned u = f32_bits(f);
  u += 0x7FFFu + ((u >> 16) & 1u);
  return (unsigned short)(u >> 16);
}
/* bf16 of a float by truncation toward zero (the planted mistake). */
static unsigned short bf16_trunc(float f) { return (unsigned short)(f32_bits(f) >> 16); }
static double bf16_val(unsigned short h) { return (double)bits_f32((unsigned)h << 16); }

void ora_split_bf16(const double *x, long n, int truncate_lo, unsigned short *hi,
                    unsigned short *lo) {
  for (long i = 0; i < n; i++) {
    /* hi = bf16(fl32(x)); lo = bf16(fl32(x - hi)), x - hi exact in FP64 */
    unsigned short h = bf16_rne((float)x[i]);
    float r = (float)(x[i] - bf16_val(h));
    hi[i] = h;
    lo[i] = truncate_lo ? bf16_trunc(r) : bf16_rne(r);
  }
}

void ora_split3_bf16(const double *x, long n, unsigned short *hi, unsigned short *mid,
                     unsigned short *lo) {
  for (long i = 0; i < n; i++) {
    /* hi = bf16(fl32(x)), mid = bf16(fl32(x - hi)), lo = bf16(fl32(x - hi -
     * mid)); both remainders exact in FP64 */
    unsigned short h = bf16_rne((float)x[i]);
    double r1 = x[i] - bf16_val(h);
    unsigned short m = bf16_rne((float)r1);
    double r2 = r1 - bf16_val(m);
    hi[i] = h;
    mid[i] = m;
    lo[i] = bf16_rne((float)r2);
  }
}

/* Split operand: value = h + m + l (FP32: h = fl32(x), m = l = 0; BF16X3:
 * m = 0; ABS: h = |x|). */
static void split_operand(const double *x, long n, double *h, double *m, double *l) {
  for (long i = 0; i < n; i++) {
    m[i] = 0.0;
    if (g_arith == ORA_ARITH_FP32) {
      h[i] = (double)(float)x[i];
      l[i] = 0.0;
    } else if (g_arith == ORA_ARITH_ABS) {
      h[i] = fabs(x[i]);
      l[i] = 0.0;
    } else if (g_arith == ORA_ARITH_BF16X6) {
      unsigned short hi, mi, lo;
      ora_split3_bf16(x + i, 1, &hi, &mi, &lo);
      h[i] = bf16_val(hi);
      m[i] = bf16_val(mi);
      l[i] = bf16_val(lo);
    } else {
      unsigned short hi, lo;
      ora_split_bf16(x + i, 1, g_arith == ORA_ARITH_BF16X3_TRUNC, &hi, &lo);
      h[i] = bf16_val(hi);
      l[i] = bf16_val(lo);
    }
  }
}

/* The product a*b of the current mode, exact in FP64 (8- or 24-bit
 * significands, at most 6 terms spanning < 53 bits): hi*hi' + hi*lo' +
 * lo*hi' (the 2-term mistake drops lo*hi'); BF16X6 also hi*mid' + mid*hi' +
 * mid*mid' (mid = 0 in the other modes). */
static inline double emul_prod(double ha, double ma, double la, double hb, double mb,
                               double lb) {
  if (g_arith == ORA_ARITH_BF16X2) return ha * hb + ha * lb;
  if (g_arith == ORA_ARITH_BF16X6)
    return ha * hb + (ha * mb + ma * hb) + (ha * lb + ma * mb + la * hb);
  return ha * hb + ha * lb + la * hb;
}

/* Strides of the six loop variables inside triplesx for layout L. */
static void out_strides(const int L[6], const long dims[6], long st[6]) {
  long s = 1;
  for (int q = 0; q < 6; q++) {
    st[L[q]] = s;
    s *= dims[L[q]];
  }
}

/* t3(L) (+/-)= t1sub(p4,h1) * v2sub(h3,h2,p6,p5) */
void ora_sd_t_s1(int n, long h3d, long h2d, long h1d, long p6d, long p5d,
                 long p4d, double *t3, const double *t1, const double *v2) {
  long dims[6] = {p4d, p5d, p6d, h1d, h2d, h3d}, st[6];
  out_strides(S1_L[n - 1], dims, st);
  int sg = mut_sign(0, n, S1_SIGN[n - 1]);
#pragma omp parallel for collapse(3) schedule(static)
  for (long p4 = 0; p4 < p4d; p4++)
    for (long p5 = 0; p5 < p5d; p5++)
      for (long p6 = 0; p6 < p6d; p6++)
        for (long h1 = 0; h1 < h1d; h1++)
          for (long h2 = 0; h2 < h2d; h2++)
            for (long h3 = 0; h3 < h3d; h3++) {
              long o = p4 * st[P4] + p5 * st[P5] + p6 * st[P6] + h1 * st[H1] +
                       h2 * st[H2] + h3 * st[H3];
              double a = t1[p4 + p4d * h1];
              double b = v2[h3 + h3d * (h2 + h2d * (p6 + p6d * p5))];
              if (sg > 0)
                t3[o] = t3[o] + a * b;
              else
                t3[o] = t3[o] - a * b;
            }
}

/* t3(L) (+/-)= sum_h7 t2sub(h7,p4,p5,h1) * v2sub(h3,h2,p6,h7) */
static void emul_d1(int n, long h3d, long h2d, long h1d, long p6d, long p5d, long p4d,
                    long h7d, double *t3, const double *t2, const double *v2);
static void emul_d2(int n, long h3d, long h2d, long h1d, long p6d, long p5d, long p4d,
                    long p7d, double *t3, const double *t2, const double *v2);

void ora_sd_t_d1(int n, long h3d, long h2d, long h1d, long p6d, long p5d,
                 long p4d, long h7d, double *t3, const double *t2,
                 const double *v2) {
  if (g_arith != ORA_ARITH_FP64) {
    emul_d1(n, h3d, h2d, h1d, p6d, p5d, p4d, h7d, t3, t2, v2);
    return;
  }
  long dims[6] = {p4d, p5d, p6d, h1d, h2d, h3d}, st[6];
  out_strides(D1_L[n - 1], dims, st);
  int sg = mut_sign(1, n, D1_SIGN[n - 1]);
#pragma omp parallel for collapse(3) schedule(static)
  for (long p4 = 0; p4 < p4d; p4++)
    for (long p5 = 0; p5 < p5d; p5++)
      for (long p6 = 0; p6 < p6d; p6++)
        for (long h1 = 0; h1 < h1d; h1++)
          for (long h2 = 0; h2 < h2d; h2++)
            for (long h3 = 0; h3 < h3d; h3++) {
              long o = p4 * st[P4] + p5 * st[P5] + p6 * st[P6] + h1 * st[H1] +
                       h2 * st[H2] + h3 * st[H3];
              double acc = t3[o];
              const double *ta = t2 + h7d * (p4 + p4d * (p5 + p5d * h1));
              for (long h7 = 0; h7 < h7d; h7++) {
                double b = v2[h3 + h3d * (h2 + h2d * (p6 + p6d * h7))];
                if (sg > 0)
                  acc = acc + ta[h7] * b;
                else
                  acc = acc - ta[h7] * b;
              }
              t3[o] = acc;
            }
}

/* t3(L) (+/-)= sum_p7 t2sub(p7,p4,h1,h2) * v2sub(p7,h3,p6,p5) */
void ora_sd_t_d2(int n, long h3d, long h2d, long h1d, long p6d, long p5d,
                 long p4d, long p7d, double *t3, const double *t2,
                 const double *v2) {
  if (g_arith != ORA_ARITH_FP64) {
    emul_d2(n, h3d, h2d, h1d, p6d, p5d, p4d, p7d, t3, t2, v2);
    return;
  }
  long dims[6] = {p4d, p5d, p6d, h1d, h2d, h3d}, st[6];
  out_strides(D2_L[n - 1], dims, st);
  int sg = mut_sign(2, n, D2_SIGN[n - 1]);
#pragma omp parallel for collapse(3) schedule(static)
  for (long p4 = 0; p4 < p4d; p4++)
    for (long p5 = 0; p5 < p5d; p5++)
      for (long p6 = 0; p6 < p6d; p6++)
        for (long h1 = 0; h1 < h1d; h1++)
          for (long h2 = 0; h2 < h2d; h2++)
            for (long h3 = 0; h3 < h3d; h3++) {
              long o = p4 * st[P4] + p5 * st[P5] + p6 * st[P6] + h1 * st[H1] +
                       h2 * st[H2] + h3 * st[H3];
              double acc = t3[o];
              const double *ta = t2 + p7d * (p4 + p4d * (h1 + h1d * h2));
              const double *vb = v2 + p7d * (h3 + h3d * (p6 + p6d * p5));
              for (long p7 = 0; p7 < p7d; p7++) {
                if (sg > 0)
                  acc = acc + ta[p7] * vb[p7];
                else
                  acc = acc - ta[p7] * vb[p7];
              }
              t3[o] = acc;
            }
}

/* The d1 and d2 loops of the reduced-precision modes: the same element
 * order, operands split once per call, each product exact. */
static void emul_d1(int n, long h3d, long h2d, long h1d, long p6d, long p5d, long p4d,
                    long h7d, double *t3, const double *t2, const double *v2) {
  long dims[6] = {p4d, p5d, p6d, h1d, h2d, h3d}, st[6];
  out_strides(D1_L[n - 1], dims, st);
  int sg = g_arith == ORA_ARITH_ABS ? 1 : mut_sign(1, n, D1_SIGN[n - 1]);
  long na = h7d * p4d * p5d * h1d, nb = h3d * h2d * p6d * h7d;
  double *ha = malloc(sizeof(double) * (3 * na + 3 * nb)), *ma = ha + na, *la = ma + na,
         *hb = la + na, *mb = hb + nb, *lb = mb + nb;
  split_operand(t2, na, ha, ma, la);
  split_operand(v2, nb, hb, mb, lb);
#pragma omp parallel for collapse(3) schedule(static)
  for (long p4 = 0; p4 < p4d; p4++)
    for (long p5 = 0; p5 < p5d; p5++)
      for (long p6 = 0; p6 < p6d; p6++)
        for (long h1 = 0; h1 < h1d; h1++)
          for (long h2 = 0; h2 < h2d; h2++)
            for (long h3 = 0; h3 < h3d; h3++) {
              long o = p4 * st[P4] + p5 * st[P5] + p6 * st[P6] + h1 * st[H1] +
                       h2 * st[H2] + h3 * st[H3];
              long ia = h7d * (p4 + p4d * (p5 + p5d * h1));
              double acc = t3[o];
              for (long h7 = 0; h7 < h7d; h7++) {
                long ib = h3 + h3d * (h2 + h2d * (p6 + p6d * h7));
                double pr = emul_prod(ha[ia + h7], ma[ia + h7], la[ia + h7], hb[ib], mb[ib], lb[ib]);
                acc = sg > 0 ? acc + pr : acc - pr;
              }
              t3[o] = acc;
            }
  free(ha);
}

static void emul_d2(int n, long h3d, long h2d, long h1d, long p6d, long p5d, long p4d,
                    long p7d, double *t3, const double *t2, const double *v2) {
  long dims[6] = {p4d, p5d, p6d, h1d, h2d, h3d}, st[6];
  out_strides(D2_L[n - 1], dims, st);
  int sg = g_arith == ORA_ARITH_ABS ? 1 : mut_sign(2, n, D2_SIGN[n - 1]);
  long na = p7d * p4d * h1d * h2d, nb = p7d * h3d * p6d * p5d;
  double *ha = malloc(sizeof(double) * (3 * na + 3 * nb)), *ma = ha + na, *la = ma + na,
         *hb = la + na, *mb = hb + nb, *lb = mb + nb;
  split_operand(t2, na, ha, ma, la);
  split_operand(v2, nb, hb, mb, lb);
#pragma omp parallel for collapse(3) schedule(static)
  for (long p4 = 0; p4 < p4d; p4++)
    for (long p5 = 0; p5 < p5d; p5++)
      for (long p6 = 0; p6 < p6d; p6++)
        for (long h1 = 0; h1 < h1d; h1++)
          for (long h2 = 0; h2 < h2d; h2++)
            for (long h3 = 0; h3 < h3d; h3++) {
              long o = p4 * st[P4] + p5 * st[P5] + p6 * st[P6] + h1 * st[H1] +
                       h2 * st[H2] + h3 * st[H3];
              long ia = p7d * (p4 + p4d * (h1 + h1d * h2));
              long ib = p7d * (h3 + h3d * (p6 + p6d * p5));
              double acc = t3[o];
              for (long p7 = 0; p7 < p7d; p7++) {
                double pr = emul_prod(ha[ia + p7], ma[ia + p7], la[ia + p7], hb[ib + p7], mb[ib + p7],
                                      lb[ib + p7]);
                acc = sg > 0 ? acc + pr : acc - pr;
              }
              t3[o] = acc;
            }
  free(ha);
}

/* Adapters so the built-in kernels share the by-reference back-end type. */
#define S1_ADAPT(N)                                                             \
  static void bi_s1_##N(long *a, long *b, long *c, long *d, long *e, long *f,   \
                        double *t3, double *x, double *y) {                     \
    ora_sd_t_s1(N, *a, *b, *c, *d, *e, *f, t3, x, y);                          \
  }
#define D1_ADAPT(N)                                                             \
  static void bi_d1_##N(long *a, long *b, long *c, long *d, long *e, long *f,   \
                        long *g, double *t3, double *x, double *y) {            \
    ora_sd_t_d1(N, *a, *b, *c, *d, *e, *f, *g, t3, x, y);                      \
  }
#define D2_ADAPT(N)                                                             \
  static void bi_d2_##N(long *a, long *b, long *c, long *d, long *e, long *f,   \
                        long *g, double *t3, double *x, double *y) {            \
    ora_sd_t_d2(N, *a, *b, *c, *d, *e, *f, *g, t3, x, y);                      \
  }
#define ALL9(M) M(1) M(2) M(3) M(4) M(5) M(6) M(7) M(8) M(9)
ALL9(S1_ADAPT)
ALL9(D1_ADAPT)
ALL9(D2_ADAPT)
static const ora_backend_t builtin_backend = {
    0,
    {bi_s1_1, bi_s1_2, bi_s1_3, bi_s1_4, bi_s1_5, bi_s1_6, bi_s1_7, bi_s1_8, bi_s1_9},
    {bi_d1_1, bi_d1_2, bi_d1_3, bi_d1_4, bi_d1_5, bi_d1_6, bi_d1_7, bi_d1_8, bi_d1_9},
    {bi_d2_1, bi_d2_2, bi_d2_3, bi_d2_4, bi_d2_5, bi_d2_6, bi_d2_7, bi_d2_8, bi_d2_9},
    0, 0, 0, 0, 0, 0};

/* --------------------------------------------------------------- energy -- */

static double task_factor(int restricted, long p4b, long p5b, long p6b, long h1b,
                          long h2b, long h3b) {
  double factor = restricted ? 2.0 : 1.0;
  if (p4b == p5b && p5b == p6b)
    factor = factor / 6.0;
  else if (p4b == p5b || p5b == p6b)
    factor = factor / 2.0;
  if (h1b == h2b && h2b == h3b)
    factor = factor / 6.0;
  else if (h1b == h2b || h2b == h3b)
    factor = factor / 2.0;
  return factor;
}

/* ccsd_t_dot.F:20-125. */
void ora_ccsd_t_dot(const double *singles, const double *doubles, int restricted,
                    long h1b, long h2b, long h3b, long p4b, long p5b, long p6b,
                    const double *o_h1, const double *o_h2, const double *o_h3,
                    const double *o_p4, const double *o_p5, const double *o_p6,
                    long r_h1, long r_h2, long r_h3, long r_p4, long r_p5,
                    long r_p6, double *energy1, double *energy2) {
  double factor = task_factor(restricted, p4b, p5b, p6b, h1b, h2b, h3b);
  double e1 = *energy1, e2 = *energy2;
  for (long p4 = 0; p4 < r_p4; p4++)
    for (long p5 = 0; p5 < r_p5; p5++)
      for (long p6 = 0; p6 < r_p6; p6++) {
        double denom_0 = -(o_p4[p4] + o_p5[p5] + o_p6[p6]);
        for (long h1 = 0; h1 < r_h1; h1++)
          for (long h2 = 0; h2 < r_h2; h2++)
            for (long h3 = 0; h3 < r_h3; h3++) {
              long i = h3 + r_h3 * (h2 + r_h2 * (h1 + r_h1 * (p6 + r_p6 * (p5 + r_p5 * p4))));
              double sing = singles[i], doub = doubles[i];
              double denom = doub * factor / (o_h1[h1] + o_h2[h2] + o_h3[h3] + denom_0);
              e1 = e1 + denom * doub;
              e2 = e2 + denom * (doub + sing);
            }
      }
  *energy1 = e1;
  *energy2 = e2;
}

/* ------------------------------------------------------------- dispatch -- */

/* Row tables of the three drivers: entries are indices into the task tuple
 * (t_p4b,t_p5b,t_p6b,t_h1b,t_h2b,t_h3b) = slots 0..5. */
static const int S1_ROWS[9][6] = {  /* ccsd_t_singles_gpu.F:101-152 */
    {0, 1, 2, 3, 4, 5}, {0, 1, 2, 4, 3, 5}, {0, 1, 2, 5, 3, 4},
    {1, 0, 2, 3, 4, 5}, {1, 0, 2, 4, 3, 5}, {1, 0, 2, 5, 3, 4},
    {2, 0, 1, 3, 4, 5}, {2, 0, 1, 4, 3, 5}, {2, 0, 1, 5, 3, 4}};
static const int D1_ROWS[9][6] = {  /* ccsd_t_doubles_gpu.F:133-187 */
    {0, 1, 2, 3, 4, 5}, {0, 1, 2, 4, 3, 5}, {0, 1, 2, 5, 3, 4},
    {1, 2, 0, 3, 4, 5}, {1, 2, 0, 4, 3, 5}, {1, 2, 0, 5, 3, 4},
    {0, 2, 1, 3, 4, 5}, {0, 2, 1, 4, 3, 5}, {0, 2, 1, 5, 3, 4}};
static const int D2_ROWS[9][6] = {  /* ccsd_t_doubles_gpu.F:816-870 */
    {0, 1, 2, 3, 4, 5}, {0, 1, 2, 4, 5, 3}, {0, 1, 2, 3, 5, 4},
    {1, 0, 2, 3, 4, 5}, {1, 0, 2, 4, 5, 3}, {1, 0, 2, 3, 5, 4},
    {2, 0, 1, 3, 4, 5}, {2, 0, 1, 4, 5, 3}, {2, 0, 1, 3, 5, 4}};

/* Kernel-firing patterns: kernel N fires for a row when
 * task[q] == row[PAT[N][q]] for every slot q (e.g. ccsd_t_singles_gpu.F:283,
 * 313, ...).  For all three drivers pattern N is the row-N permutation. */
static void build_rows(const long task[6], const int R[9][6], long a3[9][6]) {
  for (int r = 0; r < 9; r++)
    for (int q = 0; q < 6; q++) a3[r][q] = task[R[r][q]];
  /* de-duplication, e.g. ccsd_t_singles_gpu.F:161-177 */
  for (int ia = 0; ia < 8; ia++) {
    if (a3[ia][0] == 0) continue;
    for (int ja = ia + 1; ja < 9; ja++) {
      int same = 1;
      for (int q = 0; q < 6; q++) same &= (a3[ia][q] == a3[ja][q]);
      if (same)
        for (int q = 0; q < 6; q++) a3[ja][q] = 0;
    }
  }
}

/* Pattern N matches when the task tuple equals the row tuple permuted by
 * row table N, i.e. row == task o R[N] viewed from the row side: the reference
 * tests task[p] == row[...] slot by slot; with row = task o R, that is
 * row[q] == task[R[N][q]] for all q. */
static int fires(const long task[6], const long row[6], const int R[9][6], int n) {
  for (int q = 0; q < 6; q++)
    if (row[q] != task[R[n][q]]) return 0;
  return 1;
}

typedef struct {
  const ora_tiles_t *t;
  const double *t1, *t2, *v2;
  const long *t1_hash, *t2_hash, *v2_hash;
  const ora_backend_t *be;
  double *scratch[4]; /* a, a_sort, b_sort */
  long scratch_len;
  int err;
} dispatch_t;

static double *buf(dispatch_t *D, int i) { return D->scratch[i]; }

/* ccsd_t_singles_gpu_1 (ccsd_t_singles_gpu.F:36-576). */
static void singles(dispatch_t *D, double *a_c, const long task[6]) {
  const ora_tiles_t *t = D->t;
  long noab = t->noab, N = t->noab + t->nvab;
  long a3[9][6];
  build_rows(task, S1_ROWS, a3);
  int notset = 1;
  for (int r = 0; r < 9; r++) {
    long p4b = a3[r][0], p5b = a3[r][1], p6b = a3[r][2];
    long h1b = a3[r][3], h2b = a3[r][4], h3b = a3[r][5];
    if (D->be->usedevice && notset) {
      long d[6] = {RNG(t, task[3]), RNG(t, task[4]), RNG(t, task[5]),
                   RNG(t, task[0]), RNG(t, task[1]), RNG(t, task[2])};
      D->be->dev_mem_s(&d[0], &d[1], &d[2], &d[3], &d[4], &d[5]);
      notset = 0;
    }
    if (!(p5b <= p6b && h2b <= h3b && p4b != 0)) continue;
    if (t->restricted && SPN(t, p4b) + SPN(t, p5b) + SPN(t, p6b) + SPN(t, h1b) +
                                 SPN(t, h2b) + SPN(t, h3b) == 12)
      continue;
    if (SPN(t, p4b) + SPN(t, p5b) + SPN(t, p6b) != SPN(t, h1b) + SPN(t, h2b) + SPN(t, h3b))
      continue;
    if ((SYM(t, p4b) ^ SYM(t, p5b) ^ SYM(t, p6b) ^ SYM(t, h1b) ^ SYM(t, h2b) ^ SYM(t, h3b)) != 0)
      continue;
    if (SPN(t, p4b) != SPN(t, h1b)) continue;
    if ((SYM(t, p4b) ^ SYM(t, h1b)) != 0) continue;
    long p4b_1, h1b_1, p5b_2, p6b_2, h2b_2, h3b_2;
    restricted_2(t, p4b, h1b, &p4b_1, &h1b_1);
    restricted_4(t, p5b, p6b, h2b, h3b, &p5b_2, &p6b_2, &h2b_2, &h3b_2);
    long dima = RNG(t, p4b) * RNG(t, h1b);
    long dimb = RNG(t, p5b) * RNG(t, p6b) * RNG(t, h2b) * RNG(t, h3b);
    if (!(dima > 0 && dimb > 0)) continue;
    double *k_a = buf(D, 0), *k_a_sort = buf(D, 1), *k_b_sort = buf(D, 2);
    if (get_hash_block(D->t1, k_a, dima, D->t1_hash, h1b_1 - 1 + noab * (p4b_1 - noab - 1))) {
      D->err = -1;
      return;
    }
    ora_sort_2(k_a, k_a_sort, RNG(t, p4b), RNG(t, h1b), 2, 1, 1.0);
    if (get_hash_block(D->v2, k_b_sort, dimb, D->v2_hash,
                       h3b_2 - 1 + N * (h2b_2 - 1 + N * (p6b_2 - 1 + N * (p5b_2 - 1))))) {
      D->err = -2;
      return;
    }
    long rh3 = RNG(t, h3b), rh2 = RNG(t, h2b), rh1 = RNG(t, h1b);
    long rp6 = RNG(t, p6b), rp5 = RNG(t, p5b), rp4 = RNG(t, p4b);
    for (int n = 0; n < 9; n++) {
      if (!fires(task, a3[r], S1_ROWS, n)) continue;
      if (D->be->usedevice)
        D->be->s1[n](&rh1, &rh2, &rh3, &rp4, &rp5, &rp6, a_c, k_a_sort, k_b_sort);
      else
        D->be->s1[n](&rh3, &rh2, &rh1, &rp6, &rp5, &rp4, a_c, k_a_sort, k_b_sort);
    }
  }
}

/* ccsd_t_doubles_gpu_1 (ccsd_t_doubles_gpu.F:48-742): d1, loop over h7b. */
static void doubles_1(dispatch_t *D, double *a_c, const long task[6]) {
  const ora_tiles_t *t = D->t;
  long noab = t->noab, nvab = t->nvab, N = noab + nvab;
  long a3[9][6];
  build_rows(task, D1_ROWS, a3);
  int notset = 1;
  for (int r = 0; r < 9; r++) {
    long p4b = a3[r][0], p5b = a3[r][1], p6b = a3[r][2];
    long h1b = a3[r][3], h2b = a3[r][4], h3b = a3[r][5];
    if (D->be->usedevice && notset) {
      long d[6] = {RNG(t, task[3]), RNG(t, task[4]), RNG(t, task[5]),
                   RNG(t, task[0]), RNG(t, task[1]), RNG(t, task[2])};
      D->be->dev_mem_d(&d[0], &d[1], &d[2], &d[3], &d[4], &d[5]);
      notset = 0;
    }
    if (!(p4b <= p5b && h2b <= h3b && p4b != 0)) continue;
    if (t->restricted && SPN(t, p4b) + SPN(t, p5b) + SPN(t, p6b) + SPN(t, h1b) +
                                 SPN(t, h2b) + SPN(t, h3b) == 12)
      continue;
    if (SPN(t, p4b) + SPN(t, p5b) + SPN(t, p6b) != SPN(t, h1b) + SPN(t, h2b) + SPN(t, h3b))
      continue;
    if ((SYM(t, p4b) ^ SYM(t, p5b) ^ SYM(t, p6b) ^ SYM(t, h1b) ^ SYM(t, h2b) ^ SYM(t, h3b)) != 0)
      continue;
    for (long h7b = 1; h7b <= noab; h7b++) {
      if (SPN(t, p4b) + SPN(t, p5b) != SPN(t, h1b) + SPN(t, h7b)) continue;
      if ((SYM(t, p4b) ^ SYM(t, p5b) ^ SYM(t, h1b) ^ SYM(t, h7b)) != 0) continue;
      long p4b_1, p5b_1, h1b_1, h7b_1, p6b_2, h7b_2, h2b_2, h3b_2;
      restricted_4(t, p4b, p5b, h1b, h7b, &p4b_1, &p5b_1, &h1b_1, &h7b_1);
      restricted_4(t, p6b, h7b, h2b, h3b, &p6b_2, &h7b_2, &h2b_2, &h3b_2);
      long dim_common = RNG(t, h7b);
      long dima = dim_common * RNG(t, p4b) * RNG(t, p5b) * RNG(t, h1b);
      long dimb = dim_common * RNG(t, p6b) * RNG(t, h2b) * RNG(t, h3b);
      if (!(dima > 0 && dimb > 0)) continue;
      double *k_a = buf(D, 0), *k_a_sort = buf(D, 1), *k_b_sort = buf(D, 2);
      if (h7b < h1b) {
        if (get_hash_block(D->t2, k_a, dima, D->t2_hash,
                           h1b_1 - 1 + noab * (h7b_1 - 1 + noab * (p5b_1 - noab - 1 +
                                                                   nvab * (p4b_1 - noab - 1))))) {
          D->err = -3;
          return;
        }
        ora_sort_4(k_a, k_a_sort, RNG(t, p4b), RNG(t, p5b), RNG(t, h7b), RNG(t, h1b),
                   4, 2, 1, 3, -1.0);
      }
      if (h1b <= h7b) {
        if (get_hash_block(D->t2, k_a, dima, D->t2_hash,
                           h7b_1 - 1 + noab * (h1b_1 - 1 + noab * (p5b_1 - noab - 1 +
                                                                   nvab * (p4b_1 - noab - 1))))) {
          D->err = -4;
          return;
        }
        ora_sort_4(k_a, k_a_sort, RNG(t, p4b), RNG(t, p5b), RNG(t, h1b), RNG(t, h7b),
                   3, 2, 1, 4, 1.0);
      }
      if (h7b <= p6b) {
        if (get_hash_block(D->v2, k_b_sort, dimb, D->v2_hash,
                           h3b_2 - 1 + N * (h2b_2 - 1 + N * (p6b_2 - 1 + N * (h7b_2 - 1))))) {
          D->err = -5;
          return;
        }
      }
      long rh3 = RNG(t, h3b), rh2 = RNG(t, h2b), rh1 = RNG(t, h1b);
      long rp6 = RNG(t, p6b), rp5 = RNG(t, p5b), rp4 = RNG(t, p4b), rh7 = RNG(t, h7b);
      for (int n = 0; n < 9; n++) {
        if (!fires(task, a3[r], D1_ROWS, n)) continue;
        if (D->be->usedevice)
          D->be->d1[n](&rh1, &rh2, &rh3, &rh7, &rp4, &rp5, &rp6, a_c, k_a_sort, k_b_sort);
        else
          D->be->d1[n](&rh3, &rh2, &rh1, &rp6, &rp5, &rp4, &rh7, a_c, k_a_sort, k_b_sort);
      }
    }
  }
}

/* ccsd_t_doubles_gpu_2 (ccsd_t_doubles_gpu.F:743-1345): d2, loop over p7b. */
static void doubles_2(dispatch_t *D, double *a_c, const long task[6]) {
  const ora_tiles_t *t = D->t;
  long noab = t->noab, nvab = t->nvab, N = noab + nvab;
  long a3[9][6];
  build_rows(task, D2_ROWS, a3);
  for (int r = 0; r < 9; r++) {
    long p4b = a3[r][0], p5b = a3[r][1], p6b = a3[r][2];
    long h1b = a3[r][3], h2b = a3[r][4], h3b = a3[r][5];
    if (!(p5b <= p6b && h1b <= h2b && p4b != 0)) continue;
    if (t->restricted && SPN(t, p4b) + SPN(t, p5b) + SPN(t, p6b) + SPN(t, h1b) +
                                 SPN(t, h2b) + SPN(t, h3b) == 12)
      continue;
    if (SPN(t, p4b) + SPN(t, p5b) + SPN(t, p6b) != SPN(t, h1b) + SPN(t, h2b) + SPN(t, h3b))
      continue;
    if ((SYM(t, p4b) ^ SYM(t, p5b) ^ SYM(t, p6b) ^ SYM(t, h1b) ^ SYM(t, h2b) ^ SYM(t, h3b)) != 0)
      continue;
    for (long p7b = noab + 1; p7b <= noab + nvab; p7b++) {
      if (SPN(t, p4b) + SPN(t, p7b) != SPN(t, h1b) + SPN(t, h2b)) continue;
      if ((SYM(t, p4b) ^ SYM(t, p7b) ^ SYM(t, h1b) ^ SYM(t, h2b)) != 0) continue;
      long p4b_1, p7b_1, h1b_1, h2b_1, p5b_2, p6b_2, h3b_2, p7b_2;
      restricted_4(t, p4b, p7b, h1b, h2b, &p4b_1, &p7b_1, &h1b_1, &h2b_1);
      restricted_4(t, p5b, p6b, h3b, p7b, &p5b_2, &p6b_2, &h3b_2, &p7b_2);
      long dim_common = RNG(t, p7b);
      long dima = dim_common * RNG(t, p4b) * RNG(t, h1b) * RNG(t, h2b);
      long dimb = dim_common * RNG(t, p5b) * RNG(t, p6b) * RNG(t, h3b);
      if (!(dima > 0 && dimb > 0)) continue;
      double *k_a = buf(D, 0), *k_a_sort = buf(D, 1), *k_b_sort = buf(D, 2);
      if (p7b < p4b) {
        if (get_hash_block(D->t2, k_a, dima, D->t2_hash,
                           h2b_1 - 1 + noab * (h1b_1 - 1 + noab * (p4b_1 - noab - 1 +
                                                                   nvab * (p7b_1 - noab - 1))))) {
          D->err = -6;
          return;
        }
        ora_sort_4(k_a, k_a_sort, RNG(t, p7b), RNG(t, p4b), RNG(t, h1b), RNG(t, h2b),
                   4, 3, 2, 1, -1.0);
      }
      if (p4b <= p7b) {
        if (get_hash_block(D->t2, k_a, dima, D->t2_hash,
                           h2b_1 - 1 + noab * (h1b_1 - 1 + noab * (p7b_1 - noab - 1 +
                                                                   nvab * (p4b_1 - noab - 1))))) {
          D->err = -7;
          return;
        }
        ora_sort_4(k_a, k_a_sort, RNG(t, p4b), RNG(t, p7b), RNG(t, h1b), RNG(t, h2b),
                   4, 3, 1, 2, 1.0);
      }
      if (h3b <= p7b) {
        if (get_hash_block(D->v2, k_b_sort, dimb, D->v2_hash,
                           p7b_2 - 1 + N * (h3b_2 - 1 + N * (p6b_2 - 1 + N * (p5b_2 - 1))))) {
          D->err = -8;
          return;
        }
      }
      long rh3 = RNG(t, h3b), rh2 = RNG(t, h2b), rh1 = RNG(t, h1b);
      long rp6 = RNG(t, p6b), rp5 = RNG(t, p5b), rp4 = RNG(t, p4b), rp7 = RNG(t, p7b);
      for (int n = 0; n < 9; n++) {
        if (!fires(task, a3[r], D2_ROWS, n)) continue;
        if (D->be->usedevice)
          D->be->d2[n](&rh1, &rh2, &rh3, &rp4, &rp5, &rp6, &rp7, a_c, k_a_sort, k_b_sort);
        else
          D->be->d2[n](&rh3, &rh2, &rh1, &rp6, &rp5, &rp4, &rp7, a_c, k_a_sort, k_b_sort);
      }
    }
  }
}

/* Task filter of ccsd_t_gpu.F:94-112. */
static int task_ok(const ora_tiles_t *t, long p4b, long p5b, long p6b, long h1b,
                   long h2b, long h3b) {
  if (SPN(t, p4b) + SPN(t, p5b) + SPN(t, p6b) != SPN(t, h1b) + SPN(t, h2b) + SPN(t, h3b))
    return 0;
  if (t->restricted && SPN(t, p4b) + SPN(t, p5b) + SPN(t, p6b) + SPN(t, h1b) +
                               SPN(t, h2b) + SPN(t, h3b) > 8)
    return 0;
  if ((SYM(t, p4b) ^ SYM(t, p5b) ^ SYM(t, p6b) ^ SYM(t, h1b) ^ SYM(t, h2b) ^ SYM(t, h3b)) != 0)
    return 0;
  return 1;
}

long ora_count_tasks(const ora_tiles_t *t) {
  long noab = t->noab, nvab = t->nvab, count = 0;
  for (long p4b = noab + 1; p4b <= noab + nvab; p4b++)
    for (long p5b = p4b; p5b <= noab + nvab; p5b++)
      for (long p6b = p5b; p6b <= noab + nvab; p6b++)
        for (long h1b = 1; h1b <= noab; h1b++)
          for (long h2b = h1b; h2b <= noab; h2b++)
            for (long h3b = h2b; h3b <= noab; h3b++)
              if (task_ok(t, p4b, p5b, p6b, h1b, h2b, h3b)) count++;
  return count;
}

/* Dispatch state over the caller's blocks, with scratch for the largest
 * block any fetch can return (4 tile ranges).  0 or -100 (no memory). */
static int dispatch_init(dispatch_t *D, const ora_tiles_t *t, const double *t1,
                         const long *t1_hash, const double *t2, const long *t2_hash,
                         const double *v2, const long *v2_hash, const ora_backend_t *be) {
  memset(D, 0, sizeof(*D));
  D->t = t;
  D->t1 = t1;
  D->t2 = t2;
  D->v2 = v2;
  D->t1_hash = t1_hash;
  D->t2_hash = t2_hash;
  D->v2_hash = v2_hash;
  D->be = be ? be : &builtin_backend;
  long maxr = 0;
  for (long b = 1; b <= t->noab + t->nvab; b++)
    if (RNG(t, b) > maxr) maxr = RNG(t, b);
  D->scratch_len = maxr * maxr * maxr * maxr;
  for (int i = 0; i < 3; i++) {
    D->scratch[i] = (double *)malloc((size_t)D->scratch_len * sizeof(double));
    if (!D->scratch[i]) return -100;
  }
  return 0;
}

long ora_task_tiles(const ora_tiles_t *t, const double *t1, const long *t1_hash,
                    const double *t2, const long *t2_hash, const double *v2,
                    const long *v2_hash, const long tiles[6], double *t3s,
                    double *t3d) {
  const long p4b = tiles[0], p5b = tiles[1], p6b = tiles[2];
  const long h1b = tiles[3], h2b = tiles[4], h3b = tiles[5];
  const long noab = t->noab, nvab = t->nvab;
  if (!(1 <= h1b && h1b <= h2b && h2b <= h3b && h3b <= noab && noab < p4b && p4b <= p5b &&
        p5b <= p6b && p6b <= noab + nvab))
    return -1;
  if (!task_ok(t, p4b, p5b, p6b, h1b, h2b, h3b)) return -1;
  dispatch_t D;
  if (dispatch_init(&D, t, t1, t1_hash, t2, t2_hash, v2, v2_hash, NULL)) return -100;
  const long task[6] = {p4b, p5b, p6b, h1b, h2b, h3b};
  const long size = RNG(t, p4b) * RNG(t, p5b) * RNG(t, p6b) * RNG(t, h1b) * RNG(t, h2b) *
                    RNG(t, h3b);
  memset(t3s, 0, (size_t)size * sizeof(double));
  memset(t3d, 0, (size_t)size * sizeof(double));
  singles(&D, t3s, task);
  doubles_1(&D, t3d, task);
  doubles_2(&D, t3d, task);
  for (int i = 0; i < 3; i++) free(D.scratch[i]);
  return D.err ? D.err : size;
}

long ora_ccsd_t_gpu(const ora_tiles_t *t, const double *t1, const long *t1_hash,
                    const double *t2, const long *t2_hash, const double *v2,
                    const long *v2_hash, const ora_backend_t *be, long task_begin,
                    long task_end, double *energy1, double *energy2, double *task_e) {
  long noab = t->noab, nvab = t->nvab;
  dispatch_t D;
  if (dispatch_init(&D, t, t1, t1_hash, t2, t2_hash, v2, v2_hash, be)) return -100;
  double e1 = 0.0, e2 = 0.0;
  long count = 0, done = 0;
  double *singles_buf = NULL, *doubles_buf = NULL;
  size_t cap = 0;
  for (long p4b = noab + 1; p4b <= noab + nvab && !D.err; p4b++)
    for (long p5b = p4b; p5b <= noab + nvab && !D.err; p5b++)
      for (long p6b = p5b; p6b <= noab + nvab && !D.err; p6b++)
        for (long h1b = 1; h1b <= noab && !D.err; h1b++)
          for (long h2b = h1b; h2b <= noab && !D.err; h2b++)
            for (long h3b = h2b; h3b <= noab && !D.err; h3b++) {
              if (!task_ok(t, p4b, p5b, p6b, h1b, h2b, h3b)) continue;
              long idx = count++;
              if (idx < task_begin || (task_end >= 0 && idx >= task_end)) continue;
              long size = RNG(t, p4b) * RNG(t, p5b) * RNG(t, p6b) * RNG(t, h1b) *
                          RNG(t, h2b) * RNG(t, h3b);
              if ((size_t)size > cap) {
                free(singles_buf);
                free(doubles_buf);
                cap = (size_t)size;
                singles_buf = (double *)malloc(cap * sizeof(double));
                doubles_buf = (double *)malloc(cap * sizeof(double));
                if (!singles_buf || !doubles_buf) {
                  D.err = -101;
                  break;
                }
              }
              memset(singles_buf, 0, (size_t)size * sizeof(double));
              memset(doubles_buf, 0, (size_t)size * sizeof(double));
              long task[6] = {p4b, p5b, p6b, h1b, h2b, h3b};
              if (D.be->usedevice && D.be->initmemmodule) D.be->initmemmodule();
              singles(&D, singles_buf, task);
              doubles_1(&D, doubles_buf, task);
              doubles_2(&D, doubles_buf, task);
              if (D.err) break;
              double factor = task_factor(t->restricted, p4b, p5b, p6b, h1b, h2b, h3b);
              double te1 = 0.0, te2 = 0.0;
              if (!D.be->usedevice) {
                /* inline CPU energy of ccsd_t_gpu.F:170-200 */
                const double *ev = t->evl;
                long i = 0;
                for (long p4 = 0; p4 < RNG(t, p4b); p4++)
                  for (long p5 = 0; p5 < RNG(t, p5b); p5++)
                    for (long p6 = 0; p6 < RNG(t, p6b); p6++)
                      for (long h1 = 0; h1 < RNG(t, h1b); h1++)
                        for (long h2 = 0; h2 < RNG(t, h2b); h2++)
                          for (long h3 = 0; h3 < RNG(t, h3b); h3++) {
                            double d = doubles_buf[i], s = singles_buf[i];
                            i++;
                            double den = -ev[OFS(t, p4b) + p4] - ev[OFS(t, p5b) + p5] -
                                         ev[OFS(t, p6b) + p6] + ev[OFS(t, h1b) + h1] +
                                         ev[OFS(t, h2b) + h2] + ev[OFS(t, h3b) + h3];
                            te1 = te1 + factor * d * d / den;
                            te2 = te2 + factor * d * (s + d) / den;
                          }
              } else {
                double energy_l[2] = {0.0, 0.0}, factor_l = factor;
                long r1 = RNG(t, h1b), r2 = RNG(t, h2b), r3 = RNG(t, h3b);
                long r4 = RNG(t, p4b), r5 = RNG(t, p5b), r6 = RNG(t, p6b);
                const double *ev = t->evl;
                D.be->compute_en(&factor_l, energy_l, (double *)ev + OFS(t, h1b),
                                 (double *)ev + OFS(t, h2b), (double *)ev + OFS(t, h3b),
                                 (double *)ev + OFS(t, p4b), (double *)ev + OFS(t, p5b),
                                 (double *)ev + OFS(t, p6b), &r1, &r2, &r3, &r4, &r5, &r6,
                                 doubles_buf, singles_buf);
                te1 = energy_l[0];
                te2 = energy_l[1];
                D.be->dev_release();
                D.be->finalizememmodule();
              }
              e1 += te1;
              e2 += te2;
              if (task_e) {
                task_e[2 * done] = te1;
                task_e[2 * done + 1] = te2;
              }
              done++;
            }
  free(singles_buf);
  free(doubles_buf);
  for (int i = 0; i < 3; i++) free(D.scratch[i]);
  if (D.err) return D.err;
  *energy1 = e1;
  *energy2 = e2;
  return done;
}

/* ----------------------------------------------------------- brute force -- */

/* Standard spin-orbital (T) with the textbook antisymmetrisers
 *   t3d = P(a/bc) P(k/ij) [ -sum_e t(a e i j) <b c || k e> ]
 *       + P(c/ab) P(i/jk) [ -sum_m t(a b i m) <m c || j k> ]
 *   t3s = P(a/bc) P(i/jk) [ t(a i) <b c || j k> ]
 * with P(x/yz) f = f - f(x<->y) - f(x<->z), and
 *   E[T] = 1/36 sum t3d^2 / D,  E(T) = 1/36 sum t3d (t3d + t3s) / D. */
void ora_brute_force(long no, long nv, const double *T1, const double *T2,
                     const double *W, const double *V, const double *S,
                     const double *eo, const double *ev, double *e4, double *e5) {
#define T2_(a, b, i, j) T2[(((a) * nv + (b)) * no + (i)) * no + (j)]
#define W_(m, c, j, k) W[(((m) * nv + (c)) * no + (j)) * no + (k)]
#define V_(b, c, k, e) V[(((b) * nv + (c)) * no + (k)) * nv + (e)]
#define S_(b, c, j, k) S[(((b) * nv + (c)) * no + (j)) * no + (k)]
  double E4 = 0.0, E5 = 0.0;
#pragma omp parallel for collapse(2) reduction(+ : E4, E5) schedule(dynamic)
  for (long a = 0; a < nv; a++)
    for (long b = 0; b < nv; b++)
      for (long c = 0; c < nv; c++)
        for (long i = 0; i < no; i++)
          for (long j = 0; j < no; j++)
            for (long k = 0; k < no; k++) {
              /* d2 part: X(a,b,c,i,j,k) = -sum_e t(a e i j) v(b c k e) */
              const long pa[3][3] = {{a, b, c}, {b, a, c}, {c, b, a}};
              const long hk[3][3] = {{i, j, k}, {k, j, i}, {i, k, j}};
              const double sg[3] = {1.0, -1.0, -1.0};
              double d = 0.0;
              for (int x = 0; x < 3; x++)
                for (int y = 0; y < 3; y++) {
                  long A = pa[x][0], B = pa[x][1], C = pa[x][2];
                  long I = hk[y][0], J = hk[y][1], K = hk[y][2];
                  double s = 0.0;
                  for (long e = 0; e < nv; e++) s += T2_(A, e, I, J) * V_(B, C, K, e);
                  d -= sg[x] * sg[y] * s;
                }
              /* d1 part: Y(a,b,c,i,j,k) = -sum_m t(a b i m) v(m c j k),
                 antisymmetrised by P(c/ab) P(i/jk) */
              const long pc[3][3] = {{a, b, c}, {c, b, a}, {a, c, b}};
              const long hi[3][3] = {{i, j, k}, {j, i, k}, {k, j, i}};
              for (int x = 0; x < 3; x++)
                for (int y = 0; y < 3; y++) {
                  long A = pc[x][0], B = pc[x][1], C = pc[x][2];
                  long I = hi[y][0], J = hi[y][1], K = hi[y][2];
                  double s = 0.0;
                  for (long m = 0; m < no; m++) s += T2_(A, B, I, m) * W_(m, C, J, K);
                  d -= sg[x] * sg[y] * s;
                }
              /* singles: P(a/bc) P(i/jk) t(a i) v(b c j k) */
              double sv = 0.0;
              for (int x = 0; x < 3; x++)
                for (int y = 0; y < 3; y++) {
                  long A = pa[x][0], B = pa[x][1], C = pa[x][2];
                  long I = hi[y][0], J = hi[y][1], K = hi[y][2];
                  sv += sg[x] * sg[y] * T1[A * no + I] * S_(B, C, J, K);
                }
              double D = eo[i] + eo[j] + eo[k] - ev[a] - ev[b] - ev[c];
              E4 += d * d / D;
              E5 += d * (d + sv) / D;
            }
  *e4 = E4 / 36.0;
  *e5 = E5 / 36.0;
#undef T2_
#undef W_
#undef V_
#undef S_
}
