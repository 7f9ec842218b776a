/*
 * intorb_oracle.c -- CPU oracle for the orbital-form integral input (intorb)
 * of the TCE (T) path.  TEST INFRASTRUCTURE ONLY (see sdt_oracle.h): loaded
 * by tests/, never by the product library.  Literal restatements of
 *
 *   ora_v2_alpha_offset   tce_mo2e_offset_intorb.F:1-110 (the segmented
 *                         block table k_v2_alpha_offset)
 *   ora_hash_v2           tce_hash.F:1-140 (tce_hash_v2: the segment search
 *                         and the walk through the segment's blocks)
 *   ora_v2orb_pack        tce_mo2e_trans.F:686-745 (where the four-index
 *                         transformation puts (g3 g4|g1 g2) of each block)
 *   ora_sortacc_4         sort/tce_sortacc.F:50-125 (accumulating sort)
 *   ora_get_block_ind_i   get_block_ind.F:818-1470 (ioalg 2: the six spin
 *                         cases and the sixteen accumulating sorts, as
 *                         written)
 *
 * Pinning (tests/test_intorb.py):
 *   ora_v2_alpha_offset .... block count and total size by brute-force
 *                            enumeration; entries at the segment bounds
 *   ora_hash_v2 ............ every stored key's offset equals the running
 *                            sum of block sizes in storage order
 *   ora_get_block_ind_i .... on integrals with the 8-fold symmetry of real
 *                            orbitals, every spin-orbital block equals the
 *                            dense <pq||rs> = (pr|qs) d(sp,sr) d(sq,ss)
 *                            - (ps|qr) d(sp,ss) d(sq,sr) built with numpy;
 *                            a planted sign flip of the second term fails it
 */
#include <stdlib.h>
#include <string.h>

#include "sdt_oracle.h"

#define ASP(io, g) ((io)->spin_alpha[(g) - 1])
#define ASY(io, g) ((io)->sym_alpha[(g) - 1])
#define ARG(io, g) ((io)->range_alpha[(g) - 1])

static long index_pair(long i, long j) { return (i * (i - 1)) / 2 + j; }

/* tce_mo2e_offset_intorb.F:1-110. */
long ora_v2_alpha_offset(const ora_intorb_t *io, long idiv2e, long *hash, long *size) {
  const long N = io->noa + io->nva;
  long length = 0;
  for (long g3b = 1; g3b <= N; g3b++)
    for (long g4b = g3b; g4b <= N; g4b++)
      for (long g1b = 1; g1b <= N; g1b++)
        for (long g2b = g1b; g2b <= N; g2b++)
          if (ASP(io, g3b) + ASP(io, g4b) == ASP(io, g1b) + ASP(io, g2b) &&
              (ASY(io, g3b) ^ ASY(io, g4b) ^ ASY(io, g1b) ^ ASY(io, g2b)) == io->irrep_v &&
              index_pair(g4b, g3b) >= index_pair(g2b, g1b))
            length++;
  long length1 = 0;
  const long ipiece_l = length / idiv2e;
  if (ipiece_l * idiv2e == length)
    length1 = idiv2e;
  else if (ipiece_l * idiv2e < length)
    length1 = idiv2e + 1;
  if (length1 == 0) return -1;
  const long words = 6 * (length1 + 1) + 1;
  if (!hash) return words;
  memset(hash, 0, words * sizeof(long));
  hash[0] = length1;
  long addr = 0, sz = 0, i_counter = 0, ipos = 1, size_last = 0;
  long g3a = 0, g4a = 0, g1a = 0, g2a = 0;
#define PUT(K, O, A, B, C, D)                  \
  do {                                         \
    hash[ipos] = (K);                          \
    hash[(length1 + 1) + ipos] = (O);          \
    hash[2 * (length1 + 1) + ipos] = (A);      \
    hash[3 * (length1 + 1) + ipos] = (B);      \
    hash[4 * (length1 + 1) + ipos] = (C);      \
    hash[5 * (length1 + 1) + ipos] = (D);      \
    ipos++;                                    \
  } while (0)
  for (long g3b = 1; g3b <= N; g3b++)
    for (long g4b = g3b; g4b <= N; g4b++)
      for (long g1b = 1; g1b <= N; g1b++)
        for (long g2b = g1b; g2b <= N; g2b++) {
          if (ASP(io, g3b) + ASP(io, g4b) != ASP(io, g1b) + ASP(io, g2b)) continue;
          if ((ASY(io, g3b) ^ ASY(io, g4b) ^ ASY(io, g1b) ^ ASY(io, g2b)) != io->irrep_v) continue;
          if (index_pair(g4b, g3b) < index_pair(g2b, g1b)) continue;
          const long key = g2b - 1 + N * (g1b - 1 + N * (g4b - 1 + N * (g3b - 1)));
          i_counter++;
          if (addr == 0) PUT(key, sz, g3b, g4b, g1b, g2b);
          if (i_counter == ipiece_l) {
            PUT(key, sz, g3b, g4b, g1b, g2b);
            i_counter = 0;
          }
          addr += 2;
          size_last = sz;
          g3a = g3b, g4a = g4b, g1a = g1b, g2a = g2b;
          sz += ARG(io, g3b) * ARG(io, g4b) * ARG(io, g1b) * ARG(io, g2b);
        }
  if (i_counter != 0) {
    const long key = g2a - 1 + N * (g1a - 1 + N * (g4a - 1 + N * (g3a - 1)));
    PUT(key, size_last, g3a, g4a, g1a, g2a);
  }
#undef PUT
  if (size) *size = sz;
  return words;
}

/* tce_hash.F:1-140 (tce_hash_v2): the segment whose key range holds key,
 * then the reference's walk from the segment's first block, adding block
 * sizes until the key.  -1 when not found (the reference stops there). */
long ora_hash_v2(const ora_intorb_t *io, const long *hash, long key) {
  const long N = io->noa + io->nva;
  const long length = hash[0];
  long middle = -1;
  for (long i = 1; i <= length; i++)
    if (hash[i] <= key && key <= hash[i + 1]) {
      middle = i;
      break;
    }
  if (middle == -1) return -1;
  /* hash(k) of the reference is hash[k - 1] here */
  const long i_start = hash[2 * (length + 1) + middle];
  const long j_start = hash[3 * (length + 1) + middle];
  const long k_start = hash[4 * (length + 1) + middle];
  const long l_start = hash[5 * (length + 1) + middle];
  const long i_stop = hash[2 * (length + 1) + middle + 1];
  const long j_stop = hash[3 * (length + 1) + middle + 1];
  const long k_stop = hash[4 * (length + 1) + middle + 1];
  const long l_stop = hash[5 * (length + 1) + middle + 1];
  long offset = hash[(length + 1) + middle];
#define INDX_POINT(I, J, n) (((I) * (2 * (n) + 1 - (I))) / 2 - (n) + (J))
  const long pos1_l = INDX_POINT(i_start, j_start, N);
  const long pos1_u = INDX_POINT(i_stop, j_stop, N);
  long i = i_start, j = j_start, k = k_start, l = l_start;
  for (long pos1 = pos1_l; pos1 <= pos1_u; pos1++) {
    const long sa_ij = ASP(io, i) + ASP(io, j);
    for (;;) {
      const long sa_kl = ASP(io, k) + ASP(io, l);
      const long ieo_kl = ASY(io, k) ^ ASY(io, l);
      if (sa_ij == sa_kl && (ASY(io, i) ^ ASY(io, j) ^ ieo_kl) == io->irrep_v) {
        const long irow = index_pair(j, i), icol = index_pair(l, k);
        if (irow >= icol) {
          const long key_loop = l - 1 + N * (k - 1 + N * (j - 1 + N * (i - 1)));
          if (key_loop == key) return offset;
          offset += ARG(io, i) * ARG(io, j) * ARG(io, k) * ARG(io, l);
        }
      }
      if (i == i_stop && j == j_stop && k == k_stop && l == l_stop) return -1;
      if (k == N && l == N) break;
      if (l == N) {
        k++;
        l = k;
      } else {
        l++;
      }
    }
    k = 1;
    l = 1;
    if (pos1 + 1 == INDX_POINT(i + 1, i + 1, N)) {
      i++;
      j = i;
    } else {
      j++;
    }
  }
#undef INDX_POINT
  return -1;
}

/* tce_mo2e_trans.F:686-745: block (g3b <= g4b, g1b <= g2b) of the storage
 * order holds (k l | i j) at [j][i][l][k], k an orbital of g4b, l of g3b, i
 * of g2b, j of g1b; orb is the dense (pq|rs) over the n spatial orbitals
 * (C order, alpha-tile order). */
void ora_v2orb_pack(const ora_intorb_t *io, const double *orb, long n, double *v2orb) {
  const long N = io->noa + io->nva;
  long *ofs = (long *)malloc(N * sizeof(long));
  long o = 0;
  for (long g = 1; g <= N; g++) {
    ofs[g - 1] = o;
    o += ARG(io, g);
  }
  long pos = 0;
  for (long g3b = 1; g3b <= N; g3b++)
    for (long g4b = g3b; g4b <= N; g4b++)
      for (long g1b = 1; g1b <= N; g1b++)
        for (long g2b = g1b; g2b <= N; g2b++) {
          if (ASP(io, g3b) + ASP(io, g4b) != ASP(io, g1b) + ASP(io, g2b)) continue;
          if ((ASY(io, g3b) ^ ASY(io, g4b) ^ ASY(io, g1b) ^ ASY(io, g2b)) != io->irrep_v) continue;
          if (index_pair(g4b, g3b) < index_pair(g2b, g1b)) continue;
          const long size_k = ARG(io, g4b), size_l = ARG(io, g3b);
          const long size_i = ARG(io, g2b), size_j = ARG(io, g1b);
          for (long k = 1; k <= size_k; k++)
            for (long l = 1; l <= size_l; l++)
              for (long i = 1; i <= size_i; i++)
                for (long j = 1; j <= size_j; j++) {
                  const long ipos1 = (((j - 1) * size_i + i - 1) * size_l + l - 1) * size_k + k;
                  const long K = ofs[g4b - 1] + k - 1, L = ofs[g3b - 1] + l - 1;
                  const long I = ofs[g2b - 1] + i - 1, J = ofs[g1b - 1] + j - 1;
                  v2orb[pos + ipos1 - 1] = orb[((K * n + L) * n + I) * n + J];
                }
          pos += size_k * size_l * size_i * size_j;
        }
  free(ofs);
}

/* sort/tce_sortacc.F:50-125. */
void ora_sortacc_4(const double *unsorted, double *sorted, long a, long b, long c, long d,
                   long i, long j, long k, long l, double factor) {
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
          sorted[ib - 1] = sorted[ib - 1] + unsorted[ia - 1] * factor;
        }
      }
    }
  }
}

/* get_block_ind.F:818-1470, ioalg 2 (w2b, w1b, w4b, w3b = g2b, g1b, g4b,
 * g3b of the spin-orbital block v(g3b g4b g1b g2b)).  Returns 0, or -1 when
 * an orbital-form key is not in the table. */
int ora_get_block_ind_i(const ora_tiles_t *t, const ora_intorb_t *io, const double *v2orb,
                        const long *hash_alpha, long w2b, long w1b, long w4b, long w3b,
                        double *array) {
  const long N = io->noa + io->nva;
  const long *b2am = io->b2am;
  long i, j, k, l, i1, j1, k1, l1;
  /* first_h: pairs (w1b, w3b) and (w2b, w4b) */
  j1 = b2am[w3b - 1], i1 = b2am[w1b - 1], l1 = b2am[w4b - 1], k1 = b2am[w2b - 1];
  if (i1 >= j1) i = i1, j = j1; else i = j1, j = i1;
  if (k1 >= l1) k = k1, l = l1; else k = l1, l = k1;
  long irow = index_pair(i, j), icol = index_pair(k, l);
  const long first_h = irow >= icol ? k - 1 + N * (l - 1 + N * (i - 1 + N * (j - 1)))
                                    : i - 1 + N * (j - 1 + N * (k - 1 + N * (l - 1)));
  /* second_h: pairs (w2b, w3b) and (w1b, w4b) */
  j1 = b2am[w3b - 1], i1 = b2am[w2b - 1], l1 = b2am[w4b - 1], k1 = b2am[w1b - 1];
  if (i1 >= j1) i = i1, j = j1; else i = j1, j = i1;
  if (k1 >= l1) k = k1, l = l1; else k = l1, l = k1;
  irow = index_pair(i, j), icol = index_pair(k, l);
  const long second_h = irow >= icol ? k - 1 + N * (l - 1 + N * (i - 1 + N * (j - 1)))
                                     : i - 1 + N * (j - 1 + N * (k - 1 + N * (l - 1)));

  const long g3b = w3b, g4b = w4b, g1b = w1b, g2b = w2b;
  const long size = t->range[g1b - 1] * t->range[g2b - 1] * t->range[g3b - 1] * t->range[g4b - 1];
  const long s3 = t->spin[g3b - 1], s4 = t->spin[g4b - 1], s1 = t->spin[g1b - 1],
             s2 = t->spin[g2b - 1];
  const long ispin = s3 + s4 + s1 + s2;
  const int uaadaa = ispin == 4, ubbdbb = ispin == 8;
  const int uabdab = s3 == 1 && s4 == 2 && s1 == 1 && s2 == 2;
  const int ubadba = s3 == 2 && s4 == 1 && s1 == 2 && s2 == 1;
  const int ubadab = s3 == 2 && s4 == 1 && s1 == 1 && s2 == 2;
  const int uabdba = s3 == 1 && s4 == 2 && s1 == 2 && s2 == 1;
  const long size1 = t->range[g1b - 1], size2 = t->range[g2b - 1];
  const long size3 = t->range[g3b - 1], size4 = t->range[g4b - 1];
  const long ig2b = b2am[g2b - 1], ig1b = b2am[g1b - 1], ig3b = b2am[g3b - 1],
             ig4b = b2am[g4b - 1];

  if (uaadaa || ubbdbb || uabdab || ubadba) {
    const long off_a = ora_hash_v2(io, hash_alpha, first_h);
    if (off_a < 0) return -1;
    const double *f_a = v2orb + off_a;
    for (long x = 0; x < size; x++) array[x] = 0.0;
    int l31s, l42s, lp31p42;
    if (ig3b >= ig1b) l31s = 0, irow = index_pair(ig3b, ig1b);
    else l31s = 1, irow = index_pair(ig1b, ig3b);
    if (ig4b >= ig2b) l42s = 0, icol = index_pair(ig4b, ig2b);
    else l42s = 1, icol = index_pair(ig2b, ig4b);
    lp31p42 = irow >= icol ? 0 : 1;
    /* ( g3 g1 | g4 g2 ) ... ( g2 g4 | g1 g3 ) */
    if (!lp31p42 && !l31s && !l42s) ora_sortacc_4(f_a, array, size2, size1, size4, size3, 4, 2, 3, 1, 1.0);
    if (!lp31p42 && !l31s && l42s) ora_sortacc_4(f_a, array, size4, size2, size1, size3, 4, 1, 3, 2, 1.0);
    if (!lp31p42 && l31s && !l42s) ora_sortacc_4(f_a, array, size2, size4, size3, size1, 3, 2, 4, 1, 1.0);
    if (!lp31p42 && l31s && l42s) ora_sortacc_4(f_a, array, size4, size2, size3, size1, 3, 1, 4, 2, 1.0);
    if (lp31p42 && !l31s && !l42s) ora_sortacc_4(f_a, array, size1, size3, size2, size4, 2, 4, 1, 3, 1.0);
    if (lp31p42 && l31s && !l42s) ora_sortacc_4(f_a, array, size3, size1, size2, size4, 1, 4, 2, 3, 1.0);
    if (lp31p42 && !l31s && l42s) ora_sortacc_4(f_a, array, size1, size3, size4, size2, 2, 3, 1, 4, 1.0);
    if (lp31p42 && l31s && l42s) ora_sortacc_4(f_a, array, size3, size1, size4, size2, 1, 3, 2, 4, 1.0);
  }
  if (uaadaa || ubbdbb || uabdba || ubadab) {
    int l32s, l41s, lp32p41;
    if (ig3b >= ig2b) l32s = 0, irow = index_pair(ig3b, ig2b);
    else l32s = 1, irow = index_pair(ig2b, ig3b);
    if (ig4b >= ig1b) l41s = 0, icol = index_pair(ig4b, ig1b);
    else l41s = 1, icol = index_pair(ig1b, ig4b);
    lp32p41 = irow >= icol ? 0 : 1;
    if (uabdba || ubadab)
      for (long x = 0; x < size; x++) array[x] = 0.0;
    const long off_a = ora_hash_v2(io, hash_alpha, second_h);
    if (off_a < 0) return -1;
    const double *f_a = v2orb + off_a;
    /* ( g3 g2 | g4 g1 ) ... ( g1 g4 | g2 g3 ) */
    if (!lp32p41 && !l32s && !l41s) ora_sortacc_4(f_a, array, size1, size4, size2, size3, 4, 2, 1, 3, -1.0);
    if (!lp32p41 && !l32s && l41s) ora_sortacc_4(f_a, array, size2, size1, size4, size3, 4, 1, 2, 3, -1.0);
    if (!lp32p41 && l32s && !l41s) ora_sortacc_4(f_a, array, size1, size4, size3, size2, 3, 2, 1, 4, -1.0);
    if (!lp32p41 && l32s && l41s) ora_sortacc_4(f_a, array, size1, size4, size3, size2, 3, 1, 2, 4, -1.0);
    if (lp32p41 && !l32s && !l41s) ora_sortacc_4(f_a, array, size2, size3, size1, size4, 2, 4, 3, 1, -1.0);
    if (lp32p41 && l32s && !l41s) ora_sortacc_4(f_a, array, size3, size2, size1, size4, 1, 4, 3, 2, -1.0);
    if (lp32p41 && !l32s && l41s) ora_sortacc_4(f_a, array, size2, size1, size4, size3, 2, 3, 4, 1, -1.0);
    if (lp32p41 && l32s && l41s) ora_sortacc_4(f_a, array, size3, size2, size1, size4, 1, 3, 4, 2, -1.0);
  }
  return 0;
}

/* Every block of a tce_mo2e_offset table through ora_get_block_ind_i, as
 * the (T) drivers fetch them with intorb (ccsd_t_singles_gpu.F:245,
 * ccsd_t_doubles_gpu.F:316,965).  Returns 0 or -1. */
int ora_v2_from_intorb(const ora_tiles_t *t, const ora_intorb_t *io, const double *v2orb,
                       const long *hash_alpha, const long *v2_hash, double *v2) {
  const long N = t->noab + t->nvab;
  const long L = v2_hash[0];
  for (long x = 0; x < L; x++) {
    long key = v2_hash[1 + x];
    const long off = v2_hash[1 + L + x];
    const long g2 = key % N + 1;
    key /= N;
    const long g1 = key % N + 1;
    key /= N;
    const long g4 = key % N + 1;
    const long g3 = key / N + 1;
    if (ora_get_block_ind_i(t, io, v2orb, hash_alpha, g2, g1, g4, g3, v2 + off) != 0) return -1;
  }
  return 0;
}
