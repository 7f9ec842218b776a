/* Lambda-CCSD(T) inputs of the oracle -- TEST INFRASTRUCTURE ONLY.
 *
 * C restatement of the reference's block tables for the Lambda amplitudes
 * and the one-electron operator, and of its six-index accumulating sort
 * (the energy composition itself is numpy, oracle/__init__.py):
 *   ora_y1_offset    tce_y1_offset.F:1-55
 *   ora_y2_offset    tce_y2_offset.F:1-69
 *   ora_f1_offset    tce_mo1e_offset.F:1-56
 *   ora_sortacc_6    sort/tce_sortacc.F:126-280 (the loop order of its
 *                    n = 6 or 5 branch; the other branches walk the same
 *                    elements in another order, and each element of the
 *                    sorted array receives exactly one product per call)
 *
 * Pinning (tests/test_lambda.py):
 *   ora_y1/y2/f1_offset ... brute-force enumeration of the blocks in Python
 *                           (keys, order, running offsets)
 *   ora_sortacc_6 ......... numpy.transpose of tagged arrays for every
 *                           permutation of six axes, factor and accumulation
 */
#include "sdt_oracle.h"

#define SPN(t, b) ((t)->spin[(b) - 1])
#define SYM(t, b) ((t)->sym[(b) - 1])
#define RNG(t, b) ((t)->range[(b) - 1])

long ora_y1_offset(const ora_tiles_t *t, long *hash, long *size) {
  const long noab = t->noab, nvab = t->nvab;
  long length = 0, sz = 0;
  for (int pass = 0; pass < 2; pass++) {
    long addr = 0;
    for (long h4b = 1; h4b <= noab; h4b++)
      for (long p1b = noab + 1; p1b <= noab + nvab; p1b++) {
        if (SPN(t, h4b) != SPN(t, p1b) || (SYM(t, h4b) ^ SYM(t, p1b)) != 0) continue;
        if (t->restricted && SPN(t, h4b) + SPN(t, p1b) == 4) continue;
        if (pass == 0) {
          length++;
          continue;
        }
        addr++;
        if (hash) {
          hash[addr] = p1b - noab - 1 + nvab * (h4b - 1);
          hash[length + addr] = sz;
        }
        sz += RNG(t, h4b) * RNG(t, p1b);
      }
  }
  if (hash) hash[0] = length;
  if (size) *size = sz;
  return length;
}

long ora_y2_offset(const ora_tiles_t *t, long *hash, long *size) {
  const long noab = t->noab, nvab = t->nvab;
  long length = 0, sz = 0;
  for (int pass = 0; pass < 2; pass++) {
    long addr = 0;
    for (long h7b = 1; h7b <= noab; h7b++)
      for (long h8b = h7b; h8b <= noab; h8b++)
        for (long p1b = noab + 1; p1b <= noab + nvab; p1b++)
          for (long p2b = p1b; p2b <= noab + nvab; p2b++) {
            if (SPN(t, h7b) + SPN(t, h8b) != SPN(t, p1b) + SPN(t, p2b)) continue;
            if ((SYM(t, h7b) ^ SYM(t, h8b) ^ SYM(t, p1b) ^ SYM(t, p2b)) != 0) continue;
            if (t->restricted && SPN(t, h7b) + SPN(t, h8b) + SPN(t, p1b) + SPN(t, p2b) == 8)
              continue;
            if (pass == 0) {
              length++;
              continue;
            }
            addr++;
            if (hash) {
              hash[addr] = p2b - noab - 1 + nvab * (p1b - noab - 1 + nvab * (h8b - 1 + noab * (h7b - 1)));
              hash[length + addr] = sz;
            }
            sz += RNG(t, h7b) * RNG(t, h8b) * RNG(t, p1b) * RNG(t, p2b);
          }
  }
  if (hash) hash[0] = length;
  if (size) *size = sz;
  return length;
}

long ora_f1_offset(const ora_tiles_t *t, long *hash, long *size) {
  const long N = t->noab + t->nvab;
  long length = 0, sz = 0;
  for (int pass = 0; pass < 2; pass++) {
    long addr = 0;
    for (long g6b = 1; g6b <= N; g6b++)
      for (long g5b = 1; g5b <= N; g5b++) {
        if (SPN(t, g6b) != SPN(t, g5b) || (SYM(t, g6b) ^ SYM(t, g5b)) != 0) continue;
        if (t->restricted && SPN(t, g6b) + SPN(t, g5b) == 4) continue;
        if (pass == 0) {
          length++;
          continue;
        }
        addr++;
        if (hash) {
          hash[addr] = g5b - 1 + N * (g6b - 1);
          hash[length + addr] = sz;
        }
        sz += RNG(t, g6b) * RNG(t, g5b);
      }
  }
  if (hash) hash[0] = length;
  if (size) *size = sz;
  return length;
}

void ora_sortacc_6(const double *unsorted, double *sorted, long a, long b, long c, long d,
                   long e, long f, long i, long j, long k, long l, long m, long n,
                   double factor) {
  long jd[7] = {0, a, b, c, d, e, f};
  long id[7];
  for (long j1 = 1; j1 <= a; j1++) {
    id[1] = j1;
    for (long j2 = 1; j2 <= b; j2++) {
      id[2] = j2;
      for (long j3 = 1; j3 <= c; j3++) {
        id[3] = j3;
        for (long j4 = 1; j4 <= d; j4++) {
          id[4] = j4;
          for (long j5 = 1; j5 <= e; j5++) {
            id[5] = j5;
            for (long j6 = 1; j6 <= f; j6++) {
              id[6] = j6;
              long ia = id[6] + jd[6] * (id[5] - 1 + jd[5] * (id[4] - 1 + jd[4] *
                        (id[3] - 1 + jd[3] * (id[2] - 1 + jd[2] * (id[1] - 1)))));
              long ib = id[n] + jd[n] * (id[m] - 1 + jd[m] * (id[l] - 1 + jd[l] *
                        (id[k] - 1 + jd[k] * (id[j] - 1 + jd[j] * (id[i] - 1)))));
              sorted[ib - 1] = sorted[ib - 1] + unsorted[ia - 1] * factor;
            }
          }
        }
      }
    }
  }
}
