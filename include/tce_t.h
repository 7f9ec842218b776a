/*
 * tce_t.h -- C-ABI of the MI355X (gfx950) CCSD(T) perturbative-triples library
 * (libtce_t.so).  Plain C types only: 64-bit `long` for every Fortran INTEGER
 * (the reference is built with -fdefault-integer-8), `double` for REAL*8.
 *
 * Three groups of entry points:
 *
 *  1. The reference's own GPU kernel surface (Fortran-mangled names, every
 *     argument by reference).  NWChem's src/tce/ccsd_t drivers
 *     (ccsd_t_gpu.F, ccsd_t_singles_gpu.F, ccsd_t_doubles_gpu.F) link against
 *     these unchanged.  Each replaces the function of the same name in
 *     src/tce/ccsd_t/sd_t_total.cu, memory.cu and hybrid.c (cited per entry).
 *
 *  2. The whole-job API (tce_t_job_*): the tile-sextuple loop of
 *     ccsd_t_gpu.F:87-241 as one device-resident job -- T1/T2/V2 blocks
 *     uploaded once through the reference's offset hash tables, every task
 *     contracted by fused HIP kernels on FP64 MFMA, energies reduced on the
 *     device.  Returns status codes, never exits.
 *
 *  3. Host helpers that mirror the reference's integer bookkeeping
 *     (tce_tile.F ranges, tce_*_offset.F hash tables, tce_hash.F) and the
 *     seeded synthetic inputs used by tests and benchmarks.
 */
#ifndef TCE_T_H
#define TCE_T_H
#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------ status ---- */
#define TCE_T_OK 0
#define TCE_T_ERR_ARG (-1)     /* invalid argument or unsupported shape     */
#define TCE_T_ERR_HIP (-2)     /* HIP runtime error (see tce_t_last_error)  */
#define TCE_T_ERR_NOMEM (-3)   /* device or host allocation failed          */
#define TCE_T_ERR_KEY (-4)     /* a block key is missing from a hash table  */
#define TCE_T_ERR_STATE (-5)   /* call out of order (e.g. run before upload) */

/* Text of the last error on the calling thread ("" if none). */
const char *tce_t_last_error(void);

/* ================================================================ 1 ===== */
/* Reference GPU kernel surface.  Dimensions are the ranges of the dispatch
 * row's tiles; `t3` is ignored (as in the reference): results accumulate in
 * the library-owned device tiles created by dev_mem_s_/dev_mem_d_. Operand
 * arrays are host buffers owned by the caller and valid for the call only. */

/* sd_t_total.cu:5435-5526 and siblings -- singles, triplesx_s(L_N) +-=
 * t1sub(p4,h1) * v2sub(h3,h2,p6,p5). */
void sd_t_s1_1_cuda_(long *h1d, long *h2d, long *h3d, long *p4d, long *p5d, long *p6d, double *t3, double *t1, double *v2);
void sd_t_s1_2_cuda_(long *h1d, long *h2d, long *h3d, long *p4d, long *p5d, long *p6d, double *t3, double *t1, double *v2);
void sd_t_s1_3_cuda_(long *h1d, long *h2d, long *h3d, long *p4d, long *p5d, long *p6d, double *t3, double *t1, double *v2);
void sd_t_s1_4_cuda_(long *h1d, long *h2d, long *h3d, long *p4d, long *p5d, long *p6d, double *t3, double *t1, double *v2);
void sd_t_s1_5_cuda_(long *h1d, long *h2d, long *h3d, long *p4d, long *p5d, long *p6d, double *t3, double *t1, double *v2);
void sd_t_s1_6_cuda_(long *h1d, long *h2d, long *h3d, long *p4d, long *p5d, long *p6d, double *t3, double *t1, double *v2);
void sd_t_s1_7_cuda_(long *h1d, long *h2d, long *h3d, long *p4d, long *p5d, long *p6d, double *t3, double *t1, double *v2);
void sd_t_s1_8_cuda_(long *h1d, long *h2d, long *h3d, long *p4d, long *p5d, long *p6d, double *t3, double *t1, double *v2);
void sd_t_s1_9_cuda_(long *h1d, long *h2d, long *h3d, long *p4d, long *p5d, long *p6d, double *t3, double *t1, double *v2);

/* sd_t_total.cu:261-314 and siblings -- d1, triplesx(L_N) -+=
 * sum_h7 t2sub(h7,p4,p5,h1) * v2sub(h3,h2,p6,h7). */
void sd_t_d1_1_cuda_(long *h1d, long *h2d, long *h3d, long *h7d, long *p4d, long *p5d, long *p6d, double *t3, double *t2, double *v2);
void sd_t_d1_2_cuda_(long *h1d, long *h2d, long *h3d, long *h7d, long *p4d, long *p5d, long *p6d, double *t3, double *t2, double *v2);
void sd_t_d1_3_cuda_(long *h1d, long *h2d, long *h3d, long *h7d, long *p4d, long *p5d, long *p6d, double *t3, double *t2, double *v2);
void sd_t_d1_4_cuda_(long *h1d, long *h2d, long *h3d, long *h7d, long *p4d, long *p5d, long *p6d, double *t3, double *t2, double *v2);
void sd_t_d1_5_cuda_(long *h1d, long *h2d, long *h3d, long *h7d, long *p4d, long *p5d, long *p6d, double *t3, double *t2, double *v2);
void sd_t_d1_6_cuda_(long *h1d, long *h2d, long *h3d, long *h7d, long *p4d, long *p5d, long *p6d, double *t3, double *t2, double *v2);
void sd_t_d1_7_cuda_(long *h1d, long *h2d, long *h3d, long *h7d, long *p4d, long *p5d, long *p6d, double *t3, double *t2, double *v2);
void sd_t_d1_8_cuda_(long *h1d, long *h2d, long *h3d, long *h7d, long *p4d, long *p5d, long *p6d, double *t3, double *t2, double *v2);
void sd_t_d1_9_cuda_(long *h1d, long *h2d, long *h3d, long *h7d, long *p4d, long *p5d, long *p6d, double *t3, double *t2, double *v2);

/* sd_t_total.cu:2831-2888 and siblings -- d2, triplesx(L_N) -+=
 * sum_p7 t2sub(p7,p4,h1,h2) * v2sub(p7,h3,p6,p5). */
void sd_t_d2_1_cuda_(long *h1d, long *h2d, long *h3d, long *p4d, long *p5d, long *p6d, long *p7d, double *t3, double *t2, double *v2);
void sd_t_d2_2_cuda_(long *h1d, long *h2d, long *h3d, long *p4d, long *p5d, long *p6d, long *p7d, double *t3, double *t2, double *v2);
void sd_t_d2_3_cuda_(long *h1d, long *h2d, long *h3d, long *p4d, long *p5d, long *p6d, long *p7d, double *t3, double *t2, double *v2);
void sd_t_d2_4_cuda_(long *h1d, long *h2d, long *h3d, long *p4d, long *p5d, long *p6d, long *p7d, double *t3, double *t2, double *v2);
void sd_t_d2_5_cuda_(long *h1d, long *h2d, long *h3d, long *p4d, long *p5d, long *p6d, long *p7d, double *t3, double *t2, double *v2);
void sd_t_d2_6_cuda_(long *h1d, long *h2d, long *h3d, long *p4d, long *p5d, long *p6d, long *p7d, double *t3, double *t2, double *v2);
void sd_t_d2_7_cuda_(long *h1d, long *h2d, long *h3d, long *p4d, long *p5d, long *p6d, long *p7d, double *t3, double *t2, double *v2);
void sd_t_d2_8_cuda_(long *h1d, long *h2d, long *h3d, long *p4d, long *p5d, long *p6d, long *p7d, double *t3, double *t2, double *v2);
void sd_t_d2_9_cuda_(long *h1d, long *h2d, long *h3d, long *p4d, long *p5d, long *p6d, long *p7d, double *t3, double *t2, double *v2);

/* The reference's CPU kernel surface (ccsd_t_kernels_omp.F:5-1198, argument
 * order h3d, h2d, h1d, p6d, p5d, p4d[, h7d | p7d], triplesx, t1sub | t2sub,
 * v2sub): the same 27 contractions accumulated into the caller's HOST
 * triplesx (not zeroed; its layout the variant's triplesx declaration), for
 * the CPU drivers ccsd_t_singles_l.F / offl_ccsd_t_doubles_l.F and their
 * other callers (cr_ccsd_t.F, lambda_ccsd_t*.F, emb/emb2.F, the MRCC (T)).
 * Each call stages triplesx and the operands through the device and back
 * (PCIe-bound for large tiles); the whole-job and task-tile entry points
 * below are the fast paths.  These names replace ccsd_t_kernels_omp.o in
 * a link. */
void sd_t_s1_1_(long *h3d, long *h2d, long *h1d, long *p6d, long *p5d, long *p4d, double *triplesx, double *t1sub, double *v2sub);
void sd_t_s1_2_(long *h3d, long *h2d, long *h1d, long *p6d, long *p5d, long *p4d, double *triplesx, double *t1sub, double *v2sub);
void sd_t_s1_3_(long *h3d, long *h2d, long *h1d, long *p6d, long *p5d, long *p4d, double *triplesx, double *t1sub, double *v2sub);
void sd_t_s1_4_(long *h3d, long *h2d, long *h1d, long *p6d, long *p5d, long *p4d, double *triplesx, double *t1sub, double *v2sub);
void sd_t_s1_5_(long *h3d, long *h2d, long *h1d, long *p6d, long *p5d, long *p4d, double *triplesx, double *t1sub, double *v2sub);
void sd_t_s1_6_(long *h3d, long *h2d, long *h1d, long *p6d, long *p5d, long *p4d, double *triplesx, double *t1sub, double *v2sub);
void sd_t_s1_7_(long *h3d, long *h2d, long *h1d, long *p6d, long *p5d, long *p4d, double *triplesx, double *t1sub, double *v2sub);
void sd_t_s1_8_(long *h3d, long *h2d, long *h1d, long *p6d, long *p5d, long *p4d, double *triplesx, double *t1sub, double *v2sub);
void sd_t_s1_9_(long *h3d, long *h2d, long *h1d, long *p6d, long *p5d, long *p4d, double *triplesx, double *t1sub, double *v2sub);
void sd_t_d1_1_(long *h3d, long *h2d, long *h1d, long *p6d, long *p5d, long *p4d, long *h7d, double *triplesx, double *t2sub, double *v2sub);
void sd_t_d1_2_(long *h3d, long *h2d, long *h1d, long *p6d, long *p5d, long *p4d, long *h7d, double *triplesx, double *t2sub, double *v2sub);
void sd_t_d1_3_(long *h3d, long *h2d, long *h1d, long *p6d, long *p5d, long *p4d, long *h7d, double *triplesx, double *t2sub, double *v2sub);
void sd_t_d1_4_(long *h3d, long *h2d, long *h1d, long *p6d, long *p5d, long *p4d, long *h7d, double *triplesx, double *t2sub, double *v2sub);
void sd_t_d1_5_(long *h3d, long *h2d, long *h1d, long *p6d, long *p5d, long *p4d, long *h7d, double *triplesx, double *t2sub, double *v2sub);
void sd_t_d1_6_(long *h3d, long *h2d, long *h1d, long *p6d, long *p5d, long *p4d, long *h7d, double *triplesx, double *t2sub, double *v2sub);
void sd_t_d1_7_(long *h3d, long *h2d, long *h1d, long *p6d, long *p5d, long *p4d, long *h7d, double *triplesx, double *t2sub, double *v2sub);
void sd_t_d1_8_(long *h3d, long *h2d, long *h1d, long *p6d, long *p5d, long *p4d, long *h7d, double *triplesx, double *t2sub, double *v2sub);
void sd_t_d1_9_(long *h3d, long *h2d, long *h1d, long *p6d, long *p5d, long *p4d, long *h7d, double *triplesx, double *t2sub, double *v2sub);
void sd_t_d2_1_(long *h3d, long *h2d, long *h1d, long *p6d, long *p5d, long *p4d, long *p7d, double *triplesx, double *t2sub, double *v2sub);
void sd_t_d2_2_(long *h3d, long *h2d, long *h1d, long *p6d, long *p5d, long *p4d, long *p7d, double *triplesx, double *t2sub, double *v2sub);
void sd_t_d2_3_(long *h3d, long *h2d, long *h1d, long *p6d, long *p5d, long *p4d, long *p7d, double *triplesx, double *t2sub, double *v2sub);
void sd_t_d2_4_(long *h3d, long *h2d, long *h1d, long *p6d, long *p5d, long *p4d, long *p7d, double *triplesx, double *t2sub, double *v2sub);
void sd_t_d2_5_(long *h3d, long *h2d, long *h1d, long *p6d, long *p5d, long *p4d, long *p7d, double *triplesx, double *t2sub, double *v2sub);
void sd_t_d2_6_(long *h3d, long *h2d, long *h1d, long *p6d, long *p5d, long *p4d, long *p7d, double *triplesx, double *t2sub, double *v2sub);
void sd_t_d2_7_(long *h3d, long *h2d, long *h1d, long *p6d, long *p5d, long *p4d, long *p7d, double *triplesx, double *t2sub, double *v2sub);
void sd_t_d2_8_(long *h3d, long *h2d, long *h1d, long *p6d, long *p5d, long *p4d, long *p7d, double *triplesx, double *t2sub, double *v2sub);
void sd_t_d2_9_(long *h3d, long *h2d, long *h1d, long *p6d, long *p5d, long *p4d, long *p7d, double *triplesx, double *t2sub, double *v2sub);

/* sd_t_total.cu:5373-5379 -- energy of the current task's device tiles:
 * energy[0] = E[T] part, energy[1] = E(T) part; eval1..6 point at the
 * orbital energies of tiles h1,h2,h3,p4,p5,p6 (host); host1/host2 unused. */
void compute_en_(double *factor, double *energy, double *eval1, double *eval2, double *eval3, double *eval4, double *eval5, double *eval6,
                 long *h1d, long *h2d, long *h3d, long *p4d, long *p5d, long *p6d, double *host1, double *host2);

/* No counterpart in the reference: compute_en_ for callers that cannot
 * allocate device memory (code under HIP graph capture, serving loops and
 * other harnesses that forbid hipMalloc).  Same energy as compute_en_ (same
 * kernels and reduction order): t3d / t3s are the task's doubles / singles
 * tiles in DEVICE memory (h3 fastest, h2, h1, p6, p5, p4 slowest), eval1..6
 * the orbital energies of tiles h1, h2, h3, p4, p5, p6 in DEVICE memory, and
 * energy a DEVICE array of 2 doubles receiving {E[T], E(T)}.  The kernels are
 * enqueued on `stream` (a hipStream_t; NULL = the null stream) and nothing
 * is synchronised or allocated: the 1024 block partials go to a fixed
 * 16 KiB device array of the library (one per device, held until the
 * process ends), which serves tiles of any size.  Calls that overlap on one
 * device share that array and race: order them on one stream.  Returns
 * TCE_T_ERR_ARG for NULL pointers or extents <= 0. */
int tce_t_compute_en_noalloc(double factor, const double *t3d, const double *t3s,
                             const double *eval1, const double *eval2, const double *eval3,
                             const double *eval4, const double *eval5, const double *eval6,
                             long h1d, long h2d, long h3d, long p4d, long p5d, long p6d,
                             double *energy, void *stream);

/* sd_t_total.cu:5-16, 5382-5396 -- allocate and zero the task's doubles /
 * singles device tiles (h3 fastest ... p4 slowest). */
void dev_mem_d_(long *h1d, long *h2d, long *h3d, long *p4d, long *p5d, long *p6d);
void dev_mem_s_(long *h1d, long *h2d, long *h3d, long *p4d, long *p5d, long *p6d);
/* sd_t_total.cu:17-28 -- release both device tiles. */
void dev_release_(void);
/* memory.cu:69-71, 164-176 -- device pool lifecycle. */
void initmemmodule_(void);
void finalizememmodule_(void);
/* hybrid.c:25-29 -- 1 if the calling process owns a GPU (local rank < icuda).
 * The local rank is read from the launcher environment (OMPI_COMM_WORLD_LOCAL_RANK,
 * MPI_LOCALRANKID, SLURM_LOCALID, LOCAL_RANK; 0 when none is set). */
int check_device_(long *icuda);
/* hybrid.c:32-57 -- bind the process to device `local rank`; sets
 * *cuda_device_number = 30 when fewer than *icuda devices are visible. */
int device_init_(long *icuda, long *cuda_device_number);

/* ================================================================ 2 ===== */
/* Tile model: the tce.fh/MA arrays the path reads.  All arrays are indexed by
 * tile-1; holes are tiles 1..noab, particles noab+1..noab+nvab. */
typedef struct {
  long noab, nvab;
  const long *spin;   /* k_spin       */
  const long *sym;    /* k_sym        */
  const long *range;  /* k_range      */
  const long *offset; /* k_offset     */
  const long *alpha;  /* k_alpha (used when restricted), may be NULL */
  const double *evl;  /* k_evl_sorted, length = sum of ranges         */
  int restricted;
} tce_t_tiles;

/* Arithmetic of the d1/d2 contractions (s1, denominators, their division
 * and the energy sums are FP64 in every mode). */
#define TCE_T_FP64 0        /* FP64 MFMA                                 */
#define TCE_T_FP32 1        /* FP32 operands, FP32 MFMA and accumulation */
#define TCE_T_BF16X3 2      /* each operand x = hi + lo (two bf16), the  */
                            /* products hi*hi + hi*lo + lo*hi on BF16    */
                            /* MFMA with FP32 accumulation               */
#define TCE_T_BF16X6 3      /* x = hi + mid + lo (three bf16), the six   */
                            /* products down to 2^-24 relative (hi*hi,   */
                            /* hi*mid, mid*hi, hi*lo, mid*mid, lo*hi) on */
                            /* BF16 MFMA with FP32 accumulation          */

typedef struct tce_t_job tce_t_job;

/* Bind the library to a device (call once per process). */
int tce_t_init(int device);
/* Free and total device memory in bytes (hipMemGetInfo) of the current device. */
int tce_t_device_memory(long *free_bytes, long *total_bytes);

/* Plan a job: enumerate the tile sextuples of ccsd_t_gpu.F:87-112 and size
 * the device tensors.  The tile arrays are copied.  Kernel tasks: the
 * default tiling (TCE_T_TILING_SUBTILE, or TCE_T_SUBTILE=0|1|2 in the
 * environment for REFERENCE | SUBTILE | ORBITAL). */
int tce_t_job_create(const tce_t_tiles *tiles, tce_t_job **job);
/* Tiling of the kernel tasks (no reference counterpart; the energies are
 * the same sums, up to FP64 rounding, in every mode):
 *   REFERENCE  the reference tiles, boxes over each (the reference's work);
 *   SUBTILE    box-sized sub-tiles of each reference tile (4 occupied, 4 or
 *              8 virtual orbitals) in canonical order: no symmetric
 *              duplicates inside equal tiles; per-task energies and restart
 *              blocks as REFERENCE;
 *   ORBITAL    box-sized pieces of each run of consecutive tiles of one spin
 *              and irrep, ignoring the reference tile edges (less padding
 *              with ragged tiles); E[T] and E(T) only: task energies and
 *              restart blocks are TCE_T_ERR_STATE. */
#define TCE_T_TILING_REFERENCE 0
#define TCE_T_TILING_SUBTILE 1
#define TCE_T_TILING_ORBITAL 2
int tce_t_job_create_tiling(const tce_t_tiles *tiles, int tiling, tce_t_job **job);
int tce_t_job_tiling(const tce_t_job *job);
long tce_t_job_num_tasks(const tce_t_job *job);
/* Device bytes the job holds once uploaded (admission control), the tile
 * export buffer included. */
long tce_t_job_device_bytes(const tce_t_job *job);
/* Additional device bytes the first run in a reduced precision allocates
 * (converted copies of the four contraction tensors); 0 for TCE_T_FP64. */
long tce_t_job_precision_bytes(const tce_t_job *job, int precision);

/* Upload T1, T2 and V2 (again: replaces every device buffer of the job,
 * reduced-precision copies included).  Arguments mirror d_t1/k_t1_offset,
 * d_t2/k_t2_offset, d_v2/k_v2_offset of ccsd_t_gpu.F:2-3: host copies of the
 * canonical TCE blocks and their hash tables [L, keys(1..L), offsets(1..L)].
 * Every block the (T) path reads is fetched through the hash table exactly as
 * get_hash_block (+ tce_sort_2/4 and their signs) would, and re-laid out on
 * the device: only the file ranges of those blocks are copied (the v2 file
 * may hold every integral class), the sort runs on the GPU.  TCE_T_PACK=host
 * in the environment sorts on the host instead (same values).
 * TCE_T_ERR_KEY if a needed key is absent.  Synchronous. */
int tce_t_job_upload(tce_t_job *job, const double *t1, const long *k_t1_offset,
                     const double *t2, const long *k_t2_offset,
                     const double *v2, const long *k_v2_offset);

/* Run every task; *e_bt = E[T] (energy1), *e_pt = E(T) (energy2) of
 * ccsd_t_gpu.F:243-244.  precision: TCE_T_FP64, TCE_T_FP32, TCE_T_BF16X3 or
 * TCE_T_BF16X6
 * (the first run in a reduced precision converts the contraction tensors on
 * the device; see tce_t_job_precision_bytes).  Synchronous. */
int tce_t_job_run(tce_t_job *job, int precision, double *e_bt, double *e_pt);
/* Restartable run (ccsd_t_restart.F:58-262): the tasks in blocks of one
 * outer virtual tile p4b = noab + b, for b = *begin .. nvab (1-based).
 * Each block's E(T) goes into table[b-1] (nvab entries, kept by the caller
 * like tce:restart_triples_table) and *begin advances past it (like
 * tce:ccsd_t_restart_begin), at most max_tiles blocks per call (<= 0: all),
 * so a caller that saves (*begin, table) after each call can resume after a
 * stop.  Once *begin > nvab, *t_energy = sum(table).  Table entries of
 * blocks without allowed tasks are 0. */
int tce_t_job_run_restart(tce_t_job *job, int precision, long *begin, double *table,
                          long max_tiles, double *t_energy);
/* Triples tiles of one task, for the consumers of the reference's CPU
 * drivers other than the (T) energy (cr_ccsd_t.F:139-144, lambda_ccsd_t.F,
 * emb/emb2.F:293, mrcc/ccsd_t/tce_mrcc_ccsdpt_subg.F:202).  Replaces
 * ccsd_t_singles_l(singles, ..., t_h1b..t_p6b, 2) (ccsd_t_singles_l.F:1-28)
 * and ccsd_t_doubles_l(doubles, ..., t_h1b..t_p6b, 2) (ccsd_t_doubles_l.F:
 * 1-100) on the job's resident tensors: adds t3s to singles and t3d to
 * doubles, host arrays of r(h1b) r(h2b) r(h3b) r(p4b) r(p5b) r(p6b) doubles in
 * triplesx(h3,h2,h1,p6,p5,p4) order (Fortran), as the drivers accumulate
 * into the caller's zeroed buffers.  Either array may be NULL.  Tiles are
 * 1-based (holes 1..noab, particles noab+1..), h1b <= h2b <= h3b,
 * p4b <= p5b <= p6b, and the sextuple must pass the task loop's spin, irrep
 * and restricted tests (else TCE_T_ERR_ARG).  Any tiling; d1/d2 products in
 * `precision`.  Synchronous.  The job keeps a device buffer of the largest
 * tile pair exported (16 bytes per element) until it is released or
 * uploaded again; tce_t_job_device_bytes counts it. */
int tce_t_job_triples_tiles(tce_t_job *job, int precision, long h1b, long h2b, long h3b,
                            long p4b, long p5b, long p6b, double *singles, double *doubles);
/* The memory-sliced form of the same (ccsd_t_6dts.F:238-266, the xmem path):
 * ccsd_t_singles_l_6dts / ccsd_t_doubles_l_6dts for the slice p4 in
 * [istart, istop], p5 in [jstart, jstop] (1-based within tiles p4b, p5b) of
 * the task's t3s / t3d, added into singles / doubles laid out
 * triplesx(h3,h2,h1,p6,maxp5,maxp4) as the 6DTS kernels
 * (ccsd_t_dts_kernels_omp.F) leave it.  istop - istart + 1 <= maxp4 and
 * jstop - jstart + 1 <= maxp5 (else TCE_T_ERR_ARG).  Same elements, bit for
 * bit, as the corresponding part of tce_t_job_triples_tiles. */
int tce_t_job_triples_tiles_6dts(tce_t_job *job, int precision, long h1b, long h2b, long h3b,
                                 long p4b, long p5b, long p6b, long istart, long istop,
                                 long jstart, long jstop, long maxp4, long maxp5,
                                 double *singles, double *doubles);
/* The same tiles left in device memory, for consumers that run their own
 * functional on the GPU (no copy over PCIe): *doubles and *singles point to
 * the task's t3d and t3s in HBM, *elements (may be NULL) doubles each, in
 * the layout of tce_t_job_triples_tiles (assigned, not accumulated).  The
 * arrays belong to the job: valid until its next tile, Lambda or upload
 * call, or its destruction; the tiles are complete when this returns.  No
 * counterpart in the reference (its drivers only fill host buffers). */
int tce_t_job_triples_tiles_device(tce_t_job *job, int precision, long h1b, long h2b, long h3b,
                                   long p4b, long p5b, long p6b, const double **singles,
                                   const double **doubles, long *elements);
/* {E[T], E(T)} of each task in the order of ccsd_t_gpu.F's task loop, from
 * the last run.  out has 2*ntasks doubles. */
int tce_t_job_task_energies(const tce_t_job *job, double *out, long ntasks);

/* ---- Lambda-CCSD(T) (lambda_ccsd_t.F) ---------------------------------- */
/* Upload the Lambda amplitudes as a job's amplitudes: d_y1/k_l1_offset,
 * d_y2/k_l2_offset of lambda_ccsd_t.F:1-3 (host copies; the block tables of
 * tce_y1_offset.F and tce_y2_offset.F, blocks y(h p) as [h][p] and
 * y(h h' p p') as [h][h'][p][p']) and the same d_v2/k_v2_offset as the
 * right-hand job.  Each block is placed where tce_t_job_upload expects the
 * t1/t2 block of the same tiles, then the job is uploaded.  Such a job's
 * singles tiles are P(9) y(h4 p1) v(h5 h6 p2 p3) and its doubles tiles the
 * two v2 terms of lambda_ccsd_t_left.F:7-10 (real orbitals: the (T)
 * orientation of each integral block serves).  TCE_T_ERR_ARG if a y table
 * lacks a block the t tables hold. */
int tce_t_job_upload_lambda(tce_t_job *job, const double *y1, const long *k_y1_offset,
                            const double *y2, const long *k_y2_offset, const double *v2,
                            const long *k_v2_offset);
/* Element pairing of tce_t_lambda_ccsd_t: the t3 and y3 elements of the
 * same orbitals, or element i of tdoubles (order [p4][p5][p6][h1][h2][h3])
 * with element i of the left tiles (order [h1][h2][h3][p4][p5][p6]), as the
 * energy loop of lambda_ccsd_t.F:156-178 pairs them.  The two agree when
 * every tile holds one orbital. */
#define TCE_T_LAMBDA_ORBITAL 0
#define TCE_T_LAMBDA_REFERENCE_ORDER 1
/* Replaces lambda_ccsd_t(..., energy1, energy2) (lambda_ccsd_t.F:1-214):
 * right holds t1/t2/v2 (tce_t_job_upload), left the Lambda amplitudes
 * (tce_t_job_upload_lambda) over the same tiles.  f1/k_f1_offset: the
 * one-electron operator d_f1/k_f1_offset (block table of tce_mo1e_offset.F),
 * whose hole-particle blocks enter P(9) y(h4 h5 p1 p2) f(h6 p3); both NULL
 * drops that term.  Per task of the (T) loop, in `precision`:
 *   energy1 += factor * tdoubles * ydoubles / D
 *   energy2 += factor * tdoubles * (ysingles + ydoubles) / D
 * task_energies (2 * ntasks doubles, or NULL) receives {energy1, energy2}
 * of each task in loop order.  Synchronous.  Device memory: the orbital
 * pairing (same tiling on both jobs, not TCE_T_TILING_ORBITAL) runs two
 * passes of the contraction kernel over the jobs' kernel tasks with a
 * window of at most 16 GiB (a quarter of the free memory) allocated for
 * the call; the reference order (and ORBITAL-tiled jobs) keep each job's
 * tile buffer, 16 bytes per element of the largest task tile, as
 * tce_t_job_triples_tiles does. */
int tce_t_lambda_ccsd_t(tce_t_job *right, tce_t_job *left, const double *f1,
                        const long *k_f1_offset, int pairing, int precision, double *energy1,
                        double *energy2, double *task_energies, long ntasks);

/* ---- CR-CCSD(T) (cr_ccsd_t.F) -------------------------------------------- */
/* A CR-CCSD(T) job: d_t1/k_t1_offset and d_t2/k_t2_offset as for
 * tce_t_job_upload, and the two intermediates cr_ccsd_t_N builds before its
 * task loop (toggle 1, cr_ccsd_t.F:62-66): d_i1_1/k_i1_offset_1, i1(h11 p4
 * h1 h2) under the table of OFFSET_cr_ccsd_t_N_1_1 (blocks [p4][h11][h1][h2],
 * h1 tile <= h2 tile), and d_i1_2/k_i1_offset_2, i1(p4 p5 h1 p12) under
 * OFFSET_cr_ccsd_t_N_2_1 (blocks [p4][p5][h1][p12], p4 tile <= p5 tile).  The
 * job's t3 doubles are then the moment tile of cr_ccsd_t_N (toggle 2:
 * -P(9) t(p4 p5 h1 h11) i1(h11 p6 h2 h3) - P(9) t(p4 p12 h1 h2) i1(p5 p6 h3
 * p12)) and its singles the denominator tile of cr_ccsd_t_E (P(9) t(p4 p5
 * h1 h2) t(p6 h3) - 2/3 P(9) t(p4 h1) i1(p5 p6 h2 h3), i1 = -1/4 P(4)
 * t(p4 h1) t(p5 h2)).  Always packed on the device.  TCE_T_ERR_KEY if a
 * table lacks a block the spin and irrep rules require. */
int tce_t_job_upload_cr(tce_t_job *job, const double *t1, const long *k_t1_offset,
                        const double *t2, const long *k_t2_offset, const double *i1_1,
                        const long *k_i1_offset_1, const double *i1_2,
                        const long *k_i1_offset_2);
/* Replaces cr_ccsd_t(..., energy1, energy2) (cr_ccsd_t.F:1-269) after its
 * intermediates are built: triples holds t1/t2/v2 (tce_t_job_upload), cr
 * the CR job of the same tiles and tiling (not TCE_T_TILING_ORBITAL).  The
 * d1/d2 contractions of both passes run in `precision` (TCE_T_FP64 as the
 * reference), the sums in FP64; over the tasks of the (T) loop with
 * w = factor / D:
 *   num1 = sum w * moment * tdoubles,  num2 = sum w * moment * (tsingles + tdoubles)
 *   den1 = sum w * den * tdoubles,     den2 = sum w * den * (tsingles + tdoubles)
 * den0 = the overlap of cr_ccsd_t_D with c = t, and
 *   energy1 = num1 / (1 + den1 + den0), energy2 = num2 / (1 + den2 + den0).
 * terms (5 doubles or NULL) receives num1, num2, den0, den1, den2.
 * Synchronous; two passes of the contraction kernel with the window of
 * tce_t_lambda_ccsd_t. */
int tce_t_cr_ccsd_t(tce_t_job *triples, tce_t_job *cr, int precision, double *energy1,
                    double *energy2, double *terms);
/* Host packer of tce_t_job_upload, for inspection and tests: the dense tensors
 * the job places in HBM, built from the same blocks and hash tables.  With
 * Ep = no. of virtuals and Mp = no. of occupieds rounded up to a multiple of 8:
 *   T1[nv][no], T2e[no][no][nv][Ep] = t(p e h h'), T2m[no][nv][nv][Mp] =
 *   t(p p' h m), Wd1[no][no][nv][Mp] = <m p||h h'>, Vd2[no][nv][nv][Ep] =
 *   <p p'||w e>, S[no][no][nv][nv] = <p p'||h h'>.  Any output may be NULL. */
int tce_t_pack_dense(const tce_t_tiles *tiles, const double *t1, const long *k_t1_offset,
                     const double *t2, const long *k_t2_offset, const double *v2,
                     const long *k_v2_offset, double *T1, double *T2e, double *T2m,
                     double *Wd1, double *Vd2, double *S);

/* Inspection (no reference counterpart): the job's reduced-precision copy
 * of one contraction tensor, as the kernels read it.  tensor: 0 = T2e,
 * 1 = Vd2 (d2 operands), 2 = Wd1, 3 = T2m (d1 operands), laid out as
 * tce_t_pack_dense's tensors with the contraction extent zero-padded to
 * *kout (a multiple of the kernel's chunk).  FP32: out0 gets *rows x *kout
 * floats (out1, out2 must be NULL); BF16X3: out0 the bf16 hi and out1 the
 * bf16 lo parts (*rows x *kout 16-bit patterns each; out2 must be NULL);
 * BF16X6: hi, mid, lo in out0..2.  Any output may be NULL; all NULL: only
 * *rows and *kout are set (no upload needed).  Converts the tensors first if
 * no run in `precision` has yet.  Synchronous. */
int tce_t_job_operands(tce_t_job *job, int precision, int tensor, void *out0, void *out1,
                       void *out2, long *rows, long *kout);

/* Device time (ms) of the last run's contraction kernel(s), HIP events. */
double tce_t_job_last_kernel_ms(const tce_t_job *job);
/* Which FP64 contraction kernel the job runs: TCE_T_KERNEL_WAVE (one
 * wavefront per 4x4x4 particle cube, operands from L1/L2), TCE_T_KERNEL_WG
 * (8-wavefront workgroup per 8x8x8 particle box, operands staged in LDS,
 * 4x4x4 MFMA) or TCE_T_KERNEL_WG16 (the same boxes on 16x16x4 MFMA); chosen
 * at tce_t_job_create from the tile shapes.  Negative on a NULL job. */
#define TCE_T_KERNEL_WAVE 0
#define TCE_T_KERNEL_WG 1
#define TCE_T_KERNEL_WG16 2
int tce_t_job_kernel(const tce_t_job *job);
/* Work a run in `precision` performs: *flops = sum over the kernel tasks of
 * |t3 tile| * 18 * (no + nv + 1) (the reference's per-task count of
 * ccsd_t_kernels_omp.F: 9 d1 calls x 2 no, 9 d2 x 2 nv, 9 s1 x 2) and
 * *items = work items (either may be NULL).  By default the kernel tasks are
 * the box-sized sub-tiles of the reference tiles (4 occupied, 4 or 8
 * virtual orbitals) in canonical order, which leaves out the symmetric
 * duplicates a reference task computes inside equal tiles: the energies
 * are the same sums (per reference task, up to FP64 rounding) at less work
 * (see tce_t_job_create_tiling).  No reference counterpart. */
int tce_t_job_work(const tce_t_job *job, int precision, double *flops, long *items);
int tce_t_job_destroy(tce_t_job *job);

/* ---- batches: many uploaded jobs in one persistent launch -------------- */
/* No counterpart in the reference (its task farm runs one job per call).
 * The jobs stay owned by the caller and must outlive the batch; a job
 * uploaded again after tce_t_batch_create makes tce_t_batch_run return
 * TCE_T_ERR_STATE (the batch's plans hold the old device buffers). */
typedef struct tce_t_batch tce_t_batch;
int tce_t_batch_create(tce_t_job *const *jobs, int njobs, tce_t_batch **batch);
/* Run every job of the batch in `precision`: energies[2j] = E[T],
 * energies[2j+1] = E(T) of job j (tce_t_job_task_energies of each job is
 * updated too).  The first run in a precision autotunes the launch (kernel
 * family, segment order) and captures it in a hipGraph; later runs replay
 * the graph.  Synchronous. */
int tce_t_batch_run(tce_t_batch *batch, int precision, double *energies);
/* Device time (ms) of the last run (graph launch to completion). */
double tce_t_batch_last_ms(const tce_t_batch *batch);
/* The plan chosen for `precision`: family 0 = fused_triples (4-wide boxes),
 * 1 = fused_triples_wg, 2 = fused_triples_mixed (reduced precision), 3 =
 * fused_triples_mixed on FP64 (16x16x4); cost_order 1 = costliest work
 * items first; tuned_ms[6] = candidate times of the tuning runs (family x
 * order in trial order, -1 where not tried, all -1 when the plan came from
 * the process-wide cache).  The tuning runs cover the first eighth of every
 * job's work items when the batch has at least 64 per CU (all items
 * otherwise; TCE_T_BATCH_TUNE_FRACTION overrides), so their times are of
 * that sample. */
int tce_t_batch_plan(const tce_t_batch *batch, int precision, int *family, int *cost_order,
                     double *tuned_ms);
/* The work-queue slice width the plan chose for fused_triples_mixed (the
 * consecutive items each XCD takes at a time) and slice_ms[3], the tuning
 * runs at widths 32, 64, 128 (-1 where not tried: other kernel families, or
 * a plan from the process-wide cache).  No reference counterpart. */
int tce_t_batch_plan_schedule(const tce_t_batch *batch, int precision, int *slice,
                              double *slice_ms);
int tce_t_batch_destroy(tce_t_batch *batch);

/* ---- serving: continuous batching of (T) requests on one GPU ---------- */
/* No counterpart in the reference.  A worker thread owns the device.
 * Systems are registered once (host blocks owned by the caller, valid until
 * the server is destroyed); the first request for a system uploads it and it
 * stays resident until memory is needed (idle systems are evicted LRU).
 * Requests are admitted first come first served when a slot (max_active) is
 * free and free HBM minus reserve_bytes covers what they add; every
 * iteration launches, per precision, one persistent kernel over the next
 * work items of all active requests (a launch lasts about slice_ms), so
 * requests arriving while others run share the device at the next slice.
 * A request's energies do not depend on what it shared the device with. */
typedef struct tce_t_server tce_t_server;
int tce_t_server_create(int device, double slice_ms, int max_active, long reserve_bytes,
                        tce_t_server **server);
/* Returns the system id (>= 0) or a negative status. */
int tce_t_server_add_system(tce_t_server *server, const tce_t_tiles *tiles, const double *t1,
                            const long *k_t1_offset, const double *t2, const long *k_t2_offset,
                            const double *v2, const long *k_v2_offset);
/* Queue a request; returns its ticket (>= 0) or a negative status. */
long tce_t_server_submit(tce_t_server *server, int system, int precision);
/* Wait for a ticket (timeout_s < 0: forever; TCE_T_ERR_STATE on timeout).
 * times[4] = submit, admit, first slice, finish (seconds since creation).
 * Returns the request's status. */
int tce_t_server_wait(tce_t_server *server, long ticket, double timeout_s, double *e_bt,
                      double *e_pt, double *times);
/* Slices launched so far and systems resident now. */
int tce_t_server_stats(tce_t_server *server, long *slices, int *resident);
/* Finish every queued and active request, then free everything. */
int tce_t_server_destroy(tce_t_server *server);

/* ================================================================ 3 ===== */
/* tce_tile.F:330-355 -- tiles of n orbitals, sizes differing by <= 1. */
long tce_t_tile_ranges(long n, long tilesize, long *ranges);
/* tce_t1_offset.F, tce_t2_offset.F, tce_mo2e_offset.F (restricted to the
 * three integral classes the (T) path reads).  hash may be NULL to query L.
 * Returns L; *size = doubles in the block array. */
long tce_t_t1_offset(const tce_t_tiles *tiles, long *hash, long *size);
long tce_t_t2_offset(const tce_t_tiles *tiles, long *hash, long *size);
long tce_t_v2_offset(const tce_t_tiles *tiles, long *hash, long *size);
/* tce_y1_offset.F, tce_y2_offset.F (Lambda amplitudes, irrep 0) and
 * tce_mo1e_offset.F (one-electron operator, irrep 0), same conventions. */
long tce_t_y1_offset(const tce_t_tiles *tiles, long *hash, long *size);
long tce_t_y2_offset(const tce_t_tiles *tiles, long *hash, long *size);
long tce_t_f1_offset(const tce_t_tiles *tiles, long *hash, long *size);
/* tce_hash.F:271-316 -- offset of key, or -1. */
long tce_t_hash(const long *hash, long key);

/* Orbital-form two-electron integrals (intorb: a closed-shell run whose
 * integrals are stored over spatial orbitals, get_hash_block_i).  The alpha
 * tile list and the map of the spin-orbital tiles onto it, as
 * tce_tile.F:1156-1212,1370-1380 set them (k_b2am, k_spin_alpha,
 * k_sym_alpha, k_range_alpha), and the irrep of V (irrep_v). */
typedef struct {
  long noa, nva;             /* alpha occupied and virtual tiles            */
  const long *b2am;          /* [noab+nvab]: alpha tile (1-based) of each    */
                             /* spin-orbital tile                           */
  const long *spin_alpha;    /* [noa+nva] (1 in a closed-shell run)         */
  const long *sym_alpha;     /* [noa+nva]                                   */
  const long *range_alpha;   /* [noa+nva]                                   */
  long irrep_v;
} tce_t_intorb;
/* tce_mo2e_offset_intorb.F:1-110 -- the orbital-form block table
 * k_v2_alpha_offset: hash[0] = segments L1, then per entry (L1+1 of them,
 * the first block, every (blocks/idiv2e)-th block and the last) its key,
 * offset and tiles g3, g4, g1, g2 (six columns).  hash may be NULL to query
 * the length.  Returns the table's length in longs (-1: invalid, or fewer
 * blocks than idiv2e); *size = doubles of the orbital-form array. */
long tce_t_v2_alpha_offset(const tce_t_intorb *io, long idiv2e, long *hash, long *size);
/* get_hash_block_i / get_block_ind_i (get_block_ind.F:818-1470, ioalg 2)
 * over every block of v2_hash (tce_t_v2_offset's table, what
 * tce_t_job_upload reads): the antisymmetrised spin-orbital blocks
 * v(g3 g4 g1 g2) = (g3 g1|g4 g2) - (g3 g2|g4 g1), each term kept where its
 * spins pair up, from the orbital-form integrals v2orb (the layout
 * tce_mo2e_trans.F writes).  v2_alpha_offset (the caller's
 * k_v2_alpha_offset, may be NULL) is checked against the alpha tiles.
 * Host threads; TCE_T_ERR_KEY if a source block is not stored. */
int tce_t_v2_from_intorb(const tce_t_tiles *tiles, const tce_t_intorb *io, const double *v2orb,
                         const long *v2_alpha_offset, const long *v2_hash, double *v2);

/* Seeded synthetic C1 spin-orbital inputs (Philox4x32-10 counters keyed by
 * global orbital indices, Box-Muller normals):
 *   t1 ~ N(0,0.01), t2 ~ N(0,0.02) antisymmetrised in both pairs,
 *   <h p||h h>, <p p||h p>, <p p||h h> ~ N(0,0.05) antisymmetrised in their
 *   same-space pairs, occupied energies sorted U[-2,-0.3], virtual sorted
 *   U[0.2,4.0].  Values depend only on (seed, global indices), never on the
 *   tiling. */
int tce_t_synth_evl(unsigned long long seed, long no, long nv, double *evl);
int tce_t_synth_blocks(const tce_t_tiles *tiles, unsigned long long seed,
                       double *t1, const long *t1_hash,
                       double *t2, const long *t2_hash,
                       double *v2, const long *v2_hash);
/* Dense forms (tests): T1[nv][no], T2[nv][nv][no][no], W[no][nv][no][no]
 * = <m p||h h'>, V[nv][nv][no][nv] = <p p'||h e>, S[nv][nv][no][no]. */
int tce_t_synth_dense(unsigned long long seed, long no, long nv, double *T1,
                      double *T2, double *W, double *V, double *S);
/* Back end of the per-call d1/d2 contractions (groups 1 and the CPU
 * surface): DIRECT, one thread per t3 element with the contraction loop
 * (default), or TTGT, the alternative of sd_t_total_ttlg.cu: one DGEMM on
 * the FP64 matrix cores (rocBLAS) into a scratch matrix, then a transpose
 * kernel adding it into the tile in the variant's layout.  Same results up
 * to FP64 summation order.  TCE_T_LEGACY=ttgt in the environment selects
 * TTGT at the first call.  Returns the back end in force (QUERY: only that),
 * or TCE_T_ERR_ARG. */
#define TCE_T_LEGACY_DIRECT 0
#define TCE_T_LEGACY_TTGT 1
#define TCE_T_LEGACY_QUERY (-1)
int tce_t_legacy_backend(int backend);
/* Sticky failure flag of the per-call surface (group 1), which has no return
 * codes: TCE_T_ERR_HIP if any call failed since the last query, then reset. */
int tce_t_legacy_status(void);

/* Philox4x32-10 block function (known-answer tests). */
void tce_t_philox4x32(const unsigned int ctr[4], const unsigned int key[2],
                      unsigned int out[4]);

#ifdef __cplusplus
}
#endif
#endif
