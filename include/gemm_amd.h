/*
 * gemm_amd.h -- C ABI of libgemm_amd.so, an MI355X (gfx950) BLAS-3 library.
 *
 * Part 1 is the drop-in: the Fortran-convention, column-major entry points
 * of the reference (wjc404/GEMM_AVX512F) with the same argument meaning.
 * They take HOST pointers; the library copies operands to HBM, runs the HIP
 * kernels on one GPU and copies C back before returning.
 *
 * Part 2 is the device-pointer, stream-ordered variant of every entry point
 * (arguments by value, pointers into device memory, work enqueued on
 * `stream`, NULL = the null stream).  They return 0 on success, a positive
 * netlib-style parameter index when an argument is illegal (the same index
 * xerbla would print), or a negative HIP error code.
 *
 * Part 3 is runtime plumbing (device memory, copies, events) so that a C or
 * Python caller needs nothing but this library.
 *
 * Integer arguments are 32-bit (LP64), as in the reference.  Complex values
 * are interleaved (re, im) pairs; complex leading dimensions count complex
 * elements.
 */
#ifndef GEMM_AMD_H
#define GEMM_AMD_H
#include <stddef.h>
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

/* ===================================================== Part 1: drop-in (host pointers) */

/* C <- alpha*op(A)*op(B) + beta*C, op(X) = X for 'N'/'n', X^T otherwise.
 * Replaces sgemm_ of SGEMM.so (src/gemm_driver.c:116, float build; Makefile:14-15).
 * beta == 0 overwrites C without reading it (netlib); the reference multiplies
 * by zero instead (src/gemm_driver.c:110) -- identical for finite C. */
void sgemm_(const char *transa, const char *transb, const int *m, const int *n, const int *k,
            const float *alpha, const float *a, const int *lda, const float *b, const int *ldb,
            const float *beta, float *c, const int *ldc);

/* Replaces dgemm_ of DGEMM.so (src/gemm_driver.c:116 with -DDOUBLE; Makefile:11-12). */
void dgemm_(const char *transa, const char *transb, const int *m, const int *n, const int *k,
            const double *alpha, const double *a, const int *lda, const double *b, const int *ldb,
            const double *beta, double *c, const int *ldc);

/* B <- alpha * op(A)^-1 * B (side 'L') or alpha * B * op(A)^-1 (side 'R'),
 * netlib dtrsm_ contract.  The reference ships only the kernel-level solvers
 * for this path (dtrsm_solve_*: OpenBLAS-like_implementation/dtrsm_16x4/
 * dtrsm_kernel_16x4_skylakex.S:1882-2536, driven by OpenBLAS's trsm driver);
 * this is the BLAS-level routine those kernels serve.  Only the uplo triangle
 * of A is read; diag 'U' never reads the diagonal. */
void dtrsm_(const char *side, const char *uplo, const char *transa, const char *diag, const int *m, const int *n,
            const double *alpha, const double *a, const int *lda, double *b, const int *ldb);

/* Complex C <- alpha*op(A)*op(B) + beta*C by the 3M method (three real
 * products), op in {N, T, C}; alpha and beta are (re, im) pairs.  Argument
 * list of the OpenBLAS extensions cgemm3m_/zgemm3m_ (= that of cgemm_/zgemm_).
 * The reference ships the 3M real-product kernels for this path
 * (cgemm3m_kernel_16x4_skylakex.c:416, zgemm3m_kernel_2x8_skylakex.c:337,
 * zgemm3m_kernel_8x4_skylakex.c:310) whose driver lives in OpenBLAS. */
void cgemm3m_(const char *transa, const char *transb, const int *m, const int *n, const int *k,
              const float *alpha, const float *a, const int *lda, const float *b, const int *ldb,
              const float *beta, float *c, const int *ldc);
void zgemm3m_(const char *transa, const char *transb, const int *m, const int *n, const int *k,
              const double *alpha, const double *a, const int *lda, const double *b, const int *ldb,
              const double *beta, double *c, const int *ldc);

/* P*A = L*U with partial pivoting, LAPACK dgetrf_ contract: column-major m x n
 * A overwritten by L (unit lower, below the diagonal) and U; ipiv[i] (1-based,
 * min(m,n) entries) = row interchanged with row i; info = 0, -i for an illegal
 * argument i, or i > 0 when U(i,i) is exactly zero (the factorisation is still
 * completed).  No counterpart in the reference, which has no LAPACK layer: it
 * is the first caller of the DTRSM and DGEMM paths above, which carry all but
 * the panel columns.  Any m: panels of up to 16384 rows are factored in
 * registers, taller ones in global memory.  A failure of the GPU runtime
 * (initialisation, memory, a launch
 * or a copy) gives info = GEMM_AMD_INFO_RUNTIME_ERROR, and a and ipiv are then
 * not a factorisation. */
#define GEMM_AMD_INFO_RUNTIME_ERROR (-1000)
void dgetrf_(const int *m, const int *n, double *a, const int *lda, int *ipiv, int *info);

/* ----------------------------------------------------- Part 1b: CBLAS and ILP64
 * Wrappers over the routines above with no counterpart in the reference,
 * which exports only the Fortran LP64 symbols (src/gemm_driver.c:116): the
 * CBLAS interface for C callers (row-major or column-major; the enum values
 * of cblas.h: RowMajor 101, ColMajor 102, NoTrans 111, Trans 112, ConjTrans
 * 113, Upper 121, Lower 122, NonUnit 131, Unit 132, Left 141, Right 142;
 * illegal arguments are reported with CBLAS numbering) and the ILP64 Fortran
 * symbols (64-bit integers, suffix 64_, as OpenBLAS names them), which reject
 * values outside the 32-bit range the kernels index with.  Host pointers. */
void cblas_sgemm(int layout, int transa, int transb, int m, int n, int k, float alpha, const float *a, int lda,
                 const float *b, int ldb, float beta, float *c, int ldc);
void cblas_dgemm(int layout, int transa, int transb, int m, int n, int k, double alpha, const double *a, int lda,
                 const double *b, int ldb, double beta, double *c, int ldc);
void cblas_cgemm3m(int layout, int transa, int transb, int m, int n, int k, const void *alpha, const void *a,
                   int lda, const void *b, int ldb, const void *beta, void *c, int ldc);
void cblas_zgemm3m(int layout, int transa, int transb, int m, int n, int k, const void *alpha, const void *a,
                   int lda, const void *b, int ldb, const void *beta, void *c, int ldc);
void cblas_dtrsm(int layout, int side, int uplo, int transa, int diag, int m, int n, double alpha, const double *a,
                 int lda, double *b, int ldb);
void sgemm_64_(const char *transa, const char *transb, const int64_t *m, const int64_t *n, const int64_t *k,
               const float *alpha, const float *a, const int64_t *lda, const float *b, const int64_t *ldb,
               const float *beta, float *c, const int64_t *ldc);
void dgemm_64_(const char *transa, const char *transb, const int64_t *m, const int64_t *n, const int64_t *k,
               const double *alpha, const double *a, const int64_t *lda, const double *b, const int64_t *ldb,
               const double *beta, double *c, const int64_t *ldc);
void cgemm3m_64_(const char *transa, const char *transb, const int64_t *m, const int64_t *n, const int64_t *k,
                 const float *alpha, const float *a, const int64_t *lda, const float *b, const int64_t *ldb,
                 const float *beta, float *c, const int64_t *ldc);
void zgemm3m_64_(const char *transa, const char *transb, const int64_t *m, const int64_t *n, const int64_t *k,
                 const double *alpha, const double *a, const int64_t *lda, const double *b, const int64_t *ldb,
                 const double *beta, double *c, const int64_t *ldc);
void dtrsm_64_(const char *side, const char *uplo, const char *transa, const char *diag, const int64_t *m,
               const int64_t *n, const double *alpha, const double *a, const int64_t *lda, double *b,
               const int64_t *ldb);
void dgetrf_64_(const int64_t *m, const int64_t *n, double *a, const int64_t *lda, int64_t *ipiv, int64_t *info);

/* ===================================================== Part 2: device pointers */

int gemm_amd_cgemm3m(char transa, char transb, int m, int n, int k, const float *alpha,
                     const float *a, int lda, const float *b, int ldb, const float *beta,
                     float *c, int ldc, void *stream);
int gemm_amd_zgemm3m(char transa, char transb, int m, int n, int k, const double *alpha,
                     const double *a, int lda, const double *b, int ldb, const double *beta,
                     double *c, int ldc, void *stream);

int gemm_amd_dtrsm(char side, char uplo, char transa, char diag, int m, int n, double alpha,
                   const double *a, int lda, double *b, int ldb, void *stream);

/* gemm_amd_dtrsm that allocates nothing: no counterpart in the reference, for
 * callers that cannot allocate device memory (code under graph capture, tight
 * serving loops, harnesses that forbid it).  Same arguments, result and
 * kernels as gemm_amd_dtrsm; its scratch (inverted diagonal blocks, split-K
 * partials of the trailing updates) is a fixed 64 MiB device array per GPU,
 * held until the process ends.  The triangle's order (m for side 'L', n for
 * 'R') must be at most GEMM_AMD_DTRSM_NOALLOC_MAX_NA, otherwise nothing runs
 * and -hipErrorInvalidValue (-1) is returned with a message.  Calls share
 * their GPU's array: calls that overlap in time on one GPU race and must be
 * ordered by the caller. */
#define GEMM_AMD_DTRSM_NOALLOC_MAX_NA 32768
int gemm_amd_dtrsm_noalloc(char side, char uplo, char transa, char diag, int m, int n, double alpha,
                           const double *a, int lda, double *b, int ldb, void *stream);

/* Kernel-level solve with the reference's packed-operand contract, on device
 * pointers: replaces dtrsm_solve_{ln,lt,rn,rt}_m{M}n{N}(sa, sb, C, LDB, LDC)
 * (dtrsm_kernel_16x4_skylakex.S:1882-2536; semantics as its tester's
 * solve_*_reference, dtrsm_kernel_16x4_tester.c:97-199), for any M, N.
 * variant 0 = ln (backward), 1 = lt (forward), 2 = rn, 3 = rt.  The triangle's
 * diagonal holds inverse pivots; the solution is written to C and back into
 * sb (L) or sa (R). */
int gemm_amd_dtrsm_kernel(int variant, long M, long N, double *sa, double *sb, double *c, long LDB, long LDC,
                          void *stream);

/* Device-pointer dgetrf_: a, ipiv and info in device memory, stream-ordered.
 * Returns 0, the index of an illegal argument (m, n, -, lda = 1, 2, -, 4), or a
 * negative HIP error. */
int gemm_amd_dgetrf(int m, int n, double *a, int lda, int *ipiv, int *info, void *stream);

int gemm_amd_sgemm(char transa, char transb, int m, int n, int k, float alpha,
                   const float *a, int lda, const float *b, int ldb, float beta,
                   float *c, int ldc, void *stream);
int gemm_amd_dgemm(char transa, char transb, int m, int n, int k, double alpha,
                   const double *a, int lda, const double *b, int ldb, double beta,
                   double *c, int ldc, void *stream);

/* Same products with an explicit kernel configuration (tile id 0..3 =
 * 128x128, 64x64, 128x64, and 256x64 (FP64) or 256x128 (FP32); splitk >= 1), bypassing the picker.
 * Used by the autotuner; the result is identical up to summation order.  These always run the one-group
 * kernel of the named tile (the calls above may split each slab's k over groups of waves on grids of at most
 * one workgroup per CU). */
int gemm_amd_sgemm_cfg(char transa, char transb, int m, int n, int k, float alpha,
                       const float *a, int lda, const float *b, int ldb, float beta,
                       float *c, int ldc, int tile, int splitk, void *stream);
int gemm_amd_dgemm_cfg(char transa, char transb, int m, int n, int k, double alpha,
                       const double *a, int lda, const double *b, int ldb, double beta,
                       double *c, int ldc, int tile, int splitk, void *stream);

/* Emulated DGEMM on the int8 matrix cores (no reference counterpart: the
 * reference's DGEMM is FP64 FMA, src/kernel/gemm_kernel.S:11-23).  Same
 * contract as gemm_amd_dgemm; each row of op(A) / column of op(B) is scaled by
 * a power of two and cut into nslices (1..8) round-to-nearest int8 slices of
 * 7 bits, the nslices*(nslices+1)/2 leading slice products run as int8 GEMMs
 * with exact int32 accumulation, and FP64 recombines them (Ozaki-Ogita-Oishi
 * splitting; int8 form after Ootomo-Ozaki-Yokota 2024).  Accuracy ~2^(-7*nslices)
 * relative to the row/column maxima; nslices = 8 matches or beats the native
 * FP64 kernel's error on uniform inputs.  nslices = 0 takes the count the
 * emulation autotuner found for this K (gemm_amd_emulation_counts). */
int gemm_amd_dgemm_ozaki(char transa, char transb, int m, int n, int k, double alpha,
                         const double *a, int lda, const double *b, int ldb, double beta,
                         double *c, int ldc, int nslices, void *stream);

/* The split step of gemm_amd_dgemm_ozaki alone (no reference counterpart;
 * exposed so the byte-exact slicing can be checked): rows of the rows x k view
 * of x (element (r, p) at x[r*row_stride + p*k_stride]) are cut into
 * nslices int8 planes written to `slices` as [t][rows_pad][k_pad] (zero
 * padding; the buffer holds nslices*rows_pad*k_pad bytes) and their exponents
 * to exps[rows_pad].  rows_pad >= rows and k_pad >= k must be multiples of 64. */
int gemm_amd_ozaki_split(const double *x, int64_t row_stride, int64_t k_stride, int rows, int k,
                         int nslices, int8_t *slices, int rows_pad, int k_pad, int *exps, void *stream);

/* Emulated DGEMM on the int8 matrix cores, modular form (no reference
 * counterpart; Ozaki-Uchino-Imamura "Ozaki scheme II", 2025).  Same contract
 * as gemm_amd_dgemm.  Rows of op(A) / columns of op(B) are scaled by powers of
 * two to integers of `bits` = gemm_amd_ozaki_crt_bits(nmoduli, k) bits, reduced
 * modulo nmoduli (2..16) pairwise coprime moduli <= 256 (256, 255, 253, 251,
 * 247, 241, 239, 233, 229, 227, 223, 217, 211, 199, 197, 193), multiplied as
 * one exact int8 GEMM per modulus, and rebuilt by the Chinese remainder
 * theorem in exact FP64/int64 limb arithmetic.  nmoduli int8 GEMMs in all;
 * 16 keeps 55 bits per operand at K = 8192.  nmoduli = 0 takes the count the
 * emulation autotuner found for this K (gemm_amd_emulation_counts). */
int gemm_amd_dgemm_ozaki_crt(char transa, char transb, int m, int n, int k, double alpha,
                             const double *a, int lda, const double *b, int ldb, double beta,
                             double *c, int ldc, int nmoduli, void *stream);
/* Bits kept per row/column: the largest b with k*2^(2b) <= M/4 (M = product of
 * the moduli), at most 56; -1 for invalid arguments. */
/* Moduli (CRT form) and slices (slice form) used when the emulated DGEMMs
 * are called with 0 for that count: per ceil(log2 K), the smallest counts
 * whose max error the emulation autotuner (gemm_avx512f_amd.tuning,
 * tune-emulation) measured at or below the native FP64 kernel's on inputs of
 * dynamic range 2^0, 2^8 and 2^32 (no reference counterpart: the reference
 * has no emulated product).  _set installs a table (33 entries each, index
 * ceil(log2 K)); the library starts from the one compiled in. */
int gemm_amd_emulation_counts(int k, int *nmoduli, int *nslices);
int gemm_amd_emulation_counts_set(const int *moduli, const int *slices, int n);

int gemm_amd_ozaki_crt_bits(int nmoduli, int k);
/* The residue step alone (exposed for byte-exact checks): rows of the view of
 * x become integers rint(x * 2^(bits - e_r)) (e_r from frexp of the row's max),
 * whose symmetric residues modulo the first nmoduli moduli are written as
 * [l][rows_pad][k_pad] int8 planes; exps[r] = e_r - bits. */
int gemm_amd_ozaki_crt_split(const double *x, int64_t row_stride, int64_t k_stride, int rows, int k,
                             int nmoduli, int bits, int8_t *residues, int rows_pad, int k_pad, int *exps,
                             void *stream);

/* SGEMM on the bf16 matrix cores from exact bf16 splits of the FP32 operands
 * (no reference counterpart: the reference's SGEMM is FP32 FMA,
 * src/kernel/gemm_kernel.S:11-23).  Same contract as gemm_amd_sgemm; nprod = 3
 * ("bf16x3": x = x0 + x1, products x0y0 + x0y1 + x1y0, ~16 bits) or 6 ("bf16x6":
 * 3-way split, all products of weight >= 2^-16, FP32 level), FP32
 * accumulation (Henry, Tang, Heinecke, ARITH 2019). */
int gemm_amd_sgemm_bf16x(char transa, char transb, int m, int n, int k, float alpha,
                         const float *a, int lda, const float *b, int ldb, float beta,
                         float *c, int ldc, int nprod, void *stream);

/* The split step of gemm_amd_sgemm_bf16x alone (no reference counterpart;
 * exposed so the exact splitting can be checked): rows of the rows x k view of
 * x are cut into nsplit bf16 parts (round to nearest even, each part taken from
 * the exact FP32 remainder), written as [part][rows_pad][k_pad] uint16 with
 * zero padding; rows_pad >= rows and k_pad >= k must be multiples of 64. */
int gemm_amd_bf16_split(const float *x, int64_t row_stride, int64_t k_stride, int rows, int k, int nsplit,
                        uint16_t *planes, int rows_pad, int k_pad, void *stream);

/* Grouped GEMM: `count` independent column-major problems
 * C_p <- alpha_p*op(A_p)*op(B_p) + beta_p*C_p of arbitrary shapes and
 * transposes, run as ONE persistent launch (no reference counterpart; the
 * reference runs one GEMM per call, src/gemm_driver.c:116).  All pointer
 * arrays live on the host and point to device memory.  dtype 0 = FP32,
 * 1 = bf16 operands (uint16 bit patterns) with FP32 C and accumulation,
 * 2 = FP64.  A plan holds the device problem table (built once); _run launches
 * it, _capture records that launch into a hipGraph on a non-null stream and
 * _replay relaunches the graph with no host-side work.  Problems are
 * scheduled largest first; results do not depend on the schedule.
 * _create returns NULL on an illegal argument (message in
 * gemm_amd_last_error()); _destroy synchronises the device first.
 * Each problem runs on one of two tile classes in the same launch: large
 * tiles (FP32 and bf16 128x128, FP64 128x64) when M >= large_min and
 * N >= large_min (FP64: large_min / 2), small 64x64 tiles otherwise.
 * _create uses the default per dtype (bf16 96, FP64 128, FP32 no large
 * tiles); _create_ex takes large_min (0 = that default, a value above every
 * M keeps all problems on small tiles), which is what the grouped autotuner
 * (gemm_avx512f_amd.tuning.tune_grouped) sweeps. */
void *gemm_amd_grouped_create(int dtype, int count, const char *transa, const char *transb, const int *m,
                              const int *n, const int *k, const double *alpha, const void *const *a,
                              const int *lda, const void *const *b, const int *ldb, const double *beta,
                              void *const *c, const int *ldc);
void *gemm_amd_grouped_create_ex(int dtype, int count, const char *transa, const char *transb, const int *m,
                                 const int *n, const int *k, const double *alpha, const void *const *a,
                                 const int *lda, const void *const *b, const int *ldb, const double *beta,
                                 void *const *c, const int *ldc, int large_min);
int gemm_amd_grouped_run(void *plan, void *stream);
int gemm_amd_grouped_capture(void *plan, void *stream);
int gemm_amd_grouped_replay(void *plan, void *stream);
int64_t gemm_amd_grouped_tiles(void *plan);
/* Before capture (no reference counterpart; the reference's closest is the
 * CAS tile queue of OpenBLAS-like_implementation/dgemm.c:464-482): dynamic = 1
 * makes the workgroups claim tiles from an atomic counter (reset by a small
 * kernel ahead of each launch, graph-capturable) after their first one,
 * dynamic = 0 gives the static stride (workgroup w takes tiles w, w + grid,
 * ...); both compute the same bits.  A new plan is dynamic when it has more
 * than two tiles per workgroup.  timeline (device memory, 3 x gemm_amd_grouped_grid(plan) uint64, or
 * NULL) receives per workgroup its start and end on the 100 MHz real-time
 * counter and its tile count, on every later launch.  One stream at a time
 * per plan. */
int gemm_amd_grouped_set_schedule(void *plan, int dynamic, unsigned long long *timeline);
int gemm_amd_grouped_grid(void *plan);   /* workgroups of the persistent launch */
int gemm_amd_grouped_destroy(void *plan);

/* A batch of independent complex-double 3M products (zgemm3m_ contract per
 * problem; alpha/beta hold 2*count doubles): the three real products of every
 * problem run together in one grouped FP64 launch.  Returns when the batch has
 * run.  The library keeps the plans of the last 4 distinct argument lists, so
 * repeating a batch with the same arguments (shapes, pointers, scalars)
 * builds no new plan. */
int gemm_amd_zgemm3m_batched(int count, const char *transa, const char *transb, const int *m, const int *n,
                             const int *k, const double *alpha, const double *const *a, const int *lda,
                             const double *const *b, const int *ldb, const double *beta, double *const *c,
                             const int *ldc, void *stream);

/* ============================================ Part 2b: autotuner and picker
 * Replaces the reference's compile-time tuning macros (sgemm_tune.h:6-15,
 * dgemm_tune.h:6-15, src/gemm_set_parameters.h:66-82) with run-time
 * configuration choice.  op: 0 = SGEMM, 1 = DGEMM, 2 = CGEMM3M.  The 20
 * candidates are tile shape (0..3) x split-K (1, 2, 4, 8, 16), index
 * tile*5 + log2(splitk).  lda_pad = lda - rows of A's storage. */
int gemm_amd_tune_candidates(int *tile, int *splitk);   /* returns 20 */
/* Times every valid candidate on library-owned seeded inputs (alpha 1, beta 0):
 * ms[i] for candidate i (-1 when invalid), ms[20] for the built-in heuristic's
 * own choice.  Best of 2-3 batches after a warm-up run. */
int gemm_amd_tune_sweep(int op, char transa, char transb, int m, int n, int k, int lda_pad, float *ms);
/* mode 0 = heuristic, 1 = learned picker, 2 = the library's own order
 * (environment override, sweep table, picker when enabled, heuristic). */
int gemm_amd_pick_config(int op, char transa, char transb, int m, int n, int k, int lda_pad, int mode, int *tile,
                         int *splitk);
/* The learned picker: a 34-64-64-20 ReLU perceptron (7700 FP32 parameters,
 * flat layout W1 b1 W2 b2 W3 b3, weights [in][out] row-major) run by a HIP
 * inference kernel.  Trained weights are compiled in; _load replaces them. */
int gemm_amd_picker_load(const float *params, int nparams);
int gemm_amd_picker_set_mode(int mode);   /* 0 heuristic, 1 learned picker */
int gemm_amd_picker_choose(int count, const int *op, const char *transa, const char *transb, const int *m,
                           const int *n, const int *k, const int *lda_pad, int *choice);
/* The network's 34 input features of each shape (row-major out[count][34]) for
 * a device of `cus` compute units, with ldb = rows of op(B) as in the sweep.
 * Host-only computation: needs no GPU. */
int gemm_amd_picker_features(int count, const int *op, const char *transa, const char *transb, const int *m,
                             const int *n, const int *k, const int *lda_pad, int cus, float *out);
/* Exact-shape results of the exhaustive sweep; they take precedence over the picker. */
int gemm_amd_tune_table_set(int count, const int *op, const char *transa, const char *transb, const int *m,
                            const int *n, const int *k, const int *lda_pad, const int *choice);
int gemm_amd_tune_table_clear(void);
/* Training kernels (device pointers, row-major [batch][34] inputs, [batch][20]
 * targets and masks).  loss = sum(mask*(y-t)^2)/(batch*20). */
int gemm_amd_mlp_forward(const float *params, const float *x, int batch, float *y, void *stream);
int gemm_amd_mlp_grad(const float *params, const float *x, const float *t, const float *mask, int batch, float *grad,
                      float *loss, void *stream);
int gemm_amd_mlp_adam(float *params, const float *grad, float *m, float *v, int n, float lr, float beta1, float beta2,
                      float eps, int step, void *stream);
/* Minibatch Adam over `epochs` shuffled passes (cosine step-size decay);
 * epoch_loss (host, `epochs` floats) receives each epoch's mean batch loss. */
int gemm_amd_mlp_train(float *params, float *adam_m, float *adam_v, const float *x, const float *t, const float *mask,
                       int nsamples, int batch, int epochs, float lr, uint64_t seed, float *epoch_loss);

/* ===================================================== Part 3: runtime */

int gemm_amd_init(void);                         /* 0 or negative HIP error; idempotent */
int gemm_amd_device_count(void);
/* fills: [0]=compute units, [1]=max clock kHz, [2]=LDS bytes per workgroup, [3]=L2 bytes */
int gemm_amd_device_info(int64_t *info4);
const char *gemm_amd_last_error(void);            /* message of the last failure on this thread */

void *gemm_amd_malloc(size_t bytes);              /* device memory, NULL on failure */
int gemm_amd_free(void *dptr);
/* kind: 1 host->device, 2 device->host, 3 device->device; synchronous */
int gemm_amd_memcpy(void *dst, const void *src, size_t bytes, int kind);
int gemm_amd_memset(void *dptr, int value, size_t bytes);   /* synchronous */
int gemm_amd_synchronize(void);

void *gemm_amd_stream_create(void);               /* default-priority stream */
int gemm_amd_stream_destroy(void *stream);
int gemm_amd_stream_synchronize(void *stream);
void *gemm_amd_event_create(void);
int gemm_amd_event_destroy(void *event);
int gemm_amd_event_record(void *event, void *stream);
int gemm_amd_event_elapsed_ms(void *start, void *stop, float *ms);   /* waits for stop */

/* Seeded synthetic input in device memory (no reference counterpart: the
 * benchmark and parity inputs).  Element (i, j) of the rows x cols column-major
 * matrix gets draw number first_draw + j*rows + i of the splitmix64 stream
 * seeded with `seed` (two draws per element for kind 2).  kind 0: FP64
 * uniform[-1,1) = ((x>>11)*2^-53)*2-1; kind 1: FP32 ((x>>40)*2^-24)*2-1;
 * kind 2: FP64 u*2^e, e uniform integer in [-range_r, range_r].  Padding rows
 * (i >= rows) are not written. */
int gemm_amd_fill_uniform(void *dptr, int kind, int64_t rows, int64_t cols, int64_t ld, uint64_t seed,
                          uint64_t first_draw, int range_r, void *stream);

#ifdef __cplusplus
}
#endif
#endif /* GEMM_AMD_H */
