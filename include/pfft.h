/*
 * pfft.h -- public C ABI of the MI355X-native PFFT transform pipeline
 * (library: pfft_amd/lib/libpfft.so).
 *
 * This header is the drop-in boundary for code written against the
 * reference's api/pfft.h on a 1x1 process mesh (one process, one GPU).
 * Names, argument meaning, flag values and error behaviour follow the
 * reference; each group below cites the reference declaration it
 * replaces.  Arrays passed to plans and execute calls may be device
 * memory (pfft_malloc/alloc_* return hipMalloc'ed memory) or host memory
 * (staged through device buffers; the PCIe-inclusive path).
 *
 * Two precisions are exported, as in the reference (api/pfft.h:562-564):
 *   pfft_   double        complex = double[2] (double _Complex after <complex.h>)
 *   pfftf_  float         complex = float[2]  (float _Complex after <complex.h>)
 * The long-double pfftl_ family has no GPU arithmetic and is not exported.
 */
#ifndef PFFT_H
#define PFFT_H 1

#include <stdarg.h>
#include <stddef.h>
#include <stdio.h>

/* MPI communicator handles.  The library never calls MPI: on a 1x1 mesh
 * every communicator has one rank.  If the caller included an MPICH-ABI
 * <mpi.h> first (MPI_VERSION defined), its MPI_Comm (an int handle) is
 * used; otherwise a minimal single-rank definition with the same ABI is
 * provided here. */
#ifndef MPI_VERSION
typedef int MPI_Comm;
#  define MPI_COMM_WORLD ((MPI_Comm)0x44000000)
#  define MPI_COMM_SELF  ((MPI_Comm)0x44000001)
#  define MPI_COMM_NULL  ((MPI_Comm)0x04000000)
#  define PFFT_SINGLE_RANK_MPI 1
#endif

#ifdef __cplusplus
extern "C" {
#endif

/* FFTW constants restated (FFTW 3.3 fftw3.h; reference api/pfft.h:44-59) */
#define PFFT_FORWARD   (-1)
#define PFFT_BACKWARD  (+1)
typedef enum {
  PFFT_R2HC = 0, PFFT_HC2R = 1, PFFT_DHT = 2,
  PFFT_REDFT00 = 3, PFFT_REDFT01 = 4, PFFT_REDFT10 = 5, PFFT_REDFT11 = 6,
  PFFT_RODFT00 = 7, PFFT_RODFT01 = 8, PFFT_RODFT10 = 9, PFFT_RODFT11 = 10
} pfft_r2r_kind;
typedef pfft_r2r_kind pfftf_r2r_kind;

/* Complex element types as FFTW 3 defines fftw_complex (reference
 * api/pfft.h:515-516 takes them from fftw3.h): the C99 complex types when
 * <complex.h> was included first, so that programs written for the
 * reference (which scale with `in[l] /= N`) compile unchanged; otherwise
 * two-element arrays.  Both have the same layout, so the library does not
 * depend on the choice. */
#if !defined(FFTW_NO_Complex) && defined(_Complex_I) && defined(complex) && defined(I)
typedef double _Complex pfft_complex;
typedef float _Complex pfftf_complex;
#else
typedef double pfft_complex[2];
typedef float pfftf_complex[2];
#endif

/* Planner flags: same bit values as the reference (api/pfft.h:528-545). */
#define PFFT_TRANSPOSED_NONE      (0U)
#define PFFT_TRANSPOSED_IN        (1U<< 0)
#define PFFT_TRANSPOSED_OUT       (1U<< 1)
#define PFFT_SHIFTED_NONE         (0U)
#define PFFT_SHIFTED_IN           (1U<< 2)
#define PFFT_SHIFTED_OUT          (1U<< 3)
#define PFFT_MEASURE              (0U)
#define PFFT_ESTIMATE             (1U<< 4)
#define PFFT_PATIENT              (1U<< 5)
#define PFFT_EXHAUSTIVE           (1U<< 6)
#define PFFT_NO_TUNE              (0U)
#define PFFT_TUNE                 (1U<< 7)
#define PFFT_PRESERVE_INPUT       (1U<< 8)
#define PFFT_DESTROY_INPUT        (1U<< 9)
#define PFFT_BUFFERED_INPLACE     (1U<<10)
#define PFFT_PADDED_R2C           (1U<<11)
#define PFFT_PADDED_C2R           (PFFT_PADDED_R2C)

/* Extension bits (>= 1<<12; unused by the reference): arithmetic form of
 * the pfftf_ (complex64) plans.  Default: plain FP32 Stockham. */
#define PFFT_AMD_MFMA_F32         (1U<<12)  /* radix-16 stages as FP32 MFMA products */
#define PFFT_AMD_MFMA_BF16X2      (1U<<13)  /* split hi+lo BF16 MFMA, FP32 accumulate */

#define PFFT_DEFAULT_BLOCK        (0)       /* FFTW_MPI_DEFAULT_BLOCK */
#define PFFT_DEFAULT_BLOCKS       NULL
#define PFFT_NO_GCELLS            NULL

/* Ghost-cell planner flags: same bit values as the reference
 * (api/pfft.h:563-572).  PFFT_GC_SENDRECV and PFFT_GC_RMA choose an MPI
 * algorithm in the reference and are accepted and ignored here. */
/* value types of pfft_get_args (api/pfft.h:555-561) */
#define PFFT_INT                  (1U)
#define PFFT_PTRDIFF_T            (2U)
#define PFFT_FLOAT                (3U)
#define PFFT_DOUBLE               (4U)
#define PFFT_LDOUBLE              (5U)
#define PFFT_UNSIGNED             (6U)
#define PFFT_SWITCH               (7U)

#define PFFT_GC_TRANSPOSED_NONE   (0U)
#define PFFT_GC_TRANSPOSED        (1U<< 0)
#define PFFT_GC_SENDRECV          (1U<< 1)
#define PFFT_GC_RMA               (1U<< 2)
#define PFFT_GC_R2C               (1U<< 3)
#define PFFT_GC_PADDED            (1U<< 4)
#define PFFT_GC_C2R               (PFFT_GC_R2C)
#define PFFT_GC_PADDED_R2C        (PFFT_GC_R2C | PFFT_GC_PADDED)
#define PFFT_GC_PADDED_C2R        (PFFT_GC_C2R | PFFT_GC_PADDED)

/*
 * One declaration list for both precisions.
 *   P: name prefix   R: real type   C: complex type   INT: ptrdiff_t
 */
#define PFFT_AMD_DECLARE_API(P, R, C, INT)                                          \
  typedef struct P##plan_s *P##plan;                                                \
  typedef struct P##gcplan_s *P##gcplan;                                            \
  typedef struct P##gctimer_s *P##gctimer;                                          \
  typedef struct P##timer_s *P##timer;                                              \
  /* setup: api/pfft.h:77-78 (init, cleanup), 79-80 (threads) */                   \
  void P##init(void);                                                               \
  void P##cleanup(void);                                                            \
  void P##plan_with_nthreads(int nthreads);                                         \
  int  P##get_nthreads(void);                                                       \
  /* memory: api/pfft.h:82-85 */                                                    \
  void *P##malloc(size_t n);                                                        \
  R    *P##alloc_real(size_t n);                                                    \
  C    *P##alloc_complex(size_t n);                                                 \
  void  P##free(void *p);                                                           \
  /* execution: api/pfft.h:87-92 */                                                 \
  void P##execute(const P##plan plan);                                              \
  void P##execute_dft(const P##plan plan, C *in, C *out);                           \
  void P##execute_dft_r2c(const P##plan plan, R *in, C *out);                       \
  void P##execute_dft_c2r(const P##plan plan, C *in, R *out);                       \
  void P##execute_r2r(const P##plan plan, R *in, R *out);                           \
  void P##destroy_plan(P##plan plan);                                               \
  /* test-data tools: api/pfft.h:94-149 */                                          \
  void P##init_input_complex_3d(const INT *n, const INT *local_n,                   \
      const INT *local_n_start, C *data);                                           \
  void P##init_input_complex(int rnk_n, const INT *n, const INT *local_n,           \
      const INT *local_start, C *data);                                             \
  void P##init_input_complex_hermitian_3d(const INT *n, const INT *local_n,         \
      const INT *local_n_start, C *data);                                           \
  void P##init_input_complex_hermitian(int rnk_n, const INT *n,                     \
      const INT *local_n, const INT *local_start, C *data);                         \
  void P##init_input_real_3d(const INT *n, const INT *local_n,                      \
      const INT *local_n_start, R *data);                                           \
  void P##init_input_real(int rnk_n, const INT *n, const INT *local_n,              \
      const INT *local_start, R *data);                                             \
  void P##clear_input_complex_3d(const INT *n, const INT *local_n,                  \
      const INT *local_n_start, C *data);                                           \
  void P##clear_input_complex(int rnk_n, const INT *n, const INT *local_n,          \
      const INT *local_start, C *data);                                             \
  void P##clear_input_complex_hermitian_3d(const INT *n, const INT *local_n,        \
      const INT *local_n_start, C *data);                                           \
  void P##clear_input_complex_hermitian(int rnk_n, const INT *n,                    \
      const INT *local_n, const INT *local_start, C *data);                         \
  void P##clear_input_real_3d(const INT *n, const INT *local_n,                     \
      const INT *local_n_start, R *data);                                           \
  void P##clear_input_real(int rnk_n, const INT *n, const INT *local_n,             \
      const INT *local_start, R *data);                                             \
  R P##check_output_complex_3d(const INT *n, const INT *local_n,                    \
      const INT *local_n_start, const C *data, MPI_Comm comm);                      \
  R P##check_output_complex(int rnk_n, const INT *n, const INT *local_n,            \
      const INT *local_start, const C *data, MPI_Comm comm);                        \
  R P##check_output_complex_hermitian_3d(const INT *n, const INT *local_n,          \
      const INT *local_n_start, const C *data, MPI_Comm comm);                      \
  R P##check_output_complex_hermitian(int rnk_n, const INT *n,                      \
      const INT *local_n, const INT *local_start, const C *data, MPI_Comm comm);    \
  R P##check_output_real_3d(const INT *n, const INT *local_n,                       \
      const INT *local_n_start, const R *data, MPI_Comm comm);                      \
  R P##check_output_real(int rnk_n, const INT *n, const INT *local_n,               \
      const INT *local_start, const R *data, MPI_Comm comm);                        \
  /* layout: api/pfft.h:151-206 (local_size_*), 214-275 (local_block_*) */          \
  INT P##local_size_dft_3d(const INT *n, MPI_Comm comm_cart, unsigned pfft_flags,   \
      INT *local_ni, INT *local_i_start, INT *local_no, INT *local_o_start);        \
  INT P##local_size_dft_r2c_3d(const INT *n, MPI_Comm comm_cart,                    \
      unsigned pfft_flags, INT *local_ni, INT *local_i_start, INT *local_no,        \
      INT *local_o_start);                                                          \
  INT P##local_size_dft_c2r_3d(const INT *n, MPI_Comm comm_cart,                    \
      unsigned pfft_flags, INT *local_ni, INT *local_i_start, INT *local_no,        \
      INT *local_o_start);                                                          \
  INT P##local_size_dft(int rnk_n, const INT *n, MPI_Comm comm_cart,                \
      unsigned pfft_flags, INT *local_ni, INT *local_i_start, INT *local_no,        \
      INT *local_o_start);                                                          \
  INT P##local_size_dft_r2c(int rnk_n, const INT *n, MPI_Comm comm_cart,            \
      unsigned pfft_flags, INT *local_ni, INT *local_i_start, INT *local_no,        \
      INT *local_o_start);                                                          \
  INT P##local_size_dft_c2r(int rnk_n, const INT *n, MPI_Comm comm_cart,            \
      unsigned pfft_flags, INT *local_ni, INT *local_i_start, INT *local_no,        \
      INT *local_o_start);                                                          \
  INT P##local_size_many_dft(int rnk_n, const INT *n, const INT *ni,                \
      const INT *no, INT howmany, const INT *iblock, const INT *oblock,             \
      MPI_Comm comm_cart, unsigned pfft_flags, INT *local_ni,                       \
      INT *local_i_start, INT *local_no, INT *local_o_start);                       \
  INT P##local_size_many_dft_r2c(int rnk_n, const INT *n, const INT *ni,            \
      const INT *no, INT howmany, const INT *iblock, const INT *oblock,             \
      MPI_Comm comm_cart, unsigned pfft_flags, INT *local_ni,                       \
      INT *local_i_start, INT *local_no, INT *local_o_start);                       \
  INT P##local_size_many_dft_c2r(int rnk_n, const INT *n, const INT *ni,            \
      const INT *no, INT howmany, const INT *iblock, const INT *oblock,             \
      MPI_Comm comm_cart, unsigned pfft_flags, INT *local_ni,                       \
      INT *local_i_start, INT *local_no, INT *local_o_start);                       \
  void P##local_block_dft_3d(const INT *n, MPI_Comm comm_cart, int pid,             \
      unsigned pfft_flags, INT *local_ni, INT *local_i_start, INT *local_no,        \
      INT *local_o_start);                                                          \
  void P##local_block_dft_r2c_3d(const INT *n, MPI_Comm comm_cart, int pid,         \
      unsigned pfft_flags, INT *local_ni, INT *local_i_start, INT *local_no,        \
      INT *local_o_start);                                                          \
  void P##local_block_dft_c2r_3d(const INT *n, MPI_Comm comm_cart, int pid,         \
      unsigned pfft_flags, INT *local_ni, INT *local_i_start, INT *local_no,        \
      INT *local_o_start);                                                          \
  void P##local_block_dft(int rnk_n, const INT *n, MPI_Comm comm_cart, int pid,     \
      unsigned pfft_flags, INT *local_ni, INT *local_i_start, INT *local_no,        \
      INT *local_o_start);                                                          \
  void P##local_block_dft_r2c(int rnk_n, const INT *n, MPI_Comm comm_cart,          \
      int pid, unsigned pfft_flags, INT *local_ni, INT *local_i_start,              \
      INT *local_no, INT *local_o_start);                                           \
  void P##local_block_dft_c2r(int rnk_n, const INT *n, MPI_Comm comm_cart,          \
      int pid, unsigned pfft_flags, INT *local_ni, INT *local_i_start,              \
      INT *local_no, INT *local_o_start);                                           \
  void P##local_block_many_dft(int rnk_n, const INT *ni, const INT *no,             \
      const INT *iblock, const INT *oblock, MPI_Comm comm_cart, int pid,            \
      unsigned pfft_flags, INT *local_ni, INT *local_i_start, INT *local_no,        \
      INT *local_o_start);                                                          \
  void P##local_block_many_dft_r2c(int rnk_n, const INT *ni, const INT *no,         \
      const INT *iblock, const INT *oblock, MPI_Comm comm_cart, int pid,            \
      unsigned pfft_flags, INT *local_ni, INT *local_i_start, INT *local_no,        \
      INT *local_o_start);                                                          \
  void P##local_block_many_dft_c2r(int rnk_n, const INT *ni, const INT *no,         \
      const INT *iblock, const INT *oblock, MPI_Comm comm_cart, int pid,            \
      unsigned pfft_flags, INT *local_ni, INT *local_i_start, INT *local_no,        \
      INT *local_o_start);                                                          \
  /* r2r layout: api/pfft.h:163-166, 183-187, 207-212, 226-229, 246-250, 270-275 */ \
  INT P##local_size_r2r_3d(const INT *n, MPI_Comm comm_cart, unsigned pfft_flags,   \
      INT *local_ni, INT *local_i_start, INT *local_no, INT *local_o_start);        \
  INT P##local_size_r2r(int rnk_n, const INT *n, MPI_Comm comm_cart,                \
      unsigned pfft_flags, INT *local_ni, INT *local_i_start, INT *local_no,        \
      INT *local_o_start);                                                          \
  INT P##local_size_many_r2r(int rnk_n, const INT *n, const INT *ni,                \
      const INT *no, INT howmany, const INT *iblock, const INT *oblock,             \
      MPI_Comm comm_cart, unsigned pfft_flags, INT *local_ni,                       \
      INT *local_i_start, INT *local_no, INT *local_o_start);                       \
  void P##local_block_r2r_3d(const INT *n, MPI_Comm comm_cart, int pid,             \
      unsigned pfft_flags, INT *local_ni, INT *local_i_start, INT *local_no,        \
      INT *local_o_start);                                                          \
  void P##local_block_r2r(int rnk_n, const INT *n, MPI_Comm comm_cart, int pid,     \
      unsigned pfft_flags, INT *local_ni, INT *local_i_start, INT *local_no,        \
      INT *local_o_start);                                                          \
  void P##local_block_many_r2r(int rnk_n, const INT *ni, const INT *no,             \
      const INT *iblock, const INT *oblock, MPI_Comm comm_cart, int pid,            \
      unsigned pfft_flags, INT *local_ni, INT *local_i_start, INT *local_no,        \
      INT *local_o_start);                                                          \
  /* planning: api/pfft.h:277-317 (basic/many), 319-343 (skipped) */                \
  P##plan P##plan_dft_3d(const INT *n, C *in, C *out, MPI_Comm comm_cart,           \
      int sign, unsigned pfft_flags);                                               \
  P##plan P##plan_dft_r2c_3d(const INT *n, R *in, C *out, MPI_Comm comm_cart,       \
      int sign, unsigned pfft_flags);                                               \
  P##plan P##plan_dft_c2r_3d(const INT *n, C *in, R *out, MPI_Comm comm_cart,       \
      int sign, unsigned pfft_flags);                                               \
  P##plan P##plan_dft(int rnk_n, const INT *n, C *in, C *out, MPI_Comm comm_cart,   \
      int sign, unsigned pfft_flags);                                               \
  P##plan P##plan_dft_r2c(int rnk_n, const INT *n, R *in, C *out,                   \
      MPI_Comm comm_cart, int sign, unsigned pfft_flags);                           \
  P##plan P##plan_dft_c2r(int rnk_n, const INT *n, C *in, R *out,                   \
      MPI_Comm comm_cart, int sign, unsigned pfft_flags);                           \
  P##plan P##plan_many_dft(int rnk_n, const INT *n, const INT *ni, const INT *no,   \
      INT howmany, const INT *iblock, const INT *oblock, C *in, C *out,             \
      MPI_Comm comm_cart, int sign, unsigned pfft_flags);                           \
  P##plan P##plan_many_dft_r2c(int rnk_n, const INT *n, const INT *ni,              \
      const INT *no, INT howmany, const INT *iblock, const INT *oblock, R *in,      \
      C *out, MPI_Comm comm_cart, int sign, unsigned pfft_flags);                   \
  P##plan P##plan_many_dft_c2r(int rnk_n, const INT *n, const INT *ni,              \
      const INT *no, INT howmany, const INT *iblock, const INT *oblock, C *in,      \
      R *out, MPI_Comm comm_cart, int sign, unsigned pfft_flags);                   \
  P##plan P##plan_many_dft_skipped(int rnk_n, const INT *n, const INT *ni,          \
      const INT *no, INT howmany, const INT *iblock, const INT *oblock,             \
      const int *skip_trafos, C *in, C *out, MPI_Comm comm_cart, int sign,          \
      unsigned pfft_flags);                                                         \
  P##plan P##plan_many_dft_r2c_skipped(int rnk_n, const INT *n, const INT *ni,      \
      const INT *no, INT howmany, const INT *iblock, const INT *oblock,             \
      const int *skip_trafos, R *in, C *out, MPI_Comm comm_cart, int sign,          \
      unsigned pfft_flags);                                                         \
  P##plan P##plan_many_dft_c2r_skipped(int rnk_n, const INT *n, const INT *ni,      \
      const INT *no, INT howmany, const INT *iblock, const INT *oblock,             \
      const int *skip_trafos, C *in, R *out, MPI_Comm comm_cart, int sign,          \
      unsigned pfft_flags);                                                         \
  /* r2r planning: api/pfft.h:286-288, 299-301, 318-322, 339-343.  Kinds:        \
   * PFFT_DHT, PFFT_REDFT00..11, PFFT_RODFT00..11 */                                \
  P##plan P##plan_r2r_3d(const INT *n, R *in, R *out, MPI_Comm comm_cart,           \
      const P##r2r_kind *kinds, unsigned pfft_flags);                               \
  P##plan P##plan_r2r(int rnk_n, const INT *n, R *in, R *out, MPI_Comm comm_cart,   \
      const P##r2r_kind *kinds, unsigned pfft_flags);                               \
  P##plan P##plan_many_r2r(int rnk_n, const INT *n, const INT *ni, const INT *no,   \
      INT howmany, const INT *iblock, const INT *oblock, R *in, R *out,             \
      MPI_Comm comm_cart, const P##r2r_kind *kinds, unsigned pfft_flags);           \
  P##plan P##plan_many_r2r_skipped(int rnk_n, const INT *n, const INT *ni,          \
      const INT *no, INT howmany, const INT *iblock, const INT *oblock,             \
      const int *skip_trafos, R *in, R *out, MPI_Comm comm_cart,                    \
      const P##r2r_kind *kinds, unsigned pfft_flags);                               \
  /* integer vector helpers: api/pfft.h:345-359 */                                  \
  INT  P##prod_INT(int d, const INT *vec);                                          \
  INT  P##sum_INT(int d, const INT *vec);                                           \
  int  P##equal_INT(int d, const INT *vec1, const INT *vec2);                       \
  void P##vcopy_INT(int d, const INT *vec1, INT *vec2);                             \
  void P##vadd_INT(int d, const INT *vec1, const INT *vec2, INT *sum);              \
  void P##vsub_INT(int d, const INT *vec1, const INT *vec2, INT *sum);              \
  /* rank-0 printing: api/pfft.h:416-421 */                                         \
  void P##vfprintf(MPI_Comm comm, FILE *stream, const char *format, va_list ap);    \
  void P##fprintf(MPI_Comm comm, FILE *stream, const char *format, ...);            \
  void P##printf(MPI_Comm comm, const char *format, ...);                           \
  /* process mesh: api/pfft.h:423-431 */                                            \
  int P##create_procmesh(int rnk_n, MPI_Comm comm, const int *np,                   \
      MPI_Comm *comm_cart);                                                         \
  int P##create_procmesh_1d(MPI_Comm comm, int np0, MPI_Comm *comm_cart_1d);        \
  int P##create_procmesh_2d(MPI_Comm comm, int np0, int np1,                        \
      MPI_Comm *comm_cart_2d);                                                      \
  /* array printing and command-line options: api/pfft.h:361-383 (one rank;  \
   * device arrays are copied to the host for printing) */                     \
  void P##apr_complex_3d(const C *data, const INT *local_n,                        \
      const INT *local_start, const char *name, MPI_Comm comm);                     \
  void P##apr_complex_permuted_3d(const C *data, const INT *local_n,               \
      const INT *local_start, int perm0, int perm1, int perm2, const char *name,    \
      MPI_Comm comm);                                                               \
  void P##apr_real_3d(const R *data, const INT *local_n,                           \
      const INT *local_start, const char *name, MPI_Comm comm);                     \
  void P##apr_real_permuted_3d(const R *data, const INT *local_n,                  \
      const INT *local_start, int perm0, int perm1, int perm2, const char *name,    \
      MPI_Comm comm);                                                               \
  void P##get_args(int argc, char **argv, const char *name, int neededArgs,        \
      unsigned type, void *parameter);                                              \
  /* plan timers: api/pfft.h:386-412.  Each execute adds its wall time to     \
   * `whole` and each pass's device time to the trafo slot of the reference  \
   * step that transforms that dimension; remaps and twiddles are fused into \
   * the passes and read 0.  convert_timer2vec returns malloc'd memory. */     \
  void P##reset_timer(P##plan ths);                                                 \
  P##timer P##get_timer(const P##plan ths);                                         \
  void P##print_average_timer(const P##plan ths, MPI_Comm comm);                    \
  void P##print_average_timer_adv(const P##plan ths, MPI_Comm comm);                \
  void P##write_average_timer(const P##plan ths, const char *name,                  \
      MPI_Comm comm);                                                               \
  void P##write_average_timer_adv(const P##plan ths, const char *name,              \
      MPI_Comm comm);                                                               \
  P##timer P##copy_timer(const P##timer orig);                                      \
  void P##average_timer(P##timer ths);                                              \
  P##timer P##add_timers(const P##timer sum1, const P##timer sum2);                 \
  P##timer P##reduce_max_timer(const P##timer ths, MPI_Comm comm);                  \
  double *P##convert_timer2vec(const P##timer ths);                                 \
  P##timer P##convert_vec2timer(const double *times);                               \
  void P##destroy_timer(P##timer ths);                                              \
  /* ghost cells: api/pfft.h:433-501.  Three-dimensional arrays; on one         \
   * process the ghost cells are periodic images of the local block.  The       \
   * array must hold local_size_*gc reals / complex values. */                  \
  INT P##local_size_gc_3d(const INT *local_n, const INT *local_start,              \
      const INT *gc_below, const INT *gc_above, INT *local_ngc,                     \
      INT *local_gc_start);                                                         \
  INT P##local_size_gc(int rnk_n, const INT *local_n, const INT *local_start,      \
      const INT *gc_below, const INT *gc_above, INT *local_ngc,                     \
      INT *local_gc_start);                                                         \
  INT P##local_size_many_gc(int rnk_n, const INT *local_n,                         \
      const INT *local_start, INT howmany, const INT *gc_below,                     \
      const INT *gc_above, INT *local_ngc, INT *local_gc_start);                    \
  P##gcplan P##plan_rgc_3d(const INT *n, const INT *gc_below,                       \
      const INT *gc_above, R *data, MPI_Comm comm_cart, unsigned gc_flags);         \
  P##gcplan P##plan_cgc_3d(const INT *n, const INT *gc_below,                       \
      const INT *gc_above, C *data, MPI_Comm comm_cart, unsigned gc_flags);         \
  P##gcplan P##plan_rgc(int rnk_n, const INT *n, const INT *gc_below,               \
      const INT *gc_above, R *data, MPI_Comm comm_cart, unsigned gc_flags);         \
  P##gcplan P##plan_cgc(int rnk_n, const INT *n, const INT *gc_below,               \
      const INT *gc_above, C *data, MPI_Comm comm_cart, unsigned gc_flags);         \
  P##gcplan P##plan_many_rgc(int rnk_n, const INT *n, INT howmany,                  \
      const INT *block, const INT *gc_below, const INT *gc_above, R *data,          \
      MPI_Comm comm_cart, unsigned gc_flags);                                       \
  P##gcplan P##plan_many_cgc(int rnk_n, const INT *n, INT howmany,                  \
      const INT *block, const INT *gc_below, const INT *gc_above, C *data,          \
      MPI_Comm comm_cart, unsigned gc_flags);                                       \
  void P##exchange(P##gcplan ths);                                                  \
  void P##reduce(P##gcplan ths);                                                    \
  void P##destroy_gcplan(P##gcplan ths);                                            \
  void P##reset_gctimers(P##gcplan ths);                                            \
  P##gctimer P##get_gctimer_exg(const P##gcplan ths);                               \
  P##gctimer P##get_gctimer_red(const P##gcplan ths);                               \
  void P##print_average_gctimer(const P##gcplan ths, MPI_Comm comm);                \
  void P##print_average_gctimer_adv(const P##gcplan ths, MPI_Comm comm);            \
  void P##write_average_gctimer(const P##gcplan ths, const char *name,              \
      MPI_Comm comm);                                                               \
  void P##write_average_gctimer_adv(const P##gcplan ths, const char *name,          \
      MPI_Comm comm);                                                               \
  P##gctimer P##copy_gctimer(const P##gctimer orig);                                \
  void P##average_gctimer(P##gctimer ths);                                          \
  P##gctimer P##add_gctimers(const P##gctimer sum1, const P##gctimer sum2);         \
  P##gctimer P##reduce_max_gctimer(const P##gctimer ths, MPI_Comm comm);            \
  void P##convert_gctimer2vec(const P##gctimer ths, double *times);                 \
  P##gctimer P##convert_vec2gctimer(const double *times);                           \
  void P##destroy_gctimer(P##gctimer ths);

PFFT_AMD_DECLARE_API(pfft_, double, pfft_complex, ptrdiff_t)
PFFT_AMD_DECLARE_API(pfftf_, float, pfftf_complex, ptrdiff_t)

/* ---------------------------------------------------------------------
 * Extensions (no counterpart in the reference; they change no existing
 * symbol).
 * ------------------------------------------------------------------ */

/* Execute the adjoint (conjugate-transposed) pipeline of a plan: maps an
 * array shaped like the plan's output to an array shaped like its input.
 * For r2c plans the adjoint includes the Hermitian weighting (factor 2 on
 * every half-spectrum mode except the DC and Nyquist planes).  Blocking. */
void pfft_execute_adjoint(const pfft_plan plan, void *in, void *out);
void pfftf_execute_adjoint(const pfftf_plan plan, void *in, void *out);

/* Execute `count` plans as one batch: every plan's pass 0, then every
 * pass 1, and so on, each level as persistent device launches (one per
 * kernel) over a work-list of the level's tiles, all captured into a
 * hipGraph on the first call with these plans and arrays and replayed
 * later.  Plans the batch kernels do not cover run one after another.
 * Blocking.  Replaces pfft_execute for many small plans; the reference has
 * no batch call (api/api-basic.c:1109 executes one plan). */
void pfftf_execute_batch(const pfftf_plan *plans, int count);
void pfft_execute_batch(const pfft_plan *plans, int count);

/* The batch without any allocation, for callers that cannot allocate
 * (code under graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  No counterpart in the reference.  Same kernels and result
 * as pfft_execute_batch (launched directly, not through a graph), with the
 * work-list in a fixed device array of the library (1.1 MiB, 2048 task
 * records of 560 bytes, per GPU, kept until the process ends): at most 2048 passes in total (a 3-d plan has 3), plans of
 * one precision (pfft_plan or pfftf_plan handles) on device arrays, and no
 * long-line, r2r or MFMA-form passes.  Blocking; returns 0 on success and 1
 * with a message otherwise.  Calls that overlap on one GPU share the array
 * and race. */
int pfft_amd_execute_batch_noalloc(const void *const *plans, int count);

/* Copies between host and device memory (blocking), for callers that
 * fill arrays from the host. 0 on success. */
int pfft_amd_memcpy_to_device(void *dst_dev, const void *src_host, size_t bytes);
int pfft_amd_memcpy_to_host(void *dst_host, const void *src_dev, size_t bytes);
/* Measurement hooks: number of device passes of a plan; mean duration
 * (ms) of each pass over `iters` runs on device arrays in/out (after one
 * untimed run) together with the bytes each pass reads plus writes; a text
 * description of the passes. */
int pfft_amd_plan_npasses(const void *plan);
int pfft_amd_time_passes(const void *plan, void *in, void *out, int iters, double *ms,
                         double *bytes);
int pfft_amd_describe_plan(const void *plan, char *buf, size_t len);
/* Attainable HBM bandwidth: the best of `iters` timed launches of a
 * streaming kernel over `bytes` per buffer, in GB/s of bytes read plus
 * written.  mode 0 read only, 1 write only, 2 copy; width 16, 32 or 64
 * contiguous bytes per lane and access; nt nonzero for nontemporal
 * accesses; grid = 256-thread workgroups of a persistent grid, or 0 for
 * one workgroup per 4 KiB..16 KiB chunk.  Negative on error. */
double pfft_amd_bw_probe(int mode, int width, int nt, long long grid, size_t bytes, int iters);
/* Wisdom: the launch variants the planner chose by timing (flags
 * PFFT_MEASURE, PFFT_PATIENT, PFFT_EXHAUSTIVE; none with PFFT_ESTIMATE),
 * keyed by pass geometry.  Export writes the process-wide cache to a text
 * file, import merges a file into it (0 on success), forget empties it;
 * count is the number of cached selections and tune_timings the number of
 * candidate timings performed so far.  PFFT_AMD_WISDOM=<file> imports the
 * file before the first tuned plan and rewrites it as selections are
 * added.  FFTW offers the same services as fftw_export_wisdom_to_filename
 * and friends; the reference exposes none (kernel/ipfft.h:44-50 only
 * forgets FFTW's wisdom). */
int pfft_amd_export_wisdom(const char *filename);
int pfft_amd_import_wisdom(const char *filename);
void pfft_amd_forget_wisdom(void);
int pfft_amd_wisdom_count(void);
int pfft_amd_tune_timings(void);
/* Library version string. */
const char *pfft_amd_version(void);
/* Last error message of the device runtime (empty string if none). */
const char *pfft_amd_last_error(void);

#ifdef __cplusplus
}
#endif
#endif /* PFFT_H */
