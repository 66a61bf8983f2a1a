/*
 * sweeper.h -- drop-in C API of the MI355X sweep (libminisweep.so).
 *
 * Source-compatible replacement for the reference's sweeper header and the
 * types it pulls in.  A caller written against the reference (its only caller
 * is Runner_run_case, src/4_driver/runner.c:51-182) compiles against this
 * header unchanged and runs the sweep on the GPU.  Each declaration cites the
 * reference interface it replaces.
 *
 * Conventions kept from the reference:
 *   - the caller owns Sweeper by value, and owns vi/vo (Pointer), Quantities,
 *     Env and Arguments (src/4_driver/runner.c:55-137);
 *   - configuration errors print `Insist error: "<cond>". At file F, line L.`
 *     and exit(EXIT_FAILURE) (src/1_base/env_assert.c:27-32);
 *   - arguments are consumed once; leftovers are a fatal error at the caller
 *     (src/1_base/arguments.c:136-198, src/4_driver/runner.c:141-142).
 */
#ifndef MSW_SWEEPER_H
#define MSW_SWEEPER_H

#include <stddef.h>

#include "msw_plan.h"

#ifdef __cplusplus
extern "C" {
#endif

/* ---- base types (src/1_base/types_kernels.h:39-49) ---------------------- */
typedef double P;
typedef int Bool_t;
enum { Bool_false = 0, Bool_true = 1 };
static inline P P_zero(void) { return (P)0; }
static inline P P_one(void) { return (P)1; }

/* src/2_sweeper_base/definitions_kernels.h:25-64 */
enum { NDIM = 3 };
enum { NOCTANT = 8 };
enum { DIR_UP = +1, DIR_DN = -1 };
static inline int Dir_x(int octant) { return octant & (1 << 0) ? DIR_DN : DIR_UP; }
static inline int Dir_y(int octant) { return octant & (1 << 1) ? DIR_DN : DIR_UP; }
static inline int Dir_z(int octant) { return octant & (1 << 2) ? DIR_DN : DIR_UP; }
static inline int Dir_inc(int dir) { return dir; }

/* Unknowns per cell and default moment count
 * (src/2_sweeper_base/dimensions_kernels.h:24-36).  NU is fixed at 4 in this
 * library build; nm is a runtime value (--nm), defaulting to NM. */
enum { NU = 4 };
#ifdef NM_VALUE
enum { NM = NM_VALUE };
#else
enum { NM = 16 };
#endif

/* ---- assertions (src/1_base/env_assert_kernels.h:27-34) ------------------ */
void insist_(const char* condition_string, const char* file, int line);
#ifndef Insist
#define Insist(condition) \
  (void)((condition) || (insist_(#condition, __FILE__, __LINE__), 0))
#endif

/* ---- Arguments (src/1_base/arguments.h:36-88) ---------------------------- */
typedef struct {
  int argc;
  char** argv_unconsumed;
  char* argstring;
} Arguments;

Arguments Arguments_null(void);
void Arguments_create(Arguments* args, int argc, char** argv);
void Arguments_create_from_string(Arguments* args, const char* argstring);
void Arguments_destroy(Arguments* args);
Bool_t Arguments_exists(const Arguments* args, const char* arg_name);
int Arguments_consume_int_(Arguments* args, const char* arg_name);
int Arguments_consume_int_or_default(Arguments* args, const char* arg_name,
                                     int default_value);
Bool_t Arguments_are_all_consumed(const Arguments* args);

/* ---- Env (src/1_base/env.h:46-80, env_types.h:54-70) -------------------- */
typedef double Timer;
typedef struct {
  Bool_t is_using_device_;   /* --is_using_device: accepted; the sweep always
                                runs on the GPU in this library */
  Bool_t is_state_resident_; /* vi/vo device copies are authoritative across
                                sweeps (no per-sweep host<->device copies) */
  int precision_;            /* MSW_PREC_* (--precision, this library) */
  int pad_;
  void* stream_;             /* hipStream_t of all sweep work */
} Env;

Env Env_null(void);
void Env_initialize(Env* env, int argc, char** argv);
void Env_set_values(Env* env, Arguments* args);
void Env_finalize(Env* env);
Bool_t Env_is_proc_master(Env* env);
Bool_t Env_is_proc_active(const Env* env);
Timer Env_get_time(void);
Timer Env_get_synced_time(Env* env);
/* one process, one GPU: the reference's MPI facade reduces to these */
int Env_nproc(const Env* env);
int Env_nproc_x(const Env* env);
int Env_nproc_y(const Env* env);
int Env_proc_x_this(const Env* env);
int Env_proc_y_this(const Env* env);
double Env_sum_d(Env* env, double value);
P Env_sum_P(Env* env, P value);
Bool_t Env_cuda_is_using_device(const Env* env);
void Env_set_state_resident(Env* env, Bool_t resident);
/* hipStream_t of the sweep work, created on first use (this library). */
void* Env_stream(Env* env);

/* ---- Dimensions (src/2_sweeper_base/dimensions_kernels.h:41-56) ---------- */
typedef struct {
  int ncell_x, ncell_y, ncell_z;
  int ne;
  int nm;
  int na;
} Dimensions;

/* src/2_sweeper_base/dimensions.c:36-101 */
size_t Dimensions_size_state(const Dimensions dims, int nu);
size_t Dimensions_size_state_angles(const Dimensions dims, int nu);
size_t Dimensions_size_facexy(const Dimensions dims, int nu, int noct);
size_t Dimensions_size_facexz(const Dimensions dims, int nu, int noct);
size_t Dimensions_size_faceyz(const Dimensions dims, int nu, int noct);

/* ---- Pointer: mirrored host/device array (src/1_base/pointer_kernels.h:33-41,
 * pointer.h:28-112).  Host memory is pinned when requested; device memory is
 * HBM. --------------------------------------------------------------------- */
typedef struct {
  size_t n_;
  P* h_;
  P* d_;
  Bool_t is_using_device_;
  Bool_t is_pinned_;
  Bool_t is_alias_;
} Pointer;

Pointer Pointer_null(void);
void Pointer_create(Pointer* p, size_t n, Bool_t is_using_device);
void Pointer_create_alias(Pointer* p, Pointer* source, size_t base, size_t n);
void Pointer_set_pinned(Pointer* p, Bool_t is_pinned);
void Pointer_destroy(Pointer* p);
void Pointer_allocate(Pointer* p);
void Pointer_deallocate(Pointer* p);
void Pointer_update_h(Pointer* p);
void Pointer_update_d(Pointer* p);
void Pointer_update_h_stream(Pointer* p, void* stream);
void Pointer_update_d_stream(Pointer* p, void* stream);
static inline P* Pointer_h(Pointer* p) { return p->h_; }
static inline const P* Pointer_const_h(const Pointer* p) { return p->h_; }
static inline P* Pointer_d(Pointer* p) { return p->d_; }
static inline const P* Pointer_const_d(const Pointer* p) { return p->d_; }
static inline P* Pointer_active(Pointer* p) {
  return p->is_using_device_ ? p->d_ : p->h_;
}

/* ---- Quantities: the manufactured problem
 * (src/3_sweeper/quantities_testing_kernels.h:35-46, quantities_testing.h:27-90) */
typedef struct {
  Pointer a_from_m; /* [octant][im][ia], im stride dims.nm */
  Pointer m_from_a; /* [octant][ia][im], im stride dims.nm */
  int* ix_base_vals;
  int* iy_base_vals;
  int ix_base;
  int iy_base;
  int ncell_x_g;
  int ncell_y_g;
  int ncell_z_g;
} Quantities;

void Quantities_create(Quantities* quan, const Dimensions dims, Env* env);
void Quantities_destroy(Quantities* quan);
double Quantities_flops_per_solve(const Dimensions dims);

static inline int Quantities_scalefactor_energy_(int ie, Dimensions dims) {
  (void)dims;
  return 1 << (((ie * 1366 + 150889) % 714025) & 3);
}
static inline int Quantities_scalefactor_unknown_(int iu) {
  return 1 << (((iu * 741 + 66037) % 312500) & 3);
}
static inline int Quantities_scalefactor_space_(const Quantities* quan, int ix_g,
                                                int iy_g, int iz_g) {
  (void)quan;
  return 1 << ((ix_g + 3 * iy_g + 7 * iz_g + 2) & 3);
}
/* Quantities_init_state (src/3_sweeper/quantities_testing.h:60-90) */
static inline P Quantities_init_state(const Quantities* quan, int ix, int iy,
                                      int iz, int ie, int im, int iu,
                                      const Dimensions dims) {
  return ((P)(1 + im)) *
         (P)Quantities_scalefactor_space_(quan, ix + quan->ix_base,
                                          iy + quan->iy_base, iz) *
         (P)Quantities_scalefactor_energy_(ie, dims) *
         (P)Quantities_scalefactor_unknown_(iu);
}

/* ---- Sweeper (src/3_sweeper/sweeper_kba.h:34-81, 252-257) --------------- */
typedef struct {
  Dimensions dims;
  Dimensions dims_b;
  Dimensions dims_g;
  /* KBA knobs: consumed and validated as the reference does
   * (src/3_sweeper/sweeper_kba_c.h:53-232); on the GPU they are tuning
   * hints only, the hyperplane schedule is chosen by the plan. */
  int nthread_e;
  int nthread_octant;
  int nthread_x;
  int nthread_y;
  int nthread_z;
  int nblock_z;
  int nblock_octant;
  int noctant_per_block;
  int nsemiblock;
  int ncell_x_per_subblock;
  int ncell_y_per_subblock;
  int ncell_z_per_subblock;
  int is_face_comm_async;
  int precision;
  MswPlan* plan; /* device schedule, face buffers */
} Sweeper;

Sweeper Sweeper_null(void);
void Sweeper_create(Sweeper* sweeper, Dimensions dims, const Quantities* quan,
                    Env* env, Arguments* args);
void Sweeper_destroy(Sweeper* sweeper, Env* env);
/* vo <- S(vi).  With env resident: reads Pointer_d(vi), writes Pointer_d(vo).
 * Otherwise host arrays are authoritative, as in the reference: vi is copied
 * to the device first and vo is readable at Pointer_h(vo) on return
 * (src/3_sweeper/sweeper_kba_c.h:682-813). */
void Sweeper_sweep(Sweeper* sweeper, Pointer* vo, Pointer* vi,
                   const Quantities* quan, Env* env);
static inline int Sweeper_noctant_per_block(const Sweeper* sweeper) {
  return sweeper->noctant_per_block;
}

#ifdef __cplusplus
}
#endif

#endif
